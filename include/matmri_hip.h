/*
 * matmri_hip.h -- C ABI of the MI355X (gfx950) implementation of MatMRI's
 * iterative-SENSE operator chain.
 *
 * Every entry point is plain C: opaque handles, plain pointers, 64-bit sizes,
 * int return codes.  Complex data is interleaved (re, im) single precision
 * ("c64") and column-major, exactly the memory layout of a MATLAB complex
 * single array under the interleaved-complex MEX API.  Pointers passed with
 * `on_device = 1` are device addresses (HBM); with `on_device = 0` they are
 * host addresses and the call copies in and out and synchronises.
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * MatMRI repository root):
 *
 *   mm_nufft_plan_create    nufftOp(imN,kloc,dcf,useGPU,os,kwidth)
 *                           operators/@nufftOp/nufftOp.m:89-136 (+ doPrep :138-274)
 *   mm_rcvr_plan_create     rcvrOp(maps,false)      operators/@rcvrOp/rcvrOp.m:47-94
 *   mm_hio_plan_create      sampHighOrder(b0,sampTimes,phs_spha,phs_conc,phs_grid,b0mask,...)
 *                           operators/@sampHighOrder/sampHighOrder.m:129-291, :496-588
 *   mm_hio_set_phi_div      sampHighOrder.setPhiDiv  sampHighOrder.m:590-622
 *   mm_hio_interp_create    sampHighOrder(...,approach=1,opt)  sampHighOrder.m:314-359, :624-902
 *   mm_nufft_set_toeplitz   nufftOp.prepToep          nufftOp.m:410-432 (applied :348-352)
 *   mm_sense_create         mrSampFunc(x,transp,samplingOp,R,opt)  operators/mrSampFunc.m:82-159
 *   mm_cart_plan_create     mrSampFunc's Cartesian-mask samplingOp  mrSampFunc.m:108-114, :137-143
 *   mm_sumdims_plan_create  mrSampFunc's opt.sumDims                mrSampFunc.m:100-105
 *   mm_interp_traj_time     interpTrajTime(phs_spha,phs_conc,tdwell,delTime,datatime)
 *                           findDel/interpTrajTime.m:1-27
 *   mm_hio_dense_normal     mrSampFuncMat(y,Crcvr,b0,sampTimes,phs_spha,phs_conc,phs_grid,imMask)
 *                           operators/mrSampFuncMat.m:1-47
 *   mm_op_apply             A*x / A'*y  (mtimes/ctranspose overloads:
 *                           nufftOp.m:276-408, rcvrOp.m:96-159,
 *                           sampHighOrder.m:293-391, mrSampFunc.m:82-159)
 *   mm_cgne_solve           cgne(A,b,x0,NitMax,opt,gam,R,W)  iterativeSolvers/cgne.m:1-341
 *   mm_cgne_solve_ex        the same with the gam/R Tikhonov terms and opt.gtruth
 *                           (cgne.m:171-179, :219-227, :201-204, :271-274)
 *
 * `bfista`/`powermethod` (iterativeSolvers/bfista.m, powermethod.m) stay on
 * the host side and call mm_op_apply; their regulariser R is host code.
 *
 *   mm_wsnet_*              no counterpart: the learned warm start that
 *                           produces cgne's x0 argument (cgne.m:1, :143-145)
 */
#ifndef MATMRI_HIP_H
#define MATMRI_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ------------------------------------------------------ */
#define MM_OK              0
#define MM_ERR_ARG         1   /* bad argument / shape mismatch (MATLAB error('...')) */
#define MM_ERR_OOM         2   /* device allocation failed (nufftOp.m:336-345 hint)    */
#define MM_ERR_DEVICE      3   /* no usable gfx950 device                               */
#define MM_ERR_HIP         4   /* other HIP runtime error                               */
#define MM_ERR_UNSUPPORTED 5   /* valid in the reference, not implemented here          */

typedef struct mm_ctx mm_ctx;   /* one device, one HIP stream; not thread-safe */
typedef struct mm_op  mm_op;    /* a linear operator (NUFFT, coils, high-order, SENSE) */
typedef struct mm_wsnet mm_wsnet; /* learned warm-start network (one image size) */

/* Message of the last failing call on this thread ("" if none). */
const char* mm_last_error(void);
/* Library version string. */
const char* mm_version(void);
/* Number of visible HIP devices (0 when none); never fails. */
int mm_device_count(void);

/* Context: `stream` is a hipStream_t to run on (NULL = the null stream). */
int mm_ctx_create(int device, void* stream, mm_ctx** out);
int mm_ctx_destroy(mm_ctx* ctx);
int mm_ctx_sync(mm_ctx* ctx);

/* ---- operators ---------------------------------------------------------- */
/* Kaiser-Bessel gridding NUFFT over `nslice` independent 1-D or 2-D slices
 * (nslice = 1 is nufftOp).  Slice s has image size imN[2s], imN[2s+1]
 * (imN[2s+1] = 1 for 1-D), M[s] samples, kloc[s] = column-major [M x ndim]
 * cycles/pixel in [-0.5, 0.5) (wrapped), dcf[s] = [M] or NULL.  `os` and
 * `kwidth` follow nufftOp (kwidth even, <= 8).  Data of all slices is
 * concatenated: images [V_s x B] then samples [M_s x B] per slice, where B is
 * the number of transforms per slice passed to mm_op_apply as `nrep`.
 * ndim = 3 (one volume, nslice = 1): imN[0..2] = N1, N2, N3, kloc[0] =
 * column-major [M x 3], grid sides up to 512 (nufftOp.m:195-197, :244-245,
 * :268-271).  Odd sides follow the reference's ramp exactly. */
int mm_nufft_plan_create(mm_ctx* ctx, int ndim, int nslice, const int64_t* imN,
                         const int64_t* M, const double* const* kloc,
                         const double* const* dcf, double os, int kwidth,
                         mm_op** out);

/* Coil sensitivities: maps = column-major [V x C] c64 (image dims flattened). */
int mm_rcvr_plan_create(mm_ctx* ctx, int64_t V, int64_t C, const float* maps,
                        mm_op** out);

/* Coil sensitivities of `nslice` slices for a multi-slice SENSE operator:
 * maps = concatenation over slices of column-major [V_s x C]. */
int mm_rcvr_plan_create_multi(mm_ctx* ctx, int nslice, const int64_t* V, int64_t C,
                              const float* maps, mm_op** out);

/* Expanded-encoding operator.  x,y,z = voxel positions [V] (m); b0 = [V]
 * rad/s or NULL; t = [K] s; h = column-major [Lh x K] (phs_spha);
 * g = column-major [Lc x K] (phs_conc) or NULL with Lc = 0; mask = [V] or NULL
 * (b0mask).  Phase terms, their masking and their order follow
 * sampHighOrder.m:496-588; the B0 term is present when K > 1. */
int mm_hio_plan_create(mm_ctx* ctx, int64_t V, int64_t K, int Lh, int Lc,
                       const double* x, const double* y, const double* z,
                       const double* b0, const double* t, const double* h,
                       const double* g, const double* mask, mm_op** out);
/* sampHighOrder approach 1 (interpolated, sampHighOrder.m:314-359,
 * :624-902): y = phsShft .* sum_n svdTime(:,n) .* NUFFT(x .* svdSpace(:,n)) and
 * its adjoint, over a single-slice NUFFT `nufft` of the V-voxel image onto
 * the K samples (kept by the caller for the operator's lifetime);
 * svdSpace [V x Ns], svdTime [K x Ns], phsShft [K], complex64 host arrays
 * (the truncated SVD of the non-linear phase, from the plan-time host
 * code).  The Ns modulated images of each input run as one NUFFT batch. */
int mm_hio_interp_create(mm_ctx* ctx, mm_op* nufft, int64_t V, int64_t K, int Ns, const float* svdSpace,
                         const float* svdTime, const float* phsShft, mm_op** out);
/* setPhiDiv: replaces the forward by (1i*exp(1i*Phi).*dPhi).' * x / sqrt(V)
 * with dPhi from hd [Lh x K] and gd [Lc x K] (no B0 term); NULL,NULL clears. */
int mm_hio_set_phi_div(mm_op* op, const double* hd, const double* gd);
/* prepToep (operators/@nufftOp/nufftOp.m:410-432, applied at :348-352): turns
 * a NUFFT plan whose grid is at least 2N - 1 per side (os = 2) into the
 * Toeplitz form of its normal operator.  psf: host complex64 (interleaved),
 * per slice the point-spread image S2'*(dcf or ones) of a NUFFT on the doubled
 * image size (2 N1 x 2 N2, or the 2 N1 x 2 N2 x 2 N3 volume of a 3-D plan,
 * column-major), slices concatenated.  Afterwards
 * both directions of mm_op_apply map an image to an image (n_data = n_image):
 * crop(fftnc(convMat .* ifftnc(pad(x)))), convMat = 2^ndim ifftnc(psf).  A
 * SENSE operator built on it maps, like mrSampFunc on such a nufftOp,
 * forward x -> T(R x) (coil images) and adjoint y -> R'(T y); its normal
 * operator is R'(T(R x)).  mm_cgne_solve rejects it (MM_ERR_UNSUPPORTED):
 * b = E'(W.*y) needs the gridding form. */
int mm_nufft_set_toeplitz(mm_op* op, const float* psf);

/* Cartesian sampling: mrSampFunc with a k-space mask as samplingOp
 * (operators/mrSampFunc.m:108-114, :137-143): forward y = mask .* ifftnc(x),
 * adjoint x = fftnc(mask .* y), the centred unitary transforms of
 * general/ifftnc.m and general/fftnc.m over the ndim (1 or 2) image
 * dimensions; imN[0] (x imN[1]) is the image and k-space size, mask a real
 * column-major [V] array.  Dimensions up to 2730 with prime factors <= 7 run
 * as FFTs, other lengths up to 65536 as direct DFTs (MM_ERR_UNSUPPORTED
 * above).  Usable on its own or as the S of mm_sense_create. */
int mm_cart_plan_create(mm_ctx* ctx, int ndim, const int64_t* imN, const float* mask, mm_op** out);

/* opt.sumDims of mrSampFunc (operators/mrSampFunc.m:100-105): forward sums a
 * column-major array of nd <= 7 dimensions `dims` over the dimensions whose
 * sum_flags[d] is nonzero (kept as singletons); adjoint broadcasts back along
 * them, as MATLAB's implicit expansion does in the 'transp' branch.  Trailing
 * repetitions (nrep of mm_op_apply) are kept. */
int mm_sumdims_plan_create(mm_ctx* ctx, int nd, const int64_t* dims, const int* sum_flags, mm_op** out);

/* [AtA,Aty] = mrSampFuncMat(y,Crcvr,b0,sampTimes,phs_spha,phs_conc,phs_grid,
 * imMask) (operators/mrSampFuncMat.m:1-47), the AtA part, for the high-order
 * operator `op` (mm_hio_plan_create without b0mask): over the Vm voxels with
 * mask[v] != 0 (mask [V] or NULL = all, column-major order),
 * AtA[v,w] = G[v,w] * sum_c conj(maps_c(v)) maps_c(w) with G = S'*S and
 * S = exp(1i*single(kbase))/sqrt(Vm) the dense [K x Vm] encoding matrix in
 * single precision, as the reference forms it.  maps: host [V x C] c64;
 * AtA: [Vm x Vm] c64, host or device per on_device; *Vm_out receives Vm
 * (AtA == NULL: size query only).  The Gram product runs on rocBLAS (cherk, lower triangle, mirrored);
 * workspace K*Vm + (host output) Vm*Vm complex values.  Aty is the adjoint
 * of the SENSE operator (mm_op_apply) restricted to the mask. */
int mm_hio_dense_normal(mm_op* op, int64_t C, const float* maps, const unsigned char* mask, void* AtA,
                        int64_t* Vm_out, int on_device);

/* [phs_spha,phs_conc] = interpTrajTime(...) (findDel/interpTrajTime.m:1-27)
 * for one array: out [nq x ncol] = interp1(trajTim, phs, datatime, 'makima',
 * NaN) along the first dimension of phs [L x ncol] (column-major), with
 * trajTim = -delTime + tdwell*(0:L-1); FP64.  Pointers host or device per
 * on_device. */
int mm_interp_traj_time(mm_ctx* ctx, const double* phs, int64_t L, int ncol, double tdwell, double delTime,
                        const double* datatime, int64_t nq, double* out, int on_device);

/* SENSE encoding E = S o R with optional real data weights W ([M x C], used
 * only by mm_cgne_solve and by mm_op_apply_normal). */
int mm_sense_create(mm_ctx* ctx, mm_op* S, mm_op* R, const float* W, mm_op** out);

/* Element counts (complex) per repetition of the operator input/output. */
int mm_op_shape(const mm_op* op, int64_t* n_image, int64_t* n_data);
/* Number of independent slices an operator spans (0 for NULL). */
int mm_op_nslice(const mm_op* op);

/* y = A x (adjoint = 0) or x = A' y (adjoint = 1) for `nrep` repetitions
 * (column-major trailing dimension).  Out of place; c64 interleaved. */
int mm_op_apply(mm_op* op, int adjoint, const void* in, void* out, int64_t nrep,
                int on_device);
/* x_out = A'(W .* (A x)) (the normal operator cgne iterates). */
int mm_op_apply_normal(mm_op* op, const void* in, void* out, int64_t nrep,
                       int on_device);

/* ---- solver -------------------------------------------------------------- */
/* Options of cgne.m:21-89.  A threshold < 0 means "empty" (disabled). */
typedef struct mm_cg_opts {
    double resThresh;        /* absolute |r|^2 (default 1e-10)          */
    double resThreshRel;     /* |r|^2 < resThreshRel*|b|^2, b = E'(W.*y);
                                not in cgne.m: relative-residual helper
                                (default off; combined with resThresh by max) */
    double resChangeThresh;  /* relative change of |r|^2 (default off)  */
    double xChangeThresh;    /* |x - xold|^2 (default 1e-10)            */
    int    expResN;          /* explicit residual period (default 100)  */
    int    stopOnResInc;     /* default 0                               */
    double resIncThresh;     /* default 1                               */
    int    stopOnXdifInc;    /* default 0                               */
    double stopThresh;       /* noise discrepancy |A' n|^2 (< 0: off)   */
    int    nseExtraIt;       /* default 4                               */
    int    x0Fallback;       /* not in cgne.m: with x0 given, a slice whose
                                initial |b - N x0|^2 exceeds |b|^2 (the
                                residual of x0 = 0) starts from x0 = 0
                                instead (default 0)                     */
} mm_cg_opts;

void mm_cg_opts_default(mm_cg_opts* o);

/* cgne on E'(W.*E) x = E'(W.*y) for `nslice` slices at once (nslice must be
 * the slice count of E).  y: data [sum_s M_s*C]; x0: image or NULL (NULL is
 * cgne's default x0 = b); x: image out.  Per slice s the histories
 * resSqAll/xnormAll/xdiffAll hold (nit_s+1) values at offset s*(NitMax+1);
 * nit/status are per slice.  All buffers are host or device per on_device
 * (histories, nit and status are always host). */
int mm_cgne_solve(mm_op* E, const void* y, const void* x0, void* x, int NitMax,
                  const mm_cg_opts* opt, double* resSqAll, double* xnormAll,
                  double* xdiffAll, int* nit, int* status, int on_device);
/* cgne with the Tikhonov terms and the ground-truth history of cgne.m:
 * the normal operator becomes E'(W.*E x) + sum_n gam[n] R[n]'(R[n] x)
 * (cgne.m:171-179, :219-227), R[n] a device operator on E's image (e.g. a
 * coil, Cartesian or NUFFT operator) or NULL for the identity, gam[n] = 0
 * dropping the term (cgne.m:91-93); gtruth (image, host/device per
 * on_device, or NULL) fills mseAll with |gtruth - x_n|^2 per iteration
 * (cgne.m:201-204, :271-274), laid out like resSqAll (zeros without it). */
int mm_cgne_solve_ex(mm_op* E, const void* y, const void* x0, void* x, int NitMax,
                     const mm_cg_opts* opt, int nreg, const double* gam, mm_op* const* R,
                     const void* gtruth, double* resSqAll, double* mseAll, double* xnormAll,
                     double* xdiffAll, int* nit, int* status, int on_device);

/* Measurement hook (no counterpart in the reference): times the operator's
 * dominant kernel with HIP events on the context stream (device pointers
 * only) and returns the mean milliseconds per launch and the kernel's name.
 * adjoint = 0/1: the forward/adjoint kernel of a high-order SENSE operator,
 * launched `iters` times back to back; adjoint = 2: `iters` applications of
 * the normal operator E'(W.*E x) of a NUFFT SENSE operator (x: image, out:
 * image) with an event after every kernel -- the dominant kernel is the one
 * with the largest time per application. */
int mm_op_profile(mm_op* op, int adjoint, const void* in, void* out, int64_t nrep, int iters,
                  double* ms_per_launch, char* kernel_name, int name_len);
/* Stage `idx` of the last adjoint = 2 profile of `op`: its mean
 * milliseconds per launch, launches per application and kernel name;
 * MM_ERR_ARG past the last stage. */
int mm_op_profile_stage(mm_op* op, int idx, double* ms_per_launch, int* launches, char* kernel_name,
                        int name_len);
/* Tuning knob (no counterpart in the reference): "hio_variant" (0..7: bit 0
 * phase on FP64 MFMA, bit 1 hardware sine/cosine, bit 2 the software-pipelined
 * form of variant 3) and "hio_mfma" (0/1) of a
 * high-order operator or a SENSE operator built on one; for a NUFFT operator
 * (or a SENSE operator built on one) "nu_stage_samples" (16..256),
 * "nu_stage_entries" (64..2048), "nu_degrid_block" (16..256),
 * "nu_slice_group" (slices per group of the normal operator, 0: one), and the 0/1
 * switches "nu_pair_cols", "nu_xcd_order", "nu_grid_pairs", "nu_grid_region" (with "nu_xcd_order": the
 * staged gridding items of one spatial region of a trajectory's rectangles on one XCD, for every slice),
 * "nu_snake", "nu_degrid_merge", "nu_grid_stage", "nu_degrid_stage", "nu_degrid_t", "nu_toep_fused" (the
 * Toeplitz form's 2-D middle stage in one pass), "nu_fft_mid" (FFT kernels that hold 32 values per thread
 * for grid sides the default ones cannot run in one LDS buffer: 257..512 at 16 transforms). */
int mm_op_set_option(mm_op* op, const char* key, int value);
/* Current value of a tuning option of mm_op_set_option (no counterpart in
 * the reference). */
int mm_op_get_option(mm_op* op, const char* key, int* value);
/* The project's autotuner (no counterpart in the reference; the batched
 * configuration asks for tile and bin sizes chosen by one): for a SENSE
 * operator on a 2-D NUFFT, times candidate staged-gridding rectangle caps,
 * degridding block sizes and work orders on this device with the normal
 * operator's stage timers (`iters` applications each) and keeps the
 * fastest; the choice can be read back with mm_op_get_option.  For the
 * interpolated high-order operator (mm_hio_interp_create), or a SENSE
 * operator on one, the same for its NUFFT at the batch it runs (images x
 * SVD terms transforms; the SENSE operator's images are its coils).
 * MM_ERR_UNSUPPORTED for other operators. */
int mm_op_autotune(mm_op* op, int iters);
/* The hipStream_t the context runs on. */
void* mm_ctx_stream(mm_ctx* ctx);

/* Enable (1, default) or disable (0) hipGraph capture of the regular cgne
 * iteration for operator E. */
int mm_cg_set_graph(mm_op* E, int use_graph);
/* Per slice (nslice ints): 1 where the last mm_cgne_solve on E replaced x0
 * by 0 (mm_cg_opts.x0Fallback), else 0. */
int mm_cg_last_fallback(mm_op* E, int* fellback);

int mm_op_destroy(mm_op* op);

/* ---- learned warm start (no counterpart in the reference) ---------------- */
/* A 6-layer 3x3 convolutional network (2 -> 32 -> 32 -> 32 -> 32 -> 32 -> 2
 * channels, ReLU between layers, residual skip) on complex N1 x N2 images
 * (column-major, c64), computing x0 = x + s * net(x / s) with
 * s = sqrt(mean |x|^2) per image; bf16 weights, FP32 accumulation.  Its input
 * is the density-compensated adjoint image R'F'(dcf .* y); its output is the
 * x0 passed to mm_cgne_solve.  Parameters are a flat FP32 vector: for each
 * layer W[cout][cin][3][3] then b[cout] (38178 values).  A new network has
 * all parameters zero, i.e. x0 = x. */
int mm_wsnet_create(mm_ctx* ctx, int64_t N1, int64_t N2, mm_wsnet** out);
/* Network input normalisation (no counterpart in the reference; used with
 * the warm start): out = x / (ssq + eps * max(ssq)) per pixel, ssq = sum over
 * coils of |maps|^2 of the coil operator R (mm_rcvr_plan_create*), the
 * maximum over every slice; x and out are image-side vectors of R. */
int mm_rcvr_normalize_ssq(mm_op* R, const void* x, void* out, float eps, int on_device);
/* The same normalisation with no allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code under
 * hipGraph stream capture, serving loops, harnesses that forbid hipMalloc.
 * x and out are device pointers.  Its scratch (per-pixel sums of squares and
 * per-block maxima: n_image + 1024 floats per run of equal-size slices) is a
 * static 16 MiB array per device (4194304 floats); larger operators return
 * MM_ERR_ARG.  Every call on one device shares that array, so calls that
 * overlap on one device (different streams, or host threads) race. */
int mm_rcvr_normalize_ssq_noalloc(mm_op* R, const void* x, void* out, float eps);
int mm_wsnet_num_params(const mm_wsnet* net, int64_t* n);
/* Sets the FP32 master parameters and resets the optimiser state. */
int mm_wsnet_set_params(mm_wsnet* net, const float* theta, int on_device);
int mm_wsnet_get_params(mm_wsnet* net, float* theta, int on_device);
/* y = network(x) for `batch` images [batch][N1*N2]. */
int mm_wsnet_forward(mm_wsnet* net, const void* x, void* y, int64_t batch, int on_device);
/* Training loss mean((x/s + z - t/s)^2) over `batch` (input, target) image
 * pairs and its gradient w.r.t. the parameters (`grad`: host or device per
 * on_device, or NULL; `loss` may be NULL). */
int mm_wsnet_loss_grad(mm_wsnet* net, const void* x, const void* target, int64_t batch,
                       double* loss, float* grad, int on_device);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on that gradient.  With
 * loss == NULL the call does not synchronise with the device. */
int mm_wsnet_train_step(mm_wsnet* net, const void* x, const void* target, int64_t batch,
                        float lr, double* loss, int on_device);
int mm_wsnet_destroy(mm_wsnet* net);

#ifdef __cplusplus
}
#endif
#endif /* MATMRI_HIP_H */

// Interpolated expanded-encoding operator (sampHighOrder approach 1,
// operators/@sampHighOrder/sampHighOrder.m:314-359): the linear phase terms
// are a NUFFT, the rest a truncated SVD exp(1i*Phi_res) ~ U S V' of Ns terms,
//   forward  y = phsShft .* sum_n svdTime_n .* NUFFT(x .* svdSpace_n)
//   adjoint  x = sum_n conj(svdSpace_n) .* NUFFT'(conj(phsShft .* svdTime_n) .* y)
// The Ns modulated images of every input image go through the NUFFT as one
// batch of R*Ns transforms (image r, term n at transform r + R*n); the
// spatial modulation and the time-domain combination are one kernel each.
#include "hio_interp.hpp"

namespace mm {

// imgs[v, r + R n] = x[v, r] * sp[v, n].  Transform b = r + R n on the
// grid's y dimension (stride gridDim.y), voxels on x: the index split is one
// 32-bit division per transform, not a 64-bit one per element (900 -> 739 us
// at V = 220^2, 5824 transforms: the 2.25 GB of stores at 3.0 TB/s).
__global__ __launch_bounds__(256) void hin_modulate_k(const float2* __restrict__ x, const float2* __restrict__ sp,
                                                      float2* __restrict__ imgs, int64_t V, int R, int Ns) {
    const int nb = R * Ns;
    for (int b = blockIdx.y; b < nb; b += gridDim.y) {
        const int r = b % R, t = b / R;
        const float2* xr = x + (int64_t)r * V;
        const float2* st = sp + (int64_t)t * V;
        float2* o = imgs + (int64_t)b * V;
        for (int64_t v = blockIdx.x * (int64_t)256 + threadIdx.x; v < V; v += (int64_t)gridDim.x * 256)
            o[v] = cmulf(xr[v], st[v]);
    }
}

// y[k, r] = ph[k] * sum_n st[k, n] * Y[k, r + R n]
__global__ __launch_bounds__(256) void hin_combine_k(const float2* __restrict__ Y, const float2* __restrict__ st,
                                                     const float2* __restrict__ ph, float2* __restrict__ y, int64_t K,
                                                     int64_t R, int Ns) {
    const int64_t n = K * R;
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int64_t k = i % K, r = i / K;
        float2 s = make_float2(0.f, 0.f);
        for (int t = 0; t < Ns; ++t) {
            const float2 p = cmulf(st[(int64_t)t * K + k], Y[(r + R * t) * K + k]);
            s.x += p.x;
            s.y += p.y;
        }
        y[i] = cmulf(ph[k], s);
    }
}

// Y[k, r + R n] = conj(ph[k] * st[k, n]) * y[k, r] (a flat grid-stride
// loop measured faster here than the per-transform launch of hin_modulate_k:
// 378 against 407 us at K = 24000, 5824 transforms)
__global__ __launch_bounds__(256) void hin_spread_k(const float2* __restrict__ y, const float2* __restrict__ st,
                                                    const float2* __restrict__ ph, float2* __restrict__ Y, int64_t K,
                                                    int64_t R, int Ns) {
    const int64_t n = K * R * Ns;
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int64_t k = i % K, b = i / K;
        const int64_t r = b % R, t = b / R;
        const float2 w = cmulf(ph[k], st[t * K + k]);
        const float2 v = y[r * K + k];
        Y[i] = make_float2(w.x * v.x + w.y * v.y, w.x * v.y - w.y * v.x);   // conj(w) * v
    }
}

// x[v, r] = sum_n conj(sp[v, n]) * imgs[v, r + R n]
__global__ __launch_bounds__(256) void hin_demodulate_k(const float2* __restrict__ imgs, const float2* __restrict__ sp,
                                                        float2* __restrict__ x, int64_t V, int64_t R, int Ns) {
    const int64_t n = V * R;
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int64_t v = i % V, r = i / V;
        float2 s = make_float2(0.f, 0.f);
        for (int t = 0; t < Ns; ++t) {
            const float2 a = sp[(int64_t)t * V + v], b = imgs[(r + R * t) * V + v];
            s.x += a.x * b.x + a.y * b.y;
            s.y += a.x * b.y - a.y * b.x;
        }
        x[i] = s;
    }
}

static int hin_grid(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536)); }
// (voxel or sample blocks, transforms) of the per-transform kernels: two
// elements per thread along a transform, at most 512 rows of transforms (each
// block then loops over every 512th), so blocks live long enough to keep
// their stores streaming
static dim3 hin_grid2(int64_t n, int64_t nb) {
    return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 511) / 512, 65535)),
                (unsigned)std::max<int64_t>(1, std::min<int64_t>(nb, 512)));
}

HioInterpOp::HioInterpOp(mm_ctx* c, mm_op* nufft, int64_t V, int64_t K, int Ns, const float* svdSpace,
                         const float* svdTime, const float* phsShft)
    : nu_(nufft), V_(V), K_(K), Ns_(Ns) {
    ctx = c;
    if (!nufft || !svdSpace || !svdTime || !phsShft) fail(MM_ERR_ARG, "sampHighOrder: NULL argument");
    if (Ns < 1 || V < 1 || K < 1) fail(MM_ERR_ARG, "sampHighOrder: empty interpolation terms");
    if (nufft->nslice() != 1 || nufft->n_image() != V || nufft->n_data() != K)
        fail(MM_ERR_ARG, "sampHighOrder: the NUFFT does not map the image to the samples");
    sp_.upload(reinterpret_cast<const float2*>(svdSpace), (size_t)(V * Ns), ctx->stream);
    st_.upload(reinterpret_cast<const float2*>(svdTime), (size_t)(K * Ns), ctx->stream);
    ph_.upload(reinterpret_cast<const float2*>(phsShft), (size_t)K, ctx->stream);
    HIPCHECK(hipStreamSynchronize(ctx->stream));
}

bool HioInterpOp::profile_dominant(int adjoint, const float2* in, float2* out, int64_t nrep, int iters, double* ms,
                                   std::string* name) {
    if ((adjoint != 0 && adjoint != 1) || nrep < 1 || iters < 1) return false;
    hipStream_t st = ctx->stream;
    const int64_t R = nrep;
    apply(adjoint, in, out, nrep);   // sizes the buffers, first-launch costs
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    auto time_of = [&](auto launch) {
        HIPCHECK(hipEventRecord(e0, st));
        for (int i = 0; i < iters; ++i) launch();
        HIPCHECK(hipEventRecord(e1, st));
        HIPCHECK(hipEventSynchronize(e1));
        float t = 0.f;
        HIPCHECK(hipEventElapsedTime(&t, e0, e1));
        return (double)t / iters;
    };
    double t_in, t_out;
    if (!adjoint) {
        t_in = time_of([&] {
            hipLaunchKernelGGL(hin_modulate_k, hin_grid2(V_, R * Ns_), dim3(256), 0, st, in, sp_.p, imgs_.p, V_,
                               (int)R, Ns_);
        });
        t_out = time_of([&] {
            hipLaunchKernelGGL(hin_combine_k, dim3(hin_grid(K_ * R)), dim3(256), 0, st, dat_.p, st_.p, ph_.p, out, K_,
                               R, Ns_);
        });
    } else {
        t_in = time_of([&] {
            hipLaunchKernelGGL(hin_spread_k, dim3(hin_grid(K_ * R * Ns_)), dim3(256), 0, st, in, st_.p, ph_.p, dat_.p,
                               K_, R, Ns_);
        });
        t_out = time_of([&] {
            hipLaunchKernelGGL(hin_demodulate_k, dim3(hin_grid(V_ * R)), dim3(256), 0, st, imgs_.p, sp_.p, out, V_,
                               R, Ns_);
        });
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    LAUNCHCHECK();
    double t_nu = 0;
    std::string nu_name;
    const bool have_nu = adjoint ? nu_->profile_dominant(1, dat_.p, imgs_.p, R * Ns_, iters, &t_nu, &nu_name)
                                 : nu_->profile_dominant(0, imgs_.p, dat_.p, R * Ns_, iters, &t_nu, &nu_name);
    *ms = t_in;
    *name = adjoint ? "hin_spread_k" : "hin_modulate_k";
    if (t_out > *ms) {
        *ms = t_out;
        *name = adjoint ? "hin_demodulate_k" : "hin_combine_k";
    }
    if (have_nu && t_nu > *ms) {
        *ms = t_nu;
        *name = nu_name;
    }
    return true;
}

void HioInterpOp::apply(int adjoint, const float2* in, float2* out, int64_t nrep) {
    if (nrep < 1) fail(MM_ERR_ARG, "sampHighOrder: nrep must be >= 1");
    if (nrep * Ns_ >= INT32_MAX) fail(MM_ERR_UNSUPPORTED, "sampHighOrder: too many images x SVD terms");
    hipStream_t st = ctx->stream;
    const int64_t R = nrep;
    imgs_.ensure((size_t)(V_ * R * Ns_));
    dat_.ensure((size_t)(K_ * R * Ns_));
    if (!adjoint) {
        hipLaunchKernelGGL(hin_modulate_k, hin_grid2(V_, R * Ns_), dim3(256), 0, st, in, sp_.p, imgs_.p, V_, (int)R,
                           Ns_);
        LAUNCHCHECK();
        nu_->apply(0, imgs_.p, dat_.p, R * Ns_);
        hipLaunchKernelGGL(hin_combine_k, dim3(hin_grid(K_ * R)), dim3(256), 0, st, dat_.p, st_.p, ph_.p, out, K_, R,
                           Ns_);
        LAUNCHCHECK();
    } else {
        hipLaunchKernelGGL(hin_spread_k, dim3(hin_grid(K_ * R * Ns_)), dim3(256), 0, st, in, st_.p, ph_.p, dat_.p, K_,
                           R, Ns_);
        LAUNCHCHECK();
        nu_->apply(1, dat_.p, imgs_.p, R * Ns_);
        hipLaunchKernelGGL(hin_demodulate_k, dim3(hin_grid(V_ * R)), dim3(256), 0, st, imgs_.p, sp_.p, out, V_, R,
                           Ns_);
        LAUNCHCHECK();
    }
}

}  // namespace mm

// extern "C" boundary of libmatmri_hip (declared in include/matmri_hip.h).
// Every entry point converts exceptions into MM_ERR_* codes and keeps the
// message for mm_last_error().
#include <cstring>
#include <mutex>

#include "cart.hpp"
#include "cg.hpp"
#include "hio.hpp"
#include "hio_interp.hpp"
#include "nufft3.hpp"
#include "ops.hpp"
#include "traj.hpp"
#include "wsnet.hpp"

using namespace mm;

namespace {
thread_local std::string g_err;

template <class F>
int guarded(F&& f) {
    try {
        g_err.clear();
        f();
        return MM_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return MM_ERR_OOM;
    } catch (const std::exception& e) {
        g_err = e.what();
        return MM_ERR_HIP;
    }
}

void set_device(mm_ctx* ctx) { HIPCHECK(hipSetDevice(ctx->device)); }

struct SolverHolder {
    mm_op* E;
    std::unique_ptr<CgSolver> s;
};
// one cached solver per operator (its buffers and captured graph are reused)
std::map<mm_op*, std::unique_ptr<CgSolver>>& solvers() {
    static std::map<mm_op*, std::unique_ptr<CgSolver>> m;
    return m;
}

}  // namespace

extern "C" {

const char* mm_last_error(void) { return g_err.c_str(); }

const char* mm_version(void) { return "matmri-hip 0.1 (gfx950)"; }

int mm_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int mm_ctx_create(int device, void* stream, mm_ctx** out) {
    return guarded([&] {
        if (!out) fail(MM_ERR_ARG, "out is NULL");
        *out = nullptr;
        const int n = mm_device_count();
        if (n <= 0) fail(MM_ERR_DEVICE, "no HIP device visible");
        if (device < 0 || device >= n) fail(MM_ERR_DEVICE, "device index out of range");
        HIPCHECK(hipSetDevice(device));
        hipDeviceProp_t prop;
        HIPCHECK(hipGetDeviceProperties(&prop, device));
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            fail(MM_ERR_DEVICE, std::string("libmatmri_hip is built for gfx950, device is ") + prop.gcnArchName);
        auto* c = new mm_ctx();
        c->device = device;
        if (stream) {
            c->stream = (hipStream_t)stream;
            c->own_stream = false;
        } else {
            HIPCHECK(hipStreamCreate(&c->stream));
            c->own_stream = true;
        }
        *out = c;
    });
}

int mm_ctx_destroy(mm_ctx* ctx) {
    return guarded([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        blas_handle_destroy(ctx->blas);
        if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

int mm_ctx_sync(mm_ctx* ctx) {
    return guarded([&] {
        if (!ctx) fail(MM_ERR_ARG, "ctx is NULL");
        HIPCHECK(hipStreamSynchronize(ctx->stream));
    });
}

int mm_nufft_plan_create(mm_ctx* ctx, int ndim, int nslice, const int64_t* imN, const int64_t* M,
                         const double* const* kloc, const double* const* dcf, double os, int kwidth,
                         mm_op** out) {
    return guarded([&] {
        if (!ctx || !out || !imN || !M || !kloc) fail(MM_ERR_ARG, "nufftOp: NULL argument");
        set_device(ctx);
        if (ndim == 3) {
            if (nslice != 1) fail(MM_ERR_UNSUPPORTED, "nufftOp: 3-D plans hold one volume");
            *out = new Nufft3Op(ctx, imN, M[0], kloc[0], dcf ? dcf[0] : nullptr, os, kwidth);
            return;
        }
        *out = new NufftOp(ctx, ndim, nslice, imN, M, kloc, dcf, os, kwidth);
    });
}

int mm_rcvr_plan_create(mm_ctx* ctx, int64_t V, int64_t C, const float* maps, mm_op** out) {
    return guarded([&] {
        if (!ctx || !out) fail(MM_ERR_ARG, "rcvrOp: NULL argument");
        set_device(ctx);
        *out = new RcvrOp(ctx, 1, &V, C, maps);
    });
}

int mm_rcvr_plan_create_multi(mm_ctx* ctx, int nslice, const int64_t* V, int64_t C,
                              const float* maps, mm_op** out) {
    return guarded([&] {
        if (!ctx || !out || !V) fail(MM_ERR_ARG, "rcvrOp: NULL argument");
        set_device(ctx);
        *out = new RcvrOp(ctx, nslice, V, C, maps);
    });
}

int mm_rcvr_normalize_ssq(mm_op* R, const void* x, void* out, float eps, int on_device) {
    return guarded([&] {
        if (!R || !x || !out) fail(MM_ERR_ARG, "rcvrOp: NULL argument");
        auto* rc = dynamic_cast<RcvrOp*>(R);
        if (!rc) fail(MM_ERR_ARG, "rcvrOp: not a coil operator");
        if (!(eps >= 0.f)) fail(MM_ERR_ARG, "rcvrOp: eps must be >= 0");
        mm_ctx* ctx = rc->ctx;
        set_device(ctx);
        const int64_t n = rc->n_image();
        if (on_device) {
            rc->normalize_by_coil_ssq((const float2*)x, (float2*)out, eps);
            return;
        }
        DevBuf<float2> dx, dy;
        dx.upload((const float2*)x, (size_t)n, ctx->stream);
        dy.alloc((size_t)n);
        rc->normalize_by_coil_ssq(dx.p, dy.p, eps);
        HIPCHECK(hipMemcpyAsync(out, dy.p, sizeof(float2) * n, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
    });
}

}  // extern "C"

namespace {
// Static scratch of mm_rcvr_normalize_ssq_noalloc: per-pixel coil sums of
// squares and per-block maxima, kNoAllocScratchFloats floats (16 MiB) on
// every device the library runs on.  Host code uses only the per-device
// addresses from hipGetSymbolAddress, never the symbol itself.
constexpr int64_t kNoAllocScratchFloats = int64_t(1) << 22;
__device__ __align__(256) float scratch[kNoAllocScratchFloats];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on `device` (the current device),
// looked up once per device
float* scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) fail(MM_ERR_DEVICE, "no-alloc scratch: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        float* p = nullptr;
        HIPCHECK(hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}
}  // namespace

extern "C" {

int mm_rcvr_normalize_ssq_noalloc(mm_op* R, const void* x, void* out, float eps) {
    return guarded([&] {
        if (!R || !x || !out) fail(MM_ERR_ARG, "rcvrOp: NULL argument");
        auto* rc = dynamic_cast<RcvrOp*>(R);
        if (!rc) fail(MM_ERR_ARG, "rcvrOp: not a coil operator");
        if (!(eps >= 0.f)) fail(MM_ERR_ARG, "rcvrOp: eps must be >= 0");
        const int64_t need = rc->ssq_scratch_floats();
        if (need > kNoAllocScratchFloats)
            fail(MM_ERR_ARG, "rcvrOp: no-alloc normalisation needs " + std::to_string(need) +
                                 " scratch floats, above its static maximum of " +
                                 std::to_string(kNoAllocScratchFloats) + " (use mm_rcvr_normalize_ssq)");
        set_device(rc->ctx);
        float* sp = scratch_on(rc->ctx->device);
        rc->normalize_by_coil_ssq_with((const float2*)x, (float2*)out, eps, sp, sp + rc->n_image());
    });
}

int mm_hio_plan_create(mm_ctx* ctx, int64_t V, int64_t K, int Lh, int Lc, const double* x,
                       const double* y, const double* z, const double* b0, const double* t,
                       const double* h, const double* g, const double* mask, mm_op** out) {
    return guarded([&] {
        if (!ctx || !out) fail(MM_ERR_ARG, "sampHighOrder: NULL argument");
        set_device(ctx);
        *out = new HioOp(ctx, V, K, Lh, Lc, x, y, z, b0, t, h, g, mask);
    });
}

int mm_cart_plan_create(mm_ctx* ctx, int ndim, const int64_t* imN, const float* mask, mm_op** out) {
    return guarded([&] {
        if (!ctx || !out || !imN || !mask) fail(MM_ERR_ARG, "mrSampFunc: NULL argument");
        *out = nullptr;
        set_device(ctx);
        *out = new CartOp(ctx, ndim, imN, mask);
    });
}

int mm_sumdims_plan_create(mm_ctx* ctx, int nd, const int64_t* dims, const int* sum_flags, mm_op** out) {
    return guarded([&] {
        if (!ctx || !out) fail(MM_ERR_ARG, "mrSampFunc: NULL argument");
        *out = nullptr;
        set_device(ctx);
        *out = new SumDimsOp(ctx, nd, dims, sum_flags);
    });
}

int mm_interp_traj_time(mm_ctx* ctx, const double* phs, int64_t L, int ncol, double tdwell, double delTime,
                        const double* datatime, int64_t nq, double* out, int on_device) {
    return guarded([&] {
        if (!ctx || !phs || !datatime || !out) fail(MM_ERR_ARG, "interpTrajTime: NULL argument");
        set_device(ctx);
        if (on_device) {
            interp_traj_time(ctx->stream, phs, L, ncol, tdwell, delTime, datatime, nq, out);
            return;
        }
        if (L < 1 || ncol < 1 || nq < 0) fail(MM_ERR_ARG, "interpTrajTime: bad sizes");
        DevBuf<double> dp, dt, dout;
        dp.upload(phs, (size_t)L * ncol, ctx->stream);
        dt.upload(datatime, (size_t)nq, ctx->stream);
        dout.alloc((size_t)nq * ncol);
        interp_traj_time(ctx->stream, dp.p, L, ncol, tdwell, delTime, dt.p, nq, dout.p);
        if (nq) HIPCHECK(hipMemcpyAsync(out, dout.p, sizeof(double) * nq * ncol, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
    });
}

int mm_hio_dense_normal(mm_op* op, int64_t C, const float* maps, const unsigned char* mask, void* AtA,
                        int64_t* Vm_out, int on_device) {
    return guarded([&] {
        auto* h = dynamic_cast<HioOp*>(op);
        if (!h) fail(MM_ERR_ARG, "mrSampFuncMat: not a high-order operator");
        if (!maps || C < 1) fail(MM_ERR_ARG, "mrSampFuncMat: NULL maps");
        mm_ctx* ctx = op->ctx;
        set_device(ctx);
        const int64_t V = h->n_image(), K = h->samples();
        std::vector<int> vidx;
        for (int64_t v = 0; v < V; ++v)
            if (!mask || mask[v]) vidx.push_back((int)v);
        const int64_t Vm = (int64_t)vidx.size();
        if (Vm_out) *Vm_out = Vm;
        if (!AtA) return;   // size query
        if (Vm == 0) fail(MM_ERR_ARG, "mrSampFuncMat: empty image mask");
        // masked maps [Vm x C]
        std::vector<float2> mm((size_t)Vm * C);
        const float2* mp = reinterpret_cast<const float2*>(maps);
        for (int64_t c = 0; c < C; ++c)
            for (int64_t j = 0; j < Vm; ++j) mm[(size_t)(c * Vm + j)] = mp[c * V + vidx[(size_t)j]];
        if (!ctx->blas) ctx->blas = blas_handle_create();
        hipStream_t st = ctx->stream;
        DevBuf<int> dv;
        DevBuf<float2> dm, S, G;
        dv.upload(vidx, st);
        dm.upload(mm, st);
        S.alloc((size_t)(K * Vm));
        float2* out = (float2*)AtA;
        if (!on_device) {
            G.alloc((size_t)(Vm * Vm));
            out = G.p;
        }
        h->dense_normal(dv.p, Vm, dm.p, (int)C, S.p, out, ctx->blas);
        if (!on_device)
            HIPCHECK(hipMemcpyAsync(AtA, out, sizeof(float2) * Vm * Vm, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    });
}

int mm_hio_interp_create(mm_ctx* ctx, mm_op* nufft, int64_t V, int64_t K, int Ns, const float* svdSpace,
                         const float* svdTime, const float* phsShft, mm_op** out) {
    return guarded([&] {
        if (!ctx || !out) fail(MM_ERR_ARG, "sampHighOrder: NULL argument");
        *out = nullptr;
        set_device(ctx);
        *out = new HioInterpOp(ctx, nufft, V, K, Ns, svdSpace, svdTime, phsShft);
    });
}

int mm_hio_set_phi_div(mm_op* op, const double* hd, const double* gd) {
    return guarded([&] {
        auto* h = dynamic_cast<HioOp*>(op);
        if (!h) fail(MM_ERR_ARG, "setPhiDiv: not a high-order operator");
        set_device(op->ctx);
        h->set_phi_div(hd, gd);
    });
}

int mm_nufft_set_toeplitz(mm_op* op, const float* psf) {
    return guarded([&] {
        set_device(op->ctx);
        if (auto* n3 = dynamic_cast<Nufft3Op*>(op)) {
            n3->set_toeplitz(reinterpret_cast<const float2*>(psf));
            return;
        }
        auto* n = dynamic_cast<NufftOp*>(op);
        if (!n) fail(MM_ERR_ARG, "prepToep: not a NUFFT operator");
        n->set_toeplitz(reinterpret_cast<const float2*>(psf));
    });
}

int mm_sense_create(mm_ctx* ctx, mm_op* S, mm_op* R, const float* W, mm_op** out) {
    return guarded([&] {
        if (!ctx || !out) fail(MM_ERR_ARG, "mrSampFunc: NULL argument");
        set_device(ctx);
        *out = new SenseOp(ctx, S, R, W);
    });
}

int mm_op_shape(const mm_op* op, int64_t* n_image, int64_t* n_data) {
    return guarded([&] {
        if (!op) fail(MM_ERR_ARG, "op is NULL");
        if (n_image) *n_image = op->n_image();
        if (n_data) *n_data = op->n_data();
    });
}

int mm_op_nslice(const mm_op* op) { return op ? op->nslice() : 0; }

static void apply_common(mm_op* op, int mode, const void* in, void* out, int64_t nrep, int on_device) {
    if (!op || !in || !out) fail(MM_ERR_ARG, "mtimes: NULL argument");
    if (nrep < 1) fail(MM_ERR_ARG, "mtimes: nrep must be >= 1");
    if (mode == 2 && op->W && nrep != 1) fail(MM_ERR_ARG, "normal operator with weights needs nrep = 1");
    mm_ctx* ctx = op->ctx;
    set_device(ctx);
    const int64_t ni = op->n_image() * nrep, nd = op->n_data() * nrep;
    const int64_t nin = mode == 1 ? nd : ni;
    const int64_t nout = mode == 0 ? nd : ni;
    const float2* din = (const float2*)in;
    float2* dout = (float2*)out;
    if (!on_device) {
        ctx->stage_in.upload((const float2*)in, (size_t)nin, ctx->stream);
        ctx->stage_out.ensure((size_t)nout);
        din = ctx->stage_in.p;
        dout = ctx->stage_out.p;
    }
    if (mode == 2) {
        ctx->stage_aux.ensure((size_t)nd);
        op->apply_normal(din, dout, nrep, ctx->stage_aux.p);
    } else {
        op->apply(mode, din, dout, nrep);
    }
    if (!on_device) {
        HIPCHECK(hipMemcpyAsync(out, dout, sizeof(float2) * nout, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
    }
}

int mm_op_apply(mm_op* op, int adjoint, const void* in, void* out, int64_t nrep, int on_device) {
    return guarded([&] { apply_common(op, adjoint ? 1 : 0, in, out, nrep, on_device); });
}

int mm_op_apply_normal(mm_op* op, const void* in, void* out, int64_t nrep, int on_device) {
    return guarded([&] { apply_common(op, 2, in, out, nrep, on_device); });
}

void mm_cg_opts_default(mm_cg_opts* o) {
    if (!o) return;
    o->resThresh = 1e-10;
    o->resThreshRel = -1;
    o->resChangeThresh = -1;
    o->xChangeThresh = 1e-10;
    o->expResN = 100;
    o->stopOnResInc = 0;
    o->resIncThresh = 1;
    o->stopOnXdifInc = 0;
    o->stopThresh = -1;
    o->nseExtraIt = 4;
    o->x0Fallback = 0;
}

int mm_cgne_solve_ex(mm_op* E, const void* y, const void* x0, void* x, int NitMax, const mm_cg_opts* opt,
                     int nreg, const double* gam, mm_op* const* R, const void* gtruth, double* resSqAll,
                     double* mseAll, double* xnormAll, double* xdiffAll, int* nit, int* status, int on_device) {
    return guarded([&] {
        if (!E || !y || !x) fail(MM_ERR_ARG, "cgne: NULL argument");
        if (nreg < 0 || (nreg > 0 && !gam)) fail(MM_ERR_ARG, "cgne: nreg > 0 needs gam");
        if (auto* se = dynamic_cast<SenseOp*>(E); se && se->toeplitz())
            fail(MM_ERR_UNSUPPORTED,
                 "cgne: the SENSE operator is on a Toeplitz-form NUFFT (prepToep), whose forward maps images "
                 "to coil images; solve with the gridding-form NUFFT");
        mm_cg_opts o;
        mm_cg_opts_default(&o);
        if (opt) o = *opt;
        std::vector<std::pair<double, mm_op*>> regs;
        for (int i = 0; i < nreg; ++i) {
            mm_op* r = R ? R[i] : nullptr;
            if (r && r->ctx->device != E->ctx->device) fail(MM_ERR_ARG, "cgne: R on another device");
            if (gam[i] != 0.0) regs.emplace_back(gam[i], r);   // gam == 0 drops the term (cgne.m:91-93)
        }
        mm_ctx* ctx = E->ctx;
        set_device(ctx);
        auto& m = solvers();
        auto it = m.find(E);
        if (it == m.end()) it = m.emplace(E, std::make_unique<CgSolver>(E)).first;
        CgSolver& s = *it->second;
        const int64_t ni = E->n_image(), nd = E->n_data();
        const float2 *dy = (const float2*)y, *dx0 = (const float2*)x0, *dgt = (const float2*)gtruth;
        float2* dx = (float2*)x;
        DevBuf<float2> hy, hx0, hx, hgt;
        if (!on_device) {
            hy.upload((const float2*)y, (size_t)nd, ctx->stream);
            dy = hy.p;
            if (x0) {
                hx0.upload((const float2*)x0, (size_t)ni, ctx->stream);
                dx0 = hx0.p;
            }
            if (gtruth) {
                hgt.upload((const float2*)gtruth, (size_t)ni, ctx->stream);
                dgt = hgt.p;
            }
            hx.alloc((size_t)ni);
            dx = hx.p;
        }
        s.solve(dy, dx0, dx, NitMax, o, resSqAll, xnormAll, xdiffAll, nit, status, regs, dgt, mseAll);
        if (!on_device) {
            HIPCHECK(hipMemcpyAsync(x, dx, sizeof(float2) * ni, hipMemcpyDeviceToHost, ctx->stream));
            HIPCHECK(hipStreamSynchronize(ctx->stream));
        }
    });
}

int mm_cgne_solve(mm_op* E, const void* y, const void* x0, void* x, int NitMax, const mm_cg_opts* opt,
                  double* resSqAll, double* xnormAll, double* xdiffAll, int* nit, int* status,
                  int on_device) {
    return mm_cgne_solve_ex(E, y, x0, x, NitMax, opt, 0, nullptr, nullptr, nullptr, resSqAll, nullptr, xnormAll,
                            xdiffAll, nit, status, on_device);
}

int mm_cg_set_graph(mm_op* E, int use_graph) {
    return guarded([&] {
        if (!E) fail(MM_ERR_ARG, "op is NULL");
        auto& m = solvers();
        auto it = m.find(E);
        if (it == m.end()) it = m.emplace(E, std::make_unique<CgSolver>(E)).first;
        it->second->use_graph_ = use_graph != 0;
    });
}

int mm_cg_last_fallback(mm_op* E, int* fellback) {
    return guarded([&] {
        if (!E || !fellback) fail(MM_ERR_ARG, "cgne: NULL argument");
        auto& m = solvers();
        auto it = m.find(E);
        const int ns = E->nslice();
        for (int s = 0; s < ns; ++s) fellback[s] = 0;
        if (it == m.end()) return;
        const auto& v = it->second->last_fallback_;
        for (int s = 0; s < ns && s < (int)v.size(); ++s) fellback[s] = v[s];
    });
}

int mm_op_profile(mm_op* op, int adjoint, const void* in, void* out, int64_t nrep, int iters,
                  double* ms_per_launch, char* kernel_name, int name_len) {
    return guarded([&] {
        if (!op || !in || !out || !ms_per_launch || iters < 1) fail(MM_ERR_ARG, "profile: bad argument");
        set_device(op->ctx);
        std::string name;
        if (!op->profile_dominant(adjoint, (const float2*)in, (float2*)out, nrep, iters, ms_per_launch, &name))
            fail(MM_ERR_UNSUPPORTED, "profile: operator has no dominant-kernel timer");
        if (kernel_name && name_len > 0) {
            std::strncpy(kernel_name, name.c_str(), (size_t)name_len - 1);
            kernel_name[name_len - 1] = 0;
        }
    });
}

int mm_op_profile_stage(mm_op* op, int idx, double* ms_per_launch, int* launches, char* kernel_name,
                        int name_len) {
    return guarded([&] {
        auto* se = dynamic_cast<SenseOp*>(op);
        if (!se) fail(MM_ERR_ARG, "profile: not a SENSE operator");
        if (idx < 0 || idx >= (int)se->stages_.size()) fail(MM_ERR_ARG, "profile: no such stage");
        const auto& e = se->stages_[(size_t)idx];
        if (ms_per_launch) *ms_per_launch = e.second.first;
        if (launches) *launches = e.second.second;
        if (kernel_name && name_len > 0) {
            std::strncpy(kernel_name, e.first.c_str(), (size_t)name_len - 1);
            kernel_name[name_len - 1] = 0;
        }
    });
}

int mm_op_set_option(mm_op* op, const char* key, int value) {
    return guarded([&] {
        if (!op || !key) fail(MM_ERR_ARG, "set_option: NULL argument");
        if (!op->set_option(key, value)) fail(MM_ERR_ARG, std::string("set_option: unknown key ") + key);
        // captured solver graphs hold the old kernels
        solvers().erase(op);
    });
}

int mm_op_get_option(mm_op* op, const char* key, int* value) {
    return guarded([&] {
        if (!op || !key || !value) fail(MM_ERR_ARG, "get_option: NULL argument");
        if (!op->get_option(key, value)) fail(MM_ERR_ARG, std::string("get_option: unknown key ") + key);
    });
}

int mm_op_autotune(mm_op* op, int iters) {
    return guarded([&] {
        if (!op) fail(MM_ERR_ARG, "autotune: NULL operator");
        (void)hipSetDevice(op->ctx->device);
        if (!op->autotune(iters)) fail(MM_ERR_UNSUPPORTED, "autotune: nothing to tune for this operator");
        solvers().erase(op);
    });
}

void* mm_ctx_stream(mm_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int mm_op_destroy(mm_op* op) {
    return guarded([&] {
        if (!op) return;
        (void)hipSetDevice(op->ctx->device);
        (void)hipStreamSynchronize(op->ctx->stream);
        solvers().erase(op);
        delete op;
    });
}

// ---- learned warm start ----------------------------------------------------

int mm_wsnet_create(mm_ctx* ctx, int64_t N1, int64_t N2, mm_wsnet** out) {
    return guarded([&] {
        if (!ctx || !out) fail(MM_ERR_ARG, "warm-start net: NULL argument");
        *out = nullptr;
        set_device(ctx);
        *out = new mm_wsnet(ctx, N1, N2);
    });
}

int mm_wsnet_num_params(const mm_wsnet* net, int64_t* n) {
    return guarded([&] {
        if (!net || !n) fail(MM_ERR_ARG, "warm-start net: NULL argument");
        *n = net->num_params();
    });
}

int mm_wsnet_set_params(mm_wsnet* net, const float* theta, int on_device) {
    return guarded([&] {
        if (!net || !theta) fail(MM_ERR_ARG, "warm-start net: NULL argument");
        set_device(net->ctx);
        net->set_params(theta, on_device != 0);
    });
}

int mm_wsnet_get_params(mm_wsnet* net, float* theta, int on_device) {
    return guarded([&] {
        if (!net || !theta) fail(MM_ERR_ARG, "warm-start net: NULL argument");
        set_device(net->ctx);
        net->get_params(theta, on_device != 0);
    });
}

namespace {
// host-pointer calls: stage n_img images per pointer in device memory
struct WsStage {
    DevBuf<float2> x, t, y;
};
}  // namespace

int mm_wsnet_forward(mm_wsnet* net, const void* x, void* y, int64_t batch, int on_device) {
    return guarded([&] {
        if (!net || !x || !y) fail(MM_ERR_ARG, "warm-start net: NULL argument");
        if (batch < 1) fail(MM_ERR_ARG, "warm-start net: batch must be >= 1");
        mm_ctx* ctx = net->ctx;
        set_device(ctx);
        int64_t ni = 0;
        const float2* dx = (const float2*)x;
        float2* dy = (float2*)y;
        WsStage st;
        if (!on_device) {
            ni = batch * net->n_pixels();
            st.x.upload((const float2*)x, (size_t)ni, ctx->stream);
            st.y.alloc((size_t)ni);
            dx = st.x.p;
            dy = st.y.p;
        }
        net->forward(dx, dy, batch);
        if (!on_device) {
            HIPCHECK(hipMemcpyAsync(y, dy, sizeof(float2) * ni, hipMemcpyDeviceToHost, ctx->stream));
            HIPCHECK(hipStreamSynchronize(ctx->stream));
        }
    });
}

int mm_wsnet_loss_grad(mm_wsnet* net, const void* x, const void* target, int64_t batch, double* loss,
                       float* grad, int on_device) {
    return guarded([&] {
        if (!net || !x || !target) fail(MM_ERR_ARG, "warm-start net: NULL argument");
        if (batch < 1) fail(MM_ERR_ARG, "warm-start net: batch must be >= 1");
        mm_ctx* ctx = net->ctx;
        set_device(ctx);
        const float2* dx = (const float2*)x;
        const float2* dt = (const float2*)target;
        WsStage st;
        if (!on_device) {
            const int64_t ni = batch * net->n_pixels();
            st.x.upload((const float2*)x, (size_t)ni, ctx->stream);
            st.t.upload((const float2*)target, (size_t)ni, ctx->stream);
            dx = st.x.p;
            dt = st.t.p;
        }
        const double L = net->loss_grad(dx, dt, batch, loss != nullptr || !on_device);
        if (loss) *loss = L;
        if (grad) {
            HIPCHECK(hipMemcpyAsync(grad, net->grad(), sizeof(float) * net->num_params(),
                                    on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx->stream));
        }
        if (!on_device) HIPCHECK(hipStreamSynchronize(ctx->stream));
    });
}

int mm_wsnet_train_step(mm_wsnet* net, const void* x, const void* target, int64_t batch, float lr,
                        double* loss, int on_device) {
    return guarded([&] {
        if (!net || !x || !target) fail(MM_ERR_ARG, "warm-start net: NULL argument");
        if (batch < 1) fail(MM_ERR_ARG, "warm-start net: batch must be >= 1");
        if (!(lr > 0.f)) fail(MM_ERR_ARG, "warm-start net: learning rate must be > 0");
        mm_ctx* ctx = net->ctx;
        set_device(ctx);
        const float2* dx = (const float2*)x;
        const float2* dt = (const float2*)target;
        WsStage st;
        if (!on_device) {
            const int64_t ni = batch * net->n_pixels();
            st.x.upload((const float2*)x, (size_t)ni, ctx->stream);
            st.t.upload((const float2*)target, (size_t)ni, ctx->stream);
            dx = st.x.p;
            dt = st.t.p;
        }
        const double L = net->loss_grad(dx, dt, batch, loss != nullptr);
        net->adam_step(lr);
        if (loss) *loss = L;
        if (!on_device) HIPCHECK(hipStreamSynchronize(ctx->stream));
    });
}

int mm_wsnet_destroy(mm_wsnet* net) {
    return guarded([&] {
        if (!net) return;
        (void)hipSetDevice(net->ctx->device);
        (void)hipStreamSynchronize(net->ctx->stream);
        delete net;
    });
}

}  // extern "C"

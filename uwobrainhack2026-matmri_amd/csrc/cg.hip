// cgne on the device: CG on the normal equations E'(W.*E) x = E'(W.*y) for
// one or many independent slices at once, following iterativeSolvers/cgne.m.
//
// Vectors are complex64 in HBM; every inner product is accumulated in FP64
// (per-block partial sums reduced in a fixed order, so results do not depend
// on scheduling) and the scalars alpha, beta, |r|^2 and the stopping logic
// live in FP64 device memory.  The host never reads a scalar inside the
// loop: a slice that has stopped sets its `done` flag and every later kernel
// skips it; the host only polls the flags every few iterations to stop
// launching.  The regular iteration (everything except the explicit-residual
// restarts, cgne.m:230-246) is captured once into a hipGraph and replayed.
#include <algorithm>
#include <cstring>

#include "cg.hpp"

namespace mm {

struct CgScal {
    double rho_re, rho_im;       // res_sq_old
    double al_re, al_im;         // alpha
    double beta;                 // betaFact * res_sq_new / res_sq_old (real; cgne.m:328)
    double stopThresh;           // < 0: off
    double resThresh;            // effective absolute threshold (< 0: off)
    // stopping options of this solve (cgne.m:21-89), set by cg_init: kernel
    // arguments would be frozen into the captured iteration graph
    double xChangeThresh, resChangeThresh, resIncThresh;
    int stopOnResInc, stopOnXdifInc, nseExtraIt;
    int nit, nitmax, done, status;
    int resIncs, xdifIncs, lastXdifInc, fallback;  // fallback: x0 replaced by 0
};

struct CgSeg {
    int64_t off, len;   // vector segment of the slice
    int blk0, nblk;     // partial-sum blocks of the slice
};

constexpr int kCgBlock = 256;
// threads of the per-slice scalar kernels: one wave per slice in cg_beta and
// cg_init (a lane per partial-sum block), one slice per lane in zero_scal_k;
// a compile-time constant, as blockDim.x would be a memory load
constexpr int kCgScalBlock = 64;
constexpr int kCgMaxBlocksPerSlice = kCgScalBlock;
// elements per lane whose loads the per-iteration vector kernels issue
// together (a lane's loop iterations would each wait for their loads)
constexpr int kCgUnroll = 4;

__device__ __forceinline__ double2 warp_sum2(double2 v) {
    for (int o = 32; o > 0; o >>= 1) {
        v.x += __shfl_down(v.x, o, 64);
        v.y += __shfl_down(v.y, o, 64);
    }
    return v;
}

// Block reduction of NV double2 values; thread 0 gets the totals.
template <int NV>
__device__ __forceinline__ void block_sum(double2 (&v)[NV]) {
    __shared__ double2 red[NV][kCgBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        v[k] = warp_sum2(v[k]);
        if (lane == 0) red[k][wid] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double2 s = red[k][0];
            for (int w = 1; w < kCgBlock / 64; ++w) {
                s.x += red[k][w].x;
                s.y += red[k][w].y;
            }
            v[k] = s;
        }
    }
}

// conj(a) * b accumulated in FP64
__device__ __forceinline__ void acc_cdot(double2& s, float2 a, float2 b) {
    s.x += (double)a.x * b.x + (double)a.y * b.y;
    s.y += (double)a.x * b.y - (double)a.y * b.x;
}

// Sum over the slice's partial-sum blocks: one wave per slice, one block
// per lane (at most kCgMaxBlocksPerSlice = 64), an xor butterfly in a fixed
// order, so every lane holds the same deterministic total.  The loads are
// issued together (a loop over the partials would wait on each).
__device__ __forceinline__ double wave_total(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// block -> slice via the block-to-slice table
__device__ __forceinline__ int find_slice(const CgSeg*, const int* blk2slice) {
    return blk2slice[blockIdx.x];
}

// part[blk][0] = sum conj(p) q, part[blk][1] = sum conj(p) p
__global__ __launch_bounds__(kCgBlock) void cg_dot_pq(const CgSeg* __restrict__ seg,
                                                      const int* __restrict__ b2s,
                                                      const CgScal* __restrict__ sc,
                                                      const float2* __restrict__ p,
                                                      const float2* __restrict__ q,
                                                      double2* __restrict__ part) {
    const int s = find_slice(seg, b2s);
    if (sc[s].done) return;
    const CgSeg g = seg[s];
    const int bl = blockIdx.x - g.blk0;
    double2 v[2] = {make_double2(0, 0), make_double2(0, 0)};
    const int64_t S = (int64_t)g.nblk * kCgBlock;
    for (int64_t i0 = (int64_t)bl * kCgBlock + threadIdx.x; i0 < g.len; i0 += kCgUnroll * S) {
        float2 a[kCgUnroll], c[kCgUnroll];
        float w[kCgUnroll];
#pragma unroll
        for (int u = 0; u < kCgUnroll; ++u) {
            const int64_t i = i0 + u * S;
            w[u] = i < g.len ? 1.f : 0.f;
            const int64_t k = g.off + min(i, g.len - 1);
            a[u] = p[k];
            c[u] = q[k];
        }
#pragma unroll
        for (int u = 0; u < kCgUnroll; ++u) {
            const float2 aw = make_float2(a[u].x * w[u], a[u].y * w[u]);
            acc_cdot(v[0], aw, c[u]);
            acc_cdot(v[1], aw, a[u]);
        }
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = v[0];
        part[2 * blockIdx.x + 1] = v[1];
    }
}

// alpha = rho / (p'Ap) (cgne.m:228), then x += alpha p and, if `r_update`,
// r -= alpha q.  Every block reduces the slice's p'Ap partial sums (dpart,
// from cg_dot_pq) itself -- one wave, a fixed order, so all blocks hold the
// same alpha -- and the slice's first block records alpha and |p|^2 (for
// |x - xold|^2).  Partials out: [0] = r'r, [1] = x'x.
__global__ __launch_bounds__(kCgBlock) void cg_update(const CgSeg* __restrict__ seg,
                                                      const int* __restrict__ b2s,
                                                      CgScal* __restrict__ sc,
                                                      const double2* __restrict__ dpart,
                                                      double* __restrict__ pp_out,
                                                      float2* __restrict__ x,
                                                      const float2* __restrict__ p,
                                                      float2* __restrict__ r,
                                                      const float2* __restrict__ q, int r_update,
                                                      const float2* __restrict__ gt,
                                                      double2* __restrict__ part) {
    const int s = find_slice(seg, b2s);
    if (sc[s].done) return;
    const CgSeg g = seg[s];
    const int bl = blockIdx.x - g.blk0;
    __shared__ double al[2];
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        const bool in = lane < g.nblk;
        const int pb = g.blk0 + (in ? lane : 0);
        const double2 a = dpart[2 * pb];
        const double b = dpart[2 * pb + 1].x;
        const double pqx = wave_total(in ? a.x : 0.0), pqy = wave_total(in ? a.y : 0.0);
        const double pp = wave_total(in ? b : 0.0);
        if (lane == 0) {
            // complex division rho / pq
            const double den = pqx * pqx + pqy * pqy;
            const double rr = sc[s].rho_re, ri = sc[s].rho_im;
            const double are = (rr * pqx + ri * pqy) / den, aim = (ri * pqx - rr * pqy) / den;
            al[0] = are;
            al[1] = aim;
            if (bl == 0) {
                sc[s].al_re = are;
                sc[s].al_im = aim;
                pp_out[s] = pp;
            }
        }
    }
    __syncthreads();
    const double are = al[0], aim = al[1];
    double2 v[2] = {make_double2(0, 0), make_double2(0, 0)};
    const int64_t S = (int64_t)g.nblk * kCgBlock;
    for (int64_t i0 = (int64_t)bl * kCgBlock + threadIdx.x; i0 < g.len; i0 += kCgUnroll * S) {
        float2 pv[kCgUnroll], xv[kCgUnroll], qv[kCgUnroll], rv[kCgUnroll];
        int64_t k[kCgUnroll];
#pragma unroll
        for (int u = 0; u < kCgUnroll; ++u) {
            k[u] = g.off + min(i0 + u * S, g.len - 1);
            pv[u] = p[k[u]];
            xv[u] = x[k[u]];
            qv[u] = q[k[u]];   // read also when r is not updated (a valid vector)
            rv[u] = r[k[u]];
        }
#pragma unroll
        for (int u = 0; u < kCgUnroll; ++u) {
            const bool in = i0 + u * S < g.len;
            const float w = in ? 1.f : 0.f;
            float2 xn;
            xn.x = (float)(xv[u].x + (are * pv[u].x - aim * pv[u].y));
            xn.y = (float)(xv[u].y + (are * pv[u].y + aim * pv[u].x));
            if (in) x[k[u]] = xn;
            // |x|^2 in .x; |gtruth - x|^2 (cgne.m:271-274) in .y, the slot
            // the imaginary part of x'x (identically zero) would take
            v[1].x += ((double)xn.x * xn.x + (double)xn.y * xn.y) * w;
            if (gt) {
                const float2 g = gt[k[u]];
                const double dr = (double)g.x - xn.x, di = (double)g.y - xn.y;
                v[1].y += (dr * dr + di * di) * w;
            }
            if (r_update) {
                float2 rn;
                rn.x = (float)(rv[u].x - (are * qv[u].x - aim * qv[u].y));
                rn.y = (float)(rv[u].y - (are * qv[u].y + aim * qv[u].x));
                if (in) r[k[u]] = rn;
                acc_cdot(v[0], make_float2(rn.x * w, rn.y * w), rn);
            }
        }
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = v[0];
        part[2 * blockIdx.x + 1] = v[1];
    }
}

// r = b - Nx (explicit residual, cgne.m:165-170 / :232-245); partial r'r
// into part[..][0].  `force` ignores the done flags (initial residual).
__global__ __launch_bounds__(kCgBlock) void cg_resid(const CgSeg* __restrict__ seg,
                                                     const int* __restrict__ b2s,
                                                     const CgScal* __restrict__ sc,
                                                     const float2* __restrict__ b,
                                                     const float2* __restrict__ nx,
                                                     float2* __restrict__ r, float2* __restrict__ p,
                                                     double2* __restrict__ part) {
    const int s = find_slice(seg, b2s);
    if (sc[s].done) return;
    const CgSeg g = seg[s];
    const int bl = blockIdx.x - g.blk0;
    double2 v[1] = {make_double2(0, 0)};
    for (int64_t i = (int64_t)bl * kCgBlock + threadIdx.x; i < g.len; i += (int64_t)g.nblk * kCgBlock) {
        const int64_t k = g.off + i;
        const float2 bv = b[k], nv = nx[k];
        const float2 rv = make_float2(bv.x - nv.x, bv.y - nv.y);
        r[k] = rv;
        if (p) p[k] = rv;
        acc_cdot(v[0], rv, rv);
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) part[2 * blockIdx.x] = v[0];
}

// |b|^2 (for the relative threshold) and x0 = b when requested
__global__ __launch_bounds__(kCgBlock) void cg_norm_b(const CgSeg* __restrict__ seg,
                                                      const int* __restrict__ b2s,
                                                      const float2* __restrict__ b,
                                                      float2* __restrict__ x, int copy,
                                                      const float2* __restrict__ gt,
                                                      double2* __restrict__ part) {
    const int s = find_slice(seg, b2s);
    const CgSeg g = seg[s];
    const int bl = blockIdx.x - g.blk0;
    double2 v[2] = {make_double2(0, 0), make_double2(0, 0)};
    for (int64_t i = (int64_t)bl * kCgBlock + threadIdx.x; i < g.len; i += (int64_t)g.nblk * kCgBlock) {
        const int64_t k = g.off + i;
        const float2 bv = b[k];
        acc_cdot(v[0], bv, bv);
        const float2 xv = copy ? bv : x[k];
        if (copy) x[k] = bv;
        v[1].x += (double)xv.x * xv.x + (double)xv.y * xv.y;
        if (gt) {
            const double dr = (double)gt[k].x - xv.x, di = (double)gt[k].y - xv.y;
            v[1].y += dr * dr + di * di;
        }
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = v[0];
        part[2 * blockIdx.x + 1] = v[1];
    }
}

// Initial scalars after r0 = b - N x0 (cgne.m:180-207).
__global__ __launch_bounds__(kCgScalBlock) void cg_init(const CgSeg* __restrict__ seg, CgScal* __restrict__ sc,
                        const double2* __restrict__ part_b, const double2* __restrict__ part_r,
                        double* __restrict__ hres, double* __restrict__ hxn, double* __restrict__ hxd,
                        double* __restrict__ hmse,
                        int nslice, int nitmax, double resThresh, double resThreshRel,
                        double stopThresh, int x0fallback, double xChangeThresh,
                        double resChangeThresh, int stopOnResInc, double resIncThresh,
                        int stopOnXdifInc, int nseExtraIt) {
    const int s = blockIdx.x, lane = threadIdx.x;   // one wave per slice
    if (s >= nslice) return;
    const CgSeg g = seg[s];
    const bool in = lane < g.nblk;
    const int pb = g.blk0 + (in ? lane : 0);
    const double b0 = part_b[2 * pb].x, x0 = part_b[2 * pb + 1].x, r0 = part_r[2 * pb].x;
    const double m0 = part_b[2 * pb + 1].y;
    double bb = wave_total(in ? b0 : 0.0), xx = wave_total(in ? x0 : 0.0), rr = wave_total(in ? r0 : 0.0);
    const double me = wave_total(in ? m0 : 0.0);
    if (lane != 0) return;
    CgScal c{};
    // warm-start fallback: a start worse than x0 = 0 (whose residual is b)
    // is replaced by x0 = 0 (cg_fallback_k resets x, r and p)
    if (x0fallback && rr > bb) {
        c.fallback = 1;
        rr = bb;
        xx = 0;
    }
    c.rho_re = rr;
    c.rho_im = 0;
    c.nit = 0;
    c.nitmax = nitmax;
    c.stopThresh = stopThresh;
    c.xChangeThresh = xChangeThresh;
    c.resChangeThresh = resChangeThresh;
    c.resIncThresh = resIncThresh;
    c.stopOnResInc = stopOnResInc;
    c.stopOnXdifInc = stopOnXdifInc;
    c.nseExtraIt = nseExtraIt;
    double th = resThresh;
    if (resThreshRel >= 0) th = th >= 0 ? fmax(th, resThreshRel * bb) : resThreshRel * bb;
    c.resThresh = th;
    c.done = nitmax <= 0;
    sc[s] = c;
    const int64_t h = (int64_t)s * (nitmax + 1);
    hres[h] = rr;
    hxn[h] = xx;
    hxd[h] = xx;
    if (hmse) hmse[h] = me;
}

// x = 0, r = p = b for the slices cg_init marked as falling back
__global__ __launch_bounds__(kCgBlock) void cg_fallback_k(const CgSeg* __restrict__ seg,
                                                          const int* __restrict__ b2s,
                                                          const CgScal* __restrict__ sc,
                                                          const float2* __restrict__ b, float2* __restrict__ x,
                                                          float2* __restrict__ r, float2* __restrict__ p) {
    const int s = find_slice(seg, b2s);
    if (!sc[s].fallback) return;
    const CgSeg g = seg[s];
    const int bl = blockIdx.x - g.blk0;
    for (int64_t i = (int64_t)bl * kCgBlock + threadIdx.x; i < g.len; i += (int64_t)g.nblk * kCgBlock) {
        const int64_t k = g.off + i;
        const float2 bv = b[k];
        x[k] = make_float2(0.f, 0.f);
        r[k] = bv;
        p[k] = bv;
    }
}

// res_sq_new, histories, stopping tests and beta (cgne.m:252-335).
__global__ __launch_bounds__(kCgScalBlock) void cg_beta(const CgSeg* __restrict__ seg, CgScal* __restrict__ sc,
                        const double2* __restrict__ part, const double* __restrict__ pp,
                        double* __restrict__ hres, double* __restrict__ hxn, double* __restrict__ hxd,
                        double* __restrict__ hmse,
                        int nslice, int nitmax_alloc, int explicit_res, int* __restrict__ done_host,
                        int* __restrict__ done_map) {
    const int s = blockIdx.x, lane = threadIdx.x;   // one wave per slice
    if (s >= nslice) return;
    if (sc[s].done) {
        if (lane == 0) {
            done_host[s] = 1;
            done_map[s] = 1;
        }
        return;
    }
    const CgSeg g = seg[s];
    const bool in = lane < g.nblk;
    const int pb = g.blk0 + (in ? lane : 0);
    const double r0 = part[2 * pb].x, x0 = part[2 * pb + 1].x, m0 = part[2 * pb + 1].y;
    const double rr = wave_total(in ? r0 : 0.0), xx = wave_total(in ? x0 : 0.0);
    const double me = wave_total(in ? m0 : 0.0);
    if (lane != 0) return;
    CgScal c = sc[s];
    const double a2 = c.al_re * c.al_re + c.al_im * c.al_im;
    const double xd = a2 * pp[s];
    const int64_t h = (int64_t)s * (nitmax_alloc + 1);
    const int n = c.nit;
    hres[h + n + 1] = rr;
    hxn[h + n + 1] = xx;
    hxd[h + n + 1] = xd;
    if (hmse) hmse[h + n + 1] = me;
    int finished = 0;
    const double xChangeThresh = c.xChangeThresh, resChangeThresh = c.resChangeThresh;
    const double resIncThresh = c.resIncThresh;
    const int stopOnResInc = c.stopOnResInc, stopOnXdifInc = c.stopOnXdifInc, nseExtraIt = c.nseExtraIt;
    if (xChangeThresh >= 0 && xd < xChangeThresh) { finished = 1; c.status = 10; }
    if (c.resThresh >= 0 && rr < c.resThresh) { finished = 1; c.status = 1; }
    if (resChangeThresh >= 0 && fabs((hres[h + n + 1] - hres[h + n]) / hres[h + n]) < resChangeThresh) {
        finished = 1;
        c.status = 2;
    }
    if (c.stopThresh >= 0 && rr < c.stopThresh) {
        c.nitmax = n + nseExtraIt;
        c.stopThresh = 0;
        c.status = 3;
    }
    if (stopOnResInc > 0 && hres[h + n + 1] / hres[h + n] > resIncThresh) {
        c.resIncs++;
        if (c.resIncs >= stopOnResInc) { finished = 1; c.status = 4; }
    }
    if (stopOnXdifInc > 0 && hxd[h + n + 1] > hxd[h + n]) {
        if (!c.lastXdifInc) {
            c.xdifIncs++;
            c.lastXdifInc = 1;
            if (c.xdifIncs >= stopOnXdifInc) { finished = 1; c.status = 5; }
        }
    } else {
        c.lastXdifInc = 0;
    }
    const double betaFact = explicit_res ? 0.0 : 1.0;
    c.beta = betaFact * rr / c.rho_re;
    c.rho_re = rr;
    c.rho_im = 0;
    c.nit = n + 1;
    if (c.nit >= c.nitmax) { finished = 1; c.status = 0; }
    // the p update of this iteration still runs (beta above); `done` takes
    // effect from the next kernel that checks it
    c.done = finished ? 2 : 0;
    sc[s] = c;
    done_host[s] = finished;
    done_map[s] = finished;
}

// p = r + beta p (cgne.m:328); a slice that finished this iteration (done == 2)
// is promoted to done == 1.
__global__ __launch_bounds__(kCgBlock) void cg_update_p(const CgSeg* __restrict__ seg,
                                                        const int* __restrict__ b2s,
                                                        CgScal* __restrict__ sc,
                                                        float2* __restrict__ p,
                                                        const float2* __restrict__ r) {
    const int s = find_slice(seg, b2s);
    const int dn = sc[s].done;
    if (dn == 1) return;
    const CgSeg g = seg[s];
    const int bl = blockIdx.x - g.blk0;
    if (dn == 2) {
        if (bl == 0 && threadIdx.x == 0) sc[s].done = 1;
        return;
    }
    const double be = sc[s].beta;
    const int64_t S = (int64_t)g.nblk * kCgBlock;
    for (int64_t i0 = (int64_t)bl * kCgBlock + threadIdx.x; i0 < g.len; i0 += kCgUnroll * S) {
        float2 rv[kCgUnroll], pv[kCgUnroll];
        int64_t k[kCgUnroll];
#pragma unroll
        for (int u = 0; u < kCgUnroll; ++u) {
            k[u] = g.off + min(i0 + u * S, g.len - 1);
            rv[u] = r[k[u]];
            pv[u] = p[k[u]];
        }
#pragma unroll
        for (int u = 0; u < kCgUnroll; ++u)
            if (i0 + u * S < g.len)
                p[k[u]] = make_float2((float)(rv[u].x + be * pv[u].x), (float)(rv[u].y + be * pv[u].y));
    }
}

// out += g * in (the Tikhonov terms gam * R'R v of cgne.m:171-179, :219-227)
__global__ __launch_bounds__(kCgBlock) void cg_axpy_k(float2* __restrict__ out, const float2* __restrict__ in,
                                                      float g, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)kCgBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kCgBlock) {
        const float2 a = in[i];
        float2 o = out[i];
        o.x += g * a.x;
        o.y += g * a.y;
        out[i] = o;
    }
}

__global__ __launch_bounds__(kCgBlock) void copy_rr_k(const double2* __restrict__ src, double2* __restrict__ dst, int nblk) {
    const int b = blockIdx.x * kCgBlock + threadIdx.x;
    if (b < nblk) dst[2 * b] = src[2 * b];
}

__global__ __launch_bounds__(kCgBlock) void scale_copy_k(const float2* __restrict__ y, const float* __restrict__ w,
                             float2* __restrict__ out, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)kCgBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kCgBlock)
        out[i] = make_float2(y[i].x * w[i], y[i].y * w[i]);
}

__global__ __launch_bounds__(kCgScalBlock) void zero_scal_k(CgScal* sc, int ns) {
    const int s = blockIdx.x * kCgScalBlock + threadIdx.x;
    if (s < ns) sc[s] = CgScal{};
}

// ---------------------------------------------------------------------------
CgSolver::CgSolver(mm_op* E) : E_(E) {
    const int ns = E->nslice();
    std::vector<CgSeg> seg(ns);
    std::vector<int> b2s;
    for (int s = 0; s < ns; ++s) {
        const int64_t off = E->slice_image_off(s);
        const int64_t end = s + 1 < ns ? E->slice_image_off(s + 1) : E->n_image();
        seg[s].off = off;
        seg[s].len = end - off;
        // kCgUnroll elements per lane, up to kCgMaxBlocksPerSlice blocks
        const int64_t want = (seg[s].len + kCgBlock * kCgUnroll - 1) / (kCgBlock * kCgUnroll);
        seg[s].nblk = (int)std::max<int64_t>(1, std::min<int64_t>(want, kCgMaxBlocksPerSlice));
        seg[s].blk0 = (int)b2s.size();
        for (int b = 0; b < seg[s].nblk; ++b) b2s.push_back(s);
    }
    nblk_ = (int)b2s.size();
    n_ = E->n_image();
    hipStream_t st = E->ctx->stream;
    seg_.upload(seg, st);
    b2s_.upload(b2s, st);
    sc_.alloc(ns);
    part_.alloc((size_t)2 * nblk_);
    part2_.alloc((size_t)2 * nblk_);
    partu_.alloc((size_t)2 * nblk_);
    pp_.alloc(ns);
    b_.alloc(n_);
    r_.alloc(n_);
    p_.alloc(n_);
    q_.alloc(n_);
    tmp_.alloc(E->n_data());
    // the done flags also go to coherent, device-mapped host memory, which
    // cg_beta writes directly: the host polls them without a copy
    HIPCHECK(hipHostMalloc((void**)&done_h_, sizeof(int) * ns, hipHostMallocMapped | hipHostMallocCoherent));
    HIPCHECK(hipHostGetDevicePointer((void**)&done_map_, done_h_, 0));
    HIPCHECK(hipEventCreateWithFlags(&poll_ev_, hipEventDisableTiming));
    HIPCHECK(hipStreamSynchronize(st));
}

CgSolver::~CgSolver() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (done_h_) (void)hipHostFree(done_h_);
    if (poll_ev_) (void)hipEventDestroy(poll_ev_);
}

void CgSolver::drop_graph() {
    if (graph_exec_) { (void)hipGraphExecDestroy(graph_exec_); graph_exec_ = nullptr; }
    if (graph_) { (void)hipGraphDestroy(graph_); graph_ = nullptr; }
}

// out = E'(W.*E v) + sum_n gam_n R_n'(R_n v)  (R_n NULL: the identity)
uint64_t CgSolver::config_gen() const {
    uint64_t g = E_->config_gen();
    for (const auto& rg : regs_)
        if (rg.second) g += rg.second->config_gen();
    return g;
}

void CgSolver::normal(const float2* v, float2* out) {
    hipStream_t st = E_->ctx->stream;
    E_->apply_normal(v, out, 1, tmp_.p);
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((n_ + kCgBlock - 1) / kCgBlock, 4096));
    for (const auto& rg : regs_) {
        const float2* t = v;
        if (rg.second) {
            rg.second->apply_normal(v, rbuf_.p, 1, rtmp_.p);
            t = rbuf_.p;
        }
        hipLaunchKernelGGL(cg_axpy_k, dim3(grid), dim3(kCgBlock), 0, st, out, t, (float)rg.first, n_);
        LAUNCHCHECK();
    }
}

void CgSolver::launch_iteration(int explicit_res, const mm_cg_opts& o, float2* x) {
    hipStream_t st = E_->ctx->stream;
    const int ns = E_->nslice();
    normal(p_.p, q_.p);
    hipLaunchKernelGGL(cg_dot_pq, dim3(nblk_), dim3(kCgBlock), 0, st, seg_.p, b2s_.p, sc_.p, p_.p, q_.p,
                       part_.p);
    hipLaunchKernelGGL(cg_update, dim3(nblk_), dim3(kCgBlock), 0, st, seg_.p, b2s_.p, sc_.p, part_.p, pp_.p, x,
                       p_.p, r_.p, q_.p, explicit_res ? 0 : 1, gt_, partu_.p);
    if (explicit_res) {
        // r = b - N x; r'r into part2, x'x stays in part_[..][1]
        normal(x, q_.p);
        hipLaunchKernelGGL(cg_resid, dim3(nblk_), dim3(kCgBlock), 0, st, seg_.p, b2s_.p, sc_.p, b_.p,
                           q_.p, r_.p, (float2*)nullptr, part2_.p);
        hipLaunchKernelGGL(copy_rr_k, dim3((nblk_ + 255) / 256), dim3(256), 0, st, part2_.p, partu_.p,
                           nblk_);
    }
    hipLaunchKernelGGL(cg_beta, dim3(ns), dim3(kCgScalBlock), 0, st, seg_.p, sc_.p, partu_.p, pp_.p, hres_.p, hxn_.p,
                       hxd_.p, gt_ ? hmse_.p : (double*)nullptr, ns, nitmax_, explicit_res, done_d_.p, done_map_);
    hipLaunchKernelGGL(cg_update_p, dim3(nblk_), dim3(kCgBlock), 0, st, seg_.p, b2s_.p, sc_.p, p_.p, r_.p);
    LAUNCHCHECK();
}

void CgSolver::solve(const float2* y, const float2* x0, float2* x, int NitMax, const mm_cg_opts& o,
                     double* resSqAll, double* xnormAll, double* xdiffAll, int* nit, int* status,
                     const std::vector<std::pair<double, mm_op*>>& regs, const float2* gtruth, double* mseAll) {
    hipStream_t st = E_->ctx->stream;
    const int ns = E_->nslice();
    for (const auto& rg : regs)
        if (rg.second && (rg.second->n_image() != n_ || rg.second->W))
            fail(MM_ERR_ARG, "cgne: a regulariser R must map the image of E (and carry no data weights)");
    if (regs != regs_ || gtruth != gt_) drop_graph();
    regs_ = regs;
    gt_ = gtruth;
    int64_t rmax = 0;
    for (const auto& rg : regs_)
        if (rg.second) rmax = std::max(rmax, rg.second->n_data());
    if (rmax) {
        rbuf_.ensure((size_t)n_);
        rtmp_.ensure((size_t)rmax);
    }
    const int sb = (ns + kCgScalBlock - 1) / kCgScalBlock;
    if (NitMax < 0) fail(MM_ERR_ARG, "cgne: NitMax must be >= 0");
    if (NitMax != nitmax_) {
        nitmax_ = NitMax;
        hres_.alloc((size_t)ns * (NitMax + 1));
        hxn_.alloc((size_t)ns * (NitMax + 1));
        hxd_.alloc((size_t)ns * (NitMax + 1));
        hmse_.alloc((size_t)ns * (NitMax + 1));
        done_d_.alloc(ns);
        drop_graph();
    }
    if (graph_x_ != x) drop_graph();
    // b = E'(W.*y)  (cgne.m:136-140)
    if (E_->W) {
        hipLaunchKernelGGL(scale_copy_k, dim3(256), dim3(256), 0, st, y, E_->W, tmp_.p, E_->n_data());
        E_->apply(1, tmp_.p, b_.p, 1);
    } else {
        E_->apply(1, y, b_.p, 1);
    }
    // x0 (cgne.m:143-145)
    if (x0) HIPCHECK(hipMemcpyAsync(x, x0, sizeof(float2) * n_, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(cg_norm_b, dim3(nblk_), dim3(kCgBlock), 0, st, seg_.p, b2s_.p, b_.p, x,
                       x0 ? 0 : 1, gt_, part2_.p);
    // r = b - N x0, p = r  (cgne.m:165-180)
    normal(x, q_.p);
    hipLaunchKernelGGL(zero_scal_k, dim3(sb), dim3(kCgScalBlock), 0, st, sc_.p, ns);
    hipLaunchKernelGGL(cg_resid, dim3(nblk_), dim3(kCgBlock), 0, st, seg_.p, b2s_.p, sc_.p, b_.p, q_.p,
                       r_.p, p_.p, part_.p);
    hipLaunchKernelGGL(cg_init, dim3(ns), dim3(kCgScalBlock), 0, st, seg_.p, sc_.p, part2_.p, part_.p, hres_.p,
                       hxn_.p, hxd_.p, gt_ ? hmse_.p : (double*)nullptr, ns, NitMax, o.resThresh, o.resThreshRel,
                       o.stopThresh,
                       (x0 && o.x0Fallback) ? 1 : 0, o.xChangeThresh, o.resChangeThresh, o.stopOnResInc,
                       o.resIncThresh, o.stopOnXdifInc, o.nseExtraIt);
    LAUNCHCHECK();
    if (x0 && o.x0Fallback) {
        hipLaunchKernelGGL(cg_fallback_k, dim3(nblk_), dim3(kCgBlock), 0, st, seg_.p, b2s_.p, sc_.p, b_.p, x,
                           r_.p, p_.p);
        LAUNCHCHECK();
    }
    const int expResN = o.expResN > 0 ? o.expResN : 100;
    // every iteration for large batches (an iteration outlasts the host's
    // wait and relaunch by far), every 4th for small ones
    const int poll = n_ >= (int64_t)1 << 20 ? 1 : 4;
    // finished slices are left out of the operator from the next iteration on
    HIPCHECK(hipMemsetAsync(done_d_.p, 0, sizeof(int) * ns, st));
    std::memset(done_h_, 0, sizeof(int) * ns);
    struct SkipGuard {
        mm_op* E;
        ~SkipGuard() { E->set_skip(nullptr); }
    } guard{E_};
    E_->set_skip(done_d_.p);
    // cg_beta writes the done flags to mapped host memory as well; the host
    // reads them after every `poll` iterations, once the next iteration has
    // been queued, so the device never waits for the host (a solve that has
    // finished runs one more iteration whose kernels all return at once).
    // the captured iteration holds the operator's kernels, work tables and
    // buffers: recapture after a configuration change of E or any device
    // free since the capture (the preamble above may itself have reallocated)
    if (graph_exec_ && (config_gen() != graph_gen_ || free_epoch().load() != graph_epoch_)) drop_graph();
    bool pending = false;
    for (int it = 0; it < NitMax; ++it) {
        const int explicit_res = (it % expResN) == 0;
        if (explicit_res || !use_graph_) {
            launch_iteration(explicit_res, o, x);
        } else {
            if (!graph_exec_) {
                HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                launch_iteration(0, o, x);
                HIPCHECK(hipStreamEndCapture(st, &graph_));
                HIPCHECK(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
                graph_x_ = x;
                graph_gen_ = config_gen();
                graph_epoch_ = free_epoch().load();
            }
            HIPCHECK(hipGraphLaunch(graph_exec_, st));
        }
        if (pending) {
            pending = false;
            HIPCHECK(hipEventSynchronize(poll_ev_));
            bool all = true;
            for (int s = 0; s < ns; ++s) all &= done_h_[s] != 0;
            if (all) break;
        }
        if ((it + 1) % poll == 0 && it + 2 < NitMax) {
            HIPCHECK(hipEventRecord(poll_ev_, st));
            pending = true;
        }
    }
    // results
    std::vector<CgScal> hs(ns);
    HIPCHECK(hipMemcpyAsync(hs.data(), sc_.p, sizeof(CgScal) * ns, hipMemcpyDeviceToHost, st));
    const size_t hn = (size_t)ns * (NitMax + 1);
    std::vector<double> a(hn), b(hn), c(hn), d(gt_ ? hn : 0);
    if (gt_) HIPCHECK(hipMemcpyAsync(d.data(), hmse_.p, sizeof(double) * hn, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(a.data(), hres_.p, sizeof(double) * hn, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(b.data(), hxn_.p, sizeof(double) * hn, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(c.data(), hxd_.p, sizeof(double) * hn, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    last_fallback_.assign(ns, 0);
    for (int s = 0; s < ns; ++s) {
        last_fallback_[s] = hs[s].fallback;
        if (nit) nit[s] = hs[s].nit;
        if (status) status[s] = hs[s].status;
    }
    if (resSqAll) std::memcpy(resSqAll, a.data(), sizeof(double) * hn);
    if (xnormAll) std::memcpy(xnormAll, b.data(), sizeof(double) * hn);
    if (xdiffAll) std::memcpy(xdiffAll, c.data(), sizeof(double) * hn);
    if (mseAll) {
        if (gt_) std::memcpy(mseAll, d.data(), sizeof(double) * hn);
        else std::memset(mseAll, 0, sizeof(double) * hn);
    }
}

}  // namespace mm

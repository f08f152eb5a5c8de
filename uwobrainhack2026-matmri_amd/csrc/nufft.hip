// Kaiser-Bessel gridding NUFFT for gfx950.
//
// Plan (host, FP64, once): the closed forms and the per-sample kernel weights
// of nufftOp.doPrep (operators/@nufftOp/nufftOp.m:138-274) -- grid size
// (:146-152), KB beta and LUT (:154-165), deapodisation (:171-186), linearly
// interpolated LUT weights (:216-227) and wrapped grid indices (:230-246).
// The sparse matrix convMat is never formed: each sample keeps its first
// grid index per dimension and its W weights per dimension.
//
// Forward (nufftOp.m:378-394): y = dcf .* C * (P .* ifftnc(comp .* x, pad G)).
// The k-space phase ramp P = fftShiftArray (:258-273) equals an image-domain
// circular shift by N/2 times the constant (-1)^(N1/2+N2/2), so the image is
// written wrapped into the zero-padded FFT input and no ramp is multiplied.
// Adjoint (:354-377) is the exact Hermitian transpose: gridding, FFT, crop,
// deapodisation.
//
// Grid layout in HBM: Z[g2][g1][b], b = transform index within a chunk of
// BCT transforms (coils or repetitions), so one kernel tap touches BCT
// consecutive complex values.
//
// Kernels (one launch per stage covers every slice and every chunk):
//   nu_fwd_rows   coil multiply + deapod + wrap/zero-pad + FFT along g1
//   nu_cols       FFT along g2 (fwd: zero-pad from N2 rows; adj: crop to N2)
//   nu_degrid     gather of W^2 taps per sample (dcf, optional weights fused)
//   nu_grid       LDS-tiled gridding; a tile is owned by one work-group and
//                 written once, heavy tiles are split and merged by the
//                 last-arriving work-group (no global float atomics)
//   nu_adj_rows   FFT along g1 + crop + deapod + conj-coil combine
#include <algorithm>
#include <cmath>
#include <numeric>

#include "nufft.hpp"

namespace mm {

// ---------------------------------------------------------------------------
// host-side plan constants
static int max_prime_factor(int64_t n) {
    int64_t best = 1;
    for (int64_t p = 2; p * p <= n; ++p)
        while (n % p == 0) { best = p; n /= p; }
    if (n > 1) best = n;
    return (int)best;
}

int nufft_grid_size(int64_t n, double os) {
    int64_t g = (int64_t)std::ceil(n * os / 2.0) * 2;
    while (max_prime_factor(g) > 7) g += 2;
    return (int)g;
}

double nufft_kb_beta(int kwidth, double os) {
    return M_PI * std::sqrt((double)kwidth * kwidth / (os * os) * (os - 0.5) * (os - 0.5) - 0.8);
}

int nufft_kb_nkern(int kwidth, double os) {
    const double precision = 1e-6;
    return (int)std::ceil(0.5 * (kwidth + 1) * std::sqrt(0.37 / (os * os) / precision)) * 2;
}

// Modified Bessel function I0 by its power series (converges for all x;
// the KB arguments here are below 30).
static double bessel_i0(double x) {
    const double q = 0.25 * x * x;
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 500; ++k) {
        term *= q / ((double)k * k);
        sum += term;
        if (term < 1e-17 * sum) break;
    }
    return sum;
}

// KB LUT with the zero sentinel, normalised as nufftOp.m:163-164.
static std::vector<double> kb_lut(int kwidth, double os) {
    const double bet = nufft_kb_beta(kwidth, os);
    const int nk = nufft_kb_nkern(kwidth, os);
    std::vector<double> k(nk + 1);
    double sum = 0;
    for (int i = 0; i < nk; ++i) {
        const double kx = -1.0 + 2.0 * i / (nk - 1);
        const double arg = std::max(0.0, 1.0 - kx * kx);
        k[i] = bessel_i0(bet * std::sqrt(arg));
        sum += k[i];
    }
    k[nk] = 0.0;
    const double scale = (double)(nk - 1) / kwidth / sum;  // (length(kern)-2)/kwidth
    for (auto& v : k) v *= scale;
    return k;
}

// Deapodisation comp{d} of nufftOp.m:171-182 (real; the reference's
// x./sin(x) with imaginary x equals a/sinh(a), a = sqrt(beta^2 - x^2)).
static std::vector<double> deapod(int N, int G, int kwidth, double os) {
    const double bet = nufft_kb_beta(kwidth, os);
    std::vector<double> c(N);
    double cmin = 0;
    bool first = true;
    for (int i = 0; i < N; ++i) {
        const double j = -N / 2.0 + i;
        const double x = j * M_PI * kwidth / G;
        const double d = bet * bet - x * x;
        double v;
        if (d > 0) {
            const double a = std::sqrt(d);
            v = a / std::sinh(a);
        } else if (d < 0) {
            const double a = std::sqrt(-d);
            v = a / std::sin(a);
        } else {
            v = 1.0;
        }
        c[i] = v;
        if (first || std::fabs(v) < std::fabs(cmin)) { cmin = v; first = false; }
    }
    for (auto& v : c) v = v / cmin * std::sqrt((double)G / N);
    return c;
}

static int next_pow2(int64_t v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

__host__ __device__ static inline int ilog2(int v) {
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

// ---------------------------------------------------------------------------
// device code
struct NuRun {
    int sense;          // 1: image side is one image combined with coil maps
    int C;              // coils (sense) -- lanes per chunk
    int64_t B;          // transforms per slice (plain)
    const float2* maps;
    const float* W;     // optional forward-output weights (data layout)
};

__device__ __forceinline__ int lanes_of(const NuRun& r, int BCT, int ch) {
    if (r.sense) return r.C;
    const int64_t left = r.B - (int64_t)ch * BCT;
    return left < BCT ? (int)left : BCT;
}

__device__ __forceinline__ int64_t chunk_stride(const NuRun& r, int BCT) { return r.sense ? r.C : BCT; }

template <int BCT>
__global__ __launch_bounds__(256) void nu_fwd_rows(const NuSliceDev* __restrict__ sl,
                                                   const int4* __restrict__ work,
                                                   const FftFactors* __restrict__ facs,
                                                   const float2* __restrict__ twp,
                                                   const float* __restrict__ deap,
                                                   const float2* __restrict__ x,
                                                   float2* __restrict__ ws, NuRun run) {
    extern __shared__ float2 lds[];
    const int4 wk = work[blockIdx.x];
    const NuSliceDev& S = sl[wk.x];
    const int ch = wk.y, r0 = wk.z;
    const int N1 = S.N1, N2 = S.N2, G1 = S.G1, rpw = S.rpw;
    const int T = rpw * BCT, log2T = ilog2(T);
    const int64_t V = (int64_t)N1 * N2;
    float2* A = lds;
    float2* Bb = lds + G1 * T;
    float2* tw = lds + 2 * G1 * T;
    for (int i = threadIdx.x; i < G1; i += blockDim.x) tw[i] = twp[S.tw1 + i];
    for (int i = threadIdx.x; i < G1 * T; i += blockDim.x) A[i] = make_float2(0.f, 0.f);
    __syncthreads();
    const int lanes = lanes_of(run, BCT, ch);
    const int total = T * N1;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int i1 = idx % N1;
        const int rest = idx / N1;
        const int b = rest & (BCT - 1);
        const int r = rest / BCT;
        const int i2 = r0 + r;
        if (i2 >= N2 || b >= lanes) continue;
        const int64_t v = (int64_t)i2 * N1 + i1;
        float2 val;
        if (run.sense) {
            const float2 im = x[S.img_off + (int64_t)ch * V + v];
            const float2 mp = run.maps[S.maps_off + (int64_t)b * V + v];
            val = cmulf(im, mp);
        } else {
            val = x[S.img_off + ((int64_t)ch * BCT + b) * V + v];
        }
        const float f = deap[S.deap1 + i1] * deap[S.deap2 + i2];
        int pos = i1 - N1 / 2;
        if (pos < 0) pos += G1;
        A[(pos * rpw + r) * BCT + b] = make_float2(val.x * f, val.y * f);
    }
    __syncthreads();
    const float2* R = fft_lds(A, Bb, facs[S.fac1], log2T, tw, 1.f);
    float2* out = ws + S.ws_off + (int64_t)ch * S.ws_chunk;
    const int tot2 = T * G1;
    for (int idx = threadIdx.x; idx < tot2; idx += blockDim.x) {
        const int b = idx & (BCT - 1);
        const int rest = idx / BCT;
        const int g1 = rest % G1;
        const int r = rest / G1;
        const int i2 = r0 + r;
        if (i2 >= N2) continue;
        out[((int64_t)i2 * G1 + g1) * BCT + b] = R[(g1 * rpw + r) * BCT + b];
    }
}

// FFT along g2 for `cpw` columns.  Forward (adj = 0): input rows 0..N2-1 are
// placed at (i2 - N2/2) mod G2, output G2 rows.  Adjoint: input G2 rows,
// output rows i2 < N2 read from (i2 - N2/2) mod G2.  In place in `ws`.
template <int BCT>
__global__ __launch_bounds__(256) void nu_cols(const NuSliceDev* __restrict__ sl,
                                               const int4* __restrict__ work,
                                               const FftFactors* __restrict__ facs,
                                               const float2* __restrict__ twp,
                                               float2* __restrict__ ws, int adj) {
    extern __shared__ float2 lds[];
    const int4 wk = work[blockIdx.x];
    const NuSliceDev& S = sl[wk.x];
    const int ch = wk.y, c0 = wk.z;
    const int N2 = S.N2, G1 = S.G1, G2 = S.G2, cpw = S.cpw;
    const int T = cpw * BCT, log2T = ilog2(T);
    float2* A = lds;
    float2* Bb = lds + G2 * T;
    float2* tw = lds + 2 * G2 * T;
    float2* g = ws + S.ws_off + (int64_t)ch * S.ws_chunk;
    for (int i = threadIdx.x; i < G2; i += blockDim.x) tw[i] = twp[S.tw2 + i];
    const int ncol = min(cpw, G1 - c0);
    if (!adj) {
        for (int i = threadIdx.x; i < G2 * T; i += blockDim.x) A[i] = make_float2(0.f, 0.f);
        __syncthreads();
        const int total = N2 * T;
        for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
            const int t = idx & (T - 1);
            const int i2 = idx >> log2T;
            const int col = t / BCT;
            if (col >= ncol) continue;
            int pos = i2 - N2 / 2;
            if (pos < 0) pos += G2;
            A[pos * T + t] = g[((int64_t)i2 * G1 + c0) * BCT + t];
        }
    } else {
        const int total = G2 * T;
        for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
            const int t = idx & (T - 1);
            const int g2 = idx >> log2T;
            const int col = t / BCT;
            A[idx] = col < ncol ? g[((int64_t)g2 * G1 + c0) * BCT + t] : make_float2(0.f, 0.f);
        }
    }
    __syncthreads();
    const float2* R = fft_lds(A, Bb, facs[S.fac2], log2T, tw, adj ? -1.f : 1.f);
    if (!adj) {
        const int total = G2 * T;
        for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
            const int t = idx & (T - 1);
            const int g2 = idx >> log2T;
            if (t / BCT >= ncol) continue;
            g[((int64_t)g2 * G1 + c0) * BCT + t] = R[idx];
        }
    } else {
        const int total = N2 * T;
        for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
            const int t = idx & (T - 1);
            const int i2 = idx >> log2T;
            if (t / BCT >= ncol) continue;
            int pos = i2 - N2 / 2;
            if (pos < 0) pos += G2;
            g[((int64_t)i2 * G1 + c0) * BCT + t] = R[pos * T + t];
        }
    }
}

template <int BCT>
__global__ __launch_bounds__(256) void nu_degrid(const NuSliceDev* __restrict__ sl,
                                                 const int4* __restrict__ work,
                                                 const uint32_t* __restrict__ sbase,
                                                 const float* __restrict__ swt,
                                                 const float* __restrict__ sdcf,
                                                 const float2* __restrict__ ws,
                                                 float2* __restrict__ y, NuRun run) {
    const int4 wk = work[blockIdx.x];
    const NuSliceDev& S = sl[wk.x];
    const int ch = wk.y;
    const int m = wk.z + threadIdx.x;
    if (m >= S.M) return;
    const int G1 = S.G1, G2 = S.G2, W1 = S.W1, W2 = S.W2;
    const int64_t sm = S.samp_off + m;
    const uint32_t base = sbase[sm];
    const int g1_0 = base & 0xffff, g2_0 = base >> 16;
    const float* wt = swt + sm * (W1 + W2);
    float w1[8], w2[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        w1[s] = s < W1 ? wt[s] : 0.f;
        w2[s] = s < W2 ? wt[W1 + s] : 0.f;
    }
    float2 acc[BCT];
#pragma unroll
    for (int b = 0; b < BCT; ++b) acc[b] = make_float2(0.f, 0.f);
    const float2* zg = ws + S.ws_off + (int64_t)ch * S.ws_chunk;
    for (int s2 = 0; s2 < W2; ++s2) {
        int g2 = g2_0 + s2;
        if (g2 >= G2) g2 -= G2;
        const float2* zrow = zg + (int64_t)g2 * G1 * BCT;
        for (int s1 = 0; s1 < W1; ++s1) {
            int g1 = g1_0 + s1;
            if (g1 >= G1) g1 -= G1;
            const float w = w1[s1] * w2[s2];
            const float2* z = zrow + (int64_t)g1 * BCT;
            if constexpr (BCT >= 2) {
#pragma unroll
                for (int b = 0; b < BCT; b += 2) {
                    const float4 q = *reinterpret_cast<const float4*>(z + b);
                    acc[b].x += w * q.x;
                    acc[b].y += w * q.y;
                    acc[b + 1].x += w * q.z;
                    acc[b + 1].y += w * q.w;
                }
            } else {
                const float2 q = z[0];
                acc[0].x += w * q.x;
                acc[0].y += w * q.y;
            }
        }
    }
    const float d = sdcf[sm];
    const int lanes = lanes_of(run, BCT, ch);
    const int64_t stride = chunk_stride(run, BCT);
#pragma unroll
    for (int b = 0; b < BCT; ++b) {
        if (b < lanes) {
            const int64_t o = S.dat_off + ((int64_t)ch * stride + b) * S.M + m;
            float f = d;
            if (run.W) f *= run.W[o];
            y[o] = make_float2(acc[b].x * f, acc[b].y * f);
        }
    }
}

template <int BCT>
__global__ __launch_bounds__(256) void nu_grid(const NuSliceDev* __restrict__ sl,
                                               const int4* __restrict__ work,
                                               const TileWork* __restrict__ tiles,
                                               const int* __restrict__ tlist,
                                               const uint32_t* __restrict__ sbase,
                                               const float* __restrict__ swt,
                                               const float* __restrict__ sdcf,
                                               const float2* __restrict__ y, float2* __restrict__ ws,
                                               float2* __restrict__ parts, int* __restrict__ counters,
                                               NuRun run) {
    __shared__ float acc[kTile * kTile * BCT * 2];
    __shared__ int s_last;
    const int4 wk = work[blockIdx.x];
    const NuSliceDev& S = sl[wk.x];
    const int ch = wk.y;
    const TileWork tw = tiles[wk.z];
    const int G1 = S.G1, G2 = S.G2, W1 = S.W1, W2 = S.W2, M = S.M;
    for (int i = threadIdx.x; i < kTile * kTile * BCT * 2; i += blockDim.x) acc[i] = 0.f;
    __syncthreads();
    const int t1_0 = tw.t1 * kTile, t2_0 = tw.t2 * kTile;
    const int lanes = lanes_of(run, BCT, ch);
    const int64_t stride = chunk_stride(run, BCT);
    const int* list = tlist + S.tlist_off + tw.start;
    const int total = tw.count * BCT;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int b = idx & (BCT - 1);
        if (b >= lanes) continue;
        const int m = list[idx / BCT];
        const int64_t sm = S.samp_off + m;
        float2 v = y[S.dat_off + ((int64_t)ch * stride + b) * M + m];
        const float d = sdcf[sm];
        v.x *= d;
        v.y *= d;
        const uint32_t base = sbase[sm];
        const int g1_0 = base & 0xffff, g2_0 = base >> 16;
        const float* wt = swt + sm * (W1 + W2);
        for (int s2 = 0; s2 < W2; ++s2) {
            int g2 = g2_0 + s2;
            if (g2 >= G2) g2 -= G2;
            const int l2 = g2 - t2_0;
            if (l2 < 0 || l2 >= kTile) continue;
            const float w2 = wt[W1 + s2];
            for (int s1 = 0; s1 < W1; ++s1) {
                int g1 = g1_0 + s1;
                if (g1 >= G1) g1 -= G1;
                const int l1 = g1 - t1_0;
                if (l1 < 0 || l1 >= kTile) continue;
                const float w = wt[s1] * w2;
                float* a = acc + ((l2 * kTile + l1) * BCT + b) * 2;
                atomicAdd(a, w * v.x);
                atomicAdd(a + 1, w * v.y);
            }
        }
    }
    __syncthreads();
    float2* zg = ws + S.ws_off + (int64_t)ch * S.ws_chunk;
    const int cells = kTile * kTile * BCT;
    if (tw.slot < 0) {
        for (int i = threadIdx.x; i < cells; i += blockDim.x) {
            const int b = i & (BCT - 1);
            const int l1 = (i / BCT) % kTile, l2 = i / (BCT * kTile);
            const int g1 = t1_0 + l1, g2 = t2_0 + l2;
            if (g1 < G1 && g2 < G2)
                zg[((int64_t)g2 * G1 + g1) * BCT + b] = make_float2(acc[2 * i], acc[2 * i + 1]);
        }
        return;
    }
    // split tile: publish this part, the last arriving work-group sums all
    // parts in slot order and writes the tile
    float2* pbase = parts + S.part_off + (int64_t)ch * S.part_chunk;
    float2* mine = pbase + (int64_t)tw.slot * cells;
    for (int i = threadIdx.x; i < cells; i += blockDim.x) mine[i] = make_float2(acc[2 * i], acc[2 * i + 1]);
    __threadfence();   // every wave's partial is visible device-wide before the count
    __syncthreads();
    if (threadIdx.x == 0) {
        int* cnt = counters + S.cnt_off + (int64_t)ch * S.cnt_chunk + tw.counter;
        const int prev = __hip_atomic_fetch_add(cnt, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (prev == tw.nparts - 1);
        if (s_last) __hip_atomic_store(cnt, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    for (int i = threadIdx.x; i < cells; i += blockDim.x) {
        float2 s = make_float2(0.f, 0.f);
        for (int p = 0; p < tw.nparts; ++p) {
            const float2 q = pbase[(int64_t)(tw.first_slot + p) * cells + i];
            s.x += q.x;
            s.y += q.y;
        }
        const int b = i & (BCT - 1);
        const int l1 = (i / BCT) % kTile, l2 = i / (BCT * kTile);
        const int g1 = t1_0 + l1, g2 = t2_0 + l2;
        if (g1 < G1 && g2 < G2) zg[((int64_t)g2 * G1 + g1) * BCT + b] = s;
    }
}

template <int BCT>
__global__ __launch_bounds__(256) void nu_adj_rows(const NuSliceDev* __restrict__ sl,
                                                   const int4* __restrict__ work,
                                                   const FftFactors* __restrict__ facs,
                                                   const float2* __restrict__ twp,
                                                   const float* __restrict__ deap,
                                                   const float2* __restrict__ ws,
                                                   float2* __restrict__ x, NuRun run) {
    extern __shared__ float2 lds[];
    const int4 wk = work[blockIdx.x];
    const NuSliceDev& S = sl[wk.x];
    const int ch = wk.y, r0 = wk.z;
    const int N1 = S.N1, N2 = S.N2, G1 = S.G1, rpw = S.rpw;
    const int T = rpw * BCT, log2T = ilog2(T);
    const int64_t V = (int64_t)N1 * N2;
    float2* A = lds;
    float2* Bb = lds + G1 * T;
    float2* tw = lds + 2 * G1 * T;
    for (int i = threadIdx.x; i < G1; i += blockDim.x) tw[i] = twp[S.tw1 + i];
    const float2* g = ws + S.ws_off + (int64_t)ch * S.ws_chunk;
    const int total = T * G1;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int b = idx & (BCT - 1);
        const int rest = idx / BCT;
        const int g1 = rest % G1;
        const int r = rest / G1;
        const int i2 = r0 + r;
        A[(g1 * rpw + r) * BCT + b] =
            i2 < N2 ? g[((int64_t)i2 * G1 + g1) * BCT + b] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    const float2* R = fft_lds(A, Bb, facs[S.fac1], log2T, tw, -1.f);
    const int lanes = lanes_of(run, BCT, ch);
    if (run.sense) {
        const int tot = rpw * N1;
        for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
            const int i1 = idx % N1;
            const int r = idx / N1;
            const int i2 = r0 + r;
            if (i2 >= N2) continue;
            int pos = i1 - N1 / 2;
            if (pos < 0) pos += G1;
            const int64_t v = (int64_t)i2 * N1 + i1;
            const float2* src = R + (pos * rpw + r) * BCT;
            float2 s = make_float2(0.f, 0.f);
            for (int b = 0; b < lanes; ++b) {
                const float2 mp = run.maps[S.maps_off + (int64_t)b * V + v];
                const float2 q = src[b];
                s.x += mp.x * q.x + mp.y * q.y;   // conj(maps) * q
                s.y += mp.x * q.y - mp.y * q.x;
            }
            const float f = deap[S.deap1 + i1] * deap[S.deap2 + i2];
            x[S.img_off + (int64_t)ch * V + v] = make_float2(s.x * f, s.y * f);
        }
    } else {
        const int tot = T * N1;
        for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
            const int i1 = idx % N1;
            const int rest = idx / N1;
            const int b = rest & (BCT - 1);
            const int r = rest / BCT;
            const int i2 = r0 + r;
            if (i2 >= N2 || b >= lanes) continue;
            int pos = i1 - N1 / 2;
            if (pos < 0) pos += G1;
            const float2 q = R[(pos * rpw + r) * BCT + b];
            const float f = deap[S.deap1 + i1] * deap[S.deap2 + i2];
            x[S.img_off + ((int64_t)ch * BCT + b) * V + (int64_t)i2 * N1 + i1] =
                make_float2(q.x * f, q.y * f);
        }
    }
}

// ---------------------------------------------------------------------------
// host: plan
NufftOp::NufftOp(mm_ctx* c, int ndim, int nslice, const int64_t* imN, const int64_t* M,
                 const double* const* kloc, const double* const* dcf, double os, int kwidth) {
    ctx = c;
    if (ndim != 1 && ndim != 2)
        fail(MM_ERR_UNSUPPORTED, "nufftOp: only 1-D and 2-D images are implemented on gfx950");
    if (nslice < 1) fail(MM_ERR_ARG, "nufftOp: nslice must be >= 1");
    if (kwidth < 2 || kwidth > 8 || (kwidth & 1))
        fail(MM_ERR_ARG, "nufftOp: kwidth must be even and in [2, 8]");
    if (!(os > 1.0)) fail(MM_ERR_ARG, "nufftOp: os must be > 1");
    ndim_ = ndim;
    W_ = kwidth;
    os_ = os;
    beta = nufft_kb_beta(kwidth, os);
    nkern = nufft_kb_nkern(kwidth, os);
    const std::vector<double> lut = kb_lut(kwidth, os);
    const int Nk = nkern;

    std::vector<FftFactors> facs;
    std::vector<float2> twid;
    std::map<int, std::pair<int, int>> fft_of;  // G -> (factor idx, twiddle offset)
    auto fft_for = [&](int G) {
        auto it = fft_of.find(G);
        if (it != fft_of.end()) return it->second;
        FftFactors f;
        if (!fft_factorize(G, f)) fail(MM_ERR_ARG, "nufftOp: grid size with a prime factor > 7");
        std::pair<int, int> r((int)facs.size(), (int)twid.size());
        facs.push_back(f);
        for (int t = 0; t < G; ++t)
            twid.push_back(make_float2((float)std::cos(2 * M_PI * t / G), (float)std::sin(2 * M_PI * t / G)));
        fft_of[G] = r;
        return r;
    };
    std::vector<float> deap;
    std::vector<uint32_t> sbase;
    std::vector<float> swt, sdcf;
    std::vector<int> tlist;
    int64_t voff = 0, moff = 0;
    for (int s = 0; s < nslice; ++s) {
        NuSliceDev d{};
        const int64_t n1 = imN[2 * s], n2 = ndim == 2 ? imN[2 * s + 1] : 1;
        if (n1 < 2 || n2 < 1 || (ndim == 2 && n2 < 2)) fail(MM_ERR_ARG, "nufftOp: image size too small");
        if ((n1 & 1) || (ndim == 2 && (n2 & 1)))
            fail(MM_ERR_UNSUPPORTED, "nufftOp: odd image sizes are not implemented on gfx950");
        if (M[s] < 1) fail(MM_ERR_ARG, "nufftOp: no samples");
        if (!kloc || !kloc[s]) fail(MM_ERR_ARG, "nufftOp: kloc is NULL");
        d.N1 = (int)n1;
        d.N2 = (int)n2;
        d.G1 = nufft_grid_size(n1, os);
        d.G2 = ndim == 2 ? nufft_grid_size(n2, os) : 1;
        if (d.G1 > 65535 || d.G2 > 65535) fail(MM_ERR_ARG, "nufftOp: grid too large");
        d.M = (int)M[s];
        d.W1 = kwidth;
        d.W2 = ndim == 2 ? kwidth : 1;
        auto f1 = fft_for(d.G1);
        auto f2 = fft_for(d.G2);
        d.fac1 = f1.first;
        d.tw1 = f1.second;
        d.fac2 = f2.first;
        d.tw2 = f2.second;
        // deapodisation with 1/sqrt(G1*G2) and the ramp constant folded in
        const double sgn = (((n1 / 2) + (ndim == 2 ? n2 / 2 : 0)) & 1) ? -1.0 : 1.0;
        const double scl = sgn / std::sqrt((double)d.G1 * d.G2);
        auto c1 = deapod((int)n1, d.G1, kwidth, os);
        d.deap1 = (int)deap.size();
        for (double v : c1) deap.push_back((float)(v * scl));
        d.deap2 = (int)deap.size();
        if (ndim == 2) {
            auto c2 = deapod((int)n2, d.G2, kwidth, os);
            for (double v : c2) deap.push_back((float)v);
        } else {
            deap.push_back(1.f);
        }
        // samples (nufftOp.m:187-246)
        d.samp_off = (int64_t)sbase.size();
        const double* kl = kloc[s];
        const double* dc = dcf ? dcf[s] : nullptr;
        const int Ws = d.W1 + d.W2;
        swt.resize(swt.size() + (size_t)d.M * Ws);
        float* wdst = swt.data() + (size_t)d.samp_off * Ws;
        std::vector<uint16_t> g0[2];
        g0[0].resize(d.M);
        g0[1].resize(d.M, 0);
        for (int m = 0; m < d.M; ++m) {
            for (int dim = 0; dim < ndim; ++dim) {
                const int G = dim == 0 ? d.G1 : d.G2;
                const double ka = (kl[(int64_t)dim * d.M + m] + 0.5) * G;
                if (!std::isfinite(ka)) fail(MM_ERR_ARG, "nufftOp: non-finite kloc");
                const double fl = std::floor(ka);
                const double fr = fl - ka;
                for (int si = 0; si < kwidth; ++si) {
                    const int soff = -kwidth / 2 + 1 + si;
                    double ki = (fr + soff) * (Nk - 1) / kwidth + Nk / 2.0 + 0.5;
                    const double fk = std::floor(ki);
                    const double frac = ki - fk;
                    int ii = (int)fk;  // 1-based
                    ii = std::min(std::max(ii, 1), Nk);
                    const double w = (1 - frac) * lut[ii - 1] + frac * lut[ii];
                    wdst[(size_t)m * Ws + (dim == 0 ? 0 : d.W1) + si] = (float)w;
                }
                int64_t gfirst = (int64_t)fl - kwidth / 2 + 1 + G / 2;
                gfirst %= G;
                if (gfirst < 0) gfirst += G;
                g0[dim][m] = (uint16_t)gfirst;
            }
            if (ndim == 1) wdst[(size_t)m * Ws + d.W1] = 1.f;
            sbase.push_back((uint32_t)g0[0][m] | ((uint32_t)g0[1][m] << 16));
            sdcf.push_back(dc ? (float)dc[m] : 1.f);
        }
        // gridding tiles: each sample joins every tile its footprint touches
        const int nt1 = (d.G1 + kTile - 1) / kTile, nt2 = (d.G2 + kTile - 1) / kTile;
        std::vector<std::vector<int>> lists((size_t)nt1 * nt2);
        for (int m = 0; m < d.M; ++m) {
            int tl1[8], tl2[8], n1t = 0, n2t = 0;
            for (int si = 0; si < d.W1; ++si) {
                int g = g0[0][m] + si;
                if (g >= d.G1) g -= d.G1;
                const int t = g / kTile;
                bool seen = false;
                for (int q = 0; q < n1t; ++q) seen |= tl1[q] == t;
                if (!seen) tl1[n1t++] = t;
            }
            for (int si = 0; si < d.W2; ++si) {
                int g = g0[1][m] + si;
                if (g >= d.G2) g -= d.G2;
                const int t = g / kTile;
                bool seen = false;
                for (int q = 0; q < n2t; ++q) seen |= tl2[q] == t;
                if (!seen) tl2[n2t++] = t;
            }
            for (int a = 0; a < n2t; ++a)
                for (int b = 0; b < n1t; ++b) lists[(size_t)tl2[a] * nt1 + tl1[b]].push_back(m);
        }
        tlist_off_.push_back((int64_t)tlist.size());
        d.tlist_off = tlist_off_.back();
        tw_first_.push_back((int)tw_host_.size());
        int nslots = 0, ncnt = 0;
        int64_t lstart = 0;
        for (int t2 = 0; t2 < nt2; ++t2)
            for (int t1 = 0; t1 < nt1; ++t1) {
                const auto& L = lists[(size_t)t2 * nt1 + t1];
                const int cnt = (int)L.size();
                const int np = std::max(1, (cnt + kTileChunk - 1) / kTileChunk);
                const int first = np > 1 ? nslots : -1;
                const int ctr = np > 1 ? ncnt++ : -1;
                for (int p = 0; p < np; ++p) {
                    TileWork w{};
                    w.t1 = t1;
                    w.t2 = t2;
                    w.start = (int)(lstart + (int64_t)p * kTileChunk);
                    w.count = std::min(kTileChunk, cnt - p * kTileChunk);
                    if (w.count < 0) w.count = 0;
                    w.nparts = np;
                    w.slot = np > 1 ? first + p : -1;
                    w.first_slot = first;
                    w.counter = ctr;
                    tw_host_.push_back(w);
                }
                if (np > 1) nslots += np;
                tlist.insert(tlist.end(), L.begin(), L.end());
                lstart += cnt;
            }
        tw_count_.push_back((int)tw_host_.size() - tw_first_.back());
        nslots_.push_back(nslots);
        ncounters_.push_back(ncnt);
        d.tile_first = tw_first_.back();
        sl_.push_back(d);
        Vs_.push_back(n1 * n2);
        Ms_.push_back(d.M);
        img_start_.push_back(voff);
        dat_start_.push_back(moff);
        voff += n1 * n2;
        moff += d.M;
    }
    V_total_ = voff;
    M_total_ = moff;
    hipStream_t st = ctx->stream;
    facs_.upload(facs, st);
    twid_.upload(twid, st);
    deap_.upload(deap, st);
    sbase_.upload(sbase, st);
    swt_.upload(swt, st);
    sdcf_.upload(sdcf, st);
    tlist_.upload(tlist.empty() ? std::vector<int>{0} : tlist, st);
    tiles_.upload(tw_host_, st);
    HIPCHECK(hipStreamSynchronize(st));
}

static size_t fft_lds_bytes(int G, int T) { return (size_t)(2 * G * T + G) * sizeof(float2); }

NufftOp::Cfg& NufftOp::config(int64_t nrep, int sense, int C) {
    auto key = std::make_pair(nrep, sense ? C : 0);
    auto it = cfgs_.find(key);
    if (it != cfgs_.end()) return *it->second;
    auto cp = std::make_unique<Cfg>();
    Cfg& c = *cp;
    c.nrep = nrep;
    c.sense = sense;
    c.C = C;
    int nchunk;
    if (sense) {
        if (C < 1 || C > 16) fail(MM_ERR_UNSUPPORTED, "fused SENSE NUFFT supports 1..16 coils");
        c.BCT = next_pow2(C);
        nchunk = (int)nrep;
    } else {
        c.BCT = std::min<int>(16, next_pow2(nrep));
        nchunk = (int)((nrep + c.BCT - 1) / c.BCT);
    }
    const int BCT = c.BCT;
    const int ns = (int)sl_.size();
    c.hs = sl_;
    int64_t wsoff = 0, poff = 0, coff = 0;
    std::vector<int4> wr, wc, wd, wg;
    for (int s = 0; s < ns; ++s) {
        NuSliceDev& d = c.hs[s];
        // rows / columns per work-group: aim for 8-16 interleaved transforms
        // within 64 KB of LDS
        int rpw = std::max(1, 8 / BCT);
        while (rpw > 1 && fft_lds_bytes(d.G1, rpw * BCT) > 65536) rpw >>= 1;
        int cpw = std::max(1, 16 / BCT);
        while (cpw > 1 && fft_lds_bytes(d.G2, cpw * BCT) > 65536) cpw >>= 1;
        d.rpw = rpw;
        d.cpw = cpw;
        c.lds_rows = std::max(c.lds_rows, fft_lds_bytes(d.G1, rpw * BCT));
        c.lds_cols = std::max(c.lds_cols, fft_lds_bytes(d.G2, cpw * BCT));
        if (sense) {
            d.img_off = img_start_[s] * nrep;
            d.maps_off = img_start_[s] * C;
            d.dat_off = dat_start_[s] * C * nrep;
        } else {
            d.img_off = img_start_[s] * nrep;
            d.maps_off = 0;
            d.dat_off = dat_start_[s] * nrep;
        }
        d.ws_chunk = (int64_t)d.G1 * d.G2 * BCT;
        d.ws_off = wsoff;
        wsoff += d.ws_chunk * nchunk;
        d.part_chunk = (int64_t)nslots_[s] * kTile * kTile * BCT;
        d.part_off = poff;
        poff += d.part_chunk * nchunk;
        d.cnt_chunk = ncounters_[s];
        d.cnt_off = coff;
        coff += d.cnt_chunk * nchunk;
        for (int ch = 0; ch < nchunk; ++ch) {
            for (int r = 0; r < d.N2; r += rpw) wr.push_back(make_int4(s, ch, r, 0));
            if (d.G2 > 1)
                for (int q = 0; q < d.G1; q += cpw) wc.push_back(make_int4(s, ch, q, 0));
            for (int m = 0; m < d.M; m += 256) wd.push_back(make_int4(s, ch, m, 0));
            for (int t = 0; t < tw_count_[s]; ++t)
                wg.push_back(make_int4(s, ch, tw_first_[s] + t, tw_host_[tw_first_[s] + t].count));
        }
    }
    // heaviest gridding items first
    std::stable_sort(wg.begin(), wg.end(), [](const int4& a, const int4& b) { return a.w > b.w; });
    hipStream_t st = ctx->stream;
    c.sl.upload(c.hs, st);
    c.w_rows.upload(wr, st);
    c.w_cols.upload(wc, st);
    c.w_deg.upload(wd, st);
    c.w_grid.upload(wg, st);
    c.n_rows = (int)wr.size();
    c.n_cols = (int)wc.size();
    c.n_deg = (int)wd.size();
    c.n_grid = (int)wg.size();
    c.ws.alloc(std::max<int64_t>(wsoff, 1));
    c.parts.alloc(std::max<int64_t>(poff, 1));
    c.counters.alloc(std::max<int64_t>(coff, 1));
    HIPCHECK(hipMemsetAsync(c.counters.p, 0, sizeof(int) * std::max<int64_t>(coff, 1), st));
    HIPCHECK(hipStreamSynchronize(st));
    cfgs_[key] = std::move(cp);
    return c;
}

template <int BCT>
static void launch_all(int adjoint, hipStream_t st, const NuRun& run, const float2* in, float2* out,
                       const NuSliceDev* sl, int ndim, const int4* w_rows, int n_rows, size_t lds_rows,
                       const int4* w_cols, int n_cols, size_t lds_cols, const int4* w_deg, int n_deg,
                       const int4* w_grid, int n_grid, const FftFactors* facs, const float2* twid,
                       const float* deap, const uint32_t* sbase, const float* swt, const float* sdcf,
                       const TileWork* tiles, const int* tlist, float2* ws, float2* parts, int* counters) {
    static bool attr_set = false;
    if (!attr_set) {
        // allow FFT passes above 64 KB of LDS where the device permits it
        (void)hipFuncSetAttribute((const void*)nu_fwd_rows<BCT>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void*)nu_adj_rows<BCT>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void*)nu_cols<BCT>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipGetLastError();
        attr_set = true;
    }
    if (!adjoint) {
        hipLaunchKernelGGL(nu_fwd_rows<BCT>, dim3(n_rows), dim3(256), lds_rows, st, sl, w_rows, facs,
                           twid, deap, in, ws, run);
        LAUNCHCHECK();
        if (ndim == 2 && n_cols > 0) {
            hipLaunchKernelGGL(nu_cols<BCT>, dim3(n_cols), dim3(256), lds_cols, st, sl, w_cols, facs,
                               twid, ws, 0);
            LAUNCHCHECK();
        }
        hipLaunchKernelGGL(nu_degrid<BCT>, dim3(n_deg), dim3(256), 0, st, sl, w_deg, sbase, swt, sdcf,
                           (const float2*)ws, out, run);
        LAUNCHCHECK();
    } else {
        hipLaunchKernelGGL(nu_grid<BCT>, dim3(n_grid), dim3(256), 0, st, sl, w_grid, tiles, tlist,
                           sbase, swt, sdcf, in, ws, parts, counters, run);
        LAUNCHCHECK();
        if (ndim == 2 && n_cols > 0) {
            hipLaunchKernelGGL(nu_cols<BCT>, dim3(n_cols), dim3(256), lds_cols, st, sl, w_cols, facs,
                               twid, ws, 1);
            LAUNCHCHECK();
        }
        hipLaunchKernelGGL(nu_adj_rows<BCT>, dim3(n_rows), dim3(256), lds_rows, st, sl, w_rows, facs,
                           twid, deap, (const float2*)ws, out, run);
        LAUNCHCHECK();
    }
}

void NufftOp::run(Cfg& c, int adjoint, const float2* in, float2* out, const float2* maps,
                  const float* W) {
    NuRun r;
    r.sense = c.sense;
    r.C = c.C;
    r.B = c.nrep;
    r.maps = maps;
    r.W = adjoint ? nullptr : W;
    hipStream_t st = ctx->stream;
#define MM_NU_LAUNCH(BC)                                                                          \
    launch_all<BC>(adjoint, st, r, in, out, c.sl.p, ndim_, c.w_rows.p, c.n_rows, c.lds_rows,      \
                   c.w_cols.p, c.n_cols, c.lds_cols, c.w_deg.p, c.n_deg, c.w_grid.p, c.n_grid,    \
                   facs_.p, twid_.p, deap_.p, sbase_.p, swt_.p, sdcf_.p, tiles_.p, tlist_.p,       \
                   c.ws.p, c.parts.p, c.counters.p)
    switch (c.BCT) {
        case 1: MM_NU_LAUNCH(1); break;
        case 2: MM_NU_LAUNCH(2); break;
        case 4: MM_NU_LAUNCH(4); break;
        case 8: MM_NU_LAUNCH(8); break;
        default: MM_NU_LAUNCH(16); break;
    }
#undef MM_NU_LAUNCH
}

void NufftOp::apply(int adjoint, const float2* in, float2* out, int64_t nrep) {
    if (nrep < 1) fail(MM_ERR_ARG, "nufftOp: nrep must be >= 1");
    run(config(nrep, 0, 0), adjoint, in, out, nullptr, nullptr);
}

void NufftOp::apply_sense(int adjoint, const float2* in, float2* out, int64_t nrep,
                          const float2* maps, int C, const float* Wt) {
    if (nrep < 1) fail(MM_ERR_ARG, "nufftOp: nrep must be >= 1");
    run(config(nrep, 1, C), adjoint, in, out, maps, Wt);
}

}  // namespace mm

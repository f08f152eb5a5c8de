// 3-D Kaiser-Bessel gridding NUFFT for gfx950 (nufftOp with a 3-D imN).
//
// Plan (host, FP64, once): the closed forms of nufftOp.doPrep
// (operators/@nufftOp/nufftOp.m:138-274) for three sides -- grid sizes,
// KB LUT, deapodisation per side, W weights per side and sample by linear
// LUT interpolation (:216-227), wrapped first grid index per side
// (:230-246) -- and the transposed interpolation matrix as CSR rows per grid
// cell, (sample, w1 w2 w3 dcf) in sample order.
//
// Forward (:378-394), per chunk of BCT transforms:
//   n3_rows_fwd   coil image times deapodisation, wrapped into the
//                 zero-padded row, FFT along g1 (rows i2 < N2 of planes i3 < N3)
//   n3_axis<0>    FFT along g2 (input rows wrapped by floor(N2/2)), then
//                 along g3 (input planes wrapped by floor(N3/2)), in place
//   n3_odd_ramp   odd sides only: the ramp factor that is not a shift
//   n3_degrid     W^3 taps per (sample, transform), dcf; transposed through
//                 LDS so the data side is written sample-fastest
// Adjoint (:354-377): a transpose to [M][BCT] (n3_transpose_y), a
// deterministic gather per grid cell over its CSR row with the dcf in the
// row weights (n3_grid_csr), the conjugate ramp, inverse FFTs along g3 and
// g2 that keep only the N3 / N2 outputs the crop reads, and the row FFT with
// crop and deapodisation (n3_rows_adj).
//
// Grid layout in HBM: Z[g3][g2][g1][b], b = transform within the chunk, so a
// tap and every FFT position of a pencil group touch BCT (or T) consecutive
// complex values.  The ramp fftShiftArray (:258-273) of an even side is the
// circular shift by N/2 times (-1)^(N/2) (the wrapped placement and a sign
// in the deapodisation); an odd side N = 2s + 1 is placed wrapped by s and
// the grid takes i exp(-i pi g / G) per odd side (n3_odd_ramp).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

#include "nufft.hpp"
#include "nufft3.hpp"

namespace mm {

// transforms per work-group of the g2 / g3 passes: a chunk's BCT
// transforms of T / BCT pencils (16 transforms for chunks of 8 or 16: each
// grid position a whole 128-byte line, one work-group per line; else 8)
template <int BCT>
__host__ __device__ constexpr int n3_T() {
    return BCT >= 8 ? 16 : 8;
}
// ... and of the row passes (16 only for 16-transform chunks: two rows of 8
// transforms per work-group measured slower than one, 64 against 55 us)
template <int BCT>
__host__ __device__ constexpr int n3_Tr() {
    return BCT >= 16 ? 16 : 8;
}

__device__ __forceinline__ int wrap_pos(int i, int s, int n) {
    const int p = i - s;
    return p < 0 ? p + n : p;
}

// Values a thread holds in registers in a single-buffer pass: the default
// cap (16 at 16 transforms, 8 at 8), or, in the kernels for sides the
// default cannot hold in one LDS buffer (MID = 1: above 256 points at 16
// transforms, 512 at 8), twice that, so every pass over a side of up to 512
// points fits one buffer unless its radices leave the butterflies short of
// it (420, 486..504: then two buffers).  Those kernels of 16 transforms take
// up to 168 registers (3 waves per SIMD), of 8 up to 128.
template <int T, int MID>
__host__ __device__ constexpr int n3_cap() {
    return MID ? 2 * fft_reg_cap(T, 0) : fft_reg_cap(T, 0);
}
template <int T, int MID>
__host__ __device__ constexpr int n3_waves() {
    return MID ? (T >= 16 ? 3 : 4) : 1;
}

// LDS of a pass over sides of length n: one (single-buffer passes) or two
// buffers of n*T values and the n - 1 pass twiddles
static size_t n3_lds(int n, int T, bool ip) {
    return (size_t)((ip ? 1 : 2) * n * T + n) * sizeof(float2);
}

// Rows, forward: lines L = i2 + N2 i3 (i2 < N2, i3 < N3), T / BCT of them per
// work-group.  Image reads run along i1 (coalesced), grid writes along
// (g1, b).
template <int BCT, int MID>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(n3_waves<n3_Tr<BCT>(), MID>(), 8))) void n3_rows_fwd(Nu3Geom g, const FftFactors* __restrict__ fac,
                                                   const float2* __restrict__ twp,
                                                   const float* __restrict__ deap,
                                                   const float2* __restrict__ x, float2* __restrict__ ws,
                                                   int ch, int lanes) {
    constexpr int T = n3_Tr<BCT>(), LPG = T / BCT;
    extern __shared__ float2 lds[];
    const FftFactors& F = *fac;
    const int n = g.G1;
    const bool ip = g.ip & 1;
    float2* A = lds;
    float2* B = ip ? nullptr : lds + n * T;
    float2* tw = lds + (ip ? 1 : 2) * n * T;
    const int L0 = (int)blockIdx.x * LPG;
    const int nl = g.N2 * g.N3;
    for (int i = threadIdx.x; i < n * T; i += kFftThreads) A[i] = make_float2(0.f, 0.f);
    for (int i = threadIdx.x; i < n - 1; i += kFftThreads) tw[i] = twp[i];
    __syncthreads();
    const float* d1 = deap;
    const float* d2 = deap + g.N1;
    const float* d3 = d2 + g.N2;
    for (int idx = threadIdx.x; idx < g.N1 * T; idx += kFftThreads) {
        const int i1 = idx % g.N1, t = idx / g.N1;
        const int b = t % BCT, L = L0 + t / BCT;
        if (L >= nl || b >= lanes) continue;
        const int i2 = L % g.N2, i3 = L / g.N2;
        const float2 v = x[(int64_t)(ch * BCT + b) * g.V + (int64_t)L * g.N1 + i1];
        const float f = d1[i1] * d2[i2] * d3[i3];
        A[wrap_pos(i1, g.s1, n) * T + t] = make_float2(v.x * f, v.y * f);
    }
    __syncthreads();
    const float2* R = fft_lds<T, 1, false, false, n3_cap<T, MID>()>(A, B, F, tw);
    for (int idx = threadIdx.x; idx < n * T; idx += kFftThreads) {
        const int b = idx % BCT, q = (idx / BCT) % n, j = idx / (BCT * n);
        const int L = L0 + j;
        if (L >= nl) continue;
        const int i2 = L % g.N2, i3 = L / g.N2;
        ws[(((int64_t)i3 * g.G2 + i2) * g.G1 + q) * BCT + b] = R[q * T + j * BCT + b];
    }
}

// Global -> LDS staging with U loads in flight per thread: element i < tot
// is ld(i) stored by st(i, v); the U loads of a round are issued before any
// of their stores, so a thread does not wait on each load in turn.
// loads in flight per thread in the axis passes' staging (against one at a
// time on the nufft3 workload: g3 pass 126 -> 98 us adjoint, 103 -> 86
// forward; 4 gave 115 / 88, 12 the same as 8).  The row passes keep one
// load at a time: staged, they measured slower (54 -> 57 us adjoint with
// eight in flight, 57 -> 59 forward with four).
constexpr int kN3StageU = 8;
template <int U, class Ld, class St>
__device__ inline void n3_stage(int tot, Ld ld, St st) {
    for (int i0 = threadIdx.x; i0 < tot; i0 += U * kFftThreads) {
        float2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * kFftThreads;
            v[u] = i < tot ? ld(i) : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * kFftThreads;
            if (i < tot) st(i, v[u]);
        }
    }
}

// FFT along g2 (ax = 2: pencils (g1, i3), i3 < N3) or g3 (ax = 3: pencils
// (g1, g2)), in place, T / BCT consecutive g1 per work-group so every
// position is T consecutive values.  Forward: inputs q < N_ax placed at
// (q - s) mod G, all G outputs.  Adjoint: all G inputs, outputs q < N_ax
// read from (q - s) mod G.  A work-group reads its pencils whole (into LDS)
// before writing any of them.
// TX = 2: twice the pencils per work-group (the g3 pass where its side fits
// one LDS buffer at 32 transforms: 256 bytes per grid position, the
// positions being a whole g1-g2 plane apart)
template <int BCT, int ADJ, int MID, int TX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(n3_waves<n3_T<BCT>() * TX, MID>(), 8))) void n3_axis(Nu3Geom g, int ax, const FftFactors* __restrict__ fac,
                                               const float2* __restrict__ twp, float2* __restrict__ ws) {
    constexpr int T = n3_T<BCT>() * TX, LPG = T / BCT, SGN = ADJ ? -1 : 1;
    extern __shared__ float2 lds[];
    const FftFactors& F = *fac;
    const int n = ax == 2 ? g.G2 : g.G3;
    const int na = ax == 2 ? g.N2 : g.N3;
    const int s = ax == 2 ? g.s2 : g.s3;
    const int gx = (g.G1 + LPG - 1) / LPG;
    const int gb = (int)blockIdx.x % gx, outer = (int)blockIdx.x / gx;
    const int g1_0 = gb * LPG;
    const int nt = min(T, (g.G1 - g1_0) * BCT);
    int64_t base, qs;
    if (ax == 2) {
        base = ((int64_t)outer * g.G2 * g.G1 + g1_0) * BCT;
        qs = (int64_t)g.G1 * BCT;
    } else {
        base = ((int64_t)outer * g.G1 + g1_0) * BCT;
        qs = (int64_t)g.G2 * g.G1 * BCT;
    }
    const bool ip = (g.ip >> (ax - 1)) & 1;
    float2* A = lds;
    float2* B = ip ? nullptr : lds + n * T;
    float2* tw = lds + (ip ? 1 : 2) * n * T;
    for (int i = threadIdx.x; i < n - 1; i += kFftThreads) tw[i] = twp[i];
    if (!ADJ) {
        for (int i = threadIdx.x; i < n * T; i += kFftThreads) A[i] = make_float2(0.f, 0.f);
        __syncthreads();
        n3_stage<kN3StageU>(
            na * T,
            [&](int idx) {
                const int t = idx % T, q = idx / T;
                return t < nt ? ws[base + q * qs + t] : make_float2(0.f, 0.f);
            },
            [&](int idx, float2 v) {
                const int t = idx % T, q = idx / T;
                if (t < nt) A[wrap_pos(q, s, n) * T + t] = v;
            });
    } else {
        n3_stage<kN3StageU>(
            n * T,
            [&](int idx) {
                const int t = idx % T, q = idx / T;
                return t < nt ? ws[base + q * qs + t] : make_float2(0.f, 0.f);
            },
            [&](int idx, float2 v) { A[idx] = v; });
    }
    __syncthreads();
    const float2* R = fft_lds<T, SGN, false, false, n3_cap<T, MID>()>(A, B, F, tw);
    if (!ADJ) {
        for (int idx = threadIdx.x; idx < n * T; idx += kFftThreads) {
            const int t = idx % T, q = idx / T;
            if (t < nt) ws[base + q * qs + t] = R[idx];
        }
    } else {
        for (int idx = threadIdx.x; idx < na * T; idx += kFftThreads) {
            const int t = idx % T, q = idx / T;
            if (t < nt) ws[base + q * qs + t] = R[wrap_pos(q, s, n) * T + t];
        }
    }
}

// Rows, adjoint: inverse FFT along g1 of rows i2 < N2 of planes i3 < N3,
// crop (pixel i1 from (i1 - s1) mod G1) and deapodisation.
template <int BCT, int MID>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(n3_waves<n3_Tr<BCT>(), MID>(), 8))) void n3_rows_adj(Nu3Geom g, const FftFactors* __restrict__ fac,
                                                   const float2* __restrict__ twp,
                                                   const float* __restrict__ deap,
                                                   const float2* __restrict__ ws, float2* __restrict__ x,
                                                   int ch, int lanes) {
    constexpr int T = n3_Tr<BCT>(), LPG = T / BCT;
    extern __shared__ float2 lds[];
    const FftFactors& F = *fac;
    const int n = g.G1;
    const bool ip = g.ip & 1;
    float2* A = lds;
    float2* B = ip ? nullptr : lds + n * T;
    float2* tw = lds + (ip ? 1 : 2) * n * T;
    const int L0 = (int)blockIdx.x * LPG;
    const int nl = g.N2 * g.N3;
    for (int i = threadIdx.x; i < n - 1; i += kFftThreads) tw[i] = twp[i];
    for (int idx = threadIdx.x; idx < n * T; idx += kFftThreads) {
        const int b = idx % BCT, q = (idx / BCT) % n, j = idx / (BCT * n);
        const int L = L0 + j;
        float2 v = make_float2(0.f, 0.f);
        if (L < nl) {
            const int i2 = L % g.N2, i3 = L / g.N2;
            v = ws[(((int64_t)i3 * g.G2 + i2) * g.G1 + q) * BCT + b];
        }
        A[q * T + j * BCT + b] = v;
    }
    __syncthreads();
    const float2* R = fft_lds<T, -1, false, false, n3_cap<T, MID>()>(A, B, F, tw);
    const float* d1 = deap;
    const float* d2 = deap + g.N1;
    const float* d3 = d2 + g.N2;
    for (int idx = threadIdx.x; idx < g.N1 * T; idx += kFftThreads) {
        const int i1 = idx % g.N1, t = idx / g.N1;
        const int b = t % BCT, L = L0 + t / BCT;
        if (L >= nl || b >= lanes) continue;
        const int i2 = L % g.N2, i3 = L / g.N2;
        const float2 v = R[wrap_pos(i1, g.s1, n) * T + t];
        const float f = d1[i1] * d2[i2] * d3[i3];
        x[(int64_t)(ch * BCT + b) * g.V + (int64_t)L * g.N1 + i1] = make_float2(v.x * f, v.y * f);
    }
}

// Odd sides: i exp(-i pi g_d / G_d) per odd side d (CONJ: the conjugate).
template <int BCT, int CONJ>
__global__ __launch_bounds__(256) void n3_odd_ramp(Nu3Geom g, float2* __restrict__ ws) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= g.chunk) return;
    const int64_t cell = e / BCT;
    const int g1 = (int)(cell % g.G1);
    const int64_t r = cell / g.G1;
    const int g2 = (int)(r % g.G2), g3 = (int)(r / g.G2);
    float a = 0.f;   // half-turns
    if (g.odd & 1) a += 0.5f - (float)g1 / (float)g.G1;
    if (g.odd & 2) a += 0.5f - (float)g2 / (float)g.G2;
    if (g.odd & 4) a += 0.5f - (float)g3 / (float)g.G3;
    float sn, cs;
    sincospif(CONJ ? -a : a, &sn, &cs);
    const float2 v = ws[e];
    ws[e] = make_float2(v.x * cs - v.y * sn, v.x * sn + v.y * cs);
}

// Toeplitz middle stage: the grid of one chunk times the kernel spectrum
// (one value per cell, shared by the chunk's BCT transforms).
template <int BCT>
__global__ __launch_bounds__(256) void n3_toep_mul(Nu3Geom g, const float2* __restrict__ km,
                                                   float2* __restrict__ ws) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= g.chunk) return;
    const float2 k = km[e / BCT];
    const float2 v = ws[e];
    ws[e] = make_float2(v.x * k.x - v.y * k.y, v.x * k.y + v.y * k.x);
}

// Degridding: lane = (sample, transform), transform fastest, so the BCT
// lanes of a sample read each tap's BCT consecutive values (four transforms
// per lane with 16-byte loads measured slower, 173 against 131 us: fewer
// waves for loads that do not depend on each other); W^3 taps in (s3, s2,
// s1) order with weight w3 w2 w1, then dcf (WT = W at compile time, 0:
// runtime W).  The block's results go through LDS and are written
// sample-fastest.
template <int BCT, int WT>
__global__ __launch_bounds__(256) void n3_degrid(Nu3Geom g, const uint32_t* __restrict__ base,
                                                 const float* __restrict__ wt, const float* __restrict__ dcf,
                                                 const float2* __restrict__ ws, float2* __restrict__ y, int ch,
                                                 int lanes) {
    constexpr int TPL = 1, LPS = BCT / TPL, SPB = 256 / LPS;
    __shared__ float2 st[SPB * BCT];
    const int b0 = (int)(threadIdx.x % LPS) * TPL, ml = threadIdx.x / LPS;
    const int m = (int)blockIdx.x * SPB + ml;
    float2 acc[TPL];
#pragma unroll
    for (int t = 0; t < TPL; ++t) acc[t] = make_float2(0.f, 0.f);
    if (m < g.M) {
        const uint32_t bs = base[m];
        const int g10 = (int)(bs & 1023u), g20 = (int)((bs >> 10) & 1023u), g30 = (int)(bs >> 20);
        const int W = WT > 0 ? WT : g.W;
        const float* w = wt + (int64_t)m * 3 * W;
        int c3 = g30;
        for (int s3 = 0; s3 < W; ++s3) {
            const float w3 = w[2 * W + s3];
            int c2 = g20;
            for (int s2 = 0; s2 < W; ++s2) {
                const float w23 = w3 * w[W + s2];
                const float2* row = ws + (((int64_t)c3 * g.G2 + c2) * g.G1) * BCT + b0;
                int c1 = g10;
#pragma unroll
                for (int s1 = 0; s1 < (WT > 0 ? WT : 8); ++s1) {
                    if (WT == 0 && s1 >= W) break;
                    const float wk = w23 * w[s1];
                    const float2* z = row + (int64_t)c1 * BCT;
                    if constexpr (TPL >= 2) {
#pragma unroll
                        for (int t = 0; t < TPL; t += 2) {
                            const float4 q = *reinterpret_cast<const float4*>(z + t);
                            acc[t].x += wk * q.x;
                            acc[t].y += wk * q.y;
                            acc[t + 1].x += wk * q.z;
                            acc[t + 1].y += wk * q.w;
                        }
                    } else {
                        acc[0].x += wk * z[0].x;
                        acc[0].y += wk * z[0].y;
                    }
                    if (++c1 == g.G1) c1 = 0;
                }
                if (++c2 == g.G2) c2 = 0;
            }
            if (++c3 == g.G3) c3 = 0;
        }
        const float d = dcf[m];
#pragma unroll
        for (int t = 0; t < TPL; ++t) {
            acc[t].x *= d;
            acc[t].y *= d;
        }
    }
#pragma unroll
    for (int t = 0; t < TPL; ++t) st[(b0 + t) * SPB + ml] = acc[t];
    __syncthreads();
    for (int i = threadIdx.x; i < SPB * BCT; i += 256) {
        const int b2 = i / SPB, m2 = (int)blockIdx.x * SPB + i % SPB;
        if (m2 < g.M && b2 < lanes) y[(int64_t)(ch * BCT + b2) * g.M + m2] = st[i];
    }
}

// y of one chunk into the transposed layout yT[m][b] (read sample-fastest,
// written transform-fastest through LDS); the CSR weights carry the dcf.
template <int BCT>
__global__ __launch_bounds__(256) void n3_transpose_y(Nu3Geom g, const float2* __restrict__ y,
                                                      float2* __restrict__ yT, int ch, int lanes) {
    constexpr int SPB = 256 / BCT;
    __shared__ float2 st[256];
    const int b = threadIdx.x / SPB, ml = threadIdx.x % SPB;
    const int m = (int)blockIdx.x * SPB + ml;
    float2 v = make_float2(0.f, 0.f);
    if (m < g.M && b < lanes) v = y[(int64_t)(ch * BCT + b) * g.M + m];
    st[threadIdx.x] = v;
    __syncthreads();
    const int b2 = threadIdx.x % BCT, ml2 = threadIdx.x / BCT, m2 = (int)blockIdx.x * SPB + ml2;
    if (m2 < g.M) yT[(int64_t)m2 * BCT + b2] = st[b2 * SPB + ml2];
}

// Rows longer than this go to n3_grid_long (one work-group per cell).
constexpr int kN3LongRow = 64;
// transforms per lane of n3_grid_csr (eight measured slower than four: 219
// against 163 us on a 192^2 x 96 grid at 8 transforms)
template <int BCT>
__host__ __device__ constexpr int n3_grid_tpl() {
    return BCT < 4 ? BCT : 4;
}

// Gridding as an output-stationary gather: each (cell, transform) sum runs
// over its cell's CSR row in sample order; every cell written once (here, or
// by n3_grid_long when its row exceeds kN3LongRow entries).  A lane carries
// TPL = min(BCT, 4) transforms of one cell (16-byte loads and stores), so a
// wave serves 64 TPL / BCT cells: the kernel waits on the rows' load chains
// (entry, then sample) once per 16-32 cells instead of once per 8 (806 ->
// 272 us -> this on a 192^2 x 96 grid).  The row is walked U entries at a
// time, their entry loads then their sample loads issued together.
template <int BCT>
__global__ __launch_bounds__(256) void n3_grid_csr(const int* __restrict__ rp, const int2* __restrict__ ent,
                                                   const float2* __restrict__ yT, float2* __restrict__ ws,
                                                   int64_t ncell) {
    constexpr int TPL = n3_grid_tpl<BCT>(), LPC = BCT / TPL, CPB = 256 / LPC, U = 4;
    const int64_t cell = (int64_t)blockIdx.x * CPB + threadIdx.x / LPC;
    const int b0 = (int)(threadIdx.x % LPC) * TPL;
    if (cell >= ncell) return;
    const int e0 = rp[cell], e1 = rp[cell + 1];
    if (e1 - e0 > kN3LongRow) return;
    float2 acc[TPL];
#pragma unroll
    for (int t = 0; t < TPL; ++t) acc[t] = make_float2(0.f, 0.f);
    // the TPL values of entry en for this lane
    auto load = [&](const int2 en, float2* v) {
        const float2* src = yT + (int64_t)en.x * BCT + b0;
        if constexpr (TPL >= 2) {
#pragma unroll
            for (int t = 0; t < TPL; t += 2) {
                const float4 q = *reinterpret_cast<const float4*>(src + t);
                v[t] = make_float2(q.x, q.y);
                v[t + 1] = make_float2(q.z, q.w);
            }
        } else {
            v[0] = src[0];
        }
    };
    int e = e0;
    for (; e + U <= e1; e += U) {
        int2 en[U];
        float2 v[U][TPL];
#pragma unroll
        for (int u = 0; u < U; ++u) en[u] = ent[e + u];
#pragma unroll
        for (int u = 0; u < U; ++u) load(en[u], v[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const float w = __int_as_float(en[u].y);
#pragma unroll
            for (int t = 0; t < TPL; ++t) {
                acc[t].x += w * v[u][t].x;
                acc[t].y += w * v[u][t].y;
            }
        }
    }
    for (; e < e1; ++e) {
        const int2 en = ent[e];
        const float w = __int_as_float(en.y);
        float2 v[TPL];
        load(en, v);
#pragma unroll
        for (int t = 0; t < TPL; ++t) {
            acc[t].x += w * v[t].x;
            acc[t].y += w * v[t].y;
        }
    }
    float2* dst = ws + cell * BCT + b0;
    if constexpr (TPL >= 2) {
#pragma unroll
        for (int t = 0; t < TPL; t += 2)
            *reinterpret_cast<float4*>(dst + t) = make_float4(acc[t].x, acc[t].y, acc[t + 1].x, acc[t + 1].y);
    } else {
        dst[0] = acc[0];
    }
}

// Gridding of the long rows (the k-space centre): one work-group per listed
// cell, lane = (phase j, transform b) with P = 256 / BCT phases; lane j sums
// entries e0 + j, e0 + j + P, ... in order (four entries' loads issued
// together), the phases of each wave are added by a fixed shuffle tree and
// the four waves' sums in wave order, so every cell's sum is the same from
// run to run (44 us on the nufft3 workload's ~1000 cells; one wave per cell
// took 55).
template <int BCT>
__global__ __launch_bounds__(256) void n3_grid_long(const int* __restrict__ cells, int nlong,
                                                    const int* __restrict__ rp, const int2* __restrict__ ent,
                                                    const float2* __restrict__ yT, float2* __restrict__ ws) {
    constexpr int P = 256 / BCT, U = 4;
    __shared__ float2 red[4][BCT];
    if ((int)blockIdx.x >= nlong) return;   // uniform over the block
    const int j = threadIdx.x / BCT, b = threadIdx.x % BCT;
    const int64_t cell = cells[blockIdx.x];
    const int e0 = rp[cell], e1 = rp[cell + 1];
    float2 acc = make_float2(0.f, 0.f);
    int e = e0 + j;
    for (; e + (U - 1) * P < e1; e += U * P) {
        int2 en[U];
        float2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) en[u] = ent[e + u * P];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = yT[(int64_t)en[u].x * BCT + b];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const float w = __int_as_float(en[u].y);
            acc.x += w * v[u].x;
            acc.y += w * v[u].y;
        }
    }
    for (; e < e1; e += P) {
        const int2 en = ent[e];
        const float w = __int_as_float(en.y);
        const float2 v = yT[(int64_t)en.x * BCT + b];
        acc.x += w * v.x;
        acc.y += w * v.y;
    }
#pragma unroll
    for (int off = BCT; off < 64; off <<= 1) {
        acc.x += __shfl_xor(acc.x, off);
        acc.y += __shfl_xor(acc.y, off);
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) < BCT) red[wv][b] = acc;
    __syncthreads();
    if ((int)threadIdx.x < BCT) {
        float2 s = red[0][threadIdx.x];
#pragma unroll
        for (int w = 1; w < 4; ++w) {
            s.x += red[w][threadIdx.x].x;
            s.y += red[w][threadIdx.x].y;
        }
        ws[cell * BCT + threadIdx.x] = s;
    }
}

// ---------------------------------------------------------------------------
// host

Nufft3Op::Nufft3Op(mm_ctx* c, const int64_t* imN, int64_t M, const double* kloc, const double* dcf, double os,
                   int kwidth) {
    ctx = c;
    if (kwidth < 2 || kwidth > 8 || (kwidth & 1))
        fail(MM_ERR_ARG, "nufftOp: kwidth must be even and in [2, 8]");
    if (!(os > 1.0)) fail(MM_ERR_ARG, "nufftOp: os must be > 1");
    if (M < 1) fail(MM_ERR_ARG, "nufftOp: no samples");
    if (M > INT32_MAX) fail(MM_ERR_ARG, "nufftOp: too many samples");
    if (!kloc) fail(MM_ERR_ARG, "nufftOp: kloc is NULL");
    int N[3], G[3];
    for (int d = 0; d < 3; ++d) {
        if (imN[d] < 2) fail(MM_ERR_ARG, "nufftOp: image size too small");
        if (imN[d] > 100000) fail(MM_ERR_ARG, "nufftOp: image size too large");
        N[d] = (int)imN[d];
        G[d] = nufft_grid_size(N[d], os);
        if (G[d] > 512)
            fail(MM_ERR_UNSUPPORTED, "nufftOp: 3-D grid sides above 512 are not implemented (LDS FFT passes)");
        if (G[d] < kwidth) fail(MM_ERR_ARG, "nufftOp: grid side smaller than the kernel width");
    }
    const int W = kwidth;
    const int64_t ncell = (int64_t)G[0] * G[1] * G[2];
    const int64_t nent = M * W * W * W;
    if (nent > (int64_t)INT32_MAX) fail(MM_ERR_ARG, "nufftOp: too many samples for the 3-D gridding rows");
    g_.N1 = N[0];
    g_.N2 = N[1];
    g_.N3 = N[2];
    g_.G1 = G[0];
    g_.G2 = G[1];
    g_.G3 = G[2];
    g_.s1 = N[0] / 2;
    g_.s2 = N[1] / 2;
    g_.s3 = N[2] / 2;
    g_.M = (int)M;
    g_.W = W;
    g_.odd = (N[0] & 1) | ((N[1] & 1) << 1) | ((N[2] & 1) << 2);
    g_.V = (int64_t)N[0] * N[1] * N[2];
    hipStream_t st = ctx->stream;

    std::vector<FftFactors> facs(3);
    std::vector<float2> twid;
    for (int d = 0; d < 3; ++d) {
        if (!fft_factorize(G[d], facs[d], false)) fail(MM_ERR_ARG, "nufftOp: grid size with a prime factor > 7");
        fach_[d] = facs[d];
        tw_off_[d] = (int)twid.size();
        fft_pass_twiddles(facs[d], [&](float cs, float sn) { twid.push_back(make_float2(cs, sn)); });
    }
    // deapodisation per side; side 1 carries 1/sqrt(G1 G2 G3) (unitary
    // ifftnc) and (-1)^(s1+s2+s3), the real part of the ramp constant
    // exp(i pi (N1+N2+N3)/2) (an odd side's i goes with its ramp)
    std::vector<float> deap;
    const double sgn = ((g_.s1 + g_.s2 + g_.s3) & 1) ? -1.0 : 1.0;
    const double scl = sgn / std::sqrt((double)ncell);
    for (int d = 0; d < 3; ++d) {
        const auto cd = nufft_deapod(N[d], G[d], kwidth, os);
        for (double v : cd) deap.push_back((float)(d == 0 ? v * scl : v));
    }
    // per-sample weights and first taps (nufftOp.m:187-246)
    const std::vector<double> lut = nufft_kb_lut(kwidth, os);
    const int Nk = nufft_kb_nkern(kwidth, os);
    std::vector<float> wt((size_t)M * 3 * W), dcfv((size_t)M);
    std::vector<double> wd((size_t)M * 3 * W);
    std::vector<uint32_t> base((size_t)M);
    std::vector<int> g0((size_t)M * 3);
    for (int64_t m = 0; m < M; ++m) {
        uint32_t packed = 0;
        for (int d = 0; d < 3; ++d) {
            const double ka = (kloc[(int64_t)d * M + m] + 0.5) * G[d];
            if (!std::isfinite(ka)) fail(MM_ERR_ARG, "nufftOp: non-finite kloc");
            const double fl = std::floor(ka);
            const double fr = fl - ka;
            for (int si = 0; si < W; ++si) {
                const int soff = -W / 2 + 1 + si;
                const double ki = (fr + soff) * (Nk - 1) / W + Nk / 2.0 + 0.5;
                const double fk = std::floor(ki);
                const double frac = ki - fk;
                int ii = (int)fk;   // 1-based
                ii = std::min(std::max(ii, 1), Nk);
                const double w = (1 - frac) * lut[(size_t)ii - 1] + frac * lut[(size_t)ii];
                wd[(size_t)m * 3 * W + (size_t)d * W + si] = w;
                wt[(size_t)m * 3 * W + (size_t)d * W + si] = (float)w;
            }
            int64_t gfirst = ((int64_t)fl - W / 2 + 1 + G[d] / 2) % G[d];
            if (gfirst < 0) gfirst += G[d];
            g0[(size_t)m * 3 + d] = (int)gfirst;
            packed |= (uint32_t)gfirst << (10 * d);
        }
        base[(size_t)m] = packed;
        dcfv[(size_t)m] = dcf ? (float)dcf[m] : 1.f;
    }
    // CSR rows of the transposed interpolation matrix, each in sample order
    std::vector<int> rp((size_t)ncell + 1, 0);
    auto for_taps = [&](int64_t m, auto fn) {
        int c3 = g0[(size_t)m * 3 + 2];
        for (int s3 = 0; s3 < W; ++s3) {
            int c2 = g0[(size_t)m * 3 + 1];
            for (int s2 = 0; s2 < W; ++s2) {
                int c1 = g0[(size_t)m * 3];
                for (int s1 = 0; s1 < W; ++s1) {
                    fn(((int64_t)c3 * G[1] + c2) * G[0] + c1, s1, s2, s3);
                    if (++c1 == G[0]) c1 = 0;
                }
                if (++c2 == G[1]) c2 = 0;
            }
            if (++c3 == G[2]) c3 = 0;
        }
    };
    for (int64_t m = 0; m < M; ++m) for_taps(m, [&](int64_t cell, int, int, int) { ++rp[(size_t)cell + 1]; });
    for (int64_t i = 0; i < ncell; ++i) rp[(size_t)i + 1] += rp[(size_t)i];
    std::vector<int> fillp(rp.begin(), rp.end() - 1);
    std::vector<int2> ent((size_t)nent);
    for (int64_t m = 0; m < M; ++m) {
        const double* w = wd.data() + (size_t)m * 3 * W;
        const double dm = dcf ? dcf[m] : 1.0;
        for_taps(m, [&](int64_t cell, int s1, int s2, int s3) {
            const float v = (float)(w[s1] * w[W + s2] * w[2 * W + s3] * dm);
            int bits;
            std::memcpy(&bits, &v, sizeof bits);
            ent[(size_t)fillp[(size_t)cell]++] = make_int2((int)m, bits);
        });
    }
    facs_.upload(facs, st);
    twid_.upload(twid, st);
    deap_.upload(deap, st);
    base_.upload(base, st);
    wt_.upload(wt, st);
    dcf_.upload(dcfv, st);
    rp_.upload(rp, st);
    {
        std::vector<int> lc;
        for (int64_t i = 0; i < ncell; ++i)
            if (rp[(size_t)i + 1] - rp[(size_t)i] > kN3LongRow) lc.push_back((int)i);
        nlong_ = (int)lc.size();
        long_.upload(lc.empty() ? std::vector<int>{0} : lc, st);
    }
    ent_.upload(ent, st);
    HIPCHECK(hipStreamSynchronize(st));
}

template <int BCT, int MID>
void Nufft3Op::launch(int adjoint, const float2* in, float2* out, int64_t nrep) {
    hipStream_t st = ctx->stream;
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void*)n3_rows_fwd<BCT, MID>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        (void)hipFuncSetAttribute((const void*)n3_rows_adj<BCT, MID>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        (void)hipFuncSetAttribute((const void*)n3_axis<BCT, 0, MID, 2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        (void)hipFuncSetAttribute((const void*)n3_axis<BCT, 1, MID, 2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        (void)hipFuncSetAttribute((const void*)n3_axis<BCT, 0, MID, 1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        (void)hipFuncSetAttribute((const void*)n3_axis<BCT, 1, MID, 1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        (void)hipGetLastError();
        attr_set = true;
    }
    constexpr int T = n3_T<BCT>(), TR = n3_Tr<BCT>(), LPG = T / BCT, LPR = TR / BCT, SPB = 256 / BCT;
    Nu3Geom g = g_;
    g.ip = 0;
    if (fft_inplace_ok(fach_[0], TR, n3_cap<TR, MID>())) g.ip |= 1;
    for (int d = 1; d < 3; ++d)
        if (fft_inplace_ok(fach_[d], T, n3_cap<T, MID>())) g.ip |= 1 << d;
    // the g3 pass with twice the pencils where its side fits one buffer so
    constexpr int T3W = 2 * T;
    const bool wide3 = fft_inplace_ok(fach_[2], T3W, n3_cap<T3W, MID>());
    if (wide3) g.ip |= 4;
    const size_t lds1 = n3_lds(g.G1, TR, g.ip & 1), lds2 = n3_lds(g.G2, T, (g.ip >> 1) & 1),
                 lds3 = wide3 ? n3_lds(g.G3, T3W, true) : n3_lds(g.G3, T, (g.ip >> 2) & 1);
    // after each launch: the launch check and, while profiling, an event
    auto done = [&](const char* stage) {
        LAUNCHCHECK();
        if (marks_ && nmark_ < marks_->size()) {
            (*marks_)[nmark_].first = stage;
            HIPCHECK(hipEventRecord((*marks_)[nmark_].second, ctx->stream));
            ++nmark_;
        }
    };
    const int64_t ncell = (int64_t)g.G1 * g.G2 * g.G3;
    g.chunk = ncell * BCT;
    ws_.ensure((size_t)g.chunk);
    if (adjoint && !toep_) yT_.ensure((size_t)g.M * BCT);
    const unsigned rows_blocks = (unsigned)((g.N2 * g.N3 + LPR - 1) / LPR);
    const unsigned gx = (unsigned)((g.G1 + LPG - 1) / LPG);
    const unsigned gx3 = (unsigned)((g.G1 + 2 * LPG - 1) / (2 * LPG));
    auto axis3 = [&](int adj) {
        if (wide3) {
            if (adj)
                hipLaunchKernelGGL((n3_axis<BCT, 1, MID, 2>), dim3(gx3 * (unsigned)g.G2), dim3(256), lds3, st, g, 3,
                                   facs_.p + 2, twid_.p + tw_off_[2], ws_.p);
            else
                hipLaunchKernelGGL((n3_axis<BCT, 0, MID, 2>), dim3(gx3 * (unsigned)g.G2), dim3(256), lds3, st, g, 3,
                                   facs_.p + 2, twid_.p + tw_off_[2], ws_.p);
        } else {
            if (adj)
                hipLaunchKernelGGL((n3_axis<BCT, 1, MID, 1>), dim3(gx * (unsigned)g.G2), dim3(256), lds3, st, g, 3,
                                   facs_.p + 2, twid_.p + tw_off_[2], ws_.p);
            else
                hipLaunchKernelGGL((n3_axis<BCT, 0, MID, 1>), dim3(gx * (unsigned)g.G2), dim3(256), lds3, st, g, 3,
                                   facs_.p + 2, twid_.p + tw_off_[2], ws_.p);
        }
    };
    const unsigned samp_blocks = (unsigned)((g.M + SPB - 1) / SPB);
    const unsigned elem_blocks = (unsigned)((g.chunk + 255) / 256);
    // n3_grid_csr: BCT / n3_grid_tpl lanes per cell
    const unsigned cell_blocks = (unsigned)((ncell * (BCT / n3_grid_tpl<BCT>()) + 255) / 256);
    const int64_t nchunk = (nrep + BCT - 1) / BCT;
    for (int64_t ch = 0; ch < nchunk; ++ch) {
        const int lanes = (int)std::min<int64_t>(BCT, nrep - ch * BCT);
        if (toep_) {
            // both directions: forward FFTs, spectrum multiply, inverse FFTs
            hipLaunchKernelGGL((n3_rows_fwd<BCT, MID>), dim3(rows_blocks), dim3(256), lds1, st, g, facs_.p,
                               twid_.p + tw_off_[0], toep_deap_.p, in, ws_.p, (int)ch, lanes);
            done("n3_rows_fwd");
            hipLaunchKernelGGL((n3_axis<BCT, 0, MID, 1>), dim3(gx * (unsigned)g.N3), dim3(256), lds2, st, g, 2,
                               facs_.p + 1, twid_.p + tw_off_[1], ws_.p);
            done("n3_axis2_fwd");
            axis3(0);
            done("n3_axis3_fwd");
            hipLaunchKernelGGL(n3_toep_mul<BCT>, dim3(elem_blocks), dim3(256), 0, st, g, (const float2*)toep_k_.p,
                               ws_.p);
            done("n3_toep_mul");
            axis3(1);
            done("n3_axis3_adj");
            hipLaunchKernelGGL((n3_axis<BCT, 1, MID, 1>), dim3(gx * (unsigned)g.N3), dim3(256), lds2, st, g, 2,
                               facs_.p + 1, twid_.p + tw_off_[1], ws_.p);
            done("n3_axis2_adj");
            hipLaunchKernelGGL((n3_rows_adj<BCT, MID>), dim3(rows_blocks), dim3(256), lds1, st, g, facs_.p,
                               twid_.p + tw_off_[0], toep_deap_.p, (const float2*)ws_.p, out, (int)ch, lanes);
            done("n3_rows_adj");
            continue;
        }
        if (!adjoint) {
            hipLaunchKernelGGL((n3_rows_fwd<BCT, MID>), dim3(rows_blocks), dim3(256), lds1, st, g, facs_.p,
                               twid_.p + tw_off_[0], deap_.p, in, ws_.p, (int)ch, lanes);
            done("n3_rows_fwd");
            hipLaunchKernelGGL((n3_axis<BCT, 0, MID, 1>), dim3(gx * (unsigned)g.N3), dim3(256), lds2, st, g, 2,
                               facs_.p + 1, twid_.p + tw_off_[1], ws_.p);
            done("n3_axis2_fwd");
            axis3(0);
            done("n3_axis3_fwd");
            if (g.odd) {
                hipLaunchKernelGGL((n3_odd_ramp<BCT, 0>), dim3(elem_blocks), dim3(256), 0, st, g, ws_.p);
                done("n3_odd_ramp");
            }
            {
                // n3_degrid: BCT lanes per sample
                const unsigned deg_blocks = (unsigned)((g.M * (int64_t)BCT + 255) / 256);
                if (g.W == 4)
                    hipLaunchKernelGGL((n3_degrid<BCT, 4>), dim3(deg_blocks), dim3(256), 0, st, g, base_.p, wt_.p,
                                       dcf_.p, (const float2*)ws_.p, out, (int)ch, lanes);
                else
                    hipLaunchKernelGGL((n3_degrid<BCT, 0>), dim3(deg_blocks), dim3(256), 0, st, g, base_.p, wt_.p,
                                       dcf_.p, (const float2*)ws_.p, out, (int)ch, lanes);
            }
            done("n3_degrid");
        } else {
            hipLaunchKernelGGL(n3_transpose_y<BCT>, dim3(samp_blocks), dim3(256), 0, st, g, in, yT_.p, (int)ch,
                               lanes);
            done("n3_transpose_y");
            hipLaunchKernelGGL(n3_grid_csr<BCT>, dim3(cell_blocks), dim3(256), 0, st, rp_.p, ent_.p,
                               (const float2*)yT_.p, ws_.p, ncell);
            done("n3_grid_csr");
            if (nlong_ > 0) {
                hipLaunchKernelGGL(n3_grid_long<BCT>, dim3((unsigned)nlong_), dim3(256), 0, st, long_.p, nlong_,
                                   rp_.p, ent_.p, (const float2*)yT_.p, ws_.p);
                done("n3_grid_csr");   // the gridding stage: both launches
            }
            if (g.odd) {
                hipLaunchKernelGGL((n3_odd_ramp<BCT, 1>), dim3(elem_blocks), dim3(256), 0, st, g, ws_.p);
                done("n3_odd_ramp");
            }
            axis3(1);
            done("n3_axis3_adj");
            hipLaunchKernelGGL((n3_axis<BCT, 1, MID, 1>), dim3(gx * (unsigned)g.N3), dim3(256), lds2, st, g, 2,
                               facs_.p + 1, twid_.p + tw_off_[1], ws_.p);
            done("n3_axis2_adj");
            hipLaunchKernelGGL((n3_rows_adj<BCT, MID>), dim3(rows_blocks), dim3(256), lds1, st, g, facs_.p,
                               twid_.p + tw_off_[0], deap_.p, (const float2*)ws_.p, out, (int)ch, lanes);
            done("n3_rows_adj");
        }
    }
}

void Nufft3Op::set_toeplitz(const float2* psf) {
    if (!psf) fail(MM_ERR_ARG, "nufftOp: NULL point-spread image");
    const int N[3] = {g_.N1, g_.N2, g_.N3}, G[3] = {g_.G1, g_.G2, g_.G3};
    for (int d = 0; d < 3; ++d)
        if (G[d] < 2 * N[d] - 1) fail(MM_ERR_ARG, "nufftOp: Toeplitz form needs a grid of at least 2N - 1");
    const int P[3] = {2 * N[0], 2 * N[1], 2 * N[2]};
    const int L[3] = {2 * N[0] - 1, 2 * N[1] - 1, 2 * N[2] - 1};
    const double alpha = 8.0 / std::sqrt((double)P[0] * P[1] * P[2]);
    // h(D) = alpha psf[N + D] for |D_d| <= N_d - 1 (the lags between two
    // voxels of the cropped volume); K = sum_D h(D) exp(+2 pi i g.D / G), the
    // forward passes' sign, one side at a time in FP64
    auto ex = [](int Gd) {
        std::vector<std::pair<double, double>> e((size_t)Gd);
        for (int k = 0; k < Gd; ++k) e[(size_t)k] = {std::cos(2 * M_PI * k / Gd), std::sin(2 * M_PI * k / Gd)};
        return e;
    };
    const auto e1 = ex(G[0]), e2 = ex(G[1]), e3 = ex(G[2]);
    auto tw = [](const std::vector<std::pair<double, double>>& e, int gd, int D, int Gd) {
        return e[(size_t)((((int64_t)gd * D) % Gd + Gd) % Gd)];
    };
    // side 1: A1[j3][j2][g1]
    std::vector<double> ar((size_t)L[2] * L[1] * G[0]), ai(ar.size());
    for (int j3 = 0; j3 < L[2]; ++j3)
        for (int j2 = 0; j2 < L[1]; ++j2) {
            const int64_t col = (int64_t)(N[1] + j2 - (N[1] - 1)) * P[0] + (int64_t)(N[2] + j3 - (N[2] - 1)) * P[0] * P[1];
            for (int g1 = 0; g1 < G[0]; ++g1) {
                double sr = 0, si = 0;
                for (int j1 = 0; j1 < L[0]; ++j1) {
                    const int D1 = j1 - (N[0] - 1);
                    const float2 h = psf[col + N[0] + D1];
                    const auto w = tw(e1, g1, D1, G[0]);
                    sr += h.x * w.first - h.y * w.second;
                    si += h.x * w.second + h.y * w.first;
                }
                const size_t o = ((size_t)j3 * L[1] + j2) * G[0] + g1;
                ar[o] = sr;
                ai[o] = si;
            }
        }
    // side 2: A2[j3][g2][g1]
    std::vector<double> br((size_t)L[2] * G[1] * G[0]), bi(br.size());
    for (int j3 = 0; j3 < L[2]; ++j3)
        for (int g2 = 0; g2 < G[1]; ++g2)
            for (int g1 = 0; g1 < G[0]; ++g1) {
                double sr = 0, si = 0;
                for (int j2 = 0; j2 < L[1]; ++j2) {
                    const auto w = tw(e2, g2, j2 - (N[1] - 1), G[1]);
                    const size_t o = ((size_t)j3 * L[1] + j2) * G[0] + g1;
                    sr += ar[o] * w.first - ai[o] * w.second;
                    si += ar[o] * w.second + ai[o] * w.first;
                }
                const size_t q = ((size_t)j3 * G[1] + g2) * G[0] + g1;
                br[q] = sr;
                bi[q] = si;
            }
    // side 3: K[g3][g2][g1]
    std::vector<float2> km((size_t)G[2] * G[1] * G[0]);
    for (int g3 = 0; g3 < G[2]; ++g3)
        for (int g2 = 0; g2 < G[1]; ++g2)
            for (int g1 = 0; g1 < G[0]; ++g1) {
                double sr = 0, si = 0;
                for (int j3 = 0; j3 < L[2]; ++j3) {
                    const auto w = tw(e3, g3, j3 - (N[2] - 1), G[2]);
                    const size_t o = ((size_t)j3 * G[1] + g2) * G[0] + g1;
                    sr += br[o] * w.first - bi[o] * w.second;
                    si += br[o] * w.second + bi[o] * w.first;
                }
                km[((size_t)g3 * G[1] + g2) * G[0] + g1] = make_float2((float)(alpha * sr), (float)(alpha * si));
            }
    // flat scaling 1/sqrt(G1 G2 G3) on both sides (unitary FFT pair)
    std::vector<float> dflat((size_t)(N[0] + N[1] + N[2]), 1.f);
    const float f = (float)(1.0 / std::sqrt((double)G[0] * G[1] * G[2]));
    for (int i = 0; i < N[0]; ++i) dflat[(size_t)i] = f;
    hipStream_t st = ctx->stream;
    toep_k_.upload(km, st);
    toep_deap_.upload(dflat, st);
    HIPCHECK(hipStreamSynchronize(st));
    toep_ = true;
    ++gen_;
}

bool Nufft3Op::profile_dominant(int adjoint, const float2* in, float2* out, int64_t nrep, int iters, double* ms,
                                std::string* name) {
    if ((adjoint != 0 && adjoint != 1) || nrep < 1 || iters < 1) return false;
    hipStream_t st = ctx->stream;
    apply(adjoint, in, out, nrep);   // workspace, first-launch costs
    // chunks of apply's transform count, at most seven marked launches each
    int64_t bct = 1;
    while (bct < nrep && bct < 16) bct <<= 1;
    const int64_t nchunk = (nrep + bct - 1) / bct;
    std::vector<std::pair<const char*, hipEvent_t>> mk((size_t)(8 * nchunk));
    hipEvent_t e0;
    HIPCHECK(hipEventCreate(&e0));
    for (auto& m : mk) HIPCHECK(hipEventCreate(&m.second));
    std::map<std::string, double> acc;
    try {
        for (int it = 0; it < iters; ++it) {
            nmark_ = 0;
            HIPCHECK(hipEventRecord(e0, st));
            marks_ = &mk;
            apply(adjoint, in, out, nrep);
            marks_ = nullptr;
            if (nmark_ == 0) break;
            HIPCHECK(hipEventSynchronize(mk[nmark_ - 1].second));
            hipEvent_t prev = e0;
            for (size_t j = 0; j < nmark_; ++j) {
                float t = 0.f;
                HIPCHECK(hipEventElapsedTime(&t, prev, mk[j].second));
                acc[mk[j].first] += t;
                prev = mk[j].second;
            }
        }
    } catch (...) {
        marks_ = nullptr;
        (void)hipEventDestroy(e0);
        for (auto& m : mk) (void)hipEventDestroy(m.second);
        throw;
    }
    (void)hipEventDestroy(e0);
    for (auto& m : mk) (void)hipEventDestroy(m.second);
    if (acc.empty()) return false;
    auto best = std::max_element(acc.begin(), acc.end(),
                                 [](const auto& a, const auto& b) { return a.second < b.second; });
    *ms = best->second / iters;
    *name = best->first;
    return true;
}

void Nufft3Op::apply(int adjoint, const float2* in, float2* out, int64_t nrep) {
    if (nrep < 1) fail(MM_ERR_ARG, "nufftOp: nrep must be >= 1");
    int bct = 1;
    while (bct < nrep && bct < 16) bct <<= 1;
    // the kernels with twice the register cap where a side's passes do not
    // fit one LDS buffer with the default one
    const int T = bct >= 8 ? 16 : 8, TR = bct >= 16 ? 16 : 8;   // n3_T, n3_Tr
    bool mid = !fft_inplace_ok(fach_[0], TR, fft_reg_cap(TR, 0));
    for (int d = 1; d < 3; ++d) mid = mid || !fft_inplace_ok(fach_[d], T, fft_reg_cap(T, 0));
    auto go = [&](auto m) {
        constexpr int MID = decltype(m)::value;
        switch (bct) {
            case 1: launch<1, MID>(adjoint, in, out, nrep); break;
            case 2: launch<2, MID>(adjoint, in, out, nrep); break;
            case 4: launch<4, MID>(adjoint, in, out, nrep); break;
            case 8: launch<8, MID>(adjoint, in, out, nrep); break;
            default: launch<16, MID>(adjoint, in, out, nrep); break;
        }
    };
    if (mid)
        go(std::integral_constant<int, 1>());
    else
        go(std::integral_constant<int, 0>());
}

}  // namespace mm

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
// For even sides the k-space phase ramp P = fftShiftArray (:258-273) equals
// an image-domain circular shift by N/2 times the constant
// (-1)^(N1/2+N2/2), so the image is written wrapped into the zero-padded FFT
// input and no ramp is multiplied.  An odd side N = 2s + 1 is written
// wrapped by s and the rest of its ramp, i exp(-i pi g / G), is multiplied
// into the grid (nu_odd_ramp).
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
//   nu_grid_tile  2-D gridding from LDS-staged rectangles of the CSR rows
//                 (+ nu_grid_merge for the chunks of dense rectangles)
//   nu_grid_csr   gridding as an output-stationary gather from the global
//                 CSR rows (1-D always, 2-D as an option): one lane per
//                 (cell, transform) sums, in sample order, the row of
//                 (sample, w1*w2*dcf) pairs that reach the cell -- every cell
//                 written once, no atomics, deterministic
//   nu_adj_rows   FFT along g1 + crop + deapod + conj-coil combine
//   nu_odd_ramp   odd image sides only: the part of the ramp P that is not a
//                 circular shift, multiplied into the grid
#include <algorithm>
#include <cstring>
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

// Rectangles of the staged 2-D gridding for one trajectory, given its CSR
// rows (rp: G1*G2+1 pointers from 0, en: (sample, weight) entries): 16x16
// tiles, cut into quadrants while the union of their samples exceeds cap_s
// or their entries cap_e.  A rectangle of at most 4x4
// cells that still exceeds them (dense k-space centre) is cut into chunks of
// consecutive samples of its sorted union, each within both limits (kind 2).
// Staged rows keep the sample order of the global rows (local sample ids are
// ranks in the sorted union, or in the chunk).  npcells / ndense count the
// cells of the partial blocks and the dense rectangles of the trajectory.
static void build_grid_tiles(int G1, int G2, int M, const int* rp, const int2* en, int cap_s, int cap_e,
                             std::vector<GridTile>& tiles, std::vector<int>& stage,
                             std::vector<int>& ptr, std::vector<int2>& lent, int& npcells, int& ndense) {
    std::vector<int> stamp((size_t)M, -1), loc((size_t)M, 0), uni;
    std::vector<int4> todo;
    for (int a2 = G2 - ((G2 - 1) % 16 + 1); a2 >= 0; a2 -= 16)
        for (int a1 = G1 - ((G1 - 1) % 16 + 1); a1 >= 0; a1 -= 16)
            todo.push_back(make_int4(a1, a2, std::min(16, G1 - a1), std::min(16, G2 - a2)));
    npcells = 0;
    ndense = 0;
    int id = 0;
    // appends a staged tile over the samples uni[u0, u1) (sorted) of rectangle r
    auto emit = [&](const int4& r, int kind, int u0, int u1) {
        GridTile t{};
        t.g1 = r.x;
        t.g2 = r.y;
        t.w = r.z;
        t.h = r.w;
        t.kind = kind;
        t.n_stage = u1 - u0;
        t.stage_off = (int)stage.size();
        t.ptr_off = (int)ptr.size();
        t.ent_off = (int)lent.size();
        for (int i = u0; i < u1; ++i) loc[(size_t)uni[i]] = i - u0;
        stage.insert(stage.end(), uni.begin() + u0, uni.begin() + u1);
        const int mlo = uni[u0], mhi = uni[u1 - 1];
        int run = 0;
        ptr.push_back(0);
        for (int y = r.y; y < r.y + r.w; ++y)
            for (int x = r.x; x < r.x + r.z; ++x) {
                const int c = y * G1 + x;
                for (int q = rp[c]; q < rp[c + 1]; ++q) {
                    const int m = en[q].x;
                    if (m < mlo || m > mhi) continue;
                    lent.push_back(make_int2(loc[(size_t)m], en[q].y));
                    ++run;
                }
                ptr.push_back(run);
            }
        t.n_ent = run;
        tiles.push_back(t);
    };
    while (!todo.empty()) {
        const int4 r = todo.back();
        todo.pop_back();
        uni.clear();
        int64_t ne = 0;
        for (int y = r.y; y < r.y + r.w; ++y)
            for (int x = r.x; x < r.x + r.z; ++x) {
                const int c = y * G1 + x;
                ne += rp[c + 1] - rp[c];
                for (int q = rp[c]; q < rp[c + 1]; ++q)
                    if (stamp[en[q].x] != id) {
                        stamp[en[q].x] = id;
                        uni.push_back(en[q].x);
                    }
            }
        ++id;
        std::sort(uni.begin(), uni.end());
        if ((int)uni.size() <= cap_s && ne <= cap_e) {
            if (!uni.empty()) emit(r, 0, 0, (int)uni.size());
            else {
                GridTile t{};   // no samples: the cells are written as zeros
                t.g1 = r.x;
                t.g2 = r.y;
                t.w = r.z;
                t.h = r.w;
                t.ptr_off = (int)ptr.size();
                t.ent_off = (int)lent.size();
                for (int i = 0; i <= r.z * r.w; ++i) ptr.push_back(0);
                tiles.push_back(t);
            }
        } else if (r.z <= 4 && r.w <= 4) {
            // entries per sample inside the rectangle
            std::vector<int> cnt(uni.size(), 0);
            for (int i = 0; i < (int)uni.size(); ++i) loc[(size_t)uni[i]] = i;
            for (int y = r.y; y < r.y + r.w; ++y)
                for (int x = r.x; x < r.x + r.z; ++x) {
                    const int c = y * G1 + x;
                    for (int q = rp[c]; q < rp[c + 1]; ++q) cnt[(size_t)loc[(size_t)en[q].x]]++;
                }
            const size_t first = tiles.size();
            int u0 = 0, cs = 0, ce = 0;
            for (int i = 0; i < (int)uni.size(); ++i) {
                if (cs + 1 > cap_s || ce + cnt[(size_t)i] > cap_e) {
                    emit(r, 2, u0, i);
                    u0 = i;
                    cs = ce = 0;
                }
                ++cs;
                ce += cnt[(size_t)i];
            }
            emit(r, 2, u0, (int)uni.size());
            const int np = (int)(tiles.size() - first);
            for (int k = 0; k < np; ++k) {
                GridTile& t = tiles[first + (size_t)k];
                t.pofs0 = npcells;
                t.pofs = npcells + k * r.z * r.w;
                t.nparts = np;
            }
            npcells += np * r.z * r.w;
            ++ndense;
        } else {
            const int hw = (r.z + 1) / 2, hh = (r.w + 1) / 2;
            if (r.z > hw && r.w > hh) todo.push_back(make_int4(r.x + hw, r.y + hh, r.z - hw, r.w - hh));
            if (r.w > hh) todo.push_back(make_int4(r.x, r.y + hh, hw, r.w - hh));
            if (r.z > hw) todo.push_back(make_int4(r.x + hw, r.y, r.z - hw, hh));
            todo.push_back(make_int4(r.x, r.y, hw, hh));
        }
    }
    if (stage.size() > (size_t)INT32_MAX || ptr.size() > (size_t)INT32_MAX || lent.size() > (size_t)INT32_MAX)
        fail(MM_ERR_ARG, "nufftOp: staged gridding tables exceed 2^31 entries");
}

// FNV-1a over the bytes of a trajectory (and its dcf): finds slices that can
// share their sample data.
static uint64_t traj_hash(const double* kl, const double* dc, int M, int ndim) {
    uint64_t h = 1469598103934665603ull;
    auto eat = [&](const void* p, size_t n) {
        const unsigned char* c = (const unsigned char*)p;
        for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ull;
    };
    eat(kl, sizeof(double) * (size_t)M * ndim);
    if (dc) eat(dc, sizeof(double) * (size_t)M);
    h ^= (uint64_t)M * 0x9E3779B97F4A7C15ull;
    return h;
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
std::vector<double> nufft_kb_lut(int kwidth, double os) {
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
std::vector<double> nufft_deapod(int N, int G, int kwidth, double os) {
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
    int ylayout;        // 1: data side in the transposed [M][BCT] layout (yT)
    int sense;          // 1: image side is one image combined with coil maps
    int C;              // coils (sense) -- lanes per chunk
    int64_t B;          // transforms per slice (plain)
    const float2* maps;
    const float* W;     // optional forward-output weights (data layout)
    const int* skip;    // per-slice flags: nonzero = leave the slice out (normal operator)
};

__device__ __forceinline__ bool skipped(const NuRun& r, int slice) { return r.skip && r.skip[slice]; }

__device__ __forceinline__ int lanes_of(const NuRun& r, int BCT, int ch) {
    if (r.sense) return r.C;
    const int64_t left = r.B - (int64_t)ch * BCT;
    return left < BCT ? (int)left : BCT;
}

__device__ __forceinline__ int64_t chunk_stride(const NuRun& r, int BCT) { return r.sense ? r.C : BCT; }

// Transforms per work-group of the FFT passes (compile-time): the row
// passes take all BCT transforms of a chunk for max(1, 8/BCT) rows, the
// column passes 8 transforms (8/CL columns of CL = min(BCT, 8) transforms)
// for chunks of up to 4 transforms, else 16: two columns of an 8-transform
// chunk or one of a 16-transform chunk, so each 128-byte line of a grid row
// is read and written by one work-group (6 % faster than one 8-transform
// column).  Grids with a side above kFftSmallMax (BIG) take one row / one
// column per work-group so that the two LDS buffers fit.
__host__ __device__ constexpr int clog2(int v) { return v <= 1 ? 0 : 1 + clog2(v / 2); }
__host__ __device__ constexpr int rows_T(int BCT, bool BIG) { return BIG || BCT >= 8 ? BCT : 8; }
__host__ __device__ constexpr int cols_CL(int BCT) { return BCT < 16 ? (BCT < 8 ? BCT : 8) : 16; }
__host__ __device__ constexpr int cols_T(int BCT, bool BIG) { return BIG ? cols_CL(BCT) : (BCT >= 8 ? 16 : 8); }

// FFT mode of a configuration's row / column kernels (template FM):
// kFftSmall -- every pass of every slice fits one LDS buffer with the
// default register cap (fft_reg_cap: 16 values per thread at 16 transforms,
// 8 otherwise); kFftMid -- a grid side above that (257..512), where the
// kernels hold twice as many values per thread (32 at 16 transforms, 16 at
// 8) so the passes still run in one LDS buffer (3 work-groups per CU
// instead of 1 at G = 336 and 16 transforms, which is also why those may
// take 168 registers: fft_waves), and the row passes of 16 transforms may
// take the radix-16/12 factorization; kFftBig -- a side above kFftSmallMax
// (wide index modulo).
constexpr int kFftSmall = 0, kFftMid = 1, kFftBig = 2;
__host__ __device__ constexpr int fft_mode_cap(int FM, int T) { return FM == kFftMid ? 2 * fft_reg_cap(T, 0) : 0; }
__host__ __device__ constexpr bool rows_wide(int T, int FM) { return FM == kFftMid && T >= 16; }
// minimum waves per SIMD of an FFT kernel with T transforms per work-group
// (the register budget): mid-size kernels of 16 transforms 3 (their LDS
// allows no more), others 4 at 16 transforms; `dflt` otherwise
__host__ __device__ constexpr int fft_waves(int FM, int T, int dflt) {
    return FM == kFftMid ? (T >= 16 ? 3 : 4) : (T >= 16 ? 4 : dflt);
}

// Forward rows: coil multiply + deapodisation, image written wrapped into
// the zero-padded row (positions (i1 - N1/2) mod G1), FFT along g1.  Lane
// index = (position, row, transform) with the transform fastest.  Offsets
// inside a slice are 32-bit (the plan checks V*BCT and G1*G2*BCT < 2^31).
template <int BCT, int FM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(fft_waves(FM, rows_T(BCT, FM == kFftBig), 8), 8))) void nu_fwd_rows(const NuSliceDev* __restrict__ sl,
                                                   const int4* __restrict__ work,
                                                   const FftFactors* __restrict__ facs,
                                                   const float2* __restrict__ twp,
                                                   const float* __restrict__ deap,
                                                   const float2* __restrict__ x,
                                                   float2* __restrict__ ws, NuRun run) {
    constexpr bool BIG = FM == kFftBig;
    constexpr int T = rows_T(BCT, BIG), LT = clog2(T), LB = clog2(BCT);
    constexpr int CAP = fft_mode_cap(FM, T);
    constexpr bool kWide = rows_wide(T, FM);
    extern __shared__ float2 lds[];
    const int4 wk = work[blockIdx.x];
    if (skipped(run, wk.x)) return;
    const NuSliceDev& S = sl[wk.x];
    const int ch = wk.y, r0 = wk.z;
    const int N1 = S.N1, N2 = S.N2, G1 = S.G1;
    const int V = N1 * N2;
    const bool ip = S.fip & 1;   // single-buffer row FFTs
    float2* A = lds;
    float2* Bb = ip ? nullptr : lds + G1 * T;
    float2* tw = lds + (ip ? 1 : 2) * G1 * T;
    const int lanes = lanes_of(run, BCT, ch);
    const float* d1 = deap + S.deap1;
    const float* d2 = deap + S.deap2;
    const float2* xs = x + S.img_off + (run.sense ? (int64_t)ch * V : (int64_t)ch * BCT * V);
    const float2* mp = run.maps + S.maps_off;
    const FftFactors& F = facs[S.fac1];
    float2* out = ws + S.ws_off + (int64_t)ch * S.ws_chunk;
    if (F.nf >= 2) {
        // The first pass takes its inputs (coil image times deapodization,
        // image column i1 at position (i1 - N1/2) mod G1, zeros elsewhere)
        // straight into the butterflies' registers and the last pass writes
        // the rows straight to the workspace: two LDS round trips fewer.
        // Loads use clamped (valid) addresses and zero by a factor.
        auto direct = [&](auto value) {
            auto src = [&](int q, int t) {
                int i1 = q + N1 / 2;
                if (i1 >= G1) i1 -= G1;
                const int bl = t & (BCT - 1);
                const int i2 = r0 + (t >> LB);
                const float ok = i1 < N1 && i2 < N2 && bl < lanes ? 1.f : 0.f;
                const int i1c = min(i1, N1 - 1), i2c = min(i2, N2 - 1);
                const float2 v = value(i2c * N1 + i1c, min(bl, lanes - 1));
                const float f = d1[i1c] * d2[i2c] * ok;
                return make_float2(v.x * f, v.y * f);
            };
            fft_first_pass_src<T, 1, kWide>(F.r[0], A, G1, src);
        };
        if (run.sense)
            direct([&](int v, int b) { return cmulf(xs[v], mp[b * V + v]); });
        else
            direct([&](int v, int b) { return xs[b * V + v]; });
        // the pass twiddles after the data loads, which they then do not delay
        for (int i = threadIdx.x; i < G1 - 1; i += kFftThreads) tw[i] = twp[S.tw1 + i];
        __syncthreads();
        const float2* Rb = fft_lds_range<T, 1, BIG, kWide, CAP>(A, Bb, F, tw, 1, F.nf - 1);
        int p = 1;
        for (int k = 0; k + 1 < F.nf; ++k) p *= F.r[k];
        fft_last_pass_dst<T, 1, BIG, kWide>(F.r[F.nf - 1], Rb, G1, p, tw + (p - 1), [&](int q, int t, float2 v) {
            const int i2 = r0 + (t >> LB);
            if (i2 < N2) out[(i2 * G1 + q) * BCT + (t & (BCT - 1))] = v;
        });
        return;
    }
    for (int i = (N1 - N1 / 2) * T + threadIdx.x; i < (G1 - N1 / 2) * T; i += kFftThreads)
        A[i] = make_float2(0.f, 0.f);
    const int total = N1 * T;
    // four elements per thread in flight: every load of a round is issued
    // before the first LDS store.  The loads use clamped (valid) addresses
    // for every element and elements outside the image are scaled by 0, so
    // no load sits behind a branch (a branch would make it wait).
    auto stage = [&](auto value) {
        // full rounds (every element in range) store without a predicate, so
        // no load can be sunk behind one; then at most one partial round
        auto round = [&](int i0, bool full) {
            float2 val[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int idx = i0 + u * 256;
                const int t = idx & (T - 1);
                const int i1 = min(idx >> LT, N1 - 1);
                const int bl = t & (BCT - 1);
                const int i2 = r0 + (t >> LB);
                const float ok = idx < total && i2 < N2 && bl < lanes ? 1.f : 0.f;
                const int i2c = min(i2, N2 - 1);
                const float2 q = value(i2c * N1 + i1, min(bl, lanes - 1));
                const float f = d1[i1] * d2[i2c] * ok;
                val[u] = make_float2(q.x * f, q.y * f);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int idx = i0 + u * 256;
                const int i1 = idx >> LT;
                int pos = i1 - N1 / 2;
                if (pos < 0) pos += G1;
                if (full || idx < total) A[pos * T + (idx & (T - 1))] = val[u];
            }
        };
        int i0 = threadIdx.x;
        for (; i0 + 3 * 256 < total; i0 += 4 * 256) round(i0, true);
        if (i0 < total) round(i0, false);
    };
    if (run.sense)
        stage([&](int v, int b) { return cmulf(xs[v], mp[b * V + v]); });
    else
        stage([&](int v, int b) { return xs[b * V + v]; });
    // the pass twiddles after the data loads, which they then do not delay
    for (int i = threadIdx.x; i < G1 - 1; i += kFftThreads) tw[i] = twp[S.tw1 + i];
    __syncthreads();
    const float2* R = fft_lds<T, 1, BIG, kWide, CAP>(A, Bb, F, tw);
    const int tot2 = G1 * T;
    for (int idx = threadIdx.x; idx < tot2; idx += kFftThreads) {
        const int t = idx & (T - 1);
        const int i2 = r0 + (t >> LB);
        if (i2 >= N2) continue;
        out[(i2 * G1 + (idx >> LT)) * BCT + (t & (BCT - 1))] = R[idx];
    }
}

// FFT along g2 for 8/CL columns of CL transforms.  Forward (adj = 0): input
// rows 0..N2-1 are placed at (i2 - N2/2) mod G2, output G2 rows.  Adjoint:
// input G2 rows, output rows i2 < N2 read from (i2 - N2/2) mod G2.  In place
// in `ws`.
template <int BCT, int FM, int ADJ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(fft_waves(FM, cols_T(BCT, FM == kFftBig), 4), 8))) void nu_cols(const NuSliceDev* __restrict__ sl,
                                               const int4* __restrict__ work,
                                               const FftFactors* __restrict__ facs,
                                               const float2* __restrict__ twp,
                                               float2* __restrict__ ws, const int* __restrict__ skip) {
    constexpr bool BIG = FM == kFftBig;
    constexpr int T = cols_T(BCT, BIG), LT = clog2(T), CL = cols_CL(BCT), LCL = clog2(CL), CPW = T / CL;
    constexpr int CAP = fft_mode_cap(FM, T);
    // passes of 16 transforms may run a radix-16/12 factorization (fac2)
    constexpr bool kWide = T >= 16 && !BIG;
    extern __shared__ float2 lds[];
    const int4 wk = work[blockIdx.x];
    if (skip && skip[wk.x]) return;
    const NuSliceDev& S = sl[wk.x];
    const int ch = wk.y, c0 = wk.z, b0 = wk.w;
    const int N2 = S.N2, G1 = S.G1, G2 = S.G2;
    const bool ip = (S.fip >> 1) & 1;   // single-buffer column FFTs
    float2* A = lds;
    float2* Bb = ip ? nullptr : lds + G2 * T;
    float2* tw = lds + (ip ? 1 : 2) * G2 * T;
    float2* g = ws + S.ws_off + (int64_t)ch * S.ws_chunk + b0;
    const int ncol = min(CPW, G1 - c0);
    const FftFactors& F = facs[S.fac2];
    if (F.nf >= 2) {
        // The first pass takes its inputs straight from the grid into the
        // butterflies' registers and the last pass writes its outputs
        // straight to the grid: two LDS round trips fewer than staging the
        // columns and storing them from LDS.  Forward: image row n2 sits at
        // position (n2 - N2/2) mod G2, the others are zero; adjoint: rows
        // n2 < N2 are read back from those positions.  Loads use clamped
        // (valid) addresses and zero by a factor, so none waits on a branch.
        constexpr int SG = ADJ ? -1 : 1;
        auto src = [&](int q, int t) {
            const int col = t >> LCL;
            int r = q;
            float ok = col < ncol ? 1.f : 0.f;
            if (!ADJ) {
                r = q + N2 / 2;
                if (r >= G2) r -= G2;
                ok = r < N2 ? ok : 0.f;
                r = min(r, N2 - 1);
            }
            const float2 v = g[((int64_t)r * G1 + c0 + min(col, ncol - 1)) * BCT + (t & (CL - 1))];
            return make_float2(v.x * ok, v.y * ok);
        };
        fft_first_pass_src<T, SG, kWide>(F.r[0], A, G2, src);
        // the pass twiddles after the data loads, which they then do not delay
        for (int i = threadIdx.x; i < G2 - 1; i += kFftThreads) tw[i] = twp[S.tw2 + i];
        __syncthreads();
        const float2* Rb = fft_lds_range<T, SG, BIG, kWide, CAP>(A, Bb, F, tw, 1, F.nf - 1);
        int p = 1;
        for (int k = 0; k + 1 < F.nf; ++k) p *= F.r[k];
        auto dst = [&](int q, int t, float2 v) {
            const int col = t >> LCL;
            int r = q;
            bool keep = col < ncol;
            if (ADJ) {
                r = q + N2 / 2;
                if (r >= G2) r -= G2;
                keep = keep && r < N2;
            }
            if (keep) g[((int64_t)r * G1 + c0 + col) * BCT + (t & (CL - 1))] = v;
        };
        fft_last_pass_dst<T, SG, BIG, kWide>(F.r[F.nf - 1], Rb, G2, p, tw + (p - 1), dst);
        return;
    }
    // four loads per thread in flight, from clamped (valid) addresses; a
    // column past the grid's last is scaled by 0, and full rounds store
    // without a predicate, so no load sits behind a branch
    auto stage = [&](int rows, auto dst_of) {
        const int total = rows * T;
        auto round = [&](int i0, bool full) {
            float2 v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int idx = i0 + u * 256;
                const int t = idx & (T - 1);
                const int col = t >> LCL;
                const float ok = col < ncol ? 1.f : 0.f;
                const int r = min(idx >> LT, rows - 1);
                const float2 q = g[((int64_t)r * G1 + c0 + min(col, ncol - 1)) * BCT + (t & (CL - 1))];
                v[u] = make_float2(q.x * ok, q.y * ok);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int idx = i0 + u * 256;
                if (full || idx < total) A[dst_of(idx)] = v[u];
            }
        };
        int i0 = threadIdx.x;
        for (; i0 + 3 * 256 < total; i0 += 4 * 256) round(i0, true);
        if (i0 < total) round(i0, false);
    };
    if (!ADJ) {
        for (int i = (N2 - N2 / 2) * T + threadIdx.x; i < (G2 - N2 / 2) * T; i += kFftThreads)
            A[i] = make_float2(0.f, 0.f);
        stage(N2, [&](int idx) {
            int pos = (idx >> LT) - N2 / 2;
            if (pos < 0) pos += G2;
            return pos * T + (idx & (T - 1));
        });
    } else {
        stage(G2, [](int idx) { return idx; });
    }
    // the pass twiddles after the data loads, which they then do not delay
    for (int i = threadIdx.x; i < G2 - 1; i += kFftThreads) tw[i] = twp[S.tw2 + i];
    __syncthreads();
    const float2* R = fft_lds<T, ADJ ? -1 : 1, BIG, kWide, CAP>(A, Bb, facs[S.fac2], tw);
    if (!ADJ) {
        const int total = G2 * T;
        for (int idx = threadIdx.x; idx < total; idx += kFftThreads) {
            const int t = idx & (T - 1);
            const int col = t >> LCL;
            if (col >= ncol) continue;
            g[((idx >> LT) * G1 + c0 + col) * BCT + (t & (CL - 1))] = R[idx];
        }
    } else {
        const int total = N2 * T;
        for (int idx = threadIdx.x; idx < total; idx += kFftThreads) {
            const int t = idx & (T - 1);
            const int i2 = idx >> LT;
            const int col = t >> LCL;
            if (col >= ncol) continue;
            int pos = i2 - N2 / 2;
            if (pos < 0) pos += G2;
            g[(i2 * G1 + c0 + col) * BCT + (t & (CL - 1))] = R[pos * T + t];
        }
    }
}

// Kernel weights of one sample: W1 values for dim 1, W2 for dim 2.  With a
// compile-time width (WT > 0, 2-D) they come in as 16-byte vector loads.
template <int WT>
__device__ __forceinline__ void load_weights(const float* __restrict__ wt, int W1, int W2, float* w1,
                                             float* w2) {
    if constexpr (WT == 4) {
        const float4 a = *reinterpret_cast<const float4*>(wt);
        const float4 b = *reinterpret_cast<const float4*>(wt + 4);
        w1[0] = a.x; w1[1] = a.y; w1[2] = a.z; w1[3] = a.w;
        w2[0] = b.x; w2[1] = b.y; w2[2] = b.z; w2[3] = b.w;
    } else {
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            w1[s] = s < W1 ? wt[s] : 0.f;
            w2[s] = s < W2 ? wt[W1 + s] : 0.f;
        }
    }
}

// Degridding with BCT lanes per sample (one transform each): a load
// instruction reads BCT*8 contiguous bytes for each of 64/BCT samples, each
// lane keeps one accumulator.  Taps are summed in the same order as
// nu_degrid.  Work item = (slice, chunk, first sample) covering 256 samples.
template <int BCT, int WT>
__global__ __launch_bounds__(256) void nu_degrid_t(const NuSliceDev* __restrict__ sl,
                                                   const int4* __restrict__ work,
                                                   const uint32_t* __restrict__ sbase,
                                                   const float* __restrict__ swt,
                                                   const float* __restrict__ sdcf,
                                                   const float2* __restrict__ ws,
                                                   float2* __restrict__ y, NuRun run) {
    constexpr int SPP = 256 / BCT;
    const int4 wk = work[blockIdx.x];
    if (skipped(run, wk.x)) return;
    const NuSliceDev& S = sl[wk.x];
    const int ch = wk.y;
    const int b = threadIdx.x & (BCT - 1);
    const int G1 = S.G1, G2 = S.G2;
    const int W1 = WT > 0 ? WT : S.W1, W2 = WT > 0 ? WT : S.W2;
    const int lanes = lanes_of(run, BCT, ch);
    const int64_t stride = chunk_stride(run, BCT);
    // uniform base of the chunk's grid; a chunk is far below 2^31 elements,
    // so taps are addressed with 32-bit element offsets from it
    const float2* zg = ws + S.ws_off + (int64_t)ch * S.ws_chunk;
    const int mend = min(S.M, wk.z + 256);
    for (int m = wk.z + (int)(threadIdx.x / BCT); m < mend; m += SPP) {
        const int64_t sm = S.samp_off + m;
        const uint32_t base = sbase[sm];
        const int g1_0 = base & 0xffff, g2_0 = base >> 16;
        float w1[8], w2[8];
        load_weights<WT>(swt + sm * (W1 + W2), W1, W2, w1, w2);
        float2 acc = make_float2(0.f, 0.f);
        int c1[8];
#pragma unroll
        for (int s1 = 0; s1 < (WT > 0 ? WT : 8); ++s1) {
            int g1 = g1_0 + s1;
            if (g1 >= G1) g1 -= G1;
            c1[s1] = g1 * BCT + b;
        }
#pragma unroll
        for (int s2 = 0; s2 < (WT > 0 ? WT : 8); ++s2) {
            if (s2 >= W2) break;
            int g2 = g2_0 + s2;
            if (g2 >= G2) g2 -= G2;
            const int row = g2 * G1 * BCT;
#pragma unroll
            for (int s1 = 0; s1 < (WT > 0 ? WT : 8); ++s1) {
                if (s1 >= W1) break;
                const float w = w1[s1] * w2[s2];
                const float2 q = zg[(unsigned)(row + c1[s1])];
                acc.x += w * q.x;
                acc.y += w * q.y;
            }
        }
        const float d = sdcf[sm];
        if (run.ylayout) {
            float f = b < lanes ? d : 0.f;
            if (run.W && b < lanes) f *= run.W[S.dat_off + ((int64_t)ch * stride + b) * S.M + m];
            y[S.yt_off + ((int64_t)ch * S.M + m) * BCT + b] = make_float2(acc.x * f, acc.y * f);
        } else if (b < lanes) {
            const int64_t o = S.dat_off + ((int64_t)ch * stride + b) * S.M + m;
            float f = d;
            if (run.W) f *= run.W[o];
            y[o] = make_float2(acc.x * f, acc.y * f);
        }
    }
}

// Toeplitz normal operator, 2-D middle stage in one pass: per column, the
// forward FFT along g2 of the N2 image rows (zero elsewhere, wrapped as in
// nu_cols), the multiply by the kernel spectrum K[g2][g1] in LDS, the
// inverse FFT and the store of the N2 rows the adjoint rows read back; the
// grid rows beyond N2 never reach HBM.  Items and LDS as nu_cols.
template <int BCT, int FM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(fft_waves(FM, cols_T(BCT, FM == kFftBig), 4), 8))) void nu_cols_toep(
    const NuSliceDev* __restrict__ sl, const int4* __restrict__ work, const FftFactors* __restrict__ facs,
    const float2* __restrict__ twp, const float2* __restrict__ km, const int64_t* __restrict__ koff,
    float2* __restrict__ ws, const int* __restrict__ skip) {
    constexpr bool BIG = FM == kFftBig;
    constexpr int T = cols_T(BCT, BIG), LT = clog2(T), CL = cols_CL(BCT), LCL = clog2(CL), CPW = T / CL;
    constexpr int CAP = fft_mode_cap(FM, T);
    constexpr bool kWide = T >= 16 && !BIG;
    extern __shared__ float2 lds[];
    const int4 wk = work[blockIdx.x];
    if (skip && skip[wk.x]) return;
    const NuSliceDev& S = sl[wk.x];
    const int ch = wk.y, c0 = wk.z, b0 = wk.w;
    const int N2 = S.N2, G1 = S.G1, G2 = S.G2;
    const bool ip = (S.fip >> 1) & 1;
    float2* A = lds;
    float2* Bb = ip ? nullptr : lds + G2 * T;
    float2* tw = lds + (ip ? 1 : 2) * G2 * T;
    float2* g = ws + S.ws_off + (int64_t)ch * S.ws_chunk + b0;
    const float2* K = km + koff[wk.x];
    const int ncol = min(CPW, G1 - c0);
    const FftFactors& F = facs[S.fac2];
    auto src = [&](int q, int t) {
        const int col = t >> LCL;
        int r = q + N2 / 2;
        if (r >= G2) r -= G2;
        const float ok = col < ncol && r < N2 ? 1.f : 0.f;
        r = min(r, N2 - 1);
        const float2 v = g[((int64_t)r * G1 + c0 + min(col, ncol - 1)) * BCT + (t & (CL - 1))];
        return make_float2(v.x * ok, v.y * ok);
    };
    fft_first_pass_src<T, 1, kWide>(F.r[0], A, G2, src);
    for (int i = threadIdx.x; i < G2 - 1; i += kFftThreads) tw[i] = twp[S.tw2 + i];
    __syncthreads();
    float2* R = const_cast<float2*>(fft_lds_range<T, 1, BIG, kWide, CAP>(A, Bb, F, tw, 1, F.nf));
    __syncthreads();
    for (int idx = threadIdx.x; idx < G2 * T; idx += kFftThreads) {
        const int t = idx & (T - 1);
        const float2 k = K[(idx >> LT) * G1 + c0 + min(t >> LCL, ncol - 1)];
        const float2 v = R[idx];
        R[idx] = make_float2(v.x * k.x - v.y * k.y, v.x * k.y + v.y * k.x);
    }
    __syncthreads();
    const float2* Rb = fft_lds_range<T, -1, BIG, kWide, CAP>(R, R == A ? Bb : A, F, tw, 0, F.nf - 1);
    int p = 1;
    for (int k = 0; k + 1 < F.nf; ++k) p *= F.r[k];
    auto dst = [&](int q, int t, float2 v) {
        const int col = t >> LCL;
        int r = q + N2 / 2;
        if (r >= G2) r -= G2;
        if (col < ncol && r < N2) g[((int64_t)r * G1 + c0 + col) * BCT + (t & (CL - 1))] = v;
    };
    fft_last_pass_dst<T, -1, BIG, kWide>(F.r[F.nf - 1], Rb, G2, p, tw + (p - 1), dst);
}

// Toeplitz normal operator, middle stage: each slice's grid Z[g2][g1][b]
// (after the forward row and column FFTs) times the spectrum of its
// point-spread kernel, K[g2][g1], in place.  Grid: (cells * BCT / 1024,
// slice, chunk); four elements per lane, every load issued before the
// stores.
template <int BCT>
__global__ __launch_bounds__(256) void nu_toep_mul(const NuSliceDev* __restrict__ sl,
                                                   const float2* __restrict__ km,
                                                   const int64_t* __restrict__ koff,
                                                   float2* __restrict__ ws, const int* __restrict__ skip) {
    const int s = blockIdx.y;
    if (skip && skip[s]) return;
    const NuSliceDev& S = sl[s];
    const int n = (int)S.ws_chunk;
    const int e0 = (int)blockIdx.x * 1024 + (int)threadIdx.x;
    if (e0 >= n) return;
    const float2* K = km + koff[s];
    float2* z = ws + S.ws_off + (int64_t)blockIdx.z * S.ws_chunk;
    float2 v[4], k[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int e = min(e0 + u * 256, n - 1);
        v[u] = z[e];
        k[u] = K[e / BCT];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (e0 + u * 256 < n)
            z[e0 + u * 256] = make_float2(v[u].x * k[u].x - v[u].y * k[u].y, v[u].x * k[u].y + v[u].y * k[u].x);
}

// Odd image sides (fftShiftArray, nufftOp.m:258-273).  For N = 2s + 1 the
// ramp exp(-i pi (g - G/2) N / G), g = 0..G-1, is the circular shift by s
// (the wrapped placement of the row and column passes), the constant (-1)^s
// (folded into the deapodisation) and the factor i exp(-i pi g / G), which
// this kernel multiplies into the grid of every chunk (CONJ: the conjugate,
// for the adjoint).  That factor is not G-periodic: taps that wrap past
// g = 0 see it with the other sign, exactly as the reference's grid does.
// Grid [g2][g1][b] of the slice's chunk blockIdx.z; even slices return.
template <int BCT, int CONJ>
__global__ __launch_bounds__(256) void nu_odd_ramp(const NuSliceDev* __restrict__ sl, float2* __restrict__ ws,
                                                   const int* __restrict__ skip) {
    const int s = blockIdx.y;
    if (skip && skip[s]) return;
    const NuSliceDev& S = sl[s];
    const int odd = S.odd;
    if (!odd) return;
    const int n = (int)S.ws_chunk;
    const int e = (int)blockIdx.x * 256 + (int)threadIdx.x;
    if (e >= n) return;
    float2* z = ws + S.ws_off + (int64_t)blockIdx.z * S.ws_chunk;
    const int cell = e / BCT;
    const int g1 = cell % S.G1, g2 = cell / S.G1;
    // angle in half-turns: i per odd side, minus g / G per odd side
    float a = 0.f;
    if (odd & 1) a += 0.5f - (float)g1 / (float)S.G1;
    if (odd & 2) a += 0.5f - (float)g2 / (float)S.G2;
    float sn, cs;
    sincospif(CONJ ? -a : a, &sn, &cs);
    const float2 v = z[e];
    z[e] = make_float2(v.x * cs - v.y * sn, v.x * sn + v.y * cs);
}

// Degridding from an LDS-staged grid region (2-D): work item = (slice,
// chunk, sample block); a block holds up to bcap samples whose first
// tap lies in one kDegridTile x kDegridTile tile (or two such tiles side by
// side along g1 whose samples together fit one block), in sample order, with
// their first taps, weights, dcf and ids stored contiguously in block order
// (plan time).  The work-group copies the block's sample data and the tile's
// (kDegridTile + W - 1)^2 cells (wrapped) into LDS; each (sample, transform)
// lane then sums its W^2 taps from LDS in the order and with the products of
// nu_degrid_t, so the two give identical results.
template <int BCT, int WT>
__global__ __launch_bounds__(256) void nu_degrid_tile(const NuSliceDev* __restrict__ sl,
                                                      const int4* __restrict__ work,
                                                      const int4* __restrict__ blocks,
                                                      const int* __restrict__ bid,
                                                      const uint32_t* __restrict__ bbase,
                                                      const float* __restrict__ bwt,
                                                      const float* __restrict__ bdcf,
                                                      const float2* __restrict__ ws,
                                                      float2* __restrict__ y, int bcap, NuRun run) {
    constexpr int SPP = 256 / BCT;
    extern __shared__ float2 lds[];
    const int4 wk = work[blockIdx.x];
    if (skipped(run, wk.x)) return;
    const NuSliceDev& S = sl[wk.x];
    const int4 blk = blocks[wk.z];   // (t1 | t2 << 16, first block-order entry, samples, tile width)
    const int ch = wk.y;
    const int b = threadIdx.x & (BCT - 1);
    const int G1 = S.G1, G2 = S.G2;
    const int W1 = WT > 0 ? WT : S.W1, W2 = WT > 0 ? WT : S.W2, WS = W1 + W2;
    // staged region: the tile (kDegridTile, or twice that along g1 for two
    // merged sparse tiles: blk.w) plus the kernel's reach
    const int RW1 = blk.w + (WT > 0 ? WT : (W1 > W2 ? W1 : W2)) - 1;
    const int RW2 = kDegridTile + (WT > 0 ? WT : (W1 > W2 ? W1 : W2)) - 1;
    const int t1 = blk.x & 0xffff, t2 = blk.x >> 16;
    const int n = blk.z;
    const int lanes = lanes_of(run, BCT, ch);
    const int64_t stride = chunk_stride(run, BCT);
    const float2* zg = ws + S.ws_off + (int64_t)ch * S.ws_chunk;
    // LDS: grid cells [RW2][RW1][BCT], then per sample: id, first taps, dcf, weights
    const int nst = RW1 * RW2 * BCT;
    int* s_id = (int*)(lds + nst);
    uint32_t* s_base = (uint32_t*)(s_id + bcap);
    float* s_dcf = (float*)(s_base + bcap);
    float* s_w = s_dcf + bcap;
    // every staging load of a pass is issued before its LDS stores, with no
    // branch between them: the block's first taps, dcf and ids (n <= 256),
    // its weights (four per lane per pass) and the grid region (four per
    // lane per pass; wrapped cells by a select unless the region is wider
    // than the grid, a uniform case)
    const bool wide = RW1 > G1 || RW2 > G2;
    auto cell = [&](int i) {
        const int cl = i / BCT;
        const int l2 = cl / RW1, l1 = cl - l2 * RW1;
        int g1 = t1 + l1, g2 = t2 + l2;
        if (wide) {
            g1 %= G1;
            g2 %= G2;
        } else {
            g1 = g1 >= G1 ? g1 - G1 : g1;
            g2 = g2 >= G2 ? g2 - G2 : g2;
        }
        return (g2 * G1 + g1) * BCT + (i & (BCT - 1));
    };
    {
        const int i = min((int)threadIdx.x, n - 1);
        const int pid = bid[blk.y + i];
        const uint32_t pb = bbase[blk.y + i];
        const float pd = bdcf[blk.y + i];
        const int nw = n * WS;
        const float* wsrc = bwt + (int64_t)blk.y * WS;
        float w[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) w[u] = wsrc[min((int)threadIdx.x + u * 256, nw - 1)];
        // NV values per lane in the first round: the whole region of two
        // merged 8x8 tiles (19 x 11 cells x 8 transforms) in one round trip
        constexpr int NV = 8;
        float2 v[NV];
#pragma unroll
        for (int u = 0; u < NV; ++u) v[u] = zg[cell(min((int)threadIdx.x + u * 256, nst - 1))];
        if ((int)threadIdx.x < n) {
            s_id[threadIdx.x] = pid;
            s_base[threadIdx.x] = pb;
            s_dcf[threadIdx.x] = pd;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if ((int)threadIdx.x + u * 256 < nw) s_w[threadIdx.x + u * 256] = w[u];
#pragma unroll
        for (int u = 0; u < NV; ++u)
            if ((int)threadIdx.x + u * 256 < nst) lds[threadIdx.x + u * 256] = v[u];
        for (int i0 = threadIdx.x + 4 * 256; i0 < nw; i0 += 256) s_w[i0] = wsrc[i0];
        for (int i0 = threadIdx.x + NV * 256; i0 < nst; i0 += 4 * 256) {
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = zg[cell(min(i0 + u * 256, nst - 1))];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (i0 + u * 256 < nst) lds[i0 + u * 256] = v[u];
        }
    }
    __syncthreads();
    for (int k = (int)(threadIdx.x / BCT); k < n; k += SPP) {
        const int m = s_id[k];
        const uint32_t base = s_base[k];
        const int l1_0 = (int)(base & 0xffff) - t1, l2_0 = (int)(base >> 16) - t2;
        float w1[8], w2[8];
        const float* wk8 = s_w + k * WS;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            w1[q] = q < W1 ? wk8[q] : 0.f;
            w2[q] = q < W2 ? wk8[W1 + q] : 0.f;
        }
        float2 acc = make_float2(0.f, 0.f);
#pragma unroll
        for (int s2 = 0; s2 < (WT > 0 ? WT : 8); ++s2) {
            if (s2 >= W2) break;
            const float2* row = lds + ((l2_0 + s2) * RW1 + l1_0) * BCT + b;
#pragma unroll
            for (int s1 = 0; s1 < (WT > 0 ? WT : 8); ++s1) {
                if (s1 >= W1) break;
                const float w = w1[s1] * w2[s2];
                const float2 q = row[s1 * BCT];
                acc.x += w * q.x;
                acc.y += w * q.y;
            }
        }
        const float d = s_dcf[k];
        if (run.ylayout) {
            float f = b < lanes ? d : 0.f;
            if (run.W && b < lanes) f *= run.W[S.dat_off + ((int64_t)ch * stride + b) * S.M + m];
            y[S.yt_off + ((int64_t)ch * S.M + m) * BCT + b] = make_float2(acc.x * f, acc.y * f);
        } else if (b < lanes) {
            const int64_t o = S.dat_off + ((int64_t)ch * stride + b) * S.M + m;
            float f = d;
            if (run.W) f *= run.W[o];
            y[o] = make_float2(acc.x * f, acc.y * f);
        }
    }
}

template <int BCT, int WT>
__global__ __launch_bounds__(256, 2) void nu_degrid(const NuSliceDev* __restrict__ sl,
                                                 const int4* __restrict__ work,
                                                 const uint32_t* __restrict__ sbase,
                                                 const float* __restrict__ swt,
                                                 const float* __restrict__ sdcf,
                                                 const float2* __restrict__ ws,
                                                 float2* __restrict__ y, NuRun run) {
    const int4 wk = work[blockIdx.x];
    if (skipped(run, wk.x)) return;
    const NuSliceDev& S = sl[wk.x];
    const int ch = wk.y;
    const int m = wk.z + threadIdx.x;
    if (m >= S.M) return;
    const int G1 = S.G1, G2 = S.G2;
    const int W1 = WT > 0 ? WT : S.W1, W2 = WT > 0 ? WT : S.W2;
    const int64_t sm = S.samp_off + m;
    const uint32_t base = sbase[sm];
    const int g1_0 = base & 0xffff, g2_0 = base >> 16;
    float w1[8], w2[8];
    load_weights<WT>(swt + sm * (W1 + W2), W1, W2, w1, w2);
    float2 acc[BCT];
#pragma unroll
    for (int b = 0; b < BCT; ++b) acc[b] = make_float2(0.f, 0.f);
    const float2* zg = ws + S.ws_off + (int64_t)ch * S.ws_chunk;
#pragma unroll 1
    for (int s2 = 0; s2 < (WT > 0 ? WT : 8); ++s2) {
        if (s2 >= W2) break;
        int g2 = g2_0 + s2;
        if (g2 >= G2) g2 -= G2;
        const float2* zrow = zg + (int64_t)g2 * G1 * BCT;
#pragma unroll
        for (int s1 = 0; s1 < (WT > 0 ? WT : 8); ++s1) {
            if (s1 >= W1) break;
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
    if (run.ylayout) {
        // transposed layout: the BCT values of a sample are contiguous
        float2* yt = y + S.yt_off + ((int64_t)ch * S.M + m) * BCT;
#pragma unroll
        for (int b = 0; b < BCT; ++b) {
            float f = b < lanes ? d : 0.f;
            if (run.W && b < lanes) f *= run.W[S.dat_off + ((int64_t)ch * stride + b) * S.M + m];
            yt[b] = make_float2(acc[b].x * f, acc[b].y * f);
        }
        return;
    }
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

// Standard data layout -> transposed [M][BCT] layout (LDS transpose of 256
// samples x BCT lanes per work-group; lanes beyond the valid count are 0).
template <int BCT>
__global__ __launch_bounds__(256) void nu_transpose_y(const NuSliceDev* __restrict__ sl,
                                                      const int4* __restrict__ work,
                                                      const float2* __restrict__ y,
                                                      float2* __restrict__ yt, NuRun run) {
    __shared__ float2 buf[256 * BCT];
    const int4 wk = work[blockIdx.x];
    const NuSliceDev& S = sl[wk.x];
    const int ch = wk.y, m0 = wk.z;
    const int n = min(256, S.M - m0);
    const int lanes = lanes_of(run, BCT, ch);
    const int64_t stride = chunk_stride(run, BCT);
    for (int b = 0; b < BCT; ++b) {
        const int t = threadIdx.x;
        buf[t * BCT + b] = (t < n && b < lanes) ? y[S.dat_off + ((int64_t)ch * stride + b) * S.M + m0 + t]
                                                : make_float2(0.f, 0.f);
    }
    __syncthreads();
    float2* dst = yt + S.yt_off + ((int64_t)ch * S.M + m0) * BCT;
    for (int i = threadIdx.x; i < n * BCT; i += kFftThreads) dst[i] = buf[i];
}

// CSR gridding: entries per work item (lanes x entries stays balanced)
constexpr int kCsrItemEntries = 2048;
// accelerator complex dies (each with its own L2); work-groups are dealt to
// them round-robin by workgroup index
constexpr int kXcds = 8;

// CSR gridding: work item = (slice, chunk, first cell, cell count | log2(P)
// << 24).  Lanes are (cell, part, transform) triples, BCT consecutive lanes
// per (cell, part), so the BCT values a sample holds for the chunk are one
// contiguous read of the transposed data [M][BCT].  With P > 1 (rows of the
// dense k-space centre) each of the P parts sums a contiguous 1/P of the
// row and the parts are added in order through LDS; every cell's sum has a
// fixed order for a given plan.
template <int BCT>
__global__ __launch_bounds__(256) void nu_grid_csr(const NuSliceDev* __restrict__ sl,
                                                   const int4* __restrict__ work,
                                                   const int* __restrict__ rowptr,
                                                   const int2* __restrict__ ent,
                                                   const float2* __restrict__ yt, float2* __restrict__ ws,
                                                   NuRun run) {
    constexpr int NG = 256 / BCT;  // lane groups
    __shared__ float2 red[256];
    const int4 wk = work[blockIdx.x];
    if (skipped(run, wk.x)) return;
    const NuSliceDev& S = sl[wk.x];
    const int ch = wk.y;
    const int lp = wk.w >> 24;
    const int P = 1 << lp;
    const int cpi = NG >> lp;      // cells per pass
    const int b = threadIdx.x & (BCT - 1);
    const int grp = threadIdx.x / BCT;
    const int part = grp & (P - 1);
    const int cl = grp >> lp;
    const int lanes = lanes_of(run, BCT, ch);
    const int* rp = rowptr + S.csr_off;
    const int2* en = ent + S.ent_off;
    const float2* ych = yt + S.yt_off + (int64_t)ch * S.M * BCT + b;
    float2* zg = ws + S.ws_off + (int64_t)ch * S.ws_chunk + b;
    const int cend = wk.z + (wk.w & 0xFFFFFF);
    for (int cbase = wk.z; cbase < cend; cbase += cpi) {
        const int cell = cbase + cl;
        float2 acc = make_float2(0.f, 0.f);
        if (cell < cend && b < lanes) {
            const int r0 = rp[cell], len = rp[cell + 1] - r0;
            int e = r0 + (int)(((int64_t)len * part) >> lp);
            const int e1 = r0 + (int)(((int64_t)len * (part + 1)) >> lp);
            for (; e + 4 <= e1; e += 4) {
                int2 q[4];
                float2 v[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) q[i] = en[e + i];
#pragma unroll
                for (int i = 0; i < 4; ++i) v[i] = ych[(int64_t)q[i].x * BCT];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float w = __int_as_float(q[i].y);
                    acc.x = fmaf(w, v[i].x, acc.x);
                    acc.y = fmaf(w, v[i].y, acc.y);
                }
            }
            for (; e < e1; ++e) {
                const int2 q = en[e];
                const float2 v = ych[(int64_t)q.x * BCT];
                const float w = __int_as_float(q.y);
                acc.x = fmaf(w, v.x, acc.x);
                acc.y = fmaf(w, v.y, acc.y);
            }
        }
        if (P == 1) {
            if (cell < cend) zg[(int64_t)cell * BCT] = acc;
            continue;
        }
        red[threadIdx.x] = acc;
        __syncthreads();
        if (part == 0 && cell < cend) {
            float2 t = acc;
            for (int q = 1; q < P; ++q) {
                const float2 u = red[threadIdx.x + q * BCT];
                t.x += u.x;
                t.y += u.y;
            }
            zg[(int64_t)cell * BCT] = t;
        }
        __syncthreads();
    }
}

// Staged CSR gridding: one work item per rectangle (or chunk of a dense
// rectangle) of one or (SG = 2) two (slice, chunk) units of one trajectory.
// The work-group first copies, with every lane, the transposed data of its
// samples ([n_stage][BCT * SG]: the second unit's transforms after the
// first's), its (local sample, weight) rows and row pointers into LDS; each
// (cell, transform) lane then sums its row from LDS in sample order -- the
// products and the order of a sum over the global CSR row.  Kind 0 writes the
// grid cells; kind 2 (a chunk of a dense rectangle) writes a partial block,
// which nu_grid_merge adds in chunk order into the grid (deterministic for a
// given plan).  A unit whose slice has finished is neither read nor written.
//
// GridStage holds what one lane loads for an item before its LDS stores:
// NPRE staged samples (one transform each), the first 1024 entries and the
// first 512 row pointers (four and two per lane).  Loads use clamped, valid
// indices, so no load sits behind a branch.
template <int BCT, int SG>
struct GridStage {
    static constexpr int TT = BCT * SG;   // transforms per staged sample
    static constexpr int SPR = 256 / TT;  // staged samples per pass of the work-group
    static constexpr int NPRE = 8;        // sample ids each lane loads up front
    float2 v[NPRE];
    int2 e[4];
    int pr[2];
};

// sample ids of item i that lane (q0) stages, addressed by the item index
// alone (so they load together with the item record, not after it)
template <int NPRE, int SPR>
__device__ __forceinline__ void grid_item_ids(const int* __restrict__ isid, int sid_stride, int i, int q0,
                                              int (&m)[NPRE]) {
    const int* sid = isid + (int64_t)i * sid_stride;
#pragma unroll
    for (int k = 0; k < NPRE; ++k) m[k] = sid[min(q0 + k * SPR, sid_stride - 1)];
}

// The loads of one item into registers.  act0 / act1: the units whose slice
// is still active; a lane of a finished unit loads sample 0 (its value is
// never read), an item with no active unit entry 0 and row pointer 0.
template <int BCT, int SG>
__device__ __forceinline__ void grid_item_load(GridStage<BCT, SG>& g, const GridItem& T, bool act0, bool act1,
                                               const int (&m)[GridStage<BCT, SG>::NPRE],
                                               const int* __restrict__ tptr, const int2* __restrict__ tent,
                                               const float2* __restrict__ yt) {
    constexpr int TT = GridStage<BCT, SG>::TT;
    const int jb = (int)threadIdx.x & (TT - 1);
    const bool any = act0 || act1;
    const int cells = T.w * T.h;
    const int2* te = tent + T.ent_off;
    const int* tp = tptr + T.ptr_off;
    const int ne_last = any ? max(T.n_ent - 1, 0) : 0;
    const int pr_last = any ? cells : 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) g.e[u] = te[min((int)threadIdx.x + u * 256, ne_last)];
#pragma unroll
    for (int u = 0; u < 2; ++u) g.pr[u] = tp[min((int)threadIdx.x + u * 256, pr_last)];
    const int u1 = SG > 1 ? jb / BCT : 0;
    const bool act = u1 ? act1 : act0;
    const float2* ych = yt + (u1 ? T.yt2 : T.yt) + (jb & (BCT - 1));
#pragma unroll
    for (int k = 0; k < GridStage<BCT, SG>::NPRE; ++k) g.v[k] = ych[(int64_t)(act ? m[k] : 0) * BCT];
}

// The LDS stores of one item (registers from grid_item_load; samples,
// entries and row pointers past the first round load here, straight from
// global memory: only with caps above 8 samples per lane, 1024 entries or
// 511 cells).  LDS: sv [n_stage][TT], se [n_ent], sp [cells + 1].
template <int BCT, int SG>
__device__ __forceinline__ void grid_item_store(const GridStage<BCT, SG>& g, const GridItem& T, bool act0,
                                                bool act1, const int* __restrict__ sid,
                                                const int* __restrict__ tptr, const int2* __restrict__ tent,
                                                const float2* __restrict__ yt, float2* lds) {
    constexpr int TT = GridStage<BCT, SG>::TT, SPR = GridStage<BCT, SG>::SPR, NPRE = GridStage<BCT, SG>::NPRE;
    const int jb = (int)threadIdx.x & (TT - 1);
    const int q0 = (int)threadIdx.x / TT;
    const int cells = T.w * T.h;
    float2* sv = lds;
    int2* se = (int2*)(lds + T.n_stage * TT);
    int* sp = (int*)(se + T.n_ent);
#pragma unroll
    for (int k = 0; k < NPRE; ++k)
        if (q0 + k * SPR < T.n_stage) sv[(q0 + k * SPR) * TT + jb] = g.v[k];
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if ((int)threadIdx.x + u * 256 < T.n_ent) se[threadIdx.x + u * 256] = g.e[u];
#pragma unroll
    for (int u = 0; u < 2; ++u)
        if ((int)threadIdx.x + u * 256 <= cells) sp[threadIdx.x + u * 256] = g.pr[u];
    if (T.n_stage > NPRE * SPR) {
        const int u1 = SG > 1 ? jb / BCT : 0;
        const bool act = u1 ? act1 : act0;
        const float2* ych = yt + (u1 ? T.yt2 : T.yt) + (jb & (BCT - 1));
        for (int q = q0 + NPRE * SPR; q < T.n_stage; q += SPR)
            sv[q * TT + jb] = ych[(int64_t)(act ? sid[q] : 0) * BCT];
    }
    const int2* te = tent + T.ent_off;
    for (int i = threadIdx.x + 4 * 256; i < T.n_ent; i += 256) se[i] = te[i];
    const int* tp = tptr + T.ptr_off;
    for (int i = threadIdx.x + 2 * 256; i <= cells; i += 256) sp[i] = tp[i];
}

// The gather of one staged item: CPL consecutive transforms of one unit per
// lane (one 8*CPL-byte LDS read of the sample's values per entry), TT/CPL
// lanes per cell.
template <int BCT, int SG>
__device__ __forceinline__ void grid_item_gather(const GridItem& T, bool act0, bool act1, const float2* lds,
                                                 float2* __restrict__ ws, float2* __restrict__ parts) {
    constexpr int TT = GridStage<BCT, SG>::TT;
    constexpr int CPL = BCT >= 2 ? 2 : 1;
    constexpr int LPC = TT / CPL;
    constexpr int NGC = 256 / LPC;
    const int cells = T.w * T.h;
    const float2* sv = lds;
    const int2* se = (const int2*)(lds + T.n_stage * TT);
    const int* sp = (const int*)(se + T.n_ent);
    const int bc = (int)(threadIdx.x & (LPC - 1)) * CPL;
    const int gc = (int)threadIdx.x / LPC;
    const int u1 = SG > 1 ? bc / BCT : 0;
    const int bl = bc & (BCT - 1);
    const int lanes = u1 ? T.lanes2 : T.lanes;
    const bool act = u1 ? act1 : act0;
    float2* zg = ws + (u1 ? T.zg2 : T.zg) + bl;
    float2* pb = parts + (u1 ? T.part2 : T.part) + (int64_t)T.pofs * BCT + bl;   // kind 2: this chunk's block
    const float2* svb = sv + bc;
    for (int c = gc; c < cells; c += NGC) {
        float2 acc[CPL];
#pragma unroll
        for (int k = 0; k < CPL; ++k) acc[k] = make_float2(0.f, 0.f);
        if (act && bl < lanes) {
            int e = sp[c];
            const int e1 = sp[c + 1];
            for (; e + 4 <= e1; e += 4) {
                int2 q[4];
                float2 v[4][CPL];
#pragma unroll
                for (int i = 0; i < 4; ++i) q[i] = se[e + i];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if constexpr (CPL == 2) {
                        const float4 u = *reinterpret_cast<const float4*>(svb + q[i].x * TT);
                        v[i][0] = make_float2(u.x, u.y);
                        v[i][CPL - 1] = make_float2(u.z, u.w);
                    } else {
                        v[i][0] = svb[q[i].x * TT];
                    }
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float w = __int_as_float(q[i].y);
#pragma unroll
                    for (int k = 0; k < CPL; ++k) {
                        acc[k].x = fmaf(w, v[i][k].x, acc[k].x);
                        acc[k].y = fmaf(w, v[i][k].y, acc[k].y);
                    }
                }
            }
            for (; e < e1; ++e) {
                const int2 q = se[e];
                const float w = __int_as_float(q.y);
#pragma unroll
                for (int k = 0; k < CPL; ++k) {
                    const float2 v = svb[q.x * TT + k];
                    acc[k].x = fmaf(w, v.x, acc[k].x);
                    acc[k].y = fmaf(w, v.y, acc[k].y);
                }
            }
            // a transform past the chunk's last one stays zero
#pragma unroll
            for (int k = 1; k < CPL; ++k)
                if (bl + k >= lanes) acc[k] = make_float2(0.f, 0.f);
        }
        if (!act) continue;
        float2* dst;
        if (T.kind == 0) {
            const int r = c / T.w;
            dst = zg + ((int64_t)(T.g2 + r) * T.G1 + T.g1 + (c - r * T.w)) * BCT;
        } else {
            dst = pb + c * BCT;
        }
        if constexpr (CPL == 2)
            *reinterpret_cast<float4*>(dst) = make_float4(acc[0].x, acc[0].y, acc[CPL - 1].x, acc[CPL - 1].y);
        else
            dst[0] = acc[0];
    }
}

// One work-group per item.
template <int BCT, int SG>
__global__ __launch_bounds__(256) void nu_grid_tile(const GridItem* __restrict__ work,
                                                    const int* __restrict__ isid, int sid_stride,
                                                    const int* __restrict__ tptr,
                                                    const int2* __restrict__ tent,
                                                    const float2* __restrict__ yt, float2* __restrict__ ws,
                                                    float2* __restrict__ parts, const int* __restrict__ skip) {
    using GS = GridStage<BCT, SG>;
    static_assert(256 % GS::TT == 0, "transforms per sample must divide the work-group");
    extern __shared__ float2 lds[];
    const int q0 = (int)threadIdx.x / GS::TT;
    int m[GS::NPRE];
    grid_item_ids<GS::NPRE, GS::SPR>(isid, sid_stride, blockIdx.x, q0, m);
    const GridItem T = work[blockIdx.x];
    const bool act0 = !(skip && skip[T.slice]);
    const bool act1 = SG > 1 && T.slice2 >= 0 && !(skip && skip[T.slice2]);
    if (!act0 && !act1) return;
    GS g;
    grid_item_load<BCT, SG>(g, T, act0, act1, m, tptr, tent, yt);
    grid_item_store<BCT, SG>(g, T, act0, act1, isid + (int64_t)blockIdx.x * sid_stride, tptr, tent, yt, lds);
    __syncthreads();
    grid_item_gather<BCT, SG>(T, act0, act1, lds, ws, parts);
}

// Sum of the partial blocks of each dense rectangle in chunk order (one item
// per rectangle and (slice, chunk): its first chunk's record), written to the
// grid.
template <int BCT>
__global__ __launch_bounds__(256) void nu_grid_merge(const GridItem* __restrict__ work,
                                                     const float2* __restrict__ parts, float2* __restrict__ ws,
                                                     const int* __restrict__ skip) {
    const GridItem T = work[blockIdx.x];
    if (skip && skip[T.slice]) return;
    const float2* p0 = parts + T.part + (int64_t)T.pofs0 * BCT;
    const int nv = T.w * T.h * BCT;
    for (int i = threadIdx.x; i < nv; i += kFftThreads) {
        float2 t = make_float2(0.f, 0.f);
        for (int k0 = 0; k0 < T.nparts; k0 += 8) {
            float2 u[8];
#pragma unroll
            for (int k = 0; k < 8; ++k)
                u[k] = k0 + k < T.nparts ? p0[(int64_t)(k0 + k) * nv + i] : make_float2(0.f, 0.f);
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (k0 + k < T.nparts) {
                    t.x += u[k].x;
                    t.y += u[k].y;
                }
        }
        const int c = i / BCT;
        const int r = c / T.w;
        ws[T.zg + ((int64_t)(T.g2 + r) * T.G1 + T.g1 + (c - r * T.w)) * BCT + (i & (BCT - 1))] = t;
    }
}

// Adjoint rows: FFT along g1, crop to N1 positions, deapodisation and (SENSE)
// the conjugate-coil combine.  In the combine two lanes share an output
// pixel, each summing half of the coils; the halves are added through a
// lane shuffle.
template <int BCT, int FM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(fft_waves(FM, rows_T(BCT, FM == kFftBig), 8), 8))) void nu_adj_rows(const NuSliceDev* __restrict__ sl,
                                                   const int4* __restrict__ work,
                                                   const FftFactors* __restrict__ facs,
                                                   const float2* __restrict__ twp,
                                                   const float* __restrict__ deap,
                                                   const float2* __restrict__ ws,
                                                   float2* __restrict__ x, NuRun run) {
    constexpr bool BIG = FM == kFftBig;
    constexpr int T = rows_T(BCT, BIG), LT = clog2(T), LB = clog2(BCT), RPW = T / BCT, LR = clog2(RPW);
    constexpr int CAP = fft_mode_cap(FM, T);
    constexpr bool kWide = rows_wide(T, FM);
    extern __shared__ float2 lds[];
    const int4 wk = work[blockIdx.x];
    if (skipped(run, wk.x)) return;
    const NuSliceDev& S = sl[wk.x];
    const int ch = wk.y, r0 = wk.z;
    const int N1 = S.N1, N2 = S.N2, G1 = S.G1;
    const int V = N1 * N2;
    const bool ip = S.fip & 1;   // single-buffer row FFTs
    float2* A = lds;
    float2* Bb = ip ? nullptr : lds + G1 * T;
    float2* tw = lds + (ip ? 1 : 2) * G1 * T;
    const float2* g = ws + S.ws_off + (int64_t)ch * S.ws_chunk;
    const FftFactors& F = facs[S.fac1];
    const float2* R;
    if (F.nf >= 2) {
        // the first pass takes its inputs straight from the workspace rows
        // into the butterflies' registers (one LDS round trip fewer); rows
        // past the image are scaled by 0
        fft_first_pass_src<T, -1, kWide>(F.r[0], A, G1, [&](int q, int t) {
            const int i2 = r0 + (t >> LB);
            const float ok = i2 < N2 ? 1.f : 0.f;
            const float2 v = g[(min(i2, N2 - 1) * G1 + q) * BCT + (t & (BCT - 1))];
            return make_float2(v.x * ok, v.y * ok);
        });
        // the pass twiddles after the data loads, which they then do not delay
        for (int i = threadIdx.x; i < G1 - 1; i += kFftThreads) tw[i] = twp[S.tw1 + i];
        __syncthreads();
        R = fft_lds_range<T, -1, BIG, kWide, CAP>(A, Bb, F, tw, 1, F.nf);
    } else {
    const int total = G1 * T;
    // four loads per thread in flight, from clamped (valid) addresses; rows
    // past the image are scaled by 0, and full rounds store without a
    // predicate, so no load sits behind a branch
    auto round = [&](int i0, bool full) {
        float2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int idx = i0 + u * 256;
            const int t = idx & (T - 1);
            const int i2 = r0 + (t >> LB);
            const float ok = i2 < N2 ? 1.f : 0.f;
            const float2 q = g[(min(i2, N2 - 1) * G1 + min(idx >> LT, G1 - 1)) * BCT + (t & (BCT - 1))];
            v[u] = make_float2(q.x * ok, q.y * ok);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (full || i0 + u * 256 < total) A[i0 + u * 256] = v[u];
    };
    {
        int i0 = threadIdx.x;
        for (; i0 + 3 * 256 < total; i0 += 4 * 256) round(i0, true);
        if (i0 < total) round(i0, false);
    }
    // the pass twiddles after the data loads, which they then do not delay
    for (int i = threadIdx.x; i < G1 - 1; i += kFftThreads) tw[i] = twp[S.tw1 + i];
    __syncthreads();
    R = fft_lds<T, -1, BIG, kWide, CAP>(A, Bb, F, tw);
    }
    const int lanes = lanes_of(run, BCT, ch);
    const float* d1 = deap + S.deap1;
    const float* d2 = deap + S.deap2;
    if (run.sense) {
        constexpr int H = BCT >= 2 ? 2 : 1, HB = BCT / H;
        const float2* mp = run.maps + S.maps_off;
        float2* xo = x + S.img_off + (int64_t)ch * V;
        const int tot = N1 * RPW * H;
        // whole waves run every iteration (the shuffle needs both lanes)
        for (int idx0 = 0; idx0 < tot; idx0 += kFftThreads) {
            const int idx = idx0 + threadIdx.x;
            const int h = idx & (H - 1);
            const int r = (idx / H) & (RPW - 1);
            const int i1 = (idx / H) >> LR;
            const int i2 = r0 + r;
            const bool live = idx < tot && i2 < N2;
            // every lane issues its map and deapodization loads at once, from
            // clamped (valid) addresses, and uses them unconditionally (a
            // factor 0 drops coils past the last and dead lanes), so no load
            // sits behind a branch
            const int i1c = min(i1, N1 - 1), i2c = min(i2, N2 - 1);
            const int v = i2c * N1 + i1c;
            const float f1 = d1[i1c], f2 = d2[i2c];
            float2 m[HB];
#pragma unroll
            for (int q = 0; q < HB; ++q) m[q] = mp[min(h * HB + q, lanes - 1) * V + v];
            const float f = f1 * f2;
            int pos = i1c - N1 / 2;
            if (pos < 0) pos += G1;
            const float2* src = R + pos * T + r * BCT;
            float2 s = make_float2(0.f, 0.f);
#pragma unroll
            for (int q = 0; q < HB; ++q) {
                const int b = h * HB + q;
                const float k = live && b < lanes ? 1.f : 0.f;
                const float2 u = src[b];
                const float2 mk = make_float2(m[q].x * k, m[q].y * k);
                s.x += mk.x * u.x + mk.y * u.y;   // conj(maps) * u
                s.y += mk.x * u.y - mk.y * u.x;
            }
            // each half is scaled before the lane sum (the factor is the
            // pixel's on both lanes), so the factor's load is used on every
            // lane and is issued with the map loads
            s.x *= f;
            s.y *= f;
            if constexpr (H == 2) {
                s.x += __shfl_xor(s.x, 1);
                s.y += __shfl_xor(s.y, 1);
            }
            if (live && h == 0) xo[i2 * N1 + i1] = s;
        }
    } else {
        const int tot = N1 * T;
        float2* xo = x + S.img_off + (int64_t)ch * BCT * V;
        for (int idx = threadIdx.x; idx < tot; idx += kFftThreads) {
            const int t = idx & (T - 1);
            const int i1 = idx >> LT;
            const int b = t & (BCT - 1);
            const int i2 = r0 + (t >> LB);
            if (i2 >= N2 || b >= lanes) continue;
            int pos = i1 - N1 / 2;
            if (pos < 0) pos += G1;
            const float2 q = R[pos * T + t];
            const float f = d1[i1] * d2[i2];
            xo[b * V + i2 * N1 + i1] = make_float2(q.x * f, q.y * f);
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
    const std::vector<double> lut = nufft_kb_lut(kwidth, os);
    const int Nk = nkern;

    std::vector<FftFactors> facs;
    std::vector<float2> twid;
    // (G, wide) -> (factor idx, twiddle offset)
    std::map<std::pair<int, bool>, std::pair<int, int>> fft_of;
    auto fft_for = [&](int G, bool wide) {
        auto it = fft_of.find({G, wide});
        if (it != fft_of.end()) return it->second;
        FftFactors f;
        if (!fft_factorize(G, f, wide)) fail(MM_ERR_ARG, "nufftOp: grid size with a prime factor > 7");
        std::pair<int, int> r((int)facs.size(), (int)twid.size());
        facs.push_back(f);
        fft_pass_twiddles(f, [&](float c, float sn) { twid.push_back(make_float2(c, sn)); });
        fft_of[{G, wide}] = r;
        return r;
    };
    std::vector<float> deap;
    std::vector<uint32_t> sbase;
    std::vector<float> swt, sdcf;
    std::vector<int2>& csr_ent_h = csr_ent_host_;
    std::vector<int2> gent_h;
    std::map<uint64_t, int> traj_seen;   // trajectory hash -> first slice with it
    int64_t voff = 0, moff = 0;
    for (int s = 0; s < nslice; ++s) {
        NuSliceDev d{};
        const int64_t n1 = imN[2 * s], n2 = ndim == 2 ? imN[2 * s + 1] : 1;
        if (n1 < 2 || n2 < 1 || (ndim == 2 && n2 < 2)) fail(MM_ERR_ARG, "nufftOp: image size too small");
        d.odd = (int)(n1 & 1) | (ndim == 2 && (n2 & 1) ? 2 : 0);
        if (d.odd) any_odd_ = true;
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
        auto f1 = fft_for(d.G1, false);
        auto f2 = fft_for(d.G2, false);
        auto f2w = fft_for(d.G2, true);
        auto f1w = fft_for(d.G1, true);
        d.fac1 = f1.first;
        d.tw1 = f1.second;
        d.fac2 = f2.first;
        d.tw2 = f2.second;
        d.fac2w = f2w.first;
        d.tw2w = f2w.second;
        d.fac1w = f1w.first;
        d.tw1w = f1w.second;
        // deapodisation with 1/sqrt(G1*G2) and the real part (-1)^(s1+s2),
        // s = floor(N/2), of the ramp constant exp(i pi (N1+N2)/2) folded in
        // (an odd side's factor i goes with its ramp, nu_odd_ramp)
        const double sgn = (((n1 / 2) + (ndim == 2 ? n2 / 2 : 0)) & 1) ? -1.0 : 1.0;
        const double scl = sgn / std::sqrt((double)d.G1 * d.G2);
        auto c1 = nufft_deapod((int)n1, d.G1, kwidth, os);
        d.deap1 = (int)deap.size();
        for (double v : c1) deap.push_back((float)(v * scl));
        d.deap2 = (int)deap.size();
        if (ndim == 2) {
            auto c2 = nufft_deapod((int)n2, d.G2, kwidth, os);
            for (double v : c2) deap.push_back((float)v);
        } else {
            deap.push_back(1.f);
        }
        const double* kl = kloc[s];
        const double* dc = dcf ? dcf[s] : nullptr;
        if (ndim == 2) {
            // a 2-D slice with the same size and the same trajectory and dcf
            // as an earlier one shares its sample data and CSR gridding rows
            const uint64_t h = traj_hash(kl, dc, d.M, ndim) ^ ((uint64_t)d.N1 << 40) ^ ((uint64_t)d.N2 << 20);
            auto it = traj_seen.find(h);
            int twin = -1;
            if (it != traj_seen.end()) {
                const int t = it->second;
                const NuSliceDev& e = sl_[t];
                if (e.N1 == d.N1 && e.N2 == d.N2 && e.M == d.M &&
                    std::memcmp(kloc[t], kl, sizeof(double) * (size_t)d.M * ndim) == 0 &&
                    ((dcf ? dcf[t] : nullptr) == nullptr) == (dc == nullptr) &&
                    (dc == nullptr || std::memcmp(dcf[t], dc, sizeof(double) * (size_t)d.M) == 0))
                    twin = t;
            }
            if (twin >= 0) {
                const NuSliceDev& e = sl_[twin];
                d.samp_off = e.samp_off;
                d.csr_off = e.csr_off;
                d.ent_off = e.ent_off;
                sl_.push_back(d);
                Vs_.push_back(n1 * n2);
                Ms_.push_back(d.M);
                img_start_.push_back(voff);
                dat_start_.push_back(moff);
                voff += n1 * n2;
                moff += d.M;
                continue;
            }
            traj_seen.emplace(h, s);
        }
        // samples (nufftOp.m:187-246)
        d.samp_off = (int64_t)sbase.size();
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
        {
            // CSR gridding: row pointers (per cell) and (sample, weight) entries
            const int64_t ncell = (int64_t)d.G1 * d.G2;
            const int64_t nent = (int64_t)d.M * d.W1 * d.W2;
            if (nent > (int64_t)INT32_MAX) fail(MM_ERR_ARG, "nufftOp: too many samples in one slice");
            d.csr_off = (int64_t)csr_ptr_host_.size();
            d.ent_off = (int64_t)csr_ent_h.size();
            std::vector<int> cnt((size_t)ncell + 1, 0);
            for (int m = 0; m < d.M; ++m)
                for (int s2 = 0; s2 < d.W2; ++s2) {
                    int g2 = g0[1][m] + s2;
                    g2 -= g2 >= d.G2 ? d.G2 : 0;
                    for (int s1 = 0; s1 < d.W1; ++s1) {
                        int g1 = g0[0][m] + s1;
                        g1 -= g1 >= d.G1 ? d.G1 : 0;
                        cnt[(size_t)g2 * d.G1 + g1 + 1]++;
                    }
                }
            for (int64_t i = 0; i < ncell; ++i) cnt[(size_t)i + 1] += cnt[(size_t)i];
            csr_ptr_host_.insert(csr_ptr_host_.end(), cnt.begin(), cnt.end());
            std::vector<int> pos(cnt.begin(), cnt.end() - 1);
            const size_t e0 = csr_ent_h.size();
            csr_ent_h.resize(e0 + (size_t)nent);
            const float* wrow = swt.data() + (size_t)d.samp_off * Ws;
            for (int m = 0; m < d.M; ++m) {
                const double dm = dc ? dc[m] : 1.0;
                for (int s2 = 0; s2 < d.W2; ++s2) {
                    int g2 = g0[1][m] + s2;
                    g2 -= g2 >= d.G2 ? d.G2 : 0;
                    const double w2 = wrow[(size_t)m * Ws + d.W1 + s2];
                    for (int s1 = 0; s1 < d.W1; ++s1) {
                        int g1 = g0[0][m] + s1;
                        g1 -= g1 >= d.G1 ? d.G1 : 0;
                        const float w = (float)(wrow[(size_t)m * Ws + s1] * w2 * dm);
                        int2 q;
                        q.x = m;
                        q.y = __builtin_bit_cast(int, w);
                        csr_ent_h[e0 + (size_t)pos[(size_t)g2 * d.G1 + g1]++] = q;
                    }
                }
            }
            sl_.push_back(d);
            Vs_.push_back(n1 * n2);
            Ms_.push_back(d.M);
            img_start_.push_back(voff);
            dat_start_.push_back(moff);
            voff += n1 * n2;
            moff += d.M;
        }
    }
    V_total_ = voff;
    M_total_ = moff;
    for (const auto& q : sl_) big_ |= q.G1 > kFftSmallMax || q.G2 > kFftSmallMax;
    hipStream_t st = ctx->stream;
    facs_.upload(facs, st);
    twid_.upload(twid, st);
    deap_.upload(deap, st);
    sbase_.upload(sbase, st);
    swt_.upload(swt, st);
    sdcf_.upload(sdcf, st);
    sbase_host_ = sbase;
    swt_host_ = swt;
    sdcf_host_ = sdcf;
    csr_ptr_.upload(csr_ptr_host_, st);
    csr_ent_.upload(csr_ent_h.empty() ? std::vector<int2>{make_int2(0, 0)} : csr_ent_h, st);
    build_partitions();
    HIPCHECK(hipStreamSynchronize(st));
}

// LDS of one FFT work-group: one (single-buffer passes) or two buffers of
// G*T values and the G-1 pass twiddles
static size_t fft_lds_bytes(int G, int T, bool inplace) {
    return (size_t)((inplace ? 1 : 2) * G * T + G) * sizeof(float2);
}

// Plan-time partitions of every distinct 2-D trajectory (slices that share
// one share its partitions): the staged gridding rectangles (caps
// stage_samples_ / stage_entries_) and the staged degridding blocks
// (degrid_block_ samples per kDegridTile^2 tile), with their pools uploaded.
// Rebuilt when a tuning option changes them; every work table is rebuilt
// with the next configuration.
void NufftOp::build_partitions() {
    ++gen_;
    const int ns = (int)sl_.size();
    gtile_host_.clear();
    dblk_host_.clear();
    gtile_first_.assign((size_t)ns, 0);
    gtile_count_.assign((size_t)ns, 0);
    gpcells_.assign((size_t)ns, 0);
    dblk_first_.assign((size_t)ns, 0);
    dblk_count_.assign((size_t)ns, 0);
    cfgs_.clear();
    if (ndim_ != 2) return;
    std::vector<int> gstage_h, gptr_h, dsamp_h;
    std::vector<int2> gent_h;
    std::vector<uint32_t> dbase_h;
    std::vector<float> ddcf_h, dwt_h;
    std::map<int64_t, int> first_of;   // CSR offset -> first slice with that trajectory
    for (int s = 0; s < ns; ++s) {
        const NuSliceDev& d = sl_[(size_t)s];
        auto it = first_of.find(d.csr_off);
        if (it != first_of.end()) {
            const size_t t = (size_t)it->second;
            gtile_first_[(size_t)s] = gtile_first_[t];
            gtile_count_[(size_t)s] = gtile_count_[t];
            gpcells_[(size_t)s] = gpcells_[t];
            dblk_first_[(size_t)s] = dblk_first_[t];
            dblk_count_[(size_t)s] = dblk_count_[t];
            continue;
        }
        first_of.emplace(d.csr_off, s);
        const int64_t ncell = (int64_t)d.G1 * d.G2;
        const int* rp = csr_ptr_host_.data() + d.csr_off;
        const int2* en = csr_ent_host_.data() + d.ent_off;
        const int64_t nent = rp[ncell];
        gtile_first_[(size_t)s] = (int)gtile_host_.size();
        int npc = 0, nde = 0;
        build_grid_tiles(d.G1, d.G2, d.M, rp, en, stage_samples_, stage_entries_, gtile_host_, gstage_h, gptr_h,
                         gent_h, npc, nde);
        gtile_count_[(size_t)s] = (int)gtile_host_.size() - gtile_first_[(size_t)s];
        gpcells_[(size_t)s] = npc;
        {
            // regions: the rectangles in partition order (spatially coherent)
            // cut into kXcds runs of equal staging cost
            int64_t total = 0, acc = 0;
            for (int t = gtile_first_[(size_t)s]; t < (int)gtile_host_.size(); ++t)
                total += (int64_t)gtile_host_[(size_t)t].n_ent + 2 * (int64_t)gtile_host_[(size_t)t].n_stage + 1;
            for (int t = gtile_first_[(size_t)s]; t < (int)gtile_host_.size(); ++t) {
                GridTile& q = gtile_host_[(size_t)t];
                const int64_t cost = (int64_t)q.n_ent + 2 * (int64_t)q.n_stage + 1;
                q.region = (int)std::min<int64_t>(kXcds - 1, (int64_t)kXcds * (2 * acc + cost) / (2 * total));
                acc += cost;
            }
        }
        {
            // the rectangles (chunks counted once) tile the grid and hold
            // every entry once
            int64_t cells = 0, ents = 0;
            for (int t = gtile_first_[(size_t)s]; t < (int)gtile_host_.size(); ++t) {
                const GridTile& q = gtile_host_[(size_t)t];
                if (q.kind == 0 || q.pofs == q.pofs0) cells += (int64_t)q.w * q.h;
                ents += q.n_ent;
            }
            if (cells != ncell || ents != nent)
                fail(MM_ERR_HIP, "nufftOp: internal: gridding rectangles do not tile the grid");
        }
        // degridding blocks: samples bucketed by the kDegridTile^2 tile of
        // their first tap, in sample order, cut into blocks of at most
        // degrid_block_; ids, first taps, weights and dcf copied in that order
        const int Ws = d.W1 + d.W2;
        const uint32_t* sb = sbase_host_.data() + d.samp_off;
        const int ntx = (d.G1 + kDegridTile - 1) / kDegridTile;
        const int nty = (d.G2 + kDegridTile - 1) / kDegridTile;
        std::vector<int> tcnt((size_t)ntx * nty + 1, 0);
        auto tile_of = [&](int m) {
            return (int)((sb[m] >> 16) / kDegridTile) * ntx + (int)((sb[m] & 0xffff) / kDegridTile);
        };
        for (int m = 0; m < d.M; ++m) tcnt[(size_t)tile_of(m) + 1]++;
        for (size_t i = 0; i + 1 < tcnt.size(); ++i) tcnt[i + 1] += tcnt[i];
        const int64_t base = (int64_t)dsamp_h.size();
        dsamp_h.resize(dsamp_h.size() + (size_t)d.M);
        std::vector<int> pos(tcnt.begin(), tcnt.end() - 1);
        for (int m = 0; m < d.M; ++m) dsamp_h[(size_t)(base + pos[(size_t)tile_of(m)]++)] = m;
        for (int i = 0; i < d.M; ++i) {
            const int m = dsamp_h[(size_t)(base + i)];
            dbase_h.push_back(sb[m]);
            ddcf_h.push_back(sdcf_host_[(size_t)(d.samp_off + m)]);
            for (int q = 0; q < Ws; ++q) dwt_h.push_back(swt_host_[(size_t)(d.samp_off + m) * Ws + q]);
        }
        dblk_first_[(size_t)s] = (int)dblk_host_.size();
        for (int t = 0; t < ntx * nty; ++t) {
            const int corner = ((t % ntx) * kDegridTile) | (((t / ntx) * kDegridTile) << 16);
            // two sparse tiles side by side (their samples are consecutive in
            // bucket order) share one block and one staged region (option
            // "nu_degrid_merge")
            if (degrid_merge_ && (t % ntx) % 2 == 0 && (t % ntx) + 1 < ntx) {
                const int n2 = tcnt[(size_t)t + 2] - tcnt[(size_t)t];
                if (n2 > 0 && n2 <= degrid_block_) {
                    dblk_host_.push_back(make_int4(corner, (int)(base + tcnt[(size_t)t]), n2, 2 * kDegridTile));
                    ++t;
                    continue;
                }
            }
            for (int k = tcnt[(size_t)t]; k < tcnt[(size_t)t + 1]; k += degrid_block_) {
                const int nb = std::min(degrid_block_, tcnt[(size_t)t + 1] - k);
                dblk_host_.push_back(make_int4(corner, (int)(base + k), nb, kDegridTile));
            }
        }
        dblk_count_[(size_t)s] = (int)dblk_host_.size() - dblk_first_[(size_t)s];
        if (dwt_h.size() > (size_t)INT32_MAX) fail(MM_ERR_ARG, "nufftOp: too many samples");
    }
    hipStream_t st = ctx->stream;
    gstage_host_.swap(gstage_h);
    gptr_.upload(gptr_h.empty() ? std::vector<int>{0} : gptr_h, st);
    gent_.upload(gent_h.empty() ? std::vector<int2>{make_int2(0, 0)} : gent_h, st);
    dblk_.upload(dblk_host_.empty() ? std::vector<int4>{make_int4(0, 0, 0, 0)} : dblk_host_, st);
    dsamp_.upload(dsamp_h.empty() ? std::vector<int>{0} : dsamp_h, st);
    dbase_.upload(dbase_h.empty() ? std::vector<uint32_t>{0u} : dbase_h, st);
    ddcf_.upload(ddcf_h.empty() ? std::vector<float>{0.f} : ddcf_h, st);
    dwt_.upload(dwt_h.empty() ? std::vector<float>{0.f} : dwt_h, st);
    HIPCHECK(hipStreamSynchronize(st));
}

bool NufftOp::set_option(const std::string& key, int value) {
    ++gen_;   // captured solver graphs may hold the old kernels or tables
    auto flag = [&](bool& f, bool clear) {
        f = value != 0;
        if (clear) cfgs_.clear();
        return true;
    };
    if (key == "nu_degrid_t") return flag(degrid_t_, false);
    if (key == "nu_pair_cols") return flag(pair_cols_, true);
    if (key == "nu_fft_mid") return flag(fft_mid_, true);
    if (key == "nu_degrid_stage") return flag(degrid_stage_, false);
    if (key == "nu_grid_stage") return flag(grid_stage_, false);
    if (key == "nu_xcd_order") return flag(xcd_order_, true);
    if (key == "nu_grid_pairs") return flag(grid_pairs_, true);
    if (key == "nu_grid_region") return flag(grid_region_, true);
    if (key == "nu_toep_fused") return flag(toep_fused_ok_, false);
    if (key == "nu_snake") return flag(snake_, true);
    if (key == "nu_degrid_merge") {
        const bool v = value != 0;
        if (v != degrid_merge_) {
            degrid_merge_ = v;
            build_partitions();
        }
        return true;
    }
    if (key == "nu_slice_group") {
        if (value < 0) return false;
        if (slice_group_ != value) {
            slice_group_ = value;
            cfgs_.clear();
        }
        return true;
    }
    // caps keep every staged work-group within 64 KB of LDS at 16 transforms
    if (key == "nu_stage_samples" || key == "nu_stage_entries" || key == "nu_degrid_block") {
        int* dst = key == "nu_stage_samples" ? &stage_samples_
                   : key == "nu_stage_entries" ? &stage_entries_ : &degrid_block_;
        const int lo = key == "nu_stage_entries" ? 64 : 16, hi = key == "nu_stage_entries" ? 2048 : 256;
        if (value < lo || value > hi) return false;
        if (*dst != value) {
            *dst = value;
            build_partitions();
        }
        return true;
    }
    return false;
}

bool NufftOp::get_option(const std::string& key, int* value) const {
    if (!value) return false;
    if (key == "nu_degrid_t") *value = degrid_t_;
    else if (key == "nu_pair_cols") *value = pair_cols_;
    else if (key == "nu_fft_mid") *value = fft_mid_;
    else if (key == "nu_degrid_stage") *value = degrid_stage_;
    else if (key == "nu_grid_stage") *value = grid_stage_;
    else if (key == "nu_xcd_order") *value = xcd_order_;
    else if (key == "nu_grid_pairs") *value = grid_pairs_;
    else if (key == "nu_grid_region") *value = grid_region_;
    else if (key == "nu_toep_fused") *value = toep_fused_ok_;
    else if (key == "nu_snake") *value = snake_;
    else if (key == "nu_degrid_merge") *value = degrid_merge_;
    else if (key == "nu_slice_group") *value = slice_group_;
    else if (key == "nu_stage_samples") *value = stage_samples_;
    else if (key == "nu_stage_entries") *value = stage_entries_;
    else if (key == "nu_degrid_block") *value = degrid_block_;
    else return false;
    return true;
}

bool NufftOp::autotune_batch(int64_t nrep, const float2* maps, int C, const float* W, int iters) {
    if (ndim_ != 2 || nrep < 1 || toep_) return false;
    iters = std::max(iters, 2);
    const size_t nimg = (size_t)std::max<int64_t>(V_total_ * nrep, 1);
    DevBuf<float2> in(nimg), out(nimg);
    HIPCHECK(hipMemsetAsync(in.p, 0, sizeof(float2) * in.n, ctx->stream));
    // time of the named stages per application (ms); one untimed warm-up
    auto measure = [&](std::initializer_list<const char*> stages) {
        std::vector<std::pair<std::string, std::pair<double, int>>> k;
        profile_normal(in.p, out.p, nrep, maps, C, W, 1, &k);
        profile_normal(in.p, out.p, nrep, maps, C, W, iters, &k);
        double t = 0;
        for (const auto& e : k)
            for (const char* nm : stages)
                if (e.first == nm) t += e.second.first * e.second.second;
        return t;
    };
    // one option at a time: candidate values, the stages they move; the
    // first-listed candidate within 2 % of the fastest wins (timing noise
    // does not flip the choice, and the first-listed is the preferred one)
    auto tune = [&](std::vector<std::vector<std::pair<const char*, int>>> cands,
                    std::initializer_list<const char*> stages) {
        std::vector<double> t(cands.size());
        for (size_t i = 0; i < cands.size(); ++i) {
            for (const auto& kv : cands[i]) set_option(kv.first, kv.second);
            t[i] = measure(stages);
        }
        const double best = *std::min_element(t.begin(), t.end());
        size_t bi = 0;
        while (t[bi] > 1.02 * best) ++bi;
        for (const auto& kv : cands[bi]) set_option(kv.first, kv.second);
    };
    {
        // rectangle caps and unit pairing together: the best cap moves with
        // the pairing (W = 6 on a 336-point grid: 64 samples / 1024 entries
        // paired, 2.42 ms, against 2.85 ms for the best cap unpaired)
        const int caps[][2] = {{128, 1024}, {64, 512}, {96, 768}, {192, 1536}, {256, 2048},
                               {64, 1024}, {48, 1024}, {96, 2048}};
        std::vector<std::vector<std::pair<const char*, int>>> cand;
        for (int pairs = 1; pairs >= 0; --pairs)
            for (const auto& cp : caps)
                cand.push_back({{"nu_grid_pairs", pairs}, {"nu_stage_samples", cp[0]}, {"nu_stage_entries", cp[1]}});
        tune(cand, {"nu_grid"});
    }
    tune({{{"nu_degrid_merge", 1}}, {{"nu_degrid_merge", 0}}}, {"nu_degrid"});
    tune({{{"nu_degrid_block", 64}}, {{"nu_degrid_block", 32}}, {{"nu_degrid_block", 128}}}, {"nu_degrid"});
    tune({{{"nu_pair_cols", 1}}, {{"nu_pair_cols", 0}}}, {"nu_cols"});
    // the region order first: at equal time it moves the fewest HBM bytes
    tune({{{"nu_xcd_order", 1}, {"nu_grid_region", 1}},
          {{"nu_xcd_order", 1}, {"nu_grid_region", 0}},
          {{"nu_xcd_order", 0}}},
         {"nu_grid", "nu_degrid"});
    tune({{{"nu_snake", 1}}, {{"nu_snake", 0}}},
         {"nu_fwd_rows", "nu_cols", "nu_degrid", "nu_grid", "nu_adj_rows"});
    {
        // one group, or two or three groups of slices
        const int ns = (int)sl_.size();
        std::vector<std::vector<std::pair<const char*, int>>> cand{{{"nu_slice_group", 0}}};
        if (ns >= 16) {
            cand.push_back({{"nu_slice_group", (ns + 1) / 2}});
            cand.push_back({{"nu_slice_group", (ns + 2) / 3}});
        }
        tune(cand, {"nu_fwd_rows", "nu_cols", "nu_degrid", "nu_grid", "nu_adj_rows"});
    }
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    return true;
}

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
    const int ns = (int)sl_.size();
    c.hs = sl_;
    // Row / column FFT passes of one slice in mode (bigm, fm) with bct
    // transforms per chunk: the radix-16/12 factorization where it pays, the
    // single-buffer flags and the LDS bytes (d's factor and twiddle indices
    // are set accordingly).  Column passes of 16 transforms use the
    // radix-16/12 factorization where it takes fewer passes -- in the small
    // mode only where it is made of those radices alone (192 = 16 x 12: two
    // passes instead of three; 0.93 -> 0.77 us per 8-coil slice), in the
    // mid-size mode also with others (336 = 16 x 3 x 7 instead of 8 x 2 x 3 x
    // 7), where row passes of 16 transforms take it too.
    auto fft_plan = [&](NuSliceDev& d, bool bigm, int fm, int bct, size_t* lr, size_t* lc) {
        const int TR = rows_T(bct, bigm), TC = cols_T(bct, bigm);
        auto use_wide = [&](int G, bool allowed) {
            if (!allowed) return false;
            FftFactors fw, fn;
            fft_factorize(G, fw, true);
            fft_factorize(G, fn, false);
            bool w = fw.nf < fn.nf;
            if (fm == kFftSmall)
                for (int k = 0; k < fw.nf; ++k) w = w && (fw.r[k] == 16 || fw.r[k] == 12);
            return w;
        };
        const bool wide = use_wide(d.G2, !bigm && TC >= 16);
        const bool wide1 = use_wide(d.G1, rows_wide(TR, fm));
        if (wide) {
            d.fac2 = d.fac2w;
            d.tw2 = d.tw2w;
        }
        if (wide1) {
            d.fac1 = d.fac1w;
            d.tw1 = d.tw1w;
        }
        FftFactors f1, f2;
        fft_factorize(d.G1, f1, wide1);
        fft_factorize(d.G2, f2, wide);
        const bool ip1 = fft_inplace_ok(f1, TR, fft_reg_cap(TR, fft_mode_cap(fm, TR))),
                   ip2 = fft_inplace_ok(f2, TC, fft_reg_cap(TC, fft_mode_cap(fm, TC)));
        d.fip = (ip1 ? 1 : 0) | (ip2 ? 2 : 0);
        *lr = fft_lds_bytes(d.G1, TR, ip1);
        *lc = fft_lds_bytes(d.G2, TC, ip2);
    };
    // FFT mode: the small kernels where every slice's passes fit one LDS
    // buffer at their register cap, else the mid-size kernels (32 values per
    // thread) below kFftSmallMax
    auto mode_of = [&](bool bigm, int bct) {
        if (bigm) return kFftBig;
        const int TR = rows_T(bct, false), TC = cols_T(bct, false);
        bool small_ok = true;
        for (int s = 0; s < ns; ++s) {
            FftFactors f1, f2;
            fft_factorize(c.hs[s].G1, f1);
            fft_factorize(c.hs[s].G2, f2);
            small_ok = small_ok && fft_inplace_ok(f1, TR, fft_reg_cap(TR, 0)) &&
                       fft_inplace_ok(f2, TC, fft_reg_cap(TC, 0));
        }
        return small_ok || !fft_mid_ ? kFftSmall : kFftMid;
    };
    // Where those passes would need more than 160 KB of LDS (a side above
    // about 620 points at 16 transforms, 1240 at 8), the large-grid kernels
    // (one column of at most 8 transforms per work-group), and for plain
    // transforms chunks of 8; a fused SENSE chunk of 9..16 coils has no such
    // fallback (its adjoint combines all coils in one work-group).
    bool bigc = big_;
    for (;;) {
        const int fm = mode_of(bigc, c.BCT);
        size_t mr = 0, mc = 0;
        for (int s = 0; s < ns; ++s) {
            NuSliceDev d = c.hs[s];
            size_t lr = 0, lc = 0;
            fft_plan(d, bigc, fm, c.BCT, &lr, &lc);
            mr = std::max(mr, lr);
            mc = std::max(mc, lc);
        }
        if (mr <= 160 * 1024 && mc <= 160 * 1024) break;
        if (!bigc) {
            bigc = true;
            continue;
        }
        if (!sense && c.BCT > 8) {
            c.BCT = 8;
            nchunk = (int)((nrep + 7) / 8);
            continue;
        }
        fail(MM_ERR_UNSUPPORTED, sense ? "nufftOp: grid side too large for the LDS FFT passes with more than 8 "
                                         "coils fused (a SENSE operator on it needs at most 8 coils)"
                                       : "nufftOp: grid side too large for the LDS FFT passes");
    }
    const int BCT = c.BCT;
    c.fm = mode_of(bigc, BCT);
    // slice groups (2-D staged paths only)
    const int gsz = ndim_ == 2 && grid_stage_ && degrid_stage_ && slice_group_ > 0 &&
                            slice_group_ < ns ? slice_group_ : ns;
    const int ngroups = (ns + gsz - 1) / gsz;
    auto group_of = [&](int s) { return s / gsz; };
    int64_t wsoff = 0, poff = 0, ytoff = 0;
    std::vector<int4> wr, wfr, wc, wd, wcsr;
    std::vector<GridItem> wgt, wgm;
    std::vector<int4> wdb;
    std::vector<int64_t> wdb_cost;
    std::vector<int64_t> wcsr_cost, wgt_cost;
    std::vector<int> wgt_region;
    struct GridUnit {
        int s = 0, lanes = 0;
        int64_t yt = 0, zg = 0, part = 0;
    };
    std::vector<GridUnit> gunits;
    c.lds_gt = 16;
    for (int s = 0; s < ns; ++s) {
        NuSliceDev& d = c.hs[s];
        // rows / columns per work-group (the kernels' compile-time tiling):
        // rows take all BCT transforms of max(1, 8/BCT) rows, columns 8
        // transforms (8/CL columns of CL = min(BCT, 8) transforms)
        const int TR = rows_T(BCT, bigc), TC = cols_T(BCT, bigc);
        const int rpw = TR / BCT;
        const int cl = cols_CL(BCT);
        const int cpw = TC / cl;
        d.rpw = rpw;
        d.frpw = rpw;
        d.lfl = ilog2(BCT);
        d.cpw = cpw;
        d.lcl = ilog2(cl);
        {
            size_t lr = 0, lc = 0;
            fft_plan(d, bigc, c.fm, BCT, &lr, &lc);
            c.lds_rows = std::max(c.lds_rows, lr);
            c.lds_cols = std::max(c.lds_cols, lc);
        }
        if (c.lds_rows > 160 * 1024 || c.lds_cols > 160 * 1024)
            fail(MM_ERR_UNSUPPORTED, "nufftOp: grid side too large for the LDS FFT passes");
        if ((int64_t)d.N1 * d.N2 * BCT >= INT32_MAX || (int64_t)d.G1 * d.G2 * BCT >= INT32_MAX)
            fail(MM_ERR_UNSUPPORTED, "nufftOp: slice too large for 32-bit offsets");
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
        d.part_chunk = (int64_t)gpcells_[s] * BCT;
        d.part_off = poff;
        poff += d.part_chunk * nchunk;
        d.yt_off = ytoff;
        ytoff += (int64_t)d.M * BCT * nchunk;
        for (int ch = 0; ch < nchunk; ++ch) {
            for (int r = 0; r < d.N2; r += rpw) wr.push_back(make_int4(s, ch, r, 0));
            for (int r = 0; r < d.N2; r += rpw) wfr.push_back(make_int4(s, ch, r, 0));
            if (d.G2 > 1)
                for (int q = 0; q < d.G1; q += cpw)
                    for (int b0 = 0; b0 < BCT; b0 += cl) wc.push_back(make_int4(s, ch, q, b0));
            for (int m = 0; m < d.M; m += 256) wd.push_back(make_int4(s, ch, m, 0));
            if (ndim_ == 2)
                for (int k = dblk_first_[s]; k < dblk_first_[s] + dblk_count_[s]; ++k) {
                    wdb.push_back(make_int4(s, ch, k, 0));
                    wdb_cost.push_back(dblk_host_[(size_t)k].z);
                }
            {
                // blocks of NG = 256/BCT cells; a block whose longest row
                // exceeds 64 entries gets P lane groups per cell (P parts of
                // each row) and is cut into items of NG/P cells; light blocks
                // are merged into items of up to 4 blocks and
                // kCsrItemEntries entries
                const int NG = 256 / BCT;
                const int64_t ncell = (int64_t)d.G1 * d.G2;
                const int* rp = csr_ptr_host_.data() + d.csr_off;
                auto block_P = [&](int64_t c0) {
                    const int64_t c1 = std::min<int64_t>(ncell, c0 + NG);
                    int mx = 0;
                    for (int64_t q = c0; q < c1; ++q) mx = std::max(mx, rp[q + 1] - rp[q]);
                    int lp = 0;
                    while ((1 << lp) < NG && mx > 64 * (1 << lp)) ++lp;
                    return lp;
                };
                int64_t c0 = 0;
                while (c0 < ncell) {
                    const int lp = block_P(c0);
                    if (lp > 0) {
                        const int64_t c1 = std::min<int64_t>(ncell, c0 + NG);
                        const int cpi = NG >> lp;
                        for (int64_t q = c0; q < c1; q += cpi) {
                            const int64_t qe = std::min<int64_t>(c1, q + cpi);
                            wcsr.push_back(make_int4(s, ch, (int)q, (int)(qe - q) | (lp << 24)));
                            wcsr_cost.push_back((int64_t)(rp[qe] - rp[q]) * 4);
                        }
                        c0 = c1;
                        continue;
                    }
                    int64_t c1 = std::min<int64_t>(ncell, c0 + NG);
                    while (c1 < ncell && c1 - c0 < 4 * NG && block_P(c1) == 0 &&
                           rp[std::min<int64_t>(ncell, c1 + NG)] - rp[c0] <= kCsrItemEntries)
                        c1 = std::min<int64_t>(ncell, c1 + NG);
                    wcsr.push_back(make_int4(s, ch, (int)c0, (int)(c1 - c0)));
                    wcsr_cost.push_back(rp[c1] - rp[c0]);
                    c0 = c1;
                }
                // staged rectangles: the (slice, chunk) unit; items are made
                // below, after the units of one trajectory have been paired
                GridUnit u;
                u.s = s;
                u.yt = d.yt_off + (int64_t)ch * d.M * BCT;
                u.zg = d.ws_off + (int64_t)ch * d.ws_chunk;
                u.part = d.part_off + (int64_t)ch * d.part_chunk;
                u.lanes = sense ? C : (int)std::min<int64_t>(BCT, nrep - (int64_t)ch * BCT);
                gunits.push_back(u);
            }
        }
    }
    if (ndim_ == 2) {
        // Units of one trajectory are paired in slice order (option
        // "nu_grid_pairs"): an item stages its rectangle's entries once for
        // both.  Not when the staged samples of two units would need more
        // than 96 KB of LDS.
        size_t lds1 = 16, lds2 = 16;
        for (const GridTile& T : gtile_host_) {
            const size_t rest = (size_t)T.n_ent * sizeof(int2) + (size_t)(T.w * T.h + 1) * sizeof(int);
            lds1 = std::max(lds1, (size_t)T.n_stage * BCT * sizeof(float2) + rest);
            lds2 = std::max(lds2, (size_t)T.n_stage * 2 * BCT * sizeof(float2) + rest);
        }
        std::vector<std::pair<int, int>> groups;
        if (grid_pairs_ && lds2 <= 96 * 1024) {
            // (first rectangle of a trajectory, slice group) -> unpaired unit
            std::map<int64_t, int> open;
            for (int k = 0; k < (int)gunits.size(); ++k) {
                const int su = gunits[(size_t)k].s;
                const int64_t key = (int64_t)gtile_first_[(size_t)su] * 65536 + group_of(su);
                auto o = open.find(key);
                if (o == open.end()) {
                    open.emplace(key, k);
                } else {
                    groups.emplace_back(o->second, k);
                    open.erase(o);
                }
            }
            for (const auto& o : open) groups.emplace_back(o.second, -1);
            std::sort(groups.begin(), groups.end());
        } else {
            for (int k = 0; k < (int)gunits.size(); ++k) groups.emplace_back(k, -1);
        }
        c.gt_pairs = 0;
        for (const auto& g : groups) c.gt_pairs |= g.second >= 0;
        c.lds_gt = c.gt_pairs ? lds2 : lds1;
        for (const auto& g : groups) {
            const GridUnit& u = gunits[(size_t)g.first];
            const int s = u.s;
            for (int t = gtile_first_[s]; t < gtile_first_[s] + gtile_count_[s]; ++t) {
                const GridTile& T = gtile_host_[(size_t)t];
                GridItem it{};
                it.yt = u.yt;
                it.zg = u.zg;
                it.part = u.part;
                it.slice = s;
                it.G1 = c.hs[s].G1;
                it.kind = T.kind;
                it.lanes = u.lanes;
                it.slice2 = -1;
                it.g1 = T.g1;
                it.g2 = T.g2;
                it.w = T.w;
                it.h = T.h;
                it.stage_off = T.stage_off;
                it.n_stage = T.n_stage;
                it.ptr_off = T.ptr_off;
                it.ent_off = T.ent_off;
                it.n_ent = T.n_ent;
                it.pofs = T.pofs;
                it.pofs0 = T.pofs0;
                it.nparts = T.nparts;
                if (T.kind == 2 && T.pofs == T.pofs0) wgm.push_back(it);
                if (g.second >= 0) {
                    const GridUnit& v = gunits[(size_t)g.second];
                    GridItem m2 = it;   // the merge of the second unit's dense rectangle
                    m2.yt = v.yt;
                    m2.zg = v.zg;
                    m2.part = v.part;
                    m2.slice = v.s;
                    m2.lanes = v.lanes;
                    if (T.kind == 2 && T.pofs == T.pofs0) wgm.push_back(m2);
                    it.yt2 = v.yt;
                    it.zg2 = v.zg;
                    it.part2 = v.part;
                    it.slice2 = v.s;
                    it.lanes2 = v.lanes;
                }
                wgt.push_back(it);
                wgt_cost.push_back(((int64_t)T.n_ent + 2 * (int64_t)T.n_stage) * (g.second >= 0 ? 2 : 1));
                wgt_region.push_back(T.region);
            }
        }
    }
    if (xcd_order_) {
        // XCD-aware order: work-group i runs on XCD i % kXcds, so the items of
        // slice s are placed at positions = s (mod kXcds): one slice's data
        // stays in one XCD's L2.  Within an XCD, slice by slice, each
        // slice's items heaviest first.
        std::vector<std::vector<size_t>> q(kXcds);
        std::vector<size_t> ord(wcsr.size());
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) {
            if (wcsr[a].x != wcsr[b].x) return wcsr[a].x < wcsr[b].x;
            if (wcsr[a].y != wcsr[b].y) return wcsr[a].y < wcsr[b].y;
            return wcsr_cost[a] > wcsr_cost[b];
        });
        for (size_t i : ord) q[(size_t)(wcsr[i].x + (int64_t)wcsr[i].y * 3) % kXcds].push_back(i);
        std::vector<int4> sorted;
        sorted.reserve(wcsr.size());
        std::vector<size_t> pos(kXcds, 0);
        for (bool any = true; any;) {
            any = false;
            for (int x = 0; x < kXcds; ++x)
                if (pos[x] < q[x].size()) {
                    sorted.push_back(wcsr[q[x][pos[x]++]]);
                    any = true;
                }
        }
        wcsr.swap(sorted);
    } else {
        // heaviest CSR items first
        std::vector<size_t> ord(wcsr.size());
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return wcsr_cost[a] > wcsr_cost[b]; });
        std::vector<int4> sorted(wcsr.size());
        for (size_t i = 0; i < ord.size(); ++i) sorted[i] = wcsr[ord[i]];
        wcsr.swap(sorted);
    }
    // Order of the staged items: heaviest first; with the XCD-aware order
    // (option "nu_xcd_order") the items of (slice, chunk) q go to the
    // positions = q (mod kXcds), heaviest first within each XCD, so a slice's
    // data stays in one XCD's L2.
    // With the alternating slice order (option "nu_snake") the items are
    // sorted by slice first, ascending (dir > 0) or descending (dir < 0).
    auto order = [&](auto& items, const std::vector<int64_t>& cost, auto key, auto slice_of, int dir) {
        std::vector<size_t> ord(items.size());
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) {
            if (dir != 0 && slice_of(items[a]) != slice_of(items[b]))
                return dir > 0 ? slice_of(items[a]) < slice_of(items[b]) : slice_of(items[a]) > slice_of(items[b]);
            return cost[a] > cost[b];
        });
        std::vector<size_t> pick;
        if (xcd_order_) {
            std::vector<std::vector<size_t>> q(kXcds);
            for (size_t i : ord) q[(size_t)(key(i) % kXcds)].push_back(i);
            std::vector<size_t> pos(kXcds, 0);
            for (bool any = true; any;) {
                any = false;
                for (int x = 0; x < kXcds; ++x)
                    if (pos[x] < q[x].size()) {
                        pick.push_back(q[x][pos[x]++]);
                        any = true;
                    }
            }
        } else {
            pick = ord;
        }
        std::remove_reference_t<decltype(items)> sorted(items.size());
        for (size_t i = 0; i < pick.size(); ++i) sorted[i] = items[pick[i]];
        items.swap(sorted);
    };
    // Alternating slice order: each stage of the normal operator walks the
    // slices in the opposite direction to the stage before it (forward rows
    // ascending, forward columns descending, degridding ascending, gridding
    // descending, adjoint columns ascending, adjoint rows descending), so it
    // starts on the slices whose data the previous stage touched last and
    // which are still in the memory-side cache.
    const int sd = snake_ ? 1 : 0;
    auto slice4 = [](const int4& w) { return w.x; };
    order(wdb, wdb_cost, [&](size_t i) { return (int64_t)wdb[i].x * 131 + wdb[i].y; }, slice4, sd);
    // Staged gridding: with option "nu_grid_region" the items of one region
    // of a trajectory's rectangles go to one XCD for every slice, so the
    // region's shared tables (sample ids, row pointers, entries) and the
    // samples its neighbouring rectangles both stage stay in that XCD's L2;
    // otherwise one (slice, chunk)'s items go to one XCD.
    order(wgt, wgt_cost,
          [&](size_t i) {
              const GridItem& g = wgt[i];
              if (grid_region_) return (int64_t)wgt_region[i];
              return (int64_t)g.slice * 131 + (g.zg - c.hs[g.slice].ws_off) / std::max<int64_t>(1, c.hs[g.slice].ws_chunk);
          },
          [](const GridItem& g) { return g.slice; }, -sd);
    auto by_slice_desc = [&](auto& items, auto slice_of) {
        std::stable_sort(items.begin(), items.end(),
                         [&](const auto& a, const auto& b) { return slice_of(a) > slice_of(b); });
    };
    std::vector<int4> wcf = wc;   // forward column items
    if (snake_) {
        by_slice_desc(wr, slice4);
        by_slice_desc(wcf, slice4);
        by_slice_desc(wgm, [](const GridItem& g) { return g.slice; });
    }
    if (pair_cols_) {
        // neighbouring columns share cache lines (a column touches 64 of each
        // 128 bytes): place items 2k and 2k + 1 of every 16 at positions k and
        // k + 8, which the dispatcher sends to the same XCD at about the same
        // time, so each line is fetched into one L2
        auto pair = [](std::vector<int4>& v) {
            std::vector<int4> r(v.size());
            for (size_t j = 0; j < v.size(); ++j) {
                const size_t q = j / 16, rr = j % 16;
                size_t pos = q * 16 + rr / 2 + 8 * (rr % 2);
                if (pos >= v.size()) pos = j;   // incomplete last group keeps its order
                r[pos] = v[j];
            }
            if (v.size() % 16) {
                const size_t q0 = v.size() / 16 * 16;
                for (size_t j = q0; j < v.size(); ++j) r[j] = v[j];
            }
            v.swap(r);
        };
        pair(wc);
        pair(wcf);
    }
    // slice groups: each list ordered group by group (the order within a
    // group kept, so this is the last reordering), with the span of every
    // group
    c.goff.assign((size_t)ngroups, {});
    c.gcnt.assign((size_t)ngroups, {});
    auto split = [&](auto& items, auto slice_of, int L) {
        std::stable_sort(items.begin(), items.end(), [&](const auto& a, const auto& b) {
            return group_of(slice_of(a)) < group_of(slice_of(b));
        });
        int off = 0;
        for (int g = 0; g < ngroups; ++g) {
            int n = 0;
            while (off + n < (int)items.size() && group_of(slice_of(items[(size_t)(off + n)])) == g) ++n;
            c.goff[(size_t)g][(size_t)L] = off;
            c.gcnt[(size_t)g][(size_t)L] = n;
            off += n;
        }
    };
    {
        auto s4 = [](const int4& w) { return w.x; };
        auto sg = [](const GridItem& g) { return g.slice; };
        split(wr, s4, kLRows);
        split(wfr, s4, kLFrows);
        split(wc, s4, kLCols);
        split(wcf, s4, kLFcols);
        split(wdb, s4, kLDb);
        split(wgt, sg, kLGt);
        split(wgm, sg, kLGm);
    }
    hipStream_t st = ctx->stream;
    c.sl.upload(c.hs, st);
    c.w_rows.upload(wr, st);
    c.w_frows.upload(wfr, st);
    c.n_frows = (int)wfr.size();
    c.w_cols.upload(wc, st);
    c.w_fcols.upload(wcf, st);
    c.w_deg.upload(wd, st);
    c.w_csr.upload(wcsr.empty() ? std::vector<int4>{make_int4(0, 0, 0, 0)} : wcsr, st);
    c.n_csr = (int)wcsr.size();
    {
        int mx = 1;
        for (const GridItem& g : wgt) mx = std::max(mx, g.n_stage);
        c.gt_sid_stride = (mx + 15) & ~15;
        std::vector<int> rows(std::max<size_t>(1, wgt.size() * (size_t)c.gt_sid_stride), 0);
        for (size_t i = 0; i < wgt.size(); ++i)
            std::copy(gstage_host_.begin() + wgt[i].stage_off,
                      gstage_host_.begin() + wgt[i].stage_off + wgt[i].n_stage,
                      rows.begin() + (int64_t)i * c.gt_sid_stride);
        c.gt_sid.upload(rows, st);
    }
    c.w_gt.upload(wgt.empty() ? std::vector<GridItem>(1) : wgt, st);
    c.n_gt = (int)wgt.size();
    c.w_gm.upload(wgm.empty() ? std::vector<GridItem>(1) : wgm, st);
    c.n_gm = (int)wgm.size();
    c.lds_gt = (c.lds_gt + 15) & ~(size_t)15;
    c.n_rows = (int)wr.size();
    c.n_cols = (int)wc.size();
    c.n_deg = (int)wd.size();
    c.w_db.upload(wdb.empty() ? std::vector<int4>{make_int4(0, 0, 0, 0)} : wdb, st);
    c.n_db = (int)wdb.size();
    c.lds_db = (size_t)((degrid_merge_ ? 2 : 1) * kDegridTile + W_ - 1) * (kDegridTile + W_ - 1) * BCT * sizeof(float2) +
               (size_t)degrid_block_ * (3 + 2 * W_) * sizeof(float);
    c.yT.alloc(std::max<int64_t>(ytoff, 1));
    c.ws.alloc(std::max<int64_t>(wsoff, 1));
    c.parts.alloc(std::max<int64_t>(poff, 1));
    HIPCHECK(hipStreamSynchronize(st));
    cfgs_[key] = std::move(cp);
    return c;
}

template <int BCT, int FM>
void NufftOp::launch(Cfg& c, int adjoint, const NuRun& run, const float2* in, float2* out, int mode) {
    hipStream_t st = ctx->stream;
    static bool attr_set = false;
    if (!attr_set) {
        // allow FFT passes above 64 KB of LDS where the device permits it
        (void)hipFuncSetAttribute((const void*)nu_fwd_rows<BCT, FM>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void*)nu_adj_rows<BCT, FM>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void*)nu_cols<BCT, FM, 0>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void*)nu_cols<BCT, FM, 1>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void*)nu_degrid_tile<BCT, 4>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void*)nu_degrid_tile<BCT, 6>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void*)nu_degrid_tile<BCT, 0>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void*)nu_grid_tile<BCT, 1>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void*)nu_grid_tile<BCT, 2>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipGetLastError();
        attr_set = true;
    }
    const bool wt4 = ndim_ == 2 && W_ == 4;
    // the staged degridding also has W = 6 at compile time (the interpolated
    // high-order operator's NUFFT)
    const bool wt6 = ndim_ == 2 && W_ == 6;
    const NuSliceDev* sl = c.sl.p;
    // (first item, items) of list L in slice group g; g < 0: the whole list
    auto span = [&](int L, int g, int n_all) -> std::pair<int, int> {
        if (g < 0) return {0, n_all};
        return {c.goff[(size_t)g][(size_t)L], c.gcnt[(size_t)g][(size_t)L]};
    };
    // the odd-size ramp over the grids of every slice and chunk (slices
    // with even sides return at once)
    auto odd_ramp = [&](bool conj, const int* skip) {
        if (!any_odd_) return;
        int64_t mx = 1;
        for (const NuSliceDev& d : c.hs) mx = std::max(mx, d.ws_chunk);
        const int nchunk = (int)(c.sense ? c.nrep : (c.nrep + BCT - 1) / BCT);
        const dim3 grid((unsigned)((mx + 255) / 256), (unsigned)c.hs.size(), (unsigned)nchunk);
        if (conj)
            hipLaunchKernelGGL((nu_odd_ramp<BCT, 1>), grid, dim3(256), 0, st, sl, c.ws.p, skip);
        else
            hipLaunchKernelGGL((nu_odd_ramp<BCT, 0>), grid, dim3(256), 0, st, sl, c.ws.p, skip);
        LAUNCHCHECK();
        mark("nu_odd_ramp");
    };
    auto forward = [&](const float2* x, float2* y, NuRun r, int g) {
        const auto fr = span(kLFrows, g, c.n_frows), fc = span(kLFcols, g, c.n_cols), db = span(kLDb, g, c.n_db);
        if (fr.second > 0)
            hipLaunchKernelGGL((nu_fwd_rows<BCT, FM>), dim3(fr.second), dim3(256), c.lds_rows, st, sl,
                               c.w_frows.p + fr.first, facs_.p, twid_.p, deap_.p, x, c.ws.p, r);
        LAUNCHCHECK();
        mark("nu_fwd_rows");
        if (ndim_ == 2 && fc.second > 0) {
            hipLaunchKernelGGL((nu_cols<BCT, FM, 0>), dim3(fc.second), dim3(256), c.lds_cols, st, sl,
                               c.w_fcols.p + fc.first, facs_.p, twid_.p, c.ws.p, r.skip);
            LAUNCHCHECK();
            mark("nu_cols");
        }
        odd_ramp(false, r.skip);
        if (ndim_ == 2 && degrid_stage_ && db.second > 0) {
            if (wt4)
                hipLaunchKernelGGL((nu_degrid_tile<BCT, 4>), dim3(db.second), dim3(256), c.lds_db, st, sl, c.w_db.p + db.first,
                                   dblk_.p, dsamp_.p, dbase_.p, dwt_.p, ddcf_.p, (const float2*)c.ws.p, y,
                                   degrid_block_, r);
            else if (wt6)
                hipLaunchKernelGGL((nu_degrid_tile<BCT, 6>), dim3(db.second), dim3(256), c.lds_db, st, sl, c.w_db.p + db.first,
                                   dblk_.p, dsamp_.p, dbase_.p, dwt_.p, ddcf_.p, (const float2*)c.ws.p, y,
                                   degrid_block_, r);
            else
                hipLaunchKernelGGL((nu_degrid_tile<BCT, 0>), dim3(db.second), dim3(256), c.lds_db, st, sl, c.w_db.p + db.first,
                                   dblk_.p, dsamp_.p, dbase_.p, dwt_.p, ddcf_.p, (const float2*)c.ws.p, y,
                                   degrid_block_, r);
        } else if (g >= 0) {
            // a slice group without degridding blocks: nothing to degrid
        } else if (BCT >= 2 && degrid_t_) {
            if (wt4)
                hipLaunchKernelGGL((nu_degrid_t<BCT, 4>), dim3(c.n_deg), dim3(256), 0, st, sl, c.w_deg.p,
                                   sbase_.p, swt_.p, sdcf_.p, (const float2*)c.ws.p, y, r);
            else
                hipLaunchKernelGGL((nu_degrid_t<BCT, 0>), dim3(c.n_deg), dim3(256), 0, st, sl, c.w_deg.p,
                                   sbase_.p, swt_.p, sdcf_.p, (const float2*)c.ws.p, y, r);
        } else if (wt4)
            hipLaunchKernelGGL((nu_degrid<BCT, 4>), dim3(c.n_deg), dim3(256), 0, st, sl, c.w_deg.p, sbase_.p,
                               swt_.p, sdcf_.p, (const float2*)c.ws.p, y, r);
        else
            hipLaunchKernelGGL((nu_degrid<BCT, 0>), dim3(c.n_deg), dim3(256), 0, st, sl, c.w_deg.p, sbase_.p,
                               swt_.p, sdcf_.p, (const float2*)c.ws.p, y, r);
        LAUNCHCHECK();
        mark("nu_degrid");
    };
    auto adjoint_from = [&](const float2* y, float2* x, NuRun r, int g) {
        const auto gt = span(kLGt, g, c.n_gt), gm = span(kLGm, g, c.n_gm), cc = span(kLCols, g, c.n_cols),
                   rw = span(kLRows, g, c.n_rows);
        if (grid_stage_ && ndim_ == 2) {
            // the item's sample-id row is addressed by its block index:
            // a group's rows start at its first item's
            const int* sid = c.gt_sid.p + (int64_t)gt.first * c.gt_sid_stride;
            if (gt.second > 0 && c.gt_pairs)
                hipLaunchKernelGGL((nu_grid_tile<BCT, 2>), dim3(gt.second), dim3(256), c.lds_gt, st,
                                   c.w_gt.p + gt.first, sid, c.gt_sid_stride, gptr_.p, gent_.p, y, c.ws.p, c.parts.p,
                                   r.skip);
            else if (gt.second > 0)
                hipLaunchKernelGGL((nu_grid_tile<BCT, 1>), dim3(gt.second), dim3(256), c.lds_gt, st,
                                   c.w_gt.p + gt.first, sid, c.gt_sid_stride, gptr_.p, gent_.p, y, c.ws.p, c.parts.p,
                                   r.skip);
            if (gm.second > 0) {
                LAUNCHCHECK();
                hipLaunchKernelGGL(nu_grid_merge<BCT>, dim3(gm.second), dim3(256), 0, st, c.w_gm.p + gm.first,
                                   c.parts.p, c.ws.p, r.skip);
            }
        } else {
            hipLaunchKernelGGL(nu_grid_csr<BCT>, dim3(c.n_csr), dim3(256), 0, st, sl, c.w_csr.p, csr_ptr_.p,
                               csr_ent_.p, y, c.ws.p, r);
        }
        LAUNCHCHECK();
        mark("nu_grid");
        odd_ramp(true, r.skip);
        if (ndim_ == 2 && cc.second > 0) {
            hipLaunchKernelGGL((nu_cols<BCT, FM, 1>), dim3(cc.second), dim3(256), c.lds_cols, st, sl,
                               c.w_cols.p + cc.first, facs_.p, twid_.p, c.ws.p, r.skip);
            LAUNCHCHECK();
            mark("nu_cols");
        }
        if (rw.second > 0)
            hipLaunchKernelGGL((nu_adj_rows<BCT, FM>), dim3(rw.second), dim3(256), c.lds_rows, st, sl,
                               c.w_rows.p + rw.first, facs_.p, twid_.p, deap_.p, (const float2*)c.ws.p, x, r);
        LAUNCHCHECK();
        mark("nu_adj_rows");
    };
    if (mode == 2) {
        // Toeplitz normal operator: forward row and column FFTs, the kernel
        // spectrum multiply, adjoint column and row FFTs
        if (c.n_frows > 0)
            hipLaunchKernelGGL((nu_fwd_rows<BCT, FM>), dim3(c.n_frows), dim3(256), c.lds_rows, st, sl, c.w_frows.p,
                               facs_.p, twid_.p, deap_.p, in, c.ws.p, run);
        LAUNCHCHECK();
        mark("nu_fwd_rows");
        if (ndim_ == 2 && toep_fused_ok_) {
            if (c.n_cols > 0)
                hipLaunchKernelGGL((nu_cols_toep<BCT, FM>), dim3(c.n_cols), dim3(256), c.lds_cols, st, sl, c.w_cols.p,
                                   facs_.p, twid_.p, (const float2*)toep_m_.p, (const int64_t*)toep_off_.p, c.ws.p,
                                   run.skip);
            LAUNCHCHECK();
            mark("nu_cols_toep");
        } else {
        if (ndim_ == 2 && c.n_cols > 0) {
            hipLaunchKernelGGL((nu_cols<BCT, FM, 0>), dim3(c.n_cols), dim3(256), c.lds_cols, st, sl, c.w_fcols.p,
                               facs_.p, twid_.p, c.ws.p, run.skip);
            LAUNCHCHECK();
            mark("nu_cols");
        }
        int64_t mx = 1;
        for (const NuSliceDev& d : c.hs) mx = std::max(mx, d.ws_chunk);
        const int nchunk = (int)(c.sense ? c.nrep : (c.nrep + BCT - 1) / BCT);
        hipLaunchKernelGGL(nu_toep_mul<BCT>, dim3((unsigned)((mx + 1023) / 1024), (unsigned)c.hs.size(), (unsigned)nchunk),
                           dim3(256), 0, st, sl, (const float2*)toep_m_.p, (const int64_t*)toep_off_.p, c.ws.p,
                           run.skip);
        LAUNCHCHECK();
        mark("nu_toep_mul");
        if (ndim_ == 2 && c.n_cols > 0) {
            hipLaunchKernelGGL((nu_cols<BCT, FM, 1>), dim3(c.n_cols), dim3(256), c.lds_cols, st, sl, c.w_cols.p,
                               facs_.p, twid_.p, c.ws.p, run.skip);
            LAUNCHCHECK();
            mark("nu_cols");
        }
        }
        if (c.n_rows > 0)
            hipLaunchKernelGGL((nu_adj_rows<BCT, FM>), dim3(c.n_rows), dim3(256), c.lds_rows, st, sl, c.w_rows.p,
                               facs_.p, twid_.p, deap_.p, (const float2*)c.ws.p, out, run);
        LAUNCHCHECK();
        mark("nu_adj_rows");
        return;
    }
    if (mode == 1) {
        NuRun rf = run;
        rf.ylayout = 1;
        float2* ybuf = c.yT.p;
        NuRun ra = run;
        ra.W = nullptr;
        ra.ylayout = rf.ylayout;
        // slice groups one after another (one group: the whole lists)
        const int ng = (int)c.goff.size();
        for (int g = 0; g < ng; ++g) {
            forward(in, ybuf, rf, ng > 1 ? g : -1);
            adjoint_from(ybuf, out, ra, ng > 1 ? g : -1);
        }
        return;
    }
    if (!adjoint) {
        NuRun rf = run;
        rf.ylayout = 0;
        forward(in, out, rf, -1);
        return;
    }
    NuRun ra = run;
    ra.W = nullptr;
    ra.ylayout = 0;
    hipLaunchKernelGGL(nu_transpose_y<BCT>, dim3(c.n_deg), dim3(256), 0, st, sl, c.w_deg.p, in, c.yT.p, ra);
    LAUNCHCHECK();
    mark("nu_transpose_y");
    ra.ylayout = 1;
    adjoint_from(c.yT.p, out, ra, -1);
}

void NufftOp::run(Cfg& c, int adjoint, const float2* in, float2* out, const float2* maps, const float* W,
                  int mode) {
    NuRun r;
    r.ylayout = 0;
    r.sense = c.sense;
    r.C = c.C;
    r.B = c.nrep;
    r.maps = maps;
    r.W = W;
    if (toep_) {
        // both directions, and the normal operator, are the Toeplitz form
        if (W) fail(MM_ERR_UNSUPPORTED, "nufftOp: data weights with the Toeplitz form");
        mode = mode == 1 ? 3 : 2;
    }
    r.skip = mode == 1 || mode == 3 ? skip_ : nullptr;
    if (mode == 3) mode = 2;
    auto go = [&](auto fm) {
        constexpr int FM = decltype(fm)::value;
        switch (c.BCT) {
            case 1: launch<1, FM>(c, adjoint, r, in, out, mode); break;
            case 2: launch<2, FM>(c, adjoint, r, in, out, mode); break;
            case 4: launch<4, FM>(c, adjoint, r, in, out, mode); break;
            case 8: launch<8, FM>(c, adjoint, r, in, out, mode); break;
            default: launch<16, FM>(c, adjoint, r, in, out, mode); break;
        }
    };
    if (c.fm == kFftBig)
        go(std::integral_constant<int, kFftBig>());
    else if (c.fm == kFftMid)
        go(std::integral_constant<int, kFftMid>());
    else
        go(std::integral_constant<int, kFftSmall>());
}

void NufftOp::set_toeplitz(const float2* psf) {
    if (!psf) fail(MM_ERR_ARG, "nufftOp: NULL point-spread image");
    const int ns = (int)sl_.size();
    hipStream_t st = ctx->stream;
    std::vector<float> deap(deap_.n);
    if (deap_.n) HIPCHECK(hipMemcpy(deap.data(), deap_.p, deap_.n * sizeof(float), hipMemcpyDeviceToHost));
    std::vector<float2> km;
    std::vector<int64_t> koff;
    int64_t pofs = 0;
    for (int s = 0; s < ns; ++s) {
        NuSliceDev& d = sl_[(size_t)s];
        const int N1 = d.N1, N2 = ndim_ == 2 ? d.N2 : 1, G1 = d.G1, G2 = ndim_ == 2 ? d.G2 : 1;
        if (G1 < 2 * N1 - 1 || G2 < 2 * N2 - 1) fail(MM_ERR_ARG, "nufftOp: Toeplitz form needs a grid of at least 2N - 1");
        const int P1 = 2 * N1, P2 = ndim_ == 2 ? 2 * N2 : 1;   // point-spread image size
        {
            // slices with the same grid and a byte-identical point-spread
            // image (one trajectory and dcf) share one spectrum
            int same = -1;
            int64_t q = 0;
            for (int t = 0; t < s && same < 0; ++t) {
                const NuSliceDev& e = sl_[(size_t)t];
                const int64_t pt = (int64_t)2 * e.N1 * (ndim_ == 2 ? 2 * e.N2 : 1);
                if (e.N1 == d.N1 && e.N2 == d.N2 && e.G1 == d.G1 && e.G2 == d.G2 &&
                    std::memcmp(psf + q, psf + pofs, (size_t)pt * sizeof(float2)) == 0)
                    same = t;
                q += pt;
            }
            if (same >= 0) {
                koff.push_back(koff[(size_t)same]);
                pofs += (int64_t)P1 * P2;
                d.deap1 = sl_[(size_t)same].deap1;
                d.deap2 = sl_[(size_t)same].deap2;
                continue;
            }
        }
        const double alpha = std::pow(2.0, ndim_) / std::sqrt((double)P1 * P2);
        // h(D1, D2) = alpha psf[N1 + D1, N2 + D2] for |D| <= N - 1 (the only
        // lags between a pixel and a pixel of the cropped output);
        // K[g2][g1] = sum_D h(D) exp(+2 pi i (g1 D1 / G1 + g2 D2 / G2)), the
        // forward passes' sign, separably in FP64
        const int L1 = 2 * N1 - 1, L2 = 2 * N2 - 1;
        auto ex = [](int G) {
            std::vector<std::pair<double, double>> e((size_t)G);
            for (int k = 0; k < G; ++k) e[(size_t)k] = {std::cos(2 * M_PI * k / G), std::sin(2 * M_PI * k / G)};
            return e;
        };
        const auto e1 = ex(G1), e2 = ex(G2);
        std::vector<double> ar((size_t)G1 * L2), ai((size_t)G1 * L2);   // [g1][D2]
        for (int j2 = 0; j2 < L2; ++j2) {
            const int col = P2 > 1 ? N2 + j2 - (N2 - 1) : 0;   // psf column of D2 = j2 - (N2 - 1)
            for (int g1 = 0; g1 < G1; ++g1) {
                double sr = 0, si = 0;
                for (int j1 = 0; j1 < L1; ++j1) {
                    const int D1 = j1 - (N1 - 1);
                    const float2 h = psf[pofs + (int64_t)(N1 + D1) + (int64_t)P1 * col];
                    const auto& w = e1[(size_t)((((int64_t)g1 * D1) % G1 + G1) % G1)];
                    sr += h.x * w.first - h.y * w.second;
                    si += h.x * w.second + h.y * w.first;
                }
                ar[(size_t)g1 * L2 + j2] = sr;
                ai[(size_t)g1 * L2 + j2] = si;
            }
        }
        koff.push_back((int64_t)km.size());
        const size_t base = km.size();
        km.resize(base + (size_t)G1 * G2);
        for (int g2 = 0; g2 < G2; ++g2)
            for (int g1 = 0; g1 < G1; ++g1) {
                double sr = 0, si = 0;
                for (int j2 = 0; j2 < L2; ++j2) {
                    const int D2 = j2 - (N2 - 1);
                    const auto& w = e2[(size_t)((((int64_t)g2 * D2) % G2 + G2) % G2)];
                    const double xr = ar[(size_t)g1 * L2 + j2], xi = ai[(size_t)g1 * L2 + j2];
                    sr += xr * w.first - xi * w.second;
                    si += xr * w.second + xi * w.first;
                }
                km[base + (size_t)g2 * G1 + g1] = make_float2((float)(alpha * sr), (float)(alpha * si));
            }
        pofs += (int64_t)P1 * P2;
        // flat scaling 1/sqrt(G1 G2) on both sides (unitary FFT pair)
        d.deap1 = (int)deap.size();
        for (int i = 0; i < d.N1; ++i) deap.push_back((float)(1.0 / std::sqrt((double)G1 * G2)));
        d.deap2 = (int)deap.size();
        for (int i = 0; i < std::max(1, ndim_ == 2 ? d.N2 : 1); ++i) deap.push_back(1.f);
    }
    deap_.upload(deap, st);
    toep_m_.upload(km, st);
    toep_off_.upload(koff, st);
    HIPCHECK(hipStreamSynchronize(st));
    toep_ = true;
    cfgs_.clear();
    ++gen_;
}

void NufftOp::apply(int adjoint, const float2* in, float2* out, int64_t nrep) {
    if (nrep < 1) fail(MM_ERR_ARG, "nufftOp: nrep must be >= 1");
    run(config(nrep, 0, 0), adjoint, in, out, nullptr, nullptr, 0);
}

void NufftOp::apply_sense(int adjoint, const float2* in, float2* out, int64_t nrep,
                          const float2* maps, int C, const float* Wt) {
    if (nrep < 1) fail(MM_ERR_ARG, "nufftOp: nrep must be >= 1");
    if (toep_) fail(MM_ERR_HIP, "nufftOp: internal: fused SENSE apply on the Toeplitz form");
    run(config(nrep, 1, C), adjoint, in, out, maps, adjoint ? nullptr : Wt);
}

bool NufftOp::sense_fits(int C) {
    try {
        (void)config(1, 1, C);
        return true;
    } catch (const Error& e) {
        if (e.code == MM_ERR_UNSUPPORTED) return false;
        throw;
    }
}

void NufftOp::apply_sense_normal(const float2* in, float2* out, int64_t nrep, const float2* maps, int C,
                                 const float* Wt) {
    if (nrep < 1) fail(MM_ERR_ARG, "nufftOp: nrep must be >= 1");
    run(config(nrep, 1, C), 0, in, out, maps, Wt, 1);
}

void NufftOp::mark(const char* name) {
    if (!prof_) return;
    Prof& p = *prof_;
    if (p.k + 1 >= (int)p.ev.size()) return;
    HIPCHECK(hipEventRecord(p.ev[p.k + 1], ctx->stream));
    p.names[p.k] = name;
    ++p.k;
}

void NufftOp::profile_normal(const float2* in, float2* out, int64_t nrep, const float2* maps, int C,
                             const float* Wt, int iters,
                             std::vector<std::pair<std::string, std::pair<double, int>>>* per_kernel) {
    profile_run(0, 1, in, out, nrep, maps, C, Wt, iters, per_kernel);
}

bool NufftOp::profile_dominant(int adjoint, const float2* in, float2* out, int64_t nrep, int iters, double* ms,
                               std::string* name) {
    if (adjoint != 0 && adjoint != 1) return false;
    std::vector<std::pair<std::string, std::pair<double, int>>> k;
    profile_run(adjoint, 0, in, out, nrep, nullptr, 0, nullptr, iters, &k);
    double best = -1;
    for (auto& e : k) {
        const double tot = e.second.first * e.second.second;
        if (tot > best) {
            best = tot;
            *ms = e.second.first;
            *name = e.first;
        }
    }
    return best >= 0;
}

void NufftOp::profile_run(int adjoint, int mode, const float2* in, float2* out, int64_t nrep, const float2* maps,
                          int C, const float* Wt, int iters,
                          std::vector<std::pair<std::string, std::pair<double, int>>>* per_kernel) {
    if (nrep < 1 || iters < 1) fail(MM_ERR_ARG, "profile: bad argument");
    Cfg& c = config(nrep, maps ? 1 : 0, maps ? C : 0);
    // one untimed application (plan set-up, first-launch costs)
    run(c, adjoint, in, out, maps, Wt, mode);
    Prof p;
    // marks per application: up to 7 stages per slice group
    const size_t nev = 8 * std::max<size_t>(1, c.goff.size()) + 8;
    p.ev.resize(nev);
    p.names.resize(nev);
    for (auto& e : p.ev) HIPCHECK(hipEventCreate(&e));
    std::map<std::string, std::pair<double, int>> acc;
    int launches_per_app = 0;
    try {
        for (int it = 0; it < iters; ++it) {
            p.k = 0;
            HIPCHECK(hipEventRecord(p.ev[0], ctx->stream));
            prof_ = &p;
            run(c, adjoint, in, out, maps, Wt, mode);
            prof_ = nullptr;
            HIPCHECK(hipEventSynchronize(p.ev[p.k]));
            launches_per_app = p.k;
            for (int j = 0; j < p.k; ++j) {
                float ms = 0.f;
                HIPCHECK(hipEventElapsedTime(&ms, p.ev[j], p.ev[j + 1]));
                auto& a = acc[p.names[j]];
                a.first += ms;
                a.second += 1;
            }
        }
    } catch (...) {
        prof_ = nullptr;
        for (auto& e : p.ev) (void)hipEventDestroy(e);
        throw;
    }
    for (auto& e : p.ev) (void)hipEventDestroy(e);
    (void)launches_per_app;
    per_kernel->clear();
    for (auto& kv : acc)
        per_kernel->push_back({kv.first, {kv.second.first / kv.second.second, kv.second.second / iters}});
}

}  // namespace mm

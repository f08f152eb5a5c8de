// Mixed-radix (16,12,8,4,2,3,5,7) Stockham FFT on data resident in LDS.
//
// Layout: element (position q, transform t) of `T` interleaved transforms of
// length n lives at buf[q*T + t], so consecutive lanes work on consecutive
// transforms (conflict-free LDS access).  A pass of radix R with P = the
// product of the radices already done reads R inputs spaced n/R apart,
// multiplies input j by exp(sgn*2*pi*i*j*k/(P*R)) (k = butterfly index mod
// P), does a length-R DFT and writes the R outputs spaced P apart, which
// leaves the result in natural order after the last pass.  sgn = +1 is
// MATLAB's ifft kernel (unscaled), sgn = -1 its fft kernel.  T and sgn are
// compile-time constants of the kernels, so every index is a shift, a mask or
// a multiply-add.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace mm {

constexpr int kMaxFactors = 12;

// threads of every work-group that runs these transforms (a compile-time
// constant: blockDim.x would be a memory load)
constexpr int kFftThreads = 256;

struct FftFactors {
    int n = 1;
    int nf = 0;
    int r[kMaxFactors] = {};
};

// Radix order: 8s first, then 4, 2, 3, 5, 7; `wide` (passes of 16
// interleaved transforms, whose registers hold 16 values per thread):
// 16s and 12s before those (192 = 16 x 12: two passes instead of three).
// Returns false when n has a prime factor above 7 (the grid-size rule of
// nufftOp.m:146-152 never produces one).
inline bool fft_factorize(int n, FftFactors& f, bool wide = false) {
    f = FftFactors();
    f.n = n;
    int m = n;
    if (wide) {
        while (m % 16 == 0 && f.nf < kMaxFactors) { f.r[f.nf++] = 16; m /= 16; }
        while (m % 12 == 0 && f.nf < kMaxFactors) { f.r[f.nf++] = 12; m /= 12; }
    }
    while (m % 8 == 0 && f.nf < kMaxFactors) { f.r[f.nf++] = 8; m /= 8; }
    while (m % 4 == 0 && f.nf < kMaxFactors) { f.r[f.nf++] = 4; m /= 4; }
    const int primes[4] = {2, 3, 5, 7};
    for (int pi = 0; pi < 4; ++pi) {
        while (m % primes[pi] == 0) {
            if (f.nf >= kMaxFactors) return false;
            f.r[f.nf++] = primes[pi];
            m /= primes[pi];
        }
    }
    return m == 1;
}

// Pass twiddles of one transform length, n - 1 values: pass s (P = product
// of the earlier radices, radix R) holds exp(+2*pi*i*j*k/(P*R)) for k < P,
// j = 1..R-1 at offset (P - 1) + k*(R - 1) + (j - 1).  Pass 0 (P = 1) needs
// none.  Host side, FP64 angles rounded once.
template <class Push>
inline void fft_pass_twiddles(const FftFactors& f, Push push) {
    int p = 1;
    for (int s = 0; s < f.nf; ++s) {
        const int R = f.r[s];
        for (int k = 0; k < p; ++k)
            for (int j = 1; j < R; ++j) {
                const double a = 2.0 * 3.14159265358979323846 * (double)j * (double)k / ((double)p * R);
                push((float)std::cos(a), (float)std::sin(a));
            }
        p *= R;
    }
}

__host__ __device__ inline float2 cmulf(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <int R>
__host__ __device__ inline void dft_small(float2 (&v)[R], float sgn);

// Length A*B DFT as B length-A DFTs of the stride-B subsequences, the
// factors exp(sgn*2*pi*i*j2*q1/(A*B)) and A length-B DFTs:
// X_{q1 + A q2} = sum_j2 w^(j2 q1) [sum_j1 v_{B j1 + j2} w_A^(j1 q1)] w_B^(j2 q2).
// cs/sn: cos/sin(2*pi*m/(A*B)) for m <= (A-1)(B-1).
template <int A, int B>
__host__ __device__ __forceinline__ void dft_split(float2 (&v)[A * B], float sgn, const float* cs, const float* sn) {
    float2 y[B][A];
#pragma unroll
    for (int j2 = 0; j2 < B; ++j2) {
#pragma unroll
        for (int j1 = 0; j1 < A; ++j1) y[j2][j1] = v[B * j1 + j2];
        dft_small<A>(y[j2], sgn);
    }
#pragma unroll
    for (int j2 = 1; j2 < B; ++j2)
#pragma unroll
        for (int q1 = 1; q1 < A; ++q1) {
            const int m = j2 * q1;
            const float c = cs[m], sw = sgn * sn[m];
            const float2 t = y[j2][q1];
            y[j2][q1] = make_float2(t.x * c - t.y * sw, t.x * sw + t.y * c);
        }
#pragma unroll
    for (int q1 = 0; q1 < A; ++q1) {
        float2 z[B];
#pragma unroll
        for (int j2 = 0; j2 < B; ++j2) z[j2] = y[j2][q1];
        dft_small<B>(z, sgn);
#pragma unroll
        for (int q2 = 0; q2 < B; ++q2) v[q1 + A * q2] = z[q2];
    }
}

// In-register length-R DFT: X_q = sum_j v_j exp(sgn*2*pi*i*j*q/R).
template <int R>
__host__ __device__ inline void dft_small(float2 (&v)[R], float sgn) {
    if constexpr (R == 16) {
        // 4 x 4; cos/sin(2*pi*m/16), m = 0..9
        const float c1 = 0.92387953251128676f, s1 = 0.38268343236508977f, h = 0.70710678118654752f;
        const float cs[10] = {1.f, c1, h, s1, 0.f, -s1, -h, -c1, -1.f, -c1};
        const float sn[10] = {0.f, s1, h, c1, 1.f, c1, h, s1, 0.f, -s1};
        dft_split<4, 4>(v, sgn, cs, sn);
    } else if constexpr (R == 12) {
        // 4 x 3; cos/sin(2*pi*m/12), m = 0..6
        const float r = 0.86602540378443865f;
        const float cs[7] = {1.f, r, 0.5f, 0.f, -0.5f, -r, -1.f};
        const float sn[7] = {0.f, 0.5f, r, 1.f, r, 0.5f, 0.f};
        dft_split<4, 3>(v, sgn, cs, sn);
    } else if constexpr (R == 8) {
        // two length-4 DFTs (even / odd inputs) and the w8^q factors,
        // w8 = exp(sgn*2*pi*i/8) = (1 + sgn*i)/sqrt(2)
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        dft_small<4>(e, sgn);
        dft_small<4>(o, sgn);
        const float h = 0.70710678118654752f;
        const float2 o1 = make_float2(h * (o[1].x - sgn * o[1].y), h * (o[1].y + sgn * o[1].x));
        const float2 o2 = make_float2(-sgn * o[2].y, sgn * o[2].x);
        const float2 o3 = make_float2(h * (-o[3].x - sgn * o[3].y), h * (-o[3].y + sgn * o[3].x));
        v[0] = make_float2(e[0].x + o[0].x, e[0].y + o[0].y);
        v[4] = make_float2(e[0].x - o[0].x, e[0].y - o[0].y);
        v[1] = make_float2(e[1].x + o1.x, e[1].y + o1.y);
        v[5] = make_float2(e[1].x - o1.x, e[1].y - o1.y);
        v[2] = make_float2(e[2].x + o2.x, e[2].y + o2.y);
        v[6] = make_float2(e[2].x - o2.x, e[2].y - o2.y);
        v[3] = make_float2(e[3].x + o3.x, e[3].y + o3.y);
        v[7] = make_float2(e[3].x - o3.x, e[3].y - o3.y);
    } else if constexpr (R == 2) {
        float2 a = v[0], b = v[1];
        v[0] = make_float2(a.x + b.x, a.y + b.y);
        v[1] = make_float2(a.x - b.x, a.y - b.y);
    } else if constexpr (R == 4) {
        float2 s02 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
        float2 d02 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
        float2 s13 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
        float2 d13 = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
        // (sgn*i) * d13
        float2 jd = make_float2(-sgn * d13.y, sgn * d13.x);
        v[0] = make_float2(s02.x + s13.x, s02.y + s13.y);
        v[2] = make_float2(s02.x - s13.x, s02.y - s13.y);
        v[1] = make_float2(d02.x + jd.x, d02.y + jd.y);
        v[3] = make_float2(d02.x - jd.x, d02.y - jd.y);
    } else {
        // odd radix: pair j with R-j
        static_assert(R == 3 || R == 5 || R == 7, "radix");
        constexpr int H = (R - 1) / 2;
        // cos/sin(2*pi*m/R), m = 1..R-1 (exact decimal constants)
        const float c3[2] = {-0.5f, -0.5f};
        const float s3[2] = {0.86602540378443865f, -0.86602540378443865f};
        const float c5[4] = {0.30901699437494742f, -0.80901699437494742f, -0.80901699437494742f,
                             0.30901699437494742f};
        const float s5[4] = {0.95105651629515357f, 0.58778525229247313f, -0.58778525229247313f,
                             -0.95105651629515357f};
        const float c7[6] = {0.62348980185873353f, -0.22252093395631440f, -0.90096886790241913f,
                             -0.90096886790241913f, -0.22252093395631440f, 0.62348980185873353f};
        const float s7[6] = {0.78183148246802981f, 0.97492791218182361f, 0.43388373911755812f,
                             -0.43388373911755812f, -0.97492791218182361f, -0.78183148246802981f};
        const float* cs = R == 3 ? c3 : (R == 5 ? c5 : c7);
        const float* sn = R == 3 ? s3 : (R == 5 ? s5 : s7);
        float2 sum[H], dif[H];
        float2 x0 = v[0];
        float2 tot = x0;
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            sum[j - 1] = make_float2(v[j].x + v[R - j].x, v[j].y + v[R - j].y);
            dif[j - 1] = make_float2(v[j].x - v[R - j].x, v[j].y - v[R - j].y);
            tot.x += sum[j - 1].x;
            tot.y += sum[j - 1].y;
        }
        v[0] = tot;
#pragma unroll
        for (int q = 1; q <= H; ++q) {
            float2 a = x0, b = make_float2(0.f, 0.f);
#pragma unroll
            for (int j = 1; j <= H; ++j) {
                const int m = (j * q) % R;  // 1..R-1
                const float c = cs[m - 1], s = sn[m - 1];
                a.x += sum[j - 1].x * c;
                a.y += sum[j - 1].y * c;
                b.x += dif[j - 1].x * s;
                b.y += dif[j - 1].y * s;
            }
            // X_q = a + sgn*i*b, X_{R-q} = a - sgn*i*b
            v[q] = make_float2(a.x - sgn * b.y, a.y + sgn * b.x);
            v[R - q] = make_float2(a.x + sgn * b.y, a.y - sgn * b.x);
        }
    }
}

// i % p via a fixed-point reciprocal.  Small transforms (n <= 1024, so every
// index i < 512 and p <= 512): pinv = ceil(2^20/p), quotient (i*pinv) >> 20,
// exact because the rounding error i*(pinv - 2^20/p)/2^20 < 512/2^20 stays
// below 1/p.  Large transforms (BIG): pinv = ceil(2^32/p) (0 for p = 1) and
// the high half of the 64-bit product, exact while i*p < 2^32 (every side up
// to 65535).  The small form is the one the common kernels are compiled with
// (it keeps their register count low).
template <bool BIG>
__host__ __device__ inline unsigned mod_small_inv(int p) {
    if constexpr (BIG)
        return p <= 1 ? 0u : (unsigned)(((1ull << 32) + (unsigned long long)p - 1) / (unsigned long long)p);
    else
        return ((1u << 20) + (unsigned)p - 1) / (unsigned)p;
}
template <bool BIG>
__host__ __device__ inline int mod_small(int i, int p, unsigned pinv) {
    if constexpr (BIG) {
        const unsigned q = (unsigned)(((unsigned long long)(unsigned)i * (unsigned long long)pinv) >> 32);
        const int r = i - (int)q * p;
        return p == 1 ? 0 : r;
    } else {
        const int q = (int)(((unsigned)i * pinv) >> 20);
        return i - q * p;
    }
}
// largest transform length the small form handles
constexpr int kFftSmallMax = 1024;

// One butterfly of a Stockham pass of radix R over T interleaved transforms
// (w = butterfly index; nb = n/R; p = product of the earlier radices, pinv
// its mod_small reciprocal; twp = this pass's twiddles).  TW = false for the
// first pass (p = 1: no twiddles).
template <int R, int T, int SGN, bool TW, bool BIG>
__host__ __device__ __forceinline__ void fft_butterfly(const float2* __restrict__ src, float2* __restrict__ dst, int nb,
                                              int p, unsigned pinv, const float2* __restrict__ twp, int w) {
    constexpr int LT = T == 1 ? 0 : T == 2 ? 1 : T == 4 ? 2 : T == 8 ? 3 : T == 16 ? 4 : T == 32 ? 5 : 6;
    static_assert((1 << LT) == T, "T must be a power of two <= 64");
    const int i = w >> LT;
    const int t = w & (T - 1);
    float2 v[R];
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = src[w + j * nb * T];
    int k = 0;
    if constexpr (TW) {
        k = mod_small<BIG>(i, p, pinv);
        const float2* c = twp + k * (R - 1);
#pragma unroll
        for (int j = 1; j < R; ++j) {
            const float2 cj = c[j - 1];
            const float cy = SGN > 0 ? cj.y : -cj.y;
            v[j] = make_float2(v[j].x * cj.x - v[j].y * cy, v[j].x * cy + v[j].y * cj.x);
        }
    }
    dft_small<R>(v, (float)SGN);
    float2* d = dst + ((i - k) * R + k) * T + t;
#pragma unroll
    for (int j = 0; j < R; ++j) d[j * p * T] = v[j];
}

template <int R, int T, int SGN, bool TW, bool BIG>
__device__ __forceinline__ void fft_pass(const float2* __restrict__ a, float2* __restrict__ b, int n, int p,
                                         const float2* __restrict__ twp) {
    const int nb = n / R;
    const int total = nb * T;
    const unsigned pinv = TW ? mod_small_inv<BIG>(p) : 0u;
    for (int w = threadIdx.x; w < total; w += kFftThreads) fft_butterfly<R, T, SGN, TW, BIG>(a, b, nb, p, pinv, twp, w);
}

// Butterflies a thread holds in registers in a single-buffer pass of radix
// R: 8 complex values at most, 16 in kernels with 16 interleaved transforms
// (which run at fewer waves per SIMD), or CAP where the kernel sets it (32
// in the kernels for transform lengths 257..512, whose passes then fit one
// LDS buffer); at least one butterfly (radix 12 and 16 take one).
__host__ __device__ constexpr int fft_reg_cap(int T, int CAP) { return CAP > 0 ? CAP : (T >= 16 ? 16 : 8); }
__host__ __device__ constexpr int fft_inplace_maxb(int R, int T, int CAP = 0) {
    return fft_reg_cap(T, CAP) / R > 0 ? fft_reg_cap(T, CAP) / R : 1;
}

// Whether every pass of the transform (length n, T interleaved transforms)
// can run in a single buffer: each thread's butterflies (at most cap / R, at
// least one)
// fit its registers; cap = fft_reg_cap(T, CAP) of the kernel.
inline bool fft_inplace_ok(const FftFactors& f, int T, int cap) {
    for (int s = 0; s < f.nf; ++s)
        if ((f.n / f.r[s]) * T > (cap / f.r[s] > 0 ? cap / f.r[s] : 1) * kFftThreads) return false;
    return true;
}

// Single-buffer Stockham pass: every thread reads the inputs of its
// butterflies (at most fft_inplace_maxb(R)) into registers, the block
// synchronises, and the outputs overwrite the same buffer.
template <int R, int T, int SGN, bool TW, bool BIG, int CAP = 0>
__device__ __forceinline__ void fft_pass_inplace(float2* __restrict__ a, int n, int p,
                                                 const float2* __restrict__ twp) {
    constexpr int MAXB = fft_inplace_maxb(R, T, CAP);
    constexpr int LT = T == 1 ? 0 : T == 2 ? 1 : T == 4 ? 2 : T == 8 ? 3 : T == 16 ? 4 : T == 32 ? 5 : 6;
    const int nb = n / R;
    const int total = nb * T;
    const unsigned pinv = TW ? mod_small_inv<BIG>(p) : 0u;
    float2 v[MAXB][R];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int w = threadIdx.x + q * kFftThreads;
        if (w < total) {
#pragma unroll
            for (int j = 0; j < R; ++j) v[q][j] = a[w + j * nb * T];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int w = threadIdx.x + q * kFftThreads;
        if (w < total) {
            const int i = w >> LT;
            const int t = w & (T - 1);
            int k = 0;
            if constexpr (TW) {
                k = mod_small<BIG>(i, p, pinv);
                const float2* c = twp + k * (R - 1);
#pragma unroll
                for (int j = 1; j < R; ++j) {
                    const float2 cj = c[j - 1];
                    const float cy = SGN > 0 ? cj.y : -cj.y;
                    v[q][j] = make_float2(v[q][j].x * cj.x - v[q][j].y * cy, v[q][j].x * cy + v[q][j].y * cj.x);
                }
            }
            dft_small<R>(v[q], (float)SGN);
            float2* d = a + ((i - k) * R + k) * T + t;
#pragma unroll
            for (int j = 0; j < R; ++j) d[j * p * T] = v[q][j];
        }
    }
}

// Passes s0 .. s1 - 1 of a transform in LDS; tw = the pass twiddles of
// fft_pass_twiddles.  With b == nullptr every pass runs in the single buffer
// a (the caller checked fft_inplace_ok), else the passes ping-pong between a
// and b.  All threads of the block take part; returns the buffer holding the
// result.  The caller must __syncthreads() after filling `a` (and tw).
template <int T, int SGN, bool BIG = false, bool WIDE = false, int CAP = 0>
__device__ __forceinline__ float2* fft_lds_range(float2* a, float2* b, const FftFactors& f, const float2* tw, int s0,
                                                 int s1) {
    const int n = f.n;
    int p = 1;
    for (int s = 0; s < s0; ++s) p *= f.r[s];
    for (int s = s0; s < s1; ++s) {
        const int R = f.r[s];
        const float2* twp = tw + (p - 1);
        if (b == nullptr) {
            if (s == 0) {
                switch (R) {
                    case 16: if constexpr (WIDE) fft_pass_inplace<16, T, SGN, false, BIG, CAP>(a, n, p, twp); break;
                    case 12: if constexpr (WIDE) fft_pass_inplace<12, T, SGN, false, BIG, CAP>(a, n, p, twp); break;
                    case 8: fft_pass_inplace<8, T, SGN, false, BIG, CAP>(a, n, p, twp); break;
                    case 4: fft_pass_inplace<4, T, SGN, false, BIG, CAP>(a, n, p, twp); break;
                    case 2: fft_pass_inplace<2, T, SGN, false, BIG, CAP>(a, n, p, twp); break;
                    case 3: fft_pass_inplace<3, T, SGN, false, BIG, CAP>(a, n, p, twp); break;
                    case 5: fft_pass_inplace<5, T, SGN, false, BIG, CAP>(a, n, p, twp); break;
                    default: fft_pass_inplace<7, T, SGN, false, BIG, CAP>(a, n, p, twp); break;
                }
            } else {
                switch (R) {
                    case 16: if constexpr (WIDE) fft_pass_inplace<16, T, SGN, true, BIG, CAP>(a, n, p, twp); break;
                    case 12: if constexpr (WIDE) fft_pass_inplace<12, T, SGN, true, BIG, CAP>(a, n, p, twp); break;
                    case 8: fft_pass_inplace<8, T, SGN, true, BIG, CAP>(a, n, p, twp); break;
                    case 4: fft_pass_inplace<4, T, SGN, true, BIG, CAP>(a, n, p, twp); break;
                    case 2: fft_pass_inplace<2, T, SGN, true, BIG, CAP>(a, n, p, twp); break;
                    case 3: fft_pass_inplace<3, T, SGN, true, BIG, CAP>(a, n, p, twp); break;
                    case 5: fft_pass_inplace<5, T, SGN, true, BIG, CAP>(a, n, p, twp); break;
                    default: fft_pass_inplace<7, T, SGN, true, BIG, CAP>(a, n, p, twp); break;
                }
            }
            __syncthreads();
            p *= R;
            continue;
        }
        if (s == 0) {
            switch (R) {
                case 16: if constexpr (WIDE) fft_pass<16, T, SGN, false, BIG>(a, b, n, p, twp); break;
                case 12: if constexpr (WIDE) fft_pass<12, T, SGN, false, BIG>(a, b, n, p, twp); break;
                case 8: fft_pass<8, T, SGN, false, BIG>(a, b, n, p, twp); break;
                case 4: fft_pass<4, T, SGN, false, BIG>(a, b, n, p, twp); break;
                case 2: fft_pass<2, T, SGN, false, BIG>(a, b, n, p, twp); break;
                case 3: fft_pass<3, T, SGN, false, BIG>(a, b, n, p, twp); break;
                case 5: fft_pass<5, T, SGN, false, BIG>(a, b, n, p, twp); break;
                default: fft_pass<7, T, SGN, false, BIG>(a, b, n, p, twp); break;
            }
        } else {
            switch (R) {
                case 16: if constexpr (WIDE) fft_pass<16, T, SGN, true, BIG>(a, b, n, p, twp); break;
                case 12: if constexpr (WIDE) fft_pass<12, T, SGN, true, BIG>(a, b, n, p, twp); break;
                case 8: fft_pass<8, T, SGN, true, BIG>(a, b, n, p, twp); break;
                case 4: fft_pass<4, T, SGN, true, BIG>(a, b, n, p, twp); break;
                case 2: fft_pass<2, T, SGN, true, BIG>(a, b, n, p, twp); break;
                case 3: fft_pass<3, T, SGN, true, BIG>(a, b, n, p, twp); break;
                case 5: fft_pass<5, T, SGN, true, BIG>(a, b, n, p, twp); break;
                default: fft_pass<7, T, SGN, true, BIG>(a, b, n, p, twp); break;
            }
        }
        __syncthreads();
        p *= R;
        float2* t = a;
        a = b;
        b = t;
    }
    return a;
}

// Full transform in LDS (fft_lds_range over every pass).
template <int T, int SGN, bool BIG = false, bool WIDE = false, int CAP = 0>
__device__ __forceinline__ float2* fft_lds(float2* a, float2* b, const FftFactors& f, const float2* tw) {
    return fft_lds_range<T, SGN, BIG, WIDE, CAP>(a, b, f, tw, 0, f.nf);
}

// First pass (radix R, P = 1, no twiddles) with its inputs taken from
// src(q, t) -- the value at position q of interleaved transform t -- instead
// of LDS, so the data are loaded straight into the butterflies' registers
// (every load of a round issued before any use; rounds of the default
// register cap, whatever the kernel's single-buffer cap); the outputs go to
// `a`.
// The caller synchronises before the next pass reads `a`.
template <int R, int T, int SGN, class Src>
__device__ __forceinline__ void fft_first_pass_src(float2* __restrict__ a, int n, Src src) {
    constexpr int MAXB = fft_inplace_maxb(R, T);
    constexpr int LT = T == 1 ? 0 : T == 2 ? 1 : T == 4 ? 2 : T == 8 ? 3 : T == 16 ? 4 : T == 32 ? 5 : 6;
    const int nb = n / R;
    const int total = nb * T;
    for (int w0 = 0; w0 < total; w0 += MAXB * kFftThreads) {
        float2 v[MAXB][R];
#pragma unroll
        for (int q = 0; q < MAXB; ++q) {
            const int w = min(w0 + (int)threadIdx.x + q * kFftThreads, total - 1);
#pragma unroll
            for (int j = 0; j < R; ++j) v[q][j] = src((w >> LT) + j * nb, w & (T - 1));
        }
#pragma unroll
        for (int q = 0; q < MAXB; ++q) {
            const int w = w0 + (int)threadIdx.x + q * kFftThreads;
            if (w < total) {
                dft_small<R>(v[q], (float)SGN);
                float2* d = a + (w >> LT) * R * T + (w & (T - 1));
#pragma unroll
                for (int j = 0; j < R; ++j) d[j * T] = v[q][j];
            }
        }
    }
}

template <int T, int SGN, bool WIDE = false, class Src>
__device__ __forceinline__ void fft_first_pass_src(int R, float2* a, int n, Src src) {
    switch (R) {
        case 16: if constexpr (WIDE) fft_first_pass_src<16, T, SGN>(a, n, src); break;
        case 12: if constexpr (WIDE) fft_first_pass_src<12, T, SGN>(a, n, src); break;
        case 8: fft_first_pass_src<8, T, SGN>(a, n, src); break;
        case 4: fft_first_pass_src<4, T, SGN>(a, n, src); break;
        case 2: fft_first_pass_src<2, T, SGN>(a, n, src); break;
        case 3: fft_first_pass_src<3, T, SGN>(a, n, src); break;
        case 5: fft_first_pass_src<5, T, SGN>(a, n, src); break;
        default: fft_first_pass_src<7, T, SGN>(a, n, src); break;
    }
}

// Last pass (radix R after the radices of product p > 1) reading its inputs
// from LDS `a` and handing each output to dst(q, t, value) -- position q of
// transform t -- instead of writing LDS back.  No synchronisation needed
// after it.
template <int R, int T, int SGN, bool BIG, class Dst>
__device__ __forceinline__ void fft_last_pass_dst(const float2* __restrict__ a, int n, int p,
                                                  const float2* __restrict__ twp, Dst dst) {
    constexpr int MAXB = fft_inplace_maxb(R, T);
    constexpr int LT = T == 1 ? 0 : T == 2 ? 1 : T == 4 ? 2 : T == 8 ? 3 : T == 16 ? 4 : T == 32 ? 5 : 6;
    const int nb = n / R;
    const int total = nb * T;
    const unsigned pinv = mod_small_inv<BIG>(p);
    for (int w0 = 0; w0 < total; w0 += MAXB * kFftThreads) {
        float2 v[MAXB][R];
#pragma unroll
        for (int q = 0; q < MAXB; ++q) {
            const int w = min(w0 + (int)threadIdx.x + q * kFftThreads, total - 1);
#pragma unroll
            for (int j = 0; j < R; ++j) v[q][j] = a[w + j * nb * T];
        }
#pragma unroll
        for (int q = 0; q < MAXB; ++q) {
            const int w = w0 + (int)threadIdx.x + q * kFftThreads;
            if (w < total) {
                const int i = w >> LT;
                const int t = w & (T - 1);
                const int k = mod_small<BIG>(i, p, pinv);
                const float2* c = twp + k * (R - 1);
#pragma unroll
                for (int j = 1; j < R; ++j) {
                    const float2 cj = c[j - 1];
                    const float cy = SGN > 0 ? cj.y : -cj.y;
                    v[q][j] = make_float2(v[q][j].x * cj.x - v[q][j].y * cy, v[q][j].x * cy + v[q][j].y * cj.x);
                }
                dft_small<R>(v[q], (float)SGN);
#pragma unroll
                for (int j = 0; j < R; ++j) dst((i - k) * R + k + j * p, t, v[q][j]);
            }
        }
    }
}

template <int T, int SGN, bool BIG, bool WIDE = false, class Dst>
__device__ __forceinline__ void fft_last_pass_dst(int R, const float2* a, int n, int p, const float2* twp, Dst dst) {
    switch (R) {
        case 16: if constexpr (WIDE) fft_last_pass_dst<16, T, SGN, BIG>(a, n, p, twp, dst); break;
        case 12: if constexpr (WIDE) fft_last_pass_dst<12, T, SGN, BIG>(a, n, p, twp, dst); break;
        case 8: fft_last_pass_dst<8, T, SGN, BIG>(a, n, p, twp, dst); break;
        case 4: fft_last_pass_dst<4, T, SGN, BIG>(a, n, p, twp, dst); break;
        case 2: fft_last_pass_dst<2, T, SGN, BIG>(a, n, p, twp, dst); break;
        case 3: fft_last_pass_dst<3, T, SGN, BIG>(a, n, p, twp, dst); break;
        case 5: fft_last_pass_dst<5, T, SGN, BIG>(a, n, p, twp, dst); break;
        default: fft_last_pass_dst<7, T, SGN, BIG>(a, n, p, twp, dst); break;
    }
}

}  // namespace mm

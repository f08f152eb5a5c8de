// Mixed-radix (2,3,4,5,7) Stockham FFT on data resident in LDS.
//
// Layout: element (position p, transform t) of `T` interleaved transforms of
// length n lives at buf[p*T + t], so consecutive lanes work on consecutive
// transforms (conflict-free LDS access).  Each pass of radix R reads R
// inputs spaced n/R apart, applies the twiddles exp(sgn*2*pi*i*j*k/(P*R)),
// does a length-R DFT and writes the R outputs spaced P apart (P = product of
// the radices already done), which leaves the result in natural order after
// the last pass.  sgn = +1 is MATLAB's ifft kernel (unscaled), sgn = -1 its
// fft kernel.
#pragma once
#include <hip/hip_runtime.h>

namespace mm {

constexpr int kMaxFactors = 12;

struct FftFactors {
    int n = 1;
    int nf = 0;
    int r[kMaxFactors] = {};
};

// Radix order: 4s first, then 2, 3, 5, 7.  Returns false when n has a prime
// factor above 7 (the grid-size rule of nufftOp.m:146-152 never produces one).
inline bool fft_factorize(int n, FftFactors& f) {
    f = FftFactors();
    f.n = n;
    int m = n;
    while (m % 4 == 0) { f.r[f.nf++] = 4; m /= 4; }
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

__host__ __device__ inline float2 cmulf(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// In-register length-R DFT: X_q = sum_j v_j exp(sgn*2*pi*i*j*q/R).
template <int R>
__host__ __device__ inline void dft_small(float2 (&v)[R], float sgn) {
    if constexpr (R == 2) {
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

// One butterfly of a Stockham pass over T interleaved transforms
// (T = 1 << log2T); tw[t] = exp(+2*pi*i*t/n), t < n; step = n/(p*R).
template <int R, bool BIG = false>
__host__ __device__ inline void stockham_butterfly(const float2* __restrict__ src,
                                                   float2* __restrict__ dst, int n, int log2T,
                                                   int p, unsigned pinv, int step,
                                                   const float2* __restrict__ tw, float sgn, int w) {
    const int T = 1 << log2T;
    const int nb = n / R;
    const int t = w & (T - 1);
    const int i = w >> log2T;
    const int k = mod_small<BIG>(i, p, pinv);
    float2 v[R];
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = src[(i + j * nb) * T + t];
#pragma unroll
    for (int j = 1; j < R; ++j) {
        float2 c = tw[j * k * step];
        c.y *= sgn;
        v[j] = cmulf(v[j], c);
    }
    dft_small<R>(v, sgn);
    const int o = (i - k) * R + k;
#pragma unroll
    for (int j = 0; j < R; ++j) dst[(o + j * p) * T + t] = v[j];
}

// Full transform in LDS, ping-ponging between a and b.  All threads of the
// block take part; returns the buffer holding the result.  The caller must
// __syncthreads() after filling `a`.
template <bool BIG = false>
__device__ inline float2* fft_lds(float2* a, float2* b, const FftFactors& f, int log2T,
                                  const float2* tw, float sgn) {
    const int n = f.n;
    const int T = 1 << log2T;
    int p = 1;
    for (int s = 0; s < f.nf; ++s) {
        const int R = f.r[s];
        const int total = (n / R) * T;
        const unsigned pinv = mod_small_inv<BIG>(p);
        const int step = n / (p * R);
        for (int w = threadIdx.x; w < total; w += blockDim.x) {
            switch (R) {
                case 4: stockham_butterfly<4, BIG>(a, b, n, log2T, p, pinv, step, tw, sgn, w); break;
                case 2: stockham_butterfly<2, BIG>(a, b, n, log2T, p, pinv, step, tw, sgn, w); break;
                case 3: stockham_butterfly<3, BIG>(a, b, n, log2T, p, pinv, step, tw, sgn, w); break;
                case 5: stockham_butterfly<5, BIG>(a, b, n, log2T, p, pinv, step, tw, sgn, w); break;
                default: stockham_butterfly<7, BIG>(a, b, n, log2T, p, pinv, step, tw, sgn, w); break;
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

}  // namespace mm

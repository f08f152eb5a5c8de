// MFMA path of the expanded-encoding operator (gfx950).
//
// Both directions are one real GEMM whose left operand is generated on the
// fly:  out[row, n] = sum_p A'[row, 2p + {0,1}] * B'[2p + {0,1}, n]  with
//   A'[row, 2p] = cos(2*pi*Phi), A'[row, 2p+1] = sin(2*pi*Phi),
//   Phi(row, p) = sum_j rowfac[row][j] * redfac[p][j]   (revolutions, FP64)
// forward: row = sample k, p = voxel v, B' rows (Xr | Xi), (-Xi | Xr)
// adjoint: row = voxel v, p = sample k, B' rows (Yr | Yi), (Yi | -Yr)
// and the output columns hold the real parts (first half) and imaginary parts
// (second half) of the complex result.
//
// Precision: the phase is contracted in FP64 and reduced to a quarter
// revolution before an FP32 polynomial sine/cosine.  The MFMA is
// v_mfma_f32_16x16x32_f16 with split operands a = hi + lo/2048 (hi, lo fp16,
// lo scaled by 2^11 so it stays a normal number); hi*hi goes to one FP32
// accumulator and hi*lo + lo*hi to a second one, combined at the end.  B' is
// normalised by a power of two taken from its max magnitude (computed on the
// device, no host round trip) and the scale is undone in the reduction.
//
// Work decomposition: a 512-thread work-group owns 256 rows (8 waves x 2
// row tiles of 16) and one range of the reduction index; per 16-wide
// reduction block the B' operands (pre-split, in MFMA lane order) and the
// 16 reduction factors are staged through double-buffered LDS.  Partial sums
// of the reduction ranges are added by the reduce kernels.
#include <cmath>

#include "hio_mfma.hpp"

namespace mm {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr float kLoScale = 2048.f;
constexpr float kLoInv = 1.f / 2048.f;

// cos, sin of 2*pi*ph (ph in revolutions, FP64) via quarter-turn reduction
// and Taylor polynomials on [-pi/4, pi/4] (truncation below 2e-9).
__device__ __forceinline__ void sincos_rev(double ph, float& c, float& s) {
    const double q4 = rint(4.0 * ph);
    const double red = fma(q4, -0.25, ph);
    const int q = (int)q4 & 3;   // |ph| < 2^29 revolutions
    const float t = (float)red * 6.28318530717958647692f;
    const float t2 = t * t;
    float sp = fmaf(t2, 1.f / 362880.f, -1.f / 5040.f);
    sp = fmaf(t2, sp, 1.f / 120.f);
    sp = fmaf(t2, sp, -1.f / 6.f);
    sp = t * fmaf(t2, sp, 1.f);
    float cp = fmaf(t2, -1.f / 3628800.f, 1.f / 40320.f);
    cp = fmaf(t2, cp, -1.f / 720.f);
    cp = fmaf(t2, cp, 1.f / 24.f);
    cp = fmaf(t2, cp, -0.5f);
    cp = fmaf(t2, cp, 1.f);
    const bool sw = q & 1;
    const float a = sw ? sp : cp, b = sw ? cp : sp;
    c = (q == 1 || q == 2) ? -a : a;
    s = (q >= 2) ? -b : b;
}

__device__ __forceinline__ void split_f16(float v, _Float16& hi, _Float16& lo) {
    hi = (_Float16)v;
    lo = (_Float16)((v - (float)hi) * kLoScale);
}

// power-of-two scale that maps [0, amax] into [0, 1)
__device__ __forceinline__ int amax_exp(const unsigned* amax_bits) {
    const float m = __uint_as_float(*amax_bits);
    int e = 0;
    if (m > 0.f) (void)frexpf(m, &e);
    return e;
}

typedef double double4_t __attribute__((ext_vector_type(4)));

// cos/sin of 2*pi*ph with the reduction to [-1/2, 1/2] revolution in FP64
// and the hardware sine/cosine (v_sin_f32 / v_cos_f32 take revolutions).
__device__ __forceinline__ void sincos_rev_hw(double ph, float& c, float& s) {
    const float r = (float)(ph - rint(ph));
    s = __builtin_amdgcn_sinf(r);
    c = __builtin_amdgcn_cosf(r);
}

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// cos/sin of 2*pi*ph (revolutions, FP64) with the hardware sine/cosine on
// the fractional revolution, split into fp16 hi and scaled lo parts for the
// A operand: (ah.x, ah.y) = hi(cos, sin), (al.x, al.y) = lo(cos, sin) * 2048.
// The hi pair is one packed round-toward-zero conversion, and each lo is
// 2048*v - 2048*hi in one mixed-precision FMA (hi read back as fp16), both
// exact: v - hi is representable and the scaling is a power of two.
__device__ __forceinline__ void sincos_split_hw(double ph, half2_t& ah, half2_t& al) {
    const float r = (float)__builtin_amdgcn_fract(ph);   // [0, 1) revolution
    const float s = __builtin_amdgcn_sinf(r);
    const float c = __builtin_amdgcn_cosf(r);
    ah = __builtin_bit_cast(half2_t, __builtin_amdgcn_cvt_pkrtz(c, s));
    const float lc = fmaf((float)ah.x, -kLoScale, c * kLoScale);
    const float ls = fmaf((float)ah.y, -kLoScale, s * kLoScale);
    al = __builtin_bit_cast(half2_t, __builtin_amdgcn_cvt_pkrtz(lc, ls));
}

// Variants: MPHASE computes the phase tile on the FP64 MFMA (16x16x4: lane l
// holds rows l/16 + 4i of column l%16, so the A rows are permuted to deliver
// the 4 consecutive reduction indices the f16 16x16x32 A operand needs),
// otherwise on FP64 VALU;
// HWSIN selects the hardware sine/cosine over the polynomial.
template <int NCT, int RT, int RP, bool MPHASE, bool HWSIN>
__global__ __launch_bounds__(512) void hio_mfma_gemm(const double* __restrict__ rowfac, int nrows,
                                                     const double* __restrict__ redfac, int nblk_total,
                                                     const half8_t* __restrict__ Bs,
                                                     float* __restrict__ part, int blk_per_split) {
    constexpr int NB = NCT * 2 * 64;   // half8 operands per reduction block
    constexpr int NR = 16 * RP;        // reduction factors per block
    constexpr int NS = RP / 4;         // FP64 MFMA K-steps
    // padded to one entry per thread: every thread loads and stores each
    // block (no load or store under a branch)
    static_assert(NB <= 512 && NR <= 512, "operands per block exceed the work-group");
    __shared__ half8_t sB[2][512];
    __shared__ double sR[2][512];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, l16 = lane & 15;
    const int rowbase = blockIdx.x * (8 * RT * 16) + wave * RT * 16;
    const int split = blockIdx.y;
    const int b0 = split * blk_per_split;
    const int b1 = min(nblk_total, b0 + blk_per_split);
    // row factors: VALU variant keeps the whole row, MFMA variant its B operand
    double rf[RT][MPHASE ? NS : RP];
#pragma unroll
    for (int t = 0; t < RT; ++t) {
        const int row = rowbase + 16 * t + l16;
        if constexpr (MPHASE) {
#pragma unroll
            for (int q = 0; q < NS; ++q) rf[t][q] = row < nrows ? rowfac[(size_t)row * RP + 4 * q + g] : 0.0;
        } else {
#pragma unroll
            for (int j = 0; j < RP; ++j) rf[t][j] = row < nrows ? rowfac[(size_t)row * RP + j] : 0.0;
        }
    }
    float4_t accm[RT][NCT], accl[RT][NCT];
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int c = 0; c < NCT; ++c) {
            accm[t][c] = float4_t{0.f, 0.f, 0.f, 0.f};
            accl[t][c] = float4_t{0.f, 0.f, 0.f, 0.f};
        }
    half8_t pfB;
    double pfR = 0.0;
    const int tid = threadIdx.x;
    if (b0 < b1) {
        sB[0][tid] = Bs[(size_t)b0 * NB + (tid % NB)];
        sR[0][tid] = redfac[(size_t)b0 * NR + min(tid, NR - 1)];
    }
    __syncthreads();
    for (int b = b0; b < b1; ++b) {
        const int buf = (b - b0) & 1;
        const bool more = b + 1 < b1;
        {
            // unconditional loads from clamped, always-valid indices: a load
            // under a branch is waited for at the join, which would stall every
            // block on its own prefetch
            const size_t bn = (size_t)(more ? b + 1 : b);
            pfB = Bs[bn * NB + (tid % NB)];
            pfR = redfac[bn * NR + min(tid, NR - 1)];
        }
        double ph[RT][4];
        if constexpr (MPHASE) {
            double a[NS];
#pragma unroll
            // The FP64 MFMA returns row g + 4i of column l16 in lane item i;
            // loading reduction index 4*(m%4) + m/4 into A row m puts index
            // 4g + i there, the order the f16 A operand needs.
            for (int q = 0; q < NS; ++q) a[q] = sR[buf][(4 * (l16 & 3) + (l16 >> 2)) * RP + 4 * q + g];
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                double4_t d = double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int q = 0; q < NS; ++q) d = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], rf[t][q], d, 0, 0, 0);
#pragma unroll
                for (int i = 0; i < 4; ++i) ph[t][i] = d[i];
            }
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                double rs[RP];
#pragma unroll
                for (int j = 0; j < RP; ++j) rs[j] = sR[buf][(4 * g + i) * RP + j];
#pragma unroll
                for (int t = 0; t < RT; ++t) {
                    double acc = 0.0;
#pragma unroll
                    for (int j = 0; j < RP; ++j) acc = fma(rf[t][j], rs[j], acc);
                    ph[t][i] = acc;
                }
            }
        }
        half8_t Ah[RT], Al[RT];
#pragma unroll
        for (int t = 0; t < RT; ++t) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if constexpr (HWSIN) {
                    half2_t h2, l2;
                    sincos_split_hw(ph[t][i], h2, l2);
                    Ah[t][2 * i] = h2.x;
                    Ah[t][2 * i + 1] = h2.y;
                    Al[t][2 * i] = l2.x;
                    Al[t][2 * i + 1] = l2.y;
                    continue;
                }
                float cv, sv;
                sincos_rev(ph[t][i], cv, sv);
                _Float16 h, l;
                split_f16(cv, h, l);
                Ah[t][2 * i] = h;
                Al[t][2 * i] = l;
                split_f16(sv, h, l);
                Ah[t][2 * i + 1] = h;
                Al[t][2 * i + 1] = l;
            }
        }
#pragma unroll
        for (int c = 0; c < NCT; ++c) {
            const half8_t Bh = sB[buf][(c * 2 + 0) * 64 + lane];
            const half8_t Bl = sB[buf][(c * 2 + 1) * 64 + lane];
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                accm[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[t], Bh, accm[t][c], 0, 0, 0);
                accl[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[t], Bl, accl[t][c], 0, 0, 0);
                accl[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al[t], Bh, accl[t][c], 0, 0, 0);
            }
        }
        // the other buffer was last read before the previous barrier; in the
        // last iteration this stores a copy of the current block, never read
        sB[buf ^ 1][tid] = pfB;
        sR[buf ^ 1][tid] = pfR;
        __syncthreads();
    }
    // D layout of the 16x16 MFMA: lane holds rows 4*(lane/16)+i, column lane%16
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int c = 0; c < NCT; ++c)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = rowbase + 16 * t + 4 * g + i;
                if (row < nrows)
                    part[((size_t)split * nrows + row) * (16 * NCT) + 16 * c + l16] =
                        accm[t][c][i] + accl[t][c][i] * kLoInv;
            }
}

// Software-pipelined form of the FP64-MFMA-phase / hardware-sine variant.
// Iteration b issues the fp16 MFMAs of reduction block b and, independent of
// them in the same basic block, the FP64 phase MFMAs, sine/cosine and fp16
// split of block b+1, so one wave keeps the matrix pipe busy while its VALU
// prepares the next A operands.  The reduction factors come straight from
// global memory (L2) two blocks ahead; only the pre-split B' operands go
// through double-buffered LDS.  Loads use clamped, always-valid indices (a
// load under a branch would be waited for at the join); the extra phase
// work for the clamped block b+1 of the last iteration is discarded.  Same
// products, same accumulation order as hio_mfma_gemm<.., true, true>.
template <int NCT, int RT, int RP>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 4))) void hio_mfma_gemm_pipe(const double* __restrict__ rowfac, int nrows,
                                                          const double* __restrict__ redfac, int nblk_total,
                                                          const half8_t* __restrict__ Bs,
                                                          float* __restrict__ part, int blk_per_split) {
    constexpr int NB = NCT * 2 * 64;   // half8 operands per reduction block
    constexpr int NS = RP / 4;         // FP64 MFMA K-steps
    __shared__ half8_t sB[2][NB];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, l16 = lane & 15;
    const int rowbase = blockIdx.x * (8 * RT * 16) + wave * RT * 16;
    const int b0 = blockIdx.y * blk_per_split;
    const int b1 = min(nblk_total, b0 + blk_per_split);
    const int tid = threadIdx.x;
    // FP64 A row m holds reduction index 4*(m%4) + m/4, so the MFMA result
    // lane item i is index 4g + i, the order the f16 A operand needs
    const int arow = 4 * (l16 & 3) + (l16 >> 2);
    double rf[RT][NS];
#pragma unroll
    for (int t = 0; t < RT; ++t) {
        const int row = rowbase + 16 * t + l16;
#pragma unroll
        for (int q = 0; q < NS; ++q) rf[t][q] = row < nrows ? rowfac[(size_t)row * RP + 4 * q + g] : 0.0;
    }
    auto load_a = [&](int b, double (&a)[NS]) {
#pragma unroll
        for (int q = 0; q < NS; ++q) a[q] = redfac[((size_t)b * 16 + arow) * RP + 4 * q + g];
    };
    auto make_a = [&](const double (&a)[NS], half8_t (&Ah)[RT], half8_t (&Al)[RT]) {
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            double4_t d = double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < NS; ++q) d = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], rf[t][q], d, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                half2_t h2, l2;
                sincos_split_hw(d[i], h2, l2);
                Ah[t][2 * i] = h2.x;
                Ah[t][2 * i + 1] = h2.y;
                Al[t][2 * i] = l2.x;
                Al[t][2 * i + 1] = l2.y;
            }
        }
    };
    float4_t accm[RT][NCT], accl[RT][NCT];
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int c = 0; c < NCT; ++c) {
            accm[t][c] = float4_t{0.f, 0.f, 0.f, 0.f};
            accl[t][c] = float4_t{0.f, 0.f, 0.f, 0.f};
        }
    if (b0 < b1) {
        const int blast = b1 - 1;
        half8_t Ah[RT], Al[RT];
        double an[NS];
        {
            double a0[NS];
            load_a(b0, a0);
            load_a(min(b0 + 1, blast), an);
            if (tid < NB) sB[0][tid] = Bs[(size_t)b0 * NB + tid];
            make_a(a0, Ah, Al);
        }
        __syncthreads();
        for (int b = b0; b < b1; ++b) {
            const int buf = (b - b0) & 1;
            const half8_t pfB = Bs[(size_t)min(b + 1, blast) * NB + (tid % NB)];
            double an2[NS];
            load_a(min(b + 2, blast), an2);
            half8_t Nh[RT], Nl[RT];
            make_a(an, Nh, Nl);
#pragma unroll
            for (int c = 0; c < NCT; ++c) {
                const half8_t Bh = sB[buf][(c * 2 + 0) * 64 + lane];
                const half8_t Bl = sB[buf][(c * 2 + 1) * 64 + lane];
#pragma unroll
                for (int t = 0; t < RT; ++t) {
                    accm[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[t], Bh, accm[t][c], 0, 0, 0);
                    accl[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[t], Bl, accl[t][c], 0, 0, 0);
                    accl[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al[t], Bh, accl[t][c], 0, 0, 0);
                }
            }
            if (b + 1 < b1 && tid < NB) sB[buf ^ 1][tid] = pfB;
            __syncthreads();
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                Ah[t] = Nh[t];
                Al[t] = Nl[t];
            }
#pragma unroll
            for (int q = 0; q < NS; ++q) an[q] = an2[q];
        }
    }
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int c = 0; c < NCT; ++c)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = rowbase + 16 * t + 4 * g + i;
                if (row < nrows)
                    part[((size_t)blockIdx.y * nrows + row) * (16 * NCT) + 16 * c + l16] =
                        accm[t][c][i] + accl[t][c][i] * kLoInv;
            }
}

// value of the right-hand operand z(p, column) before normalisation
struct PrepArgs {
    const float2* in;
    const float2* maps;
    int mode;        // 0 forward + coil maps, 1 forward plain, 2 adjoint
    int C;           // coils (mode 0)
    int col0, gw;    // first column and width of this column group
    int64_t nred;    // reduction length (V forward, K adjoint)
};

__device__ __forceinline__ float2 prep_value(const PrepArgs& a, int64_t p, int c) {
    const int cc = a.col0 + c;
    if (a.mode == 0) {
        const int rep = cc / a.C, coil = cc % a.C;
        return cmulf(a.in[(int64_t)rep * a.nred + p], a.maps[(int64_t)coil * a.nred + p]);
    }
    return a.in[(int64_t)cc * a.nred + p];
}

__global__ void hio_zero_u32(unsigned* p) { *p = 0u; }

__global__ __launch_bounds__(256) void hio_absmax(PrepArgs a, unsigned* __restrict__ amax) {
    __shared__ float red[4];
    float m = 0.f;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < a.nred;
         p += (int64_t)gridDim.x * blockDim.x)
        for (int c = 0; c < a.gw; ++c) {
            const float2 z = prep_value(a, p, c);
            m = fmaxf(m, fmaxf(fabsf(z.x), fabsf(z.y)));
        }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_down(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        atomicMax(amax, __float_as_uint(m));
    }
}

// B' operands in MFMA lane order: Bs[blk][ct][hi/lo][lane] (half8).
// Lane l of column tile ct holds column n = 16*ct + l%16 and reduction rows
// 8*(l/16) + jj, i.e. p = 16*blk + 4*(l/16) + jj/2 with jj%2 selecting the
// row pair (2p, 2p+1).
template <int NCT>
__global__ void hio_prep_b(PrepArgs a, const unsigned* __restrict__ amax, half8_t* __restrict__ Bs,
                           int nblk) {
    const int64_t total = (int64_t)nblk * NCT * 64;
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int lane = (int)(idx & 63);
    const int ct = (int)((idx >> 6) % NCT);
    const int64_t blk = idx / (64 * NCT);
    const int g = lane >> 4, l16 = lane & 15;
    const int n = 16 * ct + l16;
    const int nh = 8 * NCT;
    const bool imag = n >= nh;
    const int c = imag ? n - nh : n;
    const float sc = ldexpf(1.f, -amax_exp(amax));
    const float s1 = a.mode == 2 ? 1.f : -1.f;   // sign of the second row, real half
    const float s2 = -s1;                         // sign of the second row, imaginary half
    half8_t H, L;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t p = blk * 16 + 4 * g + i;
        float2 z = make_float2(0.f, 0.f);
        if (p < a.nred && c < a.gw) z = prep_value(a, p, c);
        z.x *= sc;
        z.y *= sc;
        const float r0 = imag ? z.y : z.x;
        const float r1 = imag ? s2 * z.x : s1 * z.y;
        _Float16 h, l;
        split_f16(r0, h, l);
        H[2 * i] = h;
        L[2 * i] = l;
        split_f16(r1, h, l);
        H[2 * i + 1] = h;
        L[2 * i + 1] = l;
    }
    Bs[(blk * NCT + ct) * 128 + lane] = H;
    Bs[(blk * NCT + ct) * 128 + 64 + lane] = L;
}

// One wave per row; lane j < NC16 sums column j of the real GEMM over the
// splits (coalesced 4*NC16-byte rows), lane c pairs its real part with the
// imaginary part held by lane c + NC16/2.
__device__ __forceinline__ float sum_splits(const float* __restrict__ part, int nsplit, int nrows, int NC16,
                                            int row, int lane) {
    float v = 0.f;
    if (lane < NC16)
        for (int s = 0; s < nsplit; ++s) v += part[((size_t)s * nrows + row) * NC16 + lane];
    return v;
}

// forward: y[(col0+c)*K + k] = scale * (re, im)  (* W)
__global__ __launch_bounds__(256) void hio_mfma_reduce_fwd(const float* __restrict__ part, int nsplit,
                                                           int nrows, int NC16, int col0, int gw,
                                                           const unsigned* __restrict__ amax, float post,
                                                           float2* __restrict__ y, const float* __restrict__ W) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nrows) return;
    const float sc = ldexpf(post, amax_exp(amax));
    const int nh = NC16 / 2;
    const float v = sum_splits(part, nsplit, nrows, NC16, row, lane);
    const float im = __shfl(v, (lane + nh) & 63, 64);
    if (lane < gw) {
        const int64_t o = (int64_t)(col0 + lane) * nrows + row;
        float f = sc;
        if (W) f *= W[o];
        y[o] = make_float2(v * f, im * f);
    }
}

// adjoint: x per column, or combined with conj(maps) over the C coils of each
// repetition (gw is a multiple of C)
__global__ __launch_bounds__(256) void hio_mfma_reduce_adj(const float* __restrict__ part, int nsplit,
                                                           int nrows, int NC16, int col0, int gw,
                                                           const unsigned* __restrict__ amax, float post,
                                                           const float2* __restrict__ maps, int C,
                                                           float2* __restrict__ x) {
    __shared__ float2 terms[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row = blockIdx.x * 4 + w;
    const bool valid = row < nrows;
    const float sc = ldexpf(post, amax_exp(amax));
    const int nh = NC16 / 2;
    const float v = valid ? sum_splits(part, nsplit, nrows, NC16, row, lane) : 0.f;
    const float im = __shfl(v, (lane + nh) & 63, 64);
    if (!maps) {
        if (valid && lane < gw) x[(int64_t)(col0 + lane) * nrows + row] = make_float2(v * sc, im * sc);
        return;
    }
    float2 t = make_float2(0.f, 0.f);
    if (valid && lane < gw) {
        const float2 m = maps[(int64_t)(lane % C) * nrows + row];
        t = make_float2(m.x * v + m.y * im, m.x * im - m.y * v);   // conj(maps) * (re, im)
    }
    terms[w][lane] = t;
    __syncthreads();
    if (valid && lane < gw / C) {
        float2 acc = make_float2(0.f, 0.f);
        for (int c = 0; c < C; ++c) {
            acc.x += terms[w][lane * C + c].x;
            acc.y += terms[w][lane * C + c].y;
        }
        x[(int64_t)(col0 / C + lane) * nrows + row] = make_float2(acc.x * sc, acc.y * sc);
    }
}

static int grid_1d(int64_t n, int bs = 256) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + bs - 1) / bs, 65536));
}

template <int NCT, int RP>
static void launch_gemm(hipStream_t st, const HioMfmaCall& c, int gx, int nsplit, int bps) {
    if (c.variant & 4) {
        hipLaunchKernelGGL((hio_mfma_gemm_pipe<NCT, 2, RP>), dim3(gx, nsplit), dim3(512), 0, st, c.rowfac, c.nrows,
                           c.redfac, c.nblk, (const half8_t*)c.Bs, c.part, bps);
        return;
    }
    switch (c.variant & 3) {
        case 0:
            hipLaunchKernelGGL((hio_mfma_gemm<NCT, 2, RP, false, false>), dim3(gx, nsplit), dim3(512), 0, st,
                               c.rowfac, c.nrows, c.redfac, c.nblk, (const half8_t*)c.Bs, c.part, bps);
            break;
        case 1:
            hipLaunchKernelGGL((hio_mfma_gemm<NCT, 2, RP, true, false>), dim3(gx, nsplit), dim3(512), 0, st,
                               c.rowfac, c.nrows, c.redfac, c.nblk, (const half8_t*)c.Bs, c.part, bps);
            break;
        case 2:
            hipLaunchKernelGGL((hio_mfma_gemm<NCT, 2, RP, false, true>), dim3(gx, nsplit), dim3(512), 0, st,
                               c.rowfac, c.nrows, c.redfac, c.nblk, (const half8_t*)c.Bs, c.part, bps);
            break;
        default:
            hipLaunchKernelGGL((hio_mfma_gemm<NCT, 2, RP, true, true>), dim3(gx, nsplit), dim3(512), 0, st,
                               c.rowfac, c.nrows, c.redfac, c.nblk, (const half8_t*)c.Bs, c.part, bps);
            break;
    }
}

void hio_mfma_apply(hipStream_t st, const HioMfmaCall& c, int prof_iters, double* prof_ms) {
    const int NCT = c.gw <= 16 ? 2 : 4;
    const int NC16 = 16 * NCT;
    PrepArgs a;
    a.in = c.in;
    a.maps = c.maps;
    a.mode = c.mode;
    a.C = c.C;
    a.col0 = c.col0;
    a.gw = c.gw;
    a.nred = c.nred;
    hipLaunchKernelGGL(hio_zero_u32, dim3(1), dim3(1), 0, st, c.amax);
    hipLaunchKernelGGL(hio_absmax, dim3(grid_1d(c.nred)), dim3(256), 0, st, a, c.amax);
    const int64_t nprep = (int64_t)c.nblk * NCT * 64;
    if (NCT == 2)
        hipLaunchKernelGGL(hio_prep_b<2>, dim3(grid_1d(nprep)), dim3(256), 0, st, a, (const unsigned*)c.amax,
                           (half8_t*)c.Bs, c.nblk);
    else
        hipLaunchKernelGGL(hio_prep_b<4>, dim3(grid_1d(nprep)), dim3(256), 0, st, a, (const unsigned*)c.amax,
                           (half8_t*)c.Bs, c.nblk);
    const int gx = (c.nrows + 255) / 256;
    auto gemm = [&] {
        switch (c.rp * 10 + NCT) {
            case 42: launch_gemm<2, 4>(st, c, gx, c.nsplit, c.bps); break;
            case 44: launch_gemm<4, 4>(st, c, gx, c.nsplit, c.bps); break;
            case 82: launch_gemm<2, 8>(st, c, gx, c.nsplit, c.bps); break;
            case 84: launch_gemm<4, 8>(st, c, gx, c.nsplit, c.bps); break;
            case 122: launch_gemm<2, 12>(st, c, gx, c.nsplit, c.bps); break;
            case 124: launch_gemm<4, 12>(st, c, gx, c.nsplit, c.bps); break;
            case 162: launch_gemm<2, 16>(st, c, gx, c.nsplit, c.bps); break;
            default: launch_gemm<4, 16>(st, c, gx, c.nsplit, c.bps); break;
        }
    };
    if (prof_iters > 0) {
        hipEvent_t e0, e1;
        HIPCHECK(hipEventCreate(&e0));
        HIPCHECK(hipEventCreate(&e1));
        gemm();
        HIPCHECK(hipEventRecord(e0, st));
        for (int i = 0; i < prof_iters; ++i) gemm();
        HIPCHECK(hipEventRecord(e1, st));
        HIPCHECK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
        *prof_ms = (double)ms / prof_iters;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    } else {
        gemm();
    }
    if (c.mode == 2)
        hipLaunchKernelGGL(hio_mfma_reduce_adj, dim3((c.nrows + 3) / 4), dim3(256), 0, st,
                           (const float*)c.part, c.nsplit, c.nrows, NC16, c.col0, c.gw,
                           (const unsigned*)c.amax, c.post, c.maps, c.C, c.out);
    else
        hipLaunchKernelGGL(hio_mfma_reduce_fwd, dim3((c.nrows + 3) / 4), dim3(256), 0, st,
                           (const float*)c.part, c.nsplit, c.nrows, NC16, c.col0, c.gw,
                           (const unsigned*)c.amax, c.post, c.out, c.W);
    LAUNCHCHECK();
}

size_t hio_mfma_bs_bytes(int nblk, int gw) { return (size_t)nblk * (gw <= 16 ? 2 : 4) * 128 * 16; }

// Reduction splits: a launch is gx x nsplit work-groups of equal work
// (bps blocks each), resident `slots` at a time (two 512-thread work-groups
// per CU at 4 waves per SIMD), so it runs ceil(gx*nsplit/slots) rounds of
// bps blocks.  Pick the split count minimising rounds * bps plus the cost of
// writing and re-reading one more set of partial sums (~1.3 blocks of work
// per split).  A count just above a multiple of the slots leaves a nearly
// empty last round: 94 x 11 = 1034 work-groups on 512 slots take three
// rounds of 276 blocks where 94 x 16 takes three of 190.
void hio_mfma_split(int nrows, int nblk, int slots, int* nsplit, int* bps) {
    const int gx = (nrows + 255) / 256;
    slots = std::max(slots, 1);
    const int nsmax = std::max(1, std::min(64, nblk / 8));
    int best = 1;
    double best_cost = 1e300;
    for (int ns = 1; ns <= nsmax; ++ns) {
        const int b = (nblk + ns - 1) / ns;
        const int nsu = (nblk + b - 1) / b;   // splits actually holding blocks
        const int64_t wgs = (int64_t)gx * nsu;
        const double cost = (double)((wgs + slots - 1) / slots) * b + 1.3 * nsu;
        if (cost < best_cost - 1e-9) {
            best_cost = cost;
            best = ns;
        }
    }
    const int b = (nblk + best - 1) / best;
    *nsplit = (nblk + b - 1) / b;
    *bps = b;
}

}  // namespace mm

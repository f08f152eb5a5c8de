// Learned warm-start network on MFMA (v_mfma_f32_16x16x32_bf16).
//
// Convolution as implicit GEMM, D[cout][pixel] = sum_k W[cout][k] X[k][pixel]:
//   * a workgroup (4 waves) owns a 16x16 output tile; its 18x18 input halo is
//     staged in LDS once (32-channel pixels as four 16-byte chunks, chunk
//     index XOR-swizzled by pixel so the 16 lanes of a read hit distinct
//     banks);
//   * each wave owns 4 rows of 16 pixels; per 3x3 tap the K = 32 input
//     channels of one pixel are one MFMA B operand (one ds_read_b128);
//   * the packed weights (A operands, lane order of the instruction) stay in
//     VGPRs for the whole tile: 9 taps x 2 cout blocks x 4 VGPRs;
//   * 2-channel inputs (the network input, and the last layer's gradient in
//     the backward pass) pack K = 9 taps x 2 channels = 18 (padded to 32)
//     into a single MFMA.
// The same kernel runs the backward data pass: a forward convolution of the
// output gradient with channel-transposed, tap-flipped weights whose
// epilogue applies the ReLU mask of the layer input.
//
// Weight gradients, dW[cout][cin][tap] = sum_pixel d[pixel][cout] *
// a[pixel + tap][cin], run a persistent kernel over 16x16 tiles that stages
// the gradient tile and the activation halo channel-major in LDS (so the K
// dimension -- pixels -- is contiguous per lane) and builds the x-shifted
// operands of the three horizontal taps from one aligned 16-byte read with
// two funnel shifts.  Each workgroup writes its partial sums; a fixed-order
// reduction over workgroups gives a deterministic gradient.
#include <cmath>

#include "wsnet.hpp"

namespace mm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kChan[kWsLayers + 1] = {2, 32, 32, 32, 32, 32, 2};
constexpr int kHalo = kWsTile + 2;           // 18
constexpr int kDtStride = 256 + 8;           // gradient tile row (elements), 528 bytes
constexpr int kAtRow = 32;                   // activation halo row: slots 7..24 used
constexpr int kAtStride = kHalo * kAtRow + 8;  // per channel, 1168 bytes
constexpr int kMaxDwBlocks = 256;

__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ bool bf_pos(unsigned h16) { return (h16 & 0x8000u) == 0 && (h16 & 0x7FFFu) != 0; }

__device__ __forceinline__ floatx4 mfma_bf16(uint4 a, uint4 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// input scaling and packing

// one block per image: s = sqrt(mean |x|^2) (FP64 sum), 1 for an all-zero image
__global__ __launch_bounds__(256) void ws_scale_k(const float2* __restrict__ x, int64_t V,
                                                  float* __restrict__ sc, float* __restrict__ isc) {
    const float2* xb = x + (int64_t)blockIdx.x * V;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < V; i += 256) {
        const float2 v = xb[i];
        s += (double)v.x * v.x + (double)v.y * v.y;
    }
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float r = (float)sqrt(red[0] / (double)V);
        if (!(r > 0.f)) r = 1.f;
        sc[blockIdx.x] = r;
        isc[blockIdx.x] = 1.f / r;
    }
}

__global__ void ws_pack_input_k(const float2* __restrict__ x, const float* __restrict__ isc, int64_t V,
                                int64_t n, unsigned* __restrict__ a0) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float f = isc[i / V];
        const float2 v = x[i];
        a0[i] = (unsigned)f2bf(v.x * f) | ((unsigned)f2bf(v.y * f) << 16);
    }
}

// ---------------------------------------------------------------------------
// convolution (forward, and backward data)

enum { kModeHidden = 0, kModeOut = 1, kModeLoss = 2, kModeBackData = 3 };

struct WsConvArgs {
    const void* in;       // NHWC bf16, 32 channels (16-byte chunks) or 2 channels (one dword)
    const uint4* wpk;     // packed A operands
    const float* bias;    // 32 floats (bf16-rounded)
    void* out;            // hidden / back-data: NHWC32 bf16; out: float2 image; loss: 2-ch bf16 gradient
    const uint2* mask;    // back-data: NHWC32 bf16 activation whose positivity masks the result
    const float2* x;      // out / loss: network input image
    const float2* t;      // loss: target image
    const float* sc;      // out: per-image scale s
    const float* isc;     // loss: per-image 1/s
    double* lpart;        // loss: sum of squared residuals per workgroup
    float gscale;         // loss: 2 / (number of output elements)
    int H, W, tilesX;
};

template <bool CIN2, int MB, int MODE>
__global__ __launch_bounds__(256) void ws_conv_k(WsConvArgs a) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int c = lane & 15, g = lane >> 4;
    const int b = blockIdx.y;
    const int ty0 = (blockIdx.x / a.tilesX) * kWsTile, tx0 = (blockIdx.x % a.tilesX) * kWsTile;
    const int H = a.H, W = a.W;
    const int64_t V = (int64_t)H * W;
    constexpr int NS32 = CIN2 ? 1 : kHalo * kHalo * 4;
    constexpr int NS2 = CIN2 ? kHalo * kHalo : 1;
    __shared__ uint4 s32[NS32];
    __shared__ unsigned s2[NS2];

    if constexpr (!CIN2) {
        const uint4* in = (const uint4*)a.in + (int64_t)b * V * 4;
        for (int idx = tid; idx < kHalo * kHalo * 4; idx += 256) {
            const int q = idx >> 2, gg = idx & 3;
            const int hy = q / kHalo, hx = q - hy * kHalo;
            const int y = ty0 + hy - 1, x = tx0 + hx - 1;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (y >= 0 && y < H && x >= 0 && x < W) v = in[((int64_t)y * W + x) * 4 + gg];
            s32[q * 4 + (gg ^ ((q >> 2) & 3))] = v;
        }
    } else {
        const unsigned* in = (const unsigned*)a.in + (int64_t)b * V;
        for (int q = tid; q < kHalo * kHalo; q += 256) {
            const int hy = q / kHalo, hx = q - hy * kHalo;
            const int y = ty0 + hy - 1, x = tx0 + hx - 1;
            unsigned v = 0u;
            if (y >= 0 && y < H && x >= 0 && x < W) v = in[(int64_t)y * W + x];
            s2[q] = v;
        }
    }
    constexpr int NT = CIN2 ? 1 : 9;
    uint4 wa[NT][MB];
#pragma unroll
    for (int tp = 0; tp < NT; ++tp)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) wa[tp][mb] = a.wpk[(tp * MB + mb) * 64 + lane];
    __syncthreads();

    floatx4 acc[4][MB];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) acc[r][mb] = floatx4{0.f, 0.f, 0.f, 0.f};

    if constexpr (!CIN2) {
#pragma unroll
        for (int tp = 0; tp < 9; ++tp) {
            const int ky = tp / 3, kx = tp % 3;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int q = (4 * w + r + ky) * kHalo + c + kx;
                const uint4 bv = s32[q * 4 + (g ^ ((q >> 2) & 3))];
#pragma unroll
                for (int mb = 0; mb < MB; ++mb) acc[r][mb] = mfma_bf16(wa[tp][mb], bv, acc[r][mb]);
            }
        }
    } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            unsigned d[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int tp = 4 * g + j;
                d[j] = tp < 9 ? s2[(4 * w + r + tp / 3) * kHalo + c + tp % 3] : 0u;
            }
            const uint4 bv = make_uint4(d[0], d[1], d[2], d[3]);
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) acc[r][mb] = mfma_bf16(wa[0][mb], bv, acc[r][mb]);
        }
    }

    // D lane layout: rows (output channels) 4g+i, column (pixel) c
    double lsum = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int y = ty0 + 4 * w + r, x = tx0 + c;
        const bool inb = y < H && x < W;
        const int64_t p = (int64_t)b * V + (int64_t)y * W + x;
        if constexpr (MODE == kModeHidden || MODE == kModeBackData) {
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) {
                const int co0 = 16 * mb + 4 * g;
                unsigned short h[4];
                if constexpr (MODE == kModeHidden) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) h[i] = f2bf(fmaxf(acc[r][mb][i] + a.bias[co0 + i], 0.f));
                } else {
                    uint2 mk = make_uint2(0u, 0u);
                    if (inb) mk = a.mask[p * 8 + (co0 >> 2)];
                    const unsigned mv[4] = {mk.x & 0xFFFFu, mk.x >> 16, mk.y & 0xFFFFu, mk.y >> 16};
#pragma unroll
                    for (int i = 0; i < 4; ++i) h[i] = bf_pos(mv[i]) ? f2bf(acc[r][mb][i]) : (unsigned short)0;
                }
                if (inb)
                    ((uint2*)a.out)[p * 8 + (co0 >> 2)] =
                        make_uint2((unsigned)h[0] | ((unsigned)h[1] << 16), (unsigned)h[2] | ((unsigned)h[3] << 16));
            }
        } else {
            if (g == 0 && inb) {
                const float z0 = acc[r][0][0] + a.bias[0], z1 = acc[r][0][1] + a.bias[1];
                const float2 xv = a.x[p];
                if constexpr (MODE == kModeOut) {
                    const float s = a.sc[b];
                    ((float2*)a.out)[p] = make_float2(xv.x + s * z0, xv.y + s * z1);
                } else {
                    const float is = a.isc[b];
                    const float2 tv = a.t[p];
                    const float d0 = xv.x * is + z0 - tv.x * is;
                    const float d1 = xv.y * is + z1 - tv.y * is;
                    lsum += (double)d0 * d0 + (double)d1 * d1;
                    ((unsigned*)a.out)[p] = (unsigned)f2bf(a.gscale * d0) | ((unsigned)f2bf(a.gscale * d1) << 16);
                }
            }
        }
    }
    if constexpr (MODE == kModeLoss) {
        __shared__ double red[256];
        red[tid] = lsum;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (tid < o) red[tid] += red[tid + o];
            __syncthreads();
        }
        if (tid == 0) a.lpart[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
    }
}

// ---------------------------------------------------------------------------
// weight gradients

struct WsDwArgs {
    const void* d;    // output gradient: NHWC32 bf16 (16-byte chunks), or 2 channels (dword) if COUT == 2
    const void* act;  // layer input: NHWC32 bf16, or 2 channels if CIN2
    float* part;      // [gridDim.x][np] partial sums
    int64_t np;
    int woff, boff;   // flat offsets of the layer's weights and biases
    int H, W, tilesX, tiles_per_img;
    int64_t ntiles;
};

// waves: COUT 32 -> (cout block, cin block) = (w>>1, w&1); COUT 2 -> waves 0, 1
// take cin blocks 0, 1 (waves 2, 3 only stage).  CIN2: the 18 (tap, channel)
// columns are cin blocks 0 and 1 of one tap-less tile.
template <bool CIN2, int COUT>
__global__ __launch_bounds__(256) void ws_dw_k(WsDwArgs a) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int c = lane & 15, g = lane >> 4;
    const int H = a.H, W = a.W;
    const int64_t V = (int64_t)H * W;
    const int mb = COUT == 32 ? (w >> 1) : 0;
    const int nb = w & 1;
    const bool active = COUT == 32 || w < 2;
    const bool do_bias = nb == 0;
    constexpr int NTAP = CIN2 ? 1 : 9;

    __shared__ __attribute__((aligned(16))) unsigned short dT[kWsCh * kDtStride];
    constexpr int NAT = CIN2 ? 1 : kWsCh * kAtStride;
    constexpr int NA2 = CIN2 ? kHalo * kHalo : 1;
    __shared__ __attribute__((aligned(16))) unsigned short aT[NAT];
    __shared__ unsigned a2[NA2];

    floatx4 acc[NTAP + 1];
#pragma unroll
    for (int j = 0; j <= NTAP; ++j) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
    const uint4 ones = make_uint4(0x3F803F80u, 0x3F803F80u, 0x3F803F80u, 0x3F803F80u);

    for (int64_t t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
        const int64_t b = t / a.tiles_per_img;
        const int tt = (int)(t - b * a.tiles_per_img);
        const int ty0 = (tt / a.tilesX) * kWsTile, tx0 = (tt % a.tilesX) * kWsTile;
        __syncthreads();  // previous tile's reads are done
        if constexpr (COUT == 32) {
            const uint4* d = (const uint4*)a.d + b * V * 4;
            for (int idx = tid; idx < 256 * 4; idx += 256) {
                const int p = idx >> 2, gg = idx & 3;
                const int y = ty0 + (p >> 4), x = tx0 + (p & 15);
                uint4 v = make_uint4(0u, 0u, 0u, 0u);
                if (y < H && x < W) v = d[((int64_t)y * W + x) * 4 + gg];
                const unsigned vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    dT[(8 * gg + 2 * j) * kDtStride + p] = (unsigned short)(vv[j] & 0xFFFFu);
                    dT[(8 * gg + 2 * j + 1) * kDtStride + p] = (unsigned short)(vv[j] >> 16);
                }
            }
        } else {
            const unsigned* d = (const unsigned*)a.d + b * V;
            for (int p = tid; p < 256; p += 256) {
                const int y = ty0 + (p >> 4), x = tx0 + (p & 15);
                unsigned v = 0u;
                if (y < H && x < W) v = d[(int64_t)y * W + x];
                dT[p] = (unsigned short)(v & 0xFFFFu);
                dT[kDtStride + p] = (unsigned short)(v >> 16);
            }
        }
        if constexpr (!CIN2) {
            const uint4* ac = (const uint4*)a.act + b * V * 4;
            for (int idx = tid; idx < kHalo * kHalo * 4; idx += 256) {
                const int q = idx >> 2, gg = idx & 3;
                const int hy = q / kHalo, hx = q - hy * kHalo;
                const int y = ty0 + hy - 1, x = tx0 + hx - 1;
                uint4 v = make_uint4(0u, 0u, 0u, 0u);
                if (y >= 0 && y < H && x >= 0 && x < W) v = ac[((int64_t)y * W + x) * 4 + gg];
                const unsigned vv[4] = {v.x, v.y, v.z, v.w};
                const int base = hy * kAtRow + hx + 7;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    aT[(8 * gg + 2 * j) * kAtStride + base] = (unsigned short)(vv[j] & 0xFFFFu);
                    aT[(8 * gg + 2 * j + 1) * kAtStride + base] = (unsigned short)(vv[j] >> 16);
                }
            }
        } else {
            const unsigned* ac = (const unsigned*)a.act + b * V;
            for (int q = tid; q < kHalo * kHalo; q += 256) {
                const int hy = q / kHalo, hx = q - hy * kHalo;
                const int y = ty0 + hy - 1, x = tx0 + hx - 1;
                unsigned v = 0u;
                if (y >= 0 && y < H && x >= 0 && x < W) v = ac[(int64_t)y * W + x];
                a2[q] = v;
            }
        }
        __syncthreads();
        if (active) {
#pragma unroll 1
            for (int s = 0; s < 8; ++s) {
                // A: gradient rows (output channels), K = 32 pixels of rows 2s, 2s+1
                uint4 av = make_uint4(0u, 0u, 0u, 0u);
                const int co = 16 * mb + c;
                if (COUT == 32 || co < COUT) av = *(const uint4*)&dT[co * kDtStride + 32 * s + 8 * g];
                const int row = 2 * s + (g >> 1), col0 = 8 * (g & 1);
                if constexpr (!CIN2) {
#pragma unroll
                    for (int ky = 0; ky < 3; ++ky) {
                        const int base = (16 * nb + c) * kAtStride + (row + ky) * kAtRow + col0 + 8;
                        const uint4 D = *(const uint4*)&aT[base];
                        const unsigned em1 = (unsigned)aT[base - 1] << 16, e8 = aT[base + 8];
                        const uint4 Bm = make_uint4(__builtin_amdgcn_alignbit(D.x, em1, 16),
                                                    __builtin_amdgcn_alignbit(D.y, D.x, 16),
                                                    __builtin_amdgcn_alignbit(D.z, D.y, 16),
                                                    __builtin_amdgcn_alignbit(D.w, D.z, 16));
                        const uint4 Bp = make_uint4(__builtin_amdgcn_alignbit(D.y, D.x, 16),
                                                    __builtin_amdgcn_alignbit(D.z, D.y, 16),
                                                    __builtin_amdgcn_alignbit(D.w, D.z, 16),
                                                    __builtin_amdgcn_alignbit(e8, D.w, 16));
                        acc[3 * ky + 0] = mfma_bf16(av, Bm, acc[3 * ky + 0]);
                        acc[3 * ky + 1] = mfma_bf16(av, D, acc[3 * ky + 1]);
                        acc[3 * ky + 2] = mfma_bf16(av, Bp, acc[3 * ky + 2]);
                    }
                } else {
                    const int n = 16 * nb + c;  // column = 2*tap + channel
                    unsigned e[8];
#pragma unroll
                    for (int i = 0; i < 8; ++i) e[i] = 0u;
                    if (n < 18) {
                        const int tp = n >> 1, ch = n & 1;
                        const int ky = tp / 3, kx = tp % 3;
#pragma unroll
                        for (int i = 0; i < 8; ++i) {
                            const unsigned v = a2[(row + ky) * kHalo + col0 + i + kx];
                            e[i] = ch ? (v >> 16) : (v & 0xFFFFu);
                        }
                    }
                    const uint4 bv = make_uint4(e[0] | (e[1] << 16), e[2] | (e[3] << 16), e[4] | (e[5] << 16),
                                                e[6] | (e[7] << 16));
                    acc[0] = mfma_bf16(av, bv, acc[0]);
                }
                if (do_bias) acc[NTAP] = mfma_bf16(av, ones, acc[NTAP]);
            }
        }
    }

    if (!active) return;
    float* part = a.part + (int64_t)blockIdx.x * a.np;
    constexpr int CIN = CIN2 ? 2 : 32;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int co = 16 * mb + 4 * g + i;
        if (co >= COUT) continue;
        if constexpr (!CIN2) {
            const int ci = 16 * nb + c;
#pragma unroll
            for (int tp = 0; tp < 9; ++tp) part[a.woff + (co * CIN + ci) * 9 + tp] = acc[tp][i];
        } else {
            const int n = 16 * nb + c;
            if (n < 18) part[a.woff + (co * CIN + (n & 1)) * 9 + (n >> 1)] = acc[0][i];
        }
        if (do_bias && c == 0) part[a.boff + co] = acc[NTAP][i];
    }
}

// grad[j] = sum over workgroups of part[blk][j], fixed order, FP64
__global__ void ws_reduce_k(const float* __restrict__ part, int nblk, int64_t np, float* __restrict__ grad) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < np; j += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int b = 0; b < nblk; ++b) s += part[(int64_t)b * np + j];
        grad[j] = (float)s;
    }
}

__global__ __launch_bounds__(256) void ws_loss_k(const double* __restrict__ lpart, int64_t n, double inv_count,
                                                 double* __restrict__ loss) {
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) s += lpart[i];
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[0] = red[0] * inv_count;
}

// Adam (Kingma & Ba 2015); c1 = 1/(1-beta1^t), c2 = 1/(1-beta2^t)
__global__ void ws_adam_k(float* __restrict__ th, float* __restrict__ m, float* __restrict__ v,
                          const float* __restrict__ g, int64_t n, float lr, float c1, float c2) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        const float gj = g[j];
        const float mj = 0.9f * m[j] + 0.1f * gj;
        const float vj = 0.999f * v[j] + 0.001f * gj * gj;
        m[j] = mj;
        v[j] = vj;
        th[j] -= lr * (mj * c1) / (sqrtf(vj * c2) + 1e-8f);
    }
}

struct WsPackArgs {
    int boff[kWsLayers];
    int cout[kWsLayers];
};

__global__ void ws_pack_k(const float* __restrict__ th, const int* __restrict__ idx, int64_t n,
                          unsigned short* __restrict__ packed, float* __restrict__ bias, WsPackArgs pa) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n + kWsLayers * kWsCh;
         e += (int64_t)gridDim.x * blockDim.x) {
        if (e < n) {
            const int j = idx[e];
            packed[e] = j >= 0 ? f2bf(th[j]) : (unsigned short)0;
        } else {
            const int k = (int)(e - n), l = k / kWsCh, co = k % kWsCh;
            bias[k] = co < pa.cout[l] ? __uint_as_float((unsigned)f2bf(th[pa.boff[l] + co]) << 16) : 0.f;
        }
    }
}

static int grid1d(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096)); }

}  // namespace mm

using namespace mm;

mm_wsnet::mm_wsnet(mm_ctx* c, int64_t N1, int64_t N2) : ctx(c), N1_(N1), N2_(N2) {
    if (N1 < 1 || N2 < 1) fail(MM_ERR_ARG, "warm-start net: empty image");
    if (N1 > (1 << 15) || N2 > (1 << 15)) fail(MM_ERR_ARG, "warm-start net: image too large");
    V_ = N1 * N2;
    tilesX_ = (int)((N1 + kWsTile - 1) / kWsTile);
    tilesY_ = (int)((N2 + kWsTile - 1) / kWsTile);
    int64_t o = 0;
    for (int l = 0; l < kWsLayers; ++l) {
        woff_[l] = (int)o;
        o += (int64_t)kChan[l + 1] * kChan[l] * 9;
        boff_[l] = (int)o;
        o += kChan[l + 1];
    }
    np_ = o;
    // packed operand blocks (bf16 elements); 8 per lane, 64 lanes per block
    int64_t p = 0;
    pk_f0_ = p; p += 2 * 512;
    for (int j = 0; j < 4; ++j) { pk_f_[j] = p; p += 9 * 2 * 512; }
    pk_f5_ = p; p += 9 * 1 * 512;
    for (int j = 0; j < 4; ++j) { pk_t_[j] = p; p += 9 * 2 * 512; }
    pk_t5_ = p; p += 2 * 512;
    pk_total_ = p;
    std::vector<int> idx((size_t)pk_total_, -1);
    // 2-channel-input form: column k = 2*tap + channel (k < 18)
    auto cin2 = [&](int64_t off, bool transposed) {
        for (int mb = 0; mb < 2; ++mb)
            for (int ln = 0; ln < 64; ++ln)
                for (int i = 0; i < 8; ++i) {
                    const int row = 16 * mb + (ln & 15), k = 8 * (ln >> 4) + i;
                    const int tp = k >> 1, ch = k & 1;
                    if (tp >= 9) continue;
                    const int ky = tp / 3, kx = tp % 3;
                    const int64_t e = off + (mb * 64 + ln) * 8 + i;
                    if (!transposed)  // layer 0: W[row][ch][ky][kx]
                        idx[e] = woff_[0] + ((row * 2 + ch) * 3 + ky) * 3 + kx;
                    else  // layer 5, transposed and flipped: W[ch][row][2-ky][2-kx]
                        idx[e] = woff_[5] + ((ch * 32 + row) * 3 + (2 - ky)) * 3 + (2 - kx);
                }
    };
    // 32-channel-input form: [tap][cout block][lane][8]
    auto c32 = [&](int64_t off, int layer, int MB, bool transposed) {
        const int cout = kChan[layer + 1];
        for (int tp = 0; tp < 9; ++tp)
            for (int mb = 0; mb < MB; ++mb)
                for (int ln = 0; ln < 64; ++ln)
                    for (int i = 0; i < 8; ++i) {
                        const int row = 16 * mb + (ln & 15), col = 8 * (ln >> 4) + i;
                        const int ky = tp / 3, kx = tp % 3;
                        const int64_t e = off + ((tp * MB + mb) * 64 + ln) * 8 + i;
                        if (!transposed) {
                            if (row >= cout) continue;
                            idx[e] = woff_[layer] + ((row * 32 + col) * 3 + ky) * 3 + kx;
                        } else {
                            idx[e] = woff_[layer] + ((col * 32 + row) * 3 + (2 - ky)) * 3 + (2 - kx);
                        }
                    }
    };
    cin2(pk_f0_, false);
    for (int j = 0; j < 4; ++j) c32(pk_f_[j], j + 1, 2, false);
    c32(pk_f5_, 5, 1, false);
    for (int j = 0; j < 4; ++j) c32(pk_t_[j], j + 1, 2, true);
    cin2(pk_t5_, true);
    pidx_.upload(idx, ctx->stream);
    theta_.alloc((size_t)np_);
    m_.alloc((size_t)np_);
    v_.alloc((size_t)np_);
    grad_.alloc((size_t)np_);
    bias_.alloc(kWsLayers * kWsCh);
    packed_.alloc((size_t)pk_total_);
    loss_.alloc(1);
    HIPCHECK(hipMemsetAsync(theta_.p, 0, sizeof(float) * np_, ctx->stream));
    HIPCHECK(hipMemsetAsync(m_.p, 0, sizeof(float) * np_, ctx->stream));
    HIPCHECK(hipMemsetAsync(v_.p, 0, sizeof(float) * np_, ctx->stream));
    pack();
    HIPCHECK(hipStreamSynchronize(ctx->stream));
}

void mm_wsnet::pack() {
    WsPackArgs pa;
    for (int l = 0; l < kWsLayers; ++l) {
        pa.boff[l] = boff_[l];
        pa.cout[l] = kChan[l + 1];
    }
    hipLaunchKernelGGL(ws_pack_k, dim3(grid1d(pk_total_ + kWsLayers * kWsCh)), dim3(256), 0, ctx->stream,
                       theta_.p, pidx_.p, pk_total_, packed_.p, bias_.p, pa);
    LAUNCHCHECK();
}

void mm_wsnet::set_params(const float* th, bool on_device) {
    HIPCHECK(hipMemcpyAsync(theta_.p, th, sizeof(float) * np_,
                            on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(hipMemsetAsync(m_.p, 0, sizeof(float) * np_, ctx->stream));
    HIPCHECK(hipMemsetAsync(v_.p, 0, sizeof(float) * np_, ctx->stream));
    adam_t_ = 0;
    pack();
    if (!on_device) HIPCHECK(hipStreamSynchronize(ctx->stream));
}

void mm_wsnet::get_params(float* th, bool on_device) {
    HIPCHECK(hipMemcpyAsync(th, theta_.p, sizeof(float) * np_,
                            on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx->stream));
    if (!on_device) HIPCHECK(hipStreamSynchronize(ctx->stream));
}

void mm_wsnet::ensure_batch(int64_t B, bool training) {
    if (B <= Bcap_ && (!training || train_cap_)) return;
    const int64_t Bn = std::max(B, Bcap_);
    const bool tr = training || train_cap_;
    const size_t n = (size_t)(Bn * V_);
    sc_.alloc((size_t)Bn);
    isc_.alloc((size_t)Bn);
    a0_.alloc(n);
    act_[0].alloc(n * kWsCh);
    act_[1].alloc(n * kWsCh);
    if (tr) {
        for (int k = 2; k < 5; ++k) act_[k].alloc(n * kWsCh);
        dbuf_[0].alloc(n * kWsCh);
        dbuf_[1].alloc(n * kWsCh);
        g6_.alloc(n);
        lpart_.alloc((size_t)(Bn * tilesX_ * tilesY_));
        const int64_t ntiles = Bn * tilesX_ * tilesY_;
        part_.alloc((size_t)(std::min<int64_t>(ntiles, kMaxDwBlocks) * np_));
    }
    Bcap_ = Bn;
    train_cap_ = tr;
}

void mm_wsnet::scale_and_pack_input(const float2* x, int64_t B) {
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(ws_scale_k, dim3((unsigned)B), dim3(256), 0, st, x, V_, sc_.p, isc_.p);
    LAUNCHCHECK();
    hipLaunchKernelGGL(ws_pack_input_k, dim3(grid1d(B * V_)), dim3(256), 0, st, x, isc_.p, V_, B * V_, a0_.p);
    LAUNCHCHECK();
}

// layers 0..4 (conv + bias + ReLU); training keeps every activation
void mm_wsnet::run_hidden(int64_t B, bool training) {
    hipStream_t st = ctx->stream;
    const dim3 grid((unsigned)(tilesX_ * tilesY_), (unsigned)B);
    WsConvArgs a{};
    a.H = (int)N2_;
    a.W = (int)N1_;
    a.tilesX = tilesX_;
    a.in = a0_.p;
    a.wpk = (const uint4*)(packed_.p + pk_f0_);
    a.bias = bias_.p;
    a.out = act_[0].p;
    hipLaunchKernelGGL((ws_conv_k<true, 2, kModeHidden>), grid, dim3(256), 0, st, a);
    LAUNCHCHECK();
    for (int l = 1; l < 5; ++l) {
        const int src = training ? l - 1 : (l - 1) & 1;
        const int dst = training ? l : l & 1;
        a.in = act_[src].p;
        a.wpk = (const uint4*)(packed_.p + pk_f_[l - 1]);
        a.bias = bias_.p + l * kWsCh;
        a.out = act_[dst].p;
        hipLaunchKernelGGL((ws_conv_k<false, 2, kModeHidden>), grid, dim3(256), 0, st, a);
        LAUNCHCHECK();
    }
}

void mm_wsnet::forward(const float2* x, float2* y, int64_t B) {
    if (B < 1) fail(MM_ERR_ARG, "warm-start net: batch must be >= 1");
    if (B > 65535) fail(MM_ERR_ARG, "warm-start net: batch must be <= 65535");
    ensure_batch(B, false);
    scale_and_pack_input(x, B);
    run_hidden(B, false);
    WsConvArgs a{};
    a.H = (int)N2_;
    a.W = (int)N1_;
    a.tilesX = tilesX_;
    a.in = act_[0].p;  // a5 (layer 4 output) lands in buffer 0 when ping-ponging
    a.wpk = (const uint4*)(packed_.p + pk_f5_);
    a.bias = bias_.p + 5 * kWsCh;
    a.out = y;
    a.x = x;
    a.sc = sc_.p;
    hipLaunchKernelGGL((ws_conv_k<false, 1, kModeOut>), dim3((unsigned)(tilesX_ * tilesY_), (unsigned)B),
                       dim3(256), 0, ctx->stream, a);
    LAUNCHCHECK();
}

double mm_wsnet::loss_grad(const float2* x, const float2* t, int64_t B, bool want_loss) {
    if (B < 1) fail(MM_ERR_ARG, "warm-start net: batch must be >= 1");
    if (B > 65535) fail(MM_ERR_ARG, "warm-start net: batch must be <= 65535");
    ensure_batch(B, true);
    hipStream_t st = ctx->stream;
    scale_and_pack_input(x, B);
    run_hidden(B, true);
    const dim3 grid((unsigned)(tilesX_ * tilesY_), (unsigned)B);
    const double count = 2.0 * (double)B * (double)V_;
    WsConvArgs a{};
    a.H = (int)N2_;
    a.W = (int)N1_;
    a.tilesX = tilesX_;
    a.in = act_[4].p;
    a.wpk = (const uint4*)(packed_.p + pk_f5_);
    a.bias = bias_.p + 5 * kWsCh;
    a.out = g6_.p;
    a.x = x;
    a.t = t;
    a.isc = isc_.p;
    a.lpart = lpart_.p;
    a.gscale = (float)(2.0 / count);
    hipLaunchKernelGGL((ws_conv_k<false, 1, kModeLoss>), grid, dim3(256), 0, st, a);
    LAUNCHCHECK();

    const int64_t ntiles = B * tilesX_ * tilesY_;
    nblk_dw_ = (int)std::min<int64_t>(ntiles, kMaxDwBlocks);
    WsDwArgs d{};
    d.part = part_.p;
    d.np = np_;
    d.H = (int)N2_;
    d.W = (int)N1_;
    d.tilesX = tilesX_;
    d.tiles_per_img = tilesX_ * tilesY_;
    d.ntiles = ntiles;
    // layer 5: gradient g6 (2 channels), input a5
    d.d = g6_.p;
    d.act = act_[4].p;
    d.woff = woff_[5];
    d.boff = boff_[5];
    hipLaunchKernelGGL((ws_dw_k<false, 2>), dim3(nblk_dw_), dim3(256), 0, st, d);
    LAUNCHCHECK();
    WsConvArgs bd{};
    bd.H = (int)N2_;
    bd.W = (int)N1_;
    bd.tilesX = tilesX_;
    bd.in = g6_.p;
    bd.wpk = (const uint4*)(packed_.p + pk_t5_);
    bd.mask = (const uint2*)act_[4].p;
    bd.out = dbuf_[0].p;
    hipLaunchKernelGGL((ws_conv_k<true, 2, kModeBackData>), grid, dim3(256), 0, st, bd);
    LAUNCHCHECK();
    int cur = 0;
    for (int l = 4; l >= 1; --l) {
        // dW of layer l from delta_l (dbuf[cur]) and its input a_l = act_[l-1]
        d.d = dbuf_[cur].p;
        d.act = act_[l - 1].p;
        d.woff = woff_[l];
        d.boff = boff_[l];
        hipLaunchKernelGGL((ws_dw_k<false, 32>), dim3(nblk_dw_), dim3(256), 0, st, d);
        LAUNCHCHECK();
        bd.in = dbuf_[cur].p;
        bd.wpk = (const uint4*)(packed_.p + pk_t_[l - 1]);
        bd.mask = (const uint2*)act_[l - 1].p;
        bd.out = dbuf_[cur ^ 1].p;
        hipLaunchKernelGGL((ws_conv_k<false, 2, kModeBackData>), grid, dim3(256), 0, st, bd);
        LAUNCHCHECK();
        cur ^= 1;
    }
    d.d = dbuf_[cur].p;
    d.act = a0_.p;
    d.woff = woff_[0];
    d.boff = boff_[0];
    hipLaunchKernelGGL((ws_dw_k<true, 32>), dim3(nblk_dw_), dim3(256), 0, st, d);
    LAUNCHCHECK();
    hipLaunchKernelGGL(ws_reduce_k, dim3(grid1d(np_)), dim3(256), 0, st, part_.p, nblk_dw_, np_, grad_.p);
    LAUNCHCHECK();
    hipLaunchKernelGGL(ws_loss_k, dim3(1), dim3(256), 0, st, lpart_.p, ntiles, 1.0 / count, loss_.p);
    LAUNCHCHECK();
    if (!want_loss) return std::nan("");
    double L = 0;
    HIPCHECK(hipMemcpyAsync(&L, loss_.p, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    return L;
}

void mm_wsnet::adam_step(float lr) {
    ++adam_t_;
    const double c1 = 1.0 / (1.0 - std::pow(0.9, (double)adam_t_));
    const double c2 = 1.0 / (1.0 - std::pow(0.999, (double)adam_t_));
    hipLaunchKernelGGL(ws_adam_k, dim3(grid1d(np_)), dim3(256), 0, ctx->stream, theta_.p, m_.p, v_.p, grad_.p, np_,
                       lr, (float)c1, (float)c2);
    LAUNCHCHECK();
    pack();
}

// Cartesian sampling of mrSampFunc (operators/mrSampFunc.m:108-114, :137-143):
// when samplingOp is a k-space mask rather than an operator,
//   forward  y = mask .* ifftnc(x, nDim)
//   adjoint  x = fftnc(mask .* y, nDim)
// with the centred unitary transforms of general/ifftnc.m and general/fftnc.m
// (ifftshift / fftshift around each 1-D transform, scale 1/sqrt(n) each).
//
// One launch per dimension: a work-group stages T interleaved 1-D transforms
// in LDS (the shift folded into the load and store indices), runs the
// Stockham passes of fft.hpp and writes them back; the mask (and the
// 1/sqrt(n) scale) is applied on the load of the adjoint's first dimension
// or on the store of the forward's last.
#include <algorithm>
#include <cmath>

#include "cart.hpp"

namespace mm {

// Pass over the transforms along one dimension.  Transform j starts at
// (j % inner) + (j / inner) * inner * n and steps by `inner` (inner = 1 for
// dimension 1, N1 for dimension 2).  Slot q of transform j is read from
// position (q + sh_in) mod n and slot k is written to (k + sh_out) mod n.
// mask_mode 1 multiplies the loaded values by mask[e mod V], 2 the stored.
template <int T, int SGN, bool BIG>
__global__ __launch_bounds__(kFftThreads) void cart_pass(const float2* in, float2* out,
                                                         int n, int64_t inner, int64_t ntrans,
                                                         const FftFactors* __restrict__ fac,
                                                         const float2* __restrict__ twg, int sh_in, int sh_out,
                                                         const float* __restrict__ mask, int64_t V, int mask_mode,
                                                         float scale) {
    extern __shared__ float2 lds[];
    float2* A = lds;
    float2* Bb = lds + (size_t)n * T;
    float2* tw = lds + (size_t)2 * n * T;
    const FftFactors F = *fac;
    const int64_t j0 = (int64_t)blockIdx.x * T;
    const int total = n * T;
    for (int idx = threadIdx.x; idx < total; idx += kFftThreads) {
        const int q = idx / T, t = idx % T;
        const int64_t j = min(j0 + t, ntrans - 1);
        const int64_t base = (j % inner) + (j / inner) * inner * n;
        int pos = q + sh_in;
        if (pos >= n) pos -= n;
        const int64_t e = base + (int64_t)pos * inner;
        float2 v = in[e];
        if (mask_mode == 1) {
            const float m = mask[e % V];
            v = make_float2(v.x * m, v.y * m);
        }
        A[idx] = v;
    }
    for (int i = threadIdx.x; i < n - 1; i += kFftThreads) tw[i] = twg[i];
    __syncthreads();
    const float2* R = fft_lds<T, SGN, BIG>(A, Bb, F, tw);
    for (int idx = threadIdx.x; idx < total; idx += kFftThreads) {
        const int k = idx / T, t = idx % T;
        const int64_t j = j0 + t;
        if (j >= ntrans) continue;
        const int64_t base = (j % inner) + (j / inner) * inner * n;
        int pos = k + sh_out;
        if (pos >= n) pos -= n;
        const int64_t e = base + (int64_t)pos * inner;
        float f = scale;
        if (mask_mode == 2) f *= mask[e % V];
        const float2 v = R[idx];
        out[e] = make_float2(v.x * f, v.y * f);
    }
}

// Lengths with a prime factor above 7 (MATLAB's fft takes any length): a
// direct DFT, one output value per thread, w[m] = exp(sgn 2 pi i m / n)
// precomputed in FP64 and rounded once; same index maps, mask and scale.
template <int SGN>
__global__ __launch_bounds__(256) void cart_dft_pass(const float2* in, float2* out, int n, int64_t inner,
                                                     int64_t ntrans, const float2* __restrict__ w, int sh_in,
                                                     int sh_out, const float* __restrict__ mask, int64_t V,
                                                     int mask_mode, float scale) {
    const int64_t total = ntrans * n;
    for (int64_t idx = blockIdx.x * (int64_t)256 + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
        const int k = (int)(idx % n);
        const int64_t j = idx / n;
        const int64_t base = (j % inner) + (j / inner) * inner * n;
        float2 acc = make_float2(0.f, 0.f);
        int m = 0;   // (q * k) mod n
        for (int q = 0; q < n; ++q) {
            int pos = q + sh_in;
            if (pos >= n) pos -= n;
            const int64_t e = base + (int64_t)pos * inner;
            float2 v = in[e];
            if (mask_mode == 1) {
                const float mk = mask[e % V];
                v = make_float2(v.x * mk, v.y * mk);
            }
            const float2 c = w[m];
            const float cy = SGN > 0 ? c.y : -c.y;
            acc.x += v.x * c.x - v.y * cy;
            acc.y += v.x * cy + v.y * c.x;
            m += k;
            if (m >= n) m -= n;
        }
        int pos = k + sh_out;
        if (pos >= n) pos -= n;
        const int64_t e = base + (int64_t)pos * inner;
        float f = scale;
        if (mask_mode == 2) f *= mask[e % V];
        out[e] = make_float2(acc.x * f, acc.y * f);
    }
}

// transforms per work-group: LDS (two buffers of n*T values + n - 1
// twiddles) within 64 KB
static int cart_T(int n) {
    for (int T : {8, 4, 2, 1})
        if ((size_t)(2 * n * T + n) * sizeof(float2) <= 65536) return T;
    return 0;
}

CartOp::CartOp(mm_ctx* c, int ndim, const int64_t* imN, const float* mask) {
    ctx = c;
    if (ndim != 1 && ndim != 2) fail(MM_ERR_UNSUPPORTED, "mrSampFunc: Cartesian masks of 1 or 2 dimensions only");
    if (!imN || !mask) fail(MM_ERR_ARG, "mrSampFunc: NULL mask");
    ndim_ = ndim;
    n_[0] = imN[0];
    n_[1] = ndim == 2 ? imN[1] : 1;
    if (n_[0] < 1 || n_[1] < 1) fail(MM_ERR_ARG, "mrSampFunc: empty mask");
    V_ = n_[0] * n_[1];
    std::vector<FftFactors> facs;
    std::vector<float2> tw;
    for (int d = 0; d < ndim; ++d) {
        const int n = (int)n_[d];
        if (n_[d] > 65536) fail(MM_ERR_UNSUPPORTED, "mrSampFunc: Cartesian transform length above 65536");
        FftFactors f;
        tw_off_[d] = (int)tw.size();
        direct_[d] = n_[d] > 2730 || !fft_factorize(n, f) || cart_T(n) == 0;
        if (direct_[d]) {
            // direct DFT: the n roots of unity
            f = FftFactors();
            f.n = n;
            for (int m = 0; m < n; ++m) {
                const double a = 2.0 * 3.14159265358979323846 * (double)m / (double)n;
                tw.push_back(make_float2((float)std::cos(a), (float)std::sin(a)));
            }
        } else {
            fft_pass_twiddles(f, [&](float cr, float ci) { tw.push_back(make_float2(cr, ci)); });
        }
        facs.push_back(f);
        tw.push_back(make_float2(0.f, 0.f));   // keeps every slice non-empty
    }
    hipStream_t st = ctx->stream;
    fac_.upload(facs, st);
    tw_.upload(tw, st);
    mask_.upload(mask, (size_t)V_, st);
    HIPCHECK(hipStreamSynchronize(st));
}

template <int T, int SGN>
static void launch_cart(hipStream_t st, const float2* in, float2* out, int n, int64_t inner, int64_t ntrans,
                        const FftFactors* fac, const float2* tw, int sh_in, int sh_out, const float* mask,
                        int64_t V, int mm, float scale) {
    const unsigned grid = (unsigned)((ntrans + T - 1) / T);
    const size_t lds = (size_t)(2 * n * T + n) * sizeof(float2);
    if (n > kFftSmallMax)
        hipLaunchKernelGGL((cart_pass<T, SGN, true>), dim3(grid), dim3(kFftThreads), lds, st, in, out, n, inner,
                           ntrans, fac, tw, sh_in, sh_out, mask, V, mm, scale);
    else
        hipLaunchKernelGGL((cart_pass<T, SGN, false>), dim3(grid), dim3(kFftThreads), lds, st, in, out, n, inner,
                           ntrans, fac, tw, sh_in, sh_out, mask, V, mm, scale);
    LAUNCHCHECK();
}

void CartOp::pass(int d, int adjoint, const float2* in, float2* out, int64_t nrep, int mask_mode) {
    const int n = (int)n_[d];
    const int64_t inner = d == 0 ? 1 : n_[0];
    const int64_t ntrans = V_ / n * nrep;
    const int s = n / 2;   // floor(n/2): the fftshift / ifftshift amount
    // ifftnc: ifftshift before and after (read (q + s), write to (k - s));
    // fftnc: fftshift before and after (read (q - s), write to (k + s))
    const int sh_in = adjoint ? (n - s) % n : s;
    const int sh_out = adjoint ? s : (n - s) % n;
    const float scale = (float)(1.0 / std::sqrt((double)n));
    const FftFactors* fac = fac_.p + d;
    const float2* tw = tw_.p + tw_off_[d];
    hipStream_t st = ctx->stream;
    if (direct_[d]) {
        const int64_t total = ntrans * n;
        const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65536));
        if (adjoint)
            hipLaunchKernelGGL(cart_dft_pass<-1>, dim3(grid), dim3(256), 0, st, in, out, n, inner, ntrans, tw, sh_in,
                               sh_out, mask_.p, V_, mask_mode, scale);
        else
            hipLaunchKernelGGL(cart_dft_pass<1>, dim3(grid), dim3(256), 0, st, in, out, n, inner, ntrans, tw, sh_in,
                               sh_out, mask_.p, V_, mask_mode, scale);
        LAUNCHCHECK();
        return;
    }
    const int T = cart_T(n);
    auto go = [&](auto tconst) {
        constexpr int TT = decltype(tconst)::value;
        if (adjoint)
            launch_cart<TT, -1>(st, in, out, n, inner, ntrans, fac, tw, sh_in, sh_out, mask_.p, V_, mask_mode, scale);
        else
            launch_cart<TT, 1>(st, in, out, n, inner, ntrans, fac, tw, sh_in, sh_out, mask_.p, V_, mask_mode, scale);
    };
    switch (T) {
        case 8: go(std::integral_constant<int, 8>{}); break;
        case 4: go(std::integral_constant<int, 4>{}); break;
        case 2: go(std::integral_constant<int, 2>{}); break;
        default: go(std::integral_constant<int, 1>{}); break;
    }
}

void CartOp::apply(int adjoint, const float2* in, float2* out, int64_t nrep) {
    if (nrep < 1) fail(MM_ERR_ARG, "mrSampFunc: nrep must be >= 1");
    // a staged (FFT) pass reads its transforms whole before writing them, so
    // the second dimension can run in place on the first one's output; a
    // direct-DFT pass reads every value of a transform for each output, so
    // its input goes through a separate buffer
    float2* mid = out;
    if (ndim_ == 2 && direct_[1]) {
        tmp_.ensure((size_t)(V_ * nrep));
        mid = tmp_.p;
    }
    if (!adjoint) {
        // y = mask .* ifftnc(x): the mask on the store of the last dimension
        if (ndim_ == 1) {
            pass(0, 0, in, out, nrep, 2);
        } else {
            pass(0, 0, in, mid, nrep, 0);
            pass(1, 0, mid, out, nrep, 2);
        }
    } else {
        // x = fftnc(mask .* y): the mask on the load of the first dimension
        if (ndim_ == 1) {
            pass(0, 1, in, out, nrep, 1);
        } else {
            pass(0, 1, in, mid, nrep, 1);
            pass(1, 1, mid, out, nrep, 0);
        }
    }
}

}  // namespace mm

namespace mm {

// Sum over dimensions (mrSampFunc.m:100-105) of a column-major array:
// forward out = sum over the flagged dimensions (kept as singletons),
// adjoint = broadcast back along them (MATLAB's implicit expansion in the
// 'transp' branch).  Dimension strides of the full and reduced arrays are
// kernel arguments; one thread per output element.
struct SumDimsArgs {
    int nd;
    int64_t full[kSumMaxDims];     // full extents
    int64_t red[kSumMaxDims];      // reduced extents (1 where summed)
    int64_t fstride[kSumMaxDims];  // strides of the full array
};

__global__ __launch_bounds__(256) void sumdims_fwd_k(const float2* __restrict__ in, float2* __restrict__ out,
                                                     SumDimsArgs a, int64_t nout, int64_t nsum) {
    for (int64_t o = blockIdx.x * (int64_t)256 + threadIdx.x; o < nout; o += (int64_t)gridDim.x * 256) {
        // base offset of this output element in the full array
        int64_t r = o, base = 0;
        for (int d = 0; d < a.nd; ++d) {
            const int64_t i = r % a.red[d];
            r /= a.red[d];
            base += i * a.fstride[d];
        }
        float2 s = make_float2(0.f, 0.f);
        for (int64_t k = 0; k < nsum; ++k) {
            int64_t q = k, off = base;
            for (int d = 0; d < a.nd; ++d) {
                if (a.red[d] != 1 || a.full[d] == 1) continue;
                off += (q % a.full[d]) * a.fstride[d];
                q /= a.full[d];
            }
            const float2 v = in[off];
            s.x += v.x;
            s.y += v.y;
        }
        out[o] = s;
    }
}

__global__ __launch_bounds__(256) void sumdims_adj_k(const float2* __restrict__ in, float2* __restrict__ out,
                                                     SumDimsArgs a, int64_t nfull) {
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < nfull; i += (int64_t)gridDim.x * 256) {
        int64_t r = i, o = 0, ostride = 1;
        for (int d = 0; d < a.nd; ++d) {
            const int64_t c = r % a.full[d];
            r /= a.full[d];
            if (a.red[d] != 1) o += c * ostride;
            ostride *= a.red[d];
        }
        out[i] = in[o];
    }
}

SumDimsOp::SumDimsOp(mm_ctx* c, int nd, const int64_t* dims, const int* sum_flags) {
    ctx = c;
    // one more dimension is kept for the repetitions
    if (nd < 1 || nd > kSumMaxDims - 1 || !dims || !sum_flags)
        fail(MM_ERR_ARG, "mrSampFunc: sumDims over 1.." + std::to_string(kSumMaxDims - 1) + " dimensions");
    nd_ = nd;
    nfull_ = nred_ = nsum_ = 1;
    int64_t st = 1;
    for (int d = 0; d < nd; ++d) {
        if (dims[d] < 1) fail(MM_ERR_ARG, "mrSampFunc: empty dimension");
        full_[d] = dims[d];
        red_[d] = sum_flags[d] ? 1 : dims[d];
        if (sum_flags[d]) nsum_ *= dims[d];
        fstride_[d] = st;
        st *= dims[d];
        nfull_ *= full_[d];
        nred_ *= red_[d];
    }
}

void SumDimsOp::apply(int adjoint, const float2* in, float2* out, int64_t nrep) {
    if (nrep < 1) fail(MM_ERR_ARG, "mrSampFunc: nrep must be >= 1");
    SumDimsArgs a{};
    a.nd = nd_ + 1;   // repetitions as one more (kept) dimension
    for (int d = 0; d < nd_; ++d) {
        a.full[d] = full_[d];
        a.red[d] = red_[d];
        a.fstride[d] = fstride_[d];
    }
    a.full[nd_] = a.red[nd_] = nrep;
    a.fstride[nd_] = nfull_;
    const int64_t n = adjoint ? nfull_ * nrep : nred_ * nrep;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
    if (!adjoint)
        hipLaunchKernelGGL(sumdims_fwd_k, dim3(grid), dim3(256), 0, ctx->stream, in, out, a, n, nsum_);
    else
        hipLaunchKernelGGL(sumdims_adj_k, dim3(grid), dim3(256), 0, ctx->stream, in, out, a, n);
    LAUNCHCHECK();
}

}  // namespace mm

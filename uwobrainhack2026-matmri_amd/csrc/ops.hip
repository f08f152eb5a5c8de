// Coil-sensitivity operator (rcvrOp) and the SENSE composer (mrSampFunc).
//
// rcvrOp maps path, operators/@rcvrOp/rcvrOp.m:145-157:
//   forward  y(:,c) = x .* maps(:,c)
//   adjoint  x = sum_c conj(maps(:,c)) .* y(:,c)
// mrSampFunc, operators/mrSampFunc.m:82-159:  E = S o R,  E' = R' o S' (W
// applied between them by the normal operator as cgne.m:165-170 does).
#include "ops.hpp"

namespace mm {

void mm_op_default_normal(mm_op* op, const float2* in, float2* out, int64_t nrep, float2* tmp);

// Slices of one image size V run in one launch: image j = slice*nrep + rep,
// image j's coil data at j*C*V, slice s's maps at s*C*V.
__global__ __launch_bounds__(256) void rcvr_fwd_k(const float2* __restrict__ x, const float2* __restrict__ maps,
                           float2* __restrict__ y, int64_t V, int C, int64_t nrep, int64_t nimg) {
    const int64_t n = V * C * nimg;
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256) {
        const int64_t v = i % V;
        const int64_t rest = i / V;
        const int c = (int)(rest % C);
        const int64_t j = rest / C;
        y[i] = cmulf(x[j * V + v], maps[((j / nrep) * C + c) * V + v]);
    }
}

__global__ __launch_bounds__(256) void rcvr_adj_k(const float2* __restrict__ y, const float2* __restrict__ maps,
                           float2* __restrict__ x, int64_t V, int C, int64_t nrep, int64_t nimg) {
    const int64_t n = V * nimg;
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256) {
        const int64_t v = i % V;
        const int64_t j = i / V;
        const float2* mp = maps + (j / nrep) * C * V + v;
        const float2* q = y + j * C * V + v;
        float2 s = make_float2(0.f, 0.f);
        for (int c = 0; c < C; ++c) {
            const float2 m = mp[(int64_t)c * V];
            const float2 d = q[(int64_t)c * V];
            s.x += m.x * d.x + m.y * d.y;
            s.y += m.x * d.y - m.y * d.x;
        }
        x[i] = s;
    }
}

__global__ __launch_bounds__(256) void scale_by_weights_k(float2* __restrict__ y, const float* __restrict__ w, int64_t n,
                                   int64_t period) {
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256) {
        const float f = w[i % period];
        y[i].x *= f;
        y[i].y *= f;
    }
}

static int grid_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, 65536));
}

RcvrOp::RcvrOp(mm_ctx* c, int nslice, const int64_t* V, int64_t C, const float* maps) {
    ctx = c;
    if (nslice < 1 || C < 1) fail(MM_ERR_ARG, "rcvrOp: empty maps");
    C_ = (int)C;
    int64_t off = 0;
    for (int s = 0; s < nslice; ++s) {
        if (V[s] < 1) fail(MM_ERR_ARG, "rcvrOp: empty image");
        Vs_.push_back(V[s]);
        start_.push_back(off);
        off += V[s];
    }
    V_total_ = off;
    if (!maps) fail(MM_ERR_ARG, "rcvrOp: maps is NULL");
    maps_.upload(reinterpret_cast<const float2*>(maps), (size_t)(off * C), ctx->stream);
    HIPCHECK(hipStreamSynchronize(ctx->stream));
}

void RcvrOp::apply(int adjoint, const float2* in, float2* out, int64_t nrep) {
    hipStream_t st = ctx->stream;
    // Runs of consecutive slices with one image size go in one launch.
    for (size_t s0 = 0; s0 < Vs_.size();) {
        size_t s1 = s0 + 1;
        while (s1 < Vs_.size() && Vs_[s1] == Vs_[s0]) ++s1;
        const int64_t V = Vs_[s0], nimg = (int64_t)(s1 - s0) * nrep;
        const float2* mp = maps_.p + start_[s0] * C_;
        const int64_t io = start_[s0] * nrep, dof = start_[s0] * C_ * nrep;
        if (!adjoint) {
            hipLaunchKernelGGL(rcvr_fwd_k, dim3(grid_for(V * C_ * nimg)), dim3(256), 0, st, in + io, mp,
                               out + dof, V, C_, nrep, nimg);
        } else {
            hipLaunchKernelGGL(rcvr_adj_k, dim3(grid_for(V * nimg)), dim3(256), 0, st, in + dof, mp,
                               out + io, V, C_, nrep, nimg);
        }
        LAUNCHCHECK();
        s0 = s1;
    }
}

// Sum of squares of the coil maps per pixel (slices of one image size V in
// one launch, as rcvr_fwd_k) and, per block, its maximum.
constexpr int kSsqBlocks = 1024;
__global__ __launch_bounds__(256) void rcvr_ssq_k(const float2* __restrict__ maps, float* __restrict__ ssq,
                                                  float* __restrict__ bmax, int64_t V, int C, int64_t nimg) {
    __shared__ float red[256];
    const int64_t n = V * nimg;
    float mx = 0.f;
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int64_t v = i % V;
        const float2* mp = maps + (i / V) * C * V + v;
        float s = 0.f;
        for (int c = 0; c < C; ++c) {
            const float2 m = mp[(int64_t)c * V];
            s += m.x * m.x + m.y * m.y;
        }
        ssq[i] = s;
        mx = fmaxf(mx, s);
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) bmax[blockIdx.x] = red[0];
}

// out = x / (ssq + eps * max), the maximum taken over the kSsqBlocks block
// maxima (every block reduces them itself)
__global__ __launch_bounds__(256) void rcvr_norm_k(const float2* __restrict__ x, const float* __restrict__ ssq,
                                                   const float* __restrict__ bmax, int nb, float eps,
                                                   float2* __restrict__ out, int64_t n) {
    __shared__ float red[256];
    float mx = 0.f;
    for (int i = threadIdx.x; i < nb; i += 256) mx = fmaxf(mx, bmax[i]);
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    const float d0 = eps * red[0];
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float d = ssq[i] + d0;
        out[i] = make_float2(x[i].x / d, x[i].y / d);
    }
}

int RcvrOp::ssq_runs() const {
    int nruns = 0;
    for (size_t s0 = 0; s0 < Vs_.size(); ++nruns) {
        size_t s1 = s0 + 1;
        while (s1 < Vs_.size() && Vs_[s1] == Vs_[s0]) ++s1;
        s0 = s1;
    }
    return nruns;
}

int64_t RcvrOp::ssq_scratch_floats() const { return V_total_ + (int64_t)ssq_runs() * kSsqBlocks; }

__global__ __launch_bounds__(256) void fill_zero_f_k(float* __restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = 0.f;
}

void RcvrOp::normalize_by_coil_ssq(const float2* x, float2* out, float eps) {
    ssq_.ensure((size_t)V_total_);
    ssq_max_.ensure((size_t)ssq_runs() * kSsqBlocks);
    normalize_by_coil_ssq_with(x, out, eps, ssq_.p, ssq_max_.p);
}

// ssq: V_total_ floats; bmax: ssq_runs() * kSsqBlocks floats (cleared here
// by a fill kernel, so the sequence stays capturable into a graph)
void RcvrOp::normalize_by_coil_ssq_with(const float2* x, float2* out, float eps, float* ssq, float* bmax) {
    hipStream_t st = ctx->stream;
    const int nruns = ssq_runs();
    hipLaunchKernelGGL(fill_zero_f_k, dim3(grid_for((int64_t)nruns * kSsqBlocks)), dim3(256), 0, st, bmax,
                       (int64_t)nruns * kSsqBlocks);
    LAUNCHCHECK();
    int r = 0;
    for (size_t s0 = 0; s0 < Vs_.size(); ++r) {
        size_t s1 = s0 + 1;
        while (s1 < Vs_.size() && Vs_[s1] == Vs_[s0]) ++s1;
        const int64_t V = Vs_[s0], nimg = (int64_t)(s1 - s0);
        const int nb = (int)std::min<int64_t>(kSsqBlocks, (V * nimg + 255) / 256);
        hipLaunchKernelGGL(rcvr_ssq_k, dim3(nb), dim3(256), 0, st, maps_.p + start_[s0] * C_,
                           ssq + start_[s0], bmax + (size_t)r * kSsqBlocks, V, C_, nimg);
        LAUNCHCHECK();
        s0 = s1;
    }
    hipLaunchKernelGGL(rcvr_norm_k, dim3(grid_for(V_total_)), dim3(256), 0, st, x, ssq, bmax,
                       nruns * kSsqBlocks, eps, out, V_total_);
    LAUNCHCHECK();
}

SenseOp::SenseOp(mm_ctx* c, mm_op* S, mm_op* R, const float* Wh) : S_(S), R_(R) {
    ctx = c;
    if (!S || !R) fail(MM_ERR_ARG, "mrSampFunc: samplingOp and R are required");
    auto* rc = dynamic_cast<RcvrOp*>(R);
    if (!rc) fail(MM_ERR_ARG, "mrSampFunc: R must be a coil operator");
    if (rc->n_image() != S->n_image() || rc->nslice() != S->nslice())
        fail(MM_ERR_ARG, "mrSampFunc: image sizes of S and R differ");
    for (int s = 0; s < S->nslice(); ++s)
        if (rc->slice_image_off(s) != S->slice_image_off(s))
            fail(MM_ERR_ARG, "mrSampFunc: slice layouts of S and R differ");
    C_ = rc->coils();
    nu_ = dynamic_cast<NufftOp*>(S);
    // the unfused path S(R x): more than 16 coils, or a grid on which the
    // fused coil chunk's FFT passes do not fit the LDS (9..16 coils on sides
    // above about 620 points); the plain NUFFT then runs chunks of 8
    if (nu_ && (C_ > 16 || !nu_->sense_fits(C_))) nu_ = nullptr;
    hio_ = dynamic_cast<HioOp*>(S);
    nu3_ = dynamic_cast<Nufft3Op*>(S);
    if (Wh) {
        Wd_.upload(Wh, (size_t)n_data(), ctx->stream);
        HIPCHECK(hipStreamSynchronize(ctx->stream));
        W = Wd_.p;
    }
}

void SenseOp::apply_impl(int adjoint, const float2* in, float2* out, int64_t nrep, const float* Wf) {
    RcvrOp* rc = static_cast<RcvrOp*>(R_);
    // On the Toeplitz form (nufftOp.m:348-352) S maps images to images in
    // both directions: forward T(R x) gives coil images, adjoint R'(T y)
    // takes them.  That goes through the composition below.
    if (nu_ && !nu_->is_toeplitz()) {
        nu_->apply_sense(adjoint, in, out, nrep, rc->maps(), C_, Wf);
        return;
    }
    if (hio_) {
        hio_->apply_sense(adjoint, in, out, nrep, rc->maps(), C_, Wf);
        return;
    }
    const int64_t nimgC = S_->n_image() * C_ * nrep;
    tmp_.ensure((size_t)nimgC);
    if (!adjoint) {
        R_->apply(0, in, tmp_.p, nrep);
        S_->apply(0, tmp_.p, out, nrep * C_);
        if (Wf) {
            hipLaunchKernelGGL(scale_by_weights_k, dim3(grid_for(n_data() * nrep)), dim3(256), 0,
                               ctx->stream, out, Wf, n_data() * nrep, n_data());
            LAUNCHCHECK();
        }
    } else {
        S_->apply(1, in, tmp_.p, nrep * C_);
        R_->apply(1, tmp_.p, out, nrep);
    }
}

void SenseOp::apply(int adjoint, const float2* in, float2* out, int64_t nrep) {
    apply_impl(adjoint, in, out, nrep, nullptr);
}

bool SenseOp::profile_dominant(int adjoint, const float2* in, float2* out, int64_t nrep, int iters,
                               double* ms, std::string* name) {
    RcvrOp* rc = static_cast<RcvrOp*>(R_);
    if (hio_ && adjoint != 2) return hio_->profile_sense(adjoint, in, out, nrep, rc->maps(), C_, iters, ms, name);
    if (!hio_ && (!nu_ || nu_->is_toeplitz()) && (adjoint == 0 || adjoint == 1)) {
        // the composition S(R x): the coil operator is elementwise, the
        // dominant kernel is the sampling operator's over nrep * C images
        const int64_t nr = nrep * C_;
        DevBuf<float2> a((size_t)(S_->n_image() * nr)), b((size_t)(S_->n_data() * nr));
        HIPCHECK(hipMemsetAsync(a.p, 0, sizeof(float2) * a.n, ctx->stream));
        HIPCHECK(hipMemsetAsync(b.p, 0, sizeof(float2) * b.n, ctx->stream));
        return adjoint ? S_->profile_dominant(1, b.p, a.p, nr, iters, ms, name)
                       : S_->profile_dominant(0, a.p, b.p, nr, iters, ms, name);
    }
    if (nu_ && adjoint == 2) {
        // the dominant kernel of the normal operator cgne iterates
        std::vector<std::pair<std::string, std::pair<double, int>>> k;
        nu_->profile_normal(in, out, nrep, rc->maps(), C_, W, iters, &k);
        double best = -1;
        for (auto& e : k) {
            const double tot = e.second.first * e.second.second;
            if (tot > best) {
                best = tot;
                *ms = e.second.first;
                *name = e.first;
            }
        }
        stages_ = k;
        return best >= 0;
    }
    return false;
}

bool SenseOp::autotune(int iters) {
    if (!nu_) return S_->autotune_reps(C_, iters);
    return nu_->autotune_sense(static_cast<RcvrOp*>(R_)->maps(), C_, W, iters);
}

void SenseOp::apply_normal(const float2* in, float2* out, int64_t nrep, float2* tmp) {
    if (nu_) {
        nu_->apply_sense_normal(in, out, nrep, static_cast<RcvrOp*>(R_)->maps(), C_, W);
        return;
    }
    apply_impl(0, in, tmp, nrep, W);
    apply_impl(1, tmp, out, nrep, nullptr);
}

void mm_op_default_normal(mm_op* op, const float2* in, float2* out, int64_t nrep, float2* tmp) {
    op->apply(0, in, tmp, nrep);
    if (op->W) {
        hipLaunchKernelGGL(scale_by_weights_k, dim3(grid_for(op->n_data() * nrep)), dim3(256), 0,
                           op->ctx->stream, tmp, op->W, op->n_data() * nrep, op->n_data());
        LAUNCHCHECK();
    }
    op->apply(1, tmp, out, nrep);
}

}  // namespace mm

void mm_op::apply_normal(const float2* in, float2* out, int64_t nrep, float2* tmp) {
    mm::mm_op_default_normal(this, in, out, nrep, tmp);
}

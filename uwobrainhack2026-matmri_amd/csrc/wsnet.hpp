// Learned warm-start network: 6 layers of 3x3 convolutions (2 -> 32 x5 -> 2
// channels, ReLU between layers, residual skip), bf16 weights, FP32
// accumulation on MFMA.  It maps the density-compensated adjoint image to the
// initial guess x0 that cgne (iterativeSolvers/cgne.m:143-145) accepts; the
// reference has no learned component, so this class has no counterpart there.
//
// Function computed for one complex image x (H = N2 rows, W = N1 columns,
// pixel (i1, i2) at row i2, column i1):
//   s   = sqrt(mean |x|^2)            (1 if x == 0)
//   a0  = bf16(x / s)                  2 channels (re, im)
//   a_l = bf16(relu(conv_l(a_{l-1}) + b_l)),  l = 1..5, 32 channels
//   z   = conv_6(a_5) + b_6            2 channels, FP32
//   out = x + s * z                    (= s * (x/s + z): residual skip)
// with 'same' zero padding and weights/biases rounded to bf16 (round to
// nearest even) from FP32 master copies.  Training minimises
// mean((x/s + z - t/s)^2) with Adam; the backward pass keeps the same
// rounding points (gradients fed to MFMA are bf16, accumulation FP32).
//
// Activations are NHWC bf16: pixel-major, 32 channels = 64 bytes per pixel
// (2-channel tensors: 4 bytes per pixel).
#pragma once
#include "common.hpp"

namespace mm {

constexpr int kWsLayers = 6;
constexpr int kWsCh = 32;
constexpr int kWsTile = 16;  // output tile edge in pixels

}  // namespace mm

struct mm_wsnet {
    mm_wsnet(mm_ctx* ctx, int64_t N1, int64_t N2);
    int64_t num_params() const { return np_; }
    int64_t n_pixels() const { return V_; }
    // theta: flat parameters, per layer W[cout][cin][3][3] then b[cout]
    void set_params(const float* theta, bool on_device);
    void get_params(float* theta, bool on_device);
    void forward(const float2* x, float2* y, int64_t B);
    // loss (returned only when `want_loss`; else NaN) and gradient into grad_
    double loss_grad(const float2* x, const float2* t, int64_t B, bool want_loss);
    void adam_step(float lr);
    const float* grad() const { return grad_.p; }

    mm_ctx* ctx;

   private:
    void ensure_batch(int64_t B, bool training);
    void pack();
    void scale_and_pack_input(const float2* x, int64_t B);
    void run_hidden(int64_t B, bool training);

    int64_t N1_, N2_, V_;
    int tilesX_, tilesY_;
    int64_t np_ = 0;
    int woff_[mm::kWsLayers], boff_[mm::kWsLayers];
    int64_t adam_t_ = 0;
    // packed A operands (bf16) and the index table that builds them
    int64_t pk_f0_, pk_f_[4], pk_f5_, pk_t_[4], pk_t5_, pk_total_;
    mm::DevBuf<float> theta_, m_, v_, grad_, bias_;
    mm::DevBuf<unsigned short> packed_;
    mm::DevBuf<int> pidx_;
    // per-batch workspace
    int64_t Bcap_ = 0;
    bool train_cap_ = false;
    mm::DevBuf<float> sc_, isc_;
    mm::DevBuf<unsigned> a0_, g6_;
    mm::DevBuf<unsigned short> act_[5], dbuf_[2];
    mm::DevBuf<double> lpart_, loss_;
    mm::DevBuf<float> part_;
    int nblk_dw_ = 0;
};

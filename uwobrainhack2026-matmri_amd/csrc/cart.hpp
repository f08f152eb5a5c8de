// Cartesian k-space sampling operator (mrSampFunc with a mask as
// samplingOp) -- declarations.
#pragma once
#include "common.hpp"
#include "fft.hpp"

namespace mm {

class CartOp : public mm_op {
   public:
    // ndim 1 or 2; imN[0] (x imN[1]) image and k-space size; mask [V] real.
    // Lengths up to 2730 with prime factors <= 7 run as LDS Stockham FFTs,
    // others (up to 65536) as direct DFTs
    CartOp(mm_ctx* ctx, int ndim, const int64_t* imN, const float* mask);
    int64_t n_image() const override { return V_; }
    int64_t n_data() const override { return V_; }
    void apply(int adjoint, const float2* in, float2* out, int64_t nrep) override;

   private:
    void pass(int d, int adjoint, const float2* in, float2* out, int64_t nrep, int mask_mode);
    int ndim_ = 2;
    int64_t n_[2] = {1, 1};
    int64_t V_ = 1;
    int tw_off_[2] = {0, 0};
    bool direct_[2] = {false, false};   // direct DFT (a prime factor above 7)
    DevBuf<FftFactors> fac_;
    DevBuf<float2> tw_;
    DevBuf<float> mask_;
    DevBuf<float2> tmp_;   // between the dimensions when the second is a direct DFT
};

// mrSampFunc's opt.sumDims (operators/mrSampFunc.m:100-105): the image side
// is the full column-major array (nd dimensions), the data side the array
// summed over the flagged dimensions; the adjoint broadcasts back.
constexpr int kSumMaxDims = 8;
class SumDimsOp : public mm_op {
   public:
    SumDimsOp(mm_ctx* ctx, int nd, const int64_t* dims, const int* sum_flags);
    int64_t n_image() const override { return nfull_; }
    int64_t n_data() const override { return nred_; }
    void apply(int adjoint, const float2* in, float2* out, int64_t nrep) override;

   private:
    int nd_ = 1;
    int64_t full_[kSumMaxDims] = {}, red_[kSumMaxDims] = {}, fstride_[kSumMaxDims] = {};
    int64_t nfull_ = 1, nred_ = 1, nsum_ = 1;
};

}  // namespace mm

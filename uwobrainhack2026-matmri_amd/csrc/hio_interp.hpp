// Interpolated expanded-encoding operator (sampHighOrder approach 1) --
// declarations.
#pragma once
#include "common.hpp"
#include "fft.hpp"

namespace mm {

class HioInterpOp : public mm_op {
   public:
    // nufft: a single-slice NUFFT of the image onto the K samples (not
    // owned); svdSpace [V x Ns], svdTime [K x Ns], phsShft [K] complex64
    HioInterpOp(mm_ctx* ctx, mm_op* nufft, int64_t V, int64_t K, int Ns, const float* svdSpace,
                const float* svdTime, const float* phsShft);
    int64_t n_image() const override { return V_; }
    int64_t n_data() const override { return K_; }
    void apply(int adjoint, const float2* in, float2* out, int64_t nrep) override;
    // The kernel of a forward (adjoint 0) or adjoint (1) application of nrep
    // images with the largest time: the modulation, the combination, or the
    // NUFFT's dominant stage over its nrep * Ns transforms.
    bool profile_dominant(int adjoint, const float2* in, float2* out, int64_t nrep, int iters, double* ms,
                          std::string* name) override;
    uint64_t config_gen() const override { return gen_ + nu_->config_gen(); }
    // the NUFFT's options for applications of nrep images (nrep * Ns transforms)
    bool autotune(int iters) override { return autotune_reps(1, iters); }
    bool autotune_reps(int64_t nrep, int iters) override { return nu_->autotune_reps(nrep * Ns_, iters); }
    bool set_option(const std::string& key, int value) override { ++gen_; return nu_->set_option(key, value); }
    bool get_option(const std::string& key, int* value) const override { return nu_->get_option(key, value); }

   private:
    mm_op* nu_;
    int64_t V_, K_;
    int Ns_;
    DevBuf<float2> sp_, st_, ph_, imgs_, dat_;
};

}  // namespace mm

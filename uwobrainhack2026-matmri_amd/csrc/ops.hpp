// Coil operator and SENSE composer declarations.
#pragma once
#include "common.hpp"
#include "fft.hpp"
#include "hio.hpp"
#include "nufft.hpp"
#include "nufft3.hpp"

namespace mm {

class RcvrOp : public mm_op {
   public:
    RcvrOp(mm_ctx* ctx, int nslice, const int64_t* V, int64_t C, const float* maps);
    int64_t n_image() const override { return V_total_; }
    int64_t n_data() const override { return V_total_ * C_; }
    int nslice() const override { return (int)Vs_.size(); }
    int64_t slice_image_off(int s) const override { return start_[s]; }
    int64_t slice_data_off(int s) const override { return start_[s] * C_; }
    void apply(int adjoint, const float2* in, float2* out, int64_t nrep) override;
    int coils() const { return C_; }
    const float2* maps() const { return maps_.p; }
    // out = x / (ssq + eps * max(ssq)), ssq = sum_c |maps_c|^2 per pixel,
    // max over every slice (the warm-start network's input)
    void normalize_by_coil_ssq(const float2* x, float2* out, float eps);
    // the same with caller-provided scratch: `ssq` holds n_image() floats,
    // `bmax` ssq_scratch_floats() - n_image() floats; allocates nothing
    void normalize_by_coil_ssq_with(const float2* x, float2* out, float eps, float* ssq, float* bmax);
    // floats of scratch normalize_by_coil_ssq_with needs in all
    int64_t ssq_scratch_floats() const;

   private:
    int ssq_runs() const;   // runs of consecutive slices of one image size
    int C_ = 1;
    std::vector<int64_t> Vs_, start_;
    int64_t V_total_ = 0;
    DevBuf<float2> maps_;
    DevBuf<float> ssq_, ssq_max_;
};

class SenseOp : public mm_op {
   public:
    SenseOp(mm_ctx* ctx, mm_op* S, mm_op* R, const float* W);
    int64_t n_image() const override { return S_->n_image(); }
    int64_t n_data() const override { return S_->n_data() * C_; }
    int nslice() const override { return S_->nslice(); }
    int64_t slice_image_off(int s) const override { return S_->slice_image_off(s); }
    int64_t slice_data_off(int s) const override { return S_->slice_data_off(s) * C_; }
    void apply(int adjoint, const float2* in, float2* out, int64_t nrep) override;
    void apply_normal(const float2* in, float2* out, int64_t nrep, float2* tmp) override;
    bool profile_dominant(int adjoint, const float2* in, float2* out, int64_t nrep, int iters,
                          double* ms, std::string* name) override;
    bool set_option(const std::string& key, int value) override { return S_->set_option(key, value); }
    bool get_option(const std::string& key, int* value) const override { return S_->get_option(key, value); }
    bool autotune(int iters) override;
    void set_skip(const int* flags) override { S_->set_skip(flags); }
    uint64_t config_gen() const override { return gen_ + S_->config_gen() + R_->config_gen(); }
    // the sampling operator is a NUFFT in the Toeplitz form (prepToep)
    // the sampling operator itself (nu_ is NULL on the unfused path)
    bool toeplitz() const {
        const auto* n2 = dynamic_cast<const NufftOp*>(S_);
        return (n2 && n2->is_toeplitz()) || (nu3_ && nu3_->is_toeplitz());
    }
    mm_op* sampling() const { return S_; }
    int coils() const { return C_; }
    // per-kernel times of the last normal-operator profile
    std::vector<std::pair<std::string, std::pair<double, int>>> stages_;

   private:
    void apply_impl(int adjoint, const float2* in, float2* out, int64_t nrep, const float* Wf);
    mm_op* S_;
    mm_op* R_;
    int C_ = 1;
    NufftOp* nu_ = nullptr;
    Nufft3Op* nu3_ = nullptr;   // 3-D NUFFT: composed through the generic path
    HioOp* hio_ = nullptr;
    DevBuf<float2> tmp_;
    DevBuf<float> Wd_;
};

}  // namespace mm

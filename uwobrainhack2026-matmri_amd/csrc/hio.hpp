// Expanded-encoding (high-order) operator, sampHighOrder direct model.
#pragma once
#include "common.hpp"

namespace mm {

class HioOp : public mm_op {
   public:
    HioOp(mm_ctx* ctx, int64_t V, int64_t K, int Lh, int Lc, const double* x, const double* y,
          const double* z, const double* b0, const double* t, const double* h, const double* g,
          const double* mask);
    int64_t n_image() const override { return V_; }
    int64_t n_data() const override { return K_; }
    void apply(int adjoint, const float2* in, float2* out, int64_t nrep) override;
    // SENSE-fused: image side one image per repetition times `maps` [V x C];
    // data side [K x C x nrep]; W (nullable) weights the forward output.
    void apply_sense(int adjoint, const float2* in, float2* out, int64_t nrep, const float2* maps,
                     int C, const float* W);
    void set_phi_div(const double* hd, const double* gd);
    bool profile_dominant(int adjoint, const float2* in, float2* out, int64_t nrep, int iters,
                          double* ms, std::string* name) override;
    bool profile_sense(int adjoint, const float2* in, float2* out, int64_t nrep, const float2* maps,
                       int C, int iters, double* ms, std::string* name);
    int phase_terms() const { return L_; }
    int rank() const { return r_; }
    // max |Phi| in radians over the plan (host FP64, computed at creation)
    double max_abs_phase() const { return maxphi_; }

   private:
    void run(int adjoint, const float2* in, float2* out, int64_t ncol, const float2* maps, int C,
             const float* W, int prof_iters = 0, double* prof_ms = nullptr);
    int64_t V_, K_;
    int L_ = 0;      // phase terms before compression
    int r_ = 0;      // rank after compression
    int Lnb_ = 0;    // terms without the B0 column (phiDiv basis)
    int Lh_ = 0, Lc_ = 0;
    double maxphi_ = 0;
    DevBuf<double> Q_;    // [V][r] compressed voxel basis
    DevBuf<double> Hc_;   // [K][r] compressed coefficients, revolutions
    // phiDiv: dPhi = Bnb * Hd  (FP64 basis without B0, coefficients in rad)
    bool phidiv_ = false;
    std::vector<double> Bnb_host_;   // [V][Lnb]
    DevBuf<double> Bnb_, Hd_;        // [V][Lnb], [K][Lnb]
    DevBuf<float2> X_, part_;
};

}  // namespace mm

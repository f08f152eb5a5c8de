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
    bool set_option(const std::string& key, int value) override {
        ++gen_;
        if (key == "hio_variant") { variant_ = value & 7; return true; }
        if (key == "hio_mfma") { use_mfma_ = value != 0; return true; }
        return false;
    }
    bool get_option(const std::string& key, int* value) const override {
        if (key == "hio_variant") { *value = variant_; return true; }
        if (key == "hio_mfma") { *value = use_mfma_ ? 1 : 0; return true; }
        return false;
    }
    bool profile_sense(int adjoint, const float2* in, float2* out, int64_t nrep, const float2* maps,
                       int C, int iters, double* ms, std::string* name);
    int phase_terms() const { return L_; }
    int rank() const { return r_; }
    // mrSampFuncMat's A'A over the masked voxels vidx [Vm] (device): S gets
    // the dense [K x Vm] encoding matrix (workspace), AtA the [Vm x Vm] result;
    // mmaps = masked coil maps [Vm x C]; blas = a rocBLAS handle
    void dense_normal(const int* vidx, int64_t Vm, const float2* mmaps, int C, float2* S, float2* AtA, void* blas);
    int64_t samples() const { return K_; }

    // the MFMA path is used when the padded rank is <= 16 and there is no phiDiv
    bool use_mfma_ = true;
    int variant_ = 3;  // MFMA kernel variant (see HioMfmaCall::variant)
    const char* kernel_name(int adjoint) const {
        if (!phidiv_ && rp_ <= 16 && use_mfma_) return "hio_mfma_gemm";
        return adjoint ? "hio_adj_simple" : "hio_fwd_simple";
    }

   private:
    template <int NC>
    void run_simple(int adjoint, const float2* in, float2* out, int ncol, const float2* maps, int C,
                    const float* W, int prof_iters, double* prof_ms);
    void run_mfma(int adjoint, const float2* in, float2* out, int ncol, const float2* maps, int C,
                  const float* W, int prof_iters, double* prof_ms);
    void run(int adjoint, const float2* in, float2* out, int64_t ncol, const float2* maps, int C,
             const float* W, int prof_iters = 0, double* prof_ms = nullptr);
    int64_t V_, K_;
    int L_ = 0;      // phase terms before compression
    int r_ = 0;      // rank after compression
    int rp_ = 4;     // rank padded to a multiple of 4 (row stride of Q_/Hc_)
    int64_t Vp_ = 0, Kp_ = 0;   // row counts padded to a multiple of 16
    int Lnb_ = 0;    // terms without the B0 column (phiDiv basis)
    int Lh_ = 0, Lc_ = 0;
    int num_cu_ = 256;   // compute units of the device (work-group slots of the MFMA GEMM)
    DevBuf<double> Q_;    // [V][r] compressed voxel basis
    DevBuf<double> Hc_;   // [K][r] compressed coefficients, revolutions
    // phiDiv: dPhi = Bnb * Hd  (FP64 basis without B0, coefficients in rad)
    bool phidiv_ = false;
    std::vector<double> Bnb_host_;   // [V][Lnb]
    DevBuf<double> Bnb_, Hd_;        // [V][Lnb], [K][Lnb]
    DevBuf<float2> X_, part_;
    DevBuf<unsigned char> Bs_;
    DevBuf<float> mpart_;
    DevBuf<unsigned> amax_;
};

void* blas_handle_create();
void blas_handle_destroy(void* h);
void dense_gram(void* blas, hipStream_t st, const float2* S, int64_t m, int64_t n, float2* G);

}  // namespace mm

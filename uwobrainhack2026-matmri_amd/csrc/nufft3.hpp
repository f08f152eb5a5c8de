// 3-D Kaiser-Bessel gridding NUFFT on gfx950: nufftOp with numel(imN) == 3
// (operators/@nufftOp/nufftOp.m:138-274 for the plan, :354-394 for the two
// directions; the 3-D cases of :195-197, :244-245, :268-271).
#pragma once
#include <array>

#include "common.hpp"
#include "fft.hpp"

namespace mm {

// Geometry every 3-D kernel takes by value.
struct Nu3Geom {
    int N1, N2, N3, G1, G2, G3;
    int s1, s2, s3;     // floor(N_d / 2): the wrapped placement of each side
    int M, W, odd;      // odd: bit d set when side d is odd
    int ip;             // bit d set when the FFT passes along side d run in one LDS buffer
    int64_t V;          // N1 N2 N3
    int64_t chunk;      // grid elements of one chunk: G1 G2 G3 BCT
};

class Nufft3Op : public mm_op {
   public:
    // imN = {N1, N2, N3}; kloc = column-major [M x 3] cycles/pixel in
    // [-0.5, 0.5) (wrapped); dcf = [M] or NULL.
    Nufft3Op(mm_ctx* ctx, const int64_t* imN, int64_t M, const double* kloc, const double* dcf, double os,
             int kwidth);
    int64_t n_image() const override { return g_.V; }
    int64_t n_data() const override { return toep_ ? g_.V : g_.M; }
    void apply(int adjoint, const float2* in, float2* out, int64_t nrep) override;
    // The stage of a forward (adjoint 0) or adjoint (1) application of nrep
    // transforms with the largest time (HIP events after every launch, `iters`
    // applications after one untimed one).
    bool profile_dominant(int adjoint, const float2* in, float2* out, int64_t nrep, int iters, double* ms,
                          std::string* name) override;
    // Toeplitz form (nufftOp.prepToep, operators/@nufftOp/nufftOp.m:410-432,
    // applied at :348-352): `psf` is the point-spread volume A2'(dcf) on the
    // doubled grid 2N1 x 2N2 x 2N3 (column-major).  Afterwards both
    // directions apply crop(h * pad(x)), h the psf scaled by
    // 2^3 / sqrt(8 N1 N2 N3): FFTs along the three sides on this plan's grid
    // (G_d >= 2 N_d - 1), a multiply by the spectrum of h (formed on the host
    // in FP64), inverse FFTs.
    void set_toeplitz(const float2* psf);
    bool is_toeplitz() const { return toep_; }

   private:
    template <int BCT, int MID>
    void launch(int adjoint, const float2* in, float2* out, int64_t nrep);
    Nu3Geom g_{};
    // stage marks of profile_dominant: (stage, event after it); inactive when NULL
    std::vector<std::pair<const char*, hipEvent_t>>* marks_ = nullptr;
    size_t nmark_ = 0;
    std::array<int, 3> tw_off_{};        // per side: first twiddle in twid_
    DevBuf<FftFactors> facs_;            // per side
    std::array<FftFactors, 3> fach_{};   // host copies (single-buffer checks per transform count)
    DevBuf<float2> twid_;
    DevBuf<float> deap_;                 // N1 + N2 + N3 values, scale and sign in side 1
    DevBuf<uint32_t> base_;              // first taps g1 | g2 << 10 | g3 << 20 (wrapped)
    DevBuf<float> wt_, dcf_;             // per sample: W weights per side; dcf (1 without)
    DevBuf<int> rp_;                     // CSR gridding: row pointers per grid cell
    DevBuf<int> long_;                   // cells whose rows exceed kN3LongRow entries (one wave each)
    int nlong_ = 0;
    DevBuf<int2> ent_;                   // (sample, w1 w2 w3 dcf as float bits), sample order
    DevBuf<float2> ws_, yT_;             // one chunk's grid; transposed data
    bool toep_ = false;
    DevBuf<float2> toep_k_;              // spectrum of h, [g3][g2][g1]
    DevBuf<float> toep_deap_;            // flat 1 / sqrt(G1 G2 G3) on side 1, ones elsewhere
};

}  // namespace mm

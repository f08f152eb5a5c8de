// Kaiser-Bessel gridding NUFFT on gfx950 (declarations shared with the SENSE
// composer).  The plan follows nufftOp.doPrep (operators/@nufftOp/nufftOp.m:
// 138-274); forward and adjoint follow nufftOp.mtimes (:354-394).
#pragma once
#include <array>
#include <map>

#include "common.hpp"
#include "fft.hpp"

namespace mm {

// Per-slice constants seen by every NUFFT kernel.
struct NuSliceDev {
    int N1, N2, G1, G2, M, W1, W2;
    int fac1, fac2;     // index into the FFT factor pool
    int tw1, tw2;       // offset into the twiddle pool (float2)
    int deap1, deap2;   // offset into the deapodisation pool (float)
    int rpw, cpw;       // rows / columns per work-group in the FFT passes
    int frpw;           // rows per work-group of the forward row FFT
    int lfl, lcl;       // log2 transforms (coils) per work-group: forward rows, columns
    int fip;            // bit 0: row FFTs, bit 1: column FFTs in a single LDS buffer
    int fac2w, tw2w;    // the column length's radix-16/12 factorization (passes of
                        // 16 transforms): used for fac2/tw2 when the column pass has them
    int fac1w, tw1w;    // the row length's, used for fac1/tw1 by row passes of 16
                        // transforms in the mid-size FFT mode
    int odd;            // bit 0: N1 odd, bit 1: N2 odd (the grid takes the odd-size ramp)
    int64_t samp_off;   // first sample of this slice in the sample pools
    int64_t img_off;    // image-side offset (elements) for the current apply
    int64_t maps_off;   // coil-map offset (elements)
    int64_t dat_off;    // data-side offset (elements)
    int64_t ws_off;     // workspace offset of chunk 0 (elements)
    int64_t ws_chunk;   // workspace elements per chunk = G2*G1*BCT
    int64_t part_off;   // partial blocks of the dense rectangles: offset of chunk 0
    int64_t part_chunk; // partial-block elements per chunk
    int64_t yt_off;     // transposed data ([M][BCT] per chunk) offset of chunk 0
    int64_t csr_off;    // first row pointer of this slice (CSR gridding), G1*G2 + 1 of them
    int64_t ent_off;    // first CSR entry of this slice
};

// One rectangle of grid cells for 2-D CSR gridding, with the union of the
// samples that reach it (sorted, at most stage_samples_) and the cells'
// (local sample, weight) rows (at most stage_entries_ entries), staged in
// LDS by its work-group.  kind 0: the whole rectangle, written to the grid;
// kind 2: one chunk of the samples of a dense rectangle (k-space centre),
// written to a partial block; a second launch sums the blocks in chunk order
// into the grid.  Offsets are absolute in the pools; partial offsets (pofs,
// pofs0, in cells) are relative to the slice.
struct GridTile {
    int g1, g2, w, h;    // first cell and extent (cells, g1 fastest)
    int kind;
    int n_stage, n_ent;  // staged samples / entries
    int stage_off;       // first sample id in the stage pool
    int ptr_off;         // first of the w*h+1 row pointers (relative to ent_off)
    int ent_off;         // first (local sample, weight) entry
    int pofs, pofs0;     // kind 2: this chunk's partial block, the rectangle's first
    int nparts;          // kind 2: chunks of the rectangle
    int region;          // one of kXcds spatially contiguous groups of equal staging cost
};
// One work item of the staged gridding: a rectangle of one (slice, chunk)
// with every offset its work-group needs, so that one scalar load precedes
// the data loads.
struct GridItem {
    int64_t yt;          // transposed data of the (slice, chunk) in yT
    int64_t zg;          // grid of the (slice, chunk) in the workspace
    int64_t part;        // partial blocks of the (slice, chunk) (kind 2)
    // second (slice, chunk) of the same trajectory handled by the item
    // (slice2 < 0: none); its transforms follow the first's in the lanes
    int64_t yt2, zg2, part2;
    int slice, G1, kind, lanes;
    int g1, g2, w, h;
    int stage_off, n_stage, ptr_off, ent_off;
    int n_ent, pofs, pofs0, nparts;
    int slice2, lanes2;
};
constexpr int kDegridTile = 8;      // tile side of the staged degridding
// work lists of the normal operator that the slice groups split
enum NuList { kLRows = 0, kLFrows, kLCols, kLFcols, kLDb, kLGt, kLGm, kNuLists };

class NufftOp : public mm_op {
   public:
    NufftOp(mm_ctx* ctx, int ndim, int nslice, const int64_t* imN, const int64_t* M,
            const double* const* kloc, const double* const* dcf, double os, int kwidth);
    int64_t n_image() const override { return V_total_; }
    int64_t n_data() const override { return toep_ ? V_total_ : M_total_; }
    int nslice() const override { return (int)sl_.size(); }
    // Toeplitz form of the normal operator (nufftOp.prepToep,
    // operators/@nufftOp/nufftOp.m:410-432, applied at :348-352): `psf`
    // holds each slice's point-spread image A2'(dcf.^2) on the doubled
    // image grid (2 N1 x 2 N2, column-major, slices concatenated).  After the
    // call both directions apply crop(h * pad(x)), h the psf scaled by
    // 2^ndim / sqrt(prod 2 N_d): row and column FFTs on this plan's grid
    // (G_d >= 2 N_d), a multiply by the spectrum of h, inverse column and
    // row FFTs.  The spectrum is formed on the host in FP64.
    void set_toeplitz(const float2* psf);
    bool toeplitz() const { return toep_; }
    int64_t slice_image_off(int s) const override { return img_start_[s]; }
    int64_t slice_data_off(int s) const override { return dat_start_[s]; }
    void apply(int adjoint, const float2* in, float2* out, int64_t nrep) override;
    // SENSE-fused application: image side is one image per repetition,
    // combined with `maps` ([V_s x C] per slice, concatenated); data side is
    // [M_s x C x nrep] per slice.  W (nullable) multiplies the forward output.
    void apply_sense(int adjoint, const float2* in, float2* out, int64_t nrep, const float2* maps,
                     int C, const float* W);
    // x_out = E'(W .* E x) for the SENSE-fused operator; the data side stays in
    // the transposed [M][coil] layout between the two halves.
    void apply_sense_normal(const float2* in, float2* out, int64_t nrep, const float2* maps, int C,
                            const float* W);
    // Whether the fused SENSE configuration of C coils can be planned (its
    // FFT passes fit the LDS); false where it is refused as unsupported.
    bool sense_fits(int C);
    // Per-kernel time of the normal operator E'(W.*E x): `iters`
    // applications on the context stream with a HIP event after every
    // kernel; per kernel name: (ms per launch, launches per application).
    void profile_normal(const float2* in, float2* out, int64_t nrep, const float2* maps, int C,
                        const float* W, int iters,
                        std::vector<std::pair<std::string, std::pair<double, int>>>* per_kernel);
    // The stage of a plain forward (adjoint 0) or adjoint (1) application of
    // nrep transforms with the largest time per application.
    bool profile_dominant(int adjoint, const float2* in, float2* out, int64_t nrep, int iters, double* ms,
                          std::string* name) override;
    // Tuning options (mm_op_set_option); the plan-time partitions are rebuilt
    // when their caps change:
    //   nu_stage_samples / nu_stage_entries  staged gridding rectangle caps
    //   nu_degrid_block                      samples per staged degridding block
    //   nu_pair_cols        XCD-paired order of the column-pass items (0/1)
    //   nu_xcd_order        XCD-aware order of the staged items (0/1)
    //   nu_grid_stage       2-D gridding from staged rectangles (1) or the global CSR rows (0)
    //   nu_degrid_stage     2-D degridding from staged tiles (1) or the grid in global memory (0)
    //   nu_degrid_t         unstaged degridding with one lane per (sample, coil) (1)
    bool set_option(const std::string& key, int value) override;
    bool get_option(const std::string& key, int* value) const override;
    // The autotuner of the normal operator -- SENSE-fused (maps of C coils,
    // nrep 1) or, with maps NULL, nrep plain transforms (the gridding form):
    // times the candidate gridding caps, degridding block sizes, column-pass
    // order and staged-item order with profile_normal (`iters` applications
    // each), one option at a time, and keeps the fastest.
    bool autotune_batch(int64_t nrep, const float2* maps, int C, const float* W, int iters);
    bool autotune_sense(const float2* maps, int C, const float* W, int iters) {
        return autotune_batch(1, maps, C, W, iters);
    }
    bool autotune_reps(int64_t nrep, int iters) override { return autotune_batch(nrep, nullptr, 0, nullptr, iters); }
    void set_skip(const int* flags) override { skip_ = flags; }
    int ndim() const { return ndim_; }
    const std::vector<int64_t>& slice_V() const { return Vs_; }
    const std::vector<int64_t>& slice_M() const { return Ms_; }

    // host copies of the plan constants (exposed for tests)
    double beta = 0;
    int nkern = 0;

   private:
    struct Cfg {
        int BCT = 1, sense = 0, C = 0;
        int64_t nrep = 0;
        std::vector<NuSliceDev> hs;
        DevBuf<NuSliceDev> sl;
        DevBuf<int4> w_rows, w_frows, w_cols, w_deg;
        DevBuf<int4> w_fcols;    // forward column items (w_cols: adjoint)
        int n_rows = 0, n_frows = 0, n_cols = 0, n_deg = 0;
        size_t lds_rows = 0, lds_cols = 0;
        int fm = 0;   // FFT mode of the row / column kernels (kFftSmall, kFftMid, kFftBig)
        DevBuf<float2> ws, parts, yT;
        DevBuf<int4> w_csr;      // CSR gridding items: (slice, chunk, first cell, cells)
        int n_csr = 0;
        // staged CSR gridding items: (slice, chunk, tile, lp | c0 << 8 | cells << 16)
        DevBuf<int4> w_db;             // staged degridding: (slice, chunk, block)
        int n_db = 0;
        size_t lds_db = 0;
        DevBuf<GridItem> w_gt, w_gm;   // staged items; merges of dense rectangles
        int n_gt = 0, n_gm = 0;
        size_t lds_gt = 0;
        int gt_pairs = 0;              // items carry two (slice, chunk) units
        // slice groups of the normal operator (option "nu_slice_group"): the
        // work lists are ordered group by group; group g's items of list L
        // are [goff[g][L], goff[g][L] + gcnt[g][L])
        std::vector<std::array<int, kNuLists>> goff, gcnt;
        // staged sample ids of item i at gt_sid[i * gt_sid_stride]: read
        // from the block index alone, before the item record arrives
        DevBuf<int> gt_sid;
        int gt_sid_stride = 16;
    };
    Cfg& config(int64_t nrep, int sense, int C);
    // mode 0: the requested direction with the standard data layout;
    // mode 1: the normal operator (forward writes the transposed data layout
    // that the gathering adjoint reads)
    void run(Cfg& c, int adjoint, const float2* in, float2* out, const float2* maps, const float* W,
             int mode = 0);
    template <int BCT, int FM>
    void launch(Cfg& c, int adjoint, const struct NuRun& run, const float2* in, float2* out, int mode);
    // per-kernel times of `iters` runs of (adjoint, mode) with an event after every kernel
    void profile_run(int adjoint, int mode, const float2* in, float2* out, int64_t nrep, const float2* maps, int C,
                     const float* W, int iters,
                     std::vector<std::pair<std::string, std::pair<double, int>>>* per_kernel);
    bool big_ = false;
    bool toep_ = false;
    bool any_odd_ = false;   // some slice has an odd image side: the ramp stage runs

   public:
    bool is_toeplitz() const { return toep_; }

   private:
    // 2-D Toeplitz middle stage as one column kernel (option "nu_toep_fused",
    // default on); off: forward columns, nu_toep_mul, adjoint columns
    bool toep_fused_ok_ = true;
    DevBuf<float2> toep_m_;      // per slice the spectrum of h on the grid, Z[g2][g1] order
    DevBuf<int64_t> toep_off_;   // per slice: first entry in toep_m_
    const int* skip_ = nullptr;   // per-slice skip flags of the normal operator   // a grid side above kFftSmallMax: FFT kernels with the wide modulo
    // event marks for profile_normal (inactive unless prof_ is set)
    struct Prof {
        std::vector<hipEvent_t> ev;
        std::vector<const char*> names;
        int k = 0;
    };
    Prof* prof_ = nullptr;
    void mark(const char* name);

    int ndim_ = 2;
    int W_ = 4;
    double os_ = 1.5;
    std::vector<NuSliceDev> sl_;          // static part of each slice
    std::vector<int64_t> Vs_, Ms_, img_start_, dat_start_;
    int64_t V_total_ = 0, M_total_ = 0;
    // static device pools
    DevBuf<FftFactors> facs_;
    DevBuf<float2> twid_;
    DevBuf<float> deap_;
    DevBuf<uint32_t> sbase_;
    DevBuf<float> swt_, sdcf_;
    std::map<std::pair<int64_t, int>, std::unique_ptr<Cfg>> cfgs_;
    // CSR gridding (1-D and 2-D): per cell the (sample, weight) pairs that
    // reach it, in sample order; row pointers are relative to the slice's
    // ent_off
    bool degrid_t_ = true;   // BCT lanes per sample in degridding (option "nu_degrid_t")
    // XCD-aware order of the gridding items (option "nu_xcd_order"); off: on
    // the MI355X it measured slower than the heaviest-first order (318 vs 297 us
    // per launch for 100 slices x 8 coils)
    bool xcd_order_ = false;
    std::vector<int> csr_ptr_host_;
    DevBuf<int> csr_ptr_;
    DevBuf<int2> csr_ent_;
    // staged CSR gridding (option "nu_grid_stage", default on): rectangles
    // of each distinct 2-D trajectory; slices that share a trajectory share
    // its rectangles
    bool grid_stage_ = true;
    int stage_samples_ = 128, stage_entries_ = 1024;   // staged gridding rectangle caps
    // one staged gridding item serves two (slice, chunk) units of one
    // trajectory, so its rectangle's CSR entries are read once for both
    // (option "nu_grid_pairs")
    bool grid_pairs_ = true;
    bool grid_region_ = true;   // with xcd_order_: staged gridding items of a region on one XCD
    // alternating slice order of the normal operator's stages (option
    // "nu_snake"): each stage starts on the slices the previous one touched
    // last
    bool snake_ = true;
    // slices per group of the normal operator (option "nu_slice_group"; 0:
    // one group): every stage runs over one group before the next group
    // starts, so a group's intermediate data stay in the memory-side cache
    int slice_group_ = 0;
    int degrid_block_ = 64;                             // samples per staged degridding block
    // two sparse tiles side by side share a degridding block (option
    // "nu_degrid_merge")
    bool degrid_merge_ = true;
    void build_partitions();
    // host copies the partitions are rebuilt from
    std::vector<int2> csr_ent_host_;
    std::vector<uint32_t> sbase_host_;
    std::vector<float> swt_host_, sdcf_host_;
    bool pair_cols_ = true;   // XCD-paired order of the column-pass items (option "nu_pair_cols")
    bool fft_mid_ = true;     // mid-size FFT mode for grid sides the small one cannot hold in one LDS buffer ("nu_fft_mid")
    // staged degridding (option "nu_degrid_stage", default on): sample
    // blocks of each distinct 2-D trajectory, (g1 | g2 << 16 of the tile's
    // corner, first entry of dsamp_, samples, tile width along g1)
    bool degrid_stage_ = true;
    std::vector<int4> dblk_host_;
    std::vector<int> dblk_first_, dblk_count_;   // per slice
    DevBuf<int4> dblk_;
    DevBuf<int> dsamp_;          // sample ids in block order
    DevBuf<uint32_t> dbase_;     // first taps in block order
    DevBuf<float> ddcf_, dwt_;   // dcf, kernel weights (W1 + W2 a sample) in block order
    std::vector<GridTile> gtile_host_;
    std::vector<int> gtile_first_, gtile_count_;   // per slice
    std::vector<int> gpcells_;                     // per slice: cells of the partial blocks
    DevBuf<int> gptr_;
    std::vector<int> gstage_host_;   // staged sample ids of the rectangles (host)
    DevBuf<int2> gent_;
};

// Plan constants from the closed forms of nufftOp.m:146-182 (host, FP64).
int nufft_grid_size(int64_t n, double os);
double nufft_kb_beta(int kwidth, double os);
int nufft_kb_nkern(int kwidth, double os);
// KB lookup table with its zero sentinel (nufftOp.m:157-165), nkern + 1 values
std::vector<double> nufft_kb_lut(int kwidth, double os);
// deapodisation comp{d} of one side (nufftOp.m:171-182), N values
std::vector<double> nufft_deapod(int N, int G, int kwidth, double os);

}  // namespace mm

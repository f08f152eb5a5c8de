// Device cgne solver (declarations).
#pragma once
#include "common.hpp"

namespace mm {

struct CgScal;
struct CgSeg;

class CgSolver {
   public:
    explicit CgSolver(mm_op* E);
    ~CgSolver();
    // all vectors are device pointers; histories/nit/status are host
    // regs: Tikhonov terms (gam, R) of cgne.m:171-179 / :219-227, R = NULL
    // for the identity; gtruth (device, or NULL) and mseAll: cgne.m's mse
    // history |gtruth - x_n|^2 (:201-204, :271-274)
    void solve(const float2* y, const float2* x0, float2* x, int NitMax, const mm_cg_opts& o,
               double* resSqAll, double* xnormAll, double* xdiffAll, int* nit, int* status,
               const std::vector<std::pair<double, mm_op*>>& regs = {}, const float2* gtruth = nullptr,
               double* mseAll = nullptr);
    bool use_graph_ = true;
    // per slice: 1 if the last solve replaced x0 by 0 (mm_cg_opts.x0Fallback)
    std::vector<int> last_fallback_;

   private:
    void launch_iteration(int explicit_res, const mm_cg_opts& o, float2* x);
    void drop_graph();
    void normal(const float2* v, float2* out);
    uint64_t config_gen() const;   // E's and the regularisers' generations
    std::vector<std::pair<double, mm_op*>> regs_;
    const float2* gt_ = nullptr;
    DevBuf<float2> rbuf_, rtmp_;
    DevBuf<double> hmse_;
    mm_op* E_;
    int64_t n_ = 0;
    int nblk_ = 0;
    int nitmax_ = -1;
    DevBuf<CgSeg> seg_;
    DevBuf<int> b2s_;
    DevBuf<CgScal> sc_;
    DevBuf<double2> part_, part2_;
    DevBuf<double2> partu_;   // partial sums of cg_update (r'r, x'x)
    DevBuf<double> pp_, hres_, hxn_, hxd_;
    DevBuf<int> done_d_;
    DevBuf<float2> b_, r_, p_, q_, tmp_;
    int* done_h_ = nullptr;     // host view of the done flags (coherent, mapped)
    int* done_map_ = nullptr;   // device view of the same memory
    hipEvent_t poll_ev_ = nullptr;   // recorded after the copy of the done flags
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    float2* graph_x_ = nullptr;
    uint64_t graph_gen_ = 0;     // E's configuration generation at capture
    uint64_t graph_epoch_ = 0;   // device free count at capture
};

}  // namespace mm

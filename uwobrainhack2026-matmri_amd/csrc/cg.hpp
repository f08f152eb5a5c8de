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
    void solve(const float2* y, const float2* x0, float2* x, int NitMax, const mm_cg_opts& o,
               double* resSqAll, double* xnormAll, double* xdiffAll, int* nit, int* status);
    bool use_graph_ = true;
    // per slice: 1 if the last solve replaced x0 by 0 (mm_cg_opts.x0Fallback)
    std::vector<int> last_fallback_;

   private:
    void launch_iteration(int explicit_res, const mm_cg_opts& o, float2* x);
    void drop_graph();
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

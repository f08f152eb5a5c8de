// Trajectory resampling (interpTrajTime) -- declarations.
#pragma once
#include "common.hpp"

namespace mm {

// out[m + nq*c] = makima interpolation of column c of phs ([L x ncol],
// column-major, samples at -delTime + tdwell*k) at time tq[m], NaN outside;
// all pointers device, asynchronous on `st`.
void interp_traj_time(hipStream_t st, const double* phs, int64_t L, int ncol, double tdwell, double delTime,
                      const double* tq, int64_t nq, double* out);

}  // namespace mm

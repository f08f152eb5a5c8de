// Trajectory resampling of findDel/interpTrajTime.m:1-27 on the device:
// interp1(trajTim, phs, datatime, 'makima', nan) along the time dimension
// of each column, trajTim = -delTime + tdwell*(0:L-1) (uniform).
//
// Modified Akima (makima) piecewise cubic Hermite interpolation: slopes
// delta_i = (y_{i+1} - y_i)/h, extended past both ends by linear
// extrapolation of the slopes (delta_{-1} = 2 delta_0 - delta_1, ...); node
// derivative d_i = (w1 delta_{i-1} + w2 delta_i)/(w1 + w2) with
// w1 = |delta_{i+1} - delta_i| + |delta_{i+1} + delta_i|/2,
// w2 = |delta_{i-1} - delta_{i-2}| + |delta_{i-1} + delta_{i-2}|/2 (d_i = 0
// when both vanish); queries outside [trajTim(1), trajTim(end)] give NaN.
// FP64 throughout; one thread per (query time, column).
#include <cmath>

#include "traj.hpp"

namespace mm {

// slope delta_j of the column, j in [-2, L], with the end extensions
__device__ __forceinline__ double mk_slope(const double* __restrict__ y, int64_t L, double h, int64_t j) {
    auto d = [&](int64_t i) { return (y[i + 1] - y[i]) / h; };
    if (L < 2) return 0.0;
    if (j >= 0 && j <= L - 2) return d(j);
    if (L == 2) return d(0);   // one interval: every slope is the same
    if (j == -1) return 2.0 * d(0) - d(1);
    if (j == -2) return 2.0 * (2.0 * d(0) - d(1)) - d(0);
    if (j == L - 1) return 2.0 * d(L - 2) - d(L - 3);
    // j == L
    return 2.0 * (2.0 * d(L - 2) - d(L - 3)) - d(L - 2);
}

__device__ __forceinline__ double mk_deriv(const double* __restrict__ y, int64_t L, double h, int64_t i) {
    const double dm2 = mk_slope(y, L, h, i - 2), dm1 = mk_slope(y, L, h, i - 1);
    const double d0 = mk_slope(y, L, h, i), dp1 = mk_slope(y, L, h, i + 1);
    const double w1 = fabs(dp1 - d0) + 0.5 * fabs(dp1 + d0);
    const double w2 = fabs(dm1 - dm2) + 0.5 * fabs(dm1 + dm2);
    const double s = w1 + w2;
    return s > 0.0 ? (w1 * dm1 + w2 * d0) / s : 0.0;
}

__global__ __launch_bounds__(256) void interp_makima_k(const double* __restrict__ y, int64_t L, int ncol, double t0,
                                                       double h, const double* __restrict__ tq, int64_t nq,
                                                       double* __restrict__ out) {
    const int64_t n = nq * ncol;
    for (int64_t idx = blockIdx.x * (int64_t)256 + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * 256) {
        const int64_t m = idx % nq;
        const int c = (int)(idx / nq);
        const double* yc = y + (int64_t)c * L;
        const double t = tq[m];
        const double tend = t0 + h * (double)(L - 1);
        double v;
        if (!(t >= t0 && t <= tend)) {
            v = NAN;
        } else if (L == 1) {
            v = yc[0];
        } else {
            const double u = (t - t0) / h;
            int64_t i = (int64_t)floor(u);
            i = i < 0 ? 0 : (i > L - 2 ? L - 2 : i);
            const double s = u - (double)i;
            const double di = mk_deriv(yc, L, h, i), di1 = mk_deriv(yc, L, h, i + 1);
            const double s2 = s * s, s3 = s2 * s;
            const double h00 = 2 * s3 - 3 * s2 + 1, h10 = s3 - 2 * s2 + s;
            const double h01 = -2 * s3 + 3 * s2, h11 = s3 - s2;
            v = h00 * yc[i] + h10 * h * di + h01 * yc[i + 1] + h11 * h * di1;
        }
        out[idx] = v;
    }
}

void interp_traj_time(hipStream_t st, const double* phs, int64_t L, int ncol, double tdwell, double delTime,
                      const double* tq, int64_t nq, double* out) {
    if (L < 1 || ncol < 1 || nq < 0 || !(tdwell > 0)) fail(MM_ERR_ARG, "interpTrajTime: bad sizes or dwell time");
    if (nq == 0) return;
    const int64_t n = nq * ncol;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
    hipLaunchKernelGGL(interp_makima_k, dim3(grid), dim3(256), 0, st, phs, L, ncol, -delTime, tdwell, tq, nq, out);
    LAUNCHCHECK();
}

}  // namespace mm

// The one plain library product of the library: the Gram matrix S^H S of
// the dense encoding matrix (mrSampFuncMat.m:37), complex64 on rocBLAS.  The
// matrix is Hermitian, so rocBLAS forms its lower triangle (a rank-K
// Hermitian update: half the flops of the full product) and a tiled kernel
// writes the conjugate transpose into the upper triangle.
#include <rocblas/rocblas.h>

#include "hio.hpp"

namespace mm {

static void blas_check(rocblas_status s, const char* what) {
    if (s == rocblas_status_success) return;
    if (s == rocblas_status_memory_error) fail(MM_ERR_OOM, std::string(what) + ": rocBLAS out of memory");
    fail(MM_ERR_HIP, std::string(what) + ": rocBLAS status " + std::to_string((int)s));
}

void* blas_handle_create() {
    rocblas_handle h = nullptr;
    blas_check(rocblas_create_handle(&h), "rocblas_create_handle");
    return h;
}

void blas_handle_destroy(void* h) {
    if (h) (void)rocblas_destroy_handle((rocblas_handle)h);
}

// Upper triangle of the column-major Hermitian [n x n] matrix G from its
// lower triangle, G(j, i) = conj(G(i, j)) for i > j: a 32 x 32 lower tile
// (bi >= bj) is read down its columns and written, conjugated, along the
// rows of the mirror tile, both coalesced, through an LDS transpose.
__global__ __launch_bounds__(256) void herm_mirror_k(float2* __restrict__ G, int64_t n) {
    __shared__ float2 t[32][33];
    const int64_t bi = blockIdx.x, bj = blockIdx.y;
    if (bi < bj) return;
    const int tx = (int)threadIdx.x & 31, ty = (int)threadIdx.x >> 5;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t i = bi * 32 + tx, j = bj * 32 + ty + 8 * k;
        if (i < n && j < n) t[ty + 8 * k][tx] = G[i + j * n];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t j = bj * 32 + tx, i = bi * 32 + ty + 8 * k;
        if (i < n && j < n && i > j) {
            const float2 v = t[tx][ty + 8 * k];
            G[j + i * n] = make_float2(v.x, -v.y);
        }
    }
}

// G (column-major [n x n]) = S^H S, S column-major [m x n]
void dense_gram(void* blas, hipStream_t st, const float2* S, int64_t m, int64_t n, float2* G) {
    rocblas_handle h = (rocblas_handle)blas;
    blas_check(rocblas_set_stream(h, st), "rocblas_set_stream");
    const float one = 1.f, zero = 0.f;
    blas_check(rocblas_cherk(h, rocblas_fill_lower, rocblas_operation_conjugate_transpose, (rocblas_int)n,
                             (rocblas_int)m, &one, (const rocblas_float_complex*)S, (rocblas_int)m, &zero,
                             (rocblas_float_complex*)G, (rocblas_int)n),
               "rocblas_cherk");
    const unsigned nt = (unsigned)((n + 31) / 32);
    hipLaunchKernelGGL(herm_mirror_k, dim3(nt, nt), dim3(256), 0, st, G, n);
    LAUNCHCHECK();
}

}  // namespace mm

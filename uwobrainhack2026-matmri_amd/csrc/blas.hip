// The one plain library GEMM of the library: the Gram matrix S^H S of the
// dense encoding matrix (mrSampFuncMat.m:37), complex64 on rocBLAS.
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

// G (column-major [n x n]) = S^H S, S column-major [m x n]
void dense_gram(void* blas, hipStream_t st, const float2* S, int64_t m, int64_t n, float2* G) {
    rocblas_handle h = (rocblas_handle)blas;
    blas_check(rocblas_set_stream(h, st), "rocblas_set_stream");
    const rocblas_float_complex one{1.f, 0.f}, zero{0.f, 0.f};
    blas_check(rocblas_cgemm(h, rocblas_operation_conjugate_transpose, rocblas_operation_none, (rocblas_int)n,
                             (rocblas_int)n, (rocblas_int)m, &one, (const rocblas_float_complex*)S, (rocblas_int)m,
                             (const rocblas_float_complex*)S, (rocblas_int)m, &zero, (rocblas_float_complex*)G,
                             (rocblas_int)n),
               "rocblas_cgemm");
}

}  // namespace mm

// Shared runtime pieces of libmatmri_hip: error reporting, the device
// context, device buffers and the abstract linear operator.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/matmri_hip.h"

namespace mm {

// Error carrying one of the MM_ERR_* codes; converted to a return code at the
// C boundary (capi.cpp) and its message kept for mm_last_error().
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

inline void hip_check(hipError_t e, const char* what, const char* file, int line) {
    if (e == hipSuccess) return;
    (void)hipGetLastError();
    char buf[512];
    snprintf(buf, sizeof buf, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        throw Error(MM_ERR_OOM, std::string(buf) +
                    " -- reduce the batch (slices/repetitions) passed per call");
    throw Error(MM_ERR_HIP, buf);
}
#define HIPCHECK(x) ::mm::hip_check((x), #x, __FILE__, __LINE__)
#define LAUNCHCHECK() ::mm::hip_check(hipGetLastError(), "kernel launch", __FILE__, __LINE__)

// Counts device frees made through DevBuf.  A captured hipGraph holds raw
// device addresses; its owner records this count at capture and recaptures
// when it has moved (some buffer the graph may point at was freed).
inline std::atomic<uint64_t>& free_epoch() {
    static std::atomic<uint64_t> e{0};
    return e;
}

// Owning device allocation (hipMalloc / hipFree).
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        HIPCHECK(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    // grow-only reallocation (contents are not kept)
    void ensure(size_t count) { if (count > n) alloc(count); }
    void release() {
        if (p) {
            (void)hipFree(p);
            free_epoch().fetch_add(1, std::memory_order_relaxed);
        }
        p = nullptr;
        n = 0;
    }
    void upload(const T* h, size_t count, hipStream_t s) {
        ensure(count);
        if (count) HIPCHECK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& h, hipStream_t s) { upload(h.data(), h.size(), s); }
    T* get() const { return p; }
};

}  // namespace mm

struct mm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // staging for host-pointer calls
    mm::DevBuf<float2> stage_in, stage_out, stage_aux;
    // rocBLAS handle (created on first use by the dense-matrix path)
    void* blas = nullptr;
};

// Abstract linear operator.  Device entry points are asynchronous on the
// context stream; input and output never alias.
struct mm_op {
    mm_ctx* ctx = nullptr;
    virtual ~mm_op() = default;
    // complex elements per repetition: image side and data side
    virtual int64_t n_image() const = 0;
    virtual int64_t n_data() const = 0;
    // slices (independent problems) this operator spans; image/data vectors
    // are concatenations over slices
    virtual int nslice() const { return 1; }
    virtual int64_t slice_image_off(int s) const { (void)s; return 0; }
    virtual int64_t slice_data_off(int s) const { (void)s; return 0; }
    virtual void apply(int adjoint, const float2* in, float2* out, int64_t nrep) = 0;
    // data-side weights applied between forward and adjoint by apply_normal
    const float* W = nullptr;
    // x_out = A'(W .* A x); `tmp` holds n_data()*nrep elements.  Operators
    // with a fused normal path override this.
    virtual void apply_normal(const float2* in, float2* out, int64_t nrep, float2* tmp);
    // optional: accumulate per-slice partial FP64 dot products of
    // conj(in) . out into `dot` (nslice x 2 doubles, already zeroed) during
    // apply_normal.  Returns false if the operator cannot fuse it.
    // Per-slice skip flags (device, nslice ints) for apply_normal: slices
    // whose flag is nonzero may be left out (their output is not defined).
    // The solver installs its "slice finished" flags here; NULL = none.
    virtual void set_skip(const int* flags) { (void)flags; }
    // Configuration generation: bumped whenever a tuning option, the
    // Toeplitz form or the autotuner changes which kernels run or rebuilds
    // work tables.  Composite operators report the sum over their parts.
    uint64_t gen_ = 0;
    virtual uint64_t config_gen() const { return gen_; }
    // Tuning knobs (kernel variants); false if the key is unknown.
    virtual bool set_option(const std::string& key, int value) { (void)key; (void)value; return false; }
    virtual bool get_option(const std::string& key, int* value) const { (void)key; (void)value; return false; }
    // Time candidate tuning options of the normal operator on the device
    // (`iters` applications each) and keep the fastest; false if the
    // operator has nothing to tune.
    virtual bool autotune(int iters) { (void)iters; return false; }
    // The same for applications of nrep images at once, as a composite
    // operator applies this one (a SENSE operator: its coils); false if
    // there is nothing to tune.
    virtual bool autotune_reps(int64_t nrep, int iters) { (void)nrep; (void)iters; return false; }
    // Time the operator's dominant kernel: `iters` back-to-back launches on
    // the context stream between two HIP events.  Returns ms per launch and
    // the kernel's name; false if the operator has no single dominant kernel.
    virtual bool profile_dominant(int adjoint, const float2* in, float2* out, int64_t nrep,
                                  int iters, double* ms, std::string* name) {
        (void)adjoint; (void)in; (void)out; (void)nrep; (void)iters; (void)ms; (void)name;
        return false;
    }
};

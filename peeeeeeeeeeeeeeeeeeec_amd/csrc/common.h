// Shared host/device helpers for the gfx950 PEEC library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/complex_matrix.h"

namespace peec {

// Errors are reported the reference's way: an "ERROR: ..." line on stderr
// and a non-zero status to the caller.
#define PEEC_ERR(...)                                   \
    do {                                                \
        std::fprintf(stderr, "ERROR: " __VA_ARGS__);    \
        std::fprintf(stderr, "\n");                     \
    } while (0)

// Evaluate a HIP call; on failure print it and return `ret` from the caller.
#define HIP_OR_RETURN(call, ret)                                              \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            std::fprintf(stderr, "ERROR: HIP %s failed at %s:%d: %s\n", #call, \
                         __FILE__, __LINE__, hipGetErrorString(e_));          \
            return ret;                                                       \
        }                                                                     \
    } while (0)

// Throwing variant used inside C++ helpers; the C ABI catches and converts.
struct HipError {
    std::string what;
};

#define HIP_CHECK(call)                                                        \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            char buf_[512];                                                    \
            std::snprintf(buf_, sizeof buf_, "HIP %s failed at %s:%d: %s", #call, \
                          __FILE__, __LINE__, hipGetErrorString(e_));         \
            throw ::peec::HipError{buf_};                                      \
        }                                                                      \
    } while (0)

// Owning device buffer (hipMalloc / hipFree).
template <typename T>
struct DeviceBuffer {
    T *ptr = nullptr;
    size_t count = 0;
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t n) { resize(n); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    DeviceBuffer(DeviceBuffer &&o) noexcept : ptr(o.ptr), count(o.count) {
        o.ptr = nullptr;
        o.count = 0;
    }
    DeviceBuffer &operator=(DeviceBuffer &&o) noexcept {
        if (this != &o) {
            release();
            ptr = o.ptr;
            count = o.count;
            o.ptr = nullptr;
            o.count = 0;
        }
        return *this;
    }
    ~DeviceBuffer() { release(); }
    void resize(size_t n) {
        if (n == count && ptr) return;
        release();
        if (n) HIP_CHECK(hipMalloc(&ptr, n * sizeof(T)));
        count = n;
    }
    // grow only (keeps the allocation when it is large enough)
    void reserve(size_t n) {
        if (n > count) resize(n);
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    size_t bytes() const { return count * sizeof(T); }
    void upload(const T *host, size_t n, hipStream_t s) {
        resize(n);
        if (n) HIP_CHECK(hipMemcpyAsync(ptr, host, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T> &v, hipStream_t s) { upload(v.data(), v.size(), s); }
    void download(T *host, size_t n, hipStream_t s) const {
        if (n) HIP_CHECK(hipMemcpyAsync(host, ptr, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

// One stream per call keeps concurrent drop-in calls on distinct outputs
// independent (the reference functions are reentrant).
struct Stream {
    hipStream_t s = nullptr;
    Stream() { HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
    ~Stream() {
        if (s) (void)hipStreamDestroy(s);
    }
    Stream(const Stream &) = delete;
    Stream &operator=(const Stream &) = delete;
    operator hipStream_t() const { return s; }
    void sync() const { HIP_CHECK(hipStreamSynchronize(s)); }
};

// Device bytes an engine allocates per frequency point beyond its static
// Lp/P: the retarded complex Lp(w)/P(w), the FP64 MNA matrix and its
// unscaled copy, the FP32 factor, and a margin for the vectors, packed LU
// operands and block inverses.
inline size_t engine_point_bytes(size_t ne, size_t nv, bool retarded) {
    const size_t n = ne + nv, sq = ne * ne + nv * nv;
    size_t b = 32 * n * n + 8 * n * n;
    if (retarded) b += 16 * sq;
    return b + n * (size_t)64 * 1024 + ((size_t)64 << 20);
}

// Per-kernel timing with HIP events recorded on the launching stream; the
// elapsed times are resolved after the caller's next synchronisation.
enum KernelKind {
    kKernelLuUpdate = 0,
    kKernelPanel = 1,
    kKernelSmoothTile = 2,
    kKernelFarTile = 3,
    kKernelSmoothMoments = 4,
    kKernelLuUpdateSplit16 = 5,
    kKernelLuUpdateThin = 6,  // first panel of a pair: update of the next panel's rows only
    kKernelNearEntry = 7,     // static near-field entries
    kKernelLuUpdateMerged = 8,  // trailing updates of two panel pairs in one pass
    kKernelKinds = 9
};

struct KernelTotals {
    long long launches = 0;
    double ms = 0.0, flops = 0.0, bytes = 0.0;
};

class KernelTimer {
  public:
    bool enabled = false;
    KernelTotals totals[kKernelKinds];
    ~KernelTimer() {
        for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
    }
    // returns a ticket, or -1 when disabled
    int start(hipStream_t s) {
        if (!enabled) return -1;
        const size_t need = used_ + 2;
        while (pool_.size() < need) {
            hipEvent_t e;
            // timestamps only, resolved after a stream synchronisation
            HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
            pool_.push_back(e);
        }
        HIP_CHECK(hipEventRecord(pool_[used_], s));
        const int t = (int)used_;
        used_ += 2;
        return t;
    }
    void stop(int ticket, hipStream_t s, int kind, double flops, double bytes) {
        if (ticket < 0) return;
        HIP_CHECK(hipEventRecord(pool_[ticket + 1], s));
        pending_.push_back({ticket, kind, flops, bytes});
    }
    // call after the stream has been synchronised
    void resolve() {
        for (const Rec &r : pending_) {
            float ms = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms, pool_[r.ticket], pool_[r.ticket + 1]));
            KernelTotals &t = totals[r.kind];
            t.launches += 1;
            t.ms += ms;
            t.flops += r.flops;
            t.bytes += r.bytes;
        }
        pending_.clear();
        used_ = 0;
    }
    void reset() {
        pending_.clear();
        used_ = 0;
        for (KernelTotals &t : totals) t = KernelTotals{};
    }

  private:
    struct Rec {
        int ticket, kind;
        double flops, bytes;
    };
    std::vector<hipEvent_t> pool_;
    size_t used_ = 0;
    std::vector<Rec> pending_;
};

// Fails loudly when no GPU is visible: there is no CPU fallback.
bool require_device();

// Gauss-Legendre nodes/weights for the supported orders, computed once on
// the host from the Legendre three-term recurrence in binary128.
bool gauss_legendre(int order, double *x, double *w);
inline bool gl_order_supported(int o) {
    return o == 2 || o == 3 || o == 4 || o == 5 || o == 6 || o == 8 || o == 10 || o == 12 ||
           o == 16;
}

// Device constant table: nodes at [order][i], weights likewise.
struct GLTable {
    double x[17][16];
    double w[17][16];
};
const GLTable &gl_table();

}  // namespace peec

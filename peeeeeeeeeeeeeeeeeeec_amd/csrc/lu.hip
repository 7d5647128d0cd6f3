// Blocked right-looking LU on gfx950 (see lu.h). Per block column k0:
//   panel_kernel    cooperative grid: each workgroup keeps a slab of the
//                   panel's rows in LDS; per column one grid barrier decides
//                   the pivot (largest |a_ik|, first index on ties, as the
//                   reference's strict '>' scan in complex_lu.c:341-360),
//                   exchanges the pivot row through a small HBM slot and
//                   applies the rank-1 update to the LDS slab.
//   swap_*_kernel   apply the panel's row swaps to the other columns.
//   trsm_kernel     U12 = L11^{-1} A12, columns in registers, 8 lanes per
//                   column.
//   gemm_kernel     A22 -= L21 U12 on MFMA (v_mfma_f64_16x16x4f64 for FP64,
//                   v_mfma_f32_16x16x4f32 for the mixed-precision factor).
// Triangular solves are blocked (GEMV + diagonal-block kernels).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <utility>
#include <vector>

#include "lu.h"
#include "mna.h"

namespace peec {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kMaxNB = 128;
constexpr unsigned kSpinLimit = 1u << 22;  // bounded grid-barrier wait (a few seconds)
constexpr size_t kBarWords = 64;


// ---------------------------------------------------------------------------
// element helpers: storage S, arithmetic type Elem<S>::A (always FP64)

template <class S> struct Elem;
template <> struct Elem<double2> { using A = double2; static constexpr bool cplx = true; };
template <> struct Elem<float2> { using A = double2; static constexpr bool cplx = true; };
template <> struct Elem<double> { using A = double; static constexpr bool cplx = false; };

__device__ inline double2 ld(const double2 *p) { return *p; }
__device__ inline double2 ld(const float2 *p) {
    const float2 v = *p;
    return make_double2((double)v.x, (double)v.y);
}
__device__ inline double ld(const double *p) { return *p; }
__device__ inline void st(double2 *p, double2 v) { *p = v; }
__device__ inline void st(float2 *p, double2 v) { *p = make_float2((float)v.x, (float)v.y); }
__device__ inline void st(double *p, double v) { *p = v; }
__device__ inline double mag(double2 v) { return hypot(v.x, v.y); }
__device__ inline double mag(double v) { return fabs(v); }
// complex_lu.c:162-184 (naive formula)
__device__ inline double2 ediv(double2 a, double2 b) {
    const double den = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den);
}
__device__ inline double ediv(double a, double b) { return a / b; }
// c - l*u with the product formed first (complex_lu.c:482-498)
__device__ inline double2 efms(double2 c, double2 l, double2 u) {
    const double pr = l.x * u.x - l.y * u.y;
    const double pi = l.x * u.y + l.y * u.x;
    return make_double2(c.x - pr, c.y - pi);
}
__device__ inline double efms(double c, double l, double u) { return c - l * u; }
__device__ inline double2 emul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double emul(double a, double b) { return a * b; }
__device__ inline double2 eadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double eadd(double a, double b) { return a + b; }
__device__ inline double2 esub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double esub(double a, double b) { return a - b; }
template <class T> __device__ inline T ezero();
template <> __device__ inline double2 ezero<double2>() { return make_double2(0.0, 0.0); }
template <> __device__ inline double ezero<double>() { return 0.0; }
__device__ inline double2 eshfl_xor(double2 v, int m) {
    return make_double2(__shfl_xor(v.x, m), __shfl_xor(v.y, m));
}
__device__ inline double eshfl_xor(double v, int m) { return __shfl_xor(v, m); }
// Register array of N arithmetic elements. Complex values are kept as two
// plain double arrays: arrays of HIP vector types are not split into
// registers by the compiler and end up in scratch memory.
template <class T, int N> struct Regs;
template <int N> struct Regs<double2, N> {
    double re[N], im[N];
    __device__ double2 get(int i) const { return make_double2(re[i], im[i]); }
    __device__ void set(int i, double2 v) {
        re[i] = v.x;
        im[i] = v.y;
    }
};
template <int N> struct Regs<double, N> {
    double v[N];
    __device__ double get(int i) const { return v[i]; }
    __device__ void set(int i, double x) { v[i] = x; }
};

template <int N> struct Regs<float2, N> {
    float re[N], im[N];
    __device__ float2 get(int i) const { return make_float2(re[i], im[i]); }
    __device__ void set(int i, float2 v) {
        re[i] = v.x;
        im[i] = v.y;
    }
};

// Working precision of the triangular-solve and diagonal-block kernels: the
// storage precision (the FP32 factor is formed in FP32 arithmetic, like its
// MFMA trailing update), with fused multiply-adds.
template <class S> struct Work { using T = S; };
__device__ inline float2 wld(const float2 *p) { return *p; }
__device__ inline double2 wld(const double2 *p) { return *p; }
__device__ inline double wld(const double *p) { return *p; }
__device__ inline void wst(float2 *p, float2 v) { *p = v; }
__device__ inline void wst(double2 *p, double2 v) { *p = v; }
__device__ inline void wst(double *p, double v) { *p = v; }
template <class T> __device__ inline T wzero();
template <> __device__ inline float2 wzero<float2>() { return make_float2(0.f, 0.f); }
template <> __device__ inline double2 wzero<double2>() { return make_double2(0.0, 0.0); }
template <> __device__ inline double wzero<double>() { return 0.0; }
template <class T> __device__ inline T wone();
template <> __device__ inline float2 wone<float2>() { return make_float2(1.f, 0.f); }
template <> __device__ inline double2 wone<double2>() { return make_double2(1.0, 0.0); }
template <> __device__ inline double wone<double>() { return 1.0; }
// c - a b
__device__ inline float2 wfms(float2 c, float2 a, float2 b) {
    return make_float2(fmaf(-a.x, b.x, fmaf(a.y, b.y, c.x)), fmaf(-a.x, b.y, fmaf(-a.y, b.x, c.y)));
}
__device__ inline double2 wfms(double2 c, double2 a, double2 b) {
    return make_double2(fma(-a.x, b.x, fma(a.y, b.y, c.x)), fma(-a.x, b.y, fma(-a.y, b.x, c.y)));
}
__device__ inline double wfms(double c, double a, double b) { return fma(-a, b, c); }
__device__ inline float2 wmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ inline double2 wmul(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ inline double wmul(double a, double b) { return a * b; }
// a / b by the naive formula (complex_lu.c:162-184)
__device__ inline float2 wdiv(float2 a, float2 b) {
    const float den = b.x * b.x + b.y * b.y;
    return make_float2((a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den);
}
__device__ inline double2 wdiv(double2 a, double2 b) { return ediv(a, b); }
__device__ inline double wdiv(double a, double b) { return a / b; }
__device__ inline double wmag(float2 v) { return hypot((double)v.x, (double)v.y); }
__device__ inline double wmag(double2 v) { return hypot(v.x, v.y); }
__device__ inline double wmag(double v) { return fabs(v); }
__device__ inline float2 wshfl(float2 v, int src) { return make_float2(__shfl(v.x, src), __shfl(v.y, src)); }
__device__ inline double2 wshfl(double2 v, int src) { return make_double2(__shfl(v.x, src), __shfl(v.y, src)); }
__device__ inline double wshfl(double v, int src) { return __shfl(v, src); }
// value of lane src (wave-uniform) of a wavefront
__device__ inline float rdlane(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ inline double rdlane(double v, int src) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ inline float2 wreadlane(float2 v, int src) { return make_float2(rdlane(v.x, src), rdlane(v.y, src)); }
__device__ inline double2 wreadlane(double2 v, int src) { return make_double2(rdlane(v.x, src), rdlane(v.y, src)); }
__device__ inline double wreadlane(double v, int src) { return rdlane(v, src); }

__device__ inline double2 eshfl(double2 v, int src) { return make_double2(__shfl(v.x, src), __shfl(v.y, src)); }
__device__ inline double eshfl(double v, int src) { return __shfl(v, src); }

// ---------------------------------------------------------------------------
// equilibration (complex_lu.c:275-320)

// Entry sources of the read-only passes over a complex matrix: a dense
// row-major array, or the MNA matrix evaluated from its blocks (the same
// value per entry as the dense copy the fill kernel would write).
struct DenseSource {
    const double2 *A;
    size_t lda;
    __device__ double2 operator()(int i, int j) const { return A[(size_t)i * lda + j]; }
};
struct MnaSource {
    MnaOperands m;
    __device__ double2 operator()(int i, int j) const { return mna_entry(i, j, m); }
};

template <class Src>
__global__ void row_max_kernel(Src src, int n, double *row, int *bad) {
    __shared__ double red[256];
    const int i = blockIdx.x;
    double m = 0.0;
    bool finite = true;
    for (int j = threadIdx.x; j < n; j += 256) {
        const double v = mag(src(i, j));
        if (!isfinite(v)) finite = false;
        m = fmax(m, v);
    }
    red[threadIdx.x] = finite ? m : INFINITY;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double r = red[0];
        if (!isfinite(r) || r == 0.0) atomicOr(bad, 1);
        row[i] = r;
    }
}

// positive doubles order like their bit patterns: atomic max on the bits
__device__ inline void atomic_max_pos(double *addr, double v) {
    atomicMax(reinterpret_cast<unsigned long long *>(addr), (unsigned long long)__double_as_longlong(v));
}

// divide rows by r_i; column maxima of the row-scaled matrix
__global__ void scale_rows_colmax_kernel(double2 *A, int n, size_t lda, const double *row, double *col) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const int i0 = blockIdx.y * 64, i1 = min(n, i0 + 64);
    double m = 0.0;
    for (int i = i0; i < i1; ++i) {
        double2 v = A[(size_t)i * lda + j];
        const double r = row[i];
        v.x = v.x / r;
        v.y = v.y / r;
        A[(size_t)i * lda + j] = v;
        m = fmax(m, mag(v));
    }
    atomic_max_pos(&col[j], m);
}

__global__ void scale_cols_kernel(double2 *A, int n, size_t lda, const double *col, double *maxv, int *bad) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const double c = col[j];
    if (blockIdx.y == 0 && c == 0.0) atomicOr(bad, 1);
    const int i0 = blockIdx.y * 64, i1 = min(n, i0 + 64);
    double m = 0.0;
    for (int i = i0; i < i1; ++i) {
        double2 v = A[(size_t)i * lda + j];
        v.x /= c;
        v.y /= c;
        A[(size_t)i * lda + j] = v;
        m = fmax(m, mag(v));
    }
    atomic_max_pos(maxv, m);
}

// column maxima of the row-scaled matrix, |a_ij / r_i| (the arithmetic of
// scale_rows_colmax_kernel), without writing the matrix
template <class Src>
__global__ void colmax_scaled_kernel(Src src, int n, const double *row, double *col) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const int i0 = blockIdx.y * 64, i1 = min(n, i0 + 64);
    double m = 0.0;
    for (int i = i0; i < i1; ++i) {
        double2 v = src(i, j);
        const double r = row[i];
        v.x = v.x / r;
        v.y = v.y / r;
        m = fmax(m, mag(v));
    }
    atomic_max_pos(&col[j], m);
}

// the equilibrated entry (a_ij / r_i) / c_j, divided in the order of
// scale_rows_colmax_kernel and scale_cols_kernel
__device__ inline double2 equilibrated(double2 v, double r, double c) {
    v.x = v.x / r;
    v.y = v.y / r;
    v.x /= c;
    v.y /= c;
    return v;
}

// dst row i (FP32) = equilibrated row perm[i] (or i) of src, formed as
// (a_ij * (1 / r_i)) * (1 / c_j) with the scales' reciprocals: within a few
// FP64 ulps of the divided value (a_ij / r_i) / c_j, which the rounding to
// FP32 absorbs, at a fraction of four FP64 divisions per entry. The largest
// |equilibrated entry| (the root of the largest squared magnitude; entries
// are at most 1 in magnitude) into maxv[0 .. kMaxSlots). One workgroup per
// (row, 1024 columns).
constexpr int kMaxSlots = 1024;
template <class Src>
__global__ __launch_bounds__(256) void to_f32_equilibrated_kernel(Src src, float2 *dst, const int *perm,
                                                                  const double *row, const double *inv_col, int n,
                                                                  double *maxv) {
    __shared__ double red[4];
    // four entries per thread (1024 per workgroup), loads issued first
    const int i = blockIdx.y, c0 = blockIdx.x * 1024 + threadIdx.x;
    const int si = perm ? perm[i] : i;
    const double ri = 1.0 / row[si];
    double2 a[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) a[u] = c0 + 256 * u < n ? src(si, c0 + 256 * u) : double2{};
    double m = 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int c = c0 + 256 * u;
        if (c < n) {
            const double ic = inv_col[c];
            const double vx = (a[u].x * ri) * ic, vy = (a[u].y * ri) * ic;
            dst[(size_t)i * n + c] = make_float2((float)vx, (float)vy);
            m = fmax(m, vx * vx + vy * vy);
        }
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    // one of kMaxSlots slots per workgroup: no single contended address
    if (threadIdx.x == 0)
        atomic_max_pos(&maxv[(blockIdx.x + (size_t)blockIdx.y * gridDim.x) % kMaxSlots],
                       sqrt(fmax(fmax(red[0], red[1]), fmax(red[2], red[3]))));
}

// inv[j] = 1 / col[j]; a zero scale is reported in *bad
__global__ void inverse_scale_kernel(const double *col, int n, double *inv, int *bad) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    if (col[j] == 0.0) atomicOr(bad, 1);
    inv[j] = 1.0 / col[j];
}

// dst (FP64) = the equilibrated src
__global__ void equilibrated_copy_kernel(const double2 *src, double2 *dst, const double *row, const double *col,
                                         int n) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)n * n) return;
    const int i = (int)(idx / n), c = (int)(idx % n);
    dst[idx] = equilibrated(src[idx], row[i], col[c]);
}

__global__ void fill_zero_kernel(double *p, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0.0;
}
__global__ void fill_int_kernel(int *p, int n, int v) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void fill_uint_kernel(unsigned *p, int n, unsigned v) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

LUContext::LUContext(hipStream_t s) : stream(s) {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    barrier.resize(kBarWords);
    status.resize(2);
    solve_bad.resize(1);
    slots.resize((size_t)2 * (cus + 1) * kMaxNB * 2);
    cand.resize((size_t)2 * cus * 2);
    tags.resize((size_t)2 * cus);
    HIP_CHECK(hipMemset(tags.ptr, 0, tags.bytes()));
    swap_plan.resize((size_t)1 + 4 * kMaxNB);
    HIP_CHECK(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking));
    // ordering between the two streams of one device only: no timing, no
    // system-scope fence (device-scope release/acquire orders the work)
    HIP_CHECK(hipEventCreateWithFlags(&ev_next, hipEventDisableTiming | hipEventDisableSystemFence));
    HIP_CHECK(hipEventCreateWithFlags(&ev_diag, hipEventDisableTiming | hipEventDisableSystemFence));
    HIP_CHECK(hipEventCreateWithFlags(&ev_prev, hipEventDisableTiming | hipEventDisableSystemFence));
    HIP_CHECK(hipEventCreateWithFlags(&ev_u12, hipEventDisableTiming | hipEventDisableSystemFence));
    HIP_CHECK(hipDeviceGetAttribute(&lds_per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
}

LUContext::~LUContext() {
    if (ev_next) (void)hipEventDestroy(ev_next);
    if (ev_diag) (void)hipEventDestroy(ev_diag);
    if (ev_prev) (void)hipEventDestroy(ev_prev);
    if (ev_u12) (void)hipEventDestroy(ev_u12);
    if (aux) (void)hipStreamDestroy(aux);
}

int equilibrate_complex(double2 *dA, int n, size_t lda, double *d_row, double *d_col,
                        double *max_scaled, LUContext &ctx) {
    hipStream_t s = ctx.stream;
    DeviceBuffer<int> bad(1);
    DeviceBuffer<double> mx(1);
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, s, bad.ptr, 1, 0);
    hipLaunchKernelGGL(fill_zero_kernel, dim3(1), dim3(256), 0, s, mx.ptr, (size_t)1);
    hipLaunchKernelGGL(fill_zero_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_col, (size_t)n);
    hipLaunchKernelGGL(row_max_kernel<DenseSource>, dim3(n), dim3(256), 0, s, DenseSource{dA, lda}, n, d_row, bad.ptr);
    int h_bad = 0;
    bad.download(&h_bad, 1, s);
    HIP_CHECK(hipStreamSynchronize(s));
    if (h_bad) return -2;
    dim3 grid((n + 255) / 256, (n + 63) / 64);
    hipLaunchKernelGGL(scale_rows_colmax_kernel, grid, dim3(256), 0, s, dA, n, lda, d_row, d_col);
    hipLaunchKernelGGL(scale_cols_kernel, grid, dim3(256), 0, s, dA, n, lda, d_col, mx.ptr, bad.ptr);
    HIP_CHECK(hipGetLastError());
    bad.download(&h_bad, 1, s);
    mx.download(max_scaled, 1, s);
    HIP_CHECK(hipStreamSynchronize(s));
    if (h_bad) return -2;
    return 0;
}

__global__ void row_check_kernel(const double *row, int n, int *bad) {
    const int i = (int)blockIdx.x * 256 + threadIdx.x;
    if (i < n && (!isfinite(row[i]) || row[i] == 0.0)) atomicOr(bad, 1);
}

void zero_doubles(double *p, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
}

template <class Src>
static int equilibration_scales_of(Src src, int n, double *d_row, double *d_col, LUContext &ctx, bool rows_ready) {
    hipStream_t s = ctx.stream;
    DeviceBuffer<int> bad(1);
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, s, bad.ptr, 1, 0);
    hipLaunchKernelGGL(fill_zero_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_col, (size_t)n);
    if (rows_ready)
        hipLaunchKernelGGL(row_check_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_row, n, bad.ptr);
    else
        hipLaunchKernelGGL(row_max_kernel<Src>, dim3(n), dim3(256), 0, s, src, n, d_row, bad.ptr);
    dim3 grid((n + 255) / 256, (n + 63) / 64);
    hipLaunchKernelGGL(colmax_scaled_kernel<Src>, grid, dim3(256), 0, s, src, n, d_row, d_col);
    HIP_CHECK(hipGetLastError());
    int h_bad = 0;
    bad.download(&h_bad, 1, s);
    HIP_CHECK(hipStreamSynchronize(s));
    return h_bad ? -2 : 0;
}
int equilibration_scales(const double2 *dA, int n, size_t lda, double *d_row, double *d_col, LUContext &ctx,
                         bool rows_ready) {
    return equilibration_scales_of(DenseSource{dA, lda}, n, d_row, d_col, ctx, rows_ready);
}
int equilibration_scales(const MnaOperands &m, double *d_row, double *d_col, LUContext &ctx) {
    return equilibration_scales_of(MnaSource{m}, m.ne + m.nv, d_row, d_col, ctx, false);
}

template <class Src>
static int convert_equilibrated_to_f32_of(Src src, float2 *dst, const int *d_perm, const double *d_row,
                                          const double *d_col, int n, double *max_scaled, LUContext &ctx) {
    hipStream_t s = ctx.stream;
    // slot maxima, then the column scales' reciprocals
    DeviceBuffer<double> mx(kMaxSlots + (size_t)n);
    hipLaunchKernelGGL(fill_zero_kernel, dim3(kMaxSlots / 256), dim3(256), 0, s, mx.ptr, (size_t)kMaxSlots);
    // zero column maxima would divide by zero: reported like the in-place path
    DeviceBuffer<int> bad(1);
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, s, bad.ptr, 1, 0);
    hipLaunchKernelGGL(inverse_scale_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_col, n, mx.ptr + kMaxSlots,
                       bad.ptr);
    hipLaunchKernelGGL(to_f32_equilibrated_kernel<Src>, dim3((unsigned)((n + 1023) / 1024), (unsigned)n), dim3(256), 0,
                       s, src, dst, d_perm, d_row, mx.ptr + kMaxSlots, n, mx.ptr);
    HIP_CHECK(hipGetLastError());
    int h_bad = 0;
    std::vector<double> slots(kMaxSlots);
    bad.download(&h_bad, 1, s);
    mx.download(slots.data(), kMaxSlots, s);
    HIP_CHECK(hipStreamSynchronize(s));
    *max_scaled = *std::max_element(slots.begin(), slots.end());
    return h_bad ? -2 : 0;
}
int convert_equilibrated_to_f32(const double2 *dA, float2 *dst, const int *d_perm, const double *d_row,
                                const double *d_col, int n, double *max_scaled, LUContext &ctx) {
    return convert_equilibrated_to_f32_of(DenseSource{dA, (size_t)n}, dst, d_perm, d_row, d_col, n, max_scaled, ctx);
}
int convert_equilibrated_to_f32(const MnaOperands &m, float2 *dst, const int *d_perm, const double *d_row,
                                const double *d_col, double *max_scaled, LUContext &ctx) {
    return convert_equilibrated_to_f32_of(MnaSource{m}, dst, d_perm, d_row, d_col, m.ne + m.nv, max_scaled, ctx);
}

void copy_equilibrated(const double2 *src, double2 *dst, const double *d_row, const double *d_col, int n,
                       hipStream_t s) {
    const size_t total = (size_t)n * n;
    hipLaunchKernelGGL(equilibrated_copy_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, src, dst,
                       d_row, d_col, n);
    HIP_CHECK(hipGetLastError());
}

// Grid barrier for cooperative kernels (called by one thread per
// workgroup): one release fence, one arrival on a flat counter, polling with
// relaxed agent-scope loads (no cache maintenance per poll) and a single
// acquire fence once the count is reached. The counter restarts at every
// launch; barrier t completes at t * G arrivals.
__device__ inline bool grid_barrier(unsigned *bar, unsigned target) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned spins = 0;
    while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > kSpinLimit) return false;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return true;
}

// ---------------------------------------------------------------------------
// panel factorization (cooperative)
//
// Pivot exchange without fences: workgroup g publishes, for column j, its
// candidate row into a double-buffered slot and its candidate magnitude with
// agent-scope atomic (cache-bypassing) stores, waits for those stores to
// complete, then publishes a 64-bit tag (round << 32 | row). Readers poll all
// G tags in parallel (one per thread) and then read the winner's row. The
// owner of row k publishes that row too (needed by the owner of the pivot
// row for the swap). Rounds increase monotonically per LU context, so tags
// never need resetting; the double buffer is safe because nobody can write
// column j+2 before every workgroup has published column j+1, which it does
// only after reading column j.

__device__ inline void pub(double *p, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long *>(p), (unsigned long long)__double_as_longlong(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double sub(const double *p) {
    return __longlong_as_double((long long)__hip_atomic_load(reinterpret_cast<const unsigned long long *>(p),
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ inline void pub_elem(double *p, double2 v) {
    pub(p, v.x);
    pub(p + 1, v.y);
}
__device__ inline void pub_elem(double *p, double v) { pub(p, v); }
__device__ inline void sub_elem(const double *p, double2 &v) {
    v.x = sub(p);
    v.y = sub(p + 1);
}
__device__ inline void sub_elem(const double *p, double &v) { v = sub(p); }

// Butterfly arg-max over a wavefront: largest magnitude, lowest row on
// ties (the order of complex_lu.c:341-360); every lane ends with the result.
__device__ inline void wave_argmax(double &m, int &i, int &g) {
    for (int off = 32; off > 0; off >>= 1) {
        const double om = __shfl_xor(m, off);
        const int oi = __shfl_xor(i, off);
        const int og = __shfl_xor(g, off);
        if (om > m || (om == m && oi < i)) {
            m = om;
            i = oi;
            g = og;
        }
    }
}

template <class S>
__global__ __launch_bounds__(256) void panel_kernel(S *A, size_t lda, int n, int k0, int nb, int rpw, double *slots,
                                                    double *mags, unsigned long long *tags, int *ipiv,
                                                    unsigned round0, double limit, int *status) {
    using T = typename Elem<S>::A;
    constexpr int W = sizeof(T) / sizeof(double);  // doubles per element
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    T *rows = reinterpret_cast<T *>(smem_raw);
    __shared__ T u[kMaxNB];
    __shared__ double wl_m[4], wg_m[4];
    __shared__ int wl_i[4], wg_i[4], wg_g[4];
    const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    if (status[0] != 0) return;
    const int r0 = k0 + g * rpw;
    const int r1 = min(n, r0 + rpw);
    const int nr = max(0, r1 - r0);
    for (int idx = tid; idx < nr * nb; idx += 256) {
        const int rr = idx / nb, c = idx - rr * nb;
        rows[idx] = ld(A + (size_t)(r0 + rr) * lda + k0 + c);
    }
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        const int k = k0 + j, buf = j & 1;
        const unsigned round = round0 + (unsigned)j + 1u;
        // local candidate: largest magnitude, lowest row on ties
        double bm = -1.0;
        int bi = INT_MAX;
        for (int rr = tid; rr < nr; rr += 256) {
            const int row = r0 + rr;
            if (row < k) continue;
            const double m = mag(rows[rr * nb + j]);
            if (m > bm || (m == bm && row < bi)) {
                bm = m;
                bi = row;
            }
        }
        {
            int unused = 0;
            wave_argmax(bm, bi, unused);
            if (lane == 0) {
                wl_m[wave] = bm;
                wl_i[wave] = bi;
            }
            __syncthreads();
            bm = wl_m[0];
            bi = wl_i[0];
            for (int q = 1; q < 4; ++q)
                if (wl_m[q] > bm || (wl_m[q] == bm && wl_i[q] < bi)) {
                    bm = wl_m[q];
                    bi = wl_i[q];
                }
        }
        // publish candidate row (and row k if owned), then the tag
        double *slot_base = slots + (size_t)buf * (G + 1) * kMaxNB * W;
        if (bi != INT_MAX)
            for (int c = tid; c < nb; c += 256) pub_elem(slot_base + ((size_t)g * kMaxNB + c) * W, rows[(bi - r0) * nb + c]);
        if (k >= r0 && k < r1)
            for (int c = tid; c < nb; c += 256) pub_elem(slot_base + ((size_t)G * kMaxNB + c) * W, rows[(k - r0) * nb + c]);
        if (tid == 0) pub(&mags[(size_t)buf * G + g], bm);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0)
            __hip_atomic_store(&tags[(size_t)buf * G + g],
                               ((unsigned long long)round << 32) | (unsigned)(bi == INT_MAX ? 0xffffffffu : (unsigned)bi),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // gather every workgroup's candidate (one per thread), bounded wait
        double best;
        int p, gw;
        {
            double m = -1.0;
            int idx = INT_MAX;
            int timeout = 0;
            if (tid < G) {
                unsigned long long tg;
                unsigned spins = 0;
                while (((tg = __hip_atomic_load(&tags[(size_t)buf * G + tid], __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)) >> 32) != round) {
                    __builtin_amdgcn_s_sleep(1);
                    if (++spins > kSpinLimit) {
                        timeout = 1;
                        break;
                    }
                }
                const unsigned low = (unsigned)(tg & 0xffffffffu);
                if (!timeout && low != 0xffffffffu) {
                    idx = (int)low;
                    m = sub(&mags[(size_t)buf * G + tid]);
                }
            }
            int src = tid;
            wave_argmax(m, idx, src);
            if (lane == 0) {
                wg_m[wave] = m;
                wg_i[wave] = idx;
                wg_g[wave] = src;
            }
            if (__syncthreads_or(timeout)) {
                if (tid == 0) status[0] = -3;
                return;
            }
            best = wg_m[0];
            p = wg_i[0];
            gw = wg_g[0];
            for (int q = 1; q < 4; ++q)
                if (wg_m[q] > best || (wg_m[q] == best && wg_i[q] < p)) {
                    best = wg_m[q];
                    p = wg_i[q];
                    gw = wg_g[q];
                }
        }
        if (!(best > limit) || !isfinite(best) || p == INT_MAX) {
            if (g == 0 && tid == 0) {
                status[0] = -2;
                status[1] = k;
            }
            return;
        }
        for (int c = tid; c < nb; c += 256) sub_elem(slot_base + ((size_t)gw * kMaxNB + c) * W, u[c]);
        __syncthreads();
        if (p != k) {
            if (p >= r0 && p < r1)
                for (int c = tid; c < nb; c += 256) sub_elem(slot_base + ((size_t)G * kMaxNB + c) * W, rows[(p - r0) * nb + c]);
            if (k >= r0 && k < r1)
                for (int c = tid; c < nb; c += 256) rows[(k - r0) * nb + c] = u[c];
        }
        if (g == 0 && tid == 0) ipiv[k] = p;
        __syncthreads();
        // multipliers l_ik = a_ik / u_kk and rank-1 update of the slab
        const int first = max(0, k + 1 - r0);
        for (int rr = first + tid; rr < nr; rr += 256) {
            const T l = ediv(rows[rr * nb + j], u[j]);
            rows[rr * nb + j] = l;
        }
        __syncthreads();
        const int width = nb - j - 1;
        if (width > 0) {
            for (int idx = tid; idx < (nr - first) * width; idx += 256) {
                const int rr = first + idx / width, c = j + 1 + idx % width;
                rows[rr * nb + c] = efms(rows[rr * nb + c], rows[rr * nb + j], u[c]);
            }
        }
        __syncthreads();
    }
    for (int idx = tid; idx < nr * nb; idx += 256) {
        const int rr = idx / nb, c = idx - rr * nb;
        st(A + (size_t)(r0 + rr) * lda + k0 + c, rows[idx]);
    }
}

// Row swaps of one panel on every other column, in two parallel passes.
// swap_plan_kernel replays the panel's sequential swaps (ipiv[k0 + t] with
// k0 + t, t = 0..w-1, as complex_lu.c:365-380 applies them) on the at most
// 2w affected positions and emits the resulting moves (dst <- src);
// swap_gather_kernel copies every moved row's source to a staging buffer and
// swap_scatter_kernel writes it to its destination, so no row is read after
// it has been overwritten.
__global__ __launch_bounds__(256) void swap_plan_kernel(const int *ipiv, int k0, int w, int *plan,
                                                        const int *status) {
    __shared__ int pos[2 * kMaxNB], content[2 * kMaxNB], pslot[kMaxNB], fo[kMaxNB], slot_first[kMaxNB];
    __shared__ int piv[kMaxNB];
    __shared__ int s_count, s_extra;
    if (status[0] != 0) return;
    const int tid = threadIdx.x;
    if (tid < w) piv[tid] = ipiv[k0 + tid];
    __syncthreads();
    int p = 0;
    if (tid < w) {
        p = piv[tid];
        int f = -1;
        if (p >= k0 + w) {
            f = tid;
            for (int t = 0; t < tid; ++t)
                if (piv[t] == p) {
                    f = t;
                    break;
                }
        }
        fo[tid] = f;
        pos[tid] = k0 + tid;
    }
    __syncthreads();
    if (tid == 0) {
        int extra = 0;
        for (int t = 0; t < w; ++t)
            if (fo[t] == t) {
                slot_first[t] = w + extra;
                pos[w + extra] = piv[t];
                ++extra;
            }
        s_extra = extra;
        s_count = 0;
    }
    __syncthreads();
    if (tid < w) pslot[tid] = p < k0 + w ? p - k0 : slot_first[fo[tid]];
    const int slots = w + s_extra;
    for (int q = tid; q < slots; q += 256) content[q] = pos[q];
    __syncthreads();
    if (tid == 0)
        for (int t = 0; t < w; ++t) {
            const int q = pslot[t];
            const int c = content[t];
            content[t] = content[q];
            content[q] = c;
        }
    __syncthreads();
    for (int q = tid; q < slots; q += 256)
        if (content[q] != pos[q]) {
            const int i = atomicAdd(&s_count, 1);
            plan[1 + 2 * i] = pos[q];
            plan[2 + 2 * i] = content[q];
        }
    __syncthreads();
    if (tid == 0) plan[0] = s_count;
}

// column c of the n - w columns outside the panel [k0, k0 + w)
__device__ inline int swap_column(int c, int k0, int w) { return c < k0 ? c : c + w; }

template <class S>
__global__ void swap_gather_kernel(const S *A, size_t lda, int n, int k0, int w, const int *plan, S *stage,
                                   const int *status) {
    if (status[0] != 0) return;
    const int q = blockIdx.y;
    if (q >= plan[0]) return;
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n - w) return;
    stage[(size_t)q * n + c] = A[(size_t)plan[2 + 2 * q] * lda + swap_column(c, k0, w)];
}

template <class S>
__global__ void swap_scatter_kernel(S *A, size_t lda, int n, int k0, int w, const int *plan, const S *stage,
                                    const int *status) {
    if (status[0] != 0) return;
    const int q = blockIdx.y;
    if (q >= plan[0]) return;
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n - w) return;
    A[(size_t)plan[1 + 2 * q] * lda + swap_column(c, k0, w)] = stage[(size_t)q * n + c];
}

// U12 = L11^{-1} A12 by forward substitution with the unit-lower L11
// (complex_lu.c:598-632 order per column). The strictly lower part of L11
// is staged once per workgroup in LDS (packed by rows, in the storage
// type). Sixteen lanes share one column, lane `sub` keeping rows
// 8 sub .. 8 sub + 7 of it in registers; at step r the owner of row r
// broadcasts x_r with a shuffle and every lane updates its rows below r.
constexpr int kTrsmTPC = 16;                  // lanes per column
constexpr int kTrsmRPT = kMaxNB / kTrsmTPC;  // rows per lane

__host__ __device__ constexpr size_t trsm_lds_elems(int w) { return (size_t)w * (w - 1) / 2; }

template <class S>
__global__ __launch_bounds__(256) void trsm_kernel(S *A, size_t lda, int c0, int c1, int k0, int w,
                                                   const int *status) {
    using T = typename Work<S>::T;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    S *lw = reinterpret_cast<S *>(smem_raw);  // L11[i][t], t < i, at i(i-1)/2 + t
    if (status[0] != 0) return;
    {
        // flat pass over the packed triangle, 8 independent loads in flight
        const int total = (int)trsm_lds_elems(w);
        for (int base = threadIdx.x; base < total; base += 256 * 8) {
            Regs<T, 8> v;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int idx = base + u * 256;
                v.set(u, wzero<T>());
                if (idx < total) {
                    int i = (int)((1.0 + sqrt(1.0 + 8.0 * idx)) * 0.5);
                    while (i * (i - 1) / 2 > idx) --i;
                    while ((i + 1) * i / 2 <= idx) ++i;
                    v.set(u, wld(A + (size_t)(k0 + i) * lda + k0 + (idx - i * (i - 1) / 2)));
                }
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (base + u * 256 < total) wst(lw + base + u * 256, v.get(u));
        }
    }
    const int lane = threadIdx.x & 63, sub = lane & (kTrsmTPC - 1);
    const int col = c0 + (int)blockIdx.x * (256 / kTrsmTPC) + (int)threadIdx.x / kTrsmTPC;
    const bool valid = col < c1;
    const int base = lane & ~(kTrsmTPC - 1);
    Regs<T, kTrsmRPT> a;
#pragma unroll
    for (int t = 0; t < kTrsmRPT; ++t) {
        const int i = sub * kTrsmRPT + t;
        a.set(t, (valid && i < w) ? wld(A + (size_t)(k0 + i) * lda + col) : wzero<T>());
    }
    __syncthreads();
    const int last = max(0, (int)trsm_lds_elems(w) - 1);
    const int nblk = (w + kTrsmRPT - 1) / kTrsmRPT;
    for (int rb = 0; rb < nblk; ++rb) {
#pragma unroll
        for (int rr = 0; rr < kTrsmRPT; ++rr) {
            const int r = min(rb * kTrsmRPT + rr, w - 1);
            const T x = wshfl(a.get(rr), base + rb);
            // lanes whose rows all lie above r have nothing left to do
            if (sub * kTrsmRPT + kTrsmRPT - 1 <= r) continue;
            // branch-free inside: every lane reads an in-bounds element and
            // keeps the update only for its rows below r
#pragma unroll
            for (int t = 0; t < kTrsmRPT; ++t) {
                const int i = sub * kTrsmRPT + t;
                const int ic = min(i, w - 1);
                const int idx = min(ic * (ic - 1) / 2 + r, last);
                const T cur = a.get(t);
                const T upd = wfms(cur, wld(lw + idx), x);
                a.set(t, (i > r && i < w) ? upd : cur);
            }
        }
    }
    if (valid) {
#pragma unroll
        for (int t = 0; t < kTrsmRPT; ++t) {
            const int i = sub * kTrsmRPT + t;
            if (i < w) wst(A + (size_t)(k0 + i) * lda + col, a.get(t));
        }
    }
}

// ---------------------------------------------------------------------------
// trailing update A22 -= L21 U12 on MFMA.
//
// The operands are first packed in the MFMA lane order (pack_l_kernel,
// pack_u_kernel): element (i, k) of L21 goes to chunk [k/4][i/16], lane
// (i%16) + 16(k%4); element (k, j) of U12 to chunk [k/4][j/16], lane
// (j%16) + 16(k%4). A chunk is one wavefront's operand of one 16x16x4 MFMA,
// so every operand load is a single contiguous 64-lane access (512 B for
// complex FP32). Rows and columns beyond the matrix are zero in the packed
// copies. Accumulator element v of lane l holds row 4v + l/16 for the FP64
// instruction and row 4(l/16) + v for the FP32 one, column l % 16 (checked
// on gfx950 by mfma_layout_probe_kernel).

template <class S> struct Mfma;
template <> struct Mfma<double2> {
    using acc = v4d;
    using op = double;
    static constexpr bool cplx = true;
    static constexpr int wave_col_blocks = 2;  // 64 x 32 wave tile
    __device__ static acc mma(op a, op b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static int row(int lk, int v) { return 4 * v + lk; }
    __device__ static op re(const double2 &v) { return v.x; }
    __device__ static op im(const double2 &v) { return v.y; }
};
template <> struct Mfma<float2> {
    using acc = v4f;
    using op = float;
    static constexpr bool cplx = true;
    static constexpr int wave_col_blocks = 4;  // 64 x 64 wave tile
    __device__ static acc mma(op a, op b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static int row(int lk, int v) { return 4 * lk + v; }
    __device__ static op re(const float2 &v) { return v.x; }
    __device__ static op im(const float2 &v) { return v.y; }
};
template <> struct Mfma<double> {
    using acc = v4d;
    using op = double;
    static constexpr bool cplx = false;
    static constexpr int wave_col_blocks = 4;
    __device__ static acc mma(op a, op b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static int row(int lk, int v) { return 4 * v + lk; }
    __device__ static op re(const double &v) { return v; }
    __device__ static op im(const double &) { return 0.0; }
};
constexpr int kWaveRowBlocks = 4;  // 64 rows per wave, 128 per workgroup (2 x 2 waves)

template <class S> __device__ inline S szero();
template <> __device__ inline double2 szero<double2>() { return make_double2(0.0, 0.0); }
template <> __device__ inline float2 szero<float2>() { return make_float2(0.f, 0.f); }
template <> __device__ inline double szero<double>() { return 0.0; }

// L21 rows [row0, row0 + m), columns [k0, k0 + kw) -> chunks [kw/4][rb_pad]
template <class S>
__global__ void pack_l_kernel(const S *A, size_t lda, int row0, int m, int k0, int kw, int rb_pad, S *out,
                              const int *status) {
    if (status[0] != 0) return;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = (size_t)((kw + 3) / 4) * rb_pad * 64;
    if (idx >= total) return;
    const int lane = (int)(idx & 63);
    const size_t chunk = idx >> 6;
    const int rb = (int)(chunk % rb_pad), ks = (int)(chunk / rb_pad);
    const int row = rb * 16 + (lane & 15), k = ks * 4 + (lane >> 4);
    out[idx] = (row < m && k < kw) ? A[(size_t)(row0 + row) * lda + k0 + k] : szero<S>();
}

// U12 rows [k0, k0 + kw), columns [col0, col0 + ncols) -> chunks
// [kw/4][cb_count] of a buffer whose k steps are cb_stride chunks apart
template <class S>
__global__ void pack_u_kernel(const S *A, size_t lda, int k0, int kw, int col0, int ncols, int cb_count,
                              int cb_stride, S *out, const int *status) {
    if (status[0] != 0) return;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = (size_t)((kw + 3) / 4) * cb_count * 64;
    if (idx >= total) return;
    const int lane = (int)(idx & 63);
    const size_t chunk = idx >> 6;
    const int cb = (int)(chunk % cb_count), ks = (int)(chunk / cb_count);
    const int col = cb * 16 + (lane & 15), k = ks * 4 + (lane >> 4);
    out[((size_t)ks * cb_stride + cb) * 64 + lane] =
        (col < ncols && k < kw) ? A[(size_t)(k0 + k) * lda + col0 + col] : szero<S>();
}

// f(std::integral_constant<int, 0>) ... f(std::integral_constant<int, N - 1>)
template <class F, int... I> __device__ inline void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F> __device__ inline void static_for(F &&f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Workgroup tile index -> (tile row, tile column). Workgroups are dispatched
// round-robin over the 8 XCDs; consecutive ids on one XCD are given a
// contiguous range of tiles, walked in groups of 8 tile rows (column-major
// inside a group), so the tiles one XCD runs at a time share L21 rows and
// U12 columns in that XCD's L2.
__device__ inline bool gemm_tile(int tiles_m, int tiles_n, int &tm, int &tn) {
    const int T = tiles_m * tiles_n, per = (T + 7) / 8;
    const int nid = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (nid >= T) return false;
    constexpr int kGroup = 8;
    const int in_group = kGroup * tiles_n;
    const int grp = nid / in_group, first = grp * kGroup;
    const int gsz = min(tiles_m - first, kGroup);
    const int loc = nid - grp * in_group;
    tm = first + loc % gsz;
    tn = loc / gsz;
    return true;
}

template <class S>
__global__ __launch_bounds__(256, 2) void gemm_packed_kernel(S *A, size_t lda, int row0, int col0, int m, int ncols,
                                                          int ks_count, const S *__restrict__ Lp,
                                                          const S *__restrict__ Up, int rb_pad, int cb_pad,
                                                          const int *status) {
    using M = Mfma<S>;
    using acc_t = typename M::acc;
    using op = typename M::op;
    constexpr int MB = kWaveRowBlocks, NB = M::wave_col_blocks;
    constexpr int BM = 32 * MB, BN = 32 * NB;
    if (status[0] != 0) return;
    int tm, tn;
    if (!gemm_tile((m + BM - 1) / BM, (ncols + BN - 1) / BN, tm, tn)) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int rb0 = tm * (2 * MB) + (wave >> 1) * MB;
    const int cb0 = tn * (2 * NB) + (wave & 1) * NB;
    acc_t cre[MB][NB], cim[MB][NB];
    for (int a = 0; a < MB; ++a)
        for (int b = 0; b < NB; ++b) {
            cre[a][b] = acc_t{0, 0, 0, 0};
            cim[a][b] = acc_t{0, 0, 0, 0};
        }
    const S *lp = Lp + (size_t)rb0 * 64 + lane;
    const S *up = Up + (size_t)cb0 * 64 + lane;
    const size_t lstride = (size_t)rb_pad * 64, ustride = (size_t)cb_pad * 64;
    // the loads of step ks + 1 are issued before the MFMAs of step ks (the
    // packed copies hold spare k steps, so the load past the last step stays
    // in bounds). Operands are kept as plain scalars (arrays of vector types
    // would go to scratch).
    op ar0[MB], ai0[MB], br0[NB], bi0[NB], ar1[MB], ai1[MB], br1[NB], bi1[NB];
#define PEEC_GEMM_LOAD(AR, AI, BR, BI, KS)                                      \
    {                                                                          \
        const S *l_ = lp + (size_t)(KS) * lstride, *u_ = up + (size_t)(KS) * ustride; \
        for (int i = 0; i < MB; ++i) {                                         \
            const S v_ = l_[i * 64];                                           \
            AR[i] = Mfma<S>::re(v_);                                           \
            AI[i] = Mfma<S>::im(v_);                                           \
        }                                                                      \
        for (int j = 0; j < NB; ++j) {                                         \
            const S v_ = u_[j * 64];                                           \
            BR[j] = Mfma<S>::re(v_);                                           \
            BI[j] = Mfma<S>::im(v_);                                           \
        }                                                                      \
    }
#define PEEC_GEMM_STEP(AR, AI, BR, BI)                                         \
    {                                                                          \
        if constexpr (M::cplx) {                                               \
            for (int i = 0; i < MB; ++i)                                       \
                for (int j = 0; j < NB; ++j) {                                 \
                    cre[i][j] = M::mma(AR[i], BR[j], cre[i][j]);               \
                    cim[i][j] = M::mma(AR[i], BI[j], cim[i][j]);               \
                }                                                              \
            for (int i = 0; i < MB; ++i) {                                     \
                const op nai_ = -AI[i];                                        \
                for (int j = 0; j < NB; ++j) {                                 \
                    cre[i][j] = M::mma(nai_, BI[j], cre[i][j]);                \
                    cim[i][j] = M::mma(AI[i], BR[j], cim[i][j]);               \
                }                                                              \
            }                                                                  \
        } else {                                                               \
            for (int i = 0; i < MB; ++i)                                       \
                for (int j = 0; j < NB; ++j) cre[i][j] = M::mma(AR[i], BR[j], cre[i][j]); \
        }                                                                      \
    }
    PEEC_GEMM_LOAD(ar0, ai0, br0, bi0, 0);
    for (int ks = 0; ks < ks_count; ++ks) {
        PEEC_GEMM_LOAD(ar1, ai1, br1, bi1, ks + 1);
        PEEC_GEMM_STEP(ar0, ai0, br0, bi0);
        for (int i = 0; i < MB; ++i) {
            ar0[i] = ar1[i];
            ai0[i] = ai1[i];
        }
        for (int j = 0; j < NB; ++j) {
            br0[j] = br1[j];
            bi0[j] = bi1[j];
        }
        // issue order for the scheduler: the next step's loads first, then
        // this step's MFMAs (otherwise the loads sink behind the MFMAs and
        // every step waits for a full memory latency)
        __builtin_amdgcn_sched_group_barrier(0x020, MB + NB, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, (M::cplx ? 4 : 1) * MB * NB, 0);
    }
#undef PEEC_GEMM_LOAD
#undef PEEC_GEMM_STEP
    const int li = lane & 15, lk = lane >> 4;
    for (int i = 0; i < MB; ++i)
        for (int j = 0; j < NB; ++j) {
            const int col = (cb0 + j) * 16 + li;
            if (col >= ncols) continue;
            for (int v = 0; v < 4; ++v) {
                const int row = (rb0 + i) * 16 + M::row(lk, v);
                if (row >= m) continue;
                S *c = A + (size_t)(row0 + row) * lda + col0 + col;
                if constexpr (M::cplx) {
                    S cv = *c;
                    cv.x -= cre[i][j][v];
                    cv.y -= cim[i][j][v];
                    *c = cv;
                } else {
                    *c -= cre[i][j][v];
                }
            }
        }
}

// Tile of virtual id vid under the mapping of gemm_tile (vid plays the role
// of blockIdx.x); false for the padding ids past the tile count.
__device__ inline bool gemm_tile_at(int vid, int tiles_m, int tiles_n, int &tm, int &tn) {
    const int T = tiles_m * tiles_n, per = (T + 7) / 8;
    const int nid = (vid & 7) * per + (vid >> 3);
    if (nid >= T) return false;
    constexpr int kGroup = 8;
    const int in_group = kGroup * tiles_n;
    const int grp = nid / in_group, first = grp * kGroup;
    const int gsz = min(tiles_m - first, kGroup);
    const int loc = nid - grp * in_group;
    tm = first + loc % gsz;
    tn = loc / gsz;
    return true;
}

// The same update for the block widths the factorizations use (KS k steps of
// 4, i.e. 64 or 128 columns of L21), persistent: one workgroup per CU walks
// the tiles vid = blockIdx.x, + gridDim.x, ... (gridDim.x a multiple of 8,
// so every tile of a workgroup stays on its XCD). With every tile taking the
// same number of steps, reading the C tile only after the last MFMA would
// make all CUs wait on HBM together; here the C tile is read into registers
// a few elements per k step, so its traffic runs under the MFMAs and the end
// of a tile only subtracts and stores. The k loop is unrolled; operands are
// loaded PD = 3 steps ahead into four register sets, and the last steps of a
// tile load the first steps of the workgroup's next tile, so the stores of
// one tile overlap the start of the next. Per step scheduling barriers keep
// the loads ahead of the step's MFMAs.
template <class S, int KS>
__global__ __launch_bounds__(256, 1) void gemm_stream_kernel(S *A, size_t lda, int row0, int col0, int m, int ncols,
                                                             const S *__restrict__ Lp, const S *__restrict__ Up,
                                                             int rb_pad, int cb_pad, const int *status) {
    using M = Mfma<S>;
    using acc_t = typename M::acc;
    using op = typename M::op;
    constexpr int MB = kWaveRowBlocks, NB = M::wave_col_blocks;
    constexpr int BM = 32 * MB, BN = 32 * NB;
    constexpr int CL = MB * NB * 4;  // C elements per lane
    constexpr int PD = 3, NBUF = 4;  // operand prefetch distance (k steps), register sets
    static_assert(KS % NBUF == 0, "a tile's steps must cycle the register sets exactly");
    if (status[0] != 0) return;
    const int tiles_m = (m + BM - 1) / BM, tiles_n = (ncols + BN - 1) / BN;
    const int V = 8 * ((tiles_m * tiles_n + 7) / 8);
    int vt = blockIdx.x, tm = 0, tn = 0;
    while (vt < V && !gemm_tile_at(vt, tiles_m, tiles_n, tm, tn)) vt += gridDim.x;
    if (vt >= V) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const size_t lstride = (size_t)rb_pad * 64, ustride = (size_t)cb_pad * 64;
    auto lp_of = [&](int tm_) { return Lp + (size_t)(tm_ * (2 * MB) + (wave >> 1) * MB) * 64 + lane; };
    auto up_of = [&](int tn_) { return Up + (size_t)(tn_ * (2 * NB) + (wave & 1) * NB) * 64 + lane; };
    op ar[NBUF][MB], ai[NBUF][MB], br[NBUF][NB], bi[NBUF][NB];
    auto load_ops = [&](const S *lp_, const S *up_, int buf, int ks) {
        const S *l_ = lp_ + (size_t)ks * lstride, *u_ = up_ + (size_t)ks * ustride;
#pragma unroll
        for (int i = 0; i < MB; ++i) {
            const S v_ = l_[i * 64];
            ar[buf][i] = M::re(v_);
            ai[buf][i] = M::im(v_);
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const S v_ = u_[j * 64];
            br[buf][j] = M::re(v_);
            bi[buf][j] = M::im(v_);
        }
    };
    const S *lp = lp_of(tm), *up = up_of(tn);
#pragma unroll
    for (int q = 0; q < PD && q < KS; ++q) load_ops(lp, up, q, q);
    for (;;) {
        // the workgroup's next tile (its first steps are loaded during this one)
        int vn = vt + gridDim.x, tmn = tm, tnn = tn;
        while (vn < V && !gemm_tile_at(vn, tiles_m, tiles_n, tmn, tnn)) vn += gridDim.x;
        const bool more = vn < V;
        const S *lpn = more ? lp_of(tmn) : lp, *upn = more ? up_of(tnn) : up;
        const int rb0 = tm * (2 * MB) + (wave >> 1) * MB;
        const int cb0 = tn * (2 * NB) + (wave & 1) * NB;
        acc_t cre[MB][NB], cim[MB][NB];
#pragma unroll
        for (int a = 0; a < MB; ++a)
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                cre[a][b] = acc_t{0, 0, 0, 0};
                cim[a][b] = acc_t{0, 0, 0, 0};
            }
        // C element e = (i * NB + j) * 4 + v of this lane: row (rb0 + i) * 16 +
        // row(lk, v), column (cb0 + j) * 16 + li; rows and columns past the
        // matrix read a clamped in-bounds address and are not stored
        S cv[CL];
        auto c_ptr = [&](int e) -> S * {
            const int i = e / (NB * 4), j = (e / 4) % NB, v = e % 4;
            const int row = min((rb0 + i) * 16 + M::row(lk, v), m - 1);
            const int col = min((cb0 + j) * 16 + li, ncols - 1);
            return A + (size_t)(row0 + row) * lda + col0 + col;
        };
        static_for<KS>([&](auto step) {
            constexpr int s = decltype(step)::value;
            constexpr int e0 = s * CL / KS, e1 = (s + 1) * CL / KS;
            if constexpr (s + PD < KS) load_ops(lp, up, (s + PD) % NBUF, s + PD);
            else load_ops(lpn, upn, (s + PD) % NBUF, s + PD - KS);
#pragma unroll
            for (int e = e0; e < e1; ++e) cv[e] = *c_ptr(e);
            __builtin_amdgcn_sched_barrier(0);
            constexpr int b = s % NBUF;
            if constexpr (M::cplx) {
#pragma unroll
                for (int i = 0; i < MB; ++i)
#pragma unroll
                    for (int j = 0; j < NB; ++j) {
                        cre[i][j] = M::mma(ar[b][i], br[b][j], cre[i][j]);
                        cim[i][j] = M::mma(ar[b][i], bi[b][j], cim[i][j]);
                    }
#pragma unroll
                for (int i = 0; i < MB; ++i) {
                    const op nai = -ai[b][i];
#pragma unroll
                    for (int j = 0; j < NB; ++j) {
                        cre[i][j] = M::mma(nai, bi[b][j], cre[i][j]);
                        cim[i][j] = M::mma(ai[b][i], br[b][j], cim[i][j]);
                    }
                }
            } else {
#pragma unroll
                for (int i = 0; i < MB; ++i)
#pragma unroll
                    for (int j = 0; j < NB; ++j) cre[i][j] = M::mma(ar[b][i], br[b][j], cre[i][j]);
            }
            __builtin_amdgcn_sched_barrier(0);
        });
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const int col = (cb0 + j) * 16 + li;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = (rb0 + i) * 16 + M::row(lk, v);
                    const int e = (i * NB + j) * 4 + v;
                    S r = cv[e];
                    if constexpr (M::cplx) {
                        r.x -= cre[i][j][v];
                        r.y -= cim[i][j][v];
                    } else {
                        r -= cre[i][j][v];
                    }
                    if (row < m && col < ncols) A[(size_t)(row0 + row) * lda + col0 + col] = r;
                }
            }
        if (!more) break;
        vt = vn;
        tm = tmn;
        tn = tnn;
        lp = lpn;
        up = upn;
    }
}

// Narrow update on the look-ahead path: 64 x 64 tiles, the k range split
// over the four waves of a workgroup and the partial sums added through
// LDS, so one tile takes a quarter of the sequential MFMA chain.
template <class S>
__global__ __launch_bounds__(256) void gemm_splitk_kernel(S *A, size_t lda, int row0, int col0, int m, int ncols,
                                                          int ks_count, const S *__restrict__ Lp,
                                                          const S *__restrict__ Up, int rb_pad, int cb_pad,
                                                          const int *status) {
    using M = Mfma<S>;
    using acc_t = typename M::acc;
    using op = typename M::op;
    static_assert(M::cplx, "complex storage only");
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    op *part = reinterpret_cast<op *>(smem_raw);  // [wave][re/im][16 blocks][4][64 lanes]
    if (status[0] != 0) return;
    const int tiles_n = (ncols + 63) / 64;
    const int tm = blockIdx.x / tiles_n, tn = blockIdx.x % tiles_n;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int rb0 = tm * 4, cb0 = tn * 4;
    const int per = (ks_count + 3) / 4, ks0 = wave * per, ks1 = min(ks_count, ks0 + per);
    acc_t cre[4][4], cim[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            cre[i][j] = acc_t{0, 0, 0, 0};
            cim[i][j] = acc_t{0, 0, 0, 0};
        }
    if (ks0 < ks1) {
        const S *lp = Lp + ((size_t)ks0 * rb_pad + rb0) * 64 + lane;
        const S *up = Up + ((size_t)ks0 * cb_pad + cb0) * 64 + lane;
        const size_t lstride = (size_t)rb_pad * 64, ustride = (size_t)cb_pad * 64;
        S a[4], b[4];
        for (int i = 0; i < 4; ++i) a[i] = lp[i * 64];
        for (int j = 0; j < 4; ++j) b[j] = up[j * 64];
        for (int ks = ks0; ks < ks1; ++ks) {
            lp += lstride;
            up += ustride;
            S an[4], bn[4];
            for (int i = 0; i < 4; ++i) an[i] = lp[i * 64];
            for (int j = 0; j < 4; ++j) bn[j] = up[j * 64];
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) {
                    cre[i][j] = M::mma((op)a[i].x, (op)b[j].x, cre[i][j]);
                    cim[i][j] = M::mma((op)a[i].x, (op)b[j].y, cim[i][j]);
                }
            for (int i = 0; i < 4; ++i) {
                const op nai = -(op)a[i].y;
                for (int j = 0; j < 4; ++j) {
                    cre[i][j] = M::mma(nai, (op)b[j].y, cre[i][j]);
                    cim[i][j] = M::mma((op)a[i].y, (op)b[j].x, cim[i][j]);
                }
            }
            for (int i = 0; i < 4; ++i) a[i] = an[i];
            for (int j = 0; j < 4; ++j) b[j] = bn[j];
        }
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            for (int v = 0; v < 4; ++v) {
                const int e = ((i * 4 + j) * 4 + v) * 64 + lane;
                part[(size_t)(wave * 2 + 0) * 4096 + e] = cre[i][j][v];
                part[(size_t)(wave * 2 + 1) * 4096 + e] = cim[i][j][v];
            }
    __syncthreads();
    for (int e = threadIdx.x; e < 4096; e += 256) {
        op re = 0, im = 0;
        for (int q = 0; q < 4; ++q) {
            re += part[(size_t)(q * 2 + 0) * 4096 + e];
            im += part[(size_t)(q * 2 + 1) * 4096 + e];
        }
        const int blk = e >> 8, v = (e >> 6) & 3, ln = e & 63;
        const int row = (rb0 + (blk >> 2)) * 16 + M::row(ln >> 4, v);
        const int col = (cb0 + (blk & 3)) * 16 + (ln & 15);
        if (row < m && col < ncols) {
            S *c = A + (size_t)(row0 + row) * lda + col0 + col;
            S cv = *c;
            cv.x -= re;
            cv.y -= im;
            *c = cv;
        }
    }
}

// Packed-operand buffer sizes (elements) for a trailing update of m rows
// and ncols columns with k width kw (two spare k steps for the prefetch).
template <class S>
void gemm_pads(int m, int ncols, int &rb_pad, int &cb_pad) {
    constexpr int BM = 32 * kWaveRowBlocks;
    int nbk = 4;
    if constexpr (std::is_same<S, double2>::value) nbk = 2;
    const int BN = 32 * nbk;
    rb_pad = (m + BM - 1) / BM * (BM / 16);
    cb_pad = (ncols + BN - 1) / BN * (BN / 16);
}

// Pack the operands of A[row0.., col0..] -= L21 U12 (m rows, ncols columns,
// k width kw) into the context's buffers; returns the chunk counts.
template <class S>
void pack_operands(S *A, size_t lda, int row0, int col0, int m, int ncols, int k0, int kw, const int *status,
                   LUContext &ctx, hipStream_t s, int &rb_pad, int &cb_pad) {
    gemm_pads<S>(m, ncols, rb_pad, cb_pad);
    const int ks_count = (kw + 3) / 4;
    const size_t nl = (size_t)(ks_count + 2) * rb_pad * 64, nu = (size_t)(ks_count + 2) * cb_pad * 64;
    ctx.pack_l.reserve(nl * sizeof(S));
    ctx.pack_u.reserve(nu * sizeof(S));
    S *Lp = reinterpret_cast<S *>(ctx.pack_l.ptr);
    S *Up = reinterpret_cast<S *>(ctx.pack_u.ptr);
    const size_t tl = (size_t)ks_count * rb_pad * 64, tu = (size_t)ks_count * cb_pad * 64;
    hipLaunchKernelGGL(pack_l_kernel<S>, dim3((unsigned)((tl + 255) / 256)), dim3(256), 0, s, A, lda, row0, m, k0,
                       kw, rb_pad, Lp, status);
    hipLaunchKernelGGL(pack_u_kernel<S>, dim3((unsigned)((tu + 255) / 256)), dim3(256), 0, s, A, lda, k0, kw, col0,
                       ncols, cb_pad, cb_pad, Up, status);
}

// C = A[row0.., col0 + c_off ..] (m x nc) -= L21 U12[:, c_off .. c_off + nc)
// from the packed operands; c_off is a multiple of the tile width.
template <class S>
void gemm_packed(S *A, size_t lda, int row0, int col0, int m, int c_off, int nc, int kw, int rb_pad, int cb_pad,
                 const int *status, LUContext &ctx, hipStream_t s, bool leave_cus = false) {
    if (m <= 0 || nc <= 0) return;
    const int ks_count = (kw + 3) / 4;
    const S *Lp = reinterpret_cast<const S *>(ctx.pack_l.ptr);
    const S *Up = reinterpret_cast<const S *>(ctx.pack_u.ptr) + (size_t)(c_off / 16) * 64;
    int rb_t, cb_t;
    gemm_pads<S>(m, nc, rb_t, cb_t);
    const int tiles = (rb_t / (2 * kWaveRowBlocks)) * (cb_t / (std::is_same<S, double2>::value ? 4 : 8));
    const int grid = (tiles + 7) / 8 * 8;
    const int tg = ctx.timer ? ctx.timer->start(s) : -1;
    // persistent streamed kernel: one workgroup per CU (with leave_cus, one
    // CU per XCD stays free for the look-ahead's diagonal factor), a
    // multiple of 8 so that each workgroup keeps its XCD
    const int pgrid = std::min(grid, std::max(8, (ctx.cus - (leave_cus ? 8 : 0)) / 8 * 8));
    if (ks_count == 32)
        hipLaunchKernelGGL((gemm_stream_kernel<S, 32>), dim3(pgrid), dim3(256), 0, s, A, lda, row0, col0 + c_off, m,
                           nc, Lp, Up, rb_pad, cb_pad, status);
    else if (ks_count == 16)
        hipLaunchKernelGGL((gemm_stream_kernel<S, 16>), dim3(pgrid), dim3(256), 0, s, A, lda, row0, col0 + c_off, m,
                           nc, Lp, Up, rb_pad, cb_pad, status);
    else
        hipLaunchKernelGGL(gemm_packed_kernel<S>, dim3(grid), dim3(256), 0, s, A, lda, row0, col0 + c_off, m, nc,
                           ks_count, Lp, Up, rb_pad, cb_pad, status);
    if (ctx.timer) {
        const double mnk = (double)m * nc * kw;
        ctx.timer->stop(tg, s, kKernelLuUpdate, (Elem<S>::cplx ? 8.0 : 2.0) * mnk,
                        (2.0 * m * nc + (double)(m + nc) * kw) * sizeof(S));
    }
}

template <class S>
void launch_gemm(S *A, size_t lda, int row0, int col0, int m, int ncols, int k0, int kw, const int *status,
                 LUContext &ctx, hipStream_t s) {
    int rb_pad, cb_pad;
    pack_operands<S>(A, lda, row0, col0, m, ncols, k0, kw, status, ctx, s, rb_pad, cb_pad);
    gemm_packed<S>(A, lda, row0, col0, m, 0, ncols, kw, rb_pad, cb_pad, status, ctx, s);
}

// Self-test of the 16x16x4 MFMA operand layouts used by the GEMM kernels:
// one wavefront, A[i][k] = i + 16k + 1 in lane i + 16k, B[k][j] = (j == k)
// and B2[k][j] = (k == 0) in lane j + 16k; D = A padded and D2[i][j] =
// A[i][0]. Writes every lane's 4 accumulators for both products, first for
// the FP64 instruction, then for the FP32 one (4 x 256 doubles).
__global__ void mfma_layout_probe_kernel(double *out) {
    const int lane = threadIdx.x;
    const int li = lane & 15, lk = lane >> 4;
    const double a = (double)(li + 16 * lk + 1);
    const double b = (li == lk) ? 1.0 : 0.0;
    const double b2 = (lk == 0) ? 1.0 : 0.0;
    v4d d = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, v4d{0, 0, 0, 0}, 0, 0, 0);
    v4d d2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b2, v4d{0, 0, 0, 0}, 0, 0, 0);
    v4f f = __builtin_amdgcn_mfma_f32_16x16x4f32((float)a, (float)b, v4f{0, 0, 0, 0}, 0, 0, 0);
    v4f f2 = __builtin_amdgcn_mfma_f32_16x16x4f32((float)a, (float)b2, v4f{0, 0, 0, 0}, 0, 0, 0);
    for (int v = 0; v < 4; ++v) {
        out[lane * 4 + v] = d[v];
        out[256 + lane * 4 + v] = d2[v];
        out[512 + lane * 4 + v] = f[v];
        out[768 + lane * 4 + v] = f2[v];
    }
}

// ---------------------------------------------------------------------------
// host driver

// The trailing update of a partial-pivoting FP32 factor on split-FP16 MFMA
// (defined with the split-FP16 kernels below).
static void pivoted_update_h3(float2 *A, size_t lda, int k0, int w, int n, LUContext &ctx, hipStream_t s);

template <class S>
LUStatus lu_factor(S *dA, int n, size_t lda, int *d_ipiv, double pivot_limit, LUContext &ctx, int block_size) {
    using T = typename Elem<S>::A;
    hipStream_t s = ctx.stream;
    LUStatus out;
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, s, ctx.status.ptr, 2, 0);
    hipLaunchKernelGGL(fill_uint_kernel, dim3((unsigned)((kBarWords + 255) / 256)), dim3(256), 0, s, ctx.barrier.ptr,
                       (int)kBarWords, 0u);
    HIP_CHECK(hipGetLastError());
    // dynamic LDS of the panel kernel: the device's per-workgroup limit (or
    // the context's cap) less the kernel's static arrays
    size_t lds_cap = (size_t)ctx.lds_per_block;
    if (ctx.panel_lds_limit > 0 && ctx.panel_lds_limit < lds_cap) lds_cap = ctx.panel_lds_limit;
    const size_t lds_budget = lds_cap > 16 * 1024 ? lds_cap - 8 * 1024 : lds_cap / 2;
    // the cooperative panel over rows j0.. and columns j0 .. j0 + wj - 1
    // (its slab in the LDS of one workgroup per CU), then its row swaps on
    // every other column
    auto panel = [&](int j0, int wj) -> bool {
        const int mj = n - j0;
        const int G = std::max(1, std::min(ctx.cus, (mj + 7) / 8));
        const int rpw = (mj + G - 1) / G;
        double *slots = ctx.slots.ptr;
        double *mags = ctx.cand.ptr;
        unsigned long long *tags = ctx.tags.ptr;
        const size_t shm = (size_t)rpw * wj * sizeof(T);
        if (shm > lds_budget) {
            out.code = -1;
            PEEC_ERR("LU panel of %d rows exceeds the LDS budget.", mj);
            return false;
        }
        int nn = n, kk0 = j0, ww = wj, rr = rpw;
        size_t ld = lda;
        unsigned round0 = ctx.round;
        ctx.round += (unsigned)wj + 1u;
        double lim = pivot_limit;
        void *args[] = {&dA, &ld, &nn, &kk0, &ww, &rr, &slots, &mags, &tags, &d_ipiv, &round0, &lim, &ctx.status.ptr};
        if (shm > 64 * 1024)
            HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(panel_kernel<S>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        const int tp = ctx.timer ? ctx.timer->start(s) : -1;
        HIP_CHECK(hipLaunchCooperativeKernel(reinterpret_cast<void *>(panel_kernel<S>), dim3(G), dim3(256),
                                             args, (unsigned)shm, s));
        if (ctx.timer) ctx.timer->stop(tp, s, kKernelPanel, 0.0, 2.0 * mj * wj * sizeof(S));
        if (n - wj > 0) {
            ctx.swap_stage.reserve((size_t)2 * wj * n * sizeof(S));
            S *stage = reinterpret_cast<S *>(ctx.swap_stage.ptr);
            hipLaunchKernelGGL(swap_plan_kernel, dim3(1), dim3(256), 0, s, d_ipiv, j0, wj, ctx.swap_plan.ptr,
                               ctx.status.ptr);
            const dim3 sg((unsigned)((n - wj + 255) / 256), (unsigned)(2 * wj));
            hipLaunchKernelGGL(swap_gather_kernel<S>, sg, dim3(256), 0, s, dA, lda, n, j0, wj, ctx.swap_plan.ptr,
                               stage, ctx.status.ptr);
            hipLaunchKernelGGL(swap_scatter_kernel<S>, sg, dim3(256), 0, s, dA, lda, n, j0, wj, ctx.swap_plan.ptr,
                               stage, ctx.status.ptr);
        }
        return true;
    };
    // U12 = L11^-1 A12 of the w columns from k0, over columns c0 .. c1 - 1
    auto trsm = [&](int k0, int w, int c0, int c1) -> bool {
        const size_t tshm = trsm_lds_elems(w) * sizeof(S);
        if (tshm > (size_t)ctx.lds_per_block) {
            out.code = -1;
            PEEC_ERR("LU block of %d columns needs %zu bytes of LDS per workgroup.", w, tshm);
            return false;
        }
        if (tshm > 64 * 1024)
            HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(trsm_kernel<S>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)tshm));
        const int cols = c1 - c0;
        hipLaunchKernelGGL(trsm_kernel<S>, dim3((cols + 256 / kTrsmTPC - 1) / (256 / kTrsmTPC)), dim3(256), tshm, s,
                           dA, lda, c0, c1, k0, w, ctx.status.ptr);
        return true;
    };
    int k0 = 0;
    while (k0 < n) {
        const int m = n - k0;
        // workgroups: one per CU at most, at least 8 rows each
        const int G = std::max(1, std::min(ctx.cus, (m + 7) / 8));
        int nb = block_size > 0 ? block_size : 128;
        nb = std::min(nb, kMaxNB);
        const int w = std::min(nb, m);
        // Block columns of w = 128: when the slab of a w-wide panel does not
        // fit the LDS of the CUs (tall matrices), the block column is
        // factored as narrower sub-panels that fit, each followed by its
        // U12 and rank-ws update within the block column only; the trailing
        // matrix then still takes one rank-w update per block column.
        const int rpw = (m + G - 1) / G;
        int ws = w;
        while (ws > 16 && (size_t)rpw * ws * sizeof(T) > lds_budget) ws /= 2;
        for (int j0 = k0; j0 < k0 + w; j0 += ws) {
            const int wj = std::min(ws, k0 + w - j0);
            if (!panel(j0, wj)) return out;
            const int right = k0 + w - (j0 + wj);
            if (right > 0) {
                if (!trsm(j0, wj, j0 + wj, k0 + w)) return out;
                launch_gemm<S>(dA, lda, j0 + wj, j0 + wj, n - j0 - wj, right, j0, wj, ctx.status.ptr, ctx, s);
            }
        }
        const int rest = n - k0 - w;
        if (rest > 0) {
            if (!trsm(k0, w, k0 + w, n)) return out;
            bool split = false;
            if constexpr (std::is_same<S, float2>::value) split = ctx.pivoted_split16 && (w == 128 || w == 64);
            if (split)
                pivoted_update_h3(reinterpret_cast<float2 *>(dA), lda, k0, w, n, ctx, s);
            else
                launch_gemm<S>(dA, lda, k0 + w, k0 + w, rest, rest, k0, w, ctx.status.ptr, ctx, s);
        }
        HIP_CHECK(hipGetLastError());
        k0 += w;
    }
    int st[2] = {0, 0};
    ctx.status.download(st, 2, s);
    HIP_CHECK(hipStreamSynchronize(s));
    out.code = st[0];
    out.column = st[1];
    return out;
}

template LUStatus lu_factor<double2>(double2 *, int, size_t, int *, double, LUContext &, int);
template LUStatus lu_factor<float2>(float2 *, int, size_t, int *, double, LUContext &, int);
template LUStatus lu_factor<double>(double *, int, size_t, int *, double, LUContext &, int);

// ---------------------------------------------------------------------------
// split-FP16 trailing update of the ordered factor
//
// C -= L21 U12 in complex FP32 storage on v_mfma_f32_16x16x32_f16. Each
// FP32 operand x is scaled by a power of two 2^e (per row of L21, per column
// of U12, so that the row's / column's largest part lies in [2^13, 2^14)) and
// split as x 2^e = hi + lo with hi = fp16(x 2^e), lo = fp16(x 2^e - hi);
// hi hi' + hi lo' + lo hi' carries the product to ~2^-22 relative (the lo lo'
// term is below that), accumulated in FP32 like the FP32 MFMA update, and
// the scales come off exactly at the store. The complex product is two real
// dot products over interleaved parts k^ = 2k + p:
//   Re C = sum (Ar, -Ai) . (Br, Bi),   Im C = sum (Ai, Ar) . (Br, Bi),
// so L21 is packed once as the pairs (Ar, -Ai) and the pairs (Ai, Ar) are
// formed in registers (swap the halves, flip one sign).
//
// Packed layout, in 16-byte lane vectors (8 halves): a chunk is 16 complex k
// (32 k^); for a 16-row (16-column) block b and chunk c the hi vectors of
// its 64 lanes, then the lo vectors, at ((c * blocks + b) * 2 + h) * 64 +
// lane. Lane l holds row (column) l & 15 and complex k = 16 c + 4 (l >> 4)
// .. + 3. The MFMA pairs A and B lane slots by the same k^ order, so only the
// row / column of a lane matters.

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// half pair (x, y) -> (-y, x)
__device__ inline unsigned turn_pair(unsigned d) { return ((d >> 16) ^ 0x8000u) | (d << 16); }
typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));
__device__ inline u32x4_t turn_pairs(u32x4_t v) {
    return u32x4_t{turn_pair(v.x), turn_pair(v.y), turn_pair(v.z), turn_pair(v.w)};
}
__device__ inline unsigned pack_half2(_Float16 a, _Float16 b) {
    return (unsigned)__builtin_bit_cast(unsigned short, a) | ((unsigned)__builtin_bit_cast(unsigned short, b) << 16);
}
// 2^(13 - e) for the largest magnitude mx (0 or non-finite: 1)
__device__ inline float split_scale(float mx) {
    if (!(mx > 0.f) || !isfinite(mx)) return 1.f;
    return ldexpf(1.f, 13 - ilogbf(mx));
}
// x -> hi, lo halves
__device__ inline void split_half(float x, _Float16 &hi, _Float16 &lo) {
    hi = (_Float16)x;
    lo = (_Float16)(x - (float)hi);
}

// Split-FP16 operand packs (written by the blocked triangular solves,
// pack_h3_regs): per chunk of 16 complex k and per 16-line block (rows of
// L21, columns of U12), 64 lanes of uint4 hi halves then 64 of lo halves;
// lane l holds line l & 15 and k = 16 c + 4 (l >> 4) .. + 3 as (re, im)
// half pairs, L21 as (re, -im). Each line is scaled by a power of two
// (split_scale of its largest magnitude over all k of the update), whose
// inverse is stored beside the pack.

// The operands of one segment of an update: split-FP16 packs of L21 (from
// the update's first row) and U12 (from its first column), their per-row /
// per-column inverse scales, and the row / column blocks per chunk.
struct H3Seg {
    const uint4 *L = nullptr, *U = nullptr;
    const float *linv = nullptr, *uinv = nullptr;
    int rb_pad = 0, cb_count = 0;
};

// C -= L21 U12 on the split-FP16 operands, one tile of (32 MB) x (32 NB),
// NB = 8 / MB, per workgroup of four waves in a 2 x 2 grid (each wave MB x NB
// blocks of 16 x 16), two workgroups per CU so that one's waits fall under
// the other's MFMAs. No persistent loop: no loads stay in flight across
// tiles. A chunk of a tile's operands is 4 MB + 4 NB units of 1 KB (the hi
// and lo vectors of its 2 MB row blocks and 2 NB column blocks); each wave
// loads MB L21 and NB U12 units into one of four register sets four chunks
// ahead and stores them to one of two LDS slots one chunk ahead, so every
// operand is fetched once per workgroup. C is read after the last chunk.
// Per chunk and 16 x 16 block: three MFMAs for the real and three for the
// imaginary part. The update uses MB = 2 (64 x 128 tiles): rows of C are
// then read and written in 1 KB runs.
// With SEG = 2 the tile takes two updates of KC chunks each (two panel
// pairs, each packed with its own scales) in one pass over C: after the
// first segment's chunks the sums are brought to the second segment's
// scales, (linv_a / linv_b)(uinv_a / uinv_b) per entry, a power of two and
// so exact, and the second segment's products accumulate on top.
// MB x NB blocks per wave (2 x 2 waves: tiles of 32 MB x 32 NB); ND staging
// register sets (chunks loaded ahead).
template <int KC, int MB, int NB, int ND, int SEG>
__device__ __forceinline__ void gemm_tile_h3_body(float2 *A, size_t lda, int row0, int col0, int m, int ncols,
                                                  const H3Seg &sa, const H3Seg &sb, const int *status) {
    constexpr int BM = 32 * MB, BN = 32 * NB;
    constexpr int UA = 4 * MB, UNITS = UA + 4 * NB;  // L21 units, all units per chunk
    constexpr int KT = SEG * KC;                     // chunks of the tile
    // C is read and written EG block rows of the wave at a time (at most 32
    // values in registers beside the accumulators)
    constexpr int EG = MB * NB <= 8 ? MB : 8 / NB, CL = EG * NB * 4;
    static_assert(KC % 2 == 0 && KC >= 2 && MB % EG == 0, "chunks alternate between two LDS slots");
    static_assert(SEG == 1 || SEG == 2, "one or two segments");
    __shared__ u32x4_t slots_lds[2 * UNITS * 64];
    if (status[0] != 0) return;
    int tm, tn;
    if (!gemm_tile((m + BM - 1) / BM, (ncols + BN - 1) / BN, tm, tn)) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    // wave w loads L21 units MB w .. MB w + MB - 1 and U12 units NB w .. of
    // each segment
    const u32x4_t *as[SEG], *bs[SEG];
    size_t astride[SEG], bstride[SEG];
    {
        const H3Seg *sg[2] = {&sa, &sb};
#pragma unroll
        for (int g = 0; g < SEG; ++g) {
            as[g] = reinterpret_cast<const u32x4_t *>(sg[g]->L) + (size_t)(tm * (BM / 8) + wave * MB) * 64 + lane;
            bs[g] = reinterpret_cast<const u32x4_t *>(sg[g]->U) + (size_t)(tn * (BN / 8) + wave * NB) * 64 + lane;
            astride[g] = (size_t)sg[g]->rb_pad * 128;
            bstride[g] = (size_t)sg[g]->cb_count * 128;
        }
    }
    static_assert(KT % ND == 0, "a tile's chunks must cycle the staging sets exactly");
    u32x4_t ga[ND][MB], gb[ND][NB];
    auto load_g = [&](auto set, auto chunk) {
        constexpr int t = decltype(set)::value, c = decltype(chunk)::value;
        constexpr int g = c / KC, cc = c % KC;
#pragma unroll
        for (int q = 0; q < MB; ++q) ga[t][q] = as[g][(size_t)cc * astride[g] + q * 64];
#pragma unroll
        for (int q = 0; q < NB; ++q) gb[t][q] = bs[g][(size_t)cc * bstride[g] + q * 64];
    };
    auto store_g = [&](auto set, int slot) {
        constexpr int t = decltype(set)::value;
        u32x4_t *base = slots_lds + (size_t)slot * UNITS * 64 + lane;
#pragma unroll
        for (int q = 0; q < MB; ++q) base[(wave * MB + q) * 64] = ga[t][q];
#pragma unroll
        for (int q = 0; q < NB; ++q) base[(UA + wave * NB + q) * 64] = gb[t][q];
    };
    static_for<ND>([&](auto q) {
        load_g(std::integral_constant<int, decltype(q)::value>{}, std::integral_constant<int, decltype(q)::value>{});
    });
    store_g(std::integral_constant<int, 0>{}, 0);
    __syncthreads();
    const int ra = (wave >> 1) * MB, cbl = (wave & 1) * NB;
    const int rb0 = tm * (2 * MB) + ra, cb0 = tn * (2 * NB) + cbl;
    float4_t cre[MB][NB], cim[MB][NB];
#pragma unroll
    for (int a = 0; a < MB; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            cre[a][b] = float4_t{0.f, 0.f, 0.f, 0.f};
            cim[a][b] = float4_t{0.f, 0.f, 0.f, 0.f};
        }
    float2 cv[CL];
    auto c_ptr = [&](int i, int j, int v) -> float2 * {
        const int row = min((rb0 + i) * 16 + 4 * lk + v, m - 1);
        const int col = min((cb0 + j) * 16 + li, ncols - 1);
        return A + (size_t)(row0 + row) * lda + col0 + col;
    };
    static_for<KT>([&](auto step) {
        constexpr int s = decltype(step)::value;
        using cur = std::integral_constant<int, s % ND>;
        using nxt = std::integral_constant<int, (s + 1) % ND>;
        if constexpr (s + ND < KT) load_g(cur{}, std::integral_constant<int, s + ND>{});
        if constexpr (SEG == 2 && s == KC) {
            // the first segment's sums in the second segment's scales
            float fc[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const int col = (cb0 + j) * 16 + li;
                fc[j] = sa.uinv[col] / sb.uinv[col];
            }
#pragma unroll
            for (int i = 0; i < MB; ++i)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = (rb0 + i) * 16 + 4 * lk + v;
                    const float fr = sa.linv[row] / sb.linv[row];
#pragma unroll
                    for (int j = 0; j < NB; ++j) {
                        cre[i][j][v] *= fr * fc[j];
                        cim[i][j][v] *= fr * fc[j];
                    }
                }
        }
        const u32x4_t *sl = slots_lds + (size_t)(s % 2) * UNITS * 64 + lane;
        half8_t uh[NB], ul[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            uh[j] = __builtin_bit_cast(half8_t, sl[(UA + 2 * (cbl + j)) * 64]);
            ul[j] = __builtin_bit_cast(half8_t, sl[(UA + 2 * (cbl + j) + 1) * 64]);
        }
#pragma unroll
        for (int i = 0; i < MB; ++i) {
            const u32x4_t h = sl[(2 * (ra + i)) * 64], l = sl[(2 * (ra + i) + 1) * 64];
            const half8_t rh = __builtin_bit_cast(half8_t, h), rl = __builtin_bit_cast(half8_t, l);
            const half8_t ih = __builtin_bit_cast(half8_t, turn_pairs(h)), il = __builtin_bit_cast(half8_t, turn_pairs(l));
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                cre[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(rh, uh[j], cre[i][j], 0, 0, 0);
                cim[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ih, uh[j], cim[i][j], 0, 0, 0);
            }
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                cre[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(rh, ul[j], cre[i][j], 0, 0, 0);
                cim[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ih, ul[j], cim[i][j], 0, 0, 0);
            }
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                cre[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(rl, uh[j], cre[i][j], 0, 0, 0);
                cim[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(il, uh[j], cim[i][j], 0, 0, 0);
            }
        }
        if constexpr (s + 1 < KT) {
            store_g(nxt{}, (s + 1) % 2);
            __syncthreads();
        }
    });
    const H3Seg &last = SEG == 2 ? sb : sa;
    float fc[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) fc[j] = last.uinv[(cb0 + j) * 16 + li];
#pragma unroll
    for (int g = 0; g < MB; g += EG) {
#pragma unroll
        for (int i = 0; i < EG; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) cv[(i * NB + j) * 4 + v] = *c_ptr(g + i, j, v);
#pragma unroll
        for (int i = 0; i < EG; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = (rb0 + g + i) * 16 + 4 * lk + v;
                const float fr = last.linv[row];
#pragma unroll
                for (int j = 0; j < NB; ++j) {
                    const int col = (cb0 + j) * 16 + li;
                    const float f = fr * fc[j];
                    float2 r = cv[(i * NB + j) * 4 + v];
                    r.x -= cre[g + i][j][v] * f;
                    r.y -= cim[g + i][j][v] * f;
                    if (row < m && col < ncols) A[(size_t)(row0 + row) * lda + col0 + col] = r;
                }
            }
    }
}

// the segments' pointers moved to batch member blockIdx.z
__device__ inline void h3_seg_member(H3Seg &g, const BatchStride &bs) {
    bs.apply_pack(g.L, g.linv, bs.pl, bs.sl);
    bs.apply_pack(g.U, g.uinv, bs.pu, bs.su);
}

// The trailing update, and the same code under its own name for the other
// updates (a separate entry in kernel traces).
template <int KC, int MB, int NB, int ND, int SEG>
__global__ __launch_bounds__(256, 2) void gemm_tile_h3_kernel(float2 *A, size_t lda, int row0, int col0, int m,
                                                              int ncols, H3Seg sa, H3Seg sb, const int *status,
                                                              BatchStride bs) {
    bs.apply(A, status);
    h3_seg_member(sa, bs);
    if constexpr (SEG == 2) h3_seg_member(sb, bs);
    gemm_tile_h3_body<KC, MB, NB, ND, SEG>(A, lda, row0, col0, m, ncols, sa, sb, status);
}
template <int KC, int MB, int NB, int ND>
__global__ __launch_bounds__(256, 2) void gemm_next_h3_kernel(float2 *A, size_t lda, int row0, int col0, int m,
                                                              int ncols, H3Seg sa, H3Seg sb, const int *status,
                                                              BatchStride bs) {
    bs.apply(A, status);
    h3_seg_member(sa, bs);
    gemm_tile_h3_body<KC, MB, NB, ND, 1>(A, lda, row0, col0, m, ncols, sa, sb, status);
}

// A[row0.., col0..] (m x ncols) -= L21 U12 from the split-FP16 packs of
// segment a (kw = 64, 128 or 256 columns of L21), then, with b, those of a
// second segment of the same k width in the same pass; kinds other than the
// trailing update run under the next-column kernel's name (one segment).
static void gemm_h3_segs(float2 *A, size_t lda, int row0, int col0, int m, int ncols, int kw, const H3Seg &a,
                         const H3Seg *b, KernelKind kind, const int *status, LUContext &ctx, hipStream_t s,
                         const BatchStride &bs, int batch) {
    if (m <= 0 || ncols <= 0) return;
    const int tg = ctx.timer ? ctx.timer->start(s) : -1;
    const bool next = kind != kKernelLuUpdateSplit16 && kind != kKernelLuUpdateMerged;
    if (b && next) throw HipError{"split-FP16 update: two segments only in the trailing update"};
    const H3Seg sb = b ? *b : H3Seg{};
    // 64 x 128 tiles by default; 32 x 128 tiles (twice the workgroups for a
    // small update) when the context's tuner picks them; trailing updates
    // large enough for two workgroups per CU take 128 x 128 tiles (a third
    // less operand traffic and LDS reads per output; two staging register
    // sets): measured 5 % faster over the C2 factor's trailing updates
    const bool short_tiles = ctx.h3_tile_rows == 32;
    const bool tall_tiles = !short_tiles && !next &&
                            (long long)((m + 127) / 128) * ((ncols + 127) / 128) * batch >= 2LL * ctx.cus;
    const int BM = short_tiles ? 32 : tall_tiles ? 128 : 64, BN = 128;
    const int tiles = ((m + BM - 1) / BM) * ((ncols + BN - 1) / BN);
    const int grid = (tiles + 7) / 8 * 8;
#define PEEC_H3(KC, MB, NB, ND)                                                                                     \
    hipLaunchKernelGGL((next ? gemm_next_h3_kernel<KC, MB, NB, ND>                                                 \
                        : b  ? gemm_tile_h3_kernel<KC, MB, NB, ND, 2>                                               \
                             : gemm_tile_h3_kernel<KC, MB, NB, ND, 1>),                                             \
                       dim3(grid, 1, batch), dim3(256), 0, s, A, lda, row0, col0, m, ncols, a, sb, status, bs)
#define PEEC_H3_TALL(KC)                                                                                            \
    hipLaunchKernelGGL((b ? gemm_tile_h3_kernel<KC, 4, 4, 2, 2> : gemm_tile_h3_kernel<KC, 4, 4, 2, 1>),            \
                       dim3(grid, 1, batch), dim3(256), 0, s, A, lda, row0, col0, m, ncols, a, sb, status, bs)
    if (tall_tiles && kw == 256)
        PEEC_H3_TALL(16);
    else if (tall_tiles && kw == 128)
        PEEC_H3_TALL(8);
    else if (tall_tiles && kw == 64)
        PEEC_H3_TALL(4);
    else if (kw == 256 && short_tiles)
        PEEC_H3(16, 1, 4, 4);
    else if (kw == 256)
        PEEC_H3(16, 2, 4, 4);
    else if (kw == 128 && short_tiles)
        PEEC_H3(8, 1, 4, 4);
    else if (kw == 128)
        PEEC_H3(8, 2, 4, 4);
    else if (kw == 64 && short_tiles)
        PEEC_H3(4, 1, 4, 4);
    else if (kw == 64)
        PEEC_H3(4, 2, 4, 4);
    else
        throw HipError{"split-FP16 update needs a k width of 64, 128 or 256"};
#undef PEEC_H3
#undef PEEC_H3_TALL
    HIP_CHECK(hipGetLastError());
    const int kt = b ? 2 * kw : kw;
    if (ctx.timer)
        ctx.timer->stop(tg, s, kind, 8.0 * batch * m * ncols * kt,
                        batch * (2.0 * m * ncols + (double)(m + ncols) * kt) * sizeof(float2));
}

// One segment from the context's L21 pack set `set` (rows from row0 are its
// row blocks rb_first..) and the given U12 pack.
static H3Seg h3_seg(LUContext &ctx, int set, int rb_first, int rb_pad, int cb_count, const uint4 *Uh,
                    const float *us) {
    H3Seg g;
    g.L = reinterpret_cast<const uint4 *>(set ? ctx.pack_l16b.ptr : ctx.pack_l16.ptr) + (size_t)rb_first * 128;
    g.linv = (set ? ctx.scale_l16b.ptr : ctx.scale_l16.ptr) + (size_t)rb_first * 16;
    g.U = Uh;
    g.uinv = us;
    g.rb_pad = rb_pad;
    g.cb_count = cb_count;
    return g;
}

static void gemm_h3(float2 *A, size_t lda, int row0, int col0, int m, int ncols, int kw, int rb_pad, int cb_count,
                    const uint4 *Uh, const float *us, KernelKind kind, const int *status, LUContext &ctx,
                    hipStream_t s, int rb_first = 0, const BatchStride &bs = BatchStride{}, int batch = 1,
                    int set = 0) {
    const H3Seg a = h3_seg(ctx, set, rb_first, rb_pad, cb_count, Uh, us);
    gemm_h3_segs(A, lda, row0, col0, m, ncols, kw, a, nullptr, kind, status, ctx, s, bs, batch);
}

// Split-FP16 operand pack of `nlines` lines of the k range k0 .. k0 + 16 CH - 1
// of A, in the layout pack_h3_regs writes: rows of L21 (ROWS, as (re, -im))
// or columns of U12 from line0, one wavefront per 16-line block, each line
// scaled by split_scale of its largest part over the k range; the blocks
// past nlines up to `blocks` (the pack's padding) are written as zeros.
template <bool ROWS, int CH>
__global__ __launch_bounds__(256) void pack_split16_kernel(const float2 *__restrict__ A, size_t lda, int line0,
                                                           int nlines, int k0, int blocks, uint4 *pk, float *pinv,
                                                           const int *status) {
    if (status[0] != 0) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const int blk = (int)blockIdx.x * 4 + wave;
    if (blk >= blocks) return;
    const int line = 16 * blk + li;
    const bool in = line < nlines;
    float vr[CH][4], vi[CH][4];
    float mx = 0.f;
#pragma unroll
    for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int k = 16 * c + 4 * lk + v;
            float2 x = make_float2(0.f, 0.f);
            if (in) x = ROWS ? A[(size_t)(line0 + line) * lda + k0 + k] : A[(size_t)(k0 + k) * lda + line0 + line];
            vr[c][v] = x.x;
            vi[c][v] = ROWS ? -x.y : x.y;
            mx = fmaxf(mx, fmaxf(fabsf(x.x), fabsf(x.y)));
        }
    mx = fmaxf(mx, __shfl_xor(mx, 16));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    const float sc = split_scale(mx);
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        unsigned hw[4], lw[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            _Float16 xh, xl, yh, yl;
            split_half(vr[c][v] * sc, xh, xl);
            split_half(vi[c][v] * sc, yh, yl);
            hw[v] = pack_half2(xh, yh);
            lw[v] = pack_half2(xl, yl);
        }
        uint4 *o = pk + ((size_t)(c * blocks + blk) * 2) * 64 + lane;
        o[0] = make_uint4(hw[0], hw[1], hw[2], hw[3]);
        o[64] = make_uint4(lw[0], lw[1], lw[2], lw[3]);
    }
    if (lk == 0) pinv[16 * blk + li] = 1.f / sc;
}

// A[k0 + w.., k0 + w..] -= L21 U12 of the partial-pivoting FP32 factor's
// block column at k0 (w = 64 or 128), on the split-FP16 operands (the same
// update kernel as the ordered factor's, k width w).
static void pivoted_update_h3(float2 *A, size_t lda, int k0, int w, int n, LUContext &ctx, hipStream_t s) {
    const int r0 = k0 + w, rest = n - r0;
    if (rest <= 0) return;
    int rb_pad, cb_pad;
    gemm_pads<float2>(rest, rest, rb_pad, cb_pad);
    const int cb = (rest + 127) / 128 * 8, kc = w / 16;
    ctx.pack_l16.reserve((size_t)kc * rb_pad * 128 * sizeof(uint4));
    ctx.scale_l16.reserve((size_t)rb_pad * 16);
    ctx.pack_u16.reserve((size_t)kc * cb * 128 * sizeof(uint4));
    ctx.scale_u16.reserve((size_t)cb * 16);
    uint4 *pl = reinterpret_cast<uint4 *>(ctx.pack_l16.ptr), *pu = reinterpret_cast<uint4 *>(ctx.pack_u16.ptr);
    if (w == 128) {
        hipLaunchKernelGGL((pack_split16_kernel<true, 8>), dim3((unsigned)((rb_pad + 3) / 4)), dim3(256), 0, s, A, lda,
                           r0, rest, k0, rb_pad, pl, ctx.scale_l16.ptr, ctx.status.ptr);
        hipLaunchKernelGGL((pack_split16_kernel<false, 8>), dim3((unsigned)((cb + 3) / 4)), dim3(256), 0, s, A, lda,
                           r0, rest, k0, cb, pu, ctx.scale_u16.ptr, ctx.status.ptr);
    } else {
        hipLaunchKernelGGL((pack_split16_kernel<true, 4>), dim3((unsigned)((rb_pad + 3) / 4)), dim3(256), 0, s, A, lda,
                           r0, rest, k0, rb_pad, pl, ctx.scale_l16.ptr, ctx.status.ptr);
        hipLaunchKernelGGL((pack_split16_kernel<false, 4>), dim3((unsigned)((cb + 3) / 4)), dim3(256), 0, s, A, lda,
                           r0, rest, k0, cb, pu, ctx.scale_u16.ptr, ctx.status.ptr);
    }
    HIP_CHECK(hipGetLastError());
    gemm_h3(A, lda, r0, r0, rest, rest, w, rb_pad, cb, pu, ctx.scale_u16.ptr, kKernelLuUpdateSplit16, ctx.status.ptr,
            ctx, s);
}

// ---------------------------------------------------------------------------
// factorization with a given row order (no pivot search)
//
// The rows have been permuted beforehand (convert_equilibrated_to_f32) into
// the order a previous partial-pivoting factorization chose; the factor is then a
// plain blocked LU: the w x w diagonal block is factored by one workgroup in
// LDS, L21 = A21 U11^{-1} and U12 = L11^{-1} A12 are formed in parallel, and
// the trailing update is the same MFMA GEMM. There is no cross-workgroup
// exchange. A pivot with |u_kk| <= limit stops it with status -2.
// Block width w <= 128 (the 32 x 32 tile grid of the diagonal-block kernels).

// |u| > limit (compared squared, in FP64) and 1 / u of a diagonal pivot. The
// FP32 reciprocal is conj(u) / |u|^2 with the hardware reciprocal of |u|^2
// and one Newton step (about as accurate as the division, a few dependent
// instructions instead of the division sequence); a pivot whose |u|^2 is not
// a normal FP32 number is rejected like a small one.
__device__ inline void pivot_check(float2 u, double limit, int &ok, float2 &rcp) {
    const double d2 = (double)u.x * u.x + (double)u.y * u.y;
    const float den = fmaf(u.x, u.x, u.y * u.y);
    ok = d2 > limit * limit && d2 < HUGE_VAL && den >= 1.17549435e-38f;
    float r = __builtin_amdgcn_rcpf(den);
    r = fmaf(r, fmaf(-den, r, 1.f), r);
    rcp = make_float2(u.x * r, -u.y * r);
}
__device__ inline void pivot_check(double2 u, double limit, int &ok, double2 &rcp) {
    const double m = wmag(u);
    ok = (m > limit) && isfinite(m);
    rcp = wdiv(wone<double2>(), u);
}
__device__ inline void pivot_check(double u, double limit, int &ok, double &rcp) {
    ok = fabs(u) > limit && isfinite(u);
    rcp = 1.0 / u;
}

// The ordered factor's w x w diagonal block (w <= 128), one workgroup of
// 1024 threads, in panels of 4 columns (FP32 complex storage and
// arithmetic): thread (ti, tc) keeps the 4 x 4 tile
// at rows 4 ti.., columns 4 tc.. (w x w block, padded to the tile grid with
// the identity). For panel q the owner of tile (q, q) factors it alone, the
// owners of the tiles left of and above the trailing square form their L21 =
// A21 U11^-1 and U12 = L11^-1 A12 tiles from the published L11, U11, and
// every trailing tile takes the rank-4 update A22 -= L21 U12 from LDS: two
// barriers per 4 columns, no per-element masks (panels align with tiles).
// Same operations as the column-by-column elimination, grouped by panel.
// With dinv (w a multiple of 16), the inverses of the factor's 16 x 16
// diagonal blocks follow, for the blocked triangular solves: U_b^-1 row-major
// at dinv[256 b], and (L_b^-1)^T (L_b unit lower) row-major at
// dinv[2048 + 256 b], each column by substitution in FP32.
__global__ __launch_bounds__(1024) void diag_lu_tiled_kernel(float2 *A, size_t lda, int k0, int w, double limit,
                                                             int *status, float2 *dinv, BatchStride bs) {
    // double-buffered by panel parity: L11 (unit lower) and U11 of the
    // diagonal tile, reciprocals of its pivots, L21 tiles by tile row and
    // U12 tiles by tile column
    __shared__ float2 s_l11[2][16], s_u11[2][16], s_rd[2][4];
    __shared__ float2 s_l21[2][32][16], s_u12[2][32][16];
    __shared__ int s_ok[2];
    bs.apply(A, status, dinv, limit);
    if (status[0] != 0) return;
    const int tid = threadIdx.x, wv = tid >> 6, ln = tid & 63;
    const int ti = 8 * (wv >> 2) + (ln >> 3), tc = 8 * (wv & 3) + (ln & 7);
    const int nt = (w + 3) / 4;
    // the tile as plain floats (arrays of vector types would go to scratch)
    float ar[4][4], ai[4][4];
    auto get = [&](int r, int c) { return make_float2(ar[r][c], ai[r][c]); };
    auto put = [&](int r, int c, float2 v) {
        ar[r][c] = v.x;
        ai[r][c] = v.y;
    };
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int i = 4 * ti + r, cc = 4 * tc + c;
            put(r, c, (i < w && cc < w) ? A[(size_t)(k0 + i) * lda + k0 + cc] : make_float2(i == cc ? 1.f : 0.f, 0.f));
        }
    for (int q = 0; q < nt; ++q) {
        const int b = q & 1;
        if (ti == q && tc == q) {
            // the diagonal tile: unpivoted LU in registers
            int ok = 1;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                int okc;
                float2 rd;
                // padding pivots are exactly 1: any positive magnitude passes
                pivot_check(get(c, c), 4 * q + c < w ? limit : 0.0, okc, rd);
                ok &= okc;
                s_rd[b][c] = rd;
#pragma unroll
                for (int r = c + 1; r < 4; ++r) {
                    const float2 m = wmul(get(r, c), rd);
                    put(r, c, m);
#pragma unroll
                    for (int x = c + 1; x < 4; ++x) put(r, x, wfms(get(r, x), m, get(c, x)));
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    s_l11[b][4 * r + c] = r > c ? get(r, c) : make_float2(0.f, 0.f);
                    s_u11[b][4 * r + c] = r <= c ? get(r, c) : make_float2(0.f, 0.f);
                }
            s_ok[b] = ok;
        }
        __syncthreads();
        if (!s_ok[b]) {
            if (tid == 0) {
                status[0] = -2;
                status[1] = k0 + 4 * q;
            }
            return;
        }
        if (tc == q && ti > q && ti < nt) {
            // L21 tile: row by row, x_c = (a_c - sum_{k<c} x_k u_kc) / u_cc
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    float2 v = get(r, c);
#pragma unroll
                    for (int k = 0; k < c; ++k) v = wfms(v, get(r, k), s_u11[b][4 * k + c]);
                    v = wmul(v, s_rd[b][c]);
                    put(r, c, v);
                    s_l21[b][ti][4 * r + c] = v;
                }
        } else if (ti == q && tc > q && tc < nt) {
            // U12 tile: column by column, y_r = a_r - sum_{k<r} l_rk y_k
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    float2 v = get(r, c);
#pragma unroll
                    for (int k = 0; k < r; ++k) v = wfms(v, s_l11[b][4 * r + k], get(k, c));
                    put(r, c, v);
                    s_u12[b][tc][4 * r + c] = v;
                }
        }
        __syncthreads();
        if (ti > q && tc > q && ti < nt && tc < nt) {
            // rank-4 update of the trailing tile, k = 0..3 outermost as in
            // the column-by-column elimination; one k's operands live at a
            // time (all four hoisted exceed the 128 registers a thread of a
            // 1024-thread workgroup has, and spill)
#pragma unroll 1
            for (int k = 0; k < 4; ++k) {
                float lr[4], li[4], ur[4], ui[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float2 v = s_l21[b][ti][4 * r + k];
                    lr[r] = v.x;
                    li[r] = v.y;
                }
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const float2 v = s_u12[b][tc][4 * k + c];
                    ur[c] = v.x;
                    ui[c] = v.y;
                }
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        put(r, c, wfms(get(r, c), make_float2(lr[r], li[r]), make_float2(ur[c], ui[c])));
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int i = 4 * ti + r, cc = 4 * tc + c;
            if (i < w && cc < w) A[(size_t)(k0 + i) * lda + k0 + cc] = get(r, c);
        }
    if (!dinv) return;
    // the 16 x 16 diagonal blocks, staged from the tiles that hold them
    __shared__ float2 s_db[8][16][17];
    const int nbk = w / 16;
    if ((ti >> 2) == (tc >> 2) && (ti >> 2) < nbk) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) s_db[ti >> 2][4 * (ti & 3) + r][4 * (tc & 3) + c] = get(r, c);
    }
    __syncthreads();
    if (tid >= 256) return;
    const int lower = tid >> 7, b = (tid >> 4) & 7, col = tid & 15;
    if (b >= nbk) return;
    float2 x[16];
    if (!lower) {
        // column col of U_b^-1: back substitution from row 15 (rows past
        // col come out exactly zero)
#pragma unroll
        for (int r = 15; r >= 0; --r) {
            float2 acc = make_float2(r == col ? 1.f : 0.f, 0.f);
#pragma unroll
            for (int k = r + 1; k < 16; ++k) acc = wfms(acc, s_db[b][r][k], x[k]);
            x[r] = wmul(acc, wdiv(wone<float2>(), s_db[b][r][r]));
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) dinv[256 * b + 16 * r + col] = x[r];
    } else {
        // column col of L_b^-1 (unit diagonal), forward substitution
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            float2 acc = make_float2(r == col ? 1.f : 0.f, 0.f);
#pragma unroll
            for (int k = 0; k < r; ++k) acc = wfms(acc, s_db[b][r][k], x[k]);
            x[r] = acc;
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) dinv[2048 + 256 * b + 16 * col + r] = x[r];
    }
}

// Blocked triangular solves of the ordered factor on FP32 MFMA
// (v_mfma_f32_16x16x4f32), for a block width w = 16 NBK: a wavefront holds
// 16 rows of L21 (16 columns of U12) as NBK accumulator tiles and walks the
// 16-wide diagonal blocks: X_j = (A_j - sum_{i<j} X_i U_ij) U_jj^-1 for
// L21 = A21 U11^-1, X_j = L_jj^-1 (A_j - sum_{i<j} L_ji X_i) for U12 =
// L11^-1 A12, with the diagonal-block inverses of diag_lu_tiled_kernel.
//
// Register layout: lane (li, lk) = (lane % 16, lane / 16) holds value v of
// tile j at L21 row li, column 16 j + 4 lk + v (U12 row 16 j + 4 lk + v,
// column li). That is the MFMA accumulator layout of the transposed product
// (row 4 lk + v, column li), and at the same time a B operand whose k index
// runs over {4 lk + t} in the t-th MFMA of a 16-deep product: the solved
// tile feeds the next products with no data movement. The A operands (U11
// blocks transposed, L11 blocks, the inverses) are read from global memory
// in the matching order (k = 4 lk + t at lane lk of the t-th MFMA).
//
// Per complex product C += A B: re += ar br - ai bi, im += ar bi + ai br,
// four real 16 x 16 x 4 MFMAs per t.
__device__ __forceinline__ void cmfma4(float ar, float ai, float br, float bi, v4f &cr, v4f &ci) {
    cr = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, br, cr, 0, 0, 0);
    cr = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, bi, cr, 0, 0, 0);
    ci = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, bi, ci, 0, 0, 0);
    ci = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, br, ci, 0, 0, 0);
}

// The ordered factor's diagonal block for w = 16 NBK (64 or 128), blocked by
// 16 columns, in one workgroup of eight wavefronts with the block in LDS
// (rows of W + 4: the MFMA operand and tile accesses below are free of bank
// conflicts). Per 16-column step p:
//  1. wavefront 0 factors the 16 x 16 diagonal tile in registers (lane
//     (li, lk) holds row li, columns 4 lk .. + 3; pivots and row entries
//     are passed between lanes, no barriers), with the pivot check of the
//     column-by-column elimination;
//  2. wavefronts 0 and 1 form U_pp^-1 and L_pp^-1 (L unit lower) the same
//     way; they also go to dinv (U^-1 row-major at 256 p, (L^-1)^T row-major
//     at 2048 + 256 p) for the blocked triangular solves;
//  3. the tiles below and right of it: L_ip = A_ip U_pp^-1 and U_pc =
//     L_pp^-1 A_pc, one FP32 MFMA product each (L_ip also kept transposed);
//  4. the trailing tiles: A_ic -= L_ip U_pc, one product each, spread over
//     the wavefronts.
// The same operations as the column-by-column elimination within a tile;
// across tiles the updates are grouped 16 columns at a time.
// lane N of each 16-lane row to the whole row (DPP row_newbcast)
template <int N> __device__ inline float row_bcast(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x150 + N, 0xF, 0xF, false));
}
template <int N> __device__ inline float2 row_bcast(float2 v) { return make_float2(row_bcast<N>(v.x), row_bcast<N>(v.y)); }

template <int NBK> __host__ __device__ constexpr size_t diag_blocked_lds() {
    return ((size_t)16 * NBK * (16 * NBK + 4) + (size_t)(NBK - 1) * 256 + 512) * sizeof(float2);
}
template <int NBK>
__global__ __launch_bounds__(512) void diag_lu_blocked_kernel(float2 *A, size_t lda, int k0, double limit,
                                                              int *status, float2 *dinv, BatchStride bs) {
    constexpr int W = 16 * NBK, S = W + 4;
    // the look-ahead stream's critical path: ahead of the trailing update's
    // waves on the CUs it shares with them
    __builtin_amdgcn_s_setprio(2);
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    float2 *sa = reinterpret_cast<float2 *>(smem_raw);  // the block, W x S
    float2 *slt = sa + (size_t)W * S;                   // L_ip transposed, tile i - p - 1
    float2 *suinv = slt + (NBK - 1) * 256;              // U_pp^-1, row-major
    float2 *slinvt = suinv + 256;                       // (L_pp^-1)^T, row-major
    __shared__ int s_bad;
    bs.apply(A, status, dinv, limit);
    if (status[0] != 0) return;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
    if (tid == 0) s_bad = -1;
    {
        // all loads in flight before the LDS stores
        constexpr int kPer = W * W / 512;
        float2 g[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int idx = tid + 512 * u, r = idx / W, c = idx % W;
            g[u] = A[(size_t)(k0 + r) * lda + k0 + c];
        }
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int idx = tid + 512 * u, r = idx / W, c = idx % W;
            sa[r * S + c] = g[u];
        }
    }
    __syncthreads();
    for (int p = 0; p < NBK; ++p) {
        const int d = 16 * p, nb = NBK - 1 - p;
        float2 *tile = sa + (size_t)(d + li) * S + d + 4 * lk;  // row layout of the diagonal tile
        float2 rcps[16];  // wavefront 0: reciprocals of the pivots, for U^-1
        if (wave == 0) {
            float2 a[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) a[v] = tile[v];
            int bad = -1;
            static_for<16>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                const int kq = k >> 2, kv = k & 3;
                const float2 u = make_float2(rdlane(a[kv].x, k + 16 * kq), rdlane(a[kv].y, k + 16 * kq));
                int okk;
                float2 rcp;
                pivot_check(u, limit, okk, rcp);
                rcps[k] = rcp;
                if (!okk && bad < 0) bad = k;
                const float2 colk = make_float2(__shfl(a[kv].x, li + 16 * kq), __shfl(a[kv].y, li + 16 * kq));
                const float2 mult = wmul(colk, rcp);
                float2 rowk[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) rowk[v] = row_bcast<k>(a[v]);
                if (li > k) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int col = 4 * lk + v;
                        if (col > k)
                            a[v] = wfms(a[v], mult, rowk[v]);
                        else if (col == k)
                            a[v] = mult;
                    }
                }
            });
#pragma unroll
            for (int v = 0; v < 4; ++v) tile[v] = a[v];
            if (bad >= 0 && lane == 0) s_bad = d + bad;
        }
        __syncthreads();
        if (s_bad >= 0) {
            if (tid == 0) {
                status[0] = -2;
                status[1] = k0 + s_bad;
            }
            return;
        }
        if (wave < 2) {
            float2 a[4], b[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                a[v] = tile[v];
                b[v] = make_float2(li == 4 * lk + v ? 1.f : 0.f, 0.f);
            }
            // column k of the factor in row li, for every k (not on the
            // dependency chain of the substitutions)
            float2 colv[16];
#pragma unroll
            for (int k = 0; k < 16; ++k)
                colv[k] = make_float2(__shfl(a[k & 3].x, li + 16 * (k >> 2)), __shfl(a[k & 3].y, li + 16 * (k >> 2)));
            if (wave == 0) {
                // U^-1 by rows from the last: row k final, then taken out of
                // the rows above
                static_for<16>([&](auto kc) {
                    constexpr int k = 15 - decltype(kc)::value;
                    float2 xk[4];
#pragma unroll
                    for (int v = 0; v < 4; ++v) xk[v] = wmul(row_bcast<k>(b[v]), rcps[k]);
                    const float2 uik = colv[k];
                    if (li == k) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) b[v] = xk[v];
                    } else if (li < k) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) b[v] = wfms(b[v], uik, xk[v]);
                    }
                });
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    suinv[li * 16 + 4 * lk + v] = b[v];
                    if (dinv) dinv[256 * p + 16 * li + 4 * lk + v] = b[v];
                }
            } else {
                // L^-1 (unit diagonal) by rows from the first
                static_for<16>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    float2 xk[4];
#pragma unroll
                    for (int v = 0; v < 4; ++v) xk[v] = row_bcast<k>(b[v]);
                    const float2 lik = colv[k];
                    if (li > k) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) b[v] = wfms(b[v], lik, xk[v]);
                    }
                });
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    slinvt[(4 * lk + v) * 16 + li] = b[v];
                    if (dinv) dinv[2048 + 256 * p + 16 * (4 * lk + v) + li] = b[v];
                }
            }
        }
        __syncthreads();
        // L_ip = A_ip U_pp^-1 (row layout) and U_pc = L_pp^-1 A_pc (column
        // layout)
        for (int it = wave; it < 2 * nb; it += 8) {
            v4f xr, xi, nr = {0.f, 0.f, 0.f, 0.f}, ni = {0.f, 0.f, 0.f, 0.f};
            if (it < nb) {
                const int i = p + 1 + it;
                float2 *x = sa + (size_t)(16 * i + li) * S + d + 4 * lk;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    xr[v] = x[v].x;
                    xi[v] = x[v].y;
                }
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const float2 m = suinv[(4 * lk + t) * 16 + li];
                    cmfma4(m.x, m.y, xr[t], xi[t], nr, ni);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const float2 y = make_float2(nr[v], ni[v]);
                    x[v] = y;
                    slt[(size_t)(i - p - 1) * 256 + (4 * lk + v) * 16 + li] = y;
                }
            } else {
                const int c = p + 1 + (it - nb);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const float2 e = sa[(size_t)(d + 4 * lk + v) * S + 16 * c + li];
                    xr[v] = e.x;
                    xi[v] = e.y;
                }
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const float2 m = slinvt[(4 * lk + t) * 16 + li];
                    cmfma4(m.x, m.y, xr[t], xi[t], nr, ni);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) sa[(size_t)(d + 4 * lk + v) * S + 16 * c + li] = make_float2(nr[v], ni[v]);
            }
        }
        __syncthreads();
        // A_ic -= L_ip U_pc (column layout)
        for (int it = wave; it < nb * nb; it += 8) {
            const int i = p + 1 + it / nb, c = p + 1 + it % nb;
            v4f cr, ci;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float2 e = sa[(size_t)(16 * i + 4 * lk + v) * S + 16 * c + li];
                cr[v] = e.x;
                ci[v] = e.y;
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float2 l = slt[(size_t)(i - p - 1) * 256 + (4 * lk + t) * 16 + li];
                const float2 u = sa[(size_t)(d + 4 * lk + t) * S + 16 * c + li];
                cmfma4(-l.x, -l.y, u.x, u.y, cr, ci);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) sa[(size_t)(16 * i + 4 * lk + v) * S + 16 * c + li] = make_float2(cr[v], ci[v]);
        }
        __syncthreads();
    }
    {
        constexpr int kPer = W * W / 512;
        float2 g[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int idx = tid + 512 * u, r = idx / W, c = idx % W;
            g[u] = sa[r * S + c];
        }
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int idx = tid + 512 * u, r = idx / W, c = idx % W;
            A[(size_t)(k0 + r) * lda + k0 + c] = g[u];
        }
    }
}

// 16 x 16 operand blocks in LDS: element (k, m) at ((k % 4) 4 + k / 4) 16 + m,
// so that the t-th MFMA's operand read (k = 4 lk + t at lane (m, lk)) is one
// contiguous 512-byte run of the wavefront.
__device__ __forceinline__ int blk_slot(int k, int m) { return (((k & 3) << 2) + (k >> 2)) * 16 + m; }
template <int NBK> __host__ __device__ constexpr size_t trsm_mfma_lds() {
    // the 16 x 16 operand blocks, then (split-FP16 solves) 16 row scales per block
    return (size_t)(NBK + NBK * (NBK - 1) / 2) * (256 * sizeof(float2) + 16 * sizeof(float));
}

// Split-FP16 products in the MFMA triangular solves. A 16 x 16 operand block
// staged in LDS in the FP32 layout above (operand entry A[m][k] at
// blk_slot(k, m)) is rewritten in place by one wavefront as the A operand of
// v_mfma_f32_16x16x32_f16: lane (m, lk) holds k = 4 lk .. + 3 as (re, -im)
// half pairs, the hi vectors at [lane] and the lo vectors at [64 + lane]
// (uint4 units, the same 2 KB), each row m scaled by split_scale of its
// largest part; sinv[m] = 1 / scale.
__device__ __forceinline__ void split_block_lds(float2 *blk, float *sinv, int lane) {
    const int li = lane & 15, lk = lane >> 4;
    float2 a[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) a[v] = blk[blk_slot(4 * lk + v, li)];
    float mx = 0.f;
#pragma unroll
    for (int v = 0; v < 4; ++v) mx = fmaxf(mx, fmaxf(fabsf(a[v].x), fabsf(a[v].y)));
    mx = fmaxf(mx, __shfl_xor(mx, 16));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    const float sc = split_scale(mx);
    unsigned hw[4], lw[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        _Float16 xh, xl, yh, yl;
        split_half(a[v].x * sc, xh, xl);
        split_half(-a[v].y * sc, yh, yl);
        hw[v] = pack_half2(xh, yh);
        lw[v] = pack_half2(xl, yl);
    }
    u32x4_t *o = reinterpret_cast<u32x4_t *>(blk);
    o[lane] = u32x4_t{hw[0], hw[1], hw[2], hw[3]};
    o[64 + lane] = u32x4_t{lw[0], lw[1], lw[2], lw[3]};
    if (lk == 0) sinv[li] = 1.f / sc;
}

// The B operand of a wavefront's 16 lines x 16 k (lane (li, lk): line li, k =
// 4 lk + v in re[v], im[v]) as split-FP16 halves, the line scaled by
// split_scale of its largest part; returns 1 / scale.
__device__ __forceinline__ float split_lines(const v4f &re, const v4f &im, u32x4_t &hi, u32x4_t &lo) {
    float mx = 0.f;
#pragma unroll
    for (int v = 0; v < 4; ++v) mx = fmaxf(mx, fmaxf(fabsf(re[v]), fabsf(im[v])));
    mx = fmaxf(mx, __shfl_xor(mx, 16));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    const float sc = split_scale(mx);
    unsigned hw[4], lw[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        _Float16 xh, xl, yh, yl;
        split_half(re[v] * sc, xh, xl);
        split_half(im[v] * sc, yh, yl);
        hw[v] = pack_half2(xh, yh);
        lw[v] = pack_half2(xl, yl);
    }
    hi = u32x4_t{hw[0], hw[1], hw[2], hw[3]};
    lo = u32x4_t{lw[0], lw[1], lw[2], lw[3]};
    return 1.f / sc;
}

// (cr, ci) = A B of one 16 x 16 block pair on split operands (A from a block
// rewritten by split_block_lds, B from split_lines): hi hi + hi lo + lo hi
// for the real part and for the imaginary part (A's pairs turned).
__device__ __forceinline__ void cmul_h3(const u32x4_t *ablk, int lane, const u32x4_t &bh, const u32x4_t &bl, v4f &cr,
                                        v4f &ci) {
    const u32x4_t ah = ablk[lane], al = ablk[64 + lane];
    const half8_t rh = __builtin_bit_cast(half8_t, ah), rl = __builtin_bit_cast(half8_t, al);
    const half8_t ih = __builtin_bit_cast(half8_t, turn_pairs(ah)), il = __builtin_bit_cast(half8_t, turn_pairs(al));
    const half8_t hb = __builtin_bit_cast(half8_t, bh), lb = __builtin_bit_cast(half8_t, bl);
    const v4f z = {0.f, 0.f, 0.f, 0.f};
    cr = __builtin_amdgcn_mfma_f32_16x16x32_f16(rh, hb, z, 0, 0, 0);
    ci = __builtin_amdgcn_mfma_f32_16x16x32_f16(ih, hb, z, 0, 0, 0);
    cr = __builtin_amdgcn_mfma_f32_16x16x32_f16(rh, lb, cr, 0, 0, 0);
    ci = __builtin_amdgcn_mfma_f32_16x16x32_f16(ih, lb, ci, 0, 0, 0);
    cr = __builtin_amdgcn_mfma_f32_16x16x32_f16(rl, hb, cr, 0, 0, 0);
    ci = __builtin_amdgcn_mfma_f32_16x16x32_f16(il, hb, ci, 0, 0, 0);
}

// The block solve of a wavefront's 16 lines on split-FP16 products: for j in
// order, X_j = A_j B_j^-1 (inverse block j of s_inv, as the operand of the
// transposed product), then A_c -= X_j B_jc (block p of s_off) for c > j; the
// products are unscaled, (row scale of the operand block)(line scale of X),
// exact powers of two.
template <int NBK>
__device__ __forceinline__ void solve_lines_h3(v4f (&xr)[NBK], v4f (&xi)[NBK], const float2 *s_inv, const float2 *s_off,
                                               const float *sc_inv, const float *sc_off, int lane) {
    const int lk = lane >> 4;
#pragma unroll
    for (int j = 0; j < NBK; ++j) {
        u32x4_t bh, bl;
        float fb = split_lines(xr[j], xi[j], bh, bl);
        v4f cr, ci;
        cmul_h3(reinterpret_cast<const u32x4_t *>(s_inv + 256 * j), lane, bh, bl, cr, ci);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const float f = sc_inv[16 * j + 4 * lk + v] * fb;
            xr[j][v] = cr[v] * f;
            xi[j][v] = ci[v] * f;
        }
        if (j + 1 < NBK) {
            fb = split_lines(xr[j], xi[j], bh, bl);
#pragma unroll
            for (int c = j + 1; c < NBK; ++c) {
                const int p = j * (NBK - 1) - j * (j - 1) / 2 + (c - j - 1);
                cmul_h3(reinterpret_cast<const u32x4_t *>(s_off + 256 * p), lane, bh, bl, cr, ci);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const float f = sc_off[16 * p + 4 * lk + v] * fb;
                    xr[c][v] -= cr[v] * f;
                    xi[c][v] -= ci[v] * f;
                }
            }
        }
    }
}

// The split-FP16 operand block of a solved 16-row (16-column) block straight
// from the solve's registers: lane (li, lk) holds entries k = 16 j + 4 lk + v
// of line li (L21 row / U12 column), the lane order of the operand pack
// (chunk j, quarter lk, entry v). With NP > 0, the NP chunks of the previous
// panel of a pair (pre, already solved and stored) come first, and the scale
// of a line covers both panels. Lines past the matrix (in false) are packed
// as zeros; NEG packs the pairs (re, -im) (the L21 operand). Chunk c goes to
// pk[((c stride + blk) 2) 64 + lane] (hi) and + 64 (lo); pinv[16 blk + li] =
// 1 / scale.
template <int NBK, int NP, bool NEG>
__device__ __forceinline__ void pack_h3_regs(const v4f (&xr)[NBK], const v4f (&xi)[NBK],
                                             const float2 (&pre)[NP > 0 ? NP : 1][4], bool in, uint4 *pk,
                                             float *pinv, int stride, int blk, int lane) {
    const int li = lane & 15, lk = lane >> 4;
    float mx = 0.f;
#pragma unroll
    for (int j = 0; j < NBK; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) mx = fmaxf(mx, fmaxf(fabsf(xr[j][v]), fabsf(xi[j][v])));
    if constexpr (NP > 0) {
#pragma unroll
        for (int j = 0; j < NP; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) mx = fmaxf(mx, fmaxf(fabsf(pre[j][v].x), fabsf(pre[j][v].y)));
    }
    if (!in) mx = 0.f;
    mx = fmaxf(mx, __shfl_xor(mx, 16));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    const float sc = split_scale(mx);
    auto emit = [&](int c, const float2 (&val)[4]) {
        unsigned hw[4], lw[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            _Float16 xh, xl, yh, yl;
            split_half(in ? val[v].x * sc : 0.f, xh, xl);
            split_half(in ? (NEG ? -val[v].y : val[v].y) * sc : 0.f, yh, yl);
            hw[v] = pack_half2(xh, yh);
            lw[v] = pack_half2(xl, yl);
        }
        uint4 *o = pk + ((size_t)(c * stride + blk) * 2) * 64 + lane;
        o[0] = make_uint4(hw[0], hw[1], hw[2], hw[3]);
        o[64] = make_uint4(lw[0], lw[1], lw[2], lw[3]);
    };
    if constexpr (NP > 0) {
#pragma unroll
        for (int j = 0; j < NP; ++j) emit(j, pre[j]);
    }
#pragma unroll
    for (int j = 0; j < NBK; ++j) {
        const float2 val[4] = {make_float2(xr[j][0], xi[j][0]), make_float2(xr[j][1], xi[j][1]),
                               make_float2(xr[j][2], xi[j][2]), make_float2(xr[j][3], xi[j][3])};
        emit(NP + j, val);
    }
    if (lk == 0) pinv[16 * blk + li] = 1.f / sc;
}

// L21 rows r_begin .. r_end - 1 (below the diagonal block at k0), in place;
// workgroups of four wavefronts, 16 rows each. The workgroup stages the
// inverses U_jj^-1 and the blocks U_jc (j < c) of U11 in LDS first, all loads
// in flight together. With pk, the solved rows are also written as the
// split-FP16 L21 operand of the update (row block (r_begin + 16 t) / 16 for
// the t-th wavefront, rb_pad blocks per chunk); with PRE, the k range of the
// operand starts at k0 - 16 NBK (the first panel of a pair, whose L21 is in
// place) and covers both panels.
template <int NBK, bool PRE, bool H3 = true>
__global__ __launch_bounds__(256) void trsm_right_mfma_kernel(float2 *A, size_t lda, int k0, int r_begin, int r_end,
                                                              const float2 *__restrict__ dinv, const int *status,
                                                              uint4 *pk, float *pinv, int rb_pad, BatchStride bs) {
    constexpr int W = 16 * NBK, NU = NBK * (NBK - 1) / 2, NP = PRE ? NBK : 0;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    float2 *s_inv = reinterpret_cast<float2 *>(smem_raw), *s_u = s_inv + NBK * 256;
    bs.apply(A, status, dinv);
    bs.apply_pack(pk, pinv, bs.pl, bs.sl);
    if (status[0] != 0) return;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
    const int rb = r_begin + ((int)blockIdx.x * 4 + wave) * 16;
    const bool live = rb < r_end, in = rb + li < r_end;
    // lanes past r_end work on a copy of row rb and store nothing
    float2 *row = A + (size_t)(k0 + W + (in ? rb + li : live ? rb : r_begin)) * lda + k0;
    const float2 *U = A + (size_t)k0 * lda + k0;
    const int er = tid >> 4, ec = tid & 15;  // the element of each block this thread stages
    float2 g[NBK + NU];
#pragma unroll
    for (int j = 0; j < NBK; ++j) g[j] = dinv[256 * j + tid];  // U_jj^-1 [er][ec]
    {
        int q = NBK;
#pragma unroll
        for (int j = 0; j < NBK; ++j)
#pragma unroll
            for (int c = j + 1; c < NBK; ++c) g[q++] = U[(size_t)(16 * j + er) * lda + 16 * c + ec];
    }
    v4f xr[NBK], xi[NBK];
#pragma unroll
    for (int j = 0; j < NBK; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const float2 a = row[16 * j + 4 * lk + v];
            xr[j][v] = a.x;
            xi[j][v] = a.y;
        }
    float2 pre[NP > 0 ? NP : 1][4];
    if constexpr (NP > 0) {
        if (pk) {
#pragma unroll
            for (int j = 0; j < NP; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) pre[j][v] = row[16 * j + 4 * lk + v - W];
        }
    }
#pragma unroll
    for (int q = 0; q < NBK + NU; ++q) s_inv[q * 256 + blk_slot(er, ec)] = g[q];
    __syncthreads();
    float *s_sc = reinterpret_cast<float *>(s_inv + (NBK + NU) * 256);
    if constexpr (H3) {
        // the blocks as split-FP16 operands (one wavefront per block)
        for (int q = wave; q < NBK + NU; q += 4) split_block_lds(s_inv + 256 * q, s_sc + 16 * q, lane);
        __syncthreads();
    }
    if (!live) return;
    const int lo = lk * 16 + li;  // + 64 t: operand of the t-th MFMA
    if constexpr (H3) solve_lines_h3<NBK>(xr, xi, s_inv, s_u, s_sc, s_sc + 16 * NBK, lane);
#pragma unroll
    for (int j = 0; j < (H3 ? 0 : NBK); ++j) {
        // X_j^T = (U_jj^-1)^T (A_j - ...)^T
        v4f nr = {0.f, 0.f, 0.f, 0.f}, ni = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const float2 a = s_inv[256 * j + 64 * t + lo];
            cmfma4(a.x, a.y, xr[j][t], xi[j][t], nr, ni);
        }
        xr[j] = nr;
        xi[j] = ni;
        // A_c -= X_j U_jc for the tiles right of it: (U_jc)^T X_j^T
#pragma unroll
        for (int c = j + 1; c < NBK; ++c) {
            const int p = j * (NBK - 1) - j * (j - 1) / 2 + (c - j - 1);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float2 a = s_u[256 * p + 64 * t + lo];
                cmfma4(-a.x, -a.y, xr[j][t], xi[j][t], xr[c], xi[c]);
            }
        }
    }
    if (in) {
#pragma unroll
        for (int j = 0; j < NBK; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) row[16 * j + 4 * lk + v] = make_float2(xr[j][v], xi[j][v]);
    }
    if (pk) pack_h3_regs<NBK, NP, true>(xr, xi, pre, in, pk, pinv, rb_pad, rb / 16, lane);
}

// U12 columns c0 .. c1 - 1 (right of the diagonal block at k0), in place;
// one wavefront per 16 columns, four per workgroup, with (L_jj^-1)^T and the
// blocks L_cj (c > j) of L11, transposed, staged in LDS. With pk, the solved
// columns are also written as the split-FP16 U12 operand of the update
// (block `tile` = the wavefront's 16-column block from c0, pcb blocks per
// chunk); with PRE its k range starts at k0 - 16 NBK (both panels of a pair).
template <int NBK, bool PRE, bool H3 = true>
__global__ __launch_bounds__(256) void trsm_left_mfma_kernel(float2 *A, size_t lda, int k0, int c0, int c1,
                                                             const float2 *__restrict__ dinv, const int *status,
                                                             uint4 *pk, float *pinv, int pcb, BatchStride bs) {
    constexpr int W = 16 * NBK, NU = NBK * (NBK - 1) / 2, NP = PRE ? NBK : 0;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    float2 *s_inv = reinterpret_cast<float2 *>(smem_raw), *s_l = s_inv + NBK * 256;
    bs.apply(A, status, dinv);
    bs.apply_pack(pk, pinv, bs.pu, bs.su);
    if (status[0] != 0) return;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
    const int tile = (int)blockIdx.x * 4 + wave, cb = c0 + 16 * tile;
    const bool live = cb < c1, in = cb + li < c1;
    const int col = in ? cb + li : live ? cb : c0;
    float2 *colp = A + (size_t)k0 * lda + col;
    const float2 *L = A + (size_t)k0 * lda + k0;
    const int er = tid >> 4, ec = tid & 15;
    float2 g[NBK + NU];
#pragma unroll
    for (int j = 0; j < NBK; ++j) g[j] = dinv[2048 + 256 * j + tid];  // (L_jj^-1)^T [er][ec]
    {
        int q = NBK;
#pragma unroll
        for (int j = 0; j < NBK; ++j)
#pragma unroll
            for (int c = j + 1; c < NBK; ++c) g[q++] = L[(size_t)(16 * c + er) * lda + 16 * j + ec];  // L_cj[er][ec]
    }
    v4f xr[NBK], xi[NBK];
#pragma unroll
    for (int j = 0; j < NBK; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const float2 a = colp[(size_t)(16 * j + 4 * lk + v) * lda];
            xr[j][v] = a.x;
            xi[j][v] = a.y;
        }
    float2 pre[NP > 0 ? NP : 1][4];
    if constexpr (NP > 0) {
        if (pk) {
            const float2 *pp = A + (size_t)(k0 - W) * lda + col;
#pragma unroll
            for (int j = 0; j < NP; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) pre[j][v] = pp[(size_t)(16 * j + 4 * lk + v) * lda];
        }
    }
#pragma unroll
    for (int j = 0; j < NBK; ++j) s_inv[j * 256 + blk_slot(er, ec)] = g[j];
    // L_cj stored transposed: the A operand's (k, m) = L_cj[m][k]
#pragma unroll
    for (int q = 0; q < NU; ++q) s_l[q * 256 + blk_slot(ec, er)] = g[NBK + q];
    __syncthreads();
    float *s_sc = reinterpret_cast<float *>(s_inv + (NBK + NU) * 256);
    if constexpr (H3) {
        // the blocks as split-FP16 operands (one wavefront per block)
        for (int q = wave; q < NBK + NU; q += 4) split_block_lds(s_inv + 256 * q, s_sc + 16 * q, lane);
        __syncthreads();
    }
    if (!live) return;
    const int lo = lk * 16 + li;
    if constexpr (H3) solve_lines_h3<NBK>(xr, xi, s_inv, s_l, s_sc, s_sc + 16 * NBK, lane);
#pragma unroll
    for (int j = 0; j < (H3 ? 0 : NBK); ++j) {
        v4f nr = {0.f, 0.f, 0.f, 0.f}, ni = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const float2 a = s_inv[256 * j + 64 * t + lo];
            cmfma4(a.x, a.y, xr[j][t], xi[j][t], nr, ni);
        }
        xr[j] = nr;
        xi[j] = ni;
        // A_c -= L_cj X_j for the tiles below it
#pragma unroll
        for (int c = j + 1; c < NBK; ++c) {
            const int p = j * (NBK - 1) - j * (j - 1) / 2 + (c - j - 1);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float2 a = s_l[256 * p + 64 * t + lo];
                cmfma4(-a.x, -a.y, xr[j][t], xi[j][t], xr[c], xi[c]);
            }
        }
    }
    if (in) {
#pragma unroll
        for (int j = 0; j < NBK; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) colp[(size_t)(16 * j + 4 * lk + v) * lda] = make_float2(xr[j][v], xi[j][v]);
    }
    if (pk) pack_h3_regs<NBK, NP, false>(xr, xi, pre, in, pk, pinv, pcb, tile, lane);
}

// Triangular solves of the ordered factor, one wavefront per row of
// L21 = A21 U11^{-1} or per column of U12 = L11^{-1} A12 (column-by-column
// substitution in the order of trsm_kernel: x_j = a_j / u_jj once the
// updates of every earlier j are in, then a_c -= x_j u_jc for c > j, and
// likewise a_i -= l_ir x_r for i > r). Lane l keeps entries l and l + 64 of
// the vector; the finished entry is broadcast with a lane read (a scalar
// register, no LDS round trip) and every lane updates its entries past it
// against row j of U11 (column r of L11, staged transposed). Workgroups of
// 16 wavefronts stage the triangle once in LDS and walk their share of the
// rows (columns).
template <class S>
__global__ __launch_bounds__(1024) void trsm_right_wave_kernel(S *A, size_t lda, int k0, int w, int r_begin,
                                                               int r_end, const int *status) {
    using T = typename Work<S>::T;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    S *blk = reinterpret_cast<S *>(smem_raw);  // U11 row-major, w x w
    __shared__ T rcp[kMaxNB];
    if (status[0] != 0) return;
    {
        constexpr int kPer = kMaxNB * kMaxNB / 1024;
        Regs<S, kPer> v;
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int idx = (int)threadIdx.x + u * 1024;
            v.set(u, idx < w * w ? A[(size_t)(k0 + idx / w) * lda + k0 + idx % w] : S{});
        }
        // only the strict upper triangle: the zeros on and below the
        // diagonal make every step's update a no-op for the entries already
        // solved, so the steps need no per-lane selects for them
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int idx = (int)threadIdx.x + u * 1024;
            if (idx < w * w) blk[idx] = idx % w > idx / w ? v.get(u) : S{};
        }
    }
    if ((int)threadIdx.x < w)
        rcp[threadIdx.x] = wdiv(wone<T>(), wld(A + (size_t)(k0 + threadIdx.x) * lda + k0 + threadIdx.x));
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ca = lane, cb = lane + 64;
    const int cla = min(ca, w - 1), clb = min(cb, w - 1);
    // rows k0 + w + r_begin .. k0 + w + r_end - 1 of L21
    for (int rr = r_begin + (int)blockIdx.x * 16 + wave; rr < r_end; rr += (int)gridDim.x * 16) {
        S *row = A + (size_t)(k0 + w + rr) * lda + k0;
        T x0 = ca < w ? wld(row + ca) : wzero<T>();
        T x1 = cb < w ? wld(row + cb) : wzero<T>();
        // four steps per iteration (unrolled by hand: the lane reads are
        // convergent, which keeps the compiler from unrolling), so that their
        // LDS reads, independent of x, are in flight ahead of the dependent
        // lane-read chain; steps past w - 1 change nothing. Columns 0..63
        // (x0) are final after step 63 and columns 64.. (x1) are never
        // solved before it, so the steps run in two phases that leave out
        // the other half's updates (same values as one loop). Lanes past w
        // compute values that are never read or stored.
        const int wa = min(w, 64);
        for (int j0 = 0; j0 < wa; j0 += 4) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = j0 + u, jc = min(j, w - 1);
                const T xj = wmul(wreadlane(x0, j & 63), rcp[jc]);
                x0 = (ca == j) ? xj : x0;
                const T u0 = wld(blk + jc * w + cla), u1 = wld(blk + jc * w + clb);
                x0 = wfms(x0, xj, u0);
                x1 = wfms(x1, xj, u1);
            }
        }
        for (int j0 = 64; j0 < w; j0 += 4) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = j0 + u, jc = min(j, w - 1);
                const T xj = wmul(wreadlane(x1, j & 63), rcp[jc]);
                x1 = (cb == j) ? xj : x1;
                const T u1 = wld(blk + jc * w + clb);
                x1 = wfms(x1, xj, u1);
            }
        }
        if (ca < w) wst(row + ca, x0);
        if (cb < w) wst(row + cb, x1);
    }
}

constexpr int kTrsmWaveCols = 16;  // columns of A12 per workgroup pass

template <class S>
__global__ __launch_bounds__(1024) void trsm_wave_kernel(S *A, size_t lda, int c0, int c1, int k0, int w,
                                                         const int *status, uint4 *pk, float *pinv, int pcb) {
    using T = typename Work<S>::T;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    const int wp = w + 1;                      // padded rows: transposed writes spread over the banks
    S *lt = reinterpret_cast<S *>(smem_raw);   // lt[r * wp + i] = L11[i][r]
    S *tb = lt + (size_t)w * wp;               // tb[c * wp + i] = A12[i][c] of the current tile
    if (status[0] != 0) return;
    {
        constexpr int kPer = kMaxNB * kMaxNB / 1024;
        Regs<S, kPer> v;
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int idx = (int)threadIdx.x + u * 1024;
            v.set(u, idx < w * w ? A[(size_t)(k0 + idx / w) * lda + k0 + idx % w] : S{});
        }
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int idx = (int)threadIdx.x + u * 1024;
            // the strict lower triangle only (zeros elsewhere make every
            // step a no-op for the rows already solved: no per-lane selects)
            if (idx < w * w) lt[(idx % w) * wp + idx / w] = idx / w > idx % w ? v.get(u) : S{};
        }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ra = lane, rb = lane + 64;
    const int rla = min(ra, w - 1), rlb = min(rb, w - 1);
    const int ncols = c1 - c0, tiles = (ncols + kTrsmWaveCols - 1) / kTrsmWaveCols;
    constexpr int kTileElems = kMaxNB * kTrsmWaveCols / 1024;
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const int cbase = c0 + tile * kTrsmWaveCols;
        __syncthreads();  // the triangle is staged; the previous tile is stored
#pragma unroll
        for (int u = 0; u < kTileElems; ++u) {
            const int idx = (int)threadIdx.x + u * 1024, i = idx / kTrsmWaveCols, c = idx % kTrsmWaveCols;
            if (i < w && cbase + c < c1) tb[c * wp + i] = A[(size_t)(k0 + i) * lda + cbase + c];
        }
        __syncthreads();
        if (cbase + wave < c1) {
            S *col = tb + wave * wp;
            T x0 = ra < w ? wld(col + ra) : wzero<T>();
            T x1 = rb < w ? wld(col + rb) : wzero<T>();
            // four steps per iteration, in two phases, as in
            // trsm_right_wave_kernel (rows 0..63 are final after step 63)
            const int wa = min(w - 1, 64);
            for (int r0 = 0; r0 < wa; r0 += 4) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int r = r0 + u, rc = min(r, w - 1);
                    const T xr = wreadlane(x0, r & 63);
                    const T l0 = wld(lt + rc * wp + rla), l1 = wld(lt + rc * wp + rlb);
                    x0 = wfms(x0, l0, xr);
                    x1 = wfms(x1, l1, xr);
                }
            }
            for (int r0 = 64; r0 + 1 < w; r0 += 4) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int r = r0 + u, rc = min(r, w - 1);
                    const T xr = wreadlane(x1, r & 63);
                    const T l1 = wld(lt + rc * wp + rlb);
                    x1 = wfms(x1, l1, xr);
                }
            }
            if (ra < w) wst(col + ra, x0);
            if (rb < w) wst(col + rb, x1);
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kTileElems; ++u) {
            const int idx = (int)threadIdx.x + u * 1024, i = idx / kTrsmWaveCols, c = idx % kTrsmWaveCols;
            if (i < w && cbase + c < c1) A[(size_t)(k0 + i) * lda + cbase + c] = tb[c * wp + i];
        }
        if constexpr (std::is_same<S, float2>::value) {
            // with pk: the solved tile is also the split-FP16 operand block
            // `tile` of the trailing update (as pack_u_h3_wave would write
            // it), packed by the first wavefront straight from LDS
            if (pk && wave == 0) {
                const int col = lane & 15, q = lane >> 4, kc = w / 16;
                const bool cv = cbase + col < c1;
                float2 v[8][4];
                float mx = 0.f;
#pragma unroll
                for (int c = 0; c < 8; ++c)
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        v[c][t] = (c < kc && cv) ? tb[col * wp + 16 * c + 4 * q + t] : make_float2(0.f, 0.f);
                        mx = fmaxf(mx, fmaxf(fabsf(v[c][t].x), fabsf(v[c][t].y)));
                    }
                mx = fmaxf(mx, __shfl_xor(mx, 16));
                mx = fmaxf(mx, __shfl_xor(mx, 32));
                const float sc = split_scale(mx);
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    if (c >= kc) break;
                    unsigned hw[4], lw[4];
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        _Float16 xh, xl, yh, yl;
                        split_half(v[c][t].x * sc, xh, xl);
                        split_half(v[c][t].y * sc, yh, yl);
                        hw[t] = pack_half2(xh, yh);
                        lw[t] = pack_half2(xl, yl);
                    }
                    uint4 *o = pk + ((size_t)(c * pcb + tile) * 2) * 64 + lane;
                    o[0] = make_uint4(hw[0], hw[1], hw[2], hw[3]);
                    o[64] = make_uint4(lw[0], lw[1], lw[2], lw[3]);
                }
                if (q == 0) pinv[tile * 16 + col] = 1.f / sc;
            }
        }
    }
}

__host__ __device__ constexpr size_t trsm_wave_lds(int w) {
    return (size_t)w * (w + 1) + (size_t)kTrsmWaveCols * (w + 1);
}

static void set_dynamic_lds(const void *fn, size_t bytes) {
    if (bytes > 64 * 1024) HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

// Look-ahead: after the panel's triangular solves, the next block column is
// updated first and its diagonal block is factored on the context's second
// stream while the rest of the trailing update runs on the first; the next
// panel waits for that factor through an event.
LUStatus lu_factor_ordered(float2 *dA, int n, size_t lda, double pivot_limit, LUContext &ctx, int block_size,
                           int batch, size_t batch_stride, const double *d_limits, LUStatus *member_status) {
    using S = float2;
    hipStream_t s = ctx.stream, aux = ctx.aux;
    LUStatus out;
    if (batch < 1) throw HipError{"ordered factor: empty batch"};
    ctx.status.reserve((size_t)2 * batch);
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, s, ctx.status.ptr, 2 * batch, 0);
    int nb = block_size > 0 ? std::min(block_size, kMaxNB) : kMaxNB;
    while (nb > 16 && (size_t)nb * nb * sizeof(S) > (size_t)ctx.lds_per_block) nb /= 2;
    const size_t split_lds = (size_t)4 * 2 * 4096 * sizeof(float);
    if (split_lds > (size_t)ctx.lds_per_block) {
        PEEC_ERR("the ordered factor needs %zu bytes of LDS per workgroup.", split_lds);
        out.code = -1;
        return out;
    }
    set_dynamic_lds(reinterpret_cast<const void *>(trsm_right_wave_kernel<S>), (size_t)nb * nb * sizeof(S));
    set_dynamic_lds(reinterpret_cast<const void *>(trsm_wave_kernel<S>), trsm_wave_lds(nb) * sizeof(S));
    set_dynamic_lds(reinterpret_cast<const void *>(gemm_splitk_kernel<S>), split_lds);
    const bool split16 = ctx.split_fp16_update;
    // triangular solves: one wavefront per row of L21 / column of U12, at
    // most one workgroup of 16 per CU
    // rows r_begin .. r_end - 1 below the diagonal block (all of them by default)
    // blocked MFMA solves for block widths of 32, 64 and 128 (the diagonal
    // factor then also leaves its 16 x 16 block inverses, in the set of the
    // panel's parity); substitution kernels otherwise
    const bool blocked_solves = nb % 16 == 0 && (nb / 16 == 2 || nb / 16 == 4 || nb / 16 == 8);
    ctx.dblk_inv.reserve((size_t)2 * 4096 * batch);
    // members of a batch: the matrices batch_stride elements apart, each
    // with its own status pair, inverse sets, operand packs and pivot limit
    BatchStride bs;
    if (batch > 1) {
        bs.a = batch_stride;
        bs.status = 2;
        bs.dinv = 2 * 4096;
        bs.limits = d_limits;
    }
    // the same with the strides of the L21 pack and of the next block
    // column's / the rest's U12 pack (set with the packs below)
    BatchStride bs_n = bs, bs_r = bs;
    if (blocked_solves && nb == 128) {
        set_dynamic_lds(reinterpret_cast<const void *>(trsm_right_mfma_kernel<8, false>), trsm_mfma_lds<8>());
        set_dynamic_lds(reinterpret_cast<const void *>(trsm_right_mfma_kernel<8, true>), trsm_mfma_lds<8>());
        set_dynamic_lds(reinterpret_cast<const void *>(trsm_left_mfma_kernel<8, false>), trsm_mfma_lds<8>());
        set_dynamic_lds(reinterpret_cast<const void *>(trsm_left_mfma_kernel<8, true>), trsm_mfma_lds<8>());
    }
    auto dinv_of = [&](int k0) { return ctx.dblk_inv.ptr + (size_t)((k0 / nb) & 1) * 4096; };
    // With pk (block widths 64 and 128), the solved rows of L21 are also
    // written as the split-FP16 operand (stride row blocks per chunk), with
    // pre over the previous panel's k range as well.
    auto solve_l21 = [&](int k0, int w, int rest, int r_begin = 0, int r_end = -1, uint4 *pk = nullptr,
                         float *pinv = nullptr, int stride = 0, bool pre = false) {
        if (r_end < 0) r_end = rest;
        if (blocked_solves && w == nb) {
            const unsigned grid = (unsigned)((r_end - r_begin + 63) / 64);
#define PEEC_TRSM_R(K, P)                                                                                  \
    hipLaunchKernelGGL((trsm_right_mfma_kernel<K, P>), dim3(grid, 1, batch), dim3(256), trsm_mfma_lds<K>(), s, dA, \
                       lda, k0, r_begin, r_end, dinv_of(k0), ctx.status.ptr, pk, pinv, stride, bs_n)
            if (w == 128 && pre)
                PEEC_TRSM_R(8, true);
            else if (w == 128)
                PEEC_TRSM_R(8, false);
            else if (w == 64 && pre)
                PEEC_TRSM_R(4, true);
            else if (w == 64)
                PEEC_TRSM_R(4, false);
            else if (!pk)
                PEEC_TRSM_R(2, false);
            else
                throw HipError{"ordered factor: packed L21 solves need a block width of 64 or 128"};
#undef PEEC_TRSM_R
            return;
        }
        if (pk) throw HipError{"ordered factor: packed L21 solves need the blocked solve kernels"};
        if (batch > 1) throw HipError{"ordered factor: a batch needs the blocked solve kernels"};
        const int grid = std::max(1, std::min((r_end - r_begin + 15) / 16, ctx.cus));
        hipLaunchKernelGGL(trsm_right_wave_kernel<S>, dim3((unsigned)grid), dim3(1024), (size_t)w * w * sizeof(S), s,
                           dA, lda, k0, w, r_begin, r_end, ctx.status.ptr);
    };
    // with pk, U12 is also packed as the split-FP16 operand of pcb blocks
    // (and with pre, over the previous panel's k range as well)
    auto solve_u12 = [&](int k0, int w, int c0, int c1, hipStream_t st, uint4 *pk = nullptr, float *pinv = nullptr,
                         int pcb = 0, bool pre = false) {
        if (blocked_solves && w == nb) {
            const unsigned grid = (unsigned)((c1 - c0 + 63) / 64);
            const BatchStride &bu = pk == reinterpret_cast<uint4 *>(ctx.pack_u16n.ptr) ? bs_n : bs_r;
#define PEEC_TRSM_L(K, P)                                                                                  \
    hipLaunchKernelGGL((trsm_left_mfma_kernel<K, P>), dim3(grid, 1, batch), dim3(256), trsm_mfma_lds<K>(), st, dA, \
                       lda, k0, c0, c1, dinv_of(k0), ctx.status.ptr, pk, pinv, pcb, bu)
            if (w == 128 && pre)
                PEEC_TRSM_L(8, true);
            else if (w == 128)
                PEEC_TRSM_L(8, false);
            else if (w == 64 && pre)
                PEEC_TRSM_L(4, true);
            else if (w == 64)
                PEEC_TRSM_L(4, false);
            else if (!pk)
                PEEC_TRSM_L(2, false);
            else
                throw HipError{"ordered factor: packed U12 solves need a block width of 64 or 128"};
#undef PEEC_TRSM_L
            return;
        }
        if (pre) throw HipError{"ordered factor: two-panel U12 packs need the blocked solve kernels"};
        if (batch > 1) throw HipError{"ordered factor: a batch needs the blocked solve kernels"};
        const int tiles = (c1 - c0 + kTrsmWaveCols - 1) / kTrsmWaveCols;
        const int grid = std::max(1, std::min(tiles, ctx.cus));
        hipLaunchKernelGGL(trsm_wave_kernel<S>, dim3((unsigned)grid), dim3(1024), trsm_wave_lds(w) * sizeof(S), st,
                           dA, lda, c0, c1, k0, w, ctx.status.ptr, pk, pinv, pcb);
    };
    if (blocked_solves && nb == 128)
        set_dynamic_lds(reinterpret_cast<const void *>(diag_lu_blocked_kernel<8>), diag_blocked_lds<8>());
    auto diag = [&](int k0, int w, hipStream_t st) {
        const int tp = ctx.timer ? ctx.timer->start(st) : -1;
        if (blocked_solves && w == nb && (w == 128 || w == 64)) {
            if (w == 128)
                hipLaunchKernelGGL(diag_lu_blocked_kernel<8>, dim3(1, 1, batch), dim3(512), diag_blocked_lds<8>(), st,
                                   dA, lda, k0, pivot_limit, ctx.status.ptr, dinv_of(k0), bs);
            else
                hipLaunchKernelGGL(diag_lu_blocked_kernel<4>, dim3(1, 1, batch), dim3(512), diag_blocked_lds<4>(), st,
                                   dA, lda, k0, pivot_limit, ctx.status.ptr, dinv_of(k0), bs);
        } else {
            hipLaunchKernelGGL(diag_lu_tiled_kernel, dim3(1, 1, batch), dim3(1024), 0, st, dA, lda, k0, w, pivot_limit,
                               ctx.status.ptr, blocked_solves && w == nb ? dinv_of(k0) : nullptr, bs);
        }
        if (ctx.timer) ctx.timer->stop(tp, st, kKernelPanel, 0.0, 2.0 * (double)w * w * sizeof(S));
    };
    diag(0, std::min(nb, n), s);
    bool pending = false;    // a diagonal factor is in flight on aux
    bool pair_open = false;  // the previous panel began a pair of split-FP16 panels
    int wprev = 0;
    // Deferred trailing updates: the second panel of pair A (panels P0 P1)
    // updates, of its trailing matrix, only the rows of the next pair (P2 P3;
    // A1) and the columns of P3 and P4 below them (A2); the rest (rows below
    // P3, columns beyond P4) is applied with pair B's own trailing update in
    // one pass over C (two segments), from pair A's packs, kept in the other
    // pack set. Pairs alternate between the two sets.
    int pair_index = 0;
    bool deferred = false;
    int def_set = 0, def_rb_pad = 0, def_cb = 0;
    // split-FP16 operand buffers at their largest (the first panels), so
    // that no buffer moves while the other stream may use it
    const bool split_panels = split16 && blocked_solves && (nb == 128 || nb == 64) && n > nb;
    if (batch > 1 && n > nb && !split_panels)
        throw HipError{"ordered factor: a batch needs the split-FP16 update with a block width of 64 or 128"};
    if (split_panels) {
        int rbp, cbp;
        gemm_pads<S>(n, n, rbp, cbp);
        const int kcm = 2 * nb / 16, cbm = (n + 127) / 128 * 8;
        // per member (in uint4 units / floats)
        const size_t pl = (size_t)kcm * rbp * 128, pun = (size_t)kcm * 8 * 128, pu = (size_t)kcm * cbm * 128;
        ctx.pack_l16.reserve(pl * batch * sizeof(uint4));
        ctx.scale_l16.reserve((size_t)rbp * 16 * batch);
        ctx.pack_u16n.reserve(pun * batch * sizeof(uint4));
        ctx.scale_u16n.reserve((size_t)8 * 16 * batch);
        ctx.pack_u16.reserve(pu * batch * sizeof(uint4));
        ctx.scale_u16.reserve((size_t)cbm * 16 * batch);
        if (ctx.defer_updates) {
            ctx.pack_l16b.reserve(pl * batch * sizeof(uint4));
            ctx.scale_l16b.reserve((size_t)rbp * 16 * batch);
            ctx.pack_u16b.reserve(pu * batch * sizeof(uint4));
            ctx.scale_u16b.reserve((size_t)cbm * 16 * batch);
        }
        if (batch > 1) {
            bs_n.pl = bs_r.pl = pl;
            bs_n.sl = bs_r.sl = (size_t)rbp * 16;
            bs_n.pu = pun;
            bs_n.su = (size_t)8 * 16;
            bs_r.pu = pu;
            bs_r.su = (size_t)cbm * 16;
        }
    }
    for (int k0 = 0; k0 < n; k0 += nb) {
        const int w = std::min(nb, n - k0), rest = n - k0 - w;
        if (rest <= 0) {
            if (pending) HIP_CHECK(hipStreamWaitEvent(s, ctx.ev_diag, 0));
            pending = false;
            break;
        }
        const int w2 = std::min(nb, rest), r0 = k0 + w;
        int rb_pad, cb_pad;
        gemm_pads<S>(rest, rest, rb_pad, cb_pad);
        const int rest2 = rest - w2;
        // split-FP16 pair state of this panel (the second panel of a pair
        // packs both panels' L21 / U12 with one scale per row / column)
        const bool h3 = split_panels && w == nb;
        const bool second = h3 && pair_open && wprev == w;
        const int cbn = (w2 + 127) / 128 * 8, cb16 = (rest2 + 127) / 128 * 8;
        const int set = ctx.defer_updates ? (pair_index & 1) : 0;
        uint4 *pln = reinterpret_cast<uint4 *>(set ? ctx.pack_l16b.ptr : ctx.pack_l16.ptr);
        float *sln = set ? ctx.scale_l16b.ptr : ctx.scale_l16.ptr;
        uint4 *pun = reinterpret_cast<uint4 *>(ctx.pack_u16n.ptr);
        uint4 *pu = reinterpret_cast<uint4 *>(set ? ctx.pack_u16b.ptr : ctx.pack_u16.ptr);
        float *su = set ? ctx.scale_u16b.ptr : ctx.scale_u16.ptr;
        // U12 of the next block column needs this block's diagonal factor
        // and the previous trailing update; with a look-ahead factor in
        // flight it runs on the second stream behind both, beside L21
        const bool u12_aux = pending;
        if (pending) {
            HIP_CHECK(hipEventRecord(ctx.ev_prev, s));
            HIP_CHECK(hipStreamWaitEvent(aux, ctx.ev_prev, 0));
            if (h3)
                solve_u12(k0, w, r0, r0 + w2, aux, pun, ctx.scale_u16n.ptr, cbn, second);
            else
                solve_u12(k0, w, r0, r0 + w2, aux);
            HIP_CHECK(hipEventRecord(ctx.ev_u12, aux));
            HIP_CHECK(hipStreamWaitEvent(s, ctx.ev_diag, 0));
            pending = false;
        }
        const int ks_count = (w + 3) / 4;
        if (h3) {
            // Split-FP16 updates, in pairs of block columns (the first panel of
            // the matrix starts a pair). The first panel of a pair updates the
            // next block column (look-ahead) and, of the rest of the matrix,
            // only the next panel's w2 rows (the thin update: they become that
            // panel's U12); the second applies both panels at once, rank 2w,
            // to its next block column and to the trailing matrix, so the
            // trailing matrix is read and written once per pair. L21 and U12
            // stay in A between the two panels and are packed together with
            // one scale per row / column over all 2w of k.
            // The solves write the operands of the updates as they go (the
            // split-FP16 packs of L21 rows and U12 columns; pad rows and
            // columns of the packs only feed results that are never stored).
            const int kw = second ? 2 * w : w;
            auto start_next_diag = [&]() {
                HIP_CHECK(hipEventRecord(ctx.ev_next, s));
                HIP_CHECK(hipStreamWaitEvent(aux, ctx.ev_next, 0));
                diag(r0, w2, aux);
                HIP_CHECK(hipEventRecord(ctx.ev_diag, aux));
                pending = true;
            };
            if (!second && w2 % 16 == 0) {
                // first panel of a pair (nothing long follows it on this
                // stream): the next diagonal block first (its rows of L21,
                // U12 of its columns, the w2 x w2 update), so that its factor
                // on the second stream overlaps the rest of L21 and of the
                // next block column's update. The split falls on a 16-row
                // block of the L21 pack (a last, narrower next panel takes
                // the path below).
                const int rbd = w2 / 16;
                solve_l21(k0, w, rest, 0, w2, pln, sln, rb_pad, false);
                if (!u12_aux) solve_u12(k0, w, r0, r0 + w2, s, pun, ctx.scale_u16n.ptr, cbn, false);
                if (u12_aux) HIP_CHECK(hipStreamWaitEvent(s, ctx.ev_u12, 0));
                gemm_h3(dA, lda, r0, r0, w2, w2, kw, rb_pad, cbn, pun, ctx.scale_u16n.ptr, kKernelLuUpdate,
                        ctx.status.ptr, ctx, s, 0, bs_n, batch, set);
                start_next_diag();
                if (rest > w2) {
                    solve_l21(k0, w, rest, w2, rest, pln, sln, rb_pad, false);
                    gemm_h3(dA, lda, r0 + w2, r0, rest - w2, w2, kw, rb_pad, cbn, pun, ctx.scale_u16n.ptr,
                            kKernelLuUpdate, ctx.status.ptr, ctx, s, rbd, bs_n, batch, set);
                }
            } else {
                // L21 (rows r0..) and U12 of the next block column, then its
                // update (rank 2w for a pair's second panel; critical path)
                solve_l21(k0, w, rest, 0, rest, pln, sln, rb_pad, second);
                if (!u12_aux) solve_u12(k0, w, r0, r0 + w2, s, pun, ctx.scale_u16n.ptr, cbn, second);
                if (u12_aux) HIP_CHECK(hipStreamWaitEvent(s, ctx.ev_u12, 0));
                gemm_h3(dA, lda, r0, r0, rest, w2, kw, rb_pad, cbn, pun, ctx.scale_u16n.ptr, kKernelLuUpdate,
                        ctx.status.ptr, ctx, s, 0, bs_n, batch, set);
                start_next_diag();
            }
            if (rest2 > 0) {
                if (!second) {
                    // the rest of U12, packed by the solve itself (blocks past
                    // the last column are never read into a stored result);
                    // then only the next panel's rows
                    solve_u12(k0, w, r0 + w2, n, s, pu, su, cb16);
                    gemm_h3(dA, lda, r0, r0 + w2, w2, rest2, w, rb_pad, cb16, pu, su, kKernelLuUpdateThin,
                            ctx.status.ptr, ctx, s, 0, bs_r, batch, set);
                } else {
                    // the rest of this panel's U12, packed with the first
                    // panel's U12 rows, and the rank-2w trailing update:
                    // deferred in part (pair A), merged with the deferred
                    // pair's (pair B), or whole
                    solve_u12(k0, w, r0 + w2, n, s, pu, su, cb16, true);
                    const int span = 2 * nb;  // the next pair's rows / columns
                    if (!deferred && ctx.defer_updates && w2 == nb && rest2 >= span + 16) {
                        // A1: the next pair's rows, every column beyond the next panel
                        gemm_h3(dA, lda, r0, r0 + w2, span, rest2, kw, rb_pad, cb16, pu, su, kKernelLuUpdateThin,
                                ctx.status.ptr, ctx, s, 0, bs_r, batch, set);
                        // A2: the rows below it, columns of its second panel and the one after
                        gemm_h3(dA, lda, r0 + span, r0 + w2, rest - span, span, kw, rb_pad, cb16, pu, su,
                                kKernelLuUpdateThin, ctx.status.ptr, ctx, s, span / 16, bs_r, batch, set);
                        deferred = true;
                        def_set = set;
                        def_rb_pad = rb_pad;
                        def_cb = cb16;
                    } else if (deferred) {
                        // pair A's packs start a span above / to the left of this update
                        const int off = span / 16;
                        const uint4 *pua = reinterpret_cast<const uint4 *>(def_set ? ctx.pack_u16b.ptr
                                                                                   : ctx.pack_u16.ptr);
                        const float *sua = def_set ? ctx.scale_u16b.ptr : ctx.scale_u16.ptr;
                        const H3Seg ga = h3_seg(ctx, def_set, off, def_rb_pad, def_cb, pua + (size_t)off * 128,
                                                sua + (size_t)off * 16);
                        const H3Seg gb = h3_seg(ctx, set, 0, rb_pad, cb16, pu, su);
                        gemm_h3_segs(dA, lda, r0, r0 + w2, rest, rest2, kw, ga, &gb, kKernelLuUpdateMerged,
                                     ctx.status.ptr, ctx, s, bs_r, batch);
                        deferred = false;
                    } else {
                        gemm_h3(dA, lda, r0, r0 + w2, rest, rest2, kw, rb_pad, cb16, pu, su, kKernelLuUpdateSplit16,
                                ctx.status.ptr, ctx, s, 0, bs_r, batch, set);
                    }
                }
            }
            if (second) ++pair_index;
            pair_open = !second;
            wprev = w;
            HIP_CHECK(hipGetLastError());
            continue;
        }
        // a pair's second panel always takes the split-FP16 branch (every
        // panel but the last has the full width, and the last has no rows
        // below it): the first panel's thin update left the rows below the
        // pair without its contribution
        if (pair_open || deferred)
            throw HipError{"ordered factor: a split-FP16 panel pair was left unfinished"};
        ctx.pack_l.reserve((size_t)(ks_count + 2) * rb_pad * 64 * sizeof(S));
        ctx.pack_u.reserve((size_t)(ks_count + 2) * cb_pad * 64 * sizeof(S));
        S *Lp = reinterpret_cast<S *>(ctx.pack_l.ptr);
        S *Up = reinterpret_cast<S *>(ctx.pack_u.ptr);
        // critical path: L21, U12 of the next block column, its update
        solve_l21(k0, w, rest);
        if (!u12_aux) solve_u12(k0, w, r0, r0 + w2, s);
        const size_t tl = (size_t)ks_count * rb_pad * 64;
        hipLaunchKernelGGL(pack_l_kernel<S>, dim3((unsigned)((tl + 255) / 256)), dim3(256), 0, s, dA, lda, r0, rest,
                           k0, w, rb_pad, Lp, ctx.status.ptr);
        if (u12_aux) HIP_CHECK(hipStreamWaitEvent(s, ctx.ev_u12, 0));
        const int cb_next = 4 * ((w2 + 63) / 64);
        const size_t tu1 = (size_t)ks_count * cb_next * 64;
        hipLaunchKernelGGL(pack_u_kernel<S>, dim3((unsigned)((tu1 + 255) / 256)), dim3(256), 0, s, dA, lda, k0, w,
                           r0, w2, cb_next, cb_pad, Up, ctx.status.ptr);
        {
            const int tiles = ((rest + 63) / 64) * ((w2 + 63) / 64);
            const int tg = ctx.timer ? ctx.timer->start(s) : -1;
            hipLaunchKernelGGL(gemm_splitk_kernel<S>, dim3((unsigned)tiles), dim3(256), split_lds, s, dA, lda, r0, r0,
                               rest, w2, ks_count, Lp, Up, rb_pad, cb_pad, ctx.status.ptr);
            if (ctx.timer)
                ctx.timer->stop(tg, s, kKernelLuUpdate, 8.0 * rest * w2 * w,
                                (2.0 * rest * w2 + (double)(rest + w2) * w) * sizeof(S));
        }
        HIP_CHECK(hipEventRecord(ctx.ev_next, s));
        HIP_CHECK(hipStreamWaitEvent(aux, ctx.ev_next, 0));
        diag(r0, w2, aux);
        HIP_CHECK(hipEventRecord(ctx.ev_diag, aux));
        pending = true;
        // off the critical path: the rest of U12 and of the trailing update
        if (rest2 > 0) {
            solve_u12(k0, w, r0 + w2, n, s);
            const int cb_rest = cb_pad - w2 / 16;
            const size_t tu2 = (size_t)ks_count * cb_rest * 64;
            hipLaunchKernelGGL(pack_u_kernel<S>, dim3((unsigned)((tu2 + 255) / 256)), dim3(256), 0, s, dA, lda, k0,
                               w, r0 + w2, rest2, cb_rest, cb_pad, Up + (size_t)(w2 / 16) * 64, ctx.status.ptr);
            gemm_packed<S>(dA, lda, r0, r0, rest, w2, rest2, w, rb_pad, cb_pad, ctx.status.ptr, ctx, s, true);
        }
        HIP_CHECK(hipGetLastError());
    }
    if (pending) HIP_CHECK(hipStreamWaitEvent(s, ctx.ev_diag, 0));
    if (deferred) throw HipError{"ordered factor: a deferred trailing update was not applied"};
    std::vector<int> st((size_t)2 * batch, 0);
    ctx.status.download(st.data(), st.size(), s);
    HIP_CHECK(hipStreamSynchronize(s));
    for (int b = 0; b < batch; ++b) {
        LUStatus m;
        m.code = st[2 * b];
        m.column = st[2 * b + 1];
        if (member_status) member_status[b] = m;
        if (out.code == 0 && m.code != 0) out = m;
    }
    return out;
}

// ---------------------------------------------------------------------------
// solves

template <class V> struct VecT;
template <> struct VecT<double2> { using A = double2; };
template <> struct VecT<double> { using A = double; };

// y[r][i] = b[r][perm[i]] / row[perm[i]]
template <class V>
__global__ void permute_scale_kernel(V *X, const V *B, const int *perm, const double *row, int n, int nrhs) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)n * nrhs) return;
    const int r = (int)(idx / n), i = (int)(idx % n);
    const int p = perm[i];
    V v = B[(size_t)r * n + p];
    if (row) {
        if constexpr (Elem<V>::cplx) {
            v.x = v.x / row[p];
            v.y = v.y / row[p];
        } else {
            v = v / row[p];
        }
    }
    X[idx] = v;
}

template <class V>
__global__ void col_scale_kernel(V *X, const double *col, int n, int nrhs) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)n * nrhs || !col) return;
    const int i = (int)(idx % n);
    if constexpr (Elem<V>::cplx) {
        X[idx].x /= col[i];
        X[idx].y /= col[i];
    } else {
        X[idx] /= col[i];
    }
}

// X[r][i] -= sum_{j in [j0,j1)} M[i][j] X[r][j] for rows i in [i0, i0+T)
template <class S, class V>
__global__ __launch_bounds__(256) void trsv_gemv_kernel(const S *M, size_t lda, V *X, int n, int i0, int rows,
                                                        int j0, int j1) {
    using T = typename VecT<V>::A;
    __shared__ T red[256];
    const int i = i0 + blockIdx.x, r = blockIdx.y;
    if (blockIdx.x >= rows) return;
    T acc = ezero<T>();
    const S *mrow = M + (size_t)i * lda;
    const V *x = X + (size_t)r * n;
    for (int j = j0 + threadIdx.x; j < j1; j += 256) acc = eadd(acc, emul(ld(mrow + j), x[j]));
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if (threadIdx.x < s2) red[threadIdx.x] = eadd(red[threadIdx.x], red[threadIdx.x + s2]);
        __syncthreads();
    }
    if (threadIdx.x == 0) X[(size_t)r * n + i] = esub(X[(size_t)r * n + i], red[0]);
}

// diagonal block [i0, i0+rows): unit-lower forward (upper=0) or upper
// backward (upper=1) substitution, one wavefront per right-hand side
template <class S, class V>
__global__ __launch_bounds__(64) void trsv_diag_kernel(const S *M, size_t lda, V *X, int n, int i0, int rows,
                                                       int upper, int *bad) {
    using T = typename VecT<V>::A;
    const int r = blockIdx.x, lane = threadIdx.x;
    V *x = X + (size_t)r * n;
    for (int step = 0; step < rows; ++step) {
        const int i = upper ? i0 + rows - 1 - step : i0 + step;
        T part = ezero<T>();
        const int lo = upper ? i + 1 : i0, hi = upper ? i0 + rows : i;
        for (int j = lo + lane; j < hi; j += 64) part = eadd(part, emul(ld(M + (size_t)i * lda + j), x[j]));
        for (int m = 32; m > 0; m >>= 1) part = eadd(part, eshfl_xor(part, m));
        if (lane == 0) {
            T v = esub(x[i], part);
            if (upper) {
                const T d = ld(M + (size_t)i * lda + i);
                if constexpr (Elem<V>::cplx) {
                    const double den = d.x * d.x + d.y * d.y;
                    if (!isfinite(den) || den == 0.0) atomicOr(bad, 1);
                } else {
                    if (d == 0.0 || !isfinite(d)) atomicOr(bad, 1);
                }
                v = ediv(v, d);
            }
            x[i] = v;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __syncthreads();
    }
}

constexpr int kChainTB = 64;
// LDS layout of a 64-wide row in the chain kernel: column c at
// c + 4 (c / 16) in rows of 77, so that the 16 lanes of one LDS access group
// (4 rows x the 4 column quarters of a thread quad) fall in distinct banks
constexpr int kChainRow = 77;
__host__ __device__ constexpr int chain_col(int c) { return c + 4 * (c >> 4); }

// sum over the 4 lanes of a quad (lanes 4j .. 4j + 3), every lane ends
// with it: (v0 + v1) + (v2 + v3) by DPP quad permutations, no LDS traffic
template <int CTRL>
__device__ inline double dpp_quad(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ inline double quad_sum(double v) {
    v = v + dpp_quad<0xB1>(v);  // lanes 1, 0, 3, 2
    return v + dpp_quad<0x4E>(v);  // lanes 2, 3, 0, 1
}
__device__ inline double2 quad_sum(double2 v) { return make_double2(quad_sum(v.x), quad_sum(v.y)); }

// Inverses of the 64 x 64 diagonal blocks of a factor, in FP64: the
// unit-lower L part to linv, then the upper U part to uinv, both
// identity-padded past row n. One workgroup per block; thread t owns row
// t / 4 and columns 16 (t % 4) .. + 16 of X in registers, and the block is
// inverted by substitution against the identity (L X = I forward, U X = I
// backward): at each step the finished row goes through LDS to the rows
// still to update (one barrier per step, the row buffer double-buffered).
// A zero or non-finite diagonal of U sets bad |= 1.
template <class S, class V>
__global__ __launch_bounds__(256) void diag_inverse_kernel(const S *M, size_t lda, int n,
                                                           typename VecT<V>::A *linv, typename VecT<V>::A *uinv,
                                                           int *bad, size_t m_stride = 0, size_t d_stride = 0) {
    using T = typename VecT<V>::A;
    // member blockIdx.y of a batch of factors (m_stride apart), its inverses
    // d_stride apart
    M += (size_t)blockIdx.y * m_stride;
    linv += (size_t)blockIdx.y * d_stride;
    uinv += (size_t)blockIdx.y * d_stride;
    constexpr int TB = kChainTB;
    __shared__ T blk[TB][TB];
    __shared__ T xrow[2][TB];  // the finished row of X at each step (double-buffered)
    const int b = blockIdx.x, r0 = b * TB, rows = min(TB, n - r0);
    T one = ezero<T>();
    if constexpr (Elem<V>::cplx) one.x = 1.0; else one = 1.0;
    const int tid = threadIdx.x, i = tid >> 2, c0 = (tid & 3) * 16;
    for (int idx = tid; idx < TB * TB; idx += 256) {
        const int r = idx / TB, c = idx % TB;
        blk[r][c] = (r < rows && c < rows) ? ld(M + (size_t)(r0 + r) * lda + r0 + c) : (r == c ? one : ezero<T>());
    }
    // this thread's 16 entries of row i of X, in registers
    Regs<T, 16> x;
#pragma unroll
    for (int c = 0; c < 16; ++c) x.set(c, (i == c0 + c) ? one : ezero<T>());
    __syncthreads();
    // L X = I: row k is final at step k; rows below subtract L[i][k] X[k][:]
    for (int k = 0; k < TB - 1; ++k) {
        T *xk = xrow[k & 1];
        if (i == k) {
#pragma unroll
            for (int c = 0; c < 16; ++c) xk[c0 + c] = x.get(c);
        }
        __syncthreads();
        if (i > k) {
            const T l = blk[i][k];
#pragma unroll
            for (int c = 0; c < 16; ++c) x.set(c, esub(x.get(c), emul(l, xk[c0 + c])));
        }
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) linv[((size_t)b * TB + i) * TB + c0 + c] = x.get(c);
#pragma unroll
    for (int c = 0; c < 16; ++c) x.set(c, (i == c0 + c) ? one : ezero<T>());
    if (tid < TB) {
        const T d = blk[tid][tid];
        if constexpr (Elem<V>::cplx) {
            const double den = d.x * d.x + d.y * d.y;
            if (!isfinite(den) || den == 0.0) atomicOr(bad, 1);
        } else {
            if (d == 0.0 || !isfinite(d)) atomicOr(bad, 1);
        }
    }
    // U X = I backward: row k is scaled by 1 / U[k][k] and is then final;
    // rows above subtract U[i][k] X[k][:]
    for (int k = TB - 1; k >= 0; --k) {
        T *xk = xrow[k & 1];
        if (i == k) {
            const T r = ediv(one, blk[k][k]);
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                x.set(c, emul(r, x.get(c)));
                xk[c0 + c] = x.get(c);
            }
        }
        __syncthreads();
        if (i < k) {
            const T u = blk[i][k];
#pragma unroll
            for (int c = 0; c < 16; ++c) x.set(c, esub(x.get(c), emul(u, xk[c0 + c])));
        }
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) uinv[((size_t)b * TB + i) * TB + c0 + c] = x.get(c);
}

// Marker of a solved value not yet written: a signalling NaN, which no
// arithmetic result can be (operations on NaNs return quiet NaNs).
constexpr unsigned long long kChainEmpty = 0x7FF4DEADBEEF0001ull;

__global__ void fill_u64_kernel(unsigned long long *p, size_t n, unsigned long long v) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

__device__ inline bool chain_empty(const double2 &v) {
    return (unsigned long long)__double_as_longlong(v.x) == kChainEmpty ||
           (unsigned long long)__double_as_longlong(v.y) == kChainEmpty;
}
__device__ inline bool chain_empty(double v) { return (unsigned long long)__double_as_longlong(v) == kChainEmpty; }

// Triangular solve in one cooperative launch over 64-row blocks, processed
// in solve order (b = w forward, b = nblk-1-w backward, w = 0, 1, ...);
// workgroup g takes w = g, g + G, g + 2G, ... (G = grid size, at most what
// fits on the device at once). Every block depends only on blocks of lower
// w, each owned by a workgroup that reaches it after its own lower blocks,
// so the chain always completes. The right-hand side comes from B; the
// solution goes to Y, which holds kChainEmpty everywhere when the launch
// starts. For its block, a workgroup accumulates the off-diagonal products
// of the blocks it depends on in order, each as soon as that block's values
// in Y are no longer empty (the values are their own ready flags: one memory
// round trip per dependency, no fences), then applies the precomputed
// inverse of its diagonal block (one 64 x 64 product instead of 64 dependent
// substitution steps) and writes its rows to Y with cache-bypassing stores.
// Thread t works on row t / 4 and the 16 columns 16 (t % 4) .. + 16 of each
// block, so the four threads of a row read 64 consecutive elements; the
// factor entries of a dependency are loaded before its values are polled.
// Every workgroup waits at the front of the chain, so the time per block is
// the poll round trip plus what follows it: the right-hand sides are a
// template parameter (NQ = 1, or up to 8) so that those steps are straight
// code whose LDS reads are issued together, the row sums use DPP, and the
// LDS rows are skewed against bank conflicts. All arithmetic in FP64 (FP32
// factors are widened on load).
template <class S, class V, bool UPPER, int NQ>
__global__ __launch_bounds__(256) void trsv_chain_kernel(const S *M, size_t lda, const V *B, V *Y, int n, int nrhs,
                                                         const typename VecT<V>::A *dinv, int *bad,
                                                         size_t m_stride, size_t v_stride, size_t d_stride) {
    using T = typename VecT<V>::A;
    // member blockIdx.y of a batch: its factor, vectors and block inverses
    // (strides zero for a single solve); each member's chain has the grid's
    // x dimension to itself
    M += (size_t)blockIdx.y * m_stride;
    B += (size_t)blockIdx.y * v_stride;
    Y += (size_t)blockIdx.y * v_stride;
    dinv += (size_t)blockIdx.y * d_stride;
    constexpr int TB = kChainTB;
    __shared__ T dblk[TB * kChainRow];
    __shared__ T xs[NQ][kChainRow];
    __shared__ T bs[NQ][TB];
    __shared__ int s_abort;
    const int nblk = (n + TB - 1) / TB;
    const int nq = NQ == 1 ? 1 : nrhs;
    const int tid = threadIdx.x;
    const int i = tid >> 2, qt = tid & 3, c0 = qt * 16;
    const int cs = chain_col(c0);  // this thread's 16 columns start here in an LDS row
    if (tid == 0) s_abort = 0;
    for (int w = blockIdx.x; w < nblk; w += gridDim.x) {
        const int b = UPPER ? nblk - 1 - w : w;
        const int r0 = b * TB, rows = min(TB, n - r0);
        const int ri = min(i, rows - 1);  // clamped row for loads
        Regs<T, NQ> acc;
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc.set(q, ezero<T>());
        __syncthreads();  // the previous block is done with dblk, xs, bs
        for (int idx = tid; idx < TB * TB; idx += 256)
            dblk[(idx / TB) * kChainRow + chain_col(idx % TB)] = dinv[(size_t)b * TB * TB + idx];
        for (int idx = tid; idx < nq * TB; idx += 256) {
            const int q = idx / TB, c = idx % TB;
            bs[q][c] = c < rows ? B[(size_t)q * n + r0 + c] : ezero<T>();
        }
        const S *mrow = M + (size_t)(r0 + ri) * lda;
        for (int d = 0; d < w; ++d) {
            const int jb = UPPER ? nblk - 1 - d : d;
            const int cb = jb * TB, cols = min(TB, n - cb);
            // this thread's 16 factor entries of the dependency (clamped
            // columns past n are multiplied by zero below)
            Regs<T, 16> m;
#pragma unroll
            for (int c = 0; c < 16; ++c) m.set(c, ld(mrow + cb + min(c0 + c, cols - 1)));
            for (int idx = tid; idx < nq * TB; idx += 256) {
                const int q = idx / TB, c = idx % TB;
                T xv = ezero<T>();
                if (c < cols) {
                    const double *p = reinterpret_cast<const double *>(Y + (size_t)q * n + cb + c);
                    unsigned spins = 0;
                    for (;;) {
                        sub_elem(p, xv);
                        if (!chain_empty(xv)) break;
                        if (++spins > kSpinLimit) {
                            atomicOr(bad, 2);
                            s_abort = 1;
                            xv = ezero<T>();
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                xs[q][chain_col(c)] = xv;
            }
            __syncthreads();
            if (s_abort) return;
#pragma unroll
            for (int q = 0; q < NQ; ++q)
                if (NQ == 1 || q < nq) {
                    Regs<T, 16> xr;
#pragma unroll
                    for (int c = 0; c < 16; ++c) xr.set(c, xs[q][cs + c]);
                    T a = acc.get(q);
#pragma unroll
                    for (int c = 0; c < 16; ++c) a = eadd(a, emul(m.get(c), xr.get(c)));
                    acc.set(q, a);
                }
            __syncthreads();
        }
        __syncthreads();  // bs and dblk of a block without dependencies
        // row sums over the four threads of a row; z = b - sum into xs
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            if (NQ == 1 || q < nq) {
                const T v = quad_sum(acc.get(q));
                if (qt == 0) xs[q][chain_col(i)] = esub(bs[q][i], v);
            }
        __syncthreads();
        // x_b = D^{-1} z, published with cache-bypassing stores
        Regs<T, 16> drow;
#pragma unroll
        for (int k = 0; k < 16; ++k) drow.set(k, dblk[i * kChainRow + cs + k]);
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            if (NQ == 1 || q < nq) {
                Regs<T, 16> zr;
#pragma unroll
                for (int k = 0; k < 16; ++k) zr.set(k, xs[q][cs + k]);
                T v = ezero<T>();
#pragma unroll
                for (int k = 0; k < 16; ++k) v = eadd(v, emul(drow.get(k), zr.get(k)));
                v = quad_sum(v);
                if (qt == (q & 3) && i < rows) pub_elem(reinterpret_cast<double *>(Y + (size_t)q * n + r0 + i), v);
            }
    }
}

// The one-right-hand-side chain with the same order of work per block as
// trsv_chain_kernel<S, V, UPPER, 1>, restructured for the workgroups that
// run behind the front of the chain (all but the last few dependencies of
// their block are already solved when they reach them): up to K
// dependencies are taken per round trip (their factor entries and polled
// values loaded together, then summed in dependency order), and the block
// inverse is staged as its packed triangle (L_bb^-1 is lower, U_bb^-1 upper
// triangular; the zero triangle only added zeros), which also lets two
// workgroups share a CU. Products and sums are those of the one-at-a-time
// kernel, in the same order, so the solutions are the same. Member
// blockIdx.y of a batch at the given strides.
template <class S, bool UPPER, int K>
__global__ __launch_bounds__(256) void trsv_chain1_kernel(const S *M, size_t lda, const double2 *B, double2 *Y, int n,
                                                          const double2 *dinv, int *bad, size_t m_stride,
                                                          size_t v_stride, size_t d_stride) {
    using T = double2;
    M += (size_t)blockIdx.y * m_stride;
    B += (size_t)blockIdx.y * v_stride;
    Y += (size_t)blockIdx.y * v_stride;
    dinv += (size_t)blockIdx.y * d_stride;
    constexpr int TB = kChainTB, NT = TB * (TB + 1) / 2;
    __shared__ T dtri[NT];
    __shared__ T xs[K][kChainRow];
    __shared__ T bs[TB];
    __shared__ int s_abort;
    const int nblk = (n + TB - 1) / TB;
    const int tid = threadIdx.x;
    const int i = tid >> 2, qt = tid & 3, c0 = qt * 16;
    const int cs = chain_col(c0);
    // packed position of inverse entry (r, c) of the stored triangle
    auto tri = [](int r, int c) { return UPPER ? c * (c + 1) / 2 + r : r * (r + 1) / 2 + c; };
    if (tid == 0) s_abort = 0;
    for (int w = blockIdx.x; w < nblk; w += gridDim.x) {
        const int b = UPPER ? nblk - 1 - w : w;
        const int r0 = b * TB, rows = min(TB, n - r0);
        const int ri = min(i, rows - 1);
        T acc = ezero<T>();
        __syncthreads();  // the previous block is done with dtri, xs, bs
        for (int idx = tid; idx < TB * TB; idx += 256) {
            const int r = idx / TB, c = idx % TB;
            if (UPPER ? c >= r : c <= r) dtri[tri(r, c)] = dinv[(size_t)b * TB * TB + idx];
        }
        if (tid < TB) bs[tid] = tid < rows ? B[r0 + tid] : ezero<T>();
        const S *mrow = M + (size_t)(r0 + ri) * lda;
        __builtin_amdgcn_s_setprio(0);
        for (int d0 = 0; d0 < w; d0 += K) {
            const int kk = min(K, w - d0);
            // the last round is the front of the chain: its waves go first on
            // a CU shared with workgroups still summing earlier dependencies
            if (d0 + K >= w) __builtin_amdgcn_s_setprio(3);
            // this thread's 16 factor entries of each dependency of the round
            // (clamped columns past n meet zero values below)
            S mf[K][16];
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (k < kk) {
                    const int jb = UPPER ? nblk - 1 - (d0 + k) : d0 + k;
                    const int cb = jb * TB, cols = min(TB, n - cb);
#pragma unroll
                    for (int c = 0; c < 16; ++c) mf[k][c] = mrow[cb + min(c0 + c, cols - 1)];
                }
            if (tid < TB) {
                T xv[K];
                // every dependency's value first, then wait on those not yet there
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    xv[k] = ezero<T>();
                    const int jb = UPPER ? nblk - 1 - (d0 + k) : d0 + k;
                    if (k < kk && tid < min(TB, n - jb * TB))
                        sub_elem(reinterpret_cast<const double *>(Y + (size_t)jb * TB + tid), xv[k]);
                }
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int jb = UPPER ? nblk - 1 - (d0 + k) : d0 + k;
                    if (k < kk && tid < min(TB, n - jb * TB)) {
                        const double *p = reinterpret_cast<const double *>(Y + (size_t)jb * TB + tid);
                        unsigned spins = 0;
                        while (chain_empty(xv[k])) {
                            if (++spins > kSpinLimit) {
                                atomicOr(bad, 2);
                                s_abort = 1;
                                xv[k] = ezero<T>();
                                break;
                            }
                            __builtin_amdgcn_s_sleep(1);
                            sub_elem(p, xv[k]);
                        }
                    }
                    xs[k][chain_col(tid)] = xv[k];
                }
            }
            __syncthreads();
            if (s_abort) return;
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (k < kk) {
#pragma unroll
                    for (int c = 0; c < 16; ++c) acc = eadd(acc, emul(ld(&mf[k][c]), xs[k][cs + c]));
                }
            __syncthreads();
        }
        __syncthreads();  // bs and dtri of a block without dependencies
        {
            const T v = quad_sum(acc);
            if (qt == 0) xs[0][chain_col(i)] = esub(bs[i], v);
        }
        __syncthreads();
        T v = ezero<T>();
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int c = c0 + k;
            if (UPPER ? c >= i : c <= i) v = eadd(v, emul(dtri[tri(i, c)], xs[0][cs + k]));
        }
        v = quad_sum(v);
        if (qt == 0 && i < rows) pub_elem(reinterpret_cast<double *>(Y + r0 + i), v);
    }
}

// dependencies per round trip of trsv_chain1_kernel: two for FP32 factors
// (fastest measured of 2 / 3 / 4: three workgroups fit a CU), one for FP64
template <class S> constexpr int chain_deps() { return sizeof(S) == sizeof(float2) ? 2 : 1; }
constexpr int kChainDeps = chain_deps<float2>();  // the FP32 factors of a batch

// number of chain workgroups that can be resident at once (the chain
// kernel's workgroups wait on each other)
template <class S, class V>
int chain_grid(int n, const LUContext &ctx) {
    const int nblk = (n + kChainTB - 1) / kChainTB;
    // the eight-right-hand-side variants use the most LDS
    int f = 0, bw = 0;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &f, reinterpret_cast<const void *>(trsv_chain_kernel<S, V, false, 8>), 256, 0));
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &bw, reinterpret_cast<const void *>(trsv_chain_kernel<S, V, true, 8>), 256, 0));
    long long g = std::min<long long>(nblk, (long long)std::min(f, bw) * ctx.cus);
    // PEEC_CHAIN_WORKGROUPS caps the grid (tests use it to run several
    // blocks per workgroup on small systems)
    if (const char *cap = std::getenv("PEEC_CHAIN_WORKGROUPS")) {
        const long long c = std::atoll(cap);
        if (c > 0) g = std::min(g, c);
    }
    return (int)g;
}

// one triangular solve through the chain kernel: B -> Y (Y is first filled
// with kChainEmpty)
template <class S, class V, bool UPPER>
void chain_solve(const S *dLU, int n, size_t lda, const V *B, V *Y, int nrhs, const typename VecT<V>::A *dinv,
                 int *bad, int grid, const LUContext &ctx) {
    const size_t words = (size_t)n * nrhs * sizeof(V) / sizeof(double);
    hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, ctx.stream,
                       reinterpret_cast<unsigned long long *>(Y), words, kChainEmpty);
    for (int q0 = 0; q0 < nrhs; q0 += 8) {
        const V *Bq = B + (size_t)q0 * n;
        V *Yq = Y + (size_t)q0 * n;
        int nq = std::min(8, nrhs - q0), nn = n;
        size_t ld = lda, zero = 0;
        void *args[] = {(void *)&dLU, &ld, (void *)&Bq, &Yq, &nn, &nq, (void *)&dinv, &bad, &zero, &zero, &zero};
        if constexpr (std::is_same<V, double2>::value) {
            if (nq == 1) {
                // one right-hand side: the dependency-batched chain
                void *args1[] = {(void *)&dLU, &ld, (void *)&Bq, &Yq, &nn, (void *)&dinv, &bad, &zero, &zero, &zero};
                HIP_CHECK(hipLaunchCooperativeKernel(
                    reinterpret_cast<const void *>(trsv_chain1_kernel<S, UPPER, chain_deps<S>()>), dim3(grid),
                    dim3(256), args1, 0, ctx.stream));
                continue;
            }
        }
        void *fn = nq == 1 ? reinterpret_cast<void *>(trsv_chain_kernel<S, V, UPPER, 1>)
                           : reinterpret_cast<void *>(trsv_chain_kernel<S, V, UPPER, 8>);
        HIP_CHECK(hipLaunchCooperativeKernel(fn, dim3(grid), dim3(256), args, 0, ctx.stream));
    }
}

// ---------------------------------------------------------------------------
// many right-hand sides: blocked triangular solves whose off-diagonal work
// is the MFMA trailing-update GEMM
//
// With X the n x nrhs right-hand sides (rhs-major), the forward solve takes
// the 64-row blocks in order: X_j <- L_jj^{-1} X_j (the diagonal-block
// inverse), then X_i -= L_ij X_j for every block i below. Transposed, that
// update is C (nrhs x rows below) -= X_j^T (nrhs x 64) L_.j^T (64 x rows
// below): the packed-operand GEMM with X_j^T as its left operand and the
// block column of L as its right one (element (k, i) of L_.j^T is the
// element (i, k) of L_.j, so the left-operand packing of L_.j is already the
// right-operand layout). The backward solve is the same with U from the
// bottom block up. The arithmetic is in the factor's precision.
constexpr int kBlockedSolveMinRhs = 16;

// X_j <- D_j X_j for the 64-row block j0 of each right-hand side (one
// workgroup per right-hand side); D_j is the block's FP64 inverse
template <class S, class T>
__global__ __launch_bounds__(64) void block_apply_kernel(S *X, int n, int j0, int rows, const T *D) {
    __shared__ T xs[kChainTB];
    const int q = blockIdx.x, i = threadIdx.x;
    S *x = X + (size_t)q * n + j0;
    xs[i] = i < rows ? ld(x + i) : ezero<T>();
    __syncthreads();
    if (i >= rows) return;
    T y = ezero<T>();
    for (int k = 0; k < rows; ++k) y = eadd(y, emul(D[i * kChainTB + k], xs[k]));
    if constexpr (Elem<S>::cplx) {
        S o;
        o.x = y.x;
        o.y = y.y;
        x[i] = o;
    } else {
        x[i] = y;
    }
}

__global__ void narrow_kernel(const double2 *src, float2 *dst, size_t count) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < count) dst[i] = make_float2((float)src[i].x, (float)src[i].y);
}
__global__ void widen_kernel(const float2 *src, double2 *dst, size_t count) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < count) dst[i] = make_double2(src[i].x, src[i].y);
}

template <class S, class T>
void blocked_solve(const S *dLU, int n, size_t lda, S *X, int nrhs, const T *linv, const T *uinv, LUContext &ctx) {
    hipStream_t s = ctx.stream;
    const int TB = kChainTB, nblk = (n + TB - 1) / TB;
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, s, ctx.status.ptr, 2, 0);
    // C (nrhs x nc at column c0 of X^T) -= X_j^T (k0 = j0, kw rows) x (rows
    // [r0, r0 + nc) of the factor, columns [j0, j0 + kw))^T
    auto update = [&](int j0, int kw, int r0, int c0, int nc) {
        if (nc <= 0) return;
        int rb_pad, cb_pad;
        gemm_pads<S>(nrhs, nc, rb_pad, cb_pad);
        const int ks = (kw + 3) / 4;
        ctx.pack_l.reserve((size_t)(ks + 2) * rb_pad * 64 * sizeof(S));
        ctx.pack_u.reserve((size_t)(ks + 2) * cb_pad * 64 * sizeof(S));
        S *Lp = reinterpret_cast<S *>(ctx.pack_l.ptr);
        S *Up = reinterpret_cast<S *>(ctx.pack_u.ptr);
        const size_t tl = (size_t)ks * rb_pad * 64, tu = (size_t)ks * cb_pad * 64;
        hipLaunchKernelGGL(pack_l_kernel<S>, dim3((unsigned)((tl + 255) / 256)), dim3(256), 0, s, X, (size_t)n, 0,
                           nrhs, j0, kw, rb_pad, Lp, ctx.status.ptr);
        hipLaunchKernelGGL(pack_l_kernel<S>, dim3((unsigned)((tu + 255) / 256)), dim3(256), 0, s, dLU, lda, r0, nc,
                           j0, kw, cb_pad, Up, ctx.status.ptr);
        gemm_packed<S>(X, (size_t)n, 0, c0, nrhs, 0, nc, kw, rb_pad, cb_pad, ctx.status.ptr, ctx, s);
    };
    for (int b = 0; b < nblk; ++b) {
        const int j0 = b * TB, rows = std::min(TB, n - j0);
        hipLaunchKernelGGL((block_apply_kernel<S, T>), dim3(nrhs), dim3(64), 0, s, X, n, j0, rows,
                           linv + (size_t)b * TB * TB);
        update(j0, rows, j0 + rows, j0 + rows, n - j0 - rows);
    }
    for (int b = nblk - 1; b >= 0; --b) {
        const int j0 = b * TB, rows = std::min(TB, n - j0);
        hipLaunchKernelGGL((block_apply_kernel<S, T>), dim3(nrhs), dim3(64), 0, s, X, n, j0, rows,
                           uinv + (size_t)b * TB * TB);
        update(j0, rows, 0, 0, j0);
    }
    HIP_CHECK(hipGetLastError());
}

template <class S, class V>
void lu_solve(const S *dLU, int n, size_t lda, const int *d_ipiv, const double *d_row, const double *d_col, V *dX,
              int nrhs, LUContext &ctx) {
    hipStream_t s = ctx.stream;
    // row permutation from the sequential swaps
    std::vector<int> ipiv(n), perm(n);
    HIP_CHECK(hipMemcpyAsync(ipiv.data(), d_ipiv, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < n; ++k) std::swap(perm[k], perm[ipiv[k]]);
    DeviceBuffer<int> d_perm;
    d_perm.upload(perm, s);
    lu_solve_permuted(dLU, n, lda, d_perm.ptr, d_row, d_col, dX, nrhs, ctx, false);
}

template <class S, class V>
void lu_solve_permuted(const S *dLU, int n, size_t lda, const int *d_perm, const double *d_row,
                       const double *d_col, V *dX, int nrhs, LUContext &ctx, bool same_factor) {
    hipStream_t s = ctx.stream;
    const size_t total = (size_t)n * nrhs;
    // two n x nrhs vectors: the permuted right-hand side and the forward
    // solve's output
    const size_t words = total * sizeof(V) / sizeof(double);
    if (ctx.solve_tmp.count < 2 * words) ctx.solve_tmp.resize(2 * words);
    V *tmp = reinterpret_cast<V *>(ctx.solve_tmp.ptr);
    V *ybuf = tmp + total;
    hipLaunchKernelGGL(permute_scale_kernel<V>, dim3((total + 255) / 256), dim3(256), 0, s, tmp, dX, d_perm, d_row, n,
                       nrhs);
    const int nblk = (n + kChainTB - 1) / kChainTB;
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, s, ctx.solve_bad.ptr, 1, 0);
    using T = typename VecT<V>::A;
    if (ctx.diag_inv.count < (size_t)2 * nblk * kChainTB * kChainTB * sizeof(T)) ctx.dinv_factor = nullptr;
    ctx.diag_inv.reserve((size_t)2 * nblk * kChainTB * kChainTB * sizeof(T));
    T *linv = reinterpret_cast<T *>(ctx.diag_inv.ptr);
    T *uinv = linv + (size_t)nblk * kChainTB * kChainTB;
    // the block inverses depend on the factor only: computed again unless the
    // caller solves with the factor of the previous call
    if (!(same_factor && ctx.dinv_factor == dLU && ctx.dinv_n == n && ctx.dinv_members == 1)) {
        hipLaunchKernelGGL((diag_inverse_kernel<S, V>), dim3(nblk), dim3(256), 0, s, dLU, lda, n, linv, uinv,
                           ctx.solve_bad.ptr);
        ctx.dinv_factor = dLU;
        ctx.dinv_n = n;
        ctx.dinv_members = 1;
    }
    const int grid = chain_grid<S, V>(n, ctx);
    if (nrhs >= kBlockedSolveMinRhs) {
        if constexpr (std::is_same<S, float2>::value) {
            // the FP32 factor's solves in FP32 (refinement corrects them)
            DeviceBuffer<float2> &xs = ctx.solve_f32;
            xs.reserve(total);
            const unsigned g = (unsigned)((total + 255) / 256);
            hipLaunchKernelGGL(narrow_kernel, dim3(g), dim3(256), 0, s, tmp, xs.ptr, total);
            blocked_solve<S, T>(dLU, n, lda, xs.ptr, nrhs, linv, uinv, ctx);
            hipLaunchKernelGGL(widen_kernel, dim3(g), dim3(256), 0, s, xs.ptr, tmp, total);
        } else {
            blocked_solve<S, T>(dLU, n, lda, tmp, nrhs, linv, uinv, ctx);
        }
    } else if (grid > 0) {
        chain_solve<S, V, false>(dLU, n, lda, tmp, ybuf, nrhs, linv, ctx.solve_bad.ptr, grid, ctx);
        chain_solve<S, V, true>(dLU, n, lda, ybuf, tmp, nrhs, uinv, ctx.solve_bad.ptr, grid, ctx);
    } else {
        const int T = 128;
        const int nb2 = (n + T - 1) / T;
        for (int b = 0; b < nb2; ++b) {
            const int i0 = b * T, rows = std::min(T, n - i0);
            if (i0 > 0)
                hipLaunchKernelGGL((trsv_gemv_kernel<S, V>), dim3(rows, nrhs), dim3(256), 0, s, dLU, lda, tmp, n, i0,
                                   rows, 0, i0);
            hipLaunchKernelGGL((trsv_diag_kernel<S, V>), dim3(nrhs), dim3(64), 0, s, dLU, lda, tmp, n, i0, rows, 0,
                               ctx.solve_bad.ptr);
        }
        for (int b = nb2 - 1; b >= 0; --b) {
            const int i0 = b * T, rows = std::min(T, n - i0);
            if (i0 + rows < n)
                hipLaunchKernelGGL((trsv_gemv_kernel<S, V>), dim3(rows, nrhs), dim3(256), 0, s, dLU, lda, tmp, n, i0,
                                   rows, i0 + rows, n);
            hipLaunchKernelGGL((trsv_diag_kernel<S, V>), dim3(nrhs), dim3(64), 0, s, dLU, lda, tmp, n, i0, rows, 1,
                               ctx.solve_bad.ptr);
        }
    }
    if (d_col) hipLaunchKernelGGL(col_scale_kernel<V>, dim3((total + 255) / 256), dim3(256), 0, s, tmp, d_col, n, nrhs);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(dX, tmp, sizeof(V) * total, hipMemcpyDeviceToDevice, s));
    int h_bad = 0;
    HIP_CHECK(hipMemcpyAsync(&h_bad, ctx.solve_bad.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (h_bad & 2) throw HipError{"triangular solve chain timed out"};
    if (h_bad) throw HipError{"singular U diagonal in solve"};
}

// X[b][i] = B[b][perm[i]] / row[b][perm[i]] for the members b of a batch
__global__ void permute_scale_members_kernel(double2 *X, const double2 *B, const int *perm, const double *row, int n,
                                             int members) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)n * members) return;
    const size_t b = idx / n;
    const int p = perm[idx % n];
    const double2 v = B[b * n + p];
    const double r = row[b * n + p];
    X[idx] = make_double2(v.x / r, v.y / r);
}
// X[b][i] /= col[b][i]
__global__ void col_scale_members_kernel(double2 *X, const double *col, size_t total) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= total) return;
    X[idx].x /= col[idx];
    X[idx].y /= col[idx];
}

void lu_solve_members(const float2 *dLU, size_t lu_stride, int n, const int *d_perm, const double *d_row,
                      const double *d_col, double2 *dX, int members, LUContext &ctx, bool same_factors) {
    using S = float2;
    using V = double2;
    using T = double2;
    hipStream_t s = ctx.stream;
    if (members < 1) return;
    const size_t total = (size_t)n * members;
    if (ctx.solve_tmp.count < 4 * total) ctx.solve_tmp.resize(4 * total);
    V *tmp = reinterpret_cast<V *>(ctx.solve_tmp.ptr);
    V *ybuf = tmp + total;
    const unsigned g = (unsigned)((total + 255) / 256);
    hipLaunchKernelGGL(permute_scale_members_kernel, dim3(g), dim3(256), 0, s, tmp, dX, d_perm, d_row, n, members);
    const int nblk = (n + kChainTB - 1) / kChainTB;
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, s, ctx.solve_bad.ptr, 1, 0);
    // the block inverses of every member: L's of all members, then U's
    const size_t dstride = (size_t)nblk * kChainTB * kChainTB;
    const size_t need = 2 * (size_t)members * dstride * sizeof(T);
    if (ctx.diag_inv.count < need) ctx.dinv_factor = nullptr;
    ctx.diag_inv.reserve(need);
    T *linv = reinterpret_cast<T *>(ctx.diag_inv.ptr);
    T *uinv = linv + (size_t)members * dstride;
    if (!(same_factors && ctx.dinv_factor == dLU && ctx.dinv_n == n && ctx.dinv_members == members)) {
        hipLaunchKernelGGL((diag_inverse_kernel<S, V>), dim3(nblk, members), dim3(256), 0, s, dLU, (size_t)n, n, linv,
                           uinv, ctx.solve_bad.ptr, lu_stride, dstride);
        ctx.dinv_factor = dLU;
        ctx.dinv_n = n;
        ctx.dinv_members = members;
    }
    // every member's chain gets an equal share of the co-resident workgroups
    const void *fwd = reinterpret_cast<const void *>(trsv_chain1_kernel<S, false, kChainDeps>);
    const void *bwd = reinterpret_cast<const void *>(trsv_chain1_kernel<S, true, kChainDeps>);
    int f = 0, bw = 0;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&f, fwd, 256, 0));
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bw, bwd, 256, 0));
    const long long cap = (long long)std::min(f, bw) * ctx.cus / members;
    if (cap < 1) throw HipError{"batched triangular solve: too many members for co-resident chains"};
    const int G = (int)std::min<long long>(nblk, cap);
    auto chain = [&](bool upper, const V *B, V *Y, const T *dinv) {
        const size_t words = total * 2;
        hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s,
                           reinterpret_cast<unsigned long long *>(Y), words, kChainEmpty);
        int nn = n;
        size_t ld = (size_t)n, ms = lu_stride, vs = (size_t)n, ds = dstride;
        int *bad = ctx.solve_bad.ptr;
        void *args[] = {(void *)&dLU, &ld, (void *)&B, &Y, &nn, (void *)&dinv, &bad, &ms, &vs, &ds};
        HIP_CHECK(hipLaunchCooperativeKernel(upper ? bwd : fwd, dim3(G, members), dim3(256), args, 0, s));
    };
    chain(false, tmp, ybuf, linv);
    chain(true, ybuf, tmp, uinv);
    if (d_col) hipLaunchKernelGGL(col_scale_members_kernel, dim3(g), dim3(256), 0, s, tmp, d_col, total);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(dX, tmp, sizeof(V) * total, hipMemcpyDeviceToDevice, s));
    int h_bad = 0;
    HIP_CHECK(hipMemcpyAsync(&h_bad, ctx.solve_bad.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (h_bad & 2) throw HipError{"triangular solve chain timed out"};
    if (h_bad) throw HipError{"singular U diagonal in solve"};
}

template void lu_solve<double2, double2>(const double2 *, int, size_t, const int *, const double *, const double *,
                                         double2 *, int, LUContext &);
template void lu_solve<float2, double2>(const float2 *, int, size_t, const int *, const double *, const double *,
                                        double2 *, int, LUContext &);
template void lu_solve<double, double>(const double *, int, size_t, const int *, const double *, const double *,
                                       double *, int, LUContext &);
template void lu_solve_permuted<double2, double2>(const double2 *, int, size_t, const int *, const double *,
                                                  const double *, double2 *, int, LUContext &, bool);
template void lu_solve_permuted<float2, double2>(const float2 *, int, size_t, const int *, const double *,
                                                 const double *, double2 *, int, LUContext &, bool);
template void lu_solve_permuted<double, double>(const double *, int, size_t, const int *, const double *,
                                                const double *, double *, int, LUContext &, bool);

void permutation_from_pivots(const int *d_ipiv, int n, int *d_perm, hipStream_t s) {
    std::vector<int> ipiv(n), perm(n);
    HIP_CHECK(hipMemcpyAsync(ipiv.data(), d_ipiv, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < n; ++k) std::swap(perm[k], perm[ipiv[k]]);
    HIP_CHECK(hipMemcpyAsync(d_perm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
}


// residual and backward-error terms (harmonic.c:899-969), one block per row
// of a rows x cols matrix (cols summed per row)
__global__ __launch_bounds__(256) void residual_kernel(const double2 *A, size_t lda, int cols, const double2 *x,
                                                      const double *ax, const double2 *b, double2 *r,
                                                      double *terms) {
    __shared__ double sre[256], sim[256], sden[256];
    const int i = blockIdx.x;
    double yr = 0.0, yi = 0.0, den = 0.0;
    const double2 *row = A + (size_t)i * lda;
    for (int j = threadIdx.x; j < cols; j += 256) {
        const double2 z = row[j];
        const double2 v = x[j];
        yr += z.x * v.x - z.y * v.y;
        yi += z.x * v.y + z.y * v.x;
        den += hypot(z.x, z.y) * ax[j];
    }
    sre[threadIdx.x] = yr;
    sim[threadIdx.x] = yi;
    sden[threadIdx.x] = den;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if (threadIdx.x < s2) {
            sre[threadIdx.x] += sre[threadIdx.x + s2];
            sim[threadIdx.x] += sim[threadIdx.x + s2];
            sden[threadIdx.x] += sden[threadIdx.x + s2];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double2 bi = b ? b[i] : make_double2(0.0, 0.0);
        const double rr = sre[0] - bi.x, ri = sim[0] - bi.y;
        if (r) r[i] = make_double2(rr, ri);
        const double babs = hypot(bi.x, bi.y);
        terms[3 * (size_t)i] = hypot(rr, ri);
        terms[3 * (size_t)i + 1] = sden[0] + babs;
        terms[3 * (size_t)i + 2] = babs;
    }
}

__global__ void abs_kernel(const double2 *x, double *ax, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) ax[i] = hypot(x[i].x, x[i].y);
}

void complex_residual(const double2 *dA, int rows, int cols, size_t lda, const double2 *dx, const double2 *db,
                      double2 *dr, double *dterms, hipStream_t s, double *d_ax) {
    DeviceBuffer<double> own;
    if (!d_ax) {
        own.resize(cols);
        d_ax = own.ptr;
    }
    hipLaunchKernelGGL(abs_kernel, dim3((cols + 255) / 256), dim3(256), 0, s, dx, d_ax, cols);
    hipLaunchKernelGGL(residual_kernel, dim3(rows), dim3(256), 0, s, dA, lda, cols, dx, d_ax, db, dr, dterms);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));
}

// Residuals of up to kResidualRhs right-hand sides per pass, one workgroup
// per row: R[q] = A X[q] - B[q] and the component-wise backward error
// eta[q] = max_i |r_qi| / (sum_j |a_ij| |x_qj| + |b_qi|) (harmonic.c:932-969),
// reduced with an integer max on the bit patterns of the non-negative
// values, so the result does not depend on the order of the rows.
constexpr int kResidualRhs = 8;

__global__ void abs_multi_kernel(const double2 *x, double *ax, size_t count) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < count) ax[i] = hypot(x[i].x, x[i].y);
}

__global__ void zero_u64_kernel(unsigned long long *p, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// One workgroup per row (a wavefront per row left most of the device idle
// in the last of the ~1.5 rounds of rows); NQ = 1, or up to 8 right-hand
// sides, so the one-right-hand-side case keeps one set of sums.
template <int NQ, class Src>
__global__ __launch_bounds__(256) void residual_multi_kernel(Src src, int n, const double2 *X, const double *AX,
                                                             const double2 *B, double2 *R, int nq,
                                                             unsigned long long *eta_bits) {
    __shared__ double red[4][3 * NQ];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = (int)blockIdx.x;
    const int nqq = NQ == 1 ? 1 : nq;
    double yr[NQ], yi[NQ], den[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) yr[q] = yi[q] = den[q] = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) {
        const double2 z = src(i, j);
        const double za = hypot(z.x, z.y);
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            if (NQ == 1 || q < nqq) {
                const double2 v = X[(size_t)q * n + j];
                yr[q] += z.x * v.x - z.y * v.y;
                yi[q] += z.x * v.y + z.y * v.x;
                den[q] += za * AX[(size_t)q * n + j];
            }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q)
        if (NQ == 1 || q < nqq) {
            for (int off = 32; off > 0; off >>= 1) {
                yr[q] += __shfl_xor(yr[q], off);
                yi[q] += __shfl_xor(yi[q], off);
                den[q] += __shfl_xor(den[q], off);
            }
            if (lane == 0) {
                red[wave][3 * q] = yr[q];
                red[wave][3 * q + 1] = yi[q];
                red[wave][3 * q + 2] = den[q];
            }
        }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            if (NQ == 1 || q < nqq) {
                const double sr = (red[0][3 * q] + red[1][3 * q]) + (red[2][3 * q] + red[3][3 * q]);
                const double si = (red[0][3 * q + 1] + red[1][3 * q + 1]) + (red[2][3 * q + 1] + red[3][3 * q + 1]);
                const double sd = (red[0][3 * q + 2] + red[1][3 * q + 2]) + (red[2][3 * q + 2] + red[3][3 * q + 2]);
                const double2 bi = B[(size_t)q * n + i];
                const double rr = sr - bi.x, ri = si - bi.y;
                R[(size_t)q * n + i] = make_double2(rr, ri);
                const double ra = hypot(rr, ri), d = sd + hypot(bi.x, bi.y);
                const double eta = d > 0.0 ? ra / d : ra;
                atomicMax(&eta_bits[q], (unsigned long long)__double_as_longlong(eta));
            }
    }
}

// The same for many right-hand sides on FP64 MFMA: Z X (complex, four real
// products per block pair) and |Z| |X| (the denominators, a fifth) in one
// pass over Z, which is read directly (16 rows x 4 consecutive columns per
// operand load) rather than re-read by every group of right-hand sides
// through the caches. Workgroup tile 128 rows x 64 right-hand sides (2 x 2
// waves of 64 x 32); the epilogue forms r = Z x - b, stores it, and takes
// the maximum of |r_i| / (den_i + |b_i|) over the wave's rows per right-hand
// side with one atomic each.
constexpr int kResidualMfmaMinRhs = kResidualDenseMinRhs;

__global__ __launch_bounds__(256) void residual_mfma_kernel(const double2 *A, size_t lda, int n, const double2 *X,
                                                            const double2 *B, double2 *R, int nrhs,
                                                            unsigned long long *eta_bits) {
    constexpr int MB = 4, NB = 2;
    const int tiles_q = (nrhs + 63) / 64;
    const int tm = (int)blockIdx.x / tiles_q, tq = (int)blockIdx.x % tiles_q;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int r0 = tm * 128 + (wave >> 1) * 64, q0 = tq * 64 + (wave & 1) * 32;
    v4d cre[MB][NB], cim[MB][NB], cden[MB][NB];
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) cre[i][j] = cim[i][j] = cden[i][j] = v4d{0, 0, 0, 0};
    // operand rows / right-hand sides of this lane (clamped in bounds)
    const double2 *arow[MB];
    const double2 *xcol[NB];
#pragma unroll
    for (int i = 0; i < MB; ++i) arow[i] = A + (size_t)min(r0 + 16 * i + li, n - 1) * lda;
#pragma unroll
    for (int j = 0; j < NB; ++j) xcol[j] = X + (size_t)min(q0 + 16 * j + li, nrhs - 1) * n;
    // raw operands of two k steps in flight (loaded two steps ahead of their
    // products); each step's MFMAs take the same operands in the same order
    // as a load-then-multiply loop, so the sums are the same
    double2 zr[2][MB], xr[2][NB];
    auto load = [&](int st, int k0) {
        const int kc = min(k0 + lk, n - 1);
#pragma unroll
        for (int i = 0; i < MB; ++i) zr[st][i] = arow[i][kc];
#pragma unroll
        for (int j = 0; j < NB; ++j) xr[st][j] = xcol[j][kc];
    };
    auto step = [&](int st, int k0) {
        const bool in = k0 + lk < n;
        double ar[MB], ai[MB], am[MB], br[NB], bi[NB], bm[NB];
#pragma unroll
        for (int i = 0; i < MB; ++i) {
            ar[i] = in ? zr[st][i].x : 0.0;
            ai[i] = in ? zr[st][i].y : 0.0;
            am[i] = sqrt(fma(ar[i], ar[i], ai[i] * ai[i]));
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            br[j] = in ? xr[st][j].x : 0.0;
            bi[j] = in ? xr[st][j].y : 0.0;
            bm[j] = sqrt(fma(br[j], br[j], bi[j] * bi[j]));
        }
        // this stage's registers are free again: the step after next
        if (k0 + 8 < n) load(st, k0 + 8);
#pragma unroll
        for (int i = 0; i < MB; ++i) {
            const double nai = -ai[i];
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                cre[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], cre[i][j], 0, 0, 0);
                cim[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], cim[i][j], 0, 0, 0);
                cre[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai, bi[j], cre[i][j], 0, 0, 0);
                cim[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br[j], cim[i][j], 0, 0, 0);
                cden[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(am[i], bm[j], cden[i][j], 0, 0, 0);
            }
        }
    };
    load(0, 0);
    if (4 < n) load(1, 4);
    int k0 = 0;
    for (; k0 + 4 < n; k0 += 8) {
        step(0, k0);
        step(1, k0 + 4);
    }
    if (k0 < n) step(0, k0);
    // accumulator element v of lane (li, lk): row 4 v + lk of the 16-row
    // block, right-hand side li of the 16-column block
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int q = q0 + 16 * j + li;
        double emax = 0.0;
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = r0 + 16 * i + 4 * v + lk;
                if (row < n && q < nrhs) {
                    const double2 b = B[(size_t)q * n + row];
                    const double rr = cre[i][j][v] - b.x, ri = cim[i][j][v] - b.y;
                    R[(size_t)q * n + row] = make_double2(rr, ri);
                    const double ra = hypot(rr, ri), d = cden[i][j][v] + hypot(b.x, b.y);
                    emax = fmax(emax, d > 0.0 ? ra / d : ra);
                }
            }
        emax = fmax(emax, __shfl_xor(emax, 16));
        emax = fmax(emax, __shfl_xor(emax, 32));
        if (lk == 0 && q < nrhs) atomicMax(&eta_bits[q], (unsigned long long)__double_as_longlong(emax));
    }
}

// The one-right-hand-side residual of the MNA matrix evaluated from its
// blocks, visiting only the entries that can be nonzero: row i of an edge
// has the Lp (R + jwL) part and two incidence entries (-1, +1 at its end
// nodes), row i of a node has the -P A^T part and jw on the diagonal. Each
// thread sums the same columns in the same order as residual_multi_kernel
// (whose other terms are exact zeros), with four columns' loads issued
// together, so the sums and the result are the same.
__global__ __launch_bounds__(256) void residual_mna1_kernel(MnaOperands m, const double2 *__restrict__ X,
                                                            const double *__restrict__ AX,
                                                            const double2 *__restrict__ B, double2 *__restrict__ R,
                                                            unsigned long long *eta_bits) {
    __shared__ double red[4][3];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, t = threadIdx.x;
    const int i = (int)blockIdx.x, ne = m.ne, n = m.ne + m.nv;
    double yr = 0.0, yi = 0.0, den = 0.0;
    auto add = [&](double2 z, int j) {
        const double za = hypot(z.x, z.y);
        const double2 v = X[j];
        yr += z.x * v.x - z.y * v.y;
        yi += z.x * v.y + z.y * v.x;
        den += za * AX[j];
    };
    if (i < ne) {
        const double2 *lrow = m.Lw + (size_t)i * ne;
        const double rd = m.R ? m.R[i] : 0.0;
        int j = t;
        for (; j + 768 < ne; j += 1024) {
            double2 l[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) l[u] = lrow[j + 256 * u];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double r = (j + 256 * u == i) ? rd : 0.0;
                add(make_double2(r - m.omega * l[u].y, m.omega * l[u].x), j + 256 * u);
            }
        }
        for (; j < ne; j += 256) {
            const double2 l = lrow[j];
            const double r = j == i ? rd : 0.0;
            add(make_double2(r - m.omega * l.y, m.omega * l.x), j);
        }
        // incidence entries, in column order within the thread that owns them
        const int ja = ne + m.edges[2 * i], jb = ne + m.edges[2 * i + 1];
        const int j1 = min(ja, jb), j2 = max(ja, jb);
        if ((j1 & 255) == t) add(make_double2(j1 == jb ? 1.0 : -1.0, 0.0), j1);
        if ((j2 & 255) == t) add(make_double2(j2 == jb ? 1.0 : -1.0, 0.0), j2);
    } else {
        const int p = i - ne;
        const double2 *prow = m.Pw + (size_t)p * m.nv;
        int j = t;
        for (; j + 768 < ne; j += 1024) {
            int a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a[u] = m.edges[2 * (j + 256 * u)];
                b[u] = m.edges[2 * (j + 256 * u) + 1];
            }
            double2 pa[4], pb[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                pb[u] = prow[b[u]];
                pa[u] = prow[a[u]];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) add(make_double2(-(pb[u].x - pa[u].x), -(pb[u].y - pa[u].y)), j + 256 * u);
        }
        for (; j < ne; j += 256) {
            const int a = m.edges[2 * j], b = m.edges[2 * j + 1];
            const double2 pb = prow[b], pa = prow[a];
            add(make_double2(-(pb.x - pa.x), -(pb.y - pa.y)), j);
        }
        const int jd = ne + p;
        if ((jd & 255) == t) add(make_double2(0.0, m.omega), jd);
    }
    for (int off = 32; off > 0; off >>= 1) {
        yr += __shfl_xor(yr, off);
        yi += __shfl_xor(yi, off);
        den += __shfl_xor(den, off);
    }
    if (lane == 0) {
        red[wave][0] = yr;
        red[wave][1] = yi;
        red[wave][2] = den;
    }
    __syncthreads();
    if (t == 0) {
        const double sr = (red[0][0] + red[1][0]) + (red[2][0] + red[3][0]);
        const double si = (red[0][1] + red[1][1]) + (red[2][1] + red[3][1]);
        const double sd = (red[0][2] + red[1][2]) + (red[2][2] + red[3][2]);
        const double2 bi = B[i];
        const double rr = sr - bi.x, ri = si - bi.y;
        R[i] = make_double2(rr, ri);
        const double ra = hypot(rr, ri), d = sd + hypot(bi.x, bi.y);
        const double eta = d > 0.0 ? ra / d : ra;
        atomicMax(eta_bits, (unsigned long long)__double_as_longlong(eta));
    }
    (void)n;
}

// the row-wise residual passes (up to kResidualRhs right-hand sides each)
template <class Src>
static void residual_rows(Src src, int n, const double2 *dX, const double2 *dB, double2 *dR, int nrhs, double *d_ax,
                          unsigned long long *d_eta_bits, hipStream_t s) {
    const size_t total = (size_t)n * nrhs;
    hipLaunchKernelGGL(abs_multi_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, dX, d_ax, total);
    for (int q0 = 0; q0 < nrhs; q0 += kResidualRhs) {
        const int nq = std::min(kResidualRhs, nrhs - q0);
        const size_t off = (size_t)q0 * n;
        bool mna1 = false;
        if constexpr (std::is_same<Src, MnaSource>::value)
            mna1 = nq == 1 && src.m.Lw && src.m.Pw && !src.m.PAT;
        if (mna1) {
            if constexpr (std::is_same<Src, MnaSource>::value)
                hipLaunchKernelGGL(residual_mna1_kernel, dim3((unsigned)n), dim3(256), 0, s, src.m, dX + off,
                                   d_ax + off, dB + off, dR + off, d_eta_bits + q0);
        } else if (nq == 1)
            hipLaunchKernelGGL((residual_multi_kernel<1, Src>), dim3((unsigned)n), dim3(256), 0, s, src, n, dX + off,
                               d_ax + off, dB + off, dR + off, nq, d_eta_bits + q0);
        else
            hipLaunchKernelGGL((residual_multi_kernel<kResidualRhs, Src>), dim3((unsigned)n), dim3(256), 0, s, src, n,
                               dX + off, d_ax + off, dB + off, dR + off, nq, d_eta_bits + q0);
    }
    HIP_CHECK(hipGetLastError());
}

void residual_multi(const double2 *dA, int n, size_t lda, const double2 *dX, const double2 *dB, double2 *dR,
                    int nrhs, double *d_ax, unsigned long long *d_eta_bits, hipStream_t s) {
    const size_t total = (size_t)n * nrhs;
    hipLaunchKernelGGL(zero_u64_kernel, dim3((nrhs + 255) / 256), dim3(256), 0, s, d_eta_bits, nrhs);
    if (nrhs >= kResidualMfmaMinRhs) {
        const int grid = ((n + 127) / 128) * ((nrhs + 63) / 64);
        hipLaunchKernelGGL(residual_mfma_kernel, dim3((unsigned)grid), dim3(256), 0, s, dA, lda, n, dX, dB, dR, nrhs,
                           d_eta_bits);
        HIP_CHECK(hipGetLastError());
        return;
    }
    residual_rows(DenseSource{dA, lda}, n, dX, dB, dR, nrhs, d_ax, d_eta_bits, s);
}

void residual_multi(const MnaOperands &m, const double2 *dX, const double2 *dB, double2 *dR, int nrhs, double *d_ax,
                    unsigned long long *d_eta_bits, hipStream_t s) {
    if (nrhs >= kResidualMfmaMinRhs)
        throw HipError{"residual_multi: many right-hand sides need the dense matrix (MFMA residual)"};
    hipLaunchKernelGGL(zero_u64_kernel, dim3((nrhs + 255) / 256), dim3(256), 0, s, d_eta_bits, nrhs);
    residual_rows(MnaSource{m}, m.ne + m.nv, dX, dB, dR, nrhs, d_ax, d_eta_bits, s);
}

void mfma_layout_probe(double *host_out) {
    DeviceBuffer<double> d(1024);
    hipLaunchKernelGGL(mfma_layout_probe_kernel, dim3(1), dim3(64), 0, 0, d.ptr);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(host_out, d.ptr, 1024 * sizeof(double), hipMemcpyDeviceToHost));
}

}  // namespace peec

// Blocked right-looking LU on gfx950 (see lu.h). Per block column k0:
//   panel_kernel    cooperative grid: each workgroup keeps a slab of the
//                   panel's rows in LDS; per column one grid barrier decides
//                   the pivot (largest |a_ik|, first index on ties, as the
//                   reference's strict '>' scan in complex_lu.c:341-360),
//                   exchanges the pivot row through a small HBM slot and
//                   applies the rank-1 update to the LDS slab.
//   laswp_kernel    applies the panel's row swaps to the other columns.
//   trsm_kernel     U12 = L11^{-1} A12, 8 lanes per column with a shuffle
//                   reduction.
//   gemm_kernel     A22 -= L21 U12 on MFMA (v_mfma_f64_16x16x4f64 for FP64,
//                   v_mfma_f32_16x16x4f32 for the mixed-precision factor).
// Triangular solves are blocked (GEMV + diagonal-block kernels).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "lu.h"

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

// ---------------------------------------------------------------------------
// equilibration (complex_lu.c:275-320)

__global__ void row_max_kernel(const double2 *A, int n, size_t lda, double *row, int *bad) {
    __shared__ double red[256];
    const int i = blockIdx.x;
    double m = 0.0;
    bool finite = true;
    for (int j = threadIdx.x; j < n; j += 256) {
        const double v = mag(A[(size_t)i * lda + j]);
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
}

int equilibrate_complex(double2 *dA, int n, size_t lda, double *d_row, double *d_col,
                        double *max_scaled, LUContext &ctx) {
    hipStream_t s = ctx.stream;
    DeviceBuffer<int> bad(1);
    DeviceBuffer<double> mx(1);
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, s, bad.ptr, 1, 0);
    hipLaunchKernelGGL(fill_zero_kernel, dim3(1), dim3(256), 0, s, mx.ptr, (size_t)1);
    hipLaunchKernelGGL(fill_zero_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_col, (size_t)n);
    hipLaunchKernelGGL(row_max_kernel, dim3(n), dim3(256), 0, s, dA, n, lda, d_row, bad.ptr);
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

template <class S>
__global__ __launch_bounds__(256) void panel_kernel(S *A, size_t lda, int n, int k0, int nb, int rpw, double *slots,
                                                    double *mags, unsigned long long *tags, int *ipiv,
                                                    unsigned round0, double limit, int *status) {
    using T = typename Elem<S>::A;
    constexpr int W = sizeof(T) / sizeof(double);  // doubles per element
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    T *rows = reinterpret_cast<T *>(smem_raw);
    __shared__ T u[kMaxNB];
    __shared__ double red_m[256];
    __shared__ int red_i[256], red_g[256];
    __shared__ int s_p, s_gw, s_fail;
    const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
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
        red_m[tid] = bm;
        red_i[tid] = bi;
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if (tid < st) {
                const double om = red_m[tid + st];
                const int oi = red_i[tid + st];
                if (om > red_m[tid] || (om == red_m[tid] && oi < red_i[tid])) {
                    red_m[tid] = om;
                    red_i[tid] = oi;
                }
            }
            __syncthreads();
        }
        bm = red_m[0];
        bi = red_i[0];
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
            red_m[tid] = m;
            red_i[tid] = idx;
            red_g[tid] = timeout ? -1 : tid;
            __syncthreads();
            if (tid == 0) s_fail = 0;
            __syncthreads();
            if (timeout) atomicOr(&s_fail, 1);
            for (int st = 128; st > 0; st >>= 1) {
                if (tid < st) {
                    const double om = red_m[tid + st];
                    const int oi = red_i[tid + st];
                    if (om > red_m[tid] || (om == red_m[tid] && oi < red_i[tid])) {
                        red_m[tid] = om;
                        red_i[tid] = oi;
                        red_g[tid] = red_g[tid + st];
                    }
                }
                __syncthreads();
            }
            if (tid == 0) {
                const double best = red_m[0];
                s_p = red_i[0];
                s_gw = red_g[0];
                if (s_fail) {
                    status[0] = -3;
                } else if (!(best > limit) || !isfinite(best) || red_i[0] == INT_MAX) {
                    if (g == 0) {
                        status[0] = -2;
                        status[1] = k;
                    }
                    s_fail = 1;
                }
            }
        }
        __syncthreads();
        if (s_fail) return;
        const int p = s_p, gw = s_gw;
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

// apply the swaps of one panel to every other column
template <class S>
__global__ void laswp_kernel(S *A, size_t lda, int n, int k0, int w, const int *ipiv, const int *status) {
    if (status[0] != 0) return;
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n - w) return;
    const int col = c < k0 ? c : c + w;
    for (int t = 0; t < w; ++t) {
        const int p = ipiv[k0 + t];
        if (p != k0 + t) {
            S *a = A + (size_t)(k0 + t) * lda + col;
            S *b = A + (size_t)p * lda + col;
            const S tmp = *a;
            *a = *b;
            *b = tmp;
        }
    }
}

// U12 = L11^{-1} A12; 8 lanes per column (32 columns per block)
template <class S>
__global__ __launch_bounds__(256) void trsm_kernel(S *A, size_t lda, int n, int k0, int w, const int *status) {
    using T = typename Elem<S>::A;
    __shared__ T xs[32][kMaxNB];
    if (status[0] != 0) return;
    const int cl = threadIdx.x >> 3, sub = threadIdx.x & 7;
    const int col = k0 + w + blockIdx.x * 32 + cl;
    const bool valid = col < n;
    for (int r = 0; r < w; ++r) {
        T part = ezero<T>();
        const S *lrow = A + (size_t)(k0 + r) * lda + k0;
        for (int t = sub; t < r; t += 8) part = eadd(part, emul(ld(lrow + t), xs[cl][t]));
        part = eadd(part, eshfl_xor(part, 1));
        part = eadd(part, eshfl_xor(part, 2));
        part = eadd(part, eshfl_xor(part, 4));
        T x = ezero<T>();
        if (valid) x = esub(ld(A + (size_t)(k0 + r) * lda + col), part);
        if (sub == 0) {
            xs[cl][r] = x;
            if (valid) st(A + (size_t)(k0 + r) * lda + col, x);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// ---------------------------------------------------------------------------
// trailing update on MFMA. Block tile 64x64, 4 waves of 32x32, each wave a
// 2x2 grid of 16x16 MFMA blocks. Operand layout of the 16x16x4 MFMAs (checked
// on gfx950 by mfma_layout_probe_kernel): A[i][k] in lane i + 16k, B[k][j] in
// lane j + 16k; accumulator element v of lane l holds row 4v + l/16 for the
// FP64 instruction and row 4(l/16) + v for the FP32 one, column l % 16.

template <class S> struct Mfma;
template <> struct Mfma<double2> {
    using acc = v4d;
    using op = double;
    __device__ static acc mma(op a, op b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static int row(int lk, int v) { return 4 * v + lk; }
    __device__ static void load(const double2 *p, op &re, op &im) {
        const double2 v = *p;
        re = v.x;
        im = v.y;
    }
};
template <> struct Mfma<float2> {
    using acc = v4f;
    using op = float;
    __device__ static acc mma(op a, op b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static int row(int lk, int v) { return 4 * lk + v; }
    __device__ static void load(const float2 *p, op &re, op &im) {
        const float2 v = *p;
        re = v.x;
        im = v.y;
    }
};

template <class S>
__global__ __launch_bounds__(256) void gemm_complex_kernel(S *A, size_t lda, int row0, int col0, int m,
                                                           int ncols, int k0, int kw, const int *status) {
    using M = Mfma<S>;
    using acc_t = typename M::acc;
    using op = typename M::op;
    if (status[0] != 0) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = blockIdx.y * 64 + (wave >> 1) * 32;
    const int wc = blockIdx.x * 64 + (wave & 1) * 32;
    acc_t cre[2][2], cim[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            cre[a][b] = acc_t{0, 0, 0, 0};
            cim[a][b] = acc_t{0, 0, 0, 0};
        }
    const int li = lane & 15, lk = lane >> 4;
    for (int kk = 0; kk < kw; kk += 4) {
        const int t = kk + lk;
        op are[2], aim[2], bre[2], bim[2];
        for (int mi = 0; mi < 2; ++mi) {
            const int row = wr + mi * 16 + li;
            are[mi] = aim[mi] = 0;
            if (row < m && t < kw) M::load(A + (size_t)(row0 + row) * lda + k0 + t, are[mi], aim[mi]);
        }
        for (int ni = 0; ni < 2; ++ni) {
            const int col = wc + ni * 16 + li;
            bre[ni] = bim[ni] = 0;
            if (col < ncols && t < kw) M::load(A + (size_t)(k0 + t) * lda + col0 + col, bre[ni], bim[ni]);
        }
        for (int mi = 0; mi < 2; ++mi)
            for (int ni = 0; ni < 2; ++ni) {
                cre[mi][ni] = M::mma(are[mi], bre[ni], cre[mi][ni]);
                cre[mi][ni] = M::mma(-aim[mi], bim[ni], cre[mi][ni]);
                cim[mi][ni] = M::mma(are[mi], bim[ni], cim[mi][ni]);
                cim[mi][ni] = M::mma(aim[mi], bre[ni], cim[mi][ni]);
            }
    }
    for (int mi = 0; mi < 2; ++mi)
        for (int ni = 0; ni < 2; ++ni) {
            const int col = wc + ni * 16 + li;
            if (col >= ncols) continue;
            for (int v = 0; v < 4; ++v) {
                const int row = wr + mi * 16 + M::row(lk, v);
                if (row >= m) continue;
                S *c = A + (size_t)(row0 + row) * lda + col0 + col;
                S cv = *c;
                cv.x -= cre[mi][ni][v];
                cv.y -= cim[mi][ni][v];
                *c = cv;
            }
        }
}

__global__ __launch_bounds__(256) void gemm_real_kernel(double *A, size_t lda, int row0, int col0, int m,
                                                        int ncols, int k0, int kw, const int *status) {
    if (status[0] != 0) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = blockIdx.y * 64 + (wave >> 1) * 32;
    const int wc = blockIdx.x * 64 + (wave & 1) * 32;
    v4d c[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) c[a][b] = v4d{0, 0, 0, 0};
    const int li = lane & 15, lk = lane >> 4;
    for (int kk = 0; kk < kw; kk += 4) {
        const int t = kk + lk;
        double a[2], b[2];
        for (int mi = 0; mi < 2; ++mi) {
            const int row = wr + mi * 16 + li;
            a[mi] = (row < m && t < kw) ? A[(size_t)(row0 + row) * lda + k0 + t] : 0.0;
        }
        for (int ni = 0; ni < 2; ++ni) {
            const int col = wc + ni * 16 + li;
            b[ni] = (col < ncols && t < kw) ? A[(size_t)(k0 + t) * lda + col0 + col] : 0.0;
        }
        for (int mi = 0; mi < 2; ++mi)
            for (int ni = 0; ni < 2; ++ni) c[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mi], b[ni], c[mi][ni], 0, 0, 0);
    }
    for (int mi = 0; mi < 2; ++mi)
        for (int ni = 0; ni < 2; ++ni) {
            const int col = wc + ni * 16 + li;
            if (col >= ncols) continue;
            for (int v = 0; v < 4; ++v) {
                const int row = wr + mi * 16 + 4 * v + lk;  // FP64 accumulator layout
                if (row < m) A[(size_t)(row0 + row) * lda + col0 + col] -= c[mi][ni][v];
            }
        }
}

template <class S>
void launch_gemm(S *A, size_t lda, int row0, int col0, int m, int ncols, int k0, int kw, const int *status,
                 hipStream_t s) {
    dim3 grid((ncols + 63) / 64, (m + 63) / 64);
    if constexpr (Elem<S>::cplx)
        hipLaunchKernelGGL(gemm_complex_kernel<S>, grid, dim3(256), 0, s, A, lda, row0, col0, m, ncols, k0, kw, status);
    else
        hipLaunchKernelGGL(gemm_real_kernel, grid, dim3(256), 0, s, A, lda, row0, col0, m, ncols, k0, kw, status);
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

template <class S>
LUStatus lu_factor(S *dA, int n, size_t lda, int *d_ipiv, double pivot_limit, LUContext &ctx, int block_size) {
    using T = typename Elem<S>::A;
    hipStream_t s = ctx.stream;
    LUStatus out;
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, s, ctx.status.ptr, 2, 0);
    hipLaunchKernelGGL(fill_uint_kernel, dim3((unsigned)((kBarWords + 255) / 256)), dim3(256), 0, s, ctx.barrier.ptr,
                       (int)kBarWords, 0u);
    HIP_CHECK(hipGetLastError());
    const size_t lds_budget = 60 * 1024;
    int k0 = 0;
    while (k0 < n) {
        const int m = n - k0;
        // workgroups: one per CU at most, at least 8 rows each
        int G = std::max(1, std::min(ctx.cus, (m + 7) / 8));
        int nb = block_size > 0 ? block_size : 128;
        nb = std::min(nb, kMaxNB);
        int rpw = (m + G - 1) / G;
        while (nb > 16 && (size_t)rpw * nb * sizeof(T) > lds_budget) nb /= 2;
        while ((size_t)rpw * nb * sizeof(T) > lds_budget) {
            // very tall panels: more rows than the LDS budget allows per CU
            out.code = -1;
            PEEC_ERR("LU panel of %d rows exceeds the LDS budget.", m);
            return out;
        }
        const int w = std::min(nb, m);
        rpw = (m + G - 1) / G;
        double *slots = ctx.slots.ptr;
        double *mags = ctx.cand.ptr;
        unsigned long long *tags = ctx.tags.ptr;
        const size_t shm = (size_t)rpw * w * sizeof(T);
        int nn = n, kk0 = k0, ww = w, rr = rpw;
        size_t ld = lda;
        unsigned round0 = ctx.round;
        ctx.round += (unsigned)w + 1u;
        double lim = pivot_limit;
        void *args[] = {&dA, &ld, &nn, &kk0, &ww, &rr, &slots, &mags, &tags, &d_ipiv, &round0, &lim, &ctx.status.ptr};
        const int tp = ctx.timer ? ctx.timer->start(s) : -1;
        HIP_CHECK(hipLaunchCooperativeKernel(reinterpret_cast<void *>(panel_kernel<S>), dim3(G), dim3(256),
                                             args, (unsigned)shm, s));
        if (ctx.timer) ctx.timer->stop(tp, s, kKernelPanel, 0.0, 2.0 * m * w * sizeof(S));
        if (n - w > 0)
            hipLaunchKernelGGL(laswp_kernel<S>, dim3((n - w + 255) / 256), dim3(256), 0, s, dA, lda, n, k0, w,
                               d_ipiv, ctx.status.ptr);
        const int rest = n - k0 - w;
        if (rest > 0) {
            hipLaunchKernelGGL(trsm_kernel<S>, dim3((rest + 31) / 32), dim3(256), 0, s, dA, lda, n, k0, w,
                               ctx.status.ptr);
            const int tg = ctx.timer ? ctx.timer->start(s) : -1;
            launch_gemm<S>(dA, lda, k0 + w, k0 + w, rest, rest, k0, w, ctx.status.ptr, s);
            if (ctx.timer) {
                const double mnk = (double)rest * rest * w;
                ctx.timer->stop(tg, s, kKernelLuUpdate, (Elem<S>::cplx ? 8.0 : 2.0) * mnk,
                                (2.0 * rest * rest + 2.0 * rest * w) * sizeof(S));
            }
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

// Triangular solve in one cooperative launch: workgroup w owns the 32-row
// block b (b = w forward, b = nblk-1-w backward). It accumulates the
// off-diagonal products of every block it depends on as soon as that block's
// flag shows it final, then solves its diagonal block from LDS, writes its
// rows and raises its flag. Up to 8 right-hand sides per launch; all
// arithmetic in FP64 (FP32 factors are widened on load).
constexpr int kChainTB = 32;

template <class S, class V, bool UPPER>
__global__ __launch_bounds__(256) void trsv_chain_kernel(const S *M, size_t lda, V *X, int n, int nrhs,
                                                         int *flags, int epoch, int *bad) {
    using T = typename VecT<V>::A;
    constexpr int TB = kChainTB;
    __shared__ T dblk[TB][TB];
    __shared__ T z[8][TB];
    __shared__ T red[8][8][TB];
    const int nblk = (n + TB - 1) / TB;
    const int b = UPPER ? nblk - 1 - (int)blockIdx.x : (int)blockIdx.x;
    const int r0 = b * TB, rows = min(TB, n - r0);
    const int tid = threadIdx.x;
    const int r = tid % TB, cg = tid / TB;  // row in block, column group (8 groups of 4 columns)
    T acc[8];
    for (int q = 0; q < 8; ++q) acc[q] = ezero<T>();
    // diagonal block into LDS
    for (int idx = tid; idx < TB * TB; idx += 256) {
        const int i = idx / TB, j = idx % TB;
        dblk[i][j] = (i < rows && j < rows) ? ld(M + (size_t)(r0 + i) * lda + r0 + j) : ezero<T>();
    }
    // off-diagonal contributions, in the order the blocks become final; each
    // round takes every dependency already final (at least one)
    __shared__ int s_ready, s_abort;
    const int ndeps = UPPER ? nblk - 1 - b : b;
    int d = 0;
    while (d < ndeps) {
        if (tid == 0) {
            int ready = d, abort = 0;
            unsigned spins = 0;
            const int jb0 = UPPER ? nblk - 1 - d : d;
            while (__hip_atomic_load(&flags[jb0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > kSpinLimit) {
                    atomicOr(bad, 2);
                    abort = 1;
                    break;
                }
            }
            if (!abort) {
                ready = d + 1;
                while (ready < ndeps && ready < d + 16) {
                    const int jb = UPPER ? nblk - 1 - ready : ready;
                    if (__hip_atomic_load(&flags[jb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) break;
                    ++ready;
                }
            }
            s_ready = ready;
            s_abort = abort;
        }
        __syncthreads();
        if (s_abort) return;
        const int upto = s_ready;
        for (; d < upto; ++d) {
            const int jb = UPPER ? nblk - 1 - d : d;
            const int c0 = jb * TB, cols = min(TB, n - c0);
            if (r < rows) {
                const S *mrow = M + (size_t)(r0 + r) * lda + c0;
                for (int c = cg * 4; c < cg * 4 + 4 && c < cols; ++c) {
                    const T m = ld(mrow + c);
                    for (int q = 0; q < nrhs; ++q) {
                        T xv;
                        sub_elem(reinterpret_cast<const double *>(X + (size_t)q * n + c0 + c), xv);
                        acc[q] = eadd(acc[q], emul(m, xv));
                    }
                }
            }
        }
        __syncthreads();
    }
    // reduce the 8 column groups; z = x_b - sum
    for (int q = 0; q < nrhs; ++q) red[q][cg][r] = acc[q];
    __syncthreads();
    if (tid < nrhs * TB) {
        const int q = tid / TB, i = tid % TB;
        T sum = ezero<T>();
        for (int g2 = 0; g2 < 8; ++g2) sum = eadd(sum, red[q][g2][i]);
        z[q][i] = i < rows ? esub(X[(size_t)q * n + r0 + i], sum) : ezero<T>();
    }
    __syncthreads();
    // diagonal block: unit lower (forward) or upper with division (backward)
    for (int step = 0; step < rows; ++step) {
        const int i = UPPER ? rows - 1 - step : step;
        if (UPPER && tid < nrhs) {
            const T d = dblk[i][i];
            if constexpr (Elem<V>::cplx) {
                const double den = d.x * d.x + d.y * d.y;
                if (!isfinite(den) || den == 0.0) atomicOr(bad, 1);
            } else {
                if (d == 0.0 || !isfinite(d)) atomicOr(bad, 1);
            }
            z[tid][i] = ediv(z[tid][i], d);
        }
        __syncthreads();
        if (tid < nrhs * TB) {
            const int q = tid / TB, k = tid % TB;
            if (UPPER ? (k < i) : (k > i && k < rows)) z[q][k] = esub(z[q][k], emul(dblk[k][i], z[q][i]));
        }
        __syncthreads();
    }
    // publish the solved rows with cache-bypassing stores, wait for them to
    // complete, then raise the flag (no fences needed)
    if (tid < nrhs * TB) {
        const int q = tid / TB, i = tid % TB;
        if (i < rows) pub_elem(reinterpret_cast<double *>(X + (size_t)q * n + r0 + i), z[q][i]);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0) __hip_atomic_store(&flags[b], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// one triangular solve through the chain kernel; false if the device cannot
// keep every block resident (caller falls back to the blocked path)
template <class S, class V, bool UPPER>
bool chain_solve(const S *dLU, int n, size_t lda, V *X, int nrhs, int *flags, int epoch, int *bad,
                 const LUContext &ctx) {
    const int nblk = (n + kChainTB - 1) / kChainTB;
    int per_cu = 0;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void *>(trsv_chain_kernel<S, V, UPPER>), 256, 0));
    if ((long long)per_cu * ctx.cus < nblk) return false;
    for (int q0 = 0; q0 < nrhs; q0 += 8) {
        V *Xq = X + (size_t)q0 * n;
        int nq = std::min(8, nrhs - q0), nn = n, ep = epoch + q0;
        size_t ld = lda;
        void *args[] = {(void *)&dLU, &ld, &Xq, &nn, &nq, &flags, &ep, &bad};
        HIP_CHECK(hipLaunchCooperativeKernel(reinterpret_cast<void *>(trsv_chain_kernel<S, V, UPPER>), dim3(nblk),
                                             dim3(256), args, 0, ctx.stream));
    }
    return true;
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
    lu_solve_permuted(dLU, n, lda, d_perm.ptr, d_row, d_col, dX, nrhs, ctx);
}

template <class S, class V>
void lu_solve_permuted(const S *dLU, int n, size_t lda, const int *d_perm, const double *d_row,
                       const double *d_col, V *dX, int nrhs, LUContext &ctx) {
    hipStream_t s = ctx.stream;
    const size_t total = (size_t)n * nrhs;
    if (ctx.solve_tmp.count < total * sizeof(V) / sizeof(double) + 1) ctx.solve_tmp.resize(total * sizeof(V) / sizeof(double) + 1);
    V *tmp = reinterpret_cast<V *>(ctx.solve_tmp.ptr);
    hipLaunchKernelGGL(permute_scale_kernel<V>, dim3((total + 255) / 256), dim3(256), 0, s, tmp, dX, d_perm, d_row, n,
                       nrhs);
    const int nblk = (n + kChainTB - 1) / kChainTB;
    if ((int)ctx.flags.count < 2 * nblk) {
        ctx.flags.resize(2 * (size_t)nblk);
        hipLaunchKernelGGL(fill_int_kernel, dim3((2 * nblk + 255) / 256), dim3(256), 0, s, ctx.flags.ptr, 2 * nblk, 0);
        ctx.epoch = 0;
    }
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, s, ctx.solve_bad.ptr, 1, 0);
    const int ep = ctx.epoch + 1;
    ctx.epoch += 8 * ((nrhs + 7) / 8) + 1;
    bool chained = chain_solve<S, V, false>(dLU, n, lda, tmp, nrhs, ctx.flags.ptr, ep, ctx.solve_bad.ptr, ctx) &&
                   chain_solve<S, V, true>(dLU, n, lda, tmp, nrhs, ctx.flags.ptr + nblk, ep, ctx.solve_bad.ptr, ctx);
    if (!chained) {
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

template void lu_solve<double2, double2>(const double2 *, int, size_t, const int *, const double *, const double *,
                                         double2 *, int, LUContext &);
template void lu_solve<float2, double2>(const float2 *, int, size_t, const int *, const double *, const double *,
                                        double2 *, int, LUContext &);
template void lu_solve<double, double>(const double *, int, size_t, const int *, const double *, const double *,
                                       double *, int, LUContext &);
template void lu_solve_permuted<double2, double2>(const double2 *, int, size_t, const int *, const double *,
                                                  const double *, double2 *, int, LUContext &);
template void lu_solve_permuted<float2, double2>(const float2 *, int, size_t, const int *, const double *,
                                                 const double *, double2 *, int, LUContext &);
template void lu_solve_permuted<double, double>(const double *, int, size_t, const int *, const double *,
                                                const double *, double *, int, LUContext &);

void permutation_from_pivots(const int *d_ipiv, int n, int *d_perm, hipStream_t s) {
    std::vector<int> ipiv(n), perm(n);
    HIP_CHECK(hipMemcpyAsync(ipiv.data(), d_ipiv, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < n; ++k) std::swap(perm[k], perm[ipiv[k]]);
    HIP_CHECK(hipMemcpyAsync(d_perm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
}

__global__ void to_f32_kernel(const double2 *src, float2 *dst, size_t count) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < count) dst[i] = make_float2((float)src[i].x, (float)src[i].y);
}

void convert_to_f32(const double2 *src, float2 *dst, size_t count, hipStream_t s) {
    hipLaunchKernelGGL(to_f32_kernel, dim3((count + 255) / 256), dim3(256), 0, s, src, dst, count);
    HIP_CHECK(hipGetLastError());
}

// residual and backward-error terms (harmonic.c:899-969), one block per row
__global__ __launch_bounds__(256) void residual_kernel(const double2 *A, size_t lda, int n, const double2 *x,
                                                      const double *ax, const double2 *b, double2 *r,
                                                      double *terms) {
    __shared__ double sre[256], sim[256], sden[256];
    const int i = blockIdx.x;
    double yr = 0.0, yi = 0.0, den = 0.0;
    const double2 *row = A + (size_t)i * lda;
    for (int j = threadIdx.x; j < n; j += 256) {
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

void complex_residual(const double2 *dA, int n, size_t lda, const double2 *dx, const double2 *db, double2 *dr,
                      double *dterms, hipStream_t s) {
    DeviceBuffer<double> ax(n);
    hipLaunchKernelGGL(abs_kernel, dim3((n + 255) / 256), dim3(256), 0, s, dx, ax.ptr, n);
    hipLaunchKernelGGL(residual_kernel, dim3(n), dim3(256), 0, s, dA, lda, n, dx, ax.ptr, db, dr, dterms);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));
}

void mfma_layout_probe(double *host_out) {
    DeviceBuffer<double> d(1024);
    hipLaunchKernelGGL(mfma_layout_probe_kernel, dim3(1), dim3(64), 0, 0, d.ptr);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(host_out, d.ptr, 1024 * sizeof(double), hipMemcpyDeviceToHost));
}

}  // namespace peec

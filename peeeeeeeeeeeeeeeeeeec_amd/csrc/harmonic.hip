// MNA system kernels (harmonic.c:175-552) and the device-resident
// frequency-point engine declared in include/peec_gpu.h.
#include <hip/hip_runtime.h>

#include <chrono>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/peec_gpu.h"
#include "assembly.h"
#include "harmonic_kernels.h"
#include "lu.h"
#include "mna.h"
#include "rcs_kernels.h"

namespace peec {

constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kC0 = 299792458.0;  // main.c:47-54

// Z = [[R + jwL, A], [-P A^T, jw I]] written row-major with leading
// dimension ldz; A is implicit in the edge end points (incidence.c:34-115).
// When Lw/Pw are given (retarded form) they replace L/P.
// With rowmax, each workgroup also folds max |z| of its part of the row into
// rowmax[row] (zero beforehand; a non-finite entry counts as infinity), the
// row maxima of the equilibration without another pass over Z0.
__global__ __launch_bounds__(256) void mna_fill_kernel(double2 *Z, double2 *Z0, size_t ldz, int ne, int nv,
                                                       const double *L, const double2 *Lw, const double *R,
                                                       const double *P, const double2 *Pw, const double *PAT,
                                                       const int *edges, double omega, int chunks, double *rowmax) {
    const int n = ne + nv;
    const int row = blockIdx.x / chunks;
    MnaOperands mo;
    mo.ne = ne;
    mo.nv = nv;
    mo.L = L;
    mo.Lw = Lw;
    mo.R = R;
    mo.P = P;
    mo.Pw = Pw;
    mo.PAT = PAT;
    mo.edges = edges;
    mo.omega = omega;
    // four entries per thread, 1024 columns per workgroup
    const int col0 = (blockIdx.x % chunks) * 1024 + threadIdx.x;
    double2 z[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int col = col0 + 256 * u;
        z[u] = col < n ? mna_entry(row, col, mo) : double2{};
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int col = col0 + 256 * u;
        if (col < n) {
            if (Z) Z[(size_t)row * ldz + col] = z[u];
            if (Z0) Z0[(size_t)row * ldz + col] = z[u];
        }
    }
    if (!rowmax) return;
    __shared__ double red[4];
    double m = 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (col0 + 256 * u < n) {
            const double v = hypot(z[u].x, z[u].y);
            m = fmax(m, isfinite(v) ? v : INFINITY);
        }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        // non-negative doubles order like their bit patterns
        atomicMax(reinterpret_cast<unsigned long long *>(rowmax + row), (unsigned long long)__double_as_longlong(m));
    }
}

void launch_mna_fill(double2 *Z, double2 *Z0, size_t ldz, int ne, int nv, const double *L, const double2 *Lw,
                     const double *R, const double *P, const double2 *Pw, const double *PAT, const int *edges,
                     double omega, hipStream_t s, double *rowmax) {
    const int n = ne + nv;
    const int chunks = (n + 1023) / 1024;
    hipLaunchKernelGGL(mna_fill_kernel, dim3((unsigned)((size_t)n * chunks)), dim3(256), 0, s, Z, Z0, ldz, ne, nv,
                       L, Lw, R, P, Pw, PAT, edges, omega, chunks, rowmax);
    HIP_CHECK(hipGetLastError());
}

// PAT[j, e] = P[j, b_e] - P[j, a_e] (incidence.c:488-560)
__global__ void pat_kernel(const double *P, const int *edges, double *PAT, int nv, int ne) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)nv * ne) return;
    const int j = (int)(idx / ne), e = (int)(idx % ne);
    PAT[idx] = P[(size_t)j * nv + edges[2 * e + 1]] - P[(size_t)j * nv + edges[2 * e]];
}

void launch_pat(const double *P, const int *edges, double *PAT, int nv, int ne, hipStream_t s) {
    const size_t total = (size_t)nv * ne;
    hipLaunchKernelGGL(pat_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, P, edges, PAT, nv, ne);
    HIP_CHECK(hipGetLastError());
}

// rhs_k = [0 ; P(w) I_s,k] with I_s,k = +1 at a_k, -1 at b_k
__global__ void port_rhs_kernel(double2 *B, int ne, int nv, const double *P, const double2 *Pw, const int *ports,
                                int nports) {
    const int n = ne + nv;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)n * nports) return;
    const int k = (int)(idx / n), i = (int)(idx % n);
    double2 v = make_double2(0.0, 0.0);
    if (i >= ne) {
        const int row = i - ne, a = ports[2 * k], b = ports[2 * k + 1];
        if (Pw) {
            const double2 pa = Pw[(size_t)row * nv + a], pb = Pw[(size_t)row * nv + b];
            v = make_double2(pa.x - pb.x, pa.y - pb.y);
        } else {
            v = make_double2(P[(size_t)row * nv + a] - P[(size_t)row * nv + b], 0.0);
        }
    }
    B[idx] = v;
}

// general rhs = [-U_e ; P I_s] (harmonic.c:436-552), one block per node row
__global__ void general_rhs_kernel(double2 *B, int ne, int nv, const double *P, const double2 *U,
                                   const double2 *Is) {
    __shared__ double sr[256], si[256];
    const int row = blockIdx.x;
    if (row < ne) {
        if (threadIdx.x == 0) B[row] = U ? make_double2(-U[row].x, -U[row].y) : make_double2(0.0, 0.0);
        return;
    }
    const int i = row - ne;
    double a = 0.0, b = 0.0;
    if (Is)
        for (int j = threadIdx.x; j < nv; j += 256) {
            const double p = P[(size_t)i * nv + j];
            a += p * Is[j].x;
            b += p * Is[j].y;
        }
    sr[threadIdx.x] = a;
    si[threadIdx.x] = b;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if (threadIdx.x < s2) {
            sr[threadIdx.x] += sr[threadIdx.x + s2];
            si[threadIdx.x] += si[threadIdx.x + s2];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) B[row] = make_double2(sr[0], si[0]);
}

void launch_general_rhs(double2 *B, int ne, int nv, const double *P, const double2 *U, const double2 *Is,
                        hipStream_t s) {
    hipLaunchKernelGGL(general_rhs_kernel, dim3(ne + nv), dim3(256), 0, s, B, ne, nv, P, U, Is);
    HIP_CHECK(hipGetLastError());
}

__global__ void axpy_sub_kernel(double2 *X, const double2 *D, size_t count) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < count) {
        X[i].x -= D[i].x;
        X[i].y -= D[i].y;
    }
}

// reference's seven residual figures from per-row terms (harmonic.c:899-1060)
void residual_summary(const std::vector<double> &terms, size_t n, size_t ne, double out[7]) {
    double absr = 0, binf = 0, be = 0, eabs = 0, nabs = 0, ebe = 0, nbe = 0;
    for (size_t i = 0; i < n; ++i) {
        const double ra = terms[3 * i], den = terms[3 * i + 1], ba = terms[3 * i + 2];
        const double eta = den > 0.0 ? ra / den : ra;
        absr = std::fmax(absr, ra);
        binf = std::fmax(binf, ba);
        be = std::fmax(be, eta);
        if (i < ne) {
            eabs = std::fmax(eabs, ra);
            ebe = std::fmax(ebe, eta);
        } else {
            nabs = std::fmax(nabs, ra);
            nbe = std::fmax(nbe, eta);
        }
    }
    out[0] = absr;
    out[1] = binf > 0.0 ? absr / binf : absr;
    out[2] = be;
    out[3] = eabs;
    out[4] = nabs;
    out[5] = ebe;
    out[6] = nbe;
}

}  // namespace peec

using namespace peec;

// ---------------------------------------------------------------------------
// engine

// Autotuner of the ordered factor's shape: block width (128 or 64) and the
// rows of the split-FP16 update's tiles (64 or 32; with the FP32 update only
// the width varies). Each candidate is tried on one real factor, timed per
// member of its launches, and the fastest is kept; single factors and
// batches are tuned apart (they fill the device differently).
struct FactorShape {
    int nb, rows;
};
static constexpr FactorShape kShapes[4] = {{128, 64}, {128, 32}, {64, 64}, {64, 32}};
struct ShapeTuner {
    int tried = 0, best = -1;
    double seconds[4] = {0.0, 0.0, 0.0, 0.0};
    int candidates(bool split16) const { return split16 ? 4 : 2; }
    // the candidate to try next (split16 off: 128 and 64 with 64-row tiles)
    FactorShape trial(bool split16) const { return split16 ? kShapes[tried] : kShapes[2 * tried]; }
    void record(double s, bool split16) {
        seconds[tried++] = s;
        if (tried == candidates(split16)) {
            best = 0;
            for (int i = 1; i < tried; ++i)
                if (seconds[i] < seconds[best]) best = i;
        }
    }
    FactorShape chosen(bool split16) const { return split16 ? kShapes[best] : kShapes[2 * best]; }
};

struct PeecEngine {
    PeecEngineOptions opt{};
    int ne = 0, nv = 0, n = 0;
    Stream stream;
    std::unique_ptr<RegionSet> nodes, edges;
    std::unique_ptr<LUContext> lu;
    DeviceBuffer<int> d_edges, d_ports, d_ipiv, d_perm, d_err;
    DeviceBuffer<int> d_reuse_perm;  // row order of the last partial-pivoting factor
    bool have_reuse = false;
    DeviceBuffer<double> d_L, d_P, d_row, d_col, d_terms;
    DeviceBuffer<double2> d_Lw, d_Pw, d_Z, d_Z0, d_B, d_X, d_R, d_Xbest;
    DeviceBuffer<double> d_ax;                  // |x| scratch of the residuals
    DeviceBuffer<unsigned long long> d_eta;     // per right-hand side backward error (bits)
    DeviceBuffer<int> d_flags;                  // per right-hand side masks
    std::unique_ptr<RcsGeometry> rcs;           // edge geometry for excitation and far field
    double limit = 0.0;                         // pivot limit of the current point
    bool ordered = false;                       // current FP32 factor uses the reused row order
    bool split16_pivoted = false;               // ... or is a partial-pivoting one with split-FP16 updates
    bool retarded_now = false;
    // the factor held for (cached_f, cached_mode) may serve further
    // right-hand sides at that point without a new assembly or factor
    bool z_equilibrated = false;  // d_Z holds this point's equilibrated FP64 matrix
    // the MNA matrix of the current point as a function of its blocks; its
    // dense copy d_Z0 is written only when a pass needs it (the FP64
    // fallback, residuals of many right-hand sides)
    MnaOperands mna;
    bool z0_ready = false;
    bool shift_ok = true;         // shifted-center moments may still be built
    bool cache_ok = false;
    double cached_f = 0.0;
    int cached_mode = -1;
    // shape autotuners of the ordered factor (block_size = 0): single
    // factors and batched ones
    ShapeTuner tune_single, tune_batch;
    DeviceBuffer<float2> d_Z32;  // FP32 factors: one, or a batch side by side
    float2 *z32 = nullptr;       // the current point's FP32 factor (in d_Z32)
    // mixed points factored together (peec_engine_solve_batch): each one's
    // retarded blocks and equilibration scales; the engine's own buffers are
    // swapped with a point's while that point is prepared or solved
    struct PointSlot {
        DeviceBuffer<double2> Lw, Pw;
        DeviceBuffer<double> row, col;
        MnaOperands mna;
        double limit = 0.0;
    };
    std::vector<PointSlot> slots;
    DeviceBuffer<double> d_limits;
    DeviceBuffer<int> d_batch_perm;  // the row order a batch was factored in
    DeviceBuffer<double> d_mrow, d_mcol;  // a batch's equilibration scales side by side
    // GMRES-IR workspace: Krylov basis, work vector, zero right-hand side,
    // partial dot products, projections / solution coefficients, |x|, eta
    DeviceBuffer<double2> d_gV, d_gw, d_gzero, d_gpart, d_gh;
    DeviceBuffer<double> d_gax;
    DeviceBuffer<unsigned long long> d_geta;
    hipEvent_t ev[6] = {};
    long long far_evals = 0, smooth_evals = 0;
    PeecStaticStats static_stats{};
    KernelTimer timer;
    ~PeecEngine() {
        for (hipEvent_t &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

// ---------------------------------------------------------------------------
// one frequency point: factor, then solves with refinement

__global__ void fill_zero2_kernel(double2 *p, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = make_double2(0.0, 0.0);
}
static void fill_zero(double2 *p, size_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(fill_zero2_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
}

// X[q] -= D[q] (or dst[q] = src[q] for the copy) for the right-hand sides
// whose flag is set: flags[q] in device memory, or with flags null bit q of
// mask (up to 64 right-hand sides, no upload)
__device__ inline bool rhs_on(const int *flags, unsigned long long mask, size_t q) {
    return flags ? flags[q] != 0 : ((mask >> q) & 1ull) != 0;
}
__global__ void masked_axpy_kernel(double2 *X, const double2 *D, const int *flags, unsigned long long mask, int n,
                                   int nrhs) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)n * nrhs || !rhs_on(flags, mask, i / n)) return;
    X[i].x -= D[i].x;
    X[i].y -= D[i].y;
}
__global__ void masked_copy_kernel(double2 *dst, const double2 *src, const int *flags, unsigned long long mask, int n,
                                   int nrhs) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)n * nrhs || !rhs_on(flags, mask, i / n)) return;
    dst[i] = src[i];
}

static int assemble_retarded(PeecEngine *e, double k, double *center = nullptr);

// ---------------------------------------------------------------------------
// GMRES-IR (refinement whose correction solves A d = r by GMRES on the
// system preconditioned with the FP32 factor, Carson & Higham): complex
// FP64 vector kernels over n entries

constexpr int kGmresDotBlocks = 128;
constexpr int kGmresOuter = 10;  // outer GMRES-IR steps at most
// partial[i * gridDim.x + b] = sum over the entries of block b of conj(V_i) w
__global__ __launch_bounds__(256) void gmres_dots_kernel(const double2 *V, size_t ldv, const double2 *w, int n,
                                                         double2 *partial) {
    __shared__ double sr[256], si[256];
    const int i = blockIdx.y, tid = threadIdx.x;
    const double2 *v = V + (size_t)i * ldv;
    double ar = 0.0, ai = 0.0;
    for (int t = blockIdx.x * 256 + tid; t < n; t += gridDim.x * 256) {
        const double2 a = v[t], b = w[t];
        ar += a.x * b.x + a.y * b.y;
        ai += a.x * b.y - a.y * b.x;
    }
    sr[tid] = ar;
    si[tid] = ai;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) {
            sr[tid] += sr[tid + h];
            si[tid] += si[tid + h];
        }
        __syncthreads();
    }
    if (tid == 0) partial[(size_t)i * gridDim.x + blockIdx.x] = make_double2(sr[0], si[0]);
}
// w -= sum_{i < k} h_i V_i
__global__ void gmres_axpy_kernel(const double2 *V, size_t ldv, const double2 *h, int k, double2 *w, int n) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    double2 acc = w[t];
    for (int i = 0; i < k; ++i) {
        const double2 a = h[i], v = V[(size_t)i * ldv + t];
        acc.x -= a.x * v.x - a.y * v.y;
        acc.y -= a.x * v.y + a.y * v.x;
    }
    w[t] = acc;
}
// y = a x
__global__ void gmres_scale_kernel(const double2 *x, double2 a, double2 *y, int n) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const double2 v = x[t];
    y[t] = make_double2(a.x * v.x - a.y * v.y, a.x * v.y + a.y * v.x);
}

// the dense copy of the current point's MNA matrix in d_Z0
static void ensure_z0(PeecEngine *e) {
    if (e->z0_ready) return;
    const MnaOperands &m = e->mna;
    const int n = m.ne + m.nv;
    e->d_Z0.resize((size_t)n * n);
    launch_mna_fill(nullptr, e->d_Z0.ptr, n, m.ne, m.nv, m.L, m.Lw, m.R, m.P, m.Pw, m.PAT, m.edges, m.omega, e->stream);
    e->z0_ready = true;
}

static int factor_error(const LUStatus &st) {
    if (st.code == -2) PEEC_ERR("complex LU singular at column %lld.", st.column);
    else PEEC_ERR("LU failed (code %d).", st.code);
    return st.code == -2 ? -2 : -1;
}

// Partial-pivoting factor (FP64 in d_Z, or FP32 in d_Z32 from the
// equilibrated d_Z); its row order becomes the order reused by the next
// mixed-precision point. The FP32 factor's trailing updates run on
// split-FP16 MFMA (with split_fp16_update) unless `accurate` asks for the
// FP32 MFMA update: the split keeps 22 bits of each operand relative to its
// row's largest part, FP32 24 bits of every entry, which is what a refactor
// after a stalled refinement needs at badly scaled (low-frequency) points.
static int factor_pivoted(PeecEngine *e, bool fp64, bool accurate = false) {
    hipStream_t s = e->stream;
    const int n = e->n;
    LUStatus st;
    // the held factor (if any) is overwritten here: a caller that keeps this
    // one marks it cached again afterwards
    e->cache_ok = false;
    if (fp64) {
        if (!e->z_equilibrated) {
            // mixed point falling back to FP64: form the equilibrated matrix
            ensure_z0(e);
            e->d_Z.resize((size_t)n * n);
            copy_equilibrated(e->d_Z0.ptr, e->d_Z.ptr, e->d_row.ptr, e->d_col.ptr, n, s);
            e->z_equilibrated = true;
        }
        st = lu_factor<double2>(e->d_Z.ptr, n, n, e->d_ipiv.ptr, e->limit, *e->lu, e->opt.block_size);
    } else {
        double max_scaled = 0.0;
        if (convert_equilibrated_to_f32(e->mna, e->z32, nullptr, e->d_row.ptr, e->d_col.ptr, &max_scaled, *e->lu) !=
            0)
            return -2;
        e->limit = 1.0e-14 * max_scaled;
        e->lu->pivoted_split16 = e->lu->split_fp16_update && !accurate;
        st = lu_factor<float2>(e->z32, n, n, e->d_ipiv.ptr, e->limit, *e->lu, e->opt.block_size);
        e->split16_pivoted = e->lu->pivoted_split16;
    }
    if (st.code != 0) return factor_error(st);
    permutation_from_pivots(e->d_ipiv.ptr, n, e->d_perm.ptr, s);
    HIP_CHECK(hipMemcpyAsync(e->d_reuse_perm.ptr, e->d_perm.ptr, sizeof(int) * n, hipMemcpyDeviceToDevice, s));
    e->have_reuse = true;
    e->ordered = false;
    return 0;
}

// Retarded assembly, MNA fill, equilibration and the factor (FP64, or FP32
// in the previous point's row order / with partial pivoting). 0, -1 or -2.
// Retarded blocks, MNA operands and equilibration of point f in the
// engine's current buffers: the scales only (mixed mode: the FP32 matrix is
// formed from the blocks later), or the equilibrated FP64 matrix in d_Z.
// Events ev[1] / ev[2] mark the end of the assembly / of the equilibration.
// assembled_terms > 0: the retarded blocks of this point are already in
// d_Lw / d_Pw, from that many stored moments (a batch evaluates its members'
// series together).
static int prepare_point(PeecEngine *e, double f, bool mixed, PeecPointReport &report, int assembled_terms = 0) {
    hipStream_t s = e->stream;
    const int n = e->n, ne = e->ne, nv = e->nv;
    const size_t nn = (size_t)n * n;
    const double omega = kTwoPi * f;
    const bool ret = e->opt.retarded != 0;
    e->retarded_now = ret;
    if (ret && assembled_terms > 0) {
        report.moment_terms = assembled_terms;
        report.moment_center = 0.0;
    } else if (ret) {
        report.moment_terms = assemble_retarded(e, omega / kC0, &report.moment_center);
    }
    HIP_CHECK(hipEventRecord(e->ev[1], s));
    MnaOperands &m = e->mna;
    m = MnaOperands{};
    m.ne = ne;
    m.nv = nv;
    m.L = e->d_L.ptr;
    m.Lw = ret ? e->d_Lw.ptr : nullptr;
    m.P = e->d_P.ptr;
    m.Pw = ret ? e->d_Pw.ptr : nullptr;
    m.edges = e->d_edges.ptr;
    m.omega = omega;
    e->z0_ready = false;
    e->d_row.resize(n);
    e->d_col.resize(n);
    e->z_equilibrated = false;
    if (mixed) {
        // the FP32 factor is formed from the MNA blocks with the
        // equilibration's scales (read-only passes that evaluate the
        // entries); the equilibrated FP64 matrix is only built if the point
        // falls back to FP64
        if (equilibration_scales(m, e->d_row.ptr, e->d_col.ptr, *e->lu) != 0) return -2;
    } else {
        e->d_Z.resize(nn);
        launch_mna_fill(e->d_Z.ptr, nullptr, n, ne, nv, m.L, m.Lw, nullptr, m.P, m.Pw, nullptr, m.edges, omega, s);
        double max_scaled = 0.0;
        if (equilibrate_complex(e->d_Z.ptr, n, n, e->d_row.ptr, e->d_col.ptr, &max_scaled, *e->lu) != 0) return -2;
        e->z_equilibrated = true;
        e->limit = 1.0e-14 * max_scaled;
    }
    HIP_CHECK(hipEventRecord(e->ev[2], s));
    return 0;
}

static int engine_factor(PeecEngine *e, double f, int mode, PeecPointReport &report) {
    hipStream_t s = e->stream;
    const int n = e->n;
    const size_t nn = (size_t)n * n;
    HIP_CHECK(hipEventRecord(e->ev[0], s));
    if (e->cache_ok && f == e->cached_f && mode == e->cached_mode) {
        // the factor of this point is still held: nothing to assemble,
        // fill or factor
        for (int i = 1; i <= 3; ++i) HIP_CHECK(hipEventRecord(e->ev[i], s));
        report.pivot_reuse = e->ordered ? 1 : 0;
        report.factor_cached = 1;
        return 0;
    }
    e->cache_ok = false;
    const bool mixed = mode == PEEC_SOLVE_MIXED;
    const int prep = prepare_point(e, f, mixed, report);
    if (prep) return prep;
    e->d_ipiv.resize(n);
    e->d_perm.resize(n);
    e->d_reuse_perm.resize(n);
    LUStatus st;
    e->ordered = false;
    if (mixed) {
        e->d_Z32.reserve(nn);
        e->z32 = e->d_Z32.ptr;
        if (e->opt.reuse_pivots && e->have_reuse) {
            double max_scaled = 0.0;
            if (convert_equilibrated_to_f32(e->mna, e->z32, e->d_reuse_perm.ptr, e->d_row.ptr, e->d_col.ptr,
                                            &max_scaled, *e->lu) != 0)
                return -2;
            e->limit = 1.0e-14 * max_scaled;
            const bool split16 = e->lu->split_fp16_update;
            ShapeTuner &tu = e->tune_single;
            const bool trial = e->opt.block_size == 0 && tu.best < 0;
            FactorShape shape = e->opt.block_size > 0 ? FactorShape{e->opt.block_size, 64}
                                : trial                ? tu.trial(split16)
                                                       : tu.chosen(split16);
            e->lu->h3_tile_rows = shape.rows;
            if (trial) e->stream.sync();
            const auto t0 = std::chrono::steady_clock::now();
            st = lu_factor_ordered(e->z32, n, n, e->limit, *e->lu, shape.nb);
            report.block_size = shape.nb;
            report.update_tile_rows = shape.rows;
            if (trial && st.code == 0)  // lu_factor_ordered returns after its stream has drained
                tu.record(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(), split16);
            if (st.code == 0) {
                e->ordered = true;
                HIP_CHECK(hipMemcpyAsync(e->d_perm.ptr, e->d_reuse_perm.ptr, sizeof(int) * n, hipMemcpyDeviceToDevice,
                                         s));
            } else if (st.code != -2) {
                return factor_error(st);
            }
        }
        if (!e->ordered) {
            const int code = factor_pivoted(e, false);
            if (code) return code;
        }
    } else {
        const int code = factor_pivoted(e, true);
        if (code) return code;
    }
    HIP_CHECK(hipEventRecord(e->ev[3], s));
    report.pivot_reuse = e->ordered ? 1 : 0;
    return 0;
}

// Solve the nrhs right-hand sides in d_B into d_X: FP64 directly, or the
// FP32 factor with refinement against the FP64 matrix (x <- x - A^{-1} r)
// and the fallbacks (a rejected reused order is refactored with pivoting; a
// point that stays above the gate is refactored in FP64). The component-wise
// backward error need not fall at every step (the first correction often
// barely changes it at low frequencies), so a right-hand side stops after
// two steps without a new best, and its best iterate is kept.
// 0, -1, -2, or -3 (backward error above the gate).
static int engine_solve(PeecEngine *e, int mode, int nrhs, PeecPointReport &report) {
    hipStream_t s = e->stream;
    const int n = e->n;
    const size_t tot = (size_t)n * nrhs;
    e->d_X.resize(tot);
    e->d_R.resize(tot);
    e->d_Xbest.resize(tot);
    e->d_ax.resize(tot);
    e->d_eta.resize((size_t)nrhs);
    e->d_flags.resize((size_t)nrhs);
    std::vector<double> eta(nrhs);
    std::vector<unsigned long long> bits(nrhs);
    auto measure = [&]() {
        if (nrhs < kResidualDenseMinRhs) {
            residual_multi(e->mna, e->d_X.ptr, e->d_B.ptr, e->d_R.ptr, nrhs, e->d_ax.ptr, e->d_eta.ptr, s);
        } else {
            ensure_z0(e);
            residual_multi(e->d_Z0.ptr, n, n, e->d_X.ptr, e->d_B.ptr, e->d_R.ptr, nrhs, e->d_ax.ptr, e->d_eta.ptr, s);
        }
        e->d_eta.download(bits.data(), nrhs, s);
        e->stream.sync();
        for (int q = 0; q < nrhs; ++q) {
            double v;
            std::memcpy(&v, &bits[q], sizeof v);
            // a NaN or infinite residual (overflowing or NaN iterate or
            // right-hand side) is a failure: +inf keeps every later
            // comparison and maximum treating it as one
            eta[q] = std::isfinite(v) ? v : INFINITY;
        }
    };
    const unsigned grid = (unsigned)((tot + 255) / 256);
    // the per-right-hand-side flags as a kernel argument when they fit,
    // else uploaded
    auto flag_args = [&](const std::vector<int> &fl, const int *&dflags, unsigned long long &mask) {
        dflags = nullptr;
        mask = 0ull;
        if (nrhs <= 64) {
            for (int q = 0; q < nrhs; ++q)
                if (fl[q]) mask |= 1ull << q;
        } else {
            e->d_flags.upload(fl.data(), nrhs, s);
            dflags = e->d_flags.ptr;
        }
    };
    auto refine = [&]() {
        HIP_CHECK(hipMemcpyAsync(e->d_X.ptr, e->d_B.ptr, sizeof(double2) * tot, hipMemcpyDeviceToDevice, s));
        lu_solve_permuted<float2, double2>(e->z32, n, n, e->d_perm.ptr, e->d_row.ptr, e->d_col.ptr, e->d_X.ptr, nrhs,
                                           *e->lu);
        measure();
        std::vector<double> best = eta;
        std::vector<int> since(nrhs, 0), active(nrhs), flags(nrhs);
        for (int q = 0; q < nrhs; ++q) active[q] = eta[q] > e->opt.refinement_target;
        HIP_CHECK(hipMemcpyAsync(e->d_Xbest.ptr, e->d_X.ptr, sizeof(double2) * tot, hipMemcpyDeviceToDevice, s));
        while (report.refinement_steps < e->opt.max_refinement_steps) {
            bool any = false;
            for (int q = 0; q < nrhs; ++q) any = any || active[q];
            if (!any) break;
            lu_solve_permuted<float2, double2>(e->z32, n, n, e->d_perm.ptr, e->d_row.ptr, e->d_col.ptr, e->d_R.ptr,
                                               nrhs, *e->lu, true);
            const int *dfl;
            unsigned long long msk;
            flag_args(active, dfl, msk);
            hipLaunchKernelGGL(masked_axpy_kernel, dim3(grid), dim3(256), 0, s, e->d_X.ptr, e->d_R.ptr, dfl, msk, n,
                               nrhs);
            ++report.refinement_steps;
            measure();
            bool improved = false;
            for (int q = 0; q < nrhs; ++q) {
                flags[q] = 0;
                if (!active[q]) continue;
                if (eta[q] < best[q]) {
                    best[q] = eta[q];
                    since[q] = 0;
                    flags[q] = 1;
                    improved = true;
                } else if (++since[q] >= 2) {
                    active[q] = 0;
                }
                if (best[q] <= e->opt.refinement_target) active[q] = 0;
            }
            if (improved) {
                flag_args(flags, dfl, msk);
                hipLaunchKernelGGL(masked_copy_kernel, dim3(grid), dim3(256), 0, s, e->d_Xbest.ptr, e->d_X.ptr, dfl,
                                   msk, n, nrhs);
            }
        }
        bool restore = false;
        double w = 0.0;
        for (int q = 0; q < nrhs; ++q) {
            flags[q] = eta[q] > best[q];
            restore = restore || flags[q];
            w = std::fmax(w, best[q]);
        }
        if (restore) {
            const int *dfl;
            unsigned long long msk;
            flag_args(flags, dfl, msk);
            hipLaunchKernelGGL(masked_copy_kernel, dim3(grid), dim3(256), 0, s, e->d_X.ptr, e->d_Xbest.ptr, dfl, msk,
                               n, nrhs);
        }
        return w;
    };
    // GMRES-IR on the current FP32 factor: per right-hand side, outer steps
    // x <- x - d with A d = r (r = A x - b in FP64) solved by restarted-free
    // GMRES of Krylov dimension m on M^-1 A d = M^-1 r (M the FP32 factor,
    // applied in FP64 arithmetic), classical Gram-Schmidt applied twice,
    // Givens rotations on the host; each right-hand side keeps its best
    // iterate and stops at the refinement target or after two outer steps
    // without a new best. Returns the worst best backward error.
    auto gmres_refine = [&]() {
        const int m = e->opt.gmres_krylov;
        const int nblk = std::min(kGmresDotBlocks, (n + 255) / 256);
        const unsigned vg = (unsigned)((n + 255) / 256);
        e->d_gV.resize((size_t)(m + 1) * n);
        e->d_gw.resize(n);
        e->d_gzero.resize(n);
        e->d_gpart.resize((size_t)nblk * (m + 1));
        e->d_gh.resize(m + 1);
        e->d_gax.resize(n);
        e->d_geta.resize(1);
        fill_zero(e->d_gzero.ptr, n, s);
        std::vector<double2> part((size_t)nblk * (m + 1));
        // conj(V_i)^T w for i < k, summed on the host in a fixed order
        auto dots = [&](const double2 *V, int k, const double2 *w, std::vector<std::complex<double>> &out) {
            hipLaunchKernelGGL(gmres_dots_kernel, dim3((unsigned)nblk, (unsigned)k), dim3(256), 0, s, V, (size_t)n, w,
                               n, e->d_gpart.ptr);
            e->d_gpart.download(part.data(), (size_t)nblk * k, s);
            e->stream.sync();
            out.assign(k, std::complex<double>(0.0, 0.0));
            for (int i = 0; i < k; ++i)
                for (int b = 0; b < nblk; ++b) out[i] += std::complex<double>(part[(size_t)i * nblk + b].x,
                                                                              part[(size_t)i * nblk + b].y);
        };
        // w <- M^-1 w
        auto precond = [&](double2 *w) {
            lu_solve_permuted<float2, double2>(e->z32, n, n, e->d_perm.ptr, e->d_row.ptr, e->d_col.ptr, w, 1, *e->lu,
                                               true);
        };
        measure();
        std::vector<double> best = eta;
        std::vector<int> since(nrhs, 0);
        std::vector<double2> hbuf(m + 1);
        HIP_CHECK(hipMemcpyAsync(e->d_Xbest.ptr, e->d_X.ptr, sizeof(double2) * tot, hipMemcpyDeviceToDevice, s));
        for (int outer = 0; outer < kGmresOuter; ++outer) {
            bool any = false;
            for (int q = 0; q < nrhs; ++q) {
                if (best[q] <= e->opt.refinement_target || since[q] >= 2) continue;
                any = true;
                double2 *r = e->d_R.ptr + (size_t)q * n, *x = e->d_X.ptr + (size_t)q * n;
                double2 *V = e->d_gV.ptr, *w = e->d_gw.ptr;
                // v_0 = M^-1 r / beta
                HIP_CHECK(hipMemcpyAsync(w, r, sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
                precond(w);
                std::vector<std::complex<double>> hv;
                dots(w, 1, w, hv);
                const double beta = std::sqrt(std::fabs(hv[0].real()));
                if (!(beta > 0.0) || !std::isfinite(beta)) continue;
                hipLaunchKernelGGL(gmres_scale_kernel, dim3(vg), dim3(256), 0, s, w, make_double2(1.0 / beta, 0.0), V,
                                   n);
                std::vector<std::complex<double>> H((size_t)(m + 1) * m), g(m + 1), cs(m), sn(m);
                g[0] = beta;
                int kdim = 0;
                for (int j = 0; j < m; ++j) {
                    // w = M^-1 A v_j (A from its blocks: the residual kernel with b = 0)
                    residual_multi(e->mna, V + (size_t)j * n, e->d_gzero.ptr, w, 1, e->d_gax.ptr, e->d_geta.ptr, s);
                    precond(w);
                    for (int pass = 0; pass < 2; ++pass) {
                        dots(V, j + 1, w, hv);
                        for (int i = 0; i <= j; ++i) {
                            H[(size_t)i * m + j] += hv[i];
                            hbuf[i] = make_double2(hv[i].real(), hv[i].imag());
                        }
                        e->d_gh.upload(hbuf.data(), j + 1, s);
                        hipLaunchKernelGGL(gmres_axpy_kernel, dim3(vg), dim3(256), 0, s, V, (size_t)n, e->d_gh.ptr,
                                           j + 1, w, n);
                    }
                    dots(w, 1, w, hv);
                    const double hn = std::sqrt(std::fabs(hv[0].real()));
                    H[(size_t)(j + 1) * m + j] = hn;
                    // previous rotations, then this column's
                    for (int i = 0; i < j; ++i) {
                        const std::complex<double> a = H[(size_t)i * m + j], b = H[(size_t)(i + 1) * m + j];
                        H[(size_t)i * m + j] = std::conj(cs[i]) * a + std::conj(sn[i]) * b;
                        H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * b;
                    }
                    const std::complex<double> a = H[(size_t)j * m + j], b = H[(size_t)(j + 1) * m + j];
                    const double den = std::sqrt(std::norm(a) + std::norm(b));
                    if (den == 0.0) break;
                    cs[j] = a / den;
                    sn[j] = b / den;
                    H[(size_t)j * m + j] = den;
                    H[(size_t)(j + 1) * m + j] = 0.0;
                    g[j + 1] = -sn[j] * g[j];
                    g[j] = std::conj(cs[j]) * g[j];
                    kdim = j + 1;
                    // the preconditioned residual is |g_{j+1}|
                    if (std::abs(g[j + 1]) <= 1e-15 * beta || !(hn > 0.0)) break;
                    hipLaunchKernelGGL(gmres_scale_kernel, dim3(vg), dim3(256), 0, s, w, make_double2(1.0 / hn, 0.0),
                                       V + (size_t)(j + 1) * n, n);
                }
                report.gmres_iterations += kdim;
                // y = H^-1 g (upper triangular), x <- x - V y
                std::vector<std::complex<double>> y(kdim);
                for (int i = kdim - 1; i >= 0; --i) {
                    std::complex<double> t = g[i];
                    for (int c = i + 1; c < kdim; ++c) t -= H[(size_t)i * m + c] * y[c];
                    y[i] = t / H[(size_t)i * m + i];
                }
                for (int i = 0; i < kdim; ++i) hbuf[i] = make_double2(y[i].real(), y[i].imag());
                e->d_gh.upload(hbuf.data(), kdim, s);
                hipLaunchKernelGGL(gmres_axpy_kernel, dim3(vg), dim3(256), 0, s, V, (size_t)n, e->d_gh.ptr, kdim, x,
                                   n);
            }
            if (!any) break;
            ++report.gmres_steps;
            measure();
            std::vector<int> flags(nrhs, 0);
            bool improved = false;
            for (int q = 0; q < nrhs; ++q) {
                if (eta[q] < best[q]) {
                    best[q] = eta[q];
                    since[q] = 0;
                    flags[q] = 1;
                    improved = true;
                } else {
                    ++since[q];
                }
            }
            if (improved) {
                const int *dfl;
                unsigned long long msk;
                flag_args(flags, dfl, msk);
                hipLaunchKernelGGL(masked_copy_kernel, dim3(grid), dim3(256), 0, s, e->d_Xbest.ptr, e->d_X.ptr, dfl,
                                   msk, n, nrhs);
            }
        }
        HIP_CHECK(hipMemcpyAsync(e->d_X.ptr, e->d_Xbest.ptr, sizeof(double2) * tot, hipMemcpyDeviceToDevice, s));
        measure();
        double w = 0.0;
        for (int q = 0; q < nrhs; ++q) w = std::fmax(w, eta[q]);
        return w;
    };
    double worst = 0.0;
    if (mode == PEEC_SOLVE_MIXED) {
        worst = refine();
        // a reused row order is kept when refinement reaches ten times its
        // target (and at least 1e-3 of the gate): beyond that it stalled
        const double keep = std::fmax(1e-3 * e->opt.backward_error_gate, 10.0 * e->opt.refinement_target);
        if (e->ordered && !(worst <= keep)) {
            // the reused row order is not good enough at this point:
            // refactor with partial pivoting and refine again
            report.pivot_refactor = 1;
            report.reuse_backward_error = worst;
            const int code = factor_pivoted(e, false, true);
            if (code) return code;
            worst = refine();
        } else if (!e->ordered && e->split16_pivoted && !(worst <= e->opt.backward_error_gate)) {
            // a partial-pivoting factor with split-FP16 updates that stalled
            // above the gate: the same factor with the FP32 update first
            report.pivot_refactor = 1;
            report.reuse_backward_error = worst;
            const int code = factor_pivoted(e, false, true);
            if (code) return code;
            worst = refine();
        }
        if (!(worst <= e->opt.backward_error_gate) && e->opt.gmres_krylov > 0 && std::isfinite(worst)) {
            // plain refinement stalled above the gate: GMRES-IR on the same
            // factor before refactoring in FP64
            worst = gmres_refine();
        }
        if (!(worst <= e->opt.backward_error_gate)) {
            // refinement stalled above the gate: refactor in FP64
            report.used_fp64_fallback = 1;
            const int code = factor_pivoted(e, true);
            if (code) return code;
            HIP_CHECK(hipMemcpyAsync(e->d_X.ptr, e->d_B.ptr, sizeof(double2) * tot, hipMemcpyDeviceToDevice, s));
            lu_solve_permuted<double2, double2>(e->d_Z.ptr, n, n, e->d_perm.ptr, e->d_row.ptr, e->d_col.ptr,
                                                e->d_X.ptr, nrhs, *e->lu);
            measure();
            worst = 0.0;
            for (int q = 0; q < nrhs; ++q) worst = std::fmax(worst, eta[q]);
        }
        HIP_CHECK(hipEventRecord(e->ev[4], s));
    } else {
        HIP_CHECK(hipMemcpyAsync(e->d_X.ptr, e->d_B.ptr, sizeof(double2) * tot, hipMemcpyDeviceToDevice, s));
        lu_solve_permuted<double2, double2>(e->d_Z.ptr, n, n, e->d_perm.ptr, e->d_row.ptr, e->d_col.ptr, e->d_X.ptr,
                                            nrhs, *e->lu);
        HIP_CHECK(hipEventRecord(e->ev[4], s));
        measure();
        for (int q = 0; q < nrhs; ++q) worst = std::fmax(worst, eta[q]);
    }
    HIP_CHECK(hipEventRecord(e->ev[5], s));
    report.backward_error = worst;
    return worst <= e->opt.backward_error_gate ? 0 : -3;
}

// After the solves of a point: the factor now held serves later calls at
// (f, mode) only if the point succeeded on the factor of that mode (a mixed
// point that fell back to FP64 holds the FP64 factor under the FP64 row
// order, and a failed point holds nothing reusable).
static void note_cached_factor(PeecEngine *e, double f, int mode, int code, const PeecPointReport &report) {
    e->cache_ok = code == 0 && !report.used_fp64_fallback;
    e->cached_f = f;
    e->cached_mode = mode;
}

static void engine_timings(PeecEngine *e, PeecPointReport &report) {
    e->stream.sync();
    e->timer.resolve();
    float ms[5];
    for (int i = 0; i < 5; ++i) HIP_CHECK(hipEventElapsedTime(&ms[i], e->ev[i], e->ev[i + 1]));
    report.ms_assembly = ms[0];
    report.ms_fill = ms[1];
    report.ms_factor = ms[2];
    report.ms_solve = ms[3];
    report.ms_check = ms[4];
}

extern "C" {

void peec_engine_options_default(PeecEngineOptions *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    inductance_options_default(&o->inductance, 1);
    potential_options_default(&o->potential, 1);
    o->retarded = 0;
    o->smooth_order = 2;
    o->near_smooth_order = 4;
    o->block_size = 0;
    o->max_refinement_steps = 30;
    o->refinement_target = 1e-12;
    o->backward_error_gate = 1e-10;
    o->reuse_pivots = 1;
    o->smooth_moments = -1;
    o->split_fp16_update = 1;
    o->shifted_moments = 0;
    o->gmres_krylov = 20;
}

// Moment terms stored by default: the series then covers k * diameter up to
// about 10.5 (f up to 200 MHz on a body of 2.45 m); the alternating sum loses
// up to a factor e^(k diameter) ~ 4e4 to cancellation there, ~1e-12 of the
// entries, far below the 2e-5 error of the remainder's quadrature rule
// itself. Fewer than 8 are not stored.
static constexpr int kAutoMomentTerms = 56, kMinAutoMomentTerms = 8;

static int engine_create(const Mesh *mesh, const DualMesh *dual, const PeecEngineOptions *o, const double *L_host,
                         const double *P_host, PeecEngine **out);

int peec_engine_create(const Mesh *mesh, const DualMesh *dual, const PeecEngineOptions *o, PeecEngine **out) {
    return engine_create(mesh, dual, o, nullptr, nullptr, out);
}

int peec_engine_create_with_matrices(const Mesh *mesh, const DualMesh *dual, const PeecEngineOptions *o,
                                     const double *L, const double *P, PeecEngine **out) {
    if (!L || !P) return -1;
    return engine_create(mesh, dual, o, L, P, out);
}

static int engine_create(const Mesh *mesh, const DualMesh *dual, const PeecEngineOptions *o, const double *L_host,
                         const double *P_host, PeecEngine **out) {
    if (!mesh || !dual || !o || !out || !mesh->edges || mesh->n_edges == 0 || mesh->n_nodes == 0) return -1;
    if (!gl_order_supported(o->smooth_order) || !gl_order_supported(o->near_smooth_order)) {
        PEEC_ERR("unsupported smooth_order %d / near_smooth_order %d.", o->smooth_order, o->near_smooth_order);
        return -1;
    }
    if (!require_device()) return -1;
    try {
        auto e = std::make_unique<PeecEngine>();
        e->opt = *o;
        e->ne = (int)mesh->n_edges;
        e->nv = (int)mesh->n_nodes;
        e->n = e->ne + e->nv;
        hipStream_t s = e->stream;
        QuadratureOrders ql{o->inductance.far_order, o->inductance.near_order, o->inductance.touching_order,
                            o->inductance.self_order, o->inductance.near_factor};
        QuadratureOrders qp{o->potential.far_order, o->potential.near_order, o->potential.touching_order,
                            o->potential.self_order, o->potential.near_factor};
        e->edges = std::make_unique<RegionSet>(mesh, dual, RegionKind::Edge, ql, o->smooth_order,
                                               o->inductance.direction_tolerance, s, true, o->near_smooth_order);
        e->nodes = std::make_unique<RegionSet>(mesh, dual, RegionKind::Node, qp, o->smooth_order, 0.0, s, true,
                                               o->near_smooth_order);
        e->lu = std::make_unique<LUContext>(s);
        e->lu->timer = &e->timer;
        e->lu->split_fp16_update = o->split_fp16_update != 0;
        if (o->panel_lds_kib > 0) e->lu->panel_lds_limit = (size_t)o->panel_lds_kib * 1024;
        if (mesh->edge_centers && mesh->edge_vectors) e->rcs = std::make_unique<RcsGeometry>(mesh, dual, s);
        std::vector<int> edges(mesh->edges, mesh->edges + 2 * e->ne);
        e->d_edges.upload(edges, s);
        e->d_L.resize((size_t)e->ne * e->ne);
        e->d_P.resize((size_t)e->nv * e->nv);
        e->d_err.resize(1);
        HIP_CHECK(hipMemsetAsync(e->d_err.ptr, 0, sizeof(int), s));
        AssemblyWork w;
        DeviceBuffer<unsigned long long> counts;
        if (L_host && P_host) {
            // static matrices supplied by the caller (e.g. read back from
            // Matrix Market files): no assembly
            HIP_CHECK(hipMemcpyAsync(e->d_L.ptr, L_host, sizeof(double) * e->ne * e->ne, hipMemcpyHostToDevice, s));
            HIP_CHECK(hipMemcpyAsync(e->d_P.ptr, P_host, sizeof(double) * e->nv * e->nv, hipMemcpyHostToDevice, s));
        } else {
            // timed, with the near-field evaluation counts (the work of the
            // static assembly in flop-model terms)
            counts.resize(2);
            HIP_CHECK(hipMemsetAsync(counts.ptr, 0, 2 * sizeof(unsigned long long), s));
            e->timer.enabled = true;
            e->edges->assemble_static(e->d_L.ptr, s, e->d_err.ptr, &w, &e->timer, counts.ptr);
            e->nodes->assemble_static(e->d_P.ptr, s, e->d_err.ptr, &w, &e->timer, counts.ptr);
        }
        int err = 0;
        e->d_err.download(&err, 1, s);
        unsigned long long hc[2] = {0, 0};
        if (counts.ptr) counts.download(hc, 2, s);
        e->stream.sync();
        if (e->timer.enabled) {
            e->timer.resolve();
            PeecStaticStats &st = e->static_stats;
            st.far_evaluations = w.far_evaluations;
            st.near_potential_evaluations = (long long)hc[0];
            st.near_pair_evaluations = (long long)hc[1];
            st.far_ms = e->timer.totals[kKernelFarTile].ms;
            st.near_ms = e->timer.totals[kKernelNearEntry].ms;
            e->timer.reset();
            e->timer.enabled = false;
        }
        if (err) {
            PEEC_ERR("static assembly failed (code %d).", err);
            return -1;
        }
        if (o->retarded && o->smooth_moments != 0) {
            // moments of the smooth remainder for frequency sweeps: the
            // requested count, or (-1) kAutoMomentTerms when they fit in a
            // quarter of the device memory left free by the engine's
            // per-point working set
            int terms = std::min(o->smooth_moments, RegionSet::kMaxMomentTerms);
            if (o->smooth_moments < 0) {
                size_t free_b = 0, total_b = 0;
                HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
                const size_t point = engine_point_bytes(e->ne, e->nv, true);
                const size_t room = free_b > point ? (free_b - point) / 4 : 0;
                terms = kAutoMomentTerms;
                while (terms >= kMinAutoMomentTerms &&
                       RegionSet::moment_bytes(e->ne, terms) + RegionSet::moment_bytes(e->nv, terms) > room)
                    --terms;
                if (terms < kMinAutoMomentTerms) terms = 0;
            }
            if (terms > 0) {
                e->edges->build_moments(terms, s);
                e->nodes->build_moments(terms, s);
            }
        }
        e->far_evals = w.far_evaluations;
        e->smooth_evals = e->edges->smooth_evaluations() + e->nodes->smooth_evaluations();
        // stage timestamps only (read after the stream has been synchronised):
        // no system-scope fence
        for (hipEvent_t &ev : e->ev) HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableSystemFence));
        *out = e.release();
        return 0;
    } catch (const HipError &err) {
        PEEC_ERR("%s", err.what.c_str());
        return -1;
    } catch (const std::bad_alloc &) {
        PEEC_ERR("out of host memory.");
        return -1;
    }
}

void peec_engine_destroy(PeecEngine *e) { delete e; }

int peec_device_count(void) {
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess ? n : 0;
}

int peec_device_synchronize(void) {
    const hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        PEEC_ERR("hipDeviceSynchronize failed: %s", hipGetErrorString(e));
        return -1;
    }
    return 0;
}

int peec_engine_set_profiling(PeecEngine *e, int enable) {
    if (!e) return -1;
    e->timer.reset();
    e->timer.enabled = enable != 0;
    return 0;
}

int peec_engine_kernel_stats(const PeecEngine *e, int kernel, PeecKernelStats *out) {
    if (!e || !out || kernel < 0 || kernel >= kKernelKinds) return -1;
    const KernelTotals &t = e->timer.totals[kernel];
    out->launches = t.launches;
    out->total_ms = t.ms;
    out->flops = t.flops;
    out->bytes = t.bytes;
    return 0;
}

size_t peec_engine_edges(const PeecEngine *e) { return e ? (size_t)e->ne : 0; }
size_t peec_engine_nodes(const PeecEngine *e) { return e ? (size_t)e->nv : 0; }
long long peec_engine_far_evaluations(const PeecEngine *e) { return e ? e->far_evals : 0; }
int peec_engine_static_stats(const PeecEngine *e, PeecStaticStats *out) {
    if (!e || !out) return -1;
    *out = e->static_stats;
    return 0;
}
long long peec_engine_smooth_evaluations(const PeecEngine *e) { return e ? e->smooth_evals : 0; }

int peec_engine_static_matrices(const PeecEngine *e, double *L, double *P) {
    if (!e) return -1;
    try {
        if (L) e->d_L.download(L, (size_t)e->ne * e->ne, e->stream);
        if (P) e->d_P.download(P, (size_t)e->nv * e->nv, e->stream);
        e->stream.sync();
        return 0;
    } catch (const HipError &err) {
        PEEC_ERR("%s", err.what.c_str());
        return -1;
    }
}

// Shifted-center moments: terms per center (the series then covers
// |k - k0| * diameter up to ~3.2 each side of the center).
static constexpr int kAutoShiftTerms = 30;
static constexpr double kShiftHalfWidth = 3.2;

// Moments around k0 for a point above the range of the moments at 0: kept
// if they cover k, else (re)built with the window starting at k (a sweep
// upwards reuses them), when they fit in half the free device memory.
static bool ensure_shifted(PeecEngine *e, double k) {
    if (e->opt.shifted_moments == 0 || !e->shift_ok) return false;
    RegionSet &ed = *e->edges, &nd = *e->nodes;
    if (ed.shifted_terms_for(k) <= ed.shifted_terms() && nd.shifted_terms_for(k) <= nd.shifted_terms()) return true;
    const int terms = e->opt.shifted_moments > 0 ? std::min(e->opt.shifted_moments, RegionSet::kMaxShiftTerms)
                                                 : kAutoShiftTerms;
    const size_t bytes = 2 * (RegionSet::moment_bytes(e->ne, terms) + RegionSet::moment_bytes(e->nv, terms));
    size_t free_b = 0, total_b = 0;
    HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    const size_t held = ed.shifted_bytes() + nd.shifted_bytes();
    if (bytes > (free_b + held) / 2) {
        e->shift_ok = false;  // never fits on this device: quadrature from now on
        return false;
    }
    const double diam = std::max(ed.diameter(), nd.diameter());
    double half = kShiftHalfWidth;
    // the half width these terms cover: (half)^terms / terms! <= 1e-17
    for (;;) {
        double t = 1.0;
        for (int m = 1; m <= terms; ++m) t *= half / m;
        if (t <= 1e-17 || half < 0.05) break;
        half *= 0.9;
    }
    const double k0 = k + half / diam;
    ed.build_shifted_moments(terms, k0, e->stream);
    nd.build_shifted_moments(terms, k0, e->stream);
    return ed.shifted_terms_for(k) <= ed.shifted_terms() && nd.shifted_terms_for(k) <= nd.shifted_terms();
}

// Retarded Lp(w) and P(w): from the stored moments at 0 when the series at
// this wave number needs no more terms than stored, else from the moments
// around a shifted center, else by quadrature. Returns the moment terms used
// for Lp (0 = quadrature); *center = the center's wave number (0 for the
// series at 0).
static int assemble_retarded(PeecEngine *e, double k, double *center) {
    e->d_Lw.resize((size_t)e->ne * e->ne);
    e->d_Pw.resize((size_t)e->nv * e->nv);
    AssemblyWork w;
    if (center) *center = 0.0;
    const int tl = e->edges->series_terms(k), tp = e->nodes->series_terms(k);
    if (tl <= e->edges->moment_terms() && tp <= e->nodes->moment_terms()) {
        e->edges->assemble_retarded_moments(e->d_L.ptr, k, tl, e->d_Lw.ptr, e->stream, &e->timer);
        e->nodes->assemble_retarded_moments(e->d_P.ptr, k, tp, e->d_Pw.ptr, e->stream, &e->timer);
        return tl;
    }
    if (ensure_shifted(e, k)) {
        const int sl = e->edges->shifted_terms_for(k), sp = e->nodes->shifted_terms_for(k);
        e->edges->assemble_retarded_shifted(e->d_L.ptr, k, sl, e->d_Lw.ptr, e->stream, &e->timer);
        e->nodes->assemble_retarded_shifted(e->d_P.ptr, k, sp, e->d_Pw.ptr, e->stream, &e->timer);
        if (center) *center = e->edges->shifted_k0();
        return sl;
    }
    e->edges->assemble_retarded(e->d_L.ptr, k, e->d_Lw.ptr, e->stream, &w, &e->timer);
    e->nodes->assemble_retarded(e->d_P.ptr, k, e->d_Pw.ptr, e->stream, &w, &e->timer);
    return 0;
}

int peec_engine_retarded_matrices(PeecEngine *e, double k, Complex *L, Complex *P) {
    if (!e || !(k >= 0.0)) return -1;
    try {
        assemble_retarded(e, k);
        if (L) e->d_Lw.download(reinterpret_cast<double2 *>(L), (size_t)e->ne * e->ne, e->stream);
        if (P) e->d_Pw.download(reinterpret_cast<double2 *>(P), (size_t)e->nv * e->nv, e->stream);
        e->stream.sync();
        return 0;
    } catch (const HipError &err) {
        PEEC_ERR("%s", err.what.c_str());
        return -1;
    }
}

int peec_engine_solve(PeecEngine *e, double f, int mode, const int *ports, size_t nports, Complex *z_in,
                      Complex *edge_current, Complex *node_voltage, PeecPointReport *rep) {
    if (!e || !ports || nports == 0 || !(f > 0.0) || !std::isfinite(f)) return -1;
    if (mode != PEEC_SOLVE_FP64 && mode != PEEC_SOLVE_MIXED) return -1;
    for (size_t k = 0; k < 2 * nports; ++k)
        if (ports[k] < 0 || ports[k] >= e->nv) return -1;
    PeecPointReport report{};
    report.frequency = f;
    try {
        hipStream_t s = e->stream;
        const int n = e->n, ne = e->ne, nv = e->nv, nr = (int)nports;
        int code = engine_factor(e, f, mode, report);
        if (code != 0) {
            if (rep) *rep = report;
            return code;
        }
        e->d_ports.upload(ports, 2 * nports, s);
        e->d_B.resize((size_t)n * nr);
        const size_t tot = (size_t)n * nr;
        hipLaunchKernelGGL(port_rhs_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, e->d_B.ptr, ne, nv,
                           e->d_P.ptr, e->retarded_now ? e->d_Pw.ptr : nullptr, e->d_ports.ptr, nr);
        code = engine_solve(e, mode, nr, report);
        note_cached_factor(e, f, mode, code, report);
        if (code == -1 || code == -2) {
            if (rep) *rep = report;
            return code;
        }
        std::vector<double2> x(tot);
        e->d_X.download(x.data(), tot, s);
        e->stream.sync();
        for (int k = 0; k < nr; ++k) {
            const double2 *xk = x.data() + (size_t)k * n;
            const double2 va = xk[ne + ports[2 * k]], vb = xk[ne + ports[2 * k + 1]];
            if (z_in) z_in[k] = Complex{va.x - vb.x, va.y - vb.y};
            if (edge_current) std::memcpy(edge_current + (size_t)k * ne, xk, sizeof(Complex) * ne);
            if (node_voltage) std::memcpy(node_voltage + (size_t)k * nv, xk + ne, sizeof(Complex) * nv);
        }
        engine_timings(e, report);
        if (rep) *rep = report;
        if (code == -3)
            PEEC_ERR("backward error %.3e above the gate %.1e at f = %.6g Hz.", report.backward_error,
                     e->opt.backward_error_gate, f);
        return code;
    } catch (const HipError &err) {
        PEEC_ERR("%s", err.what.c_str());
        if (rep) *rep = report;
        return -1;
    }
}

// The engine's per-point buffers and a batch slot's trade places (the
// engine then works on that point's blocks and scales).
static void swap_point(PeecEngine *e, PeecEngine::PointSlot &p) {
    std::swap(e->d_Lw, p.Lw);
    std::swap(e->d_Pw, p.Pw);
    std::swap(e->d_row, p.row);
    std::swap(e->d_col, p.col);
}

int peec_engine_solve_batch(PeecEngine *e, size_t n_points, const double *frequency, const int *const *ports,
                            const size_t *n_ports, Complex *const *z_in, PeecPointReport *reports, int *status) {
    if (!e || n_points == 0 || !frequency || !ports || !n_ports || !z_in || !status) return -1;
    for (size_t p = 0; p < n_points; ++p) {
        if (!(frequency[p] > 0.0) || !std::isfinite(frequency[p]) || !ports[p] || n_ports[p] == 0 || !z_in[p])
            return -1;
        for (size_t k = 0; k < 2 * n_ports[p]; ++k)
            if (ports[p][k] < 0 || ports[p][k] >= e->nv) return -1;
    }
    auto worst = [&]() {
        for (size_t p = 0; p < n_points; ++p)
            if (status[p] != 0) return status[p];
        return 0;
    };
    auto single = [&](size_t p) {
        PeecPointReport r{};
        status[p] = peec_engine_solve(e, frequency[p], PEEC_SOLVE_MIXED, ports[p], n_ports[p], z_in[p], nullptr,
                                      nullptr, &r);
        r.batch_size = 1;
        if (reports) reports[p] = r;
    };
    // points before the engine has a row order to reuse go alone (the first
    // one's partial-pivoting factor sets it)
    size_t first = 0;
    while (first < n_points && !(e->opt.reuse_pivots && e->have_reuse)) single(first++);
    const size_t m = n_points - first;
    if (m <= 1) {
        if (m == 1) single(first);
        return worst();
    }
    std::vector<PeecPointReport> rep(m);
    for (size_t b = 0; b < m; ++b) status[first + b] = -1;  // until the point is done
    try {
        hipStream_t s = e->stream;
        const int n = e->n, ne = e->ne, nv = e->nv;
        const size_t nn = (size_t)n * n;
        const bool split16 = e->lu->split_fp16_update;
        ShapeTuner &tu = e->tune_batch;
        const bool trial = e->opt.block_size == 0 && tu.best < 0;
        const FactorShape shape = e->opt.block_size > 0 ? FactorShape{e->opt.block_size, 64}
                                  : trial                ? tu.trial(split16)
                                                         : tu.chosen(split16);
        const int nb = shape.nb;
        e->cache_ok = false;
        e->slots.resize(m);
        e->d_Z32.reserve(m * nn);
        e->d_batch_perm.resize(n);
        HIP_CHECK(hipMemcpyAsync(e->d_batch_perm.ptr, e->d_reuse_perm.ptr, sizeof(int) * n, hipMemcpyDeviceToDevice,
                                 s));
        std::vector<int> prep(m, 0);
        std::vector<double> limits(m, 0.0);
        std::vector<float> ms_prep(m, 0.f), ms_asm(m, 0.f);
        // retarded members whose series the stored moments cover: their
        // blocks in shared passes over the moments, up to four at a time
        // (each member's time is its share of the pass)
        std::vector<int> pre_terms(m, 0);
        if (e->opt.retarded != 0) {
            constexpr int G = RegionSet::kMaxMomentBatch;
            for (size_t g = 0; g < m; g += G) {
                const int cnt = (int)std::min<size_t>(G, m - g);
                double ks[G];
                int tl[G], tp[G];
                double2 *ol[G], *op[G];
                bool ok = cnt > 1;
                for (int b = 0; b < cnt; ++b) {
                    ks[b] = kTwoPi * frequency[first + g + b] / kC0;
                    tl[b] = e->edges->series_terms(ks[b]);
                    tp[b] = e->nodes->series_terms(ks[b]);
                    ok = ok && tl[b] <= e->edges->moment_terms() && tp[b] <= e->nodes->moment_terms();
                }
                if (!ok) continue;
                for (int b = 0; b < cnt; ++b) {
                    PeecEngine::PointSlot &sl = e->slots[g + b];
                    sl.Lw.resize((size_t)ne * ne);
                    sl.Pw.resize((size_t)nv * nv);
                    ol[b] = sl.Lw.ptr;
                    op[b] = sl.Pw.ptr;
                }
                HIP_CHECK(hipEventRecord(e->ev[0], s));
                e->edges->assemble_retarded_moments_batch(e->d_L.ptr, ks, tl, ol, cnt, s, &e->timer);
                e->nodes->assemble_retarded_moments_batch(e->d_P.ptr, ks, tp, op, cnt, s, &e->timer);
                HIP_CHECK(hipEventRecord(e->ev[1], s));
                HIP_CHECK(hipEventSynchronize(e->ev[1]));
                float ms = 0.f;
                HIP_CHECK(hipEventElapsedTime(&ms, e->ev[0], e->ev[1]));
                for (int b = 0; b < cnt; ++b) {
                    pre_terms[g + b] = tl[b];
                    ms_asm[g + b] = ms / (float)cnt;
                }
            }
        }
        // each point's retarded blocks, scales and FP32 matrix in the batch's
        // row order, the matrices side by side in d_Z32
        for (size_t b = 0; b < m; ++b) {
            PeecEngine::PointSlot &sl = e->slots[b];
            rep[b].frequency = frequency[first + b];
            rep[b].batch_size = (int)m;
            swap_point(e, sl);
            HIP_CHECK(hipEventRecord(e->ev[0], s));
            prep[b] = prepare_point(e, frequency[first + b], true, rep[b], pre_terms[b]);
            if (prep[b] == 0) {
                double max_scaled = 0.0;
                if (convert_equilibrated_to_f32(e->mna, e->d_Z32.ptr + b * nn, e->d_batch_perm.ptr, e->d_row.ptr,
                                                e->d_col.ptr, &max_scaled, *e->lu) != 0)
                    prep[b] = -2;
                limits[b] = 1.0e-14 * max_scaled;
            }
            HIP_CHECK(hipEventRecord(e->ev[3], s));
            sl.mna = e->mna;
            sl.limit = limits[b];
            swap_point(e, sl);
            e->stream.sync();
            if (pre_terms[b] == 0) HIP_CHECK(hipEventElapsedTime(&ms_asm[b], e->ev[0], e->ev[1]));
            HIP_CHECK(hipEventElapsedTime(&ms_prep[b], e->ev[1], e->ev[3]));
        }
        // one ordered factor over the batch (a member that failed above is
        // factored too, from whatever its slot holds, and its result ignored)
        e->d_limits.upload(limits.data(), m, s);
        std::vector<LUStatus> fst(m);
        HIP_CHECK(hipEventRecord(e->ev[0], s));
        e->lu->h3_tile_rows = shape.rows;
        const LUStatus bst =
            lu_factor_ordered(e->d_Z32.ptr, n, n, 0.0, *e->lu, nb, (int)m, nn, e->d_limits.ptr, fst.data());
        HIP_CHECK(hipEventRecord(e->ev[1], s));
        e->stream.sync();
        e->timer.resolve();
        float ms_factor = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms_factor, e->ev[0], e->ev[1]));
        if (trial && bst.code == 0) tu.record(1e-3 * ms_factor / (double)m, split16);
        // one port per point and every factor good: the points are refined
        // in lock-step, their triangular solves in shared launches
        // (lu_solve_members); a point that stays above the gate goes on
        // alone below
        std::vector<char> done(m, 0);
        bool lockstep = m <= 64;  // the members' step masks are one 64-bit word
        for (size_t b = 0; b < m; ++b) lockstep = lockstep && prep[b] == 0 && fst[b].code == 0 && n_ports[first + b] == 1;
        if (lockstep) {
            const size_t tot = (size_t)m * n;
            e->d_B.resize(tot);
            e->d_X.resize(tot);
            e->d_R.resize(tot);
            e->d_Xbest.resize(tot);
            e->d_ax.resize(tot);
            e->d_eta.resize(m);
            e->d_mrow.resize(tot);
            e->d_mcol.resize(tot);
            std::vector<int> hp(2 * m);
            for (size_t b = 0; b < m; ++b) {
                hp[2 * b] = ports[first + b][0];
                hp[2 * b + 1] = ports[first + b][1];
            }
            e->d_ports.upload(hp.data(), 2 * m, s);
            HIP_CHECK(hipEventRecord(e->ev[0], s));
            const bool ret = e->opt.retarded != 0;
            for (size_t b = 0; b < m; ++b) {
                PeecEngine::PointSlot &sl = e->slots[b];
                hipLaunchKernelGGL(port_rhs_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                                   e->d_B.ptr + b * n, ne, nv, e->d_P.ptr, ret ? sl.Pw.ptr : nullptr,
                                   e->d_ports.ptr + 2 * b, 1);
                HIP_CHECK(hipMemcpyAsync(e->d_mrow.ptr + b * n, sl.row.ptr, sizeof(double) * n,
                                         hipMemcpyDeviceToDevice, s));
                HIP_CHECK(hipMemcpyAsync(e->d_mcol.ptr + b * n, sl.col.ptr, sizeof(double) * n,
                                         hipMemcpyDeviceToDevice, s));
            }
            std::vector<double> eta(m), best(m);
            std::vector<unsigned long long> bits(m);
            std::vector<int> since(m, 0), active(m), flags(m), steps(m, 0);
            // residuals of every member (all), or only of those still refined
            // (a finished member's solution, and so its eta, no longer change)
            auto measure = [&](bool all) {
                for (size_t b = 0; b < m; ++b)
                    if (all || active[b])
                        residual_multi(e->slots[b].mna, e->d_X.ptr + b * n, e->d_B.ptr + b * n, e->d_R.ptr + b * n,
                                       1, e->d_ax.ptr + b * n, e->d_eta.ptr + b, s);
                e->d_eta.download(bits.data(), m, s);
                e->stream.sync();
                for (size_t b = 0; b < m; ++b) {
                    double v;
                    std::memcpy(&v, &bits[b], sizeof v);
                    eta[b] = std::isfinite(v) ? v : INFINITY;
                }
            };
            auto mask_of = [&](const std::vector<int> &fl) {
                unsigned long long mk = 0ull;
                for (size_t b = 0; b < m; ++b)
                    if (fl[b]) mk |= 1ull << b;
                return mk;
            };
            const unsigned grid = (unsigned)((tot + 255) / 256);
            HIP_CHECK(hipMemcpyAsync(e->d_X.ptr, e->d_B.ptr, sizeof(double2) * tot, hipMemcpyDeviceToDevice, s));
            lu_solve_members(e->d_Z32.ptr, nn, n, e->d_batch_perm.ptr, e->d_mrow.ptr, e->d_mcol.ptr, e->d_X.ptr,
                             (int)m, *e->lu, false);
            measure(true);
            best = eta;
            for (size_t b = 0; b < m; ++b) active[b] = eta[b] > e->opt.refinement_target;
            HIP_CHECK(hipMemcpyAsync(e->d_Xbest.ptr, e->d_X.ptr, sizeof(double2) * tot, hipMemcpyDeviceToDevice, s));
            for (int step = 0; step < e->opt.max_refinement_steps; ++step) {
                bool any = false;
                for (size_t b = 0; b < m; ++b) any = any || active[b];
                if (!any) break;
                lu_solve_members(e->d_Z32.ptr, nn, n, e->d_batch_perm.ptr, e->d_mrow.ptr, e->d_mcol.ptr, e->d_R.ptr,
                                 (int)m, *e->lu, true);
                hipLaunchKernelGGL(masked_axpy_kernel, dim3(grid), dim3(256), 0, s, e->d_X.ptr, e->d_R.ptr,
                                   (const int *)nullptr, mask_of(active), n, (int)m);
                for (size_t b = 0; b < m; ++b) steps[b] += active[b];
                measure(false);
                bool improved = false;
                for (size_t b = 0; b < m; ++b) {
                    flags[b] = 0;
                    if (!active[b]) continue;
                    if (eta[b] < best[b]) {
                        best[b] = eta[b];
                        since[b] = 0;
                        flags[b] = 1;
                        improved = true;
                    } else if (++since[b] >= 2) {
                        active[b] = 0;
                    }
                    if (best[b] <= e->opt.refinement_target) active[b] = 0;
                }
                if (improved)
                    hipLaunchKernelGGL(masked_copy_kernel, dim3(grid), dim3(256), 0, s, e->d_Xbest.ptr, e->d_X.ptr,
                                       (const int *)nullptr, mask_of(flags), n, (int)m);
            }
            for (size_t b = 0; b < m; ++b) flags[b] = eta[b] > best[b];
            hipLaunchKernelGGL(masked_copy_kernel, dim3(grid), dim3(256), 0, s, e->d_X.ptr, e->d_Xbest.ptr,
                               (const int *)nullptr, mask_of(flags), n, (int)m);
            HIP_CHECK(hipEventRecord(e->ev[1], s));
            std::vector<double2> x(tot);
            e->d_X.download(x.data(), tot, s);
            e->stream.sync();
            float ms_solve = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms_solve, e->ev[0], e->ev[1]));
            for (size_t b = 0; b < m; ++b) {
                if (!(best[b] <= e->opt.backward_error_gate)) continue;  // alone below, with the fallbacks
                const size_t p = first + b;
                const double2 *xb = x.data() + b * n;
                const double2 va = xb[ne + ports[p][0]], vb = xb[ne + ports[p][1]];
                z_in[p][0] = Complex{va.x - vb.x, va.y - vb.y};
                PeecPointReport &r = rep[b];
                r.ms_assembly = ms_asm[b];
                r.ms_fill = ms_prep[b];
                r.ms_factor = ms_factor / (double)m;
                r.ms_solve = ms_solve / (double)m;
                r.block_size = nb;
                r.update_tile_rows = shape.rows;
                r.pivot_reuse = 1;
                r.refinement_steps = steps[b];
                r.backward_error = best[b];
                status[p] = 0;
                done[b] = 1;
            }
        }
        // each point's solves and refinement on its own factor (with the
        // single point's fallbacks: a rejected order is refactored with
        // pivoting, a point above the gate in FP64)
        for (size_t b = 0; b < m; ++b) {
            if (done[b]) continue;
            const size_t p = first + b;
            PeecPointReport &r = rep[b];
            r.ms_assembly = ms_asm[b];
            r.ms_fill = ms_prep[b];
            r.ms_factor = ms_factor / (double)m;
            r.block_size = nb;
            r.update_tile_rows = shape.rows;
            if (prep[b] != 0) {
                status[p] = prep[b];
                continue;
            }
            PeecEngine::PointSlot &sl = e->slots[b];
            swap_point(e, sl);
            e->mna = sl.mna;
            e->limit = sl.limit;
            e->z32 = e->d_Z32.ptr + b * nn;
            e->retarded_now = e->opt.retarded != 0;
            e->z0_ready = false;
            e->z_equilibrated = false;
            int code = 0;
            if (fst[b].code == 0) {
                e->ordered = true;
                r.pivot_reuse = 1;
                HIP_CHECK(hipMemcpyAsync(e->d_perm.ptr, e->d_batch_perm.ptr, sizeof(int) * n,
                                         hipMemcpyDeviceToDevice, s));
            } else if (fst[b].code == -2) {
                e->ordered = false;
                code = factor_pivoted(e, false);
            } else {
                code = factor_error(fst[b]);
            }
            if (code == 0) {
                const int nr = (int)n_ports[p];
                const size_t tot = (size_t)n * nr;
                e->d_ports.upload(ports[p], 2 * n_ports[p], s);
                e->d_B.resize(tot);
                HIP_CHECK(hipEventRecord(e->ev[3], s));
                hipLaunchKernelGGL(port_rhs_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, e->d_B.ptr,
                                   ne, nv, e->d_P.ptr, e->retarded_now ? e->d_Pw.ptr : nullptr, e->d_ports.ptr, nr);
                code = engine_solve(e, PEEC_SOLVE_MIXED, nr, r);
                if (code == 0 || code == -3) {
                    std::vector<double2> x(tot);
                    e->d_X.download(x.data(), tot, s);
                    e->stream.sync();
                    for (int k = 0; k < nr; ++k) {
                        const double2 *xk = x.data() + (size_t)k * n;
                        const double2 va = xk[ne + ports[p][2 * k]], vb = xk[ne + ports[p][2 * k + 1]];
                        z_in[p][k] = Complex{va.x - vb.x, va.y - vb.y};
                    }
                    float a = 0.f, c = 0.f;
                    HIP_CHECK(hipEventElapsedTime(&a, e->ev[3], e->ev[4]));
                    HIP_CHECK(hipEventElapsedTime(&c, e->ev[4], e->ev[5]));
                    r.ms_solve = a;
                    r.ms_check = c;
                }
                if (code == -3)
                    PEEC_ERR("backward error %.3e above the gate %.1e at f = %.6g Hz.", r.backward_error,
                             e->opt.backward_error_gate, frequency[p]);
            }
            status[p] = code;
            swap_point(e, sl);
        }
        // the engine's own buffers are back; nothing it holds is a factor of
        // a single point
        e->cache_ok = false;
        e->z32 = e->d_Z32.ptr;
        e->stream.sync();
        e->timer.resolve();
    } catch (const HipError &err) {
        PEEC_ERR("%s", err.what.c_str());
        e->cache_ok = false;
        e->z32 = e->d_Z32.ptr;
        if (reports)
            for (size_t b = 0; b < m; ++b) reports[first + b] = rep[b];
        return -1;
    }
    if (reports)
        for (size_t b = 0; b < m; ++b) reports[first + b] = rep[b];
    return worst();
}

void peec_rcs_options_default(PeecRcsOptions *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->wave_type = INCIDENT_WAVE_COMPLEX;
    o->polarization = INCIDENT_POLARIZATION_VERTICAL;
    o->E_amp = 1.0;
    o->phase = 0.0;
    o->excitation = SCATTERING_EDGE_ANALYTIC_PLANE_WAVE;
    rcs_options_default(&o->rcs, 1);
}

int peec_engine_monostatic_rcs(PeecEngine *e, double f, int mode, const PeecRcsOptions *o, const double *theta_deg,
                               const double *phi_deg, size_t n_angles, RcsResult *results, Complex *edge_current,
                               PeecPointReport *rep) {
    if (!e || !o || !theta_deg || !phi_deg || n_angles == 0 || !results || !(f > 0.0) || !std::isfinite(f))
        return -1;
    if (mode != PEEC_SOLVE_FP64 && mode != PEEC_SOLVE_MIXED) return -1;
    if (o->wave_type != INCIDENT_WAVE_COMPLEX && o->wave_type != INCIDENT_WAVE_COS &&
        o->wave_type != INCIDENT_WAVE_SIN)
        return -1;
    if ((o->excitation != SCATTERING_EDGE_MIDPOINT && o->excitation != SCATTERING_EDGE_ANALYTIC_PLANE_WAVE) ||
        (o->rcs.integration != RCS_CURRENT_MIDPOINT && o->rcs.integration != RCS_CURRENT_DUAL_PATCH_QUADRATURE) ||
        !gl_order_supported(o->rcs.surface_order) || !(o->E_amp > 0.0) || !std::isfinite(o->E_amp))
        return -1;
    if (!e->rcs) {
        PEEC_ERR("the engine was created without the mesh data RCS needs.");
        return -1;
    }
    const int na = (int)n_angles;
    // incident waves and the monostatic observation directions (back toward
    // the source, computed through the angles as the reference driver does,
    // main_rcs.c:1380-1440)
    std::vector<PlaneWave> waves(na);
    std::vector<double> dirs(3 * (size_t)na);
    std::vector<double> obs_theta(na), obs_phi(na);
    double k = 0.0;
    for (int a = 0; a < na; ++a) {
        IncidentField fld;
        if (incident_field_configure_harmonic(&fld, o->wave_type, o->polarization, o->E_amp, f, theta_deg[a],
                                              phi_deg[a], o->phase) != 0)
            return -1;
        waves[a] = plane_wave_of(fld);
        k = fld.wave_number;
        const double bx = -fld.k_hat[0], by = -fld.k_hat[1], bz = -fld.k_hat[2];
        double bt = std::acos(std::fmax(-1.0, std::fmin(1.0, bz))) * 180.0 / M_PI;
        double bp = std::atan2(by, bx) * 180.0 / M_PI;
        if (bp < 0.0) bp += 360.0;
        obs_theta[a] = bt;
        obs_phi[a] = bp;
        if (rcs_build_observation_direction(bt, bp, &dirs[3 * (size_t)a]) != 0) return -1;
    }
    PeecPointReport report{};
    report.frequency = f;
    try {
        hipStream_t s = e->stream;
        const int n = e->n, ne = e->ne;
        int code = engine_factor(e, f, mode, report);
        if (code != 0) {
            if (rep) *rep = report;
            return code;
        }
        // right-hand sides [-U ; 0] for every incidence
        e->d_B.resize((size_t)n * na);
        fill_zero(e->d_B.ptr, (size_t)n * na, s);
        DeviceBuffer<PlaneWave> dw;
        dw.upload(waves.data(), (size_t)na, s);
        launch_excitation(*e->rcs, dw.ptr, na, o->excitation, -1.0, e->d_B.ptr, (size_t)n, s);
        code = engine_solve(e, mode, na, report);
        note_cached_factor(e, f, mode, code, report);
        if (code == -1 || code == -2) {
            if (rep) *rep = report;
            return code;
        }
        DeviceBuffer<double> dd;
        dd.upload(dirs, s);
        DeviceBuffer<double> work(radiation_work_size(ne, na));
        DeviceBuffer<double2> dF(3 * (size_t)na);
        launch_radiation(*e->rcs, e->d_X.ptr, (size_t)n, dd.ptr, na, k, o->rcs.integration, o->rcs.surface_order,
                         work.ptr, dF.ptr, s);
        std::vector<double2> F(3 * (size_t)na);
        dF.download(F.data(), F.size(), s);
        if (edge_current)
            for (int a = 0; a < na; ++a)
                HIP_CHECK(hipMemcpyAsync(edge_current + (size_t)a * ne, e->d_X.ptr + (size_t)a * n,
                                         sizeof(Complex) * ne, hipMemcpyDeviceToHost, s));
        e->stream.sync();
        for (int a = 0; a < na; ++a) {
            RcsResult &r = results[a];
            std::memset(&r, 0, sizeof(r));
            r.theta_deg = obs_theta[a];
            r.phi_deg = obs_phi[a];
            for (int c = 0; c < 3; ++c) {
                r.direction[c] = dirs[3 * (size_t)a + c];
                r.radiation_vector[c] = Complex{F[3 * (size_t)a + c].x, F[3 * (size_t)a + c].y};
            }
            rcs_result_finish(r.direction, k, o->E_amp, &r);
        }
        engine_timings(e, report);
        if (rep) *rep = report;
        if (code == -3)
            PEEC_ERR("backward error %.3e above the gate %.1e at f = %.6g Hz.", report.backward_error,
                     e->opt.backward_error_gate, f);
        return code;
    } catch (const HipError &err) {
        PEEC_ERR("%s", err.what.c_str());
        if (rep) *rep = report;
        return -1;
    }
}

}  // extern "C"

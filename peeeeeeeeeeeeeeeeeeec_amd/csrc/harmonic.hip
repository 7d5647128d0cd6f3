// MNA system kernels (harmonic.c:175-552) and the device-resident
// frequency-point engine declared in include/peec_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/peec_gpu.h"
#include "assembly.h"
#include "harmonic_kernels.h"
#include "lu.h"

namespace peec {

constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kC0 = 299792458.0;  // main.c:47-54

// Z = [[R + jwL, A], [-P A^T, jw I]] written row-major with leading
// dimension ldz; A is implicit in the edge end points (incidence.c:34-115).
// When Lw/Pw are given (retarded form) they replace L/P.
__global__ void mna_fill_kernel(double2 *Z, double2 *Z0, size_t ldz, int ne, int nv, const double *L,
                                const double2 *Lw, const double *R, const double *P, const double2 *Pw,
                                const double *PAT, const int *edges, double omega, int chunks) {
    const int n = ne + nv;
    const int row = blockIdx.x / chunks;
    const int col = (blockIdx.x % chunks) * 256 + threadIdx.x;
    if (row >= n || col >= n) return;
    double2 z;
    if (row < ne) {
        if (col < ne) {
            const size_t k = (size_t)row * ne + col;
            const double r = (row == col && R) ? R[row] : 0.0;
            if (Lw) {
                const double2 l = Lw[k];
                z = make_double2(r - omega * l.y, omega * l.x);
            } else {
                z = make_double2(r, omega * L[k]);
            }
        } else {
            const int node = col - ne;
            double a = 0.0;
            if (node == edges[2 * row]) a = -1.0;
            if (node == edges[2 * row + 1]) a = 1.0;
            z = make_double2(a, 0.0);
        }
    } else {
        const int i = row - ne;
        if (col < ne) {
            const int a = edges[2 * col], b = edges[2 * col + 1];
            if (Pw) {
                const double2 pb = Pw[(size_t)i * nv + b], pa = Pw[(size_t)i * nv + a];
                z = make_double2(-(pb.x - pa.x), -(pb.y - pa.y));
            } else if (PAT) {
                z = make_double2(-PAT[(size_t)i * ne + col], 0.0);
            } else {
                z = make_double2(-(P[(size_t)i * nv + b] - P[(size_t)i * nv + a]), 0.0);
            }
        } else {
            z = make_double2(0.0, (col - ne) == i ? omega : 0.0);
        }
    }
    Z[(size_t)row * ldz + col] = z;
    if (Z0) Z0[(size_t)row * ldz + col] = z;
}

void launch_mna_fill(double2 *Z, double2 *Z0, size_t ldz, int ne, int nv, const double *L, const double2 *Lw,
                     const double *R, const double *P, const double2 *Pw, const double *PAT, const int *edges,
                     double omega, hipStream_t s) {
    const int n = ne + nv;
    const int chunks = (n + 255) / 256;
    hipLaunchKernelGGL(mna_fill_kernel, dim3((unsigned)((size_t)n * chunks)), dim3(256), 0, s, Z, Z0, ldz, ne, nv,
                       L, Lw, R, P, Pw, PAT, edges, omega, chunks);
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
    DeviceBuffer<float2> d_Z32;
    hipEvent_t ev[6] = {};
    long long far_evals = 0, smooth_evals = 0;
    KernelTimer timer;
    ~PeecEngine() {
        for (hipEvent_t &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

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
    o->refinement_target = 1e-14;
    o->backward_error_gate = 1e-10;
    o->reuse_pivots = 1;
}

int peec_engine_create(const Mesh *mesh, const DualMesh *dual, const PeecEngineOptions *o, PeecEngine **out) {
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
        std::vector<int> edges(mesh->edges, mesh->edges + 2 * e->ne);
        e->d_edges.upload(edges, s);
        e->d_L.resize((size_t)e->ne * e->ne);
        e->d_P.resize((size_t)e->nv * e->nv);
        e->d_err.resize(1);
        HIP_CHECK(hipMemsetAsync(e->d_err.ptr, 0, sizeof(int), s));
        AssemblyWork w;
        e->edges->assemble_static(e->d_L.ptr, s, e->d_err.ptr, &w);
        e->nodes->assemble_static(e->d_P.ptr, s, e->d_err.ptr, &w);
        int err = 0;
        e->d_err.download(&err, 1, s);
        e->stream.sync();
        if (err) {
            PEEC_ERR("static assembly failed (code %d).", err);
            return -1;
        }
        e->far_evals = w.far_evaluations;
        e->smooth_evals = e->edges->smooth_evaluations() + e->nodes->smooth_evaluations();
        for (hipEvent_t &ev : e->ev) HIP_CHECK(hipEventCreate(&ev));
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

static void assemble_retarded(PeecEngine *e, double k) {
    e->d_Lw.resize((size_t)e->ne * e->ne);
    e->d_Pw.resize((size_t)e->nv * e->nv);
    AssemblyWork w;
    e->edges->assemble_retarded(e->d_L.ptr, k, e->d_Lw.ptr, e->stream, &w, &e->timer);
    e->nodes->assemble_retarded(e->d_P.ptr, k, e->d_Pw.ptr, e->stream, &w, &e->timer);
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
        const size_t nn = (size_t)n * n;
        const double omega = kTwoPi * f;
        HIP_CHECK(hipEventRecord(e->ev[0], s));
        const bool ret = e->opt.retarded != 0;
        if (ret) assemble_retarded(e, omega / kC0);
        HIP_CHECK(hipEventRecord(e->ev[1], s));
        e->d_Z.resize(nn);
        e->d_Z0.resize(nn);
        launch_mna_fill(e->d_Z.ptr, e->d_Z0.ptr, n, ne, nv, e->d_L.ptr, ret ? e->d_Lw.ptr : nullptr, nullptr,
                        e->d_P.ptr, ret ? e->d_Pw.ptr : nullptr, nullptr, e->d_edges.ptr, omega, s);
        e->d_row.resize(n);
        e->d_col.resize(n);
        double max_scaled = 0.0;
        if (equilibrate_complex(e->d_Z.ptr, n, n, e->d_row.ptr, e->d_col.ptr, &max_scaled, *e->lu) != 0)
            return -2;
        HIP_CHECK(hipEventRecord(e->ev[2], s));
        const double limit = 1.0e-14 * max_scaled;
        e->d_ipiv.resize(n);
        e->d_perm.resize(n);
        e->d_reuse_perm.resize(n);
        auto keep_order = [&]() {
            // the row order of the partial-pivoting factor just made
            permutation_from_pivots(e->d_ipiv.ptr, n, e->d_perm.ptr, s);
            HIP_CHECK(hipMemcpyAsync(e->d_reuse_perm.ptr, e->d_perm.ptr, sizeof(int) * n, hipMemcpyDeviceToDevice, s));
            e->have_reuse = true;
        };
        auto factor_error = [&](const LUStatus &st) {
            if (st.code == -2) PEEC_ERR("complex LU singular at column %lld.", st.column);
            else PEEC_ERR("LU failed (code %d).", st.code);
            return st.code == -2 ? -2 : -1;
        };
        LUStatus st;
        bool ordered = false;
        if (mode == PEEC_SOLVE_MIXED) {
            e->d_Z32.resize(nn);
            if (e->opt.reuse_pivots && e->have_reuse) {
                convert_rows_to_f32(e->d_Z.ptr, e->d_Z32.ptr, e->d_reuse_perm.ptr, n, s);
                st = lu_factor_ordered(e->d_Z32.ptr, n, n, limit, *e->lu, e->opt.block_size);
                if (st.code == 0) {
                    ordered = true;
                    HIP_CHECK(hipMemcpyAsync(e->d_perm.ptr, e->d_reuse_perm.ptr, sizeof(int) * n,
                                             hipMemcpyDeviceToDevice, s));
                } else if (st.code != -2) {
                    return factor_error(st);
                }
            }
            if (!ordered) {
                convert_to_f32(e->d_Z.ptr, e->d_Z32.ptr, nn, s);
                st = lu_factor<float2>(e->d_Z32.ptr, n, n, e->d_ipiv.ptr, limit, *e->lu, e->opt.block_size);
                if (st.code != 0) return factor_error(st);
                keep_order();
            }
        } else {
            st = lu_factor<double2>(e->d_Z.ptr, n, n, e->d_ipiv.ptr, limit, *e->lu, e->opt.block_size);
            if (st.code != 0) return factor_error(st);
            keep_order();
        }
        HIP_CHECK(hipEventRecord(e->ev[3], s));
        report.pivot_reuse = ordered ? 1 : 0;
        e->d_ports.upload(ports, 2 * nports, s);
        e->d_B.resize((size_t)n * nr);
        e->d_X.resize((size_t)n * nr);
        e->d_R.resize((size_t)n * nr);
        e->d_terms.resize((size_t)3 * n);
        const size_t tot = (size_t)n * nr;
        hipLaunchKernelGGL(port_rhs_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, e->d_B.ptr, ne, nv,
                           e->d_P.ptr, ret ? e->d_Pw.ptr : nullptr, e->d_ports.ptr, nr);
        std::vector<double> terms((size_t)3 * n);
        double worst = 0.0;
        auto backward_error = [&](int k) {
            complex_residual(e->d_Z0.ptr, n, n, e->d_X.ptr + (size_t)k * n, e->d_B.ptr + (size_t)k * n,
                             e->d_R.ptr + (size_t)k * n, e->d_terms.ptr, s);
            e->d_terms.download(terms.data(), terms.size(), s);
            e->stream.sync();
            double summary[7];
            residual_summary(terms, n, ne, summary);
            return summary[2];
        };
        // solve with the FP32 factor and refine against the FP64 matrix:
        // x <- x - A^{-1} r, r = A x - b, until the target or the step cap.
        // The component-wise backward error need not fall at every step (the
        // first correction often barely changes it at low frequencies), so
        // refinement stops only after two steps without a new best, and the
        // best iterate is kept. Returns the worst backward error over the
        // right-hand sides.
        e->d_Xbest.resize(n);
        auto solve_refine = [&]() {
            HIP_CHECK(hipMemcpyAsync(e->d_X.ptr, e->d_B.ptr, sizeof(double2) * tot, hipMemcpyDeviceToDevice, s));
            lu_solve_permuted<float2, double2>(e->d_Z32.ptr, n, n, e->d_perm.ptr, e->d_row.ptr, e->d_col.ptr,
                                               e->d_X.ptr, nr, *e->lu);
            double w = 0.0;
            for (int k = 0; k < nr; ++k) {
                double2 *xk = e->d_X.ptr + (size_t)k * n;
                double eta = backward_error(k), best = eta;
                int since_best = 0;
                bool saved = false;
                while (eta > e->opt.refinement_target && report.refinement_steps < e->opt.max_refinement_steps) {
                    lu_solve_permuted<float2, double2>(e->d_Z32.ptr, n, n, e->d_perm.ptr, e->d_row.ptr,
                                                       e->d_col.ptr, e->d_R.ptr + (size_t)k * n, 1, *e->lu);
                    if (since_best == 0 && !saved) {
                        // keep the current best before stepping away from it
                        HIP_CHECK(hipMemcpyAsync(e->d_Xbest.ptr, xk, sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
                        saved = true;
                    }
                    hipLaunchKernelGGL(axpy_sub_kernel, dim3((n + 255) / 256), dim3(256), 0, s, xk,
                                       e->d_R.ptr + (size_t)k * n, (size_t)n);
                    ++report.refinement_steps;
                    eta = backward_error(k);
                    if (eta < best) {
                        best = eta;
                        since_best = 0;
                        saved = false;
                    } else if (++since_best >= 2) {
                        break;
                    }
                }
                if (eta > best) {
                    HIP_CHECK(hipMemcpyAsync(xk, e->d_Xbest.ptr, sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
                    eta = best;
                }
                w = std::fmax(w, eta);
            }
            return w;
        };
        if (mode == PEEC_SOLVE_MIXED) {
            worst = solve_refine();
            if (ordered && !(worst <= 1e-3 * e->opt.backward_error_gate)) {
                // the reused row order is not good enough at this point:
                // refactor with partial pivoting and refine again
                report.pivot_refactor = 1;
                report.reuse_backward_error = worst;
                convert_to_f32(e->d_Z.ptr, e->d_Z32.ptr, nn, s);
                st = lu_factor<float2>(e->d_Z32.ptr, n, n, e->d_ipiv.ptr, limit, *e->lu, e->opt.block_size);
                if (st.code != 0) return factor_error(st);
                keep_order();
                worst = solve_refine();
            }
            if (!(worst <= e->opt.backward_error_gate)) {
                // refinement stalled above the gate: refactor in FP64
                report.used_fp64_fallback = 1;
                st = lu_factor<double2>(e->d_Z.ptr, n, n, e->d_ipiv.ptr, limit, *e->lu, e->opt.block_size);
                if (st.code != 0) return factor_error(st);
                keep_order();
                HIP_CHECK(hipMemcpyAsync(e->d_X.ptr, e->d_B.ptr, sizeof(double2) * tot, hipMemcpyDeviceToDevice, s));
                lu_solve_permuted<double2, double2>(e->d_Z.ptr, n, n, e->d_perm.ptr, e->d_row.ptr, e->d_col.ptr,
                                                    e->d_X.ptr, nr, *e->lu);
                worst = 0.0;
                for (int k = 0; k < nr; ++k) worst = std::fmax(worst, backward_error(k));
            }
            HIP_CHECK(hipEventRecord(e->ev[4], s));
        } else {
            HIP_CHECK(hipMemcpyAsync(e->d_X.ptr, e->d_B.ptr, sizeof(double2) * tot, hipMemcpyDeviceToDevice, s));
            lu_solve_permuted<double2, double2>(e->d_Z.ptr, n, n, e->d_perm.ptr, e->d_row.ptr, e->d_col.ptr,
                                                e->d_X.ptr, nr, *e->lu);
            HIP_CHECK(hipEventRecord(e->ev[4], s));
            for (int k = 0; k < nr; ++k) worst = std::fmax(worst, backward_error(k));
        }
        HIP_CHECK(hipEventRecord(e->ev[5], s));
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
        e->timer.resolve();
        float ms[5];
        for (int i = 0; i < 5; ++i) HIP_CHECK(hipEventElapsedTime(&ms[i], e->ev[i], e->ev[i + 1]));
        report.ms_assembly = ms[0];
        report.ms_fill = ms[1];
        report.ms_factor = ms[2];
        report.ms_solve = ms[3];
        report.ms_check = ms[4];
        report.backward_error = worst;
        if (rep) *rep = report;
        if (!(worst <= e->opt.backward_error_gate)) {
            PEEC_ERR("backward error %.3e above the gate %.1e at f = %.6g Hz.", worst, e->opt.backward_error_gate, f);
            return -3;
        }
        return 0;
    } catch (const HipError &err) {
        PEEC_ERR("%s", err.what.c_str());
        if (rep) *rep = report;
        return -1;
    }
}

}  // extern "C"

// Time-domain PEEC by the theta scheme (transient.h, transient.c:385-2225 of
// the reference) and the operators it needs: the self capacitances and node
// potentials of potential.h, A S and A P A^T of incidence.h, and the
// resistive shunt of internal_circuit.h.
//
// A transient system keeps device copies of its operators (L, R, A S, F,
// the edge end points read from A, the node -> edge incidence as CSR and the
// shunt vectors) with its own stream, in a side table keyed by the caller's
// TransientSystem. The O(N^2) system matrix is filled on the GPU and copied
// to the caller-visible host array; the factor and the solves are the GPU
// real LU of lu.h; a step's right-hand side (the L I, A S V and A^T I
// products) is formed on the GPU from the uploaded state vectors.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/incidence.h"
#include "../../include/internal_circuit.h"
#include "../../include/lu.h"
#include "../../include/matrix.h"
#include "../../include/potential.h"
#include "../../include/transient.h"
#include "common.h"

using namespace peec;

namespace {

// ---------------------------------------------------------------------------
// kernels

// Edge end points from the rows of the dense incidence matrix: the last -1
// and the last +1 of a row (as incidence.c:653-682 scans it); a row with any
// other nonzero value, or without both, sets *bad.
__global__ __launch_bounds__(256) void edge_ends_kernel(const double *A, int ne, int nv, int *ends, int *bad) {
    __shared__ int s_minus, s_plus, s_other;
    const int e = blockIdx.x;
    if (threadIdx.x == 0) {
        s_minus = -1;
        s_plus = -1;
        s_other = 0;
    }
    __syncthreads();
    const double *row = A + (size_t)e * nv;
    int mi = -1, pl = -1, other = 0;
    for (int j = threadIdx.x; j < nv; j += 256) {
        const double a = row[j];
        if (a == -1.0) mi = j;
        else if (a == 1.0) pl = j;
        else if (a != 0.0) other = 1;
    }
    if (mi >= 0) atomicMax(&s_minus, mi);
    if (pl >= 0) atomicMax(&s_plus, pl);
    if (other) atomicOr(&s_other, 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        ends[2 * e] = s_minus;
        ends[2 * e + 1] = s_plus;
        if (s_other || s_minus < 0 || s_plus < 0) atomicOr(bad, 1);
    }
}

// AS[e, a] = (P[b_e, a] - P[a_e, a]) / P_aa (incidence.c:696-715); a
// non-positive or non-finite P_aa gives NaN and sets *bad
__global__ void as_kernel(const double *P, const int *ends, int ne, int nv, double *AS, int *bad) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)ne * nv) return;
    const int e = (int)(idx / nv), a = (int)(idx % nv);
    const int mi = ends[2 * e], pl = ends[2 * e + 1];
    const double paa = P[(size_t)a * nv + a];
    if (!isfinite(paa) || paa <= 0.0 || mi < 0 || pl < 0) {
        AS[idx] = NAN;
        atomicOr(bad, 1);
        return;
    }
    AS[idx] = (P[(size_t)pl * nv + a] - P[(size_t)mi * nv + a]) / paa;
}

// K[e, f] = P[a,c] - P[a,d] - P[b,c] + P[b,d] for e <= f, mirrored below the
// diagonal (incidence.c:242-330 forms the upper triangle and copies it)
__global__ void apat_kernel(const double *P, const int *edges, int ne, int nv, double *K) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)ne * ne) return;
    const int r = (int)(idx / ne), q = (int)(idx % ne), e = min(r, q), f = max(r, q);
    const size_t a = edges[2 * e], b = edges[2 * e + 1], c = edges[2 * f], d = edges[2 * f + 1];
    K[idx] = P[a * nv + c] - P[a * nv + d] - P[b * nv + c] + P[b * nv + d];
}

// phi_j = sum_a P[j, a] V_c[a] / P_aa (potential.c), one workgroup per row
__global__ __launch_bounds__(256) void full_potential_kernel(const double *P, const double *vc, int nv, double *phi,
                                                             int *bad) {
    __shared__ double red[256];
    const int j = blockIdx.x;
    double s = 0.0;
    for (int a = threadIdx.x; a < nv; a += 256) {
        const double paa = P[(size_t)a * nv + a];
        if (!isfinite(paa) || paa <= 0.0) atomicOr(bad, 1);
        s += P[(size_t)j * nv + a] * vc[a] / paa;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) phi[j] = red[0];
}

// The theta-scheme system matrix (transient.c:1060-1419, stages 1 and 2):
//   (I, I)   L/dt + theta diag(R)      (I, V)  theta A S
//   (V, I)   -theta A^T                (V, V)  diag(F)/dt
//   stage 2: (V, s) = -theta b_R, (s, V) = shunt voltage operator, (s, s) = R_shunt
__global__ __launch_bounds__(256) void transient_fill_kernel(double *M, int ne, int nv, int nt, const double *L,
                                                             const double *R, const double *AS, const double *F,
                                                             const int *ends, const double *bR, const double *cS,
                                                             double r_shunt, double dt, double theta, int chunks) {
    const int row = blockIdx.x / chunks;
    const int col = (blockIdx.x % chunks) * 256 + threadIdx.x;
    if (row >= nt || col >= nt) return;
    double v = 0.0;
    if (row < ne) {
        if (col < ne) {
            v = L[(size_t)row * ne + col] / dt;
            if (row == col) v += theta * R[row];
        } else if (col < ne + nv) {
            v = theta * AS[(size_t)row * nv + (col - ne)];
        }
    } else if (row < ne + nv) {
        const int j = row - ne;
        if (col < ne) {
            const double a = j == ends[2 * col + 1] ? 1.0 : (j == ends[2 * col] ? -1.0 : 0.0);
            v = -theta * a;
        } else if (col < ne + nv) {
            v = (col - ne) == j ? F[j] / dt : 0.0;
        } else {
            v = -theta * bR[j];
        }
    } else {
        if (col >= ne && col < ne + nv) v = cS[col - ne];
        else if (col == ne + nv) v = r_shunt;
    }
    M[(size_t)row * nt + col] = v;
}

// The step's right-hand side (transient.c:1421-1749), one workgroup per row:
//   edge rows  sum_j L_ij I_j / dt - (1-theta) (R_i I_i + sum_a AS_ia V_a + U_n,i) - theta U_np1,i
//   node rows  F_j V_j / dt + (1-theta) (sum_e A_ej I_e + s_n,j + b_R,j I_shunt) + theta s_np1,j
//   shunt row  0
// A^T I from the node -> edge CSR (signed, edges in ascending order).
__global__ __launch_bounds__(256) void transient_rhs_kernel(double *rhs, int ne, int nv, int nt, const double *L,
                                                            const double *R, const double *AS, const double *F,
                                                            const int *nptr, const int *nedge, const double *nsign,
                                                            const double *I, const double *V, const double *sn,
                                                            const double *snp1, const double *un, const double *unp1,
                                                            const double *bR, double ishunt, double dt,
                                                            double theta) {
    __shared__ double red[2][256];
    const int row = blockIdx.x;
    const double omt = 1.0 - theta;
    double a = 0.0, b = 0.0;
    if (row < ne) {
        for (int j = threadIdx.x; j < ne; j += 256) a += L[(size_t)row * ne + j] * I[j];
        for (int j = threadIdx.x; j < nv; j += 256) b += AS[(size_t)row * nv + j] * V[j];
    } else if (row < ne + nv) {
        const int j = row - ne;
        for (int k = nptr[j] + threadIdx.x; k < nptr[j + 1]; k += 256) a += nsign[k] * I[nedge[k]];
    }
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = b;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) {
            red[0][threadIdx.x] += red[0][threadIdx.x + k];
            red[1][threadIdx.x] += red[1][threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double v;
    if (row < ne) {
        v = red[0][0] / dt - omt * R[row] * I[row] - omt * red[1][0];
        if (un) v -= omt * un[row];
        if (unp1) v -= theta * unp1[row];
    } else if (row < ne + nv) {
        const int j = row - ne;
        v = F[j] * V[j] / dt + omt * red[0][0];
        if (sn) v += omt * sn[j];
        if (snp1) v += theta * snp1[j];
        if (bR) v += omt * bR[j] * ishunt;
    } else {
        v = 0.0;
    }
    rhs[row] = v;
}

// ---------------------------------------------------------------------------
// device side of a transient system

struct TransientDevice {
    Stream stream;
    int ne = 0, nv = 0, nt = 0;
    DeviceBuffer<double> L, R, AS, F, bR, cS, I, V, sn, snp1, un, unp1, rhs;
    DeviceBuffer<int> ends, nptr, nedge;
    DeviceBuffer<double> nsign;
};

std::mutex g_tmu;
std::unordered_map<const TransientSystem *, std::unique_ptr<TransientDevice>> g_transients;

TransientDevice *device_of(const TransientSystem *s) {
    std::lock_guard<std::mutex> lock(g_tmu);
    auto it = g_transients.find(s);
    return it == g_transients.end() ? nullptr : it->second.get();
}

void forget_device(const TransientSystem *s) {
    std::lock_guard<std::mutex> lock(g_tmu);
    g_transients.erase(s);
}

double theta_of(TransientScheme s) {
    return s == TRANSIENT_BACKWARD_EULER ? 1.0 : s == TRANSIENT_TRAPEZOIDAL ? 0.5 : NAN;
}

bool scheme_ok(TransientScheme s) { return s == TRANSIENT_BACKWARD_EULER || s == TRANSIENT_TRAPEZOIDAL; }

// transient.c:210-284
int validate_common(const InductanceMatrix *L, const ResistanceMatrix *R, const IncidenceMatrix *A,
                    const PotentialMatrix *P, double dt) {
    if (!L || !R || !A || !P) {
        PEEC_ERR("transient PEEC matrix pointer is NULL.");
        return -1;
    }
    if (!std::isfinite(dt) || dt <= 0.0) {
        PEEC_ERR("transient dt must be positive.");
        return -1;
    }
    if (!L->data || !R->diagonal || !A->data || !P->data) {
        PEEC_ERR("transient PEEC matrix data is NULL.");
        return -1;
    }
    if (L->n == 0 || P->n == 0) return -1;
    if (R->n != L->n) {
        PEEC_ERR("R dimension does not match L dimension.");
        return -1;
    }
    if (A->rows != L->n || A->cols != P->n) {
        PEEC_ERR("A dimensions are inconsistent with L and P.");
        return -1;
    }
    return 0;
}

// Edge end points of A on the device (and *bad for an invalid row)
void edge_ends(const double *dA, int ne, int nv, int *d_ends, int *d_bad, hipStream_t s) {
    hipLaunchKernelGGL(edge_ends_kernel, dim3(ne), dim3(256), 0, s, dA, ne, nv, d_ends, d_bad);
    HIP_CHECK(hipGetLastError());
}

// The operators of a new system on its device: L, R, P -> A S, F, ends and
// the node -> edge CSR; the host F and AS of the caller's struct are filled.
int build_device(TransientSystem *sys, TransientDevice &d) {
    const int ne = (int)sys->n_edges, nv = (int)sys->n_nodes;
    hipStream_t s = d.stream;
    d.ne = ne;
    d.nv = nv;
    d.nt = (int)sys->n_total;
    DeviceBuffer<double> dA, dP;
    DeviceBuffer<int> bad(1);
    HIP_CHECK(hipMemsetAsync(bad.ptr, 0, sizeof(int), s));
    dA.upload(sys->A->data, (size_t)ne * nv, s);
    dP.upload(sys->P->data, (size_t)nv * nv, s);
    d.ends.resize(2 * (size_t)ne);
    edge_ends(dA.ptr, ne, nv, d.ends.ptr, bad.ptr, s);
    d.AS.resize((size_t)ne * nv);
    const size_t tot = (size_t)ne * nv;
    hipLaunchKernelGGL(as_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, dP.ptr, d.ends.ptr, ne, nv,
                       d.AS.ptr, bad.ptr);
    HIP_CHECK(hipGetLastError());
    int b = 0;
    bad.download(&b, 1, s);
    std::vector<int> ends(2 * (size_t)ne);
    d.ends.download(ends.data(), ends.size(), s);
    d.stream.sync();
    if (b) {
        PEEC_ERR("cannot build transient A S operator.");
        return -1;
    }
    if (matrix_alloc(&sys->AS, (size_t)ne, (size_t)nv) != 0) return -1;
    d.AS.download(sys->AS.data, tot, s);
    d.L.upload(sys->L->data, (size_t)ne * ne, s);
    d.R.upload(sys->R->diagonal, (size_t)ne, s);
    d.F.upload(sys->F, (size_t)nv, s);
    // node -> edge incidence (A^T) as signed CSR, edges ascending per node
    std::vector<int> cnt(nv + 1, 0);
    for (int e = 0; e < ne; ++e) {
        ++cnt[ends[2 * e] + 1];
        ++cnt[ends[2 * e + 1] + 1];
    }
    for (int j = 0; j < nv; ++j) cnt[j + 1] += cnt[j];
    std::vector<int> edge(2 * (size_t)ne), fill(cnt.begin(), cnt.end() - 1);
    std::vector<double> sign(2 * (size_t)ne);
    for (int e = 0; e < ne; ++e) {
        const int k0 = fill[ends[2 * e]]++;
        edge[k0] = e;
        sign[k0] = -1.0;
        const int k1 = fill[ends[2 * e + 1]]++;
        edge[k1] = e;
        sign[k1] = 1.0;
    }
    d.nptr.upload(cnt, s);
    d.nedge.upload(edge, s);
    d.nsign.upload(sign, s);
    d.I.resize((size_t)ne);
    d.V.resize((size_t)nv);
    d.rhs.resize((size_t)d.nt);
    d.stream.sync();
    return 0;
}

int create_common(TransientSystem *sys, const InductanceMatrix *L, const ResistanceMatrix *R,
                  const IncidenceMatrix *A, const PotentialMatrix *P, const InternalCircuitShunt *shunt,
                  TransientStage stage, TransientScheme scheme, double dt) {
    transient_system_free(sys);
    sys->stage = stage;
    sys->scheme = scheme;
    sys->n_edges = L->n;
    sys->n_nodes = P->n;
    sys->n_total = L->n + P->n + (stage == TRANSIENT_STAGE_2 ? 1 : 0);
    sys->dt = dt;
    sys->L = L;
    sys->R = R;
    sys->A = A;
    sys->P = P;
    sys->shunt = stage == TRANSIENT_STAGE_2 ? shunt : nullptr;
    const size_t nt = sys->n_total, nv = sys->n_nodes;
    if (stage == TRANSIENT_STAGE_2) {
        sys->b_R = static_cast<double *>(std::calloc(nv, sizeof(double)));
        sys->shunt_voltage_operator = static_cast<double *>(std::calloc(nv, sizeof(double)));
        if (!sys->b_R || !sys->shunt_voltage_operator) {
            transient_system_free(sys);
            return -1;
        }
    }
    sys->F = static_cast<double *>(std::calloc(nv, sizeof(double)));
    sys->system_matrix = static_cast<double *>(std::calloc(nt * nt, sizeof(double)));
    sys->rhs = static_cast<double *>(std::calloc(nt, sizeof(double)));
    sys->solution = static_cast<double *>(std::calloc(nt, sizeof(double)));
    if (!sys->F || !sys->system_matrix || !sys->rhs || !sys->solution) {
        PEEC_ERR("cannot allocate transient system storage.");
        transient_system_free(sys);
        return -1;
    }
    if (potential_build_self_capacitance(P, sys->F) != 0) {
        PEEC_ERR("cannot build transient F operator.");
        transient_system_free(sys);
        return -1;
    }
    if (!require_device()) {
        transient_system_free(sys);
        return -1;
    }
    try {
        auto d = std::make_unique<TransientDevice>();
        if (build_device(sys, *d) != 0) {
            transient_system_free(sys);
            return -1;
        }
        {
            std::lock_guard<std::mutex> lock(g_tmu);
            g_transients[sys] = std::move(d);
        }
        if ((stage == TRANSIENT_STAGE_2 && transient_build_shunt_operators(sys) != 0) ||
            transient_build_system_matrix(sys) != 0) {
            transient_system_free(sys);
            return -1;
        }
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        transient_system_free(sys);
        return -1;
    }
    sys->initialized = 1;
    sys->factorized = 0;
    return 0;
}

int rhs_common(TransientSystem *sys, const double *I, const double *V, const double *sn, const double *snp1,
               const double *un, const double *unp1, const double *bR, double ishunt) {
    TransientDevice *d = device_of(sys);
    if (!d) {
        PEEC_ERR("transient system has no device state (not created by this library).");
        return -1;
    }
    try {
        hipStream_t s = d->stream;
        const int ne = d->ne, nv = d->nv;
        d->I.upload(I, (size_t)ne, s);
        d->V.upload(V, (size_t)nv, s);
        auto opt = [&](DeviceBuffer<double> &buf, const double *h, int n) -> const double * {
            if (!h) return nullptr;
            buf.upload(h, (size_t)n, s);
            return buf.ptr;
        };
        const double *dsn = opt(d->sn, sn, nv), *dsnp1 = opt(d->snp1, snp1, nv);
        const double *dun = opt(d->un, un, ne), *dunp1 = opt(d->unp1, unp1, ne);
        hipLaunchKernelGGL(transient_rhs_kernel, dim3((unsigned)d->nt), dim3(256), 0, s, d->rhs.ptr, ne, nv, d->nt,
                           d->L.ptr, d->R.ptr, d->AS.ptr, d->F.ptr, d->nptr.ptr, d->nedge.ptr, d->nsign.ptr, d->I.ptr,
                           d->V.ptr, dsn, dsnp1, dun, dunp1, bR ? d->bR.ptr : nullptr, ishunt, sys->dt,
                           theta_of(sys->scheme));
        HIP_CHECK(hipGetLastError());
        d->rhs.download(sys->rhs, (size_t)d->nt, s);
        d->stream.sync();
        return 0;
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        return -1;
    }
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// potential.h additions

int potential_build_self_capacitance(const PotentialMatrix *P, double *F) {
    if (!P || !P->data || P->n == 0 || !F) return -1;
    for (size_t j = 0; j < P->n; ++j) {
        const double d = P->data[j * P->n + j];
        if (!std::isfinite(d) || d <= 0.0) {
            PEEC_ERR("invalid diagonal P[%zu,%zu] in potential_build_self_capacitance().", j, j);
            return -1;
        }
        F[j] = 1.0 / d;
    }
    return 0;
}

int potential_self_voltage_to_charge(const PotentialMatrix *P, const double *V_c, double *charge) {
    if (!P || !P->data || P->n == 0 || !V_c || !charge) return -1;
    for (size_t j = 0; j < P->n; ++j) {
        const double d = P->data[j * P->n + j];
        if (!std::isfinite(d) || d <= 0.0) return -1;
        charge[j] = V_c[j] / d;
    }
    return 0;
}

int potential_self_voltage_to_full_potential(const PotentialMatrix *P, const double *V_c, double *phi) {
    if (!P || !P->data || P->n == 0 || !V_c || !phi) return -1;
    if (P->n > (size_t)INT32_MAX) return -1;
    if (!require_device()) return -1;
    try {
        Stream s;
        const int nv = (int)P->n;
        DeviceBuffer<double> dP, dv, dphi((size_t)nv);
        DeviceBuffer<int> bad(1);
        HIP_CHECK(hipMemsetAsync(bad.ptr, 0, sizeof(int), s));
        dP.upload(P->data, (size_t)nv * nv, s);
        dv.upload(V_c, (size_t)nv, s);
        hipLaunchKernelGGL(full_potential_kernel, dim3(nv), dim3(256), 0, s, dP.ptr, dv.ptr, nv, dphi.ptr, bad.ptr);
        HIP_CHECK(hipGetLastError());
        int b = 0;
        bad.download(&b, 1, s);
        dphi.download(phi, (size_t)nv, s);
        s.sync();
        return b ? -1 : 0;
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        return -1;
    }
}

// ---------------------------------------------------------------------------
// incidence.h additions

double incidence_check_row_sum(const IncidenceMatrix *A) {
    if (!A || !A->data) return -1.0;
    double worst = 0.0;
    for (size_t r = 0; r < A->rows; ++r) {
        double sum = 0.0;
        for (size_t c = 0; c < A->cols; ++c) sum += A->data[r * A->cols + c];
        worst = std::fmax(worst, std::fabs(sum));
    }
    return worst;
}

void incidence_print_info(const IncidenceMatrix *A) {
    if (!A) return;
    size_t nz = 0;
    if (A->data)
        for (size_t i = 0; i < A->rows * A->cols; ++i) nz += A->data[i] != 0.0;
    std::printf("\n========================================\nINCIDENCE MATRIX\n========================================\n");
    std::printf("Rows Ne : %zu\nCols Nv : %zu\nNonzero : %zu\nExpected: %zu\n", A->rows, A->cols, nz, 2 * A->rows);
    std::printf("========================================\n");
}

int incidence_build_APAT(const Mesh *mesh, const PotentialMatrix *P, Matrix *K, int threads) {
    if (!mesh || !P || !K || !P->data || threads < 1) return -1;
    if (P->n != mesh->n_nodes) {
        PEEC_ERR("incompatible P matrix size in incidence_build_APAT.");
        return -1;
    }
    const size_t ne = mesh->n_edges, nv = mesh->n_nodes;
    if (ne == 0 || nv == 0 || !mesh->edges) return -1;
    if (matrix_alloc(K, ne, ne) != 0) return -1;
    if (!require_device()) {
        matrix_free(K);
        return -1;
    }
    try {
        Stream s;
        DeviceBuffer<double> dP, dK(ne * ne);
        DeviceBuffer<int> de;
        dP.upload(P->data, nv * nv, s);
        de.upload(mesh->edges, 2 * ne, s);
        hipLaunchKernelGGL(apat_kernel, dim3((unsigned)((ne * ne + 255) / 256)), dim3(256), 0, s, dP.ptr, de.ptr,
                           (int)ne, (int)nv, dK.ptr);
        HIP_CHECK(hipGetLastError());
        dK.download(K->data, ne * ne, s);
        s.sync();
        return 0;
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        matrix_free(K);
        return -1;
    }
}

int incidence_build_AS(const IncidenceMatrix *A, const PotentialMatrix *P, Matrix *AS, int threads) {
    if (!A || !P || !AS || !A->data || !P->data || threads < 1) return -1;
    if (A->cols != P->n || A->rows == 0 || A->cols == 0) {
        PEEC_ERR("incompatible dimensions in incidence_build_AS().");
        return -1;
    }
    const size_t ne = A->rows, nv = A->cols;
    if (ne > (size_t)INT32_MAX || nv > (size_t)INT32_MAX) return -1;
    if (matrix_alloc(AS, ne, nv) != 0) {
        PEEC_ERR("cannot allocate A S matrix.");
        return -1;
    }
    if (!require_device()) {
        matrix_free(AS);
        return -1;
    }
    try {
        Stream s;
        DeviceBuffer<double> dA, dP, dAS(ne * nv);
        DeviceBuffer<int> ends(2 * ne), bad(1);
        HIP_CHECK(hipMemsetAsync(bad.ptr, 0, sizeof(int), s));
        dA.upload(A->data, ne * nv, s);
        dP.upload(P->data, nv * nv, s);
        edge_ends(dA.ptr, (int)ne, (int)nv, ends.ptr, bad.ptr, s);
        hipLaunchKernelGGL(as_kernel, dim3((unsigned)((ne * nv + 255) / 256)), dim3(256), 0, s, dP.ptr, ends.ptr,
                           (int)ne, (int)nv, dAS.ptr, bad.ptr);
        HIP_CHECK(hipGetLastError());
        int b = 0;
        bad.download(&b, 1, s);
        dAS.download(AS->data, ne * nv, s);
        s.sync();
        if (b) {
            PEEC_ERR("cannot build A S matrix.");
            matrix_free(AS);
            return -1;
        }
        return 0;
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        matrix_free(AS);
        return -1;
    }
}

// ---------------------------------------------------------------------------
// internal_circuit.h (host)

void internal_circuit_shunt_default(InternalCircuitShunt *sh) {
    if (!sh) return;
    std::memset(sh, 0, sizeof(*sh));
    sh->type = INTERNAL_CIRCUIT_RESISTIVE_SHUNT;
    sh->resistance = 1000.0;
    sh->mapping_error_a = NAN;
    sh->mapping_error_b = NAN;
    sh->initialized = 1;
}

int internal_circuit_find_nearest_node(const Mesh *mesh, const double p[3], size_t *node, double *distance) {
    if (!mesh || !p || !node || !mesh->xyz || mesh->n_nodes == 0) return -1;
    size_t best = 0;
    double bd = 0.0;
    for (size_t j = 0; j < mesh->n_nodes; ++j) {
        const double *r = mesh->xyz + 3 * j;
        const double dx = r[0] - p[0], dy = r[1] - p[1], dz = r[2] - p[2];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (j == 0 || d2 < bd) {
            bd = d2;
            best = j;
        }
    }
    *node = best;
    if (distance) *distance = std::sqrt(bd);
    return 0;
}

int internal_circuit_shunt_set_nodes(InternalCircuitShunt *sh, const Mesh *mesh, size_t a, size_t b) {
    if (!sh || !mesh || !mesh->xyz || a >= mesh->n_nodes || b >= mesh->n_nodes) return -1;
    if (a == b) {
        PEEC_ERR("internal shunt terminals map to the same node.");
        return -1;
    }
    sh->node_a = a;
    sh->node_b = b;
    for (int c = 0; c < 3; ++c) {
        sh->requested_a[c] = sh->mapped_a[c] = mesh->xyz[3 * a + c];
        sh->requested_b[c] = sh->mapped_b[c] = mesh->xyz[3 * b + c];
    }
    sh->mapping_error_a = 0.0;
    sh->mapping_error_b = 0.0;
    sh->nodes_mapped = 1;
    return 0;
}

int internal_circuit_shunt_map_coordinates(InternalCircuitShunt *sh, const Mesh *mesh, const double pa[3],
                                           const double pb[3]) {
    if (!sh || !mesh || !pa || !pb) return -1;
    size_t a = 0, b = 0;
    double ea = 0.0, eb = 0.0;
    if (internal_circuit_find_nearest_node(mesh, pa, &a, &ea) != 0 ||
        internal_circuit_find_nearest_node(mesh, pb, &b, &eb) != 0)
        return -1;
    if (a == b) {
        PEEC_ERR("the two internal shunt terminals mapped to the same mesh node %zu.", a);
        return -1;
    }
    sh->node_a = a;
    sh->node_b = b;
    for (int c = 0; c < 3; ++c) {
        sh->requested_a[c] = pa[c];
        sh->requested_b[c] = pb[c];
        sh->mapped_a[c] = mesh->xyz[3 * a + c];
        sh->mapped_b[c] = mesh->xyz[3 * b + c];
    }
    sh->mapping_error_a = ea;
    sh->mapping_error_b = eb;
    sh->nodes_mapped = 1;
    return 0;
}

int internal_circuit_shunt_validate(const InternalCircuitShunt *sh, size_t n_nodes) {
    if (!sh || !sh->initialized || n_nodes == 0) return -1;
    if (sh->type != INTERNAL_CIRCUIT_RESISTIVE_SHUNT) {
        PEEC_ERR("unsupported internal circuit type.");
        return -1;
    }
    if (!sh->nodes_mapped) {
        PEEC_ERR("internal shunt nodes have not been mapped.");
        return -1;
    }
    if (sh->node_a >= n_nodes || sh->node_b >= n_nodes) {
        PEEC_ERR("internal shunt node is outside mesh.");
        return -1;
    }
    if (sh->node_a == sh->node_b) {
        PEEC_ERR("internal shunt terminals must be different.");
        return -1;
    }
    if (!std::isfinite(sh->resistance) || sh->resistance < 0.0) {
        PEEC_ERR("internal shunt resistance must be >= 0 Ohm.");
        return -1;
    }
    return 0;
}

int internal_circuit_build_incidence_vector(const InternalCircuitShunt *sh, size_t n_nodes, double *bR) {
    if (!sh || !bR || internal_circuit_shunt_validate(sh, n_nodes) != 0) return -1;
    std::memset(bR, 0, n_nodes * sizeof(double));
    bR[sh->node_a] = -1.0;
    bR[sh->node_b] = +1.0;
    return 0;
}

double internal_circuit_shunt_voltage(const InternalCircuitShunt *sh, const double *v, size_t n_nodes) {
    if (!sh || !v || internal_circuit_shunt_validate(sh, n_nodes) != 0) return NAN;
    return v[sh->node_a] - v[sh->node_b];
}

double internal_circuit_shunt_current_from_voltage(const InternalCircuitShunt *sh, const double *v, size_t n_nodes) {
    if (!sh || !v || internal_circuit_shunt_validate(sh, n_nodes) != 0 || sh->resistance == 0.0) return NAN;
    return internal_circuit_shunt_voltage(sh, v, n_nodes) / sh->resistance;
}

double internal_circuit_shunt_residual(const InternalCircuitShunt *sh, const double *v, size_t n_nodes,
                                       double current) {
    if (!sh || !v || !std::isfinite(current) || internal_circuit_shunt_validate(sh, n_nodes) != 0) return NAN;
    return internal_circuit_shunt_voltage(sh, v, n_nodes) - sh->resistance * current;
}

void internal_circuit_shunt_print_info(const InternalCircuitShunt *sh) {
    if (!sh) return;
    std::printf("\n============================================================\nINTERNAL CIRCUIT / SHUNT\n");
    std::printf("============================================================\n");
    std::printf("Type : RESISTIVE SHUNT\nPositive current : A -> B\n");
    std::printf("A node : %zu\nA requested xyz : %.9e %.9e %.9e m\nA mesh xyz : %.9e %.9e %.9e m\n", sh->node_a,
                sh->requested_a[0], sh->requested_a[1], sh->requested_a[2], sh->mapped_a[0], sh->mapped_a[1],
                sh->mapped_a[2]);
    std::printf("A mapping error : %.9e m\n\n", sh->mapping_error_a);
    std::printf("B node : %zu\nB requested xyz : %.9e %.9e %.9e m\nB mesh xyz : %.9e %.9e %.9e m\n", sh->node_b,
                sh->requested_b[0], sh->requested_b[1], sh->requested_b[2], sh->mapped_b[0], sh->mapped_b[1],
                sh->mapped_b[2]);
    std::printf("B mapping error : %.9e m\n\nResistance : %.9e Ohm\n", sh->mapping_error_b, sh->resistance);
    std::printf("Shunt model : %s\n\n", sh->resistance == 0.0 ? "IDEAL SHORT CIRCUIT" : "RESISTIVE");
    std::printf("Branch equation : b_R^T V + R I_shunt = 0\nVoltage definition : V_shunt = V_A - V_B\n");
    std::printf("============================================================\n");
}

// ---------------------------------------------------------------------------
// transient.h

void transient_system_init(TransientSystem *sys) {
    if (!sys) return;
    std::memset(sys, 0, sizeof(*sys));
    lu_init(&sys->lu);
}

void transient_system_free(TransientSystem *sys) {
    if (!sys) return;
    forget_device(sys);
    std::free(sys->F);
    matrix_free(&sys->AS);
    std::free(sys->b_R);
    std::free(sys->shunt_voltage_operator);
    std::free(sys->slot_mass_matrix);
    std::free(sys->slot_voltage_operators);
    std::free(sys->system_matrix);
    std::free(sys->rhs);
    std::free(sys->solution);
    lu_free(&sys->lu);
    std::memset(sys, 0, sizeof(*sys));
    matrix_init(&sys->AS);
    lu_init(&sys->lu);
}

int transient_system_create_stage1(TransientSystem *sys, const InductanceMatrix *L, const ResistanceMatrix *R,
                                   const IncidenceMatrix *A, const PotentialMatrix *P, TransientScheme scheme,
                                   double dt) {
    if (!sys || !scheme_ok(scheme)) return -1;
    if (validate_common(L, R, A, P, dt) != 0) return -1;
    return create_common(sys, L, R, A, P, nullptr, TRANSIENT_STAGE_1, scheme, dt);
}

int transient_system_create_stage2(TransientSystem *sys, const InductanceMatrix *L, const ResistanceMatrix *R,
                                   const IncidenceMatrix *A, const PotentialMatrix *P,
                                   const InternalCircuitShunt *shunt, TransientScheme scheme, double dt) {
    if (!sys || !shunt || !scheme_ok(scheme)) return -1;
    if (validate_common(L, R, A, P, dt) != 0) return -1;
    if (internal_circuit_shunt_validate(shunt, P->n) != 0) return -1;
    return create_common(sys, L, R, A, P, shunt, TRANSIENT_STAGE_2, scheme, dt);
}

// transient.c:940-1018: b_R and c_a = (P[B, a] - P[A, a]) / P_aa
int transient_build_shunt_operators(TransientSystem *sys) {
    if (!sys || (sys->stage != TRANSIENT_STAGE_2 && sys->stage != TRANSIENT_SLOT_CELLS) || !sys->P || !sys->P->data ||
        !sys->shunt || !sys->b_R || !sys->shunt_voltage_operator)
        return -1;
    const size_t nv = sys->n_nodes;
    if (internal_circuit_build_incidence_vector(sys->shunt, nv, sys->b_R) != 0) return -1;
    const size_t a = sys->shunt->node_a, b = sys->shunt->node_b;
    const double *P = sys->P->data;
    for (size_t k = 0; k < nv; ++k) {
        const double pkk = P[k * nv + k];
        if (!std::isfinite(pkk) || pkk <= 0.0) return -1;
        sys->shunt_voltage_operator[k] = (P[b * nv + k] - P[a * nv + k]) / pkk;
    }
    TransientDevice *d = device_of(sys);
    if (d) {
        try {
            d->bR.upload(sys->b_R, nv, d->stream);
            d->cS.upload(sys->shunt_voltage_operator, nv, d->stream);
            d->stream.sync();
        } catch (const HipError &e) {
            PEEC_ERR("%s", e.what.c_str());
            return -1;
        }
    }
    return 0;
}

int transient_build_system_matrix(TransientSystem *sys) {
    if (!sys || !sys->L || !sys->R || !sys->A || !sys->P || !sys->F || !sys->AS.data || !sys->system_matrix ||
        !scheme_ok(sys->scheme) || !std::isfinite(sys->dt) || sys->dt <= 0.0)
        return -1;
    if (sys->stage == TRANSIENT_SLOT_CELLS) {
        PEEC_ERR("the slot-cell transient stage is not provided by this library.");
        return -1;
    }
    if (sys->stage == TRANSIENT_STAGE_2 && (!sys->shunt || !sys->b_R || !sys->shunt_voltage_operator)) return -1;
    TransientDevice *d = device_of(sys);
    if (!d) {
        PEEC_ERR("transient system has no device state (not created by this library).");
        return -1;
    }
    try {
        hipStream_t s = d->stream;
        const int nt = d->nt;
        DeviceBuffer<double> M((size_t)nt * nt);
        const int chunks = (nt + 255) / 256;
        const bool st2 = sys->stage == TRANSIENT_STAGE_2;
        hipLaunchKernelGGL(transient_fill_kernel, dim3((unsigned)((size_t)nt * chunks)), dim3(256), 0, s, M.ptr, d->ne,
                           d->nv, nt, d->L.ptr, d->R.ptr, d->AS.ptr, d->F.ptr, d->ends.ptr,
                           st2 ? d->bR.ptr : nullptr, st2 ? d->cS.ptr : nullptr, st2 ? sys->shunt->resistance : 0.0,
                           sys->dt, theta_of(sys->scheme), chunks);
        HIP_CHECK(hipGetLastError());
        M.download(sys->system_matrix, (size_t)nt * nt, s);
        d->stream.sync();
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        return -1;
    }
    sys->factorized = 0;
    return 0;
}

size_t transient_current_offset(const TransientSystem *) { return 0; }

size_t transient_voltage_offset(const TransientSystem *sys) { return sys ? sys->n_edges : 0; }

size_t transient_slot_current_offset(const TransientSystem *sys) {
    if (!sys || sys->stage != TRANSIENT_SLOT_CELLS) return (size_t)-1;
    return sys->n_edges + sys->n_nodes;
}

// The slot-cell stage (slot inductors as extra unknowns) is not built by this
// library. Its entry points exist so that a program linked against it
// without the reference's transient.c still links, and a slot-cell run fails
// with -1 and an error line instead of running something else.
int transient_system_create_slot_cells(TransientSystem *, const InductanceMatrix *, const ResistanceMatrix *,
                                       const IncidenceMatrix *, const PotentialMatrix *, const SlotCircuit *,
                                       const InternalCircuitShunt *, TransientScheme, double) {
    PEEC_ERR("transient: the slot-cell stage is not provided by this library.");
    return -1;
}

int transient_build_rhs_slot_cells(TransientSystem *, const double *, const double *, const double *, double,
                                   const double *, const double *, const double *, const double *) {
    PEEC_ERR("transient: the slot-cell stage is not provided by this library.");
    return -1;
}

int transient_extract_slot_cells_solution(const TransientSystem *, double *, double *, double *, double *) {
    PEEC_ERR("transient: the slot-cell stage is not provided by this library.");
    return -1;
}

size_t transient_shunt_index(const TransientSystem *sys) {
    if (!sys) return (size_t)-1;
    if (sys->stage == TRANSIENT_STAGE_2) return sys->n_edges + sys->n_nodes;
    if (sys->stage == TRANSIENT_SLOT_CELLS && sys->shunt)
        return transient_slot_current_offset(sys) + sys->n_slot_cells;
    return (size_t)-1;
}

int transient_factorize(TransientSystem *sys, int threads) {
    if (!sys || !sys->initialized || !sys->system_matrix || sys->n_total == 0 || threads < 1) return -1;
    lu_free(&sys->lu);
    lu_init(&sys->lu);
    LUOptions o;
    lu_options_default(&o, threads);
    if (lu_factorize(sys->system_matrix, sys->n_total, &o, &sys->lu) != 0) {
        PEEC_ERR("transient LU factorization failed.");
        sys->factorized = 0;
        return -1;
    }
    sys->factorized = 1;
    return 0;
}

int transient_build_rhs_stage1(TransientSystem *sys, const double *I, const double *V, const double *sn,
                               const double *snp1, const double *un, const double *unp1) {
    if (!sys || !sys->initialized || sys->stage != TRANSIENT_STAGE_1) return -1;
    if (!I || !V || !sys->rhs || !sys->F || !sys->AS.data || !scheme_ok(sys->scheme)) return -1;
    return rhs_common(sys, I, V, sn, snp1, un, unp1, nullptr, 0.0);
}

int transient_build_rhs_stage2(TransientSystem *sys, const double *I, const double *V, double ishunt,
                               const double *sn, const double *snp1, const double *un, const double *unp1) {
    if (!sys || !sys->initialized || sys->stage != TRANSIENT_STAGE_2 || !sys->b_R || !scheme_ok(sys->scheme) ||
        !std::isfinite(ishunt))
        return -1;
    if (!I || !V || !sys->rhs || !sys->F || !sys->AS.data) return -1;
    return rhs_common(sys, I, V, sn, snp1, un, unp1, sys->b_R, ishunt);
}

int transient_solve_step(TransientSystem *sys) {
    if (!sys || !sys->initialized || !sys->factorized || !sys->rhs || !sys->solution) return -1;
    if (lu_solve(&sys->lu, sys->rhs, sys->solution) != 0) {
        PEEC_ERR("transient linear solve failed.");
        return -1;
    }
    return 0;
}

int transient_extract_stage1_solution(const TransientSystem *sys, double *I, double *V) {
    if (!sys || !sys->initialized || sys->stage != TRANSIENT_STAGE_1 || !sys->solution || !I || !V) return -1;
    std::memcpy(I, sys->solution, sys->n_edges * sizeof(double));
    std::memcpy(V, sys->solution + sys->n_edges, sys->n_nodes * sizeof(double));
    return 0;
}

int transient_extract_stage2_solution(const TransientSystem *sys, double *I, double *V, double *ishunt) {
    if (!sys || !sys->initialized || sys->stage != TRANSIENT_STAGE_2 || !sys->solution || !I || !V || !ishunt)
        return -1;
    std::memcpy(I, sys->solution, sys->n_edges * sizeof(double));
    std::memcpy(V, sys->solution + sys->n_edges, sys->n_nodes * sizeof(double));
    *ishunt = sys->solution[transient_shunt_index(sys)];
    return 0;
}

int transient_recover_node_state(const TransientSystem *sys, const double *V, double *phi, double *q) {
    if (!sys || !sys->initialized || !sys->P || !V) return -1;
    if (!phi && !q) return 0;
    if (q && potential_self_voltage_to_charge(sys->P, V, q) != 0) return -1;
    if (phi && potential_self_voltage_to_full_potential(sys->P, V, phi) != 0) return -1;
    return 0;
}

void transient_print_info(const TransientSystem *sys) {
    if (!sys) return;
    static const char *stages[] = {"STAGE 1", "STAGE 2 / SHUNT", "ONE-STAGE / SLOT CELLS"};
    std::printf("\n============================================================\nTRANSIENT PEEC / CES SYSTEM\n");
    std::printf("============================================================\n");
    std::printf("Stage : %s\n", (int)sys->stage >= 0 && (int)sys->stage < 3 ? stages[sys->stage] : "UNKNOWN");
    if (sys->scheme == TRANSIENT_BACKWARD_EULER) std::printf("Time scheme : BACKWARD EULER\nTemporal order : 1\n");
    else if (sys->scheme == TRANSIENT_TRAPEZOIDAL) std::printf("Time scheme : TRAPEZOIDAL RULE\nTemporal order : 2\n");
    else std::printf("Time scheme : UNKNOWN\n");
    std::printf("Node unknown : V_c (self voltage)\nElectrostatic form : F_j=1/P_jj, phi=S*V_c, AS=A*S\n");
    std::printf("dt : %.9e s\nEdge unknowns : %zu\nNode unknowns : %zu\n", sys->dt, sys->n_edges, sys->n_nodes);
    std::printf("Slot unknowns : %zu\nShunt unknowns : %d\n",
                sys->stage == TRANSIENT_SLOT_CELLS ? sys->n_slot_cells : (size_t)0, sys->shunt ? 1 : 0);
    std::printf("System size : %zu x %zu\nLU factorized : %s\n", sys->n_total, sys->n_total,
                sys->factorized ? "YES" : "NO");
    std::printf("============================================================\n");
}

}  // extern "C"

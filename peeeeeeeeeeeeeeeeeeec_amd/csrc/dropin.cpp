// The reference's C call surface for the hot path (inductance.h,
// potential.h, quadrature.h, incidence.h, resistance.h, lu.h, complex_lu.h,
// complex_matrix.h, matrix.h, harmonic.h) plus the new retardation.h.
// Same names, argument meaning, ownership and return codes as the
// reference; every O(N^2)/O(N^3) step runs on the GPU. Host-side work is
// limited to argument checks, allocation of the caller-visible results and
// O(N) bookkeeping.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/complex_lu.h"
#include "../../include/harmonic.h"
#include "../../include/incidence.h"
#include "../../include/inductance.h"
#include "../../include/lu.h"
#include "../../include/potential.h"
#include "../../include/quadrature.h"
#include "../../include/resistance.h"
#include "../../include/retardation.h"
#include "assembly.h"
#include "harmonic_kernels.h"
#include "lu.h"

using namespace peec;

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

bool orders_ok(int f, int n, int t, int s, double nf) {
    return gl_order_supported(f) && gl_order_supported(n) && gl_order_supported(t) && gl_order_supported(s) &&
           std::isfinite(nf) && nf >= 0.0;
}

QuadratureOrders orders_of(const InductanceOptions *o) {
    return {o->far_order, o->near_order, o->touching_order, o->self_order, o->near_factor};
}
QuadratureOrders orders_of(const PotentialOptions *o) {
    return {o->far_order, o->near_order, o->touching_order, o->self_order, o->near_factor};
}

// Device copies of live factorizations, keyed by the host data pointer the
// caller holds (released by *_free).
struct DeviceFactor {
    size_t n = 0;
    DeviceBuffer<double2> zc;
    DeviceBuffer<double> zr;
    DeviceBuffer<int> ipiv;
    DeviceBuffer<double> row, col;
};
std::mutex g_mu;
std::unordered_map<const void *, std::unique_ptr<DeviceFactor>> g_factors;

void forget_factor(const void *key) {
    if (!key) return;
    std::lock_guard<std::mutex> lock(g_mu);
    g_factors.erase(key);
}

std::vector<double> region_pieces(const PatchList &l) {
    std::vector<double> v(12 * l.count);
    for (size_t k = 0; k < l.count; ++k) std::memcpy(&v[12 * k], l.patches[k].points, 12 * sizeof(double));
    return v;
}

template <class T>
T *host_calloc(size_t n) {
    return static_cast<T *>(std::calloc(n ? n : 1, sizeof(T)));
}

}  // namespace

namespace {

// Z from dense host inputs exactly as harmonic.c:240-365 lays it out
__global__ void dense_mna_kernel(double2 *Z, int ne, int nv, const double *L, const double *R, const double *A,
                                 const double *PAT, double omega, int chunks) {
    const int n = ne + nv;
    const int row = blockIdx.x / chunks;
    const int col = (blockIdx.x % chunks) * 256 + threadIdx.x;
    if (row >= n || col >= n) return;
    double2 z;
    if (row < ne) {
        if (col < ne) z = make_double2(row == col ? R[row] : 0.0, omega * L[(size_t)row * ne + col]);
        else z = make_double2(A[(size_t)row * nv + (col - ne)], 0.0);
    } else {
        const int i = row - ne;
        if (col < ne) z = make_double2(-PAT[(size_t)i * ne + col], 0.0);
        else z = make_double2(0.0, (col - ne) == i ? omega : 0.0);
    }
    Z[(size_t)row * n + col] = z;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// inductance.h

void inductance_options_default(InductanceOptions *o, int threads) {
    if (!o) return;
    o->far_order = 4;  // inductance.c:46-97
    o->near_order = 10;
    o->touching_order = 12;
    o->self_order = 8;
    o->near_factor = 0.5;
    o->parallel_threads = threads;
    o->direction_tolerance = 1.0e-10;
}

void inductance_init(InductanceMatrix *L) {
    if (L) std::memset(L, 0, sizeof(*L));
}

void inductance_free(InductanceMatrix *L) {
    if (!L) return;
    std::free(L->data);
    inductance_init(L);
}

void inductance_print_info(const InductanceMatrix *L) {
    if (!L) return;
    std::printf("Inductance matrix: %zu x %zu\n", L->n, L->n);
}

int inductance_compute_pair(const Mesh *mesh, const DualMesh *dual, size_t a, size_t b,
                            const InductanceOptions *o, double *value) {
    if (!mesh || !dual || !o || !value) return -1;
    if (a >= mesh->n_edges || b >= mesh->n_edges) return -1;
    const double *da = &mesh->edge_directions[3 * a], *db = &mesh->edge_directions[3 * b];
    const double dot = da[0] * db[0] + da[1] * db[1] + da[2] * db[2];
    if (std::fabs(dot) < o->direction_tolerance) {  // inductance.c:272-277
        *value = 0.0;
        return 0;
    }
    auto width = [&](size_t e) {
        const double s = dual->edge_region_areas[e], l = mesh->edge_lengths[e];
        return (s <= 0.0 || l <= 0.0) ? 0.0 : s / l;
    };
    const double wa = width(a), wb = width(b);
    if (wa <= 0.0 || wb <= 0.0) return -1;
    if (dual->edge_patches[a].count == 0 || dual->edge_patches[b].count == 0) return -1;
    if (!orders_ok(o->far_order, o->near_order, o->touching_order, o->self_order, o->near_factor)) return -1;
    if (!require_device()) return -1;
    try {
        const double g = region_pair_integral(region_pieces(dual->edge_patches[a]),
                                              region_pieces(dual->edge_patches[b]), a == b, orders_of(o));
        if (!std::isfinite(g)) return -1;
        *value = 1.0e-7 * dot * g / (wa * wb);
        return std::isfinite(*value) ? 0 : -1;
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        return -1;
    }
}

static int assemble_matrix(const Mesh *mesh, const DualMesh *dual, RegionKind kind, const QuadratureOrders &q,
                           double dir_tol, double **out, size_t *n_out) {
    const bool nodes = kind == RegionKind::Node;
    const size_t R = nodes ? mesh->n_nodes : mesh->n_edges;
    for (size_t r = 0; r < R; ++r) {
        const PatchList &l = nodes ? dual->node_patches[r] : dual->edge_patches[r];
        double aux;
        if (nodes) {
            aux = dual->node_region_areas[r];
        } else {
            const double s = dual->edge_region_areas[r], len = mesh->edge_lengths[r];
            aux = (s <= 0.0 || len <= 0.0) ? 0.0 : s / len;
        }
        if (l.count == 0 || !(aux > 0.0)) {
            PEEC_ERR("%s calculation failed for %s %zu, %zu.", nodes ? "potential" : "inductance",
                     nodes ? "nodes" : "edges", r, r);
            return -1;
        }
    }
    if (!require_device()) return -1;
    try {
        Stream s;
        RegionSet rs(mesh, dual, kind, q, 0, dir_tol, s);
        DeviceBuffer<double> d(R * R);
        DeviceBuffer<int> err(1);
        HIP_CHECK(hipMemsetAsync(err.ptr, 0, sizeof(int), s));
        rs.assemble_static(d.ptr, s, err.ptr, nullptr);
        double *h = host_calloc<double>(R * R);
        if (!h) {
            PEEC_ERR("cannot allocate %s matrix.", nodes ? "potential" : "inductance");
            return -1;
        }
        int e = 0;
        d.download(h, R * R, s);
        err.download(&e, 1, s);
        s.sync();
        if (e) {
            std::free(h);
            PEEC_ERR("%s calculation failed (non-finite near-field integral).", nodes ? "potential" : "inductance");
            return -1;
        }
        *out = h;
        *n_out = R;
        return 0;
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        return -1;
    }
}

int inductance_compute_matrix(const Mesh *mesh, const DualMesh *dual, const InductanceOptions *o,
                              InductanceMatrix *L) {
    if (!mesh || !dual || !o || !L) return -1;
    if (mesh->n_edges == 0) return -1;
    inductance_init(L);
    if (!orders_ok(o->far_order, o->near_order, o->touching_order, o->self_order, o->near_factor)) {
        PEEC_ERR("inductance calculation failed for edges 0, 0.");
        return -1;
    }
    return assemble_matrix(mesh, dual, RegionKind::Edge, orders_of(o), o->direction_tolerance, &L->data, &L->n);
}

// ---------------------------------------------------------------------------
// potential.h

void potential_options_default(PotentialOptions *o, int threads) {
    if (!o) return;
    o->far_order = 4;  // potential.c:38-58
    o->near_order = 10;
    o->touching_order = 12;
    o->self_order = 8;
    o->near_factor = 0.5;
    o->parallel_threads = threads;
}

void potential_init(PotentialMatrix *P) {
    if (P) std::memset(P, 0, sizeof(*P));
}

void potential_free(PotentialMatrix *P) {
    if (!P) return;
    std::free(P->data);
    potential_init(P);
}

void potential_print_info(const PotentialMatrix *P) {
    if (!P) return;
    std::printf("Potential matrix: %zu x %zu\n", P->n, P->n);
}

int potential_compute_pair(const Mesh *mesh, const DualMesh *dual, size_t i, size_t j, const PotentialOptions *o,
                           double *value) {
    if (!mesh || !dual || !o || !value) return -1;
    if (i >= mesh->n_nodes || j >= mesh->n_nodes) return -1;
    const double si = dual->node_region_areas[i], sj = dual->node_region_areas[j];
    if (si <= 0.0 || sj <= 0.0) return -1;
    if (dual->node_patches[i].count == 0 || dual->node_patches[j].count == 0) return -1;
    if (!orders_ok(o->far_order, o->near_order, o->touching_order, o->self_order, o->near_factor)) return -1;
    if (!require_device()) return -1;
    try {
        const double g = region_pair_integral(region_pieces(dual->node_patches[i]),
                                              region_pieces(dual->node_patches[j]), i == j, orders_of(o));
        if (!std::isfinite(g)) return -1;
        *value = 8.987551792261171e9 * g / (si * sj);  // potential.c:242-245
        return std::isfinite(*value) ? 0 : -1;
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        return -1;
    }
}

int potential_compute_matrix(const Mesh *mesh, const DualMesh *dual, const PotentialOptions *o,
                             PotentialMatrix *P) {
    if (!mesh || !dual || !o || !P) return -1;
    if (mesh->n_nodes == 0) return -1;
    potential_init(P);
    if (!orders_ok(o->far_order, o->near_order, o->touching_order, o->self_order, o->near_factor)) {
        PEEC_ERR("potential calculation failed for nodes 0, 0.");
        return -1;
    }
    return assemble_matrix(mesh, dual, RegionKind::Node, orders_of(o), 0.0, &P->data, &P->n);
}

// ---------------------------------------------------------------------------
// quadrature.h entry point on the GPU

double quad_patch_kernel_1_over_r(const QuadPatch *a, const QuadPatch *b, int far_order, int near_order,
                                  int touching_order, int self_order, double near_factor) {
    if (!a || !b) return NAN;
    if (!orders_ok(far_order, near_order, touching_order, self_order, near_factor)) return NAN;
    if (!require_device()) return NAN;
    try {
        return patch_pair_integral(a, b, {far_order, near_order, touching_order, self_order, near_factor});
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        return NAN;
    }
}

// ---------------------------------------------------------------------------
// retardation.h (new interface; see the header)

void retardation_options_default(RetardationOptions *o, int threads) {
    if (!o) return;
    inductance_options_default(&o->inductance, threads);
    potential_options_default(&o->potential, threads);
    o->smooth_order = 2;
    o->near_smooth_order = 4;
}

int retardation_compute_matrices(const Mesh *mesh, const DualMesh *dual, double k, const RetardationOptions *o,
                                 ComplexMatrix *L, ComplexMatrix *P) {
    if (!mesh || !dual || !o || !std::isfinite(k) || k < 0.0) return -1;
    if (!gl_order_supported(o->smooth_order) || !gl_order_supported(o->near_smooth_order)) return -1;
    const InductanceOptions *li = &o->inductance;
    const PotentialOptions *pi = &o->potential;
    if (!orders_ok(li->far_order, li->near_order, li->touching_order, li->self_order, li->near_factor) ||
        !orders_ok(pi->far_order, pi->near_order, pi->touching_order, pi->self_order, pi->near_factor))
        return -1;
    if (!require_device()) return -1;
    try {
        Stream s;
        DeviceBuffer<int> err(1);
        HIP_CHECK(hipMemsetAsync(err.ptr, 0, sizeof(int), s));
        struct Job {
            ComplexMatrix *out;
            RegionKind kind;
            QuadratureOrders q;
            double tol;
            size_t R;
        };
        Job jobs[2] = {{L, RegionKind::Edge, orders_of(li), li->direction_tolerance, mesh->n_edges},
                       {P, RegionKind::Node, orders_of(pi), 0.0, mesh->n_nodes}};
        for (Job &j : jobs) {
            if (!j.out) continue;
            j.out->rows = j.out->cols = 0;
            j.out->data = nullptr;
            RegionSet rs(mesh, dual, j.kind, j.q, o->smooth_order, j.tol, s, false, o->near_smooth_order);
            DeviceBuffer<double> stat(j.R * j.R);
            DeviceBuffer<double2> ret(j.R * j.R);
            rs.assemble_static(stat.ptr, s, err.ptr, nullptr);
            rs.assemble_retarded(stat.ptr, k, ret.ptr, s, nullptr);
            Complex *h = host_calloc<Complex>(j.R * j.R);
            if (!h) return -1;
            ret.download(reinterpret_cast<double2 *>(h), j.R * j.R, s);
            s.sync();
            j.out->rows = j.out->cols = j.R;
            j.out->data = h;
        }
        int e = 0;
        err.download(&e, 1, s);
        s.sync();
        if (e) {
            if (L) complex_matrix_free(L);
            if (P) complex_matrix_free(P);
            return -1;
        }
        return 0;
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        return -1;
    }
}

// ---------------------------------------------------------------------------
// matrix.h / complex_matrix.h storage helpers

void matrix_init(Matrix *M) {
    if (M) std::memset(M, 0, sizeof(*M));
}
int matrix_alloc(Matrix *M, size_t rows, size_t cols) {
    if (!M || rows == 0 || cols == 0) return -1;
    M->data = host_calloc<double>(rows * cols);
    if (!M->data) return -1;
    M->rows = rows;
    M->cols = cols;
    return 0;
}
void matrix_free(Matrix *M) {
    if (!M) return;
    std::free(M->data);
    matrix_init(M);
}
void complex_matrix_init(ComplexMatrix *M) {
    if (M) std::memset(M, 0, sizeof(*M));
}
int complex_matrix_alloc(ComplexMatrix *M, size_t rows, size_t cols) {
    if (!M || rows == 0 || cols == 0) return -1;
    M->data = host_calloc<Complex>(rows * cols);
    if (!M->data) return -1;
    M->rows = rows;
    M->cols = cols;
    return 0;
}
void complex_matrix_free(ComplexMatrix *M) {
    if (!M) return;
    std::free(M->data);
    complex_matrix_init(M);
}

// complex_matrix.c:542: y = A x for any rows x cols
int complex_matrix_vector_multiply(const ComplexMatrix *A, const Complex *x, Complex *y, int threads) {
    if (!A || !A->data || !x || !y || threads < 1) return -1;
    if (A->cols == 0) {
        for (size_t i = 0; i < A->rows; ++i) y[i] = Complex{0.0, 0.0};
        return 0;
    }
    if (A->rows == 0) return 0;
    if (A->rows > (size_t)INT32_MAX || A->cols > (size_t)INT32_MAX) {
        PEEC_ERR("complex_matrix_vector_multiply: dimensions above 2^31 - 1.");
        return -1;
    }
    if (!require_device()) return -1;
    try {
        Stream s;
        const int rows = (int)A->rows, cols = (int)A->cols;
        DeviceBuffer<double2> dA, dx, dy(rows);
        DeviceBuffer<double> terms(3 * (size_t)rows);
        dA.upload(reinterpret_cast<const double2 *>(A->data), (size_t)rows * cols, s);
        dx.upload(reinterpret_cast<const double2 *>(x), cols, s);
        complex_residual(dA.ptr, rows, cols, (size_t)cols, dx.ptr, nullptr, dy.ptr, terms.ptr, s);
        dy.download(reinterpret_cast<double2 *>(y), rows, s);
        s.sync();
        return 0;
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        return -1;
    }
}

// ---------------------------------------------------------------------------
// incidence.h

void incidence_init(IncidenceMatrix *A) {
    if (A) std::memset(A, 0, sizeof(*A));
}
void incidence_free(IncidenceMatrix *A) {
    if (!A) return;
    std::free(A->data);
    incidence_init(A);
}

int incidence_build(const Mesh *mesh, IncidenceMatrix *A) {
    if (!mesh || !A) return -1;
    if (!mesh->edges || mesh->n_edges == 0 || mesh->n_nodes == 0) {
        PEEC_ERR("mesh topology is not available.");
        return -1;
    }
    incidence_init(A);
    A->rows = mesh->n_edges;
    A->cols = mesh->n_nodes;
    A->data = host_calloc<double>(A->rows * A->cols);
    if (!A->data) return -1;
    for (size_t e = 0; e < mesh->n_edges; ++e) {
        const int a = mesh->edges[2 * e], b = mesh->edges[2 * e + 1];
        if (a < 0 || b < 0 || (size_t)a >= mesh->n_nodes || (size_t)b >= mesh->n_nodes) {
            PEEC_ERR("node index out of range in edge %zu.", e);
            incidence_free(A);
            return -1;
        }
        A->data[e * A->cols + a] = -1.0;
        A->data[e * A->cols + b] = 1.0;
    }
    return 0;
}

int incidence_build_PAT(const Mesh *mesh, const PotentialMatrix *P, Matrix *PAT, int threads) {
    if (!mesh || !P || !PAT || threads < 1) return -1;
    if (!mesh->edges || !P->data) return -1;
    if (P->n != mesh->n_nodes) {
        PEEC_ERR("P size does not match mesh nodes in incidence_build_PAT().");
        return -1;
    }
    const int nv = (int)mesh->n_nodes, ne = (int)mesh->n_edges;
    for (int e = 0; e < ne; ++e)
        if (mesh->edges[2 * e] < 0 || mesh->edges[2 * e + 1] < 0 || mesh->edges[2 * e] >= nv ||
            mesh->edges[2 * e + 1] >= nv) {
            PEEC_ERR("invalid value in P A^T at edge %d.", e);
            return -1;
        }
    if (!require_device()) return -1;
    if (matrix_alloc(PAT, nv, ne) != 0) return -1;
    try {
        Stream s;
        DeviceBuffer<double> dP, dPAT((size_t)nv * ne);
        DeviceBuffer<int> dE;
        dP.upload(P->data, (size_t)nv * nv, s);
        dE.upload(mesh->edges, 2 * (size_t)ne, s);
        launch_pat(dP.ptr, dE.ptr, dPAT.ptr, nv, ne, s);
        dPAT.download(PAT->data, (size_t)nv * ne, s);
        s.sync();
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        matrix_free(PAT);
        return -1;
    }
    for (size_t k = 0; k < (size_t)nv * ne; ++k)
        if (!std::isfinite(PAT->data[k])) {
            matrix_free(PAT);
            return -1;
        }
    return 0;
}

// ---------------------------------------------------------------------------
// resistance.h (resistance.c:87-260)

void resistance_options_default(ResistanceOptions *o, int threads) {
    if (!o) return;
    o->model = PEEC_CONDUCTOR_PEC;
    o->sheet_resistance = 0.0;
    o->parallel_threads = threads;
}
void resistance_init(ResistanceMatrix *R) {
    if (R) std::memset(R, 0, sizeof(*R));
}
void resistance_free(ResistanceMatrix *R) {
    if (!R) return;
    std::free(R->diagonal);
    resistance_init(R);
}
double resistance_get(const ResistanceMatrix *R, size_t e) {
    if (!R || !R->diagonal || e >= R->n) return 0.0;
    return R->diagonal[e];
}

int resistance_compute_matrix(const Mesh *mesh, const DualMesh *dual, const ResistanceOptions *o,
                              ResistanceMatrix *R) {
    if (!mesh || !dual || !o || !R || o->parallel_threads < 1) return -1;
    if (mesh->n_edges == 0 || dual->n_edges != mesh->n_edges || !mesh->edge_lengths || !dual->edge_region_areas) {
        PEEC_ERR("invalid mesh or dual mesh in resistance calculation.");
        return -1;
    }
    if (o->model == PEEC_CONDUCTOR_SHEET_RESISTANCE) {
        if (!std::isfinite(o->sheet_resistance) || o->sheet_resistance < 0.0) {
            PEEC_ERR("invalid sheet resistance %.9e Ohm.", o->sheet_resistance);
            return -1;
        }
    } else if (o->model != PEEC_CONDUCTOR_PEC) {
        PEEC_ERR("unknown conductor model.");
        return -1;
    }
    resistance_free(R);
    R->n = mesh->n_edges;
    R->diagonal = host_calloc<double>(R->n);
    if (!R->diagonal) {
        resistance_init(R);
        return -1;
    }
    if (o->model == PEEC_CONDUCTOR_PEC) return 0;
    for (size_t i = 0; i < R->n; ++i) {
        const double l = mesh->edge_lengths[i], s = dual->edge_region_areas[i];
        if (l <= 0.0 || s <= 0.0 || !std::isfinite(l) || !std::isfinite(s)) {
            resistance_free(R);
            return -1;
        }
        R->diagonal[i] = o->sheet_resistance * l * l / s;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// lu.h

void lu_options_default(LUOptions *o, int threads) {
    if (!o) return;
    o->parallel_threads = threads;
    o->pivot_tolerance = 1.0e-14;
    o->parallel_threshold = 64;
}
void lu_init(LUFactorization *lu) {
    if (lu) std::memset(lu, 0, sizeof(*lu));
}
void lu_free(LUFactorization *lu) {
    if (!lu) return;
    forget_factor(lu->data);
    std::free(lu->data);
    std::free(lu->pivots);
    lu_init(lu);
}

int lu_factorize(const double *input, size_t n, const LUOptions *o, LUFactorization *lu) {
    if (!input || !o || !lu || n == 0 || o->parallel_threads < 1) return -1;
    if (n > (size_t)1 << 31) return -1;
    lu_init(lu);
    lu->n = n;
    lu->data = static_cast<double *>(std::malloc(n * n * sizeof(double)));
    lu->pivots = static_cast<size_t *>(std::malloc(n * sizeof(size_t)));
    if (!lu->data || !lu->pivots) {
        PEEC_ERR("cannot allocate LU factorization.");
        lu_free(lu);
        return -1;
    }
    double scale = 0.0;
    for (size_t k = 0; k < n * n; ++k) scale = std::fmax(scale, std::fabs(input[k]));  // lu.c:100-126
    for (size_t k = 0; k < n * n; ++k)
        if (!std::isfinite(input[k])) scale = NAN;
    lu->matrix_scale = scale;
    if (!std::isfinite(scale) || scale <= 0.0) {
        PEEC_ERR("LU input matrix is zero or contains invalid values.");
        lu_free(lu);
        return -2;
    }
    if (!require_device()) {
        lu_free(lu);
        return -1;
    }
    try {
        Stream s;
        LUContext ctx(s);
        auto f = std::make_unique<DeviceFactor>();
        f->n = n;
        f->zr.upload(input, n * n, s);
        f->ipiv.resize(n);
        LUStatus st = lu_factor<double>(f->zr.ptr, (int)n, n, f->ipiv.ptr, o->pivot_tolerance * scale, ctx);
        std::vector<int> ipiv(n);
        f->zr.download(lu->data, n * n, s);
        f->ipiv.download(ipiv.data(), n, s);
        s.sync();
        if (st.code != 0) {
            if (st.code == -2)
                PEEC_ERR("LU matrix is singular or nearly singular at column %lld.", st.column);
            lu->factorized = 0;
            return st.code == -2 ? -2 : -1;
        }
        lu->swap_count = 0;
        for (size_t k = 0; k < n; ++k) {
            lu->pivots[k] = (size_t)ipiv[k];
            if ((size_t)ipiv[k] != k) ++lu->swap_count;
        }
        lu->factorized = 1;
        std::lock_guard<std::mutex> lock(g_mu);
        g_factors[lu->data] = std::move(f);
        return 0;
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        return -1;
    }
}

static DeviceFactor *device_real_factor(const LUFactorization *lu, hipStream_t s, std::unique_ptr<DeviceFactor> &tmp) {
    {
        std::lock_guard<std::mutex> lock(g_mu);
        auto it = g_factors.find(lu->data);
        if (it != g_factors.end() && it->second->zr.ptr) return it->second.get();
    }
    tmp = std::make_unique<DeviceFactor>();
    tmp->n = lu->n;
    tmp->zr.upload(lu->data, lu->n * lu->n, s);
    std::vector<int> ipiv(lu->n);
    for (size_t k = 0; k < lu->n; ++k) ipiv[k] = (int)lu->pivots[k];
    tmp->ipiv.upload(ipiv, s);
    return tmp.get();
}

int lu_solve_multiple(const LUFactorization *lu, const double *B, double *X, size_t nrhs, int threads) {
    if (!lu || !B || !X || nrhs == 0 || threads < 1 || !lu->factorized || !lu->data || !lu->pivots) return -1;
    if (!require_device()) return -1;
    try {
        Stream s;
        LUContext ctx(s);
        std::unique_ptr<DeviceFactor> tmp;
        DeviceFactor *f = device_real_factor(lu, s, tmp);
        const size_t n = lu->n;
        DeviceBuffer<double> dx;
        dx.upload(B, n * nrhs, s);
        lu_solve<double, double>(f->zr.ptr, (int)n, n, f->ipiv.ptr, nullptr, nullptr, dx.ptr, (int)nrhs, ctx);
        dx.download(X, n * nrhs, s);
        s.sync();
        return 0;
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        return -1;
    }
}

int lu_solve(const LUFactorization *lu, const double *b, double *x) {
    if (!lu || !b || !x || !lu->data || !lu->pivots || !lu->factorized) return -1;
    return lu_solve_multiple(lu, b, x, 1, 1);
}

double lu_determinant(const LUFactorization *lu) {
    if (!lu || !lu->data || !lu->factorized) return NAN;
    double d = (lu->swap_count % 2 == 0) ? 1.0 : -1.0;  // lu.c:732-750
    for (size_t i = 0; i < lu->n; ++i) d *= lu->data[i * lu->n + i];
    return d;
}

void lu_print_info(const LUFactorization *lu) {
    if (!lu) return;
    std::printf("LU factorization: %zu x %zu, swaps %zu, scale %.9e, factorized %s\n", lu->n, lu->n,
                lu->swap_count, lu->matrix_scale, lu->factorized ? "yes" : "no");
}

// ---------------------------------------------------------------------------
// complex_lu.h

void complex_lu_options_default(ComplexLUOptions *o, int threads) {
    if (!o) return;
    o->parallel_threads = threads;
    o->pivot_tolerance = 1.0e-14;
    o->parallel_threshold = 64;
}
void complex_lu_init(ComplexLUFactorization *lu) {
    if (lu) std::memset(lu, 0, sizeof(*lu));
}
void complex_lu_free(ComplexLUFactorization *lu) {
    if (!lu) return;
    forget_factor(lu->data);
    std::free(lu->data);
    std::free(lu->pivots);
    std::free(lu->row_scale);
    std::free(lu->column_scale);
    complex_lu_init(lu);
}

int complex_lu_factorize(const Complex *input, size_t n, const ComplexLUOptions *o, ComplexLUFactorization *lu) {
    if (!input || !o || !lu || n == 0 || o->parallel_threads < 1 || !std::isfinite(o->pivot_tolerance) ||
        o->pivot_tolerance < 0.0)
        return -1;
    if (n > (size_t)1 << 31) return -1;
    complex_lu_init(lu);
    lu->n = n;
    lu->data = static_cast<Complex *>(std::malloc(n * n * sizeof(Complex)));
    lu->pivots = static_cast<size_t *>(std::malloc(n * sizeof(size_t)));
    lu->row_scale = static_cast<double *>(std::malloc(n * sizeof(double)));
    lu->column_scale = static_cast<double *>(std::malloc(n * sizeof(double)));
    if (!lu->data || !lu->pivots || !lu->row_scale || !lu->column_scale) {
        PEEC_ERR("cannot allocate complex LU factorization.");
        complex_lu_free(lu);
        return -1;
    }
    if (!require_device()) {
        complex_lu_free(lu);
        return -1;
    }
    try {
        Stream s;
        LUContext ctx(s);
        auto f = std::make_unique<DeviceFactor>();
        f->n = n;
        f->zc.upload(reinterpret_cast<const double2 *>(input), n * n, s);
        f->row.resize(n);
        f->col.resize(n);
        f->ipiv.resize(n);
        double big = 0.0;
        if (equilibrate_complex(f->zc.ptr, (int)n, n, f->row.ptr, f->col.ptr, &big, ctx) != 0) {
            complex_lu_free(lu);
            return -2;
        }
        lu->matrix_scale = big;
        if (!std::isfinite(big) || big <= 0.0) {
            PEEC_ERR("complex LU input matrix is invalid or zero.");
            complex_lu_free(lu);
            return -2;
        }
        LUStatus st = lu_factor<double2>(f->zc.ptr, (int)n, n, f->ipiv.ptr, o->pivot_tolerance * big, ctx);
        std::vector<int> ipiv(n);
        f->zc.download(reinterpret_cast<double2 *>(lu->data), n * n, s);
        f->row.download(lu->row_scale, n, s);
        f->col.download(lu->column_scale, n, s);
        f->ipiv.download(ipiv.data(), n, s);
        s.sync();
        if (st.code != 0) {
            if (st.code == -2)
                PEEC_ERR("complex LU matrix is singular or nearly singular at column %lld.", st.column);
            lu->factorized = 0;
            return st.code == -2 ? -2 : -1;
        }
        lu->swap_count = 0;
        for (size_t k = 0; k < n; ++k) {
            lu->pivots[k] = (size_t)ipiv[k];
            if ((size_t)ipiv[k] != k) ++lu->swap_count;
        }
        lu->factorized = 1;
        std::lock_guard<std::mutex> lock(g_mu);
        g_factors[lu->data] = std::move(f);
        return 0;
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        return -1;
    }
}

static DeviceFactor *device_complex_factor(const ComplexLUFactorization *lu, hipStream_t s,
                                           std::unique_ptr<DeviceFactor> &tmp) {
    {
        std::lock_guard<std::mutex> lock(g_mu);
        auto it = g_factors.find(lu->data);
        if (it != g_factors.end() && it->second->zc.ptr) return it->second.get();
    }
    tmp = std::make_unique<DeviceFactor>();
    const size_t n = lu->n;
    tmp->n = n;
    tmp->zc.upload(reinterpret_cast<const double2 *>(lu->data), n * n, s);
    std::vector<int> ipiv(n);
    for (size_t k = 0; k < n; ++k) ipiv[k] = (int)lu->pivots[k];
    tmp->ipiv.upload(ipiv, s);
    tmp->row.upload(lu->row_scale, n, s);
    tmp->col.upload(lu->column_scale, n, s);
    return tmp.get();
}

int complex_lu_solve_multiple(const ComplexLUFactorization *lu, const Complex *B, Complex *X, size_t nrhs,
                              int threads) {
    if (!lu || !B || !X || nrhs == 0 || threads < 1 || !lu->factorized || !lu->data || !lu->pivots) return -1;
    if (!require_device()) return -1;
    try {
        Stream s;
        LUContext ctx(s);
        std::unique_ptr<DeviceFactor> tmp;
        DeviceFactor *f = device_complex_factor(lu, s, tmp);
        const size_t n = lu->n;
        DeviceBuffer<double2> dx;
        dx.upload(reinterpret_cast<const double2 *>(B), n * nrhs, s);
        lu_solve<double2, double2>(f->zc.ptr, (int)n, n, f->ipiv.ptr, f->row.ptr, f->col.ptr, dx.ptr, (int)nrhs,
                                   ctx);
        dx.download(reinterpret_cast<double2 *>(X), n * nrhs, s);
        s.sync();
        return 0;
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        return -1;
    }
}

int complex_lu_solve(const ComplexLUFactorization *lu, const Complex *b, Complex *x) {
    if (!lu || !b || !x || !lu->data || !lu->pivots || !lu->factorized) return -1;
    return complex_lu_solve_multiple(lu, b, x, 1, 1);
}

void complex_lu_print_info(const ComplexLUFactorization *lu) {
    if (!lu) return;
    std::printf("Complex LU factorization: %zu x %zu, swaps %zu, scale %.9e, factorized %s\n", lu->n, lu->n,
                lu->swap_count, lu->matrix_scale, lu->factorized ? "yes" : "no");
}

// ---------------------------------------------------------------------------
// library self-test (no reference counterpart): MFMA operand layout probe

int peec_selftest_mfma_layout(double *out1024) {
    if (!out1024 || !require_device()) return -1;
    try {
        mfma_layout_probe(out1024);
        return 0;
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        return -1;
    }
}

// ---------------------------------------------------------------------------
// harmonic.h

void harmonic_system_init(HarmonicSystem *s) {
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
}
void harmonic_system_free(HarmonicSystem *s) {
    if (!s) return;
    complex_matrix_free(&s->matrix);
    complex_lu_free(&s->lu);
    std::memset(s, 0, sizeof(*s));
}
void harmonic_solution_init(HarmonicSolution *s) {
    if (s) std::memset(s, 0, sizeof(*s));
}
void harmonic_solution_free(HarmonicSolution *s) {
    if (!s) return;
    std::free(s->edge_current);
    std::free(s->node_voltage);
    harmonic_solution_init(s);
}

void harmonic_print_info(const HarmonicSystem *s) {
    if (!s) return;
    std::printf("Harmonic PEEC system: N = %zu (Ne %zu + Nv %zu), f = %.9e Hz, built %d, factorized %d\n",
                s->system_size, s->n_edges, s->n_nodes, s->frequency, s->matrix_built, s->factorized);
}

int harmonic_build_system(const InductanceMatrix *L, const ResistanceMatrix *R, const IncidenceMatrix *A,
                          const Matrix *PAT, double f, int threads, HarmonicSystem *sys) {
    if (!sys || threads < 1 || !std::isfinite(f) || f <= 0.0) return -1;
    if (!L || !R || !A || !PAT || !L->data || !R->diagonal || !A->data || !PAT->data) return -1;
    const size_t ne = L->n, nv = A->cols;
    if (ne == 0 || nv == 0 || A->rows != ne || R->n != ne || PAT->rows != nv || PAT->cols != ne) {
        PEEC_ERR("harmonic system: inconsistent L/R/A/PAT dimensions.");
        return -1;
    }
    harmonic_system_free(sys);
    harmonic_system_init(sys);
    const size_t n = ne + nv;
    const double omega = kTwoPi * f;
    sys->n_edges = ne;
    sys->n_nodes = nv;
    sys->system_size = n;
    sys->frequency = f;
    sys->omega = omega;
    if (complex_matrix_alloc(&sys->matrix, n, n) != 0) {
        PEEC_ERR("cannot allocate harmonic PEEC matrix.");
        harmonic_system_free(sys);
        return -1;
    }
    if (!require_device()) {
        harmonic_system_free(sys);
        return -1;
    }
    try {
        Stream s;
        DeviceBuffer<double> dL, dR, dA, dPAT;
        DeviceBuffer<double2> dZ(n * n);
        dL.upload(L->data, ne * ne, s);
        dR.upload(R->diagonal, ne, s);
        dA.upload(A->data, ne * nv, s);
        dPAT.upload(PAT->data, nv * ne, s);
        const int chunks = (int)((n + 255) / 256);
        hipLaunchKernelGGL(dense_mna_kernel, dim3((unsigned)(n * chunks)), dim3(256), 0, s, dZ.ptr, (int)ne,
                           (int)nv, dL.ptr, dR.ptr, dA.ptr, dPAT.ptr, omega, chunks);
        HIP_CHECK(hipGetLastError());
        dZ.download(reinterpret_cast<double2 *>(sys->matrix.data), n * n, s);
        s.sync();
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        harmonic_system_free(sys);
        return -1;
    }
    sys->matrix_built = 1;
    sys->factorized = 0;
    return 0;
}

int harmonic_factorize(HarmonicSystem *sys, int threads) {
    if (!sys || !sys->matrix_built || !sys->matrix.data || threads < 1) return -1;
    complex_lu_free(&sys->lu);
    complex_lu_init(&sys->lu);
    ComplexLUOptions o;
    complex_lu_options_default(&o, threads);
    if (complex_lu_factorize(sys->matrix.data, sys->system_size, &o, &sys->lu) != 0) {
        PEEC_ERR("harmonic PEEC LU factorization failed.");
        sys->factorized = 0;
        return -1;
    }
    sys->factorized = 1;
    return 0;
}

int harmonic_build_rhs(const PotentialMatrix *P, const HarmonicExcitation *ex, Complex *rhs, int threads) {
    if (!P || !ex || !rhs || !P->data || threads < 1) return -1;
    const size_t ne = ex->n_edges, nv = ex->n_nodes;
    if (P->n != nv) {
        PEEC_ERR("harmonic RHS: P size does not match Nv.");
        return -1;
    }
    if (!require_device()) return -1;
    try {
        Stream s;
        DeviceBuffer<double> dP;
        DeviceBuffer<double2> dU, dI, dB(ne + nv);
        dP.upload(P->data, nv * nv, s);
        if (ex->edge_voltage) dU.upload(reinterpret_cast<const double2 *>(ex->edge_voltage), ne, s);
        if (ex->node_current) dI.upload(reinterpret_cast<const double2 *>(ex->node_current), nv, s);
        launch_general_rhs(dB.ptr, (int)ne, (int)nv, dP.ptr, ex->edge_voltage ? dU.ptr : nullptr,
                           ex->node_current ? dI.ptr : nullptr, s);
        dB.download(reinterpret_cast<double2 *>(rhs), ne + nv, s);
        s.sync();
        return 0;
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        return -1;
    }
}

int harmonic_solve(const HarmonicSystem *sys, const PotentialMatrix *P, const HarmonicExcitation *ex,
                   HarmonicSolution *sol, int threads) {
    if (!sys || !P || !ex || !sol || !sys->factorized || threads < 1) return -1;
    if (ex->n_edges != sys->n_edges || ex->n_nodes != sys->n_nodes) {
        PEEC_ERR("harmonic excitation dimensions do not match system.");
        return -1;
    }
    const size_t ne = sys->n_edges, nv = sys->n_nodes, n = sys->system_size;
    std::vector<Complex> rhs(n), x(n);
    if (harmonic_build_rhs(P, ex, rhs.data(), threads) != 0) return -1;
    if (complex_lu_solve(&sys->lu, rhs.data(), x.data()) != 0) {
        PEEC_ERR("harmonic PEEC solve failed.");
        return -1;
    }
    harmonic_solution_free(sol);
    sol->n_edges = ne;
    sol->n_nodes = nv;
    sol->edge_current = static_cast<Complex *>(std::malloc(ne * sizeof(Complex)));
    sol->node_voltage = static_cast<Complex *>(std::malloc(nv * sizeof(Complex)));
    if (!sol->edge_current || !sol->node_voltage) {
        harmonic_solution_free(sol);
        return -1;
    }
    std::memcpy(sol->edge_current, x.data(), ne * sizeof(Complex));
    std::memcpy(sol->node_voltage, x.data() + ne, nv * sizeof(Complex));
    return 0;
}

int harmonic_compute_residual(const HarmonicSystem *sys, const PotentialMatrix *P, const HarmonicExcitation *ex,
                              const HarmonicSolution *sol, int threads, HarmonicResidual *res) {
    if (!sys || !P || !ex || !sol || !res || threads < 1) return -1;
    if (!sys->matrix_built || !sys->matrix.data) return -1;
    if (ex->n_edges != sys->n_edges || ex->n_nodes != sys->n_nodes) return -1;
    if (sol->n_edges != sys->n_edges || sol->n_nodes != sys->n_nodes || !sol->edge_current || !sol->node_voltage)
        return -1;
    const size_t ne = sys->n_edges, nv = sys->n_nodes, n = sys->system_size;
    std::vector<Complex> x(n), rhs(n);
    std::memcpy(x.data(), sol->edge_current, ne * sizeof(Complex));
    std::memcpy(x.data() + ne, sol->node_voltage, nv * sizeof(Complex));
    if (harmonic_build_rhs(P, ex, rhs.data(), threads) != 0) return -1;
    if (!require_device()) return -1;
    try {
        Stream s;
        DeviceBuffer<double2> dZ, dx, db;
        DeviceBuffer<double> terms(3 * n);
        dZ.upload(reinterpret_cast<const double2 *>(sys->matrix.data), n * n, s);
        dx.upload(reinterpret_cast<const double2 *>(x.data()), n, s);
        db.upload(reinterpret_cast<const double2 *>(rhs.data()), n, s);
        complex_residual(dZ.ptr, (int)n, (int)n, n, dx.ptr, db.ptr, nullptr, terms.ptr, s);
        std::vector<double> h(3 * n);
        terms.download(h.data(), 3 * n, s);
        s.sync();
        double out[7];
        residual_summary(h, n, ne, out);
        res->absolute_residual = out[0];
        res->rhs_relative_residual = out[1];
        res->backward_error = out[2];
        res->edge_absolute_residual = out[3];
        res->node_absolute_residual = out[4];
        res->edge_backward_error = out[5];
        res->node_backward_error = out[6];
        return 0;
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        return -1;
    }
}

}  // extern "C"


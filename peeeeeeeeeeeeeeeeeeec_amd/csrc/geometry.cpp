// Host-side geometry producers of the drop-in ABI: Gmsh loading, edge
// topology, dual regions and the scalar quadrature helpers. These are O(N)
// bookkeeping passes that fix the numbering and orientation of every branch;
// all O(N^2) and O(N^3) work runs in the HIP kernels.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "../../include/dual_mesh.h"
#include "../../include/mesh.h"
#include "../../include/quadrature.h"
#include "common.h"

namespace peec {

bool require_device() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
        PEEC_ERR("no HIP device is visible; this library has no CPU path.");
        return false;
    }
    return true;
}

// ---------------------------------------------------------------------------
// Gauss-Legendre rules (the reference tabulates the same orders in
// quadrature.c:94-313). Roots of P_n by Newton's method in binary128, then
// one rounding to double.

namespace {
using quad = __float128;

void legendre(int n, quad x, quad &p, quad &dp) {
    quad a = 1, b = x;
    for (int k = 2; k <= n; ++k) {
        quad c = ((2 * k - 1) * x * b - (k - 1) * a) / k;
        a = b;
        b = c;
    }
    p = b;
    dp = n * (x * b - a) / (x * x - 1);
}

GLTable build_table() {
    GLTable t{};
    for (int n = 2; n <= 16; ++n) {
        if (!gl_order_supported(n)) continue;
        for (int i = 0; i < n; ++i) {
            quad x = -std::cos(M_PI * (i + 0.75) / (n + 0.5));
            for (int it = 0; it < 100; ++it) {
                quad p, dp;
                legendre(n, x, p, dp);
                quad dx = p / dp;
                x -= dx;
                if (dx == 0) break;
            }
            quad p, dp;
            legendre(n, x, p, dp);
            t.x[n][i] = (double)x;
            t.w[n][i] = (double)(2 / ((1 - x * x) * dp * dp));
        }
        for (int i = 0; i < n / 2; ++i) {
            t.x[n][n - 1 - i] = -t.x[n][i];
            t.w[n][n - 1 - i] = t.w[n][i];
        }
        if (n % 2) t.x[n][n / 2] = 0.0;
    }
    return t;
}
}  // namespace

const GLTable &gl_table() {
    static const GLTable table = build_table();
    return table;
}

bool gauss_legendre(int order, double *x, double *w) {
    if (!x || !w || !gl_order_supported(order)) return false;
    const GLTable &t = gl_table();
    std::memcpy(x, t.x[order], sizeof(double) * order);
    std::memcpy(w, t.w[order], sizeof(double) * order);
    return true;
}

}  // namespace peec

using namespace peec;

namespace {

inline double norm3(const double *a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
inline double dist3(const double *a, const double *b) {
    double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    return norm3(d);
}
inline void cross3(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
inline bool coincide(const double *a, const double *b) { return dist3(a, b) <= 1.0e-12; }

template <typename T>
T *alloc_array(size_t n) {
    return static_cast<T *>(std::malloc(sizeof(T) * (n ? n : 1)));
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// Mesh (mesh.h:141-175)

void mesh_init(Mesh *mesh) {
    if (mesh) std::memset(mesh, 0, sizeof(*mesh));
}

void mesh_free(Mesh *m) {
    if (!m) return;
    void *arrays[] = {m->xyz, m->quads, m->edges, m->cell_edges, m->cell_edge_signs,
                      m->edge_vectors, m->edge_lengths, m->edge_directions, m->edge_centers,
                      m->cell_centers, m->cell_areas, m->cell_normals};
    for (void *p : arrays) std::free(p);
    mesh_init(m);
}

void mesh_print_info(const Mesh *m) {
    if (!m) return;
    std::printf("Mesh: %zu nodes, %zu quads, %zu edges\n", m->n_nodes, m->n_quads, m->n_edges);
}

// Unique edges keyed by the unordered node pair, numbered and oriented by
// their first appearance in cell order (topology.c:160-304); per-cell edge
// ids and signs; edge and cell geometry (topology.c:340-529).
int mesh_build_topology(Mesh *m) {
    if (!m || !m->xyz || !m->quads) return -1;
    const size_t nq = m->n_quads;
    std::free(m->edges);
    std::free(m->cell_edges);
    std::free(m->cell_edge_signs);
    m->cell_edges = alloc_array<int>(4 * nq);
    m->cell_edge_signs = alloc_array<int>(4 * nq);
    std::vector<int> edges;
    edges.reserve(4 * nq);
    std::unordered_map<uint64_t, int> seen;
    seen.reserve(4 * nq);
    for (size_t c = 0; c < nq; ++c) {
        for (int l = 0; l < 4; ++l) {
            const int from = m->quads[4 * c + l], to = m->quads[4 * c + (l + 1) % 4];
            if (from < 0 || to < 0 || (size_t)from >= m->n_nodes || (size_t)to >= m->n_nodes) {
                PEEC_ERR("quad %zu references a node out of range.", c);
                return -1;
            }
            const uint64_t lo = (uint32_t)std::min(from, to), hi = (uint32_t)std::max(from, to);
            auto ins = seen.emplace((lo << 32) | hi, (int)(edges.size() / 2));
            const int id = ins.first->second;
            int sign = 1;
            if (ins.second) {
                edges.push_back(from);
                edges.push_back(to);
            } else if (edges[2 * id] == to && edges[2 * id + 1] == from) {
                sign = -1;
            } else if (!(edges[2 * id] == from && edges[2 * id + 1] == to)) {
                PEEC_ERR("inconsistent edge orientation.");
                return -1;
            }
            m->cell_edges[4 * c + l] = id;
            m->cell_edge_signs[4 * c + l] = sign;
        }
    }
    const size_t ne = edges.size() / 2;
    m->n_edges = ne;
    m->edges = alloc_array<int>(2 * ne);
    std::memcpy(m->edges, edges.data(), sizeof(int) * 2 * ne);
    double **geo[] = {&m->edge_vectors, &m->edge_lengths, &m->edge_directions, &m->edge_centers,
                      &m->cell_centers, &m->cell_areas, &m->cell_normals};
    for (double **g : geo) std::free(*g);
    m->edge_vectors = alloc_array<double>(3 * ne);
    m->edge_lengths = alloc_array<double>(ne);
    m->edge_directions = alloc_array<double>(3 * ne);
    m->edge_centers = alloc_array<double>(3 * ne);
    m->cell_centers = alloc_array<double>(3 * nq);
    m->cell_areas = alloc_array<double>(nq);
    m->cell_normals = alloc_array<double>(3 * nq);
    for (size_t e = 0; e < ne; ++e) {
        const double *a = &m->xyz[3 * m->edges[2 * e]], *b = &m->xyz[3 * m->edges[2 * e + 1]];
        double d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double l = norm3(d);
        if (!(l > 0.0)) {
            PEEC_ERR("zero-length edge %zu.", e);
            return -1;
        }
        m->edge_lengths[e] = l;
        for (int k = 0; k < 3; ++k) {
            m->edge_vectors[3 * e + k] = d[k];
            m->edge_directions[3 * e + k] = d[k] / l;
            m->edge_centers[3 * e + k] = 0.5 * (a[k] + b[k]);
        }
    }
    for (size_t c = 0; c < nq; ++c) {
        const double *p[4];
        for (int k = 0; k < 4; ++k) p[k] = &m->xyz[3 * m->quads[4 * c + k]];
        double u[3], v[3], w[3], n1[3], n2[3];
        for (int k = 0; k < 3; ++k) {
            m->cell_centers[3 * c + k] = 0.25 * (p[0][k] + p[1][k] + p[2][k] + p[3][k]);
            u[k] = p[1][k] - p[0][k];
            v[k] = p[2][k] - p[0][k];
            w[k] = p[3][k] - p[0][k];
        }
        cross3(u, v, n1);
        cross3(v, w, n2);
        m->cell_areas[c] = 0.5 * (norm3(n1) + norm3(n2));
        double n[3] = {n1[0] + n2[0], n1[1] + n2[1], n1[2] + n2[2]};
        const double nl = norm3(n);
        for (int k = 0; k < 3; ++k) m->cell_normals[3 * c + k] = nl > 0.0 ? n[k] / nl : n[k];
    }
    return 0;
}

// Gmsh 4.x ASCII reader (mesh_gmsh.c:90-580): nodes indexed by order of
// appearance in $Nodes, quadrangles (type 3) only, other types skipped.
int mesh_load_gmsh(const char *filename, Mesh *m) {
    if (!filename || !m) return -1;
    FILE *f = std::fopen(filename, "r");
    if (!f) {
        PEEC_ERR("cannot open mesh file %s.", filename);
        return -1;
    }
    mesh_init(m);
    char line[4096];
    std::vector<int> tag_to_index;
    std::vector<double> xyz;
    std::vector<int> quads;
    bool ok = true, have_nodes = false;
    auto next = [&]() { return std::fgets(line, sizeof line, f) != nullptr; };
    while (ok && next()) {
        if (!std::strncmp(line, "$Nodes", 6)) {
            size_t blocks, count, lo, hi;
            ok = next() && std::sscanf(line, "%zu %zu %zu %zu", &blocks, &count, &lo, &hi) == 4;
            if (!ok) break;
            tag_to_index.assign(hi + 1, -1);
            xyz.reserve(3 * count);
            for (size_t b = 0; ok && b < blocks; ++b) {
                int dim, ent, par;
                size_t n;
                ok = next() && std::sscanf(line, "%d %d %d %zu", &dim, &ent, &par, &n) == 4 && !par;
                std::vector<size_t> tags(n);
                for (size_t i = 0; ok && i < n; ++i) ok = next() && std::sscanf(line, "%zu", &tags[i]) == 1;
                for (size_t i = 0; ok && i < n; ++i) {
                    double x, y, z;
                    ok = next() && std::sscanf(line, "%lf %lf %lf", &x, &y, &z) == 3 &&
                         tags[i] < tag_to_index.size();
                    if (!ok) break;
                    tag_to_index[tags[i]] = (int)(xyz.size() / 3);
                    xyz.insert(xyz.end(), {x, y, z});
                }
            }
            ok = ok && xyz.size() == 3 * count;
            have_nodes = ok;
        } else if (!std::strncmp(line, "$Elements", 9)) {
            size_t blocks, count, lo, hi;
            ok = have_nodes && next() &&
                 std::sscanf(line, "%zu %zu %zu %zu", &blocks, &count, &lo, &hi) == 4;
            for (size_t b = 0; ok && b < blocks; ++b) {
                int dim, ent, type;
                size_t n;
                ok = next() && std::sscanf(line, "%d %d %d %zu", &dim, &ent, &type, &n) == 4;
                for (size_t i = 0; ok && i < n; ++i) {
                    ok = next();
                    if (!ok || type != 3) continue;
                    size_t tag, v[4];
                    ok = std::sscanf(line, "%zu %zu %zu %zu %zu", &tag, &v[0], &v[1], &v[2], &v[3]) == 5;
                    for (int k = 0; ok && k < 4; ++k) {
                        ok = v[k] < tag_to_index.size() && tag_to_index[v[k]] >= 0;
                        if (ok) quads.push_back(tag_to_index[v[k]]);
                    }
                }
            }
        }
    }
    std::fclose(f);
    if (!ok || xyz.empty() || quads.empty()) {
        PEEC_ERR("cannot parse Gmsh quad mesh %s.", filename);
        return -1;
    }
    m->n_nodes = xyz.size() / 3;
    m->n_quads = quads.size() / 4;
    m->xyz = alloc_array<double>(xyz.size());
    m->quads = alloc_array<int>(quads.size());
    std::memcpy(m->xyz, xyz.data(), sizeof(double) * xyz.size());
    std::memcpy(m->quads, quads.data(), sizeof(int) * quads.size());
    return 0;
}

// ---------------------------------------------------------------------------
// Dual mesh (dual_mesh.c:356-692)

void dual_mesh_init(DualMesh *d) {
    if (d) std::memset(d, 0, sizeof(*d));
}

void dual_mesh_free(DualMesh *d) {
    if (!d) return;
    for (size_t i = 0; d->node_patches && i < d->n_nodes; ++i) std::free(d->node_patches[i].patches);
    for (size_t i = 0; d->edge_patches && i < d->n_edges; ++i) std::free(d->edge_patches[i].patches);
    std::free(d->node_patches);
    std::free(d->edge_patches);
    std::free(d->node_region_areas);
    std::free(d->edge_region_areas);
    std::free(d->node_region_centers);
    std::free(d->edge_region_centers);
    dual_mesh_init(d);
}

void dual_mesh_print_info(const DualMesh *d) {
    if (!d) return;
    std::printf("DualMesh: %zu node regions, %zu edge regions\n", d->n_nodes, d->n_edges);
}

static bool push_piece(PatchList &list, const double *a, const double *b, const double *c,
                       const double *e, int owner, int cell) {
    if (list.count == list.capacity) {
        size_t cap = list.capacity ? 2 * list.capacity : 4;
        auto *p = static_cast<QuadPatch *>(std::realloc(list.patches, cap * sizeof(QuadPatch)));
        if (!p) return false;
        list.patches = p;
        list.capacity = cap;
    }
    QuadPatch &q = list.patches[list.count++];
    const double *v[4] = {a, b, c, e};
    for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 3; ++d) q.points[3 * k + d] = v[k][d];
    q.owner_id = owner;
    q.cell_id = cell;
    return true;
}

// region area = sum over pieces of 0.5|t0| + 0.5|t1| (dual_mesh.c:79-108),
// centre = area-weighted piece-vertex mean (dual_mesh.c:596-681)
static void finish_region(const PatchList &list, double *area, double *centre) {
    double s = 0.0, acc[3] = {0.0, 0.0, 0.0};
    for (size_t k = 0; k < list.count; ++k) {
        const double *p = list.patches[k].points;
        double u[3], v[3], w[3], c1[3], c2[3];
        for (int d = 0; d < 3; ++d) {
            u[d] = p[3 + d] - p[d];
            v[d] = p[6 + d] - p[d];
            w[d] = p[9 + d] - p[d];
        }
        cross3(u, v, c1);
        cross3(v, w, c2);
        const double a = 0.5 * norm3(c1) + 0.5 * norm3(c2);
        s += a;
        for (int d = 0; d < 3; ++d) acc[d] += a * (0.25 * (p[d] + p[3 + d] + p[6 + d] + p[9 + d]));
    }
    *area = s;
    if (s > 0.0)
        for (int d = 0; d < 3; ++d) centre[d] = acc[d] / s;
}

int dual_mesh_build(const Mesh *m, DualMesh *d) {
    if (!m || !d) return -1;
    if (!m->xyz || !m->quads || !m->cell_edges) {
        PEEC_ERR("topology must be built before dual mesh.");
        return -1;
    }
    dual_mesh_init(d);
    d->n_nodes = m->n_nodes;
    d->n_edges = m->n_edges;
    d->node_patches = static_cast<PatchList *>(std::calloc(m->n_nodes, sizeof(PatchList)));
    d->edge_patches = static_cast<PatchList *>(std::calloc(m->n_edges, sizeof(PatchList)));
    d->node_region_areas = static_cast<double *>(std::calloc(m->n_nodes, sizeof(double)));
    d->edge_region_areas = static_cast<double *>(std::calloc(m->n_edges, sizeof(double)));
    d->node_region_centers = static_cast<double *>(std::calloc(3 * m->n_nodes, sizeof(double)));
    d->edge_region_centers = static_cast<double *>(std::calloc(3 * m->n_edges, sizeof(double)));
    if (!d->node_patches || !d->edge_patches || !d->node_region_areas || !d->edge_region_areas ||
        !d->node_region_centers || !d->edge_region_centers) {
        dual_mesh_free(d);
        return -1;
    }
    for (size_t c = 0; c < m->n_quads; ++c) {
        const int *q = &m->quads[4 * c];
        double p[4][3], mid[4][3], ctr[3];
        for (int k = 0; k < 4; ++k)
            for (int a = 0; a < 3; ++a) p[k][a] = m->xyz[3 * q[k] + a];
        for (int k = 0; k < 4; ++k)
            for (int a = 0; a < 3; ++a) mid[k][a] = 0.5 * (p[k][a] + p[(k + 1) % 4][a]);
        for (int a = 0; a < 3; ++a) ctr[a] = 0.25 * (p[0][a] + p[1][a] + p[2][a] + p[3][a]);
        bool ok = true;
        // quarter-quads [p_k, mid(k,k+1), centre, mid(k-1,k)] (dual_mesh.c:446-504)
        for (int k = 0; k < 4 && ok; ++k)
            ok = push_piece(d->node_patches[q[k]], p[k], mid[k], ctr, mid[(k + 3) % 4], q[k], (int)c);
        // half-quads [p_k, p_k+1, mid(k+1,k+2), mid(k-1,k)] (dual_mesh.c:506-588)
        for (int k = 0; k < 4 && ok; ++k) {
            const int e = m->cell_edges[4 * c + k];
            ok = push_piece(d->edge_patches[e], p[k], p[(k + 1) % 4], mid[(k + 1) % 4],
                            mid[(k + 3) % 4], e, (int)c);
        }
        if (!ok) {
            PEEC_ERR("cannot allocate dual mesh patch.");
            dual_mesh_free(d);
            return -1;
        }
    }
    for (size_t i = 0; i < d->n_nodes; ++i)
        finish_region(d->node_patches[i], &d->node_region_areas[i], &d->node_region_centers[3 * i]);
    for (size_t i = 0; i < d->n_edges; ++i)
        finish_region(d->edge_patches[i], &d->edge_region_areas[i], &d->edge_region_centers[3 * i]);
    return 0;
}

// ---------------------------------------------------------------------------
// scalar quadrature helpers (quadrature.c:94-607)

int gauss_legendre_rule(int order, double *x, double *w) {
    if (!x || !w) return -1;
    if (!peec::gauss_legendre(order, x, w)) {
        PEEC_ERR("unsupported Gauss-Legendre order: %d.", order);
        return -1;
    }
    return 0;
}

double triangle_area(const double p0[3], const double p1[3], const double p2[3]) {
    double a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    double b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    double c[3];
    cross3(a, b, c);
    return 0.5 * norm3(c);
}

double quad_patch_area(const QuadPatch *q) {
    if (!q) return 0.0;
    const double *p = q->points;
    return triangle_area(p, p + 3, p + 6) + triangle_area(p, p + 6, p + 9);
}

double quad_patch_size(const QuadPatch *q) {
    if (!q) return 0.0;
    double h = 0.0;
    for (int i = 0; i < 4; ++i) {
        const double l = dist3(&q->points[3 * i], &q->points[3 * ((i + 1) % 4)]);
        if (l > h) h = l;
    }
    return h;
}

double quad_patch_min_vertex_distance(const QuadPatch *a, const QuadPatch *b) {
    if (!a || !b) return 0.0;
    double best = HUGE_VAL;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) best = std::fmin(best, dist3(&a->points[3 * i], &b->points[3 * j]));
    return best;
}

QuadPairType quad_patch_classify_pair(const QuadPatch *a, const QuadPatch *b, double nf) {
    if (!a || !b) return QUAD_PAIR_FAR;
    if (a == b) return QUAD_PAIR_SELF;
    bool used[4] = {false, false, false, false}, all = true;
    for (int i = 0; i < 4 && all; ++i) {
        bool hit = false;
        for (int j = 0; j < 4 && !hit; ++j)
            if (!used[j] && coincide(&a->points[3 * i], &b->points[3 * j])) used[j] = hit = true;
        all = hit;
    }
    if (all) return QUAD_PAIR_SELF;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            if (coincide(&a->points[3 * i], &b->points[3 * j])) return QUAD_PAIR_TOUCHING;
    double g2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        double alo = a->points[k], ahi = alo, blo = b->points[k], bhi = blo;
        for (int i = 1; i < 4; ++i) {
            alo = std::fmin(alo, a->points[3 * i + k]);
            ahi = std::fmax(ahi, a->points[3 * i + k]);
            blo = std::fmin(blo, b->points[3 * i + k]);
            bhi = std::fmax(bhi, b->points[3 * i + k]);
        }
        const double gap = std::fmax(0.0, std::fmax(alo - bhi, blo - ahi));
        g2 += gap * gap;
    }
    const double ha = quad_patch_size(a), hb = quad_patch_size(b);
    const double h = ha > hb ? ha : hb;
    return (h > 0.0 && std::sqrt(g2) < nf * h) ? QUAD_PAIR_NEAR : QUAD_PAIR_FAR;
}

}  // extern "C"

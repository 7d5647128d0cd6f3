/* CPU oracle for the PEEC hot path — TEST INFRASTRUCTURE ONLY (see
 * peec_oracle.h). Restates the reference algorithm with the same rules,
 * orders, tolerances and floating-point evaluation order, so that built with
 * the reference's flags (ISO C17: no FP contraction) it reproduces the
 * reference bit for bit. Citations are to the reference sources. */
#include "peec_oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846264338327950288
#endif

/* ------------------------------------------------------------------ */
/* Gauss-Legendre rules (quadrature.c:94-313 tabulates the same orders). */
/* Nodes are computed here from the Legendre recurrence in binary128 and  */
/* rounded once to double, so no table is transcribed.                    */
/* ------------------------------------------------------------------ */

typedef __float128 q128;

static void legendre_q(int n, q128 x, q128 *p, q128 *dp)
{
    q128 p0 = 1, p1 = x;
    for (int k = 2; k <= n; ++k) {
        q128 pk = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = pk;
    }
    *p = p1;
    *dp = n * (x * p1 - p0) / (x * x - 1);
}

static int gl_supported(int order)
{
    return order == 2 || order == 3 || order == 4 || order == 5 || order == 6
        || order == 8 || order == 10 || order == 12 || order == 16;
}

static double gl_x[17][16], gl_w[17][16];
static int gl_ready;

static void gl_init(void)
{
    for (int n = 2; n <= 16; ++n) {
        if (!gl_supported(n)) continue;
        for (int i = 0; i < n; ++i) {
            /* i-th root in ascending order; Chebyshev-like start. */
            q128 x = -cos((double)M_PI * (i + 0.75) / (n + 0.5));
            for (int it = 0; it < 100; ++it) {
                q128 p, dp;
                legendre_q(n, x, &p, &dp);
                q128 dx = p / dp;
                x -= dx;
                if (dx == 0) break;
            }
            q128 p, dp;
            legendre_q(n, x, &p, &dp);
            q128 w = 2 / ((1 - x * x) * dp * dp);
            gl_x[n][i] = (double)x;
            gl_w[n][i] = (double)w;
        }
        /* enforce exact symmetry and the zero node like the textbook table */
        for (int i = 0; i < n / 2; ++i) {
            gl_x[n][n - 1 - i] = -gl_x[n][i];
            gl_w[n][n - 1 - i] = gl_w[n][i];
        }
        if (n % 2) gl_x[n][n / 2] = 0.0;
    }
    gl_ready = 1;
}

__attribute__((constructor)) static void oracle_ctor(void) { gl_init(); }

int oracle_gauss_legendre(int order, double *x, double *w)
{
    if (!x || !w || !gl_supported(order)) return -1;
    if (!gl_ready) gl_init();
    memcpy(x, gl_x[order], sizeof(double) * order);
    memcpy(w, gl_w[order], sizeof(double) * order);
    return 0;
}

/* ------------------------------------------------------------------ */
/* small vector helpers (quadrature.c:17-83)                           */
/* ------------------------------------------------------------------ */

static double len3(const double a[3])
{
    return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
}
static double dist3(const double a[3], const double b[3])
{
    double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    return len3(d);
}
static void crossp(const double a[3], const double b[3], double c[3])
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static double dotp(const double a[3], const double b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static int same_point(const double a[3], const double b[3])
{
    return dist3(a, b) <= 1.0e-12; /* quadrature.c:78-83 */
}

/* ------------------------------------------------------------------ */
/* topology (topology.c:111-529): edges numbered and oriented by first */
/* appearance in cell order; lookup key is the unordered node pair.    */
/* ------------------------------------------------------------------ */

typedef struct { long long key; int id; } EdgeSlot;

static size_t pow2_at_least(size_t v)
{
    size_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

static uint64_t mix64(uint64_t v)
{
    v ^= v >> 31; v *= 0x9e3779b97f4a7c15ULL; v ^= v >> 29;
    return v;
}

int oracle_build_topology(Mesh *m)
{
    if (!m || !m->xyz || !m->quads) return -1;
    size_t nq = m->n_quads;
    m->edges = malloc(sizeof(int) * 8 * nq);
    m->cell_edges = malloc(sizeof(int) * 4 * nq);
    m->cell_edge_signs = malloc(sizeof(int) * 4 * nq);
    size_t cap = pow2_at_least(8 * nq + 8);
    EdgeSlot *table = malloc(sizeof(EdgeSlot) * cap);
    if (!m->edges || !m->cell_edges || !m->cell_edge_signs || !table) {
        free(table);
        return -1;
    }
    for (size_t s = 0; s < cap; ++s) table[s].key = -1;
    size_t ne = 0;
    for (size_t c = 0; c < nq; ++c) {
        const int *q = &m->quads[4 * c];
        for (int l = 0; l < 4; ++l) {
            int from = q[l], to = q[(l + 1) % 4];
            long long lo = from < to ? from : to, hi = from < to ? to : from;
            long long key = (lo << 32) | hi;
            size_t s = mix64((uint64_t)key) & (cap - 1);
            while (table[s].key != -1 && table[s].key != key) s = (s + 1) & (cap - 1);
            if (table[s].key == -1) {
                table[s].key = key;
                table[s].id = (int)ne;
                m->edges[2 * ne] = from;
                m->edges[2 * ne + 1] = to;
                m->cell_edges[4 * c + l] = (int)ne;
                m->cell_edge_signs[4 * c + l] = 1;
                ++ne;
            } else {
                int e = table[s].id;
                int sign = (m->edges[2 * e] == from && m->edges[2 * e + 1] == to) ? 1
                         : (m->edges[2 * e] == to && m->edges[2 * e + 1] == from) ? -1 : 0;
                if (!sign) { free(table); return -1; }
                m->cell_edges[4 * c + l] = e;
                m->cell_edge_signs[4 * c + l] = sign;
            }
        }
    }
    free(table);
    m->n_edges = ne;
    m->edge_vectors = malloc(sizeof(double) * 3 * ne);
    m->edge_lengths = malloc(sizeof(double) * ne);
    m->edge_directions = malloc(sizeof(double) * 3 * ne);
    m->edge_centers = malloc(sizeof(double) * 3 * ne);
    m->cell_centers = malloc(sizeof(double) * 3 * nq);
    m->cell_areas = malloc(sizeof(double) * nq);
    m->cell_normals = malloc(sizeof(double) * 3 * nq);
    if (!m->edge_vectors || !m->edge_lengths || !m->edge_directions ||
        !m->edge_centers || !m->cell_centers || !m->cell_areas || !m->cell_normals)
        return -1;
    for (size_t e = 0; e < ne; ++e) {
        const double *A = &m->xyz[3 * m->edges[2 * e]];
        const double *B = &m->xyz[3 * m->edges[2 * e + 1]];
        double d[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
        double l = len3(d);
        if (l <= 0.0) return -1;
        for (int k = 0; k < 3; ++k) {
            m->edge_vectors[3 * e + k] = d[k];
            m->edge_directions[3 * e + k] = d[k] / l;
            m->edge_centers[3 * e + k] = 0.5 * (A[k] + B[k]);
        }
        m->edge_lengths[e] = l;
    }
    for (size_t c = 0; c < nq; ++c) {
        const double *p[4];
        for (int k = 0; k < 4; ++k) p[k] = &m->xyz[3 * m->quads[4 * c + k]];
        double d1[3], d2[3], d3[3], n1[3], n2[3];
        for (int k = 0; k < 3; ++k) {
            m->cell_centers[3 * c + k] = 0.25 * (p[0][k] + p[1][k] + p[2][k] + p[3][k]);
            d1[k] = p[1][k] - p[0][k];
            d2[k] = p[2][k] - p[0][k];
            d3[k] = p[3][k] - p[0][k];
        }
        crossp(d1, d2, n1);
        crossp(d2, d3, n2);
        m->cell_areas[c] = 0.5 * (len3(n1) + len3(n2));
        double n[3] = {n1[0] + n2[0], n1[1] + n2[1], n1[2] + n2[2]};
        double nl = len3(n);
        for (int k = 0; k < 3; ++k) m->cell_normals[3 * c + k] = nl > 0.0 ? n[k] / nl : n[k];
    }
    return 0;
}

void oracle_free_mesh(Mesh *m)
{
    if (!m) return;
    free(m->xyz); free(m->quads); free(m->edges); free(m->cell_edges);
    free(m->cell_edge_signs); free(m->edge_vectors); free(m->edge_lengths);
    free(m->edge_directions); free(m->edge_centers); free(m->cell_centers);
    free(m->cell_areas); free(m->cell_normals);
    memset(m, 0, sizeof(*m));
}

/* ------------------------------------------------------------------ */
/* Gmsh 4.x ASCII loader (mesh_gmsh.c:90-580): node index = order of   */
/* appearance in $Nodes, quads = element type 3.                        */
/* ------------------------------------------------------------------ */

int oracle_load_gmsh(const char *path, Mesh *m)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    memset(m, 0, sizeof(*m));
    char line[4096];
    int *tag2idx = NULL;
    size_t ntag = 0, nq = 0, qcap = 0;
    int status = -1;
    while (fgets(line, sizeof line, f)) {
        if (!strncmp(line, "$Nodes", 6)) {
            size_t nb, nn, mn, mx;
            if (!fgets(line, sizeof line, f) ||
                sscanf(line, "%zu %zu %zu %zu", &nb, &nn, &mn, &mx) != 4) goto done;
            m->n_nodes = nn;
            m->xyz = malloc(sizeof(double) * 3 * nn);
            ntag = mx + 1;
            tag2idx = malloc(sizeof(int) * ntag);
            for (size_t i = 0; i < ntag; ++i) tag2idx[i] = -1;
            size_t at = 0;
            for (size_t b = 0; b < nb; ++b) {
                int d, t, par; size_t cnt;
                if (!fgets(line, sizeof line, f) ||
                    sscanf(line, "%d %d %d %zu", &d, &t, &par, &cnt) != 4 || par) goto done;
                size_t *tags = malloc(sizeof(size_t) * (cnt ? cnt : 1));
                for (size_t i = 0; i < cnt; ++i) {
                    if (!fgets(line, sizeof line, f) || sscanf(line, "%zu", &tags[i]) != 1) { free(tags); goto done; }
                }
                for (size_t i = 0; i < cnt; ++i) {
                    double x, y, z;
                    if (!fgets(line, sizeof line, f) || sscanf(line, "%lf %lf %lf", &x, &y, &z) != 3 ||
                        at >= nn || tags[i] >= ntag) { free(tags); goto done; }
                    m->xyz[3 * at] = x; m->xyz[3 * at + 1] = y; m->xyz[3 * at + 2] = z;
                    tag2idx[tags[i]] = (int)at++;
                }
                free(tags);
            }
            if (at != nn) goto done;
        } else if (!strncmp(line, "$Elements", 9)) {
            size_t nb, ne, mn, mx;
            if (!fgets(line, sizeof line, f) ||
                sscanf(line, "%zu %zu %zu %zu", &nb, &ne, &mn, &mx) != 4) goto done;
            for (size_t b = 0; b < nb; ++b) {
                int d, t, type; size_t cnt;
                if (!fgets(line, sizeof line, f) ||
                    sscanf(line, "%d %d %d %zu", &d, &t, &type, &cnt) != 4) goto done;
                for (size_t i = 0; i < cnt; ++i) {
                    if (!fgets(line, sizeof line, f)) goto done;
                    if (type != 3) continue;
                    size_t et, n[4];
                    if (sscanf(line, "%zu %zu %zu %zu %zu", &et, &n[0], &n[1], &n[2], &n[3]) != 5) goto done;
                    if (nq == qcap) {
                        qcap = qcap ? 2 * qcap : 1024;
                        m->quads = realloc(m->quads, sizeof(int) * 4 * qcap);
                    }
                    for (int k = 0; k < 4; ++k) {
                        if (!tag2idx || n[k] >= ntag || tag2idx[n[k]] < 0) goto done;
                        m->quads[4 * nq + k] = tag2idx[n[k]];
                    }
                    ++nq;
                }
            }
        }
    }
    m->n_quads = nq;
    status = (m->n_nodes && nq) ? 0 : -1;
done:
    free(tag2idx);
    fclose(f);
    if (status) oracle_free_mesh(m);
    return status;
}

/* ------------------------------------------------------------------ */
/* dual mesh (dual_mesh.c:356-692)                                      */
/* ------------------------------------------------------------------ */

static int plist_push(PatchList *l, const double *p0, const double *p1,
                      const double *p2, const double *p3, int owner, int cell)
{
    if (l->count == l->capacity) {
        size_t cap = l->capacity ? 2 * l->capacity : 4;
        QuadPatch *np = realloc(l->patches, sizeof(QuadPatch) * cap);
        if (!np) return -1;
        l->patches = np;
        l->capacity = cap;
    }
    QuadPatch *q = &l->patches[l->count++];
    const double *pts[4] = {p0, p1, p2, p3};
    for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 3; ++d) q->points[3 * k + d] = pts[k][d];
    q->owner_id = owner;
    q->cell_id = cell;
    return 0;
}

/* two-triangle area, each half norm taken separately (dual_mesh.c:79-108) */
static double piece_area(const double *pt)
{
    double d1[3], d2[3], d3[3], c1[3], c2[3];
    for (int k = 0; k < 3; ++k) {
        d1[k] = pt[3 + k] - pt[k];
        d2[k] = pt[6 + k] - pt[k];
        d3[k] = pt[9 + k] - pt[k];
    }
    crossp(d1, d2, c1);
    crossp(d2, d3, c2);
    return 0.5 * len3(c1) + 0.5 * len3(c2);
}

static void region_stats(const PatchList *l, double *area, double *centre)
{
    double a = 0.0, w[3] = {0.0, 0.0, 0.0};
    for (size_t k = 0; k < l->count; ++k) {
        const double *pt = l->patches[k].points;
        double s = piece_area(pt);
        a += s;
        for (int d = 0; d < 3; ++d)
            w[d] += s * (0.25 * (pt[d] + pt[3 + d] + pt[6 + d] + pt[9 + d]));
    }
    *area = a;
    if (a > 0.0)
        for (int d = 0; d < 3; ++d) centre[d] = w[d] / a;
}

int oracle_build_dual(const Mesh *m, DualMesh *dm)
{
    memset(dm, 0, sizeof(*dm));
    if (!m || !m->cell_edges) return -1;
    dm->n_nodes = m->n_nodes;
    dm->n_edges = m->n_edges;
    dm->node_patches = calloc(m->n_nodes, sizeof(PatchList));
    dm->edge_patches = calloc(m->n_edges, sizeof(PatchList));
    dm->node_region_areas = calloc(m->n_nodes, sizeof(double));
    dm->edge_region_areas = calloc(m->n_edges, sizeof(double));
    dm->node_region_centers = calloc(3 * m->n_nodes, sizeof(double));
    dm->edge_region_centers = calloc(3 * m->n_edges, sizeof(double));
    if (!dm->node_patches || !dm->edge_patches || !dm->node_region_areas ||
        !dm->edge_region_areas || !dm->node_region_centers || !dm->edge_region_centers)
        return -1;
    for (size_t c = 0; c < m->n_quads; ++c) {
        const int *q = &m->quads[4 * c];
        double p[4][3], mid[4][3], ctr[3];
        for (int k = 0; k < 4; ++k)
            for (int d = 0; d < 3; ++d) p[k][d] = m->xyz[3 * q[k] + d];
        /* mid[k] = midpoint of side k -> k+1 */
        for (int k = 0; k < 4; ++k)
            for (int d = 0; d < 3; ++d) mid[k][d] = 0.5 * (p[k][d] + p[(k + 1) % 4][d]);
        for (int d = 0; d < 3; ++d) ctr[d] = 0.25 * (p[0][d] + p[1][d] + p[2][d] + p[3][d]);
        /* quarter-quad of corner k: [p_k, mid_k, centre, mid_{k-1}] (:446-504) */
        for (int k = 0; k < 4; ++k)
            if (plist_push(&dm->node_patches[q[k]], p[k], mid[k], ctr, mid[(k + 3) % 4], q[k], (int)c))
                return -1;
        /* half-quad of side k: [p_k, p_k+1, mid_{k+1}, mid_{k-1}] (:506-588) */
        for (int k = 0; k < 4; ++k) {
            int e = m->cell_edges[4 * c + k];
            if (plist_push(&dm->edge_patches[e], p[k], p[(k + 1) % 4], mid[(k + 1) % 4],
                           mid[(k + 3) % 4], e, (int)c))
                return -1;
        }
    }
    for (size_t v = 0; v < m->n_nodes; ++v)
        region_stats(&dm->node_patches[v], &dm->node_region_areas[v], &dm->node_region_centers[3 * v]);
    for (size_t e = 0; e < m->n_edges; ++e)
        region_stats(&dm->edge_patches[e], &dm->edge_region_areas[e], &dm->edge_region_centers[3 * e]);
    return 0;
}

void oracle_free_dual(DualMesh *dm)
{
    if (!dm) return;
    for (size_t v = 0; dm->node_patches && v < dm->n_nodes; ++v) free(dm->node_patches[v].patches);
    for (size_t e = 0; dm->edge_patches && e < dm->n_edges; ++e) free(dm->edge_patches[e].patches);
    free(dm->node_patches); free(dm->edge_patches);
    free(dm->node_region_areas); free(dm->edge_region_areas);
    free(dm->node_region_centers); free(dm->edge_region_centers);
    memset(dm, 0, sizeof(*dm));
}

/* ------------------------------------------------------------------ */
/* pair classification (quadrature.c:384-607)                           */
/* ------------------------------------------------------------------ */

static double piece_size(const double *pt)
{
    double h = 0.0;
    for (int i = 0; i < 4; ++i) {
        double l = dist3(&pt[3 * i], &pt[3 * ((i + 1) % 4)]);
        if (l > h) h = l;
    }
    return h;
}

static int pieces_identical(const double *a, const double *b)
{
    int used[4] = {0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) {
        int hit = 0;
        for (int j = 0; j < 4 && !hit; ++j)
            if (!used[j] && same_point(&a[3 * i], &b[3 * j])) used[j] = hit = 1;
        if (!hit) return 0;
    }
    return 1;
}

static int pieces_share_vertex(const double *a, const double *b)
{
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            if (same_point(&a[3 * i], &b[3 * j])) return 1;
    return 0;
}

int oracle_classify(const QuadPatch *a, const QuadPatch *b, double nf)
{
    if (!a || !b) return 3;
    if (a == b || pieces_identical(a->points, b->points)) return 0;
    if (pieces_share_vertex(a->points, b->points)) return 1;
    double g2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        double alo = a->points[k], ahi = alo, blo = b->points[k], bhi = blo;
        for (int i = 1; i < 4; ++i) {
            alo = fmin(alo, a->points[3 * i + k]); ahi = fmax(ahi, a->points[3 * i + k]);
            blo = fmin(blo, b->points[3 * i + k]); bhi = fmax(bhi, b->points[3 * i + k]);
        }
        double gap = fmax(0.0, fmax(alo - bhi, blo - ahi));
        g2 += gap * gap;
    }
    double dmin = sqrt(g2);
    double ha = piece_size(a->points), hb = piece_size(b->points);
    double h = ha > hb ? ha : hb;
    return (h > 0.0 && dmin < nf * h) ? 2 : 3;
}

/* ------------------------------------------------------------------ */
/* triangle rules (quadrature.c:632-973)                                */
/* ------------------------------------------------------------------ */

static OracleCounters *tls_counters;
#pragma omp threadprivate(tls_counters)

/* collapsed-square map: xi = u, eta = (1-u) v; J = |e1 x e2| (1-u) */
static void tri_point(const double *p0, const double *p1, const double *p2,
                      double u, double v, double r[3], double *jac)
{
    double e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    double e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    double eta = (1.0 - u) * v;
    for (int k = 0; k < 3; ++k) r[k] = p0[k] + u * e1[k] + eta * e2[k];
    double c[3];
    crossp(e1, e2, c);
    *jac = len3(c) * (1.0 - u);
}

/* regular 4-D product rule (quadrature.c:679-783) */
static double tri_pair_gl(const double (*a)[3], const double (*b)[3], int order)
{
    double x[16], w[16];
    if (oracle_gauss_legendre(order, x, w)) return NAN;
    double sum = 0.0;
    for (int i = 0; i < order; ++i) {
        double ua = 0.5 * (x[i] + 1.0), wu = 0.5 * w[i];
        for (int j = 0; j < order; ++j) {
            double va = 0.5 * (x[j] + 1.0), wv = 0.5 * w[j];
            double ra[3], ja;
            tri_point(a[0], a[1], a[2], ua, va, ra, &ja);
            for (int k = 0; k < order; ++k) {
                double ub = 0.5 * (x[k] + 1.0), wub = 0.5 * w[k];
                for (int l = 0; l < order; ++l) {
                    double vb = 0.5 * (x[l] + 1.0), wvb = 0.5 * w[l];
                    double rb[3], jb;
                    tri_point(b[0], b[1], b[2], ub, vb, rb, &jb);
                    double R = dist3(ra, rb);
                    if (tls_counters) tls_counters->far_evaluations++;
                    if (R <= 0.0) continue;
                    sum += wu * wv * wub * wvb * ja * jb / R;
                }
            }
        }
    }
    return sum;
}

/* exact potential of a uniform triangle (quadrature.c:842-898):
 * sum over edges d_i [asinh((s_i+l_i)/rho_i) - asinh(s_i/rho_i)] - |h Omega| */
typedef struct {
    double p[3][3], t[3][3], m[3][3], l[3], n[3];
    double area2;
} TriPot;

static void tripot_setup(const double (*p)[3], TriPot *T)
{
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            T->p[i][k] = p[i][k];
            T->t[i][k] = p[(i + 1) % 3][k] - p[i][k];
        }
    crossp(T->t[0], T->t[1], T->n);
    T->area2 = len3(T->n);
    if (T->area2 == 0.0) return;
    for (int k = 0; k < 3; ++k) T->n[k] /= T->area2;
    for (int i = 0; i < 3; ++i) {
        T->l[i] = len3(T->t[i]);
        for (int k = 0; k < 3; ++k) T->t[i][k] /= T->l[i];
        crossp(T->t[i], T->n, T->m[i]);
    }
}

static double tripot_eval(const TriPot *T, const double r[3])
{
    if (tls_counters) tls_counters->potential_calls++;
    double v[3][3], rad[3];
    for (int i = 0; i < 3; ++i) {
        for (int k = 0; k < 3; ++k) v[i][k] = T->p[i][k] - r[k];
        rad[i] = len3(v[i]);
    }
    double h = dotp(v[0], T->n);
    double acc = 0.0;
    for (int i = 0; i < 3; ++i) {
        double d = dotp(v[i], T->m[i]);
        if (d == 0.0) continue;
        double s = dotp(v[i], T->t[i]);
        double rho = hypot(d, h);
        acc += d * (asinh((s + T->l[i]) / rho) - asinh(s / rho));
    }
    if (h != 0.0) {
        double c[3];
        crossp(v[1], v[2], c);
        double den = rad[0] * rad[1] * rad[2] + dotp(v[0], v[1]) * rad[2]
                   + dotp(v[1], v[2]) * rad[0] + dotp(v[2], v[0]) * rad[1];
        acc -= fabs(h * (2.0 * atan2(dotp(v[0], c), den)));
    }
    return acc;
}

static double outer_rule(const double (*a)[3], const TriPot *T, int order)
{
    double x[16], w[16];
    if (oracle_gauss_legendre(order, x, w)) return NAN;
    double s = 0.0;
    for (int i = 0; i < order; ++i)
        for (int j = 0; j < order; ++j) {
            double r[3], J;
            tri_point(a[0], a[1], a[2], 0.5 * (x[i] + 1), 0.5 * (x[j] + 1), r, &J);
            s += 0.25 * w[i] * w[j] * J * tripot_eval(T, r);
        }
    return s;
}

/* adaptive outer rule with longest-edge bisection (quadrature.c:915-938) */
static double outer_adaptive(const double (*a)[3], const TriPot *T, int order,
                             double tol, int depth)
{
    double lo = outer_rule(a, T, order <= 6 ? 3 : 4);
    double hi = outer_rule(a, T, order);
    if (!isfinite(lo) || !isfinite(hi)) return NAN;
    if (fabs(hi - lo) <= tol + 2e-5 * fabs(hi)) return hi;
    if (depth >= 24) return NAN;
    int e = 0;
    double best = 0.0;
    for (int i = 0; i < 3; ++i) {
        double l = dist3(a[i], a[(i + 1) % 3]);
        if (l > best) { best = l; e = i; }
    }
    int j = (e + 1) % 3, k = (e + 2) % 3;
    double L[3][3], R[3][3];
    for (int d = 0; d < 3; ++d) {
        double mid = 0.5 * (a[e][d] + a[j][d]);
        L[0][d] = a[e][d]; L[1][d] = mid;    L[2][d] = a[k][d];
        R[0][d] = mid;     R[1][d] = a[j][d]; R[2][d] = a[k][d];
    }
    return outer_adaptive((const double(*)[3])L, T, order, tol * 0.5, depth + 1)
         + outer_adaptive((const double(*)[3])R, T, order, tol * 0.5, depth + 1);
}

/* one triangle pair (quadrature.c:940-973); cls = classification or 3 = FAR */
static double tri_pair(const double (*a)[3], const double (*b)[3], int cls,
                       int far_o, int near_o, int touch_o, int self_o)
{
    if (cls == 3) return tri_pair_gl(a, b, far_o);
    if (tls_counters) tls_counters->near_triangle_pairs++;
    TriPot Ta, Tb;
    tripot_setup(a, &Ta);
    tripot_setup(b, &Tb);
    if (Ta.area2 == 0.0 || Tb.area2 == 0.0) return 0.0;
    int same = 1;
    for (int i = 0; i < 3; ++i) {
        int hit = 0;
        for (int j = 0; j < 3; ++j) if (same_point(a[i], b[j])) hit = 1;
        if (!hit) same = 0;
    }
    if (same) {
        /* K(T,T) = (2A/3) sum_vertices Phi_T(v) */
        double s = 0.0;
        for (int i = 0; i < 3; ++i) s += tripot_eval(&Tb, a[i]);
        return Ta.area2 * s / 3.0;
    }
    int order = cls == 0 ? self_o : cls == 1 ? touch_o : near_o;
    if (order < 6) order = 6;
    double scale = Ta.area2 * Tb.area2 / sqrt(fmax(Ta.area2, Tb.area2));
    double ab = outer_adaptive(a, &Tb, order, 2e-5 * scale, 0);
    double ba = outer_adaptive(b, &Ta, order, 2e-5 * scale, 0);
    return 0.5 * (ab + ba);
}

/* split of a piece into triangles (0,1,2), (0,2,3) */
static void piece_triangles(const double *pt, double t[2][3][3])
{
    static const int idx[2][3] = {{0, 1, 2}, {0, 2, 3}};
    for (int s = 0; s < 2; ++s)
        for (int v = 0; v < 3; ++v)
            for (int d = 0; d < 3; ++d) t[s][v][d] = pt[3 * idx[s][v] + d];
}

/* quadrature.c:975-1079 */
double oracle_patch_kernel(const QuadPatch *a, const QuadPatch *b, int far_o,
                           int near_o, int touch_o, int self_o, double nf)
{
    if (!a || !b) return NAN;
    if (!gl_supported(far_o) || !gl_supported(near_o) || !gl_supported(touch_o) ||
        !gl_supported(self_o) || !isfinite(nf) || nf < 0.0)
        return NAN;
    double ta[2][3][3], tb[2][3][3];
    piece_triangles(a->points, ta);
    piece_triangles(b->points, tb);
    int cls = oracle_classify(a, b, nf);
    if (tls_counters) {
        if (cls == 3) tls_counters->far_patch_pairs++;
        else tls_counters->near_patch_pairs++;
    }
    double total = 0.0;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            double v = tri_pair((const double(*)[3])ta[i], (const double(*)[3])tb[j], cls,
                                far_o, near_o, touch_o, self_o);
            if (!isfinite(v)) return NAN;
            total += v;
        }
    return total;
}

/* ------------------------------------------------------------------ */
/* Lp and P matrices (inductance.c:221-560, potential.c:139-395)        */
/* ------------------------------------------------------------------ */

static double region_integral(const PatchList *ra, const PatchList *rb, int far_o,
                              int near_o, int touch_o, int self_o, double nf)
{
    double s = 0.0;
    for (size_t i = 0; i < ra->count; ++i)
        for (size_t j = 0; j < rb->count; ++j) {
            double v = oracle_patch_kernel(&ra->patches[i], &rb->patches[j], far_o, near_o,
                                           touch_o, self_o, nf);
            if (!isfinite(v)) return NAN;
            s += v;
        }
    return s;
}

static void add_counters(OracleCounters *dst, const OracleCounters *src)
{
    dst->far_evaluations += src->far_evaluations;
    dst->potential_calls += src->potential_calls;
    dst->near_triangle_pairs += src->near_triangle_pairs;
    dst->far_patch_pairs += src->far_patch_pairs;
    dst->near_patch_pairs += src->near_patch_pairs;
}

int oracle_inductance_matrix(const Mesh *m, const DualMesh *dm,
                             const InductanceOptions *o, double *L,
                             OracleCounters *counters)
{
    size_t ne = m->n_edges;
    int bad = 0;
    if (counters) memset(counters, 0, sizeof(*counters));
#pragma omp parallel
    {
        OracleCounters mine;
        memset(&mine, 0, sizeof mine);
        tls_counters = counters ? &mine : NULL;
#pragma omp for schedule(dynamic)
        for (long long sa = 0; sa < (long long)ne; ++sa) {
            size_t a = (size_t)sa;
            for (size_t b = a; b < ne; ++b) {
                const double *da = &m->edge_directions[3 * a], *db = &m->edge_directions[3 * b];
                double dot = dotp(da, db);
                double v = 0.0;
                if (!(fabs(dot) < o->direction_tolerance)) {
                    double sa_ = dm->edge_region_areas[a], la = m->edge_lengths[a];
                    double sb_ = dm->edge_region_areas[b], lb = m->edge_lengths[b];
                    double wa = (sa_ <= 0.0 || la <= 0.0) ? 0.0 : sa_ / la;
                    double wb = (sb_ <= 0.0 || lb <= 0.0) ? 0.0 : sb_ / lb;
                    double g = region_integral(&dm->edge_patches[a], &dm->edge_patches[b],
                                               o->far_order, o->near_order, o->touching_order,
                                               o->self_order, o->near_factor);
                    v = 1.0e-7 * dot * g / (wa * wb);
                    if (wa <= 0.0 || wb <= 0.0 || !isfinite(v)) {
#pragma omp atomic write
                        bad = 1;
                    }
                }
                L[a * ne + b] = v;
                L[b * ne + a] = v;
            }
        }
        if (counters) {
#pragma omp critical(oracle_counters)
            add_counters(counters, &mine);
        }
        tls_counters = NULL;
    }
    return bad ? -1 : 0;
}

int oracle_potential_matrix(const Mesh *m, const DualMesh *dm,
                            const PotentialOptions *o, double *P,
                            OracleCounters *counters)
{
    size_t nv = m->n_nodes;
    int bad = 0;
    if (counters) memset(counters, 0, sizeof(*counters));
#pragma omp parallel
    {
        OracleCounters mine;
        memset(&mine, 0, sizeof mine);
        tls_counters = counters ? &mine : NULL;
#pragma omp for schedule(dynamic)
        for (long long si = 0; si < (long long)nv; ++si) {
            size_t i = (size_t)si;
            for (size_t j = i; j < nv; ++j) {
                double Si = dm->node_region_areas[i], Sj = dm->node_region_areas[j];
                double g = region_integral(&dm->node_patches[i], &dm->node_patches[j],
                                           o->far_order, o->near_order, o->touching_order,
                                           o->self_order, o->near_factor);
                double v = 8.987551792261171e9 * g / (Si * Sj);
                if (Si <= 0.0 || Sj <= 0.0 || !isfinite(v)) {
#pragma omp atomic write
                    bad = 1;
                }
                P[i * nv + j] = v;
                P[j * nv + i] = v;
            }
        }
        if (counters) {
#pragma omp critical(oracle_counters)
            add_counters(counters, &mine);
        }
        tls_counters = NULL;
    }
    return bad ? -1 : 0;
}

/* incidence.c:417-590: PAT[j,e] = P[j,b_e] - P[j,a_e] */
int oracle_pat(const Mesh *m, const double *P, double *PAT)
{
    size_t nv = m->n_nodes, ne = m->n_edges;
    for (size_t j = 0; j < nv; ++j)
        for (size_t e = 0; e < ne; ++e)
            PAT[j * ne + e] = P[j * nv + m->edges[2 * e + 1]] - P[j * nv + m->edges[2 * e]];
    return 0;
}

/* harmonic.c:175-372 */
int oracle_mna(const Mesh *m, const double *L, const double *R, const double *PAT,
               const Complex *Lc, const Complex *Pc, double f, Complex *Z)
{
    size_t ne = m->n_edges, nv = m->n_nodes, n = ne + nv;
    if (!(f > 0.0) || !isfinite(f)) return -1;
    double w = 2.0 * M_PI * f;
    memset(Z, 0, sizeof(Complex) * n * n);
    for (size_t i = 0; i < ne; ++i)
        for (size_t j = 0; j < ne; ++j) {
            Complex z;
            if (Lc) {
                /* jw (Lr + j Li) */
                z.re = (i == j && R ? R[i] : 0.0) - w * Lc[i * ne + j].im;
                z.im = w * Lc[i * ne + j].re;
            } else {
                z.re = (i == j && R) ? R[i] : 0.0;
                z.im = w * L[i * ne + j];
            }
            Z[i * n + j] = z;
        }
    for (size_t e = 0; e < ne; ++e) {
        Z[e * n + ne + m->edges[2 * e]].re = -1.0;
        Z[e * n + ne + m->edges[2 * e + 1]].re = 1.0;
    }
    for (size_t i = 0; i < nv; ++i) {
        for (size_t e = 0; e < ne; ++e) {
            Complex z = {0.0, 0.0};
            if (Pc) {
                Complex pb = Pc[i * nv + m->edges[2 * e + 1]], pa = Pc[i * nv + m->edges[2 * e]];
                z.re = -(pb.re - pa.re);
                z.im = -(pb.im - pa.im);
            } else {
                z.re = -PAT[i * ne + e];
            }
            Z[(ne + i) * n + e] = z;
        }
        Z[(ne + i) * n + ne + i].im = w;
    }
    return 0;
}

/* harmonic.c:436-552 */
int oracle_rhs(size_t ne, size_t nv, const double *Pr, const Complex *Pc,
               const Complex *U, const Complex *Is, Complex *rhs)
{
    for (size_t i = 0; i < ne; ++i) {
        rhs[i].re = U ? -U[i].re : 0.0;
        rhs[i].im = U ? -U[i].im : 0.0;
    }
    for (size_t i = 0; i < nv; ++i) {
        double sr = 0.0, si = 0.0;
        if (Is) {
            for (size_t j = 0; j < nv; ++j) {
                if (Pc) {
                    Complex p = Pc[i * nv + j];
                    sr += p.re * Is[j].re - p.im * Is[j].im;
                    si += p.re * Is[j].im + p.im * Is[j].re;
                } else {
                    sr += Pr[i * nv + j] * Is[j].re;
                    si += Pr[i * nv + j] * Is[j].im;
                }
            }
        }
        rhs[ne + i].re = sr;
        rhs[ne + i].im = si;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* complex LU (complex_lu.c:208-778)                                    */
/* ------------------------------------------------------------------ */

static Complex cdiv(Complex a, Complex b)
{
    double den = b.re * b.re + b.im * b.im;
    Complex z = {(a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den};
    return z;
}

int oracle_complex_lu(const Complex *A, size_t n, double tol, Complex *LU,
                      size_t *piv, double *rs, double *cs, double *mscale)
{
    if (!A || !n || !isfinite(tol) || tol < 0.0) return -1;
    for (size_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (size_t j = 0; j < n; ++j) {
            double mag = hypot(A[i * n + j].re, A[i * n + j].im);
            if (!isfinite(mag)) return -2;
            s = fmax(s, mag);
        }
        if (s == 0.0) return -2;
        rs[i] = s;
        for (size_t j = 0; j < n; ++j) {
            LU[i * n + j].re = A[i * n + j].re / s;
            LU[i * n + j].im = A[i * n + j].im / s;
        }
    }
    for (size_t j = 0; j < n; ++j) {
        double s = 0.0;
        for (size_t i = 0; i < n; ++i) s = fmax(s, hypot(LU[i * n + j].re, LU[i * n + j].im));
        if (s == 0.0) return -2;
        cs[j] = s;
        for (size_t i = 0; i < n; ++i) {
            LU[i * n + j].re /= s;
            LU[i * n + j].im /= s;
        }
    }
    double big = 0.0;
    for (size_t k = 0; k < n * n; ++k) {
        double mag = hypot(LU[k].re, LU[k].im);
        if (mag > big) big = mag;
    }
    *mscale = big;
    if (!isfinite(big) || big <= 0.0) return -2;
    double limit = tol * big;
    for (size_t k = 0; k < n; ++k) {
        size_t p = k;
        double pa = hypot(LU[k * n + k].re, LU[k * n + k].im);
        for (size_t i = k + 1; i < n; ++i) {
            double mag = hypot(LU[i * n + k].re, LU[i * n + k].im);
            if (mag > pa) { pa = mag; p = i; }
        }
        if (!isfinite(pa) || pa <= limit) return -2;
        piv[k] = p;
        if (p != k)
            for (size_t j = 0; j < n; ++j) {
                Complex t = LU[k * n + j]; LU[k * n + j] = LU[p * n + j]; LU[p * n + j] = t;
            }
        if (k + 1 >= n) continue;
        Complex pv = LU[k * n + k];
#pragma omp parallel for schedule(static) if (n - k - 1 >= 64)
        for (long long si = (long long)k + 1; si < (long long)n; ++si) {
            size_t i = (size_t)si;
            Complex l = cdiv(LU[i * n + k], pv);
            LU[i * n + k] = l;
            for (size_t j = k + 1; j < n; ++j) {
                Complex u = LU[k * n + j];
                double pr = l.re * u.re - l.im * u.im;
                double pi = l.re * u.im + l.im * u.re;
                LU[i * n + j].re -= pr;
                LU[i * n + j].im -= pi;
            }
        }
    }
    for (size_t i = 0; i < n; ++i) {
        double mag = hypot(LU[i * n + i].re, LU[i * n + i].im);
        if (!isfinite(mag) || mag <= limit) return -2;
    }
    return 0;
}

int oracle_complex_solve(const Complex *LU, const size_t *piv, const double *rs,
                         const double *cs, size_t n, const Complex *b, Complex *x)
{
    for (size_t i = 0; i < n; ++i) {
        x[i].re = b[i].re / rs[i];
        x[i].im = b[i].im / rs[i];
    }
    for (size_t k = 0; k < n; ++k)
        if (piv[k] != k) { Complex t = x[k]; x[k] = x[piv[k]]; x[piv[k]] = t; }
    for (size_t i = 0; i < n; ++i) {
        double cr = 0.0, ci = 0.0;
        for (size_t j = 0; j < i; ++j) {
            Complex l = LU[i * n + j];
            cr += l.re * x[j].re - l.im * x[j].im;
            ci += l.re * x[j].im + l.im * x[j].re;
        }
        x[i].re -= cr;
        x[i].im -= ci;
    }
    for (size_t ii = n; ii-- > 0;) {
        double cr = 0.0, ci = 0.0;
        for (size_t j = ii + 1; j < n; ++j) {
            Complex u = LU[ii * n + j];
            cr += u.re * x[j].re - u.im * x[j].im;
            ci += u.re * x[j].im + u.im * x[j].re;
        }
        Complex r = {x[ii].re - cr, x[ii].im - ci};
        Complex d = LU[ii * n + ii];
        double den = d.re * d.re + d.im * d.im;
        if (!isfinite(den) || den == 0.0) return -1;
        x[ii].re = (r.re * d.re + r.im * d.im) / den;
        x[ii].im = (r.im * d.re - r.re * d.im) / den;
    }
    for (size_t i = 0; i < n; ++i) {
        x[i].re /= cs[i];
        x[i].im /= cs[i];
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* real LU (lu.c:194-650)                                               */
/* ------------------------------------------------------------------ */

int oracle_real_lu(const double *A, size_t n, double tol, double *LU, size_t *piv,
                   double *mscale)
{
    if (!A || !n) return -1;
    memcpy(LU, A, sizeof(double) * n * n);
    double big = 0.0;
    for (size_t k = 0; k < n * n; ++k) if (fabs(A[k]) > big) big = fabs(A[k]);
    *mscale = big;
    if (!isfinite(big) || big <= 0.0) return -2;
    double limit = tol * big;
    for (size_t k = 0; k < n; ++k) {
        size_t p = k;
        double pa = fabs(LU[k * n + k]);
        for (size_t i = k + 1; i < n; ++i)
            if (fabs(LU[i * n + k]) > pa) { pa = fabs(LU[i * n + k]); p = i; }
        if (!isfinite(pa) || pa <= limit) return -2;
        piv[k] = p;
        if (p != k)
            for (size_t j = 0; j < n; ++j) { double t = LU[k * n + j]; LU[k * n + j] = LU[p * n + j]; LU[p * n + j] = t; }
        if (k + 1 >= n) continue;
        double pv = LU[k * n + k];
        for (size_t i = k + 1; i < n; ++i) {
            double l = LU[i * n + k] / pv;
            LU[i * n + k] = l;
            for (size_t j = k + 1; j < n; ++j) LU[i * n + j] -= l * LU[k * n + j];
        }
    }
    for (size_t i = 0; i < n; ++i)
        if (!isfinite(LU[i * n + i]) || fabs(LU[i * n + i]) <= limit) return -2;
    return 0;
}

int oracle_real_solve(const double *LU, const size_t *piv, size_t n, const double *b, double *x)
{
    if (x != b) memcpy(x, b, sizeof(double) * n);
    for (size_t k = 0; k < n; ++k)
        if (piv[k] != k) { double t = x[k]; x[k] = x[piv[k]]; x[piv[k]] = t; }
    for (size_t i = 0; i < n; ++i) {
        double c = 0.0;
        for (size_t j = 0; j < i; ++j) c += LU[i * n + j] * x[j];
        x[i] -= c;
    }
    for (size_t ii = n; ii-- > 0;) {
        double c = 0.0;
        for (size_t j = ii + 1; j < n; ++j) c += LU[ii * n + j] * x[j];
        double d = LU[ii * n + ii];
        if (d == 0.0 || !isfinite(d)) return -1;
        x[ii] = (x[ii] - c) / d;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* component-wise backward error (harmonic.c:720-1080)                   */
/* ------------------------------------------------------------------ */

int oracle_backward_error(const Complex *Z, const Complex *x, const Complex *b,
                          size_t n, size_t ne, double *out)
{
    double absr = 0, binf = 0, eabs = 0, nabs = 0, be = 0, ebe = 0, nbe = 0;
    for (size_t i = 0; i < n; ++i) {
        double yr = 0.0, yi = 0.0, den = 0.0;
        for (size_t j = 0; j < n; ++j) {
            Complex z = Z[i * n + j];
            yr += z.re * x[j].re - z.im * x[j].im;
            yi += z.re * x[j].im + z.im * x[j].re;
        }
        double ra = hypot(yr - b[i].re, yi - b[i].im);
        double ba = hypot(b[i].re, b[i].im);
        den = ba;
        for (size_t j = 0; j < n; ++j)
            den += hypot(Z[i * n + j].re, Z[i * n + j].im) * hypot(x[j].re, x[j].im);
        double eta = den > 0.0 ? ra / den : ra;
        absr = fmax(absr, ra);
        binf = fmax(binf, ba);
        be = fmax(be, eta);
        if (i < ne) { eabs = fmax(eabs, ra); ebe = fmax(ebe, eta); }
        else { nabs = fmax(nabs, ra); nbe = fmax(nbe, eta); }
    }
    out[0] = absr;
    out[1] = binf > 0.0 ? absr / binf : absr;
    out[2] = be;
    out[3] = eabs;
    out[4] = nabs;
    out[5] = ebe;
    out[6] = nbe;
    return 0;
}

/* ------------------------------------------------------------------ */
/* retarded kernel exp(-jkR)/R = 1/R + (exp(-jkR)-1)/R                  */
/* (no reference implementation: include/retardation.h is empty).      */
/* The smooth remainder g(R) = (-2 sin^2(kR/2) - j sin(kR)) / R, with    */
/* g(0) = -jk, is integrated by the smooth_order GL product rule over    */
/* every triangle pair of every piece pair.                              */
/* ------------------------------------------------------------------ */

static void smooth_pair(const double *pa, const double *pb, double k, int order,
                        double *sre, double *sim)
{
    double x[16], w[16];
    oracle_gauss_legendre(order, x, w);
    double ta[2][3][3], tb[2][3][3];
    piece_triangles(pa, ta);
    piece_triangles(pb, tb);
    double accr = 0.0, acci = 0.0;
    for (int s = 0; s < 2; ++s)
        for (int t = 0; t < 2; ++t)
            for (int i = 0; i < order; ++i)
                for (int j = 0; j < order; ++j) {
                    double ra[3], ja;
                    tri_point(ta[s][0], ta[s][1], ta[s][2], 0.5 * (x[i] + 1), 0.5 * (x[j] + 1), ra, &ja);
                    double wa = 0.25 * w[i] * w[j] * ja;
                    for (int u = 0; u < order; ++u)
                        for (int v = 0; v < order; ++v) {
                            double rb[3], jb;
                            tri_point(tb[t][0], tb[t][1], tb[t][2], 0.5 * (x[u] + 1), 0.5 * (x[v] + 1), rb, &jb);
                            double wb = 0.25 * w[u] * w[v] * jb;
                            double R = dist3(ra, rb);
                            double gr, gi;
                            if (R > 0.0) {
                                double h = sin(0.5 * k * R);
                                gr = -2.0 * h * h / R;
                                gi = -sin(k * R) / R;
                            } else {
                                gr = 0.0;
                                gi = -k;
                            }
                            accr += wa * wb * gr;
                            acci += wa * wb * gi;
                        }
                }
    *sre = accr;
    *sim = acci;
}

/* Entries whose two regions come closer than near_factor x (largest piece
 * side of either region), by their axis-aligned bounding boxes, use the
 * near_order rule for the remainder (its R-kink at coincident points makes
 * the low-order rule converge slowly there); all others use `order`. */
static void region_box(const PatchList *l, double lo[3], double hi[3], double *h)
{
    for (int k = 0; k < 3; ++k) { lo[k] = HUGE_VAL; hi[k] = -HUGE_VAL; }
    *h = 0.0;
    for (size_t p = 0; p < l->count; ++p) {
        const double *pv = l->patches[p].points;
        for (int k = 0; k < 3; ++k)
            for (int i = 0; i < 4; ++i) {
                lo[k] = fmin(lo[k], pv[3 * i + k]);
                hi[k] = fmax(hi[k], pv[3 * i + k]);
            }
        *h = fmax(*h, piece_size(pv));
    }
}

static int regions_near(const PatchList *a, const PatchList *b, double nf)
{
    double alo[3], ahi[3], blo[3], bhi[3], ha, hb;
    region_box(a, alo, ahi, &ha);
    region_box(b, blo, bhi, &hb);
    double g2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        double g = fmax(0.0, fmax(alo[k] - bhi[k], blo[k] - ahi[k]));
        g2 += g * g;
    }
    double h = fmax(ha, hb);
    return h > 0.0 && sqrt(g2) < nf * h;
}

int oracle_retarded_matrices(const Mesh *m, const DualMesh *dm,
                             const InductanceOptions *lo, const PotentialOptions *po,
                             double k, int order, int near_order, const double *Ls,
                             const double *Ps, Complex *L, Complex *P)
{
    if (!gl_supported(order) || !gl_supported(near_order)) return -1;
    size_t ne = m->n_edges, nv = m->n_nodes;
    if (L) {
#pragma omp parallel for schedule(dynamic)
        for (long long sa = 0; sa < (long long)ne; ++sa) {
            size_t a = (size_t)sa;
            for (size_t b = a; b < ne; ++b) {
                double dot = dotp(&m->edge_directions[3 * a], &m->edge_directions[3 * b]);
                Complex v = {0.0, 0.0};
                if (!(fabs(dot) < lo->direction_tolerance)) {
                    double wa = dm->edge_region_areas[a] / m->edge_lengths[a];
                    double wb = dm->edge_region_areas[b] / m->edge_lengths[b];
                    double gr = 0.0, gi = 0.0;
                    const PatchList *ra = &dm->edge_patches[a], *rb = &dm->edge_patches[b];
                    int q = regions_near(ra, rb, lo->near_factor) ? near_order : order;
                    for (size_t i = 0; i < ra->count; ++i)
                        for (size_t j = 0; j < rb->count; ++j) {
                            double sr, si;
                            smooth_pair(ra->patches[i].points, rb->patches[j].points, k, q, &sr, &si);
                            gr += sr;
                            gi += si;
                        }
                    double c = 1.0e-7 * dot / (wa * wb);
                    v.re = Ls[a * ne + b] + c * gr;
                    v.im = c * gi;
                }
                L[a * ne + b] = v;
                L[b * ne + a] = v;
            }
        }
    }
    if (P) {
#pragma omp parallel for schedule(dynamic)
        for (long long si_ = 0; si_ < (long long)nv; ++si_) {
            size_t i = (size_t)si_;
            for (size_t j = i; j < nv; ++j) {
                double gr = 0.0, gi = 0.0;
                const PatchList *ra = &dm->node_patches[i], *rb = &dm->node_patches[j];
                int q = regions_near(ra, rb, po->near_factor) ? near_order : order;
                for (size_t x = 0; x < ra->count; ++x)
                    for (size_t y = 0; y < rb->count; ++y) {
                        double sr, s2;
                        smooth_pair(ra->patches[x].points, rb->patches[y].points, k, q, &sr, &s2);
                        gr += sr;
                        gi += s2;
                    }
                double c = 8.987551792261171e9 / (dm->node_region_areas[i] * dm->node_region_areas[j]);
                Complex v = {Ps[i * nv + j] + c * gr, c * gi};
                P[i * nv + j] = v;
                P[j * nv + i] = v;
            }
        }
    }
    return 0;
}


/* ======================================================================
 * plane-wave excitation and radar cross section
 * ====================================================================== */

static const double ORACLE_C0 = 299792458.0;       /* incident_field.c:18 */
static const double ORACLE_MU0 = 1.25663706212e-6; /* rcs.c:23 */
static const double ORACLE_EPS0 = 8.8541878128e-12; /* rcs.c:26 */

static double o_dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double o_norm3(const double a[3]) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

static int o_normalize(double a[3])
{
    double nrm = o_norm3(a);
    if (!isfinite(nrm) || nrm <= 1.0e-14) return -1;
    a[0] /= nrm;
    a[1] /= nrm;
    a[2] /= nrm;
    return 0;
}

/* incident_field.c:110-283,312-420 */
int oracle_plane_wave(IncidentField *f, int wave_type, int polarization, double E_amp, double frequency,
                      double theta_deg, double phi_deg, double phase)
{
    if (f == NULL || wave_type < 0 || wave_type > 2 || !(frequency > 0.0)) return -1;
    memset(f, 0, sizeof(*f));
    f->wave_type = (IncidentWaveType)wave_type;
    f->polarization = (IncidentPolarization)polarization;
    f->E_amp = E_amp;
    f->frequency = frequency;
    f->theta_deg = theta_deg;
    f->phi_deg = phi_deg;
    f->phase = phase;
    f->theta = theta_deg * M_PI / 180.0;
    f->phi = phi_deg * M_PI / 180.0;
    f->omega = 2.0 * M_PI * frequency;
    f->wave_number = f->omega / ORACLE_C0;
    const double t = f->theta, p = f->phi;
    f->k_hat[0] = -sin(t) * cos(p);
    f->k_hat[1] = -sin(t) * sin(p);
    f->k_hat[2] = -cos(t);
    f->e_theta[0] = cos(t) * cos(p);
    f->e_theta[1] = cos(t) * sin(p);
    f->e_theta[2] = -sin(t);
    f->e_phi[0] = -sin(p);
    f->e_phi[1] = cos(p);
    f->e_phi[2] = 0.0;
    if (o_normalize(f->k_hat) || o_normalize(f->e_theta) || o_normalize(f->e_phi)) return -1;
    if (fabs(o_dot3(f->k_hat, f->e_theta)) > 1.0e-12 || fabs(o_dot3(f->k_hat, f->e_phi)) > 1.0e-12) return -1;
    const double *pol;
    if (polarization == 1) pol = f->e_theta;
    else if (polarization == 0) pol = f->e_phi;
    else return -1;
    for (int c = 0; c < 3; ++c) f->E0[c] = E_amp * pol[c];
    return 0;
}

/* incident_field.c:754-858 */
int oracle_phasor(const IncidentField *f, const double r[3], Complex E[3])
{
    if (f->wave_type > 2) return -1;
    double gamma = -f->wave_number * o_dot3(f->k_hat, r) + f->phase;
    if (f->wave_type == 1) gamma -= M_PI / 2.0;
    const double c = cos(gamma), s = sin(gamma);
    for (int k = 0; k < 3; ++k) {
        E[k].re = f->E0[k] * c;
        E[k].im = f->E0[k] * s;
    }
    return 0;
}

/* scattering.c:23-43 */
static double o_sinc(double x)
{
    double ax = fabs(x);
    if (ax < 1.0e-4) {
        double x2 = x * x, x4 = x2 * x2;
        return 1.0 - x2 / 6.0 + x4 / 120.0;
    }
    return sin(x) / x;
}

/* scattering.c:180-337 (edge loop of 338-489) */
int oracle_excitation(const Mesh *mesh, const IncidentField *f, int integration, Complex *U)
{
    if (mesh == NULL || f == NULL || U == NULL || mesh->n_edges == 0) return -1;
    if (f->wave_type == 3 || (integration != 0 && integration != 1)) return -1;
    for (size_t e = 0; e < mesh->n_edges; ++e) {
        const double *rc = &mesh->edge_centers[3 * e];
        const double *l = &mesh->edge_vectors[3 * e];
        Complex E[3];
        if (oracle_phasor(f, rc, E)) return -1;
        Complex mid;
        mid.re = E[0].re * l[0] + E[1].re * l[1] + E[2].re * l[2];
        mid.im = E[0].im * l[0] + E[1].im * l[1] + E[2].im * l[2];
        if (integration == 1) {
            double q = f->wave_number * o_dot3(f->k_hat, l);
            double factor = o_sinc(0.5 * q);
            mid.re *= factor;
            mid.im *= factor;
        }
        if (!isfinite(mid.re) || !isfinite(mid.im)) return -1;
        U[e] = mid;
    }
    return 0;
}

/* rcs.c:157-362: G_T = integral over triangle (p0,p1,p2) of exp(j k d.r) dS */
static void o_triangle_phase(const double *p0, const double *p1, const double *p2, double k, const double d[3],
                             int order, const double *x, const double *w, double *gre, double *gim)
{
    double a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    double b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    double cr[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    *gre = 0.0;
    *gim = 0.0;
    for (int iu = 0; iu < order; ++iu) {
        double u = 0.5 * (x[iu] + 1.0), wu = 0.5 * w[iu];
        for (int iv = 0; iv < order; ++iv) {
            double v = 0.5 * (x[iv] + 1.0), wv = 0.5 * w[iv];
            double xi = u, eta = (1.0 - u) * v;
            double r[3] = {p0[0] + xi * a[0] + eta * b[0], p0[1] + xi * a[1] + eta * b[1],
                           p0[2] + xi * a[2] + eta * b[2]};
            double jac = o_norm3(cr) * (1.0 - u);
            double phase = k * o_dot3(d, r);
            double weight = wu * wv * jac;
            *gre += weight * cos(phase);
            *gim += weight * sin(phase);
        }
    }
}

/* rcs.c:701-995 (sequential edge order) */
int oracle_radiation_vector(const Mesh *mesh, const DualMesh *dual, const Complex *I, double k,
                            const double d[3], int integration, int order, Complex F[3])
{
    double x[16], w[16];
    if (integration == 1 && oracle_gauss_legendre(order, x, w) != 0) return -1;
    double fr[3] = {0, 0, 0}, fi[3] = {0, 0, 0};
    for (size_t e = 0; e < mesh->n_edges; ++e) {
        const double *lv = &mesh->edge_vectors[3 * e];
        double len = o_norm3(lv);
        if (len <= 0.0) continue;
        double ed[3] = {lv[0] / len, lv[1] / len, lv[2] / len};
        double sre, sim;
        if (integration == 0) {
            double phase = k * o_dot3(d, &mesh->edge_centers[3 * e]);
            double er = cos(phase), ei = sin(phase);
            sre = I[e].re * er - I[e].im * ei;
            sim = I[e].re * ei + I[e].im * er;
            sre *= len;
            sim *= len;
        } else {
            double area = dual->edge_region_areas[e];
            if (area <= 0.0) continue;
            double width = area / len;
            double gre = 0.0, gim = 0.0;
            const PatchList *list = &dual->edge_patches[e];
            for (size_t p = 0; p < list->count; ++p) {
                const double *pt = list->patches[p].points;
                double r0, i0, r1, i1;
                o_triangle_phase(pt, pt + 3, pt + 6, k, d, order, x, w, &r0, &i0);
                o_triangle_phase(pt, pt + 6, pt + 9, k, d, order, x, w, &r1, &i1);
                gre += r0 + r1;
                gim += i0 + i1;
            }
            double cre = I[e].re / width, cim = I[e].im / width;
            sre = cre * gre - cim * gim;
            sim = cre * gim + cim * gre;
        }
        for (int c = 0; c < 3; ++c) {
            fr[c] += sre * ed[c];
            fi[c] += sim * ed[c];
        }
    }
    for (int c = 0; c < 3; ++c) {
        F[c].re = fr[c];
        F[c].im = fi[c];
    }
    return 0;
}

/* rcs.c:536-598,638-700,1000-1165 */
int oracle_rcs_direction(const Mesh *mesh, const DualMesh *dual, const IncidentField *f, const Complex *I,
                         int integration, int order, double theta_deg, double phi_deg, RcsResult *out)
{
    memset(out, 0, sizeof(*out));
    out->theta_deg = theta_deg;
    out->phi_deg = phi_deg;
    double t = theta_deg * M_PI / 180.0, p = phi_deg * M_PI / 180.0;
    out->direction[0] = sin(t) * cos(p);
    out->direction[1] = sin(t) * sin(p);
    out->direction[2] = cos(t);
    double nrm = o_norm3(out->direction);
    if (!isfinite(nrm) || nrm <= 0.0) return -1;
    for (int c = 0; c < 3; ++c) out->direction[c] /= nrm;
    if (oracle_radiation_vector(mesh, dual, I, f->wave_number, out->direction, integration, order,
                                out->radiation_vector))
        return -1;
    const double *d = out->direction;
    const Complex *F = out->radiation_vector;
    double dre = d[0] * F[0].re + d[1] * F[1].re + d[2] * F[2].re;
    double dim = d[0] * F[0].im + d[1] * F[1].im + d[2] * F[2].im;
    double t2 = 0.0;
    for (int c = 0; c < 3; ++c) {
        out->transverse_vector[c].re = d[c] * dre - F[c].re;
        out->transverse_vector[c].im = d[c] * dim - F[c].im;
    }
    for (int c = 0; c < 3; ++c)
        t2 += out->transverse_vector[c].re * out->transverse_vector[c].re +
              out->transverse_vector[c].im * out->transverse_vector[c].im;
    double eta0 = sqrt(ORACLE_MU0 / ORACLE_EPS0);
    double k = f->wave_number, E0 = f->E_amp;
    out->sigma = k * k * eta0 * eta0 * t2 / (4.0 * M_PI * E0 * E0);
    out->sigma_dbsm = out->sigma > DBL_MIN ? 10.0 * log10(out->sigma) : -INFINITY;
    return 0;
}

/* ------------------------------------------------------------------ */
/* theta-scheme transient (transient.c:334-1749, stages 1 and 2)        */
/* ------------------------------------------------------------------ */

static double oracle_theta(int scheme)
{
    return scheme == 1 ? 1.0 : scheme == 2 ? 0.5 : NAN;
}

int oracle_transient_matrix(const double *L, const double *R, const double *A, const double *P, size_t ne,
                            size_t nv, int scheme, double dt, long shunt_a, long shunt_b, double r_shunt,
                            double *M, double *F, double *AS, double *bR, double *c)
{
    if (!L || !R || !A || !P || !M || ne == 0 || nv == 0 || !isfinite(dt) || dt <= 0.0) return -1;
    if (scheme != 1 && scheme != 2) return -1;
    const int st2 = shunt_b >= 0;
    const double theta = oracle_theta(scheme);
    const size_t nt = ne + nv + (st2 ? 1 : 0);
    double *f = malloc(sizeof(double) * nv), *as = malloc(sizeof(double) * ne * nv);
    double *b = calloc(nv, sizeof(double)), *cc = calloc(nv, sizeof(double));
    int rc = -1;
    if (!f || !as || !b || !cc) goto done;
    /* F_j = 1 / P_jj (potential_build_self_capacitance) */
    for (size_t j = 0; j < nv; ++j) {
        const double d = P[j * nv + j];
        if (!isfinite(d) || d <= 0.0) goto done;
        f[j] = 1.0 / d;
    }
    /* AS[e,a] = (P[plus,a] - P[minus,a]) / P_aa, the last -1 / +1 of a row */
    for (size_t e = 0; e < ne; ++e) {
        size_t mi = (size_t)-1, pl = (size_t)-1;
        for (size_t j = 0; j < nv; ++j) {
            const double a = A[e * nv + j];
            if (a == -1.0) mi = j;
            else if (a == 1.0) pl = j;
            else if (a != 0.0) goto done;
        }
        if (mi == (size_t)-1 || pl == (size_t)-1) goto done;
        for (size_t a = 0; a < nv; ++a) as[e * nv + a] = (P[pl * nv + a] - P[mi * nv + a]) / P[a * nv + a];
    }
    if (st2) {
        if (shunt_a < 0 || (size_t)shunt_a >= nv || (size_t)shunt_b >= nv || shunt_a == shunt_b) goto done;
        b[shunt_a] = -1.0;
        b[shunt_b] = +1.0;
        for (size_t a = 0; a < nv; ++a)
            cc[a] = (P[(size_t)shunt_b * nv + a] - P[(size_t)shunt_a * nv + a]) / P[a * nv + a];
    }
    memset(M, 0, sizeof(double) * nt * nt);
    for (size_t i = 0; i < ne; ++i)
        for (size_t j = 0; j < ne; ++j) {
            double v = L[i * ne + j] / dt;
            if (i == j) v += theta * R[i];
            M[i * nt + j] = v;
        }
    for (size_t e = 0; e < ne; ++e)
        for (size_t j = 0; j < nv; ++j) M[e * nt + ne + j] = theta * as[e * nv + j];
    for (size_t j = 0; j < nv; ++j)
        for (size_t e = 0; e < ne; ++e) M[(ne + j) * nt + e] = -theta * A[e * nv + j];
    for (size_t r = 0; r < nv; ++r)
        for (size_t col = 0; col < nv; ++col) {
            double v = 0.0;
            if (r == col) v += f[r];
            M[(ne + r) * nt + ne + col] = v / dt;
        }
    if (st2) {
        const size_t s = ne + nv;
        for (size_t j = 0; j < nv; ++j) M[(ne + j) * nt + s] = -theta * b[j];
        for (size_t j = 0; j < nv; ++j) M[s * nt + ne + j] = cc[j];
        M[s * nt + s] = r_shunt;
    }
    if (F) memcpy(F, f, sizeof(double) * nv);
    if (AS) memcpy(AS, as, sizeof(double) * ne * nv);
    if (bR) memcpy(bR, b, sizeof(double) * nv);
    if (c) memcpy(c, cc, sizeof(double) * nv);
    rc = 0;
done:
    free(f);
    free(as);
    free(b);
    free(cc);
    return rc;
}

int oracle_transient_rhs(const double *L, const double *R, const double *A, const double *F, const double *AS,
                         size_t ne, size_t nv, int scheme, double dt, const double *I, const double *V,
                         const double *sn, const double *snp1, const double *un, const double *unp1,
                         const double *bR, double ishunt, double *rhs)
{
    if (!L || !R || !A || !F || !AS || !I || !V || !rhs || (scheme != 1 && scheme != 2)) return -1;
    const double theta = oracle_theta(scheme), omt = 1.0 - theta;
    for (size_t i = 0; i < ne; ++i) {
        double v = 0.0;
        for (size_t j = 0; j < ne; ++j) v += L[i * ne + j] * I[j] / dt;
        v -= omt * R[i] * I[i];
        for (size_t j = 0; j < nv; ++j) v -= omt * AS[i * nv + j] * V[j];
        if (un) v -= omt * un[i];
        if (unp1) v -= theta * unp1[i];
        rhs[i] = v;
    }
    for (size_t j = 0; j < nv; ++j) {
        double v = F[j] * V[j] / dt;
        for (size_t e = 0; e < ne; ++e) v += omt * A[e * nv + j] * I[e];
        if (sn) v += omt * sn[j];
        if (snp1) v += theta * snp1[j];
        rhs[ne + j] = v;
    }
    if (bR) {
        for (size_t j = 0; j < nv; ++j) rhs[ne + j] += omt * bR[j] * ishunt;
        rhs[ne + nv] = 0.0;
    }
    return 0;
}

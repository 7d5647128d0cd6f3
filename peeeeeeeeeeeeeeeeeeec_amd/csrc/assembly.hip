// gfx950 assembly of the partial-inductance (Lp) and potential-coefficient
// (P) matrices: the reference's analytic 1/R panel integrals
// (quadrature.c:632-1079, inductance.c:221-560, potential.c:139-395) and the
// retarded exp(-jkR)/R kernel (new; the reference's retardation.c is empty).
//
// Three kernels:
//  * far_tile_kernel   one wavefront = 64 observer regions x 1 source region;
//                      the source region's Gauss points are staged in LDS and
//                      read as broadcasts, each lane keeps 4 observer points in
//                      registers; 1/R via v_rsq_f64 + one Newton step. Every
//                      entry is computed as if all its piece pairs were FAR.
//  * near_entry_kernel one wavefront per candidate near entry (host-built
//                      superset of entries with a SELF/TOUCHING/NEAR piece
//                      pair); the wavefront takes the entry's (piece pair,
//                      triangle pair) tasks one after another and runs the
//                      reference rule of each task's class with all 64 lanes:
//                      the 4-D GL product rule, the closed-form identical-
//                      triangle term, or the analytic-inner / adaptive-outer
//                      rule, whose recursion is walked one node at a time
//                      (the node's quadrature points spread over the lanes,
//                      pending parents on a stack in LDS, sums formed in
//                      post-order like the recursion). The task values are
//                      summed in the reference's order and overwrite the entry.
//  * smooth_tile_kernel same tiling; integrates (exp(-jkR)-1)/R with the
//                      smooth-order rule and adds it to the static matrix.
// This file is compiled with -ffp-contract=off so the classification and the
// near-field arithmetic round like the reference's ISO C build; the tile
// kernels use explicit fma where they want it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <unordered_map>

#include "../../include/quadrature.h"
#include "assembly.h"

namespace peec {

__constant__ double c_gl_x[17][16];
__constant__ double c_gl_w[17][16];

void ensure_device_tables() {
    static std::mutex mu;
    static bool done = false;
    std::lock_guard<std::mutex> lock(mu);
    if (done) return;
    const GLTable &t = gl_table();
    HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_gl_x), t.x, sizeof t.x));
    HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_gl_w), t.w, sizeof t.w));
    done = true;
}

constexpr double kInv4PiEps0 = 8.987551792261171e9;  // potential.c:30
constexpr double kMu0Over4Pi = 1.0e-7;               // inductance.c:24
constexpr int kTileI = 64;                           // observer regions per tile (one per lane)
constexpr int kTileJ = 16;                           // source regions per tile (4 per wave)
constexpr int kMaxTasks = 1024;                      // piece-pair x triangle-pair tasks per entry
constexpr int kMaxDepth = 24;                        // quadrature.c:921

// ---------------------------------------------------------------------------
// device geometry helpers (same evaluation order as the reference)

__device__ inline double d_norm3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
__device__ inline double d_dist3(const double *a, const double *b) {
    double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    return d_norm3(d);
}
__device__ inline void d_cross3(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline double d_dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ inline bool d_coincide(const double *a, const double *b) { return d_dist3(a, b) <= 1.0e-12; }

// collapsed-square map xi = u, eta = (1-u) v, J = |e1 x e2| (1-u)
__device__ inline void d_tri_map(const double (&t)[3][3], double u, double v, double *r, double &J) {
    double e1[3] = {t[1][0] - t[0][0], t[1][1] - t[0][1], t[1][2] - t[0][2]};
    double e2[3] = {t[2][0] - t[0][0], t[2][1] - t[0][1], t[2][2] - t[0][2]};
    double eta = (1.0 - u) * v;
    for (int k = 0; k < 3; ++k) r[k] = t[0][k] + u * e1[k] + eta * e2[k];
    double c[3];
    d_cross3(e1, e2, c);
    J = d_norm3(c) * (1.0 - u);
}

// Wave-wide sum as a butterfly over the 64 lanes: at every step a lane and
// its partner add the same two values in swapped order, and IEEE addition is
// commutative, so all lanes end with the same bits. wave_uniform() then
// takes the first lane's copy as a scalar, so the branches on it are uniform.
__device__ inline double wave_sum(double x) {
    for (int m = 32; m >= 1; m >>= 1) x += __shfl_xor(x, m);
    return x;
}
__device__ inline double wave_uniform(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// 4-D Gauss-Legendre product rule over a triangle pair (quadrature.c:679-783)
// with the order^4 points spread over the wavefront: for each point i of the
// first coordinate, lane m takes the (j, k, l) combinations m, m + 64, ...
// Each term is formed as in the reference; only the order of the sum differs.
__device__ double wave_tri_pair_gl(const double (&a)[3][3], const double (&b)[3][3], int order, int lane) {
    const int o2 = order * order, o3 = o2 * order;
    double sum = 0.0;
#pragma unroll 1
    for (int i = 0; i < order; ++i) {
        const double ua = 0.5 * (c_gl_x[order][i] + 1.0), wu = 0.5 * c_gl_w[order][i];
#pragma unroll 1
        for (int m = lane; m < o3; m += 64) {
            const int j = m / o2, kl = m - j * o2, k = kl / order, l = kl - k * order;
            const double va = 0.5 * (c_gl_x[order][j] + 1.0), wv = 0.5 * c_gl_w[order][j];
            const double ub = 0.5 * (c_gl_x[order][k] + 1.0), wub = 0.5 * c_gl_w[order][k];
            const double vb = 0.5 * (c_gl_x[order][l] + 1.0), wvb = 0.5 * c_gl_w[order][l];
            double ra[3], ja, rb[3], jb;
            d_tri_map(a, ua, va, ra, ja);
            d_tri_map(b, ub, vb, rb, jb);
            const double R = d_dist3(ra, rb);
            if (R > 0.0) sum += wu * wv * wub * wvb * ja * jb / R;
        }
    }
    return wave_uniform(wave_sum(sum));
}

// exact potential of a uniform triangle at a point (quadrature.c:842-898)
struct TriPot {
    double p[3][3], t[3][3], m[3][3], l[3], n[3];
    double area2;
};

__device__ void d_tripot_setup(const double (&p)[3][3], TriPot &T) {
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            T.p[i][k] = p[i][k];
            T.t[i][k] = p[(i + 1) % 3][k] - p[i][k];
        }
    d_cross3(T.t[0], T.t[1], T.n);
    T.area2 = d_norm3(T.n);
    if (T.area2 == 0.0) return;
    for (int k = 0; k < 3; ++k) T.n[k] /= T.area2;
    for (int i = 0; i < 3; ++i) {
        T.l[i] = d_norm3(T.t[i]);
        for (int k = 0; k < 3; ++k) T.t[i][k] /= T.l[i];
        d_cross3(T.t[i], T.n, T.m[i]);
    }
}

// the set-up values as wave-uniform scalars (every lane computed the same
// bits from the same triangle), so they can live in scalar registers
__device__ inline void d_tripot_uniform(TriPot &T) {
    double *f = &T.p[0][0];
    for (int k = 0; k < (int)(sizeof(TriPot) / sizeof(double)); ++k) f[k] = wave_uniform(f[k]);
}

__device__ double d_tripot_eval(const TriPot &T, const double *r) {
    double v[3][3], rad[3];
    for (int i = 0; i < 3; ++i) {
        for (int k = 0; k < 3; ++k) v[i][k] = T.p[i][k] - r[k];
        rad[i] = d_norm3(v[i]);
    }
    const double h = d_dot3(v[0], T.n);
    double acc = 0.0;
    for (int i = 0; i < 3; ++i) {
        const double d = d_dot3(v[i], T.m[i]);
        if (d == 0.0) continue;
        const double s = d_dot3(v[i], T.t[i]);
        const double rho = hypot(d, h);
        acc += d * (asinh((s + T.l[i]) / rho) - asinh(s / rho));
    }
    if (h != 0.0) {
        double c[3];
        d_cross3(v[1], v[2], c);
        const double den = rad[0] * rad[1] * rad[2] + d_dot3(v[0], v[1]) * rad[2] +
                           d_dot3(v[1], v[2]) * rad[0] + d_dot3(v[2], v[0]) * rad[1];
        acc -= fabs(h * (2.0 * atan2(d_dot3(v[0], c), den)));
    }
    return acc;
}

// Low and high outer rules of the analytic inner potential T over triangle t
// (quadrature.c:900-913) in one pass: the low^2 + order^2 points are spread
// over the wavefront, each lane summing its points of either rule. Each term
// is formed as in the reference; only the order of the sums differs.
__device__ void wave_outer_rules(const double (&t)[3][3], const TriPot &T, int low, int order, int lane,
                                 double &lo, double &hi) {
    const int nl = low * low, np = nl + order * order;
    double sl = 0.0, sh = 0.0;
#pragma unroll 1
    for (int p = lane; p < np; p += 64) {
        const bool is_low = p < nl;
        const int q = is_low ? low : order, idx = is_low ? p : p - nl;
        const int i = idx / q, j = idx - i * q;
        double r[3], J;
        d_tri_map(t, 0.5 * (c_gl_x[q][i] + 1), 0.5 * (c_gl_x[q][j] + 1), r, J);
        const double v = 0.25 * c_gl_w[q][i] * c_gl_w[q][j] * J * d_tripot_eval(T, r);
        if (is_low)
            sl += v;
        else
            sh += v;
    }
    lo = wave_uniform(wave_sum(sl));
    hi = wave_uniform(wave_sum(sh));
}

// longest-edge bisection (quadrature.c:923-937); child 0 = left, 1 = right
__device__ inline void d_split(const double (&a)[3][3], int child, double (&out)[3][3]) {
    int e = 0;
    double best = 0.0;
    for (int i = 0; i < 3; ++i) {
        const double l = d_dist3(a[i], a[(i + 1) % 3]);
        if (l > best) {
            best = l;
            e = i;
        }
    }
    const int j = (e + 1) % 3, k = (e + 2) % 3;
    for (int d = 0; d < 3; ++d) {
        const double mid = 0.5 * (a[e][d] + a[j][d]);
        if (child == 0) {
            out[0][d] = a[e][d];
            out[1][d] = mid;
        } else {
            out[0][d] = mid;
            out[1][d] = a[j][d];
        }
        out[2][d] = a[k][d];
    }
}

// A parent triangle of the adaptive recursion whose children are pending.
struct WaveFrame {
    double t[3][3];
    double tol;   // the parent's tolerance (each child gets half)
    double left;  // value of the left child once it is done
    int depth;
    int state;    // 1: left child in progress, 2: right child in progress
};

// Adaptive outer integration (quadrature.c:915-938) by the whole wavefront:
// one node of the recursion at a time, its two rules evaluated by all lanes,
// the accept/bisect decision uniform. The parents of the current node wait
// on a stack in LDS (every lane writes the same values); a finished node is
// combined with them in post-order, so each sum is left + right as in the
// recursive original. n counts the analytic potential evaluations.
__device__ double wave_outer_adaptive(const double (&a)[3][3], const TriPot &T, int order, double tol, int lane,
                                      WaveFrame *st, unsigned &n) {
    const int low = order <= 6 ? 3 : 4;
    double t[3][3];
    for (int r = 0; r < 3; ++r)
        for (int d = 0; d < 3; ++d) t[r][d] = a[r][d];
    double node_tol = tol;
    int depth = 0, sp = 0;
    for (;;) {
        n += low * low + order * order;
        double lo, hi;
        wave_outer_rules(t, T, low, order, lane, lo, hi);
        if (!isfinite(lo) || !isfinite(hi)) return NAN;
        if (fabs(hi - lo) <= node_tol + 2e-5 * fabs(hi)) {
            double ret = hi;
            for (;;) {
                if (sp == 0) return ret;
                WaveFrame &f = st[sp - 1];
                if (__builtin_amdgcn_readfirstlane(f.state) == 1) {
                    // left child done: keep it, evaluate the right child
                    f.left = ret;
                    f.state = 2;
                    d_split(f.t, 1, t);
                    node_tol = f.tol * 0.5;
                    depth = __builtin_amdgcn_readfirstlane(f.depth) + 1;
                    break;
                }
                ret = f.left + ret;  // right child done
                --sp;
            }
        } else {
            if (depth >= kMaxDepth) return NAN;
            WaveFrame &f = st[sp++];
            for (int r = 0; r < 3; ++r)
                for (int d = 0; d < 3; ++d) f.t[r][d] = t[r][d];
            f.tol = node_tol;
            f.depth = depth;
            f.state = 1;
            double c[3][3];
            d_split(t, 0, c);
            for (int r = 0; r < 3; ++r)
                for (int d = 0; d < 3; ++d) t[r][d] = c[r][d];
            node_tol *= 0.5;
            ++depth;
        }
    }
}

// twice the area of a triangle, formed as in the potential set-up
__device__ inline double d_twice_area(const double (&p)[3][3]) {
    double t0[3], t1[3], c[3];
    for (int k = 0; k < 3; ++k) {
        t0[k] = p[1][k] - p[0][k];
        t1[k] = p[2][k] - p[1][k];
    }
    d_cross3(t0, t1, c);
    return d_norm3(c);
}

// near-field triangle pair (quadrature.c:940-973) by the whole wavefront;
// cls 0 self, 1 touching, 2 near
__device__ double wave_tri_pair_near(const double (&a)[3][3], const double (&b)[3][3], int cls,
                                     const QuadratureOrders &q, int lane, WaveFrame *st, unsigned &n) {
    const double area_a = d_twice_area(a), area_b = d_twice_area(b);
    if (area_a == 0.0 || area_b == 0.0) return 0.0;
    bool same = true;
    for (int i = 0; i < 3; ++i) {
        bool hit = false;
        for (int j = 0; j < 3; ++j)
            if (d_coincide(a[i], b[j])) hit = true;
        if (!hit) same = false;
    }
    TriPot T;
    if (same) {
        // identical triangles: (2A/3) sum over the vertices, one per lane,
        // added in vertex order
        d_tripot_setup(b, T);
        double r[3];
        for (int d = 0; d < 3; ++d) r[d] = lane == 0 ? a[0][d] : lane == 1 ? a[1][d] : a[2][d];
        const double v = lane < 3 ? d_tripot_eval(T, r) : 0.0;
        double s = 0.0;
        for (int i = 0; i < 3; ++i) s += wave_uniform(__shfl(v, i));
        n += 3;
        return area_a * s / 3.0;
    }
    int order = cls == 0 ? q.self_order : cls == 1 ? q.touching_order : q.near_order;
    if (order < 6) order = 6;
    const double scale = area_a * area_b / sqrt(fmax(area_a, area_b));
    // a over the potential of b, then b over the potential of a (one call
    // site, so the recursion's code is emitted once)
    double half[2];
#pragma unroll 1
    for (int dn = 0; dn < 2; ++dn) {
        double outer[3][3], inner[3][3];
        for (int r = 0; r < 3; ++r)
            for (int d = 0; d < 3; ++d) {
                outer[r][d] = dn ? b[r][d] : a[r][d];
                inner[r][d] = dn ? a[r][d] : b[r][d];
            }
        d_tripot_setup(inner, T);
        d_tripot_uniform(T);
        const double v = wave_outer_adaptive(outer, T, order, 2e-5 * scale, lane, st, n);
        if (dn)
            half[1] = v;
        else
            half[0] = v;
    }
    return 0.5 * (half[0] + half[1]);
}

// piece-pair classification (quadrature.c:550-607); `same_object` is the
// reference's pointer-equality shortcut.
__device__ int d_classify(const double *a, const double *b, bool same_object, double nf) {
    if (same_object) return 0;
    bool used[4] = {false, false, false, false}, all = true;
    for (int i = 0; i < 4 && all; ++i) {
        bool hit = false;
        for (int j = 0; j < 4 && !hit; ++j)
            if (!used[j] && d_coincide(&a[3 * i], &b[3 * j])) used[j] = hit = true;
        all = hit;
    }
    if (all) return 0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            if (d_coincide(&a[3 * i], &b[3 * j])) return 1;
    double g2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        double alo = a[k], ahi = alo, blo = b[k], bhi = blo;
        for (int i = 1; i < 4; ++i) {
            alo = fmin(alo, a[3 * i + k]);
            ahi = fmax(ahi, a[3 * i + k]);
            blo = fmin(blo, b[3 * i + k]);
            bhi = fmax(bhi, b[3 * i + k]);
        }
        const double gap = fmax(0.0, fmax(alo - bhi, blo - ahi));
        g2 += gap * gap;
    }
    double ha = 0.0, hb = 0.0;
    for (int i = 0; i < 4; ++i) {
        const double la = d_dist3(&a[3 * i], &a[3 * ((i + 1) % 4)]);
        const double lb = d_dist3(&b[3 * i], &b[3 * ((i + 1) % 4)]);
        if (la > ha) ha = la;
        if (lb > hb) hb = lb;
    }
    const double h = ha > hb ? ha : hb;
    return (h > 0.0 && sqrt(g2) < nf * h) ? 2 : 3;
}

__device__ inline void d_piece_triangle(const double *pv, int which, double (&t)[3][3]) {
    const int v1 = which == 0 ? 1 : 2, v2 = which == 0 ? 2 : 3;
    for (int d = 0; d < 3; ++d) {
        t[0][d] = pv[d];
        t[1][d] = pv[3 * v1 + d];
        t[2][d] = pv[3 * v2 + d];
    }
}

// One wavefront per candidate near entry (i, j); see the file header. The
// wavefront walks the entry's (piece pair, triangle pair) tasks in the
// reference's order and evaluates each one with all lanes, so no lane waits
// on another's adaptive recursion; the task values are summed in that order.
__global__ __launch_bounds__(64) void near_entry_kernel(
    const int *__restrict__ pairs, int n_pairs, const int *__restrict__ ptr,
    const double *__restrict__ pv, QuadratureOrders q, int kind, const double *__restrict__ aux,
    const double *__restrict__ dir, double *__restrict__ out, size_t ld, int mirror,
    int *__restrict__ err, unsigned long long *__restrict__ counts = nullptr) {
    __shared__ WaveFrame st[kMaxDepth + 1];
    const int lane = threadIdx.x;
    const int e = blockIdx.x;
    if (e >= n_pairs) return;
    const int i = pairs[2 * e], j = pairs[2 * e + 1];
    const int a0 = ptr[i], na = ptr[i + 1] - a0, b0 = ptr[j], nb = ptr[j + 1] - b0;
    const unsigned far_points = (unsigned)(q.far_order * q.far_order * q.far_order * q.far_order);
    unsigned n_pot = 0, n_gl = 0;  // analytic potential evaluations, 1/R evaluations
    double integral = 0.0;
    bool finite = true;
    for (int pp = 0; pp < na * nb; ++pp) {
        const int ia = pp / nb, ib = pp - ia * nb;
        const double *pa = pv + 12 * (size_t)(a0 + ia);
        const double *pb = pv + 12 * (size_t)(b0 + ib);
        const int cls = d_classify(pa, pb, i == j && ia == ib, q.near_factor);
        double s = 0.0;
#pragma unroll 1
        for (int tp = 0; tp < 4; ++tp) {
            double ta[3][3], tb[3][3];
            d_piece_triangle(pa, tp >> 1, ta);
            d_piece_triangle(pb, tp & 1, tb);
            double x;
            if (cls == 3) {
                x = wave_tri_pair_gl(ta, tb, q.far_order, lane);
                n_gl += far_points;
            } else {
                x = wave_tri_pair_near(ta, tb, cls, q, lane, st, n_pot);
            }
            if (!isfinite(x)) finite = false;
            s += x;
        }
        integral += s;
    }
    if (lane != 0) return;
    if (counts) {
        atomicAdd(&counts[0], (unsigned long long)n_pot);
        atomicAdd(&counts[1], (unsigned long long)n_gl);
    }
    double value;
    if (kind == 2) {
        value = integral;  // raw piece-pair integral (quad_patch_kernel_1_over_r)
    } else if (kind == (int)RegionKind::Node) {
        value = kInv4PiEps0 * integral / (aux[i] * aux[j]);
    } else {
        const double dot = d_dot3(dir + 3 * i, dir + 3 * j);
        value = kMu0Over4Pi * dot * integral / (aux[i] * aux[j]);
    }
    if (!finite || !isfinite(value)) {
        atomicOr(err, 1);
        value = NAN;
    }
    if (ld == 0) {
        out[e] = value;  // compact mode: one value per listed entry
        return;
    }
    out[(size_t)i * ld + j] = value;
    if (mirror) out[(size_t)j * ld + i] = value;
}

// ---------------------------------------------------------------------------
// Tile kernels. Lane = observer region i (kernel order), wave = 4 source
// regions j in turn; the j region's points sit in LDS (broadcast reads).

__device__ inline double rsqrt_nr(double r2) {
    // v_rsq_f64 is accurate to about 2^-23; one Newton step gives ~1e-14,
    // well inside the 1e-12 entry tolerance of the static matrices.
    double y = __builtin_amdgcn_rsq(r2);
    const double t = r2 * y;
    const double e = __builtin_fma(-t, y, 1.0);
    return __builtin_fma(0.5 * y, e, y);
}

template <int KIND>
__global__ __launch_bounds__(256) void far_tile_kernel(
    const double4 *__restrict__ pts_i, const double4 *__restrict__ pts_j,
    const int *__restrict__ npts, int R, int slots, const double *__restrict__ aux,
    const double *__restrict__ dir, double dir_tol, const int *__restrict__ perm,
    const int *__restrict__ tiles, double *__restrict__ out, size_t ld) {
    extern __shared__ double4 sj[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int bi = tiles[2 * blockIdx.x], bj = tiles[2 * blockIdx.x + 1];
    const int i = bi * kTileI + lane;
    const bool iv = i < R;
    const double ai = iv ? aux[i] : 1.0;
    double di[3] = {0.0, 0.0, 0.0};
    if (KIND == 1 && iv)
        for (int k = 0; k < 3; ++k) di[k] = dir[3 * i + k];
    double4 *myj = sj + (size_t)wave * slots;
    for (int jj = 0; jj < 4; ++jj) {
        const int j = bj * kTileJ + wave * 4 + jj;
        const bool jv = j < R;
        const int nj = jv ? npts[j] : 0;
        for (int q = lane; q < nj; q += 64) myj[q] = pts_j[(size_t)j * slots + q];
        __syncthreads();
        bool active = jv && iv && i <= j;
        double dot = 1.0;
        if (KIND == 1 && active) {
            dot = di[0] * dir[3 * j] + di[1] * dir[3 * j + 1] + di[2] * dir[3 * j + 2];
            active = !(fabs(dot) < dir_tol);
        }
        double acc = 0.0;
        if (__ballot(active)) {
            for (int p = 0; p < slots; p += 4) {
                double4 P0, P1, P2, P3;
                if (iv) {
                    P0 = pts_i[(size_t)(p + 0) * R + i];
                    P1 = pts_i[(size_t)(p + 1) * R + i];
                    P2 = pts_i[(size_t)(p + 2) * R + i];
                    P3 = pts_i[(size_t)(p + 3) * R + i];
                } else {
                    P0 = P1 = P2 = P3 = make_double4(0.0, 0.0, 0.0, 0.0);
                }
                double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
                for (int q = 0; q < nj; ++q) {
                    const double4 Q = myj[q];
                    double dx, dy, dz, r2;
                    dx = P0.x - Q.x; dy = P0.y - Q.y; dz = P0.z - Q.z;
                    r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
                    s0 = __builtin_fma(Q.w, rsqrt_nr(r2), s0);
                    dx = P1.x - Q.x; dy = P1.y - Q.y; dz = P1.z - Q.z;
                    r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
                    s1 = __builtin_fma(Q.w, rsqrt_nr(r2), s1);
                    dx = P2.x - Q.x; dy = P2.y - Q.y; dz = P2.z - Q.z;
                    r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
                    s2 = __builtin_fma(Q.w, rsqrt_nr(r2), s2);
                    dx = P3.x - Q.x; dy = P3.y - Q.y; dz = P3.z - Q.z;
                    r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
                    s3 = __builtin_fma(Q.w, rsqrt_nr(r2), s3);
                }
                acc = __builtin_fma(P0.w, s0, acc);
                acc = __builtin_fma(P1.w, s1, acc);
                acc = __builtin_fma(P2.w, s2, acc);
                acc = __builtin_fma(P3.w, s3, acc);
            }
        }
        if (jv && iv && i <= j) {
            double v = 0.0;
            if (active) {
                v = KIND == 0 ? kInv4PiEps0 * acc / (ai * aux[j])
                              : kMu0Over4Pi * dot * acc / (ai * aux[j]);
            }
            const size_t oi = (size_t)perm[i], oj = (size_t)perm[j];
            out[oi * ld + oj] = v;
            out[oj * ld + oi] = v;
        }
        __syncthreads();
    }
}

// sin^2 x and sin x cos x for a moderate argument (|x| < 2^20, larger ones
// go to the library sin/cos): Cody-Waite reduction by pi/2 split in three
// parts (33+33+53 bits) to r in [-pi/4, pi/4], then the Taylor series of
// sin r to degree 15 and cos r to degree 16 (truncation below 5e-17 and
// 3e-18 of the value). With n the quadrant, sin^2 x is sin^2 r or cos^2 r
// and sin x cos x is +-sin r cos r by the parity of n.
__device__ inline void d_half_angle_terms(double x, double &s2, double &sc) {
    if (!(fabs(x) < 0x1p20)) {  // outside the reduction's exact range
        const double s = sin(x), c = cos(x);
        s2 = s * s;
        sc = s * c;
        return;
    }
    const double n = rint(x * 0x1.45f306dc9c883p-1);
    double r = __builtin_fma(-n, 0x1.921fb54400000p+0, x);
    r = __builtin_fma(-n, 0x1.0b4611a600000p-34, r);
    r = __builtin_fma(-n, 0x1.3198a2e037073p-69, r);
    const double r2 = r * r;
    // sin r = r (1 - r2/3! + r2^2/5! - ...)
    double ps = -1.0 / 1307674368000.0;                   // -1/15!
    ps = __builtin_fma(ps, r2, 1.0 / 6227020800.0);       // 1/13!
    ps = __builtin_fma(ps, r2, -1.0 / 39916800.0);        // -1/11!
    ps = __builtin_fma(ps, r2, 1.0 / 362880.0);           // 1/9!
    ps = __builtin_fma(ps, r2, -1.0 / 5040.0);            // -1/7!
    ps = __builtin_fma(ps, r2, 1.0 / 120.0);              // 1/5!
    ps = __builtin_fma(ps, r2, -1.0 / 6.0);               // -1/3!
    const double sr = __builtin_fma(ps * r2, r, r);
    double pc = 1.0 / 20922789888000.0;                   // 1/16!
    pc = __builtin_fma(pc, r2, -1.0 / 87178291200.0);     // -1/14!
    pc = __builtin_fma(pc, r2, 1.0 / 479001600.0);        // 1/12!
    pc = __builtin_fma(pc, r2, -1.0 / 3628800.0);         // -1/10!
    pc = __builtin_fma(pc, r2, 1.0 / 40320.0);            // 1/8!
    pc = __builtin_fma(pc, r2, -1.0 / 720.0);             // -1/6!
    pc = __builtin_fma(pc, r2, 1.0 / 24.0);               // 1/4!
    pc = __builtin_fma(pc, r2, -0.5);                      // -1/2!
    const double cr = __builtin_fma(pc, r2, 1.0);
    const bool odd = (((long long)n) & 1) != 0;
    const double p = sr * cr;
    s2 = odd ? cr * cr : sr * sr;
    sc = odd ? -p : p;
}

// g(R) = (exp(-jkR) - 1)/R accumulated as  re: -2 sin^2(kR/2)/R,
// im: -sin(kR)/R = -2 sin(kR/2)cos(kR/2)/R;  g(0) = -jk.
template <int KIND>
__global__ __launch_bounds__(256) void smooth_tile_kernel(
    const double4 *__restrict__ pts_i, const double4 *__restrict__ pts_j,
    const int *__restrict__ npts, int R, int slots, const double *__restrict__ aux,
    const double *__restrict__ dir, double dir_tol, const int *__restrict__ perm,
    const int *__restrict__ tiles, double k, const double *__restrict__ stat,
    double2 *__restrict__ out, size_t ld) {
    extern __shared__ double4 sj[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int bi = tiles[2 * blockIdx.x], bj = tiles[2 * blockIdx.x + 1];
    const int i = bi * kTileI + lane;
    const bool iv = i < R;
    const double ai = iv ? aux[i] : 1.0;
    const double half_k = 0.5 * k;
    double di[3] = {0.0, 0.0, 0.0};
    if (KIND == 1 && iv)
        for (int m = 0; m < 3; ++m) di[m] = dir[3 * i + m];
    double4 *myj = sj + (size_t)wave * slots;
    for (int jj = 0; jj < 4; ++jj) {
        const int j = bj * kTileJ + wave * 4 + jj;
        const bool jv = j < R;
        const int nj = jv ? npts[j] : 0;
        for (int q = lane; q < nj; q += 64) myj[q] = pts_j[(size_t)j * slots + q];
        __syncthreads();
        bool active = jv && iv && i <= j;
        double dot = 1.0;
        if (KIND == 1 && active) {
            dot = di[0] * dir[3 * j] + di[1] * dir[3 * j + 1] + di[2] * dir[3 * j + 2];
            active = !(fabs(dot) < dir_tol);
        }
        double acc_re = 0.0, acc_im = 0.0;
        if (__ballot(active)) {
            for (int p = 0; p < slots; ++p) {
                const double4 P = iv ? pts_i[(size_t)p * R + i] : make_double4(0.0, 0.0, 0.0, 0.0);
                double sre = 0.0, sim = 0.0;
                for (int q = 0; q < nj; ++q) {
                    const double4 Q = myj[q];
                    const double dx = P.x - Q.x, dy = P.y - Q.y, dz = P.z - Q.z;
                    const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
                    const bool zero = r2 == 0.0;
                    const double rinv = zero ? 0.0 : rsqrt_nr(r2);
                    const double rr = r2 * rinv;
                    double s2, sc;
                    d_half_angle_terms(half_k * rr, s2, sc);
                    const double wr = Q.w * rinv;
                    sre = __builtin_fma(wr, s2, sre);
                    sim = zero ? __builtin_fma(Q.w, half_k, sim) : __builtin_fma(wr, sc, sim);
                }
                acc_re = __builtin_fma(P.w, sre, acc_re);
                acc_im = __builtin_fma(P.w, sim, acc_im);
            }
        }
        if (jv && iv && i <= j) {
            const size_t oi = (size_t)perm[i], oj = (size_t)perm[j];
            const double base = stat[oi * ld + oj];
            double2 v = make_double2(base, 0.0);
            if (active) {
                const double sc = KIND == 0 ? kInv4PiEps0 / (ai * aux[j])
                                            : kMu0Over4Pi * dot / (ai * aux[j]);
                v.x = base + sc * (-2.0 * acc_re);
                v.y = sc * (-2.0 * acc_im);
            }
            out[oi * ld + oj] = v;
            out[oj * ld + oi] = v;
        }
        __syncthreads();
    }
}

// Remainder of near entries with the higher near_order rule (one wavefront
// per near entry, one lane per (piece pair, triangle pair)); overwrites the
// tile kernel's value: out = static + scale * remainder.
__global__ __launch_bounds__(64) void near_smooth_kernel(
    const int *__restrict__ pairs, int n_pairs, const int *__restrict__ ptr,
    const double *__restrict__ pv, int order, double k, int kind, const double *__restrict__ aux,
    const double *__restrict__ dir, double dir_tol, const double *__restrict__ stat,
    double2 *__restrict__ out, size_t ld, int *__restrict__ err) {
    __shared__ double vre[kMaxTasks], vim[kMaxTasks];
    const int e = blockIdx.x;
    if (e >= n_pairs) return;
    const int lane = threadIdx.x;
    const int i = pairs[2 * e], j = pairs[2 * e + 1];
    const int a0 = ptr[i], na = ptr[i + 1] - a0, b0 = ptr[j], nb = ptr[j + 1] - b0;
    const int tasks = na * nb * 4;
    if (tasks > kMaxTasks) {
        if (lane == 0) atomicOr(err, 2);
        return;
    }
    const double half_k = 0.5 * k;
    for (int t = lane; t < tasks; t += 64) {
        const int pp = t >> 2, tp = t & 3;
        const int ia = pp / nb, ib = pp - ia * nb;
        double ta[3][3], tb[3][3];
        d_piece_triangle(pv + 12 * (size_t)(a0 + ia), tp >> 1, ta);
        d_piece_triangle(pv + 12 * (size_t)(b0 + ib), tp & 1, tb);
        double sre = 0.0, sim = 0.0;
        for (int u = 0; u < order; ++u)
            for (int v = 0; v < order; ++v) {
                double ra[3], ja;
                d_tri_map(ta, 0.5 * (c_gl_x[order][u] + 1.0), 0.5 * (c_gl_x[order][v] + 1.0), ra, ja);
                const double wa = 0.25 * c_gl_w[order][u] * c_gl_w[order][v] * ja;
                double pre = 0.0, pim = 0.0;
                for (int x = 0; x < order; ++x)
                    for (int y = 0; y < order; ++y) {
                        double rb[3], jb;
                        d_tri_map(tb, 0.5 * (c_gl_x[order][x] + 1.0), 0.5 * (c_gl_x[order][y] + 1.0), rb, jb);
                        const double wb = 0.25 * c_gl_w[order][x] * c_gl_w[order][y] * jb;
                        const double R = d_dist3(ra, rb);
                        if (R > 0.0) {
                            double s2, sc;
                            d_half_angle_terms(half_k * R, s2, sc);
                            pre += wb * s2 / R;
                            pim += wb * sc / R;
                        } else {
                            pim += wb * half_k;
                        }
                    }
                sre += wa * pre;
                sim += wa * pim;
            }
        vre[t] = sre;
        vim[t] = sim;
    }
    __syncthreads();
    if (lane != 0) return;
    double gre = 0.0, gim = 0.0;
    for (int t = 0; t < tasks; ++t) {
        gre += vre[t];
        gim += vim[t];
    }
    double sc;
    if (kind == (int)RegionKind::Node) {
        sc = kInv4PiEps0 / (aux[i] * aux[j]);
    } else {
        const double dot = d_dot3(dir + 3 * i, dir + 3 * j);
        if (fabs(dot) < dir_tol) return;
        sc = kMu0Over4Pi * dot / (aux[i] * aux[j]);
    }
    const double base = stat[(size_t)i * ld + j];
    const double2 val = make_double2(base + sc * (-2.0 * gre), sc * (-2.0 * gim));
    if (!isfinite(val.x) || !isfinite(val.y)) atomicOr(err, 1);
    out[(size_t)i * ld + j] = val;
    out[(size_t)j * ld + i] = val;
}

// ---------------------------------------------------------------------------
// moments of the smooth remainder (frequency sweeps)
//
// U_m(i, j) = scale(i, j) * sum_w w R^(m-1), m = 1..terms, stored as
// [m][upper tile][64 x 64] in original numbering (tile (ta, tb), ta <= tb, of
// entries a = 64 ta + r, b = 64 tb + c); a diagonal tile holds both halves.

constexpr int kMomentTerms = RegionSet::kMaxMomentTerms;

__device__ inline size_t moment_slot(int a, int b, int nt) {
    const int ta = a >> 6, tb = b >> 6;
    const size_t t = (size_t)ta * nt - (size_t)ta * (ta - 1) / 2 + (size_t)(tb - ta);
    return t * 4096 + (size_t)(a & 63) * 64 + (size_t)(b & 63);
}

__device__ inline void store_moments(double *mom, size_t mstride, int nt, int oi, int oj, int terms,
                                     const double (&u)[kMomentTerms], double scale) {
    const int a = min(oi, oj), b = max(oi, oj);
    const size_t p = moment_slot(a, b, nt);
    const bool diag = (a >> 6) == (b >> 6);
    const size_t q = diag ? p - (size_t)(a & 63) * 64 - (size_t)(b & 63) + (size_t)(b & 63) * 64 + (size_t)(a & 63) : p;
#pragma unroll
    for (int m = 0; m < kMomentTerms; ++m)
        if (m < terms) {
            mom[(size_t)m * mstride + p] = scale * u[m];
            if (diag) mom[(size_t)m * mstride + q] = scale * u[m];
        }
}

// Same tiling and quadrature points as smooth_tile_kernel; each lane keeps the
// power sums of its entry in registers.
template <int KIND>
__global__ __launch_bounds__(256) void moment_tile_kernel(
    const double4 *__restrict__ pts_i, const double4 *__restrict__ pts_j,
    const int *__restrict__ npts, int R, int slots, const double *__restrict__ aux,
    const double *__restrict__ dir, double dir_tol, const int *__restrict__ perm,
    const int *__restrict__ tiles, int terms, double *__restrict__ mom, size_t mstride, int nt) {
    extern __shared__ double4 sj[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int bi = tiles[2 * blockIdx.x], bj = tiles[2 * blockIdx.x + 1];
    const int i = bi * kTileI + lane;
    const bool iv = i < R;
    const double ai = iv ? aux[i] : 1.0;
    double di[3] = {0.0, 0.0, 0.0};
    if (KIND == 1 && iv)
        for (int m = 0; m < 3; ++m) di[m] = dir[3 * i + m];
    double4 *myj = sj + (size_t)wave * slots;
    for (int jj = 0; jj < 4; ++jj) {
        const int j = bj * kTileJ + wave * 4 + jj;
        const bool jv = j < R;
        const int nj = jv ? npts[j] : 0;
        for (int q = lane; q < nj; q += 64) myj[q] = pts_j[(size_t)j * slots + q];
        __syncthreads();
        bool active = jv && iv && i <= j;
        double dot = 1.0;
        if (KIND == 1 && active) {
            dot = di[0] * dir[3 * j] + di[1] * dir[3 * j + 1] + di[2] * dir[3 * j + 2];
            active = !(fabs(dot) < dir_tol);
        }
        double u[kMomentTerms];
#pragma unroll
        for (int m = 0; m < kMomentTerms; ++m) u[m] = 0.0;
        if (__ballot(active)) {
            for (int p = 0; p < slots; ++p) {
                const double4 P = iv ? pts_i[(size_t)p * R + i] : make_double4(0.0, 0.0, 0.0, 0.0);
                for (int q = 0; q < nj; ++q) {
                    const double4 Q = myj[q];
                    const double dx = P.x - Q.x, dy = P.y - Q.y, dz = P.z - Q.z;
                    const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
                    const double rr = r2 == 0.0 ? 0.0 : r2 * rsqrt_nr(r2);
                    double w = P.w * Q.w;
#pragma unroll
                    for (int m = 0; m < kMomentTerms; ++m)
                        if (m < terms) {
                            u[m] += w;
                            w *= rr;
                        }
                }
            }
        }
        if (jv && iv && i <= j) {
            const double sc = !active ? 0.0
                              : KIND == 0 ? kInv4PiEps0 / (ai * aux[j])
                                          : kMu0Over4Pi * dot / (ai * aux[j]);
            store_moments(mom, mstride, nt, perm[i], perm[j], terms, u, sc);
        }
        __syncthreads();
    }
}

// Near entries with the near rule (overwrites the tile kernel's moments);
// mirrors near_smooth_kernel.
__global__ __launch_bounds__(64) void near_moment_kernel(
    const int *__restrict__ pairs, int n_pairs, const int *__restrict__ ptr,
    const double *__restrict__ pv, int order, int kind, const double *__restrict__ aux,
    const double *__restrict__ dir, double dir_tol, int terms, double *__restrict__ mom, size_t mstride,
    int nt, int *__restrict__ err) {
    __shared__ double part[kMomentTerms][65];
    const int e = blockIdx.x;
    if (e >= n_pairs) return;
    const int lane = threadIdx.x;
    const int i = pairs[2 * e], j = pairs[2 * e + 1];
    const int a0 = ptr[i], na = ptr[i + 1] - a0, b0 = ptr[j], nb = ptr[j + 1] - b0;
    const int tasks = na * nb * 4;
    if (tasks > kMaxTasks) {
        if (lane == 0) atomicOr(err, 2);
        return;
    }
    double u[kMomentTerms];
#pragma unroll
    for (int m = 0; m < kMomentTerms; ++m) u[m] = 0.0;
    for (int t = lane; t < tasks; t += 64) {
        const int pp = t >> 2, tp = t & 3;
        const int ia = pp / nb, ib = pp - ia * nb;
        double ta[3][3], tb[3][3];
        d_piece_triangle(pv + 12 * (size_t)(a0 + ia), tp >> 1, ta);
        d_piece_triangle(pv + 12 * (size_t)(b0 + ib), tp & 1, tb);
        for (int uu = 0; uu < order; ++uu)
            for (int vv = 0; vv < order; ++vv) {
                double ra[3], ja;
                d_tri_map(ta, 0.5 * (c_gl_x[order][uu] + 1.0), 0.5 * (c_gl_x[order][vv] + 1.0), ra, ja);
                const double wa = 0.25 * c_gl_w[order][uu] * c_gl_w[order][vv] * ja;
                for (int x = 0; x < order; ++x)
                    for (int y = 0; y < order; ++y) {
                        double rb[3], jb;
                        d_tri_map(tb, 0.5 * (c_gl_x[order][x] + 1.0), 0.5 * (c_gl_x[order][y] + 1.0), rb, jb);
                        const double wb = 0.25 * c_gl_w[order][x] * c_gl_w[order][y] * jb;
                        const double R = d_dist3(ra, rb);
                        double w = wa * wb;
#pragma unroll
                        for (int m = 0; m < kMomentTerms; ++m)
                            if (m < terms) {
                                u[m] += w;
                                w *= R;
                            }
                    }
            }
    }
#pragma unroll
    for (int m = 0; m < kMomentTerms; ++m) part[m][lane] = u[m];
    __syncthreads();
    double sc;
    if (kind == (int)RegionKind::Node) {
        sc = kInv4PiEps0 / (aux[i] * aux[j]);
    } else {
        const double dot = d_dot3(dir + 3 * i, dir + 3 * j);
        if (fabs(dot) < dir_tol) return;
        sc = kMu0Over4Pi * dot / (aux[i] * aux[j]);
    }
    if (lane >= terms) return;
    double v = 0.0;
    for (int l = 0; l < 64; ++l) v += part[lane][l];
    v *= sc;
    if (!isfinite(v)) atomicOr(err, 1);
    const int a = min(i, j), b = max(i, j);
    const size_t p = moment_slot(a, b, nt);
    mom[(size_t)lane * mstride + p] = v;
    if ((a >> 6) == (b >> 6)) mom[(size_t)lane * mstride + p - (size_t)(a & 63) * 64 - (b & 63) + (size_t)(b & 63) * 64 + (a & 63)] = v;
}

struct MomentCoefficients {
    double c[kMomentTerms];  // c[m]: coefficient of U_(m+1); odd m real, even m imaginary part
};

// out = static + sum_m c_m U_m over one 64 x 64 upper tile per workgroup:
// row-major writes of the tile, then its transpose through LDS (the real
// and imaginary parts in turn) for the lower triangle.
__global__ __launch_bounds__(256) void moment_eval_kernel(
    const double *__restrict__ mom, size_t mstride, int terms, MomentCoefficients cf,
    const int *__restrict__ tile_ids, int R, const double *__restrict__ stat, double2 *__restrict__ out,
    size_t ld) {
    __shared__ double tr[64][65];
    const int t = blockIdx.x;
    const int ta = tile_ids[2 * t], tb = tile_ids[2 * t + 1];
    const int c = threadIdx.x & 63, r0 = threadIdx.x >> 6;
    const double *base = mom + (size_t)t * 4096 + (size_t)r0 * 64 + c;
    double re[16], im[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) re[q] = im[q] = 0.0;
    // smallest terms first
    for (int m = terms - 1; m >= 0; --m) {
        const double *src = base + (size_t)m * mstride;
        const double a = cf.c[m];
        if (m & 1) {
#pragma unroll
            for (int q = 0; q < 16; ++q) re[q] = __builtin_fma(a, src[q * 256], re[q]);
        } else {
#pragma unroll
            for (int q = 0; q < 16; ++q) im[q] = __builtin_fma(a, src[q * 256], im[q]);
        }
    }
    const int j = tb * 64 + c;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int i = ta * 64 + r0 + 4 * q;
        if (i < R && j < R) {
            const size_t o = (size_t)i * ld + j;
            re[q] += stat[o];
            out[o] = make_double2(re[q], im[q]);
        }
    }
    if (ta == tb) return;  // the diagonal tile holds both halves
    // lower triangle: out[j][i] = value(i, j) for (i, j) in this tile
    double tre[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) tr[r0 + 4 * q][c] = re[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) tre[q] = tr[c][r0 + 4 * q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) tr[r0 + 4 * q][c] = im[q];
    __syncthreads();
    const int i = ta * 64 + c;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int jj = tb * 64 + r0 + 4 * q;
        if (i < R && jj < R) out[(size_t)jj * ld + i] = make_double2(tre[q], tr[c][r0 + 4 * q]);
    }
}

// The same series for up to four wave numbers per pass (a batch of sweep
// points): 1024 threads per 64 x 64 upper tile, each with four rows of the
// tile for every member, so each moment and static value is read once for
// all members. Per value and member the sum runs over the same terms in the
// same order as moment_eval_kernel (a member's coefficients past its own
// term count are zero and leave the sum at +0), so the results are
// identical.
struct MomentBatch {
    MomentCoefficients cf[RegionSet::kMaxMomentBatch];
    double2 *out[RegionSet::kMaxMomentBatch];
};

template <int B>
__global__ __launch_bounds__(1024) void moment_eval_batch_kernel(const double *__restrict__ mom, size_t mstride,
                                                                 int terms, MomentBatch mb,
                                                                 const int *__restrict__ tile_ids, int R,
                                                                 const double *__restrict__ stat, size_t ld) {
    __shared__ double tr[64][65];
    const int t = blockIdx.x;
    const int ta = tile_ids[2 * t], tb = tile_ids[2 * t + 1];
    const int c = threadIdx.x & 63, r0 = threadIdx.x >> 6;  // rows r0 + 16 q
    const double *base = mom + (size_t)t * 4096 + (size_t)r0 * 64 + c;
    double re[B][4], im[B][4];
#pragma unroll
    for (int b = 0; b < B; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) re[b][q] = im[b][q] = 0.0;
    // two terms per step (their loads issued together), still smallest first
    auto term = [&](int m, const double (&u)[4]) {
        if (m & 1) {
#pragma unroll
            for (int b = 0; b < B; ++b)
#pragma unroll
                for (int q = 0; q < 4; ++q) re[b][q] = __builtin_fma(mb.cf[b].c[m], u[q], re[b][q]);
        } else {
#pragma unroll
            for (int b = 0; b < B; ++b)
#pragma unroll
                for (int q = 0; q < 4; ++q) im[b][q] = __builtin_fma(mb.cf[b].c[m], u[q], im[b][q]);
        }
    };
    int m = terms - 1;
    for (; m >= 1; m -= 2) {
        const double *src = base + (size_t)m * mstride, *nxt = src - mstride;
        double u[4], v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            u[q] = src[q * 1024];
            v[q] = nxt[q * 1024];
        }
        term(m, u);
        term(m - 1, v);
    }
    if (m == 0) {
        double u[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) u[q] = base[q * 1024];
        term(0, u);
    }
    const int j = tb * 64 + c;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = ta * 64 + r0 + 16 * q;
        if (i < R && j < R) {
            const size_t o = (size_t)i * ld + j;
            const double sv = stat[o];
#pragma unroll
            for (int b = 0; b < B; ++b) {
                re[b][q] += sv;
                mb.out[b][o] = make_double2(re[b][q], im[b][q]);
            }
        }
    }
    if (ta == tb) return;  // the diagonal tile holds both halves
    // lower triangle of each member: out[j][i] = value(i, j), through LDS
    const int i = ta * 64 + c;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        double tre[4];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) tr[r0 + 16 * q][c] = re[b][q];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) tre[q] = tr[c][r0 + 16 * q];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) tr[r0 + 16 * q][c] = im[b][q];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int jj = tb * 64 + r0 + 16 * q;
            if (i < R && jj < R) mb.out[b][(size_t)jj * ld + i] = make_double2(tre[q], tr[c][r0 + 16 * q]);
        }
    }
}

// ---------------------------------------------------------------------------
// moments around a shifted wave number k0 (the points a sweep has above the
// range of the moments at 0): with d = k - k0,
//   (exp(-jkR) - 1)/R = (exp(-jk0R) - 1)/R + sum_{m>=1} (-jd)^m/m! exp(-jk0R) R^(m-1),
// so V_0 = scale * sum_w (exp(-jk0R) - 1)/R (the remainder at k0, formed like
// smooth_tile_kernel) and V_m = scale * sum_w exp(-jk0R) R^(m-1), complex,
// stored as [re / im][m][upper tile][64 x 64] like the moments at 0.

constexpr int kShiftTerms = RegionSet::kMaxShiftTerms;

__device__ inline void store_shifted(double *mre, double *mim, size_t mstride, int nt, int oi, int oj, int terms,
                                     const double (&ur)[kShiftTerms], const double (&ui)[kShiftTerms],
                                     double scale) {
    const int a = min(oi, oj), b = max(oi, oj);
    const size_t p = moment_slot(a, b, nt);
    const bool diag = (a >> 6) == (b >> 6);
    const size_t q = p - (size_t)(a & 63) * 64 - (size_t)(b & 63) + (size_t)(b & 63) * 64 + (size_t)(a & 63);
#pragma unroll
    for (int m = 0; m < kShiftTerms; ++m)
        if (m < terms) {
            mre[(size_t)m * mstride + p] = scale * ur[m];
            mim[(size_t)m * mstride + p] = scale * ui[m];
            if (diag) {
                mre[(size_t)m * mstride + q] = scale * ur[m];
                mim[(size_t)m * mstride + q] = scale * ui[m];
            }
        }
}

// one point pair's contributions with weight w at distance R (rinv = 1/R, 0
// at R = 0)
__device__ inline void shifted_terms(double w, double R, double rinv, bool zero, double half_k0, int terms,
                                     double (&ur)[kShiftTerms], double (&ui)[kShiftTerms]) {
    double s2, sc;
    d_half_angle_terms(half_k0 * R, s2, sc);  // sin^2(k0R/2), sin(k0R/2) cos(k0R/2)
    ur[0] = __builtin_fma(-2.0 * w * rinv, s2, ur[0]);
    ui[0] = zero ? __builtin_fma(-2.0 * w, half_k0, ui[0]) : __builtin_fma(-2.0 * w * rinv, sc, ui[0]);
    // exp(-jk0R) = (1 - 2 sin^2) - 2j sin cos
    double tr = w * __builtin_fma(-2.0, s2, 1.0), ti = -2.0 * w * sc;
#pragma unroll
    for (int m = 1; m < kShiftTerms; ++m)
        if (m < terms) {
            ur[m] += tr;
            ui[m] += ti;
            tr *= R;
            ti *= R;
        }
}

template <int KIND>
__global__ __launch_bounds__(256) void shifted_tile_kernel(
    const double4 *__restrict__ pts_i, const double4 *__restrict__ pts_j,
    const int *__restrict__ npts, int R, int slots, const double *__restrict__ aux,
    const double *__restrict__ dir, double dir_tol, const int *__restrict__ perm,
    const int *__restrict__ tiles, int terms, double k0, double *__restrict__ mre, double *__restrict__ mim,
    size_t mstride, int nt) {
    extern __shared__ double4 sj[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int bi = tiles[2 * blockIdx.x], bj = tiles[2 * blockIdx.x + 1];
    const int i = bi * kTileI + lane;
    const bool iv = i < R;
    const double ai = iv ? aux[i] : 1.0;
    const double half_k0 = 0.5 * k0;
    double di[3] = {0.0, 0.0, 0.0};
    if (KIND == 1 && iv)
        for (int m = 0; m < 3; ++m) di[m] = dir[3 * i + m];
    double4 *myj = sj + (size_t)wave * slots;
    for (int jj = 0; jj < 4; ++jj) {
        const int j = bj * kTileJ + wave * 4 + jj;
        const bool jv = j < R;
        const int nj = jv ? npts[j] : 0;
        for (int q = lane; q < nj; q += 64) myj[q] = pts_j[(size_t)j * slots + q];
        __syncthreads();
        bool active = jv && iv && i <= j;
        double dot = 1.0;
        if (KIND == 1 && active) {
            dot = di[0] * dir[3 * j] + di[1] * dir[3 * j + 1] + di[2] * dir[3 * j + 2];
            active = !(fabs(dot) < dir_tol);
        }
        double ur[kShiftTerms], ui[kShiftTerms];
#pragma unroll
        for (int m = 0; m < kShiftTerms; ++m) ur[m] = ui[m] = 0.0;
        if (__ballot(active)) {
            for (int p = 0; p < slots; ++p) {
                const double4 P = iv ? pts_i[(size_t)p * R + i] : make_double4(0.0, 0.0, 0.0, 0.0);
                for (int q = 0; q < nj; ++q) {
                    const double4 Q = myj[q];
                    const double dx = P.x - Q.x, dy = P.y - Q.y, dz = P.z - Q.z;
                    const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
                    const bool zero = r2 == 0.0;
                    const double rinv = zero ? 0.0 : rsqrt_nr(r2);
                    shifted_terms(P.w * Q.w, r2 * rinv, rinv, zero, half_k0, terms, ur, ui);
                }
            }
        }
        if (jv && iv && i <= j) {
            const double sc = !active ? 0.0
                              : KIND == 0 ? kInv4PiEps0 / (ai * aux[j])
                                          : kMu0Over4Pi * dot / (ai * aux[j]);
            store_shifted(mre, mim, mstride, nt, perm[i], perm[j], terms, ur, ui, sc);
        }
        __syncthreads();
    }
}

// near entries with the near rule (overwrites), as near_moment_kernel
__global__ __launch_bounds__(64) void near_shifted_kernel(
    const int *__restrict__ pairs, int n_pairs, const int *__restrict__ ptr,
    const double *__restrict__ pv, int order, int kind, const double *__restrict__ aux,
    const double *__restrict__ dir, double dir_tol, int terms, double k0, double *__restrict__ mre,
    double *__restrict__ mim, size_t mstride, int nt, int *__restrict__ err) {
    __shared__ double part[2][kShiftTerms][65];
    const int e = blockIdx.x;
    if (e >= n_pairs) return;
    const int lane = threadIdx.x;
    const int i = pairs[2 * e], j = pairs[2 * e + 1];
    const int a0 = ptr[i], na = ptr[i + 1] - a0, b0 = ptr[j], nb = ptr[j + 1] - b0;
    const int tasks = na * nb * 4;
    if (tasks > kMaxTasks) {
        if (lane == 0) atomicOr(err, 2);
        return;
    }
    const double half_k0 = 0.5 * k0;
    double ur[kShiftTerms], ui[kShiftTerms];
#pragma unroll
    for (int m = 0; m < kShiftTerms; ++m) ur[m] = ui[m] = 0.0;
    for (int t = lane; t < tasks; t += 64) {
        const int pp = t >> 2, tp = t & 3;
        const int ia = pp / nb, ib = pp - ia * nb;
        double ta[3][3], tb[3][3];
        d_piece_triangle(pv + 12 * (size_t)(a0 + ia), tp >> 1, ta);
        d_piece_triangle(pv + 12 * (size_t)(b0 + ib), tp & 1, tb);
        for (int uu = 0; uu < order; ++uu)
            for (int vv = 0; vv < order; ++vv) {
                double ra[3], ja;
                d_tri_map(ta, 0.5 * (c_gl_x[order][uu] + 1.0), 0.5 * (c_gl_x[order][vv] + 1.0), ra, ja);
                const double wa = 0.25 * c_gl_w[order][uu] * c_gl_w[order][vv] * ja;
                for (int x = 0; x < order; ++x)
                    for (int y = 0; y < order; ++y) {
                        double rb[3], jb;
                        d_tri_map(tb, 0.5 * (c_gl_x[order][x] + 1.0), 0.5 * (c_gl_x[order][y] + 1.0), rb, jb);
                        const double wb = 0.25 * c_gl_w[order][x] * c_gl_w[order][y] * jb;
                        const double Rd = d_dist3(ra, rb);
                        const bool zero = !(Rd > 0.0);
                        shifted_terms(wa * wb, Rd, zero ? 0.0 : 1.0 / Rd, zero, half_k0, terms, ur, ui);
                    }
            }
    }
#pragma unroll
    for (int m = 0; m < kShiftTerms; ++m) {
        part[0][m][lane] = ur[m];
        part[1][m][lane] = ui[m];
    }
    __syncthreads();
    double sc;
    if (kind == (int)RegionKind::Node) {
        sc = kInv4PiEps0 / (aux[i] * aux[j]);
    } else {
        const double dot = d_dot3(dir + 3 * i, dir + 3 * j);
        if (fabs(dot) < dir_tol) return;
        sc = kMu0Over4Pi * dot / (aux[i] * aux[j]);
    }
    if (lane >= 2 * terms) return;
    const int part_id = lane / terms, m = lane % terms;
    double v = 0.0;
    for (int l = 0; l < 64; ++l) v += part[part_id][m][l];
    v *= sc;
    if (!isfinite(v)) atomicOr(err, 1);
    double *dst = part_id ? mim : mre;
    const int a = min(i, j), b = max(i, j);
    const size_t p = moment_slot(a, b, nt);
    dst[(size_t)m * mstride + p] = v;
    if ((a >> 6) == (b >> 6)) dst[(size_t)m * mstride + p - (size_t)(a & 63) * 64 - (b & 63) + (size_t)(b & 63) * 64 + (a & 63)] = v;
}

struct ShiftCoefficients {
    double cr[kShiftTerms], ci[kShiftTerms];  // (-jd)^m / m!, m = 0..terms-1 (c_0 = 1)
};

// out = static + sum_m c_m V_m (complex), smallest terms first; layout and
// transposed store as moment_eval_kernel
__global__ __launch_bounds__(256) void shifted_eval_kernel(
    const double *__restrict__ mre, const double *__restrict__ mim, size_t mstride, int terms, ShiftCoefficients cf,
    const int *__restrict__ tile_ids, int R, const double *__restrict__ stat, double2 *__restrict__ out,
    size_t ld) {
    __shared__ double tr[64][65];
    const int t = blockIdx.x;
    const int ta = tile_ids[2 * t], tb = tile_ids[2 * t + 1];
    const int c = threadIdx.x & 63, r0 = threadIdx.x >> 6;
    const size_t base = (size_t)t * 4096 + (size_t)r0 * 64 + c;
    double re[16], im[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) re[q] = im[q] = 0.0;
    for (int m = terms - 1; m >= 0; --m) {
        const double *sr = mre + (size_t)m * mstride + base, *si = mim + (size_t)m * mstride + base;
        const double cr = cf.cr[m], ci = cf.ci[m];
        if (ci == 0.0) {
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                re[q] = __builtin_fma(cr, sr[q * 256], re[q]);
                im[q] = __builtin_fma(cr, si[q * 256], im[q]);
            }
        } else {
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                re[q] = __builtin_fma(-ci, si[q * 256], re[q]);
                im[q] = __builtin_fma(ci, sr[q * 256], im[q]);
            }
        }
    }
    const int j = tb * 64 + c;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int i = ta * 64 + r0 + 4 * q;
        if (i < R && j < R) {
            const size_t o = (size_t)i * ld + j;
            re[q] += stat[o];
            out[o] = make_double2(re[q], im[q]);
        }
    }
    if (ta == tb) return;
    double tre[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) tr[r0 + 4 * q][c] = re[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) tre[q] = tr[c][r0 + 4 * q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) tr[r0 + 4 * q][c] = im[q];
    __syncthreads();
    const int i = ta * 64 + c;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int jj = tb * 64 + r0 + 4 * q;
        if (i < R && jj < R) out[(size_t)jj * ld + i] = make_double2(tre[q], tr[c][r0 + 4 * q]);
    }
}

// ---------------------------------------------------------------------------
// host side

namespace {

void piece_aabb(const double *pv, double lo[3], double hi[3], double &h) {
    for (int k = 0; k < 3; ++k) {
        lo[k] = hi[k] = pv[k];
        for (int i = 1; i < 4; ++i) {
            lo[k] = std::fmin(lo[k], pv[3 * i + k]);
            hi[k] = std::fmax(hi[k], pv[3 * i + k]);
        }
    }
    h = 0.0;
    for (int i = 0; i < 4; ++i) {
        const double *a = &pv[3 * i], *b = &pv[3 * ((i + 1) % 4)];
        double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
        h = std::fmax(h, std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
    }
}

}  // namespace

RegionSet::RegionSet(const Mesh *mesh, const DualMesh *dual, RegionKind kind,
                     const QuadratureOrders &q, int smooth_order, double dir_tol, hipStream_t s,
                     bool count_work, int near_smooth_order)
    : kind_(kind), q_(q), smooth_order_(smooth_order), near_smooth_order_(near_smooth_order), dir_tol_(dir_tol) {
    ensure_device_tables();
    const bool nodes = kind == RegionKind::Node;
    R_ = (int)(nodes ? mesh->n_nodes : mesh->n_edges);
    const PatchList *lists = nodes ? dual->node_patches : dual->edge_patches;
    h_ptr_.assign(R_ + 1, 0);
    for (int r = 0; r < R_; ++r) {
        h_ptr_[r + 1] = h_ptr_[r] + (int)lists[r].count;
        maxp_ = std::max(maxp_, (int)lists[r].count);
    }
    h_pv_.resize(12 * (size_t)h_ptr_[R_]);
    for (int r = 0; r < R_; ++r)
        for (size_t k = 0; k < lists[r].count; ++k)
            std::memcpy(&h_pv_[12 * (h_ptr_[r] + k)], lists[r].patches[k].points, 12 * sizeof(double));
    h_aux_.resize(R_);
    h_dir_.assign(3 * (size_t)R_, 0.0);
    for (int r = 0; r < R_; ++r) {
        if (nodes) {
            h_aux_[r] = dual->node_region_areas[r];
        } else {
            const double area = dual->edge_region_areas[r], len = mesh->edge_lengths[r];
            h_aux_[r] = (area <= 0.0 || len <= 0.0) ? 0.0 : area / len;  // inductance.c:152-167
            for (int k = 0; k < 3; ++k) h_dir_[3 * r + k] = mesh->edge_directions[3 * r + k];
        }
    }
    // kernel order: edges grouped by dominant axis so tiles of mutually
    // orthogonal branches are skipped as a whole; nodes keep their order
    h_perm_.resize(R_);
    for (int r = 0; r < R_; ++r) h_perm_[r] = r;
    if (!nodes) {
        auto axis = [&](int r) {
            const double *d = &h_dir_[3 * r];
            const double ax = std::fabs(d[0]), ay = std::fabs(d[1]), az = std::fabs(d[2]);
            return ax >= ay && ax >= az ? 0 : (ay >= az ? 1 : 2);
        };
        std::stable_sort(h_perm_.begin(), h_perm_.end(), [&](int a, int b) { return axis(a) < axis(b); });
    }
    std::vector<double> aux_k(R_), dir_k(3 * (size_t)R_);
    for (int r = 0; r < R_; ++r) {
        aux_k[r] = h_aux_[h_perm_[r]];
        for (int k = 0; k < 3; ++k) dir_k[3 * r + k] = h_dir_[3 * h_perm_[r] + k];
    }
    // upper-triangle tile list in kernel order
    const int ti = (R_ + kTileI - 1) / kTileI, tj = (R_ + kTileJ - 1) / kTileJ;
    for (int bi = 0; bi < ti; ++bi)
        for (int bj = 0; bj < tj; ++bj)
            if (bj * kTileJ + kTileJ - 1 >= bi * kTileI) {
                h_tiles_.push_back(bi);
                h_tiles_.push_back(bj);
            }
    std::vector<double> pi, pj;
    std::vector<int> nf, ns;
    build_points(h_pv_, h_ptr_, q_.far_order, pi, pj, far_slots_, nf);
    d_far_i_.upload(reinterpret_cast<const double4 *>(pi.data()), pi.size() / 4, s);
    d_far_j_.upload(reinterpret_cast<const double4 *>(pj.data()), pj.size() / 4, s);
    d_far_n_.upload(nf, s);
    if (smooth_order_ > 0) {
        std::vector<double> si, sjv;
        build_points(h_pv_, h_ptr_, smooth_order_, si, sjv, smooth_slots_, ns);
        d_smooth_i_.upload(reinterpret_cast<const double4 *>(si.data()), si.size() / 4, s);
        d_smooth_j_.upload(reinterpret_cast<const double4 *>(sjv.data()), sjv.size() / 4, s);
        d_smooth_n_.upload(ns, s);
    }
    build_near_list(mesh);
    // diameter of the quadrature points (they lie in the pieces' triangles):
    // the bounding-box diagonal of every piece vertex
    {
        double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
        for (size_t v = 0; v < h_pv_.size() / 3; ++v)
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::min(lo[d], h_pv_[3 * v + d]);
                hi[d] = std::max(hi[d], h_pv_[3 * v + d]);
            }
        double d2 = 0.0;
        for (int d = 0; d < 3; ++d)
            if (hi[d] >= lo[d]) d2 += (hi[d] - lo[d]) * (hi[d] - lo[d]);
        diameter_ = std::sqrt(d2);
    }
    d_ptr_.upload(h_ptr_, s);
    d_pv_.upload(h_pv_, s);
    d_aux_.upload(h_aux_, s);
    d_dir_.upload(h_dir_, s);
    d_aux_k_.upload(aux_k, s);
    d_dir_k_.upload(dir_k, s);
    d_perm_.upload(h_perm_, s);
    d_tiles_.upload(h_tiles_, s);
    d_near_.upload(h_near_, s);
    // algorithmic evaluation counts of the tile kernels (upper triangle,
    // non-orthogonal pairs, actual points), for roofline accounting
    far_evals_ = smooth_evals_ = 0;
    if (count_work)
    for (int a = 0; a < R_; ++a)
        for (int b = a; b < R_; ++b) {
            if (!nodes) {
                const double *da = &h_dir_[3 * a], *db = &h_dir_[3 * b];
                if (std::fabs(da[0] * db[0] + da[1] * db[1] + da[2] * db[2]) < dir_tol_) continue;
            }
            const long long pa = h_ptr_[a + 1] - h_ptr_[a], pb = h_ptr_[b + 1] - h_ptr_[b];
            far_evals_ += pa * pb * 4LL * q_.far_order * q_.far_order * q_.far_order * q_.far_order;
            smooth_evals_ += pa * pb * 4LL * smooth_order_ * smooth_order_ * smooth_order_ * smooth_order_;
        }
}

// Gauss points of every piece for a given order: 2 triangles x order^2,
// (x, y, z, W) with W = (w_u/2)(w_v/2) J. Two layouts: [slot][region] for the
// lane-per-region side and [region][slot] for the LDS-staged side. Missing
// slots of regions with fewer pieces carry W = 0 at a far-away point.
void RegionSet::build_points(const std::vector<double> &pv, const std::vector<int> &ptr, int order,
                             std::vector<double> &pi, std::vector<double> &pj, int &slots,
                             std::vector<int> &npts) {
    const GLTable &t = gl_table();
    const int per_piece = 2 * order * order;
    slots = maxp_ * per_piece;
    slots = (slots + 3) / 4 * 4;
    pi.assign(4 * (size_t)slots * R_, 0.0);
    pj.assign(4 * (size_t)slots * R_, 0.0);
    npts.assign(R_, 0);
    for (int kr = 0; kr < R_; ++kr) {
        const int r = h_perm_.empty() ? kr : h_perm_[kr];
        int n = 0;
        for (int p = ptr[r]; p < ptr[r + 1]; ++p) {
            const double *v = &pv[12 * (size_t)p];
            for (int tri = 0; tri < 2; ++tri) {
                const double *p0 = v, *p1 = v + 3 * (tri == 0 ? 1 : 2), *p2 = v + 3 * (tri == 0 ? 2 : 3);
                double e1[3], e2[3], c[3];
                for (int d = 0; d < 3; ++d) {
                    e1[d] = p1[d] - p0[d];
                    e2[d] = p2[d] - p0[d];
                }
                c[0] = e1[1] * e2[2] - e1[2] * e2[1];
                c[1] = e1[2] * e2[0] - e1[0] * e2[2];
                c[2] = e1[0] * e2[1] - e1[1] * e2[0];
                const double area2 = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
                for (int a = 0; a < order; ++a) {
                    const double u = 0.5 * (t.x[order][a] + 1.0), wu = 0.5 * t.w[order][a];
                    for (int b = 0; b < order; ++b) {
                        const double vv = 0.5 * (t.x[order][b] + 1.0), wv = 0.5 * t.w[order][b];
                        const double eta = (1.0 - u) * vv;
                        double P[4];
                        for (int d = 0; d < 3; ++d) P[d] = p0[d] + u * e1[d] + eta * e2[d];
                        P[3] = wu * wv * (area2 * (1.0 - u));
                        std::memcpy(&pi[4 * ((size_t)n * R_ + kr)], P, sizeof P);
                        std::memcpy(&pj[4 * ((size_t)kr * slots + n)], P, sizeof P);
                        ++n;
                    }
                }
            }
        }
        npts[kr] = n;
        for (int m = n; m < slots; ++m) {
            // zero weight at a remote but moderate point: 1/R and exp(-jkR)
            // stay finite, so the slot contributes exactly 0
            const double far[4] = {1.0e4, 1.0e4, 1.0e4, 0.0};
            std::memcpy(&pi[4 * ((size_t)m * R_ + kr)], far, sizeof far);
            std::memcpy(&pj[4 * ((size_t)kr * slots + m)], far, sizeof far);
        }
    }
}

// Candidate near entries: region AABB gap < near_factor * (largest piece
// side of either region). Every SELF/TOUCHING/NEAR piece pair implies this
// (piece gaps >= region gaps and sides <= the region maximum, both exactly
// monotone in floating point), so the list is a superset of the entries
// the reference computes with the analytic near-field rule.
void RegionSet::build_near_list(const Mesh *mesh) {
    (void)mesh;
    const bool nodes = kind_ == RegionKind::Node;
    std::vector<double> lo(3 * (size_t)R_), hi(3 * (size_t)R_), H(R_, 0.0);
    double hmax = 0.0, ext = 0.0;
    for (int r = 0; r < R_; ++r) {
        for (int k = 0; k < 3; ++k) {
            lo[3 * r + k] = HUGE_VAL;
            hi[3 * r + k] = -HUGE_VAL;
        }
        for (int p = h_ptr_[r]; p < h_ptr_[r + 1]; ++p) {
            double a[3], b[3], h;
            piece_aabb(&h_pv_[12 * (size_t)p], a, b, h);
            for (int k = 0; k < 3; ++k) {
                lo[3 * r + k] = std::fmin(lo[3 * r + k], a[k]);
                hi[3 * r + k] = std::fmax(hi[3 * r + k], b[k]);
            }
            H[r] = std::fmax(H[r], h);
        }
        hmax = std::fmax(hmax, H[r]);
        for (int k = 0; k < 3; ++k) ext = std::fmax(ext, hi[3 * r + k] - lo[3 * r + k]);
    }
    const double reach = q_.near_factor * hmax;
    const double cell = std::fmax(ext + reach, 1e-300);
    double origin[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL};
    for (int r = 0; r < R_; ++r)
        for (int k = 0; k < 3; ++k) origin[k] = std::fmin(origin[k], lo[3 * r + k]);
    auto key = [](long long x, long long y, long long z) {
        return (uint64_t)((x & 0x1fffff) | ((y & 0x1fffff) << 21) | ((z & 0x1fffff) << 42));
    };
    std::unordered_map<uint64_t, std::vector<int>> grid;
    std::vector<long long> cellof(3 * (size_t)R_);
    for (int r = 0; r < R_; ++r) {
        for (int k = 0; k < 3; ++k) cellof[3 * r + k] = (long long)std::floor((lo[3 * r + k] - origin[k]) / cell);
        grid[key(cellof[3 * r], cellof[3 * r + 1], cellof[3 * r + 2])].push_back(r);
    }
    std::vector<std::vector<int>> partners(R_);
    for (int r = 0; r < R_; ++r) {
        for (long long dz = -1; dz <= 1; ++dz)
            for (long long dy = -1; dy <= 1; ++dy)
                for (long long dx = -1; dx <= 1; ++dx) {
                    auto it = grid.find(key(cellof[3 * r] + dx, cellof[3 * r + 1] + dy, cellof[3 * r + 2] + dz));
                    if (it == grid.end()) continue;
                    for (int o : it->second) {
                        if (o < r) continue;
                        double g2 = 0.0;
                        for (int k = 0; k < 3; ++k) {
                            const double g = std::fmax(0.0, std::fmax(lo[3 * r + k] - hi[3 * o + k],
                                                                      lo[3 * o + k] - hi[3 * r + k]));
                            g2 += g * g;
                        }
                        const double h = std::fmax(H[r], H[o]);
                        if (!(h > 0.0 && std::sqrt(g2) < q_.near_factor * h)) continue;
                        if (!nodes) {
                            const double *da = &h_dir_[3 * r], *db = &h_dir_[3 * o];
                            if (std::fabs(da[0] * db[0] + da[1] * db[1] + da[2] * db[2]) < dir_tol_) continue;
                        }
                        partners[r].push_back(o);
                    }
                }
    }
    h_near_.clear();
    for (int r = 0; r < R_; ++r) {
        std::sort(partners[r].begin(), partners[r].end());
        for (int o : partners[r]) {
            h_near_.push_back(r);
            h_near_.push_back(o);
        }
    }
}

void RegionSet::assemble_static(double *d_out, hipStream_t s, int *d_error, AssemblyWork *work,
                                KernelTimer *timer, unsigned long long *d_counts) {
    const int ntiles = (int)h_tiles_.size() / 2;
    const size_t lds = 4 * (size_t)far_slots_ * sizeof(double4);
    if (ntiles) {
        const int tk = timer ? timer->start(s) : -1;
        if (kind_ == RegionKind::Node)
            hipLaunchKernelGGL(far_tile_kernel<0>, dim3(ntiles), dim3(256), lds, s, d_far_i_.ptr,
                               d_far_j_.ptr, d_far_n_.ptr, R_, far_slots_, d_aux_k_.ptr,
                               d_dir_k_.ptr, dir_tol_, d_perm_.ptr, d_tiles_.ptr, d_out, (size_t)R_);
        else
            hipLaunchKernelGGL(far_tile_kernel<1>, dim3(ntiles), dim3(256), lds, s, d_far_i_.ptr,
                               d_far_j_.ptr, d_far_n_.ptr, R_, far_slots_, d_aux_k_.ptr,
                               d_dir_k_.ptr, dir_tol_, d_perm_.ptr, d_tiles_.ptr, d_out, (size_t)R_);
        HIP_CHECK(hipGetLastError());
        // 12 flops per 1/R evaluation; bytes: the matrix written once
        if (timer) timer->stop(tk, s, kKernelFarTile, 12.0 * far_evals_, 8.0 * R_ * R_);
    }
    const int nn = (int)h_near_.size() / 2;
    if (nn) {
        const int tk = timer ? timer->start(s) : -1;
        hipLaunchKernelGGL(near_entry_kernel, dim3(nn), dim3(64), 0, s, d_near_.ptr, nn, d_ptr_.ptr, d_pv_.ptr,
                           q_, (int)kind_, d_aux_.ptr, d_dir_.ptr, d_out, (size_t)R_, 1, d_error, d_counts);
        HIP_CHECK(hipGetLastError());
        // flops follow from the evaluation counts once the kernel is done
        if (timer) timer->stop(tk, s, kKernelNearEntry, 0.0, 16.0 * nn);
    }
    if (work) {
        work->far_evaluations += far_evals_;
        work->near_entries += nn;
    }
}

void RegionSet::assemble_retarded(const double *d_static, double k, double2 *d_out, hipStream_t s,
                                  AssemblyWork *work, KernelTimer *timer) {
    const int ntiles = (int)h_tiles_.size() / 2;
    const size_t lds = 4 * (size_t)smooth_slots_ * sizeof(double4);
    if (!ntiles) return;
    if (smooth_order_ <= 0) throw HipError{"region set built without smooth-rule points"};
    const int tk = timer ? timer->start(s) : -1;
    if (kind_ == RegionKind::Node)
        hipLaunchKernelGGL(smooth_tile_kernel<0>, dim3(ntiles), dim3(256), lds, s, d_smooth_i_.ptr,
                           d_smooth_j_.ptr, d_smooth_n_.ptr, R_, smooth_slots_, d_aux_k_.ptr,
                           d_dir_k_.ptr, dir_tol_, d_perm_.ptr, d_tiles_.ptr, k, d_static, d_out,
                           (size_t)R_);
    else
        hipLaunchKernelGGL(smooth_tile_kernel<1>, dim3(ntiles), dim3(256), lds, s, d_smooth_i_.ptr,
                           d_smooth_j_.ptr, d_smooth_n_.ptr, R_, smooth_slots_, d_aux_k_.ptr,
                           d_dir_k_.ptr, dir_tol_, d_perm_.ptr, d_tiles_.ptr, k, d_static, d_out,
                           (size_t)R_);
    HIP_CHECK(hipGetLastError());
    // per exp(-jkR) evaluation: 1/R (12) + half-angle reduction and the
    // sin (degree 15) and cos (degree 16) polynomials (~22) + complex
    // accumulate (4) = 38 flops; bytes: static matrix read + complex matrix
    // written
    if (timer) timer->stop(tk, s, kKernelSmoothTile, 38.0 * smooth_evals_, 8.0 * R_ * R_ + 16.0 * R_ * R_);
    const int nn = (int)h_near_.size() / 2;
    if (nn && near_smooth_order_ > 0) {
        if (!d_smooth_err_.ptr) d_smooth_err_.resize(1);
        hipLaunchKernelGGL(near_smooth_kernel, dim3(nn), dim3(64), 0, s, d_near_.ptr, nn, d_ptr_.ptr, d_pv_.ptr,
                           near_smooth_order_, k, (int)kind_, d_aux_.ptr, d_dir_.ptr, dir_tol_, d_static, d_out,
                           (size_t)R_, d_smooth_err_.ptr);
        HIP_CHECK(hipGetLastError());
    }
    if (work) work->smooth_evaluations += smooth_evals_;
}

size_t RegionSet::moment_bytes(int regions, int terms) {
    const size_t nt = ((size_t)regions + 63) / 64;
    return nt * (nt + 1) / 2 * 4096 * sizeof(double) * (size_t)std::max(terms, 0);
}

int RegionSet::series_terms(double k) const {
    const double x = k * diameter_;
    double t = x;  // x^m / m!
    for (int m = 1; m <= kMaxMomentTerms; ++m) {
        t *= x / (m + 1);
        if (t <= 1e-17) return m;
    }
    return kMaxMomentTerms + 1;
}

void RegionSet::build_moments(int terms, hipStream_t s) {
    terms = std::min(terms, (int)kMaxMomentTerms);
    moment_terms_ = 0;
    d_moments_.release();
    if (terms <= 0 || R_ == 0) return;
    if (smooth_order_ <= 0) throw HipError{"region set built without smooth-rule points"};
    const int nt = (R_ + 63) / 64;
    moment_tiles_ = nt * (nt + 1) / 2;
    std::vector<int> ids;
    ids.reserve(2 * (size_t)moment_tiles_);
    for (int a = 0; a < nt; ++a)
        for (int b = a; b < nt; ++b) {
            ids.push_back(a);
            ids.push_back(b);
        }
    d_moment_tiles_.upload(ids, s);
    const size_t mstride = (size_t)moment_tiles_ * 4096;
    d_moments_.resize(mstride * terms);
    HIP_CHECK(hipMemsetAsync(d_moments_.ptr, 0, mstride * terms * sizeof(double), s));
    const int ntiles = (int)h_tiles_.size() / 2;
    const size_t lds = 4 * (size_t)smooth_slots_ * sizeof(double4);
    if (kind_ == RegionKind::Node)
        hipLaunchKernelGGL(moment_tile_kernel<0>, dim3(ntiles), dim3(256), lds, s, d_smooth_i_.ptr,
                           d_smooth_j_.ptr, d_smooth_n_.ptr, R_, smooth_slots_, d_aux_k_.ptr, d_dir_k_.ptr,
                           dir_tol_, d_perm_.ptr, d_tiles_.ptr, terms, d_moments_.ptr, mstride, nt);
    else
        hipLaunchKernelGGL(moment_tile_kernel<1>, dim3(ntiles), dim3(256), lds, s, d_smooth_i_.ptr,
                           d_smooth_j_.ptr, d_smooth_n_.ptr, R_, smooth_slots_, d_aux_k_.ptr, d_dir_k_.ptr,
                           dir_tol_, d_perm_.ptr, d_tiles_.ptr, terms, d_moments_.ptr, mstride, nt);
    HIP_CHECK(hipGetLastError());
    const int nn = (int)h_near_.size() / 2;
    if (nn && near_smooth_order_ > 0) {
        if (!d_smooth_err_.ptr) d_smooth_err_.resize(1);
        HIP_CHECK(hipMemsetAsync(d_smooth_err_.ptr, 0, sizeof(int), s));
        hipLaunchKernelGGL(near_moment_kernel, dim3(nn), dim3(64), 0, s, d_near_.ptr, nn, d_ptr_.ptr, d_pv_.ptr,
                           near_smooth_order_, (int)kind_, d_aux_.ptr, d_dir_.ptr, dir_tol_, terms,
                           d_moments_.ptr, mstride, nt, d_smooth_err_.ptr);
        HIP_CHECK(hipGetLastError());
        int err = 0;
        HIP_CHECK(hipMemcpyAsync(&err, d_smooth_err_.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        if (err) throw HipError{"near-field moment pass failed (non-finite value or too many tasks)"};
    }
    moment_terms_ = terms;
}

void RegionSet::assemble_retarded_moments(const double *d_static, double k, int terms, double2 *d_out,
                                          hipStream_t s, KernelTimer *timer) {
    if (terms < 1 || terms > moment_terms_) throw HipError{"moment terms out of range"};
    MomentCoefficients cf{};
    // coefficient of U_m: (-jk)^m / m!; (-j)^m = 1, -j, -1, j for m = 0, 1, 2, 3 (mod 4)
    double km = 1.0;
    for (int m = 1; m <= terms; ++m) {
        km *= k / m;
        const int r = m & 3;
        cf.c[m - 1] = (r == 1 || r == 2) ? -km : km;
    }
    const int tk = timer ? timer->start(s) : -1;
    const size_t mstride = (size_t)moment_tiles_ * 4096;
    hipLaunchKernelGGL(moment_eval_kernel, dim3(moment_tiles_), dim3(256), 0, s, d_moments_.ptr, mstride, terms,
                       cf, d_moment_tiles_.ptr, R_, d_static, d_out, (size_t)R_);
    HIP_CHECK(hipGetLastError());
    // two flops per stored value; bytes: the moments and the static upper
    // tiles read, the complex matrix written
    if (timer)
        timer->stop(tk, s, kKernelSmoothMoments, 2.0 * terms * mstride,
                    8.0 * (terms + 1) * mstride + 16.0 * R_ * R_);
}

void RegionSet::assemble_retarded_moments_batch(const double *d_static, const double *k, const int *terms,
                                                double2 *const *d_out, int count, hipStream_t s,
                                                KernelTimer *timer) {
    if (count < 1 || count > kMaxMomentBatch) throw HipError{"moment batch size out of range"};
    MomentBatch mb{};
    int tmax = 0;
    for (int b = 0; b < count; ++b) {
        if (terms[b] < 1 || terms[b] > moment_terms_) throw HipError{"moment terms out of range"};
        // the coefficients of assemble_retarded_moments
        double km = 1.0;
        for (int m = 1; m <= terms[b]; ++m) {
            km *= k[b] / m;
            const int r = m & 3;
            mb.cf[b].c[m - 1] = (r == 1 || r == 2) ? -km : km;
        }
        mb.out[b] = d_out[b];
        tmax = std::max(tmax, terms[b]);
    }
    const int tk = timer ? timer->start(s) : -1;
    const size_t mstride = (size_t)moment_tiles_ * 4096;
#define PEEC_MB(B)                                                                                                \
    hipLaunchKernelGGL(moment_eval_batch_kernel<B>, dim3(moment_tiles_), dim3(1024), 0, s, d_moments_.ptr, mstride, \
                       tmax, mb, d_moment_tiles_.ptr, R_, d_static, (size_t)R_)
    if (count == 1)
        PEEC_MB(1);
    else if (count == 2)
        PEEC_MB(2);
    else if (count == 3)
        PEEC_MB(3);
    else
        PEEC_MB(4);
#undef PEEC_MB
    HIP_CHECK(hipGetLastError());
    if (timer)
        timer->stop(tk, s, kKernelSmoothMoments, 2.0 * count * tmax * mstride,
                    8.0 * (tmax + 1) * mstride + 16.0 * count * R_ * R_);
}

int RegionSet::shifted_terms_for(double k) const {
    if (shift_terms_ <= 0) return kMaxShiftTerms + 1;
    const double x = std::fabs(k - shift_k0_) * diameter_;
    double t = 1.0;  // x^m / m!
    for (int m = 0; m < kMaxShiftTerms; ++m) {
        t *= x / (m + 1);
        if (t <= 1e-17) return m + 1;  // terms m = 0..m
    }
    return kMaxShiftTerms + 1;
}

void RegionSet::build_shifted_moments(int terms, double k0, hipStream_t s) {
    terms = std::min(terms, (int)kMaxShiftTerms);
    shift_terms_ = 0;
    d_shifted_.release();
    if (terms <= 0 || R_ == 0) return;
    if (smooth_order_ <= 0) throw HipError{"region set built without smooth-rule points"};
    const int nt = (R_ + 63) / 64;
    if (!d_moment_tiles_.ptr) {
        std::vector<int> ids;
        for (int a = 0; a < nt; ++a)
            for (int b = a; b < nt; ++b) {
                ids.push_back(a);
                ids.push_back(b);
            }
        d_moment_tiles_.upload(ids, s);
        moment_tiles_ = nt * (nt + 1) / 2;
    }
    const size_t mstride = (size_t)moment_tiles_ * 4096;
    d_shifted_.resize(2 * mstride * terms);
    double *mre = d_shifted_.ptr, *mim = d_shifted_.ptr + mstride * terms;
    HIP_CHECK(hipMemsetAsync(d_shifted_.ptr, 0, 2 * mstride * terms * sizeof(double), s));
    const int ntiles = (int)h_tiles_.size() / 2;
    const size_t lds = 4 * (size_t)smooth_slots_ * sizeof(double4);
    if (kind_ == RegionKind::Node)
        hipLaunchKernelGGL(shifted_tile_kernel<0>, dim3(ntiles), dim3(256), lds, s, d_smooth_i_.ptr, d_smooth_j_.ptr,
                           d_smooth_n_.ptr, R_, smooth_slots_, d_aux_k_.ptr, d_dir_k_.ptr, dir_tol_, d_perm_.ptr,
                           d_tiles_.ptr, terms, k0, mre, mim, mstride, nt);
    else
        hipLaunchKernelGGL(shifted_tile_kernel<1>, dim3(ntiles), dim3(256), lds, s, d_smooth_i_.ptr, d_smooth_j_.ptr,
                           d_smooth_n_.ptr, R_, smooth_slots_, d_aux_k_.ptr, d_dir_k_.ptr, dir_tol_, d_perm_.ptr,
                           d_tiles_.ptr, terms, k0, mre, mim, mstride, nt);
    HIP_CHECK(hipGetLastError());
    const int nn = (int)h_near_.size() / 2;
    if (nn && near_smooth_order_ > 0) {
        if (!d_smooth_err_.ptr) d_smooth_err_.resize(1);
        HIP_CHECK(hipMemsetAsync(d_smooth_err_.ptr, 0, sizeof(int), s));
        hipLaunchKernelGGL(near_shifted_kernel, dim3(nn), dim3(64), 0, s, d_near_.ptr, nn, d_ptr_.ptr, d_pv_.ptr,
                           near_smooth_order_, (int)kind_, d_aux_.ptr, d_dir_.ptr, dir_tol_, terms, k0, mre, mim,
                           mstride, nt, d_smooth_err_.ptr);
        HIP_CHECK(hipGetLastError());
        int err = 0;
        HIP_CHECK(hipMemcpyAsync(&err, d_smooth_err_.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        if (err) throw HipError{"near-field shifted moment pass failed (non-finite value or too many tasks)"};
    }
    shift_terms_ = terms;
    shift_k0_ = k0;
}

void RegionSet::assemble_retarded_shifted(const double *d_static, double k, int terms, double2 *d_out,
                                          hipStream_t s, KernelTimer *timer) {
    if (terms < 1 || terms > shift_terms_) throw HipError{"shifted moment terms out of range"};
    ShiftCoefficients cf{};
    const double d = k - shift_k0_;
    double a = 1.0;  // d^m / m!
    for (int m = 0; m < terms; ++m) {
        if (m > 0) a *= d / m;
        switch (m & 3) {  // (-j)^m = 1, -j, -1, j
        case 0: cf.cr[m] = a; cf.ci[m] = 0.0; break;
        case 1: cf.cr[m] = 0.0; cf.ci[m] = -a; break;
        case 2: cf.cr[m] = -a; cf.ci[m] = 0.0; break;
        default: cf.cr[m] = 0.0; cf.ci[m] = a; break;
        }
    }
    const int tk = timer ? timer->start(s) : -1;
    const size_t mstride = (size_t)moment_tiles_ * 4096;
    const double *mre = d_shifted_.ptr, *mim = d_shifted_.ptr + mstride * shift_terms_;
    hipLaunchKernelGGL(shifted_eval_kernel, dim3(moment_tiles_), dim3(256), 0, s, mre, mim, mstride, terms, cf,
                       d_moment_tiles_.ptr, R_, d_static, d_out, (size_t)R_);
    HIP_CHECK(hipGetLastError());
    if (timer)
        timer->stop(tk, s, kKernelSmoothMoments, 4.0 * terms * mstride,
                    16.0 * terms * mstride + 8.0 * mstride + 16.0 * R_ * R_);
}

void RegionSet::assemble_pairs(const std::vector<int> &pairs, double *d_values, hipStream_t s,
                               int *d_error) {
    DeviceBuffer<int> d_pairs;
    d_pairs.upload(pairs, s);
    const int n = (int)pairs.size() / 2;
    if (n) {
        hipLaunchKernelGGL(near_entry_kernel, dim3(n), dim3(64), 0, s, d_pairs.ptr, n, d_ptr_.ptr,
                           d_pv_.ptr, q_, (int)kind_, d_aux_.ptr, d_dir_.ptr, d_values, (size_t)0,
                           0, d_error);
        HIP_CHECK(hipGetLastError());
    }
    HIP_CHECK(hipStreamSynchronize(s));
}

// Raw integral over all piece pairs of two explicit piece lists (regions),
// through the near-field kernel; `same` marks one region paired with itself.
double region_pair_integral(const std::vector<double> &pa, const std::vector<double> &pb, bool same,
                            const QuadratureOrders &q) {
    ensure_device_tables();
    Stream s;
    std::vector<double> pv(pa);
    if (!same) pv.insert(pv.end(), pb.begin(), pb.end());
    const int na = (int)(pa.size() / 12), nb = (int)(pb.size() / 12);
    std::vector<int> ptr = {0, na, same ? na : na + nb};
    std::vector<int> pairs = {0, same ? 0 : 1};
    DeviceBuffer<double> d_pv, d_val(1), d_aux(2), d_dir(6);
    DeviceBuffer<int> d_ptr, d_pairs, d_err(1);
    d_pv.upload(pv, s);
    d_ptr.upload(ptr, s);
    d_pairs.upload(pairs, s);
    HIP_CHECK(hipMemsetAsync(d_err.ptr, 0, sizeof(int), s));
    hipLaunchKernelGGL(near_entry_kernel, dim3(1), dim3(64), 0, s, d_pairs.ptr, 1, d_ptr.ptr, d_pv.ptr, q, 2,
                       d_aux.ptr, d_dir.ptr, d_val.ptr, (size_t)0, 0, d_err.ptr);
    HIP_CHECK(hipGetLastError());
    double v = NAN;
    int err = 0;
    d_val.download(&v, 1, s);
    d_err.download(&err, 1, s);
    s.sync();
    return err ? NAN : v;
}

// K(a, b) for two free-standing pieces through the near-field kernel (which
// classifies the pair and applies the FAR rule when appropriate).
double patch_pair_integral(const QuadPatch *a, const QuadPatch *b, const QuadratureOrders &q) {
    ensure_device_tables();
    Stream s;
    std::vector<double> pv(24);
    std::memcpy(&pv[0], a->points, 12 * sizeof(double));
    std::memcpy(&pv[12], b->points, 12 * sizeof(double));
    const bool same = a == b;
    std::vector<int> ptr = {0, 1, 2};
    std::vector<int> pairs = {0, same ? 0 : 1};
    DeviceBuffer<double> d_pv, d_val(1), d_aux(2), d_dir(6);
    DeviceBuffer<int> d_ptr, d_pairs, d_err(1);
    d_pv.upload(pv, s);
    d_ptr.upload(ptr, s);
    d_pairs.upload(pairs, s);
    HIP_CHECK(hipMemsetAsync(d_err.ptr, 0, sizeof(int), s));
    hipLaunchKernelGGL(near_entry_kernel, dim3(1), dim3(64), 0, s, d_pairs.ptr, 1, d_ptr.ptr,
                       d_pv.ptr, q, 2, d_aux.ptr, d_dir.ptr, d_val.ptr, (size_t)0, 0, d_err.ptr);
    HIP_CHECK(hipGetLastError());
    double v = NAN;
    int err = 0;
    d_val.download(&v, 1, s);
    d_err.download(&err, 1, s);
    s.sync();
    return err ? NAN : v;
}

}  // namespace peec

// Blocked right-looking LU with partial pivoting on the GPU (row-major,
// in place), templated on the storage element:
//   double2  complex FP64   (drop-in complex LU, FP64 oracle path)
//   float2   complex FP32   (mixed-precision factor; panel arithmetic and
//                            pivot magnitudes in FP64, trailing update on
//                            FP32 MFMA)
//   double   real FP64      (drop-in real LU)
#pragma once

#include "common.h"
#include "mna.h"

namespace peec {

struct LUStatus {
    int code = 0;        // 0 ok, -2 singular/invalid pivot, -3 internal timeout
    long long column = -1;
};

// p += blockIdx.z * stride (null pointers stay null)
template <class P> __device__ inline void batch_shift(P &p, size_t stride) {
    if (p) p += (size_t)blockIdx.z * stride;
}

// Strides between the members of a batch of equally sized factors that run
// in the same launches, one member per grid z index; all zero for a single
// factor. A kernel moves its pointers to member blockIdx.z on entry.
struct BatchStride {
    size_t a = 0;                    // float2 elements between the matrices
    size_t dinv = 0;                 // float2 elements between the 16 x 16 inverse sets
    size_t pl = 0, pu = 0;           // uint4 units between the L21 / U12 operand packs
    size_t sl = 0, su = 0;           // floats between their scales
    int status = 0;                  // ints between the status pairs
    const double *limits = nullptr;  // pivot limit of each member (null: the kernel argument)
    template <class TA, class TS> __device__ void apply(TA &A, TS &st) const {
        batch_shift(A, a);
        batch_shift(st, (size_t)status);
    }
    template <class TA, class TS, class TD> __device__ void apply(TA &A, TS &st, TD &d) const {
        apply(A, st);
        batch_shift(d, dinv);
    }
    template <class TA, class TS, class TD> __device__ void apply(TA &A, TS &st, TD &d, double &limit) const {
        apply(A, st, d);
        if (limits) limit = limits[blockIdx.z];
    }
    template <class TP, class TF> __device__ void apply_pack(TP &p, TF &f, size_t ps, size_t fs) const {
        batch_shift(p, ps);
        batch_shift(f, fs);
    }
};

// Scratch owned by one factorization context (one stream).
struct LUContext {
    DeviceBuffer<unsigned> barrier;   // grid barrier counter
    DeviceBuffer<int> status;         // [0] code, [1] column
    DeviceBuffer<double> slots;       // pivot-row exchange for the panel kernel
    DeviceBuffer<double> cand;        // per-workgroup pivot candidate magnitudes
    DeviceBuffer<unsigned long long> tags;  // per-workgroup (round, row) tags
    unsigned round = 0;               // pivot-exchange round counter
    int cus = 0;
    hipStream_t stream = nullptr;
    KernelTimer *timer = nullptr;      // optional per-kernel timing
    DeviceBuffer<int> solve_bad;
    DeviceBuffer<double> solve_tmp;
    DeviceBuffer<float2> solve_f32;              // FP32 right-hand sides of the blocked solves
    DeviceBuffer<unsigned char> pack_l, pack_u;  // MFMA-ordered trailing-update operands
    // split-FP16 operands of the ordered factor's update and their per-row /
    // per-column inverse scales
    DeviceBuffer<unsigned char> pack_l16, pack_u16;
    DeviceBuffer<float> scale_l16, scale_u16;
    DeviceBuffer<unsigned char> pack_u16n;  // the next block column's U12 (look-ahead update)
    DeviceBuffer<float> scale_u16n;
    // a second set of L21 / U12 packs: a panel pair whose trailing update is
    // deferred keeps its packs while the next pair writes its own
    DeviceBuffer<unsigned char> pack_l16b, pack_u16b;
    DeviceBuffer<float> scale_l16b, scale_u16b;
    bool defer_updates = true;                    // ordered factor: every other pair's trailing update merged
    // inverses of the ordered factor's 16 x 16 diagonal blocks, two panels'
    // worth (by panel parity): the look-ahead factor of the next panel writes
    // its set while the current panel's solves read theirs
    DeviceBuffer<float2> dblk_inv;
    bool split_fp16_update = true;                // ordered factor: FP16 hi/lo update (else FP32 MFMA)
    int h3_tile_rows = 64;                        // rows of the split-FP16 update's tiles (64 or 32)
    bool pivoted_split16 = false;                 // partial-pivoting FP32 factor: split-FP16 trailing updates
    DeviceBuffer<int> swap_plan;                  // row moves of one panel
    DeviceBuffer<unsigned char> swap_stage;       // staged source rows of those moves
    DeviceBuffer<unsigned char> diag_inv;         // inverses of the 64x64 diagonal blocks (solves)
    const void *dinv_factor = nullptr;            // factor whose inverses diag_inv holds
    int dinv_n = 0;
    int dinv_members = 1;                         // ... and of how many batch members
    int lds_per_block = 64 * 1024;
    size_t panel_lds_limit = 0;                   // cap on the pivoted panel's LDS per workgroup (0: none)
    hipStream_t aux = nullptr;                    // look-ahead stream (default priority)
    hipEvent_t ev_next = nullptr, ev_diag = nullptr, ev_prev = nullptr, ev_u12 = nullptr;
    explicit LUContext(hipStream_t s);
    ~LUContext();
    LUContext(const LUContext &) = delete;
    LUContext &operator=(const LUContext &) = delete;
};

// Per-row / per-column max equilibration (complex_lu.c:275-320) of a
// row-major complex FP64 matrix in place: A <- diag(1/r) A diag(1/c).
// Returns 0, or -2 when a row or column is zero or not finite.
// `max_scaled` receives max |A_eq|.
int equilibrate_complex(double2 *dA, int n, size_t lda, double *d_row, double *d_col,
                        double *max_scaled, LUContext &ctx);

// The same equilibration without writing the matrix (mixed path): row and
// column scales of an unscaled matrix (0 or -2 as above; rows_ready: d_row
// already holds the row maxima) ...
int equilibration_scales(const double2 *dA, int n, size_t lda, double *d_row, double *d_col, LUContext &ctx,
                         bool rows_ready = false);
// ... or of the MNA matrix of m, evaluated from its blocks (row maxima
// included) ...
int equilibration_scales(const MnaOperands &m, double *d_row, double *d_col, LUContext &ctx);
// p[0..n) = 0 on stream s
void zero_doubles(double *p, size_t n, hipStream_t s);
// ... the equilibrated entries (a_ij / r_i) / c_j in FP32, rows perm[i] (or
// i when d_perm is null), with max |equilibrated entry| in FP64; the values
// are those equilibrate_complex would produce, rounded once to FP32 ...
int convert_equilibrated_to_f32(const double2 *dA, float2 *dst, const int *d_perm, const double *d_row,
                                const double *d_col, int n, double *max_scaled, LUContext &ctx);
int convert_equilibrated_to_f32(const MnaOperands &m, float2 *dst, const int *d_perm, const double *d_row,
                                const double *d_col, double *max_scaled, LUContext &ctx);
// ... and the equilibrated FP64 matrix when it is needed after all.
void copy_equilibrated(const double2 *src, double2 *dst, const double *d_row, const double *d_col, int n,
                       hipStream_t s);

// Factor in place; d_ipiv[k] = row swapped with k at step k (complex_lu.c:330-510).
template <class S>
LUStatus lu_factor(S *dA, int n, size_t lda, int *d_ipiv, double pivot_limit, LUContext &ctx,
                   int block_size = 0);

// Factor a complex FP32 matrix whose rows are already in pivot order
// (no pivot search, no row exchange between workgroups); status -2 when a
// pivot has |u_kk| <= pivot_limit. Used by the mixed-precision path with the
// row order of an earlier partial-pivoting factorization.
// A batch of `batch` equally sized matrices batch_stride elements apart
// (block widths 64 or 128) is factored in the same launches, one member per
// grid z index, with pivot limits d_limits[b] (device); the result is the
// first failing member's status, and member_status[b] (host, optional)
// receives each one's. A single factor (batch 1) ignores the batch arguments.
LUStatus lu_factor_ordered(float2 *dA, int n, size_t lda, double pivot_limit, LUContext &ctx,
                           int block_size = 0, int batch = 1, size_t batch_stride = 0,
                           const double *d_limits = nullptr, LUStatus *member_status = nullptr);

// Solve with the factors for `nrhs` right-hand sides stored rhs-major
// (X[r*n + i]); applies the row scale, pivots, L, U and column scale like
// complex_lu_solve (complex_lu.c:724-778). Scales may be null (real LU).
// The factor may be FP32 (float2): arithmetic is FP64 throughout.
template <class S, class V>
void lu_solve(const S *dLU, int n, size_t lda, const int *d_ipiv, const double *d_row,
              const double *d_col, V *dX, int nrhs, LUContext &ctx);

// Same with a precomputed row permutation (perm[i] = original row of
// position i after all swaps), see permutation_from_pivots. same_factor:
// dLU holds the same factor as in the previous call on this context (its
// diagonal-block inverses are reused).
template <class S, class V>
void lu_solve_permuted(const S *dLU, int n, size_t lda, const int *d_perm, const double *d_row,
                       const double *d_col, V *dX, int nrhs, LUContext &ctx, bool same_factor = false);
void permutation_from_pivots(const int *d_ipiv, int n, int *d_perm, hipStream_t s);

// One right-hand side per member of a batch of FP32 factors (a batched
// ordered factor: the members lu_stride elements apart with leading dimension
// n, all in the row order d_perm; member b's scales at d_row + b n and
// d_col + b n, its vector at dX + b n): X_b <- A_b^-1 X_b as
// lu_solve_permuted does, with every member's chained solves in the same
// cooperative launches (grid y = member). same_factors: the factors are
// those of the previous call on this context (block inverses reused).
void lu_solve_members(const float2 *dLU, size_t lu_stride, int n, const int *d_perm, const double *d_row,
                      const double *d_col, double2 *dX, int members, LUContext &ctx, bool same_factors);

// y = A x for a complex FP64 rows x cols matrix (leading dimension lda),
// and the reference's component-wise backward error terms per row:
//   out[i] = (|r_i|, sum_j |A_ij||x_j| + |b_i|, |b_i|) with r = A x - b.
// d_ax: cols doubles of workspace (|x_j|), or null to allocate them.
void complex_residual(const double2 *dA, int rows, int cols, size_t lda, const double2 *dx, const double2 *db,
                      double2 *dr, double *dterms, hipStream_t s, double *d_ax = nullptr);

// R = A X - B for nrhs right-hand sides (rhs-major) and each one's
// component-wise backward error (harmonic.c:932-969) as the bit pattern of
// a non-negative double in d_eta_bits[q]. d_ax: n * nrhs doubles of scratch.
void residual_multi(const double2 *dA, int n, size_t lda, const double2 *dX, const double2 *dB, double2 *dR,
                    int nrhs, double *d_ax, unsigned long long *d_eta_bits, hipStream_t s);
// The same with the MNA matrix of m evaluated from its blocks (the same
// entries, sums and result as with its dense copy), for fewer than
// kResidualDenseMinRhs right-hand sides; more take the dense matrix.
constexpr int kResidualDenseMinRhs = 16;
void residual_multi(const MnaOperands &m, const double2 *dX, const double2 *dB, double2 *dR, int nrhs, double *d_ax,
                    unsigned long long *d_eta_bits, hipStream_t s);

// In-bounds self-test of the MFMA lane layout (1024 doubles out).
void mfma_layout_probe(double *host_out);

}  // namespace peec

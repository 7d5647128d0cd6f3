// Device-resident assembly of the static Lp/P matrices (analytic 1/R panel
// integrals) and of their retarded complex counterparts.
#pragma once

#include <memory>
#include <vector>

#include "../../include/dual_mesh.h"
#include "../../include/inductance.h"
#include "../../include/mesh.h"
#include "../../include/potential.h"
#include "common.h"

namespace peec {

struct QuadratureOrders {
    int far_order = 4, near_order = 10, touching_order = 12, self_order = 8;
    double near_factor = 0.5;
};

enum class RegionKind { Node = 0, Edge = 1 };

// Work done by one assembly pass, for roofline accounting.
struct AssemblyWork {
    long long far_evaluations = 0;      // 1/R point evaluations in the tile kernel
    long long near_entries = 0;         // entries recomputed by the near-field kernel
    long long smooth_evaluations = 0;   // exp(-jkR) evaluations (retarded pass)
};

// One set of dual regions (all node regions for P, all edge regions for
// Lp) staged in HBM for the tile kernels.
class RegionSet {
  public:
    RegionSet(const Mesh *mesh, const DualMesh *dual, RegionKind kind, const QuadratureOrders &q,
              int smooth_order, double direction_tolerance, hipStream_t stream,
              bool count_work = false, int near_smooth_order = 4);

    int count() const { return R_; }
    RegionKind kind() const { return kind_; }

    // static matrix (row-major R x R, original numbering) into d_out
    // With d_counts, the near-field kernel adds its analytic potential
    // evaluations to d_counts[0] and its 1/R evaluations to d_counts[1].
    void assemble_static(double *d_out, hipStream_t s, int *d_error, AssemblyWork *work,
                         KernelTimer *timer = nullptr, unsigned long long *d_counts = nullptr);
    // complex retarded matrix: d_out = d_static + scale * smooth remainder
    void assemble_retarded(const double *d_static, double wave_number, double2 *d_out,
                           hipStream_t s, AssemblyWork *work, KernelTimer *timer = nullptr);
    // Frequency-independent moments of the smooth remainder for a sweep:
    // (exp(-jkR)-1)/R = sum_{m>=1} (-jk)^m R^(m-1) / m!, so the remainder of an
    // entry is sum_m (-jk)^m/m! * U_m with U_m = scale * sum_w w R^(m-1) over
    // the same quadrature points (smooth rule; near entries with the near
    // rule). build_moments stores U_1..U_terms (terms <= kMaxMomentTerms) in
    // 64 x 64 tiles of the upper triangle in original numbering; a point
    // whose series needs no more terms than stored is then one streaming pass
    // over the moments instead of a quadrature pass.
    static constexpr int kMaxMomentTerms = 56;
    void build_moments(int terms, hipStream_t s);
    int moment_terms() const { return moment_terms_; }
    static size_t moment_bytes(int regions, int terms);
    // terms of the series for wave number k: the tail beyond them is below
    // 1e-17 of every static entry (k * diameter bounds k R for all points)
    int series_terms(double wave_number) const;
    double diameter() const { return diameter_; }
    // d_out = d_static + remainder from the stored moments (terms <= moment_terms())
    void assemble_retarded_moments(const double *d_static, double wave_number, int terms, double2 *d_out,
                                   hipStream_t s, KernelTimer *timer = nullptr);
    // The same for `count` wave numbers at once (count <= kMaxMomentBatch),
    // d_out[b] receiving the matrix of wave_number[b] with terms[b] terms:
    // one pass over the moments and the static matrix for all of them, each
    // result bitwise equal to assemble_retarded_moments of its wave number.
    static constexpr int kMaxMomentBatch = 4;
    void assemble_retarded_moments_batch(const double *d_static, const double *wave_number, const int *terms,
                                         double2 *const *d_out, int count, hipStream_t s,
                                         KernelTimer *timer = nullptr);
    // Complex moments around a shifted wave number k0 (points above the
    // range of the moments at 0): V_0 = scale * sum_w (exp(-jk0R) - 1)/R and
    // V_m = scale * sum_w exp(-jk0R) R^(m-1), m = 1..terms-1, so that the
    // remainder at k is sum_m (-j(k - k0))^m / m! V_m.
    static constexpr int kMaxShiftTerms = 32;
    void build_shifted_moments(int terms, double k0, hipStream_t s);
    int shifted_terms() const { return shift_terms_; }
    size_t shifted_bytes() const { return d_shifted_.bytes(); }
    double shifted_k0() const { return shift_k0_; }
    // terms for wave number k around the stored k0 (> stored: not covered)
    int shifted_terms_for(double wave_number) const;
    void assemble_retarded_shifted(const double *d_static, double wave_number, int terms, double2 *d_out,
                                   hipStream_t s, KernelTimer *timer = nullptr);
    // single entries through the near-field kernel (drop-in *_compute_pair)
    void assemble_pairs(const std::vector<int> &pairs, double *d_values, hipStream_t s,
                        int *d_error);

  private:
    void build_points(const std::vector<double> &patch_vertices, const std::vector<int> &ptr,
                      int order, std::vector<double> &pi, std::vector<double> &pj, int &slots,
                      std::vector<int> &npts);
    void build_near_list(const Mesh *mesh);

    RegionKind kind_;
    QuadratureOrders q_;
    int smooth_order_;
    int near_smooth_order_;
    double dir_tol_;
    int R_ = 0;
    int maxp_ = 0;
    int far_slots_ = 0, smooth_slots_ = 0;
    std::vector<int> h_ptr_;
    std::vector<double> h_pv_, h_aux_, h_dir_;
    std::vector<int> h_perm_;           // kernel order -> original index
    std::vector<int> h_near_;           // candidate near entries (i <= j), original ids
    std::vector<int> h_tiles_;          // upper-triangle tile list (bi, bj)
    DeviceBuffer<int> d_ptr_, d_perm_, d_near_, d_tiles_, d_far_n_, d_smooth_n_;
    DeviceBuffer<double> d_pv_, d_aux_, d_dir_, d_aux_k_, d_dir_k_;
    DeviceBuffer<int> d_smooth_err_;
    DeviceBuffer<double4> d_far_i_, d_far_j_, d_smooth_i_, d_smooth_j_;
    long long far_evals_ = 0, smooth_evals_ = 0;
    double diameter_ = 0.0;
    int moment_terms_ = 0, moment_tiles_ = 0;
    DeviceBuffer<double> d_moments_;
    int shift_terms_ = 0;
    double shift_k0_ = 0.0;
    DeviceBuffer<double> d_shifted_;
    DeviceBuffer<int> d_moment_tiles_;

  public:
    long long far_evaluations() const { return far_evals_; }
    long long smooth_evaluations() const { return smooth_evals_; }
    const std::vector<int> &near_list() const { return h_near_; }
};

// Raw integral over all piece pairs of two piece lists (12 doubles each).
double region_pair_integral(const std::vector<double> &pa, const std::vector<double> &pb, bool same,
                            const QuadratureOrders &q);

// K(a, b) of two free pieces (NaN on failure), evaluated on the GPU.
double patch_pair_integral(const QuadPatch *a, const QuadPatch *b, const QuadratureOrders &q);

// Upload the Gauss-Legendre table to __constant__ memory (once per process).
void ensure_device_tables();

}  // namespace peec

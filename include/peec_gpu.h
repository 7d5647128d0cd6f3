/* Device-resident PEEC frequency-point engine for one MI355X (gfx950).
 *
 * No counterpart in the reference: its drivers (main.c:662-900 assembly,
 * main.c:1249-1568 harmonic solve) call the drop-in functions one by one
 * through host memory. This engine keeps the mesh's static Lp/P, the MNA
 * matrix and its factors in HBM and runs every stage of a frequency point on
 * the GPU:
 *
 *   retarded Lp(w), P(w) (optional) -> MNA fill -> equilibration ->
 *   LU (FP64, or FP32 MFMA factor + FP64 iterative refinement) ->
 *   port right-hand sides -> solves -> component-wise backward error
 *   (harmonic.c:932-969).
 *
 * Ports follow the reference's harmonic excitation (harmonic.h:118-126,
 * harmonic.c:503-549): port k injects +1 A at node port_nodes[2k] and draws
 * it at port_nodes[2k+1]; Z_in[k] = V_a - V_b of its own solution.
 * Calls on one engine are not thread-safe; distinct engines are. */
#ifndef PEEC_GPU_H
#define PEEC_GPU_H

#include <stddef.h>

#include "complex_matrix.h"
#include "dual_mesh.h"
#include "inductance.h"
#include "mesh.h"
#include "potential.h"
#include "incident_field.h"
#include "lightning.h"
#include "rcs.h"
#include "scattering.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum
{
    PEEC_SOLVE_FP64 = 0,  /* FP64 LU: the drop-in / correctness path */
    PEEC_SOLVE_MIXED = 1  /* FP32 MFMA factor + FP64 iterative refinement */
} PeecSolveMode;

typedef struct
{
    InductanceOptions inductance;
    PotentialOptions potential;
    int retarded;               /* 0 quasi-static (reference model), 1 exp(-jkR)/R */
    int smooth_order;           /* GL order of the retarded remainder rule, default 2 */
    int near_smooth_order;      /* same, for entries of nearby regions, default 4 */
    int block_size;             /* LU panel width, 0 = automatic (the reused-row-order
                                   factor's width and update tile shape autotuned, see
                                   PeecPointReport); a given width uses 64-row tiles */
    int max_refinement_steps;   /* mixed mode cap, default 30 */
    double refinement_target;   /* mixed mode stops at this backward error, default 1e-12
                                   (a hundredth of the gate; the FP64 LU itself reaches
                                   ~1e-12 on systems of N ~ 10^4) */
    double backward_error_gate; /* a point is accepted at or below this, default 1e-10 */
    int reuse_pivots;           /* mixed mode: factor in the row order of the previous
                                   point (no pivot search); refactors with partial
                                   pivoting when refinement does not reach
                                   max(1e-3 x gate, 10 x refinement_target).
                                   Default 1 */
    int smooth_moments;         /* retarded sweeps: terms of the power series of the
                                   smooth remainder, (exp(-jkR)-1)/R = sum_m (-jk)^m
                                   R^(m-1)/m!, whose frequency-independent moments are
                                   assembled once with the engine (8 bytes per term per
                                   upper-triangle entry of Lp and P). A point whose
                                   series needs at most that many terms for 1e-17 of
                                   every entry (k * mesh diameter small enough) is
                                   assembled from them in one streaming pass, the
                                   others by quadrature. 0 = off, 1..56 = count,
                                   -1 (default) = 56 terms (k * diameter up to ~10.5)
                                   when they fit in a quarter of the device memory
                                   left free by the engine's per-point working set,
                                   else fewer (off below 8) */
    int shifted_moments;        /* retarded points above the range of smooth_moments:
                                   complex moments around a shifted wave number k0,
                                   (exp(-jkR)-1)/R = (exp(-jk0R)-1)/R + sum_{m>=1}
                                   (-j(k-k0))^m/m! exp(-jk0R) R^(m-1), built when a point
                                   needs them (k0 = k + ~3.2/diameter, so a sweep upwards
                                   reuses them) and kept while they cover the points,
                                   16 bytes per term per upper-triangle entry. A build
                                   costs about 8 quadrature passes and covers a window
                                   of 6.4 / diameter in k: worth it for dense sweeps
                                   (tens of points per window), not for a few points.
                                   0 (default) = off, 1..32 = terms, -1 = 30 terms when
                                   they fit in half the free device memory */
    int split_fp16_update;      /* mixed mode: trailing updates of the reused-row-order
                                   factor on FP16 MFMA with each FP32 operand split into
                                   two FP16 parts (hi + lo, power-of-two scaled per row
                                   / column; three products, FP32 accumulation, ~2^-22
                                   per product). 0 = FP32 MFMA. Default 1 */
    int panel_lds_kib;          /* cap (KiB) on the LDS a workgroup of the partial-pivoting
                                   panel may hold: a block column whose panel slab does
                                   not fit is factored as narrower sub-panels (each with
                                   its U12 and update inside the block column), as tall
                                   matrices are; 0 (default) = the device's limit */
    int gmres_krylov;           /* mixed mode: when plain refinement stalls above the gate,
                                   GMRES-IR on the same factor first (each correction
                                   A d = r solved by GMRES on the system preconditioned
                                   with the FP32 factor, this Krylov dimension, up to 10
                                   outer steps), then the FP64 refactor. 0 = off,
                                   default 20 */
} PeecEngineOptions;

typedef struct
{
    double frequency;
    double backward_error;      /* worst over the point's right-hand sides */
    int refinement_steps;       /* mixed mode */
    int used_fp64_fallback;     /* mixed mode did not reach the gate and refactored in FP64 */
    int pivot_reuse;            /* mixed mode: the factor used the previous point's row order */
    int pivot_refactor;         /* mixed mode: that factor was rejected and redone with pivoting */
    double reuse_backward_error; /* backward error the rejected reused-order factor reached */
    double ms_assembly;         /* retarded Lp(w)/P(w) */
    double ms_fill;             /* MNA fill + equilibration */
    double ms_factor;
    double ms_solve;            /* right-hand sides, solves and refinement */
    double ms_check;            /* backward error */
    int factor_cached;          /* the engine still held this point's factor (same
                                   frequency and mode as its previous call): no
                                   assembly, fill or factor */
    int moment_terms;           /* retarded Lp(w) assembled from that many stored moment
                                   terms (0: by quadrature, or not retarded) */
    int block_size;             /* block width of a reused-row-order factor (0 otherwise);
                                   with block_size = 0 in the options the engine's
                                   autotuner tries the shapes (block width 128 / 64 x
                                   update tiles of 64 / 32 rows) on its first such
                                   factors, single ones and batches apart, and keeps the
                                   fastest per factor */
    double moment_center;       /* wave number k0 of the shifted-center moments that
                                   gave this point's retarded matrices (0 otherwise) */
    int batch_size;             /* points whose factors shared this point's launches
                                   (peec_engine_solve_batch; 1 or 0 otherwise) */
    int update_tile_rows;       /* rows of the split-FP16 update tiles of that factor
                                   (64 or 32; 0 when no reused-row-order factor ran) */
    int gmres_steps;            /* outer GMRES-IR steps (0: plain refinement sufficed) */
    int gmres_iterations;       /* GMRES iterations over all of them */
} PeecPointReport;

typedef struct PeecEngine PeecEngine;

void peec_engine_options_default(PeecEngineOptions *options);

/* Builds the device copies of the mesh regions and assembles the static
 * Lp and P (FP64). 0 on success, -1 on bad input or GPU failure. */
int peec_engine_create(const Mesh *mesh, const DualMesh *dual,
                       const PeecEngineOptions *options, PeecEngine **engine);

/* Same, with the static Lp (Ne x Ne) and P (Nv x Nv) given by the caller,
 * row-major on the host (e.g. from matrix_import_dense_mtx): no assembly. */
int peec_engine_create_with_matrices(const Mesh *mesh, const DualMesh *dual,
                                     const PeecEngineOptions *options, const double *L,
                                     const double *P, PeecEngine **engine);

/* Solves one frequency point for n_ports port excitations on one factor.
 * A call at the frequency and mode of the engine's previous successful call
 * reuses the factor it still holds (report->factor_cached).
 * z_in: n_ports values; edge_current / node_voltage: optional, rhs-major
 * (n_ports x Ne, n_ports x Nv). Returns 0, -1 (bad input / GPU failure),
 * -2 (singular) or -3 (backward error above the gate). */
int peec_engine_solve(PeecEngine *engine, double frequency, int mode,
                      const int *port_nodes, size_t n_ports, Complex *z_in,
                      Complex *edge_current, Complex *node_voltage,
                      PeecPointReport *report);

/* Mixed-precision points of this engine's mesh solved together: point p at
 * frequency[p] with n_ports[p] port pairs ports[p] (as peec_engine_solve)
 * and impedances z_in[p]. Every point's retarded blocks, equilibration and
 * FP32 matrix are formed in the engine's reused row order, then one ordered
 * factor runs over all of them in shared launches (the members of a launch
 * are its grid's z index), then each point is solved and refined on its own
 * factor with the single point's fallbacks (points of one port each are
 * refined in lock-step, their triangular solves in shared launches; a
 * member's ms_factor / ms_solve are its share of the batch's). Points met
 * before the engine has a row order (its first mixed point) are solved alone
 * first. The factors and solves are those of single points in the same row
 * order and block width, so the results equal single calls'. status[p] gets each point's code (as
 * peec_engine_solve); returns 0 or the first non-zero one. */
int peec_engine_solve_batch(PeecEngine *engine, size_t n_points, const double *frequency,
                            const int *const *ports, const size_t *n_ports, Complex *const *z_in,
                            PeecPointReport *reports, int *status);

/* Monostatic radar cross section on one factor (the workload of
 * main_rcs.c:1380-1500 for many incidences at once). For incidence i the
 * plane wave arrives from (theta_deg[i], phi_deg[i]) and excites the edges as
 * scattering_build_harmonic_excitation does; the right-hand side is
 * [-U ; 0]. All incidences are solved on one factor (mixed mode refines each
 * right-hand side); results[i] is rcs_compute_direction of that solution in
 * the backscatter direction -k_hat (angles in results[i].theta_deg/phi_deg).
 * edge_current: optional, n_angles x Ne, rhs-major. Returns as
 * peec_engine_solve. */
typedef struct
{
    IncidentWaveType wave_type;           /* COMPLEX (default), COS or SIN */
    IncidentPolarization polarization;    /* default VERTICAL (e_theta) */
    double E_amp;                         /* V/m, default 1 */
    double phase;                         /* rad, default 0 */
    ScatteringEdgeIntegration excitation; /* default ANALYTIC_PLANE_WAVE */
    RcsOptions rcs;                       /* default DUAL_PATCH_QUADRATURE, order 6 */
} PeecRcsOptions;

void peec_rcs_options_default(PeecRcsOptions *options);

int peec_engine_monostatic_rcs(PeecEngine *engine, double frequency, int mode,
                               const PeecRcsOptions *options, const double *theta_deg,
                               const double *phi_deg, size_t n_angles, RcsResult *results,
                               Complex *edge_current, PeecPointReport *report);

/* Port impedance sweep: z_in[k] = Z_in(frequencies[k]) for one port
 * (port_nodes[0] -> port_nodes[1]); reports[k] optional. Mixed mode reuses
 * the row order along the sweep. Stops at the first failing point and
 * returns its status. */
int peec_engine_port_sweep(PeecEngine *engine, const double *frequencies, size_t n, int mode,
                           const int port_nodes[2], Complex *z_in, PeecPointReport *reports);

/* Fourier transform of the lightning current, I(f) = int i(t) e^{-j 2 pi f t} dt
 * = K e^{-j w delay} (1 / (alpha + j w) - 1 / (beta + j w)). */
int peec_lightning_spectrum(const LightningSource *source, double frequency, Complex *current);

/* Time response of a port to the lightning current by frequency-domain
 * superposition: V(f_k) = Z_in(f_k) I(f_k) at f_k = k df, k = 1..n_freq
 * (a power of two), Z_in(0) extrapolated linearly from the two lowest
 * points (f > 0 is required by the MNA system), and a real inverse FFT:
 * 2 n_freq samples t_i = i / (2 n_freq df); voltage[i] = v(t_i) [V],
 * current[i] = lightning_current(t_i). The window 1 / df must exceed the
 * response's decay (wrap-around otherwise). Host-only. */
int peec_port_transient_from_spectrum(const LightningSource *source, size_t n_freq, double df,
                                      const Complex *z_in, double *time, double *current,
                                      double *voltage);

/* The frequency sweep on the engine (strike -> return node) followed by
 * peec_port_transient_from_spectrum: the time-domain lightning workload.
 * time/current/voltage: 2 n_freq values; z_in (n_freq) and reports (n_freq)
 * optional. */
int peec_engine_lightning_response(PeecEngine *engine, const LightningSource *source, size_t n_freq,
                                   double df, int mode, double *time, double *current, double *voltage,
                                   Complex *z_in, PeecPointReport *reports);

/* Copies the static matrices (Ne x Ne, Nv x Nv, row-major) to the host. */
int peec_engine_static_matrices(const PeecEngine *engine, double *L, double *P);

/* Copies the retarded matrices at wave number k to the host. */
int peec_engine_retarded_matrices(PeecEngine *engine, double wave_number,
                                  Complex *L, Complex *P);

size_t peec_engine_edges(const PeecEngine *engine);
size_t peec_engine_nodes(const PeecEngine *engine);

/* FLOP and evaluation counts of the static assembly (roofline accounting). */
long long peec_engine_far_evaluations(const PeecEngine *engine);
long long peec_engine_smooth_evaluations(const PeecEngine *engine);

/* The static Lp/P assembly of the engine's construction: evaluation counts
 * and the kernels' times (HIP events on the engine stream). Work in flop
 * terms: 12 flops per 1/R evaluation (far tiles, and near pairs far enough
 * for the plain Gauss rule), 140 per analytic triangle-potential
 * evaluation. All zero when the static matrices were supplied. */
typedef struct
{
    long long far_evaluations;            /* far_tile_kernel 1/R points */
    long long near_potential_evaluations; /* near_entry_kernel analytic potentials */
    long long near_pair_evaluations;      /* near_entry_kernel 1/R points (Gauss pairs) */
    double far_ms;                        /* far_tile_kernel, Lp and P */
    double near_ms;                       /* near_entry_kernel, Lp and P */
} PeecStaticStats;
int peec_engine_static_stats(const PeecEngine *engine, PeecStaticStats *out);

/* Per-kernel timing: HIP events recorded on the engine's stream around
 * every launch of the kernel family, summed over launches. flops/bytes are
 * the algorithmic counts of those launches. */
typedef enum
{
    PEEC_KERNEL_LU_UPDATE = 0,         /* update GEMMs on MFMA (FP64, FP32, and the split-FP16
                                          next-column update): 8mnk (complex) */
    PEEC_KERNEL_PANEL = 1,             /* panel / diagonal-block factorization */
    PEEC_KERNEL_SMOOTH_TILE = 2,       /* retarded remainder tile kernel: 38 flops/eval */
    PEEC_KERNEL_FAR_TILE = 3,          /* static 1/R tile kernel: 12 flops/eval */
    PEEC_KERNEL_SMOOTH_MOMENTS = 4,    /* retarded remainder from stored moments */
    PEEC_KERNEL_LU_UPDATE_SPLIT16 = 5, /* split-FP16 trailing update of the ordered factor */
    PEEC_KERNEL_LU_UPDATE_THIN = 6,    /* split-FP16 updates of the next panel's / pair's rows and of
                                          the block columns after a deferred pair */
    PEEC_KERNEL_NEAR_ENTRY = 7,        /* static near-field entries (timed at construction) */
    PEEC_KERNEL_LU_UPDATE_MERGED = 8   /* split-FP16 trailing updates of two panel pairs in one pass */
} PeecKernel;

typedef struct
{
    long long launches;
    double total_ms;
    double flops;
    double bytes;
} PeecKernelStats;

/* Enables (1) or disables (0) the timing and clears the totals. */
int peec_engine_set_profiling(PeecEngine *engine, int enable);
int peec_engine_kernel_stats(const PeecEngine *engine, int kernel, PeecKernelStats *stats);

void peec_engine_destroy(PeecEngine *engine);

/* No counterpart in the reference. Visible HIP devices (0 without a GPU),
 * and a device-wide synchronisation of the current device (the bracket a
 * benchmark puts around its timed region), so that callers need no second
 * GPU runtime in the process. 0 on success, -1 on a HIP error. */
int peec_device_count(void);
int peec_device_synchronize(void);

/* Library self-test: runs the 16x16x4 FP64 MFMA on known operands and
 * returns every lane's accumulators (FP64 then FP32, 1024 doubles) so the operand layout the
 * GEMM kernels assume can be checked. */
int peec_selftest_mfma_layout(double *out1024);

#ifdef __cplusplus
}
#endif

#endif

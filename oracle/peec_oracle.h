/* CPU oracle for the PEEC hot path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain C restatement of the reference algorithm (quadrature, Lp, P, PAT,
 * MNA, complex/real LU, backward error) used by tests/, the smoke check and
 * bench.py's CPU-baseline leg as the checker. The product library never
 * links, loads or calls anything here.
 *
 * Every symbol is prefixed oracle_ so it can be loaded next to the product
 * library and the compiled reference without symbol clashes.
 *
 * Pinning (how each function is checked, see tests/test_oracle.py):
 *  - everything on the quasi-static path is compared bit-for-bit against the
 *    compiled reference (oracle/_ref) and against committed golden vectors,
 *    plus the closed forms of the reference's own tests (rectangle self
 *    integral, touching partition, near/far bounds, LU known answers,
 *    analytic 3-edge loop);
 *  - oracle_retarded_matrices has no reference counterpart: parity
 *    unpinned. It is pinned by the k -> 0 limit (static matrices), the exact
 *    O(k) imaginary term -k*S_a*S_b, reciprocity, and a brute-force
 *    high-order integration of tiny panel pairs. */
#ifndef PEEC_ORACLE_H
#define PEEC_ORACLE_H

#include <stddef.h>

#include "../include/complex_matrix.h"
#include "../include/dual_mesh.h"
#include "../include/inductance.h"
#include "../include/mesh.h"
#include "../include/potential.h"

#ifdef __cplusplus
extern "C" {
/* ---- plane-wave excitation and RCS (physics/incident_field.c,
 *      physics/scattering.c, physics/rcs.c) ---- */
#include "../include/incident_field.h"
#include "../include/rcs.h"

/* incident_field.c:28-283,312-420: harmonic set-up (COS/SIN/COMPLEX) */
int oracle_plane_wave(IncidentField *field, int wave_type, int polarization, double E_amp, double frequency,
                      double theta_deg, double phi_deg, double phase);
/* incident_field.c:754-858 */
int oracle_phasor(const IncidentField *field, const double r[3], Complex E[3]);
/* scattering.c:23-43,180-337,338-489: integration 0 midpoint, 1 analytic */
int oracle_excitation(const Mesh *mesh, const IncidentField *field, int integration, Complex *U);
/* rcs.c:117-531,701-995 */
int oracle_radiation_vector(const Mesh *mesh, const DualMesh *dual, const Complex *I, double k,
                            const double direction[3], int integration, int order, Complex F[3]);
/* rcs.c:536-598,638-700,1000-1165 */
int oracle_rcs_direction(const Mesh *mesh, const DualMesh *dual, const IncidentField *field, const Complex *I,
                         int integration, int order, double theta_deg, double phi_deg, RcsResult *result);

#endif

/* Work counters filled by the matrix routines (may be NULL). */
typedef struct
{
    long long far_evaluations;      /* 1/R point evaluations in FAR rules */
    long long potential_calls;      /* triangle_potential_at calls */
    long long near_triangle_pairs;  /* triangle pairs on the analytic path */
    long long far_patch_pairs;
    long long near_patch_pairs;
} OracleCounters;

int oracle_gauss_legendre(int order, double *x, double *w);

int oracle_load_gmsh(const char *path, Mesh *mesh);
int oracle_build_topology(Mesh *mesh);
int oracle_build_dual(const Mesh *mesh, DualMesh *dual);
void oracle_free_mesh(Mesh *mesh);
void oracle_free_dual(DualMesh *dual);

int oracle_classify(const QuadPatch *a, const QuadPatch *b, double near_factor);
double oracle_patch_kernel(const QuadPatch *a, const QuadPatch *b, int far_order,
                           int near_order, int touching_order, int self_order,
                           double near_factor);

int oracle_inductance_matrix(const Mesh *mesh, const DualMesh *dual,
                             const InductanceOptions *opt, double *L,
                             OracleCounters *counters);
int oracle_potential_matrix(const Mesh *mesh, const DualMesh *dual,
                            const PotentialOptions *opt, double *P,
                            OracleCounters *counters);
int oracle_pat(const Mesh *mesh, const double *P, double *PAT);

/* Z = [[R + jwL, A], [-PAT, jwI]], complex Lc/Pc optional (retarded form):
 * when Lc != NULL it replaces L, and Pc replaces P inside PAT. */
int oracle_mna(const Mesh *mesh, const double *L, const double *R_diag,
               const double *PAT, const Complex *Lc, const Complex *Pc,
               double frequency, Complex *Z);

int oracle_complex_lu(const Complex *A, size_t n, double pivot_tolerance,
                      Complex *LU, size_t *pivots, double *row_scale,
                      double *column_scale, double *matrix_scale);
int oracle_complex_solve(const Complex *LU, const size_t *pivots,
                         const double *row_scale, const double *column_scale,
                         size_t n, const Complex *b, Complex *x);
int oracle_real_lu(const double *A, size_t n, double pivot_tolerance,
                   double *LU, size_t *pivots, double *matrix_scale);
int oracle_real_solve(const double *LU, const size_t *pivots, size_t n,
                      const double *b, double *x);

/* out[0..6] = absolute, rhs_relative, backward, edge_abs, node_abs,
 * edge_backward, node_backward (HarmonicResidual order). */
int oracle_backward_error(const Complex *Z, const Complex *x, const Complex *b,
                          size_t n, size_t ne, double *out);

/* rhs = [-U_e ; P I_s]; P real (Pr) or complex (Pc). */
int oracle_rhs(size_t ne, size_t nv, const double *Pr, const Complex *Pc,
               const Complex *edge_voltage, const Complex *node_current,
               Complex *rhs);

/* Retarded Lp(k), P(k): static analytic integral + GL rule of
 * (exp(-jkR)-1)/R over every triangle pair; entries of regions closer than
 * near_factor x largest piece side use near_smooth_order (parity unpinned). */
int oracle_retarded_matrices(const Mesh *mesh, const DualMesh *dual,
                             const InductanceOptions *lopt,
                             const PotentialOptions *popt, double wave_number,
                             int smooth_order, int near_smooth_order, const double *L_static,
                             const double *P_static, Complex *L, Complex *P);

#ifdef __cplusplus
}
/* ---- plane-wave excitation and RCS (physics/incident_field.c,
 *      physics/scattering.c, physics/rcs.c) ---- */
#include "../include/incident_field.h"
#include "../include/rcs.h"

/* incident_field.c:28-283,312-420: harmonic set-up (COS/SIN/COMPLEX) */
int oracle_plane_wave(IncidentField *field, int wave_type, int polarization, double E_amp, double frequency,
                      double theta_deg, double phi_deg, double phase);
/* incident_field.c:754-858 */
int oracle_phasor(const IncidentField *field, const double r[3], Complex E[3]);
/* scattering.c:23-43,180-337,338-489: integration 0 midpoint, 1 analytic */
int oracle_excitation(const Mesh *mesh, const IncidentField *field, int integration, Complex *U);
/* rcs.c:117-531,701-995 */
int oracle_radiation_vector(const Mesh *mesh, const DualMesh *dual, const Complex *I, double k,
                            const double direction[3], int integration, int order, Complex F[3]);
/* rcs.c:536-598,638-700,1000-1165 */
int oracle_rcs_direction(const Mesh *mesh, const DualMesh *dual, const IncidentField *field, const Complex *I,
                         int integration, int order, double theta_deg, double phi_deg, RcsResult *result);

#endif

/* ---- plane-wave excitation and RCS (physics/incident_field.c,
 *      physics/scattering.c, physics/rcs.c) ---- */
#include "../include/incident_field.h"
#include "../include/rcs.h"

/* incident_field.c:28-283,312-420: harmonic set-up (COS/SIN/COMPLEX) */
int oracle_plane_wave(IncidentField *field, int wave_type, int polarization, double E_amp, double frequency,
                      double theta_deg, double phi_deg, double phase);
/* incident_field.c:754-858 */
int oracle_phasor(const IncidentField *field, const double r[3], Complex E[3]);
/* scattering.c:23-43,180-337,338-489: integration 0 midpoint, 1 analytic */
int oracle_excitation(const Mesh *mesh, const IncidentField *field, int integration, Complex *U);
/* rcs.c:117-531,701-995 */
int oracle_radiation_vector(const Mesh *mesh, const DualMesh *dual, const Complex *I, double k,
                            const double direction[3], int integration, int order, Complex F[3]);
/* rcs.c:536-598,638-700,1000-1165 */
int oracle_rcs_direction(const Mesh *mesh, const DualMesh *dual, const IncidentField *field, const Complex *I,
                         int integration, int order, double theta_deg, double phi_deg, RcsResult *result);

/* ---- theta-scheme transient (solver/transient.c stages 1 and 2) ---- */
/* transient.c:334-383, incidence.c:593-720, potential.c (self capacitance),
 * transient.c:940-1018 (shunt operators, shunt_b < 0: stage 1) and
 * transient.c:1020-1419: the nt x nt system matrix (nt = ne + nv, + 1 with a
 * shunt) of the scheme (1 backward Euler, 2 trapezoidal); F (nv), AS
 * (ne x nv), b_R and c (nv, stage 2) are returned when non-null. A is the
 * dense ne x nv incidence matrix. -1 on invalid input. */
int oracle_transient_matrix(const double *L, const double *R, const double *A, const double *P, size_t ne,
                            size_t nv, int scheme, double dt, long shunt_a, long shunt_b, double r_shunt,
                            double *M, double *F, double *AS, double *bR, double *c);
/* transient.c:1421-1749: the right-hand side of the step n -> n+1 (optional
 * sources NULL); with bR non-null the stage-2 shunt terms are added and
 * rhs[ne + nv] = 0. */
int oracle_transient_rhs(const double *L, const double *R, const double *A, const double *F, const double *AS,
                         size_t ne, size_t nv, int scheme, double dt, const double *I, const double *V,
                         const double *sn, const double *snp1, const double *un, const double *unp1,
                         const double *bR, double ishunt, double *rhs);

#endif

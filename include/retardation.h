/* Retarded (full-wave) partial elements. The reference ships this header
 * empty (include/retardation.h, src/physics/retardation.c: 0 bytes; the
 * driver rejects --physics retarded, main.c:7094-7101), so the interface
 * below is new. It follows the *_compute_matrix conventions of
 * inductance.h / potential.h: the caller owns the output structs, the
 * function allocates their data, 0 / -1 return codes.
 *
 * Kernel: exp(-jkR)/R = 1/R + (exp(-jkR) - 1)/R, time convention e^{+jwt}.
 * The 1/R part is the reference's analytic/adaptive integral (identical to
 * inductance_compute_matrix / potential_compute_matrix); the bounded smooth
 * remainder, which tends to -jk as R -> 0, is integrated by the
 * Gauss-Legendre product rule of order smooth_order on every triangle pair;
 * entries whose regions come closer (bounding boxes) than near_factor x the
 * largest piece side use near_smooth_order, because the remainder's kink at
 * coincident points (Re ~ -k^2 R / 2) converges slowly under the low rule.
 *
 *   Lp_ab(k) = 1e-7 (e_a.e_b)/(w_a w_b) * int int exp(-jkR)/R
 *   P_ij(k)  = 1/(4 pi eps0 S_i S_j)    * int int exp(-jkR)/R           */
#ifndef PEEC_RETARDATION_H
#define PEEC_RETARDATION_H

#include "complex_matrix.h"
#include "dual_mesh.h"
#include "inductance.h"
#include "mesh.h"
#include "potential.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
    InductanceOptions inductance; /* static part, reference defaults */
    PotentialOptions potential;
    int smooth_order;             /* GL order of the remainder rule, default 2 */
    int near_smooth_order;        /* order for entries of nearby regions, default 4 */
} RetardationOptions;

void retardation_options_default(RetardationOptions *options, int parallel_threads);

/* Complex Ne x Ne Lp(k) and Nv x Nv P(k), k = w/c0 in 1/m (k = 0 gives the
 * static matrices). Either output may be NULL. */
int retardation_compute_matrices(const Mesh *mesh, const DualMesh *dual,
                                 double wave_number,
                                 const RetardationOptions *options,
                                 ComplexMatrix *L, ComplexMatrix *P);

#ifdef __cplusplus
}
#endif

#endif

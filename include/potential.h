/* Potential-coefficient matrix P over mesh nodes, drop-in for the
 * reference's include/potential.h:15-241. */
#ifndef PEEC_POTENTIAL_H
#define PEEC_POTENTIAL_H

#include "dual_mesh.h"
#include "mesh.h"

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
    int far_order;
    int near_order;
    int touching_order;
    int self_order;
    double near_factor;
    int parallel_threads;
} PotentialOptions;

typedef struct
{
    size_t n;
    double *data; /* n*n row-major, 1/F */
} PotentialMatrix;

/* potential.h:84-87 (defaults potential.c:38-58) */
void potential_options_default(PotentialOptions *options, int parallel_threads);
/* potential.h:93,99 */
void potential_init(PotentialMatrix *P);
void potential_free(PotentialMatrix *P);
/* potential.h:128-135 (potential.c:139-253); one entry, on the GPU */
int potential_compute_pair(const Mesh *mesh, const DualMesh *dual,
                           size_t node_i, size_t node_j,
                           const PotentialOptions *options, double *value);
/* potential.h:151-156 (potential.c:261-395); whole matrix on the GPU */
int potential_compute_matrix(const Mesh *mesh, const DualMesh *dual,
                             const PotentialOptions *options,
                             PotentialMatrix *P);
/* potential.h:162 */
void potential_print_info(const PotentialMatrix *P);
/* potential.h:185-188 (potential.c): F[j] = 1 / P_jj, the self capacitances
 * (-1 for a non-positive or non-finite diagonal) */
int potential_build_self_capacitance(const PotentialMatrix *P, double *F);
/* potential.h:204-208: charge[j] = V_c[j] / P_jj */
int potential_self_voltage_to_charge(const PotentialMatrix *P, const double *V_c,
                                     double *charge);
/* potential.h:234-238: phi = P diag(1/P_aa) V_c, the full node potentials of
 * the self voltages (Nv x Nv product on the GPU) */
int potential_self_voltage_to_full_potential(const PotentialMatrix *P,
                                             const double *V_c, double *phi);

#ifdef __cplusplus
}
#endif

#endif

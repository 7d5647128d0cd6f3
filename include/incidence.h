/* Edge-node incidence and P*A^T, drop-in for the reference's
 * include/incidence.h:48-297. */
#ifndef PEEC_INCIDENCE_H
#define PEEC_INCIDENCE_H

#include "matrix.h"
#include "mesh.h"
#include "potential.h"

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
    size_t rows; /* Ne */
    size_t cols; /* Nv */
    double *data; /* dense row-major, A[e,a] = -1, A[e,b] = +1 */
} IncidenceMatrix;

/* incidence.h:63-95 (incidence.c:34-115) */
void incidence_init(IncidenceMatrix *A);
void incidence_free(IncidenceMatrix *A);
int incidence_build(const Mesh *mesh, IncidenceMatrix *A);
/* incidence.h:113: max over rows of |sum_j A[e,j]| (0 for a valid A; -1 for
 * a null matrix) */
double incidence_check_row_sum(const IncidenceMatrix *A);
/* incidence.h:123 */
void incidence_print_info(const IncidenceMatrix *A);
/* incidence.h:171-176 (incidence.c:242-330): K = A P A^T, Ne x Ne,
 * K[e,f] = P[a,c] - P[a,d] - P[b,c] + P[b,d] for e = (a,b), f = (c,d);
 * computed on the GPU */
int incidence_build_APAT(const Mesh *mesh, const PotentialMatrix *P, Matrix *K,
                         int parallel_threads);
/* incidence.h:240-245 (incidence.c:417-590): PAT[j,e] = P[j,b_e] - P[j,a_e],
 * Nv x Ne, computed on the GPU */
int incidence_build_PAT(const Mesh *mesh, const PotentialMatrix *P,
                        Matrix *PAT, int parallel_threads);
/* incidence.h:292-297 (incidence.c:593-720): AS = A S, Ne x Nv, with
 * S_ja = P_ja / P_aa: AS[e,a] = (P[b_e,a] - P[a_e,a]) / P_aa, the edge end
 * points read from the rows of A (exactly one -1 and one +1 each, else -1);
 * computed on the GPU */
int incidence_build_AS(const IncidenceMatrix *A, const PotentialMatrix *P,
                       Matrix *AS, int parallel_threads);

#ifdef __cplusplus
}
#endif

#endif

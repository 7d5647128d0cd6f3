/* Partial-inductance matrix Lp over all mesh edges, drop-in for the
 * reference's include/inductance.h:15-183. The matrix is assembled on the
 * GPU (FP64) and returned in host memory the caller frees with
 * inductance_free(), exactly as the reference's calloc'd matrix. */
#ifndef PEEC_INDUCTANCE_H
#define PEEC_INDUCTANCE_H

#include "dual_mesh.h"
#include "mesh.h"

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
    int far_order;               /* default 4 */
    int near_order;              /* default 10 */
    int touching_order;          /* default 12 */
    int self_order;              /* default 8 */
    double near_factor;          /* default 0.5 */
    int parallel_threads;        /* host threads; the GPU ignores it */
    double direction_tolerance;  /* |e_a.e_b| below this => L_ab = 0 */
} InductanceOptions;

typedef struct
{
    size_t n;
    double *data; /* n*n row-major, henry */
} InductanceMatrix;

/* inductance.h:95-98 (defaults inductance.c:46-97) */
void inductance_options_default(InductanceOptions *options, int parallel_threads);
/* inductance.h:106 */
void inductance_init(InductanceMatrix *L);
/* inductance.h:110 */
void inductance_free(InductanceMatrix *L);
/* inductance.h:142-149 (inductance.c:221-382); one entry, on the GPU */
int inductance_compute_pair(const Mesh *mesh, const DualMesh *dual,
                            size_t edge_a, size_t edge_b,
                            const InductanceOptions *options, double *value);
/* inductance.h:165-170 (inductance.c:390-560); whole matrix on the GPU.
 * 0 on success, -1 on bad input or a non-finite entry (output freed). */
int inductance_compute_matrix(const Mesh *mesh, const DualMesh *dual,
                              const InductanceOptions *options,
                              InductanceMatrix *L);
/* inductance.h:178 */
void inductance_print_info(const InductanceMatrix *L);

#ifdef __cplusplus
}
#endif

#endif

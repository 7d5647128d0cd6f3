/* Frequency-domain PEEC MNA system, drop-in for the reference's
 * include/harmonic.h:50-373.
 *
 *   [ R + jwL    A    ] [I]   [ -U_e  ]
 *   [ -P A^T    jw I  ] [V] = [ P I_s ]      (time convention e^{+jwt})
 *
 * The matrix is filled, factored, solved and checked on the GPU; the
 * reference's host fields are filled in too. */
#ifndef PEEC_HARMONIC_H
#define PEEC_HARMONIC_H

#include "complex_lu.h"
#include "complex_matrix.h"
#include "incidence.h"
#include "inductance.h"
#include "matrix.h"
#include "potential.h"
#include "resistance.h"

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
    size_t n_edges;
    size_t n_nodes;
    size_t system_size;
    double frequency;
    double omega;
    ComplexMatrix matrix;
    ComplexLUFactorization lu;
    int matrix_built;
    int factorized;
} HarmonicSystem;

typedef struct
{
    size_t n_edges;
    size_t n_nodes;
    const Complex *edge_voltage; /* U_e, may be NULL (zero) */
    const Complex *node_current; /* I_s, may be NULL (zero) */
} HarmonicExcitation;

typedef struct
{
    size_t n_edges;
    size_t n_nodes;
    Complex *edge_current;
    Complex *node_voltage;
} HarmonicSolution;

typedef struct
{
    double absolute_residual;
    double rhs_relative_residual;
    double backward_error;  /* max_i |r_i| / (sum_j |Z_ij||x_j| + |b_i|) */
    double edge_absolute_residual;
    double node_absolute_residual;
    double edge_backward_error;
    double node_backward_error;
} HarmonicResidual;

/* harmonic.h:252-267 */
void harmonic_system_init(HarmonicSystem *system);
void harmonic_system_free(HarmonicSystem *system);
void harmonic_solution_init(HarmonicSolution *solution);
void harmonic_solution_free(HarmonicSolution *solution);
/* harmonic.h:288 (harmonic.c:175-372) */
int harmonic_build_system(const InductanceMatrix *L, const ResistanceMatrix *R,
                          const IncidenceMatrix *A, const Matrix *PAT,
                          double frequency, int parallel_threads,
                          HarmonicSystem *system);
/* harmonic.h:304 (harmonic.c:380-423) */
int harmonic_factorize(HarmonicSystem *system, int parallel_threads);
/* harmonic.h:331 (harmonic.c:436-552) */
int harmonic_build_rhs(const PotentialMatrix *P,
                       const HarmonicExcitation *excitation, Complex *rhs,
                       int parallel_threads);
/* harmonic.h:344 (harmonic.c:560-691) */
int harmonic_solve(const HarmonicSystem *system, const PotentialMatrix *P,
                   const HarmonicExcitation *excitation,
                   HarmonicSolution *solution, int parallel_threads);
/* harmonic.h:358 (harmonic.c:720-1080) */
int harmonic_compute_residual(const HarmonicSystem *system,
                              const PotentialMatrix *P,
                              const HarmonicExcitation *excitation,
                              const HarmonicSolution *solution,
                              int parallel_threads, HarmonicResidual *residual);
/* No counterpart in the reference. harmonic_compute_residual with the same
 * arguments and results, for callers that cannot allocate (code under stream
 * capture, tight serving loops, harnesses that forbid hipMalloc): it makes no
 * device or host allocation. Its inputs and partial results are staged in a
 * fixed 32 MiB device array per GPU, so N = Ne + Nv is limited to about 1400
 * (-1 with an ERROR line above that). Calls that overlap on one GPU share
 * that array and race. */
int harmonic_compute_residual_noalloc(const HarmonicSystem *system,
                                      const PotentialMatrix *P,
                                      const HarmonicExcitation *excitation,
                                      const HarmonicSolution *solution,
                                      int parallel_threads,
                                      HarmonicResidual *residual);
/* harmonic.h:373 */
void harmonic_print_info(const HarmonicSystem *system);

#ifdef __cplusplus
}
#endif

#endif

/* Time-domain PEEC by the theta scheme, drop-in for the reference's
 * include/transient.h:1-490 (stages 1 and 2; same struct layout).
 *
 * Unknowns x = [I ; V_c] (stage 1) or [I ; V_c ; I_shunt] (stage 2), with
 * V_c the node self voltages (charge q = V_c / P_jj). Per step, with
 * theta = 1 (backward Euler) or 1/2 (trapezoidal):
 *
 *   [ L/dt + theta R     theta A S    ] [I]       [ L I_n/dt - (1-theta)(R I_n + A S V_n) ... ]
 *   [ -theta A^T         diag(F)/dt   ] [V_c]  =  [ F V_n/dt + (1-theta) A^T I_n + sources   ]
 *
 * F_j = 1 / P_jj, S_ja = P_ja / P_aa. Stage 2 adds the shunt row
 * b_R^T S V_c + R_shunt I_shunt = 0 and the column -theta b_R.
 *
 * The system matrix is filled on the GPU (and copied to the caller-visible
 * host array), factored by the real LU of lu.h on the GPU, and every step's
 * right-hand side (the L, A S and A^T products) and solve run on the GPU
 * from device copies of the operators kept with the system. The slot-cell
 * stage (aperture coupling) is outside this library; its fields keep their
 * place in the struct. */
#ifndef PEEC_TRANSIENT_H
#define PEEC_TRANSIENT_H

#include "incidence.h"
#include "inductance.h"
#include "internal_circuit.h"
#include "lu.h"
#include "matrix.h"
#include "potential.h"
#include "resistance.h"

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum
{
    TRANSIENT_STAGE_1 = 0,
    TRANSIENT_STAGE_2 = 1,
    TRANSIENT_SLOT_CELLS = 2
} TransientStage;

typedef enum
{
    TRANSIENT_BACKWARD_EULER = 1,
    TRANSIENT_TRAPEZOIDAL = 2
} TransientScheme;

typedef struct SlotCircuit SlotCircuit;

typedef struct
{
    TransientStage stage;
    TransientScheme scheme;
    size_t n_edges;
    size_t n_nodes;
    size_t n_total;
    double dt;
    /* the caller's PEEC matrices (not owned) */
    const InductanceMatrix *L;
    const ResistanceMatrix *R;
    const IncidenceMatrix *A;
    const PotentialMatrix *P;
    double *F;                      /* F_j = 1 / P_jj */
    Matrix AS;                      /* Ne x Nv, A S */
    /* stage 2 */
    const InternalCircuitShunt *shunt;
    double *b_R;                    /* b_R[A] = -1, b_R[B] = +1 */
    double *shunt_voltage_operator; /* c_a = (P[B,a] - P[A,a]) / P_aa */
    /* slot cells (not built by this library) */
    const SlotCircuit *slot_circuit;
    size_t n_slot_cells;
    double *slot_mass_matrix;
    double *slot_voltage_operators;
    /* linear system (host copies; the GPU keeps its own) */
    double *system_matrix;          /* n_total x n_total, row-major */
    double *rhs;
    double *solution;
    LUFactorization lu;
    int factorized;
    int initialized;
} TransientSystem;

/* transient.h:240,244 */
void transient_system_init(TransientSystem *system);
void transient_system_free(TransientSystem *system);
/* transient.h:260 (transient.c:385-478): 0, or -1 (bad input / GPU failure) */
int transient_system_create_stage1(TransientSystem *system, const InductanceMatrix *L,
                                   const ResistanceMatrix *R, const IncidenceMatrix *A,
                                   const PotentialMatrix *P, TransientScheme scheme, double dt);
/* transient.h:282 (transient.c:480-610) */
int transient_system_create_stage2(TransientSystem *system, const InductanceMatrix *L,
                                   const ResistanceMatrix *R, const IncidenceMatrix *A,
                                   const PotentialMatrix *P, const InternalCircuitShunt *shunt,
                                   TransientScheme scheme, double dt);
/* transient.h:303 (transient.c:940-1018): b_R and the shunt voltage operator */
int transient_build_shunt_operators(TransientSystem *system);
/* transient.h:308 (transient.c:1020-1419): fills system_matrix (GPU) */
int transient_build_system_matrix(TransientSystem *system);
/* transient.h:337: (size_t)-1 outside the slot-cell stage */
size_t transient_slot_current_offset(const TransientSystem *system);
/* transient.h:324,342,355: the slot-cell stage is not built by this library;
 * these print an ERROR line and return -1 */
int transient_system_create_slot_cells(TransientSystem *system, const InductanceMatrix *L,
                                       const ResistanceMatrix *R, const IncidenceMatrix *A,
                                       const PotentialMatrix *P, const SlotCircuit *slot_circuit,
                                       const InternalCircuitShunt *shunt, TransientScheme scheme,
                                       double dt);
int transient_build_rhs_slot_cells(TransientSystem *system, const double *edge_current_n,
                                   const double *node_self_voltage_n,
                                   const double *slot_inductor_current_n, double shunt_current_n,
                                   const double *node_source_n, const double *node_source_np1,
                                   const double *edge_voltage_n, const double *edge_voltage_np1);
int transient_extract_slot_cells_solution(const TransientSystem *system, double *edge_current,
                                          double *node_self_voltage, double *slot_inductor_current,
                                          double *shunt_current);
/* transient.h:364 (transient.c:2014-2066): real LU of system_matrix (GPU) */
int transient_factorize(TransientSystem *system, int parallel_threads);
/* transient.h:391 (transient.c:1421-1659): rhs of the step n -> n+1 from I_n,
 * V_c,n; sources may be NULL (zero) */
int transient_build_rhs_stage1(TransientSystem *system, const double *edge_current_n,
                               const double *node_self_voltage_n, const double *node_source_n,
                               const double *node_source_np1, const double *edge_voltage_n,
                               const double *edge_voltage_np1);
/* transient.h:411 (transient.c:1661-1749) */
int transient_build_rhs_stage2(TransientSystem *system, const double *edge_current_n,
                               const double *node_self_voltage_n, double shunt_current_n,
                               const double *node_source_n, const double *node_source_np1,
                               const double *edge_voltage_n, const double *edge_voltage_np1);
/* transient.h:423 (transient.c:2068-2097): solution = LU \ rhs (GPU) */
int transient_solve_step(TransientSystem *system);
/* transient.h:428-436 */
size_t transient_current_offset(const TransientSystem *system);
size_t transient_voltage_offset(const TransientSystem *system);
size_t transient_shunt_index(const TransientSystem *system);
/* transient.h:441,448 (transient.c:2099-2176) */
int transient_extract_stage1_solution(const TransientSystem *system, double *edge_current,
                                      double *node_self_voltage);
int transient_extract_stage2_solution(const TransientSystem *system, double *edge_current,
                                      double *node_self_voltage, double *shunt_current);
/* transient.h:474 (transient.c:2178-2225): node potentials phi = P F V_c and
 * charges q = F V_c (either may be NULL) */
int transient_recover_node_state(const TransientSystem *system, const double *node_self_voltage,
                                 double *node_potential, double *node_charge);
/* transient.h:482 */
void transient_print_info(const TransientSystem *system);

#ifdef __cplusplus
}
#endif

#endif

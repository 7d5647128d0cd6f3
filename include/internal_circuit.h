/* Resistive internal shunt between two mesh nodes, drop-in for the
 * reference's include/internal_circuit.h:1-320 (the stage-2 transient's
 * circuit element; same struct layout, host code). */
#ifndef PEEC_INTERNAL_CIRCUIT_H
#define PEEC_INTERNAL_CIRCUIT_H

#include "mesh.h"

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum
{
    INTERNAL_CIRCUIT_RESISTIVE_SHUNT = 0
} InternalCircuitType;

typedef struct
{
    InternalCircuitType type;
    size_t node_a;          /* positive current flows A -> B */
    size_t node_b;
    double resistance;      /* Ohm; 0 = ideal short */
    double requested_a[3];  /* coordinates asked for */
    double requested_b[3];
    double mapped_a[3];     /* coordinates of the mesh nodes used */
    double mapped_b[3];
    double mapping_error_a; /* distance requested -> mapped [m] */
    double mapping_error_b;
    int nodes_mapped;
    int initialized;
} InternalCircuitShunt;

/* internal_circuit.h:144: type resistive, R = 1000 Ohm, no nodes mapped */
void internal_circuit_shunt_default(InternalCircuitShunt *shunt);
/* internal_circuit.h:156: terminals at mesh nodes node_a != node_b */
int internal_circuit_shunt_set_nodes(InternalCircuitShunt *shunt, const Mesh *mesh,
                                     size_t node_a, size_t node_b);
/* internal_circuit.h:180: terminals at the mesh nodes nearest two points */
int internal_circuit_shunt_map_coordinates(InternalCircuitShunt *shunt, const Mesh *mesh,
                                           const double point_a[3], const double point_b[3]);
/* internal_circuit.h:193: nearest node (first of equals), optional distance */
int internal_circuit_find_nearest_node(const Mesh *mesh, const double point[3],
                                       size_t *node_id, double *distance);
/* internal_circuit.h:206 */
int internal_circuit_shunt_validate(const InternalCircuitShunt *shunt, size_t n_nodes);
/* internal_circuit.h:236: b_R[A] = -1, b_R[B] = +1, zero elsewhere */
int internal_circuit_build_incidence_vector(const InternalCircuitShunt *shunt, size_t n_nodes,
                                            double *b_R);
/* internal_circuit.h:256: V_A - V_B (NaN on invalid input) */
double internal_circuit_shunt_voltage(const InternalCircuitShunt *shunt, const double *node_voltage,
                                      size_t n_nodes);
/* internal_circuit.h:284: (V_A - V_B) / R (NaN for R = 0) */
double internal_circuit_shunt_current_from_voltage(const InternalCircuitShunt *shunt,
                                                   const double *node_voltage, size_t n_nodes);
/* internal_circuit.h:305: V_A - V_B - R I */
double internal_circuit_shunt_residual(const InternalCircuitShunt *shunt, const double *node_voltage,
                                       size_t n_nodes, double shunt_current);
/* internal_circuit.h:318 */
void internal_circuit_shunt_print_info(const InternalCircuitShunt *shunt);

#ifdef __cplusplus
}
#endif

#endif

/*
 * Plane-wave excitation of the PEEC branches (drop-in for the reference's
 * include/scattering.h, src/physics/scattering.c:44-811).
 *
 * U_e is the incident field integrated along edge e; the harmonic system is
 * (R + jwL) I + A V = -U, so the right-hand side is [-U ; 0] (no sign is
 * applied here).
 */
#ifndef PEEC_SCATTERING_H
#define PEEC_SCATTERING_H

#include <stddef.h>

#include "complex_matrix.h"
#include "incident_field.h"
#include "mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum
{
    SCATTERING_EDGE_MIDPOINT = 0,           /* U_e = E(r_c) . l_e */
    SCATTERING_EDGE_ANALYTIC_PLANE_WAVE = 1 /* U_e = [E(r_c) . l_e] sinc(k0 k_hat.l_e / 2) */
} ScatteringEdgeIntegration;

typedef struct
{
    ScatteringEdgeIntegration integration;
    int parallel_threads;
} ScatteringOptions;

/* integration = MIDPOINT */
void scattering_options_default(ScatteringOptions *options, int parallel_threads);

/* edge_voltage[n_edges]; 0, or -1 (bad input, pulse field, non-finite value). */
int scattering_build_harmonic_excitation(const Mesh *mesh, const IncidentField *field,
                                         const ScatteringOptions *options, Complex *edge_voltage);

/* Real-valued U_e(t) = E(r_c, t) . l_e. */
int scattering_build_time_excitation(const Mesh *mesh, const IncidentField *field,
                                     const ScatteringOptions *options, double time, double *edge_voltage);

void scattering_zero_node_current_real(size_t n_nodes, double *node_current);
void scattering_zero_node_current_complex(size_t n_nodes, Complex *node_current);

void scattering_print_harmonic_excitation_info(const Complex *edge_voltage, size_t n_edges);
void scattering_print_time_excitation_info(const double *edge_voltage, size_t n_edges, double time);

#ifdef __cplusplus
}
#endif

#endif

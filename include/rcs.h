/*
 * Far field and radar cross section of the PEEC branch currents (drop-in
 * for the reference's include/rcs.h, src/physics/rcs.c:600-1309).
 *
 * Radiation vector F = sum_e (I_e / w_e) e_e G_e with
 * G_e = integral over the edge's dual region of exp(+j k r_hat . r') dS'
 * (w_e = S_e / l_e), transverse part T = r_hat (r_hat . F) - F and
 *     sigma = k^2 eta0^2 |T|^2 / (4 pi E0^2).
 */
#ifndef PEEC_RCS_H
#define PEEC_RCS_H

#include <stddef.h>

#include "complex_matrix.h"
#include "dual_mesh.h"
#include "harmonic.h"
#include "incident_field.h"
#include "mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum
{
    RCS_CURRENT_MIDPOINT = 0,             /* F_e = I_e l_e exp(j k r_hat . r_c) */
    RCS_CURRENT_DUAL_PATCH_QUADRATURE = 1 /* G_e by Gauss-Legendre over the dual region */
} RcsIntegration;

typedef struct
{
    RcsIntegration integration;
    int surface_order; /* Gauss-Legendre order per triangle direction: 2,3,4,5,6,8,10,12,16 */
    int parallel_threads;
} RcsOptions;

typedef struct
{
    double theta_deg;
    double phi_deg;
    double direction[3];              /* r_hat, object to observer */
    Complex radiation_vector[3];      /* F [A m] */
    Complex transverse_vector[3];     /* T = r_hat x (r_hat x F) */
    double sigma;                     /* m^2 */
    double sigma_dbsm;                /* 10 log10(sigma), -inf when sigma <= DBL_MIN */
} RcsResult;

/* integration = DUAL_PATCH_QUADRATURE, surface_order = 6 */
void rcs_options_default(RcsOptions *options, int parallel_threads);

int rcs_build_observation_direction(double theta_deg, double phi_deg, double direction[3]);

int rcs_compute_radiation_vector(const Mesh *mesh, const DualMesh *dual, const Complex *edge_current,
                                 double wave_number, const double observation_direction[3],
                                 const RcsOptions *options, Complex radiation_vector[3]);

int rcs_compute_direction(const Mesh *mesh, const DualMesh *dual, const IncidentField *incident_field,
                          const HarmonicSolution *solution, const RcsOptions *options, double theta_deg,
                          double phi_deg, RcsResult *result);

/* E_sca = j k eta0 exp(-j k r) / (4 pi r) T at distance r. */
int rcs_compute_far_electric_field(const RcsResult *result, double wave_number, double distance,
                                   Complex E_sca[3]);

void rcs_print_result(const RcsResult *result);

#ifdef __cplusplus
}
#endif

#endif

/*
 * Incident plane wave (drop-in for the reference's include/incident_field.h;
 * same enumerators, struct layout and prototypes,
 * src/physics/incident_field.c:100-860).
 *
 * Time convention exp(+j w t); the wave travels along
 *     k_hat = -(sin t cos p, sin t sin p, cos t),
 * i.e. it arrives from the direction (theta, phi). Polarization basis
 *     e_theta = (cos t cos p, cos t sin p, -sin t),  e_phi = (-sin p, cos p, 0).
 */
#ifndef PEEC_INCIDENT_FIELD_H
#define PEEC_INCIDENT_FIELD_H

#include "complex_matrix.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum
{
    INCIDENT_POLARIZATION_HORIZONTAL = 0, /* E along e_phi */
    INCIDENT_POLARIZATION_VERTICAL = 1    /* E along e_theta */
} IncidentPolarization;

typedef enum
{
    INCIDENT_WAVE_COS = 0,     /* E0 cos(w t - k0 k_hat.r + phase) */
    INCIDENT_WAVE_SIN = 1,     /* E0 sin(w t - k0 k_hat.r + phase) */
    INCIDENT_WAVE_COMPLEX = 2, /* phasor E0 exp(-j k0 k_hat.r + j phase) */
    INCIDENT_WAVE_PULSE = 3    /* pulse_k (exp(-alpha tau) - exp(-beta tau)), tau = t - k_hat.r / c0 */
} IncidentWaveType;

typedef struct
{
    IncidentWaveType wave_type;
    IncidentPolarization polarization;
    double theta_deg;
    double phi_deg;
    double theta; /* rad */
    double phi;   /* rad */
    double k_hat[3];
    double e_theta[3];
    double e_phi[3];
    double E_amp; /* V/m */
    double E0[3]; /* E_amp times the polarization vector */
    double frequency;
    double omega;
    double wave_number; /* omega / c0 */
    double phase;       /* rad */
    double pulse_k;     /* V/m */
    double pulse_alpha; /* 1/s */
    double pulse_beta;  /* 1/s */
} IncidentField;

void incident_field_init(IncidentField *field);

/* wave_type COS, SIN or COMPLEX; 0 or -1 (invalid input). */
int incident_field_configure_harmonic(IncidentField *field, IncidentWaveType wave_type,
                                      IncidentPolarization polarization, double E_amp, double frequency,
                                      double theta_deg, double phi_deg, double phase);

/* Pulse; the angles are converted as the reference does
 * (incident_field.c:470-480: degrees / 180, without the factor pi). */
int incident_field_configure_pulse(IncidentField *field, IncidentPolarization polarization, double pulse_k,
                                   double alpha, double beta, double theta_deg, double phi_deg);

/* Real field E(r, t) [V/m] for COS, SIN, PULSE and COMPLEX. */
int incident_field_evaluate_time(const IncidentField *field, const double r[3], double time, double E[3]);

/* Phasor E_hat(r) for COS, SIN (extra -pi/2) and COMPLEX. */
int incident_field_evaluate_phasor(const IncidentField *field, const double r[3], Complex E[3]);

void incident_field_print_info(const IncidentField *field);

#ifdef __cplusplus
}
#endif

#endif

/*
 * Lightning current source (drop-in for the reference's include/lightning.h;
 * same struct layout and prototypes, src/physics/lightning.c:13-470).
 *
 *     I_L(t) = K [exp(-alpha (t - delay)) - exp(-beta (t - delay))],  t >= delay,
 *
 * injected at strike_node (+I_L) and drawn at return_node (-I_L).
 */
#ifndef PEEC_LIGHTNING_H
#define PEEC_LIGHTNING_H

#include <stddef.h>

#include "mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum
{
    LIGHTNING_WAVE_DOUBLE_EXPONENTIAL = 0
} LightningWaveType;

typedef struct
{
    LightningWaveType wave_type;
    double K;     /* A (not the peak current) */
    double alpha; /* 1/s, slow decay */
    double beta;  /* 1/s, fast front; beta > alpha */
    double delay; /* s */
    size_t strike_node;
    size_t return_node;
} LightningSource;

/* K = 102900 A, alpha = 1500 1/s, beta = 1e6 1/s, delay 0, nodes 0 */
void lightning_source_default(LightningSource *source);

/* 0, or -1 with an ERROR line */
int lightning_source_validate(const LightningSource *source, size_t n_nodes);

/* I_L(t) [A]; NaN for a null source or non-finite time */
double lightning_current(const LightningSource *source, double time);

/* t_peak = delay + ln(beta / alpha) / (beta - alpha) and I_L(t_peak) */
int lightning_peak(const LightningSource *source, double *time_peak, double *current_peak);

/* node_current[n_nodes]: zero except +I_L at strike, -I_L at return */
int lightning_build_node_current(const LightningSource *source, size_t n_nodes, double time, double *node_current);

/* nearest mesh node to point (first on ties) */
int lightning_find_nearest_node(const Mesh *mesh, const double point[3], size_t *node_id, double *distance);

void lightning_print_info(const LightningSource *source, const Mesh *mesh);

#ifdef __cplusplus
}
#endif

#endif

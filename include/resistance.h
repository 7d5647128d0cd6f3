/* Diagonal branch resistance, drop-in for the reference's
 * include/resistance.h:20-190. PEC gives zeros; a resistive sheet gives
 * R_s * l^2 / S^e per edge. */
#ifndef PEEC_RESISTANCE_H
#define PEEC_RESISTANCE_H

#include "dual_mesh.h"
#include "mesh.h"

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum
{
    PEEC_CONDUCTOR_PEC = 0,
    PEEC_CONDUCTOR_SHEET_RESISTANCE
} PeecConductorModel;

typedef struct
{
    PeecConductorModel model;
    double sheet_resistance; /* ohm per square */
    int parallel_threads;
} ResistanceOptions;

typedef struct
{
    size_t n;
    double *diagonal;
} ResistanceMatrix;

/* resistance.h:106-184 (resistance.c:87-260) */
void resistance_options_default(ResistanceOptions *options, int parallel_threads);
void resistance_init(ResistanceMatrix *R);
void resistance_free(ResistanceMatrix *R);
int resistance_compute_matrix(const Mesh *mesh, const DualMesh *dual,
                              const ResistanceOptions *options,
                              ResistanceMatrix *R);
double resistance_get(const ResistanceMatrix *R, size_t edge);

#ifdef __cplusplus
}
#endif

#endif

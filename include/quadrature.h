/* 1/R double integrals over pairs of QuadPatch regions, drop-in for the
 * reference's include/quadrature.h:20-226. Every integral here is evaluated
 * by the gfx950 assembly kernels; the geometric helpers run on the host. */
#ifndef PEEC_QUADRATURE_H
#define PEEC_QUADRATURE_H

#include "dual_mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum
{
    QUAD_PAIR_SELF = 0,  /* identical regions (singular integrand) */
    QUAD_PAIR_TOUCHING,  /* at least one shared vertex */
    QUAD_PAIR_NEAR,      /* AABB gap below near_factor * max side */
    QUAD_PAIR_FAR
} QuadPairType;

/* Gauss-Legendre orders the rules are defined for (quadrature.h:85-89). */
static inline int gauss_legendre_order_supported(int order)
{
    return order == 2 || order == 3 || order == 4 || order == 5 || order == 6
        || order == 8 || order == 10 || order == 12 || order == 16;
}

/* quadrature.h:91; nodes/weights on [-1,1]; 0 or -1 (unsupported order) */
int gauss_legendre_rule(int order, double *x, double *w);
/* quadrature.h:107 */
double triangle_area(const double p0[3], const double p1[3], const double p2[3]);
/* quadrature.h:124; split (0,1,2)+(0,2,3) */
double quad_patch_area(const QuadPatch *patch);
/* quadrature.h:136; longest side */
double quad_patch_size(const QuadPatch *patch);
/* quadrature.h:149 */
double quad_patch_min_vertex_distance(const QuadPatch *a, const QuadPatch *b);
/* quadrature.h:170 (quadrature.c:550-607) */
QuadPairType quad_patch_classify_pair(const QuadPatch *a, const QuadPatch *b,
                                      double near_factor);
/* quadrature.h:218-226 (quadrature.c:975-1079): K = int_a int_b dS dS'/|r-r'|
 * in m^3, NaN on unsupported orders or unconverged near-field quadrature.
 * Evaluated on the GPU by the same kernel that assembles Lp and P. */
double quad_patch_kernel_1_over_r(const QuadPatch *a, const QuadPatch *b,
                                  int far_order, int near_order,
                                  int touching_order, int self_order,
                                  double near_factor);

#ifdef __cplusplus
}
#endif

#endif

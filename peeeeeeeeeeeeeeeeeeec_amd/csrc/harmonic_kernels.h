// Launchers of the MNA-system kernels shared by the drop-in harmonic API and
// the frequency-point engine.
#pragma once

#include <vector>

#include "common.h"

namespace peec {

void launch_mna_fill(double2 *Z, double2 *Z0, size_t ldz, int ne, int nv, const double *L, const double2 *Lw,
                     const double *R, const double *P, const double2 *Pw, const double *PAT, const int *edges,
                     double omega, hipStream_t s, double *rowmax = nullptr);
void launch_pat(const double *P, const int *edges, double *PAT, int nv, int ne, hipStream_t s);
void launch_general_rhs(double2 *B, int ne, int nv, const double *P, const double2 *U, const double2 *Is,
                        hipStream_t s);
void residual_summary(const std::vector<double> &terms, size_t n, size_t ne, double out[7]);

}  // namespace peec

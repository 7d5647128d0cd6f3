// The MNA system matrix Z = [[R + jwL, A], [-P A^T, jw I]] (harmonic.c:175-552)
// as a function of its blocks: kernels that only read Z evaluate its entries
// from the Lp / P blocks and the edge end points instead of reading a dense
// copy (N^2 complex FP64), with the same arithmetic per entry.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace peec {

// Blocks of one frequency point. Lw / Pw (retarded, complex) replace L / P
// when given; PAT (P A^T, nv x ne) replaces the gathers from P; R (one per
// edge) may be null. A is implicit in the edge end points (incidence.c:34-115).
struct MnaOperands {
    int ne = 0, nv = 0;
    const double *L = nullptr;
    const double2 *Lw = nullptr;
    const double *R = nullptr;
    const double *P = nullptr;
    const double2 *Pw = nullptr;
    const double *PAT = nullptr;
    const int *edges = nullptr;
    double omega = 0.0;
};

// One entry of Z.
__device__ inline double2 mna_entry(int row, int col, const MnaOperands &m) {
    const int ne = m.ne, nv = m.nv;
    double2 z;
    if (row < ne) {
        if (col < ne) {
            const size_t k = (size_t)row * ne + col;
            const double r = (row == col && m.R) ? m.R[row] : 0.0;
            if (m.Lw) {
                const double2 l = m.Lw[k];
                z = make_double2(r - m.omega * l.y, m.omega * l.x);
            } else {
                z = make_double2(r, m.omega * m.L[k]);
            }
        } else {
            const int node = col - ne;
            double a = 0.0;
            if (node == m.edges[2 * row]) a = -1.0;
            if (node == m.edges[2 * row + 1]) a = 1.0;
            z = make_double2(a, 0.0);
        }
    } else {
        const int i = row - ne;
        if (col < ne) {
            const int a = m.edges[2 * col], b = m.edges[2 * col + 1];
            if (m.Pw) {
                const double2 pb = m.Pw[(size_t)i * nv + b], pa = m.Pw[(size_t)i * nv + a];
                z = make_double2(-(pb.x - pa.x), -(pb.y - pa.y));
            } else if (m.PAT) {
                z = make_double2(-m.PAT[(size_t)i * ne + col], 0.0);
            } else {
                z = make_double2(-(m.P[(size_t)i * nv + b] - m.P[(size_t)i * nv + a]), 0.0);
            }
        } else {
            z = make_double2(0.0, (col - ne) == i ? m.omega : 0.0);
        }
    }
    return z;
}

}  // namespace peec

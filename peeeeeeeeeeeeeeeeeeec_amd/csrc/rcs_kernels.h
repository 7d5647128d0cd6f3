// Plane-wave excitation and far-field kernels (physics/scattering.c,
// physics/rcs.c) shared by the drop-in functions and the engine.
#pragma once

#include "../../include/dual_mesh.h"
#include "../../include/incident_field.h"
#include "../../include/mesh.h"
#include "../../include/rcs.h"
#include "common.h"

namespace peec {

// Edge geometry and the edges' dual regions, resident on the device.
struct RcsGeometry {
    int ne = 0;
    DeviceBuffer<double> center, vec, area;  // 3 ne, 3 ne, ne
    DeviceBuffer<int> poff;                  // ne + 1 offsets into patches
    DeviceBuffer<double> patches;            // 12 doubles per quad patch
    RcsGeometry(const Mesh *mesh, const DualMesh *dual, hipStream_t s);
};

// One harmonic plane wave as the kernels use it (incident_field.c:754-858).
struct PlaneWave {
    double k_hat[3];
    double E0[3];
    double k;      // wave number
    double gamma0; // phase
    int sin_form;  // SIN wave: the phasor carries an extra -pi/2
};
PlaneWave plane_wave_of(const IncidentField &f);

// U[w * ldu + e] = excitation of edge e by wave w (scattering.c:180-489);
// sign -1 writes -U (the right-hand side's edge part).
void launch_excitation(const RcsGeometry &g, const PlaneWave *d_waves, int nw, int integration, double sign,
                       double2 *U, size_t ldu, hipStream_t s);

// F[3 a + c] = radiation vector in direction d_dirs[3 a ..] of the currents
// I[a * ldi + e] (rcs.c:701-995). d_work: 6 * na * ceil(ne / 256) doubles.
void launch_radiation(const RcsGeometry &g, const double2 *I, size_t ldi, const double *d_dirs, int na, double k,
                      int integration, int order, double *d_work, double2 *d_F, hipStream_t s);
size_t radiation_work_size(int ne, int na);

// transverse vector and sigma of a result whose direction and radiation
// vector are set (rcs.c:536-598,1000-1165)
void rcs_result_finish(const double d[3], double k, double E_amp, RcsResult *r);

}  // namespace peec

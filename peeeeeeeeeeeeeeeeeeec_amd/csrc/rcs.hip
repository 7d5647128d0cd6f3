// Plane-wave excitation, far field and radar cross section on gfx950:
// drop-in implementations of include/incident_field.h, include/scattering.h
// and include/rcs.h (reference: src/physics/incident_field.c,
// src/physics/scattering.c, src/physics/rcs.c).
//
// The per-edge work (incident phasor, edge integral; dual-region phase
// integral for the far field) runs as HIP kernels, one thread per (wave or
// direction, edge); the field set-up, angles and the sigma formula are O(1)
// host code.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/rcs.h"
#include "../../include/scattering.h"
#include "rcs_kernels.h"

namespace peec {

constexpr double kPi = 3.14159265358979323846;
constexpr double kC0 = 299792458.0;        // incident_field.c:18
constexpr double kMu0 = 1.25663706212e-6;  // rcs.c:23
constexpr double kEps0 = 8.8541878128e-12; // rcs.c:26

RcsGeometry::RcsGeometry(const Mesh *mesh, const DualMesh *dual, hipStream_t s) {
    ne = (int)mesh->n_edges;
    center.upload(mesh->edge_centers, 3 * (size_t)ne, s);
    vec.upload(mesh->edge_vectors, 3 * (size_t)ne, s);
    if (dual && dual->edge_patches && dual->edge_region_areas && dual->n_edges == mesh->n_edges) {
        area.upload(dual->edge_region_areas, (size_t)ne, s);
        std::vector<int> off((size_t)ne + 1, 0);
        for (int e = 0; e < ne; ++e) off[e + 1] = off[e] + (int)dual->edge_patches[e].count;
        std::vector<double> pts((size_t)off[ne] * 12);
        for (int e = 0; e < ne; ++e)
            for (size_t p = 0; p < dual->edge_patches[e].count; ++p)
                std::memcpy(&pts[((size_t)off[e] + p) * 12], dual->edge_patches[e].patches[p].points,
                            12 * sizeof(double));
        poff.upload(off, s);
        patches.upload(pts, s);
        HIP_CHECK(hipStreamSynchronize(s));  // host vectors go out of scope
    } else {
        HIP_CHECK(hipStreamSynchronize(s));
    }
}

PlaneWave plane_wave_of(const IncidentField &f) {
    PlaneWave w{};
    for (int c = 0; c < 3; ++c) {
        w.k_hat[c] = f.k_hat[c];
        w.E0[c] = f.E0[c];
    }
    w.k = f.wave_number;
    w.gamma0 = f.phase;
    w.sin_form = f.wave_type == INCIDENT_WAVE_SIN ? 1 : 0;
    return w;
}

__device__ inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// scattering.c:23-43
__device__ inline double sinc_of(double x) {
    const double ax = fabs(x);
    if (ax < 1.0e-4) {
        const double x2 = x * x, x4 = x2 * x2;
        return 1.0 - x2 / 6.0 + x4 / 120.0;
    }
    return sin(x) / x;
}

__global__ void excitation_kernel(const double *center, const double *vec, int ne, const PlaneWave *waves, int nw,
                                  int integration, double sign, double2 *U, size_t ldu) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)nw * ne) return;
    const int w = (int)(idx / ne), e = (int)(idx % ne);
    const PlaneWave pw = waves[w];
    const double *rc = center + 3 * (size_t)e, *l = vec + 3 * (size_t)e;
    // incident_field.c:754-858: gamma = -k k_hat.r + phase (- pi/2 for SIN)
    double gamma = -pw.k * dot3(pw.k_hat, rc) + pw.gamma0;
    if (pw.sin_form) gamma -= kPi / 2.0;
    const double c = cos(gamma), s = sin(gamma);
    double re = (pw.E0[0] * c) * l[0] + (pw.E0[1] * c) * l[1] + (pw.E0[2] * c) * l[2];
    double im = (pw.E0[0] * s) * l[0] + (pw.E0[1] * s) * l[1] + (pw.E0[2] * s) * l[2];
    if (integration == SCATTERING_EDGE_ANALYTIC_PLANE_WAVE) {
        const double factor = sinc_of(0.5 * (pw.k * dot3(pw.k_hat, l)));
        re *= factor;
        im *= factor;
    }
    U[(size_t)w * ldu + e] = make_double2(sign * re, sign * im);
}

void launch_excitation(const RcsGeometry &g, const PlaneWave *d_waves, int nw, int integration, double sign,
                       double2 *U, size_t ldu, hipStream_t s) {
    const size_t total = (size_t)nw * g.ne;
    hipLaunchKernelGGL(excitation_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g.center.ptr,
                       g.vec.ptr, g.ne, d_waves, nw, integration, sign, U, ldu);
    HIP_CHECK(hipGetLastError());
}

struct GLRule {
    double x[16], w[16];
};

// rcs.c:157-362: integral over triangle (p0, p1, p2) of exp(j k d.r) dS in
// the reference's collapsed parametrisation
__device__ inline void triangle_phase(const double *p0, const double *p1, const double *p2, double k,
                                      const double *d, int order, const GLRule &gl, double &gre, double &gim) {
    const double a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const double b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    const double cr[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double cn = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    gre = 0.0;
    gim = 0.0;
    for (int iu = 0; iu < order; ++iu) {
        const double u = 0.5 * (gl.x[iu] + 1.0), wu = 0.5 * gl.w[iu];
        for (int iv = 0; iv < order; ++iv) {
            const double v = 0.5 * (gl.x[iv] + 1.0), wv = 0.5 * gl.w[iv];
            const double xi = u, eta = (1.0 - u) * v;
            const double r[3] = {p0[0] + xi * a[0] + eta * b[0], p0[1] + xi * a[1] + eta * b[1],
                                 p0[2] + xi * a[2] + eta * b[2]};
            const double phase = k * dot3(d, r);
            const double weight = wu * wv * (cn * (1.0 - u));
            double sp, cp;
            sincos(phase, &sp, &cp);
            gre += weight * cp;
            gim += weight * sp;
        }
    }
}

// one block of 256 edges for one direction; the block's partial sum of the
// six real components of F goes to work[(a * nblk + block) * 6]
__global__ __launch_bounds__(256) void radiation_kernel(const double *center, const double *vec, const double *area,
                                                        const int *poff, const double *patches, int ne,
                                                        const double2 *I, size_t ldi, const double *dirs, double k,
                                                        int integration, int order, GLRule gl, double *work) {
    __shared__ double red[6][256];
    const int a = blockIdx.y, e = blockIdx.x * 256 + threadIdx.x;
    const double *d = dirs + 3 * (size_t)a;
    double f[6] = {0, 0, 0, 0, 0, 0};
    if (e < ne) {
        const double *lv = vec + 3 * (size_t)e;
        const double len = sqrt(lv[0] * lv[0] + lv[1] * lv[1] + lv[2] * lv[2]);
        const double2 cur = I[(size_t)a * ldi + e];
        bool use = len > 0.0;
        double sre = 0.0, sim = 0.0;
        if (use && integration == RCS_CURRENT_MIDPOINT) {
            double sp, cp;
            sincos(k * dot3(d, center + 3 * (size_t)e), &sp, &cp);
            sre = (cur.x * cp - cur.y * sp) * len;
            sim = (cur.x * sp + cur.y * cp) * len;
        } else if (use) {
            const double ar = area[e];
            if (ar <= 0.0) {
                use = false;
            } else {
                const double width = ar / len;
                double gre = 0.0, gim = 0.0;
                for (int p = poff[e]; p < poff[e + 1]; ++p) {
                    const double *pt = patches + 12 * (size_t)p;
                    double r0, i0, r1, i1;
                    triangle_phase(pt, pt + 3, pt + 6, k, d, order, gl, r0, i0);
                    triangle_phase(pt, pt + 6, pt + 9, k, d, order, gl, r1, i1);
                    gre += r0 + r1;
                    gim += i0 + i1;
                }
                const double cre = cur.x / width, cim = cur.y / width;
                sre = cre * gre - cim * gim;
                sim = cre * gim + cim * gre;
            }
        }
        if (use)
            for (int c = 0; c < 3; ++c) {
                const double ed = lv[c] / len;
                f[2 * c] = sre * ed;
                f[2 * c + 1] = sim * ed;
            }
    }
    for (int c = 0; c < 6; ++c) red[c][threadIdx.x] = f[c];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int c = 0; c < 6; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < 6) work[((size_t)a * gridDim.x + blockIdx.x) * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// F[a] = sum of the block partials in block order (deterministic)
__global__ void radiation_sum_kernel(const double *work, int nblk, int na, double2 *F) {
    const int a = blockIdx.x * 64 + threadIdx.x;
    if (a >= na) return;
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int b = 0; b < nblk; ++b)
        for (int c = 0; c < 6; ++c) s[c] += work[((size_t)a * nblk + b) * 6 + c];
    for (int c = 0; c < 3; ++c) F[3 * (size_t)a + c] = make_double2(s[2 * c], s[2 * c + 1]);
}

size_t radiation_work_size(int ne, int na) { return (size_t)6 * na * ((ne + 255) / 256); }

void launch_radiation(const RcsGeometry &g, const double2 *I, size_t ldi, const double *d_dirs, int na, double k,
                      int integration, int order, double *d_work, double2 *d_F, hipStream_t s) {
    GLRule gl{};
    if (integration == RCS_CURRENT_DUAL_PATCH_QUADRATURE && !gauss_legendre(order, gl.x, gl.w))
        throw HipError{"unsupported surface order"};
    const int nblk = (g.ne + 255) / 256;
    hipLaunchKernelGGL(radiation_kernel, dim3(nblk, na), dim3(256), 0, s, g.center.ptr, g.vec.ptr, g.area.ptr,
                       g.poff.ptr, g.patches.ptr, g.ne, I, ldi, d_dirs, k, integration, order, gl, d_work);
    hipLaunchKernelGGL(radiation_sum_kernel, dim3((na + 63) / 64), dim3(64), 0, s, d_work, nblk, na, d_F);
    HIP_CHECK(hipGetLastError());
}

}  // namespace peec

using namespace peec;

// ---------------------------------------------------------------------------
// incident field (host; incident_field.c)

static double vdot(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double vnorm(const double a[3]) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
static int vnormalize(double a[3]) {
    const double nrm = vnorm(a);
    if (!std::isfinite(nrm) || nrm <= 1.0e-14) return -1;
    a[0] /= nrm;
    a[1] /= nrm;
    a[2] /= nrm;
    return 0;
}

// basis and E0 (incident_field.c:110-283)
static int build_basis_and_E0(IncidentField *f) {
    const double t = f->theta, p = f->phi;
    f->k_hat[0] = -std::sin(t) * std::cos(p);
    f->k_hat[1] = -std::sin(t) * std::sin(p);
    f->k_hat[2] = -std::cos(t);
    f->e_theta[0] = std::cos(t) * std::cos(p);
    f->e_theta[1] = std::cos(t) * std::sin(p);
    f->e_theta[2] = -std::sin(t);
    f->e_phi[0] = -std::sin(p);
    f->e_phi[1] = std::cos(p);
    f->e_phi[2] = 0.0;
    if (vnormalize(f->k_hat) || vnormalize(f->e_theta) || vnormalize(f->e_phi)) {
        PEEC_ERR("cannot build incident-field basis.");
        return -1;
    }
    if (std::fabs(vdot(f->k_hat, f->e_theta)) > 1.0e-12 || std::fabs(vdot(f->k_hat, f->e_phi)) > 1.0e-12) {
        PEEC_ERR("incident polarization is not transverse.");
        return -1;
    }
    const double *pol;
    if (f->polarization == INCIDENT_POLARIZATION_VERTICAL) pol = f->e_theta;
    else if (f->polarization == INCIDENT_POLARIZATION_HORIZONTAL) pol = f->e_phi;
    else {
        PEEC_ERR("unknown incident polarization.");
        return -1;
    }
    for (int c = 0; c < 3; ++c) f->E0[c] = f->E_amp * pol[c];
    return 0;
}

extern "C" {

void incident_field_init(IncidentField *f) {
    if (!f) return;
    std::memset(f, 0, sizeof(*f));
    f->wave_type = INCIDENT_WAVE_COS;
    f->polarization = INCIDENT_POLARIZATION_VERTICAL;
}

int incident_field_configure_harmonic(IncidentField *f, IncidentWaveType wave_type, IncidentPolarization pol,
                                      double E_amp, double frequency, double theta_deg, double phi_deg,
                                      double phase) {
    if (!f) return -1;
    if (wave_type != INCIDENT_WAVE_COS && wave_type != INCIDENT_WAVE_SIN && wave_type != INCIDENT_WAVE_COMPLEX) {
        PEEC_ERR("invalid harmonic incident wave type.");
        return -1;
    }
    if (!std::isfinite(E_amp) || !std::isfinite(frequency) || frequency <= 0.0 || !std::isfinite(theta_deg) ||
        !std::isfinite(phi_deg) || !std::isfinite(phase)) {
        PEEC_ERR("invalid harmonic incident-field parameters.");
        return -1;
    }
    incident_field_init(f);
    f->wave_type = wave_type;
    f->polarization = pol;
    f->E_amp = E_amp;
    f->frequency = frequency;
    f->theta_deg = theta_deg;
    f->phi_deg = phi_deg;
    f->phase = phase;
    f->theta = theta_deg * kPi / 180.0;
    f->phi = phi_deg * kPi / 180.0;
    f->omega = 2.0 * kPi * frequency;
    f->wave_number = f->omega / kC0;
    return build_basis_and_E0(f);
}

int incident_field_configure_pulse(IncidentField *f, IncidentPolarization pol, double pulse_k, double alpha,
                                   double beta, double theta_deg, double phi_deg) {
    if (!f) return -1;
    if (!std::isfinite(pulse_k) || !std::isfinite(alpha) || !std::isfinite(beta) || alpha <= 0.0 || beta <= 0.0 ||
        !std::isfinite(theta_deg) || !std::isfinite(phi_deg)) {
        PEEC_ERR("invalid pulse incident-field parameters.");
        return -1;
    }
    incident_field_init(f);
    f->wave_type = INCIDENT_WAVE_PULSE;
    f->polarization = pol;
    f->pulse_k = pulse_k;
    f->pulse_alpha = alpha;
    f->pulse_beta = beta;
    f->theta_deg = theta_deg;
    f->phi_deg = phi_deg;
    // the reference converts these two without the factor pi
    // (incident_field.c:470-480); kept so the drop-in behaves the same
    f->theta = theta_deg / 180.0;
    f->phi = phi_deg / 180.0;
    if (build_basis_and_E0(f) != 0) return -1;
    f->E_amp = 1.0;
    return build_basis_and_E0(f);
}

int incident_field_evaluate_phasor(const IncidentField *f, const double r[3], Complex E[3]) {
    if (!f || !r || !E) return -1;
    if (f->wave_type != INCIDENT_WAVE_COS && f->wave_type != INCIDENT_WAVE_SIN &&
        f->wave_type != INCIDENT_WAVE_COMPLEX) {
        PEEC_ERR("phasor requested for non-harmonic field.");
        return -1;
    }
    double gamma = -f->wave_number * vdot(f->k_hat, r) + f->phase;
    if (f->wave_type == INCIDENT_WAVE_SIN) gamma -= kPi / 2.0;
    const double c = std::cos(gamma), s = std::sin(gamma);
    for (int k = 0; k < 3; ++k) {
        E[k].re = f->E0[k] * c;
        E[k].im = f->E0[k] * s;
    }
    return 0;
}

int incident_field_evaluate_time(const IncidentField *f, const double r[3], double time, double E[3]) {
    if (!f || !r || !E || !std::isfinite(time)) return -1;
    const double kr = vdot(f->k_hat, r);
    if (f->wave_type == INCIDENT_WAVE_COS || f->wave_type == INCIDENT_WAVE_SIN) {
        const double psi = f->omega * time - f->wave_number * kr + f->phase;
        const double v = f->wave_type == INCIDENT_WAVE_COS ? std::cos(psi) : std::sin(psi);
        for (int c = 0; c < 3; ++c) E[c] = f->E0[c] * v;
        return 0;
    }
    if (f->wave_type == INCIDENT_WAVE_PULSE) {
        const double tau = time - kr / kC0;
        if (tau < 0.0) {
            E[0] = E[1] = E[2] = 0.0;
            return 0;
        }
        const double env = f->pulse_k * (std::exp(-f->pulse_alpha * tau) - std::exp(-f->pulse_beta * tau));
        for (int c = 0; c < 3; ++c) E[c] = f->E0[c] * env;
        return 0;
    }
    if (f->wave_type == INCIDENT_WAVE_COMPLEX) {
        Complex ph[3];
        if (incident_field_evaluate_phasor(f, r, ph) != 0) return -1;
        const double c = std::cos(f->omega * time), s = std::sin(f->omega * time);
        for (int k = 0; k < 3; ++k) E[k] = ph[k].re * c - ph[k].im * s;
        return 0;
    }
    PEEC_ERR("unknown incident wave type.");
    return -1;
}

void incident_field_print_info(const IncidentField *f) {
    if (!f) return;
    std::printf("\nIncident field: type %d, polarization %d, theta %.6f deg, phi %.6f deg\n", (int)f->wave_type,
                (int)f->polarization, f->theta_deg, f->phi_deg);
    std::printf("  k_hat = (%.9f, %.9f, %.9f), E0 = (%.9e, %.9e, %.9e) V/m\n", f->k_hat[0], f->k_hat[1],
                f->k_hat[2], f->E0[0], f->E0[1], f->E0[2]);
    std::printf("  f = %.9e Hz, k0 = %.9e rad/m, phase = %.6f rad\n", f->frequency, f->wave_number, f->phase);
}

// ---------------------------------------------------------------------------
// excitation (scattering.c)

void scattering_options_default(ScatteringOptions *o, int parallel_threads) {
    if (!o) return;
    o->integration = SCATTERING_EDGE_MIDPOINT;
    o->parallel_threads = parallel_threads;
}

static int scattering_check(const Mesh *mesh, const ScatteringOptions *o) {
    if (!mesh || !o) return -1;
    if (mesh->n_edges == 0) {
        PEEC_ERR("scattering mesh contains no edges.");
        return -1;
    }
    if (!mesh->edge_vectors || !mesh->edge_centers) {
        PEEC_ERR("scattering requires edge vectors and edge centers.");
        return -1;
    }
    if (o->parallel_threads < 1) {
        PEEC_ERR("invalid OpenMP thread count in scattering.");
        return -1;
    }
    if (o->integration != SCATTERING_EDGE_MIDPOINT && o->integration != SCATTERING_EDGE_ANALYTIC_PLANE_WAVE) {
        PEEC_ERR("unsupported scattering edge integration method.");
        return -1;
    }
    return 0;
}

int scattering_build_harmonic_excitation(const Mesh *mesh, const IncidentField *f, const ScatteringOptions *o,
                                         Complex *U) {
    if (!mesh || !f || !o || !U) return -1;
    if (scattering_check(mesh, o) != 0) return -1;
    if (f->wave_type == INCIDENT_WAVE_PULSE) {
        PEEC_ERR("pulse incident field cannot be used directly in harmonic scattering.");
        return -1;
    }
    if (!require_device()) return -1;
    try {
        Stream s;
        RcsGeometry g(mesh, nullptr, s);
        PlaneWave pw = plane_wave_of(*f);
        DeviceBuffer<PlaneWave> dw;
        dw.upload(&pw, 1, s);
        DeviceBuffer<double2> dU((size_t)g.ne);
        launch_excitation(g, dw.ptr, 1, o->integration, 1.0, dU.ptr, (size_t)g.ne, s);
        dU.download(reinterpret_cast<double2 *>(U), (size_t)g.ne, s);
        s.sync();
    } catch (const HipError &err) {
        PEEC_ERR("%s", err.what.c_str());
        return -1;
    }
    for (size_t e = 0; e < mesh->n_edges; ++e)
        if (!std::isfinite(U[e].re) || !std::isfinite(U[e].im)) {
            PEEC_ERR("invalid harmonic incident excitation at edge %zu.", e);
            return -1;
        }
    return 0;
}

// O(N_e) host evaluation per time step (scattering.c:492-586), as the
// transient solver calls it once per step
int scattering_build_time_excitation(const Mesh *mesh, const IncidentField *f, const ScatteringOptions *o,
                                     double time, double *U) {
    if (!mesh || !f || !o || !U) return -1;
    if (scattering_check(mesh, o) != 0) return -1;
    if (!std::isfinite(time)) return -1;
    for (size_t e = 0; e < mesh->n_edges; ++e) {
        double E[3];
        if (incident_field_evaluate_time(f, &mesh->edge_centers[3 * e], time, E) != 0) return -1;
        const double *l = &mesh->edge_vectors[3 * e];
        U[e] = E[0] * l[0] + E[1] * l[1] + E[2] * l[2];
        if (!std::isfinite(U[e])) {
            PEEC_ERR("invalid time incident excitation at edge %zu.", e);
            return -1;
        }
    }
    return 0;
}

void scattering_zero_node_current_real(size_t n, double *x) {
    if (x) std::memset(x, 0, n * sizeof(double));
}
void scattering_zero_node_current_complex(size_t n, Complex *x) {
    if (x) std::memset(x, 0, n * sizeof(Complex));
}

void scattering_print_harmonic_excitation_info(const Complex *U, size_t ne) {
    if (!U || ne == 0) return;
    double mn = std::hypot(U[0].re, U[0].im), mx = mn, ss = 0.0;
    for (size_t e = 0; e < ne; ++e) {
        const double m = std::hypot(U[e].re, U[e].im);
        mn = std::fmin(mn, m);
        mx = std::fmax(mx, m);
        ss += m * m;
    }
    std::printf("\nHarmonic incident excitation: %zu edges, |U| min %.9e max %.9e rms %.9e V\n", ne, mn, mx,
                std::sqrt(ss / (double)ne));
}

void scattering_print_time_excitation_info(const double *U, size_t ne, double time) {
    if (!U || ne == 0) return;
    double mn = U[0], mx = U[0], ss = 0.0;
    for (size_t e = 0; e < ne; ++e) {
        mn = std::fmin(mn, U[e]);
        mx = std::fmax(mx, U[e]);
        ss += U[e] * U[e];
    }
    std::printf("\nTime excitation at t = %.9e s: %zu edges, U min %.9e max %.9e rms %.9e V\n", time, ne, mn, mx,
                std::sqrt(ss / (double)ne));
}

// ---------------------------------------------------------------------------
// far field (rcs.c)

void rcs_options_default(RcsOptions *o, int parallel_threads) {
    if (!o) return;
    o->integration = RCS_CURRENT_DUAL_PATCH_QUADRATURE;
    o->surface_order = 6;
    o->parallel_threads = parallel_threads;
}

int rcs_build_observation_direction(double theta_deg, double phi_deg, double d[3]) {
    if (!d || !std::isfinite(theta_deg) || !std::isfinite(phi_deg)) return -1;
    const double t = theta_deg * kPi / 180.0, p = phi_deg * kPi / 180.0;
    d[0] = std::sin(t) * std::cos(p);
    d[1] = std::sin(t) * std::sin(p);
    d[2] = std::cos(t);
    const double nrm = vnorm(d);
    if (!std::isfinite(nrm) || nrm <= 0.0) return -1;
    d[0] /= nrm;
    d[1] /= nrm;
    d[2] /= nrm;
    return 0;
}

int rcs_compute_radiation_vector(const Mesh *mesh, const DualMesh *dual, const Complex *I, double k,
                                 const double d[3], const RcsOptions *o, Complex F[3]) {
    if (!mesh || !dual || !I || !d || !o || !F) return -1;
    if (mesh->n_edges == 0 || !mesh->edge_vectors) {
        PEEC_ERR("RCS requires mesh edge vectors.");
        return -1;
    }
    if (dual->n_edges != mesh->n_edges || !dual->edge_patches || !dual->edge_region_areas) {
        PEEC_ERR("RCS requires dual edge regions.");
        return -1;
    }
    if (!std::isfinite(k) || k <= 0.0) return -1;
    if (o->parallel_threads < 1 ||
        (o->integration != RCS_CURRENT_MIDPOINT && o->integration != RCS_CURRENT_DUAL_PATCH_QUADRATURE) ||
        !gl_order_supported(o->surface_order))
        return -1;
    if (!require_device()) return -1;
    try {
        Stream s;
        RcsGeometry g(mesh, dual, s);
        DeviceBuffer<double2> dI;
        dI.upload(reinterpret_cast<const double2 *>(I), (size_t)g.ne, s);
        DeviceBuffer<double> dd;
        dd.upload(d, 3, s);
        DeviceBuffer<double> work(radiation_work_size(g.ne, 1));
        DeviceBuffer<double2> dF(3);
        launch_radiation(g, dI.ptr, (size_t)g.ne, dd.ptr, 1, k, o->integration, o->surface_order, work.ptr, dF.ptr,
                         s);
        dF.download(reinterpret_cast<double2 *>(F), 3, s);
        s.sync();
    } catch (const HipError &err) {
        PEEC_ERR("%s", err.what.c_str());
        return -1;
    }
    for (int c = 0; c < 3; ++c)
        if (!std::isfinite(F[c].re) || !std::isfinite(F[c].im)) return -1;
    return 0;
}

}  // extern "C"

// T = d (d.F) - F and sigma (rcs.c:536-598,1000-1165)
void peec::rcs_result_finish(const double d[3], double k, double E_amp, RcsResult *r) {
    const Complex *F = r->radiation_vector;
    const double dre = d[0] * F[0].re + d[1] * F[1].re + d[2] * F[2].re;
    const double dim = d[0] * F[0].im + d[1] * F[1].im + d[2] * F[2].im;
    double t2 = 0.0;
    for (int c = 0; c < 3; ++c) {
        r->transverse_vector[c].re = d[c] * dre - F[c].re;
        r->transverse_vector[c].im = d[c] * dim - F[c].im;
    }
    for (int c = 0; c < 3; ++c)
        t2 += r->transverse_vector[c].re * r->transverse_vector[c].re +
              r->transverse_vector[c].im * r->transverse_vector[c].im;
    const double eta0 = std::sqrt(kMu0 / kEps0);
    r->sigma = k * k * eta0 * eta0 * t2 / (4.0 * kPi * E_amp * E_amp);
    r->sigma_dbsm = r->sigma > DBL_MIN ? 10.0 * std::log10(r->sigma) : -INFINITY;
}

extern "C" {

int rcs_compute_direction(const Mesh *mesh, const DualMesh *dual, const IncidentField *f,
                          const HarmonicSolution *sol, const RcsOptions *o, double theta_deg, double phi_deg,
                          RcsResult *r) {
    if (!mesh || !dual || !f || !sol || !o || !r) return -1;
    if (!sol->edge_current || sol->n_edges != mesh->n_edges) {
        PEEC_ERR("RCS requires the harmonic edge currents.");
        return -1;
    }
    if (!std::isfinite(f->E_amp) || f->E_amp <= 0.0) {
        PEEC_ERR("RCS requires a positive incident amplitude.");
        return -1;
    }
    if (!std::isfinite(f->wave_number) || f->wave_number <= 0.0) {
        PEEC_ERR("RCS requires a positive wave number.");
        return -1;
    }
    std::memset(r, 0, sizeof(*r));
    r->theta_deg = theta_deg;
    r->phi_deg = phi_deg;
    if (rcs_build_observation_direction(theta_deg, phi_deg, r->direction) != 0) return -1;
    if (rcs_compute_radiation_vector(mesh, dual, sol->edge_current, f->wave_number, r->direction, o,
                                     r->radiation_vector) != 0)
        return -1;
    rcs_result_finish(r->direction, f->wave_number, f->E_amp, r);
    return 0;
}

int rcs_compute_far_electric_field(const RcsResult *r, double k, double dist, Complex E[3]) {
    if (!r || !E || !std::isfinite(k) || !std::isfinite(dist) || k <= 0.0 || dist <= 0.0) return -1;
    const double eta0 = std::sqrt(kMu0 / kEps0);
    const double amp = k * eta0 / (4.0 * kPi * dist);
    const double kr = k * dist;
    const Complex c{amp * std::sin(kr), amp * std::cos(kr)};  // j exp(-j k r) times the amplitude
    for (int i = 0; i < 3; ++i) {
        const Complex t = r->transverse_vector[i];
        E[i].re = c.re * t.re - c.im * t.im;
        E[i].im = c.re * t.im + c.im * t.re;
    }
    return 0;
}

void rcs_print_result(const RcsResult *r) {
    if (!r) return;
    std::printf("RCS theta %.6f phi %.6f deg: sigma %.9e m^2 (%.6f dBsm)\n", r->theta_deg, r->phi_deg, r->sigma,
                r->sigma_dbsm);
}

}  // extern "C"

// Lightning source (drop-in for include/lightning.h; reference
// src/physics/lightning.c) and the frequency-domain time response of a port
// to it (include/peec_gpu.h): O(N) host code over the engine's frequency
// sweep.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/lightning.h"
#include "../../include/peec_gpu.h"
#include "common.h"

extern "C" {

void lightning_source_default(LightningSource *s) {
    if (!s) return;
    s->wave_type = LIGHTNING_WAVE_DOUBLE_EXPONENTIAL;
    s->K = 102900.0;
    s->alpha = 1500.0;
    s->beta = 1.0e6;
    s->delay = 0.0;
    s->strike_node = 0;
    s->return_node = 0;
}

int lightning_source_validate(const LightningSource *s, size_t n_nodes) {
    if (!s || n_nodes == 0) return -1;
    if (s->wave_type != LIGHTNING_WAVE_DOUBLE_EXPONENTIAL) {
        PEEC_ERR("unsupported lightning waveform.");
        return -1;
    }
    if (!std::isfinite(s->K) || s->K <= 0.0) {
        PEEC_ERR("lightning K must be positive.");
        return -1;
    }
    if (!std::isfinite(s->alpha) || !std::isfinite(s->beta) || s->alpha <= 0.0 || s->beta <= 0.0) {
        PEEC_ERR("lightning alpha and beta must be positive.");
        return -1;
    }
    if (s->beta <= s->alpha) {
        PEEC_ERR("lightning requires beta > alpha.");
        return -1;
    }
    if (!std::isfinite(s->delay) || s->delay < 0.0) {
        PEEC_ERR("lightning delay must be non-negative.");
        return -1;
    }
    if (s->strike_node >= n_nodes) {
        PEEC_ERR("strike node %zu is outside mesh.", s->strike_node);
        return -1;
    }
    if (s->return_node >= n_nodes) {
        PEEC_ERR("return node %zu is outside mesh.", s->return_node);
        return -1;
    }
    if (s->strike_node == s->return_node) {
        PEEC_ERR("strike and return nodes must be different.");
        return -1;
    }
    return 0;
}

double lightning_current(const LightningSource *s, double t) {
    if (!s || !std::isfinite(t)) return NAN;
    if (t < s->delay) return 0.0;
    const double tau = t - s->delay;
    if (s->wave_type == LIGHTNING_WAVE_DOUBLE_EXPONENTIAL)
        return s->K * (std::exp(-s->alpha * tau) - std::exp(-s->beta * tau));
    return NAN;
}

int lightning_peak(const LightningSource *s, double *t_peak, double *i_peak) {
    if (!s || !t_peak || !i_peak) return -1;
    if (s->beta <= s->alpha || s->alpha <= 0.0) return -1;
    const double tau = std::log(s->beta / s->alpha) / (s->beta - s->alpha);
    *t_peak = s->delay + tau;
    *i_peak = s->K * (std::exp(-s->alpha * tau) - std::exp(-s->beta * tau));
    return 0;
}

int lightning_build_node_current(const LightningSource *s, size_t n_nodes, double t, double *node_current) {
    if (!s || !node_current || !std::isfinite(t)) return -1;
    if (lightning_source_validate(s, n_nodes) != 0) return -1;
    std::memset(node_current, 0, n_nodes * sizeof(double));
    const double i = lightning_current(s, t);
    if (!std::isfinite(i)) return -1;
    node_current[s->strike_node] = i;
    node_current[s->return_node] = -i;
    return 0;
}

int lightning_find_nearest_node(const Mesh *mesh, const double p[3], size_t *node, double *distance) {
    if (!mesh || !p || !node || !mesh->xyz || mesh->n_nodes == 0) return -1;
    size_t best = 0;
    double dx = mesh->xyz[0] - p[0], dy = mesh->xyz[1] - p[1], dz = mesh->xyz[2] - p[2];
    double bd = dx * dx + dy * dy + dz * dz;
    for (size_t j = 1; j < mesh->n_nodes; ++j) {
        const double *r = &mesh->xyz[3 * j];
        dx = r[0] - p[0];
        dy = r[1] - p[1];
        dz = r[2] - p[2];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < bd) {
            bd = d2;
            best = j;
        }
    }
    *node = best;
    if (distance) *distance = std::sqrt(bd);
    return 0;
}

void lightning_print_info(const LightningSource *s, const Mesh *mesh) {
    if (!s) return;
    double tp = 0.0, ip = 0.0;
    lightning_peak(s, &tp, &ip);
    std::printf("\nLightning source: I(t) = %.6e [exp(-%.6e t) - exp(-%.6e t)] A, delay %.6e s\n", s->K, s->alpha,
                s->beta, s->delay);
    std::printf("  peak %.6e A at %.6e s; strike node %zu, return node %zu", ip, tp, s->strike_node,
                s->return_node);
    if (mesh && mesh->xyz && s->strike_node < mesh->n_nodes && s->return_node < mesh->n_nodes)
        std::printf(" ((%.4f, %.4f, %.4f) -> (%.4f, %.4f, %.4f) m)", mesh->xyz[3 * s->strike_node],
                    mesh->xyz[3 * s->strike_node + 1], mesh->xyz[3 * s->strike_node + 2],
                    mesh->xyz[3 * s->return_node], mesh->xyz[3 * s->return_node + 1],
                    mesh->xyz[3 * s->return_node + 2]);
    std::printf("\n");
}

}  // extern "C"

// ---------------------------------------------------------------------------
// frequency-domain time response

namespace {

// in-place iterative radix-2 FFT, sign -1 forward (e^{-j}), +1 inverse
// (unscaled); n a power of two
void fft(std::vector<std::complex<double>> &a, int sign) {
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const double ang = sign * 2.0 * M_PI / (double)len;
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < len / 2; ++k) {
                const std::complex<double> w(std::cos(ang * (double)k), std::sin(ang * (double)k));
                const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
    }
}

}  // namespace

extern "C" {

int peec_lightning_spectrum(const LightningSource *s, double frequency, Complex *current) {
    if (!s || !current || !std::isfinite(frequency) || frequency < 0.0) return -1;
    const double w = 2.0 * M_PI * frequency;
    const std::complex<double> jw(0.0, w);
    const std::complex<double> I =
        s->K * std::exp(-jw * s->delay) * (1.0 / (s->alpha + jw) - 1.0 / (s->beta + jw));
    current->re = I.real();
    current->im = I.imag();
    return 0;
}

int peec_port_transient_from_spectrum(const LightningSource *s, size_t n_freq, double df, const Complex *z_in,
                                      double *time, double *current, double *voltage) {
    if (!s || !z_in || !time || !current || !voltage || n_freq < 2 || !(df > 0.0) || !std::isfinite(df))
        return -1;
    if ((n_freq & (n_freq - 1)) != 0) {
        PEEC_ERR("the number of frequency points must be a power of two (got %zu).", n_freq);
        return -1;
    }
    const size_t nt = 2 * n_freq;
    const double dt = 1.0 / (df * (double)nt);
    // V(f_k) = Z_in(f_k) I(f_k), k = 1..n_freq; Z_in(0) extrapolated from the
    // two lowest points (the MNA system needs f > 0), its imaginary part 0
    std::vector<std::complex<double>> X(nt, 0.0);
    for (size_t k = 0; k <= n_freq; ++k) {
        std::complex<double> z;
        if (k == 0) z = std::complex<double>(2.0 * z_in[0].re - z_in[1].re, 0.0);
        else z = std::complex<double>(z_in[k - 1].re, z_in[k - 1].im);
        Complex I;
        peec_lightning_spectrum(s, (double)k * df, &I);
        std::complex<double> v = z * std::complex<double>(I.re, I.im);
        if (k == n_freq) v = std::complex<double>(v.real(), 0.0);  // Nyquist bin of a real signal
        X[k] = v;
        if (k > 0 && k < n_freq) X[nt - k] = std::conj(v);
    }
    // v(t_n) = integral V(f) e^{j 2 pi f t} df ~ df sum_k X_k e^{j 2 pi k n / nt}
    fft(X, +1);
    for (size_t i = 0; i < nt; ++i) {
        time[i] = (double)i * dt;
        voltage[i] = df * X[i].real();
        current[i] = lightning_current(s, time[i]);
    }
    return 0;
}

int peec_engine_port_sweep(PeecEngine *e, const double *frequencies, size_t n, int mode, const int port_nodes[2],
                           Complex *z_in, PeecPointReport *reports) {
    if (!e || !frequencies || !port_nodes || !z_in || n == 0) return -1;
    for (size_t k = 0; k < n; ++k) {
        PeecPointReport rep{};
        const int st = peec_engine_solve(e, frequencies[k], mode, port_nodes, 1, &z_in[k], nullptr, nullptr, &rep);
        if (reports) reports[k] = rep;
        if (st != 0) return st;
    }
    return 0;
}

int peec_engine_lightning_response(PeecEngine *e, const LightningSource *s, size_t n_freq, double df, int mode,
                                   double *time, double *current, double *voltage, Complex *z_in,
                                   PeecPointReport *reports) {
    if (!e || !s || !time || !current || !voltage || n_freq < 2 || !(df > 0.0)) return -1;
    if (lightning_source_validate(s, peec_engine_nodes(e)) != 0) return -1;
    std::vector<double> f(n_freq);
    for (size_t k = 0; k < n_freq; ++k) f[k] = (double)(k + 1) * df;
    std::vector<Complex> z(n_freq);
    const int port[2] = {(int)s->strike_node, (int)s->return_node};
    const int st = peec_engine_port_sweep(e, f.data(), n_freq, mode, port, z.data(), reports);
    if (st != 0) return st;
    if (z_in) std::memcpy(z_in, z.data(), n_freq * sizeof(Complex));
    return peec_port_transient_from_spectrum(s, n_freq, df, z.data(), time, current, voltage);
}

}  // extern "C"

// Allocation-free form of harmonic_compute_residual, for callers that cannot
// allocate device memory (code under stream capture, tight serving loops,
// harnesses that forbid hipMalloc). Same arguments, checks and results as
// harmonic_compute_residual (harmonic.c:720-1080); the staged inputs (Z, P,
// U_e, I_s, x), the right-hand side, |x|, the per-row terms and the seven
// figures live in one file-scope device array of kScratchBytes per GPU, so
// the call makes no hipMalloc/hipFree and no host heap allocation.
//
// Every call on a GPU shares that GPU's array and leaves its contents
// behind: calls that overlap on one GPU (from several host threads) race.
// Work runs on the null stream and the call returns after it has finished.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>

#include "../../include/harmonic.h"
#include "common.h"
#include "harmonic_kernels.h"
#include "lu.h"

using namespace peec;

namespace {

// 32 MiB per GPU: N = Ne + Nv up to about 1400 unknowns (16 N^2 bytes of Z
// dominate).
constexpr size_t kScratchBytes = size_t(32) << 20;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of each GPU's copy of `scratch`, looked up once per device
// with that device current. Host code uses these addresses, never the
// array's name (a host-side stand-in).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

unsigned char *device_scratch() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        PEEC_ERR("harmonic_compute_residual_noalloc: no usable current HIP device.");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            PEEC_ERR("harmonic_compute_residual_noalloc: hipGetSymbolAddress of the scratch array failed: %s",
                     hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
    }
    return g_scratch_ptr[dev];
}

constexpr size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Byte offsets of the pieces inside the scratch array.
struct Layout {
    size_t z, p, u, is, x, b, ax, terms, out, total;
};

Layout layout_of(size_t ne, size_t nv) {
    const size_t n = ne + nv;
    Layout l{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off += align256(bytes);
        return at;
    };
    l.z = take(n * n * sizeof(double2));
    l.p = take(nv * nv * sizeof(double));
    l.u = take(ne * sizeof(double2));
    l.is = take(nv * sizeof(double2));
    l.x = take(n * sizeof(double2));
    l.b = take(n * sizeof(double2));
    l.ax = take(n * sizeof(double));
    l.terms = take(3 * n * sizeof(double));
    l.out = take(8 * sizeof(double));
    l.total = off;
    return l;
}

// The seven figures of harmonic.c:932-1060 from the per-row terms
// (|r_i|, sum_j |Z_ij||x_j| + |b_i|, |b_i|), as residual_summary forms them
// on the host: out = {max|r|, max|b|, max eta, edge max|r|, node max|r|,
// edge max eta, node max eta}. One workgroup; fmax is exact and order
// independent, so the figures are bit-identical to the host summary.
__global__ __launch_bounds__(256) void residual_summary_kernel(const double *terms, int n, int ne, double *out) {
    double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < n; i += 256) {
        const double ra = terms[3 * (size_t)i], den = terms[3 * (size_t)i + 1], ba = terms[3 * (size_t)i + 2];
        const double eta = den > 0.0 ? ra / den : ra;
        v[0] = fmax(v[0], ra);
        v[1] = fmax(v[1], ba);
        v[2] = fmax(v[2], eta);
        if (i < ne) {
            v[3] = fmax(v[3], ra);
            v[5] = fmax(v[5], eta);
        } else {
            v[4] = fmax(v[4], ra);
            v[6] = fmax(v[6], eta);
        }
    }
    __shared__ double red[4][7];
#pragma unroll
    for (int k = 0; k < 7; ++k)
        for (int off = 32; off > 0; off >>= 1) v[k] = fmax(v[k], __shfl_xor(v[k], off));
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 7; ++k) red[threadIdx.x >> 6][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 7) {
        const int k = threadIdx.x;
        out[k] = fmax(fmax(red[0][k], red[1][k]), fmax(red[2][k], red[3][k]));
    }
}

}  // namespace

extern "C" {

// No counterpart in the reference: harmonic_compute_residual without any
// device or host allocation (see the header comment above).
int harmonic_compute_residual_noalloc(const HarmonicSystem *sys, const PotentialMatrix *P,
                                      const HarmonicExcitation *ex, const HarmonicSolution *sol, int threads,
                                      HarmonicResidual *res) {
    if (!sys || !P || !ex || !sol || !res || threads < 1) return -1;
    if (!sys->matrix_built || !sys->matrix.data || !P->data) return -1;
    if (ex->n_edges != sys->n_edges || ex->n_nodes != sys->n_nodes) return -1;
    if (sol->n_edges != sys->n_edges || sol->n_nodes != sys->n_nodes || !sol->edge_current || !sol->node_voltage)
        return -1;
    const size_t ne = sys->n_edges, nv = sys->n_nodes, n = sys->system_size;
    if (n != ne + nv || P->n != nv) {
        PEEC_ERR("harmonic residual: inconsistent system, excitation or P dimensions.");
        return -1;
    }
    // the size check comes before anything touches the device
    const Layout l = layout_of(ne, nv);
    if (n > (size_t)INT32_MAX || l.total > kScratchBytes) {
        PEEC_ERR("harmonic_compute_residual_noalloc: N = %zu needs %zu bytes of scratch, above its maximum of "
                 "%zu bytes (N up to about 1400); use harmonic_compute_residual.",
                 n, l.total, kScratchBytes);
        return -1;
    }
    if (!require_device()) return -1;
    unsigned char *base = device_scratch();
    if (!base) return -1;
    auto at = [&](size_t off) { return base + off; };
    double2 *dZ = reinterpret_cast<double2 *>(at(l.z));
    double *dP = reinterpret_cast<double *>(at(l.p));
    double2 *dU = reinterpret_cast<double2 *>(at(l.u));
    double2 *dI = reinterpret_cast<double2 *>(at(l.is));
    double2 *dx = reinterpret_cast<double2 *>(at(l.x));
    double2 *db = reinterpret_cast<double2 *>(at(l.b));
    double *dax = reinterpret_cast<double *>(at(l.ax));
    double *dterms = reinterpret_cast<double *>(at(l.terms));
    double *dout = reinterpret_cast<double *>(at(l.out));
    hipStream_t s = nullptr;
    try {
        HIP_CHECK(hipMemcpyAsync(dZ, sys->matrix.data, n * n * sizeof(double2), hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(dP, P->data, nv * nv * sizeof(double), hipMemcpyHostToDevice, s));
        if (ex->edge_voltage)
            HIP_CHECK(hipMemcpyAsync(dU, ex->edge_voltage, ne * sizeof(double2), hipMemcpyHostToDevice, s));
        if (ex->node_current)
            HIP_CHECK(hipMemcpyAsync(dI, ex->node_current, nv * sizeof(double2), hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(dx, sol->edge_current, ne * sizeof(double2), hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(dx + ne, sol->node_voltage, nv * sizeof(double2), hipMemcpyHostToDevice, s));
        // b = [-U_e ; P I_s], as harmonic_build_rhs forms it
        launch_general_rhs(db, (int)ne, (int)nv, dP, ex->edge_voltage ? dU : nullptr,
                           ex->node_current ? dI : nullptr, s);
        complex_residual(dZ, (int)n, (int)n, n, dx, db, nullptr, dterms, s, dax);
        hipLaunchKernelGGL(residual_summary_kernel, dim3(1), dim3(256), 0, s, dterms, (int)n, (int)ne, dout);
        HIP_CHECK(hipGetLastError());
        double out[8];
        HIP_CHECK(hipMemcpyAsync(out, dout, 7 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        res->absolute_residual = out[0];
        res->rhs_relative_residual = out[1] > 0.0 ? out[0] / out[1] : out[0];
        res->backward_error = out[2];
        res->edge_absolute_residual = out[3];
        res->node_absolute_residual = out[4];
        res->edge_backward_error = out[5];
        res->node_backward_error = out[6];
        return 0;
    } catch (const HipError &e) {
        PEEC_ERR("%s", e.what.c_str());
        return -1;
    }
}

}  // extern "C"

// oqg_diag.hip -- diagnostics of the resident state on the device.
//
// The per-row partial sums and extrema behind qg_diagno's energetics and
// averages (diagnostics/qg_monitor.F90:219-383): the layer kinetic energy
// from the geostrophic velocities u = -dPdy_2bc(p)/f0, v = dPdx_2bc(p)/f0
// (glam/numerics.F90:470-513) integrated with the trapezoid weights of
// int_P_dA (numerics.F90:165-204), the averages of p and q, the interface
// displacements eta = dz_sign (p_k - p_{k+1})/gp_k and eta^2, the extrema of
// p and q; scan_cfl's velocity maxima on the staggered faces (dPdx / dPdy,
// numerics.F90:291-309; qg_monitor.F90:507-569); and a count of non-finite
// values of p and q (the check valids makes, valsubs.F90:46-139, extended to
// NaN and Inf).  One block per row; the host combines the rows in order.

#include <mutex>

#include "oqg_device.h"

namespace oqg {

// Scratch memory of oqg_diagnose_noalloc: the per-row partials of k_diag
// ([nyp][DG_LEN]) followed by their combination (DG_COMB_LEN doubles).  One
// copy per GPU, held until the process ends.  Host code never passes the
// array's name to the device: it uses the address hipGetSymbolAddress returns
// for each device (noalloc_scratch).
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};
}  // namespace

hipError_t noalloc_scratch(int dev, double** out) {
    *out = nullptr;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int prev = -1;
        hipError_t e = hipGetDevice(&prev);
        if (e != hipSuccess) return e;
        if (prev != dev && (e = hipSetDevice(dev)) != hipSuccess) return e;
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (prev != dev) {
            const hipError_t e2 = hipSetDevice(prev);
            if (e == hipSuccess) e = e2;
        }
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<double*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

// The rows' partials combined in row order, one thread per accumulator (each
// accumulator sees the rows in order, with the operations of the host
// combination in oqg_diagnose, so the two agree bit for bit): sums of the five
// summed quantities, the six extrema, p at column 1 of the first and last rows
// and the non-finite count.  Layers k >= nl are written as zeros.
__global__ __launch_bounds__(128) void k_diag_combine(const double* partial, int ny, int nl, double* comb) {
    const int t = threadIdx.x;
    if (t >= DG_COMB_LEN) return;
    if (t == DC_BAD) {
        double bad = 0.0;
        for (int j = 0; j < ny; j++) bad += partial[(long long)j * DG_LEN + DG_BAD];
        comb[DC_BAD] = bad;
        return;
    }
    const int k = t % MAXL;
    double v = 0.0;
    if (k < nl) {
        if (t < DC_EXT) {               // sums: ke, p, q, eta, eta^2
            const int slot = (t / MAXL) * MAXL + k;   // DG_KE..DG_ETA2 are 0..4
            for (int j = 0; j < ny; j++) v += partial[(long long)j * DG_LEN + slot];
        } else if (t < DC_P0FIRST) {    // extrema, seeded with row 0
            const int e = t / MAXL - 5;
            const int slot = (DG_PMIN + e) * MAXL + k;   // DG_PMIN..DG_VMAX are consecutive
            const bool is_min = (e == 0 || e == 2);
            v = partial[slot];
            for (int j = 0; j < ny; j++) {
                const double x = partial[(long long)j * DG_LEN + slot];
                v = is_min ? fmin(v, x) : fmax(v, x);
            }
        } else if (t < DC_P0LAST) {
            v = partial[DG_P0 * MAXL + k];
        } else {
            v = partial[(long long)(ny - 1) * DG_LEN + DG_P0 * MAXL + k];
        }
    }
    comb[t] = v;
}

void launch_diag_combine(const double* partial, int ny, int nl, double* comb, hipStream_t st) {
    hipLaunchKernelGGL(k_diag_combine, dim3(1), dim3(128), 0, st, partial, ny, nl, comb);
}

// One block per row, the layers one after another: per thread a few
// accumulators in registers (no per-thread arrays, so no private memory), then
// block reductions per layer.
__global__ __launch_bounds__(256) void k_diag(DiagArgs a) {
    const Geo& g = a.g;
    const int nx = g.nxp, ny = g.nyp, nl = g.nl, j = blockIdx.x;
    const long long pitch = g.pitch;
    __shared__ double red_s[6 * 16];
    __shared__ double red_m[6 * 16];
    const double wj = (j == 0 || j == ny - 1) ? 0.5 : 1.0;
    const double ca = 0.5 * a.bcco + 1.0;
    const double inf = __builtin_huge_val();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    double* out = a.partial + (long long)j * DG_LEN;
    double bad_total = 0.0;
#pragma unroll 1
    for (int k = 0; k < nl; k++) {
        const double* P = a.p + k * g.fs;
        const double* Q = a.q + k * g.fs;
        const double* P2 = (k < nl - 1) ? P + g.fs : P;
        const double gpk = (k < nl - 1) ? a.gp[k] : 1.0;
        double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};   // ke, p, q, eta, eta^2, non-finite (registers)
        double pmin = inf, pmax = -inf, qmin = inf, qmax = -inf, umax = 0.0, vmax = 0.0;
        for (int i = threadIdx.x; i < nx; i += blockDim.x) {
            const double w = wj * ((i == 0 || i == nx - 1) ? 0.5 : 1.0);
            const long long o = (long long)j * pitch + i;
            const double pc = P[o], qc = Q[o];
            double dpdx = 0.0, dpdy = 0.0;
            if (j != 0 && j != ny - 1) {
                if (i >= 1 && i <= nx - 2) dpdx = (P[o + 1] - P[o - 1]) / (2.0 * g.dx);
                else if (!g.cyclic) dpdx = (i == 0) ? (P[o + 1] - P[o]) / (g.dx * ca) : (P[o] - P[o - 1]) / (g.dx * ca);
                else dpdx = (P[(long long)j * pitch + 1] - P[(long long)j * pitch + nx - 2]) / (2.0 * g.dx);
            }
            if (g.cyclic || (i != 0 && i != nx - 1)) {
                if (j >= 1 && j <= ny - 2) dpdy = (P[o + pitch] - P[o - pitch]) / (2.0 * g.dy);
                else if (j == 0) dpdy = (P[o + pitch] - P[o]) / (g.dy * ca);
                else dpdy = (P[o] - P[o - pitch]) / (g.dy * ca);
            }
            const double u = -dpdy / g.fnot, vv = dpdx / g.fnot;
            v[0] += w * (u * u + vv * vv);
            v[1] += w * pc;
            v[2] += w * qc;
            pmin = fmin(pmin, pc);
            pmax = fmax(pmax, pc);
            qmin = fmin(qmin, qc);
            qmax = fmax(qmax, qc);
            if (j <= ny - 2) umax = fmax(umax, fabs(-((1.0 / g.dy) * (P[o + pitch] - pc)) / g.fnot));
            if (i <= nx - 2) vmax = fmax(vmax, fabs(((1.0 / g.dx) * (P[o + 1] - pc)) / g.fnot));
            v[5] += (isfinite(pc) ? 0.0 : 1.0) + (isfinite(qc) ? 0.0 : 1.0);
            if (k < nl - 1) {
                const double eta = a.dz_sign * (pc - P2[o]) / gpk;
                v[3] += w * eta;
                v[4] += w * (eta * eta);
            }
        }
        block_sum<6>(v, red_s);
        const double m0 = wave_min(pmin), m1 = wave_max(pmax), m2 = wave_min(qmin), m3 = wave_max(qmax),
                     m4 = wave_max(umax), m5 = wave_max(vmax);
        if (lane == 0) {
            red_m[0 * 16 + wv] = m0;
            red_m[1 * 16 + wv] = m1;
            red_m[2 * 16 + wv] = m2;
            red_m[3 * 16 + wv] = m3;
            red_m[4 * 16 + wv] = m4;
            red_m[5 * 16 + wv] = m5;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            out[DG_KE * MAXL + k] = v[0];
            out[DG_P * MAXL + k] = v[1];
            out[DG_Q * MAXL + k] = v[2];
            out[DG_ETA * MAXL + k] = (k < nl - 1) ? v[3] : 0.0;
            out[DG_ETA2 * MAXL + k] = (k < nl - 1) ? v[4] : 0.0;
            out[DG_P0 * MAXL + k] = P[(long long)j * pitch];
#pragma unroll
            for (int e = 0; e < 6; e++) {
                const bool is_min = (e == 0 || e == 2);
                double r = red_m[e * 16];
                for (int q = 1; q < nw; q++) r = is_min ? fmin(r, red_m[e * 16 + q]) : fmax(r, red_m[e * 16 + q]);
                out[(DG_PMIN + e) * MAXL + k] = r;
            }
            bad_total += v[5];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[DG_BAD] = bad_total;
}

void launch_diag(const DiagArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_diag, dim3(a.g.nyp), dim3(256), 0, st, a);
}

}  // namespace oqg

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
// The energetics from the lagged velocities, the stress and the entrainment
// (k_energy_d2, k_energy) follow the same row scheme.

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

// ------------------------------------------------------------------ energetics
//
// Point values of the reference's derivative operators, each written as its
// whole-field form assigns that point (the row or column zeroed last wins):
// dPdx_2bc, dPdy_2bc (numerics.F90:470-513), dP2dx2_bc, dP2dy2_bc (:427-468).
// f(i, j) gives the field's value at any point of the basin.

template <class F>
__device__ __forceinline__ double en_dpdx(F f, int i, int j, const Geo& g, double ca) {
    if (j == 0 || j == g.nyp - 1) return 0.0;
    if (i >= 1 && i <= g.nxp - 2) return (f(i + 1, j) - f(i - 1, j)) / (2.0 * g.dx);
    if (!g.cyclic) return i == 0 ? (f(1, j) - f(0, j)) / (g.dx * ca) : (f(i, j) - f(i - 1, j)) / (g.dx * ca);
    return (f(1, j) - f(g.nxp - 2, j)) / (2.0 * g.dx);
}

template <class F>
__device__ __forceinline__ double en_dpdy(F f, int i, int j, const Geo& g, double ca) {
    if (!g.cyclic && (i == 0 || i == g.nxp - 1)) return 0.0;
    if (j >= 1 && j <= g.nyp - 2) return (f(i, j + 1) - f(i, j - 1)) / (2.0 * g.dy);
    if (j == 0) return (f(i, 1) - f(i, 0)) / (g.dy * ca);
    return (f(i, j) - f(i, j - 1)) / (g.dy * ca);
}

// del2_P_bc at (i, j): dP2dx2_bc + dP2dy2_bc with alpha = bcco
template <class F>
__device__ __forceinline__ double en_del2(F f, int i, int j, const Geo& g, double alpha) {
    const double bfac = alpha / (0.5 * alpha + 1.0);
    double dxx = 0.0, dyy = 0.0;
    if (j != 0 && j != g.nyp - 1) {
        const double fac = 1.0 / (g.dx * g.dx);
        if (i >= 1 && i <= g.nxp - 2) dxx = fac * ((f(i - 1, j) + f(i + 1, j)) - 2.0 * f(i, j));
        else if (!g.cyclic) dxx = i == 0 ? fac * bfac * (f(1, j) - f(0, j)) : -fac * bfac * (f(i, j) - f(i - 1, j));
        else dxx = fac * ((f(g.nxp - 2, j) + f(1, j)) - 2.0 * f(0, j));
    }
    if (g.cyclic || (i != 0 && i != g.nxp - 1)) {
        const double fac = 1.0 / (g.dy * g.dy);
        if (j >= 1 && j <= g.nyp - 2) dyy = fac * ((f(i, j - 1) + f(i, j + 1)) - 2.0 * f(i, j));
        else if (j == 0) dyy = fac * bfac * (f(i, 1) - f(i, 0));
        else dyy = -fac * bfac * (f(i, j) - f(i, j - 1));
    }
    return dxx + dyy;
}

// del2 of the lagged velocities, one point per thread: d2u = del2(ug), d2v =
// del2(vg) with ug = -dPdy_2bc(pm)/f0, vg = dPdx_2bc(pm)/f0 evaluated at the
// stencil's points.
__global__ __launch_bounds__(256) void k_energy_d2(EnergyArgs a) {
    const Geo& g = a.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= g.nxp) return;
    const double ca = 0.5 * a.bcco + 1.0;
    const double* pm = a.pm + k * g.fs;
    const long long pitch = g.pitch;
    auto P = [&](int x, int y) { return pm[(long long)y * pitch + x]; };
    auto ug = [&](int x, int y) { return -en_dpdy(P, x, y, g, ca) / g.fnot; };
    auto vg = [&](int x, int y) { return en_dpdx(P, x, y, g, ca) / g.fnot; };
    const long long o = k * g.fs + (long long)j * pitch + i;
    a.d2u[o] = en_del2(ug, i, j, g, a.bcco);
    a.d2v[o] = en_del2(vg, i, j, g, a.bcco);
}

// One block per row j: the row's partials (EN_* slots), layers one after
// another, block sums in a fixed order.
__global__ __launch_bounds__(256) void k_energy(EnergyArgs a) {
    const Geo& g = a.g;
    const int nx = g.nxp, ny = g.nyp, nl = g.nl, j = blockIdx.x;
    const long long pitch = g.pitch;
    __shared__ double red[9 * 16];
    const double wj = (j == 0 || j == ny - 1) ? 0.5 : 1.0;
    const double ca = 0.5 * a.bcco + 1.0;
    double* out = a.partial + (long long)j * EN_LEN;
#pragma unroll 1
    for (int k = 0; k < nl; k++) {
        const double* p = a.p + k * g.fs;
        const double* pm = a.pm + k * g.fs;
        const double* d2u = a.d2u + k * g.fs;
        const double* d2v = a.d2v + k * g.fs;
        auto P = [&](int x, int y) { return p[(long long)y * pitch + x]; };
        auto PM = [&](int x, int y) { return pm[(long long)y * pitch + x]; };
        auto D2U = [&](int x, int y) { return d2u[(long long)y * pitch + x]; };
        auto D2V = [&](int x, int y) { return d2v[(long long)y * pitch + x]; };
        const bool first = k == 0, bottom = k == g.kt;
        double v[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int i = threadIdx.x; i < nx; i += blockDim.x) {
            const double wi = (i == 0 || i == nx - 1) ? 0.5 : 1.0, w = wj * wi;
            const long long o = (long long)j * pitch + i;
            const double ugeos = -en_dpdy(P, i, j, g, ca) / g.fnot;
            const double vgeos = en_dpdx(P, i, j, g, ca) / g.fnot;
            const double d4u = en_del2(D2U, i, j, g, a.bcco);
            const double d4v = en_del2(D2V, i, j, g, a.bcco);
            v[0] += w * (ugeos * d2u[o]);
            v[1] += w * (vgeos * d2v[o]);
            v[2] += w * (ugeos * d4u);
            v[3] += w * (vgeos * d4v);
            v[4] += wi * ugeos;
            if (first && a.taux) {
                v[5] += w * (a.taux[o] * ugeos);
                v[6] += w * (a.tauy[o] * vgeos);
            }
            if (bottom) {
                const double ug = -en_dpdy(PM, i, j, g, ca) / g.fnot;
                const double vg = en_dpdx(PM, i, j, g, ca) / g.fnot;
                v[7] += w * (ug * ug);
                v[8] += w * (vg * vg);
            }
        }
        block_sum<9>(v, red);
        if (threadIdx.x == 0) {
            out[EN_U2 * MAXL + k] = v[0];
            out[EN_V2 * MAXL + k] = v[1];
            out[EN_U4 * MAXL + k] = v[2];
            out[EN_V4 * MAXL + k] = v[3];
            out[EN_UJET * MAXL + k] = v[4];
            if (first) {
                out[EN_UTX] = v[5];
                out[EN_VTY] = v[6];
            }
            if (bottom) {
                out[EN_BTU] = v[7];
                out[EN_BTV] = v[8];
            }
        }
    }
    // the layer-independent sums: Ekman pumping, interface-1 entrainment and
    // eta_1 ent_1 (zero eta for one layer)
    double e[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
        const double w = wj * ((i == 0 || i == nx - 1) ? 0.5 : 1.0);
        const long long o = (long long)j * pitch + i;
        const double wk = a.wekp[o], en = a.ent1[o];
        const double eta = nl > 1 ? a.dz_sign * (a.p[o] - a.p[g.fs + o]) / a.gp0 : 0.0;
        e[0] += w * wk;
        e[1] += w * fabs(wk);
        e[2] += w * en;
        e[3] += w * fabs(en);
        e[4] += w * (eta * en);
    }
    block_sum<5>(e, red);
    if (threadIdx.x == 0) {
        out[EN_WEK] = e[0];
        out[EN_AWEK] = e[1];
        out[EN_ENT] = e[2];
        out[EN_AENT] = e[3];
        out[EN_ETAENT] = e[4];
    }
}

void launch_energy(const EnergyArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_energy_d2, dim3((a.g.nxp + 255) / 256, a.g.nyp, a.g.nl), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_energy, dim3(a.g.nyp), dim3(256), 0, st, a);
}

}  // namespace oqg

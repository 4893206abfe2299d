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

#include "oqg_device.h"

namespace oqg {

template <int NL>
__global__ __launch_bounds__(256) void k_diag(DiagArgs a) {
    const Geo& g = a.g;
    const int nx = g.nxp, ny = g.nyp, j = blockIdx.x;
    const long long pitch = g.pitch;
    constexpr int NS = 5 * NL + 1;   // ke, p, q (NL each), eta, eta^2 (NL-1 used each), non-finite count
    constexpr int NM = 6 * NL;       // min p, max p, min q, max q, max |u|, max |v|
    __shared__ double red_s[NS * 16];
    __shared__ double red_m[NM * 16];
    const double wj = (j == 0 || j == ny - 1) ? 0.5 : 1.0;
    const double ca = 0.5 * a.bcco + 1.0;
    const double inf = __builtin_huge_val();
    double s[NS], mm[NM];
#pragma unroll
    for (int t = 0; t < NS; t++) s[t] = 0.0;
#pragma unroll
    for (int k = 0; k < NL; k++) {
        mm[k] = inf;            // min p
        mm[NL + k] = -inf;      // max p
        mm[2 * NL + k] = inf;   // min q
        mm[3 * NL + k] = -inf;  // max q
        mm[4 * NL + k] = 0.0;   // max |u|
        mm[5 * NL + k] = 0.0;   // max |v|
    }
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
        const double w = wj * ((i == 0 || i == nx - 1) ? 0.5 : 1.0);
        const long long o = (long long)j * pitch + i;
        double pk[NL];
#pragma unroll
        for (int k = 0; k < NL; k++) {
            const double* P = a.p + k * g.fs;
            const double pc = P[o], qc = a.q[k * g.fs + o];
            pk[k] = pc;
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
            const double u = -dpdy / g.fnot, v = dpdx / g.fnot;
            s[k] += w * (u * u + v * v);
            s[NL + k] += w * pc;
            s[2 * NL + k] += w * qc;
            mm[k] = fmin(mm[k], pc);
            mm[NL + k] = fmax(mm[NL + k], pc);
            mm[2 * NL + k] = fmin(mm[2 * NL + k], qc);
            mm[3 * NL + k] = fmax(mm[3 * NL + k], qc);
            if (j <= ny - 2) mm[4 * NL + k] = fmax(mm[4 * NL + k], fabs(-((1.0 / g.dy) * (P[o + pitch] - pc)) / g.fnot));
            if (i <= nx - 2) mm[5 * NL + k] = fmax(mm[5 * NL + k], fabs(((1.0 / g.dx) * (P[o + 1] - pc)) / g.fnot));
            s[5 * NL] += (isfinite(pc) ? 0.0 : 1.0) + (isfinite(qc) ? 0.0 : 1.0);
        }
#pragma unroll
        for (int k = 0; k < NL - 1; k++) {
            const double eta = a.dz_sign * (pk[k] - pk[k + 1]) / a.gp[k];
            s[3 * NL + k] += w * eta;
            s[4 * NL + k] += w * (eta * eta);
        }
    }
    block_sum<NS>(s, red_s);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int t = 0; t < NM; t++) {
        const bool is_min = (t < NL) || (t >= 2 * NL && t < 3 * NL);
        const double r = is_min ? wave_min(mm[t]) : wave_max(mm[t]);
        if (lane == 0) red_m[t * 16 + wv] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double* out = a.partial + (long long)j * DG_LEN;
#pragma unroll
        for (int k = 0; k < NL; k++) {
            out[DG_KE * MAXL + k] = s[k];
            out[DG_P * MAXL + k] = s[NL + k];
            out[DG_Q * MAXL + k] = s[2 * NL + k];
            out[DG_ETA * MAXL + k] = (k < NL - 1) ? s[3 * NL + k] : 0.0;
            out[DG_ETA2 * MAXL + k] = (k < NL - 1) ? s[4 * NL + k] : 0.0;
            out[DG_P0 * MAXL + k] = a.p[k * g.fs + (long long)j * pitch];
        }
        const int kinds[6] = {DG_PMIN, DG_PMAX, DG_QMIN, DG_QMAX, DG_UMAX, DG_VMAX};
#pragma unroll
        for (int t = 0; t < NM; t++) {
            const bool is_min = (t < NL) || (t >= 2 * NL && t < 3 * NL);
            double r = red_m[t * 16];
            for (int q = 1; q < nw; q++) r = is_min ? fmin(r, red_m[t * 16 + q]) : fmax(r, red_m[t * 16 + q]);
            out[kinds[t / NL] * MAXL + (t % NL)] = r;
        }
        out[DG_BAD] = s[5 * NL];
    }
}

void launch_diag(const DiagArgs& a, hipStream_t st) {
    const dim3 grid(a.g.nyp), block(256);
    switch (a.g.nl) {
        case 2: hipLaunchKernelGGL(k_diag<2>, grid, block, 0, st, a); break;
        case 3: hipLaunchKernelGGL(k_diag<3>, grid, block, 0, st, a); break;
        case 4: hipLaunchKernelGGL(k_diag<4>, grid, block, 0, st, a); break;
        case 5: hipLaunchKernelGGL(k_diag<5>, grid, block, 0, st, a); break;
        case 6: hipLaunchKernelGGL(k_diag<6>, grid, block, 0, st, a); break;
        case 7: hipLaunchKernelGGL(k_diag<7>, grid, block, 0, st, a); break;
        case 8: hipLaunchKernelGGL(k_diag<8>, grid, block, 0, st, a); break;
        default: break;
    }
}

}  // namespace oqg

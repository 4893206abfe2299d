// oqg_mixed.hip -- the ocean mixed layer on the device: step_oml
// (mixed/omlsubs.F90:69-119) with adv_diff (mixed/mixed_layer.F90:16-95), the
// Ekman velocities of ekman_from_tau (coupler/ekman.F90:73-80) and the
// T-point stencils of glam/numerics.F90 (del2_T, map_T_to_x_face,
// map_T_to_y_face, c_grid_advection, map_T_to_P).
//
// Two launches per step:
//   k_oml_step    one thread per T point: the new mixed-layer temperature,
//                 its convective correction and the layer 1/2 entrainment at
//                 T points, plus per-block partial sums of the latter;
//   k_oml_finish  one thread per p point: the mean of the entrainment from
//                 the partials (block order), the entrainment mapped to p
//                 points into the basin's resident forcing, and the time
//                 levels shifted (sstm <- sst, sst <- new) by copies, so the
//                 buffers keep their roles in replayed graphs.
// T fields are stored like p fields (rows of the basin's pitch), nyt rows.

#include "oqg_device.h"

namespace oqg {

// A T field value at (i, j), 0-based, i wrapped in a channel.
struct TField {
    const double* f;
    long long pitch;
    int nxt;
    int cyclic;
    __device__ __forceinline__ double operator()(int i, int j) const {
        if (cyclic) i = (i < 0) ? i + nxt : (i >= nxt ? i - nxt : i);
        return f[(long long)j * pitch + i];
    }
};

// del2_T (numerics.F90:107-147) at T point (i, j) of a field given by `t`
// (a callable): zero outer differences on no-flux walls, the difference to
// the prescribed value on flux walls, periodic in x in a channel.
template <class F>
__device__ __forceinline__ double del2T_at(const F& t, int i, int j, int nxt, int nyt, int cyclic, double f, int nflux,
                                           double nval, int sflux, double sval) {
    const double c = t(i, j);
    double dxl, dxr, dyl, dyr;
    if (i >= 1) dxl = c - t(i - 1, j);
    else dxl = cyclic ? t(0, j) - t(nxt - 1, j) : 0.0;
    if (i <= nxt - 2) dxr = t(i + 1, j) - c;
    else dxr = cyclic ? t(0, j) - t(nxt - 1, j) : 0.0;
    if (j >= 1) dyl = c - t(i, j - 1);
    else dyl = sflux ? c - sval : 0.0;
    if (j <= nyt - 2) dyr = t(i, j + 1) - c;
    else dyr = nflux ? nval - c : 0.0;
    return f * (dxr - dxl + dyr - dyl);
}

__global__ __launch_bounds__(256) void k_oml_step(OmlArgs a) {
    const Geo& g = a.g;
    const int nxt = g.nxp - 1, nyt = g.nyp - 1, nxp = g.nxp, nyp = g.nyp;
    // 1-D blocks of 256 threads on 64 x 4 tiles of T points (block_sum takes 1-D blocks)
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const long long pitch = g.pitch;
    const bool in = i < nxt && j < nyt;
    double xv = 0.0;
    if (in) {
        const TField dm{a.sstm, pitch, nxt, g.cyclic}, dn{a.sst, pitch, nxt, g.cyclic};
        auto P = [&](int ii, int jj) { return a.p1[(long long)jj * pitch + ii]; };
        auto TX = [&](int ii, int jj) { return a.taux[(long long)jj * pitch + ii]; };
        auto TY = [&](int ii, int jj) { return a.tauy[(long long)jj * pitch + ii]; };
        const double f = 1.0 / (g.dx * g.dx);
        // adv_diff's diffusion: Del^2 of sstm with the flux walls, Del^4 with zero-gradient walls
        const double d2 = del2T_at(dm, i, j, nxt, nyt, g.cyclic, f, a.fluxn, a.nval, a.fluxs, a.sval);
        auto D2 = [&](int ii, int jj) {
            if (g.cyclic) ii = (ii < 0) ? ii + nxt : (ii >= nxt ? ii - nxt : ii);
            return del2T_at(dm, ii, jj, nxt, nyt, g.cyclic, f, a.fluxn, a.nval, a.fluxs, a.sval);
        };
        const double d4 = del2T_at(D2, i, j, nxt, nyt, g.cyclic, f, 0, 0.0, 0, 0.0);
        // the advective fluxes through the cell's four faces (adv_vel, mixed_layer.F90:60-93)
        const double cu = a.tau_sign / (g.fnot * a.hm), cv = -a.tau_sign / (g.fnot * a.hm);
        double ut[2], vt[2];
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const int I = i + s;   // x face, 0..nxt
            const double dpdy = (1.0 / g.dy) * (P(I, j + 1) - P(I, j));
            const double uek = cu * (0.5 * (TY(I, j + 1) + TY(I, j)));
            double vu = (-a.ycexp / g.fnot) * dpdy + uek;
            if (!g.cyclic && (I == 0 || I == nxp - 1)) vu = 0.0;
            double xf;
            if (I == 0 || I == nxt) xf = g.cyclic ? 0.5 * (dn(0, j) + dn(nxt - 1, j)) : dn(I == 0 ? 0 : nxt - 1, j);
            else xf = 0.5 * (dn(I - 1, j) + dn(I, j));
            ut[s] = vu * xf;
            const int J = j + s;   // y face, 0..nyt
            const double dpdx = (1.0 / g.dx) * (P(i + 1, J) - P(i, J));
            const double vek = cv * (0.5 * (TX(i + 1, J) + TX(i, J)));
            double vv = (a.ycexp / g.fnot) * dpdx + vek;
            if (J == 0) vv = a.fluxs ? vek : 0.0;
            if (J == nyp - 1) vv = a.fluxn ? vek : 0.0;
            double yf;
            if (J == 0) yf = a.fluxs ? dn(i, 0) : 0.5 * (dn(i, 0) + a.sval);
            else if (J == nyt) yf = a.fluxn ? dn(i, nyt - 1) : 0.5 * (dn(i, nyt - 1) + a.nval);
            else yf = 0.5 * (dn(i, J - 1) + dn(i, J));
            vt[s] = vv * yf;
        }
        const double adv = (-1.0 / g.dx) * (ut[1] - ut[0]) + (-1.0 / g.dy) * (vt[1] - vt[0]);
        const double ad = adv + a.d2 * d2 - a.d4 * d4;
        // Ekman pumping at the T point (ekman.F90:78-79)
        const double tyr = 0.5 * (TY(i + 1, j + 1) + TY(i + 1, j)), tyl = 0.5 * (TY(i, j + 1) + TY(i, j));
        const double txt = 0.5 * (TX(i + 1, j + 1) + TX(i, j + 1)), txb = 0.5 * (TX(i + 1, j) + TX(i, j));
        const double wekt = (1.0 / (g.dx * g.fnot)) * (tyr - tyl) - (1.0 / (g.dy * g.fnot)) * (txt - txb);
        // step_oml (omlsubs.F90:92-108)
        const long long o = (long long)j * pitch + i;
        const double sm = a.sstm[o];
        const double half = 0.5 * wekt * (sm - a.T1_rel);
        const double snew = sm + a.tdt * (ad + (1.0 / (a.rho_cp * a.hm)) * a.fnet[o] + (1.0 / a.hm) * (wekt * sm - half));
        const double dtc = fmax(0.0, a.T1_rel - snew);
        xv = (-1.0 / a.dT_12) * half - (a.hm / (a.tdt * a.dT_12)) * dtc;
        a.snew[o] = snew + dtc;
        a.xfo[o] = xv;
    }
    // per-block partial sum of the entrainment at T points
    __shared__ double red[16];
    double v[1] = {xv};
    block_sum<1>(v, red);
    if (threadIdx.x == 0) a.part[blockIdx.y * gridDim.x + blockIdx.x] = v[0];
}

__global__ __launch_bounds__(256) void k_oml_finish(OmlArgs a) {
    const Geo& g = a.g;
    const int nxt = g.nxp - 1, nyt = g.nyp - 1, nxp = g.nxp, nyp = g.nyp;
    const long long pitch = g.pitch;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = t / nxp, i = t - j * nxp;
    const bool in = j < nyp;
    if (a.compute_ent) {
        // avg_T (numerics.F90:155-162) from the partials, in block order
        __shared__ double red[16];
        double s = 0.0;
        for (int b = threadIdx.x; b < a.npart; b += blockDim.x) s += a.part[b];
        double v[1] = {s};
        block_sum<1>(v, red);
        const double avg = v[0] / (double)(nxt * nyt);
        if (in) {
            auto X = [&](int ii, int jj) { return a.xfo[(long long)jj * pitch + ii] - avg; };
            // map_T_to_P (numerics.F90:64-107) of xfo - avg
            double e;
            const bool xin = i >= 1 && i <= nxp - 2, yin = j >= 1 && j <= nyp - 2;
            if (xin && yin) {
                e = 0.25 * (X(i, j) + X(i - 1, j) + X(i, j - 1) + X(i - 1, j - 1));
            } else if (xin) {
                const int jt = (j == 0) ? 0 : nyt - 1;
                e = 0.5 * (X(i - 1, jt) + X(i, jt));
            } else if (yin) {
                if (g.cyclic) e = 0.25 * (X(0, j) + X(nxt - 1, j) + X(0, j - 1) + X(nxt - 1, j - 1));
                else {
                    const int it = (i == 0) ? 0 : nxt - 1;
                    e = 0.5 * (X(it, j - 1) + X(it, j));
                }
            } else {
                const int jt = (j == 0) ? 0 : nyt - 1;
                if (g.cyclic) e = 0.5 * (X(0, jt) + X(nxt - 1, jt));
                else e = X(i == 0 ? 0 : nxt - 1, jt);
            }
            a.ent1[(long long)j * pitch + i] = e;
        }
    }
    // the time levels: sstm <- sst, sst <- new
    if (in && i < nxt && j < nyt) {
        const long long o = (long long)j * pitch + i;
        const double cur = a.sst[o];
        a.sstm[o] = cur;
        a.sst[o] = a.snew[o];
    }
}

// suppress_ocean's mixed-layer part (driver.F90:165-167)
__global__ __launch_bounds__(256) void k_oml_suppress(OmlArgs a) {
    const int nxt = a.g.nxp - 1, nyt = a.g.nyp - 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = t / nxt, i = t - j * nxt;
    if (j < nyt) {
        const long long o = (long long)j * a.g.pitch + i;
        a.sst[o] = 0.5 * (a.sst[o] + a.sstm[o]);
    }
}

dim3 oml_step_grid(const Geo& g) { return dim3((g.nxp - 1 + 63) / 64, (g.nyp - 1 + 3) / 4); }

void launch_oml_step(const OmlArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_oml_step, oml_step_grid(a.g), dim3(256), 0, st, a);
}

void launch_oml_finish(const OmlArgs& a, hipStream_t st) {
    const int n = a.g.nxp * a.g.nyp;
    hipLaunchKernelGGL(k_oml_finish, dim3((n + 255) / 256), dim3(256), 0, st, a);
}

void launch_oml_suppress(const OmlArgs& a, hipStream_t st) {
    const int n = (a.g.nxp - 1) * (a.g.nyp - 1);
    hipLaunchKernelGGL(k_oml_suppress, dim3((n + 255) / 256), dim3(256), 0, st, a);
}

}  // namespace oqg

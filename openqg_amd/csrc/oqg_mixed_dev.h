// oqg_mixed_dev.h -- device stencils of the mixed layers on T points, shared by
// the ocean mixed layer (oqg_mixed.hip) and the coupling (oqg_couple.hip):
// del2_T (glam/numerics.F90:107-147) and adv_diff (mixed/mixed_layer.F90:16-95)
// with map_T_to_x_face, map_T_to_y_face and c_grid_advection
// (numerics.F90:222-266, 392-408) and the Ekman velocities of ekman_from_tau
// (coupler/ekman.F90:73-77), evaluated per point.
#pragma once
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

// del2_T at T point (i, j) of a field given by `t` (a callable): zero outer
// differences on no-flux walls, the difference to the prescribed value on flux
// walls, periodic in x in a channel.
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

// adv_diff of a mixed-layer field at T point (i, j): the advective flux
// divergence through the cell's four faces plus d2 Del^2 - d4 Del^4 of the
// previous level.  dn, dm: the current and previous levels (T accessors,
// absolute 0-based indices, wrapped in a channel); P, TX, TY: layer-1 p and
// the stress (p accessors).
template <class TA, class PA>
__device__ __forceinline__ double adv_diff_acc(const Geo& g, const AdvField& m, const TA& dn, const TA& dm, const PA& P,
                                               const PA& TX, const PA& TY, double hm, int tau_sign, int i, int j) {
    const int nxt = g.nxp - 1, nyt = g.nyp - 1, nxp = g.nxp, nyp = g.nyp;
    const double f = 1.0 / (g.dx * g.dx);
    const double d2 = del2T_at(dm, i, j, nxt, nyt, g.cyclic, f, m.fluxn, m.nval, m.fluxs, m.sval);
    auto D2 = [&](int ii, int jj) {
        if (g.cyclic) ii = (ii < 0) ? ii + nxt : (ii >= nxt ? ii - nxt : ii);
        return del2T_at(dm, ii, jj, nxt, nyt, g.cyclic, f, m.fluxn, m.nval, m.fluxs, m.sval);
    };
    const double d4 = del2T_at(D2, i, j, nxt, nyt, g.cyclic, f, 0, 0.0, 0, 0.0);
    // the advective fluxes through the cell's four faces (adv_vel, mixed_layer.F90:60-93)
    const double cu = tau_sign / (g.fnot * hm), cv = -tau_sign / (g.fnot * hm);
    double ut[2], vt[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const int I = i + s;   // x face, 0..nxt
        const double dpdy = (1.0 / g.dy) * (P(I, j + 1) - P(I, j));
        const double uek = cu * (0.5 * (TY(I, j + 1) + TY(I, j)));
        double vu = (-m.ycexp / g.fnot) * dpdy + uek;
        if (!g.cyclic && (I == 0 || I == nxp - 1)) vu = 0.0;
        double xf;
        if (I == 0 || I == nxt) xf = g.cyclic ? 0.5 * (dn(0, j) + dn(nxt - 1, j)) : dn(I == 0 ? 0 : nxt - 1, j);
        else xf = 0.5 * (dn(I - 1, j) + dn(I, j));
        ut[s] = vu * xf;
        const int J = j + s;   // y face, 0..nyt
        const double dpdx = (1.0 / g.dx) * (P(i + 1, J) - P(i, J));
        const double vek = cv * (0.5 * (TX(i + 1, J) + TX(i, J)));
        double vv = (m.ycexp / g.fnot) * dpdx + vek;
        if (J == 0) vv = m.fluxs ? vek : 0.0;
        if (J == nyp - 1) vv = m.fluxn ? vek : 0.0;
        double yf;
        if (J == 0) yf = m.fluxs ? dn(i, 0) : 0.5 * (dn(i, 0) + m.sval);
        else if (J == nyt) yf = m.fluxn ? dn(i, nyt - 1) : 0.5 * (dn(i, nyt - 1) + m.nval);
        else yf = 0.5 * (dn(i, J - 1) + dn(i, J));
        vt[s] = vv * yf;
    }
    const double adv = (-1.0 / g.dx) * (ut[1] - ut[0]) + (-1.0 / g.dy) * (vt[1] - vt[0]);
    return adv + m.d2 * d2 - m.d4 * d4;
}

// adv_diff_acc with the Del^2 of the previous level given by an accessor D2
// (del2_T of datam with the field's walls, at absolute T indices): its value at
// (i, j) and its Del^2 with zero-gradient walls make the diffusion terms.
template <class TA, class DA, class PA>
__device__ __forceinline__ double adv_diff_d2(const Geo& g, const AdvField& m, const TA& dn, const DA& D2, const PA& P,
                                              const PA& TX, const PA& TY, double hm, int tau_sign, int i, int j) {
    const int nxt = g.nxp - 1, nyt = g.nyp - 1, nxp = g.nxp, nyp = g.nyp;
    const double f = 1.0 / (g.dx * g.dx);
    const double d2 = D2(i, j);
    const double d4 = del2T_at(D2, i, j, nxt, nyt, g.cyclic, f, 0, 0.0, 0, 0.0);
    const double cu = tau_sign / (g.fnot * hm), cv = -tau_sign / (g.fnot * hm);
    double ut[2], vt[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const int I = i + s;
        const double dpdy = (1.0 / g.dy) * (P(I, j + 1) - P(I, j));
        const double uek = cu * (0.5 * (TY(I, j + 1) + TY(I, j)));
        double vu = (-m.ycexp / g.fnot) * dpdy + uek;
        if (!g.cyclic && (I == 0 || I == nxp - 1)) vu = 0.0;
        double xf;
        if (I == 0 || I == nxt) xf = g.cyclic ? 0.5 * (dn(0, j) + dn(nxt - 1, j)) : dn(I == 0 ? 0 : nxt - 1, j);
        else xf = 0.5 * (dn(I - 1, j) + dn(I, j));
        ut[s] = vu * xf;
        const int J = j + s;
        const double dpdx = (1.0 / g.dx) * (P(i + 1, J) - P(i, J));
        const double vek = cv * (0.5 * (TX(i + 1, J) + TX(i, J)));
        double vv = (m.ycexp / g.fnot) * dpdx + vek;
        if (J == 0) vv = m.fluxs ? vek : 0.0;
        if (J == nyp - 1) vv = m.fluxn ? vek : 0.0;
        double yf;
        if (J == 0) yf = m.fluxs ? dn(i, 0) : 0.5 * (dn(i, 0) + m.sval);
        else if (J == nyt) yf = m.fluxn ? dn(i, nyt - 1) : 0.5 * (dn(i, nyt - 1) + m.nval);
        else yf = 0.5 * (dn(i, J - 1) + dn(i, J));
        vt[s] = vv * yf;
    }
    const double adv = (-1.0 / g.dx) * (ut[1] - ut[0]) + (-1.0 / g.dy) * (vt[1] - vt[0]);
    return adv + m.d2 * d2 - m.d4 * d4;
}

struct PGlobal {   // a pitched p field in global memory
    const double* f;
    long long pitch;
    __device__ __forceinline__ double operator()(int i, int j) const { return f[(long long)j * pitch + i]; }
};

// adv_diff_acc on fields in global memory.
__device__ __forceinline__ double adv_diff_at(const Geo& g, const AdvField& m, const double* p1, const double* taux,
                                              const double* tauy, double hm, int tau_sign, int i, int j) {
    const long long pitch = g.pitch;
    const int nxt = g.nxp - 1;
    const TField dm{m.datam, pitch, nxt, g.cyclic}, dn{m.data, pitch, nxt, g.cyclic};
    return adv_diff_acc(g, m, dn, dm, PGlobal{p1, pitch}, PGlobal{taux, pitch}, PGlobal{tauy, pitch}, hm, tau_sign, i,
                        j);
}

// Ekman pumping at T point (i, j) (ekman.F90:78-79) from p accessors of the stress.
template <class PA>
__device__ __forceinline__ double wekt_acc(const Geo& g, const PA& TX, const PA& TY, int i, int j) {
    const double tyr = 0.5 * (TY(i + 1, j + 1) + TY(i + 1, j)), tyl = 0.5 * (TY(i, j + 1) + TY(i, j));
    const double txt = 0.5 * (TX(i + 1, j + 1) + TX(i, j + 1)), txb = 0.5 * (TX(i + 1, j) + TX(i, j));
    return (1.0 / (g.dx * g.fnot)) * (tyr - tyl) - (1.0 / (g.dy * g.fnot)) * (txt - txb);
}

__device__ __forceinline__ double wekt_at(const Geo& g, const double* taux, const double* tauy, int i, int j) {
    const long long pitch = g.pitch;
    auto TX = [&](int ii, int jj) { return taux[(long long)jj * pitch + ii]; };
    auto TY = [&](int ii, int jj) { return tauy[(long long)jj * pitch + ii]; };
    const double tyr = 0.5 * (TY(i + 1, j + 1) + TY(i + 1, j)), tyl = 0.5 * (TY(i, j + 1) + TY(i, j));
    const double txt = 0.5 * (TX(i + 1, j + 1) + TX(i, j + 1)), txb = 0.5 * (TX(i + 1, j) + TX(i, j));
    return (1.0 / (g.dx * g.fnot)) * (tyr - tyl) - (1.0 / (g.dy * g.fnot)) * (txt - txb);
}

// map_T_to_P (numerics.F90:64-107) at p point (i, j) of the T field X (callable).
template <class F>
__device__ __forceinline__ double map_T_to_P_at(const F& X, int i, int j, int nxp, int nyp, int cyclic) {
    const int nxt = nxp - 1, nyt = nyp - 1;
    const bool xin = i >= 1 && i <= nxp - 2, yin = j >= 1 && j <= nyp - 2;
    if (xin && yin) return 0.25 * (X(i, j) + X(i - 1, j) + X(i, j - 1) + X(i - 1, j - 1));
    if (xin) {
        const int jt = (j == 0) ? 0 : nyt - 1;
        return 0.5 * (X(i - 1, jt) + X(i, jt));
    }
    if (yin) {
        if (cyclic) return 0.25 * (X(0, j) + X(nxt - 1, j) + X(0, j - 1) + X(nxt - 1, j - 1));
        const int it = (i == 0) ? 0 : nxt - 1;
        return 0.5 * (X(it, j - 1) + X(it, j));
    }
    const int jt = (j == 0) ? 0 : nyt - 1;
    if (cyclic) return 0.5 * (X(0, jt) + X(nxt - 1, jt));
    return X(i == 0 ? 0 : nxt - 1, jt);
}

}  // namespace oqg

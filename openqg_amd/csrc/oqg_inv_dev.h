// oqg_inv_dev.h -- device code shared by the inverse kernels of oqg_transform.hip and
// oqg_wave.hip: the row-pair placement and the per-step correction coefficients.
#pragma once

#include "oqg_device.h"

namespace oqg {

// Row pair of workgroup b of n in the step kernels: with n a multiple of 8,
// XCD x takes row pairs [x n/8, (x+1) n/8), the rows that k_vort's tiles on
// the same XCD cover, so q_new (read by k_fwd) and p_new (read by the next
// k_vort) mostly stay in the L2 that wrote them.
__device__ __forceinline__ int xcd_row_pair(int b, int n) { return (n % NXCD == 0) ? xcd_contig(b, n) : b; }

// LU solve of the (nl-1)^2 mass-constraint system with DGERFS-style iterative
// refinement (homog.F90:356-359, linalg.F90:57-93); N = nl-1 is a compile-time
// size so the vectors stay in registers.
template <int N>
__device__ void small_lu_solve(const BoxHom& h, const double (&rhs)[N], double (&x)[N]) {
    auto lusolve = [&](double (&v)[N]) {
#pragma unroll
        for (int k = 0; k < N; k++) {
            const int p = h.ipiv[k];
#pragma unroll
            for (int q = k + 1; q < N; q++)
                if (q == p) { const double t = v[k]; v[k] = v[q]; v[q] = t; }
        }
#pragma unroll
        for (int i = 0; i < N; i++)
#pragma unroll
            for (int j = 0; j < i; j++) v[i] -= h.LU[i * MAXL + j] * v[j];
#pragma unroll
        for (int i = N - 1; i >= 0; i--) {
#pragma unroll
            for (int j = i + 1; j < N; j++) v[i] -= h.LU[i * MAXL + j] * v[j];
            v[i] /= h.LU[i * MAXL + i];
        }
    };
#pragma unroll
    for (int i = 0; i < N; i++) x[i] = rhs[i];
    lusolve(x);
    double lstres = 3.0;
    const double eps = 1.1102230246251565e-16;
    for (int it = 0; it < 5; it++) {
        double r[N], berr = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++) {
            double sres = rhs[i], den = fabs(rhs[i]);
#pragma unroll
            for (int j = 0; j < N; j++) {
                sres -= h.cdhoc[i * MAXL + j] * x[j];
                den += fabs(h.cdhoc[i * MAXL + j]) * fabs(x[j]);
            }
            r[i] = sres;
            if (den > 0.0 && fabs(sres) / den > berr) berr = fabs(sres) / den;
        }
        if (!(berr > eps && 2.0 * berr <= lstres)) break;
        lusolve(r);
#pragma unroll
        for (int i = 0; i < N; i++) x[i] += r[i];
        lstres = berr;
    }
}

// Per-step correction coefficients, computed by thread 0 of every k_inv block
// from the block-reduced integrals v[0..nl-1] (box: spectral column sums) and
// v[NL] (entrainment).  Every input sits in LDS: the precompute blocks, the
// constraint state ci, the channel's boundary integrals x and the k = 0
// spectra of the first and last interior rows spk[2m], spk[2m+1].  Writes
// coef[] and, in block 0, the new constraint state.
template <int NL>
__device__ void step_coefficients(const InvArgs& a, const Mix& mx, const BoxHom& bh, const CycHom& ch, const Phys& ph,
                                  const double* ci, const double* x, const double* spk, const double (&v)[NL + 1],
                                  double* coef, bool write_state) {
    const Geo& g = a.g;
    double dpi[NL], dpip[NL], cs[NL], cn[NL], csp[NL], cnp[NL];
#pragma unroll
    for (int l = 0; l < NL; l++) {
        dpi[l] = ci[C_DPI + l]; dpip[l] = ci[C_DPIP + l];
        cs[l] = ci[C_CS + l]; cn[l] = ci[C_CN + l]; csp[l] = ci[C_CSP + l]; cnp[l] = ci[C_CNP + l];
    }
    // update_mass_constr (constraint.F90:167-179) with ent_xn(1) = int_P_dA(ent1)
    const double exn = (v[NL] * g.dx) * g.dy;
#pragma unroll
    for (int l = 0; l < NL - 1; l++) {
        const double e = (l == 0) ? exn : 0.0;
        const double old = dpi[l];
        dpi[l] = dpip[l] - a.tdt * ph.gp[l] * e;
        dpip[l] = old;
    }
    if (!g.cyclic) {
        // box_homog (homog.F90:345-364); modal area integrals from the spectral column sums
        double im[NL], ik[NL], rhs[NL - 1], alpha[NL - 1];
#pragma unroll
        for (int m = 0; m < NL; m++) im[m] = (v[m] * g.dx) * g.dy;
#pragma unroll
        for (int k = 0; k < NL; k++) {
            double t = 0.0;
#pragma unroll
            for (int m = 0; m < NL; m++) t += mx.ctm2l[m * MAXL + k] * im[m];
            ik[k] = t;
        }
#pragma unroll
        for (int l = 0; l < NL - 1; l++) rhs[l] = dpi[l] - (ik[l + 1] - ik[l]);
        small_lu_solve<NL - 1>(bh, rhs, alpha);
#pragma unroll
        for (int l = 0; l < NL - 1; l++) coef[l] = alpha[l];
    } else {
        // step_constraints (constraint.F90:182-215)
        const double f0 = g.fnot, dy = g.dy;
        double rhss[NL], rhsn[NL];
#pragma unroll
        for (int k = 0; k < NL; k++) {
            rhss[k] = -0.5 * dy * f0 * x[A_BES + k];
            rhsn[k] = -0.5 * dy * f0 * x[A_BEN + k];
        }
        rhss[0] = rhss[0] + ph.tau_sign * ph.fohfac[0] * x[A_TXIS];
        rhsn[0] = rhsn[0] - ph.tau_sign * ph.fohfac[0] * x[A_TXIN];
        rhss[NL - 1] = rhss[NL - 1] + ph.fohfac[NL - 1] * x[A_BDRS];
        rhsn[NL - 1] = rhsn[NL - 1] - ph.fohfac[NL - 1] * x[A_BDRN];
#pragma unroll
        for (int k = 0; k < NL; k++) {
            rhss[k] = rhss[k] + x[A_AJIS + k] - x[A_AP3S + k] + x[A_AP5S + k];
            rhsn[k] = rhsn[k] + x[A_AJIN + k] + x[A_AP3N + k] - x[A_AP5N + k];
        }
#pragma unroll
        for (int k = 0; k < NL; k++) {
            const double csnew = csp[k] + a.tdt * rhss[k];
            const double cnnew = cnp[k] + a.tdt * rhsn[k];
            csp[k] = cs[k];
            cnp[k] = cn[k];
            cs[k] = csnew;
            cn[k] = cnnew;
        }
        // cyclic_homog (homog.F90:305-320); row integrals from the k = 0 spectra
        double clhss[NL], clhsn[NL];
#pragma unroll
        for (int m = 0; m < NL; m++) {
            double ts = 0.0, tn = 0.0;
#pragma unroll
            for (int k = 0; k < NL; k++) {
                ts += mx.ctl2m[k * MAXL + m] * cs[k];
                tn += mx.ctl2m[k * MAXL + m] * cn[k];
            }
            clhss[m] = ts + (ch.c0 * spk[2 * m]) * ph.rdy;
            clhsn[m] = tn + (ch.c0 * spk[2 * m + 1]) * ph.rdy;
        }
        coef[0] = clhss[0] * ch.hbsi;
#pragma unroll
        for (int m = 1; m < NL; m++) {
            coef[1 + m] = ch.hc2n[m] * clhss[m] - ch.hc2s[m] * clhsn[m];
            coef[1 + MAXL + m] = ch.hc1s[m] * clhsn[m] - ch.hc1n[m] * clhss[m];
        }
    }
    if (write_state) {
#pragma unroll
        for (int l = 0; l < NL; l++) {
            a.cout[C_DPI + l] = (l < NL - 1) ? dpi[l] : 0.0;
            a.cout[C_DPIP + l] = (l < NL - 1) ? dpip[l] : 0.0;
            a.cout[C_CS + l] = cs[l]; a.cout[C_CN + l] = cn[l];
            a.cout[C_CSP + l] = csp[l]; a.cout[C_CNP + l] = cnp[l];
        }
    }
}

}  // namespace oqg

/*
 * cpuref.c -- CPU oracle for OpenQG's QG timestep (step_qg, src/qg/core.F90:16-33).
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference Fortran,
 * the reference the HIP library's results are compared with, and the timed
 * CPU baseline.
 * Nothing in the product links or calls this file.
 *
 * Each routine cites the reference file:line it restates.  Whole-array
 * Fortran statements are restated as loops that keep the reference's order
 * of floating-point operations within each element; the build disables FMA
 * contraction (-ffp-contract=off) so every operation is one IEEE rounding.
 * Where the reference calls a third-party library the oracle uses its own
 * textbook routine:
 *   FFTW3 RODFT00 / r2c / c2r  -> mixed-radix Cooley-Tukey FFT (below)
 *   LAPACK DGEEVX              -> cyclic Jacobi on the symmetrised A matrix
 *   LAPACK DGETRF/DGETRS/DGERFS -> partial-pivot LU + iterative refinement
 * Those substitutions agree with the libraries to rounding, not bitwise.
 */
#include "cpuref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_D 3.14159265358979324    /* constants.F90: PI */
#define TWOPI_D (2.0 * PI_D)        /* constants.F90: TWOPI */
#define PI_LD 3.141592653589793238462643383279502884L

/* ------------------------------------------------------------------ */
/* FFT: recursive mixed-radix decimation in time over a twiddle table */
/* ------------------------------------------------------------------ */

typedef struct {
    int n;
    int nfac;
    int fac[48];
    double* tw; /* 2n doubles: exp(-2*pi*i*t/n), t = 0..n-1 */
} fftplan;

/* exp(-2*pi*i*t/n), exact at multiples of n/4, long-double accurate elsewhere. */
static void twiddle(long t, long n, double* c, double* s) {
    t %= n;
    if (t < 0) t += n;
    long t4 = 4 * t;
    long q = t4 / n;
    long r = t4 - q * n;
    long double phi = (PI_LD / 2.0L) * (long double)r / (long double)n;
    long double C = cosl(phi), S = sinl(phi);
    double cc, ss;
    switch (q) {
        case 0: cc = (double)C; ss = (double)S; break;
        case 1: cc = (double)-S; ss = (double)C; break;
        case 2: cc = (double)-C; ss = (double)-S; break;
        default: cc = (double)S; ss = (double)-C; break;
    }
    if (r == 0) { /* exact axis values */
        cc = (q == 0) ? 1.0 : (q == 2) ? -1.0 : 0.0;
        ss = (q == 1) ? 1.0 : (q == 3) ? -1.0 : 0.0;
    }
    *c = cc;
    *s = -ss;
}

static int fft_plan_init(fftplan* p, int n) {
    p->n = n;
    p->nfac = 0;
    int m = n;
    while (m % 4 == 0) { p->fac[p->nfac++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[p->nfac++] = 2; m /= 2; }
    for (int f = 3; m > 1; f += 2) {
        while (m % f == 0) { p->fac[p->nfac++] = f; m /= f; }
        if (f > 4096) return -1;
    }
    p->tw = (double*)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1));
    for (int t = 0; t < n; t++) twiddle(t, n, &p->tw[2 * t], &p->tw[2 * t + 1]);
    return 0;
}

static void fft_plan_free(fftplan* p) { free(p->tw); p->tw = NULL; }

/* y = DFT_n(x[0], x[xs], ...), root exp(sign*2*pi*i/n); tw has length N, n | N. */
static void fft_rec(const double* x, size_t xs, double* y, int n, const int* fac,
                    const double* tw, int N, int sign) {
    if (n == 1) { y[0] = x[0]; y[1] = x[1]; return; }
    int p = fac[0], m = n / p;
    for (int r = 0; r < p; r++) fft_rec(x + 2 * r * xs, xs * p, y + 2 * (size_t)r * m, m, fac + 1, tw, N, sign);
    int step = N / n;
    double tr[p], ti[p];
    double sg = (sign > 0) ? -1.0 : 1.0; /* conjugate table for the inverse */
    for (int k = 0; k < m; k++) {
        tr[0] = y[2 * k];
        ti[0] = y[2 * k + 1];
        for (int r = 1; r < p; r++) {
            double a = y[2 * ((size_t)r * m + k)], b = y[2 * ((size_t)r * m + k) + 1];
            long e = ((long)r * k) % n;
            double c = tw[2 * e * step], s = sg * tw[2 * e * step + 1];
            tr[r] = a * c - b * s;
            ti[r] = a * s + b * c;
        }
        if (p == 2) {
            y[2 * k] = tr[0] + tr[1];
            y[2 * k + 1] = ti[0] + ti[1];
            y[2 * (k + m)] = tr[0] - tr[1];
            y[2 * (k + m) + 1] = ti[0] - ti[1];
        } else if (p == 4) {
            /* w4 = -i (forward) or +i (inverse) */
            double a0r = tr[0] + tr[2], a0i = ti[0] + ti[2];
            double a1r = tr[0] - tr[2], a1i = ti[0] - ti[2];
            double b0r = tr[1] + tr[3], b0i = ti[1] + ti[3];
            double b1r = tr[1] - tr[3], b1i = ti[1] - ti[3];
            /* (b1) * (-i*sgn) where sgn = +1 forward */
            double jr, ji;
            if (sign < 0) { jr = b1i; ji = -b1r; } else { jr = -b1i; ji = b1r; }
            y[2 * k] = a0r + b0r;
            y[2 * k + 1] = a0i + b0i;
            y[2 * (k + m)] = a1r + jr;
            y[2 * (k + m) + 1] = a1i + ji;
            y[2 * (k + 2 * m)] = a0r - b0r;
            y[2 * (k + 2 * m) + 1] = a0i - b0i;
            y[2 * (k + 3 * m)] = a1r - jr;
            y[2 * (k + 3 * m) + 1] = a1i - ji;
        } else {
            int pstep = N / p;
            for (int q = 0; q < p; q++) {
                double sr = tr[0], si = ti[0];
                for (int r = 1; r < p; r++) {
                    long e = ((long)r * q) % p;
                    double c = tw[2 * e * pstep], s = sg * tw[2 * e * pstep + 1];
                    sr += tr[r] * c - ti[r] * s;
                    si += tr[r] * s + ti[r] * c;
                }
                y[2 * (k + (size_t)q * m)] = sr;
                y[2 * (k + (size_t)q * m) + 1] = si;
            }
        }
    }
}

static void fft_exec(const fftplan* p, const double* x, double* y, int sign) {
    fft_rec(x, 1, y, p->n, p->fac, p->tw, p->n, sign);
}

/* DST-I (FFTW RODFT00, unnormalised) of one or two real rows of length n,
 * via a complex FFT of length 2(n+1) of the odd extension of x + i*y:
 * Z_K = -i (DST(x)_K + i DST(y)_K). */
static void dst1_pair(const fftplan* p, int n, const double* x, const double* y,
                      double* X, double* Y, double* z, double* Z) {
    int N = n + 1, L = 2 * N;
    z[0] = z[1] = 0.0;
    z[2 * N] = z[2 * N + 1] = 0.0;
    for (int i = 1; i <= n; i++) {
        double xr = x[i - 1], yi = y ? y[i - 1] : 0.0;
        z[2 * i] = xr;
        z[2 * i + 1] = yi;
        z[2 * (L - i)] = -xr;
        z[2 * (L - i) + 1] = -yi;
    }
    fft_exec(p, z, Z, -1);
    for (int K = 1; K <= n; K++) {
        X[K - 1] = -Z[2 * K + 1];
        if (Y) Y[K - 1] = Z[2 * K];
    }
}

/* r2c of one or two real rows of length n (FFTW r2c convention). */
static void rfft_pair(const fftplan* p, int n, const double* x, const double* y,
                      double* X, double* Y, double* z, double* Z) {
    for (int i = 0; i < n; i++) {
        z[2 * i] = x[i];
        z[2 * i + 1] = y ? y[i] : 0.0;
    }
    fft_exec(p, z, Z, -1);
    for (int k = 0; k <= n / 2; k++) {
        int kk = (n - k) % n;
        double ar = Z[2 * k], ai = Z[2 * k + 1], br = Z[2 * kk], bi = Z[2 * kk + 1];
        X[2 * k] = 0.5 * (ar + br);
        X[2 * k + 1] = 0.5 * (ai - bi);
        if (Y) {
            Y[2 * k] = 0.5 * (ai + bi);
            Y[2 * k + 1] = 0.5 * (br - ar);
        }
    }
}

/* c2r of one or two half spectra (n/2+1 complex each), unnormalised; the
 * imaginary parts of the k = 0 and (even n) k = n/2 terms are ignored, as in FFTW. */
static void irfft_pair(const fftplan* p, int n, const double* X, const double* Y,
                       double* x, double* y, double* z, double* Z) {
    int h = n / 2;
    for (int k = 0; k < n; k++) {
        double xr, xi, yr, yi;
        if (k <= h) {
            xr = X[2 * k]; xi = X[2 * k + 1];
            yr = Y ? Y[2 * k] : 0.0; yi = Y ? Y[2 * k + 1] : 0.0;
        } else {
            int kk = n - k;
            xr = X[2 * kk]; xi = -X[2 * kk + 1];
            yr = Y ? Y[2 * kk] : 0.0; yi = Y ? -Y[2 * kk + 1] : 0.0;
        }
        if (k == 0 || (n % 2 == 0 && k == h)) { xi = 0.0; yi = 0.0; }
        z[2 * k] = xr - yi;
        z[2 * k + 1] = xi + yr;
    }
    fft_exec(p, z, Z, +1);
    for (int i = 0; i < n; i++) {
        x[i] = Z[2 * i];
        if (y) y[i] = Z[2 * i + 1];
    }
}

/* ------------------------------------------------------------------ */
/* Basin                                                              */
/* ------------------------------------------------------------------ */

struct cr_basin {
    cr_params P;
    int nxp, nyp, nl;
    size_t nxy;
    /* box.F90:126-163 derived values */
    double xl, yl, dxm2;
    double* yp;
    double* yprel;
    double* ddyn; /* topog.F90:105, (nxp,nyp) */
    /* modes.F90 */
    double amat[CR_MAXL * CR_MAXL], ctl2m[CR_MAXL * CR_MAXL], ctm2l[CR_MAXL * CR_MAXL];
    double rdm2[CR_MAXL];
    /* inhomog.F90:41-62 */
    int nx_spec, nfft, n_in;
    double a, ftnorm;
    double* bb; /* (nx_spec, nl) */
    fftplan plan;
    /* homog.F90:15-61 */
    double* hom_sol_bc; /* box (nxp,nyp,2:nl) */
    double cdhoc[CR_MAXL * CR_MAXL], LU[CR_MAXL * CR_MAXL];
    int ipiv[CR_MAXL];
    double* hom_bt;                 /* cyclic (nyp) */
    double *hom_bc1, *hom_bc2;      /* cyclic (nyp,2:nl) */
    double hc1s[CR_MAXL], hc2s[CR_MAXL], hc1n[CR_MAXL], hc2n[CR_MAXL], hbsi;
    /* state */
    double *q, *qm, *p, *pm;
    /* constraint.F90:22-67 */
    double dpi[CR_MAXL], dpip[CR_MAXL];
    double cs[CR_MAXL], cn[CR_MAXL], csp[CR_MAXL], cnp[CR_MAXL];
    double ajis[CR_MAXL], ajin[CR_MAXL], ap3s[CR_MAXL], ap3n[CR_MAXL], ap5s[CR_MAXL], ap5n[CR_MAXL];
    double int_Be_s[CR_MAXL], int_Be_n[CR_MAXL];
    double bdrins, bdrinn, txis, txin;
    /* work */
    double *w1, *w2, *w3, *w4, *w5, *w6, *w7;
};

#define NXP (b->nxp)
#define NYP (b->nyp)
/* 1-based Fortran-style element access */
#define E2(a, i, j) (a)[((size_t)(j) - 1) * (size_t)NXP + ((i) - 1)]
#define E3(a, i, j, k) (a)[(((size_t)(k) - 1) * (size_t)NYP + ((j) - 1)) * (size_t)NXP + ((i) - 1)]
#define M2(a, i, j) (a)[((j) - 1) * CR_MAXL + ((i) - 1)]

static int g_threads = 0;
void cr_set_threads(int n) { g_threads = n; }
#ifdef _OPENMP
#define OMP_THREADS() (g_threads > 0 ? g_threads : omp_get_max_threads())
#else
#define OMP_THREADS() 1
#endif

static double* dalloc(size_t n) { return (double*)calloc(n > 0 ? n : 1, sizeof(double)); }

/* ------------------------------------------------------------------ */
/* numerics.F90                                                       */
/* ------------------------------------------------------------------ */

/* int_P_dx (numerics.F90:206-212) */
double cr_int_P_dx(const cr_basin* b, const double* row) {
    double s = 0.0;
    for (int i = 2; i <= NXP - 1; i++) s += row[i - 1];
    return (0.5 * row[0] + s + 0.5 * row[NXP - 1]) * b->P.dx;
}

/* int_P_dA_2d (numerics.F90:188-204) */
double cr_int_P_dA(const cr_basin* b, const double* f) {
    double all = 0.0, s1 = 0.0, sn = 0.0, w1 = 0.0, wn = 0.0;
    for (int j = 1; j <= NYP; j++)
        for (int i = 1; i <= NXP; i++) all += E2(f, i, j);
    for (int i = 1; i <= NXP; i++) { s1 += E2(f, i, 1); sn += E2(f, i, NYP); }
    for (int j = 1; j <= NYP; j++) { w1 += E2(f, 1, j); wn += E2(f, NXP, j); }
    double r = all;
    r = r - 0.5 * (s1 + sn);
    r = r - 0.5 * (w1 + wn);
    r = r + 0.25 * (E2(f, 1, 1) + E2(f, 1, NYP) + E2(f, NXP, 1) + E2(f, NXP, NYP));
    return r * b->P.dx * b->P.dy;
}

/* del2_P_bc (numerics.F90:410-425) = dP2dx2_bc (:427-448) + dP2dy2_bc (:450-468). */
static void del2_layer(const cr_basin* b, const double* f, double* out) {
    const double al = b->P.bcco;
    const double fx = 1.0 / (b->P.dx * b->P.dx), fy = 1.0 / (b->P.dy * b->P.dy);
    const double cxl = fx * al / (0.5 * al + 1.0), cxr = -fx * al / (0.5 * al + 1.0);
    const double cyl = fy * al / (0.5 * al + 1.0), cyr = -fy * al / (0.5 * al + 1.0);
    const int cyc = b->P.cyclic;
#pragma omp parallel for num_threads(OMP_THREADS()) schedule(static)
    for (int j = 1; j <= NYP; j++) {
        for (int i = 1; i <= NXP; i++) {
            double dx2, dy2;
            /* dP2dx2_bc */
            if (j == 1 || j == NYP) {
                dx2 = 0.0;
            } else if (i >= 2 && i <= NXP - 1) {
                dx2 = fx * (E2(f, i - 1, j) + E2(f, i + 1, j) - 2.0 * E2(f, i, j));
            } else if (!cyc) {
                dx2 = (i == 1) ? cxl * (E2(f, 2, j) - E2(f, 1, j))
                               : cxr * (E2(f, NXP, j) - E2(f, NXP - 1, j));
            } else {
                dx2 = fx * (E2(f, NXP - 1, j) + E2(f, 2, j) - 2.0 * E2(f, 1, j));
            }
            /* dP2dy2_bc */
            if (!cyc && (i == 1 || i == NXP)) {
                dy2 = 0.0;
            } else if (j >= 2 && j <= NYP - 1) {
                dy2 = fy * (E2(f, i, j - 1) + E2(f, i, j + 1) - 2.0 * E2(f, i, j));
            } else if (j == 1) {
                dy2 = cyl * (E2(f, i, 2) - E2(f, i, 1));
            } else {
                dy2 = cyr * (E2(f, i, NYP) - E2(f, i, NYP - 1));
            }
            E2(out, i, j) = dx2 + dy2;
        }
    }
}

void cr_del2_bc(const cr_basin* b, const double* in, double* out, int nlayers) {
    for (int k = 0; k < nlayers; k++) del2_layer(b, in + (size_t)k * b->nxy, out + (size_t)k * b->nxy);
}

/* dPdx_2 (numerics.F90:311-325) evaluated at (i,j) of field f */
static inline double dpdx2_at(const cr_basin* b, const double* f, int i, int j) {
    if (i >= 2 && i <= NXP - 1) return E2(f, i + 1, j) - E2(f, i - 1, j);
    if (!b->P.cyclic) return 0.0;
    return E2(f, 2, j) - E2(f, NXP - 1, j); /* i == 1, and i == nxp copies i == 1 */
}
/* dPdy_2 (numerics.F90:327-338) */
static inline double dpdy2_at(const cr_basin* b, const double* f, int i, int j) {
    if (j >= 2 && j <= NYP - 1) return E2(f, i, j + 1) - E2(f, i, j - 1);
    return 0.0;
}

/* jacobian (numerics.F90:340-388) for one layer; returns ajis/ajin for it. */
static void jacobian_layer(const cr_basin* b, const double* p, const double* q, double* jac,
                           double* ajis, double* ajin, double* dpx, double* dpy,
                           double* dqx, double* dqy, double* F) {
    const double fac = 1.0 / (3.0 * (2.0 * b->P.dx) * (2.0 * b->P.dy) * b->P.fnot);
    const size_t nxy = b->nxy;
    double* F1 = F;            /* q*dp_dy */
    double* F2 = F + nxy;      /* q*dp_dx */
    double* F3 = F + 2 * nxy;  /* p*dq_dx */
    double* F4 = F + 3 * nxy;  /* p*dq_dy */
#pragma omp parallel for num_threads(OMP_THREADS()) schedule(static)
    for (int j = 1; j <= NYP; j++)
        for (int i = 1; i <= NXP; i++) {
            E2(dpx, i, j) = dpdx2_at(b, p, i, j);
            E2(dpy, i, j) = dpdy2_at(b, p, i, j);
            E2(dqx, i, j) = dpdx2_at(b, q, i, j);
            E2(dqy, i, j) = dpdy2_at(b, q, i, j);
        }
#pragma omp parallel for num_threads(OMP_THREADS()) schedule(static)
    for (int j = 1; j <= NYP; j++)
        for (int i = 1; i <= NXP; i++) {
            E2(F1, i, j) = E2(q, i, j) * E2(dpy, i, j);
            E2(F2, i, j) = E2(q, i, j) * E2(dpx, i, j);
            E2(F3, i, j) = E2(p, i, j) * E2(dqx, i, j);
            E2(F4, i, j) = E2(p, i, j) * E2(dqy, i, j);
        }
#pragma omp parallel for num_threads(OMP_THREADS()) schedule(static)
    for (int j = 1; j <= NYP; j++)
        for (int i = 1; i <= NXP; i++) {
            double t = E2(dqx, i, j) * E2(dpy, i, j) - E2(dqy, i, j) * E2(dpx, i, j);
            t = t + dpdx2_at(b, F1, i, j);
            t = t - dpdy2_at(b, F2, i, j);
            t = t + dpdy2_at(b, F3, i, j);
            t = t - dpdx2_at(b, F4, i, j);
            E2(jac, i, j) = fac * t;
        }
    if (b->P.cyclic) {
        double* row = F; /* reuse */
        for (int i = 1; i <= NXP; i++) row[i - 1] = E2(q, i, 1) * E2(dpx, i, 2);
        double aj5sms = cr_int_P_dx(b, row) / (2.0 * b->P.dx);
        for (int i = 1; i <= NXP; i++) row[i - 1] = E2(q, i, 2) * E2(dpx, i, 2);
        double aj9sms = cr_int_P_dx(b, row) / (2.0 * b->P.dx);
        *ajis = (aj5sms + 2.0 * aj9sms) / 6.0;
        for (int i = 1; i <= NXP; i++) row[i - 1] = E2(q, i, NYP) * E2(dpx, i, NYP - 1);
        double aj5smn = cr_int_P_dx(b, row) / (2.0 * b->P.dx);
        for (int i = 1; i <= NXP; i++) row[i - 1] = E2(q, i, NYP - 1) * E2(dpx, i, NYP - 1);
        double aj9smn = cr_int_P_dx(b, row) / (2.0 * b->P.dx);
        *ajin = -(aj5smn + 2.0 * aj9smn) / 6.0;
    }
}

void cr_jacobian(const cr_basin* b, const double* p, const double* q, double* jac,
                 double* ajis, double* ajin) {
    size_t nxy = b->nxy;
    double* wk = dalloc(8 * nxy);
    for (int k = 0; k < b->nl; k++) {
        double s = 0.0, n = 0.0;
        jacobian_layer(b, p + k * nxy, q + k * nxy, jac + k * nxy, &s, &n, wk, wk + nxy,
                       wk + 2 * nxy, wk + 3 * nxy, wk + 4 * nxy);
        if (ajis) ajis[k] = s;
        if (ajin) ajin[k] = n;
    }
    free(wk);
}

/* ------------------------------------------------------------------ */
/* inhomog.F90: Helmholtz solver                                      */
/* ------------------------------------------------------------------ */

/* Thomas algorithm in the Numerical Recipes form of inhomog.F90:188-199 for
 * a*u[j-1] + bb*u[j] + a*u[j+1] = f[j], j = 0..n-1, with strides fs, us. */
static void thomas_nr(int n, double a, double bb, const double* f, size_t fs, double* u, size_t us,
                      double* gam) {
    double betinv = 1.0 / bb;
    u[0] = f[0] * betinv;
    for (int j = 1; j < n; j++) {
        gam[j] = a * betinv;
        betinv = 1.0 / (bb - a * gam[j]);
        u[j * us] = (f[j * fs] - a * u[(j - 1) * us]) * betinv;
    }
    for (int j = n - 2; j >= 0; j--) u[j * us] = u[j * us] - gam[j + 1] * u[(j + 1) * us];
}

void cr_thomas(int n, double a, double bb, const double* f, double* u) {
    double* gam = dalloc((size_t)n);
    thomas_nr(n, a, bb, f, 1, u, 1, gam);
    free(gam);
}

/* solve_inhomog_box (inhomog.F90:170-212) */
static void solve_inhomog_box(cr_basin* b, int m, const double* rhs, double* out) {
    const int n = b->n_in, nyi = NYP - 2, nx = b->nx_spec;
    double* fhat = b->w4; /* (nx, nyi): transformed rows */
    const int L = b->nfft;
    /* forward DST of rows j = 2..nyp-1, interior i = 2..nxp-1, two rows per FFT */
#pragma omp parallel num_threads(OMP_THREADS())
    {
        double* z = dalloc(4 * (size_t)L);
        double* Z = z + 2 * L;
        double* x = dalloc(2 * (size_t)n);
        double* y = x + n;
#pragma omp for schedule(static)
        for (int jp = 0; jp < (nyi + 1) / 2; jp++) {
            int j0 = 2 * jp, j1 = 2 * jp + 1;
            for (int i = 0; i < n; i++) {
                x[i] = E2(rhs, i + 2, j0 + 2);
                y[i] = (j1 < nyi) ? E2(rhs, i + 2, j1 + 2) : 0.0;
            }
            dst1_pair(&b->plan, n, x, (j1 < nyi) ? y : NULL, fhat + (size_t)j0 * nx,
                      (j1 < nyi) ? fhat + (size_t)j1 * nx : NULL, z, Z);
        }
        free(z);
        free(x);
    }
    /* Thomas along y for each wavenumber (inhomog.F90:188-199) */
    double* u = b->w5;
    const double a = b->a;
    const double* bb = b->bb + (size_t)(m - 1) * nx;
#pragma omp parallel num_threads(OMP_THREADS())
    {
        double* gam = dalloc((size_t)nyi);
#pragma omp for schedule(static)
        for (int k = 0; k < nx; k++) thomas_nr(nyi, a, bb[k], fhat + k, (size_t)nx, u + k, (size_t)nx, gam);
        free(gam);
    }
    /* rescale, inverse transform (inhomog.F90:202-204), zero the ring (:207-210) */
    for (size_t t = 0; t < (size_t)nx * nyi; t++) u[t] = b->ftnorm * u[t];
#pragma omp parallel num_threads(OMP_THREADS())
    {
        double* z = dalloc(4 * (size_t)L);
        double* Z = z + 2 * L;
        double* x = dalloc(2 * (size_t)n);
        double* y = x + n;
#pragma omp for schedule(static)
        for (int jp = 0; jp < (nyi + 1) / 2; jp++) {
            int j0 = 2 * jp, j1 = 2 * jp + 1;
            dst1_pair(&b->plan, n, u + (size_t)j0 * nx, (j1 < nyi) ? u + (size_t)j1 * nx : NULL, x,
                      (j1 < nyi) ? y : NULL, z, Z);
            for (int i = 0; i < n; i++) {
                E2(out, i + 2, j0 + 2) = x[i];
                if (j1 < nyi) E2(out, i + 2, j1 + 2) = y[i];
            }
        }
        free(z);
        free(x);
    }
    for (int j = 1; j <= NYP; j++) { E2(out, 1, j) = 0.0; E2(out, NXP, j) = 0.0; }
    for (int i = 1; i <= NXP; i++) { E2(out, i, 1) = 0.0; E2(out, i, NYP) = 0.0; }
}

/* solve_inhomog_cyclic (inhomog.F90:214-255) */
static void solve_inhomog_cyclic(cr_basin* b, int m, const double* rhs, double* out) {
    const int n = b->n_in, nyi = NYP - 2, nx = b->nx_spec;
    double* fhat = b->w4; /* (2*nx, nyi) interleaved complex */
    const int L = b->nfft;
#pragma omp parallel num_threads(OMP_THREADS())
    {
        double* z = dalloc(4 * (size_t)L);
        double* Z = z + 2 * L;
#pragma omp for schedule(static)
        for (int jp = 0; jp < (nyi + 1) / 2; jp++) {
            int j0 = 2 * jp, j1 = 2 * jp + 1;
            const double* x = &E2(rhs, 1, j0 + 2);
            const double* y = (j1 < nyi) ? &E2(rhs, 1, j1 + 2) : NULL;
            rfft_pair(&b->plan, n, x, y, fhat + (size_t)j0 * 2 * nx,
                      y ? fhat + (size_t)j1 * 2 * nx : NULL, z, Z);
        }
        free(z);
    }
    double* u = b->w5;
    const double a = b->a;
    const double* bb = b->bb + (size_t)(m - 1) * nx;
#pragma omp parallel num_threads(OMP_THREADS())
    {
        double* gam = dalloc((size_t)nyi);
#pragma omp for schedule(static)
        for (int k = 0; k < nx; k++)
            for (int c = 0; c < 2; c++)  /* real and imaginary parts share gam/betinv */
                thomas_nr(nyi, a, bb[k], fhat + 2 * k + c, 2 * (size_t)nx, u + 2 * k + c, 2 * (size_t)nx, gam);
        free(gam);
    }
    for (size_t t = 0; t < 2 * (size_t)nx * nyi; t++) u[t] = b->ftnorm * u[t];
#pragma omp parallel num_threads(OMP_THREADS())
    {
        double* z = dalloc(4 * (size_t)L);
        double* Z = z + 2 * L;
#pragma omp for schedule(static)
        for (int jp = 0; jp < (nyi + 1) / 2; jp++) {
            int j0 = 2 * jp, j1 = 2 * jp + 1;
            double* x = &E2(out, 1, j0 + 2);
            double* y = (j1 < nyi) ? &E2(out, 1, j1 + 2) : NULL;
            irfft_pair(&b->plan, n, u + (size_t)j0 * 2 * nx, y ? u + (size_t)j1 * 2 * nx : NULL, x, y, z, Z);
        }
        free(z);
    }
    for (int i = 1; i <= NXP; i++) { E2(out, i, 1) = 0.0; E2(out, i, NYP) = 0.0; }
    for (int j = 1; j <= NYP; j++) E2(out, NXP, j) = E2(out, 1, j);
}

void cr_solve_inhomog(cr_basin* b, int m, const double* rhs, double* out) {
    if (b->P.cyclic) solve_inhomog_cyclic(b, m, rhs, out);
    else solve_inhomog_box(b, m, rhs, out);
}

/* ------------------------------------------------------------------ */
/* Small dense linear algebra (LAPACK substitutes)                    */
/* ------------------------------------------------------------------ */

/* Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix S (row-major). */
static void jacobi_eig(int n, double* S, double* V, double* w) {
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) V[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++) off += S[i * n + j] * S[i * n + j];
        if (off == 0.0) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = S[p * n + q];
                if (apq == 0.0) continue;
                double app = S[p * n + p], aqq = S[q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = S[k * n + p], akq = S[k * n + q];
                    S[k * n + p] = c * akp - s * akq;
                    S[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = S[p * n + k], aqk = S[q * n + k];
                    S[p * n + k] = c * apk - s * aqk;
                    S[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; i++) w[i] = S[i * n + i];
}

/* LU with partial pivoting (DGETRF semantics), column-major CR_MAXL stride. */
static int lu_factor(int n, const double* A, double* LU, int* ipiv) {
    for (int j = 1; j <= n; j++)
        for (int i = 1; i <= n; i++) M2(LU, i, j) = M2(A, i, j);
    for (int k = 1; k <= n; k++) {
        int piv = k;
        double mx = fabs(M2(LU, k, k));
        for (int i = k + 1; i <= n; i++)
            if (fabs(M2(LU, i, k)) > mx) { mx = fabs(M2(LU, i, k)); piv = i; }
        ipiv[k - 1] = piv;
        if (M2(LU, piv, k) == 0.0) return k;
        if (piv != k)
            for (int j = 1; j <= n; j++) {
                double t = M2(LU, k, j); M2(LU, k, j) = M2(LU, piv, j); M2(LU, piv, j) = t;
            }
        for (int i = k + 1; i <= n; i++) M2(LU, i, k) = M2(LU, i, k) / M2(LU, k, k);
        for (int j = k + 1; j <= n; j++)
            for (int i = k + 1; i <= n; i++) M2(LU, i, j) = M2(LU, i, j) - M2(LU, i, k) * M2(LU, k, j);
    }
    return 0;
}

static void lu_solve(int n, const double* LU, const int* ipiv, double* x) {
    for (int k = 1; k <= n; k++) {
        int p = ipiv[k - 1];
        if (p != k) { double t = x[k - 1]; x[k - 1] = x[p - 1]; x[p - 1] = t; }
    }
    for (int i = 1; i <= n; i++)
        for (int j = 1; j < i; j++) x[i - 1] -= M2(LU, i, j) * x[j - 1];
    for (int i = n; i >= 1; i--) {
        for (int j = i + 1; j <= n; j++) x[i - 1] -= M2(LU, i, j) * x[j - 1];
        x[i - 1] /= M2(LU, i, i);
    }
}

/* solve_Ax_b (linalg.F90:57-93): DGETRS, then DGERFS iterative refinement. */
static void solve_Ax_b(int n, const double* A, const double* LU, const int* ipiv,
                       const double* rhs, double* x) {
    for (int i = 0; i < n; i++) x[i] = rhs[i];
    lu_solve(n, LU, ipiv, x);
    double lstres = 3.0, eps = 1.1102230246251565e-16;
    for (int it = 0; it < 5; it++) {
        double r[CR_MAXL], berr = 0.0;
        for (int i = 1; i <= n; i++) {
            double s = rhs[i - 1], den = fabs(rhs[i - 1]);
            for (int j = 1; j <= n; j++) {
                s -= M2(A, i, j) * x[j - 1];
                den += fabs(M2(A, i, j)) * fabs(x[j - 1]);
            }
            r[i - 1] = s;
            if (den > 0.0 && fabs(s) / den > berr) berr = fabs(s) / den;
        }
        if (!(berr > eps && 2.0 * berr <= lstres)) break;
        lu_solve(n, LU, ipiv, r);
        for (int i = 0; i < n; i++) x[i] += r[i];
        lstres = berr;
    }
}

/* ------------------------------------------------------------------ */
/* Initialisation                                                     */
/* ------------------------------------------------------------------ */

/* compute_A (modes.F90:221-246) + eigmod/derive_qg_modes (:46-219). */
static int init_modes(cr_basin* b) {
    const int nl = b->nl;
    const double* gp = b->P.gp;
    const double* h = b->P.h;
    memset(b->amat, 0, sizeof(b->amat));
    if (nl == 1) {
        M2(b->ctl2m, 1, 1) = 1.0; M2(b->ctm2l, 1, 1) = 1.0; b->rdm2[0] = 0.0;
        return 0;
    }
    M2(b->amat, 1, 2) = -1.0 / (gp[0] * h[0]);
    M2(b->amat, 1, 1) = -M2(b->amat, 1, 2);
    for (int k = 2; k <= nl - 1; k++) {
        M2(b->amat, k, k - 1) = -1.0 / (gp[k - 2] * h[k - 1]);
        M2(b->amat, k, k + 1) = -1.0 / (gp[k - 1] * h[k - 1]);
        M2(b->amat, k, k) = -M2(b->amat, k, k - 1) - M2(b->amat, k, k + 1);
    }
    M2(b->amat, nl, nl - 1) = -1.0 / (gp[nl - 2] * h[nl - 1]);
    M2(b->amat, nl, nl) = -M2(b->amat, nl, nl - 1);

    /* S = H^{1/2} A H^{-1/2} is symmetric; A = H^{-1/2} V diag(w) V^T H^{1/2} */
    double S[CR_MAXL * CR_MAXL], V[CR_MAXL * CR_MAXL], w[CR_MAXL];
    for (int i = 0; i < nl; i++)
        for (int j = 0; j < nl; j++)
            S[i * nl + j] = sqrt(h[i]) * M2(b->amat, i + 1, j + 1) / sqrt(h[j]);
    for (int i = 0; i < nl; i++)
        for (int j = i + 1; j < nl; j++) {
            double s = 0.5 * (S[i * nl + j] + S[j * nl + i]);
            S[i * nl + j] = S[j * nl + i] = s;
        }
    jacobi_eig(nl, S, V, w);
    double evecr[CR_MAXL][CR_MAXL], evecl[CR_MAXL][CR_MAXL]; /* [vector][component] */
    for (int m = 0; m < nl; m++) {
        double nrm = 0.0;
        for (int k = 0; k < nl; k++) {
            evecr[m][k] = V[k * nl + m] / sqrt(h[k]);
            evecl[m][k] = V[k * nl + m] * sqrt(h[k]);
            nrm += evecr[m][k] * evecr[m][k];
        }
        nrm = sqrt(nrm);
        double sg = (evecr[m][0] < 0.0) ? -1.0 : 1.0;
        for (int k = 0; k < nl; k++) { evecr[m][k] = sg * evecr[m][k] / nrm; evecl[m][k] *= sg; }
    }
    /* insertion sort of |w| ascending (modes.F90:99-115) */
    double c2rabs[CR_MAXL];
    int index[CR_MAXL];
    for (int m = 0; m < nl; m++) { c2rabs[m] = fabs(w[m]); index[m] = m; }
    for (int m = 1; m < nl; m++) {
        int indtmp = index[m];
        double c2temp = c2rabs[indtmp];
        int i;
        for (i = m - 1; i >= 0; i--) {
            if (c2rabs[index[i]] <= c2temp) break;
            index[i + 1] = index[i];
        }
        index[i + 1] = indtmp;
    }
    b->rdm2[0] = 0.0;
    for (int m = 1; m < nl; m++) b->rdm2[m] = b->P.fnot * b->P.fnot * c2rabs[index[m]];
    for (int m = 1; m <= nl; m++) {
        int inm = index[m - 1];
        double elder = 0.0;
        for (int k = 0; k < nl; k++) elder += evecl[inm][k] * evecr[inm][k];
        for (int k = 1; k <= nl; k++) {
            M2(b->ctl2m, k, m) = evecl[inm][k - 1] / elder;
            M2(b->ctm2l, m, k) = evecr[inm][k - 1];
        }
    }
    /* identity check (modes.F90:149-161) */
    for (int l = 1; l <= nl; l++)
        for (int k = 1; k <= nl; k++) {
            double s = 0.0;
            for (int m = 1; m <= nl; m++) s += M2(b->ctm2l, m, k) * M2(b->ctl2m, l, m);
            if (fabs(s - (l == k ? 1.0 : 0.0)) > 5.0e-15) return -2;
        }
    return 0;
}

/* init_inhomog (inhomog.F90:71-144) */
static int init_inhomog(cr_basin* b) {
    int N;
    if (b->P.cyclic) {
        b->n_in = NXP - 1;
        b->nx_spec = b->n_in / 2 + 1;
        N = b->n_in;
        b->nfft = N;
    } else {
        b->n_in = NXP - 2;
        b->nx_spec = b->n_in;
        N = 2 * (b->n_in + 1);
        b->nfft = N;
    }
    if (fft_plan_init(&b->plan, b->nfft) != 0) return -3;
    b->ftnorm = 1.0 / N;
    b->a = 1.0 / (b->P.dy * b->P.dy);
    b->bb = dalloc((size_t)b->nx_spec * b->nl);
    for (int i = 1; i <= b->nx_spec; i++) {
        double arg = b->P.cyclic ? ((double)(i - 1) * TWOPI_D) / (double)N : ((double)i * TWOPI_D) / (double)N;
        double bd2 = -2.0 * b->a + 2.0 * b->dxm2 * (cos(arg) - 1.0);
        for (int m = 0; m < b->nl; m++) b->bb[(size_t)m * b->nx_spec + (i - 1)] = bd2 - b->rdm2[m];
    }
    return 0;
}

/* generate_homog_soln (inhomog.F90:146-154) */
static void generate_homog_soln(cr_basin* b, int m, const double* L, double* out) {
    cr_solve_inhomog(b, m, L, out);
    for (size_t t = 0; t < b->nxy; t++) out[t] = L[t] + b->rdm2[m - 1] * out[t];
}

/* homsol_box (homog.F90:214-278) / homsol_cyclic (:128-212) */
static int homsol(cr_basin* b) {
    const int nl = b->nl;
    size_t nxy = b->nxy;
    if (!b->P.cyclic) {
        b->hom_sol_bc = dalloc(nxy * (size_t)(nl > 1 ? nl - 1 : 1));
        double* L = dalloc(nxy);
        for (size_t t = 0; t < nxy; t++) L[t] = 1.0;
        double aipohs[CR_MAXL];
        for (int m = 2; m <= nl; m++) {
            generate_homog_soln(b, m, L, b->hom_sol_bc + (size_t)(m - 2) * nxy);
            aipohs[m - 1] = cr_int_P_dA(b, b->hom_sol_bc + (size_t)(m - 2) * nxy);
        }
        free(L);
        for (int k = 1; k <= nl - 1; k++)
            for (int m = 2; m <= nl; m++)
                M2(b->cdhoc, k, m - 1) = (M2(b->ctm2l, m, k + 1) - M2(b->ctm2l, m, k)) * aipohs[m - 1];
        if (nl > 1 && lu_factor(nl - 1, b->cdhoc, b->LU, b->ipiv) != 0) return -4;
    } else {
        b->hom_bt = dalloc((size_t)NYP);
        b->hom_bc1 = dalloc((size_t)NYP * (nl > 1 ? nl - 1 : 1));
        b->hom_bc2 = dalloc((size_t)NYP * (nl > 1 ? nl - 1 : 1));
        for (int j = 1; j <= NYP; j++) b->hom_bt[j - 1] = (double)(NYP - j) / (double)(NYP - 1);
        b->hbsi = b->yl / b->xl;
        double* r1 = dalloc(nxy);
        double* r2 = dalloc(nxy);
        double* s1 = dalloc(nxy);
        double* s2 = dalloc(nxy);
        for (int j = 1; j <= NYP; j++)
            for (int i = 1; i <= NXP; i++) {
                E2(r1, i, j) = (b->yp[NYP - 1] - b->yp[j - 1]) / b->yl;
                E2(r2, i, j) = (b->yp[j - 1] - b->yp[0]) / b->yl;
            }
        for (int m = 2; m <= nl; m++) {
            generate_homog_soln(b, m, r1, s1);
            generate_homog_soln(b, m, r2, s2);
            for (int j = 1; j <= NYP; j++) {
                b->hom_bc1[(size_t)(m - 2) * NYP + j - 1] = E2(s1, 1, j);
                b->hom_bc2[(size_t)(m - 2) * NYP + j - 1] = E2(s2, 1, j);
            }
        }
        free(r1); free(r2); free(s1); free(s2);
        const double xl = b->xl, dy = b->P.dy;
        for (int m = 2; m <= nl; m++) {
            const double* h1 = b->hom_bc1 + (size_t)(m - 2) * NYP;
            const double* h2 = b->hom_bc2 + (size_t)(m - 2) * NYP;
            double pch1ys = -xl * (h1[1] - h1[0]) / dy;
            double pch2ys = -xl * (h2[1] - h2[0]) / dy;
            double pch1yn = xl * (h1[NYP - 1] - h1[NYP - 2]) / dy;
            double pch2yn = xl * (h2[NYP - 1] - h2[NYP - 2]) / dy;
            pch1ys = pch1ys + 0.5 * dy * b->rdm2[m - 1] * h1[0] * xl;
            pch2ys = pch2ys + 0.5 * dy * b->rdm2[m - 1] * h2[0] * xl;
            pch1yn = pch1yn + 0.5 * dy * b->rdm2[m - 1] * h1[NYP - 1] * xl;
            pch2yn = pch2yn + 0.5 * dy * b->rdm2[m - 1] * h2[NYP - 1] * xl;
            double pchdet = pch1ys * pch2yn - pch2ys * pch1yn;
            b->hc1s[m - 1] = pch1ys / pchdet;
            b->hc2s[m - 1] = pch2ys / pchdet;
            b->hc1n[m - 1] = pch1yn / pchdet;
            b->hc2n[m - 1] = pch2yn / pchdet;
        }
    }
    return 0;
}

int cr_create(const cr_params* prm, const double* ddyn, cr_basin** out) {
    *out = NULL;
    if (prm->nl < 2 || prm->nl > CR_MAXL || prm->nxp < 4 || prm->nyp < 4) return -1;
    cr_basin* b = (cr_basin*)calloc(1, sizeof(cr_basin));
    b->P = *prm;
    b->nxp = prm->nxp;
    b->nyp = prm->nyp;
    b->nl = prm->nl;
    b->nxy = (size_t)b->nxp * b->nyp;
    /* compute_derived_values (box.F90:126-163), with x0 = y0 = 0 */
    b->xl = (NXP - 1) * prm->dx;
    b->yl = (NYP - 1) * prm->dy;
    b->dxm2 = 1.0 / (prm->dx * prm->dx);
    b->yp = dalloc((size_t)NYP);
    b->yprel = dalloc((size_t)NYP);
    for (int j = 1; j <= NYP; j++) {
        b->yp[j - 1] = 0.0 + (j - 1) * prm->dy;
        b->yprel[j - 1] = -0.5 * b->yl + (j - 1) * prm->dy;
    }
    b->ddyn = dalloc(b->nxy);
    if (ddyn) memcpy(b->ddyn, ddyn, sizeof(double) * b->nxy);
    size_t nf = b->nxy * b->nl;
    b->q = dalloc(nf); b->qm = dalloc(nf); b->p = dalloc(nf); b->pm = dalloc(nf);
    b->w1 = dalloc(nf); b->w2 = dalloc(nf); b->w3 = dalloc(nf);
    b->w4 = dalloc(2 * b->nxy + 64); b->w5 = dalloc(2 * b->nxy + 64);
    b->w6 = dalloc(nf); b->w7 = dalloc(8 * b->nxy);
    int rc = init_modes(b);
    if (rc == 0) rc = init_inhomog(b);
    if (rc == 0) rc = homsol(b);
    if (rc != 0) { cr_destroy(b); return rc; }
    *out = b;
    return 0;
}

void cr_destroy(cr_basin* b) {
    if (!b) return;
    free(b->yp); free(b->yprel); free(b->ddyn); free(b->bb);
    fft_plan_free(&b->plan);
    free(b->hom_sol_bc); free(b->hom_bt); free(b->hom_bc1); free(b->hom_bc2);
    free(b->q); free(b->qm); free(b->p); free(b->pm);
    free(b->w1); free(b->w2); free(b->w3); free(b->w4); free(b->w5); free(b->w6); free(b->w7);
    free(b);
}

void cr_set_state(cr_basin* b, const double* q, const double* qm, const double* p, const double* pm) {
    size_t nf = b->nxy * b->nl * sizeof(double);
    if (q) memcpy(b->q, q, nf);
    if (qm) memcpy(b->qm, qm, nf);
    if (p) memcpy(b->p, p, nf);
    if (pm) memcpy(b->pm, pm, nf);
}

void cr_get_state(const cr_basin* b, double* q, double* qm, double* p, double* pm) {
    size_t nf = b->nxy * b->nl * sizeof(double);
    if (q) memcpy(q, b->q, nf);
    if (qm) memcpy(qm, b->qm, nf);
    if (p) memcpy(p, b->p, nf);
    if (pm) memcpy(pm, b->pm, nf);
}

int cr_constr_len(const cr_basin* b) { return 2 * (b->nl - 1) + 4 * b->nl; }

void cr_get_constr(const cr_basin* b, double* o) {
    int nl = b->nl, t = 0;
    for (int k = 0; k < nl - 1; k++) o[t++] = b->dpi[k];
    for (int k = 0; k < nl - 1; k++) o[t++] = b->dpip[k];
    for (int k = 0; k < nl; k++) o[t++] = b->cs[k];
    for (int k = 0; k < nl; k++) o[t++] = b->cn[k];
    for (int k = 0; k < nl; k++) o[t++] = b->csp[k];
    for (int k = 0; k < nl; k++) o[t++] = b->cnp[k];
}

void cr_set_constr(cr_basin* b, const double* o) {
    int nl = b->nl, t = 0;
    for (int k = 0; k < nl - 1; k++) b->dpi[k] = o[t++];
    for (int k = 0; k < nl - 1; k++) b->dpip[k] = o[t++];
    for (int k = 0; k < nl; k++) b->cs[k] = o[t++];
    for (int k = 0; k < nl; k++) b->cn[k] = o[t++];
    for (int k = 0; k < nl; k++) b->csp[k] = o[t++];
    for (int k = 0; k < nl; k++) b->cnp[k] = o[t++];
}

void cr_get_modes(const cr_basin* b, double* amat, double* ctl2m, double* ctm2l, double* rdm2) {
    int nl = b->nl;
    for (int j = 1; j <= nl; j++)
        for (int i = 1; i <= nl; i++) {
            if (amat) amat[(j - 1) * nl + (i - 1)] = M2(b->amat, i, j);
            if (ctl2m) ctl2m[(j - 1) * nl + (i - 1)] = M2(b->ctl2m, i, j);
            if (ctm2l) ctm2l[(j - 1) * nl + (i - 1)] = M2(b->ctm2l, i, j);
        }
    if (rdm2) for (int m = 0; m < nl; m++) rdm2[m] = b->rdm2[m];
}

void cr_get_homog(const cr_basin* b, double* a0, double* a1, double* a2, double* coef) {
    int nl = b->nl;
    if (!b->P.cyclic) {
        if (a0) memcpy(a0, b->hom_sol_bc, sizeof(double) * b->nxy * (nl - 1));
        return;
    }
    if (a0) memcpy(a0, b->hom_bt, sizeof(double) * NYP);
    if (a1) memcpy(a1, b->hom_bc1, sizeof(double) * NYP * (nl - 1));
    if (a2) memcpy(a2, b->hom_bc2, sizeof(double) * NYP * (nl - 1));
    if (coef) {
        int t = 0;
        for (int m = 2; m <= nl; m++) {
            coef[t++] = b->hc1s[m - 1]; coef[t++] = b->hc2s[m - 1];
            coef[t++] = b->hc1n[m - 1]; coef[t++] = b->hc2n[m - 1];
        }
        coef[t] = b->hbsi;
    }
}

/* qcomp (vorsubs.F90:27-56) */
static void qcomp(const cr_basin* b, const double* p, double* q) {
    const int nl = b->nl, kt = b->P.k_topo;
    cr_del2_bc(b, p, q, nl);
    for (size_t t = 0; t < b->nxy * nl; t++) q[t] = q[t] / b->P.fnot;
    for (int j = 1; j <= NYP; j++)
        for (int i = 1; i <= NXP; i++) E3(q, i, j, kt) = E3(q, i, j, kt) + E2(b->ddyn, i, j);
    for (int j = 1; j <= NYP; j++) {
        for (int k = 1; k <= nl; k++)
            for (int i = 1; i <= NXP; i++) E3(q, i, j, k) = E3(q, i, j, k) + b->P.beta * b->yprel[j - 1];
        for (int i = 1; i <= NXP; i++)
            for (int k = 1; k <= nl; k++) {
                double s = 0.0;
                for (int l = 1; l <= nl; l++) s += M2(b->amat, k, l) * E3(p, i, j, l);
                E3(q, i, j, k) = E3(q, i, j, k) - b->P.fnot * s;
            }
    }
}

void cr_init_state(cr_basin* b) {
    const int nl = b->nl;
    qcomp(b, b->p, b->q);
    qcomp(b, b->pm, b->qm);
    double* row = b->w1;
    if (b->P.cyclic) {
        /* init_constraint (constraint.F90:80-143) */
        double pins[CR_MAXL], pinn[CR_MAXL], pinsp[CR_MAXL], pinnp[CR_MAXL];
        for (int k = 1; k <= nl; k++) {
            pinsp[k - 1] = cr_int_P_dx(b, &E3(b->pm, 1, 1, k));
            pins[k - 1] = cr_int_P_dx(b, &E3(b->p, 1, 1, k));
            pinnp[k - 1] = cr_int_P_dx(b, &E3(b->pm, 1, NYP, k));
            pinn[k - 1] = cr_int_P_dx(b, &E3(b->p, 1, NYP, k));
            for (int i = 1; i <= NXP; i++) row[i - 1] = E3(b->pm, i, 2, k) - E3(b->pm, i, 1, k);
            b->csp[k - 1] = cr_int_P_dx(b, row) / b->P.dy;
            for (int i = 1; i <= NXP; i++) row[i - 1] = E3(b->pm, i, NYP, k) - E3(b->pm, i, NYP - 1, k);
            b->cnp[k - 1] = cr_int_P_dx(b, row) / b->P.dy;
            for (int i = 1; i <= NXP; i++) row[i - 1] = E3(b->p, i, 2, k) - E3(b->p, i, 1, k);
            b->cs[k - 1] = cr_int_P_dx(b, row) / b->P.dy;
            for (int i = 1; i <= NXP; i++) row[i - 1] = E3(b->p, i, NYP, k) - E3(b->p, i, NYP - 1, k);
            b->cn[k - 1] = cr_int_P_dx(b, row) / b->P.dy;
        }
        for (int k = 1; k <= nl; k++) {
            double apsp = 0, apnp = 0, aps = 0, apn = 0;
            for (int l = 1; l <= nl; l++) {
                apsp += M2(b->amat, k, l) * pinsp[l - 1];
                apnp += M2(b->amat, k, l) * pinnp[l - 1];
                aps += M2(b->amat, k, l) * pins[l - 1];
                apn += M2(b->amat, k, l) * pinn[l - 1];
            }
            const double c = 0.5 * b->P.dy * b->P.fnot * b->P.fnot;
            b->csp[k - 1] = -b->csp[k - 1] + c * apsp;
            b->cnp[k - 1] = b->cnp[k - 1] + c * apnp;
            b->cs[k - 1] = -b->cs[k - 1] + c * aps;
            b->cn[k - 1] = b->cn[k - 1] + c * apn;
        }
        for (int k = 0; k < nl; k++)
            b->ajis[k] = b->ajin[k] = b->ap3s[k] = b->ap3n[k] = b->ap5s[k] = b->ap5n[k] =
                b->int_Be_s[k] = b->int_Be_n[k] = 0.0;
    }
    /* init_mass_constr (constraint.F90:145-165) */
    for (int k = 1; k <= nl - 1; k++) {
        double ipm_k = cr_int_P_dA(b, &E3(b->pm, 1, 1, k)), ipm_k1 = cr_int_P_dA(b, &E3(b->pm, 1, 1, k + 1));
        double ip_k = cr_int_P_dA(b, &E3(b->p, 1, 1, k)), ip_k1 = cr_int_P_dA(b, &E3(b->p, 1, 1, k + 1));
        b->dpip[k - 1] = b->P.dz_sign * (ipm_k - ipm_k1);
        b->dpi[k - 1] = b->P.dz_sign * (ip_k - ip_k1);
    }
}

/* ------------------------------------------------------------------ */
/* The step                                                           */
/* ------------------------------------------------------------------ */

/* vort_step (vorticity.F90:14-103) */
static void vort_step(cr_basin* b, double tdt, const double* ent, int n_ent, const double* wekp,
                      double txis, double txin) {
    const int nl = b->nl, kt = b->P.k_topo;
    const size_t nxy = b->nxy;
    double* dqdt = b->w1;
    double* del2p = b->w2;
    double* d4p = b->w3;
    double* wk = b->w7;
    double* d6p = b->w6;
    for (int k = 0; k < nl; k++) {
        double s = 0, n = 0;
        jacobian_layer(b, b->p + k * nxy, b->q + k * nxy, dqdt + k * nxy, &s, &n, wk, wk + nxy,
                       wk + 2 * nxy, wk + 3 * nxy, wk + 4 * nxy);
        if (b->P.cyclic) { b->ajis[k] = s; b->ajin[k] = n; }
    }
    cr_del2_bc(b, b->pm, del2p, nl);
    cr_del2_bc(b, del2p, d4p, nl);
    cr_del2_bc(b, d4p, d6p, nl);
    double* row = dalloc((size_t)NXP);
    for (int k = 1; k <= nl; k++) {
        const double c2 = b->P.ah2[k - 1] / b->P.fnot, c4 = b->P.ah4[k - 1] / b->P.fnot;
        for (size_t t = (size_t)(k - 1) * nxy; t < (size_t)k * nxy; t++)
            dqdt[t] = dqdt[t] + c2 * d4p[t] - c4 * d6p[t];
        if (b->P.cyclic) {
            for (int i = 1; i <= NXP; i++) row[i - 1] = E3(del2p, i, 2, k) - E3(del2p, i, 1, k);
            b->ap3s[k - 1] = b->P.ah2[k - 1] * cr_int_P_dx(b, row) / b->P.dy;
            for (int i = 1; i <= NXP; i++) row[i - 1] = E3(del2p, i, NYP, k) - E3(del2p, i, NYP - 1, k);
            b->ap3n[k - 1] = b->P.ah2[k - 1] * cr_int_P_dx(b, row) / b->P.dy;
            for (int i = 1; i <= NXP; i++) row[i - 1] = E3(d4p, i, 2, k) - E3(d4p, i, 1, k);
            b->ap5s[k - 1] = b->P.ah4[k - 1] * cr_int_P_dx(b, row) / b->P.dy;
            for (int i = 1; i <= NXP; i++) row[i - 1] = E3(d4p, i, NYP, k) - E3(d4p, i, NYP - 1, k);
            b->ap5n[k - 1] = b->P.ah4[k - 1] * cr_int_P_dx(b, row) / b->P.dy;
        }
    }
    double fohfac[CR_MAXL] = {0.0};
    for (int k = 0; k < nl; k++) fohfac[k] = b->P.fnot / b->P.h[k];
    const double* ent1 = ent;
    const double cent1 = b->P.dz_sign * fohfac[0], cwek = b->P.tau_sign * fohfac[0];
    const double cent2 = (nl >= 2) ? b->P.dz_sign * fohfac[1] : 0.0;
    const double cdrag = (fohfac[kt - 1] * b->P.delek / (2.0 * fabs(b->P.fnot)));
#pragma omp parallel for num_threads(OMP_THREADS()) schedule(static)
    for (size_t t = 0; t < nxy; t++) {
        dqdt[t] = dqdt[t] + cent1 * ent1[t] + cwek * wekp[t];
        if (nl >= 2) dqdt[nxy + t] = dqdt[nxy + t] - cent2 * ent1[t];
        size_t o = (size_t)(kt - 1) * nxy + t;
        dqdt[o] = dqdt[o] - cdrag * del2p[o];
    }
#pragma omp parallel for num_threads(OMP_THREADS()) schedule(static)
    for (size_t t = 0; t < nxy * nl; t++) {
        double qold = b->q[t];
        b->q[t] = b->qm[t] + tdt * dqdt[t];
        b->qm[t] = qold;
    }
    if (b->P.cyclic) {
        b->txis = txis;
        b->txin = txin;
        for (int k = 0; k < nl; k++) b->int_Be_s[k] = b->int_Be_n[k] = 0.0;
        const double* ent2 = (n_ent >= 2) ? ent + nxy : NULL;
        b->int_Be_s[0] = b->P.dz_sign * fohfac[0] * cr_int_P_dx(b, &E2(ent1, 1, 1));
        if (nl >= 2 && ent2) b->int_Be_s[1] = b->P.dz_sign * fohfac[0] * cr_int_P_dx(b, &E2(ent2, 1, 1));
        if (nl >= 2) b->int_Be_n[0] = -b->P.dz_sign * fohfac[1] * cr_int_P_dx(b, &E2(ent1, 1, NYP));
        if (nl >= 2 && ent2) b->int_Be_n[1] = -b->P.dz_sign * fohfac[1] * cr_int_P_dx(b, &E2(ent2, 1, NYP));
        for (int i = 1; i <= NXP; i++) row[i - 1] = E3(b->pm, i, 2, kt) - E3(b->pm, i, 1, kt);
        double bdsums = cr_int_P_dx(b, row) / b->P.dy;
        for (int i = 1; i <= NXP; i++) row[i - 1] = E3(b->pm, i, NYP, kt) - E3(b->pm, i, NYP - 1, kt);
        double bdsumn = cr_int_P_dx(b, row) / b->P.dy;
        const double cd = b->P.fnot * b->P.delek / (2.0 * fabs(b->P.fnot));
        b->bdrins = cd * bdsums;
        b->bdrinn = cd * bdsumn;
    }
    free(row);
}

/* step_constraints (constraint.F90:182-215) */
static void step_constraints(cr_basin* b, double tdt) {
    const int nl = b->nl;
    double rhss[CR_MAXL], rhsn[CR_MAXL];
    const double f0 = b->P.fnot, dy = b->P.dy;
    for (int k = 0; k < nl; k++) {
        rhss[k] = -0.5 * dy * f0 * b->int_Be_s[k];
        rhsn[k] = -0.5 * dy * f0 * b->int_Be_n[k];
    }
    rhss[0] = rhss[0] + b->P.tau_sign * (f0 / b->P.h[0]) * b->txis;
    rhsn[0] = rhsn[0] - b->P.tau_sign * (f0 / b->P.h[0]) * b->txin;
    rhss[nl - 1] = rhss[nl - 1] + (f0 / b->P.h[nl - 1]) * b->bdrins;
    rhsn[nl - 1] = rhsn[nl - 1] - (f0 / b->P.h[nl - 1]) * b->bdrinn;
    for (int k = 0; k < nl; k++) {
        rhss[k] = rhss[k] + b->ajis[k] - b->ap3s[k] + b->ap5s[k];
        rhsn[k] = rhsn[k] + b->ajin[k] + b->ap3n[k] - b->ap5n[k];
    }
    for (int k = 0; k < nl; k++) {
        double csnew = b->csp[k] + tdt * rhss[k];
        double cnnew = b->cnp[k] + tdt * rhsn[k];
        b->csp[k] = b->cs[k];
        b->cnp[k] = b->cn[k];
        b->cs[k] = csnew;
        b->cn[k] = cnnew;
    }
}

/* solve_pressure (pressure.F90:21-69) */
static void solve_pressure(cr_basin* b, double tdt, const double* ent) {
    const int nl = b->nl, kt = b->P.k_topo;
    const size_t nxy = b->nxy, nf = nxy * nl;
    double* rhs = b->w1;
    double* inh = b->w2;
    double* homcor = b->w3;
    memcpy(b->pm, b->p, nf * sizeof(double));
    /* p_modes_rhs (modes.F90:248-266) */
#pragma omp parallel for num_threads(OMP_THREADS()) schedule(static)
    for (int j = 1; j <= NYP; j++) {
        double ql[CR_MAXL];
        for (int i = 1; i <= NXP; i++) {
            for (int k = 1; k <= nl; k++) ql[k - 1] = E3(b->q, i, j, k) - b->P.beta * b->yprel[j - 1];
            ql[kt - 1] = ql[kt - 1] - E2(b->ddyn, i, j);
            for (int m = 1; m <= nl; m++) {
                double s = 0.0;
                for (int k = 1; k <= nl; k++) s += ql[k - 1] * M2(b->ctl2m, k, m);
                E3(rhs, i, j, m) = b->P.fnot * s;
            }
        }
    }
    for (int m = 1; m <= nl; m++) cr_solve_inhomog(b, m, rhs + (m - 1) * nxy, inh + (m - 1) * nxy);
    /* update_mass_constr (constraint.F90:167-179) with ent_xn(1) = int(ent1) */
    double ent_xn1 = cr_int_P_dA(b, ent);
    for (int l = 0; l < nl - 1; l++) {
        double aitmp = b->dpi[l];
        double exn = (l == 0) ? ent_xn1 : 0.0;
        b->dpi[l] = b->dpip[l] - tdt * b->P.gp[l] * exn;
        b->dpip[l] = aitmp;
    }
    if (b->P.cyclic) {
        /* cyclic_homog (homog.F90:280-329) */
        step_constraints(b, tdt);
        double clhss[CR_MAXL], clhsn[CR_MAXL];
        for (int m = 1; m <= nl; m++) {
            double s = 0.0, n = 0.0;
            for (int k = 1; k <= nl; k++) { s += M2(b->ctl2m, k, m) * b->cs[k - 1]; n += M2(b->ctl2m, k, m) * b->cn[k - 1]; }
            clhss[m - 1] = s + cr_int_P_dx(b, &E3(inh, 1, 2, m)) / b->P.dy;
            clhsn[m - 1] = n + cr_int_P_dx(b, &E3(inh, 1, NYP - 1, m)) / b->P.dy;
        }
        double c_bt = clhss[0] * b->hbsi;
        double c1[CR_MAXL], c2[CR_MAXL];
        for (int m = 2; m <= nl; m++) {
            c1[m - 1] = b->hc2n[m - 1] * clhss[m - 1] - b->hc2s[m - 1] * clhsn[m - 1];
            c2[m - 1] = b->hc1s[m - 1] * clhsn[m - 1] - b->hc1n[m - 1] * clhss[m - 1];
        }
        for (int j = 1; j <= NYP; j++) {
            double hc[CR_MAXL];
            hc[0] = c_bt * b->hom_bt[j - 1];
            for (int m = 2; m <= nl; m++)
                hc[m - 1] = c1[m - 1] * b->hom_bc1[(size_t)(m - 2) * NYP + j - 1] +
                            c2[m - 1] * b->hom_bc2[(size_t)(m - 2) * NYP + j - 1];
            for (int m = 1; m <= nl; m++)
                for (int i = 1; i <= NXP; i++) E3(homcor, i, j, m) = hc[m - 1];
        }
    } else {
        /* box_homog (homog.F90:331-366) */
        double im[CR_MAXL], ik[CR_MAXL], r[CR_MAXL], alpha[CR_MAXL];
        for (int m = 1; m <= nl; m++) im[m - 1] = cr_int_P_dA(b, inh + (m - 1) * nxy);
        for (int k = 1; k <= nl; k++) {
            double s = 0.0;
            for (int m = 1; m <= nl; m++) s += M2(b->ctm2l, m, k) * im[m - 1];
            ik[k - 1] = s;
        }
        for (int l = 1; l <= nl - 1; l++) r[l - 1] = b->dpi[l - 1] - (ik[l] - ik[l - 1]);
        if (nl > 1) solve_Ax_b(nl - 1, b->cdhoc, b->LU, b->ipiv, r, alpha);
        for (size_t t = 0; t < nxy; t++) homcor[t] = 0.0;
        for (int m = 2; m <= nl; m++) {
            const double* hs = b->hom_sol_bc + (size_t)(m - 2) * nxy;
            double* hcm = homcor + (size_t)(m - 1) * nxy;
            for (size_t t = 0; t < nxy; t++) hcm[t] = alpha[m - 2] * hs[t];
        }
    }
    /* modes_to_layers(inhomog + homcor) (modes.F90:268-284) */
#pragma omp parallel for num_threads(OMP_THREADS()) schedule(static)
    for (int j = 1; j <= NYP; j++)
        for (int i = 1; i <= NXP; i++) {
            double v[CR_MAXL];
            for (int m = 1; m <= nl; m++) v[m - 1] = E3(inh, i, j, m) + E3(homcor, i, j, m);
            for (int k = 1; k <= nl; k++) {
                double s = 0.0;
                for (int m = 1; m <= nl; m++) s += v[m - 1] * M2(b->ctm2l, m, k);
                E3(b->p, i, j, k) = s;
            }
        }
}

/* ocqbdy (vorsubs.F90:58-114) */
static void ocqbdy(cr_basin* b) {
    const int nl = b->nl, kt = b->P.k_topo, cyc = b->P.cyclic;
    double* q = b->q;
    const double* p = b->p;
    const double fac = b->P.bcco * b->dxm2 / (0.5 * b->P.bcco + 1.0) / b->P.fnot;
    for (int k = 1; k <= nl; k++)
        for (int i = 1; i <= NXP; i++) {
            E3(q, i, 1, k) = fac * (E3(p, i, 2, k) - E3(p, i, 1, k));
            E3(q, i, NYP, k) = fac * (E3(p, i, NYP - 1, k) - E3(p, i, NYP, k));
        }
    if (!cyc)
        for (int k = 1; k <= nl; k++)
            for (int j = 1; j <= NYP; j++) {
                E3(q, 1, j, k) = fac * (E3(p, 2, j, k) - E3(p, 1, j, k));
                E3(q, NXP, j, k) = fac * (E3(p, NXP - 1, j, k) - E3(p, NXP, j, k));
            }
#define SUB_AP(ii, jj)                                                                        \
    do {                                                                                      \
        double pv[CR_MAXL];                                                                   \
        for (int l = 1; l <= nl; l++) pv[l - 1] = E3(p, ii, jj, l);                           \
        for (int k = 1; k <= nl; k++) {                                                       \
            double s = 0.0;                                                                   \
            for (int l = 1; l <= nl; l++) s += M2(b->amat, k, l) * pv[l - 1];                 \
            E3(q, ii, jj, k) = E3(q, ii, jj, k) - b->P.fnot * s;                              \
        }                                                                                     \
    } while (0)
    for (int i = 1; i <= NXP; i++) { SUB_AP(i, 1); SUB_AP(i, NYP); }
    if (!cyc)
        for (int j = 1; j <= NYP; j++) { SUB_AP(1, j); SUB_AP(NXP, j); }
#undef SUB_AP
    for (int k = 1; k <= nl; k++)
        for (int i = 1; i <= NXP; i++) {
            E3(q, i, 1, k) = E3(q, i, 1, k) + b->P.beta * b->yprel[0];
            E3(q, i, NYP, k) = E3(q, i, NYP, k) + b->P.beta * b->yprel[NYP - 1];
        }
    if (!cyc)
        for (int j = 1; j <= NYP; j++)
            for (int k = 1; k <= nl; k++) {
                E3(q, 1, j, k) = E3(q, 1, j, k) + b->P.beta * b->yprel[j - 1];
                E3(q, NXP, j, k) = E3(q, NXP, j, k) + b->P.beta * b->yprel[j - 1];
            }
    for (int i = 1; i <= NXP; i++) {
        E3(q, i, 1, kt) = E3(q, i, 1, kt) + E2(b->ddyn, i, 1);
        E3(q, i, NYP, kt) = E3(q, i, NYP, kt) + E2(b->ddyn, i, NYP);
    }
    if (!cyc)
        for (int j = 1; j <= NYP; j++) {
            E3(q, 1, j, kt) = E3(q, 1, j, kt) + E2(b->ddyn, 1, j);
            E3(q, NXP, j, kt) = E3(q, NXP, j, kt) + E2(b->ddyn, NXP, j);
        }
}

void cr_ocqbdy(cr_basin* b) { ocqbdy(b); }

void cr_step(cr_basin* b, double tdt, const double* ent, int n_ent, const double* wekp,
             double txis, double txin) {
    vort_step(b, tdt, ent, n_ent, wekp, txis, txin);
    solve_pressure(b, tdt, ent);
    ocqbdy(b);
}

/* suppress_comp_mode (driver.F90:174-189) */
void cr_suppress(cr_basin* b) {
    size_t nf = b->nxy * b->nl;
    for (size_t t = 0; t < nf; t++) {
        b->q[t] = 0.5 * (b->q[t] + b->qm[t]);
        b->p[t] = 0.5 * (b->p[t] + b->pm[t]);
    }
    for (int l = 0; l < b->nl - 1; l++) b->dpi[l] = 0.5 * (b->dpi[l] + b->dpip[l]);
    if (b->P.cyclic)
        for (int k = 0; k < b->nl; k++) {
            b->cs[k] = 0.5 * (b->cs[k] + b->csp[k]);
            b->cn[k] = 0.5 * (b->cn[k] + b->cnp[k]);
        }
}

/* ------------------------------------------------------------------ */
/* Stand-alone transforms                                             */
/* ------------------------------------------------------------------ */

void cr_dst1(int n, int count, const double* in, double* out) {
    fftplan p;
    fft_plan_init(&p, 2 * (n + 1));
    double* z = dalloc(8 * (size_t)(n + 1));
    for (int c = 0; c < count; c++)
        dst1_pair(&p, n, in + (size_t)c * n, NULL, out + (size_t)c * n, NULL, z, z + 4 * (n + 1));
    free(z);
    fft_plan_free(&p);
}

void cr_rfft(int n, int count, const double* in, double* out_c) {
    fftplan p;
    fft_plan_init(&p, n);
    double* z = dalloc(4 * (size_t)n);
    for (int c = 0; c < count; c++)
        rfft_pair(&p, n, in + (size_t)c * n, NULL, out_c + (size_t)c * 2 * (n / 2 + 1), NULL, z, z + 2 * n);
    free(z);
    fft_plan_free(&p);
}

void cr_irfft(int n, int count, const double* in_c, double* out) {
    fftplan p;
    fft_plan_init(&p, n);
    double* z = dalloc(4 * (size_t)n);
    for (int c = 0; c < count; c++)
        irfft_pair(&p, n, in_c + (size_t)c * 2 * (n / 2 + 1), NULL, out + (size_t)c * n, NULL, z, z + 2 * n);
    free(z);
    fft_plan_free(&p);
}

void cr_cfft(int n, int sign, int count, const double* in, double* out) {
    fftplan p;
    fft_plan_init(&p, n);
    for (int c = 0; c < count; c++) fft_exec(&p, in + (size_t)c * 2 * n, out + (size_t)c * 2 * n, sign);
    fft_plan_free(&p);
}

/* ------------------------------------------------------------------ */
/* Wind forcing: ekman_from_tau (ekman.F90:53-83), map_T_to_P (numerics.F90:64-107) */
/* ------------------------------------------------------------------ */

void cr_ekman_wekp(const cr_params* prm, double hm, const double* taux, const double* tauy,
                   double* wekp, double* txis, double* txin) {
    (void)hm;
    const int nxp = prm->nxp, nyp = prm->nyp, nxt = nxp - 1, nyt = nyp - 1;
    const double dx = prm->dx, dy = prm->dy, f0 = prm->fnot;
#define P2(a, i, j) (a)[((size_t)(j) - 1) * nxp + ((i) - 1)]
#define T2(a, i, j) (a)[((size_t)(j) - 1) * nxt + ((i) - 1)]
    double* wekt = dalloc((size_t)nxt * nyt);
    for (int j = 1; j <= nyt; j++)
        for (int i = 1; i <= nxt; i++) {
            /* taux_t = map_P_to_x_face (nxp-1, nyp); tauy_t = map_P_to_y_face (nxp, nyp-1) */
            double tyt_ip = 0.5 * (P2(tauy, i + 1, j + 1) + P2(tauy, i + 1, j));
            double tyt_i = 0.5 * (P2(tauy, i, j + 1) + P2(tauy, i, j));
            double txt_jp = 0.5 * (P2(taux, i + 1, j + 1) + P2(taux, i, j + 1));
            double txt_j = 0.5 * (P2(taux, i + 1, j) + P2(taux, i, j));
            T2(wekt, i, j) = ((1.0 / (dx * f0)) * (tyt_ip - tyt_i) - (1.0 / (dy * f0)) * (txt_jp - txt_j));
        }
    for (int j = 2; j <= nyp - 1; j++)
        for (int i = 2; i <= nxp - 1; i++)
            P2(wekp, i, j) = 0.25 * (T2(wekt, i, j) + T2(wekt, i - 1, j) + T2(wekt, i, j - 1) + T2(wekt, i - 1, j - 1));
    for (int i = 2; i <= nxp - 1; i++) {
        P2(wekp, i, 1) = 0.5 * (T2(wekt, i - 1, 1) + T2(wekt, i, 1));
        P2(wekp, i, nyp) = 0.5 * (T2(wekt, i - 1, nyp - 1) + T2(wekt, i, nyp - 1));
    }
    if (prm->cyclic) {
        for (int j = 2; j <= nyp - 1; j++) {
            P2(wekp, 1, j) = 0.25 * (T2(wekt, 1, j) + T2(wekt, nxp - 1, j) + T2(wekt, 1, j - 1) + T2(wekt, nxp - 1, j - 1));
            P2(wekp, nxp, j) = P2(wekp, 1, j);
        }
        P2(wekp, 1, 1) = 0.5 * (T2(wekt, 1, 1) + T2(wekt, nxp - 1, 1));
        P2(wekp, 1, nyp) = 0.5 * (T2(wekt, 1, nyp - 1) + T2(wekt, nxp - 1, nyp - 1));
        P2(wekp, nxp, 1) = P2(wekp, 1, 1);
        P2(wekp, nxp, nyp) = P2(wekp, 1, nyp);
    } else {
        for (int j = 2; j <= nyp - 1; j++) {
            P2(wekp, 1, j) = 0.5 * (T2(wekt, 1, j - 1) + T2(wekt, 1, j));
            P2(wekp, nxp, j) = 0.5 * (T2(wekt, nxp - 1, j - 1) + T2(wekt, nxp - 1, j));
        }
        P2(wekp, 1, 1) = T2(wekt, 1, 1);
        P2(wekp, nxp, 1) = T2(wekt, nxp - 1, 1);
        P2(wekp, 1, nyp) = T2(wekt, 1, nyp - 1);
        P2(wekp, nxp, nyp) = T2(wekt, nxp - 1, nyp - 1);
    }
    free(wekt);
    double sx = 0.0, nx = 0.0;
    if (prm->cyclic) {
        double* row = dalloc((size_t)nxp);
        double s = 0.0;
        for (int i = 1; i <= nxp; i++) row[i - 1] = P2(taux, i, 1) + P2(taux, i, 2);
        for (int i = 2; i <= nxp - 1; i++) s += row[i - 1];
        sx = 0.5 * ((0.5 * row[0] + s + 0.5 * row[nxp - 1]) * dx);
        s = 0.0;
        for (int i = 1; i <= nxp; i++) row[i - 1] = P2(taux, i, nyp - 1) + P2(taux, i, nyp);
        for (int i = 2; i <= nxp - 1; i++) s += row[i - 1];
        nx = 0.5 * ((0.5 * row[0] + s + 0.5 * row[nxp - 1]) * dx);
        free(row);
    }
    if (txis) *txis = sx;
    if (txin) *txin = nx;
#undef P2
#undef T2
}

/* ------------------------------------------------------------------ */
/* Ocean mixed layer: step_oml (omlsubs.F90:69-119) with adv_diff     */
/* (mixed_layer.F90:16-95) and the Ekman velocities of ekman_from_tau */
/* (ekman.F90:53-83).  T-point fields are (nxt,nyt), p-point fields   */
/* (nxp,nyp), both x fastest.                                         */
/* ------------------------------------------------------------------ */

#define TT(a, i, j) (a)[(size_t)((j) - 1) * nxt + ((i) - 1)]
#define PP(a, i, j) (a)[(size_t)((j) - 1) * nxp + ((i) - 1)]

/* del2_T (numerics.F90:107-147): Del^2 on T points with zero-gradient or
 * flux (fixed outside value) conditions in y, cyclic or no-flux in x. */
void cr_del2_T(const cr_params* P, const double* t, int north_flux, double nval, int south_flux, double sval,
               double* out) {
    const int nxp = P->nxp, nyp = P->nyp, nxt = nxp - 1, nyt = nyp - 1;
    (void)nxp;
    const double f = 1.0 / (P->dx * P->dx);
    for (int j = 1; j <= nyt; j++)
        for (int i = 1; i <= nxt; i++) {
            double dxl, dxr, dyl, dyr;
            /* dt_x(i) = t(i) - t(i-1), dt_x(i+1) = t(i+1) - t(i) */
            if (i >= 2) dxl = TT(t, i, j) - TT(t, i - 1, j);
            else dxl = P->cyclic ? TT(t, 1, j) - TT(t, nxt, j) : 0.0;
            if (i <= nxt - 1) dxr = TT(t, i + 1, j) - TT(t, i, j);
            else dxr = P->cyclic ? TT(t, 1, j) - TT(t, nxt, j) : 0.0;
            if (j >= 2) dyl = TT(t, i, j) - TT(t, i, j - 1);
            else dyl = south_flux ? TT(t, i, 1) - sval : 0.0;
            if (j <= nyt - 1) dyr = TT(t, i, j + 1) - TT(t, i, j);
            else dyr = north_flux ? nval - TT(t, i, nyt) : 0.0;
            TT(out, i, j) = f * (dxr - dxl + dyr - dyl);
        }
}

/* Ekman velocities (ekman.F90:73-80): uek on x faces (nxp,nyt), vek on y
 * faces (nxt,nyp), wekt on T points (nxt,nyt), from the kinematic stress
 * at p points. */
void cr_ekman_uvw(const cr_params* P, double hm, const double* taux, const double* tauy, double* uek, double* vek,
                  double* wekt) {
    const int nxp = P->nxp, nyp = P->nyp, nxt = nxp - 1, nyt = nyp - 1;
    const double cu = P->tau_sign / (P->fnot * hm), cv = -P->tau_sign / (P->fnot * hm);
    for (int j = 1; j <= nyt; j++)
        for (int i = 1; i <= nxp; i++)
            uek[(size_t)(j - 1) * nxp + i - 1] = cu * (0.5 * (PP(tauy, i, j + 1) + PP(tauy, i, j)));
    for (int j = 1; j <= nyp; j++)
        for (int i = 1; i <= nxt; i++)
            vek[(size_t)(j - 1) * nxt + i - 1] = cv * (0.5 * (PP(taux, i + 1, j) + PP(taux, i, j)));
    for (int j = 1; j <= nyt; j++)
        for (int i = 1; i <= nxt; i++) {
            const double tyr = 0.5 * (PP(tauy, i + 1, j + 1) + PP(tauy, i + 1, j));
            const double tyl = 0.5 * (PP(tauy, i, j + 1) + PP(tauy, i, j));
            const double txt = 0.5 * (PP(taux, i + 1, j + 1) + PP(taux, i, j + 1));
            const double txb = 0.5 * (PP(taux, i + 1, j) + PP(taux, i, j));
            TT(wekt, i, j) = (1.0 / (P->dx * P->fnot)) * (tyr - tyl) - (1.0 / (P->dy * P->fnot)) * (txt - txb);
        }
}

/* adv_diff (mixed_layer.F90:16-57) for the mixed-layer field (data, datam)
 * advected by the geostrophic velocity of p (layer 1) plus the Ekman
 * velocity, with Del^2/Del^4 diffusion of datam. */
void cr_adv_diff(const cr_params* P, const cr_oml_params* o, const double* p, const double* uek, const double* vek,
                 const double* data, const double* datam, double* out) {
    const int nxp = P->nxp, nyp = P->nyp, nxt = nxp - 1, nyt = nyp - 1;
    double* d2 = (double*)calloc((size_t)nxt * nyt, sizeof(double));
    double* d4 = (double*)calloc((size_t)nxt * nyt, sizeof(double));
    cr_del2_T(P, datam, o->fluxn, o->nval, o->fluxs, o->sval, d2);
    cr_del2_T(P, d2, 0, 0.0, 0, 0.0, d4);
    for (int j = 1; j <= nyt; j++)
        for (int i = 1; i <= nxt; i++) {
            double ut[2], vt[2];
            for (int s = 0; s < 2; s++) {
                /* x face I = i + s (adv_vel, mixed_layer.F90:68-74; map_T_to_x_face, numerics.F90:250-266) */
                const int I = i + s;
                const double dpdy = (1.0 / P->dy) * (PP(p, I, j + 1) - PP(p, I, j));
                double vu = (-o->ycexp / P->fnot) * dpdy + uek[(size_t)(j - 1) * nxp + I - 1];
                if (!P->cyclic && (I == 1 || I == nxp)) vu = 0.0;
                double xf;
                if (I == 1) xf = P->cyclic ? 0.5 * (TT(data, 1, j) + TT(data, nxt, j)) : TT(data, 1, j);
                else if (I == nxt + 1) xf = P->cyclic ? 0.5 * (TT(data, 1, j) + TT(data, nxt, j)) : TT(data, nxt, j);
                else xf = 0.5 * (TT(data, I - 1, j) + TT(data, I, j));
                ut[s] = vu * xf;
                /* y face J = j + s (adv_vel :76-92; map_T_to_y_face, numerics.F90:222-248) */
                const int J = j + s;
                const double dpdx = (1.0 / P->dx) * (PP(p, i + 1, J) - PP(p, i, J));
                const double vk = vek[(size_t)(J - 1) * nxt + i - 1];
                double vv = (o->ycexp / P->fnot) * dpdx + vk;
                if (J == 1) vv = o->fluxs ? vk : 0.0;
                if (J == nyp) vv = o->fluxn ? vk : 0.0;
                double yf;
                if (J == 1) yf = o->fluxs ? TT(data, i, 1) : 0.5 * (TT(data, i, 1) + o->sval);
                else if (J == nyt + 1) yf = o->fluxn ? TT(data, i, nyt) : 0.5 * (TT(data, i, nyt) + o->nval);
                else yf = 0.5 * (TT(data, i, J - 1) + TT(data, i, J));
                vt[s] = vv * yf;
            }
            /* c_grid_advection (numerics.F90:392-408) */
            double a = (-1.0 / P->dx) * (ut[1] - ut[0]) + (-1.0 / P->dy) * (vt[1] - vt[0]);
            TT(out, i, j) = a + o->d2 * TT(d2, i, j) - o->d4 * TT(d4, i, j);
        }
    free(d2);
    free(d4);
}

/* map_T_to_P (numerics.F90:64-107). */
void cr_map_T_to_P(const cr_params* P, const double* t, double* out) {
    const int nxp = P->nxp, nyp = P->nyp, nxt = nxp - 1;
    for (int j = 2; j <= nyp - 1; j++)
        for (int i = 2; i <= nxp - 1; i++)
            PP(out, i, j) = 0.25 * (TT(t, i, j) + TT(t, i - 1, j) + TT(t, i, j - 1) + TT(t, i - 1, j - 1));
    for (int i = 2; i <= nxp - 1; i++) {
        PP(out, i, 1) = 0.5 * (TT(t, i - 1, 1) + TT(t, i, 1));
        PP(out, i, nyp) = 0.5 * (TT(t, i - 1, nyp - 1) + TT(t, i, nyp - 1));
    }
    for (int j = 2; j <= nyp - 1; j++) {
        if (P->cyclic) {
            PP(out, 1, j) = 0.25 * (TT(t, 1, j) + TT(t, nxp - 1, j) + TT(t, 1, j - 1) + TT(t, nxp - 1, j - 1));
            PP(out, nxp, j) = PP(out, 1, j);
        } else {
            PP(out, 1, j) = 0.5 * (TT(t, 1, j - 1) + TT(t, 1, j));
            PP(out, nxp, j) = 0.5 * (TT(t, nxp - 1, j - 1) + TT(t, nxp - 1, j));
        }
    }
    if (P->cyclic) {
        PP(out, 1, 1) = 0.5 * (TT(t, 1, 1) + TT(t, nxp - 1, 1));
        PP(out, 1, nyp) = 0.5 * (TT(t, 1, nyp - 1) + TT(t, nxp - 1, nyp - 1));
        PP(out, nxp, 1) = PP(out, 1, 1);
        PP(out, nxp, nyp) = PP(out, 1, nyp);
    } else {
        PP(out, 1, 1) = TT(t, 1, 1);
        PP(out, nxp, 1) = TT(t, nxp - 1, 1);
        PP(out, 1, nyp) = TT(t, 1, nyp - 1);
        PP(out, nxp, nyp) = TT(t, nxp - 1, nyp - 1);
    }
}

/* step_oml (omlsubs.F90:69-119): new sst, the convective correction and, with
 * compute_ent, the layer 1/2 entrainment ent1 (nxp,nyp) with zero mean. */
void cr_oml_step(const cr_params* P, const cr_oml_params* o, const double* p1, const double* taux, const double* tauy,
                 const double* fnet, double tdto, int compute_ent, double* sst, double* sstm, double* ent1) {
    const int nxp = P->nxp, nyp = P->nyp, nxt = nxp - 1, nyt = nyp - 1;
    const size_t nt = (size_t)nxt * nyt;
    double* uek = (double*)calloc((size_t)nxp * nyt, sizeof(double));
    double* vek = (double*)calloc((size_t)nxt * nyp, sizeof(double));
    double* wekt = (double*)calloc(nt, sizeof(double));
    double* ad = (double*)calloc(nt, sizeof(double));
    double* snew = (double*)calloc(nt, sizeof(double));
    double* dtc = (double*)calloc(nt, sizeof(double));
    double* xfo = (double*)calloc(nt, sizeof(double));
    cr_ekman_uvw(P, o->hm, taux, tauy, uek, vek, wekt);
    cr_adv_diff(P, o, p1, uek, vek, sst, sstm, ad);
    for (size_t t = 0; t < nt; t++) {
        const double half = 0.5 * wekt[t] * (sstm[t] - o->T1_rel);
        snew[t] = sstm[t] + tdto * (ad[t] + (1.0 / (o->rho_cp * o->hm)) * fnet[t] +
                                    (1.0 / o->hm) * (wekt[t] * sstm[t] - half));
        dtc[t] = fmax(0.0, o->T1_rel - snew[t]);
        xfo[t] = (-1.0 / o->dT_12) * half - (o->hm / (tdto * o->dT_12)) * dtc[t];
    }
    if (compute_ent) {
        /* avg_T (numerics.F90:155-162): the sum in Fortran's column order */
        double s = 0.0;
        for (size_t t = 0; t < nt; t++) s += xfo[t];
        const double avg = s / (double)(nxt * nyt);
        for (size_t t = 0; t < nt; t++) xfo[t] = xfo[t] - avg;
        cr_map_T_to_P(P, xfo, ent1);
    }
    for (size_t t = 0; t < nt; t++) {
        sstm[t] = sst[t];
        sst[t] = snew[t] + dtc[t];
    }
    free(uek); free(vek); free(wekt); free(ad); free(snew); free(dtc); free(xfo);
}
#undef TT
#undef PP

/* ------------------------------------------------------------------ */
/* Coupling: grid, coupled stress, heat-flux forcing, atmospheric mixed layer */

#define TT(a, i, j) (a)[(size_t)((j) - 1) * nxt + ((i) - 1)]
#define PP(a, i, j) (a)[(size_t)((j) - 1) * nxp + ((i) - 1)]
#define CR_PI 3.14159265358979324

/* load_grid (grid.F90:37-58): ndxr = nint(dxa/dxo), the ocean centred in y. */
void cr_load_grid(const cr_params* Po, const cr_params* Pa, cr_cpl* g) {
    const int nxto = Po->nxp - 1, nyto = Po->nyp - 1, nyta = Pa->nyp - 1;
    g->ndxr = (int)lround(Pa->dx / Po->dx);
    g->nxtoar = nxto / g->ndxr;
    g->nytoar = nyto / g->ndxr;
    g->nxpoar = g->nxtoar + 1;
    g->nypoar = g->nytoar + 1;
    g->nxaooc = nxto / g->ndxr;
    g->nyaooc = nyto / g->ndxr;
    g->nx1 = 1;
    g->ny1 = 1 + (nyta - g->nyaooc) / 2;
}

/* bcuini (bicubic.F90:97-146).  stinv maps the 16 corner data of a cell (values,
 * then x, y and cross derivatives, each at corners (0,0),(1,0),(0,1),(1,1)) to
 * the coefficients of s^i t^j (k = 1 + i + 4j).  It is the tensor product of the
 * cubic Hermite basis on [0,1]: h00 = 1 - 3s^2 + 2s^3, h01 = 3s^2 - 2s^3,
 * h10 = s - 2s^2 + s^3, h11 = -s^2 + s^3 (value at 0, value at 1, slope at 0,
 * slope at 1); st = stfn * stinv with stfn(r, k) = ss^i tt^j at fine point r. */
void cr_bcuini(int ndxr, double* st) {
    static const double herm[4][4] = {{1, 0, -3, 2}, {0, 0, 3, -2}, {0, 1, -2, 1}, {0, 0, -1, 1}};
    double stinv[16][16]; /* [k][c] */
    for (int c = 0; c < 16; c++) {
        const int kind = c / 4, corner = c % 4;          /* kind: value, d/ds, d/dt, d2/dsdt */
        const int cx = corner & 1, cy = corner >> 1;
        const int dx_ = (kind == 1 || kind == 3), dy_ = (kind == 2 || kind == 3);
        const double* bx = herm[2 * dx_ + cx];
        const double* by = herm[2 * dy_ + cy];
        for (int j = 0; j < 4; j++)
            for (int i = 0; i < 4; i++) stinv[i + 4 * j][c] = bx[i] * by[j];
    }
    const int nr = (ndxr + 1) * (ndxr + 1);
    for (int jp = 0; jp <= ndxr; jp++) {
        const double tt = (double)jp / (double)ndxr;
        for (int ip = 0; ip <= ndxr; ip++) {
            const double ss = (double)ip / (double)ndxr;
            double fn[16];
            double sp[4] = {1.0, ss, ss * ss, ss * ss * ss}, tp[4] = {1.0, tt, tt * tt, tt * tt * tt};
            for (int j = 0; j < 4; j++)
                for (int i = 0; i < 4; i++) fn[i + 4 * j] = sp[i] * tp[j];
            const int r = ip + jp * (ndxr + 1);
            for (int c = 0; c < 16; c++) {
                double s = 0.0;
                for (int k = 0; k < 16; k++) s += fn[k] * stinv[k][c];
                st[(size_t)c * nr + r] = s;
            }
        }
    }
}

/* dPdx_2bc (numerics.F90:470-492): centred, one-sided with the mixed
 * condition at solid walls, periodic in a channel, zero on the y walls. */
void cr_dPdx_2bc(const cr_params* P, double alpha, const double* p, double* out) {
    const int nxp = P->nxp, nyp = P->nyp;
    for (int j = 1; j <= nyp; j++) {
        for (int i = 2; i <= nxp - 1; i++) PP(out, i, j) = (PP(p, i + 1, j) - PP(p, i - 1, j)) / (2.0 * P->dx);
        if (!P->cyclic) {
            PP(out, 1, j) = (PP(p, 2, j) - PP(p, 1, j)) / (P->dx * (0.5 * alpha + 1.0));
            PP(out, nxp, j) = (PP(p, nxp, j) - PP(p, nxp - 1, j)) / (P->dx * (0.5 * alpha + 1.0));
        } else {
            PP(out, 1, j) = (PP(p, 2, j) - PP(p, nxp - 1, j)) / (2.0 * P->dx);
            PP(out, nxp, j) = PP(out, 1, j);
        }
    }
    for (int i = 1; i <= nxp; i++) PP(out, i, 1) = PP(out, i, nyp) = 0.0;
}

/* dPdy_2bc (numerics.F90:494-513). */
void cr_dPdy_2bc(const cr_params* P, double alpha, const double* p, double* out) {
    const int nxp = P->nxp, nyp = P->nyp;
    for (int i = 1; i <= nxp; i++) {
        for (int j = 2; j <= nyp - 1; j++) PP(out, i, j) = (PP(p, i, j + 1) - PP(p, i, j - 1)) / (2.0 * P->dy);
        PP(out, i, 1) = (PP(p, i, 2) - PP(p, i, 1)) / (P->dy * (0.5 * alpha + 1.0));
        PP(out, i, nyp) = (PP(p, i, nyp) - PP(p, i, nyp - 1)) / (P->dy * (0.5 * alpha + 1.0));
    }
    if (!P->cyclic)
        for (int j = 1; j <= nyp; j++) PP(out, 1, j) = PP(out, nxp, j) = 0.0;
}

/* dP2dx2_bc (numerics.F90:427-448). */
void cr_dP2dx2_bc(const cr_params* P, double alpha, const double* p, double* out) {
    const int nxp = P->nxp, nyp = P->nyp;
    const double fac = 1.0 / (P->dx * P->dx);
    for (int j = 1; j <= nyp; j++) {
        for (int i = 2; i <= nxp - 1; i++)
            PP(out, i, j) = fac * (PP(p, i - 1, j) + PP(p, i + 1, j) - 2.0 * PP(p, i, j));
        if (!P->cyclic) {
            PP(out, 1, j) = fac * alpha / (0.5 * alpha + 1.0) * (PP(p, 2, j) - PP(p, 1, j));
            PP(out, nxp, j) = -fac * alpha / (0.5 * alpha + 1.0) * (PP(p, nxp, j) - PP(p, nxp - 1, j));
        } else {
            PP(out, 1, j) = fac * (PP(p, nxp - 1, j) + PP(p, 2, j) - 2.0 * PP(p, 1, j));
            PP(out, nxp, j) = PP(out, 1, j);
        }
    }
    for (int i = 1; i <= nxp; i++) PP(out, i, 1) = PP(out, i, nyp) = 0.0;
}

/* dP2dy2_bc (numerics.F90:450-468). */
void cr_dP2dy2_bc(const cr_params* P, double alpha, const double* p, double* out) {
    const int nxp = P->nxp, nyp = P->nyp;
    const double fac = 1.0 / (P->dy * P->dy);
    for (int i = 1; i <= nxp; i++) {
        for (int j = 2; j <= nyp - 1; j++)
            PP(out, i, j) = fac * (PP(p, i, j - 1) + PP(p, i, j + 1) - 2.0 * PP(p, i, j));
        PP(out, i, 1) = fac * alpha / (0.5 * alpha + 1.0) * (PP(p, i, 2) - PP(p, i, 1));
        PP(out, i, nyp) = -fac * alpha / (0.5 * alpha + 1.0) * (PP(p, i, nyp) - PP(p, i, nyp - 1));
    }
    if (!P->cyclic)
        for (int j = 1; j <= nyp; j++) PP(out, 1, j) = PP(out, nxp, j) = 0.0;
}

/* new_bicubic (bicubic.F90:26-95): u, ux, uy, uxy on the (nxpoar,nypoar)
 * coarse p points over the ocean (derivatives scaled by the coarse spacing);
 * out (nxpo,nypo) on the fine grid, a point on a cell edge taking the value of
 * the last cell written (bc3's loop order: j outer, i inner). */
void cr_new_bicubic(const cr_cpl* g, int nxpo, int nypo, const double* st, const double* u, const double* ux,
                    const double* uy, const double* uxy, double* out) {
    const int n = g->ndxr, nr = (n + 1) * (n + 1), nxa = g->nxpoar;
    const double* f[4] = {u, ux, uy, uxy};
    for (int j = 1; j <= g->nytoar; j++)
        for (int i = 1; i <= g->nxtoar; i++) {
            double d[16];
            for (int t = 0; t < 4; t++)   /* bc1 (bicubic.F90:79-95) */
                for (int jq = 0; jq <= 1; jq++)
                    for (int ip = 0; ip <= 1; ip++)
                        d[4 * t + ip + 2 * jq] = f[t][(size_t)(j + jq - 1) * nxa + (i + ip - 1)];
            const int iout = 1 + (i - 1) * n, jout = 1 + (j - 1) * n;
            for (int jp = 0; jp <= n; jp++)
                for (int ip = 0; ip <= n; ip++) {
                    const int r = ip + jp * (n + 1);
                    double s = 0.0;
                    for (int k = 0; k < 16; k++) s += st[(size_t)k * nr + r] * d[k];   /* bc2 */
                    const int io = iout + ip, jo = jout + jp;
                    if (io <= nxpo && jo <= nypo) out[(size_t)(jo - 1) * nxpo + (io - 1)] = s;   /* bc3 */
                }
        }
}

/* xforc + compute_windstress + atmos_geos_vel + compute_udiff +
 * atmos_windstress (windstress.F90:102-316).  pmo1, pma1: layer 1 of the
 * basins' previous pressure.  Outputs the kinematic stresses at p points. */
void cr_xforc(const cr_params* Po, const cr_params* Pa, const cr_cpl* g, const double* st, const double* pmo1,
              const double* pma1, double* tauxo, double* tauyo, double* tauxa, double* tauya) {
    const int nxpa = Pa->nxp, nypa = Pa->nyp, nxpo = Po->nxp, nypo = Po->nyp;
    const size_t na = (size_t)nxpa * nypa, no = (size_t)nxpo * nypo;
    const double fa = Pa->fnot, fo = Po->fnot;
    double* dpdx = calloc(na, sizeof(double));
    double* dpdy = calloc(na, sizeof(double));
    double* dxx = calloc(na, sizeof(double));
    double* dyy = calloc(na, sizeof(double));
    double* dxy = calloc(na, sizeof(double));
    double* dxxy = calloc(na, sizeof(double));
    double* dxyy = calloc(na, sizeof(double));
    cr_dPdx_2bc(Pa, Pa->bcco, pma1, dpdx);
    cr_dPdy_2bc(Pa, Pa->bcco, pma1, dpdy);
    /* the atmosphere's own stress (windstress.F90:129-145) */
    {
        const double cdhfaa = (g->cdat / fa) / g->hma;
        const double qpreaa = 1.0 / (fabs(cdhfaa) * sqrt(2.0));
        for (size_t t = 0; t < na; t++) {
            const double u1 = -dpdy[t] / fa, v1 = dpdx[t] / fa;
            const double u2 = 4.0 * cdhfaa * cdhfaa * (u1 * u1 + v1 * v1);
            const double shr = qpreaa * sqrt(-1.0 + sqrt(1.0 + u2));
            const double cuohf = cdhfaa * shr;
            const double cdochi = g->cdat * shr / (1.0 + cuohf * cuohf);
            tauxa[t] = cdochi * (u1 - cuohf * v1);
            tauya[t] = cdochi * (v1 + cuohf * u1);
        }
    }
    /* atmos_geos_vel (windstress.F90:201-257) */
    cr_dP2dx2_bc(Pa, Pa->bcco, pma1, dxx);
    cr_dP2dy2_bc(Pa, Pa->bcco, pma1, dyy);
    cr_dPdx_2bc(Pa, Pa->bcco, dpdy, dxy);
    cr_dPdx_2bc(Pa, Pa->bcco, dxy, dxxy);
    cr_dPdx_2bc(Pa, Pa->bcco, dyy, dxyy);
    const int nxa = g->nxpoar, nya = g->nypoar;
    const size_t nc = (size_t)nxa * nya;
    double* c8 = calloc(8 * nc, sizeof(double));
    double *u = c8, *ux = c8 + nc, *uy = c8 + 2 * nc, *uxy = c8 + 3 * nc;
    double *v = c8 + 4 * nc, *vx = c8 + 5 * nc, *vy = c8 + 6 * nc, *vxy = c8 + 7 * nc;
    const double dxa = Pa->dx, dya = Pa->dy;
    for (int j = 1; j <= nya; j++)
        for (int i = 1; i <= nxa; i++) {
            const size_t a = (size_t)(g->ny1 + j - 2) * nxpa + (g->nx1 + i - 2), c = (size_t)(j - 1) * nxa + (i - 1);
            u[c] = -dpdy[a] / fa;
            v[c] = dpdx[a] / fa;
            ux[c] = -(dxa / fa) * dxy[a];
            uy[c] = -(dya / fa) * dyy[a];
            uxy[c] = -(dxa * dya / fa) * dxyy[a];
            vx[c] = (dxa / fa) * dxx[a];
            vy[c] = (dya / fa) * dxy[a];
            vxy[c] = (dxa * dya / fa) * dxxy[a];
        }
    double* u1o = calloc(no, sizeof(double));
    double* v1o = calloc(no, sizeof(double));
    cr_new_bicubic(g, nxpo, nypo, st, u, ux, uy, uxy, u1o);
    cr_new_bicubic(g, nxpo, nypo, st, v, vx, vy, vxy, v1o);
    if (g->tau_udiff) {   /* compute_udiff (windstress.F90:259-275) */
        double* ox = calloc(no, sizeof(double));
        double* oy = calloc(no, sizeof(double));
        cr_dPdx_2bc(Po, Po->bcco, pmo1, ox);
        cr_dPdy_2bc(Po, Po->bcco, pmo1, oy);
        for (size_t t = 0; t < no; t++) {
            u1o[t] = u1o[t] + oy[t] / fo;
            v1o[t] = v1o[t] - ox[t] / fo;
        }
        free(ox);
        free(oy);
    }
    /* atmos_windstress (windstress.F90:277-314) */
    const double raoro = g->rho_a / g->rho_o;
    const double cdhfaa = g->tau_udiff ? (g->cdat / fa) * (1.0 / g->hma + raoro / g->hmo) : (g->cdat / fa) / g->hma;
    const double qpreaa = 1.0 / (fabs(cdhfaa) * sqrt(2.0));
    for (size_t t = 0; t < no; t++) {
        const double shr = sqrt(-1.0 + sqrt(1.0 + 4.0 * cdhfaa * cdhfaa * (u1o[t] * u1o[t] + v1o[t] * v1o[t])));
        const double cuohf = (qpreaa * cdhfaa) * shr;
        const double cdchi = (qpreaa * g->cdat) * shr / (1.0 + cuohf * cuohf);
        const double tx = cdchi * (u1o[t] - cuohf * v1o[t]), ty = cdchi * (v1o[t] + cuohf * u1o[t]);
        u1o[t] = tx;   /* tauxaor */
        v1o[t] = ty;   /* tauyaor */
    }
    /* compute_windstress (windstress.F90:167-199) */
    if (g->tau_udiff)
        for (int ja = g->ny1; ja <= g->ny1 + g->nypoar - 1; ja++)
            for (int ia = g->nx1; ia <= g->nx1 + g->nxpoar - 1; ia++) {
                const size_t o = (size_t)((ja - g->ny1) * g->ndxr) * nxpo + (size_t)((ia - g->nx1) * g->ndxr);
                tauxa[(size_t)(ja - 1) * nxpa + ia - 1] = u1o[o];
                tauya[(size_t)(ja - 1) * nxpa + ia - 1] = v1o[o];
            }
    for (size_t t = 0; t < no; t++) {
        tauxo[t] = raoro * u1o[t];
        tauyo[t] = raoro * v1o[t];
    }
    free(dpdx); free(dpdy); free(dxx); free(dyy); free(dxy); free(dxxy); free(dxyy);
    free(c8); free(u1o); free(v1o);
}

/* bilint (numerics.F90:528-589): atmos T field onto the ocean T grid. */
void cr_bilint(const cr_params* Pa, const cr_params* Po, const cr_cpl* g, const double* atmos, double* ocean) {
    const int nxta = Pa->nxp - 1, nyta = Pa->nyp - 1, nxto = Po->nxp - 1, nyto = Po->nyp - 1;
    const double dxainv = 1.0 / Pa->dx, dyainv = 1.0 / Pa->dy;
    const double xta1 = (g->xa0 + 0.0 * Pa->dx) + 0.5 * Pa->dx, yta1 = (g->ya0 + 0.0 * Pa->dy) + 0.5 * Pa->dy;
    int* iam = malloc(sizeof(int) * nxto);
    int* iap = malloc(sizeof(int) * nxto);
    double* wpx = malloc(sizeof(double) * nxto);
    double* wmx = malloc(sizeof(double) * nxto);
    for (int io = 1; io <= nxto; io++) {
        const double xto = (g->xo0 + (io - 1) * Po->dx) + 0.5 * Po->dx;
        int m = (int)(1.0 + dxainv * (xto - xta1));
        int pi_ = m + 1;
        double xam;
        if (m >= 1) xam = (g->xa0 + (m - 1) * Pa->dx) + 0.5 * Pa->dx;
        else xam = xta1 - Pa->dx;
        wpx[io - 1] = dxainv * (xto - xam);
        wmx[io - 1] = 1.0 - wpx[io - 1];
        m = 1 + (m + nxta - 1) % nxta;
        pi_ = 1 + (pi_ + nxta - 1) % nxta;
        iam[io - 1] = m;
        iap[io - 1] = pi_;
    }
    for (int jo = 1; jo <= nyto; jo++) {
        const double yto = (g->yo0 + (jo - 1) * Po->dy) + 0.5 * Po->dy;
        int jam = (int)(1.0 + dyainv * (yto - yta1));
        int jap = jam + 1;
        jam = jam > 1 ? jam : 1;
        jap = jap < nyta ? jap : nyta;
        const double yam = (g->ya0 + (jam - 1) * Pa->dy) + 0.5 * Pa->dy;
        const double wpy = dyainv * (yto - yam), wmy = 1.0 - wpy;
        for (int io = 1; io <= nxto; io++) {
            const int a = iam[io - 1], b = iap[io - 1];
            const double wm = wmx[io - 1], wp = wpx[io - 1];
            ocean[(size_t)(jo - 1) * nxto + io - 1] =
                wm * wmy * atmos[(size_t)(jam - 1) * nxta + a - 1] + wp * wmy * atmos[(size_t)(jam - 1) * nxta + b - 1] +
                wm * wpy * atmos[(size_t)(jap - 1) * nxta + a - 1] + wp * wpy * atmos[(size_t)(jap - 1) * nxta + b - 1];
        }
    }
    free(iam); free(iap); free(wpx); free(wmx);
}

/* fsprim (radsubs.F90:97-114). */
static double fsprim(const cr_params* Pa, double fspco, double yrel) {
    const double yl = (Pa->nyp - 1) * Pa->dy;
    return fspco * 0.5 * sin(CR_PI * yrel / yl);
}

/* ytrel(j) (box.F90:144-150). */
static double ytrel(const cr_params* P, int j) {
    const double yl = (P->nyp - 1) * P->dy;
    return (-0.5 * yl + (j - 1) * P->dy) + 0.5 * P->dy;
}

/* compute_forcing (forcing.F90:18-99).  sstm (ocean T), astm and hmixam
 * (atmos T), eta1 and dtop (atmos p). */
void cr_compute_forcing(const cr_params* Po, const cr_params* Pa, const cr_cpl* g, const cr_aml_params* a,
                        const double* sstm, const double* astm, const double* hmixam, const double* eta1,
                        const double* dtop, double* fnetoc, double* fnetat) {
    const int nxto = Po->nxp - 1, nyto = Po->nyp - 1, nxt = Pa->nxp - 1, nyt = Pa->nyp - 1, nxp = Pa->nxp;
    const size_t no = (size_t)nxto * nyto;
    double* asto = calloc(no, sizeof(double));
    double* up = calloc(no, sizeof(double));
    double* sl = calloc(no, sizeof(double));
    cr_bilint(Pa, Po, g, astm, asto);
    for (size_t t = 0; t < no; t++) {
        up[t] = a->D0up * sstm[t];
        sl[t] = a->xlamda * (sstm[t] - asto[t]);
    }
    for (int j = 1; j <= nyt; j++) {
        const double fs = fsprim(Pa, a->fspco, ytrel(Pa, j));
        for (int i = 1; i <= nxt; i++) TT(fnetat, i, j) = -a->Dup0 * TT(astm, i, j) - fs;
    }
    for (int ja = g->ny1; ja <= g->ny1 + g->nyaooc - 1; ja++)
        for (int ia = g->nx1; ia <= g->nx1 + g->nxaooc - 1; ia++)
            TT(fnetat, ia, ja) = TT(fnetat, ia, ja) + (a->Dmdown - a->Dup0) * TT(astm, ia, ja);
    for (int jo = 1; jo <= nyto; jo++) {
        const double fsp = fsprim(Pa, a->fspco, ytrel(Po, jo));
        for (int io = 1; io <= nxto; io++) {
            const size_t t = (size_t)(jo - 1) * nxto + io - 1;
            fnetoc[t] = -fsp - a->Dmdown * asto[t] - up[t] - sl[t];
        }
    }
    const double ocfrac = Po->dx * Po->dy / (Pa->dx * Pa->dy);
    for (int jo = 1; jo <= nyto; jo++) {
        const int ja = g->ny1 + (jo - 1) / g->ndxr;
        for (int io = 1; io <= nxto; io++) {
            const int ia = g->nx1 + (io - 1) / g->ndxr;
            const size_t t = (size_t)(jo - 1) * nxto + io - 1;
            TT(fnetat, ia, ja) = TT(fnetat, ia, ja) + ocfrac * (up[t] + sl[t]);
        }
    }
    const double fmafac = a->Adown11, fmatop = a->Aup00 + a->Adown10, hmafac = -a->hmadmp - a->Aup00 - a->Adown10;
    for (int j = 1; j <= nyt; j++)
        for (int i = 1; i <= nxt; i++) {
            /* map_P_to_T (numerics.F90:52-61) */
            const double me = (fmafac * 0.25) * (PP(eta1, i, j) + PP(eta1, i + 1, j) + PP(eta1, i, j + 1) + PP(eta1, i + 1, j + 1));
            const double mt = (fmatop * 0.25) * (PP(dtop, i, j) + PP(dtop, i + 1, j) + PP(dtop, i, j + 1) + PP(dtop, i + 1, j + 1));
            TT(fnetat, i, j) = TT(fnetat, i, j) - me - mt + hmafac * (TT(hmixam, i, j) - g->hma);
        }
    free(asto); free(up); free(sl);
}

/* step_aml (amlsubs.F90:184-229) with compute_new_aml (:231-287) and
 * compute_atmos_entrainment (:289-326).  eta (nxp,nyp,nl-1) from pm, p1 the
 * current layer-1 pressure, taux/tauy the atmosphere's stress (uek, vek, wekt
 * from ekman.F90:73-80 with the atmosphere's hm), fnetat and xc1ast (T,
 * set_xc1ast amlsubs.F90:131-182).  ent1 (p) is written when compute_ent is set. */
void cr_aml_step(const cr_params* Pa, const cr_cpl* g, const cr_aml_params* a, const double* p1,
                 const double* taux, const double* tauy, const double* eta, const double* dtop,
                 const double* fnetat, const double* xc1ast, double tdta, int compute_ent, double* ast, double* astm,
                 double* hmixa, double* hmixam, double* ent1) {
    const int nxp = Pa->nxp, nyp = Pa->nyp, nxt = nxp - 1, nyt = nyp - 1, nl = Pa->nl;
    const size_t nt = (size_t)nxt * nyt;
    const double hm = g->hma;
    double* uek = calloc((size_t)nxp * nyt, sizeof(double));
    double* vek = calloc((size_t)nxt * nyp, sizeof(double));
    double* wekt = calloc(nt, sizeof(double));
    double* tm = calloc(nt, sizeof(double));
    double* dh = calloc(nt, sizeof(double));
    double* anew = calloc(nt, sizeof(double));
    double* hnew = calloc(nt, sizeof(double));
    double* dtc = calloc(nt, sizeof(double));
    double* xfa = calloc(nt, sizeof(double));
    cr_ekman_uvw(Pa, hm, taux, tauy, uek, vek, wekt);
    /* init_atmos_ml (amlsubs.F90:84-87): ast with no flux walls, hmixa with
     * flux walls holding hm */
    cr_oml_params fa = {hm, a->T1_rel, a->dT_12, a->rho_cp, a->at2d, a->at4d, 1.0, 0.0, 0.0, 0, 0};
    cr_oml_params fh = {hm, a->T1_rel, a->dT_12, a->rho_cp, a->ahmd, 0.0, 1.0, hm, hm, 1, 1};
    cr_adv_diff(Pa, &fa, p1, uek, vek, ast, astm, tm);
    cr_adv_diff(Pa, &fh, p1, uek, vek, hmixa, hmixam, dh);
    const double hdrcdt = a->hmadmp / a->rho_cp;
    const double hmainv = 1.0 / hm, rrcpat = 1.0 / a->rho_cp;
    for (size_t t = 0; t < nt; t++) {
        const double dT = a->T1_rel - astm[t];
        double dT_fix;
        if (dT >= 2.0 * hdrcdt * tdta) {
            const double diab = hdrcdt * (hmixam[t] - hm) / dT;
            double h = hmixam[t] + tdta * (dh[t] - diab);
            const double fix = fmax(a->hmamin - h, 0.0);
            h = h + fix;
            hnew[t] = h;
            dT_fix = fix * dT / hmixam[t];
        } else {
            hnew[t] = hm;
            dT_fix = 0.0;
        }
        const double r = tm[t] + rrcpat * fnetat[t] / hmixam[t] - hmainv * wekt[t] * astm[t];
        double an = astm[t] + tdta * r + dT_fix;
        dtc[t] = fmin(0.0, a->T1_rel - an);
        anew[t] = an + dtc[t];
    }
    if (compute_ent) {
        for (size_t t = 0; t < nt; t++)
            xfa[t] = a->xcexp * a->bface * (hmixam[t] - hm) + a->dface * (a->xcexp * astm[t] + (1.0 - a->xcexp) * xc1ast[t]) +
                     (a->xcexp / (tdta * a->dT_12)) * hmixa[t] * dtc[t];
        cr_map_T_to_P(Pa, xfa, ent1);
        for (int k = 1; k <= nl - 1; k++)
            for (size_t t = 0; t < (size_t)nxp * nyp; t++)
                ent1[t] = ent1[t] + a->aface[k - 1] * eta[(size_t)(k - 1) * nxp * nyp + t];
        for (size_t t = 0; t < (size_t)nxp * nyp; t++) ent1[t] = ent1[t] + a->bface * dtop[t];
    }
    for (size_t t = 0; t < nt; t++) {
        astm[t] = ast[t];
        ast[t] = anew[t];
        hmixam[t] = hmixa[t];
        hmixa[t] = hnew[t];
    }
    free(uek); free(vek); free(wekt); free(tm); free(dh); free(anew); free(hnew); free(dtc); free(xfa);
}
#undef TT
#undef PP

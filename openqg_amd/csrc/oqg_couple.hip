// oqg_couple.hip -- the ocean-atmosphere exchange of dynamic_step
// (driver.F90:36-87) on the device:
//   k_xforc_atm   one thread per atmosphere p point: the first and second
//                 derivatives of layer 1 of its previous pressure and its own
//                 stress (windstress.F90:129-145);
//   k_xforc_corner one thread per p point over the ocean: the bicubic corner
//                 data (atmos_geos_vel, windstress.F90:201-257);
//   k_xforc_ocn   one thread per ocean p point: the bicubic velocity
//                 (bicubic.F90:26-95), the ocean velocity subtracted when the
//                 stress depends on the velocity difference (:259-275), the
//                 stress (atmos_windstress :277-314) and its ocean and
//                 atmosphere copies (compute_windstress :167-199);
//   k_force_ocn   one thread per ocean T point: bilinear atmospheric
//                 temperature (numerics.F90:528-589) and the ocean's heat flux
//                 (forcing.F90:43-72);
//   k_force_atm   one thread per atmosphere T point: its heat flux
//                 (forcing.F90:52-97), the flux k_force_ocn left for each
//                 ocean cell under it gathered in the reference's loop order;
//   k_aml_step    one thread per atmosphere T point: step_aml's two adv_diff
//                 terms and compute_new_aml (amlsubs.F90:207-287);
//   k_aml_finish  one thread per atmosphere p point: the entrainment
//                 (amlsubs.F90:289-326) into the basin's resident forcing and
//                 the time levels shifted by copies.
// k_aml_step stages its tile's fields in LDS: each T point reads about forty
// neighbours across its two adv_diff terms.

#include "oqg_mixed_dev.h"

namespace oqg {

namespace {

constexpr double CPL_PI = 3.14159265358979324;

// dPdx_2bc (numerics.F90:470-492) at p point (i, j), 0-based, of field F.
template <class F>
__device__ __forceinline__ double ddx_2bc(const F& p, const Geo& g, double alpha, int i, int j) {
    const int nxp = g.nxp, nyp = g.nyp;
    if (j == 0 || j == nyp - 1) return 0.0;
    if (i >= 1 && i <= nxp - 2) return (p(i + 1, j) - p(i - 1, j)) / (2.0 * g.dx);
    if (!g.cyclic) {
        if (i == 0) return (p(1, j) - p(0, j)) / (g.dx * (0.5 * alpha + 1.0));
        return (p(nxp - 1, j) - p(nxp - 2, j)) / (g.dx * (0.5 * alpha + 1.0));
    }
    return (p(1, j) - p(nxp - 2, j)) / (2.0 * g.dx);
}

// dPdy_2bc (numerics.F90:494-513).
template <class F>
__device__ __forceinline__ double ddy_2bc(const F& p, const Geo& g, double alpha, int i, int j) {
    const int nxp = g.nxp, nyp = g.nyp;
    if (!g.cyclic && (i == 0 || i == nxp - 1)) return 0.0;
    if (j >= 1 && j <= nyp - 2) return (p(i, j + 1) - p(i, j - 1)) / (2.0 * g.dy);
    if (j == 0) return (p(i, 1) - p(i, 0)) / (g.dy * (0.5 * alpha + 1.0));
    return (p(i, nyp - 1) - p(i, nyp - 2)) / (g.dy * (0.5 * alpha + 1.0));
}

// dP2dx2_bc (numerics.F90:427-448).
template <class F>
__device__ __forceinline__ double dxx_2bc(const F& p, const Geo& g, double alpha, int i, int j) {
    const int nxp = g.nxp, nyp = g.nyp;
    if (j == 0 || j == nyp - 1) return 0.0;
    const double fac = 1.0 / (g.dx * g.dx);
    if (i >= 1 && i <= nxp - 2) return fac * (p(i - 1, j) + p(i + 1, j) - 2.0 * p(i, j));
    if (!g.cyclic) {
        if (i == 0) return fac * alpha / (0.5 * alpha + 1.0) * (p(1, j) - p(0, j));
        return -fac * alpha / (0.5 * alpha + 1.0) * (p(nxp - 1, j) - p(nxp - 2, j));
    }
    return fac * (p(nxp - 2, j) + p(1, j) - 2.0 * p(0, j));
}

// dP2dy2_bc (numerics.F90:450-468).
template <class F>
__device__ __forceinline__ double dyy_2bc(const F& p, const Geo& g, double alpha, int i, int j) {
    const int nxp = g.nxp, nyp = g.nyp;
    if (!g.cyclic && (i == 0 || i == nxp - 1)) return 0.0;
    const double fac = 1.0 / (g.dy * g.dy);
    if (j >= 1 && j <= nyp - 2) return fac * (p(i, j - 1) + p(i, j + 1) - 2.0 * p(i, j));
    if (j == 0) return fac * alpha / (0.5 * alpha + 1.0) * (p(i, 1) - p(i, 0));
    return -fac * alpha / (0.5 * alpha + 1.0) * (p(i, nyp - 1) - p(i, nyp - 2));
}

struct PF {   // a pitched p field
    const double* f;
    long long pitch;
    __device__ __forceinline__ double operator()(int i, int j) const { return f[(long long)j * pitch + i]; }
};

// bilint (numerics.F90:528-589) of the atmosphere T field `at` at ocean T point (io, jo), 0-based.
__device__ __forceinline__ double bilint_at(const CplArgs& a, const double* at, int io, int jo) {
    const Geo &ga = a.ga, &go = a.go;
    const int nxta = ga.nxp - 1, nyta = ga.nyp - 1;
    const double dxainv = 1.0 / ga.dx, dyainv = 1.0 / ga.dy;
    const double xta1 = (a.xa0 + 0.0 * ga.dx) + 0.5 * ga.dx, yta1 = (a.ya0 + 0.0 * ga.dy) + 0.5 * ga.dy;
    const double xto = (a.xo0 + io * go.dx) + 0.5 * go.dx;
    int m = (int)(1.0 + dxainv * (xto - xta1));
    int mp = m + 1;
    const double xam = (m >= 1) ? (a.xa0 + (m - 1) * ga.dx) + 0.5 * ga.dx : xta1 - ga.dx;
    const double wp = dxainv * (xto - xam), wm = 1.0 - wp;
    m = 1 + (m + nxta - 1) % nxta;
    mp = 1 + (mp + nxta - 1) % nxta;
    const double yto = (a.yo0 + jo * go.dy) + 0.5 * go.dy;
    int jam = (int)(1.0 + dyainv * (yto - yta1));
    int jap = jam + 1;
    jam = jam > 1 ? jam : 1;
    jap = jap < nyta ? jap : nyta;
    const double yam = (a.ya0 + (jam - 1) * ga.dy) + 0.5 * ga.dy;
    const double wpy = dyainv * (yto - yam), wmy = 1.0 - wpy;
    const long long p = ga.pitch;
    const double* r0 = at + (long long)(jam - 1) * p;
    const double* r1 = at + (long long)(jap - 1) * p;
    return wm * wmy * r0[m - 1] + wp * wmy * r0[mp - 1] + wm * wpy * r1[m - 1] + wp * wpy * r1[mp - 1];
}

// fsprim (radsubs.F90:97-114) at T row j (0-based) of box g, with the atmosphere's yl.
__device__ __forceinline__ double fsprim_row(const CplArgs& a, const Geo& g, int j) {
    const double yl = (g.nyp - 1) * g.dy, yla = (a.ga.nyp - 1) * a.ga.dy;
    const double yrel = (-0.5 * yl + j * g.dy) + 0.5 * g.dy;
    return a.fspco * 0.5 * sin(CPL_PI * yrel / yla);
}

}  // namespace

// The atmosphere's first derivatives and its own stress (windstress.F90:129-145).
__global__ __launch_bounds__(256) void k_xforc_atm(CplArgs a) {
    const Geo& g = a.ga;
    const int nxp = g.nxp, nyp = g.nyp;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nxp * nyp) return;
    const long long P = g.pitch, fs = g.fs;
    const int j = t / nxp, i = t - j * nxp;
    const long long o = (long long)j * P + i;
    const PF pm{a.pma, P};
    const double fa = g.fnot, al = a.bcco_a;
    const double cdhfaa = (a.cdat / fa) / a.hma;
    const double qpreaa = 1.0 / (fabs(cdhfaa) * sqrt(2.0));
    const double px = ddx_2bc(pm, g, al, i, j), py = ddy_2bc(pm, g, al, i, j);
    a.scr[o] = px;
    a.scr[fs + o] = py;
    a.scr[2 * fs + o] = dxx_2bc(pm, g, al, i, j);
    a.scr[3 * fs + o] = dyy_2bc(pm, g, al, i, j);
    const double u1 = -py / fa, v1 = px / fa;
    const double u2 = 4.0 * cdhfaa * cdhfaa * (u1 * u1 + v1 * v1);
    const double shr = qpreaa * sqrt(-1.0 + sqrt(1.0 + u2));
    const double cuohf = cdhfaa * shr;
    const double cdochi = a.cdat * shr / (1.0 + cuohf * cuohf);
    a.tauxa[o] = cdochi * (u1 - cuohf * v1);
    a.tauya[o] = cdochi * (v1 + cuohf * u1);
}

// The bicubic corner data on the (nxpoar, nypoar) p points over the ocean
// (atmos_geos_vel, windstress.F90:236-246): the mixed and third derivatives are
// x derivatives of dpdy and dP2dy2_bc taken here from the first pass's fields.
__global__ __launch_bounds__(256) void k_xforc_corner(CplArgs a) {
    const Geo& g = a.ga;
    const int nc = a.nxpoar * a.nypoar;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nc) return;
    const long long P = g.pitch, fs = g.fs;
    const PF dpdx{a.scr, P}, dpdy{a.scr + fs, P}, dxx{a.scr + 2 * fs, P}, dyy{a.scr + 3 * fs, P};
    const double fa = g.fnot, al = a.bcco_a, dxa = g.dx, dya = g.dy;
    const int jc = t / a.nxpoar, ic = t - jc * a.nxpoar;
    const int i = a.nx1 - 1 + ic, j = a.ny1 - 1 + jc;
    auto dxy_at = [&](int ii, int jj) { return ddx_2bc(dpdy, g, al, ii, jj); };   // dp2_dxy
    const double dxy = dxy_at(i, j);
    const double dxxy = ddx_2bc(dxy_at, g, al, i, j);                             // dp3_dxxy
    const double dxyy = ddx_2bc(dyy, g, al, i, j);                                // dp3_dxyy
    double* c = a.coarse + t;
    c[0 * nc] = -dpdy(i, j) / fa;
    c[1 * nc] = -(dxa / fa) * dxy;
    c[2 * nc] = -(dya / fa) * dyy(i, j);
    c[3 * nc] = -(dxa * dya / fa) * dxyy;
    c[4 * nc] = dpdx(i, j) / fa;
    c[5 * nc] = (dxa / fa) * dxx(i, j);
    c[6 * nc] = (dya / fa) * dxy;
    c[7 * nc] = (dxa * dya / fa) * dxxy;
}

__global__ __launch_bounds__(256) void k_xforc_ocn(CplArgs a) {
    const Geo& g = a.go;
    const int io = blockIdx.x * 64 + (threadIdx.x & 63), jo = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (io >= g.nxp || jo >= g.nyp) return;
    const int n = a.ndxr, nr = (n + 1) * (n + 1), nc = a.nxpoar * a.nypoar;
    // the cell whose block covers this point last (bc3's loop order)
    const int ic = min(io / n, a.nxtoar - 1), jc = min(jo / n, a.nytoar - 1);
    const int r = (io - ic * n) + (jo - jc * n) * (n + 1);
    double vel[2];
#pragma unroll
    for (int f = 0; f < 2; f++) {
        const double* base = a.coarse + 4 * f * nc;
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const int t = k >> 2, corner = k & 3;
            const int ip = corner & 1, jq = corner >> 1;
            const double d = base[t * nc + (jc + jq) * a.nxpoar + (ic + ip)];
            s += a.st[(long long)k * nr + r] * d;
        }
        vel[f] = s;
    }
    double u = vel[0], v = vel[1];
    if (a.tau_udiff) {   // compute_udiff (windstress.F90:259-275)
        const PF pm{a.pmo, g.pitch};
        const double ox = ddx_2bc(pm, g, a.bcco_o, io, jo), oy = ddy_2bc(pm, g, a.bcco_o, io, jo);
        u = u + oy / g.fnot;
        v = v - ox / g.fnot;
    }
    const double fa = a.ga.fnot;
    const double cdhfaa = a.tau_udiff ? (a.cdat / fa) * (1.0 / a.hma + a.raoro / a.hmo) : (a.cdat / fa) / a.hma;
    const double qpreaa = 1.0 / (fabs(cdhfaa) * sqrt(2.0));
    const double shr = sqrt(-1.0 + sqrt(1.0 + 4.0 * cdhfaa * cdhfaa * (u * u + v * v)));
    const double cuohf = (qpreaa * cdhfaa) * shr;
    const double cdchi = (qpreaa * a.cdat) * shr / (1.0 + cuohf * cuohf);
    const double tx = cdchi * (u - cuohf * v), ty = cdchi * (v + cuohf * u);
    const long long o = (long long)jo * g.pitch + io;
    a.tauxo[o] = a.raoro * tx;
    a.tauyo[o] = a.raoro * ty;
    if (a.tau_udiff && io % n == 0 && jo % n == 0) {
        const long long oa = (long long)(a.ny1 - 1 + jo / n) * a.ga.pitch + (a.nx1 - 1 + io / n);
        a.tauxa[oa] = tx;
        a.tauya[oa] = ty;
    }
}

__global__ __launch_bounds__(256) void k_force_ocn(CplArgs a) {
    const Geo& g = a.go;
    const int io = blockIdx.x * 64 + (threadIdx.x & 63), jo = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (io >= g.nxp - 1 || jo >= g.nyp - 1) return;
    const long long o = (long long)jo * g.pitch + io;
    const double asto = bilint_at(a, a.astm, io, jo);
    const double s = a.sstm[o];
    const double up = a.D0up * s, sl = a.xlamda * (s - asto);
    const double fsp = fsprim_row(a, g, jo);
    a.fnetoc[o] = -fsp - a.Dmdown * asto - up - sl;
    a.ocflux[o] = up + sl;
}

__global__ __launch_bounds__(256) void k_force_atm(CplArgs a) {
    const Geo &g = a.ga, &go = a.go;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nxt = g.nxp - 1, nyt = g.nyp - 1;
    const int ja = t / nxt, ia = t - ja * nxt;   // 0-based
    if (ja >= nyt) return;
    const long long P = g.pitch, o = (long long)ja * P + ia;
    const double am = a.astm[o];
    double f = -a.Dup0 * am - fsprim_row(a, g, ja);
    const int ia1 = ia - (a.nx1 - 1), ja1 = ja - (a.ny1 - 1);   // position in the cells over the ocean
    if (ia1 >= 0 && ia1 < a.nxaooc && ja1 >= 0 && ja1 < a.nyaooc) f = f + (a.Dmdown - a.Dup0) * am;
    // the ocean cells under this one, in the reference's loop order (jo outer, io inner)
    const double ocfrac = go.dx * go.dy / (g.dx * g.dy);
    const int n = a.ndxr, nxto = go.nxp - 1, nyto = go.nyp - 1;
    if (ia1 >= 0 && ja1 >= 0) {   // k_force_ocn wrote each cell's flux
        const int jo0 = ja1 * n, io0 = ia1 * n, jo1 = min(jo0 + n, nyto), io1 = min(io0 + n, nxto);
        constexpr int FR = 4, FC = 16;   // rows per batch of loads, columns of a row held
        if (io1 - io0 <= FC) {
            // FR rows loaded together, then added in the same order
            for (int jb = jo0; jb < jo1; jb += FR) {
                double v[FR][FC];
#pragma unroll
                for (int r = 0; r < FR; r++)
#pragma unroll
                    for (int u = 0; u < FC; u++) {
                        const bool ok = jb + r < jo1 && io0 + u < io1;
                        const long long q = ok ? (long long)(jb + r) * go.pitch + io0 + u : 0;
                        v[r][u] = a.ocflux[q];
                    }
#pragma unroll
                for (int r = 0; r < FR; r++)
#pragma unroll
                    for (int u = 0; u < FC; u++)
                        if (jb + r < jo1 && io0 + u < io1) f = f + ocfrac * v[r][u];
            }
        } else {
            for (int jo = jo0; jo < jo1; jo++) {
                const double* row = a.ocflux + (long long)jo * go.pitch;
                for (int io = io0; io < io1; io++) f = f + ocfrac * row[io];
            }
        }
    }
    // the thickness, topography and mixed-layer depth terms (forcing.F90:87-97)
    const double fmafac = a.Adown11, fmatop = a.Aup00 + a.Adown10, hmafac = -a.hmadmp - a.Aup00 - a.Adown10;
    auto eta1 = [&](int i, int j) {
        const long long q = (long long)j * P + i;
        return (a.pma[q] - a.pma[q + g.fs]) / a.gpa[0];
    };
    auto top = [&](int i, int j) { return a.dtop[(long long)j * P + i]; };
    const double me = (fmafac * 0.25) * (eta1(ia, ja) + eta1(ia + 1, ja) + eta1(ia, ja + 1) + eta1(ia + 1, ja + 1));
    const double mt = (fmatop * 0.25) * (top(ia, ja) + top(ia + 1, ja) + top(ia, ja + 1) + top(ia + 1, ja + 1));
    a.fnetat[o] = f - me - mt + hmafac * (a.hmixam[o] - a.hma);
}

// k_aml_step's tile: 64 x 4 T points per 256-thread block, its T fields with a
// halo of 2 (Del^4 of the previous level) and its p fields (layer-1 p, stress)
// on the tile's p points, staged in LDS.
constexpr int AML_TW = 64, AML_TH = 4, AML_H = 2, AML_LW = AML_TW + 2 * AML_H, AML_LH = AML_TH + 2 * AML_H;
constexpr int AML_PW = AML_TW + 1, AML_PH = AML_TH + 1;

struct TLds {   // a T field in the LDS tile, absolute 0-based indices (wrapped in a channel)
    const double* s;
    int i0, j0, nxt, cyclic;
    __device__ __forceinline__ double operator()(int i, int j) const {
        int lx = i - i0 + AML_H;
        if (cyclic) lx = (lx < 0) ? lx + nxt : (lx >= AML_LW ? lx - nxt : lx);
        return s[(j - j0 + AML_H) * AML_LW + lx];
    }
};

constexpr int AML_DW = AML_TW + 2, AML_DH = AML_TH + 2;   // the Del^2 tile: halo 1

struct DLds {   // Del^2 of a previous level in the LDS tile (halo 1), absolute indices
    const double* s;
    int i0, j0, nxt, cyclic;
    __device__ __forceinline__ double operator()(int i, int j) const {
        int lx = i - i0 + 1;
        if (cyclic) lx = (lx < 0) ? lx + nxt : (lx >= AML_DW ? lx - nxt : lx);
        return s[(j - j0 + 1) * AML_DW + lx];
    }
};

struct PLds {   // a p field in the LDS tile
    const double* s;
    int i0, j0;
    __device__ __forceinline__ double operator()(int i, int j) const { return s[(j - j0) * AML_PW + (i - i0)]; }
};

__global__ __launch_bounds__(256) void k_aml_step(CplArgs a) {
    const Geo& g = a.ga;
    const int nxt = g.nxp - 1, nyt = g.nyp - 1;
    const int i0 = blockIdx.x * AML_TW, j0 = blockIdx.y * AML_TH;
    const long long P = g.pitch;
    __shared__ double tl[4][AML_LH * AML_LW];   // ast, astm, hmixa, hmixam
    __shared__ double pl[3][AML_PH * AML_PW];   // p1, taux, tauy
    // every global load of the block is issued before any is used (from
    // always-valid offsets; the ones outside the basin are replaced by 0)
    constexpr int NT = (AML_LH * AML_LW + 255) / 256, NPL = (AML_PH * AML_PW + 255) / 256;
    const double* tf[4] = {a.ast, a.astm, a.hmixa, a.hmixam};
    const double* pf[3] = {a.pa1, a.tauxa, a.tauya};
    double vt[NT][4], vp[NPL][3];
    bool okt[NT], okp[NPL];
#pragma unroll
    for (int u = 0; u < NT; u++) {
        const int t = threadIdx.x + 256 * u;
        const int ly = t / AML_LW, lx = t - ly * AML_LW;
        int i = i0 - AML_H + lx;
        const int j = j0 - AML_H + ly;
        if (g.cyclic) i = (i % nxt + nxt) % nxt;
        okt[u] = t < AML_LH * AML_LW && i >= 0 && i < nxt && j >= 0 && j < nyt;
        const long long o = okt[u] ? (long long)j * P + i : 0;
#pragma unroll
        for (int f = 0; f < 4; f++) vt[u][f] = tf[f][o];
    }
#pragma unroll
    for (int u = 0; u < NPL; u++) {
        const int t = threadIdx.x + 256 * u;
        const int ly = t / AML_PW, lx = t - ly * AML_PW;
        const int i = i0 + lx, j = j0 + ly;
        okp[u] = t < AML_PH * AML_PW && i < g.nxp && j < g.nyp;
        const long long o = okp[u] ? (long long)j * P + i : 0;
#pragma unroll
        for (int f = 0; f < 3; f++) vp[u][f] = pf[f][o];
    }
    const int pi = i0 + (threadIdx.x & 63), pj = j0 + (threadIdx.x >> 6);
    const bool pout = pi < nxt && pj < nyt;
    const long long po = pout ? (long long)pj * P + pi : 0;
    const double fnet = a.fnetat[po], xc1 = a.xc1ast[po];
#pragma unroll
    for (int u = 0; u < NT; u++) {
        const int t = threadIdx.x + 256 * u;
        if (t < AML_LH * AML_LW)
#pragma unroll
            for (int f = 0; f < 4; f++) tl[f][t] = okt[u] ? vt[u][f] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < NPL; u++) {
        const int t = threadIdx.x + 256 * u;
        if (t < AML_PH * AML_PW)
#pragma unroll
            for (int f = 0; f < 3; f++) pl[f][t] = okp[u] ? vp[u][f] : 0.0;
    }
    __syncthreads();
    const double hm = a.hma;
    const TLds ast{tl[0], i0, j0, nxt, g.cyclic}, astm{tl[1], i0, j0, nxt, g.cyclic};
    const TLds hmx{tl[2], i0, j0, nxt, g.cyclic}, hmxm{tl[3], i0, j0, nxt, g.cyclic};
    // init_atmos_ml (amlsubs.F90:84-87): ast with no flux walls, hmixa with flux walls holding hm
    const AdvField fa{nullptr, nullptr, a.at2d, a.at4d, 1.0, 0.0, 0.0, 0, 0};
    const AdvField fh{nullptr, nullptr, a.ahmd, 0.0, 1.0, hm, hm, 1, 1};
    // Del^2 of both previous levels on the tile and a halo of 1 (del2_T, numerics.F90:107-147)
    __shared__ double dl[2][AML_DH * AML_DW];
    const double fx = 1.0 / (g.dx * g.dx);
    for (int t = threadIdx.x; t < AML_DH * AML_DW; t += blockDim.x) {
        const int ly = t / AML_DW, lx = t - ly * AML_DW;
        int i = i0 - 1 + lx;
        const int j = j0 - 1 + ly;
        if (g.cyclic) i = (i % nxt + nxt) % nxt;
        const bool in = i >= 0 && i < nxt && j >= 0 && j < nyt;
        dl[0][t] = in ? del2T_at(astm, i, j, nxt, nyt, g.cyclic, fx, fa.fluxn, fa.nval, fa.fluxs, fa.sval) : 0.0;
        dl[1][t] = in ? del2T_at(hmxm, i, j, nxt, nyt, g.cyclic, fx, fh.fluxn, fh.nval, fh.fluxs, fh.sval) : 0.0;
    }
    __syncthreads();
    const int i = pi, j = pj;
    if (!pout) return;
    const long long o = po;
    const PLds p1{pl[0], i0, j0}, tx{pl[1], i0, j0}, ty{pl[2], i0, j0};
    const DLds d2a{dl[0], i0, j0, nxt, g.cyclic}, d2h{dl[1], i0, j0, nxt, g.cyclic};
    const double tm = adv_diff_d2(g, fa, ast, d2a, p1, tx, ty, hm, a.tau_sign_a, i, j);
    const double dh = adv_diff_d2(g, fh, hmx, d2h, p1, tx, ty, hm, a.tau_sign_a, i, j);
    const double wekt = wekt_acc(g, tx, ty, i, j);
    // compute_new_aml (amlsubs.F90:231-287)
    const double tdta = a.tdta, astm0 = astm(i, j), hmixam = hmxm(i, j);
    const double dT = a.T1_rel - astm0;
    const double hdrcdt = a.hmadmp / a.rho_cp;
    double hnew, dT_fix;
    if (dT >= 2.0 * hdrcdt * tdta) {
        const double diab = hdrcdt * (hmixam - hm) / dT;
        double h = hmixam + tdta * (dh - diab);
        const double fix = fmax(a.hmamin - h, 0.0);
        h = h + fix;
        hnew = h;
        dT_fix = fix * dT / hmixam;
    } else {
        hnew = hm;
        dT_fix = 0.0;
    }
    const double hmainv = 1.0 / hm, rrcpat = 1.0 / a.rho_cp;
    const double r = tm + rrcpat * fnet / hmixam - hmainv * wekt * astm0;
    const double an = astm0 + tdta * r + dT_fix;
    const double dtc = fmin(0.0, a.T1_rel - an);
    a.anew[o] = an + dtc;
    a.hnew[o] = hnew;
    // compute_atmos_entrainment's T-point terms (amlsubs.F90:303-309)
    a.xfa[o] = a.xcexp * a.bface * (hmixam - hm) + a.dface * (a.xcexp * astm0 + (1.0 - a.xcexp) * xc1) +
               (a.xcexp / (tdta * a.dT_12)) * hmx(i, j) * dtc;
}

__global__ __launch_bounds__(256) void k_aml_finish(CplArgs a) {
    const Geo& g = a.ga;
    const int nxp = g.nxp, nyp = g.nyp;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = t / nxp, i = t - j * nxp;
    if (j >= nyp) return;
    const long long P = g.pitch, o = (long long)j * P + i;
    if (a.compute_ent) {
        auto X = [&](int ii, int jj) { return a.xfa[(long long)jj * P + ii]; };
        double e = map_T_to_P_at(X, i, j, nxp, nyp, g.cyclic);
        for (int k = 0; k < g.nl - 1; k++) {
            const double eta = (a.pma[o + k * g.fs] - a.pma[o + (k + 1) * g.fs]) / a.gpa[k];
            e = e + a.aface[k] * eta;
        }
        a.enta[o] = e + a.bface * a.dtop[o];
    }
    if (i < nxp - 1 && j < nyp - 1) {
        const double as = a.ast[o], hs = a.hmixa[o];
        a.astm[o] = as;
        a.ast[o] = a.anew[o];
        a.hmixam[o] = hs;
        a.hmixa[o] = a.hnew[o];
    }
}

// suppress_atmos's mixed-layer part (driver.F90:150-157)
__global__ __launch_bounds__(256) void k_aml_suppress(CplArgs a) {
    const int nxt = a.ga.nxp - 1, nyt = a.ga.nyp - 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = t / nxt, i = t - j * nxt;
    if (j >= nyt) return;
    const long long o = (long long)j * a.ga.pitch + i;
    a.ast[o] = 0.5 * (a.ast[o] + a.astm[o]);
    a.hmixa[o] = 0.5 * (a.hmixa[o] + a.hmixam[o]);
}

static dim3 tgrid(int nx, int ny) { return dim3((nx + 63) / 64, (ny + 3) / 4); }

void launch_xforc(const CplArgs& a, hipStream_t st) {
    const int na = a.ga.nxp * a.ga.nyp, nc = a.nxpoar * a.nypoar;
    hipLaunchKernelGGL(k_xforc_atm, dim3((na + 255) / 256), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_xforc_corner, dim3((nc + 255) / 256), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_xforc_ocn, tgrid(a.go.nxp, a.go.nyp), dim3(256), 0, st, a);
}

void launch_forcing(const CplArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_force_ocn, tgrid(a.go.nxp - 1, a.go.nyp - 1), dim3(256), 0, st, a);
    const int n = (a.ga.nxp - 1) * (a.ga.nyp - 1);
    hipLaunchKernelGGL(k_force_atm, dim3((n + 255) / 256), dim3(256), 0, st, a);
}

void launch_aml_step(const CplArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_aml_step, dim3((a.ga.nxp - 1 + AML_TW - 1) / AML_TW, (a.ga.nyp - 1 + AML_TH - 1) / AML_TH),
                       dim3(256), 0, st, a);
    const int n = a.ga.nxp * a.ga.nyp;
    hipLaunchKernelGGL(k_aml_finish, dim3((n + 255) / 256), dim3(256), 0, st, a);
}

void launch_aml_suppress(const CplArgs& a, hipStream_t st) {
    const int n = (a.ga.nxp - 1) * (a.ga.nyp - 1);
    hipLaunchKernelGGL(k_aml_suppress, dim3((n + 255) / 256), dim3(256), 0, st, a);
}

}  // namespace oqg

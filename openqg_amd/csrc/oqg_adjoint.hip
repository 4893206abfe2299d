// oqg_adjoint.hip -- reverse-mode (adjoint) kernels of one step_qg (box or channel).
//
// The step maps S_n = (q_n, q_{n-1}, p_n, p_{n-1}) to S_{n+1}.  Given the
// adjoint of S_{n+1} these kernels produce the adjoint of S_n (and accumulate
// the adjoint of the Ekman pumping field).  Every operator is the exact
// transpose of the forward discretisation, written out by hand:
//   k_adj_supp  suppress_comp_mode (driver.F90:174-189)
//   k_adj_vort  vort_step (vorticity.F90:14-103): Arakawa Jacobian transposes
//               (J_1(h)^T mu = -J(mu, h) for mu vanishing on the ring; the p
//               transpose through the discrete antisymmetry J(q,p) = -J(p,q)),
//               Del^2/Del^4/Del^6 transposes of del2_P_bc with its mixed
//               boundary condition, forcing
//   k_adj_add   sums of fields (state combination for the initial-state map)
// The inversion's adjoint runs through the transform kernels in their adjoint
// modes (oqg_transform.hip): k_fwd_adj forms the modal inputs from the
// adjoint of p (ocqbdy's transpose, vorsubs.F90:58-114, and the mode->layer
// mix's, modes.F90:268-284, with the homogeneous correction's inner products,
// homog.F90:331-366); the box Helmholtz solve (DST-I, tridiagonal in y,
// DST-I) is symmetric, so the forward k_ysolve is its adjoint; k_inv_adj
// applies the constraint solve's transpose and the layer->mode mix's
// (modes.F90:248-266).
// Layout and parity follow the forward state: the adjoint buffers lq[2],
// lp[2] mirror qbuf[2], pbuf[2].

#include "oqg_device.h"

namespace oqg {

// ------------------------------------------------------------------ k_adj_supp

// q <- (q + qm)/2, p <- (p + pm)/2 transposed: lqm += lq/2, lq <- lq/2;
// in a channel also cs <- (cs + csp)/2, cn <- (cn + cnp)/2.
__global__ void k_adj_supp(AdjSuppArgs a) {
    const long long n = (long long)a.g.nl * a.g.fs;
    const long long t0 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (long long t = t0; t < n; t += (long long)gridDim.x * blockDim.x) {
        const double hq = 0.5 * a.lq[t], hp = 0.5 * a.lp[t];
        a.lqm[t] += hq;
        a.lq[t] = hq;
        a.lpm[t] += hp;
        a.lp[t] = hp;
    }
    if (a.lc && t0 < LC_LEN) {
        const double h = 0.5 * a.lc[t0];
        a.lcm[t0] += h;
        a.lc[t0] = h;
    }
}

// ------------------------------------------------------------------ k_adj_add, k_copy_segs

__global__ void k_adj_add(double* dst, const double* src, long long n) {
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        dst[t] += src[t];
}

// One flat grid over all segments: segment s owns blocks [a.b0[s], a.b0[s+1]),
// each thread two 16-byte vectors, loaded together (every segment starts on a
// 16-byte boundary and holds an even number of doubles).
__global__ __launch_bounds__(256) void k_copy_segs(CopySegs a) {
    int s = 0;
    while (s + 1 < a.nseg && (int)blockIdx.x >= a.b0[s + 1]) s++;
    double* slot = a.slots ? a.slots[*a.base + a.t] : nullptr;
    const double2* src = (const double2*)a.src[s];
    double2* dst = (double2*)(slot ? slot + a.doff[s] : a.dst[s]);
    const long long n2 = a.n[s] >> 1;
    const long long t = ((long long)blockIdx.x - a.b0[s]) * 2LL * blockDim.x + threadIdx.x;
    const long long t1 = t + blockDim.x;
    if (t < n2) {
        const double2 v0 = src[t];
        const double2 v1 = src[t1 < n2 ? t1 : t];
        dst[t] = v0;
        if (t1 < n2) dst[t1] = v1;
    }
}

// ------------------------------------------------------------------ k_ke_seed

// dPdy_2bc / dPdx_2bc of a box field at (i, j) (numerics.F90:470-513).
__device__ __forceinline__ double ke_dpdx(const double* P, long long pitch, int i, int j, const Geo& g, double ca) {
    if (j == 0 || j == g.nyp - 1) return 0.0;
    const long long o = (long long)j * pitch + i;
    if (i >= 1 && i <= g.nxp - 2) return (P[o + 1] - P[o - 1]) / (2.0 * g.dx);
    return i == 0 ? (P[o + 1] - P[o]) / (g.dx * ca) : (P[o] - P[o - 1]) / (g.dx * ca);
}

__device__ __forceinline__ double ke_dpdy(const double* P, long long pitch, int i, int j, const Geo& g, double ca) {
    if (i == 0 || i == g.nxp - 1) return 0.0;
    const long long o = (long long)j * pitch + i;
    if (j >= 1 && j <= g.nyp - 2) return (P[o + pitch] - P[o - pitch]) / (2.0 * g.dy);
    return j == 0 ? (P[o + pitch] - P[o]) / (g.dy * ca) : (P[o] - P[o - pitch]) / (g.dy * ca);
}

// dJ/dp_1 for J = 0.5 rho H_1 sum w (u^2 + v^2), in gather form: point (i, j)
// collects, from every point whose derivative stencil reads it, the stencil's
// coefficient times gu = rho H_1 w u (resp. gv = rho H_1 w v) there; then
// grad = -(d/dy)^T gu / f0 + (d/dx)^T gv / f0.
__global__ __launch_bounds__(256) void k_ke_seed(KeSeedArgs a) {
    const Geo& g = a.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= g.nxp) return;
    const long long pitch = g.pitch;
    const double ca = 0.5 * a.bcco + 1.0;
    auto wgt = [&](int x, int y) {
        const double wx = (x == 0 || x == g.nxp - 1) ? 0.5 : 1.0, wy = (y == 0 || y == g.nyp - 1) ? 0.5 : 1.0;
        return wx * wy * g.dx * g.dy;
    };
    auto GU = [&](int x, int y) { return a.rhoh * wgt(x, y) * (-ke_dpdy(a.p, pitch, x, y, g, ca) / g.fnot); };
    auto GV = [&](int x, int y) { return a.rhoh * wgt(x, y) * (ke_dpdx(a.p, pitch, x, y, g, ca) / g.fnot); };
    const double cx = g.dx * ca, cy = g.dy * ca;
    double tx = 0.0, ty = 0.0;
    if (j != 0 && j != g.nyp - 1) {   // (d/dx)^T: rows 0 and nyp-1 of d/dx are zero
        if (i >= 2) tx += GV(i - 1, j) / (2.0 * g.dx);
        if (i <= g.nxp - 3) tx -= GV(i + 1, j) / (2.0 * g.dx);
        if (i == 1) tx += GV(0, j) / cx;
        if (i == 0) tx -= GV(0, j) / cx;
        if (i == g.nxp - 1) tx += GV(g.nxp - 1, j) / cx;
        if (i == g.nxp - 2) tx -= GV(g.nxp - 1, j) / cx;
    }
    if (i != 0 && i != g.nxp - 1) {   // (d/dy)^T: columns 0 and nxp-1 of d/dy are zero
        if (j >= 2) ty += GU(i, j - 1) / (2.0 * g.dy);
        if (j <= g.nyp - 3) ty -= GU(i, j + 1) / (2.0 * g.dy);
        if (j == 1) ty += GU(i, 0) / cy;
        if (j == 0) ty -= GU(i, 0) / cy;
        if (j == g.nyp - 1) ty += GU(i, g.nyp - 1) / cy;
        if (j == g.nyp - 2) ty -= GU(i, g.nyp - 1) / cy;
    }
    a.lp[(long long)j * pitch + i] = -ty / g.fnot + tx / g.fnot;
}

// Channel: on the torus of the nxp-1 distinct columns, where int_P_dA's x
// weights are 1 and dPdx_2bc is the periodic centred difference (0 on rows 0
// and nyp-1); column nxp-1 of the result is 0.
__global__ __launch_bounds__(256) void k_ke_seed_cyc(KeSeedArgs a) {
    const Geo& g = a.g;
    const int L = g.nxp - 1, N = g.nyp - 1;
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= g.nxp) return;
    const long long pitch = g.pitch;
    if (i == L) {
        a.lp[(long long)j * pitch + i] = 0.0;
        return;
    }
    const double ca = 0.5 * a.bcco + 1.0, cy = g.dy * ca;
    auto wrap = [&](int x) { return x < 0 ? x + L : (x >= L ? x - L : x); };
    auto P = [&](int x, int y) { return a.p[(long long)y * pitch + wrap(x)]; };
    auto wgt = [&](int y) { return ((y == 0 || y == N) ? 0.5 : 1.0) * g.dx * g.dy; };
    auto dpdx = [&](int x, int y) { return (y == 0 || y == N) ? 0.0 : (P(x + 1, y) - P(x - 1, y)) / (2.0 * g.dx); };
    auto dpdy = [&](int x, int y) {
        if (y >= 1 && y <= N - 1) return (P(x, y + 1) - P(x, y - 1)) / (2.0 * g.dy);
        return y == 0 ? (P(x, 1) - P(x, 0)) / cy : (P(x, N) - P(x, N - 1)) / cy;
    };
    auto GU = [&](int x, int y) { return a.rhoh * wgt(y) * (-dpdy(x, y) / g.fnot); };
    auto GV = [&](int x, int y) { return a.rhoh * wgt(y) * (dpdx(x, y) / g.fnot); };
    double tx = 0.0, ty = 0.0;
    if (j != 0 && j != N) tx = (GV(i - 1, j) - GV(i + 1, j)) / (2.0 * g.dx);
    if (j >= 2) ty += GU(i, j - 1) / (2.0 * g.dy);
    if (j <= N - 2) ty -= GU(i, j + 1) / (2.0 * g.dy);
    if (j == 1) ty += GU(i, 0) / cy;
    if (j == 0) ty -= GU(i, 0) / cy;
    if (j == N) ty += GU(i, N) / cy;
    if (j == N - 1) ty -= GU(i, N) / cy;
    a.lp[(long long)j * pitch + i] = -ty / g.fnot + tx / g.fnot;
}

void launch_ke_seed(const KeSeedArgs& a, hipStream_t st) {
    if (a.g.cyclic) hipLaunchKernelGGL(k_ke_seed_cyc, dim3((a.g.nxp + 255) / 256, a.g.nyp), dim3(256), 0, st, a);
    else hipLaunchKernelGGL(k_ke_seed, dim3((a.g.nxp + 255) / 256, a.g.nyp), dim3(256), 0, st, a);
}

// ------------------------------------------------------------------ k_adj_vort

// Forward (box interior points x): q_{n+1} = q_{n-1} + tdt F with
//   F = jfac J(q_n, p_n) + c2 L^2 p_{n-1} - c4 L^3 p_{n-1} - [k=kt] cdrag L p_{n-1}
//       + [k=0] (cent1 ent + cwek wekp) - [k=1] cent2 ent.
// With mu = tdt lq_{n+1} on the interior (0 elsewhere):
//   lq_n += jfac J_1^T(mu; p_n),  lp_n += -jfac J_1^T(mu; q_n)   (J(q,p) = -J(p,q))
//   lp_{n-1} = c2 (L^T)^2 mu - c4 (L^T)^3 mu - [k=kt] cdrag L^T mu
//   lwekp += cwek mu_0,  lq_{n-1} = lq_{n+1} on the interior, 0 on the ring.
// The kernel is latency-bound: what counts is each thread's chain of loads
// and LDS passes while every block stays resident.  k_adj_vort (box): 64 x 8
// points of one layer per 512-thread block, one point per thread, 30 KB of
// LDS, 76 VGPRs, three blocks per CU.  Reverse sweeps of the 257^2 gradient:
// 190.1 ms with these, 191.0 ms with 64 x 4 tiles of 256 threads, 195.2 ms with
// 64 x 2 of 128; two points per thread (64 x 8 tiles of 256 threads) were
// slower still.  k_adj_vort_cyc (channel): 64 x 4 points per 256 threads.
constexpr int ADJ_TX = 64, ADJ_TY = 4;
constexpr int ADJB_TY = 8, ADJB_NTH = 512;

// (L^T v)(a) of del2_P_bc (box, numerics.F90:427-468): L = Dxx + Dyy, Dxx
// zero on the rows 0 and nyp-1, Dyy zero on the columns 0 and nxp-1, so the
// transpose at a gathers v(a) and v(a -/+ 1) with the coefficients Dxx has at
// the columns a_i, a_i -/+ 1 towards a (when a's row is not a boundary row):
// -2/dx^2 and 1/dx^2 inside, the mixed condition's cxl, cxr on columns 0 and
// nxp-1; likewise in y.  Written as selects, so every point takes one path
// (a per-point choice between the interior Laplacian and a general gather
// cost 2.2 us per adjoint step at 257^2).  Neighbours outside the basin (a
// tile's zero-filled halo) get coefficient 0.
__device__ __forceinline__ double lapT_sel(const double* s, int W, int r, int c, int ai, int aj, const Geo& g) {
    const bool xin = ai >= 1 && ai <= g.nxp - 2, yin = aj >= 1 && aj <= g.nyp - 2;
    const bool xwin = ai - 1 >= 1 && ai - 1 <= g.nxp - 2, xein = ai + 1 >= 1 && ai + 1 <= g.nxp - 2;
    const bool ysin = aj - 1 >= 1 && aj - 1 <= g.nyp - 2, ynin = aj + 1 >= 1 && aj + 1 <= g.nyp - 2;
    const double sx = xin ? -2.0 * g.fx : (ai == 0 ? -g.cxl : g.cxr);
    const double wx = (ai - 1 == 0) ? g.cxl : (xwin ? g.fx : 0.0);
    const double ex = (ai + 1 == g.nxp - 1) ? -g.cxr : (xein ? g.fx : 0.0);
    const double sy = yin ? -2.0 * g.fy : (aj == 0 ? -g.cyl : g.cyr);
    const double ws = (aj - 1 == 0) ? g.cyl : (ysin ? g.fy : 0.0);
    const double en = (aj + 1 == g.nyp - 1) ? -g.cyr : (ynin ? g.fy : 0.0);
    const double* x = s + r * W + c;
    const double tx = sx * x[0] + wx * x[-1] + ex * x[1];
    const double ty = sy * x[0] + ws * x[-W] + en * x[W];
    return (yin ? tx : 0.0) + (xin ? ty : 0.0);
}

// The Jacobian transposes: mu vanishes on the ring, so every nonzero term of
// the gather comes from an interior point x, whose Arakawa coefficients are
// the full ones, and the antisymmetry of the stencil (the coefficient of f(a)
// in J(f, h)(x) is minus that of f(x) in J(f, h)(a), both built from the h
// values the two 3x3 neighbourhoods share) gives J_1(h)^T mu = -J(mu, h) at
// every point, the ring included, whatever h holds outside the domain.

// Arakawa sum J(f, h) before the factor jfac (J++ + J+x + Jx+, each times 3),
// f and h LDS tiles of widths Wf, Wh with the point at (rf, cf) and (rh, ch).
__device__ __forceinline__ double arakawa_lds(const double* f, int Wf, int rf, int cf, const double* h, int Wh,
                                              int rh, int ch) {
    const double* F = f + rf * Wf + cf;
    const double* H = h + rh * Wh + ch;
    const double fE = F[1], fW = F[-1], fN = F[Wf], fS = F[-Wf];
    const double fNE = F[Wf + 1], fNW = F[Wf - 1], fSE = F[-Wf + 1], fSW = F[-Wf - 1];
    const double hE = H[1], hW = H[-1], hN = H[Wh], hS = H[-Wh];
    const double hNE = H[Wh + 1], hNW = H[Wh - 1], hSE = H[-Wh + 1], hSW = H[-Wh - 1];
    const double jpp = (fE - fW) * (hN - hS) - (fN - fS) * (hE - hW);
    const double jpx = fE * (hNE - hSE) - fW * (hNW - hSW) - fN * (hNE - hNW) + fS * (hSE - hSW);
    const double jxp = fNE * (hN - hE) - fSW * (hW - hS) - fNW * (hN - hW) + fSE * (hE - hS);
    return jpp + jpx + jxp;
}

__global__ __launch_bounds__(ADJB_NTH) void k_adj_vort(AdjVortArgs a) {
    constexpr int TX = ADJ_TX, TY = ADJB_TY, NTH = ADJB_NTH, TR = NTH / 64;
    constexpr int W3 = TX + 6, H3 = TY + 6, W2 = TX + 4, H2 = TY + 4, W1 = TX + 2, H1 = TY + 2;
    constexpr int NMU = (H3 * W3 + NTH - 1) / NTH, NQP = (H1 * W1 + NTH - 1) / NTH, NOUT = TY / TR;
    __shared__ double s_mu[H3 * W3];
    __shared__ double s_n1[H2 * W2];
    __shared__ double s_n2[H1 * W1];
    __shared__ double s_q[H1 * W1];
    __shared__ double s_p[H1 * W1];
    const Geo& g = a.g;
    const int tid = threadIdx.x, tx = tid & 63, ty = tid >> 6;
    int bx, by, k;
    xcd_block(bx, by, k);   // neighbouring tiles, which share halo lines, on one XCD
    const int i0 = bx * TX, j0 = by * TY;
    const long long lo = (long long)k * g.fs;
    const double* ln = a.lq_np1 + lo;
    const double* qn = a.q + lo;
    const double* pn = a.p + lo;
    if (a.slots) {
        const double* sb = a.slots[*a.base + a.t];
        qn = sb + a.qoff + lo;
        pn = sb + a.poff + lo;
    }
    // every global load of the block is issued before any is used: the tile
    // of mu = tdt lq_{n+1} (interior only) with halo 3, q_n and p_n with halo
    // 1 (from clamped, always valid offsets), and the outputs updated in place
    double vmu[NMU], vq[NQP], vp[NQP];
#pragma unroll
    for (int u = 0; u < NMU; u++) {
        const int idx = tid + NTH * u, r = idx / W3, c = idx - r * W3;
        const int gi = i0 - 3 + c, gj = j0 - 3 + r;
        const bool ok = idx < H3 * W3 && gi >= 1 && gi <= g.nxp - 2 && gj >= 1 && gj <= g.nyp - 2;
        const double v = ln[ok ? (long long)gj * g.pitch + gi : 0];
        vmu[u] = ok ? a.tdt * v : 0.0;
    }
#pragma unroll
    for (int u = 0; u < NQP; u++) {
        const int idx = tid + NTH * u, r = idx / W1, c = idx - r * W1;
        const int gi = i0 - 1 + c, gj = j0 - 1 + r;
        const bool ok = idx < H1 * W1 && gi >= 0 && gi < g.nxp && gj >= 0 && gj < g.nyp;
        const long long o = ok ? (long long)gj * g.pitch + gi : 0;
        const double q = qn[o], p = pn[o];
        vq[u] = ok ? q : 0.0;
        vp[u] = ok ? p : 0.0;
    }
    const int gi = i0 + tx;
    double olq[NOUT], olp[NOUT], olw[NOUT];
#pragma unroll
    for (int u = 0; u < NOUT; u++) {
        const int gj = j0 + ty + TR * u;
        const bool ok = gi < g.nxp && gj < g.nyp;
        const long long o = ok ? (long long)gj * g.pitch + gi : 0;
        olq[u] = a.lq_n[lo + o];
        olp[u] = a.lp_n[lo + o];
        olw[u] = a.lwek[o];
    }
#pragma unroll
    for (int u = 0; u < NMU; u++) {
        const int idx = tid + NTH * u;
        if (idx < H3 * W3) s_mu[idx] = vmu[u];
    }
#pragma unroll
    for (int u = 0; u < NQP; u++) {
        const int idx = tid + NTH * u;
        if (idx < H1 * W1) {
            s_q[idx] = vq[u];
            s_p[idx] = vp[u];
        }
    }
    __syncthreads();
    // nu1 = L^T mu on the tile with halo 2
    for (int idx = tid; idx < H2 * W2; idx += NTH) {
        const int r = idx / W2, c = idx - r * W2;
        const int gi2 = i0 - 2 + c, gj2 = j0 - 2 + r;
        const bool in = gi2 >= 0 && gi2 < g.nxp && gj2 >= 0 && gj2 < g.nyp;
        double v = 0.0;
        if (in) v = lapT_sel(s_mu, W3, r + 1, c + 1, gi2, gj2, g);
        s_n1[idx] = v;
    }
    __syncthreads();
    // nu2 = L^T nu1 on the tile with halo 1
    for (int idx = tid; idx < H1 * W1; idx += NTH) {
        const int r = idx / W1, c = idx - r * W1;
        const int gi1 = i0 - 1 + c, gj1 = j0 - 1 + r;
        const bool in = gi1 >= 0 && gi1 < g.nxp && gj1 >= 0 && gj1 < g.nyp;
        double v = 0.0;
        if (in) v = lapT_sel(s_n1, W2, r + 1, c + 1, gi1, gj1, g);
        s_n2[idx] = v;
    }
    __syncthreads();
    const double c2 = a.ph->c2[k], c4 = a.ph->c4[k];
#pragma unroll
    for (int u = 0; u < NOUT; u++) {
        const int r = ty + TR * u, gj = j0 + r;
        if (gi >= g.nxp || gj >= g.nyp) continue;
        const long long o = lo + (long long)gj * g.pitch + gi;
        const double nu3 = lapT_sel(s_n2, W1, r + 1, tx + 1, gi, gj, g);
        // J_1(h)^T mu = -J(mu, h) at every point, the ring included (mu is 0
        // there; see the note above arakawa_lds)
        const double jq = -arakawa_lds(s_mu, W3, r + 3, tx + 3, s_p, W1, r + 1, tx + 1);
        const double jp = -arakawa_lds(s_mu, W3, r + 3, tx + 3, s_q, W1, r + 1, tx + 1);
        const double nu2 = s_n2[(r + 1) * W1 + tx + 1];
        const double nu1 = s_n1[(r + 2) * W2 + tx + 2];
        double lpm = c2 * nu2 - c4 * nu3;
        if (k == g.kt) lpm -= a.ph->cdrag * nu1;
        a.lq_n[o] = olq[u] + g.jfac * jq;
        a.lp_n[o] = olp[u] - g.jfac * jp;
        a.lp_nm1[o] = lpm;
        const double mu = s_mu[(r + 3) * W3 + tx + 3];
        if (k == 0) a.lwek[(long long)gj * g.pitch + gi] = olw[u] + a.ph->cwek * mu;
        // lq_{n-1}: the interior of lq_{n+1} (unchanged), 0 on the ring
        if (gi == 0 || gi == g.nxp - 1 || gj == 0 || gj == g.nyp - 1) a.lq_np1[o] = 0.0;
    }
}

// ------------------------------------------------------------------ k_adj_vort_cyc

// Channel.  The state is periodic with period nxp-1 in x (column nxp-1 repeats
// column 0), so the transposes are taken on the torus of the nxp-1 distinct
// columns: every adjoint field holds the sensitivity of a periodic point in
// its column 0..nxp-2, and 0 in column nxp-1.  mu = tdt lq_{n+1} vanishes on
// the zonal boundary rows, which ocqbdy writes; on the torus the Jacobian
// transposes are J_1(h)^T mu = -J(mu, h) as in a box.
//
// The momentum constraints add (vorticity.F90:57-62,84-101,
// constraint.F90:182-215, numerics.F90:374-386), with ls = tdt lcs_{n+1},
// ln = tdt lcn_{n+1} and sums over the torus:
//   ajis = (1/12) sum_i (q(i,0) + 2 q(i,1)) (p(i+1,1) - p(i-1,1))
//   ajin = -(1/12) sum_i (q(i,N) + 2 q(i,N-1)) (p(i+1,N-1) - p(i-1,N-1))
//   ap3s, ap5s, ap3n, ap5n and the drag integrals: functions of the row sums
//   of pm only (the periodic x difference sums to zero along a row), whose
//   gradients vs3, vs5, vsb, vn3, vn5, vnb the host computes from the y part
//   of del2_P_bc
//   txis, txin: rhss_1 += tau_sign f0/h_1 txis, rhsn_1 -= tau_sign f0/h_1 txin.

__device__ __forceinline__ int torus_col(int gi, int L) { return ((gi % L) + L) % L; }

// (L^T v)(a) of the channel's del2_P_bc (numerics.F90:427-468): the periodic
// x difference on rows 1..nyp-2, symmetric on the torus, and the y part with
// its mixed condition on every column (as for a box).
__device__ __forceinline__ double lapT_cyc(const double* s, int W, int r, int c, int aj, const Geo& g) {
    const bool yin = aj >= 1 && aj <= g.nyp - 2;
    const bool ysin = aj - 1 >= 1 && aj - 1 <= g.nyp - 2, ynin = aj + 1 >= 1 && aj + 1 <= g.nyp - 2;
    const double sy = yin ? -2.0 * g.fy : (aj == 0 ? -g.cyl : g.cyr);
    const double ws = (aj - 1 == 0) ? g.cyl : (ysin ? g.fy : 0.0);
    const double en = (aj + 1 == g.nyp - 1) ? -g.cyr : (ynin ? g.fy : 0.0);
    const double* x = s + r * W + c;
    const double tx = g.fx * (x[-1] + x[1]) - 2.0 * g.fx * x[0];
    const double ty = sy * x[0] + ws * x[-W] + en * x[W];
    return (yin ? tx : 0.0) + ty;
}

__global__ __launch_bounds__(256) void k_adj_vort_cyc(AdjVortArgs a) {
    constexpr int TX = ADJ_TX, TY = ADJ_TY;
    constexpr int W3 = TX + 6, H3 = TY + 6, W2 = TX + 4, H2 = TY + 4, W1 = TX + 2, H1 = TY + 2;
    constexpr int NMU = (H3 * W3 + 255) / 256, NQP = (H1 * W1 + 255) / 256, NOUT = TY / 4;
    __shared__ double s_mu[H3 * W3];
    __shared__ double s_n1[H2 * W2];
    __shared__ double s_n2[H1 * W1];
    __shared__ double s_q[H1 * W1];
    __shared__ double s_p[H1 * W1];
    const Geo& g = a.g;
    const int L = g.nxp - 1, N = g.nyp - 1;
    const int tid = threadIdx.x, tx = tid & 63, ty = tid >> 6;
    const int k = blockIdx.z, i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
    const long long lo = (long long)k * g.fs;
    const double* ln = a.lq_np1 + lo;
    const double* qn = a.q + lo;
    const double* pn = a.p + lo;
    if (a.slots) {
        const double* sb = a.slots[*a.base + a.t];
        qn = sb + a.qoff + lo;
        pn = sb + a.poff + lo;
    }
    double vmu[NMU], vq[NQP], vp[NQP];
#pragma unroll
    for (int u = 0; u < NMU; u++) {
        const int idx = tid + 256 * u, r = idx / W3, c = idx - r * W3;
        const int gi = torus_col(i0 - 3 + c, L), gj = j0 - 3 + r;
        const bool ok = idx < H3 * W3 && gj >= 1 && gj <= N - 1;
        const double v = ln[ok ? (long long)gj * g.pitch + gi : 0];
        vmu[u] = ok ? a.tdt * v : 0.0;
    }
#pragma unroll
    for (int u = 0; u < NQP; u++) {
        const int idx = tid + 256 * u, r = idx / W1, c = idx - r * W1;
        const int gi = torus_col(i0 - 1 + c, L), gj = j0 - 1 + r;
        const bool ok = idx < H1 * W1 && gj >= 0 && gj <= N;
        const long long o = ok ? (long long)gj * g.pitch + gi : 0;
        const double q = qn[o], p = pn[o];
        vq[u] = ok ? q : 0.0;
        vp[u] = ok ? p : 0.0;
    }
    const int gi = i0 + tx;
    double olq[NOUT], olp[NOUT], olw[NOUT];
#pragma unroll
    for (int u = 0; u < NOUT; u++) {
        const int gj = j0 + ty + 4 * u;
        const bool ok = gi < L && gj <= N;
        const long long o = ok ? (long long)gj * g.pitch + gi : 0;
        olq[u] = a.lq_n[lo + o];
        olp[u] = a.lp_n[lo + o];
        olw[u] = a.lwek[o];
    }
    const int kb = g.nl - 1;
    const double ls = a.tdt * a.lc[k], lnn = a.tdt * a.lc[MAXL + k];
    const double lsb = a.tdt * a.lc[kb], lnb = a.tdt * a.lc[MAXL + kb];
#pragma unroll
    for (int u = 0; u < NMU; u++) {
        const int idx = tid + 256 * u;
        if (idx < H3 * W3) s_mu[idx] = vmu[u];
    }
#pragma unroll
    for (int u = 0; u < NQP; u++) {
        const int idx = tid + 256 * u;
        if (idx < H1 * W1) {
            s_q[idx] = vq[u];
            s_p[idx] = vp[u];
        }
    }
    __syncthreads();
    for (int idx = tid; idx < H2 * W2; idx += 256) {
        const int r = idx / W2, c = idx - r * W2;
        const int gj2 = j0 - 2 + r;
        s_n1[idx] = (gj2 >= 0 && gj2 <= N) ? lapT_cyc(s_mu, W3, r + 1, c + 1, gj2, g) : 0.0;
    }
    __syncthreads();
    for (int idx = tid; idx < H1 * W1; idx += 256) {
        const int r = idx / W1, c = idx - r * W1;
        const int gj1 = j0 - 1 + r;
        s_n2[idx] = (gj1 >= 0 && gj1 <= N) ? lapT_cyc(s_n1, W2, r + 1, c + 1, gj1, g) : 0.0;
    }
    __syncthreads();
    const Phys& ph = *a.ph;
    const double c2 = ph.c2[k], c4 = ph.c4[k], rxy = g.dx / g.dy;
    if (tid == 0 && blockIdx.x == 0 && blockIdx.y == 0 && k == 0) {
        a.ltxi[0] += ph.tau_sign * ph.fohfac[0] * ls;
        a.ltxi[1] -= ph.tau_sign * ph.fohfac[0] * lnn;
    }
#pragma unroll
    for (int u = 0; u < NOUT; u++) {
        const int r = ty + 4 * u, gj = j0 + r;
        if (gi >= L || gj > N) continue;
        const long long o = lo + (long long)gj * g.pitch + gi;
        const double nu3 = lapT_cyc(s_n2, W1, r + 1, tx + 1, gj, g);
        const double jq = -arakawa_lds(s_mu, W3, r + 3, tx + 3, s_p, W1, r + 1, tx + 1);
        const double jp = -arakawa_lds(s_mu, W3, r + 3, tx + 3, s_q, W1, r + 1, tx + 1);
        const double nu2 = s_n2[(r + 1) * W1 + tx + 1];
        const double nu1 = s_n1[(r + 2) * W2 + tx + 2];
        double lpm = c2 * nu2 - c4 * nu3;
        if (k == g.kt) lpm -= ph.cdrag * nu1;
        double dq = g.jfac * jq, dp = -g.jfac * jp;
        // the momentum constraints' boundary integrals
        const double* P = s_p + (r + 1) * W1 + tx + 1;   // the point in the q, p tiles
        const double* Q = s_q + (r + 1) * W1 + tx + 1;
        if (gj == 0) dq += (ls / 12.0) * (P[W1 + 1] - P[W1 - 1]);
        if (gj == 1) {
            dq += (2.0 * ls / 12.0) * (P[1] - P[-1]);
            dp += (ls / 12.0) * ((Q[-W1 - 1] + 2.0 * Q[-1]) - (Q[-W1 + 1] + 2.0 * Q[1]));
        }
        if (gj == N) dq -= (lnn / 12.0) * (P[-W1 + 1] - P[-W1 - 1]);
        if (gj == N - 1) {
            dq -= (2.0 * lnn / 12.0) * (P[1] - P[-1]);
            dp -= (lnn / 12.0) * ((Q[W1 - 1] + 2.0 * Q[-1]) - (Q[W1 + 1] + 2.0 * Q[1]));
        }
        if (gj <= 3) {
            lpm += rxy * ls * (ph.ah4[k] * a.vs5[gj] - ph.ah2[k] * a.vs3[gj]);
            if (k == g.kt) lpm += ph.fohfac[kb] * ph.cd_bdr * rxy * lsb * a.vsb[gj];
        }
        if (N - gj <= 3) {
            const int d = N - gj;
            lpm += rxy * lnn * (ph.ah2[k] * a.vn3[d] - ph.ah4[k] * a.vn5[d]);
            if (k == g.kt) lpm -= ph.fohfac[kb] * ph.cd_bdr * rxy * lnb * a.vnb[d];
        }
        a.lq_n[o] = olq[u] + dq;
        a.lp_n[o] = olp[u] + dp;
        a.lp_nm1[o] = lpm;
        const double mu = s_mu[(r + 3) * W3 + tx + 3];
        if (k == 0) a.lwek[(long long)gj * g.pitch + gi] = olw[u] + ph.cwek * mu;
        if (gj == 0 || gj == N) a.lq_np1[o] = 0.0;
    }
}

// ------------------------------------------------------------------ launchers

static int blocks_for(long long n) {
    const long long b = (n + 255) / 256;
    return (int)(b < 2048 ? (b < 1 ? 1 : b) : 2048);
}

void launch_adj_supp(const AdjSuppArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_supp, dim3(blocks_for((long long)a.g.nl * a.g.fs)), dim3(256), 0, st, a);
}

void launch_adj_add(double* dst, const double* src, long long n, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_add, dim3(blocks_for(n)), dim3(256), 0, st, dst, src, n);
}

// Zero the boundary ring of nl fields (a channel: rows 0 and nyp-1).
__global__ void k_zero_ring(double* f, Geo g) {
    const int nx = g.nxp, ny = g.nyp, per = g.cyclic ? 2 * nx : 2 * nx + 2 * (ny - 2);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per * g.nl) return;
    const int k = t / per, r = t - k * per;
    int i, j;
    if (r < nx) { i = r; j = 0; }
    else if (r < 2 * nx) { i = r - nx; j = ny - 1; }
    else if (r < 2 * nx + ny - 2) { i = 0; j = 1 + (r - 2 * nx); }
    else { i = nx - 1; j = 1 + (r - 2 * nx - (ny - 2)); }
    f[(long long)k * g.fs + (long long)j * g.pitch + i] = 0.0;
}

void launch_zero_ring(double* f, const Geo& g, hipStream_t st) {
    const int n = (g.cyclic ? 2 * g.nxp : 2 * g.nxp + 2 * (g.nyp - 2)) * g.nl;
    hipLaunchKernelGGL(k_zero_ring, dim3((n + 255) / 256), dim3(256), 0, st, f, g);
}

// Channel: an adjoint given on all nxp columns as the sensitivity to each
// stored value becomes that of the periodic field: column nxp-1 (a copy of
// column 0) is added into column 0 and cleared.
__global__ void k_fold_dup(double* f, Geo g, int nl) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nl * g.nyp) return;
    double* row = f + (long long)(t / g.nyp) * g.fs + (long long)(t % g.nyp) * g.pitch;
    row[0] += row[g.nxp - 1];
    row[g.nxp - 1] = 0.0;
}

void launch_fold_dup(double* f, const Geo& g, int nl, hipStream_t st) {
    const int n = nl * g.nyp;
    hipLaunchKernelGGL(k_fold_dup, dim3((n + 255) / 256), dim3(256), 0, st, f, g, nl);
}

__global__ void k_set_int(int* p, int v) {
    if (threadIdx.x == 0) *p = v;
}

void launch_set_int(int* p, int v, hipStream_t st) { hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, st, p, v); }

void launch_copy_segs(const CopySegs& a0, hipStream_t st) {
    CopySegs a = a0;
    int nb = 0;
    for (int s = 0; s < a.nseg; s++) {
        a.b0[s] = nb;
        nb += (int)((a.n[s] / 2 + 511) / 512);   // 512 vectors per block
    }
    a.b0[a.nseg] = nb;
    hipLaunchKernelGGL(k_copy_segs, dim3(nb), dim3(256), 0, st, a);
}

void launch_adj_vort(const AdjVortArgs& a, hipStream_t st) {
    if (a.g.cyclic) {
        const dim3 grid((a.g.nxp - 1 + ADJ_TX - 1) / ADJ_TX, (a.g.nyp + ADJ_TY - 1) / ADJ_TY, a.g.nl);
        hipLaunchKernelGGL(k_adj_vort_cyc, grid, dim3(256), 0, st, a);
        return;
    }
    const dim3 grid((a.g.nxp + ADJ_TX - 1) / ADJ_TX, (a.g.nyp + ADJB_TY - 1) / ADJB_TY, a.g.nl);
    hipLaunchKernelGGL(k_adj_vort, grid, dim3(ADJB_NTH), 0, st, a);
}

}  // namespace oqg

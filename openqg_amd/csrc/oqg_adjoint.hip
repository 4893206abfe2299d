// oqg_adjoint.hip -- reverse-mode (adjoint) kernels of one box step_qg.
//
// The step maps S_n = (q_n, q_{n-1}, p_n, p_{n-1}) to S_{n+1}.  Given the
// adjoint of S_{n+1} these kernels produce the adjoint of S_n (and accumulate
// the adjoint of the Ekman pumping field).  Every operator is the exact
// transpose of the forward discretisation, written out by hand:
//   k_adj_supp  suppress_comp_mode (driver.F90:174-189)
//   k_adj_bdy   ocqbdy (vorsubs.F90:58-114): boundary-ring PV -> p
//   k_adj_inv1  mode->layer mix (modes.F90:268-284) and the homogeneous
//               correction's inner products (homog.F90:331-366)
//   k_adj_inv2  the (nl-1)^2 constraint solve, transposed
//   k_adj_mix   layer->mode mix (modes.F90:248-266) of the Helmholtz adjoint
//   k_adj_vort  vort_step (vorticity.F90:14-103): Arakawa Jacobian transposes
//               in gather form (the p transpose through the discrete
//               antisymmetry J(q,p) = -J(p,q)), Del^2/Del^4/Del^6 transposes
//               of del2_P_bc with its mixed boundary condition, forcing
//   k_adj_add   sums of fields (state combination for the initial-state map)
// The Helmholtz solve of the box (DST-I, tridiagonal in y, DST-I) is
// symmetric, so its adjoint is the forward raw solve.
// Layout and parity follow the forward state: the adjoint buffers lq[2],
// lp[2] mirror qbuf[2], pbuf[2].

#include "oqg_device.h"

namespace oqg {

// ------------------------------------------------------------------ k_adj_supp

// q <- (q + qm)/2, p <- (p + pm)/2 transposed: lqm += lq/2, lq <- lq/2.
__global__ void k_adj_supp(AdjSuppArgs a) {
    const long long n = (long long)a.g.nl * a.g.fs;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const double hq = 0.5 * a.lq[t], hp = 0.5 * a.lp[t];
        a.lqm[t] += hq;
        a.lq[t] = hq;
        a.lpm[t] += hp;
        a.lp[t] = hp;
    }
}

// ------------------------------------------------------------------ k_adj_bdy

// Forward (box), at a ring point x with inner point x_in:
//   q_k(x) = qbfac (p_k(x_in) - p_k(x)) - r f0 sum_l A_kl p_l(x) + r beta y + ...
// with r = 2 at the corners (the row and column terms both applied) and 1
// otherwise.  Transpose gathered at every point a of the two outer bands:
//   lp_l(a) += sum_k lq_k(a) (-qbfac d_kl - r f0 A_kl)        (a on the ring)
//   lp_k(a) += qbfac lq_k(x)   for each ring point x whose inner point is a.
__device__ __forceinline__ bool ring_inner(const Geo& g, int xi, int xj, int& ii, int& ij) {
    const int nx = g.nxp, ny = g.nyp;
    const bool ex = (xi == 0 || xi == nx - 1), ey = (xj == 0 || xj == ny - 1);
    if (!ex && !ey) return false;
    if (ex && ey) {   // corner: x difference along its row
        ii = (xi == 0) ? 1 : nx - 2;
        ij = xj;
    } else if (ey) {
        ii = xi;
        ij = (xj == 0) ? 1 : ny - 2;
    } else {
        ii = (xi == 0) ? 1 : nx - 2;
        ij = xj;
    }
    return true;
}

__global__ void k_adj_bdy(AdjBdyArgs a) {
    const Geo& g = a.g;
    const int nx = g.nxp, ny = g.nyp, nl = g.nl;
    // points of the two outer bands: rows 0,1,ny-2,ny-1 and columns 0,1,nx-2,nx-1
    const long long npts = (long long)nx * ny;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < npts; t += (long long)gridDim.x * blockDim.x) {
        const int aj = (int)(t / nx), ai = (int)(t - (long long)aj * nx);
        const bool band = aj <= 1 || aj >= ny - 2 || ai <= 1 || ai >= nx - 2;
        if (!band) continue;
        const long long oa = (long long)aj * g.pitch + ai;
        double acc[MAXL];
#pragma unroll
        for (int k = 0; k < MAXL; k++) acc[k] = 0.0;
        int ii, ij;
        if (ring_inner(g, ai, aj, ii, ij)) {
            const double r = ((ai == 0 || ai == nx - 1) && (aj == 0 || aj == ny - 1)) ? 2.0 : 1.0;
            for (int k = 0; k < nl; k++) {
                const double lqk = a.lq[k * g.fs + oa];
                for (int l = 0; l < nl; l++) {
                    double c = -r * g.fnot * a.mx.amat[k * MAXL + l];
                    if (l == k) c -= a.ph.qbfac;
                    acc[l] += lqk * c;
                }
            }
        }
        // ring neighbours whose inner point is a
        const int cand[4][2] = {{ai - 1, aj}, {ai + 1, aj}, {ai, aj - 1}, {ai, aj + 1}};
        for (int c = 0; c < 4; c++) {
            const int xi = cand[c][0], xj = cand[c][1];
            if (xi < 0 || xi >= nx || xj < 0 || xj >= ny) continue;
            if (ring_inner(g, xi, xj, ii, ij) && ii == ai && ij == aj) {
                const long long ox = (long long)xj * g.pitch + xi;
                for (int k = 0; k < nl; k++) acc[k] += a.ph.qbfac * a.lq[k * g.fs + ox];
            }
        }
        for (int k = 0; k < nl; k++) a.lp[k * g.fs + oa] += acc[k];
    }
}

// ------------------------------------------------------------------ k_adj_inv1

// lv_m = sum_k ctm2l(m,k) lp_k at every point (the adjoint of the modal field
// inhomog_m + homcor_m), written to w[m]; per-block partial sums of
// hom_m . lv_m (m >= 1) for the adjoint of the correction coefficients.
__global__ __launch_bounds__(256) void k_adj_inv1(AdjInv1Args a) {
    const Geo& g = a.g;
    const int nl = g.nl;
    __shared__ double red[MAXL * 16];
    double part[MAXL];
#pragma unroll
    for (int m = 0; m < MAXL; m++) part[m] = 0.0;
    const long long npts = (long long)g.nxp * g.nyp;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < npts; t += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(t / g.nxp), i = (int)(t - (long long)j * g.nxp);
        const long long o = (long long)j * g.pitch + i;
        double lp[MAXL];
        for (int k = 0; k < nl; k++) lp[k] = a.lp[k * g.fs + o];
        for (int m = 0; m < nl; m++) {
            double v = 0.0;
            for (int k = 0; k < nl; k++) v += a.mx.ctm2l[m * MAXL + k] * lp[k];
            a.w[m * g.fs + o] = v;
            if (m >= 1) part[m] += a.hom[(long long)(m - 1) * g.fs + o] * v;
        }
    }
    block_sum<MAXL>(part, red);
    if (threadIdx.x == 0)
        for (int m = 0; m < nl; m++) a.partial[(long long)blockIdx.x * MAXL + m] = part[m];
}

// ------------------------------------------------------------------ k_adj_inv2

// One block.  la_{m-1} = sum of the partials of mode m; the forward solves
// cdhoc alpha = dpi - (ik_{l+1} - ik_l), ik_k = sum_m ctm2l(m,k) I_m, I_m the
// area integral of inhomog_m (interior weight dx dy).  Transposed:
// lr = cdhoc^-T la, lI_m = -sum_l lr_l (ctm2l(m,l+1) - ctm2l(m,l)),
// s_m = lI_m dx dy (the constant added to the interior of lv_m).
__global__ __launch_bounds__(256) void k_adj_inv2(AdjInv2Args a) {
    const Geo& g = a.g;
    const int nl = g.nl;
    __shared__ double red[MAXL * 16];
    double v[MAXL];
#pragma unroll
    for (int m = 0; m < MAXL; m++) v[m] = 0.0;
    for (int b = threadIdx.x; b < a.nblocks; b += blockDim.x)
        for (int m = 0; m < nl; m++) v[m] += a.partial[(long long)b * MAXL + m];
    block_sum<MAXL>(v, red);
    if (threadIdx.x == 0) {
        double lr[MAXL];
        for (int l = 0; l < nl - 1; l++) {
            double s = 0.0;
            for (int j = 0; j < nl - 1; j++) s += a.cinvT[l * MAXL + j] * v[j + 1];
            lr[l] = s;
        }
        for (int m = 0; m < nl; m++) {
            double s = 0.0;
            for (int l = 0; l < nl - 1; l++)
                s -= lr[l] * (a.mx.ctm2l[m * MAXL + l + 1] - a.mx.ctm2l[m * MAXL + l]);
            a.s[m] = (s * g.dx) * g.dy;
        }
    }
}

// ------------------------------------------------------------------ k_adj_mix

// lq_k += f0 sum_m ctl2m(k,m) (R_m + s_m G_m) on the interior, where R_m is the
// Helmholtz solve of lv_m and G_m that of the interior indicator.
__global__ void k_adj_mix(AdjMixArgs a) {
    const Geo& g = a.g;
    const int nl = g.nl, nxi = g.nxp - 2, nyi = g.nyp - 2;
    const long long npts = (long long)nxi * nyi;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < npts; t += (long long)gridDim.x * blockDim.x) {
        const int j = 1 + (int)(t / nxi), i = 1 + (int)(t - (long long)(j - 1) * nxi);
        const long long o = (long long)j * g.pitch + i;
        double r[MAXL];
        for (int m = 0; m < nl; m++) r[m] = a.R[m * g.fs + o] + a.s[m] * a.G[m * g.fs + o];
        for (int k = 0; k < nl; k++) {
            double s = 0.0;
            for (int m = 0; m < nl; m++) s += a.mx.ctl2m[k * MAXL + m] * r[m];
            a.lq[k * g.fs + o] += g.fnot * s;
        }
    }
}

// ------------------------------------------------------------------ k_adj_add

__global__ void k_adj_add(double* dst, const double* src, long long n) {
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        dst[t] += src[t];
}

// ------------------------------------------------------------------ k_adj_vort

// Coefficient that del2_P_bc at x (box) gives to f at x + o, for the five
// offsets o of its stencil; zero elsewhere.  Mirrors del2_lds case by case.
__device__ __forceinline__ double lap_coef(int xi, int xj, int oi, int oj, const Geo& g) {
    if (xi < 0 || xi >= g.nxp || xj < 0 || xj >= g.nyp) return 0.0;
    const bool xin = xi >= 1 && xi <= g.nxp - 2;
    const bool yin = xj >= 1 && xj <= g.nyp - 2;
    double c = 0.0;
    if (oj == 0 && yin) {   // dP2dx2_bc, zero on rows 1 and nyp
        if (xin) c += (oi == 0) ? -2.0 * g.fx : ((oi == 1 || oi == -1) ? g.fx : 0.0);
        else if (xi == 0) c += (oi == 0) ? -g.cxl : ((oi == 1) ? g.cxl : 0.0);
        else c += (oi == 0) ? g.cxr : ((oi == -1) ? -g.cxr : 0.0);
    }
    if (oi == 0 && xin) {   // dP2dy2_bc, zero on columns 1 and nxp
        if (yin) c += (oj == 0) ? -2.0 * g.fy : ((oj == 1 || oj == -1) ? g.fy : 0.0);
        else if (xj == 0) c += (oj == 0) ? -g.cyl : ((oj == 1) ? g.cyl : 0.0);
        else c += (oj == 0) ? g.cyr : ((oj == -1) ? -g.cyr : 0.0);
    }
    return c;
}

// (L^T v)(a) = sum over the stencil neighbours x of a of v(x) * coef(x -> a).
// s is an LDS tile of width W with a at (r, c) and global (ai, aj).
__device__ __forceinline__ double lapT_lds(const double* s, int W, int r, int c, int ai, int aj, const Geo& g) {
    double acc = 0.0;
    const int off[5][2] = {{0, 0}, {-1, 0}, {1, 0}, {0, -1}, {0, 1}};   // x = a + off
#pragma unroll
    for (int n = 0; n < 5; n++) {
        const int di = off[n][0], dj = off[n][1];
        acc += s[(r + dj) * W + c + di] * lap_coef(ai + di, aj + dj, -di, -dj, g);
    }
    return acc;
}

// Derivative of the Arakawa sum at x (before the factor jfac) with respect to
// its first argument f at x + o, o != 0, for second argument h around x (the
// nine values hv[(dj+1)*3 + di+1]) and x's flags.
__device__ __forceinline__ double jac_coef(int oi, int oj, const double (&h)[9], bool xin, bool yin) {
    const double hN = h[7], hS = h[1], hE = h[5], hW = h[3];
    const double hNE = h[8], hNW = h[6], hSE = h[2], hSW = h[0];
    const double dhy = yin ? hN - hS : 0.0, dhx = xin ? hE - hW : 0.0;
    const double dhyE = yin ? hNE - hSE : 0.0, dhyW = yin ? hNW - hSW : 0.0;
    const double dhxN = xin ? hNE - hNW : 0.0, dhxS = xin ? hSE - hSW : 0.0;
    const double b = (xin && yin) ? 1.0 : 0.0;
    if (oj == 0) {
        if (oi == 1) return xin ? dhy + dhyE : 0.0;
        return xin ? -dhy - dhyW : 0.0;
    }
    if (oi == 0) {
        if (oj == 1) return yin ? -dhx - dhxN : 0.0;
        return yin ? dhx + dhxS : 0.0;
    }
    if (oi == 1 && oj == 1) return b * (hN - hE);
    if (oi == -1 && oj == 1) return b * (hW - hN);
    if (oi == 1 && oj == -1) return b * (hE - hS);
    return b * (hS - hW);
}

// Transpose of f -> J(f, h) gathered at a: sum over x = a - o of
// mu(x) * jac_coef(o; h around x).  mu is an LDS tile of width Wm with a at
// (rm, cm); h an LDS tile of width Wh with a at (rh, ch).
__device__ __forceinline__ double jacT_lds(const double* mu, int Wm, int rm, int cm, const double* h, int Wh, int rh,
                                           int ch, int ai, int aj, const Geo& g) {
    double acc = 0.0;
#pragma unroll
    for (int oj = -1; oj <= 1; oj++)
#pragma unroll
        for (int oi = -1; oi <= 1; oi++) {
            if (oi == 0 && oj == 0) continue;
            const int xi = ai - oi, xj = aj - oj;
            if (xi < 0 || xi >= g.nxp || xj < 0 || xj >= g.nyp) continue;
            const double m = mu[(rm - oj) * Wm + cm - oi];
            double hv[9];
#pragma unroll
            for (int dj = -1; dj <= 1; dj++)
#pragma unroll
                for (int di = -1; di <= 1; di++) hv[(dj + 1) * 3 + di + 1] = h[(rh - oj + dj) * Wh + ch - oi + di];
            const bool xin = xi >= 1 && xi <= g.nxp - 2, yin = xj >= 1 && xj <= g.nyp - 2;
            acc += m * jac_coef(oi, oj, hv, xin, yin);
        }
    return acc;
}

// Forward (box interior points x): q_{n+1} = q_{n-1} + tdt F with
//   F = jfac J(q_n, p_n) + c2 L^2 p_{n-1} - c4 L^3 p_{n-1} - [k=kt] cdrag L p_{n-1}
//       + [k=0] (cent1 ent + cwek wekp) - [k=1] cent2 ent.
// With mu = tdt lq_{n+1} on the interior (0 elsewhere):
//   lq_n += jfac J_1^T(mu; p_n),  lp_n += -jfac J_1^T(mu; q_n)   (J(q,p) = -J(p,q))
//   lp_{n-1} = c2 (L^T)^2 mu - c4 (L^T)^3 mu - [k=kt] cdrag L^T mu
//   lwekp += cwek mu_0,  lq_{n-1} = lq_{n+1} on the interior, 0 on the ring.
__global__ __launch_bounds__(256) void k_adj_vort(AdjVortArgs a) {
    constexpr int TX = VORT_TX, TY = VORT_TY;
    constexpr int W3 = TX + 6, H3 = TY + 6, W2 = TX + 4, H2 = TY + 4, W1 = TX + 2, H1 = TY + 2;
    __shared__ double s_mu[H3 * W3];
    __shared__ double s_n1[H2 * W2];
    __shared__ double s_n2[H1 * W1];
    __shared__ double s_q[H2 * W2];
    __shared__ double s_p[H2 * W2];
    const Geo& g = a.g;
    const int tid = threadIdx.x, tx = tid & 63, ty = tid >> 6;
    const int k = blockIdx.z;
    const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
    const long long lo = (long long)k * g.fs;
    const double* ln = a.lq_np1 + lo;
    auto mu_at = [&](int gi, int gj) -> double {   // tdt * lq_{n+1} on the interior only
        if (gi < 1 || gi > g.nxp - 2 || gj < 1 || gj > g.nyp - 2) return 0.0;
        return a.tdt * ln[(long long)gj * g.pitch + gi];
    };
    for (int idx = tid; idx < H3 * W3; idx += 256) {
        const int r = idx / W3, c = idx - r * W3;
        s_mu[idx] = mu_at(i0 - 3 + c, j0 - 3 + r);
    }
    for (int idx = tid; idx < H2 * W2; idx += 256) {
        const int r = idx / W2, c = idx - r * W2;
        const int gi = i0 - 2 + c, gj = j0 - 2 + r;
        const bool in = gi >= 0 && gi < g.nxp && gj >= 0 && gj < g.nyp;
        const long long o = lo + (long long)gj * g.pitch + gi;
        s_q[idx] = in ? a.q[o] : 0.0;
        s_p[idx] = in ? a.p[o] : 0.0;
    }
    __syncthreads();
    // nu1 = L^T mu on the tile with halo 2
    for (int idx = tid; idx < H2 * W2; idx += 256) {
        const int r = idx / W2, c = idx - r * W2;
        const int gi = i0 - 2 + c, gj = j0 - 2 + r;
        const bool in = gi >= 0 && gi < g.nxp && gj >= 0 && gj < g.nyp;
        s_n1[idx] = in ? lapT_lds(s_mu, W3, r + 1, c + 1, gi, gj, g) : 0.0;
    }
    __syncthreads();
    // nu2 = L^T nu1 on the tile with halo 1
    for (int idx = tid; idx < H1 * W1; idx += 256) {
        const int r = idx / W1, c = idx - r * W1;
        const int gi = i0 - 1 + c, gj = j0 - 1 + r;
        const bool in = gi >= 0 && gi < g.nxp && gj >= 0 && gj < g.nyp;
        s_n2[idx] = in ? lapT_lds(s_n1, W2, r + 1, c + 1, gi, gj, g) : 0.0;
    }
    __syncthreads();
    const double c2 = a.ph.c2[k], c4 = a.ph.c4[k];
    const int gi = i0 + tx;
    for (int u = 0; u < TY / 4; u++) {
        const int r = ty + 4 * u, gj = j0 + r;
        if (gi >= g.nxp || gj >= g.nyp) continue;
        const long long o = lo + (long long)gj * g.pitch + gi;
        const double nu3 = lapT_lds(s_n2, W1, r + 1, tx + 1, gi, gj, g);
        const double nu2 = s_n2[(r + 1) * W1 + tx + 1];
        const double nu1 = s_n1[(r + 2) * W2 + tx + 2];
        double lpm = c2 * nu2 - c4 * nu3;
        if (k == g.kt) lpm -= a.ph.cdrag * nu1;
        const double jq = jacT_lds(s_mu, W3, r + 3, tx + 3, s_p, W2, r + 2, tx + 2, gi, gj, g);
        const double jp = jacT_lds(s_mu, W3, r + 3, tx + 3, s_q, W2, r + 2, tx + 2, gi, gj, g);
        a.lq_n[o] += g.jfac * jq;
        a.lp_n[o] += -(g.jfac * jp);
        a.lp_nm1[o] = lpm;
        const double mu = s_mu[(r + 3) * W3 + tx + 3];
        if (k == 0) a.lwek[(long long)gj * g.pitch + gi] += a.ph.cwek * mu;
        // lq_{n-1}: the interior of lq_{n+1} (unchanged), 0 on the ring
        if (gi == 0 || gi == g.nxp - 1 || gj == 0 || gj == g.nyp - 1) a.lq_np1[o] = 0.0;
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

void launch_adj_bdy(const AdjBdyArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_bdy, dim3(blocks_for((long long)a.g.nxp * a.g.nyp)), dim3(256), 0, st, a);
}

int adj_inv1_blocks(const Geo& g) { return blocks_for((long long)g.nxp * g.nyp) > 256 ? 256 : blocks_for((long long)g.nxp * g.nyp); }

void launch_adj_inv1(const AdjInv1Args& a, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_inv1, dim3(a.nblocks), dim3(256), 0, st, a);
}

void launch_adj_inv2(const AdjInv2Args& a, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_inv2, dim3(1), dim3(256), 0, st, a);
}

void launch_adj_mix(const AdjMixArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_mix, dim3(blocks_for((long long)(a.g.nxp - 2) * (a.g.nyp - 2))), dim3(256), 0, st, a);
}

void launch_adj_add(double* dst, const double* src, long long n, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_add, dim3(blocks_for(n)), dim3(256), 0, st, dst, src, n);
}

void launch_adj_vort(const AdjVortArgs& a, hipStream_t st) {
    const dim3 grid((a.g.nxp + VORT_TX - 1) / VORT_TX, (a.g.nyp + VORT_TY - 1) / VORT_TY, a.g.nl);
    hipLaunchKernelGGL(k_adj_vort, grid, dim3(256), 0, st, a);
}

}  // namespace oqg

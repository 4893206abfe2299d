// oqg_kernels.hip -- gfx950 FP64 kernels of the QG timestep.
//
// One step of a basin is five launches on one stream:
//   k_vort    Arakawa Jacobian + Del^2/Del^4/Del^6 friction + forcing + leapfrog
//             (vorticity.F90:14-103, numerics.F90:340-468), 64x16 LDS tiles with
//             a halo of 3; writes q_new over qm.
//   k_cyc     (cyclic only) boundary line integrals of the momentum constraints
//             (vorticity.F90:57-62,84-101, numerics.F90:374-386).
//   k_fwd     layer->mode mix (modes.F90:248-266) fused with the x transform of
//             two interior rows per workgroup (DST-I via an odd-extended complex
//             FFT in the box, r2c in the channel; inhomog.F90:184-185,228-229),
//             Stockham mixed-radix FFT in LDS; writes column-major spectra.
//   k_ysolve  tridiagonal solves along y (inhomog.F90:188-199,232-243): one
//             wavefront per (mode, wavenumber) column, each lane eliminating a
//             block of rows, then parallel cyclic reduction across the 64 lanes
//             on the block separators; also the column sums for the box mass
//             integral.
//   k_inv     constraint/mass correction (homog.F90:280-366, constraint.F90:167-215),
//             inverse x transform, mode->layer mix (modes.F90:268-284) and the
//             boundary PV (vorsubs.F90:58-114); writes p_new over pm.
// All arithmetic is FP64 and the library is built with -ffp-contract=off.

#include "oqg_internal.h"

namespace oqg {

// ------------------------------------------------------------------ helpers

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ int wrapx(int gi, const Geo& g) {
    if (gi < 0) return gi + (g.nxp - 1);
    if (gi > g.nxp - 1) return gi - (g.nxp - 1);
    return gi;
}

// Value of a (pitch-layout) layer at global (gi,gj); cyclic wrap in x, 0 outside.
__device__ __forceinline__ double ldg_field(const double* f, int gi, int gj, const Geo& g) {
    if (gj < 0 || gj >= g.nyp) return 0.0;
    if (g.cyclic) gi = wrapx(gi, g);
    else if (gi < 0 || gi >= g.nxp) return 0.0;
    return f[(long long)gj * g.pitch + gi];
}

// del2_P_bc at one point (numerics.F90:410-468); s holds the field on an LDS
// tile of width W with the point at (r,c).
__device__ __forceinline__ double del2_lds(const double* s, int W, int r, int c, int gi, int gj,
                                           const Geo& g) {
    const double C = s[r * W + c];
    double dx2, dy2;
    if (gj == 0 || gj == g.nyp - 1) dx2 = 0.0;
    else if (g.cyclic || (gi >= 1 && gi <= g.nxp - 2)) dx2 = g.fx * ((s[r * W + c - 1] + s[r * W + c + 1]) - 2.0 * C);
    else if (gi == 0) dx2 = g.cxl * (s[r * W + c + 1] - C);
    else dx2 = g.cxr * (C - s[r * W + c - 1]);
    if (!g.cyclic && (gi == 0 || gi == g.nxp - 1)) dy2 = 0.0;
    else if (gj >= 1 && gj <= g.nyp - 2) dy2 = g.fy * ((s[(r - 1) * W + c] + s[(r + 1) * W + c]) - 2.0 * C);
    else if (gj == 0) dy2 = g.cyl * (s[(r + 1) * W + c] - C);
    else dy2 = g.cyr * (C - s[(r - 1) * W + c]);
    return dx2 + dy2;
}

// dPdx_2 / dPdy_2 (numerics.F90:311-338) on an LDS tile.
__device__ __forceinline__ double dX(const double* s, int W, int r, int c, int gi, const Geo& g) {
    return (g.cyclic || (gi >= 1 && gi <= g.nxp - 2)) ? s[r * W + c + 1] - s[r * W + c - 1] : 0.0;
}
__device__ __forceinline__ double dY(const double* s, int W, int r, int c, int gj, const Geo& g) {
    return (gj >= 1 && gj <= g.nyp - 2) ? s[(r + 1) * W + c] - s[(r - 1) * W + c] : 0.0;
}

// Arakawa J(q,p) (numerics.F90:361-372) at tile point (r,c), including the factor.
__device__ __forceinline__ double jac_lds(const double* q, const double* p, int W, int r, int c, int gi,
                                          int gj, const Geo& g) {
    const bool xin = g.cyclic || (gi >= 1 && gi <= g.nxp - 2);
    const bool yin = gj >= 1 && gj <= g.nyp - 2;
    const double dqx = dX(q, W, r, c, gi, g), dqy = dY(q, W, r, c, gj, g);
    const double dpx = dX(p, W, r, c, gi, g), dpy = dY(p, W, r, c, gj, g);
    double t = dqx * dpy - dqy * dpx;
    double v;
    v = xin ? q[r * W + c + 1] * dY(p, W, r, c + 1, gj, g) - q[r * W + c - 1] * dY(p, W, r, c - 1, gj, g) : 0.0;
    t = t + v;
    v = yin ? q[(r + 1) * W + c] * dX(p, W, r + 1, c, gi, g) - q[(r - 1) * W + c] * dX(p, W, r - 1, c, gi, g) : 0.0;
    t = t - v;
    v = yin ? p[(r + 1) * W + c] * dX(q, W, r + 1, c, gi, g) - p[(r - 1) * W + c] * dX(q, W, r - 1, c, gi, g) : 0.0;
    t = t + v;
    v = xin ? p[r * W + c + 1] * dY(q, W, r, c + 1, gj, g) - p[r * W + c - 1] * dY(q, W, r, c - 1, gj, g) : 0.0;
    t = t - v;
    return g.jfac * t;
}

// Deterministic block sum of NV values per thread (blockDim.x == 256): a fixed
// butterfly inside each wavefront, then the four wavefront sums in fixed order.
// Every thread returns the same value.  red holds NV*4 doubles.
template <int NV>
__device__ void block_sum(double (&v)[NV], double* red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_xor(v[i], off);
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < NV; i++) red[i * 4 + w] = v[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; i++) v[i] = (red[i * 4 + 0] + red[i * 4 + 1]) + (red[i * 4 + 2] + red[i * 4 + 3]);
    __syncthreads();
}

// ------------------------------------------------------------------ FFT in LDS

template <int R>
__device__ __forceinline__ void dft_small(double2 (&v)[R], const double2* tw, int L, int sign) {
    if constexpr (R == 2) {
        double2 a = v[0], b = v[1];
        v[0] = make_double2(a.x + b.x, a.y + b.y);
        v[1] = make_double2(a.x - b.x, a.y - b.y);
    } else if constexpr (R == 4) {
        double2 a0 = make_double2(v[0].x + v[2].x, v[0].y + v[2].y);
        double2 a1 = make_double2(v[0].x - v[2].x, v[0].y - v[2].y);
        double2 b0 = make_double2(v[1].x + v[3].x, v[1].y + v[3].y);
        double2 b1 = make_double2(v[1].x - v[3].x, v[1].y - v[3].y);
        double2 j = (sign < 0) ? make_double2(b1.y, -b1.x) : make_double2(-b1.y, b1.x);
        v[0] = make_double2(a0.x + b0.x, a0.y + b0.y);
        v[1] = make_double2(a1.x + j.x, a1.y + j.y);
        v[2] = make_double2(a0.x - b0.x, a0.y - b0.y);
        v[3] = make_double2(a1.x - j.x, a1.y - j.y);
    } else {
        double2 in[R];
#pragma unroll
        for (int r = 0; r < R; r++) in[r] = v[r];
        const int pstep = L / R;
        const double sg = (sign > 0) ? -1.0 : 1.0;
#pragma unroll
        for (int q = 0; q < R; q++) {
            double sr = in[0].x, si = in[0].y;
#pragma unroll
            for (int r = 1; r < R; r++) {
                const double2 w = tw[((r * q) % R) * pstep];
                const double c = w.x, s = sg * w.y;
                sr += in[r].x * c - in[r].y * s;
                si += in[r].x * s + in[r].y * c;
            }
            v[q] = make_double2(sr, si);
        }
    }
}

template <int R>
__device__ void fft_stage(const double2* __restrict__ in, double2* __restrict__ out, int L, int Ns,
                          const double2* __restrict__ tw, int sign) {
    const int LR = L / R;
    const int tstep = L / (Ns * R);
    for (int t = threadIdx.x; t < LR; t += blockDim.x) {
        const int k = t % Ns;
        double2 v[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = in[t + r * LR];
        if (k != 0) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                double2 w = tw[r * k * tstep];
                if (sign > 0) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        dft_small<R>(v, tw, L, sign);
        const int base = (t / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) out[base + r * Ns] = v[r];
    }
}

// Transforms buf0 (length L) with scratch buf1; returns the buffer holding the result.
__device__ double2* fft_lds(double2* buf0, double2* buf1, const Spec& s, const double2* tw, int sign) {
    double2* a = buf0;
    double2* b = buf1;
    int Ns = 1;
    for (int st = 0; st < s.nstages; st++) {
        const int R = s.radix[st];
        switch (R) {
            case 4: fft_stage<4>(a, b, s.L, Ns, tw, sign); break;
            case 2: fft_stage<2>(a, b, s.L, Ns, tw, sign); break;
            case 3: fft_stage<3>(a, b, s.L, Ns, tw, sign); break;
            default: fft_stage<5>(a, b, s.L, Ns, tw, sign); break;
        }
        __syncthreads();
        double2* t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}

// ------------------------------------------------------------------ k_vort

__global__ __launch_bounds__(256) void k_vort(VortArgs a) {
    constexpr int TX = VORT_TX, TY = VORT_TY;
    constexpr int W3 = TX + 6, H3 = TY + 6, W2 = TX + 4, H2 = TY + 4, W1 = TX + 2, H1 = TY + 2;
    __shared__ double s_pm[H3 * W3];
    __shared__ double s_d2[H2 * W2];
    __shared__ double s_d4[H1 * W1];
    __shared__ double s_q[H1 * W1];
    __shared__ double s_p[H1 * W1];
    __shared__ double s_red[4];
    const Geo& g = a.g;
    const int tid = threadIdx.x;
    const int k = blockIdx.z;
    const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
    const long long lo = (long long)k * g.fs;
    const double* pm = a.pm + lo;
    const double* q = a.q + lo;
    const double* p = a.p + lo;

    for (int idx = tid; idx < H3 * W3; idx += 256) {
        const int r = idx / W3, c = idx - (idx / W3) * W3;
        s_pm[idx] = ldg_field(pm, i0 - 3 + c, j0 - 3 + r, g);
    }
    for (int idx = tid; idx < H1 * W1; idx += 256) {
        const int r = idx / W1, c = idx - (idx / W1) * W1;
        s_q[idx] = ldg_field(q, i0 - 1 + c, j0 - 1 + r, g);
        s_p[idx] = ldg_field(p, i0 - 1 + c, j0 - 1 + r, g);
    }
    __syncthreads();
    // Del^2 of pm on the tile with halo 2
    for (int idx = tid; idx < H2 * W2; idx += 256) {
        const int r = idx / W2, c = idx - (idx / W2) * W2;
        int gi = i0 - 2 + c;
        const int gj = j0 - 2 + r;
        double v = 0.0;
        if (gj >= 0 && gj < g.nyp) {
            if (g.cyclic) {
                gi = wrapx(gi, g);
                v = del2_lds(s_pm, W3, r + 1, c + 1, gi, gj, g);
            } else if (gi >= 0 && gi < g.nxp) {
                v = del2_lds(s_pm, W3, r + 1, c + 1, gi, gj, g);
            }
        }
        s_d2[idx] = v;
    }
    __syncthreads();
    // Del^4 on the tile with halo 1
    for (int idx = tid; idx < H1 * W1; idx += 256) {
        const int r = idx / W1, c = idx - (idx / W1) * W1;
        int gi = i0 - 1 + c;
        const int gj = j0 - 1 + r;
        double v = 0.0;
        if (gj >= 0 && gj < g.nyp) {
            if (g.cyclic) {
                gi = wrapx(gi, g);
                v = del2_lds(s_d2, W2, r + 1, c + 1, gi, gj, g);
            } else if (gi >= 0 && gi < g.nxp) {
                v = del2_lds(s_d2, W2, r + 1, c + 1, gi, gj, g);
            }
        }
        s_d4[idx] = v;
    }
    __syncthreads();
    const double c2 = a.ph.c2[k], c4 = a.ph.c4[k];
    double epart = 0.0;
    for (int idx = tid; idx < TX * TY; idx += 256) {
        const int r = idx / TX, c = idx - (idx / TX) * TX;
        const int gi = i0 + c, gj = j0 + r;
        if (gi >= g.nxp || gj >= g.nyp) continue;
        const double d6 = del2_lds(s_d4, W1, r + 1, c + 1, gi, gj, g);
        const double d4 = s_d4[(r + 1) * W1 + c + 1];
        const double d2 = s_d2[(r + 2) * W2 + c + 2];
        double dqdt = jac_lds(s_q, s_p, W1, r + 1, c + 1, gi, gj, g);
        dqdt = dqdt + c2 * d4 - c4 * d6;
        const long long o2 = (long long)gj * g.pitch + gi;
        if (k == 0) {
            const double e = a.ent1[o2];
            dqdt = dqdt + a.ph.cent1 * e + a.ph.cwek * a.wekp[o2];
            const bool ex = (gi == 0 || gi == g.nxp - 1), ey = (gj == 0 || gj == g.nyp - 1);
            const double w = (ex && ey) ? 0.25 : ((ex || ey) ? 0.5 : 1.0);
            epart += w * e;
        }
        if (k == 1) dqdt = dqdt - a.ph.cent2 * a.ent1[o2];
        if (k == g.kt) dqdt = dqdt - a.ph.cdrag * d2;
        double* qn = a.qnew + lo + o2;
        *qn = *qn + a.tdt * dqdt;
    }
    if (k == 0) {
        double v[1] = {epart};
        block_sum<1>(v, s_red);
        if (tid == 0) a.entpart[blockIdx.y * a.tiles_x + blockIdx.x] = v[0];
    }
}

void launch_vort(const VortArgs& a, hipStream_t st) {
    dim3 grid(a.tiles_x, a.tiles_y, a.g.nl);
    hipLaunchKernelGGL(k_vort, grid, dim3(256), 0, st, a);
}

// ------------------------------------------------------------------ k_cyc

// Sum over a row with the int_P_dx weights (numerics.F90:206-212), times dx.
__global__ __launch_bounds__(256) void k_cyc(CycArgs a) {
    extern __shared__ double smem[];
    const Geo& g = a.g;
    const int nx = g.nxp, ny = g.nyp;
    const int k = blockIdx.x;
    const long long lo = (long long)k * g.fs;
    double* d2 = smem;            // rows 0,1,2 | ny-3,ny-2,ny-1  (6 x nx)
    double* d4 = smem + 6 * nx;   // rows 0,1 | ny-2,ny-1         (4 x nx)
    __shared__ double red[10 * 4];
    const int d2rows[6] = {0, 1, 2, ny - 3, ny - 2, ny - 1};
    const double* pm = a.pm + lo;
    for (int idx = threadIdx.x; idx < 6 * nx; idx += 256) {
        const int rr = idx / nx, i = idx - rr * nx;
        const int j = d2rows[rr];
        // Del^2 of pm at (i,j) straight from global memory (x periodic)
        const int il = (i == 0) ? nx - 2 : i - 1, ir = (i == nx - 1) ? 1 : i + 1;
        const double C = pm[(long long)j * g.pitch + i];
        double dx2, dy2;
        if (j == 0 || j == ny - 1) dx2 = 0.0;
        else dx2 = g.fx * ((pm[(long long)j * g.pitch + il] + pm[(long long)j * g.pitch + ir]) - 2.0 * C);
        if (j >= 1 && j <= ny - 2) dy2 = g.fy * ((pm[(long long)(j - 1) * g.pitch + i] + pm[(long long)(j + 1) * g.pitch + i]) - 2.0 * C);
        else if (j == 0) dy2 = g.cyl * (pm[(long long)1 * g.pitch + i] - C);
        else dy2 = g.cyr * (C - pm[(long long)(ny - 2) * g.pitch + i]);
        d2[idx] = dx2 + dy2;
    }
    __syncthreads();
    const int d4rows[4] = {0, 1, ny - 2, ny - 1};
    const int d4src[4] = {0, 1, 4, 5};   // index of the same row in d2
    for (int idx = threadIdx.x; idx < 4 * nx; idx += 256) {
        const int rr = idx / nx, i = idx - rr * nx;
        const int j = d4rows[rr], sr = d4src[rr];
        const int il = (i == 0) ? nx - 2 : i - 1, ir = (i == nx - 1) ? 1 : i + 1;
        const double C = d2[sr * nx + i];
        double dx2, dy2;
        if (j == 0 || j == ny - 1) dx2 = 0.0;
        else dx2 = g.fx * ((d2[sr * nx + il] + d2[sr * nx + ir]) - 2.0 * C);
        if (j >= 1 && j <= ny - 2) dy2 = g.fy * ((d2[(sr - 1) * nx + i] + d2[(sr + 1) * nx + i]) - 2.0 * C);
        else if (j == 0) dy2 = g.cyl * (d2[(sr + 1) * nx + i] - C);
        else dy2 = g.cyr * (C - d2[(sr - 1) * nx + i]);
        d4[idx] = dx2 + dy2;
    }
    __syncthreads();
    const double* q = a.q + lo;
    const double* p = a.p + lo;
    const double* pkt = a.pm + (long long)g.kt * g.fs;
    double v[10];
    for (int t = 0; t < 10; t++) v[t] = 0.0;
    for (int i = threadIdx.x; i < nx; i += 256) {
        const double w = (i == 0 || i == nx - 1) ? 0.5 : 1.0;
        const int il = (i == 0) ? nx - 2 : i - 1, ir = (i == nx - 1) ? 1 : i + 1;
        v[0] += w * (d2[1 * nx + i] - d2[0 * nx + i]);               // ap3s
        v[1] += w * (d2[5 * nx + i] - d2[4 * nx + i]);               // ap3n
        v[2] += w * (d4[1 * nx + i] - d4[0 * nx + i]);               // ap5s
        v[3] += w * (d4[3 * nx + i] - d4[2 * nx + i]);               // ap5n
        const double dpx2 = p[(long long)1 * g.pitch + ir] - p[(long long)1 * g.pitch + il];
        const double dpxn = p[(long long)(ny - 2) * g.pitch + ir] - p[(long long)(ny - 2) * g.pitch + il];
        v[4] += w * (q[i] * dpx2);                                   // aj5sms
        v[5] += w * (q[(long long)1 * g.pitch + i] * dpx2);          // aj9sms
        v[6] += w * (q[(long long)(ny - 1) * g.pitch + i] * dpxn);   // aj5smn
        v[7] += w * (q[(long long)(ny - 2) * g.pitch + i] * dpxn);   // aj9smn
        v[8] += w * (pkt[(long long)1 * g.pitch + i] - pkt[i]);      // bdsums
        v[9] += w * (pkt[(long long)(ny - 1) * g.pitch + i] - pkt[(long long)(ny - 2) * g.pitch + i]);  // bdsumn
    }
    block_sum<10>(v, red);
    double e[4] = {0.0, 0.0, 0.0, 0.0};
    if (k == 0) {
        for (int i = threadIdx.x; i < nx; i += 256) {
            const double w = (i == 0 || i == nx - 1) ? 0.5 : 1.0;
            e[0] += w * a.ent[i];
            e[1] += w * a.ent[(long long)(ny - 1) * g.pitch + i];
            if (a.ent2s) e[2] += w * a.ent2s[i];
            if (a.ent2n) e[3] += w * a.ent2n[i];
        }
        block_sum<4>(e, red);
    }
    if (threadIdx.x == 0) {
        const double dx = g.dx, dy = g.dy;
        double* o = a.aux;
        o[A_AP3S + k] = a.ph.ah2[k] * (v[0] * dx) / dy;
        o[A_AP3N + k] = a.ph.ah2[k] * (v[1] * dx) / dy;
        o[A_AP5S + k] = a.ph.ah4[k] * (v[2] * dx) / dy;
        o[A_AP5N + k] = a.ph.ah4[k] * (v[3] * dx) / dy;
        const double aj5s = (v[4] * dx) / (2.0 * dx), aj9s = (v[5] * dx) / (2.0 * dx);
        const double aj5n = (v[6] * dx) / (2.0 * dx), aj9n = (v[7] * dx) / (2.0 * dx);
        o[A_AJIS + k] = (aj5s + 2.0 * aj9s) / 6.0;
        o[A_AJIN + k] = -(aj5n + 2.0 * aj9n) / 6.0;
        if (k == g.kt) {
            o[A_BDRS] = a.ph.cd_bdr * ((v[8] * dx) / dy);
            o[A_BDRN] = a.ph.cd_bdr * ((v[9] * dx) / dy);
        }
        if (k == 0) {
            const double dz = (double)a.ph.dz_sign;
            for (int l = 0; l < g.nl; l++) { o[A_BES + l] = 0.0; o[A_BEN + l] = 0.0; }
            o[A_BES + 0] = dz * a.ph.fohfac[0] * (e[0] * dx);
            o[A_BES + 1] = dz * a.ph.fohfac[0] * (e[2] * dx);
            o[A_BEN + 0] = -dz * a.ph.fohfac[1] * (e[1] * dx);
            o[A_BEN + 1] = -dz * a.ph.fohfac[1] * (e[3] * dx);
            o[A_TXIS] = a.txis;
            o[A_TXIN] = a.txin;
        }
    }
}

void launch_cyc(const CycArgs& a, hipStream_t st) {
    const size_t lds = sizeof(double) * 10 * (size_t)a.g.nxp;
    hipLaunchKernelGGL(k_cyc, dim3(a.g.nl), dim3(256), lds, st, a);
}

// ------------------------------------------------------------------ k_fwd

__device__ __forceinline__ double rhs_mode(const FwdArgs& a, int m, int gi, int gj) {
    // p_modes_rhs (modes.F90:259-264)
    const Geo& g = a.g;
    const long long o = (long long)gj * g.pitch + gi;
    const double by = g.beta * a.yprel[gj];
    double s = 0.0;
    for (int k = 0; k < g.nl; k++) {
        double ql = a.q[k * g.fs + o] - by;
        if (k == g.kt && a.ddyn) ql = ql - a.ddyn[o];
        s += ql * a.mx.ctl2m[k * MAXL + m];
    }
    return g.fnot * s;
}

__global__ __launch_bounds__(256) void k_fwd(FwdArgs a) {
    extern __shared__ double2 zbuf[];
    const Geo& g = a.g;
    const Spec& s = a.s;
    const int L = s.L;
    double2* z0 = zbuf;
    double2* z1 = zbuf + L;
    const int jp = blockIdx.x;
    const int m = (a.raw_mode >= 0) ? a.raw_mode : (int)blockIdx.y;
    const int r0 = 2 * jp, r1 = 2 * jp + 1;
    const bool has1 = r1 < s.nyi;
    const int gj0 = r0 + 1, gj1 = r1 + 1;
    const double2* tw = (const double2*)a.tw;
    if (!g.cyclic) {
        const int N = L / 2, n = s.nin;
        for (int t = threadIdx.x; t <= N; t += blockDim.x) {
            double2 v = make_double2(0.0, 0.0);
            if (t >= 1 && t <= n) {
                double x, y = 0.0;
                if (a.raw_mode >= 0) {
                    x = a.q[(long long)gj0 * g.pitch + t];
                    if (has1) y = a.q[(long long)gj1 * g.pitch + t];
                } else {
                    x = rhs_mode(a, m, t, gj0);
                    if (has1) y = rhs_mode(a, m, t, gj1);
                }
                v = make_double2(x, y);
            }
            z0[t] = v;
            if (t >= 1 && t <= n) z0[L - t] = make_double2(-v.x, -v.y);
        }
    } else {
        for (int t = threadIdx.x; t < L; t += blockDim.x) {
            double x, y = 0.0;
            if (a.raw_mode >= 0) {
                x = a.q[(long long)gj0 * g.pitch + t];
                if (has1) y = a.q[(long long)gj1 * g.pitch + t];
            } else {
                x = rhs_mode(a, m, t, gj0);
                if (has1) y = rhs_mode(a, m, t, gj1);
            }
            z0[t] = make_double2(x, y);
        }
    }
    __syncthreads();
    const double2* Z = fft_lds(z0, z1, s, tw, -1);
    double* spec = a.spec + (long long)m * s.ncol * s.nypad;
    if (!g.cyclic) {
        for (int K = threadIdx.x + 1; K <= s.nin; K += blockDim.x) {
            const double2 v = Z[K];
            // DST(x)_K = -Im Z_K (row r0), DST(y)_K = Re Z_K (row r1)
            *(double2*)(spec + (long long)(K - 1) * s.nypad + r0) = make_double2(-v.y, v.x);
        }
    } else {
        const int N = L;
        for (int kk = threadIdx.x; kk <= N / 2; kk += blockDim.x) {
            const double2 A = Z[kk], B = Z[(N - kk) % N];
            const double xr = 0.5 * (A.x + B.x), xi = 0.5 * (A.y - B.y);
            const double yr = 0.5 * (A.y + B.y), yi = 0.5 * (B.x - A.x);
            *(double2*)(spec + (long long)(2 * kk) * s.nypad + r0) = make_double2(xr, yr);
            *(double2*)(spec + (long long)(2 * kk + 1) * s.nypad + r0) = make_double2(xi, yi);
        }
    }
}

void launch_fwd(const FwdArgs& a, int nmodes, hipStream_t st) {
    const int npairs = (a.s.nyi + 1) / 2;
    const size_t lds = sizeof(double2) * 2 * (size_t)a.s.L;
    hipLaunchKernelGGL(k_fwd, dim3(npairs, nmodes), dim3(256), lds, st, a);
}

// ------------------------------------------------------------------ k_ysolve

// Wavefront tridiagonal solver: a*u[j-1] + b*u[j] + a*u[j+1] = f[j], j = 0..nyi-1,
// u[-1] = u[nyi] = 0.  Lane l owns rows [l*R, l*R+R).  The last row of each block
// is its separator; interior rows are eliminated locally in terms of the two
// neighbouring separators, and the separators' tridiagonal system is solved by
// parallel cyclic reduction across the 64 lanes.
template <int R>
__global__ __launch_bounds__(256) void k_ysolve(YArgs a) {
    const Spec& s = a.s;
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= a.nmodes * s.ncol) return;   // whole wavefront leaves together
    const int m = a.mode0 + c / s.ncol;
    const int col = c - (c / s.ncol) * s.ncol;
    double* colp = a.spec + ((long long)m * s.ncol + col) * s.nypad;
    const double b = a.bcol[m * s.ncol + col];
    const double aa = s.a;
    const int rs = lane * R;
    int L = s.nyi - rs;
    L = L < 0 ? 0 : (L > R ? R : L);

    double f[R];
#pragma unroll
    for (int t = 0; t < R; t++) f[t] = (t < L) ? colp[rs + t] : 0.0;

    double v[R], x[R], y[R], gam[R];
#pragma unroll
    for (int t = 0; t < R; t++) { v[t] = 0.0; x[t] = 0.0; y[t] = 0.0; gam[t] = 0.0; }
    if (L >= 2) {
        double betinv = 1.0 / b;
        v[0] = f[0] * betinv;
        x[0] = (-aa) * betinv;
        y[0] = (L == 2) ? (-aa) * betinv : 0.0;
#pragma unroll
        for (int t = 1; t < R - 1; t++) {
            if (t <= L - 2) {
                gam[t] = aa * betinv;
                betinv = 1.0 / (b - aa * gam[t]);
                v[t] = (f[t] - aa * v[t - 1]) * betinv;
                x[t] = (0.0 - aa * x[t - 1]) * betinv;
                y[t] = (((t == L - 2) ? -aa : 0.0) - aa * y[t - 1]) * betinv;
            }
        }
#pragma unroll
        for (int t = R - 3; t >= 0; t--) {
            if (t <= L - 3) {
                v[t] = v[t] - gam[t + 1] * v[t + 1];
                x[t] = x[t] - gam[t + 1] * x[t + 1];
                y[t] = y[t] - gam[t + 1] * y[t + 1];
            }
        }
    }
    // first/last interior-row responses (v, x: coefficient of left separator, y: of own separator)
    double vf = 0.0, xf = 0.0, yf = 0.0, vl = 0.0, xl = 0.0, yl = 0.0, fs = 0.0;
    if (L >= 2) {
        vf = v[0]; xf = x[0]; yf = y[0];
#pragma unroll
        for (int t = 0; t < R - 1; t++)
            if (t == L - 2) { vl = v[t]; xl = x[t]; yl = y[t]; }
    } else if (L == 1) {
        yf = 1.0;   // next block's "first interior row" is its separator
        xl = 1.0;   // own "last interior row" is the left separator
    }
#pragma unroll
    for (int t = 0; t < R; t++)
        if (t == L - 1) fs = f[t];
    const double vfn = __shfl(vf, (lane + 1) & 63), xfn = __shfl(xf, (lane + 1) & 63), yfn = __shfl(yf, (lane + 1) & 63);
    const bool nxt = (lane < 63);
    double A, B, C, D;
    if (L >= 1) {
        A = (lane == 0) ? 0.0 : aa * xl;
        B = b + aa * yl + (nxt ? aa * xfn : 0.0);
        C = nxt ? aa * yfn : 0.0;
        D = fs - aa * vl - (nxt ? aa * vfn : 0.0);
    } else {
        A = 0.0; B = 1.0; C = 0.0; D = 0.0;
    }
#pragma unroll
    for (int st = 1; st < 64; st <<= 1) {
        const int lm = lane - st, lp = lane + st;
        const double Am = __shfl(A, lm & 63), Bm = __shfl(B, lm & 63), Cm = __shfl(C, lm & 63), Dm = __shfl(D, lm & 63);
        const double Ap = __shfl(A, lp & 63), Bp = __shfl(B, lp & 63), Cp = __shfl(C, lp & 63), Dp = __shfl(D, lp & 63);
        const bool hm = lm >= 0, hp = lp < 64;
        const double k1 = hm ? A / Bm : 0.0;
        const double k2 = hp ? C / Bp : 0.0;
        const double An = hm ? -Am * k1 : 0.0;
        const double Cn = hp ? -Cp * k2 : 0.0;
        const double Bn = B - (hm ? Cm * k1 : 0.0) - (hp ? Ap * k2 : 0.0);
        const double Dn = D - (hm ? Dm * k1 : 0.0) - (hp ? Dp * k2 : 0.0);
        A = An; B = Bn; C = Cn; D = Dn;
    }
    const double S = D / B;
    double Sm = __shfl(S, (lane - 1) & 63);
    if (lane == 0) Sm = 0.0;
    double csum = 0.0;
#pragma unroll
    for (int t = 0; t < R; t++) {
        if (t < L) {
            const double u = (t == L - 1) ? S : v[t] + x[t] * Sm + y[t] * S;
            colp[rs + t] = u;
            csum += u;
        }
    }
    if (a.colsum) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) csum += __shfl_xor(csum, off);
        if (lane == 0) a.colsum[m * s.ncol + col] = csum;
    }
}

// Reference-order Thomas (Numerical Recipes form, inhomog.F90:188-199), one
// thread per (mode, wavenumber) column, walking the column in y.
__global__ __launch_bounds__(256) void k_ythomas(YArgs a) {
    const Spec& s = a.s;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nmodes * s.ncol) return;
    const int m = a.mode0 + c / s.ncol;
    const int col = c - (c / s.ncol) * s.ncol;
    double* u = a.spec + ((long long)m * s.ncol + col) * s.nypad;
    double* gam = a.gam + ((long long)m * s.ncol + col) * s.nypad;
    const double b = a.bcol[m * s.ncol + col];
    const double aa = s.a;
    double betinv = 1.0 / b;
    double prev = u[0] * betinv;
    u[0] = prev;
    for (int j = 1; j < s.nyi; j++) {
        const double gj = aa * betinv;
        gam[j] = gj;
        betinv = 1.0 / (b - aa * gj);
        prev = (u[j] - aa * prev) * betinv;
        u[j] = prev;
    }
    double csum = prev;
    for (int j = s.nyi - 2; j >= 0; j--) {
        prev = u[j] - gam[j + 1] * prev;
        u[j] = prev;
        csum += prev;
    }
    if (a.colsum) a.colsum[m * s.ncol + col] = csum;
}

void launch_ysolve(const YArgs& a, int variant, hipStream_t st) {
    const int ncols = a.nmodes * a.s.ncol;
    if (variant == 1) {
        hipLaunchKernelGGL(k_ythomas, dim3((ncols + 255) / 256), dim3(256), 0, st, a);
        return;
    }
    const int blocks = (ncols + 3) / 4;
    switch (a.s.R) {
#define YCASE(RR) case RR: hipLaunchKernelGGL(k_ysolve<RR>, dim3(blocks), dim3(256), 0, st, a); break;
        YCASE(1) YCASE(2) YCASE(3) YCASE(4) YCASE(5) YCASE(6) YCASE(7) YCASE(8)
        YCASE(9) YCASE(10) YCASE(11) YCASE(12) YCASE(13) YCASE(14) YCASE(15) YCASE(16)
#undef YCASE
        default: break;
    }
}

// ------------------------------------------------------------------ k_inv

// LU solve with DGERFS-style refinement of the (nl-1)^2 mass-constraint system
// (homog.F90:356-359, linalg.F90:57-93), run by one thread.
__device__ void small_lu_solve(const BoxHom& h, int n, const double* rhs, double* x) {
    auto lusolve = [&](double* v) {
        for (int k = 0; k < n; k++) {
            const int p = h.ipiv[k];
            if (p != k) { const double t = v[k]; v[k] = v[p]; v[p] = t; }
        }
        for (int i = 0; i < n; i++)
            for (int j = 0; j < i; j++) v[i] -= h.LU[i * MAXL + j] * v[j];
        for (int i = n - 1; i >= 0; i--) {
            for (int j = i + 1; j < n; j++) v[i] -= h.LU[i * MAXL + j] * v[j];
            v[i] /= h.LU[i * MAXL + i];
        }
    };
    for (int i = 0; i < n; i++) x[i] = rhs[i];
    lusolve(x);
    double lstres = 3.0;
    const double eps = 1.1102230246251565e-16;
    for (int it = 0; it < 5; it++) {
        double r[MAXL], berr = 0.0;
        for (int i = 0; i < n; i++) {
            double sres = rhs[i], den = fabs(rhs[i]);
            for (int j = 0; j < n; j++) {
                sres -= h.cdhoc[i * MAXL + j] * x[j];
                den += fabs(h.cdhoc[i * MAXL + j]) * fabs(x[j]);
            }
            r[i] = sres;
            if (den > 0.0 && fabs(sres) / den > berr) berr = fabs(sres) / den;
        }
        if (!(berr > eps && 2.0 * berr <= lstres)) break;
        lusolve(r);
        for (int i = 0; i < n; i++) x[i] += r[i];
        lstres = berr;
    }
}

__global__ __launch_bounds__(256) void k_inv(InvArgs a) {
    extern __shared__ double smem[];
    const Geo& g = a.g;
    const Spec& s = a.s;
    const int L = s.L;
    const int nl = g.nl;
    const bool raw = a.raw_mode >= 0;
    const int nm = raw ? 1 : nl;
    double2* z0 = (double2*)smem;
    double2* z1 = z0 + L;
    double* inh = (double*)(z1 + L);     // [mi][2][pitch]
    __shared__ double red[(MAXL + 1) * 4];
    __shared__ double coef[3 * MAXL + 4];
    __shared__ double s_ctm2l[MAXL * MAXL];
    __shared__ double s_amat[MAXL * MAXL];
    for (int t = threadIdx.x; t < MAXL * MAXL; t += blockDim.x) {
        s_ctm2l[t] = a.mx->ctm2l[t];
        s_amat[t] = a.mx->amat[t];
    }
    __syncthreads();
    const int jp = blockIdx.x;
    const int r0 = 2 * jp, r1 = 2 * jp + 1;
    const bool has1 = r1 < s.nyi;
    const int gj0 = r0 + 1, gj1 = r1 + 1;
    const bool first = (jp == 0), last = (jp == (int)gridDim.x - 1);
    const double2* tw = (const double2*)a.tw;

    // ---- constraint / mass correction coefficients (identical in every block)
    if (!raw) {
        double v[MAXL + 1];
        for (int t = 0; t <= MAXL; t++) v[t] = 0.0;
        for (int t = threadIdx.x; t < a.nentpart; t += 256) v[MAXL] += a.entpart[t];
        if (!g.cyclic) {
            for (int col = threadIdx.x; col < s.ncol; col += 256) {
                const double w = a.colw[col];
                for (int m = 0; m < nl; m++) v[m] += w * a.colsum[m * s.ncol + col];
            }
        }
        block_sum<MAXL + 1>(v, red);
        if (threadIdx.x == 0) {
            const double* ci = a.cin;
            double cn_[C_LEN];
            for (int t = 0; t < C_LEN; t++) cn_[t] = ci[t];
            // update_mass_constr (constraint.F90:167-179), ent_xn(1) = int_P_dA(ent1)
            const double exn = (v[MAXL] * g.dx) * g.dy;
            for (int l = 0; l < nl - 1; l++) {
                const double e = (l == 0) ? exn : 0.0;
                cn_[C_DPI + l] = ci[C_DPIP + l] - a.tdt * a.ph.gp[l] * e;
                cn_[C_DPIP + l] = ci[C_DPI + l];
            }
            if (!g.cyclic) {
                // box_homog (homog.F90:345-364)
                double im[MAXL], ik[MAXL], rhs[MAXL], alpha[MAXL];
                for (int m = 0; m < nl; m++) im[m] = (v[m] * g.dx) * g.dy;
                for (int k = 0; k < nl; k++) {
                    double t = 0.0;
                    for (int m = 0; m < nl; m++) t += s_ctm2l[m * MAXL + k] * im[m];
                    ik[k] = t;
                }
                for (int l = 0; l < nl - 1; l++) rhs[l] = cn_[C_DPI + l] - (ik[l + 1] - ik[l]);
                small_lu_solve(*a.bh, nl - 1, rhs, alpha);
                for (int l = 0; l < nl - 1; l++) coef[l] = alpha[l];
            } else {
                // step_constraints (constraint.F90:182-215)
                const double* x = a.aux;
                const double f0 = g.fnot, dy = g.dy;
                double rhss[MAXL], rhsn[MAXL];
                for (int k = 0; k < nl; k++) {
                    rhss[k] = -0.5 * dy * f0 * x[A_BES + k];
                    rhsn[k] = -0.5 * dy * f0 * x[A_BEN + k];
                }
                rhss[0] = rhss[0] + a.ph.tau_sign * (f0 / a.ph.h[0]) * x[A_TXIS];
                rhsn[0] = rhsn[0] - a.ph.tau_sign * (f0 / a.ph.h[0]) * x[A_TXIN];
                rhss[nl - 1] = rhss[nl - 1] + (f0 / a.ph.h[nl - 1]) * x[A_BDRS];
                rhsn[nl - 1] = rhsn[nl - 1] - (f0 / a.ph.h[nl - 1]) * x[A_BDRN];
                for (int k = 0; k < nl; k++) {
                    rhss[k] = rhss[k] + x[A_AJIS + k] - x[A_AP3S + k] + x[A_AP5S + k];
                    rhsn[k] = rhsn[k] + x[A_AJIN + k] + x[A_AP3N + k] - x[A_AP5N + k];
                }
                for (int k = 0; k < nl; k++) {
                    cn_[C_CS + k] = ci[C_CSP + k] + a.tdt * rhss[k];
                    cn_[C_CN + k] = ci[C_CNP + k] + a.tdt * rhsn[k];
                    cn_[C_CSP + k] = ci[C_CS + k];
                    cn_[C_CNP + k] = ci[C_CN + k];
                }
                // cyclic_homog (homog.F90:305-320); row integrals from the k = 0 spectra
                double clhss[MAXL], clhsn[MAXL];
                for (int m = 0; m < nl; m++) {
                    double ts = 0.0, tn = 0.0;
                    for (int k = 0; k < nl; k++) {
                        ts += a.mx->ctl2m[k * MAXL + m] * cn_[C_CS + k];
                        tn += a.mx->ctl2m[k * MAXL + m] * cn_[C_CN + k];
                    }
                    const double* c0p = a.spec + (long long)m * s.ncol * s.nypad;
                    clhss[m] = ts + (a.ch->c0 * c0p[0]) / dy;
                    clhsn[m] = tn + (a.ch->c0 * c0p[s.nyi - 1]) / dy;
                }
                coef[0] = clhss[0] * a.ch->hbsi;
                for (int m = 1; m < nl; m++) {
                    coef[1 + m] = a.ch->hc2n[m] * clhss[m] - a.ch->hc2s[m] * clhsn[m];
                    coef[1 + MAXL + m] = a.ch->hc1s[m] * clhsn[m] - a.ch->hc1n[m] * clhss[m];
                }
            }
            if (blockIdx.x == 0)
                for (int t = 0; t < C_LEN; t++) a.cout[t] = cn_[t];
        }
        __syncthreads();
    }

    // ---- inverse x transforms, one mode at a time
    for (int mi = 0; mi < nm; mi++) {
        const int m = raw ? a.raw_mode : mi;
        const double* sp = a.spec + (long long)m * s.ncol * s.nypad;
        const double fn = s.ftnorm;
        if (!g.cyclic) {
            const int N = L / 2, n = s.nin;
            for (int t = threadIdx.x; t <= N; t += blockDim.x) {
                double2 v = make_double2(0.0, 0.0);
                if (t >= 1 && t <= n) {
                    const double2 u = *(const double2*)(sp + (long long)(t - 1) * s.nypad + r0);
                    v = make_double2(fn * u.x, has1 ? fn * u.y : 0.0);
                }
                z0[t] = v;
                if (t >= 1 && t <= n) z0[L - t] = make_double2(-v.x, -v.y);
            }
        } else {
            const int N = L, h = N / 2;
            for (int t = threadIdx.x; t < N; t += blockDim.x) {
                const int kk = (t <= h) ? t : N - t;
                const double2 re = *(const double2*)(sp + (long long)(2 * kk) * s.nypad + r0);
                const double2 im = *(const double2*)(sp + (long long)(2 * kk + 1) * s.nypad + r0);
                double xr = fn * re.x, yr = has1 ? fn * re.y : 0.0;
                double xi = fn * im.x, yi = has1 ? fn * im.y : 0.0;
                if (t > h) { xi = -xi; yi = -yi; }
                if (t == 0 || (N % 2 == 0 && t == h)) { xi = 0.0; yi = 0.0; }
                z0[t] = make_double2(xr - yi, xi + yr);
            }
        }
        __syncthreads();
        const double2* Z = fft_lds(z0, z1, s, tw, g.cyclic ? +1 : -1);
        double* row0 = inh + (long long)mi * 2 * g.pitch;
        double* row1 = row0 + g.pitch;
        for (int i = threadIdx.x; i < g.nxp; i += blockDim.x) {
            double x0 = 0.0, x1 = 0.0;
            if (!g.cyclic) {
                if (i >= 1 && i <= s.nin) { const double2 v = Z[i]; x0 = -v.y; x1 = v.x; }
            } else {
                const double2 v = Z[(i == g.nxp - 1) ? 0 : i];
                x0 = v.x;
                x1 = v.y;
            }
            row0[i] = x0;
            row1[i] = x1;
        }
        __syncthreads();
    }

    if (raw) {
        double* o = a.out_raw;
        for (int i = threadIdx.x; i < g.nxp; i += blockDim.x) {
            o[(long long)gj0 * g.pitch + i] = inh[i];
            if (has1) o[(long long)gj1 * g.pitch + i] = inh[g.pitch + i];
            if (first) o[i] = 0.0;
            if (last) o[(long long)(g.nyp - 1) * g.pitch + i] = 0.0;
        }
        return;
    }

    // ---- p = modes_to_layers(inhomog + homcor), then ocqbdy on the boundary ring
    // row selector: 0 -> gj0, 1 -> gj1 (interior, from LDS), 2 -> boundary row (inhomog = 0)
    auto hc = [&](int m, int i, int gj) -> double {
        if (!g.cyclic) {
            if (m == 0) return 0.0;
            return coef[m - 1] * a.hom[(long long)(m - 1) * g.fs + (long long)gj * g.pitch + i];
        }
        if (m == 0) return coef[0] * a.hom[gj];
        const double* b1 = a.hom + g.nyp;
        const double* b2 = b1 + (long long)(nl - 1) * g.nyp;
        return coef[1 + m] * b1[(m - 1) * g.nyp + gj] + coef[1 + MAXL + m] * b2[(m - 1) * g.nyp + gj];
    };
    auto pval = [&](int i, int sel, int gj, double* pk) {
        double vm[MAXL];
        for (int m = 0; m < nl; m++) {
            const double iv = (sel == 2) ? 0.0 : inh[(long long)m * 2 * g.pitch + sel * g.pitch + i];
            vm[m] = iv + hc(m, i, gj);
        }
        for (int k = 0; k < nl; k++) {
            double t = 0.0;
            for (int m = 0; m < nl; m++) t += vm[m] * s_ctm2l[m * MAXL + k];
            pk[k] = t;
        }
    };
    const double qfac = a.ph.qbfac;
    // boundary q at (i,gj) from p there (pb) and at the inner neighbour (pi)
    auto qbdy = [&](int i, int gj, int kind /*0 row,1 col,2 corner*/, const double* pb, const double* pin) {
        const long long o = (long long)gj * g.pitch + i;
        const int reps = (kind == 2) ? 2 : 1;
        for (int k = 0; k < nl; k++) {
            double qv = qfac * (pin[k] - pb[k]);
            double ap = 0.0;
            for (int l = 0; l < nl; l++) ap += s_amat[k * MAXL + l] * pb[l];
            for (int r = 0; r < reps; r++) qv = qv - g.fnot * ap;
            for (int r = 0; r < reps; r++) qv = qv + g.beta * a.yprel[gj];
            if (k == g.kt && a.ddyn)
                for (int r = 0; r < reps; r++) qv = qv + a.ddyn[o];
            a.qnew[k * g.fs + o] = qv;
        }
    };
    const int nsel = has1 ? 2 : 1;
    for (int idx = threadIdx.x; idx < nsel * g.nxp; idx += blockDim.x) {
        const int sel = idx / g.nxp, i = idx - sel * g.nxp;
        const int gj = sel == 0 ? gj0 : gj1;
        double pk[MAXL];
        pval(i, sel, gj, pk);
        const long long o = (long long)gj * g.pitch + i;
        for (int k = 0; k < nl; k++) a.pnew[k * g.fs + o] = pk[k];
        if (!g.cyclic && (i == 0 || i == g.nxp - 1)) {
            double pin[MAXL];
            pval(i == 0 ? 1 : g.nxp - 2, sel, gj, pin);
            qbdy(i, gj, 1, pk, pin);
        }
    }
    // boundary rows j = 0 (first block) and j = nyp-1 (last block)
    for (int side = 0; side < 2; side++) {
        if ((side == 0 && !first) || (side == 1 && !last)) continue;
        const int gj = side == 0 ? 0 : g.nyp - 1;
        const int gin = side == 0 ? 1 : g.nyp - 2;
        const int selin = side == 0 ? 0 : ((gin == gj1 && has1) ? 1 : 0);
        for (int i = threadIdx.x; i < g.nxp; i += blockDim.x) {
            double pb[MAXL], pin[MAXL];
            pval(i, 2, gj, pb);
            const long long o = (long long)gj * g.pitch + i;
            for (int k = 0; k < nl; k++) a.pnew[k * g.fs + o] = pb[k];
            const bool corner = !g.cyclic && (i == 0 || i == g.nxp - 1);
            if (corner) {
                // corners take the x-difference term (vorsubs.F90:83-84 overwrite :79-80)
                double pbx[MAXL];
                pval(i == 0 ? 1 : g.nxp - 2, 2, gj, pbx);
                qbdy(i, gj, 2, pb, pbx);
            } else {
                pval(i, selin, gin, pin);
                qbdy(i, gj, 0, pb, pin);
            }
        }
    }
}

int inv_block_count(const Spec& s) { return (s.nyi + 1) / 2; }

void launch_inv(const InvArgs& a, hipStream_t st) {
    const int nm = (a.raw_mode >= 0) ? 1 : a.g.nl;
    const size_t lds = sizeof(double2) * 2 * (size_t)a.s.L + sizeof(double) * 2 * (size_t)nm * a.g.pitch;
    hipLaunchKernelGGL(k_inv, dim3(inv_block_count(a.s)), dim3(256), lds, st, a);
}

// ------------------------------------------------------------------ small kernels

// suppress_comp_mode (driver.F90:174-189)
__global__ void k_suppress(SuppArgs a) {
    const Geo& g = a.g;
    const long long n = (long long)g.nl * g.nyp * g.pitch;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        a.q[t] = 0.5 * (a.q[t] + a.qm[t]);
        a.p[t] = 0.5 * (a.p[t] + a.pm[t]);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        for (int l = 0; l < g.nl - 1; l++) a.c[C_DPI + l] = 0.5 * (a.c[C_DPI + l] + a.c[C_DPIP + l]);
        if (g.cyclic)
            for (int k = 0; k < g.nl; k++) {
                a.c[C_CS + k] = 0.5 * (a.c[C_CS + k] + a.c[C_CSP + k]);
                a.c[C_CN + k] = 0.5 * (a.c[C_CN + k] + a.c[C_CNP + k]);
            }
    }
}

void launch_suppress(const SuppArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_suppress, dim3(1024), dim3(256), 0, st, a);
}

// qcomp (vorsubs.F90:27-56): q = Del^2_bc(p)/f0 (+ddyn at k_topo) + beta*y - f0*A p
__global__ void k_qcomp(QcompArgs a) {
    const Geo& g = a.g;
    const long long npts = (long long)g.nyp * g.nxp;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < npts; t += (long long)gridDim.x * blockDim.x) {
        const int gj = (int)(t / g.nxp), gi = (int)(t - (long long)gj * g.nxp);
        const long long o = (long long)gj * g.pitch + gi;
        double pv[MAXL];
        for (int l = 0; l < g.nl; l++) pv[l] = a.p[l * g.fs + o];
        const int il = g.cyclic ? wrapx(gi - 1, g) : gi - 1;
        const int ir = g.cyclic ? wrapx(gi + 1, g) : gi + 1;
        for (int k = 0; k < g.nl; k++) {
            const double* f = a.p + k * g.fs;
            const double C = pv[k];
            double dx2, dy2;
            if (gj == 0 || gj == g.nyp - 1) dx2 = 0.0;
            else if (g.cyclic || (gi >= 1 && gi <= g.nxp - 2)) dx2 = g.fx * ((f[(long long)gj * g.pitch + il] + f[(long long)gj * g.pitch + ir]) - 2.0 * C);
            else if (gi == 0) dx2 = g.cxl * (f[(long long)gj * g.pitch + 1] - C);
            else dx2 = g.cxr * (C - f[(long long)gj * g.pitch + g.nxp - 2]);
            if (!g.cyclic && (gi == 0 || gi == g.nxp - 1)) dy2 = 0.0;
            else if (gj >= 1 && gj <= g.nyp - 2) dy2 = g.fy * ((f[(long long)(gj - 1) * g.pitch + gi] + f[(long long)(gj + 1) * g.pitch + gi]) - 2.0 * C);
            else if (gj == 0) dy2 = g.cyl * (f[(long long)g.pitch + gi] - C);
            else dy2 = g.cyr * (C - f[(long long)(g.nyp - 2) * g.pitch + gi]);
            double qv = (dx2 + dy2) / g.fnot;
            if (k == g.kt && a.ddyn) qv = qv + a.ddyn[o];
            qv = qv + g.beta * a.yprel[gj];
            double ap = 0.0;
            for (int l = 0; l < g.nl; l++) ap += a.mx.amat[k * MAXL + l] * pv[l];
            qv = qv - g.fnot * ap;
            a.q[k * g.fs + o] = qv;
        }
    }
}

void launch_qcomp(const QcompArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_qcomp, dim3(512), dim3(256), 0, st, a);
}

// Allow the LDS-staged transforms more than the default 64 KB of dynamic LDS.
void set_kernel_attrs(size_t fwd_lds, size_t inv_lds) {
    hipFuncSetAttribute((const void*)k_fwd, hipFuncAttributeMaxDynamicSharedMemorySize, (int)fwd_lds);
    hipFuncSetAttribute((const void*)k_inv, hipFuncAttributeMaxDynamicSharedMemorySize, (int)inv_lds);
}

}  // namespace oqg

// oqg_device.h -- device helpers shared by the QG step kernels: stencils on LDS
// tiles (numerics.F90), deterministic block reductions and the LDS FFT.
#pragma once

#include "oqg_internal.h"

namespace oqg {

constexpr int FFT_GROUP = 256;   // threads cooperating on one transform
constexpr int FFT_MAXB = 4;      // most butterflies per thread per stage (L/R <= 1024)

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

// del2_P_bc at one point (numerics.F90:410-468): dP2dx2_bc + dP2dy2_bc with the
// mixed boundary condition.  s holds the field on an LDS tile of width W with
// the point at (r,c); (gi,gj) is its global position.  In a channel the tile
// holds the x-periodic neighbours.
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

// Arakawa J(q,p) (numerics.F90:361-372) at tile point (r,c), times the factor of :361.
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

// Deterministic block sum of NV values per thread: a fixed butterfly inside
// each wavefront, then the wavefront sums in wavefront order.  Every thread
// returns the same value.  red holds NV*16 doubles; blockDim.x <= 1024.
template <int NV>
__device__ void block_sum(double (&v)[NV], double* red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_xor(v[i], off);
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < NV; i++) red[i * 16 + w] = v[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; i++) {
        double s = red[i * 16];
        for (int k = 1; k < nw; k++) s += red[i * 16 + k];
        v[i] = s;
    }
    __syncthreads();
}

// ---------------------------------------------------------------- FFT in LDS
// Stockham autosort (Govindaraju et al. 2008) with mixed radix 2/3/4/5, one
// complex buffer of length L per transform; each of the FFT_GROUP threads of a
// group holds up to FFT_MAXB butterflies in registers across the barrier that
// separates a stage's reads from its writes.  tw[t] = exp(-2 pi i t / L).

template <int R>
__device__ __forceinline__ void dft_small(double2 (&v)[R], const double2* tw, int L, int sign) {
    if constexpr (R == 2) {
        const double2 a = v[0], b = v[1];
        v[0] = make_double2(a.x + b.x, a.y + b.y);
        v[1] = make_double2(a.x - b.x, a.y - b.y);
    } else if constexpr (R == 4) {
        const double2 a0 = make_double2(v[0].x + v[2].x, v[0].y + v[2].y);
        const double2 a1 = make_double2(v[0].x - v[2].x, v[0].y - v[2].y);
        const double2 b0 = make_double2(v[1].x + v[3].x, v[1].y + v[3].y);
        const double2 b1 = make_double2(v[1].x - v[3].x, v[1].y - v[3].y);
        const double2 j = (sign < 0) ? make_double2(b1.y, -b1.x) : make_double2(-b1.y, b1.x);
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

template <int R, int MB>
__device__ void fft_stage(double2* buf, int L, int Ns, const double2* __restrict__ tw, int sign, int gt) {
    const int LR = L / R;
    const int tstep = L / (Ns * R);
    double2 v[MB][R];
#pragma unroll
    for (int b = 0; b < MB; b++) {
        const int t = gt + b * FFT_GROUP;
        if (t < LR) {
            const int k = t % Ns;
#pragma unroll
            for (int r = 0; r < R; r++) v[b][r] = buf[t + r * LR];
            if (k != 0) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    double2 w = tw[r * k * tstep];
                    if (sign > 0) w.y = -w.y;
                    v[b][r] = cmul(v[b][r], w);
                }
            }
            dft_small<R>(v[b], tw, L, sign);
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MB; b++) {
        const int t = gt + b * FFT_GROUP;
        if (t < LR) {
            const int k = t % Ns;
            const int base = (t / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) buf[base + r * Ns] = v[b][r];
        }
    }
    __syncthreads();
}

// In-place transform of buf (length s.L); gt = thread index within its group;
// MB >= the most butterflies a thread owns in any stage.  Every thread of the
// block must call this (it contains block barriers).
template <int MB>
__device__ void fft_lds(double2* buf, const Spec& s, const double2* tw, int sign, int gt) {
    int Ns = 1;
    for (int st = 0; st < s.nstages; st++) {
        const int R = s.radix[st];
        switch (R) {
            case 4: fft_stage<4, MB>(buf, s.L, Ns, tw, sign, gt); break;
            case 2: fft_stage<2, MB>(buf, s.L, Ns, tw, sign, gt); break;
            case 3: fft_stage<3, MB>(buf, s.L, Ns, tw, sign, gt); break;
            default: fft_stage<5, MB>(buf, s.L, Ns, tw, sign, gt); break;
        }
        Ns *= R;
    }
}

// Butterflies per thread needed by the plan of s (1, 2 or 4).
inline int fft_mb(const Spec& s) {
    int mx = 1;
    for (int st = 0; st < s.nstages; st++) {
        const int b = (s.L / s.radix[st] + FFT_GROUP - 1) / FFT_GROUP;
        if (b > mx) mx = b;
    }
    return mx <= 1 ? 1 : (mx <= 2 ? 2 : 4);
}

}  // namespace oqg

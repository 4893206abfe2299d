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

// Offset of global (gi,gj) in a pitch-layout layer and whether it lies in the
// basin (cyclic wrap in x).  Out-of-basin points get offset 0, so a load from
// the returned offset is always legal: callers load unconditionally and zero
// the value afterwards, which keeps a block's loads free of branches and
// lets them all issue before any is waited on.
__device__ __forceinline__ int field_offset(int gi, int gj, const Geo& g, bool& ok) {
    if (g.cyclic) gi = wrapx(gi, g);
    ok = gj >= 0 && gj < g.nyp && gi >= 0 && gi < g.nxp;
    return ok ? gj * g.pitch + gi : 0;
}

// del2_P_bc at one point (numerics.F90:410-468): dP2dx2_bc + dP2dy2_bc with the
// mixed boundary condition.  s holds the field on an LDS tile of width W with
// the point at (r,c); (gi,gj) is its global position in the basin.  In a
// channel the tile holds the x-periodic neighbours.  Branch-free: the four
// neighbours are always in the tile, every case of the reference is evaluated
// as written there and the applicable one is selected.
__device__ __forceinline__ double del2_lds(const double* s, int W, int r, int c, int gi, int gj,
                                           const Geo& g) {
    const int o = r * W + c;
    const double C = s[o], Wv = s[o - 1], Ev = s[o + 1], Sv = s[o - W], Nv = s[o + W];
    const bool xin = g.cyclic || (gi >= 1 && gi <= g.nxp - 2);
    const bool yin = gj >= 1 && gj <= g.nyp - 2;
    const double dxi = g.fx * ((Wv + Ev) - 2.0 * C);
    const double dxl = g.cxl * (Ev - C), dxr = g.cxr * (C - Wv);
    const double dyi = g.fy * ((Sv + Nv) - 2.0 * C);
    const double dyl = g.cyl * (Nv - C), dyr = g.cyr * (C - Sv);
    const double dx2 = !yin ? 0.0 : (xin ? dxi : (gi == 0 ? dxl : dxr));   // dx2 = 0 on rows 1, nyp
    const double dy2 = !xin ? 0.0 : (yin ? dyi : (gj == 0 ? dyl : dyr));   // dy2 = 0 on box columns 1, nxp
    return dx2 + dy2;
}

// Arakawa J(q,p) (numerics.F90:361-372) at tile point (r,c), times the factor
// of :361, with dPdx_2 / dPdy_2 (numerics.F90:311-338) zero on the boundary
// columns (box) and rows.  Branch-free like del2_lds: the 3x3 neighbourhoods
// are read and the zero cases selected, with the reference's operations.
__device__ __forceinline__ double jac_lds(const double* q, const double* p, int W, int r, int c, int gi,
                                          int gj, const Geo& g) {
    const bool xin = g.cyclic || (gi >= 1 && gi <= g.nxp - 2);
    const bool yin = gj >= 1 && gj <= g.nyp - 2;
    const int o = r * W + c;
    const double qW = q[o - 1], qE = q[o + 1], qS = q[o - W], qN = q[o + W];
    const double pW = p[o - 1], pE = p[o + 1], pS = p[o - W], pN = p[o + W];
    const double qNE = q[o + W + 1], qNW = q[o + W - 1], qSE = q[o - W + 1], qSW = q[o - W - 1];
    const double pNE = p[o + W + 1], pNW = p[o + W - 1], pSE = p[o - W + 1], pSW = p[o - W - 1];
    const double dqx = xin ? qE - qW : 0.0, dqy = yin ? qN - qS : 0.0;
    const double dpx = xin ? pE - pW : 0.0, dpy = yin ? pN - pS : 0.0;
    double t = dqx * dpy - dqy * dpx;
    // dY at (r, c+-1), dX at (r+-1, c): the point's own row/column decides
    const double dpyE = yin ? pNE - pSE : 0.0, dpyW = yin ? pNW - pSW : 0.0;
    const double dpxN = xin ? pNE - pNW : 0.0, dpxS = xin ? pSE - pSW : 0.0;
    const double dqxN = xin ? qNE - qNW : 0.0, dqxS = xin ? qSE - qSW : 0.0;
    const double dqyE = yin ? qNE - qSE : 0.0, dqyW = yin ? qNW - qSW : 0.0;
    double v;
    v = xin ? qE * dpyE - qW * dpyW : 0.0;
    t = t + v;
    v = yin ? qN * dpxN - qS * dpxS : 0.0;
    t = t - v;
    v = yin ? pN * dqxN - pS * dqxS : 0.0;
    t = t + v;
    v = xin ? pE * dqyE - pW * dqyW : 0.0;
    t = t - v;
    return g.jfac * t;
}

// Global -> LDS copies split in two halves so that a block issues all of its
// global loads before it waits for any: stage_load reads this thread's first
// NPT words of `src` into registers (straight-line, predicated), stage_store
// later writes them to LDS and copies any remainder (n > NPT * blockDim) with a
// strided loop.  The caller's next barrier publishes the copy.
template <int NPT>
struct Staged {
    unsigned long long w[NPT];
};

template <int NPT>
__device__ __forceinline__ void stage_load(Staged<NPT>& r, const void* src, int nwords) {
    const unsigned long long* s = static_cast<const unsigned long long*>(src);
#pragma unroll
    for (int p = 0; p < NPT; p++) {
        const int idx = threadIdx.x + p * blockDim.x;
        r.w[p] = (idx < nwords) ? s[idx] : 0ull;
    }
}

template <int NPT>
__device__ __forceinline__ void stage_store(const Staged<NPT>& r, void* dst, const void* src, int nwords) {
    unsigned long long* d = static_cast<unsigned long long*>(dst);
#pragma unroll
    for (int p = 0; p < NPT; p++) {
        const int idx = threadIdx.x + p * blockDim.x;
        if (idx < nwords) d[idx] = r.w[p];
    }
    const unsigned long long* s = static_cast<const unsigned long long*>(src);
    for (int idx = threadIdx.x + NPT * blockDim.x; idx < nwords; idx += blockDim.x) d[idx] = s[idx];
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
// separates a stage's reads from its writes.
//
// Twiddles come from an LDS table: when 4 | L it holds the first quadrant only,
// tw[r] = exp(-2 pi i r / L) for r < L/4, and exp(-2 pi i (q L/4 + r) / L) is
// tw[r] times (-i)^q, an exact swap and sign change (the host table is built
// by the same quadrant reduction, so the values are identical); otherwise it
// holds all L entries.

// Entries of the LDS twiddle table for a transform of length L.
__host__ __device__ inline int twiddle_table_len(int L) { return (L % 4 == 0) ? L / 4 : L; }

// Branch-free: when 4 does not divide L the quadrant width is L itself and q = 0.
__device__ __forceinline__ double2 twid(const double2* tw, int t, int L) {
    const int L4 = (L & 3) ? L : (L >> 2);
    const int q = (t >= L4) + (t >= 2 * L4) + (t >= 3 * L4);
    const double2 w = tw[t - q * L4];
    // q = 0: (x, y); 1: (y, -x); 2: (-x, -y); 3: (-y, x)
    const double u = (q & 1) ? w.y : w.x, v = (q & 1) ? w.x : w.y;
    return make_double2((q >= 2) ? -u : u, (q == 1 || q == 2) ? -v : v);
}

// 8-byte words of the LDS twiddle table of a length-L transform.
__host__ __device__ inline int twiddle_words(int L) { return 2 * twiddle_table_len(L); }

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
                const double2 w = twid(tw, ((r * q) % R) * pstep, L);
                const double c = w.x, s = sg * w.y;
                sr += in[r].x * c - in[r].y * s;
                si += in[r].x * s + in[r].y * c;
            }
            v[q] = make_double2(sr, si);
        }
    }
}

template <int R, int MB>
__device__ __forceinline__ void fft_stage(double2* buf, int L, int Ns, const double2* tw, int sign, int gt) {
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
                    double2 w = twid(tw, r * k * tstep, L);
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

// In-place transform of buf (length s.L) with the LDS twiddle table tw;
// gt = thread index within its group; MB >= the most butterflies a thread owns
// in any stage.  Every thread of the block must call this (it contains block
// barriers).
template <int MB>
__device__ __forceinline__ void fft_lds(double2* buf, const Spec& s, const double2* tw, int sign, int gt) {
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

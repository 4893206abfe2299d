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
// NPT elements (of type T) of `src` into registers (straight-line,
// predicated), stage_store later writes them to LDS and copies any remainder
// (n > NPT * blockDim) with a strided loop.  The caller's next barrier
// publishes the copy.
template <int NPT, class T = unsigned long long>
struct Staged {
    T w[NPT];
};

template <int NPT, class T>
__device__ __forceinline__ void stage_load(Staged<NPT, T>& r, const void* src, int n) {
    const T* s = static_cast<const T*>(src);
#pragma clang loop unroll(full)
    for (int p = 0; p < NPT; p++) {
        const int idx = threadIdx.x + p * blockDim.x;
        r.w[p] = s[(idx < n) ? idx : 0];
    }
}

template <int NPT, class T>
__device__ __forceinline__ void stage_store(const Staged<NPT, T>& r, void* dst, const void* src, int n) {
    T* d = static_cast<T*>(dst);
#pragma clang loop unroll(full)   // constant indices keep r in registers
    for (int p = 0; p < NPT; p++) {
        const int idx = threadIdx.x + p * blockDim.x;
        if (idx < n) d[idx] = r.w[p];
    }
    const T* s = static_cast<const T*>(src);
    for (int idx = threadIdx.x + NPT * blockDim.x; idx < n; idx += blockDim.x) d[idx] = s[idx];
}

// x from another lane through a DPP (data-parallel primitive) move of both
// 32-bit halves; CTRL is the gfx9 dpp_ctrl code.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double x) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double read_lane(double x, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
    return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wavefront, the same value in every lane and the
// same order of additions in every lane (each pairwise step adds the partner's
// partial sum, and addition commutes): lane pairs, quads (quad_perm), half
// rows (row_half_mirror), rows of 16 (row_mirror), then the four row sums.
// Pure VALU cross-lane moves: no LDS traffic.
__device__ __forceinline__ double wave_sum(double x) {
    x = x + dpp_move<0xB1>(x);    // quad_perm [1,0,3,2]
    x = x + dpp_move<0x4E>(x);    // quad_perm [2,3,0,1]
    x = x + dpp_move<0x141>(x);   // row_half_mirror
    x = x + dpp_move<0x140>(x);   // row_mirror
    return (read_lane(x, 0) + read_lane(x, 16)) + (read_lane(x, 32) + read_lane(x, 48));
}

// Maximum and minimum over the 64 lanes, by the same moves (fmax/fmin drop a
// NaN operand; callers count non-finite values separately).
__device__ __forceinline__ double wave_max(double x) {
    x = fmax(x, dpp_move<0xB1>(x));
    x = fmax(x, dpp_move<0x4E>(x));
    x = fmax(x, dpp_move<0x141>(x));
    x = fmax(x, dpp_move<0x140>(x));
    return fmax(fmax(read_lane(x, 0), read_lane(x, 16)), fmax(read_lane(x, 32), read_lane(x, 48)));
}

__device__ __forceinline__ double wave_min(double x) {
    x = fmin(x, dpp_move<0xB1>(x));
    x = fmin(x, dpp_move<0x4E>(x));
    x = fmin(x, dpp_move<0x141>(x));
    x = fmin(x, dpp_move<0x140>(x));
    return fmin(fmin(read_lane(x, 0), read_lane(x, 16)), fmin(read_lane(x, 32), read_lane(x, 48)));
}

// Deterministic block sum of NV values per thread: wave_sum inside each
// wavefront, then the wavefront sums in wavefront order.  Every thread
// returns the same value.  red holds NV*16 doubles; blockDim.x <= 1024.
// In two halves: block_sum_post (the wavefront sums into red, no barrier) and
// block_sum_finish (barrier, then the sum of the slots), so that a kernel can
// release the per-thread values before other work and finish later.
template <int NV>
__device__ __forceinline__ void block_sum_post(const double (&v)[NV], double* red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double s[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) s[i] = wave_sum(v[i]);
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < NV; i++) red[i * 16 + w] = s[i];
}

template <int NV>
__device__ __forceinline__ void block_sum_finish(double (&v)[NV], double* red) {
    const int nw = blockDim.x >> 6;
    __syncthreads();
    // all 16 slots are read at once (the ones of absent wavefronts are
    // replaced by exact zeros), then summed in wavefront order
#pragma unroll
    for (int i = 0; i < NV; i++) {
        double part[16];
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const double r = red[i * 16 + k];
            part[k] = (k < nw) ? r : 0.0;
        }
        double s = part[0];
#pragma unroll
        for (int k = 1; k < 16; k++) s += part[k];
        v[i] = s;
    }
    __syncthreads();
}

template <int NV>
__device__ void block_sum(double (&v)[NV], double* red) {
    block_sum_post<NV>(v, red);
    block_sum_finish<NV>(v, red);
}

// ---------------------------------------------------------------- FFT in LDS
// Stockham autosort (Govindaraju et al. 2008) with mixed radix 2/3/4/5 (and
// 8/16, and 7/11/13 in the PR instantiations), one
// complex buffer of length L per transform; each of the FFT_GROUP threads of a
// group holds up to FFT_MAXB butterflies in registers across the barrier that
// separates a stage's reads from its writes.  Twiddles come from an LDS copy of
// the plan's table: with W = exp(-2 pi i / L), stage s (Ns = R_0...R_{s-1},
// tstep = L/(Ns R_s)) holds W^(r k tstep) at off_s + (r-1) Ns + k for
// r = 1..R_s-1, k < Ns (off_s = sum of (R_i - 1) N_i over earlier stages;
// L - 1 entries in all), so a wave reads consecutive entries; then W^(m L/3)
// at L-1+(m-1), m = 1, 2, and W^(m L/5) at L+1+(m-1), m = 1..4, for the
// radix-3 and radix-5 butterflies.  SIGN = -1 is the forward transform, +1 the
// backward one (conjugate twiddles).
//
// Each butterfly index t stays with its thread through all stages; the stage
// with Ns = R_0...R_{s-1} needs k = t mod Ns and j = t / Ns, which are carried
// from stage to stage (k += Ns (j mod R), j /= R with R a compile-time radix)
// instead of being recomputed with runtime divisions.


// exp(-2 pi i m / 16) for the forward transform (SIGN < 0), its conjugate for
// the backward one, m = 0..9: the twiddles inside a radix-16 butterfly.
__device__ __forceinline__ double2 w16(int m, int sign) {
    constexpr double C[10] = {1.0, 0.92387953251128674, 0.70710678118654752, 0.38268343236508977, 0.0,
                              -0.38268343236508977, -0.70710678118654752, -0.92387953251128674, -1.0,
                              -0.92387953251128674};
    constexpr double S[10] = {0.0, 0.38268343236508977, 0.70710678118654752, 0.92387953251128674, 1.0,
                              0.92387953251128674, 0.70710678118654752, 0.38268343236508977, 0.0,
                              -0.38268343236508977};
    return make_double2(C[m], sign < 0 ? -S[m] : S[m]);
}

template <int R, int SIGN>
__device__ __forceinline__ void dft_small(double2 (&v)[R], const double2* tw, int L);

// Radix 8 as two radix-4 DFTs of the even and odd inputs and one twiddle
// layer: X[k] = E[k] + W8^k O[k], X[k+4] = E[k] - W8^k O[k].
template <int SIGN>
__device__ __forceinline__ void dft8(double2 (&v)[8]) {
    double2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft_small<4, SIGN>(e, nullptr, 4);
    dft_small<4, SIGN>(o, nullptr, 4);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double2 t = (k == 0) ? o[0] : cmul(o[k], w16(2 * k, SIGN));
        v[k] = make_double2(e[k].x + t.x, e[k].y + t.y);
        v[k + 4] = make_double2(e[k].x - t.x, e[k].y - t.y);
    }
}

// Radix 16 as 4 x 4: Y_n2 = DFT4 of x[4 n1 + n2] over n1, Z_n2[k1] =
// W16^(n2 k1) Y_n2[k1], X[k1 + 4 k2] = DFT4 of Z_n2[k1] over n2.
template <int SIGN>
__device__ __forceinline__ void dft16(double2 (&v)[16]) {
    double2 y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) {
#pragma unroll
        for (int n1 = 0; n1 < 4; n1++) y[n2][n1] = v[4 * n1 + n2];
        dft_small<4, SIGN>(y[n2], nullptr, 4);
#pragma unroll
        for (int k1 = 1; k1 < 4; k1++)
            if (n2 > 0) y[n2][k1] = cmul(y[n2][k1], w16(n2 * k1, SIGN));
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
        double2 z[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
        dft_small<4, SIGN>(z, nullptr, 4);
#pragma unroll
        for (int k2 = 0; k2 < 4; k2++) v[k1 + 4 * k2] = z[k2];
    }
}

template <int R, int SIGN>
__device__ __forceinline__ void dft_small(double2 (&v)[R], const double2* tw, int L) {
    if constexpr (R == 8) {
        dft8<SIGN>(v);
    } else if constexpr (R == 16) {
        dft16<SIGN>(v);
    } else if constexpr (R == 2) {
        const double2 a = v[0], b = v[1];
        v[0] = make_double2(a.x + b.x, a.y + b.y);
        v[1] = make_double2(a.x - b.x, a.y - b.y);
    } else if constexpr (R == 4) {
        const double2 a0 = make_double2(v[0].x + v[2].x, v[0].y + v[2].y);
        const double2 a1 = make_double2(v[0].x - v[2].x, v[0].y - v[2].y);
        const double2 b0 = make_double2(v[1].x + v[3].x, v[1].y + v[3].y);
        const double2 b1 = make_double2(v[1].x - v[3].x, v[1].y - v[3].y);
        const double2 j = (SIGN < 0) ? make_double2(b1.y, -b1.x) : make_double2(-b1.y, b1.x);
        v[0] = make_double2(a0.x + b0.x, a0.y + b0.y);
        v[1] = make_double2(a1.x + j.x, a1.y + j.y);
        v[2] = make_double2(a0.x - b0.x, a0.y - b0.y);
        v[3] = make_double2(a1.x - j.x, a1.y - j.y);
    } else {
        // radix 3 and 5: tw holds the roots W^(m L/R) at m-1, m = 1..R-1
        double2 in[R];
#pragma unroll
        for (int r = 0; r < R; r++) in[r] = v[r];
        constexpr double sg = (SIGN > 0) ? -1.0 : 1.0;
#pragma unroll
        for (int q = 0; q < R; q++) {
            double sr = in[0].x, si = in[0].y;
#pragma unroll
            for (int r = 1; r < R; r++) {
                const int mq = (r * q) % R;
                const double2 w = (mq == 0) ? make_double2(1.0, 0.0) : tw[mq - 1];
                const double c = w.x, s = sg * w.y;
                sr += in[r].x * c - in[r].y * s;
                si += in[r].x * s + in[r].y * c;
            }
            v[q] = make_double2(sr, si);
        }
    }
}

// Workgroups are dealt to the NXCD dies round-robin by linear id, and each die
// has its own L2.  xcd_contig gives workgroup w of n a position such that the
// workgroups of every die take one contiguous run of positions (die x: ids
// x, x + NXCD, ...; run start x*q + min(x, r) with n = q*NXCD + r).
__device__ __forceinline__ int xcd_contig(int w, int n) {
    const int q = n / NXCD, r = n % NXCD, x = w % NXCD, i = w / NXCD;
    return x * q + (x < r ? x : r) + i;
}

// This workgroup's (x, y, z) block, remapped so that neighbouring blocks,
// which share halo and row lines, mostly run on one die: when a z slice
// (layer or mode) holds a multiple of NXCD blocks, every die takes one
// contiguous run of each slice (the same band of every layer).  Other grids
// keep the hardware order: runs over the whole grid measured slower for
// k_vort at 513^2 (1 us) and for k_adj_vort at 257^2 (0.4 us).
__device__ __forceinline__ void xcd_block(int& bx, int& by, int& bz) {
    const int gx = gridDim.x, gy = gridDim.y, slice = gx * gy;
    bz = blockIdx.z;
    if (slice % NXCD != 0) {
        bx = blockIdx.x;
        by = blockIdx.y;
        return;
    }
    const int t = xcd_contig(blockIdx.x + gx * blockIdx.y, slice);
    by = t / gx;
    bx = t - by * gx;
}

// LDS slot of element x of a transform buffer: one 16-byte slot of padding
// after every 16, so that the strided writes of a first stage (stride R slots,
// R <= 16) and the contiguous reads of every stage fall in distinct banks.
// A buffer of length L takes zpad_len(L) slots.
__host__ __device__ constexpr int zpad(int x) { return x + (x >> 4); }
__host__ __device__ constexpr int zpad_len(int L) { return L + (L >> 4); }

// One radix-R stage.  LR = L/R butterflies; tws = the stage's twiddles
// (W^(r k tstep) at (r-1) Ns + k), roots = the radix-3/5 roots; kk/jj carry
// k = t mod Ns and j = t / Ns of this thread's butterflies t = gt + b*FFT_GROUP
// and are advanced to the next stage on return.
template <int R, int MB, int SIGN>
__device__ __forceinline__ void fft_stage(double2* buf, int LR, int Ns, const double2* tws, const double2* roots,
                                          int gt, int (&kk)[MB], int (&jj)[MB]) {
    double2 v[MB][R];
#pragma unroll
    for (int b = 0; b < MB; b++) {
        const int t = gt + b * FFT_GROUP;
        if (t < LR) {
            const int k = kk[b];
#pragma unroll
            for (int r = 0; r < R; r++) v[b][r] = buf[zpad(t + r * LR)];
#pragma unroll
            for (int r = 1; r < R; r++) {
                double2 w = tws[(r - 1) * Ns + k];   // k = 0 gives exactly (1, 0)
                if (SIGN > 0) w.y = -w.y;
                v[b][r] = cmul(v[b][r], w);
            }
            dft_small<R, SIGN>(v[b], roots, R);
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MB; b++) {
        const int t = gt + b * FFT_GROUP;
        const int j = jj[b], k = kk[b];
        if (t < LR) {
            const int base = j * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) buf[zpad(base + r * Ns)] = v[b][r];
        }
        kk[b] = k + Ns * (j % R);
        jj[b] = j / R;
    }
    __syncthreads();
}

// In-place transform of buf (length L, stored through zpad) with the LDS copy
// tw of the twiddle table of the plan radix[0..nst), whose radices are all
// <= MAXR (a larger one is never
// instantiated, so it costs the kernel no registers); gt = thread index within
// its group; MB >= the most butterflies a thread owns in any stage.  Every
// thread of the block must call this (it contains block barriers).
template <int MB, int SIGN, int MAXR, bool PR = false>
__device__ __forceinline__ void fft_lds(double2* buf, int L, int nst, const int* radix, const double2* tw, int gt) {
    int kk[MB], jj[MB];
#pragma unroll
    for (int b = 0; b < MB; b++) {
        kk[b] = 0;
        jj[b] = gt + b * FFT_GROUP;
    }
    int Ns = 1;
    const double2* tws = tw;   // this stage's twiddles
    for (int st = 0; st < nst; st++) {
        const int R = radix[st];
        const int LR = L / R;   // uniform (scalar) arithmetic
        if constexpr (MAXR >= 16) {
            if (R == 16) {
                fft_stage<16, MB, SIGN>(buf, LR, Ns, tws, nullptr, gt, kk, jj);
                tws += 15 * Ns;
                Ns *= R;
                continue;
            }
        }
        if constexpr (MAXR >= 8) {
            if (R == 8) {
                fft_stage<8, MB, SIGN>(buf, LR, Ns, tws, nullptr, gt, kk, jj);
                tws += 7 * Ns;
                Ns *= R;
                continue;
            }
        }
        if constexpr (PR) {   // the odd primes 7, 11, 13 (plans with prime = 1 only)
            if (R == 7 || R == 11 || R == 13) {
                if (R == 7) fft_stage<7, MB, SIGN>(buf, LR, Ns, tws, tw + twiddle_roots7(L), gt, kk, jj);
                else if (R == 11) fft_stage<11, MB, SIGN>(buf, LR, Ns, tws, tw + twiddle_roots11(L), gt, kk, jj);
                else fft_stage<13, MB, SIGN>(buf, LR, Ns, tws, tw + twiddle_roots13(L), gt, kk, jj);
                tws += (R - 1) * Ns;
                Ns *= R;
                continue;
            }
        }
        switch (R) {
            case 4: fft_stage<4, MB, SIGN>(buf, LR, Ns, tws, nullptr, gt, kk, jj); break;
            case 2: fft_stage<2, MB, SIGN>(buf, LR, Ns, tws, nullptr, gt, kk, jj); break;
            case 3: fft_stage<3, MB, SIGN>(buf, LR, Ns, tws, tw + twiddle_roots3(L), gt, kk, jj); break;
            default: fft_stage<5, MB, SIGN>(buf, LR, Ns, tws, tw + twiddle_roots5(L), gt, kk, jj); break;
        }
        tws += (R - 1) * Ns;
        Ns *= R;
    }
}

// ---------------------------------------------------------------- FFT per wavefront
// The same Stockham stages, twiddle layout and buffer padding as fft_lds, with
// one wavefront per transform: lane l owns the butterflies t = l + 64 b of
// every stage, b < ceil(P / R) for a plan whose transform has at most 64 P
// points.  A wavefront's LDS reads and writes are performed in the order it
// issues them, so a compiler-level fence separates a stage's reads from its
// writes and its writes from the next stage's reads: the transforms of the
// wavefronts of a block need no block barrier, and a block may give other
// wavefronts other work meanwhile.

// a*b with fused multiply-adds (two roundings per component): the wavefront
// transforms' twiddle products, where it saves a third of the instructions.
__device__ __forceinline__ double2 cmul_fma(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int R, int P, int SIGN>
__device__ __forceinline__ void fftw_stage(double2* buf, int L, int Ns, const double2* tws, const double2* roots,
                                           int lane) {
    constexpr int MB = (P + R - 1) / R;
    // The stage's addresses are formed here from an opaque copy of L: hoisted
    // out of the stage loop for every radix the plan might use, they held
    // about a hundred registers and the stages' LDS reads were then issued one
    // at a time.
    int Lx = L;
    __asm__ volatile("" : "+s"(Lx));
    const int LR = Lx / R;
    const double rNs = 1.0 / (double)Ns;   // t < 2^13: (t + 1/2)/Ns rounds to the right floor
    double2 v[MB][R], w[MB][R > 1 ? R - 1 : 1];
    int kk[MB], jj[MB];
    const bool twiddled = Ns > 1;   // the first stage's twiddles are all exactly 1
#pragma unroll
    for (int b = 0; b < MB; b++) {
        const int t = lane + 64 * b;
        jj[b] = (int)(((double)t + 0.5) * rNs);
        kk[b] = t - jj[b] * Ns;
        if (t < LR) {
            if (twiddled)
#pragma unroll
                for (int r = 1; r < R; r++) w[b][r - 1] = tws[(r - 1) * Ns + kk[b]];
#pragma unroll
            for (int r = 0; r < R; r++) v[b][r] = buf[zpad(t + r * LR)];
        }
    }
#pragma unroll
    for (int b = 0; b < MB; b++) {
        const int t = lane + 64 * b;
        if (t < LR) {
            if (twiddled)
#pragma unroll
                for (int r = 1; r < R; r++) {
                    double2 wr = w[b][r - 1];
                    if (SIGN > 0) wr.y = -wr.y;
                    v[b][r] = cmul_fma(v[b][r], wr);
                }
            dft_small<R, SIGN>(v[b], roots, R);
        }
    }
    wave_lds_sync();
#pragma unroll
    for (int b = 0; b < MB; b++) {
        const int t = lane + 64 * b;
        if (t < LR) {
            const int base = jj[b] * Ns * R + kk[b];
#pragma unroll
            for (int r = 0; r < R; r++) buf[zpad(base + r * Ns)] = v[b][r];
        }
    }
    wave_lds_sync();
}

// In-place transform of buf (length L <= 64 P, stored through zpad) by the
// calling wavefront, with the LDS copy tw of the plan's twiddle table.
template <int P, int SIGN>
__device__ __forceinline__ void fft_wave(double2* buf, int L, int nst, const int* radix, const double2* tw, int lane) {
    int Ns = 1;
    const double2* tws = tw;
    for (int st = 0; st < nst; st++) {
        const int R = radix[st];
        switch (R) {
            case 16: fftw_stage<16, P, SIGN>(buf, L, Ns, tws, nullptr, lane); break;
            case 8: fftw_stage<8, P, SIGN>(buf, L, Ns, tws, nullptr, lane); break;
            case 4: fftw_stage<4, P, SIGN>(buf, L, Ns, tws, nullptr, lane); break;
            case 2: fftw_stage<2, P, SIGN>(buf, L, Ns, tws, nullptr, lane); break;
            case 3: fftw_stage<3, P, SIGN>(buf, L, Ns, tws, tw + twiddle_roots3(L), lane); break;
            default: fftw_stage<5, P, SIGN>(buf, L, Ns, tws, tw + twiddle_roots5(L), lane); break;
        }
        tws += (R - 1) * Ns;
        Ns *= R;
    }
}

// The same transform with the plan fixed at compile time, for the lengths of
// the usual basins: every stride, twiddle offset and butterfly count is a
// constant (LDS offsets fold into the instructions), so a stage costs about
// two thirds of the run-time plan's.  The plan is the one the library builds
// on the host (radices 16, 8, 4, 2, 3, 5 taken greedily), so the stage-ordered
// twiddle table is the same.
struct FixedPlan {
    int n;
    int r[16];
};
constexpr FixedPlan fixed_plan(int L) {
    FixedPlan p{0, {}};
    int m = L;
    const int rs[6] = {16, 8, 4, 2, 3, 5};
    for (int q = 0; q < 6; q++)
        while (m % rs[q] == 0) {
            p.r[p.n++] = rs[q];
            m /= rs[q];
        }
    return p;
}

template <int R, int L, int NS, int SIGN>
__device__ __forceinline__ void fftf_stage(double2* buf, const double2* tws, const double2* roots, int lane) {
    constexpr int LR = L / R, MB = (LR + 63) / 64;
    double2 v[MB][R], w[MB][R > 1 ? R - 1 : 1];
#pragma unroll
    for (int b = 0; b < MB; b++) {
        const int t = lane + 64 * b;
        if (LR % 64 == 0 || t < LR) {
            const int k = t % NS;
            if constexpr (NS > 1)
#pragma unroll
                for (int r = 1; r < R; r++) w[b][r - 1] = tws[(r - 1) * NS + k];
#pragma unroll
            for (int r = 0; r < R; r++) v[b][r] = buf[zpad(t + r * LR)];
        }
    }
#pragma unroll
    for (int b = 0; b < MB; b++) {
        const int t = lane + 64 * b;
        if (LR % 64 == 0 || t < LR) {
            if constexpr (NS > 1)
#pragma unroll
                for (int r = 1; r < R; r++) {
                    double2 wr = w[b][r - 1];
                    if (SIGN > 0) wr.y = -wr.y;
                    v[b][r] = cmul_fma(v[b][r], wr);
                }
            dft_small<R, SIGN>(v[b], roots, R);
        }
    }
    wave_lds_sync();
#pragma unroll
    for (int b = 0; b < MB; b++) {
        const int t = lane + 64 * b;
        if (LR % 64 == 0 || t < LR) {
            const int base = (t / NS) * NS * R + t % NS;
#pragma unroll
            for (int r = 0; r < R; r++) buf[zpad(base + r * NS)] = v[b][r];
        }
    }
    wave_lds_sync();
}

template <int L, int SIGN, int ST = 0, int NS = 1, int OFF = 0>
__device__ __forceinline__ void fft_wave_fixed(double2* buf, const double2* tw, int lane) {
    constexpr FixedPlan pl = fixed_plan(L);
    if constexpr (ST < pl.n) {
        constexpr int R = pl.r[ST];
        const double2* roots = (R == 3) ? tw + twiddle_roots3(L) : (R == 5 ? tw + twiddle_roots5(L) : nullptr);
        fftf_stage<R, L, NS, SIGN>(buf, tw + OFF, roots, lane);
        fft_wave_fixed<L, SIGN, ST + 1, NS * R, OFF + (R - 1) * NS>(buf, tw, lane);
    }
}

// fft_wave with the fixed plan when L is one of the lengths compiled for a
// kernel's points per lane P (L <= 64 P), else the run-time plan.
template <int P, int SIGN>
__device__ __forceinline__ void fft_wave_any(double2* buf, int L, int nst, const int* radix, const double2* tw,
                                             int lane) {
    if constexpr (P == 16) {
        if (L == 1024) return fft_wave_fixed<1024, SIGN>(buf, tw, lane);
        if (L == 768) return fft_wave_fixed<768, SIGN>(buf, tw, lane);
        if (L == 512) return fft_wave_fixed<512, SIGN>(buf, tw, lane);
    } else if constexpr (P == 4) {
        if (L == 256) return fft_wave_fixed<256, SIGN>(buf, tw, lane);
        if (L == 128) return fft_wave_fixed<128, SIGN>(buf, tw, lane);
        if (L == 96) return fft_wave_fixed<96, SIGN>(buf, tw, lane);
        if (L == 64) return fft_wave_fixed<64, SIGN>(buf, tw, lane);
    }
    fft_wave<P, SIGN>(buf, L, nst, radix, tw, lane);
}

// ---------------------------------------------------------------- ocqbdy, transposed
// Forward (box), at a ring point x with inner point x_in (vorsubs.F90:58-114):
//   q_k(x) = qbfac (p_k(x_in) - p_k(x)) - r f0 sum_l A_kl p_l(x) + r beta y + ...
// with r = 2 at the corners (the row and column terms both applied) and 1
// otherwise; a corner's inner point is along its row.
__device__ __forceinline__ bool ring_inner(const Geo& g, int xi, int xj, int& ii, int& ij) {
    const int nx = g.nxp, ny = g.nyp;
    const bool ex = (xi == 0 || xi == nx - 1), ey = (xj == 0 || xj == ny - 1);
    if (!ex && !ey) return false;
    if (ex && ey) {
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

// The transpose gathered at point a of the two outer bands (rows 0, 1,
// nyp-2, nyp-1 and columns 0, 1, nxp-2, nxp-1), added to lp[0..NL-1]:
//   lp_l(a) += sum_k lq_k(a) (-qbfac d_kl - r f0 A_kl)        (a on the ring)
//   lp_k(a) += qbfac lq_k(x)   for each ring point x whose inner point is a.
template <int NL>
__device__ __forceinline__ void bdy_gather(const Geo& g, double qbfac, const Mix& mx, const double* lq, int ai,
                                           int aj, double (&lp)[NL]) {
    const int nx = g.nxp, ny = g.nyp;
    const long long oa = (long long)aj * g.pitch + ai;
    int ii, ij;
    if (ring_inner(g, ai, aj, ii, ij)) {
        const double r = ((ai == 0 || ai == nx - 1) && (aj == 0 || aj == ny - 1)) ? 2.0 : 1.0;
#pragma unroll
        for (int k = 0; k < NL; k++) {
            const double lqk = lq[k * g.fs + oa];
#pragma unroll
            for (int l = 0; l < NL; l++) {
                double c = -r * g.fnot * mx.amat[k * MAXL + l];
                if (l == k) c -= qbfac;
                lp[l] += lqk * c;
            }
        }
    }
    const int cand[4][2] = {{ai - 1, aj}, {ai + 1, aj}, {ai, aj - 1}, {ai, aj + 1}};
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int xi = cand[c][0], xj = cand[c][1];
        if (xi < 0 || xi >= nx || xj < 0 || xj >= ny) continue;
        if (ring_inner(g, xi, xj, ii, ij) && ii == ai && ij == aj) {
            const long long ox = (long long)xj * g.pitch + xi;
#pragma unroll
            for (int k = 0; k < NL; k++) lp[k] += qbfac * lq[k * g.fs + ox];
        }
    }
}

__device__ __forceinline__ bool outer_band(const Geo& g, int i, int j) {
    return j <= 1 || j >= g.nyp - 2 || i <= 1 || i >= g.nxp - 2;
}

}  // namespace oqg

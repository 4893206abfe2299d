// oqg_wave.hip -- the wavefront-transform forms of k_fwd and k_inv for one layer
// count.  The file is compiled once per layer count (-DOQG_WAVE_NL=2..8), each
// object defining the launchers launch_*_nl<NL>; oqg_transform.hip dispatches
// to them.

#include "oqg_inv_dev.h"

#ifndef OQG_WAVE_NL
#error "compile oqg_wave.hip with -DOQG_WAVE_NL=<layer count>"
#endif

namespace oqg {

// k_fwdw and k_invw do the work of k_fwd and k_inv for a whole row pair in one
// workgroup with one wavefront per mode: each wavefront runs its mode's
// transform on its own (fft_wave, radix-16/8 plan, no block barrier inside),
// which takes the block barriers out of the transform stages and the modes'
// redundant reads of q out of k_fwd.  In k_invw one more wavefront computes
// the correction coefficients from the column sums and the constraint state
// while the others transform, and then writes the zonal boundary rows of the
// first and last row pairs.  The arithmetic per point is that of k_fwd and
// k_inv; only the transform's stage plan and the sums' order differ.

constexpr int WFPT = 4;   // points per thread whose inputs k_fwdw loads before any is used

template <int NL, int P>
__device__ __forceinline__ void fwdw_block(const FwdArgs& a, int jp) {
    extern __shared__ double2 wsm[];   // [NL][zpad_len(L)] transforms, then the twiddle table
    const Geo& g = a.g;
    const Spec& s = a.s;
    const int L = s.L, ZL = zpad_len(L);
    const int nth = (int)blockDim.x;   // 64 NL
    const int wv = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    double2* tw = wsm + (size_t)NL * ZL;
    const int r0 = 2 * jp, r1 = r0 + 1;
    const bool has1 = r1 < s.nyi;
    const int gj0 = r0 + 1, gj1 = r1 + 1;
    Staged<12, double> stw;
    stage_load(stw, a.tw, 2 * twiddle_table_len(L));
    const int first = g.cyclic ? 0 : 1, npts = s.nin;
    double q0[WFPT][NL], q1[WFPT][NL], d0[WFPT], d1[WFPT];
#pragma unroll
    for (int u = 0; u < WFPT; u++) {
        const int t = first + (int)threadIdx.x + nth * u;
        const bool in = t < first + npts;
        const long long o0 = (long long)gj0 * g.pitch + t, o1 = (long long)gj1 * g.pitch + t;
#pragma unroll
        for (int k = 0; k < NL; k++) {
            q0[u][k] = in ? a.q[k * g.fs + o0] : 0.0;
            q1[u][k] = (in && has1) ? a.q[k * g.fs + o1] : 0.0;
        }
        d0[u] = (in && a.ddyn) ? a.ddyn[o0] : 0.0;
        d1[u] = (in && has1 && a.ddyn) ? a.ddyn[o1] : 0.0;
    }
    const double by0 = g.beta * a.yprel[gj0], by1 = has1 ? g.beta * a.yprel[gj1] : 0.0;
    double cm[NL][NL];   // ctl2m(k, m), the same in every thread
#pragma unroll
    for (int k = 0; k < NL; k++)
#pragma unroll
        for (int m = 0; m < NL; m++) cm[k][m] = a.dc->mx.ctl2m[k * MAXL + m];
    // p_modes_rhs (modes.F90:259-264) of every mode at point t of both rows,
    // into the transform buffers: the odd extension in a box, the packed rows
    // in a channel
    auto put = [&](int t, const double (&qa)[NL], const double (&qb)[NL], double da, double db) {
        double la[NL], lb[NL];
#pragma unroll
        for (int k = 0; k < NL; k++) {
            la[k] = qa[k] - by0;
            lb[k] = qb[k] - by1;
            if (k == g.kt) {
                la[k] = la[k] - da;
                lb[k] = lb[k] - db;
            }
        }
#pragma unroll
        for (int m = 0; m < NL; m++) {
            double sa = 0.0, sb = 0.0;
#pragma unroll
            for (int k = 0; k < NL; k++) {
                sa += la[k] * cm[k][m];
                sb += lb[k] * cm[k][m];
            }
            const double2 w = make_double2(g.fnot * sa, has1 ? g.fnot * sb : 0.0);
            double2* z = wsm + (size_t)m * ZL;
            z[zpad(t)] = w;
            if (!g.cyclic) z[zpad(L - t)] = make_double2(-w.x, -w.y);
        }
    };
#pragma unroll
    for (int u = 0; u < WFPT; u++) {
        const int t = first + (int)threadIdx.x + nth * u;
        if (t < first + npts) put(t, q0[u], q1[u], d0[u], d1[u]);
    }
    for (int t = first + (int)threadIdx.x + nth * WFPT; t < first + npts; t += nth) {
        const long long o0 = (long long)gj0 * g.pitch + t, o1 = (long long)gj1 * g.pitch + t;
        double qa[NL], qb[NL];
#pragma unroll
        for (int k = 0; k < NL; k++) {
            qa[k] = a.q[k * g.fs + o0];
            qb[k] = has1 ? a.q[k * g.fs + o1] : 0.0;
        }
        put(t, qa, qb, a.ddyn ? a.ddyn[o0] : 0.0, (has1 && a.ddyn) ? a.ddyn[o1] : 0.0);
    }
    if (!g.cyclic && (int)threadIdx.x < NL) {
        double2* z = wsm + (size_t)threadIdx.x * ZL;
        z[zpad(0)] = make_double2(0.0, 0.0);
        z[zpad(L / 2)] = make_double2(0.0, 0.0);
    }
    stage_store(stw, (double*)tw, a.tw, 2 * twiddle_table_len(L));
    __syncthreads();
    double2* z = wsm + (size_t)wv * ZL;
    fft_wave_any<P, -1>(z, L, s.nstages, s.radix, tw, lane);
    double* row0 = a.spec + ((long long)wv * s.nrow + r0) * s.kp;
    double* row1 = row0 + s.kp;
    if (!g.cyclic) {
        // DST(x)_K = -Im Z_K (row r0), DST(y)_K = Re Z_K (row r1)
        for (int K = lane + 1; K <= s.nin; K += 64) {
            const double2 v = z[zpad(K)];
            row0[K - 1] = -v.y;
            if (has1) row1[K - 1] = v.x;
        }
    } else {
        const int N = L;
        for (int kk = lane; kk <= N / 2; kk += 64) {
            const double2 A = z[zpad(kk)], B = z[zpad((N - kk) % N)];
            const double xr = 0.5 * (A.x + B.x), xi = 0.5 * (A.y - B.y);
            const double yr = 0.5 * (A.y + B.y), yi = 0.5 * (B.x - A.x);
            *(double2*)(row0 + 2 * kk) = make_double2(xr, xi);
            if (has1) *(double2*)(row1 + 2 * kk) = make_double2(yr, yi);
        }
    }
}

template <int NL, int P>
__global__ __launch_bounds__(64 * NL) void k_fwdw(FwdArgs a) {
    fwdw_block<NL, P>(a, xcd_row_pair(blockIdx.x, gridDim.x));
}

// Ensemble form: block b works on row pair map[b].y of member map[b].x.
template <int NL, int P>
__global__ __launch_bounds__(64 * NL) void k_fwdw_ens(const FwdArgs* __restrict__ args, const int2* __restrict__ map) {
    const int2 e = map[blockIdx.x];
    fwdw_block<NL, P>(args[e.x], e.y);
}

// The homogeneous correction of every mode at one point (homog.F90:323-327,
// 362-364): box, from the point's homogeneous solutions h[m-1] (m >= 1);
// channel, from its row's profiles hr = (bt, bc1(2..nl), bc2(2..nl)).
template <int NL>
__device__ __forceinline__ void homcor(bool cyc, const double* coef, const double* h, double (&hc)[NL]) {
    if (!cyc) {
        hc[0] = 0.0;
#pragma unroll
        for (int m = 1; m < NL; m++) hc[m] = coef[m - 1] * h[m - 1];
    } else {
        hc[0] = coef[0] * h[0];
#pragma unroll
        for (int m = 1; m < NL; m++) hc[m] = coef[1 + m] * h[m] + coef[1 + MAXL + m] * h[NL - 1 + m];
    }
}

// p_k = sum_m inh_m ctm2l(m,k) + sum_m hc_m ctm2l(m,k) (modes.F90:268-284), in
// the order of pmix.
template <int NL>
__device__ __forceinline__ void mix_point(const double (&inh)[NL], const double (&hc)[NL], const double* sctm2l,
                                          double (&pk)[NL]) {
#pragma unroll
    for (int k = 0; k < NL; k++) {
        double t = 0.0;
#pragma unroll
        for (int m = 0; m < NL; m++) t += inh[m] * sctm2l[m * MAXL + k];
        pk[k] = t;
    }
#pragma unroll
    for (int k = 0; k < NL; k++) {
        double t = 0.0;
#pragma unroll
        for (int m = 0; m < NL; m++) t += hc[m] * sctm2l[m * MAXL + k];
        pk[k] = pk[k] + t;
    }
}

// ocqbdy (vorsubs.F90:58-114) at one boundary point, with beta*yprel(gj)
// given: kind 0 zonal row, 1 box meridional column, 2 box corner.
template <int NL>
__device__ __forceinline__ void qbdy_at(const InvArgs& a, const double* samat, double qbfac, double by, int i, int gj,
                                        int kind, const double (&pb)[NL], const double (&pin)[NL]) {
    const Geo& g = a.g;
    const long long o = (long long)gj * g.pitch + i;
    const int reps = (kind == 2) ? 2 : 1;
#pragma unroll
    for (int k = 0; k < NL; k++) {
        double qv = qbfac * (pin[k] - pb[k]);
        double ap = 0.0;
#pragma unroll
        for (int l = 0; l < NL; l++) ap += samat[k * MAXL + l] * pb[l];
        for (int r = 0; r < reps; r++) qv = qv - g.fnot * ap;
        for (int r = 0; r < reps; r++) qv = qv + by;
        if (k == g.kt && a.ddyn)
            for (int r = 0; r < reps; r++) qv = qv + a.ddyn[o];
        a.qnew[k * g.fs + o] = qv;
    }
}

// The inverse transform's value of mode buffer z at column i of row sel,
// without branches (every read is from a valid slot).
template <bool CYC>
__device__ __forceinline__ double inh_at(const Geo& g, const Spec& s, const double2* z, int i, int sel) {
    if constexpr (!CYC) {
        const bool in = i >= 1 && i <= s.nin;
        const double2 v = z[zpad(in ? i : 0)];
        return in ? (sel ? v.x : -v.y) : 0.0;
    } else {
        const double2 v = z[zpad(i == g.nxp - 1 ? 0 : i)];
        return sel ? v.y : v.x;
    }
}

constexpr int INV_EPU = 3;              // epilogue points per thread and row with prefetched inputs
constexpr int HR = 1 + 2 * (MAXL - 1);   // a channel row's homogeneous profile values

// Row pair jp.  CYC: a channel (else a box).
template <int NL, int P, bool CYC>
__device__ __forceinline__ void invw_block(const InvArgs& a, int jp) {
    extern __shared__ double2 wsm[];   // [NL][zpad_len(L)] transforms, then the twiddle table
    constexpr int NH = NL > 1 ? NL - 1 : 1;
    const Geo& g = a.g;
    const Spec& s = a.s;
    const int L = s.L, ZL = zpad_len(L);
    const int nth = (int)blockDim.x;   // 64 (NL + 1)
    const int wv = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    const bool helper = wv == NL;
    double2* tw = wsm + (size_t)NL * ZL;
    __shared__ DevConsts s_dc;
    __shared__ double s_cin[C_LEN];
    __shared__ double s_aux[A_LEN];
    __shared__ double s_spk[2 * MAXL];
    __shared__ double s_csum[2 * MAXL * CQ_N];
    __shared__ double coef[3 * MAXL + 4];
    __shared__ double s_hin[4 * MAXL];   // box: hom at columns 1 and nxp-2 of the block's rows, [sel][side][m]
    __shared__ double s_hbc[4 * MAXL];   // box: hom at columns 1 and nxp-2 of the zonal rows, [zone][side][m]
    __shared__ double s_hr[4][HR];       // channel: the profiles of rows gj0, gj1, 0 and nyp-1
    const int r0 = 2 * jp, r1 = r0 + 1;
    const bool has1 = r1 < s.nyi;
    const int gj0 = r0 + 1, gj1 = r1 + 1;
    const int npairs = (s.nyi + 1) / 2;
    const bool first = jp == 0, last = jp == npairs - 1;
    const int zsel1 = has1 ? 1 : 0;   // the row of the last pair next to row nyp-1

    // ---- every global load of the block, issued before any is used
    const int nw_dc = (int)(sizeof(DevConsts) / 8);
    Staged<2> st_dc;
    stage_load(st_dc, a.dc, nw_dc);
    Staged<12, double> stw;
    stage_load(stw, a.twf, 2 * twiddle_table_len(L));
    // transform waves: the spectra of mode wv on both rows
    constexpr int SP = P / 2 + 1;
    double sx0[CYC ? 1 : SP], sx1[CYC ? 1 : SP];
    double2 cx0[CYC ? SP : 1], cx1[CYC ? SP : 1];
    if (!helper) {
        const double* row0 = a.spec + ((long long)wv * s.nrow + r0) * s.kp;
        const double* row1 = row0 + s.kp;
#pragma unroll
        for (int u = 0; u < SP; u++) {
            const int t = lane + 64 * u;
            if constexpr (!CYC) {
                const int tc = (t < s.nin) ? t : 0;
                sx0[u] = row0[tc];
                sx1[u] = row1[tc];
            } else {
                const int tc = (t <= L / 2) ? t : 0;
                cx0[u] = *(const double2*)(row0 + 2 * tc);
                cx1[u] = *(const double2*)(row1 + 2 * tc);
            }
        }
    }
    // helper wave: the integrals behind the correction coefficients and the
    // small per-block inputs of the epilogue
    double v[NL + 1];
#pragma unroll
    for (int t = 0; t <= NL; t++) v[t] = 0.0;
    double cinv = 0.0, spkv = 0.0, hinv = 0.0, hbcv = 0.0, hrv[4] = {0.0, 0.0, 0.0, 0.0};
    double txi_s = 0.0, txi_n = 0.0;
    if (helper) {
        if (lane < C_LEN) cinv = a.cin[lane];
        if constexpr (CYC) {
            txi_s = a.fcon[F_TXIS];
            txi_n = a.fcon[F_TXIN];
            if (lane < 2 * NL) {
                const int m = lane >> 1, r = (lane & 1) ? s.nyi - 1 : 0;
                spkv = a.spec[((long long)m * s.nrow + r) * s.kp];
            }
            // the row profiles (bt, bc1(2..nl), bc2(2..nl)) of rows gj0, gj1, 0, nyp-1
            if (lane < 2 * NL - 1) {
                const int rows[4] = {gj0, has1 ? gj1 : gj0, 0, g.nyp - 1};
                const double* src = lane == 0 ? a.hom
                                              : (lane < NL ? a.hom + g.nyp + (long long)(lane - 1) * g.nyp
                                                           : a.hom + g.nyp + (long long)(NL - 1 + lane - NL) * g.nyp);
#pragma unroll
                for (int r = 0; r < 4; r++) hrv[r] = src[rows[r]];
            }
        } else {
            if (lane < 4 * (NL - 1)) {
                const int sel = lane / (2 * (NL - 1)), side = (lane / (NL - 1)) & 1, m = lane % (NL - 1);
                const int gj = sel == 0 ? gj0 : gj1, i = side == 0 ? 1 : g.nxp - 2;
                if (sel == 0 || has1) hinv = a.hom[(long long)m * g.fs + (long long)gj * g.pitch + i];
                const int gz = sel == 0 ? 0 : g.nyp - 1;   // here sel is the zone
                if ((sel == 0 && first) || (sel == 1 && last))
                    hbcv = a.hom[(long long)m * g.fs + (long long)gz * g.pitch + i];
            }
            // ncol <= 32 P: at most P/2 columns per lane, all loaded before any is used
            constexpr int NC = P / 2;
            double cw[NC], cs[NC][NL];
#pragma unroll
            for (int u = 0; u < NC; u++) {
                const int col = lane + 64 * u;
                const bool ok = col < s.ncol;
                cw[u] = ok ? a.colw[col] : 0.0;
#pragma unroll
                for (int m = 0; m < NL; m++) cs[u][m] = ok ? a.colsum[m * s.ncol + col] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < NC; u++)
#pragma unroll
                for (int m = 0; m < NL; m++) v[m] += cw[u] * cs[u][m];
        }
        constexpr int NE = 4;   // entrainment partials per lane loaded together; more by the loop
        double ev[NE];
#pragma unroll
        for (int u = 0; u < NE; u++) ev[u] = (lane + 64 * u < a.nentpart) ? a.entpart[lane + 64 * u] : 0.0;
#pragma unroll
        for (int u = 0; u < NE; u++) v[NL] += ev[u];
        for (int t = lane + 64 * NE; t < a.nentpart; t += 64) v[NL] += a.entpart[t];
    }
    // every thread (box): the homogeneous solutions under its first epilogue
    // points of both rows (i = tid + nth u) and, in the first and last row
    // pairs, on the zonal boundary row next to the block
    const int zrow = first ? 0 : (last ? g.nyp - 1 : -1);   // nyp >= 8: never both
    const int zsel = first ? 0 : zsel1;
    double hv[2][INV_EPU][NH], hz[INV_EPU][NH];
#pragma unroll
    for (int u = 0; u < INV_EPU; u++) {
        const int i = (int)threadIdx.x + nth * u;
        const bool in = !CYC && i < g.nxp;
#pragma unroll
        for (int m = 1; m < NL; m++) {
            const double* hm = a.hom + (long long)(m - 1) * g.fs + i;
            hv[0][u][m - 1] = in ? hm[(long long)gj0 * g.pitch] : 0.0;
            hv[1][u][m - 1] = (in && has1) ? hm[(long long)gj1 * g.pitch] : 0.0;
            hz[u][m - 1] = (in && zrow >= 0) ? hm[(long long)(zrow > 0 ? zrow : 0) * g.pitch] : 0.0;
        }
    }
    // box: the ocqbdy work at the walls (items 0..3: row sel = e / 2, side
    // e % 2) and at the corners of the zonal row next to the first or last
    // row pair (items 4, 5) is done by the last six threads of the block, out
    // of the unrolled per-point code (a large branch there, duplicated per
    // point, cost 2.7 us of 13.7 at 513^2); each loads the homogeneous
    // solutions at its point here
    const int witem = nth - 1 - (int)threadIdx.x;
    const bool wact = !CYC && witem < 6 && (witem < 2 || (witem < 4 ? has1 : zrow >= 0));
    const int wsel = witem < 4 ? (witem >> 1) : 0, wside = witem & 1;
    const int wi = wside ? g.nxp - 1 : 0, wj = witem < 4 ? (wsel ? gj1 : gj0) : (zrow > 0 ? zrow : 0);
    double hw[NH];
#pragma unroll
    for (int m = 1; m < NL; m++) hw[m - 1] = wact ? a.hom[(long long)(m - 1) * g.fs + (long long)wj * g.pitch + wi] : 0.0;
    const double by0 = g.beta * a.yprel[gj0], by1 = has1 ? g.beta * a.yprel[gj1] : 0.0;
    const double byz0 = g.beta * a.yprel[0], byz1 = g.beta * a.yprel[g.nyp - 1];

    // ---- LDS: constants, twiddles, the transforms' inputs
    stage_store(st_dc, &s_dc, a.dc, nw_dc);
    stage_store(stw, (double*)tw, a.twf, 2 * twiddle_table_len(L));
    const double fn = s.ftnorm;
    if (!helper) {
        double2* z = wsm + (size_t)wv * ZL;
        if constexpr (!CYC) {
            // odd extension of ftnorm*u for the two rows (inhomog.F90:202-203)
            if (lane == 0) {
                z[zpad(0)] = make_double2(0.0, 0.0);
                z[zpad(L / 2)] = make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int u = 0; u < SP; u++) {
                const int t = lane + 64 * u;   // spectral index K-1
                if (t < s.nin) {
                    const double2 w = make_double2(fn * sx0[u], has1 ? fn * sx1[u] : 0.0);
                    z[zpad(t + 1)] = w;
                    z[zpad(L - t - 1)] = make_double2(-w.x, -w.y);
                }
            }
        } else {
            // Hermitian spectra of the two packed rows (inhomog.F90:246-247)
            const int N = L, h = N / 2;
#pragma unroll
            for (int u = 0; u < SP; u++) {
                const int kk = lane + 64 * u;
                if (kk <= h) {
                    double xr = fn * cx0[u].x, xi = fn * cx0[u].y;
                    double yr = has1 ? fn * cx1[u].x : 0.0, yi = has1 ? fn * cx1[u].y : 0.0;
                    if (kk == 0 || (N % 2 == 0 && kk == h)) { xi = 0.0; yi = 0.0; }
                    z[zpad(kk)] = make_double2(xr - yi, xi + yr);
                    if (kk != 0 && !(N % 2 == 0 && kk == h)) z[zpad(N - kk)] = make_double2(xr + yi, -xi + yr);
                }
            }
        }
    } else {
        if (lane < C_LEN) s_cin[lane] = cinv;
        if constexpr (CYC) {
            if (lane < 2 * NL) s_spk[lane] = spkv;
            if (lane < 2 * NL - 1) {
                // lane 0: bt; lanes 1..NL-1: bc1 of mode lane; NL..2NL-2: bc2 of mode lane-NL+1 (0-based modes)
#pragma unroll
                for (int r = 0; r < 4; r++) s_hr[r][lane] = hrv[r];
            }
        } else if (lane < 4 * (NL - 1)) {
            s_hin[lane] = hinv;
            s_hbc[lane] = hbcv;
        }
    }
    __syncthreads();

    if (!helper) {
        // ---- the inverse transform of mode wv
        double2* z = wsm + (size_t)wv * ZL;
        if constexpr (CYC) fft_wave_any<P, +1>(z, L, s.nstages, s.radix, tw, lane);   // c2r (inhomog.F90:247)
        else fft_wave_any<P, -1>(z, L, s.nstages, s.radix, tw, lane);
    } else {
        // ---- correction coefficients (the same in every block)
#pragma unroll
        for (int t = 0; t <= NL; t++) v[t] = wave_sum(v[t]);
        if constexpr (CYC) {
            // the boundary integrals of each layer from k_vort's per-tile partials, in tile order
            for (int idx = lane; idx < 2 * NL * CQ_N; idx += 64) {
                const int ks = idx / CQ_N, qy = idx - ks * CQ_N;
                const double* cp = a.cpart + (long long)ks * a.ctiles * CYC_SLOTS + qy;
                double t = cp[0];
                for (int u = 1; u < a.ctiles; u++) t += cp[u * CYC_SLOTS];
                s_csum[idx] = t;
            }
            wave_lds_sync();
            if (lane < NL) {
                const int k = lane;
                const double dx = g.dx;
                const Phys& ph = s_dc.ph;
                auto S = [&](int side, int q) { return s_csum[(k * 2 + side) * CQ_N + q]; };
                s_aux[A_AP3S + k] = ph.ah2[k] * (S(0, CQ_D2) * dx) * ph.rdy;
                s_aux[A_AP3N + k] = ph.ah2[k] * (S(1, CQ_D2) * dx) * ph.rdy;
                s_aux[A_AP5S + k] = ph.ah4[k] * (S(0, CQ_D4) * dx) * ph.rdy;
                s_aux[A_AP5N + k] = ph.ah4[k] * (S(1, CQ_D4) * dx) * ph.rdy;
                const double aj5s = (S(0, CQ_AJ5) * dx) * ph.rdx2, aj9s = (S(0, CQ_AJ9) * dx) * ph.rdx2;
                const double aj5n = (S(1, CQ_AJ5) * dx) * ph.rdx2, aj9n = (S(1, CQ_AJ9) * dx) * ph.rdx2;
                s_aux[A_AJIS + k] = (aj5s + 2.0 * aj9s) * (1.0 / 6.0);
                s_aux[A_AJIN + k] = -((aj5n + 2.0 * aj9n) * (1.0 / 6.0));
                if (k >= 2) {   // layers 1 and 2 are written below
                    s_aux[A_BES + k] = 0.0;
                    s_aux[A_BEN + k] = 0.0;
                }
                if (k == g.kt) {
                    s_aux[A_BDRS] = ph.cd_bdr * ((S(0, CQ_BD) * dx) * ph.rdy);
                    s_aux[A_BDRN] = ph.cd_bdr * ((S(1, CQ_BD) * dx) * ph.rdy);
                }
                if (k == 0) {
                    // int_Be_s/n (vorticity.F90:86-93): fohfac(2) in both northern terms
                    const double dz = (double)ph.dz_sign;
                    s_aux[A_BES + 0] = dz * ph.fohfac[0] * (S(0, CQ_E1) * dx);
                    s_aux[A_BES + 1] = dz * ph.fohfac[0] * (S(0, CQ_E2) * dx);
                    s_aux[A_BEN + 0] = -dz * ph.fohfac[1] * (S(1, CQ_E1) * dx);
                    s_aux[A_BEN + 1] = -dz * ph.fohfac[1] * (S(1, CQ_E2) * dx);
                    s_aux[A_TXIS] = txi_s;
                    s_aux[A_TXIN] = txi_n;
                }
            }
            wave_lds_sync();
        }
        if (lane == 0) step_coefficients<NL>(a, s_dc.mx, s_dc.bh, s_dc.ch, s_dc.ph, s_cin, s_aux, s_spk, v, coef, jp == 0);
    }
    __syncthreads();

    // ---- p_new on the block's rows and boundary q on their meridional ends;
    // the point's thread also writes the zonal boundary row next to it (first
    // and last row pairs)
    const double* s_ctm2l = s_dc.mx.ctm2l;
    const double* s_amat = s_dc.mx.amat;
    const double qbfac = s_dc.ph.qbfac;
    const int nxp = g.nxp;
    double* const pnew = a.pnew;
    const int zn = zrow == 0 ? 0 : 1;
    const double byz = zn == 0 ? byz0 : byz1;
    double zero[NL];
#pragma unroll
    for (int m = 0; m < NL; m++) zero[m] = 0.0;
    // p_new at point i of row sel (every point, walls included)
    auto main_point = [&](int sel, int i, const double (&hp)[NH], double (&pk)[NL]) {
        const int gj = sel == 0 ? gj0 : gj1;
        double inh[NL], hc[NL];
#pragma unroll
        for (int m = 0; m < NL; m++) inh[m] = inh_at<CYC>(g, s, wsm + (size_t)m * ZL, i, sel);
        if constexpr (CYC) homcor<NL>(true, coef, s_hr[sel], hc);
        else homcor<NL>(false, coef, hp, hc);
        mix_point<NL>(inh, hc, s_ctm2l, pk);
        const long long o = (long long)gj * g.pitch + i;
#pragma unroll
        for (int k = 0; k < NL; k++) pnew[k * g.fs + o] = pk[k];
    };
    // the zonal row next to the first and last row pairs: p (the homogeneous
    // correction alone) and, away from a box's corners, its boundary q from the
    // inner row's p at the same column (pin)
    auto zone_point = [&](int i, const double (&hzp)[NH], const double (&pin)[NL]) {
        double hc[NL], pb[NL];
        if constexpr (CYC) homcor<NL>(true, coef, s_hr[2 + zn], hc);
        else homcor<NL>(false, coef, hzp, hc);
        mix_point<NL>(zero, hc, s_ctm2l, pb);
        const long long oz = (long long)zrow * g.pitch + i;
#pragma unroll
        for (int k = 0; k < NL; k++) pnew[k * g.fs + oz] = pb[k];
        if (CYC || (i != 0 && i != nxp - 1)) qbdy_at<NL>(a, s_amat, qbfac, byz, i, zrow, 0, pb, pin);
    };
    double pks[2][INV_EPU][NL] = {};
#pragma unroll
    for (int sel = 0; sel < 2; sel++) {
        if (sel == 1 && !has1) break;
#pragma unroll
        for (int u = 0; u < INV_EPU; u++) {
            const int i = (int)threadIdx.x + nth * u;
            if (i < nxp) main_point(sel, i, hv[sel][u], pks[sel][u]);
        }
    }
    if (zrow >= 0) {
#pragma unroll
        for (int u = 0; u < INV_EPU; u++) {
            const int i = (int)threadIdx.x + nth * u;
            double pin[NL];
#pragma unroll
            for (int k = 0; k < NL; k++) pin[k] = zsel ? pks[1][u][k] : pks[0][u][k];
            if (i < nxp) zone_point(i, hz[u], pin);
        }
    }
    // wide basins: the points past the prefetched ones
    for (int i = (int)threadIdx.x + nth * INV_EPU; i < nxp; i += nth) {
        double h[2][NH], hzr[NH], pk[2][NL];
#pragma unroll
        for (int m = 1; m < NL; m++) {
            const double* hm = a.hom + (long long)(m - 1) * g.fs + i;
            h[0][m - 1] = CYC ? 0.0 : hm[(long long)gj0 * g.pitch];
            h[1][m - 1] = (CYC || !has1) ? 0.0 : hm[(long long)gj1 * g.pitch];
            hzr[m - 1] = (CYC || zrow < 0) ? 0.0 : hm[(long long)zrow * g.pitch];
        }
        main_point(0, i, h[0], pk[0]);
        if (has1) main_point(1, i, h[1], pk[1]);
        if (zrow >= 0) {
            double pz[NL];
#pragma unroll
            for (int k = 0; k < NL; k++) pz[k] = zsel ? pk[1][k] : pk[0][k];
            zone_point(i, hzr, pz);
        }
    }
    // box walls and corners (vorsubs.F90:79-84: the corners take the x
    // difference): p at the wall point from the homogeneous correction (the
    // transforms vanish there), p at its inner neighbour (columns 1, nxp-2)
    if (wact) {
        double hc[NL], pb[NL], pin[NL], inh[NL];
        homcor<NL>(false, coef, hw, hc);
        mix_point<NL>(zero, hc, s_ctm2l, pb);
        if (witem < 4) {
            const int ii = wside ? nxp - 2 : 1;
#pragma unroll
            for (int m = 0; m < NL; m++) inh[m] = inh_at<CYC>(g, s, wsm + (size_t)m * ZL, ii, wsel);
            homcor<NL>(false, coef, s_hin + (wsel * 2 + wside) * (NL - 1), hc);
            mix_point<NL>(inh, hc, s_ctm2l, pin);
            qbdy_at<NL>(a, s_amat, qbfac, wsel == 0 ? by0 : by1, wi, wj, 1, pb, pin);
        } else {
            homcor<NL>(false, coef, s_hbc + (zn * 2 + wside) * (NL - 1), hc);
            mix_point<NL>(zero, hc, s_ctm2l, pin);
            qbdy_at<NL>(a, s_amat, qbfac, byz, wi, zrow, 2, pb, pin);
        }
    }
}

template <int NL, int P, bool CYC>
__global__ __launch_bounds__(64 * (NL + 1)) void k_invw(InvArgs a) {
    invw_block<NL, P, CYC>(a, xcd_row_pair(blockIdx.x, gridDim.x));
}

template <int NL, int P, bool CYC>
__global__ __launch_bounds__(64 * (NL + 1)) void k_invw_ens(const InvArgs* __restrict__ args,
                                                            const int2* __restrict__ map) {
    const int2 e = map[blockIdx.x];
    invw_block<NL, P, CYC>(args[e.x], e.y);
}

// ------------------------------------------------------------------ adjoint forms (box)
//
// The adjoint of the box inversion in the same layout.  k_fwdw_adj: the input
// of mode m is w_m = sum_k ctm2l(m,k) lp_k, lp the adjoint of p completed by
// the transpose of ocqbdy from the adjoint of q's boundary ring; the row
// pair's sums of hom_m w_m (over its interior points, the ends of its rows and
// its share of rows 0 and nyp-1) go to adj_partial[jp][m].  k_invw_adj: the
// inverse transforms u_m, the constants s_m from the partial sums and
// cdhoc^-T, and lq_k += f0 sum_m ctl2m(k,m) (u_m + s_m G_m) on the interior.

template <int NL, int P>
__device__ __forceinline__ void fwdw_adj_block(const FwdArgs& a, int jp) {
    extern __shared__ double2 wsm[];
    constexpr int NH = NL > 1 ? NL - 1 : 1;
    const Geo& g = a.g;
    const Spec& s = a.s;
    const int L = s.L, ZL = zpad_len(L);
    const int nth = (int)blockDim.x;   // 64 NL
    const int wv = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    double2* tw = wsm + (size_t)NL * ZL;
    __shared__ double s_red[(NL + 1) * 16];
    const int r0 = 2 * jp, r1 = r0 + 1;
    const bool has1 = r1 < s.nyi;
    const int gj0 = r0 + 1, gj1 = r1 + 1;
    Staged<12, double> stw;
    stage_load(stw, a.tw, 2 * twiddle_table_len(L));
    const double qb = a.adj_qbfac;
    const double* lq = a.adj_lq;
    double cm[NL][NL];   // ctm2l(m, k)
#pragma unroll
    for (int m = 0; m < NL; m++)
#pragma unroll
        for (int k = 0; k < NL; k++) cm[m][k] = a.dc->mx.ctm2l[m * MAXL + k];
    // lp, hom and, on the rows next to rows 0 and nyp-1, the adjoint of q on
    // those rows (the ring neighbours the transpose of ocqbdy gathers), at this
    // thread's first points of both rows, all loaded before any is used
    const bool s0 = gj0 == 1, n0 = gj0 == g.nyp - 2, n1 = has1 && gj1 == g.nyp - 2;
    double p0[WFPT][NL], p1[WFPT][NL], h0[WFPT][NH], h1[WFPT][NH], rq0[WFPT][NL], rq1[WFPT][NL];
#pragma unroll
    for (int u = 0; u < WFPT; u++) {
        const int t = 1 + (int)threadIdx.x + nth * u;
        const bool in = t <= s.nin;
        const long long o0 = (long long)gj0 * g.pitch + t, o1 = (long long)gj1 * g.pitch + t;
#pragma unroll
        for (int k = 0; k < NL; k++) {
            p0[u][k] = in ? a.adj_lp[k * g.fs + o0] : 0.0;
            p1[u][k] = (in && has1) ? a.adj_lp[k * g.fs + o1] : 0.0;
            rq0[u][k] = (in && (s0 || n0)) ? lq[k * g.fs + o0 + (s0 ? -g.pitch : g.pitch)] : 0.0;
            rq1[u][k] = (in && n1) ? lq[k * g.fs + o1 + g.pitch] : 0.0;
        }
#pragma unroll
        for (int m = 1; m < NL; m++) {
            h0[u][m - 1] = in ? a.adj_hom[(long long)(m - 1) * g.fs + o0] : 0.0;
            h1[u][m - 1] = (in && has1) ? a.adj_hom[(long long)(m - 1) * g.fs + o1] : 0.0;
        }
    }
    // ocqbdy's x difference at the walls, transposed: the inner points t = 1
    // and nxp-2 of both rows take qb times the adjoint of the ring's q beside
    // them.  Two threads (one per wall) load those ring values here and add
    // their part after the points are written, out of the unrolled per-point
    // code (a dependent load there sat on two wavefronts' paths).
    const int cw = (int)threadIdx.x - (nth - 2);   // 0: west wall, 1: east wall, else none
    const bool cact = cw >= 0;
    const int ct = cw == 1 ? g.nxp - 2 : 1;
    double cr0[NL], cr1[NL], ch0[NH], ch1[NH];
#pragma unroll
    for (int k = 0; k < NL; k++) {
        const long long o0 = (long long)gj0 * g.pitch + (cw == 1 ? g.nxp - 1 : 0);
        const long long o1 = (long long)gj1 * g.pitch + (cw == 1 ? g.nxp - 1 : 0);
        cr0[k] = cact ? lq[k * g.fs + o0] : 0.0;
        cr1[k] = (cact && has1) ? lq[k * g.fs + o1] : 0.0;
    }
#pragma unroll
    for (int m = 1; m < NL; m++) {
        ch0[m - 1] = cact ? a.adj_hom[(long long)(m - 1) * g.fs + (long long)gj0 * g.pitch + ct] : 0.0;
        ch1[m - 1] = (cact && has1) ? a.adj_hom[(long long)(m - 1) * g.fs + (long long)gj1 * g.pitch + ct] : 0.0;
    }
    double part[NL];
#pragma unroll
    for (int m = 0; m < NL; m++) part[m] = 0.0;
    // ring points of the inner products: columns 0 and nxp-1 of the block's
    // rows, and its share of rows 0 and nyp-1 (2 nxp points dealt out evenly).
    // A thread's first one is loaded with the batch above: lp and lq at the
    // point, lq at the corner beside it when its inner point is this point
    // (columns 1 and nxp-2 of rows 0 and nyp-1), and the homogeneous solutions.
    const int npairs = (s.nyi + 1) / 2, nring = 2 * g.nxp;
    const int rper = (nring + npairs - 1) / npairs;
    const int rlo = min(nring, jp * rper), rhi = min(nring, rlo + rper);
    const int nside = has1 ? 4 : 2;
    const int nextra = nside + (rhi - rlo);
    auto extra_point = [&](int e, int& i, int& gj) {
        if (e < nside) {
            gj = (e >> 1) ? gj1 : gj0;
            i = (e & 1) ? g.nxp - 1 : 0;
        } else {
            const int idx = rlo + (e - nside);
            gj = idx < g.nxp ? 0 : g.nyp - 1;
            i = idx < g.nxp ? idx : idx - g.nxp;
        }
    };
    const bool ehas = (int)threadIdx.x < nextra;
    int ei = 0, ej = 0;
    if (ehas) extra_point(threadIdx.x, ei, ej);
    const bool erow = ej == 0 || ej == g.nyp - 1;
    const bool ewc = ehas && erow && ei == 1, eec = ehas && erow && ei == g.nxp - 2;
    const long long eo = (long long)ej * g.pitch + ei;
    double elp[NL], elq[NL], elc[NL], ehm[NH];
#pragma unroll
    for (int k = 0; k < NL; k++) {
        elp[k] = ehas ? a.adj_lp[k * g.fs + eo] : 0.0;
        elq[k] = ehas ? lq[k * g.fs + eo] : 0.0;
        elc[k] = (ewc || eec) ? lq[k * g.fs + eo + (ewc ? -1 : 1)] : 0.0;
    }
#pragma unroll
    for (int m = 1; m < NL; m++) ehm[m - 1] = ehas ? a.adj_hom[(long long)(m - 1) * g.fs + eo] : 0.0;
    // the transpose of ocqbdy at an interior point: the ring points whose
    // inner point it is are its west (t = 1), east (t = nxp-2), south (gj = 1)
    // and north (gj = nyp-2) neighbours, added in that order; rs/rn: the
    // loaded south or north neighbours
    auto gather = [&](int t, int gj, double (&lp)[NL], const double (&rs)[NL]) {
        if (gj == 1 || gj == g.nyp - 2)
#pragma unroll
            for (int k = 0; k < NL; k++) lp[k] += qb * rs[k];
    };
    auto put = [&](int t, double (&la)[NL], double (&lb)[NL], const double (&ha)[NH], const double (&hb)[NH],
                   const double (&ra)[NL], const double (&rb)[NL]) {
        gather(t, gj0, la, ra);
        if (has1) gather(t, gj1, lb, rb);
#pragma unroll
        for (int m = 0; m < NL; m++) {
            double wa = 0.0, wb = 0.0;
#pragma unroll
            for (int k = 0; k < NL; k++) {
                wa += cm[m][k] * la[k];
                wb += cm[m][k] * lb[k];
            }
            if (!has1) wb = 0.0;
            double2* z = wsm + (size_t)m * ZL;
            z[zpad(t)] = make_double2(wa, wb);
            z[zpad(L - t)] = make_double2(-wa, -wb);
            if (m > 0) part[m] += ha[m - 1] * wa + hb[m - 1] * wb;
        }
    };
#pragma unroll
    for (int u = 0; u < WFPT; u++) {
        const int t = 1 + (int)threadIdx.x + nth * u;
        if (t <= s.nin) put(t, p0[u], p1[u], h0[u], h1[u], rq0[u], rq1[u]);
    }
    for (int t = 1 + (int)threadIdx.x + nth * WFPT; t <= s.nin; t += nth) {
        const long long o0 = (long long)gj0 * g.pitch + t, o1 = (long long)gj1 * g.pitch + t;
        double la[NL], lb[NL], ha[NH], hb[NH], ra[NL], rb[NL];
#pragma unroll
        for (int k = 0; k < NL; k++) {
            la[k] = a.adj_lp[k * g.fs + o0];
            lb[k] = has1 ? a.adj_lp[k * g.fs + o1] : 0.0;
            ra[k] = (s0 || n0) ? lq[k * g.fs + o0 + (s0 ? -g.pitch : g.pitch)] : 0.0;
            rb[k] = n1 ? lq[k * g.fs + o1 + g.pitch] : 0.0;
        }
#pragma unroll
        for (int m = 1; m < NL; m++) {
            ha[m - 1] = a.adj_hom[(long long)(m - 1) * g.fs + o0];
            hb[m - 1] = has1 ? a.adj_hom[(long long)(m - 1) * g.fs + o1] : 0.0;
        }
        put(t, la, lb, ha, hb, ra, rb);
    }
    // (the ring points' arithmetic after the block's own points)
    if (ehas) {
        // bdy_gather at a ring point from the loaded values: the point's
        // own term (rows and columns of ocqbdy twice at a corner), then
        // the corner whose inner point it is
        const Mix& mx = a.dc->mx;
        const double r = ((ei == 0 || ei == g.nxp - 1) && erow) ? 2.0 : 1.0;
#pragma unroll
        for (int k = 0; k < NL; k++) {
#pragma unroll
            for (int l = 0; l < NL; l++) {
                double c = -r * g.fnot * mx.amat[k * MAXL + l];
                if (l == k) c -= qb;
                elp[l] += elq[k] * c;
            }
        }
        if (ewc || eec)
#pragma unroll
            for (int k = 0; k < NL; k++) elp[k] += qb * elc[k];
#pragma unroll
        for (int m = 1; m < NL; m++) {
            double w = 0.0;
#pragma unroll
            for (int k = 0; k < NL; k++) w += cm[m][k] * elp[k];
            part[m] += ehm[m - 1] * w;
        }
    }
    for (int e = (int)threadIdx.x + nth; e < nextra; e += nth) {   // very wide, short basins
        int i, gj;
        extra_point(e, i, gj);
        const long long o = (long long)gj * g.pitch + i;
        double lp[NL];
#pragma unroll
        for (int k = 0; k < NL; k++) lp[k] = a.adj_lp[k * g.fs + o];
        bdy_gather<NL>(g, qb, a.dc->mx, lq, i, gj, lp);   // a ring point
#pragma unroll
        for (int m = 1; m < NL; m++) {
            double w = 0.0;
#pragma unroll
            for (int k = 0; k < NL; k++) w += cm[m][k] * lp[k];
            part[m] += a.adj_hom[(long long)(m - 1) * g.fs + o] * w;
        }
    }
    if (threadIdx.x < NL) {
        double2* z = wsm + (size_t)threadIdx.x * ZL;
        z[zpad(0)] = make_double2(0.0, 0.0);
        z[zpad(L / 2)] = make_double2(0.0, 0.0);
    }
    stage_store(stw, (double*)tw, a.tw, 2 * twiddle_table_len(L));
    // the walls' part: its modal input and its share of the inner products
    double dw0[NL], dw1[NL];
#pragma unroll
    for (int m = 0; m < NL; m++) {
        double s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int k = 0; k < NL; k++) {
            s0 += cm[m][k] * (qb * cr0[k]);
            s1 += cm[m][k] * (qb * cr1[k]);
        }
        dw0[m] = s0;
        dw1[m] = has1 ? s1 : 0.0;
        if (m > 0) part[m] += ch0[m - 1] * dw0[m] + ch1[m - 1] * dw1[m];
    }
    block_sum<NL>(part, s_red);   // its barriers also publish the transforms' inputs
    if (threadIdx.x == 0)
#pragma unroll
        for (int m = 0; m < NL; m++) a.adj_partial[(long long)jp * MAXL + m] = part[m];
    if (cact) {
#pragma unroll
        for (int m = 0; m < NL; m++) {
            double2* zm = wsm + (size_t)m * ZL;
            const double2 v = zm[zpad(ct)];
            zm[zpad(ct)] = make_double2(v.x + dw0[m], v.y + dw1[m]);
            zm[zpad(L - ct)] = make_double2(-(v.x + dw0[m]), -(v.y + dw1[m]));
        }
    }
    __syncthreads();
    double2* z = wsm + (size_t)wv * ZL;
    fft_wave_any<P, -1>(z, L, s.nstages, s.radix, tw, lane);
    double* row0 = a.spec + ((long long)wv * s.nrow + r0) * s.kp;
    double* row1 = row0 + s.kp;
    for (int K = lane + 1; K <= s.nin; K += 64) {
        const double2 v = z[zpad(K)];
        row0[K - 1] = -v.y;
        if (has1) row1[K - 1] = v.x;
    }
}

template <int NL, int P>
__global__ __launch_bounds__(64 * NL) void k_fwdw_adj(FwdArgs a) {
    fwdw_adj_block<NL, P>(a, blockIdx.x);
}

template <int NL, int P>
__device__ __forceinline__ void invw_adj_block(const InvArgs& a, int jp) {
    extern __shared__ double2 wsm[];
    const Geo& g = a.g;
    const Spec& s = a.s;
    const int L = s.L, ZL = zpad_len(L);
    const int nth = (int)blockDim.x;   // 64 (NL + 1)
    const int wv = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    const bool helper = wv == NL;
    double2* tw = wsm + (size_t)NL * ZL;
    __shared__ double coef[MAXL];
    __shared__ double s_ctl2m[MAXL * MAXL];
    const int r0 = 2 * jp, r1 = r0 + 1;
    const bool has1 = r1 < s.nyi;
    const int gj0 = r0 + 1, gj1 = r1 + 1;
    Staged<12, double> stw;
    stage_load(stw, a.twf, 2 * twiddle_table_len(L));
    constexpr int SP = P / 2 + 1;
    double sx0[SP], sx1[SP];
    if (!helper) {
        const double* row0 = a.spec + ((long long)wv * s.nrow + r0) * s.kp;
        const double* row1 = row0 + s.kp;
#pragma unroll
        for (int u = 0; u < SP; u++) {
            const int t = lane + 64 * u;
            const int tc = (t < s.nin) ? t : 0;
            sx0[u] = row0[tc];
            sx1[u] = row1[tc];
        }
    }
    // helper wave: the sums over row pairs of the forward kernel's partials, in row-pair order per lane
    double la[NL];
#pragma unroll
    for (int m = 0; m < NL; m++) la[m] = 0.0;
    double ctlv = 0.0;
    if (helper) {
        if (lane < MAXL * MAXL) ctlv = a.dc->mx.ctl2m[lane];
        // the first four per lane loaded together, then added in row-pair order
        constexpr int NB = 4;
        double pb[NB][NL];
#pragma unroll
        for (int u = 0; u < NB; u++) {
            const int b = lane + 64 * u;
#pragma unroll
            for (int m = 0; m < NL; m++) pb[u][m] = b < a.adj_npart ? a.adj_partial[(long long)b * MAXL + m] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < NB; u++)
            if (lane + 64 * u < a.adj_npart)
#pragma unroll
                for (int m = 0; m < NL; m++) la[m] += pb[u][m];
        for (int b = lane + 64 * NB; b < a.adj_npart; b += 64)
#pragma unroll
            for (int m = 0; m < NL; m++) la[m] += a.adj_partial[(long long)b * MAXL + m];
    }
    // every thread: G and lq at its first interior points of both rows
    double gv[2][INV_EPU][NL], lqv[2][INV_EPU][NL];
#pragma unroll
    for (int sel = 0; sel < 2; sel++)
#pragma unroll
        for (int u = 0; u < INV_EPU; u++) {
            const int i = (int)threadIdx.x + nth * u;
            const bool in = i >= 1 && i <= g.nxp - 2 && (sel == 0 || has1);
            const long long o = (long long)(sel == 0 ? gj0 : gj1) * g.pitch + (in ? i : 0);
#pragma unroll
            for (int m = 0; m < NL; m++) {
                gv[sel][u][m] = in ? a.adj_G[m * g.fs + o] : 0.0;
                lqv[sel][u][m] = in ? a.adj_lq[m * g.fs + o] : 0.0;
            }
        }
    stage_store(stw, (double*)tw, a.twf, 2 * twiddle_table_len(L));
    const double fn = s.ftnorm;
    if (!helper) {
        double2* z = wsm + (size_t)wv * ZL;
        if (lane == 0) {
            z[zpad(0)] = make_double2(0.0, 0.0);
            z[zpad(L / 2)] = make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < SP; u++) {
            const int t = lane + 64 * u;
            if (t < s.nin) {
                const double2 w = make_double2(fn * sx0[u], has1 ? fn * sx1[u] : 0.0);
                z[zpad(t + 1)] = w;
                z[zpad(L - t - 1)] = make_double2(-w.x, -w.y);
            }
        }
    } else if (lane < MAXL * MAXL) {
        s_ctl2m[lane] = ctlv;
    }
    __syncthreads();
    if (!helper) {
        fft_wave_any<P, -1>(wsm + (size_t)wv * ZL, L, s.nstages, s.radix, tw, lane);
    } else {
#pragma unroll
        for (int m = 0; m < NL; m++) la[m] = wave_sum(la[m]);
        if (lane == 0) {
            const Mix& mx = a.dc->mx;
            double lr[NL];
#pragma unroll
            for (int l = 0; l < NL - 1; l++) {
                double sacc = 0.0;
#pragma unroll
                for (int j = 0; j < NL - 1; j++) sacc += a.adj_cinvT[l * MAXL + j] * la[j + 1];
                lr[l] = sacc;
            }
#pragma unroll
            for (int mm = 0; mm < NL; mm++) {
                double sacc = 0.0;
#pragma unroll
                for (int l = 0; l < NL - 1; l++) sacc -= lr[l] * (mx.ctm2l[mm * MAXL + l + 1] - mx.ctm2l[mm * MAXL + l]);
                coef[mm] = (sacc * g.dx) * g.dy;
            }
        }
    }
    __syncthreads();
    // lq_k += f0 sum_m ctl2m(k,m) (u_m + s_m G_m) on the block's interior points
    auto point = [&](int sel, int i, const double (&G)[NL], const double (&lqk)[NL]) {
        const long long o = (long long)(sel == 0 ? gj0 : gj1) * g.pitch + i;
        double r[NL];
#pragma unroll
        for (int m = 0; m < NL; m++) r[m] = inh_at<false>(g, s, wsm + (size_t)m * ZL, i, sel) + coef[m] * G[m];
#pragma unroll
        for (int k = 0; k < NL; k++) {
            double acc = 0.0;
#pragma unroll
            for (int m = 0; m < NL; m++) acc += s_ctl2m[k * MAXL + m] * r[m];
            a.adj_lq[k * g.fs + o] = lqk[k] + g.fnot * acc;
        }
    };
#pragma unroll
    for (int sel = 0; sel < 2; sel++) {
        if (sel == 1 && !has1) break;
#pragma unroll
        for (int u = 0; u < INV_EPU; u++) {
            const int i = (int)threadIdx.x + nth * u;
            if (i >= 1 && i <= g.nxp - 2) point(sel, i, gv[sel][u], lqv[sel][u]);
        }
        for (int i = (int)threadIdx.x + nth * INV_EPU; i <= g.nxp - 2; i += nth) {
            const long long o = (long long)(sel == 0 ? gj0 : gj1) * g.pitch + i;
            double G[NL], lqk[NL];
#pragma unroll
            for (int m = 0; m < NL; m++) {
                G[m] = a.adj_G[m * g.fs + o];
                lqk[m] = a.adj_lq[m * g.fs + o];
            }
            point(sel, i, G, lqk);
        }
    }
}

template <int NL, int P>
__global__ __launch_bounds__(64 * (NL + 1)) void k_invw_adj(InvArgs a) {
    invw_adj_block<NL, P>(a, blockIdx.x);
}

// ---- the channel's inversion adjoint, on the torus of the nxp-1 distinct
// columns (column nxp-1 of every adjoint field stays 0).  Forward: p_k =
// sum_m ctm2l(m,k) (u_m + hc_m(j)), u_m = H_m^-1 r_m the periodic Helmholtz
// solve of r_m = f0 sum_k ctl2m(k,m) (q_k - beta y - ddyn) on rows 1..nyp-2,
// hc_m(j) = c_bt bt(j) (m = 0) or c_bc1(m) bc1_m(j) + c_bc2(m) bc2_m(j), the
// coefficients from clhss_m = sum_k ctl2m(k,m) cs_k + (dx/dy) sum_i u_m(i,1)
// and clhsn_m likewise (homog.F90:280-329), and ocqbdy on rows 0 and nyp-1.
// Adjoint: w_m = sum_k ctm2l(m,k) lp_k (lp completed by ocqbdy's transpose);
// the profile inner products sum_j prof(j) sum_i w_m(i,j) give the adjoints of
// the coefficients and through them lclhss_m, lclhsn_m; lcs_k += sum_m
// ctl2m(k,m) lclhss_m (lcn_k likewise); the adjoint of u_m is w_m plus
// (dx/dy) lclhss_m on row 1 and (dx/dy) lclhsn_m on row nyp-2, and H_m is
// symmetric, so lr_m = H_m^-1 w_m + (dx/dy)(lclhss_m Gs_m + lclhsn_m Gn_m)
// with Gs_m, Gn_m = H_m^-1 of the unit rows; lq_k += f0 sum_m ctl2m(k,m) lr_m.

// k_fwdw_adj, channel: the inputs w_m of both rows of the pair through the
// forward real-to-complex transform, and the pair's profile inner products
// (the pairs also share out the points of rows 0 and nyp-1) at
// adj_partial[row pair * ADJ_CYC_NPROF + profile].
template <int NL, int P>
__device__ __forceinline__ void fwdw_adj_cyc_block(const FwdArgs& a, int jp) {
    extern __shared__ double2 wsm[];
    constexpr int NP = 2 * NL - 1;
    const Geo& g = a.g;
    const Spec& s = a.s;
    const int L = s.L, ZL = zpad_len(L), N = g.nyp - 1;
    const int nth = (int)blockDim.x;   // 64 NL
    const int wv = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    double2* tw = wsm + (size_t)NL * ZL;
    __shared__ double s_red[NP * 16];
    const int r0 = 2 * jp, r1 = r0 + 1;
    const bool has1 = r1 < s.nyi;
    const int gj0 = r0 + 1, gj1 = r1 + 1;
    Staged<12, double> stw;
    stage_load(stw, a.tw, 2 * twiddle_table_len(L));
    const double qb = a.adj_qbfac;
    const double* lq = a.adj_lq;
    const Mix& mx = a.dc->mx;
    double cm[NL][NL];   // ctm2l(m, k)
#pragma unroll
    for (int m = 0; m < NL; m++)
#pragma unroll
        for (int k = 0; k < NL; k++) cm[m][k] = mx.ctm2l[m * MAXL + k];
    double h0[NP], h1[NP];
#pragma unroll
    for (int q = 0; q < NP; q++) {
        h0[q] = a.adj_hom[(long long)q * g.nyp + gj0];
        h1[q] = has1 ? a.adj_hom[(long long)q * g.nyp + gj1] : 0.0;
    }
    double part[NP];
#pragma unroll
    for (int q = 0; q < NP; q++) part[q] = 0.0;
    auto prof_add = [&](const double (&w)[NL], const double (&h)[NP]) {
        part[0] += h[0] * w[0];
#pragma unroll
        for (int m = 1; m < NL; m++) {
            part[m] += h[m] * w[m];
            part[NL - 1 + m] += h[NL - 1 + m] * w[m];
        }
    };
    for (int t = (int)threadIdx.x; t < L; t += nth) {
        const long long o0 = (long long)gj0 * g.pitch + t, o1 = (long long)gj1 * g.pitch + t;
        double la[NL], lb[NL];
#pragma unroll
        for (int k = 0; k < NL; k++) {
            la[k] = a.adj_lp[k * g.fs + o0];
            lb[k] = has1 ? a.adj_lp[k * g.fs + o1] : 0.0;
        }
        // ocqbdy's fac (p(inner row) - p(boundary row)), transposed at the inner rows
        if (gj0 == 1)
#pragma unroll
            for (int k = 0; k < NL; k++) la[k] += qb * lq[k * g.fs + t];
        if (gj0 == N - 1)
#pragma unroll
            for (int k = 0; k < NL; k++) la[k] += qb * lq[k * g.fs + (long long)N * g.pitch + t];
        if (has1 && gj1 == N - 1)
#pragma unroll
            for (int k = 0; k < NL; k++) lb[k] += qb * lq[k * g.fs + (long long)N * g.pitch + t];
        double wa[NL], wb[NL];
#pragma unroll
        for (int m = 0; m < NL; m++) {
            double sa = 0.0, sb = 0.0;
#pragma unroll
            for (int k = 0; k < NL; k++) {
                sa += cm[m][k] * la[k];
                sb += cm[m][k] * lb[k];
            }
            wa[m] = sa;
            wb[m] = has1 ? sb : 0.0;
            wsm[(size_t)m * ZL + zpad(t)] = make_double2(wa[m], wb[m]);
        }
        prof_add(wa, h0);
        if (has1) prof_add(wb, h1);
    }
    // rows 0 and nyp-1: 2 (nxp-1) points dealt out over the row pairs; p there
    // is the homogeneous correction alone
    const int npairs = (s.nyi + 1) / 2, nring = 2 * L;
    const int rper = (nring + npairs - 1) / npairs;
    const int rlo = min(nring, jp * rper), rhi = min(nring, rlo + rper);
    for (int e = rlo + (int)threadIdx.x; e < rhi; e += nth) {
        const int gj = e < L ? 0 : N, i = e < L ? e : e - L;
        const long long o = (long long)gj * g.pitch + i;
        double lp[NL], lqb[NL], hb[NP], w[NL];
#pragma unroll
        for (int k = 0; k < NL; k++) {
            lp[k] = a.adj_lp[k * g.fs + o];
            lqb[k] = lq[k * g.fs + o];
        }
#pragma unroll
        for (int q = 0; q < NP; q++) hb[q] = a.adj_hom[(long long)q * g.nyp + gj];
        // ocqbdy at the point: q_k = fac (p_k(inner) - p_k) - f0 sum_l amat(k,l) p_l + ...
#pragma unroll
        for (int k = 0; k < NL; k++)
#pragma unroll
            for (int l = 0; l < NL; l++) {
                double c = -g.fnot * mx.amat[k * MAXL + l];
                if (l == k) c -= qb;
                lp[l] += lqb[k] * c;
            }
#pragma unroll
        for (int m = 0; m < NL; m++) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < NL; k++) t += cm[m][k] * lp[k];
            w[m] = t;
        }
        prof_add(w, hb);
    }
    stage_store(stw, (double*)tw, a.tw, 2 * twiddle_table_len(L));
    block_sum<NP>(part, s_red);   // its barriers also publish the transforms' inputs
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NP; q++) a.adj_partial[(long long)jp * ADJ_CYC_NPROF + q] = part[q];
    double2* z = wsm + (size_t)wv * ZL;
    fft_wave_any<P, -1>(z, L, s.nstages, s.radix, tw, lane);
    double* row0 = a.spec + ((long long)wv * s.nrow + r0) * s.kp;
    double* row1 = row0 + s.kp;
    for (int kk = lane; kk <= L / 2; kk += 64) {
        const double2 A = z[zpad(kk)], B = z[zpad((L - kk) % L)];
        const double xr = 0.5 * (A.x + B.x), xi = 0.5 * (A.y - B.y);
        const double yr = 0.5 * (A.y + B.y), yi = 0.5 * (B.x - A.x);
        *(double2*)(row0 + 2 * kk) = make_double2(xr, xi);
        if (has1) *(double2*)(row1 + 2 * kk) = make_double2(yr, yi);
    }
}

template <int NL, int P>
__global__ __launch_bounds__(64 * NL) void k_fwdw_adj_cyc(FwdArgs a) {
    fwdw_adj_cyc_block<NL, P>(a, blockIdx.x);
}

// k_invw_adj, channel: the complex-to-real transforms of the solved spectra
// while the helper wavefront turns the profile sums into lclhss, lclhsn (and
// block 0 adds the constraint adjoints), then lq on the pair's rows.
template <int NL, int P>
__device__ __forceinline__ void invw_adj_cyc_block(const InvArgs& a, int jp) {
    extern __shared__ double2 wsm[];
    constexpr int NP = 2 * NL - 1;
    const Geo& g = a.g;
    const Spec& s = a.s;
    const int L = s.L, ZL = zpad_len(L);
    const int nth = (int)blockDim.x;   // 64 (NL + 1)
    const int wv = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    const bool helper = wv == NL;
    double2* tw = wsm + (size_t)NL * ZL;
    __shared__ double coef[2 * MAXL];
    __shared__ double s_ctl2m[MAXL * MAXL];
    const int r0 = 2 * jp, r1 = r0 + 1;
    const bool has1 = r1 < s.nyi;
    const int gj0 = r0 + 1, gj1 = r1 + 1;
    Staged<12, double> stw;
    stage_load(stw, a.twf, 2 * twiddle_table_len(L));
    constexpr int SP = P / 2 + 1;
    double2 cx0[SP], cx1[SP];
    if (!helper) {
        const double* row0 = a.spec + ((long long)wv * s.nrow + r0) * s.kp;
        const double* row1 = row0 + s.kp;
#pragma unroll
        for (int u = 0; u < SP; u++) {
            const int kk = lane + 64 * u;
            const int tc = (kk <= L / 2) ? kk : 0;
            cx0[u] = *(const double2*)(row0 + 2 * tc);
            cx1[u] = *(const double2*)(row1 + 2 * tc);
        }
    }
    double pv[NP];
#pragma unroll
    for (int q = 0; q < NP; q++) pv[q] = 0.0;
    double ctlv = 0.0;
    if (helper) {
        if (lane < MAXL * MAXL) ctlv = a.dc->mx.ctl2m[lane];
        constexpr int NB = 2;   // loaded together, then added in row-pair order
        double pb[NB][NP];
#pragma unroll
        for (int u = 0; u < NB; u++) {
            const int b = lane + 64 * u;
#pragma unroll
            for (int q = 0; q < NP; q++) pb[u][q] = b < a.adj_npart ? a.adj_partial[(long long)b * ADJ_CYC_NPROF + q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < NB; u++)
            if (lane + 64 * u < a.adj_npart)
#pragma unroll
                for (int q = 0; q < NP; q++) pv[q] += pb[u][q];
        for (int b = lane + 64 * NB; b < a.adj_npart; b += 64)
#pragma unroll
            for (int q = 0; q < NP; q++) pv[q] += a.adj_partial[(long long)b * ADJ_CYC_NPROF + q];
    }
    // every thread: Gs, Gn of both rows (row-constant fields, column 0)
    double gs[2][NL], gn[2][NL];
#pragma unroll
    for (int m = 0; m < NL; m++) {
        gs[0][m] = a.adj_G[(long long)m * g.fs + (long long)gj0 * g.pitch];
        gn[0][m] = a.adj_G[(long long)(g.nl + m) * g.fs + (long long)gj0 * g.pitch];
        gs[1][m] = has1 ? a.adj_G[(long long)m * g.fs + (long long)gj1 * g.pitch] : 0.0;
        gn[1][m] = has1 ? a.adj_G[(long long)(g.nl + m) * g.fs + (long long)gj1 * g.pitch] : 0.0;
    }
    stage_store(stw, (double*)tw, a.twf, 2 * twiddle_table_len(L));
    const double fn = s.ftnorm;
    if (!helper) {
        // Hermitian spectra of the two packed rows, as the forward inverse kernel
        double2* z = wsm + (size_t)wv * ZL;
        const int h = L / 2;
#pragma unroll
        for (int u = 0; u < SP; u++) {
            const int kk = lane + 64 * u;
            if (kk <= h) {
                double xr = fn * cx0[u].x, xi = fn * cx0[u].y;
                double yr = has1 ? fn * cx1[u].x : 0.0, yi = has1 ? fn * cx1[u].y : 0.0;
                if (kk == 0 || (L % 2 == 0 && kk == h)) { xi = 0.0; yi = 0.0; }
                z[zpad(kk)] = make_double2(xr - yi, xi + yr);
                if (kk != 0 && !(L % 2 == 0 && kk == h)) z[zpad(L - kk)] = make_double2(xr + yi, -xi + yr);
            }
        }
    } else if (lane < MAXL * MAXL) {
        s_ctl2m[lane] = ctlv;
    }
    __syncthreads();
    if (!helper) {
        fft_wave_any<P, +1>(wsm + (size_t)wv * ZL, L, s.nstages, s.radix, tw, lane);
    } else {
#pragma unroll
        for (int q = 0; q < NP; q++) pv[q] = wave_sum(pv[q]);
        if (lane == 0) {
            const CycHom& ch = a.dc->ch;
            double ls[NL], ln[NL];
            ls[0] = ch.hbsi * pv[0];
            ln[0] = 0.0;
#pragma unroll
            for (int m = 1; m < NL; m++) {
                const double b1 = pv[m], b2 = pv[NL - 1 + m];
                ls[m] = ch.hc2n[m] * b1 - ch.hc1n[m] * b2;
                ln[m] = ch.hc1s[m] * b2 - ch.hc2s[m] * b1;
            }
            const double rxy = g.dx / g.dy;
#pragma unroll
            for (int m = 0; m < NL; m++) {
                coef[m] = rxy * ls[m];
                coef[MAXL + m] = rxy * ln[m];
            }
            if (jp == 0) {
                const Mix& mx = a.dc->mx;
#pragma unroll
                for (int k = 0; k < NL; k++) {
                    double ts = 0.0, tn = 0.0;
#pragma unroll
                    for (int m = 0; m < NL; m++) {
                        ts += mx.ctl2m[k * MAXL + m] * ls[m];
                        tn += mx.ctl2m[k * MAXL + m] * ln[m];
                    }
                    a.adj_lc[k] += ts;
                    a.adj_lc[MAXL + k] += tn;
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int sel = 0; sel < 2; sel++) {
        if (sel == 1 && !has1) break;
        const long long orow = (long long)(sel == 0 ? gj0 : gj1) * g.pitch;
        for (int i = (int)threadIdx.x; i < L; i += nth) {
            double r[NL];
#pragma unroll
            for (int m = 0; m < NL; m++)
                r[m] = inh_at<true>(g, s, wsm + (size_t)m * ZL, i, sel) + coef[m] * gs[sel][m] + coef[MAXL + m] * gn[sel][m];
#pragma unroll
            for (int k = 0; k < NL; k++) {
                double acc = 0.0;
#pragma unroll
                for (int m = 0; m < NL; m++) acc += s_ctl2m[k * MAXL + m] * r[m];
                double* d = a.adj_lq + k * g.fs + orow + i;
                *d = *d + g.fnot * acc;
            }
        }
    }
}

template <int NL, int P>
__global__ __launch_bounds__(64 * (NL + 1)) void k_invw_adj_cyc(InvArgs a) {
    invw_adj_cyc_block<NL, P>(a, blockIdx.x);
}

// Launchers of this object's layer count, over the points per lane (4, 16, 32).
#define OQG_WCAT2(a, b) a##b
#define OQG_WCAT(a, b) OQG_WCAT2(a, b)
#define OQG_WFN(name) OQG_WCAT(name, OQG_WAVE_NL)
constexpr int WNL = OQG_WAVE_NL;

#define OQG_P_SWITCH(PV, CALL)      \
    switch (PV) {                   \
        case 4: CALL(4); break;     \
        case 16: CALL(16); break;   \
        case 32: CALL(32); break;   \
        default: break;             \
    }

void OQG_WFN(launch_fwdw_nl)(const FwdArgs& a, hipStream_t st) {
    const int npairs = (a.s.nyi + 1) / 2;
    const size_t lds = fftw_lds_bytes(a.s, WNL);
#define FWDW_CALL(Pc) hipLaunchKernelGGL((k_fwdw<WNL, Pc>), dim3(npairs), dim3(64 * WNL), lds, st, a)
    OQG_P_SWITCH(fftw_points(a.s), FWDW_CALL)
#undef FWDW_CALL
}

void OQG_WFN(launch_invw_nl)(const InvArgs& a, hipStream_t st) {
    const int npairs = (a.s.nyi + 1) / 2;
    const size_t lds = fftw_lds_bytes(a.s, WNL);
#define INVW_CALL(Pc)                                                                                        \
    if (a.g.cyclic) hipLaunchKernelGGL((k_invw<WNL, Pc, true>), dim3(npairs), dim3(64 * (WNL + 1)), lds, st, a); \
    else hipLaunchKernelGGL((k_invw<WNL, Pc, false>), dim3(npairs), dim3(64 * (WNL + 1)), lds, st, a);
    OQG_P_SWITCH(fftw_points(a.s), INVW_CALL)
#undef INVW_CALL
}

void OQG_WFN(launch_fwdw_ens_nl)(const FwdArgs* args, const int2* map, int nblocks, int p, size_t lds, hipStream_t st) {
#define FWDWE_CALL(Pc) hipLaunchKernelGGL((k_fwdw_ens<WNL, Pc>), dim3(nblocks), dim3(64 * WNL), lds, st, args, map)
    OQG_P_SWITCH(p, FWDWE_CALL)
#undef FWDWE_CALL
}

void OQG_WFN(launch_invw_ens_nl)(const InvArgs* args, const int2* map, int nblocks, int p, bool cyc, size_t lds,
                                 hipStream_t st) {
#define INVWE_CALL(Pc)                                                                                          \
    if (cyc) hipLaunchKernelGGL((k_invw_ens<WNL, Pc, true>), dim3(nblocks), dim3(64 * (WNL + 1)), lds, st, args, map); \
    else hipLaunchKernelGGL((k_invw_ens<WNL, Pc, false>), dim3(nblocks), dim3(64 * (WNL + 1)), lds, st, args, map);
    OQG_P_SWITCH(p, INVWE_CALL)
#undef INVWE_CALL
}

void OQG_WFN(launch_fwdw_adj_nl)(const FwdArgs& a, hipStream_t st) {
    const int npairs = (a.s.nyi + 1) / 2;
    const size_t lds = fftw_lds_bytes(a.s, WNL);
#define FWDWA_CALL(Pc)                                                                                      \
    if (a.g.cyclic) hipLaunchKernelGGL((k_fwdw_adj_cyc<WNL, Pc>), dim3(npairs), dim3(64 * WNL), lds, st, a); \
    else hipLaunchKernelGGL((k_fwdw_adj<WNL, Pc>), dim3(npairs), dim3(64 * WNL), lds, st, a);
    OQG_P_SWITCH(fftw_points(a.s), FWDWA_CALL)
#undef FWDWA_CALL
}

void OQG_WFN(launch_invw_adj_nl)(const InvArgs& a, hipStream_t st) {
    const int npairs = (a.s.nyi + 1) / 2;
    const size_t lds = fftw_lds_bytes(a.s, WNL);
#define INVWA_CALL(Pc)                                                                                              \
    if (a.g.cyclic) hipLaunchKernelGGL((k_invw_adj_cyc<WNL, Pc>), dim3(npairs), dim3(64 * (WNL + 1)), lds, st, a); \
    else hipLaunchKernelGGL((k_invw_adj<WNL, Pc>), dim3(npairs), dim3(64 * (WNL + 1)), lds, st, a);
    OQG_P_SWITCH(fftw_points(a.s), INVWA_CALL)
#undef INVWA_CALL
}

void OQG_WFN(set_wave_attrs_nl)(int lim) {
#define ATTR_P(Pc)                                                                                                    \
    (void)hipFuncSetAttribute((const void*)k_fwdw<WNL, Pc>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);     \
    (void)hipFuncSetAttribute((const void*)k_fwdw_ens<WNL, Pc>, hipFuncAttributeMaxDynamicSharedMemorySize, lim); \
    (void)hipFuncSetAttribute((const void*)k_invw<WNL, Pc, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);     \
    (void)hipFuncSetAttribute((const void*)k_invw<WNL, Pc, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);      \
    (void)hipFuncSetAttribute((const void*)k_invw_ens<WNL, Pc, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim); \
    (void)hipFuncSetAttribute((const void*)k_invw_ens<WNL, Pc, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    ATTR_P(4) ATTR_P(16) ATTR_P(32)
#undef ATTR_P
#define ATTR_P(Pc)                                                                                                    \
    (void)hipFuncSetAttribute((const void*)k_fwdw_adj<WNL, Pc>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);  \
    (void)hipFuncSetAttribute((const void*)k_invw_adj<WNL, Pc>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);  \
    (void)hipFuncSetAttribute((const void*)k_fwdw_adj_cyc<WNL, Pc>, hipFuncAttributeMaxDynamicSharedMemorySize, lim); \
    (void)hipFuncSetAttribute((const void*)k_invw_adj_cyc<WNL, Pc>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    ATTR_P(4) ATTR_P(16) ATTR_P(32)
#undef ATTR_P
}

}  // namespace oqg

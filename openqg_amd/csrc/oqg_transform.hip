// oqg_transform.hip -- PV inversion kernels of the QG step (pressure.F90:21-69).
//
//   k_fwd     layer->mode mix (modes.F90:248-266) fused with the x transform of two
//             interior rows per workgroup: DST-I via the odd-extended complex FFT in
//             the box, r2c in the channel (inhomog.F90:184-185,228-229).  Spectra are
//             stored row-major, spec[m][j][k], so both row kernels move whole rows.
//   k_ysolve  tridiagonal solve along y for each (mode, wavenumber) column
//             (inhomog.F90:188-199,232-243).  A workgroup stages CT columns through
//             LDS (transposing them), then one wavefront per column: each lane
//             eliminates its own block of rows, and the block separators are solved
//             by parallel cyclic reduction across the 64 lanes.  Box: also the
//             column sums that give the area integral of each modal solution.
//   k_ythomas the reference's sequential Thomas recurrence, one thread per column
//             (selected with OQG_FLAG_YSOLVE_THOMAS).
//   k_inv     mass/momentum-constraint correction (homog.F90:280-366,
//             constraint.F90:167-215), inverse x transforms of all modes of two rows
//             (one 256-thread group per mode), mode->layer mix (modes.F90:268-284)
//             and the boundary PV (vorsubs.F90:58-114); writes p_new over pm.

#include <mutex>

#include "oqg_device.h"
#include "oqg_inv_dev.h"

namespace oqg {

// Dispatch to the wavefront-transform objects of a layer count (oqg_wave.hip).
#define OQG_WAVE_NL_SWITCH(NLV, CALL) \
    switch (NLV) {                    \
        case 2: CALL(2); break;       \
        case 3: CALL(3); break;       \
        case 4: CALL(4); break;       \
        case 5: CALL(5); break;       \
        case 6: CALL(6); break;       \
        case 7: CALL(7); break;       \
        case 8: CALL(8); break;       \
        default: break;               \
    }


// ------------------------------------------------------------------ k_fwd

// p_modes_rhs (modes.F90:259-264) at one point from its loaded inputs:
// f0 * sum_k (q_k - beta*y - [k=k_topo] ddyn) ctl2m(k,m), with cm the
// column ctl2m(:,m).
template <int NL>
__device__ __forceinline__ double rhs_mode(const FwdArgs& a, const double (&cm)[NL], const double (&qk)[NL], double by,
                                           double dd) {
    const Geo& g = a.g;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NL; k++) {
        double ql = qk[k] - by;
        if (k == g.kt) ql = ql - dd;
        s += ql * cm[k];
    }
    return g.fnot * s;
}

// w = sum_k cm[k] lp_k at the interior point (i, gj) with lp completed by the
// transpose of ocqbdy (adjoint mode).  The ring points whose inner point it is
// are its west (i = 1), east (i = nxp-2), south (gj = 1) and north
// (gj = nyp-2) neighbours; they are added in bdy_gather's order, without its
// general case analysis (which kept the kernel's registers spilling).
template <int NL>
__device__ __forceinline__ double adj_modal_input(const FwdArgs& a, const double (&cm)[NL], double (&lp)[NL], int i,
                                                  int gj) {
    const Geo& g = a.g;
    const double* lq = a.adj_lq;
    const double qb = a.adj_qbfac;
    const long long o = (long long)gj * g.pitch + i;
    if (i == 1) {
#pragma unroll
        for (int k = 0; k < NL; k++) lp[k] += qb * lq[k * g.fs + o - 1];
    }
    if (i == g.nxp - 2) {
#pragma unroll
        for (int k = 0; k < NL; k++) lp[k] += qb * lq[k * g.fs + o + 1];
    }
    if (gj == 1) {
#pragma unroll
        for (int k = 0; k < NL; k++) lp[k] += qb * lq[k * g.fs + o - g.pitch];
    }
    if (gj == g.nyp - 2) {
#pragma unroll
        for (int k = 0; k < NL; k++) lp[k] += qb * lq[k * g.fs + o + g.pitch];
    }
    double w = 0.0;
#pragma unroll
    for (int k = 0; k < NL; k++) w += cm[k] * lp[k];
    return w;
}

// Extra point e of a row pair's adjoint inner products: the first nside are
// columns 0 and nxp-1 of its rows gj0 (and gj1), the rest its share of the
// ring rows 0 and nyp-1 from ring index rlo on.
__device__ __forceinline__ void extra_point(const Geo& g, int e, int nside, int rlo, int gj0, int gj1, int& i, int& gj) {
    if (e < nside) {
        gj = (e >> 1) ? gj1 : gj0;
        i = (e & 1) ? g.nxp - 1 : 0;
    } else {
        const int idx = rlo + (e - nside);
        gj = idx < g.nxp ? 0 : g.nyp - 1;
        i = idx < g.nxp ? idx : idx - g.nxp;
    }
}

// Row pair jp of mode mb (or of the raw field).  ADJ: the adjoint mode of
// the box inversion (FwdArgs::adj_*), whose input of mode m is
// w_m = sum_k ctm2l(m,k) lp_k.
template <int NL, int MB, bool ADJ = false, bool STAGE_DOUBLES = ADJ, bool PR = false>
__device__ __forceinline__ void fwd_block(const FwdArgs& a, int jp, int mb) {
    extern __shared__ double2 zbuf[];   // [zpad_len(L)] transform, then the twiddle table
    const Geo& g = a.g;
    const Spec& s = a.s;
    const int L = s.L;
    const int m = (!ADJ && a.raw_mode >= 0) ? a.raw_mode : mb;
    const int r0 = 2 * jp, r1 = 2 * jp + 1;
    const bool has1 = r1 < s.nyi;
    const int gj0 = r0 + 1, gj1 = r1 + 1;
    double2* tw = zbuf + zpad_len(L);
    // the twiddle table, staged in double2 (the forward kernel) or as doubles
    // (the adjoint and ensemble forms, where the compiler kept a staged double2
    // array in scratch memory)
    Staged<4, double2> stw2;
    Staged<8, double> stw1;
    if constexpr (STAGE_DOUBLES) stage_load(stw1, a.tw, 2 * twiddle_table_len(L));
    else stage_load(stw2, a.tw, twiddle_table_len(L));
    const bool raw = !ADJ && a.raw_mode >= 0;
    const double* qin = ADJ ? a.adj_lp : a.q;
    // inputs of this thread's (at most four) points of both rows, loaded together
    const int first = g.cyclic ? 0 : 1;
    const int npts = s.nin;   // box: interior points 1..nin; channel: 0..nin-1
    // every load is issued before any is used: q of all layers (one layer in
    // raw mode), ddyn and beta*y of both rows (adjoint mode: lp of all layers
    // and the homogeneous solution of mode m)
    constexpr int NQ = 4;
    const int nlq = raw ? 1 : NL;
    double cm[NL];   // ctl2m(:, m) from the device constants; adjoint mode: ctm2l(m, :)
#pragma unroll
    for (int k = 0; k < NL; k++) cm[k] = ADJ ? a.dc->mx.ctm2l[m * MAXL + k] : a.dc->mx.ctl2m[k * MAXL + m];
    double q0[NQ][NL], q1[NQ][NL], d0[NQ], d1[NQ];
#pragma unroll
    for (int u = 0; u < NQ; u++) {
        const int t = first + threadIdx.x + 256 * u;
        const bool in = t < first + npts;
        const long long o0 = (long long)gj0 * g.pitch + t, o1 = (long long)gj1 * g.pitch + t;
#pragma unroll
        for (int k = 0; k < NL; k++) {
            q0[u][k] = (in && k < nlq) ? qin[k * g.fs + o0] : 0.0;
            q1[u][k] = (in && has1 && k < nlq) ? qin[k * g.fs + o1] : 0.0;
        }
        if (ADJ) {
            const double* hm = a.adj_hom + (long long)(m > 0 ? m - 1 : 0) * g.fs;
            d0[u] = (in && m > 0) ? hm[o0] : 0.0;
            d1[u] = (in && has1 && m > 0) ? hm[o1] : 0.0;
        } else {
            d0[u] = (in && !raw && a.ddyn) ? a.ddyn[o0] : 0.0;
            d1[u] = (in && has1 && !raw && a.ddyn) ? a.ddyn[o1] : 0.0;
        }
    }
    const double by0 = g.beta * a.yprel[gj0], by1 = has1 ? g.beta * a.yprel[gj1] : 0.0;
    // adjoint mode: the other points of the inner products, columns 0 and
    // nxp-1 of the block's rows and a share of the ring rows 0 and nyp-1 (2
    // nxp points dealt out evenly over the row pairs).  All lie on the ring,
    // where the transpose of ocqbdy needs lq at the point and, for the points
    // next to a corner in rows 0 and nyp-1, at that corner; the first such
    // point of each thread is loaded here with the rest.
    const int npairs = (s.nyi + 1) / 2, nring = 2 * g.nxp;
    const int rper = (nring + npairs - 1) / npairs;
    const int rlo = min(nring, jp * rper), rhi = min(nring, rlo + rper);
    const int nside = has1 ? 4 : 2;
    const int nextra = ADJ ? nside + (rhi - rlo) : 0;
    int ei = 0, ej = 0;
    const bool ehas = (int)threadIdx.x < nextra;
    if (ehas) extra_point(g, threadIdx.x, nside, rlo, gj0, gj1, ei, ej);
    const bool erow = ej == 0 || ej == g.nyp - 1;
    const bool ewc = ehas && erow && ei == 1, eec = ehas && erow && ei == g.nxp - 2;
    double elp[NL], elq[NL], elw[NL], ele[NL], ehm = 0.0;
    if constexpr (ADJ) {
        const long long o = (long long)ej * g.pitch + ei;
#pragma unroll
        for (int k = 0; k < NL; k++) {
            elp[k] = ehas ? a.adj_lp[k * g.fs + o] : 0.0;
            elq[k] = ehas ? a.adj_lq[k * g.fs + o] : 0.0;
            elw[k] = ewc ? a.adj_lq[k * g.fs + o - 1] : 0.0;
            ele[k] = eec ? a.adj_lq[k * g.fs + o + 1] : 0.0;
        }
        ehm = (ehas && m > 0) ? a.adj_hom[(long long)(m - 1) * g.fs + o] : 0.0;
    }
    double xv[NQ], yv[NQ];
    double part = 0.0;   // adjoint mode: this thread's share of hom_m . w_m
#pragma unroll
    for (int u = 0; u < NQ; u++) {
        const int t = first + threadIdx.x + 256 * u;
        if (ADJ) {
            const bool in = t < first + npts;
            xv[u] = in ? adj_modal_input<NL>(a, cm, q0[u], t, gj0) : 0.0;
            yv[u] = (in && has1) ? adj_modal_input<NL>(a, cm, q1[u], t, gj1) : 0.0;
        } else if (raw) {
            xv[u] = q0[u][0];
            yv[u] = q1[u][0];
        } else {
            xv[u] = rhs_mode<NL>(a, cm, q0[u], by0, d0[u]);
            yv[u] = has1 ? rhs_mode<NL>(a, cm, q1[u], by1, d1[u]) : 0.0;
        }
        if (t >= first + npts) {
            xv[u] = 0.0;
            yv[u] = 0.0;
        }
        if (ADJ) part += d0[u] * xv[u] + d1[u] * yv[u];
    }
    if constexpr (ADJ) {
        __shared__ double s_red[16];
        if (ehas) {
            // bdy_gather at a ring point, from the loaded values: the point's
            // own term, then its west and east neighbours when they are the
            // corner whose inner point it is
            const Mix& mx = a.dc->mx;
            const double r = ((ei == 0 || ei == g.nxp - 1) && erow) ? 2.0 : 1.0;
#pragma unroll
            for (int k = 0; k < NL; k++) {
#pragma unroll
                for (int l = 0; l < NL; l++) {
                    double c = -r * g.fnot * mx.amat[k * MAXL + l];
                    if (l == k) c -= a.adj_qbfac;
                    elp[l] += elq[k] * c;
                }
            }
#pragma unroll
            for (int k = 0; k < NL; k++)
                if (ewc) elp[k] += a.adj_qbfac * elw[k];
#pragma unroll
            for (int k = 0; k < NL; k++)
                if (eec) elp[k] += a.adj_qbfac * ele[k];
            double w = 0.0;
#pragma unroll
            for (int k = 0; k < NL; k++) w += cm[k] * elp[k];
            if (m > 0) part += ehm * w;
        }
        // points beyond one per thread (very wide, short basins)
        for (int e = threadIdx.x + blockDim.x; e < nextra; e += blockDim.x) {
            int i, gj;
            extra_point(g, e, nside, rlo, gj0, gj1, i, gj);
            const long long o = (long long)gj * g.pitch + i;
            double lp[NL];
#pragma unroll
            for (int k = 0; k < NL; k++) lp[k] = a.adj_lp[k * g.fs + o];
            bdy_gather<NL>(g, a.adj_qbfac, a.dc->mx, a.adj_lq, i, gj, lp);   // a ring point
            double w = 0.0;
#pragma unroll
            for (int k = 0; k < NL; k++) w += cm[k] * lp[k];
            if (m > 0) part += a.adj_hom[(long long)(m - 1) * g.fs + o] * w;
        }
        double pv[1] = {part};
        block_sum<1>(pv, s_red);
        if (threadIdx.x == 0) a.adj_partial[(long long)jp * MAXL + m] = pv[0];
    }
    if constexpr (STAGE_DOUBLES) stage_store(stw1, (double*)tw, a.tw, 2 * twiddle_table_len(L));
    else stage_store(stw2, tw, a.tw, twiddle_table_len(L));
    if (!g.cyclic) {
        // odd extension of the two rows: z_t = x_t + i y_t, z_{L-t} = -z_t, z_0 = z_N = 0
        if (threadIdx.x == 0) {
            zbuf[zpad(0)] = make_double2(0.0, 0.0);
            zbuf[zpad(L / 2)] = make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int t = 1 + threadIdx.x + 256 * u;
            if (t <= npts) {
                zbuf[zpad(t)] = make_double2(xv[u], yv[u]);
                zbuf[zpad(L - t)] = make_double2(-xv[u], -yv[u]);
            }
        }
    } else {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int t = threadIdx.x + 256 * u;
            if (t < npts) zbuf[zpad(t)] = make_double2(xv[u], yv[u]);
        }
    }
    __syncthreads();
    fft_lds<MB, -1, FWD_MAX_RADIX, PR>(zbuf, s.L, s.nstages, s.radix, tw, threadIdx.x);
    double* row0 = a.spec + ((long long)m * s.nrow + r0) * s.kp;
    double* row1 = row0 + s.kp;
    if (!g.cyclic) {
        // DST(x)_K = -Im Z_K (row r0), DST(y)_K = Re Z_K (row r1)
        for (int K = threadIdx.x + 1; K <= s.nin; K += blockDim.x) {
            const double2 v = zbuf[zpad(K)];
            row0[K - 1] = -v.y;
            if (has1) row1[K - 1] = v.x;
        }
    } else {
        const int N = L;
        for (int kk = threadIdx.x; kk <= N / 2; kk += blockDim.x) {
            const double2 A = zbuf[zpad(kk)], B = zbuf[zpad((N - kk) % N)];
            // split the spectra of the two packed real rows
            const double xr = 0.5 * (A.x + B.x), xi = 0.5 * (A.y - B.y);
            const double yr = 0.5 * (A.y + B.y), yi = 0.5 * (B.x - A.x);
            *(double2*)(row0 + 2 * kk) = make_double2(xr, xi);
            if (has1) *(double2*)(row1 + 2 * kk) = make_double2(yr, yi);
        }
    }
}


template <int NL, int MB>
__global__ __launch_bounds__(256) void k_fwd(FwdArgs a) {
    fwd_block<NL, MB>(a, xcd_row_pair(blockIdx.x, gridDim.x), blockIdx.y);
}

template <int NL, int MB>
__global__ __launch_bounds__(256) void k_fwd_adj(FwdArgs a) { fwd_block<NL, MB, true>(a, blockIdx.x, blockIdx.y); }

// Plans with a radix 7, 11 or 13 (transform lengths the wavefront forms and
// the other instantiations do not take): four butterflies per thread cover
// every such plan of the lengths the LDS FFT accepts.
template <int NL>
__global__ __launch_bounds__(256) void k_fwd_pr(FwdArgs a) {
    fwd_block<NL, 4, false, false, true>(a, xcd_row_pair(blockIdx.x, gridDim.x), blockIdx.y);
}

// Ensemble form: block b works on block map[b].y (row pairs fastest, then
// modes) of member map[b].x.
template <int NL, int MB>
__global__ __launch_bounds__(256) void k_fwd_ens(const FwdArgs* __restrict__ args, const int2* __restrict__ map) {
    const int2 e = map[blockIdx.x];
    const FwdArgs& a = args[e.x];
    const int npairs = (a.s.nyi + 1) / 2;
    const int mb = e.y / npairs;
    fwd_block<NL, MB, false, true>(a, e.y - mb * npairs, mb);
}

// Instantiations over the layer count (2..8) and butterflies per thread (1, 2, 4).
#define OQG_NL_MB_SWITCH(NLV, MBV, CALL)                                      \
    switch ((NLV) * 8 + (MBV)) {                                              \
        case 2 * 8 + 1: CALL(2, 1); break; case 2 * 8 + 2: CALL(2, 2); break; case 2 * 8 + 4: CALL(2, 4); break; \
        case 3 * 8 + 1: CALL(3, 1); break; case 3 * 8 + 2: CALL(3, 2); break; case 3 * 8 + 4: CALL(3, 4); break; \
        case 4 * 8 + 1: CALL(4, 1); break; case 4 * 8 + 2: CALL(4, 2); break; case 4 * 8 + 4: CALL(4, 4); break; \
        case 5 * 8 + 1: CALL(5, 1); break; case 5 * 8 + 2: CALL(5, 2); break; case 5 * 8 + 4: CALL(5, 4); break; \
        case 6 * 8 + 1: CALL(6, 1); break; case 6 * 8 + 2: CALL(6, 2); break; case 6 * 8 + 4: CALL(6, 4); break; \
        case 7 * 8 + 1: CALL(7, 1); break; case 7 * 8 + 2: CALL(7, 2); break; case 7 * 8 + 4: CALL(7, 4); break; \
        case 8 * 8 + 1: CALL(8, 1); break; case 8 * 8 + 2: CALL(8, 2); break; case 8 * 8 + 4: CALL(8, 4); break; \
        default: break;                                                       \
    }

int fft_mb(const Spec& s, bool inverse) {
    const int nst = inverse ? s.instages : s.nstages;
    const int* radix = inverse ? s.iradix : s.radix;
    int mx = 1;
    for (int st = 0; st < nst; st++) {
        const int b = (s.L / radix[st] + FFT_GROUP - 1) / FFT_GROUP;
        if (b > mx) mx = b;
    }
    return mx <= 1 ? 1 : (mx <= 2 ? 2 : 4);
}

size_t fwd_lds_bytes(const Spec& s) { return sizeof(double2) * (size_t)(zpad_len(s.L) + twiddle_table_len(s.L)); }

void launch_fwd(const FwdArgs& a, int nmodes, hipStream_t st) {
    if (a.raw_mode < 0 && nmodes == a.g.nl && fftw_fits(a.s, a.g.nl)) {
        launch_fwdw(a, st);
        return;
    }
    const int npairs = (a.s.nyi + 1) / 2;
    const size_t lds = fwd_lds_bytes(a.s);
    if (a.s.prime) {
#define C(N) hipLaunchKernelGGL((k_fwd_pr<N>), dim3(npairs, nmodes), dim3(256), lds, st, a)
        OQG_WAVE_NL_SWITCH(a.g.nl, C)
#undef C
        return;
    }
#define FWD_CALL(NLc, MBc) hipLaunchKernelGGL((k_fwd<NLc, MBc>), dim3(npairs, nmodes), dim3(256), lds, st, a)
    OQG_NL_MB_SWITCH(a.g.nl, fft_mb(a.s, false), FWD_CALL)
#undef FWD_CALL
}

void launch_fwd_adj(const FwdArgs& a, hipStream_t st) {
    if (fftw_fits(a.s, a.g.nl)) {
#define C(N) launch_fwdw_adj_nl##N(a, st)
        OQG_WAVE_NL_SWITCH(a.g.nl, C)
#undef C
        return;
    }
    const int npairs = (a.s.nyi + 1) / 2;
    const size_t lds = fwd_lds_bytes(a.s);
#define FWDA_CALL(NLc, MBc) hipLaunchKernelGGL((k_fwd_adj<NLc, MBc>), dim3(npairs, a.g.nl), dim3(256), lds, st, a)
    OQG_NL_MB_SWITCH(a.g.nl, fft_mb(a.s, false), FWDA_CALL)
#undef FWDA_CALL
}

void launch_fwd_ens(const FwdArgs* args, const int2* map, int nblocks, int nl, int mb, size_t lds, hipStream_t st) {
#define FWDE_CALL(NLc, MBc) hipLaunchKernelGGL((k_fwd_ens<NLc, MBc>), dim3(nblocks), dim3(256), lds, st, args, map)
    OQG_NL_MB_SWITCH(nl, mb, FWDE_CALL)
#undef FWDE_CALL
}

// ------------------------------------------------------------------ k_ysolve

// a*u[j-1] + b*u[j] + a*u[j+1] = f[j], j = 0..nyi-1, u[-1] = u[nyi] = 0.
// Lane l owns rows [l*R, l*R+R).  The last row of each lane's block is its
// separator; the other rows are eliminated locally in terms of the two
// neighbouring separators, and the separators' tridiagonal system is solved
// by parallel cyclic reduction across the 64 lanes.
template <int R>
__device__ void wave_tridiag(double (&f)[R], int L, double aa, double b, int lane) {
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
    // responses of the first / last interior row (v: particular part, x: per unit
    // left separator, y: per unit own separator)
    double vf = 0.0, xf = 0.0, yf = 0.0, vl = 0.0, xl = 0.0, yl = 0.0, fs = 0.0;
    if (L >= 2) {
        vf = v[0]; xf = x[0]; yf = y[0];
#pragma unroll
        for (int t = 0; t < R - 1; t++)
            if (t == L - 2) { vl = v[t]; xl = x[t]; yl = y[t]; }
    } else if (L == 1) {
        yf = 1.0;   // the next block's "first interior row" is its separator
        xl = 1.0;   // the own "last interior row" is the left separator
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
#pragma unroll
    for (int t = 0; t < R; t++)
        if (t < L) f[t] = (t == L - 1) ? S : v[t] + x[t] * Sm + y[t] * S;
}

// LDS slot of row j of a y-solver column: row j = lane*R + t sits at
// t*65 + lane, so a wave's accesses to its rows t are contiguous, and the
// transposing loads and stores (8 columns by consecutive rows) spread over
// the banks through the odd row stride and the column pitch ys_pitch(R).
template <int R>
__device__ __forceinline__ int ys_slot(int j) { return (j % R) * 65 + j / R; }

// Columns [cb*YS_CT, cb*YS_CT + YS_CT) of mode mode0 + mb.
template <int R>
__device__ __forceinline__ void ysolve_block(const YArgs& a, int cb, int mb) {
    extern __shared__ double tile[];   // [YS_CT][ys_pitch(R)]
    const Spec& s = a.s;
    constexpr int P = ys_pitch(R);
    const int m = a.mode0 + mb;
    const int c0 = cb * YS_CT;
    double* S = a.spec + (long long)m * s.nrow * s.kp;
    // nyi <= 64 R, so each of the 64*YS_CT threads loads at most R values: issue them together
    double tmp[R];
#pragma unroll
    for (int t = 0; t < R; t++) {
        const int idx = threadIdx.x + 64 * YS_CT * t;
        const int j = idx / YS_CT, c = idx - j * YS_CT;
        const int col = c0 + c;
        tmp[t] = (j < s.nyi && col < s.ncol) ? S[(long long)j * s.kp + col] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < R; t++) {
        const int idx = threadIdx.x + 64 * YS_CT * t;
        const int j = idx / YS_CT, c = idx - j * YS_CT;
        if (j < s.nyi) tile[c * P + ys_slot<R>(j)] = tmp[t];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, c = threadIdx.x >> 6;
    const int col = c0 + c;
    if (col < s.ncol) {
        const double b = a.bcol[m * s.ncol + col];
        const int rs = lane * R;
        int L = s.nyi - rs;
        L = L < 0 ? 0 : (L > R ? R : L);
        double f[R];
#pragma unroll
        for (int t = 0; t < R; t++) f[t] = (t < L) ? tile[c * P + t * 65 + lane] : 0.0;
        wave_tridiag<R>(f, L, s.a, b, lane);
        double csum = 0.0;
#pragma unroll
        for (int t = 0; t < R; t++)
            if (t < L) {
                tile[c * P + t * 65 + lane] = f[t];
                csum += f[t];
            }
        if (a.colsum) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) csum += __shfl_xor(csum, off);
            if (lane == 0) a.colsum[m * s.ncol + col] = csum;
        }
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < s.nyi * YS_CT; idx += blockDim.x) {
        const int j = idx / YS_CT, cc = idx - j * YS_CT;
        const int cl = c0 + cc;
        if (cl < s.ncol) S[(long long)j * s.kp + cl] = tile[cc * P + ys_slot<R>(j)];
    }
}

template <int R>
__global__ __launch_bounds__(64 * YS_CT) void k_ysolve(YArgs a) {
    // neighbouring column blocks share 128-byte lines of the spectral rows:
    // each XCD takes a contiguous run of column blocks, so a shared line is
    // fetched into one L2
    int cb, mb, bz;
    xcd_block(cb, mb, bz);
    ysolve_block<R>(a, cb, mb);
}

// Ensemble form: block b works on column block map[b].y (column blocks
// fastest, then modes) of member map[b].x.
template <int R>
__global__ __launch_bounds__(64 * YS_CT) void k_ysolve_ens(const YArgs* __restrict__ args,
                                                           const int2* __restrict__ map) {
    const int2 e = map[blockIdx.x];
    const YArgs& a = args[e.x];
    const int ncb = (a.s.ncol + YS_CT - 1) / YS_CT;
    const int mb = e.y / ncb;
    ysolve_block<R>(a, e.y - mb * ncb, mb);
}

// Reference-order Thomas (Numerical Recipes form, inhomog.F90:188-199), one
// thread per column walking down the rows; gam is kept in a scratch array of
// the spectra's shape.
__global__ __launch_bounds__(256) void k_ythomas(YArgs a) {
    const Spec& s = a.s;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nmodes * s.ncol) return;
    const int m = a.mode0 + c / s.ncol;
    const int col = c - (c / s.ncol) * s.ncol;
    double* u = a.spec + (long long)m * s.nrow * s.kp + col;
    double* gam = a.gam + (long long)m * s.nrow * s.kp + col;
    const double b = a.bcol[m * s.ncol + col];
    const double aa = s.a;
    double betinv = 1.0 / b;
    double prev = u[0] * betinv;
    u[0] = prev;
    for (int j = 1; j < s.nyi; j++) {
        const double gj = aa * betinv;
        gam[(long long)j * s.kp] = gj;
        betinv = 1.0 / (b - aa * gj);
        prev = (u[(long long)j * s.kp] - aa * prev) * betinv;
        u[(long long)j * s.kp] = prev;
    }
    double csum = prev;
    for (int j = s.nyi - 2; j >= 0; j--) {
        prev = u[(long long)j * s.kp] - gam[(long long)(j + 1) * s.kp] * prev;
        u[(long long)j * s.kp] = prev;
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
    const dim3 grid((a.s.ncol + YS_CT - 1) / YS_CT, a.nmodes);
    const size_t lds = ysolve_lds_bytes(a.s);
    switch (a.s.R) {
#define YCASE(RR) case RR: hipLaunchKernelGGL(k_ysolve<RR>, grid, dim3(64 * YS_CT), lds, st, a); break;
        YCASE(1) YCASE(2) YCASE(3) YCASE(4) YCASE(5) YCASE(6) YCASE(7) YCASE(8)
        YCASE(9) YCASE(10) YCASE(11) YCASE(12) YCASE(13) YCASE(14) YCASE(15) YCASE(16)
#undef YCASE
        default: break;
    }
}

size_t ysolve_lds_bytes(const Spec& s) { return sizeof(double) * YS_CT * (size_t)ys_pitch(s.R); }

int ysolve_blocks(const Spec& s, int nmodes) { return ((s.ncol + YS_CT - 1) / YS_CT) * nmodes; }

void launch_ysolve_ens(const YArgs* args, const int2* map, int nblocks, int R, size_t lds, hipStream_t st) {
    switch (R) {
#define YCASE(RR) case RR: hipLaunchKernelGGL(k_ysolve_ens<RR>, dim3(nblocks), dim3(64 * YS_CT), lds, st, args, map); break;
        YCASE(1) YCASE(2) YCASE(3) YCASE(4) YCASE(5) YCASE(6) YCASE(7) YCASE(8)
        YCASE(9) YCASE(10) YCASE(11) YCASE(12) YCASE(13) YCASE(14) YCASE(15) YCASE(16)
#undef YCASE
        default: break;
    }
}

// ------------------------------------------------------------------ k_inv

// p(i,gj,:) = sum_m (inhomog + homcor)(i,gj,m) ctm2l(m,:)  (modes.F90:268-284),
// evaluated as sum_m inhomog_m ctm2l(m,:) + sum_m homcor_m ctm2l(m,:): the
// first part needs no correction coefficient, so it can be formed while they
// are computed.  sel: 0/1 interior row from LDS, 2: boundary row (inhomog = 0).
template <int NL>
__device__ __forceinline__ void pmix_inh(const InvArgs& a, const double* inh, const double* sctm2l, int i, int sel,
                                         double (&pk)[NL]) {
    const Geo& g = a.g;
    double vm[NL];
#pragma unroll
    for (int m = 0; m < NL; m++) vm[m] = (sel == 2) ? 0.0 : inh[(m * 2 + sel) * g.pitch + i];
#pragma unroll
    for (int k = 0; k < NL; k++) {
        double t = 0.0;
#pragma unroll
        for (int m = 0; m < NL; m++) t += vm[m] * sctm2l[m * MAXL + k];
        pk[k] = t;
    }
}

// pk += sum_m homcor_m(i,gj) ctm2l(m,:).  hpre: the box homogeneous solutions
// at (i,gj) when already loaded (use_pre).
template <int NL>
__device__ __forceinline__ void pmix_hom(const InvArgs& a, const double* coef, const double* sctm2l, int i, int gj,
                                         double (&pk)[NL], const double (&hpre)[NL > 1 ? NL - 1 : 1], bool use_pre) {
    const Geo& g = a.g;
    double vm[NL];
#pragma unroll
    for (int m = 0; m < NL; m++) {
        double h;
        if (!g.cyclic) {
            h = (m == 0) ? 0.0
                         : coef[m - 1] * (use_pre ? hpre[m - 1]
                                                  : a.hom[(long long)(m - 1) * g.fs + (long long)gj * g.pitch + i]);
        } else if (m == 0) {
            h = coef[0] * a.hom[gj];
        } else {
            const double* b1 = a.hom + g.nyp;
            const double* b2 = b1 + (long long)(NL - 1) * g.nyp;
            h = coef[1 + m] * b1[(m - 1) * g.nyp + gj] + coef[1 + MAXL + m] * b2[(m - 1) * g.nyp + gj];
        }
        vm[m] = h;
    }
#pragma unroll
    for (int k = 0; k < NL; k++) {
        double t = 0.0;
#pragma unroll
        for (int m = 0; m < NL; m++) t += vm[m] * sctm2l[m * MAXL + k];
        pk[k] = pk[k] + t;
    }
}

template <int NL>
__device__ __forceinline__ void pmix(const InvArgs& a, const double* inh, const double* coef,
                                     const double* sctm2l, int i, int sel, int gj, double (&pk)[NL],
                                     const double (&hpre)[NL > 1 ? NL - 1 : 1], bool use_pre) {
    pmix_inh<NL>(a, inh, sctm2l, i, sel, pk);
    pmix_hom<NL>(a, coef, sctm2l, i, gj, pk, hpre, use_pre);
}

// ocqbdy (vorsubs.F90:58-114) at one boundary point: kind 0 zonal row, 1 box
// meridional column, 2 box corner (x difference, the row and column terms twice).
template <int NL>
__device__ __forceinline__ void qbdy(const InvArgs& a, const double* samat, double qbfac, int i, int gj, int kind,
                                     const double (&pb)[NL], const double (&pin)[NL]) {
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
        for (int r = 0; r < reps; r++) qv = qv + g.beta * a.yprel[gj];
        if (k == g.kt && a.ddyn)
            for (int r = 0; r < reps; r++) qv = qv + a.ddyn[o];
        a.qnew[k * g.fs + o] = qv;
    }
}

// Transforms in flight per k_inv block: all modes for nl <= 3, else two at a time
// (keeps the register budget of a 1024-thread block out of the way).
constexpr int inv_group_count(int nm) { return nm <= 3 ? nm : 2; }

// This thread's spectral inputs of the row pair (r0, r0+1) of mode m, for
// points t = gt + FFT_GROUP*u: a box row holds real values (bx0/bx1 from the
// two rows), a channel row complex pairs (cx/cy).  Every load is unconditional
// from a valid address (clamped to the first entry of mode 0 when out of
// range; row r0+1 exists even when it is not a real row, as the padding row)
// and nothing is computed from the values here, so that the block's loads all
// issue before any result is waited on.  Entries outside the row, and the
// second row when it is padding, hold junk that the caller must not use.
struct SpecIn {
    double bx0[4], bx1[4];
    double2 cx[4], cy[4];
};

__device__ __forceinline__ void load_spec_pair(const InvArgs& a, int m, bool ok, int r0, int gt, SpecIn& in) {
    const Spec& s = a.s;
    const double* row0 = a.spec + ((long long)(ok ? m : 0) * s.nrow + (ok ? r0 : 0)) * s.kp;
    const double* row1 = row0 + s.kp;
    if (!a.g.cyclic) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int t = gt + FFT_GROUP * u;
            const int tc = (t < s.nin) ? t : 0;
            in.bx0[u] = row0[tc];
            in.bx1[u] = row1[tc];
        }
    } else {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int t = gt + FFT_GROUP * u;
            const int tc = (t <= s.L / 2) ? t : 0;
            in.cx[u] = *(const double2*)(row0 + 2 * tc);
            in.cy[u] = *(const double2*)(row1 + 2 * tc);
        }
    }
}

// Row pair jp.  ADJ: the adjoint mode of the box inversion (InvArgs::adj_*).
template <int NL, int MB, bool ADJ = false, bool PR = false>
__device__ __forceinline__ void inv_block(const InvArgs& a, int jp) {
    extern __shared__ double smem[];
    const Geo& g = a.g;
    const Spec& s = a.s;
    const int L = s.L;
    const bool raw1 = !ADJ && a.raw_mode >= 0, raw = raw1 || ADJ;
    const int nm = raw1 ? 1 : NL;
    const int G = blockDim.x / FFT_GROUP;          // transforms in flight
    const int grp = threadIdx.x / FFT_GROUP, gt = threadIdx.x - grp * FFT_GROUP;
    double2* fbuf = (double2*)smem;                           // [G][zpad_len(L)]
    double* inh = (double*)(fbuf + (size_t)G * zpad_len(L));  // [nm][2][pitch]
    double2* tw = (double2*)(inh + (size_t)nm * 2 * g.pitch);   // twiddle table
    __shared__ double red[(NL + 1) * 16];
    __shared__ double coef[3 * MAXL + 4];
    __shared__ DevConsts s_dc;
    __shared__ double s_cin[C_LEN];
    __shared__ double s_aux[A_LEN];
    __shared__ double s_spk[2 * MAXL];
    // the precompute, the twiddles and the constraint inputs of the correction
    // coefficients are staged through registers with the block's other global
    // loads (one memory round trip) and land in LDS, so the serial
    // coefficient code below reads no global memory
    const int nw_dc = (int)(sizeof(DevConsts) / 8);
    const int nw_cin = raw ? 0 : C_LEN;
    Staged<1> st_dc, st_cin;
    Staged<2, double2> st_tw;   // the whole table for L <= 2*blockDim
    stage_load(st_dc, a.dc, nw_dc);
    stage_load(st_tw, a.tw, twiddle_table_len(L));
    stage_load(st_cin, nw_cin ? (const void*)a.cin : (const void*)a.dc, nw_cin);   // a.cin is null in raw mode
    // channel: thread t < 2 NL CQ_N sums quantity t of k_vort's per-tile partial
    // boundary integrals over the tiles of a row, in tile order
    __shared__ double s_csum[2 * MAXL * CQ_N];
    __shared__ double s_hin[4 * (NL > 1 ? NL - 1 : 1)];   // [row][side][m], see the correction step
    const int ncsum = (!raw && g.cyclic) ? 2 * NL * CQ_N : 0;
    const bool do_csum = (int)threadIdx.x < ncsum;
    constexpr int CSTAGE = 4;   // tiles loaded with the first batch; wider channels loop for the rest
    double cpv[CSTAGE];
    const double* cp = a.cpart;
    if (do_csum) {
        const int ks = (int)threadIdx.x / CQ_N, qy = (int)threadIdx.x - ks * CQ_N;
        cp = a.cpart + (long long)ks * a.ctiles * CYC_SLOTS + qy;
    }
#pragma unroll
    for (int t = 0; t < CSTAGE; t++) cpv[t] = (do_csum && t < a.ctiles) ? cp[t * CYC_SLOTS] : 0.0;
    const double txi_s = (!raw && g.cyclic) ? a.fcon[F_TXIS] : 0.0, txi_n = (!raw && g.cyclic) ? a.fcon[F_TXIN] : 0.0;
    const bool want_spk = !raw && g.cyclic && threadIdx.x < 2 * NL;
    const int m_spk = want_spk ? (int)(threadIdx.x >> 1) : 0;
    const int r_spk = (want_spk && (threadIdx.x & 1)) ? s.nyi - 1 : 0;
    const double spk = a.spec[((long long)m_spk * s.nrow + r_spk) * s.kp];   // used only when want_spk
    const double* s_ctm2l = s_dc.mx.ctm2l;
    const double* s_amat = s_dc.mx.amat;
    const int r0 = 2 * jp, r1 = 2 * jp + 1;
    const bool has1 = r1 < s.nyi;
    const int gj0 = r0 + 1, gj1 = r1 + 1;
    const bool first = (jp == 0), last = (jp == (s.nyi + 1) / 2 - 1);
    const int nsel = has1 ? 2 : 1;

    // ---- issue the block's global loads together: the integrals behind the
    // correction coefficients, the spectra of the first round of modes, and the
    // homogeneous solutions under this thread's first two output points
    double v[NL + 1];
#pragma unroll
    for (int t = 0; t <= NL; t++) v[t] = 0.0;
    double entv[2], cw[2], csv[2][NL];
#pragma unroll
    for (int u = 0; u < 2; u++) {
        const int t = threadIdx.x + blockDim.x * u;
        const bool oke = !raw && t < a.nentpart, okc = !raw && !g.cyclic && t < s.ncol;
        entv[u] = oke ? a.entpart[t] : 0.0;
        cw[u] = okc ? a.colw[t] : 0.0;
#pragma unroll
        for (int m = 0; m < NL; m++) csv[u][m] = okc ? a.colsum[m * s.ncol + t] : 0.0;
    }
    SpecIn sin;
    load_spec_pair(a, raw1 ? a.raw_mode : grp, grp < nm, r0, gt, sin);
    double hv[2][NL > 1 ? NL - 1 : 1];
#pragma unroll
    for (int u = 0; u < 2; u++) {
        const int idx = threadIdx.x + blockDim.x * u;
        const int sel = idx / g.nxp, i = idx - sel * g.nxp;
        const int gj = sel == 0 ? gj0 : gj1;
#pragma unroll
        for (int m = 1; m < NL; m++)
            hv[u][m - 1] = (!raw && !g.cyclic && idx < nsel * g.nxp)
                               ? a.hom[(long long)(m - 1) * g.fs + (long long)gj * g.pitch + i] : 0.0;
    }
    stage_store(st_dc, &s_dc, a.dc, nw_dc);
    stage_store(st_tw, tw, a.tw, twiddle_table_len(L));
    stage_store(st_cin, s_cin, a.cin, nw_cin);
    if (do_csum) {
        double t = cpv[0];
#pragma unroll
        for (int u = 1; u < CSTAGE; u++)
            if (u < a.ctiles) t += cpv[u];
        for (int u = CSTAGE; u < a.ctiles; u++) t += cp[u * CYC_SLOTS];
        s_csum[threadIdx.x] = t;
    }
    if (want_spk) s_spk[threadIdx.x] = spk;

    // this thread's share of the integrals behind the correction coefficients,
    // summed over each wavefront now (which frees the registers during the
    // transforms); the block's sum of the wavefront sums follows them
    if (!raw) {
        v[NL] = entv[0] + entv[1];
        for (int t = threadIdx.x + 2 * blockDim.x; t < a.nentpart; t += blockDim.x) v[NL] += a.entpart[t];
        if (!g.cyclic) {
#pragma unroll
            for (int m = 0; m < NL; m++) v[m] = cw[0] * csv[0][m] + cw[1] * csv[1][m];
            for (int col = threadIdx.x + 2 * blockDim.x; col < s.ncol; col += blockDim.x) {
                const double w = a.colw[col];
#pragma unroll
                for (int m = 0; m < NL; m++) v[m] += w * a.colsum[m * s.ncol + col];
            }
        }
        block_sum_post<NL + 1>(v, red);
    }

    // ---- inverse x transforms, G modes at a time (their barriers publish the LDS copies above)
    const double fn = s.ftnorm;
    for (int m0 = 0; m0 < nm; m0 += G) {
        const int mi = m0 + grp;
        const bool valid = mi < nm;
        const int m = raw1 ? a.raw_mode : mi;
        double2* z = fbuf + (size_t)grp * zpad_len(L);
        if (m0 > 0) load_spec_pair(a, m, valid, r0, gt, sin);   // later rounds (nl > 3)
        if (valid) {
            if (!g.cyclic) {
                // odd extension of ftnorm*u for the two rows (inhomog.F90:202-203)
                if (gt == 0) {
                    z[zpad(0)] = make_double2(0.0, 0.0);
                    z[zpad(L / 2)] = make_double2(0.0, 0.0);
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int t = gt + FFT_GROUP * u;   // spectral index K-1
                    if (t < s.nin) {
                        const double2 w = make_double2(fn * sin.bx0[u], has1 ? fn * sin.bx1[u] : 0.0);
                        z[zpad(t + 1)] = w;
                        z[zpad(L - t - 1)] = make_double2(-w.x, -w.y);
                    }
                }
            } else {
                // Hermitian spectra of the two packed rows (inhomog.F90:246-247)
                const int N = L, h = N / 2;
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int kk = gt + FFT_GROUP * u;
                    if (kk <= h) {
                        double xr = fn * sin.cx[u].x, xi = fn * sin.cx[u].y;
                        double yr = has1 ? fn * sin.cy[u].x : 0.0, yi = has1 ? fn * sin.cy[u].y : 0.0;
                        if (kk == 0 || (N % 2 == 0 && kk == h)) { xi = 0.0; yi = 0.0; }
                        z[zpad(kk)] = make_double2(xr - yi, xi + yr);
                        if (kk != 0 && !(N % 2 == 0 && kk == h)) z[zpad(N - kk)] = make_double2(xr + yi, -xi + yr);
                    }
                }
            }
        }
        __syncthreads();
        if (g.cyclic) fft_lds<MB, +1, INV_MAX_RADIX, PR>(z, s.L, s.instages, s.iradix, tw, gt);   // c2r (inhomog.F90:247)
        else fft_lds<MB, -1, INV_MAX_RADIX, PR>(z, s.L, s.instages, s.iradix, tw, gt);
        if (valid) {
            double* o0 = inh + (size_t)mi * 2 * g.pitch;
            double* o1 = o0 + g.pitch;
            for (int i = gt; i < g.nxp; i += FFT_GROUP) {
                double x0 = 0.0, x1 = 0.0;
                if (!g.cyclic) {
                    if (i >= 1 && i <= s.nin) { const double2 v = z[zpad(i)]; x0 = -v.y; x1 = v.x; }
                } else {
                    const double2 v = z[zpad((i == g.nxp - 1) ? 0 : i)];
                    x0 = v.x;
                    x1 = v.y;
                }
                o0[i] = x0;
                o1[i] = x1;
            }
        }
        __syncthreads();
    }

    // The first and last row pairs also write the zonal boundary rows.  The
    // homogeneous-solution values these need (boundary row, and the row or
    // corner column that ocqbdy pairs with it) are fetched into the transform
    // buffers, now free, by waves 1.. while thread 0 computes the correction
    // coefficients.  Box, nl <= 3, where the buffers hold both sides:
    // zh[side][0: boundary, 1: inner][m][i].
    const bool zone = !raw && !g.cyclic && NL <= 3 && (first || last);
    double* zh = (double*)fbuf;
    auto zone_fetch = [&]() {
        for (int t = (int)threadIdx.x - WAVE; t >= 0 && t < 2 * g.nxp; t += (int)blockDim.x - WAVE) {
            const int side = t / g.nxp, i = t - side * g.nxp;
            if ((side == 0 && !first) || (side == 1 && !last)) continue;
            const int gj = side == 0 ? 0 : g.nyp - 1, gin = side == 0 ? 1 : g.nyp - 2;
            const bool corner = i == 0 || i == g.nxp - 1;
            const long long ob = (long long)gj * g.pitch + i;
            const long long oi = corner ? (long long)gj * g.pitch + (i == 0 ? 1 : g.nxp - 2) : (long long)gin * g.pitch + i;
            double hb[NL > 1 ? NL - 1 : 1], hi[NL > 1 ? NL - 1 : 1];
#pragma unroll
            for (int m = 0; m < NL - 1; m++) {
                hb[m] = a.hom[(long long)m * g.fs + ob];
                hi[m] = a.hom[(long long)m * g.fs + oi];
            }
#pragma unroll
            for (int m = 0; m < NL - 1; m++) {
                zh[((side * 2 + 0) * (NL - 1) + m) * g.nxp + i] = hb[m];
                zh[((side * 2 + 1) * (NL - 1) + m) * g.nxp + i] = hi[m];
            }
        }
    };

    // ---- correction coefficients (identical in every block)
    if (!raw && g.cyclic && (int)threadIdx.x < NL) {
        // the boundary integrals of layer k from the summed partials (k_vort)
        const int k = threadIdx.x;
        const double dx = g.dx;
        auto S = [&](int side, int q) { return s_csum[(k * 2 + side) * CQ_N + q]; };
        s_aux[A_AP3S + k] = s_dc.ph.ah2[k] * (S(0, CQ_D2) * dx) * s_dc.ph.rdy;
        s_aux[A_AP3N + k] = s_dc.ph.ah2[k] * (S(1, CQ_D2) * dx) * s_dc.ph.rdy;
        s_aux[A_AP5S + k] = s_dc.ph.ah4[k] * (S(0, CQ_D4) * dx) * s_dc.ph.rdy;
        s_aux[A_AP5N + k] = s_dc.ph.ah4[k] * (S(1, CQ_D4) * dx) * s_dc.ph.rdy;
        const double aj5s = (S(0, CQ_AJ5) * dx) * s_dc.ph.rdx2, aj9s = (S(0, CQ_AJ9) * dx) * s_dc.ph.rdx2;
        const double aj5n = (S(1, CQ_AJ5) * dx) * s_dc.ph.rdx2, aj9n = (S(1, CQ_AJ9) * dx) * s_dc.ph.rdx2;
        s_aux[A_AJIS + k] = (aj5s + 2.0 * aj9s) * (1.0 / 6.0);
        s_aux[A_AJIN + k] = -((aj5n + 2.0 * aj9n) * (1.0 / 6.0));
        if (k >= 2) {   // layers 1 and 2 are written by thread 0 below
            s_aux[A_BES + k] = 0.0;
            s_aux[A_BEN + k] = 0.0;
        }
        if (k == g.kt) {
            s_aux[A_BDRS] = s_dc.ph.cd_bdr * ((S(0, CQ_BD) * dx) * s_dc.ph.rdy);
            s_aux[A_BDRN] = s_dc.ph.cd_bdr * ((S(1, CQ_BD) * dx) * s_dc.ph.rdy);
        }
        if (k == 0) {
            // int_Be_s/n (vorticity.F90:86-93): fohfac(2) in both northern terms
            const double dz = (double)s_dc.ph.dz_sign;
            s_aux[A_BES + 0] = dz * s_dc.ph.fohfac[0] * (S(0, CQ_E1) * dx);
            s_aux[A_BES + 1] = dz * s_dc.ph.fohfac[0] * (S(0, CQ_E2) * dx);
            s_aux[A_BEN + 0] = -dz * s_dc.ph.fohfac[1] * (S(1, CQ_E1) * dx);
            s_aux[A_BEN + 1] = -dz * s_dc.ph.fohfac[1] * (S(1, CQ_E2) * dx);
            s_aux[A_TXIS] = txi_s;
            s_aux[A_TXIN] = txi_n;
        }
    }
    // thread 0 computes them while the other threads form the coefficient-free
    // part of the mode->layer mix at their first two points
    double pinh[2][NL];
    if (!raw) {
        block_sum_finish<NL + 1>(v, red);
        if (threadIdx.x == 0) {
            step_coefficients<NL>(a, s_dc.mx, s_dc.bh, s_dc.ch, s_dc.ph, s_cin, s_aux, s_spk, v, coef, jp == 0);
        } else {
            // box: the homogeneous solutions at the inner points (columns 1 and
            // nxp-2) of the block's rows, which ocqbdy pairs with the side walls
            const int t = (int)threadIdx.x - WAVE;
            if (!g.cyclic && t >= 0 && t < 4 * (NL - 1)) {
                const int sel = t / (2 * (NL - 1)), side = (t / (NL - 1)) & 1, m = t % (NL - 1);
                const int gj = sel == 0 ? gj0 : gj1, i = side == 0 ? 1 : g.nxp - 2;
                if (sel == 0 || has1) s_hin[t] = a.hom[(long long)m * g.fs + (long long)gj * g.pitch + i];
            }
            if (zone) zone_fetch();
        }
        if (!ADJ) {
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const int idx = threadIdx.x + blockDim.x * u;
                const int sel = idx / g.nxp, i = idx - sel * g.nxp;
                if (idx < nsel * g.nxp) pmix_inh<NL>(a, inh, s_ctm2l, i, sel, pinh[u]);
            }
        }
        __syncthreads();
    }

    if constexpr (ADJ) {
        // The forward solves cdhoc alpha = dpi - (ik_{l+1} - ik_l), ik_k =
        // sum_m ctm2l(m,k) I_m, I_m the area integral of inhomog_m (interior
        // weight dx dy).  Transposed, with la_{m-1} the sum over row pairs of
        // the partials hom_m . w_m: lr = cdhoc^-T la and the constant
        // s_m = -dx dy sum_l lr_l (ctm2l(m,l+1) - ctm2l(m,l)) on the interior
        // of w_m, whose Helmholtz solve is s_m G_m.  The same sums in the same
        // order in every block.
        double la[NL];
#pragma unroll
        for (int mm = 0; mm < NL; mm++) la[mm] = 0.0;
        for (int b = threadIdx.x; b < a.adj_npart; b += blockDim.x)
#pragma unroll
            for (int mm = 0; mm < NL; mm++) la[mm] += a.adj_partial[(long long)b * MAXL + mm];
        block_sum<NL>(la, red);
        if (threadIdx.x == 0) {
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
                for (int l = 0; l < NL - 1; l++)
                    sacc -= lr[l] * (s_ctm2l[mm * MAXL + l + 1] - s_ctm2l[mm * MAXL + l]);
                coef[mm] = (sacc * g.dx) * g.dy;
            }
        }
        __syncthreads();
        // lq_k += f0 sum_m ctl2m(k,m) (u_m + s_m G_m) on the block's interior points
        for (int idx = threadIdx.x; idx < nsel * g.nxp; idx += blockDim.x) {
            const int sel = idx / g.nxp, i = idx - sel * g.nxp;
            if (i < 1 || i > g.nxp - 2) continue;
            const int gj = sel == 0 ? gj0 : gj1;
            const long long o = (long long)gj * g.pitch + i;
            double r[NL], lq[NL];
#pragma unroll
            for (int mm = 0; mm < NL; mm++) {
                r[mm] = a.adj_G[mm * g.fs + o];
                lq[mm] = a.adj_lq[mm * g.fs + o];
            }
#pragma unroll
            for (int mm = 0; mm < NL; mm++) r[mm] = inh[(size_t)mm * 2 * g.pitch + sel * g.pitch + i] + coef[mm] * r[mm];
#pragma unroll
            for (int k = 0; k < NL; k++) {
                double acc = 0.0;
#pragma unroll
                for (int mm = 0; mm < NL; mm++) acc += s_dc.mx.ctl2m[k * MAXL + mm] * r[mm];
                a.adj_lq[k * g.fs + o] = lq[k] + g.fnot * acc;
            }
        }
        return;
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

    // ---- p_new on the interior rows, boundary q on their meridional ends
    double hnone[NL > 1 ? NL - 1 : 1];
#pragma unroll
    for (int m = 0; m < (NL > 1 ? NL - 1 : 1); m++) hnone[m] = 0.0;
    auto interior_point = [&](int idx, const double (&hp)[NL > 1 ? NL - 1 : 1], bool use_pre) {
        const int sel = idx / g.nxp, i = idx - sel * g.nxp;
        const int gj = sel == 0 ? gj0 : gj1;
        double pk[NL];
        pmix<NL>(a, inh, coef, s_ctm2l, i, sel, gj, pk, hp, use_pre);
        const long long o = (long long)gj * g.pitch + i;
#pragma unroll
        for (int k = 0; k < NL; k++) a.pnew[k * g.fs + o] = pk[k];
        if (!g.cyclic && (i == 0 || i == g.nxp - 1)) {
            double pin[NL];
            pmix<NL>(a, inh, coef, s_ctm2l, i == 0 ? 1 : g.nxp - 2, sel, gj, pin, hnone, false);
            qbdy<NL>(a, s_amat, s_dc.ph.qbfac, i, gj, 1, pk, pin);
        }
    };
#pragma unroll
    for (int u = 0; u < 2; u++) {
        const int idx = threadIdx.x + blockDim.x * u;
        if (idx < nsel * g.nxp) {
            const int sel = idx / g.nxp, i = idx - sel * g.nxp;
            const int gj = sel == 0 ? gj0 : gj1;
            double pk[NL];
#pragma unroll
            for (int k = 0; k < NL; k++) pk[k] = pinh[u][k];
            pmix_hom<NL>(a, coef, s_ctm2l, i, gj, pk, hv[u], true);
            const long long o = (long long)gj * g.pitch + i;
#pragma unroll
            for (int k = 0; k < NL; k++) a.pnew[k * g.fs + o] = pk[k];
            if (!g.cyclic && (i == 0 || i == g.nxp - 1)) {
                double pin[NL], hin[NL > 1 ? NL - 1 : 1];
                const int side = i == 0 ? 0 : 1;
#pragma unroll
                for (int m = 0; m < NL - 1; m++) hin[m] = s_hin[(sel * 2 + side) * (NL - 1) + m];
                pmix<NL>(a, inh, coef, s_ctm2l, i == 0 ? 1 : g.nxp - 2, sel, gj, pin, hin, true);
                qbdy<NL>(a, s_amat, s_dc.ph.qbfac, i, gj, 1, pk, pin);
            }
        }
    }
    for (int idx = threadIdx.x + 2 * blockDim.x; idx < nsel * g.nxp; idx += blockDim.x)
        interior_point(idx, hnone, false);
    // ---- zonal boundary rows j = 0 (first block) and j = nyp-1 (last block)
    for (int side = 0; side < 2; side++) {
        if ((side == 0 && !first) || (side == 1 && !last)) continue;
        const int gj = side == 0 ? 0 : g.nyp - 1;
        const int gin = side == 0 ? 1 : g.nyp - 2;
        const int selin = side == 0 ? 0 : ((gin == gj1 && has1) ? 1 : 0);
        for (int i = threadIdx.x; i < g.nxp; i += blockDim.x) {
            double pb[NL], pin[NL];
            double hb[NL > 1 ? NL - 1 : 1], hi[NL > 1 ? NL - 1 : 1];   // fetched above (zone)
#pragma unroll
            for (int m = 0; m < (NL > 1 ? NL - 1 : 1); m++) {
                hb[m] = zone ? zh[((side * 2 + 0) * (NL - 1) + m) * g.nxp + i] : 0.0;
                hi[m] = zone ? zh[((side * 2 + 1) * (NL - 1) + m) * g.nxp + i] : 0.0;
            }
            pmix<NL>(a, inh, coef, s_ctm2l, i, 2, gj, pb, hb, zone);
            const long long o = (long long)gj * g.pitch + i;
#pragma unroll
            for (int k = 0; k < NL; k++) a.pnew[k * g.fs + o] = pb[k];
            if (!g.cyclic && (i == 0 || i == g.nxp - 1)) {
                // corners take the x difference (vorsubs.F90:83-84 overwrite :79-80)
                pmix<NL>(a, inh, coef, s_ctm2l, i == 0 ? 1 : g.nxp - 2, 2, gj, pin, hi, zone);
                qbdy<NL>(a, s_amat, s_dc.ph.qbfac, i, gj, 2, pb, pin);
            } else {
                pmix<NL>(a, inh, coef, s_ctm2l, i, selin, gin, pin, hi, zone);
                qbdy<NL>(a, s_amat, s_dc.ph.qbfac, i, gj, 0, pb, pin);
            }
        }
    }
}

template <int NL, int MB>
__global__ __launch_bounds__(FFT_GROUP * inv_group_count(NL)) void k_inv(InvArgs a) {
    inv_block<NL, MB>(a, xcd_row_pair(blockIdx.x, gridDim.x));
}

template <int NL, int MB>
__global__ __launch_bounds__(FFT_GROUP * inv_group_count(NL)) void k_inv_adj(InvArgs a) {
    inv_block<NL, MB, true>(a, blockIdx.x);
}

template <int NL>
__global__ __launch_bounds__(FFT_GROUP * inv_group_count(NL)) void k_inv_pr(InvArgs a) {
    inv_block<NL, 4, false, true>(a, xcd_row_pair(blockIdx.x, gridDim.x));
}

// Ensemble form: block b works on row pair map[b].y of member map[b].x.
template <int NL, int MB>
__global__ __launch_bounds__(FFT_GROUP * inv_group_count(NL)) void k_inv_ens(const InvArgs* __restrict__ args,
                                                                              const int2* __restrict__ map) {
    const int2 e = map[blockIdx.x];
    inv_block<NL, MB>(args[e.x], e.y);
}

// ------------------------------------------------------------------ wavefront-transform forms (oqg_wave.hip)

int fftw_points(const Spec& s) { return s.L <= 256 ? 4 : (s.L <= 1024 ? 16 : 32); }

size_t fftw_lds_bytes(const Spec& s, int nl) {
    return sizeof(double2) * ((size_t)nl * zpad_len(s.L) + twiddle_table_len(s.L));
}

bool fftw_fits(const Spec& s, int nl) { return !s.prime && fftw_lds_bytes(s, nl) <= MAX_DYN_LDS; }

void launch_fwdw(const FwdArgs& a, hipStream_t st) {
#define C(N) launch_fwdw_nl##N(a, st)
    OQG_WAVE_NL_SWITCH(a.g.nl, C)
#undef C
}

void launch_invw(const InvArgs& a, hipStream_t st) {
#define C(N) launch_invw_nl##N(a, st)
    OQG_WAVE_NL_SWITCH(a.g.nl, C)
#undef C
}

void launch_fwdw_ens(const FwdArgs* args, const int2* map, int nblocks, int nl, int p, size_t lds, hipStream_t st) {
#define C(N) launch_fwdw_ens_nl##N(args, map, nblocks, p, lds, st)
    OQG_WAVE_NL_SWITCH(nl, C)
#undef C
}

void launch_invw_ens(const InvArgs* args, const int2* map, int nblocks, int nl, int p, bool cyc, size_t lds,
                     hipStream_t st) {
#define C(N) launch_invw_ens_nl##N(args, map, nblocks, p, cyc, lds, st)
    OQG_WAVE_NL_SWITCH(nl, C)
#undef C
}

int inv_block_count(const Spec& s) { return (s.nyi + 1) / 2; }

static int inv_groups(const InvArgs& a) { return inv_group_count((a.raw_mode >= 0) ? 1 : a.g.nl); }

size_t inv_lds_bytes(const Spec& s, const Geo& g, int nm) {
    return sizeof(double2) * ((size_t)inv_group_count(nm) * zpad_len(s.L) + twiddle_table_len(s.L)) +
           sizeof(double) * 2 * (size_t)nm * g.pitch;
}

void launch_inv(const InvArgs& a, hipStream_t st) {
    if (a.raw_mode < 0 && fftw_fits(a.s, a.g.nl)) {
        launch_invw(a, st);
        return;
    }
    const int nm = (a.raw_mode >= 0) ? 1 : a.g.nl;
    const int G = inv_groups(a);
    const dim3 grid(inv_block_count(a.s)), block(FFT_GROUP * G);
    const size_t lds = inv_lds_bytes(a.s, a.g, nm);
    if (a.s.prime) {
#define C(N) hipLaunchKernelGGL((k_inv_pr<N>), grid, block, lds, st, a)
        OQG_WAVE_NL_SWITCH(a.g.nl, C)
#undef C
        return;
    }
#define INV_CALL(NLc, MBc) hipLaunchKernelGGL((k_inv<NLc, MBc>), grid, block, lds, st, a)
    OQG_NL_MB_SWITCH(a.g.nl, fft_mb(a.s, true), INV_CALL)
#undef INV_CALL
}

void launch_inv_adj(const InvArgs& a, hipStream_t st) {
    if (fftw_fits(a.s, a.g.nl)) {
#define C(N) launch_invw_adj_nl##N(a, st)
        OQG_WAVE_NL_SWITCH(a.g.nl, C)
#undef C
        return;
    }
    const dim3 grid(inv_block_count(a.s)), block(FFT_GROUP * inv_group_count(a.g.nl));
    const size_t lds = inv_lds_bytes(a.s, a.g, a.g.nl);
#define INVA_CALL(NLc, MBc) hipLaunchKernelGGL((k_inv_adj<NLc, MBc>), grid, block, lds, st, a)
    OQG_NL_MB_SWITCH(a.g.nl, fft_mb(a.s, true), INVA_CALL)
#undef INVA_CALL
}

void launch_inv_ens(const InvArgs* args, const int2* map, int nblocks, int nl, int mb, size_t lds, hipStream_t st) {
    const dim3 block(FFT_GROUP * inv_group_count(nl));
#define INVE_CALL(NLc, MBc) hipLaunchKernelGGL((k_inv_ens<NLc, MBc>), dim3(nblocks), block, lds, st, args, map)
    OQG_NL_MB_SWITCH(nl, mb, INVE_CALL)
#undef INVE_CALL
}

// Every instantiation of the LDS-staged kernels may use up to MAX_DYN_LDS of
// dynamic LDS (each launch asks for what its basins need; the setup of a basin
// checks that this fits).  Set once per process, for all basins alike.
void set_kernel_attrs() {
    static std::once_flag once;
    std::call_once(once, [] {
        const int lim = (int)MAX_DYN_LDS;
#define ATTR_NLMB(NLc, MBc)                                                                                           \
    do {                                                                                                              \
        (void)hipFuncSetAttribute((const void*)k_fwd<NLc, MBc>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);     \
        (void)hipFuncSetAttribute((const void*)k_fwd_ens<NLc, MBc>, hipFuncAttributeMaxDynamicSharedMemorySize, lim); \
        (void)hipFuncSetAttribute((const void*)k_inv<NLc, MBc>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);     \
        (void)hipFuncSetAttribute((const void*)k_inv_ens<NLc, MBc>, hipFuncAttributeMaxDynamicSharedMemorySize, lim); \
        (void)hipFuncSetAttribute((const void*)k_fwd_adj<NLc, MBc>, hipFuncAttributeMaxDynamicSharedMemorySize, lim); \
        (void)hipFuncSetAttribute((const void*)k_inv_adj<NLc, MBc>, hipFuncAttributeMaxDynamicSharedMemorySize, lim); \
    } while (0)
#define ATTR_NL(NLc) ATTR_NLMB(NLc, 1); ATTR_NLMB(NLc, 2); ATTR_NLMB(NLc, 4);
        ATTR_NL(2) ATTR_NL(3) ATTR_NL(4) ATTR_NL(5) ATTR_NL(6) ATTR_NL(7) ATTR_NL(8)
#undef ATTR_NL
#undef ATTR_NLMB
#define C(N) set_wave_attrs_nl##N(lim);                                                             \
    (void)hipFuncSetAttribute((const void*)k_fwd_pr<N>, hipFuncAttributeMaxDynamicSharedMemorySize, lim); \
    (void)hipFuncSetAttribute((const void*)k_inv_pr<N>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
        C(2) C(3) C(4) C(5) C(6) C(7) C(8)
#undef C
#define ATTR_R(RR)                                                                                                 \
    (void)hipFuncSetAttribute((const void*)k_ysolve<RR>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);        \
    (void)hipFuncSetAttribute((const void*)k_ysolve_ens<RR>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
        ATTR_R(1) ATTR_R(2) ATTR_R(3) ATTR_R(4) ATTR_R(5) ATTR_R(6) ATTR_R(7) ATTR_R(8)
        ATTR_R(9) ATTR_R(10) ATTR_R(11) ATTR_R(12) ATTR_R(13) ATTR_R(14) ATTR_R(15) ATTR_R(16)
#undef ATTR_R
    });
}

}  // namespace oqg

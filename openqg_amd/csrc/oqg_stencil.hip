// oqg_stencil.hip -- the grid-point kernels of the QG step:
//   k_vort     Arakawa Jacobian + Del^2/Del^4/Del^6 friction + forcing + leapfrog
//              (vorticity.F90:14-103, numerics.F90:340-468) on 64x16 LDS tiles with a
//              halo of 3; writes q_new over qm on the points that ocqbdy does not
//              rewrite (the tiles start at row 1, and at column 1 in a box);
//              in a channel its first and last tile rows also write the per-tile
//              partials of the momentum constraints' boundary line integrals
//              (vorticity.F90:57-62,84-101, numerics.F90:374-386);
//   k_suppress suppress_comp_mode (driver.F90:174-189);
//   k_qcomp    qcomp (vorsubs.F90:27-56), used by init_pv.
// All arithmetic is FP64; the library is built with -ffp-contract=off.

#include "oqg_device.h"

namespace oqg {

// ------------------------------------------------------------------ k_vort

// The tile with its halo of 3 lies inside the domain and holds no boundary point
// (every stencil is the plain interior formula, identical in value to del2_lds / jac_lds).
__device__ __forceinline__ double lap_int(const double* s, int W, int o, const Geo& g) {
    const double C = s[o];
    return g.fx * ((s[o - 1] + s[o + 1]) - 2.0 * C) + g.fy * ((s[o - W] + s[o + W]) - 2.0 * C);
}

__device__ __forceinline__ double jac_int(const double* q, const double* p, int W, int o, const Geo& g) {
    const double dqx = q[o + 1] - q[o - 1], dqy = q[o + W] - q[o - W];
    const double dpx = p[o + 1] - p[o - 1], dpy = p[o + W] - p[o - W];
    double t = dqx * dpy - dqy * dpx;
    t = t + (q[o + 1] * (p[o + 1 + W] - p[o + 1 - W]) - q[o - 1] * (p[o - 1 + W] - p[o - 1 - W]));
    t = t - (q[o + W] * (p[o + W + 1] - p[o + W - 1]) - q[o - W] * (p[o - W + 1] - p[o - W - 1]));
    t = t + (p[o + W] * (q[o + W + 1] - q[o + W - 1]) - p[o - W] * (q[o - W + 1] - q[o - W - 1]));
    t = t - (p[o + 1] * (q[o + 1 + W] - q[o + 1 - W]) - p[o - 1] * (q[o - 1 + W] - q[o - 1 - W]));
    return g.jfac * t;
}

// Threads of a tile's workgroup: 512 for 16-row tiles (two output rows per
// thread), 256 (four rows of 64) otherwise.
__host__ __device__ constexpr int vort_threads(int ty) { return ty >= 16 ? 512 : 256; }

// Tile (bx, by) of layer k, TY rows of 64 points.
template <int TY, bool EDGE>
__device__ __forceinline__ void vort_tile(const VortArgs& a, int bx, int by, int k, double* s_pm, double* s_d2,
                                          double* s_d4, double* s_q, double* s_p, double* s_red) {
    constexpr int TX = VORT_TX;
    constexpr int NTH = vort_threads(TY), TR = NTH / 64;   // TR rows of 64 threads
    static_assert(TY % TR == 0 && TY <= 64 && TR >= 4, "rows of 64 threads; the ring logic takes waves 0..3 and tx < TY");
    constexpr int W3 = TX + 6, H3 = TY + 6, W2 = TX + 4, H2 = TY + 4, W1 = TX + 2, H1 = TY + 2;
    const Geo& g = a.g;
    const int tid = threadIdx.x;
    const int tx = tid & 63, ty = tid >> 6;
    const int i0 = a.org_x + bx * TX, j0 = 1 + by * TY;
    const int lim_x = vort_lim_x(g), lim_y = g.nyp - 1;
    const long long lo = (long long)k * g.fs;
    const double* pm = a.pm + lo;
    const double* q = a.q + lo;
    const double* p = a.p + lo;

    // Issue every global load of the block at once (one memory round trip):
    // the pm tile with halo 3, the q and p tiles with halo 1, and the pointwise
    // inputs of this thread's four output points.
    // Edge tiles load from clamped offsets and zero the out-of-basin values
    // when they are stored, so their loads carry no branches either.
    constexpr int NPM = (H3 * W3 + NTH - 1) / NTH, NQP = (H1 * W1 + NTH - 1) / NTH, NOUT = TY / TR;
    double rpm[NPM], rq[NQP], rp[NQP], rqm[NOUT], rent[NOUT], rwek[NOUT];
    bool okpm[NPM], okqp[NQP];
#pragma unroll
    for (int t = 0; t < NPM; t++) {
        const int idx = tid + NTH * t;
        const int r = idx / W3, c = idx - r * W3;
        if constexpr (EDGE) {
            bool ok;
            const int o = field_offset(i0 - 3 + c, j0 - 3 + r, g, ok);
            okpm[t] = ok && idx < H3 * W3;
            rpm[t] = pm[okpm[t] ? o : 0];
        } else {
            rpm[t] = (idx < H3 * W3) ? pm[(j0 - 3 + r) * g.pitch + (i0 - 3 + c)] : 0.0;
        }
    }
#pragma unroll
    for (int t = 0; t < NQP; t++) {
        const int idx = tid + NTH * t;
        const int r = idx / W1, c = idx - r * W1;
        const bool in = idx < H1 * W1;
        if constexpr (EDGE) {
            bool ok;
            const int o = field_offset(i0 - 1 + c, j0 - 1 + r, g, ok);
            okqp[t] = ok && in;
            rq[t] = q[okqp[t] ? o : 0];
            rp[t] = p[okqp[t] ? o : 0];
        } else {
            const int o = (j0 - 1 + r) * g.pitch + (i0 - 1 + c);
            rq[t] = in ? q[o] : 0.0;
            rp[t] = in ? p[o] : 0.0;
        }
    }
    const int gi = i0 + tx;
#pragma unroll
    for (int u = 0; u < NOUT; u++) {
        const int gj = j0 + ty + TR * u;
        const bool in = !EDGE || (gi < lim_x && gj < lim_y);
        const long long o2 = in ? (long long)gj * g.pitch + gi : 0;
        rqm[u] = a.qm[lo + o2];
        rent[u] = (k <= 1) ? a.ent1[o2] : 0.0;
        rwek[u] = (k == 0) ? a.wekp[o2] : 0.0;
    }
    // Edge tiles of layer 1 also carry the trapezoid weights of the boundary
    // ring of ent1 (int_P_dA, numerics.F90:165-204), whose points no tile
    // computes: the rows j = 0, nyp-1 come with the first / last tile row, and
    // in a box the columns i = 0, nxp-1 and the corners with the first / last
    // tile column.  At most one ring point per thread.
    double ring_w = 0.0;
    long long ring_o = 0;
    if (EDGE && k == 0) {
        const bool frow = by == 0, lrow = by == a.tiles_y - 1;
        const bool fcol = bx == 0, lcol = bx == a.tiles_x - 1;
        if (((ty == 0 && frow) || (ty == 1 && lrow)) && gi < lim_x) {
            const bool ex = (gi == 0 || gi == g.nxp - 1);
            ring_w = ex ? 0.25 : 0.5;
            ring_o = (long long)((ty == 0) ? 0 : g.nyp - 1) * g.pitch + gi;
        }
        if (!g.cyclic && ((ty == 2 && fcol) || (ty == 3 && lcol))) {
            const int gic = (ty == 2) ? 0 : g.nxp - 1;
            if (tx < TY) {
                if (j0 + tx < lim_y) {
                    ring_w = 0.5;
                    ring_o = (long long)(j0 + tx) * g.pitch + gic;
                }
            } else if ((tx == TY && frow) || (tx == TY + 1 && lrow)) {
                ring_w = 0.25;
                ring_o = (long long)((tx == TY) ? 0 : g.nyp - 1) * g.pitch + gic;
            }
        }
    }
    const double ring_e = (EDGE && k == 0) ? a.ent1[ring_o] : 0.0;
    // channel: wave 0 of the first tile row and wave 1 of the last one sum the
    // south and north boundary integrals over the tile's columns; layer 1 also
    // needs the layer-2 edge rows of ent there (ring_e is ent1 on the boundary row)
    const bool cyc_s = EDGE && g.cyclic && a.cpart && ty == 0 && by == 0;
    const bool cyc_n = EDGE && g.cyclic && a.cpart && ty == 1 && by == a.tiles_y - 1;
    double e2v = 0.0, e2on = 0.0;
    if (EDGE && k == 0 && (cyc_s || cyc_n)) {
        const int gic = gi < g.nxp ? gi : 0;
        e2v = cyc_s ? a.ent2s[gic] : a.ent2n[gic];
        e2on = a.fcon[F_ENT2];
    }
#pragma unroll
    for (int t = 0; t < NPM; t++) {
        const int idx = tid + NTH * t;
        if (idx < H3 * W3) s_pm[idx] = (!EDGE || okpm[t]) ? rpm[t] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < NQP; t++) {
        const int idx = tid + NTH * t;
        if (idx < H1 * W1) {
            s_q[idx] = (!EDGE || okqp[t]) ? rq[t] : 0.0;
            s_p[idx] = (!EDGE || okqp[t]) ? rp[t] : 0.0;
        }
    }
    __syncthreads();
    // Del^2 of pm on the tile with halo 2
    if constexpr (EDGE) {
#pragma unroll
        for (int t = 0; t < (H2 * W2 + NTH - 1) / NTH; t++) {
            const int idx = tid + NTH * t;
            const int r = idx / W2, c = idx - r * W2;
            const int gj = j0 - 2 + r, gx = i0 - 2 + c;
            const int gxw = g.cyclic ? wrapx(gx, g) : gx;
            const bool in = gj >= 0 && gj < g.nyp && gxw >= 0 && gxw < g.nxp;
            const double v = del2_lds(s_pm, W3, r + 1, c + 1, gxw, gj, g);
            if (idx < H2 * W2) s_d2[idx] = in ? v : 0.0;
        }
    } else {
#pragma unroll
        for (int u = 0; u < (H2 + TR - 1) / TR; u++) {
            const int r = ty + TR * u;
            if (r < H2) {
                s_d2[r * W2 + tx] = lap_int(s_pm, W3, (r + 1) * W3 + tx + 1, g);
                if (tx < W2 - 64) s_d2[r * W2 + tx + 64] = lap_int(s_pm, W3, (r + 1) * W3 + tx + 65, g);
            }
        }
    }
    __syncthreads();
    // Del^4 on the tile with halo 1
    if constexpr (EDGE) {
#pragma unroll
        for (int t = 0; t < (H1 * W1 + NTH - 1) / NTH; t++) {
            const int idx = tid + NTH * t;
            const int r = idx / W1, c = idx - r * W1;
            const int gj = j0 - 1 + r, gx = i0 - 1 + c;
            const int gxw = g.cyclic ? wrapx(gx, g) : gx;
            const bool in = gj >= 0 && gj < g.nyp && gxw >= 0 && gxw < g.nxp;
            const double v = del2_lds(s_d2, W2, r + 1, c + 1, gxw, gj, g);
            if (idx < H1 * W1) s_d4[idx] = in ? v : 0.0;
        }
    } else {
#pragma unroll
        for (int u = 0; u < (H1 + TR - 1) / TR; u++) {
            const int r = ty + TR * u;
            if (r < H1) {
                s_d4[r * W1 + tx] = lap_int(s_d2, W2, (r + 1) * W2 + tx + 1, g);
                if (tx < W1 - 64) s_d4[r * W1 + tx + 64] = lap_int(s_d2, W2, (r + 1) * W2 + tx + 65, g);
            }
        }
    }
    __syncthreads();
    if (EDGE && (cyc_s || cyc_n)) {
        // rows gb0 (boundary) and gb1 (inner) of this side in the LDS tiles
        const int gb0 = cyc_s ? 0 : g.nyp - 1, gb1 = cyc_s ? 1 : g.nyp - 2;
        const int r2a = gb0 - (j0 - 2), r2b = gb1 - (j0 - 2);   // s_d2 rows
        const int r1a = gb0 - (j0 - 1), r1b = gb1 - (j0 - 1);   // s_d4, s_q, s_p rows
        const int r3a = gb0 - (j0 - 3), r3b = gb1 - (j0 - 3);   // s_pm rows
        const double w = gi < g.nxp ? ((gi == 0 || gi == g.nxp - 1) ? 0.5 : 1.0) : 0.0;
        const double d2a = s_d2[r2a * W2 + tx + 2], d2b = s_d2[r2b * W2 + tx + 2];
        const double d4a = s_d4[r1a * W1 + tx + 1], d4b = s_d4[r1b * W1 + tx + 1];
        const double dpx = s_p[r1b * W1 + tx + 2] - s_p[r1b * W1 + tx];
        const double pma = s_pm[r3a * W3 + tx + 3], pmb = s_pm[r3b * W3 + tx + 3];
        double v[CQ_N];
        v[CQ_D2] = w * (cyc_s ? d2b - d2a : d2a - d2b);
        v[CQ_D4] = w * (cyc_s ? d4b - d4a : d4a - d4b);
        v[CQ_AJ5] = w * (s_q[r1a * W1 + tx + 1] * dpx);
        v[CQ_AJ9] = w * (s_q[r1b * W1 + tx + 1] * dpx);
        v[CQ_BD] = (k == g.kt) ? w * (cyc_s ? pmb - pma : pma - pmb) : 0.0;
        v[CQ_E1] = (k == 0) ? w * ring_e : 0.0;
        v[CQ_E2] = (k == 0 && e2on != 0.0) ? w * e2v : 0.0;
        double* o = a.cpart + ((long long)(k * 2 + (cyc_s ? 0 : 1)) * a.tiles_x + bx) * CYC_SLOTS;
#pragma unroll
        for (int t = 0; t < CQ_N; t++) {
            const double r = wave_sum(v[t]);
            if (tx == 0) o[t] = r;
        }
    }
    const double c2 = a.ph->c2[k], c4 = a.ph->c4[k];
    double epart = 0.0;
#pragma unroll
    for (int u = 0; u < NOUT; u++) {
        const int r = ty + TR * u;
        const int gj = j0 + r;
        const bool out = !EDGE || (gi < lim_x && gj < lim_y);
        const int o1 = (r + 1) * W1 + tx + 1;
        double d6, dqdt;
        if constexpr (EDGE) {
            d6 = del2_lds(s_d4, W1, r + 1, tx + 1, gi, gj, g);
            dqdt = jac_lds(s_q, s_p, W1, r + 1, tx + 1, gi, gj, g);
        } else {
            d6 = lap_int(s_d4, W1, o1, g);
            dqdt = jac_int(s_q, s_p, W1, o1, g);
        }
        const double d4 = s_d4[o1];
        const double d2 = s_d2[(r + 2) * W2 + tx + 2];
        dqdt = dqdt + c2 * d4 - c4 * d6;
        if (k == 0) {
            const double e = rent[u];
            dqdt = dqdt + a.ph->cent1 * e + a.ph->cwek * rwek[u];
            if constexpr (EDGE) {
                // interior points weigh 1; in a channel the columns 0 and nxp-1 weigh 1/2
                const double w = (gi == 0 || gi == g.nxp - 1) ? 0.5 : 1.0;
                if (out) epart += w * e;
            } else {
                epart += e;
            }
        }
        if (k == 1) dqdt = dqdt - a.ph->cent2 * rent[u];
        if (k == g.kt) dqdt = dqdt - a.ph->cdrag * d2;
        if (out) a.qnew[lo + (long long)gj * g.pitch + gi] = rqm[u] + a.tdt * dqdt;
    }
    if (EDGE && k == 0) epart += ring_w * ring_e;
    if (k == 0) {
        double v[1] = {epart};
        block_sum<1>(v, s_red);
        if (threadIdx.x == 0) a.entpart[by * a.tiles_x + bx] = v[0];
    }
}

// Tiles whose halo of 3 stays inside the domain without touching a boundary
// point or crossing the channel's periodic seam take the branch-free body.
template <int TY>
__device__ __forceinline__ void vort_block(const VortArgs& a, int bx, int by, int k) {
    constexpr int TX = VORT_TX;
    __shared__ double s_pm[(TY + 6) * (TX + 6)];
    __shared__ double s_d2[(TY + 4) * (TX + 4)];
    __shared__ double s_d4[(TY + 2) * (TX + 2)];
    __shared__ double s_q[(TY + 2) * (TX + 2)];
    __shared__ double s_p[(TY + 2) * (TX + 2)];
    __shared__ double s_red[16];
    const int i0 = a.org_x + bx * TX, j0 = 1 + by * TY;
    const bool inner = i0 >= 3 && i0 + TX + 3 <= a.g.nxp && j0 >= 3 && j0 + TY + 3 <= a.g.nyp;
    if (inner) vort_tile<TY, false>(a, bx, by, k, s_pm, s_d2, s_d4, s_q, s_p, s_red);
    else vort_tile<TY, true>(a, bx, by, k, s_pm, s_d2, s_d4, s_q, s_p, s_red);
}

// Each XCD takes a contiguous run of tiles (whole rows of tiles), so the halo
// lines that neighbouring tiles share are fetched from HBM into one L2.
template <int TY>
__global__ __launch_bounds__(vort_threads(TY), TY >= 16 ? 6 : 1) void k_vort(VortArgs a) {
    int bx, by, bz;
    xcd_block(bx, by, bz);
    vort_block<TY>(a, bx, by, bz);
}

// Ensemble form: block b works on tile map[b].y of member map[b].x, whose
// arguments are args[map[b].x] (tiles numbered x fastest, then y, then layer).
template <int TY>
__global__ __launch_bounds__(vort_threads(TY), TY >= 16 ? 6 : 1) void k_vort_ens(const VortArgs* __restrict__ args, const int2* __restrict__ map) {
    const int2 e = map[blockIdx.x];
    const VortArgs& a = args[e.x];
    const int per_layer = a.tiles_x * a.tiles_y;
    const int k = e.y / per_layer, r = e.y - k * per_layer;
    const int by = r / a.tiles_x, bx = r - by * a.tiles_x;
    vort_block<TY>(a, bx, by, k);
}

// A tile's LDS (bytes): pm with halo 3, del2 with halo 2, del4, q and p with
// halo 1, the reduction slots.
static constexpr size_t vort_lds(int ty) {
    return sizeof(double) * ((size_t)(ty + 6) * (VORT_TX + 6) + (size_t)(ty + 4) * (VORT_TX + 4) +
                             3 * (size_t)(ty + 2) * (VORT_TX + 2) + 16);
}

// Shorter tiles shorten each block's chain of loads and LDS passes (the
// kernel is latency-bound), at the price of more halo reads; they pay while
// every block is resident at once: 64 x 4 tiles took k_vort from 10.4 to 6.1
// us on the 97 x 49 channel and from 9.5 to 6.5 us on the 257^2 box, while
// the 513^2 box stays with 64 x 16 (14.2 us against 14.7 us with 64 x 4).
int vort_rows(int tiles_x, int nyp, int nl) {
    constexpr int CUS = 256;
    constexpr size_t LDS_PER_CU = 160 * 1024;
    for (int ty : {4, 8}) {
        const long long blocks = (long long)tiles_x * ((nyp - 2 + ty - 1) / ty) * nl;
        const long long slots = (long long)CUS * (long long)(LDS_PER_CU / vort_lds(ty));
        if (blocks <= slots) return ty;
    }
    return 16;
}

void launch_vort(const VortArgs& a, hipStream_t st) {
    dim3 grid(a.tiles_x, a.tiles_y, a.g.nl);
    if (a.ty == 4) hipLaunchKernelGGL(k_vort<4>, grid, dim3(256), 0, st, a);
    else if (a.ty == 8) hipLaunchKernelGGL(k_vort<8>, grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL(k_vort<16>, grid, dim3(vort_threads(16)), 0, st, a);
}

void launch_vort_ens(const VortArgs* args, const int2* map, int nblocks, int ty, hipStream_t st) {
    if (ty == 4) hipLaunchKernelGGL(k_vort_ens<4>, dim3(nblocks), dim3(256), 0, st, args, map);
    else if (ty == 8) hipLaunchKernelGGL(k_vort_ens<8>, dim3(nblocks), dim3(256), 0, st, args, map);
    else hipLaunchKernelGGL(k_vort_ens<16>, dim3(nblocks), dim3(vort_threads(16)), 0, st, args, map);
}

// ------------------------------------------------------------------ k_suppress

// Block b of nb blocks sweeping one basin.
__device__ __forceinline__ void suppress_block(const SuppArgs& a, int b, int nb) {
    const Geo& g = a.g;
    const long long n = (long long)g.nl * g.nyp * g.pitch;
    for (long long t = b * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)nb * blockDim.x) {
        a.q[t] = 0.5 * (a.q[t] + a.qm[t]);
        a.p[t] = 0.5 * (a.p[t] + a.pm[t]);
    }
    if (b == 0 && threadIdx.x == 0) {
        for (int l = 0; l < g.nl - 1; l++) a.c[C_DPI + l] = 0.5 * (a.c[C_DPI + l] + a.c[C_DPIP + l]);
        if (g.cyclic)
            for (int k = 0; k < g.nl; k++) {
                a.c[C_CS + k] = 0.5 * (a.c[C_CS + k] + a.c[C_CSP + k]);
                a.c[C_CN + k] = 0.5 * (a.c[C_CN + k] + a.c[C_CNP + k]);
            }
    }
}

__global__ void k_suppress(SuppArgs a) { suppress_block(a, blockIdx.x, gridDim.x); }

// Ensemble form: blockIdx.y is the member.
__global__ void k_suppress_ens(const SuppArgs* __restrict__ args) {
    suppress_block(args[blockIdx.y], blockIdx.x, gridDim.x);
}

void launch_suppress(const SuppArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_suppress, dim3(1024), dim3(256), 0, st, a);
}

void launch_suppress_ens(const SuppArgs* args, int nmembers, hipStream_t st) {
    hipLaunchKernelGGL(k_suppress_ens, dim3(256, nmembers), dim3(256), 0, st, args);
}

// ------------------------------------------------------------------ k_qcomp

// q = Del^2_bc(p)/f0 (+ddyn at k_topo) + beta*yprel - f0*A p  (vorsubs.F90:46-54)
__global__ void k_qcomp(QcompArgs a) {
    const Geo& g = a.g;
    const long long npts = (long long)g.nyp * g.nxp;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < npts; t += (long long)gridDim.x * blockDim.x) {
        const int gj = (int)(t / g.nxp), gi = (int)(t - (long long)gj * g.nxp);
        const long long o = (long long)gj * g.pitch + gi;
        double pv[MAXL];
#pragma unroll
        for (int l = 0; l < MAXL; l++) pv[l] = (l < g.nl) ? a.p[l * g.fs + o] : 0.0;
        const int il = g.cyclic ? wrapx(gi - 1, g) : gi - 1;
        const int ir = g.cyclic ? wrapx(gi + 1, g) : gi + 1;
#pragma unroll
        for (int k = 0; k < MAXL; k++) {
            if (k >= g.nl) break;
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
#pragma unroll
            for (int l = 0; l < MAXL; l++)
                if (l < g.nl) ap += a.mx.amat[k * MAXL + l] * pv[l];
            qv = qv - g.fnot * ap;
            a.q[k * g.fs + o] = qv;
        }
    }
}

void launch_qcomp(const QcompArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_qcomp, dim3(512), dim3(256), 0, st, a);
}

}  // namespace oqg

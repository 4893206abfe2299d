// oqg_mixed.hip -- the ocean mixed layer on the device: step_oml
// (mixed/omlsubs.F90:69-119) with adv_diff (mixed/mixed_layer.F90:16-95), the
// Ekman velocities of ekman_from_tau (coupler/ekman.F90:73-80) and the
// T-point stencils of glam/numerics.F90 (del2_T, map_T_to_x_face,
// map_T_to_y_face, c_grid_advection, map_T_to_P).
//
// Two launches per step:
//   k_oml_step    one thread per T point: the new mixed-layer temperature,
//                 its convective correction and the layer 1/2 entrainment at
//                 T points, plus per-block partial sums of the latter;
//   k_oml_finish  one thread per p point: the mean of the entrainment from
//                 the partials (block order), the entrainment mapped to p
//                 points into the basin's resident forcing, and the time
//                 levels shifted (sstm <- sst, sst <- new) by copies, so the
//                 buffers keep their roles in replayed graphs.
// T fields are stored like p fields (rows of the basin's pitch), nyt rows.

#include "oqg_mixed_dev.h"

namespace oqg {

__global__ __launch_bounds__(256) void k_oml_step(OmlArgs a) {
    const Geo& g = a.g;
    const int nxt = g.nxp - 1, nyt = g.nyp - 1, nxp = g.nxp, nyp = g.nyp;
    // 1-D blocks of 256 threads on 64 x 4 tiles of T points (block_sum takes 1-D blocks)
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const long long pitch = g.pitch;
    const bool in = i < nxt && j < nyt;
    double xv = 0.0;
    if (in) {
        const AdvField m{a.sst, a.sstm, a.d2, a.d4, a.ycexp, a.nval, a.sval, a.fluxn, a.fluxs};
        const double ad = adv_diff_at(g, m, a.p1, a.taux, a.tauy, a.hm, a.tau_sign, i, j);
        const double wekt = wekt_at(g, a.taux, a.tauy, i, j);
        // step_oml (omlsubs.F90:92-108)
        const long long o = (long long)j * pitch + i;
        const double sm = a.sstm[o];
        const double half = 0.5 * wekt * (sm - a.T1_rel);
        const double snew = sm + a.tdt * (ad + (1.0 / (a.rho_cp * a.hm)) * a.fnet[o] + (1.0 / a.hm) * (wekt * sm - half));
        const double dtc = fmax(0.0, a.T1_rel - snew);
        xv = (-1.0 / a.dT_12) * half - (a.hm / (a.tdt * a.dT_12)) * dtc;
        a.snew[o] = snew + dtc;
        a.xfo[o] = xv;
    }
    // per-block partial sum of the entrainment at T points
    __shared__ double red[16];
    double v[1] = {xv};
    block_sum<1>(v, red);
    if (threadIdx.x == 0) a.part[blockIdx.y * gridDim.x + blockIdx.x] = v[0];
}

__global__ __launch_bounds__(256) void k_oml_finish(OmlArgs a) {
    const Geo& g = a.g;
    const int nxt = g.nxp - 1, nyt = g.nyp - 1, nxp = g.nxp, nyp = g.nyp;
    const long long pitch = g.pitch;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = t / nxp, i = t - j * nxp;
    const bool in = j < nyp;
    if (a.compute_ent) {
        // avg_T (numerics.F90:155-162) from the partials, in block order
        __shared__ double red[16];
        double s = 0.0;
        for (int b = threadIdx.x; b < a.npart; b += blockDim.x) s += a.part[b];
        double v[1] = {s};
        block_sum<1>(v, red);
        const double avg = v[0] / (double)(nxt * nyt);
        if (in) {
            auto X = [&](int ii, int jj) { return a.xfo[(long long)jj * pitch + ii] - avg; };
            // map_T_to_P (numerics.F90:64-107) of xfo - avg
            const double e = map_T_to_P_at(X, i, j, nxp, nyp, g.cyclic);
            a.ent1[(long long)j * pitch + i] = e;
        }
    }
    // the time levels: sstm <- sst, sst <- new
    if (in && i < nxt && j < nyt) {
        const long long o = (long long)j * pitch + i;
        const double cur = a.sst[o];
        a.sstm[o] = cur;
        a.sst[o] = a.snew[o];
    }
}

// suppress_ocean's mixed-layer part (driver.F90:165-167)
__global__ __launch_bounds__(256) void k_oml_suppress(OmlArgs a) {
    const int nxt = a.g.nxp - 1, nyt = a.g.nyp - 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = t / nxt, i = t - j * nxt;
    if (j < nyt) {
        const long long o = (long long)j * a.g.pitch + i;
        a.sst[o] = 0.5 * (a.sst[o] + a.sstm[o]);
    }
}

dim3 oml_step_grid(const Geo& g) { return dim3((g.nxp - 1 + 63) / 64, (g.nyp - 1 + 3) / 4); }

void launch_oml_step(const OmlArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_oml_step, oml_step_grid(a.g), dim3(256), 0, st, a);
}

void launch_oml_finish(const OmlArgs& a, hipStream_t st) {
    const int n = a.g.nxp * a.g.nyp;
    hipLaunchKernelGGL(k_oml_finish, dim3((n + 255) / 256), dim3(256), 0, st, a);
}

void launch_oml_suppress(const OmlArgs& a, hipStream_t st) {
    const int n = (a.g.nxp - 1) * (a.g.nyp - 1);
    hipLaunchKernelGGL(k_oml_suppress, dim3((n + 255) / 256), dim3(256), 0, st, a);
}

}  // namespace oqg

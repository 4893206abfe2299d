// oqg_forcing.hip -- the wind-stress forcing producer on the device.
//
// ekman_from_tau (coupler/ekman.F90:53-83): face averages of the kinematic
// stress (map_P_to_x_face / map_P_to_y_face, numerics.F90:271-289), the Ekman
// pumping at T points and map_T_to_P (numerics.F90:64-107) to p points, with
// the channel's boundary stress integrals txis/txin (int_P_dx of the two
// outer row pairs).  One thread per p point recomputes the (at most four)
// T-point values it averages, in the reference's order of operations.  The
// transposes (for gradients with respect to the stress) are gathers in the
// same style.

#include "oqg_device.h"

namespace oqg {

// Ekman pumping at T point (I, J), 0-based (0 <= I < nxp-1, 0 <= J < nyp-1),
// from the stress at p points (pitch layout).
__device__ __forceinline__ double wek_t(const EkmanArgs& a, int I, int J) {
    const Geo& g = a.g;
    const long long p = g.pitch;
    auto tx = [&](int i, int j) { return a.taux[(long long)j * p + i]; };
    auto ty = [&](int i, int j) { return a.tauy[(long long)j * p + i]; };
    // tauy on the y faces at columns I and I+1, taux on the x faces at rows J and J+1
    const double tyE = 0.5 * (ty(I + 1, J + 1) + ty(I + 1, J)), tyW = 0.5 * (ty(I, J + 1) + ty(I, J));
    const double txN = 0.5 * (tx(I + 1, J + 1) + tx(I, J + 1)), txS = 0.5 * (tx(I + 1, J) + tx(I, J));
    return (1.0 / (g.dx * g.fnot)) * (tyE - tyW) - (1.0 / (g.dy * g.fnot)) * (txN - txS);
}

__global__ void k_ekman(EkmanArgs a) {
    const Geo& g = a.g;
    const int nx = g.nxp, ny = g.nyp;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nx * ny) return;
    const int j = t / nx, i = t - j * nx;
    const int mi = nx - 2, mj = ny - 2;   // last T index in x and y
    double w;
    const bool xin = i >= 1 && i <= nx - 2, yin = j >= 1 && j <= ny - 2;
    if (xin && yin) {
        w = 0.25 * (((wek_t(a, i, j) + wek_t(a, i - 1, j)) + wek_t(a, i, j - 1)) + wek_t(a, i - 1, j - 1));
    } else if (xin) {   // N/S boundary rows
        const int J = (j == 0) ? 0 : mj;
        w = 0.5 * (wek_t(a, i - 1, J) + wek_t(a, i, J));
    } else if (g.cyclic) {
        // column 0 wraps to the last T column; column nxp-1 repeats column 0
        if (yin) w = 0.25 * (((wek_t(a, 0, j) + wek_t(a, mi, j)) + wek_t(a, 0, j - 1)) + wek_t(a, mi, j - 1));
        else {
            const int J = (j == 0) ? 0 : mj;
            w = 0.5 * (wek_t(a, 0, J) + wek_t(a, mi, J));
        }
    } else if (yin) {   // E/W boundary columns of a box
        const int I = (i == 0) ? 0 : mi;
        w = 0.5 * (wek_t(a, I, j - 1) + wek_t(a, I, j));
    } else {            // box corners
        w = wek_t(a, (i == 0) ? 0 : mi, (j == 0) ? 0 : mj);
    }
    a.wekp[(long long)j * g.pitch + i] = w;
}

// Channel: txis = 0.5 int_P_dx(taux row 1 + row 2), txin the same of the last
// two rows (ekman.F90:61-70); one block, the sums in a fixed order.
__global__ __launch_bounds__(256) void k_ekman_txi(EkmanArgs a) {
    const Geo& g = a.g;
    __shared__ double red[2 * 16];
    const long long p = g.pitch;
    const double* r0 = a.taux;
    const double* r1 = a.taux + p;
    const double* r2 = a.taux + (long long)(g.nyp - 2) * p;
    const double* r3 = a.taux + (long long)(g.nyp - 1) * p;
    double v[2] = {0.0, 0.0};
    for (int i = threadIdx.x + 1; i <= g.nxp - 2; i += blockDim.x) {
        v[0] += r0[i] + r1[i];
        v[1] += r2[i] + r3[i];
    }
    block_sum<2>(v, red);
    if (threadIdx.x == 0) {
        const int e = g.nxp - 1;
        const double s = ((0.5 * (r0[0] + r1[0]) + v[0]) + 0.5 * (r0[e] + r1[e])) * g.dx;
        const double n = ((0.5 * (r2[0] + r3[0]) + v[1]) + 0.5 * (r2[e] + r3[e])) * g.dx;
        a.fcon[F_TXIS] = 0.5 * s;
        a.fcon[F_TXIN] = 0.5 * n;
    }
}

// ---------------------------------------------------------------- transpose (box)

// Weight with which map_T_to_P at p point (i, j) takes T point (I, J) (box).
__device__ __forceinline__ double t2p_weight(const Geo& g, int i, int j, int I, int J) {
    const int nx = g.nxp, ny = g.nyp;
    if (i < 0 || i >= nx || j < 0 || j >= ny) return 0.0;
    if (!(I == i || I == i - 1) || !(J == j || J == j - 1)) return 0.0;
    const bool xin = i >= 1 && i <= nx - 2, yin = j >= 1 && j <= ny - 2;
    if (xin && yin) return 0.25;
    if (xin) return 0.5;   // N/S rows use T row 0 / nyp-2, which (I, J) is when it neighbours
    if (yin) return 0.5;
    return 1.0;            // a corner uses its single T point
}

// Adjoint of the T-point pumping, lt(I, J) = sum over p points of lw * weight.
__device__ __forceinline__ double lwek_t(const EkmanArgs& a, int I, int J) {
    const Geo& g = a.g;
    if (I < 0 || I > g.nxp - 2 || J < 0 || J > g.nyp - 2) return 0.0;
    double s = 0.0;
#pragma unroll
    for (int dj = 0; dj <= 1; dj++)
#pragma unroll
        for (int di = 0; di <= 1; di++) {
            const int i = I + di, j = J + dj;
            const double w = t2p_weight(g, i, j, I, J);
            if (w != 0.0) s += w * a.lwek[(long long)j * g.pitch + i];
        }
    return s;
}

// ltaux(i, j) and ltauy(i, j): the face averages and the T-point differences transposed.
__global__ void k_ekman_adj(EkmanArgs a) {
    const Geo& g = a.g;
    const int nx = g.nxp, ny = g.nyp;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nx * ny) return;
    const int j = t / nx, i = t - j * nx;
    const double cx = 1.0 / (g.dx * g.fnot), cy = 1.0 / (g.dy * g.fnot);
    // x-face value (I, j') of taux feeds T points (I, j'-1) with -cy and (I, j') with +cy
    auto ltaux_face = [&](int I, int jf) {
        if (I < 0 || I > nx - 2) return 0.0;
        double v = 0.0;
        if (jf >= 1) v -= cy * lwek_t(a, I, jf - 1);
        if (jf <= ny - 2) v += cy * lwek_t(a, I, jf);
        return v;
    };
    // y-face value (i', J) of tauy feeds T points (i'-1, J) with +cx and (i', J) with -cx
    auto ltauy_face = [&](int ifc, int J) {
        if (J < 0 || J > ny - 2) return 0.0;
        double v = 0.0;
        if (ifc >= 1) v += cx * lwek_t(a, ifc - 1, J);
        if (ifc <= nx - 2) v -= cx * lwek_t(a, ifc, J);
        return v;
    };
    const double lx = 0.5 * ltaux_face(i - 1, j) + 0.5 * ltaux_face(i, j);
    const double ly = 0.5 * ltauy_face(i, j - 1) + 0.5 * ltauy_face(i, j);
    a.taux[(long long)j * g.pitch + i] = lx;
    a.tauy[(long long)j * g.pitch + i] = ly;
}

// Channel transpose, on the torus of the nxp-1 distinct columns (column nxp-1
// of the stress repeats column 0; the adjoint holds 0 there): p column i takes
// T columns i-1 and i (mod nxp-1), T column I its x faces from taux columns I
// and I+1 and its y faces from tauy columns I and I+1 (mod nxp-1); txis and
// txin take 0.5 dx of taux on their two rows (ltxi[0], ltxi[1]).
__global__ void k_ekman_adj_cyc(EkmanArgs a) {
    const Geo& g = a.g;
    const int L = g.nxp - 1, ny = g.nyp, mj = ny - 2;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nxp * ny) return;
    const int j = t / g.nxp, i = t - j * g.nxp;
    const long long o = (long long)j * g.pitch + i;
    if (i == L) {
        a.taux[o] = 0.0;
        a.tauy[o] = 0.0;
        return;
    }
    auto wrap = [&](int x) { return x < 0 ? x + L : (x >= L ? x - L : x); };
    auto lwek_t = [&](int I, int J) {
        if (J < 0 || J > mj) return 0.0;
        double s = 0.0;
#pragma unroll
        for (int dj = 0; dj <= 1; dj++) {
            const int jj = J + dj;
            const double w = (jj >= 1 && jj <= ny - 2) ? 0.25 : ((jj == 0 || jj == ny - 1) ? 0.5 : 0.0);
#pragma unroll
            for (int di = 0; di <= 1; di++) s += w * a.lwek[(long long)jj * g.pitch + wrap(I + di)];
        }
        return s;
    };
    const double cx = 1.0 / (g.dx * g.fnot), cy = 1.0 / (g.dy * g.fnot);
    auto ltaux_face = [&](int I, int jf) {
        double v = 0.0;
        if (jf >= 1) v -= cy * lwek_t(I, jf - 1);
        if (jf <= ny - 2) v += cy * lwek_t(I, jf);
        return v;
    };
    auto ltauy_face = [&](int ifc, int J) {
        if (J < 0 || J > mj) return 0.0;
        return cx * lwek_t(wrap(ifc - 1), J) - cx * lwek_t(ifc, J);
    };
    double lx = 0.5 * ltaux_face(wrap(i - 1), j) + 0.5 * ltaux_face(i, j);
    const double ly = 0.5 * ltauy_face(i, j - 1) + 0.5 * ltauy_face(i, j);
    if (j <= 1) lx += 0.5 * g.dx * a.ltxi[0];
    if (j >= ny - 2) lx += 0.5 * g.dx * a.ltxi[1];
    a.taux[o] = lx;
    a.tauy[o] = ly;
}

void launch_ekman(const EkmanArgs& a, hipStream_t st) {
    const int n = a.g.nxp * a.g.nyp;
    hipLaunchKernelGGL(k_ekman, dim3((n + 255) / 256), dim3(256), 0, st, a);
    if (a.g.cyclic) hipLaunchKernelGGL(k_ekman_txi, dim3(1), dim3(256), 0, st, a);
}

void launch_ekman_adj(const EkmanArgs& a, hipStream_t st) {
    const int n = a.g.nxp * a.g.nyp;
    if (a.g.cyclic) hipLaunchKernelGGL(k_ekman_adj_cyc, dim3((n + 255) / 256), dim3(256), 0, st, a);
    else hipLaunchKernelGGL(k_ekman_adj, dim3((n + 255) / 256), dim3(256), 0, st, a);
}

}  // namespace oqg

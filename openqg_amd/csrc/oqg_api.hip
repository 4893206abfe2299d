// oqg_api.hip -- host side of liboqg: the C ABI of include/oqg.h, the init-time
// precompute (modes, Helmholtz tables, homogeneous solutions) and step scheduling
// (eager launches or replayed HIP graphs) on one stream.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "oqg_internal.h"

using namespace oqg;

namespace {

thread_local std::string g_create_err;

// Streams are created once per process and reused: a basin, an ensemble or a
// coupler takes one from its device's pool and gives it back, synchronised,
// when it is destroyed.  No stream is destroyed while the process runs, so the
// hardware queues behind the library's streams keep their lifetimes (two runs
// under rocprofv3's kernel trace that aborted with a malformed dispatch packet
// had both destroyed a stream of the same process earlier).
std::mutex g_pool_mu;
std::map<int, std::vector<hipStream_t>> g_stream_pool;

hipError_t pool_take(hipStream_t* out) {
    int dev = 0;
    if (hipError_t e = hipGetDevice(&dev); e != hipSuccess) return e;
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        auto& v = g_stream_pool[dev];
        if (!v.empty()) {
            *out = v.back();
            v.pop_back();
            return hipSuccess;
        }
    }
    return hipStreamCreateWithFlags(out, hipStreamNonBlocking);
}

// dev: the device the stream was taken on.
void pool_give(int dev, hipStream_t s) {
    if (!s) return;
    (void)hipStreamSynchronize(s);
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_stream_pool[dev].push_back(s);
}

constexpr double TWOPI_REF = 2.0 * 3.14159265358979324;   // constants.F90: TWOPI

// Graph kinds: plain steps (oqg_run), steps with snapshot saves
// (oqg_adj_record), adjoint steps (oqg_adj_sweep).
// G_SEG: a stretch of oqg_run shorter than a whole period, captured ahead by
// oqg_plan_run (phase = the stretch's first nt-1 modulo the period).
enum { G_RUN = 0, G_RECORD = 1, G_SWEEP = 2, G_SEG = 3 };

struct GraphKey {
    double tdt;
    int nsteps, supp, cur, kind;
    int phase = 0;
    bool operator<(const GraphKey& o) const {
        return std::tie(tdt, nsteps, supp, cur, kind, phase) < std::tie(o.tdt, o.nsteps, o.supp, o.cur, o.kind, o.phase);
    }
};

}  // namespace

struct oqg_ctx {
    oqg_desc d{};
    Geo g{};
    Spec s{};
    Mix mx{};
    Phys ph{};
    BoxHom bh{};
    CycHom ch{};
    int variant = 0;       // y-solver: 0 wavefront partition+PCR, 1 per-thread Thomas
    int cur = 0;           // time-level parity
    double* hst = nullptr; // pinned staging of host field transfers (hst_words doubles)
    size_t hst_words = 0;
    int tiles_x = 0, tiles_y = 0, vort_ty = VORT_TY;
    size_t fbytes = 0;     // bytes of one nl-layer field (pitch layout)
    // device memory
    // the dynamic state: q[0], q[1], p[0], p[1], c[0], c[1] in one allocation
    // laid out as a snapshot slot (set_work), so that a checkpoint can take
    // it over instead of copying it
    double* qbuf[2] = {nullptr, nullptr};
    double* pbuf[2] = {nullptr, nullptr};
    double* cbuf[2] = {nullptr, nullptr};
    double* wbase = nullptr;        // the allocation qbuf/pbuf/cbuf point into
    double* wbase_own = nullptr;    // the context's own one (wbase except inside a checkpointed sweep)
    // Levels of the state that a step reads from elsewhere than qbuf/pbuf/cbuf
    // (a checkpoint it starts from): the current level (cur) and the previous
    // one; null q: the level is in the working buffers.
    struct LevelRef {
        const double* q = nullptr;
        const double* p = nullptr;
        const double* c = nullptr;
    } ext_cur, ext_old;
    double* spec = nullptr;
    double* gamw = nullptr;
    double* colsum = nullptr;
    double* colw = nullptr;
    double* entpart = nullptr;
    double* cpart = nullptr;   // channel: per-tile partial boundary integrals (k_vort -> k_inv)
    double* ent1 = nullptr;
    double* ent2s = nullptr;
    double* ent2n = nullptr;
    double* wekp = nullptr;
    double* ddyn = nullptr;
    double* yprel = nullptr;
    double* hom = nullptr;
    double* bcol = nullptr;
    double* tw = nullptr;    // twiddles of the forward transform plan (k_fwd)
    double* itw = nullptr;   // twiddles of the inverse transform plan (k_inv)
    double* raw[2] = {nullptr, nullptr};
    DevConsts* dconst = nullptr;
    std::vector<void*> allocs;
    bool has_ent2 = false;
    double* fcon = nullptr;   // txis, txin, layer-2 entrainment flag (device)
    double* taux_d = nullptr;   // kinematic wind stress at p points (oqg_set_wind / oqg_adj_wind)
    double* tauy_d = nullptr;
    double* diag_part = nullptr;   // per-row diagnostics partials (oqg_diagnose)
    double* en_d2 = nullptr;       // oqg_diagnose_energy: del2 of the lagged velocities (2 nl fields)
    double* en_part = nullptr;     // and its per-row partials
    int dev = 0;                   // the HIP device the context was created on
    double* scratch = nullptr;     // this device's address of the no-alloc scratch array
    hipStream_t own = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t ens_done = nullptr;   // end of the last ensemble run this basin took part in, not yet joined
    std::map<GraphKey, hipGraphExec_t> graphs;
    std::string err;
    // host copies
    std::vector<double> yprel_h, ddyn_h;
    // adjoint (allocated by oqg_adj_init): adjoints of q/p at both time-level
    // parities (mirroring qbuf/pbuf), of the Ekman pumping, and scratch
    bool adj = false;
    double* lq[2] = {nullptr, nullptr};
    double* lp[2] = {nullptr, nullptr};
    double* lwek = nullptr;
    double* adjW = nullptr;       // one field: the interior indicator
    double* adjG = nullptr;       // box: nl fields, H_m^-1 of the interior indicator; channel: 2 nl (unit rows)
    double* adjPart = nullptr;
    double* adjCinvT = nullptr;
    // channel: adjoint constraint blocks (lcs, lcn) at both parities, a
    // scratch block, the adjoints of txis and txin, and the gradients of the
    // boundary integrals with respect to the row sums of pm (AdjVortArgs)
    double* lc[2] = {nullptr, nullptr};
    double* lc_scratch = nullptr;
    double* ltxi = nullptr;
    double adj_rows[6][4] = {};
    // state snapshots (oqg_snap_*): q[0..1], p[0..1], c[0..1] and the parity per slot
    std::vector<double*> snaps;
    std::vector<int> snap_cur;
    // reverse sweep: a slot's level held by another slot (q null: its own),
    // see rv_save
    std::vector<std::array<LevelRef, 2>> snap_ref;
    // device copies for graph-replayed record/sweep stretches: the slot
    // pointers and the first slot of the stretch being replayed
    double** snap_ptrs_dev = nullptr;
    int* snap_base_dev = nullptr;
    // binomial checkpointing (oqg_adj_revolve_*): the pending left segments
    // (lo, hi, slot, free) of the forward sweep's rightmost path, its tail
    std::vector<std::array<int, 4>> rv_left;
    std::array<int, 3> rv_tail{};
    int rv_n = -1, rv_supp = 0;
    bool rv_ref = false;   // reverse sweep: checkpoints taken and restored by reference
    double rv_tdt = 0.0;
    long long rv_steps = 0;
    std::map<std::pair<int, int>, int> rv_split;   // memo of revolve_split
};

// ------------------------------------------------------------------ error helpers

// Orders the basin's stream after the ensemble run it last took part in
// (oqg_ens_run leaves the event for the member's next call to wait for).
static void ctx_join(oqg_ctx* c) {
    if (c && c->ens_done) {
        (void)hipStreamWaitEvent(c->st, c->ens_done, 0);
        c->ens_done = nullptr;
    }
}

static int fail(oqg_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    else g_create_err = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                    \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(ctx, OQG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
static int dalloc(oqg_ctx* c, T** p, size_t n) {
    void* v = nullptr;
    HIPCHK(c, hipMalloc(&v, std::max<size_t>(n, 1) * sizeof(T)));
    HIPCHK(c, hipMemsetAsync(v, 0, std::max<size_t>(n, 1) * sizeof(T), c->st));
    c->allocs.push_back(v);
    *p = static_cast<T*>(v);
    return 0;
}

// ------------------------------------------------------------------ host precompute

// exp(-2 pi i t / n): quadrant reduction, long-double sin/cos of the reduced angle.
static void twiddle_ld(long t, long n, double* re, double* im) {
    t %= n;
    const long q = (4 * t) / n, r = 4 * t - q * n;   // angle = (q + r/n) pi/2
    const long double ang = 1.5707963267948966192313216916397514L * (long double)r / (long double)n;
    long double c = cosl(ang), s = sinl(ang);
    if (r == 0) { c = 1.0L; s = 0.0L; }
    long double C, S;   // cos, sin of the full angle
    switch (q) {
        case 0: C = c; S = s; break;
        case 1: C = -s; S = c; break;
        case 2: C = -c; S = -s; break;
        default: C = s; S = -c; break;
    }
    *re = (double)C;
    *im = (double)-S;
}

// Mode matrices: compute_A (modes.F90:221-246) and derive_qg_modes (:76-161).
// Eigenpairs of A come from the similar symmetric tridiagonal H^1/2 A H^-1/2:
// eigenvalues by Sturm-sequence bisection, eigenvectors by inverse iteration.
static int derive_modes(oqg_ctx* c) {
    const int nl = c->d.nl;
    const double* h = c->d.h;
    const double* gp = c->d.gp;
    double A[MAXL][MAXL] = {};
    A[0][1] = -1.0 / (gp[0] * h[0]);
    A[0][0] = -A[0][1];
    for (int k = 1; k < nl - 1; k++) {
        A[k][k - 1] = -1.0 / (gp[k - 1] * h[k]);
        A[k][k + 1] = -1.0 / (gp[k] * h[k]);
        A[k][k] = -A[k][k - 1] - A[k][k + 1];
    }
    A[nl - 1][nl - 2] = -1.0 / (gp[nl - 2] * h[nl - 1]);
    A[nl - 1][nl - 1] = -A[nl - 1][nl - 2];
    double dg[MAXL], of[MAXL];
    for (int k = 0; k < nl; k++) dg[k] = A[k][k];
    for (int k = 0; k < nl - 1; k++) of[k] = -1.0 / (gp[k] * std::sqrt(h[k] * h[k + 1]));
    double lo = 0, hi = 0;
    for (int k = 0; k < nl; k++) {
        const double rad = (k > 0 ? std::fabs(of[k - 1]) : 0.0) + (k < nl - 1 ? std::fabs(of[k]) : 0.0);
        lo = std::min(lo, dg[k] - rad);
        hi = std::max(hi, dg[k] + rad);
    }
    auto count_below = [&](double x) {
        int cnt = 0;
        double dd = dg[0] - x;
        if (dd < 0) cnt++;
        for (int k = 1; k < nl; k++) {
            if (dd == 0.0) dd = 1e-300;
            dd = (dg[k] - x) - of[k - 1] * of[k - 1] / dd;
            if (dd < 0) cnt++;
        }
        return cnt;
    };
    double lam[MAXL], vec[MAXL][MAXL];   // vec[m][k] eigenvector of the symmetric matrix
    for (int i = 0; i < nl; i++) {
        double a = lo - 1.0, b = hi + 1.0;
        for (int it = 0; it < 200; it++) {
            const double mid = 0.5 * (a + b);
            if (mid <= a || mid >= b) break;
            if (count_below(mid) > i) b = mid;
            else a = mid;
        }
        lam[i] = 0.5 * (a + b);
        // inverse iteration on (T - sigma I)
        const double scale = std::max(std::fabs(lo), std::fabs(hi));
        const double sigma = lam[i] + 1e-13 * scale;
        double v[MAXL];
        for (int k = 0; k < nl; k++) v[k] = 1.0 + 0.01 * k;
        for (int it = 0; it < 4; it++) {
            double M[MAXL][MAXL + 1] = {};
            for (int k = 0; k < nl; k++) {
                M[k][k] = dg[k] - sigma;
                if (k > 0) M[k][k - 1] = of[k - 1];
                if (k < nl - 1) M[k][k + 1] = of[k];
                M[k][nl] = v[k];
            }
            for (int col = 0; col < nl; col++) {
                int piv = col;
                for (int r = col + 1; r < nl; r++)
                    if (std::fabs(M[r][col]) > std::fabs(M[piv][col])) piv = r;
                if (piv != col)
                    for (int t = 0; t <= nl; t++) std::swap(M[col][t], M[piv][t]);
                if (M[col][col] == 0.0) M[col][col] = 1e-300;
                for (int r = col + 1; r < nl; r++) {
                    const double f = M[r][col] / M[col][col];
                    for (int t = col; t <= nl; t++) M[r][t] -= f * M[col][t];
                }
            }
            for (int r = nl - 1; r >= 0; r--) {
                double t = M[r][nl];
                for (int q = r + 1; q < nl; q++) t -= M[r][q] * v[q];
                v[r] = t / M[r][r];
            }
            double nrm = 0.0;
            for (int k = 0; k < nl; k++) nrm += v[k] * v[k];
            nrm = std::sqrt(nrm);
            for (int k = 0; k < nl; k++) v[k] /= nrm;
        }
        for (int k = 0; k < nl; k++) vec[i][k] = v[k];
    }
    // right eigenvectors of A: H^-1/2 v (unit 2-norm, first component positive); left: H^1/2 v
    double er[MAXL][MAXL], el[MAXL][MAXL];
    for (int i = 0; i < nl; i++) {
        double nrm = 0.0;
        for (int k = 0; k < nl; k++) {
            er[i][k] = vec[i][k] / std::sqrt(h[k]);
            el[i][k] = vec[i][k] * std::sqrt(h[k]);
            nrm += er[i][k] * er[i][k];
        }
        nrm = std::sqrt(nrm);
        const double sg = er[i][0] < 0 ? -1.0 : 1.0;
        for (int k = 0; k < nl; k++) { er[i][k] = sg * er[i][k] / nrm; el[i][k] *= sg; }
    }
    // order by |lambda| ascending, stable (modes.F90:99-115)
    int idx[MAXL];
    double ab[MAXL];
    for (int i = 0; i < nl; i++) { idx[i] = i; ab[i] = std::fabs(lam[i]); }
    for (int m = 1; m < nl; m++) {
        const int t = idx[m];
        int i = m - 1;
        while (i >= 0 && ab[idx[i]] > ab[t]) { idx[i + 1] = idx[i]; i--; }
        idx[i + 1] = t;
    }
    std::memset(&c->mx, 0, sizeof(c->mx));
    c->mx.rdm2[0] = 0.0;
    for (int m = 1; m < nl; m++) c->mx.rdm2[m] = c->d.fnot * c->d.fnot * ab[idx[m]];
    for (int m = 0; m < nl; m++) {
        const int im = idx[m];
        double elder = 0.0;
        for (int k = 0; k < nl; k++) elder += el[im][k] * er[im][k];
        for (int k = 0; k < nl; k++) {
            c->mx.ctl2m[k * MAXL + m] = el[im][k] / elder;
            c->mx.ctm2l[m * MAXL + k] = er[im][k];
        }
    }
    for (int k = 0; k < nl; k++)
        for (int l = 0; l < nl; l++) c->mx.amat[k * MAXL + l] = A[k][l];
    return 0;
}

static int check_identity(oqg_ctx* c) {
    const int nl = c->d.nl;
    for (int l = 0; l < nl; l++)
        for (int k = 0; k < nl; k++) {
            double s = 0.0;
            for (int m = 0; m < nl; m++) s += c->mx.ctm2l[m * MAXL + k] * c->mx.ctl2m[l * MAXL + m];
            if (std::fabs(s - (l == k ? 1.0 : 0.0)) > 5.0e-15) {
                char buf[160];
                std::snprintf(buf, sizeof buf, "mode conversion matrices are not inverse: (%d,%d) = %.17g", k + 1, l + 1, s);
                return fail(c, OQG_ERR_NUMERIC, buf);
            }
        }
    return 0;
}

// Host trapezoid integrals (numerics.F90:188-212) on a dense (nxp,nyp) array.
static double host_int_dA(const oqg_ctx* c, const double* f) {
    const int nx = c->g.nxp, ny = c->g.nyp;
    double all = 0, s1 = 0, sn = 0, w1 = 0, wn = 0;
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) all += f[(size_t)j * nx + i];
    for (int i = 0; i < nx; i++) { s1 += f[i]; sn += f[(size_t)(ny - 1) * nx + i]; }
    for (int j = 0; j < ny; j++) { w1 += f[(size_t)j * nx]; wn += f[(size_t)j * nx + nx - 1]; }
    double r = all;
    r = r - 0.5 * (s1 + sn);
    r = r - 0.5 * (w1 + wn);
    r = r + 0.25 * (f[0] + f[(size_t)(ny - 1) * nx] + f[nx - 1] + f[(size_t)(ny - 1) * nx + nx - 1]);
    return r * c->g.dx * c->g.dy;
}

static double host_int_dx(const oqg_ctx* c, const double* row) {
    const int nx = c->g.nxp;
    double s = 0.0;
    for (int i = 1; i < nx - 1; i++) s += row[i];
    return (0.5 * row[0] + s + 0.5 * row[nx - 1]) * c->g.dx;
}

// Partial-pivot LU of the (nl-1)^2 matrix (LU_factor, linalg.F90:24-55).
static int host_lu(oqg_ctx* c, int n) {
    BoxHom& b = c->bh;
    for (int i = 0; i < MAXL * MAXL; i++) b.LU[i] = b.cdhoc[i];
    for (int k = 0; k < n; k++) {
        int piv = k;
        for (int i = k + 1; i < n; i++)
            if (std::fabs(b.LU[i * MAXL + k]) > std::fabs(b.LU[piv * MAXL + k])) piv = i;
        b.ipiv[k] = piv;
        if (b.LU[piv * MAXL + k] == 0.0) return fail(c, OQG_ERR_NUMERIC, "mass constraint matrix cdhoc is singular");
        if (piv != k)
            for (int j = 0; j < n; j++) std::swap(b.LU[k * MAXL + j], b.LU[piv * MAXL + j]);
        for (int i = k + 1; i < n; i++) b.LU[i * MAXL + k] /= b.LU[k * MAXL + k];
        for (int j = k + 1; j < n; j++)
            for (int i = k + 1; i < n; i++) b.LU[i * MAXL + j] -= b.LU[i * MAXL + k] * b.LU[k * MAXL + j];
    }
    return 0;
}

// ------------------------------------------------------------------ copies

// Host fields move through a pinned staging buffer: a pitched copy between
// pageable memory and the device goes row by row (milliseconds per field),
// one between pinned memory and the device is a single DMA.  Host transfers
// are synchronous (the staging buffer is free again on return).
static int host_stage(oqg_ctx* c, size_t words) {
    if (c->hst_words >= words) return 0;
    if (c->hst) HIPCHK(c, hipHostFree(c->hst));
    c->hst = nullptr;
    c->hst_words = 0;
    HIPCHK(c, hipHostMalloc((void**)&c->hst, words * sizeof(double), hipHostMallocDefault));
    c->hst_words = words;
    return 0;
}

static int upload_field(oqg_ctx* c, double* dst, const double* src, int nlayers, int kind) {
    const size_t w = (size_t)c->g.nxp * sizeof(double), rows = (size_t)nlayers * c->g.nyp;
    const size_t dp = (size_t)c->g.pitch * sizeof(double);
    if (kind == OQG_DEVICE) {
        HIPCHK(c, hipMemcpy2DAsync(dst, dp, src, w, w, rows, hipMemcpyDeviceToDevice, c->st));
        return 0;
    }
    if (int r = host_stage(c, rows * c->g.nxp)) return r;
    std::memcpy(c->hst, src, rows * w);
    HIPCHK(c, hipMemcpy2DAsync(dst, dp, c->hst, w, w, rows, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

static int download_field(oqg_ctx* c, double* dst, const double* src, int nlayers, int kind) {
    const size_t w = (size_t)c->g.nxp * sizeof(double), rows = (size_t)nlayers * c->g.nyp;
    const size_t sp = (size_t)c->g.pitch * sizeof(double);
    if (kind == OQG_DEVICE) {
        HIPCHK(c, hipMemcpy2DAsync(dst, w, src, sp, w, rows, hipMemcpyDeviceToDevice, c->st));
        return 0;
    }
    if (int r = host_stage(c, rows * c->g.nxp)) return r;
    HIPCHK(c, hipMemcpy2DAsync(c->hst, w, src, sp, w, rows, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    std::memcpy(dst, c->hst, rows * w);
    return 0;
}

// ------------------------------------------------------------------ launches

static void fill_inv_common(oqg_ctx* c, InvArgs& ia) {
    std::memset(&ia, 0, sizeof ia);
    ia.g = c->g;
    ia.s = c->s;
    ia.ph = &c->dconst->ph;
    ia.dc = c->dconst;
    ia.spec = c->spec;
    ia.tw = c->itw;
    ia.twf = c->tw;
    ia.colsum = c->colsum;
    ia.colw = c->colw;
    ia.entpart = c->entpart;
    ia.nentpart = c->tiles_x * c->tiles_y;
    ia.hom = c->hom;
    ia.cpart = c->cpart;
    ia.ctiles = c->tiles_x;
    ia.fcon = c->fcon;
    ia.yprel = c->yprel;
    ia.ddyn = c->ddyn;
    ia.raw_mode = -1;
}

// Raw Helmholtz solve of mode m (0-based): out = H_m^-1 in on the interior,
// 0 on the boundary (default buffers raw[0] -> raw[1]).
static int enqueue_raw_solve(oqg_ctx* c, int m, const double* in = nullptr, double* out = nullptr) {
    FwdArgs fa{};
    fa.g = c->g;
    fa.s = c->s;
    fa.dc = c->dconst;
    fa.q = in ? in : c->raw[0];
    fa.yprel = c->yprel;
    fa.tw = c->tw;
    fa.spec = c->spec;
    fa.raw_mode = m;
    launch_fwd(fa, 1, c->st);
    YArgs ya{};
    ya.s = c->s;
    ya.spec = c->spec;
    ya.bcol = c->bcol;
    ya.gam = c->gamw;
    ya.mode0 = m;
    ya.nmodes = 1;
    launch_ysolve(ya, c->variant, c->st);
    InvArgs ia;
    fill_inv_common(c, ia);
    ia.raw_mode = m;
    ia.out_raw = out ? out : c->raw[1];
    launch_inv(ia, c->st);
    HIPCHK(c, hipGetLastError());
    return 0;
}


enum { KV = 0, KF, KY, KI, NK };
static const char* kNames[NK] = {"k_vort", "k_fwd", "k_ysolve", "k_inv"};

// Arguments of the five kernels of one step_qg from the current time level.
struct StepArgs {
    VortArgs va;
    FwdArgs fa;
    YArgs ya;
    InvArgs ia;
};

// The working buffers inside the slot-shaped allocation base.
static void set_work(oqg_ctx* c, double* base) {
    const long long fl = (long long)c->g.nl * c->g.fs;
    c->wbase = base;
    c->qbuf[0] = base;
    c->qbuf[1] = base + fl;
    c->pbuf[0] = base + 2 * fl;
    c->pbuf[1] = base + 3 * fl;
    c->cbuf[0] = base + 4 * fl;
    c->cbuf[1] = base + 4 * fl + C_LEN;
}

static void make_step_args(oqg_ctx* c, double tdt, StepArgs& a) {
    const int cur = c->cur, nxt = 1 - cur;
    std::memset(&a, 0, sizeof a);
    VortArgs& va = a.va;
    va.g = c->g;
    va.ph = &c->dconst->ph;
    // inputs: the current and previous levels (from a checkpoint when a step
    // starts from one); outputs: the working buffers of the next level
    va.q = c->ext_cur.q ? c->ext_cur.q : c->qbuf[cur];
    va.p = c->ext_cur.q ? c->ext_cur.p : c->pbuf[cur];
    va.pm = c->ext_old.q ? c->ext_old.p : c->pbuf[nxt];
    va.qm = c->ext_old.q ? c->ext_old.q : c->qbuf[nxt];
    va.qnew = c->qbuf[nxt];
    va.ent1 = c->ent1;
    va.wekp = c->wekp;
    va.entpart = c->entpart;
    va.tdt = tdt;
    va.tiles_x = c->tiles_x;
    va.tiles_y = c->tiles_y;
    va.ty = c->vort_ty;
    va.org_x = vort_org_x(c->g);
    va.cpart = c->g.cyclic ? c->cpart : nullptr;
    va.ent2s = c->ent2s;
    va.ent2n = c->ent2n;
    va.fcon = c->fcon;
    FwdArgs& fa = a.fa;
    fa.g = c->g;
    fa.s = c->s;
    fa.dc = c->dconst;
    fa.q = c->qbuf[nxt];
    fa.ddyn = c->ddyn;
    fa.yprel = c->yprel;
    fa.tw = c->tw;
    fa.spec = c->spec;
    fa.raw_mode = -1;
    YArgs& ya = a.ya;
    ya.s = c->s;
    ya.spec = c->spec;
    ya.bcol = c->bcol;
    ya.colsum = c->g.cyclic ? nullptr : c->colsum;
    ya.gam = c->gamw;
    ya.mode0 = 0;
    ya.nmodes = c->g.nl;
    InvArgs& ia = a.ia;
    fill_inv_common(c, ia);
    ia.cin = c->ext_cur.q ? c->ext_cur.c : c->cbuf[cur];
    ia.cout = c->cbuf[nxt];
    ia.pnew = c->pbuf[nxt];
    ia.qnew = c->qbuf[nxt];
    ia.tdt = tdt;
}

static void launch_kernel(oqg_ctx* c, const StepArgs& a, int k) {
    switch (k) {
        case KV: launch_vort(a.va, c->st); break;
        case KF: launch_fwd(a.fa, c->g.nl, c->st); break;
        case KY: launch_ysolve(a.ya, c->variant, c->st); break;
        default: launch_inv(a.ia, c->st); break;
    }
}

// One step_qg on resident forcing.  The new level goes to the working
// buffers; the level that was current becomes the previous one, wherever it is.
static int enqueue_step(oqg_ctx* c, double tdt) {
    StepArgs a;
    make_step_args(c, tdt, a);
    for (int k = 0; k < NK; k++) launch_kernel(c, a, k);
    HIPCHK(c, hipGetLastError());
    c->cur = 1 - c->cur;
    c->ext_old = c->ext_cur;
    c->ext_cur = {};
    return 0;
}

static int enqueue_suppress(oqg_ctx* c) {
    if (c->ext_cur.q) return fail(c, OQG_ERR_STATE, "suppression of a level the working buffers do not hold");
    SuppArgs sa{};
    sa.g = c->g;
    sa.q = c->qbuf[c->cur];
    sa.qm = c->ext_old.q ? c->ext_old.q : c->qbuf[1 - c->cur];
    sa.p = c->pbuf[c->cur];
    sa.pm = c->ext_old.q ? c->ext_old.p : c->pbuf[1 - c->cur];
    sa.c = c->cbuf[c->cur];
    launch_suppress(sa, c->st);
    HIPCHK(c, hipGetLastError());
    return 0;
}

// ------------------------------------------------------------------ create

static int setup(oqg_ctx* c) {
    const oqg_desc& d = c->d;
    const int nl = d.nl, nxp = d.nxp, nyp = d.nyp;
    Geo& g = c->g;
    g.nxp = nxp;
    g.nyp = nyp;
    g.nl = nl;
    g.pitch = (nxp + 15) / 16 * 16;
    g.cyclic = d.cyclic ? 1 : 0;
    const int kt1 = d.k_topo > 0 ? d.k_topo : (d.dz_sign == 1 ? 1 : nl);
    g.kt = kt1 - 1;
    g.fs = (long long)g.pitch * nyp;
    g.dx = d.dx;
    g.dy = d.dy;
    g.fnot = d.fnot;
    g.beta = d.beta;
    g.fx = 1.0 / (d.dx * d.dx);
    g.fy = 1.0 / (d.dy * d.dy);
    const double al = d.bcco;
    g.cxl = g.fx * al / (0.5 * al + 1.0);
    g.cxr = -g.fx * al / (0.5 * al + 1.0);
    g.cyl = g.fy * al / (0.5 * al + 1.0);
    g.cyr = -g.fy * al / (0.5 * al + 1.0);
    g.jfac = 1.0 / (3.0 * (2.0 * d.dx) * (2.0 * d.dy) * d.fnot);
    c->fbytes = (size_t)nl * g.fs * sizeof(double);

    Phys& ph = c->ph;
    for (int k = 0; k < nl; k++) {
        ph.c2[k] = d.ah2[k] / d.fnot;
        ph.c4[k] = d.ah4[k] / d.fnot;
        ph.fohfac[k] = d.fnot / d.h[k];
        ph.ah2[k] = d.ah2[k];
        ph.ah4[k] = d.ah4[k];
        ph.h[k] = d.h[k];
    }
    for (int k = 0; k < nl - 1; k++) ph.gp[k] = d.gp[k];
    ph.cent1 = d.dz_sign * ph.fohfac[0];
    ph.cwek = d.tau_sign * ph.fohfac[0];
    ph.cent2 = d.dz_sign * ph.fohfac[1];
    ph.cdrag = (ph.fohfac[g.kt] * d.delek / (2.0 * std::fabs(d.fnot)));
    ph.cd_bdr = d.fnot * d.delek / (2.0 * std::fabs(d.fnot));
    ph.qbfac = d.bcco * (1.0 / (d.dx * d.dx)) / (0.5 * d.bcco + 1.0) / d.fnot;
    ph.rdy = 1.0 / d.dy;
    ph.rdx2 = 1.0 / (2.0 * d.dx);
    ph.dz_sign = d.dz_sign;
    ph.tau_sign = d.tau_sign;

    // spectral layout and FFT plan (init_inhomog, inhomog.F90:83-124)
    Spec& s = c->s;
    int Nref;
    if (g.cyclic) {
        s.nin = nxp - 1;
        s.L = s.nin;
        s.ncol = 2 * (s.nin / 2 + 1);
        Nref = s.nin;
    } else {
        s.nin = nxp - 2;
        s.L = 2 * (s.nin + 1);
        s.ncol = s.nin;
        Nref = s.L;
    }
    if (s.L > MAX_FFT) return fail(c, OQG_ERR_ARG, "x transform longer than 4096 points");
    s.nyi = nyp - 2;
    s.nrow = s.nyi + 1;
    s.kp = (s.ncol + 15) / 16 * 16;
    s.R = (s.nyi + WAVE - 1) / WAVE;
    if (s.R > MAX_YROWS_PER_LANE) return fail(c, OQG_ERR_ARG, "nyp too large for the y solver (nyp-2 > 1024)");
    s.nypad = WAVE * s.R;
    s.a = 1.0 / (d.dy * d.dy);
    s.ftnorm = 1.0 / Nref;
    {
        // the fewest LDS passes within each kernel's largest butterfly
        auto plan = [&](int maxr, int& nst, int* radix) {
            int m = s.L;
            nst = 0;
            for (int r : {16, 8, 4, 2, 3, 5, 7, 11, 13})
                if (r <= maxr || r == 2 || r % 2 == 1)   // 16, 8, 4 up to the kernel's largest; all primes
                    while (m % r == 0) {
                        if (nst == MAX_RADIX_STAGES) return fail(c, OQG_ERR_ARG, "x transform plan too long");
                        radix[nst++] = r;
                        m /= r;
                    }
            if (m != 1)
                return fail(c, OQG_ERR_ARG, "x transform length " + std::to_string(s.L) +
                                                " must factor into 2, 3, 5, 7, 11 and 13");
            for (int st = 0; st < nst; st++)   // each thread holds at most 4 butterflies per stage
                if (s.L / radix[st] > 4 * 256) return fail(c, OQG_ERR_ARG, "x transform too long for the LDS FFT");
            return 0;
        };
        if (int r = plan(FWD_MAX_RADIX, s.nstages, s.radix)) return r;
        if (int r = plan(INV_MAX_RADIX, s.instages, s.iradix)) return r;
        s.prime = (s.L % 7 == 0 || s.L % 11 == 0 || s.L % 13 == 0) ? 1 : 0;
    }
    if (inv_lds_bytes(s, g, nl) > MAX_DYN_LDS || fwd_lds_bytes(s) > MAX_DYN_LDS || ysolve_lds_bytes(s) > MAX_DYN_LDS)
        return fail(c, OQG_ERR_ARG, "grid too large for the LDS-staged transforms");

    // modes
    if (d.ctl2m && d.ctm2l && d.rdm2 && d.amat) {
        std::memset(&c->mx, 0, sizeof(c->mx));
        for (int k = 0; k < nl; k++)
            for (int m = 0; m < nl; m++) {
                c->mx.ctl2m[k * MAXL + m] = d.ctl2m[m * nl + k];   // Fortran (k,m)
                c->mx.ctm2l[m * MAXL + k] = d.ctm2l[k * nl + m];   // Fortran (m,k)
                c->mx.amat[k * MAXL + m] = d.amat[m * nl + k];
            }
        for (int m = 0; m < nl; m++) c->mx.rdm2[m] = d.rdm2[m];
    } else {
        int rc = derive_modes(c);
        if (rc) return rc;
    }
    if (int rc = check_identity(c)) return rc;

    // device buffers
    const size_t fl = (size_t)nl * g.fs, f1 = (size_t)g.fs;
    const size_t nspec = (size_t)nl * s.nrow * s.kp;
    c->tiles_x = (vort_lim_x(g) - vort_org_x(g) + VORT_TX - 1) / VORT_TX;
    if (g.cyclic && c->tiles_x > CYC_MAX_TX) return fail(c, OQG_ERR_ARG, "channel too wide for the boundary-integral partials");
    c->vort_ty = vort_rows(c->tiles_x, nyp, nl);
    c->tiles_y = (nyp - 2 + c->vort_ty - 1) / c->vort_ty;
    int rc = dalloc(c, &c->wbase, 4 * fl + 2 * (size_t)C_LEN);
    if (!rc) {
        c->wbase_own = c->wbase;
        set_work(c, c->wbase);
    }
    for (int t = 0; t < 2 && !rc; t++) rc |= dalloc(c, &c->raw[t], f1);
    if (!rc) rc = dalloc(c, &c->spec, nspec);
    if (!rc && c->variant == 1) rc = dalloc(c, &c->gamw, nspec);
    if (!rc) rc = dalloc(c, &c->colsum, (size_t)nl * s.ncol);
    if (!rc) rc = dalloc(c, &c->colw, (size_t)s.ncol);
    // sized for the shortest tiles, so that the tile height can change (oqg_set_tile_rows)
    if (!rc) rc = dalloc(c, &c->entpart, (size_t)c->tiles_x * ((nyp - 2 + 3) / 4));
    if (!rc && g.cyclic) rc = dalloc(c, &c->cpart, (size_t)2 * nl * c->tiles_x * CYC_SLOTS);
    if (!rc) rc = dalloc(c, &c->fcon, F_LEN);
    if (!rc) rc = dalloc(c, &c->ent1, f1);
    if (!rc) rc = dalloc(c, &c->ent2s, (size_t)nxp);
    if (!rc) rc = dalloc(c, &c->ent2n, (size_t)nxp);
    if (!rc) rc = dalloc(c, &c->wekp, f1);
    if (!rc) rc = dalloc(c, &c->yprel, (size_t)nyp);
    if (!rc) rc = dalloc(c, &c->bcol, (size_t)nl * s.ncol);
    if (!rc) rc = dalloc(c, &c->tw, 2 * (size_t)twiddle_table_len(s.L));
    if (!rc) rc = dalloc(c, &c->itw, 2 * (size_t)twiddle_table_len(s.L));
    if (!rc && d.ddyn) rc = dalloc(c, &c->ddyn, f1);
    if (!rc) rc = dalloc(c, &c->hom, g.cyclic ? (size_t)nyp * (1 + 2 * (nl - 1)) : (size_t)(nl - 1) * f1);
    if (!rc) rc = dalloc(c, &c->dconst, 1);
    if (rc) return rc;
    auto upload_consts = [&]() -> int {
        DevConsts h;
        h.mx = c->mx;
        h.bh = c->bh;
        h.ch = c->ch;
        h.ph = c->ph;
        HIPCHK(c, hipMemcpyAsync(c->dconst, &h, sizeof h, hipMemcpyHostToDevice, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        return 0;
    };
    if (int r = upload_consts()) return r;

    // constants
    const double xl = (nxp - 1) * d.dx, yl = (nyp - 1) * d.dy;   // box.F90:135-136
    c->yprel_h.resize(nyp);
    for (int j = 0; j < nyp; j++) c->yprel_h[j] = -0.5 * yl + j * d.dy;
    HIPCHK(c, hipMemcpyAsync(c->yprel, c->yprel_h.data(), nyp * sizeof(double), hipMemcpyHostToDevice, c->st));
    if (d.ddyn) {
        c->ddyn_h.assign(d.ddyn, d.ddyn + (size_t)nxp * nyp);
        if (int r = upload_field(c, c->ddyn, d.ddyn, 1, OQG_HOST)) return r;
    }
    for (int inv = 0; inv < 2; inv++) {
        // the plan's twiddle table in the stage-by-stage layout the LDS FFT reads
        const int L = s.L, nst = inv ? s.instages : s.nstages;
        const int* radix = inv ? s.iradix : s.radix;
        std::vector<double> tw(2 * (size_t)twiddle_table_len(L), 0.0);
        size_t e = 0;
        int Ns = 1;
        for (int st = 0; st < nst; st++) {
            const int R = radix[st], tstep = L / (Ns * R);
            for (int r = 1; r < R; r++)
                for (int k = 0; k < Ns; k++, e++) twiddle_ld((long)r * k * tstep, L, &tw[2 * e], &tw[2 * e + 1]);
            Ns *= R;
        }
        for (int m = 1; m < 3 && L % 3 == 0; m++) {
            const size_t o = twiddle_roots3(L) + m - 1;
            twiddle_ld((long)m * (L / 3), L, &tw[2 * o], &tw[2 * o + 1]);
        }
        for (int m = 1; m < 5 && L % 5 == 0; m++) {
            const size_t o = twiddle_roots5(L) + m - 1;
            twiddle_ld((long)m * (L / 5), L, &tw[2 * o], &tw[2 * o + 1]);
        }
        for (int R : {7, 11, 13}) {
            if (L % R) continue;
            const int base = R == 7 ? twiddle_roots7(L) : (R == 11 ? twiddle_roots11(L) : twiddle_roots13(L));
            for (int m = 1; m < R; m++) {
                const size_t o = base + m - 1;
                twiddle_ld((long)m * (L / R), L, &tw[2 * o], &tw[2 * o + 1]);
            }
        }
        HIPCHK(c, hipMemcpy(inv ? c->itw : c->tw, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    // tridiagonal diagonals bb(k,m) (inhomog.F90:127-141)
    std::vector<double> bcol((size_t)nl * s.ncol);
    const double dxm2 = 1.0 / (d.dx * d.dx);
    for (int col = 0; col < s.ncol; col++) {
        double arg;
        if (g.cyclic) arg = ((double)(col / 2) * TWOPI_REF) / (double)Nref;
        else arg = ((double)(col + 1) * TWOPI_REF) / (double)Nref;
        const double bd2 = -2.0 * s.a + 2.0 * dxm2 * (std::cos(arg) - 1.0);
        for (int m = 0; m < nl; m++) bcol[(size_t)m * s.ncol + col] = bd2 - c->mx.rdm2[m];
    }
    HIPCHK(c, hipMemcpyAsync(c->bcol, bcol.data(), bcol.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
    if (!g.cyclic) {
        // sum over interior i of the inverse DST basis: 2*ftnorm*cot(pi K / (2N)) for odd K
        std::vector<double> colw(s.ncol);
        const long double N = (long double)(nxp - 1);
        for (int K = 1; K <= s.ncol; K++) {
            long double v = 0.0L;
            if (K % 2 == 1) {
                const long double th = 3.14159265358979323846264338327950288L * K / (2.0L * N);
                v = 2.0L * (long double)s.ftnorm * cosl(th) / sinl(th);
            }
            colw[K - 1] = (double)v;
        }
        HIPCHK(c, hipMemcpyAsync(c->colw, colw.data(), colw.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
    }
    set_kernel_attrs();

    // homogeneous solutions (homsol, homog.F90:108-278) from this library's own solver
    const size_t nxy = (size_t)nxp * nyp;
    std::vector<double> L1(nxy), sol(nxy);
    if (!g.cyclic) {
        std::vector<double> homh((size_t)(nl - 1) * nxy);
        double aipohs[MAXL] = {};
        if (d.hom_sol_bc) {
            std::memcpy(homh.data(), d.hom_sol_bc, homh.size() * sizeof(double));
        } else {
            std::fill(L1.begin(), L1.end(), 1.0);
            if (int r = upload_field(c, c->raw[0], L1.data(), 1, OQG_HOST)) return r;
            for (int m = 1; m < nl; m++) {
                if (int r = enqueue_raw_solve(c, m)) return r;
                if (int r = download_field(c, sol.data(), c->raw[1], 1, OQG_HOST)) return r;
                HIPCHK(c, hipStreamSynchronize(c->st));
                double* hm = homh.data() + (size_t)(m - 1) * nxy;
                for (size_t t = 0; t < nxy; t++) hm[t] = L1[t] + c->mx.rdm2[m] * sol[t];
            }
        }
        for (int m = 1; m < nl; m++) aipohs[m] = host_int_dA(c, homh.data() + (size_t)(m - 1) * nxy);
        std::memset(&c->bh, 0, sizeof(c->bh));
        for (int k = 0; k < nl - 1; k++)
            for (int m = 1; m < nl; m++)
                c->bh.cdhoc[k * MAXL + (m - 1)] =
                    (c->mx.ctm2l[m * MAXL + k + 1] - c->mx.ctm2l[m * MAXL + k]) * aipohs[m];
        if (int r = host_lu(c, nl - 1)) return r;
        if (int r = upload_field(c, c->hom, homh.data(), nl - 1, OQG_HOST)) return r;
    } else {
        std::vector<double> hb((size_t)nyp * (1 + 2 * (nl - 1)));
        double* bt = hb.data();
        double* b1 = bt + nyp;
        double* b2 = b1 + (size_t)(nl - 1) * nyp;
        std::memset(&c->ch, 0, sizeof(c->ch));
        if (d.hom_bt && d.hom_bc1 && d.hom_bc2 && d.hom_coef) {
            std::memcpy(bt, d.hom_bt, nyp * sizeof(double));
            std::memcpy(b1, d.hom_bc1, (size_t)(nl - 1) * nyp * sizeof(double));
            std::memcpy(b2, d.hom_bc2, (size_t)(nl - 1) * nyp * sizeof(double));
            for (int m = 1; m < nl; m++) {
                c->ch.hc1s[m] = d.hom_coef[4 * (m - 1) + 0];
                c->ch.hc2s[m] = d.hom_coef[4 * (m - 1) + 1];
                c->ch.hc1n[m] = d.hom_coef[4 * (m - 1) + 2];
                c->ch.hc2n[m] = d.hom_coef[4 * (m - 1) + 3];
            }
            c->ch.hbsi = d.hom_coef[4 * (nl - 1)];
        } else {
            for (int j = 0; j < nyp; j++) bt[j] = (double)(nyp - 1 - j) / (double)(nyp - 1);
            c->ch.hbsi = yl / xl;
            std::vector<double> r2(nxy);
            const double ypn = (nyp - 1) * d.dy, yp1 = 0.0;
            for (int j = 0; j < nyp; j++)
                for (int i = 0; i < nxp; i++) {
                    const double ypj = j * d.dy;
                    L1[(size_t)j * nxp + i] = (ypn - ypj) / yl;
                    r2[(size_t)j * nxp + i] = (ypj - yp1) / yl;
                }
            for (int m = 1; m < nl; m++) {
                for (int which = 0; which < 2; which++) {
                    const std::vector<double>& rr = which ? r2 : L1;
                    if (int r = upload_field(c, c->raw[0], rr.data(), 1, OQG_HOST)) return r;
                    if (int r = enqueue_raw_solve(c, m)) return r;
                    if (int r = download_field(c, sol.data(), c->raw[1], 1, OQG_HOST)) return r;
                    HIPCHK(c, hipStreamSynchronize(c->st));
                    double* dst = (which ? b2 : b1) + (size_t)(m - 1) * nyp;
                    for (int j = 0; j < nyp; j++)
                        dst[j] = rr[(size_t)j * nxp] + c->mx.rdm2[m] * sol[(size_t)j * nxp];
                }
            }
            const double dy = d.dy;
            for (int m = 1; m < nl; m++) {
                const double* h1 = b1 + (size_t)(m - 1) * nyp;
                const double* h2 = b2 + (size_t)(m - 1) * nyp;
                double p1s = -xl * (h1[1] - h1[0]) / dy;
                double p2s = -xl * (h2[1] - h2[0]) / dy;
                double p1n = xl * (h1[nyp - 1] - h1[nyp - 2]) / dy;
                double p2n = xl * (h2[nyp - 1] - h2[nyp - 2]) / dy;
                p1s = p1s + 0.5 * dy * c->mx.rdm2[m] * h1[0] * xl;
                p2s = p2s + 0.5 * dy * c->mx.rdm2[m] * h2[0] * xl;
                p1n = p1n + 0.5 * dy * c->mx.rdm2[m] * h1[nyp - 1] * xl;
                p2n = p2n + 0.5 * dy * c->mx.rdm2[m] * h2[nyp - 1] * xl;
                const double det = p1s * p2n - p2s * p1n;
                c->ch.hc1s[m] = p1s / det;
                c->ch.hc2s[m] = p2s / det;
                c->ch.hc1n[m] = p1n / det;
                c->ch.hc2n[m] = p2n / det;
            }
        }
        c->ch.c0 = d.dx * (s.ftnorm * (double)Nref);
        HIPCHK(c, hipMemcpyAsync(c->hom, hb.data(), hb.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
    }
    return upload_consts();
}

// ------------------------------------------------------------------ ABI

extern "C" {

int oqg_create(const oqg_desc* desc, oqg_ctx** out) {
    g_create_err.clear();
    if (!out) return fail(nullptr, OQG_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!desc) return fail(nullptr, OQG_ERR_ARG, "desc is NULL");
    const oqg_desc& d = *desc;
    if (d.nl < 2 || d.nl > OQG_MAX_LAYERS) return fail(nullptr, OQG_ERR_ARG, "nl must be in 2..8");
    if (d.nxp < 5 || d.nyp < 8) return fail(nullptr, OQG_ERR_ARG, "grid too small (nxp >= 5, nyp >= 8)");
    if (!(d.dx > 0) || !(d.dy > 0) || d.fnot == 0.0) return fail(nullptr, OQG_ERR_ARG, "dx, dy must be > 0 and fnot != 0");
    for (int k = 0; k < d.nl; k++)
        if (!(d.h[k] > 0)) return fail(nullptr, OQG_ERR_ARG, "layer thicknesses must be > 0");
    for (int k = 0; k < d.nl - 1; k++)
        if (!(d.gp[k] > 0)) return fail(nullptr, OQG_ERR_ARG, "reduced gravities must be > 0");
    if (d.k_topo < 0 || d.k_topo > d.nl) return fail(nullptr, OQG_ERR_ARG, "k_topo out of range");
    if (d.cyclic && d.nxp - 1 > 1024) return fail(nullptr, OQG_ERR_ARG, "cyclic nxp-1 must be <= 1024");
    if (!d.cyclic && d.nxp > 1025) return fail(nullptr, OQG_ERR_ARG, "box nxp must be <= 1025");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(nullptr, OQG_ERR_HIP, "no HIP device available");
    auto c = std::make_unique<oqg_ctx>();
    if (hipGetDevice(&c->dev) != hipSuccess) return fail(nullptr, OQG_ERR_HIP, "hipGetDevice failed");
    // the device address of oqg_diagnose_noalloc's scratch array, once per device
    if (hipError_t e = noalloc_scratch(c->dev, &c->scratch); e != hipSuccess)
        return fail(nullptr, OQG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    c->d = d;
    c->d.ddyn = nullptr;   // borrowed pointers are not kept
    c->d.amat = c->d.ctl2m = c->d.ctm2l = c->d.rdm2 = nullptr;
    c->d.hom_sol_bc = c->d.hom_bt = c->d.hom_bc1 = c->d.hom_bc2 = c->d.hom_coef = nullptr;
    c->variant = (d.flags & OQG_FLAG_YSOLVE_THOMAS) ? 1 : 0;
    if (pool_take(&c->own) != hipSuccess) return fail(nullptr, OQG_ERR_HIP, "hipStreamCreate failed");
    c->st = c->own;
    // setup() may read the caller's optional arrays through a temporary copy of desc
    oqg_desc keep = c->d;
    c->d = d;
    int rc = setup(c.get());
    c->d = keep;
    if (rc) {
        g_create_err = c->err;
        oqg_destroy(c.release());
        return rc;
    }
    *out = c.release();
    return 0;
}

void oqg_destroy(oqg_ctx* c) {
    if (!c) return;
    if (c->st) hipStreamSynchronize(c->st);
    for (auto& kv : c->graphs) hipGraphExecDestroy(kv.second);
    for (void* p : c->allocs) hipFree(p);
    if (c->hst) hipHostFree(c->hst);
    pool_give(c->dev, c->own);
    delete c;
}

const char* oqg_last_error(const oqg_ctx* c) { return c ? c->err.c_str() : g_create_err.c_str(); }

int oqg_set_state(oqg_ctx* c, const double* q, const double* qm, const double* p, const double* pm, int kind) {
    ctx_join(c);
    if (!c) return OQG_ERR_ARG;
    if (c->g.cyclic && kind == OQG_HOST) {
        const double* fs[4] = {q, qm, p, pm};
        const int nx = c->g.nxp;
        for (int f = 0; f < 4; f++) {
            if (!fs[f]) continue;
            for (size_t r = 0; r < (size_t)c->g.nl * c->g.nyp; r++)
                if (fs[f][r * nx] != fs[f][r * nx + nx - 1])
                    return fail(c, OQG_ERR_ARG, "cyclic field: column nxp differs from column 1");
        }
    }
    const int cur = c->cur;
    if (q) if (int r = upload_field(c, c->qbuf[cur], q, c->g.nl, kind)) return r;
    if (qm) if (int r = upload_field(c, c->qbuf[1 - cur], qm, c->g.nl, kind)) return r;
    if (p) if (int r = upload_field(c, c->pbuf[cur], p, c->g.nl, kind)) return r;
    if (pm) if (int r = upload_field(c, c->pbuf[1 - cur], pm, c->g.nl, kind)) return r;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

int oqg_get_state(oqg_ctx* c, double* q, double* qm, double* p, double* pm, int kind) {
    ctx_join(c);
    if (!c) return OQG_ERR_ARG;
    const int cur = c->cur;
    if (q) if (int r = download_field(c, q, c->qbuf[cur], c->g.nl, kind)) return r;
    if (qm) if (int r = download_field(c, qm, c->qbuf[1 - cur], c->g.nl, kind)) return r;
    if (p) if (int r = download_field(c, p, c->pbuf[cur], c->g.nl, kind)) return r;
    if (pm) if (int r = download_field(c, pm, c->pbuf[1 - cur], c->g.nl, kind)) return r;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

int oqg_get_layer(oqg_ctx* c, int which, int k, double* out, int kind) {
    ctx_join(c);
    if (!c || !out) return OQG_ERR_ARG;
    if (k < 1 || k > c->g.nl) return fail(c, OQG_ERR_ARG, "layer index out of range");
    const double* src;
    switch (which) {
        case OQG_FIELD_Q: src = c->qbuf[c->cur]; break;
        case OQG_FIELD_QM: src = c->qbuf[1 - c->cur]; break;
        case OQG_FIELD_P: src = c->pbuf[c->cur]; break;
        case OQG_FIELD_PM: src = c->pbuf[1 - c->cur]; break;
        default: return fail(c, OQG_ERR_ARG, "unknown field selector");
    }
    if (int r = download_field(c, out, src + (size_t)(k - 1) * c->g.fs, 1, kind)) return r;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

int oqg_constr_len(const oqg_ctx* c) { return c ? 2 * (c->g.nl - 1) + 4 * c->g.nl : 0; }

int oqg_get_constr(oqg_ctx* c, double* v) {
    ctx_join(c);
    if (!c || !v) return OQG_ERR_ARG;
    double h[C_LEN];
    HIPCHK(c, hipMemcpyAsync(h, c->cbuf[c->cur], sizeof h, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    const int nl = c->g.nl;
    int t = 0;
    for (int k = 0; k < nl - 1; k++) v[t++] = h[C_DPI + k];
    for (int k = 0; k < nl - 1; k++) v[t++] = h[C_DPIP + k];
    for (int k = 0; k < nl; k++) v[t++] = h[C_CS + k];
    for (int k = 0; k < nl; k++) v[t++] = h[C_CN + k];
    for (int k = 0; k < nl; k++) v[t++] = h[C_CSP + k];
    for (int k = 0; k < nl; k++) v[t++] = h[C_CNP + k];
    return 0;
}

int oqg_set_constr(oqg_ctx* c, const double* v) {
    ctx_join(c);
    if (!c || !v) return OQG_ERR_ARG;
    double h[C_LEN] = {};
    const int nl = c->g.nl;
    int t = 0;
    for (int k = 0; k < nl - 1; k++) h[C_DPI + k] = v[t++];
    for (int k = 0; k < nl - 1; k++) h[C_DPIP + k] = v[t++];
    for (int k = 0; k < nl; k++) h[C_CS + k] = v[t++];
    for (int k = 0; k < nl; k++) h[C_CN + k] = v[t++];
    for (int k = 0; k < nl; k++) h[C_CSP + k] = v[t++];
    for (int k = 0; k < nl; k++) h[C_CNP + k] = v[t++];
    HIPCHK(c, hipMemcpyAsync(c->cbuf[c->cur], h, sizeof h, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

int oqg_init_state(oqg_ctx* c) {
    ctx_join(c);
    if (!c) return OQG_ERR_ARG;
    const int cur = c->cur;
    for (int t = 0; t < 2; t++) {
        QcompArgs qa{};
        qa.g = c->g;
        qa.mx = c->mx;
        qa.p = t == 0 ? c->pbuf[cur] : c->pbuf[1 - cur];
        qa.ddyn = c->ddyn;
        qa.yprel = c->yprel;
        qa.q = t == 0 ? c->qbuf[cur] : c->qbuf[1 - cur];
        launch_qcomp(qa, c->st);
    }
    HIPCHK(c, hipGetLastError());
    // constraint integrals on the host (init only)
    const int nl = c->g.nl, nx = c->g.nxp, ny = c->g.nyp;
    const size_t nxy = (size_t)nx * ny;
    std::vector<double> p(nxy * nl), pm(nxy * nl);
    if (int r = download_field(c, p.data(), c->pbuf[cur], nl, OQG_HOST)) return r;
    if (int r = download_field(c, pm.data(), c->pbuf[1 - cur], nl, OQG_HOST)) return r;
    HIPCHK(c, hipStreamSynchronize(c->st));
    double h[C_LEN] = {};
    if (c->g.cyclic) {
        // init_constraint (constraint.F90:114-141)
        double pins[MAXL], pinn[MAXL], pinsp[MAXL], pinnp[MAXL];
        std::vector<double> row(nx);
        const double dy = c->g.dy, f0 = c->g.fnot;
        for (int k = 0; k < nl; k++) {
            const double* P = p.data() + k * nxy;
            const double* PM = pm.data() + k * nxy;
            pinsp[k] = host_int_dx(c, PM);
            pins[k] = host_int_dx(c, P);
            pinnp[k] = host_int_dx(c, PM + (size_t)(ny - 1) * nx);
            pinn[k] = host_int_dx(c, P + (size_t)(ny - 1) * nx);
            for (int i = 0; i < nx; i++) row[i] = PM[nx + i] - PM[i];
            h[C_CSP + k] = host_int_dx(c, row.data()) / dy;
            for (int i = 0; i < nx; i++) row[i] = PM[(size_t)(ny - 1) * nx + i] - PM[(size_t)(ny - 2) * nx + i];
            h[C_CNP + k] = host_int_dx(c, row.data()) / dy;
            for (int i = 0; i < nx; i++) row[i] = P[nx + i] - P[i];
            h[C_CS + k] = host_int_dx(c, row.data()) / dy;
            for (int i = 0; i < nx; i++) row[i] = P[(size_t)(ny - 1) * nx + i] - P[(size_t)(ny - 2) * nx + i];
            h[C_CN + k] = host_int_dx(c, row.data()) / dy;
        }
        for (int k = 0; k < nl; k++) {
            double apsp = 0, apnp = 0, aps = 0, apn = 0;
            for (int l = 0; l < nl; l++) {
                const double a = c->mx.amat[k * MAXL + l];
                apsp += a * pinsp[l];
                apnp += a * pinnp[l];
                aps += a * pins[l];
                apn += a * pinn[l];
            }
            const double cf = 0.5 * dy * f0 * f0;
            h[C_CSP + k] = -h[C_CSP + k] + cf * apsp;
            h[C_CNP + k] = h[C_CNP + k] + cf * apnp;
            h[C_CS + k] = -h[C_CS + k] + cf * aps;
            h[C_CN + k] = h[C_CN + k] + cf * apn;
        }
    }
    // init_mass_constr (constraint.F90:157-163)
    for (int k = 0; k < nl - 1; k++) {
        const double ipm0 = host_int_dA(c, pm.data() + k * nxy), ipm1 = host_int_dA(c, pm.data() + (k + 1) * nxy);
        const double ip0 = host_int_dA(c, p.data() + k * nxy), ip1 = host_int_dA(c, p.data() + (k + 1) * nxy);
        h[C_DPIP + k] = c->d.dz_sign * (ipm0 - ipm1);
        h[C_DPI + k] = c->d.dz_sign * (ip0 - ip1);
    }
    HIPCHK(c, hipMemcpyAsync(c->cbuf[cur], h, sizeof h, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

int oqg_set_forcing(oqg_ctx* c, const double* ent, int n_ent, const double* wekp, double txis, double txin, int kind) {
    ctx_join(c);
    if (!c) return OQG_ERR_ARG;
    const hipMemcpyKind mk = kind == OQG_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const int nx = c->g.nxp, ny = c->g.nyp;
    if (ent) {
        if (n_ent < 1) return fail(c, OQG_ERR_ARG, "n_ent must be >= 1");
        if (int r = upload_field(c, c->ent1, ent, 1, kind)) return r;
        c->has_ent2 = (n_ent >= 2);
        if (c->has_ent2) {
            const double* e2 = ent + (size_t)nx * ny;
            HIPCHK(c, hipMemcpyAsync(c->ent2s, e2, nx * sizeof(double), mk, c->st));
            HIPCHK(c, hipMemcpyAsync(c->ent2n, e2 + (size_t)(ny - 1) * nx, nx * sizeof(double), mk, c->st));
        }
    }
    if (wekp)
        if (int r = upload_field(c, c->wekp, wekp, 1, kind)) return r;
    double f[F_LEN] = {};
    f[F_TXIS] = txis;
    f[F_TXIN] = txin;
    f[F_ENT2] = c->has_ent2 ? 1.0 : 0.0;
    HIPCHK(c, hipMemcpyAsync(c->fcon, f, sizeof f, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

static int wind_buffers(oqg_ctx* c) {
    if (!c->taux_d)
        if (int r = dalloc(c, &c->taux_d, (size_t)c->g.fs)) return r;
    if (!c->tauy_d)
        if (int r = dalloc(c, &c->tauy_d, (size_t)c->g.fs)) return r;
    return 0;
}

int oqg_set_wind(oqg_ctx* c, const double* taux, const double* tauy, int kind) {
    ctx_join(c);
    if (!c || !taux || !tauy) return c ? fail(c, OQG_ERR_ARG, "oqg_set_wind: null stress") : OQG_ERR_ARG;
    if (int r = wind_buffers(c)) return r;
    if (int r = upload_field(c, c->taux_d, taux, 1, kind)) return r;
    if (int r = upload_field(c, c->tauy_d, tauy, 1, kind)) return r;
    EkmanArgs a{};
    a.g = c->g;
    a.taux = c->taux_d;
    a.tauy = c->tauy_d;
    a.wekp = c->wekp;
    a.fcon = c->fcon;
    launch_ekman(a, c->st);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

int oqg_get_forcing(oqg_ctx* c, double* wekp, double* txi, int kind) {
    ctx_join(c);
    if (!c) return OQG_ERR_ARG;
    if (wekp)
        if (int r = download_field(c, wekp, c->wekp, 1, kind)) return r;
    if (txi) {
        double f[F_LEN];
        HIPCHK(c, hipMemcpyAsync(f, c->fcon, sizeof f, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        txi[0] = f[F_TXIS];
        txi[1] = f[F_TXIN];
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

int oqg_get_wind(oqg_ctx* c, double* taux, double* tauy, int kind) {
    ctx_join(c);
    if (!c) return OQG_ERR_ARG;
    if (!c->taux_d) return fail(c, OQG_ERR_ARG, "oqg_get_wind: the basin has no resident stress");
    if (taux)
        if (int r = download_field(c, taux, c->taux_d, 1, kind)) return r;
    if (tauy)
        if (int r = download_field(c, tauy, c->tauy_d, 1, kind)) return r;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

int oqg_get_ent(oqg_ctx* c, double* ent1, int kind) {
    ctx_join(c);
    if (!c || !ent1) return c ? fail(c, OQG_ERR_ARG, "oqg_get_ent: null field") : OQG_ERR_ARG;
    if (int r = download_field(c, ent1, c->ent1, 1, kind)) return r;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

}  // extern "C"

static void diag_args(oqg_ctx* c, double* partial, DiagArgs& a) {
    a = DiagArgs{};
    a.g = c->g;
    for (int k = 0; k < c->g.nl - 1; k++) a.gp[k] = c->d.gp[k];
    a.dz_sign = c->d.dz_sign;
    a.bcco = c->d.bcco;
    a.q = c->qbuf[c->cur];
    a.p = c->pbuf[c->cur];
    a.partial = partial;
}

// The rows' partials [nyp][DG_LEN] combined in row order into comb[DG_COMB_LEN]
// (the same operations, in the same order, as k_diag_combine).
static void diag_combine_host(const double* h, int ny, int nl, double* comb) {
    const int kinds[6] = {DG_PMIN, DG_PMAX, DG_QMIN, DG_QMAX, DG_UMAX, DG_VMAX};
    double sum[5][MAXL] = {}, mn[6][MAXL] = {}, bad = 0.0;
    for (int t = 0; t < 6; t++)
        for (int k = 0; k < nl; k++) mn[t][k] = h[kinds[t] * MAXL + k];
    for (int j = 0; j < ny; j++) {
        const double* r = h + (size_t)j * DG_LEN;
        for (int k = 0; k < nl; k++) {
            sum[0][k] += r[DG_KE * MAXL + k];
            sum[1][k] += r[DG_P * MAXL + k];
            sum[2][k] += r[DG_Q * MAXL + k];
            sum[3][k] += r[DG_ETA * MAXL + k];
            sum[4][k] += r[DG_ETA2 * MAXL + k];
            for (int t = 0; t < 6; t++) {
                const double v = r[kinds[t] * MAXL + k];
                mn[t][k] = (t == 0 || t == 2) ? std::fmin(mn[t][k], v) : std::fmax(mn[t][k], v);
            }
        }
        bad += r[DG_BAD];
    }
    for (int k = 0; k < MAXL; k++) {
        for (int t = 0; t < 5; t++) comb[DC_SUM + t * MAXL + k] = sum[t][k];
        for (int t = 0; t < 6; t++) comb[DC_EXT + t * MAXL + k] = mn[t][k];
        comb[DC_P0FIRST + k] = h[DG_P0 * MAXL + k];
        comb[DC_P0LAST + k] = h[(size_t)(ny - 1) * DG_LEN + DG_P0 * MAXL + k];
    }
    comb[DC_BAD] = bad;
}

// qg_diagno's state energetics and averages and scan_cfl's maxima from the
// combined rows.
static void diag_finish(const oqg_ctx* c, const double* comb, double dt, double rho, oqg_diag* out) {
    const Geo& g = c->g;
    const int nl = g.nl, ny = g.nyp;
    const double dA = g.dx * g.dy, area = ((g.nxp - 1) * g.dx) * ((ny - 1) * g.dy), f0 = g.fnot;
    std::memset(out, 0, sizeof *out);
    const double* sum0 = comb + DC_SUM;
    auto sum = [&](int t, int k) { return sum0[t * MAXL + k]; };
    auto ext = [&](int t, int k) { return comb[DC_EXT + t * MAXL + k]; };
    const double* p_first = comb + DC_P0FIRST;   // row 0, column 1
    const double* p_last = comb + DC_P0LAST;     // row nyp, column 1
    for (int k = 0; k < nl; k++) {
        out->ekei[k] = 0.5 * rho * c->d.h[k] * (sum(0, k) * dA);
        out->pavg[k] = (sum(1, k) * dA) / area;
        out->qavg[k] = (sum(2, k) * dA) / area;
        out->pmin[k] = ext(0, k);
        out->pmax[k] = ext(1, k);
        out->qmin[k] = ext(2, k);
        out->qmax[k] = ext(3, k);
        out->umax[k] = ext(4, k);
        out->vmax[k] = ext(5, k);
        out->cflu[k] = ext(4, k) * dt / g.dx;
        out->cflv[k] = ext(5, k) * dt / g.dx;
        out->circ[k] = 1.0e-6 * ((p_first[k] - p_last[k]) / f0);
        out->ctot += out->circ[k];
        const double poref = f0 > 0.0 ? p_first[k] : p_last[k];
        const double lo = std::fmin(ext(0, k) / f0, ext(1, k) / f0), hi = std::fmax(ext(0, k) / f0, ext(1, k) / f0);
        out->sfmin[k] = 1.0e-6 * (c->d.h[k] * (lo - poref / f0));
        out->sfmax[k] = 1.0e-6 * (c->d.h[k] * (hi - poref / f0));
        if (k < nl - 1) {
            out->etam[k] = (sum(3, k) * dA) / area;
            out->et2m[k] = (sum(4, k) * dA) / area;
        }
        out->pint[k] = sum(1, k) * dA;
    }
    out->nonfinite = comb[DC_BAD];
}

// qg_diagno's energetics from the rows' partials, combined in row order.
static void energy_finish(const oqg_ctx* c, const double* h, double rho, oqg_energy* out) {
    const Geo& g = c->g;
    const int nl = g.nl, ny = g.nyp;
    const double dA = g.dx * g.dy, xl = (g.nxp - 1) * g.dx, area = xl * ((ny - 1) * g.dy);
    std::memset(out, 0, sizeof *out);
    auto total = [&](int slot) {
        double s = 0.0;
        for (int j = 0; j < ny; j++) s += h[(size_t)j * EN_LEN + slot];
        return s;
    };
    for (int k = 0; k < nl; k++) {
        const double u2diss = total(EN_U2 * MAXL + k) * dA, v2diss = total(EN_V2 * MAXL + k) * dA;
        const double u4diss = total(EN_U4 * MAXL + k) * dA, v4diss = total(EN_V4 * MAXL + k) * dA;
        out->ah2d[k] = -rho * c->d.ah2[k] * c->d.h[k] * (u2diss + v2diss);
        out->ah4d[k] = rho * c->d.ah4[k] * c->d.h[k] * (u4diss + v4diss);
        // maxloc / maxval of |int_P_dx(ugeos(:,j))| / xl over rows 1..nyt
        double best = -1.0;
        int pos = 0;
        for (int j = 0; j < ny - 1; j++) {
            const double ujet = h[(size_t)j * EN_LEN + EN_UJET * MAXL + k] * g.dx;
            const double v = std::fabs(ujet) / xl;
            if (v > best) {
                best = v;
                pos = j + 1;
            }
        }
        out->val[k] = best;
        out->pos[k] = pos;
    }
    out->btdg = 0.5 * rho * c->d.delek * std::fabs(g.fnot) * (total(EN_BTU) * dA + total(EN_BTV) * dA);
    out->utau = rho * (total(EN_VTY) * dA + total(EN_UTX) * dA);
    out->wepm = (total(EN_WEK) * dA) / area;
    out->wapm = (total(EN_AWEK) * dA) / area;
    out->entm = (total(EN_ENT) * dA) / area;
    out->enam = (total(EN_AENT) * dA) / area;
    out->pken = nl > 1 ? rho * c->d.gp[0] * ((total(EN_ETAENT) * dA) / area) : 0.0;
}

extern "C" {

int oqg_diagnose(oqg_ctx* c, double dt, double rho, oqg_diag* out) {
    ctx_join(c);
    if (!c || !out) return OQG_ERR_ARG;
    const int nl = c->g.nl, ny = c->g.nyp;
    if (!c->diag_part)
        if (int r = dalloc(c, &c->diag_part, (size_t)ny * DG_LEN)) return r;
    DiagArgs a;
    diag_args(c, c->diag_part, a);
    launch_diag(a, c->st);
    HIPCHK(c, hipGetLastError());
    std::vector<double> h((size_t)ny * DG_LEN);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->diag_part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    double comb[DG_COMB_LEN];
    diag_combine_host(h.data(), ny, nl, comb);   // the rows in order (deterministic)
    diag_finish(c, comb, dt, rho, out);
    return 0;
}

int oqg_diagnose_energy(oqg_ctx* c, double rho, oqg_energy* out) {
    ctx_join(c);
    if (!c || !out) return OQG_ERR_ARG;
    const Geo& g = c->g;
    if (!c->en_d2)
        if (int r = dalloc(c, &c->en_d2, 2 * (size_t)g.nl * g.fs)) return r;
    if (!c->en_part)
        if (int r = dalloc(c, &c->en_part, (size_t)g.nyp * EN_LEN)) return r;
    EnergyArgs a{};
    a.g = g;
    a.bcco = c->d.bcco;
    a.gp0 = g.nl > 1 ? c->d.gp[0] : 1.0;
    a.dz_sign = c->d.dz_sign;
    a.p = c->pbuf[c->cur];
    a.pm = c->pbuf[1 - c->cur];
    a.wekp = c->wekp;
    a.ent1 = c->ent1;
    a.taux = c->taux_d;
    a.tauy = c->taux_d ? c->tauy_d : nullptr;
    a.d2u = c->en_d2;
    a.d2v = c->en_d2 + (size_t)g.nl * g.fs;
    a.partial = c->en_part;
    launch_energy(a, c->st);
    HIPCHK(c, hipGetLastError());
    std::vector<double> h((size_t)g.nyp * EN_LEN);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->en_part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    energy_finish(c, h.data(), rho, out);
    return 0;
}

int oqg_diagnose_noalloc(oqg_ctx* c, double dt, double rho, oqg_diag* out) {
    ctx_join(c);
    if (!c || !out) return OQG_ERR_ARG;
    const int nl = c->g.nl, ny = c->g.nyp;
    const long long need = (long long)ny * DG_LEN + DG_COMB_LEN;
    if (need > NOALLOC_SCRATCH_DOUBLES)
        return fail(c, OQG_ERR_ARG, "oqg_diagnose_noalloc: nyp = " + std::to_string(ny) + " needs " +
                                        std::to_string(need) + " doubles of scratch, more than its " +
                                        std::to_string(NOALLOC_SCRATCH_DOUBLES));
    double* scr = c->scratch;
    if (!scr) {
        HIPCHK(c, noalloc_scratch(c->dev, &scr));
        c->scratch = scr;
    }
    double* comb_d = scr + (long long)ny * DG_LEN;
    DiagArgs a;
    diag_args(c, scr, a);
    launch_diag(a, c->st);   // writes every slot that the combination reads
    launch_diag_combine(scr, ny, nl, comb_d, c->st);
    HIPCHK(c, hipGetLastError());
    double comb[DG_COMB_LEN];
    HIPCHK(c, hipMemcpyAsync(comb, comb_d, sizeof comb, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    diag_finish(c, comb, dt, rho, out);
    return 0;
}

int oqg_adj_wind(oqg_ctx* c, double* ltaux, double* ltauy, int kind) {
    ctx_join(c);
    if (!c || !ltaux || !ltauy) return c ? fail(c, OQG_ERR_ARG, "oqg_adj_wind: null output") : OQG_ERR_ARG;
    if (!c->adj) return fail(c, OQG_ERR_STATE, "call oqg_adj_init first");
    if (int r = wind_buffers(c)) return r;
    EkmanArgs a{};
    a.g = c->g;
    a.taux = c->taux_d;
    a.tauy = c->tauy_d;
    a.lwek = c->lwek;
    a.ltxi = c->ltxi;
    launch_ekman_adj(a, c->st);
    HIPCHK(c, hipGetLastError());
    if (int r = download_field(c, ltaux, c->taux_d, 1, kind)) return r;
    if (int r = download_field(c, ltauy, c->tauy_d, 1, kind)) return r;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

int oqg_step(oqg_ctx* c, double tdt, const double* ent, int n_ent, const double* wekp, double txis, double txin) {
    ctx_join(c);
    if (!c) return OQG_ERR_ARG;
    if ((ent == nullptr) != (wekp == nullptr))
        return fail(c, OQG_ERR_ARG, "pass both ent and wekp, or neither to use the resident forcing");
    if (ent)
        if (int r = oqg_set_forcing(c, ent, n_ent, wekp, txis, txin, OQG_HOST)) return r;
    return enqueue_step(c, tdt);
}

int oqg_suppress(oqg_ctx* c) {
    ctx_join(c);
    if (!c) return OQG_ERR_ARG;
    return enqueue_suppress(c);
}

}  // extern "C"

// Steps per replayed graph: a whole number of suppression periods, even so
// that the time-level parity is the same at both ends.
static int graph_period(int supp) { return supp > 0 ? (supp % 2 == 0 ? supp : 2 * supp) : 2; }

// The graph cached under `key`, captured from what `enqueue` puts on the
// stream the first time.  Capture leaves the host's parity as it found it.
template <class F>
static int cached_graph(oqg_ctx* c, const GraphKey& key, F&& enqueue, hipGraphExec_t* out) {
    auto it = c->graphs.find(key);
    if (it == c->graphs.end()) {
        const int cur0 = c->cur;
        hipGraph_t gr;
        HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue();
        hipError_t ce = hipStreamEndCapture(c->st, &gr);
        c->cur = cur0;
        if (rc) {
            if (ce == hipSuccess) hipGraphDestroy(gr);
            return rc;
        }
        if (ce != hipSuccess) return fail(c, OQG_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
        hipGraphExec_t ex;
        hipError_t ie = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
        hipGraphDestroy(gr);
        if (ie != hipSuccess) return fail(c, OQG_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
        it = c->graphs.emplace(key, ex).first;
    }
    *out = it->second;
    return 0;
}

extern "C" {

}  // extern "C"

// Steps nt0 .. nt0+n-1 of a run, from the current time level, eagerly.
static int run_eager(oqg_ctx* c, double tdt, int nt0, int n, int supp) {
    for (int nt = nt0; nt < nt0 + n; nt++) {
        if (int r = enqueue_step(c, tdt)) return r;
        if (supp > 0 && (nt - 1) % supp == 0)
            if (int r = enqueue_suppress(c)) return r;
    }
    return 0;
}

// The stretches oqg_run splits steps step0+1 .. step0+nsteps into: a head up
// to the first step nt with (nt-1) % P == 0, whole periods, and a tail.
struct RunSplit {
    int head, periods, tail;
};
static RunSplit run_split(int nsteps, int step0, int P) {
    RunSplit r{};
    const int first = step0 + 1;
    int head = (P - (first - 1) % P) % P;
    if (head > nsteps) head = nsteps;
    r.head = head;
    r.periods = (nsteps - head) / P;
    r.tail = nsteps - head - r.periods * P;
    return r;
}

// A head or tail stretch: its planned graph when oqg_plan_run captured one,
// otherwise eager launches.  A one-step graph measured slower than its four
// launches, and graphs of 2^k suppression-free steps for the stretches of a
// checkpointed sweep measured no faster than eager launches (the sweep's
// steps are bound by the kernels' own time, not by their launches).  Leaves
// c->cur advanced by n.
static int run_stretch(oqg_ctx* c, double tdt, int nt0, int n, int supp, int P, bool capture) {
    if (n <= 0) return 0;
    const GraphKey key{tdt, n, supp, c->cur, G_SEG, (nt0 - 1) % P};
    if (capture || c->graphs.count(key)) {
        hipGraphExec_t ex;
        if (int r = cached_graph(c, key, [&]() { return run_eager(c, tdt, nt0, n, supp); }, &ex)) return r;
        if (!capture) HIPCHK(c, hipGraphLaunch(ex, c->st));
        if (n & 1) c->cur = 1 - c->cur;
        return 0;
    }
    return run_eager(c, tdt, nt0, n, supp);
}

static int run_steps(oqg_ctx* c, double tdt, int nsteps, int step0, int supp, bool plan) {
    if (!c || nsteps < 0) return OQG_ERR_ARG;
    const int P = graph_period(supp);
    const int cur0 = c->cur;
    const RunSplit sp = run_split(nsteps, step0, P);
    int nt = step0 + 1;
    if (int r = run_stretch(c, tdt, nt, sp.head, supp, P, plan)) return r;
    nt += sp.head;
    for (int b = 0; b < sp.periods; b++, nt += P) {
        hipGraphExec_t ex;
        const int rc = cached_graph(c, GraphKey{tdt, P, supp, c->cur, G_RUN}, [&]() {
            for (int t = 0; t < P; t++) {
                if (int r = enqueue_step(c, tdt)) return r;
                if (supp > 0 && (t % supp) == 0)
                    if (int r = enqueue_suppress(c)) return r;
            }
            return 0;
        }, &ex);
        if (rc) return rc;
        if (plan) break;   // one period's graph serves every period
        HIPCHK(c, hipGraphLaunch(ex, c->st));
        // P is even, so the time-level parity is unchanged
    }
    if (plan) nt = step0 + 1 + sp.head + sp.periods * P;
    if (int r = run_stretch(c, tdt, nt, sp.tail, supp, P, plan)) return r;
    if (plan) c->cur = cur0;   // planning runs nothing
    return 0;
}

extern "C" {

int oqg_run(oqg_ctx* c, double tdt, int nsteps, int step0, int supp) {
    ctx_join(c);
    return run_steps(c, tdt, nsteps, step0, supp, false);
}

}  // extern "C"

// The vorticity kernel's tile height (4, 8 or 16 rows; 0: the occupancy rule
// of vort_rows).  Captured graphs hold the launch shape, so they go.  Only the
// order in which the per-tile entrainment integrals are summed depends on it.
static int set_tile_rows(oqg_ctx* c, int rows) {
    if (rows == 0) rows = vort_rows(c->tiles_x, c->g.nyp, c->g.nl);
    if (rows != 4 && rows != 8 && rows != 16) return fail(c, OQG_ERR_ARG, "tile rows: 4, 8 or 16 (0: the default rule)");
    if (rows == c->vort_ty) return 0;
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (auto& kv : c->graphs) hipGraphExecDestroy(kv.second);
    c->graphs.clear();
    c->vort_ty = rows;
    c->tiles_y = (c->g.nyp - 2 + rows - 1) / rows;
    return 0;
}

// Device copies of a context's state (both time levels of q, p and the
// constraint blocks, and the parity), for trial runs.
struct StateKeep {
    std::vector<void*> bufs;
    int cur = 0;
    int save(oqg_ctx* c) {
        cur = c->cur;
        for (int t = 0; t < 2; t++) {
            const std::pair<const void*, size_t> parts[3] = {
                {c->qbuf[t], c->fbytes}, {c->pbuf[t], c->fbytes}, {c->cbuf[t], C_LEN * sizeof(double)}};
            for (const auto& pr : parts) {
                void* d = nullptr;
                if (hipMalloc(&d, pr.second) != hipSuccess) return OQG_ERR_HIP;
                bufs.push_back(d);
                if (hipMemcpyAsync(d, pr.first, pr.second, hipMemcpyDeviceToDevice, c->st) != hipSuccess) return OQG_ERR_HIP;
            }
        }
        return 0;
    }
    void restore(oqg_ctx* c, hipStream_t st) {
        for (int t = 0; t < 2; t++) {
            hipMemcpyAsync(c->qbuf[t], bufs[t * 3 + 0], c->fbytes, hipMemcpyDeviceToDevice, st);
            hipMemcpyAsync(c->pbuf[t], bufs[t * 3 + 1], c->fbytes, hipMemcpyDeviceToDevice, st);
            hipMemcpyAsync(c->cbuf[t], bufs[t * 3 + 2], C_LEN * sizeof(double), hipMemcpyDeviceToDevice, st);
        }
        c->cur = cur;
    }
    void release(hipStream_t st) {
        hipStreamSynchronize(st);
        for (void* b : bufs) hipFree(b);
        bufs.clear();
    }
};

extern "C" {

int oqg_set_tile_rows(oqg_ctx* c, int rows) {
    ctx_join(c);
    if (!c) return OQG_ERR_ARG;
    return set_tile_rows(c, rows);
}

int oqg_get_tile_rows(const oqg_ctx* c) { return c ? c->vort_ty : -OQG_ERR_ARG; }

int oqg_tune(oqg_ctx* c, double tdt, int trial_steps, int supp_every, double* ms) {
    ctx_join(c);
    if (!c || trial_steps < 1) return -OQG_ERR_ARG;
    if (c->ext_cur.q || c->ext_old.q) return -fail(c, OQG_ERR_STATE, "oqg_tune: a checkpoint is loaded by reference");
    StateKeep keep;
    if (keep.save(c)) return (keep.release(c->st), -fail(c, OQG_ERR_HIP, "oqg_tune: could not keep the state"));
    const int rows0 = c->vort_ty;
    hipEvent_t ev[2];
    hipEventCreate(&ev[0]);
    hipEventCreate(&ev[1]);
    int best = rows0, rc = 0;
    double best_ms = 1e300;
    const int opts[3] = {4, 8, 16};
    for (int t = 0; t < 3 && !rc; t++) {
        rc = set_tile_rows(c, opts[t]);
        if (!rc) rc = run_steps(c, tdt, trial_steps, 0, supp_every, false);   // captures the graphs
        keep.restore(c, c->st);
        if (rc) break;
        hipEventRecord(ev[0], c->st);
        rc = run_steps(c, tdt, trial_steps, 0, supp_every, false);
        hipEventRecord(ev[1], c->st);
        hipEventSynchronize(ev[1]);
        keep.restore(c, c->st);
        float el = 0.f;
        hipEventElapsedTime(&el, ev[0], ev[1]);
        if (ms) ms[t] = el;
        if (el < best_ms) {
            best_ms = el;
            best = opts[t];
        }
    }
    hipEventDestroy(ev[0]);
    hipEventDestroy(ev[1]);
    keep.release(c->st);
    if (rc) {
        set_tile_rows(c, rows0);
        return -rc;
    }
    if (int r = set_tile_rows(c, best)) return -r;
    return best;
}

int oqg_plan_run(oqg_ctx* c, double tdt, int nsteps, int step0, int supp) {
    ctx_join(c);
    return run_steps(c, tdt, nsteps, step0, supp, true);
}

int oqg_sync(oqg_ctx* c) {
    ctx_join(c);
    if (!c) return OQG_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

int oqg_solve_inhomog(oqg_ctx* c, int m, const double* rhs, double* out, int kind) {
    ctx_join(c);
    if (!c || !rhs || !out) return OQG_ERR_ARG;
    if (m < 1 || m > c->g.nl) return fail(c, OQG_ERR_ARG, "mode index out of range");
    if (int r = upload_field(c, c->raw[0], rhs, 1, kind)) return r;
    if (int r = enqueue_raw_solve(c, m - 1)) return r;
    if (int r = download_field(c, out, c->raw[1], 1, kind)) return r;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

int oqg_get_modes(oqg_ctx* c, double* amat, double* ctl2m, double* ctm2l, double* rdm2) {
    if (!c) return OQG_ERR_ARG;
    const int nl = c->g.nl;
    for (int i = 0; i < nl; i++)
        for (int j = 0; j < nl; j++) {
            // Fortran column-major (i,j) at [j*nl + i]
            if (amat) amat[j * nl + i] = c->mx.amat[i * MAXL + j];
            if (ctl2m) ctl2m[j * nl + i] = c->mx.ctl2m[i * MAXL + j];   // ctl2m(k=i, m=j)
            if (ctm2l) ctm2l[j * nl + i] = c->mx.ctm2l[i * MAXL + j];   // ctm2l(m=i, k=j)
        }
    if (rdm2)
        for (int m = 0; m < nl; m++) rdm2[m] = c->mx.rdm2[m];
    return 0;
}

int oqg_set_stream(oqg_ctx* c, void* stream) {
    ctx_join(c);
    if (!c) return OQG_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->st = stream ? static_cast<hipStream_t>(stream) : c->own;
    return 0;
}

void* oqg_get_stream(oqg_ctx* c) {
    ctx_join(c);
    return c ? static_cast<void*>(c->st) : nullptr;
}

int oqg_state_ptrs(oqg_ctx* c, double** q, double** qm, double** p, double** pm, size_t* pitch) {
    ctx_join(c);
    if (!c) return OQG_ERR_ARG;
    if (q) *q = c->qbuf[c->cur];
    if (qm) *qm = c->qbuf[1 - c->cur];
    if (p) *p = c->pbuf[c->cur];
    if (pm) *pm = c->pbuf[1 - c->cur];
    if (pitch) *pitch = (size_t)c->g.pitch;
    return 0;
}

// Times each kernel of one step over `reps` back-to-back launches between two
// events, then launches it once more on the real buffers so the next kernel
// sees the inputs of a normal step.  The in-place kernels (k_vort writes q_new
// over qm, the y solve overwrites the spectra) run each timed launch on its own
// copy of that operand, so every timed launch does the work of a real step.
static int time_kernels(oqg_ctx* c, double tdt, int reps, hipEvent_t* ev, double* ms,
                        std::vector<void*>& tmp) {
    const size_t sbytes = (size_t)c->g.nl * c->s.nrow * c->s.kp * sizeof(double);
    for (int r = 0; r < 2 * reps; r++) {
        void* p = nullptr;
        HIPCHK(c, hipMalloc(&p, r < reps ? c->fbytes : sbytes));
        tmp.push_back(p);
    }
    StepArgs a;
    make_step_args(c, tdt, a);
    for (int k = 0; k < NK; k++) {
        ms[k] = 0.0;
        for (int r = 0; r < reps; r++) {
            if (k == KV) HIPCHK(c, hipMemcpyAsync(tmp[r], a.va.qnew, c->fbytes, hipMemcpyDeviceToDevice, c->st));
            if (k == KY) HIPCHK(c, hipMemcpyAsync(tmp[reps + r], a.ya.spec, sbytes, hipMemcpyDeviceToDevice, c->st));
        }
        // the burst is replayed from a graph, so that no host submission gap
        // falls inside the timed region (as in oqg_run)
        hipGraph_t gr;
        HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
        for (int r = 0; r < reps; r++) {
            StepArgs b = a;
            b.va.qnew = static_cast<double*>(tmp[r]);
            b.va.qm = b.va.qnew;
            b.ya.spec = static_cast<double*>(tmp[reps + r]);
            launch_kernel(c, (k == KV || k == KY) ? b : a, k);
        }
        hipError_t ce = hipStreamEndCapture(c->st, &gr);
        if (ce != hipSuccess) return fail(c, OQG_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
        hipGraphExec_t ex;
        hipError_t ie = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
        hipGraphDestroy(gr);
        if (ie != hipSuccess) return fail(c, OQG_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
        hipError_t ue = hipGraphUpload(ex, c->st);
        if (ue == hipSuccess) ue = hipEventRecord(ev[0], c->st);
        if (ue == hipSuccess) ue = hipGraphLaunch(ex, c->st);
        if (ue == hipSuccess) ue = hipEventRecord(ev[1], c->st);
        if (ue == hipSuccess) ue = hipStreamSynchronize(c->st);
        hipGraphExecDestroy(ex);
        HIPCHK(c, ue);
        launch_kernel(c, a, k);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventSynchronize(ev[1]));
        float t = 0.f;
        HIPCHK(c, hipEventElapsedTime(&t, ev[0], ev[1]));
        ms[k] = (double)t / reps;
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->cur = 1 - c->cur;
    return 0;
}

int oqg_kernel_times(oqg_ctx* c, double tdt, int reps, double* ms, int max, char* names, int names_len) {
    ctx_join(c);
    if (!c || reps < 1 || reps > 256 || !ms) return -OQG_ERR_ARG;
    hipEvent_t ev[2];
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) return -fail(c, OQG_ERR_HIP, "hipEventCreate failed");
    double acc[NK] = {};
    std::vector<void*> tmp;
    const int rc = time_kernels(c, tdt, reps, ev, acc, tmp);
    hipStreamSynchronize(c->st);
    for (void* p : tmp) hipFree(p);
    for (auto& e : ev) hipEventDestroy(e);
    if (rc) return -rc;
    const int n = std::min(max, (int)NK);
    std::string nm;
    for (int k = 0; k < n; k++) {
        ms[k] = acc[k];
        nm += kNames[k];
        nm += '\n';
    }
    if (names && names_len > 0) {
        std::strncpy(names, nm.c_str(), names_len - 1);
        names[names_len - 1] = '\0';
    }
    return n;
}

}  // extern "C"

// ------------------------------------------------------------------ adjoint (box)
//
// Reverse mode of step_qg for box basins.  The adjoint buffers mirror the
// forward state's time-level parity: with the forward state S_n loaded (q_n at
// qbuf[cur]), lq[cur]/lq[1-cur] hold the adjoints of q_n/q_{n-1}, and lp
// likewise.  One adjoint step takes the adjoints of S_{n+1} (current parity
// 1 - cur) to those of S_n, reading only q_n and p_n of the forward state.

static int adj_require(oqg_ctx* c) {
    if (!c->adj) return fail(c, OQG_ERR_STATE, "call oqg_adj_init first");
    return 0;
}

// p_new = Inv(q) with the current parity's q as input (interior), the
// inversion of one step: pnew over p, q's boundary ring by ocqbdy, both time
// levels set equal.  tdt = 0 leaves the mass constraint's dpi as it is.
static int enqueue_invert(oqg_ctx* c) {
    const int cur = c->cur;
    StepArgs a;
    make_step_args(c, 0.0, a);
    a.fa.q = c->qbuf[cur];
    a.ia.cin = c->cbuf[cur];
    a.ia.cout = c->cbuf[1 - cur];
    a.ia.pnew = c->pbuf[cur];
    a.ia.qnew = c->qbuf[cur];
    a.ia.tdt = 0.0;
    launch_fwd(a.fa, c->g.nl, c->st);
    launch_ysolve(a.ya, c->variant, c->st);
    launch_inv(a.ia, c->st);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->qbuf[1 - cur], c->qbuf[cur], c->fbytes, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->pbuf[1 - cur], c->pbuf[cur], c->fbytes, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->cbuf[cur], c->cbuf[1 - cur], C_LEN * sizeof(double), hipMemcpyDeviceToDevice, c->st));
    return 0;
}

// lq (interior) += M^T lp, where p = M(q) is the inversion (mix, Helmholtz
// solves, homogeneous correction, mix), with lp already holding the adjoint of
// p including what the boundary PV passed on.
// The boundary ring of q_new is written from p_new (ocqbdy), so its adjoint
// in lq's ring is passed on to lp first (inside k_fwd_adj; lp itself is left
// as it is).  Three launches: the transform kernels in their adjoint modes
// around the forward y solve.
static int enqueue_adj_inversion(oqg_ctx* c, double* lq, const double* lp, double* lc) {
    FwdArgs fa{};
    fa.g = c->g;
    fa.s = c->s;
    fa.dc = c->dconst;
    fa.yprel = c->yprel;
    fa.tw = c->tw;
    fa.spec = c->spec;
    fa.raw_mode = -1;
    fa.adj_lp = lp;
    fa.adj_lq = lq;
    fa.adj_hom = c->hom;
    fa.adj_partial = c->adjPart;
    fa.adj_qbfac = c->ph.qbfac;
    launch_fwd_adj(fa, c->st);
    YArgs ya{};
    ya.s = c->s;
    ya.spec = c->spec;
    ya.bcol = c->bcol;
    ya.gam = c->gamw;
    ya.mode0 = 0;
    ya.nmodes = c->g.nl;
    launch_ysolve(ya, c->variant, c->st);
    InvArgs ia;
    fill_inv_common(c, ia);
    ia.adj_G = c->adjG;
    ia.adj_cinvT = c->adjCinvT;
    ia.adj_partial = c->adjPart;
    ia.adj_npart = (c->s.nyi + 1) / 2;
    ia.adj_lq = lq;
    ia.adj_lc = lc;
    launch_inv_adj(ia, c->st);
    HIPCHK(c, hipGetLastError());
    return 0;
}

// Where an adjoint step reads q_n and p_n of the forward state before the
// step: the loaded state's current time level (all null), given pointers, or
// a snapshot slot found on the device (slots[*base + t], at qoff and poff).
struct SnapRef {
    const double* q = nullptr;
    const double* p = nullptr;
    const double* const* slots = nullptr;
    const int* base = nullptr;
    int t = 0;
    long long qoff = 0, poff = 0;
};

static int enqueue_adj_step(oqg_ctx* c, double tdt, bool suppressed, const SnapRef& src = SnapRef{}) {
    const int cur = c->cur, cn = 1 - cur;
    if (suppressed) {
        AdjSuppArgs s{};
        s.g = c->g;
        s.lq = c->lq[cn];
        s.lqm = c->lq[cur];
        s.lp = c->lp[cn];
        s.lpm = c->lp[cur];
        if (c->g.cyclic) {
            s.lc = c->lc[cn];
            s.lcm = c->lc[cur];
        }
        launch_adj_supp(s, c->st);
    }
    if (int r = enqueue_adj_inversion(c, c->lq[cn], c->lp[cn], c->lc[cn])) return r;
    AdjVortArgs v{};
    v.g = c->g;
    v.ph = &c->dconst->ph;
    v.q = src.q ? src.q : c->qbuf[cur];
    v.p = src.p ? src.p : c->pbuf[cur];
    v.slots = src.slots;
    v.base = src.base;
    v.t = src.t;
    v.qoff = src.qoff;
    v.poff = src.poff;
    v.lq_np1 = c->lq[cn];
    v.lq_n = c->lq[cur];
    v.lp_n = c->lp[cur];
    v.lp_nm1 = c->lp[cn];
    v.lwek = c->lwek;
    v.tdt = tdt;
    if (c->g.cyclic) {
        v.lc = c->lc[cn];
        v.ltxi = c->ltxi;
        for (int d = 0; d < 4; d++) {
            v.vs3[d] = c->adj_rows[0][d];
            v.vs5[d] = c->adj_rows[1][d];
            v.vsb[d] = c->adj_rows[2][d];
            v.vn3[d] = c->adj_rows[3][d];
            v.vn5[d] = c->adj_rows[4][d];
            v.vnb[d] = c->adj_rows[5][d];
        }
    }
    launch_adj_vort(v, c->st);
    HIPCHK(c, hipGetLastError());
    return 0;
}

// Channel precompute of the adjoint: Gs_m, Gn_m = H_m^-1 of the unit rows 1
// and nyp-2 (row-constant fields at adjG[side * nl + m]), and the gradients
// of the boundary integrals with respect to the row sums P_j of pm
// (vorticity.F90:57-62,95-101): with D the y part of del2_P_bc acting on a
// column (numerics.F90:455-468) and es = e_1 - e_0, en = e_N - e_{N-1},
//   ap3s = ah2 (dx/dy) es.D P, ap5s = ah4 (dx/dy) es.D D P, bdrins = cd (dx/dy) es.P
// (north likewise with en), so the gradients are D^T es, D^T D^T es and es
// (and with en), of which rows 0..3 (south) and N..N-3 (north) are nonzero.
static int adj_init_channel(oqg_ctx* c) {
    const Geo& g = c->g;
    const int ny = g.nyp, N = ny - 1;
    const size_t f1 = (size_t)g.fs;
    std::vector<double> unit(f1, 0.0);
    for (int side = 0; side < 2; side++) {
        std::fill(unit.begin(), unit.end(), 0.0);
        const int row = side == 0 ? 1 : N - 1;
        for (int i = 0; i < g.nxp; i++) unit[(size_t)row * g.pitch + i] = 1.0;
        HIPCHK(c, hipMemcpyAsync(c->adjW, unit.data(), f1 * sizeof(double), hipMemcpyHostToDevice, c->st));
        for (int m = 0; m < g.nl; m++)
            if (int r = enqueue_raw_solve(c, m, c->adjW, c->adjG + (size_t)(side * g.nl + m) * f1)) return r;
        HIPCHK(c, hipStreamSynchronize(c->st));
    }
    // (D^T v)(a) = sum_j D[j][a] v[j]
    auto DT = [&](const std::vector<double>& v) {
        std::vector<double> r(ny, 0.0);
        auto D = [&](int j, int a) -> double {
            if (j == 0) return a == 0 ? -g.cyl : (a == 1 ? g.cyl : 0.0);
            if (j == N) return a == N ? g.cyr : (a == N - 1 ? -g.cyr : 0.0);
            return a == j ? -2.0 * g.fy : ((a == j - 1 || a == j + 1) ? g.fy : 0.0);
        };
        for (int a = 0; a < ny; a++) {
            double t = 0.0;
            for (int j = std::max(0, a - 1); j <= std::min(N, a + 1); j++) t += D(j, a) * v[j];
            r[a] = t;
        }
        return r;
    };
    std::vector<double> es(ny, 0.0), en(ny, 0.0);
    es[0] = -1.0;
    es[1] = 1.0;
    en[N] = 1.0;
    en[N - 1] = -1.0;
    const std::vector<double> s3 = DT(es), s5 = DT(s3), n3 = DT(en), n5 = DT(n3);
    for (int d = 0; d < 4; d++) {
        c->adj_rows[0][d] = s3[d];
        c->adj_rows[1][d] = s5[d];
        c->adj_rows[2][d] = es[d];
        c->adj_rows[3][d] = n3[N - d];
        c->adj_rows[4][d] = n5[N - d];
        c->adj_rows[5][d] = en[N - d];
    }
    c->adj = true;
    return 0;
}

extern "C" {

int oqg_adj_init(oqg_ctx* c) {
    ctx_join(c);
    if (!c) return OQG_ERR_ARG;
    if (c->s.prime)
        return fail(c, OQG_ERR_ARG, "the adjoint takes x transform lengths that factor into 2, 3 and 5");
    if (c->g.cyclic && (!fftw_fits(c->s, c->g.nl) || c->g.nyp < 9))
        return fail(c, OQG_ERR_ARG, "the channel adjoint takes nyp >= 9 and transforms that fit the wavefront kernels");
    if (c->adj) return 0;
    const size_t fl = (size_t)c->g.nl * c->g.fs, f1 = (size_t)c->g.fs;
    const size_t npairs = (size_t)((c->s.nyi + 1) / 2);
    int rc = 0;
    for (int t = 0; t < 2 && !rc; t++) {
        rc = dalloc(c, &c->lq[t], fl);
        if (!rc) rc = dalloc(c, &c->lp[t], fl);
    }
    if (!rc) rc = dalloc(c, &c->lwek, f1);
    if (!rc) rc = dalloc(c, &c->adjW, f1);
    if (!rc) rc = dalloc(c, &c->adjG, c->g.cyclic ? 2 * fl : fl);
    // [row pair][mode] (box), [row pair][profile] (channel)
    if (!rc) rc = dalloc(c, &c->adjPart, npairs * (c->g.cyclic ? ADJ_CYC_NPROF : MAXL));
    if (!rc) rc = dalloc(c, &c->adjCinvT, MAXL * MAXL);
    double* cblk = nullptr;
    if (!rc) rc = dalloc(c, &cblk, 3 * LC_LEN + 8);
    if (rc) return rc;
    c->lc[0] = cblk;
    c->lc[1] = cblk + LC_LEN;
    c->lc_scratch = cblk + 2 * LC_LEN;
    c->ltxi = cblk + 3 * LC_LEN;
    if (c->g.cyclic) return adj_init_channel(c);
    // (cdhoc^-1)^T by Gauss-Jordan in long double
    const int n = c->g.nl - 1;
    long double A[MAXL][2 * MAXL] = {};
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) A[i][j] = c->bh.cdhoc[i * MAXL + j];
        A[i][n + i] = 1.0L;
    }
    for (int col = 0; col < n; col++) {
        int piv = col;
        for (int r = col + 1; r < n; r++)
            if (fabsl(A[r][col]) > fabsl(A[piv][col])) piv = r;
        if (A[piv][col] == 0.0L) return fail(c, OQG_ERR_NUMERIC, "singular mass-constraint matrix");
        if (piv != col)
            for (int j = 0; j < 2 * n; j++) std::swap(A[piv][j], A[col][j]);
        const long double d = A[col][col];
        for (int j = 0; j < 2 * n; j++) A[col][j] /= d;
        for (int r = 0; r < n; r++)
            if (r != col) {
                const long double f = A[r][col];
                for (int j = 0; j < 2 * n; j++) A[r][j] -= f * A[col][j];
            }
    }
    std::vector<double> cinvT(MAXL * MAXL, 0.0);
    for (int l = 0; l < n; l++)
        for (int j = 0; j < n; j++) cinvT[l * MAXL + j] = (double)A[j][n + l];   // (C^-1)^T[l][j] = C^-1[j][l]
    HIPCHK(c, hipMemcpyAsync(c->adjCinvT, cinvT.data(), cinvT.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
    // G_m = H_m^-1 (1 on the interior)
    std::vector<double> one(f1, 0.0);
    for (int j = 1; j < c->g.nyp - 1; j++)
        for (int i = 1; i < c->g.nxp - 1; i++) one[(size_t)j * c->g.pitch + i] = 1.0;
    HIPCHK(c, hipMemcpyAsync(c->adjW, one.data(), f1 * sizeof(double), hipMemcpyHostToDevice, c->st));
    for (int m = 0; m < c->g.nl; m++)
        if (int r = enqueue_raw_solve(c, m, c->adjW, c->adjG + (size_t)m * f1)) return r;
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->adj = true;
    return 0;
}

int oqg_adj_set(oqg_ctx* c, const double* lq, const double* lqm, const double* lp, const double* lpm, int kind) {
    ctx_join(c);
    if (!c) return OQG_ERR_ARG;
    if (int r = adj_require(c)) return r;
    double* dst[4] = {c->lq[c->cur], c->lq[1 - c->cur], c->lp[c->cur], c->lp[1 - c->cur]};
    const double* src[4] = {lq, lqm, lp, lpm};
    for (int t = 0; t < 4; t++) {
        if (src[t]) {
            if (int r = upload_field(c, dst[t], src[t], c->g.nl, kind)) return r;
            if (c->g.cyclic) launch_fold_dup(dst[t], c->g, c->g.nl, c->st);
        } else {
            HIPCHK(c, hipMemsetAsync(dst[t], 0, c->fbytes, c->st));
        }
    }
    HIPCHK(c, hipMemsetAsync(c->lc[0], 0, 2 * LC_LEN * sizeof(double), c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

// The adjoint constraint scalars in oqg_get_constr's packing; the dpi and dpip
// entries are not state of the adjoint (the step's p does not depend on them
// in a channel, and a box's dpi depends on the entrainment alone) and read 0.
int oqg_adj_set_constr(oqg_ctx* c, const double* v) {
    ctx_join(c);
    if (!c || !v) return OQG_ERR_ARG;
    if (int r = adj_require(c)) return r;
    const int nl = c->g.nl, o = 2 * (nl - 1);
    double h[2][LC_LEN] = {};
    for (int k = 0; k < nl; k++) {
        h[0][k] = v[o + k];
        h[0][MAXL + k] = v[o + nl + k];
        h[1][k] = v[o + 2 * nl + k];
        h[1][MAXL + k] = v[o + 3 * nl + k];
    }
    HIPCHK(c, hipMemcpyAsync(c->lc[c->cur], h[0], sizeof h[0], hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->lc[1 - c->cur], h[1], sizeof h[1], hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

int oqg_adj_get_constr(oqg_ctx* c, double* v) {
    ctx_join(c);
    if (!c || !v) return OQG_ERR_ARG;
    if (int r = adj_require(c)) return r;
    const int nl = c->g.nl, o = 2 * (nl - 1);
    double h[2][LC_LEN];
    HIPCHK(c, hipMemcpyAsync(h[0], c->lc[c->cur], sizeof h[0], hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(h[1], c->lc[1 - c->cur], sizeof h[1], hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (int t = 0; t < o; t++) v[t] = 0.0;
    for (int k = 0; k < nl; k++) {
        v[o + k] = h[0][k];
        v[o + nl + k] = h[0][MAXL + k];
        v[o + 2 * nl + k] = h[1][k];
        v[o + 3 * nl + k] = h[1][MAXL + k];
    }
    return 0;
}

int oqg_adj_get_txi(oqg_ctx* c, double* out) {
    ctx_join(c);
    if (!c || !out) return OQG_ERR_ARG;
    if (int r = adj_require(c)) return r;
    HIPCHK(c, hipMemcpyAsync(out, c->ltxi, 2 * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

int oqg_adj_get(oqg_ctx* c, double* lq, double* lqm, double* lp, double* lpm, int kind) {
    ctx_join(c);
    if (!c) return OQG_ERR_ARG;
    if (int r = adj_require(c)) return r;
    const double* src[4] = {c->lq[c->cur], c->lq[1 - c->cur], c->lp[c->cur], c->lp[1 - c->cur]};
    double* dst[4] = {lq, lqm, lp, lpm};
    for (int t = 0; t < 4; t++)
        if (dst[t])
            if (int r = download_field(c, dst[t], src[t], c->g.nl, kind)) return r;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

int oqg_adj_step(oqg_ctx* c, double tdt, int suppressed) {
    ctx_join(c);
    if (!c) return OQG_ERR_ARG;
    if (int r = adj_require(c)) return r;
    return enqueue_adj_step(c, tdt, suppressed != 0);
}

int oqg_adj_get_wekp(oqg_ctx* c, double* out, int kind) {
    ctx_join(c);
    if (!c || !out) return OQG_ERR_ARG;
    if (int r = adj_require(c)) return r;
    if (int r = download_field(c, out, c->lwek, 1, kind)) return r;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

int oqg_adj_zero_wekp(oqg_ctx* c) {
    ctx_join(c);
    if (!c) return OQG_ERR_ARG;
    if (int r = adj_require(c)) return r;
    HIPCHK(c, hipMemsetAsync(c->lwek, 0, (size_t)c->g.fs * sizeof(double), c->st));
    HIPCHK(c, hipMemsetAsync(c->ltxi, 0, 2 * sizeof(double), c->st));
    return 0;
}

int oqg_adj_seed_ke(oqg_ctx* c, double rho, double* J) {
    ctx_join(c);
    if (!c || !J) return OQG_ERR_ARG;
    if (int r = adj_require(c)) return r;
    oqg_diag d;
    if (int r = oqg_diagnose(c, 1.0, rho, &d)) return r;   // ekei from the device's row sums
    *J = d.ekei[0];
    const int cur = c->cur;
    for (double* f : {c->lq[cur], c->lq[1 - cur], c->lp[cur], c->lp[1 - cur]})
        HIPCHK(c, hipMemsetAsync(f, 0, c->fbytes, c->st));
    HIPCHK(c, hipMemsetAsync(c->lc[0], 0, 2 * LC_LEN * sizeof(double), c->st));
    KeSeedArgs a{};
    a.g = c->g;
    a.bcco = c->d.bcco;
    a.rhoh = rho * c->d.h[0];
    a.p = c->pbuf[cur];
    a.lp = c->lp[cur];
    launch_ke_seed(a, c->st);
    HIPCHK(c, hipGetLastError());
    return 0;
}

int oqg_invert(oqg_ctx* c) {
    ctx_join(c);
    if (!c) return OQG_ERR_ARG;
    return enqueue_invert(c);
}

int oqg_adj_invert(oqg_ctx* c) {
    ctx_join(c);
    if (!c) return OQG_ERR_ARG;
    if (int r = adj_require(c)) return r;
    const int cur = c->cur;
    const long long n = (long long)c->g.nl * c->g.fs;
    launch_adj_add(c->lq[cur], c->lq[1 - cur], n, c->st);
    launch_adj_add(c->lp[cur], c->lp[1 - cur], n, c->st);
    // the constraint state is an input of its own, not a function of q: its
    // adjoint goes to scratch
    if (int r = enqueue_adj_inversion(c, c->lq[cur], c->lp[cur], c->lc_scratch)) return r;
    // only the interior of q is an input of the inversion
    launch_zero_ring(c->lq[cur], c->g, c->st);
    HIPCHK(c, hipMemsetAsync(c->lq[1 - cur], 0, c->fbytes, c->st));
    HIPCHK(c, hipMemsetAsync(c->lp[cur], 0, c->fbytes, c->st));
    HIPCHK(c, hipMemsetAsync(c->lp[1 - cur], 0, c->fbytes, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

}  // extern "C"

// The slot pointers on the device, for graph-replayed record and sweep
// stretches.  A new array each time slots are added (the old one stays
// allocated until oqg_destroy); the graphs captured with the old one go.
static int snap_sync_dev(oqg_ctx* c) {
    HIPCHK(c, hipStreamSynchronize(c->st));
    double** p = nullptr;
    if (int r = dalloc(c, &p, c->snaps.size())) return r;
    HIPCHK(c, hipMemcpyAsync(p, c->snaps.data(), c->snaps.size() * sizeof(double*), hipMemcpyHostToDevice, c->st));
    if (!c->snap_base_dev)
        if (int r = dalloc(c, &c->snap_base_dev, 1)) return r;
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->snap_ptrs_dev = p;
    for (auto it = c->graphs.begin(); it != c->graphs.end();) {
        const int k = it->first.kind;
        if (k == G_RECORD || k == G_SWEEP) {
            hipGraphExecDestroy(it->second);
            it = c->graphs.erase(it);
        } else {
            ++it;
        }
    }
    return 0;
}

extern "C" {

int oqg_snap_alloc(oqg_ctx* c, int nslots) {
    ctx_join(c);
    if (!c || nslots < 0) return OQG_ERR_ARG;
    const size_t words = 4 * (size_t)c->g.nl * c->g.fs + 2 * C_LEN;
    const size_t had = c->snaps.size();
    while ((int)c->snaps.size() < nslots) {
        double* p = nullptr;
        if (int r = dalloc(c, &p, words)) return r;
        c->snaps.push_back(p);
        c->snap_cur.push_back(0);
    }
    if (c->snaps.size() != had)
        if (int r = snap_sync_dev(c)) return r;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

}  // extern "C"

// Snapshot slot layout: q[0], q[1], p[0], p[1] (nl fields each), c[0], c[1].
static void snap_copy(oqg_ctx* c, int slot, bool save) {
    double* s = c->snaps[slot];
    const long long fl = (long long)c->g.nl * c->g.fs;
    double* state[6] = {c->qbuf[0], c->qbuf[1], c->pbuf[0], c->pbuf[1], c->cbuf[0], c->cbuf[1]};
    CopySegs a{};
    a.nseg = 6;
    long long off = 0;
    for (int t = 0; t < 6; t++) {
        const long long n = t < 4 ? fl : (long long)C_LEN;
        a.src[t] = save ? state[t] : s + off;
        a.dst[t] = save ? s + off : state[t];
        a.n[t] = n;
        off += n;
    }
    launch_copy_segs(a, c->st);
    if (save) c->snap_cur[slot] = c->cur;
    else c->cur = c->snap_cur[slot];
}

// Inside a captured stretch: save the state in slot (*snap_base_dev + t).
static void snap_save_dev(oqg_ctx* c, int t) {
    const long long fl = (long long)c->g.nl * c->g.fs;
    const double* state[6] = {c->qbuf[0], c->qbuf[1], c->pbuf[0], c->pbuf[1], c->cbuf[0], c->cbuf[1]};
    CopySegs a{};
    a.nseg = 6;
    long long off = 0;
    for (int s = 0; s < 6; s++) {
        const long long n = s < 4 ? fl : (long long)C_LEN;
        a.src[s] = state[s];
        a.doff[s] = off;
        a.n[s] = n;
        off += n;
    }
    a.slots = c->snap_ptrs_dev;
    a.base = c->snap_base_dev;
    a.t = t;
    launch_copy_segs(a, c->st);
}

static bool suppressed_at(int nt, int supp) { return supp > 0 && (nt - 1) % supp == 0; }

extern "C" {

int oqg_snap_save(oqg_ctx* c, int slot) {
    ctx_join(c);
    if (!c || slot < 0 || slot >= (int)c->snaps.size()) return c ? fail(c, OQG_ERR_ARG, "snapshot slot out of range") : OQG_ERR_ARG;
    snap_copy(c, slot, true);
    HIPCHK(c, hipGetLastError());
    return 0;
}

int oqg_snap_load(oqg_ctx* c, int slot) {
    ctx_join(c);
    if (!c || slot < 0 || slot >= (int)c->snaps.size()) return c ? fail(c, OQG_ERR_ARG, "snapshot slot out of range") : OQG_ERR_ARG;
    snap_copy(c, slot, false);
    HIPCHK(c, hipGetLastError());
    return 0;
}

int oqg_adj_record(oqg_ctx* c, double tdt, int nsteps, int step0, int supp, int slot0) {
    ctx_join(c);
    if (!c || nsteps < 0 || slot0 < 0) return OQG_ERR_ARG;
    if ((long long)slot0 + nsteps > (long long)c->snaps.size())
        return fail(c, OQG_ERR_ARG, "oqg_adj_record: fewer snapshot slots than steps");
    // aligned stretches of P steps replay one graph (as oqg_run), the rest is eager
    const int P = graph_period(supp);
    int nt = step0 + 1;
    const int end = step0 + nsteps;
    while (nt <= end) {
        const int slot = slot0 + (nt - 1 - step0);
        if ((nt - 1) % P == 0 && end - nt + 1 >= P) {
            const int cur0 = c->cur;
            hipGraphExec_t ex;
            const int rc = cached_graph(c, GraphKey{tdt, P, supp, cur0, G_RECORD}, [&]() {
                for (int t = 0; t < P; t++) {
                    snap_save_dev(c, t);
                    if (int r = enqueue_step(c, tdt)) return r;
                    if (suppressed_at(t + 1, supp))
                        if (int r = enqueue_suppress(c)) return r;
                }
                return 0;
            }, &ex);
            if (rc) return rc;
            launch_set_int(c->snap_base_dev, slot, c->st);
            HIPCHK(c, hipGraphLaunch(ex, c->st));
            for (int t = 0; t < P; t++) c->snap_cur[slot + t] = cur0 ^ (t & 1);
            nt += P;   // P even: the parity is back to cur0
        } else {
            snap_copy(c, slot, true);
            if (int r = enqueue_step(c, tdt)) return r;
            if (suppressed_at(nt, supp))
                if (int r = enqueue_suppress(c)) return r;
            nt++;
        }
    }
    HIPCHK(c, hipGetLastError());
    return 0;
}

int oqg_adj_sweep(oqg_ctx* c, double tdt, int nsteps, int step0, int supp, int slot0) {
    ctx_join(c);
    if (!c || nsteps < 0 || slot0 < 0) return OQG_ERR_ARG;
    if (int r = adj_require(c)) return r;
    if ((long long)slot0 + nsteps > (long long)c->snaps.size())
        return fail(c, OQG_ERR_ARG, "oqg_adj_sweep: fewer snapshot slots than steps");
    const long long fl = (long long)c->g.nl * c->g.fs;
    // the adjoint step reads only q_n and p_n, straight from the snapshot;
    // aligned stretches of P steps (steps b+1 .. b+P, b a multiple of P)
    // replay one graph, the rest is eager
    const int P = graph_period(supp);
    int nt = step0 + nsteps;   // the next step to take back
    while (nt > step0) {
        const int b = nt - P;
        if (b >= step0 && b % P == 0) {
            const int base = slot0 + (b - step0), cb = c->snap_cur[base];
            bool alternating = true;
            for (int t = 0; t < P; t++) alternating = alternating && c->snap_cur[base + t] == (cb ^ (t & 1));
            if (alternating) {
                hipGraphExec_t ex;
                const int rc = cached_graph(c, GraphKey{tdt, P, supp, cb, G_SWEEP}, [&]() {
                    for (int t = P - 1; t >= 0; t--) {
                        c->cur = cb ^ (t & 1);
                        SnapRef src;
                        src.slots = c->snap_ptrs_dev;
                        src.base = c->snap_base_dev;
                        src.t = t;
                        src.qoff = c->cur ? fl : 0;
                        src.poff = 2 * fl + (c->cur ? fl : 0);
                        if (int r = enqueue_adj_step(c, tdt, suppressed_at(t + 1, supp), src)) return r;
                    }
                    return 0;
                }, &ex);
                if (rc) return rc;
                launch_set_int(c->snap_base_dev, base, c->st);
                HIPCHK(c, hipGraphLaunch(ex, c->st));
                c->cur = cb;
                nt = b;
                continue;
            }
        }
        const int slot = slot0 + (nt - 1 - step0), sc = c->snap_cur[slot];
        const double* s = c->snaps[slot];
        c->cur = sc;
        SnapRef src;
        src.q = s + (sc ? fl : 0);
        src.p = s + 2 * fl + (sc ? fl : 0);
        if (int r = enqueue_adj_step(c, tdt, suppressed_at(nt, supp), src)) return r;
        nt--;
    }
    if (nsteps > 0) snap_copy(c, slot0, false);
    HIPCHK(c, hipGetLastError());
    return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ binomial checkpointing
//
// Reversing n steps from a stored state with s further free snapshot slots:
// advance m steps, store that state, reverse the last n - m steps with s - 1
// free slots, then the first m with s.  With no free slot, every step is
// recomputed from the stored state.  The forward steps this costs, at the best
// m on every level, are Griewank's t(n, c) = r n - C(c + r, r - 1) with c = s + 1
// checkpoints and r the least repetition number with C(c + r, r) >= n
// (Griewank and Walther, ACM TOMS 26, 2000).

static long long binom(int a, int b) {   // C(a, b), exact for the sizes used here
    if (b < 0 || b > a) return 0;
    b = std::min(b, a - b);
    long long v = 1;
    for (int i = 1; i <= b; i++) v = v * (a - b + i) / i;
    return v;
}

static long long rev_cost(int n, int s) {
    if (n <= 1) return 0;
    if (s == 0) return (long long)n * (n - 1) / 2;
    const int c = s + 1;
    int r = 0;
    while (binom(c + r, c) < n) r++;
    return (long long)r * n - binom(c + r, r - 1);
}

static int rev_split(oqg_ctx* c, int n, int s) {
    if (n <= 2) return 1;
    if (c) {
        auto it = c->rv_split.find({n, s});
        if (it != c->rv_split.end()) return it->second;
    }
    int best_m = 1;
    long long best = -1;
    for (int m = 1; m < n; m++) {
        const long long v = m + rev_cost(n - m, s - 1) + rev_cost(m, s);
        if (best < 0 || v < best) {
            best = v;
            best_m = m;
        }
    }
    if (c) c->rv_split[{n, s}] = best_m;
    return best_m;
}

// One adjoint step (eager launches): q_n, p_n from snapshot `slot`, or from
// the loaded state when slot < 0.  c->cur must be the parity of that state.
// Level t of the snapshot slot at s (q, p and the constraint block).
static oqg_ctx::LevelRef slot_level(const oqg_ctx* c, const double* s, int t) {
    const long long fl = (long long)c->g.nl * c->g.fs;
    oqg_ctx::LevelRef r;
    r.q = s + t * fl;
    r.p = s + 2 * fl + t * fl;
    r.c = s + 4 * fl + t * C_LEN;
    return r;
}

// Where level t of checkpoint `slot` is: its own allocation, or (reverse
// sweep) the slot that held it when the checkpoint was taken.
static oqg_ctx::LevelRef slot_lvl(const oqg_ctx* c, int slot, int t) {
    if (slot < (int)c->snap_ref.size() && c->snap_ref[slot][t].q) return c->snap_ref[slot][t];
    return slot_level(c, c->snaps[slot], t);
}

static int adj_step_at(oqg_ctx* c, double tdt, bool suppressed, int slot) {
    SnapRef src;
    if (slot >= 0) {
        const oqg_ctx::LevelRef l = slot_lvl(c, slot, c->cur);
        src.q = l.q;
        src.p = l.p;
    } else if (c->ext_cur.q) {   // the loaded state's current level is a checkpoint's
        src.q = c->ext_cur.q;
        src.p = c->ext_cur.p;
    }
    return enqueue_adj_step(c, tdt, suppressed, src);
}

// Copies into the working buffers the levels a step would read from a
// checkpoint, so that they hold the whole state again.
static void materialize(oqg_ctx* c) {
    if (!c->ext_cur.q && !c->ext_old.q) return;
    const long long fl = (long long)c->g.nl * c->g.fs;
    CopySegs a{};
    int n = 0;
    auto add = [&](const double* src, double* dst, long long cnt) {
        a.src[n] = src;
        a.dst[n] = dst;
        a.n[n] = cnt;
        n++;
    };
    const int cur = c->cur;
    if (c->ext_cur.q) {
        add(c->ext_cur.q, c->qbuf[cur], fl);
        add(c->ext_cur.p, c->pbuf[cur], fl);
        add(c->ext_cur.c, c->cbuf[cur], C_LEN);
    }
    if (c->ext_old.q) {
        add(c->ext_old.q, c->qbuf[1 - cur], fl);
        add(c->ext_old.p, c->pbuf[1 - cur], fl);
        add(c->ext_old.c, c->cbuf[1 - cur], C_LEN);
    }
    a.nseg = n;
    launch_copy_segs(a, c->st);
    c->ext_cur = {};
    c->ext_old = {};
}

// Forward sweep: checkpoints are copies and the advances replay graphs.
// Reverse sweep (rv_ref): a checkpoint takes over the working allocation
// (the working buffers move to the slot's), a restore only points the next
// step's inputs at the slot, and the advances launch their steps one by one
// (the graphs hold the context's own buffers).  The first two steps after a
// checkpoint read their inputs from it and write the working buffers, so no
// state is copied: a copy of the 257^2 x 3 state took 4.4 us, about 2100 of
// them a reverse sweep of 2160 steps with 32 slots.  Most checkpoints follow a
// single step from the previous one, whose current level is then the new
// checkpoint's previous level: the new slot refers to it there.  That slot
// is not written while the new one is in use (revolve's sub-sweeps use only
// higher slots), so the reference stays valid.
static int rv_advance(oqg_ctx* c, int lo, int hi) {
    if (hi <= lo) return 0;
    c->rv_steps += hi - lo;
    if (c->rv_ref) return run_eager(c, c->rv_tdt, lo + 1, hi - lo, c->rv_supp);
    return oqg_run(c, c->rv_tdt, hi - lo, lo, c->rv_supp);
}

static void rv_save(oqg_ctx* c, int slot) {
    if (!c->rv_ref) {
        snap_copy(c, slot, true);
        return;
    }
    if (c->ext_cur.q) materialize(c);   // (every checkpoint follows a step)
    const oqg_ctx::LevelRef old = c->ext_old;   // the previous level, if a checkpoint holds it
    std::swap(c->wbase, c->snaps[slot]);
    set_work(c, c->wbase);
    c->snap_cur[slot] = c->cur;
    c->snap_ref[slot][c->cur] = {};
    c->snap_ref[slot][1 - c->cur] = old;
    c->ext_cur = slot_lvl(c, slot, c->cur);
    c->ext_old = slot_lvl(c, slot, 1 - c->cur);
}

static void rv_restore(oqg_ctx* c, int slot) {
    if (!c->rv_ref) {
        snap_copy(c, slot, false);
        return;
    }
    c->cur = c->snap_cur[slot];
    c->ext_cur = slot_lvl(c, slot, c->cur);
    c->ext_old = slot_lvl(c, slot, 1 - c->cur);
}

// The state lo is in `slot`; the adjoint holds that of state hi; leave it at state lo.
static int rv_reverse(oqg_ctx* c, int lo, int hi, int slot, int free) {
    const int n = hi - lo;
    const double tdt = c->rv_tdt;
    const int supp = c->rv_supp;
    if (n == 1) {
        c->cur = c->snap_cur[slot];
        return adj_step_at(c, tdt, suppressed_at(hi, supp), slot);
    }
    if (free == 0) {
        for (int k = hi - 1; k >= lo; k--) {
            if (k == lo) {
                c->cur = c->snap_cur[slot];
                if (int r = adj_step_at(c, tdt, suppressed_at(k + 1, supp), slot)) return r;
                continue;
            }
            rv_restore(c, slot);
            if (int r = rv_advance(c, lo, k)) return r;
            if (int r = adj_step_at(c, tdt, suppressed_at(k + 1, supp), -1)) return r;
        }
        return 0;
    }
    const int m = lo + rev_split(c, n, free);
    rv_restore(c, slot);
    if (int r = rv_advance(c, lo, m)) return r;
    if (hi - m == 1) {
        // the checkpoint at m would serve this one adjoint step only: it
        // reads the loaded state instead
        if (int r = adj_step_at(c, tdt, suppressed_at(hi, supp), -1)) return r;
        return rv_reverse(c, lo, m, slot, free);
    }
    rv_save(c, slot + 1);
    if (int r = rv_reverse(c, m, hi, slot + 1, free - 1)) return r;
    return rv_reverse(c, lo, m, slot, free);
}

extern "C" {

long long oqg_revolve_cost(int nsteps, int free_slots) {
    if (nsteps < 0 || free_slots < 0) return -1;
    return rev_cost(nsteps, free_slots);
}

int oqg_revolve_split(int nsteps, int free_slots) {
    if (nsteps < 2 || free_slots < 1) return -1;
    return rev_split(nullptr, nsteps, free_slots);
}

int oqg_adj_revolve_forward(oqg_ctx* c, double tdt, int nsteps, int supp, int free_slots) {
    ctx_join(c);
    if (!c || nsteps < 1 || free_slots < 0) return OQG_ERR_ARG;
    if ((int)c->snaps.size() < free_slots + 1)
        return fail(c, OQG_ERR_ARG, "oqg_adj_revolve_forward: needs free_slots + 1 snapshot slots");
    c->rv_tdt = tdt;
    c->rv_supp = supp;
    c->rv_n = nsteps;
    c->rv_steps = 0;
    c->rv_left.clear();
    snap_copy(c, 0, true);
    int lo = 0, slot = 0, free = free_slots;
    while (nsteps - lo > 1 && free > 0) {
        const int m = lo + rev_split(c, nsteps - lo, free);
        if (int r = rv_advance(c, lo, m)) return r;
        snap_copy(c, slot + 1, true);
        c->rv_left.push_back({lo, m, slot, free});
        lo = m;
        slot++;
        free--;
    }
    if (int r = rv_advance(c, lo, nsteps)) return r;
    c->rv_tail = {lo, slot, free};
    HIPCHK(c, hipGetLastError());
    return 0;
}

int oqg_adj_revolve_reverse(oqg_ctx* c, long long* forward_steps) {
    ctx_join(c);
    if (!c) return OQG_ERR_ARG;
    if (int r = adj_require(c)) return r;
    if (c->rv_n < 1) return fail(c, OQG_ERR_STATE, "oqg_adj_revolve_reverse: no forward sweep recorded");
    c->rv_ref = true;
    c->snap_ref.assign(c->snaps.size(), std::array<oqg_ctx::LevelRef, 2>{});
    int rc = rv_reverse(c, c->rv_tail[0], c->rv_n, c->rv_tail[1], c->rv_tail[2]);
    for (auto it = c->rv_left.rbegin(); !rc && it != c->rv_left.rend(); ++it)
        rc = rv_reverse(c, (*it)[0], (*it)[1], (*it)[2], (*it)[3]);
    c->rv_ref = false;
    c->snap_ref.clear();
    // the context takes its own allocation back (slot 0, the initial state,
    // never changed hands) and loads the initial state into it
    c->ext_cur = {};
    c->ext_old = {};
    if (c->wbase != c->wbase_own) {
        for (double*& sl : c->snaps)
            if (sl == c->wbase_own) {
                std::swap(sl, c->wbase);
                break;
            }
        set_work(c, c->wbase);
        if (c->snap_ptrs_dev) {
            HIPCHK(c, hipMemcpyAsync(c->snap_ptrs_dev, c->snaps.data(), c->snaps.size() * sizeof(double*),
                                     hipMemcpyHostToDevice, c->st));
            HIPCHK(c, hipStreamSynchronize(c->st));
        }
    }
    if (rc) return rc;
    snap_copy(c, 0, false);
    c->rv_n = -1;
    if (forward_steps) *forward_steps = c->rv_steps;
    HIPCHK(c, hipGetLastError());
    return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ ensembles
//
// A batched ensemble advances many basins together.  Each phase of the step is
// one launch over the blocks of all basins that share its template parameters
// (layer count and FFT butterflies for the x transforms, rows per lane for the
// y solve), driven by a block map (member, member's own block).  The members'
// argument blocks live in device memory, one table per time-level parity.  The
// set of members with steps left and its launch schedule form a plan; lcm(2,
// supp) steps of a plan are captured in a HIP graph and replayed.

namespace {

enum { ENS_VORT = 0, ENS_FWD, ENS_YS, ENS_INV, ENS_FWDW, ENS_INVW };

struct EnsLaunch {
    int kind = ENS_VORT;
    int nl = 0, mb = 0, R = 0;   // mb: butterflies per thread, or points per lane (wavefront forms)
    bool cyc = false;            // wavefront inverse form: channel members
    size_t lds = 0;
    int nblocks = 0;
    int2* map = nullptr;
};

struct EnsPlan {
    std::vector<int> members;   // indices into oqg_ens::m, in launch order
    VortArgs* va[2] = {nullptr, nullptr};
    FwdArgs* fa[2] = {nullptr, nullptr};
    YArgs* ya = nullptr;
    InvArgs* ia[2] = {nullptr, nullptr};
    SuppArgs* sa[2] = {nullptr, nullptr};
    std::vector<EnsLaunch> launches;
    std::vector<void*> allocs;
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;   // (period, supp)
};

}  // namespace

struct oqg_ens {
    std::vector<oqg_ctx*> m;
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;   // recorded after each run; the members' streams wait for it
    int dev = 0;
    std::string err;
    int schedule = 0;          // OQG_ENS_* bits
    std::vector<double> tdt;   // per member, baked into the plans' argument tables
    std::map<std::vector<int>, EnsPlan*> plans;   // key: members, then their parities at odd steps, then schedule
};

static int ens_fail(oqg_ens* e, int code, const std::string& msg) {
    e->err = msg;
    return code;
}

#define ENSCHK(e, expr)                                                                          \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return ens_fail(e, OQG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static void ens_free_plan(EnsPlan* p) {
    for (auto& kv : p->graphs) hipGraphExecDestroy(kv.second);
    for (void* a : p->allocs) hipFree(a);
    delete p;
}

static void ens_clear_plans(oqg_ens* e) {
    if (e->st) hipStreamSynchronize(e->st);
    for (auto& kv : e->plans) ens_free_plan(kv.second);
    e->plans.clear();
}

template <class T>
static int ens_upload(oqg_ens* e, EnsPlan* p, T** dst, const std::vector<T>& v) {
    void* d = nullptr;
    ENSCHK(e, hipMalloc(&d, std::max<size_t>(1, v.size()) * sizeof(T)));
    p->allocs.push_back(d);
    if (!v.empty()) ENSCHK(e, hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, e->st));
    *dst = static_cast<T*>(d);
    return 0;
}

// Plan for the members `act` whose time-level parity before an odd-numbered
// step (nt - 1 even) is base[i].
static int ens_build_plan(oqg_ens* e, const std::vector<int>& act, const std::vector<int>& base, EnsPlan** out) {
    std::unique_ptr<EnsPlan, void (*)(EnsPlan*)> P(new EnsPlan, ens_free_plan);
    const int n = (int)act.size();
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    if (e->schedule & OQG_ENS_LARGEST_FIRST) {
        auto work = [&](int i) {
            const oqg_ctx* c = e->m[act[i]];
            return (long long)c->g.nxp * c->g.nyp * c->g.nl;
        };
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
    }
    for (int i = 0; i < n; i++) P->members.push_back(act[order[i]]);
    std::vector<int> pbase(n);
    for (int i = 0; i < n; i++) pbase[i] = base[order[i]];

    // argument tables for both parities
    for (int par = 0; par < 2; par++) {
        std::vector<VortArgs> va;
        std::vector<FwdArgs> fa;
        std::vector<YArgs> ya;
        std::vector<InvArgs> ia;
        std::vector<SuppArgs> sa;
        for (int i = 0; i < n; i++) {
            oqg_ctx* c = e->m[P->members[i]];
            const int saved = c->cur;
            c->cur = pbase[i] ^ par;
            StepArgs a;
            make_step_args(c, e->tdt[P->members[i]], a);
            va.push_back(a.va);
            fa.push_back(a.fa);
            ya.push_back(a.ya);
            ia.push_back(a.ia);
            SuppArgs s{};   // suppression follows the step: the new level is 1 - cur
            const int cn = 1 - c->cur;
            s.g = c->g;
            s.q = c->qbuf[cn];
            s.qm = c->qbuf[1 - cn];
            s.p = c->pbuf[cn];
            s.pm = c->pbuf[1 - cn];
            s.c = c->cbuf[cn];
            sa.push_back(s);
            c->cur = saved;
        }
        if (int r = ens_upload(e, P.get(), &P->va[par], va)) return r;
        if (int r = ens_upload(e, P.get(), &P->fa[par], fa)) return r;
        if (int r = ens_upload(e, P.get(), &P->ia[par], ia)) return r;
        if (int r = ens_upload(e, P.get(), &P->sa[par], sa)) return r;
        if (par == 0)
            if (int r = ens_upload(e, P.get(), &P->ya, ya)) return r;
    }

    // block maps: one launch for the stencil, one per (layer count, butterflies)
    // class for each x transform, one per rows-per-lane class (or one for all)
    // for the y solve
    auto add_launch = [&](EnsLaunch L, const std::vector<int2>& map) -> int {
        L.nblocks = (int)map.size();
        if (!L.nblocks) return 0;
        if (int r = ens_upload(e, P.get(), &L.map, map)) return r;
        P->launches.push_back(L);
        return 0;
    };
    for (int ty : {4, 8, 16}) {   // one stencil launch per tile height
        std::vector<int2> map;
        for (int i = 0; i < n; i++) {
            const oqg_ctx* c = e->m[P->members[i]];
            if (c->vort_ty != ty) continue;
            const int nb = c->tiles_x * c->tiles_y * c->g.nl;
            for (int b = 0; b < nb; b++) map.push_back(make_int2(i, b));
        }
        EnsLaunch L;
        L.kind = ENS_VORT;
        L.mb = ty;
        if (int r = add_launch(L, map)) return r;
    }
    // x-transform classes: (wavefront form?, layer count, points per lane or butterflies per thread)
    std::map<std::array<int, 4>, std::pair<std::vector<int2>, size_t>> fwd, inv;
    std::map<int, std::pair<std::vector<int2>, size_t>> ys;
    int rmax = 1;
    for (int i = 0; i < n; i++) rmax = std::max(rmax, e->m[P->members[i]]->s.R);
    for (int i = 0; i < n; i++) {
        const oqg_ctx* c = e->m[P->members[i]];
        const bool w = fftw_fits(c->s, c->g.nl);
        const std::array<int, 4> key{w ? 1 : 0, c->g.nl, w ? fftw_points(c->s) : fft_mb(c->s, false), 0};
        const std::array<int, 4> ikey{w ? 1 : 0, c->g.nl, w ? fftw_points(c->s) : fft_mb(c->s, true),
                                      w ? c->g.cyclic : 0};
        auto& f = fwd[key];
        const int np = inv_block_count(c->s);
        for (int b = 0; b < np * (w ? 1 : c->g.nl); b++) f.first.push_back(make_int2(i, b));
        f.second = std::max(f.second, w ? fftw_lds_bytes(c->s, c->g.nl) : fwd_lds_bytes(c->s));
        auto& v = inv[ikey];
        for (int b = 0; b < np; b++) v.first.push_back(make_int2(i, b));
        v.second = std::max(v.second, w ? fftw_lds_bytes(c->s, c->g.nl) : inv_lds_bytes(c->s, c->g, c->g.nl));
        auto& y = ys[(e->schedule & OQG_ENS_ONE_YSOLVE) ? rmax : c->s.R];
        for (int b = 0; b < ysolve_blocks(c->s, c->g.nl); b++) y.first.push_back(make_int2(i, b));
        y.second = std::max(y.second, ysolve_lds_bytes(c->s));
    }
    for (auto& kv : fwd) {
        EnsLaunch L;
        L.kind = kv.first[0] ? ENS_FWDW : ENS_FWD;
        L.nl = kv.first[1];
        L.mb = kv.first[2];
        L.lds = kv.second.second;
        if (int r = add_launch(L, kv.second.first)) return r;
    }
    for (auto& kv : ys) {
        EnsLaunch L;
        L.kind = ENS_YS;
        L.R = kv.first;
        L.lds = std::max(kv.second.second, sizeof(double) * YS_CT * (size_t)ys_pitch(kv.first));
        if (int r = add_launch(L, kv.second.first)) return r;
    }
    for (auto& kv : inv) {
        EnsLaunch L;
        L.kind = kv.first[0] ? ENS_INVW : ENS_INV;
        L.nl = kv.first[1];
        L.mb = kv.first[2];
        L.cyc = kv.first[3] != 0;
        L.lds = kv.second.second;
        if (int r = add_launch(L, kv.second.first)) return r;
    }
    ENSCHK(e, hipStreamSynchronize(e->st));   // the host tables above go out of scope
    *out = P.release();
    return 0;
}

// One step of every member of plan P; par = (nt - 1) & 1.
static int ens_enqueue_step(oqg_ens* e, EnsPlan* P, int par, bool supp) {
    for (const EnsLaunch& L : P->launches) {
        switch (L.kind) {
            case ENS_VORT:
                launch_vort_ens(P->va[par], L.map, L.nblocks, L.mb, e->st);
                break;
            case ENS_FWD: launch_fwd_ens(P->fa[par], L.map, L.nblocks, L.nl, L.mb, L.lds, e->st); break;
            case ENS_FWDW: launch_fwdw_ens(P->fa[par], L.map, L.nblocks, L.nl, L.mb, L.lds, e->st); break;
            case ENS_INVW: launch_invw_ens(P->ia[par], L.map, L.nblocks, L.nl, L.mb, L.cyc, L.lds, e->st); break;
            case ENS_YS: launch_ysolve_ens(P->ya, L.map, L.nblocks, L.R, L.lds, e->st); break;
            default: launch_inv_ens(P->ia[par], L.map, L.nblocks, L.nl, L.mb, L.lds, e->st); break;
        }
    }
    if (supp) launch_suppress_ens(P->sa[par], (int)P->members.size(), e->st);
    ENSCHK(e, hipGetLastError());
    return 0;
}

static int ens_get_plan(oqg_ens* e, const std::vector<int>& act, const std::vector<int>& base, EnsPlan** out) {
    std::vector<int> key = act;
    key.insert(key.end(), base.begin(), base.end());
    key.push_back(e->schedule);
    auto it = e->plans.find(key);
    if (it != e->plans.end()) {
        *out = it->second;
        return 0;
    }
    EnsPlan* P = nullptr;
    if (int r = ens_build_plan(e, act, base, &P)) return r;
    e->plans[key] = P;
    *out = P;
    return 0;
}

static int ens_run(oqg_ens* e, const double* tdt, const int* nsteps, int step0, int supp) {
    const int n = (int)e->m.size();
    bool same = (int)e->tdt.size() == n;
    for (int i = 0; same && i < n; i++) same = e->tdt[i] == tdt[i];
    if (!same) {
        ens_clear_plans(e);
        e->tdt.assign(tdt, tdt + n);
    }
    int maxs = 0;
    std::vector<int> cur0(n);
    for (int i = 0; i < n; i++) {
        if (nsteps[i] < 0) return ens_fail(e, OQG_ERR_ARG, "negative step count");
        maxs = std::max(maxs, nsteps[i]);
        oqg_ctx* c = e->m[i];
        if (c->ens_done && c->ens_done != e->done) ENSCHK(e, hipStreamWaitEvent(e->st, c->ens_done, 0));   // another ensemble's run
        c->ens_done = nullptr;
        ENSCHK(e, hipStreamSynchronize(c->st));
        cur0[i] = e->m[i]->cur;
    }
    const int Pd = supp > 0 ? (supp % 2 == 0 ? supp : 2 * supp) : 2;
    int nt = step0 + 1;
    const int end = step0 + maxs;
    while (nt <= end) {
        std::vector<int> act, base;
        int seg_end = end;
        for (int i = 0; i < n; i++)
            if (nt <= step0 + nsteps[i]) {
                act.push_back(i);
                base.push_back(cur0[i] ^ (step0 & 1));
                seg_end = std::min(seg_end, step0 + nsteps[i]);
            }
        EnsPlan* P = nullptr;
        if (int r = ens_get_plan(e, act, base, &P)) return r;
        while (nt <= seg_end) {
            if ((nt - 1) % Pd == 0 && seg_end - nt + 1 >= Pd) {
                const std::pair<int, int> gk(Pd, supp);
                auto it = P->graphs.find(gk);
                if (it == P->graphs.end()) {
                    hipGraph_t gr;
                    ENSCHK(e, hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal));
                    int rc = 0;
                    for (int t = 0; t < Pd && !rc; t++) rc = ens_enqueue_step(e, P, t & 1, supp > 0 && t % supp == 0);
                    hipError_t ce = hipStreamEndCapture(e->st, &gr);
                    if (rc) return rc;
                    if (ce != hipSuccess) return ens_fail(e, OQG_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
                    hipGraphExec_t ex;
                    hipError_t ie = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
                    hipGraphDestroy(gr);
                    if (ie != hipSuccess)
                        return ens_fail(e, OQG_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
                    it = P->graphs.emplace(gk, ex).first;
                }
                ENSCHK(e, hipGraphLaunch(it->second, e->st));
                nt += Pd;
            } else {
                if (int r = ens_enqueue_step(e, P, (nt - 1) & 1, supp > 0 && (nt - 1) % supp == 0)) return r;
                nt++;
            }
        }
    }
    for (int i = 0; i < n; i++) e->m[i]->cur = cur0[i] ^ (nsteps[i] & 1);
    // work later enqueued on a member's own stream starts after the ensemble's:
    // the member's next call makes its stream wait for this event (ctx_join).
    // Waiting here instead would put a barrier packet in a hardware queue that
    // other streams may share, holding their work behind this run's end.
    ENSCHK(e, hipEventRecord(e->done, e->st));
    for (int i = 0; i < n; i++) e->m[i]->ens_done = e->done;
    return 0;
}

extern "C" {

int oqg_ens_create(oqg_ctx* const* members, int n, oqg_ens** out) {
    if (!members || n < 1 || !out) return fail(nullptr, OQG_ERR_ARG, "oqg_ens_create: members and n >= 1 required");
    *out = nullptr;
    for (int i = 0; i < n; i++) {
        if (!members[i]) return fail(nullptr, OQG_ERR_ARG, "oqg_ens_create: null member");
        for (int j = 0; j < i; j++)
            if (members[j] == members[i]) return fail(nullptr, OQG_ERR_ARG, "oqg_ens_create: a basin appears twice");
        if (members[i]->s.prime)
            return fail(nullptr, OQG_ERR_ARG,
                        "oqg_ens_create: batched members take x transform lengths that factor into 2, 3 and 5");
    }
    std::unique_ptr<oqg_ens> e(new oqg_ens);
    e->m.assign(members, members + n);
    e->dev = members[0]->dev;
    if (pool_take(&e->st) != hipSuccess) return fail(nullptr, OQG_ERR_HIP, "hipStreamCreate failed");
    if (hipEventCreateWithFlags(&e->done, hipEventDisableTiming) != hipSuccess) {
        pool_give(e->dev, e->st);
        return fail(nullptr, OQG_ERR_HIP, "hipEventCreate failed");
    }
    *out = e.release();
    return 0;
}

void oqg_ens_destroy(oqg_ens* e) {
    if (!e) return;
    ens_clear_plans(e);   // synchronises the ensemble stream
    for (oqg_ctx* c : e->m)
        if (c->ens_done == e->done) c->ens_done = nullptr;
    pool_give(e->dev, e->st);
    if (e->done) hipEventDestroy(e->done);
    delete e;
}

const char* oqg_ens_last_error(const oqg_ens* e) { return e ? e->err.c_str() : g_create_err.c_str(); }

int oqg_ens_run(oqg_ens* e, const double* tdt, const int* nsteps, int step0, int supp_every) {
    if (!e || !tdt || !nsteps || step0 < 0) return OQG_ERR_ARG;
    return ens_run(e, tdt, nsteps, step0, supp_every);
}

int oqg_ens_sync(oqg_ens* e) {
    if (!e) return OQG_ERR_ARG;
    ENSCHK(e, hipStreamSynchronize(e->st));
    return 0;
}

int oqg_ens_set_schedule(oqg_ens* e, int schedule) {
    if (!e || schedule < 0 || schedule > (OQG_ENS_ONE_YSOLVE | OQG_ENS_LARGEST_FIRST)) return OQG_ERR_ARG;
    if (schedule != e->schedule) {
        ens_clear_plans(e);
        e->schedule = schedule;
    }
    return 0;
}

int oqg_ens_get_schedule(const oqg_ens* e) { return e ? e->schedule : -OQG_ERR_ARG; }

int oqg_ens_tune(oqg_ens* e, const double* tdt, int trial_steps, int supp_every, double* ms) {
    if (!e || !tdt || trial_steps < 1) return -OQG_ERR_ARG;
    const int n = (int)e->m.size();
    // keep the members' state: both time levels of q, p and the constraint blocks
    std::vector<void*> keep;
    auto release = [&]() {
        hipStreamSynchronize(e->st);
        for (void* p : keep) hipFree(p);
    };
    std::vector<int> cur0(n), steps(n, trial_steps);
    for (int i = 0; i < n; i++) {
        oqg_ctx* c = e->m[i];
        if (hipStreamSynchronize(c->st) != hipSuccess) return (release(), -ens_fail(e, OQG_ERR_HIP, "member stream"));
        cur0[i] = c->cur;
        for (int t = 0; t < 2; t++) {
            void* a = nullptr;
            void* b = nullptr;
            void* cc = nullptr;
            if (hipMalloc(&a, c->fbytes) != hipSuccess || (keep.push_back(a), hipMalloc(&b, c->fbytes)) != hipSuccess ||
                (keep.push_back(b), hipMalloc(&cc, C_LEN * sizeof(double))) != hipSuccess)
                return (release(), -ens_fail(e, OQG_ERR_HIP, "hipMalloc failed while tuning"));
            keep.push_back(cc);
            hipMemcpyAsync(a, c->qbuf[t], c->fbytes, hipMemcpyDeviceToDevice, e->st);
            hipMemcpyAsync(b, c->pbuf[t], c->fbytes, hipMemcpyDeviceToDevice, e->st);
            hipMemcpyAsync(cc, c->cbuf[t], C_LEN * sizeof(double), hipMemcpyDeviceToDevice, e->st);
        }
    }
    auto restore = [&]() {
        for (int i = 0; i < n; i++) {
            oqg_ctx* c = e->m[i];
            for (int t = 0; t < 2; t++) {
                hipMemcpyAsync(c->qbuf[t], keep[(i * 2 + t) * 3 + 0], c->fbytes, hipMemcpyDeviceToDevice, e->st);
                hipMemcpyAsync(c->pbuf[t], keep[(i * 2 + t) * 3 + 1], c->fbytes, hipMemcpyDeviceToDevice, e->st);
                hipMemcpyAsync(c->cbuf[t], keep[(i * 2 + t) * 3 + 2], C_LEN * sizeof(double), hipMemcpyDeviceToDevice, e->st);
            }
            c->cur = cur0[i];
        }
    };
    hipEvent_t ev[2];
    hipEventCreate(&ev[0]);
    hipEventCreate(&ev[1]);
    int best = 0;
    double best_ms = 1e300;
    const int nsched = OQG_ENS_ONE_YSOLVE | OQG_ENS_LARGEST_FIRST;
    const int tile_opt[4] = {0, 4, 8, 16};   // 0: each member's own rule
    std::vector<int> rows0(n);
    for (int i = 0; i < n; i++) rows0[i] = e->m[i]->vort_ty;
    auto set_tiles = [&](int t) {
        int r = 0;
        for (int i = 0; i < n && !r; i++) r = set_tile_rows(e->m[i], tile_opt[t]);
        ens_clear_plans(e);
        return r;
    };
    int rc = 0;
    for (int t = 0; t < 4 && !rc; t++) {
        rc = set_tiles(t);
        for (int sched = 0; sched <= nsched && !rc; sched++) {
            oqg_ens_set_schedule(e, sched);
            rc = ens_run(e, tdt, steps.data(), 0, supp_every);   // builds the plan and its graphs
            restore();
            if (rc) break;
            hipEventRecord(ev[0], e->st);
            rc = ens_run(e, tdt, steps.data(), 0, supp_every);
            hipEventRecord(ev[1], e->st);
            hipEventSynchronize(ev[1]);
            restore();
            float el = 0.f;
            hipEventElapsedTime(&el, ev[0], ev[1]);
            if (ms) ms[t * (nsched + 1) + sched] = el;
            if (el < best_ms) {
                best_ms = el;
                best = t * (nsched + 1) + sched;
            }
        }
    }
    hipEventDestroy(ev[0]);
    hipEventDestroy(ev[1]);
    release();
    if (rc) {
        for (int i = 0; i < n; i++) set_tile_rows(e->m[i], rows0[i]);
        ens_clear_plans(e);
        return -rc;
    }
    if (int r = set_tiles(best / (nsched + 1))) return -r;
    oqg_ens_set_schedule(e, best % (nsched + 1));
    return best;
}

}  // extern "C"

// ------------------------------------------------------------------ ocean mixed layer

// The ocean mixed layer of a basin (ocean_mixed_type, mixed/omlsubs.F90:13-20):
// its temperature at T points (two time levels and the scratch of a step),
// the resident net heat flux, and the parameters of step_oml.
struct oqg_oml {
    oqg_ctx* c = nullptr;
    oqg_oml_desc d{};
    double* sst = nullptr;    // current level (T layout: the basin's pitch, nyp-1 rows)
    double* sstm = nullptr;   // previous level
    double* snew = nullptr;
    double* xfo = nullptr;
    double* fnet = nullptr;
    double* part = nullptr;
    int npart = 0;
    std::vector<void*> allocs;
    std::map<GraphKey, hipGraphExec_t> graphs;   // oqg_run_ocean
    std::string err;
};

static int oml_fail(oqg_oml* m, int code, const std::string& msg) {
    if (m) m->err = msg;
    else g_create_err = msg;
    return code;
}

#define OMLCHK(m, expr)                                                                          \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return oml_fail(m, OQG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static int oml_alloc(oqg_oml* m, double** p, size_t n) {
    void* v = nullptr;
    OMLCHK(m, hipMalloc(&v, n * sizeof(double)));
    OMLCHK(m, hipMemsetAsync(v, 0, n * sizeof(double), m->c->st));
    m->allocs.push_back(v);
    *p = static_cast<double*>(v);
    return 0;
}

static OmlArgs oml_args(oqg_oml* m, double tdt, int compute_ent) {
    oqg_ctx* c = m->c;
    OmlArgs a{};
    a.g = c->g;
    a.hm = m->d.hm;
    a.T1_rel = m->d.T1_rel;
    a.dT_12 = m->d.dT_12;
    a.rho_cp = m->d.rho_cp;
    a.d2 = m->d.st2d;
    a.d4 = m->d.st4d;
    a.ycexp = m->d.ycexp;
    a.nval = m->d.nval;
    a.sval = m->d.sval;
    a.fluxn = m->d.fluxn ? 1 : 0;
    a.fluxs = m->d.fluxs ? 1 : 0;
    a.tau_sign = c->d.tau_sign;
    a.p1 = c->pbuf[c->cur];   // layer 1 of the current p (driver.F90:98-101)
    a.taux = c->taux_d;
    a.tauy = c->tauy_d;
    a.fnet = m->fnet;
    a.sst = m->sst;
    a.sstm = m->sstm;
    a.snew = m->snew;
    a.xfo = m->xfo;
    a.part = m->part;
    a.npart = m->npart;
    a.ent1 = c->ent1;
    a.tdt = tdt;
    a.compute_ent = compute_ent;
    return a;
}

static int enqueue_oml_step(oqg_oml* m, double tdt, int compute_ent) {
    const OmlArgs a = oml_args(m, tdt, compute_ent);
    launch_oml_step(a, m->c->st);
    launch_oml_finish(a, m->c->st);
    OMLCHK(m, hipGetLastError());
    return 0;
}

static int enqueue_oml_suppress(oqg_oml* m) {
    launch_oml_suppress(oml_args(m, 0.0, 0), m->c->st);
    OMLCHK(m, hipGetLastError());
    return 0;
}

// step_oml sets the whole entrainment field (omlsubs.F90:100-104): layer 1 from
// the mixed layer, the deeper interfaces zero.  The basin's step then reads the
// layer 1 field the mixed layer writes and no deeper entrainment.
static int oml_own_ent(oqg_oml* m) {
    oqg_ctx* c = m->c;
    if (!c->has_ent2) return 0;
    c->has_ent2 = false;
    static const double zero = 0.0;
    OMLCHK(m, hipMemcpyAsync(c->fcon + F_ENT2, &zero, sizeof zero, hipMemcpyHostToDevice, c->st));
    OMLCHK(m, hipStreamSynchronize(c->st));
    return 0;
}

// T fields between (nxt, nyt) arrays and the pitched layout.
static int oml_copy(oqg_oml* m, double* dst, size_t dpitch, const double* src, size_t spitch, hipMemcpyKind kind) {
    const Geo& g = m->c->g;
    OMLCHK(m, hipMemcpy2DAsync(dst, dpitch, src, spitch, (size_t)(g.nxp - 1) * sizeof(double), (size_t)(g.nyp - 1),
                               kind, m->c->st));
    OMLCHK(m, hipStreamSynchronize(m->c->st));
    return 0;
}

extern "C" {

int oqg_oml_create(oqg_ctx* c, const oqg_oml_desc* d, oqg_oml** out) {
    if (!out) return oml_fail(nullptr, OQG_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!c || !d) return oml_fail(nullptr, OQG_ERR_ARG, "oqg_oml_create: null basin or description");
    if (!(d->hm > 0) || !(d->rho_cp > 0) || d->dT_12 == 0.0)
        return oml_fail(nullptr, OQG_ERR_ARG, "oqg_oml_create: need hm > 0, rho_cp > 0, dT_12 != 0");
    auto m = std::make_unique<oqg_oml>();
    m->c = c;
    m->d = *d;
    const size_t ft = (size_t)c->g.pitch * (c->g.nyp - 1);
    const dim3 grid = oml_step_grid(c->g);
    m->npart = (int)(grid.x * grid.y);
    int rc = 0;
    for (double** p : {&m->sst, &m->sstm, &m->snew, &m->xfo, &m->fnet})
        if (!rc) rc = oml_alloc(m.get(), p, ft);
    if (!rc) rc = oml_alloc(m.get(), &m->part, (size_t)m->npart);
    if (!rc) rc = wind_buffers(c) ? OQG_ERR_HIP : 0;   // the stress the Ekman velocities come from
    if (!rc && hipStreamSynchronize(c->st) != hipSuccess) rc = OQG_ERR_HIP;
    if (rc) {
        g_create_err = m->err.empty() ? c->err : m->err;
        oqg_oml_destroy(m.release());
        return rc;
    }
    *out = m.release();
    return 0;
}

void oqg_oml_destroy(oqg_oml* m) {
    if (!m) return;
    if (m->c && m->c->st) hipStreamSynchronize(m->c->st);
    for (auto& kv : m->graphs) hipGraphExecDestroy(kv.second);
    for (void* p : m->allocs) hipFree(p);
    delete m;
}

const char* oqg_oml_last_error(const oqg_oml* m) { return m ? m->err.c_str() : g_create_err.c_str(); }

int oqg_oml_set_state(oqg_oml* m, const double* sst, const double* sstm, int kind) {
    if (!m) return OQG_ERR_ARG;
    const size_t pb = (size_t)m->c->g.pitch * sizeof(double), w = (size_t)(m->c->g.nxp - 1) * sizeof(double);
    const hipMemcpyKind k = kind == OQG_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (sst)
        if (int r = oml_copy(m, m->sst, pb, sst, w, k)) return r;
    if (sstm)
        if (int r = oml_copy(m, m->sstm, pb, sstm, w, k)) return r;
    return 0;
}

int oqg_oml_get_state(oqg_oml* m, double* sst, double* sstm, int kind) {
    if (!m) return OQG_ERR_ARG;
    const size_t pb = (size_t)m->c->g.pitch * sizeof(double), w = (size_t)(m->c->g.nxp - 1) * sizeof(double);
    const hipMemcpyKind k = kind == OQG_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (sst)
        if (int r = oml_copy(m, sst, w, m->sst, pb, k)) return r;
    if (sstm)
        if (int r = oml_copy(m, sstm, w, m->sstm, pb, k)) return r;
    return 0;
}

int oqg_oml_set_fnet(oqg_oml* m, const double* fnet, int kind) {
    if (!m || !fnet) return m ? oml_fail(m, OQG_ERR_ARG, "oqg_oml_set_fnet: null field") : OQG_ERR_ARG;
    const size_t pb = (size_t)m->c->g.pitch * sizeof(double), w = (size_t)(m->c->g.nxp - 1) * sizeof(double);
    return oml_copy(m, m->fnet, pb, fnet, w, kind == OQG_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice);
}

int oqg_oml_step(oqg_oml* m, double tdto, int compute_ent) {
    if (!m) return OQG_ERR_ARG;
    if (compute_ent)
        if (int r = oml_own_ent(m)) return r;
    return enqueue_oml_step(m, tdto, compute_ent);
}

int oqg_oml_suppress(oqg_oml* m) {
    if (!m) return OQG_ERR_ARG;
    return enqueue_oml_suppress(m);
}

int oqg_run_ocean(oqg_oml* m, double tdto, int nsteps, int step0, int supp, int compute_ent) {
    if (!m || nsteps < 0) return OQG_ERR_ARG;
    oqg_ctx* c = m->c;
    if (compute_ent)
        if (int r = oml_own_ent(m)) return r;
    const int P = graph_period(supp);
    auto one = [&](int nt) -> int {
        // step_ocn (driver.F90:89-110), then suppress_ocean (driver.F90:160-171)
        if (int r = enqueue_oml_step(m, tdto, compute_ent)) return r;
        if (int r = enqueue_step(c, tdto)) return oml_fail(m, r, c->err);
        if (supp > 0 && (nt - 1) % supp == 0) {
            if (int r = enqueue_suppress(c)) return oml_fail(m, r, c->err);
            if (int r = enqueue_oml_suppress(m)) return r;
        }
        return 0;
    };
    int nt = step0 + 1;
    const int end = step0 + nsteps;
    while (nt <= end) {
        if ((nt - 1) % P == 0 && end - nt + 1 >= P) {
            const GraphKey key{tdto, P, supp, c->cur, G_RUN, compute_ent};
            auto it = m->graphs.find(key);
            if (it == m->graphs.end()) {
                const int cur0 = c->cur;
                hipGraph_t gr;
                OMLCHK(m, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
                int rc = 0;
                for (int t = 0; t < P && !rc; t++) rc = one(nt + t);
                const hipError_t ce = hipStreamEndCapture(c->st, &gr);
                c->cur = cur0;
                if (rc) {
                    if (ce == hipSuccess) hipGraphDestroy(gr);
                    return rc;
                }
                OMLCHK(m, ce);
                hipGraphExec_t ex;
                const hipError_t ie = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
                hipGraphDestroy(gr);
                OMLCHK(m, ie);
                it = m->graphs.emplace(key, ex).first;
            }
            OMLCHK(m, hipGraphLaunch(it->second, c->st));
            nt += P;   // P is even: the time-level parity is unchanged
        } else {
            if (int r = one(nt)) return r;
            nt++;
        }
    }
    return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ coupling
// The ocean-atmosphere exchange of dynamic_step (driver.F90:36-87) between an
// ocean basin with its mixed layer and an atmosphere basin, whose atmospheric
// mixed layer the coupler holds.  Everything is enqueued on the ocean basin's
// stream (the atmosphere's work joins it through events).

struct oqg_cpl {
    oqg_oml* om = nullptr;   // the ocean mixed layer; its basin is om->c
    oqg_ctx* ca = nullptr;   // the atmosphere basin
    oqg_cpl_desc d{};
    int ndxr = 0, nxtoar = 0, nytoar = 0, nxpoar = 0, nypoar = 0, nxaooc = 0, nyaooc = 0, nx1 = 0, ny1 = 0;
    double *ast = nullptr, *astm = nullptr, *hmixa = nullptr, *hmixam = nullptr, *anew = nullptr, *hnew = nullptr;
    double *xfa = nullptr, *fnetat = nullptr, *xc1ast = nullptr, *dtop = nullptr;
    double *scr = nullptr, *coarse = nullptr, *st = nullptr, *ocflux = nullptr;
    std::vector<void*> allocs;
    std::map<std::array<double, 10>, hipGraphExec_t> graphs;   // oqg_cpl_run
    hipEvent_t ev = nullptr;
    // the ocean's step runs on a side stream beside the atmosphere's steps, from
    // the heat-flux forcing of an ocean step to the next ocean step's stress
    hipStream_t side = nullptr;
    int dev = 0;   // the device of the side stream
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool pending = false;   // ocean work forked and not yet joined
    std::string err;
};

static int cpl_fail(oqg_cpl* p, int code, const std::string& msg) {
    if (p) p->err = msg;
    else g_create_err = msg;
    return code;
}

#define CPLCHK(p, expr)                                                                          \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return cpl_fail(p, OQG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static int cpl_alloc(oqg_cpl* p, double** out, size_t n) {
    void* v = nullptr;
    CPLCHK(p, hipMalloc(&v, n * sizeof(double)));
    CPLCHK(p, hipMemsetAsync(v, 0, n * sizeof(double), p->om->c->st));
    p->allocs.push_back(v);
    *out = static_cast<double*>(v);
    return 0;
}

// The bicubic weights of bcuini (bicubic.F90:97-146): for fine point r = ip +
// jp (ndxr+1) of a cell and corner datum c (values, then x, y and cross
// derivatives, each at corners (0,0), (1,0), (0,1), (1,1)), st[c][r] is the sum
// over the monomials s^i t^j of their value at (ip, jp)/ndxr times the
// datum's coefficient, which is a product of cubic Hermite basis coefficients.
static void bicubic_weights(int n, std::vector<double>& st) {
    static const double herm[4][4] = {{1, 0, -3, 2}, {0, 0, 3, -2}, {0, 1, -2, 1}, {0, 0, -1, 1}};
    const int nr = (n + 1) * (n + 1);
    double coef[16][16];   // [monomial][datum]
    for (int c = 0; c < 16; c++) {
        const int kind = c / 4, corner = c % 4;
        const int sx = (kind == 1 || kind == 3) ? 2 : 0, sy = (kind == 2 || kind == 3) ? 2 : 0;
        const double* bx = herm[sx + (corner & 1)];
        const double* by = herm[sy + (corner >> 1)];
        for (int j = 0; j < 4; j++)
            for (int i = 0; i < 4; i++) coef[i + 4 * j][c] = bx[i] * by[j];
    }
    st.assign((size_t)16 * nr, 0.0);
    for (int jp = 0; jp <= n; jp++) {
        const double tt = (double)jp / (double)n;
        const double tp[4] = {1.0, tt, tt * tt, tt * tt * tt};
        for (int ip = 0; ip <= n; ip++) {
            const double ss = (double)ip / (double)n;
            const double sp[4] = {1.0, ss, ss * ss, ss * ss * ss};
            const int r = ip + jp * (n + 1);
            for (int c = 0; c < 16; c++) {
                double s = 0.0;
                for (int j = 0; j < 4; j++)
                    for (int i = 0; i < 4; i++) s += (sp[i] * tp[j]) * coef[i + 4 * j][c];
                st[(size_t)c * nr + r] = s;
            }
        }
    }
}

static CplArgs cpl_args(oqg_cpl* p, double tdta, int compute_ent) {
    oqg_ctx *co = p->om->c, *ca = p->ca;
    const oqg_cpl_desc& d = p->d;
    CplArgs a{};
    a.go = co->g;
    a.ga = ca->g;
    a.ndxr = p->ndxr;
    a.nxtoar = p->nxtoar;
    a.nytoar = p->nytoar;
    a.nxpoar = p->nxpoar;
    a.nypoar = p->nypoar;
    a.nxaooc = p->nxaooc;
    a.nyaooc = p->nyaooc;
    a.nx1 = p->nx1;
    a.ny1 = p->ny1;
    a.xo0 = d.xo0;
    a.yo0 = d.yo0;
    a.xa0 = d.xa0;
    a.ya0 = d.ya0;
    a.hmo = d.hmo;
    a.hma = d.hma;
    a.raoro = d.rho_a / d.rho_o;
    a.cdat = d.cdat;
    a.bcco_o = co->d.bcco;
    a.bcco_a = ca->d.bcco;
    a.tau_udiff = d.tau_udiff ? 1 : 0;
    a.tau_sign_a = ca->d.tau_sign;
    a.D0up = d.D0up;
    a.Dmdown = d.Dmdown;
    a.Dup0 = d.Dup0;
    a.xlamda = d.xlamda;
    a.fspco = d.fspco;
    a.Aup00 = d.Aup00;
    a.Adown10 = d.Adown10;
    a.Adown11 = d.Adown11;
    a.hmamin = d.hmamin;
    a.hmadmp = d.hmadmp;
    a.xcexp = d.xcexp;
    a.T1_rel = d.T1_rel;
    a.dT_12 = d.dT_12;
    a.rho_cp = d.rho_cp;
    a.at2d = d.at2d;
    a.at4d = d.at4d;
    a.ahmd = d.ahmd;
    for (int k = 0; k < MAXL; k++) {
        a.aface[k] = d.aface[k];
        a.gpa[k] = ca->d.gp[k];
    }
    a.bface = d.bface;
    a.dface = d.dface;
    a.pmo = co->pbuf[1 - co->cur];
    a.pma = ca->pbuf[1 - ca->cur];
    a.pa1 = ca->pbuf[ca->cur];
    a.tauxo = co->taux_d;
    a.tauyo = co->tauy_d;
    a.tauxa = ca->taux_d;
    a.tauya = ca->tauy_d;
    a.scr = p->scr;
    a.coarse = p->coarse;
    a.st = p->st;
    a.sstm = p->om->sstm;
    a.fnetoc = p->om->fnet;
    a.ocflux = p->ocflux;
    a.ast = p->ast;
    a.astm = p->astm;
    a.hmixa = p->hmixa;
    a.hmixam = p->hmixam;
    a.anew = p->anew;
    a.hnew = p->hnew;
    a.xfa = p->xfa;
    a.fnetat = p->fnetat;
    a.xc1ast = p->xc1ast;
    a.dtop = p->dtop;
    a.enta = ca->ent1;
    a.tdta = tdta;
    a.compute_ent = compute_ent;
    return a;
}

// ekman_from_tau (ekman.F90:53-83) of a basin's resident stress.
static int enqueue_ekman(oqg_ctx* c) {
    EkmanArgs a{};
    a.g = c->g;
    a.taux = c->taux_d;
    a.tauy = c->tauy_d;
    a.wekp = c->wekp;
    a.fcon = c->fcon;
    launch_ekman(a, c->st);
    HIPCHK(c, hipGetLastError());
    return 0;
}

// The atmosphere's work goes on the ocean's stream for the duration of a call.
struct OnStream {
    oqg_ctx* c;
    hipStream_t saved;
    OnStream(oqg_ctx* c_, hipStream_t s) : c(c_), saved(c_->st) { c->st = s; }
    ~OnStream() { c->st = saved; }
};

// Order the ocean stream after the atmosphere's pending work (begin) and the
// atmosphere's stream after what the call enqueued (end).
static int cpl_join(oqg_cpl* p, bool begin) {
    hipStream_t so = p->om->c->st, sa = p->ca->st;
    if (so == sa) return 0;
    CPLCHK(p, hipEventRecord(p->ev, begin ? sa : so));
    CPLCHK(p, hipStreamWaitEvent(begin ? so : sa, p->ev, 0));
    return 0;
}

static int cpl_own_ent(oqg_cpl* p, int ent_o, int ent_a) {
    if (ent_o)
        if (int r = oml_own_ent(p->om)) return cpl_fail(p, r, p->om->err);
    oqg_ctx* ca = p->ca;
    if (ent_a && ca->has_ent2) {
        ca->has_ent2 = false;
        static const double zero = 0.0;
        CPLCHK(p, hipMemcpyAsync(ca->fcon + F_ENT2, &zero, sizeof zero, hipMemcpyHostToDevice, ca->st));
        CPLCHK(p, hipStreamSynchronize(ca->st));
    }
    return 0;
}

static int enqueue_xforc(oqg_cpl* p) {
    launch_xforc(cpl_args(p, 0.0, 0), p->om->c->st);
    CPLCHK(p, hipGetLastError());
    if (int r = enqueue_ekman(p->om->c)) return cpl_fail(p, r, p->om->c->err);
    if (int r = enqueue_ekman(p->ca)) return cpl_fail(p, r, p->ca->err);
    return 0;
}

// Wait, on the main (ocean basin's) stream, for the forked ocean work.
static int cpl_join_side(oqg_cpl* p, hipStream_t main) {
    if (!p->pending) return 0;
    CPLCHK(p, hipStreamWaitEvent(main, p->ev_join, 0));
    p->pending = false;
    return 0;
}

// One atmospheric step nt of dynamic_step (driver.F90:36-87) on stream `main`.
// On an ocean step the exchange (stress, heat flux) runs on `main`, then the
// ocean's step (mixed layer, QG, its suppression) forks onto the side stream:
// nothing the atmosphere does before the next ocean step reads or writes what
// the ocean's step does, and the next ocean step's exchange joins it first.
static int enqueue_dynamic(oqg_cpl* p, hipStream_t main, int nt, double tdto, double tdta, int nstr, int osupp,
                           int asupp, int ent_o, int ent_a) {
    oqg_ctx *co = p->om->c, *ca = p->ca;
    const bool ocean_step = (nt - 1) % nstr == 0;
    if (ocean_step) {
        if (int r = cpl_join_side(p, main)) return r;
        if (int r = enqueue_xforc(p)) return r;
        // step_atm's compute_forcing (driver.F90:130-133)
        launch_forcing(cpl_args(p, tdta, ent_a), main);
        CPLCHK(p, hipGetLastError());
        // step_ocn (driver.F90:89-110) and suppress_ocean (:160-171), forked
        CPLCHK(p, hipEventRecord(p->ev_fork, main));
        CPLCHK(p, hipStreamWaitEvent(p->side, p->ev_fork, 0));
        {
            OnStream on(co, p->side);
            if (int r = enqueue_oml_step(p->om, tdto, ent_o)) return cpl_fail(p, r, p->om->err);
            if (int r = enqueue_step(co, tdto)) return cpl_fail(p, r, co->err);
            if (osupp > 0 && (nt - 1) % osupp == 0) {
                if (int r = enqueue_suppress(co)) return cpl_fail(p, r, co->err);
                if (int r = enqueue_oml_suppress(p->om)) return cpl_fail(p, r, p->om->err);
            }
        }
        CPLCHK(p, hipEventRecord(p->ev_join, p->side));
        p->pending = true;
    }
    // the rest of step_atm (driver.F90:134-146)
    launch_aml_step(cpl_args(p, tdta, ent_a), main);
    CPLCHK(p, hipGetLastError());
    if (int r = enqueue_step(ca, tdta)) return cpl_fail(p, r, ca->err);
    // suppress_atmos (driver.F90:149-158)
    if (asupp > 0 && (nt - 1) % asupp == 0) {
        if (int r = enqueue_suppress(ca)) return cpl_fail(p, r, ca->err);
        launch_aml_suppress(cpl_args(p, tdta, 0), main);
        CPLCHK(p, hipGetLastError());
    }
    return 0;
}

static int lcm(int a, int b) {
    int x = a, y = b;
    while (y) {
        const int t = x % y;
        x = y;
        y = t;
    }
    return a / x * b;
}

static int cpl_copy(oqg_cpl* p, double* dst, size_t dpitch, const double* src, size_t spitch, size_t w, size_t rows,
                    hipMemcpyKind kind) {
    hipStream_t s = p->om->c->st;
    CPLCHK(p, hipMemcpy2DAsync(dst, dpitch, src, spitch, w, rows, kind, s));
    CPLCHK(p, hipStreamSynchronize(s));
    return 0;
}

static int cpl_run_steps(oqg_cpl* p, int nsteps, int step0, double tdto, double tdta, int nstr, int osupp, int asupp,
                         int flags, int ent_o, int ent_a);

extern "C" {

int oqg_cpl_create(oqg_oml* om, oqg_ctx* ca, const oqg_cpl_desc* d, oqg_cpl** out) {
    if (!out) return cpl_fail(nullptr, OQG_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!om || !ca || !d) return cpl_fail(nullptr, OQG_ERR_ARG, "oqg_cpl_create: null ocean mixed layer, basin or description");
    oqg_ctx* co = om->c;
    if (co->g.cyclic || !ca->g.cyclic)
        return cpl_fail(nullptr, OQG_ERR_ARG, "oqg_cpl_create: the ocean must be a box and the atmosphere a channel");
    if (ca->g.nl < 2) return cpl_fail(nullptr, OQG_ERR_ARG, "oqg_cpl_create: the atmosphere needs >= 2 layers");
    if (!(d->hma > 0) || !(d->hmo > 0) || !(d->rho_cp > 0) || !(d->rho_o > 0) || !(d->rho_a > 0) || d->dT_12 == 0.0)
        return cpl_fail(nullptr, OQG_ERR_ARG, "oqg_cpl_create: need hmo, hma, rho_o, rho_a, rho_cp > 0 and dT_12 != 0");
    auto p = std::make_unique<oqg_cpl>();
    p->om = om;
    p->ca = ca;
    p->d = *d;
    p->d.xc1ast = nullptr;
    p->d.dtop = nullptr;
    // load_grid (grid.F90:37-58)
    const int nxto = co->g.nxp - 1, nyto = co->g.nyp - 1, nxta = ca->g.nxp - 1, nyta = ca->g.nyp - 1;
    p->ndxr = (int)std::lround(ca->g.dx / co->g.dx);
    if (p->ndxr < 1 || nxto % p->ndxr || nyto % p->ndxr)
        return cpl_fail(nullptr, OQG_ERR_ARG, "oqg_cpl_create: the ocean's cells must tile whole atmosphere cells "
                                              "(nxt and nyt of the ocean multiples of dxa/dxo)");
    p->nxtoar = nxto / p->ndxr;
    p->nytoar = nyto / p->ndxr;
    p->nxpoar = p->nxtoar + 1;
    p->nypoar = p->nytoar + 1;
    p->nxaooc = nxto / p->ndxr;
    p->nyaooc = nyto / p->ndxr;
    p->nx1 = 1;
    p->ny1 = 1 + (nyta - p->nyaooc) / 2;
    if (p->nxaooc > nxta || p->ny1 < 1 || p->nx1 + p->nxpoar - 1 > ca->g.nxp || p->ny1 + p->nypoar - 1 > ca->g.nyp)
        return cpl_fail(nullptr, OQG_ERR_ARG, "oqg_cpl_create: the ocean does not fit inside the atmosphere");
    int rc = 0;
    const size_t ta = (size_t)ca->g.pitch * (ca->g.nyp - 1);
    for (double** f : {&p->ast, &p->astm, &p->hmixa, &p->hmixam, &p->anew, &p->hnew, &p->xfa, &p->fnetat, &p->xc1ast})
        if (!rc) rc = cpl_alloc(p.get(), f, ta);
    if (!rc) rc = cpl_alloc(p.get(), &p->dtop, (size_t)ca->g.fs);
    if (!rc) rc = cpl_alloc(p.get(), &p->scr, 4 * (size_t)ca->g.fs);
    if (!rc) rc = cpl_alloc(p.get(), &p->ocflux, (size_t)co->g.pitch * (co->g.nyp - 1));
    if (!rc) rc = cpl_alloc(p.get(), &p->coarse, 8 * (size_t)p->nxpoar * p->nypoar);
    std::vector<double> st;
    bicubic_weights(p->ndxr, st);
    if (!rc) rc = cpl_alloc(p.get(), &p->st, st.size());
    if (!rc) rc = wind_buffers(co) ? OQG_ERR_HIP : 0;
    if (!rc) rc = wind_buffers(ca) ? OQG_ERR_HIP : 0;
    if (!rc && hipEventCreateWithFlags(&p->ev, hipEventDisableTiming) != hipSuccess) rc = OQG_ERR_HIP;
    if (!rc && hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming) != hipSuccess) rc = OQG_ERR_HIP;
    if (!rc && hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming) != hipSuccess) rc = OQG_ERR_HIP;
    p->dev = co->dev;
    if (!rc && pool_take(&p->side) != hipSuccess) rc = OQG_ERR_HIP;
    hipStream_t s = co->st;
    if (!rc) rc = hipMemcpyAsync(p->st, st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice, s) ? OQG_ERR_HIP : 0;
    const size_t wa = (size_t)nxta * sizeof(double), pa = (size_t)ca->g.pitch * sizeof(double);
    if (!rc && d->xc1ast)
        rc = hipMemcpy2DAsync(p->xc1ast, pa, d->xc1ast, wa, wa, nyta, hipMemcpyHostToDevice, s) ? OQG_ERR_HIP : 0;
    if (!rc && d->dtop)
        rc = hipMemcpy2DAsync(p->dtop, pa, d->dtop, wa + sizeof(double), wa + sizeof(double), ca->g.nyp,
                              hipMemcpyHostToDevice, s) ? OQG_ERR_HIP : 0;
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = OQG_ERR_HIP;
    if (rc) {
        g_create_err = p->err.empty() ? std::string("oqg_cpl_create: device allocation or copy failed") : p->err;
        oqg_cpl_destroy(p.release());
        return rc;
    }
    *out = p.release();
    return 0;
}

void oqg_cpl_destroy(oqg_cpl* p) {
    if (!p) return;
    if (p->om && p->om->c && p->om->c->st) hipStreamSynchronize(p->om->c->st);
    for (auto& kv : p->graphs) hipGraphExecDestroy(kv.second);
    if (p->side) hipStreamSynchronize(p->side);
    for (void* v : p->allocs) hipFree(v);
    if (p->ev) hipEventDestroy(p->ev);
    if (p->ev_fork) hipEventDestroy(p->ev_fork);
    if (p->ev_join) hipEventDestroy(p->ev_join);
    pool_give(p->dev, p->side);
    delete p;
}

const char* oqg_cpl_last_error(const oqg_cpl* p) { return p ? p->err.c_str() : g_create_err.c_str(); }

int oqg_cpl_set_aml_state(oqg_cpl* p, const double* ast, const double* astm, const double* hmixa,
                          const double* hmixam, int kind) {
    if (!p) return OQG_ERR_ARG;
    const Geo& g = p->ca->g;
    const size_t w = (size_t)(g.nxp - 1) * sizeof(double), pb = (size_t)g.pitch * sizeof(double);
    const hipMemcpyKind k = kind == OQG_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const double* src[4] = {ast, astm, hmixa, hmixam};
    double* dst[4] = {p->ast, p->astm, p->hmixa, p->hmixam};
    for (int f = 0; f < 4; f++)
        if (src[f])
            if (int r = cpl_copy(p, dst[f], pb, src[f], w, w, g.nyp - 1, k)) return r;
    return 0;
}

int oqg_cpl_get_aml_state(oqg_cpl* p, double* ast, double* astm, double* hmixa, double* hmixam, int kind) {
    if (!p) return OQG_ERR_ARG;
    const Geo& g = p->ca->g;
    const size_t w = (size_t)(g.nxp - 1) * sizeof(double), pb = (size_t)g.pitch * sizeof(double);
    const hipMemcpyKind k = kind == OQG_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (int r = cpl_join(p, true)) return r;
    double* dst[4] = {ast, astm, hmixa, hmixam};
    const double* src[4] = {p->ast, p->astm, p->hmixa, p->hmixam};
    for (int f = 0; f < 4; f++)
        if (dst[f])
            if (int r = cpl_copy(p, dst[f], w, src[f], pb, w, g.nyp - 1, k)) return r;
    return 0;
}

int oqg_cpl_get_fnet(oqg_cpl* p, double* fnetoc, double* fnetat, int kind) {
    if (!p) return OQG_ERR_ARG;
    const hipMemcpyKind k = kind == OQG_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (int r = cpl_join(p, true)) return r;
    if (fnetoc) {
        const Geo& g = p->om->c->g;
        const size_t w = (size_t)(g.nxp - 1) * sizeof(double), pb = (size_t)g.pitch * sizeof(double);
        if (int r = cpl_copy(p, fnetoc, w, p->om->fnet, pb, w, g.nyp - 1, k)) return r;
    }
    if (fnetat) {
        const Geo& g = p->ca->g;
        const size_t w = (size_t)(g.nxp - 1) * sizeof(double), pb = (size_t)g.pitch * sizeof(double);
        if (int r = cpl_copy(p, fnetat, w, p->fnetat, pb, w, g.nyp - 1, k)) return r;
    }
    return 0;
}

int oqg_cpl_xforc(oqg_cpl* p) {
    if (!p) return OQG_ERR_ARG;
    if (int r = cpl_join(p, true)) return r;
    {
        OnStream on(p->ca, p->om->c->st);
        if (int r = enqueue_xforc(p)) return r;
    }
    return cpl_join(p, false);
}

int oqg_cpl_forcing(oqg_cpl* p) {
    if (!p) return OQG_ERR_ARG;
    if (int r = cpl_join(p, true)) return r;
    launch_forcing(cpl_args(p, 0.0, 0), p->om->c->st);
    CPLCHK(p, hipGetLastError());
    return cpl_join(p, false);
}

int oqg_cpl_aml_step(oqg_cpl* p, double tdta, int compute_ent) {
    if (!p) return OQG_ERR_ARG;
    if (int r = cpl_own_ent(p, 0, compute_ent)) return r;
    if (int r = cpl_join(p, true)) return r;
    launch_aml_step(cpl_args(p, tdta, compute_ent), p->om->c->st);
    CPLCHK(p, hipGetLastError());
    return cpl_join(p, false);
}

int oqg_cpl_run(oqg_cpl* p, int nsteps, int step0, double tdto, double tdta, int nstr, int osupp, int asupp,
                int flags) {
    if (!p || nsteps < 0 || nstr < 1 || osupp < 0 || asupp < 0) return p ? cpl_fail(p, OQG_ERR_ARG, "oqg_cpl_run: bad arguments") : OQG_ERR_ARG;
    const int ent_o = (flags & OQG_CPL_OCEAN_ENT) ? 1 : 0, ent_a = (flags & OQG_CPL_ATMOS_ENT) ? 1 : 0;
    if (int r = cpl_own_ent(p, ent_o, ent_a)) return r;
    if (int r = cpl_join(p, true)) return r;
    if (int r = cpl_run_steps(p, nsteps, step0, tdto, tdta, nstr, osupp, asupp, flags, ent_o, ent_a)) return r;
    return cpl_join(p, false);
}

}  // extern "C"

static int cpl_run_steps(oqg_cpl* p, int nsteps, int step0, double tdto, double tdta, int nstr, int osupp, int asupp,
                         int flags, int ent_o, int ent_a) {
    oqg_ctx *co = p->om->c, *ca = p->ca;
    hipStream_t main = co->st;
    OnStream on(ca, main);
    // the cadence repeats every lcm(nstr, osupp, asupp) steps; a graph also
    // needs both basins back at their starting time-level parity
    int P = nstr;
    if (osupp > 0) P = lcm(P, osupp);
    if (asupp > 0) P = lcm(P, asupp);
    if (P % 2) P *= 2;
    if ((P / nstr) % 2) P *= 2;
    int nt = step0 + 1;
    const int end = step0 + nsteps;
    while (nt <= end) {
        if ((nt - 1) % P == 0 && end - nt + 1 >= P && P <= 4096) {
            const std::array<double, 10> key{tdto, tdta, (double)P, (double)nstr, (double)osupp, (double)asupp,
                                             (double)co->cur, (double)ca->cur, (double)flags, 0.0};
            auto it = p->graphs.find(key);
            if (it == p->graphs.end()) {
                const int c0 = co->cur, a0 = ca->cur;
                hipGraph_t gr;
                // work forked before the capture is joined outside it
                if (int r = cpl_join_side(p, main)) return r;
                CPLCHK(p, hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal));
                int rc = 0;
                for (int t = 0; t < P && !rc; t++)
                    rc = enqueue_dynamic(p, main, nt + t, tdto, tdta, nstr, osupp, asupp, ent_o, ent_a);
                if (!rc) rc = cpl_join_side(p, main);   // every forked stream rejoins before the capture ends
                p->pending = false;
                const hipError_t ce = hipStreamEndCapture(main, &gr);
                co->cur = c0;
                ca->cur = a0;
                if (rc) {
                    if (ce == hipSuccess) hipGraphDestroy(gr);
                    return rc;
                }
                CPLCHK(p, ce);
                hipGraphExec_t ex;
                const hipError_t ie = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
                hipGraphDestroy(gr);
                CPLCHK(p, ie);
                it = p->graphs.emplace(key, ex).first;
            }
            if (int r = cpl_join_side(p, main)) return r;
            CPLCHK(p, hipGraphLaunch(it->second, main));
            nt += P;
        } else {
            if (int r = enqueue_dynamic(p, main, nt, tdto, tdta, nstr, osupp, asupp, ent_o, ent_a)) return r;
            nt++;
        }
    }
    return cpl_join_side(p, main);
}

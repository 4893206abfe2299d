// oqg_internal.h -- shared definitions of the HIP QG step library (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>
#include <vector>

#include "../../include/oqg.h"

namespace oqg {

constexpr int MAXL = OQG_MAX_LAYERS;
constexpr int WAVE = 64;
constexpr int NXCD = 8;              // accelerator dies (XCDs) of an MI355X, one L2 each
constexpr int MAX_RADIX_STAGES = 16;
constexpr int FWD_MAX_RADIX = 16;    // radix-16/8 butterflies: fewest LDS passes
constexpr int INV_MAX_RADIX = 4;     // the inverse kernel's register budget keeps radix 4
constexpr int MAX_FFT = 4096;        // longest x transform (complex points)
constexpr int VORT_TX = 64;          // vorticity tile (x)
constexpr int VORT_TY = 16;          // vorticity tile (y), the most rows; a basin takes 4, 8 or 16 (vort_rows)
constexpr int MAX_YROWS_PER_LANE = 16;  // y-solver: nyp-2 <= 64*16
// Columns per y-solver workgroup (one wavefront each).  Two per workgroup
// (768 workgroups at 513^2 instead of 192) measured slower: 12.1 against
// 10.7 us.
constexpr int YS_CT = 8;

// Geometry and physics shared by every kernel (by value in kernel arguments).
struct Geo {
    int nxp, nyp, nl, pitch;
    int cyclic, kt;          // kt: 0-based k_topo
    long long fs;            // layer stride in doubles = pitch*nyp
    double dx, dy, fnot, beta;
    double fx, fy;           // 1/dx^2, 1/dy^2 as in dP2dx2_bc / dP2dy2_bc
    double cxl, cxr, cyl, cyr;   // mixed-BC coefficients (numerics.F90:439-440,461-462)
    double jfac;             // Arakawa factor (numerics.F90:361)
};

// Spectral (x-transformed) layout: spec[m][j][k], row-major, rows of kp doubles,
// nrow = nyi+1 rows per mode (the last one only padding).
struct Spec {
    int nin;        // transform input points per row (box nxp-2, cyclic nxp-1)
    int L;          // complex FFT length (box 2(nxp-1), cyclic nxp-1)
    int ncol;       // real tridiagonal columns per mode
    int nyi;        // interior rows nyp-2
    int nrow;       // nyi + 1
    int kp;         // spectral row pitch (doubles), multiple of 16
    int R;          // rows per lane in the wavefront y-solver
    int nypad;      // 64*R
    int nstages;    // x transform plan of the forward (spectral) kernels
    int radix[MAX_RADIX_STAGES];
    int instages;   // plan of the inverse kernels, radix <= INV_MAX_RADIX
    int iradix[MAX_RADIX_STAGES];
    int prime;      // the plans hold a radix 7, 11 or 13 (block kernels k_fwd_pr / k_inv_pr only)
    double a;       // 1/dy^2 (inhomog.F90:127)
    double ftnorm;  // 1/N (inhomog.F90:124)
};

struct Mix {
    double ctl2m[MAXL * MAXL];   // [k*MAXL + m]
    double ctm2l[MAXL * MAXL];   // [m*MAXL + k]
    double amat[MAXL * MAXL];    // [k*MAXL + l]
    double rdm2[MAXL];
};

// Per-layer forcing and friction coefficients (vorticity.F90:47-74).
struct Phys {
    double c2[MAXL], c4[MAXL];   // ah2/f0, ah4/f0
    double fohfac[MAXL];         // f0/h
    double cent1, cwek, cent2, cdrag;
    double ah2[MAXL], ah4[MAXL];
    double h[MAXL], gp[MAXL];
    double cd_bdr;               // f0*delek/(2|f0|)  (vorticity.F90:134)
    double qbfac;                // bcco*dxm2/(0.5*bcco+1)/f0 (vorsubs.F90:76)
    double rdy, rdx2;            // 1/dy, 1/(2 dx): the per-step correction code multiplies
    int dz_sign, tau_sign;
};

// Box mass-constraint precompute (homog.F90:214-278).
struct BoxHom {
    double cdhoc[MAXL * MAXL];   // [(k-1)*MAXL + (m-1)]  (nl-1)^2
    double LU[MAXL * MAXL];
    int ipiv[MAXL];
};

// Cyclic precompute (homog.F90:128-212).
struct CycHom {
    double hc1s[MAXL], hc2s[MAXL], hc1n[MAXL], hc2n[MAXL];
    double hbsi;
    double c0;      // int_P_dx of a transformed row per unit U_0: dx * ftnorm * N
};

// Constraint state block (one per time-level parity).
//   [0, MAXL)  dpi   [MAXL, 2MAXL) dpip  then cs, cn, csp, cnp
constexpr int C_DPI = 0, C_DPIP = MAXL, C_CS = 2 * MAXL, C_CN = 3 * MAXL, C_CSP = 4 * MAXL,
              C_CNP = 5 * MAXL, C_LEN = 6 * MAXL;
// Cyclic per-step boundary integrals of the momentum constraints, as the
// correction code reads them (built in LDS by k_inv from the tiles' partials):
//   ajis, ajin, ap3s, ap3n, ap5s, ap5n, int_Be_s, int_Be_n (per layer), then
//   bdrins, bdrinn, txis, txin
constexpr int A_AJIS = 0, A_AJIN = MAXL, A_AP3S = 2 * MAXL, A_AP3N = 3 * MAXL, A_AP5S = 4 * MAXL,
              A_AP5N = 5 * MAXL, A_BES = 6 * MAXL, A_BEN = 7 * MAXL, A_BDRS = 8 * MAXL,
              A_BDRN = 8 * MAXL + 1, A_TXIS = 8 * MAXL + 2, A_TXIN = 8 * MAXL + 3, A_LEN = 8 * MAXL + 4;
// Channel: k_vort's first and last tile rows write per-tile partial row sums
// (int_P_dx weights, before the factor dx) of the constraint integrals
// (vorticity.F90:57-62,84-101, numerics.F90:374-386) for the zonal boundary
// at side 0 (south, rows 1-2) and side 1 (north, rows nyp-1, nyp), at
// cpart[((k * 2 + side) * tiles_x + tile) * CYC_SLOTS + quantity]:
//   CQ_D2: Del^2 pm difference across the boundary row pair (ap3)
//   CQ_D4: Del^4 pm difference (ap5)
//   CQ_AJ5, CQ_AJ9: q of the boundary / inner row times dPdx_2 p of the inner row
//   CQ_BD: pm difference (layer k_topo; bottom drag)
//   CQ_E1, CQ_E2: layer-1 entrainment on the boundary row, layer-2 edge row (layer 1)
enum { CQ_D2 = 0, CQ_D4, CQ_AJ5, CQ_AJ9, CQ_BD, CQ_E1, CQ_E2, CQ_N };
constexpr int CYC_SLOTS = 8;
constexpr int CYC_MAX_TX = 17;   // tiles per row of a channel of nxp <= 1025

// Arguments of the vorticity kernel.
struct VortArgs {
    Geo g;
    const Phys* ph;   // DevConsts::ph (kept out of the kernel arguments)
    const double* q;      // current q
    const double* p;      // current p
    const double* pm;     // previous p
    const double* qm;     // previous q (normally qnew's buffer: q_new is written over it)
    double* qnew;         // out: q_new
    const double* ent1;   // (pitch layout) layer-1 entrainment
    const double* wekp;
    double* entpart;      // per-tile partial integrals of ent1 (layer-0 tiles)
    double tdt;
    int tiles_x, tiles_y;
    int ty;               // tile rows (4, 8 or 16), the k_vort instantiation
    int org_x;            // first column of the tiles: 1 in a box, 0 in a channel
    // channel: the constraint integrals' partials (cpart, see CQ_*), the
    // layer-2 edge rows of ent and the resident forcing scalars (F_*)
    double* cpart;
    const double* ent2s;
    const double* ent2n;
    const double* fcon;
};

// Entries of the twiddle table of a length-L transform plan (the stages'
// twiddles, then the radix-3 and radix-5 roots at the two offsets below, then
// those of the radices 7, 11 and 13 that divide L).
__host__ __device__ inline int twiddle_roots3(int L) { return L - 1; }
__host__ __device__ inline int twiddle_roots5(int L) { return L + 1; }
__host__ __device__ inline int twiddle_roots7(int L) { return L + 8; }
__host__ __device__ inline int twiddle_roots11(int L) { return twiddle_roots7(L) + (L % 7 == 0 ? 6 : 0); }
__host__ __device__ inline int twiddle_roots13(int L) { return twiddle_roots11(L) + (L % 11 == 0 ? 10 : 0); }
__host__ __device__ inline int twiddle_table_len(int L) { return twiddle_roots13(L) + (L % 13 == 0 ? 12 : 0); }

// Column pitch (doubles) of the y solver's LDS tile for R rows per lane: room
// for R row groups of stride 65, rounded up to 2 mod 16, which puts the 8
// columns of a transposing access 4 banks apart.
__host__ __device__ constexpr int ys_pitch(int R) { return 65 * R + ((2 - 65 * R) % 16 + 16) % 16; }

// k_vort computes q_new on columns [org_x, lim_x) and rows [1, nyp-1): the
// zonal boundary rows, and in a box the meridional boundary columns, are
// rewritten by ocqbdy (vorsubs.F90:58-114) in the same step before anything
// reads them.
__host__ __device__ inline int vort_org_x(const Geo& g) { return g.cyclic ? 0 : 1; }
__host__ __device__ inline int vort_lim_x(const Geo& g) { return g.cyclic ? g.nxp : g.nxp - 1; }

// Device block holding the precompute read by k_fwd, k_inv and the adjoint
// kernels (kept out of the kernel arguments, which every launch uploads).
struct DevConsts {
    Mix mx;
    BoxHom bh;
    CycHom ch;
    Phys ph;
};

struct FwdArgs {
    Geo g;
    Spec s;
    const DevConsts* dc;
    const double* q;      // q_new (step) or rhs field (raw)
    const double* ddyn;   // may be null
    const double* yprel;
    const double* tw;     // twiddle table of the forward plan (stage order, interleaved)
    double* spec;
    int raw_mode;         // >=0: raw single field solved for this mode (0-based)
    // adjoint mode (k_fwd_adj, box): the input of mode m is
    // w_m = sum_k ctm2l(m,k) lp_k, lp being adj_lp completed by the transpose
    // of ocqbdy from adj_lq's boundary ring; per-block sums of hom_m . w_m over
    // the block's rows (the first and last blocks add rows 0 and nyp-1) go to
    // adj_partial[row pair][m]
    const double* adj_lp;
    const double* adj_lq;
    const double* adj_hom;
    double* adj_partial;
    double adj_qbfac;
};

struct YArgs {
    Spec s;
    double* spec;
    const double* bcol;   // [m][col]
    double* colsum;       // [m][col] or null
    double* gam;          // Thomas variant scratch, same layout as spec
    int mode0, nmodes;    // range of modes solved
};

struct InvArgs {
    Geo g;
    Spec s;
    const Phys* ph;   // DevConsts::ph (kept out of the kernel arguments)
    const DevConsts* dc;    // device copy of the precompute blocks
    const double* spec;
    const double* tw;       // twiddle table of the inverse plan
    const double* twf;      // twiddle table of the forward plan (k_invw)
    const double* colsum;   // box: [m][col]
    const double* colw;     // box: closed-form x-sum weights per column
    const double* entpart;
    int nentpart;
    const double* hom;      // box: (nl-1) pitch fields; cyclic: bt(nyp), bc1(nl-1,nyp), bc2(nl-1,nyp)
    const double* cin;      // constraint state in
    double* cout;           // constraint state out
    const double* cpart;    // cyclic: k_vort's partial boundary integrals (CQ_*)
    int ctiles;             // cyclic: tiles per row in cpart
    const double* fcon;     // cyclic: resident forcing scalars (F_TXIS, F_TXIN)
    const double* yprel;
    const double* ddyn;
    double* pnew;           // p_new (pm slot)
    double* qnew;           // q_new (boundary ring written)
    double* out_raw;        // raw mode output field
    int raw_mode;           // >=0 raw mode
    // adjoint mode (k_inv_adj, box): every mode's inverse transform u_m, the
    // constants s_m from k_fwd_adj's partial sums (adj_npart row pairs) and
    // adj_cinvT, then adj_lq_k += f0 sum_m ctl2m(k,m) (u_m + s_m G_m) on the
    // block's interior points
    const double* adj_G;
    const double* adj_cinvT;
    const double* adj_partial;
    int adj_npart;
    double* adj_lq;
    // channel: adj_G holds H_m^-1 of the unit rows 1 and nyp-2 at fields
    // [side * nl + m]; the adjoint of the new constraint state (cs at
    // [0, MAXL), cn at [MAXL, 2 MAXL)) is accumulated into adj_lc by block 0
    double* adj_lc;
    double tdt;
};

// Channel adjoints: stride of the per-row-pair partial sums of the inversion's
// homogeneous profiles (bt, bc1 of modes 1..nl-1, bc2 of modes 1..nl-1).
constexpr int ADJ_CYC_NPROF = 2 * MAXL;
// the adjoint constraint block: lcs at [0, MAXL), lcn at [MAXL, 2 MAXL)
constexpr int LC_LEN = 2 * MAXL;

// Resident forcing scalars of a channel, read at run time so that a replayed
// graph sees the latest forcing: [F_TXIS] txis, [F_TXIN] txin, [F_ENT2] 1 when
// the layer-2 edge rows of ent hold data.
constexpr int F_TXIS = 0, F_TXIN = 1, F_ENT2 = 2, F_LEN = 4;

struct SuppArgs {
    Geo g;
    double* q;
    const double* qm;
    double* p;
    const double* pm;
    double* c;
};

struct QcompArgs {
    Geo g;
    Mix mx;
    const double* p;
    const double* ddyn;
    const double* yprel;
    double* q;
};

// Adjoint kernels (oqg_adjoint.hip).
struct AdjSuppArgs {
    Geo g;
    double *lq, *lqm, *lp, *lpm;
    double *lc, *lcm;   // channel: adjoint constraint blocks of cs/cn and csp/cnp (LC_LEN), else null
};

struct KeSeedArgs {
    Geo g;
    double bcco;
    double rhoh;        // rho H_1
    const double* p;    // layer 1 of the loaded p
    double* lp;         // layer 1 of its adjoint (written)
};
void launch_ke_seed(const KeSeedArgs& a, hipStream_t st);

struct AdjVortArgs {
    Geo g;
    const Phys* ph;   // DevConsts::ph (kept out of the kernel arguments)
    const double* q;    // forward state before the step: q_n, p_n
    const double* p;
    double* lq_np1;     // in: adjoint of q_{n+1}; out: adjoint of q_{n-1} (ring zeroed)
    double* lq_n;       // adjoint of q_n (+=)
    double* lp_n;       // adjoint of p_n (+=)
    double* lp_nm1;     // out: adjoint of p_{n-1}
    double* lwek;       // adjoint of the Ekman pumping field (+=)
    double tdt;
    // with `slots` set, q_n and p_n are read from snapshot slot
    // slots[*base + t] at offsets qoff, poff instead of q, p
    const double* const* slots;
    const int* base;
    int t;
    long long qoff, poff;
    // channel: the adjoint of the new constraint state (LC_LEN, read), the
    // adjoints of txis and txin (+=, block 0), and the gradients of the
    // boundary integrals ap3, ap5 and the drag integral with respect to the
    // row sums of pm next to each zonal boundary ([d]: row d (south) or
    // nyp-1-d (north)), before the factors ah2, ah4, cd_bdr and dx/dy
    const double* lc;
    double* ltxi;
    double vs3[4], vs5[4], vsb[4], vn3[4], vn5[4], vnb[4];
};

void launch_adj_supp(const AdjSuppArgs& a, hipStream_t st);

// Wind-stress forcing producer (oqg_forcing.hip).
struct EkmanArgs {
    Geo g;
    double* taux;         // kinematic stress at p points (pitch layout); the adjoint writes its result here
    double* tauy;
    double* wekp;         // out (k_ekman)
    double* fcon;         // out, channel: txis, txin (k_ekman_txi)
    const double* lwek;   // in (k_ekman_adj): adjoint of wekp
    const double* ltxi;   // in (k_ekman_adj, channel): adjoints of txis, txin
};
void launch_ekman(const EkmanArgs& a, hipStream_t st);
void launch_ekman_adj(const EkmanArgs& a, hipStream_t st);

// Diagnostics (oqg_diag.hip): per-row partials, MAXL slots per quantity.
enum { DG_KE = 0, DG_P, DG_Q, DG_ETA, DG_ETA2, DG_PMIN, DG_PMAX, DG_QMIN, DG_QMAX, DG_UMAX, DG_VMAX, DG_P0, DG_NQ };
constexpr int DG_BAD = DG_NQ * MAXL;    // count of non-finite p and q values
constexpr int DG_LEN = DG_NQ * MAXL + 1;
struct DiagArgs {
    Geo g;
    double gp[MAXL];
    int dz_sign;
    double bcco;
    const double* q;   // current time level
    const double* p;
    double* partial;   // [nyp][DG_LEN]
};
void launch_diag(const DiagArgs& a, hipStream_t st);

// The energetics of qg_diagno (k_energy_d2, k_energy): per-row partials
// [nyp][EN_LEN]: per layer (MAXL slots each) the row sums of ugeos del2(ug),
// vgeos del2(vg), ugeos del4(ug), vgeos del4(vg) with int_P_dA's weights and
// the row's int_P_dx weights of ugeos (times dx on the host); then the
// weighted row sums of taux ugeos_1, tauy vgeos_1, ug^2 and vg^2 in layer
// k_topo, wekp, |wekp|, ent_1, |ent_1| and eta_1 ent_1.
enum { EN_U2 = 0, EN_V2, EN_U4, EN_V4, EN_UJET, EN_NL };
enum { EN_UTX = EN_NL * MAXL, EN_VTY, EN_BTU, EN_BTV, EN_WEK, EN_AWEK, EN_ENT, EN_AENT, EN_ETAENT, EN_LEN };
struct EnergyArgs {
    Geo g;
    double bcco;
    double gp0;          // gp_1 (eta_1), 1 for one layer
    int dz_sign;
    const double* p;     // current time level
    const double* pm;    // previous time level
    const double* wekp;
    const double* ent1;
    const double* taux;  // null: no resident stress
    const double* tauy;
    double* d2u;         // nl fields: del2_P_bc(ug), ug = -dPdy_2bc(pm)/f0
    double* d2v;         // nl fields: del2_P_bc(vg), vg = dPdx_2bc(pm)/f0
    double* partial;     // [nyp][EN_LEN]
};
void launch_energy(const EnergyArgs& a, hipStream_t st);

// Ocean mixed layer (oqg_mixed.hip).  T fields are stored with the basin's
// pitch, nyp-1 rows.
struct OmlArgs {
    Geo g;
    double hm, T1_rel, dT_12, rho_cp, d2, d4, ycexp, nval, sval;
    int fluxn, fluxs, tau_sign;
    const double* p1;     // layer 1 of the basin's current p
    const double* taux;   // kinematic wind stress at p points
    const double* tauy;
    const double* fnet;   // net heat flux into the layer at T points
    double* sst;          // current level (in); the new level after k_oml_finish
    double* sstm;         // previous level (in); the current one after k_oml_finish
    double* snew;         // new level (scratch between the two launches)
    double* xfo;          // entrainment at T points (scratch)
    double* part;         // per-block partial sums of xfo (k_oml_step's blocks)
    int npart;
    double* ent1;         // the basin's resident layer-1 entrainment (p points)
    double tdt;
    int compute_ent;
};
dim3 oml_step_grid(const Geo& g);
// adv_diff (mixed_layer.F90:16-95) of one mixed-layer field at T point (i, j),
// 0-based: its C-grid advection by the layer-1 geostrophic plus Ekman velocity
// and its Del^2/Del^4 diffusion.  The Ekman velocity comes from the stress at
// p points with depth hm (ekman.F90:73-77).
struct AdvField {
    const double* data;    // current level (advected)
    const double* datam;   // previous level (diffused)
    double d2, d4, ycexp, nval, sval;
    int fluxn, fluxs;
};
// Ocean-atmosphere coupling (windstress.F90, forcing.F90, amlsubs.F90): the
// coupling grid, the fields either side and the coefficients one step reads.
struct CplArgs {
    Geo go, ga;                       // ocean and atmosphere geometry
    int ndxr, nxtoar, nytoar, nxpoar, nypoar, nxaooc, nyaooc, nx1, ny1;   // grid_type, nx1/ny1 1-based
    double xo0, yo0, xa0, ya0;        // box origins
    double hmo, hma, raoro, cdat, bcco_o, bcco_a;
    int tau_udiff, tau_sign_a;
    double D0up, Dmdown, Dup0, xlamda, fspco, Aup00, Adown10, Adown11;
    double hmamin, hmadmp, xcexp, T1_rel, dT_12, rho_cp, at2d, at4d, ahmd;
    double aface[MAXL], gpa[MAXL];
    double bface, dface;
    const double* pmo;                // ocean previous p (layer 1 at offset 0)
    const double* pma;                // atmosphere previous p, all layers
    const double* pa1;                // atmosphere current p, layer 1
    double *tauxo, *tauyo, *tauxa, *tauya;   // the basins' resident stress
    double* scr;                      // atmosphere p scratch: dpdx, dpdy, dxx, dyy
    double* coarse;                   // u, ux, uy, uxy, v, vx, vy, vxy on the (nxpoar, nypoar) points
    const double* st;                 // bicubic weights st[c][r], ((ndxr+1)^2, 16)
    const double* sstm;               // ocean mixed layer, previous level (T)
    double* fnetoc;                   // the ocean mixed layer's heat flux (T)
    double* ocflux;                   // ocean T scratch: the flux each ocean cell gives the atmosphere
    double *ast, *astm, *hmixa, *hmixam, *anew, *hnew, *xfa, *fnetat;   // atmosphere T fields
    const double* xc1ast;             // T
    const double* dtop;               // atmosphere topography (p)
    double* enta;                     // the atmosphere's resident layer-1 entrainment (p)
    double tdta;
    int compute_ent;
};
void launch_xforc(const CplArgs& a, hipStream_t st);
void launch_forcing(const CplArgs& a, hipStream_t st);
void launch_aml_step(const CplArgs& a, hipStream_t st);
void launch_aml_suppress(const CplArgs& a, hipStream_t st);
void launch_oml_step(const OmlArgs& a, hipStream_t st);
void launch_oml_finish(const OmlArgs& a, hipStream_t st);
void launch_oml_suppress(const OmlArgs& a, hipStream_t st);
// The rows combined (row order): sums [5][MAXL] (ke, p, q, eta, eta^2), extrema
// [6][MAXL] (min p, max p, min q, max q, max |u|, max |v|), p at column 1 of
// the first and last rows, the non-finite count.
constexpr int DC_SUM = 0, DC_EXT = 5 * MAXL, DC_P0FIRST = 11 * MAXL, DC_P0LAST = 12 * MAXL, DC_BAD = 13 * MAXL,
              DG_COMB_LEN = 13 * MAXL + 1;
void launch_diag_combine(const double* partial, int ny, int nl, double* comb, hipStream_t st);
// Scratch array of oqg_diagnose_noalloc (a file-scope device array, one per
// GPU): its device address on device `dev`, looked up once per device.
constexpr long long NOALLOC_SCRATCH_DOUBLES = 262144;   // 2 MiB
hipError_t noalloc_scratch(int dev, double** out);
// Up to six contiguous segments copied in one launch (state snapshots); every
// length is even and every pointer 16-byte aligned.  With `slots` set, the
// destinations are offsets doff[] into snapshot slot slots[*base + t] (the
// base slot is read from device memory, so a captured graph can be replayed
// for any stretch of slots).
struct CopySegs {
    const double* src[6];
    double* dst[6];
    long long n[6];
    int nseg;
    double* const* slots;
    const int* base;
    int t;
    long long doff[6];
    int b0[7];   // first block of each segment (set by launch_copy_segs)
};
void launch_copy_segs(const CopySegs& a, hipStream_t st);
void launch_set_int(int* p, int v, hipStream_t st);
void launch_zero_ring(double* f, const Geo& g, hipStream_t st);
void launch_fold_dup(double* f, const Geo& g, int nl, hipStream_t st);

void launch_adj_add(double* dst, const double* src, long long n, hipStream_t st);
void launch_adj_vort(const AdjVortArgs& a, hipStream_t st);

// Most dynamic LDS a launch of the LDS-staged kernels may ask for (the rest of
// the 160 KB of a CU is left to their static LDS).
constexpr size_t MAX_DYN_LDS = 150 * 1024;

// Kernel launchers (oqg_stencil.hip, oqg_transform.hip).
void launch_vort(const VortArgs& a, hipStream_t st);
void launch_fwd(const FwdArgs& a, int nmodes, hipStream_t st);
void launch_ysolve(const YArgs& a, int variant, hipStream_t st);
void launch_inv(const InvArgs& a, hipStream_t st);
void launch_fwd_adj(const FwdArgs& a, hipStream_t st);   // adjoint modes, every mode
void launch_inv_adj(const InvArgs& a, hipStream_t st);
void launch_suppress(const SuppArgs& a, hipStream_t st);
void launch_qcomp(const QcompArgs& a, hipStream_t st);
int inv_block_count(const Spec& s);
int ysolve_blocks(const Spec& s, int nmodes);
int fft_mb(const Spec& s, bool inverse);   // butterflies per thread of a plan (1, 2 or 4)
size_t fwd_lds_bytes(const Spec& s);
size_t ysolve_lds_bytes(const Spec& s);
size_t inv_lds_bytes(const Spec& s, const Geo& g, int nmodes);
void set_kernel_attrs();

// Ensemble launchers: one launch covers blocks of many basins; map[b] =
// (member index into args, that member's block index).  The caller groups
// members so that one launch shares the template parameters (layer count,
// butterflies per thread, y-solver rows per lane); lds is the largest need of
// the members in the launch.
void launch_vort_ens(const VortArgs* args, const int2* map, int nblocks, int ty, hipStream_t st);
// Tile rows of k_vort for a basin with tiles_x tile columns, nyp rows and nl
// layers: the fewest (4, 8, 16) whose blocks all fit on the device at once.
int vort_rows(int tiles_x, int nyp, int nl);
void launch_fwd_ens(const FwdArgs* args, const int2* map, int nblocks, int nl, int mb, size_t lds, hipStream_t st);
void launch_ysolve_ens(const YArgs* args, const int2* map, int nblocks, int R, size_t lds, hipStream_t st);
void launch_inv_ens(const InvArgs* args, const int2* map, int nblocks, int nl, int mb, size_t lds, hipStream_t st);
// The wavefront-transform forms (one wavefront per mode and row pair): used for
// the step whenever their LDS fits (fftw_fits); p = fftw_points of the basin.
int fftw_points(const Spec& s);
size_t fftw_lds_bytes(const Spec& s, int nl);
bool fftw_fits(const Spec& s, int nl);
void launch_fwdw(const FwdArgs& a, hipStream_t st);
void launch_invw(const InvArgs& a, hipStream_t st);
void launch_fwdw_ens(const FwdArgs* args, const int2* map, int nblocks, int nl, int p, size_t lds, hipStream_t st);
void launch_invw_ens(const InvArgs* args, const int2* map, int nblocks, int nl, int p, bool cyc, size_t lds,
                     hipStream_t st);
// Their per-layer-count objects (oqg_wave.hip built with -DOQG_WAVE_NL=N).
#define OQG_WAVE_DECL(N)                                                                                         \
    void launch_fwdw_nl##N(const FwdArgs& a, hipStream_t st);                                                    \
    void launch_invw_nl##N(const InvArgs& a, hipStream_t st);                                                    \
    void launch_fwdw_ens_nl##N(const FwdArgs* args, const int2* map, int nblocks, int p, size_t lds, hipStream_t st); \
    void launch_invw_ens_nl##N(const InvArgs* args, const int2* map, int nblocks, int p, bool cyc, size_t lds, hipStream_t st); \
    void launch_fwdw_adj_nl##N(const FwdArgs& a, hipStream_t st);                                                \
    void launch_invw_adj_nl##N(const InvArgs& a, hipStream_t st);                                                \
    void set_wave_attrs_nl##N(int lim);
OQG_WAVE_DECL(2)
OQG_WAVE_DECL(3)
OQG_WAVE_DECL(4)
OQG_WAVE_DECL(5)
OQG_WAVE_DECL(6)
OQG_WAVE_DECL(7)
OQG_WAVE_DECL(8)
#undef OQG_WAVE_DECL
void launch_suppress_ens(const SuppArgs* args, int nmembers, hipStream_t st);

}  // namespace oqg

/*
 * cpuref.h -- CPU oracle for OpenQG's QG timestep (step_qg).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * Fortran algorithm, used to check the HIP library.  Only tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py may load
 * it; the product library never links or calls it.
 *
 * Array layout: every (nxp,nyp,nl) field is Fortran column-major, i.e. the C
 * array a[k][j][i] with i (x) fastest, exactly as src/qg/qg.F90 stores
 * q, qm, p, pm.  All arithmetic is IEEE FP64 with no contraction
 * (built with -ffp-contract=off).
 */
#ifndef OQG_CPUREF_H
#define OQG_CPUREF_H

#ifdef __cplusplus
extern "C" {
#endif

#define CR_MAXL 8

/* Basin parameters: the subset of box_type / qg_type / topog_type
 * (src/glam/box.F90:11-41, src/qg/qg.F90:17-53, src/glam/topog.F90:16-28)
 * that step_qg reads. */
typedef struct {
    int nxp, nyp, nl;
    int cyclic;               /* box.F90:75 */
    double dx, dy;
    double fnot, beta;        /* mesh.F90:58-70 (set directly here) */
    double h[CR_MAXL];        /* layer thicknesses (m) */
    double gp[CR_MAXL];       /* reduced gravities, nl-1 used */
    double ah2[CR_MAXL], ah4[CR_MAXL];
    double bcco, delek;
    int dz_sign, tau_sign;
    int k_topo;               /* 1-based, topog.F90:52-56 */
} cr_params;

typedef struct cr_basin cr_basin;

/* Lifecycle: init_qg (qg.F90:60) -- modes, Helmholtz tables, homogeneous
 * solutions.  ddyn may be NULL (flat bottom). */
int  cr_create(const cr_params* prm, const double* ddyn, cr_basin** out);
void cr_destroy(cr_basin* b);

/* State access (q, qm, p, pm), all (nxp,nyp,nl). */
void cr_set_state(cr_basin* b, const double* q, const double* qm,
                  const double* p, const double* pm);
void cr_get_state(const cr_basin* b, double* q, double* qm, double* p, double* pm);

/* Constraint scalars, packed as dpi[nl-1], dpip[nl-1], cs[nl], cn[nl],
 * csp[nl], cnp[nl]  (length 2(nl-1)+4nl). */
int  cr_constr_len(const cr_basin* b);
void cr_get_constr(const cr_basin* b, double* out);
void cr_set_constr(cr_basin* b, const double* in);

/* init_pv (vorsubs.F90:17) + init_constraint (cyclic, constraint.F90:80)
 * + init_mass_constr (constraint.F90:145), as openqg.F90:239-243. */
void cr_init_state(cr_basin* b);

/* step_qg (core.F90:16).  ent is (nxp,nyp,n_ent) with n_ent >= 1; layer 2
 * edge rows are read in the cyclic case when n_ent >= 2. */
void cr_step(cr_basin* b, double tdt, const double* ent, int n_ent,
             const double* wekp, double txis, double txin);

/* suppress_comp_mode (driver.F90:174-189). */
void cr_suppress(cr_basin* b);

/* ocqbdy (vorsubs.F90:58-114) on the basin's state: boundary q from p. */
void cr_ocqbdy(cr_basin* b);

/* Thomas algorithm of inhomog.F90:188-199 (Numerical Recipes form) for
 * a*u[j-1] + bb*u[j] + a*u[j+1] = f[j], j = 0..n-1, zero outside. */
void cr_thomas(int n, double a, double bb, const double* f, double* u);

/* Derived data for pinning tests. */
void cr_get_modes(const cr_basin* b, double* amat, double* ctl2m,
                  double* ctm2l, double* rdm2);
/* Box: hom_sol_bc (nxp,nyp,nl-1).  Cyclic: hom_bt(nyp), bc1(nyp,nl-1),
 * bc2(nyp,nl-1), coef[4*(nl-1)+1] = hc1s,hc2s,hc1n,hc2n per mode, hbsi. */
void cr_get_homog(const cr_basin* b, double* a0, double* a1, double* a2, double* coef);

/* Building blocks, exposed so tests can pin each against closed forms and
 * library routines.  All use the basin's geometry. */
void cr_solve_inhomog(cr_basin* b, int m, const double* rhs, double* out); /* inhomog.F90:156 */
void cr_del2_bc(const cr_basin* b, const double* in, double* out, int nlayers); /* numerics.F90:410 */
void cr_jacobian(const cr_basin* b, const double* p, const double* q, double* jac,
                 double* ajis, double* ajin);                               /* numerics.F90:340 */
double cr_int_P_dA(const cr_basin* b, const double* f);                     /* numerics.F90:188 */
double cr_int_P_dx(const cr_basin* b, const double* row);                   /* numerics.F90:206 */

/* Stand-alone transforms. */
void cr_dst1(int n, int count, const double* in, double* out); /* FFTW RODFT00, unnormalised */
void cr_rfft(int n, int count, const double* in, double* out_c);  /* r2c: n/2+1 complex out */
void cr_irfft(int n, int count, const double* in_c, double* out); /* c2r, unnormalised */
void cr_cfft(int n, int sign, int count, const double* in, double* out); /* complex, interleaved */

/* Wind-stress forcing: ekman_from_tau (ekman.F90:53-83) with map_T_to_P
 * (numerics.F90:64-107).  taux, tauy (nxp,nyp) kinematic (m^2/s^2);
 * outputs wekp (nxp,nyp), txis, txin (cyclic only, else 0). */
void cr_ekman_wekp(const cr_params* prm, double hm, const double* taux, const double* tauy,
                   double* wekp, double* txis, double* txin);

/* Ocean mixed layer (omlsubs.F90, mixed_layer.F90, mixed.F90).  T fields are
 * (nxt,nyt) = (nxp-1,nyp-1), x fastest. */
typedef struct {
    double hm;            /* mixed layer depth (m) */
    double T1_rel, dT_12; /* temp_type (mixed.F90:34-37) */
    double rho_cp;
    double d2, d4;        /* st2d, st4d */
    double ycexp;
    double nval, sval;    /* boundary temperatures (zonal_temp at nyt and 1) */
    int fluxn, fluxs;
} cr_oml_params;

void cr_del2_T(const cr_params* P, const double* t, int north_flux, double nval, int south_flux, double sval,
               double* out);                                                  /* numerics.F90:107 */
void cr_ekman_uvw(const cr_params* P, double hm, const double* taux, const double* tauy, double* uek,
                  double* vek, double* wekt);                                 /* ekman.F90:73-80 */
void cr_adv_diff(const cr_params* P, const cr_oml_params* o, const double* p, const double* uek,
                 const double* vek, const double* data, const double* datam, double* out); /* mixed_layer.F90:16 */
void cr_map_T_to_P(const cr_params* P, const double* t, double* out);        /* numerics.F90:64 */
void cr_oml_step(const cr_params* P, const cr_oml_params* o, const double* p1, const double* taux,
                 const double* tauy, const double* fnet, double tdto, int compute_ent, double* sst,
                 double* sstm, double* ent1);                                 /* omlsubs.F90:69 */

/* Coupling of the ocean and the atmosphere: the coupling grid, the coupled
 * wind stress (windstress.F90:102-316 with bicubic.F90), the heat-flux forcing
 * (forcing.F90:18-99) and the atmospheric mixed layer (amlsubs.F90:184-326). */
typedef struct {
    /* grid_type (grid.F90:11-27), filled by cr_load_grid (grid.F90:37-58) */
    int ndxr, nxtoar, nytoar, nxpoar, nypoar, nxaooc, nyaooc, nx1, ny1;
    /* box origins x0, y0 (box.F90:18, 62-63) and mixed-layer depths hm */
    double xo0, yo0, xa0, ya0, hmo, hma;
    double rho_o, rho_a;  /* qgo%rho, qga%rho */
    double cdat;          /* windstress_type drag coefficient */
    int tau_udiff;        /* stress from the atmosphere-ocean velocity difference */
} cr_cpl;

typedef struct {
    /* the radiate_type coefficients a step reads (radsubs.F90:14-53) */
    double D0up, Dmdown, Dup0, xlamda, fspco, Aup00, Adown10, Adown11;
    /* atmos_mixed_type (amlsubs.F90:30-53) and its temp_type */
    double hmamin, hmadmp, xcexp, T1_rel, dT_12, rho_cp;
    double at2d, at4d, ahmd;
    double aface[CR_MAXL];  /* aface(l), l = 1..nl-1, at [l-1] */
    double bface, dface;
} cr_aml_params;

void cr_load_grid(const cr_params* Po, const cr_params* Pa, cr_cpl* g);
void cr_bcuini(int ndxr, double* st);          /* bicubic.F90:97-146: st((ndxr+1)^2, 16), column-major */
void cr_dPdx_2bc(const cr_params* P, double alpha, const double* p, double* out);   /* numerics.F90:470 */
void cr_dPdy_2bc(const cr_params* P, double alpha, const double* p, double* out);   /* numerics.F90:494 */
void cr_dP2dx2_bc(const cr_params* P, double alpha, const double* p, double* out);  /* numerics.F90:427 */
void cr_dP2dy2_bc(const cr_params* P, double alpha, const double* p, double* out);  /* numerics.F90:450 */
void cr_new_bicubic(const cr_cpl* g, int nxpo, int nypo, const double* st, const double* u, const double* ux,
                    const double* uy, const double* uxy, double* out);               /* bicubic.F90:26-46 */
void cr_xforc(const cr_params* Po, const cr_params* Pa, const cr_cpl* g, const double* st, const double* pmo1,
              const double* pma1, double* tauxo, double* tauyo, double* tauxa, double* tauya); /* windstress.F90:102 */
void cr_bilint(const cr_params* Pa, const cr_params* Po, const cr_cpl* g, const double* atmos,
               double* ocean);                                                  /* numerics.F90:528-589 */
void cr_compute_forcing(const cr_params* Po, const cr_params* Pa, const cr_cpl* g, const cr_aml_params* a,
                        const double* sstm, const double* astm, const double* hmixam, const double* eta1,
                        const double* dtop, double* fnetoc, double* fnetat);     /* forcing.F90:18-99 */
void cr_aml_step(const cr_params* Pa, const cr_cpl* g, const cr_aml_params* a, const double* p1,
                 const double* taux, const double* tauy, const double* eta, const double* dtop,
                 const double* fnetat, const double* xc1ast, double tdta, int compute_ent, double* ast,
                 double* astm, double* hmixa, double* hmixam, double* ent1);                                 /* amlsubs.F90:184-326 */

/* OpenMP thread count used by the oracle (0 = runtime default). */
void cr_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif

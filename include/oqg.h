/*
 * oqg.h -- C ABI of the MI355X (gfx950) QG timestep library, liboqg.so.
 *
 * Drop-in boundary for OpenQG's step_qg (src/qg/core.F90:16-33).  A context
 * holds one basin (ocean box or atmosphere channel) resident in device memory;
 * the caller steps it and reads fields back when its diagnostics, coupling or
 * restart code needs them.
 *
 * Layout: every field argument is a Fortran (nxp,nyp,nl) array, i.e. the C
 * array a[k][j][i] with x fastest, exactly as qg%q / qg%p are stored
 * (src/qg/qg.F90:84-87).  All arithmetic is IEEE FP64.
 *
 * Conventions: functions return 0 on success and a nonzero OQG_ERR_* code on
 * failure; oqg_last_error() gives the message.  A context is not thread safe;
 * use one context per basin.  Host pointers are borrowed only for the call.
 * `kind` selects host (OQG_HOST) or device (OQG_DEVICE) memory for pointer
 * arguments that accept either.
 */
#ifndef OQG_H
#define OQG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OQG_MAX_LAYERS 8

enum { OQG_HOST = 0, OQG_DEVICE = 1 };
enum { OQG_FIELD_Q = 0, OQG_FIELD_QM = 1, OQG_FIELD_P = 2, OQG_FIELD_PM = 3 };
enum {
    OQG_OK = 0,
    OQG_ERR_ARG = 1,         /* invalid argument or unsupported shape */
    OQG_ERR_HIP = 2,         /* HIP runtime error (message has the HIP error string) */
    OQG_ERR_NUMERIC = 3,     /* failed precompute check (mode identity, singular matrix) */
    OQG_ERR_STATE = 4        /* call made in the wrong state (e.g. step before state set) */
};
/* oqg_desc.flags */
enum {
    OQG_FLAG_YSOLVE_THOMAS = 1   /* y-tridiagonal by per-thread Thomas instead of the
                                    wavefront partition + PCR solver */
};
/* Ensemble launch schedules (bits), see oqg_ens_set_schedule. */
enum {
    OQG_ENS_ONE_YSOLVE = 1,      /* one y-solve launch for all members (rows per lane of
                                    the largest) instead of one per rows-per-lane class */
    OQG_ENS_LARGEST_FIRST = 2    /* members ordered by size, largest first, in every launch */
};

/*
 * Basin description: the fields of box_type (src/glam/box.F90:11-41),
 * qg_type (src/qg/qg.F90:17-53) and topog_type (src/glam/topog.F90:16-28)
 * that the step reads.  Fixed-size arrays keep the struct interoperable with
 * Fortran bind(C).  The optional pointers let a Fortran caller hand over the
 * precompute its unchanged init_qg already did (qg%mod, qg%hom); when NULL the
 * library derives them itself.
 */
typedef struct oqg_desc {
    int nxp, nyp, nl;             /* p-grid points and layers (box.F90:132-133) */
    int cyclic;                   /* zonally periodic channel (box.F90:75) */
    double dx, dy;                /* grid spacing (m) */
    double fnot, beta;            /* f0 (1/s), beta (1/(m s)) */
    double h[OQG_MAX_LAYERS];     /* layer thicknesses (m) */
    double gp[OQG_MAX_LAYERS];    /* reduced gravities, nl-1 used (m/s^2) */
    double ah2[OQG_MAX_LAYERS];   /* Del^4 coefficients (m^2/s) */
    double ah4[OQG_MAX_LAYERS];   /* Del^6 coefficients (m^4/s) */
    double bcco;                  /* mixed boundary coefficient */
    double delek;                 /* bottom Ekman layer thickness (m) */
    int dz_sign, tau_sign;        /* basin.F90:36,62 (ocean -1,+1) / :129,158 (atmos +1,-1) */
    int k_topo;                   /* 1-based; 0 = derive from dz_sign (topog.F90:52-56) */
    int flags;                    /* OQG_FLAG_* */
    const double* ddyn;           /* (nxp,nyp) dynamic topography, NULL = flat */
    /* optional precompute, Fortran (nl,nl) column-major; NULL = derive */
    const double* amat;           /* modes_type%amat  (modes.F90:13) */
    const double* ctl2m;          /* modes_type%ctl2m (modes.F90:17) */
    const double* ctm2l;          /* modes_type%ctm2l (modes.F90:18) */
    const double* rdm2;           /* modes_type%rdm2 (nl)           */
    /* optional homogeneous solutions; NULL = derive (requires the modes above
       to be consistent with them when given) */
    const double* hom_sol_bc;     /* box: (nxp,nyp,nl-1)  homog_box_type%hom_sol_bc */
    const double* hom_bt;         /* cyclic: (nyp)        hom_sol_bt */
    const double* hom_bc1;        /* cyclic: (nyp,nl-1)   hom_sol_bc1 */
    const double* hom_bc2;        /* cyclic: (nyp,nl-1)   hom_sol_bc2 */
    const double* hom_coef;       /* cyclic: hc1s,hc2s,hc1n,hc2n per mode 2..nl, then hbsi */
} oqg_desc;

typedef struct oqg_ctx oqg_ctx;

/* init_qg (src/qg/qg.F90:60-87): allocate device state, derive modes
 * (eigmod, modes.F90:46), Helmholtz tables (init_inhomog, inhomog.F90:71)
 * and homogeneous solutions (homsol, homog.F90:108).  State starts at zero.
 * The x transform length (box 2(nxp-1), channel nxp-1) must factor into 2, 3,
 * 5, 7, 11 and 13 (FFTW takes any, inhomog.F90:103-122) and fit the LDS
 * transforms (box nxp <= 1025, channel nxp <= 1025); lengths with a factor 7,
 * 11 or 13 run the plain step only (no batched ensembles, no adjoint). */
int oqg_create(const oqg_desc* desc, oqg_ctx** out);
void oqg_destroy(oqg_ctx* ctx);

/* Message for the last failure on ctx (or of oqg_create when ctx is NULL). */
const char* oqg_last_error(const oqg_ctx* ctx);

/* Replace or read the four time levels q, qm, p, pm (qg%q .. qg%pm).  NULL
 * pointers are skipped.  In a cyclic basin column nxp must equal column 1. */
int oqg_set_state(oqg_ctx* ctx, const double* q, const double* qm, const double* p,
                  const double* pm, int kind);
int oqg_get_state(oqg_ctx* ctx, double* q, double* qm, double* p, double* pm, int kind);

/* One layer (k = 1..nl) of q/qm/p/pm (OQG_FIELD_*): the cheap exchanges of
 * driver.F90:100,127 and windstress.F90:129-130. */
int oqg_get_layer(oqg_ctx* ctx, int which, int k, double* out, int kind);

/* Constraint scalars (constraint.F90:22-67), packed as
 * dpi[nl-1], dpip[nl-1], cs[nl], cn[nl], csp[nl], cnp[nl]. */
int oqg_constr_len(const oqg_ctx* ctx);
int oqg_set_constr(oqg_ctx* ctx, const double* v);
int oqg_get_constr(oqg_ctx* ctx, double* v);

/* init_pv (vorsubs.F90:17-25), init_constraint (constraint.F90:80-143, cyclic)
 * and init_mass_constr (constraint.F90:145-165) from the current p and pm,
 * in the order of openqg.F90:239-243. */
int oqg_init_state(oqg_ctx* ctx);

/* Upload forcing that stays resident for later steps: ent (nxp,nyp,n_ent),
 * n_ent >= 1 (layer 1 everywhere; layer 2 edge rows in the cyclic case),
 * wekp (nxp,nyp), and the cyclic boundary stress integrals txis, txin. */
int oqg_set_forcing(oqg_ctx* ctx, const double* ent, int n_ent, const double* wekp,
                    double txis, double txin, int kind);

/* step_qg (core.F90:16-33): vort_step, solve_pressure, ocqbdy.  ent/wekp are
 * host arrays as in the reference call; pass NULL for both to use the
 * resident forcing (then txis/txin are also taken from oqg_set_forcing).
 * Asynchronous on the context's stream when the forcing is resident. */
int oqg_step(oqg_ctx* ctx, double tdt, const double* ent, int n_ent, const double* wekp,
             double txis, double txin);

/* suppress_comp_mode (driver.F90:174-189). */
int oqg_suppress(oqg_ctx* ctx);

/* Ocean-only time loop (openqg.F90:483-509 with driver.F90:89-110) on the
 * resident forcing: for nt = step0+1 .. step0+nsteps, one step_qg, then
 * suppress_comp_mode when mod(nt-1, supp_every) == 0 (supp_every <= 0: never).
 * Replays cached HIP graphs; asynchronous. */
int oqg_run(oqg_ctx* ctx, double tdt, int nsteps, int step0, int supp_every);

/* Planning for oqg_run (no counterpart in the reference): captures, without
 * running anything, the HIP graphs that oqg_run with the same arguments, from
 * the current time level, replays.  oqg_run splits a run into whole
 * suppression periods (replayed from one cached graph each) and a shorter
 * head and tail; those are replayed from graphs only when planned, and
 * launched kernel by kernel otherwise.  Plan once, run many times. */
int oqg_plan_run(oqg_ctx* ctx, double tdt, int nsteps, int step0, int supp_every);

/* Wait for all work queued on the context's stream. */
int oqg_sync(oqg_ctx* ctx);

/* Tile height of the vorticity kernel (no counterpart in the reference, whose
 * vort_step has no tiling): 4, 8 or 16 rows of 64 points; 0 restores the
 * default rule (the fewest rows whose workgroups the device holds at once).
 * The results do not depend on it, except for the summation order of the
 * layer-1 entrainment integral.  oqg_tune times trial_steps steps of oqg_run
 * (from step 1, suppression every supp_every) with each height, restores the
 * state, keeps the fastest and returns it (negative on error); ms[] (3
 * entries: 4, 8, 16 rows; may be NULL) receives the times.  It allocates a
 * copy of the state (both time levels) while it runs and recaptures the
 * context's graphs for each height. */
int oqg_set_tile_rows(oqg_ctx* ctx, int rows);
int oqg_get_tile_rows(const oqg_ctx* ctx);
int oqg_tune(oqg_ctx* ctx, double tdt, int trial_steps, int supp_every, double* ms);

/* solve_inhomog_eqn (inhomog.F90:156-168) for mode m = 1..nl: rhs and out are
 * (nxp,nyp) arrays.  Synchronous. */
int oqg_solve_inhomog(oqg_ctx* ctx, int m, const double* rhs, double* out, int kind);

/* The mode matrices in use, Fortran (nl,nl) column-major, and rdm2 (nl). */
int oqg_get_modes(oqg_ctx* ctx, double* amat, double* ctl2m, double* ctm2l, double* rdm2);

/* Stream used for all device work of ctx (a hipStream_t). Setting NULL
 * restores the context's own stream. */
int oqg_set_stream(oqg_ctx* ctx, void* stream);
void* oqg_get_stream(oqg_ctx* ctx);

/* Device pointers of the current q, qm, p, pm and the row pitch (in doubles)
 * of the resident layout a[k][j][i] at a + (k*nyp + j)*pitch + i.  Valid until
 * the next step or suppression, which rotate the time levels. */
int oqg_state_ptrs(oqg_ctx* ctx, double** q, double** qm, double** p, double** pm,
                   size_t* pitch);

/* Advances the basin by one step of step_qg on the resident forcing (the same
 * result as oqg_step) and returns in ms[] the average duration of each kernel
 * of that step, measured with two HIP events on the context's stream around
 * `reps` (1..256) back-to-back launches of the kernel on the step's inputs
 * (names in `names`, '\n' separated; a kernel a basin does not use reports
 * 0).  Allocates 2*reps temporary copies of a field and of the spectra while
 * it runs.  Returns the number of kernels reported (<= max) or a negative
 * error. */
int oqg_kernel_times(oqg_ctx* ctx, double tdt, int reps, double* ms, int max, char* names,
                     int names_len);

/*
 * Adjoint (reverse mode) of step_qg, box and channel basins.  No counterpart
 * in the reference, which has no derivative code; these serve gradient
 * studies of the model (e.g. of a kinetic-energy diagnostic,
 * qg_monitor.F90:331-333, with respect to the wind stress and the initial PV).
 *
 * The adjoint state mirrors the forward state: with the forward state
 * S = (q, qm, p, pm) loaded, oqg_adj_set/get address the adjoints of q, qm,
 * p, pm (NULL arguments: zero / skip).  oqg_adj_step takes the adjoints of
 * the state after one step_qg (with its suppress_comp_mode when `suppressed`)
 * to those of the state before it; the forward state before the step must be
 * loaded (e.g. from a snapshot).  The adjoint of the resident Ekman pumping
 * field accumulates across steps.  Asynchronous on the context's stream.
 *
 * A channel's state is periodic in x with column nxp-1 a copy of column 0:
 * its adjoints are those of the periodic fields.  oqg_adj_set adds column
 * nxp-1 of its inputs into column 0, and every adjoint the library returns
 * holds 0 in column nxp-1 (a derivative with respect to a periodic point sits
 * in its column 0..nxp-2).  A channel's adjoint state also holds the
 * adjoints of the momentum-constraint scalars cs, cn, csp, cnp
 * (constraint.F90:182-215): oqg_adj_set zeroes them, oqg_adj_set_constr /
 * oqg_adj_get_constr address them in oqg_get_constr's packing (the dpi,
 * dpip entries are ignored and read 0), and the adjoints of the forcing
 * scalars txis, txin accumulate across steps beside that of wekp
 * (oqg_adj_get_txi, 2 doubles on the host; cleared by oqg_adj_zero_wekp).
 * oqg_adj_init rejects channels with nyp < 9.
 */
int oqg_adj_init(oqg_ctx* ctx);
int oqg_adj_set(oqg_ctx* ctx, const double* lq, const double* lqm, const double* lp, const double* lpm,
                int kind);
int oqg_adj_get(oqg_ctx* ctx, double* lq, double* lqm, double* lp, double* lpm, int kind);
int oqg_adj_set_constr(oqg_ctx* ctx, const double* v);
int oqg_adj_get_constr(oqg_ctx* ctx, double* v);
int oqg_adj_step(oqg_ctx* ctx, double tdt, int suppressed);
int oqg_adj_get_wekp(oqg_ctx* ctx, double* out, int kind);
int oqg_adj_get_txi(oqg_ctx* ctx, double* out);
int oqg_adj_zero_wekp(oqg_ctx* ctx);
/* Seeds the adjoint with the layer-1 kinetic energy of the loaded state,
 * J = ekei(1) = 0.5 rho H_1 int_P_dA(u^2 + v^2), u = -dPdy_2bc(p_1)/f0,
 * v = dPdx_2bc(p_1)/f0 (qg_monitor.F90:297-302,331-333): the adjoints of q,
 * qm, pm and the constraint scalars are zeroed and that of p is dJ/dp
 * (layer 1; 0 below), all on the device; *J (host) receives the objective. */
int oqg_adj_seed_ke(oqg_ctx* ctx, double rho, double* J);

/* Wind-stress forcing on the device: ekman_from_tau (src/coupler/ekman.F90:53-83)
 * with map_P_to_x_face / map_P_to_y_face and map_T_to_P
 * (src/glam/numerics.F90:64-107,271-289).  oqg_set_wind takes the kinematic
 * stress (taux, tauy) at p points, each (nxp,nyp), and sets the resident
 * Ekman pumping wekp (and, for a channel, txis/txin); the resident
 * entrainment is left as it is.  oqg_adj_wind applies the transpose to the
 * accumulated adjoints of wekp (oqg_adj_get_wekp) and, for a channel, of
 * txis and txin, and returns the adjoints of taux and tauy. */
int oqg_set_wind(oqg_ctx* ctx, const double* taux, const double* tauy, int kind);
int oqg_adj_wind(oqg_ctx* ctx, double* ltaux, double* ltauy, int kind);
/* The resident forcing: wekp (nxp,nyp) and txi = {txis, txin} (host, 2
 * doubles); either may be NULL. */
int oqg_get_forcing(oqg_ctx* ctx, double* wekp, double* txi, int kind);
/* The resident layer 1/2 entrainment (nxp,nyp), as the coupler holds it in
 * ocn_cpl%ent(:,:,1) between step_oml and step_qg (driver.F90:103-107). */
int oqg_get_ent(oqg_ctx* ctx, double* ent1, int kind);
/* The resident kinematic wind stress (nxp,nyp) each; either may be NULL.
 * The basin's eko/eka%taux, %tauy (coupler/ekman.F90:16-30). */
int oqg_get_wind(oqg_ctx* ctx, double* taux, double* tauy, int kind);

/*
 * Diagnostics of the resident state, computed on the device: the energetics
 * and averages of qg_diagno (src/diagnostics/qg_monitor.F90:219-383) that
 * depend on the state alone, scan_cfl's velocity maxima (:507-569) and a
 * count of non-finite values of p and q (the check of valids,
 * src/diagnostics/valsubs.F90:46-139, extended to NaN and Inf).  Per layer k
 * (nl used) and interface (nl-1 used), in the units of qg_monitor.
 */
typedef struct oqg_diag {
    double ekei[OQG_MAX_LAYERS];    /* 0.5 rho H_k int (u^2 + v^2) dA, u = -dPdy_2bc(p)/f0, v = dPdx_2bc(p)/f0 */
    double pavg[OQG_MAX_LAYERS];    /* avg_P(p_k) */
    double qavg[OQG_MAX_LAYERS];    /* avg_P(q_k) */
    double pmin[OQG_MAX_LAYERS], pmax[OQG_MAX_LAYERS];
    double qmin[OQG_MAX_LAYERS], qmax[OQG_MAX_LAYERS];
    double sfmin[OQG_MAX_LAYERS], sfmax[OQG_MAX_LAYERS];   /* transport stream function extrema (Sv) */
    double circ[OQG_MAX_LAYERS];    /* (p(1,1,k) - p(1,nyp,k))/f0 in Sv */
    double umax[OQG_MAX_LAYERS], vmax[OQG_MAX_LAYERS];     /* max |u|, |v| on the staggered faces */
    double cflu[OQG_MAX_LAYERS], cflv[OQG_MAX_LAYERS];     /* umax dt/dx, vmax dt/dx */
    double etam[OQG_MAX_LAYERS];    /* avg_P(eta_k), eta_k = dz_sign (p_k - p_{k+1})/gp_k */
    double et2m[OQG_MAX_LAYERS];    /* avg_P(eta_k^2) */
    double ctot;                    /* sum of circ */
    double nonfinite;               /* number of non-finite values of p and q */
    double pint[OQG_MAX_LAYERS];    /* int_P_dA(p_k) (check_continuity's int_p, qg_monitor.F90:379) */
} oqg_diag;

int oqg_diagnose(oqg_ctx* ctx, double dt, double rho, oqg_diag* out);

/* oqg_diagnose without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (serving loops, harnesses
 * that forbid hipMalloc).  oqg_diagnose allocates its per-row partials on its
 * first call; this variant keeps them, and their combination, in a
 * library-wide device array of 2 MiB (262144 doubles; one per GPU, held until
 * the process ends), so it serves nyp * 97 + 105 <= 262144, i.e. nyp <= 2701
 * (every basin the library accepts), and returns OQG_ERR_ARG otherwise.  It
 * calls no hipMalloc, hipFree or other allocating routine and gives the same
 * result as oqg_diagnose, bit for bit.  All calls on one GPU share the array:
 * calls that overlap on one GPU (from several threads or streams) race, so
 * serialise them. */
int oqg_diagnose_noalloc(oqg_ctx* ctx, double dt, double rho, oqg_diag* out);

/*
 * The rest of qg_diagno's energetics (src/diagnostics/qg_monitor.F90:219-383),
 * computed on the device from the resident state and forcing: the lagged
 * geostrophic velocities ug = -dPdy_2bc(pm)/f0, vg = dPdx_2bc(pm)/f0, the
 * current ones ugeos, vgeos from p, del2 = del2_P_bc and del4 = del2_P_bc of
 * del2 (numerics.F90:410-513), integrals by int_P_dA / int_P_dx.  The
 * reference declares vg (nxt,nyp,nl) (qg_monitor.F90:228) and assigns and
 * passes (nxp,nyp) fields through it, so its v terms of ah2d, ah4d and btdg
 * read past the array; here vg is (nxp,nyp), the shape of every field it is
 * combined with.  ujeto is taken over rows 1..nyt as its declaration
 * (:242) gives maxloc.  pket needs the previous sample's et2m and time: a
 * caller forms it from two oqg_diagnose results (0.5 rho gp_k (et2m - et2m')
 * / (tsec - tsec'), :266-272).
 */
typedef struct oqg_energy {
    double ah2d[OQG_MAX_LAYERS];    /* -rho ah2_k H_k int(ugeos del2(ug) + vgeos del2(vg)) dA */
    double ah4d[OQG_MAX_LAYERS];    /* rho ah4_k H_k int(ugeos del4(ug) + vgeos del4(vg)) dA */
    double val[OQG_MAX_LAYERS];     /* max over rows j = 1..nyt of |int_P_dx(ugeos(:,j))| / xl */
    double pos[OQG_MAX_LAYERS];     /* the first row j (1-based) where it is reached */
    double btdg;                    /* 0.5 rho delek |f0| int(ug^2 + vg^2) dA in layer k_topo */
    double utau;                    /* rho int(taux ugeos_1 + tauy vgeos_1) dA, the resident
                                       stress of oqg_set_wind (0 before any) */
    double wepm, wapm;              /* avg_P(wekp), avg_P(|wekp|) */
    double entm, enam;              /* avg_P(ent_1), avg_P(|ent_1|), the resident entrainment */
    double pken;                    /* rho gp_1 avg_P(eta_1 ent_1) (0 for one layer) */
} oqg_energy;

int oqg_diagnose_energy(oqg_ctx* ctx, double rho, oqg_energy* out);

/* The inversion of one step applied to the current q (interior): p from q
 * (modes, Helmholtz solves, mass correction with the current dpi), q's
 * boundary PV by ocqbdy, and qm = q, pm = p.  oqg_adj_invert is its adjoint:
 * from the adjoints of (q, qm, p, pm) it leaves the adjoint of the interior q
 * in the adjoint of q (ring and the other three zero). */
int oqg_invert(oqg_ctx* ctx);
int oqg_adj_invert(oqg_ctx* ctx);

/* Device-resident snapshots of the whole dynamic state (both time levels of
 * q and p, the constraint blocks and the parity), for checkpointing. */
int oqg_snap_alloc(oqg_ctx* ctx, int nslots);
int oqg_snap_save(oqg_ctx* ctx, int slot);
int oqg_snap_load(oqg_ctx* ctx, int slot);

/* Store-all trajectories.  oqg_adj_record runs steps step0+1 .. step0+nsteps
 * (suppress_comp_mode after those with mod(nt-1, supp) == 0 when supp > 0,
 * as oqg_run) and saves the state before each step in slots slot0 ..
 * slot0+nsteps-1.  oqg_adj_sweep, with the adjoint of the state after the
 * last step set, takes it back through those steps to the adjoint of the
 * state in slot slot0, which it leaves loaded.  Both are asynchronous. */
int oqg_adj_record(oqg_ctx* ctx, double tdt, int nsteps, int step0, int supp, int slot0);
int oqg_adj_sweep(oqg_ctx* ctx, double tdt, int nsteps, int step0, int supp, int slot0);

/* Binomial checkpointing (Griewank's revolve) with free_slots + 1 snapshot
 * slots (slot 0 keeps the initial state).  oqg_adj_revolve_forward runs steps
 * 1 .. nsteps from the loaded state (suppression as oqg_run with step0 = 0),
 * placing the schedule's first checkpoints on the way, and leaves the final
 * state loaded.  oqg_adj_revolve_reverse, with the adjoint of that state set,
 * takes it back to the adjoint of the initial state, which it leaves loaded;
 * forward_steps (may be NULL) receives the forward steps of both calls.
 * oqg_revolve_cost gives the forward steps that reversing nsteps from a stored
 * state costs with free_slots further slots, oqg_revolve_split the first
 * advance of that schedule (host only; -1 on bad arguments). */
int oqg_adj_revolve_forward(oqg_ctx* ctx, double tdt, int nsteps, int supp, int free_slots);
int oqg_adj_revolve_reverse(oqg_ctx* ctx, long long* forward_steps);
long long oqg_revolve_cost(int nsteps, int free_slots);
int oqg_revolve_split(int nsteps, int free_slots);

/*
 * Batched ensembles.  No counterpart in the reference, whose driver steps one
 * basin per step_qg call (src/driver.F90:107,143); an ensemble of independent
 * basins there is a set of separate runs.  Here the members, ordinary contexts
 * of any sizes and layer counts, advance together: each phase of the step is
 * one launch over the blocks of every member (per class of template
 * parameters), on the ensemble's own stream, replayed from HIP graphs.  A
 * member's results are those of stepping it alone.  While an ensemble runs,
 * its members must not be used otherwise.
 */
typedef struct oqg_ens oqg_ens;

/* Members must stay alive until oqg_ens_destroy. */
int oqg_ens_create(oqg_ctx* const* members, int n, oqg_ens** out);
void oqg_ens_destroy(oqg_ens* ens);
const char* oqg_ens_last_error(const oqg_ens* ens);

/* Advances member i by nsteps[i] steps of step_qg with time step tdt[i] on its
 * resident forcing, all members in lockstep from global step step0 + 1 (a
 * member stops after its own count), with suppress_comp_mode after the steps
 * nt with mod(nt-1, supp_every) = 0 (0: never) as src/openqg.F90:485-487.
 * Asynchronous on the ensemble stream; see oqg_ens_sync. */
int oqg_ens_run(oqg_ens* ens, const double* tdt, const int* nsteps, int step0, int supp_every);
int oqg_ens_sync(oqg_ens* ens);

/* Launch schedule (OQG_ENS_* bits).  oqg_ens_tune times trial_steps steps
 * on the members (whose state it restores) of every launch schedule with
 * every tile height (the members' own default rule, then 4, 8 and 16 rows for
 * all; oqg_set_tile_rows), keeps the fastest pair (the tile height set on
 * the members) and returns its index t*4 + schedule (negative on error);
 * ms[] (16 entries, [t*4 + schedule], may be NULL) receives the times.  It
 * allocates a copy of every member's state while it runs. */
int oqg_ens_set_schedule(oqg_ens* ens, int schedule);
int oqg_ens_get_schedule(const oqg_ens* ens);
int oqg_ens_tune(oqg_ens* ens, const double* tdt, int trial_steps, int supp_every, double* ms);

/*
 * Ocean mixed layer (ocean_mixed_type, src/mixed/omlsubs.F90:13-20) of an
 * ocean basin, on the device.  Its temperature lives at T points: (nxt,nyt) =
 * (nxp-1,nyp-1) arrays, x fastest.  It reads the basin's current layer-1
 * pressure and the resident wind stress of oqg_set_wind (its Ekman
 * velocities and pumping, src/coupler/ekman.F90:73-80), and writes the
 * basin's resident layer-1 entrainment, which the next step_qg uses.
 */
typedef struct oqg_oml_desc {
    double hm;            /* mixed layer depth (box_type%hm, m) */
    double T1_rel;        /* temp_type%T1_rel: T_bar - T(1) (K) */
    double dT_12;         /* temp_type%dT_12: T(1) - T(2) (K) */
    double rho_cp;        /* temp_type%rho_cp (J/(m^3 K)) */
    double st2d, st4d;    /* Del^2 / Del^4 diffusivities (mixed_type%d2, %d4) */
    double ycexp;         /* geostrophic advection coefficient (mixed_type%ycexp) */
    double nval, sval;    /* northern / southern boundary temperatures (zonal_temp at nyt, 1) */
    int fluxn, fluxs;     /* boundary heat-flux switches (basin flux_n, flux_s) */
} oqg_oml_desc;

typedef struct oqg_oml oqg_oml;

/* init_ocean_ml (omlsubs.F90:27-62) for basin ctx; temperature and net heat
 * flux start at zero.  The mixed layer must not outlive its basin. */
int oqg_oml_create(oqg_ctx* ctx, const oqg_oml_desc* desc, oqg_oml** out);
void oqg_oml_destroy(oqg_oml* oml);
const char* oqg_oml_last_error(const oqg_oml* oml);
/* The two time levels sst%data, sst%datam (NULL skips), (nxt,nyt). */
int oqg_oml_set_state(oqg_oml* oml, const double* sst, const double* sstm, int kind);
int oqg_oml_get_state(oqg_oml* oml, double* sst, double* sstm, int kind);
/* The resident net heat flux into the layer, fnetoc (nxt,nyt) (W/m^2). */
int oqg_oml_set_fnet(oqg_oml* oml, const double* fnet, int kind);
/* step_oml (omlsubs.F90:69-119): the new temperature with its convective
 * correction and, when compute_ent, the layer 1/2 entrainment written to the
 * basin's resident forcing.  Asynchronous on the basin's stream. */
int oqg_oml_step(oqg_oml* oml, double tdto, int compute_ent);
/* The mixed-layer part of suppress_ocean (driver.F90:165-170). */
int oqg_oml_suppress(oqg_oml* oml);
/* The ocean's run loop with its mixed layer: for nt = step0+1 .. step0+nsteps,
 * step_ocn (driver.F90:89-110: step_oml, then step_qg on the entrainment it
 * wrote) and, when mod(nt-1, supp_every) == 0, suppress_ocean (both the basin
 * and the mixed layer).  Replays cached HIP graphs; asynchronous. */
int oqg_run_ocean(oqg_oml* oml, double tdto, int nsteps, int step0, int supp_every, int compute_ent);

/*
 * Ocean-atmosphere coupling: the exchange of dynamic_step (driver.F90:36-87)
 * between an ocean box with its mixed layer (oqg_oml) and an atmosphere
 * channel, with the atmospheric mixed layer held by the coupler.  The coupler
 * computes on the device the coupled wind stress (xforc, windstress.F90:102-316,
 * with the bicubic interpolation of bicubic.F90), the heat-flux forcing
 * (compute_forcing, forcing.F90:18-99) and step_aml (amlsubs.F90:184-326), and
 * writes the stress, the ocean's heat flux and the atmosphere's entrainment
 * into the basins' resident buffers.  The coupling grid is load_grid's
 * (grid.F90:37-58): ndxr = nint(dxa/dxo), the ocean starting at atmosphere p
 * point (1, 1 + (nyta - nyto/ndxr)/2); the ocean's nxt and nyt must be
 * multiples of ndxr.  The radiation coefficients (radiate_type,
 * radsubs.F90:14-53) and the entrainment coefficients (init_entrain_coeff,
 * amlsubs.F90:93-129) are inputs.
 */
typedef struct oqg_cpl oqg_cpl;

typedef struct {
    double xo0, yo0, xa0, ya0;   /* box origins x0, y0 (box.F90:18, 62-63) (m) */
    double hmo, hma;             /* ocean and atmosphere mixed-layer depths go%hm, ga%hm (m) */
    double rho_o, rho_a;         /* qgo%rho, qga%rho */
    double cdat;                 /* drag coefficient (windstress_type) */
    int tau_udiff;               /* stress from the atmosphere-ocean velocity difference */
    double D0up, Dmdown, Dup0, xlamda, fspco, Aup00, Adown10, Adown11;   /* rad%D0up, %Dmdown, %Dup(0), %xlamda,
                                    %fspco, %Aup(0,0), %Adown(1,0), %Adown(1,1) */
    double hmamin, hmadmp, xcexp;          /* atmos_mixed_type */
    double T1_rel, dT_12, rho_cp;          /* the atmosphere's temp_type */
    double at2d, at4d, ahmd;               /* ast Del^2, Del^4 and hmixa Del^2 diffusivities */
    double aface[OQG_MAX_LAYERS];          /* aface(l), l = 1..nl-1 at [l-1] */
    double bface, dface;
    const double* xc1ast;        /* (nxta, nyta) host field (set_xc1ast); NULL: zero */
    const double* dtop;          /* (nxpa, nypa) atmosphere topography, host; NULL: flat */
} oqg_cpl_desc;

#define OQG_CPL_OCEAN_ENT 1      /* step_oml computes the ocean's entrainment (ent_coupled) */
#define OQG_CPL_ATMOS_ENT 2      /* step_aml computes the atmosphere's entrainment */

int oqg_cpl_create(oqg_oml* ocean_ml, oqg_ctx* atmos, const oqg_cpl_desc* d, oqg_cpl** out);
void oqg_cpl_destroy(oqg_cpl* cpl);
const char* oqg_cpl_last_error(const oqg_cpl* cpl);
/* The atmospheric mixed layer: aml%ast%data, %datam, aml%hmixa%data, %datam (nxta, nyta). */
int oqg_cpl_set_aml_state(oqg_cpl* cpl, const double* ast, const double* astm, const double* hmixa,
                          const double* hmixam, int kind);
int oqg_cpl_get_aml_state(oqg_cpl* cpl, double* ast, double* astm, double* hmixa, double* hmixam, int kind);
/* The last heat fluxes: fnetoc (ocean T, = the ocean mixed layer's fnet) and fnetat (atmosphere T). */
int oqg_cpl_get_fnet(oqg_cpl* cpl, double* fnetoc, double* fnetat, int kind);
/* xforc + ekman_from_tau of both basins (driver.F90:47-53) from their previous pressure. */
int oqg_cpl_xforc(oqg_cpl* cpl);
/* compute_forcing (forcing.F90:18-99): the ocean's fnet and the atmosphere's fnetat. */
int oqg_cpl_forcing(oqg_cpl* cpl);
/* step_aml (amlsubs.F90:184-229) on the atmosphere's current p and resident stress. */
int oqg_cpl_aml_step(oqg_cpl* cpl, double tdta, int compute_ent);
/* nsteps atmospheric steps of dynamic_step from step0 (openqg.F90:483-488):
 * step nt is an ocean step when mod(nt-1, nstr) == 0, suppresses the ocean when
 * mod(nt-1, osupp) == 0 and the atmosphere when mod(nt-1, asupp) == 0 (0: never).
 * Replays cached HIP graphs; asynchronous. */
int oqg_cpl_run(oqg_cpl* cpl, int nsteps, int step0, double tdto, double tdta, int nstr, int osupp, int asupp,
                int flags);

#ifdef __cplusplus
}
#endif
#endif

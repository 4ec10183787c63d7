/*
 * fdns.h -- C-ABI of the MI355X time-stepping core for FAlCon-DNS.
 *
 * The reference keeps all state in Fortran module arrays and advances it with
 * timeloop_imex / timeloop_imexrk (solver.F90:68-88).  This library replaces
 * that seam: a host driver (the Fortran driver in falcon-dns_amd/fortran, or
 * any C/ctypes caller) reads the same namelist, creates a run and advances it
 * on the GPU.  All arithmetic is FP64.
 *
 * Array convention: a spectral field is the reference's complex(dp) array
 * (Nm_max+1, Nr_max) in Fortran column-major order, passed as interleaved
 * re/im doubles: element (m, r) at [2*(m + (Nm+1)*r)], [2*(...)+1].
 *
 * Status codes: 0 = ok, negative = error (fdns_strerror).  Each handle is
 * used from one host thread; device work runs on the handle's own stream.
 */
#ifndef FDNS_H
#define FDNS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FDNS_OK            0
#define FDNS_EINVAL       -1   /* invalid parameter                         */
#define FDNS_ESCHEME      -2   /* scheme not in the registry, or not runnable */
#define FDNS_ESIZE        -3   /* Nr/Nm outside what the kernels support     */
#define FDNS_ENOMEM       -4   /* device or host allocation failed           */
#define FDNS_EHIP         -5   /* HIP runtime error                          */
#define FDNS_ESTATE       -6   /* call out of order (e.g. step before state) */
#define FDNS_ENOTFOUND    -7   /* unknown job id                             */
#define FDNS_EBUSY        -8   /* pool cannot admit the job (budget)          */

/* fdns_params.solver: how the per-mode implicit systems are solved.
 *   LU      packed LU factors, solved in dgetrs' operation order
 *           (algebra_lapack.F90:28-53): the reference's arithmetic.
 *   INVERSE resident inverses formed once from the LU factors; a solve is a
 *           matrix-vector product (the m = 0 temperature keeps the LU order).
 *           Same operators, rounding differs from dgetrs at ~1e-14 per solve.
 *   AUTO    INVERSE for Nr_max <= 129, LU above. */
#define FDNS_SOLVE_AUTO    0
#define FDNS_SOLVE_LU      1
#define FDNS_SOLVE_INVERSE 2

/* Namelist values of the reference that reach the time stepper
 * (namelists.F90:65-104; &grid, &physics, &timecontrol). */
typedef struct fdns_params {
    int32_t Nr;              /* Nr_max                                        */
    int32_t Nm;              /* Nm_max (Np_max = 3*Nm_max)                    */
    double  eta;             /* radius ratio                                  */
    double  Ra;              /* Rayleigh number                               */
    double  Pr;              /* Prandtl number                                */
    double  dt;              /* time step (the first one with l_vartimestep)  */
    char    mBC[8];          /* "NS" | "SF"                                   */
    char    buo_tscheme[8];  /* "IMP" | "EXP"                                 */
    char    time_scheme_type[16];  /* "IMEX" | "IMEXRK" | "RK"              */
    char    time_scheme_imp[16];   /* timeschemes.F90 names: IMEX "CN",
                                      "BDF2", "BDF3", "BDF4"; IMEXRK with
                                      diag_diff = .FALSE.: stiffly accurate
                                      "ARS222", "PC2", "SSP332", "SSP543",
                                      "ARS443", "BPR353", "FW53", "CK222",
                                      "BPR442"; ending with Assembly_stage
                                      "EUL111", "ARS232", "ARS233", "ARS343",
                                      "MARS343", "IJ3", "PR433", "DBM453",
                                      "KC343", "KC564", "KC774", "KC885",
                                      "CFN343", "BHR553", "CFN564"; each
                                      stage with its own diagonal (LU
                                      solves) "SMR333", "CB3e", "CB3f",
                                      "CB4", "LZ3L1", "LZ4A1" ("LZ3A4a"
                                      gives FDNS_ESCHEME: its zero
                                      diagonals make the vorticity-
                                      streamfunction operator singular);
                                      RK (explicit) "NONE"                 */
    char    time_scheme_exp[16];   /* "AB2", "AB3", "AB4"; "exp" + the
                                      IMEXRK name; RK "RK2", "RK4"         */
    int32_t l_restart;       /* 0: start at t=0 (zero histories, init.F90:193);
                                1: start as a restart from identical
                                checkpoints of the state given to
                                fdns_set_state (l_imexrk_started = .TRUE.)   */
    int32_t n_KE;            /* record (t, KE, Nu_t, Nu_b) every n_KE steps,
                                and keep the inputs of fdns_budget;
                                0 = never                                    */
    int32_t solver;          /* FDNS_SOLVE_* (no reference counterpart)      */
    /* variable dt (&timecontrol, namelists.F90:65-104): with l_vartimestep,
     * dt is the first step's dt and later steps take the CFL-limited dt of
     * compute_new_dt (steptime.F90:313-370, steptime_imexrk.F90:248-308),
     * re-factoring the operators whenever it changes (fourierloop.F90:86-94,
     * fourierloop_imexrk.F90:108-118).  Such runs step alone (no ensemble or
     * pool), launched one step at a time, and solver AUTO picks LU. */
    int32_t l_vartimestep;
    int32_t l_optimizedt;
    double  CFL;
    double  dt_coef;
    double  dt_max;
    /* lagpts (&physics): Lagrange points of the wall vorticity of
     * Assembly_stage (init.F90:268-271, fourierloop_imexrk.F90:601-627), used
     * by the IMEXRK schemes that are not stiffly accurate; 2..min(64, Nr-2) */
    int32_t lagpts;
} fdns_params;

typedef struct fdns_run fdns_run;

/* ---- one run ------------------------------------------------------------
 * fdns_create replaces solver_init's allocation and operator set-up
 * (solver.F90:38-66, mat_assembly.F90:33-50): it validates the scheme names
 * against the timeschemes.F90 registry, allocates device state, and builds and
 * factors the per-mode operators (mat_build, algebra_lapack.F90:14-23) on the
 * GPU, sharing factor sets with other runs of the same key. */
int fdns_create(const fdns_params* p, fdns_run** out);
int fdns_destroy(fdns_run* run);

/* State in the reference's arrays tFR, omgFR, urFR, upFR (init.F90:21-24).
 * set_state replaces init_fields/init_perturbation or get_checkpoint_data
 * (init.F90:164-275, 447-863) and resets time to 0. */
int fdns_set_state(fdns_run* run, const double* tFR, const double* omgFR,
                   const double* urFR, const double* upFR);
/* Any output pointer may be NULL. psi is psii (fourierloop.F90:204). */
int fdns_get_state(fdns_run* run, double* tFR, double* omgFR, double* urFR,
                   double* upFR, double* psi);
/* Fresh-start initial state of the reference, replacing init_fields and
 * init_perturbation (init.F90:164-399): the conductive temperature plus the
 * perturbation of amplitude ampT that n_init selects (&physics), u = omega =
 * 0.  n_init = 0: a compact Gaspari-Cohn bubble at the mid radius and
 * phi = pi; n_init < 0: a Gaussian at the mid radius and phi = 2, cut to
 * 1 < phi < 3; n_init > 0: ampT (rmax - r)(rmin - r) cos(n_init phi).  Built in
 * physical space on the Np x Nr grid and forward-transformed (forfft) as the
 * reference does; resets time to 0. */
int fdns_init_perturbation(fdns_run* run, int n_init, double ampT);
/* Build-defined seeded initial state (the reference has no RNG): conductive
 * T_0(r); T_m = amp*(a_m + i b_m)*sin(pi (r - rmin)) for m = 1..Nm with
 * a_m, b_m uniform in [-1,1) from SplitMix64(seed, m); u = omega = 0. */
int fdns_seed_state(fdns_run* run, uint64_t seed, double amp);

/* Advance nsteps time steps: the body of timeloop_imex (steptime.F90:196-251)
 * or timeloop_imexrk (steptime_imexrk.F90:100-190).  Synchronous. */
int fdns_step(fdns_run* run, int nsteps);

/* dt_new of the last step (the first column of the KE file, output.F90:256):
 * the constant dt, or the CFL-limited one of a variable-dt run; and, unless
 * NULL, dt_array (max(n_imp, n_exp) entries, newest first), as a checkpoint
 * stores them (output.F90:381-382). */
int fdns_get_dt(fdns_run* run, double* dt_new, double* dt_array);

/* writeKE_spectral's KE and Nusselt numbers (output.F90:232-265) for the
 * current state: out = {t, KE, Nu_t, Nu_b, KE_r, KE_phi}. */
int fdns_diag(fdns_run* run, double* out6);
/* Series recorded every n_KE steps: rows of {t, KE, Nu_t, Nu_b}. */
int fdns_get_series(fdns_run* run, double* out, int cap_rows, int* nrows);

/* writeKE_spectral's further outputs (output.F90:122-324) after the last
 * step: out4 = {the KE file's sixth column 2 pi radInt(<u_phi>^2), the
 * viscous dissipation and the buoyancy power of power_<tag> (columns 2, 3;
 * stress free: less the wall terms), t}; upbal (3 Nr doubles, level-major,
 * may be NULL): the upbalance_<tag> columns 2-4 per level, Re((u_phi,0 -
 * upFR_prev)/dt_new), the advective term -Re(<u_r du_phi/dr> + <u_r
 * u_phi>/r), the viscous term.  As the reference, the buoyancy power takes
 * the temperature of the step's last nonlinear pass and upFR_prev is the
 * state at the start of the step (IMEX) or the initial state (IMEXRK, RK).
 * Needs a run created with n_KE > 0 (it keeps those inputs) and a completed
 * step.  No counterpart as one call in the reference (it writes the files
 * from writeKE_spectral every n_KE steps). */
int fdns_budget(fdns_run* run, double* out4, double* upbal);

/* fdns_budget that allocates nothing, for callers that cannot allocate
 * device memory (tight serving or output loops, code that forbids
 * hipMalloc): no counterpart in the reference.  The outputs are staged in
 * one file-scope device array per GPU (4 + 3 * 513 doubles, held until the
 * process ends) and copied into out4 and upbal; sizes beyond it (Nr_max >
 * 513) give FDNS_ESIZE.  Every call on a GPU shares that GPU's array, so
 * calls that overlap on one GPU race; each call writes every staged value
 * before reading it back.  Same results as fdns_budget, bit for bit. */
int fdns_budget_noalloc(fdns_run* run, double* out4, double* upbal);

/* Restart with a change of resolution (init.F90:722-861): the four fields
 * (tFR, omgFR, urFR, upFR; (Nm+1, Nr) complex each) of an Nm_old x Nr_old
 * checkpoint on Nm_new x Nr_new >= the old sizes, for fdns_restart (one
 * state).  Host arithmetic, no run needed: REDFT00 of each mode, zero
 * padding, REDFT00 on the new grid, scaled as the reference (its highest old
 * Chebyshev mode enters doubled); new azimuthal modes zero. */
int fdns_refine_state(int Nm_old, int Nr_old, const double* tFR, const double* omgFR, const double* urFR,
                      const double* upFR, int Nm_new, int Nr_new, double* tN, double* omgN, double* urN,
                      double* upN);

/* calculate_spectra (output.F90:327-355) of the current state:
 * KEspec[m] = 4 pi radInt |ur_m^2 + up_m^2|, Tspec[m] = 4 pi radInt |T_m|,
 * m = 0..Nm_max (the KEspec_<tag> and Temp_spec_<tag> rows). */
int fdns_spectra(fdns_run* run, double* KEspec, double* Tspec);

/* ---- restart and checkpoint / snapshot files ------------------------------
 * n_order_tscheme_imp / n_order_tscheme_exp of the scheme (timeschemes.F90:
 * 42-973): the history depth of an IMEX scheme, the stage count of an IMEXRK
 * one.  An IMEX restart reads *n_imp checkpoints (init.F90:464-559). */
int fdns_scheme_orders(const fdns_params* p, int* n_imp, int* n_exp);
/* Restart start-up from checkpoint states, replacing get_checkpoint_data
 * (init.F90:447-592) and the history rebuild of steptime.F90:87-193.  The run
 * must have l_restart = 1.  tFR..upFR hold nstates states back to back, oldest
 * first, each an (Nm+1, Nr) array; the newest becomes the current state and
 * time is its tot_time.  IMEX: nstates = n_order_tscheme_imp (the reference's
 * l_imexrk_started = .TRUE.; CN and BDF2 read 2 checkpoints, BDF3 3, BDF4 4):
 * state j fills history slot n_imp - j of the implicit arrays and n_exp - j
 * of the explicit ones; or nstates = 1 (l_imexrk_started = .FALSE.): the
 * histories start at zero.  IMEXRK: nstates = 1. */
int fdns_restart(fdns_run* run, int nstates, const double* tFR, const double* omgFR,
                 const double* urFR, const double* upFR, double time);
/* dt_array of a restart (n = max(n_order_tscheme_imp, n_order_tscheme_exp)
 * entries, newest first) as get_checkpoint_data restores it from the
 * checkpoints (init.F90:464-612): with l_imexrk_started the IMEX schemes take
 * entry k from the first dt_array entry of the k-th newest checkpoint, else
 * the newest checkpoint's dt_array carries over (cut, or padded with its last
 * entry, when the orders differ).  Call after fdns_restart and before the
 * first step.  Only the variable-dt IMEX loop reads it: its first step keeps
 * entries 2..n and puts the CFL dt in entry 1 (steptime.F90:344-348); the
 * IMEXRK and RK loops reset dt_array to dt at their first step
 * (steptime_imexrk.F90:102-106, steptime_rk.F90:96-99), so for them it is
 * kept only until then. */
int fdns_restart_dt(fdns_run* run, int n, const double* dt_array);
/* The history arrays a checkpoint stores, in the reference's layout with the
 * order index fastest and slot 1 the newest: rhs_imp_temp, rhs_exp_temp,
 * rhs_imp_vort, rhs_exp_vort (n, Nm+1, Nr) and rhs_imp_uphi_bar,
 * rhs_exp_uphi_bar (n, Nr), n = n_imp or n_exp (timeschemes.F90:42-150).
 * IMEXRK runs: n = 1 and zeros (those arrays are (1, Nm+1, Nr) there and the
 * IMEXRK step never assigns them, rhs_create_imexrk.F90:7-8). */
int fdns_get_history(fdns_run* run, double* imp_temp, double* exp_temp, double* imp_vort,
                     double* exp_vort, double* imp_ub, double* exp_ub);
/* store_checkpoint / store_checkpoint_imexrk (output.F90:359-450): Fortran
 * unformatted sequential records (little-endian int32 length before and
 * after each record): (dt, tot_time, Nm, Nr, n_imp, n_exp, n_max), dt_array,
 * tFR, omgFR, urFR, upFR, then the six history arrays above.  The caller
 * names the file (the reference's checkpoint_NNNNN, I5.5). */
int fdns_write_checkpoint(fdns_run* run, const char* path);
/* store_snapshot / store_snapshot_imexrk (output.F90:451-516): (Ra, Pr, eta),
 * (dt, tot_time), (Nm, Nr), radius, tFR, omgFR, urFR, upFR, and for IMEXRK
 * runs omgFR_check (never assigned by the reference: zeros).  The file
 * python/annulus/snapshot.py:24-65 reads (snapshot_plot_NNNNN). */
int fdns_write_snapshot(fdns_run* run, const char* path);

/* ---- ensemble: many runs advanced by one batched step --------------------
 * No reference counterpart (the reference runs one simulation per process);
 * it batches independent runs of any mix of (Nr, Nm, Ra, Pr, scheme). */
typedef struct fdns_ens fdns_ens;
int fdns_ens_create(fdns_run* const* runs, int nruns, fdns_ens** out);
/* one batched step advances every member by one time step; use_graph != 0
 * replays a captured hipGraph of the batched step */
int fdns_ens_step(fdns_ens* ens, int nsteps, int use_graph);
/* the same batched steps launched without a graph and with the groups one
 * after another on the ensemble's stream, with a HIP event pair around every
 * kernel launch: kernel_ms[c] and launches[c] (FDNS_NKCLASS entries each) are
 * the summed duration and count per kernel class */
#define FDNS_KC_NONLIN    0   /* nonlinear pass (Nr_maxLOOP) */
#define FDNS_KC_IMEX_PRE  1   /* IMEX temperature right-hand sides */
#define FDNS_KC_IMEX_MID  2   /* IMEX new temperature, vorticity right-hand sides */
#define FDNS_KC_IMEX_POST 3   /* IMEX new vorticity, streamfunction, velocities */
#define FDNS_KC_RK_PRE    4   /* IMEXRK stage: explicit terms, temperature rhs */
#define FDNS_KC_RK_MID    5   /* IMEXRK stage: temperature, vorticity rhs */
#define FDNS_KC_RK_POST   6   /* IMEXRK stage: vorticity, psi, velocities */
#define FDNS_KC_SOLVE_T   7   /* temperature solves */
#define FDNS_KC_SOLVE_F   8   /* vorticity-streamfunction and mean-flow solves */
#define FDNS_KC_OTHER     9   /* first-step set-up, step counters, diagnostics */
#define FDNS_NKCLASS     10
int fdns_ens_step_timed(fdns_ens* ens, int nsteps, double* kernel_ms,
                        int* launches, double* wall_ms);
int fdns_ens_destroy(fdns_ens* ens);

/* ---- launch tuning --------------------------------------------------------
 * No reference counterpart (the reference leaves launch shapes to its
 * libraries).  Launch choices per (Nr_max, Nm_max); every choice gives bitwise
 * the same results, only the time changes.  A shape without a stored choice
 * starts from the library's table of choices measured for the benchmark
 * shapes, else from {256, 8, 0, 0}.  Runs, ensembles and pools pick up a
 * change at their next step (captured graphs and solve plans are made again). */
typedef struct fdns_tuning {
    int32_t nl_tpb;         /* threads per nonlinear-pass workgroup: 128 or 256
                               (256 above Nm_max = 768) */
    /* per column-tile kernel, in the order of kernel classes 1..6
     * (FDNS_KC_IMEX_PRE .. FDNS_KC_RK_POST): */
    int32_t tile_waves[6];  /* wavefronts per workgroup: 4 or 8 */
    int32_t tile_stage[6];  /* 1: stage the kernel's right-hand blocks in LDS
                               (where they fit in 160 KiB) */
    int32_t tile_pf[6];     /* 1: prefetch the kernel's epilogue inputs into
                               LDS during its MFMA pass (where they fit) */
    int32_t solve_chunk;    /* runs per batched-solve chunk: 8, 16, 24 or 32
                               (capped by the LDS the chunk stages), or 0 for
                               the largest chunk that fits */
    int32_t lane_split;     /* 1: past their first step, ensembles of this
                               shape split their groups with the most work per
                               step (runs x stages) over free stream lanes */
} fdns_tuning;
int fdns_tuning_get(int Nr, int Nm, fdns_tuning* out);
int fdns_tuning_set(int Nr, int Nm, const fdns_tuning* t);
/* Measure the choices on scratch copies of the ensemble's members (same
 * parameters, seeded states; the members are not touched): eight passes of
 * `steps` timed batched steps (fdns_ens_step_timed) after one untimed step,
 * pass i using option i of every family (cycled), and keep per family -- per
 * kernel for the column tiles -- the option with the least kernel time per
 * step; then `steps` graph-replayed steps with lane_split off and on, keeping
 * the faster; last, 4 `steps` graph-replayed steps with the choice so found
 * and with the choice in force before the call, keeping the faster (the
 * per-kernel times do not see how groups share the chip).  Needs members of
 * one (Nr_max, Nm_max).  *chosen = the stored
 * choice; report (may be NULL) gets FDNS_TUNE_REPORT per-step milliseconds,
 * -1 for an option not tried or a kernel not launched:
 *   [0..1]        nonlinear pass, nl_tpb 128 / 256
 *   [2 + 8c + o]  column-tile kernel c (0..5, classes 1..6), option o:
 *                 (tile_waves, tile_stage, tile_pf) = (4 or 8, o>>1 & 1,
 *                 o>>2 & 1) with waves 4 for even o
 *   [50..52]      solves, solve_chunk 0 (largest) / 16 / 8
 *   [53..54]      wall time per graph-replayed step, lane_split 0 / 1
 *   [55..56]      wall time per graph-replayed step, the choice before the
 *                 call / the one found (the faster is kept) */
#define FDNS_TUNE_REPORT 57
int fdns_ens_autotune(fdns_ens* ens, int steps, fdns_tuning* chosen, double* report);

/* ---- serving: continuous batching of simulation jobs --------------------
 * No reference counterpart.  A pool admits jobs while their HBM footprint
 * fits the budget, keeps their state and factors resident, and advances all
 * resident jobs together; finished jobs free their slot for waiting ones. */
typedef struct fdns_pool fdns_pool;
int fdns_pool_create(size_t hbm_budget_bytes, fdns_pool** out);
int fdns_pool_submit(fdns_pool* pool, const fdns_params* p, uint64_t seed,
                     double amp, int nsteps, int* job_id);
/* admit waiting jobs that fit, advance every resident job by one step,
 * retire finished jobs; *n_active = resident jobs after the call */
int fdns_pool_step(fdns_pool* pool, int* n_active, int* n_waiting);
/* the same with up to max_steps batched steps per call and one host
 * synchronisation: newly admitted jobs take their first step on their own;
 * the resident ones advance by k = min(max_steps, the fewest steps any of
 * them has left) replays of one captured graph of every group (captured
 * again only when jobs join or leave), so the call ends at the first step
 * that finishes a job; *steps_taken = k (1 when only new jobs stepped) */
int fdns_pool_run(fdns_pool* pool, int max_steps, int* steps_taken, int* n_active, int* n_waiting);
/* status: 0 waiting, 1 running, 2 done */
int fdns_pool_poll(fdns_pool* pool, int job_id, int* status, int* steps_done);
int fdns_pool_result(fdns_pool* pool, int job_id, double* tFR, double* omgFR,
                     double* urFR, double* upFR, double* diag6);
int fdns_pool_release(fdns_pool* pool, int job_id);
int fdns_pool_destroy(fdns_pool* pool);
/* bytes of device memory a job of these parameters needs when resident
 * (state, histories and its share of factor sets not yet resident) */
size_t fdns_job_footprint(const fdns_params* p);

/* ---- misc ---------------------------------------------------------------- */
const char* fdns_strerror(int code);
/* Measurement support, no reference counterpart: a plain device copy of
 * `bytes` (rounded down to 16) between two buffers, launched reps + 1 times
 * (kernel k_copy, the first launch untimed); *ms = mean duration of the timed
 * launches by HIP events.  Profilers' HBM byte counters are scaled by what
 * they read for this kernel. */
int fdns_device_copy(size_t bytes, int reps, double* ms);

#ifdef __cplusplus
}
#endif
#endif

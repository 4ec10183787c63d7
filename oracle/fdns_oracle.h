/*
 * fdns_oracle.h -- CPU restatement of the FAlCon-DNS time-stepping core.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product path: it may be loaded by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg, and by nothing else.  The product library never
 * links or calls it.
 *
 * Every function follows a named routine of the reference (file:line given at
 * each definition in fdns_oracle.c).  Arrays use the reference's Fortran
 * layout: a spectral field is (Nm+1, Nr) column-major complex(dp), i.e. element
 * (m, r) sits at index m + (Nm+1)*r (0-based), stored as interleaved re/im.
 */
#ifndef FDNS_ORACLE_H
#define FDNS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Run description: the namelist values of the reference that reach the hot
 * path (namelists.F90:65-104), plus the start-up mode. */
typedef struct orc_params {
    int32_t Nr;            /* Nr_max: Chebyshev-Gauss-Lobatto points           */
    int32_t Nm;            /* Nm_max: Fourier modes 0..Nm, Np = 3*Nm            */
    double  eta;           /* radius ratio                                      */
    double  Ra, Pr;        /* Rayleigh, Prandtl                                 */
    double  dt;            /* constant time step                                */
    int32_t mBC;           /* 0 = "NS" (no slip), 1 = "SF" (stress free)         */
    int32_t buo_imp;       /* 1 = buo_tscheme "IMP", 0 = "EXP" (IMEXRK only)     */
    int32_t restart;       /* 0: fresh start (l_restart=F, zero histories);
                              1: restart (l_restart=T): from identical
                              checkpoints of the state given to orc_set_state
                              (l_imexrk_started=T), or from the checkpoint
                              states given to orc_restart                     */
    int32_t nthreads;      /* OpenMP threads for the mode loops (0 = default)   */
    char    scheme_type[16];  /* "IMEX" | "IMEXRK"                           */
    char    scheme_imp[16];   /* e.g. "CN", "BDF3", "ARS443", "BPR353"       */
    char    scheme_exp[16];   /* e.g. "AB2", "AB3", "expARS443"             */
    /* &timecontrol of variable dt: l_vartimestep, l_optimizedt, CFL, dt_coef,
     * dt_max (dt above is then the first step's dt) */
    int32_t vardt, optimizedt;
    double  CFL, dt_coef, dt_max;
    /* lagpts (&physics): Lagrange points of the Johnston vorticity wall values
     * in Assembly_stage (IMEXRK schemes that are not stiffly accurate) */
    int32_t lagpts;
} orc_params;

typedef struct orc_run orc_run;

/* ---- whole-run API ---- */
int      orc_create(const orc_params* p, orc_run** out);
void     orc_destroy(orc_run* R);
int      orc_set_state(orc_run* R, const double* tFR, const double* omgFR,
                       const double* urFR, const double* upFR);
int      orc_get_state(const orc_run* R, double* tFR, double* omgFR,
                       double* urFR, double* upFR, double* psi);
int      orc_step(orc_run* R, int nsteps);
/* restart start-up from K checkpoint states, oldest first, back to back;
 * t = tot_time of the newest (see fdns_oracle.c) */
int      orc_restart(orc_run* R, int K, const double* tFR, const double* omgFR,
                     const double* urFR, const double* upFR, double t);
/* checkpoint history arrays, reference layout (n, Nm+1, Nr) / (n, Nr) */
int      orc_get_history(const orc_run* R, double* imp_temp, double* exp_temp,
                         double* imp_vort, double* exp_vort, double* imp_ub,
                         double* exp_ub);
/* dt_new of the last completed step */
double   orc_dt(const orc_run* R);
/* out: [t, KE, Nu_t, Nu_b, KE_r, KE_phi] after the last completed step */
int      orc_diag(orc_run* R, double* out);
int      orc_get_TFC(const orc_run* R, double* TFC);
/* writeKE_spectral's mean-flow energy, power budget {KE_mean, vis, buo, t}
 * and the per-level mean-flow balance (3 Nr, may be NULL) */
int      orc_budget(orc_run* R, double* out4, double* upbal);
/* calculate_spectra: KEspec(m), Tspec(m), m = 0..Nm */
int      orc_spectra(orc_run* R, double* KEspec, double* Tspec);
/* the RK path's mean-flow right-hand side (rhs_construct_uphi_bar) of the
 * current state, N complex values */
int      orc_rk_mean_rhs(orc_run* R, double* out);

/* restart with a change of resolution (init.F90:722-861): the four fields
 * of an Nm_o x Nr_o checkpoint on Nm_n x Nr_n */
int      orc_refine(int Nm_o, int Nr_o, int Nm_n, int Nr_n, const double* const in[4],
                    double* const out[4]);

/* ---- seeded initial condition (build-defined; see fdns_oracle.c) ---- */
void     orc_seed_state(int Nr, int Nm, double eta, uint64_t seed, double amp,
                        double* tFR, double* omgFR, double* urFR, double* upFR);

/* ---- the reference's fresh start (init_perturbation) and restart dt ---- */
void     orc_perturbation_field(int N, int Nm, double eta, int n_init, double ampT, double* tt);
void     orc_init_perturbation(int N, int Nm, double eta, int n_init, double ampT, double* tFR,
                               double* omgFR, double* urFR, double* upFR);
/* Gaspari piece: 0 middle (r < c), 1 edge (c <= r < 2c), 2 centre value */
double   orc_gaspari(int piece, double r, double c, double L);
int      orc_restart_dt(orc_run* R, int n, const double* dt_array);

/* ---- building blocks exposed for pinning tests ---- */
void     orc_dcheb(int N, double* t, double* D, double* D2, double* xp);
void     orc_grid(int N, double eta, double* radius, double* r_radius,
                  double* r_radius2);
void     orc_fft(int n, const double* in, double* out, int sign);
void     orc_chebtransform(int N, const double* f, double* ft);
void     orc_chebinvtran(int N, const double* ft, double* f);
void     orc_chebinvtranD1(int N, const double* ft, double* df);
void     orc_chebinvtranD2(int N, const double* ft, double* d2f);
void     orc_forfft(int Nm, int Np, const double* x, double* f);
void     orc_invfft(int Nm, int Np, const double* f, double* x);
double   orc_radint(int N, double eta, const double* f);
void     orc_mat_build(int N, double eta, int n, double wt_dt, double Pr, int mBC,
                       int irk_order, double* AT, double* AF, double* Auphi);
/* LAPpsi of mode n (mat_build_rk), column-major N x N */
void     orc_mat_build_lap(int N, double eta, int n, double* L);
int      orc_dgetrf(int n, double* A, int32_t* ipiv);
void     orc_dgetrs(int n, const double* A, const int32_t* ipiv, double* b);
void     orc_nonlin(int Nm, int Nr, const double* tFR, const double* omgFR,
                    const double* upFR, const double* urFR, double* upT,
                    double* urT, double* upW, double* urW, double* urw_mean);
/* scheme registry: fills weights / tableaux; returns stage or order info */
int      orc_scheme_info(const char* type, const char* imp, const char* exp,
                         int32_t* n_imp, int32_t* n_exp, double* aD, double* bD,
                         double* aA, double* bA, int32_t* sa);

/* ---- planted mistakes, for the tests that pin the coupled right-hand side:
 * 0 = the restatement (default); 1 = buoyancy sign flipped; 2 = buoyancy
 * scaled by Ra instead of Ra/Pr; 3 = (u_r f)/r of the flux-form advection
 * with the wrong sign.  Process-wide; tests set it back to 0. */
void     orc_set_mutation(int k);
int      orc_get_mutation(void);

#ifdef __cplusplus
}
#endif
#endif

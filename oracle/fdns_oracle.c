/*
 * fdns_oracle.c -- CPU restatement of the FAlCon-DNS time-stepping core.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library.
 *
 * It restates, routine by routine, the reference's hot path:
 *   grid and Chebyshev matrices     init.F90:120-162, chebyshev_fftw.F90:88-121
 *   Chebyshev transforms (DCT-I)    chebyshev_fftw.F90:124-427
 *   azimuthal real FFTs             fourier_fftw.F90:39-97
 *   nonlinear products              nonlin.F90:134-190
 *   operator matrices               mat_assembly.F90:68-413
 *   LU factor / solve               algebra_lapack.F90:14-53 (dgetrf/dgetrs)
 *   scheme weights and tableaux     timeschemes.F90:37-973, 993-2388
 *   IMEX step                       steptime.F90:196-251, fourierloop.F90:48-238,
 *                                   rhs_create.F90:24-385
 *   IMEXRK step, every tableau      steptime_imexrk.F90:100-190,
 *   (stiffly accurate, Assembly_    fourierloop_imexrk.F90:75-665,
 *   stage, own diagonals)           rhs_create_imexrk.F90:20-289
 *   LAPpsi, Lagrange wall weights   mat_assembly.F90:415-462,
 *                                   chebyshev_fftw.F90:429-517
 *   explicit RK step (RK2, RK4)     steptime_rk.F90:85-236, fourierloop_rk.F90,
 *                                   rhs_create_rk.F90
 *   variable dt                     steptime.F90:313-370, steptime_imexrk.F90:
 *                                   248-308, steptime_rk.F90:181-236
 *   KE / Nusselt diagnostics        output.F90:232-265, integ.F90:36-76
 *   power budget, mean-flow         output.F90:122-355
 *   balance, spectra
 *   multistep restart start-up      steptime.F90:87-193, rhs_create.F90:387-573,
 *                                   init.F90:447-592
 *   resolution-change restart       init.F90:722-861
 *   checkpoint history arrays       output.F90:359-450
 *
 * Each is pinned by the tests named in its comment block (tests/test_oracle_*,
 * tests/test_vardt.py, tests/test_checkpoint.py): closed forms, numpy
 * restatements, invariants and the reference's golden series.
 *
 * Third-party arithmetic the reference delegates: FFTW3 (DFTs) and LAPACK
 * dgetrf/dgetrs.  Neither is present here nor vendored by the reference, so
 * this file restates their published definitions: a mixed-radix DFT with
 * exactly computed twiddles, and the unblocked right-looking LU with partial
 * pivoting of LAPACK's dgetf2 (pivot = first max |a|, column scaled by the
 * reciprocal pivot) followed by dgetrs (row interchanges, unit-lower then upper
 * column-oriented triangular solves).  Multiply-adds in the LU are written as
 * explicit fma() so that the order and rounding of every factor element is
 * fixed.
 *
 * Layout: spectral fields are (Nm+1, Nr) column-major complex, element (m, r)
 * at m + (Nm+1)*r, interleaved re/im -- the reference's module arrays.
 */
#include "fdns_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cplx;

static inline cplx cx(double re, double im) { cplx z; z.re = re; z.im = im; return z; }
static inline cplx cadd(cplx a, cplx b) { return cx(a.re + b.re, a.im + b.im); }
static inline cplx csub(cplx a, cplx b) { return cx(a.re - b.re, a.im - b.im); }
static inline cplx cmul(cplx a, cplx b) { return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cplx cscale(double s, cplx a) { return cx(s * a.re, s * a.im); }
static inline cplx cconj(cplx a) { return cx(a.re, -a.im); }
/* i*m*r*z as the reference writes it: (ii*real(m))*r_radius*z (rhs_create.F90:99) */
static inline cplx cimr(double m, double r, cplx z) { return cmul(cx(0.0, m * r), z); }

#ifndef _OPENMP
static int omp_get_max_threads(void) { return 1; }
#endif

/* Planted mistakes in the coupled right-hand side, for the tests that pin it
 * (tests/test_oracle_coupling.py shows that each of its checks fails on them).
 * 0, the default, is the restatement itself; every other value is wrong on
 * purpose:
 *   1  buoyancy with the wrong sign:  +(Ra/Pr) i m T / r
 *   2  buoyancy scaled by Ra instead of Ra/Pr
 *   3  flux-form advection with the wrong sign on (u_r f)/r, for T and omega
 * With 0 the helpers return the reference's own expressions, bit for bit. */
static int g_mutation = 0;
void orc_set_mutation(int k) { g_mutation = k; }
int orc_get_mutation(void) { return g_mutation; }
/* Ra/Pr of the buoyancy term -(Ra/Pr) i m T/r (rhs_create.F90:216-218, 244-254,
 * fourierloop_imexrk.F90:298-300, rhs_create_rk.F90:112-115) */
static inline double buo_ra_pr(double Ra, double Pr)
{
    if (g_mutation == 1) return -(Ra / Pr);
    if (g_mutation == 2) return Ra;
    return Ra / Pr;
}
/* (1/r) (u_r f) of the flux-form advection i m/r (u_phi f) + d/dr (u_r f) +
 * (u_r f)/r (rhs_create.F90:99-100, 223-224; rhs_create_imexrk.F90:49-50,
 * 224-225; rhs_create_rk.F90:57-60) */
static inline cplx flux_over_r(double rr, cplx urf)
{
    return cscale(g_mutation == 3 ? -rr : rr, urf);
}

/* ------------------------------------------------------------------------ */
/* DFT: the arithmetic FFTW provides to fourier_fftw.F90 / chebyshev_fftw.F90 */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n;
    int nf;
    int f[40];
    cplx* tw;   /* tw[t] = exp(-2 pi i t / n) */
    cplx* scratch;
} fftplan;

static int plan_init(fftplan* P, int n)
{
    P->n = n;
    P->nf = 0;
    int m = n;
    for (int p = 2; m > 1;) {
        if (m % p == 0) { P->f[P->nf++] = p; m /= p; }
        else p = (p == 2) ? 3 : p + 2;
    }
    P->tw = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    if (!P->tw) return -1;
    for (int t = 0; t < n; t++) {
        long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n;
        P->tw[t] = cx((double)cosl(a), -(double)sinl(a));
    }
    P->scratch = NULL;
    return 0;
}

static void plan_free(fftplan* P)
{
    free(P->tw);
    P->tw = NULL;
}

/* Recursive decimation in time: out[k] = sum_j in[j*stride] * w^(sign*j*k).
 * sign = -1 is FFTW_FORWARD, +1 is FFTW_BACKWARD; no normalisation. */
static void fft_rec(const fftplan* P, int n, int fi, const cplx* in, int stride,
                    cplx* out, int sign, int twstep)
{
    if (n == 1) { out[0] = in[0]; return; }
    const int p = P->f[fi];
    const int m = n / p;
    for (int s = 0; s < p; s++)
        fft_rec(P, m, fi + 1, in + (size_t)s * stride, stride * p, out + (size_t)s * m, sign, twstep * p);
    cplx tmp[64];
    cplx* t = tmp;
    cplx* heap = NULL;
    if (p > 64) { heap = (cplx*)malloc(sizeof(cplx) * (size_t)p); t = heap; }
    for (int k = 0; k < m; k++) {
        for (int s = 0; s < p; s++) t[s] = out[(size_t)s * m + k];
        for (int q = 0; q < p; q++) {
            const int kk = k + q * m;
            cplx acc = t[0];
            for (int s = 1; s < p; s++) {
                const long e = ((long)s * kk) % n;
                cplx w = P->tw[(size_t)e * twstep];
                if (sign > 0) w.im = -w.im;
                acc = cadd(acc, cmul(w, t[s]));
            }
            out[kk] = acc;
        }
    }
    free(heap);
}

static void fft_exec(const fftplan* P, const cplx* in, cplx* out, int sign)
{
    fft_rec(P, P->n, 0, in, 1, out, sign, 1);
}

void orc_fft(int n, const double* in, double* out, int sign)
{
    fftplan P;
    plan_init(&P, n);
    fft_exec(&P, (const cplx*)in, (cplx*)out, sign);
    plan_free(&P);
}

/* ------------------------------------------------------------------------ */
/* Chebyshev machinery                                                        */
/* ------------------------------------------------------------------------ */

/* dcheb, chebyshev_fftw.F90:88-121.  Outputs column-major N x N matrices with
 * row = collocation point i, column = polynomial j (after the transposes at
 * :116-120): t(i,j) = T_j(x_i), D = 2 dT/dx, D2 = 4 d2T/dx2. */
void orc_dcheb(int N, double* t, double* D, double* D2, double* xp)
{
    const double pi = 4.0 * atan(1.0);
    double* T0 = (double*)malloc(sizeof(double) * (size_t)N * N);
    double* T1 = (double*)malloc(sizeof(double) * (size_t)N * N);
    double* T2 = (double*)malloc(sizeof(double) * (size_t)N * N);
    for (int i = 0; i < N; i++) xp[i] = cos(((double)(N - 1 - i) * pi) / (double)(N - 1));
    /* T0(j,i) in the reference's pre-transpose orientation: index j + N*i */
    for (int i = 0; i < N; i++) {
        double* a = T0 + (size_t)N * i;
        double* b = T1 + (size_t)N * i;
        double* c = T2 + (size_t)N * i;
        a[0] = 1.0; a[1] = xp[i];
        b[0] = 0.0; b[1] = 1.0;
        c[0] = 0.0; c[1] = 0.0;
        for (int j = 2; j < N; j++) {
            a[j] = 2.0 * xp[i] * a[j - 1] - a[j - 2];
            b[j] = 2.0 * a[j - 1] + 2.0 * xp[i] * b[j - 1] - b[j - 2];
            c[j] = 4.0 * b[j - 1] + 2.0 * xp[i] * c[j - 1] - c[j - 2];
        }
    }
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            t[i + (size_t)N * j] = T0[j + (size_t)N * i];
            D[i + (size_t)N * j] = 2.0 * T1[j + (size_t)N * i];
            D2[i + (size_t)N * j] = (2.0 * 2.0) * T2[j + (size_t)N * i];
        }
    free(T0); free(T1); free(T2);
}

/* init_grid, init.F90:120-140 */
void orc_grid(int N, double eta, double* radius, double* r_radius, double* r_radius2)
{
    const double pi = 4.0 * atan(1.0);
    const double rmin = eta / (1.0 - eta);
    for (int i = 0; i < N; i++) {
        radius[i] = rmin + (0.5 * (1.0 + cos(((double)(N - 1 - i) * pi) / (double)(N - 1))));
        r_radius[i] = 1.0 / radius[i];
        r_radius2[i] = r_radius[i] * r_radius[i];
    }
}

typedef struct {
    int N, M;          /* M = 2N-2 */
    fftplan P;
    double fact;       /* sqrt(2/(N-1)) */
} chebctx;

static int cheb_init(chebctx* C, int N)
{
    C->N = N;
    C->M = 2 * N - 2;
    C->fact = sqrt(2.0 / (double)(N - 1));
    return plan_init(&C->P, C->M);
}

/* chebtransform, chebyshev_fftw.F90:124-155.  w: 2*M cplx workspace. */
static void cheb_fwd(const chebctx* C, const cplx* f, cplx* ft, cplx* w)
{
    const int N = C->N, M = C->M;
    cplx* f00 = w;
    cplx* ff = w + M;
    for (int i = 0; i < N; i++) f00[i] = f[N - 1 - i];
    for (int i = 1; i < N - 1; i++) f00[N + i - 1] = f[i];
    fft_exec(&C->P, f00, ff, +1);
    const double dM = (double)M;
    for (int k = 0; k < M; k++) { ff[k].re = ff[k].re / dM; ff[k].im = ff[k].im / dM; }
    const double fac = C->fact;
    ft[0] = cx(2.0 * (ff[0].re / fac), 2.0 * (ff[0].im / fac));
    ft[N - 1] = cx(2.0 * (ff[N - 1].re / fac), 2.0 * (ff[N - 1].im / fac));
    for (int k = 1; k < N - 1; k++) ft[k] = cx((2.0 * ff[k].re) / fac, (2.0 * ff[k].im) / fac);
}

/* chebinvtran, chebyshev_fftw.F90:157-194: REDFT00 of Re and Im (here both at
 * once through the even extension), times 0.5*fact, then reversed. */
static void cheb_inv(const chebctx* C, const cplx* ft, cplx* f, cplx* w)
{
    const int N = C->N, M = C->M;
    cplx* x = w;
    cplx* y = w + M;
    for (int i = 0; i < N; i++) x[i] = ft[i];
    for (int i = 1; i < N - 1; i++) x[M - i] = ft[i];
    fft_exec(&C->P, x, y, +1);
    const double fact = C->fact;
    for (int i = 0; i < N; i++) f[N - 1 - i] = cx((y[i].re * 0.5) * fact, (y[i].im * 0.5) * fact);
}

/* Backward recurrence for derivative coefficients, chebyshev_fftw.F90:262-267 */
static void cheb_deriv_coef(int N, const cplx* f, cplx* beta)
{
    beta[N - 1] = cx(0.0, 0.0);
    beta[N - 2] = cscale(2.0 * (double)(N - 1), f[N - 1]);
    for (int i = N - 2; i >= 1; i--) {
        /* beta(i-1) = beta(i+1) + 4(i-1) f(i)  with 1-based i -> i+1 here */
        const double c = 4.0 * (double)i;
        beta[i - 1] = cadd(beta[i + 1], cscale(c, f[i]));
    }
}

/* Evaluate a derivative-coefficient vector on the grid: scaling, halving and
 * the length-2N-2 DFT of chebinvtranD1, chebyshev_fftw.F90:269-288 */
static void cheb_eval_beta(const chebctx* C, cplx* beta, cplx* out, cplx* w)
{
    const int N = C->N, M = C->M;
    const double fact = C->fact;
    for (int i = 0; i < N; i++) beta[i] = cscale(fact, beta[i]);
    beta[0] = cx(beta[0].re / 2.0, beta[0].im / 2.0);
    beta[N - 1] = cx(beta[N - 1].re / 2.0, beta[N - 1].im / 2.0);
    cplx* x = w;
    cplx* y = w + M;
    x[0] = beta[0];
    for (int i = 1; i < N; i++) x[i] = cx(beta[i].re / 2.0, beta[i].im / 2.0);
    for (int i = 1; i < N - 1; i++) x[N - 1 + i] = cx(beta[N - 1 - i].re / 2.0, beta[N - 1 - i].im / 2.0);
    fft_exec(&C->P, x, y, -1);
    for (int i = 0; i < N; i++) out[N - 1 - i] = y[i];
}

/* chebinvtranD1, chebyshev_fftw.F90:247-291.  w: 2*M + N cplx */
static void cheb_d1(const chebctx* C, const cplx* ft, cplx* df, cplx* w)
{
    cplx* beta = w + 2 * C->M;
    cheb_deriv_coef(C->N, ft, beta);
    cheb_eval_beta(C, beta, df, w);
}

/* chebinvtranD2, chebyshev_fftw.F90:293-352.  w: 2*M + 2N cplx */
static void cheb_d2(const chebctx* C, const cplx* ft, cplx* d2f, cplx* w)
{
    const int N = C->N;
    cplx* beta1 = w + 2 * C->M;
    cplx* beta2 = beta1 + N;
    cheb_deriv_coef(N, ft, beta1);
    cheb_deriv_coef(N, beta1, beta2);
    cheb_eval_beta(C, beta2, d2f, w);
}

/* chebinvtranD1D2, chebyshev_fftw.F90:354-427.  w: 2*M + 2N cplx */
static void cheb_d1d2(const chebctx* C, const cplx* ft, cplx* df, cplx* d2f, cplx* w)
{
    const int N = C->N;
    cplx* beta1 = w + 2 * C->M;
    cplx* beta2 = beta1 + N;
    cheb_deriv_coef(N, ft, beta1);
    cheb_deriv_coef(N, beta1, beta2);
    cheb_eval_beta(C, beta1, df, w);
    cheb_eval_beta(C, beta2, d2f, w);
}

void orc_chebtransform(int N, const double* f, double* ft)
{
    chebctx C; cheb_init(&C, N);
    cplx* w = (cplx*)malloc(sizeof(cplx) * (size_t)(2 * C.M + 2 * N));
    cheb_fwd(&C, (const cplx*)f, (cplx*)ft, w);
    free(w); plan_free(&C.P);
}

void orc_chebinvtran(int N, const double* ft, double* f)
{
    chebctx C; cheb_init(&C, N);
    cplx* w = (cplx*)malloc(sizeof(cplx) * (size_t)(2 * C.M + 2 * N));
    cheb_inv(&C, (const cplx*)ft, (cplx*)f, w);
    free(w); plan_free(&C.P);
}

void orc_chebinvtranD1(int N, const double* ft, double* df)
{
    chebctx C; cheb_init(&C, N);
    cplx* w = (cplx*)malloc(sizeof(cplx) * (size_t)(2 * C.M + 2 * N));
    cheb_d1(&C, (const cplx*)ft, (cplx*)df, w);
    free(w); plan_free(&C.P);
}

void orc_chebinvtranD2(int N, const double* ft, double* d2f)
{
    chebctx C; cheb_init(&C, N);
    cplx* w = (cplx*)malloc(sizeof(cplx) * (size_t)(2 * C.M + 2 * N));
    cheb_d2(&C, (const cplx*)ft, (cplx*)d2f, w);
    free(w); plan_free(&C.P);
}

/* radInt, integ.F90:36-76: Clenshaw-Curtis integral of f(r)*r over [rmin,rmax] */
static double radint_ctx(const chebctx* C, double eta, const double* f, cplx* w, cplx* ff, cplx* ft)
{
    const int N = C->N;
    const double pi = 4.0 * atan(1.0);
    const double rmin = eta / (1.0 - eta);
    for (int i = 0; i < N; i++) {
        const double rad = rmin + (0.5 * (1.0 + cos(((double)(N - 1 - i) * pi) / (double)(N - 1))));
        ff[i] = cx(f[i] * rad, 0.0);
    }
    cheb_fwd(C, ff, ft, w);
    double ftot = 0.0;
    for (int i = 0; i < N; i++) {
        const double Qm = (i % 2 == 0) ? -1.0 / (double)(i * i - 1) : 0.0;
        if (i == 0 || i == N - 1) ftot = ftot + 0.5 * ft[i].re * Qm;
        else ftot = ftot + ft[i].re * Qm;
    }
    return sqrt(2.0 / (double)(N - 1)) * ftot;
}

double orc_radint(int N, double eta, const double* f)
{
    chebctx C; cheb_init(&C, N);
    cplx* w = (cplx*)malloc(sizeof(cplx) * (size_t)(2 * C.M + 4 * N));
    double r = radint_ctx(&C, eta, f, w, w + 2 * C.M, w + 2 * C.M + N);
    free(w); plan_free(&C.P);
    return r;
}

/* ------------------------------------------------------------------------ */
/* Azimuthal real transforms, fourier_fftw.F90:39-97                          */
/* ------------------------------------------------------------------------ */
typedef struct { int Nm, Np; fftplan P; } fourctx;

/* forfft: f_m = (1/Np) sum_j x_j exp(-2 pi i j m / Np), m = 0..Nm. w: 2*Np */
static void four_fwd(const fourctx* F, const double* x, cplx* f, cplx* w)
{
    const int Np = F->Np;
    cplx* a = w;
    cplx* b = w + Np;
    for (int j = 0; j < Np; j++) a[j] = cx(x[j], 0.0);
    fft_exec(&F->P, a, b, -1);
    const double d = (double)Np;
    for (int m = 0; m <= F->Nm; m++) f[m] = cx(b[m].re / d, b[m].im / d);
}

/* invfft: unnormalised complex-to-real of the zero-padded half spectrum; the
 * imaginary part of f_0 is ignored as FFTW's c2r does. */
static void four_inv(const fourctx* F, const cplx* f, double* x, cplx* w)
{
    const int Np = F->Np, Nm = F->Nm;
    cplx* a = w;
    cplx* b = w + Np;
    for (int j = 0; j < Np; j++) a[j] = cx(0.0, 0.0);
    a[0] = cx(f[0].re, 0.0);
    for (int m = 1; m <= Nm; m++) {
        if (2 * m == Np) { a[m] = cx(f[m].re, 0.0); continue; }
        a[m] = f[m];
        a[Np - m] = cx(f[m].re, -f[m].im);
    }
    fft_exec(&F->P, a, b, +1);
    for (int j = 0; j < Np; j++) x[j] = b[j].re;
}

void orc_forfft(int Nm, int Np, const double* x, double* f)
{
    fourctx F; F.Nm = Nm; F.Np = Np; plan_init(&F.P, Np);
    cplx* w = (cplx*)malloc(sizeof(cplx) * (size_t)(2 * Np));
    four_fwd(&F, x, (cplx*)f, w);
    free(w); plan_free(&F.P);
}

void orc_invfft(int Nm, int Np, const double* f, double* x)
{
    fourctx F; F.Nm = Nm; F.Np = Np; plan_init(&F.P, Np);
    cplx* w = (cplx*)malloc(sizeof(cplx) * (size_t)(2 * Np));
    four_inv(&F, (const cplx*)f, x, w);
    free(w); plan_free(&F.P);
}

/* ------------------------------------------------------------------------ */
/* LU: dgetrf / dgetrs semantics (algebra_lapack.F90:14-53)                    */
/* ------------------------------------------------------------------------ */

/* Unblocked right-looking LU with partial pivoting (LAPACK dgetf2 order).
 * A: n x n column-major, overwritten by L (unit, below diag) and U.
 * ipiv: 0-based row interchanged with row j at step j.  Returns info (0 ok,
 * j+1 if U(j,j) is exactly zero; INFO is ignored by the reference). */
int orc_dgetrf(int n, double* A, int32_t* ipiv)
{
    int info = 0;
    const double sfmin = DBL_MIN;
    for (int j = 0; j < n; j++) {
        double* cj = A + (size_t)n * j;
        int p = j;
        double amax = fabs(cj[j]);
        for (int i = j + 1; i < n; i++) {
            const double v = fabs(cj[i]);
            if (v > amax) { amax = v; p = i; }
        }
        ipiv[j] = p;
        if (cj[p] != 0.0) {
            if (p != j)
                for (int k = 0; k < n; k++) {
                    double* c = A + (size_t)n * k;
                    const double tmp = c[j]; c[j] = c[p]; c[p] = tmp;
                }
            if (fabs(cj[j]) >= sfmin) {
                const double r = 1.0 / cj[j];
                for (int i = j + 1; i < n; i++) cj[i] = cj[i] * r;
            } else {
                for (int i = j + 1; i < n; i++) cj[i] = cj[i] / cj[j];
            }
        } else if (info == 0) {
            info = j + 1;
        }
        for (int k = j + 1; k < n; k++) {
            double* ck = A + (size_t)n * k;
            const double u = ck[j];
            for (int i = j + 1; i < n; i++) ck[i] = fma(-cj[i], u, ck[i]);
        }
    }
    return info;
}

/* dgetrs('N', n, 1, ...): interchanges, unit-lower solve, upper solve; each
 * b(i) accumulates column updates in column order. */
void orc_dgetrs(int n, const double* A, const int32_t* ipiv, double* b)
{
    for (int k = 0; k < n; k++) {
        const int p = ipiv[k];
        if (p != k) { const double t = b[k]; b[k] = b[p]; b[p] = t; }
    }
    for (int k = 0; k < n; k++) {
        const double x = b[k];
        const double* c = A + (size_t)n * k;
        for (int i = k + 1; i < n; i++) b[i] = fma(-x, c[i], b[i]);
    }
    for (int k = n - 1; k >= 0; k--) {
        const double* c = A + (size_t)n * k;
        b[k] = b[k] / c[k];
        const double x = b[k];
        for (int i = 0; i < k; i++) b[i] = fma(-x, c[i], b[i]);
    }
}

/* ------------------------------------------------------------------------ */
/* Operator matrices, mat_assembly.F90                                        */
/* ------------------------------------------------------------------------ */
typedef struct {
    int N;
    double *t, *D, *D2, *xp;
    double *radius, *r_radius, *r_radius2;
} chebgrid;

/* mat_build (:242-359, IMEX BC row order) and mat_build_IRK (:68-187, IMEXRK
 * order: the psi' rows first).  wt_dt = wt_lhs*dt.  Any output may be NULL.
 * Column-major, AT: N x N, AF: 2N x 2N. */
static void mat_build_g(const chebgrid* G, int n, double wt, double dt, double Pr, int mBC,
                        int irk_order, double* AT, double* AF)
{
    const int N = G->N;
    const double C = sqrt(2.0 / (double)(N - 1));
    const double* t = G->t;
    const double* D = G->D;
    const double* D2 = G->D2;
    const double* rr = G->r_radius;
    const double* rr2 = G->r_radius2;
    const double nn = (double)n;
#define IX(i, j, ld) ((size_t)(i) + (size_t)(ld) * (size_t)(j))
    if (AT) {
        for (int j = 0; j < N; j++) {
            for (int i = 0; i < N; i++)
                AT[IX(i, j, N)] = t[IX(i, j, N)] - ((wt * dt) * (1.0 / Pr)) *
                    ((rr[i] * D[IX(i, j, N)] + D2[IX(i, j, N)]) - ((nn * nn) * rr2[i]) * t[IX(i, j, N)]);
            AT[IX(0, j, N)] = t[IX(0, j, N)];
            AT[IX(N - 1, j, N)] = t[IX(N - 1, j, N)];
        }
        for (int i = 0; i < N; i++) {
            AT[IX(i, 0, N)] = 0.5 * AT[IX(i, 0, N)];
            AT[IX(i, N - 1, N)] = 0.5 * AT[IX(i, N - 1, N)];
        }
        for (size_t k = 0; k < (size_t)N * N; k++) AT[k] = C * AT[k];
    }
    if (AF) {
        const int L = 2 * N;
        double bc1[2], bc2[2];
        for (int j = 0; j < N; j++) {
            for (int i = 0; i < N; i++) {
                const double lap = (rr[i] * D[IX(i, j, N)] + D2[IX(i, j, N)]) - ((nn * nn) * rr2[i]) * t[IX(i, j, N)];
                AF[IX(i, j, L)] = lap;
                AF[IX(i, N + j, L)] = t[IX(i, j, N)];
                AF[IX(N + i, j, L)] = 0.0;
                AF[IX(N + i, N + j, L)] = t[IX(i, j, N)] - (wt * dt) * lap;
            }
            bc1[0] = t[IX(0, j, N)];
            bc1[1] = t[IX(N - 1, j, N)];
            if (mBC == 0) {
                bc2[0] = D[IX(0, j, N)];
                bc2[1] = D[IX(N - 1, j, N)];
            } else {
                bc2[0] = D2[IX(0, j, N)] - rr[0] * D[IX(0, j, N)];
                bc2[1] = D2[IX(N - 1, j, N)] - rr[N - 1] * D[IX(N - 1, j, N)];
            }
            const double* top = irk_order ? bc2 : bc1;   /* rows 1 and N   */
            const double* bot = irk_order ? bc1 : bc2;   /* rows N+1 and 2N */
            AF[IX(0, j, L)] = top[0];
            AF[IX(N - 1, j, L)] = top[1];
            AF[IX(0, N + j, L)] = 0.0;
            AF[IX(N - 1, N + j, L)] = 0.0;
            AF[IX(N, j, L)] = bot[0];
            AF[IX(2 * N - 1, j, L)] = bot[1];
            AF[IX(N, N + j, L)] = 0.0;
            AF[IX(2 * N - 1, N + j, L)] = 0.0;
        }
        for (int i = 0; i < N; i++) {
            const int cols[4] = {0, N - 1, N, 2 * N - 1};
            for (int c = 0; c < 4; c++) {
                AF[IX(i, cols[c], L)] = 0.5 * AF[IX(i, cols[c], L)];
                AF[IX(N + i, cols[c], L)] = 0.5 * AF[IX(N + i, cols[c], L)];
            }
        }
        for (size_t k = 0; k < (size_t)L * L; k++) AF[k] = C * AF[k];
    }
}

/* mat_build_uphibar (:361-413) / mat_build_uphibar_IRK (:189-240): identical */
static void mat_build_uphibar_g(const chebgrid* G, double wt, double dt, int mBC, double* A)
{
    const int N = G->N;
    const double C = sqrt(2.0 / (double)(N - 1));
    const double* t = G->t;
    const double* D = G->D;
    const double* D2 = G->D2;
    const double* rr = G->r_radius;
    const double* rr2 = G->r_radius2;
    for (int j = 0; j < N; j++) {
        for (int i = 0; i < N; i++)
            A[IX(i, j, N)] = t[IX(i, j, N)] - (wt * dt) * ((rr[i] * D[IX(i, j, N)] + D2[IX(i, j, N)]) - rr2[i] * t[IX(i, j, N)]);
        if (mBC == 0) {
            A[IX(0, j, N)] = t[IX(0, j, N)];
            A[IX(N - 1, j, N)] = t[IX(N - 1, j, N)];
        } else {
            A[IX(0, j, N)] = D[IX(0, j, N)] - rr[0] * t[IX(0, j, N)];
            A[IX(N - 1, j, N)] = D[IX(N - 1, j, N)] - rr[N - 1] * t[IX(N - 1, j, N)];
        }
    }
    for (int i = 0; i < N; i++) {
        A[IX(i, 0, N)] = 0.5 * A[IX(i, 0, N)];
        A[IX(i, N - 1, N)] = 0.5 * A[IX(i, N - 1, N)];
    }
    for (size_t k = 0; k < (size_t)N * N; k++) A[k] = C * A[k];
}

static int grid_init(chebgrid* G, int N, double eta)
{
    G->N = N;
    G->t = (double*)malloc(sizeof(double) * (size_t)N * N);
    G->D = (double*)malloc(sizeof(double) * (size_t)N * N);
    G->D2 = (double*)malloc(sizeof(double) * (size_t)N * N);
    G->xp = (double*)malloc(sizeof(double) * (size_t)N);
    G->radius = (double*)malloc(sizeof(double) * (size_t)N);
    G->r_radius = (double*)malloc(sizeof(double) * (size_t)N);
    G->r_radius2 = (double*)malloc(sizeof(double) * (size_t)N);
    if (!G->t || !G->D || !G->D2 || !G->xp || !G->radius || !G->r_radius || !G->r_radius2) return -1;
    orc_dcheb(N, G->t, G->D, G->D2, G->xp);
    orc_grid(N, eta, G->radius, G->r_radius, G->r_radius2);
    return 0;
}

static void grid_free(chebgrid* G)
{
    free(G->t); free(G->D); free(G->D2); free(G->xp);
    free(G->radius); free(G->r_radius); free(G->r_radius2);
}

void orc_mat_build(int N, double eta, int n, double wt_dt, double Pr, int mBC, int irk_order,
                   double* AT, double* AF, double* Auphi)
{
    chebgrid G;
    grid_init(&G, N, eta);
    mat_build_g(&G, n, wt_dt, 1.0, Pr, mBC, irk_order, AT, AF);
    if (Auphi) mat_build_uphibar_g(&G, wt_dt, 1.0, mBC, Auphi);
    grid_free(&G);
}

/* ------------------------------------------------------------------------ */
/* Scheme registry, timeschemes.F90                                           */
/* ------------------------------------------------------------------------ */
enum { ST_IMEX = 0, ST_IMEXRK = 1, ST_RK = 2 };
enum { IMP_CN = 0, IMP_BDF2, IMP_BDF3, IMP_BDF4 };
#define MAXS 8

typedef struct {
    int type;
    int imp;              /* IMEX implicit family */
    int n_imp, n_exp;     /* history depths (IMEX) or stage counts (IMEXRK) */
    char expname[16];
    /* IMEXRK */
    int S;
    double aD[MAXS * MAXS], aA[MAXS * MAXS], bD[MAXS], bA[MAXS];
    int diag_index;       /* 1-based, as the reference */
    int diag_diff;        /* stages solved with their own diagonal's operator */
    int sa;               /* both ars_eqn_check flags set */
    int nl_col[MAXS];     /* 1 if sum(aA(:,k)) + bA(k) /= 0 */
} scheme;

#define AD(s, i) aD[(s) * MAXS + (i)]
#define AA(s, i) aA[(s) * MAXS + (i)]

static void set_tab(double* dst, const double* rows, int S)
{
    for (int i = 0; i < S; i++)
        for (int j = 0; j < S; j++) dst[i * MAXS + j] = rows[i * S + j];
}

/* The stiffly accurate IMEXRK tableaux with diag_diff = .FALSE. and
 * diag_index = 2 (rhs_update_wts_imp / rhs_update_wts_exp of timeschemes.F90,
 * implicit / explicit case at the lines given): published constants, entered
 * row by row, with the reference's own arithmetic for the irrational ones. */
typedef struct {
    const char* name;
    int S;
    void (*fill)(double* aD, double* bD, double* aA, double* bA);
} rk_tableau;

static void tab_copy(double* dst, const double* src, int n) { memcpy(dst, src, sizeof(double) * (size_t)n); }

/* ARS222 (:1066 / :1741) */
static void fill_ars222(double* aD, double* bD, double* aA, double* bA)
{
    const double g = (2.0 - sqrt(2.0)) / 2.0, d = 1.0 - 1.0 / g / 2.0;
    const double D[9] = {0, 0, 0, 0, g, 0, 0, 1.0 - g, g};
    const double A[9] = {0, 0, 0, g, 0, 0, d, 1.0 - d, 0};
    tab_copy(aD, D, 9); tab_copy(aA, A, 9);
    tab_copy(bD, D + 6, 3); tab_copy(bA, A + 6, 3);
}
/* PC2 (:1168 / :1866) */
static void fill_pc2(double* aD, double* bD, double* aA, double* bA)
{
    const double D[16] = {0, 0, 0, 0, 0.5, 0.5, 0, 0, 0.5, 0, 0.5, 0, 0.5, 0, 0, 0.5};
    const double A[16] = {0, 0, 0, 0, 1.0, 0, 0, 0, 0.5, 0.5, 0, 0, 0.5, 0, 0.5, 0};
    tab_copy(aD, D, 16); tab_copy(aA, A, 16);
    tab_copy(bD, D + 12, 4); tab_copy(bA, A + 12, 4);
}
/* SSP332 (:1183 / :1884) */
static void fill_ssp332(double* aD, double* bD, double* aA, double* bA)
{
    const double D[16] = {0, 0, 0, 0, 0, 0.5, 0, 0, 0, 0.5, 0.5, 0, 0, 1.0, -0.5, 0.5};
    const double A[16] = {0, 0, 0, 0, 0.5, 0, 0, 0, 0.5, 0.5, 0, 0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 0};
    tab_copy(aD, D, 16); tab_copy(aA, A, 16);
    tab_copy(bD, D + 12, 4); tab_copy(bA, A + 12, 4);
}
/* SSP543 (:1198 / :1902) */
static void fill_ssp543(double* aD, double* bD, double* aA, double* bA)
{
    const double D[36] = {0, 0, 0, 0, 0, 0,
                          0, 0.5, 0, 0, 0, 0,
                          0, 0.5, 0.5, 0, 0, 0,
                          0, 0.5, -0.5, 0.5, 0, 0,
                          0, 2.0, 0.5, -3.0, 0.5, 0,
                          0, 2.0 / 3.0, -1.0 / 3.0, 0, 1.0 / 6.0, 0.5};
    const double A[36] = {0, 0, 0, 0, 0, 0,
                          0.5, 0, 0, 0, 0, 0,
                          0.5, 0.5, 0, 0, 0, 0,
                          1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 0, 0, 0,
                          0, 0, 0, 0, 0, 0,
                          1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 0.5, 0, 0};
    tab_copy(aD, D, 36); tab_copy(aA, A, 36);
    tab_copy(bD, D + 30, 6); tab_copy(bA, A + 30, 6);
}
/* ARS443 (:1215 / :1919) */
static void fill_ars443(double* aD, double* bD, double* aA, double* bA)
{
    const double D[25] = {0, 0, 0, 0, 0,
                          0, 0.5, 0, 0, 0,
                          0, 1.0 / 6.0, 0.5, 0, 0,
                          0, -0.5, 0.5, 0.5, 0,
                          0, 1.5, -1.5, 0.5, 0.5};
    const double A[25] = {0, 0, 0, 0, 0,
                          0.5, 0, 0, 0, 0,
                          11.0 / 18.0, 1.0 / 18.0, 0, 0, 0,
                          5.0 / 6.0, -5.0 / 6.0, 0.5, 0, 0,
                          0.25, 7.0 / 4.0, 0.75, -7.0 / 4.0, 0};
    tab_copy(aD, D, 25); tab_copy(aA, A, 25);
    tab_copy(bD, D + 20, 5); tab_copy(bA, A + 20, 5);
}
/* BPR353 (:1277 / :1995) */
static void fill_bpr353(double* aD, double* bD, double* aA, double* bA)
{
    const double D[25] = {0, 0, 0, 0, 0,
                          0.5, 0.5, 0, 0, 0,
                          5.0 / 18.0, -1.0 / 9.0, 0.5, 0, 0,
                          0.5, 0, 0, 0.5, 0,
                          0.25, 0, 0.75, -0.5, 0.5};
    const double A[25] = {0, 0, 0, 0, 0,
                          1.0, 0, 0, 0, 0,
                          4.0 / 9.0, 2.0 / 9.0, 0, 0, 0,
                          0.25, 0, 0.75, 0, 0,
                          0.25, 0, 0.75, 0, 0};
    tab_copy(aD, D, 25); tab_copy(aA, A, 25);
    tab_copy(bD, D + 20, 5); tab_copy(bA, A + 20, 5);
}
/* FW53 (:1410 / :2151): the diagonal is not uniform; with diag_diff = .FALSE.
 * every stage is solved with the aD(2,2) operator, as the reference does */
static void fill_fw53(double* aD, double* bD, double* aA, double* bA)
{
    const double D[25] = {0, 0, 0, 0, 0,
                          0, 1.0, 0, 0, 0,
                          0.5, 0, 0.5, 0, 0,
                          0.25, -0.25, 0, 0.5, 0,
                          1.0 / 6.0, 1.0 / 6.0, -1.0 / 6.0, 4.0 / 6.0, 1.0 / 6.0};
    const double A[25] = {0, 0, 0, 0, 0,
                          1.0, 0, 0, 0, 0,
                          0.5, 0.5, 0, 0, 0,
                          0.25, 0.25, 0, 0, 0,
                          1.0 / 6.0, 1.0 / 6.0, 0, 4.0 / 6.0, 0};
    tab_copy(aD, D, 25); tab_copy(aA, A, 25);
    tab_copy(bD, D + 20, 5); tab_copy(bA, A + 20, 5);
}
/* CK222 (:1430 / :2173) */
static void fill_ck222(double* aD, double* bD, double* aA, double* bA)
{
    const double r = sqrt(2.0), third = 1.0 / 3.0, tq = 3.0 / 4.0;
    const double D[9] = {0, 0, 0,
                         -third + r * 0.5, 1.0 - r * 0.5, 0,
                         tq - r * 0.25, -tq + r * 0.75, 1.0 - r * 0.5};
    const double A[9] = {0, 0, 0, 2.0 / 3.0, 0, 0, 0.25, 0.75, 0};
    tab_copy(aD, D, 9); tab_copy(aA, A, 9);
    tab_copy(bD, D + 6, 3); tab_copy(bA, A + 6, 3);
}
/* BPR442 (:1449 / :2194) */
static void fill_bpr442(double* aD, double* bD, double* aA, double* bA)
{
    const double D[25] = {0, 0, 0, 0, 0,
                          0, 0.25, 0, 0, 0,
                          0, 0, 0.25, 0, 0,
                          0, 1.0 / 24.0, 11.0 / 24.0, 0.25, 0,
                          0, 11.0 / 24.0, 1.0 / 6.0, 1.0 / 8.0, 0.25};
    const double A[25] = {0, 0, 0, 0, 0,
                          0.25, 0, 0, 0, 0,
                          13.0 / 4.0, -3.0, 0, 0, 0,
                          0.25, 0, 0.5, 0, 0,
                          0, 1.0 / 3.0, 1.0 / 6.0, 0.5, 0};
    tab_copy(aD, D, 25); tab_copy(aA, A, 25);
    tab_copy(bD, D + 20, 5); tab_copy(bA, A + 20, 5);
}

static const rk_tableau RK_TABLEAUX[] = {
    {"ARS222", 3, fill_ars222}, {"PC2", 4, fill_pc2},       {"SSP332", 4, fill_ssp332},
    {"SSP543", 6, fill_ssp543}, {"ARS443", 5, fill_ars443}, {"BPR353", 5, fill_bpr353},
    {"FW53", 5, fill_fw53},     {"CK222", 3, fill_ck222},   {"BPR442", 5, fill_bpr442},
};

/* The other IMEXRK tableaux, as doubles (row-major aD, bD, aA, bA): the
 * reference's constant expressions evaluated in double arithmetic, printed to
 * 17 digits.  ckA / ckD are the ars_eqn_check flags the reference ends up with
 * (set literally, or by its any-element loop): both set, the step ends with
 * Assembly_stage_SA, otherwise with Assembly_stage.  diag_diff / diag_index
 * as the reference sets them. */
typedef struct {
    const char* name;
    int S, ckA, ckD, diag_diff, diag_index;
    double aD[MAXS * MAXS], bD[MAXS], aA[MAXS * MAXS], bA[MAXS];
} rk_numeric;

static const rk_numeric RK_NUMERIC[] = {
    /* EUL111: rhs_update_wts_imp :1055, rhs_update_wts_exp :1727, init_time_schemes :190 */
    {"EUL111", 2, 0, 0, 0, 2,
     {0, 0,
      0, 1},
     {0, 1},
     {0, 0,
      1, 0},
     {1, 0}
    },
    /* ARS232: rhs_update_wts_imp :1080, rhs_update_wts_exp :1758, init_time_schemes :236 */
    {"ARS232", 3, 0, 1, 0, 2,
     {0, 0, 0,
      0, 0.29289321881345243, 0,
      0, 0.70710678118654757, 0.29289321881345243},
     {0, 0.70710678118654757, 0.29289321881345243},
     {0, 0, 0,
      0.29289321881345243, 0, 0,
      -0.94280904158206347, 1.9428090415820636, 0},
     {0, 0.70710678118654757, 0.29289321881345243}
    },
    /* ARS233: rhs_update_wts_imp :1094, rhs_update_wts_exp :1775, init_time_schemes :260 */
    {"ARS233", 3, 0, 0, 0, 2,
     {0, 0, 0,
      0, 0.78867513459481275, 0,
      0, -0.57735026918962551, 0.78867513459481275},
     {0, 0.5, 0.5},
     {0, 0, 0,
      0.78867513459481275, 0, 0,
      -0.21132486540518725, 0.42264973081037449, 0},
     {0, 0.5, 0.5}
    },
    /* ARS343: rhs_update_wts_imp :1107, rhs_update_wts_exp :1792, init_time_schemes :283 */
    {"ARS343", 4, 0, 1, 0, 2,
     {0, 0, 0, 0,
      0, 0.435866521508459, 0, 0,
      0, 0.28206673924577053, 0.435866521508459, 0,
      0, 1.2084966491760101, -0.64436317068446924, 0.435866521508459},
     {0, 1.2084966491760101, -0.64436317068446924, 0.435866521508459},
     {0, 0, 0, 0,
      0.435866521508459, 0, 0, 0,
      0.32127888602862781, 0.39665437472560172, 0, 0,
      -0.1058582960718797, 0.55292914803593984, 0.55292914803593984, 0},
     {0, 1.2084966491760101, -0.64436317068446924, 0.435866521508459}
    },
    /* MARS343: rhs_update_wts_imp :1139, rhs_update_wts_exp :1830, init_time_schemes :341 */
    {"MARS343", 4, 0, 1, 0, 2,
     {0, 0, 0, 0,
      0, 0.435866521508458, 0, 0,
      0, 0.28206673924577003, 0.435866521508458, 0,
      0, 1.2084966491760101, -0.64436317068446003, 0.435866521508458},
     {0, 1.2084966491760101, -0.64436317068446003, 0.435866521508458},
     {0, 0, 0, 0,
      0.435866521508458, 0, 0, 0,
      0.53539654030735395, 0.18253672044687499, 0, 0,
      0.63041255815287001, -0.83193390106307996, 1.2015213429102101, 0},
     {0, 1.2084966491760101, -0.64436317068446003, 0.435866521508458}
    },
    /* IJ3: rhs_update_wts_imp :1154, rhs_update_wts_exp :1848, init_time_schemes :363 */
    {"IJ3", 4, 0, 0, 0, 2,
     {0, 0, 0, 0,
      0, 0.78868665109985225, 0, 0,
      0, -0.45533028521586527, 0.78866361854919842, 0,
      0, 1.3659908556475957, -0.36599085564759576, 0},
     {0, 0, 0.75, 0.25},
     {0, 0, 0, 0,
      0, 0, 0, 0,
      0, 0.33333333333333331, 0, 0,
      0, -1, 2, 0},
     {0, 0, 0.75, 0.25}
    },
    /* PR433: rhs_update_wts_imp :1235, rhs_update_wts_exp :1946, init_time_schemes :474 */
    {"PR433", 5, 0, 0, 0, 2,
     {0, 0, 0, 0, 0,
      0, 0.24169426078820999, 0, 0, 0,
      0, -0.24169426078820999, 0.24169426078820999, 0, 0,
      0, 0, 0.75830573921179001, 0.24169426078820999, 0,
      0, 0.06042356519705, 0.1291528696059, 0.068729304408840008, 0.24169426078820999},
     {0, 0, 0.16666666666666666, 0.16666666666666666, 0.66666666666666663},
     {0, 0, 0, 0, 0,
      0, 0, 0, 0, 0,
      0, 0, 0, 0, 0,
      0, 0, 1, 0, 0,
      0, 0, 0.25, 0.25, 0},
     {0, 0, 0.16666666666666666, 0.16666666666666666, 0.66666666666666663}
    },
    /* DBM453: rhs_update_wts_imp :1258, rhs_update_wts_exp :1978, init_time_schemes :499 */
    {"DBM453", 5, 0, 1, 0, 2,
     {0, 0, 0, 0, 0,
      -0.2228498531852541, 0.32591194130117246, 0, 0, 0,
      -0.46801347074080546, 0.8634928422571696, 0.32591194130117246, 0, 0,
      -0.46509906651927418, 0.81063103116959556, 0.6103672675683236, 0.32591194130117246, 0,
      0.87795339639076675, -0.72692641526151547, 0.75204137157372719, -0.22898029400415087, 0.32591194130117246},
     {0.87795339639076675, -0.72692641526151547, 0.75204137157372719, -0.2289802940041509, 0.32591194130117246},
     {0, 0, 0, 0, 0,
      0.10306208811591838, 0, 0, 0, 0,
      -0.94124866143519892, 1.6626399742527356, 0, 0, 0,
      -1.3670975201437765, 1.3815852911016873, 1.2673234025619065, 0, 0,
      -0.81287582068772446, 0.8122373906050574, 0.90644429603699306, 0.094194134045674116, 0},
     {0.87795339639076675, -0.72692641526151547, 0.75204137157372719, -0.2289802940041509, 0.32591194130117246}
    },
    /* KC343: rhs_update_wts_imp :1297, rhs_update_wts_exp :2022, init_time_schemes :544 */
    {"KC343", 4, 0, 1, 0, 2,
     {0, 0, 0, 0,
      0.435866521508459, 0.435866521508459, 0, 0,
      0.25764824606642722, -0.093514767574886248, 0.435866521508459, 0,
      0.18764102434672383, -0.59529747357695495, 0.97178992772177208, 0.435866521508459},
     {0.18764102434672383, -0.59529747357695495, 0.97178992772177208, 0.435866521508459},
     {0, 0, 0, 0,
      0.87173304301691801, 0, 0, 0,
      0.52758901197630037, 0.072410988023699593, 0, 0,
      0.39909600767607012, -0.43755765461351942, 1.0384616469374492, 0},
     {0.18764102434672383, -0.59529747357695495, 0.97178992772177208, 0.435866521508459}
    },
    /* KC564: rhs_update_wts_imp :1323, rhs_update_wts_exp :2052, init_time_schemes :566 */
    {"KC564", 6, 0, 1, 0, 2,
     {0, 0, 0, 0, 0, 0,
      0.25, 0.25, 0, 0, 0, 0,
      0.13777600000000001, -0.055775999999999999, 0.25, 0, 0, 0,
      0.14463686602698217, -0.22393190761334475, 0.44929504158636258, 0.25, 0, 0,
      0.098258783283564771, -0.59154424281967044, 0.81012105382829958, 0.28316440570780599, 0.25, 0,
      0.15791629516167136, 0, 0.18675894052400077, 0.68056529530933463, -0.27524053099500667, 0.25},
     {0.15791629516167136, 0, 0.18675894052400077, 0.68056529530933463, -0.27524053099500667, 0.25},
     {0, 0, 0, 0, 0, 0,
      0.5, 0, 0, 0, 0, 0,
      0.221776, 0.110224, 0, 0, 0, 0,
      -0.04884659515311858, -0.177720652326401, 0.84656724747951961, 0, 0, 0,
      -0.15541685842491548, -0.3567050098221991, 1.0587258798684427, 0.30339598837867193, 0, 0,
      0.20142435067267633, 0.0087420578429041849, 0.15993995707168115, 0.40382906052207751, 0.22606457389066084, 0},
     {0.15791629516167136, 0, 0.18675894052400077, 0.68056529530933463, -0.27524053099500667, 0.25}
    },
    /* KC774: rhs_update_wts_imp :1342, rhs_update_wts_exp :2076, init_time_schemes :588 */
    {"KC774", 7, 0, 1, 0, 2,
     {0, 0, 0, 0, 0, 0, 0,
      0.1235, 0.1235, 0, 0, 0, 0, 0,
      0.14907768747653863, 0.14907768747653863, 0.1235, 0, 0, 0, 0,
      0.12483442871739439, 0.12483442871739439, -0.038168857434788782, 0.1235, 0, 0, 0,
      -0.073031940302180909, -0.073031940302180909, -0.24343568716014671, 0.34099956776450852, 0.1235, 0, 0,
      -0.15296500088128806, -0.15296500088128806, 0.072205620474335874, 0.40430630248551713, 0.40591807880272318, 0.1235, 0,
      0, 0, 0.51611072831742366, -0.14606356393857081, 0.23473048589019332, 0.27172234973095377, 0.1235},
     {0, 0, 0.51611072831742366, -0.14606356393857081, 0.23473048589019332, 0.27172234973095377, 0.1235},
     {0, 0, 0, 0, 0, 0, 0,
      0.247, 0, 0, 0, 0, 0, 0,
      0.061749999999999999, 0.35990537495307723, 0, 0, 0, 0, 0,
      0.05301658458687121, 0.35949264529328429, -0.077509229880155461, 0, 0, 0, 0,
      0.058417159447800002, -0.16313824817772324, -0.19732090979798411, 0.37704199852790737, 0, 0, 0,
      0.53853032270810797, -0.45497746895916669, 1.2562905623429941, -0.47828452721130055, -0.16155888888063494, 0, 0,
      0.23221715782277083, 0.23221715782277083, -6.809994375038098, 7.3618585524244216, -1.3748790779406981, 1.3585805849088326, 0},
     {0, 0, 0.51611072831742366, -0.14606356393857081, 0.23473048589019332, 0.27172234973095377, 0.1235}
    },
    /* KC885: rhs_update_wts_imp :1373, rhs_update_wts_exp :2110, init_time_schemes :610 */
    {"KC885", 8, 0, 1, 0, 2,
     {0, 0, 0, 0, 0, 0, 0, 0,
      0.22222222222222221, 0.22222222222222221, 0, 0, 0, 0, 0, 0,
      0.26824595137478835, 0.26824595137478835, 0.22222222222222221, 0, 0, 0, 0, 0,
      -0.057945592237231995, -0.057945592237231995, 0.0089383968162837328, 0.22222222222222221, 0, 0, 0, 0,
      -0.043305287723547685, -0.043305287723547685, -0.034013891077568637, 0.25515937270676026, 0.22222222222222221, 0, 0, 0,
      0.13179599023759678, 0.13179599023759678, -0.032376726277862332, 0.12385474427672251, 0.14270777930372408, 0.22222222222222221, 0, 0,
      0.30932282100434261, 0.30932282100434261, -0.68291992723367922, -0.058822756149695461, -0.041308613833499437, 0.89718343298596659, 0.22222222222222221, 0,
      0, 0, 0.17366253573581261, 0.25479166260812353, 0.24190176845094791, 0.30740485830222825, -0.19998304731933453, 0.22222222222222221},
     {0, 0, 0.17366253573581261, 0.25479166260812353, 0.24190176845094791, 0.30740485830222825, -0.19998304731933453, 0.22222222222222221},
     {0, 0, 0, 0, 0, 0, 0, 0,
      0.44444444444444442, 0, 0, 0, 0, 0, 0, 0,
      0.1111111111111111, 0.64760301386068775, 0, 0, 0, 0, 0, 0,
      0.091829866647747907, 0.035448567517799241, -0.012008999601505184, 0, 0, 0, 0, 0,
      -0.34252354516023137, -0.26767785943050182, 0.11056894178117282, 0.85638959121387881, 0, 0, 0, 0,
      -0.0097722828790043955, 0.21070865398661751, 0.075924120912175361, 0.20765518596381696, 0.23548432201639455, 0, 0, 0,
      0.46686370681500694, 1.2903598800650855, 0.37840596884419414, -0.56345584032826157, -0.2883238346202236, -0.32884988077580141, 0, 0,
      0.61439671625166914, 0.61439671625166914, 0.31747780106686158, -0.71215206239529361, 0.11498708015310211, 0.09139031575415682, -0.040496567082165244, 0},
     {0, 0, 0.17366253573581261, 0.25479166260812353, 0.24190176845094791, 0.30740485830222825, -0.19998304731933453, 0.22222222222222221}
    },
    /* CFN343: rhs_update_wts_imp :1465, rhs_update_wts_exp :2217, init_time_schemes :698 */
    {"CFN343", 4, 0, 1, 0, 2,
     {0, 0, 0, 0,
      0, 0.435866521508459, 0, 0,
      0, 0.28206673924577053, 0.435866521508459, 0,
      0, 1.2084966491760101, -0.64436317068446902, 0.435866521508459},
     {0, 1.2084966491760101, -0.64436317068446902, 0.435866521508459},
     {0, 0, 0, 0,
      0.435866521508459, 0, 0, 0,
      1.0679332607542293, -0.34999999999999998, 0, 0,
      0, 1.989175724679846, -0.98917572467984605, 0},
     {0, 1.2084966491760101, -0.64436317068446902, 0.435866521508459}
    },
    /* BHR553: rhs_update_wts_imp :1480, rhs_update_wts_exp :2235, init_time_schemes :721 */
    {"BHR553", 5, 0, 1, 0, 2,
     {0, 0, 0, 0, 0,
      0.43586652150848199, 0.43586652150848199, 0, 0, 0,
      0.43586652150848199, -6.9635084214665835e-14, 0.43586652150848199, 0, 0,
      -0.066758687019583762, -6.1296596594174833e-13, 1.9711047406203948, 0.43586652150848199, 0,
      0.41289804281247428, 0, 0.19733989037886918, -0.04610445469982568, 0.43586652150848199},
     {0.41289804281247428, 0, 0.19733989037886918, -0.04610445469982568, 0.43586652150848199},
     {0, 0, 0, 0, 0,
      0.87173304301696397, 0, 0, 0, 0,
      0.43586652150848199, 0.43586652150848199, 0, 0, 0,
      -0.80099845306562878, 0, 3.1412110281743093, 0, 0,
      0.35675320777963954, -0.19733989037886918, 0.88194884139379082, -0.041362158794624231, 0},
     {0.41289804281247428, 0, 0.19733989037886918, -0.04610445469982568, 0.43586652150848199}
    },
    /* CFN564: rhs_update_wts_imp :1575, rhs_update_wts_exp :2349, init_time_schemes :918 */
    {"CFN564", 6, 0, 1, 0, 2,
     {0, 0, 0, 0, 0, 0,
      0, 0.25, 0, 0, 0, 0,
      0, 0.5, 0.25, 0, 0, 0,
      0, 0.34000000000000002, -0.040000000000000001, 0.25, 0, 0,
      0, 0.2727941176470588, -0.050367647058823531, 0.027573529411764705, 0.25, 0,
      0, 1.0416666666666667, -1.0208333333333333, 7.8125, -7.083333333333333, 0.25},
     {0, 1.0416666666666667, -1.0208333333333333, 7.8125, -7.083333333333333, 0.25},
     {0, 0, 0, 0, 0, 0,
      0.25, 0, 0, 0, 0, 0,
      -0.25, 1, 0, 0, 0, 0,
      -0.13, 0.57333333333333336, 0.10666666666666667, 0, 0, 0,
      -0.070588235294117646, 0.49411764705882355, 0.13161764705882353, -0.05514705882352941, 0, 0,
      0, 3.2916666666666665, -0.625, 12.5, -14.166666666666666, 0},
     {0, 1.0416666666666667, -1.0208333333333333, 7.8125, -7.083333333333333, 0.25}
    },
    /* stages solved with their own diagonal (diag_diff = .TRUE.,
     * fourierloop_imexrk.F90:134-158) */
    /* SMR333: rhs_update_wts_imp :1122, rhs_update_wts_exp :1810, init_time_schemes :308 */
    {"SMR333", 4, 1, 1, 1, 2,
     {0, 0, 0, 0,
      0.30208333333333331, 0.23125000000000001, 0, 0,
      0.30208333333333331, 0.15625, 0.20833333333333334, 0,
      0.30208333333333331, 0.15625, 0.375, 0.16666666666666666},
     {0.30208333333333331, 0.15625, 0.375, 0.16666666666666666},
     {0, 0, 0, 0,
      0.53333333333333333, 0, 0, 0,
      0.25, 0.41666666666666669, 0, 0,
      0.25, 0, 0.75, 0},
     {0.25, 0, 0.75, 0}
    },
    /* CB3e: rhs_update_wts_imp :1496, rhs_update_wts_exp :2254, init_time_schemes :765 */
    {"CB3e", 4, 0, 1, 1, 3,
     {0, 0, 0, 0,
      0, 0.33333333333333331, 0, 0,
      0, 0.5, 0.5, 0,
      0, 0.75, -0.25, 0.5},
     {0, 0.75, -0.25, 0.5},
     {0, 0, 0, 0,
      0.33333333333333331, 0, 0, 0,
      0, 1, 0, 0,
      0, 0.75, 0.25, 0},
     {0, 0.75, -0.25, 0.5}
    },
    /* CB3f: rhs_update_wts_imp :1511, rhs_update_wts_exp :2272, init_time_schemes :787 */
    {"CB3f", 4, 0, 1, 1, 2,
     {0, 0, 0, 0,
      0.48999999999999999, 0.48999999999999999, 0, 0,
      -0.71803538544242407, -0.031405761082656344, 0.78944114652508046, 0,
      -3.6143737489279602, 0.11126152710363123, 4.2834880399852997, 0.21962418183902943},
     {-3.6143737489279602, 0.11126152710363123, 4.2834880399852997, 0.21962418183902943},
     {0, 0, 0, 0,
      0.97999999999999998, 0, 0, 0,
      0.020449687951736302, 0.019550312048263699, 0, 0,
      -3.6143737489279602, 0.21342496834978161, 4.4009487805781786, 0},
     {-3.6143737489279602, 0.11126152710363123, 4.2834880399852997, 0.21962418183902943}
    },
    /* CB4: rhs_update_wts_imp :1526, rhs_update_wts_exp :2290, init_time_schemes :824 */
    {"CB4", 6, 0, 1, 1, 2,
     {0, 0, 0, 0, 0, 0,
      0.125, 0.125, 0, 0, 0, 0,
      0.22486299228804568, 0.22520551156793148, 0.29993149614402281, 0, 0, 0,
      0.1685018686835717, -0.29862732345987331, -0.1185405724697194, 0.62366602724602105, 0, 0,
      0.1685018686835717, 0.14353694052880869, -0.11698581780682539, -0.19505299140555502, 0.5, 0,
      0.1685018686835717, 0.14353694052880869, -0.15821855666404905, -0.48280879550857242, 1.1303111245573836, 0.19867741840285749},
     {0.1685018686835717, 0.14353694052880869, -0.15821855666404905, -0.48280879550857242, 1.1303111245573836, 0.19867741840285749},
     {0, 0, 0, 0, 0, 0,
      0.25, 0, 0, 0, 0, 0,
      0.15321917472916377, 0.5967808252708362, 0, 0, 0, 0,
      0.1685018686835717, 0.12100822974370064, 0.08548990157272765, 0, 0, 0,
      0.1685018686835717, 0.14353694052880869, 0.075161004492142164, 0.11280018629547743, 0, 0,
      0.1685018686835717, 0.14353694052880869, -0.15821855666404905, -0.64963614327365304, 1.4958158907253216, 0},
     {0.1685018686835717, 0.14353694052880869, -0.15821855666404905, -0.48280879550857242, 1.1303111245573836, 0.19867741840285749}
    },
    /* LZ3L1: rhs_update_wts_imp :1543, rhs_update_wts_exp :2309, init_time_schemes :873 */
    {"LZ3L1", 5, 1, 1, 1, 2,
     {0, 0, 0, 0, 0,
      0.14999999999999999, 0.10000000000000001, 0, 0, 0,
      0.90000000000000002, -1.3, 0.90000000000000002, 0, 0,
      1.7, -2.75, 1.5, 0.29999999999999999, 0,
      1, -3.3333333333333335, 5.666666666666667, -3.3333333333333335, 1},
     {1, -3.3333333333333335, 5.666666666666667, -3.3333333333333335, 1},
     {0, 0, 0, 0, 0,
      0.25, 0, 0, 0, 0,
      0, 0.5, 0, 0, 0,
      -0.5, 1.25, 0, 0, 0,
      0, 0.66666666666666663, -0.33333333333333331, 0.66666666666666663, 0},
     {0, 0.66666666666666663, -0.33333333333333331, 0.66666666666666663, 0}
    },
    /* LZ3A4a: rhs_update_wts_imp :1559, rhs_update_wts_exp :2328, init_time_schemes :895 */
    {"LZ3A4a", 5, 1, 1, 1, 3,
     {0, 0, 0, 0, 0,
      0.5, 0, 0, 0, 0,
      0.25, -0.75, 1, 0, 0,
      0, -3, 4, 0, 0,
      0.16666666666666666, 0, 0.66666666666666663, -0.5, 0.66666666666666663},
     {0.16666666666666666, 0, 0.66666666666666663, -0.5, 0.66666666666666663},
     {0, 0, 0, 0, 0,
      0.5, 0, 0, 0, 0,
      0.25, 0.25, 0, 0, 0,
      0, 1, 0, 0, 0,
      0.16666666666666666, 0, 0.66666666666666663, 0.16666666666666666, 0},
     {0.16666666666666666, 0, 0.66666666666666663, 0.16666666666666666, 0}
    },
    /* LZ4A1: rhs_update_wts_imp :1592, rhs_update_wts_exp :2369, init_time_schemes :940 */
    {"LZ4A1", 7, 1, 1, 1, 2,
     {0, 0, 0, 0, 0, 0, 0,
      -0.16666666666666666, 0.5, 0, 0, 0, 0, 0,
      0.16666666666666666, -0.33333333333333331, 0.5, 0, 0, 0, 0,
      0.375, -0.375, 0, 0.5, 0, 0, 0,
      0.125, 0, 0.375, -0.5, 0.5, 0, 0,
      -0.5, 0, 3, -2, 0, 0.5, 0,
      0.16666666666666666, 0, 0, 0, 0.66666666666666663, -0.5, 0.66666666666666663},
     {0.16666666666666666, 0, 0, 0, 0.66666666666666663, -0.5, 0.66666666666666663},
     {0, 0, 0, 0, 0, 0, 0,
      0.33333333333333331, 0, 0, 0, 0, 0, 0,
      0.16666666666666666, 0.16666666666666666, 0, 0, 0, 0, 0,
      0.125, 0, 0.375, 0, 0, 0, 0,
      0.125, 0, 0.375, 0, 0, 0, 0,
      0.5, 0, -1.5, 1, 1, 0, 0,
      0.16666666666666666, 0, 0, 0, 0.66666666666666663, 0.16666666666666666, 0},
     {0.16666666666666666, 0, 0, 0, 0.66666666666666663, 0.16666666666666666, 0}
    },
};

/* the tableau of an implicit / explicit name pair ("X", "expX") */
static const rk_tableau* rk_lookup(const char* imp, const char* exp)
{
    if (strncmp(exp, "exp", 3) || strcmp(exp + 3, imp)) return NULL;
    for (size_t k = 0; k < sizeof(RK_TABLEAUX) / sizeof(RK_TABLEAUX[0]); k++)
        if (!strcmp(RK_TABLEAUX[k].name, imp)) return &RK_TABLEAUX[k];
    return NULL;
}

static const rk_numeric* rk_lookup_numeric(const char* imp, const char* exp)
{
    if (strncmp(exp, "exp", 3) || strcmp(exp + 3, imp)) return NULL;
    for (size_t k = 0; k < sizeof(RK_NUMERIC) / sizeof(RK_NUMERIC[0]); k++)
        if (!strcmp(RK_NUMERIC[k].name, imp)) return &RK_NUMERIC[k];
    return NULL;
}

/* init_time_schemes (timeschemes.F90:42-973) and the tableaux above */
static int scheme_setup(scheme* sc, const char* type, const char* imp, const char* exp)
{
    memset(sc, 0, sizeof(*sc));
    strncpy(sc->expname, exp, 15);
    if (!strcmp(type, "IMEX")) {
        sc->type = ST_IMEX;
        if (!strcmp(imp, "CN")) {
            sc->imp = IMP_CN; sc->n_imp = 2;
            if (!strcmp(exp, "AB2")) sc->n_exp = 2;
            else if (!strcmp(exp, "AB3")) sc->n_exp = 3;
            else return -2;
        } else if (!strcmp(imp, "BDF2")) {
            sc->imp = IMP_BDF2; sc->n_imp = 2;
            if (!strcmp(exp, "AB2")) sc->n_exp = 2; else return -2;
        } else if (!strcmp(imp, "BDF3")) {
            sc->imp = IMP_BDF3; sc->n_imp = 3;
            if (!strcmp(exp, "AB3")) sc->n_exp = 3; else return -2;
        } else if (!strcmp(imp, "BDF4")) {
            sc->imp = IMP_BDF4; sc->n_imp = 4;
            if (!strcmp(exp, "AB4")) sc->n_exp = 4; else return -2;
        } else return -2;
        return 0;
    }
    if (!strcmp(type, "IMEXRK")) {
        sc->type = ST_IMEXRK;
        const rk_tableau* t = rk_lookup(imp, exp);
        const rk_numeric* q = t ? NULL : rk_lookup_numeric(imp, exp);
        int ckA = 0, ckD = 0;
        if (t) {
            double aD[MAXS * MAXS], aA[MAXS * MAXS];
            t->fill(aD, sc->bD, aA, sc->bA);
            sc->S = t->S;
            set_tab(sc->aD, aD, t->S); set_tab(sc->aA, aA, t->S);
            /* ars_eqn_check_A/D: set if any element of the last row equals b
             * (timeschemes.F90:1226-1230, 1934-1938) */
            for (int i = 0; i < sc->S; i++) {
                if (sc->AD(sc->S - 1, i) == sc->bD[i]) ckD = 1;
                if (sc->AA(sc->S - 1, i) == sc->bA[i]) ckA = 1;
            }
        } else if (q) {
            sc->S = q->S;
            set_tab(sc->aD, q->aD, q->S); set_tab(sc->aA, q->aA, q->S);
            memcpy(sc->bD, q->bD, sizeof(double) * (size_t)q->S);
            memcpy(sc->bA, q->bA, sizeof(double) * (size_t)q->S);
            ckA = q->ckA; ckD = q->ckD;
            sc->diag_diff = q->diag_diff;
        } else {
            return -3;   /* not in the registry (diag_diff schemes are not restated) */
        }
        sc->diag_index = q ? q->diag_index : 2;
        sc->n_imp = sc->S; sc->n_exp = sc->S;
        sc->sa = ckA && ckD;
        for (int k = 0; k < sc->S; k++) {
            double s = 0.0;
            for (int i = 0; i < sc->S; i++) s += sc->AA(i, k);
            sc->nl_col[k] = (s + sc->bA[k] != 0.0);
        }
        return 0;
    }
    if (!strcmp(type, "RK")) {
        /* the explicit path (time_scheme_imp = 'NONE', timeschemes.F90:158-186,
         * 1701-1721): butcher_a / butcher_b kept in aA / bA.  RK4's a(2,3) =
         * 1/2 above the diagonal is the reference's (its reshape); the stage
         * sums stop below the diagonal, so it is never read.
         * n_order_tscheme_imp = n_order_tscheme_exp = the stage count. */
        if (strcmp(imp, "NONE")) return -2;
        sc->type = ST_RK;
        if (!strcmp(exp, "RK2")) {
            sc->S = 2;
            sc->AA(1, 0) = 0.5;
            sc->bA[0] = 0.0; sc->bA[1] = 1.0;
        } else if (!strcmp(exp, "RK4")) {
            sc->S = 4;
            sc->AA(1, 0) = 0.5; sc->AA(1, 2) = 0.5;
            sc->AA(2, 1) = 0.5;
            sc->AA(3, 2) = 1.0;
            sc->bA[0] = 1.0 / 6.0; sc->bA[1] = 1.0 / 3.0; sc->bA[2] = 1.0 / 3.0; sc->bA[3] = 1.0 / 6.0;
        } else return -2;
        sc->n_imp = sc->S; sc->n_exp = sc->S;
        sc->sa = 0;
        return 0;
    }
    return -1;
}

int orc_scheme_info(const char* type, const char* imp, const char* exp,
                    int32_t* n_imp, int32_t* n_exp, double* aD, double* bD,
                    double* aA, double* bA, int32_t* sa)
{
    scheme sc;
    int rc = scheme_setup(&sc, type, imp, exp);
    if (rc) return rc;
    if (n_imp) *n_imp = sc.n_imp;
    if (n_exp) *n_exp = sc.n_exp;
    if (sc.type != ST_IMEX) {
        for (int i = 0; i < sc.S; i++) {
            for (int j = 0; j < sc.S; j++) {
                if (aD) aD[i * sc.S + j] = sc.AD(i, j);
                if (aA) aA[i * sc.S + j] = sc.AA(i, j);
            }
            if (bD) bD[i] = sc.bD[i];
            if (bA) bA[i] = sc.bA[i];
        }
        if (sa) *sa = sc.sa;
    }
    return 0;
}

/* rhs_update_wts_imp / rhs_update_wts_exp (timeschemes.F90:993-1052,
 * 1614-1695): variable-step formulas evaluated with dt_array. */
static void imex_weights(const scheme* sc, const double* dta, double* wl, double* wi, double* we)
{
    if (sc->imp == IMP_CN) {
        wi[0] = 0.5;
        *wl = 0.5;
        if (!strcmp(sc->expname, "AB2")) {
            we[0] = 1.0 + 0.5 * dta[0] / dta[1];
            we[1] = -0.5 * dta[0] / dta[1];
        } else { /* AB3 */
            we[1] = (dta[0] * dta[2] * 0.5 + (1.0 / 6.0) * dta[0] * dta[0]) / (dta[1] * dta[1] * 0.5 - dta[1] * dta[2]);
            we[2] = (-dta[0] * 0.5 - dta[1] * we[1]) * 0.5 / (dta[2]);
            we[0] = 1.0 - we[1] - we[2];
        }
    } else if (sc->imp == IMP_BDF2) {
        const double r1 = dta[1] / dta[0];
        const double a0 = (2.0 + r1) / (1.0 + r1);
        const double a1 = 1.0 + 1.0 / r1;
        const double a2 = -1.0 * (1.0 / (r1) / (1.0 + r1));
        *wl = 1.0 / a0; wi[0] = a1 / a0; wi[1] = a2 / a0;
        const double b0 = 1.0 + 1.0 / r1, b1 = -1.0 / r1;
        we[0] = b0 / a0; we[1] = b1 / a0;
    } else if (sc->imp == IMP_BDF3) {
        const double r1 = dta[1] / dta[0];
        const double r2 = dta[2] / dta[0];
        const double a0 = (1.0 + 1.0 / (1.0 + r1) + 1.0 / (1.0 + r1 + r2));
        const double a1 = (((1.0 + r1) * (1.0 + r1 + r2)) / (r1 * (r1 + r2)));
        const double a2 = (-1.0 * (1.0 + r1 + r2) / (r1 * r2 * (1.0 + r1)));
        const double a3 = ((1.0 + r1) / (r2 * (r1 + r2) * (1.0 + r1 + r2)));
        *wl = 1.0 / a0; wi[0] = a1 / a0; wi[1] = a2 / a0; wi[2] = a3 / a0;
        const double b0 = 1.0 * ((1.0 + r1) * (1.0 + r1 + r2)) / (r1 * (r1 + r2));
        const double b1 = 1.0 * (-(1.0 + r1 + r2) / (r1 * r2));
        const double b2 = 1.0 * (1.0 + r1) / (r2 * (r1 + r2));
        we[0] = b0 / a0; we[1] = b1 / a0; we[2] = b2 / a0;
    } else { /* BDF4 / AB4 */
        const double one = 1.0;
        const double r1 = dta[0] / dta[1], r2 = dta[1] / dta[2], r3 = dta[2] / dta[3];
        const double c1 = one + r3 * (one + r2);
        const double c2 = one + r2 * (one + r1);
        const double c3 = one + r3 * c2;
        const double a0 = one + r1 / (one + r1) + r2 * r1 / c2 + r3 * r2 * r1 / c3;
        const double a1 = -one - r1 * (one + r2 * (one + r1) * (one + r3 * c2 / c1) / (one + r2));
        const double a2 = r1 * (r1 / (one + r1) + r2 * r1 * (c3 + r3) / (one + r3));
        const double a3 = -pow(r2, 3) * pow(r1, 2) * (one + r1) / (one + r2) * c3 / c2;
        const double a4 = (one + r1) / (one + r3) * c2 / c1 / c3 * (pow(r3, 4) * pow(r2, 3) * pow(r1, 2));
        *wl = 1.0 / a0;
        wi[0] = -a1 / a0; wi[1] = -a2 / a0; wi[2] = -a3 / a0; wi[3] = -a4 / a0;
        const double b3 = -pow(r3, 3) * pow(r2, 2) * r1 * (one + r1) / (one + r3) * c2 / c1;
        const double b2 = pow(r2, 2) * r1 * (one + r1) / (one + r2) * c3;
        const double b1 = -c2 * c3 * r1 / (one + r3);
        const double b0 = r2 * (one + r1) / (one + r2) * ((one + r1) * (c3 + r3) + (one + r3) / r2) / c1;
        we[0] = b0 / a0; we[1] = b1 / a0; we[2] = b2 / a0; we[3] = b3 / a0;
    }
}

/* ------------------------------------------------------------------------ */
/* Run state                                                                  */
/* ------------------------------------------------------------------------ */
struct orc_run {
    orc_params p;
    scheme sc;
    int N, Nm, Nm1, Np;
    double rmin, rmax;
    chebgrid G;
    chebctx C;
    fourctx F;
    /* state, (Nm+1, Nr) */
    cplx *tFR, *omgFR, *upFR, *urFR, *psii;
    cplx *temp_spec, *omg_spec, *upFC;
    cplx *uphi_bar_spec, *TFC;
    /* products and the physical mean <ur omega> of the last nonlinear pass */
    cplx *upT, *urT, *upW, *urW;
    double* urw_mean;
    /* IMEX histories [order][Nm+1][Nr] and the shift copies */
    cplx *rhs_imp_temp, *rhs_exp_temp, *rhs_imp_vort, *rhs_exp_vort, *rhs_buo_term;
    cplx *tmp_imp_temp, *tmp_exp_temp, *tmp_imp_vort, *tmp_exp_vort;
    cplx *rhs_imp_ub, *rhs_exp_ub, *tmp_imp_ub, *tmp_exp_ub;  /* [order][Nr] */
    /* IMEXRK stage terms [stage][Nm+1][Nr] */
    cplx *dT_a, *dT_d, *dW_a, *dW_d, *buo;
    /* factors */
    double *AT, *AF, *Au;
    int32_t *pAT, *pAF, *pAu;
    int mats_built;
    double mat_wt, mat_dt;
    /* counters */
    int kstep;            /* n_step - n_restart after the last step */
    double tot_time;
    double dt_array[8];
    double dt_new;
    double *dtval_r, *dtval_p, *dr;   /* get_dtvals per radial level; dr of init_grid */
    /* Assembly_stage: LAPpsi factors per mode (mat_build_rk) and the Lagrange
     * derivative weights at r_min / r_max (init.F90:268-271) */
    double* LAP;
    int32_t* pLAP;
    int lap_built;
    double dw_rmin[64], d2w_rmin[64], dw_rmax[64], d2w_rmax[64];
    int nthreads;
    int hist_given;       /* histories set by orc_restart */
    /* writeKE_spectral's hidden inputs: the temperature of the last nonlinear
     * pass (the module array tt, nonlin.F90:150) and upFR_prev (steptime.F90:
     * 212, init.F90:253), both spectral here */
    cplx *tt, *up_prev;
};

#define SPEC(R) ((size_t)(R)->Nm1 * (size_t)(R)->N)

static cplx* calloc_c(size_t n) { return (cplx*)calloc(n ? n : 1, sizeof(cplx)); }

void orc_destroy(orc_run* R)
{
    if (!R) return;
    cplx* arrs[] = {R->tFR, R->omgFR, R->upFR, R->urFR, R->psii, R->temp_spec, R->omg_spec, R->upFC,
                    R->uphi_bar_spec, R->TFC, R->upT, R->urT, R->upW, R->urW,
                    R->rhs_imp_temp, R->rhs_exp_temp, R->rhs_imp_vort, R->rhs_exp_vort, R->rhs_buo_term,
                    R->tmp_imp_temp, R->tmp_exp_temp, R->tmp_imp_vort, R->tmp_exp_vort,
                    R->rhs_imp_ub, R->rhs_exp_ub, R->tmp_imp_ub, R->tmp_exp_ub,
                    R->dT_a, R->dT_d, R->dW_a, R->dW_d, R->buo};
    for (size_t i = 0; i < sizeof(arrs) / sizeof(arrs[0]); i++) free(arrs[i]);
    free(R->urw_mean);
    free(R->tt); free(R->up_prev);
    free(R->dtval_r); free(R->dtval_p); free(R->dr);
    free(R->LAP); free(R->pLAP);
    free(R->AT); free(R->AF); free(R->Au);
    free(R->pAT); free(R->pAF); free(R->pAu);
    grid_free(&R->G);
    plan_free(&R->C.P);
    plan_free(&R->F.P);
    free(R);
}

static int assembly_setup(orc_run* R);

int orc_create(const orc_params* p, orc_run** out)
{
    *out = NULL;
    if (p->Nr < 4 || p->Nm < 1) return -1;
    orc_run* R = (orc_run*)calloc(1, sizeof(orc_run));
    if (!R) return -4;
    R->p = *p;
    int rc = scheme_setup(&R->sc, p->scheme_type, p->scheme_imp, p->scheme_exp);
    if (rc) { free(R); return rc; }
    R->N = p->Nr; R->Nm = p->Nm; R->Nm1 = p->Nm + 1; R->Np = 3 * p->Nm;
    R->rmin = p->eta / (1.0 - p->eta);
    R->rmax = 1.0 / (1.0 - p->eta);
    if (grid_init(&R->G, R->N, p->eta)) { orc_destroy(R); return -4; }
    cheb_init(&R->C, R->N);
    R->F.Nm = R->Nm; R->F.Np = R->Np; plan_init(&R->F.P, R->Np);
    const size_t S = SPEC(R);
    R->tFR = calloc_c(S); R->omgFR = calloc_c(S); R->upFR = calloc_c(S); R->urFR = calloc_c(S);
    R->psii = calloc_c(S); R->temp_spec = calloc_c(S); R->omg_spec = calloc_c(S); R->upFC = calloc_c(S);
    R->uphi_bar_spec = calloc_c(R->N); R->TFC = calloc_c(R->N);
    R->upT = calloc_c(S); R->urT = calloc_c(S); R->upW = calloc_c(S); R->urW = calloc_c(S);
    R->urw_mean = (double*)calloc((size_t)R->N, sizeof(double));
    R->tt = calloc_c(S); R->up_prev = calloc_c(R->N);
    if (!R->tt || !R->up_prev) { orc_destroy(R); return -4; }
    R->dtval_r = (double*)calloc((size_t)R->N, sizeof(double));
    R->dtval_p = (double*)calloc((size_t)R->N, sizeof(double));
    R->dr = (double*)calloc((size_t)R->N, sizeof(double));
    if (!R->urw_mean || !R->dtval_r || !R->dtval_p || !R->dr) { orc_destroy(R); return -4; }
    /* dr(i+1) = radius(i+1) - radius(i), dr(1) = dr(2) (init.F90:133-137) */
    for (int i = 1; i < R->N; i++) R->dr[i] = R->G.radius[i] - R->G.radius[i - 1];
    R->dr[0] = R->dr[1];
    const int ni = R->sc.n_imp, ne = R->sc.n_exp;
    if (R->sc.type == ST_IMEX) {
        R->rhs_imp_temp = calloc_c(ni * S); R->rhs_exp_temp = calloc_c(ne * S);
        R->rhs_imp_vort = calloc_c(ni * S); R->rhs_exp_vort = calloc_c(ne * S);
        R->rhs_buo_term = calloc_c(ni * S);
        R->tmp_imp_temp = calloc_c(ni * S); R->tmp_exp_temp = calloc_c(ne * S);
        R->tmp_imp_vort = calloc_c(ni * S); R->tmp_exp_vort = calloc_c(ne * S);
        R->rhs_imp_ub = calloc_c(ni * R->N); R->rhs_exp_ub = calloc_c(ne * R->N);
        R->tmp_imp_ub = calloc_c(ni * R->N); R->tmp_exp_ub = calloc_c(ne * R->N);
    } else {
        const int Sg = R->sc.S;
        R->dT_a = calloc_c(Sg * S); R->dT_d = calloc_c(Sg * S);
        R->dW_a = calloc_c(Sg * S); R->dW_d = calloc_c(Sg * S); R->buo = calloc_c(Sg * S);
    }
    const size_t n2 = (size_t)R->N * R->N;
    R->AT = (double*)malloc(sizeof(double) * n2 * R->Nm1);
    R->AF = (double*)malloc(sizeof(double) * 4 * n2 * R->Nm1);
    R->Au = (double*)malloc(sizeof(double) * n2);
    R->pAT = (int32_t*)malloc(sizeof(int32_t) * (size_t)R->N * R->Nm1);
    R->pAF = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)R->N * R->Nm1);
    R->pAu = (int32_t*)malloc(sizeof(int32_t) * (size_t)R->N);
    if (!R->AT || !R->AF || !R->Au || !R->pAT || !R->pAF || !R->pAu) { orc_destroy(R); return -4; }
    R->nthreads = p->nthreads;
    if ((R->sc.type == ST_IMEXRK && !R->sc.sa) || R->sc.type == ST_RK) {
        if (p->lagpts < 2 || p->lagpts > 64 || p->lagpts > p->Nr - 2) { orc_destroy(R); return -1; }
        if (assembly_setup(R)) { orc_destroy(R); return -4; }
    }
    for (int i = 0; i < 8; i++) R->dt_array[i] = p->dt;
    R->dt_new = p->dt;
    *out = R;
    return 0;
}

int orc_set_state(orc_run* R, const double* tFR, const double* omgFR, const double* urFR, const double* upFR)
{
    const size_t S = SPEC(R);
    memcpy(R->tFR, tFR, sizeof(cplx) * S);
    memcpy(R->omgFR, omgFR, sizeof(cplx) * S);
    memcpy(R->urFR, urFR, sizeof(cplx) * S);
    memcpy(R->upFR, upFR, sizeof(cplx) * S);
    /* init_perturbation / init_fields (init.F90:167-174, 393-397): the spec
     * arrays start as copies; IMEX converts them at its first step. */
    memcpy(R->temp_spec, tFR, sizeof(cplx) * S);
    memcpy(R->omg_spec, omgFR, sizeof(cplx) * S);
    memcpy(R->upFC, upFR, sizeof(cplx) * S);
    for (int r = 0; r < R->N; r++) R->uphi_bar_spec[r] = R->upFR[(size_t)R->Nm1 * r];
    /* init_fields / init_perturbation: tt is the initial temperature,
     * upFR_prev the initial u_phi (init.F90:253) */
    memcpy(R->tt, tFR, sizeof(cplx) * S);
    for (int r = 0; r < R->N; r++) R->up_prev[r] = R->upFR[(size_t)R->Nm1 * r];
    R->kstep = 0;
    R->tot_time = 0.0;
    R->mats_built = 0;
    R->hist_given = 0;
    R->dt_new = R->p.dt;
    for (int i = 0; i < 8; i++) R->dt_array[i] = R->p.dt;
    return 0;
}

int orc_get_state(const orc_run* R, double* tFR, double* omgFR, double* urFR, double* upFR, double* psi)
{
    const size_t S = SPEC(R);
    if (tFR) memcpy(tFR, R->tFR, sizeof(cplx) * S);
    if (omgFR) memcpy(omgFR, R->omgFR, sizeof(cplx) * S);
    if (urFR) memcpy(urFR, R->urFR, sizeof(cplx) * S);
    if (upFR) memcpy(upFR, R->upFR, sizeof(cplx) * S);
    if (psi) memcpy(psi, R->psii, sizeof(cplx) * S);
    return 0;
}

int orc_get_TFC(const orc_run* R, double* TFC)
{
    memcpy(TFC, R->TFC, sizeof(cplx) * (size_t)R->N);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Nonlinear pass: Nr_maxLOOP, nonlin.F90:134-190                             */
/* ------------------------------------------------------------------------ */
static void nonlin_run(orc_run* R, const cplx* tFR, const cplx* omgFR, const cplx* upFR, const cplx* urFR)
{
    if (R->tt) memcpy(R->tt, tFR, sizeof(cplx) * SPEC(R));   /* tt of this pass */
    const int Nm1 = R->Nm1, Np = R->Np, N = R->N;
    double* tt = (double*)malloc(sizeof(double) * (size_t)Np * 6);
    double* om = tt + Np;
    double* up = om + Np;
    double* ur = up + Np;
    double* pa = ur + Np;
    double* pb = pa + Np;
    cplx* w = (cplx*)malloc(sizeof(cplx) * 2 * (size_t)Np);
    for (int r = 0; r < N; r++) {
        const size_t o = (size_t)Nm1 * r;
        four_inv(&R->F, tFR + o, tt, w);
        four_inv(&R->F, omgFR + o, om, w);
        four_inv(&R->F, upFR + o, up, w);
        four_inv(&R->F, urFR + o, ur, w);
        if (R->dtval_r) {
            /* get_dtvals (nonlin.F90:93-106): grid spacing over the largest
             * |u_r| and |u_phi| of the level (x/0 = +Inf for a fluid at rest) */
            double mr = 0.0, mp = 0.0;
            for (int i = 0; i < Np; i++) { mr = fmax(mr, fabs(ur[i])); mp = fmax(mp, fabs(up[i])); }
            const double pi = 4.0 * atan(1.0);
            R->dtval_r[r] = R->dr[r] / mr;
            R->dtval_p[r] = (2.0 / 3.0) * (R->G.radius[r] * pi) / ((double)R->Nm * mp);
        }
        for (int i = 0; i < Np; i++) { pa[i] = up[i] * tt[i]; pb[i] = ur[i] * tt[i]; }
        four_fwd(&R->F, pa, R->upT + o, w);
        four_fwd(&R->F, pb, R->urT + o, w);
        for (int i = 0; i < Np; i++) { pa[i] = up[i] * om[i]; pb[i] = ur[i] * om[i]; }
        four_fwd(&R->F, pa, R->upW + o, w);
        four_fwd(&R->F, pb, R->urW + o, w);
        /* <ur omega>(r) from the physical arrays, rhs_create.F90:329-335 */
        double s = 0.0;
        for (int i = 0; i < Np; i++) s = s + ur[i] * om[i];
        R->urw_mean[r] = s * (1.0 / (double)Np);
    }
    free(w);
    free(tt);
}

void orc_nonlin(int Nm, int Nr, const double* tFR, const double* omgFR, const double* upFR, const double* urFR,
                double* upT, double* urT, double* upW, double* urW, double* urw_mean)
{
    orc_run R;
    memset(&R, 0, sizeof(R));
    R.N = Nr; R.Nm = Nm; R.Nm1 = Nm + 1; R.Np = 3 * Nm;
    R.F.Nm = Nm; R.F.Np = R.Np; plan_init(&R.F.P, R.Np);
    R.upT = (cplx*)upT; R.urT = (cplx*)urT; R.upW = (cplx*)upW; R.urW = (cplx*)urW;
    R.urw_mean = urw_mean;
    nonlin_run(&R, (const cplx*)tFR, (const cplx*)omgFR, (const cplx*)upFR, (const cplx*)urFR);
    plan_free(&R.F.P);
}

/* ------------------------------------------------------------------------ */
/* Helpers                                                                    */
/* ------------------------------------------------------------------------ */
static void get_rad(const orc_run* R, const cplx* F, int m, cplx* v)
{
    for (int r = 0; r < R->N; r++) v[r] = F[m + (size_t)R->Nm1 * r];
}
static void put_rad(const orc_run* R, cplx* F, int m, const cplx* v)
{
    for (int r = 0; r < R->N; r++) F[m + (size_t)R->Nm1 * r] = v[r];
}
static cplx* hist(cplx* H, const orc_run* R, int order, int m, int r)
{
    return H + (size_t)order * SPEC(R) + m + (size_t)R->Nm1 * r;
}

/* Complex solve as the reference does it: Re and Im through dgetrs. */
static void solve_cplx(int n, const double* A, const int32_t* piv, cplx* b, double* br, double* bi)
{
    for (int i = 0; i < n; i++) { br[i] = b[i].re; bi[i] = b[i].im; }
    orc_dgetrs(n, A, piv, br);
    orc_dgetrs(n, A, piv, bi);
    for (int i = 0; i < n; i++) b[i] = cx(br[i], bi[i]);
}

static void build_factors(orc_run* R, double wt, double dt, int irk_order)
{
    const int N = R->N;
    const size_t n2 = (size_t)N * N;
#pragma omp parallel for schedule(dynamic) num_threads(R->nthreads > 0 ? R->nthreads : omp_get_max_threads())
    for (int m = 0; m < R->Nm1; m++) {
        double* AT = R->AT + n2 * m;
        double* AF = R->AF + 4 * n2 * m;
        mat_build_g(&R->G, m, wt, dt, R->p.Pr, R->p.mBC, irk_order, AT, m > 0 ? AF : NULL);
        orc_dgetrf(N, AT, R->pAT + (size_t)N * m);
        if (m > 0) orc_dgetrf(2 * N, AF, R->pAF + (size_t)2 * N * m);
    }
    mat_build_uphibar_g(&R->G, wt, dt, R->p.mBC, R->Au);
    orc_dgetrf(N, R->Au, R->pAu);
    R->mats_built = 1;
    R->mat_wt = wt;
    R->mat_dt = dt;
}

/* ------------------------------------------------------------------------ */
/* IMEX step: timeloop_imex body (steptime.F90:196-251) + Nm_maxLOOP          */
/* ------------------------------------------------------------------------ */
typedef struct {
    cplx *a, *b, *c, *d, *e, *f, *g, *h, *rhs, *rhsf, *w;
    double *br, *bi;
} wsp;

static void wsp_alloc(wsp* W, int N, int M)
{
    cplx* base = (cplx*)malloc(sizeof(cplx) * (size_t)(9 * N + 2 * N + 2 * M + 2 * N));
    W->a = base; W->b = W->a + N; W->c = W->b + N; W->d = W->c + N; W->e = W->d + N;
    W->f = W->e + N; W->g = W->f + N; W->h = W->g + N; W->rhs = W->h + N;
    W->rhsf = W->rhs + N; W->w = W->rhsf + 2 * N;
    W->br = (double*)malloc(sizeof(double) * 4 * (size_t)N);
    W->bi = W->br + 2 * N;
}
static void wsp_free(wsp* W) { free(W->a); free(W->br); }

static void imex_mode(orc_run* R, int m, wsp* W, double dt, double wl, const double* wi, const double* we)
{
    const int N = R->N, Nm1 = R->Nm1;
    const int ni = R->sc.n_imp, ne = R->sc.n_exp;
    const chebctx* C = &R->C;
    const double* rr = R->G.r_radius;
    const double* rr2 = R->G.r_radius2;
    const double mm = (double)m;
    const double Pr = R->p.Pr, Ra = R->p.Ra;
    const int cn = (R->sc.imp == IMP_CN);
    (void)Nm1;

    /* --- rhs_construct_temp (rhs_create.F90:24-140) --- */
    if (R->kstep > 1) {
        for (int r = 0; r < N; r++) {
            for (int o = 0; o < ne - 1; o++) *hist(R->rhs_exp_temp, R, o + 1, m, r) = *hist(R->tmp_exp_temp, R, o, m, r);
            for (int o = 0; o < ni - 1; o++) *hist(R->rhs_imp_temp, R, o + 1, m, r) = *hist(R->tmp_imp_temp, R, o, m, r);
        }
    }
    cplx* upT = W->a; cplx* urT = W->b; cplx* urTs = W->c; cplx* Tr = W->d; cplx* dT = W->e; cplx* d2T = W->f;
    cplx* dUrT = W->g; cplx* spec = W->h; cplx* rhs = W->rhs;
    get_rad(R, R->upT, m, upT);
    get_rad(R, R->urT, m, urT);
    get_rad(R, R->temp_spec, m, spec);
    cheb_fwd(C, urT, urTs, W->w);
    cheb_inv(C, spec, Tr, W->w);
    cheb_d1d2(C, spec, dT, d2T, W->w);
    cheb_d1(C, urTs, dUrT, W->w);
    for (int i = 0; i < N; i++) {
        cplx r0;
        *hist(R->rhs_imp_temp, R, 0, m, i) = Tr[i];
        if (cn) {
            const double c = ((1.0 - wi[0]) * dt) * (1.0 / Pr);
            const cplx lap = csub(cadd(cscale(rr[i], dT[i]), d2T[i]), cscale((mm * mm) * rr2[i], Tr[i]));
            r0 = cadd(Tr[i], cscale(c, lap));
        } else {
            r0 = cx(0.0, 0.0);
            for (int o = 0; o < ni; o++) r0 = cadd(r0, cscale(wi[o], *hist(R->rhs_imp_temp, R, o, m, i)));
        }
        *hist(R->rhs_exp_temp, R, 0, m, i) = cadd(cadd(cimr(mm, rr[i], upT[i]), dUrT[i]), flux_over_r(rr[i], urT[i]));
        for (int o = 0; o < ne; o++) r0 = csub(r0, cscale(dt, cscale(we[o], *hist(R->rhs_exp_temp, R, o, m, i))));
        rhs[i] = r0;
    }
    rhs[0] = cx(m == 0 ? 1.0 : 0.0, 0.0);
    rhs[N - 1] = cx(0.0, 0.0);
    /* solve AT (fourierloop.F90:126-142) */
    const size_t n2 = (size_t)N * N;
    solve_cplx(N, R->AT + n2 * m, R->pAT + (size_t)N * m, rhs, W->br, W->bi);
    if (m == 0) memcpy(R->TFC, rhs, sizeof(cplx) * N);
    cplx* Tn = W->a;   /* reuse */
    cheb_inv(C, rhs, Tn, W->w);
    put_rad(R, R->temp_spec, m, rhs);
    put_rad(R, R->tFR, m, Tn);

    if (m == 0) {
        /* --- rhs_construct_uphi_bar (rhs_create.F90:273-385) --- */
        if (R->kstep > 1) {
            for (int r = 0; r < N; r++) {
                for (int o = 0; o < ne - 1; o++) R->rhs_exp_ub[(size_t)(o + 1) * N + r] = R->tmp_exp_ub[(size_t)o * N + r];
                for (int o = 0; o < ni - 1; o++) R->rhs_imp_ub[(size_t)(o + 1) * N + r] = R->tmp_imp_ub[(size_t)o * N + r];
            }
        }
        cplx* ufc = W->b; cplx* d1 = W->c; cplx* d2 = W->d; cplx* up0 = W->e;
        get_rad(R, R->upFC, 0, ufc);
        get_rad(R, R->upFR, 0, up0);
        cheb_d1(C, ufc, d1, W->w);
        cheb_d2(C, ufc, d2, W->w);
        double* ru = W->br + N;   /* real rhs */
        for (int i = 0; i < N; i++) {
            const cplx ud = cadd(csub(d2[i], cscale(rr2[i], up0[i])), cscale(rr[i], d1[i]));
            R->rhs_imp_ub[i] = up0[i];
            double v;
            if (cn) v = R->rhs_imp_ub[i].re + ((1.0 - wi[0]) * dt) * ud.re;
            else {
                v = 0.0;
                for (int o = 0; o < ni; o++) v = v + wi[o] * R->rhs_imp_ub[(size_t)o * N + i].re;
            }
            R->rhs_exp_ub[i] = cx(R->urw_mean[i], 0.0);
            for (int o = 0; o < ne; o++) v = v - dt * (we[o] * R->rhs_exp_ub[(size_t)o * N + i].re);
            ru[i] = v;
        }
        ru[0] = 0.0;
        ru[N - 1] = 0.0;
        orc_dgetrs(N, R->Au, R->pAu, ru);
        for (int i = 0; i < N; i++) ufc[i] = cx(ru[i], 0.0);
        put_rad(R, R->upFC, 0, ufc);
        cplx* ub = W->c; cplx* dub = W->d;
        cheb_inv(C, ufc, ub, W->w);
        cheb_d1(C, ufc, dub, W->w);
        put_rad(R, R->upFR, 0, ub);
        for (int i = 0; i < N; i++) {
            R->urFR[(size_t)Nm1 * i] = cx(0.0, 0.0);
            R->omgFR[(size_t)Nm1 * i] = cadd(cscale(rr[i], ub[i]), dub[i]);
        }
        return;
    }

    /* --- rhs_construct_vort (rhs_create.F90:142-271) --- */
    if (R->kstep > 1) {
        for (int r = 0; r < N; r++) {
            for (int o = 0; o < ne - 1; o++) *hist(R->rhs_exp_vort, R, o + 1, m, r) = *hist(R->tmp_exp_vort, R, o, m, r);
            for (int o = 0; o < ni - 1; o++) *hist(R->rhs_imp_vort, R, o + 1, m, r) = *hist(R->tmp_imp_vort, R, o, m, r);
        }
    }
    cplx* upW = W->b; cplx* urW = W->c; cplx* urWs = W->d; cplx* Wr = W->e; cplx* dW = W->f; cplx* d2W = W->g;
    cplx* dUrW = W->h;
    cplx* wsp_ = W->rhsf;       /* omega spec, then the 2N rhs */
    get_rad(R, R->upW, m, upW);
    get_rad(R, R->urW, m, urW);
    cplx* os = W->rhs;          /* omega spectral vector (rhs of T no longer needed) */
    get_rad(R, R->omg_spec, m, os);
    cheb_fwd(C, urW, urWs, W->w);
    cheb_inv(C, os, Wr, W->w);
    cheb_d1d2(C, os, dW, d2W, W->w);
    cheb_d1(C, urWs, dUrW, W->w);
    cplx* rhsf = wsp_;
    for (int i = 0; i < N; i++) {
        cplx r2;
        *hist(R->rhs_imp_vort, R, 0, m, i) = Wr[i];
        const cplx Tnew = Tn[i];
        if (cn) {
            const cplx lap = csub(cadd(cscale(rr[i], dW[i]), d2W[i]), cscale((mm * mm) * rr2[i], Wr[i]));
            const cplx b1 = cimr(mm, rr[i], Tnew);
            const cplx b2 = cimr(mm, rr[i], *hist(R->rhs_imp_temp, R, 0, m, i));
            r2 = csub(cadd(Wr[i], cscale((1.0 - wi[0]) * dt, lap)), cscale((buo_ra_pr(Ra, Pr) * dt) * 0.5, cadd(b1, b2)));
            *hist(R->rhs_exp_vort, R, 0, m, i) = cadd(cadd(cimr(mm, rr[i], upW[i]), dUrW[i]), flux_over_r(rr[i], urW[i]));
            for (int o = 0; o < ne; o++) r2 = csub(r2, cscale(dt, cscale(we[o], *hist(R->rhs_exp_vort, R, o, m, i))));
        } else {
            r2 = cx(0.0, 0.0);
            for (int o = 0; o < ni; o++) r2 = cadd(r2, cscale(wi[o], *hist(R->rhs_imp_vort, R, o, m, i)));
            const cplx bt = cscale(-buo_ra_pr(Ra, Pr), cimr(mm, rr[i], Tnew));
            *hist(R->rhs_buo_term, R, 0, m, i) = bt;
            *hist(R->rhs_exp_vort, R, 0, m, i) = cadd(cadd(cimr(mm, rr[i], upW[i]), dUrW[i]), flux_over_r(rr[i], urW[i]));
            for (int o = 0; o < ne; o++) r2 = csub(r2, cscale(dt, cscale(we[o], *hist(R->rhs_exp_vort, R, o, m, i))));
            r2 = cadd(r2, cscale(dt * wl, bt));
        }
        rhsf[N + i] = r2;
        rhsf[i] = cx(0.0, 0.0);
    }
    rhsf[N] = cx(0.0, 0.0);
    rhsf[2 * N - 1] = cx(0.0, 0.0);
    /* solve AF (fourierloop.F90:179-207) */
    solve_cplx(2 * N, R->AF + 4 * n2 * m, R->pAF + (size_t)2 * N * m, rhsf, W->br, W->bi);
    cplx* wphys = W->a; cplx* psi = W->b; cplx* dpsi = W->c;
    put_rad(R, R->omg_spec, m, rhsf + N);
    cheb_inv(C, rhsf + N, wphys, W->w);
    put_rad(R, R->omgFR, m, wphys);
    cheb_inv(C, rhsf, psi, W->w);
    cheb_d1(C, rhsf, dpsi, W->w);
    for (int i = 0; i < N; i++) {
        const size_t o = (size_t)m + (size_t)Nm1 * i;
        R->psii[o] = psi[i];
        R->upFR[o] = cscale(-1.0, dpsi[i]);
        R->urFR[o] = cimr(mm, rr[i], psi[i]);
    }
    cplx* upv = W->d; cplx* upc = W->e;
    get_rad(R, R->upFR, m, upv);
    cheb_fwd(C, upv, upc, W->w);
    put_rad(R, R->upFC, m, upc);
}

/* rhs_temp_from_restart / rhs_vort_from_restart / rhs_uphibar_from_restart
 * (rhs_create.F90:387-573) for the current state, after a nonlinear pass over
 * it: history slot ki of the implicit arrays and ke of the explicit ones
 * (0-based; the reference's index n_order - j for its j-th checkpoint,
 * steptime.F90:88-186), or every slot when ki < 0 (identical checkpoints). */
static void restart_fill(orc_run* R, int ki, int ke)
{
    const int N = R->N, ni = R->sc.n_imp, ne = R->sc.n_exp;
    const int i0 = ki < 0 ? 0 : ki, i1 = ki < 0 ? ni : ki + 1;
    const int e0 = ke < 0 ? 0 : ke, e1 = ke < 0 ? ne : ke + 1;
    const double* rr = R->G.r_radius;
    wsp W;
    wsp_alloc(&W, N, R->C.M);
    for (int m = 0; m < R->Nm1; m++) {
        const double mm = (double)m;
        get_rad(R, R->urT, m, W.b);
        cheb_fwd(&R->C, W.b, W.c, W.w);
        cheb_d1(&R->C, W.c, W.d, W.w);
        get_rad(R, R->upT, m, W.a);
        get_rad(R, R->urW, m, W.e);
        cheb_fwd(&R->C, W.e, W.f, W.w);
        cheb_d1(&R->C, W.f, W.g, W.w);
        get_rad(R, R->upW, m, W.h);
        for (int i = 0; i < N; i++) {
            const size_t o = (size_t)m + (size_t)R->Nm1 * i;
            const cplx NT = cadd(cadd(cimr(mm, rr[i], W.a[i]), W.d[i]), flux_over_r(rr[i], W.b[i]));
            const cplx NW = cadd(cadd(cimr(mm, rr[i], W.h[i]), W.g[i]), flux_over_r(rr[i], W.e[i]));
            const cplx bt = cscale(-buo_ra_pr(R->p.Ra, R->p.Pr), cimr(mm, rr[i], R->tFR[o]));
            for (int k = i0; k < i1; k++) {
                *hist(R->rhs_imp_temp, R, k, m, i) = R->tFR[o];
                *hist(R->rhs_imp_vort, R, k, m, i) = R->omgFR[o];
                *hist(R->rhs_buo_term, R, k, m, i) = bt;
            }
            for (int k = e0; k < e1; k++) {
                *hist(R->rhs_exp_temp, R, k, m, i) = NT;
                *hist(R->rhs_exp_vort, R, k, m, i) = NW;
            }
        }
    }
    for (int i = 0; i < N; i++) {
        for (int k = i0; k < i1; k++) R->rhs_imp_ub[(size_t)k * N + i] = R->upFR[(size_t)R->Nm1 * i];
        for (int k = e0; k < e1; k++) R->rhs_exp_ub[(size_t)k * N + i] = cx(R->urw_mean[i], 0.0);
    }
    wsp_free(&W);
}

/* Multistep restart start-up (steptime.F90:87-193) with K identical
 * checkpoints of the current state: every history slot holds the state and
 * its explicit term. */
static void imex_restart_histories(orc_run* R)
{
    nonlin_run(R, R->tFR, R->omgFR, R->upFR, R->urFR);
    restart_fill(R, -1, -1);
}

/* compute_new_dt (steptime.F90:313-370 for IMEX, steptime_imexrk.F90:248-308
 * for IMEXRK): the CFL-limited dt from the get_dtvals of the last nonlinear
 * pass, optionally "optimised", then the dt_array update.  R->kstep is
 * n_step - n_restart of the step being taken.  IMEX shifts every slot,
 * IMEXRK only slot 2 (its loop runs i_order = 2, 2, -1). */
static void new_dt(orc_run* R, int imexrk)
{
    const orc_params* p = &R->p;
    const int ks = R->kstep;
    double dt_cal = fmin(R->dtval_r[0], R->dtval_p[0]);
    for (int r = 1; r < R->N; r++) dt_cal = fmin(dt_cal, fmin(R->dtval_r[r], R->dtval_p[r]));
    double dt_n;
    if (p->Pr == 1.0 || p->Pr > 1.0) dt_n = fmax(1.0 / p->Pr, 1.0) * p->CFL * dt_cal;
    else dt_n = fmin(1.0 / p->Pr, 1.0) * p->CFL * dt_cal;
    const double dt2 = fmin(0.5 * (1.0 / p->dt_coef + 1.0) * dt_n, p->dt_max);
    if (p->optimizedt) {
        if (R->dt_array[1] > p->dt_max && ks > 1) R->dt_new = p->dt_max;
        else if (R->dt_array[1] >= dt_cal && ks > 1) R->dt_new = dt2;
        else if (p->dt_coef * R->dt_array[1] < dt_cal && R->dt_array[1] < p->dt_max && ks > 1)
            R->dt_new = (R->dt_array[0] < dt2) ? dt2 : R->dt_array[0];
        /* no branch taken: dt_new keeps the previous step's value */
    } else {
        R->dt_new = dt2;
    }
    const int nmax = R->sc.n_imp > R->sc.n_exp ? R->sc.n_imp : R->sc.n_exp;
    if (!p->restart && ks == 1) for (int i = 0; i < nmax; i++) R->dt_array[i] = R->dt_new;
    else if (p->restart && ks == 1) R->dt_array[0] = R->dt_new;
    if (ks > 1) {
        if (nmax > 1) {
            if (imexrk) R->dt_array[1] = R->dt_array[0];
            else for (int i = nmax - 1; i >= 1; i--) R->dt_array[i] = R->dt_array[i - 1];
        }
        R->dt_array[0] = R->dt_new;
    }
}

/* compute_new_dt of the RK path (steptime_rk.F90:181-236) after dt_array(:)
 * = dt at the step's start (:96-99): the CFL limits of the last pass and the
 * diffusive ones, dr^2 and ((2/3) pi r / Nm)^2 (get_dtvals, nonlin.F90:
 * 102-104); from the second step the optimize branches, with dt_array(1) =
 * dt_array(2) = dt; no branch taken keeps dt. */
static void new_dt_rk(orc_run* R)
{
    const orc_params* p = &R->p;
    const int ks = R->kstep;
    const double pi = 4.0 * atan(1.0);
    R->dt_new = p->dt;
    if (ks > 1) {
        double dt_cal = INFINITY;
        for (int r = 0; r < R->N; r++) {
            const double rad = R->G.radius[r];
            const double rkr = R->dr[r] * R->dr[r];
            const double rkp = (2.0 / 3.0) * (rad * pi) / (double)R->Nm * (2.0 / 3.0) * (rad * pi) / (double)R->Nm;
            dt_cal = fmin(dt_cal, fmin(fmin(R->dtval_r[r], R->dtval_p[r]), fmin(rkr, rkp)));
        }
        double dt_n;
        if (p->Pr == 1.0 || p->Pr > 1.0) dt_n = fmax(1.0 / p->Pr, 1.0) * p->CFL * dt_cal;
        else dt_n = fmin(1.0 / p->Pr, 1.0) * p->CFL * dt_cal;
        const double dt2 = fmin(0.5 * (1.0 / p->dt_coef + 1.0) * dt_n, p->dt_max);
        const double d = p->dt;   /* dt_array(1) = dt_array(2) = dt */
        if (d > p->dt_max) R->dt_new = p->dt_max;
        else if (d >= dt_cal) R->dt_new = dt2;
        else if (p->dt_coef * d < dt_cal && d < p->dt_max) R->dt_new = (d < dt2) ? dt2 : d;
    }
    for (int i = 0; i < 8; i++) R->dt_array[i] = p->dt;
    R->dt_array[0] = R->dt_new;
}

double orc_dt(const orc_run* R) { return R->dt_new; }

static void imex_step(orc_run* R)
{
    const int first = (R->kstep == 0);
    if (first && R->p.restart && !R->hist_given) imex_restart_histories(R);
    R->kstep++;
    if (R->kstep == 1) R->dt_new = R->p.dt;
    /* Nr_maxLOOP unless it is the first step of a fresh start (steptime.F90:202) */
    if (R->kstep > 1 || R->p.restart) {
        nonlin_run(R, R->tFR, R->omgFR, R->upFR, R->urFR);
        for (int r = 0; r < R->N; r++) R->up_prev[r] = R->upFR[(size_t)R->Nm1 * r];   /* :212 */
    }
    /* steptime.F90:224-231 */
    if (R->p.vardt && (R->kstep > 1 || R->p.restart)) {
        new_dt(R, 0);
    } else {
        R->dt_new = R->p.dt;
        for (int i = 0; i < 8; i++) R->dt_array[i] = R->p.dt;
    }
    const double dt = R->dt_new;
    const size_t S = SPEC(R);
    const int ni = R->sc.n_imp, ne = R->sc.n_exp;
    if (R->kstep > 1) {   /* steptime.F90:233-241 */
        memcpy(R->tmp_imp_temp, R->rhs_imp_temp, sizeof(cplx) * S * ni);
        memcpy(R->tmp_exp_temp, R->rhs_exp_temp, sizeof(cplx) * S * ne);
        memcpy(R->tmp_imp_vort, R->rhs_imp_vort, sizeof(cplx) * S * ni);
        memcpy(R->tmp_exp_vort, R->rhs_exp_vort, sizeof(cplx) * S * ne);
        memcpy(R->tmp_imp_ub, R->rhs_imp_ub, sizeof(cplx) * R->N * ni);
        memcpy(R->tmp_exp_ub, R->rhs_exp_ub, sizeof(cplx) * R->N * ne);
    }
    double wl, wi[4], we[4];
    imex_weights(&R->sc, R->dt_array, &wl, wi, we);
    if (R->kstep == 1) {   /* fourierloop.F90:78-84 */
        wsp W; wsp_alloc(&W, R->N, R->C.M);
        for (int m = 0; m < R->Nm1; m++) {
            get_rad(R, R->tFR, m, W.a); cheb_fwd(&R->C, W.a, W.b, W.w); put_rad(R, R->temp_spec, m, W.b);
            get_rad(R, R->omgFR, m, W.a); cheb_fwd(&R->C, W.a, W.b, W.w); put_rad(R, R->omg_spec, m, W.b);
            get_rad(R, R->upFR, m, W.a); cheb_fwd(&R->C, W.a, W.b, W.w); put_rad(R, R->upFC, m, W.b);
        }
        wsp_free(&W);
    }
    /* mat_build only on the first step or when dt changed (fourierloop.F90:86-94);
     * wt_lhs is the current one, updated just before (:74) */
    if (R->kstep == 1 || R->dt_array[0] != R->dt_array[1])
        build_factors(R, wl, R->dt_array[0], 0);
    const int nt = R->nthreads > 0 ? R->nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
    {
        wsp W;
        wsp_alloc(&W, R->N, R->C.M);
#pragma omp for schedule(static)
        for (int m = 0; m < R->Nm1; m++) imex_mode(R, m, &W, dt, wl, wi, we);
        wsp_free(&W);
    }
    R->tot_time = R->tot_time + dt;
}

/* ------------------------------------------------------------------------ */
/* IMEXRK step: timeloop_imexrk body (steptime_imexrk.F90:100-190)            */
/* ------------------------------------------------------------------------ */
#define STG(A, s) ((A) + (size_t)(s) * SPEC(R))

/* RHS_construct_stage for one mode (fourierloop_imexrk.F90:429-488) */
static void rk_rhs_mode(orc_run* R, int s, int m, wsp* W)
{
    const int N = R->N, Nm1 = R->Nm1;
    const chebctx* C = &R->C;
    const double* rr = R->G.r_radius;
    const double* rr2 = R->G.r_radius2;
    const double mm = (double)m;
    (void)Nm1;
    /* rhs_construct_temp_a, rhs_create_imexrk.F90:20-57 */
    get_rad(R, R->upT, m, W->a);
    get_rad(R, R->urT, m, W->b);
    cheb_fwd(C, W->b, W->c, W->w);
    cheb_d1(C, W->c, W->d, W->w);
    for (int i = 0; i < N; i++) {
        const cplx v = cadd(cadd(cimr(mm, rr[i], W->a[i]), W->d[i]), flux_over_r(rr[i], W->b[i]));
        W->e[i] = cx(-v.re, -v.im);
    }
    put_rad(R, STG(R->dT_a, s), m, W->e);
    /* rhs_construct_temp_d, :59-93 */
    get_rad(R, R->tFR, m, W->a);
    cheb_fwd(C, W->a, W->b, W->w);
    cheb_d1d2(C, W->b, W->c, W->d, W->w);
    for (int i = 0; i < N; i++) {
        const cplx lap = csub(cadd(cscale(rr[i], W->c[i]), W->d[i]), cscale((mm * mm) * rr2[i], W->a[i]));
        W->e[i] = cscale(1.0 / R->p.Pr, lap);
    }
    put_rad(R, STG(R->dT_d, s), m, W->e);
    if (m == 0) {
        /* rhs_construct_uphibar_a (:95-160): -<ur omega> from the physical arrays */
        for (int i = 0; i < N; i++) W->e[i] = cx(-R->urw_mean[i], 0.0);
        put_rad(R, STG(R->dW_a, s), 0, W->e);
        /* rhs_construct_uphibar_d (:162-191) */
        get_rad(R, R->upFR, 0, W->a);
        cheb_fwd(C, W->a, W->b, W->w);
        cheb_d1d2(C, W->b, W->c, W->d, W->w);
        for (int i = 0; i < N; i++) {
            const cplx v = cadd(csub(W->d[i], cscale(rr2[i], W->a[i])), cscale(rr[i], W->c[i]));
            W->e[i] = cx(v.re, 0.0);
        }
        put_rad(R, STG(R->dW_d, s), 0, W->e);
        return;
    }
    /* rhs_construct_vort_a (:194-242) */
    get_rad(R, R->upW, m, W->a);
    get_rad(R, R->urW, m, W->b);
    cheb_fwd(C, W->b, W->c, W->w);
    cheb_d1(C, W->c, W->d, W->w);
    for (int i = 0; i < N; i++) {
        const cplx v = cadd(cadd(cimr(mm, rr[i], W->a[i]), W->d[i]), flux_over_r(rr[i], W->b[i]));
        W->e[i] = cx(-v.re, -v.im);
    }
    put_rad(R, STG(R->dW_a, s), m, W->e);
    /* rhs_construct_vort_d (:244-289) */
    get_rad(R, R->omgFR, m, W->a);
    cheb_fwd(C, W->a, W->b, W->w);
    cheb_d1d2(C, W->b, W->c, W->d, W->w);
    for (int i = 0; i < N; i++)
        W->e[i] = csub(cadd(cscale(rr[i], W->c[i]), W->d[i]), cscale((mm * mm) * rr2[i], W->a[i]));
    put_rad(R, STG(R->dW_d, s), m, W->e);
    if (s == 0) {
        get_rad(R, R->tFR, m, W->a);
        for (int i = 0; i < N; i++) W->e[i] = cscale(-buo_ra_pr(R->p.Ra, R->p.Pr), cimr(mm, rr[i], W->a[i]));
        put_rad(R, STG(R->buo, 0), m, W->e);
    }
}

/* Get_stage_var for one mode (fourierloop_imexrk.F90:169-393); s is 0-based */
static void rk_stage_mode(orc_run* R, int s, int m, wsp* W)
{
    const int N = R->N, Nm1 = R->Nm1;
    const scheme* sc = &R->sc;
    const chebctx* C = &R->C;
    const double* rr = R->G.r_radius;
    const double mm = (double)m;
    const double dt = R->dt_array[0];
    const size_t n2 = (size_t)N * N;
    cplx* Fd = W->a; cplx* Fa = W->b; cplx* rhs = W->rhs;
    for (int i = 0; i < N; i++) { Fd[i] = cx(0, 0); Fa[i] = cx(0, 0); }
    for (int k = 0; k < s; k++)
        for (int i = 0; i < N; i++) {
            Fd[i] = cadd(Fd[i], cscale(sc->AD(s, k), STG(R->dT_d, k)[m + (size_t)Nm1 * i]));
        }
    for (int k = 0; k < s; k++)
        for (int i = 0; i < N; i++) {
            Fa[i] = cadd(Fa[i], cscale(sc->AA(s, k), STG(R->dT_a, k)[m + (size_t)Nm1 * i]));
        }
    for (int i = 0; i < N; i++)
        rhs[i] = cadd(R->temp_spec[m + (size_t)Nm1 * i], cscale(dt, cadd(Fd[i], Fa[i])));
    rhs[0] = cx(m == 0 ? 1.0 : 0.0, 0.0);
    rhs[N - 1] = cx(0.0, 0.0);
    solve_cplx(N, R->AT + n2 * m, R->pAT + (size_t)N * m, rhs, W->br, W->bi);
    cplx* Tk = W->c;
    cheb_inv(C, rhs, Tk, W->w);
    put_rad(R, R->tFR, m, Tk);
    if (m == 0) {
        for (int i = 0; i < N; i++) { Fd[i] = cx(0, 0); Fa[i] = cx(0, 0); }
        for (int k = 0; k < s; k++)
            for (int i = 0; i < N; i++) Fd[i] = cadd(Fd[i], cscale(sc->AD(s, k), STG(R->dW_d, k)[(size_t)Nm1 * i]));
        for (int k = 0; k < s; k++)
            for (int i = 0; i < N; i++) Fa[i] = cadd(Fa[i], cscale(sc->AA(s, k), STG(R->dW_a, k)[(size_t)Nm1 * i]));
        for (int i = 0; i < N; i++) rhs[i] = cadd(R->uphi_bar_spec[i], cscale(dt, cadd(Fd[i], Fa[i])));
        rhs[0] = cx(0.0, 0.0);
        rhs[N - 1] = cx(0.0, 0.0);
        for (int i = 0; i < N; i++) { W->br[i] = rhs[i].re; W->bi[i] = 0.0; }
        orc_dgetrs(N, R->Au, R->pAu, W->br);
        orc_dgetrs(N, R->Au, R->pAu, W->bi);
        cplx* ufc = W->d; cplx* ub = W->e; cplx* dub = W->f;
        for (int i = 0; i < N; i++) ufc[i] = cx(W->br[i], 0.0);
        put_rad(R, R->upFC, 0, ufc);
        cheb_inv(C, ufc, ub, W->w);
        cheb_d1(C, ufc, dub, W->w);
        put_rad(R, R->upFR, 0, ub);
        for (int i = 0; i < N; i++) {
            R->urFR[(size_t)Nm1 * i] = cx(0.0, 0.0);
            R->omgFR[(size_t)Nm1 * i] = cadd(cscale(rr[i], ub[i]), dub[i]);
        }
        return;
    }
    for (int i = 0; i < N; i++)
        STG(R->buo, s)[m + (size_t)Nm1 * i] = cscale(-buo_ra_pr(R->p.Ra, R->p.Pr), cimr(mm, rr[i], Tk[i]));
    cplx* Fb = W->d;
    for (int i = 0; i < N; i++) { Fd[i] = cx(0, 0); Fa[i] = cx(0, 0); Fb[i] = cx(0, 0); }
    for (int k = 0; k < s; k++)
        for (int i = 0; i < N; i++) Fd[i] = cadd(Fd[i], cscale(sc->AD(s, k), STG(R->dW_d, k)[m + (size_t)Nm1 * i]));
    for (int k = 0; k < s; k++)
        for (int i = 0; i < N; i++) Fa[i] = cadd(Fa[i], cscale(sc->AA(s, k), STG(R->dW_a, k)[m + (size_t)Nm1 * i]));
    if (R->p.buo_imp) {
        for (int k = 0; k <= s; k++)
            for (int i = 0; i < N; i++) Fb[i] = cadd(Fb[i], cscale(sc->AD(s, k), STG(R->buo, k)[m + (size_t)Nm1 * i]));
    } else {
        for (int k = 0; k < s; k++)
            for (int i = 0; i < N; i++) Fb[i] = cadd(Fb[i], cscale(sc->AA(s, k), STG(R->buo, k)[m + (size_t)Nm1 * i]));
    }
    cplx* rhsf = W->rhsf;
    for (int i = 0; i < N; i++) {
        rhsf[N + i] = cadd(R->omg_spec[m + (size_t)Nm1 * i], cscale(dt, cadd(cadd(Fd[i], Fa[i]), Fb[i])));
        rhsf[i] = cx(0.0, 0.0);
    }
    rhsf[N] = cx(0.0, 0.0);
    rhsf[2 * N - 1] = cx(0.0, 0.0);
    solve_cplx(2 * N, R->AF + 4 * n2 * m, R->pAF + (size_t)2 * N * m, rhsf, W->br, W->bi);
    cplx* wphys = W->e; cplx* psi = W->f; cplx* dpsi = W->g;
    cheb_inv(C, rhsf + N, wphys, W->w);
    put_rad(R, R->omgFR, m, wphys);
    cheb_inv(C, rhsf, psi, W->w);
    cheb_d1(C, rhsf, dpsi, W->w);
    for (int i = 0; i < N; i++) {
        const size_t o = (size_t)m + (size_t)Nm1 * i;
        R->psii[o] = psi[i];
        R->upFR[o] = cscale(-1.0, dpsi[i]);
        R->urFR[o] = cimr(mm, rr[i], psi[i]);
    }
    get_rad(R, R->upFR, m, W->a);
    cheb_fwd(C, W->a, W->b, W->w);
    put_rad(R, R->upFC, m, W->b);
}

/* dlagrange / d2lagrange (chebyshev_fftw.F90:449-517): weights of the first
 * and second derivative at x of the Lagrange interpolant through xp[0..n) */
static void lagrange_d1(int n, const double* xp, double x, double* dw)
{
    for (int j = 0; j < n; j++) dw[j] = 0.0;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            if (i == j) continue;
            double w = 1.0;
            for (int m = 0; m < n; m++)
                if (m != i && m != j) w = w * (x - xp[m]) / (xp[j] - xp[m]);
            dw[j] = dw[j] + w / (xp[j] - xp[i]);
        }
}

static void lagrange_d2(int n, const double* xp, double x, double* d2w)
{
    for (int j = 0; j < n; j++) d2w[j] = 0.0;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            if (i == j) continue;
            double mw = 0.0;
            for (int m = 0; m < n; m++) {
                if (m == i || m == j) continue;
                double w = 1.0;
                for (int l = 0; l < n; l++)
                    if (l != j && l != i && l != m) w = w * (x - xp[l]) / (xp[j] - xp[l]);
                mw = mw + w / (xp[j] - xp[m]);
            }
            d2w[j] = d2w[j] + mw / (xp[j] - xp[i]);
        }
}

/* mat_build_rk (mat_assembly.F90:415-462): LAPpsi of mode n, the psi rows of
 * the Laplacian with psi = 0 at both walls (the same rows for NS and SF) */
static void mat_build_lap(const chebgrid* G, int n, double* L)
{
    const int N = G->N;
    const double C = sqrt(2.0 / (double)(N - 1));
    const double nn = (double)n;
    for (int j = 0; j < N; j++) {
        for (int i = 0; i < N; i++)
            L[IX(i, j, N)] = (G->r_radius[i] * G->D[IX(i, j, N)] + G->D2[IX(i, j, N)]) -
                             ((nn * nn) * G->r_radius2[i]) * G->t[IX(i, j, N)];
        L[IX(0, j, N)] = G->t[IX(0, j, N)];
        L[IX(N - 1, j, N)] = G->t[IX(N - 1, j, N)];
    }
    for (int i = 0; i < N; i++) {
        L[IX(i, 0, N)] = 0.5 * L[IX(i, 0, N)];
        L[IX(i, N - 1, N)] = 0.5 * L[IX(i, N - 1, N)];
    }
    for (size_t k = 0; k < (size_t)N * N; k++) L[k] = C * L[k];
}

void orc_mat_build_lap(int N, double eta, int n, double* L)
{
    chebgrid G;
    grid_init(&G, N, eta);
    mat_build_lap(&G, n, L);
    grid_free(&G);
}

/* LAPpsi factors of every mode (steptime_imexrk.F90:92-94) and the Lagrange
 * weights (init.F90:268-271): nodes radius(2 .. lm+1) for r_min and
 * radius(Nr_max-1 .. Nr_max-lm) for r_max */
static int assembly_setup(orc_run* R)
{
    const int N = R->N, lm = R->p.lagpts;
    const size_t n2 = (size_t)N * N;
    R->LAP = (double*)malloc(sizeof(double) * n2 * R->Nm1);
    R->pLAP = (int32_t*)malloc(sizeof(int32_t) * (size_t)N * R->Nm1);
    if (!R->LAP || !R->pLAP) return -4;
    for (int m = 0; m < R->Nm1; m++) {
        mat_build_lap(&R->G, m, R->LAP + n2 * m);
        orc_dgetrf(N, R->LAP + n2 * m, R->pLAP + (size_t)N * m);
    }
    double xmin[64], xmax[64];
    for (int i = 0; i < lm; i++) {
        xmin[i] = R->G.radius[1 + i];
        xmax[i] = R->G.radius[N - 2 - i];
    }
    lagrange_d1(lm, xmin, R->rmin, R->dw_rmin);
    lagrange_d2(lm, xmin, R->rmin, R->d2w_rmin);
    lagrange_d1(lm, xmax, R->rmax, R->dw_rmax);
    lagrange_d2(lm, xmax, R->rmax, R->d2w_rmax);
    R->lap_built = 1;
    return 0;
}

/* The m >= 1 tail shared by Assembly_stage (fourierloop_imexrk.F90:590-651)
 * and the RK path's Get_stage_var and Assembly_stage (fourierloop_rk.F90:
 * 158-210, 363-411): psi from LAPpsi psi = -omega with psi = 0 at the walls
 * (omega's wall values do not enter), u_phi = -dpsi/dr, u_r = i m psi / r,
 * omega at the walls from the Lagrange extrapolation of psi's interior values
 * (Johnston's strategy, no slip) or sf[k] dpsi/dr (stress free), and
 * upFC(m) = chebtransform(u_phi).  om: the array that holds omega (omg_spec
 * or omgFR); its wall values are set.  psii, if given, receives psi. */
static void psi_tail(orc_run* R, int m, cplx* om, const double sf[2], cplx* psii, wsp* W)
{
    const int N = R->N, lm = R->p.lagpts;
    const chebctx* C = &R->C;
    const double* rr = R->G.r_radius;
    const double mm = (double)m;
    cplx* psi = W->b;
    for (int r = 0; r < N; r++) {
        const cplx w = om[m + (size_t)R->Nm1 * r];
        psi[r] = cx(-1.0 * w.re, -1.0 * w.im);
    }
    psi[0] = cx(0.0, 0.0);
    psi[N - 1] = cx(0.0, 0.0);
    const size_t n2 = (size_t)N * N;
    solve_cplx(N, R->LAP + n2 * m, R->pLAP + (size_t)N * m, psi, W->br, W->bi);
    cplx* pv = W->c; cplx* dp = W->d; cplx* d2p = W->e;
    cheb_inv(C, psi, pv, W->w);
    cheb_d1(C, psi, dp, W->w);
    cheb_d2(C, psi, d2p, W->w);
    cplx d2min = cx(0.0, 0.0), dmin = cx(0.0, 0.0), d2max = cx(0.0, 0.0), dmax = cx(0.0, 0.0);
    for (int i = 0; i < lm; i++) {
        d2min = cadd(d2min, cscale(R->d2w_rmin[i], pv[1 + i]));
        dmin = cadd(dmin, cscale(R->dw_rmin[i], pv[1 + i]));
    }
    for (int i = 0; i < lm; i++) {
        d2max = cadd(d2max, cscale(R->d2w_rmax[i], pv[N - 2 - i]));
        dmax = cadd(dmax, cscale(R->dw_rmax[i], pv[N - 2 - i]));
    }
    for (int r = 0; r < N; r++) {
        const size_t o = m + (size_t)R->Nm1 * r;
        if (psii) psii[o] = pv[r];
        R->upFR[o] = cx(-1.0 * dp[r].re, -1.0 * dp[r].im);
        R->urFR[o] = cimr(mm, rr[r], pv[r]);
    }
    cplx* w0 = om + m;
    cplx* w1 = om + m + (size_t)R->Nm1 * (N - 1);
    if (R->p.mBC == 0) {
        const cplx a = csub(d2min, cscale(2.0 * R->dw_rmin[0], dmin));
        const cplx b = csub(d2max, cscale(2.0 * R->dw_rmax[0], dmax));
        *w0 = cx(-1.0 * a.re, -1.0 * a.im);
        *w1 = cx(-1.0 * b.re, -1.0 * b.im);
    } else {
        *w0 = cscale(sf[0], dmin);
        *w1 = cscale(sf[1], dmax);
    }
    get_rad(R, R->upFR, m, W->f);
    cheb_fwd(C, W->f, W->g, W->w);
    put_rad(R, R->upFC, m, W->g);
}

/* Assembly_stage for one mode (fourierloop_imexrk.F90:499-651): the new state
 * from the stage start and the b-weighted stage terms; omega's interior, then
 * psi from LAPpsi psi = -omega, the velocities, and omega at the walls from
 * the Lagrange extrapolation of psi (Johnston's strategy). */
static void assemble_mode(orc_run* R, int m, wsp* W, double dt)
{
    const int N = R->N, S = R->sc.S;
    const scheme* sc = &R->sc;
    const chebctx* C = &R->C;
    const double* rr = R->G.r_radius;
    for (int r = 0; r < N; r++) {
        cplx F = cx(0.0, 0.0);
        for (int i = 0; i < S; i++)
            F = cadd(cadd(F, cscale(sc->bA[i], *hist(R->dT_a, R, i, m, r))), cscale(sc->bD[i], *hist(R->dT_d, R, i, m, r)));
        cplx* T = R->temp_spec + m + (size_t)R->Nm1 * r;
        *T = cadd(*T, cscale(dt, F));
    }
    R->temp_spec[m] = cx(m == 0 ? 1.0 : 0.0, 0.0);
    R->temp_spec[m + (size_t)R->Nm1 * (N - 1)] = cx(0.0, 0.0);
    if (m == 0) {
        for (int r = 0; r < N; r++) {
            cplx F = cx(0.0, 0.0);
            for (int i = 0; i < S; i++)
                F = cadd(cadd(F, cscale(sc->bA[i], *hist(R->dW_a, R, i, 0, r))), cscale(sc->bD[i], *hist(R->dW_d, R, i, 0, r)));
            R->uphi_bar_spec[r] = cadd(R->uphi_bar_spec[r], cscale(dt, F));
        }
        R->uphi_bar_spec[0] = cx(0.0, 0.0);
        R->uphi_bar_spec[N - 1] = cx(0.0, 0.0);
        put_rad(R, R->upFR, 0, R->uphi_bar_spec);
        cheb_fwd(C, R->uphi_bar_spec, W->a, W->w);
        put_rad(R, R->upFC, 0, W->a);
        cheb_d1(C, W->a, W->b, W->w);
        for (int r = 0; r < N; r++) {
            R->urFR[(size_t)R->Nm1 * r] = cx(0.0, 0.0);
            R->omg_spec[(size_t)R->Nm1 * r] = cadd(cscale(rr[r], R->uphi_bar_spec[r]), W->b[r]);
        }
        return;
    }
    for (int i = 0; i < S; i++)
        for (int r = 1; r < N - 1; r++) {
            cplx* F = W->a + r;
            if (i == 0) *F = cx(0.0, 0.0);
            *F = cadd(cadd(cadd(*F, cscale(sc->bA[i], *hist(R->dW_a, R, i, m, r))),
                           cscale(sc->bD[i], *hist(R->dW_d, R, i, m, r))),
                      cscale(sc->bD[i], *hist(R->buo, R, i, m, r)));
        }
    for (int r = 1; r < N - 1; r++) {
        cplx* w = R->omg_spec + m + (size_t)R->Nm1 * r;
        *w = cadd(*w, cscale(dt, W->a[r]));
    }
    /* -2 / radius at the walls for stress free (fourierloop_imexrk.F90:634-635) */
    const double sf[2] = {-2.0 / R->G.radius[0], -2.0 / R->G.radius[N - 1]};
    psi_tail(R, m, R->omg_spec, sf, NULL, W);
}

static void imexrk_step(orc_run* R)
{
    const scheme* sc = &R->sc;
    R->kstep++;
    /* steptime_imexrk.F90:102-138 */
    if (R->kstep == 1) {
        R->dt_new = R->p.dt;
        for (int i = 0; i < 8; i++) R->dt_array[i] = R->p.dt;
    }
    if (R->p.vardt) {
        if (R->kstep == 1 && R->p.restart) {
            nonlin_run(R, R->tFR, R->omgFR, R->upFR, R->urFR);
            new_dt(R, 1);
        }
        if (R->kstep > 1) {
            /* stiffly accurate schemes: get_dtvals of the previous step's last
             * pass; the others take a pass over the state first (:118-125) */
            if (!R->sc.sa) nonlin_run(R, R->tFR, R->omgFR, R->upFR, R->urFR);
            new_dt(R, 1);
        }
    } else {
        R->dt_new = R->p.dt;
        for (int i = 0; i < 8; i++) R->dt_array[i] = R->p.dt;
    }
    const double dt = R->dt_new;
    const int nt = R->nthreads > 0 ? R->nthreads : omp_get_max_threads();
    for (int s = 0; s < sc->S; s++) {
        if (s > 0) {
            /* diag_diff: the stage's own diagonal (fourierloop_imexrk.F90:134-158);
             * the reference keeps one factor set per stage, rebuilt when dt
             * changes; rebuilding here when the weight changes gives the same
             * factors */
            const double wt = sc->diag_diff ? sc->AD(s, s) : sc->AD(sc->diag_index - 1, sc->diag_index - 1);
            if (!R->mats_built || R->mat_dt != R->dt_array[0] || R->mat_wt != wt)
                build_factors(R, wt, R->dt_array[0], 1);
#pragma omp parallel num_threads(nt)
            {
                wsp W;
                wsp_alloc(&W, R->N, R->C.M);
#pragma omp for schedule(static)
                for (int m = 0; m < R->Nm1; m++) rk_stage_mode(R, s, m, &W);
                wsp_free(&W);
            }
        }
        if ((R->kstep > 1 || s > 0 || R->p.restart) && sc->nl_col[s])
            nonlin_run(R, R->tFR, R->omgFR, R->upFR, R->urFR);
        {   /* RHS_construct_stage: serial over modes in the reference */
            wsp W;
            wsp_alloc(&W, R->N, R->C.M);
            for (int m = 0; m < R->Nm1; m++) rk_rhs_mode(R, s, m, &W);
            wsp_free(&W);
        }
    }
    if (sc->sa) {
        /* Assembly_stage_SA (fourierloop_imexrk.F90:653-665) */
        memcpy(R->temp_spec, R->tFR, sizeof(cplx) * SPEC(R));
        memcpy(R->omg_spec, R->omgFR, sizeof(cplx) * SPEC(R));
        for (int r = 0; r < R->N; r++) R->uphi_bar_spec[r] = R->upFR[(size_t)R->Nm1 * r];
    } else {
        /* Assembly_stage (fourierloop_imexrk.F90:499-651), then tFR = temp_spec,
         * omgFR = omg_spec */
#pragma omp parallel num_threads(nt)
        {
            wsp W;
            wsp_alloc(&W, R->N, R->C.M);
#pragma omp for schedule(static)
            for (int m = 0; m < R->Nm1; m++) assemble_mode(R, m, &W, dt);
            wsp_free(&W);
        }
        memcpy(R->tFR, R->temp_spec, sizeof(cplx) * SPEC(R));
        memcpy(R->omgFR, R->omg_spec, sizeof(cplx) * SPEC(R));
    }
    R->tot_time = R->tot_time + dt;
    /* TFC for the Nusselt number (steptime_imexrk.F90:190) */
    wsp W;
    wsp_alloc(&W, R->N, R->C.M);
    get_rad(R, R->tFR, 0, W.a);
    cheb_fwd(&R->C, W.a, R->TFC, W.w);
    wsp_free(&W);
}

/* ------------------------------------------------------------------------ */
/* The explicit RK path: steptime_rk.F90, fourierloop_rk.F90,                 */
/* rhs_create_rk.F90                                                          */
/* ------------------------------------------------------------------------ */

/* Get_stage_var (s < S, weights a(s,:), dt_array(1)) or Assembly_stage
 * (s == S, weights b, dt_new) of the RK path for one mode
 * (fourierloop_rk.F90:69-224, 285-424): the stage value, or the new state,
 * from the step's start (temp_spec, omg_spec, uphi_bar_spec: physical
 * values) plus dt times the weighted stage terms, all of them explicit. */
static void rkx_update_mode(orc_run* R, int s, int m, wsp* W, double dt)
{
    const int N = R->N, S = R->sc.S, Nm1 = R->Nm1;
    const scheme* sc = &R->sc;
    const chebctx* C = &R->C;
    const double* rr = R->G.r_radius;
    const int fin = (s == S);
    const int K = fin ? S : s;
    double wk[MAXS];
    for (int k = 0; k < K; k++) wk[k] = fin ? sc->bA[k] : sc->AA(s, k);
    cplx* T = fin ? R->temp_spec : R->tFR;
    for (int r = 0; r < N; r++) {
        cplx F = cx(0.0, 0.0);
        for (int k = 0; k < K; k++) F = cadd(F, cscale(wk[k], STG(R->dT_d, k)[m + (size_t)Nm1 * r]));
        T[m + (size_t)Nm1 * r] = cadd(R->temp_spec[m + (size_t)Nm1 * r], cscale(dt, F));
    }
    T[m] = cx(m == 0 ? 1.0 : 0.0, 0.0);
    T[m + (size_t)Nm1 * (N - 1)] = cx(0.0, 0.0);
    cplx* om = fin ? R->omg_spec : R->omgFR;
    if (m == 0) {
        cplx* u = W->h;
        for (int r = 0; r < N; r++) {
            cplx F = cx(0.0, 0.0);
            for (int k = 0; k < K; k++) F = cadd(F, cscale(wk[k], STG(R->dW_d, k)[(size_t)Nm1 * r]));
            u[r] = cadd(R->uphi_bar_spec[r], cscale(dt, F));
        }
        u[0] = cx(0.0, 0.0);
        u[N - 1] = cx(0.0, 0.0);
        if (fin) memcpy(R->uphi_bar_spec, u, sizeof(cplx) * (size_t)N);
        else put_rad(R, R->upFR, 0, u);
        cheb_fwd(C, u, W->a, W->w);
        put_rad(R, R->upFC, 0, W->a);
        cheb_d1(C, W->a, W->b, W->w);
        for (int r = 0; r < N; r++) {
            R->urFR[(size_t)Nm1 * r] = cx(0.0, 0.0);
            om[(size_t)Nm1 * r] = cadd(cscale(rr[r], u[r]), W->b[r]);
        }
        return;
    }
    for (int r = 0; r < N; r++) {
        cplx F = cx(0.0, 0.0);
        for (int k = 0; k < K; k++) F = cadd(F, cscale(wk[k], STG(R->dW_d, k)[m + (size_t)Nm1 * r]));
        om[m + (size_t)Nm1 * r] = cadd(R->omg_spec[m + (size_t)Nm1 * r], cscale(dt, F));
    }
    /* stress free: -2 r_radius at the walls (fourierloop_rk.F90:204-208) */
    const double sf[2] = {-2.0 * rr[0], -2.0 * rr[N - 1]};
    psi_tail(R, m, om, sf, R->psii, W);
}

/* The Fourier-space sums of rhs_construct_uphi_bar (rhs_create_rk.F90:
 * 160-190), per radial level: su = sum over m of Re(ur_m conj(u_m) +
 * conj(ur_m) u_m) (Re(ur_0 u_0) at m = 0) with u = u_phi, and sd the same
 * with u = d u_phi / dr = chebinvtranD1(upFC(m)).  The physical-space
 * products that routine also forms do not enter its result. */
static void rkx_mean_sums(orc_run* R, double* su, double* sd, wsp* W)
{
    const int N = R->N, Nm1 = R->Nm1;
    cplx* d1 = calloc_c((size_t)Nm1 * N);
    for (int m = 0; m < Nm1; m++) {
        get_rad(R, R->upFC, m, W->a);
        cheb_d1(&R->C, W->a, W->b, W->w);
        for (int r = 0; r < N; r++) d1[m + (size_t)Nm1 * r] = W->b[r];
    }
    for (int r = 0; r < N; r++) {
        double a = 0.0, b = 0.0;
        for (int m = 0; m < Nm1; m++) {
            const size_t o = m + (size_t)Nm1 * r;
            const cplx ur = R->urFR[o], up = R->upFR[o], du = d1[o];
            if (m == 0) {
                a = a + cmul(ur, up).re;
                b = b + cmul(ur, du).re;
            } else {
                a = a + cadd(cmul(ur, cconj(up)), cmul(cconj(ur), up)).re;
                b = b + cadd(cmul(ur, cconj(du)), cmul(cconj(ur), du)).re;
            }
        }
        su[r] = a;
        sd[r] = b;
    }
    free(d1);
}

/* RHS_construct_stage of the RK path for one mode (fourierloop_rk.F90:226-283):
 * the whole right-hand side of stage s into dT_d / dW_d slot s --
 * rhs_construct_temp (rhs_create_rk.F90:19-67), rhs_construct_uphi_bar at
 * m = 0 (:131-224), rhs_construct_vort at m >= 1 (:69-129). */
static void rkx_rhs_mode(orc_run* R, int s, int m, wsp* W, const double* su, const double* sd)
{
    const int N = R->N;
    const chebctx* C = &R->C;
    const double* rr = R->G.r_radius;
    const double* rr2 = R->G.r_radius2;
    const double mm = (double)m;
    const double Pr = R->p.Pr;
    get_rad(R, R->upT, m, W->a);
    get_rad(R, R->urT, m, W->b);
    get_rad(R, R->tFR, m, W->f);
    cheb_fwd(C, W->b, W->c, W->w);
    cheb_fwd(C, W->f, W->g, W->w);
    cheb_d1d2(C, W->g, W->d, W->e, W->w);
    cheb_d1(C, W->c, W->g, W->w);
    for (int i = 0; i < N; i++) {
        const cplx lap = csub(cadd(cscale(rr[i], W->d[i]), W->e[i]), cscale((mm * mm) * rr2[i], W->f[i]));
        const cplx adv = cadd(cadd(cimr(mm, rr[i], W->a[i]), W->g[i]), flux_over_r(rr[i], W->b[i]));
        W->h[i] = csub(cscale(1.0 / Pr, lap), adv);
    }
    put_rad(R, STG(R->dT_d, s), m, W->h);
    if (m == 0) {
        get_rad(R, R->upFC, 0, W->a);
        cheb_d1(C, W->a, W->b, W->w);
        cheb_d2(C, W->a, W->c, W->w);
        get_rad(R, R->upFR, 0, W->d);
        for (int i = 0; i < N; i++) {
            const cplx ud = cadd(csub(W->c[i], cscale(rr2[i], W->d[i])), cscale(rr[i], W->b[i]));
            W->h[i] = cx(ud.re - (sd[i] + rr[i] * su[i]), 0.0);
        }
        put_rad(R, STG(R->dW_d, s), 0, W->h);
        return;
    }
    get_rad(R, R->upW, m, W->a);
    get_rad(R, R->urW, m, W->b);
    get_rad(R, R->omgFR, m, W->f);
    cheb_fwd(C, W->b, W->c, W->w);
    cheb_fwd(C, W->f, W->g, W->w);
    cheb_d1d2(C, W->g, W->d, W->e, W->w);
    cheb_d1(C, W->c, W->g, W->w);
    const double raPr = buo_ra_pr(R->p.Ra, Pr);
    for (int i = 0; i < N; i++) {
        const cplx lap = csub(cadd(cscale(rr[i], W->d[i]), W->e[i]), cscale((mm * mm) * rr2[i], W->f[i]));
        const cplx adv = cadd(cadd(cimr(mm, rr[i], W->a[i]), W->g[i]), flux_over_r(rr[i], W->b[i]));
        const cplx buo = cscale(raPr, cimr(mm, rr[i], R->tFR[m + (size_t)R->Nm1 * i]));
        W->h[i] = csub(csub(lap, adv), buo);
    }
    put_rad(R, STG(R->dW_d, s), m, W->h);
}

/* The mean-flow right-hand side rhs_construct_uphi_bar of the current state
 * (the RK path; out: N complex values, imaginary parts zero).  Uses the
 * stage-1 slot of the stage terms as scratch; a step rebuilds it. */
int orc_rk_mean_rhs(orc_run* R, double* out)
{
    if (R->sc.type != ST_RK) return -1;
    double* su = (double*)calloc((size_t)2 * R->N, sizeof(double));
    if (!su) return -4;
    wsp W;
    wsp_alloc(&W, R->N, R->C.M);
    rkx_mean_sums(R, su, su + R->N, &W);
    rkx_rhs_mode(R, 0, 0, &W, su, su + R->N);
    wsp_free(&W);
    free(su);
    for (int r = 0; r < R->N; r++) ((cplx*)out)[r] = STG(R->dW_d, 0)[(size_t)R->Nm1 * r];
    return 0;
}

/* timeloop_rk (steptime_rk.F90:86-149): dt_array = dt every step; per stage
 * the stage values (from stage 2), the nonlinear pass (except at stage 1 of a
 * fresh start's first step) and the stage's right-hand sides; then
 * Assembly_stage, tFR = temp_spec, omgFR = omg_spec, upFR(1,:) =
 * uphi_bar_spec, and TFC. */
static void rkx_step(orc_run* R)
{
    const scheme* sc = &R->sc;
    R->kstep++;
    R->dt_new = R->p.dt;
    for (int i = 0; i < 8; i++) R->dt_array[i] = R->p.dt;
    if (R->p.vardt) new_dt_rk(R);
    const double dt = R->dt_new;
    const int nt = R->nthreads > 0 ? R->nthreads : omp_get_max_threads();
    double* su = (double*)calloc((size_t)2 * R->N, sizeof(double));
    double* sd = su + R->N;
    for (int s = 0; s < sc->S; s++) {
        if (s > 0) {
#pragma omp parallel num_threads(nt)
            {
                wsp W;
                wsp_alloc(&W, R->N, R->C.M);
#pragma omp for schedule(static)
                for (int m = 0; m < R->Nm1; m++) rkx_update_mode(R, s, m, &W, R->dt_array[0]);
                wsp_free(&W);
            }
        }
        if (R->kstep > 1 || s > 0 || R->p.restart) nonlin_run(R, R->tFR, R->omgFR, R->upFR, R->urFR);
        {
            wsp W;
            wsp_alloc(&W, R->N, R->C.M);
            rkx_mean_sums(R, su, sd, &W);
            wsp_free(&W);
        }
#pragma omp parallel num_threads(nt)
        {
            wsp W;
            wsp_alloc(&W, R->N, R->C.M);
#pragma omp for schedule(static)
            for (int m = 0; m < R->Nm1; m++) rkx_rhs_mode(R, s, m, &W, su, sd);
            wsp_free(&W);
        }
    }
    free(su);
#pragma omp parallel num_threads(nt)
    {
        wsp W;
        wsp_alloc(&W, R->N, R->C.M);
#pragma omp for schedule(static)
        for (int m = 0; m < R->Nm1; m++) rkx_update_mode(R, sc->S, m, &W, dt);
        wsp_free(&W);
    }
    memcpy(R->tFR, R->temp_spec, sizeof(cplx) * SPEC(R));
    memcpy(R->omgFR, R->omg_spec, sizeof(cplx) * SPEC(R));
    for (int r = 0; r < R->N; r++) R->upFR[(size_t)R->Nm1 * r] = R->uphi_bar_spec[r];
    R->tot_time = R->tot_time + dt;
    wsp W;
    wsp_alloc(&W, R->N, R->C.M);
    get_rad(R, R->tFR, 0, W.a);
    cheb_fwd(&R->C, W.a, R->TFC, W.w);
    wsp_free(&W);
}

int orc_step(orc_run* R, int nsteps)
{
    for (int k = 0; k < nsteps; k++) {
        if (R->sc.type == ST_IMEX) imex_step(R);
        else if (R->sc.type == ST_RK) rkx_step(R);
        else imexrk_step(R);
    }
    return 0;
}

/* Restart start-up from the states of K consecutive checkpoints, oldest
 * first (get_checkpoint_data, init.F90:447-592; steptime.F90:87-193).
 * IMEX with l_imexrk_started: K = n_order_tscheme_imp (CN, BDF2: 2; BDF3: 3;
 * BDF4: 4); the j-th state fills implicit slot n_imp - j and explicit slot
 * n_exp - j (1-based), the newest becomes the current state.  IMEX without
 * it, and IMEXRK: K = 1 and the histories start at zero (the reference copies
 * history arrays that read_checkpoint never reads, init.F90:616-716, 926-931).
 * t is the newest checkpoint's tot_time. */
int orc_restart(orc_run* R, int K, const double* tFR, const double* omgFR, const double* urFR,
                const double* upFR, double t)
{
    const size_t S = SPEC(R);
    const int ni = R->sc.n_imp, ne = R->sc.n_exp;
    if (!R->p.restart || K < 1) return -1;
    if (R->sc.type == ST_IMEX ? (K != 1 && K != ni) : K != 1) return -1;
    if (R->sc.type == ST_IMEX) {
        memset(R->rhs_imp_temp, 0, sizeof(cplx) * S * ni); memset(R->rhs_exp_temp, 0, sizeof(cplx) * S * ne);
        memset(R->rhs_imp_vort, 0, sizeof(cplx) * S * ni); memset(R->rhs_exp_vort, 0, sizeof(cplx) * S * ne);
        memset(R->rhs_buo_term, 0, sizeof(cplx) * S * ni);
        memset(R->rhs_imp_ub, 0, sizeof(cplx) * R->N * ni); memset(R->rhs_exp_ub, 0, sizeof(cplx) * R->N * ne);
        if (K > 1) {
            for (int j = 0; j < K; j++) {
                memcpy(R->tFR, tFR + 2 * S * j, sizeof(cplx) * S);
                memcpy(R->omgFR, omgFR + 2 * S * j, sizeof(cplx) * S);
                memcpy(R->urFR, urFR + 2 * S * j, sizeof(cplx) * S);
                memcpy(R->upFR, upFR + 2 * S * j, sizeof(cplx) * S);
                nonlin_run(R, R->tFR, R->omgFR, R->upFR, R->urFR);
                restart_fill(R, ni - 1 - j, ne - 1 - j);
            }
        }
    }
    const size_t last = 2 * S * (size_t)(K - 1);
    orc_set_state(R, tFR + last, omgFR + last, urFR + last, upFR + last);
    R->tot_time = t;
    R->hist_given = 1;
    return 0;
}

/* The history arrays a checkpoint stores (store_checkpoint, output.F90:359-404)
 * in the reference's layout: rhs_*_temp / rhs_*_vort (n, Nm+1, Nr) and
 * rhs_*_uphi_bar (n, Nr), the order index fastest, slot 1 the newest.  IMEXRK
 * runs: one zero slot each (those arrays are (1, Nm+1, Nr) on that path and
 * never assigned, timeschemes.F90:452-545, rhs_create_imexrk.F90:7-8). */
int orc_get_history(const orc_run* R, double* imp_temp, double* exp_temp, double* imp_vort, double* exp_vort,
                    double* imp_ub, double* exp_ub)
{
    const size_t S = SPEC(R);
    const int N = R->N;
    if (R->sc.type != ST_IMEX) {
        /* RK: n_order_tscheme slots, as store_checkpoint declares them; zero
         * (the one-slot rhs_imp_* are read past their end there, and only the
         * first rhs_exp_* slot is ever assigned, fourierloop_rk.F90:120, 147,
         * 213; restarts never read them) */
        const size_t n = R->sc.type == ST_RK ? (size_t)R->sc.S : 1;
        memset(imp_temp, 0, sizeof(cplx) * S * n); memset(exp_temp, 0, sizeof(cplx) * S * n);
        memset(imp_vort, 0, sizeof(cplx) * S * n); memset(exp_vort, 0, sizeof(cplx) * S * n);
        memset(imp_ub, 0, sizeof(cplx) * N * n); memset(exp_ub, 0, sizeof(cplx) * N * n);
        return 0;
    }
    const int ni = R->sc.n_imp, ne = R->sc.n_exp;
    cplx *it = (cplx*)imp_temp, *et = (cplx*)exp_temp, *iw = (cplx*)imp_vort, *ew = (cplx*)exp_vort;
    for (size_t e = 0; e < S; e++) {
        for (int k = 0; k < ni; k++) {
            it[(size_t)k + ni * e] = R->rhs_imp_temp[(size_t)k * S + e];
            iw[(size_t)k + ni * e] = R->rhs_imp_vort[(size_t)k * S + e];
        }
        for (int k = 0; k < ne; k++) {
            et[(size_t)k + ne * e] = R->rhs_exp_temp[(size_t)k * S + e];
            ew[(size_t)k + ne * e] = R->rhs_exp_vort[(size_t)k * S + e];
        }
    }
    for (int i = 0; i < N; i++) {
        for (int k = 0; k < ni; k++) ((cplx*)imp_ub)[(size_t)k + ni * (size_t)i] = R->rhs_imp_ub[(size_t)k * N + i];
        for (int k = 0; k < ne; k++) ((cplx*)exp_ub)[(size_t)k + ne * (size_t)i] = R->rhs_exp_ub[(size_t)k * N + i];
    }
    return 0;
}

/* writeKE_spectral KE and Nusselt parts (output.F90:232-265) */
int orc_diag(orc_run* R, double* out)
{
    const int N = R->N, Nm1 = R->Nm1;
    double* ur2 = (double*)malloc(sizeof(double) * 2 * (size_t)N);
    double* up2 = ur2 + N;
    for (int j = 0; j < N; j++) {
        double a = 0.0, b = 0.0;
        for (int i = 0; i < Nm1; i++) {
            const cplx zr = R->urFR[i + (size_t)Nm1 * j];
            const cplx zp = R->upFR[i + (size_t)Nm1 * j];
            const cplx sr = cmul(zr, zr), sp = cmul(zp, zp);
            a = a + hypot(sr.re, sr.im);
            b = b + hypot(sp.re, sp.im);
        }
        ur2[j] = a;
        up2[j] = b;
    }
    cplx* w = (cplx*)malloc(sizeof(cplx) * (size_t)(2 * R->C.M + 4 * N));
    const double pi = 4.0 * atan(1.0);
    const double urk = radint_ctx(&R->C, R->p.eta, ur2, w, w + 2 * R->C.M, w + 2 * R->C.M + N);
    const double upk = radint_ctx(&R->C, R->p.eta, up2, w, w + 2 * R->C.M, w + 2 * R->C.M + N);
    cplx* dT = w + 2 * R->C.M + 2 * N;
    cheb_d1(&R->C, R->TFC, dT, w);
    const double eta = R->p.eta;
    out[0] = R->tot_time;
    out[1] = 2.0 * pi * (urk + upk);
    out[2] = dT[N - 1].re / ((1.0 - eta) / (log(eta)));
    out[3] = dT[0].re / ((1.0 - eta) / (eta * log(eta)));
    out[4] = 2.0 * pi * urk;
    out[5] = 2.0 * pi * upk;
    free(w);
    free(ur2);
    return 0;
}

/* writeKE_spectral's further outputs (output.F90:122-324) for the current
 * state: out4 = {2 pi radInt(upavg^2) (the KE file's sixth column), the
 * viscous dissipation and the buoyancy power (power_ columns 2 and 3),
 * tot_time}; upbal (3 N, may be NULL): per level the upbalance_ columns 2-4,
 * Re((u_phi,0 - upFR_prev)/dt_new), Re(uFa_eqn), Re(u_diff).  upavg and
 * <u_r T> are azimuthal means of the physical fields (invfft, Np points);
 * T is the temperature of the last nonlinear pass, as the module array tt the
 * reference multiplies is. */
int orc_budget(orc_run* R, double* out4, double* upbal)
{
    const int N = R->N, Nm1 = R->Nm1, Np = R->Np;
    const double pi = 4.0 * atan(1.0);
    const double* rr = R->G.r_radius;
    const double* rr2 = R->G.r_radius2;
    double* lev = (double*)calloc((size_t)5 * N + 3 * (size_t)Np, sizeof(double));
    cplx* w = (cplx*)malloc(sizeof(cplx) * (size_t)(2 * R->C.M + 8 * N + 2 * Np));
    cplx* sp = calloc_c((size_t)Nm1 * N);
    if (!lev || !w || !sp) { free(lev); free(w); free(sp); return -4; }
    double *upbar2 = lev, *urt = lev + N, *w2 = lev + 2 * N, *up2 = lev + 3 * N;
    double *xu = lev + 5 * N, *xr = xu + Np, *xt = xr + Np;
    cplx* col = w + 2 * R->C.M;                 /* one mode's or level's spectral values */
    cplx* fw = col + 2 * N + Np;                /* four_inv scratch: 2 Np */
    for (int j = 0; j < N; j++) {
        cplx* f = col;
        for (int m = 0; m < Nm1; m++) f[m] = R->upFR[m + (size_t)Nm1 * j];
        four_inv(&R->F, f, xu, fw);
        for (int m = 0; m < Nm1; m++) f[m] = R->urFR[m + (size_t)Nm1 * j];
        four_inv(&R->F, f, xr, fw);
        for (int m = 0; m < Nm1; m++) f[m] = R->tt[m + (size_t)Nm1 * j];
        four_inv(&R->F, f, xt, fw);
        double ua = 0.0, ut = 0.0;
        for (int q = 0; q < Np; q++) {
            ua = ua + xu[q];
            ut = ut + xr[q] * xt[q];
        }
        ua = ua * (1.0 / Np);
        ut = ut * (1.0 / Np);
        upbar2[j] = ua * ua;
        urt[j] = ut;
        double a = 0.0, b = 0.0;
        for (int m = 0; m < Nm1; m++) {
            const cplx o = R->omgFR[m + (size_t)Nm1 * j], u = R->upFR[m + (size_t)Nm1 * j];
            if (m == 0) a = a + cmul(o, o).re;
            else a = a + cadd(cmul(o, cconj(o)), cmul(cconj(o), o)).re;
            const cplx u2 = cmul(u, u);
            b = b + hypot(u2.re, u2.im);
        }
        w2[j] = a;
        up2[j] = b;
    }
    cplx* ra = col + N;
    const double upbark = radint_ctx(&R->C, R->p.eta, upbar2, w, col, ra);
    const double vis1 = radint_ctx(&R->C, R->p.eta, w2, w, col, ra);
    const double buo = radint_ctx(&R->C, R->p.eta, urt, w, col, ra);
    out4[0] = 2.0 * pi * upbark;
    if (R->p.mBC == 1) {
        const double ro = 4.0 * pi * 2.0 * up2[N - 1], ri = 4.0 * pi * 2.0 * up2[0];
        out4[1] = 2.0 * pi * vis1 - (ro - ri);
    } else {
        out4[1] = 2.0 * pi * vis1;
    }
    out4[2] = 2.0 * pi * R->p.Ra / R->p.Pr * buo;
    out4[3] = R->tot_time;
    if (upbal) {
        /* chebinvtranD1 of upFC(m), and D2 of upFC(1,:) */
        wsp W;
        wsp_alloc(&W, N, R->C.M);
        for (int m = 0; m < Nm1; m++) {
            get_rad(R, R->upFC, m, W.a);
            cheb_d1(&R->C, W.a, W.b, W.w);
            put_rad(R, sp, m, W.b);
        }
        get_rad(R, R->upFC, 0, W.a);
        cheb_d2(&R->C, W.a, W.c, W.w);
        for (int j = 0; j < N; j++) {
            cplx su = cx(0.0, 0.0), sd = cx(0.0, 0.0);
            for (int m = 0; m < Nm1; m++) {
                const size_t o = m + (size_t)Nm1 * j;
                const cplx ur = R->urFR[o], up = R->upFR[o], du = sp[o];
                if (m == 0) {
                    su = cadd(su, cmul(ur, up));
                    sd = cadd(sd, cmul(ur, du));
                } else {
                    su = cadd(su, cadd(cmul(ur, cconj(up)), cmul(cconj(ur), up)));
                    sd = cadd(sd, cadd(cmul(ur, cconj(du)), cmul(cconj(ur), du)));
                }
            }
            const cplx u0 = R->upFR[(size_t)Nm1 * j];
            const cplx ufa = cadd(sd, cscale(rr[j], su));
            const cplx ud = cadd(csub(W.c[j], cscale(rr2[j], u0)), cscale(rr[j], sp[(size_t)Nm1 * j]));
            const cplx du = csub(u0, R->up_prev[j]);
            upbal[3 * j] = du.re / R->dt_new;
            upbal[3 * j + 1] = -ufa.re;
            upbal[3 * j + 2] = ud.re;
        }
        wsp_free(&W);
    }
    free(sp);
    free(w);
    free(lev);
    return 0;
}

/* calculate_spectra (output.F90:327-355): per azimuthal mode m,
 * KEspec(m) = 4 pi radInt |ur_m^2 + up_m^2| (the modulus of the sum of the
 * complex squares, as written) and Tspec(m) = 4 pi radInt |T_m|. */
int orc_spectra(orc_run* R, double* KEspec, double* Tspec)
{
    const int N = R->N, Nm1 = R->Nm1;
    const double pi = 4.0 * atan(1.0);
    double* e = (double*)malloc(sizeof(double) * 2 * (size_t)N);
    cplx* w = (cplx*)malloc(sizeof(cplx) * (size_t)(2 * R->C.M + 2 * N));
    if (!e || !w) { free(e); free(w); return -4; }
    double* t = e + N;
    for (int m = 0; m < Nm1; m++) {
        for (int r = 0; r < N; r++) {
            const size_t o = m + (size_t)Nm1 * r;
            const cplx k = cadd(cmul(R->urFR[o], R->urFR[o]), cmul(R->upFR[o], R->upFR[o]));
            e[r] = hypot(k.re, k.im);
            t[r] = hypot(R->tFR[o].re, R->tFR[o].im);
        }
        KEspec[m] = 4.0 * pi * radint_ctx(&R->C, R->p.eta, e, w, w + 2 * R->C.M, w + 2 * R->C.M + N);
        Tspec[m] = 4.0 * pi * radint_ctx(&R->C, R->p.eta, t, w, w + 2 * R->C.M, w + 2 * R->C.M + N);
    }
    free(w);
    free(e);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Restart with a change of resolution (init.F90:722-861)                     */
/* ------------------------------------------------------------------------ */

/* FFTW's REDFT00 of n real values (chebtransform_redft / chebinvtran_redft,
 * chebyshev_fftw.F90:196-245, without their scaling): y_k = x_0 +
 * (-1)^k x_{n-1} + 2 sum_{j=1}^{n-2} x_j cos(pi j k / (n-1)), summed directly. */
static void redft00(int n, const double* x, double* y)
{
    const double pi = 4.0 * atan(1.0);
    const long L = 2 * (long)(n - 1);
    for (int k = 0; k < n; k++) {
        double s = x[0] + ((k % 2) ? -x[n - 1] : x[n - 1]);
        for (int j = 1; j < n - 1; j++) s = s + 2.0 * x[j] * cos(pi * (double)(((long)j * k) % L) / (double)(n - 1));
        y[k] = s;
    }
}

/* The state of a checkpoint of Nm_o x Nr_o (in[0..3]: tFR, omgFR, urFR,
 * upFR) on Nm_n x Nr_n (out[0..3]): per mode, REDFT00 on the old grid, the
 * transform padded with zeros to Nr_n, REDFT00 on the new grid divided by
 * 2 (Nr_n - 1), times fac = 2 (Nr_n - 1) / (2 (Nr_o - 1)); modes above Nm_o
 * zero.  As written there, the highest old Chebyshev mode enters twice its
 * weight (the halving of the last DCT-I term is not undone), and coefficients
 * the reference leaves unassigned (modes above Nm_o below Nr_o when both
 * grow) are zero here.  Equal sizes copy; a smaller grid is not handled by
 * the reference (returns -1). */
int orc_refine(int Nm_o, int Nr_o, int Nm_n, int Nr_n, const double* const in[4], double* const out[4])
{
    if (Nm_n < Nm_o || Nr_n < Nr_o || Nr_o < 2) return -1;
    const size_t So = (size_t)(Nm_o + 1) * Nr_o, Sn = (size_t)(Nm_n + 1) * Nr_n;
    if (Nm_n == Nm_o && Nr_n == Nr_o) {
        for (int f = 0; f < 4; f++) memcpy(out[f], in[f], sizeof(cplx) * So);
        return 0;
    }
    const double fac = (2.0 * (double)(Nr_n - 1)) / (2.0 * (double)(Nr_o - 1));
    double* b = (double*)calloc((size_t)4 * Nr_n, sizeof(double));
    if (!b) return -4;
    double *xo = b, *co = b + Nr_n, *cn = b + 2 * Nr_n, *yn = b + 3 * Nr_n;
    for (int f = 0; f < 4; f++) {
        const cplx* a = (const cplx*)in[f];
        cplx* z = (cplx*)out[f];
        memset(z, 0, sizeof(cplx) * Sn);
        for (int m = 0; m <= Nm_o; m++)
            for (int part = 0; part < 2; part++) {
                for (int r = 0; r < Nr_o; r++) {
                    const cplx v = a[m + (size_t)(Nm_o + 1) * r];
                    xo[r] = part ? v.im : v.re;
                }
                redft00(Nr_o, xo, co);
                for (int r = 0; r < Nr_n; r++) cn[r] = r < Nr_o ? co[r] : 0.0;
                redft00(Nr_n, cn, yn);
                for (int r = 0; r < Nr_n; r++) {
                    const double v = (yn[r] / (2.0 * (double)(Nr_n - 1))) * fac;
                    cplx* o = z + m + (size_t)(Nm_n + 1) * r;
                    if (part) o->im = v; else o->re = v;
                }
            }
    }
    free(b);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Seeded initial condition (build-defined; the reference has no RNG)         */
/* ------------------------------------------------------------------------ */
static uint64_t splitmix64(uint64_t* s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* T_0(r) = conductive profile (init.F90:207-211); for m = 1..Nm:
 * T_m(r) = amp*(a_m + i b_m)*sin(pi (r - rmin)) at interior points, 0 at the
 * walls; a_m, b_m uniform in [-1, 1) from SplitMix64 seeded by (seed, m).
 * Velocity and vorticity are zero. */
void orc_seed_state(int Nr, int Nm, double eta, uint64_t seed, double amp,
                    double* tFR, double* omgFR, double* urFR, double* upFR)
{
    const int Nm1 = Nm + 1;
    const double rmin = eta / (1.0 - eta), rmax = 1.0 / (1.0 - eta);
    double* radius = (double*)malloc(sizeof(double) * 3 * (size_t)Nr);
    orc_grid(Nr, eta, radius, radius + Nr, radius + 2 * Nr);
    const size_t S = (size_t)Nm1 * Nr;
    memset(tFR, 0, sizeof(double) * 2 * S);
    memset(omgFR, 0, sizeof(double) * 2 * S);
    memset(urFR, 0, sizeof(double) * 2 * S);
    memset(upFR, 0, sizeof(double) * 2 * S);
    const double pi = 4.0 * atan(1.0);
    for (int r = 0; r < Nr; r++) {
        const double tc = log(radius[r]) / log(rmin / rmax) - log(rmax) / log(rmin / rmax);
        tFR[2 * ((size_t)Nm1 * r)] = tc;
    }
    for (int m = 1; m <= Nm; m++) {
        uint64_t st = seed * 0xD1B54A32D192ED03ULL + (uint64_t)m;
        (void)splitmix64(&st);
        const double a = (double)(splitmix64(&st) >> 11) * 0x1.0p-53 * 2.0 - 1.0;
        const double b = (double)(splitmix64(&st) >> 11) * 0x1.0p-53 * 2.0 - 1.0;
        for (int r = 1; r < Nr - 1; r++) {
            const double env = sin(pi * (radius[r] - rmin));
            tFR[2 * ((size_t)m + (size_t)Nm1 * r)] = amp * a * env;
            tFR[2 * ((size_t)m + (size_t)Nm1 * r) + 1] = amp * b * env;
        }
    }
    free(radius);
}

/* ------------------------------------------------------------------------ */
/* init_perturbation (init.F90:277-399) and its Gaspari pieces (:937-985)     */
/* ------------------------------------------------------------------------ */

/* gausslike_compact_middle (init.F90:937-953, Gaspari & Cohn 1999 eq. 4.7b) */
static double gasp_middle(double r, double c, double L)
{
    const double pi = 4.0 * atan(1.0);
    return pi / 3.0 * L * r * (r + 3.0 * L) * exp(-r / L) +
           (2.0 * pi * L * L * ((c + L) * (c + L))) / r * exp(-(2.0 * c) / L) * (1.0 - (1.0 - r / (c + L)) * exp(r / L)) +
           pi * L * L * L * (exp(-r / L) - exp((r - 2.0 * c) / L));
}

/* gausslike_compact_edge (init.F90:955-970, eq. 4.7a) */
static double gasp_edge(double r, double c, double L)
{
    const double pi = 4.0 * atan(1.0);
    return 2.0 * pi * L / r * exp(-r / L) *
           (0.5 * (r * (r + L) * (2.0 * c - r)) + 1.0 / 3.0 * ((r - c) * (r - c) * (r - c) - c * c * c) -
            L * (c + L) * (r - c + L) + L * ((c + L) * (c + L)) * exp((r - 2.0 * c) / L));
}

/* gausslike_compact_center (init.F90:972-985, eq. 4.7c) */
static double gasp_center(double c, double L)
{
    const double pi = 4.0 * atan(1.0);
    return pi * L * L * L * (1.0 - exp(-(2.0 * c) / L)) - 2.0 * pi * c * L * (c + L) * exp(-(2.0 * c) / L);
}

double orc_gaspari(int piece, double r, double c, double L)
{
    if (piece == 0) return gasp_middle(r, c, L);
    if (piece == 1) return gasp_edge(r, c, L);
    return gasp_center(c, L);
}

/* The physical temperature tt(Np, Nr), column-major, of init_perturbation on
 * top of the conductive profile (init.F90:207-211, 277-390):
 *   n_init = 0: Gaspari pieces in r about the grid point nearest the mid
 *     radius (sigma_r = 0.1/sqrt 2, L = sigma_r/2), the inner branch (built
 *     with c2) divided by the centre value of c1 and the outer one by that of
 *     c2, times exp(-(phi - pi)^2/sigma_p^2) with sigma_p the default-real
 *     0.2 and phi = (i-1) 2 pi/Np;
 *   n_init < 0: for 1 < phi < 3 a Gaussian at (rmax - 1/2, 2) with sigma
 *     0.25, times -(rmax - r)(rmin - r) ampT/(2 pi sigma_p sigma_r);
 *   n_init > 0: ampT (rmax - r)(rmin - r) cos(n_init phi);
 * phi of the last two is init_grid's accumulated grid (init.F90:141-147). */
void orc_perturbation_field(int N, int Nm, double eta, int n_init, double ampT, double* tt)
{
    const int Np = 3 * Nm;
    const double pi = 4.0 * atan(1.0);
    const double rmin = eta / (1.0 - eta), rmax = 1.0 / (1.0 - eta);
    double* rad = (double*)malloc(sizeof(double) * (3 * (size_t)N + (size_t)Np + (size_t)N));
    double* phi = rad + 3 * (size_t)N;
    double* g = phi + Np;
    orc_grid(N, eta, rad, rad + N, rad + 2 * N);
    phi[0] = 0.0;
    for (int i = 1; i < Np; i++) phi[i] = phi[i - 1] + 2.0 * pi / (double)Np;
#define TCON(j) (log(rad[j]) / log(rmin / rmax) - log(rmax) / log(rmin / rmax))
    if (n_init == 0) {
        const double sigma_r = 0.1 / sqrt(2.0);
        const double sigma_p = (double)0.2f;
        double rc = 0.5 * (rmax + rmin);
        int ir = 0;
        for (int j = 1; j < N; j++)
            if (fabs(rad[j] - rc) < fabs(rad[ir] - rc)) ir = j;
        rc = rad[ir];
        const double c1 = 0.5 * (rad[N - 1] - rc), c2 = 0.5 * (rc - rad[0]), L = 0.5 * sigma_r;
        for (int j = 0; j < N; j++) {
            g[j] = 0.0;
            if (j == ir) g[j] = 1.0;
            else if (rad[j] < rc - c2) g[j] = gasp_edge(rc - rad[j], c2, L);
            else if (rad[j] >= rc - c2 && rad[j] < rc) g[j] = gasp_middle(rc - rad[j], c2, L);
            else if (rad[j] > rc && rad[j] <= rc + c1) g[j] = gasp_middle(rad[j] - rc, c1, L);
            else if (rad[j] > rc + c1) g[j] = gasp_edge(rad[j] - rc, c1, L);
        }
        for (int j = 0; j < N; j++) {
            if (j < ir) g[j] = g[j] / gasp_center(c1, L);
            else if (j > ir) g[j] = g[j] / gasp_center(c2, L);
        }
        for (int j = 0; j < N; j++) {
            const double c_r = ampT * g[j];
            for (int i = 0; i < Np; i++) {
                const double ph = (double)i * 2.0 * pi / (double)Np;
                tt[i + (size_t)Np * j] = c_r * exp(-((ph - pi) * (ph - pi)) / (sigma_p * sigma_p)) + TCON(j);
            }
        }
    } else if (n_init < 0) {
        const double sigma_r = 0.25, sigma_p = 0.25;
        const double rm = rmax - (rmax - rmin) / 2.0;
        for (int j = 0; j < N; j++)
            for (int i = 0; i < Np; i++) {
                double v = TCON(j);
                if (phi[i] > 1.0 && phi[i] < 3.0)
                    v = -1.0 * (rmax - rad[j]) * (rmin - rad[j]) * ampT * (1.0 / (2.0 * pi * sigma_p * sigma_r)) *
                            exp(-1.0 * ((rad[j] - rm) * (rad[j] - rm) / (sigma_r * sigma_r) +
                                        (phi[i] - 2.0) * (phi[i] - 2.0) / (sigma_p * sigma_p))) +
                        TCON(j);
                tt[i + (size_t)Np * j] = v;
            }
    } else {
        for (int j = 0; j < N; j++)
            for (int i = 0; i < Np; i++)
                tt[i + (size_t)Np * j] = ampT * (rmax - rad[j]) * (rmin - rad[j]) * cos((double)n_init * phi[i]) + TCON(j);
    }
#undef TCON
    free(rad);
}

/* init_perturbation's spectral state: forfft of every level of the physical
 * temperature (init.F90:355, 373-375, 386-388); u = omega = 0. */
void orc_init_perturbation(int N, int Nm, double eta, int n_init, double ampT, double* tFR, double* omgFR,
                           double* urFR, double* upFR)
{
    const int Np = 3 * Nm, Nm1 = Nm + 1;
    const size_t S = (size_t)Nm1 * N;
    double* tt = (double*)malloc(sizeof(double) * (size_t)Np * N);
    orc_perturbation_field(N, Nm, eta, n_init, ampT, tt);
    for (int j = 0; j < N; j++) orc_forfft(Nm, Np, tt + (size_t)Np * j, tFR + 2 * (size_t)Nm1 * j);
    memset(omgFR, 0, sizeof(double) * 2 * S);
    memset(urFR, 0, sizeof(double) * 2 * S);
    memset(upFR, 0, sizeof(double) * 2 * S);
    free(tt);
}

/* dt_array as get_checkpoint_data restores it (init.F90:464-612), n entries,
 * newest first; after orc_restart.  Only the IMEX variable-dt path reads it
 * (its first step keeps entries 2..n, steptime.F90:344-348); the IMEXRK and
 * RK loops reset it at their first step. */
int orc_restart_dt(orc_run* R, int n, const double* dta)
{
    if (n < 1 || n > 8) return -1;
    for (int i = 0; i < n; i++) R->dt_array[i] = dta[i];
    return 0;
}

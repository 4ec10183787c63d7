// fdns_internal.h -- device-side descriptors shared by the kernels and the
// host orchestration of the MI355X time-stepping core.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fdns {

constexpr int NT = 256;        // threads per workgroup of every kernel
constexpr int MAXS = 8;        // max IMEXRK stages
constexpr int MAXH = 4;        // max IMEX history depth (BDF4/AB4)
constexpr int LAGMAX = 64;     // max Lagrange points of the wall vorticity (lagpts)
constexpr int MAXFAC = 20;     // max radix factors of one DFT length
constexpr int MAX_N = 513;     // largest Nr_max the radial kernels take
constexpr int MAX_NP = 4608;   // largest Np_max = 3*Nm_max of the nonlinear pass

enum SchemeKind : int { SK_IMEX = 0, SK_IMEXRK = 1, SK_RK = 2 };

#define GLOBAL_AS __attribute__((address_space(1)))
#define LDS_AS __attribute__((address_space(3)))

// Device loads and stores through the global address space.  Pointers read
// out of a run descriptor are otherwise generic ("flat") to the compiler,
// which then orders every such access after all outstanding LDS traffic.
__device__ __forceinline__ double gload(const double* p) { return *(const GLOBAL_AS double*)p; }
__device__ __forceinline__ double2 gload(const double2* p)
{
    const GLOBAL_AS double* q = (const GLOBAL_AS double*)p;
    double2 v;
    v.x = q[0];
    v.y = q[1];
    return v;
}
__device__ __forceinline__ int gload(const int* p) { return *(const GLOBAL_AS int*)p; }
__device__ __forceinline__ void gstore(double* p, double v) { *(GLOBAL_AS double*)p = v; }
__device__ __forceinline__ void gstore(double2* p, double2 v)
{
    GLOBAL_AS double* q = (GLOBAL_AS double*)p;
    q[0] = v.x;
    q[1] = v.y;
}

// The same for LDS (the FFT work buffers, always in LDS).
__device__ __forceinline__ double2 lload(const double2* p)
{
    const LDS_AS double* q = (const LDS_AS double*)p;
    double2 v;
    v.x = q[0];
    v.y = q[1];
    return v;
}
__device__ __forceinline__ void lstore(double2* p, double2 v)
{
    LDS_AS double* q = (LDS_AS double*)p;
    q[0] = v.x;
    q[1] = v.y;
}

template <class T> struct GRef {
    T* p;
    __device__ __forceinline__ operator T() const { return gload(p); }
    __device__ __forceinline__ GRef& operator=(const T& v) { gstore(p, v); return *this; }
    __device__ __forceinline__ GRef& operator=(const GRef& o) { gstore(p, gload(o.p)); return *this; }
};

// A device array in global memory inside a descriptor: indexes to GRef on the
// device, converts to a plain pointer on the host.
template <class T> struct GPtr {
    T* p;
    __host__ __device__ GPtr& operator=(T* q) { p = q; return *this; }
    __host__ __device__ operator T*() const { return p; }
    __device__ __forceinline__ GRef<T> operator[](size_t i) const { return GRef<T>{p + i}; }
};
enum ImpKind : int { IMP_CN = 0, IMP_BDF2 = 1, IMP_BDF3 = 2, IMP_BDF4 = 3 };

// One DFT length: radix sequence (4 bits per pass, first pass lowest) and
// table tw[t] = exp(-2 pi i t / n).  Radices 2, 3, 4, 5 have their own
// butterflies; a code of 0 is the last pass, a direct DFT of radix rg (the
// cofactor of n without 2, 3, 5, e.g. 31 for the length 62 of Nr_max = 32).
struct FftPlan {
    int n;
    int nf;
    unsigned long long code;
    const double2* tw;
    int rg;
};

// Everything fixed by (Nr_max, Nm_max); passed by value to kernels.
struct Shape {
    int N;      // Nr_max
    int Nm;     // Nm_max
    int Nm1;    // Nm_max + 1
    int Np;     // 3 * Nm_max
    int M;      // 2 * Nr_max - 2
    double fact;        // sqrt(2/(N-1))
    double eta;
    FftPlan pm;         // length M (radial DCT-I through the even extension)
    FftPlan pp;         // length Np (azimuthal real transforms, two per DFT)
    FftPlan pq;         // length Np for the nonlinear pass: the factors 3 last; the
                        // twiddles are pp's
    const double* rr;   // 1/r  (r_radius)
    const double* rr2;  // 1/r^2
    const double* rad;  // r
    // the radial transforms as N x N matrices, column-major ([row + N col]):
    const double* E0;   // chebinvtran    coefficients -> values
    const double* E1;   // chebinvtranD1  coefficients -> d/dr values
    const double* E2;   // chebinvtranD2  coefficients -> d2/dr2 values
    const double* Fw;   // chebtransform  values -> coefficients
    const double* G;    // E1 Fw          values -> d/dr values
    // launch choices of this shape (set by the tuner; results do not depend
    // on them): threads per nonlinear-pass workgroup (128 | 256); per
    // column-tile kernel (imex_pre, imex_mid, imex_post, rk_pre, rk_mid,
    // rk_post) the wavefronts per workgroup (4 | 8), whether its right-hand
    // blocks are staged in LDS and whether its epilogue inputs are
    // prefetched into LDS; runs per batched solve chunk
    int nl_tpb;
    int tile_waves[6];
    int tile_stage[6];
    int tile_pf[6];
    int solve_chunk;
    int lane_split;   // an ensemble splits its longest groups over free stream lanes
};

// Per-run descriptor, resident in device memory.
struct RunDesc {
    // state, (Nm+1, N) column-major complex
    GPtr<double2> tFR, omgFR, upFR, urFR, psi;
    GPtr<double2> tspec, wspec;   // IMEX: temp_spec / omg_spec (Chebyshev coefficients);
                                  // the m = 0 column of wspec holds upFC(1,:), the
                                  // coefficients of the mean flow
                                  // IMEXRK: the stage start value Y0 (physical in r)
    GPtr<double2> ubar;           // IMEXRK: uphi_bar_spec (N)
    GPtr<double2> TFC;            // N
    GPtr<double2> upT, urT, upW, urW;   // nonlinear products (Nm+1, N)
    GPtr<double> urw;                   // <ur omega>(r), N
    GPtr<double> cfl;                   // variable dt: max |u_r|, max |u_phi| of the
                                        // last nonlinear pass per radial level (2N)
    // IMEX histories: ring of L slots, each (Nm+1, N); mean flow: L slots of N
    GPtr<double2> hiT, heT, hiW, heW;
    GPtr<double2> hiU, heU;
    // IMEXRK stage terms: S slots, each (Nm+1, N); the RK path keeps each
    // stage's whole right-hand side in FdT / FdW
    GPtr<double2> FaT, FdT, FaW, FdW, Fb;
    // RK path: upFC (Chebyshev coefficients of u_phi, (Nm+1, N)) and
    // chebinvtranD1 of them (d u_phi / dr)
    GPtr<double2> upC, d1u;
    // runs that record (n_KE > 0): the temperature of the last nonlinear pass
    // ((Nm+1, N)) and u_phi,0 at upFR_prev (N), inputs of the power budget
    // and the mean-flow balance
    GPtr<double2> tpass, up0prev;
    int keep;
    // solve areas, (Nm+1) x 2N complex each, mode-major (rows 0..2N-1 of mode
    // m at +2N m): sx the right-hand sides of a step's implicit solves, sy
    // their solutions (temperature or psi in rows 0..N-1, omega or mean flow
    // in rows N..2N-1)
    GPtr<double2> sx, sy;
    // resident implicit operators (shared sets), column-major.  Inverse mode:
    // Ti + m*N*N = AT^{-1} of mode m >= 1 and, at Ti, the packed LU factor of
    // AT of m = 0 (row permutation pT); Fi + (m-1)*2N*(N-2) = columns
    // N+1..2N-2 of AF^{-1} of mode m >= 1; Ui = A_uphi^{-1}.  LU mode: packed
    // LU factors with net row permutations, Ti + m*N*N / pT + m*N (AT),
    // Fi + (m-1)*4N*N / pF + (m-1)*2N (AF), Ui / pU (A_uphi).
    const double *Ti, *Fi, *Ui;
    const int *pT, *pF, *pU;
    // LU mode, per IMEXRK stage (index s; IMEX uses 0): the factors of the
    // stage's operators -- the same set for every stage unless the scheme
    // solves each stage with its own diagonal (diag_diff)
    const double *TiS[MAXS], *FiS[MAXS], *UiS[MAXS];
    const int *pTS[MAXS], *pFS[MAXS], *pUS[MAXS];
    int solve_lu;
    // physics and scheme
    double Ra, Pr, dt;
    int kind, imp, n_imp, n_exp, L;  // L = history ring length
    int head;                         // slot written by the previous IMEX step
    int hist_ready;                   // IMEX restart: histories already filled
                                      // from checkpoint states (no constant fill)
    int buo_imp;
    int vardt;                        // record the velocity maxima (l_vartimestep)
    double wl, wi[MAXH], we[MAXH];   // IMEX weights (rhs_update_wts_imp/exp)
    int S;
    double aD[MAXS * MAXS], aA[MAXS * MAXS];
    double bD[MAXS], bA[MAXS];
    int sa;                           // stiffly accurate: the last stage is the new state;
                                      // otherwise Assembly_stage (k_rk_asm)
    // Assembly_stage: LAPpsi LU factors (Li + m*N*N, net row permutation
    // pL + m*N), the Lagrange derivative weights lagw = [dw_rmin, d2w_rmin,
    // dw_rmax, d2w_rmax] (LAGMAX each, lm used) and, at 4*LAGMAX, the
    // stress-free wall factors: -2/r_min, -2/r_max (IMEXRK) or -2 (1/r_min),
    // -2 (1/r_max) (RK), as the two paths round them
    const double* Li;
    const int* pL;
    GPtr<const double> lagw;
    int lm, mbc;
    int nl_col[MAXS];                 // nonlinear pass after stage k
    int fa_col[MAXS];                 // stage-k explicit terms are read later
    int need_fd[MAXS];                // stage-k diffusion term is read later
    // counters and diagnostics
    int kstep;
    double time;
    GPtr<double> series;              // rows {t, KE, Nu_t, Nu_b}
    int nrec, cap;
};

// Batched solves of one group of runs: the runs that share an operator set
// form chunks of at most SOLVE_RHS right-hand sides; one workgroup applies the
// set's inverse of one mode to a whole chunk.
constexpr int SOLVE_RHS = 32;
struct SolvePlan {
    const int2* chT;   // temperature sets: chunks (first entry of idxT, count)
    const int* idxT;   // group run indices, ordered by temperature set
    int nchT;
    const int2* chF;   // vorticity / mean-flow sets
    const int* idxF;
    int nchF;
    int wT, wF;        // real columns of the widest chunk, 16 per 8 runs
};

// ---- kernel launchers (fdns_kernels.hip) ----
void launch_nonlin(const RunDesc* runs, int nrun, const Shape& sh, hipStream_t st);
void launch_forfft(const double* x, double2* f, const Shape& sh, hipStream_t st);
void launch_copy(const double* a, double* b, size_t n16, hipStream_t st);
void launch_imex_prep(const RunDesc* runs, int nrun, const Shape& sh, hipStream_t st);
void launch_imex_pre(const RunDesc* runs, int nrun, const Shape& sh, hipStream_t st);
void launch_imex_mid(const RunDesc* runs, int nrun, const Shape& sh, hipStream_t st);
void launch_imex_post(const RunDesc* runs, int nrun, const Shape& sh, hipStream_t st);
// ki, ke: 0-based history slot (counted back from the first step's slot) of
// the implicit and explicit arrays to fill from the current state, or -1 for
// every slot of the runs whose histories are not already filled
void launch_imex_restart(const RunDesc* runs, int nrun, const Shape& sh, hipStream_t st, int ki = -1, int ke = -1);
void launch_rk_pre(const RunDesc* runs, int nrun, const Shape& sh, int s, hipStream_t st);
void launch_rk_mid(const RunDesc* runs, int nrun, const Shape& sh, int s, hipStream_t st);
void launch_rk_post(const RunDesc* runs, int nrun, const Shape& sh, int s, hipStream_t st);
void launch_rk_asm(const RunDesc* runs, int nrun, const Shape& sh, hipStream_t st);
void launch_asm_LAP(double* A, const double* t, const double* D, const double* D2, const Shape& sh, int m0, int count,
                    hipStream_t st);
void launch_rk_init(const RunDesc* runs, int nrun, const Shape& sh, hipStream_t st);
void launch_rkx_init(const RunDesc* runs, int nrun, const Shape& sh, hipStream_t st);
void launch_rkx_update(const RunDesc* runs, int nrun, const Shape& sh, int s, hipStream_t st);
void launch_rkx_rhs(const RunDesc* runs, int nrun, const Shape& sh, int s, hipStream_t st);
void launch_keep(const RunDesc* runs, int nrun, const Shape& sh, int with_up, hipStream_t st);
// out: per run 4 + 3N doubles (k_budget); uses the runs' sy as scratch
void launch_budget(const RunDesc* runs, int nrun, const Shape& sh, double* out, hipStream_t st);
// the current device's address of the budget scratch array and its capacity
// in doubles; -1 if the address cannot be resolved
int budget_scratch(double** out, size_t* cap);
// out: per run 2 (Nm+1) doubles, KEspec then Tspec
void launch_spectra(const RunDesc* runs, int nrun, const Shape& sh, double* out, hipStream_t st);
void launch_solve(const RunDesc* runs, int nrun, const SolvePlan& pl, const Shape& sh, int phase, int lu, int stage,
                  hipStream_t st, int m0n = 0);
// inverse mode, Nr_max small enough: the temperature solves and, in the same
// launch, the m = 0 temperature solves of every run (k_m0's work); false
// (nothing launched) above that size
bool launch_solve_T_m0(const RunDesc* runs, int nrun, const SolvePlan& pl, const Shape& sh, int stage, hipStream_t st);
void launch_m0(const RunDesc* runs, int nrun, const Shape& sh, hipStream_t st);
size_t solve_lds(const Shape& sh, int w);
void launch_asm_T(double* A, const double* t, const double* D, const double* D2, const Shape& sh, int m0, int count,
                  double wtdt, double Pr, hipStream_t st);
void launch_asm_F(double* A, const double* t, const double* D, const double* D2, const Shape& sh, int m0, int count,
                  double wtdt, int mBC, int irk, hipStream_t st);
void launch_asm_U(double* A, const double* t, const double* D, const double* D2, const Shape& sh, double wtdt, int mBC,
                  hipStream_t st);
void launch_lu(double* A, int* perm, int n, size_t stride, int nmat, hipStream_t st);
void launch_inv(const double* LU, const int* perm, int n, size_t stride, int nmat, int col0, int ncols, double* out,
                size_t ostride, hipStream_t st);
void launch_diag(const RunDesc* runs, int nrun, const Shape& sh, int record, double* out6, hipStream_t st);
void launch_advance(RunDesc* runs, int nrun, hipStream_t st);
int  kernels_init();   // raise dynamic LDS limits once

}  // namespace fdns

// fdns_kernels.hip -- gfx950 kernels of the FAlCon-DNS time-stepping core.
//
// Work decomposition (one run or a batch of runs of one (Nr, Nm) shape):
//   k_nonlin    one workgroup per (run, radial level): the four inverse real
//               transforms, the four products and the four forward transforms
//               of Nr_maxLOOP (nonlin.F90:134-190), all in LDS.
//   k_imex_pre / k_imex_mid / k_imex_post
//               one wavefront per (run, Fourier mode): the IMEX mode update of
//               Nm_maxLOOP (fourierloop.F90:108-230 with rhs_create.F90:24-385)
//               split around the two solve phases: Chebyshev transforms, RHS
//               assembly into the run's solve area, the new fields.
//   k_rk_pre / k_rk_mid / k_rk_post
//               the same for one IMEXRK stage (Get_stage_var,
//               fourierloop_imexrk.F90:169-393, with the stage terms of
//               RHS_construct_stage, :429-488); k_rk_init the stage-0 terms.
//   k_solve     one solve phase (temperature, or vorticity-streamfunction and
//               mean flow) of a group of runs: FP64 MFMA products of the
//               resident operator inverses with the right-hand sides of every
//               run sharing the operator set, or dgetrs per (run, mode).
//   k_asm_*     operator matrices of mat_build / mat_build_IRK /
//               mat_build_uphibar (mat_assembly.F90:68-413), one workgroup per
//               (mode, column), a chunk of modes per launch.
//   k_lu        one workgroup per matrix: blocked right-looking LU with
//               partial pivoting whose per-element arithmetic is dgetf2's.
//   k_inv       the needed columns of each operator's inverse from its LU
//               factor (dgetrs on unit vectors), a wavefront per block of
//               columns.
//   k_diag      KE and Nusselt (output.F90:232-265), one workgroup per run.
//
// All DFTs are LDS-resident Stockham transforms with radices 2, 3, 4, 5.
#include "fdns_internal.h"

#include <float.h>

#include <algorithm>
#include <mutex>
#include <type_traits>

namespace fdns {

// ------------------------------------------------------------------ complex
__device__ __forceinline__ double2 mk(double x, double y) { double2 r; r.x = x; r.y = y; return r; }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return mk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return mk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ double2 csc(double s, double2 a) { return mk(s * a.x, s * a.y); }
// i*m*(1/r)*z written as the reference does: (ii*m*r_radius)*z (rhs_create.F90:99)
__device__ __forceinline__ double2 cimr(double m, double r, double2 z) { return cmul(mk(0.0, m * r), z); }
__device__ __forceinline__ double2 cneg(double2 a) { return mk(-a.x, -a.y); }
// Component-wise select: a ternary on double2 makes the compiler pick between
// stack addresses, which puts both operands in scratch memory.
__device__ __forceinline__ double2 csel(bool c, double2 a, double2 b) { return mk(c ? a.x : b.x, c ? a.y : b.y); }

// Map a workgroup to a work item so that consecutive work items (neighbouring
// modes or radial levels of one run) land on the same XCD and share its L2.
// gridDim.x is padded to a multiple of 8 by the launcher.
__device__ __forceinline__ bool map_work(int total, int& w)
{
    const int per = gridDim.x >> 3;
    const int b = blockIdx.x;
    w = (b & 7) * per + (b >> 3);
    return w < total;
}

// map_work over the blocks from b0 on (b0 a multiple of 8, so that a block
// keeps its XCD): the blocks before b0 do other work of the same launch.
__device__ __forceinline__ bool map_work_at(int b0, int total, int& w)
{
    const int per = ((int)gridDim.x - b0) >> 3;
    const int b = (int)blockIdx.x - b0;
    w = (b & 7) * per + (b >> 3);
    return w < total;
}

static inline int padded_grid(int total) { return ((total + 7) / 8) * 8; }

// Radial kernels: work item w -> (mode, run) with the runs of one mode
// adjacent, so that runs sharing a factor set read it through one XCD's L2.
__device__ __forceinline__ void split_mode_major(int w, int nrun, int& m, int& ri)
{
    m = w / nrun;
    ri = w - m * nrun;
}

// ------------------------------------------------------------------ DFT
template <int R> __device__ __forceinline__ void dft_small(double2* v, int sign);

template <> __device__ __forceinline__ void dft_small<2>(double2* v, int)
{
    const double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <> __device__ __forceinline__ void dft_small<3>(double2* v, int sign)
{
    const double s = (double)sign * 0.86602540378443864676372317075294;
    const double2 t1 = cadd(v[1], v[2]);
    const double2 t2 = mk(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
    const double2 t3 = csc(s, csub(v[1], v[2]));
    v[0] = cadd(v[0], t1);
    v[1] = mk(t2.x - t3.y, t2.y + t3.x);
    v[2] = mk(t2.x + t3.y, t2.y - t3.x);
}

template <> __device__ __forceinline__ void dft_small<4>(double2* v, int sign)
{
    const double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    const double2 a2 = cadd(v[1], v[3]), a3 = csub(v[1], v[3]);
    const double2 wa3 = csel(sign > 0, mk(-a3.y, a3.x), mk(a3.y, -a3.x));   // (i*sign)*a3
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, wa3);
    v[3] = csub(a1, wa3);
}

template <> __device__ __forceinline__ void dft_small<5>(double2* v, int sign)
{
    const double c1 = 0.30901699437494742410229341718282;    // cos(2pi/5)
    const double c2 = -0.80901699437494742410229341718282;   // cos(4pi/5)
    const double s1 = (double)sign * 0.95105651629515357211643933337938;  // sin(2pi/5)
    const double s2 = (double)sign * 0.58778525229247312916870595463907;  // sin(4pi/5)
    const double2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const double2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const double2 a1 = mk(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
    const double2 a2 = mk(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
    const double2 b1 = mk(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
    const double2 b2 = mk(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = mk(a1.x - b1.y, a1.y + b1.x);
    v[4] = mk(a1.x + b1.y, a1.y - b1.x);
    v[2] = mk(a2.x - b2.y, a2.y + b2.x);
    v[3] = mk(a2.x + b2.y, a2.y - b2.x);
}

// One in-place Stockham pass of radix R over s[0..n) (in LDS): every
// butterfly's inputs are read into registers before any output is written.  TWL: tw is
// the table's copy in LDS (radial kernels); else tw is the global table and
// the pass's twiddle loads are all issued before the data is read.
template <int NB_T, int R, int NMAX, bool TWL>
__device__ __forceinline__ void fft_pass(double2* s, int n, int p, const double2* __restrict__ tw, int sign)
{
    constexpr int NT = NB_T;
    constexpr int MAXB = (NMAX + R * NT - 1) / (R * NT);
    constexpr int NW = TWL ? 1 : MAXB;     // hoisted twiddle rows (global table only)
    const int tid = threadIdx.x;
    const int nb = n / R;
    const int tws = n / (p * R);
    auto twiddle = [&](int idx) {
        double2 w;
        if (TWL) {
            const LDS_AS double* t = (const LDS_AS double*)tw;
            w.x = t[2 * idx];
            w.y = t[2 * idx + 1];
        } else {
            const GLOBAL_AS double* t = (const GLOBAL_AS double*)tw;
            w.x = t[2 * idx];
            w.y = t[2 * idx + 1];
        }
        return w;
    };
    double2 w[NW][R - 1];
    if (!TWL) {
#pragma unroll
        for (int q = 0; q < MAXB; ++q) {
            const int b = tid + q * NT;
            if (b < nb) {
                const int k = ((p & (p - 1)) == 0) ? (b & (p - 1)) : (b % p);
#pragma unroll
                for (int r = 1; r < R; ++r) w[q < NW ? q : 0][r - 1] = twiddle(r * k * tws);
            }
        }
    }
    double2 v[MAXB][R];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = tid + q * NT;
        if (b < nb) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = lload(s + b + r * nb);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = tid + q * NT;
        if (b < nb) {
            const int k = ((p & (p - 1)) == 0) ? (b & (p - 1)) : (b % p);
#pragma unroll
            for (int r = 1; r < R; ++r) {
                double2 wr = TWL ? twiddle(r * k * tws) : w[q < NW ? q : 0][r - 1];
                if (sign > 0) wr.y = -wr.y;
                v[q][r] = cmul(v[q][r], wr);
            }
            dft_small<R>(v[q], sign);
            const int j = (b - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) lstore(s + j + r * p, v[q][r]);
        }
    }
    __syncthreads();
}

// One in-place Stockham pass of a radix R that has no butterfly of its own
// (a prime above 5, or their product): each output of the pass is a direct
// R-term sum, its twiddle exp(-+2 pi i (r k (n/(pR)) + (r q mod R) n/R) / n)
// one entry of the length-n table.  Each thread keeps its outputs in
// registers until every input has been read.
template <int NT, int NMAX, bool TWL>
__device__ __forceinline__ void fft_pass_gen(double2* s, int n, int p, int R, const double2* __restrict__ tw, int sign)
{
    constexpr int MO = (NMAX + NT - 1) / NT;
    const int tid = threadIdx.x;
    const int nb = n / R;
    const int tws = n / (p * R);
    double2 y[MO];
#pragma unroll
    for (int u = 0; u < MO; ++u) {
        const int o = tid + u * NT;     // output o = (b, q)
        if (o < n) {
            const int b = o % nb, q = o / nb;
            const int k = b % p;
            double2 acc = mk(0.0, 0.0);
            int rq = 0;                  // (r q) mod R
            for (int r = 0; r < R; ++r) {
                const int e = (int)(((long long)r * k * tws + (long long)rq * nb) % n);
                double2 w;
                if (TWL) {
                    const LDS_AS double* t = (const LDS_AS double*)tw;
                    w.x = t[2 * e];
                    w.y = t[2 * e + 1];
                } else {
                    w = gload(tw + e);
                }
                if (sign > 0) w.y = -w.y;
                const double2 v = lload(s + b + r * nb);
                acc = cadd(acc, cmul(v, w));
                rq += q;
                if (rq >= R) rq -= R;
            }
            y[u] = acc;
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < MO; ++u) {
        const int o = tid + u * NT;
        if (o < n) {
            const int b = o % nb, q = o / nb;
            const int k = b % p;
            lstore(s + (b - k) * R + k + q * p, y[u]);
        }
    }
    __syncthreads();
}

// s[k] <- sum_j s[j] exp(sign 2 pi i j k / n), natural order in and out,
// twiddles from the LDS table tw (exp(-2 pi i t / n), t < n).  Kept out of
// line: it is called a dozen times per mode kernel, and the plan travels as
// scalars so no kernel argument ever needs an address.
template <int NT, int NMAX>
__device__ __noinline__ void fft_run(double2* s, int n, int nf, unsigned long long code, int rg,
                                     const double2* __restrict__ tw, int sign)
{
    int p = 1;
    for (int f = 0; f < nf; ++f) {
        const int R = (int)((code >> (4 * f)) & 15ull);
        if (R == 4) fft_pass<NT, 4, NMAX, true>(s, n, p, tw, sign);
        else if (R == 2) fft_pass<NT, 2, NMAX, true>(s, n, p, tw, sign);
        else if (R == 3) fft_pass<NT, 3, NMAX, true>(s, n, p, tw, sign);
        else if (R == 5) fft_pass<NT, 5, NMAX, true>(s, n, p, tw, sign);
        else fft_pass_gen<NT, NMAX, true>(s, n, p, rg, tw, sign);
        p *= (R == 0 ? rg : R);
    }
}

template <int NT, int NMAX>
__device__ __forceinline__ void fft(double2* s, const FftPlan& P, int sign)
{
    fft_run<NT, NMAX>(s, P.n, P.nf, P.code, P.rg, P.tw, sign);
}

// Inlined variant with the global twiddle table, for the nonlinear pass.
// GEN: the plan may end with a direct-DFT pass (kept out of the common
// instantiations, whose registers it would otherwise cost).
template <int NT, int NMAX, bool GEN = true>
__device__ __forceinline__ void fft_inl(double2* s, const FftPlan& P, int sign)
{
    int p = 1;
    for (int f = 0; f < P.nf; ++f) {
        const int R = (int)((P.code >> (4 * f)) & 15ull);
        if (R == 4) fft_pass<NT, 4, NMAX, false>(s, P.n, p, P.tw, sign);
        else if (R == 2) fft_pass<NT, 2, NMAX, false>(s, P.n, p, P.tw, sign);
        else if (R == 3) fft_pass<NT, 3, NMAX, false>(s, P.n, p, P.tw, sign);
        else if (!GEN || R == 5) fft_pass<NT, 5, NMAX, false>(s, P.n, p, P.tw, sign);
        else fft_pass_gen<NT, NMAX, false>(s, P.n, p, P.rg, P.tw, sign);
        p *= (R == 0 ? P.rg : R);
    }
}

// Two independent transforms of the same plan in one sweep of passes (LDS
// buffers s1, s2; twiddles from the LDS table, or from the global table with
// the reads hoisted ahead of the pass): the butterflies of both share each
// pass's barriers and twiddle reads.
template <int NB_T, int R, int NMAX, bool TWL>
__device__ __forceinline__ void fft_pass2(double2* s1, double2* s2, int n, int p, const double2* __restrict__ tw,
                                          int sign)
{
    constexpr int NT = NB_T;
    constexpr int MAXB = (NMAX + R * NT - 1) / (R * NT);
    constexpr int NW = TWL ? 1 : MAXB;
    const int tid = threadIdx.x;
    const int nb = n / R;
    const int tws = n / (p * R);
    auto twiddle = [&](int idx) {
        double2 w;
        if (TWL) {
            const LDS_AS double* t = (const LDS_AS double*)tw;
            w.x = t[2 * idx];
            w.y = t[2 * idx + 1];
        } else {
            const GLOBAL_AS double* t = (const GLOBAL_AS double*)tw;
            w.x = t[2 * idx];
            w.y = t[2 * idx + 1];
        }
        return w;
    };
    double2 w[NW][R - 1];
    if (!TWL) {
#pragma unroll
        for (int q = 0; q < MAXB; ++q) {
            const int b = tid + q * NT;
            if (b < nb) {
                const int k = ((p & (p - 1)) == 0) ? (b & (p - 1)) : (b % p);
#pragma unroll
                for (int r = 1; r < R; ++r) w[q < NW ? q : 0][r - 1] = twiddle(r * k * tws);
            }
        }
    }
    double2 v[MAXB][R], u[MAXB][R];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = tid + q * NT;
        if (b < nb) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                v[q][r] = lload(s1 + b + r * nb);
                u[q][r] = lload(s2 + b + r * nb);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = tid + q * NT;
        if (b < nb) {
            const int k = ((p & (p - 1)) == 0) ? (b & (p - 1)) : (b % p);
#pragma unroll
            for (int r = 1; r < R; ++r) {
                double2 wr = TWL ? twiddle(r * k * tws) : w[q < NW ? q : 0][r - 1];
                if (sign > 0) wr.y = -wr.y;
                v[q][r] = cmul(v[q][r], wr);
                u[q][r] = cmul(u[q][r], wr);
            }
            dft_small<R>(v[q], sign);
            dft_small<R>(u[q], sign);
            const int j = (b - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                lstore(s1 + j + r * p, v[q][r]);
                lstore(s2 + j + r * p, u[q][r]);
            }
        }
    }
    __syncthreads();
}

// The first pass (p = 1) of the paired transforms with its inputs from
// in(j, a, c) -- a = s1[j], c = s2[j] -- instead of LDS: fft_pass2's
// butterflies, each with the same twiddle reads and arithmetic, writing LDS.
// Nothing in LDS is read, so no barrier comes before the writes.
template <int NT, int R, int NMAX, class In>
__device__ __forceinline__ void fft_pass2_in(double2* s1, double2* s2, int n, const double2* __restrict__ tw, int sign,
                                             In in)
{
    constexpr int MAXB = (NMAX + R * NT - 1) / (R * NT);
    const int tid = threadIdx.x;
    const int nb = n / R;
    const int tws = n / R;   // n / (p R), p = 1
    const GLOBAL_AS double* t = (const GLOBAL_AS double*)tw;
    double2 w[MAXB][R - 1], v[MAXB][R], u[MAXB][R];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = tid + q * NT;
        if (b < nb) {
#pragma unroll
            for (int r = 1; r < R; ++r) w[q][r - 1] = mk(t[0], t[1]);   // k = 0: twiddle index r k tws = 0
#pragma unroll
            for (int r = 0; r < R; ++r) in(b + r * nb, v[q][r], u[q][r]);
        }
    }
    (void)tws;
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = tid + q * NT;
        if (b < nb) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                double2 wr = w[q][r - 1];
                if (sign > 0) wr.y = -wr.y;
                v[q][r] = cmul(v[q][r], wr);
                u[q][r] = cmul(u[q][r], wr);
            }
            dft_small<R>(v[q], sign);
            dft_small<R>(u[q], sign);
            const int j = b * R;   // (b - k) R + k, k = 0
#pragma unroll
            for (int r = 0; r < R; ++r) {
                lstore(s1 + j + r, v[q][r]);
                lstore(s2 + j + r, u[q][r]);
            }
        }
    }
    __syncthreads();
}

// fft2_inl (NMAX <= 768, first factor 2-5) with the first pass's inputs from
// in(j, a, c) (fft_pass2_in), the other passes in LDS.
template <int NT, int NMAX, bool GEN, class In>
__device__ __forceinline__ void fft2_inl_in(double2* s1, double2* s2, const FftPlan& P, int sign, In in)
{
    static_assert(NMAX <= 768, "paired passes only");
    const int R0 = (int)(P.code & 15ull);
    if (R0 == 4) fft_pass2_in<NT, 4, NMAX>(s1, s2, P.n, P.tw, sign, in);
    else if (R0 == 2) fft_pass2_in<NT, 2, NMAX>(s1, s2, P.n, P.tw, sign, in);
    else if (R0 == 3) fft_pass2_in<NT, 3, NMAX>(s1, s2, P.n, P.tw, sign, in);
    else fft_pass2_in<NT, 5, NMAX>(s1, s2, P.n, P.tw, sign, in);
    int p = R0;
    for (int f = 1; f < P.nf; ++f) {
        const int R = (int)((P.code >> (4 * f)) & 15ull);
        if (R == 4) fft_pass2<NT, 4, NMAX, false>(s1, s2, P.n, p, P.tw, sign);
        else if (R == 2) fft_pass2<NT, 2, NMAX, false>(s1, s2, P.n, p, P.tw, sign);
        else if (R == 3) fft_pass2<NT, 3, NMAX, false>(s1, s2, P.n, p, P.tw, sign);
        else if (!GEN || R == 5) fft_pass2<NT, 5, NMAX, false>(s1, s2, P.n, p, P.tw, sign);
        else {
            fft_pass_gen<NT, NMAX, false>(s1, P.n, p, P.rg, P.tw, sign);
            fft_pass_gen<NT, NMAX, false>(s2, P.n, p, P.rg, P.tw, sign);
        }
        p *= (R == 0 ? P.rg : R);
    }
}

// A paired pass (both transforms' butterflies in one thread, global twiddle
// table) whose outputs go to out(j, z1_j, z2_j) instead of back to LDS: the
// last inverse pass of the nonlinear pass, which forms the products there.
template <int NT, int R, int NMAX, class F>
__device__ __forceinline__ void fft_pass2f(double2* s1, double2* s2, int n, int p, const double2* __restrict__ tw,
                                           int sign, F out)
{
    constexpr int MAXB = (NMAX + R * NT - 1) / (R * NT);
    const int tid = threadIdx.x;
    const int nb = n / R;
    const int tws = n / (p * R);
    double2 w[MAXB][R - 1], v[MAXB][R], u[MAXB][R];
    const GLOBAL_AS double* t = (const GLOBAL_AS double*)tw;
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = tid + q * NT;
        if (b < nb) {
            const int k = ((p & (p - 1)) == 0) ? (b & (p - 1)) : (b % p);
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const int idx = r * k * tws;
                w[q][r - 1] = mk(t[2 * idx], t[2 * idx + 1]);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                v[q][r] = lload(s1 + b + r * nb);
                u[q][r] = lload(s2 + b + r * nb);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = tid + q * NT;
        if (b < nb) {
            const int k = ((p & (p - 1)) == 0) ? (b & (p - 1)) : (b % p);
#pragma unroll
            for (int r = 1; r < R; ++r) {
                double2 wr = w[q][r - 1];
                if (sign > 0) wr.y = -wr.y;
                v[q][r] = cmul(v[q][r], wr);
                u[q][r] = cmul(u[q][r], wr);
            }
            dft_small<R>(v[q], sign);
            dft_small<R>(u[q], sign);
            const int j = (b - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) out(j + r * p, v[q][r], u[q][r]);
        }
    }
    __syncthreads();
}

// The nonlinear pass's paired transforms of z1 and z2 by its plan pq (the
// factors 3 last); fused: the last pass hands its outputs to out instead of
// LDS.
template <int NT, int NMAX, class F>
__device__ __forceinline__ void fft2_nl(double2* s1, double2* s2, const FftPlan& P, int sign, bool fused, F out)
{
    int p = 1;
    for (int f = 0; f < P.nf; ++f) {
        const int R = (int)((P.code >> (4 * f)) & 15ull);
        if (fused && f == P.nf - 1) {
            if (R == 3) fft_pass2f<NT, 3, NMAX>(s1, s2, P.n, p, P.tw, sign, out);
            else if (R == 2) fft_pass2f<NT, 2, NMAX>(s1, s2, P.n, p, P.tw, sign, out);
            else if (R == 4) fft_pass2f<NT, 4, NMAX>(s1, s2, P.n, p, P.tw, sign, out);
            else fft_pass2f<NT, 5, NMAX>(s1, s2, P.n, p, P.tw, sign, out);
        } else if (R == 4) fft_pass2<NT, 4, NMAX, false>(s1, s2, P.n, p, P.tw, sign);
        else if (R == 2) fft_pass2<NT, 2, NMAX, false>(s1, s2, P.n, p, P.tw, sign);
        else if (R == 3) fft_pass2<NT, 3, NMAX, false>(s1, s2, P.n, p, P.tw, sign);
        else fft_pass2<NT, 5, NMAX, false>(s1, s2, P.n, p, P.tw, sign);
        p *= R;
    }
}

template <int NT, int NMAX>
__device__ __noinline__ void fft2_run(double2* s1, double2* s2, int n, int nf, unsigned long long code, int rg,
                                      const double2* __restrict__ tw, int sign)
{
    int p = 1;
    for (int f = 0; f < nf; ++f) {
        const int R = (int)((code >> (4 * f)) & 15ull);
        if (R == 4) fft_pass2<NT, 4, NMAX, true>(s1, s2, n, p, tw, sign);
        else if (R == 2) fft_pass2<NT, 2, NMAX, true>(s1, s2, n, p, tw, sign);
        else if (R == 3) fft_pass2<NT, 3, NMAX, true>(s1, s2, n, p, tw, sign);
        else if (R == 5) fft_pass2<NT, 5, NMAX, true>(s1, s2, n, p, tw, sign);
        else {
            fft_pass_gen<NT, NMAX, true>(s1, n, p, rg, tw, sign);
            fft_pass_gen<NT, NMAX, true>(s2, n, p, rg, tw, sign);
        }
        p *= (R == 0 ? rg : R);
    }
}

// Two transforms in one sweep of passes with the global twiddle table (the
// nonlinear pass); up to length 768, beyond which the paired registers would
// cost occupancy and the transforms run one after the other.
template <int NT, int NMAX, bool GEN = true>
__device__ __forceinline__ void fft2_inl(double2* s1, double2* s2, const FftPlan& P, int sign)
{
    if constexpr (NMAX <= 768) {
        int p = 1;
        for (int f = 0; f < P.nf; ++f) {
            const int R = (int)((P.code >> (4 * f)) & 15ull);
            if (R == 4) fft_pass2<NT, 4, NMAX, false>(s1, s2, P.n, p, P.tw, sign);
            else if (R == 2) fft_pass2<NT, 2, NMAX, false>(s1, s2, P.n, p, P.tw, sign);
            else if (R == 3) fft_pass2<NT, 3, NMAX, false>(s1, s2, P.n, p, P.tw, sign);
            else if (!GEN || R == 5) fft_pass2<NT, 5, NMAX, false>(s1, s2, P.n, p, P.tw, sign);
            else {
                fft_pass_gen<NT, NMAX, false>(s1, P.n, p, P.rg, P.tw, sign);
                fft_pass_gen<NT, NMAX, false>(s2, P.n, p, P.rg, P.tw, sign);
            }
            p *= (R == 0 ? P.rg : R);
        }
    } else {
        fft_inl<NT, NMAX, GEN>(s1, P, sign);
        fft_inl<NT, NMAX, GEN>(s2, P, sign);
    }
}

// Radial kernels: copy the length-M twiddle table into LDS (twl, M entries)
// and return a shape whose radial plan reads it there.
template <int NT>
__device__ __forceinline__ Shape lds_twiddles(const Shape& sh, double2* twl)
{
    for (int i = threadIdx.x; i < sh.M; i += NT) twl[i] = gload(sh.pm.tw + i);
    __syncthreads();
    Shape s2 = sh;
    s2.pm.tw = twl;
    return s2;
}

// ------------------------------------------------------------------ Chebyshev
// chebtransform (chebyshev_fftw.F90:124-155): coefficients out[0..N) of the
// grid values src(i), through the even extension of length M = 2N-2.
template <int NT, int MM, class F>
__device__ __forceinline__ void cheb_fwd(double2* buf, double2* out, const Shape& sh, F src)
{
    const int N = sh.N;
    for (int i = threadIdx.x; i < N; i += NT) {
        const double2 v = src(i);
        buf[N - 1 - i] = v;
        if (i >= 1 && i <= N - 2) buf[N - 1 + i] = v;
    }
    __syncthreads();
    fft<NT, MM>(buf, sh.pm, +1);
    const double dM = (double)sh.M, fac = sh.fact;
    for (int k = threadIdx.x; k < N; k += NT) {
        double2 ff = buf[k];
        ff.x = ff.x / dM;
        ff.y = ff.y / dM;
        if (k == 0 || k == N - 1) out[k] = mk(2.0 * (ff.x / fac), 2.0 * (ff.y / fac));
        else out[k] = mk((2.0 * ff.x) / fac, (2.0 * ff.y) / fac);
    }
    __syncthreads();
}

// chebinvtran (:157-194): grid values from coefficients c(k); dst(i, value).
template <int NT, int MM, class F, class G>
__device__ __forceinline__ void cheb_inv(double2* buf, const Shape& sh, F c, G dst)
{
    const int N = sh.N, M = sh.M;
    for (int i = threadIdx.x; i < N; i += NT) {
        const double2 v = c(i);
        buf[i] = v;
        if (i >= 1 && i <= N - 2) buf[M - i] = v;
    }
    __syncthreads();
    fft<NT, MM>(buf, sh.pm, +1);
    const double fact = sh.fact;
    for (int i = threadIdx.x; i < N; i += NT) {
        const double2 y = buf[i];
        dst(N - 1 - i, mk((y.x * 0.5) * fact, (y.y * 0.5) * fact));
    }
    __syncthreads();
}

// Derivative coefficients of the backward recurrence of chebinvtranD1
// (:262-267): beta(k) = sum over j > k, j - k odd, of c_j f_j with c_j = 4j
// (2(N-1) for the last index); evaluated as alternating suffix sums by a
// log-depth scan.  beta is an LDS vector of N.
template <int NT, int MM, class F>
__device__ __forceinline__ void cheb_beta(double2* beta, const Shape& sh, F c)
{
    const int N = sh.N;
    constexpr int Q = (MM / 2 + 1 + NT - 1) / NT;   // N <= MM/2 + 1
    for (int k = threadIdx.x; k < N; k += NT) {
        double2 h = mk(0.0, 0.0);
        if (k < N - 1) {
            const int j = k + 1;
            const double cj = (j == N - 1) ? 2.0 * (double)(N - 1) : 4.0 * (double)j;
            h = csc(cj, c(j));
        }
        beta[k] = h;
    }
    __syncthreads();
    for (int d = 2; d < N; d *= 2) {
        double2 t[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int k = threadIdx.x + q * NT;
            if (k < N) {
                t[q] = beta[k];
                if (k + d < N) t[q] = cadd(t[q], beta[k + d]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int k = threadIdx.x + q * NT;
            if (k < N) beta[k] = t[q];
        }
        __syncthreads();
    }
}

// Scaling, halving and the length-M DFT of chebinvtranD1 (:269-288).
template <int NT, int MM, class G>
__device__ __forceinline__ void cheb_eval_beta(double2* buf, const double2* beta, const Shape& sh, G dst)
{
    const int N = sh.N;
    const double fact = sh.fact;
    for (int i = threadIdx.x; i < N; i += NT) {
        double2 s = csc(fact, beta[i]);
        if (i == 0 || i == N - 1) s = mk(s.x / 2.0, s.y / 2.0);
        const double2 h = mk(s.x / 2.0, s.y / 2.0);
        buf[i] = csel(i == 0, s, h);
        if (i >= 1 && i <= N - 2) buf[2 * N - 2 - i] = h;
    }
    __syncthreads();
    fft<NT, MM>(buf, sh.pm, -1);
    for (int i = threadIdx.x; i < N; i += NT) dst(N - 1 - i, buf[i]);
    __syncthreads();
}

// cheb_eval_beta of two derivative-coefficient vectors at once (buffers buf
// and buf2, each of at least M entries); dst1/dst2 receive the two results.
// Up to M = 256 the two DFTs share their passes.
template <int NT, int MM, class G1, class G2>
__device__ __forceinline__ void cheb_eval_beta2(double2* buf, double2* buf2, const double2* beta1,
                                                const double2* beta2, const Shape& sh, G1 dst1, G2 dst2)
{
    const int N = sh.N;
    const double fact = sh.fact;
    auto prep = [&](double2* b, const double2* beta, int i) {
        double2 s = csc(fact, beta[i]);
        if (i == 0 || i == N - 1) s = mk(s.x / 2.0, s.y / 2.0);
        const double2 h = mk(s.x / 2.0, s.y / 2.0);
        b[i] = csel(i == 0, s, h);
        if (i >= 1 && i <= N - 2) b[2 * N - 2 - i] = h;
    };
    for (int i = threadIdx.x; i < N; i += NT) {
        prep(buf, beta1, i);
        prep(buf2, beta2, i);
    }
    __syncthreads();
    if constexpr (MM <= 256) {
        fft2_run<NT, MM>(buf, buf2, sh.pm.n, sh.pm.nf, sh.pm.code, sh.pm.rg, sh.pm.tw, -1);
    } else {   // larger tiers: the paired registers would not fit
        fft<NT, MM>(buf, sh.pm, -1);
        fft<NT, MM>(buf2, sh.pm, -1);
    }
    for (int i = threadIdx.x; i < N; i += NT) {
        dst1(N - 1 - i, buf[i]);
        dst2(N - 1 - i, buf2[i]);
    }
    __syncthreads();
}

// ------------------------------------------------------------------ solves
// Every implicit solve of a step is a kernel of its own between the radial
// kernels that build the right-hand sides and those that use the solutions.
// Each run keeps its right-hand sides and solutions in its solve area sx
// (2N complex per mode: temperature or psi rows 0..N-1, omega or mean-flow
// rows N..2N-1).  Inverse mode: x = A^{-1} b with the inverse formed once from
// the LU factors (dgetrf, then the columns of A^{-1} by dgetrs on unit
// vectors) and kept resident, applied to all runs of one operator set at once
// as an FP64 MFMA product.  Only the columns that meet non-zero rows of the
// right-hand side are stored: all N for AT and A_uphi, the N-2 interior omega
// rows for AF, whose psi rows and boundary rows are zero
// (fourierloop.F90:171-178, fourierloop_imexrk.F90:341-350).  LU mode:
// dgetrs on the packed factors, one workgroup per (run, mode).

// Broadcast lane l's double to the wavefront (v_readlane on both halves).
__device__ __forceinline__ double rdl(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

// Buffer resources (raw, stride 0): loads outside [0, bytes) return zero, so
// the K padding of a product's last k-step and rows past the end of an array
// read zeros without a bounds test.
typedef __amdgpu_buffer_rsrc_t BRes;
__device__ __forceinline__ BRes bres(const void* base, size_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ double bload(BRes r, int off)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ double2 bload2(BRes r, int off)
{
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
// LDS-direct load (buffer_load_dwordx4 ... lds): each lane's 16 bytes at
// byte offset off of r land in LDS at lds + 16 * lane (lds uniform over the
// wavefront); out-of-range lanes write zeros.  Counted by vmcnt: the compiler
// does not order these against LDS reads, so callers wait (vmcnt(0) with a
// memory clobber) before reading what they loaded.
__device__ __forceinline__ void lds_dma(BRes r, double2* lds, int off)
{
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)lds, 16, off, 0, 0, 0);
}

// dgetrs('N', n, nrhs=1) on the real and imaginary parts of x (LDS, n complex)
// with a resident packed LU factor A (column-major) and its net row
// permutation.  Column blocks of NB: the diagonal block is solved by one
// wavefront, the rest of the block's columns are applied as a GEMV; every x(i)
// receives its column updates in dgetrs' column order.  Each thread's first
// GEMV row of a block is loaded before the diagonal solve, so that the two
// factor reads of a block overlap.
template <int NT, int NB, int KR>
__device__ void lu_solve_k(double2* x, double2* xs, const double* __restrict__ A, const int* __restrict__ perm,
                           const int n)
{
    const int tid = threadIdx.x;
    // dgetrs' row interchanges as one gather through the LDS scratch xs
    for (int i = tid; i < n; i += NT) lstore(xs + i, lload(x + gload(perm + i)));
    __syncthreads();
    for (int i = tid; i < n; i += NT) lstore(x + i, lload(xs + i));
    __syncthreads();
    const size_t ld = (size_t)n;
    // unit lower triangle
    for (int j0 = 0; j0 < n; j0 += NB) {
        const int jb = min(NB, n - j0);
        // the factor entries of this thread's first KR GEMV rows (i0 + k NT),
        // all loaded while wavefront 0 solves the diagonal block; further rows
        // in groups of KR, each group's loads issued together
        const int i0 = j0 + jb + tid;
        double l0[KR][NB];
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int i = i0 + k * NT;
#pragma unroll
            for (int j = 0; j < NB; ++j) l0[k][j] = (i < n && j < jb) ? gload(A + (size_t)i + ld * (size_t)(j0 + j)) : 0.0;
        }
        if (tid < 64) {
            const int l = tid;
            double a[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) a[j] = (l < jb && j < l) ? gload(A + (size_t)(j0 + l) + ld * (size_t)(j0 + j)) : 0.0;
            double2 xi = mk(0.0, 0.0);
            if (l < jb) xi = lload(x + j0 + l);
#pragma unroll
            for (int j = 0; j < NB - 1; ++j) {
                if (j < jb - 1) {
                    const double xr = rdl(xi.x, j), xm = rdl(xi.y, j);   // v_readlane: j is a constant
                    if (l > j) { xi.x = fma(-xr, a[j], xi.x); xi.y = fma(-xm, a[j], xi.y); }
                }
            }
            if (l < jb) lstore(x + j0 + l, xi);
        }
        __syncthreads();
        auto lrow = [&](int i, const double (&lv)[NB]) {
            double2 xi = lload(x + i);
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                if (j < jb) {
                    const double2 xj = lload(x + j0 + j);
                    xi.x = fma(-xj.x, lv[j], xi.x);
                    xi.y = fma(-xj.y, lv[j], xi.y);
                }
            }
            lstore(x + i, xi);
        };
#pragma unroll
        for (int k = 0; k < KR; ++k)
            if (i0 + k * NT < n) lrow(i0 + k * NT, l0[k]);
#pragma unroll 1
        for (int ib = i0 + KR * NT; ib < n; ib += KR * NT) {
            double lv[KR][NB];
#pragma unroll
            for (int k = 0; k < KR; ++k) {
                const int i = ib + k * NT;
#pragma unroll
                for (int j = 0; j < NB; ++j) lv[k][j] = (i < n && j < jb) ? gload(A + (size_t)i + ld * (size_t)(j0 + j)) : 0.0;
            }
#pragma unroll
            for (int k = 0; k < KR; ++k)
                if (ib + k * NT < n) lrow(ib + k * NT, lv[k]);
        }
        __syncthreads();
    }
    // upper triangle
    const int nblk = (n + NB - 1) / NB;
    for (int bi = nblk - 1; bi >= 0; --bi) {
        const int j0 = bi * NB;
        const int jb = min(NB, n - j0);
        double u0[KR][NB];
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int i = tid + k * NT;
#pragma unroll
            for (int j = 0; j < NB; ++j) u0[k][j] = (i < j0 && j < jb) ? gload(A + (size_t)i + ld * (size_t)(j0 + j)) : 0.0;
        }
        if (tid < 64) {
            const int l = tid;
            double a[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) a[j] = (l < jb && j >= l && j < jb) ? gload(A + (size_t)(j0 + l) + ld * (size_t)(j0 + j)) : 1.0;
            double2 xi = mk(0.0, 0.0);
            if (l < jb) xi = lload(x + j0 + l);
#pragma unroll
            for (int j = NB - 1; j >= 0; --j) {
                if (j < jb) {
                    if (l == j) { xi.x = xi.x / a[j]; xi.y = xi.y / a[j]; }
                    const double xr = rdl(xi.x, j), xm = rdl(xi.y, j);   // v_readlane: j is a constant
                    if (l < j) { xi.x = fma(-xr, a[j], xi.x); xi.y = fma(-xm, a[j], xi.y); }
                }
            }
            if (l < jb) lstore(x + j0 + l, xi);
        }
        __syncthreads();
        auto urow = [&](int i, const double (&uv)[NB]) {
            double2 xi = lload(x + i);
#pragma unroll
            for (int j = NB - 1; j >= 0; --j) {
                if (j < jb) {
                    const double2 xj = lload(x + j0 + j);
                    xi.x = fma(-xj.x, uv[j], xi.x);
                    xi.y = fma(-xj.y, uv[j], xi.y);
                }
            }
            lstore(x + i, xi);
        };
#pragma unroll
        for (int k = 0; k < KR; ++k)
            if (tid + k * NT < j0) urow(tid + k * NT, u0[k]);
#pragma unroll 1
        for (int ib = tid + KR * NT; ib < j0; ib += KR * NT) {
            double uv[KR][NB];
#pragma unroll
            for (int k = 0; k < KR; ++k) {
                const int i = ib + k * NT;
#pragma unroll
                for (int j = 0; j < NB; ++j) uv[k][j] = (i < j0 && j < jb) ? gload(A + (size_t)i + ld * (size_t)(j0 + j)) : 0.0;
            }
#pragma unroll
            for (int k = 0; k < KR; ++k)
                if (ib + k * NT < j0) urow(ib + k * NT, uv[k]);
        }
        __syncthreads();
    }
}

// lu_solve_k with the GEMV rows of a thread loaded one at a time (KR = 1,
// the radial kernels, whose registers are spoken for) or two at a time.
template <int NT, int NB, int KR = 1>
__device__ void lu_solve(double2* x, double2* xs, const double* __restrict__ A, const int* __restrict__ perm, const int n)
{
    if (KR == 1 || n <= NT + NB) lu_solve_k<NT, NB, 1>(x, xs, A, perm, n);
    else lu_solve_k<NT, NB, 2>(x, xs, A, perm, n);
}

constexpr int SOLVE_NB = 8;    // column block of lu_solve
// dgetrs('N', n, nrhs=1) on one real vector (the real or the imaginary part
// of a right-hand side; dgetrs solves them separately) by one wavefront
// (n <= 64 SLOTS) with the packed LU factor read by col(p, j) = row
// lane + 64 p of column j (column-major, staged in LDS; rows past n read the
// next column or padding): lane l keeps rows l, l + 64, ... of x in xr; the
// forward sweep takes the columns of the unit lower factor in ascending
// order, the backward sweep those of the upper factor in descending order
// (x_j / U_jj by its lane, then broadcast), each row receiving its updates
// in dgetrs' column order with lu_solve's arithmetic, so the results are
// lu_solve's.  The sweeps go one 64-row slot Q at a time, so that x_j comes
// from a slot known at compile time (v_readlane, no select): slot Q's own
// rows take the update where below (above) the diagonal, the later
// (earlier) slots' rows always, the others never.  A slot reads only the
// columns 64 Q .. 64 Q + 63 (D ahead of their use, within the slot), so a
// caller may stage one slot's columns at a time.

// forward sweep over the columns of slot Q
template <int Q, int SLOTS, int D, class Col>
__device__ __forceinline__ void lu_wave_fwd(double (&xr)[SLOTS], Col col, const int n)
{
    static_assert(64 % D == 0, "the column ring steps through 64-row slots evenly");
    const int lane = threadIdx.x & 63;
    const int jlo = 64 * Q, jhi = min(64 * Q + 63, n - 1);
    if (jlo >= n) return;
    double c[D][SLOTS];   // ring of prefetched columns
    auto column = [&](double (&cv)[SLOTS], int j) {
        j = min(max(j, jlo), jhi);
#pragma unroll
        for (int p = 0; p < SLOTS; ++p) cv[p] = col(p, j);
    };
#pragma unroll
    for (int d = 0; d < D; ++d) column(c[d], jlo + d);
    for (int l0 = 0; l0 < 64 && jlo + l0 < n; l0 += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int jl = l0 + d, j = jlo + jl;
            if (j < n) {
                const double xj = rdl(xr[Q], jl);
                if (lane > jl) xr[Q] = fma(-xj, c[d][Q], xr[Q]);
#pragma unroll
                for (int p = Q + 1; p < SLOTS; ++p) xr[p] = fma(-xj, c[d][p], xr[p]);
                column(c[d], j + D);
            }
        }
    }
}

// backward sweep over the columns of slot Q
template <int Q, int SLOTS, int D, class Col>
__device__ __forceinline__ void lu_wave_bwd(double (&xr)[SLOTS], Col col, const int n)
{
    const int lane = threadIdx.x & 63;
    const int qtop = (n - 1) >> 6, ltop = (n - 1) & 63;
    if (Q > qtop) return;
    const int jlo = 64 * Q, jhi = min(64 * Q + 63, n - 1);
    double c[D][SLOTS];
    auto column = [&](double (&cv)[SLOTS], int j) {
        j = min(max(j, jlo), jhi);
#pragma unroll
        for (int p = 0; p < SLOTS; ++p) cv[p] = col(p, j);
    };
    // column j is step n-1-j of the sweep and sits in ring slot (n-1-j) mod D:
    // the slot's first group starts `phase` entries in, so ring slots
    // d < phase first serve the second group
    const int jtop = (Q == qtop) ? ltop : 63;
    const int phase = (n - 1 - (64 * Q + jtop)) % D;
#pragma unroll
    for (int d = 0; d < D; ++d) column(c[d], jlo + jtop + phase - d - (d < phase ? D : 0));
    for (int jh = jtop + phase; jh >= 0; jh -= D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int jl = jh - d, j = jlo + jl;
            if (jl >= 0 && jl <= jtop) {
                if (lane == jl) xr[Q] = xr[Q] / c[d][Q];   // x_j / U_jj
                const double xj = rdl(xr[Q], jl);
                if (lane < jl) xr[Q] = fma(-xj, c[d][Q], xr[Q]);
#pragma unroll
                for (int p = 0; p < Q; ++p) xr[p] = fma(-xj, c[d][p], xr[p]);
                column(c[d], j - D);
            }
        }
    }
}

// The whole solve with every column readable by col.
template <int SLOTS, int D, class Col>
__device__ void lu_solve_wave(double (&xr)[SLOTS], Col col, const int n)
{
    lu_wave_fwd<0, SLOTS, D>(xr, col, n);
    if constexpr (SLOTS > 1) lu_wave_fwd<1, SLOTS, D>(xr, col, n);
    if constexpr (SLOTS > 2) lu_wave_fwd<2, SLOTS, D>(xr, col, n);
    if constexpr (SLOTS > 2) lu_wave_bwd<2, SLOTS, D>(xr, col, n);
    if constexpr (SLOTS > 1) lu_wave_bwd<1, SLOTS, D>(xr, col, n);
    lu_wave_bwd<0, SLOTS, D>(xr, col, n);
}

constexpr int SOLVE_KC = 16;   // MFMA k-steps (4 columns of A each) loaded ahead per wavefront

// Workgroups per (chunk, mode) of a batched solve phase: groups of four 16-row
// strips of the 2N-row (phase 1) or N-row (phase 0) inverse.
__host__ __device__ inline int solve_groups(int N, int phase)
{
    const int rows = phase == 0 ? N : 2 * N;
    return (((rows + 15) >> 4) + 3) >> 2;
}

typedef double dbl4 __attribute__((ext_vector_type(4)));

// Y = A X for the right-hand sides staged in LDS, with v_mfma_f64_16x16x4f64.
// A is column-major n_out x ncols (leading dimension n_out), resource Ar of
// exactly n_out * ncols doubles; Xs[k * W + j] holds real column j of X
// (k < K4 = 4 ceil(ncols/4), zero rows and columns padded up to K4 and
// W = 16 nct).  Wavefront wv of the workgroup takes the 16-row strips
// s0 + wv, s0 + wv + 4, ... below s1 as one stream of (strip, k-chunk)
// iterations: the A operands of
// the next SOLVE_KC k-steps are in flight while the current ones are applied,
// across strip boundaries.  Operand layout of the instruction: lane l holds
// A(l%16, l/16) and X(l/16, l%16) of a 16x4 by 4x16 step and accumulates
// Y(4r + l/16, l%16) in its register r = 0..3.  Columns k >= ncols read zero
// from the resource; rows >= n_out reach only discarded outputs.  Every
// output is the sum over k in ascending steps of four, whatever the other
// columns hold.
template <class ST>
__device__ __forceinline__ void mfma_apply(const BRes Ar, const int n_out, const int ncols, const double* Xs,
                                           const int W, const int nct, const int s0, const int s1, ST store)
{
    const int lane = threadIdx.x & 63;
    const int wv = s0 + (threadIdx.x >> 6);
    const int li = lane & 15, lk = lane >> 4;
    const int nk = (ncols + 3) >> 2;
    const int nc = (nk + SOLVE_KC - 1) / SOLVE_KC;            // k-chunks per strip
    const int ns = wv < s1 ? (s1 - wv + NT / 64 - 1) / (NT / 64) : 0;   // strips of this wavefront
    const int T = ns * nc;
    const LDS_AS double* X = (const LDS_AS double*)Xs;
    dbl4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
    double a0[SOLVE_KC], a1[SOLVE_KC];
    auto load = [&](double (&a)[SOLVE_KC], int it) {
        if (it >= T) return;
        const int strip = wv + (NT / 64) * (it / nc), k0 = SOLVE_KC * (it % nc);
        const int off = 8 * (strip * 16 + li + n_out * (4 * k0 + lk));
#pragma unroll
        for (int c = 0; c < SOLVE_KC; ++c) a[c] = bload(Ar, off + 32 * n_out * c);
    };
    auto step = [&](const double (&a)[SOLVE_KC], int it) {
        if (it >= T) return;
        const int strip = wv + (NT / 64) * (it / nc), kc = it % nc, k0 = SOLVE_KC * kc;
#pragma unroll
        for (int c = 0; c < SOLVE_KC; ++c) {
            if (k0 + c < nk) {
                const LDS_AS double* xr = X + (size_t)(4 * (k0 + c) + lk) * W + li;
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    if (t < nct) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[c], xr[16 * t], acc[t], 0, 0, 0);
            }
        }
        if (kc == nc - 1) {   // strip complete
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (t < nct) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) store(strip * 16 + 4 * r + lk, 16 * t + li, acc[t][r]);
                    acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
                }
        }
    };
    load(a0, 0);
    for (int it = 0; it < T; it += 2) {
        load(a1, it + 1);
        step(a0, it);
        load(a0, it + 2);
        step(a1, it + 1);
    }
}

// The m = 0 temperature system of run R (N <= M0_FUSE_MAXN) in dgetrs order,
// by a workgroup of a batched temperature-solve launch (no launch, stream,
// fork or join of its own): the factor is staged in LDS one slot's columns
// at a time (0..63, then 64..N-1 for the second slot's forward and backward
// sweeps, then 0..63 again), so the workgroup needs N complex and
// 64 N + 128 doubles of LDS; two wavefronts sweep the real and the
// imaginary parts (lu_wave_fwd / lu_wave_bwd).
constexpr int M0_FUSE_MAXN = 100;
__device__ void m0_solve(const RunDesc& R, const Shape& sh, double2* smem)
{
    const int N = sh.N, tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, im = wv & 1;
    double2* x = smem;
    double* half = (double*)(smem + N);
    const LDS_AS double* fl = (const LDS_AS double*)half;
    // columns [c0, c1) of the factor into half by LDS-direct loads (the last
    // chunk reads zeros past the factor's end into the padding)
    auto stage = [&](int c0, int c1) {
        __syncthreads();   // every read of the columns staged before is done
        const int cnt = (c1 - c0) * N;
        const BRes fr = bres(R.Ti + (size_t)c0 * N, sizeof(double) * (size_t)cnt);
        for (int c = wv; 128 * c < cnt; c += NT / 64) lds_dma(fr, (double2*)(half + 128 * c), 16 * (64 * c + lane));
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
    };
    for (int i = tid; i < N; i += NT) lstore(x + i, gload((const double2*)R.sx + i));
    stage(0, min(64, N));
    double xr[2] = {0.0, 0.0};
    if (tid < 128) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int i = lane + 64 * q;
            const double2 v = (i < N) ? lload(x + gload(R.pT + i)) : mk(0.0, 0.0);
            xr[q] = im ? v.y : v.x;
        }
    }
    auto colA = [fl, lane, N](int p, int j) { return fl[lane + 64 * p + N * j]; };
    auto colB = [fl, lane, N](int p, int j) { return fl[lane + 64 * p + N * (j - 64)]; };
    if (tid < 128) lu_wave_fwd<0, 2, 4>(xr, colA, N);
    if (N > 64) {
        stage(64, N);
        if (tid < 128) {
            lu_wave_fwd<1, 2, 4>(xr, colB, N);
            lu_wave_bwd<1, 2, 4>(xr, colB, N);
        }
        stage(0, 64);
    }
    if (tid < 128) {
        lu_wave_bwd<0, 2, 4>(xr, colA, N);
        // every gather of x was before the stagings' barriers
        LDS_AS double* xd = (LDS_AS double*)x + im;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int i = lane + 64 * q;
            if (i < N) xd[2 * i] = xr[q];
        }
    }
    __syncthreads();
    for (int i = tid; i < N; i += NT) gstore((double2*)R.sy + i, lload(x + i));
}

// The implicit solves of one phase of a step for a group of runs:
// phase 0 the temperature (AT), phase 1 the vorticity-streamfunction (AF) of
// modes m >= 1 and the mean flow (A_uphi) of m = 0.  Right-hand sides come
// from each run's solve area sx, solutions go to its solution area sy, both
// mode-major: rows 0..2N-1 of mode m at +2N m (temperature or psi in rows
// 0..N-1, omega or mean flow in rows N..2N-1).
//   LU mode: one workgroup per (run, mode), dgetrs in its operation order.
//   Inverse mode: one workgroup per (chunk of up to SOLVE_RHS runs that share
//   an operator set, mode, group of four 16-row strips of the inverse), the
//   MFMA product over the chunk's right-hand sides.  The m = 0 temperature,
//   whose wall derivative is the Nusselt number, keeps dgetrs: with m0n > 0
//   (phase 0, inverse mode) the launch's first blocks solve it for runs
//   0..m0n-1 (m0_solve), the rest do the batched solves; otherwise k_m0
//   solves it beside this kernel.
template <int phase, bool LU>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(LU ? 3 : 1))) void k_solve(const RunDesc* __restrict__ runs, const int nrun, const SolvePlan pl,
                                              const Shape sh, const int stage, const int m0n)
{
    extern __shared__ double2 smem[];
    __shared__ const double2* sxp[SOLVE_RHS];
    __shared__ double2* syp[SOLVE_RHS];
    const int N = sh.N, Nm1 = sh.Nm1, tid = threadIdx.x;
    const size_t MS = 2 * (size_t)N;     // solve-area entries per mode
    int w;
    auto lu_one = [&](const RunDesc& R, const int m) {
        double2* x = smem;
        double2* xs = smem + 2 * N;
        const double2* v = (const double2*)R.sx + (size_t)m * MS;
        double2* y = (double2*)R.sy + (size_t)m * MS;
        if (phase == 0) {
            for (int i = tid; i < N; i += NT) lstore(x + i, gload(v + i));
            __syncthreads();
            lu_solve<NT, SOLVE_NB, 2>(x, xs, R.TiS[stage] + (size_t)m * N * N, R.pTS[stage] + (size_t)m * N, N);
            for (int i = tid; i < N; i += NT) gstore(y + i, lload(x + i));
        } else if (m == 0) {
            for (int i = tid; i < N; i += NT) lstore(x + i, gload(v + N + i));
            __syncthreads();
            lu_solve<NT, SOLVE_NB, 2>(x, xs, R.UiS[stage], R.pUS[stage], N);
            for (int i = tid; i < N; i += NT) gstore(y + N + i, mk(lload(x + i).x, 0.0));
        } else {
            for (int i = tid; i < 2 * N; i += NT) lstore(x + i, i < N ? mk(0.0, 0.0) : gload(v + i));
            __syncthreads();
            lu_solve<NT, SOLVE_NB, 2>(x, xs, R.FiS[stage] + (size_t)(m - 1) * 4 * N * N,
                                      R.pFS[stage] + (size_t)(m - 1) * 2 * N, 2 * N);
            for (int i = tid; i < 2 * N; i += NT) gstore(y + i, lload(x + i));
        }
    };
    if constexpr (LU) {
        if (!map_work(nrun * Nm1, w)) return;
        int m, ri;
        split_mode_major(w, nrun, m, ri);
        lu_one(runs[ri], m);
        return;
    }
    int b0 = 0;   // blocks before b0: the m = 0 solves (a multiple of 8)
    if constexpr (phase == 0) {
        if (m0n > 0) {
            b0 = (m0n + 7) & ~7;
            if ((int)blockIdx.x < b0) {
                if ((int)blockIdx.x < m0n) m0_solve(runs[blockIdx.x], sh, smem);
                return;
            }
        }
    }
    const int nch = phase == 0 ? pl.nchT : pl.nchF;
    const int ngrp = solve_groups(N, phase);
    const int nb = nch * (phase == 0 ? sh.Nm : Nm1) * ngrp;
    if (!map_work_at(b0, nb, w)) return;
    const int g = w % ngrp;
    w /= ngrp;
    const int mi = w / nch, c = w - mi * nch;
    const int m = phase == 0 ? mi + 1 : mi;
    const int2 ch = phase == 0 ? pl.chT[c] : pl.chF[c];
    const int* idx = (phase == 0 ? pl.idxT : pl.idxF) + ch.x;
    const int nr = ch.y;
    const RunDesc& R0 = runs[gload(idx)];
    const double* A;
    int n_out, ncols, in0, out0;   // rows of the solution, columns of A, first rhs row, first solution row
    if (phase == 0) { A = R0.Ti + (size_t)m * N * N; n_out = N; ncols = N; in0 = 0; out0 = 0; }
    else if (m == 0) { A = R0.Ui; n_out = N; ncols = N; in0 = N; out0 = N; }
    else { A = R0.Fi + (size_t)(m - 1) * 2 * N * (N - 2); n_out = 2 * N; ncols = N - 2; in0 = N + 1; out0 = 0; }
    const int s0 = 4 * g, s1 = min((n_out + 15) >> 4, s0 + 4);
    if (s0 >= s1) return;   // the mean flow has fewer strips than the 2N-row systems
    if (tid < nr) {
        const RunDesc& Rj = runs[gload(idx + tid)];
        sxp[tid] = (const double2*)Rj.sx + (size_t)m * MS;
        syp[tid] = (double2*)Rj.sy + (size_t)m * MS;
    }
    __syncthreads();
    const int nct = (2 * nr + 15) >> 4, W = 16 * nct, half = W >> 1;
    const int K4 = 4 * ((ncols + 3) >> 2);
    double* Xs = (double*)smem;
#pragma unroll 4
    for (int e = tid; e < K4 * half; e += NT) {
        const int j = e / K4, k = e - j * K4;
        double2 v = mk(0.0, 0.0);
        if (j < nr && k < ncols) v = gload(sxp[j] + in0 + k);
        Xs[(size_t)k * W + 2 * j] = v.x;
        Xs[(size_t)k * W + 2 * j + 1] = v.y;
    }
    __syncthreads();
    const bool real_out = (phase == 1 && m == 0);
    mfma_apply(bres(A, sizeof(double) * (size_t)n_out * ncols), n_out, ncols, Xs, W, nct, s0, s1,
               [&](int i, int j, double v) {
                   const int jj = j >> 1;
                   if (i < n_out && jj < nr)
                       gstore((double*)(syp[jj] + out0 + i) + (j & 1), (real_out && (j & 1)) ? 0.0 : v);
               });
}

// Inverse mode: the m = 0 temperature system of each run (one workgroup per
// run), in dgetrs order with its packed LU factor, which the workgroup first
// stages in LDS (up to M0_STAGE_MAXN); one wavefront then sweeps it.
constexpr int M0_STAGE_MAXN = 139;   // N^2 + 192 doubles and N complex within 160 KiB
__global__ __launch_bounds__(NT) void k_m0(const RunDesc* __restrict__ runs, const Shape sh, const int nrun)
{
    extern __shared__ double2 smem[];
    int w;
    if (!map_work(nrun, w)) return;
    const RunDesc& R = runs[w];
    const int N = sh.N, tid = threadIdx.x;
    double2* x = smem;
    double2* xs = smem + 2 * N;
    for (int i = tid; i < N; i += NT) lstore(x + i, gload((const double2*)R.sx + i));
    if (N <= M0_STAGE_MAXN) {
        double* f = (double*)(smem + N);
        const int nn = N * N;
        // the factor into LDS by LDS-direct loads (16 B per lane, 1 KiB per
        // wavefront instruction), all of them in flight at once; the last
        // chunk may read past the factor (zeros from the buffer resource)
        // into the padding that follows it
        {
            const int wv = tid >> 6, lane = tid & 63;
            const BRes fr = bres(R.Ti, sizeof(double) * (size_t)nn);
            for (int c = wv; 128 * c < nn; c += NT / 64)
                lds_dma(fr, (double2*)(f + 128 * c), 16 * (64 * c + lane));
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        }
        __syncthreads();
        // wavefront 0 solves the real parts, wavefront 1 the imaginary parts:
        // rows gathered through dgetrs' row interchanges (every gather ahead
        // of any store), swept, stored
        const int lane = tid & 63, im = (tid >> 6) & 1;
        double xr[3] = {0.0, 0.0, 0.0};
        if (tid < 128) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const int i = lane + 64 * q;
                const double2 v = (i < N) ? lload(x + gload(R.pT + i)) : mk(0.0, 0.0);
                xr[q] = im ? v.y : v.x;
            }
        }
        __syncthreads();
        if (tid < 128) {
            const LDS_AS double* fl = (const LDS_AS double*)f;
            auto col = [fl, lane, N](int p, int j) { return fl[lane + 64 * p + N * j]; };
            if (N <= 128) {
                double x2[2] = {xr[0], xr[1]};
                lu_solve_wave<2, 4>(x2, col, N);
                xr[0] = x2[0];
                xr[1] = x2[1];
            } else {
                lu_solve_wave<3, 4>(xr, col, N);
            }
            LDS_AS double* xd = (LDS_AS double*)x + im;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const int i = lane + 64 * q;
                if (i < N) xd[2 * i] = xr[q];
            }
        }
        __syncthreads();
    } else {
        __syncthreads();
        lu_solve<NT, SOLVE_NB>(x, xs, R.Ti, R.pT, N);
    }
    for (int i = tid; i < N; i += NT) gstore((double2*)R.sy + i, lload(x + i));
}

size_t m0_lds(const Shape& sh)
{
    const size_t N = (size_t)sh.N;
    return N <= (size_t)M0_STAGE_MAXN ? sizeof(double2) * N + sizeof(double) * (N * N + 192) : sizeof(double2) * 4 * N;
}

// ------------------------------------------------------------------ nonlinear
// Nr_maxLOOP for one radial level: T + i*omega and u_phi + i*u_r are each one
// complex inverse DFT (two real c2r transforms per DFT); the products
// u_phi*T + i*u_r*T and u_phi*omega + i*u_r*omega are each one forward DFT.
// TPB threads per workgroup (a tuning choice): every butterfly's arithmetic
// and the summation order of <ur omega> (256 fixed partial sums, point j in
// partial j mod 256, combined by one fixed tree) do not depend on it, so the
// results are the same bit for bit at every TPB.
template <int NPMAX, bool GEN, int TPB>
__global__ __launch_bounds__(TPB) void k_nonlin(const RunDesc* __restrict__ runs, const Shape sh, const int nrun)
{
    static_assert(TPB == 128 || TPB == 256, "k_nonlin takes 128 or 256 threads");
    constexpr int VL = 256 / TPB;     // partial sums per thread
    extern __shared__ double2 smem[];
    const int N = sh.N, Nm = sh.Nm, Nm1 = sh.Nm1, Np = sh.Np;
    int w;
    if (!map_work(nrun * N, w)) return;
    const int ri = w / N, r = w - ri * N;
    const RunDesc& R = runs[ri];
    double2* z1 = smem;
    double2* z2 = smem + Np;
    double* red = (double*)(smem + 2 * Np);
    const size_t o = (size_t)Nm1 * r;
    const int tid = threadIdx.x;
    constexpr bool FUSE = !GEN && NPMAX > 768;
    // Np <= 768 with a first factor 2-5: the first inverse pass reads the
    // packed spectra straight from the fields (no zero fill, no scatter)
    const bool first_in = NPMAX <= 768 && (sh.pp.code & 15ull) != 0;
    if (!first_in) {
    for (int j = tid; j < Np; j += TPB) { z1[j] = mk(0.0, 0.0); z2[j] = mk(0.0, 0.0); }
    __syncthreads();
    for (int m = tid; m <= Nm; m += TPB) {
        const double2 t = R.tFR[m + o], wv = R.omgFR[m + o], u = R.upFR[m + o], v = R.urFR[m + o];
        if (m == 0) {
            z1[0] = mk(t.x, wv.x);   // FFTW's c2r ignores Im f_0
            z2[0] = mk(u.x, v.x);
        } else {
            z1[m] = mk(t.x - wv.y, t.y + wv.x);
            z1[Np - m] = mk(t.x + wv.y, wv.x - t.y);
            z2[m] = mk(u.x - v.y, u.y + v.x);
            z2[Np - m] = mk(u.x + v.y, v.x - u.y);
        }
    }
    __syncthreads();
    }
    double s[VL], mu = 0.0, mv = 0.0;
#pragma unroll
    for (int q = 0; q < VL; ++q) s[q] = 0.0;
    // above Np = 768: plan pq, the last inverse pass forming the products in
    // registers (its outputs j of this thread in ascending r: the partial sum
    // of <ur omega>), one LDS sweep fewer; at Np <= 768 the separate products
    // sweep is faster (the fused pass's registers cost occupancy)
    if constexpr (FUSE) {
        fft2_nl<TPB, NPMAX>(z1, z2, sh.pq, +1, true, [&](int j, double2 a, double2 c) {
            const double T = a.x, W = a.y, U = c.x, V = c.y;
            lstore(z1 + j, mk(U * T, V * T));
            lstore(z2 + j, mk(U * W, V * W));
            s[0] = s[0] + V * W;
            mu = fmax(mu, fabs(U));
            mv = fmax(mv, fabs(V));
        });
    } else {
        if constexpr (NPMAX <= 768) {
            if (first_in) {
                // z1[j] = T + i omega, z2[j] = u_phi + i u_r of the packed
                // spectra: modes j <= Nm, their mirrors j >= Np - Nm, zero
                // between (FFTW's c2r ignores Im f_0)
                fft2_inl_in<TPB, NPMAX, GEN>(z1, z2, sh.pp, +1, [&](int j, double2& a, double2& c) {
                    const int m = (j <= Nm) ? j : Np - j;
                    if (j > Nm && j < Np - Nm) {
                        a = mk(0.0, 0.0);
                        c = mk(0.0, 0.0);
                        return;
                    }
                    const double2 t = R.tFR[m + o], wv = R.omgFR[m + o], u = R.upFR[m + o], v = R.urFR[m + o];
                    if (j == 0) {
                        a = mk(t.x, wv.x);
                        c = mk(u.x, v.x);
                    } else if (j <= Nm) {
                        a = mk(t.x - wv.y, t.y + wv.x);
                        c = mk(u.x - v.y, u.y + v.x);
                    } else {
                        a = mk(t.x + wv.y, wv.x - t.y);
                        c = mk(u.x + v.y, v.x - u.y);
                    }
                });
            } else {
                fft2_inl<TPB, NPMAX, GEN>(z1, z2, sh.pp, +1);
            }
        } else {
            fft2_inl<TPB, NPMAX, GEN>(z1, z2, sh.pp, +1);
        }
#pragma unroll
        for (int q = 0; q < VL; ++q)
            for (int j = tid + q * TPB; j < Np; j += 256) {
                const double T = z1[j].x, W = z1[j].y, U = z2[j].x, V = z2[j].y;
                z1[j] = mk(U * T, V * T);
                z2[j] = mk(U * W, V * W);
                s[q] = s[q] + V * W;
                mu = fmax(mu, fabs(U));
                mv = fmax(mv, fabs(V));
            }
    }
    // <ur omega>: the 64-lane sums of partials 0..63, 64..127, 128..191,
    // 192..255, then those four in order
#pragma unroll
    for (int q = 0; q < VL; ++q) {
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) s[q] = s[q] + __shfl_xor(s[q], d);
        if ((tid & 63) == 0) red[(tid + q * TPB) >> 6] = s[q];
    }
    const bool cfl = R.vardt != 0;   // one run per workgroup: uniform
    if (cfl) {
        // get_dtvals (nonlin.F90:93-106): the level's largest |u_r|, |u_phi|
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) {
            mu = fmax(mu, __shfl_xor(mu, d));
            mv = fmax(mv, __shfl_xor(mv, d));
        }
        if ((tid & 63) == 0) {
            red[4 + (tid >> 6)] = mu;
            red[8 + (tid >> 6)] = mv;
        }
    }
    __syncthreads();
    if (tid == 0) {
        R.urw[r] = (((red[0] + red[1]) + red[2]) + red[3]) * (1.0 / (double)Np);
        if (cfl) {
            double a = red[8], b = red[4];
            for (int k = 1; k < TPB / 64; ++k) { a = fmax(a, red[8 + k]); b = fmax(b, red[4 + k]); }
            R.cfl[2 * (size_t)r] = a;
            R.cfl[2 * (size_t)r + 1] = b;
        }
    }
    if constexpr (FUSE) fft2_nl<TPB, NPMAX>(z1, z2, sh.pq, -1, false, [](int, double2, double2) {});
    else fft2_inl<TPB, NPMAX, GEN>(z1, z2, sh.pp, -1);
    const double dNp = (double)Np;
    for (int m = tid; m <= Nm; m += TPB) {
        const int mq = (m == 0) ? 0 : Np - m;
        {
            const double2 P = z1[m], Qv = z1[mq];
            const double2 X = mk((P.x + Qv.x) * 0.5, (P.y - Qv.y) * 0.5);
            const double2 Y = mk((P.y + Qv.y) * 0.5, (Qv.x - P.x) * 0.5);
            R.upT[m + o] = mk(X.x / dNp, X.y / dNp);
            R.urT[m + o] = mk(Y.x / dNp, Y.y / dNp);
        }
        {
            const double2 P = z2[m], Qv = z2[mq];
            const double2 X = mk((P.x + Qv.x) * 0.5, (P.y - Qv.y) * 0.5);
            const double2 Y = mk((P.y + Qv.y) * 0.5, (Qv.x - P.x) * 0.5);
            R.upW[m + o] = mk(X.x / dNp, X.y / dNp);
            R.urW[m + o] = mk(Y.x / dNp, Y.y / dNp);
        }
    }
}

// forfft (fourier_fftw.F90:39-57) of every radial level of a physical field
// x(Np, Nr) (the reference's tt, column-major): f(m, r) = X_m / Np for
// m = 0..Nm with X the forward DFT of the level.  One workgroup per level;
// used once, for the initial perturbations of init_perturbation.
template <int NPMAX>
__global__ __launch_bounds__(NT) void k_forfft(const double* __restrict__ x, double2* __restrict__ f, const Shape sh)
{
    extern __shared__ double2 smem[];
    const int r = blockIdx.x, tid = threadIdx.x;
    const int Np = sh.Np, Nm1 = sh.Nm1;
    for (int j = tid; j < Np; j += NT) smem[j] = mk(gload(x + j + (size_t)Np * r), 0.0);
    __syncthreads();
    fft_inl<NT, NPMAX>(smem, sh.pp, -1);
    const double dNp = (double)Np;
    for (int m = tid; m < Nm1; m += NT) {
        const double2 X = smem[m];
        gstore(f + m + (size_t)Nm1 * r, mk(X.x / dNp, X.y / dNp));
    }
}

// ------------------------------------------------------------------ IMEX
// Radial derivative of the flux term: d/dr of the values given by src, i.e.
// chebinvtranD1(chebtransform(src)) (rhs_create.F90:79-85).
template <int NT, int MM, class F>
__device__ __forceinline__ void flux_dr(double2* buf, double2* c, double2* beta, double2* out, const Shape& sh, F src)
{
    cheb_fwd<NT, MM>(buf, c, sh, src);
    cheb_beta<NT, MM>(beta, sh, [&](int k) { return c[k]; });
    cheb_eval_beta<NT, MM>(buf, beta, sh, [&](int i, double2 v) { out[i] = v; });
}

// Shared set-up of the (run, mode) radial kernels: LDS twiddles, the work
// item, and the LDS vectors buf (M), x (2N) and V0..V5 (N each).
#define RADIAL_SETUP                                                      \
    extern __shared__ double2 smem[];                                     \
    const int N = sh0.N, Nm1 = sh0.Nm1;                                   \
    int w;                                                                \
    if (!map_work(nrun * Nm1, w)) return;                                 \
    const Shape sh = lds_twiddles<NT>(sh0, smem);                         \
    int m, ri;                                                            \
    split_mode_major(w, nrun, m, ri);                                     \
    const RunDesc& R = runs[ri];                                          \
    double2* buf = smem + sh.M;                                           \
    double2* x = buf + sh.M;                                              \
    double2* V0 = x + 2 * N;                                              \
    double2* V1 = V0 + N;                                                 \
    double2* V2 = V1 + N;                                                 \
    double2* V3 = V2 + N;                                                 \
    double2* V4 = V3 + N;                                                 \
    double2* V5 = V4 + N;                                                 \
    const GPtr<const double> rr = {sh.rr};                                \
    const GPtr<const double> rr2 = {sh.rr2};                              \
    const double mm = (double)m;                                          \
    const size_t S = (size_t)Nm1 * N;                                     \
    const int tid = threadIdx.x;                                          \
    const size_t xo = (size_t)m * 2 * N; /* this mode's solve area */     \
    auto at = [&](int i) { return (size_t)m + (size_t)Nm1 * (size_t)i; }; \
    (void)V1; (void)V2; (void)V3; (void)V4; (void)V5; (void)rr2; (void)S; (void)mm

// First IMEX step of a fresh or restarted run: temp_spec, omg_spec and the
// mean-flow upFC(1,:) by chebtransform of the physical fields
// (fourierloop.F90:78-84), and tFR, omgFR replaced by chebinvtran of those
// coefficients, which is what the right-hand sides read as the current
// fields (rhs_create.F90:81, 201); the first step then runs as any other.
template <int NT, int MM>
__global__ __launch_bounds__(NT) void k_imex_prep(const RunDesc* __restrict__ runs, const Shape sh0, const int nrun)
{
    RADIAL_SETUP;
    cheb_fwd<NT, MM>(buf, V0, sh, [&](int i) { return R.tFR[at(i)]; });
    for (int i = tid; i < N; i += NT) R.tspec[at(i)] = V0[i];
    cheb_inv<NT, MM>(buf, sh, [&](int k) { return V0[k]; }, [&](int i, double2 v) { R.tFR[at(i)] = v; });
    if (m == 0) {
        cheb_fwd<NT, MM>(buf, V2, sh, [&](int i) { return R.upFR[at(i)]; });
        for (int i = tid; i < N; i += NT) R.wspec[at(i)] = V2[i];   // upFC(1,:)
    } else {
        cheb_fwd<NT, MM>(buf, V2, sh, [&](int i) { return R.omgFR[at(i)]; });
        for (int i = tid; i < N; i += NT) R.wspec[at(i)] = V2[i];
        cheb_inv<NT, MM>(buf, sh, [&](int k) { return V2[k]; }, [&](int i, double2 v) { R.omgFR[at(i)] = v; });
    }
}

// rhs_temp_from_restart / rhs_vort_from_restart / rhs_uphibar_from_restart
// (rhs_create.F90:387-573) of the current state, after a nonlinear pass over
// it.  ki, ke >= 0: fill history slot ki of the implicit arrays and ke of the
// explicit ones, counted back from the slot the first step writes (ring slot
// 0, as head = L - 1 before the first step): the reference's index
// n_order - j for its j-th checkpoint (steptime.F90:88-186).  ki < 0: fill
// every slot (identical checkpoints), except for runs whose histories were
// already filled from checkpoint states.
template <int NT, int MM>
__global__ __launch_bounds__(NT) void k_imex_restart(const RunDesc* __restrict__ runs, const Shape sh0, const int nrun,
                                                     const int ki, const int ke)
{
    extern __shared__ double2 smem[];
    const int N = sh0.N, Nm1 = sh0.Nm1;
    int w;
    if (!map_work(nrun * Nm1, w)) return;
    int m, ri;
    split_mode_major(w, nrun, m, ri);
    const RunDesc& R = runs[ri];
    if (ki < 0 && R.hist_ready) return;              // one work item per block: the whole block leaves
    const Shape sh = lds_twiddles<NT>(sh0, smem);     // twiddle table: smem[0..M)
    double2* buf = smem + sh.M;
    double2* V0 = buf + sh.M + 4 * N;
    double2* V1 = V0 + N;
    double2* V2 = V1 + N;
    double2* V3 = V2 + N;
    const GPtr<const double> rr = {sh.rr};
    const double mm = (double)m;
    const size_t S = (size_t)Nm1 * N;
    const int tid = threadIdx.x;
    auto at = [&](int i) { return (size_t)m + (size_t)Nm1 * (size_t)i; };
    flux_dr<NT, MM>(buf, V0, V1, V2, sh, [&](int i) { return R.urT[at(i)]; });
    flux_dr<NT, MM>(buf, V0, V1, V3, sh, [&](int i) { return R.urW[at(i)]; });
    const int L = R.L;
    const int i0 = ki < 0 ? 0 : ki, i1 = ki < 0 ? L : ki + 1;
    const int e0 = ke < 0 ? 0 : ke, e1 = ke < 0 ? L : ke + 1;
    for (int i = tid; i < N; i += NT) {
        const double2 NT_ = cadd(cadd(cimr(mm, rr[i], R.upT[at(i)]), V2[i]), csc(rr[i], R.urT[at(i)]));
        const double2 NW = cadd(cadd(cimr(mm, rr[i], R.upW[at(i)]), V3[i]), csc(rr[i], R.urW[at(i)]));
        for (int k = i0; k < i1; ++k) {
            const size_t o = (size_t)((L - k) % L);
            R.hiT[o * S + at(i)] = R.tFR[at(i)];
            R.hiW[o * S + at(i)] = R.omgFR[at(i)];
            if (m == 0) R.hiU[o * N + i] = R.upFR[at(i)];
        }
        for (int k = e0; k < e1; ++k) {
            const size_t o = (size_t)((L - k) % L);
            R.heT[o * S + at(i)] = NT_;
            R.heW[o * S + at(i)] = NW;
            if (m == 0) R.heU[o * N + i] = mk(R.urw[i], 0.0);
        }
    }
}

// ------------------------------------------------------------------ IMEXRK
// The stage-0 terms of the current state before the first step (or after a
// new state): the buoyancy and the diffusion terms rhs_construct_temp_d /
// uphibar_d / vort_d (rhs_create_imexrk.F90:59-93, 162-191, 244-289) from the
// physical fields.
template <int NT, int MM>
__global__ __launch_bounds__(NT) void k_rk_init(const RunDesc* __restrict__ runs, const Shape sh0, const int nrun)
{
    RADIAL_SETUP;
    const double Pr = R.Pr;
    const double raPr = R.Ra / R.Pr;
    const bool fd = R.need_fd[0] != 0;
    for (int i = tid; i < N; i += NT) V0[i] = R.tFR[at(i)];
    __syncthreads();
    if (m == 0) {
        for (int i = tid; i < N; i += NT) V1[i] = R.upFR[at(i)];
    } else {
        for (int i = tid; i < N; i += NT) {
            V1[i] = R.omgFR[at(i)];
            R.Fb[at(i)] = csc(-raPr, cimr(mm, rr[i], V0[i]));
        }
    }
    __syncthreads();
    if (!fd) return;
    cheb_fwd<NT, MM>(buf, V2, sh, [&](int i) { return V0[i]; });
    cheb_beta<NT, MM>(V4, sh, [&](int k) { return V2[k]; });
    cheb_beta<NT, MM>(V5, sh, [&](int k) { return V4[k]; });
    cheb_eval_beta<NT, MM>(buf, V4, sh, [&](int i, double2 v) { x[i] = v; });
    cheb_eval_beta<NT, MM>(buf, V5, sh, [&](int i, double2 v) { x[N + i] = v; });
    for (int i = tid; i < N; i += NT) {
        const double2 lap = csub(cadd(csc(rr[i], x[i]), x[N + i]), csc((mm * mm) * rr2[i], V0[i]));
        R.FdT[at(i)] = csc(1.0 / Pr, lap);
    }
    __syncthreads();
    cheb_fwd<NT, MM>(buf, V2, sh, [&](int i) { return V1[i]; });
    cheb_beta<NT, MM>(V4, sh, [&](int k) { return V2[k]; });
    cheb_beta<NT, MM>(V5, sh, [&](int k) { return V4[k]; });
    cheb_eval_beta<NT, MM>(buf, V4, sh, [&](int i, double2 v) { x[i] = v; });
    cheb_eval_beta<NT, MM>(buf, V5, sh, [&](int i, double2 v) { x[N + i] = v; });
    for (int i = tid; i < N; i += NT) {
        double2 f;
        if (m == 0) {
            const double2 v = cadd(csub(x[N + i], csc(rr2[i], V1[i])), csc(rr[i], x[i]));
            f = mk(v.x, 0.0);
        } else {
            f = csub(cadd(csc(rr[i], x[i]), x[N + i]), csc((mm * mm) * rr2[i], V1[i]));
        }
        R.FdW[at(i)] = f;
    }
}


// Assembly_stage (fourierloop_imexrk.F90:499-651), the end of a step of an
// IMEXRK scheme that is not stiffly accurate, for one (run, mode): the new
// state is the stage start plus dt times the b-weighted stage terms (those of
// the last stage computed here: its diffusion terms were stored by the stage
// kernels, its explicit terms come from the nonlinear pass over the last stage
// when its column is used).  Mode 0: the mean flow and its vorticity.  Modes
// m >= 1: omega's interior, psi from LAPpsi psi = -omega (psi = 0 at the
// walls), u_phi = -dpsi/dr, u_r = i m psi / r, and omega at the walls from the
// Lagrange derivative weights applied to psi's interior values (Johnston's
// strategy: -(d2psi - 2 dw_1 dpsi) no slip, -2 dpsi / r stress free).  The
// stage start values are physical (temp_spec / omg_spec / uphi_bar_spec).
template <int NT, int MM>
__global__ __launch_bounds__(NT) void k_rk_asm(const RunDesc* __restrict__ runs, const Shape sh0, const int nrun)
{
    RADIAL_SETUP;
    const int SS = R.S;
    const bool fa = R.fa_col[SS - 1] != 0;
    const double dt = R.dt;
    if (fa) {
        // rhs_construct_temp_a / vort_a of the last stage: d(ur T)/dr, d(ur omega)/dr
        flux_dr<NT, MM>(buf, V0, V1, V2, sh, [&](int i) { return R.urT[at(i)]; });
        __syncthreads();
        if (m > 0) flux_dr<NT, MM>(buf, V0, V1, V3, sh, [&](int i) { return R.urW[at(i)]; });
        __syncthreads();
    }
    // temperature
    for (int i = tid; i < N; i += NT) {
        double2 F = mk(0.0, 0.0);
        for (int k = 0; k < SS; ++k) {
            double2 a;
            if (k == SS - 1 && fa)
                a = cneg(cadd(cadd(cimr(mm, rr[i], R.upT[at(i)]), V2[i]), csc(rr[i], R.urT[at(i)])));
            else
                a = R.FaT[(size_t)k * S + at(i)];
            F = cadd(cadd(F, csc(R.bA[k], a)), csc(R.bD[k], R.FdT[(size_t)k * S + at(i)]));
        }
        double2 y = cadd(R.tspec[at(i)], csc(dt, F));
        if (i == 0) y = mk(m == 0 ? 1.0 : 0.0, 0.0);
        if (i == N - 1) y = mk(0.0, 0.0);
        V4[i] = y;
        R.tFR[at(i)] = y;
        R.tspec[at(i)] = y;
    }
    __syncthreads();
    if (m == 0) {
        // TFC = chebtransform(tFR(1,:)) for the Nusselt number (steptime_imexrk.F90:190)
        cheb_fwd<NT, MM>(buf, V0, sh, [&](int i) { return V4[i]; });
        for (int i = tid; i < N; i += NT) R.TFC[i] = V0[i];
        __syncthreads();
        // the mean flow (uphi_bar_spec)
        for (int i = tid; i < N; i += NT) {
            double2 F = mk(0.0, 0.0);
            for (int k = 0; k < SS; ++k) {
                const double2 a = (k == SS - 1 && fa) ? mk(-R.urw[i], 0.0) : R.FaW[(size_t)k * S + at(i)];
                F = cadd(cadd(F, csc(R.bA[k], a)), csc(R.bD[k], R.FdW[(size_t)k * S + at(i)]));
            }
            double2 u = cadd(R.ubar[i], csc(dt, F));
            if (i == 0 || i == N - 1) u = mk(0.0, 0.0);
            V5[i] = u;
            R.ubar[i] = u;
            R.upFR[at(i)] = u;
            R.urFR[at(i)] = mk(0.0, 0.0);
        }
        __syncthreads();
        // omega of the mean flow: u/r + du/dr
        cheb_fwd<NT, MM>(buf, V0, sh, [&](int i) { return V5[i]; });
        cheb_beta<NT, MM>(V1, sh, [&](int k) { return V0[k]; });
        cheb_eval_beta<NT, MM>(buf, V1, sh, [&](int i, double2 v) { V2[i] = v; });
        __syncthreads();
        for (int i = tid; i < N; i += NT) {
            const double2 w = cadd(csc(rr[i], V5[i]), V2[i]);
            R.omgFR[at(i)] = w;
            R.wspec[at(i)] = w;
        }
        return;
    }
    // vorticity interior, then -omega as the LAPpsi right-hand side in x
    for (int i = tid; i < N; i += NT) {
        double2 w = R.wspec[at(i)];
        if (i > 0 && i < N - 1) {
            double2 F = mk(0.0, 0.0);
            for (int k = 0; k < SS; ++k) {
                double2 a;
                if (k == SS - 1 && fa)
                    a = cneg(cadd(cadd(cimr(mm, rr[i], R.upW[at(i)]), V3[i]), csc(rr[i], R.urW[at(i)])));
                else
                    a = R.FaW[(size_t)k * S + at(i)];
                F = cadd(cadd(cadd(F, csc(R.bA[k], a)), csc(R.bD[k], R.FdW[(size_t)k * S + at(i)])),
                         csc(R.bD[k], R.Fb[(size_t)k * S + at(i)]));
            }
            w = cadd(w, csc(dt, F));
        }
        V5[i] = w;
        x[i] = (i == 0 || i == N - 1) ? mk(0.0, 0.0) : mk(-1.0 * w.x, -1.0 * w.y);
    }
    __syncthreads();
    lu_solve<NT, SOLVE_NB>(x, x + N, R.Li + (size_t)m * N * N, R.pL + (size_t)m * N, N);
    __syncthreads();
    // psi values (V0) and d/dr (V2) from its coefficients x
    cheb_inv<NT, MM>(buf, sh, [&](int k) { return x[k]; }, [&](int i, double2 v) { V0[i] = v; });
    cheb_beta<NT, MM>(V1, sh, [&](int k) { return x[k]; });
    cheb_eval_beta<NT, MM>(buf, V1, sh, [&](int i, double2 v) { V2[i] = v; });
    __syncthreads();
    if (tid == 0) {
        const int lm = R.lm;
        const GPtr<const double> lw = R.lagw;
        double2 d2min = mk(0.0, 0.0), dmin = mk(0.0, 0.0), d2max = mk(0.0, 0.0), dmax = mk(0.0, 0.0);
        for (int i = 0; i < lm; ++i) {
            d2min = cadd(d2min, csc(lw[LAGMAX + i], V0[1 + i]));
            dmin = cadd(dmin, csc(lw[i], V0[1 + i]));
        }
        for (int i = 0; i < lm; ++i) {
            d2max = cadd(d2max, csc(lw[3 * LAGMAX + i], V0[N - 2 - i]));
            dmax = cadd(dmax, csc(lw[2 * LAGMAX + i], V0[N - 2 - i]));
        }
        if (R.mbc == 0) {
            const double2 a = csub(d2min, csc(2.0 * lw[0], dmin));
            const double2 b = csub(d2max, csc(2.0 * lw[2 * LAGMAX], dmax));
            V5[0] = mk(-1.0 * a.x, -1.0 * a.y);
            V5[N - 1] = mk(-1.0 * b.x, -1.0 * b.y);
        } else {
            V5[0] = csc(lw[4 * LAGMAX], dmin);
            V5[N - 1] = csc(lw[4 * LAGMAX + 1], dmax);
        }
    }
    __syncthreads();
    for (int i = tid; i < N; i += NT) {
        R.omgFR[at(i)] = V5[i];
        R.wspec[at(i)] = V5[i];
        R.upFR[at(i)] = mk(-1.0 * V2[i].x, -1.0 * V2[i].y);
        R.urFR[at(i)] = cimr(mm, rr[i], V0[i]);
    }
}

// ------------------------------------------------------------------ RK path
// The explicit RK path (time_scheme_type = 'RK', steptime_rk.F90): every term
// of a stage's right-hand side is explicit; each stage (and the step's end)
// rebuilds omega, psi and the velocities as Assembly_stage does.  One
// workgroup per (run, mode), as k_rk_asm.

// First stage of a run: upFC is the state's u_phi values (init.F90:252, 395
// copy upFR into upFC before any transform), and d1u = chebinvtranD1 of it.
template <int NT, int MM>
__global__ __launch_bounds__(NT) void k_rkx_init(const RunDesc* __restrict__ runs, const Shape sh0, const int nrun)
{
    RADIAL_SETUP;
    for (int i = tid; i < N; i += NT) {
        const double2 u = R.upFR[at(i)];
        V0[i] = u;
        R.upC[at(i)] = u;
    }
    __syncthreads();
    cheb_beta<NT, MM>(V1, sh, [&](int k) { return V0[k]; });
    cheb_eval_beta<NT, MM>(buf, V1, sh, [&](int i, double2 v) { R.d1u[at(i)] = v; });
}

// Get_stage_var (s < S: weights a(s,:), fourierloop_rk.F90:69-224) or
// Assembly_stage (s == S: weights b, the new state, :285-424) for one
// (run, mode): the value is the step's start (tspec / wspec / ubar, physical)
// plus dt times the weighted stage right-hand sides.  Mode 0: the mean flow,
// its coefficients upFC, its derivative and omega = u/r + du/dr.  Modes
// m >= 1: LAPpsi psi = -omega, u_phi = -dpsi/dr, u_r = i m psi / r, omega at
// the walls (Johnston / stress free), upFC = chebtransform(u_phi) and its
// derivative d1u for the mean-flow right-hand side.
template <int NT, int MM>
__global__ __launch_bounds__(NT) void k_rkx_update(const RunDesc* __restrict__ runs, const Shape sh0, const int nrun,
                                                   const int s)
{
    RADIAL_SETUP;
    const int SS = R.S;
    const bool fin = (s == SS);
    const int K = fin ? SS : s;
    const double* wk = fin ? R.bA : R.aA + s * MAXS;
    const double dt = R.dt;
    // temperature
    for (int i = tid; i < N; i += NT) {
        double2 F = mk(0.0, 0.0);
        for (int k = 0; k < K; ++k) F = cadd(F, csc(wk[k], R.FdT[(size_t)k * S + at(i)]));
        double2 y = cadd(R.tspec[at(i)], csc(dt, F));
        if (i == 0) y = mk(m == 0 ? 1.0 : 0.0, 0.0);
        if (i == N - 1) y = mk(0.0, 0.0);
        V4[i] = y;
        R.tFR[at(i)] = y;
        if (fin) R.tspec[at(i)] = y;
    }
    __syncthreads();
    if (m == 0) {
        if (fin) {
            // TFC = chebtransform(tFR(1,:)) for the Nusselt number (steptime_rk.F90:138)
            cheb_fwd<NT, MM>(buf, V0, sh, [&](int i) { return V4[i]; });
            for (int i = tid; i < N; i += NT) R.TFC[i] = V0[i];
            __syncthreads();
        }
        for (int i = tid; i < N; i += NT) {
            double2 F = mk(0.0, 0.0);
            for (int k = 0; k < K; ++k) F = cadd(F, csc(wk[k], R.FdW[(size_t)k * S + at(i)]));
            double2 u = cadd(R.ubar[i], csc(dt, F));
            if (i == 0 || i == N - 1) u = mk(0.0, 0.0);
            V5[i] = u;
            if (fin) R.ubar[i] = u;
            R.upFR[at(i)] = u;
            R.urFR[at(i)] = mk(0.0, 0.0);
        }
        __syncthreads();
        cheb_fwd<NT, MM>(buf, V0, sh, [&](int i) { return V5[i]; });
        for (int i = tid; i < N; i += NT) R.upC[at(i)] = V0[i];
        cheb_beta<NT, MM>(V1, sh, [&](int k) { return V0[k]; });
        cheb_eval_beta<NT, MM>(buf, V1, sh, [&](int i, double2 v) { V2[i] = v; });
        for (int i = tid; i < N; i += NT) {
            const double2 w = cadd(csc(rr[i], V5[i]), V2[i]);
            R.d1u[at(i)] = V2[i];
            R.omgFR[at(i)] = w;
            if (fin) R.wspec[at(i)] = w;
        }
        return;
    }
    // omega, then -omega as the LAPpsi right-hand side in x (psi = 0 rows)
    for (int i = tid; i < N; i += NT) {
        double2 F = mk(0.0, 0.0);
        for (int k = 0; k < K; ++k) F = cadd(F, csc(wk[k], R.FdW[(size_t)k * S + at(i)]));
        const double2 w = cadd(R.wspec[at(i)], csc(dt, F));
        V5[i] = w;
        x[i] = (i == 0 || i == N - 1) ? mk(0.0, 0.0) : mk(-1.0 * w.x, -1.0 * w.y);
    }
    __syncthreads();
    lu_solve<NT, SOLVE_NB>(x, x + N, R.Li + (size_t)m * N * N, R.pL + (size_t)m * N, N);
    __syncthreads();
    cheb_inv<NT, MM>(buf, sh, [&](int k) { return x[k]; }, [&](int i, double2 v) { V0[i] = v; });
    cheb_beta<NT, MM>(V1, sh, [&](int k) { return x[k]; });
    cheb_eval_beta<NT, MM>(buf, V1, sh, [&](int i, double2 v) { V2[i] = mk(-1.0 * v.x, -1.0 * v.y); });
    if (tid == 0) {
        const int lm = R.lm;
        const GPtr<const double> lw = R.lagw;
        double2 d2min = mk(0.0, 0.0), dmin = mk(0.0, 0.0), d2max = mk(0.0, 0.0), dmax = mk(0.0, 0.0);
        for (int i = 0; i < lm; ++i) {
            d2min = cadd(d2min, csc(lw[LAGMAX + i], V0[1 + i]));
            dmin = cadd(dmin, csc(lw[i], V0[1 + i]));
        }
        for (int i = 0; i < lm; ++i) {
            d2max = cadd(d2max, csc(lw[3 * LAGMAX + i], V0[N - 2 - i]));
            dmax = cadd(dmax, csc(lw[2 * LAGMAX + i], V0[N - 2 - i]));
        }
        if (R.mbc == 0) {
            const double2 a = csub(d2min, csc(2.0 * lw[0], dmin));
            const double2 b = csub(d2max, csc(2.0 * lw[2 * LAGMAX], dmax));
            V5[0] = mk(-1.0 * a.x, -1.0 * a.y);
            V5[N - 1] = mk(-1.0 * b.x, -1.0 * b.y);
        } else {
            V5[0] = csc(lw[4 * LAGMAX], dmin);
            V5[N - 1] = csc(lw[4 * LAGMAX + 1], dmax);
        }
    }
    __syncthreads();
    for (int i = tid; i < N; i += NT) {
        R.omgFR[at(i)] = V5[i];
        if (fin) R.wspec[at(i)] = V5[i];
        R.psi[at(i)] = V0[i];
        R.upFR[at(i)] = V2[i];
        R.urFR[at(i)] = cimr(mm, rr[i], V0[i]);
    }
    // upFC = chebtransform(u_phi) (fourierloop_rk.F90:210, 411) and d/dr of it
    cheb_fwd<NT, MM>(buf, V3, sh, [&](int i) { return V2[i]; });
    for (int i = tid; i < N; i += NT) R.upC[at(i)] = V3[i];
    cheb_beta<NT, MM>(V1, sh, [&](int k) { return V3[k]; });
    cheb_eval_beta<NT, MM>(buf, V1, sh, [&](int i, double2 v) { R.d1u[at(i)] = v; });
}

// RHS_construct_stage of the RK path for one (run, mode) (fourierloop_rk.F90:
// 226-283): the whole right-hand side of stage s into FdT / FdW slot s.
// Temperature: (1/Pr) lap T - (i m/r up_T + d(ur_T)/dr + ur_T/r)
// (rhs_create_rk.F90:19-67).  Mode 0: the mean flow, Re(d2u - u/r^2 + du/dr
// / r) from upFC(1,:), minus the sums over modes of 2 Re(ur_m conj(du_m)) +
// 2 Re(ur_m conj(u_m)) / r (Re(ur_0 x_0) at m = 0), u = u_phi, du its
// derivative from upFC (:131-224).  Modes m >= 1: lap omega - the advection
// - (Ra/Pr) i m T / r (:69-129).
template <int NT, int MM>
__global__ __launch_bounds__(NT) void k_rkx_rhs(const RunDesc* __restrict__ runs, const Shape sh0, const int nrun,
                                                const int s)
{
    RADIAL_SETUP;
    const double Pr = R.Pr;
    for (int i = tid; i < N; i += NT) V0[i] = R.tFR[at(i)];
    flux_dr<NT, MM>(buf, V3, V4, V2, sh, [&](int i) { return R.urT[at(i)]; });
    cheb_fwd<NT, MM>(buf, V1, sh, [&](int i) { return V0[i]; });
    cheb_beta<NT, MM>(V4, sh, [&](int k) { return V1[k]; });
    cheb_beta<NT, MM>(V5, sh, [&](int k) { return V4[k]; });
    cheb_eval_beta<NT, MM>(buf, V4, sh, [&](int i, double2 v) { x[i] = v; });
    cheb_eval_beta<NT, MM>(buf, V5, sh, [&](int i, double2 v) { x[N + i] = v; });
    for (int i = tid; i < N; i += NT) {
        const double2 lap = csub(cadd(csc(rr[i], x[i]), x[N + i]), csc((mm * mm) * rr2[i], V0[i]));
        const double2 adv = cadd(cadd(cimr(mm, rr[i], R.upT[at(i)]), V2[i]), csc(rr[i], R.urT[at(i)]));
        R.FdT[(size_t)s * S + at(i)] = csub(csc(1.0 / Pr, lap), adv);
    }
    __syncthreads();
    if (m == 0) {
        cheb_beta<NT, MM>(V4, sh, [&](int k) { return R.upC[at(k)]; });
        cheb_beta<NT, MM>(V5, sh, [&](int k) { return V4[k]; });
        cheb_eval_beta<NT, MM>(buf, V4, sh, [&](int i, double2 v) { x[i] = v; });
        cheb_eval_beta<NT, MM>(buf, V5, sh, [&](int i, double2 v) { x[N + i] = v; });
        for (int i = tid; i < N; i += NT) {
            double su = 0.0, sd = 0.0;
            for (int q = 0; q < Nm1; ++q) {
                const size_t o = (size_t)q + (size_t)Nm1 * (size_t)i;
                const double2 ur = R.urFR[o], up = R.upFR[o], du = R.d1u[o];
                if (q == 0) {
                    su = su + cmul(ur, up).x;
                    sd = sd + cmul(ur, du).x;
                } else {
                    su = su + cadd(cmul(ur, mk(up.x, -up.y)), cmul(mk(ur.x, -ur.y), up)).x;
                    sd = sd + cadd(cmul(ur, mk(du.x, -du.y)), cmul(mk(ur.x, -ur.y), du)).x;
                }
            }
            const double2 u = R.upFR[at(i)];
            const double2 ud = cadd(csub(x[N + i], csc(rr2[i], u)), csc(rr[i], x[i]));
            R.FdW[(size_t)s * S + at(i)] = mk(ud.x - (sd + rr[i] * su), 0.0);
        }
        return;
    }
    for (int i = tid; i < N; i += NT) V1[i] = R.omgFR[at(i)];
    __syncthreads();
    flux_dr<NT, MM>(buf, V3, V4, V2, sh, [&](int i) { return R.urW[at(i)]; });
    cheb_fwd<NT, MM>(buf, V3, sh, [&](int i) { return V1[i]; });
    cheb_beta<NT, MM>(V4, sh, [&](int k) { return V3[k]; });
    cheb_beta<NT, MM>(V5, sh, [&](int k) { return V4[k]; });
    cheb_eval_beta<NT, MM>(buf, V4, sh, [&](int i, double2 v) { x[i] = v; });
    cheb_eval_beta<NT, MM>(buf, V5, sh, [&](int i, double2 v) { x[N + i] = v; });
    const double raPr = R.Ra / Pr;
    for (int i = tid; i < N; i += NT) {
        const double2 lap = csub(cadd(csc(rr[i], x[i]), x[N + i]), csc((mm * mm) * rr2[i], V1[i]));
        const double2 adv = cadd(cadd(cimr(mm, rr[i], R.upW[at(i)]), V2[i]), csc(rr[i], R.urW[at(i)]));
        const double2 buo = csc(raPr, cimr(mm, rr[i], V0[i]));
        R.FdW[(size_t)s * S + at(i)] = csub(csub(lap, adv), buo);
    }
}

// mat_build_rk (mat_assembly.F90:415-462): LAPpsi of modes m0 + blockIdx.y,
// the Laplacian's psi rows with psi = 0 at both walls (the same rows for no
// slip and stress free); one column per workgroup.
__global__ __launch_bounds__(NT) void k_asm_LAP(double* __restrict__ A, const double* __restrict__ t,
                                                const double* __restrict__ D, const double* __restrict__ D2,
                                                const Shape sh, const int m0)
{
#pragma clang fp contract(off)
    const int N = sh.N;
    const int j = blockIdx.x;
    const double nn = (double)(m0 + (int)blockIdx.y);
    const double C = sqrt(2.0 / (double)(N - 1));
    const GPtr<const double> rr = {sh.rr};
    const GPtr<const double> rr2 = {sh.rr2};
    auto IX = [&](int i, int jj) { return (size_t)i + (size_t)N * (size_t)jj; };
    const bool halfj = (j == 0 || j == N - 1);
    double* Al = A + (size_t)blockIdx.y * N * N;
    for (int i = threadIdx.x; i < N; i += NT) {
        double v = (rr[i] * D[IX(i, j)] + D2[IX(i, j)]) - ((nn * nn) * rr2[i]) * t[IX(i, j)];
        if (i == 0 || i == N - 1) v = t[IX(i, j)];
        if (halfj) v = 0.5 * v;
        Al[IX(i, j)] = C * v;
    }
}

// ------------------------------------------------------------------ column tiles
// The per-mode radial work of a step (Nm_maxLOOP, fourierloop.F90:108-230, and
// Get_stage_var, fourierloop_imexrk.F90:169-393) as dense products over
// column tiles.  A tile is 16 consecutive Fourier modes of one run: its
// columns are the modes' complex radial vectors, the real and imaginary parts
// forming two 16-column blocks of one v_mfma_f64_16x16x4f64 step.  The
// operators are the shape's N x N Chebyshev matrices (chebinvtran E0,
// chebinvtranD1 E1, chebinvtranD2 E2, and the flux derivative G = E1 Fw of
// rhs_create.F90:79-85), read from global memory (every tile of every run of
// the shape shares them); the right-hand vectors are read straight from the
// run's fields.  Wavefront wv of the 64 TW-thread workgroup takes the 16-row
// strips wv, wv + TW, ...; lane (lk = l/16, li = l%16) accumulates rows
// i_r = 16 st + 4 r + lk (r = 0..3) of mode m = 16 mb + li and finishes those
// outputs elementwise: the right-hand sides, the new fields and the stage
// terms, written as the reference writes them.  Each epilogue first loads
// every input of its four rows (history slots outermost), then computes, then
// stores, so that its loads are in flight together.

// One right-hand block of a column-tile product: resource, byte offset of this
// lane's column at row 0, byte stride between rows, byte stride between
// columns (16 for a field, where the 16 modes of a block are adjacent; 32 N
// for the mode-major solve area).
struct ColIn {
    BRes r;
    int o, st, cs;
};

// acc[2q + t] += Op_q X_{b(q)} for the P0 + P1 + P2 operators in ops (N x N,
// column-major), the first P0 applied to block x0, the next P1 to x1, the last
// P2 to x2 (t = 0 real, 1 imaginary parts).  A operand rows
// arow = 16 st + li.  KC k-steps of every operand are loaded ahead of the
// products.  Rows k >= N lie past the end of every resource and load as zero;
// A rows >= N and the columns of modes past Nm reach only outputs that are
// discarded.  Each output is the sum over k in ascending steps of four,
// independent of the other columns.
template <int P0, int P1, int P2, int KCMAX = 4>
__device__ __forceinline__ void col_pass(const BRes* ops, const int N, const int arow, const ColIn x0, const ColIn x1,
                                         const ColIn x2, dbl4* acc)
{
    constexpr int NP = P0 + P1 + P2;
    constexpr int NX = (P0 > 0) + (P1 > 0) + (P2 > 0);
    constexpr int KC = (NP + 2 * NX <= 6 && KCMAX >= 4) ? 4 : 2;
    const int lk = (threadIdx.x & 63) >> 4;
    const int nk = (N + 3) >> 2;
    const int ao = 8 * (arow + N * lk), astep = 32 * N;         // bytes
    const BRes xs[3] = {x0.r, x1.r, x2.r};
    const int bo[3] = {x0.o + x0.st * lk, x1.o + x1.st * lk, x2.o + x2.st * lk};
    const int bstep[3] = {4 * x0.st, 4 * x1.st, 4 * x2.st};
    double a0[KC][NP], a1[KC][NP];
    double2 b0[KC][3], b1[KC][3];
    auto load = [&](double (&a)[KC][NP], double2 (&b)[KC][3], int k0) {
#pragma unroll
        for (int c = 0; c < KC; ++c) {
#pragma unroll
            for (int q = 0; q < NP; ++q) a[c][q] = bload(ops[q], ao + astep * (k0 + c));
            if (P0 > 0) b[c][0] = bload2(xs[0], bo[0] + bstep[0] * (k0 + c));
            if (P1 > 0) b[c][1] = bload2(xs[1], bo[1] + bstep[1] * (k0 + c));
            if (P2 > 0) b[c][2] = bload2(xs[2], bo[2] + bstep[2] * (k0 + c));
        }
    };
    auto mac = [&](const double (&a)[KC][NP], const double2 (&b)[KC][3], int k0) {
#pragma unroll
        for (int c = 0; c < KC; ++c) {
            if (k0 + c < nk) {
#pragma unroll
                for (int q = 0; q < NP; ++q) {
                    const int x = q < P0 ? 0 : (q < P0 + P1 ? 1 : 2);
                    acc[2 * q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[c][q], b[c][x].x, acc[2 * q], 0, 0, 0);
                    acc[2 * q + 1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[c][q], b[c][x].y, acc[2 * q + 1], 0, 0, 0);
                }
            }
        }
    };
    load(a0, b0, 0);
    for (int k0 = 0; k0 < nk; k0 += 2 * KC) {
        load(a1, b1, k0 + KC);
        mac(a0, b0, k0);
        load(a0, b0, k0 + 2 * KC);
        mac(a1, b1, k0 + KC);
    }
}

// Staged right-hand blocks: the workgroup loads the blocks a column-tile
// product reads -- K4 = 4 ceil(N/4) rows x 16 columns each, rows >= N zero --
// into LDS once, [block][k][col] with a row pitch of 17 so that row- and
// column-order accesses are free of bank conflicts, and every wavefront then
// reads its B operands there.  All the workgroup's loads are in flight
// together instead of each wavefront's few k-steps ahead.
constexpr int STG_PITCH = 17;
__host__ __device__ constexpr int stg_blocks(int P0, int P1, int P2) { return (P0 > 0) + (P1 > 0) + (P2 > 0); }

template <int P0, int P1, int P2, int NTH>
__device__ __forceinline__ void stage_cols(double2* bs, const int N, const int K4, const ColIn x0, const ColIn x1,
                                           const ColIn x2)
{
    const int li = threadIdx.x & 15, tid = threadIdx.x;
    const ColIn xs[3] = {x0, x1, x2};
    const bool use[3] = {P0 > 0, P1 > 0, P2 > 0};
    int slot = 0;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        if (!use[x]) continue;
        const ColIn X = xs[x];
        const int base = X.o - X.cs * li;       // column 0, row 0 of the block
        double2* dst = bs + (size_t)slot * K4 * STG_PITCH;
        const int n = K4 * 16;
        const bool field = X.cs == 16;          // columns adjacent: columns fastest; else rows fastest
        for (int e0 = tid; e0 < n; e0 += 4 * NTH) {
            double2 v[4];
            int at[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = e0 + u * NTH;
                const int k = field ? (e >> 4) : (e % K4), c = field ? (e & 15) : (e / K4);
                at[u] = e < n ? k * STG_PITCH + c : -1;
                v[u] = (e < n && k < N) ? bload2(X.r, base + X.st * k + X.cs * c) : mk(0.0, 0.0);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (at[u] >= 0) lstore(dst + at[u], v[u]);
        }
        ++slot;
    }
    __syncthreads();
}

// col_pass with the B operands from the staged blocks (stage_cols): the same
// products, each output summed over k in the same ascending steps of four.
template <int P0, int P1, int P2, int KCMAX = 4>
__device__ __forceinline__ void col_pass_lds(const BRes* ops, const int N, const int K4, const int arow,
                                             const double2* bs, dbl4* acc)
{
    constexpr int NP = P0 + P1 + P2;
    constexpr int NX = stg_blocks(P0, P1, P2);
    constexpr int KC = (NP <= 3 && KCMAX >= 4) ? 4 : 2;
    const int lane = threadIdx.x & 63, lk = lane >> 4, li = lane & 15;
    const int nk = (N + 3) >> 2;
    const int ao = 8 * (arow + N * lk), astep = 32 * N;
    const int bo = lk * STG_PITCH + li, bstep = 4 * STG_PITCH, bslot = K4 * STG_PITCH;
    double a0[KC][NP], a1[KC][NP];
    double2 b0[KC][NX], b1[KC][NX];
    auto load = [&](double (&a)[KC][NP], double2 (&b)[KC][NX], int k0) {
#pragma unroll
        for (int c = 0; c < KC; ++c) {
#pragma unroll
            for (int q = 0; q < NP; ++q) a[c][q] = bload(ops[q], ao + astep * (k0 + c));
            const int kk = k0 + c < nk ? k0 + c : nk - 1;   // stay inside the staged rows
#pragma unroll
            for (int x = 0; x < NX; ++x) b[c][x] = lload(bs + x * bslot + bo + bstep * kk);
        }
    };
    auto mac = [&](const double (&a)[KC][NP], const double2 (&b)[KC][NX], int k0) {
#pragma unroll
        for (int c = 0; c < KC; ++c) {
            if (k0 + c < nk) {
#pragma unroll
                for (int q = 0; q < NP; ++q) {
                    const int x = q < P0 ? 0 : (q < P0 + P1 ? (P0 > 0 ? 1 : 0) : NX - 1);
                    acc[2 * q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[c][q], b[c][x].x, acc[2 * q], 0, 0, 0);
                    acc[2 * q + 1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[c][q], b[c][x].y, acc[2 * q + 1], 0, 0, 0);
                }
            }
        }
    };
    load(a0, b0, 0);
    for (int k0 = 0; k0 < nk; k0 += 2 * KC) {
        load(a1, b1, k0 + KC);
        mac(a0, b0, k0);
        load(a0, b0, k0 + 2 * KC);
        mac(a1, b1, k0 + KC);
    }
}

template <int NA>
__device__ __forceinline__ void acc_zero(dbl4 (&a)[NA])
{
#pragma unroll
    for (int q = 0; q < NA; ++q) a[q] = dbl4{0.0, 0.0, 0.0, 0.0};
}

// complex value r of product q
#define ACC(a, q, r) mk((a)[2 * (q)][r], (a)[2 * (q) + 1][r])
#define UNROLL _Pragma("unroll")

// Sum over history slots or stages k0 <= k < n (n runtime): the loads of G
// slots are issued before any of them is used, then op(k, values of the
// lane's four rows) folds them in ascending k -- the order of the plain loop,
// so the result is bitwise the same, with G loads per row in flight instead of
// one.
template <int G, class Ld, class Op>
__device__ __forceinline__ void slot_sum(const int k0, const int n, Ld ld, Op op)
{
    for (int kb = k0; kb < n; kb += G) {
        double2 v[G][4];
        UNROLL for (int g = 0; g < G; ++g)
            if (kb + g < n) UNROLL for (int r = 0; r < 4; ++r) v[g][r] = ld(kb + g, r);
        UNROLL for (int g = 0; g < G; ++g)
            if (kb + g < n) op(kb + g, v[g]);
    }
}

// Wavefronts per workgroup of the column-tile kernels (TW, 4 or 8, a tuning
// choice): wavefront wv takes the 16-row strips wv, wv + TW, ... of the tile,
// so that a tile's strips run side by side; each output's arithmetic does not
// depend on TW.

// Workgroups per column tile: two (each with half of the tile's 16-row
// strips) when the kernel allows it, runs four wavefronts per workgroup and
// the launch has fewer than TILE_SPLIT_BELOW tiles, so that a small group's
// launch puts one strip on each wavefront instead of two; one otherwise
// (eight wavefronts already take one strip each).  Each strip's arithmetic
// is the same either way.
constexpr int TILE_SPLIT_BELOW = 1024;
constexpr int TILE_SPLIT_MASK = (1 << 2) | (1 << 3) | (1 << 5);   // k_imex_post, k_rk_pre, k_rk_post
__host__ __device__ inline int tile_split(bool allowed, int tiles) { return (allowed && tiles < TILE_SPLIT_BELOW) ? 2 : 1; }

// TILE_SETUP_S(SPLIT): SPLIT lets a small launch give each tile two
// workgroups (tile_split)
#define TILE_SETUP TILE_SETUP_S(false)
#define TILE_SETUP_S(SPLIT)                                                                            \
    const int N = sh.N, Nm1 = sh.Nm1;                                                                  \
    const int nb = (Nm1 + 15) >> 4;                                                                    \
    const int SP = tile_split((SPLIT) && TW == 4, nrun * nb);                                          \
    int w;                                                                                             \
    if (!map_work(nrun * nb * SP, w)) return;                                                          \
    const int ri = w / (nb * SP), wr_ = w - ri * nb * SP;                                              \
    const int mb = wr_ / SP, hf_ = wr_ - mb * SP;   /* a run's mode blocks adjacent (shared L2 lines) */ \
    const RunDesc& R = runs[ri];                                                                       \
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;                                          \
    const int li = lane & 15, lk = lane >> 4;                                                          \
    const int m = 16 * mb + li;                                                                        \
    const bool ok = m < Nm1;                                                                           \
    const double mm = (double)m;                                                                       \
    const size_t S = (size_t)Nm1 * N;                                                                  \
    const int nstrip = (N + 15) >> 4;                                                                  \
    const int sper_ = (nstrip + SP - 1) / SP;                                                          \
    const int st0 = hf_ * sper_, st1 = min(nstrip, st0 + sper_);                                       \
    const size_t fb = sizeof(double2) * S;   /* bytes of one (Nm+1, N) field */                        \
    auto fres = [&](const double2* f, int nslot = 1) { return bres(f, fb * (size_t)nslot); };         \
    const int xo = 16 * m, xst = 16 * Nm1;   /* this lane's column, row stride (bytes) */              \
    const int sst = (int)fb;                 /* bytes between history / stage slots */                 \
    const size_t ob = sizeof(double) * (size_t)N * N;                                                  \
    const BRes E0 = bres(sh.E0, ob), E1 = bres(sh.E1, ob), E2 = bres(sh.E2, ob), G = bres(sh.G, ob);   \
    /* right-hand sides (sx) and solutions (sy), mode-major: rows 0..2N-1 of mode m at +2N m */        \
    const BRes syR = fres((const double2*)R.sy, 2);                                                    \
    const int sxo = 32 * N * m;                                                                        \
    const ColIn sya = {syR, sxo, 16, 32 * N}, syb = {syR, sxo + 16 * N, 16, 32 * N};                   \
    const size_t sxe = (size_t)2 * N * m;   /* element index of row 0 of this mode */                 \
    auto fin = [&](BRes f) { return ColIn{f, xo, xst, 16}; };                                          \
    const BRes rrR = bres(sh.rr, 8 * (size_t)N), rr2R = bres(sh.rr2, 8 * (size_t)N);                   \
    extern __shared__ double2 stg_[];        /* staged right-hand blocks (STG) */                      \
    const int K4 = 4 * ((N + 3) >> 2);                                                                 \
    (void)mm; (void)sst; (void)E0; (void)E1; (void)E2; (void)G; (void)sya; (void)syb; (void)rr2R;      \
    (void)sxe; (void)fin; (void)stg_; (void)K4

// The B operands of a column pass: staged in LDS (STG) or loaded per wavefront.
#define COL_STAGE(P0, P1, P2, X0, X1, X2) \
    if (STG) stage_cols<P0, P1, P2, 64 * TW>(stg_, N, K4, X0, X1, X2)
#define COL_PASS(OPS, P0, P1, P2, X0, X1, X2) COL_PASS_KC(4, OPS, P0, P1, P2, X0, X1, X2)
// KC: at most KC k-steps of operands loaded ahead (4 or 2; fewer registers)
#define COL_PASS_KC(KC, OPS, P0, P1, P2, X0, X1, X2)                            \
    do {                                                                       \
        if (STG) col_pass_lds<P0, P1, P2, KC>(OPS, N, K4, arow, stg_, acc);    \
        else col_pass<P0, P1, P2, KC>(OPS, N, arow, X0, X1, X2, acc);          \
    } while (0)

// this lane's rows of strip st: index, validity, field byte offset; then
// 1/r, 1/r^2
#define STRIP_IDX(st)                                                  \
    int ir[4], fo[4];                                                  \
    bool vr[4];                                                        \
    UNROLL for (int r = 0; r < 4; ++r) {                               \
        ir[r] = 16 * (st) + 4 * r + lk;                                \
        vr[r] = ok && ir[r] < N;                                       \
        fo[r] = xo + xst * ir[r];                                      \
    }
#define STRIP_RR()                                                     \
    double rri[4], rr2i[4];                                            \
    UNROLL for (int r = 0; r < 4; ++r) {                               \
        rri[r] = bload(rrR, 8 * ir[r]);                                \
        rr2i[r] = bload(rr2R, 8 * ir[r]);                              \
    }                                                                  \
    (void)rr2i
#define STRIP_ROWS(st) \
    STRIP_IDX(st);     \
    STRIP_RR()

// Epilogue prefetch (PF): before a strip's column pass each wavefront issues
// LDS-direct buffer loads (buffer_load_dwordx4 ... lds) of the epilogue inputs
// that do not depend on the products -- NPF vectors x the lane's four rows,
// 16 B per lane, lane-ordered at pf + (4 f + r) * 64 -- so that their latency
// passes under the MFMA loop instead of following it.  pf is the wavefront's
// own region of the dynamic LDS, after the staged blocks.
#define PF_BASE(NX, NPF) \
    double2* const pf = stg_ + (STG ? (NX) * K4 * STG_PITCH : 0) + wv * ((NPF) * 256)
// the previous strip's reads of pf are done before it is written again, and
// the loads have landed in LDS before it is read (the compiler does not order
// LDS-direct loads against LDS reads by itself; the memory clobber keeps the
// reads below the wait)
#define PF_FENCE() \
    if (PF && st != st0 + wv) asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory")
#define PF_WAIT() \
    if (PF) asm volatile("s_waitcnt vmcnt(0)" ::: "memory")
#define PF_LOAD(f, RES, OFF) \
    if (PF) UNROLL for (int r = 0; r < 4; ++r) lds_dma(RES, pf + (4 * (f) + r) * 64, OFF)
// input f of row r: from the prefetch, else loaded now
#define EPV(f, r, RES, OFF) (PF ? pf[(4 * (f) + (r)) * 64 + lane] : bload2(RES, OFF))

// ---- IMEX step in three tile kernels around the two solve phases
//   k_imex_pre   rhs_construct_temp (rhs_create.F90:24-140) -> solve rows 0..N-1
//   k_imex_mid   the new temperature; rhs_construct_vort (:142-271), or
//                rhs_construct_uphi_bar (:273-385) for m = 0 -> rows N..2N-1
//   k_imex_post  the new vorticity, streamfunction and velocities, or the
//                new mean flow for m = 0.
template <int TW, bool STG, bool PF>
__global__ __launch_bounds__(64 * TW) __attribute__((amdgpu_waves_per_eu(2))) void k_imex_pre(const RunDesc* __restrict__ runs, const Shape sh, const int nrun)
{
    TILE_SETUP;
    // the run's array pointers, read once before any store (stores through
    // them could otherwise alias the descriptor and force a reload per store)
    const GPtr<double2> hiT_ = R.hiT;
    const GPtr<double2> heT_ = R.heT;
    const GPtr<double2> sx_ = R.sx;
    const int L = R.L;
    const int head = (R.head + 1) % L;
    auto slot = [&](int o) { return ((head - o + L) % L) * sst; };
    const double dt = R.dt, Pr = R.Pr;
    const bool cn = (R.imp == IMP_CN);
    const int ni = R.n_imp, ne = R.n_exp;
    const BRes ops[3] = {E1, E2, G};
    const BRes rTs = fres(R.tspec), rUr = fres(R.urT), rUp = fres(R.upT), rT = fres(R.tFR);
    const BRes rHi = fres(R.hiT, L), rHe = fres(R.heT, L);
    PF_BASE(2, 3);
    COL_STAGE(2, 1, 0, fin(rTs), fin(rUr), fin(rUr));
    for (int st = st0 + wv; st < st1; st += TW) {
        const int arow = 16 * st + li;
        STRIP_IDX(st);
        PF_FENCE();
        PF_LOAD(0, rT, fo[r]);
        PF_LOAD(1, rUp, fo[r]);
        PF_LOAD(2, rUr, fo[r]);
        dbl4 acc[6];
        acc_zero(acc);
        COL_PASS(ops, 2, 1, 0, fin(rTs), fin(rUr), fin(rUr));   // dT/dr, d2T/dr2; d(ur T)/dr
        PF_WAIT();
        STRIP_RR();
        double2 Tn[4], NT_[4], r0[4];
        UNROLL for (int r = 0; r < 4; ++r) {
            Tn[r] = EPV(0, r, rT, fo[r]);
            const double2 upT = EPV(1, r, rUp, fo[r]), urT = EPV(2, r, rUr, fo[r]);
            NT_[r] = cadd(cadd(cimr(mm, rri[r], upT), ACC(acc, 2, r)), csc(rri[r], urT));
        }
        if (cn) {
            const double c = ((1.0 - R.wi[0]) * dt) * (1.0 / Pr);
            UNROLL for (int r = 0; r < 4; ++r) {
                const double2 lap = csub(cadd(csc(rri[r], ACC(acc, 0, r)), ACC(acc, 1, r)), csc((mm * mm) * rr2i[r], Tn[r]));
                r0[r] = cadd(Tn[r], csc(c, lap));
            }
        } else {
            UNROLL for (int r = 0; r < 4; ++r) r0[r] = csc(R.wi[0], Tn[r]);
            slot_sum<MAXH - 1>(1, ni, [&](int o, int r) { return bload2(rHi, slot(o) + fo[r]); },
                               [&](int o, const double2 (&x)[4]) {
                                   const double w = R.wi[o];
                                   UNROLL for (int r = 0; r < 4; ++r) r0[r] = cadd(r0[r], csc(w, x[r]));
                               });
        }
        UNROLL for (int r = 0; r < 4; ++r) r0[r] = csub(r0[r], csc(dt, csc(R.we[0], NT_[r])));
        slot_sum<MAXH - 1>(1, ne, [&](int o, int r) { return bload2(rHe, slot(o) + fo[r]); },
                           [&](int o, const double2 (&x)[4]) {
                               const double w = R.we[o];
                               UNROLL for (int r = 0; r < 4; ++r) r0[r] = csub(r0[r], csc(dt, csc(w, x[r])));
                           });
        UNROLL for (int r = 0; r < 4; ++r) {
            if (!vr[r]) continue;
            const int i = ir[r];
            const size_t e = (size_t)(fo[r] >> 4);
            hiT_[(size_t)(slot(0) >> 4) + e] = Tn[r];
            heT_[(size_t)(slot(0) >> 4) + e] = NT_[r];
            // boundary rows: T = 1 (m = 0) or 0 at r_min, 0 at r_max
            double2 v = r0[r];
            if (i == 0) v = mk(m == 0 ? 1.0 : 0.0, 0.0);
            if (i == N - 1) v = mk(0.0, 0.0);
            sx_[sxe + i] = v;
        }
    }
}

template <int TW, bool STG, bool PF>
__global__ __launch_bounds__(64 * TW) __attribute__((amdgpu_waves_per_eu(2))) void k_imex_mid(const RunDesc* __restrict__ runs, const Shape sh, const int nrun)
{
    TILE_SETUP;
    // the run's array pointers, read once before any store (stores through
    // them could otherwise alias the descriptor and force a reload per store)
    const GPtr<double2> tFR_ = R.tFR;
    const GPtr<double2> tspec_ = R.tspec;
    const GPtr<double2> TFC_ = R.TFC;
    const GPtr<double2> hiW_ = R.hiW;
    const GPtr<double2> heW_ = R.heW;
    const GPtr<double2> hiU_ = R.hiU;
    const GPtr<double2> heU_ = R.heU;
    const GPtr<double2> sx_ = R.sx;
    const int L = R.L;
    const int head = (R.head + 1) % L;
    auto slot = [&](int o) { return ((head - o + L) % L) * sst; };
    auto uslot = [&](int o) { return ((head - o + L) % L) * N * 16; };
    const double dt = R.dt, Pr = R.Pr, Ra = R.Ra;
    const bool cn = (R.imp == IMP_CN);
    const int ni = R.n_imp, ne = R.n_exp;
    const double raPr = Ra / Pr;
    const BRes ops[4] = {E0, E1, E2, G};
    // omg_spec; in its m = 0 column the mean-flow coefficients upFC(1,:)
    const BRes rW = fres(R.wspec), rUr = fres(R.urW), rUp = fres(R.upW), rWn = fres(R.omgFR), rUb = fres(R.upFR);
    const BRes rHT = fres(R.hiT, L), rHi = fres(R.hiW, L), rHe = fres(R.heW, L);
    const BRes rHU = bres(R.hiU, 16 * (size_t)N * L), rEU = bres(R.heU, 16 * (size_t)N * L), rUw = bres(R.urw, 8 * (size_t)N);
    PF_BASE(3, 4);
    COL_STAGE(1, 2, 1, sya, fin(rW), fin(rUr));
    for (int st = st0 + wv; st < st1; st += TW) {
        const int arow = 16 * st + li;
        STRIP_IDX(st);
        PF_FENCE();
        PF_LOAD(0, syR, sxo + 16 * ir[r]);
        PF_LOAD(1, rWn, fo[r]);
        PF_LOAD(2, rUp, fo[r]);
        PF_LOAD(3, rUr, fo[r]);
        dbl4 acc[8];
        acc_zero(acc);
        // T^{n+1} from its coefficients; d/dr, d2/dr2 of omega (u_phi mean); d(ur omega)/dr
        COL_PASS(ops, 1, 2, 1, sya, fin(rW), fin(rUr));
        PF_WAIT();
        STRIP_RR();
        double2 xc[4];
        UNROLL for (int r = 0; r < 4; ++r) xc[r] = EPV(0, r, syR, sxo + 16 * ir[r]);
        if (m == 0) {
            // mean flow: rhs_construct_uphi_bar (rhs_create.F90:273-385)
            double2 ub[4];
            double v[4], uw[4];
            UNROLL for (int r = 0; r < 4; ++r) {
                ub[r] = bload2(rUb, fo[r]);
                uw[r] = bload(rUw, 8 * ir[r]);
                const double2 ud = cadd(csub(ACC(acc, 2, r), csc(rr2i[r], ub[r])), csc(rri[r], ACC(acc, 1, r)));
                v[r] = cn ? ub[r].x + ((1.0 - R.wi[0]) * dt) * ud.x : R.wi[0] * ub[r].x;
            }
            if (!cn)
                slot_sum<MAXH - 1>(1, ni, [&](int o, int r) { return bload2(rHU, uslot(o) + 16 * ir[r]); },
                                   [&](int o, const double2 (&x)[4]) {
                                       const double w = R.wi[o];
                                       UNROLL for (int r = 0; r < 4; ++r) v[r] = v[r] + w * x[r].x;
                                   });
            UNROLL for (int r = 0; r < 4; ++r) v[r] = v[r] - dt * (R.we[0] * uw[r]);
            slot_sum<MAXH - 1>(1, ne, [&](int o, int r) { return bload2(rEU, uslot(o) + 16 * ir[r]); },
                               [&](int o, const double2 (&x)[4]) {
                                   const double w = R.we[o];
                                   UNROLL for (int r = 0; r < 4; ++r) v[r] = v[r] - dt * (w * x[r].x);
                               });
            UNROLL for (int r = 0; r < 4; ++r) {
                if (!vr[r]) continue;
                const int i = ir[r];
                const size_t e = (size_t)(fo[r] >> 4);
                tspec_[e] = xc[r];
                TFC_[i] = xc[r];
                tFR_[e] = ACC(acc, 0, r);
                hiU_[(size_t)(uslot(0) >> 4) + i] = ub[r];
                heU_[(size_t)(uslot(0) >> 4) + i] = mk(uw[r], 0.0);
                sx_[sxe + N + i] = mk((i == 0 || i == N - 1) ? 0.0 : v[r], 0.0);
            }
            continue;
        }
        // vorticity: rhs_construct_vort (rhs_create.F90:142-271)
        double2 Wn[4], NW[4], r2[4];
        UNROLL for (int r = 0; r < 4; ++r) {
            Wn[r] = EPV(1, r, rWn, fo[r]);
            const double2 upW = EPV(2, r, rUp, fo[r]), urW = EPV(3, r, rUr, fo[r]);
            NW[r] = cadd(cadd(cimr(mm, rri[r], upW), ACC(acc, 3, r)), csc(rri[r], urW));
        }
        if (cn) {
            UNROLL for (int r = 0; r < 4; ++r) {
                const double2 lap = csub(cadd(csc(rri[r], ACC(acc, 1, r)), ACC(acc, 2, r)), csc((mm * mm) * rr2i[r], Wn[r]));
                const double2 b1 = cimr(mm, rri[r], ACC(acc, 0, r));
                const double2 b2 = cimr(mm, rri[r], bload2(rHT, slot(0) + fo[r]));
                r2[r] = csub(cadd(Wn[r], csc((1.0 - R.wi[0]) * dt, lap)), csc((raPr * dt) * 0.5, cadd(b1, b2)));
                r2[r] = csub(r2[r], csc(dt, csc(R.we[0], NW[r])));
            }
            slot_sum<MAXH - 1>(1, ne, [&](int o, int r) { return bload2(rHe, slot(o) + fo[r]); },
                               [&](int o, const double2 (&x)[4]) {
                                   const double w = R.we[o];
                                   UNROLL for (int r = 0; r < 4; ++r) r2[r] = csub(r2[r], csc(dt, csc(w, x[r])));
                               });
        } else {
            UNROLL for (int r = 0; r < 4; ++r) r2[r] = csc(R.wi[0], Wn[r]);
            slot_sum<MAXH - 1>(1, ni, [&](int o, int r) { return bload2(rHi, slot(o) + fo[r]); },
                               [&](int o, const double2 (&x)[4]) {
                                   const double w = R.wi[o];
                                   UNROLL for (int r = 0; r < 4; ++r) r2[r] = cadd(r2[r], csc(w, x[r]));
                               });
            UNROLL for (int r = 0; r < 4; ++r) r2[r] = csub(r2[r], csc(dt, csc(R.we[0], NW[r])));
            slot_sum<MAXH - 1>(1, ne, [&](int o, int r) { return bload2(rHe, slot(o) + fo[r]); },
                               [&](int o, const double2 (&x)[4]) {
                                   const double w = R.we[o];
                                   UNROLL for (int r = 0; r < 4; ++r) r2[r] = csub(r2[r], csc(dt, csc(w, x[r])));
                               });
            UNROLL for (int r = 0; r < 4; ++r) {
                const double2 bt = csc(-raPr, cimr(mm, rri[r], ACC(acc, 0, r)));
                r2[r] = cadd(r2[r], csc(dt * R.wl, bt));
            }
        }
        UNROLL for (int r = 0; r < 4; ++r) {
            if (!vr[r]) continue;
            const int i = ir[r];
            const size_t e = (size_t)(fo[r] >> 4);
            tspec_[e] = xc[r];
            tFR_[e] = ACC(acc, 0, r);
            hiW_[(size_t)(slot(0) >> 4) + e] = Wn[r];
            heW_[(size_t)(slot(0) >> 4) + e] = NW[r];
            sx_[sxe + N + i] = csel(i == 0 || i == N - 1, mk(0.0, 0.0), r2[r]);
        }
    }
}

template <int TW, bool STG, bool PF>
__global__ __launch_bounds__(64 * TW) __attribute__((amdgpu_waves_per_eu(2))) void k_imex_post(const RunDesc* __restrict__ runs, const Shape sh, const int nrun)
{
    TILE_SETUP_S(true);
    // the run's array pointers, read once before any store (stores through
    // them could otherwise alias the descriptor and force a reload per store)
    const GPtr<double2> omgFR_ = R.omgFR;
    const GPtr<double2> upFR_ = R.upFR;
    const GPtr<double2> urFR_ = R.urFR;
    const GPtr<double2> psi_ = R.psi;
    const GPtr<double2> wspec_ = R.wspec;
    const BRes ops[4] = {E0, E1, E0, E1};
    PF_BASE(2, 1);
    COL_STAGE(2, 2, 0, syb, sya, sya);
    for (int st = st0 + wv; st < st1; st += TW) {
        const int arow = 16 * st + li;
        STRIP_IDX(st);
        PF_FENCE();
        PF_LOAD(0, syR, sxo + 16 * (N + ir[r]));
        dbl4 acc[8];
        acc_zero(acc);
        // omega (u_phi mean for m = 0) values and d/dr; psi values and d/dr
        COL_PASS(ops, 2, 2, 0, syb, sya, sya);
        PF_WAIT();
        STRIP_RR();
        double2 xw[4];
        UNROLL for (int r = 0; r < 4; ++r) xw[r] = EPV(0, r, syR, sxo + 16 * (N + ir[r]));
        UNROLL for (int r = 0; r < 4; ++r) {
            if (!vr[r]) continue;
            const size_t e = (size_t)(fo[r] >> 4);
            wspec_[e] = xw[r];   // omg_spec, or upFC(1,:) for m = 0
            if (m == 0) {
                const double2 u = ACC(acc, 0, r);
                upFR_[e] = u;
                urFR_[e] = mk(0.0, 0.0);
                omgFR_[e] = cadd(csc(rri[r], u), ACC(acc, 1, r));
                continue;
            }
            omgFR_[e] = ACC(acc, 0, r);
            const double2 ps = ACC(acc, 2, r);
            psi_[e] = ps;
            upFR_[e] = cneg(ACC(acc, 3, r));
            urFR_[e] = cimr(mm, rri[r], ps);
        }
    }
}

// ---- IMEXRK stage s (1 <= s < S) of a stiffly accurate scheme in three tile
// kernels around the two solve phases (with the stage terms of
// RHS_construct_stage, fourierloop_imexrk.F90:429-488):
//   k_rk_pre   the explicit terms of stage s-1 (rhs_construct_temp_a /
//              vort_a) and the temperature stage right-hand side
//   k_rk_mid   the stage temperature and its diffusion term
//              (rhs_construct_temp_d, rhs_create_imexrk.F90:59-93); the
//              vorticity (or mean flow) stage right-hand side
//   k_rk_post  the stage vorticity, streamfunction and velocities (or mean
//              flow) and their diffusion terms (rhs_construct_vort_d /
//              uphibar_d, :162-191, 244-289); after the last stage the step's
//              new state (Assembly_stage_SA, fourierloop_imexrk.F90:653-665).
// The diffusion terms of stage s go to slot s; those of the last stage are
// the next step's first-stage terms (slot 0).
template <int TW, bool STG, bool PF>
__global__ __launch_bounds__(64 * TW) __attribute__((amdgpu_waves_per_eu(3))) void k_rk_pre(const RunDesc* __restrict__ runs, const Shape sh, const int nrun,
                                               const int s)
{
    TILE_SETUP_S(true);
    // the run's array pointers, read once before any store (stores through
    // them could otherwise alias the descriptor and force a reload per store)
    const GPtr<double2> FaT_ = R.FaT;
    const GPtr<double2> FaW_ = R.FaW;
    const GPtr<double2> sx_ = R.sx;
    const double dt = R.dt;
    const double* aD = R.aD + s * MAXS;
    const double* aA = R.aA + s * MAXS;
    const bool fa_prev = R.fa_col[s - 1] != 0;
    const int SS = R.S;
    const BRes ops[2] = {G, G};
    const BRes rUrT = fres(R.urT), rUrW = fres(R.urW), rUpT = fres(R.upT), rUpW = fres(R.upW);
    const BRes rFdT = fres(R.FdT, SS), rFaT = fres(R.FaT, SS), rTs = fres(R.tspec);
    const BRes rUw = bres(R.urw, 8 * (size_t)N);
    PF_BASE(2, 5);
    if (fa_prev) { COL_STAGE(1, 1, 0, fin(rUrT), fin(rUrW), fin(rUrW)); }
    for (int st = st0 + wv; st < st1; st += TW) {
        const int arow = 16 * st + li;
        STRIP_IDX(st);
        PF_FENCE();
        PF_LOAD(0, rTs, fo[r]);
        if (fa_prev) {
            PF_LOAD(1, rUpT, fo[r]);
            PF_LOAD(2, rUrT, fo[r]);
            PF_LOAD(3, rUpW, fo[r]);
            PF_LOAD(4, rUrW, fo[r]);
        }
        dbl4 acc[4];
        acc_zero(acc);
        if (fa_prev) COL_PASS_KC(2, ops, 1, 1, 0, fin(rUrT), fin(rUrW), fin(rUrW));   // d(ur T)/dr, d(ur omega)/dr
        PF_WAIT();
        STRIP_RR();
        double2 faT[4], faW[4], Fd[4], Fa[4];
        if (fa_prev) {
            UNROLL for (int r = 0; r < 4; ++r) {
                faT[r] = cneg(cadd(cadd(cimr(mm, rri[r], EPV(1, r, rUpT, fo[r])), ACC(acc, 0, r)),
                                   csc(rri[r], EPV(2, r, rUrT, fo[r]))));
                if (m == 0) faW[r] = mk(-bload(rUw, 8 * ir[r]), 0.0);
                else faW[r] = cneg(cadd(cadd(cimr(mm, rri[r], EPV(3, r, rUpW, fo[r])), ACC(acc, 1, r)),
                                        csc(rri[r], EPV(4, r, rUrW, fo[r]))));
            }
        }
        UNROLL for (int r = 0; r < 4; ++r) { Fd[r] = mk(0.0, 0.0); Fa[r] = mk(0.0, 0.0); }
        slot_sum<2>(0, s, [&](int k, int r) { return bload2(rFdT, k * sst + fo[r]); },
                    [&](int k, const double2 (&x)[4]) {
                        const double w = aD[k];
                        UNROLL for (int r = 0; r < 4; ++r) Fd[r] = cadd(Fd[r], csc(w, x[r]));
                    });
        // the terms of stage s - 1 are the ones just formed when fa_prev
        slot_sum<2>(0, fa_prev ? s - 1 : s, [&](int k, int r) { return bload2(rFaT, k * sst + fo[r]); },
                    [&](int k, const double2 (&x)[4]) {
                        const double w = aA[k];
                        UNROLL for (int r = 0; r < 4; ++r) Fa[r] = cadd(Fa[r], csc(w, x[r]));
                    });
        if (fa_prev) UNROLL for (int r = 0; r < 4; ++r) Fa[r] = cadd(Fa[r], csc(aA[s - 1], faT[r]));
        double2 v[4];
        UNROLL for (int r = 0; r < 4; ++r) v[r] = cadd(EPV(0, r, rTs, fo[r]), csc(dt, cadd(Fd[r], Fa[r])));
        UNROLL for (int r = 0; r < 4; ++r) {
            if (!vr[r]) continue;
            const int i = ir[r];
            const size_t e = (size_t)(fo[r] >> 4);
            if (fa_prev) {
                FaT_[(size_t)(s - 1) * S + e] = faT[r];
                FaW_[(size_t)(s - 1) * S + e] = faW[r];
            }
            double2 x = v[r];
            if (i == 0) x = mk(m == 0 ? 1.0 : 0.0, 0.0);
            if (i == N - 1) x = mk(0.0, 0.0);
            sx_[sxe + i] = x;
        }
    }
}

template <int TW, bool STG, bool PF>
__global__ __launch_bounds__(64 * TW) __attribute__((amdgpu_waves_per_eu(2))) void k_rk_mid(const RunDesc* __restrict__ runs, const Shape sh, const int nrun,
                                               const int s)
{
    TILE_SETUP;
    // the run's array pointers, read once before any store (stores through
    // them could otherwise alias the descriptor and force a reload per store)
    const GPtr<double2> tFR_ = R.tFR;
    const GPtr<double2> tspec_ = R.tspec;
    const GPtr<double2> TFC_ = R.TFC;
    const GPtr<double2> FdT_ = R.FdT;
    const GPtr<double2> Fb_ = R.Fb;
    const GPtr<double2> sx_ = R.sx;
    const int SS = R.S;
    // the last stage of a stiffly accurate scheme is the new state, and its
    // terms are the next step's first-stage terms (slot 0); otherwise the
    // stage keeps its own slot and k_rk_asm assembles the state
    const bool last = (s == SS - 1) && R.sa != 0;
    const double dt = R.dt, Pr = R.Pr;
    const double raPr = R.Ra / R.Pr;
    const double* aD = R.aD + s * MAXS;
    const double* aA = R.aA + s * MAXS;
    const int fslot = last ? 0 : s;
    const bool fd = R.need_fd[fslot] != 0;
    const bool bimp = R.buo_imp != 0;
    const BRes ops[3] = {E0, E1, E2};
    const BRes rFdW = fres(R.FdW, SS), rFaW = fres(R.FaW, SS), rFb = fres(R.Fb, SS), rWs = fres(R.wspec);
    const BRes rUb = bres(R.ubar, 16 * (size_t)N);
    PF_BASE(1, 1);
    COL_STAGE(1, 0, 0, sya, sya, sya);
    for (int st = st0 + wv; st < st1; st += TW) {
        const int arow = 16 * st + li;
        STRIP_IDX(st);
        PF_FENCE();
        PF_LOAD(0, rWs, fo[r]);
        dbl4 acc[6];
        acc_zero(acc);
        if (fd) COL_PASS(ops, 3, 0, 0, sya, sya, sya);   // stage T values, d/dr, d2/dr2
        else COL_PASS(ops, 1, 0, 0, sya, sya, sya);
        PF_WAIT();
        STRIP_RR();
        // stage start value (omega, or the mean flow for m = 0) and the weighted
        // stage terms of the vorticity or mean-flow equation
        double2 y0[4], Fd[4], Fa[4], Fb[4], bu[4];
        UNROLL for (int r = 0; r < 4; ++r) {
            y0[r] = (m == 0) ? bload2(rUb, 16 * ir[r]) : EPV(0, r, rWs, fo[r]);
            Fd[r] = mk(0.0, 0.0); Fa[r] = mk(0.0, 0.0); Fb[r] = mk(0.0, 0.0);
            bu[r] = csc(-raPr, cimr(mm, rri[r], ACC(acc, 0, r)));   // buoyancy of this stage
        }
        slot_sum<2>(0, s, [&](int k, int r) { return bload2(rFdW, k * sst + fo[r]); },
                    [&](int k, const double2 (&x)[4]) {
                        const double w = aD[k];
                        UNROLL for (int r = 0; r < 4; ++r) Fd[r] = cadd(Fd[r], csc(w, x[r]));
                    });
        slot_sum<2>(0, s, [&](int k, int r) { return bload2(rFaW, k * sst + fo[r]); },
                    [&](int k, const double2 (&x)[4]) {
                        const double w = aA[k];
                        UNROLL for (int r = 0; r < 4; ++r) Fa[r] = cadd(Fa[r], csc(w, x[r]));
                    });
        if (m > 0) {
            // the buoyancy terms: weights aD (implicit buoyancy) or aA
            const double* wb = bimp ? aD : aA;
            slot_sum<2>(0, s, [&](int k, int r) { return bload2(rFb, k * sst + fo[r]); },
                        [&](int k, const double2 (&x)[4]) {
                            const double w = wb[k];
                            UNROLL for (int r = 0; r < 4; ++r) Fb[r] = cadd(Fb[r], csc(w, x[r]));
                        });
            if (bimp) UNROLL for (int r = 0; r < 4; ++r) Fb[r] = cadd(Fb[r], csc(aD[s], bu[r]));
        }
        double2 xc[4];
        if (last && m == 0) UNROLL for (int r = 0; r < 4; ++r) xc[r] = bload2(syR, sxo + 16 * ir[r]);
        UNROLL for (int r = 0; r < 4; ++r) {
            if (!vr[r]) continue;
            const int i = ir[r];
            const size_t e = (size_t)(fo[r] >> 4);
            const double2 T1 = ACC(acc, 0, r);
            tFR_[e] = T1;
            if (fd) {
                const double2 lap = csub(cadd(csc(rri[r], ACC(acc, 1, r)), ACC(acc, 2, r)), csc((mm * mm) * rr2i[r], T1));
                FdT_[(size_t)fslot * S + e] = csc(1.0 / Pr, lap);
            }
            if (last) {   // Assembly_stage_SA: the stage value is the new state
                tspec_[e] = T1;
                // TFC = chebtransform(tFR(1,:)) (steptime_imexrk.F90:190): the
                // coefficients the stage temperature was evaluated from
                if (m == 0) TFC_[i] = xc[r];
            }
            if (m == 0) {
                // mean flow stage (fourierloop_imexrk.F90:233-273)
                const double2 v = cadd(y0[r], csc(dt, cadd(Fd[r], Fa[r])));
                sx_[sxe + N + i] = mk((i == 0 || i == N - 1) ? 0.0 : v.x, 0.0);
                continue;
            }
            // vorticity stage (fourierloop_imexrk.F90:276-388)
            const double2 v = cadd(y0[r], csc(dt, cadd(cadd(Fd[r], Fa[r]), Fb[r])));
            sx_[sxe + N + i] = csel(i == 0 || i == N - 1, mk(0.0, 0.0), v);
            // this stage's buoyancy; after the last stage it is the buoyancy of
            // the next step's start state (slot 0, read above by this lane)
            Fb_[(size_t)fslot * S + e] = bu[r];
        }
    }
}

template <int TW, bool STG, bool PF>
__global__ __launch_bounds__(64 * TW) __attribute__((amdgpu_waves_per_eu(2))) void k_rk_post(const RunDesc* __restrict__ runs, const Shape sh, const int nrun,
                                                const int s)
{
    TILE_SETUP_S(true);
    // the run's array pointers, read once before any store (stores through
    // them could otherwise alias the descriptor and force a reload per store)
    const GPtr<double2> omgFR_ = R.omgFR;
    const GPtr<double2> upFR_ = R.upFR;
    const GPtr<double2> urFR_ = R.urFR;
    const GPtr<double2> psi_ = R.psi;
    const GPtr<double2> wspec_ = R.wspec;
    const GPtr<double2> FdW_ = R.FdW;
    const GPtr<double2> ubar_ = R.ubar;
    const bool last = (s == R.S - 1) && R.sa != 0;
    const int fslot = last ? 0 : s;
    const bool fd = R.need_fd[fslot] != 0;
    const BRes ops[5] = {E0, E1, E2, E0, E1};
    const BRes ops2[4] = {E0, E1, E0, E1};
    COL_STAGE(2, 2, 0, syb, sya, sya);
    for (int st = st0 + wv; st < st1; st += TW) {
        const int arow = 16 * st + li;
        dbl4 acc[10];
        acc_zero(acc);
        // omega (u_phi mean for m = 0) values, d/dr and, for the diffusion
        // term, d2/dr2; psi values and d/dr
        int qp;   // first psi product
        if (fd) { COL_PASS(ops, 3, 2, 0, syb, sya, sya); qp = 3; }
        else { COL_PASS(ops2, 2, 2, 0, syb, sya, sya); qp = 2; }
        STRIP_ROWS(st);
        UNROLL for (int r = 0; r < 4; ++r) {
            if (!vr[r]) continue;
            const int i = ir[r];
            const size_t e = (size_t)(fo[r] >> 4);
            const double2 V1 = ACC(acc, 0, r);
            if (m == 0) {
                upFR_[e] = V1;
                if (fd) {
                    // rhs_construct_uphibar_d (rhs_create_imexrk.F90:162-191)
                    const double2 v = cadd(csub(ACC(acc, 2, r), csc(rr2i[r], V1)), csc(rri[r], ACC(acc, 1, r)));
                    FdW_[(size_t)fslot * S + e] = mk(v.x, 0.0);
                }
                const double2 om = cadd(csc(rri[r], V1), ACC(acc, 1, r));
                urFR_[e] = mk(0.0, 0.0);
                omgFR_[e] = om;
                if (last) {
                    wspec_[e] = om;
                    ubar_[i] = V1;
                }
                continue;
            }
            omgFR_[e] = V1;
            const double2 ps = qp == 3 ? ACC(acc, 3, r) : ACC(acc, 2, r);
            const double2 dps = qp == 3 ? ACC(acc, 4, r) : ACC(acc, 3, r);
            psi_[e] = ps;
            upFR_[e] = cneg(dps);
            urFR_[e] = cimr(mm, rri[r], ps);
            if (fd)   // rhs_construct_vort_d (rhs_create_imexrk.F90:244-289)
                FdW_[(size_t)fslot * S + e] =
                    csub(cadd(csc(rri[r], ACC(acc, 1, r)), ACC(acc, 2, r)), csc((mm * mm) * rr2i[r], V1));
            if (last) wspec_[e] = V1;
        }
    }
}

// ------------------------------------------------------------------ matrices
// mat_build / mat_build_IRK / mat_build_uphibar (mat_assembly.F90:68-413), one
// column j per workgroup; every + and * rounds separately, as written there.

// AT of modes m0 + blockIdx.y (mat_assembly.F90:287-301): A + blockIdx.y*N*N.
__global__ __launch_bounds__(NT) void k_asm_T(double* __restrict__ A, const double* __restrict__ t,
                                              const double* __restrict__ D, const double* __restrict__ D2,
                                              const Shape sh, const double wtdt, const double Pr, const int m0)
{
#pragma clang fp contract(off)
    const int N = sh.N;
    const int j = blockIdx.x;
    const double nn = (double)(m0 + (int)blockIdx.y);
    const double C = sqrt(2.0 / (double)(N - 1));
    const GPtr<const double> rr = {sh.rr};
    const GPtr<const double> rr2 = {sh.rr2};
    auto IX = [&](int i, int jj) { return (size_t)i + (size_t)N * (size_t)jj; };
    const bool halfj = (j == 0 || j == N - 1);
    const double cT = wtdt * (1.0 / Pr);
    double* At = A + (size_t)blockIdx.y * N * N;
    for (int i = threadIdx.x; i < N; i += NT) {
        double v = t[IX(i, j)] - cT * ((rr[i] * D[IX(i, j)] + D2[IX(i, j)]) - ((nn * nn) * rr2[i]) * t[IX(i, j)]);
        if (i == 0 || i == N - 1) v = t[IX(i, j)];
        if (halfj) v = 0.5 * v;
        At[IX(i, j)] = C * v;
    }
}

// AF of modes m0 + blockIdx.y >= 1: A + blockIdx.y*4N*N, with mat_build's
// boundary rows (irk = 0, mat_assembly.F90:304-345) or mat_build_IRK's
// (irk = 1, :143-150), which swap rows 1,N with N+1,2N.  The swap permutes
// only the inverse's columns 1, N, N+1, 2N; the interior columns that the
// inverse-mode solves use are the same for both orders.
__global__ __launch_bounds__(NT) void k_asm_F(double* __restrict__ A, const double* __restrict__ t,
                                              const double* __restrict__ D, const double* __restrict__ D2,
                                              const Shape sh, const double wtdt, const int mBC, const int irk,
                                              const int m0)
{
#pragma clang fp contract(off)
    const int N = sh.N;
    const int j = blockIdx.x;
    const double nn = (double)(m0 + (int)blockIdx.y);
    const double C = sqrt(2.0 / (double)(N - 1));
    const GPtr<const double> rr = {sh.rr};
    const GPtr<const double> rr2 = {sh.rr2};
    auto IX = [&](int i, int jj) { return (size_t)i + (size_t)N * (size_t)jj; };
    const bool halfj = (j == 0 || j == N - 1);
    double* Af = A + (size_t)blockIdx.y * 4 * N * N;
    const size_t L = 2 * (size_t)N;
    for (int i = threadIdx.x; i < N; i += NT) {
        const double lap = (rr[i] * D[IX(i, j)] + D2[IX(i, j)]) - ((nn * nn) * rr2[i]) * t[IX(i, j)];
        double pp = lap;                       // (i, j)
        double pw = t[IX(i, j)];               // (i, N+j)
        double wp = 0.0;                       // (N+i, j)
        double ww = t[IX(i, j)] - wtdt * lap;  // (N+i, N+j)
        if (i == 0 || i == N - 1) {
            const double bc1 = t[IX(i, j)];
            const double bc2 = (mBC == 0) ? D[IX(i, j)] : D2[IX(i, j)] - rr[i] * D[IX(i, j)];
            pp = irk ? bc2 : bc1;
            pw = 0.0;
            wp = irk ? bc1 : bc2;
            ww = 0.0;
        }
        if (halfj) { pp = 0.5 * pp; pw = 0.5 * pw; wp = 0.5 * wp; ww = 0.5 * ww; }
        Af[(size_t)i + L * j] = C * pp;
        Af[(size_t)i + L * (N + j)] = C * pw;
        Af[(size_t)(N + i) + L * j] = C * wp;
        Af[(size_t)(N + i) + L * (N + j)] = C * ww;
    }
}

// A_uphi of mat_build_uphibar (mat_assembly.F90:393-406).
__global__ __launch_bounds__(NT) void k_asm_U(double* __restrict__ A, const double* __restrict__ t,
                                              const double* __restrict__ D, const double* __restrict__ D2,
                                              const Shape sh, const double wtdt, const int mBC)
{
#pragma clang fp contract(off)
    const int N = sh.N;
    const int j = blockIdx.x;
    const double C = sqrt(2.0 / (double)(N - 1));
    const GPtr<const double> rr = {sh.rr};
    const GPtr<const double> rr2 = {sh.rr2};
    auto IX = [&](int i, int jj) { return (size_t)i + (size_t)N * (size_t)jj; };
    const bool halfj = (j == 0 || j == N - 1);
    for (int i = threadIdx.x; i < N; i += NT) {
        double v = t[IX(i, j)] - wtdt * ((rr[i] * D[IX(i, j)] + D2[IX(i, j)]) - rr2[i] * t[IX(i, j)]);
        if (i == 0 || i == N - 1) {
            if (mBC == 0) v = t[IX(i, j)];
            else v = D[IX(i, j)] - rr[i] * t[IX(i, j)];
        }
        if (halfj) v = 0.5 * v;
        A[IX(i, j)] = C * v;
    }
}

// ------------------------------------------------------------------ LU factor
constexpr int LU_NB = 8;
constexpr int LU_MAXN = 2 * MAX_N;

// Blocked right-looking LU with partial pivoting of one n x n column-major
// matrix per workgroup.  Each element sees exactly dgetf2's sequence of
// operations (pivot = first max |a|, column scaled by the reciprocal pivot,
// a(i,k) = fma(-l(i,j), u(j,k), a(i,k)) for j = 0, 1, ...); the panel of
// LU_NB columns lives in LDS and the trailing matrix is streamed once per
// panel.  Writes the net row permutation of dgetrs' interchanges.
__global__ __launch_bounds__(NT) void k_lu(double* __restrict__ A_all, int* __restrict__ perm_all, const int n,
                                           const size_t stride)
{
    extern __shared__ double P[];          // rows x LU_NB
    __shared__ int ipiv[LU_MAXN];
    __shared__ double redv[NT];
    __shared__ int redi[NT];
    double* A = A_all + (size_t)blockIdx.x * stride;
    int* perm = perm_all + (size_t)blockIdx.x * n;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const size_t ld = (size_t)n;
    for (int k0 = 0; k0 < n; k0 += LU_NB) {
        const int kb = min(LU_NB, n - k0);
        const int rows = n - k0;
        for (int idx = tid; idx < rows * kb; idx += NT) {
            const int c = idx / rows, r = idx - c * rows;
            P[r + rows * c] = A[(size_t)(k0 + r) + ld * (size_t)(k0 + c)];
        }
        __syncthreads();
        for (int c = 0; c < kb; ++c) {
            double best = -1.0;
            int bidx = 0x7fffffff;
            for (int r = c + tid; r < rows; r += NT) {
                const double v = fabs(P[r + rows * c]);
                if (v > best) { best = v; bidx = r; }
            }
            redv[tid] = best;
            redi[tid] = bidx;
            __syncthreads();
            for (int k = NT / 2; k > 0; k >>= 1) {
                if (tid < k) {
                    const double v2 = redv[tid + k];
                    const int i2 = redi[tid + k];
                    if (v2 > redv[tid] || (v2 == redv[tid] && i2 < redi[tid])) { redv[tid] = v2; redi[tid] = i2; }
                }
                __syncthreads();
            }
            const int p = redi[0];
            if (tid == 0) ipiv[k0 + c] = k0 + p;
            const double pv = P[p + rows * c];
            __syncthreads();
            if (pv != 0.0) {
                if (p != c && tid < kb) {
                    const double a = P[c + rows * tid];
                    P[c + rows * tid] = P[p + rows * tid];
                    P[p + rows * tid] = a;
                }
                __syncthreads();
                const double piv = P[c + rows * c];
                if (fabs(piv) >= DBL_MIN) {
                    const double rcp = 1.0 / piv;
                    for (int r = c + 1 + tid; r < rows; r += NT) P[r + rows * c] = P[r + rows * c] * rcp;
                } else {
                    for (int r = c + 1 + tid; r < rows; r += NT) P[r + rows * c] = P[r + rows * c] / piv;
                }
            }
            __syncthreads();
            for (int cc = c + 1; cc < kb; ++cc) {
                const double u = P[c + rows * cc];
                for (int r = c + 1 + tid; r < rows; r += NT) P[r + rows * cc] = fma(-P[r + rows * c], u, P[r + rows * cc]);
            }
            __syncthreads();
        }
        for (int idx = tid; idx < rows * kb; idx += NT) {
            const int c = idx / rows, r = idx - c * rows;
            A[(size_t)(k0 + r) + ld * (size_t)(k0 + c)] = P[r + rows * c];
        }
        // this panel's interchanges on the columns left and right of it
        for (int col = tid; col < n; col += NT) {
            if (col >= k0 && col < k0 + kb) continue;
            double* cp = A + ld * (size_t)col;
            for (int c = 0; c < kb; ++c) {
                const int p = ipiv[k0 + c];
                if (p != k0 + c) { const double a = cp[k0 + c]; cp[k0 + c] = cp[p]; cp[p] = a; }
            }
        }
        __syncthreads();
        // U12 = L11^{-1} A12 and A22 -= L21 U12, one wavefront per column
        for (int col = k0 + kb + wave; col < n; col += NT / 64) {
            double* cp = A + ld * (size_t)col;
            double u = (lane < kb) ? cp[k0 + lane] : 0.0;
#pragma unroll
            for (int j = 0; j < LU_NB - 1; ++j) {
                if (j < kb - 1) {
                    const double uj = __shfl(u, j);
                    if (lane > j && lane < kb) u = fma(-P[lane + rows * j], uj, u);
                }
            }
            if (lane < kb) cp[k0 + lane] = u;
            double ub[LU_NB];
#pragma unroll
            for (int j = 0; j < LU_NB; ++j) ub[j] = __shfl(u, j);
            for (int r = k0 + kb + lane; r < n; r += 64) {
                double a = cp[r];
#pragma unroll
                for (int j = 0; j < LU_NB; ++j)
                    if (j < kb) a = fma(-P[(r - k0) + rows * j], ub[j], a);
                cp[r] = a;
            }
        }
        __syncthreads();
    }
    // net permutation of the interchanges: perm[i] = original row of row i
    __shared__ int pidx[LU_MAXN];
    for (int i = tid; i < n; i += NT) pidx[i] = i;
    __syncthreads();
    if (tid == 0) {
        for (int k = 0; k < n; ++k) {
            const int p = ipiv[k];
            if (p != k) { const int a = pidx[k]; pidx[k] = pidx[p]; pidx[p] = a; }
        }
    }
    __syncthreads();
    for (int i = tid; i < n; i += NT) perm[i] = pidx[i];
}

// ------------------------------------------------------------------ inverse
// Columns col0 .. col0+ncols-1 of A^{-1} from the packed LU factor and net row
// permutation of k_lu: dgetrs on the unit vectors, G of them per wavefront.
// Lane l keeps rows l, l+64, ... of the G right-hand sides in registers; the
// sweep broadcasts the finished x(j) with readlane while the factor columns
// D ahead are in flight.  blockIdx.y: matrix of the chunk (stride `stride`,
// permutation stride n); output column-major n x ncols at out + y*ostride.
template <int SLOTS, int G, int D>
__global__ __launch_bounds__(64) void k_inv(const double* __restrict__ LU, const int* __restrict__ perm, const int n,
                                            const size_t stride, const int col0, const int ncols,
                                            double* __restrict__ out, const size_t ostride)
{
    const double* A = LU + (size_t)blockIdx.y * stride;
    const int* pm = perm + (size_t)blockIdx.y * n;
    double* O = out + (size_t)blockIdx.y * ostride;
    const int lane = threadIdx.x;
    const int cbase = (int)blockIdx.x * G;     // first column of this block, relative to col0
    const size_t ld = (size_t)n;
    double xv[G][SLOTS];
#pragma unroll
    for (int q = 0; q < SLOTS; ++q) {
        const int r = lane + 64 * q;
        const int src = (r < n) ? pm[r] : -1;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int c = col0 + cbase + g;
            xv[g][q] = (cbase + g < ncols && src == c) ? 1.0 : 0.0;
        }
    }
    double pf[D][SLOTS];
    // unit lower triangle: column j updates rows r > j
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
        for (int q = 0; q < SLOTS; ++q) {
            const int r = lane + 64 * q;
            pf[d][q] = (d < n && r > d && r < n) ? A[(size_t)r + ld * (size_t)d] : 0.0;
        }
    for (int j0 = 0; j0 < n; j0 += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int j = j0 + d;
            if (j < n) {
                const int qj = j >> 6, lj = j & 63;
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    double v = xv[g][0];
#pragma unroll
                    for (int q = 1; q < SLOTS; ++q)
                        if (qj == q) v = xv[g][q];
                    const double xj = rdl(v, lj);
#pragma unroll
                    for (int q = 0; q < SLOTS; ++q) {
                        if (64 * q + 63 > j) {           // slot holds rows below j
                            const int r = lane + 64 * q;
                            if (r > j && r < n) xv[g][q] = fma(-xj, pf[d][q], xv[g][q]);
                        }
                    }
                }
                const int jn = j + D;
#pragma unroll
                for (int q = 0; q < SLOTS; ++q) {
                    const int r = lane + 64 * q;
                    if (jn < n && r > jn && r < n) pf[d][q] = A[(size_t)r + ld * (size_t)jn];
                }
            }
        }
    }
    // upper triangle: x(j) /= u(j,j), then column j updates rows r < j
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
        for (int q = 0; q < SLOTS; ++q) {
            const int r = lane + 64 * q;
            const int jc = n - 1 - d;
            pf[d][q] = (jc >= 0 && r <= jc) ? A[(size_t)r + ld * (size_t)jc] : 1.0;
        }
    for (int j0 = n - 1; j0 >= 0; j0 -= D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int j = j0 - d;
            if (j >= 0) {
                const int qj = j >> 6, lj = j & 63;
#pragma unroll
                for (int g = 0; g < G; ++g) {
#pragma unroll
                    for (int q = 0; q < SLOTS; ++q)
                        if (q == qj && lane == lj) xv[g][q] = xv[g][q] / pf[d][q];
                    double v = xv[g][0];
#pragma unroll
                    for (int q = 1; q < SLOTS; ++q)
                        if (qj == q) v = xv[g][q];
                    const double xj = rdl(v, lj);
#pragma unroll
                    for (int q = 0; q < SLOTS; ++q) {
                        if (64 * q < j) {                // slot holds rows above j
                            const int r = lane + 64 * q;
                            if (r < j) xv[g][q] = fma(-xj, pf[d][q], xv[g][q]);
                        }
                    }
                }
                const int jn = j - D;
#pragma unroll
                for (int q = 0; q < SLOTS; ++q) {
                    const int r = lane + 64 * q;
                    if (jn >= 0 && r <= jn) pf[d][q] = A[(size_t)r + ld * (size_t)jn];
                }
            }
        }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
        if (cbase + g >= ncols) break;
#pragma unroll
        for (int q = 0; q < SLOTS; ++q) {
            const int r = lane + 64 * q;
            if (r < n) O[(size_t)r + ld * (size_t)(cbase + g)] = xv[g][q];
        }
    }
}

// ------------------------------------------------------------------ diagnostics
// writeKE_spectral: KE = 2 pi (radInt(sum_m |ur_m^2|) + radInt(sum_m |up_m^2|))
// and the Nusselt numbers from d/dr of TFC (output.F90:232-265, integ.F90:36-76)
template <int NT, int MM>
__global__ __launch_bounds__(NT) void k_diag(RunDesc* __restrict__ runs, const Shape sh0, const int record,
                                             double* __restrict__ out6)
{
    extern __shared__ double2 smem[];
    const Shape sh = lds_twiddles<NT>(sh0, smem);     // twiddle table: smem[0..M)
    const int N = sh.N, Nm1 = sh.Nm1;
    RunDesc& R = runs[blockIdx.x];
    double2* buf = smem + sh.M;
    double2* V0 = buf + sh.M + 4 * N;
    double2* V1 = V0 + N;
    double2* V2 = V1 + N;
    double* red = (double*)(V2 + N);
    const int tid = threadIdx.x;
    for (int j = tid; j < N; j += NT) {
        double a = 0.0, b = 0.0;
        for (int i = 0; i < Nm1; ++i) {
            const double2 zr = R.urFR[i + (size_t)Nm1 * j], zp = R.upFR[i + (size_t)Nm1 * j];
            const double2 sr = cmul(zr, zr), sp = cmul(zp, zp);
            a = a + hypot(sr.x, sr.y);
            b = b + hypot(sp.x, sp.y);
        }
        V1[j] = mk(a * sh.rad[j], b * sh.rad[j]);
    }
    __syncthreads();
    cheb_fwd<NT, MM>(buf, V0, sh, [&](int i) { return V1[i]; });
    double sa = 0.0, sb = 0.0;
    for (int i = tid; i < N; i += NT) {
        const double Qm = (i % 2 == 0) ? -1.0 / (double)(i * i - 1) : 0.0;
        const double wgt = (i == 0 || i == N - 1) ? 0.5 : 1.0;
        sa = sa + wgt * V0[i].x * Qm;
        sb = sb + wgt * V0[i].y * Qm;
    }
    red[tid] = sa;
    red[NT + tid] = sb;
    __syncthreads();
    for (int k = NT / 2; k > 0; k >>= 1) {
        if (tid < k) { red[tid] = red[tid] + red[tid + k]; red[NT + tid] = red[NT + tid] + red[NT + tid + k]; }
        __syncthreads();
    }
    const double urk = sh.fact * red[0], upk = sh.fact * red[NT];
    cheb_beta<NT, MM>(V2, sh, [&](int k) { return R.TFC[k]; });
    cheb_eval_beta<NT, MM>(buf, V2, sh, [&](int i, double2 v) { V1[i] = v; });
    if (tid == 0) {
        const double pi = 3.14159265358979323846;
        const double eta = sh.eta;
        const double KE = 2.0 * pi * (urk + upk);
        const double Nut = V1[N - 1].x / ((1.0 - eta) / (log(eta)));
        const double Nub = V1[0].x / ((1.0 - eta) / (eta * log(eta)));
        if (out6) {
            double* o = out6 + 6 * blockIdx.x;
            o[0] = R.time; o[1] = KE; o[2] = Nut; o[3] = Nub; o[4] = 2.0 * pi * urk; o[5] = 2.0 * pi * upk;
        }
        if (record && R.series && R.nrec < R.cap) {
            double* row = R.series + 4 * (size_t)R.nrec;
            row[0] = R.time; row[1] = KE; row[2] = Nut; row[3] = Nub;
            R.nrec = R.nrec + 1;
        }
    }
}

// writeKE_spectral's further outputs need, besides the state, the
// temperature of the last nonlinear pass (tpass) and u_phi,0 at upFR_prev
// (up0prev).  Runs that record (n_KE > 0) keep them: before a step's last
// pass (tpass), and for IMEX at the step's start (up0prev, steptime.F90:212).
__global__ void k_keep(const RunDesc* __restrict__ runs, const int Nm1, const int N, const int with_up)
{
    const RunDesc& R = runs[blockIdx.y];
    if (!R.keep) return;
    const size_t S = (size_t)Nm1 * N;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < S; e += (size_t)gridDim.x * blockDim.x) {
        R.tpass[e] = R.tFR[e];
        if (with_up && e % (size_t)Nm1 == 0) R.up0prev[e / (size_t)Nm1] = R.upFR[e];
    }
}

// d/dr of chebtransform(u_phi) for every mode into sy (upFC's derivative,
// fourierloop.F90:226, fourierloop_imexrk.F90:390, 639) and, for m = 0, the
// second derivative into sy + S.
template <int NT, int MM>
__global__ __launch_bounds__(NT) void k_budget_dr(const RunDesc* __restrict__ runs, const Shape sh0, const int nrun)
{
    RADIAL_SETUP;
    cheb_fwd<NT, MM>(buf, V0, sh, [&](int i) { return R.upFR[at(i)]; });
    cheb_beta<NT, MM>(V1, sh, [&](int k) { return V0[k]; });
    cheb_eval_beta<NT, MM>(buf, V1, sh, [&](int i, double2 v) { R.sy[at(i)] = v; });
    if (m == 0) {
        cheb_beta<NT, MM>(V2, sh, [&](int k) { return V1[k]; });
        cheb_eval_beta<NT, MM>(buf, V2, sh, [&](int i, double2 v) { R.sy[S + at(i)] = v; });
    }
}

// The mean-flow energy, power budget and mean-flow balance of
// writeKE_spectral (output.F90:122-324) for one run per workgroup:
// out[0..3] = {2 pi radInt(<u_phi>^2), viscous dissipation (stress free:
// less the wall terms 8 pi sum_m |up_m^2|), 2 pi Ra/Pr radInt(<u_r tt>), t},
// out[4 + 3 j + c] the upbalance_ columns of level j.  Azimuthal means by
// Parseval over the modes (Np = 3 Nm > 2 Nm: exact for the products):
// <x> = Re x_0, <x y> = Re x_0 Re y_0 + 2 sum_{m>=1} Re(x_m conj y_m).
template <int NT, int MM>
__global__ __launch_bounds__(NT) void k_budget(const RunDesc* __restrict__ runs, const Shape sh0, double* __restrict__ out)
{
    extern __shared__ double2 smem[];
    const Shape sh = lds_twiddles<NT>(sh0, smem);
    const int N = sh.N, Nm1 = sh.Nm1;
    const RunDesc& R = runs[blockIdx.x];
    double2* buf = smem + sh.M;
    double2* V0 = buf + sh.M + 4 * N;
    double2* V1 = V0 + N;
    double2* V2 = V1 + N;
    double* red = (double*)(V2 + N);
    const GPtr<const double> rr = {sh.rr};
    const GPtr<const double> rr2 = {sh.rr2};
    const int tid = threadIdx.x;
    const size_t S = (size_t)Nm1 * N;
    double* o = out + (size_t)blockIdx.x * (4 + 3 * (size_t)N);
    __shared__ double wall_up2[2];
    for (int j = tid; j < N; j += NT) {
        double ut = 0.0, w2 = 0.0, up2 = 0.0;
        double2 su = mk(0.0, 0.0), sd = mk(0.0, 0.0);
        for (int q = 0; q < Nm1; ++q) {
            const size_t e = (size_t)q + (size_t)Nm1 * (size_t)j;
            const double2 ur = R.urFR[e], up = R.upFR[e], om = R.omgFR[e], tp = R.tpass[e], du = R.sy[e];
            const double2 u2 = cmul(up, up);
            up2 = up2 + hypot(u2.x, u2.y);
            if (q == 0) {
                ut = ut + ur.x * tp.x;
                w2 = w2 + cmul(om, om).x;
                su = cadd(su, cmul(ur, up));
                sd = cadd(sd, cmul(ur, du));
            } else {
                ut = ut + 2.0 * (ur.x * tp.x + ur.y * tp.y);
                w2 = w2 + cadd(cmul(om, mk(om.x, -om.y)), cmul(mk(om.x, -om.y), om)).x;
                su = cadd(su, cadd(cmul(ur, mk(up.x, -up.y)), cmul(mk(ur.x, -ur.y), up)));
                sd = cadd(sd, cadd(cmul(ur, mk(du.x, -du.y)), cmul(mk(ur.x, -ur.y), du)));
            }
        }
        const double2 u0 = R.upFR[(size_t)Nm1 * j];
        const double ua = u0.x;
        V0[j] = mk((ua * ua) * sh.rad[j], w2 * sh.rad[j]);
        V1[j] = mk(ut * sh.rad[j], 0.0);
        if (j == 0) wall_up2[0] = up2;
        if (j == N - 1) wall_up2[1] = up2;
        const double2 ufa = cadd(sd, csc(rr[j], su));
        const double2 d1 = R.sy[(size_t)Nm1 * j], d2 = R.sy[S + (size_t)Nm1 * j];
        const double2 ud = cadd(csub(d2, csc(rr2[j], u0)), csc(rr[j], d1));
        const double2 u0p = R.up0prev[j];
        o[4 + 3 * j] = (u0.x - u0p.x) / R.dt;
        o[4 + 3 * j + 1] = -ufa.x;
        o[4 + 3 * j + 2] = ud.x;
    }
    __syncthreads();
    cheb_fwd<NT, MM>(buf, V2, sh, [&](int i) { return V0[i]; });
    double a = 0.0, b = 0.0;
    for (int i = tid; i < N; i += NT) {
        const double Qm = (i % 2 == 0) ? -1.0 / (double)(i * i - 1) : 0.0;
        const double wgt = (i == 0 || i == N - 1) ? 0.5 : 1.0;
        a = a + wgt * V2[i].x * Qm;
        b = b + wgt * V2[i].y * Qm;
    }
    cheb_fwd<NT, MM>(buf, V2, sh, [&](int i) { return V1[i]; });
    double c = 0.0;
    for (int i = tid; i < N; i += NT) {
        const double Qm = (i % 2 == 0) ? -1.0 / (double)(i * i - 1) : 0.0;
        const double wgt = (i == 0 || i == N - 1) ? 0.5 : 1.0;
        c = c + wgt * V2[i].x * Qm;
    }
    red[tid] = a;
    red[NT + tid] = b;
    red[2 * NT + tid] = c;
    __syncthreads();
    for (int k = NT / 2; k > 0; k >>= 1) {
        if (tid < k) {
            red[tid] = red[tid] + red[tid + k];
            red[NT + tid] = red[NT + tid] + red[NT + tid + k];
            red[2 * NT + tid] = red[2 * NT + tid] + red[2 * NT + tid + k];
        }
        __syncthreads();
    }
    if (tid == 0) {
        const double pi = 3.14159265358979323846;
        const double upbark = sh.fact * red[0], vis1 = sh.fact * red[NT], buo = sh.fact * red[2 * NT];
        o[0] = 2.0 * pi * upbark;
        o[1] = R.mbc == 1 ? 2.0 * pi * vis1 - (4.0 * pi * 2.0 * wall_up2[1] - 4.0 * pi * 2.0 * wall_up2[0])
                          : 2.0 * pi * vis1;
        o[2] = 2.0 * pi * R.Ra / R.Pr * buo;
        o[3] = R.time;
    }
}

// calculate_spectra (output.F90:327-355) for one (run, mode):
// out[m] = 4 pi radInt |ur_m^2 + up_m^2|, out[Nm1 + m] = 4 pi radInt |T_m|.
template <int NT, int MM>
__global__ __launch_bounds__(NT) void k_spectra(const RunDesc* __restrict__ runs, const Shape sh0, const int nrun,
                                                double* __restrict__ out)
{
    RADIAL_SETUP;
    for (int i = tid; i < N; i += NT) {
        const double2 ur = R.urFR[at(i)], up = R.upFR[at(i)], t = R.tFR[at(i)];
        const double2 k = cadd(cmul(ur, ur), cmul(up, up));
        V0[i] = mk(hypot(k.x, k.y) * sh.rad[i], hypot(t.x, t.y) * sh.rad[i]);
    }
    __syncthreads();
    cheb_fwd<NT, MM>(buf, V1, sh, [&](int i) { return V0[i]; });
    if (tid == 0) {
        double a = 0.0, b = 0.0;
        for (int i = 0; i < N; ++i) {
            const double Qm = (i % 2 == 0) ? -1.0 / (double)(i * i - 1) : 0.0;
            const double wgt = (i == 0 || i == N - 1) ? 0.5 : 1.0;
            a = a + wgt * V1[i].x * Qm;
            b = b + wgt * V1[i].y * Qm;
        }
        const double pi = 3.14159265358979323846;
        double* o = out + (size_t)ri * 2 * Nm1;
        o[m] = 4.0 * pi * (sh.fact * a);
        o[Nm1 + m] = 4.0 * pi * (sh.fact * b);
    }
}

__global__ void k_advance(RunDesc* __restrict__ runs, const int nrun)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrun) return;
    RunDesc& R = runs[i];
    R.head = (R.head + 1) % R.L;
    R.kstep = R.kstep + 1;
    R.time = R.time + R.dt;
}

// A plain device copy of known size (measurement support: the byte counters
// of a profiler are scaled by what they read for this kernel).  Grid-stride,
// 16-byte loads and stores.
__global__ __launch_bounds__(256) void k_copy(const double2* __restrict__ a, double2* __restrict__ b, const size_t n)
{
    // four 16-byte loads in flight per thread, then their stores
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        double2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = gload(a + i + u * stride);
#pragma unroll
        for (int u = 0; u < 4; ++u) gstore(b + i + u * stride, v[u]);
    }
    for (; i < n; i += stride) gstore(b + i, gload(a + i));
}

void launch_copy(const double* a, double* b, size_t n16, hipStream_t st)
{
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, st, (const double2*)a, (double2*)b, n16);
}

// ------------------------------------------------------------------ launchers
static size_t radial_lds(const Shape& sh) { return sizeof(double2) * (2 * (size_t)sh.M + 8 * (size_t)sh.N); }
static size_t diag_lds(const Shape& sh) { return sizeof(double2) * (2 * (size_t)sh.M + 7 * (size_t)sh.N) + 2 * 256 * sizeof(double); }
static size_t budget_lds(const Shape& sh) { return sizeof(double2) * (2 * (size_t)sh.M + 7 * (size_t)sh.N) + 3 * 64 * sizeof(double); }
static size_t nonlin_lds(const Shape& sh) { return sizeof(double2) * 2 * (size_t)sh.Np + NT * sizeof(double); }

// (threads per workgroup, radial DFT size tier) by M = 2N-2: one wavefront
// per (run, mode) at every size; the tier fixes the register row slots.
#define TIER_DISPATCH(M, BODY)                                                        \
    do {                                                                              \
        if ((M) <= 128) { constexpr int B_ = 64; constexpr int T_ = 128; BODY; }      \
        else if ((M) <= 256) { constexpr int B_ = 64; constexpr int T_ = 256; BODY; } \
        else if ((M) <= 512) { constexpr int B_ = 64; constexpr int T_ = 512; BODY; } \
        else { constexpr int B_ = 64; constexpr int T_ = 1024; BODY; }                \
    } while (0)

#define NP_DISPATCH(NP, BODY)                       \
    do {                                            \
        if ((NP) <= 768) { constexpr int T_ = 768; BODY; }       \
        else if ((NP) <= 2304) { constexpr int T_ = 2304; BODY; }\
        else { constexpr int T_ = 4608; BODY; }                  \
    } while (0)

// Allow each kernel the whole 160 KiB of LDS (minus its static LDS) as
// dynamic LDS.  A runtime that refuses the attribute keeps its default limit;
// the launchers never ask for more than 160 KiB in total.
static void allow_lds(const void* fn)
{
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, fn) != hipSuccess) { (void)hipGetLastError(); return; }
    const int want = 163840 - (int)a.sharedSizeBytes;
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, want) != hipSuccess) (void)hipGetLastError();
}

int kernels_init()
{
#define SETLDS(fn) allow_lds((const void*)(fn));
#define SETNL(G, T) SETLDS((k_nonlin<768, G, T>)) SETLDS((k_nonlin<2304, G, T>)) SETLDS((k_nonlin<4608, G, T>))
    SETNL(false, 256) SETNL(true, 256) SETNL(false, 128) SETNL(true, 128)
#undef SETNL
    SETLDS(k_forfft<768>) SETLDS(k_forfft<2304>) SETLDS(k_forfft<4608>)
#define SETLDS4(k) SETLDS((k<64, 128>)) SETLDS((k<64, 256>)) SETLDS((k<64, 512>)) SETLDS((k<64, 1024>))
    SETLDS4(k_imex_prep) SETLDS4(k_imex_restart) SETLDS4(k_rk_init) SETLDS4(k_rk_asm) SETLDS4(k_diag)
    SETLDS4(k_rkx_init) SETLDS4(k_rkx_update) SETLDS4(k_rkx_rhs) SETLDS4(k_budget_dr) SETLDS4(k_budget)
    SETLDS4(k_spectra)
#undef SETLDS4
    SETLDS(k_lu) SETLDS((k_solve<0, false>)) SETLDS((k_solve<1, false>)) SETLDS((k_solve<0, true>))
    SETLDS((k_solve<1, true>)) SETLDS(k_m0)
#define SETTILE(k) SETLDS((k<4, true, false>)) SETLDS((k<8, true, false>))
#define SETTILEPF(k) SETTILE(k) SETLDS((k<4, true, true>)) SETLDS((k<8, true, true>)) \
    SETLDS((k<4, false, true>)) SETLDS((k<8, false, true>))
    SETTILEPF(k_imex_pre) SETTILEPF(k_imex_mid) SETTILEPF(k_imex_post) SETTILEPF(k_rk_pre) SETTILEPF(k_rk_mid)
    SETTILE(k_rk_post)
#undef SETTILEPF
#undef SETTILE
#undef SETLDS
    return 0;
}

void launch_nonlin(const RunDesc* runs, int nrun, const Shape& sh, hipStream_t st)
{
    const int total = nrun * sh.N;
    // the direct-DFT pass only in the instantiations of lengths that need it
#define NL_LAUNCH(G, TPB)                                                                                          \
    NP_DISPATCH(sh.Np, hipLaunchKernelGGL((k_nonlin<T_, G, TPB>), dim3(padded_grid(total)), dim3(TPB), nonlin_lds(sh), \
                                          st, runs, sh, nrun))
    const bool gen = sh.pq.rg > 1;
    // 128 threads only up to Np = 2304: beyond it a thread's share of the
    // level would not fit its registers
    if (sh.nl_tpb == 128 && sh.Np <= 2304) {
        if (gen) NL_LAUNCH(true, 128);
        else NL_LAUNCH(false, 128);
    } else {
        if (gen) NL_LAUNCH(true, 256);
        else NL_LAUNCH(false, 256);
    }
#undef NL_LAUNCH
}

void launch_forfft(const double* x, double2* f, const Shape& sh, hipStream_t st)
{
    NP_DISPATCH(sh.Np, hipLaunchKernelGGL(k_forfft<T_>, dim3(sh.N), dim3(NT), sizeof(double2) * (size_t)sh.Np, st, x,
                                          f, sh));
}

// The (run, mode) radial kernels: one wavefront per work item.
#define RADIAL_LAUNCH(KERNEL, ...)                                                                         \
    TIER_DISPATCH(sh.M, hipLaunchKernelGGL((KERNEL<B_, T_>), dim3(padded_grid(nrun * sh.Nm1)), dim3(B_),   \
                                           radial_lds(sh), st, runs, sh, nrun, ##__VA_ARGS__))

// The column-tile kernels: one workgroup of sh.tile_waves[KI] wavefronts per
// (run, block of 16 modes); with sh.tile_stage[KI] the NX right-hand blocks
// the kernel reads are staged in LDS where they fit in one CU's 160 KiB.
size_t tile_stage_bytes(const Shape& sh, int nx)
{
    return (size_t)nx * 4 * ((sh.N + 3) / 4) * STG_PITCH * sizeof(double2);
}

// With NPF > 0 the kernel's epilogue inputs are prefetched into LDS (PF)
// where the wavefronts' regions (NPF x 4 KiB each) fit beside the staged
// blocks.
#define TILE_LAUNCH_TW(KERNEL, TW_, NPF, ...)                                                                    \
    do {                                                                                                         \
        const size_t pb_ = (size_t)(TW_) * (NPF) * 4096;                                                         \
        if constexpr ((NPF) > 0) {                                                                               \
            if (pfo_ && stg_ && lb_ + pb_ <= 163840) {                                                                   \
                hipLaunchKernelGGL((KERNEL<TW_, true, true>), g_, dim3(64 * (TW_)), lb_ + pb_, st, runs, sh, nrun, \
                                   ##__VA_ARGS__);                                                               \
                break;                                                                                           \
            }                                                                                                    \
            if (pfo_ && !stg_ && pb_ <= 163840) {                                                                        \
                hipLaunchKernelGGL((KERNEL<TW_, false, true>), g_, dim3(64 * (TW_)), pb_, st, runs, sh, nrun,     \
                                   ##__VA_ARGS__);                                                               \
                break;                                                                                           \
            }                                                                                                    \
        }                                                                                                        \
        if (stg_)                                                                                                \
            hipLaunchKernelGGL((KERNEL<TW_, true, false>), g_, dim3(64 * (TW_)), lb_, st, runs, sh, nrun,         \
                               ##__VA_ARGS__);                                                                   \
        else                                                                                                     \
            hipLaunchKernelGGL((KERNEL<TW_, false, false>), g_, dim3(64 * (TW_)), 0, st, runs, sh, nrun,          \
                               ##__VA_ARGS__);                                                                   \
    } while (0)
#define TILE_LAUNCH(KERNEL, KI, NX, NPF, ...)                                                                    \
    do {                                                                                                         \
        const int tiles_ = nrun * ((sh.Nm1 + 15) / 16);                                                          \
        const bool spl_ = ((TILE_SPLIT_MASK >> (KI)) & 1) && sh.tile_waves[KI] == 4;                            \
        const dim3 g_(padded_grid(tiles_ * tile_split(spl_, tiles_)));                                           \
        const size_t lb_ = tile_stage_bytes(sh, NX);                                                             \
        const bool stg_ = sh.tile_stage[KI] && lb_ <= 163840;                                                   \
        const bool pfo_ = sh.tile_pf[KI] != 0;                                                                   \
        if (sh.tile_waves[KI] == 4) TILE_LAUNCH_TW(KERNEL, 4, NPF, ##__VA_ARGS__);                               \
        else TILE_LAUNCH_TW(KERNEL, 8, NPF, ##__VA_ARGS__);                                                      \
    } while (0)


void launch_imex_prep(const RunDesc* runs, int nrun, const Shape& sh, hipStream_t st)
{
    RADIAL_LAUNCH(k_imex_prep);
}

void launch_imex_pre(const RunDesc* runs, int nrun, const Shape& sh, hipStream_t st)
{
    TILE_LAUNCH(k_imex_pre, 0, 2, 3);
}

void launch_imex_mid(const RunDesc* runs, int nrun, const Shape& sh, hipStream_t st)
{
    TILE_LAUNCH(k_imex_mid, 1, 3, 4);
}

void launch_imex_post(const RunDesc* runs, int nrun, const Shape& sh, hipStream_t st)
{
    TILE_LAUNCH(k_imex_post, 2, 2, 1);
}

void launch_imex_restart(const RunDesc* runs, int nrun, const Shape& sh, hipStream_t st, int ki, int ke)
{
    RADIAL_LAUNCH(k_imex_restart, ki, ke);
}

void launch_rk_pre(const RunDesc* runs, int nrun, const Shape& sh, int s, hipStream_t st)
{
    TILE_LAUNCH(k_rk_pre, 3, 2, 5, s);
}

void launch_rk_mid(const RunDesc* runs, int nrun, const Shape& sh, int s, hipStream_t st)
{
    TILE_LAUNCH(k_rk_mid, 4, 1, 1, s);
}

void launch_rk_post(const RunDesc* runs, int nrun, const Shape& sh, int s, hipStream_t st)
{
    TILE_LAUNCH(k_rk_post, 5, 2, 0, s);
}

void launch_rk_asm(const RunDesc* runs, int nrun, const Shape& sh, hipStream_t st)
{
    RADIAL_LAUNCH(k_rk_asm);
}

void launch_asm_LAP(double* A, const double* t, const double* D, const double* D2, const Shape& sh, int m0, int count,
                    hipStream_t st)
{
    hipLaunchKernelGGL(k_asm_LAP, dim3(sh.N, count), dim3(NT), 0, st, A, t, D, D2, sh, m0);
}

void launch_rk_init(const RunDesc* runs, int nrun, const Shape& sh, hipStream_t st)
{
    RADIAL_LAUNCH(k_rk_init);
}

void launch_rkx_init(const RunDesc* runs, int nrun, const Shape& sh, hipStream_t st)
{
    RADIAL_LAUNCH(k_rkx_init);
}

void launch_rkx_update(const RunDesc* runs, int nrun, const Shape& sh, int s, hipStream_t st)
{
    RADIAL_LAUNCH(k_rkx_update, s);
}

void launch_rkx_rhs(const RunDesc* runs, int nrun, const Shape& sh, int s, hipStream_t st)
{
    RADIAL_LAUNCH(k_rkx_rhs, s);
}

// LDS of k_solve: the two 2N vectors of dgetrs, or the staged right-hand
// sides of the widest chunk (w real columns) of a batched phase.
size_t solve_lds(const Shape& sh, int w)
{
    const size_t k4 = 4 * (((size_t)sh.N + 3) / 4);
    return std::max(sizeof(double2) * 4 * (size_t)sh.N, sizeof(double) * k4 * (size_t)w);
}

void launch_m0(const RunDesc* runs, int nrun, const Shape& sh, hipStream_t st)
{
    hipLaunchKernelGGL(k_m0, dim3(padded_grid(nrun)), dim3(NT), m0_lds(sh), st, runs, sh, nrun);
}

void launch_solve(const RunDesc* runs, int nrun, const SolvePlan& pl, const Shape& sh, int phase, int lu,
                  int stage, hipStream_t st, int m0n)
{
    const int b0 = (m0n + 7) & ~7;
    int total;
    if (lu) total = nrun * sh.Nm1;
    else if (phase == 0) total = pl.nchT * sh.Nm * solve_groups(sh.N, 0);
    else total = pl.nchF * sh.Nm1 * solve_groups(sh.N, 1);
    size_t lds = solve_lds(sh, lu ? 0 : (phase == 0 ? pl.wT : pl.wF));
    if (m0n > 0) lds = std::max(lds, sizeof(double2) * (size_t)sh.N + sizeof(double) * (64 * (size_t)sh.N + 128));
    // the dgetrs-order solves in a kernel of their own, so that their
    // registers do not cost the batched inverse products occupancy
    const dim3 g(b0 + padded_grid(total));
    if (lu) {
        if (phase == 0) hipLaunchKernelGGL((k_solve<0, true>), g, dim3(NT), lds, st, runs, nrun, pl, sh, stage, 0);
        else hipLaunchKernelGGL((k_solve<1, true>), g, dim3(NT), lds, st, runs, nrun, pl, sh, stage, 0);
    } else {
        if (phase == 0) hipLaunchKernelGGL((k_solve<0, false>), g, dim3(NT), lds, st, runs, nrun, pl, sh, stage, m0n);
        else hipLaunchKernelGGL((k_solve<1, false>), g, dim3(NT), lds, st, runs, nrun, pl, sh, stage, 0);
    }
}

bool launch_solve_T_m0(const RunDesc* runs, int nrun, const SolvePlan& pl, const Shape& sh, int stage, hipStream_t st)
{
    if (sh.N > M0_FUSE_MAXN) return false;
    launch_solve(runs, nrun, pl, sh, 0, 0, stage, st, nrun);
    return true;
}

void launch_asm_T(double* A, const double* t, const double* D, const double* D2, const Shape& sh, int m0, int count,
                  double wtdt, double Pr, hipStream_t st)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_asm_T, dim3(sh.N, count), dim3(NT), 0, st, A, t, D, D2, sh, wtdt, Pr, m0);
}

void launch_asm_F(double* A, const double* t, const double* D, const double* D2, const Shape& sh, int m0, int count,
                  double wtdt, int mBC, int irk, hipStream_t st)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_asm_F, dim3(sh.N, count), dim3(NT), 0, st, A, t, D, D2, sh, wtdt, mBC, irk, m0);
}

void launch_asm_U(double* A, const double* t, const double* D, const double* D2, const Shape& sh, double wtdt, int mBC,
                  hipStream_t st)
{
    hipLaunchKernelGGL(k_asm_U, dim3(sh.N), dim3(NT), 0, st, A, t, D, D2, sh, wtdt, mBC);
}

// (row slots, right-hand sides, prefetch depth) of k_inv by matrix order n
#define INV_DISPATCH(n, BODY)                                                                              \
    do {                                                                                                   \
        if ((n) <= 128) { constexpr int S_ = 2, G_ = 16, D_ = 8; BODY; }                                  \
        else if ((n) <= 256) { constexpr int S_ = 4, G_ = 8, D_ = 8; BODY; }                              \
        else if ((n) <= 512) { constexpr int S_ = 8, G_ = 4, D_ = 4; BODY; }                              \
        else { constexpr int S_ = 17, G_ = 2, D_ = 2; BODY; }                                             \
    } while (0)

void launch_inv(const double* LU, const int* perm, int n, size_t stride, int nmat, int col0, int ncols, double* out,
                size_t ostride, hipStream_t st)
{
    if (nmat <= 0 || ncols <= 0) return;
    INV_DISPATCH(n, hipLaunchKernelGGL((k_inv<S_, G_, D_>), dim3((ncols + G_ - 1) / G_, nmat), dim3(64), 0, st, LU, perm,
                                       n, stride, col0, ncols, out, ostride));
}

void launch_lu(double* A, int* perm, int n, size_t stride, int nmat, hipStream_t st)
{
    if (nmat <= 0) return;
    const size_t lds = sizeof(double) * (size_t)n * LU_NB;
    hipLaunchKernelGGL(k_lu, dim3(nmat), dim3(NT), lds, st, A, perm, n, stride);
}

void launch_diag(const RunDesc* runs, int nrun, const Shape& sh, int record, double* out6, hipStream_t st)
{
    TIER_DISPATCH(sh.M, hipLaunchKernelGGL((k_diag<B_, T_>), dim3(nrun), dim3(B_), diag_lds(sh), st, (RunDesc*)runs, sh, record, out6));
}

void launch_keep(const RunDesc* runs, int nrun, const Shape& sh, int with_up, hipStream_t st)
{
    const int blocks = (int)std::min<size_t>(((size_t)sh.Nm1 * sh.N + 255) / 256, 1024);
    hipLaunchKernelGGL(k_keep, dim3(blocks, nrun), dim3(256), 0, st, runs, sh.Nm1, sh.N, with_up);
}

void launch_budget(const RunDesc* runs, int nrun, const Shape& sh, double* out, hipStream_t st)
{
    RADIAL_LAUNCH(k_budget_dr);
    TIER_DISPATCH(sh.M, hipLaunchKernelGGL((k_budget<B_, T_>), dim3(nrun), dim3(B_), budget_lds(sh), st, runs, sh, out));
}

// The allocation-free budget entry point stages k_budget's outputs (4 + 3 Nr
// doubles per run) in this file-scope device array instead of a buffer
// allocated per call.  Each GPU holds its own copy until the process ends;
// host code uses only the address budget_scratch() resolves per device, never
// the array's name.
constexpr size_t BUDGET_SCRATCH_DOUBLES = 4 + 3 * (size_t)MAX_N;
static __device__ __align__(256) double scratch[BUDGET_SCRATCH_DOUBLES];

int budget_scratch(double** out, size_t* cap)
{
    static std::mutex mu;
    static double* ptrs[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return -1;
    std::lock_guard<std::mutex> lk(mu);
    if (!ptrs[dev]) {
        // with the device current: its own copy of the array
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return -1;
        }
        ptrs[dev] = (double*)p;
    }
    *out = ptrs[dev];
    *cap = BUDGET_SCRATCH_DOUBLES;
    return 0;
}

void launch_spectra(const RunDesc* runs, int nrun, const Shape& sh, double* out, hipStream_t st)
{
    RADIAL_LAUNCH(k_spectra, out);
}

void launch_advance(RunDesc* runs, int nrun, hipStream_t st)
{
    hipLaunchKernelGGL(k_advance, dim3((nrun + 63) / 64), dim3(64), 0, st, runs, nrun);
}

}  // namespace fdns

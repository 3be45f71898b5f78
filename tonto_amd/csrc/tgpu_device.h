// Device-side building blocks shared by the per-class translation units and the host
// translation unit of libtonto_gpu (see tgpu_host.hip for the data layout).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/tonto_gpu.h"
#include "gen/classes.inc"
#include "gen/sph.inc"

#define RYS_QUAL static __device__ const __attribute__((aligned(16)))
#include "rys_table.inc"

#define TWOPI5ON2 34.9868366552497256925  // include/macros.in:446

// Rys coefficient loads of the lane-per-quartet kernels: batches of about this many
// (root, weight) pairs issued together (0: one step of the recurrence at a time)
#ifndef TGPU_RYS_BATCH
#define TGPU_RYS_BATCH 7
#endif

// ----------------------------------------------------------------------------------
// Device helpers
// ----------------------------------------------------------------------------------
// One primitive pair of a shell pair: exponent sum, its inverse, prefactor and product
// centre, packed in one 64-byte record so that a lane gathering its ket's primitive pair
// touches one cache line (one scalar load sequence for a wavefront-uniform bra).
struct __attribute__((aligned(64))) PrimRec {
    double zeta, zinv, K, Px, Py, Pz, pad0, pad1;
};

struct DevPrims {
    const PrimRec *rec;
};

// Load from memory that no kernel of the build writes (pair and primitive tables) through
// the constant address space: a wave-uniform address then becomes a scalar load into
// SGPRs, a per-lane address stays an ordinary vector load.
template <typename T>
__device__ __forceinline__ T ldc(const T *p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const __attribute__((address_space(4))) T *)p;
#else
    return *p;  // the host pass never runs device functions
#endif
}

// One shell pair as the integral code sees it: a run of n primitive pairs starting at
// off in the primitive arrays (DevPrims), and the pair's A and A - B.  A 32-bit offset
// rather than six pointers keeps the per-lane state small.
struct PairView {
    int off, n;
    double Ax, Ay, Az, ABx, ABy, ABz;
};

// Rys roots (u = t^2) and weights, FP64: Chebyshev pieces of width RYS_H below
// rys_xasym[N], Gauss-Hermite limit above (table made by gen_rys_table.py).
template <int N>
__device__ __forceinline__ void rys_roots(double X, double *u, double *w)
{
    if (X >= rys_xasym[N]) {
        const double ix = 1.0 / X, isx = sqrt(ix);
#pragma unroll
        for (int i = 0; i < N; ++i) {
            u[i] = rys_herm_s2[(N - 1) * RYS_NMAX + i] * ix;
            w[i] = rys_herm_w[(N - 1) * RYS_NMAX + i] * isx;
        }
        return;
    }
    const int k = (int)(X * (1.0 / RYS_H));
    const double t = (X - k * RYS_H) * (2.0 / RYS_H) - 1.0, tt = 2.0 * t;
    const double *c = rys_table_data + rys_table_offset[N] + k * (RYS_NCHEB * 2 * N);
    double b1[2 * N], b2[2 * N];
#pragma unroll
    for (int f = 0; f < 2 * N; ++f) {
        b1[f] = 0.0;
        b2[f] = 0.0;
    }
#pragma unroll 2
    for (int m = RYS_NCHEB - 1; m >= 1; --m) {
#pragma unroll
        for (int f = 0; f < 2 * N; ++f) {
            const double b0 = tt * b1[f] - b2[f] + c[m * 2 * N + f];
            b2[f] = b1[f];
            b1[f] = b0;
        }
    }
#pragma unroll
    for (int f = 0; f < N; ++f) {
        u[f] = t * b1[f] - b2[f] + c[f];
        w[f] = t * b1[N + f] - b2[N + f] + c[N + f];
    }
}

// Root r alone (u_r, w_r) of the n-root rule: two Clenshaw chains instead of 2N, for
// code where each lane needs only its own root.
template <int N>
__device__ __forceinline__ void rys_root_one(double X, int r, double &u, double &w)
{
    if (X >= rys_xasym[N]) {
        const double ix = 1.0 / X;
        u = rys_herm_s2[(N - 1) * RYS_NMAX + r] * ix;
        w = rys_herm_w[(N - 1) * RYS_NMAX + r] * sqrt(ix);
        return;
    }
    const int k = (int)(X * (1.0 / RYS_H));
    const double t = (X - k * RYS_H) * (2.0 / RYS_H) - 1.0, tt = 2.0 * t;
    const double *c = rys_table_data + rys_table_offset[N] + k * (RYS_NCHEB * 2 * N);
    // the coefficients are loaded half a chain at a time, each half before the
    // recurrence uses any of it, so their latencies overlap instead of adding up
    constexpr int H = RYS_NCHEB / 2;
    double cu[H], cw[H];
    double bu1 = 0.0, bu2 = 0.0, bw1 = 0.0, bw2 = 0.0;
#pragma unroll
    for (int h = 1; h >= 0; --h) {
#pragma unroll
        for (int i = 0; i < H; ++i) {
            cu[i] = c[(h * H + i) * 2 * N + r];
            cw[i] = c[(h * H + i) * 2 * N + N + r];
        }
#pragma unroll
        for (int i = H - 1; i >= (h ? 0 : 1); --i) {
            const double bu0 = tt * bu1 - bu2 + cu[i];
            const double bw0 = tt * bw1 - bw2 + cw[i];
            bu2 = bu1;
            bu1 = bu0;
            bw2 = bw1;
            bw1 = bw0;
        }
    }
    u = t * bu1 - bu2 + cu[0];
    w = t * bw1 - bw2 + cw[0];
}

// Same evaluation with the n-root coefficient block passed in (LDS copy or global):
// tab points at rys_table_data + rys_table_offset[N]; pairs (root, weight) are read as
// 16-byte vectors.
template <int N>
__device__ __forceinline__ void rys_roots_tab(double X, double *u, double *w, const double *__restrict__ tab)
{
    if (X >= rys_xasym[N]) {
        const double ix = 1.0 / X, isx = sqrt(ix);
#pragma unroll
        for (int i = 0; i < N; ++i) {
            u[i] = rys_herm_s2[(N - 1) * RYS_NMAX + i] * ix;
            w[i] = rys_herm_w[(N - 1) * RYS_NMAX + i] * isx;
        }
        return;
    }
    const int k = (int)(X * (1.0 / RYS_H));
    const double t = (X - k * RYS_H) * (2.0 / RYS_H) - 1.0, tt = 2.0 * t;
    const double2 *c = reinterpret_cast<const double2 *>(tab + k * (RYS_NCHEB * 2 * N));
    double b1[2 * N], b2[2 * N];
#pragma unroll
    for (int f = 0; f < 2 * N; ++f) {
        b1[f] = 0.0;
        b2[f] = 0.0;
    }
#if TGPU_RYS_BATCH
    // coefficients m = 13 .. 1 in batches of G, each batch's loads issued before the
    // recurrence uses any of them, so that their latencies overlap (2N*G doubles in
    // flight)
    constexpr int G = N == 1 ? TGPU_RYS_BATCH : (TGPU_RYS_BATCH + N - 1) / N;
#pragma unroll
    for (int m0 = RYS_NCHEB - 1; m0 >= 1; m0 -= G) {
        double2 cb[G][N];
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int h = 0; h < N; ++h)
                if (m0 - g >= 1) cb[g][h] = c[(m0 - g) * N + h];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if (m0 - g < 1) break;
#pragma unroll
            for (int h = 0; h < N; ++h) {
                const double b0x = tt * b1[2 * h] - b2[2 * h] + cb[g][h].x;
                const double b0y = tt * b1[2 * h + 1] - b2[2 * h + 1] + cb[g][h].y;
                b2[2 * h] = b1[2 * h];
                b1[2 * h] = b0x;
                b2[2 * h + 1] = b1[2 * h + 1];
                b1[2 * h + 1] = b0y;
            }
        }
    }
#else
#pragma unroll 2
    for (int m = RYS_NCHEB - 1; m >= 1; --m) {
#pragma unroll
        for (int h = 0; h < N; ++h) {
            const double2 cm = c[m * N + h];
            const double b0x = tt * b1[2 * h] - b2[2 * h] + cm.x;
            const double b0y = tt * b1[2 * h + 1] - b2[2 * h + 1] + cm.y;
            b2[2 * h] = b1[2 * h];
            b1[2 * h] = b0x;
            b2[2 * h + 1] = b1[2 * h + 1];
            b1[2 * h + 1] = b0y;
        }
    }
#endif
#pragma unroll
    for (int h = 0; h < N; ++h) {
        const double2 c0 = c[h];
        const double v0 = t * b1[2 * h] - b2[2 * h] + c0.x;
        const double v1 = t * b1[2 * h + 1] - b2[2 * h + 1] + c0.y;
        // the 2N functions are stored roots first, then weights
        if (2 * h < N) u[2 * h] = v0; else w[2 * h - N] = v0;
        if (2 * h + 1 < N) u[2 * h + 1] = v1; else w[2 * h + 1 - N] = v1;
    }
}

// number of doubles in the n-root block of the table
__host__ __device__ constexpr int rys_block_size(int n)
{
    return (n == 1 ? 20 : n == 2 ? 24 : n == 3 ? 27 : n == 4 ? 31 : n == 5 ? 34 : n == 6 ? 37 : 39) *
           RYS_NCHEB * 2 * n;
}

static __device__ void rys_roots_rt(int n, double X, double *u, double *w)
{
    switch (n) {
    case 1: rys_roots<1>(X, u, w); break;
    case 2: rys_roots<2>(X, u, w); break;
    case 3: rys_roots<3>(X, u, w); break;
    case 4: rys_roots<4>(X, u, w); break;
    case 5: rys_roots<5>(X, u, w); break;
    case 6: rys_roots<6>(X, u, w); break;
    default: rys_roots<7>(X, u, w); break;
    }
}

static constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
static constexpr int pair_class(int lp, int lq) { return lp * (lp + 1) / 2 + lq; }
// functions of an l-shell in Tonto's spherical basis: real solid harmonics for l >= 2,
// the Cartesian functions for s and p (shell1quartet.foo:505-563)
__host__ __device__ constexpr int nsph(int l) { return l >= 2 ? 2 * l + 1 : ncart(l); }

// coefficient of Cartesian component c (Tonto order) in solid harmonic m of an l-shell
__device__ __forceinline__ double sph_t(int l, int c, int m)
{
    return l == 2 ? tgpu_sph_T2[c * 5 + m] : l == 3 ? tgpu_sph_T3[c * 7 + m] : (c == m ? 1.0 : 0.0);
}

// (ab|ab) in the spherical basis for solid harmonics (alpha of shell P, beta of shell Q)
// from the Cartesian block I[((p NQ + q) NP + p') NQ + q'] of the pair (P Q|P Q)
template <int LP, int LQ>
__device__ __forceinline__ double sph_abab(const double *I, int alpha, int beta)
{
    constexpr int NP = ncart(LP), NQ = ncart(LQ);
    double v = 0.0;
    for (int p1 = 0; p1 < NP; ++p1) {
        const double t1 = sph_t(LP, p1, alpha);
        if (t1 == 0.0) continue;
        for (int q1 = 0; q1 < NQ; ++q1) {
            const double t2 = t1 * sph_t(LQ, q1, beta);
            if (t2 == 0.0) continue;
            for (int p2 = 0; p2 < NP; ++p2) {
                const double t3 = t2 * sph_t(LP, p2, alpha);
                if (t3 == 0.0) continue;
                for (int q2 = 0; q2 < NQ; ++q2)
                    v += t3 * sph_t(LQ, q2, beta) * I[((p1 * NQ + q1) * NP + p2) * NQ + q2];
            }
        }
    }
    return v;
}

template <int LP, int LQ, int LR, int LS, typename R> struct EriBody;
template <int LP, int LQ, int LR, int LS> struct WaveTab;

// Per-pair integer fields packed into two 16-byte records so that a lane fetches a
// ket pair's description with two vector loads (coalesced across the wavefront):
//   rec0 = {Tonto index, shells p<<16|q, atoms p<<16|q, first bf p<<16|q}
//   rec1 = {full-list offset, skipped-list offset, counts full<<16|skipped, molecule}
struct DevPairs {
    const int4 *rec0, *rec1;
    const double *A, *AB, *maxI;
};
// two 16-bit fields (0..65535) in one int; the upper one may set the sign bit, so it is
// packed and unpacked unsigned
__host__ __device__ __forceinline__ int pack16(int hi, int lo) { return (int)(((unsigned)hi << 16) | ((unsigned)lo & 0xffffu)); }
__host__ __device__ __forceinline__ int hi16(int v) { return (int)((unsigned)v >> 16); }
__host__ __device__ __forceinline__ int lo16(int v) { return v & 0xffff; }

__device__ __forceinline__ PairView make_view(const DevPairs &pr, const DevPrims &pm, int s,
                                              bool skip)
{
    PairView v;
    const int4 r1 = ldc(pr.rec1 + s);
    const int off = skip ? r1.y : r1.x;
    v.n = skip ? lo16(r1.z) : hi16(r1.z);
    v.off = off;
    v.Ax = ldc(pr.A + 3 * s);
    v.Ay = ldc(pr.A + 3 * s + 1);
    v.Az = ldc(pr.A + 3 * s + 2);
    v.ABx = ldc(pr.AB + 3 * s);
    v.ABy = ldc(pr.AB + 3 * s + 1);
    v.ABz = ldc(pr.AB + 3 * s + 2);
    return v;
}

__device__ __forceinline__ long long tonto_pair(int x, int y)
{
    return x >= y ? (long long)x * (x + 1) / 2 + y : (long long)y * (y + 1) / 2 + x;
}

// Pair bins: every angular-momentum class is split by the primitive count of the pair
// (full list) into TGPU_NBUCKET buckets, so that the lanes of a wavefront run primitive
// loops of similar length; bin = class * TGPU_NBUCKET + bucket.
#define TGPU_NBUCKET 4
// lane-per-quartet kernels compact the kets that pass the screening tests (1) or let
// the lanes whose ket fails idle for that round (0)
#ifndef TGPU_COMPACT
#define TGPU_COMPACT 1
#endif
#define TGPU_NBIN (TGPU_NPAIRCLASS * TGPU_NBUCKET)

// Per-molecule layout of a batch: molecule m's pairs of bin k occupy
// [seg_start[m*NBIN+k], +seg_count) of the sorted pair arrays; its matrices start at
// element mat_off[m] with leading dimension nbf[m]; its Tonto pairs start at pair_off[m]
// in the P_max array; scal[2m], scal[2m+1] are its P_max and IP_max.
struct DevMols {
    const int *seg_start, *seg_count, *nbf;
    const long long *mat_off, *pair_off;
};

struct JKArgs {
    DevPairs pr;
    DevPrims pm;
    DevMols ml;
    int bra_start, n_bra, bra_bin, ket_bin, same_bin, counter_slot, ket_chunk, n_ket_max;
    int n_ket_avg;  // ket-bin pairs of a single molecule (0 for batches)
    double *ws;    // wavefront-form workspace of ws_doubles doubles: a class uses
    long long ws_doubles;  // min(8192, ws_doubles / WaveCfg::WS) slots
    int max_nbf;   // largest molecule (basis functions)
    int krow_rows;   // KROWS form: the first krow_rows K rows of the bra pair live in LDS
    int krow_budget; // cap on a work-group's LDS in the KROWS form (bytes; 0: default)
    int krow_min_kets;  // shortest ket segment (n_ket_avg) for the KROWS form
    // KROWS form: density rows staged in LDS next to the K rows (set per launch;
    // prow_enable 0 keeps every density read global)
    int prow_rows, prow_enable;
    // Pair-ordered ket blocks (lane form): PP holds each pair's density block, XP takes
    // each pair's ket-side J (added into X after the class launches); a ket pair sk of
    // the launch's ket bin has its block at kblk_off + sk * NR * NS
    const double *PP;
    double *XP;
    long long kblk_off;
    // KROWS form, single molecules: n_shell shells; pmrow (set per launch) 1 when the bra
    // pair's two P_max rows (2 n_shell doubles) are staged in LDS after the density rows
    int n_shell, pmrow;
    // KROWS form, single molecules: the K columns a class pair touches are the functions
    // of its ket shells' two angular momenta, so an LDS row holds only those, numbered
    // within each momentum.  kcol[pair] = pack16(number of the pair's first function
    // among the functions of its first shell's momentum, the same for its second shell);
    // the functions of momentum l are kcol_map[kcol_lofs[l] .. kcol_lofs[l + 1]).  Null
    // for batches.
    const int *kcol;
    const int *kcol_map;
    int kcol_lofs[TGPU_LMAX + 2];
    // UHF pass: J from Tonto's ket side only, X(cd) += f (ab|cd) P(ab) with ab the pair of
    // larger Tonto index, for ab != cd (the updates Tonto's UHF engine puts into JA
    // through JBcd => JA(cd), molecule.fock.foo:2637)
    int j_ket_nonsame;
    // spherical basis: Tonto's make_r_JK_direct screening (Schwarz_inequality_test), one
    // decision for J and K
    int sph_screen;
    const double *P, *Pmax, *scal;
    double cutoff, jp, kp, fp32_thr;
    double *X, *Y;
    // accumulator copies (single molecules): X and Y hold TGPU_ACC_COPIES copies acc_stride
    // doubles apart and a wavefront adds into the copy of the XCD it runs on (0: one copy)
    long long acc_stride;
    unsigned long long *counters;
};

// Copies of the J/K accumulators: one per XCD (accelerator complex die).  Device-scope
// atomics from different XCDs on one cache line move the line between the XCDs' L2
// caches, one memory write per update; with a copy per XCD each copy's lines are updated
// by one XCD only.  Any copy index gives the same sums (the atomics stay device-scope),
// so the index only has to be in range.
#define TGPU_ACC_COPIES 8

// The XCD the wavefront runs on (hardware register XCC_ID), reduced to a copy index.
__device__ __forceinline__ int acc_copy()
{
    unsigned r;
    asm("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(r));
    return (int)(r & (TGPU_ACC_COPIES - 1));
}

// Everything a wavefront needs about its bra pair, and the Schwarz frontier of its kets.
struct BraInfo {
    int sb, tb, bp, bq, atp, atq, fp, fq, ld, kb, jend;
    double mIb, Pmax_b, IP_max;
    const double *P, *Pmax, *maxI;
    double *X, *Y;
};

__device__ __forceinline__ void load_bra(const JKArgs &a, int i, BraInfo &b)
{
    b.sb = a.bra_start + i;
    const int4 r0 = ldc(a.pr.rec0 + b.sb);
    const int m = ldc(a.pr.rec1 + b.sb).w;
    b.kb = ldc(a.ml.seg_start + m * TGPU_NBIN + a.ket_bin);
    const int kn = ldc(a.ml.seg_count + m * TGPU_NBIN + a.ket_bin);
    const int jlim = a.same_bin ? b.sb - ldc(a.ml.seg_start + m * TGPU_NBIN + a.bra_bin) + 1 : kn;
    const long long mo = ldc(a.ml.mat_off + m);
    b.ld = ldc(a.ml.nbf + m);
    b.P = a.P + mo;
    const long long co = a.acc_stride ? (long long)acc_copy() * a.acc_stride : 0;
    b.X = a.X + co + mo;
    b.Y = a.Y + co + mo;
    b.Pmax = a.Pmax + ldc(a.ml.pair_off + m);
    const double P_max = ldc(a.scal + 2 * m);
    b.IP_max = ldc(a.scal + 2 * m + 1);
    b.tb = r0.x;
    b.mIb = ldc(a.pr.maxI + b.sb);
    b.bp = hi16(r0.y);
    b.bq = lo16(r0.y);
    b.atp = hi16(r0.z);
    b.atq = lo16(r0.z);
    b.fp = hi16(r0.w);
    b.fq = lo16(r0.w);
    b.Pmax_b = ldc(b.Pmax + b.tb);
    b.maxI = a.pr.maxI + b.kb;
    // kets are sorted by max_I descending, so those passing
    // max_I(ab)*max_I(cd)*P_max >= cutoff (molecule.fock.foo:1430) form a prefix
    int lo = 0, hi = jlim;
    // most bra pairs far down the sorted bin pass no ket at all: one test of the largest
    // ket settles that without the search
    if (hi > 0 && b.mIb * ldc(b.maxI) * P_max < a.cutoff) hi = 0;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (b.mIb * ldc(b.maxI + mid) * P_max < a.cutoff) hi = mid;
        else lo = mid + 1;
    }
    b.jend = lo;
}

struct KetTest {
    int sk, fr, fs;
    bool skipJ, skipK, skip_lists;
    int halvings;  // permutation factor f = 2^-halvings
    double f;
};

// Tonto's remaining per-quartet tests (molecule.fock.foo:1412, 1432-1472 and
// Schwarz_test 3497-3536), evaluated in Tonto's operand order.  Returns false if the
// quartet contributes to neither J nor K.
// pmr (optional, LDS): P_max(bra shell p, x) for the molecule's nsh shells x, then
// P_max(bra shell q, x): the same values as the Tonto-pair lookups, read from LDS.
__device__ __forceinline__ bool test_ket(const JKArgs &a, const BraInfo &b, int j, KetTest &t,
                                         const double *pmr = nullptr, int nsh = 0)
{
    t.sk = b.kb + j;
    const int4 r0 = ldc(a.pr.rec0 + t.sk);
    const int tk = r0.x;
    const double mIk = ldc(b.maxI + j);
    const bool bhi = b.tb >= tk;  // Tonto's (ab| is the pair with the larger index
    const double Iab = bhi ? b.mIb : mIk, Icd = bhi ? mIk : b.mIb;
    if (Iab * b.IP_max < a.cutoff) return false;
    const int kr = hi16(r0.y), ks = lo16(r0.y);
    const double Pmax_k = ldc(b.Pmax + tk);
    const double JP = bhi ? fmax(b.Pmax_b, Pmax_k) : fmax(Pmax_k, b.Pmax_b);
    const double KP = pmr ? fmax(fmax(pmr[kr], pmr[ks]), fmax(pmr[nsh + kr], pmr[nsh + ks]))
                          : fmax(fmax(ldc(b.Pmax + tonto_pair(b.bp, kr)), ldc(b.Pmax + tonto_pair(b.bp, ks))),
                                 fmax(ldc(b.Pmax + tonto_pair(b.bq, kr)), ldc(b.Pmax + tonto_pair(b.bq, ks))));
    if (a.sph_screen) {
        // Schwarz_inequality_test (molecule.fock.foo:3454-3495): the largest of the six
        // density blocks, skip below TOL(20) or unless Pmax*I_max(ab)*I_max(cd) > cutoff
        const double Pm = fmax(JP, KP);
        const bool skip = (Pm < 1e-20) || !(Pm * Iab * Icd > a.cutoff);
        t.skipJ = t.skipK = skip;
    } else {
        t.skipJ = (JP < a.jp) || (JP * Iab * Icd < a.cutoff);
        t.skipK = (KP < a.kp) || (KP * Iab * Icd < a.cutoff);
    }
    if (t.skipJ && t.skipK) return false;
    t.halvings = (b.bp == b.bq) + (kr == ks) + (b.sb == t.sk);
    t.f = t.halvings == 0 ? 1.0 : t.halvings == 1 ? 0.5 : t.halvings == 2 ? 0.25 : 0.125;
    const int akr = hi16(r0.z), aks = lo16(r0.z);
    t.fr = hi16(r0.w);
    t.fs = lo16(r0.w);
    const bool any_same = b.atp == b.atq || akr == aks || b.atp == akr || b.atp == aks ||
                          b.atq == akr || b.atq == aks;
    t.skip_lists = !any_same;
    return true;
}

// Occupancy floor handed to the register allocator: (ss|ss) fits 8 wavefronts per SIMD
// and (pp|ss) 4 without spilling; every other class spills when pushed above what it
// reaches unconstrained, so it is left free.
__host__ __device__ constexpr int jk_min_waves(int lp, int lq, int lr, int ls)
{
#ifdef TGPU_JK_WAVES_OVERRIDE
    return TGPU_JK_WAVES_OVERRIDE;
#else
    return lp + lq + lr + ls == 0 ? 8 : (lp == 1 && lq == 1 && lr == 0 && ls == 0) ? 4 : 1;
#endif
}

// Orders LDS accesses of the lanes of one wavefront (LDS executes a wavefront's
// instructions in issue order; this keeps the compiler from moving accesses across).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ----------------------------------------------------------------------------------
// The J/K kernel for one light class pair (LP LQ | LR LS): one lane per quartet
// ----------------------------------------------------------------------------------
// KROWS: one bra pair per work-group, its four wavefronts taking interleaved rounds of the
// ket prefix; the K updates of the bra pair all fall in its NP + NQ rows, and there only
// in the columns of momentum LR and LS.  The first krow_rows rows (all, when the LDS
// budget allows), restricted to those columns, accumulate in LDS (dynamic shared memory,
// krow_rows x W doubles, W = the number of functions of momentum LR, plus those of
// momentum LS if LS != LR) and are flushed once, row by row, with atomics that skip zeros;
// the other rows are updated with global atomics.  Otherwise four bra pairs per
// work-group, one per wavefront, and every K update is a global atomic.
template <int LP, int LQ, int LR, int LS, bool MIXED, bool KROWS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(jk_min_waves(LP, LQ, LR, LS), 8))) void jk_kernel(JKArgs a)
{
    constexpr int NP = ncart(LP), NQ = ncart(LQ), NR = ncart(LR), NS = ncart(LS);
    constexpr int NPQ = NP * NQ, NRS = NR * NS, NI = NPQ * NRS;
    constexpr int NRT = EriBody<LP, LQ, LR, LS, double>::NROOT;
    // the n-root Rys table in LDS (<= 18 KB): the root/weight lookups of every primitive
    // quartet then hit LDS instead of L2
    __shared__ __attribute__((aligned(16))) double rys_lds[rys_block_size(NRT)];
#if TGPU_COMPACT
    // per-wavefront queue of kets that passed the screening tests:
    // (j << 5 | flags, fr << 16 | fs); basis functions < 65536 (tgpu_create)
    __shared__ int2 kqueue[4][128];
#endif
    extern __shared__ double krow[];  // KROWS: [NP + NQ][ld]
    const int lane = threadIdx.x & 63;
    // the bra pair is uniform over the wavefront: readfirstlane lets the compiler keep
    // everything derived from it in scalar registers
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = KROWS ? (int)blockIdx.x : (int)blockIdx.x * 4 + wave;
    BraInfo b;
    if (KROWS) {
        if (i >= a.n_bra) return;  // uniform over the work-group
        load_bra(a, i, b);
        // a bra pair with no kets left by the Schwarz prefix leaves before staging the
        // Rys table and zeroing its K rows (2.5 % of an (H2O)64/cc-pVTZ build)
        if (b.jend == 0) return;  // uniform
    }
    for (int t = threadIdx.x; t < rys_block_size(NRT); t += blockDim.x)
        rys_lds[t] = rys_table_data[rys_table_offset[NRT] + t];
    __syncthreads();
    if (!KROWS) {
        if (i >= a.n_bra) return;
        load_bra(a, i, b);
    }
    const int ld = b.ld, fp = b.fp, fq = b.fq;
    const int nkr = KROWS ? a.krow_rows : 0;  // K rows accumulated in LDS
    // LDS row: the nR functions of momentum LR, then (LS != LR) those of momentum LS
    const int nR = KROWS ? a.kcol_lofs[LR + 1] - a.kcol_lofs[LR] : 0;
    const int W = KROWS ? (LR == LS ? nR : nR + a.kcol_lofs[LS + 1] - a.kcol_lofs[LS]) : 0;
    // P rows (KROWS with (ss) kets, prefixes of at least krow_min_kets kets (256 unless
    // set), when all NP + NQ rows fit): the bra pair's density rows at the same columns as its K rows, staged in LDS
    // after them, so that the K digestion reads P(bra function, ket function) from LDS,
    // not global memory ((ps|ss) at 4160 functions: 324 -> 239 ms).  Compiled for the
    // (..|ss) class pairs only: with p or d kets the rows rarely fit twice, and the second
    // path alone slowed those kernels by 5-13 %.
    constexpr bool PR = KROWS && LR == 0 && LS == 0;
    const int npr = PR && b.jend >= a.krow_min_kets ? a.prow_rows : 0;  // 0 or NP + NQ, uniform
    double *const prow = krow + nkr * W;
    // P_max rows (KROWS, prefixes of at least half the shell count): the screening's four
    // cross-pair density maxima of each ket come from LDS
    const int nsh = a.n_shell;
    double *const pmr = KROWS && a.pmrow && 2 * b.jend >= nsh ? prow + a.prow_rows * W : nullptr;
    if (KROWS) {
        if (pmr)
            for (int x = threadIdx.x; x < nsh; x += blockDim.x) {
                pmr[x] = ldc(b.Pmax + tonto_pair(b.bp, x));
                pmr[nsh + x] = ldc(b.Pmax + tonto_pair(b.bq, x));
            }
        for (int t = threadIdx.x; t < nkr * W; t += blockDim.x) krow[t] = 0.0;
        for (int t = threadIdx.x; PR && t < npr * W; t += blockDim.x) {
            const int row = t / W, cc = t - row * W;
            const int col = ldc(a.kcol_map + (cc < nR ? a.kcol_lofs[LR] + cc : a.kcol_lofs[LS] + cc - nR));
            prow[t] = b.P[(size_t)(row < NP ? fp + row : fq + row - NP) * ld + col];
        }
        __syncthreads();
    }
    // K(row, col) += v: row = bra function fp + p (p < NP) or fq + q (NP + q); col the
    // molecule's function, ccol its LDS column
    auto kadd = [&](const int row, const int col, const int ccol, const double v) {
        if (KROWS && row < nkr) {
            __hip_atomic_fetch_add(&krow[row * W + ccol], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
            unsafeAtomicAdd(&b.Y[(size_t)(row < NP ? fp + row : fq + row - NP) * ld + col], v);
        }
    };
    const double *__restrict__ P = b.P;
    int jsplit = b.jend;
    if (MIXED) {
        int lo = 0, hi = b.jend;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (b.mIb * ldc(b.maxI + mid) < a.fp32_thr) hi = mid;
            else lo = mid + 1;
        }
        jsplit = lo;
    }
    double Ppq[NPQ], Jpq[NPQ];
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            Ppq[p * NQ + q] = P[(size_t)(fp + p) * ld + fq + q];
            Jpq[p * NQ + q] = 0.0;
        }
    unsigned long long nq = 0, npq = 0, n32 = 0;
    // one contracted quartet (bra pair, ket j of the prefix) that passed the tests
    // bv: the bra pair's view for this quartet's primitive lists
    auto quartet = [&](const int j, const int flags, const int fr, const int fs, const PairView &bv) {
        KetTest t;
        t.sk = b.kb + j;
        t.skipJ = flags & 1;
        t.skipK = flags & 2;
        t.skip_lists = flags & 4;
        t.f = (flags >> 3) == 0 ? 1.0 : (flags >> 3) == 1 ? 0.5 : (flags >> 3) == 2 ? 0.25 : 0.125;
        const PairView kv = make_view(a.pr, a.pm, t.sk, t.skip_lists);
        double I[NI];
        if (!MIXED || j < jsplit) {
            EriBody<LP, LQ, LR, LS, double>::run(a.pm, bv, kv, I, rys_lds);
        } else {
            EriBody<LP, LQ, LR, LS, float>::run(a.pm, bv, kv, I, rys_lds);
            ++n32;
        }
        ++nq;
        npq += (unsigned long long)(bv.n * kv.n);
        const double f = t.f;
        // which J sides: both, or in the UHF pass only Tonto's ket side (the pair of smaller
        // Tonto index receives), none for ab = cd
        bool j_ket = true, j_bra = true;
        if (a.j_ket_nonsame) {
            const int tk = ldc(a.pr.rec0 + t.sk).x;
            j_ket = b.tb > tk;
            j_bra = b.tb < tk;
        }
        if (!t.skipJ) {
            // the ket's density block and its ket-side J through the pair-ordered blocks:
            // neighbouring lanes hold neighbouring kets, so these loads and atomics fall on
            // few cache lines instead of one line each in P and X
            const long long kb0 = a.kblk_off + (long long)t.sk * NRS;
            double Prs[NRS];
#pragma unroll
            for (int rs = 0; rs < NRS; ++rs) Prs[rs] = a.PP[kb0 + rs];
            if (j_ket) {
#pragma unroll
                for (int rs = 0; rs < NRS; ++rs) {
                    double acc = 0.0;
#pragma unroll
                    for (int pq = 0; pq < NPQ; ++pq) acc += I[pq * NRS + rs] * Ppq[pq];
                    unsafeAtomicAdd(&a.XP[kb0 + rs], f * acc);
                }
            }
            if (j_bra) {
#pragma unroll
                for (int pq = 0; pq < NPQ; ++pq) {
                    double acc = 0.0;
#pragma unroll
                    for (int rs = 0; rs < NRS; ++rs) acc += I[pq * NRS + rs] * Prs[rs];
                    Jpq[pq] += f * acc;
                }
            }
        }
        if (!t.skipK) {
            // K(p,r) += I P(s,q); K(p,s) += I P(r,q); K(q,r) += I P(s,p); K(q,s) += I P(r,p)
            // (shell1quartet.foo:11508-11521)
            // P is symmetric: P(s,q) etc. are read as P(q,s) from the bra pair's rows,
            // which every lane of the wavefront shares (cache-resident), instead of one
            // row per lane
            // LDS columns of the ket's functions (KROWS)
            int ccr = 0, ccs = 0;
            if (KROWS) {
                const int kc = ldc(a.kcol + t.sk);
                ccr = hi16(kc);
                ccs = (LR == LS ? 0 : nR) + lo16(kc);
            }
            double Psq[NS * NQ], Prq[NR * NQ], Psp[NS * NP], Prp[NR * NP];
            if (PR && npr) {  // uniform: every row staged
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
#pragma unroll
                    for (int s = 0; s < NS; ++s) Psq[s * NQ + q] = prow[(NP + q) * W + ccs + s];
#pragma unroll
                    for (int r = 0; r < NR; ++r) Prq[r * NQ + q] = prow[(NP + q) * W + ccr + r];
                }
#pragma unroll
                for (int p = 0; p < NP; ++p) {
#pragma unroll
                    for (int s = 0; s < NS; ++s) Psp[s * NP + p] = prow[p * W + ccs + s];
#pragma unroll
                    for (int r = 0; r < NR; ++r) Prp[r * NP + p] = prow[p * W + ccr + r];
                }
            } else {
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
#pragma unroll
                    for (int s = 0; s < NS; ++s) Psq[s * NQ + q] = P[(size_t)(fq + q) * ld + fs + s];
#pragma unroll
                    for (int r = 0; r < NR; ++r) Prq[r * NQ + q] = P[(size_t)(fq + q) * ld + fr + r];
                }
#pragma unroll
                for (int p = 0; p < NP; ++p) {
#pragma unroll
                    for (int s = 0; s < NS; ++s) Psp[s * NP + p] = P[(size_t)(fp + p) * ld + fs + s];
#pragma unroll
                    for (int r = 0; r < NR; ++r) Prp[r * NP + p] = P[(size_t)(fp + p) * ld + fr + r];
                }
            }
#pragma unroll
            for (int p = 0; p < NP; ++p)
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    double acc = 0.0;
#pragma unroll
                    for (int q = 0; q < NQ; ++q)
#pragma unroll
                        for (int s = 0; s < NS; ++s)
                            acc += I[((p * NQ + q) * NR + r) * NS + s] * Psq[s * NQ + q];
                    kadd(p, fr + r, ccr + r, f * acc);
                }
#pragma unroll
            for (int p = 0; p < NP; ++p)
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    double acc = 0.0;
#pragma unroll
                    for (int q = 0; q < NQ; ++q)
#pragma unroll
                        for (int r = 0; r < NR; ++r)
                            acc += I[((p * NQ + q) * NR + r) * NS + s] * Prq[r * NQ + q];
                    kadd(p, fs + s, ccs + s, f * acc);
                }
#pragma unroll
            for (int q = 0; q < NQ; ++q)
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    double acc = 0.0;
#pragma unroll
                    for (int p = 0; p < NP; ++p)
#pragma unroll
                        for (int s = 0; s < NS; ++s)
                            acc += I[((p * NQ + q) * NR + r) * NS + s] * Psp[s * NP + p];
                    kadd(NP + q, fr + r, ccr + r, f * acc);
                }
#pragma unroll
            for (int q = 0; q < NQ; ++q)
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    double acc = 0.0;
#pragma unroll
                    for (int p = 0; p < NP; ++p)
#pragma unroll
                        for (int r = 0; r < NR; ++r)
                            acc += I[((p * NQ + q) * NR + r) * NS + s] * Prp[r * NP + p];
                    kadd(NP + q, fs + s, ccs + s, f * acc);
                }
        }
    };
#if TGPU_COMPACT
    // Compaction: each round the 64 lanes test 64 kets of the prefix; the kets that pass
    // are appended to this wavefront's queue in LDS, and the integral code runs only on
    // full groups of 64 queued kets (all lanes busy), plus one partial group at the end.
    // The queue is private to the wavefront, so ordering its LDS traffic needs no
    // work-group barrier.
    int2 *q = kqueue[threadIdx.x >> 6];
    int qn = 0;                             // wave-uniform
    constexpr int step = KROWS ? 256 : 64;  // KROWS: the four wavefronts interleave rounds
    for (int j0 = KROWS ? 64 * wave : 0;; j0 += step) {
        if (j0 < b.jend) {
            const int j = j0 + lane;
            KetTest t;
            const bool ok = j < b.jend && test_ket(a, b, j, t, pmr, nsh);
            const unsigned long long m = __ballot(ok);
            if (ok) {
                const int pos = qn + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                q[pos] = make_int2((j << 5) | (int)t.skipJ | ((int)t.skipK << 1) | ((int)t.skip_lists << 2) |
                                       (t.halvings << 3),
                                   pack16(t.fr, t.fs));
            }
            qn += __popcll(m);
            wave_lds_sync();
        }
        const bool drain = j0 + step >= b.jend;  // nothing left to test
        if (qn >= 64 || (drain && qn > 0)) {
            // take the last min(qn, 64) entries; later rounds append behind the rest
            const int take = qn < 64 ? qn : 64, base = qn - take;
            const int2 e = lane < take ? q[base + lane] : make_int2(0, 0);
            qn = base;
            wave_lds_sync();
            if (lane < take) quartet(e.x >> 5, e.x & 31, hi16(e.y), lo16(e.y),
                                     make_view(a.pr, a.pm, b.sb, e.x & 4));
        }
        if (drain && qn == 0) break;
    }
#else
    for (int j0 = KROWS ? 64 * wave : 0; j0 < b.jend; j0 += KROWS ? 256 : 64) {
        const int j = j0 + lane;
        KetTest t;
        if (j < b.jend && test_ket(a, b, j, t, pmr, nsh))
            quartet(j, (int)t.skipJ | ((int)t.skipK << 1) | ((int)t.skip_lists << 2) | (t.halvings << 3),
                    t.fr, t.fs, make_view(a.pr, a.pm, b.sb, t.skip_lists));
    }
#endif
    // wavefront reduction of the bra-pair J block
#pragma unroll
    for (int pq = 0; pq < NPQ; ++pq) {
        double v = Jpq[pq];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0 && v != 0.0)
            unsafeAtomicAdd(&b.X[(size_t)(fp + pq / NQ) * ld + fq + pq % NQ], v);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        nq += __shfl_xor(nq, off, 64);
        npq += __shfl_xor(npq, off, 64);
        n32 += __shfl_xor(n32, off, 64);
    }
    if (lane == 0 && nq) {
        atomicAdd(&a.counters[3 * a.counter_slot + 0], nq);
        atomicAdd(&a.counters[3 * a.counter_slot + 1], npq);
        atomicAdd(&a.counters[3 * a.counter_slot + 2], n32);
    }
    if (KROWS) {
        // every wavefront of the work-group is here (one bra pair, no early return)
        __syncthreads();
        for (int t = threadIdx.x; t < nkr * W; t += blockDim.x) {
            const double v = krow[t];
            const int row = t / W, cc = t - row * W;
            if (v != 0.0) {
                const int col = ldc(a.kcol_map + (cc < nR ? a.kcol_lofs[LR] + cc : a.kcol_lofs[LS] + cc - nR));
                unsafeAtomicAdd(&b.Y[(size_t)(row < NP ? fp + row : fq + row - NP) * ld + col], v);
            }
        }
    }
}

// Schwarz table: max_I = sqrt(max |(ab|ab)|) with the full primitive lists
// (molecule.fock.foo:3712-3752).  One thread per shell pair of one class bin.
template <int LP, int LQ>
__global__ __launch_bounds__(64) void schwarz_kernel(DevPairs pr, DevPrims pm, int start, int n,
                                                    double *maxI, int sph)
{
    constexpr int NP = ncart(LP), NQ = ncart(LQ);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = start + i;
    const PairView v = make_view(pr, pm, s, false);
    double I[NP * NQ * NP * NQ];
    EriBody<LP, LQ, LP, LQ, double>::run(pm, v, v, I, rys_table_data + rys_table_offset[EriBody<LP, LQ, LP, LQ, double>::NROOT]);
    double m = 0.0;
    if (sph && (LP >= 2 || LQ >= 2)) {
        // spherical basis: the largest diagonal element of the solid-harmonic block
        for (int al = 0; al < nsph(LP); ++al)
            for (int be = 0; be < nsph(LQ); ++be) m = fmax(m, fabs(sph_abab<LP, LQ>(I, al, be)));
    } else {
#pragma unroll
        for (int p = 0; p < NP; ++p)
#pragma unroll
            for (int q = 0; q < NQ; ++q) m = fmax(m, fabs(I[((p * NQ + q) * NP + p) * NQ + q]));
    }
    maxI[s] = sqrt(m);
}


// ----------------------------------------------------------------------------------
// Heavy classes: one 64-lane wavefront (= one work-group) per quartet.
// Per primitive quartet, lanes (axis, root) build the 2-D Rys tables into LDS and every
// lane accumulates its slice of the (e0|f0) entries in registers; the ket-side and then
// the bra-side transfer relation run level by level in LDS with lanes over entries;
// the digestion gives each J/K target element to one lane (one atomic per element).
// ----------------------------------------------------------------------------------
// slots of a workspace class: one per resident work-group of its persistent grid
__host__ __device__ constexpr long long ws_slot_count(long long ws_doubles, long long per_slot)
{
    return per_slot <= 0 ? 0 : (ws_doubles / per_slot < 8192 ? ws_doubles / per_slot : 8192);
}

// LDS budget of the wavefront form.  A class whose transfer buffers would take more keeps
// them in a global workspace slot instead (one slot per resident work-group, persistent
// grid), so that the 160 KiB of a CU still holds several work-groups.
#define TGPU_WAVE_LDS_BUDGET (64 * 1024)

// (e0|f0) entries a lane of the wavefront form accumulates in registers over the
// primitive quartets; with more, the accumulation goes through the transfer buffer
#ifndef TGPU_WAVE_GREG_MAX
#define TGPU_WAVE_GREG_MAX 4
#endif
// LDS of one wavefront's teams together in the wavefront form (see WaveCfg::TEAMS)
#ifndef TGPU_WAVE_TEAM_LDS
#define TGPU_WAVE_TEAM_LDS (16 * 1024)
#endif

// Class pairs whose wavefront form runs faster with up to 8 (e0|f0) entries per lane in
// registers than with 4 (FP64 build at (H2O)64/cc-pVTZ: (dp|dp) 264 -> 250 ms, (fs|dp)
// 208 -> 193, (dd|ds) 199 -> 184, (fp|pp) 135 -> 124, (fp|fs) 68 -> 63, (fd|dp) 175 ->
// 170; five other class pairs were 5-24 ms slower with 8)
__host__ __device__ constexpr bool wave_greg8(int lp, int lq, int lr, int ls)
{
    return (lp == 2 && lq == 1 && lr == 2 && ls == 1) || (lp == 3 && lq == 0 && lr == 2 && ls == 1) ||
           (lp == 2 && lq == 2 && lr == 2 && ls == 0) || (lp == 3 && lq == 1 && lr == 1 && ls == 1) ||
           (lp == 3 && lq == 1 && lr == 3 && ls == 0) || (lp == 3 && lq == 2 && lr == 2 && ls == 1);
}

template <int LP, int LQ, int LR, int LS>
struct WaveCfg {
    using T = WaveTab<LP, LQ, LR, LS>;
    static constexpr int NG = T::NE * T::NF;
    static constexpr int NGT = 3 * T::NROOT * (T::LPQ + 1) * (T::LRS + 1);
    static constexpr int NDEN = T::NP * T::NQ + T::NR * T::NS + T::NS * T::NQ + T::NR * T::NQ +
                                T::NS * T::NP + T::NR * T::NP;
    // transfer buffers in the global workspace
    static constexpr bool GB = (size_t)(NGT + 2 * T::BUF + NDEN) * 8 > TGPU_WAVE_LDS_BUDGET;
    static constexpr int LBUF = GB ? 1 : T::BUF;
    // doubles of workspace per slot (0: LDS only)
    static constexpr size_t WS = GB ? 2 * (size_t)T::BUF : 0;
    // LDS bytes of one quartet
    static constexpr size_t TEAM_BYTES = (size_t)(NGT + 2 * LBUF + NDEN) * 8;
    // Quartets one wavefront works on at once: its 64 lanes form TEAMS teams of S lanes,
    // each team on its own quartet (same bra pair, different kets) with its own LDS, all
    // in lockstep.  4 teams when their LDS fits TGPU_WAVE_TEAM_LDS and a team's lanes
    // cover the 3n (axis, root) tasks of the table phase, else 2, else 1.
    static constexpr int TEAMS = GB ? 1
                                 : (4 * TEAM_BYTES <= TGPU_WAVE_TEAM_LDS && 3 * T::NROOT <= 16) ? 4
                                 : (2 * TEAM_BYTES <= TGPU_WAVE_TEAM_LDS && 3 * T::NROOT <= 32) ? 2
                                                                                                : 1;
    static constexpr int S = 64 / TEAMS;
    // (e0|f0) entries per lane; in registers up to TGPU_WAVE_GREG_MAX (8 for the class
    // pairs of wave_greg8), else in the buffer
    static constexpr int MG = (NG + S - 1) / S;
    static constexpr bool GREG = MG <= (wave_greg8(LP, LQ, LR, LS) && TGPU_WAVE_GREG_MAX < 8 ? 8 : TGPU_WAVE_GREG_MAX);
};

// One team's 2-D tables and transfer buffers, in the arithmetic type R of its quartet
// (double, or float for the small-bound quartets of a mixed build)
template <int LP, int LQ, int LR, int LS, typename R>
struct WaveBuf {
    using C = WaveCfg<LP, LQ, LR, LS>;
    R g[C::NGT];
    R buf[2][C::LBUF];
};

// One team's LDS: its buffers (either precision in the same place) and the density blocks
// of its quartet
template <int LP, int LQ, int LR, int LS>
struct WaveLDS {
    using T = WaveTab<LP, LQ, LR, LS>;
    union {
        WaveBuf<LP, LQ, LR, LS, double> d;
        WaveBuf<LP, LQ, LR, LS, float> f;
    } u;
    double Ppq[T::NP * T::NQ], Prs[T::NR * T::NS];
    double Psq[T::NS * T::NQ], Prq[T::NR * T::NQ], Psp[T::NS * T::NP], Prp[T::NR * T::NP];
};

template <int LP, int LQ, int LR, int LS, typename R>
__device__ __forceinline__ WaveBuf<LP, LQ, LR, LS, R> &wave_buf(WaveLDS<LP, LQ, LR, LS> &L)
{
    if constexpr (sizeof(R) == sizeof(double)) return L.u.d;
    else return L.u.f;
}

// Computes (pq|rs) of one contracted quartet per team; returns the team's buffer holding
// [pq][rs] (LDS, or the work-group's workspace slot gbuf of 2*BUF values for the classes
// that keep their transfer buffers in global memory).  R is the type of the 2-D tables,
// the (e0|f0) accumulation and the transfer relations; roots, weights and the recurrence
// coefficients are computed in FP64.  All 64 lanes of the work-group call it together (it
// synchronises); tl is the lane's index in its team of S lanes, B the team's buffers.  A
// team with live == false (no quartet for it this round) follows along without
// contributing; npmax, the largest primitive-quartet count over the live teams, is
// uniform so that the loops and barriers are.
template <int LP, int LQ, int LR, int LS, int S, typename R>
__device__ const R *wave_eri(const DevPrims &pm, const PairView &bra, const PairView &ket,
                             WaveBuf<LP, LQ, LR, LS, R> &B, R *gbuf, int tl, bool live, int npmax)
{
    using T = WaveTab<LP, LQ, LR, LS>;
    using C = WaveCfg<LP, LQ, LR, LS>;
    constexpr int N = T::NROOT, W1 = T::LPQ + 1, W2 = T::LRS + 1, TW = W1 * W2;
    constexpr int NG = C::NG, MG = C::GREG ? (NG + S - 1) / S : 1;
    constexpr int NRS = T::NR * T::NS;
    R *src, *dst;
    if constexpr (C::GB) {
        src = gbuf;
        dst = gbuf + T::BUF;
    } else {
        src = B.buf[0];
        dst = B.buf[1];
        (void)gbuf;
    }
    R Gr[MG];
    // (e0|f0) entry k reads the x, y, z tables at gidx[3k..3k+2] (< TW <= 64): packed 6 bits each
    int gi[MG];
    if constexpr (C::GREG) {
#pragma unroll
        for (int m = 0; m < MG; ++m) {
            Gr[m] = (R)0;
            const int k = tl + S * m;
            const int kk = k < NG ? k : 0;
            gi[m] = T::gidx[3 * kk] | (T::gidx[3 * kk + 1] << 6) | (T::gidx[3 * kk + 2] << 12);
        }
    } else {
        // each lane owns entries tl, tl+S, ... until the transfer step
        for (int k = tl; k < NG; k += S) src[k] = (R)0;
    }
    const int ax = tl / N, rr = tl - (tl / N) * N;
    const int np = live ? bra.n * ket.n : 0;
    for (int ipq = 0; ipq < npmax; ++ipq) {
        const bool on = ipq < np;
        if (on && tl < 3 * N) {
            const int ip = ipq / ket.n, iq = ipq - ip * ket.n;
            const PrimRec pb = ldc(pm.rec + bra.off + ip);
            const PrimRec pk = ldc(pm.rec + ket.off + iq);
            const double za = pb.zeta, Ka = pb.K, Px = pb.Px, Py = pb.Py, Pz = pb.Pz;
            const double zc = pk.zeta, Kc = pk.K, Qx = pk.Px, Qy = pk.Py, Qz = pk.Pz;
            const double zs = za + zc, izs = 1.0 / zs, rho = za * zc * izs;
            const double QP[3] = {Qx - Px, Qy - Py, Qz - Pz};
            const double X = rho * (QP[0] * QP[0] + QP[1] * QP[1] + QP[2] * QP[2]);
            double ur, wr;
            rys_root_one<N>(X, rr, ur, wr);
            const double PA = (ax == 0 ? Px - bra.Ax : ax == 1 ? Py - bra.Ay : Pz - bra.Az);
            const double QC = (ax == 0 ? Qx - ket.Ax : ax == 1 ? Qy - ket.Ay : Qz - ket.Az);
            const double qp = QP[ax];
            const double rz = ur * rho / za, re = ur * rho / zc;
            const R b10 = (R)((1.0 - rz) * 0.5 / za), b01 = (R)((1.0 - re) * 0.5 / zc), b00 = (R)(ur * 0.5 * izs);
            const R c00 = (R)(PA + rz * qp), d00 = (R)(QC - re * qp);
            // the 2-D table I(e, f) column by column (two columns live), each column
            // stored to LDS as soon as it is complete
            R *gt = B.g + (ax * N + rr) * TW;
            R cp[W1], cc[W1];
            cc[0] = ax == 2 ? (R)(wr * TWOPI5ON2 * Ka * Kc * sqrt(rho)) : (R)1;
#pragma unroll
            for (int e = 0; e + 1 < W1; ++e) cc[e + 1] = c00 * cc[e] + (e ? (R)e * b10 * cc[e - 1] : (R)0);
#pragma unroll
            for (int e = 0; e < W1; ++e) gt[e * W2] = cc[e];
#pragma unroll
            for (int f = 0; f + 1 < W2; ++f) {
                R cn[W1];
#pragma unroll
                for (int e = 0; e < W1; ++e)
                    cn[e] = d00 * cc[e] + (f ? (R)f * b01 * cp[e] : (R)0) + (e ? (R)e * b00 * cc[e - 1] : (R)0);
#pragma unroll
                for (int e = 0; e < W1; ++e) {
                    gt[e * W2 + f + 1] = cn[e];
                    cp[e] = cc[e];
                    cc[e] = cn[e];
                }
            }
        }
        __syncthreads();
        if (on) {
            if constexpr (C::GREG) {
#pragma unroll
                for (int m = 0; m < MG; ++m) {
                    const int i0 = gi[m] & 63, i1 = TW * N + ((gi[m] >> 6) & 63), i2 = 2 * TW * N + (gi[m] >> 12);
                    R acc = (R)0;
#pragma unroll
                    for (int r = 0; r < N; ++r) acc += B.g[i0 + r * TW] * B.g[i1 + r * TW] * B.g[i2 + r * TW];
                    Gr[m] += acc;
                }
            } else {
                for (int k = tl; k < NG; k += S) {
                    const int i0 = T::gidx[3 * k], i1 = TW * N + T::gidx[3 * k + 1],
                              i2 = 2 * TW * N + T::gidx[3 * k + 2];
                    R acc = (R)0;
#pragma unroll
                    for (int r = 0; r < N; ++r) acc += B.g[i0 + r * TW] * B.g[i1 + r * TW] * B.g[i2 + r * TW];
                    src[k] += acc;
                }
            }
        }
        __syncthreads();
    }
    if constexpr (C::GREG) {
#pragma unroll
        for (int m = 0; m < MG; ++m) {
            const int k = tl + S * m;
            if (k < NG) src[k] = Gr[m];
        }
    }
    __syncthreads();
    // ket-side transfer, rows e (T::NE of them), entries per level
    const R CD[3] = {(R)ket.ABx, (R)ket.ABy, (R)ket.ABz};
    int nprev = T::NF;
    for (int lev = 0; lev < T::NCD; ++lev) {
        const int n = T::cd_n[lev], off = T::cd_off[lev];
        for (int idx = tl; idx < T::NE * n; idx += S) {
            const int e = idx / n, t = idx - e * n;
            dst[e * n + t] = src[e * nprev + T::cd_s1[off + t]] + CD[T::cd_ax[off + t]] * src[e * nprev + T::cd_s2[off + t]];
        }
        __syncthreads();
        R *tmp = src; src = dst; dst = tmp;
        nprev = n;
    }
    // bra-side transfer on columns rs
    const R AB[3] = {(R)bra.ABx, (R)bra.ABy, (R)bra.ABz};
    for (int lev = 0; lev < T::NAB; ++lev) {
        const int n = T::ab_n[lev], off = T::ab_off[lev];
        for (int idx = tl; idx < n * NRS; idx += S) {
            const int t = idx / NRS, c = idx - t * NRS;
            dst[t * NRS + c] = src[T::ab_s1[off + t] * NRS + c] + AB[T::ab_ax[off + t]] * src[T::ab_s2[off + t] * NRS + c];
        }
        __syncthreads();
        R *tmp = src; src = dst; dst = tmp;
    }
    return src;
}

// index of the (k+1)-th set bit of m (64 if there is none)
__device__ __forceinline__ int nth_set_bit(unsigned long long m, int k)
{
    for (int i = 0; i < k && m; ++i) m &= m - 1;
    return m ? __builtin_ctzll(m) : 64;
}

// J and K digestion of one quartet's integrals I ([pq][rs], type R) by the S lanes of its
// team: one atomic per target element; the bra-pair J block accumulates in Jpq
template <int LP, int LQ, int LR, int LS, int S, typename R>
__device__ __forceinline__ void wave_digest(const JKArgs &a, const BraInfo &b, const KetTest &t, int fr, int fs,
                                            const R *__restrict__ I, const WaveLDS<LP, LQ, LR, LS> &L, int tl,
                                            double *Jpq)
{
    using T = WaveTab<LP, LQ, LR, LS>;
    constexpr int NP = T::NP, NQ = T::NQ, NR = T::NR, NS = T::NS;
    constexpr int NPQ = NP * NQ, NRS = NR * NS, MJ = (NPQ + S - 1) / S;
    const int ld = b.ld, fp = b.fp, fq = b.fq;
    const double f = t.f;
    // J sides as in the lane form (both, or the UHF pass's Tonto ket side only)
    bool j_ket = true, j_bra = true;
    if (a.j_ket_nonsame) {
        const int tk = ldc(a.pr.rec0 + t.sk).x;
        j_ket = b.tb > tk;
        j_bra = b.tb < tk;
    }
    if (!t.skipJ) {
        if (j_ket) {
            for (int u = tl; u < NRS; u += S) {
                double acc = 0.0;
#pragma unroll 4
                for (int pq = 0; pq < NPQ; ++pq) acc += (double)I[pq * NRS + u] * L.Ppq[pq];
                unsafeAtomicAdd(&a.XP[a.kblk_off + (long long)t.sk * NRS + u], f * acc);
            }
        }
        if (j_bra) {
#pragma unroll
            for (int m = 0; m < MJ; ++m) {
                const int pq = tl + S * m;
                if (pq < NPQ) {
                    double acc = 0.0;
#pragma unroll 4
                    for (int rs = 0; rs < NRS; ++rs) acc += (double)I[pq * NRS + rs] * L.Prs[rs];
                    Jpq[m] += f * acc;
                }
            }
        }
    }
    if (!t.skipK) {
        constexpr int N1 = NP * NR, N2 = N1 + NP * NS, N3 = N2 + NQ * NR, N4 = N3 + NQ * NS;
        for (int u = tl; u < N4; u += S) {
            double acc = 0.0;
            size_t dst;
            if (u < N1) {  // K(p,r) += I P(s,q)
                const int p = u / NR, r = u % NR;
#pragma unroll 1
                for (int q = 0; q < NQ; ++q)
#pragma unroll 4
                    for (int s = 0; s < NS; ++s) acc += (double)I[(p * NQ + q) * NRS + r * NS + s] * L.Psq[s * NQ + q];
                dst = (size_t)(fp + p) * ld + fr + r;
            } else if (u < N2) {  // K(p,s) += I P(r,q)
                const int p = (u - N1) / NS, s = (u - N1) % NS;
#pragma unroll 1
                for (int q = 0; q < NQ; ++q)
#pragma unroll 4
                    for (int r = 0; r < NR; ++r) acc += (double)I[(p * NQ + q) * NRS + r * NS + s] * L.Prq[r * NQ + q];
                dst = (size_t)(fp + p) * ld + fs + s;
            } else if (u < N3) {  // K(q,r) += I P(s,p)
                const int q = (u - N2) / NR, r = (u - N2) % NR;
#pragma unroll 1
                for (int p = 0; p < NP; ++p)
#pragma unroll 4
                    for (int s = 0; s < NS; ++s) acc += (double)I[(p * NQ + q) * NRS + r * NS + s] * L.Psp[s * NP + p];
                dst = (size_t)(fq + q) * ld + fr + r;
            } else {  // K(q,s) += I P(r,p)
                const int q = (u - N3) / NS, s = (u - N3) % NS;
#pragma unroll 1
                for (int p = 0; p < NP; ++p)
#pragma unroll 4
                    for (int r = 0; r < NR; ++r) acc += (double)I[(p * NQ + q) * NRS + r * NS + s] * L.Prp[r * NP + p];
                dst = (size_t)(fq + q) * ld + fs + s;
            }
            unsafeAtomicAdd(&b.Y[dst], f * acc);
        }
    }
}

// The wavefront form's work for bra pair bi: the chunks blockIdx.y, blockIdx.y + gridDim.y,
// ... of a.ket_chunk kets of its Schwarz prefix (only chunks that hold kets are visited).
// The 64 lanes screen up to 64 kets of a chunk at once; the kets that pass are then taken
// TEAMS at a time, one per team (their test results fetched from the lanes that computed
// them).  The prefix is split where max_I(ab) max_I(cd) falls below a.fp32_thr (0 in an
// FP64 build: no split): the FP64 instantiation (R = double) takes the kets before the
// split, the FP32 one (R = float: tables, accumulation and transfers in FP32) those after
// it, in a launch of its own, so that neither carries the other's registers.
template <int LP, int LQ, int LR, int LS, typename R>
__device__ void wave_bra(const JKArgs &a, int bi, WaveLDS<LP, LQ, LR, LS> *Lt, double *gbuf)
{
    using T = WaveTab<LP, LQ, LR, LS>;
    using C = WaveCfg<LP, LQ, LR, LS>;
    constexpr bool F32 = sizeof(R) == sizeof(float);
    constexpr int NP = T::NP, NQ = T::NQ, NR = T::NR, NS = T::NS, S = C::S, NT = C::TEAMS;
    constexpr int NPQ = NP * NQ, NRS = NR * NS, MJ = (NPQ + S - 1) / S;
    const int lane = threadIdx.x, team = lane / S, tl = lane - team * S;
    WaveLDS<LP, LQ, LR, LS> &L = Lt[team];
    BraInfo b;
    load_bra(a, bi, b);
    int lo = 0, hi = b.jend;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (b.mIb * ldc(b.maxI + mid) < a.fp32_thr) hi = mid;
        else lo = mid + 1;
    }
    const int jbeg = F32 ? lo : 0, jend = F32 ? b.jend : lo;
    const int step = gridDim.y * a.ket_chunk;
    if (jbeg + (int)blockIdx.y * a.ket_chunk >= jend) return;  // uniform over the work-group
    const int ld = b.ld, fp = b.fp, fq = b.fq;
    const double *__restrict__ P = b.P;
    for (int t = tl; t < NPQ; t += S) L.Ppq[t] = P[(size_t)(fp + t / NQ) * ld + fq + t % NQ];
    double Jpq[MJ];
#pragma unroll
    for (int m = 0; m < MJ; ++m) Jpq[m] = 0.0;
    unsigned long long nq = 0, npq = 0;
    for (int jc = jbeg + blockIdx.y * a.ket_chunk; jc < jend; jc += step) {
        const int jlim = min(jend, jc + a.ket_chunk);
        KetTest tt;
        const bool ok = jc + lane < jlim && test_ket(a, b, jc + lane, tt);
        const int flags_l = ok ? ((int)tt.skipJ | ((int)tt.skipK << 1) | ((int)tt.skip_lists << 2) |
                                  (tt.halvings << 3))
                               : 0;
        const int frs_l = ok ? pack16(tt.fr, tt.fs) : 0;
        unsigned long long pass = __ballot(ok);
    while (pass) {  // uniform over the work-group
        // this round: the team-th remaining ket for each team
        const int l = nth_set_bit(pass, team);
        const bool live = l < 64;
#pragma unroll
        for (int k = 0; k < NT; ++k) pass &= pass - 1;
        const int src_lane = live ? l : 0;
        const int flags = __shfl(flags_l, src_lane, 64), frs = __shfl(frs_l, src_lane, 64);
        const int j = jc + src_lane;
        KetTest t;
        t.sk = b.kb + j;
        t.skipJ = flags & 1;
        t.skipK = flags & 2;
        t.skip_lists = flags & 4;
        t.f = (flags >> 3) == 0 ? 1.0 : (flags >> 3) == 1 ? 0.5 : (flags >> 3) == 2 ? 0.25 : 0.125;
        const PairView bv = make_view(a.pr, a.pm, b.sb, t.skip_lists);
        const PairView kv = make_view(a.pr, a.pm, t.sk, t.skip_lists);
        const int fr = hi16(frs), fs = lo16(frs);
        // stage the ket-dependent density blocks (P symmetric: the K-side blocks come from
        // the bra pair's rows); when each fits one pass of the team's lanes the five loads
        // are issued together and waited for once
        if (live) {
            if constexpr (NRS <= S && NS * NQ <= S && NR * NQ <= S && NS * NP <= S && NR * NP <= S) {
                const int u = tl;
                const double v0 = u < NRS ? a.PP[a.kblk_off + (long long)t.sk * NRS + u] : 0.0;
                const double v1 = u < NS * NQ ? P[(size_t)(fq + u % NQ) * ld + fs + u / NQ] : 0.0;
                const double v2 = u < NR * NQ ? P[(size_t)(fq + u % NQ) * ld + fr + u / NQ] : 0.0;
                const double v3 = u < NS * NP ? P[(size_t)(fp + u % NP) * ld + fs + u / NP] : 0.0;
                const double v4 = u < NR * NP ? P[(size_t)(fp + u % NP) * ld + fr + u / NP] : 0.0;
                if (u < NRS) L.Prs[u] = v0;
                if (u < NS * NQ) L.Psq[u] = v1;
                if (u < NR * NQ) L.Prq[u] = v2;
                if (u < NS * NP) L.Psp[u] = v3;
                if (u < NR * NP) L.Prp[u] = v4;
            } else {
                for (int u = tl; u < NRS; u += S) L.Prs[u] = a.PP[a.kblk_off + (long long)t.sk * NRS + u];
                for (int u = tl; u < NS * NQ; u += S) L.Psq[u] = P[(size_t)(fq + u % NQ) * ld + fs + u / NQ];
                for (int u = tl; u < NR * NQ; u += S) L.Prq[u] = P[(size_t)(fq + u % NQ) * ld + fr + u / NQ];
                for (int u = tl; u < NS * NP; u += S) L.Psp[u] = P[(size_t)(fp + u % NP) * ld + fs + u / NP];
                for (int u = tl; u < NR * NP; u += S) L.Prp[u] = P[(size_t)(fp + u % NP) * ld + fr + u / NP];
            }
        }
        // primitive quartets of the longest quartet this round (uniform loop bound)
        const int np = live ? bv.n * kv.n : 0;
        int npmax = np;
#pragma unroll
        for (int off = 32; off >= S; off >>= 1) npmax = max(npmax, __shfl_xor(npmax, off, 64));
        const R *I = wave_eri<LP, LQ, LR, LS, S, R>(a.pm, bv, kv, wave_buf<LP, LQ, LR, LS, R>(L),
                                                    reinterpret_cast<R *>(gbuf), tl, live, npmax);
        if (live) {
            wave_digest<LP, LQ, LR, LS, S, R>(a, b, t, fr, fs, I, L, tl, Jpq);
            if (tl == 0) {
                ++nq;
                npq += (unsigned long long)np;
            }
        }
        __syncthreads();
    }
    }
    // the teams' partial sums of the bra-pair J block: lanes with the same tl add up
#pragma unroll
    for (int m = 0; m < MJ; ++m) {
#pragma unroll
        for (int off = 32; off >= S; off >>= 1) Jpq[m] += __shfl_xor(Jpq[m], off, 64);
        const int pq = tl + S * m;
        if (team == 0 && pq < NPQ && Jpq[m] != 0.0)
            unsafeAtomicAdd(&b.X[(size_t)(fp + pq / NQ) * ld + fq + pq % NQ], Jpq[m]);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        nq += __shfl_xor(nq, off, 64);
        npq += __shfl_xor(npq, off, 64);
    }
    if (lane == 0 && nq) {
        atomicAdd(&a.counters[3 * a.counter_slot + 0], nq);
        atomicAdd(&a.counters[3 * a.counter_slot + 1], npq);
        if (F32) atomicAdd(&a.counters[3 * a.counter_slot + 2], nq);
    }
    __syncthreads();  // L is reused by the work-group's next bra pair
}

// Wavefront form: grid (gx, gy); work-group (x, y) takes bra pairs x, x + gx, ... and, of
// each, the ket chunks y, y + gy, ...  Workspace classes keep gx * gy <= their slot count and use
// slot y * gx + x.
#ifndef TGPU_WAVE_MIN_WAVES
#define TGPU_WAVE_MIN_WAVES 1
#endif
template <int LP, int LQ, int LR, int LS, typename R>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(TGPU_WAVE_MIN_WAVES, 8))) void jk_wave_kernel(JKArgs a)
{
    using C = WaveCfg<LP, LQ, LR, LS>;
    __shared__ WaveLDS<LP, LQ, LR, LS> L[C::TEAMS];
    double *gbuf = nullptr;
    if constexpr (C::GB) gbuf = a.ws + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * C::WS;
    for (int bi = blockIdx.x; bi < a.n_bra; bi += gridDim.x) wave_bra<LP, LQ, LR, LS, R>(a, bi, L, gbuf);
}

template <int LP, int LQ>
__global__ __launch_bounds__(64) void schwarz_wave_kernel(DevPairs pr, DevPrims pm, int start, int n,
                                                         double *maxI, double *ws, int sph)
{
    using T = WaveTab<LP, LQ, LP, LQ>;
    using C = WaveCfg<LP, LQ, LP, LQ>;
    constexpr int NPQ = T::NP * T::NQ;
    __shared__ WaveLDS<LP, LQ, LP, LQ> L;
    double *gbuf = C::GB ? ws + (size_t)blockIdx.x * C::WS : nullptr;
    // LDS classes launch one work-group per pair; workspace classes a persistent grid
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int s = start + i;
        const PairView v = make_view(pr, pm, s, false);
        const double *I = wave_eri<LP, LQ, LP, LQ, 64, double>(pm, v, v, L.u.d, gbuf, threadIdx.x, true, v.n * v.n);
        double m = 0.0;
        if (sph && (LP >= 2 || LQ >= 2)) {
            for (int t = threadIdx.x; t < nsph(LP) * nsph(LQ); t += 64)
                m = fmax(m, fabs(sph_abab<LP, LQ>(I, t / nsph(LQ), t % nsph(LQ))));
        } else {
            for (int t = threadIdx.x; t < NPQ; t += 64) m = fmax(m, fabs(I[t * NPQ + t]));
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
        if (threadIdx.x == 0) maxI[s] = sqrt(m);
        __syncthreads();
    }
}

// Launchers defined in the per-class translation units (jk_class.hip).  variant 0: one
// lane per quartet (light classes only); variant 1: one wavefront per quartet.
#define TGPU_DECLARE_JK(LP, LQ, LR, LS, H) \
    int tgpu_launch_jk_##LP##LQ##LR##LS(const JKArgs &a, int variant, bool mixed, hipStream_t s); \
    size_t tgpu_wave_ws_##LP##LQ##LR##LS();
TGPU_FOR_EACH_CLASS(TGPU_DECLARE_JK)
#undef TGPU_DECLARE_JK
#define TGPU_DECLARE_SCHWARZ(LP, LQ, H) \
    void tgpu_launch_schwarz_##LP##LQ(const DevPairs &pr, const DevPrims &pm, int start, int n, double *maxI, \
                                      double *ws, long long ws_doubles, int sph, hipStream_t s);
TGPU_FOR_EACH_PAIRCLASS(TGPU_DECLARE_SCHWARZ)
#undef TGPU_DECLARE_SCHWARZ

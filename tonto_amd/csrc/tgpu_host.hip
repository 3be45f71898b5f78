// libtonto_gpu: MI355X (gfx950) direct-SCF J/K build behind the C ABI of
// include/tonto_gpu.h.  Drop-in for Tonto's MOLECULE.FOCK:make_r_JK_engine
// (foofiles/molecule.fock.foo:1365-1505).
//
// Data layout in HBM (built once per geometry by tgpu_create / tgpu_schwarz):
//   * shell pairs in Tonto's order are binned by angular-momentum class (lp >= lq,
//     class = lp(lp+1)/2 + lq) and, after the Schwarz table is known, sorted inside
//     each bin by max_I descending; per-pair fields are structure-of-arrays;
//   * primitive pairs (zeta, prefactor, product centre) are SoA, each shell pair owning
//     two contiguous runs: the full list and the list with Tonto's 1e-6 primitive-pair
//     skip applied (shell1quartet.foo:229-273, 374-416).
// One J/K build = P_max per shell pair -> global maxima -> one launch per class pair
// (bra class >= ket class) -> post-processing J = 2(X + X^T), K = Y + Y^T.
// In the class-pair launch one 64-lane wavefront owns one bra pair and its lanes walk
// the Schwarz-sorted ket pairs of the other bin up to the exact Schwarz frontier;
// every lane applies Tonto's screening tests bit for bit (molecule.fock.foo:1412-1472,
// Schwarz_test 3497-3536), evaluates the contracted quartet (Rys + HRR) and scatters
// into J and K.  J(bra) partial sums stay in registers and are wavefront-reduced once
// per bra pair; the remaining J and K updates are FP64 hardware atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "tgpu_device.h"

// P_max(ab) = max |P(a-block, b-block)| per Tonto shell pair (molecule.fock.foo:3601-3633)
__global__ void pmax_kernel(const double *__restrict__ P, int nbf, int nshell,
                            const int *__restrict__ sh_l, const int *__restrict__ sh_bf,
                            double *__restrict__ Pmax)
{
    const long long npair = (long long)nshell * (nshell + 1) / 2;
    for (long long ab = blockIdx.x * (long long)blockDim.x + threadIdx.x; ab < npair;
         ab += (long long)gridDim.x * blockDim.x) {
        int x = (int)((sqrt(8.0 * (double)ab + 1.0) - 1.0) * 0.5);
        while ((long long)x * (x + 1) / 2 > ab) --x;
        while ((long long)(x + 1) * (x + 2) / 2 <= ab) ++x;
        const int y = (int)(ab - (long long)x * (x + 1) / 2);
        const int fa = sh_bf[x], na = ncart(sh_l[x]), fb = sh_bf[y], nb = ncart(sh_l[y]);
        double m = 0.0;
        for (int i = 0; i < na; ++i)
            for (int j = 0; j < nb; ++j) m = fmax(m, fabs(P[(size_t)(fa + i) * nbf + fb + j]));
        Pmax[ab] = m;
    }
}

// scal[0] = P_max = max over pairs; scal[1] = IP_max = max(max_I) * P_max (molecule.fock.foo:1401-1402)
__global__ __launch_bounds__(1024) void global_max_kernel(const double *__restrict__ v, long long n,
                                                          double maxI_all, double *scal)
{
    __shared__ double red[1024];
    double m = 0.0;
    for (long long i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, v[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        scal[0] = red[0];
        scal[1] = maxI_all * red[0];
    }
}

__global__ void fill_zero_kernel(double *__restrict__ a, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        a[i] = 0.0;
}

__global__ void fill_zero_u64_kernel(unsigned long long *a, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0ull;
}

// Tonto's post-processing: weight_diagonal_blocks(J,2), J = 2J, blockwise_symmetric_fold(K),
// symmetric_reflect (molecule.fock.foo:1485-1500).  With the contributions kept in X and Y
// (each unique quartet scattered once, with its permutation factor) these are
// J = 2 (X + X^T) and K = Y + Y^T.
__global__ void post_kernel(const double *__restrict__ X, const double *__restrict__ Y, int n,
                            double *__restrict__ J, double *__restrict__ K)
{
    const size_t nn = (size_t)n * n;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < nn;
         idx += (size_t)gridDim.x * blockDim.x) {
        const size_t i = idx / n, j = idx % n, t = j * n + i;
        J[idx] = 2.0 * (X[idx] + X[t]);
        K[idx] = Y[idx] + Y[t];
    }
}

// One-electron S, T, V for one Tonto shell pair per thread (molecule.ints.foo:34-415):
// Obara-Saika 1-D overlap/kinetic, Rys quadrature with X = p|PC|^2 for the nuclear term.
__device__ void os1d(double PA, double PB, double hp, int la, int lb, double S[6][6])
{
    S[0][0] = 1.0;
    for (int i = 0; i <= la; ++i)
        for (int j = 0; j <= lb; ++j) {
            if (i == 0 && j == 0) continue;
            if (i > 0)
                S[i][j] = PA * S[i - 1][j] +
                          hp * ((i > 1 ? (i - 1) * S[i - 2][j] : 0.0) + (j > 0 ? j * S[i - 1][j - 1] : 0.0));
            else
                S[i][j] = PB * S[i][j - 1] + hp * (j > 1 ? (j - 1) * S[i][j - 2] : 0.0);
        }
}

struct OneEArgs {
    int nshell, natom, nbf;
    const int *sh_l, *sh_bf, *sh_atom, *sh_off, *sh_np, *cx, *cy, *cz;  // cx..: component tables [l][i]
    const double *exps, *coefs, *xyz, *Z;
    double *S, *T, *V;
};

__global__ void one_e_kernel(OneEArgs a)
{
    const long long npair = (long long)a.nshell * (a.nshell + 1) / 2;
    const long long ab = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (ab >= npair) return;
    int x = (int)((sqrt(8.0 * (double)ab + 1.0) - 1.0) * 0.5);
    while ((long long)x * (x + 1) / 2 > ab) --x;
    while ((long long)(x + 1) * (x + 2) / 2 <= ab) ++x;
    const int y = (int)(ab - (long long)x * (x + 1) / 2);
    const int la = a.sh_l[x], lb = a.sh_l[y], na = ncart(la), nb = ncart(lb);
    const double *A = a.xyz + 3 * a.sh_atom[x], *B = a.xyz + 3 * a.sh_atom[y];
    double sb[10][10], tb[10][10], vb[10][10];
    for (int i = 0; i < 10; ++i)
        for (int j = 0; j < 10; ++j) sb[i][j] = tb[i][j] = vb[i][j] = 0.0;
    double AB2 = 0.0;
    for (int t = 0; t < 3; ++t) AB2 += (A[t] - B[t]) * (A[t] - B[t]);
    const int L = la + lb, nr = L / 2 + 1;
    for (int ia = 0; ia < a.sh_np[x]; ++ia)
        for (int ib = 0; ib < a.sh_np[y]; ++ib) {
            const double al = a.exps[a.sh_off[x] + ia], be = a.exps[a.sh_off[y] + ib];
            const double cc = a.coefs[a.sh_off[x] + ia] * a.coefs[a.sh_off[y] + ib];
            const double p = al + be, mu = al * be / p;
            double Pc[3], Sx[3][6][6];
            for (int t = 0; t < 3; ++t) {
                Pc[t] = (al * A[t] + be * B[t]) / p;
                os1d(Pc[t] - A[t], Pc[t] - B[t], 0.5 / p, la + 1, lb + 1, Sx[t]);
            }
            const double ek = exp(-mu * AB2);
            const double pre = cc * ek * pow(M_PI / p, 1.5);
            for (int i = 0; i < na; ++i)
                for (int j = 0; j < nb; ++j) {
                    const int xa[3] = {a.cx[la * 10 + i], a.cy[la * 10 + i], a.cz[la * 10 + i]};
                    const int xb[3] = {a.cx[lb * 10 + j], a.cy[lb * 10 + j], a.cz[lb * 10 + j]};
                    double s3[3], t3[3];
                    for (int t = 0; t < 3; ++t) {
                        const int u = xa[t], v = xb[t];
                        s3[t] = Sx[t][u][v];
                        t3[t] = 2.0 * al * be * Sx[t][u + 1][v + 1] +
                                (u > 0 && v > 0 ? 0.5 * u * v * Sx[t][u - 1][v - 1] : 0.0) -
                                (v > 0 ? al * v * Sx[t][u + 1][v - 1] : 0.0) -
                                (u > 0 ? be * u * Sx[t][u - 1][v + 1] : 0.0);
                    }
                    sb[i][j] += pre * s3[0] * s3[1] * s3[2];
                    tb[i][j] += pre * (t3[0] * s3[1] * s3[2] + s3[0] * t3[1] * s3[2] + s3[0] * s3[1] * t3[2]);
                }
            for (int C = 0; C < a.natom; ++C) {
                const double *Cc = a.xyz + 3 * C;
                double PC2 = 0.0;
                for (int t = 0; t < 3; ++t) PC2 += (Pc[t] - Cc[t]) * (Pc[t] - Cc[t]);
                double u[7], w[7];
                rys_roots_rt(nr, p * PC2, u, w);
                const double vpre = -a.Z[C] * 2.0 * M_PI / p * ek * cc;
                for (int r = 0; r < nr; ++r) {
                    double I1[3][4][4];
                    for (int t = 0; t < 3; ++t) {
                        const double c00 = (Pc[t] - A[t]) + u[r] * (Cc[t] - Pc[t]);
                        const double b10 = (1.0 - u[r]) * 0.5 / p;
                        double H[7][4];
                        H[0][0] = 1.0;
                        for (int e = 0; e < L; ++e)
                            H[e + 1][0] = c00 * H[e][0] + (e ? e * b10 * H[e - 1][0] : 0.0);
                        for (int j = 1; j <= lb; ++j)
                            for (int i = 0; i <= L - j; ++i)
                                H[i][j] = H[i + 1][j - 1] + (A[t] - B[t]) * H[i][j - 1];
                        for (int i = 0; i <= la; ++i)
                            for (int j = 0; j <= lb; ++j) I1[t][i][j] = H[i][j];
                    }
                    for (int i = 0; i < na; ++i)
                        for (int j = 0; j < nb; ++j) {
                            vb[i][j] += vpre * w[r] *
                                        I1[0][a.cx[la * 10 + i]][a.cx[lb * 10 + j]] *
                                        I1[1][a.cy[la * 10 + i]][a.cy[lb * 10 + j]] *
                                        I1[2][a.cz[la * 10 + i]][a.cz[lb * 10 + j]];
                        }
                }
            }
        }
    const int fa = a.sh_bf[x], fb = a.sh_bf[y];
    for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j) {
            const size_t ij = (size_t)(fa + i) * a.nbf + fb + j, ji = (size_t)(fb + j) * a.nbf + fa + i;
            a.S[ij] = a.S[ji] = sb[i][j];
            a.T[ij] = a.T[ji] = tb[i][j];
            a.V[ij] = a.V[ji] = vb[i][j];
        }
}

// ----------------------------------------------------------------------------------
// Host side
// ----------------------------------------------------------------------------------
static constexpr int NCLS = TGPU_NPAIRCLASS;
static constexpr int NQCLS = NCLS * (NCLS + 1) / 2;

template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
};

struct tgpu_ctx {
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    int n_atom = 0, n_shell = 0, n_bf = 0;
    long long n_pairs = 0;
    std::vector<double> xyz, Z, exps, coefs;
    std::vector<int> sh_l, sh_atom, sh_np, sh_off, sh_bf;
    double cut_schwarz = 1e-9, cut_jp = 1e-9, cut_kp = 1e-9, cut_pair = 1e-6, fp32_thr = 1e-6;
    // host pair tables in device order (binned by class; sorted by max_I after Schwarz)
    int bin_start[NCLS + 1] = {0};
    std::vector<int> tonto, sp, sq, bfp, bfq, atp, atq, off_full, n_full, off_skip, n_skip;
    std::vector<double> A, AB, maxI;
    std::vector<double> pz, pK, pPx, pPy, pPz;
    bool have_schwarz = false, sorted = false;
    double maxI_all = 0.0;
    std::vector<double> maxI_tonto;
    // device
    DBuf<int> d_tonto, d_sp, d_sq, d_bfp, d_bfq, d_atp, d_atq, d_off_full, d_n_full, d_off_skip,
        d_n_skip, d_sh_l, d_sh_bf;
    DBuf<double> d_A, d_AB, d_maxI, d_pz, d_pK, d_pPx, d_pPy, d_pPz;
    DBuf<double> d_X, d_Y, d_Pmax, d_scal, d_hP, d_hJ, d_hK;
    DBuf<unsigned long long> d_counters;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    // optional per-class-pair timing (tgpu_set_profiling)
    bool profiling = false;
    std::vector<hipEvent_t> cls_ev;  // 2 per class pair
    std::vector<char> cls_launched;
};

#define HIPCHK(ctx, call)                                                              \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);            \
            return TGPU_E_HIP;                                                         \
        }                                                                              \
    } while (0)

template <typename T>
static int dalloc(tgpu_ctx *c, DBuf<T> &b, size_t n)
{
    if (b.p && b.n >= n) return 0;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
    if (n == 0) n = 1;
    HIPCHK(c, hipMalloc((void **)&b.p, n * sizeof(T)));
    b.n = n;
    return 0;
}

template <typename T>
static int dupload(tgpu_ctx *c, DBuf<T> &b, const std::vector<T> &h)
{
    if (int rc = dalloc(c, b, h.size())) return rc;
    if (!h.empty()) HIPCHK(c, hipMemcpy(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

template <typename T>
static void dfree(DBuf<T> &b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
}

// Primitive-pair data of one Tonto shell pair (a >= b): full list, and the list with
// pairs exp(-a b R^2/(a+b)) < pair_cutoff dropped.  Same arithmetic order as
// SHELL1PAIR:make_precalculated_data (shell1pair.foo:128-143) and SHELL1QUARTET:set_ab_new
// (shell1quartet.foo:201-270) so that the skip decisions are bit-identical.
static void build_prims(tgpu_ctx *c)
{
    c->pz.clear(); c->pK.clear(); c->pPx.clear(); c->pPy.clear(); c->pPz.clear();
    const size_t np = c->tonto.size();
    c->off_full.assign(np, 0); c->n_full.assign(np, 0);
    c->off_skip.assign(np, 0); c->n_skip.assign(np, 0);
    for (size_t s = 0; s < np; ++s) {
        const int ab = c->tonto[s];
        int a = (int)((std::sqrt(8.0 * ab + 1.0) - 1.0) / 2.0);
        while (a * (a + 1) / 2 > ab) --a;
        while ((a + 1) * (a + 2) / 2 <= ab) ++a;
        const int b = ab - a * (a + 1) / 2;
        const double *PA = &c->xyz[3 * c->sh_atom[a]], *PB = &c->xyz[3 * c->sh_atom[b]];
        const bool same = c->sh_atom[a] == c->sh_atom[b];
        const double ABv[3] = {PA[0] - PB[0], PA[1] - PB[1], PA[2] - PB[2]};
        const double r2 = ABv[0] * ABv[0] + ABv[1] * ABv[1] + ABv[2] * ABv[2];
        for (int pass = 0; pass < (same ? 1 : 2); ++pass) {
            const bool skip = pass == 1;
            const int off = (int)c->pz.size();
            for (int bg = 0; bg < c->sh_np[b]; ++bg) {
                const double be = c->exps[c->sh_off[b] + bg], bcc = c->coefs[c->sh_off[b] + bg];
                const double b_r2 = be * r2;
                for (int ag = 0; ag < c->sh_np[a]; ++ag) {
                    const double ae = c->exps[c->sh_off[a] + ag], acc = c->coefs[c->sh_off[a] + ag];
                    const double sum = ae + be, inv = 1.0 / sum;
                    const double cc = bcc * acc * inv * std::sqrt(inv);
                    double pref, P[3];
                    if (same) {
                        pref = cc;
                        P[0] = PA[0]; P[1] = PA[1]; P[2] = PA[2];
                    } else {
                        const double e = std::exp(-ae * b_r2 * inv);
                        if (skip && e < c->cut_pair) continue;
                        pref = cc * e;
                        for (int t = 0; t < 3; ++t) P[t] = (be * PB[t] + ae * PA[t]) * inv;
                    }
                    c->pz.push_back(sum);
                    c->pK.push_back(pref);
                    c->pPx.push_back(P[0]);
                    c->pPy.push_back(P[1]);
                    c->pPz.push_back(P[2]);
                }
            }
            const int n = (int)c->pz.size() - off;
            if (!skip) { c->off_full[s] = off; c->n_full[s] = n; }
            else { c->off_skip[s] = off; c->n_skip[s] = n; }
        }
        if (same) { c->off_skip[s] = c->off_full[s]; c->n_skip[s] = c->n_full[s]; }
    }
}

static int upload_pairs(tgpu_ctx *c)
{
    int rc = 0;
    if ((rc = dupload(c, c->d_tonto, c->tonto)) || (rc = dupload(c, c->d_sp, c->sp)) ||
        (rc = dupload(c, c->d_sq, c->sq)) || (rc = dupload(c, c->d_bfp, c->bfp)) ||
        (rc = dupload(c, c->d_bfq, c->bfq)) || (rc = dupload(c, c->d_atp, c->atp)) ||
        (rc = dupload(c, c->d_atq, c->atq)) || (rc = dupload(c, c->d_off_full, c->off_full)) ||
        (rc = dupload(c, c->d_n_full, c->n_full)) || (rc = dupload(c, c->d_off_skip, c->off_skip)) ||
        (rc = dupload(c, c->d_n_skip, c->n_skip)) || (rc = dupload(c, c->d_A, c->A)) ||
        (rc = dupload(c, c->d_AB, c->AB)) || (rc = dupload(c, c->d_pz, c->pz)) ||
        (rc = dupload(c, c->d_pK, c->pK)) || (rc = dupload(c, c->d_pPx, c->pPx)) ||
        (rc = dupload(c, c->d_pPy, c->pPy)) || (rc = dupload(c, c->d_pPz, c->pPz)))
        return rc;
    if (c->maxI.size() == c->tonto.size()) rc = dupload(c, c->d_maxI, c->maxI);
    else rc = dalloc(c, c->d_maxI, c->tonto.size());
    return rc;
}

static DevPairs dev_pairs(const tgpu_ctx *c)
{
    DevPairs d;
    d.tonto = c->d_tonto.p; d.sh_p = c->d_sp.p; d.sh_q = c->d_sq.p; d.bf_p = c->d_bfp.p;
    d.bf_q = c->d_bfq.p; d.atom_p = c->d_atp.p; d.atom_q = c->d_atq.p;
    d.off_full = c->d_off_full.p; d.n_full = c->d_n_full.p; d.off_skip = c->d_off_skip.p;
    d.n_skip = c->d_n_skip.p; d.A = c->d_A.p; d.AB = c->d_AB.p; d.maxI = c->d_maxI.p;
    return d;
}

static DevPrims dev_prims(const tgpu_ctx *c)
{
    DevPrims d;
    d.zeta = c->d_pz.p; d.K = c->d_pK.p; d.Px = c->d_pPx.p; d.Py = c->d_pPy.p; d.Pz = c->d_pPz.p;
    return d;
}

// Permute all per-pair host arrays by `perm` (new position -> old position).
static void permute_pairs(tgpu_ctx *c, const std::vector<int> &perm)
{
    auto P = [&](auto &v, int stride) {
        auto old = v;
        for (size_t i = 0; i < perm.size(); ++i)
            for (int k = 0; k < stride; ++k) v[i * stride + k] = old[(size_t)perm[i] * stride + k];
    };
    P(c->tonto, 1); P(c->sp, 1); P(c->sq, 1); P(c->bfp, 1); P(c->bfq, 1); P(c->atp, 1);
    P(c->atq, 1); P(c->off_full, 1); P(c->n_full, 1); P(c->off_skip, 1); P(c->n_skip, 1);
    P(c->A, 3); P(c->AB, 3);
    if (c->maxI.size() == perm.size()) P(c->maxI, 1);
}

extern "C" int tgpu_create(const tgpu_system *sys, tgpu_ctx **out)
{
    if (!sys || !out || sys->n_atom <= 0 || sys->n_shell <= 0) return TGPU_E_ARG;
    tgpu_ctx *c = new tgpu_ctx();
    *out = c;
    if (hipGetDevice(&c->device) != hipSuccess) {
        c->err = "no HIP device";
        return TGPU_E_HIP;
    }
    c->n_atom = sys->n_atom;
    c->n_shell = sys->n_shell;
    c->xyz.assign(sys->atom_xyz, sys->atom_xyz + 3 * sys->n_atom);
    c->Z.assign(sys->atom_Z, sys->atom_Z + sys->n_atom);
    c->sh_l.assign(sys->shell_l, sys->shell_l + sys->n_shell);
    c->sh_atom.assign(sys->shell_atom, sys->shell_atom + sys->n_shell);
    c->sh_np.assign(sys->shell_nprim, sys->shell_nprim + sys->n_shell);
    c->sh_off.resize(c->n_shell);
    c->sh_bf.resize(c->n_shell);
    int off = 0, bf = 0;
    for (int s = 0; s < c->n_shell; ++s) {
        if (c->sh_l[s] < 0 || c->sh_l[s] > TGPU_LMAX) {
            c->err = "shell angular momentum " + std::to_string(c->sh_l[s]) +
                     " exceeds this build's maximum " + std::to_string(TGPU_LMAX);
            return TGPU_E_UNSUPPORTED;
        }
        if (c->sh_np[s] <= 0 || c->sh_atom[s] < 0 || c->sh_atom[s] >= c->n_atom) {
            c->err = "invalid shell " + std::to_string(s);
            return TGPU_E_ARG;
        }
        c->sh_off[s] = off;
        c->sh_bf[s] = bf;
        off += c->sh_np[s];
        bf += ncart(c->sh_l[s]);
    }
    c->exps.assign(sys->exps, sys->exps + off);
    c->coefs.assign(sys->coefs, sys->coefs + off);
    c->n_bf = bf;
    c->n_pairs = (long long)c->n_shell * (c->n_shell + 1) / 2;
    // bin Tonto's pairs by class, keeping Tonto order within a bin
    std::vector<std::vector<int>> bins(NCLS);
    for (int a = 0; a < c->n_shell; ++a)
        for (int b = 0; b <= a; ++b) {
            const int lp = std::max(c->sh_l[a], c->sh_l[b]), lq = std::min(c->sh_l[a], c->sh_l[b]);
            bins[pair_class(lp, lq)].push_back(a * (a + 1) / 2 + b);
        }
    c->tonto.clear();
    for (int k = 0; k < NCLS; ++k) {
        c->bin_start[k] = (int)c->tonto.size();
        c->tonto.insert(c->tonto.end(), bins[k].begin(), bins[k].end());
    }
    c->bin_start[NCLS] = (int)c->tonto.size();
    const size_t np = c->tonto.size();
    c->sp.resize(np); c->sq.resize(np); c->bfp.resize(np); c->bfq.resize(np);
    c->atp.resize(np); c->atq.resize(np); c->A.resize(3 * np); c->AB.resize(3 * np);
    for (size_t s = 0; s < np; ++s) {
        const int ab = c->tonto[s];
        int a = (int)((std::sqrt(8.0 * ab + 1.0) - 1.0) / 2.0);
        while (a * (a + 1) / 2 > ab) --a;
        while ((a + 1) * (a + 2) / 2 <= ab) ++a;
        const int b = ab - a * (a + 1) / 2;
        const int p = c->sh_l[a] >= c->sh_l[b] ? a : b, q = p == a ? b : a;
        c->sp[s] = p; c->sq[s] = q;
        c->bfp[s] = c->sh_bf[p]; c->bfq[s] = c->sh_bf[q];
        c->atp[s] = c->sh_atom[p]; c->atq[s] = c->sh_atom[q];
        for (int t = 0; t < 3; ++t) {
            c->A[3 * s + t] = c->xyz[3 * c->sh_atom[p] + t];
            c->AB[3 * s + t] = c->xyz[3 * c->sh_atom[p] + t] - c->xyz[3 * c->sh_atom[q] + t];
        }
    }
    build_prims(c);
    int rc;
    if ((rc = upload_pairs(c))) return rc;
    if ((rc = dupload(c, c->d_sh_l, c->sh_l)) || (rc = dupload(c, c->d_sh_bf, c->sh_bf))) return rc;
    const size_t nn = (size_t)c->n_bf * c->n_bf;
    if ((rc = dalloc(c, c->d_X, nn)) || (rc = dalloc(c, c->d_Y, nn)) ||
        (rc = dalloc(c, c->d_Pmax, (size_t)c->n_pairs)) || (rc = dalloc(c, c->d_scal, 4)) ||
        (rc = dalloc(c, c->d_counters, 3 * NQCLS)))
        return rc;
    HIPCHK(c, hipEventCreate(&c->ev0));
    HIPCHK(c, hipEventCreate(&c->ev1));
    return TGPU_OK;
}

extern "C" int tgpu_destroy(tgpu_ctx *c)
{
    if (!c) return TGPU_E_ARG;
    dfree(c->d_tonto); dfree(c->d_sp); dfree(c->d_sq); dfree(c->d_bfp); dfree(c->d_bfq);
    dfree(c->d_atp); dfree(c->d_atq); dfree(c->d_off_full); dfree(c->d_n_full);
    dfree(c->d_off_skip); dfree(c->d_n_skip); dfree(c->d_sh_l); dfree(c->d_sh_bf);
    dfree(c->d_A); dfree(c->d_AB); dfree(c->d_maxI); dfree(c->d_pz); dfree(c->d_pK);
    dfree(c->d_pPx); dfree(c->d_pPy); dfree(c->d_pPz); dfree(c->d_X); dfree(c->d_Y);
    dfree(c->d_Pmax); dfree(c->d_scal); dfree(c->d_hP); dfree(c->d_hJ); dfree(c->d_hK);
    dfree(c->d_counters);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (auto &e : c->cls_ev) (void)hipEventDestroy(e);
    delete c;
    return TGPU_OK;
}

extern "C" const char *tgpu_last_error(const tgpu_ctx *c)
{
    return c ? c->err.c_str() : "null context";
}

extern "C" int tgpu_dims(const tgpu_ctx *c, int *n_bf, long long *n_pairs)
{
    if (!c) return TGPU_E_ARG;
    if (n_bf) *n_bf = c->n_bf;
    if (n_pairs) *n_pairs = c->n_pairs;
    return TGPU_OK;
}

extern "C" int tgpu_set_stream(tgpu_ctx *c, void *s)
{
    if (!c) return TGPU_E_ARG;
    c->stream = (hipStream_t)s;
    return TGPU_OK;
}

extern "C" int tgpu_set_cutoffs(tgpu_ctx *c, double schwarz, double jp, double kp, double pair)
{
    if (!c || !(schwarz >= 0) || !(jp >= 0) || !(kp >= 0) || !(pair > 0)) {
        if (c) c->err = "cutoffs must be non-negative (pair cutoff positive)";
        return TGPU_E_ARG;
    }
    c->cut_schwarz = schwarz;
    c->cut_jp = jp;
    c->cut_kp = kp;
    if (pair != c->cut_pair) {
        c->cut_pair = pair;
        build_prims(c);
        int rc = upload_pairs(c);
        if (rc) return rc;
    }
    return TGPU_OK;
}

extern "C" int tgpu_set_fp32_threshold(tgpu_ctx *c, double thr)
{
    if (!c || !(thr >= 0)) return TGPU_E_ARG;
    c->fp32_thr = thr;
    return TGPU_OK;
}


extern "C" int tgpu_schwarz(tgpu_ctx *c, double *max_I)
{
    if (!c) return TGPU_E_ARG;
    if (!c->have_schwarz) {
        const DevPairs dpr = dev_pairs(c);
        const DevPrims dpm = dev_prims(c);
#define TGPU_SCHWARZ_CASE(LP, LQ, H)                                                       \
    {                                                                                    \
        const int cls = pair_class(LP, LQ), n = c->bin_start[cls + 1] - c->bin_start[cls]; \
        if (n) tgpu_launch_schwarz_##LP##LQ(dpr, dpm, c->bin_start[cls], n, c->d_maxI.p, c->stream); \
    }
        TGPU_FOR_EACH_PAIRCLASS(TGPU_SCHWARZ_CASE)
#undef TGPU_SCHWARZ_CASE
        HIPCHK(c, hipGetLastError());
        c->maxI.resize(c->tonto.size());
        HIPCHK(c, hipMemcpyAsync(c->maxI.data(), c->d_maxI.p, c->maxI.size() * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        // sort every bin by max_I descending (ties: Tonto order) and re-upload
        std::vector<int> perm(c->tonto.size());
        std::iota(perm.begin(), perm.end(), 0);
        for (int k = 0; k < NCLS; ++k)
            std::sort(perm.begin() + c->bin_start[k], perm.begin() + c->bin_start[k + 1],
                      [&](int x, int y) {
                          if (c->maxI[x] != c->maxI[y]) return c->maxI[x] > c->maxI[y];
                          return c->tonto[x] < c->tonto[y];
                      });
        permute_pairs(c, perm);
        int rc = upload_pairs(c);
        if (rc) return rc;
        c->maxI_tonto.assign(c->n_pairs, 0.0);
        c->maxI_all = 0.0;
        for (size_t s = 0; s < c->tonto.size(); ++s) {
            c->maxI_tonto[c->tonto[s]] = c->maxI[s];
            c->maxI_all = std::max(c->maxI_all, c->maxI[s]);
        }
        c->have_schwarz = true;
    }
    if (max_I) std::memcpy(max_I, c->maxI_tonto.data(), sizeof(double) * c->n_pairs);
    return TGPU_OK;
}

static void launch_jk(tgpu_ctx *c, const double *P, bool mixed, int slot, int X, int Y,
                      void (*launcher)(const JKArgs &, bool, hipStream_t))
{
    if (c->profiling) c->cls_launched[slot] = 0;
    JKArgs a;
    a.pr = dev_pairs(c);
    a.pm = dev_prims(c);
    a.bra_start = c->bin_start[X];
    a.n_bra = c->bin_start[X + 1] - c->bin_start[X];
    a.ket_start = c->bin_start[Y];
    a.n_ket = c->bin_start[Y + 1] - c->bin_start[Y];
    if (a.n_bra == 0 || a.n_ket == 0) return;
    a.same_bin = X == Y;
    a.nbf = c->n_bf;
    a.counter_slot = slot;
    a.P = P;
    a.Pmax = c->d_Pmax.p;
    a.scal = c->d_scal.p;
    a.cutoff = c->cut_schwarz;
    a.jp = c->cut_jp;
    a.kp = c->cut_kp;
    a.fp32_thr = c->fp32_thr;
    a.X = c->d_X.p;
    a.Y = c->d_Y.p;
    a.counters = c->d_counters.p;
    if (c->profiling) {
        (void)hipEventRecord(c->cls_ev[2 * slot], c->stream);
        launcher(a, mixed, c->stream);
        (void)hipEventRecord(c->cls_ev[2 * slot + 1], c->stream);
        c->cls_launched[slot] = 1;
    } else {
        launcher(a, mixed, c->stream);
    }
}

static int grid_for(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 16384); }

extern "C" int tgpu_make_r_JK_device(tgpu_ctx *c, const double *P, double *J, double *K, int flags)
{
    if (!c || !P || !J || !K) return TGPU_E_ARG;
    if (!c->have_schwarz) {
        int rc = tgpu_schwarz(c, nullptr);
        if (rc) return rc;
    }
    const size_t nn = (size_t)c->n_bf * c->n_bf;
    fill_zero_kernel<<<grid_for(nn), 256, 0, c->stream>>>(c->d_X.p, nn);
    fill_zero_kernel<<<grid_for(nn), 256, 0, c->stream>>>(c->d_Y.p, nn);
    fill_zero_u64_kernel<<<(3 * NQCLS + 255) / 256, 256, 0, c->stream>>>(c->d_counters.p, 3 * NQCLS);
    pmax_kernel<<<grid_for((size_t)c->n_pairs), 256, 0, c->stream>>>(
        P, c->n_bf, c->n_shell, c->d_sh_l.p, c->d_sh_bf.p, c->d_Pmax.p);
    global_max_kernel<<<1, 1024, 0, c->stream>>>(c->d_Pmax.p, c->n_pairs, c->maxI_all, c->d_scal.p);
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    const bool mixed = flags & TGPU_JK_MIXED;
    int slot = 0;
#define TGPU_JK_CASE(LP, LQ, LR, LS, H) \
    launch_jk(c, P, mixed, slot++, pair_class(LP, LQ), pair_class(LR, LS), tgpu_launch_jk_##LP##LQ##LR##LS);
    TGPU_FOR_EACH_CLASS(TGPU_JK_CASE)
#undef TGPU_JK_CASE
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    c->timed = true;
    post_kernel<<<grid_for(nn), 256, 0, c->stream>>>(c->d_X.p, c->d_Y.p, c->n_bf, J, K);
    HIPCHK(c, hipGetLastError());
    return TGPU_OK;
}

extern "C" int tgpu_make_r_JK(tgpu_ctx *c, const double *P, double *J, double *K, int flags)
{
    if (!c || !P || !J || !K) return TGPU_E_ARG;
    const size_t nn = (size_t)c->n_bf * c->n_bf;
    int rc;
    if ((rc = dalloc(c, c->d_hP, nn)) || (rc = dalloc(c, c->d_hJ, nn)) || (rc = dalloc(c, c->d_hK, nn)))
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->d_hP.p, P, nn * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if ((rc = tgpu_make_r_JK_device(c, c->d_hP.p, c->d_hJ.p, c->d_hK.p, flags))) return rc;
    HIPCHK(c, hipMemcpyAsync(J, c->d_hJ.p, nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(K, c->d_hK.p, nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TGPU_OK;
}

extern "C" int tgpu_get_stats(tgpu_ctx *c, tgpu_stats *st)
{
    if (!c || !st) return TGPU_E_ARG;
    std::memset(st, 0, sizeof(*st));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<unsigned long long> h(3 * NQCLS);
    HIPCHK(c, hipMemcpy(h.data(), c->d_counters.p, h.size() * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost));
    for (int k = 0; k < NQCLS; ++k) {
        st->quartets += h[3 * k];
        st->prim_quartets += h[3 * k + 1];
        st->fp32_quartets += h[3 * k + 2];
    }
    if (c->timed) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        st->kernel_ms = ms;
    }
    return TGPU_OK;
}

// Per class-pair counters of the last build: out[3*k + {0,1,2}] = quartets, primitive
// quartets, FP32 quartets for the k-th class pair in TGPU_FOR_EACH_CLASS order.
extern "C" int tgpu_get_class_counters(tgpu_ctx *c, unsigned long long *out, int n)
{
    if (!c || !out || n < 3 * NQCLS) return TGPU_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out, c->d_counters.p, 3 * NQCLS * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost));
    return TGPU_OK;
}

extern "C" int tgpu_n_class_pairs(void) { return NQCLS; }

// Per class-pair kernel timing with HIP events on the context's stream (off by default).
extern "C" int tgpu_set_profiling(tgpu_ctx *c, int on)
{
    if (!c) return TGPU_E_ARG;
    if (on && c->cls_ev.empty()) {
        c->cls_ev.resize(2 * NQCLS);
        c->cls_launched.assign(NQCLS, 0);
        for (auto &e : c->cls_ev) HIPCHK(c, hipEventCreate(&e));
    }
    c->profiling = on != 0;
    return TGPU_OK;
}

// ms[k] = device time of class pair k in the last build (0 if not launched).
extern "C" int tgpu_get_class_times(tgpu_ctx *c, double *ms, int n)
{
    if (!c || !ms || n < NQCLS) return TGPU_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < NQCLS; ++k) {
        ms[k] = 0.0;
        if (c->profiling && c->cls_launched[k]) {
            float t = 0.f;
            HIPCHK(c, hipEventElapsedTime(&t, c->cls_ev[2 * k], c->cls_ev[2 * k + 1]));
            ms[k] = t;
        }
    }
    return TGPU_OK;
}

extern "C" int tgpu_one_electron(tgpu_ctx *c, double *S, double *T, double *V)
{
    if (!c || !S || !T || !V) return TGPU_E_ARG;
    const size_t nn = (size_t)c->n_bf * c->n_bf;
    std::vector<int> cx(4 * 10, 0), cy(4 * 10, 0), cz(4 * 10, 0);
    // Tonto's Cartesian order (gaussian_data.foo:295-479) for l <= 3
    static const int tab[4][10][3] = {
        {{0, 0, 0}},
        {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
        {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}},
        {{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {0, 2, 1}, {1, 0, 2}, {0, 1, 2}, {1, 1, 1}}};
    for (int l = 0; l < 4; ++l)
        for (int i = 0; i < ncart(l); ++i) {
            cx[l * 10 + i] = tab[l][i][0];
            cy[l * 10 + i] = tab[l][i][1];
            cz[l * 10 + i] = tab[l][i][2];
        }
    DBuf<int> d_atom, d_off, d_np, d_cx, d_cy, d_cz;
    DBuf<double> d_e, d_c, d_xyz, d_Z, d_S, d_T, d_V;
    int rc;
    auto cleanup = [&]() {
        dfree(d_atom); dfree(d_off); dfree(d_np); dfree(d_cx); dfree(d_cy); dfree(d_cz);
        dfree(d_e); dfree(d_c); dfree(d_xyz); dfree(d_Z); dfree(d_S); dfree(d_T); dfree(d_V);
    };
    if ((rc = dupload(c, d_atom, c->sh_atom)) || (rc = dupload(c, d_off, c->sh_off)) ||
        (rc = dupload(c, d_np, c->sh_np)) || (rc = dupload(c, d_cx, cx)) || (rc = dupload(c, d_cy, cy)) ||
        (rc = dupload(c, d_cz, cz)) || (rc = dupload(c, d_e, c->exps)) || (rc = dupload(c, d_c, c->coefs)) ||
        (rc = dupload(c, d_xyz, c->xyz)) || (rc = dupload(c, d_Z, c->Z)) || (rc = dalloc(c, d_S, nn)) ||
        (rc = dalloc(c, d_T, nn)) || (rc = dalloc(c, d_V, nn))) {
        cleanup();
        return rc;
    }
    OneEArgs a;
    a.nshell = c->n_shell; a.natom = c->n_atom; a.nbf = c->n_bf;
    a.sh_l = c->d_sh_l.p; a.sh_bf = c->d_sh_bf.p; a.sh_atom = d_atom.p; a.sh_off = d_off.p;
    a.sh_np = d_np.p; a.cx = d_cx.p; a.cy = d_cy.p; a.cz = d_cz.p; a.exps = d_e.p; a.coefs = d_c.p;
    a.xyz = d_xyz.p; a.Z = d_Z.p; a.S = d_S.p; a.T = d_T.p; a.V = d_V.p;
    one_e_kernel<<<(int)((c->n_pairs + 63) / 64), 64, 0, c->stream>>>(a);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(S, d_S.p, nn * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(T, d_T.p, nn * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(V, d_V.p, nn * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    cleanup();
    if (e != hipSuccess) {
        c->err = std::string("one-electron integrals: ") + hipGetErrorString(e);
        return TGPU_E_HIP;
    }
    return TGPU_OK;
}

/* ORACLE -- test infrastructure only.
 *
 * A plain-C, FP64, CPU restatement of Tonto's restricted direct-SCF J/K build for a
 * Cartesian basis, used by tests/ (as the checker), by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg.  It is never linked into, loaded by, or called from the
 * product library (tonto_amd/csrc); the product fails loudly if its HIP library is
 * missing.
 *
 * Reference functions restated here (file:line under the Tonto tree):
 *   make_r_JK_engine ............ foofiles/molecule.fock.foo:1365-1505
 *   Schwarz_test ................ foofiles/molecule.fock.foo:3497-3536
 *   make_max_density_elements ... foofiles/molecule.fock.foo:3601-3633
 *   make_max_abab_integrals ..... foofiles/molecule.fock.foo:3712-3752
 *   weight_diagonal_blocks ...... foofiles/molecule.fock.foo:4181-4195
 *   blockwise_symmetric_fold .... foofiles/molecule.fock.foo:4245-4263
 *   symmetric_reflect ........... foofiles/mat{intrinsic}.foo:4672-4700
 *   make_precalculated_data ..... foofiles/shell1pair.foo:100-156
 *   set_ab_new / set_cd_new ..... foofiles/shell1quartet.foo:131-420 (pair skip rule)
 *   make_esfs_XX (Rys VRR) ...... foofiles/shell1quartet.foo:7420-7609
 *   transfer_cd / transfer_ab ... foofiles/shell1quartet.foo:8310-11134 (HRR)
 *   make_r_JK_engine (quartet) .. foofiles/shell1quartet.foo:11386-11528 (K scatter)
 * Rys roots/weights: either the FP64 Chebyshev table (tonto_amd/csrc/rys_table.inc,
 * itself pinned against rys_quad.c) or the quad-precision Golub-Welsch of rys_quad.c.
 *
 * J is accumulated directly from (ab|cd) instead of through Tonto's J-engine
 * (reverse_transfer / forward_transfer, molecule.fock.foo:3823-4150); the two are the
 * same linear map (the HRR is linear in the (e0| source), so this changes rounding
 * only.
 *
 * Layouts: matrices are n_bf x n_bf, row-major (they are symmetric).  Shell pairs use
 * Tonto's order: pair index ab = a(a+1)/2 + b (0-based, a >= b).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include <unistd.h>

#include "../tonto_amd/csrc/rys_table.inc"

int orq_rys(int n, double X, double *u, double *w);

#define TWOPI5ON2 34.9868366552497256925   /* include/macros.in:446 */
#define LGMAX 8                             /* component tables up to l = 8 */
#define NGMAX 165                           /* n_comp_up_to(8) */
#define MAXPRIM 16

typedef struct {
    int n_shell, n_atom, n_bf;
    const int *l, *nprim, *atom, *first_bf, *prim_off;
    const double *exps, *coefs, *atom_xyz, *atom_Z;
} osys_t;

typedef struct {
    double schwarz, jp, kp, pair;   /* ERI_Schwarz_cutoff, ERI_J/K_density_cutoff, pair_cutoff */
    int rys_quad;                   /* 1: quad-precision Golub-Welsch roots */
} ocut_t;

/* ---------------- threads (plain pthreads: no OpenMP runtime, so the oracle can share a
 * process with torch's own OpenMP copy) ------------------------------------------ */
static int g_threads = 0;

int orc_max_threads(void)
{
    if (g_threads <= 0) {
        const char *e = getenv("ORACLE_THREADS");
        g_threads = e ? atoi(e) : (int)sysconf(_SC_NPROCESSORS_ONLN);
        if (g_threads <= 0) g_threads = 1;
    }
    return g_threads;
}
void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }

typedef void (*work_fn)(void *arg, long item, int worker);
typedef struct { work_fn fn; void *arg; long n; long next; pthread_mutex_t mu; int id; } pool_t;
typedef struct { pool_t *pool; int worker; } wk_t;

static void *pool_main(void *p)
{
    wk_t *w = (wk_t *)p;
    for (;;) {
        long item = __atomic_fetch_add(&w->pool->next, 1, __ATOMIC_RELAXED);
        if (item >= w->pool->n) break;
        w->pool->fn(w->pool->arg, item, w->worker);
    }
    return NULL;
}

/* run fn(arg, item, worker) for item in [0,n) on nthr workers, dynamic schedule */
static void parallel_for(long n, int nthr, work_fn fn, void *arg)
{
    pool_t pool = {fn, arg, n, 0};
    pthread_t th[256];
    wk_t wk[256];
    if (nthr > 256) nthr = 256;
    for (int t = 0; t < nthr; t++) {
        wk[t].pool = &pool;
        wk[t].worker = t;
        if (t > 0) pthread_create(&th[t], NULL, pool_main, &wk[t]);
    }
    pool_main(&wk[0]);
    for (int t = 1; t < nthr; t++) pthread_join(th[t], NULL);
}

/* ---------------- Cartesian component tables (gaussian_data.foo:295-479) ----- */
static int g_ready = 0;
static int g_xyz[NGMAX][3], g_l[NGMAX], g_p1[NGMAX][3];
static int g_index[LGMAX + 2][LGMAX + 2][LGMAX + 2];

static int ncomp(int l) { return (l + 1) * (l + 2) / 2; }
static int ncomp_up_to(int l) { return l < 0 ? 0 : (l + 1) * (l + 2) * (l + 3) / 6; }

static void comp_tables(void)
{
    if (g_ready) return;
    int i = 0;
    for (int L = 0; L <= LGMAX; L++) {
        for (int a = L; a >= (L + 2) / 3; a--) {
            for (int b = (L - a < a ? L - a : a); b >= (L - a + 1) / 2; b--) {
                int c = L - a - b, pat[6][3], np;
                if (a == b && b == c) { np = 1; int t[1][3] = {{a, a, a}}; memcpy(pat, t, sizeof t); }
                else if (a > b && b == c) { np = 3; int t[3][3] = {{a, b, b}, {b, a, b}, {b, b, a}}; memcpy(pat, t, sizeof t); }
                else if (a == b && b > c) { np = 3; int t[3][3] = {{a, a, c}, {a, c, a}, {c, a, a}}; memcpy(pat, t, sizeof t); }
                else { np = 6; int t[6][3] = {{a, b, c}, {a, c, b}, {b, a, c}, {c, a, b}, {b, c, a}, {c, b, a}}; memcpy(pat, t, sizeof t); }
                for (int k = 0; k < np; k++, i++) {
                    g_xyz[i][0] = pat[k][0]; g_xyz[i][1] = pat[k][1]; g_xyz[i][2] = pat[k][2];
                    g_l[i] = L;
                    g_index[pat[k][0]][pat[k][1]][pat[k][2]] = i;
                }
            }
        }
    }
    for (int g = 0; g < NGMAX; g++)
        for (int t = 0; t < 3; t++) {
            int x[3] = {g_xyz[g][0], g_xyz[g][1], g_xyz[g][2]};
            x[t]++;
            g_p1[g][t] = (g_l[g] < LGMAX) ? g_index[x[0]][x[1]][x[2]] : -1;
        }
    g_ready = 1;
}

/* component (lx,ly,lz) of the i-th member of an l-shell */
void orc_component(int l, int i, int *xyz)
{
    comp_tables();
    int g = ncomp_up_to(l - 1) + i;
    xyz[0] = g_xyz[g][0]; xyz[1] = g_xyz[g][1]; xyz[2] = g_xyz[g][2];
}

/* ---------------- Rys roots/weights --------------------------------------------- */
void orc_rys_table(int n, double X, double *u, double *w)
{
    if (X >= rys_xasym[n]) {
        double ix = 1.0 / X, isx = sqrt(ix);
        for (int i = 0; i < n; i++) {
            u[i] = rys_herm_s2[(n - 1) * RYS_NMAX + i] * ix;
            w[i] = rys_herm_w[(n - 1) * RYS_NMAX + i] * isx;
        }
        return;
    }
    int k = (int)(X * (1.0 / RYS_H));
    double t = (X - k * RYS_H) * (2.0 / RYS_H) - 1.0;
    const double *c = rys_table_data + rys_table_offset[n] + (size_t)k * RYS_NCHEB * 2 * n;
    for (int f = 0; f < 2 * n; f++) {
        double b1 = 0.0, b2 = 0.0;
        for (int m = RYS_NCHEB - 1; m >= 1; m--) {
            double b0 = 2.0 * t * b1 - b2 + c[m * 2 * n + f];
            b2 = b1; b1 = b0;
        }
        double v = t * b1 - b2 + c[f];
        if (f < n) u[f] = v; else w[f - n] = v;
    }
}

static void rys(int n, double X, double *u, double *w, int quad)
{
    if (quad) orq_rys(n, X, u, w);
    else orc_rys_table(n, X, u, w);
}

/* ---------------- primitive-pair data (shell1pair.foo:100-156, shell1quartet.foo:131-275) */
typedef struct {
    int n;                 /* surviving primitive pairs */
    double zeta[MAXPRIM * MAXPRIM], K[MAXPRIM * MAXPRIM];
    double P[MAXPRIM * MAXPRIM][3], PA[MAXPRIM * MAXPRIM][3];
    double R[3];           /* big - small centre (HRR) */
    int big_is_first;      /* Tonto: At = pos_a if la > lb else pos_b */
} pairdat_t;

/* per-worker scratch of one quartet evaluation: the two pairs, the (e0|f0) block and the
 * transfer-relation buffers (sizes cover (ff|ff): 74 x 74 (e0|f0) entries, transfer levels
 * up to 150 x 100) */
#define QBUF 16384
typedef struct {
    pairdat_t pd[2];
    double G[QBUF], escd[QBUF], abcd[QBUF], t0[QBUF], t1[QBUF];
} qws_t;

static void make_pair(const osys_t *s, int a, int b, int skip, double pair_cut, pairdat_t *pd)
{
    const double *A = s->atom_xyz + 3 * s->atom[a], *B = s->atom_xyz + 3 * s->atom[b];
    int same = s->atom[a] == s->atom[b];
    double AB[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
    double r2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
    pd->big_is_first = s->l[a] > s->l[b];
    const double *At = pd->big_is_first ? A : B;
    const double *Sm = pd->big_is_first ? B : A;
    for (int t = 0; t < 3; t++) pd->R[t] = At[t] - Sm[t];
    int j = 0;
    const double *ea = s->exps + s->prim_off[a], *ca = s->coefs + s->prim_off[a];
    const double *eb = s->exps + s->prim_off[b], *cb = s->coefs + s->prim_off[b];
    for (int bg = 0; bg < s->nprim[b]; bg++) {
        double be = eb[bg], b_r2 = be * r2;
        for (int ag = 0; ag < s->nprim[a]; ag++) {
            double ae = ea[ag];
            double sum = ae + be, inv = 1.0 / sum;
            double cc = cb[bg] * ca[ag] * inv * sqrt(inv);
            double pref;
            if (same && !skip) {
                pref = cc;
            } else {
                double e = exp(-ae * b_r2 * inv);
                if (skip && e < pair_cut) continue;
                pref = cc * e;
            }
            pd->zeta[j] = sum;
            pd->K[j] = pref;
            for (int t = 0; t < 3; t++) {
                pd->P[j][t] = same && !skip ? A[t] : (be * B[t] + ae * A[t]) * inv;
                pd->PA[j][t] = pd->P[j][t] - At[t];
            }
            j++;
        }
    }
    pd->n = j;
}

/* ---------------- HRR (transfer relations) ------------------------------------ */
/* in : [lead][comps with l in lbig..lbig+lsm][trail]
 * out: [lead][ncomp(lbig)][ncomp(lsm)][trail]
 * (a, b+1_t) = (a+1_t, b) + R_t (a, b), R = big - small centre. */
static void hrr(int lbig, int lsm, const double *R, const double *in, int lead, int trail,
                double *out, double *t0, double *t1)
{
    int base = ncomp_up_to(lbig - 1);
    int nin = ncomp_up_to(lbig + lsm) - base;
    if (lsm == 0) {
        memcpy(out, in, sizeof(double) * lead * nin * trail);
        return;
    }
    /* level k: a in l range [lbig, lbig+lsm-k] (na_k comps), b comps of l = k */
    const double *cur = in;
    int na = nin, nb = 1;
    for (int k = 0; k < lsm; k++) {
        int na1 = ncomp_up_to(lbig + lsm - k - 1) - base, nb1 = ncomp(k + 1);
        double *nxt = k == lsm - 1 ? out : (k % 2 ? t1 : t0);
        if ((long)lead * na1 * nb1 * trail > QBUF) abort();
        int gb0 = ncomp_up_to(k - 1), gb1 = ncomp_up_to(k);
        for (int ib1 = 0; ib1 < nb1; ib1++) {
            int g1 = gb1 + ib1, t = 0;
            while (g_xyz[g1][t] == 0) t++;
            int xb[3] = {g_xyz[g1][0], g_xyz[g1][1], g_xyz[g1][2]};
            xb[t]--;
            int ib = g_index[xb[0]][xb[1]][xb[2]] - gb0;
            for (int L = 0; L < lead; L++)
                for (int ia = 0; ia < na1; ia++) {
                    int ga = base + ia, ia1 = g_p1[ga][t] - base;
                    for (int T = 0; T < trail; T++)
                        nxt[((L * na1 + ia) * nb1 + ib1) * trail + T] =
                            cur[((L * na + ia1) * nb + ib) * trail + T] +
                            R[t] * cur[((L * na + ia) * nb + ib) * trail + T];
                }
        }
        cur = nxt; na = na1; nb = nb1;
    }
    /* the last level wrote out: na == ncomp(lbig) */
}

/* ---------------- one shell quartet (ab|cd) ------------------------------------- */
/* out[((ia*nb + ib)*nc + ic)*nd + id]; returns number of primitive quartets */
static long quartet(const osys_t *s, int a, int b, int c, int d, int skip, const ocut_t *cut,
                    double *out, qws_t *ws)
{
    comp_tables();
    pairdat_t *pab = &ws->pd[0], *pcd = &ws->pd[1];
    make_pair(s, a, b, skip, cut->pair, pab);
    make_pair(s, c, d, skip, cut->pair, pcd);
    int la = s->l[a], lb = s->l[b], lc = s->l[c], ld = s->l[d];
    int lab = la + lb, lcd = lc + ld;
    int lbig_ab = la > lb ? la : lb, lsm_ab = la > lb ? lb : la;
    int lbig_cd = lc > ld ? lc : ld, lsm_cd = lc > ld ? ld : lc;
    int e0 = ncomp_up_to(lbig_ab - 1), ne = ncomp_up_to(lab) - e0;
    int f0 = ncomp_up_to(lbig_cd - 1), nf = ncomp_up_to(lcd) - f0;
    int nroot = (lab + lcd) / 2 + 1;
    if (ne * nf > QBUF) abort();
    double *G = ws->G;
    memset(G, 0, sizeof(double) * ne * nf);
    double Ix[9][9], Iy[9][9], Iz[9][9];
    double u[16], w[16];
    for (int p = 0; p < pab->n; p++) {
        double zeta = pab->zeta[p], zinv = 1.0 / zeta;
        for (int q = 0; q < pcd->n; q++) {
            double eta = pcd->zeta[q], einv = 1.0 / eta;
            double rho = zeta * eta / (zeta + eta);
            double QP[3] = {pcd->P[q][0] - pab->P[p][0], pcd->P[q][1] - pab->P[p][1],
                            pcd->P[q][2] - pab->P[p][2]};
            double X = rho * (QP[0] * QP[0] + QP[1] * QP[1] + QP[2] * QP[2]);
            rys(nroot, X, u, w, cut->rys_quad);
            double pref = TWOPI5ON2 * pab->K[p] * pcd->K[q] * sqrt(rho);
            for (int r = 0; r < nroot; r++) {
                double t2 = u[r];
                double rz = t2 * rho * zinv, re = t2 * rho * einv;
                double b10 = (1.0 - rz) * 0.5 * zinv, b01 = (1.0 - re) * 0.5 * einv;
                double b00 = rz * 0.5 * einv;
                double (*I[3])[9] = {Ix, Iy, Iz};
                for (int t = 0; t < 3; t++) {
                    double (*M)[9] = I[t];
                    double c00 = pab->PA[p][t] + rz * QP[t];
                    double d00 = pcd->PA[q][t] - re * QP[t];
                    M[0][0] = 1.0;
                    for (int e = 0; e < lab; e++)
                        M[e + 1][0] = c00 * M[e][0] + (e ? e * b10 * M[e - 1][0] : 0.0);
                    for (int f = 0; f < lcd; f++)
                        for (int e = 0; e <= lab; e++)
                            M[e][f + 1] = d00 * M[e][f] + (f ? f * b01 * M[e][f - 1] : 0.0) +
                                          (e ? e * b00 * M[e - 1][f] : 0.0);
                }
                double wt = w[r] * pref;
                for (int ie = 0; ie < ne; ie++) {
                    const int *xe = g_xyz[e0 + ie];
                    for (int jf = 0; jf < nf; jf++) {
                        const int *xf = g_xyz[f0 + jf];
                        G[ie * nf + jf] += Ix[xe[0]][xf[0]] * Iy[xe[1]][xf[1]] * Iz[xe[2]][xf[2]] * wt;
                    }
                }
            }
        }
    }
    long nprimq = (long)pab->n * pcd->n;
    /* HRR on cd (trailing index), then ab */
    int nc = ncomp(lc), nd = ncomp(ld), na_ = ncomp(la), nb_ = ncomp(lb);
    int nbig_cd = ncomp(lbig_cd), nsm_cd = ncomp(lsm_cd);
    double *escd = ws->escd;
    hrr(lbig_cd, lsm_cd, pcd->R, G, ne, 1, escd, ws->t0, ws->t1);   /* [e][big][small] */
    int nsm_ab = ncomp(lsm_ab);
    double *abcd = ws->abcd;
    hrr(lbig_ab, lsm_ab, pab->R, escd, 1, nbig_cd * nsm_cd, abcd, ws->t0, ws->t1); /* [big][small][cdbig][cdsmall] */
    for (int ia = 0; ia < na_; ia++)
        for (int ib = 0; ib < nb_; ib++)
            for (int ic = 0; ic < nc; ic++)
                for (int id = 0; id < nd; id++) {
                    int i1 = pab->big_is_first ? ia : ib, i2 = pab->big_is_first ? ib : ia;
                    int i3 = pcd->big_is_first ? ic : id, i4 = pcd->big_is_first ? id : ic;
                    out[((ia * nb_ + ib) * nc + ic) * nd + id] =
                        abcd[((i1 * nsm_ab + i2) * nbig_cd + i3) * nsm_cd + i4];
                }
    return nprimq;
}

/* Public: one Cartesian shell quartet, with or without primitive-pair skipping. */
long orc_eri_quartet(const osys_t *s, int a, int b, int c, int d, int skip, double pair_cut,
                     int rys_quad, double *out)
{
    ocut_t cut = {0, 0, 0, pair_cut, rys_quad};
    qws_t *ws = malloc(sizeof(qws_t));
    long n = quartet(s, a, b, c, d, skip, &cut, out, ws);
    free(ws);
    return n;
}

/* ---------------- Schwarz integrals (molecule.fock.foo:3712-3752) ---------------- */
typedef struct { const osys_t *s; ocut_t cut; double *max_I; } schwarz_job;

static void schwarz_item(void *arg, long item, int worker)
{
    (void)worker;
    schwarz_job *jb = (schwarz_job *)arg;
    const osys_t *s = jb->s;
    int a = (int)item;
    {
        double *buf = malloc(sizeof(double) * 10 * 10 * 10 * 10);
        qws_t *ws = malloc(sizeof(qws_t));
        for (int b = 0; b <= a; b++) {
            quartet(s, a, b, a, b, 0, &jb->cut, buf, ws);
            int na = ncomp(s->l[a]), nb = ncomp(s->l[b]);
            double m = 0.0;
            for (int ia = 0; ia < na; ia++)
                for (int ib = 0; ib < nb; ib++) {
                    double v = fabs(buf[((ia * nb + ib) * na + ia) * nb + ib]);
                    if (v > m) m = v;
                }
            jb->max_I[a * (a + 1) / 2 + b] = sqrt(m);
        }
        free(ws);
        free(buf);
    }
}

void orc_schwarz(const osys_t *s, int rys_quad, double *max_I)
{
    comp_tables();
    schwarz_job jb = {s, {0, 0, 0, 1e-6, rys_quad}, max_I};
    parallel_for(s->n_shell, orc_max_threads(), schwarz_item, &jb);
}

/* ---------------- max density per shell pair (molecule.fock.foo:3601-3633) ------ */
static void max_density(const osys_t *s, const double *P, double *Pmax)
{
    int ns = s->n_shell, n = s->n_bf;
    for (int a = 0; a < ns; a++)
        for (int b = 0; b <= a; b++) {
            double m = 0.0;
            for (int i = 0; i < ncomp(s->l[a]); i++)
                for (int j = 0; j < ncomp(s->l[b]); j++) {
                    double v = fabs(P[(size_t)(s->first_bf[a] + i) * n + s->first_bf[b] + j]);
                    if (v > m) m = v;
                }
            Pmax[a * (a + 1) / 2 + b] = m;
        }
}

/* UHF: max(|Pa|, |Pb|) per shell pair (molecule.fock.foo:3670-3701) */
static void max_density2(const osys_t *s, const double *Pa, const double *Pb, double *Pmax)
{
    int ns = s->n_shell, n = s->n_bf;
    for (int a = 0; a < ns; a++)
        for (int b = 0; b <= a; b++) {
            double m = 0.0;
            for (int i = 0; i < ncomp(s->l[a]); i++)
                for (int j = 0; j < ncomp(s->l[b]); j++) {
                    size_t e = (size_t)(s->first_bf[a] + i) * n + s->first_bf[b] + j;
                    double v = fabs(Pa[e]), w = fabs(Pb[e]);
                    if (w > v) v = w;
                    if (v > m) m = v;
                }
            Pmax[a * (a + 1) / 2 + b] = m;
        }
}

static inline int pidx(int x, int y) { return x >= y ? x * (x + 1) / 2 + y : y * (y + 1) / 2 + x; }

typedef struct {
    const osys_t *s;
    const ocut_t *cut;
    const double *P, *max_I, *Pmax;
    const double *Pb;             /* UHF: the B-spin density (NULL for RHF) */
    double **Jbt, **Kbt;          /* UHF: per-worker B-spin J and K */
    double P_max, IP_max;
    double **Jt, **Kt, **Ibuf;
    qws_t **ws;
    long (*st)[5];
    long sample_stride;
} jk_job;

/* Planted screening mistakes, for the test that shows the screening audit catches them
 * (tests/test_oracle_audit.py); 0 (the default) everywhere else.
 *   1: KP reads the (b,c) block twice instead of (b,c) and (b,d)
 *   2: skipped primitive-pair lists also when a and c (or b and d) share an atom
 *   3: JP is the smaller of the two pair maxima instead of the larger */
static int g_plant = 0;
void orc_set_plant(int k) { g_plant = k; }

/* Tonto's row test for the bra pair (molecule.fock.foo:1412): 1 if the row survives */
static int ab_survives(const jk_job *jb, long ab)
{
    return !(jb->max_I[ab] * jb->IP_max < jb->cut->schwarz);
}

/* Tonto uses the skipped primitive-pair lists only when no two of the four atoms
 * coincide (molecule.fock.foo:1456-1472) */
static int uses_skip_lists(const osys_t *s, int a, int b, int c, int d)
{
    const int A = s->atom[a], B = s->atom[b], C = s->atom[c], D = s->atom[d];
    if (g_plant == 2) return !(A == B || C == D || A == D || B == C);
    return !(A == B || C == D || A == C || A == D || B == C || B == D);
}

/* Tonto's tests of quartet (ab|cd) after the row test: the Schwarz test with P_max
 * (molecule.fock.foo:1430) and Schwarz_test (3497-3536, triangular indices 3516-3521).
 * Returns 0 if the quartet is dropped, else bit 0 = J digested, bit 1 = K digested. */
static int screen_quartet(const jk_job *jb, long ab, long cd, int a, int b, int c, int d)
{
    const double *max_I = jb->max_I, *Pmax = jb->Pmax;
    const ocut_t *cut = jb->cut;
    if (max_I[ab] * max_I[cd] * jb->P_max < cut->schwarz) return 0;
    double JP = Pmax[ab] > Pmax[cd] ? Pmax[ab] : Pmax[cd];
    if (g_plant == 3) JP = Pmax[ab] < Pmax[cd] ? Pmax[ab] : Pmax[cd];
    double KP = Pmax[pidx(a, c)];
    if (Pmax[pidx(a, d)] > KP) KP = Pmax[pidx(a, d)];
    if (Pmax[pidx(b, c)] > KP) KP = Pmax[pidx(b, c)];
    const long bd = g_plant == 1 ? pidx(b, c) : pidx(b, d);
    if (Pmax[bd] > KP) KP = Pmax[bd];
    int skip_J = (JP < cut->jp) || (JP * max_I[ab] * max_I[cd] < cut->schwarz);
    int skip_K = (KP < cut->kp) || (KP * max_I[ab] * max_I[cd] < cut->schwarz);
    return (!skip_J) | ((!skip_K) << 1);
}

static void decode(long ab, int *a, int *b)
{
    int x = (int)((sqrt(8.0 * ab + 1.0) - 1.0) / 2.0);
    while (x * (x + 1) / 2 > ab) x--;
    while ((x + 1) * (x + 2) / 2 <= ab) x++;
    *a = x;
    *b = (int)(ab - (long)x * (x + 1) / 2);
}

static void jk_item(void *arg, long ab, int w)
{
    jk_job *jb = (jk_job *)arg;
    const osys_t *s = jb->s;
    const ocut_t *cut = jb->cut;
    const double *P = jb->P, *Pb = jb->Pb;
    int n = s->n_bf;
    double *Jl = jb->Jt[w], *Kl = jb->Kt[w], *I = jb->Ibuf[w];
    double *Jbl = Pb ? jb->Jbt[w] : NULL, *Kbl = Pb ? jb->Kbt[w] : NULL;
    long *st = jb->st[w];
    if (!ab_survives(jb, ab)) return;
    int a, b;
    decode(ab, &a, &b);
    for (long cd = 0; cd <= ab; cd++) {
        int c, d;
        decode(cd, &c, &d);
        const int code = screen_quartet(jb, ab, cd, a, b, c, d);
        if (!code) continue;
        const int skip_J = !(code & 1), skip_K = !(code & 2);
        double f = (a == b) ? 0.5 : 1.0;
        if (c == d) f *= 0.5;
        if (a == c && b == d) f *= 0.5;
        const int any_same = !uses_skip_lists(s, a, b, c, d);
        long npq = quartet(s, a, b, c, d, !any_same, cut, I, jb->ws[w]);
        st[0]++; st[1] += npq; st[2] += skip_J; st[3] += skip_K; st[4] += !any_same;
        int na = ncomp(s->l[a]), nb = ncomp(s->l[b]), nc = ncomp(s->l[c]), nd = ncomp(s->l[d]);
        int fa = s->first_bf[a], fb = s->first_bf[b], fc = s->first_bf[c], fd = s->first_bf[d];
        for (int ia = 0; ia < na; ia++)
            for (int ib = 0; ib < nb; ib++)
                for (int ic = 0; ic < nc; ic++)
                    for (int id = 0; id < nd; id++) {
                        double v = f * I[((ia * nb + ib) * nc + ic) * nd + id];
                        int i = fa + ia, j = fb + ib, k = fc + ic, l = fd + id;
                        if (!skip_J) {
                            Jl[(size_t)i * n + j] += v * P[(size_t)k * n + l];
                            Jl[(size_t)k * n + l] += v * P[(size_t)i * n + j];
                            if (Pb) {
                                /* UHF (shell1quartet.foo:11530-11714): B-spin bra side into JB;
                                 * the B-spin ket side goes into JA, because the engine points
                                 * JBcd at JA(cd) (molecule.fock.foo:2637), except for ab = cd,
                                 * where Tonto doubles the bra update of JB instead */
                                Jbl[(size_t)i * n + j] += v * Pb[(size_t)k * n + l];
                                if (a == c && b == d) Jbl[(size_t)k * n + l] += v * Pb[(size_t)i * n + j];
                                else Jl[(size_t)k * n + l] += v * Pb[(size_t)i * n + j];
                            }
                        }
                        if (!skip_K) {
                            Kl[(size_t)i * n + k] += v * P[(size_t)l * n + j];
                            Kl[(size_t)i * n + l] += v * P[(size_t)k * n + j];
                            Kl[(size_t)j * n + k] += v * P[(size_t)l * n + i];
                            Kl[(size_t)j * n + l] += v * P[(size_t)k * n + i];
                            if (Pb) {
                                Kbl[(size_t)i * n + k] += v * Pb[(size_t)l * n + j];
                                Kbl[(size_t)i * n + l] += v * Pb[(size_t)k * n + j];
                                Kbl[(size_t)j * n + k] += v * Pb[(size_t)l * n + i];
                                Kbl[(size_t)j * n + l] += v * Pb[(size_t)k * n + i];
                            }
                        }
                    }
    }
}

static void jk_item_sample(void *arg, long item, int w)
{
    jk_job *jb = (jk_job *)arg;
    jk_item(arg, item * jb->sample_stride, w);
}

/* ---------------- the J/K build (molecule.fock.foo:1365-1505) ------------------- */
/* stats (may be NULL): [0] quartets computed, [1] primitive quartets, [2] J-skips,
 * [3] K-skips, [4] quartets using skipped primitive-pair lists */
static void jk_item_sample(void *arg, long item, int w);

static void post_jk(const osys_t *s, double *J, double *K);

static void make_r_JK_impl(const osys_t *s, const ocut_t *cut, const double *P, const double *max_I_in,
                           double *J, double *K, long *stats, long sample_stride,
                           const double *Pb, double *Jb, double *Kb)
{
    comp_tables();
    int ns = s->n_shell, n = s->n_bf;
    long npair = (long)ns * (ns + 1) / 2;
    double *max_I = malloc(sizeof(double) * npair);
    if (max_I_in) memcpy(max_I, max_I_in, sizeof(double) * npair);
    else orc_schwarz(s, cut->rys_quad, max_I);
    double *Pmax = malloc(sizeof(double) * npair);
    if (Pb) max_density2(s, P, Pb, Pmax);
    else max_density(s, P, Pmax);
    double P_max = 0.0, I_max = 0.0;
    for (long i = 0; i < npair; i++) {
        if (Pmax[i] > P_max) P_max = Pmax[i];
        if (max_I[i] > I_max) I_max = max_I[i];
    }
    int nthr = orc_max_threads();
    jk_job jb;
    jb.s = s; jb.cut = cut; jb.P = P; jb.max_I = max_I; jb.Pmax = Pmax;
    jb.Pb = Pb;
    jb.P_max = P_max; jb.IP_max = I_max * P_max;
    jb.Jt = malloc(sizeof(double *) * nthr);
    jb.Kt = malloc(sizeof(double *) * nthr);
    jb.Jbt = malloc(sizeof(double *) * nthr);
    jb.Kbt = malloc(sizeof(double *) * nthr);
    jb.Ibuf = malloc(sizeof(double *) * nthr);
    jb.ws = malloc(sizeof(qws_t *) * nthr);
    jb.st = calloc(nthr, sizeof(long[5]));
    for (int t = 0; t < nthr; t++) {
        jb.Jt[t] = calloc((size_t)n * n, sizeof(double));
        jb.Kt[t] = calloc((size_t)n * n, sizeof(double));
        jb.Jbt[t] = Pb ? calloc((size_t)n * n, sizeof(double)) : NULL;
        jb.Kbt[t] = Pb ? calloc((size_t)n * n, sizeof(double)) : NULL;
        jb.Ibuf[t] = malloc(sizeof(double) * 10 * 10 * 10 * 10);
        jb.ws[t] = malloc(sizeof(qws_t));
    }
    if (sample_stride > 1) {
        /* bounded sample of the same build: every sample_stride-th (ab| pair */
        jb.sample_stride = sample_stride;
        parallel_for((npair + sample_stride - 1) / sample_stride, nthr, jk_item_sample, &jb);
    } else {
        parallel_for(npair, nthr, jk_item, &jb);
    }
    memset(J, 0, sizeof(double) * n * n);
    memset(K, 0, sizeof(double) * n * n);
    if (Pb) {
        memset(Jb, 0, sizeof(double) * n * n);
        memset(Kb, 0, sizeof(double) * n * n);
    }
    long st[5] = {0, 0, 0, 0, 0};
    for (int t = 0; t < nthr; t++) {
        for (long i = 0; i < (long)n * n; i++) { J[i] += jb.Jt[t][i]; K[i] += jb.Kt[t][i]; }
        if (Pb)
            for (long i = 0; i < (long)n * n; i++) { Jb[i] += jb.Jbt[t][i]; Kb[i] += jb.Kbt[t][i]; }
        for (int k = 0; k < 5; k++) st[k] += jb.st[t][k];
        free(jb.Jt[t]); free(jb.Kt[t]); free(jb.Ibuf[t]); free(jb.ws[t]);
        free(jb.Jbt[t]); free(jb.Kbt[t]);
    }
    free(jb.Jt); free(jb.Kt); free(jb.Ibuf); free(jb.ws); free(jb.st); free(jb.Jbt); free(jb.Kbt);
    post_jk(s, J, K);
    if (Pb) post_jk(s, Jb, Kb);
    if (stats) memcpy(stats, st, sizeof st);
    free(max_I); free(Pmax);
}

/* post-processing (molecule.fock.foo:1485-1500, 2688-2706 for UHF):
 * weight_diagonal_blocks(J,2); blockwise_symmetric_fold(K); J = 2J; symmetric_reflect
 * (lower -> upper) */
/* post-processing over shell blocks given by first function and size (Cartesian for the
 * engine, spherical for the direct build) */
static void post_jk_blocks(int ns, const int *first, const int *cnt, int n, double *J, double *K)
{
    for (int a = 0; a < ns; a++) {
        int fa = first[a], na = cnt[a];
        for (int i = 0; i < na; i++)
            for (int j = 0; j < na; j++) J[(size_t)(fa + i) * n + fa + j] *= 2.0;
    }
    double *tmp = malloc(sizeof(double) * 100);
    for (int a = 0; a < ns; a++)
        for (int b = 0; b <= a; b++) {
            int fa = first[a], na = cnt[a], fb = first[b], nb = cnt[b];
            for (int i = 0; i < na; i++)
                for (int j = 0; j < nb; j++) tmp[i * nb + j] = K[(size_t)(fb + j) * n + fa + i];
            for (int i = 0; i < na; i++)
                for (int j = 0; j < nb; j++) K[(size_t)(fa + i) * n + fb + j] += tmp[i * nb + j];
        }
    free(tmp);
    for (long i = 0; i < (long)n * n; i++) J[i] *= 2.0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < i; j++) {
            J[(size_t)j * n + i] = J[(size_t)i * n + j];
            K[(size_t)j * n + i] = K[(size_t)i * n + j];
        }
}

static void post_jk(const osys_t *s, double *J, double *K)
{
    int *cnt = malloc(sizeof(int) * s->n_shell);
    for (int a = 0; a < s->n_shell; a++) cnt[a] = ncomp(s->l[a]);
    post_jk_blocks(s->n_shell, s->first_bf, cnt, s->n_bf, J, K);
    free(cnt);
}

/* ---------------- one-electron integrals (molecule.ints.foo:34-415) ------------ */
/* Overlap and kinetic by Obara-Saika 1-D recursions; nuclear attraction by Rys
 * quadrature with X = p|PC|^2 (the eta -> infinity limit of the ERI recursion). */
static void os1d(double PA, double PB, double p, int la, int lb, double S[][12])
{
    double h = 0.5 / p;
    S[0][0] = 1.0;
    for (int i = 0; i <= la; i++)
        for (int j = 0; j <= lb; j++) {
            if (i == 0 && j == 0) continue;
            if (i > 0)
                S[i][j] = PA * S[i - 1][j] + h * ((i > 1 ? (i - 1) * S[i - 2][j] : 0.0) +
                                                  (j > 0 ? j * S[i - 1][j - 1] : 0.0));
            else
                S[i][j] = PB * S[i][j - 1] + h * ((j > 1 ? (j - 1) * S[i][j - 2] : 0.0));
        }
}

typedef struct { const osys_t *s; double *S, *T, *V; } oe_job;

static void oe_item(void *arg, long item, int worker)
{
    (void)worker;
    oe_job *jb = (oe_job *)arg;
    const osys_t *s = jb->s;
    double *S = jb->S, *T = jb->T, *V = jb->V;
    int n = s->n_bf, a = (int)item;
    {
        for (int b = 0; b <= a; b++) {
            int la = s->l[a], lb = s->l[b], na = ncomp(la), nb = ncomp(lb);
            const double *A = s->atom_xyz + 3 * s->atom[a], *B = s->atom_xyz + 3 * s->atom[b];
            double sb[10][10], tb[10][10], vb[10][10];
            memset(sb, 0, sizeof sb); memset(tb, 0, sizeof tb); memset(vb, 0, sizeof vb);
            double AB2 = 0;
            for (int t = 0; t < 3; t++) AB2 += (A[t] - B[t]) * (A[t] - B[t]);
            for (int ia = 0; ia < s->nprim[a]; ia++)
                for (int ib = 0; ib < s->nprim[b]; ib++) {
                    double al = s->exps[s->prim_off[a] + ia], be = s->exps[s->prim_off[b] + ib];
                    double cc = s->coefs[s->prim_off[a] + ia] * s->coefs[s->prim_off[b] + ib];
                    double p = al + be, mu = al * be / p;
                    double Pc[3], Sx[3][12][12];
                    for (int t = 0; t < 3; t++) {
                        Pc[t] = (al * A[t] + be * B[t]) / p;
                        os1d(Pc[t] - A[t], Pc[t] - B[t], p, la + 1, lb + 1, Sx[t]);
                    }
                    double pre = cc * exp(-mu * AB2) * pow(M_PI / p, 1.5);
                    for (int i = 0; i < na; i++)
                        for (int j = 0; j < nb; j++) {
                            const int *xa = g_xyz[ncomp_up_to(la - 1) + i], *xb = g_xyz[ncomp_up_to(lb - 1) + j];
                            double s3[3], t3[3];
                            for (int t = 0; t < 3; t++) {
                                int u = xa[t], v = xb[t];
                                s3[t] = Sx[t][u][v];
                                t3[t] = 2.0 * al * be * Sx[t][u + 1][v + 1] +
                                        (u > 0 && v > 0 ? 0.5 * u * v * Sx[t][u - 1][v - 1] : 0.0) -
                                        (v > 0 ? al * v * Sx[t][u + 1][v - 1] : 0.0) -
                                        (u > 0 ? be * u * Sx[t][u - 1][v + 1] : 0.0);
                            }
                            sb[i][j] += pre * s3[0] * s3[1] * s3[2];
                            tb[i][j] += pre * (t3[0] * s3[1] * s3[2] + s3[0] * t3[1] * s3[2] +
                                               s3[0] * s3[1] * t3[2]);
                        }
                    /* nuclear attraction */
                    int L = la + lb, nr = L / 2 + 1;
                    for (int C = 0; C < s->n_atom; C++) {
                        const double *Cc = s->atom_xyz + 3 * C;
                        double PC2 = 0;
                        for (int t = 0; t < 3; t++) PC2 += (Pc[t] - Cc[t]) * (Pc[t] - Cc[t]);
                        double u[16], w[16];
                        rys(nr, p * PC2, u, w, 0);
                        double vpre = -s->atom_Z[C] * 2.0 * M_PI / p * exp(-mu * AB2) * cc;
                        for (int r = 0; r < nr; r++) {
                            double I1[3][8][8];
                            for (int t = 0; t < 3; t++) {
                                double c00 = (Pc[t] - A[t]) + u[r] * (Cc[t] - Pc[t]);
                                double b10 = (1.0 - u[r]) * 0.5 / p;
                                double e1d[10];
                                e1d[0] = 1.0;
                                for (int e = 0; e < L; e++)
                                    e1d[e + 1] = c00 * e1d[e] + (e ? e * b10 * e1d[e - 1] : 0.0);
                                /* HRR (i, j+1) = (i+1, j) + (A-B)(i, j) */
                                double H[10][10];
                                for (int e = 0; e <= L; e++) H[e][0] = e1d[e];
                                for (int j = 1; j <= lb; j++)
                                    for (int i = 0; i <= L - j; i++)
                                        H[i][j] = H[i + 1][j - 1] + (A[t] - B[t]) * H[i][j - 1];
                                for (int i = 0; i <= la; i++)
                                    for (int j = 0; j <= lb; j++) I1[t][i][j] = H[i][j];
                            }
                            for (int i = 0; i < na; i++)
                                for (int j = 0; j < nb; j++) {
                                    const int *xa = g_xyz[ncomp_up_to(la - 1) + i], *xb = g_xyz[ncomp_up_to(lb - 1) + j];
                                    vb[i][j] += vpre * w[r] * I1[0][xa[0]][xb[0]] * I1[1][xa[1]][xb[1]] *
                                                I1[2][xa[2]][xb[2]];
                                }
                        }
                    }
                }
            int fa = s->first_bf[a], fb = s->first_bf[b];
            for (int i = 0; i < na; i++)
                for (int j = 0; j < nb; j++) {
                    size_t ij = (size_t)(fa + i) * n + fb + j, ji = (size_t)(fb + j) * n + fa + i;
                    S[ij] = S[ji] = sb[i][j];
                    T[ij] = T[ji] = tb[i][j];
                    V[ij] = V[ji] = vb[i][j];
                }
        }
    }
}

void orc_one_electron(const osys_t *s, double *S, double *T, double *V)
{
    comp_tables();
    int n = s->n_bf;
    memset(S, 0, sizeof(double) * n * n);
    memset(T, 0, sizeof(double) * n * n);
    memset(V, 0, sizeof(double) * n * n);
    oe_job jb = {s, S, T, V};
    parallel_for(s->n_shell, orc_max_threads(), oe_item, &jb);
}

void orc_make_r_JK(const osys_t *s, const ocut_t *cut, const double *P, const double *max_I_in,
                   double *J, double *K, long *stats)
{
    make_r_JK_impl(s, cut, P, max_I_in, J, K, stats, 1, NULL, NULL, NULL);
}

/* UHF J/K (make_u_JK_engine, molecule.fock.foo:2558-2712): Ja, Jb, Ka, Kb as Tonto
 * returns them (see jk_item for the J split). */
void orc_make_u_JK(const osys_t *s, const ocut_t *cut, const double *Pa, const double *Pb,
                   const double *max_I_in, double *Ja, double *Jb, double *Ka, double *Kb, long *stats)
{
    make_r_JK_impl(s, cut, Pa, max_I_in, Ja, Ka, stats, 1, Pb, Jb, Kb);
}

/* Timing sample for the CPU baseline: the same build restricted to every stride-th
 * (ab| shell pair (J and K are then partial sums). */
void orc_make_r_JK_sample(const osys_t *s, const ocut_t *cut, const double *P, const double *max_I_in,
                          double *J, double *K, long *stats, long stride)
{
    make_r_JK_impl(s, cut, P, max_I_in, J, K, stats, stride < 1 ? 1 : stride, NULL, NULL, NULL);
}

/* ---------------- screening audit (tests only) ------------------------------------
 * For every unique quartet (ab >= cd, q = ab(ab+1)/2 + cd, Tonto's pair order) code[q]
 * is what the build above does with it (0: dropped; bit 0 J digested, bit 1 K digested,
 * bit 2 skipped primitive-pair lists used) and maxabs[q] = max |(ab|cd)| over its
 * components with full primitive lists, for every quartet, dropped or not: a test can
 * then bound what each dropped quartet would have contributed from the integrals and the
 * density themselves, without max_I or P_max. */
typedef struct {
    jk_job *jb;
    unsigned char *code;
    double *maxabs;
    double **Ibuf;
    qws_t **ws;
} audit_job;

static void audit_item(void *arg, long ab, int w)
{
    audit_job *aj = (audit_job *)arg;
    jk_job *jb = aj->jb;
    const osys_t *s = jb->s;
    int a, b;
    decode(ab, &a, &b);
    const int row = ab_survives(jb, ab);
    for (long cd = 0; cd <= ab; cd++) {
        int c, d;
        decode(cd, &c, &d);
        const long q = ab * (ab + 1) / 2 + cd;
        int code = row ? screen_quartet(jb, ab, cd, a, b, c, d) : 0;
        if (code && uses_skip_lists(s, a, b, c, d)) code |= 4;
        aj->code[q] = (unsigned char)code;
        double *I = aj->Ibuf[w];
        quartet(s, a, b, c, d, 0, jb->cut, I, aj->ws[w]);
        const int nI = ncomp(s->l[a]) * ncomp(s->l[b]) * ncomp(s->l[c]) * ncomp(s->l[d]);
        double m = 0.0;
        for (int i = 0; i < nI; i++)
            if (fabs(I[i]) > m) m = fabs(I[i]);
        aj->maxabs[q] = m;
    }
}

long orc_quartet_audit(const osys_t *s, const ocut_t *cut, const double *P, const double *max_I,
                       unsigned char *code, double *maxabs)
{
    comp_tables();
    const long npair = (long)s->n_shell * (s->n_shell + 1) / 2;
    double *Pmax = malloc(sizeof(double) * npair);
    max_density(s, P, Pmax);
    double P_max = 0.0, I_max = 0.0;
    for (long i = 0; i < npair; i++) {
        if (Pmax[i] > P_max) P_max = Pmax[i];
        if (max_I[i] > I_max) I_max = max_I[i];
    }
    jk_job jb;
    memset(&jb, 0, sizeof jb);
    jb.s = s; jb.cut = cut; jb.P = P; jb.max_I = max_I; jb.Pmax = Pmax;
    jb.P_max = P_max; jb.IP_max = I_max * P_max;
    const int nthr = orc_max_threads();
    audit_job aj = {&jb, code, maxabs, malloc(sizeof(double *) * nthr), malloc(sizeof(qws_t *) * nthr)};
    for (int t = 0; t < nthr; t++) {
        aj.Ibuf[t] = malloc(sizeof(double) * 10 * 10 * 10 * 10);
        aj.ws[t] = malloc(sizeof(qws_t));
    }
    parallel_for(npair, nthr, audit_item, &aj);
    for (int t = 0; t < nthr; t++) { free(aj.Ibuf[t]); free(aj.ws[t]); }
    free(aj.Ibuf); free(aj.ws); free(Pmax);
    return npair * (npair + 1) / 2;
}

/* Number of primitive pairs of shell pair (a, b) with or without the 1e-6 skip
 * (SHELL1QUARTET:set_ab_new, shell1quartet.foo:229-273). */
int orc_pair_prims(const osys_t *s, int a, int b, int skip, double pair_cut)
{
    pairdat_t pd;
    make_pair(s, a, b, skip, pair_cut, &pd);
    return pd.n;
}

/* ---------------- spherical basis: MOLECULE.FOCK:make_r_JK_direct ---------------------
 * (molecule.fock.foo:1235-1363) with Tonto's spherical integrals: the Cartesian block of
 * each quartet turned into real solid harmonics index by index for shells with l >= 2
 * (SHELL1QUARTET:make_ERI / change_to_spherical, shell1quartet.foo:433-458, 505-563),
 * screening by Schwarz_inequality_test (molecule.fock.foo:3454-3495: the largest of the
 * six density blocks ab, cd, ac, ad, bc, bd; skip below TOL(20) or when
 * Pmax*I_ab*I_cd > cutoff fails), one decision for J and K together, max_I and P_max over
 * spherical blocks (3601-3633, 3705-3752).  The solid harmonics are written out in closed
 * form (real solid harmonics, Molden order M = 0, +1, -1, +2, -2, ..., normalised like
 * x^l), independently of the product's recurrence. */
#define SQ3 1.7320508075688772935
static double g_T2[6][5], g_T3[10][7];
static int g_sph_ready = 0;

static void sph_tables(void)
{
    if (g_sph_ready) return;
    const double a = sqrt(3.0 / 8.0), b = sqrt(5.0 / 8.0), c = sqrt(15.0);
    memset(g_T2, 0, sizeof g_T2);
    memset(g_T3, 0, sizeof g_T3);
    /* d: xx yy zz xy xz yz */
    g_T2[0][0] = -0.5; g_T2[1][0] = -0.5; g_T2[2][0] = 1.0;  /* z^2 - (x^2+y^2)/2 */
    g_T2[4][1] = SQ3;                                        /* sqrt3 xz */
    g_T2[5][2] = SQ3;                                        /* sqrt3 yz */
    g_T2[0][3] = SQ3 / 2; g_T2[1][3] = -SQ3 / 2;             /* sqrt3/2 (x^2 - y^2) */
    g_T2[3][4] = SQ3;                                        /* sqrt3 xy */
    /* f: xxx yyy zzz xxy xxz xyy yyz xzz yzz xyz */
    g_T3[2][0] = 1.0; g_T3[4][0] = -1.5; g_T3[6][0] = -1.5;  /* z (2z^2 - 3x^2 - 3y^2)/2 */
    g_T3[0][1] = -a; g_T3[5][1] = -a; g_T3[7][1] = 4 * a;    /* sqrt(3/8) x (4z^2 - x^2 - y^2) */
    g_T3[1][2] = -a; g_T3[3][2] = -a; g_T3[8][2] = 4 * a;    /* sqrt(3/8) y (4z^2 - x^2 - y^2) */
    g_T3[4][3] = c / 2; g_T3[6][3] = -c / 2;                 /* sqrt15/2 z (x^2 - y^2) */
    g_T3[9][4] = c;                                          /* sqrt15 xyz */
    g_T3[0][5] = b; g_T3[5][5] = -3 * b;                     /* sqrt(5/8) x (x^2 - 3y^2) */
    g_T3[3][6] = 3 * b; g_T3[1][6] = -b;                     /* sqrt(5/8) y (3x^2 - y^2) */
    g_sph_ready = 1;
}

static int nsph(int l) { return l >= 2 ? 2 * l + 1 : ncomp(l); }

void orc_solid_harmonics(int l, double *out)  /* ncomp(l) x (2l+1), row-major */
{
    sph_tables();
    for (int i = 0; i < ncomp(l); i++)
        for (int m = 0; m < 2 * l + 1; m++)
            out[i * (2 * l + 1) + m] = l == 2 ? g_T2[i][m] : l == 3 ? g_T3[i][m] : (l < 2 && i == m ? 1.0 : 0.0);
}

static double tval(int l, int i, int m) { return l == 2 ? g_T2[i][m] : l == 3 ? g_T3[i][m] : (i == m); }

/* in: [na][nb][nc][nd] Cartesian; out: the same with every l >= 2 index spherical */
static void cart_to_sph4(const int *l4, const double *in, double *out, double *t0, double *t1)
{
    int dims[4], dimo[4];
    for (int k = 0; k < 4; k++) { dims[k] = ncomp(l4[k]); dimo[k] = nsph(l4[k]); }
    const double *cur = in;
    int cd[4] = {dims[0], dims[1], dims[2], dims[3]};
    for (int k = 0; k < 4; k++) {
        double *nxt = k == 3 ? out : (k % 2 ? t1 : t0);
        if (l4[k] < 2) {
            long n = (long)cd[0] * cd[1] * cd[2] * cd[3];
            memcpy(nxt, cur, sizeof(double) * n);
            cur = nxt;
            continue;
        }
        int pre = 1, post = 1;
        for (int q = 0; q < k; q++) pre *= cd[q];
        for (int q = k + 1; q < 4; q++) post *= cd[q];
        const int ni = cd[k], no = dimo[k];
        for (int p = 0; p < pre; p++)
            for (int m = 0; m < no; m++)
                for (int r = 0; r < post; r++) {
                    double acc = 0.0;
                    for (int i = 0; i < ni; i++) acc += tval(l4[k], i, m) * cur[((long)p * ni + i) * post + r];
                    nxt[((long)p * no + m) * post + r] = acc;
                }
        cd[k] = no;
        cur = nxt;
    }
}

static long sph_quartet(const osys_t *s, int a, int b, int c, int d, int skip, const ocut_t *cut, double *out,
                        qws_t *ws, double *cbuf)
{
    long np = quartet(s, a, b, c, d, skip, cut, cbuf, ws);
    const int l4[4] = {s->l[a], s->l[b], s->l[c], s->l[d]};
    cart_to_sph4(l4, cbuf, out, ws->t0, ws->t1);
    return np;
}

typedef struct {
    const osys_t *s;
    const ocut_t *cut;
    const int *sf;            /* first spherical function of each shell */
    int nsf;                  /* spherical functions */
    const double *P, *max_I, *Pmax;
    double P_max, IP_max;
    double **Jt, **Kt, **Ibuf, **Cbuf;
    qws_t **ws;
    long (*st)[5];
} direct_job;

static void schwarz_sph_item(void *arg, long item, int worker)
{
    (void)worker;
    void **A = (void **)arg;
    const osys_t *s = (const osys_t *)A[0];
    const ocut_t *cut = (const ocut_t *)A[1];
    double *max_I = (double *)A[2];
    int a = (int)item;
    double *buf = malloc(sizeof(double) * 10000), *cb = malloc(sizeof(double) * 10000);
    qws_t *ws = malloc(sizeof(qws_t));
    for (int b = 0; b <= a; b++) {
        sph_quartet(s, a, b, a, b, 0, cut, buf, ws, cb);
        const int na = nsph(s->l[a]), nb = nsph(s->l[b]);
        double m = 0.0;
        for (int ia = 0; ia < na; ia++)
            for (int ib = 0; ib < nb; ib++) {
                double v = fabs(buf[((ia * nb + ib) * na + ia) * nb + ib]);
                if (v > m) m = v;
            }
        max_I[(long)a * (a + 1) / 2 + b] = sqrt(m);
    }
    free(ws); free(buf); free(cb);
}

/* max_I over spherical (ab|ab) blocks (make_max_abab_integrals with a spherical basis) */
void orc_schwarz_sph(const osys_t *s, double *max_I)
{
    comp_tables();
    sph_tables();
    ocut_t cut = {0, 0, 0, 1e-6, 0};
    void *A[3] = {(void *)s, &cut, max_I};
    parallel_for(s->n_shell, orc_max_threads(), schwarz_sph_item, A);
}

static void direct_item(void *arg, long ab, int w)
{
    direct_job *jb = (direct_job *)arg;
    const osys_t *s = jb->s;
    const ocut_t *cut = jb->cut;
    const double *max_I = jb->max_I, *Pmax = jb->Pmax, *P = jb->P;
    const int n = jb->nsf;
    double *Jl = jb->Jt[w], *Kl = jb->Kt[w], *I = jb->Ibuf[w];
    long *st = jb->st[w];
    if (max_I[ab] * jb->IP_max < cut->schwarz) return;
    int a, b;
    decode(ab, &a, &b);
    for (long cd = 0; cd <= ab; cd++) {
        if (max_I[ab] * max_I[cd] * jb->P_max < cut->schwarz) continue;
        int c, d;
        decode(cd, &c, &d);
        /* Schwarz_inequality_test (molecule.fock.foo:3454-3495) */
        double Pm = Pmax[ab];
        const long blk[5] = {cd, pidx(a, c), pidx(a, d), pidx(b, c), pidx(b, d)};
        for (int k = 0; k < 5; k++)
            if (Pmax[blk[k]] > Pm) Pm = Pmax[blk[k]];
        if (Pm < 1e-20) continue;
        if (!(Pm * max_I[ab] * max_I[cd] > cut->schwarz)) continue;
        double f = (a == b) ? 0.5 : 1.0;
        if (c == d) f *= 0.5;
        if (a == c && b == d) f *= 0.5;
        const int any_same = !uses_skip_lists(s, a, b, c, d);
        long npq = sph_quartet(s, a, b, c, d, !any_same, cut, I, jb->ws[w], jb->Cbuf[w]);
        st[0]++; st[1] += npq; st[4] += !any_same;
        const int na = nsph(s->l[a]), nb = nsph(s->l[b]), nc = nsph(s->l[c]), nd = nsph(s->l[d]);
        const int fa = jb->sf[a], fb = jb->sf[b], fc = jb->sf[c], fd = jb->sf[d];
        for (int ia = 0; ia < na; ia++)
            for (int ib = 0; ib < nb; ib++)
                for (int ic = 0; ic < nc; ic++)
                    for (int id = 0; id < nd; id++) {
                        const double v = f * I[((ia * nb + ib) * nc + ic) * nd + id];
                        const int i = fa + ia, j = fb + ib, k = fc + ic, l = fd + id;
                        Jl[(size_t)i * n + j] += v * P[(size_t)l * n + k];
                        Jl[(size_t)k * n + l] += v * P[(size_t)j * n + i];
                        Kl[(size_t)i * n + k] += v * P[(size_t)l * n + j];
                        Kl[(size_t)i * n + l] += v * P[(size_t)k * n + j];
                        Kl[(size_t)j * n + k] += v * P[(size_t)l * n + i];
                        Kl[(size_t)j * n + l] += v * P[(size_t)k * n + i];
                    }
    }
}

/* J and K in the spherical basis (n_sph x n_sph, row-major) from the spherical density */
void orc_make_r_JK_direct(const osys_t *s, const ocut_t *cut, const double *P, const double *max_I_in,
                          double *J, double *K, long *stats)
{
    comp_tables();
    sph_tables();
    const int ns = s->n_shell;
    int *sf = malloc(sizeof(int) * (ns + 1));
    sf[0] = 0;
    for (int a = 0; a < ns; a++) sf[a + 1] = sf[a] + nsph(s->l[a]);
    const int n = sf[ns];
    const long npair = (long)ns * (ns + 1) / 2;
    double *max_I = malloc(sizeof(double) * npair);
    if (max_I_in) memcpy(max_I, max_I_in, sizeof(double) * npair);
    else orc_schwarz_sph(s, max_I);
    double *Pmax = malloc(sizeof(double) * npair);
    for (int a = 0; a < ns; a++)
        for (int b = 0; b <= a; b++) {
            double m = 0.0;
            for (int i = 0; i < nsph(s->l[a]); i++)
                for (int j = 0; j < nsph(s->l[b]); j++) {
                    const double v = fabs(P[(size_t)(sf[a] + i) * n + sf[b] + j]);
                    if (v > m) m = v;
                }
            Pmax[(long)a * (a + 1) / 2 + b] = m;
        }
    double P_max = 0.0, I_max = 0.0;
    for (long i = 0; i < npair; i++) {
        if (Pmax[i] > P_max) P_max = Pmax[i];
        if (max_I[i] > I_max) I_max = max_I[i];
    }
    const int nthr = orc_max_threads();
    direct_job jb;
    jb.s = s; jb.cut = cut; jb.sf = sf; jb.nsf = n; jb.P = P; jb.max_I = max_I; jb.Pmax = Pmax;
    jb.P_max = P_max; jb.IP_max = I_max * P_max;
    jb.Jt = malloc(sizeof(double *) * nthr); jb.Kt = malloc(sizeof(double *) * nthr);
    jb.Ibuf = malloc(sizeof(double *) * nthr); jb.Cbuf = malloc(sizeof(double *) * nthr);
    jb.ws = malloc(sizeof(qws_t *) * nthr);
    jb.st = calloc(nthr, sizeof(long[5]));
    for (int t = 0; t < nthr; t++) {
        jb.Jt[t] = calloc((size_t)n * n, sizeof(double));
        jb.Kt[t] = calloc((size_t)n * n, sizeof(double));
        jb.Ibuf[t] = malloc(sizeof(double) * 10000);
        jb.Cbuf[t] = malloc(sizeof(double) * 10000);
        jb.ws[t] = malloc(sizeof(qws_t));
    }
    parallel_for(npair, nthr, direct_item, &jb);
    memset(J, 0, sizeof(double) * n * n);
    memset(K, 0, sizeof(double) * n * n);
    long st[5] = {0, 0, 0, 0, 0};
    for (int t = 0; t < nthr; t++) {
        for (long i = 0; i < (long)n * n; i++) { J[i] += jb.Jt[t][i]; K[i] += jb.Kt[t][i]; }
        for (int k = 0; k < 5; k++) st[k] += jb.st[t][k];
        free(jb.Jt[t]); free(jb.Kt[t]); free(jb.Ibuf[t]); free(jb.Cbuf[t]); free(jb.ws[t]);
    }
    free(jb.Jt); free(jb.Kt); free(jb.Ibuf); free(jb.Cbuf); free(jb.ws); free(jb.st);
    /* post-processing over spherical shell blocks (molecule.fock.foo:1350-1361) */
    int *cnt = malloc(sizeof(int) * ns);
    for (int a = 0; a < ns; a++) cnt[a] = nsph(s->l[a]);
    post_jk_blocks(ns, sf, cnt, n, J, K);
    free(cnt);
    if (stats) memcpy(stats, st, sizeof st);
    free(max_I); free(Pmax); free(sf);
}

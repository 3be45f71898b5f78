/* ORACLE (test infrastructure only; never linked into the product library).
 *
 * Quad-precision (__float128) Boys functions and Rys roots/weights, used to pin the
 * FP64 Rys tables and the FP64 oracle.  Semantics follow `foofiles/rys.foo:89-134`:
 * roots are returned as u = t^2, weights satisfy sum_i w_i u_i^k = F_k(X) for
 * k = 0..2n-1.  Method: Boys moments by series + downward recursion, Hankel Cholesky
 * -> Jacobi matrix (Golub-Welsch), cyclic Jacobi eigen-solve.  This is a different
 * construction from Tonto's fits (`rys.foo:226-3427`) and from its n>=6 moment /
 * regula-falsi path (`rys.foo:3429-3769`).
 */
#include <quadmath.h>
#include <string.h>

typedef __float128 q_t;

/* F_0..F_mmax(X) */
static void boys_q(int mmax, q_t X, q_t *F)
{
    if (X == 0) {
        for (int m = 0; m <= mmax; m++) F[m] = (q_t)1 / (2 * m + 1);
        return;
    }
    q_t ex = expq(-X);
    /* series for F_mmax: e^{-X} sum_k (2X)^k / ((2m+1)(2m+3)...(2m+2k+1)) */
    q_t term = (q_t)1 / (2 * mmax + 1), sum = term;
    for (int k = 1; k < 100000; k++) {
        term *= 2 * X / (2 * mmax + 2 * k + 1);
        sum += term;
        if (term < sum * 1e-36Q) break;
    }
    F[mmax] = ex * sum;
    for (int m = mmax; m > 0; m--) F[m - 1] = (2 * X * F[m] + ex) / (2 * m - 1);
}

double orq_boys(int m, double X)
{
    q_t F[64];
    boys_q(m, (q_t)X, F);
    return (double)F[m];
}

/* Rys roots u (ascending) and weights w for n roots at X; also returned in quad
 * precision as long double-free doubles.  Returns 0 on success. */
int orq_rys(int n, double Xd, double *u, double *w)
{
    if (n < 1 || n > 12) return 1;
    q_t X = (q_t)Xd;
    q_t mu[32];
    boys_q(2 * n, X, mu);
    int m = n + 1;
    q_t R[16][16];
    memset(R, 0, sizeof(R));
    for (int i = 0; i < m; i++) {
        q_t s = mu[2 * i];
        for (int k = 0; k < i; k++) s -= R[k][i] * R[k][i];
        if (!(s > 0)) return 2;
        R[i][i] = sqrtq(s);
        for (int j = i + 1; j < m; j++) {
            q_t t = mu[i + j];
            for (int k = 0; k < i; k++) t -= R[k][i] * R[k][j];
            R[i][j] = t / R[i][i];
        }
    }
    q_t A[16][16], V[16][16];
    memset(A, 0, sizeof(A));
    memset(V, 0, sizeof(V));
    for (int j = 0; j < n; j++) {
        q_t a = R[j][j + 1] / R[j][j];
        if (j > 0) a -= R[j - 1][j] / R[j - 1][j - 1];
        A[j][j] = a;
        if (j < n - 1) A[j][j + 1] = A[j + 1][j] = R[j + 1][j + 1] / R[j][j];
        V[j][j] = 1;
    }
    /* cyclic Jacobi */
    for (int sweep = 0; sweep < 100; sweep++) {
        q_t off = 0, diag = 0;
        for (int p = 0; p < n; p++) {
            diag += A[p][p] * A[p][p];
            for (int q = p + 1; q < n; q++) off += A[p][q] * A[p][q];
        }
        if (off <= diag * 1e-66Q) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                if (A[p][q] == 0) continue;
                q_t theta = (A[q][q] - A[p][p]) / (2 * A[p][q]);
                q_t t = (theta >= 0 ? 1 : -1) / (fabsq(theta) + sqrtq(theta * theta + 1));
                q_t c = 1 / sqrtq(t * t + 1), s = t * c;
                for (int k = 0; k < n; k++) {
                    q_t akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    q_t apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    q_t vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    /* sort ascending */
    int idx[16];
    for (int i = 0; i < n; i++) idx[i] = i;
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++)
            if (A[idx[j]][idx[j]] < A[idx[i]][idx[i]]) { int t = idx[i]; idx[i] = idx[j]; idx[j] = t; }
    for (int i = 0; i < n; i++) {
        u[i] = (double)A[idx[i]][idx[i]];
        w[i] = (double)(mu[0] * V[0][idx[i]] * V[0][idx[i]]);
    }
    return 0;
}

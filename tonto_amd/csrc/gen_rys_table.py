"""Generate the FP64 Rys root/weight tables used by the HIP kernels.

Definition (the convention of `foofiles/rys.foo:89-134`): for n roots and argument
X = rho |PQ|^2 the Rys quadrature on u = t^2 in [0,1) satisfies

    sum_i w_i u_i^k = F_k(X)   (k = 0 .. 2n-1),   F_k(X) = int_0^1 t^{2k} e^{-X t^2} dt,

and the routine returns the roots as u = t^2 (Tonto's "t2" roots).

Method (ours, not Tonto's fits): for each n and each interval [k*h, (k+1)*h) of X below
X_asym(n) the roots and weights are computed in 60-digit arithmetic (mpmath) by
Golub-Welsch from the Boys moments (Hankel Cholesky -> Jacobi matrix -> symmetric
eigen-solve) at Chebyshev nodes, and stored as Chebyshev coefficients.  For
X >= X_asym(n) the kernels use the large-X limit u_i = s_i^2 / X, w_i = W_i / sqrt(X),
with s_i, W_i the positive nodes and weights of the 2n-point Gauss-Hermite rule; the
generator checks that this limit is FP64-exact at X_asym(n).

Output: `rys_table.inc` (C source data) next to this file.
"""
import os
import sys

import mpmath as mp

NMAX = 7            # (ff|ff) needs floor(12/2)+1 = 7 roots
H = 2.0             # interval width in X
NCHEB = 14          # Chebyshev coefficients per interval (degree 13)
XASYM = {1: 40.0, 2: 48.0, 3: 54.0, 4: 62.0, 5: 68.0, 6: 74.0, 7: 78.0}
mp.mp.dps = 60


def boys(kmax, x):
    x = mp.mpf(x)
    if x == 0:
        return [mp.mpf(1) / (2 * k + 1) for k in range(kmax + 1)]
    return [mp.hyp1f1(k + mp.mpf(0.5), k + mp.mpf(1.5), -x) / (2 * k + 1)
            for k in range(kmax + 1)]


def rys_exact(n, x):
    """Roots u_i (ascending) and weights w_i in high precision by Golub-Welsch."""
    mu = boys(2 * n, x)                      # moments mu_0 .. mu_2n
    # Hankel matrix H_ij = mu_{i+j}, i,j = 0..n ; Cholesky H = R^T R
    m = n + 1
    Hm = mp.matrix(m, m)
    for i in range(m):
        for j in range(m):
            Hm[i, j] = mu[i + j]
    R = mp.matrix(m, m)
    for i in range(m):
        s = Hm[i, i] - mp.fsum(R[k, i] ** 2 for k in range(i))
        R[i, i] = mp.sqrt(s)
        for j in range(i + 1, m):
            R[i, j] = (Hm[i, j] - mp.fsum(R[k, i] * R[k, j] for k in range(i))) / R[i, i]
    alpha = []
    beta = []
    for j in range(n):
        a = R[j, j + 1] / R[j, j]
        if j > 0:
            a -= R[j - 1, j] / R[j - 1, j - 1]
        alpha.append(a)
        if j < n - 1:
            beta.append(R[j + 1, j + 1] / R[j, j])
    J = mp.matrix(n, n)
    for j in range(n):
        J[j, j] = alpha[j]
        if j < n - 1:
            J[j, j + 1] = beta[j]
            J[j + 1, j] = beta[j]
    E, Q = mp.eigsy(J)
    pairs = sorted((E[i], mu[0] * Q[0, i] ** 2) for i in range(n))
    return [p[0] for p in pairs], [p[1] for p in pairs]


def hermite_half(n):
    """Positive nodes s_i (ascending) and weights of the 2n-point Gauss-Hermite rule."""
    # Golub-Welsch for Hermite weight e^{-t^2}: alpha=0, beta_k = sqrt(k/2)
    N = 2 * n
    J = mp.matrix(N, N)
    for k in range(1, N):
        J[k - 1, k] = J[k, k - 1] = mp.sqrt(mp.mpf(k) / 2)
    E, Q = mp.eigsy(J)
    mu0 = mp.sqrt(mp.pi)
    pts = sorted((E[i], mu0 * Q[0, i] ** 2) for i in range(N) if E[i] > 0)
    return [p[0] for p in pts], [p[1] for p in pts]


def cheb_fit(values):
    N = len(values)
    c = []
    for m in range(N):
        s = mp.fsum(values[j] * mp.cos(mp.pi * m * (j + mp.mpf(0.5)) / N) for j in range(N))
        c.append(2 * s / N)
    c[0] /= 2
    return c


def cheb_eval(c, t):
    b1 = b2 = mp.mpf(0)
    for ck in reversed(c[1:]):
        b1, b2 = 2 * t * b1 - b2 + ck, b1
    return t * b1 - b2 + c[0]


def build(check=True):
    out = {}
    herm = {}
    for n in range(1, NMAX + 1):
        nint = int(round(XASYM[n] / H))
        coefs = []                            # [interval][coef][func]
        for k in range(nint):
            x0 = k * H
            nodes = [mp.cos(mp.pi * (j + mp.mpf(0.5)) / NCHEB) for j in range(NCHEB)]
            vals = [rys_exact(n, x0 + H * (1 + t) / 2) for t in nodes]
            funcs = []
            for i in range(n):
                funcs.append(cheb_fit([v[0][i] for v in vals]))
            for i in range(n):
                funcs.append(cheb_fit([v[1][i] for v in vals]))
            coefs.append([[funcs[f][m] for f in range(2 * n)] for m in range(NCHEB)])
        out[n] = coefs
        s, W = hermite_half(n)
        herm[n] = (s, W)
        if check:
            worst = mp.mpf(0)
            for k in range(nint):
                for frac in (mp.mpf("0.137"), mp.mpf("0.5123"), mp.mpf("0.9"), mp.mpf(0)):
                    x = k * H + frac * H
                    r, w = rys_exact(n, x)
                    t = 2 * frac - 1
                    for f in range(2 * n):
                        c = [coefs[k][m][f] for m in range(NCHEB)]
                        ref = r[f] if f < n else w[f - n]
                        worst = max(worst, abs(cheb_eval(c, t) - ref) / abs(ref))
            r, w = rys_exact(n, XASYM[n])
            asym = max(max(abs(s[i] ** 2 / XASYM[n] - r[i]) / r[i] for i in range(n)),
                       max(abs(W[i] / mp.sqrt(XASYM[n]) - w[i]) / w[i] for i in range(n)))
            print(f"n={n}: intervals={nint} max rel err (fit)={mp.nstr(worst, 3)} "
                  f"(asym at X={XASYM[n]})={mp.nstr(asym, 3)}", file=sys.stderr)
    return out, herm


def write_inc(path, out, herm):
    lines = ["/* Generated by gen_rys_table.py: Chebyshev tables of Rys roots (u = t^2) and",
             " * weights, and the Gauss-Hermite large-X limit.  Do not edit. */",
             "#ifndef RYS_QUAL",
             "#define RYS_QUAL static const",
             "#endif",
             f"#define RYS_NMAX {NMAX}",
             f"#define RYS_NCHEB {NCHEB}",
             f"#define RYS_H {H!r}",
             ]
    offs = []
    off = 0
    allc = []
    for n in range(1, NMAX + 1):
        offs.append(off)
        for interval in out[n]:
            for m in range(NCHEB):
                for v in interval[m]:
                    allc.append(v)
        off += len(out[n]) * NCHEB * 2 * n
    lines.append(f"#define RYS_TABLE_SIZE {off}")
    lines.append("RYS_QUAL int rys_table_offset[RYS_NMAX + 1] = {0, "
                 + ", ".join(str(o) for o in offs) + "};")
    lines.append("RYS_QUAL double rys_xasym[RYS_NMAX + 1] = {0.0, "
                 + ", ".join(repr(float(XASYM[n])) for n in range(1, NMAX + 1)) + "};")
    hs = []
    hw = []
    for n in range(1, NMAX + 1):
        s, W = herm[n]
        hs += [mp.nstr(v * v, 20, strip_zeros=False) for v in s] + ["0.0"] * (NMAX - n)
        hw += [mp.nstr(v, 20, strip_zeros=False) for v in W] + ["0.0"] * (NMAX - n)
    lines.append("/* [n-1][i]: s_i^2 and W_i of the positive 2n-point Gauss-Hermite nodes */")
    lines.append("RYS_QUAL double rys_herm_s2[RYS_NMAX * RYS_NMAX] = {" + ", ".join(hs) + "};")
    lines.append("RYS_QUAL double rys_herm_w[RYS_NMAX * RYS_NMAX] = {" + ", ".join(hw) + "};")
    lines.append("/* layout: for n, interval k, coefficient m: 2n values (n roots then n weights) */")
    lines.append("RYS_QUAL double rys_table_data[RYS_TABLE_SIZE] = {")
    for i in range(0, len(allc), 4):
        lines.append("  " + ", ".join(mp.nstr(v, 20, strip_zeros=False) for v in allc[i:i + 4]) + ",")
    lines.append("};")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    out, herm = build(check="--nocheck" not in sys.argv)
    write_inc(os.path.join(os.path.dirname(os.path.abspath(__file__)), "rys_table.inc"), out, herm)

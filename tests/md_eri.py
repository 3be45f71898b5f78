"""Independent textbook ERIs for tests: McMurchie-Davidson Hermite expansion
(Helgaker, Jorgensen, Olsen, "Molecular Electronic-Structure Theory", ch. 9) with
Boys functions from scipy.  Shares no code or algorithm with the Rys/HRR oracle."""
import math
from functools import lru_cache

import numpy as np
from scipy import special

from tonto_amd.basis import cart_components


def boys(m, T):
    if T < 1e-14:
        return 1.0 / (2 * m + 1) - T / (2 * m + 3)
    a = m + 0.5
    return special.gamma(a) * special.gammainc(a, T) / (2.0 * T ** a)


def hermite_E(i, j, a, b, Xab):
    """E_t^{ij} for t = 0..i+j (1-D Hermite expansion coefficients)."""
    p = a + b
    q = a * b / p
    XPA = -b / p * Xab
    XPB = a / p * Xab

    @lru_cache(None)
    def E(i, j, t):
        if t < 0 or t > i + j:
            return 0.0
        if i == 0 and j == 0:
            return math.exp(-q * Xab * Xab) if t == 0 else 0.0
        if i > 0:
            return E(i - 1, j, t - 1) / (2 * p) + XPA * E(i - 1, j, t) + (t + 1) * E(i - 1, j, t + 1)
        return E(i, j - 1, t - 1) / (2 * p) + XPB * E(i, j - 1, t) + (t + 1) * E(i, j - 1, t + 1)
    return [E(i, j, t) for t in range(i + j + 1)]


def hermite_R(L, alpha, PQ):
    T = alpha * float(np.dot(PQ, PQ))
    F = [boys(n, T) for n in range(L + 1)]

    @lru_cache(None)
    def R(t, u, v, n):
        if t < 0 or u < 0 or v < 0:
            return 0.0
        if t == 0 and u == 0 and v == 0:
            return (-2 * alpha) ** n * F[n]
        if t > 0:
            return (t - 1) * R(t - 2, u, v, n + 1) + PQ[0] * R(t - 1, u, v, n + 1)
        if u > 0:
            return (u - 1) * R(t, u - 2, v, n + 1) + PQ[1] * R(t, u - 1, v, n + 1)
        return (v - 1) * R(t, u, v - 2, n + 1) + PQ[2] * R(t, u, v - 1, n + 1)
    return R


def prim_eri(la, a, A, lb, b, B, lc, c, C, ld, d, D):
    """Primitive Cartesian (ab|cd) for all component combinations, Tonto component order."""
    A, B, C, D = (np.asarray(x, float) for x in (A, B, C, D))
    p, q = a + b, c + d
    P = (a * A + b * B) / p
    Q = (c * C + d * D) / q
    alpha = p * q / (p + q)
    L = la + lb + lc + ld
    R = hermite_R(L, alpha, P - Q)
    pref = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q))
    ca, cb, cc, cd = (cart_components(l) for l in (la, lb, lc, ld))
    out = np.zeros((len(ca), len(cb), len(cc), len(cd)))
    Eab = {}
    Ecd = {}
    for ia, xa in enumerate(ca):
        for ib, xb in enumerate(cb):
            Eab[ia, ib] = [hermite_E(xa[k], xb[k], a, b, A[k] - B[k]) for k in range(3)]
    for ic, xc in enumerate(cc):
        for id_, xd in enumerate(cd):
            Ecd[ic, id_] = [hermite_E(xc[k], xd[k], c, d, C[k] - D[k]) for k in range(3)]
    for (ia, ib), e1 in Eab.items():
        for (ic, id_), e2 in Ecd.items():
            s = 0.0
            for t, et in enumerate(e1[0]):
                for u, eu in enumerate(e1[1]):
                    for v, ev in enumerate(e1[2]):
                        e123 = et * eu * ev
                        if e123 == 0.0:
                            continue
                        for tau, ft in enumerate(e2[0]):
                            for nu, fu in enumerate(e2[1]):
                                for phi, fv in enumerate(e2[2]):
                                    s += e123 * ft * fu * fv * (-1) ** (tau + nu + phi) * \
                                        R(t + tau, u + nu, v + phi, 0)
            out[ia, ib, ic, id_] = pref * s
    return out


def shell_quartet(shells, a, b, c, d):
    sa, sb, sc, sd = (shells[i] for i in (a, b, c, d))
    out = 0.0
    for ea, ka in zip(sa.exps, sa.coefs):
        for eb, kb in zip(sb.exps, sb.coefs):
            for ec, kc in zip(sc.exps, sc.coefs):
                for ed, kd in zip(sd.exps, sd.coefs):
                    out = out + ka * kb * kc * kd * prim_eri(sa.l, ea, sa.center, sb.l, eb, sb.center,
                                                             sc.l, ec, sc.center, sd.l, ed, sd.center)
    return out

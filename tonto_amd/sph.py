"""Cartesian -> real solid harmonic transformation of Tonto's spherical basis.

T[l] is the (n_cart(l) x (2l+1)) matrix with S_{lM} = sum_c T[l][c, M] x^c: the columns
in Molden/Gaussian order M = 0, +1, -1, +2, -2, ... (GAUSSIAN_DATA:into_std_S_order,
foofiles/gaussian_data.foo:1285-), the rows in Tonto's Cartesian order (basis.cart_components),
every Cartesian function carrying the x^l normalisation, so that the S_{lM} come out
normalised.  Built by the recurrences of Helgaker, Jorgensen and Olsen, Molecular
Electronic-Structure Theory, eqs. (6.4.70)-(6.4.73), in the form and with the prefactors
of GAUSSIAN_DATA:make_xyz_to_solid_harmonics (gaussian_data.foo:1143-1282).  Tonto applies
it for l >= 2 only (SHELL1QUARTET:change_to_spherical, shell1quartet.foo:505-563); s and p
shells keep their Cartesian functions.
"""
import math

import numpy as np

from .basis import cart_components


def std_index(L, M):
    """Column of M in the Molden order 0, +1, -1, +2, -2, ... (0-based)."""
    return 0 if M == 0 else (2 * M - 1 if M > 0 else -2 * M)


def solid_harmonic_tables(lmax=4):
    """[T_0, ..., T_lmax] as numpy arrays (n_cart(l), 2l+1)."""
    T = [np.zeros(((l + 1) * (l + 2) // 2, 2 * l + 1)) for l in range(lmax + 1)]
    T[0][0, 0] = 1.0
    for L in range(lmax):
        d0 = 1 if L == 0 else 0
        top = 2 * L + 1
        fac = math.sqrt((2 ** d0) * top / (top + 1))
        dl = 1 - d0
        nxt = cart_components(L + 1)
        idx = {c: i for i, c in enumerate(nxt)}
        for i, p in enumerate(cart_components(L)):
            x = idx[(p[0] + 1, p[1], p[2])]
            y = idx[(p[0], p[1] + 1, p[2])]
            z = idx[(p[0], p[1], p[2] + 1)]
            valp = fac * T[L][i, std_index(L, L)]
            valm = fac * T[L][i, std_index(L, -L)] * dl
            N = std_index(L + 1, L + 1)
            T[L + 1][x, N] += valp
            T[L + 1][y, N] -= valm
            N = std_index(L + 1, -L - 1)
            T[L + 1][x, N] += valm
            T[L + 1][y, N] += valp
            for M in range(-L, L + 1):
                bot = (L + M + 1) * (L - M + 1)
                T[L + 1][z, std_index(L + 1, M)] += top / math.sqrt(bot) * T[L][i, std_index(L, M)]
        if L >= 1:
            for i, p in enumerate(cart_components(L - 1)):
                x = idx[(p[0] + 2, p[1], p[2])]
                y = idx[(p[0], p[1] + 2, p[2])]
                z = idx[(p[0], p[1], p[2] + 2)]
                for M in range(-L + 1, L):
                    bot = (L + M + 1) * (L - M + 1)
                    val = math.sqrt((L + M) * (L - M) / bot) * T[L - 1][i, std_index(L - 1, M)]
                    N = std_index(L + 1, M)
                    T[L + 1][x, N] -= val
                    T[L + 1][y, N] -= val
                    T[L + 1][z, N] -= val
    return T


def n_sph(l):
    """Functions of an l-shell in Tonto's spherical basis (Cartesian for s and p)."""
    return 2 * l + 1 if l >= 2 else (l + 1) * (l + 2) // 2


def shell_transform(l, T=None):
    """The matrix applied to an l-shell: T_l for l >= 2, the identity for s and p."""
    if l < 2:
        return np.eye((l + 1) * (l + 2) // 2)
    return (T or solid_harmonic_tables(max(l, 2)))[l]


def basis_transform(mol):
    """Block-diagonal (n_cart x n_sph) matrix of a molecule's basis: P_cart = T P_sph T^T,
    J_sph = T^T J_cart T."""
    Ts = solid_harmonic_tables(4)
    ls = [s.l for s in mol.shells]
    nc = sum((l + 1) * (l + 2) // 2 for l in ls)
    ns = sum(n_sph(l) for l in ls)
    out = np.zeros((nc, ns))
    i = j = 0
    for l in ls:
        B = shell_transform(l, Ts)
        out[i:i + B.shape[0], j:j + B.shape[1]] = B
        i += B.shape[0]
        j += B.shape[1]
    return out

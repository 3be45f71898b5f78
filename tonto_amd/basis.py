"""Molecules, Cartesian Gaussian shells and Tonto's basis conventions.

Conventions restated from the reference (each must match for the Fock matrix to be
Tonto's):

* Cartesian component order per angular momentum l, `foofiles/gaussian_data.foo:295-479`
  (d = xx,yy,zz,xy,xz,yz; f = xxx,yyy,zzz,xxy,xxz,xyy,yyz,xzz,yzz,xyz).
* Every component of an l-shell carries the x^l primitive normalisation
  (`foofiles/gaussian_data.foo:757-767`); the library coefficients refer to normalised
  primitives and are un-normalised on read (`foofiles/shell.foo:459-507`,
  `foofiles/molecule.main.foo:803-807`).
* Angstrom -> bohr with ANGSTROM_PER_BOHR = 0.5291772108 (`include/macros.in:508`).
* Shells are ordered atom by atom, in library order within an atom; basis functions
  are numbered consecutively shell by shell.
"""
import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

ANGSTROM_PER_BOHR = 0.5291772108
ELEMENT_Z = {"H": 1, "He": 2, "C": 6, "N": 7, "O": 8}
_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "basis_sets.json")
_LIBRARY = None


def _library():
    global _LIBRARY
    if _LIBRARY is None:
        with open(_DATA) as f:
            _LIBRARY = json.load(f)
    return _LIBRARY


def n_comp(l):
    """Number of Cartesian components of an l-shell (`gaussian_data.foo:174`)."""
    return (l + 1) * (l + 2) // 2


def cart_components(l):
    """Cartesian powers (lx, ly, lz) of an l-shell in Tonto's order.

    Groups by the sorted power triple (a >= b >= c, a descending, then b descending) and
    lists the distinct permutations of each group in the fixed pattern of
    `gaussian_data.foo:295-479`.
    """
    out = []
    a = l
    while a >= (l + 2) // 3:
        b = min(l - a, a)
        while b >= (l - a + 1) // 2:
            c = l - a - b
            if a == b == c:
                out.append((a, a, a))
            elif a > b and b == c:
                out += [(a, b, b), (b, a, b), (b, b, a)]
            elif a == b and b > c:
                out += [(a, a, c), (a, c, a), (c, a, a)]
            else:
                out += [(a, b, c), (a, c, b), (b, a, c), (c, a, b), (b, c, a), (c, b, a)]
            b -= 1
        a -= 1
    assert len(out) == n_comp(l)
    return out


def double_factorial_odd(l):
    """(2l-1)!!, Tonto's `INT:double_factorial` (`foofiles/int.foo:192`)."""
    r = 1.0
    for k in range(1, 2 * l, 2):
        r *= k
    return r


def unnormalize(l, exps, coefs):
    """Tonto's `SHELL:unnormalize` applied to library coefficients.

    norm = sqrt(sum_ij c_i c_j (2 sqrt(a_i a_j)/(a_i+a_j))^(l+3/2)) (2 pi)^(3/4)
    c_i <- c_i (4 a_i)^(l/2+3/4) / (norm sqrt((2l-1)!!))
    """
    exps = [float(e) for e in exps]
    coefs = [float(c) for c in coefs]
    sm = 0.0
    for i, a in enumerate(exps):
        for j in range(i):
            b = exps[j]
            ab = 2.0 * math.sqrt(a * b) / (a + b)
            sm += 2.0 * coefs[i] * coefs[j] * ab ** (l + 1.5)
        sm += coefs[i] * coefs[i]
    norm = math.sqrt(sm) * (2.0 * math.pi) ** 0.75
    df = double_factorial_odd(l)
    return [c * (4.0 * a) ** (0.5 * l + 0.75) / (norm * math.sqrt(df))
            for a, c in zip(exps, coefs)]


@dataclass
class Shell:
    l: int
    exps: list
    coefs: list          # un-normalised (Tonto internal) contraction coefficients
    atom: int
    center: tuple        # bohr
    first_bf: int

    @property
    def n_comp(self):
        return n_comp(self.l)


@dataclass
class Molecule:
    """Atoms (bohr), charge and the shell list for one basis."""
    symbols: list
    coords: np.ndarray                     # (n_atom, 3) bohr
    basis_name: str
    charge: int = 0
    shells: list = field(default_factory=list)

    @classmethod
    def from_angstrom(cls, symbols, coords_ang, basis_name, charge=0):
        coords = np.asarray(coords_ang, dtype=np.float64) / ANGSTROM_PER_BOHR
        m = cls(list(symbols), coords, basis_name, charge)
        m.build_shells()
        return m

    def build_shells(self):
        lib = _library()[self.basis_name]
        self.shells = []
        bf = 0
        for ia, sym in enumerate(self.symbols):
            for l, exps, coefs in lib[sym]:
                self.shells.append(Shell(l, list(exps), unnormalize(l, exps, coefs), ia,
                                         tuple(float(x) for x in self.coords[ia]), bf))
                bf += n_comp(l)
        self.n_bf = bf

    @property
    def n_atom(self):
        return len(self.symbols)

    @property
    def Z(self):
        return np.array([ELEMENT_Z[s] for s in self.symbols], dtype=np.float64)

    @property
    def n_electrons(self):
        return int(round(self.Z.sum())) - self.charge

    @property
    def n_shell(self):
        return len(self.shells)

    def nuclear_repulsion(self):
        Z = self.Z
        e = 0.0
        for i in range(self.n_atom):
            for j in range(i):
                e += Z[i] * Z[j] / float(np.linalg.norm(self.coords[i] - self.coords[j]))
        return e

    def shell_arrays(self):
        """Flat shell description handed across the C ABI (see include/tonto_gpu.h)."""
        ns = self.n_shell
        l = np.array([s.l for s in self.shells], dtype=np.int32)
        nprim = np.array([len(s.exps) for s in self.shells], dtype=np.int32)
        atom = np.array([s.atom for s in self.shells], dtype=np.int32)
        first_bf = np.array([s.first_bf for s in self.shells], dtype=np.int32)
        prim_off = np.zeros(ns + 1, dtype=np.int32)
        prim_off[1:] = np.cumsum(nprim)
        exps = np.array([e for s in self.shells for e in s.exps], dtype=np.float64)
        coefs = np.array([c for s in self.shells for c in s.coefs], dtype=np.float64)
        return dict(l=l, nprim=nprim, atom=atom, first_bf=first_bf, prim_off=prim_off,
                    exps=exps, coefs=coefs,
                    atom_xyz=np.ascontiguousarray(self.coords, dtype=np.float64),
                    atom_Z=self.Z)

"""ORACLE (test infrastructure only): ctypes wrapper of oracle/liboracle.so.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module,
as the checker.  The product (tonto_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        d = ctypes.POINTER(ctypes.c_double)
        _lib.orc_make_r_JK.argtypes = [ctypes.c_void_p, ctypes.c_void_p, d, ctypes.c_void_p, d, d,
                                       ctypes.POINTER(ctypes.c_long)]
        _lib.orc_make_r_JK_sample.argtypes = _lib.orc_make_r_JK.argtypes + [ctypes.c_long]
        _lib.orc_make_u_JK.argtypes = [ctypes.c_void_p, ctypes.c_void_p, d, d, ctypes.c_void_p, d, d, d, d,
                                       ctypes.POINTER(ctypes.c_long)]
        _lib.orc_eri_quartet.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 5 + \
            [ctypes.c_double, ctypes.c_int, d]
        _lib.orc_eri_quartet.restype = ctypes.c_long
        _lib.orc_schwarz.argtypes = [ctypes.c_void_p, ctypes.c_int, d]
        _lib.orc_one_electron.argtypes = [ctypes.c_void_p, d, d, d]
        _lib.orc_rys_table.argtypes = [ctypes.c_int, ctypes.c_double, d, d]
        _lib.orq_rys.argtypes = [ctypes.c_int, ctypes.c_double, d, d]
        _lib.orq_rys.restype = ctypes.c_int
        _lib.orq_boys.argtypes = [ctypes.c_int, ctypes.c_double]
        _lib.orq_boys.restype = ctypes.c_double
        _lib.orc_set_threads.argtypes = [ctypes.c_int]
        _lib.orc_max_threads.restype = ctypes.c_int
        _lib.orc_component.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        _lib.orc_quartet_audit.argtypes = [ctypes.c_void_p, ctypes.c_void_p, d, d, ctypes.c_void_p, d]
        _lib.orc_quartet_audit.restype = ctypes.c_long
        _lib.orc_pair_prims.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 3 + [ctypes.c_double]
        _lib.orc_pair_prims.restype = ctypes.c_int
        _lib.orc_set_plant.argtypes = [ctypes.c_int]
        _lib.orc_make_r_JK_direct.argtypes = [ctypes.c_void_p, ctypes.c_void_p, d, ctypes.c_void_p, d, d,
                                              ctypes.POINTER(ctypes.c_long)]
        _lib.orc_schwarz_sph.argtypes = [ctypes.c_void_p, d]
        _lib.orc_solid_harmonics.argtypes = [ctypes.c_int, d]
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class _OSys(ctypes.Structure):
    _fields_ = [("n_shell", ctypes.c_int), ("n_atom", ctypes.c_int), ("n_bf", ctypes.c_int)] + \
        [(k, ctypes.POINTER(ctypes.c_int)) for k in ("l", "nprim", "atom", "first_bf", "prim_off")] + \
        [(k, ctypes.POINTER(ctypes.c_double)) for k in ("exps", "coefs", "atom_xyz", "atom_Z")]


class _OCut(ctypes.Structure):
    _fields_ = [("schwarz", ctypes.c_double), ("jp", ctypes.c_double), ("kp", ctypes.c_double),
                ("pair", ctypes.c_double), ("rys_quad", ctypes.c_int)]


class OracleSystem:
    """Holds a molecule's shell arrays alive for the C oracle."""

    def __init__(self, mol):
        self.mol = mol
        self.arr = mol.shell_arrays()
        a = self.arr
        self.s = _OSys(mol.n_shell, mol.n_atom, mol.n_bf, _ip(a["l"]), _ip(a["nprim"]),
                       _ip(a["atom"]), _ip(a["first_bf"]), _ip(a["prim_off"]), _dp(a["exps"]),
                       _dp(a["coefs"]), _dp(a["atom_xyz"]), _dp(a["atom_Z"]))
        self.n_bf = mol.n_bf
        self.n_pairs = mol.n_shell * (mol.n_shell + 1) // 2

    def schwarz(self, rys_quad=False):
        out = np.zeros(self.n_pairs)
        lib().orc_schwarz(ctypes.byref(self.s), int(rys_quad), _dp(out))
        return out

    def make_r_JK(self, P, schwarz=1e-9, jp=1e-9, kp=1e-9, pair=1e-6, max_I=None,
                  rys_quad=False, return_stats=False):
        P = np.ascontiguousarray(P, dtype=np.float64)
        n = self.n_bf
        J = np.zeros((n, n))
        K = np.zeros((n, n))
        cut = _OCut(schwarz, jp, kp, pair, int(rys_quad))
        st = (ctypes.c_long * 5)()
        mi = None
        if max_I is not None:
            mi = np.ascontiguousarray(max_I, dtype=np.float64)
        lib().orc_make_r_JK(ctypes.byref(self.s), ctypes.byref(cut), _dp(P),
                            None if mi is None else mi.ctypes.data, _dp(J), _dp(K), st)
        if return_stats:
            return J, K, dict(quartets=st[0], prim_quartets=st[1], skip_J=st[2], skip_K=st[3],
                              skip_lists=st[4])
        return J, K

    def quartet_audit(self, P, max_I, schwarz=1e-9, jp=1e-9, kp=1e-9, pair=1e-6):
        """Per unique quartet (ab >= cd, index ab(ab+1)/2 + cd): what the build does with
        it (0 dropped; bit 0 J, bit 1 K digested; bit 2 skipped primitive lists) and the
        exact max |(ab|cd)| (full primitive lists)."""
        P = np.ascontiguousarray(P, dtype=np.float64)
        mi = np.ascontiguousarray(max_I, dtype=np.float64)
        npair = self.n_pairs
        nq = npair * (npair + 1) // 2
        code = np.zeros(nq, dtype=np.uint8)
        maxabs = np.zeros(nq)
        cut = _OCut(schwarz, jp, kp, pair, 0)
        lib().orc_quartet_audit(ctypes.byref(self.s), ctypes.byref(cut), _dp(P), _dp(mi),
                                code.ctypes.data, _dp(maxabs))
        return code, maxabs

    @property
    def n_sph(self):
        return sum(2 * s.l + 1 if s.l >= 2 else s.n_comp for s in self.mol.shells)

    def schwarz_sph(self):
        """max_I from the spherical (ab|ab) blocks."""
        out = np.zeros(self.n_pairs)
        lib().orc_schwarz_sph(ctypes.byref(self.s), _dp(out))
        return out

    def make_r_JK_direct(self, P, schwarz=1e-9, pair=1e-6, max_I=None, return_stats=False):
        """Spherical-basis J/K (make_r_JK_direct): P, J, K are n_sph x n_sph."""
        P = np.ascontiguousarray(P, dtype=np.float64)
        n = self.n_sph
        assert P.shape == (n, n)
        J, K = np.zeros((n, n)), np.zeros((n, n))
        cut = _OCut(schwarz, 0.0, 0.0, pair, 0)
        st = (ctypes.c_long * 5)()
        mi = None if max_I is None else np.ascontiguousarray(max_I, dtype=np.float64)
        lib().orc_make_r_JK_direct(ctypes.byref(self.s), ctypes.byref(cut), _dp(P),
                                   None if mi is None else mi.ctypes.data, _dp(J), _dp(K), st)
        if return_stats:
            return J, K, dict(quartets=st[0], prim_quartets=st[1], skip_lists=st[4])
        return J, K

    def pair_prims(self, a, b, skip, pair=1e-6):
        return lib().orc_pair_prims(ctypes.byref(self.s), a, b, int(skip), pair)

    def make_u_JK(self, Pa, Pb, schwarz=1e-9, jp=1e-9, kp=1e-9, pair=1e-6, max_I=None, return_stats=False):
        """UHF J/K as Tonto's make_u_JK_engine returns them (molecule.fock.foo:2558-2712):
        (Ja, Jb, Ka, Kb); Ja + Jb = J(Pa + Pb), Ka = K(Pa), Kb = K(Pb), with the J split
        of the engine's JBcd => JA(cd) aliasing (:2637) and max(|Pa|, |Pb|) screening."""
        Pa = np.ascontiguousarray(Pa, dtype=np.float64)
        Pb = np.ascontiguousarray(Pb, dtype=np.float64)
        n = self.n_bf
        out = [np.zeros((n, n)) for _ in range(4)]
        cut = _OCut(schwarz, jp, kp, pair, 0)
        st = (ctypes.c_long * 5)()
        mi = None if max_I is None else np.ascontiguousarray(max_I, dtype=np.float64)
        lib().orc_make_u_JK(ctypes.byref(self.s), ctypes.byref(cut), _dp(Pa), _dp(Pb),
                            None if mi is None else mi.ctypes.data, *[_dp(x) for x in out], st)
        if return_stats:
            return (*out, dict(quartets=st[0], prim_quartets=st[1], skip_J=st[2], skip_K=st[3],
                               skip_lists=st[4]))
        return tuple(out)

    def make_r_JK_sample(self, P, stride, schwarz=1e-9, jp=1e-9, kp=1e-9, pair=1e-6, max_I=None):
        """Timing sample: the build over every stride-th (ab| pair; returns stats."""
        P = np.ascontiguousarray(P, dtype=np.float64)
        n = self.n_bf
        J, K = np.zeros((n, n)), np.zeros((n, n))
        cut = _OCut(schwarz, jp, kp, pair, 0)
        st = (ctypes.c_long * 5)()
        mi = None if max_I is None else np.ascontiguousarray(max_I, dtype=np.float64)
        lib().orc_make_r_JK_sample(ctypes.byref(self.s), ctypes.byref(cut), _dp(P),
                                   None if mi is None else mi.ctypes.data, _dp(J), _dp(K), st,
                                   int(stride))
        return dict(quartets=st[0], prim_quartets=st[1])

    def eri_quartet(self, a, b, c, d, skip=False, pair=1e-6, rys_quad=False):
        sh = self.mol.shells
        dims = [sh[i].n_comp for i in (a, b, c, d)]
        out = np.zeros(int(np.prod(dims)))
        lib().orc_eri_quartet(ctypes.byref(self.s), a, b, c, d, int(skip), pair, int(rys_quad),
                              _dp(out))
        return out.reshape(dims)

    def one_electron(self):
        n = self.n_bf
        S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
        lib().orc_one_electron(ctypes.byref(self.s), _dp(S), _dp(T), _dp(V))
        return S, T, V

    def full_eri(self, rys_quad=False):
        """All (pq|rs) over basis functions, no screening/skipping (tiny systems only)."""
        sh = self.mol.shells
        n = self.n_bf
        eri = np.zeros((n, n, n, n))
        for a in range(len(sh)):
            for b in range(len(sh)):
                for c in range(len(sh)):
                    for d in range(len(sh)):
                        blk = self.eri_quartet(a, b, c, d, rys_quad=rys_quad)
                        fa, fb, fc, fd = (sh[i].first_bf for i in (a, b, c, d))
                        eri[fa:fa + blk.shape[0], fb:fb + blk.shape[1], fc:fc + blk.shape[2],
                            fd:fd + blk.shape[3]] = blk
        return eri


def rys_table(n, X):
    u, w = np.zeros(n), np.zeros(n)
    lib().orc_rys_table(n, X, _dp(u), _dp(w))
    return u, w


def rys_quad(n, X):
    u, w = np.zeros(n), np.zeros(n)
    rc = lib().orq_rys(n, X, _dp(u), _dp(w))
    assert rc == 0
    return u, w


def boys_quad(m, X):
    return lib().orq_boys(m, X)


def set_threads(n):
    lib().orc_set_threads(int(n))


def solid_harmonics(l):
    """The oracle's closed-form Cartesian -> real solid harmonic matrix (ncomp(l), 2l+1)."""
    out = np.zeros(((l + 1) * (l + 2) // 2) * (2 * l + 1))
    lib().orc_solid_harmonics(int(l), _dp(out))
    return out.reshape((l + 1) * (l + 2) // 2, 2 * l + 1)


def set_plant(k):
    """Planted screening mistake (0: none); only the test of the screening audit sets it."""
    lib().orc_set_plant(int(k))


def max_threads():
    return lib().orc_max_threads()

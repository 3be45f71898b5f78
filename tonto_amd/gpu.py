"""Python host binding of libtonto_gpu.so (the C ABI in include/tonto_gpu.h).

This is the product path: it loads the in-tree HIP library and raises if it is
missing or if no GPU is present; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# the library built in-tree by tonto_amd.build; TONTO_GPU_LIB selects another build of it
LIB_PATH = os.environ.get("TONTO_GPU_LIB") or os.path.join(HERE, "libtonto_gpu.so")

TGPU_JK_MIXED = 1
TGPU_JK_GRAPH = 2
TGPU_J_ONLY = 4
TGPU_K_ONLY = 8

# exported symbols, in the order of include/tonto_gpu.h (plus the two extensions)
EXPORTS = ["tgpu_create", "tgpu_destroy", "tgpu_last_error", "tgpu_dims", "tgpu_set_stream",
           "tgpu_set_cutoffs", "tgpu_set_fp32_threshold", "tgpu_set_spherical", "tgpu_schwarz", "tgpu_make_r_JK",
           "tgpu_make_r_JK_device", "tgpu_make_r_JK_device_noalloc", "tgpu_make_u_JK", "tgpu_make_u_JK_device", "tgpu_get_stats",
           "tgpu_one_electron",
           "tgpu_get_class_counters", "tgpu_n_class_pairs", "tgpu_set_profiling",
           "tgpu_get_class_times", "tgpu_create_batch", "tgpu_batch_layout", "tgpu_set_variant",
           "tgpu_get_variant", "tgpu_set_ket_chunk", "tgpu_get_ket_chunk", "tgpu_autotune"]

_lib = None


class TGPUError(RuntimeError):
    pass


class _System(ctypes.Structure):
    _fields_ = [("n_atom", ctypes.c_int), ("n_shell", ctypes.c_int),
                ("atom_xyz", ctypes.POINTER(ctypes.c_double)),
                ("atom_Z", ctypes.POINTER(ctypes.c_double)),
                ("shell_l", ctypes.POINTER(ctypes.c_int)),
                ("shell_atom", ctypes.POINTER(ctypes.c_int)),
                ("shell_nprim", ctypes.POINTER(ctypes.c_int)),
                ("exps", ctypes.POINTER(ctypes.c_double)),
                ("coefs", ctypes.POINTER(ctypes.c_double))]


class Stats(ctypes.Structure):
    _fields_ = [("quartets", ctypes.c_ulonglong), ("prim_quartets", ctypes.c_ulonglong),
                ("fp32_quartets", ctypes.c_ulonglong), ("kernel_ms", ctypes.c_double)]


def lib():
    """Load libtonto_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TGPUError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
        L.tgpu_create.argtypes = [ctypes.POINTER(_System), ctypes.POINTER(vp)]
        L.tgpu_destroy.argtypes = [vp]
        L.tgpu_last_error.argtypes = [vp]
        L.tgpu_last_error.restype = ctypes.c_char_p
        L.tgpu_dims.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_longlong)]
        L.tgpu_set_stream.argtypes = [vp, vp]
        L.tgpu_set_cutoffs.argtypes = [vp] + [ctypes.c_double] * 4
        L.tgpu_set_fp32_threshold.argtypes = [vp, ctypes.c_double]
        L.tgpu_set_spherical.argtypes = [vp, ctypes.c_int]
        L.tgpu_schwarz.argtypes = [vp, vp]
        L.tgpu_make_r_JK.argtypes = [vp, dp, dp, dp, ctypes.c_int]
        L.tgpu_make_r_JK_device.argtypes = [vp, vp, vp, vp, ctypes.c_int]
        L.tgpu_make_r_JK_device_noalloc.argtypes = [vp, vp, vp, vp, ctypes.c_int]
        L.tgpu_make_u_JK.argtypes = [vp, dp, dp, dp, dp, dp, dp, ctypes.c_int]
        L.tgpu_make_u_JK_device.argtypes = [vp, vp, vp, vp, vp, vp, vp, ctypes.c_int]
        L.tgpu_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
        L.tgpu_one_electron.argtypes = [vp, dp, dp, dp]
        L.tgpu_get_class_counters.argtypes = [vp, ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_int]
        L.tgpu_n_class_pairs.restype = ctypes.c_int
        L.tgpu_set_profiling.argtypes = [vp, ctypes.c_int]
        L.tgpu_get_class_times.argtypes = [vp, dp, ctypes.c_int]
        L.tgpu_create_batch.argtypes = [ctypes.POINTER(_System), ctypes.c_int, ctypes.POINTER(vp)]
        L.tgpu_batch_layout.argtypes = [vp, ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_longlong), vp, vp, vp]
        L.tgpu_set_variant.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.tgpu_get_variant.argtypes = [vp, ctypes.c_int]
        L.tgpu_set_ket_chunk.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.tgpu_get_ket_chunk.argtypes = [vp, ctypes.c_int]
        L.tgpu_autotune.argtypes = [vp, vp, ctypes.c_int, vp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class JKEngine:
    """A context on the current HIP device for one molecule, or for a list of
    independent molecules built together (mirrors Tonto's MOLECULE state consumed by
    make_r_JK_engine).  For a batch, matrices are packed molecule after molecule."""

    def __init__(self, mol, schwarz=1e-9, j_density=1e-9, k_density=1e-9, pair=1e-6,
                 fp32_threshold=1e-6, spherical=False):
        mols = list(mol) if isinstance(mol, (list, tuple)) else [mol]
        self.mol = mol
        self.mols = mols
        self._keep = [m.shell_arrays() for m in mols]
        arr = (_System * len(mols))()
        for i, (m, a) in enumerate(zip(mols, self._keep)):
            arr[i] = _System(m.n_atom, m.n_shell, _dp(a["atom_xyz"]), _dp(a["atom_Z"]),
                             _ip(a["l"]), _ip(a["atom"]), _ip(a["nprim"]), _dp(a["exps"]),
                             _dp(a["coefs"]))
        self._sys = arr
        self._ctx = ctypes.c_void_p()
        rc = lib().tgpu_create_batch(arr, len(mols), ctypes.byref(self._ctx))
        self._check(rc)
        self.n_mol = len(mols)
        self.n_bf = mols[0].n_bf
        self.mat_off = np.zeros(self.n_mol, dtype=np.int64)
        self.nbf = np.zeros(self.n_mol, dtype=np.int32)
        self.pair_off = np.zeros(self.n_mol + 1, dtype=np.int64)
        nm, ne = ctypes.c_int(), ctypes.c_longlong()
        self._check(lib().tgpu_batch_layout(self._ctx, ctypes.byref(nm), ctypes.byref(ne),
                                            self.mat_off.ctypes.data, self.nbf.ctypes.data,
                                            self.pair_off.ctypes.data))
        self.n_elements = ne.value
        self.n_pairs = int(self.pair_off[-1])
        self.cutoffs = (schwarz, j_density, k_density, pair)
        self._check(lib().tgpu_set_cutoffs(self._ctx, schwarz, j_density, k_density, pair))
        self._check(lib().tgpu_set_fp32_threshold(self._ctx, fp32_threshold))
        self.spherical = bool(spherical)
        if self.spherical:
            # matrices in the solid-harmonic basis (Tonto's use_spherical_basis)
            self._check(lib().tgpu_set_spherical(self._ctx, 1))
            nb = ctypes.c_int()
            self._check(lib().tgpu_dims(self._ctx, ctypes.byref(nb), None))
            self.n_bf = nb.value
            self.n_elements = self.n_bf * self.n_bf

    def _check(self, rc):
        if rc != 0:
            msg = lib().tgpu_last_error(self._ctx).decode() if self._ctx else ""
            raise TGPUError(f"libtonto_gpu error {rc}: {msg}")

    def close(self):
        if self._ctx:
            lib().tgpu_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_cutoffs_scaled(self, scale, density_scale=None):
        """Schwarz cutoff times `scale`, the J and K density cutoffs times `density_scale`
        (default: `scale`); the pair cutoff is kept."""
        s, j, k, p = self.cutoffs
        d = scale if density_scale is None else density_scale
        self._check(lib().tgpu_set_cutoffs(self._ctx, s * scale, j * d, k * d, p))

    def set_stream(self, stream_ptr):
        self._check(lib().tgpu_set_stream(self._ctx, ctypes.c_void_p(stream_ptr)))

    def schwarz(self):
        out = np.zeros(self.n_pairs)
        self._check(lib().tgpu_schwarz(self._ctx, out.ctypes.data))
        return out

    def pack(self, mats):
        """Pack a list of per-molecule square matrices into the batch layout."""
        return np.concatenate([np.ascontiguousarray(m, dtype=np.float64).ravel() for m in mats])

    def unpack(self, flat):
        return [flat[o:o + n * n].reshape(n, n) for o, n in zip(self.mat_off, self.nbf)]

    def make_r_JK(self, P, mixed=False, only=None):
        """Host arrays in, host arrays out (synchronous).  For a batch, P is a list of
        matrices (or the packed vector) and lists are returned."""
        batch = self.n_mol > 1 or isinstance(P, (list, tuple))
        Pf = self.pack(P) if isinstance(P, (list, tuple)) else np.ascontiguousarray(P, dtype=np.float64).ravel()
        assert Pf.size == self.n_elements
        J, K = np.zeros(self.n_elements), np.zeros(self.n_elements)
        flags = (TGPU_JK_MIXED if mixed else 0) | {None: 0, "J": TGPU_J_ONLY, "K": TGPU_K_ONLY}[only]
        self._check(lib().tgpu_make_r_JK(self._ctx, _dp(Pf), _dp(J), _dp(K), flags))
        if batch:
            return self.unpack(J), self.unpack(K)
        n = self.n_bf
        return J.reshape(n, n), K.reshape(n, n)

    def make_u_JK(self, Pa, Pb, mixed=False):
        """UHF: host arrays Pa, Pb in, (Ja, Jb, Ka, Kb) out (synchronous; one molecule)."""
        assert self.n_mol == 1
        n = self.n_bf
        Pa = np.ascontiguousarray(Pa, dtype=np.float64).ravel()
        Pb = np.ascontiguousarray(Pb, dtype=np.float64).ravel()
        assert Pa.size == Pb.size == n * n
        out = [np.zeros(n * n) for _ in range(4)]
        self._check(lib().tgpu_make_u_JK(self._ctx, _dp(Pa), _dp(Pb), *[_dp(x) for x in out],
                                         TGPU_JK_MIXED if mixed else 0))
        return tuple(x.reshape(n, n) for x in out)

    def make_u_JK_device(self, Pa_ptr, Pb_ptr, Ja_ptr, Jb_ptr, Ka_ptr, Kb_ptr, mixed=False):
        """UHF on device pointers, asynchronous on the stream (one molecule)."""
        assert self.n_mol == 1
        self._check(lib().tgpu_make_u_JK_device(self._ctx, *[ctypes.c_void_p(p) for p in
                                                             (Pa_ptr, Pb_ptr, Ja_ptr, Jb_ptr, Ka_ptr, Kb_ptr)],
                                                TGPU_JK_MIXED if mixed else 0))

    def make_r_JK_device(self, P_ptr, J_ptr, K_ptr, mixed=False, graph=False):
        """Device pointers (e.g. torch tensor .data_ptr()), asynchronous on the stream."""
        flags = (TGPU_JK_MIXED if mixed else 0) | (TGPU_JK_GRAPH if graph else 0)
        self._check(lib().tgpu_make_r_JK_device(self._ctx, ctypes.c_void_p(P_ptr),
                                                ctypes.c_void_p(J_ptr), ctypes.c_void_p(K_ptr),
                                                flags))

    def make_r_JK_device_noalloc(self, P_ptr, J_ptr, K_ptr, mixed=False, only=None):
        """The allocation-free build: device pointers, asynchronous on the stream, its
        scratch buffers in the library's per-GPU array (TGPUError if they do not fit)."""
        flags = (TGPU_JK_MIXED if mixed else 0) | {None: 0, "J": TGPU_J_ONLY, "K": TGPU_K_ONLY}[only]
        self._check(lib().tgpu_make_r_JK_device_noalloc(self._ctx, ctypes.c_void_p(P_ptr),
                                                        ctypes.c_void_p(J_ptr), ctypes.c_void_p(K_ptr),
                                                        flags))

    def autotune(self, P_ptr, mixed=False):
        n = lib().tgpu_n_class_pairs()
        ms = np.zeros(2 * n)
        self._check(lib().tgpu_autotune(self._ctx, ctypes.c_void_p(P_ptr),
                                        TGPU_JK_MIXED if mixed else 0, ms.ctypes.data))
        return ms.reshape(n, 2), self.variants()

    def variants(self):
        return [lib().tgpu_get_variant(self._ctx, k) for k in range(lib().tgpu_n_class_pairs())]

    def set_variant(self, k, v):
        self._check(lib().tgpu_set_variant(self._ctx, k, v))

    def ket_chunks(self):
        """Per class pair: the wavefront form's ket chunk (0: chosen per launch)."""
        return [lib().tgpu_get_ket_chunk(self._ctx, k) for k in range(lib().tgpu_n_class_pairs())]

    def set_ket_chunk(self, k, chunk):
        self._check(lib().tgpu_set_ket_chunk(self._ctx, k, chunk))

    def stats(self):
        s = Stats()
        self._check(lib().tgpu_get_stats(self._ctx, ctypes.byref(s)))
        return dict(quartets=s.quartets, prim_quartets=s.prim_quartets,
                    fp32_quartets=s.fp32_quartets, kernel_ms=s.kernel_ms)

    def class_counters(self):
        n = lib().tgpu_n_class_pairs()
        buf = (ctypes.c_ulonglong * (3 * n))()
        self._check(lib().tgpu_get_class_counters(self._ctx, buf, 3 * n))
        return np.array(buf[:], dtype=np.uint64).reshape(n, 3)

    def set_profiling(self, on=True):
        self._check(lib().tgpu_set_profiling(self._ctx, int(on)))

    def class_times(self):
        n = lib().tgpu_n_class_pairs()
        out = np.zeros(n)
        self._check(lib().tgpu_get_class_times(self._ctx, _dp(out), n))
        return out

    def one_electron(self):
        S, T, V = (np.zeros(self.n_elements) for _ in range(3))
        self._check(lib().tgpu_one_electron(self._ctx, _dp(S), _dp(T), _dp(V)))
        if self.n_mol > 1:
            return self.unpack(S), self.unpack(T), self.unpack(V)
        n = self.n_bf
        return S.reshape(n, n), T.reshape(n, n), V.reshape(n, n)

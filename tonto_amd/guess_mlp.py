"""Learned initial density for water clusters: density-matrix blocks of atom pairs
predicted by small MLPs (tonto_amd.mlp, HIP kernels) from radial-basis descriptors.

Blocks and frames.  Every block P_AB (basis functions of atom A x those of atom B) is
predicted in a local frame built from the geometry, so the guess turns with the
molecule:
  * on-site (A = B): z along the water's bisector (O) or along H->O (H); x in the
    molecular plane (towards the molecule's first H for O, towards the other H for H);
  * pairs (A != B, distance < R_CUT): z along A->B; x from the vector sum of A's bonds
    within its own water, made perpendicular to z (B's, if A's is nearly parallel to z).
For the frame's rows Q = (e_x, e_y, e_z) a Cartesian shell of angular momentum l
transforms with T_l(Q), T[c, c'] = coefficient of the monomial c' in c(Q^T u); all
components of a Tonto shell share one normalisation (x^l), so P_local = T_A^T P T_B and
P = T_A(Q^T)^T P_local T_B(Q^T) exactly (T_l(Q) T_l(Q^T) = 1).

Descriptors (inputs): 16 Gaussians of the pair distance on [0, R_CUT] (width = spacing),
a same-molecule flag, and for each end of the pair the neighbour environment within
R_CUT: sum over neighbours C of each element of g_k(r) (1, u_x, u_y, u_z) with u the
local-frame direction to C, 16 Gaussians each -> 16 + 1 + 2 * 2 * 64 = 273 inputs.

Targets: the converged RHF density's local-frame blocks minus the promolecule (atomic
densities) blocks, which are zero off the diagonal.  Types: "O" (on-site O, upper
triangle), "H" (on-site H, upper triangle), "OO", "OH" (O rows, H columns), "HH";
same-element pairs are taken with A the lower atom index and P_BA = P_AB^T.
"""
import os

import numpy as np

from .basis import ANGSTROM_PER_BOHR, cart_components

R_CUT = 6.0          # Angstrom
N_RBF = 16
NH = 128
TYPES = ("O", "H", "OO", "OH", "HH")
N_IN = N_RBF + 1 + 2 * 2 * 4 * N_RBF

_MU = np.linspace(0.0, R_CUT, N_RBF)
_SIG = R_CUT / (N_RBF - 1)


def rbf(r):
    """16 Gaussians of r (Angstrom) times a cosine cutoff that vanishes at R_CUT."""
    r = np.asarray(r, np.float64)
    g = np.exp(-(((r[..., None] - _MU) / _SIG) ** 2))
    fc = np.where(r < R_CUT, 0.5 * (np.cos(np.pi * r / R_CUT) + 1.0), 0.0)
    return g * fc[..., None]


# ----------------------------------------------------------------------------- rotations

_MULTI = {}


def _multi(l):
    """All l-long axis sequences (3^l, l) and the component index of each one's monomial."""
    if l not in _MULTI:
        seqs = np.array(np.meshgrid(*[np.arange(3)] * l, indexing="ij")).reshape(l, -1).T
        index = {c: i for i, c in enumerate(cart_components(l))}
        mono = np.array([index[tuple(int(np.sum(s == k)) for k in range(3))] for s in seqs])
        axes = [np.array([0] * c[0] + [1] * c[1] + [2] * c[2]) for c in cart_components(l)]
        _MULTI[l] = (seqs, mono, axes)
    return _MULTI[l]


def cart_transform(l, Q):
    """T_l(Q) in Tonto's component order: row c holds the expansion of the monomial c
    evaluated at Q^T u over the monomials of u ((Q^T u)_axis = sum_k Q[k, axis] u_k)."""
    Q = np.asarray(Q, np.float64)
    n = (l + 1) * (l + 2) // 2
    if l == 0:
        return np.ones((1, 1))
    seqs, mono, axes = _multi(l)
    T = np.zeros((n, n))
    for i, ax in enumerate(axes):
        prod = np.ones(seqs.shape[0])
        for pos in range(l):
            prod = prod * Q[seqs[:, pos], ax[pos]]
        np.add.at(T[i], mono, prod)
    return T


def cart_transform_batch(l, Q):
    """cart_transform for a stack of frames Q (nb, 3, 3): (nb, n, n)."""
    Q = np.asarray(Q, np.float64)
    nb = Q.shape[0]
    n = (l + 1) * (l + 2) // 2
    if l == 0:
        return np.ones((nb, 1, 1))
    seqs, mono, axes = _multi(l)
    onehot = np.zeros((seqs.shape[0], n))
    onehot[np.arange(seqs.shape[0]), mono] = 1.0
    T = np.empty((nb, n, n))
    for i, ax in enumerate(axes):
        prod = np.ones((nb, seqs.shape[0]))
        for pos in range(l):
            prod = prod * Q[:, seqs[:, pos], ax[pos]]
        T[:, i, :] = prod @ onehot
    return T


def atom_transform_batch(ls, Q):
    """atom_transform for a stack of frames Q (nb, 3, 3)."""
    blocks = [cart_transform_batch(l, Q) for l in ls]
    n = sum(b.shape[1] for b in blocks)
    T = np.zeros((Q.shape[0], n, n))
    o = 0
    for b in blocks:
        k = b.shape[1]
        T[:, o:o + k, o:o + k] = b
        o += k
    return T


def atom_transform(ls, Q):
    """Block-diagonal transform of one atom's shells (angular momenta ls)."""
    blocks = [cart_transform(l, Q) for l in ls]
    n = sum(b.shape[0] for b in blocks)
    T = np.zeros((n, n))
    o = 0
    for b in blocks:
        k = b.shape[0]
        T[o:o + k, o:o + k] = b
        o += k
    return T


# ----------------------------------------------------------------------------- geometry

class Cluster:
    """Atoms of a water cluster in Angstrom with each H assigned to its nearest O."""

    def __init__(self, symbols, xyz_ang):
        self.symbols = list(symbols)
        self.xyz = np.asarray(xyz_ang, np.float64)
        n = len(self.symbols)
        self.is_O = np.array([s == "O" for s in self.symbols])
        if not all(s in ("O", "H") for s in self.symbols):
            raise ValueError("the learned guess covers water clusters (O and H) only")
        O = np.nonzero(self.is_O)[0]
        self.mol = np.empty(n, int)
        self.mol[O] = np.arange(len(O))
        for i in np.nonzero(~self.is_O)[0]:
            d = np.linalg.norm(self.xyz[O] - self.xyz[i], axis=1)
            self.mol[i] = int(np.argmin(d))
        self.members = [np.nonzero(self.mol == m)[0] for m in range(len(O))]
        self.D = np.linalg.norm(self.xyz[:, None, :] - self.xyz[None, :, :], axis=2)
        self.bond_sum = np.array([sum((self.xyz[b] - self.xyz[a] for b in self._mates(a)), np.zeros(3))
                                  for a in range(n)])

    @classmethod
    def from_molecule(cls, mol):
        return cls(mol.symbols, np.asarray(mol.coords) * ANGSTROM_PER_BOHR)

    def _mates(self, a):
        return [b for b in self.members[self.mol[a]] if b != a]

    def frame(self, a, b):
        """Rows e_x, e_y, e_z of the local frame of the block (a, b)."""
        if a == b:
            mates = self._mates(a)
            if self.is_O[a]:
                hs = [m for m in mates if not self.is_O[m]]
                ez = self.bond_sum[a]
                ref = self.xyz[hs[0]] - self.xyz[hs[1]] if len(hs) >= 2 else None
            else:
                o = [m for m in mates if self.is_O[m]]
                hs = [m for m in mates if not self.is_O[m]]
                ez = self.xyz[o[0]] - self.xyz[a]
                ref = self.xyz[hs[0]] - self.xyz[a] if hs else None
            cands = [ref]
        else:
            ez = self.xyz[b] - self.xyz[a]
            cands = [self.bond_sum[a], self.bond_sum[b]]
        ez = ez / np.linalg.norm(ez)
        cands = [c for c in cands if c is not None] + [np.array([1.0, 0, 0]), np.array([0, 1.0, 0])]
        for c in cands:
            perp = c - np.dot(c, ez) * ez
            nrm = np.linalg.norm(perp)
            if nrm > 0.1 * max(np.linalg.norm(c), 1e-12):
                ex = perp / nrm
                break
        ey = np.cross(ez, ex)
        return np.array([ex, ey, ez])

    def _env(self, a, Q, skip):
        """neighbour environment of atom a (all atoms within R_CUT but a and skip)"""
        f = np.zeros((2, N_RBF, 4))
        near = np.nonzero(self.D[a] < R_CUT)[0]
        near = near[(near != a) & (near != skip)]
        if near.size:
            r = self.D[a, near]
            u = ((self.xyz[near] - self.xyz[a]) / r[:, None]) @ Q.T
            g = rbf(r)
            h = np.concatenate([np.ones((near.size, 1)), u], axis=1)
            for e, mask in ((0, self.is_O[near]), (1, ~self.is_O[near])):
                if mask.any():
                    f[e] = g[mask].T @ h[mask]
        return f.ravel()

    def features(self, a, b, Q):
        r = self.D[a, b]
        x = np.empty(N_IN)
        x[:N_RBF] = rbf(r)
        x[N_RBF] = 1.0 if self.mol[a] == self.mol[b] else 0.0
        k = 4 * N_RBF * 2
        x[N_RBF + 1:N_RBF + 1 + k] = self._env(a, Q, b)
        x[N_RBF + 1 + k:] = self._env(b, Q, a)
        return x

    # ------------------------------------------------------------ whole-cluster forms
    # The same frames and descriptors as frame() and features(), for every block at once
    # (numpy over blocks; the environment sums are formed once per atom and turned into
    # each block's frame, which is linear in the directions).

    def _mate_refs(self):
        """per atom: the on-site frame's z vector and reference vector (zero if none)"""
        n = len(self.symbols)
        ez = np.zeros((n, 3))
        ref = np.zeros((n, 3))
        for a in range(n):
            mates = self._mates(a)
            hs = [m for m in mates if not self.is_O[m]]
            if self.is_O[a]:
                ez[a] = self.bond_sum[a]
                if len(hs) >= 2:
                    ref[a] = self.xyz[hs[0]] - self.xyz[hs[1]]
            else:
                o = [m for m in mates if self.is_O[m]]
                ez[a] = self.xyz[o[0]] - self.xyz[a]
                if hs:
                    ref[a] = self.xyz[hs[0]] - self.xyz[a]
        return ez, ref

    def frames(self, a, b):
        """frame(a[i], b[i]) for index arrays a, b: (nb, 3, 3)."""
        a = np.asarray(a)
        b = np.asarray(b)
        nb = a.size
        on = a == b
        ez_on, ref_on = self._mate_refs()
        ez = np.where(on[:, None], ez_on[a], self.xyz[b] - self.xyz[a])
        ez = ez / np.linalg.norm(ez, axis=1, keepdims=True)
        # candidates in frame()'s order: on-site (ref, x, y, -), pair (bond_sum a, b, x, y)
        c0 = np.where(on[:, None], ref_on[a], self.bond_sum[a])
        c1 = np.where(on[:, None], np.array([1.0, 0, 0]), self.bond_sum[b])
        c2 = np.where(on[:, None], np.array([0, 1.0, 0]), np.array([1.0, 0, 0]))
        c3 = np.where(on[:, None], np.zeros(3), np.array([0, 1.0, 0]))
        C = np.stack([c0, c1, c2, c3], axis=1)                           # (nb, 4, 3)
        perp = C - np.einsum("nkc,nc->nk", C, ez)[:, :, None] * ez[:, None, :]
        nrm = np.linalg.norm(perp, axis=2)
        ok = nrm > 0.1 * np.maximum(np.linalg.norm(C, axis=2), 1e-12)
        first = np.argmax(ok, axis=1)
        ex = perp[np.arange(nb), first] / nrm[np.arange(nb), first][:, None]
        ey = np.cross(ez, ex)
        return np.stack([ex, ey, ez], axis=1)

    def _env_sums(self):
        """per atom and element of the neighbour: sum of g_k(r) and of g_k(r) u (global
        axes) over neighbours within R_CUT: (n, 2, N_RBF), (n, 2, N_RBF, 3).  Formed once
        per cluster."""
        if getattr(self, "_sums", None) is not None:
            return self._sums
        n = len(self.symbols)
        S0 = np.zeros((n, 2, N_RBF))
        S1 = np.zeros((n, 2, N_RBF, 3))
        for a in range(n):
            near = np.nonzero(self.D[a] < R_CUT)[0]
            near = near[near != a]
            if near.size:
                r = self.D[a, near]
                u = (self.xyz[near] - self.xyz[a]) / r[:, None]
                g = rbf(r)
                for e, mask in ((0, self.is_O[near]), (1, ~self.is_O[near])):
                    S0[a, e] = g[mask].sum(axis=0)
                    S1[a, e] = g[mask].T @ u[mask]
        self._sums = (S0, S1)
        return S0, S1

    def features_batch(self, a, b, Q):
        """features(a[i], b[i], Q[i]) for every block: (nb, N_IN)."""
        a = np.asarray(a)
        b = np.asarray(b)
        S0, S1 = self._env_sums()
        r = self.D[a, b]
        X = np.empty((a.size, N_IN))
        X[:, :N_RBF] = rbf(r)
        X[:, N_RBF] = (self.mol[a] == self.mol[b]).astype(np.float64)
        k = 4 * N_RBF * 2
        for col, (x, y) in ((N_RBF + 1, (a, b)), (N_RBF + 1 + k, (b, a))):
            # environment of x without y (y is a neighbour of x unless y == x)
            s0 = S0[x].copy()
            s1 = S1[x].copy()
            drop = (y != x) & (self.D[x, y] < R_CUT)
            if drop.any():
                i = np.nonzero(drop)[0]
                ry = self.D[x[i], y[i]]
                gy = rbf(ry)
                uy = (self.xyz[y[i]] - self.xyz[x[i]]) / ry[:, None]
                ey = np.where(self.is_O[y[i]], 0, 1)
                s0[i, ey] -= gy
                s1[i, ey] -= gy[:, :, None] * uy[:, None, :]
            f = np.empty((a.size, 2, N_RBF, 4))
            f[..., 0] = s0
            f[..., 1:] = (s1.reshape(a.size, 2 * N_RBF, 3) @ np.swapaxes(Q, 1, 2)).reshape(a.size, 2, N_RBF, 3)
            X[:, col:col + k] = f.reshape(a.size, k)
        return X

    def block_index(self):
        """type -> (a, b) index arrays of its blocks, in blocks() order."""
        out = {}
        for t, a, b in self.blocks():
            out.setdefault(t, ([], []))
            out[t][0].append(a)
            out[t][1].append(b)
        return {t: (np.array(x), np.array(y)) for t, (x, y) in out.items()}

    def blocks(self):
        """[(type, a, b)] of every predicted block."""
        out = []
        n = len(self.symbols)
        for a in range(n):
            out.append(("O" if self.is_O[a] else "H", a, a))
        for a in range(n):
            for b in range(a + 1, n):
                if self.D[a, b] >= R_CUT:
                    continue
                if self.is_O[a] and self.is_O[b]:
                    out.append(("OO", a, b))
                elif self.is_O[a] != self.is_O[b]:
                    o, h = (a, b) if self.is_O[a] else (b, a)
                    out.append(("OH", o, h))
                else:
                    out.append(("HH", a, b))
        return out


# ----------------------------------------------------------------------------- blocks

def atom_layout(mol):
    """first basis function and shell angular momenta of every atom"""
    first = [None] * mol.n_atom
    ls = [[] for _ in range(mol.n_atom)]
    for s in mol.shells:
        if first[s.atom] is None:
            first[s.atom] = s.first_bf
        ls[s.atom].append(s.l)
    nbf = [sum((l + 1) * (l + 2) // 2 for l in x) for x in ls]
    return first, ls, nbf


def _pack(t, blk):
    if t in ("O", "H"):
        iu = np.triu_indices(blk.shape[0])
        return blk[iu]
    return blk.ravel()


def _unpack(t, vec, na, nb):
    if t in ("O", "H"):
        blk = np.zeros((na, na))
        iu = np.triu_indices(na)
        blk[iu] = vec
        blk = blk + blk.T - np.diag(np.diag(blk))
        return blk
    return np.asarray(vec).reshape(na, nb)


def out_dim(t, nbf_O, nbf_H):
    n = {"O": nbf_O, "H": nbf_H}
    if t in ("O", "H"):
        return n[t] * (n[t] + 1) // 2
    return n[t[0]] * n[t[1]]


def _gather_blocks(P, first, nbf, a, b, na, nb):
    """P's (na x nb) blocks at atoms a[i], b[i]: (n, na, nb)."""
    ri = np.asarray(first)[a][:, None] + np.arange(na)[None, :]
    ci = np.asarray(first)[b][:, None] + np.arange(nb)[None, :]
    return P[ri[:, :, None], ci[:, None, :]]


def training_rows(mol, P, atom_density):
    """(type -> (features, targets)) of one converged density P (numpy, Tonto order).
    atom_density: symbol -> promolecule block of that atom (numpy)."""
    cl = Cluster.from_molecule(mol)
    first, ls, nbf = atom_layout(mol)
    rows = {}
    for t, (a, b) in cl.block_index().items():
        Q = cl.frames(a, b)
        na, nb = nbf[a[0]], nbf[b[0]]
        blk = _gather_blocks(P, first, nbf, a, b, na, nb)
        TA = atom_transform_batch(ls[a[0]], Q)
        TB = atom_transform_batch(ls[b[0]], Q)
        loc = np.swapaxes(TA, 1, 2) @ blk @ TB
        if t in ("O", "H"):
            loc = loc - atom_density[mol.symbols[a[0]]][None]
            iu = np.triu_indices(na)
            Y = loc[:, iu[0], iu[1]]
        else:
            Y = loc.reshape(len(a), -1)
        rows[t] = (cl.features_batch(a, b, Q), Y)
    return rows


def assemble(mol, atom_density, predict):
    """Guess density: promolecule blocks plus predicted local-frame blocks turned back.
    predict(type, X) -> raw model outputs (denormalised) for the rows X."""
    cl = Cluster.from_molecule(mol)
    first, ls, nbf = atom_layout(mol)
    P = np.zeros((mol.n_bf, mol.n_bf))
    for t, (a, b) in cl.block_index().items():
        Q = cl.frames(a, b)
        Y = predict(t, cl.features_batch(a, b, Q))
        na, nb = nbf[a[0]], nbf[b[0]]
        if t in ("O", "H"):
            loc = np.zeros((len(a), na, na))
            iu = np.triu_indices(na)
            loc[:, iu[0], iu[1]] = Y
            loc = loc + np.swapaxes(loc, 1, 2)
            loc[:, np.arange(na), np.arange(na)] *= 0.5
            loc = loc + atom_density[mol.symbols[a[0]]][None]
        else:
            loc = np.asarray(Y).reshape(len(a), na, nb)
        QT = np.swapaxes(Q, 1, 2)
        TA = atom_transform_batch(ls[a[0]], QT)
        TB = atom_transform_batch(ls[b[0]], QT)
        blk = np.swapaxes(TA, 1, 2) @ loc @ TB
        ri = np.asarray(first)[a][:, None] + np.arange(na)[None, :]
        ci = np.asarray(first)[b][:, None] + np.arange(nb)[None, :]
        P[ri[:, :, None], ci[:, None, :]] = blk
        if t not in ("O", "H"):
            P[ci[:, :, None], ri[:, None, :]] = np.swapaxes(blk, 1, 2)
    return P


# ----------------------------------------------------------------------------- the model

class GuessModel:
    """Five networks with their input and output normalisation (a .npz file)."""

    def __init__(self, nets, meta):
        self.nets = nets          # type -> dict(params, x_mean, x_std, y_scale, n_out)
        self.meta = meta

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        nets = {}
        for t in TYPES:
            if f"{t}_params" in z:
                nets[t] = {k: z[f"{t}_{k}"] for k in ("params", "x_mean", "x_std", "y_scale")}
                nets[t]["n_out"] = int(z[f"{t}_n_out"])
        meta = {k[5:]: z[k] for k in z.files if k.startswith("meta_")}
        return cls(nets, meta)

    def save(self, path):
        d = {}
        for t, n in self.nets.items():
            for k in ("params", "x_mean", "x_std", "y_scale"):
                d[f"{t}_{k}"] = np.asarray(n[k], np.float32)
            d[f"{t}_n_out"] = np.int64(n["n_out"])
        for k, v in self.meta.items():
            d["meta_" + k] = np.asarray(v)
        np.savez_compressed(path, **d)

    def normalise(self, t, X):
        n = self.nets[t]
        return ((X - n["x_mean"]) / n["x_std"]).astype(np.float32)

    def predictor(self, forward):
        """predict(type, X) for assemble(); forward(type, params, Xn) runs the network."""
        def predict(t, X):
            n = self.nets[t]
            return forward(t, n["params"], self.normalise(t, X)).astype(np.float64) * n["y_scale"]
        return predict


DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "guess_mlp_631gd.npz")
_default = None


def default_model():
    """The committed 6-31G(d) water model, or None if it is absent."""
    global _default
    if _default is None and os.path.exists(DEFAULT_MODEL):
        _default = GuessModel.load(DEFAULT_MODEL)
    return _default


def gpu_forward(model, device="cuda"):
    """forward(type, params, Xn) on the GPU through the tmlp kernels (one DeviceNet per type,
    uploaded once per model and device)."""
    import torch
    from .mlp import DeviceNet
    cache = model.__dict__.setdefault("_device_nets", {})
    nets = cache.setdefault(str(device), {})

    def forward(t, params, Xn):
        if t not in nets:
            nets[t] = DeviceNet(N_IN, NH, model.nets[t]["n_out"], params, device)
        X = torch.tensor(np.ascontiguousarray(Xn, np.float32), device=device)
        return nets[t].forward(X).cpu().numpy()
    return forward


def normalise_electrons(P, S, n_electrons):
    """Scale P so that tr(P S) = N."""
    tr = float(np.sum(P * S))
    return P * (n_electrons / tr) if tr > 0 else P

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


def training_rows(mol, P, atom_density):
    """(type -> (features, targets)) of one converged density P (numpy, Tonto order).
    atom_density: symbol -> promolecule block of that atom (numpy)."""
    cl = Cluster.from_molecule(mol)
    first, ls, nbf = atom_layout(mol)
    rows = {t: ([], []) for t in TYPES}
    for t, a, b in cl.blocks():
        Q = cl.frame(a, b)
        blk = P[first[a]:first[a] + nbf[a], first[b]:first[b] + nbf[b]]
        loc = atom_transform(ls[a], Q).T @ blk @ atom_transform(ls[b], Q)
        if a == b:
            loc = loc - atom_density[mol.symbols[a]]
        rows[t][0].append(cl.features(a, b, Q))
        rows[t][1].append(_pack(t, loc))
    return {t: (np.array(x), np.array(y)) for t, (x, y) in rows.items() if x}


def assemble(mol, atom_density, predict):
    """Guess density: promolecule blocks plus predicted local-frame blocks turned back.
    predict(type, X) -> raw model outputs (denormalised) for the rows X."""
    cl = Cluster.from_molecule(mol)
    first, ls, nbf = atom_layout(mol)
    P = np.zeros((mol.n_bf, mol.n_bf))
    by_type = {}
    for t, a, b in cl.blocks():
        Q = cl.frame(a, b)
        by_type.setdefault(t, []).append((a, b, Q, cl.features(a, b, Q)))
    for t, items in by_type.items():
        Y = predict(t, np.array([it[3] for it in items]))
        for (a, b, Q, _), y in zip(items, Y):
            loc = _unpack(t, y, nbf[a], nbf[b])
            if a == b:
                loc = loc + atom_density[mol.symbols[a]]
            blk = atom_transform(ls[a], Q.T).T @ loc @ atom_transform(ls[b], Q.T)
            P[first[a]:first[a] + nbf[a], first[b]:first[b] + nbf[b]] = blk
            if a != b:
                P[first[b]:first[b] + nbf[b], first[a]:first[a] + nbf[a]] = blk.T
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
    """forward(type, params, Xn) on the GPU through the tmlp kernels (one DeviceNet per type)."""
    import torch
    from .mlp import DeviceNet
    nets = {}

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

"""CPU pins of the learned guess's pieces: the shell rotation matrices against the
physics (a rotated water's converged density), the frames and descriptors against
rigid motions, the assembled guess's covariance, and the numpy network checker
(oracle/mlp_ref.py) against finite differences and the Adam closed form."""
import numpy as np
import pytest
import torch

from tonto_amd import guess_mlp as G
from tonto_amd.basis import Molecule
from tonto_amd.geometry import water_cluster_xyz, _rotation
from tonto_amd.scf import run_rhf
import oracle.mlp_ref as R


def _rot(seed):
    return _rotation(np.random.default_rng(seed).standard_normal(4))


def test_p_shell_transform_is_q_transpose():
    Q = _rot(1)
    assert np.allclose(G.cart_transform(1, Q), Q.T, atol=1e-15)


@pytest.mark.parametrize("l", [0, 1, 2, 3])
def test_transform_inverse_and_identity(l):
    Q = _rot(2 + l)
    n = (l + 1) * (l + 2) // 2
    assert np.allclose(G.cart_transform(l, np.eye(3)), np.eye(n), atol=1e-15)
    assert np.allclose(G.cart_transform(l, Q) @ G.cart_transform(l, Q.T), np.eye(n), atol=1e-13)


def _scf_density(oracle_mod, mol):
    o = oracle_mod.OracleSystem(mol)
    S, T, V = o.one_electron()
    maxI = o.schwarz()

    def jk(P, mode):
        J, K = o.make_r_JK(P.numpy(), max_I=maxI, schwarz=1e-14)
        return torch.from_numpy(J), torch.from_numpy(K)
    r = run_rhf(torch.from_numpy(T + V), torch.from_numpy(S), mol.n_electrons,
                mol.nuclear_repulsion(), jk, convergence=1e-11, diis_tol=1e-9)
    assert r.converged
    return r.P.numpy(), r.energy


def test_rotated_water_density_follows_the_shell_transform(oracle_mod):
    """The converged 6-31G(d) density of a rotated water equals T^T P T of the original's
    (d shells included), which pins T against the integrals and the component order."""
    syms, xyz = water_cluster_xyz(1, seed=3)
    Q = _rot(7)
    P0, E0 = _scf_density(oracle_mod, Molecule.from_angstrom(syms, xyz, "6-31G(d)"))
    P1, E1 = _scf_density(oracle_mod, Molecule.from_angstrom(syms, xyz @ Q.T, "6-31G(d)"))
    assert abs(E0 - E1) < 1e-9
    mol = Molecule.from_angstrom(syms, xyz, "6-31G(d)")
    first, ls, nbf = G.atom_layout(mol)
    # the rotated molecule (r -> Q r) holds the original density seen in the frame with
    # rows Q: P1 = T(Q)^T P0 T(Q), shell by shell
    T = np.zeros((mol.n_bf, mol.n_bf))
    for a in range(mol.n_atom):
        T[first[a]:first[a] + nbf[a], first[a]:first[a] + nbf[a]] = G.atom_transform(ls[a], Q)
    assert np.max(np.abs(T.T @ P0 @ T - P1)) < 1e-9
    # and not with the inverse rotation
    assert np.max(np.abs(T @ P0 @ T.T - P1)) > 1e-3


def _cluster(seed, n=4, Q=None, shift=None):
    syms, xyz = water_cluster_xyz(n, seed=seed)
    if Q is not None:
        xyz = xyz @ Q.T + shift
    return syms, xyz


def test_frames_and_features_move_with_the_cluster():
    syms, xyz = _cluster(5)
    Q = _rot(9)
    syms2, xyz2 = _cluster(5, Q=Q, shift=np.array([1.0, -2.0, 0.5]))
    c1, c2 = G.Cluster(syms, xyz), G.Cluster(syms2, xyz2)
    b1, b2 = c1.blocks(), c2.blocks()
    assert b1 == b2 and len(b1) > 12
    for t, a, b in b1:
        F1, F2 = c1.frame(a, b), c2.frame(a, b)
        assert np.allclose(F1 @ Q.T, F2, atol=1e-12)       # frames turn with the cluster
        assert np.allclose(c1.features(a, b, F1), c2.features(a, b, F2), atol=1e-12)
    assert all(x.shape == (G.N_IN,) for x in [c1.features(0, 1, c1.frame(0, 1))])


def test_whole_cluster_frames_and_features_equal_the_per_block_ones():
    """Cluster.frames / features_batch (all blocks at once, environment sums per atom turned
    into each frame) against frame() / features() block by block, on (H2O)_6 with its
    pair blocks inside the 6 A cutoff."""
    syms, xyz = _cluster(3, n=6)
    cl = G.Cluster(syms, xyz)
    for t, (a, b) in cl.block_index().items():
        Q = cl.frames(a, b)
        X = cl.features_batch(a, b, Q)
        for i in range(len(a)):
            F = cl.frame(int(a[i]), int(b[i]))
            assert np.allclose(Q[i], F, atol=1e-14), (t, a[i], b[i])
            assert np.allclose(X[i], cl.features(int(a[i]), int(b[i]), F), atol=1e-12), (t, a[i], b[i])
    Q = cl.frames(np.array([0, 0]), np.array([0, 1]))
    for l in (1, 2, 3):
        Tb = G.cart_transform_batch(l, Q)
        for i in range(2):
            assert np.allclose(Tb[i], G.cart_transform(l, Q[i]), atol=1e-14)


def test_assembled_guess_is_covariant_and_symmetric():
    """Any network (here random weights through the numpy checker) gives a guess that is
    symmetric and turns exactly with the cluster."""
    syms, xyz = _cluster(11, n=3)
    Q = _rot(4)
    syms2, xyz2 = _cluster(11, n=3, Q=Q, shift=np.zeros(3))
    m1 = Molecule.from_angstrom(syms, xyz, "6-31G(d)")
    m2 = Molecule.from_angstrom(syms2, xyz2, "6-31G(d)")
    first, ls, nbf = G.atom_layout(m1)
    atom = {"O": np.eye(15) * 0.1, "H": np.eye(2) * 0.2}
    rng = np.random.default_rng(0)
    dims = {t: G.out_dim(t, 15, 2) for t in G.TYPES}
    params = {t: rng.standard_normal(
        G.N_IN * G.NH + G.NH + G.NH * G.NH + G.NH + G.NH * dims[t] + dims[t]) * 0.05 for t in G.TYPES}

    def predict(t, X):
        return R.forward(params[t], G.N_IN, G.NH, dims[t], X)[0]
    P1 = G.assemble(m1, atom, predict)
    P2 = G.assemble(m2, atom, predict)
    assert np.allclose(P1, P1.T, atol=1e-14)
    T = np.zeros((m1.n_bf, m1.n_bf))
    for a in range(m1.n_atom):
        T[first[a]:first[a] + nbf[a], first[a]:first[a] + nbf[a]] = G.atom_transform(ls[a], Q)
    assert np.max(np.abs(T.T @ P1 @ T - P2)) < 1e-9 * max(1.0, np.max(np.abs(P1)))


def test_training_rows_round_trip_through_assemble(oracle_mod):
    """Targets extracted from a density and fed back as predictions rebuild that density
    on every block within the cutoff (zero beyond it)."""
    syms, xyz = _cluster(2, n=2)
    mol = Molecule.from_angstrom(syms, xyz, "6-31G(d)")
    rng = np.random.default_rng(1)
    A = rng.standard_normal((mol.n_bf, mol.n_bf))
    P = A @ A.T
    atom = {"O": np.eye(15) * 0.3, "H": np.eye(2) * 0.1}
    rows = G.training_rows(mol, P, atom)
    feats = {t: {tuple(np.round(x, 12)): y for x, y in zip(*rows[t])} for t in rows}

    def predict(t, X):
        return np.array([feats[t][tuple(np.round(x, 12))] for x in X])
    Pg = G.assemble(mol, atom, predict)
    cl = G.Cluster(syms, xyz)
    first, ls, nbf = G.atom_layout(mol)
    for a in range(mol.n_atom):
        for b in range(mol.n_atom):
            blk = (slice(first[a], first[a] + nbf[a]), slice(first[b], first[b] + nbf[b]))
            if cl.D[a, b] < G.R_CUT:
                assert np.allclose(Pg[blk], P[blk], atol=1e-10)
            else:
                assert np.all(Pg[blk] == 0.0)


def test_mlp_checker_gradient_against_finite_differences():
    n_in, nh, n_out = 7, 5, 3
    rng = np.random.default_rng(3)
    p = rng.standard_normal(n_in * nh + nh + nh * nh + nh + nh * n_out + n_out) * 0.5
    X = rng.standard_normal((6, n_in))
    T = rng.standard_normal((6, n_out))
    Y, cache = R.forward(p, n_in, nh, n_out, X)
    loss, dY = R.mse(Y, T, 0.1)
    g = R.backward(p, n_in, nh, n_out, cache, dY)
    h = 1e-6
    num = np.empty_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        lp = R.mse(R.forward(p + e, n_in, nh, n_out, X)[0], T, 0.1)[0]
        lm = R.mse(R.forward(p - e, n_in, nh, n_out, X)[0], T, 0.1)[0]
        num[i] = (lp - lm) / (2 * h)
    assert np.max(np.abs(g - num)) < 1e-7 * max(1.0, np.max(np.abs(num)))


def test_adam_checker_closed_form():
    # first step of Adam moves every parameter by lr * sign(g) (up to eps)
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    p1, m, v = R.adam(p, g, np.zeros(3), np.zeros(3), 0.01, 0.9, 0.999, 1e-12, 1)
    assert np.allclose(p1, p - 0.01 * np.sign(g), atol=1e-9)

"""Oracle restatement of the J-only (make_r_J_engine) and K-only screening modes on the
CPU: they must reproduce the J (resp. K) of an unscreened build within the cutoff scale
and contribute nothing to the other matrix."""
import numpy as np

from tonto_amd.geometry import water_cluster


def test_j_only_k_only_modes(oracle_mod):
    mol = water_cluster(2, "6-31G(d)")
    o = oracle_mod.OracleSystem(mol)
    rng = np.random.default_rng(4)
    A = rng.standard_normal((mol.n_bf, mol.n_bf))
    P = A @ A.T / mol.n_bf
    J0, K0 = o.make_r_JK(P, schwarz=0.0, jp=0.0, kp=0.0)
    Jj, Kj = o.make_r_JK(P, jp=1e-20, kp=float("inf"))
    Jk, Kk = o.make_r_JK(P, jp=float("inf"), kp=1e-20)
    assert not Kj.any() and not Jk.any()
    assert np.max(np.abs(Jj - J0)) < 1e-7 and np.max(np.abs(Kk - K0)) < 1e-7

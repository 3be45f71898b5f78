"""The oracle's UHF J/K (make_u_JK_engine, molecule.fock.foo:2558-2712, per-quartet
digestion shell1quartet.foo:11530-11714).  Tonto's engine points the ket side's B-spin J
block at the A-spin one (JBcd => JA(cd), molecule.fock.foo:2637), so Ja and Jb are not
J(Pa) and J(Pb) separately, only their sum is J(Pa + Pb); Ka = K(Pa), Kb = K(Pb).
Pinned here by the reductions to RHF and, for the split itself, by a restatement from
the full ERI tensor in Tonto's shell-pair order (no Tonto output holds Ja and Jb
separately: the split is otherwise unpinned)."""
import numpy as np

from tonto_amd.basis import Molecule


def _system():
    return Molecule.from_angstrom(["O", "H", "H"], [[0, 0, 0], [0.76, 0.59, 0], [-0.76, 0.59, 0.1]],
                                  "6-31G(d)")


def _density(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    return A @ A.T / n


OFF = dict(schwarz=0.0, jp=0.0, kp=0.0, pair=1e-300)


def test_closed_shell_reduces_to_rhf(oracle_mod):
    m = _system()
    o = oracle_mod.OracleSystem(m)
    P = _density(m.n_bf, 1)
    J, K = o.make_r_JK(P, **OFF)
    Ja, Jb, Ka, Kb = o.make_u_JK(P / 2, P / 2, **OFF)
    s = np.max(np.abs(J))
    assert np.max(np.abs(Ja + Jb - J)) < 1e-13 * s
    assert np.max(np.abs(Ka - K / 2)) < 1e-13 * s and np.max(np.abs(Kb - K / 2)) < 1e-13 * s
    # the split is not symmetric between the spins
    assert np.max(np.abs(Ja - Jb)) > 1e-3 * s


def test_zero_beta_density_is_rhf_with_tonto_cutoffs(oracle_mod):
    m = _system()
    o = oracle_mod.OracleSystem(m)
    P = _density(m.n_bf, 2)
    J, K, st = o.make_r_JK(P, schwarz=1e-9, return_stats=True)
    Ja, Jb, Ka, Kb, su = o.make_u_JK(P, np.zeros_like(P), schwarz=1e-9, return_stats=True)
    assert st == su   # the same screening decisions
    # (the oracle's worker threads sum in a run-dependent order: equal to rounding)
    assert np.max(np.abs(Ja - J)) < 1e-13 * np.max(np.abs(J))
    assert np.max(np.abs(Ka - K)) < 1e-13 * np.max(np.abs(K))
    assert not Jb.any() and not Kb.any()


def _post_j(m, X):
    """Tonto's J post-processing (molecule.fock.foo:2688-2706): diagonal shell blocks x2,
    J = 2J, lower triangle reflected onto the upper."""
    J = X.copy()
    for sh in m.shells:
        f, n = sh.first_bf, sh.n_comp
        J[f:f + n, f:f + n] *= 2.0
    J *= 2.0
    il = np.tril_indices(J.shape[0], -1)
    J[il[1], il[0]] = J[il]
    return J


def test_alpha_beta_split_against_full_tensor(oracle_mod):
    m = _system()
    o = oracle_mod.OracleSystem(m)
    eri = o.full_eri()
    Pa, Pb = _density(m.n_bf, 3), _density(m.n_bf, 4)
    Ja, Jb, Ka, Kb = o.make_u_JK(Pa, Pb, **OFF)
    sh = m.shells
    blk = [slice(s.first_bf, s.first_bf + s.n_comp) for s in sh]
    pairs = [(a, b) for a in range(len(sh)) for b in range(a + 1)]   # Tonto's pair order
    XA = np.zeros_like(Pa)
    XB = np.zeros_like(Pa)
    for ab, (a, b) in enumerate(pairs):
        for cd in range(ab + 1):
            c, d = pairs[cd]
            f = (0.5 if a == b else 1.0) * (0.5 if c == d else 1.0) * (0.5 if ab == cd else 1.0)
            I = f * eri[blk[a], blk[b], blk[c], blk[d]]
            XA[blk[a], blk[b]] += np.einsum("ijkl,kl->ij", I, Pa[blk[c], blk[d]])
            XA[blk[c], blk[d]] += np.einsum("ijkl,ij->kl", I, Pa[blk[a], blk[b]])
            XB[blk[a], blk[b]] += np.einsum("ijkl,kl->ij", I, Pb[blk[c], blk[d]])
            # the B-spin ket side lands in JA, except for ab = cd
            tgt = XB if ab == cd else XA
            tgt[blk[c], blk[d]] += np.einsum("ijkl,ij->kl", I, Pb[blk[a], blk[b]])
    s = np.max(np.abs(Ja))
    assert np.max(np.abs(Ja - _post_j(m, XA))) < 1e-12 * s
    assert np.max(np.abs(Jb - _post_j(m, XB))) < 1e-12 * s
    K0a = np.einsum("prqs,rs->pq", eri, Pa)
    K0b = np.einsum("prqs,rs->pq", eri, Pb)
    assert np.max(np.abs(Ka - K0a)) < 1e-12 * np.max(np.abs(K0a))
    assert np.max(np.abs(Kb - K0b)) < 1e-12 * np.max(np.abs(K0b))


def test_uhf_scf_reaches_g09_cation(oracle_mod):
    """H2O+ UHF/cc-pVDZ (6D), Tonto's water geometry: g09 -75.6327038760
    (scripts/check_dft_reference.py:72; Tonto agreed to 2.0e-10 there).  The oracle's
    UHF J/K (make_u_JK_engine's split included) in tonto_amd.scf.run_uhf."""
    import torch
    from tonto_amd.geometry import tonto_water
    from tonto_amd.scf import run_uhf
    m = tonto_water("cc-pVDZ", charge=1)
    assert m.n_electrons == 9
    o = oracle_mod.OracleSystem(m)
    S, T, V = o.one_electron()
    mi = o.schwarz()

    def jk(Pa, Pb):
        return tuple(torch.from_numpy(x) for x in o.make_u_JK(Pa.numpy(), Pb.numpy(), max_I=mi))
    r = run_uhf(torch.from_numpy(T + V), torch.from_numpy(S), 5, 4, m.nuclear_repulsion(), jk,
                convergence=1e-10, diis_tol=1e-7)
    assert r.converged
    assert abs(r.energy - (-75.6327038760)) < 1e-9
    # <S^2> of a UHF doublet is slightly above 3/4 (spin contamination), and the alpha and
    # beta densities hold 5 and 4 electrons
    Sm = torch.from_numpy(S)
    assert abs(float(torch.sum(r.Pa * Sm)) - 5) < 1e-10 and abs(float(torch.sum(r.Pb * Sm)) - 4) < 1e-10

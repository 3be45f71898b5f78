"""GPU parity: libtonto_gpu (through its C ABI) against the CPU oracle on the same
inputs.  Bars: max|dF| <= 1e-10 Eh with F = h + J - K/2 (so J and K to <= 1e-10 each
on the scale used), the screening decisions identical (same number of quartets
computed), the converged RHF energy within 1e-9 Eh of the oracle's."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible: GPU tests must run on an MI355X")
    return torch


def _rand_density(n, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    return scale * A @ A.T / n


def _engine(mol, **kw):
    from tonto_amd.gpu import JKEngine
    return JKEngine(mol, **kw)


def _check_jk(oracle_mod, mol, P, cut=1e-9, tol=1e-10, mixed=False):
    o = oracle_mod.OracleSystem(mol)
    e = _engine(mol, schwarz=cut)
    mi = e.schwarz()
    mo = o.schwarz()
    assert np.max(np.abs(mi - mo) / np.maximum(mo, 1e-300)) < 1e-13
    J, K = e.make_r_JK(P, mixed=mixed)
    st = e.stats()
    Jo, Ko, so = o.make_r_JK(P, schwarz=cut, max_I=mi, return_stats=True)
    dJ, dK = np.max(np.abs(J - Jo)), np.max(np.abs(K - Ko))
    dF = np.max(np.abs((J - 0.5 * K) - (Jo - 0.5 * Ko)))
    print(f"{mol.n_bf} bf: dJ={dJ:.2e} dK={dK:.2e} dF={dF:.2e} quartets gpu={st['quartets']} "
          f"oracle={so['quartets']}")
    assert st["quartets"] == so["quartets"]
    assert st["prim_quartets"] == so["prim_quartets"]
    assert dF <= tol and dJ <= tol and dK <= tol
    assert np.array_equal(J, J.T) and np.array_equal(K, K.T)
    return e, o


@pytest.mark.parametrize("basis", ["STO-3G", "6-31G(d)", "6-311G(d,p)", "cc-pVDZ", "DZP"])
def test_water_all_classes(torch_cuda, oracle_mod, basis):
    from tonto_amd.geometry import tonto_water
    mol = tonto_water(basis)
    _check_jk(oracle_mod, mol, _rand_density(mol.n_bf, 1))


@pytest.mark.parametrize("budget", [None, 16690, 24032, 26450])
def test_lds_row_form_full_and_partial_rows(torch_cuda, oracle_mod, budget):
    """The lane-per-quartet kernels' LDS K-row form on (H2O)_8 / 6-31G(d) (152 functions;
    56 s-type, 48 p-type, 48 d-type), made to run for every ket segment length.  A row
    holds the columns of the ket shells' momenta only (56 doubles for an (..|ss) class).
    LDS caps: 16690 B leaves room next to the 14848 B two-root Rys table and ket queue for
    4 rows of 56: (pp|ss) keeps 4 of 6 rows, (ds|ss) 4 of 7, (ps|ps) (104 columns) 2 of 4 in
    LDS, the rest go to global atomics; 26450 B, next to the 22240 B of the three-root
    classes: (dp|ps) (104 columns) keeps 5 of 9 rows, (dd|ss) 9 of 12, (fp|ss) 9 of 13,
    (fd|ss) 9 of 16; 24032 B: 4 rows of 56 next to the three-root classes, fewer than half
    of them: (dd|ss) 4 of 12, (fp|ss) 4 of 13, (fd|ss) 4 of 16 (the form needs a quarter).
    Uncapped, the (..|ss) class pairs also read their density rows from LDS (every bra
    pair: the segment floor is 0)."""
    import os
    from tonto_amd.geometry import water_cluster
    g = np.load(os.path.join(HERE, "golden", "w8_631gd.npz"))
    mol = water_cluster(8, "6-31G(d)", seed=int(g["seed"]))
    old = {k: os.environ.get(k) for k in ("TGPU_KROW_BUDGET", "TGPU_KROW_MIN_KETS")}
    os.environ["TGPU_KROW_MIN_KETS"] = "0"
    if budget:
        os.environ["TGPU_KROW_BUDGET"] = str(budget)
    try:
        _check_jk(oracle_mod, mol, g["P"], cut=1e-12)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("n,seed", [(2, 0), (4, 2)])
def test_lds_row_form_every_lane_class(torch_cuda, oracle_mod, mixed, n, seed):
    """Every class pair that has the lane-per-quartet kernel, run in its LDS K-row form one
    at a time (the others in the wavefront form; (ff|ss) has only the global-atomic lane
    form), on (H2O)_2 and (H2O)_4 / cc-pVTZ (130 and 260 functions, all rows in LDS): J
    and K against the oracle to 1e-10 (FP64) and to the mixed-precision bound 1e-8, the
    same quartets."""
    from tonto_amd.build import class_list
    from tonto_amd.flops import class_pairs
    from tonto_amd.geometry import water_cluster
    mol = water_cluster(n, "cc-pVTZ", seed=seed)
    P = _rand_density(mol.n_bf, 31)
    old = os.environ.get("TGPU_KROW_MIN_KETS")
    os.environ["TGPU_KROW_MIN_KETS"] = "0"
    try:
        e = _engine(mol, schwarz=1e-11)
    finally:
        if old is None:
            os.environ.pop("TGPU_KROW_MIN_KETS", None)
        else:
            os.environ["TGPU_KROW_MIN_KETS"] = old
    mi = e.schwarz()
    o = oracle_mod.OracleSystem(mol)
    Jo, Ko, so = o.make_r_JK(P, schwarz=1e-11, max_I=mi, return_stats=True)
    heavy = {tuple(c[:4]): c[4] for c in class_list()}
    cps = class_pairs()
    light = [k for k, cp in enumerate(cps) if not heavy[tuple(cp)]]
    assert len(light) == 28
    tol = 1e-8 if mixed else 1e-10
    for k in light:
        for kk in light:
            e.set_variant(kk, 1)
        e.set_variant(k, 0)
        J, K = e.make_r_JK(P, mixed=mixed)
        st = e.stats()
        dJ, dK = np.max(np.abs(J - Jo)), np.max(np.abs(K - Ko))
        print(f"class {cps[k]} mixed={mixed}: dJ={dJ:.2e} dK={dK:.2e}")
        assert st["quartets"] == so["quartets"]
        assert dJ <= tol and dK <= tol, cps[k]


def test_f_shells_water_cc_pvtz(torch_cuda, oracle_mod):
    """cc-pVTZ water: f on O, d on H; (ff|ff) and every other f class pair, including
    the classes whose transfer buffers live in the global workspace."""
    from tonto_amd.geometry import tonto_water
    mol = tonto_water("cc-pVTZ")
    _check_jk(oracle_mod, mol, _rand_density(mol.n_bf, 11))


def test_f_shells_water_dimer_cc_pvtz_skip_path(torch_cuda, oracle_mod):
    """(H2O)_2 / cc-pVTZ: f-shell quartets over four distinct atoms (skipped primitive-pair
    lists live) at the tight Schwarz cutoff of the (H2O)_64 workload."""
    from tonto_amd.geometry import water_cluster
    mol = water_cluster(2, "cc-pVTZ")
    _check_jk(oracle_mod, mol, _rand_density(mol.n_bf, 12, 0.5), cut=1e-11)


def test_four_distinct_atoms_skip_path(torch_cuda, oracle_mod):
    from tonto_amd.geometry import cn_hcn_dimer
    mol = cn_hcn_dimer()
    _check_jk(oracle_mod, mol, _rand_density(mol.n_bf, 2, 0.2))


def test_water_dimer_random_density_zero_cutoffs(torch_cuda, oracle_mod):
    from tonto_amd.geometry import water_cluster
    mol = water_cluster(2, "6-31G(d,p)")
    _check_jk(oracle_mod, mol, _rand_density(mol.n_bf, 3), cut=0.0)


def test_config1_w8_converged_density(torch_cuda, oracle_mod):
    """(H2O)_8 / 6-31G(d), 152 Cartesian functions, Schwarz cutoff 1e-12, at the
    converged density (tests/golden/w8_631gd.npz)."""
    from tonto_amd.geometry import water_cluster
    g = np.load(os.path.join(HERE, "golden", "w8_631gd.npz"))
    mol = water_cluster(8, "6-31G(d)", seed=int(g["seed"]))
    assert mol.n_bf == 152
    _check_jk(oracle_mod, mol, g["P"], cut=1e-12)


def test_w4_6311gdp_converged_density(torch_cuda, oracle_mod):
    from tonto_amd.geometry import water_cluster
    g = np.load(os.path.join(HERE, "golden", "w4_6311gdp.npz"))
    mol = water_cluster(4, "6-311G(d,p)", seed=int(g["seed"]))
    _check_jk(oracle_mod, mol, g["P"], cut=1e-9)


def test_zero_density_gives_zero(torch_cuda):
    from tonto_amd.geometry import tonto_water
    mol = tonto_water("6-31G(d)")
    e = _engine(mol)
    J, K = e.make_r_JK(np.zeros((mol.n_bf, mol.n_bf)))
    assert not J.any() and not K.any()
    assert e.stats()["quartets"] == 0


def test_single_atom_and_single_shell(torch_cuda, oracle_mod):
    from tonto_amd.basis import Molecule
    for mol in (Molecule.from_angstrom(["O"], [[0, 0, 0]], "cc-pVDZ"),
                Molecule.from_angstrom(["H"], [[0, 0, 0]], "STO-3G")):
        _check_jk(oracle_mod, mol, _rand_density(mol.n_bf, 4))


def test_mixed_precision_error_is_small(torch_cuda, oracle_mod):
    from tonto_amd.geometry import water_cluster
    g = np.load(os.path.join(HERE, "golden", "w8_631gd.npz"))
    mol = water_cluster(8, "6-31G(d)", seed=int(g["seed"]))
    e = _engine(mol, schwarz=1e-12, fp32_threshold=1e-6)
    e.schwarz()
    J, K = e.make_r_JK(g["P"])
    Jm, Km = e.make_r_JK(g["P"], mixed=True)
    st = e.stats()
    assert st["fp32_quartets"] > 0
    dF = np.max(np.abs((Jm - 0.5 * Km) - (J - 0.5 * K)))
    print("mixed dF", dF, st)
    assert dF < 1e-7


def test_one_electron_parity(torch_cuda, oracle_mod):
    from tonto_amd.geometry import water_cluster
    mol = water_cluster(2, "6-311G(d,p)")
    e = _engine(mol)
    S, T, V = e.one_electron()
    So, To, Vo = oracle_mod.OracleSystem(mol).one_electron()
    assert np.max(np.abs(S - So)) < 1e-12 and np.max(np.abs(T - To)) < 1e-11
    assert np.max(np.abs(V - Vo)) < 1e-11


def test_device_entry_point_matches_host(torch_cuda):
    import torch
    from tonto_amd.geometry import water_cluster
    mol = water_cluster(3, "6-31G(d)")
    e = _engine(mol)
    P = _rand_density(mol.n_bf, 5)
    J, K = e.make_r_JK(P)
    dP = torch.tensor(P, device="cuda")
    dJ = torch.empty_like(dP)
    dK = torch.empty_like(dP)
    e.set_stream(torch.cuda.current_stream().cuda_stream)
    e.make_r_JK_device(dP.data_ptr(), dJ.data_ptr(), dK.data_ptr())
    torch.cuda.synchronize()
    assert np.max(np.abs(dJ.cpu().numpy() - J)) < 1e-13 and np.max(np.abs(dK.cpu().numpy() - K)) < 1e-13


def test_gpu_scf_reaches_g09(torch_cuda):
    """H2O RHF/cc-pVDZ on the GPU path: g09 -76.0232731206 (scripts/check_dft_reference.py:67)."""
    from tonto_amd.geometry import tonto_water
    from tonto_amd.driver import GPUSCF
    r = GPUSCF(tonto_water("cc-pVDZ")).run(convergence=1e-10, diis_tol=1e-7)
    assert r.converged and abs(r.energy - (-76.0232731206)) < 1e-9


def test_mixed_incremental_scf_energy(torch_cuda):
    """FP32 small-bound quartets, loose-cutoff first builds and incremental
    (delta-density) builds converge to the FP64 energy within 1e-8 Eh on
    (H2O)8 / 6-31G(d,p)."""
    from tonto_amd.geometry import water_cluster
    from tonto_amd.driver import GPUSCF
    scf = GPUSCF(water_cluster(8, "6-31G(d,p)", seed=1), schwarz=1e-11, fp32_threshold=1e-6)
    r64 = scf.run(precision="fp64", convergence=1e-9, diis_tol=1e-6)
    rmx = scf.run(precision="mixed", mixed_until=1e-5, incremental=True, rebuild_every=8,
                  convergence=1e-9, diis_tol=1e-6)
    print(r64.energy, rmx.energy - r64.energy, rmx.n_builds)
    assert r64.converged and rmx.converged
    assert abs(rmx.energy - r64.energy) <= 1e-8
    assert rmx.n_builds["delta"] + rmx.n_builds["delta_mixed"] > 0


@pytest.mark.parametrize("basis,cut,krows", [("6-31G(d)", 1e-9, False), ("6-31G(d)", 1e-12, True),
                                              ("cc-pVTZ", 1e-11, False)])
def test_uhf_against_oracle(torch_cuda, oracle_mod, basis, cut, krows):
    """UHF J/K (make_u_JK_engine) against the oracle: Ja, Jb, Ka, Kb to 1e-10 and the
    same quartets, with two different random densities; cc-pVTZ runs the f-shell
    wavefront classes, krows the lane kernels' LDS K-row form."""
    from tonto_amd.geometry import water_cluster
    mol = water_cluster(2 if basis == "cc-pVTZ" else 4, basis, seed=5)
    old = os.environ.get("TGPU_KROW_MIN_KETS")
    if krows:
        os.environ["TGPU_KROW_MIN_KETS"] = "0"
    try:
        e = _engine(mol, schwarz=cut)
    finally:
        if old is None:
            os.environ.pop("TGPU_KROW_MIN_KETS", None)
        else:
            os.environ["TGPU_KROW_MIN_KETS"] = old
    mi = e.schwarz()
    o = oracle_mod.OracleSystem(mol)
    Pa = _rand_density(mol.n_bf, 21)
    Pb = _rand_density(mol.n_bf, 22, 0.7)
    got = e.make_u_JK(Pa, Pb)
    st = e.stats()
    *ref, so = o.make_u_JK(Pa, Pb, schwarz=cut, max_I=mi, return_stats=True)
    d = [float(np.max(np.abs(x - y))) for x, y in zip(got, ref)]
    print(f"{basis} n_bf={mol.n_bf}: dJa dJb dKa dKb = {d}, quartets gpu={st['quartets']} oracle={so['quartets']}")
    assert st["quartets"] == so["quartets"] and st["prim_quartets"] == so["prim_quartets"]
    assert max(d) <= 1e-10


def _uneven_batch(n_mol=6, basis="6-31G(d,p)", seed=7):
    from tonto_amd.geometry import water_cluster
    rng = np.random.default_rng(seed)
    sizes = rng.integers(1, 6, size=n_mol)
    return [water_cluster(int(k), basis, seed=int(seed + i)) for i, k in enumerate(sizes)]


def test_batch_matches_oracle_per_molecule(torch_cuda, oracle_mod):
    mols = _uneven_batch()
    Ps = [_rand_density(m.n_bf, 10 + i, 0.3) for i, m in enumerate(mols)]
    e = _engine(mols)
    mi = e.schwarz()
    J, K = e.make_r_JK(Ps)
    tot = 0
    for i, m in enumerate(mols):
        o = oracle_mod.OracleSystem(m)
        mo = o.schwarz()
        seg = mi[e.pair_off[i]:e.pair_off[i + 1]]
        assert np.max(np.abs(seg - mo) / np.maximum(mo, 1e-300)) < 1e-13
        Jo, Ko, so = o.make_r_JK(Ps[i], max_I=seg, return_stats=True)
        dF = np.max(np.abs((J[i] - 0.5 * K[i]) - (Jo - 0.5 * Ko)))
        assert dF <= 1e-10, (i, dF)
        tot += so["quartets"]
    assert e.stats()["quartets"] == tot


def test_batch_equals_single_builds(torch_cuda):
    mols = _uneven_batch(4, "6-31G(d)", 3)
    Ps = [_rand_density(m.n_bf, 20 + i, 0.3) for i, m in enumerate(mols)]
    e = _engine(mols)
    J, K = e.make_r_JK(Ps)
    S, T, V = e.one_electron()
    for i, m in enumerate(mols):
        s = _engine(m)
        J1, K1 = s.make_r_JK(Ps[i])
        S1, T1, V1 = s.one_electron()
        assert np.max(np.abs(J[i] - J1)) < 1e-12 and np.max(np.abs(K[i] - K1)) < 1e-12
        assert np.max(np.abs(S[i] - S1)) < 1e-14 and np.max(np.abs(V[i] - V1)) < 1e-12


def test_graph_replay_matches_eager(torch_cuda):
    """Trap check: the captured build is replayed several times with eager work and
    synchronisations in between, and every replay is compared with the eager build."""
    import torch
    from tonto_amd.geometry import water_cluster
    mol = water_cluster(3, "6-31G(d)")
    e = _engine(mol)
    P = torch.tensor(_rand_density(mol.n_bf, 6), device="cuda")
    J0, K0 = (torch.empty_like(P) for _ in range(2))
    J1, K1 = (torch.empty_like(P) for _ in range(2))
    e.set_stream(torch.cuda.current_stream().cuda_stream)
    e.make_r_JK_device(P.data_ptr(), J0.data_ptr(), K0.data_ptr())
    torch.cuda.synchronize()
    q0 = e.stats()["quartets"]
    for rep in range(4):
        J1.fill_(float("nan"))
        K1.fill_(float("nan"))
        e.make_r_JK_device(P.data_ptr(), J1.data_ptr(), K1.data_ptr(), graph=True)
        torch.cuda.synchronize()
        assert torch.max(torch.abs(J1 - J0)).item() < 1e-12
        assert torch.max(torch.abs(K1 - K0)).item() < 1e-12
        assert e.stats()["quartets"] == q0
        # eager work on the same context between replays
        e.make_r_JK_device(P.data_ptr(), J0.data_ptr(), K0.data_ptr())
        torch.cuda.synchronize()


def test_autotune_keeps_results(torch_cuda):
    import torch
    from tonto_amd.geometry import water_cluster
    g = np.load(os.path.join(HERE, "golden", "w8_631gd.npz"))
    mol = water_cluster(8, "6-31G(d)", seed=int(g["seed"]))
    e = _engine(mol, schwarz=1e-12)
    J0, K0 = e.make_r_JK(g["P"])
    P = torch.tensor(g["P"], device="cuda")
    ms, var = e.autotune(P.data_ptr())
    chunks = e.ket_chunks()
    print("autotune variants", var, "ket chunks", chunks)
    assert all(v in (0, 1) for v in var)
    assert all(c in (0, 16, 32, 64) for c in chunks)
    assert all(c == 0 for v, c in zip(var, chunks) if v == 0)
    J1, K1 = e.make_r_JK(g["P"])
    assert np.max(np.abs(J1 - J0)) < 1e-12 and np.max(np.abs(K1 - K0)) < 1e-12
    # every light class pair also through the wavefront form, with every ket chunk
    for k in range(len(var)):
        e.set_variant(k, 1)
    for chunk in (16, 32, 64, 0):
        for k in range(len(var)):
            e.set_ket_chunk(k, chunk)
        J2, K2 = e.make_r_JK(g["P"])
        assert np.max(np.abs(J2 - J0)) < 1e-12 and np.max(np.abs(K2 - K0)) < 1e-12, chunk


@pytest.mark.parametrize("only", ["J", "K"])
def test_j_only_and_k_only_builds(torch_cuda, oracle_mod, only):
    """make_r_J_engine (molecule.fock.foo:1638-1757, Schwarz_J_test 3538-3560) and the K-only
    build (Schwarz_K_test): the density test has a 1e-20 floor, the other matrix is absent."""
    from tonto_amd.geometry import water_cluster
    g = np.load(os.path.join(HERE, "golden", "w4_6311gdp.npz"))
    mol = water_cluster(4, "6-311G(d,p)", seed=int(g["seed"]))
    e = _engine(mol)
    mi = e.schwarz()
    J, K = e.make_r_JK(g["P"], only=only)
    o = oracle_mod.OracleSystem(mol)
    jp, kp = (1e-20, float("inf")) if only == "J" else (float("inf"), 1e-20)
    Jo, Ko, so = o.make_r_JK(g["P"], jp=jp, kp=kp, max_I=mi, return_stats=True)
    assert e.stats()["quartets"] == so["quartets"]
    if only == "J":
        assert not K.any() and np.max(np.abs(J - Jo)) <= 1e-10
    else:
        assert not J.any() and np.max(np.abs(K - Ko)) <= 1e-10


def test_uhf_mixed_precision(torch_cuda, oracle_mod):
    """The UHF build with TGPU_JK_MIXED on (H2O)4 / 6-31G(d) at 1e-12: every pass runs its
    small-bound quartets in FP32; the Fock-relevant combinations Ja + Jb - Ka and
    Ja + Jb - Kb stay within 1e-7 of the FP64 UHF build and the quartet set is the same."""
    from tonto_amd.geometry import water_cluster
    mol = water_cluster(4, "6-31G(d)", seed=5)
    e = _engine(mol, schwarz=1e-12, fp32_threshold=1e-6)
    e.schwarz()
    Pa = _rand_density(mol.n_bf, 23, 0.5)
    Pb = _rand_density(mol.n_bf, 24, 0.4)
    Ja, Jb, Ka, Kb = e.make_u_JK(Pa, Pb)
    q64 = e.stats()
    Jam, Jbm, Kam, Kbm = e.make_u_JK(Pa, Pb, mixed=True)
    qmx = e.stats()
    assert qmx["fp32_quartets"] > 0 and qmx["quartets"] == q64["quartets"]
    dFa = np.max(np.abs((Jam + Jbm - Kam) - (Ja + Jb - Ka)))
    dFb = np.max(np.abs((Jam + Jbm - Kbm) - (Ja + Jb - Kb)))
    print("UHF mixed dF", dFa, dFb, qmx)
    assert dFa < 1e-7 and dFb < 1e-7


@pytest.mark.parametrize("nw,basis,cut", [(1, "cc-pVTZ", 1e-9), (2, "cc-pVTZ", 1e-11), (3, "6-31G(d)", 1e-10)])
def test_spherical_direct_build_against_oracle(torch_cuda, oracle_mod, nw, basis, cut):
    """Spherical basis (make_r_JK_direct): the GPU's Cartesian kernels on T P T^T with
    J, K turned back into solid harmonics, against the oracle's per-quartet solid-harmonic
    integrals digested with the spherical density; spherical Schwarz table to 1e-12
    relative (the two sum the solid-harmonic transform in different orders), the same
    quartets (Schwarz_inequality_test, both on the GPU's table), J and K to 1e-10."""
    from tonto_amd.geometry import tonto_water, water_cluster
    mol = tonto_water(basis) if nw == 1 else water_cluster(nw, basis, seed=9)
    e = _engine(mol, schwarz=cut, spherical=True)
    o = oracle_mod.OracleSystem(mol)
    assert e.n_bf == o.n_sph
    mi = e.schwarz()
    mo = o.schwarz_sph()
    assert np.max(np.abs(mi - mo) / np.maximum(mo, 1e-300)) < 1e-12
    P = _rand_density(e.n_bf, 31, 0.5)
    J, K = e.make_r_JK(P)
    st = e.stats()
    Jo, Ko, so = o.make_r_JK_direct(P, schwarz=cut, max_I=mi, return_stats=True)
    dJ, dK = np.max(np.abs(J - Jo)), np.max(np.abs(K - Ko))
    print(f"spherical {basis} n={e.n_bf}: dJ={dJ:.2e} dK={dK:.2e} quartets gpu={st['quartets']} oracle={so['quartets']}")
    assert st["quartets"] == so["quartets"] and st["prim_quartets"] == so["prim_quartets"]
    assert dJ <= 1e-10 and dK <= 1e-10


def test_spherical_scf_cc_pvtz_water(torch_cuda):
    """H2O RHF/cc-pVTZ in the spherical basis on the GPU path (58 functions): Tonto's
    tests/short/h2o_rhf_cc-pVTZ_spherical_harmonic_basis/stdout:204 prints -76.0524."""
    from tonto_amd.driver import GPUSCF
    from tonto_amd.geometry import tonto_water
    scf = GPUSCF(tonto_water("cc-pVTZ"), spherical=True)
    assert scf.engine.n_bf == 58
    r = scf.run(convergence=1e-9, diis_tol=1e-6)
    print(f"spherical cc-pVTZ water: E = {r.energy:.10f}")
    assert r.converged and round(r.energy, 4) == -76.0524

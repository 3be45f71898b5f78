"""The workloads' configurations on the GPU path, each against the oracle.

* (H2O)8 / 6-31G(d), Schwarz 1e-12: a GPU RHF whose every Fock build is checked against
  the oracle's on the same density (max |dF| <= 1e-10 Eh per iteration), converged energy
  within 1e-9 Eh of the oracle's SCF (tests/golden/w8_631gd.npz).
* 6-311G(d,p), FP32 below Schwarz bound 1e-6: the mixed build against the oracle's FP64
  build at the converged density of (H2O)32 (992 functions) and (H2O)4, with max |dF|
  reported; the mixed/incremental GPU SCF of (H2O)4 within 1e-8 Eh of the oracle's
  energy (tests/golden/w4_6311gdp.npz), per-iteration dF of its accumulated Fock matrix
  against the oracle's on the same density reported.
* cc-pVTZ (f shells), Schwarz 1e-11: the mixed build of (H2O)2 against the oracle's FP64
  build, and the mixed SCF of (H2O)4 within 1e-8 Eh of the FP64 GPU SCF.
"""
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


def _dF(J, K, Jo, Ko):
    return float(np.max(np.abs((J - 0.5 * K) - (Jo - 0.5 * Ko))))


def test_config1_every_iteration_against_oracle(torch_cuda, oracle_mod):
    from tonto_amd.driver import GPUSCF
    from tonto_amd.geometry import water_cluster
    from tonto_amd.scf import run_rhf
    g = np.load(os.path.join(HERE, "golden", "w8_631gd.npz"))
    mol = water_cluster(8, "6-31G(d)", seed=int(g["seed"]))
    assert mol.n_bf == 152
    scf = GPUSCF(mol, schwarz=1e-12)
    o = oracle_mod.OracleSystem(mol)
    dFs = []

    def jk(P, mode):
        J, K = scf.jk(P, mode)
        Jo, Ko = o.make_r_JK(P.cpu().numpy(), schwarz=1e-12, max_I=scf.max_I)
        dFs.append(_dF(J.cpu().numpy(), K.cpu().numpy(), Jo, Ko))
        return J, K
    r = run_rhf(scf.h, scf.S, mol.n_electrons, mol.nuclear_repulsion(), jk, P0=scf.guess(),
                convergence=1e-10, diis_tol=1e-7)
    print("per-iteration max|dF|:", " ".join(f"{x:.1e}" for x in dFs))
    print(f"E = {r.energy:.10f}, golden {float(g['energy']):.10f}, {r.iterations} iterations")
    assert r.converged and len(dFs) == r.iterations + 1
    assert max(dFs) <= 1e-10
    assert abs(r.energy - float(g["energy"])) <= 1e-9


def test_config2_mixed_build_992_functions(torch_cuda, oracle_mod):
    """(H2O)32 / 6-311G(d,p) at 992 Cartesian functions: the mixed build at the GPU's
    converged FP64 density against the oracle's FP64 build (same screening)."""
    from tonto_amd.driver import GPUSCF
    from tonto_amd.geometry import water_cluster
    mol = water_cluster(32, "6-311G(d,p)", seed=0)
    assert mol.n_bf == 992
    scf = GPUSCF(mol, schwarz=1e-9, fp32_threshold=1e-6)
    r = scf.run(precision="fp64", convergence=1e-8, diis_tol=1e-5)
    P = r.P.cpu().numpy()
    J, K = scf.engine.make_r_JK(P, mixed=True)
    st = scf.engine.stats()
    Jo, Ko, so = oracle_mod.OracleSystem(mol).make_r_JK(P, schwarz=1e-9, max_I=scf.max_I, return_stats=True)
    dF = _dF(J, K, Jo, Ko)
    print(f"(H2O)32 mixed vs oracle FP64: max|dF| = {dF:.2e}, {st['fp32_quartets']} of "
          f"{st['quartets']} quartets in FP32")
    assert st["quartets"] == so["quartets"] and st["fp32_quartets"] > 0
    assert dF <= 1e-6


def test_config2_mixed_incremental_scf(torch_cuda, oracle_mod):
    """(H2O)4 / 6-311G(d,p), Schwarz 1e-9: the mixed/incremental SCF's energy against the
    oracle's SCF (golden), and its accumulated Fock matrix against the oracle's at every
    iteration's density."""
    from tonto_amd.driver import GPUSCF
    from tonto_amd.geometry import water_cluster
    g = np.load(os.path.join(HERE, "golden", "w4_6311gdp.npz"))
    mol = water_cluster(4, "6-311G(d,p)", seed=int(g["seed"]))
    scf = GPUSCF(mol, schwarz=1e-9, fp32_threshold=1e-6)
    J, K = scf.engine.make_r_JK(g["P"], mixed=True)
    o = oracle_mod.OracleSystem(mol)
    Jo, Ko = o.make_r_JK(g["P"], schwarz=1e-9, max_I=scf.max_I)
    print(f"(H2O)4 mixed build at the converged density: max|dF| = {_dF(J, K, Jo, Ko):.2e}")
    assert _dF(J, K, Jo, Ko) <= 1e-7
    r = scf.run(precision="mixed", mixed_until=1e-5, incremental=True, rebuild_every=8,
                convergence=1e-9, diis_tol=1e-6)
    print(f"mixed SCF E = {r.energy:.10f}, golden {float(g['energy']):.10f}, builds {r.n_builds}")
    assert r.converged and abs(r.energy - float(g["energy"])) <= 1e-8
    # per-iteration dF of the accumulated Fock matrix (FP32 quartets, loose and delta
    # builds included) against the oracle's FP64 F(P) on the same density
    h = scf.h.cpu().numpy()
    rr = scf.run(precision="mixed", mixed_until=1e-5, incremental=True, rebuild_every=8,
                 convergence=1e-9, diis_tol=1e-6, keep_fock=True)
    dFs = []
    for it in rr.history:
        if "P" not in it:
            continue
        Jo, Ko = o.make_r_JK(it["P"], schwarz=1e-9, max_I=scf.max_I)
        dFs.append((it["mode"], float(np.max(np.abs(it["F"] - (h + Jo - 0.5 * Ko))))))
    print("per-iteration (mode, max|dF|):", ", ".join(f"{m} {d:.1e}" for m, d in dFs))
    assert dFs and dFs[-1][1] <= 1e-8


def test_config5_mixed_build_cc_pvtz(torch_cuda, oracle_mod):
    """(H2O)2 / cc-pVTZ at Schwarz 1e-11, FP32 below bound 1e-6, against the oracle's FP64
    build on the same density."""
    from tonto_amd.geometry import water_cluster
    from tonto_amd.gpu import JKEngine
    mol = water_cluster(2, "cc-pVTZ", seed=0)
    e = JKEngine(mol, schwarz=1e-11, fp32_threshold=1e-6)
    mi = e.schwarz()
    rng = np.random.default_rng(41)
    A = rng.standard_normal((mol.n_bf, mol.n_bf))
    P = 0.5 * A @ A.T / mol.n_bf
    J, K = e.make_r_JK(P, mixed=True)
    st = e.stats()
    Jo, Ko, so = oracle_mod.OracleSystem(mol).make_r_JK(P, schwarz=1e-11, max_I=mi, return_stats=True)
    dF = _dF(J, K, Jo, Ko)
    print(f"(H2O)2 cc-pVTZ mixed: max|dF| = {dF:.2e}, {st['fp32_quartets']} of {st['quartets']} in FP32")
    assert st["quartets"] == so["quartets"] and st["fp32_quartets"] > 0
    assert dF <= 1e-7


def test_config5_mixed_scf_cc_pvtz(torch_cuda):
    """(H2O)4 / cc-pVTZ at Schwarz 1e-11: the mixed SCF (loose first builds, FP32, delta
    builds, FP64 at the end) within 1e-8 Eh of the FP64 GPU SCF."""
    from tonto_amd.driver import GPUSCF
    from tonto_amd.geometry import water_cluster
    scf = GPUSCF(water_cluster(4, "cc-pVTZ", seed=0), schwarz=1e-11, fp32_threshold=1e-6)
    r64 = scf.run(precision="fp64", convergence=1e-9, diis_tol=1e-6)
    rmx = scf.run(precision="mixed", convergence=1e-9, diis_tol=1e-6)
    print(f"FP64 {r64.energy:.10f} ({r64.iterations} it), mixed dE {rmx.energy - r64.energy:+.2e} "
          f"({rmx.iterations} it, builds {rmx.n_builds})")
    assert r64.converged and rmx.converged
    assert abs(rmx.energy - r64.energy) <= 1e-8


def test_uhf_scf_cation_reaches_g09(torch_cuda):
    """H2O+ UHF/cc-pVDZ (6D) on the GPU path (make_u_JK_engine builds on the device, the
    UHF loop in torch): g09 -75.6327038760 (scripts/check_dft_reference.py:72)."""
    from tonto_amd.driver import GPUSCF
    from tonto_amd.geometry import tonto_water
    r = GPUSCF(tonto_water("cc-pVDZ", charge=1)).run_uhf(multiplicity=2, convergence=1e-10, diis_tol=1e-7)
    print(f"GPU UHF H2O+ E = {r.energy:.10f} ({r.iterations} iterations)")
    assert r.converged and abs(r.energy - (-75.6327038760)) < 1e-9


def test_uhf_scf_cluster_against_oracle(torch_cuda, oracle_mod):
    """(H2O)2+ UHF/6-31G(d) at Schwarz 1e-10: the GPU UHF SCF's energy against the same
    SCF driven by the oracle's UHF J/K (quartets over four distinct atoms, skip lists live)."""
    import torch
    from tonto_amd.driver import GPUSCF
    from tonto_amd.geometry import water_cluster
    from tonto_amd.scf import run_uhf
    mol = water_cluster(2, "6-31G(d)", seed=3)
    mol.charge = 1
    scf = GPUSCF(mol, schwarz=1e-10)
    r = scf.run_uhf(multiplicity=2, guess="core", convergence=1e-10, diis_tol=1e-7)
    o = oracle_mod.OracleSystem(mol)

    def jk(Pa, Pb):
        return tuple(torch.from_numpy(x) for x in o.make_u_JK(Pa.numpy(), Pb.numpy(), schwarz=1e-10,
                                                             max_I=scf.max_I))
    ro = run_uhf(scf.h.cpu(), scf.S.cpu(), 10, 9, mol.nuclear_repulsion(), jk, convergence=1e-10, diis_tol=1e-7)
    print(f"(H2O)2+ UHF: GPU {r.energy:.10f} oracle {ro.energy:.10f}")
    assert r.converged and ro.converged and abs(r.energy - ro.energy) < 1e-9

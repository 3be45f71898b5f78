"""The mixed/incremental SCF schedule of tonto_amd.scf.run_rhf (host logic, CPU): which
builds it asks for and at which cutoff scale, and that the schedule leaves the converged
energy where the plain FP64 run puts it.  J/K come from the oracle, with the requested
cutoff scale applied to its Schwarz and density cutoffs."""
import numpy as np
import torch

from tonto_amd.geometry import water_cluster
from tonto_amd.scf import run_rhf

CUT = 1e-10


def _system(oracle_mod):
    mol = water_cluster(2, "6-31G(d)", seed=4)
    o = oracle_mod.OracleSystem(mol)
    S, T, V = o.one_electron()
    max_I = o.schwarz()
    calls = []

    def jk(P, mode, scale=1.0):
        calls.append((mode, scale))
        c = CUT * scale
        J, K = o.make_r_JK(P.numpy(), schwarz=c, jp=c, kp=c, max_I=max_I)
        return torch.from_numpy(J), torch.from_numpy(K)
    return mol, torch.from_numpy(T + V), torch.from_numpy(S), jk, calls


def test_schedule_and_energy(oracle_mod):
    mol, h, S, jk, calls = _system(oracle_mod)
    ref = run_rhf(h, S, mol.n_electrons, mol.nuclear_repulsion(), jk, convergence=1e-9, diis_tol=1e-6)
    assert ref.converged and all(m == "full" and s == 1.0 for m, s in calls)
    calls.clear()
    r = run_rhf(h, S, mol.n_electrons, mol.nuclear_repulsion(), jk, convergence=1e-9, diis_tol=1e-6,
                mixed_until=1e-5, incremental=True, rebuild_every=8, loose_scale=100.0)
    assert r.converged
    assert abs(r.energy - ref.energy) <= 1e-8
    modes = [m for m, _ in calls]
    errs = [h_["diis_error"] for h_ in r.history]
    # the first build and every build made at DIIS error >= 1e-2 are loose full mixed builds
    assert calls[0] == ("full_mixed", 100.0)
    for k, e in enumerate(errs):
        mode, scale = calls[k + 1]
        if e >= 1e-2 and mode.endswith("_mixed"):
            assert (mode, scale) == ("full_mixed", 100.0)
        else:
            assert scale == 1.0
    # the first build after the loose ones is full, at the usual cutoffs
    first_tight = next(i for i, (m, s) in enumerate(calls) if s == 1.0)
    assert calls[first_tight][0].startswith("full")
    # the precision switch is a full FP64 build, FP64 to the end, and the last build is full FP64
    first64 = next(i for i, m in enumerate(modes) if not m.endswith("_mixed"))
    assert modes[first64] == "full"
    assert all(not m.endswith("_mixed") for m in modes[first64:])
    assert modes[-1] == "full"
    assert "delta" in modes and "delta_mixed" in modes
    # builds are timed and counted by kind
    assert sum(r.n_builds.values()) == len(calls)
    assert all(r.build_seconds[k] > 0.0 for k, n in r.n_builds.items() if n)


def test_rebuild_every_counts_builds_since_the_last_full_one(oracle_mod):
    mol, h, S, jk, calls = _system(oracle_mod)
    r = run_rhf(h, S, mol.n_electrons, mol.nuclear_repulsion(), jk, convergence=1e-11, diis_tol=1e-8,
                incremental=True, rebuild_every=3, incremental_from=1.0)
    assert r.converged
    modes = [m for m, _ in calls]
    run = 0
    for m in modes[1:]:
        run = 0 if m == "full" else run + 1
        assert run <= 2
    assert np.isfinite(r.energy)


def test_staged_schedule_switches_on_the_diis_error(oracle_mod):
    """switch_error: the low-precision phase (loose full builds, then delta builds) runs
    until the DIIS error of its own Fock matrix is below the threshold; then one full FP64
    build and FP64 delta builds to convergence, with no extra final build (final_full
    False).  Energy within 1e-8 of the plain FP64 run."""
    mol, h, S, jk, calls = _system(oracle_mod)
    ref = run_rhf(h, S, mol.n_electrons, mol.nuclear_repulsion(), jk, convergence=1e-9, diis_tol=1e-6)
    calls.clear()
    r = run_rhf(h, S, mol.n_electrons, mol.nuclear_repulsion(), jk, convergence=1e-9, diis_tol=1e-6,
                mixed_until=1e-5, incremental=True, rebuild_every=100, loose_scale=100.0,
                switch_error=1e-4, final_full=False)
    assert r.converged and abs(r.energy - ref.energy) <= 1e-8
    modes = [m for m, _ in calls]
    errs = [h_["diis_error"] for h_ in r.history]
    first64 = next(i for i, m in enumerate(modes) if not m.endswith("_mixed"))
    # the switch comes right after the first iteration whose low-precision DIIS error < 1e-4
    k_switch = next(k for k, e in enumerate(errs) if e < 1e-4)
    assert first64 == k_switch + 2 and modes[first64] == "full"
    assert all(m == "delta" for m in modes[first64 + 1:])
    assert "delta_mixed" in modes[:first64]
    assert len(modes) == r.iterations + 1

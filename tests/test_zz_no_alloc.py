"""The allocation-free build (tgpu_make_r_JK_device_noalloc) against tgpu_make_r_JK_device
through the C ABI: two calls in a row on one GPU, the second on a smaller system (so it
reuses the per-GPU scratch array left dirty by the first), each against the allocating
entry point on the same inputs: J and K to 1e-12 (FP64 atomics add in a different order),
identical quartet counts; and the size limit reported as an error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand_density(n, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    return scale * A @ A.T / n


def _pair(mol, seed, mixed=False, only=None, cut=1e-11):
    import torch
    from tonto_amd.gpu import JKEngine
    e = JKEngine(mol, schwarz=cut)
    e.schwarz()
    e.set_stream(torch.cuda.current_stream().cuda_stream)
    P = torch.tensor(_rand_density(mol.n_bf, seed, 0.5), device="cuda")
    J0, K0, J1, K1 = (torch.full_like(P, float("nan")) for _ in range(4))
    e.make_r_JK_device(P.data_ptr(), J0.data_ptr(), K0.data_ptr(), mixed=mixed)
    torch.cuda.synchronize()
    q0 = e.stats()
    e.make_r_JK_device_noalloc(P.data_ptr(), J1.data_ptr(), K1.data_ptr(), mixed=mixed)
    torch.cuda.synchronize()
    q1 = e.stats()
    dJ = torch.max(torch.abs(J1 - J0)).item()
    dK = torch.max(torch.abs(K1 - K0)).item()
    print(f"n_bf={mol.n_bf} mixed={mixed}: dJ={dJ:.2e} dK={dK:.2e} quartets {q0['quartets']} {q1['quartets']}")
    assert q0["quartets"] == q1["quartets"] and q0["prim_quartets"] == q1["prim_quartets"]
    assert q0["fp32_quartets"] == q1["fp32_quartets"]
    assert dJ <= 1e-12 and dK <= 1e-12
    return e


def test_noalloc_matches_allocating_build_twice():
    import torch
    from tonto_amd.geometry import tonto_water, water_cluster
    assert torch.cuda.is_available(), "needs the MI355X"
    # first: f shells (the wavefront-form workspace comes from the scratch array too)
    _pair(water_cluster(2, "cc-pVTZ", seed=4), 31)
    # second, smaller: the scratch holds the first call's values
    _pair(tonto_water("6-31G(d)"), 32)
    # and the mixed path once more on a third system
    _pair(water_cluster(3, "6-31G(d,p)", seed=6), 33, mixed=True, cut=1e-12)


def test_noalloc_rejects_what_does_not_fit():
    import torch
    from tonto_amd.geometry import water_cluster
    from tonto_amd.gpu import JKEngine, TGPUError
    mol = water_cluster(96, "6-31G(d,p)", seed=2)  # 2400 functions: 2 n_bf^2 doubles > 64 MiB
    e = JKEngine(mol)
    P = torch.zeros((mol.n_bf, mol.n_bf), dtype=torch.float64, device="cuda")
    with pytest.raises(TGPUError, match="maximum is"):
        e.make_r_JK_device_noalloc(P.data_ptr(), P.data_ptr(), P.data_ptr())

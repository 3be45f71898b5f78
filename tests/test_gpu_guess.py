"""GPU tests of the learned guess: the tmlp_* training and inference kernels against the
numpy checker (oracle/mlp_ref.py), and what the trained network gains over the
promolecule guess on held-out clusters at an unchanged converged energy."""
import numpy as np
import pytest

import oracle.mlp_ref as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible: GPU tests must run on an MI355X")
    return torch


@pytest.mark.parametrize("n_in,nh,n_out,M", [(273, 128, 30, 1000), (273, 128, 225, 77), (7, 5, 3, 5),
                                             (273, 128, 4, 4096)])
def test_forward_backward_against_checker(torch_cuda, n_in, nh, n_out, M):
    torch = torch_cuda
    from tonto_amd.mlp import DeviceNet, init_params
    rng = np.random.default_rng(M)
    p = init_params(n_in, nh, n_out, 1) + rng.standard_normal(
        init_params(n_in, nh, n_out, 1).size).astype(np.float32) * 0.01
    X = rng.standard_normal((M, n_in)).astype(np.float32)
    T = rng.standard_normal((M, n_out)).astype(np.float32)
    net = DeviceNet(n_in, nh, n_out, p)
    Xd, Td = torch.tensor(X, device="cuda"), torch.tensor(T, device="cuda")
    scale = 1.0 / (M * n_out)
    loss = float(net.loss_and_grad(Xd, Td, scale)[0])
    g = net.g.cpu().numpy()
    Y = net.forward(Xd).cpu().numpy()
    Yr, cache = R.forward(p, n_in, nh, n_out, X)
    lr, dY = R.mse(Yr, T, scale)
    gr = R.backward(p, n_in, nh, n_out, cache, dY)
    print(f"M={M} n_out={n_out}: max|dY| {np.max(np.abs(Y - Yr)):.2e}, loss {loss:.6e} vs {lr:.6e}, "
          f"max|dg| {np.max(np.abs(g - gr)):.2e} (max|g| {np.max(np.abs(gr)):.2e})")
    # FP32 kernels against the float64 checker
    assert np.max(np.abs(Y - Yr)) <= 1e-4 * max(1.0, np.max(np.abs(Yr)))
    assert abs(loss - lr) <= 1e-4 * abs(lr)
    assert np.max(np.abs(g - gr)) <= 1e-4 * np.max(np.abs(gr))


def test_adam_step_against_checker(torch_cuda):
    torch = torch_cuda
    from tonto_amd.mlp import DeviceNet, init_params
    n_in, nh, n_out = 9, 8, 3
    p = init_params(n_in, nh, n_out, 3)
    net = DeviceNet(n_in, nh, n_out, p)
    rng = np.random.default_rng(0)
    X = torch.tensor(rng.standard_normal((33, n_in)).astype(np.float32), device="cuda")
    T = torch.tensor(rng.standard_normal((33, n_out)).astype(np.float32), device="cuda")
    pr, m, v = p.astype(np.float64), np.zeros(p.size), np.zeros(p.size)
    for step in range(1, 4):
        net.loss_and_grad(X, T, 1.0 / 99)
        g = net.g.cpu().numpy().astype(np.float64)
        net.adam(1e-2)
        pr, m, v = R.adam(pr, g, m, v, 1e-2, 0.9, 0.999, 1e-8, step)
        net_p = net.params()
        # the checker is fed the kernel's own gradients, so the update rule alone is tested
        assert np.max(np.abs(net_p - pr)) < 1e-5
        pr = net_p.astype(np.float64)


def test_training_reduces_the_loss(torch_cuda):
    """A few hundred steps on a smooth target: the kernels' gradients point downhill."""
    torch = torch_cuda
    from tonto_amd.mlp import DeviceNet, init_params
    rng = np.random.default_rng(5)
    X = rng.standard_normal((2048, 16)).astype(np.float32)
    W = rng.standard_normal((16, 4)).astype(np.float32)
    T = np.tanh(X @ W * 0.3).astype(np.float32)
    net = DeviceNet(16, 32, 4, init_params(16, 32, 4, 0))
    Xd, Td = torch.tensor(X, device="cuda"), torch.tensor(T, device="cuda")
    first = float(net.loss_and_grad(Xd, Td, 1.0 / T.size)[0])
    for _ in range(300):
        net.loss_and_grad(Xd, Td, 1.0 / T.size)
        net.adam(3e-3)
    last = float(net.loss_and_grad(Xd, Td, 1.0 / T.size)[0])
    print(f"loss {first:.4e} -> {last:.4e}")
    assert last < 0.05 * first


def test_learned_guess_saves_iterations_and_time_on_held_out_clusters(torch_cuda):
    """(H2O)_20 / 6-31G(d) clusters not seen in training: the SCF from the learned guess
    converges to the same energy (<= 1e-8 Eh) in fewer iterations than from the
    promolecule guess, and in less wall time with making the guess included (runs in
    alternating order after a shared warm-up)."""
    from tonto_amd import guess_mlp as G
    from tonto_amd.guess_train import evaluate
    model = G.default_model()
    assert model is not None, f"{G.DEFAULT_MODEL} is missing"
    r = evaluate(6, 20, model=model)
    print(r)
    assert max(abs(x) for x in r["dE"]) <= 1e-8
    assert np.mean(r["mlp_iterations"]) < np.mean(r["sad_iterations"])
    assert np.sum(r["mlp_wall"]) < np.sum(r["sad_wall"])


def test_learned_guess_projected_onto_cc_pvtz(torch_cuda):
    """The 6-31G(d) network on a cc-pVTZ cluster: its correction to the 6-31G(d)
    promolecule density projected onto cc-pVTZ and added to cc-pVTZ's promolecule
    density.  The projector of a basis onto itself, built through the one-electron
    kernel on the union of the two shell lists, is the unit matrix; the guess holds the
    electron count, is symmetric, and the SCF from it reaches the promolecule run's
    energy (<= 1e-8 Eh) in no more iterations."""
    torch = torch_cuda
    from tonto_amd.driver import GPUSCF
    from tonto_amd.geometry import water_cluster
    mol = water_cluster(4, "cc-pVTZ", seed=2)
    s = GPUSCF(mol, schwarz=1e-11)
    small = water_cluster(4, "6-31G(d)", seed=2)
    Xs = GPUSCF(small).basis_projector(small)
    assert float(torch.max(torch.abs(Xs - torch.eye(small.n_bf, dtype=Xs.dtype, device=Xs.device)))) < 1e-8
    assert s.basis_projector(small).shape == (mol.n_bf, small.n_bf)
    P = s.mlp_guess()
    assert P is not None
    assert abs(float(torch.sum(P * s.S)) - mol.n_electrons) < 1e-10
    assert float(torch.max(torch.abs(P - P.T))) < 1e-12
    r0 = s.run(precision="fp64", convergence=1e-9, diis_tol=1e-6)
    r1 = s.run(precision="fp64", convergence=1e-9, diis_tol=1e-6, guess="mlp")
    print(r0.iterations, r1.iterations, r1.energy - r0.energy)
    assert abs(r1.energy - r0.energy) <= 1e-8
    assert r1.iterations <= r0.iterations

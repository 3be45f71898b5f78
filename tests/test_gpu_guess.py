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


def test_learned_guess_saves_iterations_on_held_out_clusters(torch_cuda):
    """(H2O)_20 / 6-31G(d) clusters not seen in training: the SCF from the learned guess
    converges to the same energy (<= 1e-8 Eh) in fewer iterations than from the
    promolecule guess."""
    from tonto_amd import guess_mlp as G
    from tonto_amd.guess_train import evaluate
    model = G.default_model()
    assert model is not None, f"{G.DEFAULT_MODEL} is missing"
    r = evaluate(2, 20, model=model)
    print(r)
    assert max(abs(x) for x in r["dE"]) <= 1e-8
    assert np.mean(r["mlp_iterations"]) < np.mean(r["sad_iterations"])

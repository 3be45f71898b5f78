"""The learned guess's network on the GPU through libtonto_gpu's tmlp_* kernels
(include/tonto_mlp.h).  torch only holds the device buffers and draws minibatch
indices; every forward, backward, loss and Adam step is a hand-written HIP kernel.

Network: n_in -> nh -> nh -> n_out, SiLU after the two hidden layers, FP32 weights and
FP32 accumulation.  Parameters are one flat float array in the layout of tonto_mlp.h.
"""
import ctypes

import numpy as np

from .gpu import TGPUError, lib as _gpu_lib

_bound = False


def lib():
    global _bound
    L = _gpu_lib()
    if not _bound:
        vp, fp, lg, i = ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_int
        L.tmlp_n_params.argtypes = [i, i, i]
        L.tmlp_n_params.restype = lg
        L.tmlp_forward_workspace.argtypes = [lg, i]
        L.tmlp_forward_workspace.restype = lg
        L.tmlp_backward_workspace.argtypes = [lg, i]
        L.tmlp_backward_workspace.restype = lg
        L.tmlp_forward.argtypes = [fp, i, i, i, fp, lg, fp, fp, vp]
        L.tmlp_backward.argtypes = [fp, i, i, i, fp, lg, fp, fp, fp, fp, vp]
        L.tmlp_mse.argtypes = [fp, fp, lg, ctypes.c_float, fp, fp, vp]
        L.tmlp_adam.argtypes = [fp, fp, fp, fp, lg, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                ctypes.c_float, i, vp]
        _bound = True
    return L


def _check(rc, what):
    if rc != 0:
        raise TGPUError(f"{what} failed with status {rc}")


def n_params(n_in, nh, n_out):
    return int(lib().tmlp_n_params(n_in, nh, n_out))


def init_params(n_in, nh, n_out, seed):
    """He-style initial weights (numpy float32, flat, tonto_mlp.h layout)."""
    rng = np.random.default_rng(seed)
    parts = []
    for a, b in ((n_in, nh), (nh, nh), (nh, n_out)):
        parts.append((rng.standard_normal((a, b)) * np.sqrt(2.0 / a)).astype(np.float32).ravel())
        parts.append(np.zeros(b, np.float32))
    return np.concatenate(parts)


class DeviceNet:
    """One network resident on the GPU, with its Adam state."""

    def __init__(self, n_in, nh, n_out, params, device="cuda"):
        import torch
        self.torch = torch
        self.n_in, self.nh, self.n_out = n_in, nh, n_out
        n = n_params(n_in, nh, n_out)
        params = np.asarray(params, np.float32)
        if params.size != n:
            raise ValueError(f"expected {n} parameters, got {params.size}")
        self.device = torch.device(device)
        self.p = torch.tensor(params, device=self.device)
        self.g = torch.zeros_like(self.p)
        self.m = torch.zeros_like(self.p)
        self.v = torch.zeros_like(self.p)
        self.step = 0
        self._ws = {}
        self.loss = torch.zeros(1, dtype=torch.float32, device=self.device)

    def _stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    def _work(self, M):
        t = self.torch
        if M not in self._ws:
            self._ws.clear()
            self._ws[M] = (t.empty(int(lib().tmlp_forward_workspace(M, self.nh)), dtype=t.float32, device=self.device),
                           t.empty(int(lib().tmlp_backward_workspace(M, self.nh)), dtype=t.float32, device=self.device),
                           t.empty((M, self.n_out), dtype=t.float32, device=self.device),
                           t.empty((M, self.n_out), dtype=t.float32, device=self.device))
        return self._ws[M]

    def forward(self, X):
        """X: contiguous float32 device tensor (M, n_in) -> new tensor (M, n_out)."""
        M = int(X.shape[0])
        ws, _, Y, _ = self._work(M)
        _check(lib().tmlp_forward(self.p.data_ptr(), self.n_in, self.nh, self.n_out, X.data_ptr(), M,
                                  ws.data_ptr(), Y.data_ptr(), self._stream()), "tmlp_forward")
        return Y.clone()

    def loss_and_grad(self, X, T, scale):
        """Forward + MSE (scale * sum (Y - T)^2) + backward; gradient left in self.g.
        Returns the loss as a device tensor (no host synchronisation)."""
        M = int(X.shape[0])
        ws, ws2, Y, dY = self._work(M)
        L = lib()
        s = self._stream()
        _check(L.tmlp_forward(self.p.data_ptr(), self.n_in, self.nh, self.n_out, X.data_ptr(), M,
                              ws.data_ptr(), Y.data_ptr(), s), "tmlp_forward")
        _check(L.tmlp_mse(Y.data_ptr(), T.data_ptr(), M * self.n_out, float(scale), dY.data_ptr(),
                          self.loss.data_ptr(), s), "tmlp_mse")
        _check(L.tmlp_backward(self.p.data_ptr(), self.n_in, self.nh, self.n_out, X.data_ptr(), M,
                               ws.data_ptr(), dY.data_ptr(), self.g.data_ptr(), ws2.data_ptr(), s),
               "tmlp_backward")
        return self.loss

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        _check(lib().tmlp_adam(self.p.data_ptr(), self.g.data_ptr(), self.m.data_ptr(), self.v.data_ptr(),
                               self.p.numel(), lr, beta1, beta2, eps, self.step, self._stream()), "tmlp_adam")

    def params(self):
        return self.p.detach().cpu().numpy().copy()

"""Restricted Hartree-Fock SCF driver around the J/K build.

Mirrors the shape of Tonto's `MOLECULE.SCF:usual_SCF` (`foofiles/molecule.scf.foo:1690-1727`):
extrapolate F (DIIS, keep 8, start at iteration 3: `include/macros.in:594-601`) ->
diagonalise in the S^{-1/2} basis (`MO_r_eigen_update`, `molecule.scf.foo:3797-3851`,
level shift 0.3 for the first iterations, `types.foo:6047-6050`) -> density
P = 2 C_occ C_occ^T with damping 1/2 on the first iterations (`molecule.base.foo:318-517`,
`types.foo:6034-6038`) -> new F = h + J - K/2 (`molecule.fock.foo:136,153`) ->
E = 1/2 tr P(F+h) + V_NN (`molecule.scf.foo:4398-4433`) -> converged when |dE| and the
DIIS error are both below their tolerances (`scf_data.foo:2283-2313`).

Matrices are torch tensors on any device; the J/K provider decides where the Fock
build runs.  The provider signature is jk(P, mode) -> (J, K) where mode is "full" or
"delta" (a build from a density difference, used by incremental builds).
"""
import time

import torch


class SCFResult(dict):
    __getattr__ = dict.__getitem__


def _sym_orth(S):
    w, U = torch.linalg.eigh(S)
    return (U * w.rsqrt()) @ U.T


def core_guess(h, X, n_occ):
    e, Cp = torch.linalg.eigh(X.T @ h @ X)
    C = X @ Cp
    Co = C[:, :n_occ]
    return 2.0 * Co @ Co.T


def run_rhf(h, S, n_electrons, e_nuc, jk, P0=None, convergence=1e-8, diis_tol=1e-6,
            max_iterations=100, min_iterations=1, diis_keep=8, diis_start=3,
            level_shift=0.3, level_shift_iters=3, damp=0.5, damp_iters=3, incremental=None,
            log=None):
    """Run RHF to convergence.

    incremental: None, or dict(start=k, rebuild_every=m, final_full=True): from
      iteration k on, the 2e Fock part is updated with a build on P - P_prev
      ("delta" mode); every m-th iteration and once converged a full build is done,
      and convergence is only accepted on a full build.
    """
    dt = h.dtype
    n_occ = n_electrons // 2
    X = _sym_orth(S)
    P = core_guess(h, X, n_occ) if P0 is None else P0.clone()
    t0 = time.perf_counter()
    J, K = jk(P, "full")
    G = J - 0.5 * K
    F = h + G
    E = 0.5 * torch.sum(P * (F + h)).item() + e_nuc
    history = []
    diis_F, diis_E = [], []
    n_builds = {"full": 1, "delta": 0}
    P_build = P
    last_full = True
    converged = False
    it = 0
    while True:
        it += 1
        # DIIS error and extrapolation
        err = X.T @ (F @ P @ S - S @ P @ F) @ X
        err_max = err.abs().max().item()
        diis_F.append(F)
        diis_E.append(err)
        if len(diis_F) > diis_keep:
            diis_F.pop(0)
            diis_E.pop(0)
        F_use = F
        if it >= diis_start and len(diis_F) >= 2:
            m = len(diis_F)
            B = torch.empty((m + 1, m + 1), dtype=dt, device=h.device)
            for i in range(m):
                for j in range(i + 1):
                    B[i, j] = B[j, i] = torch.sum(diis_E[i] * diis_E[j])
            B[m, :] = -1.0
            B[:, m] = -1.0
            B[m, m] = 0.0
            rhs = torch.zeros(m + 1, dtype=dt, device=h.device)
            rhs[m] = -1.0
            try:
                c = torch.linalg.solve(B, rhs)[:m]
                F_use = sum(ci * Fi for ci, Fi in zip(c, diis_F))
            except RuntimeError:
                F_use = F
        # diagonalise (level shift on virtual space in early iterations)
        Fp = X.T @ F_use @ X
        if level_shift and it < level_shift_iters:
            Cocc_p = torch.linalg.pinv(X) @ (P / 2.0) @ torch.linalg.pinv(X).T
            Fp = Fp + level_shift * (torch.eye(Fp.shape[0], dtype=dt, device=h.device) - Cocc_p)
        eps, Cp = torch.linalg.eigh(Fp)
        C = X @ Cp
        Co = C[:, :n_occ]
        P_new = 2.0 * Co @ Co.T
        if damp and it < damp_iters:
            P_new = damp * P + (1.0 - damp) * P_new
        # Fock build: full or incremental
        full = True
        if incremental and it >= incremental.get("start", 2):
            every = incremental.get("rebuild_every", 8)
            full = (it % every == 0)
        if full:
            J, K = jk(P_new, "full")
            G = J - 0.5 * K
            n_builds["full"] += 1
        else:
            dJ, dK = jk(P_new - P_build, "delta")
            G = G + dJ - 0.5 * dK
            n_builds["delta"] += 1
        P_build = P_new
        last_full = full
        P = P_new
        F = h + G
        E_new = 0.5 * torch.sum(P * (F + h)).item() + e_nuc
        dE = E_new - E
        E = E_new
        history.append(dict(iteration=it, energy=E, dE=dE, diis_error=err_max, full=full))
        if log:
            log(f"iter {it:3d} E={E:.12f} dE={dE:+.3e} err={err_max:.3e} {'full' if full else 'delta'}")
        done = abs(dE) < convergence and err_max < diis_tol
        if done and incremental and not last_full:
            # win back the incremental error: one full build on the current density
            J, K = jk(P, "full")
            G = J - 0.5 * K
            n_builds["full"] += 1
            F = h + G
            E_full = 0.5 * torch.sum(P * (F + h)).item() + e_nuc
            history.append(dict(iteration=it, energy=E_full, dE=E_full - E, diis_error=err_max,
                                full=True))
            done = abs(E_full - E) < convergence
            E = E_full
            last_full = True
        if (done and it >= min_iterations) or it >= max_iterations:
            converged = done
            break
    wall = time.perf_counter() - t0
    return SCFResult(energy=E, converged=converged, iterations=it, P=P, F=F, C=C, eps=eps,
                     history=history, wall=wall, n_builds=n_builds)

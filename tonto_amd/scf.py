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
build runs (see run_rhf for its signature).
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


def atomic_density(h, S, n_electrons, jk, max_iterations=60, tol=1e-9, damp=0.5):
    """Spherically averaged atomic density for the promolecule guess: restricted SCF with
    fractional aufbau occupations, the electrons of the open shell spread evenly over its
    degenerate orbitals (Tonto's promolecule guess, molecule.scf.foo:4827-4844, sums
    atomic densities)."""
    X = _sym_orth(S)
    e, Cp = torch.linalg.eigh(X.T @ h @ X)
    P = None
    for it in range(max_iterations):
        if P is not None:
            J, K = jk(P, "full")
            F = h + J - 0.5 * K
            e, Cp = torch.linalg.eigh(X.T @ F @ X)
        C = X @ Cp
        occ = torch.zeros_like(e)
        left = float(n_electrons)
        i = 0
        n = e.shape[0]
        while left > 1e-12 and i < n:
            # degenerate set starting at orbital i
            j = i + 1
            while j < n and abs(float(e[j] - e[i])) < 1e-4:
                j += 1
            cap = 2.0 * (j - i)
            put = min(cap, left)
            occ[i:j] = put / (j - i)
            left -= put
            i = j
        P_new = (C * occ) @ C.T
        if P is not None and damp and it < max_iterations // 2:
            P_new = damp * P + (1.0 - damp) * P_new
        if P is not None and (P_new - P).abs().max().item() < tol:
            return P_new
        P = P_new
    return P


def promolecule_guess(mol, atom_scf):
    """Block-diagonal superposition of spherically averaged atomic densities.
    atom_scf(symbol) -> atomic density matrix (n_bf_atom x n_bf_atom) in the same basis."""
    cache = {}
    blocks = []
    for ia, sym in enumerate(mol.symbols):
        if sym not in cache:
            cache[sym] = atom_scf(sym)
        blocks.append(cache[sym])
    return torch.block_diag(*blocks)


def _gram_append(gram, errs):
    """Gram matrix of the error vectors errs (tensors, or tuples of tensors whose inner
    products add up) after the last one was appended: only the new row is computed, on
    the device (no host round trip)."""
    def dot(a, b):
        if isinstance(a, tuple):
            return sum(torch.sum(x * y) for x, y in zip(a, b))
        return torch.sum(a * b)
    new = torch.stack([dot(e, errs[-1]) for e in errs])
    if gram is None:
        return new.reshape(1, 1)
    return torch.cat([torch.cat([gram, new[:-1, None]], 1), new[None, :]], 0)


def _diis_matrix(gram):
    """Pulay's bordered matrix [[B, -1], [-1, 0]] from the Gram matrix B."""
    m = gram.shape[0]
    B = torch.full((m + 1, m + 1), -1.0, dtype=gram.dtype, device=gram.device)
    B[:m, :m] = gram
    B[m, m] = 0.0
    return B


def run_rhf(h, S, n_electrons, e_nuc, jk, P0=None, convergence=1e-8, diis_tol=1e-6,
            max_iterations=100, min_iterations=1, diis_keep=8, diis_start=3,
            diis_start_error=None,
            level_shift=0.3, level_shift_iters=3, damp=0.5, damp_iters=3,
            mixed_until=None, incremental=False, rebuild_every=8, incremental_from=1e-2,
            loose_scale=1.0, loose_until=None, rebuild_after_loose=True, diis_reset_on_switch=False,
            switch_error=None, final_full=True, keep_fock=False, log=None):
    """Run RHF to convergence.

    jk(P, mode) returns (J, K); mode is "full" or "delta" (a build from a density
    difference), with the suffix "_mixed" while low precision is allowed; jk(P, mode, scale)
    is called for builds whose screening cutoffs are to be scale times the usual ones.
    mixed_until: None for FP64 builds throughout; otherwise builds are "_mixed" until
      |dE| < mixed_until, then a full FP64 build is forced and FP64 is kept to the end.
    loose_scale: cutoff scale of the "_mixed" builds made while the DIIS error is still
      at or above loose_until (default incremental_from; the first build included).  They
      are full builds, their screening error (~1e-5 in F at a 1e-8 cutoff) stays below the
      SCF error then, and the next full build at the usual cutoffs (at the latest the FP64
      one at the precision switch) removes it.
    rebuild_after_loose: make the first build after the loose ones a full build at the
      usual cutoffs; otherwise delta builds go on from the loose G and the FP64 build at
      the precision switch is the first one without the loose error.
    diis_reset_on_switch: drop the DIIS history at the precision switch.
    switch_error: if given, the precision switch comes when the DIIS error of the
      low-precision Fock matrix falls below it (instead of |dE| < mixed_until): the
      low-precision phase then converges the density as far as its own Fock matrix
      allows, and the FP64 phase starts close to convergence.
    keep_fock: keep each iteration's density and Fock matrix (host copies) in the history.
    final_full: once converged on delta builds, make one more full FP64 build and accept
      only its energy (False: the FP64 delta builds after the switch are trusted, so their
      cutoffs must keep the energy within the convergence tolerance).
    incremental: update the 2e Fock part with builds on P - P_prev; after rebuild_every
      builds since the last full one, at the precision switch and once converged a full
      build is done, and convergence is only accepted on a full FP64 build.  While the
      DIIS error is still at or above incremental_from, P - P_prev is as large as P and a
      difference build (run at a tighter cutoff) would cost more than a full one, so full
      builds are used.
    """
    dt = h.dtype
    n_occ = n_electrons // 2
    X = _sym_orth(S)
    P = core_guess(h, X, n_occ) if P0 is None else P0.clone()
    low = mixed_until is not None
    loose_until = incremental_from if loose_until is None else loose_until
    n_builds = {"full": 0, "delta": 0, "full_mixed": 0, "delta_mixed": 0}
    build_s = dict.fromkeys(n_builds, 0.0)  # wall time of the builds by kind

    def build(D, mode, loose=False):
        scale = loose_scale if loose and mode == "full_mixed" else 1.0
        tb = time.perf_counter()
        J, K = jk(D, mode) if scale == 1.0 else jk(D, mode, scale)
        if J.is_cuda:
            torch.cuda.synchronize(J.device)
        build_s[mode] += time.perf_counter() - tb
        n_builds[mode] += 1
        return J, K

    t0 = time.perf_counter()
    J, K = build(P, "full_mixed" if low else "full", loose=True)
    G = J - 0.5 * K
    F = h + G
    E = 0.5 * torch.sum(P * (F + h)).item() + e_nuc
    history = []
    diis_F, diis_E = [], []
    gram = None  # Gram matrix of the kept DIIS error vectors, on the device
    P_build = P
    since_full = 0  # builds since the last full one
    loose_G = low and loose_scale != 1.0  # G holds a loose-cutoff build
    last_full_fp64 = not low
    force_full = False
    converged = False
    diis_on = diis_start_error is None
    it = 0
    while True:
        it += 1
        ti = time.perf_counter()
        err = X.T @ (F @ P @ S - S @ P @ F) @ X
        err_max = err.abs().max().item()
        diis_F.append(F)
        diis_E.append(err)
        gram = _gram_append(gram, diis_E)
        if len(diis_F) > diis_keep:
            diis_F.pop(0)
            diis_E.pop(0)
            gram = gram[1:, 1:]
        F_use = F
        if diis_start_error is not None and err_max < diis_start_error:
            diis_on = True
        if it >= diis_start and len(diis_F) >= 2 and diis_on:
            m = len(diis_F)
            B = _diis_matrix(gram)
            rhs = torch.zeros(m + 1, dtype=dt, device=h.device)
            rhs[m] = -1.0
            try:
                c = torch.linalg.solve(B, rhs)[:m]
                F_use = sum(ci * Fi for ci, Fi in zip(c, diis_F))
            except RuntimeError:
                F_use = F
        Fp = X.T @ F_use @ X
        if level_shift and it < level_shift_iters:
            Xi = S @ X  # X = S^-1/2, so X^-1 = S^1/2 = S X
            Cocc_p = Xi @ (P / 2.0) @ Xi.T
            Fp = Fp + level_shift * (torch.eye(Fp.shape[0], dtype=dt, device=h.device) - Cocc_p)
        eps, Cp = torch.linalg.eigh(Fp)
        C = X @ Cp
        Co = C[:, :n_occ]
        P_new = 2.0 * Co @ Co.T
        if damp and (it < damp_iters or not diis_on):
            P_new = damp * P + (1.0 - damp) * P_new
        # a G from a loose-cutoff build is replaced by a full build before any delta
        full = ((not incremental) or force_full or since_full + 1 >= rebuild_every or
                err_max >= incremental_from or (loose_G and rebuild_after_loose))
        mode = ("full" if full else "delta") + ("_mixed" if low else "")
        tb = time.perf_counter()
        if full:
            loose_G = low and loose_scale != 1.0 and err_max >= loose_until
            J, K = build(P_new, mode, loose=loose_G)
            G = J - 0.5 * K
        else:
            dJ, dK = build(P_new - P_build, mode)
            G = G + dJ - 0.5 * dK
        tb = time.perf_counter() - tb
        since_full = 0 if full else since_full + 1
        force_full = False
        last_full_fp64 = full and not low
        P_build = P_new
        P = P_new
        F = h + G
        E_new = 0.5 * torch.sum(P * (F + h)).item() + e_nuc
        dE = E_new - E
        E = E_new
        history.append(dict(iteration=it, energy=E, dE=dE, diis_error=err_max, mode=mode,
                            build_seconds=tb, seconds=time.perf_counter() - ti))
        if keep_fock:
            history[-1].update(P=P.detach().cpu().numpy().copy(), F=F.detach().cpu().numpy().copy())
        if log:
            log(f"iter {it:3d} E={E:.12f} dE={dE:+.3e} err={err_max:.3e} {mode} "
                f"build {tb:.3f}s iteration {history[-1]['seconds']:.3f}s")
        done = abs(dE) < convergence and err_max < diis_tol
        if low:
            switch = (err_max < switch_error) if switch_error is not None else abs(dE) < mixed_until
            if switch or done:
                low = False
                force_full = True
                if diis_reset_on_switch:
                    diis_F.clear()
                    diis_E.clear()
                    gram = None
            done = False
        elif done and not last_full_fp64 and final_full:
            # win back the incremental error: a full FP64 build on the current density
            J, K = build(P, "full")
            G = J - 0.5 * K
            F = h + G
            E_full = 0.5 * torch.sum(P * (F + h)).item() + e_nuc
            history.append(dict(iteration=it, energy=E_full, dE=E_full - E, diis_error=err_max,
                                mode="full"))
            done = abs(E_full - E) < convergence
            E = E_full
            last_full_fp64 = True
        if (done and it >= min_iterations) or it >= max_iterations:
            converged = done
            break
    wall = time.perf_counter() - t0
    return SCFResult(energy=E, converged=converged, iterations=it, P=P, F=F, C=C, eps=eps,
                     history=history, wall=wall, n_builds=n_builds, build_seconds=build_s)


def run_uhf(h, S, n_alpha, n_beta, e_nuc, jk, Pa0=None, Pb0=None, convergence=1e-8, diis_tol=1e-6,
            max_iterations=100, diis_keep=8, diis_start=3, level_shift=0.3, level_shift_iters=3,
            damp=0.5, damp_iters=3, log=None):
    """Unrestricted Hartree-Fock to convergence, the spin-resolved form of run_rhf.

    jk(Pa, Pb) returns (Ja, Jb, Ka, Kb) as Tonto's make_u_JK_engine leaves them
    (molecule.fock.foo:2558-2712); the Fock matrices are Tonto's
    F.a = h + J.a + J.b - K.a, F.b = h + J.a + J.b - K.b (molecule.fock.foo:247-300),
    the energy E = 1/2 tr[(Pa + Pb) h + Pa F.a + Pb F.b] + V_NN.  One DIIS for both spins
    (the error matrices of both in one inner product, the same coefficients for F.a and
    F.b), and damping and level shift of the first iterations as in run_rhf.  Converged
    when |dE| and the larger DIIS error are below their tolerances."""
    dt = h.dtype
    X = _sym_orth(S)
    n = h.shape[0]
    eye = torch.eye(n, dtype=dt, device=h.device)

    def occupy(F, n_occ):
        e, Cp = torch.linalg.eigh(X.T @ F @ X)
        C = X @ Cp
        return C[:, :n_occ] @ C[:, :n_occ].T, e, C

    if Pa0 is None or Pb0 is None:
        Pa, _, _ = occupy(h, n_alpha)
        Pb, _, _ = occupy(h, n_beta)
    else:
        Pa, Pb = Pa0.clone(), Pb0.clone()

    def fock(Pa, Pb):
        Ja, Jb, Ka, Kb = jk(Pa, Pb)
        J = Ja + Jb
        Fa, Fb = h + J - Ka, h + J - Kb
        E = 0.5 * torch.sum((Pa + Pb) * h + Pa * Fa + Pb * Fb).item() + e_nuc
        return Fa, Fb, E

    t0 = time.perf_counter()
    Fa, Fb, E = fock(Pa, Pb)
    hist_F, hist_E = [], []
    gram = None  # Gram matrix of the kept (alpha, beta) error pairs
    history = []
    converged = False
    it = 0
    while it < max_iterations:
        it += 1
        ea = X.T @ (Fa @ Pa @ S - S @ Pa @ Fa) @ X
        eb = X.T @ (Fb @ Pb @ S - S @ Pb @ Fb) @ X
        err = max(ea.abs().max().item(), eb.abs().max().item())
        hist_F.append((Fa, Fb))
        hist_E.append((ea, eb))
        gram = _gram_append(gram, hist_E)
        if len(hist_F) > diis_keep:
            hist_F.pop(0)
            hist_E.pop(0)
            gram = gram[1:, 1:]
        Fa_u, Fb_u = Fa, Fb
        if it >= diis_start and len(hist_F) >= 2:
            m = len(hist_F)
            B = _diis_matrix(gram)
            rhs = torch.zeros(m + 1, dtype=dt, device=h.device)
            rhs[m] = -1.0
            try:
                c = torch.linalg.solve(B, rhs)[:m]
                Fa_u = sum(ci * F[0] for ci, F in zip(c, hist_F))
                Fb_u = sum(ci * F[1] for ci, F in zip(c, hist_F))
            except RuntimeError:
                pass
        Pn = []
        for F_u, P, n_occ in ((Fa_u, Pa, n_alpha), (Fb_u, Pb, n_beta)):
            if level_shift and it < level_shift_iters:
                # virtual space raised by the shift: F + shift (S - S P S) in the S^-1/2 basis
                Xi = S @ X
                Fp = X.T @ F_u @ X + level_shift * (eye - Xi @ P @ Xi.T)
                e, Cp = torch.linalg.eigh(Fp)
                C = X @ Cp
                Pnew = C[:, :n_occ] @ C[:, :n_occ].T
            else:
                Pnew, _, _ = occupy(F_u, n_occ)
            if damp and it < damp_iters:
                Pnew = damp * P + (1.0 - damp) * Pnew
            Pn.append(Pnew)
        Pa, Pb = Pn
        Fa, Fb, E_new = fock(Pa, Pb)
        dE = E_new - E
        E = E_new
        history.append(dict(iteration=it, energy=E, dE=dE, diis_error=err))
        if log:
            log(f"uhf iter {it:3d} E={E:.12f} dE={dE:+.3e} err={err:.3e}")
        if abs(dE) < convergence and err < diis_tol:
            converged = True
            break
    return SCFResult(energy=E, converged=converged, iterations=it, Pa=Pa, Pb=Pb, Fa=Fa, Fb=Fb,
                     history=history, wall=time.perf_counter() - t0)

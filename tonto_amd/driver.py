"""GPU-resident RHF: one-electron matrices, Schwarz table and every J/K build on the
device through libtonto_gpu's C ABI; the SCF algebra (DIIS, eigen-solve, density) in
torch on the same device and stream.  Mirrors MOLECULE.SCF:usual_SCF
(foofiles/molecule.scf.foo:1690-1727) with make_scf_Fock_mx (1562-1620) as the J/K call.
"""
import time

import torch

from .basis import Molecule
from .gpu import JKEngine
from .scf import atomic_density, promolecule_guess, run_rhf, run_uhf

_ATOM_DENSITY = {}


class GPUSCF:
    def __init__(self, mol, schwarz=1e-9, j_density=1e-9, k_density=1e-9, pair=1e-6,
                 fp32_threshold=1e-6, device="cuda", spherical=False):
        if not torch.cuda.is_available():
            raise RuntimeError("GPUSCF needs a HIP device")
        self.mol = mol
        self.device = torch.device(device)
        t0 = time.perf_counter()
        # spherical: Tonto's use_spherical_basis (solid harmonics for d and f, the
        # make_r_JK_direct build); every matrix below is then n_sph x n_sph
        self.spherical = bool(spherical)
        self.engine = JKEngine(mol, schwarz=schwarz, j_density=j_density, k_density=k_density,
                               pair=pair, fp32_threshold=fp32_threshold, spherical=spherical)
        S, T, V = self.engine.one_electron()
        self.S = torch.tensor(S, device=self.device)
        self.h = torch.tensor(T + V, device=self.device)
        self.max_I = self.engine.schwarz()
        self.setup_s = time.perf_counter() - t0
        n = self.engine.n_bf
        # persistent device buffers: the build reads P from and writes J, K to fixed
        # addresses, so a captured build (graph=True) can be replayed every iteration
        self._P = torch.empty((n, n), dtype=torch.float64, device=self.device)
        self._J = torch.empty((n, n), dtype=torch.float64, device=self.device)
        self._K = torch.empty((n, n), dtype=torch.float64, device=self.device)
        self.graph = False
        self.delta_cutoff = 1e-14
        self.delta_density_cutoff = 1e-12
        self.mixed_delta_cutoff = 1e-10
        self.mixed_delta_density_cutoff = 1e-12
        # Schwarz cutoff of the mixed run's first full builds (see run)
        self.loose_cutoff = 1e-8
        # mixed run: DIIS error at which the low-precision phase hands over to FP64, and
        # whether a converged run makes one more full FP64 build (see run)
        self.switch_error = 1e-4
        self.final_full = False

    def autotune(self, P, mixed=False):
        """Pick the faster kernel form per class pair on this molecule and density."""
        self._P.copy_(P)
        self.engine.set_stream(torch.cuda.current_stream(self.device).cuda_stream)
        return self.engine.autotune(self._P.data_ptr(), mixed=mixed)

    def jk(self, P, mode="full", scale=1.0):
        """One J/K build on the device; mode as in tonto_amd.scf.run_rhf, screening
        cutoffs `scale` times the engine's."""
        if scale != 1.0:
            self.engine.set_cutoffs_scaled(scale)
            try:
                return self.jk(P, mode)
            finally:
                self.engine.set_cutoffs_scaled(1.0)
        self._P.copy_(P)
        self.engine.set_stream(torch.cuda.current_stream(self.device).cuda_stream)
        self.engine.make_r_JK_device(self._P.data_ptr(), self._J.data_ptr(), self._K.data_ptr(),
                                     mixed=mode.endswith("_mixed"), graph=self.graph)
        return self._J.clone(), self._K.clone()

    def atom_scf(self, symbol, basis=None):
        """Spherically averaged density of one atom in the molecule's basis (or `basis`),
        on the GPU."""
        from .basis import ELEMENT_Z
        atom = Molecule.from_angstrom([symbol], [[0.0, 0.0, 0.0]], basis or self.mol.basis_name)
        eng = JKEngine(atom, spherical=self.spherical)
        S, T, V = eng.one_electron()
        eng.schwarz()
        n = eng.n_bf
        Pd = torch.empty((n, n), dtype=torch.float64, device=self.device)
        Jd, Kd = torch.empty_like(Pd), torch.empty_like(Pd)

        def jk(P, mode):
            Pd.copy_(P)
            eng.set_stream(torch.cuda.current_stream(self.device).cuda_stream)
            eng.make_r_JK_device(Pd.data_ptr(), Jd.data_ptr(), Kd.data_ptr())
            return Jd.clone(), Kd.clone()
        return atomic_density(torch.tensor(T + V, device=self.device), torch.tensor(S, device=self.device),
                              ELEMENT_Z[symbol], jk)

    def atom_density(self, symbol, basis=None):
        """Promolecule block of one element (numpy; cached per basis and element)."""
        key = (basis or self.mol.basis_name, symbol, self.spherical)
        if key not in _ATOM_DENSITY:
            _ATOM_DENSITY[key] = self.atom_scf(symbol, basis).cpu().numpy()
        return _ATOM_DENSITY[key]

    def guess(self, kind="promolecule", model=None):
        """Initial density: "core" (None: run_rhf diagonalises h), "promolecule", or "mlp"
        (the learned guess; falls back to the promolecule guess for molecules or bases it
        does not cover)."""
        if kind == "core":
            return None
        if kind == "mlp":
            P = self.mlp_guess(model)
            if P is not None:
                return P
        return promolecule_guess(self.mol, lambda sym: torch.tensor(self.atom_density(sym),
                                                                    device=self.device))

    def mlp_guess(self, model=None):
        """The learned guess (tonto_amd.guess_mlp) for an O/H molecule, or None.  A model
        trained in this basis gives the density directly.  A model trained in another
        basis B gives its correction to B's promolecule density, D_B = P_B - P_B^atoms;
        that is carried into this basis by the least-squares projection
        X = S^-1 S_(this, B) and added to this basis's promolecule density:
        P = P^atoms + X D_B X^T, scaled to the electron count."""
        from . import guess_mlp as G
        model = model or G.default_model()
        if model is None or self.spherical or not set(self.mol.symbols) <= {"O", "H"}:
            return None
        mb = str(model.meta.get("basis", ""))
        predict = model.predictor(G.gpu_forward(model, self.device))
        S = self.S.cpu().numpy()
        if mb == self.mol.basis_name:
            atom = {sym: self.atom_density(sym) for sym in set(self.mol.symbols)}
            P = G.assemble(self.mol, atom, predict)
            return torch.tensor(G.normalise_electrons(P, S, self.mol.n_electrons), device=self.device)
        small = Molecule(list(self.mol.symbols), self.mol.coords.copy(), mb, self.mol.charge)
        small.build_shells()
        atom_b = {sym: self.atom_density(sym, mb) for sym in set(self.mol.symbols)}
        PB = G.assemble(small, atom_b, predict)
        PB0 = promolecule_guess(small, lambda sym: torch.tensor(atom_b[sym])).numpy()
        X = self.basis_projector(small)
        D = torch.tensor(PB - PB0, device=self.device)
        P = self.guess("promolecule") + X @ D @ X.T
        tr = float(torch.sum(P * self.S))
        return P * (self.mol.n_electrons / tr) if tr > 0 else P

    def basis_projector(self, other):
        """X = S^-1 S_(this basis, other basis) (torch, on the device): the least-squares
        expansion of the other basis's functions (same atoms) in this one.  The mixed
        overlap comes from the one-electron kernel of the union of the two shell lists."""
        import dataclasses
        n = self.mol.n_bf
        u = Molecule(list(self.mol.symbols), self.mol.coords, f"{self.mol.basis_name}+{other.basis_name}",
                     self.mol.charge)
        u.shells = list(self.mol.shells) + [dataclasses.replace(s, first_bf=s.first_bf + n) for s in other.shells]
        u.n_bf = n + other.n_bf
        eng = JKEngine(u)
        try:
            Su, _, _ = eng.one_electron()
        finally:
            eng.close()
        Sx = torch.tensor(Su[:n, n:], device=self.device)
        return torch.linalg.solve(self.S, Sx)

    def jk_incremental(self, P, mode, scale=1.0):
        """Delta builds are screened with other cutoffs than full builds.  The
        contributions a full build drops are bounded by the cutoffs times |P|, a delta
        build's by the cutoffs times |dP|, and those errors accumulate over the delta
        builds; the density cutoffs (skip a block whose |dP| is below them, whatever its
        integrals) cost the most accuracy: on (H2O)64 / cc-pVTZ a difference build at the
        full-build cutoffs is off by ~1e-7 Eh, with the density tests off by ~2e-8 Eh.
        FP64 delta builds use the Schwarz cutoff delta_cutoff (default 1e-14) and the
        density cutoffs delta_density_cutoff (default 1e-12), which keeps the energy drift
        over the final iterations below the 1e-9 Eh convergence test (a Schwarz cutoff of
        1e-13 or 1e-12 leaves ~2e-9 Eh per iteration).  In the mixed phase, whose errors
        the full FP64 build at the precision switch removes, the Schwarz cutoff is
        mixed_delta_cutoff (default 1e-10) and the density cutoffs
        mixed_delta_density_cutoff (default 1e-12): ~5e-8 Eh per build on
        (H2O)64 / cc-pVTZ at half the cost of the FP64 ones.  Cutoffs are never loosened
        below the full-build ones."""
        if mode in ("delta", "delta_mixed"):
            s, j, _, _ = self.engine.cutoffs
            cs, cd = ((self.mixed_delta_cutoff, self.mixed_delta_density_cutoff) if mode == "delta_mixed"
                      else (self.delta_cutoff, self.delta_density_cutoff))
            return self.jk_scaled(P, mode, min(1.0, cs / s) if s > 0 else 1.0,
                                  min(1.0, cd / j) if j > 0 else 1.0)
        return self.jk(P, mode, scale)

    def jk_scaled(self, P, mode, schwarz_scale, density_scale):
        """A build with the Schwarz cutoff and the density cutoffs scaled separately."""
        self.engine.set_cutoffs_scaled(schwarz_scale, density_scale)
        try:
            return self.jk(P, mode)
        finally:
            self.engine.set_cutoffs_scaled(1.0)

    def run(self, precision="fp64", mixed_until=1e-5, incremental=None, guess="promolecule",
            log=None, guess_model=None, **kw):
        """RHF to convergence; the wall time includes making the initial guess.

        precision="fp64": a full FP64 build at the engine's cutoffs every iteration (the
        plain path).  precision="mixed" (defaults, each overridable through kw):
          * builds made while the DIIS error is >= 1e-2 (the first one included) are full
            builds with FP32 small-bound quartets at the Schwarz cutoff loose_cutoff (1e-8:
            ~1e-5 in F, 2.2x faster builds on (H2O)64 / cc-pVTZ);
          * then delta builds (G += G(P - P_prev), FP32 allowed, mixed_delta_* cutoffs)
            until the DIIS error of that Fock matrix is below switch_error (1e-4);
          * then a full FP64 build at the engine's cutoffs on the current density, which
            removes every loose, FP32 and delta error, and FP64 delta builds (delta_*
            cutoffs) to convergence.  With final_full (default False) a converged run
            makes one more full FP64 build and reports its energy."""
        if incremental is None:
            incremental = precision == "mixed"
        if precision == "mixed":
            if "loose_scale" not in kw:
                s = self.engine.cutoffs[0]
                kw["loose_scale"] = max(1.0, self.loose_cutoff / s) if s > 0 else 1.0
            kw.setdefault("switch_error", self.switch_error)
            kw.setdefault("final_full", self.final_full)
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        if guess == "mlp":
            P0 = self.mlp_guess(guess_model)
            if P0 is not None:
                # a learned density is already close: no damping or level shift of the
                # first iterations (they exist to tame far-off starting densities)
                kw.setdefault("damp_iters", 0)
                kw.setdefault("level_shift_iters", 0)
            else:
                P0 = self.guess("promolecule")
        else:
            P0 = self.guess(guess, guess_model)
        r = run_rhf(self.h, self.S, self.mol.n_electrons, self.mol.nuclear_repulsion(),
                    self.jk_incremental if incremental else self.jk, P0=P0,
                    mixed_until=mixed_until if precision == "mixed" else None,
                    incremental=incremental, log=log, **kw)
        torch.cuda.synchronize(self.device)
        r["wall"] = time.perf_counter() - t0
        return r

    def jk_u(self, Pa, Pb, mixed=False):
        """One UHF J/K build on the device (Tonto's make_u_JK_engine): (Ja, Jb, Ka, Kb)."""
        if not hasattr(self, "_Pb"):
            n = self.engine.n_bf
            self._Pb, self._Jb, self._Kb = (torch.empty((n, n), dtype=torch.float64, device=self.device)
                                            for _ in range(3))
        self._P.copy_(Pa)
        self._Pb.copy_(Pb)
        self.engine.set_stream(torch.cuda.current_stream(self.device).cuda_stream)
        self.engine.make_u_JK_device(self._P.data_ptr(), self._Pb.data_ptr(), self._J.data_ptr(),
                                     self._Jb.data_ptr(), self._K.data_ptr(), self._Kb.data_ptr(), mixed=mixed)
        return self._J.clone(), self._Jb.clone(), self._K.clone(), self._Kb.clone()

    def run_uhf(self, multiplicity=1, guess="promolecule", log=None, **kw):
        """UHF to convergence (tonto_amd.scf.run_uhf) with every J/K build on the device;
        n_alpha - n_beta = multiplicity - 1.  The promolecule guess gives each spin half
        of the summed atomic densities; "core" diagonalises h for both."""
        ne = self.mol.n_electrons
        n_unpaired = multiplicity - 1
        if n_unpaired < 0 or (ne - n_unpaired) % 2:
            raise ValueError(f"{ne} electrons cannot have multiplicity {multiplicity}")
        n_beta = (ne - n_unpaired) // 2
        n_alpha = n_beta + n_unpaired
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        Pa0 = Pb0 = None
        if guess == "promolecule":
            P = self.guess("promolecule")
            Pa0 = P * (n_alpha / ne)
            Pb0 = P * (n_beta / ne)
        r = run_uhf(self.h, self.S, n_alpha, n_beta, self.mol.nuclear_repulsion(), self.jk_u,
                    Pa0=Pa0, Pb0=Pb0, log=log, **kw)
        torch.cuda.synchronize(self.device)
        r["wall"] = time.perf_counter() - t0
        return r

"""Write the converged FP64 densities used as inputs by the parity tests.

Each fixture is the converged RHF density of a seeded water cluster, computed with the
oracle (oracle/oracle.c) and the SCF driver; J/K parity is then checked on the exact
density the screening sees at convergence.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tonto_amd.geometry import water_cluster  # noqa: E402
from tonto_amd.scf import run_rhf  # noqa: E402
import oracle.oracle as O  # noqa: E402

CASES = {
    # name: (n_water, basis, seed, Schwarz cutoff)
    "w8_631gd": (8, "6-31G(d)", 0, 1e-12),
    "w4_6311gdp": (4, "6-311G(d,p)", 0, 1e-9),
}


def make(name):
    n, b, seed, cut = CASES[name]
    mol = water_cluster(n, b, seed=seed)
    o = O.OracleSystem(mol)
    S, T, V = o.one_electron()
    maxI = o.schwarz()

    def jk(P, mode):
        J, K = o.make_r_JK(P.numpy(), schwarz=cut, max_I=maxI)
        return torch.from_numpy(J), torch.from_numpy(K)
    r = run_rhf(torch.from_numpy(T + V), torch.from_numpy(S), mol.n_electrons,
                mol.nuclear_repulsion(), jk, convergence=1e-9, diis_tol=1e-6)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), P=r.P.numpy(), energy=r.energy,
                        n_water=n, basis=b, seed=seed, schwarz=cut, iterations=r.iterations)
    print(name, mol.n_bf, r.energy, r.iterations)


if __name__ == "__main__":
    for k in (sys.argv[1:] or CASES):
        make(k)

"""Seeded water-cluster geometries for the direct-SCF workloads.

Water monomer: O-H 0.96 Angstrom, H-O-H 104.5 degrees.  Oxygen atoms sit on a simple
cubic lattice of 2.9 Angstrom spacing filled in raster order (x fastest, then y, then
z) over a box of nx x ny x nz sites; each molecule gets a uniformly random rotation
(a normalised Gaussian quaternion) and a rigid shift of its whole molecule drawn
uniformly from [-0.2, 0.2] Angstrom per Cartesian coordinate ("jitter per molecule").
Everything is drawn from numpy's PCG64 generator with the given seed, molecule by
molecule: 4 normals for the quaternion, then 3 uniforms for the shift.
"""
import math

import numpy as np

from .basis import Molecule

R_OH = 0.96
THETA_HOH = 104.5
SPACING = 2.9
JITTER = 0.2

BOXES = {8: (2, 2, 2), 32: (4, 4, 2), 64: (4, 4, 4)}


def _rotation(q):
    q = q / np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def monomer():
    h = math.radians(THETA_HOH) / 2
    return np.array([[0.0, 0.0, 0.0],
                     [R_OH * math.sin(h), R_OH * math.cos(h), 0.0],
                     [-R_OH * math.sin(h), R_OH * math.cos(h), 0.0]])


def box_for(n):
    if n in BOXES:
        return BOXES[n]
    if n <= 18:
        return (3, 3, 2)
    k = int(math.ceil(n ** (1.0 / 3.0)))
    return (k, k, k)


def water_cluster_xyz(n, seed=0, box=None):
    """Return (symbols, coords in Angstrom) for n waters."""
    nx, ny, nz = box or box_for(n)
    assert nx * ny * nz >= n
    rng = np.random.Generator(np.random.PCG64(seed))
    mono = monomer()
    symbols, xyz = [], []
    k = 0
    for iz in range(nz):
        for iy in range(ny):
            for ix in range(nx):
                if k == n:
                    break
                R = _rotation(rng.standard_normal(4))
                shift = rng.uniform(-JITTER, JITTER, 3)
                site = np.array([ix, iy, iz], dtype=np.float64) * SPACING + shift
                for s, r in zip(("O", "H", "H"), mono):
                    symbols.append(s)
                    xyz.append(site + R @ r)
                k += 1
    return symbols, np.array(xyz)


def water_cluster(n, basis_name, seed=0, box=None):
    symbols, xyz = water_cluster_xyz(n, seed, box)
    return Molecule.from_angstrom(symbols, xyz, basis_name)


def tonto_water(basis_name):
    """The single water of Tonto's `tests/short/h2o_rhf_*` jobs (Angstrom input)."""
    return Molecule.from_angstrom(
        ["O", "H", "H"],
        [[-0.028260, 0.000029, 0.0], [0.599917, 0.767685, 0.0], [0.599941, -0.767609, 0.0]],
        basis_name)


def cn_hcn_dimer(basis_name="cc-pVDZ"):
    """CN-...HCN of `tests/long/cn-hcn_dimer_rhf_cc-pVDZ_various_E_decompositions`."""
    return Molecule.from_angstrom(
        ["C", "N", "H", "C", "N"],
        [[-2.703329, 0, 0], [-1.543137, 0, 0], [0.344846, 0, 0], [1.439363, 0, 0],
         [2.577273, 0, 0]],
        basis_name, charge=-1)

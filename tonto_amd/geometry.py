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


def tonto_water(basis_name, charge=0):
    """The single water of Tonto's `tests/short/h2o_rhf_*` jobs (Angstrom input); the
    cation (charge 1) of scripts/check_dft_reference.py uses the same geometry."""
    return Molecule.from_angstrom(
        ["O", "H", "H"],
        [[-0.028260, 0.000029, 0.0], [0.599917, 0.767685, 0.0], [0.599941, -0.767609, 0.0]],
        basis_name, charge=charge)


def cn_hcn_dimer(basis_name="cc-pVDZ"):
    """CN-...HCN of `tests/long/cn-hcn_dimer_rhf_cc-pVDZ_various_E_decompositions`."""
    return Molecule.from_angstrom(
        ["C", "N", "H", "C", "N"],
        [[-2.703329, 0, 0], [-1.543137, 0, 0], [0.344846, 0, 0], [1.439363, 0, 0],
         [2.577273, 0, 0]],
        basis_name, charge=-1)


def _from_fractional(symbols, frac, cell_ang, basis_name):
    """Orthogonal cell: Cartesian Angstrom = fractional * cell edge."""
    xyz = np.asarray(frac, dtype=np.float64) * np.asarray(cell_ang, dtype=np.float64)
    return Molecule.from_angstrom(symbols, xyz, basis_name)


def tonto_nh3(basis_name="DZP"):
    """NH3 of `tests/short/nh3_rhf_DZP_HAF_and_structure_factors` (stdin: fractional
    coordinates in a 5.1305 Angstrom cubic cell); its first SCF is the isolated molecule
    (stdout:628-630, E = -56.2093, V_NN = 11.9664)."""
    return _from_fractional(["N", "H", "H", "H"],
                            [[0.2103, 0.2103, 0.2103], [0.3722, 0.2627, 0.1113],
                             [0.1113, 0.3722, 0.2627], [0.2627, 0.1113, 0.3722]],
                            [5.1305] * 3, basis_name)


def tonto_urea(basis_name="STO-3G"):
    """Urea of `tests/short/urea_rhf_STO-3G_HA_1e_properties` (stdin: fractional
    coordinates, tetragonal cell 5.5780 x 5.5780 x 4.6860 Angstrom; stdout: E = -221.0034,
    V_NN = 124.4944)."""
    return _from_fractional(["O", "N", "N", "C", "H", "H", "H", "H"],
                            [[0.0, 0.5, 0.596333], [0.144896, 0.644896, 0.178139],
                             [-0.144896, 0.355104, 0.178139], [0.0, 0.5, 0.328189],
                             [0.2557, 0.7557, 0.2841], [-0.2557, 0.2443, 0.2841],
                             [0.1431, 0.6431, -0.0348], [-0.1431, 0.3569, -0.0348]],
                            [5.5780, 5.5780, 4.6860], basis_name)

"""Pins of the oracle's screening and primitive-pair skip by what the mathematics fixes.

Tonto drops quartets by the row test (molecule.fock.foo:1412), the Schwarz test with the
largest density element (1430) and Schwarz_test (3497-3536: J with the larger of the two
pair density maxima, K with the four cross blocks, each against the density cutoff and
the product cutoff), and uses skipped primitive-pair lists (shell1quartet.foo:229-273)
only when the four atoms differ (molecule.fock.foo:1456-1472).  On (H2O)8 / 6-31G(d) at
its converged density (tests/golden/w8_631gd.npz), with the oracle's per-quartet
decisions and the exact max |(ab|cd)| of every quartet (full primitive lists):

* sound: every dropped J (K) part would have added, per integral-density product, less
  than the cutoff, |(ab|cd)| * max |P| over the J (K) blocks < cutoff, or its density is
  below the density cutoff.  The bound uses the integrals and the density themselves,
  not max_I or P_max; a correct Schwarz test guarantees it through the Schwarz inequality
  |(ab|cd)| <= max_I(ab) max_I(cd) (itself pinned in test_oracle_eri);
* tight: every digested J (K) part passes the tests as Tonto states them, recomputed here
  in numpy from the Schwarz table and the density in Tonto's multiplication order;
* the skipped lists are used exactly on quartets over four distinct atoms, and a pair's
  skipped list keeps exactly its primitive pairs with exp(-ab R^2/(a+b)) >= 1e-6;
* planted mistakes in the oracle (a wrong K block, the skip rule on the wrong atom
  pairs, the smaller J maximum) each fail one of these checks.
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def w8():
    from tonto_amd.geometry import water_cluster
    g = np.load(os.path.join(HERE, "golden", "w8_631gd.npz"))
    mol = water_cluster(8, "6-31G(d)", seed=int(g["seed"]))
    return mol, g["P"]


def _tri(x):
    """(x, y) with x >= y from a triangular index x(x+1)/2 + y."""
    x = np.asarray(x, dtype=np.int64)
    a = ((np.sqrt(8.0 * x + 1.0) - 1.0) / 2.0).astype(np.int64)
    a -= (a * (a + 1) // 2 > x)
    a += ((a + 1) * (a + 2) // 2 <= x)
    return a, x - a * (a + 1) // 2


def _pidx(x, y):
    hi, lo = np.maximum(x, y), np.minimum(x, y)
    return hi * (hi + 1) // 2 + lo


def _block_max(mol, P):
    """max |P| over every shell-pair block, Tonto pair order, from the density itself."""
    f = np.array([s.first_bf for s in mol.shells])
    n = np.array([s.n_comp for s in mol.shells])
    ns = len(f)
    out = np.zeros(ns * (ns + 1) // 2)
    A = np.abs(P)
    for a in range(ns):
        rows = A[f[a]:f[a] + n[a]]
        for b in range(a + 1):
            out[a * (a + 1) // 2 + b] = rows[:, f[b]:f[b] + n[b]].max()
    return out


def _audit(O, mol, P, cut, dens):
    o = O.OracleSystem(mol)
    mi = o.schwarz()
    code, maxabs = o.quartet_audit(P, mi, schwarz=cut, jp=dens, kp=dens)
    q = np.arange(code.size, dtype=np.int64)
    ab, cd = _tri(q)
    a, b = _tri(ab)
    c, d = _tri(cd)
    Pt = _block_max(mol, P)
    JP = np.maximum(Pt[ab], Pt[cd])
    KP = np.maximum(np.maximum(Pt[_pidx(a, c)], Pt[_pidx(a, d)]),
                    np.maximum(Pt[_pidx(b, c)], Pt[_pidx(b, d)]))
    return dict(o=o, mi=mi, code=code, maxabs=maxabs, ab=ab, cd=cd, a=a, b=b, c=c, d=d,
                JP=JP, KP=KP, Pt=Pt, cut=cut, dens=dens)


def _violations(mol, r):
    """Counts of quartets breaking each check (all zero for a correct screening)."""
    code, maxabs, cut, dens = r["code"], r["maxabs"], r["cut"], r["dens"]
    JP, KP, mi, Pt, ab, cd = r["JP"], r["KP"], r["mi"], r["Pt"], r["ab"], r["cd"]
    slack = 1.0 + 1e-12
    j_on, k_on = (code & 1) > 0, (code & 2) > 0
    # sound: what is dropped is below the cutoffs
    j_bad = ~j_on & ~(JP < dens) & ~(JP * maxabs < cut * slack)
    k_bad = ~k_on & ~(KP < dens) & ~(KP * maxabs < cut * slack)
    # tight: what is digested passes Tonto's tests (same operand order as Tonto)
    P_max, IP_max = Pt.max(), mi.max() * Pt.max()
    row = ~(mi[ab] * IP_max < cut)
    schw = ~(mi[ab] * mi[cd] * P_max < cut)
    j_pass = row & schw & ~((JP < dens) | (JP * mi[ab] * mi[cd] < cut))
    k_pass = row & schw & ~((KP < dens) | (KP * mi[ab] * mi[cd] < cut))
    j_extra = j_on != j_pass
    k_extra = k_on != k_pass
    # skipped lists on exactly the quartets over four distinct atoms
    at = np.array([s.atom for s in mol.shells])
    A, B, C, D = at[r["a"]], at[r["b"]], at[r["c"]], at[r["d"]]
    distinct = (A != B) & (C != D) & (A != C) & (A != D) & (B != C) & (B != D)
    skip_bad = (code > 0) & (((code & 4) > 0) != distinct)
    return dict(j_dropped_above_cutoff=int(j_bad.sum()), k_dropped_above_cutoff=int(k_bad.sum()),
                j_mismatch=int(j_extra.sum()), k_mismatch=int(k_extra.sum()),
                skip_lists_mismatch=int(skip_bad.sum()))


@pytest.mark.parametrize("cut,dens", [(1e-12, 1e-9), (1e-9, 1e-9), (1e-7, 1e-7)])
def test_screening_is_sound_and_tight(oracle_mod, w8, cut, dens):
    mol, P = w8
    r = _audit(oracle_mod, mol, P, cut, dens)
    v = _violations(mol, r)
    code = r["code"]
    counts = dict(dropped=int((code == 0).sum()), j_only=int(((code & 3) == 1).sum()),
                  k_only=int(((code & 3) == 2).sum()), skip_lists=int(((code & 4) > 0).sum()))
    print(cut, dens, counts, v)
    assert all(x == 0 for x in v.values()), v
    # not vacuous: every kind of decision occurs
    assert counts["dropped"] > 0 and counts["skip_lists"] > 0
    if cut >= 1e-9:
        assert counts["j_only"] > 0 and counts["k_only"] > 0
    # the build's counters agree with the audit's decisions
    _, _, st = r["o"].make_r_JK(P, schwarz=cut, jp=dens, kp=dens, max_I=r["mi"], return_stats=True)
    assert st["quartets"] == int((code > 0).sum())
    assert st["skip_J"] == int(((code > 0) & ((code & 1) == 0)).sum())
    assert st["skip_K"] == int(((code > 0) & ((code & 2) == 0)).sum())
    assert st["skip_lists"] == counts["skip_lists"]


def test_skipped_primitive_lists(oracle_mod, w8):
    """A pair on two atoms keeps in its skipped list exactly the primitive pairs with
    exp(-a b R^2 / (a+b)) >= 1e-6 (shell1quartet.foo:229-273); the full list keeps all."""
    mol, _ = w8
    o = oracle_mod.OracleSystem(mol)
    sh = mol.shells
    checked = dropped = 0
    for a in range(0, mol.n_shell, 3):
        for b in range(a + 1):
            A, B = mol.coords[sh[a].atom], mol.coords[sh[b].atom]
            r2 = float(np.sum((A - B) ** 2))
            ea, eb = np.asarray(sh[a].exps), np.asarray(sh[b].exps)
            ee = np.exp(-np.outer(eb, ea) * r2 / np.add.outer(eb, ea))
            full = len(ea) * len(eb)
            assert o.pair_prims(a, b, False) == full
            if sh[a].atom != sh[b].atom:
                keep = int((ee >= 1e-6).sum())
                assert o.pair_prims(a, b, True) == keep
                dropped += full - keep
            checked += 1
    assert checked > 1000 and dropped > 0


@pytest.mark.parametrize("plant,check", [(1, "k_dropped_above_cutoff"), (2, "skip_lists_mismatch"),
                                         (3, "j_dropped_above_cutoff")])
def test_planted_screening_mistakes_are_caught(oracle_mod, w8, plant, check):
    mol, P = w8
    oracle_mod.set_plant(plant)
    try:
        found = 0
        for cut, dens in ((1e-12, 1e-9), (1e-9, 1e-9)):
            v = _violations(mol, _audit(oracle_mod, mol, P, cut, dens))
            print(plant, cut, v)
            found += v[check]
    finally:
        oracle_mod.set_plant(0)
    assert found > 0

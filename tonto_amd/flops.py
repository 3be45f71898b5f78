"""Algorithmic work of the J/K build, frozen in one place.

Counts the reference algorithm (Rys quadrature + transfer + J/K digestion as in
foofiles/shell1quartet.foo), not the instructions a kernel happens to issue.  For a
quartet of class (la lb|lc ld) with L1 = la+lb, L2 = lc+ld, n = floor((L1+L2)/2)+1 Rys
roots, N_e = #Cartesian components with l in [max(la,lb), L1] (N_f likewise):

  per primitive quartet  F_p = 20 + 20 n + 9 n (L1+1)(L2+1) + 3 n N_e N_f
      (set-up, roots/weights as a fixed convention, 2-D recurrences, (e0|f0) assembly)
  per contracted quartet F_c = 2 H + 4 N_e N_f + 8 n_a n_b n_c n_d
      (H = values produced by the transfer relations, J digestion, K digestion)

The build's FLOPs are sum over class pairs of prim_quartets * F_p + quartets * F_c using
the per-class counters the library returns (tgpu_get_class_counters).
"""
from .basis import n_comp

LMAX = 3


def comps_range(l0, l1):
    return sum(n_comp(l) for l in range(l0, l1 + 1))


def hrr_outputs(lbig, lsm, lead):
    """Values written by the transfer relation from (e| (l in [lbig, lbig+lsm]) to (big, small),
    for `lead` independent rows."""
    return lead * sum(comps_range(lbig, lbig + lsm - k - 1) * n_comp(k + 1) for k in range(lsm))


def class_flops(la, lb, lc, ld):
    L1, L2 = la + lb, lc + ld
    n = (L1 + L2) // 2 + 1
    Ne = comps_range(max(la, lb), L1)
    Nf = comps_range(max(lc, ld), L2)
    na, nb, nc, nd = n_comp(la), n_comp(lb), n_comp(lc), n_comp(ld)
    Fp = 20 + 20 * n + 9 * n * (L1 + 1) * (L2 + 1) + 3 * n * Ne * Nf
    H = hrr_outputs(max(lc, ld), min(lc, ld), Ne) + hrr_outputs(max(la, lb), min(la, lb), nc * nd)
    Fc = 2 * H + 4 * Ne * Nf + 8 * na * nb * nc * nd
    return Fp, Fc


def class_pairs(lmax=LMAX):
    """Class pairs in the library's order (bra class >= ket class)."""
    pcs = [(lp, lq) for lp in range(lmax + 1) for lq in range(lp + 1)]
    return [pcs[X] + pcs[Y] for X in range(len(pcs)) for Y in range(X + 1)]


def build_flops(counters, lmax=LMAX):
    """counters: array (n_class_pairs, 3) of (quartets, prim quartets, fp32 quartets).
    Returns (total FLOPs, per-class FLOPs list)."""
    per = []
    for (la, lb, lc, ld), row in zip(class_pairs(lmax), counters):
        Fp, Fc = class_flops(la, lb, lc, ld)
        per.append(float(row[1]) * Fp + float(row[0]) * Fc)
    return sum(per), per

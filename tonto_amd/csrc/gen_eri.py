"""Generate straight-line HIP device code for contracted Cartesian ERI shell quartets.

For each class (LP LQ | LR LS) with LP >= LQ, LR >= LS the generator emits

    template <> struct EriBody<LP, LQ, LR, LS, R> { static __device__ void run(...); };

which, for one contracted quartet, accumulates the (e0|f0) integrals over primitive
quartets and Rys roots (2-D recurrences of Lindh, Ryu and Liu, the form used in
`foofiles/shell1quartet.foo:7420-7609`: e on the centre of the first (higher-l) shell
of the bra pair, f likewise on the ket), then applies the horizontal transfer
relation (a, b+1_t) = (a+1_t, b) + (A-B)_t (a, b) first on the ket side and then on the
bra side (`shell1quartet.foo:8310-11134`), writing (pq|rs) as double in the order
[p][q][r][s] with Tonto's Cartesian component order.

R is the arithmetic type of the recurrences, accumulation and transfer (double or
float); Rys roots and weights and the primitive prefactors are always FP64.

Every class also gets `WaveTab<LP, LQ, LR, LS>`: index tables driving the
wavefront-cooperative kernel (one 64-lane wave per quartet; 2-D tables and HRR levels
staged in LDS).  Classes with more than HEAVY_NI integrals (or (e0|f0) entries) get only
that form.

Output: gen/eri_<LP><LQ><LR><LS>.inc per class and gen/classes.inc, next to this file.
"""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))
from tonto_amd.basis import cart_components  # noqa: E402

LMAX = int(os.environ.get("TGPU_GEN_LMAX", "3"))
HEAVY_NI = 200


def ncart(l):
    return (l + 1) * (l + 2) // 2


def comps_range(l0, l1):
    out = []
    for l in range(l0, l1 + 1):
        out += cart_components(l)
    return out


def plus1(c, t):
    c = list(c)
    c[t] += 1
    return tuple(c)


def minus1(c, t):
    c = list(c)
    c[t] -= 1
    return tuple(c)


class Emitter:
    def __init__(self):
        self.lines = []

    def __call__(self, s, ind=1):
        self.lines.append("    " * ind + s)


def hrr_plan(lbig, lsm):
    """Symbolic HRR from (e| with |e| in [lbig, lbig+lsm] to (big, small).

    Returns (levels, outputs): levels is a list of lists of statements
    (dst_key, src_plus_key, src_key, axis) where keys are (big_comp, small_comp);
    level-0 keys are (e, ()) meaning the source array."""
    levels = []
    cur = {(e, (0, 0, 0)) for e in comps_range(lbig, lbig + lsm)}
    for k in range(lsm):
        nxt = []
        for s1 in cart_components(k + 1):
            t = next(i for i in range(3) if s1[i] > 0)
            s0 = minus1(s1, t)
            for a in comps_range(lbig, lbig + lsm - k - 1):
                nxt.append(((a, s1), (plus1(a, t), s0), (a, s0), t))
        levels.append(nxt)
    return levels


def gen_class(LP, LQ, LR, LS, R):
    LPQ, LRS = LP + LQ, LR + LS
    nroot = (LPQ + LRS) // 2 + 1
    E = comps_range(LP, LPQ)
    F = comps_range(LR, LRS)
    NE, NF = len(E), len(F)
    NP, NQ, NR, NS = ncart(LP), ncart(LQ), ncart(LR), ncart(LS)
    em = Emitter()
    em(f"template <> struct EriBody<{LP}, {LQ}, {LR}, {LS}, {R}> {{", 0)
    em("static constexpr int NP = %d, NQ = %d, NR = %d, NS = %d, NROOT = %d;" % (NP, NQ, NR, NS, nroot))
    em("static __device__ __forceinline__ void run(const DevPrims &pm, const PairView &bra, const PairView &ket,")
    em("                                           double *__restrict__ out, const double *__restrict__ rys_tab)")
    em("{")
    em(f"{R} G[{NE * NF}];", 2)
    em(f"#pragma unroll", 2)
    em(f"for (int k = 0; k < {NE * NF}; ++k) G[k] = ({R})0;", 2)
    # outer loop over the ket's primitives (per-lane data, read once each), inner loop over
    # the bra's primitives (the same for the whole wavefront when the caller's bra view
    # is uniform: ldc() then makes them scalar loads)
    em("for (int iq = 0; iq < ket.n; ++iq) {", 2)
    em("const int iqq = ket.off + iq;", 3)
    em("const PrimRec *__restrict__ pk = pm.rec + iqq;", 3)
    em("const double zc = pk->zeta, izc = pk->zinv, Kc = pk->K;", 3)
    em("const double Qx = pk->Px, Qy = pk->Py, Qz = pk->Pz;", 3)
    em(f"const {R} QCx = ({R})(Qx - ket.Ax), QCy = ({R})(Qy - ket.Ay), QCz = ({R})(Qz - ket.Az);", 3)
    em("for (int ip = 0; ip < bra.n; ++ip) {", 3)
    em("const int ipp = bra.off + ip;", 4)
    em("const PrimRec pb = ldc(pm.rec + ipp);", 4)
    em("const double za = pb.zeta, iza = pb.zinv, Ka = pb.K, Px = pb.Px, Py = pb.Py, Pz = pb.Pz;", 4)
    em(f"const {R} PAx = ({R})(Px - bra.Ax), PAy = ({R})(Py - bra.Ay), PAz = ({R})(Pz - bra.Az);", 4)
    em("const double izs = 1.0 / (za + zc), rho = za * zc * izs;", 4)
    em("const double QPx = Qx - Px, QPy = Qy - Py, QPz = Qz - Pz;", 4)
    em("const double X = rho * (QPx * QPx + QPy * QPy + QPz * QPz);", 4)
    em(f"double u[{nroot}], w[{nroot}];", 4)
    em(f"rys_roots_tab<{nroot}>(X, u, w, rys_tab);", 4)
    em("const double pref = TWOPI5ON2 * Ka * Kc * sqrt(rho);", 4)
    em("const double rz0 = zc * izs, re0 = za * izs;   // rho/zeta, rho/eta", 4)
    em(f"const {R} hz = ({R})(0.5 * iza), he = ({R})(0.5 * izc), hs = ({R})(0.5 * izs);", 4)
    em(f"const {R} qpx = ({R})QPx, qpy = ({R})QPy, qpz = ({R})QPz;", 4)
    em(f"#pragma unroll", 4)
    em(f"for (int r = 0; r < {nroot}; ++r) {{", 4)
    em(f"const {R} t2 = ({R})u[r];", 5)
    em(f"const {R} rz = ({R})(u[r] * rz0), re = ({R})(u[r] * re0);", 5)
    if LPQ > 1:
        em(f"const {R} b10 = (({R})1 - rz) * hz;", 5)
    if LRS > 1:
        em(f"const {R} b01 = (({R})1 - re) * he;", 5)
    if LPQ > 0 and LRS > 0:
        em(f"const {R} b00 = t2 * hs;", 5)
    for ax, PA, QC, qp, first in (("x", "PAx", "QCx", "qpx", "1"), ("y", "PAy", "QCy", "qpy", "1"),
                                  ("z", "PAz", "QCz", "qpz", f"({R})(w[r] * pref)")):
        # 2-D integrals I(e,f), e <= LPQ, f <= LRS
        if LPQ > 0:
            em(f"const {R} c00{ax} = {PA} + rz * {qp};", 5)
        if LRS > 0:
            em(f"const {R} d00{ax} = {QC} - re * {qp};", 5)
        em(f"const {R} {ax}_0_0 = {first};", 5)
        for e in range(LPQ):
            rhs = f"c00{ax} * {ax}_{e}_0"
            if e > 0:
                rhs += f" + ({R}){e} * b10 * {ax}_{e - 1}_0"
            em(f"const {R} {ax}_{e + 1}_0 = {rhs};", 5)
        for f in range(LRS):
            for e in range(LPQ + 1):
                rhs = f"d00{ax} * {ax}_{e}_{f}"
                if f > 0:
                    rhs += f" + ({R}){f} * b01 * {ax}_{e}_{f - 1}"
                if e > 0:
                    rhs += f" + ({R}){e} * b00 * {ax}_{e - 1}_{f}"
                em(f"const {R} {ax}_{e}_{f + 1} = {rhs};", 5)
    for ie, ce in enumerate(E):
        for jf, cf in enumerate(F):
            em(f"G[{ie * NF + jf}] += x_{ce[0]}_{cf[0]} * y_{ce[1]}_{cf[1]} * z_{ce[2]}_{cf[2]};", 5)
    em("}", 4)
    em("}", 3)
    em("}", 2)
    # ---- HRR on ket: G[e][f] -> T[e][r][s]
    Eidx = {c: i for i, c in enumerate(E)}
    Fidx = {c: i for i, c in enumerate(F)}
    em(f"const {R} CDx = ({R})ket.ABx, CDy = ({R})ket.ABy, CDz = ({R})ket.ABz;", 2)
    em(f"const {R} ABx = ({R})bra.ABx, ABy = ({R})bra.ABy, ABz = ({R})bra.ABz;", 2)
    axn = ["x", "y", "z"]

    def key_name(prefix, k):
        return prefix + "_" + "_".join("%d%d%d" % c for c in k)

    # variables for ket-level values: kv(e, rcomp, scomp)
    def ket_src(e, rc, sc):
        if sum(sc) == 0:
            return f"G[{Eidx[e] * NF + Fidx[rc]}]"
        return "k_%d%d%d_%d%d%d_%d%d%d" % (e + rc + sc)

    ket_levels = hrr_plan(LR, LS)
    for e in E:
        for lev in ket_levels:
            for (dst, sp, s0, t) in lev:
                a, s1 = dst
                em(f"const {R} {ket_src(e, a, s1)} = {ket_src(e, sp[0], sp[1])} + CD{axn[t]} * {ket_src(e, s0[0], s0[1])};", 2)

    # ---- HRR on bra: T[e][r][s] -> out[p][q][r][s]
    def bra_src(pc, qc, rc, sc):
        if sum(qc) == 0:
            return ket_src(pc, rc, sc)
        return "b_%d%d%d_%d%d%d_%d%d%d_%d%d%d" % (pc + qc + rc + sc)

    bra_levels = hrr_plan(LP, LQ)
    Rc = cart_components(LR)
    Sc = cart_components(LS)
    for rc in Rc:
        for sc in Sc:
            for lev in bra_levels:
                for (dst, sp, s0, t) in lev:
                    a, q1 = dst
                    em(f"const {R} {bra_src(a, q1, rc, sc)} = {bra_src(sp[0], sp[1], rc, sc)} + AB{axn[t]} * {bra_src(s0[0], s0[1], rc, sc)};", 2)
    Pc = cart_components(LP)
    Qc = cart_components(LQ)
    for ip, pc in enumerate(Pc):
        for iq, qc in enumerate(Qc):
            for ir, rc in enumerate(Rc):
                for is_, sc in enumerate(Sc):
                    idx = ((ip * NQ + iq) * NR + ir) * NS + is_
                    em(f"out[{idx}] = (double){bra_src(pc, qc, rc, sc)};", 2)
    em("}")
    em("};", 0)
    return "\n".join(em.lines)


def is_heavy(LP, LQ, LR, LS):
    NE = len(comps_range(LP, LP + LQ))
    NF = len(comps_range(LR, LR + LS))
    NI = ncart(LP) * ncart(LQ) * ncart(LR) * ncart(LS)
    return NI > HEAVY_NI or NE * NF > HEAVY_NI


def level_tables(lbig, lsm):
    """HRR levels as index tables.  Level 0 is the source vector over comps with l in
    [lbig, lbig+lsm]; each level k+1 entry is (src_plus_index, src_index, axis) into
    level k.  The last level is ordered big-major (big * ncart(lsm) + small)."""
    prev = {(e, (0, 0, 0)): i for i, e in enumerate(comps_range(lbig, lbig + lsm))}
    levels = []
    plan = hrr_plan(lbig, lsm)
    for k, lev in enumerate(plan):
        if k == len(plan) - 1:
            big = cart_components(lbig)
            sm = cart_components(lsm)
            order = {(a, b): ia * len(sm) + ib for ia, a in enumerate(big) for ib, b in enumerate(sm)}
            lev = sorted(lev, key=lambda st: order[st[0]])
        cur = {}
        ent = []
        for (dst, sp, s0, t) in lev:
            cur[dst] = len(ent)
            ent.append((prev[sp], prev[s0], t))
        levels.append(ent)
        prev = cur
    return levels


def gen_wave(LP, LQ, LR, LS):
    LPQ, LRS = LP + LQ, LR + LS
    nroot = (LPQ + LRS) // 2 + 1
    E = comps_range(LP, LPQ)
    F = comps_range(LR, LRS)
    NE, NF = len(E), len(F)
    NP, NQ, NR, NS = ncart(LP), ncart(LQ), ncart(LR), ncart(LS)
    W = LRS + 1
    gidx = []
    for ce in E:
        for cf in F:
            gidx += [ce[0] * W + cf[0], ce[1] * W + cf[1], ce[2] * W + cf[2]]
    cd = level_tables(LR, LS)
    ab = level_tables(LP, LQ)
    NRS = NR * NS
    buf = NE * NF
    prevn = NF
    for lev in cd:
        buf = max(buf, NE * len(lev))
    prevn = NE
    for lev in ab:
        buf = max(buf, len(lev) * NRS)
    buf = max(buf, NE * NRS)

    def arr(name, vals, typ="short"):
        if not vals:
            vals = [0]
        return f"    static constexpr {typ} {name}[{len(vals)}] = {{{', '.join(str(v) for v in vals)}}};"
    name = f"WaveTab<{LP}, {LQ}, {LR}, {LS}>"
    out = [f"template <> struct {name} {{",
           f"    static constexpr int NP = {NP}, NQ = {NQ}, NR = {NR}, NS = {NS}, NROOT = {nroot};",
           f"    static constexpr int LPQ = {LPQ}, LRS = {LRS}, NE = {NE}, NF = {NF}, BUF = {buf};",
           f"    static constexpr int NCD = {len(cd)}, NAB = {len(ab)};",
           arr("gidx", gidx),
           arr("cd_n", [len(l) for l in cd] or [0], "int"),
           arr("cd_off", [sum(len(x) for x in cd[:i]) for i in range(len(cd))] or [0], "int"),
           arr("cd_s1", [e[0] for l in cd for e in l]),
           arr("cd_s2", [e[1] for l in cd for e in l]),
           arr("cd_ax", [e[2] for l in cd for e in l], "char"),
           arr("ab_n", [len(l) for l in ab] or [0], "int"),
           arr("ab_off", [sum(len(x) for x in ab[:i]) for i in range(len(ab))] or [0], "int"),
           arr("ab_s1", [e[0] for l in ab for e in l]),
           arr("ab_s2", [e[1] for l in ab for e in l]),
           arr("ab_ax", [e[2] for l in ab for e in l], "char"),
           "};"]
    return "\n".join(out)


def classes(lmax):
    pcs = [(lp, lq) for lp in range(lmax + 1) for lq in range(lp + 1)]
    out = []
    for X in range(len(pcs)):
        for Y in range(X + 1):
            out.append(pcs[X] + pcs[Y])
    return pcs, out


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    gdir = os.path.join(here, "gen")
    os.makedirs(gdir, exist_ok=True)
    pcs, qcs = classes(LMAX)
    for (LP, LQ, LR, LS) in qcs:
        parts = ["/* Generated by gen_eri.py -- do not edit. */",
                 gen_wave(LP, LQ, LR, LS)]
        if not is_heavy(LP, LQ, LR, LS):
            for R in ("double", "float"):
                parts.append(gen_class(LP, LQ, LR, LS, R))
        with open(os.path.join(gdir, f"eri_{LP}{LQ}{LR}{LS}.inc"), "w") as f:
            f.write("\n\n".join(parts) + "\n")
    lines = ["/* Generated by gen_eri.py -- do not edit. */",
             f"#define TGPU_LMAX {LMAX}",
             f"#define TGPU_NPAIRCLASS {len(pcs)}",
             "/* M(LP, LQ, LR, LS, HEAVY) for every class pair, bra class >= ket class */",
             "#define TGPU_FOR_EACH_CLASS(M) \\\n" +
             " \\\n".join(f"    M({a}, {b}, {c}, {d}, {int(is_heavy(a, b, c, d))})" for (a, b, c, d) in qcs),
             "/* M(LP, LQ, HEAVY) for every shell-pair class (Schwarz kernels) */",
             "#define TGPU_FOR_EACH_PAIRCLASS(M) \\\n" +
             " \\\n".join(f"    M({a}, {b}, {int(is_heavy(a, b, a, b))})" for (a, b) in pcs)]
    with open(os.path.join(gdir, "classes.inc"), "w") as f:
        f.write("\n".join(lines) + "\n")
    # Cartesian -> real solid harmonic matrices of the spherical basis, l = 2 .. LMAX
    from tonto_amd.sph import solid_harmonic_tables
    T = solid_harmonic_tables(LMAX)
    out = ["/* Generated by gen_eri.py -- do not edit. */",
           "/* tgpu_sph_T<l>[c * (2l+1) + M]: Cartesian component c (Tonto order) of the real solid",
           "   harmonic M (Molden order 0, +1, -1, ...), tonto_amd/sph.py */"]
    for l in range(2, LMAX + 1):
        vals = ", ".join(repr(float(v)) for v in T[l].ravel())
        out.append(f"static __device__ const double tgpu_sph_T{l}[{T[l].size}] = {{{vals}}};")
    with open(os.path.join(gdir, "sph.inc"), "w") as f:
        f.write("\n".join(out) + "\n")
    print(f"wrote {gdir}: {len(qcs)} classes, heavy: "
          + " ".join(f"({a}{b}|{c}{d})" for (a, b, c, d) in qcs if is_heavy(a, b, c, d)))


if __name__ == "__main__":
    main()

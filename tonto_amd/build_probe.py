"""Per-class timing of one J/K build of a seeded water cluster (a profiling aid).

    python tonto_amd/build_probe.py [basis] [n_waters] [schwarz_cutoff] [scf_iterations] [mode]

Sets up the GPU SCF context, forms the promolecule guess, optionally runs a few FP64 SCF
iterations from it (the density then resembles the converged one, which screens far
fewer quartets away than the block-diagonal guess), runs two J/K builds at that density
with per-class HIP-event timing (mode "full", the default, or "full_mixed"), and prints
quartet counts (FP32 ones too), times and the algorithmic TFLOP/s (tonto_amd/flops.py) of
every class pair that has quartets.
"""
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main(argv):
    import torch
    from tonto_amd.driver import GPUSCF
    from tonto_amd.flops import build_flops, class_pairs
    from tonto_amd.geometry import water_cluster
    basis = argv[0] if len(argv) > 0 else "cc-pVTZ"
    nw = int(argv[1]) if len(argv) > 1 else 64
    cut = float(argv[2]) if len(argv) > 2 else 1e-11
    n_it = int(argv[3]) if len(argv) > 3 else 0
    mode = argv[4] if len(argv) > 4 else "full"
    mol = water_cluster(nw, basis)
    print(f"(H2O){nw} {basis}: {mol.n_bf} basis functions, {len(mol.shells)} shells", flush=True)
    t = time.perf_counter()
    s = GPUSCF(mol, schwarz=cut)
    torch.cuda.synchronize()
    print(f"setup {time.perf_counter() - t:.2f} s", flush=True)
    t = time.perf_counter()
    P0 = s.guess()
    torch.cuda.synchronize()
    print(f"guess {time.perf_counter() - t:.2f} s", flush=True)
    if n_it:
        from tonto_amd.scf import run_rhf
        t = time.perf_counter()
        r = run_rhf(s.h, s.S, mol.n_electrons, mol.nuclear_repulsion(), s.jk, P0=P0,
                    max_iterations=n_it, log=lambda m: print(m, flush=True))
        torch.cuda.synchronize()
        P0 = r.P
        print(f"{n_it} SCF iterations {time.perf_counter() - t:.1f} s, E = {r.energy:.8f}", flush=True)
    s.engine.set_profiling(True)
    for it in range(2):
        torch.cuda.synchronize()
        t = time.perf_counter()
        s.jk(P0, mode)
        torch.cuda.synchronize()
        print(f"build {it} {1e3 * (time.perf_counter() - t):.1f} ms", flush=True)
    st = s.engine.stats()
    print(st)
    ct = s.engine.class_times()
    cc = s.engine.class_counters()
    forms = s.engine.variants()
    tot, per = build_flops(cc)
    print(f"build flops {tot:.3e} -> {tot / (st['kernel_ms'] * 1e-3) / 1e12:.2f} TF")
    for k, cp in enumerate(class_pairs()):
        if cc[k][0]:
            print("(%d%d|%d%d) q=%11d fp32=%11d prim=%12d %9.3f ms %5.2f TF form=%s"
                  % (cp + (cc[k][0], cc[k][2], cc[k][1], ct[k], per[k] / max(ct[k], 1e-9) / 1e9,
                           "wave" if forms[k] else "lane")), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])

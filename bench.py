#!/usr/bin/env python3
"""Benchmark of the direct-SCF J/K hot path on one MI355X.

Workload (the headline): (H2O)64 RHF / cc-pVTZ with f on O, Cartesian (4160 basis
functions), water cluster from tonto_amd.geometry (seed 0), Schwarz cutoff 1e-11, J/K
density cutoffs 1e-9, primitive-pair cutoff 1e-6; quartets with max_I(ab) max_I(cd) < 1e-6
run in FP32 in the mixed builds.

1. Warm-up (not timed): promolecule guess (atomic SCFs) and one loose build of each
   precision, so both SCF runs start from loaded kernels and cached atomic densities.
2. FP64 SCF from the promolecule guess (full FP64 build every iteration) and the full
   form: mixed SCF (loose first builds, FP32 small-bound quartets, delta builds, FP64 from
   DIIS error 1e-4 on) from the learned guess (the 6-31G(d) network's density projected
   onto cc-pVTZ): wall times with making the guess included, iterations, energy
   difference.  The FP64 build time comes from the FP64 SCF.
3. Timed steps: one step = one mixed-precision J/K build at the converged density with P,
   J and K resident in HBM; W warm-up steps, then K timed steps between device
   synchronisations.  value = contracted shell quartets computed per second.
4. Roofline of the dominant class kernel (FP64 vector ALU bound) from per-class HIP
   events recorded on the library's stream during the timed steps and the algorithmic
   FLOPs of tonto_amd.flops.
5. CPU baseline: the oracle's FP64 J/K build (oracle/oracle.c) on a bounded sample of the
   same build (every k-th bra shell pair) at the converged density, on all host cores and
   on one.
6. Side legs (configs of the other workloads, small): (H2O)32 / 6-311G(d,p) FP64 vs mixed
   SCF, 256 batched clusters, the learned guess on held-out (H2O)20 clusters.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SCF wall-time & ERI quartets/s, mixed vs FP64 GPU at |ΔE|≤1e-8 Eh"
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector (and FP64 matrix), AMD published
N_WATER, BASIS, SEED = 64, "cc-pVTZ", 0
SCHWARZ, FP32_THR = 1e-11, 1e-6
# HBM bytes per build of the class pairs that can be the dominant kernel on this workload
# (sum over the class pair's launches of one mixed build), from rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes scaled by what the same counters read for a 256 MiB copy (x2.000 and
# x1.000): the profiles/ traffic summary of (H2O)64 / cc-pVTZ
TRAFFIC = {
    "jk class (21|21) mixed": {"bytes": 2.512824e+11, "read": 1.807234e+11, "written": 7.055901e+10, "dispatches": 6},
    "jk class (20|10) mixed": {"bytes": 9.951025e+11, "read": 8.881830e+11, "written": 1.069195e+11, "dispatches": 12},
    "jk class (10|00) mixed": {"bytes": 5.139057e+11, "read": 4.466742e+11, "written": 6.723149e+10, "dispatches": 16},
    "jk class (10|10) mixed": {"bytes": 1.251133e+12, "read": 1.139521e+12, "written": 1.116115e+11, "dispatches": 10},
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(mol, P, max_I, target_s, schwarz=SCHWARZ):
    """The oracle's FP64 J/K build on a bounded sample of the same build (every k-th bra
    shell pair, k chosen for about target_s seconds on all host cores), at the same
    cutoffs and density; timed on all cores and on one (the same sample)."""
    import oracle.oracle as O
    o = O.OracleSystem(mol)
    cores = O.max_threads()
    npair = mol.n_shell * (mol.n_shell + 1) // 2
    stride = max(1, npair // 64)
    t = time.perf_counter()
    o.make_r_JK_sample(P, stride, schwarz=schwarz, max_I=max_I)
    full_est = (time.perf_counter() - t) * stride
    stride = max(1, int(math.ceil(full_est / target_s)))
    # the cost per bra pair varies widely: refine the stride until the sample takes at
    # least half the target
    for _ in range(3):
        t = time.perf_counter()
        st = o.make_r_JK_sample(P, stride, schwarz=schwarz, max_I=max_I)
        dt = time.perf_counter() - t
        if dt >= 0.5 * target_s or stride == 1:
            break
        stride = max(1, int(stride * dt / target_s))
    # one thread on a sample a twentieth as large (every 20k-th pair of the same list)
    O.set_threads(1)
    try:
        t = time.perf_counter()
        st1 = o.make_r_JK_sample(P, 20 * stride, schwarz=schwarz, max_I=max_I)
        dt1 = time.perf_counter() - t
    finally:
        O.set_threads(cores)
    return {"value": st["quartets"] / dt, "unit": "quartets/s", "cores": cores, "kind": "port",
            "sample": f"FP64 J/K build of the same workload at the converged density restricted to "
                      f"every {stride}-th bra shell pair ({st['quartets']} quartets, "
                      f"{st['prim_quartets']} primitive quartets, {dt:.1f} s); oracle/oracle.c, "
                      f"{cores} pthreads",
            "one_thread": {"value": st1["quartets"] / dt1, "unit": "quartets/s", "cores": 1,
                           "sample": f"every {20 * stride}-th bra shell pair ({st1['quartets']} "
                                     f"quartets, {dt1:.1f} s)"},
            "full_build_estimate_s": dt * stride,
            "seconds": dt + dt1}


def batch_leg(steps, log_fn):
    """256 independent water clusters of 1-13 molecules (6-31G(d,p)), one J/K build each at
    the promolecule density: all molecules in one batched build (autotuned kernel forms,
    replayed hipGraph) against one build per molecule; builds/s and per-molecule max |dF|."""
    import torch
    from tonto_amd.driver import GPUSCF
    from tonto_amd.geometry import water_cluster
    from tonto_amd.gpu import JKEngine
    rng = np.random.Generator(np.random.PCG64(3))
    sizes = rng.integers(1, 14, size=256)
    mols = [water_cluster(int(k), "6-31G(d,p)", seed=1000 + i) for i, k in enumerate(sizes)]
    t0 = time.perf_counter()
    eng = JKEngine(mols)
    eng.schwarz()
    t_setup = time.perf_counter() - t0
    atom = GPUSCF(mols[0])
    dens = {sym: atom.atom_scf(sym).cpu().numpy() for sym in ("O", "H")}
    Ps = []
    for m in mols:
        blocks = [dens[s] for s in m.symbols]
        n = m.n_bf
        P = np.zeros((n, n))
        o = 0
        for b in blocks:
            P[o:o + b.shape[0], o:o + b.shape[0]] = b
            o += b.shape[0]
        Ps.append(P)
    Pd = torch.tensor(eng.pack(Ps), device="cuda")
    Jd, Kd = torch.empty_like(Pd), torch.empty_like(Pd)
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    ms, variants = eng.autotune(Pd.data_ptr())
    for graph in (False, True):
        eng.make_r_JK_device(Pd.data_ptr(), Jd.data_ptr(), Kd.data_ptr(), graph=graph)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.make_r_JK_device(Pd.data_ptr(), Jd.data_ptr(), Kd.data_ptr(), graph=True)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / steps
    quartets = eng.stats()["quartets"]
    Jb, Kb = eng.unpack(Jd.cpu().numpy()), eng.unpack(Kd.cpu().numpy())
    # one build per molecule, each in its own context
    singles = []
    for m, P in zip(mols, Ps):
        e = JKEngine(m)
        e.schwarz()
        Pm = torch.tensor(P, device="cuda")
        singles.append((e, Pm, torch.empty_like(Pm), torch.empty_like(Pm)))
    for e, Pm, Jm, Km in singles:
        e.set_stream(torch.cuda.current_stream().cuda_stream)
        e.make_r_JK_device(Pm.data_ptr(), Jm.data_ptr(), Km.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for e, Pm, Jm, Km in singles:
            e.make_r_JK_device(Pm.data_ptr(), Jm.data_ptr(), Km.data_ptr())
    torch.cuda.synchronize()
    t_single = (time.perf_counter() - t0) / steps
    dF = max(float(np.max(np.abs((Jb[i] - 0.5 * Kb[i]) - (Jm.cpu().numpy() - 0.5 * Km.cpu().numpy()))))
             for i, (e, Pm, Jm, Km) in enumerate(singles))
    log_fn(f"batch: 256 molecules ({int(sum(m.n_bf for m in mols))} bf total) "
           f"{t_batch * 1e3:.1f} ms per batched build, {t_single * 1e3:.1f} ms one by one; max dF {dF:.1e}")
    return {"workload": "256 water clusters, 1-13 molecules (seeded), 6-31G(d,p), one J/K "
                        "build each at the promolecule density",
            "builds_per_s": 256 / t_batch, "builds_per_s_one_by_one": 256 / t_single,
            "batched_ms": t_batch * 1e3, "quartets": quartets, "max_dF_vs_single": dF,
            "setup_s": t_setup, "autotuned_wave_forms": int(sum(variants))}


def guess_leg(n_clusters, log_fn):
    """Learned initial guess against the promolecule guess on held-out (H2O)20 / 6-31G(d)
    clusters: SCF iterations and wall time (guess construction included), converged
    energy difference."""
    from tonto_amd import guess_mlp as G
    from tonto_amd.guess_train import EVAL_SEED0, evaluate
    model = G.default_model()
    if model is None:
        return {"status": f"model file {os.path.relpath(G.DEFAULT_MODEL, ROOT)} absent"}
    r = evaluate(n_clusters, 20, model=model, log=log_fn)
    m = {k: float(np.mean(v)) for k, v in r.items() if k != "dE"}
    return {"workload": f"{n_clusters} held-out (H2O)20 clusters, RHF/6-31G(d) FP64, seeds "
                        f"{EVAL_SEED0}..{EVAL_SEED0 + n_clusters - 1}; model trained on "
                        f"{int(model.meta['n_clusters'])} clusters of "
                        f"{int(model.meta['size_min'])}-{int(model.meta['size_max'])} waters",
            "iterations_promolecule": m["sad_iterations"], "iterations_learned": m["mlp_iterations"],
            "scf_wall_promolecule_s": m["sad_wall"], "scf_wall_learned_s": m["mlp_wall"],
            "learned_guess_s": m["mlp_guess_seconds"],
            "max_abs_dE": float(np.max(np.abs(r["dE"]))), "dE_within_1e-8": bool(np.max(np.abs(r["dE"])) <= 1e-8)}


def config2_leg(log_fn):
    """(H2O)32 / 6-311G(d,p) (992 Cartesian functions), Schwarz 1e-9: FP64 SCF and mixed SCF
    wall times and energy difference, and the mixed build's max |dF| at the converged
    density against the FP64 build."""
    import torch
    from tonto_amd.driver import GPUSCF
    from tonto_amd.geometry import water_cluster
    mol = water_cluster(32, "6-311G(d,p)", seed=SEED)
    scf = GPUSCF(mol, schwarz=1e-9, fp32_threshold=FP32_THR)
    P0 = scf.guess()
    scf.jk(P0, "full"), scf.jk(P0, "full_mixed")
    r64 = scf.run(precision="fp64", convergence=1e-9, diis_tol=1e-6)
    rmx = scf.run(precision="mixed", convergence=1e-9, diis_tol=1e-6)
    J, K = scf.jk(r64.P, "full")
    Jm, Km = scf.jk(r64.P, "full_mixed")
    dF = float(torch.max(torch.abs((Jm - 0.5 * Km) - (J - 0.5 * K))).item())
    log_fn(f"(H2O)32 6-311G(d,p): FP64 SCF {r64.wall:.2f}s, mixed {rmx.wall:.2f}s, "
           f"dE {rmx.energy - r64.energy:+.1e}, mixed build dF {dF:.1e}")
    return {"workload": f"(H2O)32 RHF/6-311G(d,p) Cartesian, {mol.n_bf} basis functions, Schwarz 1e-9",
            "scf_wall_fp64_s": r64.wall, "scf_wall_mixed_s": rmx.wall,
            "scf_speedup_mixed": r64.wall / rmx.wall, "dE_mixed_vs_fp64": rmx.energy - r64.energy,
            "scf_iterations_fp64": r64.iterations, "scf_iterations_mixed": rmx.iterations,
            "mixed_build_max_dF_at_converged_density": dF}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--waters", type=int, default=N_WATER, help="cluster size (default 64)")
    ap.add_argument("--basis", default=BASIS, help="basis of the cluster (default cc-pVTZ; other "
                    "workloads are for profiling, the headline line is the default)")
    ap.add_argument("--schwarz", type=float, default=SCHWARZ, help="Schwarz cutoff (default 1e-11)")
    ap.add_argument("--density", default=None,
                    help="converged density (.npy) to time the steps at; skips both SCF runs")
    ap.add_argument("--save-density", default=None)
    ap.add_argument("--profile-iterations", type=int, default=0,
                    help="profiling runs: skip both timed SCFs, time the steps at the density "
                         "of this many FP64 SCF iterations from the promolecule guess")
    ap.add_argument("--no-side", action="store_true", help="skip the side legs (configs 2-4)")
    ap.add_argument("--guess-clusters", type=int, default=50,
                    help="held-out clusters of the learned-guess leg (0: skip)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (HBM traffic "
                         "calibration for rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE) and exit")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the J/K build would shard by bra shell pairs with one all-reduce of J and K; "
              "see DESIGN.md).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    from tonto_amd.driver import GPUSCF
    from tonto_amd.flops import build_flops, class_pairs
    from tonto_amd.geometry import water_cluster

    if not torch.cuda.is_available():
        sys.exit("bench.py needs a HIP device")
    if a.copy_calibration_mib:
        n = a.copy_calibration_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_mib": a.copy_calibration_mib, "dispatches": 3,
                          "bytes_read_per_dispatch": n * 8, "bytes_written_per_dispatch": n * 8}))
        return
    mol = water_cluster(a.waters, a.basis, seed=SEED)
    log(f"workload: (H2O){a.waters} {a.basis} n_bf={mol.n_bf} n_shell={mol.n_shell}")
    scf = GPUSCF(mol, schwarz=a.schwarz, fp32_threshold=FP32_THR)
    log(f"setup (1e ints + Schwarz) {scf.setup_s:.2f}s")
    extra = {}
    eng = scf.engine
    if a.profile_iterations:
        from tonto_amd.scf import run_rhf
        r = run_rhf(scf.h, scf.S, mol.n_electrons, mol.nuclear_repulsion(), scf.jk, P0=scf.guess(),
                    max_iterations=a.profile_iterations, log=lambda s: log("profile " + s))
        P = r.P.contiguous()
        if a.save_density:
            np.save(a.save_density, P.cpu().numpy())
    elif a.density is None:
        # warm-up for both runs alike: atomic densities cached, every kernel loaded (the
        # learned guess's network and projection included)
        t0 = time.perf_counter()
        P0 = scf.guess()
        scf.jk_scaled(P0, "full", 1e3, 1.0)
        scf.jk_scaled(P0, "full_mixed", 1e3, 1.0)
        learned = scf.mlp_guess() is not None
        torch.cuda.synchronize()
        log(f"warm-up {time.perf_counter() - t0:.1f}s")
        r64 = scf.run(precision="fp64", convergence=1e-9, diis_tol=1e-6,
                      log=lambda s: log("fp64 " + s))
        # the full form: mixed precision, incremental builds and the learned guess
        rmx = scf.run(precision="mixed", convergence=1e-9, diis_tol=1e-6,
                      guess="mlp" if learned else "promolecule", log=lambda s: log("mixed " + s))
        dE = rmx.energy - r64.energy
        n64 = r64.n_builds["full"]
        extra.update({
            "scf_energy_fp64": r64.energy, "scf_energy_mixed": rmx.energy,
            "dE_mixed_vs_fp64": dE, "dE_within_1e-8": bool(abs(dE) <= 1e-8),
            "scf_wall_fp64_s": r64.wall, "scf_wall_mixed_s": rmx.wall,
            "scf_speedup_mixed": r64.wall / rmx.wall,
            "scf_iterations_fp64": r64.iterations, "scf_iterations_mixed": rmx.iterations,
            "scf_builds_mixed": rmx.n_builds, "build_seconds_mixed": rmx.build_seconds,
            "build_seconds_fp64": r64.build_seconds,
            "fp64_build_s_mean_in_scf": r64.build_seconds["full"] / max(1, n64),
            "fp64_build_s_last_in_scf": r64.history[-1]["build_seconds"],
            "scf_converged": bool(r64.converged and rmx.converged),
            "guess_fp64": "promolecule (atomic SCF densities)",
            "guess_mixed": ("learned (6-31G(d) network, projected onto this basis)" if learned
                            else "promolecule (no learned model)"),
            "mixed_schedule": {"loose_cutoff": scf.loose_cutoff, "switch_diis_error": scf.switch_error,
                               "mixed_delta_cutoffs": [scf.mixed_delta_cutoff, scf.mixed_delta_density_cutoff],
                               "fp64_delta_cutoffs": [scf.delta_cutoff, scf.delta_density_cutoff]}})
        log(f"SCF fp64 E={r64.energy:.10f} {r64.wall:.2f}s it={r64.iterations}; "
            f"mixed E={rmx.energy:.10f} {rmx.wall:.2f}s it={rmx.iterations} dE={dE:.2e} "
            f"speedup {r64.wall / rmx.wall:.2f}")
        P = r64.P.contiguous()
        if a.save_density:
            np.save(a.save_density, P.cpu().numpy())
    else:
        P = torch.tensor(np.load(a.density), device="cuda")

    eng.set_profiling(True)
    ncls = len(class_pairs())

    def timed(mode, steps, warmup):
        for _ in range(warmup):
            scf.jk(P, mode)
        torch.cuda.synchronize()
        times = np.zeros(ncls)
        t0 = time.perf_counter()
        for _ in range(steps):
            scf.jk(P, mode)
            times += eng.class_times()        # HIP events of this step (synchronises)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return wall, eng.stats(), eng.class_counters(), times / max(1, steps)

    wallmx, stmx, cntmx, tmx = timed("full_mixed", a.steps, a.warmup)
    st64 = None
    if a.density is not None or a.profile_iterations:
        # no FP64 SCF ran: one FP64 build for the comparison
        _, st64, cnt64, t64 = timed("full", 1, 0)
    q = stmx["quartets"]
    value = q * a.steps / wallmx
    fmx, pmx = build_flops(cntmx)
    for kk, cp in enumerate(class_pairs()):
        if cntmx[kk][0]:
            log("class (%d%d|%d%d) quartets %11d prim %12d fp32 %11d | mixed %9.3f ms %6.2f TF"
                % (*cp, cntmx[kk][0], cntmx[kk][1], cntmx[kk][2], tmx[kk], pmx[kk] / max(tmx[kk], 1e-9) / 1e9))
    ms_step = wallmx / a.steps * 1e3
    log(f"mixed build {ms_step:.1f} ms ({fmx / (ms_step * 1e-3) / 1e12:.2f} TF algorithmic)")
    k = int(np.argmax(tmx))
    la, lb, lc, ld = class_pairs()[k]
    achieved = pmx[k] / (tmx[k] * 1e-3) / 1e12
    forms = eng.variants()
    roof = {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
            "kernel": f"jk class ({la}{lb}|{lc}{ld}) mixed, {'wavefront' if forms[k] else 'lane'} form",
            "kernel_ms_per_build": float(tmx[k]), "kernel_flops_per_build": pmx[k],
            "fp32_quartet_share": float(cntmx[k][2]) / max(1.0, float(cntmx[k][0])),
            "build_tflops": fmx / (ms_step * 1e-3) / 1e12,
            "build_frac": fmx / (ms_step * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
            "note": "FP64 vector-ALU bound path (Rys recurrences, (e0|f0) assembly and transfer are "
                    "FMA chains; no MFMA); MI355X FP64 vector peak 78.6 TFLOP/s; FP32 quartets counted "
                    "against the FP64 peak (non-packed FP32 issues at the FP64 rate); FLOPs by "
                    "tonto_amd/flops.py; kernel time = HIP events around the class pair's launches on "
                    "the library's stream"}
    tr = TRAFFIC.get(roof["kernel"].split(",")[0])
    if tr is not None and mol.n_bf == 4160:
        # per build, like "achieved" (the class pair's launches of one build)
        roof["traffic"] = tr["bytes"]
        roof["traffic_read_written"] = [tr["read"], tr["written"]]
        roof["traffic_per_dispatch"] = tr["bytes"] / tr["dispatches"]
        roof["traffic_note"] = ("HBM-side bytes per build from the rocprofv3 FETCH_SIZE/WRITE_SIZE "
                                "profile of this workload (memory-side requests, mostly served by the "
                                "Infinity Cache): the kets' primitive-pair records and pair data, the "
                                "density and the J/K atomics; algorithmic bytes per build are a few MB")
    cfg = {"workload": f"(H2O){a.waters} RHF/{a.basis} Cartesian, {mol.n_bf} basis functions, "
                       f"seeded cubic-lattice cluster, Schwarz {a.schwarz:g}; one step = one J/K build at the "
                       f"converged density, FP32 below Schwarz bound {FP32_THR:g}",
           "n_bf": mol.n_bf, "schwarz_cutoff": a.schwarz,
           "quartets_per_step": q, "prim_quartets_per_step": stmx["prim_quartets"],
           "fp32_quartets_per_step": stmx["fp32_quartets"],
           "build_tflops_mixed": fmx / (ms_step * 1e-3) / 1e12, **extra}
    if "fp64_build_s_last_in_scf" in extra:
        # the FP64 SCF's last full build ran at the converged density: same quartet set
        cfg["fp64_quartets_per_s"] = q / extra["fp64_build_s_last_in_scf"]
    if st64 is not None:
        cfg["fp64_build_ms"] = float(np.sum(t64))
        cfg["fp64_quartets_per_s"] = st64["quartets"] / (np.sum(t64) * 1e-3)
    out = {"metric": METRIC, "value": value, "unit": "quartets/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms_step,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "fp64+fp32", "data": "synthetic",
           "config": cfg, "roofline": roof}
    if not a.no_side and a.density is None and not a.profile_iterations:
        out["config"]["config2_h2o32_6311gdp"] = config2_leg(log)
        out["config"]["batched_builds"] = batch_leg(a.steps, log)
        if a.guess_clusters > 0:
            out["config"]["learned_guess"] = guess_leg(a.guess_clusters, log)
    if not a.no_cpu:
        cb = cpu_baseline(mol, P.cpu().numpy(), scf.max_I, a.cpu_seconds, a.schwarz)
        log(f"cpu baseline {cb['value']:.3e} quartets/s on {cb['cores']} threads, "
            f"{cb['one_thread']['value']:.3e} on one ({cb['seconds']:.1f}s)")
        cb.pop("seconds")
        out["cpu_baseline"] = cb
        if out["value"]:
            out["config"]["gpu_over_cpu_quartets_per_s"] = out["value"] / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

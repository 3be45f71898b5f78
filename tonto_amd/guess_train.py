"""Training and evaluation of the learned initial guess (runs on the GPU).

Data: converged RHF/6-31G(d) densities (FP64 GPU SCF from the promolecule guess, energy
converged to 1e-8 Eh) of seeded water clusters with 4-16 molecules (sizes drawn
uniformly, seeds seed0, seed0 + 1, ...), turned into per-block training rows by
tonto_amd.guess_mlp.training_rows.

Training: one network per block type (n_in 273 -> 128 -> 128 -> n_out, SiLU, FP32),
inputs standardised per feature, targets divided by their RMS; Adam with a cosine
learning-rate decay, minibatches drawn on the device; every forward, loss, backward and
update step is a tmlp_* HIP kernel.  5% of each type's rows are held out for the
validation loss that is printed.

Evaluation: held-out clusters (20 molecules, seeds from 900000): the same FP64 SCF from
the promolecule guess and from the learned guess, iterations, wall time and the
difference of the converged energies.

    python -m tonto_amd.guess_train train --clusters 2000 --out tonto_amd/data/guess_mlp_631gd.npz
    python -m tonto_amd.guess_train eval --clusters 200
"""
import argparse
import math
import sys
import time

import numpy as np

from . import guess_mlp as G

BASIS = "6-31G(d)"
EVAL_SEED0 = 900000


def _log(*a):
    print(*a, flush=True)


def _generate_part(args):
    """Training rows of the clusters `idx` (indices into the seeded size list), in a
    worker process of its own: (rows, SCF iterations, skipped indices); prints its
    progress every 25 clusters."""
    idx, sizes, seed0 = args
    import torch
    from .driver import GPUSCF
    from .geometry import water_cluster
    rows = {t: ([], []) for t in G.TYPES}
    its, skipped = [], []
    atom = None
    t0 = time.perf_counter()
    for k, i in enumerate(idx):
        if k and k % 25 == 0:
            print(f"data: worker at cluster {k}/{len(idx)}, {time.perf_counter() - t0:.0f}s", flush=True)
        mol = water_cluster(int(sizes[i]), BASIS, seed=seed0 + int(i))
        s = GPUSCF(mol)
        if atom is None:
            atom = {sym: s.atom_density(sym) for sym in ("O", "H")}
        r = s.run(guess="promolecule", convergence=1e-8)
        if not r.converged:
            skipped.append(int(i))
            continue
        its.append(r.iterations)
        for t, (x, y) in G.training_rows(mol, r.P.cpu().numpy(), atom).items():
            rows[t][0].append(x.astype(np.float32))
            rows[t][1].append(y.astype(np.float32))
        del s
    torch.cuda.synchronize()
    out = {t: (np.concatenate(x), np.concatenate(y)) for t, (x, y) in rows.items() if x}
    return out, its, skipped, atom


def generate(n_clusters, nmin, nmax, seed0, workers=1, log=_log):
    """Converged densities of n_clusters seeded clusters turned into training rows.  With
    workers > 1 the clusters are dealt out to that many processes sharing the GPU (each
    cluster's data depends on its seed only, so the rows are the same; they are merged
    in cluster order per worker)."""
    rng = np.random.default_rng(seed0)
    sizes = rng.integers(nmin, nmax + 1, n_clusters)
    t0 = time.perf_counter()
    parts = [np.arange(w, n_clusters, workers) for w in range(workers)]
    if workers > 1:
        import multiprocessing as mp
        with mp.get_context("spawn").Pool(workers) as pool:
            res = pool.map(_generate_part, [(p, sizes, seed0) for p in parts])
    else:
        res = [_generate_part((parts[0], sizes, seed0))]
    rows = {t: ([], []) for t in G.TYPES}
    its, skipped, atom = [], [], None
    for out, it, sk, at in res:
        its += it
        skipped += sk
        atom = atom or at
        for t, (x, y) in out.items():
            rows[t][0].append(x)
            rows[t][1].append(y)
    for i in skipped:
        log(f"cluster {i} (n={sizes[i]}) did not converge; skipped")
    log(f"data {n_clusters} clusters on {workers} worker(s) {time.perf_counter() - t0:.1f}s "
        f"(mean SCF iterations {np.mean(its):.1f})")
    data = {t: (np.concatenate(x), np.concatenate(y)) for t, (x, y) in rows.items() if x}
    return data, atom, sizes, time.perf_counter() - t0


def train_net(t, X, Y, epochs, batch, lr, seed, log=_log):
    import torch
    from .mlp import DeviceNet, init_params
    dev = torch.device("cuda")
    n_out = Y.shape[1]
    x_mean = X.mean(0)
    x_std = X.std(0)
    x_std[x_std < 1e-6] = 1.0
    y_scale = float(np.sqrt(np.mean(Y.astype(np.float64) ** 2))) or 1.0
    rng = np.random.default_rng(seed)
    perm = rng.permutation(X.shape[0])
    n_val = max(1, X.shape[0] // 20)
    val, tr = perm[:n_val], perm[n_val:]
    Xn = ((X - x_mean) / x_std).astype(np.float32)
    Yn = (Y / y_scale).astype(np.float32)
    Xt, Yt = torch.tensor(Xn[tr], device=dev), torch.tensor(Yn[tr], device=dev)
    Xv, Yv = torch.tensor(Xn[val], device=dev), torch.tensor(Yn[val], device=dev)
    batch = min(batch, Xt.shape[0])
    steps = Xt.shape[0] // batch
    net = DeviceNet(G.N_IN, G.NH, n_out, init_params(G.N_IN, G.NH, n_out, seed), dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    hist = []
    for ep in range(epochs):
        lr_ep = 0.02 * lr + 0.98 * lr * 0.5 * (1.0 + math.cos(math.pi * ep / epochs))
        order = torch.randperm(Xt.shape[0], device=dev, generator=gen)
        tot = torch.zeros((), device=dev)
        for k in range(steps):
            idx = order[k * batch:(k + 1) * batch]
            xb = Xt.index_select(0, idx).contiguous()
            yb = Yt.index_select(0, idx).contiguous()
            tot = tot + net.loss_and_grad(xb, yb, 1.0 / (batch * n_out))[0]
            net.adam(lr_ep)
        pv = net.forward(Xv)
        v = float(((pv - Yv) ** 2).mean())
        hist.append((float(tot) / max(steps, 1), v))
        if ep % 10 == 0 or ep + 1 == epochs:
            log(f"{t:2s} epoch {ep:3d} lr {lr_ep:.2e} train {hist[-1][0]:.4e} val {v:.4e}")
    return dict(params=net.params(), x_mean=x_mean.astype(np.float32), x_std=x_std.astype(np.float32),
                y_scale=np.float32(y_scale), n_out=n_out), hist


def train(args):
    t0 = time.perf_counter()
    data, atom, sizes, t_data = generate(args.clusters, args.min, args.max, args.seed0, args.workers)
    _log("rows: " + ", ".join(f"{t} {X.shape[0]}x{X.shape[1]}->{Y.shape[1]}" for t, (X, Y) in data.items()))
    nets, val = {}, {}
    t1 = time.perf_counter()
    for k, (t, (X, Y)) in enumerate(data.items()):
        nets[t], hist = train_net(t, X, Y, args.epochs, args.batch, args.lr, args.seed0 + k)
        val[t] = hist[-1][1]
    t_train = time.perf_counter() - t1
    meta = dict(basis=BASIS, n_clusters=args.clusters, size_min=args.min, size_max=args.max,
                seed0=args.seed0, epochs=args.epochs, batch=args.batch, lr=args.lr,
                data_seconds=t_data, train_seconds=t_train,
                val_loss=np.array([val[t] for t in G.TYPES if t in val]),
                val_types=np.array([t for t in G.TYPES if t in val]))
    G.GuessModel(nets, meta).save(args.out)
    _log(f"saved {args.out}: data {t_data:.0f}s, training {t_train:.0f}s, total {time.perf_counter() - t0:.0f}s")


def evaluate(n_clusters, n_waters=20, seed0=EVAL_SEED0, model=None, log=_log):
    """SCF iterations and wall time (making the guess included) from the promolecule and
    the learned guess on held-out clusters, after a warm-up shared by both and in
    alternating order; the converged energies must agree (the guess only moves the
    iteration count)."""
    import torch
    from .driver import GPUSCF
    from .geometry import water_cluster
    model = model or G.default_model()
    if model is None:
        raise FileNotFoundError(G.DEFAULT_MODEL)
    out = {"sad_iterations": [], "mlp_iterations": [], "sad_wall": [], "mlp_wall": [], "dE": [],
           "mlp_guess_seconds": []}
    for i in range(n_clusters):
        mol = water_cluster(n_waters, BASIS, seed=seed0 + i)
        s = GPUSCF(mol)
        # warm-up shared by both runs: atomic densities, the build's kernels, the network
        s.atom_density("O"), s.atom_density("H")
        s.jk(s.guess(), "full")
        s.guess("mlp", model)
        torch.cuda.synchronize()
        t = time.perf_counter()
        s.guess("mlp", model)
        torch.cuda.synchronize()
        tg = time.perf_counter() - t
        # the two runs in alternating order
        if i % 2 == 0:
            r0 = s.run(guess="promolecule")
            r1 = s.run(guess="mlp", guess_model=model)
        else:
            r1 = s.run(guess="mlp", guess_model=model)
            r0 = s.run(guess="promolecule")
        out["sad_iterations"].append(r0.iterations)
        out["mlp_iterations"].append(r1.iterations)
        out["sad_wall"].append(r0.wall)
        out["mlp_wall"].append(r1.wall)
        out["mlp_guess_seconds"].append(tg)
        out["dE"].append(r1.energy - r0.energy)
        if (i + 1) % 10 == 0 or i + 1 == n_clusters:
            log(f"eval {i + 1}/{n_clusters}: iterations promolecule {np.mean(out['sad_iterations']):.2f} "
                f"learned {np.mean(out['mlp_iterations']):.2f}; max|dE| {np.max(np.abs(out['dE'])):.1e}")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("train")
    a.add_argument("--clusters", type=int, default=2000)
    a.add_argument("--min", type=int, default=4)
    a.add_argument("--max", type=int, default=16)
    a.add_argument("--seed0", type=int, default=1000)
    a.add_argument("--epochs", type=int, default=60)
    a.add_argument("--batch", type=int, default=2048)
    a.add_argument("--lr", type=float, default=2e-3)
    a.add_argument("--out", default=G.DEFAULT_MODEL)
    a.add_argument("--workers", type=int, default=4,
                   help="processes generating the data (sharing the GPU)")
    e = sub.add_parser("eval")
    e.add_argument("--clusters", type=int, default=200)
    e.add_argument("--waters", type=int, default=20)
    args = ap.parse_args(argv)
    if args.cmd == "train":
        train(args)
    else:
        r = evaluate(args.clusters, args.waters)
        print({k: float(np.mean(np.abs(v)) if k == "dE" else np.mean(v)) for k, v in r.items()})


if __name__ == "__main__":
    sys.exit(main())

"""Training pipeline of the learned SGS closure, end to end on one GPU.

1. a channel flow between smooth walls on a DNS-resolved grid (default
   256×256×128, Re_τ = 180 forcing, float64) advanced by the engine from a
   parabolic profile plus a smooth broadband perturbation;
2. every `--every` steps, box-filtered samples of the state (filter width
   `--filter` cells, the engine's `bld_closure_sample`);
3. training on the GPU on all but the last snapshot (`closure.train`);
4. the a-priori comparison with Smagorinsky on the held-out last snapshot
   (`closure.apriori`), printed as one JSON line; the trained parameters are
   written with numpy.save when `--out` is given.

Usage: python -m <package>.train_closure [--res 256,256,128] [--snapshots 2000] ...
(or run this file directly).
"""

import argparse
import json
import os
import sys
import time

import numpy as np


def _package():
    here = os.path.dirname(os.path.abspath(__file__))
    if __package__:
        import importlib
        return importlib.import_module(__package__)
    import importlib.util
    spec = importlib.util.spec_from_file_location("bld_amd", os.path.join(here, "__init__.py"),
                                                  submodule_search_locations=[here])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["bld_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def generate(B, res, precision, dt, warmup, snapshots, every, F, seed=5, log=None, keep=None):
    """DNS run and box-filtered samples: lists of (x, t) per snapshot, each
    but the last reduced to a seeded random subset of `keep` samples (the
    data set of thousands of snapshots then stays in host memory; the last,
    held-out snapshot is kept whole)."""
    W = B.workloads
    n, dom, procs = W.dns_channel(B, res)
    m = B.Model(n, dom, procs, precision=precision)
    zc, zi = m.coordinates("vel1", 3), m.coordinates("vel3", 3)
    n1, n2 = m.default_size()
    u, v, w = W.dns_initial(zc, zi, n1, n2, seed=seed, uc=15.0, amp=0.0)
    m.set_physical("vel1", u)
    m.set_physical("vel2", v)
    m.set_physical("vel3", w)
    st, noise = m.state, W.spectral_noise(m, seed=seed, amp=0.25, k0=4.0)
    m.set_state(**{f: st[f] + noise[f] for f in st})
    m.init_problem()
    t0 = time.perf_counter()
    m.step(dt, warmup)
    t_warm = time.perf_counter() - t0
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    xs, ts = [], []
    for i in range(snapshots):
        m.step(dt, every)
        x, t, _ = m.sample_closure(F)
        if keep is not None and i < snapshots - 1 and keep < x.shape[0]:
            sel = rng.choice(x.shape[0], size=keep, replace=False)
            x, t = x[sel], t[sel]
        xs.append(x)
        ts.append(t)
        if log and (i + 1) % max(1, snapshots // 10) == 0:
            log(f"snapshot {i + 1}/{snapshots}: {x.shape[0]} samples, {time.perf_counter() - t0:.1f} s")
    steps = warmup + snapshots * every
    return xs, ts, time.perf_counter() - t0, dict(steps=steps, warmup_steps_per_s=round(warmup / max(t_warm, 1e-9), 2))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--res", default="256,256,128")
    ap.add_argument("--precision", default="fp64", choices=("fp64", "fp32"))
    ap.add_argument("--dt", type=float, default=2e-4)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--snapshots", type=int, default=2000)
    ap.add_argument("--every", type=int, default=5)
    ap.add_argument("--keep", type=int, default=2048, help="samples kept per training snapshot")
    ap.add_argument("--filter", type=int, default=4)
    ap.add_argument("--train-steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    B = _package()
    res = tuple(int(v) for v in a.res.split(","))
    log = lambda *s: print("[train_closure]", *s, file=sys.stderr, flush=True)
    xs, ts, sim_s, info = generate(B, res, a.precision, a.dt, a.warmup, a.snapshots, a.every, a.filter, log=log,
                                   keep=a.keep)
    if not all(np.all(np.isfinite(t)) for t in ts):
        raise RuntimeError("non-finite samples: the DNS run diverged")
    xtr, ttr = np.concatenate(xs[:-1]), np.concatenate(ts[:-1])
    t0 = time.perf_counter()
    params, losses = B.closure.train(xtr, ttr, steps=a.train_steps, batch=a.batch, lr=a.lr)
    train_s = time.perf_counter() - t0
    res_ap = B.closure.apriori(params, xs[-1], ts[-1])
    if a.out:
        np.save(a.out, B.closure.pack(params))
    print(json.dumps(dict(dns=dict(res=list(res), precision=a.precision, dt=a.dt, snapshots=a.snapshots,
                                   every=a.every, seconds=round(sim_s, 2), **info),
                          samples_train=int(xtr.shape[0]), samples_heldout=int(xs[-1].shape[0]),
                          filter_cells=a.filter, train_steps=a.train_steps, train_seconds=round(train_s, 2),
                          loss_first=float(losses[:10].mean()), loss_last=float(losses[-10:].mean()),
                          apriori_heldout=res_ap)), flush=True)


if __name__ == "__main__":
    main()

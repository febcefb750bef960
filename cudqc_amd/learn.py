"""Training data and training of the cost model behind DQC_SCHED_LEARNED.

Task logs come from the enumerator itself (dqc_set_task_log): one record per expanded
task with its features and the branches it ran. Two targets:
  own      log(1 + branches of the task's own loop): what a task costs inside the launch
           that runs it, the quantity batch ordering and splitting balance (default; every
           logged task is a sample);
  subtree  log(1 + branches over the task and all its descendants): known exactly once no
           descendant is left on the task stack, so only those tasks become samples.
The schedule multiplies the prediction by the task size (each branch scans lists of the
task's vertices).

Usage (on a GPU): python -m cudqc_amd.learn --out cudqc_amd/models/cost_mlp.f32
"""
import argparse
import os
import sys
import time

import numpy as np

import cudqc_amd as dq

MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models", "cost_mlp.f32")


def subtree_samples(records, remaining):
    """Features and log1p(subtree branches) of every logged task whose subtree finished.

    records: TaskLog records (pieces of a split task share its uid and add their
    branches); remaining: (uid, parent) of the tasks still on the stack."""
    own, parent, first = {}, {}, {}
    for r in records:
        own[r.uid] = own.get(r.uid, 0) + r.branches
        parent[r.uid] = r.parent
        if r.uid not in first or not (r.flags & 1):
            first[r.uid] = r
    incomplete = set()
    for uid, par in remaining:
        u = par
        while u in parent and u not in incomplete:
            incomplete.add(u)
            u = parent[u]
    total = dict(own)
    for u in sorted(own, reverse=True):  # children have larger ids than their parent
        p = parent[u]
        if p in total:
            total[p] += total[u]
    keep = [u for u in own if u not in incomplete]
    x = dq.task_features([first[u] for u in keep])
    t = np.log1p(np.array([total[u] for u in keep], dtype=np.float64)).astype(np.float32)
    return x, t


def own_samples(records):
    """Features and log1p(own branches) of every logged task (pieces excluded)."""
    keep = [r for r in records if not (r.flags & 1)]
    x = dq.task_features(keep)
    t = np.log1p(np.array([r.branches for r in keep], dtype=np.float64)).astype(np.float32)
    return x, t


def collect(graph, gammas, steps, log_capacity=1 << 23, sched=dq.SCHED_FIFO, target="own"):
    """Runs `steps` expansion batches on one graph with the task log on; returns samples."""
    go, gi, mn = gammas
    cfg = dq.Config()
    cfg.host_addfail_check = 1
    with dq.Engine(cfg) as eng:
        eng.prepare(graph, go, gi, mn)
        eng.set_task_log(log_capacity)
        for _ in range(steps):
            if eng.step(1, sched=sched):
                break
        recs, dropped = eng.task_log()
        if dropped:
            raise RuntimeError(f"task log overflow ({dropped} records dropped)")
        remaining = eng.task_ids()
    return own_samples(recs) if target == "own" else subtree_samples(recs, remaining)


def train(x, t, epochs=30, batch=1024, lr=2e-3, seed=1):
    p0 = dq.mlp_init(seed)
    return dq.mlp_train(p0, x, t, epochs=epochs, batch=batch, lr=lr, seed=seed)


def load_model(path=MODEL_PATH):
    p = np.fromfile(path, dtype=np.float32)
    if p.size != dq.MLP_PARAMS:
        raise dq.DQCError(f"{path}: {p.size} parameters, expected {dq.MLP_PARAMS}")
    return p


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--graph", default="/tmp/dqc_workload2.txt")
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--out", default=MODEL_PATH)
    ap.add_argument("--target", choices=("own", "subtree"), default="own")
    args = ap.parse_args()
    if not os.path.exists(args.graph):
        dq.generate(2, args.graph)
    t0 = time.time()
    x, t = collect(args.graph, dq.WORKLOADS[2], args.steps, target=args.target)
    print(f"samples {len(x)} ({args.target} cost), target mean {t.mean():.3f} var {t.var():.3f}, "
          f"{time.time() - t0:.1f}s", flush=True)
    rng = np.random.default_rng(0)
    perm = rng.permutation(len(x))
    cut = int(0.9 * len(x))
    tr, va = perm[:cut], perm[cut:]
    p, loss = train(x[tr], t[tr], epochs=args.epochs)
    y = dq.mlp_forward(p, x[va])
    val = float(np.mean((y - t[va]) ** 2))
    print(f"train mse {loss:.4f} validation mse {val:.4f} (target variance {t[va].var():.4f})", flush=True)
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    p.astype(np.float32).tofile(args.out)
    print(f"wrote {args.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Autotuner for the enumerator's runtime knobs on one workload.

The knobs change only when and where work runs, never the output or the tree counts:
  pool_per_wg     tasks per launch, per resident workgroup (dqc_config.pool_tasks)
  split_branches  least branches of a task the SIZE / LEARNED schedules may split
  split_chunk     least branches per piece
  resident_steps  0 (host-driven steps) or S steps per replay of one captured HIP graph
  wave_tasks      the wavefront-vs-workgroup cutoff: batches whose tasks all have at most
                  this many vertices run one task per 64-lane wavefront (-1: never)
Every candidate runs the same complete unit of work: the whole enumeration, or the
enumeration cut at max_depth (a deterministic prefix of the tree). The fastest wins. The
search is coordinate descent: one knob at a time, the others held at their best so far,
so the cost is the sum of the knob choices rather than their product. Every candidate
must reproduce the first run's counts (tasks, branches, raw cliques, truncated
children); a mismatch raises.

Usage (on a GPU): python -m cudqc_amd.tune GRAPH OUT_GAMMA IN_GAMMA MIN_SIZE [--depth D]
"""
import argparse
import sys
import time

import cudqc_amd as dq

GRID = (("pool_per_wg", (8, 16, 32, 64)),
        ("split_branches", (256, 1024, 4096)),
        ("split_chunk", (16, 32, 64)),
        ("resident_steps", (0, 16)),
        ("wave_tasks", (-1, 64, 256, 2048)))


def apply(base, knobs, workgroups):
    """A copy of `base` with the chosen knobs set (None keeps base's value)."""
    cfg = dq.Config.from_buffer_copy(base)
    if knobs.get("pool_per_wg") is not None:
        cfg.pool_tasks = int(knobs["pool_per_wg"]) * int(workgroups)
    for name in ("split_branches", "split_chunk", "resident_steps", "wave_tasks"):
        if knobs.get(name) is not None:
            setattr(cfg, name, int(knobs[name]))
    return cfg


def measure(graph, gammas, cfg, sched, model=None):
    """(wall seconds of the expansion, counts, workgroups) of one complete run. `graph` is
    a path, or a list of paths enumerated together as one batch."""
    with dq.Engine(cfg) as eng:
        if isinstance(graph, (list, tuple)):
            eng.prepare_batch(list(graph), *gammas)
        else:
            eng.prepare(graph, *gammas)
        if sched == dq.SCHED_LEARNED:
            eng.set_cost_model(model)
        t = time.perf_counter()
        eng.run(sched)
        wall = time.perf_counter() - t
        st = eng.stats()
    return wall, (st.tasks_expanded, st.branches, st.raw_cliques, st.truncated_children), st.workgroups


def tune(graph, out_gamma, in_gamma, min_size, max_depth=0, sched=dq.SCHED_SIZE, base=None, grid=GRID,
         model=None, log=None):
    """Returns (best Config, best knobs, table of (knobs, wall seconds)) for one graph or a
    batch of graphs (a list of paths). `base` supplies every field the grid does not touch
    (default: the library defaults, 4-rank add-failure checks as in dqc_open(NULL))."""
    if base is None:
        base = dq.Config()
        base.host_addfail_check = 1
    base = dq.Config.from_buffer_copy(base)
    base.max_depth = max_depth
    if sched == dq.SCHED_LEARNED and model is None:
        from cudqc_amd import learn
        model = learn.load_model()
    gammas = (out_gamma, in_gamma, min_size)
    best = {name: None for name, _ in grid}
    best_wall, counts, workgroups = measure(graph, gammas, base, sched, model)
    table = [(dict(best), best_wall)]
    if log:
        log(f"defaults: {best_wall:.3f}s counts {counts}")
    for name, values in grid:
        current = dict(best)
        for v in values:
            trial = dict(current)
            trial[name] = v
            wall, c, _ = measure(graph, gammas, apply(base, trial, workgroups), sched, model)
            if c != counts:
                raise dq.DQCError(f"knobs {trial} changed the counts: {c} != {counts}")
            table.append((trial, wall))
            if log:
                log(f"{trial}: {wall:.3f}s")
            if wall < best_wall:
                best_wall, best = wall, trial
    return apply(base, best, workgroups), best, table


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("graph")
    ap.add_argument("out_gamma", type=float)
    ap.add_argument("in_gamma", type=float)
    ap.add_argument("min_size", type=int)
    ap.add_argument("--depth", type=int, default=0, help="cut the enumeration at this depth (0 = complete)")
    ap.add_argument("--sched", type=int, default=dq.SCHED_SIZE)
    args = ap.parse_args()
    _, best, table = tune(args.graph, args.out_gamma, args.in_gamma, args.min_size, args.depth, args.sched,
                          log=lambda s: print(s, flush=True))
    wall = min(w for _, w in table)
    print(f"best {best}: {wall:.3f}s (defaults {table[0][1]:.3f}s)", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

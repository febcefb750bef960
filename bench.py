"""Benchmark: maximal directed quasi-clique enumeration on one MI355X.

Headline workload: the R-MAT 2^20 graph (scale 20, a,b,c,d = .57/.19/.19/.05, 2^24 unique
arcs plus 500 planted quasi-cliques of 20-40 vertices at density 0.95; gamma_out =
gamma_in = 0.9, min size 18; seeded generator). Level 0 condenses it to 132,279 vertices.

Its full enumeration is far beyond one run, so the unit of work is a fixed, complete slice
of it: the level-0 loop steps [R0, R0 + B), the depth-1 tasks they produce and the depth-2
tasks those produce, each expanded in full (children deeper than that are counted, not
expanded). The level-0 steps before R0 run only as far as needed to find the first child
(which takes the host level-1 role, so no child of the window does). The unit's tasks,
branches and raw cliques are the same under every schedule; the benchmark checks that.

A step is one complete expansion of that unit by the batched + learned form (the LEARNED
schedule: batches ordered by the cost model, long tasks split into pieces sized to fill
the chip), from the level-0 task already resident in HBM. W untimed steps, then K timed
steps; value = tasks expanded per second over the timed steps. The same unit under plain
FIFO (stack order, no splitting: the plain GPU form) and under SIZE is timed for the
schedule comparison. Also reported: maximal QCs/s on the unit and on the workloads that
complete (the 1,000-vertex ER graph and the 256-graph batch); the expansion kernel's
roofline over the timed steps; the CPU oracle on a bounded sample of the unit's tasks.

Prints one JSON line. --gpus N > 1 exits: this path runs on one GPU per process.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
WORKLOAD = 4
# the bounded unit: level-0 steps [R0, R0 + B), subtrees to depth DEPTH (DQC_UNIT="R0,B,DEPTH"
# overrides it, for exploring other units; the reported workload names the one used)
R0, B, DEPTH = (int(x) for x in os.environ.get("DQC_UNIT", "90112,64,2").split(","))
WORKLOAD_NAME = ("R-MAT 2^20 directed, 2^24 arcs (.57/.19/.19/.05) + 500 planted QCs of 20-40 @0.95; "
                 "gamma_out=gamma_in=0.9, min_size=18; unit: level-0 steps [%d, %d) and their subtrees to depth %d"
                 % (R0, R0 + B, DEPTH))
# HBM-side traffic of expand_kernel per step of this unit (bytes), from two rocprofv3
# --pmc passes over this command (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors a
# 256 MiB copy reads on the same counters); profiles/expand_kernel_rmat20_unit.md. It is a
# measurement of this build, not of the run that prints it.
PMC_TRAFFIC_PER_STEP = 425.0e9
PMC_TRAFFIC_SOURCE = ("profiles/expand_kernel_rmat20_unit.md: rocprofv3 --pmc FETCH_SIZE (x2.0) and "
                      "WRITE_SIZE (x1.0) passes over this unit; 355.7 GB fetched + 69.3 GB written per step")


def unit_config(dq, sched):
    cfg = dq.Config()
    cfg.host_addfail_check = 1
    cfg.schedule = sched
    cfg.max_depth = DEPTH
    cfg.root_first = R0
    cfg.root_branches = B
    return cfg


def run_unit(eng, root, ncov, sched):
    """One complete expansion of the unit from the resident level-0 task; returns
    (wall seconds, stats, kernel ms, launches) of that run."""
    eng.reset()
    eng.push_root(root, ncov)
    s0 = eng.stats()
    k0, l0 = s0.expand_kernel_ms, s0.launches
    t = time.perf_counter()
    eng.run(sched)
    wall = time.perf_counter() - t
    st = eng.stats()
    return wall, st, st.expand_kernel_ms - k0, st.launches - l0


def maximal_unit(dq, eng, tmp):
    """The unit's raw cliques through the library's maximal-set filter (RemoveNonMax)."""
    cl = eng.cliques()
    raw = os.path.join(tmp, "unit.raw")
    mx = os.path.join(tmp, "unit.max")
    with open(raw, "w") as f:
        for c in cl:
            f.write(f"{len(c)} " + "".join(f"{x} " for x in c) + "\n")
    t = time.perf_counter()
    n = dq.remove_non_max(raw, mx) if cl else 0
    return n, time.perf_counter() - t


def complete_workloads(dq, data_dir):
    """Maximal QCs/s where the enumeration completes: the 1,000-vertex ER graph (+5 planted
    sets; gamma 0.8, min 10) through dqc_search_file, and the 256-graph batch (gamma 0.85,
    min 10) in one handle under the learned schedule with device-resident steps; time =
    enumeration + maximal-set filter, preprocessing reported apart (median of 3)."""
    from cudqc_amd import learn
    out = {}
    g1 = os.path.join(data_dir, "dqc_workload1.txt")
    if not os.path.exists(g1):
        dq.generate(1, g1)
    runs = []
    for _ in range(3):
        cnt, st = dq.search_file(g1, 0.8, 0.8, 10, None, os.path.join(data_dir, "dqc_w1.max"))
        runs.append((st.expand_wall_ms + st.rmnonmax_ms) / 1e3)
    t = statistics.median(runs)
    out["er1000"] = {"maximal": cnt, "tasks": st.tasks_expanded, "seconds": t, "maximal_qcs_per_s": cnt / t,
                     "tasks_per_s": st.tasks_expanded / t, "preprocess_s": st.preprocess_ms / 1e3}
    d = os.path.join(data_dir, "dqc_batch256")
    os.makedirs(d, exist_ok=True)
    paths = []
    for g in range(256):
        p = os.path.join(d, f"g{g}.txt")
        if not os.path.exists(p):
            dq.generate(3, p, g)
        paths.append(p)
    runs = []
    for name, sched, resident in (("fifo", dq.SCHED_FIFO, 0), ("learned_resident", dq.SCHED_LEARNED, 16)):
        times = []
        for _ in range(3):
            cfg = dq.Config()
            cfg.host_addfail_check = 1
            cfg.resident_steps = resident
            with dq.Engine(cfg) as eng:
                eng.prepare_batch(paths, 0.85, 0.85, 10)
                if sched == dq.SCHED_LEARNED:
                    eng.set_cost_model(learn.load_model())
                t = time.perf_counter()
                eng.run(sched)
                wall = time.perf_counter() - t
                counts, st = eng.finish_batch([None] * len(paths))
            times.append(wall + st.rmnonmax_ms / 1e3)
            runs.append((name, st.tasks_expanded, st.branches, st.raw_cliques, st.maximal_cliques))
        t = statistics.median(times)
        out["batch256_" + name] = {"maximal": st.maximal_cliques, "tasks": st.tasks_expanded, "seconds": t,
                                   "maximal_qcs_per_s": st.maximal_cliques / t, "tasks_per_s": st.tasks_expanded / t}
    if len({r[1:] for r in runs}) != 1:
        raise RuntimeError(f"batch runs disagree: {runs}")
    out["batch256_learned_over_fifo"] = out["batch256_fifo"]["seconds"] / out["batch256_learned_resident"]["seconds"]
    return out


def cpu_baseline(graph_path, sample_tasks, threads):
    """The oracle (C++ restatement of the reference's CPU path) expanding a sample of the
    unit's tasks on the host cores, one task per thread; rate in tasks/s and branches/s.
    Its own preprocessing (load, two-hop lists, level 0) is excluded."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as orc
    t0 = time.perf_counter()
    ot = orc.Tasks(graph_path, 0.9, 0.9, 18, threads=threads)
    prep = time.perf_counter() - t0
    try:
        t = time.perf_counter()
        with ThreadPoolExecutor(max_workers=threads) as pool:
            res = list(pool.map(lambda x: ot.expand_digest(x[0], x[1], 0, 2), sample_tasks))
        wall = time.perf_counter() - t
    finally:
        ot.close()
    branches = sum(r[2] for r in res)
    cpu = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            cpu = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), cpu)
    except OSError:
        pass
    return {
        "value": len(sample_tasks) / wall,
        "unit": "tasks expanded/s",
        "branches_per_s": branches / wall,
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle on {threads} threads of {cpu}: {len(sample_tasks)} depth-1 tasks of the unit "
                   f"({branches} branches) in {wall:.1f}s, one task per thread; its preprocessing "
                   f"({prep:.0f}s) excluded"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-tasks", type=int, default=64, help="unit tasks the CPU baseline expands (64: all of them)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-schedules", action="store_true", help="skip the FIFO / SIZE runs of the unit")
    ap.add_argument("--no-complete", action="store_true", help="skip the completed workloads (ER 1000, batch)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 256 MiB device-to-device copy first (counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): sharding the enumeration over "
              "several GPUs is not built.", file=sys.stderr)
        sys.exit(2)

    import cudqc_amd as dq
    from cudqc_amd import learn
    data_dir = os.environ.get("DQC_DATA", "/tmp")
    graph = os.path.join(data_dir, "dqc_workload4.txt")
    t = time.time()
    if not os.path.exists(graph):
        dq.generate(WORKLOAD, graph)
    gen_s = time.time() - t
    go, gi, mn = dq.WORKLOADS[WORKLOAD]
    print(f"bench: R-MAT 2^20 graph ready ({gen_s:.0f}s); preparing", file=sys.stderr, flush=True)

    if args.calibrate_copy:
        import torch
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
        dst = src.clone()
        torch.cuda.synchronize()
        del src, dst

    learned = dq.SCHED_LEARNED
    eng = dq.Engine(unit_config(dq, learned))
    t = time.time()
    k = eng.prepare(graph, go, gi, mn)
    prep_s = time.time() - t
    root, _, _, ncov = eng.task_meta()[0]
    eng.set_cost_model(learn.load_model())
    print(f"bench: {k} condensed vertices ({prep_s:.0f}s); {args.warmup} warm-up steps", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        run_unit(eng, root, ncov, learned)
    ref = None
    wall = kernel_ms = 0.0
    tasks = branches = nbytes = launches = 0
    for i in range(args.steps):
        w, st, kms, nl = run_unit(eng, root, ncov, learned)
        counts = (st.tasks_expanded, st.branches, st.raw_cliques, st.truncated_children)
        if ref is None:
            ref = counts
        elif counts != ref:
            raise RuntimeError(f"unit counts changed between steps: {ref} vs {counts}")
        wall += w
        kernel_ms += kms
        tasks += st.tasks_expanded
        branches += st.branches
        nbytes += st.bytes_scanned
        launches += nl
        print(f"bench: step {i + 1}/{args.steps}: {w:.2f}s", file=sys.stderr, flush=True)
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        n_max, rm_s = maximal_unit(dq, eng, tmp)
    unit_s = wall / max(args.steps, 1)
    # a sample of the unit's depth-1 tasks for the CPU baseline (the level-0 window alone)
    sample = []
    if not args.no_cpu_baseline:
        cfg0 = unit_config(dq, dq.SCHED_FIFO)
        cfg0.max_depth = 0
        with dq.Engine(cfg0) as e0:
            e0.prepare(graph, go, gi, mn)
            e0.step(1)
            kids = e0.tasks()
        sample = [(r, nm) for r, nm, _ in sorted(kids, key=lambda x: len(x[0]))][:args.cpu_tasks]
    sched = {}
    if not args.no_schedules:
        for name, sc in (("fifo", dq.SCHED_FIFO), ("size", dq.SCHED_SIZE)):
            w, st, kms, _ = run_unit(eng, root, ncov, sc)
            counts = (st.tasks_expanded, st.branches, st.raw_cliques, st.truncated_children)
            if counts != ref:
                raise RuntimeError(f"{name} disagrees with the learned form on the unit: {counts} vs {ref}")
            sched[name] = w
            print(f"bench: {name} unit: {w:.2f}s", file=sys.stderr, flush=True)
    eng.close()

    achieved = nbytes / (kernel_ms / 1e3) / 1e9 if kernel_ms > 0 else 0.0
    out = {
        "metric": "tasks expanded/s (maximal QC enumeration, 1 MI355X, batched + learned form, bit-exact output)",
        "value": tasks / wall if wall > 0 else 0.0,
        "unit": "tasks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": unit_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded generator)",
        "config": {
            "workload": WORKLOAD_NAME,
            "step": "one complete expansion of the unit, from the level-0 task resident in HBM",
            "schedule": "LEARNED (cost-model ordered batches, long tasks split into pieces)",
            "condensed_vertices": k,
            "workgroups": st.workgroups,
            "unit_tasks": ref[0],
            "unit_branches": ref[1],
            "unit_raw_cliques": ref[2],
            "unit_maximal_cliques": n_max,
            "unit_truncated_children": ref[3],
            "branches_per_s": branches / wall if wall > 0 else 0.0,
            "maximal_qcs_per_s": n_max / (unit_s + rm_s) if unit_s > 0 else 0.0,
            "launches_per_step": launches / max(args.steps, 1),
            "kernel_ms_per_step": kernel_ms / max(args.steps, 1),
            "preprocess_s": prep_s,
            "generate_s": gen_s,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": PMC_TRAFFIC_PER_STEP,
            "traffic_source": PMC_TRAFFIC_SOURCE,
            "kernel": ("expand_kernel, HBM-bitmap variant: per step one level-0 head launch, the level-0 window, "
                       "the split passes (1024-thread workgroups) and the pieces; achieved = algorithmic bytes per "
                       "step / their HIP-event time per step"),
            "algorithmic_bytes_per_step": nbytes / max(args.steps, 1),
        },
    }
    if sched:
        out["schedules"] = {
            "workload": "the same unit, complete, identical counts under every schedule",
            "learned_s": unit_s,
            "fifo_s": sched["fifo"],
            "size_s": sched["size"],
            "fifo_tasks_per_s": ref[0] / sched["fifo"],
            "size_tasks_per_s": ref[0] / sched["size"],
            "learned_tasks_per_s": ref[0] / unit_s,
            "learned_over_fifo": sched["fifo"] / unit_s,
            "size_over_fifo": sched["fifo"] / sched["size"],
            "learned_over_size": sched["size"] / unit_s,
        }
    if not args.no_complete:
        out["complete"] = complete_workloads(dq, data_dir)
    if not args.no_cpu_baseline:
        threads = min(os.cpu_count() or 1, 16)
        out["cpu_baseline"] = cpu_baseline(graph, sample, threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

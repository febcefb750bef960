"""GPU parity at full size on the R-MAT 2^20 workload (2^24 arcs + 500 planted sets of
20-40 vertices at 0.95; gamma 0.9/0.9, min size 18). Level 0 condenses it to 132,279
vertices, so the HBM-bitmap kernel variant runs.

Its enumeration is far beyond the oracle, so parity is checked task by task, as for the
R-MAT 2^16 workload: a window of level-0 branches (including the first child, which
keeps the host level-1 semantics), and a sample of the depth-1 tasks those branches
produce, each expanded alone by the GPU (cut into pieces, so the whole chip runs it) and
by the oracle. Every child record (ids, four degrees, lvl2adj, order), member count,
clique and the branch count must be equal (children of depth-1 tasks are compared by a
SHA-1 of their record bytes, tests/test_oracle.py pins that digest form).
"""
import os
from concurrent.futures import ThreadPoolExecutor

import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GO, GI, MN = 0.9, 0.9, 18
STEPS = 40  # the first level-0 loop steps (they hold the first child, the host level-1 task)
# a window of level-0 steps where most branches yield a depth-1 task of 10-20k vertices
WINDOW = (90112, 12)


@pytest.fixture(scope="module")
def w4(tmp_path_factory):
    import cudqc_amd as dq
    path = str(tmp_path_factory.mktemp("w4") / "workload4.txt")
    print("generating the R-MAT 2^20 workload", flush=True)
    assert dq.generate(4, path) == 1 << 20
    return path


@pytest.fixture(scope="module")
def oracle_w4(w4):
    print("oracle: load, two-hop lists, level 0", flush=True)
    t = orc.Tasks(w4, GO, GI, MN, threads=min(os.cpu_count() or 1, 16))
    root, ncov = t.root()
    print(f"oracle: {len(root)} condensed vertices, {ncov} covered; expanding {STEPS} level-0 steps", flush=True)
    kids, cliques, branches = t.expand(root, 0, ncov, orc.Tasks.HOST_L0, max_steps=STEPS)
    print(f"oracle: level-0 window {WINDOW}", flush=True)
    wkids, wcl, wbr = t.expand(root, 0, ncov, orc.Tasks.HOST_L0, max_steps=WINDOW[1], first_step=WINDOW[0])
    yield t, root, ncov, kids, cliques, branches, (wkids, wcl, wbr)
    t.close()


@pytest.fixture(scope="module")
def eng4(w4):
    """One handle prepared once (45 s of host preprocessing) for every test of this file;
    each test re-aims its bounded unit (dqc_set_unit) and starts from the level-0 task or
    from the tasks it pushes. SIZE schedule, tasks of 64+ steps split into pieces of 8+."""
    import cudqc_amd as dq
    cfg = dq.Config()
    cfg.host_addfail_check = 1
    cfg.schedule = dq.SCHED_SIZE
    cfg.split_branches, cfg.split_chunk = 64, 8
    eng = dq.Engine(cfg)
    print("engine: preparing the R-MAT 2^20 graph", flush=True)
    k = eng.prepare(w4, GO, GI, MN)
    root, _, _, ncov = eng.task_meta()[0]
    yield eng, k, root, ncov
    eng.close()


def run_window(eng, root, ncov, first, count):
    eng.set_unit(first, count, 0)
    eng.reset()
    eng.push_root(root, ncov)
    eng.step(1)
    return eng.tasks(), eng.cliques(), eng.stats()


def test_level0_window_matches_oracle(dqc, eng4, oracle_w4):
    """The first level-0 steps, branch-parallel on the GPU, equal the oracle's sequential
    loop over the same steps (the bounded unit of work root_branches defines)."""
    _, root, ncov, okids, ocl, obr, _ = oracle_w4
    eng, k, groot, gncov = eng4
    assert k == len(root) == 132279
    assert gncov == ncov
    tasks, cl, st = run_window(eng, groot, gncov, 0, STEPS)
    print(f"level 0: {len(okids)} children, {obr} branches, {len(ocl)} cliques", flush=True)
    assert okids and st.branches == obr
    assert sorted((tuple(r), nm) for r, nm, _ in tasks) == sorted((tuple(r), nm) for r, nm in okids)
    first = [t for t in tasks if t[2] == 1]
    assert len(first) == 1 and first[0][0] == okids[0][0]
    assert sorted(tuple(sorted(c)) for c in cl) == ocl


def test_level0_window_in_dense_part_matches_oracle(dqc, eng4, oracle_w4):
    """Level-0 steps [90112, 90124) (the bounded unit of work root_first / root_branches
    define; the steps before run only until the first child): the same children, cliques
    and branch count as the oracle's loop over that window, all children device tasks."""
    _, root, ncov, _, _, _, (wkids, wcl, wbr) = oracle_w4
    eng, _, groot, gncov = eng4
    tasks, cl, st = run_window(eng, groot, gncov, *WINDOW)
    print(f"window: {len(wkids)} children, {wbr} branches, {len(wcl)} cliques", flush=True)
    assert len(wkids) >= 8 and st.branches == wbr
    assert sorted((tuple(r), nm) for r, nm, _ in tasks) == sorted((tuple(r), nm) for r, nm in wkids)
    assert all(m == 2 for _, _, m in tasks)
    assert sorted(tuple(sorted(c)) for c in cl) == wcl


def test_depth1_tasks_match_oracle(dqc, eng4, oracle_w4):
    """The depth-1 tasks of the dense window, each split into pieces on the GPU (so the
    whole chip runs it: the split pass on a 1024-thread workgroup, then its pieces) and
    compared with the oracle's unsplit expansion (the oracle runs them side by side on
    the host cores meanwhile)."""
    ot, _, _, _, _, _, (wkids, _, _) = oracle_w4
    sample = list(range(len(wkids)))[:10]
    assert len(sample) >= 8, len(wkids)
    print("sample sizes", [len(wkids[i][0]) for i in sample], flush=True)
    with ThreadPoolExecutor(max_workers=len(sample)) as pool:
        futs = {i: pool.submit(ot.expand_digest, wkids[i][0], wkids[i][1], 0, orc.Tasks.DEVICE) for i in sample}
        eng, _, _, _ = eng4
        eng.set_unit(0, 0, 0)
        eng.set_task_log(1 << 16)
        got = {}
        for i in sample:
            recs, nm = wkids[i]
            eng.reset()
            eng.push_task(recs, nm, 0, orc.Tasks.DEVICE)
            eng.run()
            st = eng.stats()
            log, _ = eng.task_log()
            assert sum(1 for r in log if r.flags & 1) >= 2, i  # it ran as pieces
            assert st.stack_pushbacks == 0 and st.time_slices == 0, i
            got[i] = (sorted((h, m) for h, m, _ in eng.task_digests()),
                      sorted(tuple(sorted(c)) for c in eng.cliques()), st.branches, st.tasks_expanded)
            print(f"task {i}: size {len(recs)}, {st.branches} branches, {len(got[i][0])} children", flush=True)
        for i in sample:
            kids, cl, br = futs[i].result()
            assert got[i][3] == 1, i
            assert got[i][2] == br, i
            assert got[i][0] == sorted(kids), i
            assert got[i][1] == cl, i

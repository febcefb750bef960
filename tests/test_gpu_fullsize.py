"""GPU parity at the benchmark workload's full size (R-MAT 2^16, 2^20 arcs + 50 planted
sets; gamma 0.9, min size 12).

Its full enumeration runs to tens of millions of tasks, beyond what the oracle finishes,
so parity is checked task by task: one level of the enumeration tree expanded by the GPU
must equal the oracle's expansion of the same task: every child record (ids, four
degrees, lvl2adj, order), every member count, every clique, and the branch count.
"""
import os

import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GO, GI, MN = 0.9, 0.9, 12


@pytest.fixture(scope="module")
def w2(tmp_path_factory):
    import cudqc_amd as dq
    path = str(tmp_path_factory.mktemp("w2") / "workload2.txt")
    dq.generate(2, path)
    return path


@pytest.fixture(scope="module")
def oracle_tasks(w2):
    t = orc.Tasks(w2, GO, GI, MN)
    root, ncov = t.root()
    kids, cliques, branches = t.expand(root, 0, ncov, orc.Tasks.HOST_L0)
    yield t, root, ncov, kids, cliques, branches
    t.close()


def config(dqc, split_branches=-1, split_chunk=0):
    cfg = dqc.Config()
    cfg.host_addfail_check = 1
    cfg.split_branches = split_branches
    cfg.split_chunk = split_chunk
    return cfg


def test_level0_children_match_oracle(dqc, w2, oracle_tasks):
    """Branch-parallel level 0 == the reference's sequential level-0 loop, at full size."""
    _, root, ncov, okids, ocl, obr = oracle_tasks
    with dqc.Engine() as eng:
        k = eng.prepare(w2, GO, GI, MN)
        assert k == len(root)
        eng.step(1)
        tasks = eng.tasks()
        cl = eng.cliques()
        st = eng.stats()
    print(f"level 0: {len(okids)} children, {obr} branches, {len(ocl)} cliques", flush=True)
    assert st.branches == obr
    assert sorted((tuple(r), nm) for r, nm, _ in tasks) == sorted((tuple(r), nm) for r, nm in okids)
    first = [t for t in tasks if t[2] == 1]
    assert len(first) == 1 and first[0][0] == okids[0][0]
    assert all(t[2] == 2 for t in tasks if t is not first[0])
    assert sorted(tuple(sorted(c)) for c in cl) == ocl


def test_level1_tasks_match_oracle(dqc, w2, oracle_tasks):
    """A deterministic sample of level-1 tasks (the host level-1 one, the largest ones and
    evenly spaced ones), each expanded alone on the GPU and by the oracle."""
    ot, _, _, okids, _, _ = oracle_tasks
    order = sorted(range(1, len(okids)), key=lambda i: -len(okids[i][0]))
    sample = [0] + order[:3] + list(range(1, len(okids), max(1, len(okids) // 10)))[:10]
    with dqc.Engine(config(dqc)) as eng:
        eng.prepare(w2, GO, GI, MN)
        for i in sample:
            recs, nm = okids[i]
            mode = orc.Tasks.HOST_L1 if i == 0 else orc.Tasks.DEVICE
            eng.reset()
            eng.push_task(recs, nm, 0, mode)
            eng.step(1)
            gk = sorted((tuple(r), m) for r, m, _ in eng.tasks())
            gc = sorted(tuple(sorted(c)) for c in eng.cliques())
            gbr = eng.stats().branches
            kids, cl, br = ot.expand(recs, nm, 0, mode)
            print(f"task {i}: size {len(recs)}, {len(kids)} children, {br} branches, {len(cl)} cliques", flush=True)
            assert gbr == br, i
            assert gk == sorted((tuple(r), m) for r, m in kids), i
            assert gc == cl, i


def test_split_tasks_match_oracle(dqc, w2, oracle_tasks):
    """The host level-1 task and the largest level-1 tasks cut into pieces of 16 branches:
    one step runs the split pass (pieces only, no branches), the next runs every piece;
    the children, cliques and branch count must then equal the oracle's unsplit expansion."""
    ot, _, _, okids, _, _ = oracle_tasks
    order = sorted(range(1, len(okids)), key=lambda i: -len(okids[i][0]))
    with dqc.Engine(config(dqc, split_branches=64, split_chunk=16)) as eng:
        eng.prepare(w2, GO, GI, MN)
        for i in [0] + order[:3]:
            recs, nm = okids[i]
            mode = orc.Tasks.HOST_L1 if i == 0 else orc.Tasks.DEVICE
            eng.reset()
            eng.push_task(recs, nm, 0, mode)
            eng.step(1)
            pieces = eng.tasks()
            assert eng.stats().branches == 0 and eng.stats().tasks_expanded == 1
            # one piece per 16 steps, fewer if the loop stops early (a failed removal or lookahead)
            assert 1 <= len(pieces) <= -(-(len(recs) - nm) // 16) and all(m == mode for _, _, m in pieces), i
            eng.step(1)
            gk = sorted((tuple(r), m) for r, m, _ in eng.tasks())
            gc = sorted(tuple(sorted(c)) for c in eng.cliques())
            st = eng.stats()
            kids, cl, br = ot.expand(recs, nm, 0, mode)
            print(f"task {i}: size {len(recs)}, {len(pieces)} pieces, {len(kids)} children, {br} branches", flush=True)
            assert st.tasks_expanded == 1, i
            assert st.branches == br, i
            assert gk == sorted((tuple(r), m) for r, m in kids), i
            assert gc == cl, i

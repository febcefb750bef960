"""Host side of the learned schedule, the autotuner and the benchmark driver (CPU only).

The training-sample builders of cudqc_amd/learn.py are checked on a hand-built task tree
whose subtree costs are worked out by hand below; the benchmark refuses more than one GPU.
"""
import os
import subprocess
import sys

import numpy as np

from tests.conftest import ROOT

NO_PARENT = 0xFFFFFFFF


def rec(dqc, uid, parent, branches, piece=False, tot=10, nm=2, depth=1):
    return dqc.TaskLog(uid=uid, parent=parent, depth=depth, nm=nm, tot=tot, dmin_out=1, dmax_out=5, dmin_in=1,
                       dmax_in=6, branches=branches, flags=1 if piece else 0)


def tree(dqc):
    """uid 1 (root) has children 2 and 3; 2 has child 4; 3 has child 5. Task 3 was split:
    its own record ran 7 branches and one piece ran 4 more. Task 6, a child of 4, is still
    on the stack, so the subtrees of 4, 2 and 1 are incomplete."""
    records = [rec(dqc, 1, NO_PARENT, 10, tot=40, depth=0), rec(dqc, 2, 1, 5, tot=20), rec(dqc, 3, 1, 7, tot=30),
               rec(dqc, 3, 1, 4, piece=True, tot=25), rec(dqc, 4, 2, 2, tot=8, depth=2),
               rec(dqc, 5, 3, 1, tot=6, depth=2)]
    remaining = [(6, 4)]
    return records, remaining


def test_subtree_samples_hand_built_tree(dqc):
    from cudqc_amd import learn
    records, remaining = tree(dqc)
    x, t = learn.subtree_samples(records, remaining)
    # complete subtrees: 5 alone (1 branch) and 3 with 5 (7 + 4 + 1 = 12 branches)
    assert sorted(np.round(np.expm1(t)).astype(int).tolist()) == [1, 12]
    # features come from the task's own record, not from its piece
    by_cost = dict(zip(np.round(np.expm1(t)).astype(int).tolist(), x))
    own3 = dqc.task_features([records[2]])[0]
    assert np.allclose(by_cost[12], own3)
    assert not np.allclose(by_cost[12], dqc.task_features([records[3]])[0])


def test_subtree_samples_all_complete(dqc):
    from cudqc_amd import learn
    records, _ = tree(dqc)
    _, t = learn.subtree_samples(records, [])
    # 1: 10 + (5 + 2) + (7 + 4 + 1) = 29; 2: 7; 3: 12; 4: 2; 5: 1
    assert sorted(np.round(np.expm1(t)).astype(int).tolist()) == [1, 2, 7, 12, 29]


def test_own_samples_skip_pieces(dqc):
    from cudqc_amd import learn
    records, _ = tree(dqc)
    x, t = learn.own_samples(records)
    assert len(x) == 5
    assert sorted(np.round(np.expm1(t)).astype(int).tolist()) == [1, 2, 5, 7, 10]


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_tuner_applies_knobs_to_a_copy(dqc):
    from cudqc_amd import tune
    base = dqc.Config()
    base.host_addfail_check, base.max_depth, base.spill_tasks = 1, 3, 12345
    cfg = tune.apply(base, {"pool_per_wg": 16, "split_branches": 256, "split_chunk": None, "resident_steps": 4}, 2048)
    assert (cfg.pool_tasks, cfg.split_branches, cfg.split_chunk, cfg.resident_steps) == (16 * 2048, 256, 0, 4)
    assert (cfg.host_addfail_check, cfg.max_depth, cfg.spill_tasks) == (1, 3, 12345)
    assert (base.pool_tasks, base.split_branches, base.resident_steps) == (0, 0, 0)

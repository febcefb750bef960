"""Batched form: many graphs expanded together in one handle (shared launches, one id
space). Each graph's maximal file must equal its single-graph golden output, and the
tree counts of the batch must be the sums of the single-graph counts."""
import os

import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu


def batch_cases(manifest):
    return [c for c in manifest["cases"] if c["name"].startswith("workload3_")]


def check(dqc, cases, tmp_path, config=None):
    paths = [os.path.join(GOLDEN, c["graph"]) for c in cases]
    outs = [str(tmp_path / (c["name"] + ".max")) for c in cases]
    counts, st = dqc.search_batch(paths, 0.85, 0.85, 10, outs, config=config)
    for c, n, o in zip(cases, counts, outs):
        got = open(o, "rb").read() if os.path.exists(o) else b""
        exp = open(os.path.join(GOLDEN, c["expected"]), "rb").read() if c["expected"] else b""
        assert (n, got) == (c["maximal"], exp), c["name"]
    assert st.tasks_expanded == sum(c["tasks"] for c in cases)
    assert st.branches == sum(c["branches"] for c in cases)
    assert st.raw_cliques == sum(c["raw_cliques"] for c in cases)


def test_batch_equals_single_graph_runs(dqc, manifest, tmp_path):
    cases = batch_cases(manifest)
    assert len(cases) >= 8
    check(dqc, cases, tmp_path)


def test_batch_fifo_and_split(dqc, manifest, tmp_path):
    """The same under the plain FIFO order with tiny batches, and with every task split."""
    cases = batch_cases(manifest)
    cfg = dqc.Config()
    cfg.host_addfail_check = 1
    cfg.schedule = dqc.SCHED_FIFO
    cfg.pool_tasks = 64
    check(dqc, cases, tmp_path, config=cfg)
    cfg = dqc.Config()
    cfg.host_addfail_check = 1
    cfg.split_branches = 2
    cfg.split_chunk = 3
    check(dqc, list(reversed(cases)), tmp_path, config=cfg)


def test_batch_hbm_bitmaps(dqc, manifest, tmp_path, monkeypatch):
    monkeypatch.setenv("DQC_HBM_BITMAPS", "1")
    check(dqc, batch_cases(manifest), tmp_path)


def test_batch_resident(dqc, manifest, tmp_path):
    cfg = dqc.Config()
    cfg.host_addfail_check = 1
    cfg.resident_steps = 8
    check(dqc, batch_cases(manifest), tmp_path, config=cfg)

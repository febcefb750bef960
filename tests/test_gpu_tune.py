"""The runtime-knob autotuner (cudqc_amd/tune.py): every candidate it times must give the
same tree counts, and the configuration it returns must give the golden output."""
import os

import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

# small values force splitting and tiny batches, so the candidates really differ
GRID = (("pool_per_wg", (1, 32)),
        ("split_branches", (2, 1024)),
        ("split_chunk", (3, 32)),
        ("resident_steps", (0, 4)),
        ("wave_tasks", (-1, 2048)))


def test_tuner_keeps_counts_and_output(dqc, manifest, tmp_path):
    from cudqc_amd import tune
    case = next(c for c in manifest["cases"] if c["name"] == "workload3_g2")
    graph = os.path.join(GOLDEN, case["graph"])
    go, gi, mn = case["out_gamma"], case["in_gamma"], case["min_size"]
    cfg, best, table = tune.tune(graph, go, gi, mn, grid=GRID, log=lambda s: print(s, flush=True))
    assert len(table) == 1 + sum(len(v) for _, v in GRID)
    print("best", best, flush=True)
    with dqc.Engine(cfg) as eng:
        eng.prepare(graph, go, gi, mn)
        eng.run()
        st = eng.stats()
        cl = eng.cliques()
    assert (st.tasks_expanded, st.branches, st.raw_cliques) == (case["tasks"], case["branches"], case["raw_cliques"])
    raw, mx = str(tmp_path / "t.raw"), str(tmp_path / "t.max")
    with open(raw, "w") as f:
        for s in cl:
            f.write(f"{len(s)} " + "".join(f"{x} " for x in s) + "\n")
    assert dqc.remove_non_max(raw, mx) == case["maximal"]
    assert open(mx, "rb").read() == open(os.path.join(GOLDEN, case["expected"]), "rb").read()

"""GPU parity: the HIP path (through the C-ABI) against the oracle's golden outputs.

Bit-exact bar: the maximal file (DQC-R bytes) must equal the oracle's, and so must the
scheduling-invariant counts (tasks expanded, branches, raw cliques), which pins the
shape of the enumeration tree, not just the final set.
"""
import hashlib
import os
import subprocess

import pytest

from tests.conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu


def run_case(dqc, c, tmp_path, config=None):
    mx = str(tmp_path / (c["name"] + ".max"))
    cnt, st = dqc.search_file(os.path.join(GOLDEN, c["graph"]), c["out_gamma"], c["in_gamma"], c["min_size"],
                              None, mx, config=config)
    return cnt, st, mx


# wave_tasks: 0 = the library default (tasks of up to 256 vertices one per wavefront),
# -1 = every task on a 256-thread workgroup, 2048 = every batch that fits the wavefront
# kernel (graphs of at most 2048 condensed vertices) on it
@pytest.mark.parametrize("wave_tasks", [0, -1, 2048])
def test_golden_cases_bit_exact(dqc, manifest, tmp_path, wave_tasks):
    cfg = dqc.Config()
    cfg.host_addfail_check = 1
    cfg.wave_tasks = wave_tasks
    bad = []
    waves = 0
    for c in manifest["cases"]:
        print("case", c["name"], "wave_tasks", wave_tasks, flush=True)
        cnt, st, mx = run_case(dqc, c, tmp_path, config=cfg)
        waves += st.wave_launches
        got = open(mx, "rb").read() if os.path.exists(mx) else b""
        exp = open(os.path.join(GOLDEN, c["expected"]), "rb").read() if c["expected"] else b""
        if (got != exp or cnt != c["maximal"] or st.tasks_expanded != c["tasks"] or st.branches != c["branches"]
                or st.raw_cliques != c["raw_cliques"]):
            bad.append((c["name"], cnt, c["maximal"], st.tasks_expanded, c["tasks"]))
    assert not bad, bad
    assert (waves == 0) == (wave_tasks < 0), waves


def test_small_batches_same_result(dqc, manifest, tmp_path):
    """Scheduling invariance: a tiny batch size (many launches, deep stack) in plain FIFO
    order gives the same bytes and counts as the default."""
    c = next(x for x in manifest["cases"] if x["name"] == "workload3_g0")
    cfg = dqc.Config()
    cfg.pool_tasks = 7
    cfg.workgroups = 5
    cfg.host_addfail_check = 1
    cfg.schedule = dqc.SCHED_FIFO
    cnt, st, mx = run_case(dqc, c, tmp_path, config=cfg)
    assert cnt == c["maximal"] and st.tasks_expanded == c["tasks"] and st.branches == c["branches"]
    assert hashlib.sha256(open(mx, "rb").read()).hexdigest() == c["sha256"]
    assert st.launches > 100


def test_capacity_error_is_clean(dqc, manifest, tmp_path):
    """A task stack too small for the search fails with DQC_E_CAPACITY, not a fault."""
    c = next(x for x in manifest["cases"] if x["name"] == "workload3_g2")
    cfg = dqc.Config()
    cfg.spill_tasks = 4
    cfg.spill_vertices = 1 << 16
    cfg.host_addfail_check = 1
    with pytest.raises(dqc.DQCError, match="DQC_E_CAPACITY"):
        run_case(dqc, c, tmp_path, config=cfg)


def test_cli_writes_reference_files(dqc, manifest, tmp_path):
    c = next(x for x in manifest["cases"] if x["name"] == "workload1")
    ds = tmp_path / "DS_Sizes.csv"
    ds.write_text("TASKS_SIZE,100000000\n")
    r = subprocess.run([dqc.CLI_PATH, os.path.join(GOLDEN, c["graph"]), "0.8", "0.8", "10", str(ds), "w1"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "NUMBER OF CLIQUES:   11" in r.stdout
    assert (tmp_path / "DQC-R_w1").read_bytes() == open(os.path.join(GOLDEN, c["expected"]), "rb").read()
    raw = (tmp_path / "DQC-T_w1").read_text().splitlines()
    assert len(raw) == c["raw_cliques"] and all(l.endswith(" ") for l in raw)


def test_cli_argument_errors(dqc, tmp_path):
    r = subprocess.run([dqc.CLI_PATH, "x"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Usage" in r.stdout
    g = tmp_path / "g.txt"
    g.write_text("1\n0\n")
    r = subprocess.run([dqc.CLI_PATH, str(g), "0.3", "0.8", "3", "ds", "t"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "between .5 and 1" in r.stdout


def test_serial_level0_same_result(dqc, manifest, tmp_path, monkeypatch):
    """The level-0 task run by one workgroup (sequential branch loop) and branch-parallel
    (default) give the same bytes and the same tree counts."""
    monkeypatch.setenv("DQC_SERIAL_ROOT", "1")
    for name in ("workload1", "workload3_g0", "workload3_g21"):
        c = next(x for x in manifest["cases"] if x["name"] == name)
        cnt, st, mx = run_case(dqc, c, tmp_path)
        assert cnt == c["maximal"] and st.tasks_expanded == c["tasks"] and st.branches == c["branches"], name
        if c["sha256"]:
            assert hashlib.sha256(open(mx, "rb").read()).hexdigest() == c["sha256"], name


def test_split_everything_same_result(dqc, manifest, tmp_path):
    """Every task with 2 or more branches cut into pieces of 3: same bytes and counts."""
    cfg = dqc.Config()
    cfg.host_addfail_check = 1
    cfg.split_branches = 2
    cfg.split_chunk = 3
    bad = []
    for c in manifest["cases"]:
        cnt, st, mx = run_case(dqc, c, tmp_path, config=cfg)
        got = open(mx, "rb").read() if os.path.exists(mx) else b""
        exp = open(os.path.join(GOLDEN, c["expected"]), "rb").read() if c["expected"] else b""
        if (got != exp or cnt != c["maximal"] or st.tasks_expanded != c["tasks"] or st.branches != c["branches"]
                or st.raw_cliques != c["raw_cliques"]):
            bad.append((c["name"], cnt, c["maximal"], st.tasks_expanded, c["tasks"], st.branches, c["branches"]))
    assert not bad, bad


def test_split_everything_on_workgroups(dqc, manifest, tmp_path):
    """Every task with 2 or more branches cut into pieces of 3, on 256-thread workgroups
    (no wavefront kernel): the split passes of each batch run first, on 1024-thread
    workgroups, then the pieces. Same bytes and counts."""
    cfg = dqc.Config()
    cfg.host_addfail_check = 1
    cfg.split_branches = 2
    cfg.split_chunk = 3
    cfg.wave_tasks = -1
    bad = []
    for c in manifest["cases"]:
        cnt, st, mx = run_case(dqc, c, tmp_path, config=cfg)
        got = open(mx, "rb").read() if os.path.exists(mx) else b""
        exp = open(os.path.join(GOLDEN, c["expected"]), "rb").read() if c["expected"] else b""
        if (got != exp or cnt != c["maximal"] or st.tasks_expanded != c["tasks"] or st.branches != c["branches"]
                or st.raw_cliques != c["raw_cliques"]):
            bad.append((c["name"], cnt, c["maximal"], st.tasks_expanded, c["tasks"], st.branches, c["branches"]))
    assert not bad, bad


@pytest.mark.parametrize("summary_shift", [0, 4])
def test_hbm_bitmaps_same_result(dqc, manifest, tmp_path, monkeypatch, summary_shift):
    """The kernel variant for condensed graphs above 16384 vertices keeps its id bitmaps in
    HBM, behind a live-set summary in LDS (one bit per id up to 98,304 ids, per 2^s ids
    above). Forced on every golden case, with an exact summary and with one bit per 16 ids,
    it gives the same bytes and counts."""
    monkeypatch.setenv("DQC_HBM_BITMAPS", "1")
    monkeypatch.setenv("DQC_LIVE_SUMMARY_SHIFT", str(summary_shift))
    print("summary shift", summary_shift, flush=True)
    bad = []
    for c in manifest["cases"]:
        cnt, st, mx = run_case(dqc, c, tmp_path)
        got = open(mx, "rb").read() if os.path.exists(mx) else b""
        exp = open(os.path.join(GOLDEN, c["expected"]), "rb").read() if c["expected"] else b""
        if (got != exp or cnt != c["maximal"] or st.tasks_expanded != c["tasks"] or st.branches != c["branches"]
                or st.raw_cliques != c["raw_cliques"]):
            bad.append((c["name"], cnt, c["maximal"], st.tasks_expanded, c["tasks"]))
    assert not bad, bad


DEEP = ["workload3_g0", "workload3_g2", "workload3_g8", "rand017", "rand069", "rand088", "bigk20000"]


@pytest.mark.parametrize("sched", ["SCHED_FIFO", "SCHED_SIZE"])
def test_time_slices_same_result(dqc, manifest, tmp_path, sched):
    """Launches cut into 1 ms time slices: a task still running at the end of its slice
    pushes the rest of its steps back onto the stack as a piece. Same bytes and counts,
    and the slicing happened."""
    cfg = dqc.Config()
    cfg.host_addfail_check = 1
    cfg.schedule = getattr(dqc, sched)
    cfg.launch_ms = 1
    slices = 0
    for c in manifest["cases"]:
        if c["name"] not in DEEP:
            continue
        cnt, st, mx = run_case(dqc, c, tmp_path, config=cfg)
        got = open(mx, "rb").read() if os.path.exists(mx) else b""
        exp = open(os.path.join(GOLDEN, c["expected"]), "rb").read() if c["expected"] else b""
        assert got == exp and cnt == c["maximal"], c["name"]
        assert (st.tasks_expanded, st.branches, st.raw_cliques) == (c["tasks"], c["branches"], c["raw_cliques"]), c["name"]
        slices += st.time_slices
    print("time slices", slices, flush=True)
    assert slices > 0


def level0_children(dqc, path, c, first, count, serial, monkeypatch):
    if serial:
        monkeypatch.setenv("DQC_SERIAL_ROOT", "1")
    else:
        monkeypatch.delenv("DQC_SERIAL_ROOT", raising=False)
    cfg = dqc.Config()
    cfg.host_addfail_check = 1
    cfg.root_first, cfg.root_branches = first, count
    with dqc.Engine(cfg) as eng:
        eng.prepare(path, c["out_gamma"], c["in_gamma"], c["min_size"])
        eng.step(1)
        kids = eng.tasks()
        st = eng.stats()
    return kids, st


def test_level0_window(dqc, manifest, monkeypatch):
    """A window [a, a + b) of level-0 steps (root_first / root_branches): its children are
    those of the prefix [0, a + b) minus those of the prefix [0, a); the first child keeps
    host level-1 semantics only if it lies in the window (the steps before the window run
    only until they produce a child, which is not expanded). Branch-parallel and serial
    level 0 agree."""
    from collections import Counter
    for name, a, b in (("workload3_g2", 3, 25), ("workload3_g0", 1, 30), ("rand069", 0, 12), ("bigk20000", 40, 60)):
        c = next(x for x in manifest["cases"] if x["name"] == name)
        path = os.path.join(GOLDEN, c["graph"])
        full, _ = level0_children(dqc, path, c, 0, a + b, False, monkeypatch)
        head, _ = level0_children(dqc, path, c, 0, a, False, monkeypatch) if a else ([], None)
        expect = Counter((tuple(r), nm) for r, nm, _ in full) - Counter((tuple(r), nm) for r, nm, _ in head)
        for serial in (False, True):
            win, st = level0_children(dqc, path, c, a, b, serial, monkeypatch)
            got = Counter((tuple(r), nm) for r, nm, _ in win)
            assert got == expect, (name, serial)
            assert st.truncated_children == 0 and st.tasks_expanded == 1, (name, serial)
            modes = sorted(m for _, _, m in win)
            first_in_window = not head and bool(full)
            assert modes.count(1) == (1 if first_in_window and win else 0), (name, serial, modes)

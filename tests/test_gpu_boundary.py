"""GPU tests of the in-memory host API (dqc_search_graph / dqc_prepare_graph, the
counterparts of the reference's h_search over a CPU_Graph with caller-owned minimum-
degree tables, inc/host_functions.hpp:7-10) and of the binary graph cache on the search
path: each must give exactly the bytes dqc_search_file gives from the text file."""
import os

import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

NAMES = ["workload1", "workload3_g0", "workload3_g5", "rand017", "rand061", "rand069", "rand088"]


def max_text(sets):
    return "".join(f"{len(s)} " + "".join(f"{x} " for x in s) + "\n" for s in sets).encode()


def test_search_graph_equals_search_file(dqc, manifest, tmp_path):
    """The graph loaded by the caller (adjacency lists, with and without two-hop lists),
    the reference's n + 1 entry tables from dqc_min_degrees: maximal sets, their order
    and the tree counts equal the file path's golden output; the raw file holds the same
    sets as the file path's."""
    cases = [c for c in manifest["cases"] if c["name"] in NAMES or c["out_gamma"] != c["in_gamma"]][:12]
    assert any(c["out_gamma"] != c["in_gamma"] for c in cases)
    for c in cases:
        path = os.path.join(GOLDEN, c["graph"])
        print("case", c["name"], flush=True)
        g = dqc.load_graph(path)
        n = g["n"]
        mdo, mdi = dqc.min_degrees(c["out_gamma"], n), dqc.min_degrees(c["in_gamma"], n)
        exp = open(os.path.join(GOLDEN, c["expected"]), "rb").read() if c["expected"] else b""
        for two in (g["two"], None):
            raw = str(tmp_path / "mem.raw")
            mx = str(tmp_path / "mem.max")
            if os.path.exists(mx):
                os.remove(mx)
            sets, st = dqc.search_graph(g["out"], g["in"], two, mdo, mdi, c["out_gamma"], c["in_gamma"],
                                        c["min_size"], raw_path=raw, max_path=mx)
            assert max_text(sets) == exp, c["name"]
            assert (open(mx, "rb").read() if os.path.exists(mx) else b"") == exp, c["name"]
            assert st.tasks_expanded == c["tasks"] and st.branches == c["branches"], c["name"]
            assert st.raw_cliques == c["raw_cliques"] and st.maximal_cliques == c["maximal"], c["name"]
            fraw = str(tmp_path / "file.raw")
            dqc.search_file(path, c["out_gamma"], c["in_gamma"], c["min_size"], fraw, None)
            assert sorted(open(raw).read().splitlines()) == sorted(open(fraw).read().splitlines()), c["name"]


def test_search_graph_caller_tables_are_used(dqc, manifest):
    """The caller's tables are the thresholds: tables of gamma 0.8 passed with gamma 0.8
    give the golden 0.8 output; the same tables passed with a different gamma argument
    still enumerate by the tables up to their length (gamma then only continues them and
    sets the bound tables), so the output differs from a search at that gamma alone."""
    c = next(x for x in manifest["cases"] if x["name"] == "workload1")
    g = dqc.load_graph(os.path.join(GOLDEN, c["graph"]))
    md = dqc.min_degrees(0.8, g["n"])
    sets, _ = dqc.search_graph(g["out"], g["in"], g["two"], md, md, 0.8, 0.8, 10)
    assert max_text(sets) == open(os.path.join(GOLDEN, c["expected"]), "rb").read()
    alone, _ = dqc.search_graph(g["out"], g["in"], g["two"], None, None, 0.95, 0.95, 10)
    mixed, _ = dqc.search_graph(g["out"], g["in"], g["two"], md, md, 0.95, 0.95, 10)
    assert mixed != alone


def test_search_graph_orig_ids(dqc, manifest):
    """orig_of maps each condensed id of the result back to the caller's vertex id."""
    c = next(x for x in manifest["cases"] if x["name"] == "workload1")
    path = os.path.join(GOLDEN, c["graph"])
    g = dqc.load_graph(path)
    sets, st, orig = dqc.search_graph(g["out"], g["in"], None, None, None, 0.8, 0.8, 10, want_orig=True)
    _, _, lorig = dqc.level0(path, 0.8, 0.8, 10)
    assert orig == lorig and len(orig) == st.n_condensed
    for s in sets:  # every reported set is a quasi-clique in the caller's ids
        ids = {orig[x] for x in s}
        k = len(ids)
        for v in ids:
            assert len(set(g["out"][v]) & ids) >= -(-8 * (k - 1) // 10)
            assert len(set(g["in"][v]) & ids) >= -(-8 * (k - 1) // 10)


def test_prepare_graph_equals_prepare_file(dqc, manifest):
    c = next(x for x in manifest["cases"] if x["name"] == "workload3_g8")
    path = os.path.join(GOLDEN, c["graph"])
    g = dqc.load_graph(path)
    with dqc.Engine() as eng:
        eng.prepare_graph(g, c["out_gamma"], c["in_gamma"], c["min_size"])
        eng.run()
        a = (eng.stats().tasks_expanded, eng.stats().branches, sorted(map(tuple, eng.cliques())))
    with dqc.Engine() as eng:
        eng.prepare(path, c["out_gamma"], c["in_gamma"], c["min_size"])
        eng.run()
        b = (eng.stats().tasks_expanded, eng.stats().branches, sorted(map(tuple, eng.cliques())))
    assert a == b and a[0] == c["tasks"]


def test_binary_cache_on_search_path(dqc, manifest, tmp_path):
    """dqc_search_file on the binary cache (with and without stored two-hop lists) writes
    the golden bytes."""
    for name in ("workload1", "bigk20000"):
        c = next(x for x in manifest["cases"] if x["name"] == name)
        exp = open(os.path.join(GOLDEN, c["expected"]), "rb").read()
        for two in (True, False):
            b = str(tmp_path / f"{name}_{int(two)}.bin")
            dqc.save_graph(os.path.join(GOLDEN, c["graph"]), b, with_twohop=two)
            mx = str(tmp_path / f"{name}.max")
            cnt, st = dqc.search_file(b, c["out_gamma"], c["in_gamma"], c["min_size"], None, mx)
            assert open(mx, "rb").read() == exp and cnt == c["maximal"] and st.tasks_expanded == c["tasks"]

"""CPU tests of the product library's host side: the C-ABI exports, the generator, and
the host preprocessing (loader, two-hop lists, level 0 + condense, RemoveNonMax) against
the oracle. No GPU is touched."""
import ctypes
import hashlib
import os
import random
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "dqc.h")).read()
    return sorted(set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(dqc_\w+)\s*\(", text, re.M)))


def test_exports_every_declared_symbol(dqc):
    syms = header_symbols()
    assert len(syms) >= 15
    lib = dqc.load()
    for s in syms:
        assert hasattr(lib, s), f"{s} declared in include/dqc.h but not exported"
    assert sorted(dqc.EXPORTS) == syms


def test_generator_workload1_pinned(dqc, manifest, tmp_path):
    p = str(tmp_path / "w1.txt")
    assert dqc.generate(1, p) == 1000
    assert hashlib.sha256(open(p, "rb").read()).hexdigest() == manifest["workload1_sha256"]


def test_generator_rejects_bad_workload(dqc, tmp_path):
    with pytest.raises(dqc.DQCError):
        dqc.generate(9, str(tmp_path / "x.txt"))


def rand_graph(path, rng, n, p, self_loops=True):
    with open(path, "w") as f:
        for u in range(n):
            f.write(" ".join(str(v) for v in range(n) if (self_loops or v != u) and rng.random() < p) + "\n")


def test_loader_and_twohop_match_oracle(dqc, oracle_lib, tmp_path):
    rng = random.Random(3)
    for t in range(80):
        path = str(tmp_path / f"g{t}.txt")
        rand_graph(path, rng, rng.randint(1, 45), rng.choice([0.05, 0.2, 0.4, 0.8]))
        assert dqc.load_graph(path) == oracle_lib.graph(path, threads=2), f"trial {t}"


def test_loader_rejects_out_of_range_id(dqc, tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("1\n5\n")
    with pytest.raises(dqc.DQCError):
        dqc.load_graph(str(p))


def test_level0_matches_oracle(dqc, oracle_lib, tmp_path):
    rng = random.Random(4)
    for t in range(80):
        path = str(tmp_path / f"g{t}.txt")
        rand_graph(path, rng, rng.randint(2, 50), rng.choice([0.2, 0.4, 0.7]))
        g = rng.choice([0.5, 0.6, 0.8, 1.0])
        gi = g if rng.random() < 0.7 else rng.choice([0.5, 0.7, 0.9])
        mn = rng.randint(2, 7)
        rec, ncov, orig = dqc.level0(path, g, gi, mn)
        assert orig == oracle_lib.condense_map(path, g, gi, mn), f"trial {t}"
        assert [r[0] for r in rec] == list(range(len(rec)))
        assert all(r[1] == 0 and r[3] == 0 for r in rec)  # no members at level 0
        assert 0 <= ncov <= len(rec)


def test_remove_non_max_matches_oracle(dqc, oracle_lib, tmp_path):
    rng = random.Random(5)
    for t in range(30):
        src = tmp_path / f"raw{t}.txt"
        sets = [rng.sample(range(30), rng.randint(1, 9)) for _ in range(rng.randint(1, 80))]
        src.write_text("".join(f"{len(s)} " + "".join(f"{x} " for x in s) + "\n" for s in sets))
        a, b = tmp_path / f"a{t}.txt", tmp_path / f"b{t}.txt"
        assert dqc.remove_non_max(str(src), str(a)) == oracle_lib.rmnonmax(str(src), str(b))
        assert a.read_bytes() == b.read_bytes()


def test_remove_non_max_empty_input_writes_nothing(dqc, tmp_path):
    src = tmp_path / "empty.txt"
    src.write_text("")
    out = tmp_path / "out.txt"
    assert dqc.remove_non_max(str(src), str(out)) == 0
    assert not out.exists()


def test_search_rejects_bad_arguments(dqc, tmp_path):
    """reference program.cu:44-57: gamma in [0.5, 1], min size > 1 (checked before any GPU use)."""
    p = tmp_path / "g.txt"
    p.write_text("1\n0\n")
    for go, gi, mn in ((0.4, 0.8, 3), (0.8, 1.2, 3), (0.8, 0.8, 1)):
        with pytest.raises(dqc.DQCError, match="DQC_E_ARG"):
            dqc.search_file(str(p), go, gi, mn)
    with pytest.raises(dqc.DQCError, match="DQC_E_IO"):
        dqc.search_file(str(tmp_path / "missing.txt"), 0.8, 0.8, 3)


def test_min_degrees_closed_form(dqc):
    """dqc_min_degrees = h_calculate_minimum_degrees: n + 1 entries of ceil(gamma (k - 1))
    in IEEE double (0.56 * 25 = 14.000000000000002 gives 15)."""
    import math
    for g in (0.5, 0.56, 0.8, 0.9, 0.95, 1.0):
        assert dqc.min_degrees(g, 60) == [0] + [math.ceil(g * (k - 1)) for k in range(1, 61)]
    assert dqc.min_degrees(0.56, 30)[26] == 15
    assert dqc.min_degrees(0.9, 0) == [0]
    lib = dqc.load()
    arr = (ctypes.c_int32 * 4)()
    assert lib.dqc_min_degrees(0.9, -1, arr) == 3 and lib.dqc_min_degrees(1.5, 3, arr) == 3


def test_binary_cache_round_trip(dqc, tmp_path):
    """The binary graph cache gives back exactly what the text loader (and the two-hop
    builder) produce, with or without stored two-hop lists; level 0 read from it equals
    level 0 from the text file."""
    rng = random.Random(5)
    for t in range(12):
        path = str(tmp_path / f"g{t}.txt")
        rand_graph(path, rng, rng.randint(1, 60), rng.choice([0.05, 0.2, 0.5]))
        ref = dqc.load_graph(path)
        for two in (True, False):
            b = str(tmp_path / f"g{t}_{int(two)}.bin")
            dqc.save_graph(path, b, with_twohop=two)
            assert dqc.load_graph(b) == ref, (t, two)
            nt = dqc.load_graph(b, with_twohop=False)
            assert nt["out"] == ref["out"] and nt["in"] == ref["in"]
    g = str(tmp_path / "w3.txt")
    dqc.generate(3, g, 4)
    b = str(tmp_path / "w3.bin")
    dqc.save_graph(g, b)
    assert dqc.level0(b, 0.85, 0.85, 10) == dqc.level0(g, 0.85, 0.85, 10)


def test_binary_cache_rejects_damage(dqc, tmp_path):
    g = str(tmp_path / "w3.txt")
    dqc.generate(3, g, 1)
    b = tmp_path / "w3.bin"
    dqc.save_graph(g, str(b))
    data = bytearray(b.read_bytes())
    bad = tmp_path / "bad.bin"
    for pos in (30, len(data) // 2, len(data) - 3):  # flipped bytes: checksum (or bounds) fail
        d = bytearray(data)
        d[pos] ^= 0x41
        bad.write_bytes(bytes(d))
        with pytest.raises(dqc.DQCError):
            dqc.load_graph(str(bad))
    bad.write_bytes(bytes(data[:len(data) // 3]))  # truncated
    with pytest.raises(dqc.DQCError):
        dqc.load_graph(str(bad))


def test_search_graph_checks_its_input(dqc):
    """dqc_search_graph refuses a CSR that breaks the loader's invariants (unsorted or
    duplicate ids, self loops, ids out of range, in-lists not the transpose), with
    DQC_E_ARG and without touching the GPU; a graph with no level-0 survivor returns no
    set, also without the GPU."""
    out = [[1, 2], [2], [0]]
    inn = [[2], [0], [0, 1]]
    for o, i in (([[2, 1], [2], [0]], inn), ([[1, 1], [2], [0]], inn), ([[0, 1], [2], [0]], inn),
                 ([[1, 3], [2], [0]], inn), (out, [[2], [0], [1]])):
        with pytest.raises(dqc.DQCError, match="DQC_E_ARG"):
            dqc.search_graph(o, i, None, None, None, 0.9, 0.9, 2)
    sets, st = dqc.search_graph(out, inn, None, None, None, 0.9, 0.9, 3)
    assert sets == [] and st.n_condensed == 0 and st.maximal_cliques == 0


def test_generator_workload3_all_graphs_pinned(dqc, tmp_path):
    """The batching workload's 256 graphs: the generator's files have the SHA-256 the
    per-graph expectations were made from."""
    import json
    from tests.conftest import GOLDEN
    exp = json.load(open(os.path.join(GOLDEN, "workload3_all.json")))
    assert len(exp["graphs"]) == 256
    for e in exp["graphs"][::5]:
        p = str(tmp_path / f"g{e['graph']}.txt")
        assert dqc.generate(3, p, e["graph"]) == e["n"]
        assert hashlib.sha256(open(p, "rb").read()).hexdigest() == e["graph_sha256"], e["graph"]

"""Pins level 0 (peel, cascade, cover vertex, covered set, enumeration order, condense)
of the oracle and of the product's host code against a restatement written from the
sets these steps compute, not from their loops.

Reference: src/host_functions.cu:461-686 (h_initialize_tasks), :690-934
(h_condense_graph, relabel at :879-881) and inc/host_functions.hpp:74-128
(h_comp_vert_Q). What those loops compute:
  * C0 = vertices with out-degree >= md_out[min], in-degree >= md_in[min] and a two-hop
    list of at least min - 1 entries (:495-511);
  * the peel rounds (:518-576) and the cascade (:581-622) only ever remove a vertex that
    lacks md_out[min] out-arcs or md_in[min] in-arcs inside a superset of what is left, so
    the survivors S are the largest subset of C0 in which every vertex keeps both degrees
    inside S (the greatest fixed point; removal order does not change it). The
    restatement removes every failing vertex at once per round instead;
  * cover vertex (:626-640): the survivor with the largest min(out, in) inside S; the
    scan runs over ascending ids with a strict '>' from 0, so the smallest id wins ties;
  * covered (:643-665): survivors u != cover with cover -> u and u -> cover;
  * enumeration order: covered, then other candidates, then the cover vertex; within a
    class min(out, in) descending, then the two-hop list length of the uncondensed graph
    descending (it is sorted before condensing), then id descending;
  * condensed id = rank in that order; the record is {rank, 0, |out & S|, 0, |in & S|,
    |L2 & S|} (lvl2adj recomputed by the condense, :786).

Plausible slips each fail here: a '>=' in the cover argmax (largest id wins ties), an
ascending id tie-break, a wrong class precedence, the condensed instead of the original
two-hop length as sort key, a missing cascade after the last peel round.
"""
import math
import random

import pytest


def md(g, k):
    return 0 if k == 0 else math.ceil(g * (k - 1))


def level0_restated(n, out, inn, two, go, gi, mn):
    """Returns (orig_of, records, ncov) of the level-0 task."""
    mo, mi = md(go, mn), md(gi, mn)
    S = {v for v in range(n) if len(out[v]) >= mo and len(inn[v]) >= mi and len(two[v]) >= mn - 1}
    while True:
        bad = {v for v in S if len(out[v] & S) < mo or len(inn[v] & S) < mi}
        if not bad:
            break
        S -= bad
    if not S:
        return [], [], 0
    dout = {v: len(out[v] & S) for v in S}
    din = {v: len(inn[v] & S) for v in S}
    cover = min(S, key=lambda v: (-min(dout[v], din[v]), v))
    covered = {u for u in S if u != cover and u in out[cover] and cover in out[u]}

    def cls(v):
        return 0 if v in covered else (2 if v == cover else 1)

    order = sorted(S, key=lambda v: (cls(v), -min(dout[v], din[v]), -len(two[v]), -v))
    rank = {v: i for i, v in enumerate(order)}
    records = [(rank[v], 0, dout[v], 0, din[v], len(two[v] & S)) for v in order]
    return order, records, len(covered)


def random_digraph(rng, n):
    """Sparse background plus 0-3 dense planted groups (so that peel rounds, cascades and
    degree ties all occur), or a uniform density."""
    kind = rng.random()
    arcs = [set() for _ in range(n)]
    if kind < 0.3:
        p = rng.choice([0.1, 0.3, 0.6, 0.9])
        for u in range(n):
            arcs[u] = {v for v in range(n) if v != u and rng.random() < p}
        return arcs
    pb = rng.choice([0.02, 0.05, 0.1])
    for u in range(n):
        arcs[u] = {v for v in range(n) if v != u and rng.random() < pb}
    perm = list(range(n))
    rng.shuffle(perm)
    at = 0
    for _ in range(rng.randint(1, 3)):
        k = rng.randint(4, 14)
        grp = perm[at:at + k]
        at += k
        pd = rng.choice([0.8, 0.9, 1.0])  # 1.0: complete groups, all keys tie but the id
        for u in grp:
            for v in grp:
                if u != v and rng.random() < pd:
                    arcs[u].add(v)
    return arcs


def write_graph(path, arcs):
    with open(path, "w") as f:
        for a in arcs:
            f.write(" ".join(map(str, sorted(a))) + "\n")


PARAMS = [(0.8, 0.8, 5), (0.9, 0.9, 6), (0.6, 0.6, 4), (0.9, 0.5, 5), (0.5, 0.9, 5), (0.95, 0.9, 7),
          (1.0, 0.7, 4), (0.75, 0.75, 3)]


def cases(seed, count):
    rng = random.Random(seed)
    for t in range(count):
        n = rng.randint(6, 50)
        arcs = random_digraph(rng, n)
        go, gi, mn = rng.choice(PARAMS)
        yield t, arcs, go, gi, mn


def restated_for(oracle_lib, path, arcs, go, gi, mn):
    n = len(arcs)
    inn = [set() for _ in range(n)]
    for u in range(n):
        for v in arcs[u]:
            inn[v].add(u)
    # two-hop lists from the oracle's loader; they are pinned on their own by
    # tests/test_oracle.py::test_twohop_vs_restatement
    two = [set(x) for x in oracle_lib.graph(path, with_twohop=True)["two"]]
    return level0_restated(n, arcs, inn, two, go, gi, mn)


def test_oracle_level0_vs_restatement(oracle_lib, tmp_path):
    nonempty = ties = peeled = 0
    for t, arcs, go, gi, mn in cases(41, 90):
        path = str(tmp_path / f"l{t}.txt")
        write_graph(path, arcs)
        orig, recs, ncov = restated_for(oracle_lib, path, arcs, go, gi, mn)
        got_orig = oracle_lib.condense_map(path, go, gi, mn)
        assert got_orig == orig, f"trial {t}: condense map (gammas {go}/{gi}, min {mn})"
        tk = oracle_lib.Tasks(path, go, gi, mn, threads=1)
        try:
            root, oncov = tk.root()
        finally:
            tk.close()
        assert root == recs, f"trial {t}: root records"
        assert oncov == ncov, f"trial {t}: covered count"
        if recs:
            nonempty += 1
            keys = [(min(r[2], r[4])) for r in recs]
            ties += len(keys) != len(set(keys))
            peeled += len(recs) < sum(1 for a in arcs if a) // 2
    # the sample must exercise what it pins
    assert nonempty >= 60 and ties >= 30 and peeled >= 10, (nonempty, ties, peeled)


def test_product_level0_vs_restatement(dqc, oracle_lib, tmp_path):
    """The product's host level 0 (dqc_level0, what the GPU path is fed) against the same
    restatement, on a different sample."""
    nonempty = 0
    for t, arcs, go, gi, mn in cases(43, 60):
        path = str(tmp_path / f"p{t}.txt")
        write_graph(path, arcs)
        orig, recs, ncov = restated_for(oracle_lib, path, arcs, go, gi, mn)
        got, gncov, gorig = dqc.level0(path, go, gi, mn)
        assert gorig == orig and got == recs and gncov == ncov, f"trial {t}"
        nonempty += bool(recs)
    assert nonempty >= 40


@pytest.mark.parametrize("slip", ["cover_ge", "id_ascending", "condensed_l2_key", "no_cascade"])
def test_restatement_detects_slips(oracle_lib, tmp_path, slip):
    """The restatement is sharp enough: each plausible slip, planted in a copy of it, is
    caught on this sample (it disagrees with the oracle on some trial)."""

    def slipped(n, out, inn, two, go, gi, mn):
        mo, mi = md(go, mn), md(gi, mn)
        S = {v for v in range(n) if len(out[v]) >= mo and len(inn[v]) >= mi and len(two[v]) >= mn - 1}
        if slip == "no_cascade":  # one simultaneous round only
            S -= {v for v in S if len(out[v] & S) < mo or len(inn[v] & S) < mi}
        else:
            while True:
                bad = {v for v in S if len(out[v] & S) < mo or len(inn[v] & S) < mi}
                if not bad:
                    break
                S -= bad
        if not S:
            return [], [], 0
        dout = {v: len(out[v] & S) for v in S}
        din = {v: len(inn[v] & S) for v in S}
        tie = (lambda v: -v) if slip == "cover_ge" else (lambda v: v)
        cover = min(S, key=lambda v: (-min(dout[v], din[v]), tie(v)))
        covered = {u for u in S if u != cover and u in out[cover] and cover in out[u]}
        cls = lambda v: 0 if v in covered else (2 if v == cover else 1)
        l2key = (lambda v: -len(two[v] & S)) if slip == "condensed_l2_key" else (lambda v: -len(two[v]))
        idkey = (lambda v: v) if slip == "id_ascending" else (lambda v: -v)
        order = sorted(S, key=lambda v: (cls(v), -min(dout[v], din[v]), l2key(v), idkey(v)))
        rank = {v: i for i, v in enumerate(order)}
        return order, [(rank[v], 0, dout[v], 0, din[v], len(two[v] & S)) for v in order], len(covered)

    caught = False
    for t, arcs, go, gi, mn in cases(41, 90):
        path = str(tmp_path / f"s{t}.txt")
        write_graph(path, arcs)
        n = len(arcs)
        inn = [set() for _ in range(n)]
        for u in range(n):
            for v in arcs[u]:
                inn[v].add(u)
        two = [set(x) for x in oracle_lib.graph(path, with_twohop=True)["two"]]
        if slipped(n, arcs, inn, two, go, gi, mn)[0] != oracle_lib.condense_map(path, go, gi, mn):
            caught = True
            break
    assert caught, slip

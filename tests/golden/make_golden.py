"""Regenerates the golden fixtures used by the parity tests.

Inputs are seeded random digraphs (plain text, reference adjacency format) and the
synthetic workloads the generator writes; expected outputs come from the CPU oracle
(oracle/dqc_oracle.cpp, a restatement of the reference's CPU path). Cases are kept only
when the oracle finishes them in well under a second, so the GPU tests stay short.

Writes tests/golden/graphs/*.txt, tests/golden/expected/*.max and tests/golden/manifest.json.
Usage: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import random
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402

GRAPHS = os.path.join(HERE, "graphs")
EXPECTED = os.path.join(HERE, "expected")


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def write_random(path, rng, n, p):
    with open(path, "w") as f:
        for u in range(n):
            f.write(" ".join(str(v) for v in range(n) if v != u and rng.random() < p) + "\n")


def write_reciprocal(path, rng, n, arcs_per_vertex, back):
    """n vertices, n * arcs_per_vertex random arcs, each reversed too with probability back."""
    adj = [set() for _ in range(n)]
    for _ in range(n * arcs_per_vertex):
        a, b = rng.randrange(n), rng.randrange(n)
        if a != b:
            adj[a].add(b)
            if rng.random() < back:
                adj[b].add(a)
    with open(path, "w") as f:
        for u in range(n):
            f.write(" ".join(str(v) for v in sorted(adj[u])) + "\n")


def oracle_case(name, path, go, gi, mn, cap=3.0):
    exp = os.path.join(EXPECTED, name + ".max")
    if os.path.exists(exp):
        os.remove(exp)
    t = time.time()
    r, st = orc.run(path, go, gi, mn, None, exp, threads=1, time_cap_s=cap)
    dt = time.time() - t
    if r < 0:
        return None
    return {
        "name": name, "graph": os.path.relpath(path, HERE), "out_gamma": go, "in_gamma": gi, "min_size": mn,
        "maximal": r, "tasks": st.tasks, "branches": st.branches, "raw_cliques": st.raw_cliques,
        "n_condensed": st.n_condensed, "expected": os.path.relpath(exp, HERE) if r > 0 else None,
        "sha256": sha(exp) if r > 0 else None, "oracle_s": round(dt, 3),
    }


def main():
    os.makedirs(GRAPHS, exist_ok=True)
    os.makedirs(EXPECTED, exist_ok=True)
    import cudqc_amd as dq  # generator only (writes input files)
    cases = []
    # workload 1 (ER 1000 + 5 planted sets): small enough for the oracle
    c1 = os.path.join(GRAPHS, "workload1.txt")
    dq.generate(1, c1)
    cases.append(oracle_case("workload1", c1, 0.8, 0.8, 10))
    # a few graphs of the batching workload
    for g in (0, 1, 2, 3, 5, 8, 13, 21):
        path = os.path.join(GRAPHS, f"workload3_g{g}.txt")
        dq.generate(3, path, g)
        c = oracle_case(f"workload3_g{g}", path, 0.85, 0.85, 10)
        if c:
            cases.append(c)
    # seeded random digraphs over a spread of densities, gammas and minimum sizes
    rng = random.Random(20261202)
    kept = 0
    trial = 0
    while kept < 40 and trial < 400:
        n = rng.randint(6, 70)
        p = rng.choice([0.15, 0.3, 0.45, 0.6, 0.75, 0.9])
        g = rng.choice([0.5, 0.6, 0.7, 0.75, 0.8, 0.9, 1.0])
        gi = g if rng.random() < 0.75 else rng.choice([0.5, 0.6, 0.7, 0.75, 0.8, 0.9, 1.0])
        mn = rng.randint(2, 9)
        name = f"rand{trial:03d}"
        path = os.path.join(GRAPHS, name + ".txt")
        write_random(path, rng, n, p)
        trial += 1
        c = oracle_case(name, path, g, gi, mn, cap=2.0)
        if c is None or c["oracle_s"] > 1.0 or c["tasks"] < 2:
            os.remove(path)
            exp = os.path.join(EXPECTED, name + ".max")
            if os.path.exists(exp):
                os.remove(exp)
            continue
        cases.append(c)
        kept += 1
    # a sparse graph that condenses to more than 16384 vertices: the kernels then keep
    # their id bitmaps in HBM instead of LDS
    path = os.path.join(GRAPHS, "bigk20000.txt")
    write_reciprocal(path, random.Random(7), 20000, 4, 0.7)
    c = oracle_case("bigk20000", path, 0.6, 0.6, 4, cap=60.0)
    assert c and c["n_condensed"] > 16384, c
    cases.append(c)
    cases = [c for c in cases if c]
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump({"workload1_sha256": sha(c1), "cases": cases}, f, indent=1)
    print(f"{len(cases)} cases")


if __name__ == "__main__":
    main()

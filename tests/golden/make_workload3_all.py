"""Per-graph expectations for all 256 graphs of the batching workload (directed ER
background of out-degree 8 on n log-uniform in [200, 5000], 1-10 planted sets of 10-25
vertices at density 0.92; gamma 0.85, min size 10): the SHA-256 of each graph's maximal
file and its tree counts, from the CPU oracle (oracle/dqc_oracle.cpp). The graphs
themselves are not stored: the library's seeded generator writes them (its output for
graphs 0..255 is pinned here by the SHA-256 of each file).

Writes tests/golden/workload3_all.json.
Usage: python tests/golden/make_workload3_all.py
"""
import hashlib
import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import cudqc_amd as dq  # noqa: E402  (the generator only)
from oracle import oracle as orc  # noqa: E402

GO, GI, MN = 0.85, 0.85, 10


def sha_file(p):
    return hashlib.sha256(open(p, "rb").read()).hexdigest() if os.path.exists(p) else ""


def main():
    out = {"gamma_out": GO, "gamma_in": GI, "min_size": MN, "graphs": []}
    with tempfile.TemporaryDirectory() as d:
        for g in range(256):
            path = os.path.join(d, f"g{g}.txt")
            n = dq.generate(3, path, g)
            mx = os.path.join(d, f"g{g}.max")
            r, st = orc.run(path, GO, GI, MN, None, mx, threads=os.cpu_count() or 1)
            assert r >= 0, (g, r)
            out["graphs"].append({"graph": g, "n": n, "graph_sha256": sha_file(path), "maximal": int(r),
                                  "sha256": sha_file(mx), "tasks": int(st.tasks), "branches": int(st.branches),
                                  "raw_cliques": int(st.raw_cliques), "n_condensed": int(st.n_condensed)})
            print(g, n, r, st.tasks, flush=True)
    with open(os.path.join(HERE, "workload3_all.json"), "w") as f:
        json.dump(out, f, indent=0)
        f.write("\n")


if __name__ == "__main__":
    main()

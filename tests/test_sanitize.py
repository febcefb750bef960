"""The product's host code and the oracle built with AddressSanitizer and
UndefinedBehaviorSanitizer (g++, CPU only) and run on golden inputs: the text loader, the
OpenMP two-hop builder (4 threads), the binary graph cache round trip, level 0 +
condense, the maximal-set filter, the generator and the oracle's whole CPU path. Any
sanitizer report fails the test; the results must also match the plain builds'."""
import os
import shutil
import subprocess

import pytest

from tests.conftest import GOLDEN, ROOT

CSRC = os.path.join(ROOT, "cudqc_amd", "csrc")
SOURCES = [os.path.join(ROOT, "tests", "sanitize", "driver.cpp"), os.path.join(CSRC, "graph.cpp"),
           os.path.join(CSRC, "rmnonmax.cpp"), os.path.join(CSRC, "gen.cpp"),
           os.path.join(ROOT, "oracle", "dqc_oracle.cpp")]


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("san") / "driver")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fopenmp", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-I", os.path.join(ROOT, "include")] + SOURCES + ["-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    return out


def run(driver, *args, timeout=600):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="4")
    r = subprocess.run([driver] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout, env=env)
    report = "ERROR: AddressSanitizer" in r.stderr or "runtime error:" in r.stderr or "LeakSanitizer" in r.stderr
    assert r.returncode == 0 and not report, (args, r.returncode, r.stderr[-4000:])
    return r.stdout


CASES = ["workload1", "workload3_g0", "workload3_g8", "rand017", "rand069", "bigk20000"]


@pytest.mark.parametrize("name", CASES)
def test_host_preprocessing_clean(driver, manifest, tmp_path, dqc, name):
    c = next(x for x in manifest["cases"] if x["name"] == name)
    g = os.path.join(GOLDEN, c["graph"])
    out = run(driver, "host", g, c["out_gamma"], c["in_gamma"], c["min_size"], 4, tmp_path / "g.bin")
    k = int(out.split("level0 k=")[1].split()[0])
    recs, ncov, _ = dqc.level0(g, c["out_gamma"], c["in_gamma"], c["min_size"])
    assert k == len(recs) == c["n_condensed"]
    assert f"ncov={ncov} " in out


@pytest.mark.parametrize("name", ["workload1", "workload3_g2", "rand017"])
def test_oracle_whole_path_clean(driver, manifest, tmp_path, name):
    c = next(x for x in manifest["cases"] if x["name"] == name)
    mx = tmp_path / "o.max"
    out = run(driver, "oracle", os.path.join(GOLDEN, c["graph"]), c["out_gamma"], c["in_gamma"], c["min_size"], 4, mx)
    assert f"maximal={c['maximal']} tasks={c['tasks']} branches={c['branches']}" in out
    exp = open(os.path.join(GOLDEN, c["expected"]), "rb").read() if c["expected"] else b""
    assert (mx.read_bytes() if mx.exists() else b"") == exp


def test_rmnonmax_and_generator_clean(driver, manifest, tmp_path):
    p = tmp_path / "w3.txt"
    run(driver, "gen", 3, 5, p)
    raw = tmp_path / "raw.txt"
    raw.write_text("3 1 2 3 \n2 2 1 \n4 9 8 7 6 \n3 6 7 8 \n1 5 \n3 3 2 1 \n")
    out = run(driver, "rmnonmax", raw, tmp_path / "max.txt")
    assert "rmnonmax 3" in out
    assert (tmp_path / "max.txt").read_text() == "3 1 2 3 \n1 5 \n4 6 7 8 9 \n"

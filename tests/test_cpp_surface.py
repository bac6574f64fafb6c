"""The C-ABI library and the reference-style C++ surface.

- every function declared in include/*.h (C linkage) is exported by
  libwamca_mi.so (no compute calls);
- tests/cpp/test_mlkernel restates the reference's gtest cases
  (tests/mtrand.Test.cpp, tests/mlproblem.Test.cpp, tests/mlsolution.Test.cpp,
  tests/WamcaExperiment.Test.cpp) against include/wamca/mlkernel.hpp;
- one WAMCAExperiment pass over the five kernels equals the oracle's chain of
  neighborhood steps on the same start tour.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from helpers import GOLDEN, ROOT, load_tsplib_coords
from oracle import oracle as O

wm = pytest.importorskip("wamca_mi")

INCLUDE = os.path.join(ROOT, "include")
PROG = os.path.join(ROOT, "tests", "cpp", "_build", "test_mlkernel")
INST = os.path.join(GOLDEN, "instances")


def _declared_c_functions(path):
    """Names of the functions a C header declares (comments, preprocessor lines
    and extern "C" braces stripped; one declaration per ';'-terminated statement)."""
    src = open(path).read()
    src = re.sub(r"/\*.*?\*/", " ", src, flags=re.S)
    src = re.sub(r"//[^\n]*", " ", src)
    src = re.sub(r"^\s*#[^\n]*", " ", src, flags=re.M)
    src = re.sub(r'extern\s+"C"\s*\{', " ", src)
    names = set()
    for stmt in src.split(";"):
        stmt = stmt.replace("}", " ").strip()
        if not stmt or "{" in stmt or stmt.startswith("typedef"):
            continue
        m = re.match(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\((.*)\)$", stmt, flags=re.S)
        if m:
            names.add(m.group(1))
    return names


def test_header_symbols_exported():
    lib = C.CDLL(wm.LIB_PATH)
    seen = {}
    for h in ("mlp_gpu.h", "mlp_host.h", "dvnd.h", "mlp_serve.h"):
        names = _declared_c_functions(os.path.join(INCLUDE, h))
        assert names, h
        seen[h] = names
        missing = [n for n in sorted(names) if not hasattr(lib, n)]
        assert not missing, (h, missing)
    # spot-check that the parser saw the entry points that matter
    assert {"mlp_ctx_create", "mlp_eval", "mlp_merge_apply", "mlp_dvnd"} <= seen["mlp_gpu.h"]
    assert {"bestNeighbor", "bestNeighborSimple", "noConflict", "getNoConflictMoves",
            "applyMoves", "segFault"} == seen["dvnd.h"]


def _run(*args, timeout=120):
    if not os.path.exists(PROG):
        pytest.fail("tests/cpp/_build/test_mlkernel is not built (run __graft_entry__.build())")
    return subprocess.run([PROG, *args], capture_output=True, text=True, timeout=timeout)


def test_cpp_surface_cpu_cases():
    r = _run("cpu", INST)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout


@pytest.mark.gpu
def test_cpp_surface_wamca_experiment_367403():
    r = _run("gpu", INST)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout


def _parse_chain(out):
    lines = out.strip().splitlines()
    start = np.array([int(x) for x in lines[0].split()[1:]], dtype=np.int32)
    cost = int(lines[1].split()[1])
    final = np.array([int(x) for x in lines[2].split()[1:]], dtype=np.int32)
    nm = int(lines[3].split()[1])
    moves = [tuple(int(x) for x in ln.split()) for ln in lines[4:4 + nm]]
    return start, cost, final, moves


@pytest.mark.gpu
@pytest.mark.parametrize("name,seed,rowmap", [("01_berlin52.tsp", 3, 1), ("02_kroD100.tsp", 5, 1),
                                              ("03_pr226.tsp", 2, 0), ("07_rat784.tsp", 1, 1)])
def test_cpp_surface_chain_matches_oracle(name, seed, rowmap):
    r = _run("chain", os.path.join(INST, name), str(seed), str(rowmap))
    assert r.returncode == 0, r.stdout + r.stderr
    start, cost, final, moves = _parse_chain(r.stdout)
    xs, ys = load_tsplib_coords(name)
    W = O.weights(xs, ys, 0)
    assert np.array_equal(start, O.random_tour(O.MT(seed), W, 0.5))
    t = start.copy()
    ref_moves = []
    for nb in range(5):
        rows = O.eval_rows(nb, rowmap, W, t)
        kept, _ = O.merge(nb, rows["cost"], rows["i"], rows["j"])
        ref_moves += [(nb, int(rows["i"][q]), int(rows["j"][q]), int(rows["cost"][q])) for q in kept]
        t = O.step(nb, rowmap, W, t)[0]
    assert moves == ref_moves
    assert np.array_equal(final, t)
    assert cost == O.latency_u32(W, t)


def _total_lines(out):
    res = {}
    for ln in out.splitlines():
        f = ln.split()
        if len(f) == 16 and f[2] == "elems":
            res[(int(f[0]), f[1])] = {
                "elems": int(f[3]), "best": [int(f[5]), int(f[6]), int(f[7])],
                "improving": int(f[9]), "gpu": [int(f[11]), int(f[12])],
                "cpu": [int(f[14]), int(f[15])]}
    return res


@pytest.mark.gpu
@pytest.mark.parametrize("logname", ["results-inst100.txt", "results-inst226.txt",
                                     "results-inst1000.txt"])
def test_cpp_surface_total_kernels_match_logs(logname):
    """MLKernel{Swap,2Opt,OrOpt}(problem, isTotal = true) with the logged block
    sizes reproduce each logged TOTAL record: best entry, improving entries,
    GPU-GPU and GPU-CPU merged (sum, count) -- including the runs where the
    reference's CPU graph stopped at 120000 moves and the two merges differ."""
    from helpers import results_logs
    log = next(l for l in results_logs() if l["log"] == logname)
    sols = log["solutions"]
    xs, _ = load_tsplib_coords(log["instance"])
    N = len(xs) + 1
    specs = []
    for name, k in sols[0]["kernels"].items():
        t = k.get("total", {})
        if "best" in t:
            B = t.get("block") or t["move_elems"] // O.rows(k["nb"], 0, N)
            specs.append(f"{k['nb']}:{B}")
    r = _run("total", os.path.join(INST, log["instance"]), str(log["seed"]), str(len(sols)),
             *specs, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    got = _total_lines(r.stdout)
    checked = 0
    for sol in sols:
        for name, k in sol["kernels"].items():
            p = k.get("total", {})
            if "best" not in p:
                continue
            g = got[(sol["index"], name)]
            assert g["elems"] == p["move_elems"]
            assert g["best"] == p["best"]
            assert g["improving"] == p["graph"][0] or (p["graph"][0] == 120000 and
                                                       g["improving"] > 120000)
            assert g["gpu"] == p["merge_gpu"], (sol["index"], name)
            assert g["cpu"] == p["merge_cpu"], (sol["index"], name)
            checked += 1
    assert checked == 3 * len(sols)

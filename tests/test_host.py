"""Host-side checks without a GPU: the C-ABI library loads and exports every
declared entry point, file formats round-trip, the generator keeps its
invariants, the CLI/bench argument handling."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "brainmatch.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(bm_[a-z0-9_]+)\s*\(", text,
                                 flags=re.M)))


def test_library_exports_every_declared_symbol(bm):
    import ctypes
    lib = ctypes.CDLL(bm.LIB_PATH)
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert bm.api_version() == 1


def test_no_device_error_is_reported(bm):
    if os.path.exists("/dev/kfd"):
        pytest.skip("a device is present")
    male, female, truth, pi0 = bm.synth_pair(50, 4.0, bm.WEIGHTS_GEOMETRIC, 0.3, seed=1)
    with pytest.raises(bm.BMError) as e:
        bm.Context(50, male, female)
    assert e.value.code == -6


def test_synth_invariants(bm):
    n = 3000
    male, female, truth, pi0 = bm.synth_pair(n, 20.0, bm.WEIGHTS_GEOMETRIC, 0.3, seed=4)
    for g in (male, female):
        assert (g.src != g.dst).all()
        key = g.src.astype(np.int64) * (n + 1) + g.dst
        assert np.unique(key).size == key.size
        assert (g.w >= 1).all()
        assert g.src.min() >= 1 and g.dst.max() <= n
    assert abs(male.m - 20 * n) <= n  # realised edge count hits the target within one vertex' degree
    assert female.m == male.m
    assert np.array_equal(np.sort(truth[1:]), np.arange(1, n + 1))
    assert np.array_equal(np.sort(pi0[1:]), np.arange(1, n + 1))
    moved = (pi0[1:] != truth[1:]).mean()
    assert 0.2 < moved <= 0.3
    # 80% of male edges survive the relabelling unchanged
    fk = set(zip(female.src.tolist(), female.dst.tolist(), female.w.tolist()))
    kept = sum((truth[s], truth[d], w) in fk for s, d, w in zip(male.src, male.dst, male.w))
    assert 0.78 < kept / male.m < 0.85
    # deterministic in the seed
    m2, f2, t2, p2 = bm.synth_pair(n, 20.0, bm.WEIGHTS_GEOMETRIC, 0.3, seed=4)
    assert np.array_equal(m2.dst, male.dst) and np.array_equal(p2, pi0)


def test_lognormal_weights(bm):
    male, female, truth, pi0 = bm.synth_pair(2000, 30.0, bm.WEIGHTS_LOGNORMAL, 0.3, seed=2)
    w = male.w.astype(np.float64)
    assert w.min() >= 1
    # median of round(exp(0.5 + 1.2 Z)) is round(e^0.5) = 2
    assert np.median(w) == 2
    assert w.max() < 32767


def test_csv_roundtrip(bm, tmp_path):
    male, female, truth, pi0 = bm.synth_pair(100, 5.0, bm.WEIGHTS_GEOMETRIC, 0.3, seed=3)
    p = tmp_path / "g.csv"
    bm.write_graph_csv(str(p), male)
    g = bm.load_graph_csv(str(p))
    assert g.n == male.src.max() or g.n == male.dst.max() or g.n <= 100
    assert np.array_equal(g.src, male.src) and np.array_equal(g.dst, male.dst) and np.array_equal(g.w, male.w)
    q = tmp_path / "pi.csv"
    bm.write_mapping_csv(str(q), pi0)
    assert open(q).readline().strip() == "Male Node ID,Female Node ID"
    assert np.array_equal(bm.load_mapping_csv(str(q), 100), pi0)


def test_mapping_formats(bm, tmp_path):
    p = tmp_path / "m.csv"
    p.write_text("Male ID,Female ID\n3,5\nm4,f1\n 2 , 2\nbad\n1,\n")
    pi = bm.load_mapping_csv(str(p), 6)
    assert list(pi) == [0, 0, 2, 5, 1, 0, 0]
    p.write_text("h\n9,1\n")
    with pytest.raises(bm.BMError):
        bm.load_mapping_csv(str(p), 6)


def test_graph_csv_rules(bm, tmp_path):
    p = tmp_path / "g.csv"
    p.write_text("from,to,w\nm1,m2,5\nf2,3,-4\n1,2,7\nx,y,1\n4,5\n6,7,z\n")
    g = bm.load_graph_csv(str(p))
    assert list(g.src) == [1, 2, 1] and list(g.dst) == [2, 3, 2] and list(g.w) == [5, -4, 7] and g.n == 3


def test_greedy_deterministic_and_bijective(bm):
    male, female, truth, pi0 = bm.synth_pair(400, 10.0, bm.WEIGHTS_GEOMETRIC, 0.3, seed=8)
    a = bm.greedy_match(400, male, female)
    b = bm.greedy_match(400, male, female)
    assert np.array_equal(a, b)
    assert np.array_equal(np.sort(a[1:]), np.arange(1, 401))


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_tools_usage_errors(bm):
    for tool in ("graph-re-alignment-gpu-swap", "graph-re-alignment-gpu-3-swap", "graph-re-alignment-gpu"):
        r = subprocess.run([os.path.join(bm.BIN_DIR, tool)], capture_output=True, text=True)
        assert r.returncode == 1 and "Usage" in r.stderr


def test_tool_reports_missing_input(bm, tmp_path):
    r = subprocess.run([os.path.join(bm.BIN_DIR, "graph-re-alignment-gpu-swap"), str(tmp_path / "nope.csv"),
                        str(tmp_path / "nope2.csv"), str(tmp_path / "pi.csv"), str(tmp_path / "out.csv")],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "Failed to load male graph" in r.stderr


CONVERT_CASES = sorted(os.listdir(os.path.join(ROOT, "tests", "golden", "convert")))


@pytest.mark.parametrize("case", CONVERT_CASES)
def test_matching_converters_match_reference(bm, tmp_path, case):
    """transform-solution / convert-VNC-matching write the same bytes, print
    the same lines and exit with the same status as the reference tools
    (fixtures from the compiled reference, tests/golden/make_golden_convert.py):
    CRLF lines, a last line without newline, a path without extension or with
    two dots, a header-only file, a non-digit id, a blank line, a line
    without prefix."""
    d = os.path.join(ROOT, "tests", "golden", "convert", case)
    tool, rc = open(os.path.join(d, "rc.txt")).read().split()
    inp = [f for f in os.listdir(d) if f.startswith("input_")][0]
    name = inp[len("input_"):]
    with open(os.path.join(d, inp), "rb") as f:
        (tmp_path / name).write_bytes(f.read())
    r = subprocess.run([os.path.join(bm.BIN_DIR, tool), name], cwd=tmp_path, capture_output=True)
    assert r.returncode == int(rc), r.stderr
    assert r.stdout == open(os.path.join(d, "stdout.txt"), "rb").read()
    dot = name.rfind(".")
    out = (name[:dot] if dot >= 0 else name) + ("_matching.csv" if tool == "transform-solution" else "_generic.csv")
    expected = os.path.join(d, "expected_output")
    if os.path.exists(expected):
        assert (tmp_path / out).read_bytes() == open(expected, "rb").read()
    else:
        assert not (tmp_path / out).exists()


def test_converters_round_trip(bm, tmp_path):
    """generic -> prefixed -> generic returns the same lines; the loader
    reads all three forms as the same matching."""
    pi = np.concatenate([[0], np.random.default_rng(4).permutation(np.arange(1, 51))]).astype(np.int32)
    (tmp_path / "sol.csv").write_text("Male ID,Female ID\n" + "".join(f"{i},{pi[i]}\n" for i in range(1, 51)))
    for tool, arg in (("transform-solution", "sol.csv"), ("convert-VNC-matching", "sol_matching.csv")):
        r = subprocess.run([os.path.join(bm.BIN_DIR, tool), arg], cwd=tmp_path, capture_output=True)
        assert r.returncode == 0, r.stderr
    assert (tmp_path / "sol_matching_generic.csv").read_text() == (tmp_path / "sol.csv").read_text()
    for f in ("sol.csv", "sol_matching.csv", "sol_matching_generic.csv"):
        assert np.array_equal(bm.load_mapping_csv(str(tmp_path / f), 50), pi)


def test_bench_dense_row_rule():
    """bench.py's copy of the library's dense-row rule (rows of at least a
    quarter of the stride, longest first, each taken only while every
    column's dense weights sum to at most 65,535): three female hubs of weight
    21,845 to every vertex fill the columns to 65,535 exactly, so a fourth
    hub of weight 1 is left out, and the split counts the three hubs' male
    edges and their expansions."""
    sys.path.insert(0, ROOT)
    import bench

    n = 64
    hubs = ((1, 21845), (2, 21845), (3, 21845), (4, 1))
    src = np.array([h for h, _ in hubs for v in range(1, n + 1) if v != h], np.int32)
    dst = np.array([v for h, _ in hubs for v in range(1, n + 1) if v != h], np.int32)
    fw = np.array([w for h, w in hubs for v in range(1, n + 1) if v != h], np.int32)

    class G:   # the fields dense_split reads
        pass

    female = G()
    female.src, female.dst, female.w = src, dst, fw

    class Ctx:
        def __init__(self):
            self.n = n

        def degrees(self, female=False):   # the same edges in both graphs
            return np.bincount(src, minlength=n + 1), np.bincount(dst, minlength=n + 1)

    pi = np.arange(n + 1)
    entries, pairs = bench.dense_split(Ctx(), pi, female, 3)
    assert pairs == 3 * 63 and entries == 3 * 63 * 63
    with pytest.raises(SystemExit):
        bench.dense_split(Ctx(), pi, female, 4)

"""The drop-in path end to end on the device: the search driver's
checkpointing through the C ABI and the three re-alignment tools run as
processes, as a user of the reference tools runs them
(src/graph-re-alignment-gpu-swap.cu:1469-1519, :749-763, :1114-1132;
src/best-swap/score-lookup.c:43-47 for the save on a signal).  Every saved
matching is rescored by the oracle."""
import os
import re
import signal
import subprocess
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOOLS = ("graph-re-alignment-gpu-swap", "graph-re-alignment-gpu-3-swap", "graph-re-alignment-gpu")
N = 600


@pytest.fixture(scope="module")
def pair(bm, orc, tmp_path_factory):
    male, female, truth, pi0 = bm.synth_pair(N, 10.0, bm.WEIGHTS_GEOMETRIC, 0.3, seed=41)
    d = tmp_path_factory.mktemp("tools")
    bm.write_graph_csv(str(d / "male.csv"), male)
    bm.write_graph_csv(str(d / "female.csv"), female)
    bm.write_mapping_csv(str(d / "start.csv"), pi0)
    o = orc.Oracle(N, male, female)
    return {"dir": d, "male": male, "female": female, "truth": truth, "pi0": pi0, "o": o}


def _final_score(text):
    m = re.search(r"Final score: (-?\d+)", text)
    assert m, text[-2000:]
    return int(m.group(1))


def _run(bm, tool, workdir, inp, out, seconds, env_extra=None):
    env = dict(os.environ, BM_SEED="7", **(env_extra or {}))
    return subprocess.run([os.path.join(bm.BIN_DIR, tool), "../male.csv", "../female.csv", inp, out, str(seconds)],
                          cwd=workdir, capture_output=True, text=True, env=env, timeout=90)


def test_search_checkpoint_through_abi(bm, orc, pair, tmp_path):
    """bm_search writes the best matching on every new best and rescoring the
    file gives the best score; a better matching dropped into the sync file
    is adopted."""
    o, pi0 = pair["o"], pair["pi0"]
    out = tmp_path / "best_out.csv"
    sync = tmp_path / "sync.csv"
    with bm.Context(N, pair["male"], pair["female"]) as ctx:
        ctx.set_mapping(pi0)
        best, res = ctx.search(bm.SEARCH_SWAP_UNIFORM, seed=3, max_iters=400, out_path=str(out), verify_every=1)
        assert res["verify_failures"] == 0 and res["best_score"] > res["initial_score"]
        saved = bm.load_mapping_csv(str(out), N)
        assert o.score(saved) == res["best_score"] == o.score(best)
        # a much better matching (the hidden permutation) in the sync file
        bm.write_mapping_csv(str(sync), pair["truth"])
        s_truth = o.score(pair["truth"])
        assert s_truth > res["best_score"]
        ctx.set_mapping(pi0)
        best2, res2 = ctx.search(bm.SEARCH_SWAP_UNIFORM, seed=4, time_limit_s=1.5, sync_path=str(sync),
                                 sync_interval_s=0.2, verify_every=1)
        assert res2["verify_failures"] == 0
        assert res2["best_score"] >= s_truth and o.score(best2) == res2["best_score"]


@pytest.mark.parametrize("tool", TOOLS)
def test_tool_writes_rescored_checkpoint_and_resumes(bm, orc, pair, tool, tmp_path):
    """A 4-second run writes <out minus .csv>-rank0.csv whose oracle score is
    the logged final score; a second run started from that file begins at
    that score and does not end below it."""
    o = pair["o"]
    run = tmp_path / "run"
    run.mkdir()
    for f in ("male.csv", "female.csv", "start.csv"):
        (tmp_path / f).write_bytes((pair["dir"] / f).read_bytes())
    print(f"[tools] {tool}: first run", flush=True)
    r = _run(bm, tool, run, "../start.csv", "result.csv", 4)
    assert r.returncode == 0, r.stderr[-2000:]
    ckpt = run / "result-rank0.csv"
    assert ckpt.exists()
    assert ckpt.read_text().splitlines()[0] == "Male Node ID,Female Node ID"
    s1 = _final_score(r.stdout)
    assert o.score(bm.load_mapping_csv(str(ckpt), N)) == s1
    assert s1 > o.score(pair["pi0"])
    print(f"[tools] {tool}: resume", flush=True)
    r2 = _run(bm, tool, run, "result-rank0.csv", "again.csv", 2)
    assert r2.returncode == 0, r2.stderr[-2000:]
    assert f"Initial alignment score: {s1}" in r2.stdout
    s2 = _final_score(r2.stdout)
    assert s2 >= s1 and o.score(bm.load_mapping_csv(str(run / "again-rank0.csv"), N)) == s2


def test_tool_adopts_better_best_csv(bm, orc, pair, tmp_path):
    """../data/best.csv (the reference's side channel) holding a better
    matching is adopted at the next sync."""
    o = pair["o"]
    run = tmp_path / "run"
    (tmp_path / "data").mkdir()
    run.mkdir()
    for f in ("male.csv", "female.csv", "start.csv"):
        (tmp_path / f).write_bytes((pair["dir"] / f).read_bytes())
    bm.write_mapping_csv(str(tmp_path / "data" / "best.csv"), pair["truth"])
    s_truth = o.score(pair["truth"])
    print("[tools] adopt best.csv", flush=True)
    r = _run(bm, "graph-re-alignment-gpu-swap", run, "../start.csv", "result.csv", 3, {"BM_SYNC_INTERVAL": "0.3"})
    assert r.returncode == 0, r.stderr[-2000:]
    assert f"adopted ../data/best.csv with score {s_truth}" in r.stdout
    s = _final_score(r.stdout)
    assert s >= s_truth and o.score(bm.load_mapping_csv(str(run / "result-rank0.csv"), N)) == s


def test_tool_saves_on_sigterm(bm, orc, pair, tmp_path):
    """With no time limit the tool runs until a signal and then saves the
    best matching, as the reference's handler does."""
    o = pair["o"]
    run = tmp_path / "run"
    run.mkdir()
    for f in ("male.csv", "female.csv", "start.csv"):
        (tmp_path / f).write_bytes((pair["dir"] / f).read_bytes())
    env = dict(os.environ, BM_SEED="9")
    env.pop("BM_TIME_LIMIT", None)
    print("[tools] sigterm", flush=True)
    p = subprocess.Popen([os.path.join(bm.BIN_DIR, "graph-re-alignment-gpu"), "../male.csv", "../female.csv",
                          "../start.csv", "result.csv"], cwd=run, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                         text=True, env=env)
    try:
        t0 = time.time()
        ckpt = run / "result-rank0.csv"
        while not ckpt.exists() and time.time() - t0 < 60 and p.poll() is None:
            time.sleep(0.1)
        time.sleep(1.0)
        assert p.poll() is None, "the tool ended without a signal"
        p.send_signal(signal.SIGTERM)
        out, err = p.communicate(timeout=60)
    finally:
        if p.poll() is None:
            p.kill()
            p.communicate()
    assert p.returncode == 0, err[-2000:]
    s = _final_score(out)
    assert o.score(bm.load_mapping_csv(str(run / "result-rank0.csv"), N)) == s > o.score(pair["pi0"])

"""The push autotuner (--push-kernel tune, EXC_KERNEL_TUNE; no reference counterpart: the
reference fixes the group width from the average degree, src/excent.cu:238, 629-633).

Checks: runs with the tuner -- resident push passes, and streamed chunks (its trials then
drain a staging slot's stream) -- end within the same FP64 bound, invariant and |r| <= eps as
the load-balanced kernel; at least one degree class timed its three candidates (a positive
time per edge for each) and kept the fastest; the kept widths are among the candidates.  (Bursts
push with the load-balanced kernel and run no trials: the runs here synchronise per pass.)"""
import numpy as np
import pytest

import excent_amd.native as nat
from oracle import oracle as orc
from tests.test_gpu_push import true_residual

pytestmark = pytest.mark.gpu


def check_pr(idx, adj, x, eps):
    n = len(idx) - 1
    b = np.full(n, np.float64(np.float32(1) - np.float32(0.85)))
    ref, _, _ = orc.fp64_solve_b(idx, adj, b)
    tr = true_residual(idx, adj, x, b, 0.85, True)
    assert np.abs(x.astype(np.float64) - ref).sum() <= np.abs(tr).sum() / 0.15 * (1 + 1e-6) + 1e-6
    assert np.abs(tr).max() <= eps * (1 + 1e-3) + 2e-4 * max(1.0, np.abs(x).max())


@pytest.mark.parametrize("cache", ["chunks", "none"])
def test_push_tuner(tmp_path, rmat14, cache):
    """PageRank on rmat14 by pushes only (--pull off), resident or streamed in 0.1 M-edge
    chunks: the tuner's run meets the bounds, and its decisions are consistent."""
    print("case tune", cache, flush=True)
    idx, adj = orc.egr_load(rmat14)
    eps = 1e-4
    args = [rmat14, "--measure", "pr", "--eps-pr", eps, "--pull", "off", "--push-kernel", "tune", "--cache", cache,
            "--stream-ratio", "0.01", "--chunk-medges", "0.1", "--burst", 1, "--log", str(tmp_path / "t.log")]
    if cache == "chunks":
        args.append("--preload")
    with nat.Session(args) as S:
        r = S.run(reinit=False)
        x = S.values(0, 0)
    assert r["stop_reason"] == 0
    check_pr(idx, adj, x, eps)
    k = r["kernel"]
    best, ns = k["push_tune_best"], k["push_tune_ns_edge"]
    decided = [c for c in range(7) if best[c] >= 0]
    timed = [c for c in decided if all(t > 0 for t in ns[c])]
    assert timed, (best, ns)
    for c in timed:   # the kept candidate is the fastest of the three
        assert ns[c][best[c]] == min(ns[c])
    for c in range(7):
        assert best[c] in (-1, 0, 1, 2)


def test_push_tuner_matches_lb(tmp_path, rmat12):
    """The same run with the load-balanced kernel and with the tuner: both within the bound,
    and within eps * n / (1 - d) of each other."""
    idx, adj = orc.egr_load(rmat12)
    n = len(idx) - 1
    xs = []
    for kern in ("lb", "tune"):
        with nat.Session([rmat12, "--measure", "pr", "--eps-pr", 1e-4, "--pull", "off", "--push-kernel", kern,
                          "--cache", "chunks", "--preload", "--burst", 1, "--log", str(tmp_path / f"{kern}.log")]) as S:
            r = S.run(reinit=False)
            xs.append(S.values(0, 0))
        assert r["stop_reason"] == 0
        check_pr(idx, adj, xs[-1], 1e-4)
    assert np.abs(xs[0].astype(np.float64) - xs[1]).sum() <= 2 * 1e-4 * n / 0.15
    assert nat.excent([rmat12, "--push-kernel", "fastest", "--log", str(tmp_path / "bad.log")]) == 1

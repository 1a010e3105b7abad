"""The oracle's CPU asynchronous delta push (the timed CPU baseline) against FP64
closed forms and the FP64 oracle, and the FP64 residual invariant.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.conftest import SAMPLE_EGR, SAMPLE_ENV, read_env


def run(idx, adj, kinds, eps, scale, divdeg, inits, cv0=None, threads=4):
    xs = [np.array(i[0], np.float32) for i in inits]
    rs = [np.array(i[1], np.float32) for i in inits]
    cv0 = np.array([0], np.int32) if cv0 is None else cv0
    passes, pushed, secs = orc.run_push(idx, adj, cv0, xs, rs, eps, scale, divdeg, threads=threads)
    return xs, rs, passes, pushed


@pytest.mark.parametrize("gname", ["sample", "rmat12"])
def test_cpu_push_pagerank_bound(gname, rmat12):
    idx, adj = orc.egr_load(SAMPLE_EGR if gname == "sample" else rmat12)
    cv0 = orc.plan_chunks(idx, max(1, len(adj) // 4))
    xs, rs, passes, pushed = run(idx, adj, [0], [0.01], [np.float32(0.85)], [1], [orc.init(idx, 0)], cv0)
    ref, _, _, _ = orc.fp64_reference(idx, adj, orc.ORC_PR)
    rl1 = np.abs(rs[0].astype(np.float64)).sum()
    assert np.all(np.abs(rs[0]) <= 0.01)
    assert np.abs(xs[0] - ref).sum() <= rl1 / 0.15 * 1.001 + 1e-3
    # the true FP64 residual of x equals the state's r up to float rounding
    true = orc.residual_l1(idx, adj, orc.ORC_PR, 0.85, 1.0, xs[0])
    assert abs(true - rl1) <= 1e-3 * max(1.0, rl1) + 1e-3 * len(idx)
    assert pushed >= len(adj)


def test_cpu_push_all_sample():
    env = read_env(SAMPLE_ENV)
    idx, adj = orc.egr_load(SAMPLE_EGR)
    lam = np.float32(env["LAMBDA"])
    inits = [orc.init(idx, 0), orc.init(idx, 1, beta=1.0), orc.init(idx, 2, lam=lam)]
    xs, rs, _, _ = run(idx, adj, [0, 1, 2], [0.01, env["EPS_KATZ"], env["EPS_EV"]],
                       [np.float32(0.85), np.float32(env["ALPHA"]), np.float32(1 / env["LAMBDA"])], [1, 0, 0], inits)
    # PageRank within its bound; Katz and eigenvector (eps = 1% of their means) within a
    # few per cent, with the top-100 sets nearly equal
    ref = orc.fp64_reference(idx, adj, orc.ORC_PR)[0]
    assert np.abs(xs[0] - ref).sum() <= np.abs(rs[0].astype(np.float64)).sum() / 0.15 * 1.001
    c = orc.compare(orc.fp64_reference(idx, adj, orc.ORC_KATZ, alpha=env["ALPHA"])[0], xs[1])
    assert c["l1_rel"] < 0.05 and c["top100"] >= 85
    c = orc.compare(orc.fp64_reference(idx, adj, orc.ORC_EV)[0], xs[2], unit=True)
    assert c["l1_rel"] < 0.1 and c["top100"] >= 80


def test_cpu_push_edgeless_and_cycle():
    n = 100
    idx = np.zeros(n + 1, np.uint64)
    xs, rs, passes, pushed = run(idx, np.zeros(0, np.int32), [0], [0.01], [np.float32(0.85)], [1], [orc.init(idx, 0)])
    assert np.all(xs[0] == np.float32(1) - np.float32(0.85)) and pushed == 0 and passes == 1
    idx = np.arange(n + 1, dtype=np.uint64)
    adj = ((np.arange(n) + 1) % n).astype(np.int32)
    xs, rs, _, _ = run(idx, adj, [0], [1e-5], [np.float32(0.85)], [1], [orc.init(idx, 0)])
    assert np.abs(xs[0].astype(np.float64) - 1.0).sum() <= np.abs(rs[0].astype(np.float64)).sum() / 0.15 + 1e-3


def test_cpu_push_time_limit_stops_early(rmat12):
    idx, adj = orc.egr_load(rmat12)
    xs = [np.zeros(len(idx) - 1, np.float32)]
    rs = [np.full(len(idx) - 1, 0.15, np.float32)]
    passes, pushed, secs = orc.run_push(idx, adj, np.array([0], np.int32), xs, rs, [1e-9], [np.float32(0.85)], [1],
                                        threads=2, time_limit=1e-9)
    assert passes == 1 and pushed > 0


def test_bf16_round_known_answers():
    """Round to nearest, ties to even, at 7 fraction bits (step 2^-7 at 1.0)."""
    a = np.array([1.0, 1 + 2**-8, 1 + 3 * 2**-9, 1 + 2**-7 + 2**-8, -2.5, 0.0], np.float32)
    want = np.array([1.0, 1.0, 1 + 2**-7, 1 + 2**-6, -2.5, 0.0], np.float32)
    assert np.array_equal(orc.bf16_round(a), want)
    # every output has a zero low half and is the closer of its two bf16 neighbours
    v = np.random.default_rng(2).standard_normal(10000).astype(np.float32) * 1e3
    b = orc.bf16_round(v)
    assert not np.any(b.view(np.uint32) & 0xFFFF)
    down = (v.view(np.uint32) & 0xFFFF0000).view(np.float32)
    up = ((v.view(np.uint32) & 0xFFFF0000) + 0x10000).view(np.float32)
    assert np.all(np.abs(b.astype(np.float64) - v) <= np.minimum(np.abs(down - v.astype(np.float64)),
                                                                 np.abs(up - v.astype(np.float64))))


def star_graph(L=9000):
    deg = np.concatenate([[L], np.ones(L, np.int64)])
    idx = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    adj = np.concatenate([np.arange(1, L + 1), np.zeros(L, np.int64)]).astype(np.int32)
    return idx, adj


def invariant_excess(x, r, true, eps=1e-4):
    """Largest amount by which |r - (b + W x - x)| exceeds 1e-4 |x| + eps at a vertex (the GPU
    test's tolerance): float rounding stays under it, a lost rounding remainder does not."""
    return float((np.abs(r - true) - (1e-4 * np.abs(x) + eps)).max())


def _ppr_dense_run(idx, adj, rows, feedback, eps=1e-4, S=3):
    n = len(idx) - 1
    seeds = np.random.default_rng(21).choice(n, S, replace=False)
    b = np.zeros(n)
    b[seeds] = np.float32(0.15 * n / S)
    x, r = np.zeros(n, np.float32), b.astype(np.float32)
    for _ in range(1000):
        if orc.dense_pass_rows(idx, adj, x, r, eps, rows=rows, feedback=feedback) == 0:
            break
    deg = np.diff(idx.astype(np.int64))
    src = np.repeat(np.arange(n), deg)
    acc = np.zeros(n)
    np.add.at(acc, adj, float(np.float32(0.85)) * x.astype(np.float64)[src] / deg[src])
    return x, r, b + acc - x.astype(np.float64), b


@pytest.mark.parametrize("gname", ["rmat12", "star"])
@pytest.mark.parametrize("rows", ["f32", "bf16"])
def test_dense_rows_keep_the_residual_invariant(rows, gname, rmat12):
    """Dense passes with bf16 weight rows keep r = b + W x - x as float rows do (the fold
    leaves the rounding remainder in r), so the FP64 bound holds at termination."""
    idx, adj = orc.egr_load(rmat12) if gname == "rmat12" else star_graph()
    x, r, true, b = _ppr_dense_run(idx, adj, rows, True)
    assert np.all(np.abs(r) <= 1e-4)
    assert invariant_excess(x, r, true) <= 0
    ref, _, _ = orc.fp64_solve_b(idx, adj, b)
    assert np.abs(x - ref).sum() <= np.abs(true).sum() / 0.15 * (1 + 1e-6) + 1e-6


@pytest.mark.parametrize("gname", ["rmat12", "star"])
def test_dense_rows_without_feedback_break_the_invariant(gname, rmat12):
    """Planted mistake: bf16 weights with all of d moved into x.  The invariant check at the
    GPU test's tolerance fails, and the true residual exceeds eps."""
    idx, adj = orc.egr_load(rmat12) if gname == "rmat12" else star_graph()
    x, r, true, _ = _ppr_dense_run(idx, adj, "bf16", False)
    assert np.all(np.abs(r) <= 1e-4)
    assert invariant_excess(x, r, true) > 0
    assert np.abs(true).max() > 1e-4

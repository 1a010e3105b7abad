"""FP64 global residual sums of the snapshot (exc_residual_l1; no reference counterpart:
the reference stops on the per-vertex test alone, src/excent.cu:680), the global-L1 stop
built on them (--stop global), and the state a run leaves at its pass limit.

Checks: the GPU's FP64 sum of |r| equals numpy's FP64 sum of the float32 r read back (the
same numbers summed in another order: relative difference ~1e-13), from k_snapshot and from
the snapshot fused into the dense pull; it equals the oracle's FP64 residual of x within the
float rounding of the invariant x + r = b + W x; a --stop global run ends on the first pass
whose sums are all below n * eps; a run cut by --max-passes right after dense passes keeps
x + r = b + W x (no fold of a pass that will not run)."""
import os
import re

import numpy as np
import pytest

import excent_amd.native as nat
from oracle import oracle as orc
from tests.conftest import read_env

pytestmark = pytest.mark.gpu


def _loop(ctx, n, m, check, max_passes=3000):
    for p in range(max_passes):
        cnt, _ = ctx.snapshot()
        check(p)
        if cnt[0::2].sum() == 0:
            return p
        ctx.push_resident(0, n, 0, m, False)
        ctx.pass_sync()
    raise AssertionError("did not converge")


@pytest.mark.parametrize("pull", [None, "fused"])
def test_residual_sum_matches_state(rmat12, pull, monkeypatch):
    """Three measures on rmat12, pass by pass: exc_residual_l1 against the FP64 sum of the
    r read back, with push passes (k_snapshot) and with dense passes whose next snapshot is
    fused into k_pull (EXC_FUSE_FOLD=0 keeps the state readable between passes)."""
    monkeypatch.setenv("EXC_FUSE_FOLD", "0")
    idx, adj = orc.egr_load(rmat12)
    n, m = len(idx) - 1, len(adj)
    lam = float(np.float32(orc.fp64_reference(idx, adj, orc.ORC_EV, tol=1e-10, max_iter=2000)[3]))
    assert lam > 1
    alpha = 0.5 / lam
    kinds = [nat.EXC_PR, nat.EXC_KATZ, nat.EXC_EV]
    eps = [1e-3, 1e-5, 1e-5]
    scale = [np.float32(0.85), np.float32(alpha), np.float32(1.0 / lam)]
    with nat.Context(n, m, np.array([0], np.int32), kinds, eps, scale, [1, 0, 0]) as ctx:
        ctx.upload_offsets(idx)
        ctx.init_measure(0)
        ctx.init_measure(1, beta=1.0)
        # eigenvector start x = 1, r = indeg / lambda - 1: the invariant x + r = W x on a
        # directed graph
        indeg = np.bincount(adj, minlength=n).astype(np.float64)
        ctx.set_state(2, x=np.ones(n, np.float32), r=(indeg / lam - 1.0).astype(np.float32))
        assert ctx.cache_chunks(m) == 0
        ctx.cache_upload(adj)
        if pull:
            assert ctx.enable_pull(0.0) == 0
        seen = []

        def check(p):
            got = ctx.residual_l1()
            rs = [ctx.read_state(i, 1) for i in range(3)]
            want = [np.abs(r.astype(np.float64)).sum() for r in rs]
            for g, w in zip(got, want):
                assert abs(g - w) <= 1e-12 * max(w, 1e-300) + 1e-300, (p, got, want)
            # the largest |r| per measure, exactly
            assert ctx.residual_max() == [float(np.abs(r).max()) for r in rs], p
            x = [ctx.read_state(i, 0) for i in range(3)]
            # the residual of x, recomputed from the graph in FP64 by the oracle
            true = [orc.residual_l1(idx, adj, orc.ORC_PR, float(scale[0]), 0.15, x[0]),
                    orc.residual_l1(idx, adj, orc.ORC_KATZ, float(scale[1]), 1.0, x[1]),
                    orc.residual_l1(idx, adj, orc.ORC_EV, float(scale[2]), 0.0, x[2])]
            for i in range(3):
                tol = 256 * 2.0 ** -24 * np.abs(x[i].astype(np.float64)).sum() + 1e-6 * n
                assert abs(got[i] - true[i]) <= tol, (p, i, got[i], true[i], tol)
            seen.append(got)

        _loop(ctx, n, m, check)
        if pull:
            assert ctx.stats()["pull_passes"] > 0
    # PageRank's residual mass falls pass by pass (its push is a contraction)
    pr = [s[0] for s in seen]
    assert pr[-1] < pr[0] and pr[0] == pytest.approx(0.15 * n, rel=1e-6)


def test_residual_sum_of_folded_state(rmat12):
    """With the next fold fused into the dense pull (the default), the sum is the one of the
    state the snapshot tested, before the fold moved r into x: at least the sum of the r read
    back, and equal to it once nothing was folded."""
    idx, adj = orc.egr_load(rmat12)
    n, m = len(idx) - 1, len(adj)
    with nat.Context(n, m, np.array([0], np.int32), [nat.EXC_PR], [1e-4], [0.85], [1]) as ctx:
        ctx.upload_offsets(idx)
        ctx.init_measure(0)
        assert ctx.cache_chunks(m) == 0
        ctx.cache_upload(adj)
        assert ctx.enable_pull(0.0) == 0
        folded = 0

        def check(p):
            nonlocal folded
            got = ctx.residual_l1()[0]
            back = np.abs(ctx.read_state(0, 1).astype(np.float64)).sum()
            if ctx.dense_required():
                folded += 1
                assert got >= back
            else:
                assert abs(got - back) <= 1e-12 * max(back, 1e-300) + 1e-300, (p, got, back)

        _loop(ctx, n, m, check)
        assert folded > 0


@pytest.fixture(scope="module")
def rmat14_env(rmat14, tmp_path_factory):
    out = str(tmp_path_factory.mktemp("env") / "rmat14.env")
    assert nat.run_entry("excent_params_main", [rmat14, out, "--directed", "--alpha-factor", "0.5", "--seed", "7"]) == 0
    return read_env(out)


PASS = re.compile(r"pass +(\d+): .* \| residual l1 pr ([0-9.e+-]+) katz ([0-9.e+-]+) ev ([0-9.e+-]+)")


@pytest.mark.parametrize("cache", ["chunks", "none"])
def test_global_stop(tmp_path, rmat14, rmat14_env, cache):
    """--stop global with --measure all: the loop ends at the first snapshot whose FP64 sums
    are all <= n * eps (every earlier pass line shows one above it), the report gives that
    sum and the rule, the FP64 residual of the dumped x is within float rounding of it, and
    the per-vertex rule alone needs more passes."""
    env = rmat14_env
    idx, adj = orc.egr_load(rmat14)
    n = len(idx) - 1
    eps = 1e-6
    base = [rmat14, "--measure", "all", "--alpha", env["ALPHA"], "--beta", 1, "--lambda", env["LAMBDA"],
            "--eps-pr", 0.01, "--eps-katz", eps, "--eps-ev", eps, "--ev-init", "indeg", "--cache", cache,
            "--chunk-medges", "0.05", "--stream-ratio", "0.3", "--max-passes", "3000", "--dump", str(tmp_path / "o")]
    print("case global stop", cache, flush=True)
    log = str(tmp_path / "g.log")
    assert nat.excent(base + ["--stop", "global", "--log", log]) == 0, open(log).read()
    text = open(log).read()
    bounds = np.array([n * float(np.float32(0.01)), n * float(np.float32(eps)), n * float(np.float32(eps))])
    lines = [(int(mo.group(1)), np.array([float(mo.group(i)) for i in (2, 3, 4)])) for mo in PASS.finditer(text)]
    # with the dense passes on (cache chunks), a state that meets the bounds but was folded by
    # the pass before it still has deltas in flight: the loop runs one more pass (folding
    # nothing) and tests again, so the last pass line may already be below the bounds
    last = lines[:-1] if cache == "chunks" else lines
    assert lines and all(np.any(l1 > bounds) for _, l1 in last), text
    mo = re.search(r"residual l1 \(fp64\):  pr ([0-9.e+-]+) \(n\*eps [0-9.e+-]+\)  katz ([0-9.e+-]+) \(n\*eps "
                   r"[0-9.e+-]+\)  ev ([0-9.e+-]+) \(n\*eps [0-9.e+-]+\) \| stop: global l1 <= n\*eps", text)
    assert mo, text
    final = np.array([float(mo.group(i)) for i in (1, 2, 3)])
    assert np.all(final <= bounds)
    passes = int(re.search(r"passes (\d+)  measure all", text).group(1))
    assert passes == len(lines)
    x = {k: np.fromfile(str(tmp_path / f"o.{k}.bin"), np.float32) for k in ("pr", "katz", "ev")}
    true = [orc.residual_l1(idx, adj, orc.ORC_PR, float(np.float32(0.85)), 0.15, x["pr"]),
            orc.residual_l1(idx, adj, orc.ORC_KATZ, float(np.float32(env["ALPHA"])), 1.0, x["katz"]),
            orc.residual_l1(idx, adj, orc.ORC_EV, float(np.float32(1.0 / env["LAMBDA"])), 0.0, x["ev"])]
    for i, k in enumerate(("pr", "katz", "ev")):
        tol = 256 * 2.0 ** -24 * np.abs(x[k].astype(np.float64)).sum()
        assert abs(true[i] - final[i]) <= tol + 1e-9 * n, (k, true[i], final[i], tol)
    # the reference's rule on the same run needs at least as many passes
    log2 = str(tmp_path / "v.log")
    assert nat.excent(base + ["--log", log2]) == 0
    passes_v = int(re.search(r"passes (\d+)  measure all", open(log2).read()).group(1))
    assert passes_v >= passes
    assert "stop: nothing pending" in open(log2).read()


def test_stop_rule_per_measure(tmp_path, rmat14, rmat14_env):
    """--stop katz,ev: PageRank keeps the reference's per-vertex rule while Katz and the
    eigenvector may end on their global sum; at the end every PageRank |r| <= eps, each of
    the others either has nothing pending or its sum within n * eps, and the report's
    largest |r| per measure is the dumped state's."""
    env = rmat14_env
    idx, adj = orc.egr_load(rmat14)
    n = len(idx) - 1
    log = str(tmp_path / "s.log")
    args = [rmat14, "--measure", "all", "--alpha", env["ALPHA"], "--beta", 1, "--lambda", env["LAMBDA"],
            "--eps-pr", 0.001, "--eps-katz", 1e-6, "--eps-ev", 1e-6, "--ev-init", "indeg", "--cache", "chunks",
            "--preload", "--stop", "katz,ev", "--max-passes", 3000, "--log", log]
    with nat.Session(args) as S:
        r = S.run(reinit=False)
        R = [S.values(i, 1) for i in range(3)]
    assert r["stop_reason"] in (0, 1)
    assert np.abs(R[0]).max() <= np.float32(0.001)
    for i in (1, 2):
        assert np.abs(R[i]).max() <= np.float32(1e-6) or np.abs(R[i].astype(np.float64)).sum() <= n * float(np.float32(1e-6))
        assert r["residual_l1"][i] == pytest.approx(np.abs(R[i].astype(np.float64)).sum(), rel=1e-12)
    assert nat.excent([rmat14, "--measure", "pr", "--stop", "katz", "--log", str(tmp_path / "bad.log")]) == 1


def test_pass_limit_keeps_invariant(rmat12):
    """A run cut by --max-passes while its dense passes fold the pass after them: the last
    pass folds nothing more, so the state returned keeps x + r = b + W x per vertex (within
    float rounding), and the report names the pass limit."""
    idx, adj = orc.egr_load(rmat12)
    n = len(idx) - 1
    for cap in (2, 3, 5):
        with nat.Session([rmat12, "--measure", "pr", "--eps-pr", 1e-6, "--cache", "chunks", "--preload",
                          "--max-passes", cap]) as S:
            st = S.run(reinit=False)
            x, r = S.values(0, 0), S.values(0, 1)
        assert st["passes"] == cap and st["stop_reason"] == 2
        assert st["kernel"]["pull_passes"] > 0
        deg = np.diff(idx.astype(np.int64))
        src = np.repeat(np.arange(n), deg)
        acc = np.zeros(n)
        np.add.at(acc, adj, float(np.float32(0.85)) * x.astype(np.float64)[src] / deg[src])
        true = 0.15 + acc - x.astype(np.float64)
        assert np.abs(true - r).max() <= 2e-4 * max(1.0, np.abs(x).max()), cap
        assert st["residual_l1"][0] == pytest.approx(np.abs(r.astype(np.float64)).sum(), rel=1e-12)

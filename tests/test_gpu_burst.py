"""Bursts of passes with one host synchronisation (exc_run_passes; no reference counterpart:
the reference waits for the host at every pass, src/excent.cu:664-665, 750).

Checks: on graphs where every destination receives at most one delta per pass (a directed
cycle for dense passes, a perfect matching for pushes) a pass's arithmetic does not depend on
atomic order, so a burst must reproduce pass-by-pass runs bit for bit -- every pass's snapshot
counts and FP64 sums, and the final x and r; a burst that starts from a converged state
changes nothing; a whole R-MAT run by bursts meets the FP64 bound and the invariant, and its
last snapshot equals the oracle's snapshot of the state read back; the refusals; and
bin/excent's --burst 8 against --burst 1."""
import re

import numpy as np
import pytest

import excent_amd.native as nat
from oracle import oracle as orc
from tests.test_gpu_push import true_residual

pytestmark = pytest.mark.gpu


def context(idx, adj, eps, pull, init=None):
    n, m = len(idx) - 1, len(adj)
    ctx = nat.Context(n, m, np.array([0], np.int32), [nat.EXC_PR], [eps], [0.85], [1])
    ctx.upload_offsets(idx)
    x0, r0 = orc.init(idx, orc.ORC_PR) if init is None else init
    ctx.set_state(0, x=x0, r=r0)
    assert ctx.cache_chunks(m) == 0
    if m:
        ctx.cache_upload(adj, 0)
    if pull is not None:
        assert ctx.enable_pull(pull) == 0
    return ctx


def single_passes(idx, adj, eps, pull, max_passes=500, init=None):
    """Pass by pass: the counts and FP64 sums of every snapshot, the final state."""
    n, m = len(idx) - 1, len(adj)
    snaps = []
    with context(idx, adj, eps, pull, init) as ctx:
        for _ in range(max_passes):
            cnt, _ = ctx.snapshot()
            snaps.append((cnt.copy(), ctx.residual_l1()[0], ctx.residual_max()[0]))
            if cnt[0::2].sum() == 0:
                break
            ctx.push_resident(0, n, 0, m, True)
            ctx.pass_sync()
        return snaps, ctx.read_state(0, 0), ctx.read_state(0, 1)


def burst_passes(idx, adj, eps, pull, B, dense, max_bursts=100, init=None, ahead=False):
    """The same run by bursts of B passes, ended at the first snapshot with nothing pending.
    ahead: the next burst is submitted before the current one is collected (two in flight),
    and the one still in flight when the run ends is drained."""
    snaps = []
    with context(idx, adj, eps, pull, init) as ctx:
        cnt, _ = ctx.snapshot()
        snaps.append((cnt.copy(), ctx.residual_l1()[0], ctx.residual_max()[0]))
        inflight = 0
        for _ in range(max_bursts):
            if snaps[-1][0][0::2].sum() == 0:
                break
            d = dense if isinstance(dense, bool) else dense(ctx)
            if ahead:
                while inflight < 2:
                    ctx.burst_submit(B, d)
                    inflight += 1
                cnts, l1, mx = ctx.burst_collect(B)
                inflight -= 1
                assert len(cnts) == B
            else:
                cnts, l1, mx = ctx.run_passes(B, d)
            for i in range(B):
                snaps.append((cnts[i].copy(), float(l1[i][0]), float(mx[i][0])))
                if cnts[i][0::2].sum() == 0:
                    break
        while inflight:   # submitted after the run's end: it starts from nothing pending
            cnts, _, _ = ctx.burst_collect(B)
            inflight -= 1
            assert cnts.sum() == 0
        # the next snapshot is the burst's last one, without a launch
        cnt, _ = ctx.snapshot()
        assert np.array_equal(cnt, cnts[B - 1])
        burst_passes.graph_launches = ctx.stats()["burst_graph_launches"]
        return snaps, ctx.read_state(0, 0), ctx.read_state(0, 1)


@pytest.mark.parametrize("ahead", [False, True])
@pytest.mark.parametrize("B", [2, 8, 16])
def test_burst_dense_cycle_bit_exact(B, ahead):
    """Dense passes on a directed cycle (one in-edge per destination): bursts of B passes give
    every pass's snapshot and the final state bit for bit as pass-by-pass runs, also with the
    next burst submitted before the current one is collected (ahead)."""
    print("case burst dense cycle", B, ahead, flush=True)
    n = 5000
    idx = np.arange(n + 1, dtype=np.uint64)
    adj = ((np.arange(n) + 1) % n).astype(np.int32)
    a = single_passes(idx, adj, 1e-4, 0.0)
    b = burst_passes(idx, adj, 1e-4, 0.0, B, True, ahead=ahead)
    assert len(a[0]) == len(b[0]) > 10
    for p, (sa, sb) in enumerate(zip(a[0], b[0])):
        assert np.array_equal(sa[0], sb[0]) and sa[1] == sb[1] and sa[2] == sb[2], p
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


@pytest.mark.parametrize("graph", ["1", "0", "1-ahead"])
@pytest.mark.parametrize("B", [2, 8])
def test_burst_push_paths_bit_exact(B, graph, monkeypatch):
    """Bitmap-driven pushes along 300 disjoint paths of 20 edges with the residual mass at the
    heads (r = 1 there, 0 elsewhere): each pass one vertex per path pushes into a vertex that
    was not pending, so every destination receives one delta and the arithmetic does not
    depend on atomic order; the bursts equal pass-by-pass runs bit for bit (21 passes).
    graph 1: push bursts replayed from recorded HIP graphs (every replay checked against the
    pass-by-pass run), 0: launched directly; 1-ahead: graphs, two bursts in flight (each slot
    of count regions has its own recorded graph)."""
    print("case burst push paths", B, graph, flush=True)
    graph, _, ahead = graph.partition("-")
    monkeypatch.setenv("EXC_BURST_GRAPH", graph)
    paths, L = 300, 20
    n = paths * (L + 1)
    v = np.arange(n)
    deg = np.where(v % (L + 1) < L, 1, 0).astype(np.uint64)
    idx = np.r_[0, np.cumsum(deg)].astype(np.uint64)
    adj = (v[deg > 0] + 1).astype(np.int32)
    r0 = np.where(v % (L + 1) == 0, 1.0, 0.0).astype(np.float32)
    init = (np.zeros(n, np.float32), r0)
    a = single_passes(idx, adj, 1e-3, None, init=init)
    b = burst_passes(idx, adj, 1e-3, None, B, False, init=init, ahead=ahead == "ahead")
    assert (burst_passes.graph_launches > 0) == (graph == "1")
    assert len(a[0]) == len(b[0]) == L + 2
    for p, (sa, sb) in enumerate(zip(a[0], b[0])):
        assert np.array_equal(sa[0], sb[0]) and sa[1] == sb[1] and sa[2] == sb[2], p
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def test_burst_rmat_run(rmat12):
    """A whole run by bursts of 8 (dense while the demand is at least 30 % of the edges, then
    pushes): FP64 bound, invariant, |r| <= eps, and its last snapshot equal to the oracle's
    snapshot of the state read back; a further burst from the converged state changes
    nothing."""
    idx, adj = orc.egr_load(rmat12)
    n, m = len(idx) - 1, len(adj)
    eps = 1e-4

    def kind(ctx):   # dense while the demand is at least 30 % of the edges (or the last pass folded)
        return ctx.dense_required() == 1 or last[0] >= 0.3 * m

    cv0 = np.array([0], np.int32)
    with context(idx, adj, eps, 0.3) as ctx:
        cnt, _ = ctx.snapshot()
        last = [int(cnt[1::2].sum())]
        checked = 0
        kinds = []
        for _ in range(100):
            if cnt[0::2].sum() == 0:
                break
            dense = kind(ctx)
            kinds.append(dense)
            cnts, _, _ = ctx.run_passes(8, dense)
            stop = [i for i in range(8) if cnts[i][0::2].sum() == 0]
            cnt = cnts[stop[0]] if stop else cnts[7]
            last[0] = int(cnt[1::2].sum())
            if not dense and not stop:   # the burst's last snapshot is of the state read back
                _, ocnt, _ = orc.snapshot([ctx.read_state(0, 1)], np.float32([eps]), idx, cv0)
                assert np.array_equal(cnts[7], ocnt)
                checked += 1
        else:
            raise AssertionError("did not converge")
        x, r = ctx.read_state(0, 0), ctx.read_state(0, 1)
        bm, ocnt, _ = orc.snapshot([r], np.float32([eps]), idx, cv0)
        assert np.array_equal(ctx.bitmap(), bm) and np.array_equal(ctx.snapshot()[0], ocnt)
        st = ctx.stats()
        assert st["pull_passes"] > 0 and (st["src_launches"][1] > 0) == (kinds[-1] is False or False in kinds), kinds
        # from the converged state: nothing moves
        if ctx.dense_required() != 1:
            cnts, _, _ = ctx.run_passes(4, False)
            assert cnts.sum() == 0
            assert np.array_equal(ctx.read_state(0, 0), x) and np.array_equal(ctx.read_state(0, 1), r)
        cnts, _, _ = ctx.run_passes(4, True)
        assert cnts.sum() == 0
        assert np.array_equal(ctx.read_state(0, 0), x) and np.array_equal(ctx.read_state(0, 1), r)
    b = np.full(n, np.float64(np.float32(1) - np.float32(0.85)))
    ref, _, _ = orc.fp64_solve_b(idx, adj, b)
    tr = true_residual(idx, adj, x, b, 0.85, True)
    assert np.abs(x.astype(np.float64) - ref).sum() <= np.abs(tr).sum() / 0.15 * (1 + 1e-6) + 1e-6
    assert np.abs(tr - r).max() < 2e-4 * max(1.0, np.abs(x).max())
    assert np.all(np.abs(r) <= np.float32(eps))


def test_burst_refusals(rmat12):
    idx, adj = orc.egr_load(rmat12)
    n, m = len(idx) - 1, len(adj)
    with context(idx, adj, 1e-4, None) as ctx:
        for bad in (0, 17):
            with pytest.raises(nat.ExcError, match="passes"):
                ctx.run_passes(bad, False)
        with pytest.raises(nat.ExcError, match="exc_enable_pull"):
            ctx.run_passes(2, True)
        assert ctx.enable_pull(0.0) == 0
        ctx.snapshot()
        ctx.push_resident(0, n, 0, m, False)   # a full-demand dense pass folds the next one
        ctx.pass_sync()
        if ctx.dense_required():
            with pytest.raises(nat.ExcError, match="folded"):
                ctx.run_passes(2, False)
    with context(idx, adj, 1e-4, None) as ctx:   # two bursts in flight at most; nothing else meanwhile
        ctx.snapshot()
        with pytest.raises(nat.ExcError, match="no burst in flight"):
            ctx.burst_collect(2)
        ctx.burst_submit(2, False)
        ctx.burst_submit(2, False)
        with pytest.raises(nat.ExcError, match="two bursts"):
            ctx.burst_submit(2, False)
        for call in (ctx.snapshot, lambda: ctx.run_passes(2, False), lambda: ctx.read_state(0, 1),
                     lambda: ctx.push_resident(0, n, 0, m, False)):
            with pytest.raises(nat.ExcError, match="not collected"):
                call()
        ctx.burst_collect(2)
        ctx.burst_collect(2)
        ctx.snapshot()
    with nat.Context(n, m, np.array([0], np.int32), [nat.EXC_PR], [1e-4], [0.85], [1]) as ctx:
        ctx.upload_offsets(idx)
        with pytest.raises(nat.ExcError, match="edge cache"):
            ctx.run_passes(2, False)
    with nat.Context(n, m, np.array([0], np.int32), [nat.EXC_PR], [1e-4], [0.85], [1], batch_k=4) as ctx:
        with pytest.raises(nat.ExcError, match="plain"):
            ctx.run_passes(2, False)


PASS = re.compile(r"^pass +(\d+): pending (\d+) v", re.M)


def test_cli_burst(tmp_path, rmat14):
    """bin/excent PageRank with --burst 8 (the next burst submitted ahead, and not) and
    --burst 1: each ends with nothing pending within the FP64 bound; the pass lines are
    numbered 0 .. passes-1 and the last one's pending count is nonzero (the run ended at the
    first snapshot with nothing pending)."""
    idx, adj = orc.egr_load(rmat14)
    n = len(idx) - 1
    b = np.full(n, np.float64(np.float32(1) - np.float32(0.85)))
    ref, _, _ = orc.fp64_solve_b(idx, adj, b)
    for burst, ahead in ((8, "on"), (8, "off"), (1, "on")):
        print("case cli burst", burst, ahead, flush=True)
        log, pre = str(tmp_path / f"b{burst}{ahead}.log"), str(tmp_path / f"b{burst}{ahead}")
        assert nat.excent([rmat14, "--measure", "pr", "--eps-pr", 1e-4, "--cache", "chunks", "--preload",
                           "--burst", burst, "--burst-ahead", ahead, "--dump", pre, "--log", log]) == 0, open(log).read()
        text = open(log).read()
        passes = int(re.search(r"passes (\d+)  measure pr", text).group(1))
        lines = [(int(a), int(p)) for a, p in PASS.findall(text)]
        assert [a for a, _ in lines] == list(range(passes)), text
        assert lines[-1][1] > 0
        assert "stop: nothing pending" in text
        x = np.fromfile(pre + ".pr.bin", np.float32)
        tr = true_residual(idx, adj, x, b, 0.85, True)
        assert np.abs(x.astype(np.float64) - ref).sum() <= np.abs(tr).sum() / 0.15 * (1 + 1e-6) + 1e-6
        assert np.abs(tr).max() <= 1e-4 * (1 + 1e-3) + 2e-4 * max(1.0, np.abs(x).max())
    assert nat.excent([rmat14, "--burst", "17", "--log", str(tmp_path / "bad.log")]) == 1
    assert nat.excent([rmat14, "--burst-ahead", "maybe", "--log", str(tmp_path / "bad2.log")]) == 1


def test_burst_graphs_follow_the_hub_table(rmat12):
    """Push-only bursts replayed from recorded graphs on rmat12, with the hub-destination table
    replaced between bursts (the recorded launches carry its size, so the graphs must be
    recorded again): the run meets the FP64 bound, the invariant and |r| <= eps, and graphs
    were replayed before and after the change."""
    idx, adj = orc.egr_load(rmat12)
    n, m = len(idx) - 1, len(adj)
    indeg = np.bincount(adj, minlength=n)
    eps = 1e-4
    with context(idx, adj, eps, None) as ctx:
        ctx.set_hub_destinations(np.argsort(-indeg)[:16].astype(np.int32))
        cnt, _ = ctx.snapshot()
        for k in range(200):
            if cnt[0::2].sum() == 0:
                break
            if k == 2:
                g0 = ctx.stats()["burst_graph_launches"]
                assert g0 > 0
                ctx.set_hub_destinations(np.argsort(-indeg)[:300].astype(np.int32))
            cnts, _, _ = ctx.run_passes(4, False)
            stop = [i for i in range(4) if cnts[i][0::2].sum() == 0]
            cnt = cnts[stop[0]] if stop else cnts[3]
        else:
            raise AssertionError("did not converge")
        assert k > 3 and ctx.stats()["burst_graph_launches"] > g0
        x, r = ctx.read_state(0, 0), ctx.read_state(0, 1)
    b = np.full(n, np.float64(np.float32(1) - np.float32(0.85)))
    ref, _, _ = orc.fp64_solve_b(idx, adj, b)
    tr = true_residual(idx, adj, x, b, 0.85, True)
    assert np.abs(x.astype(np.float64) - ref).sum() <= np.abs(tr).sum() / 0.15 * (1 + 1e-6) + 1e-6
    assert np.abs(tr - r).max() < 2e-4 * max(1.0, np.abs(x).max())
    assert np.all(np.abs(r) <= np.float32(eps))

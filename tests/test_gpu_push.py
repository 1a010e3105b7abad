"""Delta push on the GPU (push_vertex_g and its launch variants, src/excent.cu:100-220)
driven pass by pass through the C ABI, against the FP64 oracle and invariants.

Push results depend on atomic order, so they are checked by:
  * the FP64 invariant x + r = b + W x (the GPU's r equals the true residual of its
    x up to float rounding), after the run;
  * the PageRank error bound ||x - x*||_1 <= ||r||_1 / (1 - d) against the FP64 oracle;
  * closed forms: edgeless graph (x = b), directed cycle (PageRank x = 1), a star whose
    hub has more out-edges than the hub-split threshold (exercises hub pieces).
"""
import numpy as np
import pytest

import excent_amd.native as nat
from oracle import oracle as orc
from tests.conftest import SAMPLE_EGR

pytestmark = pytest.mark.gpu


def run_passes(idx, adj, kinds, eps, scale, divdeg, inits, cv0=None, kernel=nat.KERNEL_LB, group=0,
               list_frac=0.5, max_passes=5000, pull=None):
    """Resident pass loop: snapshot, per chunk ids-or-range push from the edge cache."""
    n = len(idx) - 1
    m = len(adj)
    cv0 = np.array([0], np.int32) if cv0 is None else cv0
    bounds = np.append(cv0, n)
    nm = len(kinds)
    with nat.Context(n, m, cv0, kinds, eps, scale, divdeg, push_kernel=kernel) as ctx:
        ctx.upload_offsets(idx)
        for i, (x0, r0) in enumerate(inits):
            ctx.set_state(i, x=x0, r=r0)
        assert ctx.cache_chunks(m) == 0   # an edgeless graph still gets a (one-element) cache
        if m:
            ctx.cache_upload(adj, 0)
        if pull is not None:
            assert ctx.enable_pull(pull) == 0
            cv0 = np.array([0], np.int32)   # one run over all vertices so the pull path is taken
            bounds = np.append(cv0, n)
        passes = 0
        for passes in range(max_passes):
            cnt, _ = ctx.snapshot()
            if cnt[0::2].sum() == 0:
                break
            for c in range(len(cv0)):
                k = int(cnt[2 * c])
                if not k:
                    continue
                v0, v1 = int(bounds[c]), int(bounds[c + 1])
                ul = k < list_frac * (v1 - v0)
                ctx.push_resident(v0, v1, int(idx[v0]), int(idx[v1]), ul, group)
            ctx.pass_sync()
        else:
            raise AssertionError("did not converge")
        xs = [ctx.read_state(i, 0) for i in range(nm)]
        rs = [ctx.read_state(i, 1) for i in range(nm)]
        pushed = ctx.pushed()
        if pull is not None:
            assert ctx.stats()["pull_passes"] > 0
    return xs, rs, passes, pushed


def true_residual(idx, adj, x, b, w, divdeg):
    """b + W x - x in FP64 (W pushes w[/outdeg] along out-edges)."""
    n = len(idx) - 1
    deg = np.diff(idx.astype(np.int64))
    src = np.repeat(np.arange(n), deg)
    coef = np.where(deg > 0, w / np.maximum(deg, 1), 0.0) if divdeg else np.full(n, w)
    acc = np.zeros(n)
    np.add.at(acc, adj, coef[src] * x.astype(np.float64)[src])
    return b + acc - x.astype(np.float64)


@pytest.mark.parametrize("kernel,group", [(nat.KERNEL_LB, 0), (nat.KERNEL_GROUP, 1), (nat.KERNEL_GROUP, 4),
                                          (nat.KERNEL_GROUP, 16), (nat.KERNEL_GROUP, 64)])
@pytest.mark.parametrize("gname", ["sample", "rmat12"])
def test_pagerank_against_fp64(kernel, group, gname, rmat12):
    idx, adj = orc.egr_load(SAMPLE_EGR if gname == "sample" else rmat12)
    n = len(idx) - 1
    print(f"case {gname} n={n} kernel={kernel} group={group}", flush=True)
    x0, r0 = orc.init(idx, orc.ORC_PR)
    cv0 = orc.plan_chunks(idx, max(1, len(adj) // 5))
    xs, rs, passes, pushed = run_passes(idx, adj, [nat.EXC_PR], [0.01], [0.85], [1], [(x0, r0)], cv0, kernel, group)
    ref, _, _, _ = orc.fp64_reference(idx, adj, orc.ORC_PR)
    x = xs[0].astype(np.float64)
    rl1 = float(np.abs(rs[0].astype(np.float64)).sum())
    assert np.all(np.abs(rs[0]) <= 0.01)
    err = np.abs(x - ref).sum()
    true_l1 = np.abs(true_residual(idx, adj, xs[0], 0.15, 0.85, True)).sum()
    assert err <= true_l1 / 0.15 * (1 + 1e-6) + 1e-6, (err, true_l1)
    assert err <= 0.01 * n / 0.15
    # invariant: the GPU residual is the true residual of its x
    tr = true_residual(idx, adj, xs[0], 0.15, 0.85, True)
    assert np.abs(tr - rs[0]).max() < 2e-4 * max(1.0, np.abs(x).max())
    assert pushed >= len(adj) * 0.5


def test_all_measures_sample():
    """--measure all fused push (3 measures per adjacency read) with the sample's constants."""
    from tests.conftest import SAMPLE_ENV, read_env
    env = read_env(SAMPLE_ENV)
    idx, adj = orc.egr_load(SAMPLE_EGR)
    lam, alpha = env["LAMBDA"], env["ALPHA"]
    eps = [0.01, env["EPS_KATZ"], env["EPS_EV"]]
    scale = [np.float32(0.85), np.float32(alpha), np.float32(1.0 / lam)]
    inits = [orc.init(idx, orc.ORC_PR), orc.init(idx, orc.ORC_KATZ, 1.0), orc.init(idx, orc.ORC_EV, lam=np.float32(lam))]
    xs, rs, passes, _ = run_passes(idx, adj, [nat.EXC_PR, nat.EXC_KATZ, nat.EXC_EV], eps, scale, [1, 0, 0], inits)
    refs = [orc.fp64_reference(idx, adj, orc.ORC_PR)[0], orc.fp64_reference(idx, adj, orc.ORC_KATZ, alpha=alpha)[0],
            orc.fp64_reference(idx, adj, orc.ORC_EV)[0]]
    c = [orc.compare(refs[i], xs[i], unit=(i == 2)) for i in range(3)]
    assert np.abs(xs[0] - refs[0]).sum() <= np.abs(rs[0].astype(np.float64)).sum() / 0.15 * 1.001
    assert c[1]["l1_rel"] < 0.05 and c[1]["top100"] >= 85
    assert c[2]["l1_rel"] < 0.1 and c[2]["top100"] >= 80
    for i, (b, w, dd) in enumerate([(0.15, 0.85, True), (1.0, float(scale[1]), False), (0.0, float(scale[2]), False)]):
        tr = true_residual(idx, adj, xs[i], b, w, dd)
        assert np.abs(tr - rs[i]).max() < 1e-3 * max(1.0, np.abs(xs[i]).max()), i


def test_edgeless_graph():
    n = 1000
    idx = np.zeros(n + 1, np.uint64)
    adj = np.zeros(0, np.int32)
    x0, r0 = orc.init(idx, orc.ORC_PR)
    xs, rs, passes, pushed = run_passes(idx, adj, [nat.EXC_PR], [0.01], [0.85], [1], [(x0, r0)])
    assert np.all(xs[0] == np.float32(1) - np.float32(0.85)) and np.all(rs[0] == 0) and pushed == 0 and passes == 1


def test_directed_cycle_closed_form():
    n = 5000
    idx = np.arange(n + 1, dtype=np.uint64)
    adj = ((np.arange(n) + 1) % n).astype(np.int32)
    x0, r0 = orc.init(idx, orc.ORC_PR)
    xs, rs, _, _ = run_passes(idx, adj, [nat.EXC_PR], [1e-4], [0.85], [1], [(x0, r0)])
    # x* = 1 everywhere; error bounded by ||r||_1 / 0.15
    err = np.abs(xs[0].astype(np.float64) - 1.0).sum()
    assert err <= np.abs(rs[0].astype(np.float64)).sum() / 0.15 + 1e-2


@pytest.mark.parametrize("kernel,group", [(nat.KERNEL_LB, 0), (nat.KERNEL_GROUP, 32)])
def test_star_hub_pieces(kernel, group):
    """Vertex 0 points at 20000 leaves (far above the hub threshold), leaves point back."""
    L = 20000
    n = L + 1
    deg = np.concatenate([[L], np.ones(L, np.int64)])
    idx = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    adj = np.concatenate([np.arange(1, n), np.zeros(L, np.int64)]).astype(np.int32)
    x0, r0 = orc.init(idx, orc.ORC_PR)
    xs, rs, _, pushed = run_passes(idx, adj, [nat.EXC_PR], [1e-5], [0.85], [1], [(x0, r0)], kernel=kernel, group=group)
    # closed form: hub h = 0.15 + 0.85 * L * l, leaf l = 0.15 + 0.85 * h / L
    h = (0.15 + 0.85 * 0.15 * L) / (1 - 0.85 * 0.85)
    leaf = 0.15 + 0.85 * h / L
    ref = np.concatenate([[h], np.full(L, leaf)])
    err = np.abs(xs[0].astype(np.float64) - ref).sum()
    # exact bound from the FP64 residual of the returned x (float32 accumulation of 20000
    # adds into a hub of value ~9190 makes the state's own r drift from the true residual)
    true_l1 = np.abs(true_residual(idx, adj, xs[0], 0.15, 0.85, True)).sum()
    assert err <= true_l1 / 0.15 * (1 + 1e-6)
    assert err / ref.sum() < 2e-3
    assert pushed >= 2 * L


@pytest.mark.parametrize("tile", ["0", "1024", "2048", "4096"])
@pytest.mark.parametrize("threshold", [0.0, 0.2])
@pytest.mark.parametrize("gname", ["sample", "rmat12", "star"])
def test_fold_pull_dense_passes(threshold, gname, tile, rmat12, monkeypatch):
    """Dense passes as fold + pull over the device transpose: same bounds and invariant."""
    monkeypatch.setenv("EXC_PULL_TILE", tile)   # 0: k_pull; else k_pull_sorted with this tile
    if gname == "star":
        L = 9000   # in-degree of the hub above the pull hub threshold
        deg = np.concatenate([[L], np.ones(L, np.int64)])
        idx = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
        adj = np.concatenate([np.arange(1, L + 1), np.zeros(L, np.int64)]).astype(np.int32)
    else:
        idx, adj = orc.egr_load(SAMPLE_EGR if gname == "sample" else rmat12)
    n = len(idx) - 1
    x0, r0 = orc.init(idx, orc.ORC_PR)
    xs, rs, passes, pushed = run_passes(idx, adj, [nat.EXC_PR], [1e-3], [0.85], [1], [(x0, r0)], pull=threshold)
    b = np.full(n, np.float64(np.float32(1) - np.float32(0.85)))
    ref, _, _ = orc.fp64_solve_b(idx, adj, b)
    err = np.abs(xs[0].astype(np.float64) - ref).sum()
    true_l1 = np.abs(true_residual(idx, adj, xs[0], b, 0.85, True)).sum()
    assert err <= true_l1 / 0.15 * (1 + 1e-6) + 1e-6
    assert np.all(np.abs(rs[0]) <= 1e-3)
    tr = true_residual(idx, adj, xs[0], b, 0.85, True)
    assert np.abs(tr - rs[0]).max() < 2e-4 * max(1.0, np.abs(xs[0]).max())
    assert pushed >= len(adj)


@pytest.mark.parametrize("tile", ["0", "2048"])
def test_fold_pull_all_measures(tile, monkeypatch):
    monkeypatch.setenv("EXC_PULL_TILE", tile)   # 0: k_pull; else k_pull_sorted with this tile
    from tests.conftest import SAMPLE_ENV, read_env
    env = read_env(SAMPLE_ENV)
    idx, adj = orc.egr_load(SAMPLE_EGR)
    lam, alpha = env["LAMBDA"], env["ALPHA"]
    eps = [0.001, env["EPS_KATZ"] / 10, env["EPS_EV"] / 10]
    scale = [np.float32(0.85), np.float32(alpha), np.float32(1.0 / lam)]
    inits = [orc.init(idx, orc.ORC_PR), orc.init(idx, orc.ORC_KATZ, 1.0), orc.init(idx, orc.ORC_EV, lam=np.float32(lam))]
    xs, rs, _, _ = run_passes(idx, adj, [nat.EXC_PR, nat.EXC_KATZ, nat.EXC_EV], eps, scale, [1, 0, 0], inits, pull=0.0)
    for i, (bv, w, dd) in enumerate([(0.15, 0.85, True), (1.0, float(scale[1]), False), (0.0, float(scale[2]), False)]):
        tr = true_residual(idx, adj, xs[i], bv, w, dd)
        assert np.abs(tr - rs[i]).max() < 1e-3 * max(1.0, np.abs(xs[i]).max()), i
    assert orc.compare(orc.fp64_reference(idx, adj, orc.ORC_KATZ, alpha=alpha)[0], xs[1])["l1_rel"] < 0.01


@pytest.mark.parametrize("live", ["0", "1"])
@pytest.mark.parametrize("tile", ["0", "2048"])
@pytest.mark.parametrize("gname", ["rmat12", "star"])
def test_dense_passes_live_rows_forced(live, gname, tile, rmat12, monkeypatch):
    """The live-row bitmap (a source's row is gathered only if it pushed since the last pass)
    forced off, and forced on also in the first passes, where nearly every source is pending:
    the same bounds and invariant as the default choice by demand."""
    monkeypatch.setenv("EXC_PULL_LIVE", live)
    test_fold_pull_dense_passes(0.0, gname, tile, rmat12, monkeypatch)

@pytest.mark.parametrize("tile", ["0", "0:256", "1024", "4096", "4096/1"])
@pytest.mark.parametrize("measures", ["pr", "all"])
def test_fused_snapshot_bit_exact(rmat12, measures, tile, monkeypatch):
    """The snapshot a dense pass computes for the next pass inside k_pull (bitmap words,
    per-chunk pending counts and degree sums) equals the oracle's snapshot of the state
    read back, bit for bit, on chunk boundaries that cut through the pull's tiles; and
    equals a second, separately launched snapshot of the same state.  (The next pass's
    fold is kept out of k_pull here, so the state read back is the one the snapshot saw.)
    k_pull runs 64-destination tiles on a graph this small; "0:256" forces 256."""
    tile, _, small = tile.partition(":")
    monkeypatch.setenv("EXC_PULL_TILE", tile)   # 0: k_pull; else k_pull_sorted with this tile
    monkeypatch.setenv("EXC_PULL_SMALL", "0" if small == "256" else "2")
    monkeypatch.setenv("EXC_FUSE_FOLD", "0")
    idx, adj = orc.egr_load(rmat12)
    n, m = len(idx) - 1, len(adj)
    cv0 = np.array([0, 1000, 1001, 2560, 3333], np.int32)
    if measures == "pr":
        kinds, eps, scale, divdeg = [nat.EXC_PR], [1e-3], [0.85], [1]
        inits = [orc.init(idx, orc.ORC_PR)]
    else:
        kinds, eps = [nat.EXC_PR, nat.EXC_KATZ, nat.EXC_EV], [1e-3, 1e-4, 1e-4]
        # contractive constants (alpha and 1/lambda well below 1/lambda1 of this graph)
        scale, divdeg = [np.float32(0.85), np.float32(0.001), np.float32(0.001)], [1, 0, 0]
        inits = [orc.init(idx, orc.ORC_PR), orc.init(idx, orc.ORC_KATZ, 1.0), orc.init(idx, orc.ORC_EV, lam=np.float32(1000))]
    with nat.Context(n, m, cv0, kinds, eps, scale, divdeg) as ctx:
        ctx.upload_offsets(idx)
        for i, (x0, r0) in enumerate(inits):
            ctx.set_state(i, x=x0, r=r0)
        assert ctx.cache_chunks(m) == 0
        ctx.cache_upload(adj, 0)
        assert ctx.enable_pull(0.0) == 0
        checked = 0
        for p in range(200):
            cnt, _ = ctx.snapshot()
            bm = ctx.bitmap()
            rs = [ctx.read_state(i, 1) for i in range(len(kinds))]
            obm, ocnt, _ = orc.snapshot(rs, np.asarray(eps, np.float32), idx, cv0)
            assert np.array_equal(bm, obm), p
            assert np.array_equal(cnt, ocnt), p
            if p > 0:   # the state is unchanged: a launched snapshot agrees with the fused one
                cnt2, _ = ctx.snapshot()
                assert np.array_equal(cnt2, cnt) and np.array_equal(ctx.bitmap(), bm), p
                checked += 1
            if cnt[0::2].sum() == 0:
                break
            ctx.push_resident(0, n, 0, m, False)
            ctx.pass_sync()
        assert checked >= 3


@pytest.mark.parametrize("live", ["0", "1"])
@pytest.mark.parametrize("tile", ["0", "4096/1"])
def test_fused_snapshot_live_rows_forced(rmat12, live, tile, monkeypatch):
    """The fused snapshot stays bit-exact with the live-row bitmap forced off and on, over
    three-measure rows (k_pull) and the one-measure passes of the source-sorted pull."""
    monkeypatch.setenv("EXC_PULL_LIVE", live)
    test_fused_snapshot_bit_exact(rmat12, "all", tile, monkeypatch)


def test_fused_fold_state_and_results(rmat12):
    """With the next pass's fold inside k_pull: the context reports that a dense pass is
    required, refuses a partial push, and the run ends within the same FP64 bound and
    invariant as with the separate fold; both runs push the same vertices' edges per pass
    count to within the atomic-order noise of the pass structure."""
    import os
    idx, adj = orc.egr_load(rmat12)
    n, m = len(idx) - 1, len(adj)
    x0, r0 = orc.init(idx, orc.ORC_PR)
    cv0 = np.array([0], np.int32)
    with nat.Context(n, m, cv0, [nat.EXC_PR], [1e-4], [0.85], [1]) as ctx:
        ctx.upload_offsets(idx)
        ctx.set_state(0, x=x0, r=r0)
        assert ctx.cache_chunks(m) == 0
        ctx.cache_upload(adj, 0)
        assert ctx.enable_pull(0.0) == 0
        saw_folded = False
        for p in range(3000):
            cnt, _ = ctx.snapshot()
            if cnt[0::2].sum() == 0:
                break
            if ctx.dense_required():
                saw_folded = True
                with pytest.raises(nat.ExcError, match="folded"):
                    ctx.push_resident(0, n // 2, 0, int(idx[n // 2]), False)
            ctx.push_resident(0, n, 0, m, False)
            ctx.pass_sync()
        assert saw_folded
        x, r = ctx.read_state(0, 0), ctx.read_state(0, 1)
        st = ctx.stats()
    b = np.full(n, np.float64(np.float32(1) - np.float32(0.85)))
    ref, _, _ = orc.fp64_solve_b(idx, adj, b)
    tr = true_residual(idx, adj, x, b, 0.85, True)
    assert np.abs(x.astype(np.float64) - ref).sum() <= np.abs(tr).sum() / 0.15 * (1 + 1e-6) + 1e-6
    assert np.all(np.abs(r) <= 1e-4)
    assert np.abs(tr - r).max() < 2e-4 * max(1.0, np.abs(x).max())
    assert st["pushed_edges"] >= m and st["pull_passes"] > 0


@pytest.mark.parametrize("tile", ["0", "0:256", "1024"])
@pytest.mark.parametrize("n", [1, 255, 257, 1000, 1025])
def test_dense_passes_edgeless_and_ragged(n, tile, monkeypatch):
    """Dense passes on an edgeless graph (no sources, empty transpose, no p rows): x = b in
    one pass, with tile counts that are not multiples of the 64- or 256-vertex tile."""
    tile, _, small = tile.partition(":")
    monkeypatch.setenv("EXC_PULL_TILE", tile)   # 0: k_pull; else k_pull_sorted with this tile
    monkeypatch.setenv("EXC_PULL_SMALL", "0" if small == "256" else "2")
    idx = np.zeros(n + 1, np.uint64)
    adj = np.zeros(0, np.int32)
    x0, r0 = orc.init(idx, orc.ORC_PR)
    xs, rs, passes, pushed = run_passes(idx, adj, [nat.EXC_PR], [0.01], [0.85], [1], [(x0, r0)], pull=0.0)
    assert np.all(xs[0] == np.float32(1) - np.float32(0.85)) and np.all(rs[0] == 0) and pushed == 0 and passes == 1


@pytest.mark.parametrize("tile", ["0", "2048"])
@pytest.mark.parametrize("gname", ["rmat12", "star"])
def test_weighted_dense_passes(gname, tile, rmat12, monkeypatch):
    """Dense passes on a weighted graph (the pull scales each in-edge's row by its
    coefficient, staged with the ids; hub pieces too): the result sits within the exact
    FP64 bound of the weighted solve, per-vertex |r| <= eps at the end, and weights set
    after the pull is enabled are refused."""
    monkeypatch.setenv("EXC_PULL_TILE", tile)   # 0: k_pull; else k_pull_sorted with this tile
    if gname == "star":
        L = 9000
        deg = np.concatenate([[L], np.ones(L, np.int64)])
        idx = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
        adj = np.concatenate([np.arange(1, L + 1), np.zeros(L, np.int64)]).astype(np.int32)
    else:
        idx, adj = orc.egr_load(rmat12)
    n, m = len(idx) - 1, len(adj)
    rng = np.random.default_rng(8)
    w = rng.uniform(0.5, 1.5, m)
    src = np.repeat(np.arange(n), np.diff(idx.astype(np.int64)))
    S = np.bincount(src, weights=w, minlength=n)
    wn = (w / S[src]).astype(np.float32)
    eps = 1e-5
    b = np.full(n, np.float64(np.float32(1) - np.float32(0.85)))
    with nat.Context(n, m, np.array([0], np.int32), [nat.EXC_PR], [eps], [0.85], [0]) as ctx:
        ctx.upload_offsets(idx)
        ctx.set_state(0, x=np.zeros(n, np.float32), r=b.astype(np.float32))
        assert ctx.cache_chunks(m) == 0
        ctx.cache_upload(adj, 0)
        check = nat.F["exc_set_edge_weights"](ctx.ctx, nat.ptr(wn), m)
        assert check == 0
        assert ctx.enable_pull(0.0) == 0
        assert nat.F["exc_set_edge_weights"](ctx.ctx, nat.ptr(wn), m) != 0   # after the pull: refused
        for p in range(5000):
            cnt, _ = ctx.snapshot()
            if cnt[0::2].sum() == 0:
                break
            ctx.push_resident(0, n, 0, m, False)
            ctx.pass_sync()
        x, r = ctx.read_state(0, 0), ctx.read_state(0, 1)
        assert ctx.stats()["pull_passes"] > 0
    sf = float(np.float32(0.85))   # the scale the GPU multiplies by
    ref, _, _ = orc.fp64_solve_w(idx, adj, wn.astype(np.float64), b, s=sf)
    acc = np.zeros(n)
    np.add.at(acc, adj, sf * wn.astype(np.float64) * x.astype(np.float64)[src])
    tr = b + acc - x.astype(np.float64)
    assert np.abs(x.astype(np.float64) - ref).sum() <= np.abs(tr).sum() / (1 - sf) * (1 + 1e-6) + 1e-6
    assert np.all(np.abs(r) <= eps)


def test_weighted_graph_refuses_several_measures(rmat12):
    """Per-edge coefficients are accepted only by a one-measure context: with two or three
    measures the call fails (the weighted dense pull is built for one measure)."""
    idx, adj = orc.egr_load(rmat12)
    n, m = len(idx) - 1, len(adj)
    wn = np.full(m, 0.5, np.float32)
    for kinds in ([nat.EXC_PR, nat.EXC_KATZ], [nat.EXC_PR, nat.EXC_KATZ, nat.EXC_EV]):
        k = len(kinds)
        with nat.Context(n, m, np.array([0], np.int32), kinds, [1e-3] * k, [0.85] * k, [1] + [0] * (k - 1)) as ctx:
            ctx.upload_offsets(idx)
            assert ctx.cache_chunks(m) == 0
            ctx.cache_upload(adj, 0)
            assert nat.F["exc_set_edge_weights"](ctx.ctx, nat.ptr(wn), m) == nat.EXC_EINVAL
            assert "one measure" in nat.F["exc_last_error"]().decode()


@pytest.mark.parametrize("nm", [1, 3])
def test_hub_destinations_privatised(nm):
    """Hub pieces whose lists converge on a few destinations (40 sources of 3000 out-edges
    into the same 3000 vertices, every such vertex pointing back to 2 sources): with those
    destinations (and some vertices no list reaches) privatised in LDS, the run meets the
    FP64 bound and the per-vertex invariant like the run without privatisation, and the two
    results agree within the push's float rounding."""
    S_, T = 40, 3000
    n = S_ + T + 10
    lists = [np.arange(S_, S_ + T) for _ in range(S_)] + [np.array([(v * 7) % S_, (v * 13 + 1) % S_])
                                                       for v in range(T)] + [np.zeros(0, np.int64)] * 10
    deg = np.array([len(l) for l in lists])
    idx = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    adj = np.concatenate(lists).astype(np.int32)
    m = len(adj)
    kinds = [nat.EXC_PR, nat.EXC_KATZ, nat.EXC_EV][:nm]
    scale = [0.85, 1e-4, 1e-4][:nm]
    hubs = np.concatenate([np.arange(S_, S_ + 1000), np.arange(n - 10, n)]).astype(np.int32)
    results = []
    for table in (None, hubs):
        with nat.Context(n, m, np.array([0], np.int32), kinds, [1e-6] * nm, scale, [1] + [0] * (nm - 1),
                         timing=True) as ctx:
            ctx.upload_offsets(idx)
            for i in range(nm):
                ctx.set_state(i, x=np.zeros(n, np.float32), r=np.full(n, 0.15, np.float32))
            assert ctx.cache_chunks(m) == 0
            ctx.cache_upload(adj, 0)
            if table is not None:
                ctx.set_hub_destinations(table)
            for p in range(5000):
                cnt, _ = ctx.snapshot()
                if cnt[0::2].sum() == 0:
                    break
                ctx.push_resident(0, n, 0, m, False)
                ctx.pass_sync()
            results.append(([ctx.read_state(i, 0) for i in range(nm)], [ctx.read_state(i, 1) for i in range(nm)]))
            assert ctx.stats()["piece_launches"] > 0
    ref, _, _, _ = orc.fp64_reference(idx, adj, orc.ORC_PR, tol=1e-13, max_iter=5000)
    for xs, rs in results:
        x = xs[0].astype(np.float64)
        true = true_residual(idx, adj, xs[0], 0.15, 0.85, True)
        assert np.abs(x - ref).sum() <= np.abs(true).sum() / 0.15 * (1 + 1e-6) + 1e-6
        assert np.abs(true - rs[0]).max() < 2e-4 * max(1.0, np.abs(x).max())
        for i in range(1, nm):   # the other measures keep their invariant r = b + W x - x too
            tr = true_residual(idx, adj, xs[i], 0.15, scale[i], False)
            assert np.abs(tr - rs[i]).max() < 1e-4 * max(1.0, np.abs(xs[i]).max())
    np.testing.assert_allclose(results[0][0][0], results[1][0][0], rtol=1e-4, atol=1e-5)

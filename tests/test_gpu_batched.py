"""Batched personalised PageRank on the GPU ([n][K] rows, one wave per vertex, lanes =
columns; no reference counterpart): the snapshot against the oracle's multi-measure
snapshot bit for bit, and every column against the FP64 personalised solution within
the exact bound ||x - x*||_1 <= ||b + W x - x||_1 / (1 - d), and against a plain
single-column run."""
import os
import re

import numpy as np
import pytest

import excent_amd.native as nat
from oracle import oracle as orc
from tests.conftest import SAMPLE_EGR

pytestmark = pytest.mark.gpu


def true_residual(idx, adj, x, b):
    n = len(idx) - 1
    deg = np.diff(idx.astype(np.int64))
    src = np.repeat(np.arange(n), deg)
    acc = np.zeros(n)
    np.add.at(acc, adj, 0.85 * x.astype(np.float64)[src] / deg[src])
    return b + acc - x.astype(np.float64)


@pytest.mark.parametrize("K", [1, 7, 64])
def test_batched_snapshot_bit_exact(K, rmat12):
    idx, adj = orc.egr_load(rmat12)
    n = len(idx) - 1
    rng = np.random.default_rng(K)
    eps = (rng.random(K) * 0.02 + 0.001).astype(np.float32)
    rs = [(rng.standard_normal(n) * 0.01 * (rng.random(n) < 0.3)).astype(np.float32) for _ in range(K)]
    cv0 = orc.plan_chunks(idx, len(adj) // 6)
    with nat.Context(n, len(adj), cv0, [nat.EXC_PR], [0.01], [0.85], [1], batch_k=K) as ctx:
        ctx.upload_offsets(idx)
        ctx.batch_eps(eps)
        for j in range(K):
            ctx.set_state(j, r=rs[j])
        cnt, _ = ctx.snapshot()
        bm = ctx.bitmap()
        colpend = ctx.batch_colpend()
        # inactive columns do not count
        ctx.batch_set_active((1 << (K // 2)) - 1 if K > 1 else 1)
        cnt2, _ = ctx.snapshot()
    obm, ocnt, _ = orc.snapshot(rs, eps, idx, cv0)
    assert np.array_equal(bm, obm) and np.array_equal(cnt, ocnt)
    assert np.array_equal(colpend, [(np.abs(r) > e).sum() for r, e in zip(rs, eps)])
    act = K // 2 if K > 1 else 1
    _, ocnt2, _ = orc.snapshot(rs[:act], eps[:act], idx, cv0)
    assert np.array_equal(cnt2, ocnt2)


@pytest.mark.parametrize("K", [7, 64])
@pytest.mark.parametrize("use_list", [True, False])
def test_batched_front_snapshot_bit_exact(K, use_list, rmat12):
    """After push passes that pushed every chunk with pending vertices, the snapshot reads
    only the vertices the pushes added to: bitmap, per-chunk counts and per-column counts
    equal the oracle's snapshot of the state read back, bit for bit, on every pass until
    nothing pends (columns dropping out as they finish).  A pass that leaves a pending chunk
    unpushed, a change of eps and a column switched back on each fall back to a full
    snapshot, which is exact as well."""
    print(f"case front snapshot K={K} list={use_list}", flush=True)
    idx, adj = orc.egr_load(rmat12)
    n, m = len(idx) - 1, len(adj)
    rng = np.random.default_rng(K + 5)
    seeds = [rng.choice(n, 4, replace=False).astype(np.int32) for _ in range(K)]
    eps = (rng.random(K) * 2e-3 + 1e-3).astype(np.float32)
    cv0 = orc.plan_chunks(idx, m // 6)
    nc = len(cv0)
    bounds = list(cv0) + [n]

    def check(ctx, cnt, active, tag):
        rs = [ctx.read_state(j, 1) for j in range(K)]
        on = [j for j in range(K) if (active >> j) & 1]
        obm, ocnt, _ = orc.snapshot([rs[j] for j in on], eps[on], idx, cv0)
        assert np.array_equal(ctx.bitmap(), obm) and np.array_equal(cnt, ocnt), tag
        cp = ctx.batch_colpend()
        assert all(cp[j] == ((np.abs(rs[j]) > eps[j]).sum() if j in on else 0) for j in range(K)), tag

    with nat.Context(n, m, cv0, [nat.EXC_PR], [1e-3], [0.85], [1], batch_k=K) as ctx:
        ctx.upload_offsets(idx)
        assert ctx.cache_chunks(m) == 0
        ctx.cache_upload(adj)
        ctx.batch_eps(eps)
        for j in range(K):
            ctx.batch_seed(j, seeds[j], np.float32(0.15 * n / 4))
        active = (1 << K) - 1 if K < 64 else ~0 & ((1 << 64) - 1)
        cnt, _ = ctx.snapshot()
        check(ctx, cnt, active, "first")
        fronts, passes, full_next, reenabled = 0, 0, False, False
        while cnt[0::2].sum() and passes < 300:
            full = full_next or passes in (3, 5)
            last = max(c for c in range(nc) if cnt[2 * c])
            for c in range(nc):   # pass 3 leaves the last pending chunk out
                if cnt[2 * c] and not (passes == 3 and c == last):
                    ctx.push_resident(bounds[c], bounds[c + 1], int(idx[bounds[c]]), int(idx[bounds[c + 1]]), use_list)
            if passes == 5:
                ctx.batch_eps(eps)   # the same values: the pending test is reset all the same
            before = ctx.stats()["front_snapshots"]
            cnt, _ = ctx.snapshot()
            took_front = ctx.stats()["front_snapshots"] - before
            assert took_front == (0 if full else 1), (passes, took_front)
            fronts += took_front
            check(ctx, cnt, active, passes)
            cp = ctx.batch_colpend()
            new = sum(1 << j for j in range(K) if cp[j])
            full_next = False
            if new != active and not reenabled and passes >= 6:
                ctx.batch_set_active(new)
                ctx.batch_set_active(active)   # the finished columns back on: a full snapshot next
                reenabled = full_next = True
            else:
                active = new
                ctx.batch_set_active(active)
            passes += 1
        assert cnt[0::2].sum() == 0 and fronts >= 3, (passes, fronts)


@pytest.mark.parametrize("K", [16, 64])
@pytest.mark.parametrize("gname", ["sample", "rmat12"])
def test_batched_ppr_against_fp64(tmp_path, gname, rmat12, K):
    """Every column of a K-column run (K = 64: the full [n][64] rows of the batched
    configuration) within the exact FP64 bound of its own personalised solution."""
    path = SAMPLE_EGR if gname == "sample" else rmat12
    idx, adj = orc.egr_load(path)
    n = len(idx) - 1
    S = 4
    print(f"case batched ppr {gname} K={K}", flush=True)
    rng = np.random.default_rng(3)
    seeds = np.stack([rng.choice(n, S, replace=False) for _ in range(K)]).astype(np.int32)
    sf = str(tmp_path / "seeds.i32")
    seeds.tofile(sf)
    eps = 1e-4
    log = str(tmp_path / "b.log")
    rc = nat.excent([path, "--measure", "ppr", "--ppr-k", K, "--ppr-seeds", sf, "--eps-ppr", eps, "--dump",
                     str(tmp_path / "o"), "--stream-ratio", "0.5", "--log", log])
    text = open(log).read()
    assert rc == 0, text
    assert "active columns" in text and re.search(rf"{K} columns, ppr sum", text)
    X = np.fromfile(str(tmp_path / "o.ppr.bin"), np.float32).reshape(K, n)
    mass = np.float32((1 - np.float64(np.float32(0.85))) * n / S)
    for j in range(K):
        b = np.zeros(n)
        np.add.at(b, seeds[j], np.float64(mass))
        ref, _, _ = orc.fp64_solve_b(idx, adj, b)
        err = np.abs(X[j] - ref).sum()
        bound = np.abs(true_residual(idx, adj, X[j], b)).sum() / 0.15
        assert err <= bound * (1 + 1e-6) + 1e-6, (j, err, bound)
        assert err <= eps * n / 0.15


def test_batched_equals_plain_runs_within_bound(rmat12):
    """Column j of a batched run and a plain run from the same teleport vector both sit
    within eps-bounds of the same FP64 solution."""
    idx, adj = orc.egr_load(rmat12)
    n = len(idx) - 1
    K = 4
    rng = np.random.default_rng(9)
    seeds = [rng.choice(n, 8, replace=False).astype(np.int32) for _ in range(K)]
    mass = np.float32(0.15 * n / 8)
    cv0 = np.array([0], np.int32)

    def run():
        with nat.Context(n, len(adj), cv0, [nat.EXC_PR], [1e-4], [0.85], [1], batch_k=K) as ctx:
            ctx.upload_offsets(idx)
            assert ctx.cache_chunks(len(adj)) == 0
            ctx.cache_upload(adj)
            for j in range(K):
                ctx.batch_seed(j, seeds[j], mass)
            for it in range(5000):
                cnt, _ = ctx.snapshot()
                cp = ctx.batch_colpend()
                ctx.batch_set_active(sum(1 << j for j in range(K) if cp[j]))
                if cnt[0] == 0:
                    break
                ctx.push_resident(0, n, 0, len(adj), cnt[0] < 0.5 * n)
                ctx.pass_sync()
            return [ctx.read_state(j, 0) for j in range(K)], ctx.stats()

    X, st = run()
    assert st["pushed_edges"] > 0
    for j in range(K):
        b = np.zeros(n)
        np.add.at(b, seeds[j], np.float64(mass))
        ref, _, _ = orc.fp64_solve_b(idx, adj, b)
        # plain run of the same column through the oracle's CPU push
        x = [np.zeros(n, np.float32)]
        r = [b.astype(np.float32)]
        orc.run_push(idx, adj, cv0, x, r, [1e-4], [np.float32(0.85)], [1], threads=4)
        for xx in (X[j], x[0]):
            err = np.abs(xx - ref).sum()
            assert err <= np.abs(true_residual(idx, adj, xx, b)).sum() / 0.15 * (1 + 1e-6) + 1e-6


def test_session_column_runs_match_batched_columns(rmat12):
    """bin/excent --measure pr --ppr-column j (the plain solver on column j's teleport set)
    and --measure ppr (all K columns at once) through the session API: the plain start is
    exactly column j's seeds at mass (1 - d) n / S, and both results sit within the exact
    FP64 bound of the product's vector-constant FP64 solve, itself pinned to the oracle's."""
    idx, adj = orc.egr_load(rmat12)
    n = len(idx) - 1
    gen, eps = "4:6:2:11", 1e-5
    common = [rmat12, "--cache", "chunks", "--preload", "--ppr-gen", gen, "--eps-ppr", eps]
    with nat.Session(common + ["--measure", "ppr"]) as S:
        S.run(reinit=False)
        xb = [S.values(j) for j in range(4)]
    mass = np.float32((1.0 - float(np.float32(0.85))) * n / 6)
    with nat.Session(common + ["--measure", "pr", "--ppr-column", 0, "--pull", "0.001"]) as S, \
            nat.Fp64Solver(idx, adj) as F:
        for j in range(4):
            S.set_column(j)
            r0 = S.values(0, which=1)
            assert np.count_nonzero(r0) == 6 and np.all(r0[r0 != 0] == mass)
            S.run(reinit=False)
            xp = S.values(0)
            c = r0.astype(np.float64)
            ref, _, _ = F.solve_vec(c, float(np.float32(0.85)), 1, tol=1e-13)
            oref, _, _ = orc.fp64_solve_b(idx, adj, c, s=float(np.float32(0.85)))
            assert np.allclose(ref, oref, rtol=1e-9, atol=1e-9)
            for x in (xp, xb[j]):
                bound = F.residual(x, c, float(np.float32(0.85)), 1) / (1 - 0.85)
                assert np.abs(x - ref).sum() <= bound * (1 + 1e-6) + 1e-6
                assert np.abs(x - ref).sum() <= eps * n / (1 - 0.85)


def test_batched_session_rerun_after_reinit(rmat12):
    """A second batched run from the restored initial state pushes again (every column
    becomes active with the restored state) and lands within the same bound."""
    idx, adj = orc.egr_load(rmat12)
    n = len(idx) - 1
    with nat.Session([rmat12, "--cache", "chunks", "--preload", "--ppr-gen", "3:5:2:4", "--eps-ppr", 1e-5,
                      "--measure", "ppr"]) as S:
        r1 = S.run(reinit=False)
        x1 = [S.values(j) for j in range(3)]
        r2 = S.run(reinit=True)
        x2 = [S.values(j) for j in range(3)]
    assert r1["passes"] > 0 and r2["passes"] > 0
    assert r2["kernel"]["pushed_edge_measures"] > 0
    for a, b in zip(x1, x2):
        assert np.abs(a - b).sum() <= 2 * 1e-5 * n / (1 - 0.85)


def test_batched_pull_rows_fixed_by_env(rmat12, monkeypatch):
    """EXC_BPULL_Q fixes the batched pull's rows in flight and turns the tuner off; the
    columns still meet their FP64 bound."""
    monkeypatch.setenv("EXC_BPULL_Q", "16")
    idx, adj = orc.egr_load(rmat12)
    n, m = len(idx) - 1, len(adj)
    K, S, eps = 3, 2, 1e-4
    rng = np.random.default_rng(5)
    seeds = [rng.choice(n, S, replace=False).astype(np.int32) for _ in range(K)]
    mass = np.float32(0.15 * n / S)
    with nat.Context(n, m, np.array([0], np.int32), [nat.EXC_PR], [eps], [0.85], [1], batch_k=K) as ctx:
        ctx.upload_offsets(idx)
        assert ctx.cache_chunks(m) == 0
        ctx.cache_upload(adj)
        for j in range(K):
            ctx.batch_seed(j, seeds[j], mass)
        assert ctx.enable_pull(0.0) == 0
        for it in range(5000):
            cnt, _ = ctx.snapshot()
            cp = ctx.batch_colpend()
            ctx.batch_set_active(sum(1 << j for j in range(K) if cp[j]))
            if cnt[0::2].sum() == 0:
                break
            ctx.push_resident(0, n, 0, m, False)
            ctx.pass_sync()
        X = [ctx.read_state(j, 0) for j in range(K)]
        st = ctx.stats()
    assert st["pull_rows_in_flight"] == 16 and list(st["pull_tune_ns_per_edge"]) == [0.0, 0.0]
    for j in range(K):
        b = np.zeros(n)
        np.add.at(b, seeds[j], np.float64(mass))
        ref, _, _ = orc.fp64_solve_b(idx, adj, b)
        bound = np.abs(true_residual(idx, adj, X[j], b)).sum() / 0.15
        assert np.abs(X[j] - ref).sum() <= bound * (1 + 1e-6) + 1e-6, j


@pytest.mark.parametrize("K", [5, 64])
@pytest.mark.parametrize("rows", ["bf16", "f32"])
@pytest.mark.parametrize("fuse", [0, 1])
@pytest.mark.parametrize("gname", ["rmat12", "star"])
def test_batched_dense_passes(gname, fuse, rows, K, rmat12, monkeypatch):
    """Batched fold + pull over the transpose (every pass dense): each column within the
    exact FP64 bound, per-vertex |r| <= eps at the end, the residual invariant
    r = b + W x - x per vertex within 1e-4 |x| + eps (float rounding reaches 2e-5 |x| at the
    star's centre; bf16 weight rows keep their rounding remainder in r, and without that the
    oracle's run misses this tolerance by 7x or more), and in-lists above the hub
    threshold (the star's centre has 9000 in-neighbours) summed by pieces.  fuse = 0: the
    fold stays in k_bfold and the snapshot after every pass is checked bit for bit against
    the oracle's; fuse = 1: k_bpull also computes the next snapshot and fold (the state read
    back after such a pass is already folded, so only the end results are checked).
    rows: the weight rows' storage (EXC_BATCH_ROWS).  K = 64 fills every lane of a wave with a
    column (bf16 rows: a gather covers two rows, lanes 0-31 and 32-63)."""
    monkeypatch.setenv("EXC_FUSE_FOLD", str(fuse))
    monkeypatch.setenv("EXC_BATCH_ROWS", rows)
    print(f"case batched dense passes {gname} fuse={fuse} rows={rows} K={K}", flush=True)
    if gname == "star":
        L = 9000
        deg = np.concatenate([[L], np.ones(L, np.int64)])
        idx = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
        adj = np.concatenate([np.arange(1, L + 1), np.zeros(L, np.int64)]).astype(np.int32)
    else:
        idx, adj = orc.egr_load(rmat12)
    n, m = len(idx) - 1, len(adj)
    S, eps = 3, 1e-4
    rng = np.random.default_rng(21)
    seeds = [rng.choice(n, S, replace=False).astype(np.int32) for _ in range(K)]
    mass = np.float32(0.15 * n / S)
    cv0 = np.array([0, n // 3, n // 2], np.int32) if n > 10 else np.array([0], np.int32)
    with nat.Context(n, m, cv0, [nat.EXC_PR], [eps], [0.85], [1], batch_k=K) as ctx:
        ctx.upload_offsets(idx)
        assert ctx.cache_chunks(m) == 0
        ctx.cache_upload(adj)
        for j in range(K):
            ctx.batch_seed(j, seeds[j], mass)
        assert ctx.enable_pull(0.0) == 0
        folded_passes = 0
        for it in range(5000):
            cnt, _ = ctx.snapshot()
            if ctx.dense_required():
                folded_passes += 1
            else:
                rs = [ctx.read_state(j, 1) for j in range(K)]
                obm, ocnt, _ = orc.snapshot(rs, np.full(K, eps, np.float32), idx, cv0)
                assert np.array_equal(ctx.bitmap(), obm) and np.array_equal(cnt, ocnt), it
            cp = ctx.batch_colpend()
            ctx.batch_set_active(sum(1 << j for j in range(K) if cp[j]))
            if cnt[0::2].sum() == 0:
                break
            ctx.push_resident(0, n, 0, m, False)
            ctx.pass_sync()
        X = [ctx.read_state(j, 0) for j in range(K)]
        R = [ctx.read_state(j, 1) for j in range(K)]
        st = ctx.stats()
    assert st["pull_passes"] > 0 and st["pushed_edge_measures"] > 0
    assert (folded_passes > 0) == bool(fuse)
    # the rows-in-flight tuner timed 8 and 16 on the first two passes over half the edges
    # (rmat12's early passes) and kept the faster; without such passes it keeps 8
    t8, t16 = st["pull_tune_ns_per_edge"]
    if gname == "rmat12":
        assert t8 > 0 and t16 > 0 and st["pull_rows_in_flight"] == (16 if t16 < t8 else 8), st
    else:
        assert st["pull_rows_in_flight"] in (8, 16)
    for j in range(K):
        assert np.all(np.abs(R[j]) <= eps)
        b = np.zeros(n)
        np.add.at(b, seeds[j], np.float64(mass))
        ref, _, _ = orc.fp64_solve_b(idx, adj, b)
        err = np.abs(X[j] - ref).sum()
        true = true_residual(idx, adj, X[j], b)
        bound = np.abs(true).sum() / 0.15
        assert err <= bound * (1 + 1e-6) + 1e-6, (j, err, bound)
        dev = np.abs(R[j] - true) - (1e-4 * np.abs(X[j]) + eps)
        assert dev.max() <= 0, (j, int(dev.argmax()), dev.max())


@pytest.mark.parametrize("mode", ["stream_only", "gather_only", "mixed", "hot"])
def test_batched_out_of_core_regimes(tmp_path, rmat12, mode):
    """Batched personalised PageRank with the edges read per chunk from the EGR file -- every
    chunk streamed, every chunk's pending lists gathered, or both -- or partly held in the
    hot-vertex cache: one read of a chunk serves all K columns, and every column lands within
    its exact FP64 bound."""
    idx, adj = orc.egr_load(rmat12)
    n = len(idx) - 1
    K, S, eps = 8, 4, 1e-4
    rng = np.random.default_rng(11)
    seeds = np.stack([rng.choice(n, S, replace=False) for _ in range(K)]).astype(np.int32)
    sf = str(tmp_path / "seeds.i32")
    seeds.tofile(sf)
    extra = {"stream_only": ["--cache", "none", "--stream-ratio", "0"],
             "gather_only": ["--cache", "none", "--stream-ratio", "2"],
             "mixed": ["--cache", "none", "--stream-ratio", "0.2", "--dedup", "off"]}.get(mode)
    if mode == "hot":
        free, total = nat.ctypes.c_size_t(0), nat.ctypes.c_size_t(0)
        nat.F["exc_mem_info"](0, nat.ctypes.byref(free), nat.ctypes.byref(total))
        extra = ["--cache", "hot", "--reserve-mb", f"{free.value / (1 << 20) - 64:.3f}", "--stream-ratio", "0.3"]
    print(f"case batched out of core {mode}", flush=True)
    log = str(tmp_path / "b.log")
    rc = nat.excent([rmat12, "--measure", "ppr", "--ppr-k", K, "--ppr-seeds", sf, "--eps-ppr", eps, "--dump",
                     str(tmp_path / "o"), "--chunk-medges", "0.01", "--max-passes", "5000", "--log", log] + extra)
    text = open(log).read()
    assert rc == 0, text
    lines = [l for l in text.splitlines() if l.startswith("pass ")]
    assert lines
    if mode == "stream_only":
        assert all(" gather 0 " in l for l in lines)
    if mode == "gather_only":
        assert all(" stream 0 " in l for l in lines)
    if mode == "hot":
        assert "GPU degree cache:" in text
    X = np.fromfile(str(tmp_path / "o.ppr.bin"), np.float32).reshape(K, n)
    mass = np.float32((1 - np.float64(np.float32(0.85))) * n / S)
    for j in range(K):
        b = np.zeros(n)
        np.add.at(b, seeds[j], np.float64(mass))
        ref, _, _ = orc.fp64_solve_b(idx, adj, b)
        err = np.abs(X[j] - ref).sum()
        bound = np.abs(true_residual(idx, adj, X[j], b)).sum() / 0.15
        assert err <= bound * (1 + 1e-6) + 1e-6, (j, err, bound)
        assert err <= eps * n / 0.15

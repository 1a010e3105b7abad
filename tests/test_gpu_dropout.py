"""Dense passes over the measures still pending (exc_set_active_measures; no reference
counterpart: the reference pushes every measure of a pending vertex, src/excent.cu:100-124,
and a converged measure has no pending vertex, so it pushes nothing there either).

Checks: a measure dropped from the dense passes keeps its x and r bit for bit from the pass
it was dropped on (its residual moves only by its own pushes, and it had none left), and its
snapshot sums stay those of that pass; the measures still active converge as without
dropping, to the exact FP64 bounds and the invariant x + r = b + W x; the passes run at the
narrower widths (per-width launch counts); slots frozen while still pending (the global
rule) stay unchanged through dense and push passes; the refusals; and bin/excent's
--drop-measures on / off end within the same bounds."""
import numpy as np
import pytest

import excent_amd.native as nat
from oracle import oracle as orc
from tests.test_gpu_push import true_residual

pytestmark = pytest.mark.gpu


def katz_slots(idx, adj):
    """PageRank (eps 1e-4) in slot 0, Katz with alpha = 0.5 / lambda1 (eps 1e-5) in slot 1 and
    Katz with alpha = 0.05 / lambda1 (eps 1e-3) in slot 2: slot 2 contracts ten times faster
    with a looser eps, so the slots converge in the order 2, 1, 0 and each drop is a prefix."""
    lam = float(orc.fp64_reference(idx, adj, orc.ORC_EV, tol=1e-10, max_iter=2000)[3])
    kinds = [nat.EXC_PR, nat.EXC_KATZ, nat.EXC_KATZ]
    eps = [1e-4, 1e-5, 1e-3]
    scale = [np.float32(0.85), np.float32(0.5 / lam), np.float32(0.05 / lam)]
    inits = [orc.init(idx, orc.ORC_PR), orc.init(idx, orc.ORC_KATZ, 1.0), orc.init(idx, orc.ORC_KATZ, 1.0)]
    return kinds, eps, scale, [1, 0, 0], inits


def run_dropping(idx, adj, kinds, eps, scale, divdeg, inits, drop=True, max_passes=3000):
    """Resident dense-pass loop that narrows the passes to the measures still pending after
    each snapshot; returns the final state, the drops (slot -> pass, x, r, residual sum at the
    drop) and the statistics."""
    n, m = len(idx) - 1, len(adj)
    nm = len(kinds)
    eps32 = [np.float32(e) for e in eps]
    drops = {}
    with nat.Context(n, m, np.array([0], np.int32), kinds, eps, scale, divdeg) as ctx:
        ctx.upload_offsets(idx)
        for i, (x0, r0) in enumerate(inits):
            ctx.set_state(i, x=x0, r=r0)
        assert ctx.cache_chunks(m) == 0
        ctx.cache_upload(adj, 0)
        assert ctx.enable_pull(0.0) == 0
        cur = nm
        for p in range(max_passes):
            cnt, _ = ctx.snapshot()
            l1, mx = ctx.residual_l1(), ctx.residual_max()
            for i, (_, _, _, l0) in drops.items():   # a dropped slot's sums are those of its drop
                assert l1[i] == l0, (p, i)
            if cnt[0::2].sum() == 0:
                break
            act = max([i + 1 for i in range(nm) if np.float32(mx[i]) > eps32[i]] or [0])
            if drop and 1 <= act < cur:
                ctx.set_active_measures(act)
                for i in range(act, cur):
                    drops[i] = (p, ctx.read_state(i, 0), ctx.read_state(i, 1), l1[i])
                cur = act
            ctx.push_resident(0, n, 0, m, False)
            ctx.pass_sync()
        else:
            raise AssertionError("did not converge")
        xs = [ctx.read_state(i, 0) for i in range(nm)]
        rs = [ctx.read_state(i, 1) for i in range(nm)]
        st = ctx.stats()
    return xs, rs, drops, st, p


def check_results(idx, adj, xs, rs, eps, scale):
    n = len(idx) - 1
    b_pr = np.full(n, np.float64(np.float32(1) - np.float32(0.85)))
    ref, _, _ = orc.fp64_solve_b(idx, adj, b_pr)
    tr = true_residual(idx, adj, xs[0], b_pr, 0.85, True)
    assert np.abs(xs[0].astype(np.float64) - ref).sum() <= np.abs(tr).sum() / 0.15 * (1 + 1e-6) + 1e-6
    assert np.abs(tr - rs[0]).max() < 2e-4 * max(1.0, np.abs(xs[0]).max())
    for i in (1, 2):
        tr = true_residual(idx, adj, xs[i], np.ones(n), float(scale[i]), False)
        assert np.abs(tr - rs[i]).max() < 2e-4 * max(1.0, np.abs(xs[i]).max()), i
        kref = orc.fp64_reference(idx, adj, orc.ORC_KATZ, alpha=float(scale[i]))[0]
        # x* >= 1 per vertex and ||x - x*||_1 <= ||r||_1 / (1 - alpha lambda1), alpha lambda1 <= 0.5
        assert orc.compare(kref, xs[i])["l1_rel"] < 2.5 * eps[i], i
    for i in range(3):
        assert np.abs(rs[i]).max() <= np.float32(eps[i]), i


@pytest.mark.parametrize("tile", ["0", "512", "4096/1"])
def test_dropped_measures_stay_and_active_converge(rmat12, tile, monkeypatch):
    """Slots drop in the order 2, 1 (widths 3 -> 2 -> 1); each dropped slot's x and r are
    unchanged to the end; the results meet the FP64 bounds; launches ran at every width.
    tile 512: the source-sorted pull at every width (EXC_PULL_TILE); 4096/1: at width 1
    only; else k_pull."""
    print("case dropout tile", tile, flush=True)
    monkeypatch.setenv("EXC_PULL_TILE", tile)
    idx, adj = orc.egr_load(rmat12)
    kinds, eps, scale, divdeg, inits = katz_slots(idx, adj)
    xs, rs, drops, st, passes = run_dropping(idx, adj, kinds, eps, scale, divdeg, inits)
    assert sorted(drops) == [1, 2] and drops[2][0] <= drops[1][0], {k: v[0] for k, v in drops.items()}
    for i, (p, x0, r0, _) in drops.items():
        assert np.array_equal(xs[i], x0) and np.array_equal(rs[i], r0), (i, p)
    check_results(idx, adj, xs, rs, eps, scale)
    w = st["pull_w_launches"]
    assert w[2] > 0 and w[0] > 0, w
    assert sum(w) == st["pull_passes"]
    # bytes per launch shrink with the width (rows of 3, 2, 1 floats)
    per = [st["pull_w_bytes"][k] / w[k] for k in range(3) if w[k]]
    assert all(b > 0 for b in per)
    if tile != "0":
        assert st["pull_w_tile"][0] == int(tile.split("/")[0])
    if tile.endswith("/1"):
        assert st["pull_w_tile"][1] == 0 and st["pull_w_tile"][2] == 0


def test_dropping_matches_full_width(rmat12):
    """The same problem with and without dropping: both within the FP64 bounds, and the
    active slot's values agree to within the run-to-run noise of the atomic sums."""
    idx, adj = orc.egr_load(rmat12)
    kinds, eps, scale, divdeg, inits = katz_slots(idx, adj)
    a = run_dropping(idx, adj, kinds, eps, scale, divdeg, inits, drop=True)
    b = run_dropping(idx, adj, kinds, eps, scale, divdeg, inits, drop=False)
    check_results(idx, adj, a[0], a[1], eps, scale)
    check_results(idx, adj, b[0], b[1], eps, scale)
    assert b[3]["pull_w_launches"][0] == 0 and b[3]["pull_w_launches"][1] == 0
    n = len(idx) - 1
    # both PageRank results lie within eps * n / (1 - d) of the fixed point, so of each other
    assert np.abs(a[0][0].astype(np.float64) - b[0][0]).sum() <= 2 * 1e-4 * n / 0.15


def test_freeze_pending_slots_and_refusals(rmat12):
    """Slots frozen while they still have vertices above eps (what the global rule does) stay
    bit for bit unchanged through dense and push passes, and resume when widened again; the
    call refuses before any snapshot, counts outside 1..nm, narrowing over a folded pass that
    holds the frozen slots' deltas, widening while a pass is folded, and batched contexts."""
    idx, adj = orc.egr_load(rmat12)
    n, m = len(idx) - 1, len(adj)
    kinds, eps, scale, divdeg, inits = katz_slots(idx, adj)
    with nat.Context(n, m, np.array([0], np.int32), kinds, eps, scale, divdeg) as ctx:
        ctx.upload_offsets(idx)
        for i, (x0, r0) in enumerate(inits):
            ctx.set_state(i, x=x0, r=r0)
        assert ctx.cache_chunks(m) == 0
        ctx.cache_upload(adj, 0)
        assert ctx.enable_pull(0.3) == 0
        with pytest.raises(nat.ExcError, match="no snapshot"):
            ctx.set_active_measures(2)
        for bad in (0, 4):
            with pytest.raises(nat.ExcError, match="1..3"):
                ctx.set_active_measures(bad)
        ctx.snapshot()
        ctx.set_active_measures(1)   # every slot is pending: slots 1 and 2 freeze as they are
        frozen = [(ctx.read_state(i, 0), ctx.read_state(i, 1)) for i in (1, 2)]
        for p in range(12):   # dense passes first (demand >= 0.3 m), then pushes
            ctx.push_resident(0, n, 0, m, p % 2 == 0)
            ctx.pass_sync()
            ctx.snapshot()
        for i, (x, r) in zip((1, 2), frozen):
            assert np.array_equal(ctx.read_state(i, 0), x) and np.array_equal(ctx.read_state(i, 1), r), i
        assert ctx.stats()["pull_w_launches"][0] > 0
        # widen again (a folded state must be continued by a dense pass first)
        while ctx.dense_required():
            ctx.push_resident(0, n, 0, m, False)
            ctx.pass_sync()
            ctx.snapshot()
        ctx.set_active_measures(3)
        # a dense pass at full demand folds the next pass over all three slots; narrowing now,
        # with slot 2 still above eps, would lose its folded deltas
        ctx.push_resident(0, n, 0, m, False)
        ctx.pass_sync()
        ctx.snapshot()
        if ctx.dense_required() and np.float32(ctx.residual_max()[2]) > np.float32(eps[2]):
            with pytest.raises(nat.ExcError, match="folded"):
                ctx.set_active_measures(2)
        # run to convergence at full width from here: the bounds hold for every slot
        for _ in range(3000):
            cnt, _ = ctx.snapshot()
            if cnt[0::2].sum() == 0:
                break
            ctx.push_resident(0, n, 0, m, False)
            ctx.pass_sync()
        xs = [ctx.read_state(i, 0) for i in range(3)]
        rs = [ctx.read_state(i, 1) for i in range(3)]
    check_results(idx, adj, xs, rs, eps, scale)
    with nat.Context(n, m, np.array([0], np.int32), [nat.EXC_PR], [1e-4], [0.85], [1], batch_k=4) as ctx:
        with pytest.raises(nat.ExcError, match="batched"):
            ctx.set_active_measures(1)


def test_cli_drop_measures(tmp_path, rmat14):
    """bin/excent --measure all with --drop-measures on and off (dense passes): both runs end
    with every |r| within eps and the PageRank error within its FP64 bound; with dropping,
    some dense passes ran narrower than three measures."""
    idx, adj = orc.egr_load(rmat14)
    n = len(idx) - 1
    lam = float(orc.fp64_reference(idx, adj, orc.ORC_EV, tol=1e-10, max_iter=2000)[3])
    out = {}
    for mode in ("on", "off"):
        print("case cli drop", mode, flush=True)
        args = [rmat14, "--measure", "all", "--alpha", 0.5 / lam, "--beta", 1, "--lambda", lam, "--eps-pr", 1e-3,
                "--eps-katz", 1e-5, "--eps-ev", 1e-4, "--ev-init", "indeg", "--cache", "chunks", "--preload",
                "--max-passes", 3000, "--drop-measures", mode, "--log", str(tmp_path / f"{mode}.log")]
        with nat.Session(args) as S:
            r = S.run(reinit=False)
            out[mode] = (r, [S.values(i, 0) for i in range(3)], [S.values(i, 1) for i in range(3)])
        assert r["stop_reason"] == 0, mode
        for i, e in enumerate((1e-3, 1e-5, 1e-4)):
            assert np.abs(out[mode][2][i]).max() <= np.float32(e), (mode, i)
    w_on, w_off = out["on"][0]["kernel"]["pull_w_launches"], out["off"][0]["kernel"]["pull_w_launches"]
    assert w_off[0] == w_off[1] == 0 and w_off[2] > 0
    assert w_on[0] + w_on[1] > 0, w_on
    b = np.full(n, np.float64(np.float32(1) - np.float32(0.85)))
    ref, _, _ = orc.fp64_solve_b(idx, adj, b)
    for mode in ("on", "off"):
        x = out[mode][1][0]
        tr = true_residual(idx, adj, x, b, 0.85, True)
        assert np.abs(x.astype(np.float64) - ref).sum() <= np.abs(tr).sum() / 0.15 * (1 + 1e-6) + 1e-6, mode
    # --stop katz,ev: Katz and the eigenvector freeze once their FP64 sums are within n eps
    args = [rmat14, "--measure", "all", "--alpha", 0.5 / lam, "--beta", 1, "--lambda", lam, "--eps-pr", 1e-3,
            "--eps-katz", 1e-6, "--eps-ev", 1e-6, "--ev-init", "indeg", "--cache", "chunks", "--preload",
            "--max-passes", 3000, "--stop", "katz,ev", "--log", str(tmp_path / "g.log")]
    with nat.Session(args) as S:
        r = S.run(reinit=False)
        R = [S.values(i, 1) for i in range(3)]
    assert r["stop_reason"] in (0, 1)
    assert np.abs(R[0]).max() <= np.float32(1e-3)
    for i in (1, 2):
        l1 = np.abs(R[i].astype(np.float64)).sum()
        assert np.abs(R[i]).max() <= np.float32(1e-6) or l1 <= n * float(np.float32(1e-6)), i
        assert r["residual_l1"][i] == pytest.approx(l1, rel=1e-12)
    assert r["kernel"]["pull_w_launches"][0] > 0
    assert nat.excent([rmat14, "--drop-measures", "maybe", "--log", str(tmp_path / "bad.log")]) == 1

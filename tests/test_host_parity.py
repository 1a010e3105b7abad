"""The product's host side (reader, chunk planner, fetch plan, hot set, tau, gather list,
prefetch counts) against the oracle -- and, where the reference's own StaticGraph reader
was compiled into oracle/_ref, against the reference code itself.  CPU only."""
import ctypes
import os

import numpy as np
import pytest

import excent_amd.native as nat
from oracle import oracle as orc
from tests.conftest import SAMPLE_EGR

ref = orc.ref()
needs_ref = pytest.mark.skipif(ref is None, reason="reference tree absent: oracle/_ref not built")


def product_chunks(path, cap):
    n = 1 << 22
    v0 = np.zeros(n, np.int32); v1 = np.zeros(n, np.int32)
    e0 = np.zeros(n, np.uint64); e1 = np.zeros(n, np.uint64)
    nc = nat.F["excent_plan_chunks"](path.encode(), cap, nat.ptr(v0), nat.ptr(v1), nat.ptr(e0), nat.ptr(e1), n)
    assert nc > 0
    return v0[:nc], v1[:nc], e0[:nc], e1[:nc]


def ref_chunks(path, cap):
    n = 1 << 22
    v0 = np.zeros(n, np.int32); v1 = np.zeros(n, np.int32)
    e0 = np.zeros(n, np.uint64); e1 = np.zeros(n, np.uint64)
    nc = ref.ref_plan_chunks(path.encode(), cap, nat.ptr(v0), nat.ptr(v1), nat.ptr(e0), nat.ptr(e1), n)
    return v0[:nc], v1[:nc], e0[:nc], e1[:nc]


@pytest.fixture(scope="module", params=["sample", "rmat12", "rmat12_narrow"])
def gpath(request, rmat12, tmp_path_factory):
    if request.param == "sample":
        return SAMPLE_EGR
    if request.param == "rmat12":
        return rmat12
    idx, adj = orc.egr_load(rmat12)
    p = str(tmp_path_factory.mktemp("narrow") / "r12n.egr")
    orc.egr_write(p, idx, adj, 0)
    return p


def test_reader_matches_oracle(gpath):
    idx, adj = orc.egr_load(gpath)
    pidx, padj = nat.load_egr(gpath)
    assert np.array_equal(idx, pidx) and np.array_equal(adj, padj)


@needs_ref
def test_reader_matches_reference(gpath):
    idx, adj = orc.egr_load(gpath)
    n, m = ctypes.c_int(0), ctypes.c_uint64(0)
    out = np.zeros(len(idx), np.uint64)
    for in_ram in (0, 1):
        assert ref.ref_open(gpath.encode(), in_ram, ctypes.byref(n), ctypes.byref(m), nat.ptr(out), len(out)) == 0
        assert n.value == len(idx) - 1 and m.value == len(adj) and np.array_equal(out, idx)
    e = np.zeros(len(adj), np.int32)
    ref.ref_read_edges(gpath.encode(), 0, len(adj), nat.ptr(e))
    assert np.array_equal(e, adj)


@pytest.mark.parametrize("cap", [1, 50, 1000, 4096, 100000, 4000000])
def test_chunk_planner_matches_oracle(gpath, cap):
    idx, _ = orc.egr_load(gpath)
    v0, v1, e0, e1 = product_chunks(gpath, cap)
    assert np.array_equal(v0, orc.plan_chunks(idx, cap))
    assert np.array_equal(e0, idx[v0]) and np.array_equal(e1, idx[v1])


@needs_ref
@pytest.mark.parametrize("cap", [1, 50, 1000, 4096, 100000, 4000000])
def test_chunk_planner_matches_reference(gpath, cap):
    a = product_chunks(gpath, cap)
    b = ref_chunks(gpath, cap)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_chunk_cap_matches_oracle():
    for m in (0, 1, 16000, 10 ** 6, 268435456, 4 * 10 ** 9, 17 * 10 ** 9):
        for cm in (0, 0.02, 3.5, 64):
            for tc in (16, 3):
                for mc in (4, 0.5):
                    assert nat.F["excent_chunk_cap"](m, cm, tc, mc) == orc.chunk_cap(m, cm, tc, mc)


def test_fetch_plan_matches_oracle():
    rng = np.random.default_rng(0)
    for trial in range(300):
        nc = int(rng.integers(1, 40))
        ce = rng.integers(0, 10 ** 6, nc).astype(np.uint64)
        cv = rng.integers(0, 3, nc) * rng.integers(0, 1000, nc)
        cnt = np.zeros(2 * nc, np.uint64)
        cnt[0::2] = cv
        cnt[1::2] = (rng.random(nc) * ce * (cv > 0)).astype(np.uint64)
        res = rng.integers(0, 2, nc).astype(np.uint8)
        ld = (res & rng.integers(0, 2, nc)).astype(np.uint8)
        hlo = rng.integers(0, 5, nc).astype(np.int32)
        hhi = (hlo + rng.integers(0, 2, nc)).astype(np.int32)
        tau = float(rng.choice([0.0, 0.005, 0.1, 0.37, 1.0, 2.0]))
        hm, ps = int(rng.integers(0, 2)), int(rng.integers(0, 3))
        mode = np.zeros(nc, np.int8)
        pv, de = ctypes.c_uint64(0), ctypes.c_uint64(0)
        need = nat.F["excent_fetch_plan"](nc, nat.ptr(cnt), nat.ptr(res), nat.ptr(ld), tau, nat.ptr(ce), hm, ps,
                                          nat.ptr(hlo), nat.ptr(hhi), nat.ptr(mode), ctypes.byref(pv), ctypes.byref(de))
        omode, oneed, opv, ode = orc.fetch_plan(cnt, res, ld, tau, ce, hm, ps, hlo, hhi)
        assert np.array_equal(mode, omode) and bool(need) == oneed and (pv.value, de.value) == (opv, ode)


def test_pick_group_matches_oracle():
    for avg in np.concatenate([np.linspace(0, 70, 701), [1e3, 1e9]]):
        assert nat.F["excent_pick_group"](float(avg), 32) == orc.pick_group(float(avg))
    assert nat.F["excent_pick_group"](50.0, 64) == 64


def test_tau_matches_oracle():
    rng = np.random.default_rng(1)
    for _ in range(500):
        ts = float(rng.random() * 2)
        tg = np.sort(rng.random(3) * 3)
        fg = np.sort(rng.random(3))
        a = nat.F["excent_tau"](ts, nat.ptr(tg), nat.ptr(fg))
        assert a == orc.tau(ts, tg, fg)


@pytest.mark.parametrize("seed", range(5))
def test_hot_select_matches_oracle(seed, rmat12):
    rng = np.random.default_rng(seed)
    idx, _ = orc.egr_load(rmat12) if seed < 3 else (np.concatenate([[0], np.cumsum(
        np.where(rng.random(5000) < 0.01, rng.integers(1 << 20, 1 << 21, 5000), rng.integers(0, 30, 5000)))]).astype(np.uint64), None)
    n = len(idx) - 1
    total = int(4 * idx[-1] + 16 * n)
    for frac in (0.0, 0.01, 0.3, 0.9, 1.5):
        budget = int(total * frac)
        score = rng.integers(0, 6, n).astype(np.uint16) if seed % 2 else None
        ids = np.zeros(n, np.int32); off = np.zeros(n + 1, np.uint64); bm = np.zeros((n + 31) // 32, np.uint32)
        dm = ctypes.c_uint64(0)
        nh = nat.F["excent_hot_select"](n, nat.ptr(idx), budget, nat.ptr(score), nat.ptr(ids), nat.ptr(off), nat.ptr(bm),
                                        ctypes.byref(dm))
        oids, ooff, obm, odm = orc.hot_select(idx, budget, score)
        assert np.array_equal(ids[:nh], oids) and np.array_equal(off[: nh + 1], ooff)
        assert np.array_equal(bm, obm) and dm.value == odm


def test_gather_list_matches_numpy():
    rng = np.random.default_rng(2)
    n = 10000
    bm = rng.integers(0, 2 ** 32, (n + 31) // 32, dtype=np.uint64).astype(np.uint32)
    bm[3] = 0
    bits = ((bm[np.arange(n) >> 5] >> (np.arange(n) & 31)) & 1).astype(bool)
    out = np.zeros(n, np.int32)
    for v0, v1 in [(0, n), (5, 77), (96, 97), (31, 33), (1000, 9999)]:
        k = nat.F["excent_gather_list"](nat.ptr(bm), v0, v1, nat.ptr(out))
        want = np.nonzero(bits[v0:v1])[0] + v0
        assert np.array_equal(out[:k], want)


@pytest.mark.parametrize("k", [5, 2000, 20000])
def test_prefetch_counts_product_oracle_reference(rmat12, k, tmp_path):
    idx, adj = orc.egr_load(rmat12)
    n = len(idx) - 1
    rng = np.random.default_rng(k)
    vids = np.sort(rng.choice(n, size=min(k, n), replace=False)).astype(np.int32)
    pg, rn = ctypes.c_uint64(0), ctypes.c_uint64(0)
    assert nat.F["excent_prefetch_counts"](rmat12.encode(), nat.ptr(vids), len(vids), 16, ctypes.byref(pg), ctypes.byref(rn)) == 0
    header = 8 + (8 if open(rmat12, "rb").read(8)[4:8] == b"\xff\xff\xff\xff" else 0)
    wide = header == 16
    nlist_off = header + (8 if wide else 4) * (n + 1)
    page = os.sysconf("SC_PAGESIZE")
    assert (pg.value, rn.value) == orc.prefetch_counts(vids, idx, nlist_off, page)
    if ref is not None:
        a, b = ctypes.c_uint64(0), ctypes.c_uint64(0)
        ref.ref_prefetch_counts(rmat12.encode(), nat.ptr(vids), len(vids), 16, ctypes.byref(a), ctypes.byref(b))
        assert (a.value, b.value) == (pg.value, rn.value)


def _hot_cases():
    rng = np.random.default_rng(17)
    big = 1 << 20
    yield "ties", rng.integers(0, 6, 3000)                                     # few classes, many ties
    yield "big_ties", np.concatenate([np.full(4, big), np.full(3, big + 7), [3 * big], rng.integers(0, 40, 500)])
    yield "isolated", np.where(rng.random(2000) < 0.7, 0, rng.integers(1, 300, 2000))
    yield "spread", (rng.pareto(1.2, 4000) * 3).astype(np.int64)              # heavy tail, wide buckets
    yield "only_big", np.array([big, 0, big + 1, big, 0])
    yield "edgeless", np.zeros(100, np.int64)


@pytest.mark.parametrize("name,deg", list(_hot_cases()))
def test_hot_select_budget_sweep(name, deg):
    """The product's selection (weighted radix select + parallel compaction) against the
    oracle's restatement of the reference's rule on budgets at, just below and just above
    every class boundary (where the split class and its quota change), with and without
    ranked scores."""
    deg = np.asarray(deg, np.int64)
    n = len(deg)
    idx = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    cost = 4 * deg + 16
    cls = np.unique(deg[deg > 0])[::-1]
    bounds = np.cumsum([cost[deg == d].sum() for d in cls]) if len(cls) else np.zeros(0)
    budgets = {0, 1, 15, 16, 17, int(cost[deg > 0].sum()) if (deg > 0).any() else 0}
    for b in bounds[:60]:
        b = int(b)
        budgets.update({b - 1, b, b + 1, b + 20, max(0, b - 3000)})
    rng = np.random.default_rng(len(budgets))
    score = rng.integers(0, 4, n).astype(np.uint16)
    for budget in sorted(x for x in budgets if x >= 0):
        for sc in (None, score):
            ids = np.zeros(max(n, 1), np.int32); off = np.zeros(n + 1, np.uint64)
            bm = np.zeros((n + 31) // 32, np.uint32); dm = ctypes.c_uint64(0)
            nh = nat.F["excent_hot_select"](n, nat.ptr(idx), budget, nat.ptr(sc), nat.ptr(ids), nat.ptr(off), nat.ptr(bm),
                                            ctypes.byref(dm))
            oids, ooff, obm, odm = orc.hot_select(idx, budget, sc)
            assert np.array_equal(ids[:nh], oids) and np.array_equal(off[: nh + 1], ooff), (name, budget, sc is None)
            assert np.array_equal(bm, obm) and dm.value == odm, (name, budget, dm.value, odm)


@pytest.mark.parametrize("threads", [1, 3, 16])
def test_stage_lists_packs_the_lists(tmp_path, threads):
    """The gather's host staging: packed columns equal the concatenated lists (numpy), the
    offsets their prefix sums, and the readahead counts those of the oracle's and the
    reference reader's prefetch -- on a graph with empty lists and a hub list longer than a
    worker's share (the copy splits it)."""
    rng = np.random.default_rng(threads)
    n = 30000
    deg = np.where(rng.random(n) < 0.2, 0, rng.integers(1, 40, n))
    deg[12345] = 400000
    idx = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    adj = rng.integers(0, n, int(idx[-1])).astype(np.int32)
    path = str(tmp_path / "g.egr")
    orc.egr_write(path, idx, adj, 1)
    header = 16
    nlist_off = header + 8 * (n + 1)
    page = os.sysconf("SC_PAGESIZE")
    for k in (0, 1, 700, 5000, 20000):
        vids = np.sort(rng.choice(n, size=k, replace=False)).astype(np.int32)
        if k >= 700:
            vids = np.unique(np.concatenate([vids, [12345]])).astype(np.int32)
        k = len(vids)
        off = np.zeros(k + 1, np.uint64)
        out = np.zeros(max(1, int(deg[vids].sum())), np.int32)
        pg, rn = ctypes.c_uint64(0), ctypes.c_uint64(0)
        assert nat.F["excent_stage_lists"](path.encode(), nat.ptr(vids), k, 1, threads, nat.ptr(off), nat.ptr(out),
                                           ctypes.byref(pg), ctypes.byref(rn)) == 0
        want = np.concatenate([adj[idx[v]:idx[v + 1]] for v in vids]) if k else np.zeros(0, np.int32)
        assert np.array_equal(off[1:], np.cumsum(deg[vids])) and off[0] == 0
        assert np.array_equal(out[: len(want)], want)
        if k:
            assert (pg.value, rn.value) == orc.prefetch_counts(vids, idx, nlist_off, page)
            if ref is not None:
                a, b = ctypes.c_uint64(0), ctypes.c_uint64(0)
                ref.ref_prefetch_counts(path.encode(), nat.ptr(vids), len(vids), 16, ctypes.byref(a), ctypes.byref(b))
                assert (a.value, b.value) == (pg.value, rn.value)


def test_top_in_degree_exact_below_sample_floor(rmat12):
    """Graphs under 16M columns are read whole: the chosen destinations are the ones with
    the largest in-degrees (ties in any order)."""
    idx, adj = orc.egr_load(rmat12)
    indeg = np.bincount(adj, minlength=len(idx) - 1)
    for count in (1, 37, 1024):
        out = np.zeros(count, np.int32)
        k = nat.F["excent_top_in_degree"](rmat12.encode(), count, 1.0 / 16, 4, nat.ptr(out))
        assert k == min(count, int((indeg > 0).sum()))
        assert np.array_equal(np.sort(indeg[out[:k]])[::-1], np.sort(indeg)[::-1][:k])

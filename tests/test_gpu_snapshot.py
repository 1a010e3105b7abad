"""Demand snapshot on the GPU (k_snapshot, src/excent.cu:69-98) against the oracle, bit
for bit: bitmap words, per-chunk pending counts and degree sums, hot-pending count
and the --trace-pending counters, on seeded residual states.  Then the fetch plan
computed from the GPU's counts by the product planner against the oracle's plan."""
import numpy as np
import pytest

import excent_amd.native as nat
from oracle import oracle as orc
from tests.conftest import SAMPLE_EGR

pytestmark = pytest.mark.gpu

KIND_SETS = {1: [nat.EXC_PR], 2: [nat.EXC_PR, nat.EXC_KATZ], 3: [nat.EXC_PR, nat.EXC_KATZ, nat.EXC_EV]}


def states(n, nm, eps, kind, seed):
    rng = np.random.default_rng(seed)
    rs = []
    for k in range(nm):
        e = np.float32(eps[k])
        if kind == "uniform":
            r = (rng.standard_normal(n) * 2 * e).astype(np.float32)
        elif kind == "near":   # exactly eps, one ulp above and below, both signs
            choices = np.array([e, np.nextafter(e, np.float32(1)), np.nextafter(e, np.float32(0)),
                                -e, -np.nextafter(e, np.float32(1)), np.float32(0)], np.float32)
            r = choices[rng.integers(0, len(choices), n)]
        elif kind == "all":
            r = np.full(n, 1.0, np.float32)
        elif kind == "none":
            r = np.zeros(n, np.float32)
        else:   # sparse: few pending
            r = np.zeros(n, np.float32)
            pick = rng.choice(n, size=max(1, n // 97), replace=False)
            r[pick] = -3 * e
        rs.append(r)
    return rs


def chunk_starts(idx, parts):
    n = len(idx) - 1
    if parts == 1:
        return np.array([0], np.int32)
    cap = max(1, int(idx[-1]) // parts)
    return orc.plan_chunks(idx, cap)


@pytest.fixture(scope="module", params=["sample", "rmat12"])
def graph(request, rmat12):
    path = SAMPLE_EGR if request.param == "sample" else rmat12
    return orc.egr_load(path)


@pytest.mark.parametrize("nm", [1, 2, 3])
@pytest.mark.parametrize("kind", ["uniform", "near", "all", "none", "sparse"])
@pytest.mark.parametrize("parts", [1, 7])
def test_snapshot_bit_exact(graph, nm, kind, parts):
    idx, adj = graph
    n = len(idx) - 1
    eps = [0.01, 0.05, 0.003][:nm]
    cv0 = chunk_starts(idx, parts)
    rs = states(n, nm, eps, kind, seed=nm * 31 + parts)
    print(f"case n={n} nm={nm} kind={kind} chunks={len(cv0)}", flush=True)
    kinds = KIND_SETS[nm]
    with nat.Context(n, len(adj), cv0, kinds, eps, [0.85] * nm, [1] * nm) as ctx:
        ctx.upload_offsets(idx)
        for k in range(nm):
            ctx.set_state(k, x=np.zeros(n, np.float32), r=rs[k])
        cnt, hp = ctx.snapshot()
        bm = ctx.bitmap()
    obm, ocnt, ohp = orc.snapshot(rs, eps, idx, cv0)
    assert np.array_equal(bm, obm)
    assert np.array_equal(cnt, ocnt)
    assert hp == 0 and ohp == 0


@pytest.mark.parametrize("nm", [1, 3])
def test_snapshot_hot_skip_and_trace(graph, nm):
    idx, adj = graph
    n = len(idx) - 1
    eps = [0.01, 0.05, 0.003][:nm]
    cv0 = chunk_starts(idx, 5)
    deg = np.diff(idx.astype(np.int64))
    budget = int(4 * deg.sum() * 0.2)   # about a fifth of the edges as hot lists
    ids, off, bitmask, _ = orc.hot_select(idx, budget)
    assert len(ids) > 0
    with nat.Context(n, len(adj), cv0, KIND_SETS[nm], eps, [0.85] * nm, [1] * nm, trace=True) as ctx:
        ctx.upload_offsets(idx)
        ctx.cache_hot(ids, off, bitmask)
        trace = np.zeros(n, np.uint16)
        for it in range(3):
            rs = states(n, nm, eps, "uniform", seed=100 + it)
            for k in range(nm):
                ctx.set_state(k, r=rs[k])
            hot_skip = it > 0
            cnt, hp = ctx.snapshot(hot_skip=hot_skip)
            bm = ctx.bitmap()
            obm, ocnt, ohp = orc.snapshot(rs, eps, idx, cv0, hot=bitmask if hot_skip else None, trace=trace)
            assert np.array_equal(bm, obm)
            assert np.array_equal(cnt, ocnt)
            assert hp == ohp
            if hot_skip:
                assert hp > 0
        assert np.array_equal(ctx.trace(), trace)


def test_plan_from_gpu_counts(graph):
    """Counts from the GPU feed the product planner; its plan equals the oracle's."""
    import ctypes
    idx, adj = graph
    n = len(idx) - 1
    cv0 = chunk_starts(idx, 7)
    nc = len(cv0)
    bounds = np.append(cv0, n)
    chunk_edges = np.array([idx[bounds[c + 1]] - idx[bounds[c]] for c in range(nc)], np.uint64)
    rs = states(n, 1, [0.01], "sparse", seed=5)
    with nat.Context(n, len(adj), cv0, [nat.EXC_PR], [0.01], [0.85], [1]) as ctx:
        ctx.upload_offsets(idx)
        ctx.set_state(0, r=rs[0])
        cnt, _ = ctx.snapshot()
    rng = np.random.default_rng(11)
    for trial in range(20):
        resident = rng.integers(0, 2, nc).astype(np.uint8)
        loaded = (resident & rng.integers(0, 2, nc)).astype(np.uint8)
        tau = float(rng.choice([0.0, 0.01, 0.05, 0.3, 1.0, 2.0]))
        hot_mode = int(rng.integers(0, 2))
        pass_ = int(rng.integers(0, 2))
        hlo = np.zeros(nc, np.int32)
        hhi = rng.integers(0, 2, nc).astype(np.int32)
        mode = np.zeros(nc, np.int8)
        pv, de = ctypes.c_uint64(0), ctypes.c_uint64(0)
        need = nat.F["excent_fetch_plan"](nc, nat.ptr(cnt), nat.ptr(resident), nat.ptr(loaded), tau,
                                          nat.ptr(chunk_edges), hot_mode, pass_, nat.ptr(hlo), nat.ptr(hhi),
                                          nat.ptr(mode), ctypes.byref(pv), ctypes.byref(de))
        omode, oneed, opv, ode = orc.fetch_plan(cnt, resident, loaded, tau, chunk_edges, hot_mode, pass_, hlo, hhi)
        assert np.array_equal(mode, omode) and bool(need) == oneed and pv.value == opv and de.value == ode

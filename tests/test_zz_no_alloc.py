"""exc_snapshot_noalloc (the allocation-free demand snapshot, its scratch in a static device
array) against exc_snapshot on a context, the entry point it stands for: bitmap, per-chunk
counts and hot-pending bit for bit, FP64 residual sums within FP64 rounding of the order of
their partial sums.  Two calls in a row, the second with a smaller graph, fewer chunks and
fewer measures (the scratch keeps the first call's contents, which the second must clear);
plus the argument checks.  Device buffers come from the HIP runtime the library itself
loaded (ctypes), so the test allocates, the entry point does not."""
import ctypes

import numpy as np
import pytest

import excent_amd.native as nat
from oracle import oracle as orc
from tests.conftest import SAMPLE_EGR

pytestmark = pytest.mark.gpu

hip = ctypes.CDLL("libamdhip64.so.7")   # the runtime libexcent_hip.so is linked against
H2D, D2H = 1, 2


class Dev:
    """Device copy of a numpy array (test plumbing)."""

    def __init__(self, a=None, nbytes=None):
        self.nbytes = int(a.nbytes if a is not None else nbytes)
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(max(self.nbytes, 4))) == 0
        if a is not None:
            a = np.ascontiguousarray(a)
            assert hip.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(self.nbytes), H2D) == 0

    def get(self, dtype, count):
        out = np.zeros(count, dtype)
        assert hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, ctypes.c_size_t(out.nbytes), D2H) == 0
        return out

    def free(self):
        hip.hipFree(self.p)


S_OF = {1: 1, 2: 2, 3: 4}


def case(idx, nm, nchunks, seed, hot_frac):
    n = len(idx) - 1
    rng = np.random.default_rng(seed)
    eps = (rng.random(nm) * 0.01 + 0.002).astype(np.float32)
    rs = [(rng.standard_normal(n) * 0.01 * (rng.random(n) < 0.4)).astype(np.float32) for _ in range(nm)]
    rs[0][: 5] = [eps[0], -eps[0], np.nextafter(eps[0], np.float32(1)), 0, -np.nextafter(eps[0], np.float32(1))]
    cv0 = np.unique(np.concatenate([[0], rng.integers(1, n, nchunks - 1)])).astype(np.int32)
    hotv = np.nonzero(rng.random(n) < hot_frac)[0].astype(np.int32) if hot_frac else np.zeros(0, np.int32)
    bm = np.zeros((n + 31) // 32, np.uint32)
    np.bitwise_or.at(bm, hotv >> 5, (np.uint32(1) << (hotv & 31).astype(np.uint32)))
    return eps, rs, cv0, hotv, bm


def with_context(idx, eps, rs, cv0, hotv, hotbm):
    """The existing entry point: a context's exc_snapshot (hot-skip when a hot set is given)."""
    n, nm = len(idx) - 1, len(rs)
    with nat.Context(n, int(idx[-1]), cv0, [nat.EXC_PR] * nm, list(eps), [0.85] * nm, [1] * nm) as ctx:
        ctx.upload_offsets(idx)
        for i in range(nm):
            ctx.set_state(i, x=np.zeros(n, np.float32), r=rs[i])
        if len(hotv):
            deg = np.diff(idx.astype(np.int64))[hotv]
            off = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
            ctx.cache_hot(hotv, off, hotbm)
        cnt, hp = ctx.snapshot(hot_skip=len(hotv) > 0)
        return ctx.bitmap(), cnt, hp, ctx.residual_l1()


def without_alloc(idx, eps, rs, cv0, hotv, hotbm):
    n, nm = len(idx) - 1, len(rs)
    S = S_OF[nm]
    rows = np.zeros((n, S), np.float32)
    for i in range(nm):
        rows[:, i] = rs[i]
    bufs = [Dev(rows), Dev(np.ascontiguousarray(idx, np.uint64)), Dev(cv0), Dev(hotbm),
            Dev(nbytes=4 * ((n + 31) // 32)), Dev(nbytes=8 * (2 * len(cv0) + 1)), Dev(nbytes=8 * nm)]
    d_r, d_idx, d_cv0, d_hot, d_bm, d_cnt, d_l1 = bufs
    try:
        # outputs start as garbage: every element must be written by the call
        assert hip.hipMemset(d_bm.p, 0xA5, ctypes.c_size_t(d_bm.nbytes)) == 0
        assert hip.hipMemset(d_cnt.p, 0xA5, ctypes.c_size_t(d_cnt.nbytes)) == 0
        rc = nat.F["exc_snapshot_noalloc"](0, None, n, nm, d_r.p, nat.ptr(eps), d_idx.p, d_cv0.p, len(cv0),
                                           d_hot.p if len(hotv) else None, d_bm.p, d_cnt.p, d_l1.p)
        assert rc == 0, nat.F["exc_last_error"]().decode()
        assert hip.hipDeviceSynchronize() == 0
        cnt = d_cnt.get(np.uint64, 2 * len(cv0) + 1)
        return d_bm.get(np.uint32, (n + 31) // 32), cnt[:-1], int(cnt[-1]), list(d_l1.get(np.float64, nm))
    finally:
        for b in bufs:
            b.free()


def test_noalloc_snapshot_matches_exc_snapshot(rmat12):
    """Call 1: rmat12, three measures, 7 chunks, a hot set.  Call 2: the 2000-vertex sample,
    one measure, 2 chunks, no hot set."""
    for path, nm, nchunks, seed, hot in ((rmat12, 3, 7, 1, 0.05), (SAMPLE_EGR, 1, 2, 2, 0.0)):
        idx, _ = orc.egr_load(path)
        eps, rs, cv0, hotv, hotbm = case(idx, nm, nchunks, seed, hot)
        print(f"case noalloc n={len(idx) - 1} nm={nm} nc={len(cv0)} hot={len(hotv)}", flush=True)
        bm0, cnt0, hp0, l10 = with_context(idx, eps, rs, cv0, hotv, hotbm)
        bm1, cnt1, hp1, l11 = without_alloc(idx, eps, rs, cv0, hotv, hotbm)
        assert np.array_equal(bm0, bm1) and np.array_equal(cnt0, cnt1) and hp0 == hp1
        assert np.allclose(l11, l10, rtol=1e-12, atol=0)
        # and both equal the oracle's restatement
        obm, ocnt, ohp = orc.snapshot(rs, eps, idx, cv0, hot=hotbm if len(hotv) else None)
        assert np.array_equal(bm1, obm) and np.array_equal(cnt1, ocnt) and hp1 == ohp
        want = [np.abs(r.astype(np.float64)).sum() for r in rs]
        assert np.allclose(l11, want, rtol=1e-12, atol=0)
        assert hp1 > 0 if len(hotv) else hp1 == 0


def test_noalloc_snapshot_rejects_sizes():
    eps = np.full(3, 0.01, np.float32)
    dummy = Dev(nbytes=64)
    try:
        p = dummy.p
        for nm, nc in ((1, 4097), (1, 0), (4, 1), (0, 1)):
            rc = nat.F["exc_snapshot_noalloc"](0, None, 10, nm, p, nat.ptr(eps), p, p, nc, None, p, p, p)
            assert rc == nat.EXC_EINVAL, (nm, nc)
        rc = nat.F["exc_snapshot_noalloc"](0, None, 10, 1, p, nat.ptr(eps), p, p, 4097, None, p, p, p)
        assert "1 to 4096" in nat.F["exc_last_error"]().decode()
    finally:
        dummy.free()

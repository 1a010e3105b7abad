"""ctypes bindings of the oracle (oracle/liboracle.so) and of the reference's own
StaticGraph reader compiled into oracle/_ref (when it was built).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product never imports this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
P = ctypes.POINTER
c_int, c_u64, c_double, c_float, c_void_p = ctypes.c_int, ctypes.c_uint64, ctypes.c_double, ctypes.c_float, ctypes.c_void_p

ORC_PR, ORC_KATZ, ORC_EV = 0, 1, 2


def _p(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        sig = {
            "orc_egr_load": (c_int, [ctypes.c_char_p, P(c_int), P(c_u64), P(c_void_p), P(c_void_p)]),
            "orc_free": (None, [c_void_p]),
            "orc_egr_write": (c_int, [ctypes.c_char_p, c_int, c_u64, c_void_p, c_void_p, c_int]),
            "orc_chunk_cap": (c_u64, [c_u64, c_double, c_double, c_double]),
            "orc_plan_chunks": (c_int, [c_int, c_void_p, c_u64, c_void_p, c_int]),
            "orc_init": (None, [c_int, c_void_p, c_int, c_float, c_float, c_void_p, c_void_p]),
            "orc_snapshot": (None, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                                    c_void_p, c_void_p, P(c_u64), c_void_p]),
            "orc_fetch_plan": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_int, c_int,
                                       c_void_p, c_void_p, c_void_p, P(c_u64), P(c_u64)]),
            "orc_pick_group": (c_int, [c_double]),
            "orc_use_list": (c_int, [c_u64, c_int, c_double]),
            "orc_hot_select": (c_int, [c_int, c_void_p, c_u64, c_void_p, c_void_p, c_void_p, c_void_p, P(c_u64)]),
            "orc_prefetch_counts": (None, [c_void_p, c_int, c_void_p, c_u64, c_u64, c_int, P(c_u64), P(c_u64)]),
            "orc_tau": (c_double, [c_double, c_void_p, c_void_p]),
            "orc_run_push": (c_int, [c_int, c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_double, c_int, c_int, c_double, P(c_u64),
                                     P(c_double)]),
            "orc_fp64_reference": (c_int, [c_int, c_void_p, c_void_p, c_int, c_double, c_double, c_int, c_double,
                                           c_void_p, P(c_double), P(c_double)]),
            "orc_fp64_solve_b": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_double, c_int, c_int, c_double,
                                         c_void_p, P(c_double)]),
            "orc_fp64_solve_w": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_int, c_double,
                                         c_void_p, P(c_double)]),
            "orc_params": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_double, P(c_double), P(c_double),
                                   P(c_double), P(c_double), P(ctypes.c_long)]),
            "orc_compare": (None, [c_int, c_void_p, c_void_p, c_int, P(c_double), P(c_double), P(c_double), P(c_int)]),
            "orc_rank": (None, [c_int, c_void_p, c_void_p, c_int, c_void_p, P(c_double)]),
            "orc_residual_l1": (c_double, [c_int, c_void_p, c_void_p, c_int, c_double, c_double, c_void_p]),
        }
        for k, (res, args) in sig.items():
            f = getattr(L, k)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def ref():
    """The reference's include/StaticGraph.h compiled in place (None when not built)."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libref_staticgraph.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        R.ref_open.restype = c_int
        R.ref_open.argtypes = [ctypes.c_char_p, c_int, P(c_int), P(c_u64), c_void_p, ctypes.c_longlong]
        R.ref_plan_chunks.restype = c_int
        R.ref_plan_chunks.argtypes = [ctypes.c_char_p, c_u64, c_void_p, c_void_p, c_void_p, c_void_p, c_int]
        R.ref_prefetch_counts.restype = c_int
        R.ref_prefetch_counts.argtypes = [ctypes.c_char_p, c_void_p, c_int, c_int, P(c_u64), P(c_u64)]
        R.ref_read_edges.restype = c_int
        R.ref_read_edges.argtypes = [ctypes.c_char_p, c_u64, c_u64, c_void_p]
        _ref = R
    return _ref


# ------------------------------------------------------------------ wrappers

def egr_load(path):
    n, m, ip, ap = c_int(0), c_u64(0), c_void_p(), c_void_p()
    rc = lib().orc_egr_load(path.encode(), ctypes.byref(n), ctypes.byref(m), ctypes.byref(ip), ctypes.byref(ap))
    if rc != 0:
        raise ValueError(f"orc_egr_load({path}) = {rc}")
    idx = np.ctypeslib.as_array(ctypes.cast(ip, P(c_u64)), shape=(n.value + 1,)).copy()
    adj = np.ctypeslib.as_array(ctypes.cast(ap, P(ctypes.c_int32)), shape=(max(1, m.value),)).copy()[: m.value]
    lib().orc_free(ip)
    lib().orc_free(ap)
    return idx, adj


def egr_write(path, idx, adj, wide):
    idx = np.ascontiguousarray(idx, np.uint64)
    adj = np.ascontiguousarray(adj, np.int32)
    rc = lib().orc_egr_write(path.encode(), len(idx) - 1, int(idx[-1]), _p(idx), _p(adj), int(wide))
    assert rc == 0


def chunk_cap(m, chunk_medges=0.0, target_chunks=16.0, min_chunk_medges=4.0):
    return lib().orc_chunk_cap(m, chunk_medges, target_chunks, min_chunk_medges)


def plan_chunks(idx, cap):
    n = len(idx) - 1
    out = np.zeros(n + 1, dtype=np.int32)
    nc = lib().orc_plan_chunks(n, _p(np.ascontiguousarray(idx, np.uint64)), cap, _p(out), n + 1)
    return out[:nc].copy()


def init(idx, kind, beta=1.0, lam=1.0):
    n = len(idx) - 1
    x = np.zeros(n, np.float32)
    r = np.zeros(n, np.float32)
    lib().orc_init(n, _p(np.ascontiguousarray(idx, np.uint64)), kind, beta, lam, _p(x), _p(r))
    return x, r


def snapshot(rs, eps, idx, cv0, hot=None, trace=None):
    """rs: list of float32[n] residual arrays.  Returns bitmap, cnt[2nc], hotpend."""
    n = len(idx) - 1
    nm = len(rs)
    rs = [np.ascontiguousarray(r, np.float32) for r in rs]
    rp = (c_void_p * nm)(*[r.ctypes.data for r in rs])
    eps_a = np.asarray(eps, np.float32)
    cv0 = np.ascontiguousarray(cv0, np.int32)
    bitmap = np.zeros((n + 31) // 32, np.uint32)
    cnt = np.zeros(2 * len(cv0), np.uint64)
    hp = c_u64(0)
    hot = None if hot is None else np.ascontiguousarray(hot, np.uint32)
    lib().orc_snapshot(n, nm, ctypes.cast(rp, c_void_p), _p(eps_a), _p(np.ascontiguousarray(idx, np.uint64)),
                       _p(cv0), len(cv0), _p(hot), _p(bitmap), _p(cnt), ctypes.byref(hp), _p(trace))
    return bitmap, cnt, hp.value


def fetch_plan(cnt, resident, loaded, tau, chunk_edges, hot_mode, pass_, hlo, hhi):
    nc = len(chunk_edges)
    mode = np.zeros(nc, np.int8)
    pv, de = c_u64(0), c_u64(0)
    need = lib().orc_fetch_plan(nc, _p(np.ascontiguousarray(cnt, np.uint64)),
                                _p(np.ascontiguousarray(resident, np.uint8)), _p(np.ascontiguousarray(loaded, np.uint8)),
                                tau, _p(np.ascontiguousarray(chunk_edges, np.uint64)), int(hot_mode), pass_,
                                _p(np.ascontiguousarray(hlo, np.int32)), _p(np.ascontiguousarray(hhi, np.int32)),
                                _p(mode), ctypes.byref(pv), ctypes.byref(de))
    return mode, bool(need), pv.value, de.value


def pick_group(avg):
    return lib().orc_pick_group(avg)


def use_list(k, nv, frac):
    return bool(lib().orc_use_list(k, nv, frac))


def hot_select(idx, budget, score=None):
    n = len(idx) - 1
    ids = np.zeros(n, np.int32)
    off = np.zeros(n + 1, np.uint64)
    bm = np.zeros((n + 31) // 32, np.uint32)
    dm = c_u64(0)
    sc = None if score is None else np.ascontiguousarray(score, np.uint16)
    nh = lib().orc_hot_select(n, _p(np.ascontiguousarray(idx, np.uint64)), budget, _p(sc), _p(ids), _p(off), _p(bm),
                              ctypes.byref(dm))
    return ids[:nh].copy(), off[: nh + 1].copy(), bm, dm.value


def prefetch_counts(vids, idx, nlist_off, page, io_threads=16):
    vids = np.ascontiguousarray(vids, np.int32)
    pg, rn = c_u64(0), c_u64(0)
    lib().orc_prefetch_counts(_p(vids), len(vids), _p(np.ascontiguousarray(idx, np.uint64)), nlist_off, page,
                              io_threads, ctypes.byref(pg), ctypes.byref(rn))
    return pg.value, rn.value


def tau(t_stream, t_g, f_got):
    return lib().orc_tau(t_stream, _p(np.asarray(t_g, np.float64)), _p(np.asarray(f_got, np.float64)))


def run_push(idx, adj, cv0, xs, rs, eps, scale, divdeg, list_frac=0.5, max_passes=100000, threads=0,
             time_limit=0.0):
    """CPU asynchronous delta push in place on xs/rs (lists of float32[n])."""
    n = len(idx) - 1
    nm = len(xs)
    xp = (c_void_p * nm)(*[x.ctypes.data for x in xs])
    rp = (c_void_p * nm)(*[r.ctypes.data for r in rs])
    pushed, secs = c_u64(0), c_double(0)
    passes = lib().orc_run_push(n, _p(np.ascontiguousarray(idx, np.uint64)), _p(np.ascontiguousarray(adj, np.int32)),
                                len(cv0), _p(np.ascontiguousarray(cv0, np.int32)), nm, ctypes.cast(xp, c_void_p),
                                ctypes.cast(rp, c_void_p), _p(np.asarray(eps, np.float32)),
                                _p(np.asarray(scale, np.float32)), _p(np.asarray(divdeg, np.int32)), list_frac,
                                max_passes, threads, time_limit, ctypes.byref(pushed), ctypes.byref(secs))
    return passes, pushed.value, secs.value


def fp64_reference(idx, adj, kind, alpha=-1.0, beta=1.0, max_iter=20000, tol=1e-9):
    n = len(idx) - 1
    out = np.zeros(n, np.float64)
    rel, lam = c_double(0), c_double(0)
    it = lib().orc_fp64_reference(n, _p(np.ascontiguousarray(idx, np.uint64)), _p(np.ascontiguousarray(adj, np.int32)),
                                  kind, alpha, beta, max_iter, tol, _p(out), ctypes.byref(rel), ctypes.byref(lam))
    return out, it, rel.value, lam.value


def fp64_solve_b(idx, adj, b, s=0.85, divdeg=1, max_iter=20000, tol=1e-12):
    n = len(idx) - 1
    out = np.zeros(n, np.float64)
    rel = c_double(0)
    it = lib().orc_fp64_solve_b(n, _p(np.ascontiguousarray(idx, np.uint64)), _p(np.ascontiguousarray(adj, np.int32)),
                                _p(np.ascontiguousarray(b, np.float64)), s, divdeg, max_iter, tol, _p(out),
                                ctypes.byref(rel))
    return out, it, rel.value


def fp64_solve_w(idx, adj, w, b, s=0.85, max_iter=100000, tol=1e-14):
    n = len(idx) - 1
    out = np.zeros(n, np.float64)
    rel = c_double(0)
    it = lib().orc_fp64_solve_w(n, _p(np.ascontiguousarray(idx, np.uint64)), _p(np.ascontiguousarray(adj, np.int32)),
                                _p(np.ascontiguousarray(w, np.float64)), _p(np.ascontiguousarray(b, np.float64)), s,
                                max_iter, tol, _p(out), ctypes.byref(rel))
    return out, it, rel.value


def params(idx, adj, afac=0.85, start=None):
    n = len(idx) - 1
    lam, al, mk, me = c_double(0), c_double(0), c_double(0), c_double(0)
    bad = ctypes.c_long(0)
    st = None if start is None else np.ascontiguousarray(start, np.float64)
    it = lib().orc_params(n, _p(np.ascontiguousarray(idx, np.uint64)), _p(np.ascontiguousarray(adj, np.int32)),
                          _p(st), afac, ctypes.byref(lam), ctypes.byref(al), ctypes.byref(mk), ctypes.byref(me),
                          ctypes.byref(bad))
    return dict(lam=lam.value, alpha=al.value, mean_katz=mk.value, mean_ev=me.value, asym=bad.value, iters=it)


def compare(ref_vals, x, unit=False):
    n = len(ref_vals)
    l1, l2, mr, top = c_double(0), c_double(0), c_double(0), c_int(0)
    lib().orc_compare(n, _p(np.ascontiguousarray(ref_vals, np.float64)), _p(np.ascontiguousarray(x, np.float32)),
                      int(unit), ctypes.byref(l1), ctypes.byref(l2), ctypes.byref(mr), ctypes.byref(top))
    return dict(l1_rel=l1.value, l2=l2.value, max_rel=mr.value, top100=top.value)


def rank(ref_vals, x, unit=False):
    n = len(ref_vals)
    p = np.zeros(5, np.float64)
    t = c_double(0)
    lib().orc_rank(n, _p(np.ascontiguousarray(ref_vals, np.float64)), _p(np.ascontiguousarray(x, np.float32)),
                   int(unit), _p(p), ctypes.byref(t))
    return p, t.value


def residual_l1(idx, adj, kind, scale, beta, x):
    n = len(idx) - 1
    return lib().orc_residual_l1(n, _p(np.ascontiguousarray(idx, np.uint64)), _p(np.ascontiguousarray(adj, np.int32)),
                                 kind, scale, beta, _p(np.ascontiguousarray(x, np.float32)))


def bf16_round(a):
    """float32 -> the nearest bfloat16 (ties to even), returned as float32: the high half of
    the float's bits after adding 0x7fff plus the lowest kept bit."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def dense_pass_rows(idx, adj, x, r, eps, damp=0.85, rows="bf16", feedback=True):
    """One dense personalised-PageRank pass (fold of every pending vertex, then every vertex
    adds the weights of its in-neighbours) on one column, in place on float32 x and r.

    The delta push of src/excent.cu:100-124 (pending iff |r| > eps, x += d, r = 0, weight
    damp d / deg to each out-neighbour, nothing pushed for deg = 0) with all folds before all
    adds.  rows="bf16": the weight is stored in bfloat16 and, with feedback, the fold moves
    only the delta the stored weight accounts for, bf16(w) deg / damp, into x and leaves the
    rounding remainder in r; feedback=False moves all of d (the remainder is lost).  Returns
    the number of pending vertices folded."""
    deg = np.diff(np.asarray(idx, np.int64))
    n = len(deg)
    damp = np.float32(damp)
    pend = np.abs(r) > np.float32(eps)
    d = np.where(pend, r, np.float32(0)).astype(np.float32)
    nz = pend & (deg > 0)
    w = np.zeros(n, np.float32)
    w[nz] = damp * d[nz] / deg[nz].astype(np.float32)
    moved = d.copy()
    if rows == "bf16":
        w = bf16_round(w)
        if feedback:
            moved[nz] = w[nz] * deg[nz].astype(np.float32) / damp
    x += moved
    r -= moved
    acc = np.zeros(n, np.float32)
    np.add.at(acc, np.asarray(adj), w[np.repeat(np.arange(n), deg)])
    r += acc
    return int(pend.sum())

"""ctypes bindings of the ExCent C ABIs (include/excent_hip.h, include/excent.h).

The libraries are built in-tree (excent_amd/lib) by ``__graft_entry__.build()``.
Importing this module fails loudly when they are missing: there is no fallback
path.  Device memory, streams and kernels are owned by the HIP library; Python
only passes host numpy buffers and sizes.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(HERE, "lib")
BINDIR = os.path.join(ROOT, "bin")

EXC_PR, EXC_KATZ, EXC_EV = 0, 1, 2
KERNEL_LB, KERNEL_GROUP, KERNEL_TUNE = 0, 1, 2
EXC_OK, EXC_EINVAL, EXC_ENOMEM, EXC_EHIP, EXC_ESTATE = 0, -1, -2, -3, -4


class NativeMissing(RuntimeError):
    pass


class ExcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _load(name):
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise NativeMissing(f"{path} is not built; run __graft_entry__.build() first")
    return ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


hip = _load("libexcent_hip.so")
host = _load("libexcent.so")

c_int, c_u64, c_double, c_float, c_void_p = ctypes.c_int, ctypes.c_uint64, ctypes.c_double, ctypes.c_float, ctypes.c_void_p
P = ctypes.POINTER


class Measures(ctypes.Structure):
    _fields_ = [("nm", c_int), ("kind", c_int * 3), ("eps", c_float * 3), ("scale", c_float * 3), ("divdeg", c_int * 3)]


class Config(ctypes.Structure):
    _fields_ = [("device", c_int), ("n", c_int), ("m", c_u64), ("nc", c_int), ("cv0", P(c_int)),
                ("max_chunk_edges", c_u64), ("max_chunk_verts", c_int), ("trace", c_int), ("zero_copy", c_int),
                ("push_kernel", c_int), ("timing", c_int), ("batch_k", c_int)]


class Chunk(ctypes.Structure):
    _fields_ = [("v0", c_int), ("v1", c_int), ("e0", c_u64), ("e1", c_u64)]


class Stats(ctypes.Structure):
    _fields_ = [("snapshot_ms", c_double), ("snapshot_launches", c_u64), ("snapshot_bytes", c_double),
                ("push_ms", c_double), ("push_launches", c_u64), ("push_bytes", c_double),
                ("piece_ms", c_double), ("piece_launches", c_u64), ("piece_bytes", c_double),
                ("other_ms", c_double), ("other_launches", c_u64),
                ("pushed_vertices", c_u64), ("pushed_edges", c_u64),
                ("pushed_edge_measures", c_u64), ("pushed_vertex_measures", c_u64),
                ("src_ms", c_double * 4), ("src_launches", c_u64 * 4), ("src_bytes", c_double * 4),
                ("fold_ms", c_double), ("pull_ms", c_double), ("pull_hub_ms", c_double), ("pull_passes", c_u64),
                ("fold_bytes", c_double), ("pull_bytes", c_double), ("pull_hub_bytes", c_double),
                ("pull_rows_in_flight", c_int), ("pull_tune_ns_per_edge", c_double * 2),
                ("pull_tile", c_int), ("pull_w_ms", c_double * 3), ("pull_w_launches", c_u64 * 3),
                ("pull_w_bytes", c_double * 3), ("pull_w_tile", c_int * 3), ("burst_graph_launches", c_u64),
                ("push_tune_ns_edge", (c_double * 3) * 7), ("push_tune_best", c_int * 7),
                ("front_snapshots", c_u64)]


class RunStats(ctypes.Structure):
    _fields_ = [("time_passes_s", c_double), ("time_snapshot_s", c_double), ("time_io_s", c_double),
                ("time_wait_s", c_double), ("passes", c_int), ("pushed_edges", c_u64), ("bytes_stream", c_u64),
                ("bytes_gather", c_u64), ("launches", c_u64), ("kernel", Stats),
                ("residual_l1", c_double * 3), ("stop_reason", c_int),
                ("time_open_s", c_double), ("time_setup_s", c_double), ("time_transpose_s", c_double)]


def stats_dict(s):
    out = {}
    for k, _ in Stats._fields_:
        v = getattr(s, k)
        if isinstance(v, (int, float)):
            out[k] = v
        else:   # arrays, and arrays of arrays
            out[k] = [list(e) if not isinstance(e, (int, float)) else e for e in v]
    return out


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


def ptr(a):
    return a.ctypes.data_as(c_void_p) if a is not None else None


# include/excent_hip.h -- every entry point declared there (checked by the tests)
HIP_SIGNATURES = {
    "exc_last_error": (ctypes.c_char_p, []),
    "exc_device_count": (c_int, [P(c_int)]),
    "exc_mem_info": (c_int, [c_int, P(ctypes.c_size_t), P(ctypes.c_size_t)]),
    "exc_device_info": (c_int, [c_int, P(c_int), P(c_int), ctypes.c_char_p, c_int]),
    "exc_hold_memory": (c_int, [c_int, ctypes.c_size_t, P(c_void_p), P(ctypes.c_size_t)]),
    "exc_release_memory": (c_int, [c_void_p]),
    "exc_device_synchronize": (c_int, [c_int]),
    "exc_create": (c_int, [P(Config), P(Measures), P(c_void_p)]),
    "exc_destroy": (c_int, [c_void_p]),
    "exc_upload_offsets": (c_int, [c_void_p, c_void_p, ctypes.c_size_t, ctypes.c_size_t]),
    "exc_init_measure": (c_int, [c_void_p, c_int, c_float, c_float]),
    "exc_set_state": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "exc_read_state": (c_int, [c_void_p, c_int, c_int, c_void_p, ctypes.c_size_t, ctypes.c_size_t]),
    "exc_state_checkpoint": (c_int, [c_void_p, c_int]),
    "exc_snapshot": (c_int, [c_void_p, c_int, c_void_p, P(c_u64)]),
    "exc_bitmap_to_host": (c_int, [c_void_p, c_void_p]),
    "exc_read_trace": (c_int, [c_void_p, c_void_p, ctypes.c_size_t, ctypes.c_size_t]),
    "exc_cache_chunks": (c_int, [c_void_p, c_u64]),
    "exc_cache_upload": (c_int, [c_void_p, c_void_p, c_u64, c_u64]),
    "exc_cache_hot": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "exc_cache_hot_upload": (c_int, [c_void_p, c_void_p, c_u64, c_u64]),
    "exc_slot_buffers": (c_int, [c_void_p, c_int, P(c_void_p), P(c_void_p), P(c_void_p)]),
    "exc_slot_acquire": (c_int, [c_void_p, c_int, P(c_double)]),
    "exc_slot_copy_probe": (c_int, [c_void_p, c_int, c_u64, P(c_double)]),
    "exc_enable_pull": (c_int, [c_void_p, c_double]),
    "exc_set_edge_weights": (c_int, [c_void_p, c_void_p, c_u64]),
    "exc_push_resident": (c_int, [c_void_p, P(Chunk), c_int, c_int, c_int]),
    "exc_stream_chunk": (c_int, [c_void_p, c_int, P(Chunk), c_int, c_int, c_int, c_int, c_int, c_int]),
    "exc_gather_chunk": (c_int, [c_void_p, c_int, c_int, c_int]),
    "exc_push_hot": (c_int, [c_void_p, c_int, c_int, c_int, c_int]),
    "exc_batch_eps": (c_int, [c_void_p, c_void_p]),
    "exc_batch_seed": (c_int, [c_void_p, c_int, c_void_p, c_int, c_float]),
    "exc_batch_colpend": (c_int, [c_void_p, c_void_p]),
    "exc_batch_set_active": (c_int, [c_void_p, c_u64]),
    "exc_pass_sync": (c_int, [c_void_p, P(c_double)]),
    "exc_pushed": (c_int, [c_void_p, P(c_u64)]),
    "exc_get_stats": (c_int, [c_void_p, P(Stats)]),
    "exc_reset_stats": (c_int, [c_void_p]),
    "exc_indegree": (c_int, [c_void_p, c_void_p]),
    "exc_fp64_create": (c_int, [c_int, c_int, c_u64, c_void_p, c_void_p, P(c_void_p)]),
    "exc_fp64_destroy": (c_int, [c_void_p]),
    "exc_fp64_solve": (c_int, [c_void_p, c_int, c_double, c_double, c_int, c_double, c_void_p, c_void_p,
                               P(c_double), P(c_double), P(c_int)]),
    "exc_fp64_lambda": (c_int, [c_void_p, c_void_p, c_int, c_double, P(c_double), P(c_double), P(c_int)]),
    "exc_dense_required": (c_int, [c_void_p]),
    "exc_residual_l1": (c_int, [c_void_p, P(c_double)]),
    "exc_residual_max": (c_int, [c_void_p, P(c_float)]),
    "exc_fold_ahead": (c_int, [c_void_p, c_int]),
    "exc_set_active_measures": (c_int, [c_void_p, c_int]),
    "exc_run_passes": (c_int, [c_void_p, c_int, c_int, P(c_u64), P(c_double), P(ctypes.c_float)]),
    "exc_burst_submit": (c_int, [c_void_p, c_int, c_int]),
    "exc_burst_collect": (c_int, [c_void_p, P(c_u64), P(c_double), P(ctypes.c_float), P(c_int)]),
    "exc_set_hub_destinations": (c_int, [c_void_p, c_void_p, c_int]),
    "exc_snapshot_noalloc": (c_int, [c_int, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                     c_void_p, c_void_p, c_void_p, c_void_p]),
    "exc_fp64_residual": (c_int, [c_void_p, c_double, c_void_p, c_double, c_int, c_void_p, P(c_double)]),
    "exc_fp64_solve_vec": (c_int, [c_void_p, c_void_p, c_double, c_int, c_int, c_double, c_void_p, P(c_double),
                                   P(c_int)]),
    "exc_fp64_katz_mean": (c_int, [c_void_p, c_double, c_double, P(c_double), P(c_int)]),
    "exc_fp64_set_coef": (c_int, [c_void_p, c_void_p, c_void_p]),
    "exc_fp64_solve_coef": (c_int, [c_void_p, c_int, c_void_p, c_double, c_int, c_double, c_void_p, P(c_double),
                                    P(c_int)]),
}

# include/excent.h
HOST_SIGNATURES = {
    "excent_main": (c_int, [c_int, P(ctypes.c_char_p)]),
    "excent_rmat_main": (c_int, [c_int, P(ctypes.c_char_p)]),
    "excent_convert_main": (c_int, [c_int, P(ctypes.c_char_p)]),
    "excent_reference_main": (c_int, [c_int, P(ctypes.c_char_p)]),
    "excent_params_main": (c_int, [c_int, P(ctypes.c_char_p)]),
    "excent_egr_info_main": (c_int, [c_int, P(ctypes.c_char_p)]),
    "excent_evict_main": (c_int, [c_int, P(ctypes.c_char_p)]),
    "excent_ballast_main": (c_int, [c_int, P(ctypes.c_char_p)]),
    "excent_session_open": (c_int, [c_int, P(ctypes.c_char_p), P(c_void_p)]),
    "excent_session_run": (c_int, [c_void_p, c_int, P(RunStats)]),
    "excent_session_results": (c_int, [c_void_p]),
    "excent_session_set_column": (c_int, [c_void_p, c_int]),
    "excent_session_values": (c_int, [c_void_p, c_int, c_int, c_void_p]),
    "excent_session_vertices": (c_int, [c_void_p]),
    "excent_session_close": (c_int, [c_void_p]),
    "excent_session_error": (ctypes.c_char_p, []),
    "excent_dpr_create": (c_int, [c_int, c_int, c_u64, c_void_p, c_void_p, c_void_p, c_float, c_float, P(c_void_p)]),
    "excent_dpr_forward": (c_int, [c_void_p, c_void_p, c_void_p]),
    "excent_dpr_backward": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "excent_dpr_passes": (c_int, [c_void_p, P(c_int), P(c_int)]),
    "excent_dpr_destroy": (c_int, [c_void_p]),
    "excent_dpr_error": (ctypes.c_char_p, []),
    "excent_last_error": (ctypes.c_char_p, []),
    "excent_chunk_cap": (c_u64, [c_u64, c_double, c_double, c_double]),
    "excent_plan_chunks": (c_int, [ctypes.c_char_p, c_u64, c_void_p, c_void_p, c_void_p, c_void_p, c_int]),
    "excent_egr_header": (c_int, [ctypes.c_char_p, P(c_int), P(c_u64), P(c_int)]),
    "excent_egr_offsets": (c_int, [ctypes.c_char_p, c_void_p]),
    "excent_egr_edges": (c_int, [ctypes.c_char_p, c_u64, c_u64, c_void_p]),
    "excent_fetch_plan": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_int, c_int,
                                  c_void_p, c_void_p, c_void_p, P(c_u64), P(c_u64)]),
    "excent_pick_group": (c_int, [c_double, c_int]),
    "excent_hot_select": (c_int, [c_int, c_void_p, c_u64, c_void_p, c_void_p, c_void_p, c_void_p, P(c_u64)]),
    "excent_tau": (c_double, [c_double, c_void_p, c_void_p]),
    "excent_prefetch_counts": (c_int, [ctypes.c_char_p, c_void_p, c_int, c_int, P(c_u64), P(c_u64)]),
    "excent_gather_list": (c_int, [c_void_p, c_int, c_int, c_void_p]),
    "excent_top_in_degree": (c_int, [ctypes.c_char_p, c_int, c_double, c_int, c_void_p]),
    "excent_stage_lists": (c_int, [ctypes.c_char_p, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, P(c_u64),
                                   P(c_u64)]),
}

F = {}
for _name, (_res, _args) in HIP_SIGNATURES.items():
    F[_name] = _sig(hip, _name, _res, _args)
for _name, (_res, _args) in HOST_SIGNATURES.items():
    F[_name] = _sig(host, _name, _res, _args)


def check(rc, what=""):
    if rc != EXC_OK:
        raise ExcError(rc, f"{what}: {F['exc_last_error']().decode()}")


def run_entry(entry, args):
    """Runs a bin/ program's entry point in this process; returns its exit code."""
    argv = [entry.encode()] + [str(a).encode() for a in args]
    arr = (ctypes.c_char_p * len(argv))(*argv)
    return F[entry](len(argv), arr)


def excent(args):
    return run_entry("excent_main", args)


class Session:
    """bin/excent's setup once, then repeated pass loops (include/excent.h excent_session_*)."""

    def __init__(self, args):
        argv = [b"excent"] + [str(a).encode() for a in args]
        arr = (ctypes.c_char_p * len(argv))(*argv)
        self.h = c_void_p()
        rc = F["excent_session_open"](len(argv), arr, ctypes.byref(self.h))
        if rc != 0:
            raise ExcError(rc, F["excent_session_error"]().decode())

    def run(self, reinit=True):
        st = RunStats()
        rc = F["excent_session_run"](self.h, int(reinit), ctypes.byref(st))
        if rc != 0:
            raise ExcError(rc, F["excent_session_error"]().decode())
        out = {k: getattr(st, k) for k, _ in RunStats._fields_ if k not in ("kernel", "residual_l1")}
        out["residual_l1"] = list(st.residual_l1)
        out["kernel"] = stats_dict(st.kernel)
        return out

    def set_column(self, j):
        rc = F["excent_session_set_column"](self.h, int(j))
        if rc != 0:
            raise ExcError(rc, F["excent_session_error"]().decode())

    def values(self, column=0, which=0):
        """x (which = 0) or r (which = 1) of a measure slot or batched column, float32[n]."""
        n = F["excent_session_vertices"](self.h)
        out = np.zeros(max(n, 1), dtype=np.float32)
        rc = F["excent_session_values"](self.h, int(column), int(which), ptr(out))
        if rc != 0:
            raise ExcError(rc, F["excent_session_error"]().decode())
        return out[:n]

    def results(self):
        rc = F["excent_session_results"](self.h)
        if rc != 0:
            raise ExcError(rc, F["excent_session_error"]().decode())

    def close(self):
        if self.h:
            F["excent_session_close"](self.h)
            self.h = c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def synchronize(device=0):
    check(F["exc_device_synchronize"](device), "device synchronize")


def device_count():
    n = c_int(0)
    rc = F["exc_device_count"](ctypes.byref(n))
    return n.value if rc == 0 else 0


class Context:
    """One device context (include/excent_hip.h exc_create..exc_destroy)."""

    def __init__(self, n, m, cv0, kinds, eps, scale, divdeg, max_chunk_edges=1, max_chunk_verts=1,
                 trace=False, push_kernel=KERNEL_LB, timing=False, device=0, zero_copy=False, batch_k=0):
        self.n, self.m, self.nm = int(n), int(m), len(kinds)
        self.cv0 = np.ascontiguousarray(cv0, dtype=np.int32)
        self.nc = len(self.cv0)
        M = Measures()
        M.nm = self.nm
        for i in range(self.nm):
            M.kind[i], M.eps[i], M.scale[i], M.divdeg[i] = kinds[i], eps[i], scale[i], divdeg[i]
        cfg = Config(device, self.n, self.m, self.nc, self.cv0.ctypes.data_as(P(c_int)),
                     max(1, int(max_chunk_edges)), max(1, int(max_chunk_verts)), int(trace), int(zero_copy),
                     int(push_kernel), int(timing), int(batch_k))
        self.K = int(batch_k)
        self.ctx = c_void_p()
        check(F["exc_create"](ctypes.byref(cfg), ctypes.byref(M), ctypes.byref(self.ctx)), "exc_create")

    def close(self):
        if self.ctx:
            F["exc_destroy"](self.ctx)
            self.ctx = c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload_offsets(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.uint64)
        check(F["exc_upload_offsets"](self.ctx, ptr(idx), 0, len(idx)), "upload_offsets")

    def init_measure(self, i, beta=1.0, lam=1.0):
        check(F["exc_init_measure"](self.ctx, i, beta, lam), "init_measure")

    def set_state(self, i, x=None, r=None):
        x = None if x is None else np.ascontiguousarray(x, dtype=np.float32)
        r = None if r is None else np.ascontiguousarray(r, dtype=np.float32)
        check(F["exc_set_state"](self.ctx, i, ptr(x), ptr(r)), "set_state")

    def read_state(self, i, which):
        out = np.empty(self.n, dtype=np.float32)
        check(F["exc_read_state"](self.ctx, i, which, ptr(out), 0, self.n), "read_state")
        return out

    def snapshot(self, hot_skip=False):
        cnt = np.zeros(2 * self.nc, dtype=np.uint64)
        hp = c_u64(0)
        check(F["exc_snapshot"](self.ctx, int(hot_skip), ptr(cnt), ctypes.byref(hp)), "snapshot")
        return cnt, hp.value

    def bitmap(self):
        out = np.zeros((self.n + 31) // 32, dtype=np.uint32)
        check(F["exc_bitmap_to_host"](self.ctx, ptr(out)), "bitmap")
        return out

    def trace(self):
        out = np.zeros(self.n, dtype=np.uint16)
        check(F["exc_read_trace"](self.ctx, ptr(out), 0, self.n), "trace")
        return out

    def cache_chunks(self, edges):
        return F["exc_cache_chunks"](self.ctx, int(edges))

    def cache_upload(self, adj, e0=0):
        adj = np.ascontiguousarray(adj, dtype=np.int32)
        check(F["exc_cache_upload"](self.ctx, ptr(adj), e0, len(adj)), "cache_upload")

    def cache_hot(self, ids, off, bitmask):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        off = np.ascontiguousarray(off, dtype=np.uint64)
        bitmask = np.ascontiguousarray(bitmask, dtype=np.uint32)
        check(F["exc_cache_hot"](self.ctx, len(ids), ptr(ids), ptr(off), ptr(bitmask)), "cache_hot")

    def cache_hot_upload(self, edges, off0=0):
        edges = np.ascontiguousarray(edges, dtype=np.int32)
        check(F["exc_cache_hot_upload"](self.ctx, ptr(edges), off0, len(edges)), "cache_hot_upload")

    def push_resident(self, v0, v1, e0, e1, use_list, group=0, hot_skip=False):
        ch = Chunk(v0, v1, e0, e1)
        check(F["exc_push_resident"](self.ctx, ctypes.byref(ch), int(use_list), group, int(hot_skip)), "push_resident")

    def enable_pull(self, threshold):
        return F["exc_enable_pull"](self.ctx, threshold)

    def residual_l1(self):
        out = (c_double * 3)()
        check(F["exc_residual_l1"](self.ctx, out), "residual_l1")
        return list(out)[: self.nm]

    def set_hub_destinations(self, ids):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        check(F["exc_set_hub_destinations"](self.ctx, ptr(ids), len(ids)), "set_hub_destinations")

    def residual_max(self):
        out = (c_float * 3)()
        check(F["exc_residual_max"](self.ctx, out), "residual_max")
        return list(out)[: self.nm]

    def fold_ahead(self, allow):
        check(F["exc_fold_ahead"](self.ctx, 1 if allow else 0), "fold_ahead")

    def run_passes(self, passes, dense):
        """A burst of passes with one synchronisation: per pass the counts [2 nc] of the snapshot
        it leaves, its FP64 residual sums and largest |r| per measure."""
        cnt = np.zeros((passes, 2 * self.nc), np.uint64)
        l1 = np.zeros((passes, self.nm), np.float64)
        mx = np.zeros((passes, self.nm), np.float32)
        check(F["exc_run_passes"](self.ctx, passes, 1 if dense else 0, cnt.ctypes.data_as(P(c_u64)),
                                  l1.ctypes.data_as(P(c_double)), mx.ctypes.data_as(P(ctypes.c_float))), "run_passes")
        return cnt, l1, mx

    def burst_submit(self, passes, dense):
        """Enqueue a burst without waiting (at most two in flight)."""
        check(F["exc_burst_submit"](self.ctx, passes, 1 if dense else 0), "burst_submit")

    def burst_collect(self, passes):
        """Wait for the oldest burst in flight: as run_passes' outputs, and its pass count."""
        cnt = np.zeros((passes, 2 * self.nc), np.uint64)
        l1 = np.zeros((passes, self.nm), np.float64)
        mx = np.zeros((passes, self.nm), np.float32)
        got = c_int(0)
        check(F["exc_burst_collect"](self.ctx, cnt.ctypes.data_as(P(c_u64)), l1.ctypes.data_as(P(c_double)),
                                     mx.ctypes.data_as(P(ctypes.c_float)), ctypes.byref(got)), "burst_collect")
        return cnt[:got.value], l1[:got.value], mx[:got.value]

    def set_active_measures(self, count):
        """Dense passes over the first `count` measure slots only (the others converged)."""
        check(F["exc_set_active_measures"](self.ctx, count), "set_active_measures")

    def dense_required(self):
        rc = F["exc_dense_required"](self.ctx)
        check(min(rc, 0), "dense_required")
        return rc == 1

    def push_hot(self, lo, hi, group=0, after_slots=False):
        check(F["exc_push_hot"](self.ctx, lo, hi, group, int(after_slots)), "push_hot")

    def slot_copy_probe(self, s, nbytes):
        """Seconds of one synchronous H2D copy of nbytes from slot s's pinned edge buffer."""
        t = c_double(0)
        check(F["exc_slot_copy_probe"](self.ctx, s, int(nbytes), ctypes.byref(t)), "slot_copy_probe")
        return t.value

    def slot_buffers(self, s, max_edges, max_verts):
        e, v, o = c_void_p(), c_void_p(), c_void_p()
        check(F["exc_slot_buffers"](self.ctx, s, ctypes.byref(e), ctypes.byref(v), ctypes.byref(o)), "slot_buffers")
        edges = np.ctypeslib.as_array(ctypes.cast(e, P(ctypes.c_int32)), shape=(max_edges,))
        vids = np.ctypeslib.as_array(ctypes.cast(v, P(ctypes.c_int32)), shape=(max_verts + 1,))
        off = np.ctypeslib.as_array(ctypes.cast(o, P(ctypes.c_uint64)), shape=(max_verts + 1,))
        return edges, vids, off

    def slot_acquire(self, s):
        w = c_double(0)
        check(F["exc_slot_acquire"](self.ctx, s, ctypes.byref(w)), "slot_acquire")

    def stream_chunk(self, s, v0, v1, e0, e1, to_resident, use_list, group=0, hot_skip=False):
        ch = Chunk(v0, v1, e0, e1)
        check(F["exc_stream_chunk"](self.ctx, s, ctypes.byref(ch), int(to_resident), 0, 0, int(use_list), group,
                                    int(hot_skip)), "stream_chunk")

    def gather_chunk(self, s, k, group=0):
        check(F["exc_gather_chunk"](self.ctx, s, k, group), "gather_chunk")

    def pass_sync(self):
        w = c_double(0)
        check(F["exc_pass_sync"](self.ctx, ctypes.byref(w)), "pass_sync")

    def pushed(self):
        v = c_u64(0)
        check(F["exc_pushed"](self.ctx, ctypes.byref(v)), "pushed")
        return v.value

    def stats(self):
        s = Stats()
        check(F["exc_get_stats"](self.ctx, ctypes.byref(s)), "stats")
        return stats_dict(s)

    def reset_stats(self):
        check(F["exc_reset_stats"](self.ctx), "reset_stats")

    def batch_eps(self, eps):
        eps = np.ascontiguousarray(eps, dtype=np.float32)
        check(F["exc_batch_eps"](self.ctx, ptr(eps)), "batch_eps")

    def batch_seed(self, col, seeds, value):
        seeds = np.ascontiguousarray(seeds, dtype=np.int32)
        check(F["exc_batch_seed"](self.ctx, col, ptr(seeds), len(seeds), value), "batch_seed")

    def batch_colpend(self):
        out = np.zeros(self.K, np.uint32)
        check(F["exc_batch_colpend"](self.ctx, ptr(out)), "batch_colpend")
        return out

    def batch_set_active(self, mask):
        check(F["exc_batch_set_active"](self.ctx, int(mask)), "batch_set_active")

    def indegree(self):
        out = np.zeros(self.n, dtype=np.uint32)
        check(F["exc_indegree"](self.ctx, ptr(out)), "indegree")
        return out


class Fp64Solver:
    """The product's FP64 pull solver on one device (include/excent_hip.h exc_fp64_*), kept
    across solves and residual evaluations of one graph."""

    def __init__(self, idx, adj, device=0):
        idx = np.ascontiguousarray(idx, dtype=np.uint64)
        adj = np.ascontiguousarray(adj, dtype=np.int32)
        self.n = len(idx) - 1
        self.h = c_void_p()
        check(F["exc_fp64_create"](device, self.n, int(idx[-1]), ptr(idx), ptr(adj), ctypes.byref(self.h)),
              "fp64_create")

    def solve(self, kind, alpha=-1.0, beta=1.0, max_iter=20000, tol=1e-9, x0=None):
        """Jacobi (PageRank, Katz) or power iteration (eigenvector) from x0 (default: the
        reference's start vector) until the relative change is at most tol."""
        out = np.zeros(self.n, dtype=np.float64)
        if x0 is not None:
            x0 = np.ascontiguousarray(x0, dtype=np.float64)
            assert len(x0) == self.n
        rel, lam, it = c_double(0), c_double(0), c_int(0)
        check(F["exc_fp64_solve"](self.h, kind, alpha, beta, max_iter, tol, None if x0 is None else ptr(x0), ptr(out),
                                  ctypes.byref(rel), ctypes.byref(lam), ctypes.byref(it)), "fp64_solve")
        return out, rel.value, lam.value, it.value

    def residual(self, x, c, s, divdeg):
        """||c + s W x - x||_1 in FP64 for a float32 state x; c a scalar or a vector[n]."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        assert len(x) == self.n
        out = c_double(0)
        cv = None
        if np.ndim(c):
            cv = np.ascontiguousarray(c, dtype=np.float64)
            assert len(cv) == self.n
        check(F["exc_fp64_residual"](self.h, 0.0 if cv is not None else float(c), None if cv is None else ptr(cv),
                                     float(s), int(divdeg), ptr(x), ctypes.byref(out)), "fp64_residual")
        return out.value

    def solve_vec(self, c, s, divdeg, max_iter=20000, tol=1e-12):
        """Jacobi x = c + s W x for a per-vertex constant c (FP64)."""
        cv = np.ascontiguousarray(c, dtype=np.float64)
        assert len(cv) == self.n
        out = np.zeros(self.n, dtype=np.float64)
        rel, it = c_double(0), c_int(0)
        check(F["exc_fp64_solve_vec"](self.h, ptr(cv), float(s), int(divdeg), int(max_iter), float(tol), ptr(out),
                                      ctypes.byref(rel), ctypes.byref(it)), "fp64_solve_vec")
        return out, rel.value, it.value

    def set_coef(self, adj, coef):
        """Per-edge FP64 coefficients (adj: the adjacency given at construction)."""
        adj = np.ascontiguousarray(adj, dtype=np.int32)
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        assert len(adj) == len(coef)
        check(F["exc_fp64_set_coef"](self.h, ptr(adj), ptr(coef)), "fp64_set_coef")

    def solve_coef(self, c, s, transpose=False, max_iter=20000, tol=1e-14):
        """Jacobi x = c + s W x (transpose: x = c + s W^T x) with the per-edge coefficients."""
        cv = np.ascontiguousarray(c, dtype=np.float64)
        assert len(cv) == self.n
        out = np.zeros(self.n, dtype=np.float64)
        rel, it = c_double(0), c_int(0)
        check(F["exc_fp64_solve_coef"](self.h, int(bool(transpose)), ptr(cv), float(s), int(max_iter), float(tol),
                                       ptr(out), ctypes.byref(rel), ctypes.byref(it)), "fp64_solve_coef")
        return out, rel.value, it.value

    def close(self):
        if self.h:
            F["exc_fp64_destroy"](self.h)
            self.h = c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def fp64_solve(idx, adj, kind, alpha=-1.0, beta=1.0, max_iter=20000, tol=1e-9, device=0):
    """FP64 solve through the product's HIP solver (bin/centrality_reference's engine)."""
    idx = np.ascontiguousarray(idx, dtype=np.uint64)
    adj = np.ascontiguousarray(adj, dtype=np.int32)
    n = len(idx) - 1
    S = c_void_p()
    check(F["exc_fp64_create"](device, n, int(idx[-1]), ptr(idx), ptr(adj), ctypes.byref(S)), "fp64_create")
    try:
        out = np.zeros(n, dtype=np.float64)
        rel, lam, it = c_double(0), c_double(0), c_int(0)
        check(F["exc_fp64_solve"](S, kind, alpha, beta, max_iter, tol, None, ptr(out), ctypes.byref(rel),
                                  ctypes.byref(lam), ctypes.byref(it)), "fp64_solve")
        return out, rel.value, lam.value, it.value
    finally:
        F["exc_fp64_destroy"](S)


def load_egr(path):
    """Offsets (uint64) and columns (int32) of an EGR file, through the product reader."""
    n, m, wide = c_int(0), c_u64(0), c_int(0)
    if F["excent_egr_header"](path.encode(), ctypes.byref(n), ctypes.byref(m), ctypes.byref(wide)) != 0:
        raise ExcError(-1, F["excent_last_error"]().decode())
    idx = np.zeros(n.value + 1, dtype=np.uint64)
    adj = np.zeros(max(1, m.value), dtype=np.int32)
    if F["excent_egr_offsets"](path.encode(), ptr(idx)) != 0:
        raise ExcError(-1, F["excent_last_error"]().decode())
    if m.value and F["excent_egr_edges"](path.encode(), 0, m.value, ptr(adj)) != 0:
        raise ExcError(-1, F["excent_last_error"]().decode())
    return idx, adj[: m.value]


class DiffPageRank:
    """Differentiable personalised PageRank through include/excent.h excent_dpr_*."""

    def __init__(self, idx, adj, w_raw, damp=0.85, eps=1e-7, device=0):
        self.idx = np.ascontiguousarray(idx, np.uint64)
        self.adj = np.ascontiguousarray(adj, np.int32)
        self.w = np.ascontiguousarray(w_raw, np.float32)
        self.n, self.m = len(self.idx) - 1, len(self.adj)
        self.h = c_void_p()
        rc = F["excent_dpr_create"](device, self.n, self.m, ptr(self.idx), ptr(self.adj), ptr(self.w), damp, eps,
                                    ctypes.byref(self.h))
        if rc != 0:
            raise ExcError(rc, F["excent_dpr_error"]().decode())

    def forward(self, s):
        s = np.ascontiguousarray(s, np.float32)
        x = np.zeros(self.n, np.float32)
        rc = F["excent_dpr_forward"](self.h, ptr(s), ptr(x))
        if rc != 0:
            raise ExcError(rc, F["excent_dpr_error"]().decode())
        return x

    def backward(self, g):
        g = np.ascontiguousarray(g, np.float32)
        gs = np.zeros(self.n, np.float32)
        gw = np.zeros(self.m, np.float32)
        gwn = np.zeros(self.m, np.float32)
        rc = F["excent_dpr_backward"](self.h, ptr(g), ptr(gs), ptr(gw), ptr(gwn))
        if rc != 0:
            raise ExcError(rc, F["excent_dpr_error"]().decode())
        return gs, gw, gwn

    def passes(self):
        a, b = c_int(0), c_int(0)
        F["excent_dpr_passes"](self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def close(self):
        if self.h:
            F["excent_dpr_destroy"](self.h)
            self.h = c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

"""Benchmark: converged GTEPS and time-to-eps of bin/excent's engine on one MI355X.

Workload (default): seeded R-MAT scale 24, edge factor 16 (a,b,c = .57,.19,.19),
self-loops and duplicates removed, --measure all (PageRank d = 0.85, eps 0.01;
Katz alpha = 0.5/lambda1, beta = 1; eigenvector with lambda1 from a seeded FP64
power iteration; eps-katz = eps-ev = 1e-6 per vertex), float32 state, the whole
graph resident in HBM before timing (--cache chunks --preload).  The graph is
directed, so the eigenvector's initial residual uses in-degrees (--ev-init indeg).

A step is one converged run: the state is restored to its initial values (a device
copy) and the pass loop runs until nothing is pending.  value = pushed adjacency
entries / timed seconds / 1e9 (converged GTEPS).  roofline: the dominant push
kernel's algorithmic bytes per launch / its average HIP-event duration.
cpu_baseline: the oracle's OpenMP restatement of the same pass loop on the same
graph and constants, on this host's cores, for a bounded number of seconds.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import glob
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "converged GTEPS + time-to-eps, in-HBM & out-of-core; batched vs plain at FP64 err"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class native_stdout_to_stderr:
    """The generator and constants tools print progress with C stdio; keep it off stdout,
    which carries only the JSON line."""

    def __enter__(self):
        import ctypes
        self.libc = ctypes.CDLL(None)
        sys.stdout.flush()
        self.libc.fflush(None)
        self.saved = os.dup(1)
        os.dup2(2, 1)

    def __exit__(self, *a):
        self.libc.fflush(None)
        os.dup2(self.saved, 1)
        os.close(self.saved)


def read_env(path):
    out = {}
    for line in open(path):
        line = line.strip()
        if line and not line.startswith("#") and "=" in line:
            k, v = line.split("=", 1)
            out[k] = float(v)
    return out


def solve_logged(S, kind, tol, max_iter, seg=20, **kw):
    """S.solve in segments of seg iterations, each resumed from the last one's vector and
    logged, so that a long FP64 solve on a large graph reports its progress."""
    x, it = None, 0
    while True:
        x, rel, lam, k = S.solve(kind, tol=tol, max_iter=min(seg, max_iter - it), x0=x, **kw)
        it += k
        if it >= max_iter or rel <= tol:
            return x, rel, lam, it
        log(f"fp64 solve kind {kind}: {it} iterations, relative change {rel:.3g}")


def top_overlap(np, a, b, k):
    """Vertices shared by the k largest entries of a and of b (k capped at the length)."""
    k = min(k, len(a))
    top = lambda v: set(np.argpartition(-np.asarray(v), k - 1)[:k].tolist())
    return len(top(a) & top(b))


def kendall_top(np, x, ref, t=1000):
    """Kendall tau over the t largest entries of ref, ranked by (value desc, index asc):
    (concordant - discordant) pairs / all pairs, ties in either vector counting neither
    (tools/centrality_reference.cu:111-136)."""
    t = min(t, len(ref))
    if t < 2:
        return 1.0
    part = np.argpartition(-ref, t - 1)[:t]
    top = part[np.lexsort((part, -ref[part]))]
    r, v = ref[top], np.asarray(x, dtype=np.float64)[top]
    s = np.sign(r[:, None] - r[None, :]) * np.sign(v[:, None] - v[None, :])
    iu = np.triu_indices(t, 1)
    return float(s[iu].sum() / (t * (t - 1) / 2))


def fp64_errors(nat, np, graph, work, measure, env):
    """Errors of the last timed run's values (the --dump files) against the product's FP64
    solver, and the FP64 global residual of the float32 state (tools/centrality_reference.cu:
    78-139 metrics; the residual bounds PageRank's L1 error by ||r||_1 / (1 - d))."""
    t0 = time.time()
    idx, adj = nat.load_egr(graph)
    f32 = np.float32
    out = {}
    with nat.Fp64Solver(idx, adj) as S:
        for name in {"pr": ["pr"], "katz": ["katz"], "ev": ["ev"], "all": ["pr", "katz", "ev"]}[measure]:
            x = np.fromfile(os.path.join(work, f"x.{name}.bin"), dtype=np.float32)
            if name == "pr":
                c, s, dd = float(f32(1) - f32(0.85)), float(f32(0.85)), 1
                ref, rel, _, it = solve_logged(S, nat.EXC_PR, 1e-12, 400)
            elif name == "katz":
                c, s, dd = 1.0, float(f32(env["ALPHA"])), 0
                ref, rel, _, it = solve_logged(S, nat.EXC_KATZ, 1e-12, 400, alpha=env["ALPHA"], beta=1.0)
            else:
                c, s, dd = 0.0, float(f32(1.0 / env["LAMBDA"])), 0
                ref, rel, _, it = solve_logged(S, nat.EXC_EV, 1e-10, 3000)
            res = S.residual(x, c, s, dd)
            xv = x.astype(np.float64)
            if name == "ev":   # unit L2 with the sign fixed by the sum
                xv = xv / np.linalg.norm(xv) * (1 if xv.sum() >= 0 else -1)
            d = np.abs(xv - ref)
            e = {"l1_rel": float(d.sum() / np.abs(ref).sum()), "l2": float(np.sqrt((d * d).sum())),
                 "max_abs": float(d.max()), "top100_overlap": top_overlap(np, xv, ref, 100),
                 "top1000_overlap": top_overlap(np, xv, ref, 1000), "tau1000": kendall_top(np, xv, ref),
                 "residual_l1_fp64": res, "ref_iters": it, "ref_rel_change": rel}
            if name == "pr":
                e["l1_bound"] = res / (1 - 0.85)
                e["l1_abs"] = float(d.sum())
            out[name] = {k2: (round(v, 9) if isinstance(v, float) else v) for k2, v in e.items()}
            log(f"fp64 {name}: {out[name]}")
    log(f"fp64 check {time.time() - t0:.1f}s")
    return out


def constants(nat, graph):
    """LAMBDA / ALPHA of the graph in the reference's .env format (tools/excent_params.cu),
    computed once per graph file by the product's FP64 power iteration."""
    envp = graph + ".env"
    t0 = time.time()
    if not os.path.exists(envp):
        with native_stdout_to_stderr():
            rc = nat.run_entry("excent_params_main", [graph, envp, "--directed", "--alpha-factor", "0.5", "--seed",
                                                      "7", "--eps-abs", "1e-6"])
        assert rc == 0, "constants failed"
    env = read_env(envp)
    log(f"constants {env} in {time.time() - t0:.1f}s")
    return env


def measure_args(measure, env):
    """Katz alpha/beta/eps and eigenvector lambda/eps flags of the all-measure workloads."""
    a = []
    if measure in ("katz", "all"):
        a += ["--alpha", env["ALPHA"], "--beta", 1, "--eps-katz", "1e-6"]
    if measure in ("ev", "all"):
        a += ["--lambda", env["LAMBDA"], "--eps-ev", "1e-6", "--ev-init", "indeg"]
    return a


def evict(path):
    """Drop the file's pages from the host page cache (what tools/evict_file.c does); dirty
    pages are written back first, since only clean ones can be dropped."""
    fd = os.open(path, os.O_RDONLY)
    try:
        os.fsync(fd)
        os.posix_fadvise(fd, 0, 0, os.POSIX_FADV_DONTNEED)
    finally:
        os.close(fd)


def page_cache_fraction(path, np):
    """Fraction of the file's pages in the host page cache (mincore over a read-only map)."""
    import ctypes
    libc = ctypes.CDLL(None, use_errno=True)
    libc.mmap.restype = ctypes.c_void_p
    libc.mmap.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_long]
    libc.mincore.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    libc.munmap.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    size = os.path.getsize(path)
    page = os.sysconf("SC_PAGE_SIZE")
    fd = os.open(path, os.O_RDONLY)
    try:
        addr = libc.mmap(None, size, 1, 1, fd, 0)   # PROT_READ, MAP_SHARED
        if addr in (None, ctypes.c_void_p(-1).value):
            return None
        vec = np.zeros((size + page - 1) // page, np.uint8)
        rc = libc.mincore(addr, size, vec.ctypes.data)
        libc.munmap(addr, size)
        return None if rc else round(float((vec & 1).mean()), 4)
    finally:
        os.close(fd)


def graph_size(nat, graph):
    n, m, wide = nat.c_int(0), nat.c_u64(0), nat.c_int(0)
    nat.F["excent_egr_header"](graph.encode(), nat.ctypes.byref(n), nat.ctypes.byref(m), nat.ctypes.byref(wide))
    return n.value, m.value


def bench_batched(args, nat, np, work, graph, sample=None):
    """K (default 64) personalised PageRank columns in one batched run against the same K
    problems solved one after another by the plain solver (--measure pr --ppr-column j),
    both from the resident graph; FP64 errors on sampled columns."""
    K, SEEDS, TOP, EPS = args.k, 16, 8, 1e-7
    gen = f"{K}:{SEEDS}:{TOP}:{args.seed + 6}"
    common = [graph, "--cache", "chunks", "--preload", "--stream-ratio", "1", "--timing", "--ppr-gen", gen,
              "--eps-ppr", EPS]
    n, m = graph_size(nat, graph)
    sample = sorted(sample or {0, K // 3, (2 * K) // 3, K - 1})
    t0 = time.time()
    sess = nat.Session(common + ["--measure", "ppr", "--pull", args.pull,
                                 "--log", os.path.join(args.logdir or work, "batched.log")])
    log(f"batched setup {time.time() - t0:.1f}s")
    for w in range(args.warmup):
        sess.run(reinit=w > 0)
        log(f"batched warmup {w} done")
    nat.synchronize()
    t0 = time.perf_counter()
    rb = sess.run(reinit=args.warmup > 0)
    nat.synchronize()
    tb = time.perf_counter() - t0
    xb = {j: sess.values(j) for j in sample}
    sess.close()
    log(f"batched: {rb['passes']} passes, {tb:.3f}s, {rb['kernel']['pushed_edge_measures'] / 1e9:.2f} G vector pushes, "
        f"{rb['kernel']['front_snapshots']} snapshots over the touched vertices only")
    kb = rb["kernel"]
    bpull = kb["pull_bytes"] / (kb["pull_ms"] / 1e3) / 1e9 if kb["pull_ms"] > 0 else 0.0
    roof = {"bound": "hbm", "achieved": round(bpull, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(bpull / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_bpull (batched dense pull)",
            "launches": kb["pull_passes"], "avg_launch_ms": round(kb["pull_ms"] / max(1, kb["pull_passes"]), 4),
            "algorithmic_bytes_per_launch": round(kb["pull_bytes"] / max(1, kb["pull_passes"]), 1)}
    if args.batched_only:
        return {"metric": "batched personalised PageRank (no baseline)", "unit": "s", "value": round(tb, 4),
                "k": K, "passes": rb["passes"], "dense_passes": kb["pull_passes"],
                "pull_rows_in_flight": kb["pull_rows_in_flight"], "front_snapshots": kb["front_snapshots"],
                "pull_hub_gbs": round(kb["pull_hub_bytes"] / (kb["pull_hub_ms"] / 1e3) / 1e9, 2) if kb["pull_hub_ms"] > 0 else None,
                "pull_tune_ns_per_edge": [round(v, 4) for v in kb["pull_tune_ns_per_edge"]],
                "vector_pushes": kb["pushed_edge_measures"], "vector_gteps": round(kb["pushed_edge_measures"] / tb / 1e9, 3),
                "kernel_ms": {x: round(kb[x], 2) for x in ("snapshot_ms", "push_ms", "piece_ms", "fold_ms", "pull_ms",
                                                            "pull_hub_ms")},
                "roofline": roof}

    sess = nat.Session(common + ["--measure", "pr", "--ppr-column", 0, "--pull", args.pull,
                                 "--log", os.path.join(args.logdir or work, "plain.log")])
    sess.set_column(0)
    sess.run(reinit=False)   # warmup
    seq_t, seq_pushes, seq_passes, xp, cvec = [], 0, [], {}, {}
    for j in range(K):
        sess.set_column(j)
        if j in sample:
            cvec[j] = sess.values(0, which=1).astype(np.float64)   # r0 = (1 - d) * teleport
        nat.synchronize()
        t0 = time.perf_counter()
        r = sess.run(reinit=False)
        nat.synchronize()
        seq_t.append(time.perf_counter() - t0)
        seq_pushes += r["pushed_edges"]
        seq_passes.append(r["passes"])
        if j in sample:
            xp[j] = sess.values(0)
        log(f"plain column {j}: {r['passes']} passes, {seq_t[-1]:.3f}s")
    sess.close()
    ts = sum(seq_t)
    log(f"plain: {K} runs {ts:.3f}s, {seq_pushes / 1e9:.2f} G pushes")

    fp64 = {}
    idx, adj = nat.load_egr(graph)
    with nat.Fp64Solver(idx, adj) as S:
        d = float(np.float32(0.85))
        for j in sample:
            ref, rel, it = S.solve_vec(cvec[j], d, 1, tol=1e-12, max_iter=1000)
            e = {"ref_iters": it}
            for tag, x in (("batched", xb[j]), ("plain", xp[j])):
                dx = np.abs(x.astype(np.float64) - ref)
                e[tag] = {"l1_rel": round(float(dx.sum() / np.abs(ref).sum()), 9),
                          "max_abs": round(float(dx.max()), 9),
                          "l1_bound_rel": round(S.residual(x, cvec[j], d, 1) / (1 - 0.85) / np.abs(ref).sum(), 9),
                          "top1000_overlap": top_overlap(np, x.astype(np.float64), ref, 1000),
                          "tau1000": round(kendall_top(np, x, ref), 6)}
            fp64[f"column {j}"] = e
            log(f"fp64 column {j}: {e}")
    vb = rb["kernel"]["pushed_edge_measures"]
    return {
        "metric": "batched vs plain personalised PageRank at FP64 error", "unit": "x (wall-time ratio)",
        "value": round(ts / tb, 3), "higher_is_better": True, "n_gpus": 1, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"R-MAT scale {args.scale} ef {args.edge_factor}, K={K} personalised PageRank "
                               f"columns ({SEEDS} seeds each, {TOP} from the top 1% by out-degree, mass n/{SEEDS} "
                               f"per seed), eps {EPS} per vertex and column, whole graph in HBM",
                   "n": n, "m": m, "ppr_gen": gen},
        "batched": {"time_s": round(tb, 4), "passes": rb["passes"], "vector_pushes": vb,
                    "vector_gteps": round(vb / tb / 1e9, 3),
                    "kernel_ms": {k: round(rb["kernel"][k], 2) for k in ("snapshot_ms", "push_ms", "piece_ms",
                                                                          "fold_ms", "pull_ms", "pull_hub_ms")},
                    "dense_passes": rb["kernel"]["pull_passes"],
                    "pull_rows_in_flight": rb["kernel"]["pull_rows_in_flight"],
                    "pull_tune_ns_per_edge": [round(v, 4) for v in rb["kernel"]["pull_tune_ns_per_edge"]]},
        "plain": {"time_s": round(ts, 4), "time_s_per_column": round(ts / K, 4),
                  "passes_mean": round(sum(seq_passes) / K, 1), "vector_pushes": seq_pushes,
                  "vector_gteps": round(seq_pushes / ts / 1e9, 3)},
        "roofline": roof,
        "wall_ratio": round(ts / tb, 3),
        "vector_push_rate_ratio": round((vb / tb) / (seq_pushes / ts), 3),
        "fp64_err": fp64,
    }


def storage_probe(path, nbytes=4 << 30):
    """Sequential read rate of the graph file from storage: the page cache is dropped first,
    then `nbytes` from the middle of the file are read in 16 MiB requests, once through the
    page cache and once with O_DIRECT (aligned buffer) when the file system supports it."""
    import mmap
    size = os.path.getsize(path)
    nbytes = min(nbytes, size // 2) // (16 << 20) * (16 << 20)
    start = (size // 4) // 4096 * 4096
    out = {"bytes": nbytes}
    buf = mmap.mmap(-1, 16 << 20)   # page-aligned
    for mode in ("buffered", "direct"):
        evict(path)
        flags = os.O_RDONLY | (getattr(os, "O_DIRECT", 0) if mode == "direct" else 0)
        try:
            fd = os.open(path, flags)
        except OSError as e:
            out[mode] = f"unavailable: {e}"
            continue
        try:
            t0 = time.perf_counter()
            done = 0
            while done < nbytes:
                got = os.preadv(fd, [buf], start + done)
                if got <= 0:
                    break
                done += got
            el = time.perf_counter() - t0
            out[f"{mode}_gb_per_s"] = round(done / el / 1e9, 3)
        except OSError as e:
            out[mode] = f"unavailable: {e}"
        finally:
            os.close(fd)
    evict(path)
    log(f"storage probe: {out}")
    return out


def bench_ooc(args, nat, np, work, graph):
    """Out-of-core run: the graph file streamed / gathered over PCIe each pass, with a capped
    hot cache and with no cache (and, with --cold, the hot-cache run again after the file
    was dropped from the host page cache); bytes shipped as the reference counts them; FP64
    errors of the hot-cache run's result."""
    n, m = graph_size(nat, graph)
    measure = args.measure
    env = constants(nat, graph) if measure in ("katz", "ev", "all") else {}
    eps = {"pr": "eps 0.01", "katz": "eps 1e-6", "ev": "eps 1e-6", "all": "eps 0.01 / 1e-6 / 1e-6"}[measure]
    out = {"metric": f"converged GTEPS out-of-core (--measure {measure})", "unit": "GTEPS", "higher_is_better": True,
           "n_gpus": 1, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"R-MAT scale {args.scale} ef {args.edge_factor} --measure {measure}, {eps}, "
                                  f"edges read from the EGR file every pass",
                      "n": n, "m": m, "graph_gb": round(m * 4 / 1e9, 3)}, "runs": {}}
    hot = f"hot {args.cache_mb} MB"
    modes = [(hot, ["--cache", "hot", "--cache-mb", args.cache_mb], False), ("no cache", ["--cache", "none"], False)]
    if args.cold:
        modes.append((hot + ", cold page cache", ["--cache", "hot", "--cache-mb", args.cache_mb], True))
    for tag, extra, cold in modes:
        sargs = [graph, "--measure", measure, "--timing", "--log",
                 os.path.join(args.logdir or work, f"ooc_{tag.split()[0]}{'_cold' if cold else ''}.log")] + extra
        sargs += measure_args(measure, env) + ["--stop", stop_rule(args, measure)]
        if args.max_passes:
            sargs += ["--max-passes", args.max_passes]
        if tag == hot:
            sargs += ["--dump", os.path.join(work, "x")]
        resident = None
        if cold:
            evict(graph)
            resident = page_cache_fraction(graph, np)
            log(f"ooc {tag}: {resident} of the file in the page cache after eviction")
        t0 = time.time()
        sess = nat.Session(sargs)
        log(f"ooc {tag}: setup {time.time() - t0:.1f}s")
        for w in range(0 if cold else args.warmup):
            sess.run(reinit=w > 0)
            log(f"ooc {tag}: warmup {w} done")
        nat.synchronize()
        t0 = time.perf_counter()
        r = sess.run(reinit=args.warmup > 0 and not cold)
        nat.synchronize()
        el = time.perf_counter() - t0
        if tag == hot:
            sess.results()
        sess.close()
        shipped = r["bytes_stream"] + r["bytes_gather"]
        out["runs"][tag] = {"time_to_eps_s": round(el, 4), "passes": r["passes"],
                            "converged": r["stop_reason"] != 2,
                            "stop": ["nothing pending", "global l1 <= n*eps", "pass limit"][r["stop_reason"]],
                            "residual_l1_fp64_at_stop": [float(f"{v:.6g}") for v in r["residual_l1"]],
                            "gteps": round(r["pushed_edges"] / el / 1e9, 3),
                            "bytes_stream_gb": round(r["bytes_stream"] / 1e9, 3),
                            "bytes_gather_gb": round(r["bytes_gather"] / 1e9, 3),
                            "shipped_gb_per_s": round(shipped / el / 1e9, 2),
                            "time_io_s": round(r["time_io_s"], 4), "time_wait_s": round(r["time_wait_s"], 4),
                            "kernel_ms": {k: round(r["kernel"][k], 1) for k in ("snapshot_ms", "push_ms", "piece_ms",
                                                                                "other_ms")}}
        if cold:
            out["runs"][tag]["page_cache_fraction_at_start"] = resident
        log(f"ooc {tag}: {out['runs'][tag]}")
    # the link peak: synchronous copies of a 256 MB pinned staging buffer to the device
    chunk = 64 << 20
    with nat.Context(1, 1, np.array([0], np.int32), [nat.EXC_PR], [0.01], [0.85], [1], max_chunk_edges=chunk,
                     max_chunk_verts=1) as ctx:
        best = min(ctx.slot_copy_probe(0, chunk * 4) for _ in range(5))
    link = chunk * 4 / best / 1e9
    out["link_h2d_gb_per_s"] = round(link, 2)
    out["storage"] = storage_probe(graph)
    out["host"] = host_info()
    for r in out["runs"].values():
        r["link_fraction"] = round(r["shipped_gb_per_s"] / link, 3)
    log(f"ooc: pinned H2D {link:.1f} GB/s")
    out["value"] = out["runs"][hot]["gteps"]
    out["time_to_eps_s"] = out["runs"][hot]["time_to_eps_s"]
    if not args.no_fp64:
        out["fp64_error"] = fp64_errors(nat, np, graph, work, measure, env)
    return out


def adjoint_fp64_checks(nat, np, idx, adj, w, s, g, x, gs, gw, gwn, seed, fd_dirs=32):
    """The float32 results against the product's FP64 solves with per-edge coefficients
    (exc_fp64_solve_coef, pinned to the oracle's weighted solve in the tests): the forward
    x, the gradients against the FP64 adjoint solve lambda = (I - d W^T)^-1 g
    (dL/ds = (1 - d) lambda; dL/dw of a raw weight by the chain rule through the per-source
    normalisation), and central differences of L = g . x in FP64 on fd_dirs seeded directions
    (even: raw edge weights, odd: teleport), step 1e-3."""
    D = 0.85
    t0 = time.time()
    n, m = len(idx) - 1, len(adj)
    deg = np.diff(idx.astype(np.int64))
    src = np.repeat(np.arange(n, dtype=np.int64), deg)
    wd = w.astype(np.float64)
    s64, g64 = s.astype(np.float64), g.astype(np.float64)

    def coefs(wr):
        S = np.bincount(src, weights=wr, minlength=n)
        return wr / S[src], S

    wn, S = coefs(wd)
    rel = lambda a, b: float(np.linalg.norm(a.astype(np.float64) - b) / max(np.linalg.norm(b), 1e-300))
    out = {}
    with nat.Fp64Solver(idx, adj) as F:
        F.set_coef(adj, wn)
        xr, _, itf = F.solve_coef((1 - D) * s64, D)
        lam, _, itb = F.solve_coef(g64, D, transpose=True)
        gs_ref = (1 - D) * lam
        t = np.bincount(src, weights=wn * lam[adj], minlength=n)
        gw_ref = D * xr[src] / S[src] * (lam[adj] - t[src])
        gwn_ref = D * lam[adj] * xr[src]
        out["x_l1_rel"] = float(np.abs(x.astype(np.float64) - xr).sum() / np.abs(xr).sum())
        out["grad_s_l2_rel"] = rel(gs, gs_ref)
        out["grad_w_l2_rel"] = rel(gw, gw_ref)
        out["grad_w_normalised_l2_rel"] = rel(gwn, gwn_ref)
        out["fp64_iters"] = [itf, itb]
        log(f"adjoint fp64: forward {itf} / adjoint {itb} iterations, {out}")
        h = 1e-3
        worst_gpu = worst_fp64 = 0.0
        for k in range(fd_dirs):
            dr = np.random.default_rng(seed + 100 + k)
            if k % 2 == 0:   # raw edge weights (stay positive: uniform(0.5, 1.5) +- 1e-3 N(0, 1))
                dw = dr.standard_normal(m)
                F.set_coef(adj, coefs(wd + h * dw)[0])
                lp = float(g64 @ F.solve_coef((1 - D) * s64, D)[0])
                F.set_coef(adj, coefs(wd - h * dw)[0])
                lm = float(g64 @ F.solve_coef((1 - D) * s64, D)[0])
                an, af = float(gw.astype(np.float64) @ dw), float(gw_ref @ dw)
            else:            # teleport vector
                ds = dr.standard_normal(n)
                F.set_coef(adj, wn)
                lp = float(g64 @ F.solve_coef((1 - D) * (s64 + h * ds), D)[0])
                lm = float(g64 @ F.solve_coef((1 - D) * (s64 - h * ds), D)[0])
                an, af = float(gs.astype(np.float64) @ ds), float(gs_ref @ ds)
            fd = (lp - lm) / (2 * h)
            worst_gpu = max(worst_gpu, abs(an - fd) / max(abs(fd), 1e-300))
            worst_fp64 = max(worst_fp64, abs(af - fd) / max(abs(fd), 1e-300))
            log(f"adjoint fd direction {k} ({'w' if k % 2 == 0 else 's'}): fd {fd:.9g}, gpu {an:.9g}, "
                f"fp64 {af:.9g}, {time.time() - t0:.1f}s")
        out["fd_directions"] = fd_dirs
        out["fd_max_rel_err_gpu"] = worst_gpu
        out["fd_max_rel_err_fp64_adjoint"] = worst_fp64
    out = {k: (float(f"{v:.4g}") if isinstance(v, float) else v) for k, v in out.items()}
    log(f"adjoint fp64 checks {time.time() - t0:.1f}s: {out}")
    return out


def bench_adjoint(args, nat, np, work, graph, fd_dirs=32):
    """Differentiable personalised PageRank on a weighted graph: forward push, adjoint push
    over the transpose and gradient assembly, timed; the adjoint identity
    g . x = (1 - d) lambda . s checks the two solves against each other."""
    idx, adj = nat.load_egr(graph)
    w = np.fromfile(graph + ".w.f32", dtype=np.float32)
    n, m = len(idx) - 1, len(adj)
    rng = np.random.default_rng(args.seed)
    s = np.ones(n, np.float32)
    g = rng.standard_normal(n).astype(np.float32)
    with nat.DiffPageRank(idx, adj, w, eps=1e-7) as P:
        P.forward(s)
        P.backward(g)   # warmup
        t0 = time.perf_counter()
        x = P.forward(s)
        tf = time.perf_counter() - t0
        t0 = time.perf_counter()
        gs, gw, gwn = P.backward(g)
        tb = time.perf_counter() - t0
        pf, pb = P.passes()
    lhs = float(g.astype(np.float64) @ x.astype(np.float64))
    rhs = float(gs.astype(np.float64) @ s.astype(np.float64))   # gs = (1 - d) lambda
    fp64 = None if args.no_fp64 else adjoint_fp64_checks(nat, np, idx, adj, w, s, g, x, gs, gw, gwn, args.seed, fd_dirs)
    return {"metric": "differentiable personalised PageRank forward + backward", "unit": "s",
            "value": round(tf + tb, 4), "higher_is_better": False, "n_gpus": 1, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"R-MAT scale {args.scale} ef {args.edge_factor}, raw weights uniform(0.5, 1.5), "
                                   f"s = 1, L = g . x with g ~ N(0, 1), eps 1e-7", "n": n, "m": m},
            "forward_s": round(tf, 4), "backward_s": round(tb, 4), "passes": [pf, pb],
            "adjoint_identity_rel_err": abs(lhs - rhs) / max(abs(lhs), 1e-30),
            "grad_norms": {"s": float(np.linalg.norm(gs)), "w": float(np.linalg.norm(gw)),
                           "w_normalised": float(np.linalg.norm(gwn))},
            "fp64_check": fp64}


def host_info():
    """The host the CPU baseline ran on: logical CPUs, those this process may use, the CPU
    model and RAM."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    ram = None
    try:
        for line in open("/proc/meminfo"):
            if line.startswith("MemTotal:"):
                ram = round(int(line.split()[1]) / 2 ** 20, 1)
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable_cpus": len(os.sched_getaffinity(0)), "cpu_model": model, "ram_gib": ram}


def cpu_threads():
    """Every host core this process may run on (the baseline's definition: all host cores)."""
    return len(os.sched_getaffinity(0))


def plain_state(orc, np, idx, adj, measure, env):
    """Initial x, r and the constants of the CPU restatement, as bin/excent sets them up
    (the eigenvector starts from in-degrees: the graphs are directed)."""
    n = len(idx) - 1
    kinds = {"pr": [0], "katz": [1], "ev": [2], "all": [0, 1, 2]}[measure]
    xs, rs, eps, scale, divdeg = [], [], [], [], []
    indeg = np.bincount(adj, minlength=n).astype(np.float64)
    for k in kinds:
        if k == 0:
            x, r = orc.init(idx, 0)
            eps.append(0.01); scale.append(np.float32(0.85)); divdeg.append(1)
        elif k == 1:
            x, r = orc.init(idx, 1, beta=1.0)
            eps.append(1e-6); scale.append(np.float32(env["ALPHA"])); divdeg.append(0)
        else:
            lam = float(np.float32(env["LAMBDA"]))
            x = np.ones(n, np.float32)
            r = (indeg / lam - 1.0).astype(np.float32)
            eps.append(1e-6); scale.append(np.float32(1.0 / env["LAMBDA"])); divdeg.append(0)
        xs.append(x); rs.append(r)
    return xs, rs, eps, scale, divdeg


def cpu_baseline_run(np, orc, idx, adj, xs, rs, eps, scale, divdeg, seconds, what):
    """The oracle's OpenMP restatement of the pass loop (snapshot, plan, push with CAS float
    atomics) on every usable host core, to convergence or for `seconds`."""
    threads = cpu_threads()
    cv0 = orc.plan_chunks(idx, orc.chunk_cap(len(adj)))
    passes, pushed, secs = orc.run_push(idx, adj, cv0, xs, rs, eps, scale, divdeg, threads=threads, time_limit=seconds)
    converged = all(float(np.abs(r).max(initial=0.0)) <= e for r, e in zip(rs, eps))
    sample = (f"{what}: converged run, {passes} passes, {secs:.1f}s" if converged else
              f"{what}: first {passes} passes of the same run, {secs:.1f}s of a {seconds:.0f}s budget")
    log(f"cpu baseline: {sample}, {pushed / 1e9:.3f} G edges")
    rec = {"value": round(pushed / secs / 1e9, 4) if secs > 0 else None, "unit": "GTEPS", "cores": threads,
           "kind": "port", "sample": sample, "converged": converged, "passes": passes}
    if converged:
        rec["time_to_eps_s"] = round(secs, 3)
    rec.update(host_info())
    return rec


def stop_rule(args, measure):
    """The loop's stop rule: --stop when given; otherwise the reference's per-vertex rule,
    except that with --measure all Katz and the eigenvector end on their global sum
    sum |r| <= 1e-6 |V| (the configuration's eps = 1e-6 |V| read as a global L1 bound), while
    PageRank keeps eps 0.01 per vertex."""
    if args.stop:
        return args.stop
    return "katz,ev" if measure == "all" else "vertex"


def bench_plain(args, nat, np, work, graph, scale, measure, steps, warmup, env, tag, timing="--timing"):
    """Converged runs of bin/excent's engine with the whole graph in HBM: steps timed
    back to back after `warmup` untimed ones, each from the initial state (a device copy);
    the dominant kernel's roofline; FP64 errors of the last run; setup times."""
    excent_log = os.path.join(args.logdir or work, f"{tag}.log")
    sargs = ([graph, "--measure", measure, "--cache", "chunks", "--preload", "--stream-ratio", "1"] +
             ([timing] if timing else []) +
             ["--push-kernel", args.kernel, "--pull", args.pull, "--hub-lds", args.hub_lds, "--log", excent_log,
              "--dump", os.path.join(work, "x"), "--stop", stop_rule(args, measure),
              "--drop-measures", args.drop_measures, "--burst", args.burst, "--burst-ahead", args.burst_ahead,
              # a run whose stop rule is never met ends here and is reported as not converged
              "--max-passes", args.max_passes or 2000] + measure_args(measure, env))
    t0 = time.time()
    sess = nat.Session(sargs)
    t_session = time.time() - t0
    log(f"{tag}: session setup {t_session:.1f}s")
    first = True
    for w in range(warmup):
        r = sess.run(reinit=not first)
        first = False
        log(f"{tag} warmup {w}: {r['passes']} passes, {r['time_passes_s']:.3f}s, {r['pushed_edges'] / 1e9:.2f} G edges")
    nat.synchronize()
    runs = []
    t_start = time.perf_counter()
    for i in range(steps):
        runs.append(sess.run(reinit=not first))
        first = False
        if i == 0 or i == steps - 1:
            log(f"{tag} step {i} done: {runs[-1]['passes']} passes")
    nat.synchronize()
    elapsed = time.perf_counter() - t_start
    sess.results()
    sess.close()
    for i, r in enumerate(runs):
        log(f"{tag} step {i}: {r['passes']} passes, {r['time_passes_s']:.4f}s, {r['pushed_edges'] / 1e9:.3f} G edges")
    pushed = sum(r["pushed_edges"] for r in runs)
    K = lambda key: sum(r["kernel"][key] for r in runs)
    names = ["range", "bitmap", "list", "hot"]
    ktag = "3,4" if measure == "all" else "1,1"
    cands = [(f"k_push<{ktag},{i},0> ({names[i]} source)", sum(r["kernel"]["src_ms"][i] for r in runs),
              sum(r["kernel"]["src_launches"][i] for r in runs), sum(r["kernel"]["src_bytes"][i] for r in runs))
             for i in range(4)]
    npull = K("pull_passes")
    # the dense pull by width: passes over every measure, and (--measure all) those over the
    # first two / first measure once the others have nothing pending
    S = ktag.split(",")[1]
    for w in range(3):
        launches = sum(r["kernel"]["pull_w_launches"][w] for r in runs)
        if not launches:
            continue
        tile = runs[-1]["kernel"]["pull_w_tile"][w]
        name = f"k_pull_sorted<{w + 1},{S},{tile}>" if tile else f"k_pull<{w + 1},{S}>"
        cands.append((name, sum(r["kernel"]["pull_w_ms"][w] for r in runs), launches,
                      sum(r["kernel"]["pull_w_bytes"][w] for r in runs)))
    cands += [(f"k_fold<{ktag}>", K("fold_ms"), npull, K("fold_bytes")),
              (f"k_pull_hub<{ktag}>", K("pull_hub_ms"), npull, K("pull_hub_bytes")),
              (f"k_snapshot<{ktag}>", K("snapshot_ms"), K("snapshot_launches"), K("snapshot_bytes")),
              (f"k_pieces<{ktag}>", K("piece_ms"), K("piece_launches"), K("piece_bytes"))]
    dom = max(cands, key=lambda c: c[1])
    achieved = dom[3] / (dom[1] / 1e3) / 1e9 if dom[1] > 0 else 0.0
    traffic = None
    want = {"scale": scale, "edge_factor": args.edge_factor,
            "seed": args.seed, "measure": measure, "kernel": args.kernel, "pull": args.pull}
    for path in sorted(glob.glob(args.traffic)):
        # per-launch HBM bytes from rocprofv3 FETCH_SIZE/WRITE_SIZE runs of the same workload
        tr = json.load(open(path))
        if tr.get("workload") == want and traffic is None:   # by full name, else the kernel's base name
            bpl = tr["bytes_per_launch"]
            traffic = bpl.get(dom[0], bpl.get(dom[0].split("<")[0]))
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom[0],
            "launches": dom[2], "avg_launch_ms": round(dom[1] / max(1, dom[2]), 5),
            "algorithmic_bytes_per_launch": round(dom[3] / max(1, dom[2]), 1)}
    last = runs[-1]
    rec = {
        "value": round(pushed / elapsed / 1e9, 4), "unit": "GTEPS", "steps": steps, "warmup": warmup,
        "ms_per_step": round(elapsed / steps * 1e3, 3),
        "time_to_eps_s": round(sum(r["time_passes_s"] for r in runs) / len(runs), 5),
        "setup_s": {"session": round(t_session, 3), "open": round(last["time_open_s"], 3),
                    "setup": round(last["time_setup_s"], 3), "of_which_transpose": round(last["time_transpose_s"], 4)},
        "passes": [r["passes"] for r in runs],
        "converged": all(r["stop_reason"] != 2 for r in runs),
        "stop": sorted({["nothing pending", "global l1 <= n*eps", "pass limit"][r["stop_reason"]] for r in runs}),
        "pushed_edges_per_run": last["pushed_edges"],
        "residual_l1_fp64_at_stop": {k: float(f"{v:.6g}") for k, v in
                                     zip(["pr", "katz", "ev"] if measure == "all" else [measure], last["residual_l1"])},
        "kernel_ms_per_run": {c[0]: round(c[1] / len(runs), 3) for c in cands if c[2] and c[1] > 0},
        "roofline": roof,
    }
    rec["fp64_err"] = None if args.no_fp64 else fp64_errors(nat, np, graph, work, measure, env)
    return rec, excent_log


def graph_file(args, nat, work, scale, weights=False):
    graph = os.path.join(work, f"rmat{scale}_ef{args.edge_factor}_s{args.seed}{'_w' if weights else ''}.egr")
    t0 = time.time()
    if not os.path.exists(graph):
        gargs = ["--scale", scale, "--edge-factor", args.edge_factor, "--seed", args.seed]
        if weights:
            gargs.append("--weights")
        with native_stdout_to_stderr():
            rc = nat.run_entry("excent_rmat_main", gargs + [graph])
        assert rc == 0, "graph generation failed"
    log(f"graph {graph} ready in {time.time() - t0:.1f}s")
    return graph


def sub_s16(args, nat, np, work):
    """R-MAT scale 16 PageRank (eps 0.01, whole graph cached) with the CPU restatement run to
    convergence beside it."""
    from oracle import oracle as orc
    graph = graph_file(args, nat, work, 16)
    # events around the dense pull only: at this size a pass is tens of microseconds and events
    # around every kernel would cost a fifth of it
    rec, _ = bench_plain(args, nat, np, work, graph, 16, "pr", 20, 3, {}, "rmat16_pr", timing="--timing-pull")
    n, m = graph_size(nat, graph)
    rec.update({"metric": "converged GTEPS + time-to-eps", "higher_is_better": True, "dtype": "f32",
                "data": "synthetic", "config": {"workload": "R-MAT scale 16 ef 16 --measure pr, eps 0.01, whole "
                                                            "graph in HBM, converged to eps", "n": n, "m": m,
                                                "burst": args.burst or "auto", "burst_ahead": args.burst_ahead}})
    if not args.no_cpu:
        idx, adj = orc.egr_load(graph)
        st = plain_state(orc, np, idx, adj, "pr", {})
        rec["cpu_baseline"] = cpu_baseline_run(np, orc, idx, adj, *st, args.cpu_seconds, "R-MAT scale 16 PageRank")
    return rec


def sub_batched(args, nat, np, work):
    """Batched K=64 personalised PageRank at scale 26 against 64 plain runs."""
    a = argparse.Namespace(**vars(args))
    a.scale, a.k, a.warmup, a.batched_only = 26, 64, 1, False
    graph = graph_file(a, nat, work, 26)
    rec = bench_batched(a, nat, np, work, graph, sample=[0, a.k - 1])
    # the batched run's baseline is the 64 plain GPU runs above; a CPU sample of one column at
    # this size (--config batched --cpu) costs half a minute of the default run
    rec["cpu_baseline"] = None
    return rec


def sub_adjoint(args, nat, np, work):
    a = argparse.Namespace(**vars(args))
    a.scale = 22
    graph = graph_file(a, nat, work, 22, weights=True)
    # 6 finite-difference directions here (3 raw-weight, 3 teleport; about 10 s each at this
    # size); --config adjoint runs the full 32
    return bench_adjoint(a, nat, np, work, graph, fd_dirs=6)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="headline", choices=["headline", "batched", "ooc", "adjoint", "s16"],
                    help="headline: the scale-24 all-measure in-HBM run (the bench line), with the scale-16 "
                         "PageRank, adjoint and batched configurations as sub-records; batched: K=64 personalised "
                         "PageRank vs 64 plain runs; ooc: out-of-core; adjoint: differentiable personalised "
                         "PageRank forward + backward; s16: the scale-16 PageRank record alone")
    ap.add_argument("--sub", default="s16,adjoint,batched",
                    help="headline: sub-records to add, in order (comma list of s16, adjoint, batched; '' for none)")
    ap.add_argument("--sub-budget", type=float, default=420.0,
                    help="headline: seconds since start after which remaining sub-records are skipped")
    ap.add_argument("--scale", type=int, default=None, help="R-MAT scale (default 24; batched/ooc 26; adjoint 22)")
    ap.add_argument("--cache-mb", type=int, default=1024, help="ooc: device cache cap for the hot-cache run")
    ap.add_argument("--batched-only", action="store_true", help="batched: time the batched run only (no baseline)")
    ap.add_argument("--k", type=int, default=64, help="batched: personalisation columns (1..64)")
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--measure", default=None, choices=["pr", "katz", "ev", "all"],
                    help="measures of the plain runs (default: all; ooc: pr)")
    ap.add_argument("--stop", default=None,
                    help="the loop's stop rule: vertex, global or a list of measures that end on their sum "
                         "|r| <= n eps (default: katz,ev with --measure all, else vertex)")
    ap.add_argument("--max-passes", type=int, default=0,
                    help="stop a run after this many passes (reported as not converged; plain runs: 2000 when 0)")
    ap.add_argument("--cold", action="store_true",
                    help="ooc: also time the hot-cache run after dropping the file from the host page cache")
    ap.add_argument("--kernel", default="lb", choices=["lb", "group", "tune"])
    ap.add_argument("--pull", default="auto", help="dense passes by pull: auto | off | demand fraction")
    ap.add_argument("--hub-lds", type=int, default=1024, help="destinations privatised in LDS by the hub-piece push")
    ap.add_argument("--drop-measures", default="on", choices=["on", "off"],
                    help="dense passes leave out the measures with nothing pending")
    ap.add_argument("--timing", choices=["full", "pull", "off"], default="pull",
                    help="headline: HIP events around the dense pulls (the roofline's kernel; the default, "
                         "~0.5%% of the run), around every kernel (~1%%), or none (off: no roofline)")
    ap.add_argument("--burst-ahead", choices=["on", "off"], default="on",
                    help="submit the next burst before reading the current one")
    ap.add_argument("--burst", type=int, default=0,
                    help="passes per host synchronisation on a cached graph (0: automatic, 1: every pass)")
    ap.add_argument("--cpu-seconds", type=float, default=30.0,
                    help="CPU baseline: time limit of the restatement's run (it stops earlier when it converges)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 error report of the last run")
    ap.add_argument("--workdir", default=None)
    ap.add_argument("--logdir", default=None, help="directory of the runs' bin/excent reports (default: the workdir)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic_*.json"),
                    help="JSON files (a glob) with HBM bytes per launch from rocprofv3 --pmc runs, one workload each")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU design is described, not built", file=sys.stderr)
        sys.exit(2)
    t_begin = time.time()
    # stdout carries exactly one JSON line: everything else written to file descriptor 1 (the
    # native tools' C stdio included) goes to stderr, and the line to the saved descriptor
    sys.stdout.flush()
    json_fd = os.dup(1)
    os.dup2(2, 1)

    def emit(rec):
        sys.stdout.flush()
        os.write(json_fd, (json.dumps(rec) + "\n").encode())

    if args.scale is None:
        args.scale = {"headline": 24, "batched": 26, "ooc": 26, "adjoint": 22, "s16": 16}[args.config]
    if args.measure is None:
        args.measure = "pr" if args.config in ("ooc", "s16") else "all"

    import numpy as np

    import excent_amd.native as nat

    work = args.workdir or tempfile.mkdtemp(prefix="excent_bench_", dir=os.environ.get("TMPDIR", "/tmp"))
    os.makedirs(work, exist_ok=True)
    if args.config == "s16":
        emit(sub_s16(args, nat, np, work))
        return
    graph = graph_file(args, nat, work, args.scale, weights=args.config == "adjoint")
    if args.config != "headline":
        out = {"batched": bench_batched, "ooc": bench_ooc, "adjoint": bench_adjoint}[args.config](args, nat, np, work, graph)
        emit(out)
        return
    env = constants(nat, graph)
    rec, excent_log = bench_plain(args, nat, np, work, graph, args.scale, args.measure, args.steps, args.warmup, env,
                                  "headline", timing={"full": "--timing", "pull": "--timing-pull", "off": None}[args.timing])

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        log("cpu baseline: loading graph")
        idx, adj = orc.egr_load(graph)
        st = plain_state(orc, np, idx, adj, args.measure, env)
        cpu = cpu_baseline_run(np, orc, idx, adj, *st, args.cpu_seconds,
                               f"R-MAT scale {args.scale} --measure {args.measure}")
        del idx, adj, st

    n_v, m_e = graph_size(nat, graph)
    out = {
        "metric": METRIC, "value": rec["value"], "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": rec["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"R-MAT scale {args.scale} ef {args.edge_factor} --measure {args.measure}, "
                               f"whole graph in HBM, converged to eps", "n": n_v, "m": m_e,
                   "lambda1": env["LAMBDA"], "alpha": env["ALPHA"], "eps_pr": 0.01, "eps_katz": 1e-6,
                   "eps_ev": 1e-6, "stop": stop_rule(args, args.measure), "push_kernel": args.kernel, "pull": args.pull,
                   "drop_measures": args.drop_measures, "burst": args.burst or "auto",
                   "burst_ahead": args.burst_ahead, "timing": args.timing, "parallelism": "single GPU"},
    }
    out.update({k: v for k, v in rec.items() if k not in ("value", "unit", "steps", "warmup", "ms_per_step")})
    out["cpu_baseline"] = cpu

    subs = {}
    for name in [x for x in args.sub.split(",") if x]:
        el = time.time() - t_begin
        if el > args.sub_budget:
            subs[name] = {"skipped": f"time budget: {el:.0f}s of {args.sub_budget:.0f}s used before it"}
            log(f"sub-record {name} skipped ({el:.0f}s elapsed)")
            continue
        t0 = time.time()
        try:
            subs[name] = {"s16": sub_s16, "adjoint": sub_adjoint, "batched": sub_batched}[name](args, nat, np, work)
        except Exception as e:   # a failed sub-record must not lose the headline line
            subs[name] = {"error": f"{type(e).__name__}: {e}"}
        subs[name]["wall_s"] = round(time.time() - t0, 1)
        log(f"sub-record {name}: {time.time() - t0:.1f}s")
    if subs:
        out["configs"] = subs
    emit(out)


if __name__ == "__main__":
    main()

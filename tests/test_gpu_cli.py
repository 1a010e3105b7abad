"""bin/excent end to end on the GPU (its engine called in process): every cache and
fetch regime on small graphs, against the FP64 oracle, plus the report and dump
formats (src/excent.cu:756-805)."""
import os
import re

import numpy as np
import pytest

import excent_amd.native as nat
from oracle import oracle as orc
from tests.conftest import SAMPLE_EGR, SAMPLE_ENV, read_env

pytestmark = pytest.mark.gpu


def run(args, tmp_path, name="run"):
    log = str(tmp_path / f"{name}.log")
    rc = nat.excent([*args, "--log", log])
    text = open(log).read() if os.path.exists(log) else ""
    return rc, text


def pr_check(idx, adj, x, eps=0.01):
    ref, _, _, _ = orc.fp64_reference(idx, adj, orc.ORC_PR)
    n = len(idx) - 1
    err = np.abs(x.astype(np.float64) - ref).sum()
    assert err <= eps * n / 0.15, err
    c = orc.compare(ref, x)
    assert c["top100"] >= 90, c
    return c


def test_sample_pr_default(tmp_path):
    rc, text = run([SAMPLE_EGR, "--measure", "pr", "--dump", str(tmp_path / "o"), "--count-pushes",
                    "--stream-ratio", "0.1"], tmp_path)
    assert rc == 0, text
    assert "cache: auto -> chunks (the whole graph fits)" in text
    assert re.search(r"pass +0: pending 2000 v", text)
    assert "pass   0:" in text and "stream 1" in text.split("pass   0:")[1].split("\n")[0]
    assert re.search(r"passes \d+  measure pr \(n=2000\):  pr sum ", text)
    assert "edges pushed:" in text
    x = np.fromfile(str(tmp_path / "o.pr.bin"), dtype=np.float32)
    assert x.shape == (2000,)
    idx, adj = orc.egr_load(SAMPLE_EGR)
    pr_check(idx, adj, x)
    s = float(re.search(r"pr sum ([0-9.e+-]+)", text).group(1))
    assert abs(s - x.astype(np.float64).sum()) <= 1e-6 * abs(s)


@pytest.mark.parametrize("kernel", ["lb", "group"])
def test_sample_all(tmp_path, kernel):
    env = read_env(SAMPLE_ENV)
    args = [SAMPLE_EGR, "--measure", "all", "--alpha", env["ALPHA"], "--beta", 1, "--lambda", env["LAMBDA"],
            "--eps-pr", 0.01, "--eps-katz", env["EPS_KATZ"], "--eps-ev", env["EPS_EV"], "--dump", str(tmp_path / "a"),
            "--push-kernel", kernel, "--stream-ratio", "0.1"]
    rc, text = run(args, tmp_path)
    assert rc == 0, text
    idx, adj = orc.egr_load(SAMPLE_EGR)
    xs = {m: np.fromfile(str(tmp_path / f"a.{m}.bin"), np.float32) for m in ("pr", "katz", "ev")}
    pr_check(idx, adj, xs["pr"])
    kref = orc.fp64_reference(idx, adj, orc.ORC_KATZ, alpha=env["ALPHA"])[0]
    assert orc.compare(kref, xs["katz"])["l1_rel"] < 0.05
    eref = orc.fp64_reference(idx, adj, orc.ORC_EV)[0]
    assert orc.compare(eref, xs["ev"], unit=True)["l1_rel"] < 0.1


@pytest.mark.parametrize("mode", ["stream_only", "gather_only", "mixed", "hot", "preload", "zero_copy"])
def test_out_of_core_regimes(tmp_path, rmat14, mode):
    idx, adj = orc.egr_load(rmat14)
    base = [rmat14, "--measure", "pr", "--dump", str(tmp_path / "o"), "--chunk-medges", "0.02"]
    if mode == "stream_only":
        extra = ["--cache", "none", "--stream-ratio", "0"]
    elif mode == "gather_only":
        extra = ["--cache", "none", "--stream-ratio", "2"]
    elif mode == "mixed":
        extra = ["--cache", "none", "--stream-ratio", "0.2", "--dedup", "off"]
    elif mode == "hot":
        free = nat.ctypes.c_size_t(0)
        total = nat.ctypes.c_size_t(0)
        nat.F["exc_mem_info"](0, nat.ctypes.byref(free), nat.ctypes.byref(total))
        # leave about 1 MB for hot lists (after the context's own allocations, which are small)
        reserve_mb = free.value / (1 << 20) - 64
        extra = ["--cache", "hot", "--reserve-mb", f"{reserve_mb:.3f}", "--stream-ratio", "0.3"]
    elif mode == "preload":
        extra = ["--cache", "chunks", "--preload", "--stream-ratio", "0.5"]
    else:
        extra = ["--cache", "none", "--zero-copy", "on", "--stream-ratio", "0.3"]
    print("case", mode, flush=True)
    rc, text = run(base + extra, tmp_path, mode)
    assert rc == 0, text
    x = np.fromfile(str(tmp_path / "o.pr.bin"), np.float32)
    pr_check(idx, adj, x)
    lines = [l for l in text.splitlines() if l.startswith("pass ")]
    assert lines
    if mode == "stream_only":
        assert all(" gather 0 " in l for l in lines)
    if mode == "gather_only":
        assert all(" stream 0 " in l for l in lines)
        m = re.search(r"bytes shipped: stream ([0-9.]+) GB, gather ([0-9.]+) GB", text)
        assert float(m.group(1)) == 0.0 and float(m.group(2)) > 0
    if mode == "hot":
        assert "GPU degree cache:" in text
    if mode == "preload":
        assert " resident " in lines[0] and "stream 0" in lines[0]


PASS_RE = re.compile(r"pass +(\d+): pending (\d+) v, demand ratio [0-9.]+ \| resident (\d+) stream (\d+) gather (\d+) \| "
                     r"cached high-degree pending (\d+)")


@pytest.mark.parametrize("when,order", [("end", "sweep"), ("start", "sweep"), ("both", "sweep"), ("end", "ranges"),
                                        ("both", "ranges")])
def test_hot_cache_partial(tmp_path, rmat14, when, order):
    """The hot-vertex cache holding only part of the edges (--cache-mb caps its budget at
    0.25 MB of rmat14's ~1 MB of lists): the highest-degree lists are pushed from HBM at the
    start, the end or both ends of each pass (src/excent.cu:683, 745), as one sweep or one
    range per chunk (612-614), while the other pending vertices are streamed or gathered."""
    idx, adj = orc.egr_load(rmat14)
    n, m = len(idx) - 1, len(adj)
    args = [rmat14, "--measure", "pr", "--dump", str(tmp_path / "o"), "--chunk-medges", "0.02", "--cache", "hot",
            "--cache-mb", "0.25", "--stream-ratio", "0.3", "--hot-when", when, "--hot-order", order]
    print("case hot cache", when, order, flush=True)
    rc, text = run(args, tmp_path)
    assert rc == 0, text
    mo = re.search(r"GPU degree cache: (\d+) vertices with degree >= (\d+) \(([0-9.]+)% of edges", text)
    assert mo, text
    nh, dmin, pct = int(mo.group(1)), int(mo.group(2)), float(mo.group(3))
    deg = np.diff(idx.astype(np.int64))
    # the cache holds the highest-degree lists: every vertex above the threshold degree
    assert 0 < nh < n and 0 < pct < 60 and (deg > dmin).sum() <= nh <= (deg >= dmin).sum()
    assert (4 * deg[deg > dmin] + 16).sum() <= 0.25 * (1 << 20)
    passes = [tuple(int(g) for g in p.groups()) for p in PASS_RE.finditer(text)]
    assert len(passes) > 2
    # after pass 0, pending hot vertices are counted apart and pushed from the cache, and the
    # other pending vertices are fetched from the file
    assert any(p[0] > 0 and p[5] > 0 for p in passes), text
    assert any(p[0] > 0 and p[3] + p[4] > 0 for p in passes), text
    assert all(p[2] == 0 for p in passes)   # nothing is resident in hot mode
    pr_check(idx, adj, np.fromfile(str(tmp_path / "o.pr.bin"), np.float32))


def test_hot_rank(tmp_path, rmat14, capfd):
    """--hot-rank (src/excent.cu:459-483, 583-585): a --trace-pending run writes the passes each
    vertex pended (uint16 per vertex), and a second run caches the lists ranked by that count
    within a 0.25 MB budget.  The cached set is the oracle's ranked selection for the same
    scores and budget; the run pushes cached vertices from HBM, fetches the rest, and
    converges to the FP64 bound.  A file shorter than n entries is refused."""
    idx, adj = orc.egr_load(rmat14)
    n = len(idx) - 1
    tf = str(tmp_path / "t.bin")
    print("case hot rank: trace run", flush=True)
    rc, text = run([rmat14, "--measure", "pr", "--trace-pending", tf, "--stream-ratio", "0.3",
                    "--chunk-medges", "0.02"], tmp_path, "trace")
    assert rc == 0, text
    score = np.fromfile(tf, np.uint16)
    assert score.shape == (n,) and score.max() > 1 and len(np.unique(score)) > 2
    print("case hot rank: ranked run", flush=True)
    args = [rmat14, "--measure", "pr", "--dump", str(tmp_path / "o"), "--chunk-medges", "0.02", "--cache", "hot",
            "--cache-mb", "0.25", "--stream-ratio", "0.3", "--hot-rank", tf]
    rc, text = run(args, tmp_path)
    assert rc == 0, text
    mo = re.search(r"GPU cache ranked by passes pending \((.+)\): (\d+) vertices \(([0-9.]+)% of edges", text)
    assert mo and mo.group(1) == tf, text
    ids, off, bm, _ = orc.hot_select(idx, 262144, score)
    assert int(mo.group(2)) == len(ids) > 0
    assert abs(float(mo.group(3)) - 100.0 * int(off[-1]) / len(adj)) < 0.06
    passes = [tuple(int(g) for g in p.groups()) for p in PASS_RE.finditer(text)]
    assert any(p[0] > 0 and p[5] > 0 for p in passes), text
    assert any(p[0] > 0 and p[3] + p[4] > 0 for p in passes), text
    pr_check(idx, adj, np.fromfile(str(tmp_path / "o.pr.bin"), np.float32))
    short = str(tmp_path / "short.bin")
    score[: n // 2].tofile(short)
    print("case hot rank: short file", flush=True)
    rc, text = run(args[:-1] + [short], tmp_path, "short")
    assert rc != 0
    assert f"--hot-rank: cannot read {n} entries from {short}" in capfd.readouterr().err + text


def test_tau_calibration(tmp_path, rmat14, monkeypatch):
    """--stream-ratio auto (the default): tau is measured on the middle chunk, written to
    ~/.cache/excent/<graph>.<n>.<m>.tau as "%.6f %.4f" (src/excent.cu:518-579), read back by
    the next run, and measured again with --recalibrate; tau stays within [0.005, 1] and the
    results within the PageRank bound."""
    monkeypatch.setenv("HOME", str(tmp_path))
    idx, adj = orc.egr_load(rmat14)
    n, m = len(idx) - 1, len(adj)
    base = [rmat14, "--measure", "pr", "--dump", str(tmp_path / "o"), "--chunk-medges", "0.02", "--cache", "none"]
    tau_file = tmp_path / ".cache" / "excent" / f"{os.path.basename(rmat14)}.{n}.{m}.tau"
    calib = re.compile(r"fetch rule: tau = ([0-9.]+) \(stream ([0-9.]+)s; gather ([0-9.]+)s / ([0-9.]+)s / ([0-9.]+)s "
                       r"at demand ([0-9.]+) / ([0-9.]+) / ([0-9.]+)\)  calibrate ([0-9.]+)s")
    for step, extra in (("first", []), ("cached", []), ("recalibrate", ["--recalibrate"])):
        print("case tau", step, flush=True)
        rc, text = run(base + extra, tmp_path, step)
        assert rc == 0, text
        if step == "cached":
            mo = re.search(r"fetch rule: tau = ([0-9.]+) \(measured earlier on this machine in ([0-9.]+)s, (.+)\)", text)
            assert mo and mo.group(3) == str(tau_file), text
            assert float(mo.group(1)) == pytest.approx(tau, abs=5e-4)
        else:
            mo = calib.search(text)
            assert mo, text
            tau = float(mo.group(1))
            assert 0.005 <= tau <= 1.0
            demands = [float(mo.group(i)) for i in (6, 7, 8)]
            assert demands == sorted(demands) and demands[2] <= 1.0
            body = tau_file.read_text()
            assert re.fullmatch(r"\d+\.\d{6} \d+\.\d{4}\n", body), body
            assert float(body.split()[0]) == pytest.approx(tau, abs=5e-4)
        pr_check(idx, adj, np.fromfile(str(tmp_path / "o.pr.bin"), np.float32))


def test_rmat16_pagerank_config(tmp_path, rmat16):
    """R-MAT scale 16 PageRank (d = 0.85, eps 0.01, r0 = 0.15, whole graph cached): with the
    dense passes (default) and with the push alone, against the FP64 solution -- the exact
    bound ||x - x*||_1 <= ||b + W x - x||_1 / (1 - d) with the residual of the returned x in
    FP64, the eps bound, and top-100 overlap."""
    idx, adj = orc.egr_load(rmat16)
    n = len(idx) - 1
    assert n == 65536 and 900_000 < len(adj) <= 1 << 20
    ref, _, _, _ = orc.fp64_reference(idx, adj, orc.ORC_PR, tol=1e-12, max_iter=3000)
    for pull in ("auto", "off"):
        print("case rmat16 pr pull", pull, flush=True)
        rc, text = run([rmat16, "--measure", "pr", "--cache", "chunks", "--preload", "--pull", pull, "--dump",
                        str(tmp_path / "o"), "--count-pushes"], tmp_path, pull)
        assert rc == 0, text
        assert "GPU edge cache: 1/1 chunks resident (100.0% of edges)" in text
        x = np.fromfile(str(tmp_path / "o.pr.bin"), np.float32)
        err = np.abs(x.astype(np.float64) - ref).sum()
        res = orc.residual_l1(idx, adj, orc.ORC_PR, 0.85, 0.15, x)
        assert err <= res / 0.15 * (1 + 1e-6) + 1e-6, (err, res)
        assert err <= 0.01 * n / 0.15
        c = orc.compare(ref, x)
        assert c["top100"] >= 95, c


@pytest.fixture(scope="module")
def rmat14_env(rmat14, tmp_path_factory):
    """rmat14's constants: lambda1 by FP64 power iteration, alpha = 0.5 / lambda1."""
    out = str(tmp_path_factory.mktemp("env") / "rmat14.env")
    assert nat.run_entry("excent_params_main", [rmat14, out, "--directed", "--alpha-factor", "0.5", "--seed", "7"]) == 0
    return read_env(out)


@pytest.mark.parametrize("mode", ["stream_only", "gather_only", "mixed", "hot", "zero_copy"])
def test_out_of_core_regimes_all_measures(tmp_path, rmat14, rmat14_env, mode):
    """The fused three-measure push (16-byte rows) in every out-of-core regime: PageRank,
    Katz and eigenvector against FP64, every |r| under its eps.  --max-passes bounds the
    run, so a regime that stopped converging fails here instead of running on."""
    idx, adj = orc.egr_load(rmat14)
    env = rmat14_env
    eps = 1e-5
    base = [rmat14, "--measure", "all", "--alpha", env["ALPHA"], "--beta", 1, "--lambda", env["LAMBDA"],
            "--eps-katz", eps, "--eps-ev", eps, "--ev-init", "indeg", "--dump", str(tmp_path / "o"),
            "--chunk-medges", "0.02", "--max-passes", "2000"]
    extra = {"stream_only": ["--cache", "none", "--stream-ratio", "0"],
             "gather_only": ["--cache", "none", "--stream-ratio", "2"],
             "mixed": ["--cache", "none", "--stream-ratio", "0.2", "--dedup", "off"],
             "zero_copy": ["--cache", "none", "--zero-copy", "on", "--stream-ratio", "0.3"]}.get(mode)
    if mode == "hot":
        free, total = nat.ctypes.c_size_t(0), nat.ctypes.c_size_t(0)
        nat.F["exc_mem_info"](0, nat.ctypes.byref(free), nat.ctypes.byref(total))
        extra = ["--cache", "hot", "--reserve-mb", f"{free.value / (1 << 20) - 64:.3f}", "--stream-ratio", "0.3"]
    print("case all measures", mode, flush=True)
    rc, text = run(base + extra, tmp_path, mode)
    assert rc == 0, text
    passes = int(re.search(r"passes (\d+)  measure all", text).group(1))
    assert passes < 2000, "did not converge"
    x = {k: np.fromfile(str(tmp_path / f"o.{k}.bin"), np.float32) for k in ("pr", "katz", "ev")}
    pr_check(idx, adj, x["pr"])
    kref = orc.fp64_reference(idx, adj, orc.ORC_KATZ, alpha=env["ALPHA"], tol=1e-12, max_iter=3000)[0]
    assert orc.compare(kref, x["katz"])["l1_rel"] < 1e-3
    eref = orc.fp64_reference(idx, adj, orc.ORC_EV, tol=1e-12, max_iter=3000)[0]
    assert orc.compare(eref, x["ev"], unit=True)["l1_rel"] < 1e-2


def test_trace_pending(tmp_path, rmat12):
    tf = str(tmp_path / "trace.u16")
    rc, text = run([rmat12, "--measure", "pr", "--trace-pending", tf, "--stream-ratio", "0.1"], tmp_path)
    assert rc == 0, text
    t = np.fromfile(tf, np.uint16)
    idx, _ = orc.egr_load(rmat12)
    passes = int(re.search(r"passes (\d+)", text).group(1))
    assert t.shape == (len(idx) - 1,) and t.max() <= passes and t.min() >= 1   # every vertex pending in pass 0


def test_reference_tool_matches_oracle(tmp_path, rmat12):
    """bin/centrality_reference's FP64 engine against the oracle's FP64 restatement."""
    for path in (SAMPLE_EGR, rmat12):
        idx, adj = orc.egr_load(path)
        for kind, kw in ((nat.EXC_PR, {}), (nat.EXC_KATZ, {"alpha": 0.05}), (nat.EXC_EV, {})):
            got, _, _, _ = nat.fp64_solve(idx, adj, kind, max_iter=3000, tol=1e-12, **kw)
            ref, _, _, _ = orc.fp64_reference(idx, adj, kind, max_iter=3000, tol=1e-12, **kw)
            assert np.allclose(got, ref, rtol=1e-9, atol=1e-12), (path, kind)


def test_fp64_residual_matches_oracle(rmat12):
    """exc_fp64_residual (the bench's FP64 ||r||_1) against the oracle's CPU residual, plus
    the closed form at x = 0 (every vertex contributes |c|)."""
    idx, adj = orc.egr_load(rmat12)
    n = len(idx) - 1
    x = np.random.default_rng(3).random(n).astype(np.float32) * 2
    with nat.Fp64Solver(idx, adj) as S:
        for kind, c, s, dd in ((orc.ORC_PR, 0.15, 0.85, 1), (orc.ORC_KATZ, 1.0, 0.01, 0), (orc.ORC_EV, 0.0, 0.02, 0)):
            got = S.residual(x, c, s, dd)
            want = orc.residual_l1(idx, adj, kind, s, c, x)
            assert abs(got - want) <= 1e-10 * want, (kind, got, want)
        assert S.residual(np.zeros(n, np.float32), 0.15, 0.85, 1) == pytest.approx(0.15 * n, rel=1e-12)


def test_fp64_solve_resumes_from_x0(rmat12):
    """A solve resumed from an earlier solve's output equals one solve of the summed
    iterations (PageRank and Katz Jacobi; eigenvector power iteration, whose resumed run
    starts from the unit vector the first one returned), and the resumed PageRank matches
    the oracle's FP64 reference."""
    idx, adj = orc.egr_load(rmat12)
    with nat.Fp64Solver(idx, adj) as S:
        for kind, kw in ((nat.EXC_PR, {}), (nat.EXC_KATZ, {"alpha": 0.002, "beta": 1.0}), (nat.EXC_EV, {})):
            whole, _, lam_w, it_w = S.solve(kind, tol=0.0, max_iter=14, **kw)
            part, _, _, it_a = S.solve(kind, tol=0.0, max_iter=6, **kw)
            rest, _, lam_r, it_b = S.solve(kind, tol=0.0, max_iter=8, x0=part, **kw)
            assert (it_w, it_a, it_b) == (14, 6, 8)
            np.testing.assert_allclose(rest, whole, rtol=1e-11, atol=1e-300)
            if kind == nat.EXC_EV:
                assert lam_r == pytest.approx(lam_w, rel=1e-11)
        x, rel, _, it = S.solve(nat.EXC_PR, tol=1e-12, max_iter=400)
        part, _, _, _ = S.solve(nat.EXC_PR, tol=0.0, max_iter=30)
        y, rel2, _, it2 = S.solve(nat.EXC_PR, tol=1e-12, max_iter=400, x0=part)
        assert it2 < it and rel2 <= 1e-12
        ref = orc.fp64_reference(idx, adj, orc.ORC_PR, tol=1e-12, max_iter=400)[0]
        np.testing.assert_allclose(y, ref, rtol=1e-9)


def test_params_tool_reproduces_sample_env(tmp_path):
    out = str(tmp_path / "s.env")
    rc = nat.run_entry("excent_params_main", [SAMPLE_EGR, out])
    assert rc == 0
    got, want = read_env(out), read_env(SAMPLE_ENV)
    for k in ("LAMBDA", "ALPHA", "EPS_KATZ", "EPS_EV"):
        assert abs(got[k] - want[k]) <= 1e-6 * abs(want[k]), k


def test_run_sh_end_to_end(tmp_path):
    """run.sh as the reference's users call it: --sample (this build's sample graph and its
    constants, excent_amd/data) with --measure=all, and a copy of the reference's sample
    fixture without a .env, for which the wrapper writes the constants next to the graph
    with bin/excent_params (the fixture's own .env digits) and passes them to bin/excent
    (--measure=katz)."""
    import shutil
    import subprocess
    from tests.conftest import ROOT
    sh = os.path.join(ROOT, "run.sh")
    data = os.path.join(ROOT, "excent_amd", "data")
    print("case run.sh sample all", flush=True)
    r = subprocess.run([sh, "--measure=all", "--sample", "--", "--dump", str(tmp_path / "a"), "--stream-ratio", "0.1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert re.search(r"measure all \(n=2048\)", r.stdout), r.stdout
    idx, adj = orc.egr_load(os.path.join(data, "sample.egr"))
    pr_check(idx, adj, np.fromfile(str(tmp_path / "a.pr.bin"), np.float32))
    env = read_env(os.path.join(data, "sample.env"))
    kref = orc.fp64_reference(idx, adj, orc.ORC_KATZ, alpha=env["ALPHA"])[0]
    assert orc.compare(kref, np.fromfile(str(tmp_path / "a.katz.bin"), np.float32))["l1_rel"] < 0.05
    g = str(tmp_path / "g.egr")
    shutil.copy(SAMPLE_EGR, g)
    print("case run.sh sample copy, katz (constants computed)", flush=True)
    r = subprocess.run([sh, "--measure=katz", f"--graph={g}", "--", "--dump", str(tmp_path / "k")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert os.path.exists(str(tmp_path / "g.env")) and "Computing the constants" in r.stdout
    genv, fenv = read_env(str(tmp_path / "g.env")), read_env(SAMPLE_ENV)
    for k in ("LAMBDA", "ALPHA", "EPS_KATZ", "EPS_EV"):
        assert abs(genv[k] - fenv[k]) <= 1e-6 * abs(fenv[k]), k
    sidx, sadj = orc.egr_load(SAMPLE_EGR)
    kref = orc.fp64_reference(sidx, sadj, orc.ORC_KATZ, alpha=genv["ALPHA"])[0]
    assert orc.compare(kref, np.fromfile(str(tmp_path / "k.katz.bin"), np.float32))["l1_rel"] < 0.05


def test_fp64_solver_strided_grid(tmpdir_factory):
    """A graph with more vertices than the FP64 pull's capped grid has 16-lane groups
    (2^19 > 256 CUs x 64 blocks x 16 groups), so every group takes several vertices: the
    solves still equal the oracle's FP64 restatement, and the residual its CPU sum."""
    from tests.conftest import rmat_graph
    g = rmat_graph(tmpdir_factory, 19, seed=5)
    idx, adj = orc.egr_load(g)
    n = len(idx) - 1
    print("case fp64 strided grid, n =", n, flush=True)
    with nat.Fp64Solver(idx, adj) as S:
        for kind, kw in ((nat.EXC_PR, {}), (nat.EXC_KATZ, {"alpha": 1e-3}), (nat.EXC_EV, {})):
            got, _, _, it = S.solve(kind, tol=1e-10, max_iter=300, **kw)
            ref, it_ref, _, _ = orc.fp64_reference(idx, adj, kind, max_iter=300, tol=1e-10, **kw)
            assert abs(it - it_ref) <= 1, (kind, it, it_ref)
            np.testing.assert_allclose(got, ref, rtol=1e-8, atol=1e-12)
        x = np.random.default_rng(4).random(n).astype(np.float32)
        want = orc.residual_l1(idx, adj, orc.ORC_PR, 0.85, 0.15, x)
        assert abs(S.residual(x, 0.15, 0.85, 1) - want) <= 1e-10 * want


def test_gpu_ballast_caps_the_budget(tmp_path, rmat14):
    """bin/gpu_ballast (tools/gpu_ballast.cu:13-41) in a second process takes 8 GiB of device
    memory and holds it until SIGTERM: the free memory this process sees drops by that much,
    and a --cache auto run whose --reserve-mb exceeds the remaining free memory falls back to
    the high-degree cache, where the same run without the ballast caches the whole graph."""
    import queue
    import signal
    import subprocess
    import threading
    from tests.conftest import ROOT

    def free_bytes():
        free, total = nat.ctypes.c_size_t(0), nat.ctypes.c_size_t(0)
        assert nat.F["exc_mem_info"](0, nat.ctypes.byref(free), nat.ctypes.byref(total)) == 0
        return free.value

    gib, mib = 1 << 30, 1 << 20
    idx, adj = orc.egr_load(rmat14)
    free0 = free_bytes()
    budget = free0 - 8 * gib
    reserve_mb = (budget + 64 * mib) // mib   # 64 MiB more than stays free under the ballast
    args = [rmat14, "--measure", "pr", "--dump", str(tmp_path / "o"), "--chunk-medges", "0.02", "--stream-ratio", "0.3",
            "--reserve-mb", reserve_mb]
    print("case ballast: control run", flush=True)
    rc, text = run(args, tmp_path, "control")
    assert rc == 0 and "cache: auto -> chunks (the whole graph fits)" in text, text
    print("case ballast: start", flush=True)
    p = subprocess.Popen([os.path.join(ROOT, "bin", "gpu_ballast"), f"{budget / gib:.6f}"], stdout=subprocess.PIPE,
                         stderr=subprocess.PIPE, text=True)
    try:
        q = queue.Queue()
        threading.Thread(target=lambda: q.put(p.stdout.readline()), daemon=True).start()
        line = q.get(timeout=120)
        mo = re.match(r"READY free=([0-9.]+) GB \(budget ([0-9.]+) GB\) of ([0-9.]+) GB", line)
        assert mo, (line, p.poll())
        assert abs(float(mo.group(2)) - budget / gib) < 0.01
        held_free = free_bytes()
        # the two processes' views differ by their own context and allocator state (35 MB seen)
        assert budget - gib <= held_free <= budget + 256 * mib, (held_free, budget)
        print("case ballast: capped run", flush=True)
        rc, text = run(args, tmp_path, "capped")
        assert rc == 0 and "cache: auto -> high-degree lists" in text, text
        pr_check(idx, adj, np.fromfile(str(tmp_path / "o.pr.bin"), np.float32))
        p.send_signal(signal.SIGTERM)
        assert p.wait(timeout=60) == 0, p.stderr.read()
    finally:
        if p.poll() is None:
            p.kill()
            p.wait()
    assert free_bytes() >= free0 - gib

"""CPU-side checks of the native product: every C-ABI symbol declared in include/*.h is
exported by the built libraries (no compute calls), the R-MAT generator, the .mtx
converter (tools/efficient_converter.py behaviour) and bin/excent's flag validation
(src/excent.cu:269-316), which runs before any device call."""
import os
import re
import subprocess

import numpy as np
import pytest

import excent_amd.native as nat
from oracle import oracle as orc
from tests.conftest import ROOT, SAMPLE_EGR, SAMPLE_MTX


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:exc|excent)_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("excent_hip.h", "libexcent_hip.so"), ("excent.h", "libexcent.so")])
def test_abi_exports_every_declared_symbol(header, lib):
    names = declared(header)
    assert len(names) > 10
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(nat.LIBDIR, lib)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [s for s in names if s not in exported]
    assert not missing, missing
    # and the bindings cover them
    table = nat.HIP_SIGNATURES if header == "excent_hip.h" else nat.HOST_SIGNATURES
    assert not [s for s in names if s not in table], [s for s in names if s not in table]


def test_abi_rejects_bad_arguments_without_a_device():
    # argument validation happens before any device call
    rc = nat.F["exc_create"](None, None, None)
    assert rc == nat.EXC_EINVAL and "null" in nat.F["exc_last_error"]().decode()


def test_binaries_built():
    for b in ("excent", "excent_rmat", "egr_convert", "centrality_reference", "excent_params"):
        assert os.access(os.path.join(nat.BINDIR, b), os.X_OK), b


def rmat(tmp_path, *args, name="g.egr"):
    p = str(tmp_path / name)
    assert nat.run_entry("excent_rmat_main", [*args, p]) == 0
    return p


def test_rmat_properties(tmp_path):
    p = rmat(tmp_path, "--scale", 12, "--seed", 5)
    idx, adj = orc.egr_load(p)
    n = len(idx) - 1
    assert n == 4096 and 0.7 * 16 * n < len(adj) <= 16 * n
    deg = np.diff(idx.astype(np.int64))
    src = np.repeat(np.arange(n), deg)
    assert np.all(adj >= 0) and np.all(adj < n) and np.all(src != adj)
    for v in range(0, n, 97):
        lst = adj[idx[v]:idx[v + 1]]
        assert np.all(np.diff(lst) > 0)   # sorted, no duplicates
    # deterministic, seed-dependent
    p2 = rmat(tmp_path, "--scale", 12, "--seed", 5, name="g2.egr")
    assert open(p, "rb").read() == open(p2, "rb").read()
    p3 = rmat(tmp_path, "--scale", 12, "--seed", 6, name="g3.egr")
    assert open(p, "rb").read() != open(p3, "rb").read()
    # skew: R-MAT degrees are heavy-tailed
    assert deg.max() > 20 * deg.mean()


def test_rmat_permutation_is_a_relabelling(tmp_path):
    a = orc.egr_load(rmat(tmp_path, "--scale", 11, "--seed", 9, name="p.egr"))
    b = orc.egr_load(rmat(tmp_path, "--scale", 11, "--seed", 9, "--no-permute", name="np.egr"))
    da, db = np.diff(a[0].astype(np.int64)), np.diff(b[0].astype(np.int64))
    assert len(a[1]) == len(b[1]) and np.array_equal(np.sort(da), np.sort(db))
    ia, ib = np.bincount(a[1], minlength=2048), np.bincount(b[1], minlength=2048)
    assert np.array_equal(np.sort(ia), np.sort(ib))
    # without the relabelling the hubs sit at the low ids
    assert np.argmax(db) < 64


def test_rmat_sidecars(tmp_path):
    p = rmat(tmp_path, "--scale", 10, "--seed", 2, "--csc", "--weights", "--wide", "on")
    idx, adj = orc.egr_load(p)
    tidx, tadj = orc.egr_load(p + ".csc.egr")
    n = len(idx) - 1
    src = np.repeat(np.arange(n), np.diff(idx.astype(np.int64)))
    fwd = set(zip(src.tolist(), adj.tolist()))
    tsrc = np.repeat(np.arange(n), np.diff(tidx.astype(np.int64)))
    assert set(zip(tadj.tolist(), tsrc.tolist())) == fwd
    w = np.fromfile(p + ".w.f32", np.float32)
    assert len(w) == len(adj) and w.min() >= 0.5 and w.max() < 1.5
    assert open(p, "rb").read(8)[4:8] == b"\xff\xff\xff\xff"


def test_converter_sample_byte_exact(tmp_path):
    out = str(tmp_path / "s.egr")
    assert nat.run_entry("excent_convert_main", [SAMPLE_MTX, out]) == 0
    assert open(out, "rb").read() == open(SAMPLE_EGR, "rb").read()


def test_converter_edge_cases(tmp_path):
    mtx = tmp_path / "e.mtx"
    # comments before the size line, no edge count, a source beyond n dropped, a
    # duplicate and a self-loop kept in file order
    mtx.write_text("%%MatrixMarket matrix coordinate pattern general\n% comment\n3 3\n1 2\n3 1\n1 2\n2 2\n4 1\n1 3\n")
    out = str(tmp_path / "e.egr")
    assert nat.run_entry("excent_convert_main", [str(mtx), out]) == 0
    idx, adj = orc.egr_load(out)
    assert idx.tolist() == [0, 3, 4, 5] and adj.tolist() == [1, 1, 2, 1, 0]


@pytest.mark.parametrize("args,msg", [
    (["--measure", "pagerank"], "--measure must be pr, katz, ev or all"),
    (["--measure", "katz"], "katz needs --alpha and --eps-katz"),
    (["--measure", "ev", "--lambda", "3"], "ev needs --lambda and --eps-ev"),
    (["--measure", "pr", "--cache", "gpu"], "--cache must be auto, chunks, hot or none"),
    (["--measure", "pr", "--dedup", "maybe"], "--dedup must be on or off"),
    (["--measure", "pr", "--frobnicate"], "unknown option --frobnicate"),
    (["--measure", "pr", "--max-passes"], "missing value for --max-passes"),
    (["--measure", "pr", "--push-kernel", "x"], "--push-kernel must be lb, group or tune"),
    (["--measure", "katz", "--ppr-column", "0"], "--ppr-column needs --measure pr"),
    (["--measure", "pr", "--ppr-column", "0"], "--ppr-column needs one of --ppr-seeds FILE or --ppr-gen"),
    (["--measure", "ppr", "--ppr-gen", "4:2:1:3"], "ppr needs --eps-ppr"),
])
def test_cli_validation(args, msg, capfd):
    rc = nat.excent([SAMPLE_EGR, *args])
    assert rc == 1
    assert msg in capfd.readouterr().err


def test_cli_bad_files(tmp_path, capfd):
    assert nat.excent([str(tmp_path / "missing.egr"), "--measure", "pr"]) == 1
    bad = tmp_path / "bad.egr"
    bad.write_bytes(np.array([0, 5], np.int32).tobytes())
    assert nat.excent([str(bad), "--measure", "pr"]) == 1
    assert "bad header" in capfd.readouterr().err
    short = tmp_path / "short.egr"
    short.write_bytes(open(SAMPLE_EGR, "rb").read()[:40000])
    assert nat.excent([str(short), "--measure", "pr"]) == 1
    assert "shorter than its header" in capfd.readouterr().err
    assert nat.excent([]) == 1


def test_cli_binary_usage():
    r = subprocess.run([os.path.join(nat.BINDIR, "excent")], capture_output=True, text=True)
    assert r.returncode == 1 and "usage:" in r.stderr


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU only: --gpus N > 1 exits with a message, before any GPU work."""
    import subprocess
    import sys
    from tests.conftest import ROOT
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "4"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU only" in r.stderr and r.stdout == ""


@pytest.mark.parametrize("args,msg", [
    ([], "--measure=pr|katz|ev|all is required"),
    (["--measure=foo", "--sample"], "--measure must be pr, katz, ev, or all, got 'foo'"),
    (["--measure=pr"], "--graph is required (or use --sample)"),
    (["--measure=pr", "--graph=/nonexistent/g.egr"], "graph file not found: /nonexistent/g.egr"),
    (["--measure=pr", "--sample", "--bogus"], "unrecognized argument: --bogus"),
])
def test_run_sh_validation(args, msg):
    """run.sh rejects what the reference's run.sh rejects (run.sh:42-67), with exit 1,
    before building or touching a GPU."""
    r = subprocess.run([os.path.join(ROOT, "run.sh"), *args], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and f"ERROR: {msg}" in r.stderr, r.stderr
    r = subprocess.run([os.path.join(ROOT, "run.sh"), "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "usage:" in r.stderr


def test_bench_top_overlap_known_answers():
    """bench.py's top-k overlap (reported at k = 100 and 1000 beside the FP64 errors)."""
    import bench
    a = np.arange(5000, dtype=np.float64)
    assert bench.top_overlap(np, a, a, 1000) == 1000
    assert bench.top_overlap(np, a, -a, 1000) == 0          # reversed ranking
    b = a.copy()
    b[[4999, 10]] = b[[10, 4999]]                            # the largest swapped out of the top
    assert bench.top_overlap(np, a, b, 1000) == 999
    assert bench.top_overlap(np, a[:50], a[:50], 1000) == 50   # k capped at the length


def test_bench_kendall_top():
    """bench.py's tau@1000 against the pairwise count over the reference's top t, on tiny
    inputs, and scipy's tau (the same as tau-a when nothing ties)."""
    import bench
    from scipy.stats import kendalltau
    rng = np.random.default_rng(7)
    ref = rng.random(3000)
    assert bench.kendall_top(np, ref, ref) == 1.0
    assert bench.kendall_top(np, -ref, ref) == -1.0
    x = ref + rng.normal(0, 0.02, ref.size)
    top = np.argsort(-ref, kind="stable")[:1000]
    assert abs(bench.kendall_top(np, x, ref) - kendalltau(ref[top], x[top])[0]) < 1e-12
    ref6, x6 = np.array([5.0, 4, 3, 2, 1, 0]), np.array([5.0, 3, 4, 2, 2, 9])
    t, c, d = 4, 0, 0   # top 4 of ref6 = indices 0..3; pairs (1,2) discordant, the rest concordant
    for i in range(t):
        for j in range(i + 1, t):
            p = (ref6[i] - ref6[j]) * (x6[i] - x6[j])
            c, d = c + (p > 0), d + (p < 0)
    assert (c, d) == (5, 1)
    assert bench.kendall_top(np, x6, ref6, t) == (c - d) / 6
    assert bench.kendall_top(np, x6[:1], ref6[:1]) == 1.0


def test_build_sample_graph_and_constants():
    """run.sh --sample's graph is this build's own: the seeded symmetric R-MAT the generator
    writes (same bytes when written again), with the constants bin/excent_params wrote for it
    (lambda1 checked against a dense symmetric eigensolve, alpha = 0.85 / lambda1)."""
    import filecmp
    import tempfile
    from tests.conftest import ROOT, read_env
    data = os.path.join(ROOT, "excent_amd", "data")
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "s.egr")
        assert nat.run_entry("excent_rmat_main", ["--scale", 11, "--edge-factor", 8, "--seed", 2026, "--symmetric",
                                                  "--wide", "on", p]) == 0
        assert filecmp.cmp(p, os.path.join(data, "sample.egr"), shallow=False)
    idx, adj = orc.egr_load(os.path.join(data, "sample.egr"))
    n = len(idx) - 1
    A = np.zeros((n, n))
    A[np.repeat(np.arange(n), np.diff(idx.astype(np.int64))), adj] = 1
    assert np.array_equal(A, A.T)
    lam = np.linalg.eigvalsh(A)[-1]
    env = read_env(os.path.join(data, "sample.env"))
    assert abs(env["LAMBDA"] - lam) <= 1e-8 * lam
    assert abs(env["ALPHA"] - 0.85 / env["LAMBDA"]) <= 1e-5 * env["ALPHA"]   # written with 6 digits

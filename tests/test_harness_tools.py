"""The harness programs around bin/excent on the CPU: bin/egr_info (tools/egr_info.cpp:10-29)
against a numpy restatement of its report, with the sampled edges drawn by a Python
MT19937-64 pinned to the C++ standard's known answer; bin/evict_file (tools/evict_file.c:
36-65) on a scratch file; the usage errors of all three, which return before any device call.
bin/gpu_ballast's hold and release run in the GPU tests."""
import os
import re
import subprocess

import numpy as np
import pytest

import excent_amd.native as nat
from oracle import oracle as orc
from tests.conftest import ROOT, SAMPLE_EGR

BIN = os.path.join(ROOT, "bin")
MASK = (1 << 64) - 1


class MT19937_64:
    """The 64-bit Mersenne Twister (Matsumoto and Nishimura, 2004) with the parameters of the
    C++ standard's std::mt19937_64."""

    def __init__(self, seed):
        self.mt = [seed & MASK]
        for i in range(1, 312):
            prev = self.mt[-1]
            self.mt.append((6364136223846793005 * (prev ^ (prev >> 62)) + i) & MASK)
        self.i = 312

    def _twist(self):
        mt = self.mt
        for k in range(312):
            x = (mt[k] & 0xFFFFFFFF80000000) | (mt[(k + 1) % 312] & 0x7FFFFFFF)
            xa = x >> 1
            if x & 1:
                xa ^= 0xB5026F5AA96619E9
            mt[k] = mt[(k + 156) % 312] ^ xa
        self.i = 0

    def __call__(self):
        if self.i >= 312:
            self._twist()
        y = self.mt[self.i]
        self.i += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & MASK


def test_mt19937_64_known_answer():
    # [rand.predef]: the 10000th invocation of a default-constructed mt19937_64 (seed 5489)
    g = MT19937_64(5489)
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


def expected_info(path):
    idx, adj = orc.egr_load(path)
    idx = idx.astype(np.int64)
    n, m = len(idx) - 1, len(adj)
    deg = np.diff(idx)
    draw = MT19937_64(7)
    missing = loops = 0
    for _ in range(20000):
        e = draw() % m
        s = int(np.searchsorted(idx, e, side="right")) - 1
        d = int(adj[e])
        loops += s == d
        missing += not np.any(adj[idx[d]:idx[d + 1]] == s)
    return (f"n={n} m={m} ({m * 4.0 / 1e9:.2f} GB edges) maxdeg={int(deg.max())}  reverse-edge missing "
            f"{missing}/20000 ({100.0 * missing / 20000:.1f}%) self-loops {loops}  m>2^32-1: no")


@pytest.fixture(scope="module")
def directed12(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("harness") / "d12.egr")
    assert nat.run_entry("excent_rmat_main", ["--scale", 12, "--edge-factor", 8, "--seed", 5, path]) == 0
    return path


@pytest.mark.parametrize("which", ["sample", "reference_sample", "directed"])
def test_egr_info(which, directed12):
    path = {"sample": os.path.join(ROOT, "excent_amd", "data", "sample.egr"), "reference_sample": SAMPLE_EGR,
            "directed": directed12}[which]
    out = subprocess.run([os.path.join(BIN, "egr_info"), path], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == expected_info(path)
    if which == "sample":   # the shipped sample is symmetric
        assert "reverse-edge missing 0/20000" in out.stdout
    if which == "directed":
        assert not re.search(r"reverse-edge missing 0/", out.stdout)


def test_evict_file(tmp_path):
    f = tmp_path / "blob.bin"
    np.arange(2 << 20, dtype=np.int32).tofile(str(f))   # 8 MiB, just written: in the page cache
    with open(f, "rb") as fh:
        fh.read()
    out = subprocess.run([os.path.join(BIN, "evict_file"), "--report", str(f)], capture_output=True, text=True)
    assert out.returncode == 0 and 0.0 <= float(out.stdout) <= 100.0
    out = subprocess.run([os.path.join(BIN, "evict_file"), str(f)], capture_output=True, text=True)
    m = re.fullmatch(re.escape(str(f)) + r": ([0-9.]+)% -> ([0-9.]+)% resident( \(fadvise failed\))?\n", out.stdout)
    assert m, out.stdout
    # exit status 1 exactly when the advice failed or more than 1 % stayed (a file system
    # that keeps its pages, such as tmpfs, reports that)
    assert out.returncode == (1 if m.group(3) or float(m.group(2)) > 1.0 else 0)
    out = subprocess.run([os.path.join(BIN, "evict_file"), str(tmp_path / "absent")], capture_output=True, text=True)
    assert out.returncode == 1 and "absent" in out.stderr


@pytest.mark.parametrize("prog,usage", [("egr_info", "<graph.egr>"), ("evict_file", "[--report] <file>"),
                                        ("gpu_ballast", "<budget_GB>")])
def test_usage(prog, usage):
    out = subprocess.run([os.path.join(BIN, prog)], capture_output=True, text=True)
    assert out.returncode == 1 and "usage:" in out.stderr and usage in out.stderr


def test_egr_info_bad_file(tmp_path):
    bad = tmp_path / "short.egr"
    np.array([100, 50], np.int32).tofile(str(bad))   # header only
    out = subprocess.run([os.path.join(BIN, "egr_info"), str(bad)], capture_output=True, text=True)
    assert out.returncode == 1 and "shorter than its header" in out.stderr

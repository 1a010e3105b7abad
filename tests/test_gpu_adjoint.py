"""Differentiable personalised PageRank (include/excent.h excent_dpr_*; no reference
counterpart): the forward weighted push against the FP64 weighted solve, the adjoint
push's gradients against the FP64 adjoint solve and against FP64 central finite
differences on 32 seeded directions (relative error <= 1e-3)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import excent_amd.native as nat
from oracle import oracle as orc
from tests.conftest import ROOT

pytestmark = pytest.mark.gpu

D = 0.85


@pytest.fixture(scope="module")
def wgraph(tmp_path_factory):
    p = str(tmp_path_factory.mktemp("wg") / "w.egr")
    assert nat.run_entry("excent_rmat_main", ["--scale", 11, "--edge-factor", 8, "--seed", 4, "--weights", p]) == 0
    idx, adj = orc.egr_load(p)
    w = np.fromfile(p + ".w.f32", np.float32)
    return idx, adj, w


def normalise(idx, w):
    n = len(idx) - 1
    deg = np.diff(idx.astype(np.int64))
    src = np.repeat(np.arange(n), deg)
    S = np.bincount(src, weights=w.astype(np.float64), minlength=n)
    return w.astype(np.float64) / S[src], S, src


def fp64_forward(idx, adj, w, s):
    wn, _, _ = normalise(idx, w)
    x, _, _ = orc.fp64_solve_w(idx, adj, wn, (1 - D) * s.astype(np.float64))
    return x


def transpose(idx, adj, vals):
    n = len(idx) - 1
    deg = np.diff(idx.astype(np.int64))
    src = np.repeat(np.arange(n), deg)
    order = np.lexsort((src, adj))
    tidx = np.concatenate([[0], np.cumsum(np.bincount(adj, minlength=n))]).astype(np.uint64)
    return tidx, src[order].astype(np.int32), vals[order]


def test_fp64_coefficient_solves_against_oracle(wgraph):
    """The product's FP64 Jacobi with per-edge coefficients (exc_fp64_solve_coef), used to
    check the full-size gradients, against the oracle's weighted FP64 solve: the forward
    system over the in-lists, and the adjoint system over the out-lists against the oracle
    on a host transpose."""
    idx, adj, w = wgraph
    n = len(idx) - 1
    wn, _, _ = normalise(idx, w)
    rng = np.random.default_rng(3)
    c = rng.random(n)
    g = rng.standard_normal(n)
    with nat.Fp64Solver(idx, adj) as F:
        F.set_coef(adj, wn)
        x, rel, it = F.solve_coef(c, D)
        lam, rel2, _ = F.solve_coef(g, D, transpose=True)
    assert rel <= 1e-14 and rel2 <= 1e-14 and it > 10
    xo, _, _ = orc.fp64_solve_w(idx, adj, wn, c, s=D)
    tidx, tadj, twn = transpose(idx, adj, wn)
    lo, _, _ = orc.fp64_solve_w(tidx, tadj, twn, g, s=D)
    assert np.abs(x - xo).max() <= 1e-10 * np.abs(xo).max()
    assert np.abs(lam - lo).max() <= 1e-10 * np.abs(lo).max()
    # the two systems are adjoint: g . x(c) = c . lambda(g)
    assert abs(g @ x - c @ lam) <= 1e-10 * abs(g @ x)


def test_forward_against_fp64(wgraph):
    idx, adj, w = wgraph
    n = len(idx) - 1
    rng = np.random.default_rng(1)
    s = np.zeros(n, np.float32)
    s[rng.choice(n, 16, replace=False)] = n / 16
    with nat.DiffPageRank(idx, adj, w, eps=1e-7) as P:
        x = P.forward(s)
    ref = fp64_forward(idx, adj, w, s)
    assert np.abs(x - ref).sum() <= 1e-7 * n / (1 - D) + 1e-4 * np.abs(ref).sum() * 1e-2


def test_gradients_against_fp64_adjoint_and_finite_differences(wgraph):
    idx, adj, w = wgraph
    n, m = len(idx) - 1, len(adj)
    rng = np.random.default_rng(2)
    s = rng.random(n).astype(np.float32) + 0.5
    g = rng.standard_normal(n).astype(np.float32)
    with nat.DiffPageRank(idx, adj, w, eps=1e-8) as P:
        x = P.forward(s)
        gs, gw, gwn = P.backward(g)
    # FP64 adjoint: lambda = (I - d A^T)^-1 g is the weighted solve on the transpose
    wn, S, src = normalise(idx, w)
    tidx, tadj, twn = transpose(idx, adj, wn)
    lam, _, _ = orc.fp64_solve_w(tidx, tadj, twn, g.astype(np.float64))
    xs = fp64_forward(idx, adj, w, s)
    gs_ref = (1 - D) * lam
    assert np.linalg.norm(gs - gs_ref) <= 1e-3 * np.linalg.norm(gs_ref)
    t = np.bincount(src, weights=wn * lam[adj], minlength=n)
    gw_ref = D * xs[src] / S[src] * (lam[adj] - t[src])
    assert np.linalg.norm(gw - gw_ref) <= 1e-3 * np.linalg.norm(gw_ref)
    assert np.linalg.norm(gwn - D * lam[adj] * xs[src]) <= 1e-3 * np.linalg.norm(D * lam[adj] * xs[src])
    # central differences of L = g . x in FP64 on 32 seeded directions
    L = lambda ww, ss: float(g.astype(np.float64) @ fp64_forward(idx, adj, ww, ss))
    h = 1e-3
    worst = 0.0
    for k in range(32):
        dr = np.random.default_rng(100 + k)
        if k % 2 == 0:   # raw edge weights (kept positive)
            dw = dr.standard_normal(m)
            fd = (L(w.astype(np.float64) + h * dw, s) - L(w.astype(np.float64) - h * dw, s)) / (2 * h)
            an = float(gw.astype(np.float64) @ dw)
        else:            # teleport vector
            ds = dr.standard_normal(n)
            fd = (L(w, s.astype(np.float64) + h * ds) - L(w, s.astype(np.float64) - h * ds)) / (2 * h)
            an = float(gs.astype(np.float64) @ ds)
        worst = max(worst, abs(an - fd) / max(abs(fd), 1e-12))
    assert worst <= 1e-3, worst


def test_autograd_function_in_subprocess(wgraph, tmp_path):
    """torch.autograd.Function wrapper (torch imported first: one HIP runtime)."""
    idx, adj, w = wgraph
    np.save(tmp_path / "idx.npy", idx)
    np.save(tmp_path / "adj.npy", adj)
    np.save(tmp_path / "w.npy", w)
    code = f"""
import sys, numpy as np, torch
sys.path.insert(0, {ROOT!r})
from excent_amd.autograd import pagerank
import excent_amd.native as nat
idx = np.load({str(tmp_path / 'idx.npy')!r}); adj = np.load({str(tmp_path / 'adj.npy')!r}); w0 = np.load({str(tmp_path / 'w.npy')!r})
n = len(idx) - 1
s = torch.ones(n, dtype=torch.float32, requires_grad=True)
w = torch.tensor(w0, requires_grad=True)
x = pagerank(s, w, idx, adj, 0.85, 1e-8)
g = torch.from_numpy(np.random.default_rng(5).standard_normal(n).astype(np.float32))
(x * g).sum().backward()
with nat.DiffPageRank(idx, adj, w0, eps=1e-8) as P:
    P.forward(np.ones(n, np.float32))
    gs, gw, _ = P.backward(g.numpy())
assert np.allclose(s.grad.numpy(), gs, rtol=1e-4, atol=1e-6)
assert np.allclose(w.grad.numpy(), gw, rtol=1e-4, atol=1e-6)
print("autograd ok")
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "autograd ok" in r.stdout, r.stdout + r.stderr

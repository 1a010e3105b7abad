"""Size-independent properties at the headline workload's full size (R-MAT scale 24, edge
factor 16, --measure all, the whole graph in HBM, dense passes on), where no CPU oracle run
finishes in seconds.  The product's FP64 solver and residual stand in for the oracle here;
both are pinned to the oracle's FP64 restatement on small graphs (test_gpu_cli.py)."""
import numpy as np
import pytest

import excent_amd.native as nat
from tests.conftest import read_env, rmat_graph

pytestmark = pytest.mark.gpu


def test_scale24_all_terminates_inside_its_bounds(tmpdir_factory):
    """The headline workload's stop rule (--stop katz,ev): PageRank until every |r| <= 0.01
    (src/excent.cu:58-64), Katz and eigenvector until their sum |r| <= 1e-6 |V| (the global
    reading of eps = 1e-6 |V|) or nothing of theirs is pending.  After the run: PageRank's
    every |r| <= eps; for every measure the FP64 residual ||c + s W x - x||_1 of the float32 x
    (computed exactly from x) is at most n * eps -- the termination guarantee holds for x,
    not only for the float32 r -- and equals sum |r| to within float32 rounding of x (the
    invariant x + r = c + s W x); PageRank's L1 error against the FP64 fixed point stays
    under ||residual||_1 / (1 - d).  Katz and eigenvector (unit vectors) against their FP64
    solutions, with margins over the bench's measured errors (1.3e-7 and 2.1e-5)."""
    g = rmat_graph(tmpdir_factory, 24)
    envp = g + ".env"
    print("case scale 24 all: constants", flush=True)
    rc = nat.run_entry("excent_params_main", [g, envp, "--directed", "--alpha-factor", "0.5", "--seed", "7",
                                              "--eps-abs", "1e-6"])
    assert rc == 0
    env = read_env(envp)
    eps = [0.01, 1e-6, 1e-6]
    args = [g, "--measure", "all", "--cache", "chunks", "--preload", "--stream-ratio", "1", "--alpha", env["ALPHA"],
            "--beta", 1, "--eps-katz", eps[1], "--lambda", env["LAMBDA"], "--eps-ev", eps[2], "--ev-init", "indeg",
            "--stop", "katz,ev", "--max-passes", 3000]
    print("case scale 24 all: run", flush=True)
    with nat.Session(args) as sess:
        r = sess.run(reinit=False)
        X = [sess.values(i, 0) for i in range(3)]
        R = [sess.values(i, 1) for i in range(3)]
    assert r["passes"] > 1 and r["kernel"]["pull_passes"] > 0 and r["stop_reason"] != 2, r["passes"]
    idx, adj = nat.load_egr(g)
    n = len(idx) - 1
    f32 = np.float32
    consts = [(float(f32(1) - f32(0.85)), float(f32(0.85)), 1), (1.0, float(f32(env["ALPHA"])), 0),
              (0.0, float(f32(1.0 / env["LAMBDA"])), 0)]
    print("case scale 24 all: FP64 checks", flush=True)
    with nat.Fp64Solver(idx, adj) as S:
        res = []
        for i, (c, s, dd) in enumerate(consts):
            rsum = float(np.abs(R[i].astype(np.float64)).sum())
            xsum = float(np.abs(X[i].astype(np.float64)).sum())
            if i == 0:
                assert np.all(np.abs(R[i]) <= f32(eps[i]))
            else:
                assert rsum <= n * float(f32(eps[i])) or np.all(np.abs(R[i]) <= f32(eps[i])), (i, rsum)
            assert abs(r["residual_l1"][i] - rsum) <= 1e-12 * rsum + 1e-300, (i, r["residual_l1"][i], rsum)
            res.append(S.residual(X[i], c, s, dd))
            assert abs(res[i] - rsum) <= 256 * 2.0 ** -24 * xsum, (i, res[i], rsum, xsum)   # 0.6 % of res for pr
            assert res[i] <= n * eps[i] + 256 * 2.0 ** -24 * xsum, (i, res[i])
        ref, _, _, _ = S.solve(nat.EXC_PR, tol=1e-12, max_iter=400)
        assert np.abs(X[0] - ref).sum() <= res[0] / (1 - 0.85)
        top = lambda v: set(np.argpartition(-v, 99)[:100].tolist())
        assert len(top(X[0]) & top(ref)) >= 95
        ref, _, _, _ = S.solve(nat.EXC_KATZ, alpha=env["ALPHA"], beta=1.0, tol=1e-12, max_iter=400)
        assert np.abs(X[1] - ref).sum() / np.abs(ref).sum() <= 1e-5
        ref, _, _, _ = S.solve(nat.EXC_EV, tol=1e-10, max_iter=3000)
        xe = X[2].astype(np.float64)
        xe = xe / np.linalg.norm(xe) * (1 if xe.sum() >= 0 else -1)
        assert np.abs(xe - ref).sum() / np.abs(ref).sum() <= 1e-3

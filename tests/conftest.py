import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
SAMPLE_EGR = os.path.join(GOLDEN, "sample.egr")
SAMPLE_MTX = os.path.join(GOLDEN, "sample.mtx")
SAMPLE_ENV = os.path.join(GOLDEN, "sample.env")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def read_env(path):
    out = {}
    for line in open(path):
        line = line.strip()
        if not line or line.startswith("#") or "=" not in line:
            continue
        k, v = line.split("=", 1)
        out[k] = float(v)
    return out


_graphs = {}


def rmat_graph(tmpdir_factory, scale, seed=1, ef=16, extra=()):
    """Path of a seeded R-MAT EGR written by bin/excent_rmat's entry point (cached per session)."""
    key = (scale, seed, ef, tuple(extra))
    if key not in _graphs:
        import excent_amd.native as nat
        d = tmpdir_factory.mktemp(f"rmat{scale}_{seed}")
        path = str(d.join(f"rmat{scale}.egr"))
        rc = nat.run_entry("excent_rmat_main", ["--scale", scale, "--edge-factor", ef, "--seed", seed, *extra, path])
        assert rc == 0
        _graphs[key] = path
    return _graphs[key]


@pytest.fixture(scope="session")
def rmat12(tmpdir_factory):
    return rmat_graph(tmpdir_factory, 12)


@pytest.fixture(scope="session")
def rmat14(tmpdir_factory):
    return rmat_graph(tmpdir_factory, 14, seed=3)


@pytest.fixture(scope="session")
def rmat16(tmpdir_factory):
    return rmat_graph(tmpdir_factory, 16)

"""ExCent for MI355X: asynchronous delta-push centrality (PageRank, Katz, eigenvector)
on storage-resident EGR graphs, with hand-written HIP kernels for gfx950.

The product is native: bin/excent (C++17/OpenMP host over the C ABI of
include/excent_hip.h) and the libraries under excent_amd/lib.  This package only
binds them (``excent_amd.native``) for tests, benchmarks and callers that drive the
C ABI from Python.
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def build(verbose=False):
    """Compiles the HIP and host libraries and the bin/ programs in-tree (gfx950)."""
    import subprocess

    jobs = str(min(16, os.cpu_count() or 4))
    r = subprocess.run(["make", "-C", os.path.join(HERE, "csrc"), "-j" + jobs, "all"],
                       stdout=None if verbose else subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + (r.stdout or ""))

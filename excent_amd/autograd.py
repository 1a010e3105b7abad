"""A thin torch.autograd.Function over the differentiable personalised PageRank of the C
ABI (include/excent.h excent_dpr_*).  It exists for gradient checks: the forward and the
adjoint pushes run in the native library on host buffers; torch only carries tensors.

Import torch before this module so the process has one HIP runtime (torch's), which the
native libraries then bind to by soname.
"""
import numpy as np
import torch

from . import native


class PageRankFunction(torch.autograd.Function):
    """x = PPR(s, w): s teleport [n], w raw edge weights [m] (CSR order of the graph)."""

    @staticmethod
    def forward(ctx, s, w, idx, adj, damp=0.85, eps=1e-7):
        solver = native.DiffPageRank(idx, adj, w.detach().cpu().numpy().astype(np.float32), damp=damp, eps=eps)
        x = solver.forward(s.detach().cpu().numpy().astype(np.float32))
        ctx.solver = solver
        return torch.from_numpy(x).to(dtype=s.dtype, device=s.device)

    @staticmethod
    def backward(ctx, gx):
        gs, gw, _ = ctx.solver.backward(gx.detach().cpu().numpy().astype(np.float32))
        ctx.solver.close()
        return (torch.from_numpy(gs).to(dtype=gx.dtype, device=gx.device),
                torch.from_numpy(gw).to(dtype=gx.dtype, device=gx.device), None, None, None, None)


def pagerank(s, w, idx, adj, damp=0.85, eps=1e-7):
    return PageRankFunction.apply(s, w, idx, adj, damp, eps)

"""torch.autograd bridge: the recovered fluxes of the mimetic hybrid solve as a
differentiable function of permeability, mobilities and sources.

forward  = computeInnerProducts + solve (IncompFlowSolverHybrid.hpp:564-578,
           :657-703) on device tensors, returning outflux per half-face;
backward = mhs_adjoint_device: one adjoint solve with the same face operator
           and preconditioner, then per-cell derivative kernels.

The gradient is exact for the discrete solve (up to the solve tolerances).
Boundary values and geometry are constants.  Periodic boundaries are not
supported by the adjoint.
"""
from __future__ import annotations

import torch

from .mhs import HybridSolver


def _sync_if_unordered(solver: HybridSolver, dev):
    # a solver on its own stream is not ordered after torch's work
    if not solver.stream:
        torch.cuda.current_stream(dev).synchronize()


class HybridFlux(torch.autograd.Function):
    """v = outflux(K, totmob, rhomob, src) of a solver whose grid and structure
    are set (set_grid / set_cell_ijk / initSystemStructure)."""

    @staticmethod
    def forward(ctx, perm, totmob, rhomob, src, solver, bc_type, bc_value, gravity, options):
        dev = perm.device
        nc = totmob.shape[0]
        perm_c = perm.detach().reshape(nc, 9).contiguous().to(torch.float64)
        tm = totmob.detach().contiguous().to(torch.float64)
        rm = rhomob.detach().contiguous().to(torch.float64)
        sr = src.detach().contiguous().to(torch.float64)
        bv = bc_value.detach().contiguous().to(torch.float64)
        nrhs = 1 if sr.dim() == 1 else sr.shape[0]
        _sync_if_unordered(solver, dev)
        solver.computeInnerProducts_device(perm_c.data_ptr(), gravity)
        solver.solve_device(tm.data_ptr(), rm.data_ptr(), sr.data_ptr(), bc_type, bv.data_ptr(),
                            nrhs=nrhs, **options)
        nhf = solver.dims().num_half_faces
        v = torch.empty((nrhs, nhf), dtype=torch.float64, device=dev)
        solver.get_solution_device(None, v.data_ptr())
        ctx.solver = solver
        ctx.count = solver.solve_count
        ctx.nrhs = nrhs
        ctx.perm_shape = perm.shape
        ctx.src_shape = src.shape
        ctx.options = {k: options[k] for k in ("tol", "chunk", "use_graph") if k in options}
        return v if nrhs > 1 else v[0]

    @staticmethod
    def backward(ctx, gv):
        solver = ctx.solver
        if solver.solve_count != ctx.count:
            raise RuntimeError("HybridFlux.backward: the solver has solved again since this "
                               "forward; its state no longer belongs to this graph")
        dev = gv.device
        nc = solver.dims().num_cells
        w = gv.detach().reshape(ctx.nrhs, -1).contiguous().to(torch.float64)
        need = ctx.needs_input_grad
        gK = torch.empty((nc, 9), dtype=torch.float64, device=dev) if need[0] else None
        gt = torch.empty(nc, dtype=torch.float64, device=dev) if need[1] else None
        gr = torch.empty(nc, dtype=torch.float64, device=dev) if need[2] else None
        gs = torch.empty((ctx.nrhs, nc), dtype=torch.float64, device=dev) if need[3] else None
        _sync_if_unordered(solver, dev)
        solver.adjoint_device(w.data_ptr(),
                              gK.data_ptr() if gK is not None else None,
                              gt.data_ptr() if gt is not None else None,
                              gr.data_ptr() if gr is not None else None,
                              gs.data_ptr() if gs is not None else None, **ctx.options)
        if gK is not None:
            gK = gK.reshape(ctx.perm_shape)
        if gs is not None:
            gs = gs.reshape(ctx.src_shape)
        return gK, gt, gr, gs, None, None, None, None, None


def hybrid_flux(solver: HybridSolver, perm, totmob, rhomob, src, bc_type, bc_value,
                gravity=(0.0, 0.0, 0.0), **options):
    """Differentiable outflux [nhf] (or [nrhs, nhf]) of one solve.  perm is
    [nc, 9] or [nc, 3, 3]; options go to HybridSolver.solve_device (tol,
    precision, chunk, use_graph, precond, ...)."""
    return HybridFlux.apply(perm, totmob, rhomob, src, solver, bc_type, bc_value,
                            tuple(float(x) for x in gravity), options)

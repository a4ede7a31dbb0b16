"""The multigrid preconditioner: a smoothed transfer from level 1 to level 2
and tentative (piecewise-constant) transfers below it, with one smoothing
sweep (the fused V-cycle: lane groups on the long coarse rows, the
workgroup-per-instance tail) and with several (the V-cycle with its own
prolongation and sweep kernels).  A batch of two uneven instances with two
right-hand sides, against the FP64 oracle at 1e-8."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

DIMS = ((24, 24, 12), (14, 10, 9))


@pytest.fixture(scope="module")
def batch(pkg, orc):
    gs = [pkg.grids.cartesian(*d) for d in DIMS]
    G, icp = pkg.grids.concat(gs)
    nc = G.num_cells
    perm = pkg.fields.perm_c1(nc, 5)
    bt, bv = cases.xdrop_bc(G)
    tm, rm = cases.unit_fluid(nc)
    bv2 = np.stack([bv, -0.25 * bv])
    refs = {}
    for t in range(len(gs)):
        c0, c1 = icp[t], icp[t + 1]
        h0, h1 = G.facepos[c0], G.facepos[c1]
        for k in range(2):
            refs[t, k] = orc.solve(gs[t], perm[c0:c1], tm[c0:c1], rm[c0:c1], np.zeros(c1 - c0),
                                   bt[h0:h1], bv2[k, h0:h1], tol=1e-14)
    return G, icp, perm, tm, rm, bt, bv2, refs


@pytest.mark.parametrize("sweeps", [1, 2, 3])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_vcycle_matches_the_oracle(pkg, batch, sweeps, precision):
    G, icp, perm, tm, rm, bt, bv2, refs = batch
    s = pkg.HybridSolver(0)
    s.init(G, perm, (0, 0, 0), bt, None, inst_cellpos=icp)
    s.solve(tm, rm, np.zeros((2, G.num_cells)), bt, bv2, residual_tolerance=1e-11,
            precision=precision, smooth_steps=sweeps)
    r = s.report
    assert r.precond == 1 and r.levels >= 4, (r.precond, r.levels)
    p, pi, v = s.getSolution()
    for t in range(len(DIMS)):
        c0, c1 = icp[t], icp[t + 1]
        h0, h1 = G.facepos[c0], G.facepos[c1]
        for k in range(2):
            ref = refs[t, k]
            assert cases.rel_err(ref.outflux, v[k, h0:h1]) < 1e-8
            assert cases.rel_err(ref.cell_pressure, p[k, c0:c1]) < 1e-8


def test_multigrid_beats_jacobi(pkg, batch):
    """The hierarchy is doing its job: far fewer iterations than Jacobi-PCG on
    the same system, and more sweeps do not need more iterations."""
    G, icp, perm, tm, rm, bt, bv2, refs = batch
    s = pkg.HybridSolver(0)
    s.init(G, perm, (0, 0, 0), bt, None, inst_cellpos=icp)
    its = {}
    for pc, sw in (("jacobi", 1), ("amg", 1), ("amg", 2)):
        s.solve(tm, rm, np.zeros((2, G.num_cells)), bt, bv2, residual_tolerance=1e-10,
                precond=pc, smooth_steps=sw)
        its[pc, sw] = s.report.iterations
    assert its["amg", 1] * 4 < its["jacobi", 1], its
    assert its["amg", 2] <= its["amg", 1], its

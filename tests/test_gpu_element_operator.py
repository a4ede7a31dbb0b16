"""The element-by-element level-0 operator of the mixed-precision solve
(the face matrix applied cell by cell from the local Schur complements after
the Dirichlet rules of addCellContrib, IncompFlowSolverHybrid.hpp:1817-1901,
one launch per cell colour): the same converged solution as the oracle and
as the assembled FP32 matrix, on box grids (two colours), faulted
corner-point grids (more colours, 6 to 14 faces per cell), batches of uneven
instances with several right-hand sides, under gravity and sources, with
the multigrid and the Jacobi preconditioner; and where it does not apply
(periodic faces, a regularised instance) the solve falls back to the
assembled matrix."""
import numpy as np
import pytest

import cases
from cases import rel_err

pytestmark = pytest.mark.gpu


def _check(pkg, orc, g, perm, tm, rm, src, bt, bv, grav=(0.0, 0.0, 0.0), inst=None, parts=None, precond=-1,
           tol=1e-12):
    nrhs = 1 if np.ndim(src) == 1 else src.shape[0]
    s = pkg.HybridSolver(0)
    s.init(g, perm, grav, bt, None, inst_cellpos=inst)
    out = {}
    for eo in (True, False):
        # 2: cell by cell whenever it applies (1 would let the autotuner pick)
        rep = s.solve(tm, rm, src, bt, bv, residual_tolerance=tol, precision="mixed", precond=precond,
                      element_operator=2 if eo else 0)
        assert rep.converged and rep.reduction <= tol
        assert rep.element_operator == int(eo)
        out[eo] = (s.getSolution(), rep.iterations, rep.colours)
    (p, pi, v), it_e, ncol = out[True]
    (p0, pi0, v0), it_a, _ = out[False]
    # the two operators agree to FP32 rounding: the converged solutions agree
    assert rel_err(v, v0) < 1e-8 and rel_err(p, p0) < 1e-8
    assert abs(it_e - it_a) <= max(2, it_a // 10)
    # against the oracle, instance by instance and right-hand side by right-hand side
    icp = inst if inst is not None else np.array([0, g.num_cells])
    for t in range(len(icp) - 1):
        c0, c1 = icp[t], icp[t + 1]
        h0, h1 = g.facepos[c0], g.facepos[c1]
        gt = g if inst is None else parts[t]
        for k in range(nrhs):
            sk = src if nrhs == 1 else src[k]
            bk = bv if nrhs == 1 else bv[k]
            ref = orc.solve(gt, perm[c0:c1], tm[c0:c1], rm[c0:c1], sk[c0:c1], bt[h0:h1], bk[h0:h1], grav=grav,
                            tol=1e-14)
            vk = v if nrhs == 1 else v[k]
            pk = p if nrhs == 1 else p[k]
            assert rel_err(vk[h0:h1], ref.outflux) < 1e-8, (t, k)
            assert rel_err(pk[c0:c1], ref.cell_pressure) < 1e-8, (t, k)
    return ncol


@pytest.mark.parametrize("precond", ["amg", "jacobi"])
def test_box_grid_two_colours(pkg, orc, precond):
    g, perm, tm, rm, src, bt, bv = cases.case_c1()
    ncol = _check(pkg, orc, g, perm, tm, rm, src, bt, bv, precond=precond)
    assert ncol == 2


def test_faulted_grid_with_gravity_and_sources(pkg, orc):
    g = pkg.grids.corner_point(14, 12, 6, 10.0, 10.0, 2.0, seed=5)
    nc = g.num_cells
    assert g.ncf().max() > 6
    perm = pkg.fields.perm_c4(nc, seed=5)
    bt, bv = cases.xdrop_bc(g, 2.0e5, 1.0e5)
    rng = np.random.default_rng(3)
    tm = rng.uniform(500.0, 1500.0, nc)
    rm = tm * rng.uniform(800.0, 1000.0, nc)
    src = np.zeros(nc)
    src[nc // 3] = 1e-4
    src[2 * nc // 3] = -0.5e-4
    ncol = _check(pkg, orc, g, perm, tm, rm, src, bt, bv, grav=(0.0, 0.0, 9.81))
    assert ncol >= 2


def test_batch_of_uneven_instances_two_rhs(pkg, orc):
    gs = [pkg.grids.cartesian(9, 7, 5), pkg.grids.cartesian(12, 4, 6, 0.5, 1.0, 2.0), pkg.grids.cartesian(5, 5, 5)]
    G, icp = pkg.grids.concat(gs)
    nc = G.num_cells
    perm = pkg.fields.perm_c1(nc, 8)
    bt, bv = cases.xdrop_bc(G)
    bv2 = np.stack([bv, -0.5 * bv])
    src = np.zeros((2, nc))
    src[1, 10] = 0.2
    tm, rm = np.ones(nc), np.zeros(nc)
    _check(pkg, orc, G, perm, tm, rm, src, bt, bv2, inst=icp, parts=gs)


def test_falls_back_where_it_does_not_apply(pkg, orc):
    # periodic faces: the constraint rows are not cell contributions
    g = pkg.grids.cartesian(6, 5, 4)
    perm = cases.fields.perm_c1(g.num_cells, seed=3)
    bt, bv, partner = cases.periodic_bc(g, (0, 1), drop_dir=0, pdrop=1.0)
    tm, rm = cases.unit_fluid(g.num_cells)
    s = pkg.HybridSolver(0)
    s.init(g, perm, (0, 0, 0), bt, partner)
    rep = s.solve(tm, rm, np.zeros(g.num_cells), bt, bv, residual_tolerance=1e-12, precision="mixed")
    assert rep.element_operator == 0 and rep.converged
    # no Dirichlet face: the regularised entry is not a cell contribution
    g = pkg.grids.cartesian(6, 5, 4)
    nc = g.num_cells
    bt = np.zeros(g.num_half_faces, np.int32)
    bv = np.where(g.side == 0, -1.0, np.where(g.side == 1, 1.0, 0.0))   # Neumann in at x-, out at x+
    s = pkg.HybridSolver(0)
    s.init(g, perm, (0, 0, 0), bt, None)
    rep = s.solve(tm, rm, np.zeros(nc), bt, bv, residual_tolerance=1e-12, precision="mixed")
    assert rep.element_operator == 0 and rep.converged
    ref = orc.solve(g, perm, tm, rm, np.zeros(nc), bt, bv, tol=1e-14)
    assert rel_err(s.getSolution()[2], ref.outflux) < 1e-8


def test_autotuner_chooses_the_operator(pkg, orc):
    """element_operator 1 (the default): the first solve times the cell-by-cell
    product against the assembled FP32 SpMV and keeps the faster for the
    structure; either way the solution is the oracle's."""
    for g in (cases.case_c1()[0], pkg.grids.corner_point(14, 12, 6, 10.0, 10.0, 2.0, seed=5)):
        nc = g.num_cells
        perm = pkg.fields.perm_c4(nc, seed=5)
        bt, bv = cases.xdrop_bc(g, 2.0e5, 1.0e5)
        tm, rm = cases.unit_fluid(nc)
        s = pkg.HybridSolver(0)
        s.init(g, perm, (0, 0, 0), bt, None)
        ref = orc.solve(g, perm, tm, rm, np.zeros(nc), bt, bv, tol=1e-14)
        chosen = None
        for _ in range(2):
            rep = s.solve(tm, rm, np.zeros(nc), bt, bv, residual_tolerance=1e-12, precision="mixed")
            assert rep.converged and rep.element_operator in (0, 1)
            if chosen is None:
                assert rep.tune_ms > 0.0
                chosen = rep.element_operator
            else:
                assert rep.tune_ms == 0.0 and rep.element_operator == chosen   # kept for the structure
            assert rel_err(s.getSolution()[2], ref.outflux) < 1e-8


@pytest.mark.parametrize("precond", ["amg", "jacobi"])
def test_solve_without_face_matrix(pkg, orc, precond):
    """Once the element operator is settled and the smoother weights are
    known, a mixed solve fills no face matrix (multigrid A P from the element
    values): it matches the oracle and a solve that filled the matrix, the
    matrix asked for afterwards equals the one an FP64 solve assembles, and a
    solve stopped early still reports its true FP64 residual."""
    g, perm, tm, rm, src, bt, bv = cases.case_c1()
    ref = orc.solve(g, perm, tm, rm, src, bt, bv, tol=1e-14)
    s = pkg.HybridSolver(0)
    s.init(g, perm, (0.0, 0.0, 0.0), bt, None)
    kw = dict(residual_tolerance=1e-12, precision="mixed", precond=precond, element_operator=2)
    first = s.solve(tm, rm, src, bt, bv, **kw)
    assert first.converged and first.element_operator == 1
    v_first = s.getSolution()[2].copy()
    # different mobilities, then the original ones again: no matrix either time
    rep = s.solve(2.0 * tm, rm, src, bt, bv, **kw)
    assert rep.converged and rep.face_matrix == 0
    rep = s.solve(tm, rm, src, bt, bv, **kw)
    assert rep.converged and rep.face_matrix == 0 and rep.reduction <= 1e-12
    p, pi, v = s.getSolution()
    assert rel_err(v, ref.outflux) < 1e-8 and rel_err(p, ref.cell_pressure) < 1e-8
    assert rel_err(pi, ref.face_pressure) < 1e-8
    assert abs(rep.iterations - first.iterations) <= 1
    assert rel_err(v, v_first) < 1e-9
    # the face matrix on request: the one a fresh FP64 solve fills
    rp, cols, vals, rhs = s.system(0)
    s64 = pkg.HybridSolver(0)
    s64.init(g, perm, (0.0, 0.0, 0.0), bt, None)
    r64 = s64.solve(tm, rm, src, bt, bv, residual_tolerance=1e-12, precision="fp64", precond=precond)
    assert r64.face_matrix == 1
    rp64, cols64, vals64, rhs64 = s64.system(0)
    np.testing.assert_array_equal(rp, rp64)
    np.testing.assert_array_equal(cols, cols64)
    np.testing.assert_array_equal(vals, vals64)
    np.testing.assert_array_equal(rhs, rhs64)
    # stopped early: the report's reduction is the true FP64 residual of the
    # iterate, ||b - A x|| / ||b - A x0|| with the Dirichlet guess x0
    with pytest.raises(RuntimeError):
        s.solve(tm, rm, src, bt, bv, linsolver_maxit=3, **kw)
    red = s.report.reduction
    x = s.getSolution()[1]
    import scipy.sparse as sp
    n = len(rp64) - 1
    A = sp.csr_matrix((vals64, cols64, rp64), shape=(n, n))
    Ad = A.copy()
    Ad.eliminate_zeros()
    dirichlet = np.diff(Ad.indptr) == 1       # rows that hold only their diagonal
    x0 = np.where(dirichlet, rhs64 / A.diagonal(), 0.0)
    true = np.linalg.norm(rhs64 - A @ x) / np.linalg.norm(rhs64 - A @ x0)
    assert red > 1e-12 and abs(red - true) <= 1e-6 * true

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


def _aniso_perm(nc, kx, kz):
    K = np.zeros((nc, 9))
    K[:, 0] = K[:, 4] = kx
    K[:, 8] = kz
    return K


@pytest.mark.parametrize("spacing,kz,first", [
    ((1.0, 1.0, 1.0), 1.0, (2, 2, 2)),     # isotropic: full coarsening
    ((1.0, 1.0, 1.0), 0.1, (2, 2, 1)),     # weak vertical coupling: coarsen in x and y first
    ((10.0, 10.0, 1.0), 1.0, (1, 1, 2)),   # thin layers: vertical coupling dominates
])
def test_semi_coarsening_follows_coupling_strength(pkg, orc, spacing, kz, first):
    """Aggregate shapes from K and the geometry (mean two-point half-
    transmissibility along i, j, k), and per-level Jacobi weights 4/(3 rho)
    with rho from Lanczos steps; the solution still matches the oracle."""
    g = pkg.grids.cartesian(24, 24, 16, *spacing)
    nc = g.num_cells
    perm = _aniso_perm(nc, 1e-12, kz * 1e-12)
    bt, bv = cases.xdrop_bc(g)
    tm, rm = cases.unit_fluid(nc)
    s = pkg.HybridSolver(0)
    s.init(g, perm, (0, 0, 0), bt, None)
    s.solve(tm, rm, np.zeros(nc), bt, bv, residual_tolerance=1e-11, precision="mixed")
    r = s.report.as_dict()
    assert r["precond"] == 1
    assert r["coarsening"][0] == first, r["coarsening"]
    # Cartesian face matrices: rho(D^-1 A) is about 2 to 2.5 on the face level,
    # below 2 on the coarse levels
    w = r["smoother_weights"]
    assert len(w) == r["levels"] - 1
    assert 0.45 < w[0] < 0.75, w
    assert all(0.2 <= x <= 1.0 for x in w), w
    ref = orc.solve(g, perm, tm, rm, np.zeros(nc), bt, bv, tol=1e-14)
    assert cases.rel_err(ref.outflux, s.getSolution()[2]) < 1e-8


def test_forced_smoother_weight(pkg):
    g, perm, tm, rm, src, bt, bv = cases.case_c1()
    s = pkg.HybridSolver(0)
    s.init(g, perm, (0, 0, 0), bt, None)
    s.solve(tm, rm, src, bt, bv, residual_tolerance=1e-10, smoother_weight=0.6)
    w = s.report.as_dict()["smoother_weights"]
    assert w and all(x == 0.6 for x in w), w
    it_forced = s.report.iterations
    s.solve(tm, rm, src, bt, bv, residual_tolerance=1e-10)
    w2 = s.report.as_dict()["smoother_weights"]
    assert w2 != w and s.report.converged
    assert it_forced > 0


@pytest.mark.parametrize("dims", [(40, 30, 1), (1, 1, 60), (1, 24, 20)])
def test_flat_and_thin_grids(pkg, orc, dims):
    """Directions of extent 1 are never coarsened; the hierarchy still reaches
    a small coarsest level and the solution matches the oracle."""
    g = pkg.grids.cartesian(*dims)
    nc = g.num_cells
    perm = pkg.fields.perm_c1(nc, 4)
    bt, bv = cases.xdrop_bc(g) if dims[0] > 1 else _zdrop(g)
    tm, rm = cases.unit_fluid(nc)
    s = pkg.HybridSolver(0)
    s.init(g, perm, (0, 0, 0), bt, None)
    s.solve(tm, rm, np.zeros(nc), bt, bv, residual_tolerance=1e-11, precision="mixed")
    r = s.report.as_dict()
    assert r["precond"] == 1
    for blk in r["coarsening"]:
        for d in range(3):
            if dims[d] == 1:
                assert blk[d] == 1, r["coarsening"]
    ref = orc.solve(g, perm, tm, rm, np.zeros(nc), bt, bv, tol=1e-14)
    assert cases.rel_err(ref.outflux, s.getSolution()[2]) < 1e-8


def _zdrop(g):
    """Dirichlet 1 on z- and 0 on z+, no-flow elsewhere."""
    bt = np.zeros(g.num_half_faces, np.int32)
    bv = np.zeros(g.num_half_faces)
    side = np.asarray(g.side)
    bt[(side == 4) | (side == 5)] = cases.DIR
    bv[side == 4] = 1.0
    return bt, bv


def test_batch_of_differently_anisotropic_instances(pkg, orc):
    """One coarsening plan for the batch (strengths averaged over all
    instances); every instance still converges to the oracle's solution."""
    gs = [pkg.grids.cartesian(20, 20, 8), pkg.grids.cartesian(12, 12, 12, 10.0, 10.0, 1.0)]
    G, icp = pkg.grids.concat(gs)
    nc = G.num_cells
    perm = np.zeros((nc, 9))
    n0 = gs[0].num_cells
    perm[:n0, 0] = perm[:n0, 4] = 1e-12
    perm[:n0, 8] = 1e-14                     # weak vertical coupling
    perm[n0:, [0, 4, 8]] = 1e-12              # isotropic K on thin layers
    bt, bv = cases.xdrop_bc(G)
    tm, rm = cases.unit_fluid(nc)
    s = pkg.HybridSolver(0)
    s.init(G, perm, (0, 0, 0), bt, None, inst_cellpos=icp)
    s.solve(tm, rm, np.zeros(nc), bt, bv, residual_tolerance=1e-11, precision="mixed")
    assert s.report.precond == 1 and s.report.converged
    v = s.getSolution()[2]
    for t in range(2):
        c0, c1 = icp[t], icp[t + 1]
        h0, h1 = G.facepos[c0], G.facepos[c1]
        ref = orc.solve(gs[t], perm[c0:c1], tm[c0:c1], rm[c0:c1], np.zeros(c1 - c0), bt[h0:h1], bv[h0:h1],
                        tol=1e-14)
        assert cases.rel_err(ref.outflux, v[h0:h1]) < 1e-8


def test_face_level_weight_from_spectral_radius(pkg, orc):
    """The face-level Jacobi weight is 4/(3 rho(D^-1 A)) with rho from the
    device Lanczos steps: against rho of the oracle's assembled face matrix
    (scipy, symmetric form D^-1/2 A D^-1/2).  Lanczos approaches rho from
    below, so the weight may only be a little larger."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    g = pkg.grids.cartesian(16, 12, 10, 1.0, 1.0, 0.5)
    nc = g.num_cells
    perm = pkg.fields.perm_c1(nc, 6)
    bt, bv = cases.xdrop_bc(g)
    tm, rm = cases.unit_fluid(nc)
    s = pkg.HybridSolver(0)
    s.init(g, perm, (0, 0, 0), bt, None)
    s.solve(tm, rm, np.zeros(nc), bt, bv, residual_tolerance=1e-10, precision="mixed")
    w0 = s.report.as_dict()["smoother_weights"][0]
    hf, nint, nf, _ = orc.enumerate_dof(g)
    pp, _ = orc.periodic_dof(g, hf, nf, bt, np.full(g.num_half_faces, -1, np.int32))
    binv, pos, gfl = orc.static_ip(g, perm)
    S = orc.assemble(g, hf, nf, pp, binv, gfl, tm, rm, np.zeros(nc), bt, bv)
    A = sp.csr_matrix((S.val, S.col, S.row_ptr), shape=(nf, nf))
    d = 1.0 / np.sqrt(A.diagonal())
    M = sp.diags(d) @ A @ sp.diags(d)
    rho = spl.eigsh(M, k=1, which="LA", return_eigenvectors=False)[0]
    w_exact = 4.0 / (3.0 * rho)
    assert w_exact * 0.999 <= w0 <= w_exact * 1.05, (w0, w_exact, rho)


def test_new_permeability_reestimates_smoother_weights(pkg, orc):
    """The Jacobi weights follow the permeability: after computeInnerProducts
    with a far more anisotropic K on the same handle (rho(D^-1 A) 2.14 ->
    2.97 here), the next solve's face-level weight is 4/(3 rho) of the new
    matrix, and its solution and iteration count equal a fresh handle's
    that was given the same two permeabilities."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    g = pkg.grids.cartesian(16, 12, 10, 1.0, 1.0, 0.5)
    nc = g.num_cells
    p1 = pkg.fields.perm_c1(nc, 6)
    p2 = pkg.fields.diag_perm(p1[:, 0], p1[:, 0], 0.01 * p1[:, 0])
    bt, bv = cases.xdrop_bc(g)
    tm, rm = cases.unit_fluid(nc)
    s = pkg.HybridSolver(0)
    s.init(g, p1, (0, 0, 0), bt, None)
    s.solve(tm, rm, np.zeros(nc), bt, bv, residual_tolerance=1e-10, precision="mixed")
    w1 = s.report.as_dict()["smoother_weights"][0]
    s.computeInnerProducts(p2)
    s.solve(tm, rm, np.zeros(nc), bt, bv, residual_tolerance=1e-10, precision="mixed")
    w2 = s.report.as_dict()["smoother_weights"][0]
    it2 = s.report.iterations
    hf, nint, nf, _ = orc.enumerate_dof(g)
    pp, _ = orc.periodic_dof(g, hf, nf, bt, np.full(g.num_half_faces, -1, np.int32))
    binv, pos, gfl = orc.static_ip(g, p2)
    S = orc.assemble(g, hf, nf, pp, binv, gfl, tm, rm, np.zeros(nc), bt, bv)
    A = sp.csr_matrix((S.val, S.col, S.row_ptr), shape=(nf, nf))
    d = 1.0 / np.sqrt(A.diagonal())
    rho = spl.eigsh(sp.diags(d) @ A @ sp.diags(d), k=1, which="LA", return_eigenvectors=False)[0]
    w_exact = 4.0 / (3.0 * rho)
    assert w_exact * 0.999 <= w2 <= w_exact * 1.05, (w1, w2, w_exact, rho)
    assert w2 < 0.8 * w1
    # the same sequence on a fresh handle
    f = pkg.HybridSolver(0)
    f.init(g, p1, (0, 0, 0), bt, None)
    f.computeInnerProducts(p2)
    f.solve(tm, rm, np.zeros(nc), bt, bv, residual_tolerance=1e-10, precision="mixed")
    assert f.report.as_dict()["smoother_weights"][0] == w2
    assert abs(f.report.iterations - it2) <= 1
    assert cases.rel_err(s.getSolution()[2], f.getSolution()[2]) < 1e-8
    ref = orc.solve(g, p2, tm, rm, np.zeros(nc), bt, bv, tol=1e-14)
    assert cases.rel_err(s.getSolution()[2], ref.outflux) < 1e-8


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_same_matrix_keeps_the_multigrid_setup(pkg, orc, precision):
    """same_matrix (IncompFlowSolverHybrid.hpp:1491): the second solve keeps
    the coarse levels of the first, as the reference keeps its AMG, and
    solves its own system (new sources and boundary values) to the oracle's
    solution.  A caller whose matrix did change (new mobilities) still gets
    the new system's solution: the PCG runs on the new assembly and only the
    preconditioner is the old one."""
    g, perm, tm, rm, src, bt, bv = cases.case_c1()
    s = pkg.HybridSolver(0)
    s.init(g, perm, (0.0, 0.0, 0.0), bt, None)
    kw = dict(residual_tolerance=1e-12, precision=precision, precond="amg")
    first = s.solve(tm, rm, src, bt, bv, **kw)
    assert first.converged and first.mg_setup_reused == 0
    rng = np.random.default_rng(3)
    src2 = np.zeros(g.num_cells)
    src2[rng.integers(0, g.num_cells, 5)] = 1e-6
    src2[rng.integers(0, g.num_cells, 5)] -= 1e-6
    bv2 = 0.5 * bv
    for tm_k, src_k, bv_k in ((tm, src2, bv2), (3.0 * tm, src2, bv2)):
        rep = s.solve(tm_k, rm, src_k, bt, bv_k, same_matrix=True, **kw)
        assert rep.converged and rep.mg_setup_reused == 1 and rep.reduction <= 1e-12
        ref = orc.solve(g, perm, tm_k, rm, src_k, bt, bv_k, tol=1e-14)
        p, pi, v = s.getSolution()
        assert cases.rel_err(pi, ref.face_pressure) < 1e-8
        assert cases.rel_err(p, ref.cell_pressure) < 1e-8
        assert cases.rel_err(v, ref.outflux) < 1e-8
    # without the promise the set-up is done again
    rep = s.solve(tm, rm, src, bt, bv, **kw)
    assert rep.converged and rep.mg_setup_reused == 0
    assert abs(rep.iterations - first.iterations) <= 1

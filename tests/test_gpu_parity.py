"""GPU parity: libmhs.so (through the C-ABI) against the FP64 CPU oracle.

Bar: face pressures, cell pressures and half-face fluxes within 1e-8
(max-norm relative) of the oracle solved to 1e-14; per-cell mass balance
sum_f v_f - src_c within 1e-12 of the largest flux; static inner products and
the assembled face matrix within 1e-12.
"""
import numpy as np
import pytest

import cases
from cases import DIR, NEU, PER, rel_err

pytestmark = pytest.mark.gpu

TOL_SOL = 1e-8


def _solver(pkg, g, perm, grav=(0.0, 0.0, 0.0), bt=None, partner=None, inst=None):
    s = pkg.HybridSolver(0)
    s.init(g, perm, grav, bt, partner, inst_cellpos=inst)
    return s


def _mass_balance(g, v, src):
    cell = g.hf_cell()
    bal = np.bincount(cell, weights=v, minlength=g.num_cells) - src
    return float(np.max(np.abs(bal)) / max(np.max(np.abs(v)), 1e-300))


def _spd_perm(nc, rng, scale=1e-13):
    A = rng.standard_normal((nc, 3, 3))
    K = np.einsum("nij,nkj->nik", A, A) + 0.5 * np.eye(3)[None]
    return (K * scale).reshape(nc, 9)


def test_static_inner_products(pkg, orc):
    rng = np.random.default_rng(7)
    g = pkg.grids.cartesian(5, 4, 3, 0.7, 1.3, 0.4)
    perm = _spd_perm(g.num_cells, rng)
    grav = (0.3, -0.2, 9.81)
    s = _solver(pkg, g, perm, grav)
    binv, gfl = s.inner_products()
    rb, _, rg = orc.static_ip(g, perm, grav)
    assert rel_err(binv, rb) < 1e-13
    assert rel_err(gfl, rg) < 1e-13


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_c1_parity(pkg, orc, precision):
    g, perm, tm, rm, src, bt, bv = cases.case_c1()
    ref = orc.solve(g, perm, tm, rm, src, bt, bv, tol=1e-14)
    s = _solver(pkg, g, perm)
    rep = s.solve(tm, rm, src, bt, bv, residual_tolerance=1e-12, precision=precision)
    assert rep.converged and rep.reduction <= 1e-12
    p, pi, v = s.getSolution()
    assert rel_err(pi, ref.face_pressure) < TOL_SOL
    assert rel_err(p, ref.cell_pressure) < TOL_SOL
    assert rel_err(v, ref.outflux) < TOL_SOL
    assert _mass_balance(g, v, src) < 1e-12


def test_c1_system_matrix(pkg, orc):
    g, perm, tm, rm, src, bt, bv = cases.case_c1()
    ref = orc.solve(g, perm, tm, rm, src, bt, bv, tol=1e-10)
    s = _solver(pkg, g, perm)
    s.solve(tm, rm, src, bt, bv, residual_tolerance=1e-10)
    rp, col, val, rhs = s.system()
    import scipy.sparse as sp
    A = sp.csr_matrix((val, col, rp), shape=(ref.nf, ref.nf))
    R = sp.csr_matrix((ref.system.val, ref.system.col, ref.system.row_ptr), shape=(ref.nf, ref.nf))
    D = (A - R).tocoo()
    assert np.max(np.abs(D.data)) / np.max(np.abs(R.data)) < 1e-13 if D.nnz else True
    assert rel_err(rhs, ref.system.rhs) < 1e-13


def test_cell_faces_match_reference_numbering(pkg, orc):
    g = pkg.grids.cartesian(6, 5, 4)
    s = _solver(pkg, g, cases.fields.perm_c1(g.num_cells))
    hf, ifp = s.cell_faces()
    ref_dof, nint, nf, _ = orc.enumerate_dof(g)
    assert np.array_equal(hf, ref_dof)
    d = s.dims()
    assert (d.num_faces, d.num_internal_faces) == (nf, nint)


@pytest.mark.parametrize("dirs", [(0, 1, 2), (0, 1)])
def test_periodic_parity(pkg, orc, dirs):
    g = pkg.grids.cartesian(6, 5, 4)
    perm = cases.fields.perm_c1(g.num_cells, seed=3)
    bt, bv, partner = cases.periodic_bc(g, dirs, drop_dir=0, pdrop=1.0)
    tm, rm = cases.unit_fluid(g.num_cells)
    src = np.zeros(g.num_cells)
    ref = orc.solve(g, perm, tm, rm, src, bt, bv, partner=partner, tol=1e-14)
    s = _solver(pkg, g, perm, bt=bt, partner=partner)
    s.solve(tm, rm, src, bt, bv, residual_tolerance=1e-13)
    p, pi, v = s.getSolution()
    assert rel_err(v, ref.outflux) < TOL_SOL
    assert rel_err(pi, ref.face_pressure) < TOL_SOL
    assert rel_err(p, ref.cell_pressure) < TOL_SOL
    mm = s.postProcessFluxes()
    _, _, v2 = s.getSolution()
    rv, rmm = orc.post_process(g, ref.hf_dof, ref.nf, ref.ppartner, True, ref.outflux)
    assert rel_err(v2, rv) < TOL_SOL
    assert abs(mm - rmm) <= 1e-8 * max(np.max(np.abs(rv)), 1e-300)


def test_gravity_sources_mobility(pkg, orc):
    rng = np.random.default_rng(11)
    g = pkg.grids.cartesian(7, 6, 5, 2.0, 1.5, 0.5)
    nc = g.num_cells
    perm = _spd_perm(nc, rng)
    grav = (0.0, 0.0, 9.81)
    tm = rng.uniform(0.5, 2.0, nc) * 1e3
    rm = tm * rng.uniform(800.0, 1000.0, nc)
    src = np.zeros(nc)
    src[0], src[-1] = 1e-3, -1e-3
    bt = np.zeros(g.num_half_faces, np.int32)
    bv = np.zeros(g.num_half_faces)
    bt[g.side == 4] = DIR; bv[g.side == 4] = 2e5
    bt[g.side == 0] = NEU; bv[g.side == 0] = -1e-4   # inflow through x-
    ref = orc.solve(g, perm, tm, rm, src, bt, bv, grav=grav, tol=1e-14)
    s = _solver(pkg, g, perm, grav)
    for prec in ("fp64", "mixed"):
        s.solve(tm, rm, src, bt, bv, residual_tolerance=1e-13, precision=prec)
        p, pi, v = s.getSolution()
        assert rel_err(pi, ref.face_pressure) < TOL_SOL, prec
        assert rel_err(p, ref.cell_pressure) < TOL_SOL, prec
        assert rel_err(v, ref.outflux) < TOL_SOL, prec
        assert _mass_balance(g, v, src) < 1e-12


def test_batch_of_instances(pkg, orc):
    dims = [(5, 4, 3), (8, 3, 2), (4, 4, 4)]
    gs = [pkg.grids.cartesian(*d) for d in dims]
    u, inst = pkg.grids.concat(gs)
    perms = [cases.fields.perm_c1(x.num_cells, seed=20 + i) for i, x in enumerate(gs)]
    bts, bvs = zip(*[cases.xdrop_bc(x, 1.0 + i, 0.0) for i, x in enumerate(gs)])
    s = _solver(pkg, u, np.concatenate(perms), inst=inst)
    tm, rm = cases.unit_fluid(u.num_cells)
    s.solve(tm, rm, np.zeros(u.num_cells), np.concatenate(bts), np.concatenate(bvs),
            residual_tolerance=1e-13)
    p, pi, v = s.getSolution()
    hf, ifp = s.cell_faces()
    c0 = h0 = 0
    for i, x in enumerate(gs):
        ref = orc.solve(x, perms[i], *cases.unit_fluid(x.num_cells), np.zeros(x.num_cells), bts[i],
                        bvs[i], tol=1e-14)
        nc, nh = x.num_cells, x.num_half_faces
        assert rel_err(p[c0:c0 + nc], ref.cell_pressure) < TOL_SOL
        assert rel_err(v[h0:h0 + nh], ref.outflux) < TOL_SOL
        assert rel_err(pi[ifp[i]:ifp[i + 1]], ref.face_pressure) < TOL_SOL
        c0 += nc
        h0 += nh


def test_multiple_rhs_periodic(pkg, orc):
    g = pkg.grids.cartesian(6, 6, 5)
    perm = cases.fields.perm_c1(g.num_cells, seed=5)
    tm, rm = cases.unit_fluid(g.num_cells)
    src = np.zeros((3, g.num_cells))
    bvs, refs = [], []
    for d in range(3):
        bt, bv, partner = cases.periodic_bc(g, (0, 1, 2), drop_dir=d, pdrop=1.0)
        bvs.append(bv)
        refs.append(orc.solve(g, perm, tm, rm, src[d], bt, bv, partner=partner, tol=1e-14))
    s = _solver(pkg, g, perm, bt=bt, partner=partner)
    for prec in ("fp64", "mixed"):
        s.solve(tm, rm, src, bt, np.stack(bvs), residual_tolerance=1e-13, precision=prec)
        p, pi, v = s.getSolution()
        for d in range(3):
            assert rel_err(v[d], refs[d].outflux) < TOL_SOL, (prec, d)
            assert rel_err(p[d], refs[d].cell_pressure) < TOL_SOL, (prec, d)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_boundary_types_change_between_solves(pkg, orc, precision):
    """One handle, boundary types changed between solves (x drop, then a y
    drop with Neumann inflow on z-, then an invalid type, then the x drop
    again): every solve matches the oracle for its own conditions, the invalid
    type raises and leaves no solution behind."""
    g, perm, tm, rm, src, bt_x, bv_x = cases.case_c1()
    bt_y = np.zeros_like(bt_x)
    bv_y = np.zeros_like(bv_x)
    bt_y[g.side == 2] = DIR
    bv_y[g.side == 2] = 2.0e5
    bt_y[g.side == 3] = DIR
    bt_y[g.side == 4] = NEU
    bv_y[g.side == 4] = -1.0e-9
    s = _solver(pkg, g, perm, bt=bt_x)
    for case in ((bt_x, bv_x), (bt_y, bv_y), None, (bt_x, bv_x), (bt_x, bv_x)):
        if case is None:
            bad = bt_y.copy()
            bad[np.flatnonzero(g.side == 4)[0]] = 7
            with pytest.raises(RuntimeError) as e:
                s.solve(tm, rm, src, bad, bv_y, residual_tolerance=1e-12, precision=precision)
            assert e.value.status == pkg.mhs.MHS_ERR_INVALID
            with pytest.raises(RuntimeError):
                s.getSolution()
            continue
        bt, bv = case
        ref = orc.solve(g, perm, tm, rm, src, bt, bv, tol=1e-14)
        rep = s.solve(tm, rm, src, bt, bv, residual_tolerance=1e-12, precision=precision)
        assert rep.converged
        p, pi, v = s.getSolution()
        assert rel_err(pi, ref.face_pressure) < TOL_SOL
        assert rel_err(p, ref.cell_pressure) < TOL_SOL
        assert rel_err(v, ref.outflux) < TOL_SOL
        assert _mass_balance(g, v, src) < 1e-12


def test_unknown_linsolver_type_raises(pkg):
    g, perm, tm, rm, src, bt, bv = cases.case_c1()
    s = _solver(pkg, g, perm)
    with pytest.raises(RuntimeError) as e:
        s.solve(tm, rm, src, bt, bv, linsolver_type=7)
    assert e.value.status == pkg.mhs.MHS_ERR_INVALID
    assert "Unknown linsolver_type" in str(e.value)


def test_not_converged_raises(pkg):
    g, perm, tm, rm, src, bt, bv = cases.case_c1()
    s = _solver(pkg, g, perm)
    with pytest.raises(RuntimeError) as e:
        s.solve(tm, rm, src, bt, bv, residual_tolerance=1e-12, linsolver_maxit=3)
    assert e.value.status == pkg.mhs.MHS_ERR_NOT_CONVERGED
    assert "failed to converge" in str(e.value)


def test_golden_fixtures(pkg):
    d = np.load(cases.golden("c1_20x20x10.npz"))
    g, perm, tm, rm, src, bt, bv = cases.case_c1()
    s = _solver(pkg, g, perm)
    s.solve(tm, rm, src, bt, bv, residual_tolerance=1e-12, precision="mixed")
    p, pi, v = s.getSolution()
    assert rel_err(pi, d["face_pressure"]) < TOL_SOL
    assert rel_err(p, d["cell_pressure"]) < TOL_SOL
    assert rel_err(v, d["outflux"]) < TOL_SOL
    d = np.load(cases.golden("periodic_4x4x4.npz"))
    g = pkg.grids.cartesian(4, 4, 4)
    perm = cases.fields.perm_c1(g.num_cells, seed=7)
    bt, bv, partner = cases.periodic_bc(g, (0, 1, 2), drop_dir=0, pdrop=1.0)
    s = _solver(pkg, g, perm, bt=bt, partner=partner)
    s.solve(*cases.unit_fluid(g.num_cells), np.zeros(g.num_cells), bt, bv,
            residual_tolerance=1e-13, precision="mixed")
    p, pi, v = s.getSolution()
    assert rel_err(v, d["outflux"]) < TOL_SOL
    assert rel_err(p, d["cell_pressure"]) < TOL_SOL


def test_full_size_properties(pkg):
    """128x128x64 (the benchmark size), size-independent checks: with uniform
    K the discrete solution is exactly linear; with the anisotropic lognormal
    field every cell balances its fluxes and the mixed-precision solve agrees
    with the FP64 solve."""
    dims = (128, 128, 64)
    g = pkg.grids.cartesian(*dims)
    nc = g.num_cells
    bt, bv = cases.xdrop_bc(g, 1.0, 0.0)
    tm, rm = cases.unit_fluid(nc)
    src = np.zeros(nc)
    s = _solver(pkg, g, pkg.fields.diag_perm(np.full(nc, 1e-13)))
    s.solve(tm, rm, src, bt, bv, residual_tolerance=1e-10, precision="mixed")
    p, pi, v = s.getSolution()
    x = g.cell_centroid[:, 0]
    assert np.max(np.abs(p - (1.0 - x / 128.0))) < 1e-7
    perm = pkg.fields.perm_c2(dims, seed=2)
    s = _solver(pkg, g, perm)
    s.solve(tm, rm, src, bt, bv, residual_tolerance=1e-10, precision="fp64")
    p64, pi64, v64 = s.getSolution()
    s.solve(tm, rm, src, bt, bv, residual_tolerance=1e-10, precision="mixed")
    p, pi, v = s.getSolution()
    assert s.report.reduction <= 1e-10
    # Recovery sums terms of size ~|p| into per-cell flux sums of size
    # ~|p|/nx (nx = 128 cells along the drop): FP64 round-off shows up
    # amplified by about nx, so the bar here is 1e-11 instead of 1e-12.
    assert _mass_balance(g, v, src) < 1e-11
    assert rel_err(pi, pi64) < 1e-6
    assert rel_err(v, v64) < 1e-6


@pytest.mark.parametrize("kind", ["cartesian", "corner_point"])
def test_gravity_term(pkg, orc, kind):
    """MimeticIPEvaluator::gravityTerm for every cell on the device against the
    oracle (rel. 1e-14)."""
    if kind == "cartesian":
        g = pkg.grids.cartesian(7, 5, 4, 2.0, 3.0, 0.5)
    else:
        g = pkg.grids.corner_point(9, 7, 5, 10.0, 10.0, 2.0, seed=4)
    rng = np.random.default_rng(11)
    om = rng.uniform(0.1, 3.0, g.num_cells)
    grav = np.array([0.4, -0.7, 9.81])
    s = pkg.HybridSolver(0)
    s.init(g, pkg.fields.perm_c1(g.num_cells, 2), (0, 0, 0))
    got = s.gravityTerm(grav, om)
    want = orc.gravity_term(g, grav, om)
    assert np.max(np.abs(got - want)) <= 1e-14 * np.max(np.abs(want))


@pytest.mark.parametrize("ltype", [0, 2, 3])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_c1_linsolver_types(pkg, orc, ltype, precision):
    """linsolver_type 0 (Jacobi-PCG, standing in for SeqILU0-CG), 2 (KAMG)
    and 3 (FastAMG) through the ctypes binding against the oracle: every type
    converges to the same discrete solution (IncompFlowSolverHybrid.hpp:671-700)."""
    g, perm, tm, rm, src, bt, bv = cases.case_c1()
    ref = orc.solve(g, perm, tm, rm, src, bt, bv, tol=1e-14)
    s = _solver(pkg, g, perm)
    rep = s.solve(tm, rm, src, bt, bv, residual_tolerance=1e-12, precision=precision, linsolver_type=ltype)
    assert rep.converged and rep.reduction <= 1e-12
    assert rep.precond == (0 if ltype == 0 else 1)
    p, pi, v = s.getSolution()
    assert rel_err(pi, ref.face_pressure) < TOL_SOL
    assert rel_err(p, ref.cell_pressure) < TOL_SOL
    assert rel_err(v, ref.outflux) < TOL_SOL

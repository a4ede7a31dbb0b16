"""The explicit upstream transport on the GPU (include/euler_upstream.h)
against the oracle (oracle/euler_upstream.py) on the fluxes of a GPU
pressure solve: the residual term by term, whole transport solves with the
CFL step count, periodic and corner-point grids, the retry and clamp rules,
the out-of-range error, and an IMPES loop (pressure, then transport, with
mobilities from the new saturation) against the oracle's own loop.

Bars: residuals to 1e-12 of their largest entry (FP64, same fluxes, only the
summation order differs); saturations to 1e-12 after a transport solve on the
same fluxes; 1e-8 through an IMPES loop whose pressure solves stop at a
relative residual of 1e-12 (GPU) and 1e-14 (oracle)."""
import numpy as np
import pytest

import cases
import euler_upstream as tr

pytestmark = pytest.mark.gpu

VISC = (1.0e-3, 3.0e-3)
RHO = (1013.9, 834.7)
DRHO = RHO[0] - RHO[1]
GRAV = (0.0, 0.0, 9.81)
BAR = 1e5


def _fluid(sat):
    lw, lo = tr.mobilities(sat, VISC)
    return lw + lo, RHO[0] * lw + RHO[1] * lo


def _flow(pkg, g, perm, bt, bv, sat, partner=None, grav=GRAV, src=None):
    s = pkg.HybridSolver(0)
    s.init(g, perm, grav, bt, partner)
    tm, rm = _fluid(sat)
    s.solve(tm, rm, np.zeros(g.num_cells) if src is None else src, bt, bv, residual_tolerance=1e-12)
    return s, np.asarray(s.getSolution()[2]).reshape(-1)


def _cart_case(pkg, dims=(12, 10, 6), seed=3):
    g = pkg.grids.cartesian(*dims, 2.0, 2.0, 1.0)
    nc, nhf = g.num_cells, g.num_half_faces
    rng = np.random.default_rng(seed)
    perm = pkg.fields.perm_c1(nc, seed)
    bt, bv = cases.xdrop_bc(g, 2 * BAR, 0.0)
    sat = rng.uniform(0.0, 1.0, nc)
    sat_bc = rng.uniform(0.0, 1.0, nhf)
    poro = rng.uniform(0.1, 0.3, nc)
    inj = np.zeros(nc)
    inj[5], inj[nc - 7] = 3e-8, -2e-8
    return g, perm, bt, bv, sat, sat_bc, poro, inj


@pytest.mark.parametrize("methods", [(1, 1, 1), (1, 0, 0), (0, 1, 0), (0, 0, 1)])
def test_residual_matches_oracle(pkg, methods):
    g, perm, bt, bv, sat, sat_bc, poro, inj = _cart_case(pkg)
    s, flux = _flow(pkg, g, perm, bt, bv, sat)
    t = pkg.EulerUpstream(s, poro, sat_bc, VISC, RHO)
    mv, mg, mc = map(bool, methods)
    r_gpu = t.computeResidual(sat, GRAV, inj, mv, mg, mc)
    r_orc = tr.residual(g, perm, sat, flux, GRAV, DRHO, VISC, inj, sat_bc, None, mv, mg, mc)
    assert np.max(np.abs(r_orc)) > 0
    assert cases.rel_err(r_gpu, r_orc) < 1e-12


def test_residual_corner_point_full_tensor(pkg):
    g = pkg.grids.corner_point(10, 8, 5, 3.0, 3.0, 1.0, seed=4)
    nc, nhf = g.num_cells, g.num_half_faces
    rng = np.random.default_rng(9)
    perm = pkg.fields.perm_c4(nc, 4)
    bt, bv = cases.xdrop_bc(g, 3 * BAR, BAR)
    sat = rng.uniform(0.0, 1.0, nc)
    sat_bc = rng.uniform(0.0, 1.0, nhf)
    s, flux = _flow(pkg, g, perm, bt, bv, sat)
    t = pkg.EulerUpstream(s, np.full(nc, 0.2), sat_bc, VISC, RHO)
    r_gpu = t.computeResidual(sat, GRAV)
    r_orc = tr.residual(g, perm, sat, flux, GRAV, DRHO, VISC, None, sat_bc)
    assert cases.rel_err(r_gpu, r_orc) < 1e-12


def test_transport_solve_matches_oracle(pkg):
    g, perm, bt, bv, sat, sat_bc, poro, inj = _cart_case(pkg, seed=5)
    sat = np.clip(sat, 0.3, 0.7)
    s, flux = _flow(pkg, g, perm, bt, bv, sat)
    t = pkg.EulerUpstream(s, poro, sat_bc, VISC, RHO)
    cfl = tr.cfl_time(g, perm, poro, flux, GRAV, DRHO, VISC)
    T = 24.5 * cfl
    s_orc, n_orc, rep_orc, cfl_orc = tr.transport_solve(g, perm, poro, sat, T, GRAV, flux, DRHO, VISC, inj, sat_bc)
    s_gpu = sat.copy()
    t.transportSolve(s_gpu, T, GRAV, inj)
    r = t.report
    assert r.cfl_dt == pytest.approx(cfl_orc, rel=1e-13)
    # the capillary CFL time is 1e100 for this fluid (as in the reference), so
    # the capillary term can push a step out of range and force a retry
    assert (r.small_steps, r.repeats) == (n_orc, rep_orc)
    assert n_orc == 25 * 2 ** rep_orc
    fac = tr.cfl_factors(VISC)
    assert (r.cfl_factor, r.cfl_factor_gravity, r.cfl_factor_capillary) == fac
    np.testing.assert_allclose(s_gpu, s_orc, rtol=0, atol=1e-12)
    assert np.max(np.abs(s_gpu - sat)) > 1e-3


def test_transport_solve_corner_point_matches_oracle(pkg):
    """A whole transport solve on a faulted corner-point grid (cells with up to
    11 faces, rotated full tensors, gravity, a well pair): CFL time, step and
    retry counts, and saturations against the oracle on the same fluxes."""
    g = pkg.grids.corner_point(10, 8, 5, 3.0, 3.0, 1.0, seed=12, fault_frac=0.3)
    nc, nhf = g.num_cells, g.num_half_faces
    assert g.ncf().max() > 6
    rng = np.random.default_rng(21)
    perm = pkg.fields.perm_c4(nc, 12)
    bt, bv = cases.xdrop_bc(g, 3 * BAR, BAR)
    sat = rng.uniform(0.3, 0.7, nc)
    sat_bc = rng.uniform(0.0, 1.0, nhf)
    poro = rng.uniform(0.1, 0.3, nc)
    inj = np.zeros(nc)
    inj[3], inj[nc - 5] = 3e-8, -2e-8
    s, flux = _flow(pkg, g, perm, bt, bv, sat)
    t = pkg.EulerUpstream(s, poro, sat_bc, VISC, RHO)
    cfl = tr.cfl_time(g, perm, poro, flux, GRAV, DRHO, VISC)
    s_orc, n_orc, rep_orc, cfl_orc = tr.transport_solve(g, perm, poro, sat, 12.5 * cfl, GRAV, flux, DRHO, VISC,
                                                        inj, sat_bc)
    s_gpu = sat.copy()
    t.transportSolve(s_gpu, 12.5 * cfl, GRAV, inj)
    r = t.report
    assert r.cfl_dt == pytest.approx(cfl_orc, rel=1e-13)
    assert (r.small_steps, r.repeats) == (n_orc, rep_orc)
    np.testing.assert_allclose(s_gpu, s_orc, rtol=0, atol=1e-12)
    assert np.max(np.abs(s_gpu - sat)) > 1e-2


def test_periodic_transport(pkg):
    """Periodic in x and y with a pressure drop along x: the saturation
    crosses the periodic faces through the partner cells."""
    g = pkg.grids.cartesian(8, 6, 4, 2.0, 2.0, 1.0)
    nc, nhf = g.num_cells, g.num_half_faces
    bt, bv, part = cases.periodic_bc(g, (0, 1), 0, BAR)
    perm = pkg.fields.perm_c1(nc, 7)
    rng = np.random.default_rng(1)
    sat = rng.uniform(0.2, 0.8, nc)
    poro = np.full(nc, 0.25)
    s, flux = _flow(pkg, g, perm, bt, bv, sat, partner=part, grav=(0.0, 0.0, 0.0))
    sat_bc = np.full(nhf, 0.5)
    t = pkg.EulerUpstream(s, poro, sat_bc, VISC, RHO)
    r_gpu = t.computeResidual(sat, (0.0, 0.0, 0.0))
    r_orc = tr.residual(g, perm, sat, flux, (0, 0, 0), DRHO, VISC, None, sat_bc, part)
    assert cases.rel_err(r_gpu, r_orc) < 1e-12
    cfl = tr.cfl_time(g, perm, poro, flux, (0, 0, 0), DRHO, VISC)
    s_orc, n_orc, _, _ = tr.transport_solve(g, perm, poro, sat, 10.2 * cfl, (0, 0, 0), flux, DRHO, VISC, None,
                                            sat_bc, part)
    s_gpu = sat.copy()
    t.transportSolve(s_gpu, 10.2 * cfl, (0.0, 0.0, 0.0))
    assert t.report.small_steps == n_orc and n_orc % 11 == 0
    np.testing.assert_allclose(s_gpu, s_orc, rtol=0, atol=1e-12)
    # closed system: water is conserved up to what the no-flow faces carry (the
    # pressure solve's residual, stopped at 1e-12)
    assert np.sum(poro * s_gpu) == pytest.approx(np.sum(poro * sat), rel=1e-10)


@pytest.mark.parametrize("clamp", [False, True])
def test_retry_and_clamp(pkg, clamp):
    """Courant number 8: without clamping the pass overshoots and restarts
    with doubled steps; with clamping it runs once.  Both as the oracle."""
    g, perm, bt, bv, sat, sat_bc, poro, inj = _cart_case(pkg, (16, 4, 2), seed=11)
    s, flux = _flow(pkg, g, perm, bt, bv, sat, grav=(0.0, 0.0, 0.0))
    t = pkg.EulerUpstream(s, poro, sat_bc, VISC, RHO)
    t.init({"courant_number": 8.0, "clamp_sat": clamp, "method_gravity": False})
    cfl = tr.cfl_time(g, perm, poro, flux, (0, 0, 0), DRHO, VISC, 8.0, use_gravity=False)
    T = 6.0 * cfl
    s_orc, n_orc, rep_orc, _ = tr.transport_solve(g, perm, poro, sat, T, (0, 0, 0), flux, DRHO, VISC, None, sat_bc,
                                                  courant_number=8.0, method_gravity=False, clamp_sat=clamp)
    s_gpu = sat.copy()
    t.transportSolve(s_gpu, T, (0.0, 0.0, 0.0))
    assert (t.report.small_steps, t.report.repeats) == (n_orc, rep_orc)
    assert (rep_orc == 0) if clamp else (rep_orc >= 1)
    np.testing.assert_allclose(s_gpu, s_orc, rtol=0, atol=1e-12)


def test_out_of_range_raises_and_restores(pkg):
    """An injection that overfills a closed box fails for every step count:
    RuntimeError after the retries, the saturation given back unchanged."""
    g = pkg.grids.cartesian(4, 3, 2)
    nc, nhf = g.num_cells, g.num_half_faces
    bt = np.zeros(nhf, np.int32)
    bv = np.zeros(nhf)
    src = np.zeros(nc)
    s, _ = _flow(pkg, g, pkg.fields.perm_c1(nc, 2), bt, bv, np.full(nc, 0.1), grav=(0.0, 0.0, 0.0), src=src)
    t = pkg.EulerUpstream(s, np.full(nc, 0.2), None, VISC, RHO)
    inj = np.zeros(nc)
    inj[0] = 1e-3
    sat = np.full(nc, 0.1)
    with pytest.raises(RuntimeError, match="Saturation out of range"):
        t.transportSolve(sat, 1e3, (0.0, 0.0, 0.0), inj)
    assert np.all(sat == 0.1)
    with pytest.raises(tr.SaturationOutOfRange):
        tr.transport_solve(g, pkg.fields.perm_c1(nc, 2), np.full(nc, 0.2), np.full(nc, 0.1), 1e3, (0, 0, 0),
                           np.zeros(nhf), DRHO, VISC, inj)


def test_impes_loop_matches_oracle(pkg, orc):
    """SimulatorTester::run (examples/SimulatorTester.hpp:59-107): pressure
    with the mobilities of the current saturation, then transport over the
    step; three steps, same structure, changing matrix."""
    g = pkg.grids.cartesian(10, 8, 4, 2.0, 2.0, 1.0)
    nc, nhf = g.num_cells, g.num_half_faces
    perm = pkg.fields.perm_c1(nc, 12)
    bt, bv = cases.xdrop_bc(g, 2 * BAR, 0.0)
    sat_bc = np.zeros(nhf)
    sat_bc[g.side == 0] = 1.0           # water enters through x-
    poro = np.full(nc, 0.2)
    sat0 = np.zeros(nc)
    s = pkg.HybridSolver(0)
    s.init(g, perm, GRAV, bt, None)
    t = pkg.EulerUpstream(s, poro, sat_bc, VISC, RHO)
    step = 2.0e5
    s_gpu = sat0.copy()
    s_orc = sat0.copy()
    for _ in range(3):
        tm, rm = _fluid(s_gpu)
        s.solve(tm, rm, np.zeros(nc), bt, bv, residual_tolerance=1e-12)
        t.transportSolve(s_gpu, step, GRAV)
        tm, rm = _fluid(s_orc)
        ref = orc.solve(g, perm, tm, rm, np.zeros(nc), bt, bv, grav=GRAV, tol=1e-14)
        s_orc, *_ = tr.transport_solve(g, perm, poro, s_orc, step, GRAV, ref.outflux, DRHO, VISC, None, sat_bc)
    assert np.max(s_orc) > 0.3
    np.testing.assert_allclose(s_gpu, s_orc, rtol=0, atol=1e-8)


def test_device_path_matches_host_path(pkg):
    """transportSolve_device and mobilities_device on torch tensors give the
    host path's saturation and the host mobilities bit for bit."""
    import torch
    g, perm, bt, bv, sat, sat_bc, poro, inj = _cart_case(pkg, (8, 8, 4), seed=13)
    sat = np.clip(sat, 0.3, 0.7)
    s, flux = _flow(pkg, g, perm, bt, bv, sat)
    t = pkg.EulerUpstream(s, poro, sat_bc, VISC, RHO)
    cfl = tr.cfl_time(g, perm, poro, flux, GRAV, DRHO, VISC)
    host = sat.copy()
    t.transportSolve(host, 7.5 * cfl, GRAV, inj)
    d_sat = torch.tensor(sat, dtype=torch.float64, device="cuda")
    d_inj = torch.tensor(inj, dtype=torch.float64, device="cuda")
    t.transportSolve_device(d_sat.data_ptr(), 7.5 * cfl, GRAV, d_inj.data_ptr())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d_sat.cpu().numpy(), host)
    tm = torch.empty_like(d_sat)
    rm = torch.empty_like(d_sat)
    t.mobilities_device(d_sat.data_ptr(), tm.data_ptr(), rm.data_ptr())
    torch.cuda.synchronize()
    tm_h, rm_h = _fluid(host)
    np.testing.assert_allclose(tm.cpu().numpy(), tm_h, rtol=1e-15, atol=0)
    np.testing.assert_allclose(rm.cpu().numpy(), rm_h, rtol=1e-15, atol=0)


def test_coefficients_follow_permeability_and_gravity(pkg):
    """The gravity and capillary coefficients are kept between calls: a new
    gravity vector, the capillary switch and a new permeability must each
    reach the residual."""
    g, perm, bt, bv, sat, sat_bc, poro, inj = _cart_case(pkg, (8, 6, 5), seed=17)
    s, flux = _flow(pkg, g, perm, bt, bv, sat)
    t = pkg.EulerUpstream(s, poro, sat_bc, VISC, RHO)
    for grav, cap in ((GRAV, True), ((0.0, 3.0, -9.81), True), ((0.0, 3.0, -9.81), False), ((0.0, 0.0, 0.0), True)):
        r_gpu = t.computeResidual(sat, grav, inj, method_capillary=cap)
        r_orc = tr.residual(g, perm, sat, flux, grav, DRHO, VISC, inj, sat_bc, None, method_capillary=cap)
        assert cases.rel_err(r_gpu, r_orc) < 1e-12, (grav, cap)
    perm2 = pkg.fields.perm_c1(g.num_cells, 23)
    s.computeInnerProducts(perm2, GRAV)
    tm, rm = _fluid(sat)
    s.solve(tm, rm, np.zeros(g.num_cells), bt, bv, residual_tolerance=1e-12)
    flux2 = np.asarray(s.getSolution()[2]).reshape(-1)
    r_gpu = t.computeResidual(sat, GRAV, inj)
    r_orc = tr.residual(g, perm2, sat, flux2, GRAV, DRHO, VISC, inj, sat_bc)
    assert cases.rel_err(r_gpu, r_orc) < 1e-12


def test_estimate_cell_velocity(pkg, orc):
    """estimateCellVelocity on the GPU against the oracle on the same fluxes
    (faulted corner-point grid), and the closed form of a uniform flow: a
    linear pressure drop on a box with uniform K gives u = lambda K dp / L in
    every cell."""
    import torch
    g = pkg.grids.corner_point(10, 8, 5, 3.0, 3.0, 1.0, seed=6)
    nc = g.num_cells
    perm = pkg.fields.perm_c4(nc, 6)
    bt, bv = cases.xdrop_bc(g, 3 * BAR, BAR)
    s, flux = _flow(pkg, g, perm, bt, bv, np.full(nc, 0.4))
    v_gpu = s.estimateCellVelocity()
    v_orc = orc.estimate_cell_velocity(g, flux)
    assert cases.rel_err(v_gpu, v_orc) < 1e-13
    d = torch.empty(3 * nc, dtype=torch.float64, device="cuda")
    s.estimate_cell_velocity_device(d.data_ptr())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.cpu().numpy().reshape(nc, 3), v_gpu)
    gb = pkg.grids.cartesian(12, 5, 4, 2.0, 1.0, 1.0)
    k = 150 * 9.869233e-16
    K = np.zeros((gb.num_cells, 9))
    K[:, [0, 4, 8]] = k
    bt, bv = cases.xdrop_bc(gb, 2 * BAR, 0.0)
    sb = pkg.HybridSolver(0)
    sb.init(gb, K, (0.0, 0.0, 0.0), bt, None)
    sb.solve(np.full(gb.num_cells, 1e3), np.zeros(gb.num_cells), np.zeros(gb.num_cells), bt, bv,
             residual_tolerance=1e-13)
    u = 1e3 * k * 2 * BAR / 24.0
    vb = sb.estimateCellVelocity()
    np.testing.assert_allclose(vb[:, 0], u, rtol=1e-9)
    assert np.max(np.abs(vb[:, 1:])) < 1e-9 * u


def test_simulator_output_fields(pkg, tmp_path):
    """writeVtkOutput's cell data after a pressure solve: the phase
    velocities split the cell velocity by the fractional flow, the pressure
    is the solution's, and the VTU file holds every field."""
    g, perm, bt, bv, sat, sat_bc, poro, inj = _cart_case(pkg, (6, 5, 4), seed=21)
    s, flux = _flow(pkg, g, perm, bt, bv, sat)
    f = pkg.output.simulator_fields(s, sat, VISC)
    np.testing.assert_allclose(f["phase velocity [water]"] + f["phase velocity [oil]"], f["velocity"],
                               rtol=1e-14, atol=1e-30)
    np.testing.assert_allclose(f["fractional flow [water]"], tr.fractional_flow(sat, VISC), rtol=1e-15)
    np.testing.assert_array_equal(f["pressure"], s.getSolution()[0])
    zc = np.broadcast_to(np.arange(5) * 1.0, (5, 6, 4, 5)).copy()
    P, conn = pkg.output.hexahedra(6, 5, 4, 2.0, 2.0, zc)
    pkg.output.write_vtu(str(tmp_path / "o.vtu"), P, conn, f)
    txt = (tmp_path / "o.vtu").read_text()
    assert all(f'Name="{k}"' in txt for k in f)


def test_transport_handle_lifetime_and_stale_structure(pkg):
    """A new structure on the bound solver makes the transport ask for a new
    init; destroying the solver before the transport handle is safe."""
    import ctypes
    g, perm, bt, bv, sat, sat_bc, poro, inj = _cart_case(pkg, (6, 4, 3), seed=31)
    s, flux = _flow(pkg, g, perm, bt, bv, sat)
    t = pkg.EulerUpstream(s, poro, sat_bc, VISC, RHO)
    t.computeResidual(sat, GRAV)
    s.initSystemStructure(bt)
    s.computeInnerProducts(perm, GRAV)
    tm, rm = _fluid(sat)
    s.solve(tm, rm, np.zeros(g.num_cells), bt, bv, residual_tolerance=1e-12)
    with pytest.raises(pkg.MhsError, match="mhs_transport_init again"):
        t.computeResidual(sat, GRAV)
    t.initObj(poro, sat_bc, VISC, RHO)
    r_gpu = t.computeResidual(sat, GRAV)
    r_orc = tr.residual(g, perm, sat, np.asarray(s.getSolution()[2]).reshape(-1), GRAV, DRHO, VISC, None, sat_bc)
    assert cases.rel_err(r_gpu, r_orc) < 1e-12
    L = pkg.transport.lib()
    h = ctypes.c_void_p()
    assert L.mhs_transport_create(ctypes.byref(h), s._h) == 0
    L.mhs_destroy(s._h)
    s._h = None
    assert L.mhs_transport_destroy(h) == 0


def test_transport_on_a_batch_with_two_rhs(pkg):
    """A handle holding two uneven instances, solved with two right-hand
    sides: the transport runs on every cell of the batch with the fluxes of
    right-hand side 0.  One CFL time (the smallest over the batch) sets the
    small steps of every instance, so the batch equals the oracle's transport
    on the concatenated grid, and an instance with that CFL time equals the
    oracle on its own grid."""
    gs = [pkg.grids.cartesian(8, 6, 3, 2.0, 2.0, 1.0), pkg.grids.cartesian(5, 7, 4, 2.0, 2.0, 1.0)]
    G, icp = pkg.grids.concat(gs)
    nc, nhf = G.num_cells, G.num_half_faces
    perm = pkg.fields.perm_c1(nc, 41)
    bt, bv = cases.xdrop_bc(G, 2 * BAR, 0.0)
    bv2 = np.stack([bv, 0.5 * bv])
    rng = np.random.default_rng(41)
    sat = rng.uniform(0.3, 0.7, nc)
    sat_bc = np.where(G.side == 0, 1.0, 0.0)
    poro = np.full(nc, 0.2)
    s = pkg.HybridSolver(0)
    s.init(G, perm, GRAV, bt, None, inst_cellpos=icp)
    tm, rm = _fluid(sat)
    s.solve(tm, rm, np.zeros((2, nc)), bt, bv2, residual_tolerance=1e-12)
    flux = np.asarray(s.getSolution()[2])[0]
    t = pkg.EulerUpstream(s, poro, sat_bc, VISC, RHO)
    t.init({"method_capillary": False})
    T = 2.0 * tr.cfl_time(G, perm, poro, flux, GRAV, DRHO, VISC, use_capillary=False)
    s_gpu = sat.copy()
    t.transportSolve(s_gpu, T, GRAV)
    whole, n, _, _ = tr.transport_solve(G, perm, poro, sat, T, GRAV, flux, DRHO, VISC, None, sat_bc,
                                        method_capillary=False)
    assert t.report.small_steps == n
    np.testing.assert_allclose(s_gpu, whole, rtol=0, atol=1e-12)
    same = 0
    for i, g in enumerate(gs):
        c0, c1 = icp[i], icp[i + 1]
        h0, h1 = G.facepos[c0], G.facepos[c1]
        ref, n_i, _, _ = tr.transport_solve(g, perm[c0:c1], poro[c0:c1], sat[c0:c1], T, GRAV, flux[h0:h1], DRHO,
                                            VISC, None, sat_bc[h0:h1], method_capillary=False)
        if n_i == n:
            np.testing.assert_allclose(s_gpu[c0:c1], ref, rtol=0, atol=1e-12)
            same += 1
    assert same >= 1

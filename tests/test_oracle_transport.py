"""Pins of the transport oracle (oracle/euler_upstream.py) against what the
mathematics fixes: the Buckley-Leverett solution, closed-form two-point
gravity and capillary fluxes, conservation, translation invariance on a
periodic ring, the CFL times of uniform flows, the sampled CFL factors
against the analytic fractional-flow derivatives, and the retry rule."""
import numpy as np
import pytest
from scipy.optimize import brentq

import euler_upstream as tr

VISC = (1.0e-3, 3.0e-3)
DRHO = 1013.9 - 834.7
MD = 9.869233e-16


def _uniform_x_flux(g, q):
    """Outflux of a uniform flow q (m^3/s per unit area) along +x."""
    return q * np.asarray(g.face_normal)[:, 0] * np.asarray(g.face_area)


def _iso(nc, k):
    K = np.zeros((nc, 9))
    K[:, [0, 4, 8]] = k
    return K


def _dfw(s, mu):
    a, b = s * s / mu[0], (1 - s) ** 2 / mu[1]
    da, db = 2 * s / mu[0], -2 * (1 - s) / mu[1]
    return (da * b - a * db) / (a + b) ** 2


def _dfg(s, mu):
    a, b = s * s / mu[0], (1 - s) ** 2 / mu[1]
    da, db = 2 * s / mu[0], -2 * (1 - s) / mu[1]
    return ((da * b + a * db) * (a + b) - a * b * (da + db)) / (a + b) ** 2


@pytest.mark.parametrize("mu", [(1.0, 1.0), VISC, (5e-3, 1e-3)])
def test_cfl_factors_are_inverse_max_slopes(mu):
    fac, fac_g, fac_c = tr.cfl_factors(mu)
    s = np.linspace(0, 1, 200001)
    assert fac_c == 0.0
    assert 1.0 / fac == pytest.approx(np.max(np.abs(_dfw(s, mu))), rel=2e-3)
    assert 1.0 / fac_g == pytest.approx(np.max(np.abs(_dfg(s, mu))), rel=2e-3)
    if mu == (1.0, 1.0):
        assert 1.0 / fac <= 2.0 and 1.0 / fac > 1.999   # f'(1/2) = 2, sampled at 1/256


def test_buckley_leverett(pkg):
    """Water injected at x = 0 into a 1-D oil column: the upstream scheme
    approaches the entropy solution (rarefaction f'(s) = x / (u t / phi) down to
    the shock at s*, f(s*) / s* = f'(s*)), and conserves water exactly."""
    n, phi, q = 400, 0.25, 1e-6
    g = pkg.grids.cartesian(n, 1, 1)
    nc = g.num_cells
    flux = _uniform_x_flux(g, q)
    K = _iso(nc, 100 * MD)
    fw = lambda s: tr.fractional_flow(s, VISC)
    s_star = brentq(lambda s: _dfw(s, VISC) * s - fw(s), 0.05, 0.99)
    speed = q / phi * fw(s_star) / s_star
    T = 0.6 * n / speed
    s, steps, rep, _ = tr.transport_solve(g, K, np.full(nc, phi), np.zeros(nc), T, (0, 0, 0), flux, DRHO, VISC,
                                          method_gravity=False, method_capillary=False)
    assert rep == 0 and steps > 100
    x = np.asarray(g.cell_centroid)[:, 0]
    xi = x / (q * T / phi)
    grid_s = np.linspace(s_star, 1.0, 20001)
    exact = np.interp(xi, _dfw(grid_s, VISC)[::-1], grid_s[::-1], left=1.0, right=0.0)
    exact[xi > _dfw(s_star, VISC)] = 0.0
    assert np.mean(np.abs(s - exact)) < 0.02
    # front position within a few cells, saturation behind it in range
    front_num = x[np.nonzero(s > 0.5 * s_star)[0].max()]
    assert abs(front_num - speed * T) < 6.0
    # conservation before breakthrough: pore volume times the saturation gain
    # equals the water injected
    assert np.sum(phi * s) == pytest.approx(q * T, rel=1e-11)


def test_gravity_segregation_closed_column(pkg):
    """No flux, uniform s = 1/2 in a vertical column: water (heavier) moves
    along gravity through every internal face with the two-point rate
    lambda_w lambda_o / lambda_t k drho g A; interior cells do not change."""
    g = pkg.grids.cartesian(1, 1, 8)
    nc = g.num_cells
    k = 200 * MD
    grav = (0.0, 0.0, 9.81)
    r = tr.residual(g, _iso(nc, k), np.full(nc, 0.5), np.zeros(g.num_half_faces), grav, DRHO, VISC, None,
                    np.ones(g.num_half_faces), method_capillary=False)
    lw, lo = tr.mobilities(0.5, VISC)
    rate = lw * lo / (lw + lo) * k * DRHO * 9.81
    z = np.asarray(g.cell_centroid)[:, 2]
    lo_c, hi_c = np.argmin(z), np.argmax(z)
    assert r[lo_c] == pytest.approx(-rate, rel=1e-13)
    assert r[hi_c] == pytest.approx(rate, rel=1e-13)
    mid = np.ones(nc, bool)
    mid[[lo_c, hi_c]] = False
    assert np.all(r[mid] == 0.0)


def test_capillary_exchange_two_cells(pkg):
    """No flux and no gravity: water moves from the wetter cell to the drier
    one at lambda_w lambda_o / lambda_t (at the mean saturation) times the
    two-point transmissibility k A / dx times the capillary pressure jump."""
    g = pkg.grids.cartesian(2, 1, 1, 2.0, 1.5, 0.5)
    k = 50 * MD
    s = np.array([0.8, 0.2])
    r = tr.residual(g, _iso(2, k), s, np.zeros(g.num_half_faces), (0, 0, 0), DRHO, VISC, None,
                    np.ones(g.num_half_faces))
    lw, lo = tr.mobilities(0.5, VISC)
    A, dx = 1.5 * 0.5, 2.0
    rate = lw * lo / (lw + lo) * k * A / dx * (tr.cap_pressure(0.2) - tr.cap_pressure(0.8))
    assert r[0] == pytest.approx(-rate, rel=1e-13)
    assert r[1] == pytest.approx(rate, rel=1e-13)


def test_periodic_ring_translation_invariance(pkg):
    """On a ring of identical cells with a uniform flux through the periodic
    boundary, shifting the initial saturation by one cell shifts the result by
    one cell; water is conserved and a uniform state stays uniform."""
    n = 12
    g = pkg.grids.cartesian(n, 1, 1)
    part = pkg.grids.periodic_partners_cartesian(g, (0,))
    nc = g.num_cells
    flux = _uniform_x_flux(g, 2e-7)
    K = _iso(nc, 100 * MD)
    poro = np.full(nc, 0.2)
    s0 = np.zeros(nc)
    s0[3] = 0.9
    s0[4] = 0.5
    kw = dict(partner=part, method_gravity=False)
    a, *_ = tr.transport_solve(g, K, poro, s0, 3e5, (0, 0, 0), flux, DRHO, VISC, **kw)
    b, *_ = tr.transport_solve(g, K, poro, np.roll(s0, 1), 3e5, (0, 0, 0), flux, DRHO, VISC, **kw)
    np.testing.assert_allclose(b, np.roll(a, 1), rtol=0, atol=1e-14)
    assert np.sum(a) == pytest.approx(np.sum(s0), rel=1e-13)
    assert not np.allclose(a, s0)
    u, *_ = tr.transport_solve(g, K, poro, np.full(nc, 0.3), 3e5, (0, 0, 0), flux, DRHO, VISC, **kw)
    np.testing.assert_allclose(u, 0.3, rtol=0, atol=1e-15)


def test_cfl_times_of_uniform_flows(pkg):
    phi, q = 0.2, 3e-7
    fac, fac_g, _ = tr.cfl_factors(VISC)
    g = pkg.grids.cartesian(10, 1, 1)
    nc = g.num_cells
    K = _iso(nc, 100 * MD)
    dt = tr.cfl_time(g, K, np.full(nc, phi), _uniform_x_flux(g, q), (0, 0, 0), DRHO, VISC, 1.0,
                     use_gravity=False, use_capillary=False)
    assert dt == pytest.approx(fac * phi / q, rel=1e-14)
    g = pkg.grids.cartesian(1, 1, 10)
    k = 300 * MD
    dt = tr.cfl_time(g, _iso(nc, k), np.full(nc, phi), np.zeros(g.num_half_faces), (0, 0, 9.81), DRHO, VISC,
                     0.5, use_viscous=False)
    assert dt == pytest.approx(0.5 * fac_g * phi / (k * 9.81 * DRHO), rel=1e-13)
    # the capillary factor is 0: that CFL time stays 1e100
    dt = tr.cfl_time(g, _iso(nc, k), np.full(nc, phi), np.zeros(g.num_half_faces), (0, 0, 0), DRHO, VISC, 1.0,
                     use_viscous=False, use_gravity=False)
    assert dt == 1e100


def test_small_steps_rule():
    assert tr.small_steps(10.0, 5.0) == 1
    assert tr.small_steps(10.0, 5.0, minimum_small_steps=3) == 3
    assert tr.small_steps(1.0, 10.5) == 11
    assert tr.small_steps(1e-9, 1.0) == 10000


def test_retry_restarts_with_doubled_steps(pkg):
    """A Courant number far above 1 overshoots; the retried pass must equal a
    pass run directly with the final number of steps."""
    n, phi, q = 60, 0.25, 1e-6
    g = pkg.grids.cartesian(n, 1, 1)
    nc = g.num_cells
    flux = _uniform_x_flux(g, q)
    K = _iso(nc, 100 * MD)
    T = 20 * phi / q
    kw = dict(method_gravity=False, method_capillary=False)
    s, steps, rep, cfl = tr.transport_solve(g, K, np.full(nc, phi), np.zeros(nc), T, (0, 0, 0), flux, DRHO, VISC,
                                            courant_number=8.0, **kw)
    assert rep >= 1 and steps == tr.small_steps(cfl, T) * 2 ** rep
    d, steps2, rep2, _ = tr.transport_solve(g, K, np.full(nc, phi), np.zeros(nc), T, (0, 0, 0), flux, DRHO, VISC,
                                            courant_number=8.0, minimum_small_steps=steps,
                                            maximum_small_steps=steps, check_sat=False, **kw)
    assert rep2 == 0 and steps2 == steps
    np.testing.assert_array_equal(s, d)
    assert np.all((s >= -0.001) & (s <= 1.001))


def test_saturation_out_of_range_raises(pkg):
    """An injection that overfills a closed box fails for every step count."""
    g = pkg.grids.cartesian(3, 1, 1)
    nc = g.num_cells
    inj = np.array([1e-6, 0.0, 0.0])
    with pytest.raises(tr.SaturationOutOfRange):
        tr.transport_solve(g, _iso(nc, 100 * MD), np.full(nc, 0.2), np.zeros(nc), 1e6, (0, 0, 0),
                           np.zeros(g.num_half_faces), DRHO, VISC, inj, method_gravity=False,
                           method_capillary=False)

"""Pins of the oracle's tensor-mobility inner products
(MimeticIPAnisoRelpermEvaluator): isotropic mobility tensors reduce to the
scalar evaluator; a linear pressure is reproduced exactly with full mobility
tensors that share K's principal axes; hydrostatic equilibrium carries no
flux."""
import numpy as np
import pytest

import cases


def _rotation(a, b, c):
    ca, sa, cb, sb, cc, sc = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    Rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    Ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    Rx = np.array([[1, 0, 0], [0, cc, -sc], [0, sc, cc]])
    return Rz @ Ry @ Rx


def _commuting(K, a, b):
    """a I + b K / |K|: a mobility tensor with K's principal axes."""
    nc = K.shape[0]
    Km = K.reshape(nc, 3, 3)
    nrm = np.linalg.norm(Km, axis=(1, 2))[:, None, None]
    return (a[:, None, None] * np.eye(3) + b[:, None, None] * Km / nrm).reshape(nc, 9)


def test_isotropic_tensors_reduce_to_scalar_mobilities(pkg, orc):
    g = pkg.grids.corner_point(6, 5, 4, 3.0, 3.0, 1.0, seed=2)
    nc = g.num_cells
    perm = pkg.fields.perm_c4(nc, 2)
    rng = np.random.default_rng(4)
    lam = rng.uniform(300.0, 1000.0, nc)
    rl = lam * rng.uniform(800.0, 1000.0, nc)
    bt, bv = cases.xdrop_bc(g, 2e5, 0.0)
    grav = (0.0, 0.0, 9.81)
    ref = orc.solve(g, perm, lam, rl, np.zeros(nc), bt, bv, grav=grav, tol=1e-14)
    ani = orc.solve_aniso(g, perm, lam[:, None] * np.eye(3).reshape(1, 9), rl[:, None] * np.eye(3).reshape(1, 9),
                          np.zeros(nc), bt, bv, grav=grav, tol=1e-14)
    assert cases.rel_err(ani.outflux, ref.outflux) < 1e-10
    assert cases.rel_err(ani.cell_pressure, ref.cell_pressure) < 1e-10


@pytest.mark.parametrize("kind", ["cartesian", "corner_point"])
def test_linear_pressure_is_exact_with_full_tensors(pkg, orc, kind):
    if kind == "cartesian":
        g = pkg.grids.cartesian(5, 4, 3, 2.0, 1.5, 1.0)
    else:
        g = pkg.grids.corner_point(5, 4, 3, 2.0, 1.5, 1.0, perturb=0.0, fault_frac=0.3, seed=5)
    nc = g.num_cells
    R = _rotation(0.4, 0.3, 0.2)
    K0 = 1e-13 * R @ np.diag([1.0, 0.3, 0.1]) @ R.T
    perm = np.tile(K0.reshape(1, 9), (nc, 1))
    lam = _commuting(perm, np.full(nc, 200.0), np.full(nc, 500.0))
    grad = np.array([-3e4, 1e4, 2e4])
    bt = np.where(g.nb < 0, 1, 0).astype(np.int32)
    bv = np.where(g.nb < 0, np.asarray(g.face_centroid) @ grad + 1e5, 0.0)
    r = orc.solve_aniso(g, perm, lam, np.zeros((nc, 9)), np.zeros(nc), bt, bv, tol=1e-15)
    lK = lam[0].reshape(3, 3) @ K0
    assert np.allclose(lK, lK.T, rtol=1e-12, atol=0)
    exact = -np.asarray(g.face_area) * (np.asarray(g.face_normal) @ (lK @ grad))
    assert cases.rel_err(r.outflux, exact) < 1e-9
    np.testing.assert_allclose(r.cell_pressure, np.asarray(g.cell_centroid) @ grad + 1e5, rtol=1e-12)


def test_hydrostatic_equilibrium_has_no_flux(pkg, orc):
    """rho_lambda = rho lambda_t and p = rho g.x on every boundary face: the
    gravity term balances the pressure gradient everywhere."""
    g = pkg.grids.cartesian(4, 3, 5, 2.0, 2.0, 1.0)
    nc = g.num_cells
    perm = pkg.fields.perm_c4(nc, 3)
    rng = np.random.default_rng(8)
    lam = _commuting(perm, rng.uniform(100, 300, nc), rng.uniform(100, 300, nc))
    rho, grav = 950.0, np.array([0.0, 0.0, 9.81])
    bt = np.where(g.nb < 0, 1, 0).astype(np.int32)
    bv = np.where(g.nb < 0, rho * (np.asarray(g.face_centroid) @ grav), 0.0)
    r = orc.solve_aniso(g, perm, lam, rho * lam, np.zeros(nc), bt, bv, grav=grav, tol=1e-15)
    scale = np.max(np.abs(orc.aniso_inner_products(g, perm, lam, rho * lam, grav)[2]))
    assert np.max(np.abs(r.outflux)) < 1e-9 * scale

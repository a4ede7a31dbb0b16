"""Pin of the oracle's estimateCellVelocity: for the fluxes of any uniform
velocity u (outflux = A u.n), the divergence theorem gives
sum_f (x_f - x_c) A_f u.n_f = |c| u on every closed polyhedral cell, so the
estimate returns u to round-off.  That needs planar faces (area centroid
and normal then integrate x and n exactly): Cartesian cells and faulted
corner-point cells with unperturbed layers, whose split fault faces are
planar rectangles.  With perturbed (non-planar) layers the identity holds
only approximately, and the estimate is compared with the oracle instead."""
import numpy as np
import pytest


@pytest.mark.parametrize("kind", ["cartesian", "corner_point"])
def test_uniform_velocity_is_reproduced(pkg, orc, kind):
    if kind == "cartesian":
        g = pkg.grids.cartesian(5, 4, 3, 2.0, 1.0, 0.5)
    else:
        g = pkg.grids.corner_point(6, 5, 4, 3.0, 2.0, 1.0, perturb=0.0, fault_frac=0.3, seed=3)
        assert np.any(g.ncf() > 6)          # split fault faces present
    u = np.array([1.5e-6, -0.7e-6, 0.3e-6])
    flux = np.asarray(g.face_area) * (np.asarray(g.face_normal) @ u)
    v = orc.estimate_cell_velocity(g, flux)
    np.testing.assert_allclose(v, np.broadcast_to(u, v.shape), rtol=0, atol=1e-12 * np.linalg.norm(u) * 1e3)

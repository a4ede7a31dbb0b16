"""Tensor mobilities (MimeticIPAnisoRelpermEvaluator) on the GPU through the
C-ABI: against the oracle on a faulted corner-point grid with rotated
permeability and gravity (FP64 and mixed precision, 1e-8); the linear
pressure closed form; isotropic tensors against the scalar solve; the
static inner products untouched for the next scalar solve; the adjoint
refused after a tensor solve."""
import numpy as np
import pytest

import cases
from test_oracle_aniso import _commuting, _rotation

pytestmark = pytest.mark.gpu

GRAV = (0.0, 0.0, 9.81)


def _case(pkg, seed=3):
    g = pkg.grids.corner_point(10, 8, 5, 3.0, 3.0, 1.0, seed=seed)
    nc = g.num_cells
    perm = pkg.fields.perm_c4(nc, seed)
    rng = np.random.default_rng(seed)
    lam = _commuting(perm, rng.uniform(200.0, 600.0, nc), rng.uniform(100.0, 800.0, nc))
    rl = _commuting(perm, rng.uniform(2e5, 6e5, nc), rng.uniform(1e5, 8e5, nc))
    bt, bv = cases.xdrop_bc(g, 3e5, 1e5)
    return g, perm, lam, rl, bt, bv


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tensor_mobility_matches_oracle(pkg, orc, precision):
    g, perm, lam, rl, bt, bv = _case(pkg)
    nc = g.num_cells
    s = pkg.HybridSolver(0)
    s.init(g, perm, GRAV, bt, None)
    s.solve_aniso(lam, rl, np.zeros(nc), bt, bv, residual_tolerance=1e-12, precision=precision)
    p, pi, v = s.getSolution()
    ref = orc.solve_aniso(g, perm, lam, rl, np.zeros(nc), bt, bv, grav=GRAV, tol=1e-14)
    assert cases.rel_err(v, ref.outflux) < 1e-8
    assert cases.rel_err(p, ref.cell_pressure) < 1e-8
    # the static inner products are untouched: a scalar solve after it
    tm = np.full(nc, 400.0)
    rm = tm * 900.0
    s.solve(tm, rm, np.zeros(nc), bt, bv, residual_tolerance=1e-12, precision=precision)
    ref2 = orc.solve(g, perm, tm, rm, np.zeros(nc), bt, bv, grav=GRAV, tol=1e-14)
    assert cases.rel_err(s.getSolution()[2], ref2.outflux) < 1e-8


def test_linear_pressure_closed_form(pkg):
    g = pkg.grids.cartesian(12, 10, 8, 2.0, 1.5, 1.0)
    nc = g.num_cells
    R = _rotation(0.4, 0.3, 0.2)
    K0 = 1e-13 * R @ np.diag([1.0, 0.3, 0.1]) @ R.T
    perm = np.tile(K0.reshape(1, 9), (nc, 1))
    lam = _commuting(perm, np.full(nc, 200.0), np.full(nc, 500.0))
    grad = np.array([-3e4, 1e4, 2e4])
    bt = np.where(g.nb < 0, 1, 0).astype(np.int32)
    bv = np.where(g.nb < 0, np.asarray(g.face_centroid) @ grad + 1e5, 0.0)
    s = pkg.HybridSolver(0)
    s.init(g, perm, (0.0, 0.0, 0.0), bt, None)
    s.solve_aniso(lam, np.zeros((nc, 9)), np.zeros(nc), bt, bv, residual_tolerance=1e-13)
    lK = lam[0].reshape(3, 3) @ K0
    exact = -np.asarray(g.face_area) * (np.asarray(g.face_normal) @ (lK @ grad))
    assert cases.rel_err(s.getSolution()[2], exact) < 1e-9


def test_isotropic_tensors_equal_scalar_solve_and_device_path(pkg):
    import torch
    g, perm, _, _, bt, bv = _case(pkg, seed=5)
    nc = g.num_cells
    rng = np.random.default_rng(2)
    lam = rng.uniform(300.0, 900.0, nc)
    rl = lam * 950.0
    s = pkg.HybridSolver(0)
    s.init(g, perm, GRAV, bt, None)
    s.solve(lam, rl, np.zeros(nc), bt, bv, residual_tolerance=1e-13)
    v0 = s.getSolution()[2]
    eye = np.eye(3).reshape(1, 9)
    s.solve_aniso(lam[:, None] * eye, rl[:, None] * eye, np.zeros(nc), bt, bv, residual_tolerance=1e-13)
    v1 = s.getSolution()[2]
    assert cases.rel_err(v1, v0) < 1e-10
    d = lambda a: torch.tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")
    L, Rl, z, B = d(lam[:, None] * eye), d(rl[:, None] * eye), d(np.zeros(nc)), d(bv)
    s.solve_aniso_device(L.data_ptr(), Rl.data_ptr(), z.data_ptr(), bt, B.data_ptr(), tol=1e-13)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(s.getSolution()[2], v1)
    with pytest.raises(pkg.MhsError, match="tensor-mobility"):
        s.adjoint(np.ones(g.num_half_faces))


def test_tensor_mobility_batch_two_rhs_and_cell_velocity(pkg, orc):
    """Two uneven instances and two right-hand sides in one tensor-mobility
    solve: every (instance, right-hand side) as the oracle on its own grid,
    and estimateCellVelocity of the second right-hand side as the oracle's."""
    gs = [pkg.grids.corner_point(7, 6, 4, 3.0, 3.0, 1.0, seed=11), pkg.grids.cartesian(6, 5, 5, 2.0, 2.0, 1.0)]
    G, icp = pkg.grids.concat(gs)
    nc = G.num_cells
    perm = pkg.fields.perm_c4(nc, 11)
    rng = np.random.default_rng(11)
    lam = _commuting(perm, rng.uniform(200.0, 600.0, nc), rng.uniform(100.0, 800.0, nc))
    rl = _commuting(perm, rng.uniform(2e5, 6e5, nc), rng.uniform(1e5, 8e5, nc))
    bt, bv = cases.xdrop_bc(G, 3e5, 1e5)
    bv2 = np.stack([bv, 0.3 * bv])
    s = pkg.HybridSolver(0)
    s.init(G, perm, GRAV, bt, None, inst_cellpos=icp)
    s.solve_aniso(lam, rl, np.zeros((2, nc)), bt, bv2, residual_tolerance=1e-12, precision="mixed")
    p, pi, v = s.getSolution()
    vel = s.estimateCellVelocity(1)
    for t, g in enumerate(gs):
        c0, c1 = icp[t], icp[t + 1]
        h0, h1 = G.facepos[c0], G.facepos[c1]
        for k in range(2):
            ref = orc.solve_aniso(g, perm[c0:c1], lam[c0:c1], rl[c0:c1], np.zeros(c1 - c0), bt[h0:h1],
                                  bv2[k, h0:h1], grav=GRAV, tol=1e-14)
            assert cases.rel_err(v[k, h0:h1], ref.outflux) < 1e-8, (t, k)
            assert cases.rel_err(p[k, c0:c1], ref.cell_pressure) < 1e-8, (t, k)
            if k == 1:
                ve = orc.estimate_cell_velocity(g, v[1, h0:h1])
                assert cases.rel_err(vel[c0:c1], ve) < 1e-13

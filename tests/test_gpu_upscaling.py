"""Batched periodic upscaling on the GPU: many uneven, fully periodic cell
problems and three pressure drops in one solve, against the oracle per
instance, and K_eff from the device flux sums against the closed forms."""
import numpy as np
import pytest

from test_oracle_upscaling import _oracle_flux_sums

pytestmark = pytest.mark.gpu


def _solve(pkg, batch, perm, precision, tol=1e-10):
    from opm_porsol_amd import upscaling as U
    s = pkg.HybridSolver(0)
    s.init(batch.grid, perm, (0, 0, 0), batch.bc_type, batch.partner, inst_cellpos=batch.inst_cellpos)
    nc = batch.grid.num_cells
    bv = U.pressure_drop_values(batch)
    s.solve(np.ones(nc), np.zeros(nc), np.zeros((3, nc)), batch.bc_type, bv,
            residual_tolerance=tol, precision=precision)
    return s


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_periodic_matches_oracle(pkg, orc, precision):
    from opm_porsol_amd import upscaling as U
    dims = [(6, 6, 6), (9, 7, 8), (4, 5, 6), (10, 10, 10), (7, 7, 7)]
    b = U.periodic_batch(dims)
    K = U.lognormal_isotropic(dims, seed=4, mean_perm=1.0)
    # pure-periodic (regularised) systems: the flux error is the residual times
    # the condition number, so parity at 1e-8 asks for a 1e-11 residual
    s = _solve(pkg, b, K, precision, tol=1e-11)
    assert s.report.precond == 1 and s.report.converged
    p, pi, v = s.getSolution()
    g = b.grid
    for t, d in enumerate(dims):
        c0, c1 = b.inst_cellpos[t], b.inst_cellpos[t + 1]
        h0, h1 = g.facepos[c0], g.facepos[c1]
        gt = pkg.grids.cartesian(*d)
        pp = pkg.grids.periodic_partners_cartesian(gt, (0, 1, 2))
        bt = np.where(gt.nb < 0, 2, 0).astype(np.int32)
        for j in range(3):
            bvt = np.zeros(gt.num_half_faces)
            bvt[gt.side == 2 * j] = 1.0
            bvt[gt.side == 2 * j + 1] = -1.0
            n = c1 - c0
            r = orc.solve(gt, K[c0:c1], np.ones(n), np.zeros(n), np.zeros(n), bt, bvt,
                          partner=pp, tol=1e-14)
            ref_v = r.outflux
            assert np.max(np.abs(v[j, h0:h1] - ref_v)) <= 1e-8 * np.abs(ref_v).max()
            assert np.max(np.abs(p[j, c0:c1] - r.cell_pressure)) <= 1e-8 * np.abs(r.cell_pressure).max()
    sums = s.sum_outflux(b.group_ptr, b.group_hf)
    Keff = U.effective_permeability(sums, b)
    Kref = U.effective_permeability(_oracle_flux_sums(pkg, orc, b, K), b)
    assert np.max(np.abs(Keff - Kref)) <= 1e-8 * np.abs(Kref).max()
    # the device sums equal host sums of the returned fluxes
    host = np.stack([[v[j, b.group_hf[b.group_ptr[q]:b.group_ptr[q + 1]]].sum()
                      for q in range(3 * b.ninst)] for j in range(3)])
    assert np.max(np.abs(sums - host)) <= 1e-13 * np.abs(host).max()


def test_homogeneous_tensor_batch_upscales_to_itself(pkg):
    from opm_porsol_amd import upscaling as U
    dims = [(8, 8, 8), (12, 9, 10), (16, 16, 16)]
    b = U.periodic_batch(dims)
    M = np.array([[2.0, 0.3, 0.1], [0.3, 1.0, 0.2], [0.1, 0.2, 0.5]]) * 1e-13
    K = np.tile(M.reshape(1, 9), (b.grid.num_cells, 1))
    s = _solve(pkg, b, K, "mixed")
    Keff = U.effective_permeability(s.sum_outflux(b.group_ptr, b.group_hf), b)
    for t in range(b.ninst):
        assert np.max(np.abs(Keff[t] - M)) <= 1e-8 * np.abs(M).max()

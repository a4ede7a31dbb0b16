"""Faulted corner-point grids on the GPU (cells with 6 to 14 faces, full
rotated permeability tensors, Dirichlet injector / producer corner columns)
against the oracle."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _case(pkg, dims=(16, 14, 8), frac=0.2, seed=5):
    g = pkg.grids.corner_point(*dims, 10.0, 10.0, 2.0, fault_frac=frac, seed=seed)
    nc = g.num_cells
    K = pkg.fields.perm_c4(nc, seed=seed)
    bt = np.zeros(g.num_half_faces, np.int32)
    bv = np.zeros(g.num_half_faces)
    ij = g.ijk[g.hf_cell()]
    inj = (ij[:, 0] == 0) & (ij[:, 1] == 0) & np.isin(g.side, [0, 2])
    prod = (ij[:, 0] == dims[0] - 1) & (ij[:, 1] == dims[1] - 1) & np.isin(g.side, [1, 3])
    bt[inj | prod] = 1
    bv[inj] = 1.0e5
    return g, K, bt, bv, inj, prod


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_faulted_grid_matches_oracle(pkg, orc, precision):
    g, K, bt, bv, inj, prod = _case(pkg)
    assert g.ncf().max() > 10          # exercises the 16-face cell kernels
    nc = g.num_cells
    tm = np.full(nc, 1.0 / 1.0e-3)
    z = np.zeros(nc)
    s = pkg.HybridSolver(0)
    s.init(g, K, (0, 0, 9.81), bt, None)
    s.solve(tm, 800.0 * tm, z, bt, bv, residual_tolerance=1e-11, precision=precision)
    assert s.report.precond == 1
    p, pi, v = s.getSolution()
    ref = orc.solve(g, K, tm, 800.0 * tm, z, bt, bv, grav=(0, 0, 9.81), tol=1e-14)
    assert cases.rel_err(ref.outflux, v) < 1e-8
    assert cases.rel_err(ref.cell_pressure, p) < 1e-8
    assert cases.rel_err(ref.face_pressure, pi) < 1e-8
    # injector and producer rates balance (no sources)
    q_in, q_out = v[inj].sum(), v[prod].sum()
    assert q_in < 0 and abs(q_in + q_out) <= 1e-9 * abs(q_in)
    binv, gfl = s.inner_products()
    rb, _, rg = orc.static_ip(g, K, (0, 0, 9.81))
    assert np.max(np.abs(binv - rb)) <= 1e-12 * np.abs(rb).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_faulted_batch_two_rhs_matches_oracle(pkg, orc, precision):
    """Two faulted grids of different sizes in one batch, each with two right-
    hand sides (the corner-column pressure drop, then twice the drop plus cell
    sources), with gravity: every instance and right-hand side against the
    oracle's solve of that instance alone."""
    cases_ = [_case(pkg, dims=(16, 14, 8), frac=0.2, seed=5), _case(pkg, dims=(12, 10, 6), frac=0.3, seed=9)]
    gs = [c[0] for c in cases_]
    u, inst = pkg.grids.concat(gs)
    K = np.concatenate([c[1] for c in cases_])
    bt = np.concatenate([c[2] for c in cases_])
    rng = np.random.default_rng(11)
    src = np.zeros((2, u.num_cells))
    src[1] = 1e-4 * rng.standard_normal(u.num_cells)
    bv = np.stack([np.concatenate([c[3] for c in cases_]), 2.0 * np.concatenate([c[3] for c in cases_])])
    tm = np.full(u.num_cells, 1.0 / 1.0e-3)
    grav = (0, 0, 9.81)
    s = pkg.HybridSolver(0)
    s.init(u, K, grav, bt, None, inst_cellpos=inst)
    s.solve(tm, 800.0 * tm, src, bt, bv, residual_tolerance=1e-11, precision=precision)
    p, pi, v = s.getSolution()
    hf, ifp = s.cell_faces()
    c0 = h0 = 0
    for i, g in enumerate(gs):
        nc, nh = g.num_cells, g.num_half_faces
        tmi = tm[c0:c0 + nc]
        for r in range(2):
            ref = orc.solve(g, K[c0:c0 + nc], tmi, 800.0 * tmi, src[r, c0:c0 + nc], bt[h0:h0 + nh],
                            bv[r, h0:h0 + nh], grav=grav, tol=1e-14)
            assert cases.rel_err(ref.outflux, v[r, h0:h0 + nh]) < 1e-8, (i, r)
            assert cases.rel_err(ref.cell_pressure, p[r, c0:c0 + nc]) < 1e-8, (i, r)
            assert cases.rel_err(ref.face_pressure, pi[r, ifp[i]:ifp[i + 1]]) < 1e-8, (i, r)
            bal = np.bincount(g.hf_cell(), weights=v[r, h0:h0 + nh], minlength=nc) - src[r, c0:c0 + nc]
            assert np.max(np.abs(bal)) < 1e-10 * np.max(np.abs(v[r, h0:h0 + nh])), (i, r)
        c0 += nc
        h0 += nh

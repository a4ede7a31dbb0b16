"""Periodic upscaling on the oracle: the K_eff formula and the batch helpers.

K_eff is not computed anywhere in the reference tree (it lives in the upscaling
codes built on it); these closed forms pin the formula the batched GPU workload
uses: a homogeneous full tensor upscales to itself, a layered medium to the
harmonic mean across the layers and the arithmetic mean along them.
"""
import numpy as np


def _oracle_flux_sums(pkg, orc, batch, perm, mobility=1.0):
    from opm_porsol_amd import upscaling as U  # noqa: F401
    sums = np.zeros((3, 3 * batch.ninst))
    bv = None
    for t, d in enumerate(batch.dims):
        c0, c1 = batch.inst_cellpos[t], batch.inst_cellpos[t + 1]
        g = pkg.grids.cartesian(*d)
        pp = pkg.grids.periodic_partners_cartesian(g, (0, 1, 2))
        bt = np.where(g.nb < 0, 2, 0).astype(np.int32)
        n = c1 - c0
        for j in range(3):
            bv = np.zeros(g.num_half_faces)
            bv[g.side == 2 * j] = 1.0
            bv[g.side == 2 * j + 1] = -1.0
            r = orc.solve(g, perm[c0:c1], np.full(n, mobility), np.zeros(n), np.zeros(n), bt, bv,
                          partner=pp, tol=1e-14)
            for i in range(3):
                sums[j, 3 * t + i] = r.outflux[g.side == 2 * i + 1].sum()
    return sums


def test_homogeneous_full_tensor_upscales_to_itself(pkg, orc):
    from opm_porsol_amd import upscaling as U
    dims = [(4, 3, 5), (3, 3, 3)]
    b = U.periodic_batch(dims)
    M = np.array([[2.0, 0.3, 0.1], [0.3, 1.0, 0.2], [0.1, 0.2, 0.5]])
    K = np.tile(M.reshape(1, 9), (b.grid.num_cells, 1))
    Keff = U.effective_permeability(_oracle_flux_sums(pkg, orc, b, K, 2.0), b, mobility=2.0)
    for t in range(b.ninst):
        assert np.allclose(Keff[t], M, rtol=0, atol=1e-11)


def test_layered_medium_harmonic_and_arithmetic_means(pkg, orc):
    from opm_porsol_amd import upscaling as U
    dims = [(6, 2, 3)]
    b = U.periodic_batch(dims)
    kx_layer = np.array([1.0, 5.0, 0.2, 3.0, 0.7, 2.0])     # varies along x
    g = b.grid
    i = np.arange(g.num_cells) % 6
    k = kx_layer[i]
    K = pkg.fields.diag_perm(k, k, k)
    Keff = U.effective_permeability(_oracle_flux_sums(pkg, orc, b, K), b)[0]
    harm = len(kx_layer) / np.sum(1.0 / kx_layer)
    arit = np.mean(kx_layer)
    assert abs(Keff[0, 0] - harm) <= 1e-11 * harm
    assert abs(Keff[1, 1] - arit) <= 1e-11 * arit
    assert abs(Keff[2, 2] - arit) <= 1e-11 * arit
    assert np.max(np.abs(Keff - np.diag(np.diag(Keff)))) <= 1e-12 * arit


def test_heterogeneous_upscaled_tensor_is_symmetric_positive(pkg, orc):
    from opm_porsol_amd import upscaling as U
    dims = [(5, 4, 4)]
    b = U.periodic_batch(dims)
    K = U.lognormal_isotropic(dims, seed=9, mean_perm=1.0)
    Keff = U.effective_permeability(_oracle_flux_sums(pkg, orc, b, K), b)[0]
    assert np.max(np.abs(Keff - Keff.T)) <= 1e-10 * np.abs(Keff).max()
    assert np.all(np.linalg.eigvalsh(0.5 * (Keff + Keff.T)) > 0)
    # bounded by the harmonic and arithmetic means of the cell values
    k = K[:, 0]
    lo, hi = 1.0 / np.mean(1.0 / k), np.mean(k)
    assert np.all(np.diag(Keff) >= lo * (1 - 1e-12)) and np.all(np.diag(Keff) <= hi * (1 + 1e-12))


def test_batch_helpers(pkg):
    from opm_porsol_amd import upscaling as U
    dims = [(4, 3, 5), (2, 2, 2)]
    b = U.periodic_batch(dims, spacing=(0.5, 1.0, 2.0))
    g = b.grid
    assert np.allclose(b.length[0], [2.0, 3.0, 10.0]) and np.allclose(b.area[0], [30.0, 20.0, 6.0])
    # every boundary half-face is periodic and paired both ways, within its instance
    bnd = np.nonzero(g.nb < 0)[0]
    assert np.all(b.bc_type[bnd] == 2) and np.all(b.partner[bnd] >= 0)
    assert np.array_equal(b.partner[b.partner[bnd]], bnd)
    cell = g.hf_cell()
    inst = np.searchsorted(b.inst_cellpos, cell, side="right") - 1
    assert np.array_equal(inst[bnd], inst[b.partner[bnd]])
    # groups: the + side of each direction of each instance, with its area
    for t in range(2):
        for i in range(3):
            h = b.group_hf[b.group_ptr[3 * t + i]:b.group_ptr[3 * t + i + 1]]
            assert np.all(g.side[h] == 2 * i + 1) and np.all(inst[h] == t)
            assert abs(g.face_area[h].sum() - b.area[t, i]) <= 1e-12
    bv = U.pressure_drop_values(b)
    for j in range(3):
        assert np.all(bv[j][g.side == 2 * j] == 1.0) and np.all(bv[j][g.side == 2 * j + 1] == -1.0)
        other = (g.side >= 0) & (g.side // 2 != j)
        assert np.all(bv[j][other] == 0.0)
    sizes = U.random_sizes(200, 16, 48, seed=1)
    n = np.array([s[0] for s in sizes])
    assert n.min() >= 16 and n.max() <= 48 and len(set(n)) > 20

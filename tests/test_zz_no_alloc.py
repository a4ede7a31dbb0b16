"""mhs_sum_outflux_noalloc (staging in the library's per-GPU scratch array)
against mhs_sum_outflux, through the C-ABI: bit-identical sums on two calls in
a row, the second with a smaller input, and the size limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch_solve(pkg, dims):
    from opm_porsol_amd import upscaling as U
    b = U.periodic_batch(dims)
    K = U.lognormal_isotropic(dims, seed=9, mean_perm=1.0)
    s = pkg.HybridSolver(0)
    s.init(b.grid, K, (0, 0, 0), b.bc_type, b.partner, inst_cellpos=b.inst_cellpos)
    nc = b.grid.num_cells
    s.solve(np.ones(nc), np.zeros(nc), np.zeros((3, nc)), b.bc_type, U.pressure_drop_values(b),
            residual_tolerance=1e-10, precision="mixed")
    return s, b


def test_noalloc_group_sums_match_allocating_entry_point(pkg):
    s, b = _batch_solve(pkg, [(6, 6, 6), (9, 7, 8), (4, 5, 6), (10, 10, 10)])
    # first call: every + side of every instance
    ref = s.sum_outflux(b.group_ptr, b.group_hf)
    got = s.sum_outflux_noalloc(b.group_ptr, b.group_hf)
    assert got.shape == ref.shape and np.array_equal(got, ref)
    # second call, smaller: the groups of the first two instances, other half-faces
    gp = b.group_ptr[:7].copy()
    gh = b.group_hf[:gp[-1]].copy()
    gh = gh[::-1].copy()               # the same members in another order inside the span
    gp2 = np.array([0, gp[-1] // 3, gp[-1]], np.int32)
    ref2 = s.sum_outflux(gp2, gh)
    got2 = s.sum_outflux_noalloc(gp2, gh)
    assert got2.shape == (3, 2) and np.array_equal(got2, ref2)
    # the sums are those of the returned fluxes
    v = s.getSolution()[2]
    host = np.stack([[v[j, gh[gp2[q]:gp2[q + 1]]].sum() for q in range(2)] for j in range(3)])
    assert np.max(np.abs(got2 - host)) <= 1e-13 * np.abs(host).max()


def test_noalloc_rejects_staging_beyond_scratch(pkg):
    s, b = _batch_solve(pkg, [(4, 4, 4)])
    n = (16 << 20) // 4           # ints alone fill the 16 MiB array
    gp = np.array([0, n], np.int32)
    gh = np.zeros(n, np.int32)
    with pytest.raises(pkg.MhsError, match="exceed the 16777216-byte scratch"):
        s.sum_outflux_noalloc(gp, gh)
    # a later call still works
    ref = s.sum_outflux(b.group_ptr, b.group_hf)
    assert np.array_equal(s.sum_outflux_noalloc(b.group_ptr, b.group_hf), ref)

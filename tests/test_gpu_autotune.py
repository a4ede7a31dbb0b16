"""The launch-shape autotuner: it picks among valid shapes, tunes once per
structure / precision / preconditioner / nrhs, and changes results only at
rounding level."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autotuned_shapes_give_the_same_solution(pkg, precision):
    g, perm, tm, rm, src, bt, bv = cases.case_c1()
    out = {}
    for tune in (False, True):
        s = pkg.HybridSolver(0)
        s.init(g, perm, (0, 0, 0), bt, None)
        s.solve(tm, rm, src, bt, bv, residual_tolerance=1e-12, precision=precision, autotune=tune)
        r = s.report
        if tune:
            assert r.spmv_threads in (256, 512, 1024) and r.spmv_xcd in (0, 1)
            assert r.sweep_threads in (256, 512, 1024) and r.sweep_xcd in (0, 1)
            assert r.tune_ms > 0.0
            s.solve(tm, rm, src, bt, bv, residual_tolerance=1e-12, precision=precision)
            assert s.report.tune_ms == 0.0                       # tuned once
            assert (s.report.spmv_threads, s.report.sweep_threads) == (r.spmv_threads, r.sweep_threads)
        else:
            assert (r.spmv_threads, r.spmv_xcd, r.sweep_threads, r.sweep_xcd) == (256, 1, 256, 1)
        out[tune] = s.getSolution()
    for a, b in zip(out[False], out[True]):
        assert cases.rel_err(a, b) < 1e-9


@pytest.mark.parametrize("threads,tail", [(256, 0), (512, 300), (1024, 1 << 30)])
def test_every_shape_matches_the_oracle(pkg, orc, threads, tail):
    """Each candidate workgroup size, forced, for the SpMV and the sweep, and
    coarse-level tails from none to every coarse level, on a batch of uneven
    instances with two right-hand sides."""
    gs = [pkg.grids.cartesian(*d) for d in ((20, 20, 10), (9, 13, 7))]
    G, icp = pkg.grids.concat(gs)
    nc = G.num_cells
    perm = pkg.fields.perm_c1(nc, 3)
    bt, bv = cases.xdrop_bc(G)
    tm, rm = cases.unit_fluid(nc)
    src = np.zeros((2, nc))
    bv2 = np.stack([bv, 0.5 * bv])
    s = pkg.HybridSolver(0)
    s.init(G, perm, (0, 0, 0), bt, None, inst_cellpos=icp)
    for prec in ("fp64", "mixed"):
        s.solve(tm, rm, src, bt, bv2, residual_tolerance=1e-11, precision=prec,
                spmv_threads=threads, sweep_threads=threads, tail_rows=tail)
        assert (s.report.spmv_threads, s.report.sweep_threads) == (threads, threads)
        assert s.report.tail_rows == tail and s.report.precond == 1
        p, pi, v = s.getSolution()
        for t in range(2):
            c0, c1 = icp[t], icp[t + 1]
            h0, h1 = G.facepos[c0], G.facepos[c1]
            for k in range(2):
                ref = orc.solve(gs[t], perm[c0:c1], tm[c0:c1], rm[c0:c1], np.zeros(c1 - c0),
                                bt[h0:h1], bv2[k, h0:h1], tol=1e-14)
                assert cases.rel_err(ref.outflux, v[k, h0:h1]) < 1e-8
        assert s.time_kernel(0, reps=3) > 0.0
    with pytest.raises(pkg.MhsError, match="spmv_threads"):
        s.solve(tm, rm, src, bt, bv2, spmv_threads=300)

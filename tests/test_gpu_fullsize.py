"""The configured workloads at their full sizes on the GPU.

128x128x64 (headline): the mixed-precision and FP64 solves at a 1e-10
relative residual against the oracle's ILU0-CG solution to 1e-14
(tests/golden/c2_128x128x64.npz: 32768 seeded samples of face pressure, cell
pressure and flux, whole-field norms and the flow rate through every x-plane),
at 1e-8 relative, and per-cell mass balance (IncompFlowSolverHybrid.hpp:
1324-1420, 1672-1730).

The others through properties that hold at any size:
- 256 periodic cell problems 16^3..48^3 x 3 drops: FP64 true residual, mass
  balance, mixed against FP64, K_eff symmetric and between the harmonic and
  arithmetic means of each instance;
- 200x200x50 faulted corner-point grid: the same residual, balance and
  precision checks, equal injection and production;
- 32 x 64x64x32 adjoint: the batched mixed-precision gradient against central
  finite differences (h = 1e-6) of FP64 solves at 20 cells of one
  realization, and against the FP64 gradient.
"""
import numpy as np
import pytest

import cases
from cases import golden

pytestmark = pytest.mark.gpu


def _mass_balance(g, v, src=None):
    bal = np.bincount(g.hf_cell(), weights=np.asarray(v).reshape(-1), minlength=g.num_cells)
    if src is not None:
        bal = bal - src
    return float(np.max(np.abs(bal)) / np.max(np.abs(v)))


def _balance_against_oracle(g, perm, tm, rm, src, bt, bv, sols):
    """Per-cell mass balance sum_f v_f - src_c of each GPU solution (p, pi, v)
    in sols.  Recovery makes it zero up to round-off, and the round-off
    scales with the terms recovery combines: p = (g + F.pi)/L carries a
    relative error of the order of the machine epsilon times sum |F_f pi_f|,
    which L p turns into an imbalance (absolute pressures of 1e5 Pa against
    cell pressure drops of a few Pa give 1e-12 of the largest flux).  So the
    bars are (a) the imbalance measured in units of sum_f |F_f pi_f| + |g_c|
    + |src_c| and (b) the oracle's own recovery (IncompFlowSolverHybrid.hpp:
    1672-1730) from the same face pressures."""
    import oracle
    nc, nhf = g.num_cells, g.num_half_faces
    hf, nint, nf, _ = oracle.enumerate_dof(g)
    pp, _ = oracle.periodic_dof(g, hf, nf, bt, np.full(nhf, -1, np.int32))
    binv, pos, gfl = oracle.static_ip(g, perm)
    S = oracle.assemble(g, hf, nf, pp, binv, gfl, tm, rm, src, bt, bv)
    cell = g.hf_cell()
    for name, (p, pi, v) in sols.items():
        pr, vr = oracle.recover(g, hf, binv, gfl, tm, rm, S, pi)
        # the same recovery in another summation order: round-off apart
        # (5.6e-12 of max |v| on the faulted grid, where F cancels)
        assert np.max(np.abs(vr - v)) <= 1e-10 * np.abs(v).max()
        bal = np.abs(np.bincount(cell, weights=v, minlength=nc) - src)
        bal_ref = np.abs(np.bincount(cell, weights=vr, minlength=nc) - src)
        unit = np.bincount(cell, weights=np.abs(S.F * pi[hf]), minlength=nc) + np.abs(S.g) + np.abs(src)
        rel_cell = float(np.max(bal / np.maximum(unit, 1e-300)))
        print(f"{name}: mass balance {bal.max() / np.abs(v).max():.2e} of max |v|, "
              f"{rel_cell:.2e} of sum |F pi| + |g|; the oracle's recovery {bal_ref.max() / np.abs(vr).max():.2e}")
        assert rel_cell <= 1e-13, (name, rel_cell)
        # Both imbalances are round-off of the same size class: the worst cell
        # of either recovery depends on the last bits of pi (2.5x apart on the
        # faulted grid between two runs), so the bar is an order of magnitude
        # of the oracle's own, far below what a wrong term would leave.
        assert bal.max() <= 10.0 * bal_ref.max() + 1e-14 * np.abs(v).max(), name


@pytest.fixture(scope="module")
def c2():
    g, perm, bt, bv = cases.c2_problem()
    fx = np.load(golden("c2_128x128x64.npz"))
    assert str(fx["perm_checksum"]) == cases.fields.checksum(perm)
    return g, perm, bt, bv, fx


# Bars against the oracle's 1e-14 solution.  A solve stopped at a 1e-10
# relative residual is as close as the conditioning allows: the oracle's own
# ILU0-CG stopped at 1e-10 lies 2.2e-8 (face pressure), 2.1e-8 (cell pressure)
# and 4.7e-8 (flux) from it (recorded in the fixture), so at 1e-10 the
# pressures are held to 1e-8 and the fluxes to 2e-8 (the GPU multigrid solves
# land at about 1.4e-8); at 1e-12 everything is held to 1e-9.
C2_BARS = {1e-10: {"face_pressure": 1e-8, "cell_pressure": 1e-8, "outflux": 2e-8},
           1e-12: {"face_pressure": 1e-9, "cell_pressure": 1e-9, "outflux": 1e-9}}


@pytest.mark.parametrize("tol", [1e-10, 1e-12])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_c2_headline_matches_oracle(pkg, c2, precision, tol):
    g, perm, bt, bv, fx = c2
    nc = g.num_cells
    s = pkg.HybridSolver(0)
    s.init(g, perm, (0.0, 0.0, 0.0), bt, None)
    rep = s.solve(np.ones(nc), np.zeros(nc), np.zeros(nc), bt, bv, residual_tolerance=tol,
                  precision=precision)
    assert rep.converged and rep.reduction <= tol and rep.precond == 1
    p, pi, v = s.getSolution()
    errs = {}
    for k, (name, a) in enumerate((("face_pressure", pi), ("cell_pressure", p), ("outflux", v))):
        idx = cases.c2_sample_index(a.shape[0], k)
        ref = fx[name + "_sample"]
        nrm = fx[name + "_norms"]
        errs[name] = float(np.max(np.abs(a[idx] - ref)) / nrm[0])
        got = np.array([np.max(np.abs(a)), np.linalg.norm(a)])
        assert np.all(np.abs(got - nrm[:2]) <= C2_BARS[tol][name] * nrm[:2]), (name, got, nrm)
    print(f"{precision} at {tol:g}: errors against the oracle {errs}")
    for name, e in errs.items():
        assert e < C2_BARS[tol][name], (name, e)
        # never further from the oracle than its own ILU0-CG at 1e-10
        assert e < float(fx[name + "_ilu0_1e10_err"])
    rates = cases.xplane_rates(g, v)
    ref_rates = fx["xplane_rates"]
    assert np.max(np.abs(rates - ref_rates)) <= C2_BARS[tol]["outflux"] * np.abs(ref_rates).max()
    # Without sources every x-plane carries the same flow.  The rates come
    # from the x- half-faces of each column, whose twins on the column before
    # agree only up to the solver residual (measured 7e-9 of the rate at a
    # 1e-10 residual), so the bar is the flux bar.
    assert np.max(np.abs(rates - rates.mean())) <= C2_BARS[tol]["outflux"] * abs(rates.mean())
    if tol == 1e-10:
        _balance_against_oracle(g, perm, np.ones(nc), np.zeros(nc), np.zeros(nc), bt, bv, {precision: (p, pi, v)})


# --- 256 periodic cell problems, 16^3 .. 48^3, three drops ---------------------------------------

@pytest.fixture(scope="module")
def c3(pkg):
    from opm_porsol_amd import upscaling as U
    sizes = U.random_sizes(256, 16, 48, seed=2)
    b = U.periodic_batch(sizes)
    K = U.lognormal_isotropic(sizes, seed=2, sigma=1.0)
    s = pkg.HybridSolver(0)
    s.init(b.grid, K, (0, 0, 0), b.bc_type, b.partner, inst_cellpos=b.inst_cellpos)
    nc = b.grid.num_cells
    bv = U.pressure_drop_values(b)
    out = {}
    for prec in ("fp64", "mixed"):
        rep = s.solve(np.ones(nc), np.zeros(nc), np.zeros((3, nc)), b.bc_type, bv,
                      residual_tolerance=1e-10, precision=prec)
        assert rep.converged and rep.reduction <= 1e-10, (prec, rep.reduction)
        p, _, v = s.getSolution()
        out[prec] = (p, v, U.effective_permeability(s.sum_outflux(b.group_ptr, b.group_hf), b))
    return b, K, out


def test_c3_mass_balance_and_precision(pkg, c3):
    """Mass balance of every right-hand side; mixed precision against FP64
    cell by cell.  The fully periodic cell problems are nearly singular in the
    constant mode (one regularised entry), so two solves stopped at 1e-10
    differ in their fluxes by up to about 2e-7 of the largest (measured
    1.6e-7 at these sizes), while their K_eff agree to 1e-8 (next test); at a
    1e-12 residual the fluxes agree to 1e-8."""
    from opm_porsol_amd import upscaling as U
    b, K, out = c3
    g = b.grid
    for prec in ("fp64", "mixed"):
        for j in range(3):
            mb = _mass_balance(g, out[prec][1][j])
            print(f"{prec} drop {j}: mass balance {mb:.2e} of max |v|")
            assert mb <= 1e-11, (prec, j, mb)
    (p64, v64, _), (p32, v32, _) = out["fp64"], out["mixed"]
    assert np.max(np.abs(v32 - v64)) <= 5e-7 * np.abs(v64).max()
    s = pkg.HybridSolver(0)
    s.init(g, K, (0, 0, 0), b.bc_type, b.partner, inst_cellpos=b.inst_cellpos)
    nc = g.num_cells
    sol = {}
    for prec in ("fp64", "mixed"):
        rep = s.solve(np.ones(nc), np.zeros(nc), np.zeros((3, nc)), b.bc_type, U.pressure_drop_values(b),
                      residual_tolerance=1e-12, precision=prec)
        assert rep.converged and rep.reduction <= 1e-12
        sol[prec] = s.getSolution()
    dv = np.max(np.abs(sol["mixed"][2] - sol["fp64"][2])) / np.abs(sol["fp64"][2]).max()
    dp = np.max(np.abs(sol["mixed"][0] - sol["fp64"][0])) / np.abs(sol["fp64"][0]).max()
    print(f"at 1e-12: mixed against FP64 flux {dv:.2e}, cell pressure {dp:.2e}")
    assert dv <= 1e-8 and dp <= 1e-8


def test_c3_effective_permeability_bounds(c3):
    b, K, out = c3
    for prec in ("fp64", "mixed"):
        Keff = out[prec][2]
        for t in range(b.ninst):
            k = K[b.inst_cellpos[t]:b.inst_cellpos[t + 1], 0]
            kh, ka = 1.0 / np.mean(1.0 / k), np.mean(k)
            d = np.diag(Keff[t])
            assert np.all(d > kh) and np.all(d < ka), (prec, t, d, kh, ka)
            assert np.max(np.abs(Keff[t] - Keff[t].T)) <= 1e-6 * np.max(d), (prec, t)
    assert np.max(np.abs(out["mixed"][2] - out["fp64"][2])) <= 1e-8 * np.abs(out["fp64"][2]).max()


# --- 200x200x50 faulted corner-point grid -------------------------------------------------------

def test_c4_faulted_full_size(pkg):
    dims = (200, 200, 50)
    g = pkg.grids.corner_point(*dims, 10.0, 10.0, 2.0, seed=2)
    perm = pkg.fields.perm_c4(g.num_cells, seed=2)
    nc, nhf = g.num_cells, g.num_half_faces
    ij = g.ijk[g.hf_cell()]
    inj = (ij[:, 0] == 0) & (ij[:, 1] == 0) & np.isin(g.side, [0, 2])
    prod = (ij[:, 0] == dims[0] - 1) & (ij[:, 1] == dims[1] - 1) & np.isin(g.side, [1, 3])
    bt = np.zeros(nhf, np.int32)
    bt[inj | prod] = 1
    bv = np.where(inj, 1.0e5, 0.0)
    s = pkg.HybridSolver(0)
    s.init(g, perm, (0.0, 0.0, 0.0), bt, None)
    tm, z = np.full(nc, 1.0e3), np.zeros(nc)
    sol = {}
    for prec in ("fp64", "mixed"):
        rep = s.solve(tm, z, z, bt, bv, residual_tolerance=1e-10, precision=prec)
        assert rep.converged and rep.reduction <= 1e-10 and rep.precond == 1, (prec, rep.reduction)
        sol[prec] = s.getSolution()
        v = sol[prec][2]
        qi, qp = -v[inj].sum(), v[prod].sum()
        assert qi > 0 and abs(qi - qp) <= 1e-9 * qi
    _balance_against_oracle(g, perm, tm, z, z, bt, bv, sol)
    (p64, pi64, v64), (p32, pi32, v32) = sol["fp64"], sol["mixed"]
    assert np.max(np.abs(pi32 - pi64)) <= 1e-8 * np.abs(pi64).max()
    assert np.max(np.abs(p32 - p64)) <= 1e-8 * np.abs(p64).max()
    assert np.max(np.abs(v32 - v64)) <= 1e-8 * np.abs(v64).max()


# --- 32 x 64x64x32 adjoint ----------------------------------------------------------------------

def test_c5_adjoint_full_size(pkg):
    dims, R = (64, 64, 32), 32
    mD, bar = 9.869233e-16, 1.0e5
    grids_, perms, refs = [], [], []
    for r in range(R):
        grids_.append(pkg.grids.cartesian(*dims))
        lnk = np.log(100 * mD) + pkg.fields.gaussian_filtered(dims, corr_len=6.0, seed=2, stream=r)
        perms.append(pkg.fields.diag_perm(np.exp(lnk)))
        lnr = np.log(100 * mD) + pkg.fields.gaussian_filtered(dims, corr_len=6.0, seed=3, stream=r)
        refs.append(pkg.fields.diag_perm(np.exp(lnr)))
    G, icp = pkg.grids.concat(grids_)
    K, Kref = np.concatenate(perms), np.concatenate(refs)
    nc, nhf = G.num_cells, G.num_half_faces
    bt = np.zeros(nhf, np.int32)
    bt[(G.side == 0) | (G.side == 1)] = 1
    bv = np.where(G.side == 0, bar, 0.0)
    cell = G.hf_cell()
    loc = cell - icp[np.searchsorted(icp, cell, side="right") - 1]
    jj, kk = (loc // dims[0]) % dims[1], loc // (dims[0] * dims[1])
    prod = np.nonzero((G.side == 1) & (np.abs(jj - 31.5) < 2) & (np.abs(kk - 15.5) < 2))[0]
    assert prod.shape[0] == 16 * R
    tm, z = np.full(nc, 1.0e3), np.zeros(nc)
    s = pkg.HybridSolver(0)
    s.init(G, Kref, (0.0, 0.0, 0.0), bt, None, inst_cellpos=icp)
    s.solve(tm, z, z, bt, bv, residual_tolerance=1e-12, precision="fp64")
    target = s.getSolution()[2][prod]
    s.computeInnerProducts(K)
    grads = {}
    for prec in ("fp64", "mixed"):
        s.solve(tm, z, z, bt, bv, residual_tolerance=1e-10, precision=prec)
        v = s.getSolution()[2]
        w = np.zeros(nhf)
        w[prod] = 2.0 * (v[prod] - target)
        gK = s.adjoint(w, tol=1e-10)["grad_perm"]
        grads[prec] = K[:, 0] * (gK[:, 0] + gK[:, 4] + gK[:, 8])
    g64, g32 = grads["fp64"], grads["mixed"]
    assert np.max(np.abs(g32 - g64)) <= 1e-6 * np.abs(g64).max()
    # central differences on realization 0 alone (its J does not see the others)
    n0 = int(icp[1])
    g0 = grids_[0]
    h0 = int(G.facepos[n0])
    prod0 = prod[prod < h0]
    s0 = pkg.HybridSolver(0)
    s0.init(g0, K[:n0], (0.0, 0.0, 0.0), bt[:h0], None)

    def J(Kc):
        s0.computeInnerProducts(Kc)
        s0.solve(tm[:n0], z[:n0], z[:n0], bt[:h0], bv[:h0], residual_tolerance=1e-13, precision="fp64")
        return float(np.sum((s0.getSolution()[2][prod0] - target[:prod0.shape[0]]) ** 2))

    rng = np.random.default_rng(20)
    h = 1e-6
    scale = np.abs(g32[:n0]).max()
    for c in rng.choice(n0, 20, replace=False):
        vals = []
        for sgn in (1, -1):
            Kp = K[:n0].copy()
            Kp[c] *= np.exp(sgn * h)
            vals.append(J(Kp))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - g32[c]) <= 1e-4 * scale, (c, fd, g32[c], scale)

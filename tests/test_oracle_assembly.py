"""Pins of the oracle's assembly, boundary rules, solver and recovery
(IncompFlowSolverHybrid.hpp:1324-1420, 1672-1901) against an independent dense
solve of the unreduced mixed-hybrid system, closed forms and invariants.

Unreduced system (the docstring of IncompFlowSolverHybrid.hpp:193-229), per
cell c with half-face fluxes v_c, cell pressure p_c, face pressures pi:
    B_c (v_c - g_c) / lambda_c = p_c 1 - pi_c        (B_c = inverse of B^-1_c)
    sum v_c = src_c
face equations: internal v_1 + v_2 = 0; Neumann v = h; Dirichlet pi = p_D;
periodic pair v_1 + v_2 = 0 and pi_1 - pi_2 = pd_1.
"""
import numpy as np
import pytest

import cases
from cases import DIR, NEU, PER


def kkt_solve(g, orc, perm, tm, rm, src, bt, bv, partner=None, grav=(0, 0, 0)):
    hf_dof, nint, nf, _ = orc.enumerate_dof(g)
    binv, pos, gfl = orc.static_ip(g, perm, grav)
    nc, nhf = g.num_cells, g.num_half_faces
    n = nhf + nc + nf
    M = np.zeros((n, n))
    rhs = np.zeros(n)
    row = 0
    for c in range(nc):
        f0, f1 = g.facepos[c], g.facepos[c + 1]
        B = np.linalg.inv(orc.cell_matrix(binv, pos, g, c) * tm[c])
        gc = gfl[f0:f1] * rm[c]
        for i in range(f1 - f0):
            M[row, f0:f1] = B[i]
            rhs[row] = B[i] @ gc
            M[row, nhf + c] = -1.0
            M[row, nhf + nc + hf_dof[f0 + i]] = 1.0
            row += 1
    for c in range(nc):
        M[row, g.facepos[c]:g.facepos[c + 1]] = 1.0
        rhs[row] = src[c]
        row += 1
    hfs = [[] for _ in range(nf)]
    for h in range(nhf):
        hfs[hf_dof[h]].append(h)
    for f in range(nf):
        r = row + f
        if len(hfs[f]) == 2:
            M[r, hfs[f]] = 1.0
            continue
        h = hfs[f][0]
        if bt[h] == DIR:
            M[r, nhf + nc + f] = 1.0
            rhs[r] = bv[h]
        elif bt[h] == NEU:
            M[r, h] = 1.0
            rhs[r] = bv[h]
        else:
            p = partner[h]
            fp = hf_dof[p]
            if f < fp:
                M[r, [h, p]] = 1.0
            else:
                M[r, nhf + nc + f] = 1.0
                M[r, nhf + nc + fp] = -1.0
                rhs[r] = bv[h]
    x = np.linalg.lstsq(M, rhs, rcond=None)[0]
    return x[:nhf], x[nhf:nhf + nc], x[nhf + nc:]


def _compare(ref, v, p, pi, gauge):
    assert cases.rel_err(ref.outflux, v) < 1e-11
    if gauge:
        p = p - p[0] + ref.cell_pressure[0]
        pi = pi - pi[0] + ref.face_pressure[0]
    assert cases.rel_err(ref.cell_pressure, p) < 1e-11
    assert cases.rel_err(ref.face_pressure, pi) < 1e-11


GRIDS = [(1, 1, 1), (2, 1, 1), (2, 2, 1), (2, 2, 2), (3, 2, 2)]


@pytest.mark.parametrize("dims", GRIDS)
def test_dirichlet_neumann_sources_gravity(pkg, orc, dims):
    rng = np.random.default_rng(sum(dims))
    g = pkg.grids.cartesian(*dims, 0.8, 1.2, 0.5)
    nc = g.num_cells
    A = rng.standard_normal((nc, 3, 3))
    perm = (np.einsum("nij,nkj->nik", A, A) + np.eye(3)).reshape(nc, 9)
    tm = rng.uniform(0.5, 2.0, nc)
    rm = tm * rng.uniform(0.5, 1.5, nc)
    src = rng.standard_normal(nc)
    bt = np.zeros(g.num_half_faces, np.int32)
    bv = np.zeros(g.num_half_faces)
    bt[g.side == 0] = DIR; bv[g.side == 0] = 2.0
    bt[g.side == 5] = DIR; bv[g.side == 5] = -1.0
    m = g.side == 2
    bv[m] = rng.standard_normal(m.sum())          # Neumann outflux on y-
    grav = (0.3, -0.2, 1.0)
    ref = orc.solve(g, perm, tm, rm, src, bt, bv, grav=grav, tol=1e-15)
    v, p, pi = kkt_solve(g, orc, perm, tm, rm, src, bt, bv, grav=grav)
    _compare(ref, v, p, pi, gauge=False)


@pytest.mark.parametrize("dims,dirs", [((2, 2, 2), (0, 1, 2)), ((3, 2, 2), (0,)), ((3, 3, 2), (0, 1))])
def test_periodic_against_kkt(pkg, orc, dims, dirs):
    rng = np.random.default_rng(11)
    g = pkg.grids.cartesian(*dims)
    nc = g.num_cells
    perm = pkg.fields.diag_perm(rng.uniform(0.5, 2.0, nc), rng.uniform(0.5, 2.0, nc),
                                rng.uniform(0.5, 2.0, nc))
    bt, bv, partner = cases.periodic_bc(g, dirs, drop_dir=dirs[0], pdrop=1.5)
    tm, rm = np.ones(nc), np.zeros(nc)
    src = np.zeros(nc)
    ref = orc.solve(g, perm, tm, rm, src, bt, bv, partner=partner, tol=1e-15)
    v, p, pi = kkt_solve(g, orc, perm, tm, rm, src, bt, bv, partner=partner)
    _compare(ref, v, p, pi, gauge=True)


def test_pure_neumann_gauge(pkg, orc):
    g = pkg.grids.cartesian(3, 2, 2)
    nc = g.num_cells
    perm = pkg.fields.perm_c1(nc, 4) * 1e13       # O(1) entries keep the dense KKT well scaled
    src = np.zeros(nc); src[0], src[-1] = 1.0, -1.0
    bt = np.zeros(g.num_half_faces, np.int32); bv = np.zeros(g.num_half_faces)
    ref = orc.solve(g, perm, np.ones(nc), np.zeros(nc), src, bt, bv, tol=1e-15)
    assert ref.system.do_regularization
    v, p, pi = kkt_solve(g, orc, perm, np.ones(nc), np.zeros(nc), src, bt, bv)
    _compare(ref, v, p, pi, gauge=True)


def test_matrix_symmetric_and_singular_without_dirichlet(pkg, orc):
    g = pkg.grids.cartesian(4, 3, 2)
    nc = g.num_cells
    perm = pkg.fields.perm_c1(nc, 2)
    hf, nint, nf, _ = orc.enumerate_dof(g)
    bt = np.zeros(g.num_half_faces, np.int32); bv = np.zeros(g.num_half_faces)
    pp, _ = orc.periodic_dof(g, hf, nf, bt, np.full(g.num_half_faces, -1, np.int32))
    binv, pos, gfl = orc.static_ip(g, perm)
    s = orc.assemble(g, hf, nf, pp, binv, gfl, np.ones(nc), np.zeros(nc), np.zeros(nc), bt, bv)
    import scipy.sparse as sp
    A = sp.csr_matrix((s.val, s.col, s.row_ptr), shape=(nf, nf)).toarray()
    assert np.allclose(A, A.T, rtol=0, atol=1e-14 * np.abs(A).max())
    # local Schur complements have zero row sums: constants are in the kernel
    assert np.max(np.abs(A @ np.ones(nf))) < 1e-13 * np.abs(A).max()
    orc.regularize(s)
    A2 = sp.csr_matrix((s.val, s.col, s.row_ptr), shape=(nf, nf)).toarray()
    assert np.min(np.linalg.eigvalsh(A2)) > 0      # S[0][0] *= 2 makes it SPD
    assert A2[0, 0] == 2 * A[0, 0]


def test_linear_pressure_and_layered_closed_forms(pkg, orc):
    nx, ny, nz = 6, 3, 2
    g = pkg.grids.cartesian(nx, ny, nz, 0.5, 1.0, 2.0)
    nc = g.num_cells
    bt, bv = cases.xdrop_bc(g, 3.0, 1.0)
    # uniform K: exact linear pressure, uniform flux
    ref = orc.solve(g, pkg.fields.diag_perm(np.full(nc, 2.0)), np.ones(nc), np.zeros(nc),
                    np.zeros(nc), bt, bv, tol=1e-15)
    x = g.cell_centroid[:, 0]
    L = nx * 0.5
    assert np.allclose(ref.cell_pressure, 3.0 - 2.0 * x / L, rtol=0, atol=1e-12)
    q = 2.0 * 2.0 / L * (1.0 * 2.0)       # k dp/L * face area, per cell row
    xm = g.side == 0
    assert np.allclose(ref.outflux[xm], -q, rtol=1e-12)
    # K layered along x: total flux dp / sum(dx / (k A_total)) (harmonic mean)
    kx = np.exp(np.linspace(-1, 2, nx))
    kc = kx[np.arange(nc) % nx]
    ref = orc.solve(g, pkg.fields.diag_perm(kc), np.ones(nc), np.zeros(nc), np.zeros(nc), bt, bv,
                    tol=1e-15)
    total = -np.sum(ref.outflux[xm])
    expect = 2.0 / np.sum(0.5 / (kx * ny * 1.0 * nz * 2.0))
    assert abs(total - expect) < 1e-11 * expect


def test_mass_balance_independent_of_solver_accuracy(pkg, orc):
    g, perm, tm, rm, src, bt, bv = cases.case_c1()
    src = np.random.default_rng(1).standard_normal(g.num_cells) * 1e-3
    ref = orc.solve(g, perm, tm, rm, src, bt, bv, tol=1e-30, maxit=5, check=False)
    bal = np.bincount(g.hf_cell(), weights=ref.outflux, minlength=g.num_cells) - src
    assert np.max(np.abs(bal)) < 1e-12 * np.max(np.abs(ref.outflux))


def test_cg_matches_dense_and_preconditioners_agree(pkg, orc):
    g, perm, tm, rm, src, bt, bv = cases.case_c1()
    sols = [orc.solve(g, perm, tm, rm, src, bt, bv, tol=1e-14, precond=pc) for pc in (0, 1, 2)]
    import scipy.sparse as sp
    s = sols[0].system
    A = sp.csr_matrix((s.val, s.col, s.row_ptr), shape=(s.nf, s.nf))
    import scipy.sparse.linalg as spl
    x = spl.spsolve(A.tocsc(), s.rhs)
    for r in sols:
        assert cases.rel_err(r.face_pressure, x) < 1e-10
    assert sols[2].iterations < sols[1].iterations      # ILU0 beats Jacobi
    for r in sols:
        assert r.reduction < 1e-14


def test_threaded_cg_agrees_and_one_thread_is_unchanged(pkg, orc):
    """The OpenMP CG (multi-core CPU figure) solves the same system; with one
    thread it is bit-identical to the single-threaded summation order."""
    g, perm, tm, rm, src, bt, bv = cases.case_c1()
    base = [orc.solve(g, perm, tm, rm, src, bt, bv, tol=1e-14, precond=pc) for pc in (0, 1)]
    orc.set_threads(4)
    try:
        thr = [orc.solve(g, perm, tm, rm, src, bt, bv, tol=1e-14, precond=pc) for pc in (0, 1)]
    finally:
        orc.set_threads(1)
    again = [orc.solve(g, perm, tm, rm, src, bt, bv, tol=1e-14, precond=pc) for pc in (0, 1)]
    for b, t, a in zip(base, thr, again):
        assert cases.rel_err(t.face_pressure, b.face_pressure) < 1e-11
        # at 1e-14 the count depends on round-off, and the threads' partial
        # sums are combined in no fixed order
        assert abs(t.iterations - b.iterations) <= max(2, b.iterations // 20)
        np.testing.assert_array_equal(a.face_pressure, b.face_pressure)
        assert a.iterations == b.iterations


def test_enumeration_rules(pkg, orc):
    g = pkg.grids.cartesian(3, 2, 2)
    hf, nint, nf, mx = orc.enumerate_dof(g)
    # 3x2x2: internal 2*2*2 + 3*1*2 + 3*2*1 = 20, boundary 2*(4 + 6 + 6) = 32
    assert (nint, nf, mx) == (20, 52, 6)
    nb = g.nb
    # internal faces numbered first, in discovery order (cell order, local order)
    seen = []
    for c in range(g.num_cells):
        for h in range(g.facepos[c], g.facepos[c + 1]):
            if nb[h] > c:
                seen.append(hf[h])
    assert seen == list(range(nint))
    bnd = [hf[h] for h in range(g.num_half_faces) if nb[h] < 0]
    assert bnd == list(range(nint, nf))
    tw = g.twin()
    inside = nb >= 0
    assert np.array_equal(hf[inside], hf[tw[inside]])


def test_post_process_antisymmetric(pkg, orc):
    g = pkg.grids.cartesian(4, 4, 3)
    nc = g.num_cells
    perm = pkg.fields.perm_c1(nc, 8)
    bt, bv, partner = cases.periodic_bc(g, (0, 1), drop_dir=0, pdrop=1.0)
    ref = orc.solve(g, perm, np.ones(nc), np.zeros(nc), np.zeros(nc), bt, bv, partner=partner,
                    tol=1e-8)
    v, mm = orc.post_process(g, ref.hf_dof, ref.nf, ref.ppartner, True, ref.outflux)
    tw = g.twin().copy()
    tw[partner >= 0] = partner[partner >= 0]
    has = tw >= 0
    assert np.max(np.abs(v[has] + v[tw[has]])) == 0.0
    expect = np.where(has, 0.5 * (ref.outflux - ref.outflux[np.where(has, tw, 0)]), ref.outflux)
    assert np.allclose(v, expect, rtol=0, atol=1e-15 * np.abs(v).max())
    assert np.isclose(mm, np.max(np.abs(ref.outflux - expect)), rtol=1e-12)

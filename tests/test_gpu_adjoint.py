"""Adjoint of the hybrid solve (differentiation of the recovered fluxes).

The reference has no gradient; the gradients are pinned by central finite
differences of the oracle's FP64 forward solve (itself pinned by closed forms,
dense KKT solves and the reference's known-answer numbers), and by central
differences of the GPU's own FP64 forward solve at larger sizes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _objective(v, c, d):
    return float(np.sum(c * v * v) + np.sum(d * v))


def _small_case(pkg, seed=4):
    g = pkg.grids.cartesian(4, 3, 3, 0.7, 1.1, 0.9)
    rng = np.random.default_rng(seed)
    nc, nhf = g.num_cells, g.num_half_faces
    M = 0.3 * rng.standard_normal((nc, 3, 3))
    K = np.einsum("nij,nkj->nik", M, M) + np.exp(rng.normal(0.0, 0.5, (nc, 1, 1))) * np.eye(3)
    tm = rng.uniform(0.5, 2.0, nc)
    rm = rng.uniform(0.2, 1.5, nc)
    src = rng.normal(0.0, 0.1, nc)
    grav = (0.3, -0.2, 1.5)
    bt = np.zeros(nhf, np.int32)
    bv = np.zeros(nhf)
    bt[g.side == 0] = 1
    bv[g.side == 0] = rng.uniform(0.5, 1.5, int(np.sum(g.side == 0)))
    bv[g.side == 3] = 0.05                      # Neumann outflux on y+
    c = rng.uniform(0.5, 1.5, nhf)
    d = rng.normal(0.0, 1.0, nhf)
    return g, K.reshape(nc, 9), tm, rm, src, grav, bt, bv, c, d


def test_adjoint_matches_oracle_finite_differences(pkg, orc):
    g, K, tm, rm, src, grav, bt, bv, c, d = _small_case(pkg)
    nc = g.num_cells
    s = pkg.HybridSolver(0)
    s.init(g, K, grav, bt, None)
    s.solve(tm, rm, src, bt, bv, residual_tolerance=1e-13, precision="fp64")
    _, _, v = s.getSolution()
    gr = s.adjoint(2.0 * c * v + d, tol=1e-13)

    def J(K_=K, tm_=tm, rm_=rm, src_=src):
        r = orc.solve(g, K_, tm_, rm_, src_, bt, bv, grav=grav, tol=1e-15)
        return _objective(r.outflux, c, d)

    assert abs(J() - _objective(v, c, d)) <= 1e-10 * abs(J())
    rel = 1e-6
    # permeability: diagonal entries and symmetric off-diagonal pairs
    fdK = np.zeros((nc, 9))
    for cell in range(nc):
        for (l, m) in ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)):
            h = rel * np.abs(K[cell]).max()
            Kp, Km = K.copy(), K.copy()
            for (a, b) in {(l, m), (m, l)}:
                Kp[cell, 3 * a + b] += h
                Km[cell, 3 * a + b] -= h
            dj = (J(K_=Kp) - J(K_=Km)) / (2 * h)
            if l == m:
                fdK[cell, 3 * l + m] = dj
            else:  # both entries moved: dJ = h (G_lm + G_ml) = 2 h G_lm
                fdK[cell, 3 * l + m] = fdK[cell, 3 * m + l] = dj / 2
    gK = gr["grad_perm"]
    assert np.allclose(gK, gK.reshape(nc, 3, 3).transpose(0, 2, 1).reshape(nc, 9), rtol=0, atol=1e-13 * np.abs(gK).max())
    assert np.max(np.abs(gK - fdK)) <= 2e-6 * np.abs(fdK).max(), np.max(np.abs(gK - fdK)) / np.abs(fdK).max()

    def fd_vec(name, base):
        out = np.zeros(nc)
        for cell in range(nc):
            h = rel * max(abs(base[cell]), 1.0)
            p, m = base.copy(), base.copy()
            p[cell] += h
            m[cell] -= h
            out[cell] = (J(**{name: p}) - J(**{name: m})) / (2 * h)
        return out

    for key, name, base in (("grad_totmob", "tm_", tm), ("grad_rhomob", "rm_", rm), ("grad_src", "src_", src)):
        fd = fd_vec(name, base)
        assert np.max(np.abs(gr[key] - fd)) <= 2e-6 * np.abs(fd).max(), (key, np.max(np.abs(gr[key] - fd)))


def test_adjoint_corner_point_matches_oracle_finite_differences(pkg, orc):
    """The same gradients on a faulted corner-point grid (cells with up to 9
    faces, non-orthogonal faces), against central differences of the oracle."""
    g = pkg.grids.corner_point(4, 3, 3, 1.0, 1.2, 0.6, seed=8, fault_frac=0.4)
    assert g.ncf().max() > 6
    nc, nhf = g.num_cells, g.num_half_faces
    rng = np.random.default_rng(4)
    M = 0.3 * rng.standard_normal((nc, 3, 3))
    K = (np.einsum("nij,nkj->nik", M, M) + np.exp(rng.normal(0.0, 0.5, (nc, 1, 1))) * np.eye(3)).reshape(nc, 9)
    tm, rm, src = rng.uniform(0.5, 2.0, nc), rng.uniform(0.2, 1.5, nc), rng.normal(0.0, 0.1, nc)
    grav = (0.3, -0.2, 1.5)
    bt = np.zeros(nhf, np.int32)
    bv = np.zeros(nhf)
    bt[g.side == 0] = 1
    bv[g.side == 0] = rng.uniform(0.5, 1.5, int(np.sum(g.side == 0)))
    bv[g.side == 3] = 0.05
    c, d = rng.uniform(0.5, 1.5, nhf), rng.normal(0.0, 1.0, nhf)
    s = pkg.HybridSolver(0)
    s.init(g, K, grav, bt, None)
    s.solve(tm, rm, src, bt, bv, residual_tolerance=1e-13, precision="fp64")
    gr = s.adjoint(2.0 * c * s.getSolution()[2] + d, tol=1e-13)

    def J(K_=K, tm_=tm, rm_=rm, src_=src):
        return _objective(orc.solve(g, K_, tm_, rm_, src_, bt, bv, grav=grav, tol=1e-15).outflux, c, d)

    rel = 1e-6
    fdK = np.zeros((nc, 9))
    for cell in range(nc):
        for (l, m) in ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)):
            h = rel * np.abs(K[cell]).max()
            Kp, Km = K.copy(), K.copy()
            for (a, b) in {(l, m), (m, l)}:
                Kp[cell, 3 * a + b] += h
                Km[cell, 3 * a + b] -= h
            dj = (J(K_=Kp) - J(K_=Km)) / (2 * h)
            fdK[cell, 3 * l + m] = fdK[cell, 3 * m + l] = dj if l == m else dj / 2
    assert np.max(np.abs(gr["grad_perm"] - fdK)) <= 2e-6 * np.abs(fdK).max()
    for key, name, base in (("grad_totmob", "tm_", tm), ("grad_rhomob", "rm_", rm), ("grad_src", "src_", src)):
        fd = np.zeros(nc)
        for cell in range(nc):
            h = rel * max(abs(base[cell]), 1.0)
            p, m = base.copy(), base.copy()
            p[cell] += h
            m[cell] -= h
            fd[cell] = (J(**{name: p}) - J(**{name: m})) / (2 * h)
        assert np.max(np.abs(gr[key] - fd)) <= 2e-6 * np.abs(fd).max(), key


def _c5_like(pkg, ninst=2, dims=(12, 10, 6), seed=11):
    grids, perms = [], []
    for t in range(ninst):
        g = pkg.grids.cartesian(*dims)
        lnk = pkg.fields.gaussian_filtered(dims, corr_len=3.0, seed=seed + t)
        perms.append(np.exp(lnk)[:, None] * np.eye(3).reshape(1, 9))
        grids.append(g)
    G, icp = pkg.grids.concat(grids)
    K = np.concatenate(perms)
    nhf = G.num_half_faces
    bt = np.zeros(nhf, np.int32)
    bv = np.zeros(nhf)
    bt[G.side == 0] = 1
    bv[G.side == 0] = 1.0
    bt[G.side == 1] = 1
    # 16 producer faces per instance on x+, targets from a perturbed field
    rng = np.random.default_rng(seed)
    prod = []
    for t in range(ninst):
        cand = np.nonzero(G.side == 1)[0]
        cand = cand[(G.hf_cell()[cand] >= icp[t]) & (G.hf_cell()[cand] < icp[t + 1])]
        prod.append(rng.choice(cand, 16, replace=False))
    prod = np.concatenate(prod)
    return G, icp, K, bt, bv, prod


def test_adjoint_lognormal_batch_vs_gpu_finite_differences(pkg):
    G, icp, K, bt, bv, prod = _c5_like(pkg)
    nc = G.num_cells
    tm, rm, src = np.ones(nc), np.zeros(nc), np.zeros(nc)
    s = pkg.HybridSolver(0)
    s.init(G, K, (0, 0, 0), bt, None, inst_cellpos=icp)
    s.solve(tm, rm, src, bt, bv, residual_tolerance=1e-13, precision="fp64")
    v0 = s.getSolution()[2]
    target = 0.9 * v0[prod]

    def J(v):
        return float(np.sum((v[prod] - target) ** 2))

    w = np.zeros(G.num_half_faces)
    w[prod] = 2.0 * (v0[prod] - target)
    gK = s.adjoint(w, tol=1e-13)["grad_perm"]
    assert s.adjoint_report.precond == 1
    # isotropic K = k I: dJ/dln k = sum_l k dJ/dK_ll
    glnk = K[:, 0] * (gK[:, 0] + gK[:, 4] + gK[:, 8])
    rng = np.random.default_rng(3)
    cells = np.concatenate([rng.choice(np.arange(icp[t], icp[t + 1]), 6, replace=False)
                            for t in range(len(icp) - 1)])
    h = 1e-6
    for cell in cells:
        vals = []
        for sgn in (1, -1):
            Kp = K.copy()
            Kp[cell] *= np.exp(sgn * h)
            s.computeInnerProducts(Kp)
            s.solve(tm, rm, src, bt, bv, residual_tolerance=1e-13, precision="fp64")
            vals.append(J(s.getSolution()[2]))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - glnk[cell]) <= 1e-5 * np.abs(glnk).max(), (cell, fd, glnk[cell])


def test_mixed_precision_adjoint_matches_fp64(pkg):
    G, icp, K, bt, bv, prod = _c5_like(pkg, seed=21)
    nc = G.num_cells
    tm, rm, src = np.ones(nc), np.zeros(nc), np.zeros(nc)
    out = {}
    for prec in ("fp64", "mixed"):
        s = pkg.HybridSolver(0)
        s.init(G, K, (0, 0, 0), bt, None, inst_cellpos=icp)
        s.solve(tm, rm, src, bt, bv, residual_tolerance=1e-10, precision=prec)
        v = s.getSolution()[2]
        w = np.zeros(G.num_half_faces)
        w[prod] = 2.0 * v[prod]
        out[prec] = s.adjoint(w, tol=1e-10)["grad_perm"]
        # the forward solution is left as it was
        assert np.array_equal(s.getSolution()[2], v)
    a, b = out["mixed"], out["fp64"]
    assert np.max(np.abs(a - b)) <= 1e-6 * np.abs(b).max()


def test_adjoint_rejects_periodic_and_missing_solve(pkg):
    g = pkg.grids.cartesian(4, 4, 4)
    pp = pkg.grids.periodic_partners_cartesian(g, (0,))
    bt = np.where(pp >= 0, 2, 0).astype(np.int32)
    s = pkg.HybridSolver(0)
    s.init(g, np.tile(np.eye(3).reshape(1, 9), (g.num_cells, 1)), (0, 0, 0), bt, pp)
    with pytest.raises(pkg.MhsError, match="no solution"):
        s.adjoint(np.zeros(g.num_half_faces))
    bv = np.where((pp >= 0) & (g.side == 0), 1.0, 0.0)
    s.solve(np.ones(g.num_cells), np.zeros(g.num_cells), np.zeros(g.num_cells), bt, bv,
            residual_tolerance=1e-10)
    with pytest.raises(pkg.MhsError, match="periodic"):
        s.adjoint(np.zeros(g.num_half_faces))


def test_torch_autograd_function(pkg):
    import torch
    from opm_porsol_amd.autograd import hybrid_flux
    dev = torch.device("cuda:0")
    G, icp, K, bt, bv, prod = _c5_like(pkg, ninst=1, dims=(8, 6, 4), seed=5)
    nc = G.num_cells
    s = pkg.HybridSolver(0, stream=torch.cuda.current_stream(dev).cuda_stream)
    s.init(G, K, (0, 0, 0), bt, None)
    lnk = torch.tensor(np.log(K[:, 0]), device=dev, requires_grad=True)
    eye = torch.eye(3, dtype=torch.float64, device=dev).reshape(1, 9)
    tm = torch.ones(nc, dtype=torch.float64, device=dev, requires_grad=True)
    rm = torch.zeros(nc, dtype=torch.float64, device=dev)
    src = torch.zeros(nc, dtype=torch.float64, device=dev)
    bvt = torch.tensor(bv, device=dev)
    pidx = torch.tensor(prod, device=dev)

    def loss_fn(lnk_, tm_):
        v = hybrid_flux(s, torch.exp(lnk_)[:, None] * eye, tm_, rm, src, bt, bvt,
                        tol=1e-13, precision="fp64")
        return torch.sum(v[pidx] ** 2)

    loss = loss_fn(lnk, tm)
    loss.backward()
    g_lnk = lnk.grad.detach().cpu().numpy()
    g_tm = tm.grad.detach().cpu().numpy()
    # central differences through the same function
    h = 1e-6
    rng = np.random.default_rng(0)
    with torch.no_grad():
        for cell in rng.choice(nc, 5, replace=False):
            e = torch.zeros(nc, dtype=torch.float64, device=dev)
            e[cell] = h
            fd = (loss_fn(lnk + e, tm).item() - loss_fn(lnk - e, tm).item()) / (2 * h)
            assert abs(fd - g_lnk[cell]) <= 1e-5 * np.abs(g_lnk).max(), (cell, fd, g_lnk[cell])
            fd = (loss_fn(lnk, tm + e).item() - loss_fn(lnk, tm - e).item()) / (2 * h)
            assert abs(fd - g_tm[cell]) <= 1e-5 * np.abs(g_tm).max(), (cell, fd, g_tm[cell])
    # a second solve invalidates the first graph's backward
    l1 = loss_fn(lnk, tm)
    loss_fn(lnk.detach(), tm.detach())
    with pytest.raises(RuntimeError, match="solved again"):
        l1.backward()

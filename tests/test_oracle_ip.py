"""Pins of the oracle's static inner products (MimeticIPEvaluator.hpp:201-270).

Each check fails for a plausible mistake: a dropped stabilisation term, t off by
a factor, a transposed K, a wrong area scaling or a missing 1/|c|.
"""
import numpy as np
import pytest

import cases


def _box(pkg, dx, dy, dz):
    return pkg.grids.cartesian(1, 1, 1, dx, dy, dz)


def _parallelepiped(pkg, A, origin=(0.0, 0.0, 0.0)):
    """One hexahedral cell spanned by the columns of A (planar faces), in the
    local face order x-, x+, y-, y+, z-, z+ (faces opposite to e_d)."""
    g = pkg.grids.cartesian(1, 1, 1)
    A = np.asarray(A, float)
    o = np.asarray(origin, float)
    cen = o + 0.5 * A.sum(axis=1)
    fc, nrm, area = [], [], []
    for d in range(3):
        e1, e2 = A[:, (d + 1) % 3], A[:, (d + 2) % 3]
        cr = np.cross(e1, e2)
        a = np.linalg.norm(cr)
        n = cr / a
        if np.dot(n, A[:, d]) < 0:
            n = -n
        for side in (0, 1):
            c = cen + (side - 0.5) * A[:, d]
            fc.append(c)
            nrm.append(n if side == 1 else -n)
            area.append(a)
    g.cell_centroid = cen[None, :]
    g.cell_volume = np.array([abs(np.linalg.det(A))])
    g.face_centroid = np.array(fc)
    g.face_normal = np.array(nrm)
    g.face_area = np.array(area)
    return g


def test_box_closed_form(pkg, orc):
    # B^-1 block-diagonal over the +- face pairs of each axis d:
    # (A_d^2/V) [[k_d + trK, -k_d + trK], [-k_d + trK, k_d + trK]]
    dx, dy, dz = 0.7, 1.9, 0.3
    k = np.array([3.0, 0.5, 0.02])
    g = _box(pkg, dx, dy, dz)
    perm = np.diag(k).reshape(1, 9)
    binv, pos, gfl = orc.static_ip(g, perm)
    B = orc.cell_matrix(binv, pos, g, 0)
    V = dx * dy * dz
    areas = [dy * dz, dx * dz, dx * dy]
    expect = np.zeros((6, 6))
    for d in range(3):
        blk = np.array([[k[d] + k.sum(), -k[d] + k.sum()], [-k[d] + k.sum(), k[d] + k.sum()]])
        expect[2 * d:2 * d + 2, 2 * d:2 * d + 2] = areas[d] ** 2 / V * blk
    assert np.max(np.abs(B - expect)) <= 1e-14 * np.max(np.abs(expect))


def test_unit_cube_isotropic(pkg, orc):
    g = _box(pkg, 1, 1, 1)
    binv, pos, _ = orc.static_ip(g, (5.0 * np.eye(3)).reshape(1, 9))
    B = orc.cell_matrix(binv, pos, g, 0)
    blk = np.array([[4.0, 2.0], [2.0, 4.0]]) * 5.0
    assert np.allclose(B, np.kron(np.eye(3), blk), rtol=0, atol=1e-13)


def _numpy_binv(g, K, c=0):
    """Independent evaluation with numpy's QR (a different factorisation)."""
    f0, f1 = g.facepos[c], g.facepos[c + 1]
    a = g.face_area[f0:f1]
    N = g.face_normal[f0:f1] * a[:, None]
    C = (g.face_centroid[f0:f1] - g.cell_centroid[c]) * a[:, None]
    Q, _ = np.linalg.qr(C)
    n = f1 - f0
    P = np.eye(n) - Q @ Q.T
    t = 2.0 * np.trace(K)
    return (N @ K @ N.T + t * np.diag(a) @ P @ np.diag(a)) / g.cell_volume[c], Q


def test_parallelepiped_against_numpy_qr_and_sign_invariance(pkg, orc):
    rng = np.random.default_rng(3)
    for _ in range(5):
        A = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
        g = _parallelepiped(pkg, A, rng.standard_normal(3))
        M = rng.standard_normal((3, 3))
        K = M @ M.T + 0.3 * np.eye(3)
        binv, pos, _ = orc.static_ip(g, K.reshape(1, 9))
        B = orc.cell_matrix(binv, pos, g, 0)
        ref, Q = _numpy_binv(g, K)
        assert np.max(np.abs(B - ref)) <= 1e-12 * np.max(np.abs(ref))
        # Q -> Q diag(+-1) leaves Q Q' and hence B^-1 unchanged
        Qs = Q * np.array([1.0, -1.0, -1.0])
        assert np.allclose(Qs @ Qs.T, Q @ Q.T, atol=1e-14)


def test_mimetic_consistency_linear_pressure(pkg, orc):
    # For planar-faced cells sum_i a_i n_i (x_i - x_c)' = |c| I, so
    # B^-1 (x_f - x_c) = N K exactly: linear pressures give exact fluxes.
    rng = np.random.default_rng(5)
    A = np.eye(3) + 0.25 * rng.standard_normal((3, 3))
    g = _parallelepiped(pkg, A)
    M = rng.standard_normal((3, 3))
    K = M @ M.T + 0.5 * np.eye(3)
    binv, pos, _ = orc.static_ip(g, K.reshape(1, 9))
    B = orc.cell_matrix(binv, pos, g, 0)
    Ct = g.face_centroid - g.cell_centroid[0]
    N = g.face_normal * g.face_area[:, None]
    assert np.max(np.abs(B @ Ct - N @ K)) <= 1e-12 * np.max(np.abs(N @ K))


def test_gravity_flux(pkg, orc):
    rng = np.random.default_rng(9)
    g = pkg.grids.cartesian(2, 2, 1, 0.5, 2.0, 1.5)
    K = rng.standard_normal((g.num_cells, 3, 3))
    K = np.einsum("nij,nkj->nik", K, K) + np.eye(3)
    grav = np.array([0.1, -0.3, 9.81])
    _, _, gfl = orc.static_ip(g, K.reshape(-1, 9), grav)
    expect = np.einsum("hi,hij,j->h", g.face_normal * g.face_area[:, None],
                       np.repeat(K, 6, axis=0), grav)
    assert np.allclose(gfl, expect, rtol=1e-14, atol=1e-14 * np.abs(expect).max())


def test_attic_fixture_matches(pkg, orc):
    """attic/mimetic/mimetic_ipeval_test.cpp prints B^-1 for K = diag(10,1,0.1)
    on 1^3, 2^3, 4^3 unit-cube grids; the committed fixture holds the oracle's
    values, which must equal the closed form."""
    data = np.load(cases.golden("ip_attic_diag10_1_01.npz"))
    for n in (1, 2, 4):
        g = pkg.grids.cartesian(n, n, n, 1.0 / n, 1.0 / n, 1.0 / n)
        perm = np.tile(np.diag([10.0, 1.0, 0.1]).reshape(1, 9), (g.num_cells, 1))
        binv, pos, _ = orc.static_ip(g, perm)
        assert np.array_equal(binv, data[f"binv_{n}"])
        k = np.array([10.0, 1.0, 0.1]); h = 1.0 / n
        blk = [h ** 4 / h ** 3 * np.array([[kd + k.sum(), -kd + k.sum()], [-kd + k.sum(), kd + k.sum()]])
               for kd in k]
        expect = np.zeros((6, 6))
        for d in range(3):
            expect[2 * d:2 * d + 2, 2 * d:2 * d + 2] = blk[d]
        assert np.allclose(orc.cell_matrix(binv, pos, g, 0), expect, rtol=1e-14, atol=1e-14)

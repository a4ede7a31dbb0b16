"""The oracle against the committed fixtures and the reference's recorded
known-answer numbers (examples/timing_results.txt)."""
import json

import numpy as np
import pytest

import cases


def test_c1_fixture(pkg, orc):
    d = np.load(cases.golden("c1_20x20x10.npz"))
    g, perm, tm, rm, src, bt, bv = cases.case_c1()
    assert str(d["perm_checksum"]) == pkg.fields.checksum(perm)
    r = orc.solve(g, perm, tm, rm, src, bt, bv, tol=1e-14)
    assert np.array_equal(r.hf_dof, d["hf_dof"])
    for k in ("face_pressure", "cell_pressure", "outflux"):
        assert cases.rel_err(getattr(r, k), d[k]) < 1e-12, k


def test_periodic_fixture(pkg, orc):
    d = np.load(cases.golden("periodic_4x4x4.npz"))
    g = pkg.grids.cartesian(4, 4, 4)
    perm = pkg.fields.perm_c1(g.num_cells, seed=7)
    bt, bv, partner = cases.periodic_bc(g, (0, 1, 2), drop_dir=0, pdrop=1.0)
    r = orc.solve(g, perm, *cases.unit_fluid(g.num_cells), np.zeros(g.num_cells), bt, bv,
                  partner=partner, tol=1e-14)
    for k in ("face_pressure", "cell_pressure", "outflux"):
        assert cases.rel_err(getattr(r, k), d[k]) < 1e-12, k


def _known_answer_norms(pkg, orc, n):
    g = pkg.grids.cartesian(n, n, n, 1.0 / n, 1.0 / n, 1.0 / n)
    u = lambda x: np.sin(2 * np.pi * x[:, 0]) * np.cos(2 * np.pi * x[:, 1]) * x[:, 2]
    lu = lambda x: -8 * np.pi ** 2 * np.sin(2 * np.pi * x[:, 0]) * np.cos(2 * np.pi * x[:, 1]) * x[:, 2]
    nc = g.num_cells
    perm = np.tile(np.eye(3).reshape(1, 9), (nc, 1))
    bt = np.where(g.nb < 0, 1, 0).astype(np.int32)
    bv = np.where(g.nb < 0, u(g.face_centroid), 0.0)
    src = -lu(g.cell_centroid) * g.cell_volume
    r = orc.solve(g, perm, np.ones(nc), np.zeros(nc), src, bt, bv, tol=1e-12)
    d = u(g.cell_centroid) - r.cell_pressure
    v = g.cell_volume
    return np.array([np.sum(np.abs(d * v)) / v.sum(), np.sqrt(np.sum(d * d * v)) / v.sum(),
                     np.max(np.abs(d))])


@pytest.mark.slow
def test_known_answer_50(pkg, orc):
    ref = json.load(open(cases.golden("known_answer.json")))["50"]
    got = _known_answer_norms(pkg, orc, 50)
    exp = np.array([ref["L1"], ref["L2"], ref["Linf"]])
    assert np.all(np.abs(got - exp) / exp < np.array([5e-5, 5e-5, 1e-4])), (got, exp)


def test_c2_fixture_is_the_headline_case_and_conserves_mass(pkg):
    """The 128x128x64 fixture (the oracle's ILU0-CG to 1e-14) belongs to the
    headline inputs (permeability checksum), was solved to a true residual
    below 1e-13, and carries the same flow through every x-plane (no sources:
    conservation fixes the rate from x- to x+)."""
    import cases
    fx = np.load(cases.golden("c2_128x128x64.npz"))
    g, perm, bt, bv = cases.c2_problem()
    assert str(fx["perm_checksum"]) == pkg.fields.checksum(perm)
    assert float(fx["true_residual"]) < 1e-13
    rates = fx["xplane_rates"]
    assert rates.shape == (129,) and rates.min() > 0
    assert np.max(np.abs(rates - rates.mean())) <= 1e-10 * rates.mean()
    assert int(fx["ilu0_iterations_1e10"]) < int(fx["ilu0_iterations_1e14"])
    for name in ("face_pressure", "cell_pressure", "outflux"):
        assert fx[name + "_sample"].shape == (cases.C2_SAMPLES,)
        assert 0 < float(fx[name + "_ilu0_1e10_err"]) < 1e-7
    # pressures lie between the two Dirichlet values (maximum principle of the
    # face and cell pressures on this grid of unit cubes)
    for name in ("face_pressure", "cell_pressure"):
        a = fx[name + "_sample"]
        assert a.min() >= -1e-9 and a.max() <= 1.0e5 * (1 + 1e-12)

"""ECLIPSE input (opm-porsol_amd/deck.py): the GRDECL reader, corner-point
geometry from COORD/ZCORN, and the reference's permeability / porosity
assembly rules (ReservoirPropertyCommon_impl.hpp:52-244, 592-611, 682-736)."""
import numpy as np
import pytest

MD = 1e-3 * 9.869233e-13


def test_reader_repeats_comments_sections(pkg):
    d = pkg.deck.read_grdecl("""
-- a comment
GRID
SPECGRID
 2 1 1 1 F /
PORO
 2*0.25   -- two cells
/
PERMX 100 3.5E2 /
""")
    assert list(d["SPECGRID"]) == [2, 1, 1]
    np.testing.assert_array_equal(d["PORO"], [0.25, 0.25])
    np.testing.assert_array_equal(d["PERMX"], [100.0, 350.0])
    assert "GRID" in d
    with pytest.raises(ValueError, match="not terminated"):
        pkg.deck.read_grdecl("PORO 0.1 0.2")
    with pytest.raises(ValueError, match="METRIC"):
        pkg.deck.read_grdecl("FIELD\nPORO 0.1 /")


def test_flat_deck_is_the_cartesian_grid(pkg):
    """A deck with flat layers on a regular lattice gives the Cartesian box:
    same cells, faces, neighbours, areas, normals, centroids and volumes."""
    nx, ny, nz, dx, dy, dz = 4, 3, 2, 2.0, 1.5, 0.5
    zc = np.broadcast_to(np.arange(nz + 1) * dz, (ny, nx, 4, nz + 1)).copy()
    g = pkg.deck.grid_from_deck(pkg.deck.read_grdecl(pkg.deck.write_grdecl(nx, ny, nz, dx, dy, zc)))
    c = pkg.grids.cartesian(nx, ny, nz, dx, dy, dz)
    np.testing.assert_array_equal(g.facepos, c.facepos)
    np.testing.assert_array_equal(g.nb, c.nb)
    for f in ("cell_volume", "cell_centroid", "face_area", "face_centroid", "face_normal"):
        np.testing.assert_allclose(getattr(g, f), getattr(c, f), rtol=1e-13, atol=1e-13, err_msg=f)


def test_faulted_deck_round_trip_and_closed_cells(pkg):
    """The faulted generator written as COORD/ZCORN (with an origin shift)
    reads back to the same geometry shifted; every cell's surface closes
    (sum of area * normal = 0) and the volumes fill the columns."""
    nx, ny, nz, dx, dy = 6, 5, 4, 3.0, 2.0
    zc = pkg.grids.corner_point(nx, ny, nz, dx, dy, 1.0, seed=7, fault_frac=0.3, return_depths=True)
    ref = pkg.grids.corner_point(nx, ny, nz, dx, dy, 1.0, seed=7, fault_frac=0.3)
    txt = pkg.deck.write_grdecl(nx, ny, nz, dx, dy, zc, origin=(100.0, 50.0))
    g = pkg.deck.grid_from_deck(pkg.deck.read_grdecl(txt))
    np.testing.assert_array_equal(g.facepos, ref.facepos)
    np.testing.assert_array_equal(g.nb, ref.nb)
    np.testing.assert_allclose(g.cell_volume, ref.cell_volume, rtol=1e-13)
    np.testing.assert_allclose(g.face_area, ref.face_area, rtol=1e-13)
    np.testing.assert_allclose(g.cell_centroid, ref.cell_centroid + [100.0, 50.0, 0.0], rtol=1e-13)
    assert np.any(g.ncf() > 6)
    an = np.asarray(g.face_normal) * np.asarray(g.face_area)[:, None]
    closure = np.zeros((g.num_cells, 3))
    np.add.at(closure, g.hf_cell(), an)
    assert np.max(np.abs(closure)) < 1e-11 * np.max(g.face_area)
    col = zc[..., -1].mean(axis=2) - zc[..., 0].mean(axis=2)          # mean column thickness
    assert g.cell_volume.sum() == pytest.approx(dx * dy * col.sum(), rel=1e-12)


def test_unsupported_decks_are_rejected(pkg):
    nx, ny, nz = 2, 2, 2
    zc = np.broadcast_to(np.arange(nz + 1) * 1.0, (ny, nx, 4, nz + 1)).copy()
    d = pkg.deck.read_grdecl(pkg.deck.write_grdecl(nx, ny, nz, 1.0, 1.0, zc))
    d2 = dict(d)
    d2["ACTNUM"] = np.array([1, 1, 0, 1, 1, 1, 1, 1])
    with pytest.raises(ValueError, match="inactive"):
        pkg.deck.grid_from_deck(d2)
    d3 = dict(d)
    c = d["COORD"].copy().reshape(-1, 6)
    c[4, 3] += 0.5                               # a tilted pillar
    d3["COORD"] = c.reshape(-1)
    with pytest.raises(ValueError, match="vertical"):
        pkg.deck.grid_from_deck(d3)
    d4 = dict(d)
    z = d["ZCORN"].copy().reshape(nz, 2, -1)
    z[1, 0] += 0.1                               # a gap between the layers
    d4["ZCORN"] = z.reshape(-1)
    with pytest.raises(ValueError, match="gaps"):
        pkg.deck.grid_from_deck(d4)


def _perm_deck(**kw):
    d = {"SPECGRID": np.array([2, 1, 1])}
    for k, v in kw.items():
        d[k] = np.asarray(v, dtype=np.float64)
    return d


def test_permeability_assembly_rules(pkg):
    rock = pkg.deck.rock_from_deck
    # PERMX alone: isotropic (setScalarPermIfNeeded copies it to yy and zz)
    K, phi, kind = rock(_perm_deck(PERMX=[100, 200]))
    assert kind == "scalar"
    np.testing.assert_allclose(K[:, [0, 4, 8]], np.array([[100] * 3, [200] * 3]) * MD)
    assert np.all(K[:, [1, 2, 3, 5, 6, 7]] == 0) and np.all(phi == 1.0)
    # PERMX and PERMZ: yy follows PERMX
    K, _, kind = rock(_perm_deck(PERMX=[100, 200], PERMZ=[10, 20]))
    assert kind == "diagonal"
    np.testing.assert_allclose(K[:, [0, 4, 8]], np.array([[100, 100, 10], [200, 200, 20]]) * MD)
    # PERMY alone fills xx and zz
    K, _, _ = rock(_perm_deck(PERMY=[5, 6]))
    np.testing.assert_allclose(K[:, [0, 4, 8]], np.array([[5] * 3, [6] * 3]) * MD)
    # cross terms are symmetric; the later keyword wins for a pair
    K, _, kind = rock(_perm_deck(PERMX=[100, 100], PERMY=[50, 50], PERMZ=[10, 10], PERMXY=[7, 8],
                                 PERMZY=[3, 4]))
    assert kind == "tensor"
    np.testing.assert_allclose(K[:, 1], [7 * MD, 8 * MD])
    np.testing.assert_allclose(K[:, 3], K[:, 1])
    np.testing.assert_allclose(K[:, 5], [3 * MD, 4 * MD])
    np.testing.assert_allclose(K[:, 7], K[:, 5])
    # diagonal threshold, porosity, global cells
    K, phi, _ = rock(dict(_perm_deck(PERMX=[0.0, 200]), PORO=np.array([0.1, 0.3])), global_cell=[1, 0],
                     perm_threshold=1e-16)
    np.testing.assert_allclose(K[:, 0], [200 * MD, 1e-16])
    np.testing.assert_allclose(phi, [0.3, 0.1])
    # no permeability keywords: zero tensors
    K, _, kind = rock(_perm_deck())
    assert kind == "none" and np.all(K == 0)


def test_permeability_errors(pkg):
    with pytest.raises(RuntimeError, match="Invalid set of permeability"):
        pkg.deck.rock_from_deck(_perm_deck(PERMY=[1, 1], PERMZ=[1, 1], PERMXY=[1, 1]))
    with pytest.raises(RuntimeError, match="same size"):
        pkg.deck.rock_from_deck(_perm_deck(PERMX=[1, 1, 1]))
    with pytest.raises(RuntimeError, match="PORO"):
        pkg.deck.rock_from_deck(dict(_perm_deck(PERMX=[1, 1]), PORO=np.array([0.2])))


def test_reader_rejects_unreadable_values(pkg):
    with pytest.raises(ValueError, match="unreadable"):
        pkg.deck.read_grdecl("PORO 0.1 abc 0.2 /")

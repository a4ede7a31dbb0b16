"""The C++ ECLIPSE set-up and output of the drop-in on the host:
setupGridAndProps (setupGridAndProps.hpp:66-130) over CornerPointGrid and
ReservoirPropertyCommon::init(deck, ...), then writeVtkOutput / writeField
(SimulatorUtilities.hpp:232-297) for a host-only solution object, compiled
from tests/dropin_deck_check.cpp and compared with the Python reader."""
import os
import re
import subprocess

import numpy as np
import pytest
from cases import read_vtu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MD = 1e-3 * 9.869233e-13


@pytest.fixture(scope="module")
def checker(pkg, tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("dropin") / "dropin_deck_check")
    here = os.path.join(ROOT, "opm-porsol_amd")
    cmd = ["g++", "-O1", "-std=c++17", "-Wall", "-I", os.path.join(here, "host"), "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "dropin_deck_check.cpp"), "-o", exe, "-L", here, "-lmhs",
           f"-Wl,-rpath,{here}"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    return exe


def _run(exe, prefix, *args):
    res = subprocess.run([exe, prefix, *args], capture_output=True, text=True, timeout=120)
    return res.returncode, res.stdout + res.stderr


def _faulted_deck(pkg, path):
    nx, ny, nz, dx, dy = 6, 5, 4, 3.0, 2.0
    zc = pkg.grids.corner_point(nx, ny, nz, dx, dy, 1.0, seed=5, fault_frac=0.35, return_depths=True)
    rng = np.random.default_rng(9)
    n = nx * ny * nz
    props = {"PERMX": rng.uniform(50, 400, n), "PERMY": rng.uniform(50, 400, n), "PERMZ": rng.uniform(5, 40, n),
             "PORO": rng.uniform(0.1, 0.3, n)}
    txt = pkg.deck.write_grdecl(nx, ny, nz, dx, dy, zc, origin=(10.0, 5.0), props=props)
    open(path, "w").write(txt)
    return txt


def test_eclipse_setup_and_vtk_output(pkg, checker, tmp_path):
    deck = str(tmp_path / "faulted.grdecl")
    txt = _faulted_deck(pkg, deck)
    prefix = str(tmp_path / "out")
    rc, out = _run(checker, prefix, "fileformat=eclipse", f"filename={deck}", "perm_threshold_md=0.5")
    assert rc == 0, out
    py = pkg.deck.read_grdecl(txt)
    g = pkg.deck.grid_from_deck(py)
    K, phi, _ = pkg.deck.rock_from_deck(py, perm_threshold=0.5 * MD)
    nc, nhf = g.num_cells, g.facepos[-1]
    m = re.search(r"cells (\d+) half_faces (\d+) faces (\d+) logical (\d+) (\d+) (\d+) kind (\d+) unique (\d+)", out)
    nb = np.asarray(g.nb)
    nfaces = int((nb >= 0).sum() // 2 + (nb < 0).sum())
    assert tuple(int(x) for x in m.groups()) == (nc, nhf, nfaces, 6, 5, 4, 2, 0)
    rows = [l.split() for l in open(prefix + "-grid.txt") if not l.startswith("bid")]
    a = np.array(rows, dtype=np.float64)
    np.testing.assert_array_equal(a[:, 0], np.arange(nc))                     # globalCell
    np.testing.assert_allclose(a[:, 1], g.cell_volume, rtol=1e-13)
    np.testing.assert_array_equal(a[:, 2:5], K[:, [0, 4, 8]])
    np.testing.assert_array_equal(a[:, 5], phi)
    # boundary ids: CpGrid's side numbering, fault-exposed faces by their normal
    bid = np.array([int(l.split()[1]) for l in open(prefix + "-grid.txt") if l.startswith("bid")])
    expect = np.asarray(g.bid).copy()
    fault = expect == 7
    assert fault.any()
    n = np.asarray(g.face_normal)[fault]
    axis = np.where(np.abs(n[:, 0]) >= np.abs(n[:, 1]), 0, 1)
    expect[fault] = 2 * axis + np.where(n[np.arange(len(n)), axis] > 0, 2, 1)
    np.testing.assert_array_equal(bid, expect)
    assert set(np.unique(bid)) == {0, 1, 2, 3, 4, 5, 6}

    v = read_vtu(prefix + ".vtu")
    assert v["_sizes"] == (8 * nc, nc)
    s = (np.arange(nc) % 10) / 10.0
    np.testing.assert_array_equal(v["saturation"], s)
    np.testing.assert_array_equal(v["pressure"], 1e5 + np.arange(nc))
    np.testing.assert_allclose(v["capillary pressure"], 1e5 * (1 - s), rtol=1e-15)
    l1, l2 = s ** 2 / 1e-3, (1 - s) ** 2 / 3e-3
    f = l1 / (l1 + l2)
    np.testing.assert_allclose(v["fractional flow [water]"], f, rtol=1e-14)
    flux = np.sin(np.arange(nhf, dtype=np.float64))
    hc = g.hf_cell()
    vel = np.zeros((nc, 3))
    np.add.at(vel, hc, (np.asarray(g.face_centroid) - np.asarray(g.cell_centroid)[hc]) * flux[:, None])
    vel /= np.asarray(g.cell_volume)[:, None]
    np.testing.assert_allclose(v["velocity"], vel, rtol=1e-12, atol=1e-14 * np.abs(vel).max())
    np.testing.assert_allclose(v["phase velocity [water]"], vel * f[:, None], rtol=1e-12, atol=1e-14 * np.abs(vel).max())
    np.testing.assert_allclose(v["phase velocity [oil]"], vel * (1 - f)[:, None], rtol=1e-12,
                               atol=1e-14 * np.abs(vel).max())
    _, corners = pkg.deck.NativeDeck(text=txt).grid()
    np.testing.assert_array_equal(v["points"], corners.reshape(-1, 3))
    np.testing.assert_array_equal(v["connectivity"], np.arange(8 * nc))
    np.testing.assert_array_equal(v["offsets"], 8 * np.arange(1, nc + 1))
    np.testing.assert_array_equal(v["types"], 12)
    # writeField: the size, then one value per line (default stream precision)
    fld = open(prefix + "-field.txt").read().split()
    assert int(fld[0]) == nc and np.allclose(np.array(fld[1:], float), 0.5 * np.arange(nc))
    assert "writeField error: Could not open file /nonexistent-dir/x.txt" in out


def test_cartesian_setup_and_unique_ids(pkg, checker, tmp_path):
    prefix = str(tmp_path / "box")
    rc, out = _run(checker, prefix, "fileformat=cartesian", "nx=4", "ny=3", "nz=2", "dx=2", "dy=1.5", "dz=0.5",
                   "default_poro=0.25", "default_perm_md=50", "use_unique_boundary_ids=true")
    assert rc == 0, out
    assert "cells 24 half_faces 144 faces 98 logical 4 3 2 kind -1 unique 1" in out
    a = np.array([l.split() for l in open(prefix + "-grid.txt") if not l.startswith("bid")], dtype=np.float64)
    np.testing.assert_allclose(a[:, 1], 2 * 1.5 * 0.5)
    np.testing.assert_allclose(a[:, 2:5], 50 * MD)
    np.testing.assert_allclose(a[:, 5], 0.25)
    bid = np.array([int(l.split()[1]) for l in open(prefix + "-grid.txt") if l.startswith("bid")])
    nb = pkg.grids.cartesian(4, 3, 2, 2.0, 1.5, 0.5).nb
    np.testing.assert_array_equal(bid[nb >= 0], 0)
    np.testing.assert_array_equal(bid[nb < 0], np.arange(1, (nb < 0).sum() + 1))
    v = read_vtu(prefix + ".vtu")
    p = v["points"].reshape(24, 8, 3)
    np.testing.assert_allclose(p.max(axis=1) - p.min(axis=1), np.tile([2.0, 1.5, 0.5], (24, 1)))


def test_setup_errors(pkg, checker, tmp_path):
    rc, out = _run(checker, str(tmp_path / "x"), "fileformat=sintef_legacy")
    assert rc == 1 and "Unknown file format string: sintef_legacy" in out
    rc, out = _run(checker, str(tmp_path / "x"), "fileformat=eclipse", "filename=/nonexistent.grdecl")
    assert rc == 1 and "Could not open file /nonexistent.grdecl" in out
    deck = str(tmp_path / "d.grdecl")
    _faulted_deck(pkg, deck)
    rc, out = _run(checker, str(tmp_path / "x"), "fileformat=eclipse", f"filename={deck}", "rock_list=rocks.txt")
    assert rc == 1 and "rock tables" in out
    rc, out = _run(checker, str(tmp_path / "x"), "fileformat=eclipse", f"filename={deck}", "periodic_extension=1")
    assert rc == 1 and "not supported" in out
    rc, out = _run(checker, str(tmp_path / "x"), "fileformat=eclipse")
    assert rc == 1 and "missing parameter filename" in out

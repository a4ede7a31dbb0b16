"""Simulator output (opm-porsol_amd/output.py): the VTU file of hexahedra
with cell data reads back (counts, names, values), the hexahedra of a flat
corner-point deck are the Cartesian cells, and writeField's format."""
import xml.etree.ElementTree as ET

import numpy as np


def test_vtu_round_trip(pkg, tmp_path):
    nx, ny, nz, dx, dy, dz = 3, 2, 2, 2.0, 1.0, 0.5
    zc = np.broadcast_to(np.arange(nz + 1) * dz, (ny, nx, 4, nz + 1)).copy()
    P, conn = pkg.output.hexahedra(nx, ny, nz, dx, dy, zc, origin=(10.0, 0.0))
    g = pkg.grids.cartesian(nx, ny, nz, dx, dy, dz)
    nc = g.num_cells
    # each hexahedron spans its cell: the corner mean is the cell centroid
    np.testing.assert_allclose(P[conn].mean(axis=1), g.cell_centroid + [10.0, 0.0, 0.0], rtol=1e-14)
    sat = np.linspace(0.0, 1.0, nc)
    vel = np.arange(3 * nc, dtype=float).reshape(nc, 3)
    path = tmp_path / "out.vtu"
    pkg.output.write_vtu(str(path), P, conn, {"saturation": sat, "velocity": vel})
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfCells")) == nc and int(piece.get("NumberOfPoints")) == 8 * nc
    arrays = {a.get("Name"): a for a in piece.find("CellData")}
    assert set(arrays) == {"saturation", "velocity"}
    np.testing.assert_array_equal(np.array(arrays["saturation"].text.split(), float), sat)
    assert arrays["velocity"].get("NumberOfComponents") == "3"
    np.testing.assert_array_equal(np.array(arrays["velocity"].text.split(), float), vel.reshape(-1))
    types = [a for a in piece.find("Cells") if a.get("Name") == "types"][0]
    assert set(types.text.split()) == {"12"}


def test_write_field(pkg, tmp_path):
    path = tmp_path / "sat.txt"
    pkg.output.write_field([0.25, 0.5, 1.0], str(path))
    assert path.read_text().split("\n")[:4] == ["3", "0.25", "0.5", "1.0"]

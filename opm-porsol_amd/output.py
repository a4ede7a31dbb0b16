"""Simulator output: writeVtkOutput and writeField of the reference
(opm/porsol/common/SimulatorUtilities.hpp:234-285, 288-297).

``simulator_fields`` gathers the cell data writeVtkOutput writes: saturation,
pressure, capillary pressure, water fractional flow, the cell velocity
(estimateCellVelocity, computed on the GPU from the fluxes in HBM) and the
phase velocities f v and (1 - f) v (computePhaseVelocities :154-170), and
the capillary pressure of computeCapPressure (:220-230), for
the fluid without rock tables (krw = s^2, kro = (1-s)^2, p_c = 1e5 (1 - s)).
``write_vtu`` writes them as an ASCII VTK unstructured grid (the format the
reference asks Dune's VTKWriter for) with one hexahedron per cell of a
corner-point description; ``write_field`` is writeField.
"""
from __future__ import annotations

import numpy as np

DEFAULT_VISCOSITIES = (1.0e-3, 3.0e-3)


def simulator_fields(solver, saturation, viscosities=DEFAULT_VISCOSITIES) -> dict:
    """Cell data of writeVtkOutput (:234-285) for the last solve of ``solver``."""
    s = np.asarray(saturation, dtype=np.float64)
    v = solver.estimateCellVelocity()
    p = np.asarray(solver.getSolution()[0]).reshape(-1)[: s.size]
    lw, lo = s * s / viscosities[0], (1.0 - s) ** 2 / viscosities[1]
    f = lw / (lw + lo)
    return {
        "saturation": s,
        "pressure": p,
        "capillary pressure": 1e5 * (1.0 - s),
        "fractional flow [water]": f,
        "velocity": v,
        "phase velocity [water]": v * f[:, None],
        "phase velocity [oil]": v * (1.0 - f)[:, None],
    }


def hexahedra(nx, ny, nz, dx, dy, zc, origin=(0.0, 0.0)):
    """Points [8 nc, 3] and connectivity [nc, 8] (VTK_HEXAHEDRON order: the
    four corners of the top surface counter-clockwise, then the bottom) of the
    cells of a corner-point description zc[j, i, c, l] (as
    grids.corner_point_from_depths), cells numbered i fastest, then j, k."""
    zc = np.asarray(zc, dtype=np.float64)
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    I, J, K = I.reshape(-1), J.reshape(-1), K.reshape(-1)
    pts = []
    # corner c of the column: 0 (i, j), 1 (i+1, j), 2 (i, j+1), 3 (i+1, j+1)
    ring = ((0, 0, 0), (1, 1, 0), (3, 1, 1), (2, 0, 1))
    for lev in (0, 1):
        for c, ox, oy in ring:
            pts.append(np.stack([origin[0] + (I + ox) * dx, origin[1] + (J + oy) * dy, zc[J, I, c, K + lev]], axis=1))
    P = np.stack(pts, axis=1).reshape(-1, 3)                            # [nc*8, 3]
    conn = np.arange(P.shape[0]).reshape(-1, 8)
    return P, conn


def write_vtu(path: str, points, conn, cell_data: dict) -> None:
    """ASCII VTK unstructured grid of hexahedra with cell data (scalars, or
    [nc, 3] vectors written with three components)."""
    nc = conn.shape[0]

    def arr(a):
        return " ".join(repr(float(x)) for x in np.asarray(a).reshape(-1))

    with open(path, "w") as f:
        f.write('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n')
        f.write(f'<UnstructuredGrid>\n<Piece NumberOfPoints="{points.shape[0]}" NumberOfCells="{nc}">\n')
        f.write('<CellData>\n')
        for name, vals in cell_data.items():
            vals = np.asarray(vals, dtype=np.float64)
            ncomp = 1 if vals.ndim == 1 else vals.shape[1]
            if vals.shape[0] != nc:
                raise ValueError(f"write_vtu: {name} has {vals.shape[0]} cells, expected {nc}")
            f.write(f'<DataArray type="Float64" Name="{name}" NumberOfComponents="{ncomp}" format="ascii">\n')
            f.write(arr(vals) + "\n</DataArray>\n")
        f.write('</CellData>\n<Points>\n<DataArray type="Float64" NumberOfComponents="3" format="ascii">\n')
        f.write(arr(points) + "\n</DataArray>\n</Points>\n<Cells>\n")
        f.write('<DataArray type="Int64" Name="connectivity" format="ascii">\n')
        f.write(" ".join(str(int(x)) for x in conn.reshape(-1)) + "\n</DataArray>\n")
        f.write('<DataArray type="Int64" Name="offsets" format="ascii">\n')
        f.write(" ".join(str(8 * (c + 1)) for c in range(nc)) + "\n</DataArray>\n")
        f.write('<DataArray type="UInt8" Name="types" format="ascii">\n')
        f.write(" ".join(["12"] * nc) + "\n</DataArray>\n</Cells>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")


def write_field(field, filename: str) -> None:
    """writeField (:288-297): the size, then one value per line."""
    field = np.asarray(field, dtype=np.float64).reshape(-1)
    with open(filename, "w") as f:
        f.write(f"{field.size}\n")
        for x in field:
            f.write(f"{float(x)!r}\n")

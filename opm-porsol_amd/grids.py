"""Flat grid descriptions consumed by the C-ABI.

A grid is the information the reference solver pulls through
``GridInterfaceEuler`` (opm/porsol/common/GridInterfaceEuler.hpp:98-367):

  per cell        volume, centroid                      (Cell::volume :346, centroid :351)
  per half-face   area, centroid, unit outer normal,    (Face::area :126, centroid :134,
                  neighbour cell (-1 on the boundary),   normal :143, neighbourCellIndex :192,
                  boundary id (0 inside)                 boundaryId :160)

Half-faces are stored cell by cell in local-face order (``Face::localIndex``
:212), ``facepos[c]:facepos[c+1]`` being cell ``c``'s range, which is also the
half-face order of ``FlowSolution::outflux(hf)`` (IncompFlowSolverHybrid.hpp:430).

``cartesian`` reproduces a Dune YaspGrid: cells lexicographic with x fastest,
local faces (x-, x+, y-, y+, z-, z+), boundary ids 1..6 by side
(setupBoundaryConditions.hpp:84-89).  ``corner_point`` builds hexahedral
pillar grids with perturbed layers and vertically offset columns, whose split
faces give cells with 6 to 10 faces.
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class Grid:
    facepos: np.ndarray          # int32 [nc+1]
    nb: np.ndarray               # int32 [nhf]  neighbour cell, -1 on the boundary
    bid: np.ndarray              # int32 [nhf]  boundary id, 0 for internal half-faces
    cell_volume: np.ndarray      # f64 [nc]
    cell_centroid: np.ndarray    # f64 [nc,3]
    face_area: np.ndarray        # f64 [nhf]
    face_centroid: np.ndarray    # f64 [nhf,3]
    face_normal: np.ndarray      # f64 [nhf,3] unit outer normal
    side: np.ndarray             # int8 [nhf] canonical side 0..5 of boundary half-faces, -1 inside
    dims: tuple | None = None    # (nx, ny, nz) for logically Cartesian grids
    ijk: np.ndarray | None = None  # int32 [nc,3] logical cell indices (instance-local)

    @property
    def num_cells(self) -> int:
        return int(self.facepos.shape[0] - 1)

    @property
    def num_half_faces(self) -> int:
        return int(self.nb.shape[0])

    def ncf(self) -> np.ndarray:
        return np.diff(self.facepos)

    def hf_cell(self) -> np.ndarray:
        return np.repeat(np.arange(self.num_cells, dtype=np.int32), self.ncf())

    def twin(self) -> np.ndarray:
        """For each internal half-face, the neighbour's half-face on the same face."""
        nc = self.num_cells
        hc = self.hf_cell().astype(np.int64)
        nb = self.nb.astype(np.int64)
        tw = np.full(self.num_half_faces, -1, dtype=np.int64)
        inside = nb >= 0
        # boundary half-faces get distinct negative keys so they never match
        key_self = np.where(inside, hc * nc + nb, -1 - np.arange(nb.shape[0]))
        key_twin = np.where(inside, nb * nc + hc, -2 - nb.shape[0])
        order = np.argsort(key_self, kind="stable")
        idx = np.searchsorted(key_self[order], key_twin)
        idx = np.clip(idx, 0, len(order) - 1)
        cand = order[idx]
        ok = inside & (key_self[cand] == key_twin)
        tw[ok] = cand[ok]
        return tw.astype(np.int32)


def cartesian(nx: int, ny: int, nz: int, dx: float = 1.0, dy: float = 1.0,
              dz: float = 1.0, unique_bids: bool = False) -> Grid:
    """Uniform box grid with YaspGrid cell and intersection ordering."""
    nc = nx * ny * nz
    c = np.arange(nc, dtype=np.int64)
    i = c % nx
    j = (c // nx) % ny
    k = c // (nx * ny)
    facepos = (6 * np.arange(nc + 1)).astype(np.int32)
    nb = np.empty((nc, 6), dtype=np.int64)
    nb[:, 0] = np.where(i > 0, c - 1, -1)
    nb[:, 1] = np.where(i < nx - 1, c + 1, -1)
    nb[:, 2] = np.where(j > 0, c - nx, -1)
    nb[:, 3] = np.where(j < ny - 1, c + nx, -1)
    nb[:, 4] = np.where(k > 0, c - nx * ny, -1)
    nb[:, 5] = np.where(k < nz - 1, c + nx * ny, -1)
    side = np.where(nb < 0, np.arange(6)[None, :], -1)
    bid = np.where(nb < 0, np.arange(1, 7)[None, :], 0)
    if unique_bids:
        flat = (nb.reshape(-1) < 0)
        ub = np.zeros(flat.shape[0], dtype=np.int64)
        ub[flat] = np.arange(1, int(flat.sum()) + 1)
        bid = ub.reshape(nc, 6)
    cen = np.stack([(i + 0.5) * dx, (j + 0.5) * dy, (k + 0.5) * dz], axis=1)
    area = np.array([dy * dz, dy * dz, dx * dz, dx * dz, dx * dy, dx * dy])
    off = np.array([[-0.5 * dx, 0, 0], [0.5 * dx, 0, 0], [0, -0.5 * dy, 0],
                    [0, 0.5 * dy, 0], [0, 0, -0.5 * dz], [0, 0, 0.5 * dz]])
    nrm = np.array([[-1.0, 0, 0], [1.0, 0, 0], [0, -1.0, 0], [0, 1.0, 0],
                    [0, 0, -1.0], [0, 0, 1.0]])
    fcen = (cen[:, None, :] + off[None, :, :]).reshape(-1, 3)
    return Grid(
        facepos=facepos,
        nb=nb.reshape(-1).astype(np.int32),
        bid=bid.reshape(-1).astype(np.int32),
        cell_volume=np.full(nc, dx * dy * dz),
        cell_centroid=cen,
        face_area=np.tile(area, nc),
        face_centroid=fcen,
        face_normal=np.tile(nrm, (nc, 1)),
        side=side.reshape(-1).astype(np.int8),
        dims=(nx, ny, nz),
        ijk=np.stack([i, j, k], axis=1).astype(np.int32),
    )


def periodic_partners_cartesian(g: Grid, directions=(0, 1, 2)) -> np.ndarray:
    """Partner half-face of every boundary half-face on the periodic sides.

    The reference pairs faces by centroid matching across opposite sides of the
    bounding box (BoundaryPeriodicity.hpp:86-175, match() in
    BoundaryPeriodicity.cpp:25-48); on a logically Cartesian grid the pairing
    is the index map (0,j,k) <-> (nx-1,j,k) and likewise in y and z.
    Returns int32 [nhf], -1 where there is no partner.
    """
    nx, ny, nz = g.dims
    nc = g.num_cells
    c = np.arange(nc, dtype=np.int64)
    i = c % nx
    j = (c // nx) % ny
    k = c // (nx * ny)
    partner = np.full(g.num_half_faces, -1, dtype=np.int64)
    for d in directions:
        if d == 0:
            lo = c[i == 0]; hi = lo + (nx - 1)
        elif d == 1:
            lo = c[j == 0]; hi = lo + (ny - 1) * nx
        else:
            lo = c[k == 0]; hi = lo + (nz - 1) * nx * ny
        hf_lo = 6 * lo + 2 * d
        hf_hi = 6 * hi + 2 * d + 1
        partner[hf_lo] = hf_hi
        partner[hf_hi] = hf_lo
    return partner.astype(np.int32)


def match_periodic_partners(g: Grid, periodic_sides, tol: float = 1e-6) -> np.ndarray:
    """Geometric pairing of boundary half-faces on opposite sides.

    Restates the rule of BoundaryPeriodicity.hpp:86-175 and
    BoundaryPeriodicity.cpp:25-48: a boundary face on side 2d (resp. 2d+1) of
    the bounding box pairs with the face on the opposite side whose area
    agrees within ``tol`` and whose centroid, with the d-th coordinate
    zeroed, agrees within ``tol``.  ``periodic_sides`` is a length-6 bool
    sequence.  Returns int32 [nhf], -1 where there is no partner.
    """
    bnd = np.nonzero(g.nb < 0)[0]
    fc = g.face_centroid[bnd]
    lo = fc.min(axis=0)
    hi = fc.max(axis=0)
    canon = np.full(bnd.shape[0], -1, dtype=np.int64)
    for d in range(3):
        unset = canon < 0
        at_lo = unset & (np.abs(fc[:, d] - lo[d]) <= tol)
        canon[at_lo] = 2 * d
        unset = canon < 0
        at_hi = unset & (np.abs(fc[:, d] - hi[d]) <= tol)
        canon[at_hi] = 2 * d + 1
    if np.any(canon < 0):
        raise ValueError("boundary face centroid not on bounding box")
    partner = np.full(g.num_half_faces, -1, dtype=np.int64)
    area = g.face_area[bnd]
    for d in range(3):
        if not (periodic_sides[2 * d] and periodic_sides[2 * d + 1]):
            continue
        a = np.nonzero(canon == 2 * d)[0]
        b = np.nonzero(canon == 2 * d + 1)[0]
        ka = fc[a].copy(); ka[:, d] = 0.0
        kb = fc[b].copy(); kb[:, d] = 0.0
        # Sort both sides by a rounded key, then verify tolerance pairwise.
        scale = 1.0 / max(tol, 1e-300)
        ra = np.round(ka * scale * 0.5).astype(np.int64)
        rb = np.round(kb * scale * 0.5).astype(np.int64)
        oa = np.lexsort(ra.T[::-1])
        ob = np.lexsort(rb.T[::-1])
        if a.shape[0] != b.shape[0]:
            raise ValueError("unequal numbers of faces on periodic sides")
        dist = np.linalg.norm(ka[oa] - kb[ob], axis=1)
        darea = np.abs(area[a[oa]] - area[b[ob]])
        if np.any(dist > tol) or np.any(darea > tol):
            raise ValueError("periodic sides do not match")
        partner[bnd[a[oa]]] = bnd[b[ob]]
        partner[bnd[b[ob]]] = bnd[a[oa]]
    return partner.astype(np.int32)


def concat(grids) -> tuple[Grid, np.ndarray]:
    """Disjoint union of several grids (a batch of independent instances).

    Cell and half-face arrays are concatenated; neighbour indices are shifted
    to the global cell numbering.  Returns the union and inst_cellpos[ninst+1].
    """
    cell_off = [0]
    hf_off = [0]
    for g in grids:
        cell_off.append(cell_off[-1] + g.num_cells)
        hf_off.append(hf_off[-1] + g.num_half_faces)
    facepos = [np.zeros(1, dtype=np.int64)]
    nb = []
    for t, g in enumerate(grids):
        facepos.append(g.facepos[1:].astype(np.int64) + hf_off[t])
        nb.append(np.where(g.nb >= 0, g.nb.astype(np.int64) + cell_off[t], -1))
    u = Grid(
        facepos=np.concatenate(facepos).astype(np.int32),
        nb=np.concatenate(nb).astype(np.int32),
        bid=np.concatenate([g.bid for g in grids]).astype(np.int32),
        cell_volume=np.concatenate([g.cell_volume for g in grids]),
        cell_centroid=np.concatenate([g.cell_centroid for g in grids]),
        face_area=np.concatenate([g.face_area for g in grids]),
        face_centroid=np.concatenate([g.face_centroid for g in grids]),
        face_normal=np.concatenate([g.face_normal for g in grids]),
        side=np.concatenate([g.side for g in grids]).astype(np.int8),
        dims=None,
        ijk=(np.concatenate([g.ijk for g in grids]).astype(np.int32)
             if all(g.ijk is not None for g in grids) else None),
    )
    return u, np.asarray(cell_off, dtype=np.int32)

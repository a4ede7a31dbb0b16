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


# --- corner-point grids with faults ---------------------------------------------
_GAUSS = (0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0))


def _bilinear_weights(xi, eta):
    """Weights of the corners (0,0), (1,0), (0,1), (1,1) at (xi, eta)."""
    return np.array([(1 - xi) * (1 - eta), xi * (1 - eta), (1 - xi) * eta, xi * eta])


def _overlap_integrals(aL1, aL2, aH1, aH2, bL1, bL2, bH1, bH2):
    """For two bands {aL(s) <= z <= aH(s)} and {bL(s) <= z <= bH(s)} over
    s in [0, 1], each bound linear (value 1 at s = 0, value 2 at s = 1), the
    integrals over their intersection of 1, s and z: (I0, Is, Iz).  The
    height h(s) = min(aH, bH) - max(aL, bL) is piecewise linear with kinks
    only where two bounds cross, so Simpson's rule between the sorted
    crossings is exact."""
    L1 = np.stack([aL1, bL1, aH1, bH1])
    L2 = np.stack([aL2, bL2, aH2, bH2])
    n = aL1.shape[0]
    pts = [np.zeros(n), np.ones(n)]
    for p in range(4):
        for q in range(p + 1, 4):
            d1, d2 = L1[p] - L1[q], L2[p] - L2[q]
            den = d1 - d2
            with np.errstate(divide="ignore", invalid="ignore"):
                s = np.where(den != 0.0, d1 / np.where(den != 0.0, den, 1.0), 0.0)
            pts.append(np.where((s > 0.0) & (s < 1.0), s, 0.0))
    B = np.sort(np.stack(pts, axis=1), axis=1)

    def band(s):
        lo = np.maximum(aL1 + s * (aL2 - aL1), bL1 + s * (bL2 - bL1))
        hi = np.minimum(aH1 + s * (aH2 - aH1), bH1 + s * (bH2 - bH1))
        return hi - lo, 0.5 * (hi * hi - lo * lo)

    I0 = np.zeros(n); Is = np.zeros(n); Iz = np.zeros(n)
    for t in range(B.shape[1] - 1):
        a, b = B[:, t], B[:, t + 1]
        m = 0.5 * (a + b)
        w = (b - a) / 6.0
        ha, qa = band(a)
        hm, qm = band(m)
        hb, qb = band(b)
        pos = (hm > 0.0) & (b > a)
        I0 += np.where(pos, w * (ha + 4 * hm + hb), 0.0)
        Is += np.where(pos, w * (a * ha + 4 * m * hm + b * hb), 0.0)
        Iz += np.where(pos, w * (qa + 4 * qm + qb), 0.0)
    return I0, Is, Iz


def corner_point(nx: int, ny: int, nz: int, dx: float = 1.0, dy: float = 1.0, dz: float = 1.0,
                 perturb: float = 0.3, fault_frac: float = 0.05, fault_shift: float = 0.5,
                 seed: int = 0, return_depths: bool = False) -> Grid:
    """Faulted corner-point grid on vertical pillars (z is depth).

    Every pillar node depth is k dz + U(-perturb, perturb) dz (seeded); a
    fraction fault_frac of the columns is shifted down by fault_shift dz.
    Between columns with different shifts, each cell side is split into the
    overlaps with the neighbour column's cells (exact polygons on the shared
    vertical plane) plus the parts that face no cell (no-flow boundary, id 7,
    side 6), so faulted cells have more than 6 faces.  Top and bottom faces
    are bilinear surfaces described by their exact vector area (area times
    unit normal) and the mean of their corner depths; lateral faces are
    planar and exact; cell volumes and centroids are exact (2x2 Gauss
    quadrature of the bilinear thickness).

    return_depths=True returns the corner depths zc[j, i, c, l] instead of
    the grid (the input of corner_point_from_depths and deck.write_grdecl).

    Local face order per cell: x- side, x+, y-, y+ (sub-faces by neighbour
    layer, exposed parts first/last), top (z-), bottom (z+).  Boundary ids
    1..6 by side of the box, 7 for fault-exposed parts.  This is the layout
    Dune::CpGrid builds from a corner-point description, up to its face
    order, which is its own (parity unpinned here).
    """
    from . import fields
    P = (nx + 1) * (ny + 1)
    u = fields.uniform(seed, 11, P * (nz + 1)).reshape(P, nz + 1)
    Z = (np.arange(nz + 1)[None, :] + perturb * (2.0 * u - 1.0)) * dz
    fu = fields.uniform(seed, 12, nx * ny).reshape(ny, nx)
    off = np.where(fu < fault_frac, fault_shift * dz, 0.0)
    J, I = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")      # [ny, nx]

    def pid(i, j):
        return i + (nx + 1) * j

    # zc[j, i, c, l]: depth of corner c (0:(i,j) 1:(i+1,j) 2:(i,j+1) 3:(i+1,j+1)) at level l
    zc = np.stack([Z[pid(I, J)], Z[pid(I + 1, J)], Z[pid(I, J + 1)], Z[pid(I + 1, J + 1)]], axis=2)
    zc = zc + off[:, :, None, None]
    if return_depths:
        return zc
    return corner_point_from_depths(nx, ny, nz, dx, dy, zc, dz)


def corner_point_from_depths(nx: int, ny: int, nz: int, dx: float, dy: float, zc: np.ndarray,
                             dz: float | None = None, origin=(0.0, 0.0)) -> Grid:
    """Corner-point geometry on vertical pillars at (x0 + i dx, y0 + j dy).

    zc[j, i, c, l] is the depth of corner c (0: pillar (i, j), 1: (i+1, j),
    2: (i, j+1), 3: (i+1, j+1)) of column (i, j) at level l (the top of
    layer l; l = nz is the bottom of the last layer); layers within a column
    have no gaps.  Two neighbouring columns whose corner depths coincide on
    their shared pillars are matching (1:1 faces); otherwise every overlap of
    their layers on the shared vertical plane is a face, and parts facing no
    cell are no-flow boundary (id 7, side 6).  dz scales the overlap
    tolerance (default: the mean layer thickness).  Faces, ids and order as
    described in corner_point().
    """
    zc = np.asarray(zc, dtype=np.float64)
    if zc.shape != (ny, nx, 4, nz + 1):
        raise ValueError(f"corner_point_from_depths: zc must have shape {(ny, nx, 4, nz + 1)}, got {zc.shape}")
    if dz is None:
        dz = float(np.mean(zc[..., -1] - zc[..., 0])) / nz
    J, I = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")      # [ny, nx]
    zmin, zmax = zc.min() - 10.0 * dz, zc.max() + 10.0 * dz
    nc = nx * ny * nz

    def cidx(i, j, k):
        return i + nx * (j + ny * k)

    # --- cells: exact volume and centroid of the column slab --------------------
    V = np.zeros((ny, nx, nz)); MX = np.zeros_like(V); MY = np.zeros_like(V); MZ = np.zeros_like(V)
    for xi in _GAUSS:
        for eta in _GAUSS:
            wgt = _bilinear_weights(xi, eta)
            zt = np.einsum("c,jick->jik", wgt, zc[:, :, :, :-1])
            zb = np.einsum("c,jick->jik", wgt, zc[:, :, :, 1:])
            h = 0.25 * dx * dy * (zb - zt)
            V += h
            MX += h * ((I[:, :, None] + xi) * dx)
            MY += h * ((J[:, :, None] + eta) * dy)
            MZ += 0.25 * dx * dy * 0.5 * (zb * zb - zt * zt)
    # [ny, nx, nz] -> cell order (x fastest, layer slowest)
    to_cells = lambda a: np.transpose(a, (2, 0, 1)).reshape(-1)
    cvol = to_cells(V)
    ccen = np.stack([to_cells(MX), to_cells(MY), to_cells(MZ)], axis=1) / cvol[:, None]
    if np.any(cvol <= 0.0):
        raise ValueError("corner_point: non-positive cell volume")

    rec = {k: [] for k in ("cell", "key", "nb", "area", "cen", "nrm", "bid", "side")}

    def add(cell, key, nb, area, cen, nrm, bid, side):
        n = cell.shape[0]
        rec["cell"].append(cell.astype(np.int64)); rec["key"].append(np.broadcast_to(key, n).astype(np.int64))
        rec["nb"].append(np.broadcast_to(nb, n).astype(np.int64)); rec["area"].append(area)
        rec["cen"].append(cen); rec["nrm"].append(np.broadcast_to(nrm, (n, 3)).astype(np.float64))
        rec["bid"].append(np.broadcast_to(bid, n).astype(np.int64))
        rec["side"].append(np.broadcast_to(side, n).astype(np.int64))

    # --- top / bottom faces: bilinear surfaces ----------------------------------
    for l in range(nz + 1):
        z = zc[:, :, :, l]                                                 # [ny, nx, 4]
        gx = dy * 0.5 * ((z[..., 1] - z[..., 0]) + (z[..., 3] - z[..., 2]))   # integral of dz/dx
        gy = dx * 0.5 * ((z[..., 2] - z[..., 0]) + (z[..., 3] - z[..., 1]))
        N = np.stack([-gx, -gy, np.full_like(gx, dx * dy)], axis=-1).reshape(-1, 3)   # +z (down)
        A = np.linalg.norm(N, axis=1)
        nrm = N / A[:, None]
        cen = np.stack([((I + 0.5) * dx).reshape(-1), ((J + 0.5) * dy).reshape(-1),
                        z.mean(axis=-1).reshape(-1)], axis=1)
        ii, jj = I.reshape(-1), J.reshape(-1)
        if l > 0:      # bottom of layer l-1
            up = cidx(ii, jj, l - 1)
            add(up, 5 * 1000, cidx(ii, jj, l) if l < nz else -1, A, cen, nrm,
                0 if l < nz else 6, -1 if l < nz else 5)
        if l < nz:     # top of layer l
            lo = cidx(ii, jj, l)
            add(lo, 4 * 1000, cidx(ii, jj, l - 1) if l > 0 else -1, A, cen, -nrm,
                0 if l > 0 else 5, -1 if l > 0 else 4)

    # --- lateral faces ---------------------------------------------------------
    def lateral(iA, jA, iB, jB, cA, cB, direction):
        """Columns A -> B (B may be absent: iB < 0) across the plane through
        A's corners cA = (p1, p2); B's matching corners cB.  direction 0: x, 1: y."""
        nrm = np.zeros(3); nrm[direction] = 1.0
        sideA, sideB = 2 * direction + 1, 2 * direction
        L = dy if direction == 0 else dx
        zA1, zA2 = zc[jA, iA, cA[0]], zc[jA, iA, cA[1]]                     # [npair, nz+1]
        if direction == 0:
            px = (iA + 1) * dx * np.ones_like(iA, dtype=np.float64); p1y = jA * dy
        else:
            py = (jA + 1) * dy * np.ones_like(jA, dtype=np.float64); p1x = iA * dx

        def point(sbar, zbar, sel):
            if direction == 0:
                return np.stack([px[sel], p1y[sel] + sbar * dy, zbar], axis=1)
            return np.stack([p1x[sel] + sbar * dx, py[sel], zbar], axis=1)

        if iB is None:   # domain boundary: whole side, outward normal along +direction
            for k in range(nz):
                I0, Is, Iz = _overlap_integrals(zA1[:, k], zA2[:, k], zA1[:, k + 1], zA2[:, k + 1],
                                                zA1[:, k], zA2[:, k], zA1[:, k + 1], zA2[:, k + 1])
                sel = np.arange(iA.shape[0])
                add(cidx(iA, jA, k), sideA * 1000 + k + 1, -1, L * I0, point(Is / I0, Iz / I0, sel),
                    nrm, sideA + 1, sideA)
            return
        zB1, zB2 = zc[jB, iB, cB[0]], zc[jB, iB, cB[1]]
        same = np.all(zA1 == zB1, axis=1) & np.all(zA2 == zB2, axis=1)
        # matching columns: 1:1 faces
        sel = np.nonzero(same)[0]
        for k in range(nz):
            I0, Is, Iz = _overlap_integrals(zA1[sel, k], zA2[sel, k], zA1[sel, k + 1], zA2[sel, k + 1],
                                            zA1[sel, k], zA2[sel, k], zA1[sel, k + 1], zA2[sel, k + 1])
            cen = point(Is / I0, Iz / I0, sel)
            a, b = cidx(iA[sel], jA[sel], k), cidx(iB[sel], jB[sel], k)
            add(a, sideA * 1000 + k + 1, b, L * I0, cen, nrm, 0, -1)
            add(b, sideB * 1000 + k + 1, a, L * I0, cen, -nrm, 0, -1)
        # faulted pairs: every overlap of A's layers with B's layers and the
        # virtual layers above (-1) and below (nz) each column
        sel = np.nonzero(~same)[0]
        if sel.shape[0] == 0:
            return

        def bounds(z1, z2, k):
            if k == -1:
                return np.full(z1.shape[0], zmin), np.full(z1.shape[0], zmin), z1[:, 0], z2[:, 0]
            if k == nz:
                return z1[:, nz], z2[:, nz], np.full(z1.shape[0], zmax), np.full(z1.shape[0], zmax)
            return z1[:, k], z2[:, k], z1[:, k + 1], z2[:, k + 1]

        a1, a2, b1, b2 = zA1[sel], zA2[sel], zB1[sel], zB2[sel]
        for ka in range(-1, nz + 1):
            for kb in range(-1, nz + 1):
                if (ka < 0 or ka >= nz) and (kb < 0 or kb >= nz):
                    continue
                if 0 <= ka < nz and 0 <= kb < nz and abs(ka - kb) > 3:
                    continue
                aL1, aL2, aH1, aH2 = bounds(a1, a2, ka)
                bL1, bL2, bH1, bH2 = bounds(b1, b2, kb)
                I0, Is, Iz = _overlap_integrals(aL1, aL2, aH1, aH2, bL1, bL2, bH1, bH2)
                keep = I0 > 1e-12 * dz
                if not np.any(keep):
                    continue
                s2 = sel[keep]
                cen = point(Is[keep] / I0[keep], Iz[keep] / I0[keep], s2)
                area = L * I0[keep]
                real_a, real_b = 0 <= ka < nz, 0 <= kb < nz
                if real_a:
                    a = cidx(iA[s2], jA[s2], ka)
                    nbr = cidx(iB[s2], jB[s2], kb) if real_b else -1
                    add(a, sideA * 1000 + kb + 1, nbr, area, cen, nrm, 0 if real_b else 7,
                        -1 if real_b else 6)
                if real_b:
                    b = cidx(iB[s2], jB[s2], kb)
                    nbr = cidx(iA[s2], jA[s2], ka) if real_a else -1
                    add(b, sideB * 1000 + ka + 1, nbr, area, cen, -nrm, 0 if real_a else 7,
                        -1 if real_a else 6)

    ii, jj = I.reshape(-1), J.reshape(-1)
    # x direction: A = (i, j), B = (i+1, j); plane through corners 1, 3 of A (0, 2 of B)
    m = ii < nx - 1
    lateral(ii[m], jj[m], ii[m] + 1, jj[m], (1, 3), (0, 2), 0)
    # y direction: A = (i, j), B = (i, j+1); corners 2, 3 of A (0, 1 of B)
    m = jj < ny - 1
    lateral(ii[m], jj[m], ii[m], jj[m] + 1, (2, 3), (0, 1), 1)
    # domain sides: x+ and y+ use the same routine with no neighbour; x- and y-
    # are the mirrored planes (corners 0, 2 resp. 0, 1 with the normal reversed)
    m = ii == nx - 1
    lateral(ii[m], jj[m], None, None, (1, 3), None, 0)
    m = jj == ny - 1
    lateral(ii[m], jj[m], None, None, (2, 3), None, 1)
    for direction, corners in ((0, (0, 2)), (1, (0, 1))):
        m = (ii == 0) if direction == 0 else (jj == 0)
        iA, jA = ii[m], jj[m]
        z1, z2 = zc[jA, iA, corners[0]], zc[jA, iA, corners[1]]
        nrm = np.zeros(3); nrm[direction] = -1.0
        side = 2 * direction
        for k in range(nz):
            I0, Is, Iz = _overlap_integrals(z1[:, k], z2[:, k], z1[:, k + 1], z2[:, k + 1],
                                            z1[:, k], z2[:, k], z1[:, k + 1], z2[:, k + 1])
            sb, zb = Is / I0, Iz / I0
            if direction == 0:
                cen = np.stack([np.zeros_like(sb), (jA + sb) * dy, zb], axis=1)
                L = dy
            else:
                cen = np.stack([(iA + sb) * dx, np.zeros_like(sb), zb], axis=1)
                L = dx
            add(cidx(iA, jA, k), side * 1000 + k + 1, -1, L * I0, cen, nrm, side + 1, side)

    cell = np.concatenate(rec["cell"]); key = np.concatenate(rec["key"])
    order = np.lexsort((key, cell))
    cell = cell[order]
    nhf_per = np.bincount(cell, minlength=nc)
    facepos = np.concatenate([[0], np.cumsum(nhf_per)]).astype(np.int32)
    cat = lambda name: np.concatenate(rec[name])[order]
    shift = np.array([origin[0], origin[1], 0.0])
    ccen = ccen + shift
    return Grid(
        facepos=facepos,
        nb=cat("nb").astype(np.int32),
        bid=cat("bid").astype(np.int32),
        cell_volume=cvol,
        cell_centroid=ccen,
        face_area=cat("area"),
        face_centroid=cat("cen") + shift,
        face_normal=cat("nrm"),
        side=cat("side").astype(np.int8),
        dims=None,
        ijk=np.stack([np.tile(np.arange(nx), ny * nz), np.tile(np.repeat(np.arange(ny), nx), nz),
                      np.repeat(np.arange(nz), nx * ny)], axis=1).astype(np.int32),
    )

"""ECLIPSE grid and rock input: the ECLIPSE branch of the reference's set-up
(setupGridAndProps, opm/porsol/common/setupGridAndProps.hpp:78-117) and
ReservoirPropertyCommon::init(deck, global_cell, perm_threshold, ...)
(opm/porsol/common/ReservoirPropertyCommon_impl.hpp:275-300).

- ``read_grdecl``: keyword / value records of a GRDECL file (values ended
  by ``/``, ``n*v`` repeats, ``--`` comments, section keywords without data).
- ``grid_from_deck``: SPECGRID (or DIMENS), COORD and ZCORN into the solver's
  grid.  The geometry is exact for vertical pillars on a regular lattice with
  gap-free layers (faults between columns allowed); other decks, inactive
  cells (ACTNUM 0) and FIELD units are rejected with a ValueError.  The
  reference reads decks through dune-cornerpoint's processEclipseFormat,
  which is not in its tree: its geometry and face order are not reproduced
  here (parity unpinned, as for the corner-point generator).
- ``rock_from_deck``: permeability tensors (classifyPermeability,
  fillTensor, setScalarPermIfNeeded, assignPermeability :52-244, 682-736)
  and porosity (assignPorosity :592-611), in SI units.
- ``write_grdecl``: the inverse of the above, for tests and examples.
"""
from __future__ import annotations

import re

import numpy as np

from . import grids

# opm-core units: milli * darcy (9.869233e-13 m^2)
MILLIDARCY = 1e-3 * 9.869233e-13

SECTION_KEYWORDS = {"RUNSPEC", "GRID", "EDIT", "PROPS", "REGIONS", "SOLUTION", "SUMMARY", "SCHEDULE",
                    "NOECHO", "ECHO", "METRIC", "FIELD", "LAB", "END", "NONNC", "OLDTRAN", "NEWTRAN"}
INT_KEYWORDS = {"SPECGRID", "DIMENS", "ACTNUM"}


def _values(tokens, keyword):
    out = []
    for t in tokens:
        if "*" in t:
            n, v = t.split("*", 1)
            if v == "":
                raise ValueError(f"{keyword}: default values ('{t}') are not supported")
            out.extend([v] * int(n))
        else:
            out.append(t)
    return out


def _parse(rest: str, kw: str, dtype):
    toks = rest.split()
    if "*" in rest:
        toks = _values(toks, kw)
    try:
        return np.array(toks, dtype=dtype)
    except ValueError as e:
        raise ValueError(f"read_grdecl: unreadable value in {kw}: {e}") from None


def read_grdecl(text: str) -> dict:
    """Keyword records of a GRDECL text: {keyword: numpy array} (raw values,
    deck units); section keywords map to an empty array."""
    deck = {}
    text = re.sub(r"--[^\n]*", " ", text)
    parts = text.split("/")
    for n, part in enumerate(parts):
        rest = part
        while True:
            head = rest.split(None, 1)
            if not head or head[0] not in SECTION_KEYWORDS:
                break
            deck[head[0]] = np.zeros(0)
            rest = head[1] if len(head) > 1 else ""
        if not head:
            continue
        kw = head[0]
        rest = head[1] if len(head) > 1 else ""
        if not re.fullmatch(r"[A-Z][A-Z0-9_]{0,7}", kw):
            raise ValueError(f"read_grdecl: expected a keyword, got '{kw}'")
        if n == len(parts) - 1:
            raise ValueError(f"read_grdecl: keyword {kw} is not terminated by '/'")
        if kw == "SPECGRID":
            deck[kw] = np.array(rest.split()[:3], dtype=np.int64)     # NX NY NZ (then NUMRES, radial flag)
        elif kw in INT_KEYWORDS:
            deck[kw] = _parse(rest, kw, np.int64)
        else:
            deck[kw] = _parse(rest, kw, np.float64)
    if "FIELD" in deck or "LAB" in deck:
        raise ValueError("read_grdecl: only METRIC units are supported")
    return deck


def deck_dims(deck: dict) -> tuple[int, int, int]:
    for kw in ("SPECGRID", "DIMENS"):
        if kw in deck:
            nx, ny, nz = (int(v) for v in deck[kw][:3])
            return nx, ny, nz
    raise ValueError("deck has neither SPECGRID nor DIMENS")


def grid_from_deck(deck: dict, rtol: float = 1e-9) -> grids.Grid:
    """Corner-point grid of a deck with vertical pillars on a regular lattice,
    gap-free layers and every cell active."""
    nx, ny, nz = deck_dims(deck)
    coord = deck.get("COORD")
    zcorn = deck.get("ZCORN")
    if coord is None or zcorn is None:
        raise ValueError("deck needs COORD and ZCORN")
    if coord.size != 6 * (nx + 1) * (ny + 1):
        raise ValueError(f"COORD has {coord.size} values, expected {6 * (nx + 1) * (ny + 1)}")
    if zcorn.size != 8 * nx * ny * nz:
        raise ValueError(f"ZCORN has {zcorn.size} values, expected {8 * nx * ny * nz}")
    if "ACTNUM" in deck and np.any(deck["ACTNUM"] == 0):
        raise ValueError("inactive cells (ACTNUM 0) are not supported")
    p = coord.reshape(ny + 1, nx + 1, 6)
    x, y = p[..., 0], p[..., 1]
    scale = max(np.ptp(x), np.ptp(y), 1.0)
    if np.any(np.abs(p[..., 3] - x) > rtol * scale) or np.any(np.abs(p[..., 4] - y) > rtol * scale):
        raise ValueError("only vertical pillars are supported")
    x0, y0 = x[0, 0], y[0, 0]
    dx = (x[0, -1] - x0) / nx if nx > 0 else 1.0
    dy = (y[-1, 0] - y0) / ny if ny > 0 else 1.0
    ii, jj = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1))
    if (np.max(np.abs(x - (x0 + ii * dx))) > rtol * scale or np.max(np.abs(y - (y0 + jj * dy))) > rtol * scale
            or dx <= 0 or dy <= 0):
        raise ValueError("pillars must lie on a regular lattice with increasing x and y")
    # ZCORN: per layer k, top then bottom surface; per surface, per row j the
    # near (y-) and far (y+) corner rows; per row, per cell i its two corners
    z = zcorn.reshape(nz, 2, ny, 2, nx, 2)
    if nz > 1 and np.any(z[:-1, 1] != z[1:, 0]):
        raise ValueError("gaps between layers are not supported")
    levels = np.concatenate([z[:, 0], z[-1:, 1]], axis=0)              # [nz+1, ny, 2, nx, 2]
    zc = np.empty((ny, nx, 4, nz + 1))
    for c, (cx, cy) in enumerate(((0, 0), (1, 0), (0, 1), (1, 1))):
        zc[:, :, c, :] = np.transpose(levels[:, :, cy, :, cx], (1, 2, 0))
    return grids.corner_point_from_depths(nx, ny, nz, dx, dy, zc, origin=(x0, y0))


def _classify(deck):
    """classifyPermeability (ReservoirPropertyCommon_impl.hpp:70-112)."""
    has = {k: (k in deck) for k in ("PERMX", "PERMXY", "PERMXZ", "PERMYX", "PERMY", "PERMYZ",
                                    "PERMZX", "PERMZY", "PERMZ")}
    xx, xy, xz = has["PERMX"], has["PERMXY"], has["PERMXZ"]
    yx, yy, yz = has["PERMYX"], has["PERMY"], has["PERMYZ"]
    zx, zy, zz = has["PERMZX"], has["PERMZY"], has["PERMZ"]
    ncross = xy + xz + yx + yz + zx + zy
    ncomp = xx + yy + zz + ncross
    kind = "none"
    if ncross > 0:
        kind = "tensor"
    elif ncomp == 1:
        kind = "scalar"
    elif ncomp >= 2:
        kind = "diagonal"
    ok = True
    if ncomp > 0:
        ok = xx or not (xy or xz or yx or zx)
        ok = ok and (yy or not (yx or yz or xy or zy))
        ok = ok and (zz or not (zx or zy or xz or yz))
    return kind if ok else "invalid"


def _fill_tensor(deck):
    """fillTensor (:178-244): kmap[i] indexes the list of component arrays
    (0: zero) for tensor entry i (row-major); cross terms are symmetric and a
    diagonal keyword fills the other diagonal entries not yet set."""
    kind = _classify(deck)
    if kind == "invalid":
        raise RuntimeError("Invalid set of permeability fields given.")
    XX, XY, XZ, YX, YY, YZ, ZX, ZY, ZZ = range(9)
    kmap = [0] * 9
    comps = [None]

    def scalar_if_needed(i, j, k):
        if kmap[j] == 0:
            kmap[j] = kmap[i]
        if kmap[k] == 0:
            kmap[k] = kmap[i]

    def push(kw):
        comps.append(deck[kw] * MILLIDARCY)
        return len(comps) - 1

    if "PERMX" in deck:
        kmap[XX] = push("PERMX")
        scalar_if_needed(XX, YY, ZZ)
    if "PERMXY" in deck:
        kmap[XY] = kmap[YX] = push("PERMXY")
    if "PERMXZ" in deck:
        kmap[XZ] = kmap[ZX] = push("PERMXZ")
    if "PERMYX" in deck:
        kmap[YX] = kmap[XY] = push("PERMYX")
    if "PERMY" in deck:
        kmap[YY] = push("PERMY")
        scalar_if_needed(YY, ZZ, XX)
    if "PERMYZ" in deck:
        kmap[YZ] = kmap[ZY] = push("PERMYZ")
    if "PERMZX" in deck:
        kmap[ZX] = kmap[XZ] = push("PERMZX")
    if "PERMZY" in deck:
        kmap[ZY] = kmap[YZ] = push("PERMZY")
    if "PERMZ" in deck:
        kmap[ZZ] = push("PERMZ")
        scalar_if_needed(ZZ, XX, YY)
    return kind, kmap, comps


def rock_from_deck(deck: dict, global_cell=None, perm_threshold: float = 0.0):
    """(perm [nc, 9] row-major in m^2, porosity [nc], kind) as
    assignPermeability (:682-736) and assignPorosity (:592-611): entries by
    global cell, the diagonal raised to perm_threshold; without permeability
    keywords the tensors stay zero; without PORO the porosity is 1."""
    nx, ny, nz = deck_dims(deck)
    nglob = nx * ny * nz
    gc = np.arange(nglob) if global_cell is None else np.asarray(global_cell, dtype=np.int64)
    kind, kmap, comps = _fill_tensor(deck)
    for a in comps[1:]:
        if a.size != nglob:
            raise RuntimeError("All permeability fields must have the same size as the logical cartesian "
                               f"size of the grid: {a.size} != {nglob}")
    perm = np.zeros((gc.size, 9))
    if len(comps) > 1:
        comps[0] = np.zeros(nglob)
        for e in range(9):
            perm[:, e] = comps[kmap[e]][gc]
        for d in (0, 4, 8):
            perm[:, d] = np.maximum(perm[:, d], perm_threshold)
    poro = np.ones(gc.size)
    if "PORO" in deck:
        if deck["PORO"].size != nglob:
            raise RuntimeError("PORO field must have the same size as the logical cartesian size of the grid: "
                               f"{deck['PORO'].size} != {nglob}")
        poro = deck["PORO"][gc].astype(np.float64)
    return perm, poro, kind


def _fmt(a, per_line=6):
    a = np.asarray(a).reshape(-1)
    if a.dtype.kind in "iu":
        items = [str(int(v)) for v in a]
    else:
        items = [repr(float(v)) for v in a]
    return "\n".join(" ".join(items[i:i + per_line]) for i in range(0, len(items), per_line))


def write_grdecl(nx, ny, nz, dx, dy, zc, origin=(0.0, 0.0), props=None) -> str:
    """GRDECL text of the corner-point grid described as in
    corner_point_from_depths (zc[j, i, c, l]), with optional property
    keywords {name: values in deck units}."""
    zc = np.asarray(zc, dtype=np.float64)
    ii, jj = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1))
    x = origin[0] + ii * dx
    y = origin[1] + jj * dy
    top, bot = float(zc.min()), float(zc.max())
    coord = np.stack([x, y, np.full_like(x, top), x, y, np.full_like(x, bot)], axis=-1)
    z = np.empty((nz, 2, ny, 2, nx, 2))
    for c, (cx, cy) in enumerate(((0, 0), (1, 0), (0, 1), (1, 1))):
        lv = np.transpose(zc[:, :, c, :], (2, 0, 1))                    # [nz+1, ny, nx]
        z[:, 0, :, cy, :, cx] = lv[:-1]
        z[:, 1, :, cy, :, cx] = lv[1:]
    out = ["-- corner-point grid", "SPECGRID", f"{nx} {ny} {nz} 1 F /", "COORD", _fmt(coord), "/",
           "ZCORN", _fmt(z), "/"]
    for name, vals in (props or {}).items():
        out += [name, _fmt(vals), "/"]
    return "\n".join(out) + "\n"

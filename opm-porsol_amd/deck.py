"""ECLIPSE grid and rock input: the ECLIPSE branch of the reference's set-up
(setupGridAndProps, opm/porsol/common/setupGridAndProps.hpp:78-117) and
ReservoirPropertyCommon::init(deck, global_cell, perm_threshold, ...)
(opm/porsol/common/ReservoirPropertyCommon_impl.hpp:275-300).

- ``read_grdecl``: keyword / value records of a GRDECL or DATA file (values
  ended by ``/``, ``n*v`` repeats, ``--`` comments; keywords the grid and
  rock do not need are skipped).
- ``grid_from_deck``: SPECGRID (or DIMENS), COORD and ZCORN into the solver's
  grid.  The geometry is exact for vertical pillars on a regular lattice with
  gap-free layers (faults between columns allowed); other decks and
  inactive cells (ACTNUM 0) are rejected with a ValueError.  FIELD lengths
  are converted from feet.  The
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

SECTION_KEYWORDS = {"RUNSPEC", "GRID", "EDIT", "PROPS", "REGIONS", "SOLUTION", "SUMMARY", "SCHEDULE"}
INT_KEYWORDS = {"SPECGRID", "DIMENS", "ACTNUM"}
# the arrays the grid and the rock need; every other keyword's data is skipped
ARRAY_KEYWORDS = {"SPECGRID", "DIMENS", "COORD", "ZCORN", "ACTNUM", "PORO", "PERMX", "PERMY", "PERMZ",
                  "PERMXY", "PERMXZ", "PERMYX", "PERMYZ", "PERMZX", "PERMZY"}
FOOT = 0.3048
_KW = re.compile(r"[A-Z][A-Z0-9_]{0,7}")


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


def _keyword_line(toks):
    """A keyword stands alone on its line; a section keyword may carry a
    banner after it (RUNSPEC ====) and an array keyword the first of its
    values."""
    if not toks or not _KW.fullmatch(toks[0]):
        return False
    return len(toks) == 1 or toks[0] in SECTION_KEYWORDS or toks[0] in ARRAY_KEYWORDS


def read_grdecl(text: str) -> dict:
    """Keyword records of a GRDECL or DATA text: {keyword: numpy array} (raw
    values, deck units) for the arrays of ARRAY_KEYWORDS; the unit keyword
    (METRIC, FIELD or LAB) maps to an empty array.  Other keywords (dataless
    ones such as OIL or NONNC, TITLE with its text line, tables, RUNSPEC
    records) are skipped up to the next keyword line; reading stops at the
    first SOLUTION, SUMMARY, SCHEDULE or REGIONS section.  LAB units are
    rejected."""
    deck = {}
    lines = [ln.split("--", 1)[0] for ln in text.splitlines()]
    li = 0
    while li < len(lines):
        toks = lines[li].split()
        if not toks:
            li += 1
            continue
        if not _keyword_line(toks):
            raise ValueError(f"read_grdecl: expected a keyword, got '{toks[0]}' (line {li + 1})")
        kw = toks[0]
        first = lines[li].split(None, 1)[1] if len(toks) > 1 and kw in ARRAY_KEYWORDS else ""
        li += 1
        if kw in SECTION_KEYWORDS:
            deck[kw] = np.zeros(0)
        if kw in ("SOLUTION", "SUMMARY", "SCHEDULE", "REGIONS"):
            break
        if kw in ("METRIC", "FIELD", "LAB"):
            if kw == "LAB":
                raise ValueError("read_grdecl: only METRIC and FIELD units are supported")
            deck[kw] = np.zeros(0)
            continue
        if kw == "TITLE":
            li += 1
            continue
        if kw in ARRAY_KEYWORDS:
            parts, done = [], False
            if first:
                if "/" in first:
                    first = first.split("/", 1)[0]
                    done = True
                parts.append(first)
            while not done and li < len(lines):
                line = lines[li]
                if "/" in line:
                    line = line.split("/", 1)[0]
                    done = True
                parts.append(line)
                li += 1
            if not done:
                raise ValueError(f"read_grdecl: keyword {kw} is not terminated by '/'")
            rest = " ".join(parts)
            if kw == "SPECGRID":
                deck[kw] = np.array(rest.split()[:3], dtype=np.int64)     # NX NY NZ (then NUMRES, radial flag)
            elif kw in INT_KEYWORDS:
                deck[kw] = _parse(rest, kw, np.int64)
            else:
                deck[kw] = _parse(rest, kw, np.float64)
            continue
        while li < len(lines) and not _keyword_line(lines[li].split()):
            li += 1
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
    if "FIELD" in deck:                      # lengths in feet
        coord, zcorn = coord * FOOT, zcorn * FOOT
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


# --- the same input through the library's C-ABI (include/eclipse_grid.h) ------------------------
EXPORTS = ("mhs_deck_read_file", "mhs_deck_read_text", "mhs_deck_destroy", "mhs_deck_last_error", "mhs_deck_dims",
           "mhs_deck_build_grid", "mhs_deck_grid_arrays", "mhs_deck_rock")


class NativeDeck:
    """A deck read by libmhs (mhs_deck_*), the reader the C++ drop-in uses."""

    def __init__(self, path: str | None = None, text: str | None = None):
        import ctypes
        from . import mhs
        L = mhs.lib()
        self._L, self._c = L, ctypes
        P = ctypes.c_void_p
        L.mhs_deck_read_file.argtypes = [ctypes.POINTER(P), ctypes.c_char_p]
        L.mhs_deck_read_text.argtypes = [ctypes.POINTER(P), ctypes.c_char_p]
        L.mhs_deck_destroy.argtypes = [P]
        L.mhs_deck_last_error.argtypes = [P]
        L.mhs_deck_last_error.restype = ctypes.c_char_p
        L.mhs_deck_dims.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
        L.mhs_deck_build_grid.argtypes = [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.mhs_deck_grid_arrays.argtypes = [P] + [P] * 10
        L.mhs_deck_rock.argtypes = [P, ctypes.c_int, P, ctypes.c_double, P, P, ctypes.POINTER(ctypes.c_int)]
        self._h = P()
        if path is not None:
            rc = L.mhs_deck_read_file(ctypes.byref(self._h), path.encode())
        else:
            rc = L.mhs_deck_read_text(ctypes.byref(self._h), text.encode())
        self._check(rc)

    def _check(self, rc):
        if rc != 0:
            msg = self._L.mhs_deck_last_error(self._h).decode() if self._h else "deck read failed"
            raise ValueError(msg)

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.mhs_deck_destroy(self._h)
            self._h = None

    def dims(self):
        d = (self._c.c_int * 3)()
        self._check(self._L.mhs_deck_dims(self._h, d))
        return tuple(d)

    def grid(self):
        """(grids.Grid, corners [nc, 8, 3]) of the deck."""
        c = self._c
        nc, nhf = c.c_int(), c.c_int()
        self._check(self._L.mhs_deck_build_grid(self._h, c.byref(nc), c.byref(nhf)))
        nc, nhf = nc.value, nhf.value
        a = {"facepos": np.empty(nc + 1, np.int32), "nb": np.empty(nhf, np.int32), "bid": np.empty(nhf, np.int32),
             "cvol": np.empty(nc), "ccen": np.empty((nc, 3)), "farea": np.empty(nhf), "fcen": np.empty((nhf, 3)),
             "fnrm": np.empty((nhf, 3)), "ijk": np.empty((nc, 3), np.int32), "corners": np.empty((nc, 8, 3))}
        ptr = lambda x: x.ctypes.data_as(c.c_void_p)
        self._check(self._L.mhs_deck_grid_arrays(self._h, *(ptr(a[k]) for k in (
            "facepos", "nb", "bid", "cvol", "ccen", "farea", "fcen", "fnrm", "ijk", "corners"))))
        side = np.where(a["bid"] == 0, -1, np.where(a["bid"] <= 6, a["bid"] - 1, 6)).astype(np.int8)
        g = grids.Grid(facepos=a["facepos"], nb=a["nb"], bid=a["bid"], cell_volume=a["cvol"], cell_centroid=a["ccen"],
                       face_area=a["farea"], face_centroid=a["fcen"], face_normal=a["fnrm"], side=side, dims=None,
                       ijk=a["ijk"])
        return g, a["corners"]

    def rock(self, global_cell=None, perm_threshold: float = 0.0):
        """(perm [nc, 9] in m^2, porosity [nc], kind) as rock_from_deck."""
        c = self._c
        nx, ny, nz = self.dims()
        gc = np.arange(nx * ny * nz, dtype=np.int32) if global_cell is None else np.asarray(global_cell, np.int32)
        perm = np.empty((gc.size, 9))
        poro = np.empty(gc.size)
        kind = c.c_int()
        self._check(self._L.mhs_deck_rock(self._h, gc.size, gc.ctypes.data_as(c.c_void_p), perm_threshold,
                                          perm.ctypes.data_as(c.c_void_p), poro.ctypes.data_as(c.c_void_p),
                                          c.byref(kind)))
        return perm, poro, ("none", "scalar", "diagonal", "tensor")[kind.value]

"""Batched single-phase periodic upscaling (effective permeability).

Each cell problem is one instance of a batch.  Its boundaries are fully
periodic, as in the reference's periodic set-up (setupBoundaryConditions.hpp:
136-155 with bct = 2, createPeriodic in PeriodicHelpers.hpp:54-214).  It is
solved for three unit pressure drops, along x, y and z, as three right-hand
sides against one matrix.

Effective permeability follows from Darcy's law on cell averages.  With a drop
dp across length L_j in direction j, the mean flux in direction i is
u_i = Q_i / A_i, where Q_i is the net flux through the + side of direction i
and A_i is that side's area.  Then K_ij = u_i L_j / (lambda dp).  The formula
belongs to the upscaling codes built on opm-porsol, which are not in the
reference tree.  The tests pin it by closed forms: a homogeneous tensor
upscales to itself, and layered media to the harmonic and arithmetic means.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import fields, grids


@dataclasses.dataclass
class PeriodicBatch:
    grid: grids.Grid                 # disjoint union of the cell problems
    inst_cellpos: np.ndarray         # [ninst+1]
    partner: np.ndarray              # [nhf] periodic partner half-face
    bc_type: np.ndarray              # [nhf] 2 (periodic) on every boundary half-face
    group_ptr: np.ndarray            # [3 ninst + 1] half-faces on the + side of each direction
    group_hf: np.ndarray
    length: np.ndarray               # [ninst, 3] domain lengths
    area: np.ndarray                 # [ninst, 3] areas of the + sides
    dims: list

    @property
    def ninst(self) -> int:
        return len(self.dims)


def periodic_batch(dims_list, spacing=(1.0, 1.0, 1.0)) -> PeriodicBatch:
    """Box cell problems of the given (nx, ny, nz), fully periodic, unit-spaced cells
    by default."""
    dx, dy, dz = spacing
    gs, parts = [], []
    for d in dims_list:
        g = grids.cartesian(*d, dx, dy, dz)
        gs.append(g)
        parts.append(grids.periodic_partners_cartesian(g, (0, 1, 2)))
    G, icp = grids.concat(gs)
    hf_off = np.cumsum([0] + [g.num_half_faces for g in gs])
    partner = np.concatenate([np.where(p >= 0, p + hf_off[t], -1) for t, p in enumerate(parts)])
    bt = np.where(G.nb < 0, 2, 0).astype(np.int32)
    ptr, hfs = [0], []
    for t, g in enumerate(gs):
        for i in range(3):
            h = np.nonzero(g.side == 2 * i + 1)[0] + hf_off[t]
            hfs.append(h)
            ptr.append(ptr[-1] + h.shape[0])
    length = np.array([[d[0] * dx, d[1] * dy, d[2] * dz] for d in dims_list], dtype=np.float64)
    area = np.stack([length[:, 1] * length[:, 2], length[:, 0] * length[:, 2],
                     length[:, 0] * length[:, 1]], axis=1)
    return PeriodicBatch(G, icp, partner.astype(np.int32), bt, np.asarray(ptr, np.int32),
                         np.concatenate(hfs).astype(np.int32), length, area, list(dims_list))


def pressure_drop_values(batch: PeriodicBatch, dp: float = 1.0) -> np.ndarray:
    """bc_value[3, nhf]: right-hand side j has pressure difference +dp on the
    - side of direction j and -dp on its + side (pressure higher on the - side,
    flow along +j), zero across the other periodic pairs."""
    g = batch.grid
    bv = np.zeros((3, g.num_half_faces))
    for j in range(3):
        bv[j, g.side == 2 * j] = dp
        bv[j, g.side == 2 * j + 1] = -dp
    return bv


def effective_permeability(flux_sums: np.ndarray, batch: PeriodicBatch, mobility: float = 1.0,
                           dp: float = 1.0) -> np.ndarray:
    """K_eff[ninst, 3, 3] from flux_sums[3 (drops), 3 ninst (+ sides)]
    (HybridSolver.sum_outflux with batch.group_ptr / group_hf)."""
    n = batch.ninst
    Q = flux_sums.reshape(3, n, 3)          # [j, t, i]
    K = np.empty((n, 3, 3))
    for j in range(3):
        K[:, :, j] = Q[j] / batch.area * batch.length[:, j:j + 1] / (mobility * dp)
    return K


def lognormal_isotropic(dims_list, seed: int, sigma: float = 1.0,
                        mean_perm: float = 100.0 * 9.869233e-16) -> np.ndarray:
    """perm[nc, 9] for a batch: i.i.d. ln k ~ N(ln mean_perm, sigma) per cell, its
    own counter-based stream per instance."""
    out = []
    for t, d in enumerate(dims_list):
        n = int(np.prod(d))
        k = np.exp(np.log(mean_perm) + sigma * fields.normal(seed, t, n))
        out.append(fields.diag_perm(k, k, k))
    return np.concatenate(out)


def random_sizes(ninst: int, lo: int, hi: int, seed: int):
    """Cubic cell-problem sizes n^3 with n uniform on [lo, hi] (seeded)."""
    u = fields.uniform(seed, 7, ninst)
    n = np.minimum(lo + np.floor(u * (hi - lo + 1)).astype(np.int64), hi)
    return [(int(v), int(v), int(v)) for v in n]

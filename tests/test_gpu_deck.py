"""A GRDECL deck (faulted corner-point grid, full permeability tensors,
porosity) read by opm-porsol_amd/deck.py and solved on the GPU through the
C-ABI, against the oracle on the same grid and rock (1e-8)."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def test_deck_to_gpu_solve(pkg, orc):
    nx, ny, nz = 12, 10, 5
    zc = pkg.grids.corner_point(nx, ny, nz, 20.0, 20.0, 2.0, seed=9, fault_frac=0.2, return_depths=True)
    rng = np.random.default_rng(9)
    n = nx * ny * nz
    props = {"PERMX": rng.lognormal(np.log(200.0), 0.8, n), "PERMY": rng.lognormal(np.log(100.0), 0.8, n),
             "PERMZ": rng.lognormal(np.log(20.0), 0.8, n), "PERMXY": rng.uniform(-5.0, 5.0, n),
             "PORO": rng.uniform(0.1, 0.3, n)}
    deck = pkg.deck.read_grdecl(pkg.deck.write_grdecl(nx, ny, nz, 20.0, 20.0, zc, props=props))
    g = pkg.deck.grid_from_deck(deck)
    perm, poro, kind = pkg.deck.rock_from_deck(deck)
    assert kind == "tensor" and np.all(poro == props["PORO"])
    bt, bv = cases.xdrop_bc(g, 2e5, 1e5)
    nc = g.num_cells
    s = pkg.HybridSolver(0)
    s.init(g, perm, (0.0, 0.0, 9.81), bt, None)
    tm, rm = np.full(nc, 1e3), np.full(nc, 1e3 * 1000.0)
    s.solve(tm, rm, np.zeros(nc), bt, bv, residual_tolerance=1e-12, precision="mixed")
    ref = orc.solve(g, perm, tm, rm, np.zeros(nc), bt, bv, grav=(0.0, 0.0, 9.81), tol=1e-14)
    assert cases.rel_err(s.getSolution()[2], ref.outflux) < 1e-8

"""Writes the committed fixtures of tests/golden/ with the FP64 oracle.

    python tests/golden/make_golden.py

ip_attic_diag10_1_01.npz  B^-1 (MimeticIPEvaluator layout, column-major per
                          cell) for K = diag(10, 1, 0.1) on 1^3, 2^3, 4^3 unit
                          cubes, the cases printed by the reference's
                          attic/mimetic/mimetic_ipeval_test.cpp
c1_20x20x10.npz           the 20x20x10 plain case (ln k ~ N(ln 100 mD, 1),
                          seed 1, x pressure drop 1 -> 0): face and cell
                          pressures and half-face fluxes at 1e-14
periodic_4x4x4.npz        4^3 fully periodic, unit pressure drop along x
known_answer.json         the error norms the reference records in
                          examples/timing_results.txt
c2_128x128x64.npz         the 128x128x64 headline case (anisotropic Gaussian-
                          filtered K, seed 2, 1 bar -> 0 along x) solved by the
                          oracle's ILU0-CG to a 1e-14 relative residual: face
                          and cell pressures and half-face fluxes at 32768
                          seeded sample positions each, whole-field max, 2-norm
                          and sum, the flow rate through every x-plane, and the
                          ILU0-CG iteration counts to 1e-10 and 1e-14

    python tests/golden/make_golden.py c2     (only the headline fixture, ~5 min)
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(HERE))

import __graft_entry__  # noqa: E402

pkg = __graft_entry__.load_package()
import oracle  # noqa: E402
import cases  # noqa: E402


def make_c2():
    g, perm, bt, bv = cases.c2_problem()
    nc = g.num_cells
    tm, rm, src = np.ones(nc), np.zeros(nc), np.zeros(nc)
    r = oracle.solve(g, perm, tm, rm, src, bt, bv, tol=1e-14, precond=2)
    res = r.system.rhs - oracle.spmv(r.system, r.face_pressure)
    true_res = float(np.linalg.norm(res) / np.linalg.norm(r.system.rhs))
    r10 = oracle.solve(g, perm, tm, rm, src, bt, bv, tol=1e-10, precond=2)
    out = {"perm_checksum": pkg.fields.checksum(perm), "ilu0_iterations_1e14": r.iterations,
           "ilu0_iterations_1e10": r10.iterations, "true_residual": true_res,
           "xplane_rates": cases.xplane_rates(g, r.outflux)}
    for k, (name, a) in enumerate((("face_pressure", r.face_pressure), ("cell_pressure", r.cell_pressure),
                                   ("outflux", r.outflux))):
        idx = cases.c2_sample_index(a.shape[0], k)
        out[name + "_sample"] = a[idx]
        out[name + "_norms"] = np.array([np.max(np.abs(a)), np.linalg.norm(a), np.sum(a)])
        e10 = np.max(np.abs(getattr(r10, name) - a)) / np.max(np.abs(a))
        out[name + "_ilu0_1e10_err"] = e10
        print(f"{name}: ILU0-CG at 1e-10 differs from 1e-14 by {e10:.2e} (max-norm relative)")
    print(f"ILU0-CG iterations: {r10.iterations} to 1e-10, {r.iterations} to 1e-14; true residual {true_res:.2e}")
    np.savez_compressed(os.path.join(HERE, "c2_128x128x64.npz"), **out)


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "c2":
        make_c2()
        return
    out = {}
    for n in (1, 2, 4):
        g = pkg.grids.cartesian(n, n, n, 1.0 / n, 1.0 / n, 1.0 / n)
        perm = np.tile(np.diag([10.0, 1.0, 0.1]).reshape(1, 9), (g.num_cells, 1))
        out[f"binv_{n}"] = oracle.static_ip(g, perm)[0]
    np.savez_compressed(os.path.join(HERE, "ip_attic_diag10_1_01.npz"), **out)

    g, perm, tm, rm, src, bt, bv = cases.case_c1()
    r = oracle.solve(g, perm, tm, rm, src, bt, bv, tol=1e-14)
    np.savez_compressed(os.path.join(HERE, "c1_20x20x10.npz"), perm_checksum=pkg.fields.checksum(perm),
                        hf_dof=r.hf_dof, face_pressure=r.face_pressure,
                        cell_pressure=r.cell_pressure, outflux=r.outflux)

    g = pkg.grids.cartesian(4, 4, 4)
    perm = pkg.fields.perm_c1(g.num_cells, seed=7)
    bt, bv, partner = cases.periodic_bc(g, (0, 1, 2), drop_dir=0, pdrop=1.0)
    tm, rm = cases.unit_fluid(g.num_cells)
    r = oracle.solve(g, perm, tm, rm, np.zeros(g.num_cells), bt, bv, partner=partner, tol=1e-14)
    np.savez_compressed(os.path.join(HERE, "periodic_4x4x4.npz"), perm_checksum=pkg.fields.checksum(perm),
                        face_pressure=r.face_pressure, cell_pressure=r.cell_pressure,
                        outflux=r.outflux)

    known = {
        "source": "examples/timing_results.txt",
        "problem": "unit cube N^3, K = 1, unit viscosities, u = sin(2 pi x) cos(2 pi y) z, "
                   "Dirichlet u(face centroid), src = -Lu vol (examples/known_answer_test.cpp)",
        "50": {"L1": 1.83583e-4, "L2": 2.57007e-4, "Linf": 8.57712e-4, "lines": "14-16"},
        "60_yaspgrid": {"L1": 1.27476e-4, "L2": 1.78521e-4, "Linf": 5.94968e-4, "lines": "35-37"},
        "60_cpgrid": {"L1": 1.27477e-4, "L2": 1.78521e-4, "Linf": 5.9499e-4, "lines": "54-56"},
    }
    with open(os.path.join(HERE, "known_answer.json"), "w") as f:
        json.dump(known, f, indent=1)
    make_c2()


if __name__ == "__main__":
    main()

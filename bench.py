#!/usr/bin/env python3
"""Face-pressure solves/sec and HBM GB/s of the mimetic hybrid solve at 128x128x64.

One step = one IncompFlowSolverHybrid::solve(): dynamic assembly, PCG to a
relative residual of 1e-10 and flux / cell-pressure recovery, on the 128x128x64
Cartesian grid with anisotropic Gaussian-filtered permeability (kz/kx = 0.1,
ln kx ~ ln 100 mD + 1.5 z, correlation length 8 cells), Dirichlet pressure
drop along x, no-flow elsewhere.  Inputs (mobilities, sources, boundary values)
are resident in HBM before the timed region.

The timed solve is the default solver type (multigrid-preconditioned CG, as the
reference's default AMG-CG) in mixed precision.  Beside it, once each: the same
solve in FP64, and the plain method in FP64 (Jacobi-PCG, no multigrid).

Prints one JSON line.  --gpus N > 1 is rejected: the solver runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

DIMS = (128, 128, 64)
TOL = 1e-10


def _problem(pkg, seed):
    g = pkg.grids.cartesian(*DIMS)
    perm = pkg.fields.perm_c2(DIMS, seed=seed)
    nhf = g.num_half_faces
    bt = np.zeros(nhf, np.int32)
    bv = np.zeros(nhf)
    bt[g.side == 0] = 1
    bv[g.side == 0] = 1.0e5          # 1 bar on x-
    bt[g.side == 1] = 1
    bv[g.side == 1] = 0.0
    return g, perm, bt, bv


def _vcycle_kernels(nf, v, lev_rows, lev_nnz, p_nnz, t_nnz, ap_nnz, sa_level):
    """Algorithmic bytes of the kernels of one multigrid V(1,1) cycle outside
    the level-0 operator pass (the pre-smoothing residual is counted by the
    caller), as [(kernel, bytes)]: each array once per kernel at its stored
    precision v, 4-byte indices, gathered vectors once.  lev_rows / lev_nnz
    hold levels 0 (faces) .. L-1; sa_level is the level whose transfer to the
    next is smoothed (-1: none)."""
    L = len(lev_rows)
    out = []
    n1 = lev_rows[1]
    # faces -> level 1 (children lists): child pointers, children, res0, b1
    out.append(("k_restrict (faces -> level 1)", nf * (4 + v) + n1 * (4 + v)))
    for l in range(1, L - 1):
        n, nnz, nc = lev_rows[l], lev_nnz[l], lev_rows[l + 1]
        # down-sweep from zero: row pointers, matrix, b, D^-1, x and res written
        out.append((f"down-sweep level {l}", 8 * nnz + n * (4 + 4 * v)))
        if l == sa_level:
            # P' rows (pointers, columns, values), res gathered, b_c written
            out.append((f"restriction P' level {l}", 8 * p_nnz + nc * (4 + v) + n * v))
            # up-sweep across the smoothed transfer: the rows of M = P - w D^-1 T
            # (T = A P's pattern), b, res, D^-1 read, x written, x_c gathered
            out.append((f"up-sweep (M = P - w D^-1 A P) level {l}", (4 + v) * t_nnz + n * (4 + 4 * v) + nc * v))
        else:
            out.append((f"restriction level {l}", n * (4 + v) + nc * (4 + v)))
            # up-sweep: matrix, aggregate map, b, D^-1 read, x written, x_c gathered
            out.append((f"up-sweep level {l}", 8 * nnz + n * (8 + 3 * v) + nc * v))
    m = lev_rows[L - 1]
    out.append(("coarsest dense solve", m * m * v + 2 * m * v))
    # post-smoothing on level 0 with A P: its SELL entries, the aggregate map,
    # x0, res0, r, D^-1 read, z written, x_1 gathered
    out.append(("k_vc0_post", ap_nnz * (v + 4) + nf * (4 + 5 * v) + n1 * v))
    return out


def _alg_bytes(g, nnz, nf, rep=None):
    """Algorithmic HBM bytes per unit of work for one right-hand side: each
    array counted once per kernel at its stored precision, 4-byte indices,
    gathered vectors counted once.  v is the value size of the PCG (4 in
    mixed precision, 8 in FP64).  With the solve report of a multigrid solve,
    "amg_iter*" are the sums over the kernels of one iteration, listed per
    kernel in "kernels*"."""
    nc, nhf = g.num_cells, g.num_half_faces
    sq = int(np.sum(g.ncf().astype(np.int64) ** 2))
    pk = int(np.sum(g.ncf().astype(np.int64) * (g.ncf() + 1) // 2))
    nint = int(np.sum(g.nb >= 0)) // 2
    out = {}
    for v, tag in ((4, "32"), (8, "64")):
        mat = nnz * (v + 4)
        # SpMV q = A p with the fused <p,q> partials: matrix, p, q
        out["spmv" + tag] = mat + 2 * nf * v
        # pre-smoothing residual res = r - A x: matrix, x, r, res
        out["resid" + tag] = mat + 3 * nf * v
        # Jacobi-PCG iteration: SpMV + x,r,p,q,dinv reads and x,r,p writes
        out["iter" + tag] = mat + 4 * (nf + 1) + 13 * nf * v
    # The same two operator passes element by element (FP32): the packed local
    # matrices and the face rows of every cell, the gathered vector once, the
    # output written once and, on every face with two cells, read and written
    # again by the second colour.
    cell_part = pk * 4 + nhf * 4
    out["ebe_spmv"] = cell_part + nf * 4 + nf * 4 + nint * 8
    out["ebe_resid"] = cell_part + 2 * nf * 4 + nf * 4 + nint * 8
    if rep is not None and rep.get("precond") == 1:
        lev_rows = [int(x) for x in rep["level_rows"]]
        lev_nnz = [int(x) for x in rep["level_nnz"]]
        p_nnz, t_nnz = (int(x) for x in rep["transfer_nnz"])
        sa_level = int(rep["smoothed_level"])
        for v, tag in ((4, "32"), (8, "64")):
            # x += a p, r -= a q with the <r, D^-1 r>, |r|^2 partials and the
            # first smoothing step x0 = w D^-1 r: x, r, p, q, D^-1 read, x, r, x0 written
            upd_xr = ("k_update_xr", 8 * nf * v)
            upd_pz = ("k_update_pz", 3 * nf * v)        # p = z + b p
            vc = _vcycle_kernels(nf, v, lev_rows, lev_nnz, p_nnz, t_nnz, int(rep["ap_nnz"]), sa_level)
            ks = [("SpMV q = A p", out["spmv" + tag]), upd_xr, ("pre-smoothing residual", out["resid" + tag])]
            out["kernels" + tag] = ks + vc + [upd_pz]
            out["amg_iter" + tag] = sum(b for _, b in out["kernels" + tag])
            if tag == "32":
                ke = [("k_ebe product (2 colour launches)", out["ebe_spmv"]), upd_xr,
                      ("k_ebe pre-smoothing residual (2 colour launches)", out["ebe_resid"])]
                out["kernels_ebe"] = ke + vc + [upd_pz]
                out["amg_iter_ebe"] = sum(b for _, b in out["kernels_ebe"])
    # FP64 residual replacement: FP64 operator + x, b read, r written
    out["refine64"] = nnz * (8 + 4) + 4 * nf * 8
    # element by element it rides on the SpMV: the FP64 values instead of
    # the FP32 ones, x64 and b read, the FP64 sums of two-cell faces kept
    # between the colours, the FP32 residual written
    out["refine64_ebe"] = pk * 4 + 2 * nf * 8 + nint * 16 + nf * 4
    # assembly (static B^-1 read packed, values written in FP64 + FP32, rhs, F, L, g)
    out["assemble"] = pk * 8 + nc * 16 + nnz * (8 + 4) + nf * 8 + nhf * 8 + nc * 16 + sq
    # recovery: pi, F, L, g, B^-1 (packed), mobility, p, v
    out["recover"] = nf * 8 + nhf * 8 + nc * 16 + pk * 8 + nc * 8 + nc * 8 + nhf * 8
    out["pk"] = pk
    return out


def _measured_traffic(tag, name="k_spmv_pq_128x128x64_hbm_traffic.txt"):
    """HBM bytes per launch of the dominant kernel (the FP32 SpMV, or the
    transport small step) as measured by the FETCH_SIZE / WRITE_SIZE counter
    runs summarised in profiles/ (calibrated against a copy of known size);
    None when no such measurement is committed."""
    path = os.path.join(ROOT, "profiles", name)
    if tag != "32" or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def _cpu_baseline(pkg, g, perm, bt, bv, iters_needed, sample_iters):
    """FP64 C restatement (the oracle) on one host core: full assembly and
    recovery plus `sample_iters` Jacobi-PCG iterations, extrapolated to the
    iteration count the identical FP64 Jacobi-PCG needs (measured on the GPU)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    nc = g.num_cells
    tm, rm, src = np.ones(nc), np.zeros(nc), np.zeros(nc)
    hf, nint, nf, _ = oracle.enumerate_dof(g)
    pp, _ = oracle.periodic_dof(g, hf, nf, bt, np.full(g.num_half_faces, -1, np.int32))
    binv, pos, gfl = oracle.static_ip(g, perm)
    t0 = time.perf_counter()
    S = oracle.assemble(g, hf, nf, pp, binv, gfl, tm, rm, src, bt, bv)
    t1 = time.perf_counter()
    x, rc, its, red = oracle.cg(S, 1e-30, maxit=sample_iters, precond=1, dirichlet_guess=True)
    t2 = time.perf_counter()
    oracle.recover(g, hf, binv, gfl, tm, rm, S, x)
    t3 = time.perf_counter()
    per_it = (t2 - t1) / max(its, 1)
    t_solve = (t1 - t0) + (t3 - t2) + per_it * iters_needed
    out = {
        "value": 1.0 / t_solve, "unit": "solves/s", "cores": 1, "kind": "port",
        "sample": (f"oracle/oracle.c FP64 on 1 host core ({platform.processor() or 'cpu'}, "
                   f"{os.cpu_count()} visible): full assembly {t1 - t0:.2f} s + recovery "
                   f"{t3 - t2:.2f} s + {its} Jacobi-PCG iterations at {per_it * 1e3:.1f} ms, "
                   f"extrapolated to the {iters_needed} iterations the same FP64 Jacobi-PCG "
                   f"takes at {g.num_cells} cells"),
        "seconds_per_solve": t_solve,
    }
    # The same Jacobi-PCG sample with its SpMV, dot products and vector
    # updates spread over the host's cores (OpenMP); assembly and recovery
    # stay the single-core times above.
    threads = min(int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1), 16)
    if threads > 1:
        oracle.set_threads(threads)
        try:
            t4 = time.perf_counter()
            _, _, its_mt, _ = oracle.cg(S, 1e-30, maxit=sample_iters, precond=1, dirichlet_guess=True)
            per_it_mt = (time.perf_counter() - t4) / max(its_mt, 1)
        finally:
            oracle.set_threads(1)
        t_mt = (t1 - t0) + (t3 - t2) + per_it_mt * iters_needed
        out["openmp"] = {"value": 1.0 / t_mt, "unit": "solves/s", "cores": threads,
                         "sample": (f"{its_mt} Jacobi-PCG iterations at {per_it_mt * 1e3:.1f} ms on "
                                    f"{threads} threads, single-core assembly and recovery")}
    return out



def _forced_shape(args):
    """Launch shapes given on the command line (the profiling runs fix them so
    that every counter pass replays the same kernels); the autotuner's choice
    otherwise."""
    out = {}
    if args.spmv_threads:
        out["spmv_threads"] = args.spmv_threads
    if args.sweep_threads:
        out["sweep_threads"] = args.sweep_threads
    if args.tail_rows >= 0:
        out["tail_rows"] = args.tail_rows
    if args.element_operator >= 0:
        out["element_operator"] = args.element_operator
    return out


def _c2_fixture():
    """The oracle's solution of the headline case (tests/golden/c2_128x128x64.npz,
    written by tests/golden/make_golden.py with the oracle's ILU0-CG to 1e-14)."""
    return np.load(os.path.join(ROOT, "tests", "golden", "c2_128x128x64.npz"))


def _err_vs_oracle(g, perm, fx, p, pi, v):
    """Max-norm relative error of the GPU solution against the oracle's at the
    fixture's 32768 seeded sample positions per field (the same positions the
    parity test checks)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from cases import c2_sample_index
    import opm_porsol_amd
    if str(fx["perm_checksum"]) != opm_porsol_amd.fields.checksum(perm):
        return None
    out = {}
    for k, (name, a) in enumerate((("face_pressure", pi), ("cell_pressure", p), ("outflux", v))):
        idx = c2_sample_index(a.shape[0], k)
        out[name] = float(np.max(np.abs(a[idx] - fx[name + "_sample"])) / fx[name + "_norms"][0])
    return out


def _cpu_baseline_ilu0(pkg, g, perm, bt, bv, fx, sample_iters):
    """The oracle (oracle/oracle.c, FP64) with ILU0-CG, the reference's
    linsolver_type 0 (SeqILU0 + CGSolver, IncompFlowSolverHybrid.hpp:1387-1420),
    on the host: init (dof numbering, static inner products), full assembly,
    the ILU0 factorisation, `sample_iters` CG iterations and full recovery,
    timed; the iterations extrapolated to the count the oracle's complete
    ILU0-CG run needed to reach 1e-10 (recorded in the fixture).  One core,
    then the CG's SpMV, dot products and vector updates on all host cores
    (OpenMP; the ILU0 triangular solves stay sequential)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    nc = g.num_cells
    tm, rm, src = np.ones(nc), np.zeros(nc), np.zeros(nc)
    its_needed = int(fx["ilu0_iterations_1e10"])
    t0 = time.perf_counter()
    hf, nint, nf, _ = oracle.enumerate_dof(g)
    pp, _ = oracle.periodic_dof(g, hf, nf, bt, np.full(g.num_half_faces, -1, np.int32))
    binv, pos, gfl = oracle.static_ip(g, perm)
    t1 = time.perf_counter()
    S = oracle.assemble(g, hf, nf, pp, binv, gfl, tm, rm, src, bt, bv)
    t2 = time.perf_counter()

    def cg_time(n):
        t = time.perf_counter()
        x, rc, its, red = oracle.cg(S, 1e-30, maxit=n, precond=2)
        return time.perf_counter() - t, its, x

    t_one, _, _ = cg_time(1)                      # factorisation + one iteration
    t_n, its, x = cg_time(1 + sample_iters)
    per_it = (t_n - t_one) / max(its - 1, 1)
    t_fact = max(t_one - per_it, 0.0)
    t3 = time.perf_counter()
    oracle.recover(g, hf, binv, gfl, tm, rm, S, x)
    t_rec = time.perf_counter() - t3
    t_solve = (t2 - t1) + t_fact + its_needed * per_it + t_rec
    out = {
        "value": 1.0 / t_solve, "unit": "solves/s", "cores": 1, "kind": "port",
        "sample": (f"oracle/oracle.c FP64 ILU0-CG (reference linsolver_type 0) on 1 host core "
                   f"({platform.processor() or 'cpu'}, {os.cpu_count()} visible): assembly {t2 - t1:.2f} s, "
                   f"ILU0 factorisation {t_fact:.2f} s, {its - 1} CG iterations at {per_it * 1e3:.1f} ms "
                   f"extrapolated to the {its_needed} the oracle's full ILU0-CG run needs to reach 1e-10 "
                   f"(tests/golden/c2_128x128x64.npz), recovery {t_rec:.2f} s"),
        "seconds_per_solve": t_solve, "init_s": round(t1 - t0, 3), "iterations": its_needed,
    }
    threads = min(int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1), 16)
    if threads > 1:
        oracle.set_threads(threads)
        try:
            t_one_mt, _, _ = cg_time(1)
            t_n_mt, its_mt, _ = cg_time(1 + sample_iters)
        finally:
            oracle.set_threads(1)
        per_it_mt = (t_n_mt - t_one_mt) / max(its_mt - 1, 1)
        t_mt = (t2 - t1) + t_fact + its_needed * per_it_mt + t_rec
        out["openmp"] = {"value": 1.0 / t_mt, "unit": "solves/s", "cores": threads,
                         "sample": (f"{its_mt - 1} ILU0-CG iterations at {per_it_mt * 1e3:.1f} ms with the SpMV, "
                                    f"dot products and updates on {threads} threads, the triangular solves, "
                                    f"assembly and recovery on one core")}
    return out


def _copy_gbps(dev, mib=1024, reps=20):
    """Read + write bandwidth of a plain device-to-device copy of `mib` MiB
    (torch's copy kernel, timed with events on torch's current stream): the
    streaming rate HBM reaches in practice, beside the 8 TB/s peak."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return 2.0 * n * 4 / (ms * 1e-3) / 1e9


def _spmv_roofline(s, nnz, nf, nrhs, tag, g=None):
    """Roofline entry of the PCG product of the last solve (HIP-event timing):
    the element-by-element product when the solve used it, else the SpMV
    with the assembled face matrix."""
    v = 4 if tag == "32" else 8
    rep = s.report
    ms = s.time_kernel(0, reps=50)
    if rep.element_operator and g is not None:
        pk = int(np.sum(g.ncf().astype(np.int64) * (g.ncf() + 1) // 2))
        nint = int(np.sum(g.nb >= 0)) // 2
        alg = pk * 4 + g.num_half_faces * 4 + (2 * nf * 4 + nint * 8) * nrhs
        kname = f"k_ebe<nrhs={nrhs}> (element-by-element product, {rep.colours} colour launches)"
    else:
        alg = nnz * (v + 4) + 2 * nf * v * nrhs
        kname = (f"{'k_spmv_pq8' if (tag == '32' and nrhs == 1) else 'k_spmv_pq'}"
                 f"<{'float' if tag == '32' else 'double'}, nrhs={nrhs}> (face-matrix SpMV with fused <p,Ap>)")
    ach = alg / (ms * 1e-3) / 1e9
    return {"kernel": kname, "bound": "hbm", "achieved": round(ach, 1), "peak": 8000.0, "unit": "GB/s",
            "frac": round(ach / 8000.0, 4), "traffic": None,
            "alg_bytes_per_launch": alg, "avg_launch_us": round(ms * 1e3, 2)}


def bench_upscale(args, pkg, torch, dev):
    """256 fully periodic cell problems, n^3 with n uniform on [16, 48], i.i.d.
    lognormal isotropic K (sigma 1, mean ln 100 mD), three unit pressure drops
    as three right-hand sides: one step = one batched solve (768 cell-problem
    solves) plus the effective-permeability reduction."""
    from opm_porsol_amd import upscaling as U
    t0 = time.perf_counter()
    sizes = U.random_sizes(256, 16, 48, seed=args.seed)
    batch = U.periodic_batch(sizes)
    perm = U.lognormal_isotropic(sizes, seed=args.seed, sigma=1.0)
    t_gen = time.perf_counter() - t0
    g = batch.grid
    nc, nhf = g.num_cells, g.num_half_faces
    s = pkg.HybridSolver(0)
    t0 = time.perf_counter()
    s.init(g, perm, (0.0, 0.0, 0.0), batch.bc_type, batch.partner, inst_cellpos=batch.inst_cellpos)
    t_init = time.perf_counter() - t0
    tm = torch.ones(nc, dtype=torch.float64, device=dev)
    rm = torch.zeros(nc, dtype=torch.float64, device=dev)
    src = torch.zeros(3 * nc, dtype=torch.float64, device=dev)
    bvd = torch.from_numpy(U.pressure_drop_values(batch).reshape(-1)).to(dev)
    torch.cuda.synchronize()

    def step(precision):
        rep = s.solve_device(tm.data_ptr(), rm.data_ptr(), src.data_ptr(), batch.bc_type, bvd.data_ptr(),
                             nrhs=3, tol=TOL, precision=precision, chunk=args.chunk,
                             use_graph=not args.no_graph)
        return rep

    cmp = None
    if not args.no_compare:
        step("fp64")
        torch.cuda.synchronize()
        t = time.perf_counter()
        r64 = step("fp64").as_dict()
        torch.cuda.synchronize()
        cmp = (time.perf_counter() - t, r64,
               U.effective_permeability(s.sum_outflux(batch.group_ptr, batch.group_hf), batch))
    for _ in range(args.warmup):
        step(args.precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(args.precision)
        keff = U.effective_permeability(s.sum_outflux(batch.group_ptr, batch.group_hf), batch)
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) / args.steps * 1e3
    rep = s.report.as_dict()
    d = s.dims()
    nnz = int(s.system(0)[0][-1])
    tag = "32" if args.precision == "mixed" else "64"
    roof = _spmv_roofline(s, nnz, d.num_faces, 3, tag, g)
    mD = 9.869233e-16
    kd = np.stack([keff[:, i, i] for i in range(3)], 1) / mD
    config = {
        "workload": "256 fully periodic cell problems n^3, n~U[16,48], lognormal isotropic K "
                    "(sigma 1, mean ln 100 mD), 3 unit pressure drops, multigrid-PCG to 1e-10",
        "instances": batch.ninst, "cells": nc, "faces": d.num_faces, "half_faces": nhf, "nnz": nnz,
        "precision": args.precision, "iterations": rep["iterations"], "levels": rep["levels"],
        "fp64_replacements": rep["replacements"], "reduction": rep["reduction"],
        "init_s": round(t_init, 2), "generate_s": round(t_gen, 2),
        "keff_mD_diag_mean": [round(float(x), 3) for x in kd.mean(0)],
        "keff_asymmetry_max": float(np.max(np.abs(keff - keff.transpose(0, 2, 1))) / np.max(np.abs(keff))),
    }
    if cmp is not None:
        t64, r64, k64 = cmp
        config["fp64_gpu"] = {"cell_solves_per_s": round(768.0 / t64, 2), "iterations": r64["iterations"],
                              "speedup": round(t64 / (ms_step * 1e-3), 3)}
        config["keff_err_vs_fp64_gpu"] = float(np.max(np.abs(keff - k64)) / np.max(np.abs(k64)))
    cpu = None
    if not args.no_cpu_baseline:
        cpu = _cpu_baseline_upscale(pkg, sizes, perm, batch)
    print(json.dumps({
        "metric": "periodic upscaling cell-problem solves/sec (256 instances 16^3..48^3 x 3 drops), "
                  "rel. residual 1e-10",
        "value": round(768.0 / (ms_step * 1e-3), 3), "unit": "cell-problem solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 PCG / f64 residual, assembly, recovery" if args.precision == "mixed" else "f64",
        "data": "synthetic (seeded i.i.d. lognormal permeability)", "config": config,
        "roofline": roof, "cpu_baseline": cpu}))


def _cpu_baseline_upscale(pkg, sizes, perm, batch, sample=2):
    """The oracle (FP64, Jacobi-PCG, one core) on the first `sample` cell
    problems of the batch, all three drops each."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    t_tot, n_solved, cells = 0.0, 0, 0
    for t in range(sample):
        c0, c1 = batch.inst_cellpos[t], batch.inst_cellpos[t + 1]
        gt = pkg.grids.cartesian(*sizes[t])
        pp = pkg.grids.periodic_partners_cartesian(gt, (0, 1, 2))
        bt = np.where(gt.nb < 0, 2, 0).astype(np.int32)
        n = c1 - c0
        for j in range(3):
            bv = np.zeros(gt.num_half_faces)
            bv[gt.side == 2 * j] = 1.0
            bv[gt.side == 2 * j + 1] = -1.0
            t0 = time.perf_counter()
            oracle.solve(gt, perm[c0:c1], np.ones(n), np.zeros(n), np.zeros(n), bt, bv, partner=pp,
                         tol=TOL, precond=1)
            t_tot += time.perf_counter() - t0
            n_solved += 1
        cells += n
    # scale by cells: the batch's mean problem is larger than the sample's
    mean_cells = batch.grid.num_cells / batch.ninst
    per_solve = t_tot / n_solved * (mean_cells / (cells / sample))
    return {"value": 1.0 / per_solve, "unit": "cell-problem solves/s", "cores": 1, "kind": "port",
            "sample": (f"oracle/oracle.c FP64 Jacobi-PCG on 1 host core: cell problems "
                       f"{[sizes[t][0] for t in range(sample)]}^3, 3 drops each, {t_tot:.2f} s; "
                       f"scaled linearly in cells to the batch mean of {mean_cells:.0f} cells "
                       f"(iteration growth with size not included, so this favours the CPU)")}


def bench_adjoint(args, pkg, torch, dev):
    """Gradients dJ/dln K of 32 realizations of a 64x64x32 grid (lognormal
    isotropic K, correlation length 6 cells, sigma 1), J = sum over 16
    producer faces per realization of (flux - target)^2, targets from a seeded
    reference field; x pressure drop 1 bar.  One step = the batched forward
    solve plus the adjoint solve and gradient kernels."""
    dims, R = (64, 64, 32), 32
    mD, bar = 9.869233e-16, 1.0e5
    grids_, perms, refs = [], [], []
    for r in range(R):
        grids_.append(pkg.grids.cartesian(*dims))
        lnk = np.log(100 * mD) + pkg.fields.gaussian_filtered(dims, corr_len=6.0, seed=args.seed, stream=r)
        perms.append(pkg.fields.diag_perm(np.exp(lnk)))
        lnr = np.log(100 * mD) + pkg.fields.gaussian_filtered(dims, corr_len=6.0, seed=args.seed + 1, stream=r)
        refs.append(pkg.fields.diag_perm(np.exp(lnr)))
    G, icp = pkg.grids.concat(grids_)
    K, Kref = np.concatenate(perms), np.concatenate(refs)
    nc, nhf = G.num_cells, G.num_half_faces
    bt = np.zeros(nhf, np.int32)
    bt[G.side == 0] = 1
    bt[G.side == 1] = 1
    bv = np.where(G.side == 0, bar, 0.0)
    # producers: the 4x4 block of x+ faces at the centre of each realization's x+ side
    cell = G.hf_cell()
    loc = cell - icp[np.searchsorted(icp, cell, side="right") - 1]
    jj, kk = (loc // dims[0]) % dims[1], loc // (dims[0] * dims[1])
    prod = np.nonzero((G.side == 1) & (np.abs(jj - 31.5) < 2) & (np.abs(kk - 15.5) < 2))[0]
    lam = 1.0 / 1.0e-3
    s = pkg.HybridSolver(0)
    t0 = time.perf_counter()
    s.init(G, Kref, (0.0, 0.0, 0.0), bt, None, inst_cellpos=icp)
    t_init = time.perf_counter() - t0
    tm = np.full(nc, lam)
    z = np.zeros(nc)
    s.solve(tm, z, z, bt, bv, residual_tolerance=1e-12, precision="fp64")
    target = s.getSolution()[2][prod]
    s.computeInnerProducts(K)
    tmd = torch.from_numpy(tm).to(dev)
    zd = torch.zeros(nc, dtype=torch.float64, device=dev)
    bvd = torch.from_numpy(bv).to(dev)
    prod_d = torch.from_numpy(prod).to(dev)
    tgt_d = torch.from_numpy(target).to(dev)
    vd = torch.empty(nhf, dtype=torch.float64, device=dev)
    wd = torch.zeros(nhf, dtype=torch.float64, device=dev)
    gK = torch.empty((nc, 9), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def step(precision):
        s.solve_device(tmd.data_ptr(), zd.data_ptr(), zd.data_ptr(), bt, bvd.data_ptr(), nrhs=1, tol=TOL,
                       precision=precision, chunk=args.chunk, use_graph=not args.no_graph)
        s.get_solution_device(None, vd.data_ptr())
        wd.zero_()
        wd[prod_d] = 2.0 * (vd[prod_d] - tgt_d)
        torch.cuda.synchronize()
        s.adjoint_device(wd.data_ptr(), gK.data_ptr(), tol=TOL, chunk=args.chunk,
                         use_graph=not args.no_graph)
        return gK

    Kd = torch.from_numpy(K).to(dev)
    def glnk():
        return (Kd[:, 0] * (gK[:, 0] + gK[:, 4] + gK[:, 8])).cpu().numpy()
    cmp = None
    if not args.no_compare:
        step("fp64")
        torch.cuda.synchronize()
        t = time.perf_counter()
        step("fp64")
        torch.cuda.synchronize()
        cmp = (time.perf_counter() - t, glnk())
    for _ in range(args.warmup):
        step(args.precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(args.precision)
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) / args.steps * 1e3
    g_mixed = glnk()
    rep, arep = s.report.as_dict(), s.adjoint_report.as_dict()
    d = s.dims()
    nnz = int(s.system(0)[0][-1])
    roof = _spmv_roofline(s, nnz, d.num_faces, 1, "32" if args.precision == "mixed" else "64", G)
    config = {
        "workload": "32 realizations of 64x64x32, lognormal isotropic K (l=6, sigma 1), J = sum over 16 "
                    "x+ producer faces per realization of (flux - target)^2, dJ/dln K by the adjoint",
        "realizations": R, "cells": nc, "producer_faces": int(prod.shape[0]), "precision": args.precision,
        "forward_iterations": rep["iterations"], "adjoint_iterations": arep["iterations"],
        "init_s": round(t_init, 2),
    }
    if cmp is not None:
        t64, g64 = cmp
        config["fp64_gpu"] = {"gradients_per_s": round(R / t64, 3), "speedup": round(t64 / (ms_step * 1e-3), 3)}
        config["grad_err_vs_fp64_gpu"] = float(np.max(np.abs(g_mixed - g64)) / np.max(np.abs(g64)))
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        n0 = int(icp[1])
        t0 = time.perf_counter()
        oracle.solve(grids_[0], K[:n0], tm[:n0], z[:n0], z[:n0], bt[:G.facepos[n0]], bv[:G.facepos[n0]],
                     tol=TOL, precond=1)
        t1 = time.perf_counter() - t0
        cpu = {"value": 1.0 / (2.0 * t1), "unit": "gradients/s", "cores": 1, "kind": "port",
               "sample": f"oracle/oracle.c FP64 Jacobi-PCG forward solve of one 64x64x32 realization on "
                         f"1 host core ({t1:.2f} s); a gradient is counted as two such solves "
                         f"(forward + adjoint), the per-cell derivative work not included"}
    print(json.dumps({
        "metric": "adjoint gradients/sec (32 realizations of 64x64x32, dJ/dln K), rel. residual 1e-10",
        "value": round(R / (ms_step * 1e-3), 3), "unit": "gradients/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 PCG / f64 residual, assembly, recovery, gradient" if args.precision == "mixed" else "f64",
        "data": "synthetic (seeded Gaussian-filtered lognormal permeability)", "config": config,
        "roofline": roof, "cpu_baseline": cpu}))

def bench_faulted(args, pkg, torch, dev):
    """200x200x50 corner-point grid (node depths perturbed +-30 % of the layer
    thickness, 5 % of columns shifted by half a layer: cells with 6 to 14
    faces), K = R diag(k, k, 0.1 k) R' with a random dip up to 20 degrees and
    lognormal k (sigma 1), Dirichlet 1 bar / 0 on the outer faces of two
    opposite corner columns, no-flow elsewhere.  One step = one solve."""
    dims = (200, 200, 50)
    t0 = time.perf_counter()
    g = pkg.grids.corner_point(*dims, 10.0, 10.0, 2.0, seed=args.seed)
    perm = pkg.fields.perm_c4(g.num_cells, seed=args.seed)
    t_gen = time.perf_counter() - t0
    nc, nhf = g.num_cells, g.num_half_faces
    ij = g.ijk[g.hf_cell()]
    inj = (ij[:, 0] == 0) & (ij[:, 1] == 0) & np.isin(g.side, [0, 2])
    prod = (ij[:, 0] == dims[0] - 1) & (ij[:, 1] == dims[1] - 1) & np.isin(g.side, [1, 3])
    bt = np.zeros(nhf, np.int32)
    bt[inj | prod] = 1
    bv = np.where(inj, 1.0e5, 0.0)
    s = pkg.HybridSolver(0)
    t0 = time.perf_counter()
    s.init(g, perm, (0.0, 0.0, 0.0), bt, None)
    t_init = time.perf_counter() - t0
    tm = torch.full((nc,), 1.0 / 1.0e-3, dtype=torch.float64, device=dev)
    rm = torch.zeros(nc, dtype=torch.float64, device=dev)
    bvd = torch.from_numpy(bv).to(dev)
    torch.cuda.synchronize()

    def step(precision):
        return s.solve_device(tm.data_ptr(), rm.data_ptr(), rm.data_ptr(), bt, bvd.data_ptr(), nrhs=1,
                              tol=TOL, precision=precision, chunk=args.chunk, use_graph=not args.no_graph,
                              **_forced_shape(args))

    cmp = None
    if not args.no_compare:
        step("fp64")
        torch.cuda.synchronize()
        t = time.perf_counter()
        r64 = step("fp64").as_dict()
        torch.cuda.synchronize()
        cmp = (time.perf_counter() - t, r64, s.getSolution())
    for _ in range(args.warmup):
        step(args.precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(args.precision)
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) / args.steps * 1e3
    rep = s.report.as_dict()
    p, pi, v = s.getSolution()
    d = s.dims()
    nnz = int(s.system(0)[0][-1])
    roof = _spmv_roofline(s, nnz, d.num_faces, 1, "32" if args.precision == "mixed" else "64", g)
    cell = g.hf_cell()
    config = {
        "workload": "200x200x50 faulted corner-point grid (6-14 faces per cell), rotated anisotropic "
                    "lognormal K (kz/kx 0.1, dip <= 20 deg), corner-column Dirichlet, multigrid-PCG to 1e-10",
        "cells": nc, "faces": d.num_faces, "half_faces": nhf, "nnz": nnz, "stored_entries": d.nnz,
        "faces_per_cell": {str(k): int(n) for k, n in enumerate(np.bincount(g.ncf())) if n},
        "precision": args.precision, "iterations": rep["iterations"], "levels": rep["levels"],
        "fp64_replacements": rep["replacements"], "reduction": rep["reduction"],
        "init_s": round(t_init, 2), "generate_s": round(t_gen, 2),
        "injection_rate": float(-v[inj].sum()), "production_rate": float(v[prod].sum()),
        "mass_balance": float(np.max(np.abs(np.bincount(cell, weights=v, minlength=nc))) / np.max(np.abs(v))),
    }
    if cmp is not None:
        t64, r64, (p64, pi64, v64) = cmp
        config["fp64_gpu"] = {"solves_per_s": round(1.0 / t64, 3), "iterations": r64["iterations"],
                              "speedup": round(t64 / (ms_step * 1e-3), 3)}
        config["err_vs_fp64_gpu"] = {"face_pressure": float(np.max(np.abs(pi - pi64)) / np.max(np.abs(pi64))),
                                     "outflux": float(np.max(np.abs(v - v64)) / np.max(np.abs(v64)))}
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        hf, nint, nf, _ = oracle.enumerate_dof(g)
        pp, _ = oracle.periodic_dof(g, hf, nf, bt, np.full(nhf, -1, np.int32))
        binv, pos, gfl = oracle.static_ip(g, perm)
        tmh, z = np.full(nc, 1.0e3), np.zeros(nc)
        t0 = time.perf_counter()
        S = oracle.assemble(g, hf, nf, pp, binv, gfl, tmh, z, z, bt, bv)
        t1 = time.perf_counter()
        x, rc, its, red = oracle.cg(S, 1e-30, maxit=args.cpu_sample_iters, precond=1, dirichlet_guess=True)
        t2 = time.perf_counter()
        oracle.recover(g, hf, binv, gfl, tmh, z, S, x)
        t3 = time.perf_counter()
        # the plain FP64 Jacobi-PCG iteration count, measured on the GPU
        s.solve_device(tm.data_ptr(), rm.data_ptr(), rm.data_ptr(), bt, bvd.data_ptr(), nrhs=1, tol=TOL,
                       precision="fp64", chunk=args.chunk, precond="jacobi")
        itj = s.report.iterations
        per_it = (t2 - t1) / max(its, 1)
        t_solve = (t1 - t0) + (t3 - t2) + per_it * itj
        cpu = {"value": 1.0 / t_solve, "unit": "solves/s", "cores": 1, "kind": "port",
               "sample": (f"oracle/oracle.c FP64 on 1 host core: assembly {t1 - t0:.2f} s + recovery "
                          f"{t3 - t2:.2f} s + {its} Jacobi-PCG iterations at {per_it * 1e3:.1f} ms, "
                          f"extrapolated to the {itj} iterations FP64 Jacobi-PCG takes on the GPU")}
    print(json.dumps({
        "metric": "face-pressure solves/sec on a 200x200x50 faulted corner-point grid, rel. residual 1e-10",
        "value": round(1000.0 / ms_step, 4), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 PCG / f64 residual, assembly, recovery" if args.precision == "mixed" else "f64",
        "data": "synthetic (seeded corner-point geometry and lognormal rotated permeability)",
        "config": config, "roofline": roof, "cpu_baseline": cpu}))


def bench_impes(args, pkg, torch, dev):
    """IMPES steps on the 128x128x64 grid (SimulatorTester::run,
    examples/SimulatorTester.hpp:59-107): per step the mobilities of the
    current saturation, one pressure solve (mixed-precision multigrid-PCG to
    1e-10) and one transport solve (explicit upstream, CFL-limited small
    steps, viscous and gravity terms) over 2 days.  Water enters through x-
    (Dirichlet 2 bar, saturation 1); x+ is held at 0 bar; porosity 0.2;
    gravity along +z; the reference's default fluid (1 / 3 cP, 1013.9 / 834.7
    kg/m^3, quadratic relative permeabilities).  The capillary term is off:
    without rock tables its CFL time is 1e100 (as in the reference) and
    p_c = 1e5 (1 - s) across 1 m cells then forces retries."""
    g, perm, bt, bv = _problem(pkg, args.seed)
    nc, nhf = g.num_cells, g.num_half_faces
    bv = np.where(g.side == 0, 2.0e5, 0.0) * (bt == 1)
    poro = np.full(nc, 0.2)
    sat_bc = np.where(g.side == 0, 1.0, 0.0)
    grav = (0.0, 0.0, 9.81)
    step_time = 2 * 86400.0
    s = pkg.HybridSolver(0)
    t0 = time.perf_counter()
    s.init(g, perm, grav, bt, None)
    t = pkg.EulerUpstream(s, poro, sat_bc)
    t.init({"method_capillary": False})
    t_init = time.perf_counter() - t0
    sat = torch.zeros(nc, dtype=torch.float64, device=dev)
    tm = torch.empty(nc, dtype=torch.float64, device=dev)
    rm = torch.empty(nc, dtype=torch.float64, device=dev)
    src = torch.zeros(nc, dtype=torch.float64, device=dev)
    bvd = torch.from_numpy(bv).to(dev)
    torch.cuda.synchronize()
    log = []

    def step():
        t.mobilities_device(sat.data_ptr(), tm.data_ptr(), rm.data_ptr())
        ta = time.perf_counter()
        s.solve_device(tm.data_ptr(), rm.data_ptr(), src.data_ptr(), bt, bvd.data_ptr(), nrhs=1, tol=TOL,
                       precision=args.precision, chunk=args.chunk, use_graph=not args.no_graph)
        tb = time.perf_counter()
        t.transportSolve_device(sat.data_ptr(), step_time, grav)
        tc = time.perf_counter()
        log.append((tb - ta, tc - tb, t.report.small_steps, t.report.repeats, s.report.iterations))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log.clear()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) / args.steps * 1e3
    small = sum(x[2] for x in log)
    t_press = sum(x[0] for x in log)
    t_trans = sum(x[1] for x in log)
    # roofline of the small-step kernel: saturation in and out, pore volume,
    # facepos, and per half-face the cell across (int) and the face's flux,
    # gravity and capillary coefficients (FP64); Dirichlet saturations
    ndir = int(np.sum(np.asarray(g.nb) < 0))
    alg = nc * (8 + 8 + 8) + (nc + 1) * 4 + nhf * (4 + 3 * 8) + ndir * 8
    ms_k = t.time_step(reps=50)
    ach = alg / (ms_k * 1e-3) / 1e9
    roof = {"kernel": "k_tr_step (explicit upstream small step, one thread per cell)", "bound": "hbm",
            "achieved": round(ach, 1), "peak": 8000.0, "unit": "GB/s", "frac": round(ach / 8000.0, 4),
            "traffic": _measured_traffic("32", "k_tr_step_128x128x64_hbm_traffic.txt"),
            "alg_bytes_per_launch": alg, "avg_launch_us": round(ms_k * 1e3, 2)}
    s_h = sat.cpu().numpy()
    config = {
        "workload": "IMPES on 128x128x64 Cartesian, anisotropic K (kz/kx=0.1, sigma ln k 1.5, l=8), water "
                    "injection through x- at 2 bar, 2-day steps: multigrid-PCG pressure to 1e-10 + explicit "
                    "upstream transport (viscous + gravity, CFL 0.5)",
        "cells": nc, "half_faces": nhf, "precision": args.precision, "init_s": round(t_init, 2),
        "step_days": step_time / 86400.0,
        "small_steps_per_step": round(small / max(args.steps, 1), 2),
        "retries": int(sum(x[3] for x in log)),
        "pcg_iterations_per_step": round(sum(x[4] for x in log) / max(args.steps, 1), 1),
        "pressure_ms_per_step": round(t_press / max(args.steps, 1) * 1e3, 3),
        "transport_ms_per_step": round(t_trans / max(args.steps, 1) * 1e3, 3),
        "transport_small_steps_per_s": round(small / max(t_trans, 1e-12), 1),
        "water_in_place": float(np.sum(s_h) * 0.2),
        "max_saturation": float(s_h.max()),
    }
    cpu = None
    if not args.no_cpu_baseline:
        # the oracle on one host core: one residual evaluation of the
        # transport at full size (numpy), the pressure solve as the default
        # workload's baseline (assembly, recovery, Jacobi-PCG iterations
        # extrapolated to the count FP64 Jacobi-PCG needs on the GPU)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import euler_upstream as orc_tr
        flux = np.asarray(s.getSolution()[2]).reshape(-1)
        tr0 = time.perf_counter()
        orc_tr.residual(g, perm, s_h, flux, grav, 1013.9 - 834.7, (1e-3, 3e-3), None, sat_bc,
                        method_capillary=False)
        t_res = time.perf_counter() - tr0
        s.solve_device(tm.data_ptr(), rm.data_ptr(), src.data_ptr(), bt, bvd.data_ptr(), nrhs=1, tol=TOL,
                       precision="fp64", chunk=args.chunk, precond="jacobi")
        itj = s.report.iterations
        base = _cpu_baseline(pkg, g, perm, bt, bv, itj, max(args.cpu_sample_iters // 4, 5))
        per_step = base["seconds_per_solve"] + t_res * small / max(args.steps, 1)
        cpu = {"value": 1.0 / per_step, "unit": "IMPES steps/s", "cores": 1, "kind": "port",
               "sample": (f"oracle on 1 host core: transport residual (oracle/euler_upstream.py, numpy) "
                          f"{t_res:.2f} s per small step x {small / max(args.steps, 1):.1f} small steps, "
                          f"plus the pressure solve {base['seconds_per_solve']:.1f} s ({base['sample']})")}
    print(json.dumps({
        "metric": "IMPES steps/sec (pressure solve + explicit upstream transport) at 128x128x64",
        "value": round(1000.0 / ms_step, 4), "unit": "IMPES steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 PCG / f64 residual, assembly, recovery, transport",
        "data": "synthetic (seeded Gaussian-filtered lognormal permeability)",
        "config": config, "roofline": roof, "cpu_baseline": cpu}))


def bench_aniso(args, pkg, torch, dev):
    """The 128x128x64 headline with tensor mobilities
    (MimeticIPAnisoRelpermEvaluator): saturation s = 0.5 + 0.3 z (z the
    standardised permeability field), lambda_t = (s^2 / 1 cP + (1-s)^2 / 3 cP)
    diag(1, 1, 0.6) and sum rho_i lambda_i likewise (a relative permeability
    reduced vertically, in K's principal axes), gravity along +z.  One step =
    one solve: the per-cell lambda_t K inner products, assembly, mixed-precision
    multigrid-PCG to 1e-10 and recovery."""
    g, perm, bt, bv = _problem(pkg, args.seed)
    nc = g.num_cells
    z = np.log(perm[:, 0])
    z = (z - z.mean()) / z.std()
    sat = np.clip(0.5 + 0.3 * z, 0.0, 1.0)
    lw, lo = sat * sat / 1e-3, (1 - sat) ** 2 / 3e-3
    scale = np.array([1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.6])
    lam = (lw + lo)[:, None] * scale
    rl = (1013.9 * lw + 834.7 * lo)[:, None] * scale
    grav = (0.0, 0.0, 9.81)
    s = pkg.HybridSolver(0)
    t0 = time.perf_counter()
    s.init(g, perm, grav, bt, None)
    t_init = time.perf_counter() - t0
    L = torch.from_numpy(np.ascontiguousarray(lam)).to(dev)
    R = torch.from_numpy(np.ascontiguousarray(rl)).to(dev)
    src = torch.zeros(nc, dtype=torch.float64, device=dev)
    bvd = torch.from_numpy(bv).to(dev)
    torch.cuda.synchronize()

    def step():
        return s.solve_aniso_device(L.data_ptr(), R.data_ptr(), src.data_ptr(), bt, bvd.data_ptr(), nrhs=1,
                                    tol=TOL, precision=args.precision, chunk=args.chunk,
                                    use_graph=not args.no_graph)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) / args.steps * 1e3
    rep = s.report.as_dict()
    v = s.getSolution()[2]
    d = s.dims()
    nnz = int(s.system(0)[0][-1])
    roof = _spmv_roofline(s, nnz, d.num_faces, 1, "32" if args.precision == "mixed" else "64", g)
    cell = g.hf_cell()
    config = {
        "workload": "mimetic hybrid solve with tensor mobilities, 128x128x64 Cartesian, anisotropic K "
                    "(kz/kx=0.1, sigma ln k 1.5, l=8), lambda_t = lambda(s) diag(1,1,0.6), gravity, "
                    "multigrid-PCG to 1e-10",
        "cells": nc, "faces": d.num_faces, "nnz": nnz, "precision": args.precision,
        "iterations": rep["iterations"], "levels": rep["levels"], "reduction": rep["reduction"],
        "init_s": round(t_init, 2),
        "mass_balance": float(np.max(np.abs(np.bincount(cell, weights=v, minlength=nc))) / np.max(np.abs(v))),
    }
    cpu = None
    if not args.no_cpu_baseline:
        # the oracle on one host core: the tensor inner products at full size,
        # plus the scalar pressure baseline (assembly, recovery, Jacobi-PCG
        # iterations extrapolated to the count FP64 Jacobi-PCG needs here)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        t0 = time.perf_counter()
        oracle.aniso_inner_products(g, perm, lam, rl, grav)
        t_ip = time.perf_counter() - t0
        s.solve_aniso_device(L.data_ptr(), R.data_ptr(), src.data_ptr(), bt, bvd.data_ptr(), nrhs=1, tol=TOL,
                             precision="fp64", chunk=args.chunk, precond="jacobi")
        itj = s.report.iterations
        base = _cpu_baseline(pkg, g, perm, bt, bv, itj, max(args.cpu_sample_iters // 4, 5))
        t_all = t_ip + base["seconds_per_solve"]
        cpu = {"value": 1.0 / t_all, "unit": "solves/s", "cores": 1, "kind": "port",
               "sample": (f"oracle on 1 host core: tensor inner products {t_ip:.2f} s (oracle.aniso_inner_products) "
                          f"+ pressure solve {base['seconds_per_solve']:.1f} s ({base['sample']}); "
                          f"FP64 Jacobi-PCG takes {itj} iterations on this system")}
    print(json.dumps({
        "metric": "face-pressure solves/sec with tensor mobilities at 128x128x64, rel. residual 1e-10",
        "value": round(1000.0 / ms_step, 4), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 PCG / f64 residual, inner products, assembly, recovery" if args.precision == "mixed" else "f64",
        "data": "synthetic (seeded Gaussian-filtered lognormal permeability)",
        "config": config, "roofline": roof, "cpu_baseline": cpu}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="solve", choices=["solve", "upscale", "adjoint", "faulted", "impes", "aniso"],
                    help="solve: the 128x128x64 headline; upscale: 256 periodic cell problems "
                         "x 3 pressure drops; adjoint: gradients of 32 realizations of 64x64x32; "
                         "faulted: 200x200x50 faulted corner-point grid; impes: pressure + "
                         "transport steps at 128x128x64; aniso: the headline with tensor mobilities")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--precond", default="auto", choices=["auto", "amg", "jacobi"])
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--chunk", type=int, default=32)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--spmv-threads", type=int, default=0,
                    help="force the SpMV workgroup size (256, 512, 1024; default: autotuned)")
    ap.add_argument("--sweep-threads", type=int, default=0,
                    help="force the post-smoothing sweep's workgroup size (default: autotuned)")
    ap.add_argument("--tail-rows", type=int, default=-1,
                    help="force the one-workgroup tail's level size limit (default: autotuned)")
    ap.add_argument("--element-operator", type=int, default=-1, choices=[-1, 0, 1, 2],
                    help="level-0 operator of the FP32 iteration: 0 assembled matrix, 1 autotuned, "
                         "2 element by element (default: the solver's, 1)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch PCG kernels one by one instead of replaying captured graphs")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the FP64 comparison solves (profiling runs)")
    ap.add_argument("--cpu-sample-iters", type=int, default=40)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run one device copy of this size (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the solver runs on one GPU; for several GPUs, run one process per GPU on "
              "separate instances (the batched workloads) -- there is no multi-GPU solve.",
              file=sys.stderr)
        sys.exit(2)

    import __graft_entry__
    pkg = __graft_entry__.load_package()
    import torch
    dev = torch.device("cuda:0")
    if args.workload == "upscale":
        return bench_upscale(args, pkg, torch, dev)
    if args.workload == "adjoint":
        return bench_adjoint(args, pkg, torch, dev)
    if args.workload == "aniso":
        return bench_aniso(args, pkg, torch, dev)
    if args.workload == "impes":
        return bench_impes(args, pkg, torch, dev)
    if args.workload == "faulted":
        return bench_faulted(args, pkg, torch, dev)
    g, perm, bt, bv = _problem(pkg, args.seed)
    nc = g.num_cells

    s = pkg.HybridSolver(0)
    t0 = time.perf_counter()
    s.init(g, perm, (0.0, 0.0, 0.0), bt, None)
    t_init = time.perf_counter() - t0
    d = s.dims()
    nf = d.num_faces

    tm = torch.ones(nc, dtype=torch.float64, device=dev)
    rm = torch.zeros(nc, dtype=torch.float64, device=dev)
    src = torch.zeros(nc, dtype=torch.float64, device=dev)
    bvd = torch.from_numpy(bv).to(dev)
    torch.cuda.synchronize()

    def step(precision, precond=args.precond):
        return s.solve_device(tm.data_ptr(), rm.data_ptr(), src.data_ptr(), bt, bvd.data_ptr(),
                              nrhs=1, tol=TOL, precision=precision, chunk=args.chunk,
                              precond=precond, use_graph=not args.no_graph, **_forced_shape(args))

    def timed(precision, precond):
        step(precision, precond)                 # first call builds graphs / hierarchy
        torch.cuda.synchronize()
        t = time.perf_counter()
        step(precision, precond)
        torch.cuda.synchronize()
        t = time.perf_counter() - t
        rep = s.report.as_dict()
        return t, rep, s.getSolution()

    cmp = {}
    if not args.no_compare:
        # FP64 with the same preconditioner: the comparison point for the error report
        cmp["fp64"] = timed("fp64", args.precond)
        # the plain method in FP64: Jacobi-PCG
        cmp["fp64_jacobi"] = timed("fp64", "jacobi")

    for _ in range(args.warmup):
        step(args.precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(args.precision)
    torch.cuda.synchronize()
    t_tot = time.perf_counter() - t0
    rep = s.report.as_dict()
    ms_step = t_tot / args.steps * 1e3
    p, pi, v = s.getSolution()
    nnz = int(s.system(0)[0][-1])

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    cell = g.hf_cell()
    mass = float(np.max(np.abs(np.bincount(cell, weights=v, minlength=nc))) / np.max(np.abs(v)))

    B = _alg_bytes(g, nnz, nf, rep)
    tag = "32" if args.precision == "mixed" else "64"
    ebe = bool(rep["element_operator"])
    if ebe:
        it_key = "amg_iter_ebe" if rep["precond"] == 1 else None
        ref_key = "refine64_ebe"
    else:
        it_key = ("amg_iter" if rep["precond"] == 1 else "iter") + tag
        ref_key = "refine64"
    n_ref = rep["replacements"]
    if it_key is None:      # element-by-element Jacobi-PCG
        it_bytes = B["ebe_spmv"] + B["iter32"] - B["spmv32"]
    else:
        it_bytes = B[it_key]
    bytes_solve = B["assemble"] + rep["iterations"] * it_bytes + n_ref * B[ref_key] + B["recover"]
    if ebe:   # the assembly also writes the FP64 and FP32 local matrices
        bytes_solve += B["pk"] * 12
    gbps = bytes_solve / (ms_step * 1e-3) / 1e9

    if args.copy_probe_mib > 0:
        n = args.copy_probe_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    kern_ms = s.time_kernel(0, reps=50)
    if ebe:
        alg = B["ebe_spmv"]
        roof = {"kernel": f"k_ebe<6, 1, EBE_SPMV, REL> (element-by-element PCG product q = A p with fused <p,Ap>, "
                          f"{rep['colours']} colour launches per product)",
                "traffic": _measured_traffic(tag, "k_ebe_spmv_128x128x64_hbm_traffic.txt")}
    else:
        alg = B["spmv" + tag]
        roof = {"kernel": f"{'k_spmv_pq8<float>' if tag == '32' else 'k_spmv_pq<double>'} (face-matrix SpMV "
                          f"with fused <p,Ap>)",
                "traffic": _measured_traffic(tag)}
    ach = alg / (kern_ms * 1e-3) / 1e9
    roof.update({"bound": "hbm", "achieved": round(ach, 1), "peak": 8000.0, "unit": "GB/s",
                 "frac": round(ach / 8000.0, 4), "alg_bytes_per_launch": alg,
                 "avg_launch_us": round(kern_ms * 1e3, 2)})
    if ebe:
        roof["alg_bytes_note"] = ("per product (both colour launches): packed FP32 local matrices and face rows "
                                  "of every cell, p gathered once, q written once and read+written again on "
                                  "faces with two cells")
        # the assembled FP32 matrix product the same solve no longer runs
        sell_ms = s.time_kernel(1, reps=20)
        roof["assembled_matrix_spmv"] = {"alg_bytes": B["spmv32"], "avg_us": round(sell_ms * 1e3, 2),
                                         "achieved": round(B["spmv32"] / (sell_ms * 1e-3) / 1e9, 1)}
    copy_gbps = _copy_gbps(dev)
    roof["device_copy_gbps"] = round(copy_gbps, 1)
    roof["frac_of_device_copy"] = round(ach / copy_gbps, 4)

    fx = _c2_fixture() if args.seed == 2 else None
    cpu = None
    if not args.no_cpu_baseline:
        if fx is not None:
            cpu = _cpu_baseline_ilu0(pkg, g, perm, bt, bv, fx, args.cpu_sample_iters)
        elif "fp64_jacobi" in cmp:
            cpu = _cpu_baseline(pkg, g, perm, bt, bv, cmp["fp64_jacobi"][1]["iterations"],
                                args.cpu_sample_iters)

    config = {
        "workload": "mimetic hybrid solve, 128x128x64 Cartesian, anisotropic K (kz/kx=0.1, "
                    "sigma ln k 1.5, l=8), x pressure drop, "
                    + ("multigrid-PCG" if rep["precond"] == 1 else "Jacobi-PCG") + " to 1e-10",
        "cells": nc, "faces": nf, "nnz": nnz, "precision": args.precision,
        "preconditioner": "multigrid" if rep["precond"] == 1 else "jacobi",
        "levels": rep["levels"], "iterations": rep["iterations"], "fp64_replacements": n_ref,
        "reduction": rep["reduction"], "init_s": round(t_init, 3),
        "hbm_gbps_per_solve": round(gbps, 1), "alg_bytes_per_solve": int(bytes_solve),
        "mass_balance": mass,
        "launch_shape": {k: rep[k] for k in ("spmv_threads", "spmv_xcd", "sweep_threads", "sweep_xcd",
                                             "tail_rows")},
        "level0_operator": ("element by element (FP32 local matrices, %d colours)" % rep["colours"]) if ebe
                           else "assembled face matrix",
        "face_matrix_filled": bool(rep["face_matrix"]),
    }
    if rep["precond"] == 1:
        config["multigrid"] = {"level_rows": [int(x) for x in rep["level_rows"] if x],
                               "level_nnz": [int(x) for x in rep["level_nnz"] if x],
                               "ap_nnz": int(rep["ap_nnz"]), "smoothed_level": int(rep["smoothed_level"]),
                               "transfer_nnz": [int(x) for x in rep["transfer_nnz"]],
                               "smoother_weights": rep["smoother_weights"],
                               "coarsening": rep["coarsening"]}
    kkey = "kernels_ebe" if (ebe and rep["precond"] == 1) else ("kernels" + tag)
    if kkey in B:
        # the per-iteration part of alg_bytes_per_solve, kernel by kernel
        config["alg_bytes_per_iteration"] = {k: int(b) for k, b in B[kkey]}
    if fx is not None:
        config["err_vs_oracle"] = _err_vs_oracle(g, perm, fx, p, pi, v)
        config["err_vs_oracle_note"] = ("max-norm relative error at 32768 seeded positions per field against "
                                        "the oracle's ILU0-CG solution to 1e-14 (tests/golden/c2_128x128x64.npz)")
    for key, (t, r, sol) in cmp.items():
        p64, pi64, v64 = sol
        config[key + "_gpu"] = {
            "solves_per_s": round(1.0 / t, 4), "iterations": r["iterations"],
            "reduction": r["reduction"], "speedup": round((1000.0 / ms_step) * t, 3),
        }
        if key == "fp64":
            config["err_vs_fp64_gpu"] = {"face_pressure": rel(pi, pi64), "cell_pressure": rel(p, p64),
                                         "outflux": rel(v, v64)}
        if fx is not None:
            config[key + "_gpu"]["err_vs_oracle"] = _err_vs_oracle(g, perm, fx, p64, pi64, v64)

    out = {
        "metric": "face-pressure solves/sec + HBM GB/s at 128x128x64, rel. residual 1e-10",
        "value": round(1000.0 / ms_step, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 PCG / f64 residual, assembly, recovery" if args.precision == "mixed" else "f64",
        "data": "synthetic (seeded Gaussian-filtered lognormal permeability)",
        "config": config,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""ctypes front end of the FP64 CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product library never loads
it.  ``solve`` chains the restated stages in the order of
IncompFlowSolverHybrid::init (IncompFlowSolverHybrid.hpp:476-488) and
IncompFlowSolverHybrid::solve (:657-703).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

NEUMANN, DIRICHLET, PERIODIC = 0, 1, 2

_i = ctypes.POINTER(ctypes.c_int)
_d = ctypes.POINTER(ctypes.c_double)
_ll = ctypes.POINTER(ctypes.c_longlong)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_enumerate_dof.argtypes = [ctypes.c_int, _i, _i, _i, _i]
        L.orc_periodic_dof.argtypes = [ctypes.c_int, _i, _i, _i, _i, ctypes.c_int, _i]
        L.orc_static_ip.argtypes = [ctypes.c_int, _i, _d, _d, _d, _d, _d, _d, _d, _d, _d]
        L.orc_assemble.argtypes = [ctypes.c_int, _i, _i, _i, ctypes.c_int, _i, _d, _d, _d, _d,
                                   _d, _i, _d, _i, _i, _d, ctypes.c_longlong, _ll, _d, _d, _d,
                                   _d, _i]
        L.orc_regularize.argtypes = [_i, _i, _d]
        L.orc_cg.argtypes = [ctypes.c_int, _i, _i, _d, _d, _d, ctypes.c_double, ctypes.c_int,
                             ctypes.c_int, ctypes.c_int, _d]
        L.orc_spmv.argtypes = [ctypes.c_int, _i, _i, _d, _d, _d]
        L.orc_recover.argtypes = [ctypes.c_int, _i, _i, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d]
        L.orc_post_process.argtypes = [ctypes.c_int, _i, _i, _i, ctypes.c_int, _i, ctypes.c_int,
                                       _d]
        L.orc_post_process.restype = ctypes.c_double
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _LIB = L
    return _LIB


def set_threads(n: int) -> None:
    """Host threads of the CG's SpMV, dot products and vector updates.  1 (the
    default) sums in the single-threaded order every pinned result uses; more
    threads reorder the dot-product sums (results then differ in the last bits)."""
    lib().orc_set_threads(int(n))


def _pi(a):
    return a.ctypes.data_as(_i)


def _pd(a):
    return a.ctypes.data_as(_d)


def _ci(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _cd(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def enumerate_dof(g):
    """IncompFlowSolverHybrid.hpp:956-1067.  Returns (hf_dof, n_internal, nf, max_ncf)."""
    facepos, nb = _ci(g.facepos), _ci(g.nb)
    hf_dof = np.empty(g.num_half_faces, dtype=np.int32)
    info = np.zeros(3, dtype=np.int32)
    rc = lib().orc_enumerate_dof(g.num_cells, _pi(facepos), _pi(nb), _pi(hf_dof), _pi(info))
    if rc != 0:
        raise ValueError("invalid grid topology")
    return hf_dof, int(info[0]), int(info[1]), int(info[2])


def periodic_dof(g, hf_dof, nf, bctype, partner):
    """IncompFlowSolverHybrid.hpp:1071-1110."""
    pp = np.empty(nf, dtype=np.int32)
    bt, pa = _ci(bctype), _ci(partner)
    n = lib().orc_periodic_dof(g.num_half_faces, _pi(_ci(g.nb)), _pi(hf_dof), _pi(bt), _pi(pa),
                               nf, _pi(pp))
    return pp, n


def static_ip(g, perm, grav=(0.0, 0.0, 0.0)):
    """MimeticIPEvaluator.hpp:201-270.  Returns (binv, binv_pos, gflux_static)."""
    ncf = g.ncf().astype(np.int64)
    pos = np.zeros(g.num_cells + 1, dtype=np.int64)
    np.cumsum(ncf * ncf, out=pos[1:])
    binv = np.empty(int(pos[-1]), dtype=np.float64)
    gfl = np.empty(g.num_half_faces, dtype=np.float64)
    perm = _cd(perm).reshape(-1)
    gr = _cd(grav)
    rc = lib().orc_static_ip(g.num_cells, _pi(_ci(g.facepos)), _pd(_cd(g.cell_volume)),
                             _pd(_cd(g.cell_centroid)), _pd(_cd(g.face_area)),
                             _pd(_cd(g.face_centroid)), _pd(_cd(g.face_normal)), _pd(perm),
                             _pd(gr), _pd(binv), _pd(gfl))
    if rc != 0:
        raise ValueError("static inner product failed")
    return binv, pos, gfl


def cell_matrix(binv, pos, g, c):
    n = int(g.facepos[c + 1] - g.facepos[c])
    return binv[pos[c]:pos[c + 1]].reshape(n, n, order="F")


class System:
    pass


def assemble(g, hf_dof, nf, ppartner, binv, gflux, totmob, rhomob, src, bctype, bcval):
    """IncompFlowSolverHybrid.hpp:1324-1382, 1736-1779, 1785-1811, 1817-1901."""
    ncf = g.ncf().astype(np.int64)
    cap = int(np.sum(ncf * ncf)) + 4 * int(np.sum(np.asarray(bctype) == PERIODIC)) + 16
    rp = np.empty(nf + 1, dtype=np.int32)
    col = np.empty(cap, dtype=np.int32)
    val = np.empty(cap, dtype=np.float64)
    rhs = np.empty(nf, dtype=np.float64)
    F = np.empty(g.num_half_faces, dtype=np.float64)
    L = np.empty(g.num_cells, dtype=np.float64)
    gg = np.empty(g.num_cells, dtype=np.float64)
    nnz = ctypes.c_longlong(0)
    reg = ctypes.c_int(0)
    rc = lib().orc_assemble(g.num_cells, _pi(_ci(g.facepos)), _pi(_ci(g.nb)), _pi(hf_dof), nf,
                            _pi(ppartner), _pd(binv), _pd(gflux), _pd(_cd(totmob)),
                            _pd(_cd(rhomob)), _pd(_cd(src)), _pi(_ci(bctype)), _pd(_cd(bcval)),
                            _pi(rp), _pi(col), _pd(val), cap, ctypes.byref(nnz), _pd(rhs), _pd(F),
                            _pd(L), _pd(gg), ctypes.byref(reg))
    if rc != 0:
        raise RuntimeError("assembly capacity exceeded")
    s = System()
    m = nnz.value
    s.row_ptr, s.col, s.val = rp, col[:m].copy(), val[:m].copy()
    s.rhs, s.F, s.L, s.g = rhs, F, L, gg
    s.do_regularization = bool(reg.value)
    s.nf = nf
    return s


def regularize(s):
    lib().orc_regularize(_pi(s.row_ptr), _pi(s.col), _pd(s.val))


def cg(s, tol, maxit=0, precond=2, dirichlet_guess=False):
    """dune-istl CGSolver as driven by IncompFlowSolverHybrid.hpp:1387-1541."""
    x = np.zeros(s.nf, dtype=np.float64)
    out = np.zeros(2, dtype=np.float64)
    rc = lib().orc_cg(s.nf, _pi(s.row_ptr), _pi(s.col), _pd(s.val), _pd(s.rhs), _pd(x), tol,
                      maxit if maxit > 0 else s.nf, precond, int(dirichlet_guess), _pd(out))
    return x, rc, int(out[0]), float(out[1])


def spmv(s, x):
    y = np.empty(s.nf, dtype=np.float64)
    lib().orc_spmv(s.nf, _pi(s.row_ptr), _pi(s.col), _pd(s.val), _pd(_cd(x)), _pd(y))
    return y


def recover(g, hf_dof, binv, gflux, totmob, rhomob, s, soln):
    """IncompFlowSolverHybrid.hpp:1672-1730."""
    p = np.empty(g.num_cells, dtype=np.float64)
    v = np.empty(g.num_half_faces, dtype=np.float64)
    lib().orc_recover(g.num_cells, _pi(_ci(g.facepos)), _pi(hf_dof), _pd(binv), _pd(gflux),
                      _pd(_cd(totmob)), _pd(_cd(rhomob)), _pd(s.F), _pd(s.L), _pd(s.g),
                      _pd(_cd(soln)), _pd(p), _pd(v))
    return p, v


def post_process(g, hf_dof, nf, ppartner, have_periodic, outflux):
    """IncompFlowSolverHybrid.hpp:707-807.  Modifies outflux in place."""
    v = _cd(outflux).copy()
    mm = lib().orc_post_process(g.num_cells, _pi(_ci(g.facepos)), _pi(_ci(g.nb)), _pi(hf_dof),
                                nf, _pi(ppartner), int(have_periodic), _pd(v))
    return v, mm


class Result:
    pass


def solve(g, perm, totmob, rhomob, src, bctype, bcval, partner=None, grav=(0.0, 0.0, 0.0),
          tol=1e-14, maxit=0, precond=2, dirichlet_guess=False, check=True):
    """init() followed by solve(): the whole reference path on the CPU, FP64."""
    t0 = time.perf_counter()
    hf_dof, nint, nf, maxf = enumerate_dof(g)
    bctype = _ci(bctype)
    if partner is None:
        partner = np.full(g.num_half_faces, -1, dtype=np.int32)
    pp, nper = periodic_dof(g, hf_dof, nf, bctype, _ci(partner))
    binv, pos, gfl = static_ip(g, perm, grav)
    t1 = time.perf_counter()
    s = assemble(g, hf_dof, nf, pp, binv, gfl, totmob, rhomob, src, bctype, bcval)
    if s.do_regularization:
        regularize(s)
    t2 = time.perf_counter()
    x, rc, its, red = cg(s, tol, maxit, precond, dirichlet_guess)
    if check and rc != 0:
        raise RuntimeError(f"oracle CG failed rc={rc} its={its} reduction={red}")
    t3 = time.perf_counter()
    p, v = recover(g, hf_dof, binv, gfl, totmob, rhomob, s, x)
    t4 = time.perf_counter()
    r = Result()
    r.hf_dof, r.nf, r.n_internal, r.max_ncf = hf_dof, nf, nint, maxf
    r.ppartner, r.n_periodic = pp, nper
    r.binv, r.binv_pos, r.gflux = binv, pos, gfl
    r.system = s
    r.face_pressure, r.cell_pressure, r.outflux = x, p, v
    r.iterations, r.reduction, r.rc = its, red, rc
    r.t_init, r.t_assemble, r.t_cg, r.t_recover = t1 - t0, t2 - t1, t3 - t2, t4 - t3
    return r


def aniso_inner_products(g, perm, lambda_t, rho_lambda, grav=(0.0, 0.0, 0.0)):
    """MimeticIPAnisoRelpermEvaluator (MimeticIPAnisoRelpermEvaluator.hpp):
    computeDynamicParams (:299-330) forms lambda K = lambda_t K and
    dyn_Kg = (sum rho_i lambda_i) (K g); getInverseMatrix (:363-400) is the
    static formula with lambda K in place of K (t = 6 tr(lambda K) / 3), so
    the static evaluator applied to lambda K gives B^-1; gravityFlux
    (:406-416) is N dyn_Kg.  Returns (binv, pos, gflux)."""
    nc = g.num_cells
    K = np.asarray(perm, dtype=np.float64).reshape(nc, 3, 3)
    L = np.asarray(lambda_t, dtype=np.float64).reshape(nc, 3, 3)
    R = np.asarray(rho_lambda, dtype=np.float64).reshape(nc, 3, 3)
    lamK = np.einsum("cij,cjk->cik", L, K)
    binv, pos, _ = static_ip(g, lamK.reshape(nc, 9))
    dyn = np.einsum("cij,cj->ci", R, np.einsum("cij,j->ci", K, np.asarray(grav, dtype=np.float64)))
    hc = g.hf_cell()
    gfl = np.einsum("hi,hi->h", np.asarray(g.face_normal) * np.asarray(g.face_area)[:, None], dyn[hc])
    return binv, pos, gfl


def solve_aniso(g, perm, lambda_t, rho_lambda, src, bctype, bcval, partner=None, grav=(0.0, 0.0, 0.0),
                tol=1e-14, maxit=0, precond=2):
    """solve() with the tensor-mobility inner products: assembly and recovery
    with unit scalar mobilities on the dynamic B^-1 and gravity flux."""
    hf_dof, nint, nf, maxf = enumerate_dof(g)
    bctype = _ci(bctype)
    if partner is None:
        partner = np.full(g.num_half_faces, -1, dtype=np.int32)
    pp, nper = periodic_dof(g, hf_dof, nf, bctype, _ci(partner))
    binv, pos, gfl = aniso_inner_products(g, perm, lambda_t, rho_lambda, grav)
    ones = np.ones(g.num_cells)
    s = assemble(g, hf_dof, nf, pp, binv, gfl, ones, ones, src, bctype, bcval)
    if s.do_regularization:
        regularize(s)
    x, rc, its, red = cg(s, tol, maxit, precond, False)
    if rc != 0:
        raise RuntimeError(f"oracle CG failed rc={rc} its={its} reduction={red}")
    p, v = recover(g, hf_dof, binv, gfl, ones, ones, s, x)
    r = Result()
    r.face_pressure, r.cell_pressure, r.outflux = x, p, v
    r.binv, r.binv_pos, r.gflux = binv, pos, gfl
    r.iterations, r.reduction, r.rc = its, red, rc
    return r


def topology_is_sane(num_cells, facepos, nb):
    """topologyIsSane (IncompFlowSolverHybrid.hpp:108-130): true when, for
    every cell, its internal neighbours (non-boundary faces) are pairwise
    distinct, i.e. no two cells share more than one face.  Plain loops."""
    for c in range(num_cells):
        seen = set()
        for h in range(int(facepos[c]), int(facepos[c + 1])):
            n = int(nb[h])
            if n < 0:
                continue
            if n in seen:
                return False
            seen.add(n)
    return True


def gravity_term(g, grav, omega):
    """MimeticIPEvaluator::gravityTerm (MimeticIPEvaluator.hpp:477-523) for
    every cell: gterm[h] = omega[c] * (x_f - x_c) . g over the half-faces h of
    cell c (the face centroid minus the cell centroid, dotted with gravity)."""
    cell = g.hf_cell()
    d = np.asarray(g.face_centroid, dtype=np.float64).reshape(-1, 3) - \
        np.asarray(g.cell_centroid, dtype=np.float64).reshape(-1, 3)[cell]
    return np.asarray(omega, dtype=np.float64)[cell] * (d @ np.asarray(grav, dtype=np.float64))


def estimate_cell_velocity(g, outflux):
    """estimateCellVelocity (opm/porsol/common/SimulatorUtilities.hpp:59-81):
    per cell the sum over its faces of (x_f - x_c) * outflux(f) / |c|."""
    hc = g.hf_cell()
    d = np.asarray(g.face_centroid) - np.asarray(g.cell_centroid)[hc]
    contrib = d * (np.asarray(outflux, dtype=np.float64) / np.asarray(g.cell_volume)[hc])[:, None]
    v = np.zeros((g.num_cells, 3))
    np.add.at(v, hc, contrib)
    return v

"""ctypes binding of libmhs.so (include/mimetic_hybrid.h).

``HybridSolver`` mirrors the public methods of opm-porsol's
IncompFlowSolverHybrid (opm/porsol/mimetic/IncompFlowSolverHybrid.hpp):
init :476, clear :498, initSystemStructure :534, computeInnerProducts :564,
solve :657, postProcessFluxes :754, getSolution :828.  Errors keep the
reference's behaviour: non-convergence and an unknown linsolver_type raise
RuntimeError (:1416-1419, :698-700).

There is no CPU fallback: if libmhs.so is missing or no GPU is present the
calls fail loudly.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libmhs.so")

MHS_OK, MHS_ERR_INVALID, MHS_ERR_STATE, MHS_ERR_NOT_CONVERGED, MHS_ERR_HIP, MHS_ERR_ALLOC, \
    MHS_ERR_BREAKDOWN, MHS_ERR_RANGE = range(8)
NEUMANN, DIRICHLET, PERIODIC = 0, 1, 2
FP64, MIXED = 0, 1

EXPORTS = (
    "mhs_default_options", "mhs_create", "mhs_destroy", "mhs_last_error", "mhs_clear",
    "mhs_set_grid", "mhs_set_cell_ijk", "mhs_init_system_structure", "mhs_compute_inner_products",
    "mhs_assemble_static", "mhs_get_inner_products", "mhs_solve", "mhs_solve_device", "mhs_get_solution",
    "mhs_solution_device", "mhs_post_process_fluxes", "mhs_get_dims", "mhs_get_cell_faces",
    "mhs_get_face_rows", "mhs_get_system", "mhs_time_kernel", "mhs_compute_inner_products_device",
    "mhs_get_solution_device", "mhs_adjoint", "mhs_adjoint_device", "mhs_sum_outflux",
    "mhs_gravity_term", "mhs_topology_is_sane", "mhs_estimate_cell_velocity",
    "mhs_estimate_cell_velocity_device", "mhs_solve_aniso", "mhs_solve_aniso_device",
    "mhs_sum_outflux_noalloc",
)


class Options(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("maxit", ctypes.c_int),
                ("linsolver_type", ctypes.c_int), ("verbosity", ctypes.c_int),
                ("same_matrix", ctypes.c_int), ("prolongate_factor", ctypes.c_double),
                ("smooth_steps", ctypes.c_int), ("precision", ctypes.c_int),
                ("nrhs", ctypes.c_int), ("chunk", ctypes.c_int),
                ("replace_factor", ctypes.c_double), ("use_graph", ctypes.c_int),
                ("precond", ctypes.c_int), ("smoother_weight", ctypes.c_double),
                ("autotune", ctypes.c_int), ("spmv_threads", ctypes.c_int),
                ("sweep_threads", ctypes.c_int), ("tail_rows", ctypes.c_int),
                ("element_operator", ctypes.c_int)]


class Report(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("replacements", ctypes.c_int),
                ("converged", ctypes.c_int), ("reduction", ctypes.c_double),
                ("precond", ctypes.c_int), ("levels", ctypes.c_int),
                ("mg_rows", ctypes.c_longlong), ("mg_nnz", ctypes.c_longlong),
                ("assemble_ms", ctypes.c_double), ("solve_ms", ctypes.c_double),
                ("recover_ms", ctypes.c_double), ("spmv_threads", ctypes.c_int),
                ("spmv_xcd", ctypes.c_int), ("sweep_threads", ctypes.c_int),
                ("sweep_xcd", ctypes.c_int), ("tail_rows", ctypes.c_int),
                ("tune_ms", ctypes.c_double), ("smoother_weights", ctypes.c_double * 8),
                ("coarsening", (ctypes.c_int * 3) * 8), ("ap_nnz", ctypes.c_longlong),
                ("element_operator", ctypes.c_int), ("colours", ctypes.c_int),
                ("element_values", ctypes.c_longlong), ("level_rows", ctypes.c_longlong * 8),
                ("level_nnz", ctypes.c_longlong * 8), ("transfer_nnz", ctypes.c_longlong * 2),
                ("smoothed_level", ctypes.c_int), ("face_matrix", ctypes.c_int),
                ("mg_setup_reused", ctypes.c_int)]

    def as_dict(self):
        out = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            if k == "smoother_weights":
                v = [x for x in v if x > 0.0]
            elif k == "coarsening":
                v = [tuple(b) for b in v if b[0] > 0]
            elif k in ("level_rows", "level_nnz"):
                v = [int(x) for x in v if x > 0]
            elif k == "transfer_nnz":
                v = [int(x) for x in v]
            out[k] = v
        return out


class Dims(ctypes.Structure):
    _fields_ = [("num_cells", ctypes.c_int), ("num_half_faces", ctypes.c_int),
                ("num_faces", ctypes.c_int), ("num_internal_faces", ctypes.c_int),
                ("max_ncf", ctypes.c_int), ("num_instances", ctypes.c_int),
                ("num_rows_padded", ctypes.c_int), ("nnz", ctypes.c_longlong)]


_LIB = None
_P = ctypes.c_void_p
_I = ctypes.POINTER(ctypes.c_int)
_D = ctypes.POINTER(ctypes.c_double)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(f"{LIBPATH} is missing: build it with __graft_entry__.build()")
        # torch ships its own HIP runtime under the same soname as the system
        # one.  Loading torch first makes libmhs bind to that runtime, so the
        # process holds one HIP runtime; two cannot both open the GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIBPATH)
        L.mhs_default_options.argtypes = [ctypes.POINTER(Options)]
        L.mhs_default_options.restype = None
        L.mhs_create.argtypes = [ctypes.POINTER(_P), ctypes.c_int, _P]
        L.mhs_destroy.argtypes = [_P]
        L.mhs_last_error.argtypes = [_P]
        L.mhs_last_error.restype = ctypes.c_char_p
        L.mhs_clear.argtypes = [_P]
        L.mhs_set_grid.argtypes = [_P, ctypes.c_int, _I, ctypes.c_int, _I, _I, _D, _D, _D, _D, _D]
        L.mhs_set_cell_ijk.argtypes = [_P, _I]
        L.mhs_init_system_structure.argtypes = [_P, _I, _I]
        L.mhs_compute_inner_products.argtypes = [_P, _D, _D]
        L.mhs_assemble_static.argtypes = [_P, _D, _D]
        L.mhs_get_inner_products.argtypes = [_P, _D, _D]
        L.mhs_solve.argtypes = [_P, _D, _D, _D, _I, _D, ctypes.POINTER(Options),
                                ctypes.POINTER(Report)]
        L.mhs_solve_device.argtypes = [_P, _P, _P, _P, _I, _P, ctypes.POINTER(Options),
                                       ctypes.POINTER(Report)]
        L.mhs_get_solution.argtypes = [_P, _D, _D, _D]
        L.mhs_solution_device.argtypes = [_P, ctypes.POINTER(_P), ctypes.POINTER(_P),
                                          ctypes.POINTER(_P)]
        L.mhs_post_process_fluxes.argtypes = [_P, _D]
        L.mhs_get_dims.argtypes = [_P, ctypes.POINTER(Dims)]
        L.mhs_get_cell_faces.argtypes = [_P, _I, _I]
        L.mhs_get_face_rows.argtypes = [_P, _I]
        L.mhs_get_system.argtypes = [_P, ctypes.c_int, _I, _I, _D, _D,
                                     ctypes.POINTER(ctypes.c_longlong)]
        L.mhs_time_kernel.argtypes = [_P, ctypes.c_int, ctypes.c_int, _D]
        L.mhs_compute_inner_products_device.argtypes = [_P, _P, _D]
        L.mhs_get_solution_device.argtypes = [_P, _P, _P]
        L.mhs_sum_outflux.argtypes = [_P, ctypes.c_int, _I, _I, _D]
        L.mhs_sum_outflux_noalloc.argtypes = [_P, ctypes.c_int, _I, _I, _D]
        L.mhs_gravity_term.argtypes = [_P, _D, _D, _D]
        L.mhs_topology_is_sane.argtypes = [ctypes.c_int, _I, _I, ctypes.POINTER(ctypes.c_int)]
        L.mhs_estimate_cell_velocity.argtypes = [_P, ctypes.c_int, _D]
        L.mhs_solve_aniso.argtypes = [_P, _D, _D, _D, _I, _D, ctypes.POINTER(Options), ctypes.POINTER(Report)]
        L.mhs_solve_aniso_device.argtypes = [_P, _P, _P, _P, _I, _P, ctypes.POINTER(Options),
                                             ctypes.POINTER(Report)]
        L.mhs_estimate_cell_velocity_device.argtypes = [_P, ctypes.c_int, _P]
        L.mhs_adjoint.argtypes = [_P, _D, ctypes.POINTER(Options), _D, _D, _D, _D,
                                  ctypes.POINTER(Report)]
        L.mhs_adjoint_device.argtypes = [_P, _P, ctypes.POINTER(Options), _P, _P, _P, _P,
                                         ctypes.POINTER(Report)]
        for name in EXPORTS:
            if name not in ("mhs_default_options", "mhs_last_error"):
                getattr(L, name).restype = ctypes.c_int
        _LIB = L
    return _LIB


class MhsError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_I)


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_D)


def topology_is_sane(g):
    """topologyIsSane (IncompFlowSolverHybrid.hpp:108-130): no two cells share
    more than one face.  Host check through the C-ABI, no GPU needed."""
    fp, pfp = _i(g.facepos)
    nb, pnb = _i(g.nb)
    ok = ctypes.c_int(-1)
    rc = lib().mhs_topology_is_sane(int(g.num_cells), pfp, pnb, ctypes.byref(ok))
    if rc != 0:
        raise MhsError(rc, "topology_is_sane: invalid arguments")
    return bool(ok.value)


def _solve_kw(kw):
    """solve()'s keyword names mapped onto options()'s."""
    names = {"residual_tolerance": "tol", "linsolver_verbosity": "linsolver_verbosity",
             "linsolver_maxit": "linsolver_maxit"}
    return {names.get(k, k): v for k, v in kw.items()}


class HybridSolver:
    """Mimetic hybrid face-pressure solver on one GPU (a handle over the C-ABI)."""

    def __init__(self, device: int = 0, stream: int | None = None):
        self._L = lib()
        h = _P()
        rc = self._L.mhs_create(ctypes.byref(h), int(device), _P(stream or 0))
        if rc != MHS_OK:
            raise MhsError(rc, "mhs_create failed (no usable GPU?)")
        self._h = h
        self.stream = stream
        self.nrhs = 1
        self.grid = None
        self.report = None
        self.solve_count = 0

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self._L.mhs_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def _check(self, rc):
        if rc != MHS_OK:
            msg = self._L.mhs_last_error(self._h).decode()
            raise MhsError(rc, msg)

    # --- set-up -------------------------------------------------------------------
    def clear(self):
        self._check(self._L.mhs_clear(self._h))
        self.grid = None

    def set_grid(self, g, inst_cellpos=None):
        keep = []
        fp, pfp = _i(g.facepos); nb, pnb = _i(g.nb)
        cv, pcv = _d(g.cell_volume); cc, pcc = _d(g.cell_centroid)
        fa, pfa = _d(g.face_area); fc, pfc = _d(g.face_centroid); fn, pfn = _d(g.face_normal)
        keep += [fp, nb, cv, cc, fa, fc, fn]
        if inst_cellpos is None:
            ninst, pic = 1, None
        else:
            ic, pic = _i(inst_cellpos)
            keep.append(ic)
            ninst = len(ic) - 1
        self._check(self._L.mhs_set_grid(self._h, ninst, pic, g.num_cells, pfp, pnb, pcv, pcc, pfa,
                                         pfc, pfn))
        self.grid = g
        self.ninst = ninst

    def set_cell_ijk(self, ijk):
        """Logical cell indices [nc, 3] (enables the multigrid preconditioner)."""
        if ijk is None:
            self._check(self._L.mhs_set_cell_ijk(self._h, None))
            return
        a, pa = _i(np.asarray(ijk).reshape(-1))
        self._check(self._L.mhs_set_cell_ijk(self._h, pa))

    def initSystemStructure(self, bc_type=None, periodic_partner=None):
        if bc_type is None or periodic_partner is None:
            self._check(self._L.mhs_init_system_structure(self._h, None, None))
        else:
            bt, pbt = _i(bc_type); pp, ppp = _i(periodic_partner)
            self._check(self._L.mhs_init_system_structure(self._h, pbt, ppp))

    def computeInnerProducts(self, perm, gravity=(0.0, 0.0, 0.0)):
        K, pK = _d(np.asarray(perm).reshape(-1))
        gr, pg = _d(gravity)
        self._check(self._L.mhs_compute_inner_products(self._h, pK, pg))

    assembleStatic = computeInnerProducts

    def computeInnerProducts_device(self, perm_ptr, gravity=(0.0, 0.0, 0.0)):
        """computeInnerProducts with perm[nc*9] in device memory (a data_ptr())."""
        gr, pg = _d(gravity)
        self._check(self._L.mhs_compute_inner_products_device(self._h, _P(perm_ptr), pg))

    def gravityTerm(self, gravity, omega):
        """MimeticIPEvaluator::gravityTerm (MimeticIPEvaluator.hpp:477-523) for every cell:
        omega[c] (x_f - x_c) . g per half-face, computed on the device."""
        g = self.grid
        om, pom = _d(np.broadcast_to(np.asarray(omega, dtype=np.float64), (g.num_cells,)))
        gr, pg = _d(gravity)
        out = np.empty(g.num_half_faces)
        self._check(self._L.mhs_gravity_term(self._h, pg, pom, out.ctypes.data_as(_D)))
        return out

    def inner_products(self):
        """(binv, gflux): per-cell B^-1 column-major, concatenated; N K g per half-face."""
        g = self.grid
        n = g.ncf().astype(np.int64)
        binv = np.empty(int(np.sum(n * n)))
        gfl = np.empty(g.num_half_faces)
        self._check(self._L.mhs_get_inner_products(self._h, binv.ctypes.data_as(_D),
                                                   gfl.ctypes.data_as(_D)))
        return binv, gfl

    def init(self, g, perm, gravity=(0.0, 0.0, 0.0), bc_type=None, periodic_partner=None,
             inst_cellpos=None, ijk="auto"):
        """IncompFlowSolverHybrid::init (:476-488).  ijk: logical cell indices
        [nc, 3] for the multigrid preconditioner ("auto": from the grid, if any)."""
        self.set_grid(g, inst_cellpos)
        if isinstance(ijk, str):
            ijk = g.ijk if getattr(g, "ijk", None) is not None else None
        if ijk is not None:
            self.set_cell_ijk(ijk)
        if g.num_cells > 0:
            self.initSystemStructure(bc_type, periodic_partner)
            self.computeInnerProducts(perm, gravity)

    # --- solve ----------------------------------------------------------------------
    def options(self, tol=1e-8, linsolver_verbosity=0, linsolver_type=1, same_matrix=False,
                linsolver_maxit=0, prolongate_factor=1.6, smooth_steps=1, precision=FP64,
                nrhs=1, chunk=32, replace_factor=0.01, use_graph=True, precond=-1,
                smoother_weight=None, autotune=True, spmv_threads=0, sweep_threads=0,
                tail_rows=-1, element_operator=True):
        o = Options()
        self._L.mhs_default_options(ctypes.byref(o))
        o.tol, o.maxit, o.linsolver_type, o.verbosity = tol, linsolver_maxit, linsolver_type, \
            linsolver_verbosity
        o.same_matrix, o.prolongate_factor, o.smooth_steps = int(same_matrix), \
            prolongate_factor, smooth_steps
        o.precision = {"fp64": FP64, "mixed": MIXED}.get(precision, precision)
        o.nrhs, o.chunk, o.replace_factor, o.use_graph = nrhs, chunk, replace_factor, \
            int(use_graph)
        o.precond = {"auto": -1, "jacobi": 0, "amg": 1}.get(precond, precond)
        if smoother_weight is not None:
            o.smoother_weight = smoother_weight
        o.autotune = int(autotune)
        o.spmv_threads, o.sweep_threads = int(spmv_threads), int(sweep_threads)
        o.tail_rows = int(tail_rows)
        o.element_operator = int(element_operator)
        return o

    def solve(self, totmob, rhomob, src, bc_type, bc_value, residual_tolerance=1e-8,
              linsolver_verbosity=0, linsolver_type=1, same_matrix=False, linsolver_maxit=0,
              prolongate_factor=1.6, smooth_steps=1, precision="fp64", chunk=32,
              replace_factor=0.01, use_graph=True, raise_on_failure=True, precond=-1,
              smoother_weight=None, autotune=True, spmv_threads=0, sweep_threads=0,
              tail_rows=-1, element_operator=True):
        """IncompFlowSolverHybrid::solve (:657-703) with per-cell mobilities."""
        g = self.grid
        src = np.asarray(src, dtype=np.float64)
        nrhs = 1 if src.ndim == 1 else src.shape[0]
        bv = np.asarray(bc_value, dtype=np.float64).reshape(nrhs, g.num_half_faces)
        tm, ptm = _d(totmob); rm, prm = _d(rhomob); sr, psr = _d(src.reshape(nrhs, -1))
        bt, pbt = _i(bc_type); bvv, pbv = _d(bv)
        o = self.options(residual_tolerance, linsolver_verbosity, linsolver_type, same_matrix,
                         linsolver_maxit, prolongate_factor, smooth_steps, precision, nrhs, chunk,
                         replace_factor, use_graph, precond, smoother_weight, autotune,
                         spmv_threads, sweep_threads, tail_rows, element_operator)
        rep = Report()
        rc = self._L.mhs_solve(self._h, ptm, prm, psr, pbt, pbv, ctypes.byref(o),
                               ctypes.byref(rep))
        self.report = rep
        self.nrhs = nrhs
        self.solve_count += 1
        if rc == MHS_ERR_NOT_CONVERGED and not raise_on_failure:
            return rep
        self._check(rc)
        return rep

    def solve_aniso(self, lambda_t, rho_lambda, src, bc_type, bc_value, **kw):
        """solve() with MimeticIPAnisoRelpermEvaluator: per-cell mobility tensors
        lambda_t [nc, 9] (total) and rho_lambda [nc, 9] (sum rho_i lambda_i),
        row-major like the permeability.  Keyword options as solve()."""
        g = self.grid
        src = np.asarray(src, dtype=np.float64)
        nrhs = 1 if src.ndim == 1 else src.shape[0]
        bv = np.asarray(bc_value, dtype=np.float64).reshape(nrhs, g.num_half_faces)
        lt, plt = _d(np.asarray(lambda_t, dtype=np.float64).reshape(-1))
        rl, prl = _d(np.asarray(rho_lambda, dtype=np.float64).reshape(-1))
        sr, psr = _d(src.reshape(nrhs, -1))
        bt, pbt = _i(bc_type); bvv, pbv = _d(bv)
        o = self.options(nrhs=nrhs, **_solve_kw(kw))
        rep = Report()
        rc = self._L.mhs_solve_aniso(self._h, plt, prl, psr, pbt, pbv, ctypes.byref(o), ctypes.byref(rep))
        self.report = rep
        self.nrhs = nrhs
        self.solve_count += 1
        self._check(rc)
        return rep

    def solve_aniso_device(self, lambda_ptr, rho_lambda_ptr, src_ptr, bc_type, bcval_ptr, nrhs=1, **kw):
        bt, pbt = _i(bc_type)
        o = self.options(nrhs=nrhs, **_solve_kw(kw))
        rep = Report()
        rc = self._L.mhs_solve_aniso_device(self._h, _P(lambda_ptr), _P(rho_lambda_ptr), _P(src_ptr), pbt,
                                            _P(bcval_ptr), ctypes.byref(o), ctypes.byref(rep))
        self.report = rep
        self.nrhs = nrhs
        self.solve_count += 1
        self._check(rc)
        return rep

    def solve_device(self, totmob_ptr, rhomob_ptr, src_ptr, bc_type, bcval_ptr, nrhs=1, **kw):
        """Same as solve with device pointers (e.g. torch tensor data_ptr())."""
        bt, pbt = _i(bc_type)
        o = self.options(nrhs=nrhs, **kw)
        rep = Report()
        rc = self._L.mhs_solve_device(self._h, _P(totmob_ptr), _P(rhomob_ptr), _P(src_ptr), pbt,
                                      _P(bcval_ptr), ctypes.byref(o), ctypes.byref(rep))
        self.report = rep
        self.nrhs = nrhs
        self.solve_count += 1
        self._check(rc)
        return rep

    def getSolution(self):
        """(cell_pressure[nrhs,nc], face_pressure[nrhs,nf], outflux[nrhs,nhf])."""
        d = self.dims()
        p = np.empty((self.nrhs, d.num_cells))
        pi = np.empty((self.nrhs, d.num_faces))
        v = np.empty((self.nrhs, d.num_half_faces))
        self._check(self._L.mhs_get_solution(self._h, p.ctypes.data_as(_D), pi.ctypes.data_as(_D),
                                             v.ctypes.data_as(_D)))
        if self.nrhs == 1:
            return p[0], pi[0], v[0]
        return p, pi, v

    def solution_device(self):
        a, b, c = _P(), _P(), _P()
        self._check(self._L.mhs_solution_device(self._h, ctypes.byref(a), ctypes.byref(b),
                                                ctypes.byref(c)))
        return a.value, b.value, c.value

    def get_solution_device(self, cell_pressure_ptr=None, outflux_ptr=None):
        """Copy the last cell pressures / outfluxes into caller device buffers."""
        self._check(self._L.mhs_get_solution_device(self._h, _P(cell_pressure_ptr or 0),
                                                    _P(outflux_ptr or 0)))

    def sum_outflux(self, group_ptr, group_hf):
        """Sums of the last outflux over half-face groups: [nrhs, ngroups]."""
        gp, pgp = _i(group_ptr)
        gh, pgh = _i(group_hf)
        ng = gp.shape[0] - 1
        out = np.empty((self.nrhs, ng))
        self._check(self._L.mhs_sum_outflux(self._h, ng, pgp, pgh, out.ctypes.data_as(_D)))
        return out

    def sum_outflux_noalloc(self, group_ptr, group_hf):
        """sum_outflux through the entry point that allocates no device memory
        (its staging lives in the library's per-GPU scratch array)."""
        gp, pgp = _i(group_ptr)
        gh, pgh = _i(group_hf)
        ng = gp.shape[0] - 1
        out = np.empty((self.nrhs, ng))
        self._check(self._L.mhs_sum_outflux_noalloc(self._h, ng, pgp, pgh, out.ctypes.data_as(_D)))
        return out

    # --- adjoint --------------------------------------------------------------------
    def adjoint(self, dJ_dv, tol=1e-10, **kw):
        """Gradients of J through the last solve, given dJ/dv per half-face
        ([nrhs, nhf] or [nhf]).  Returns a dict with grad_perm [nc, 9],
        grad_totmob [nc], grad_rhomob [nc] and grad_src ([nrhs, nc] or [nc])."""
        d = self.dims()
        w, pw = _d(np.asarray(dJ_dv).reshape(-1))
        if w.shape[0] != self.nrhs * d.num_half_faces:
            raise ValueError("dJ_dv must have nrhs * num_half_faces entries")
        o = self.options(tol=tol, nrhs=self.nrhs, **kw)
        gK = np.empty((d.num_cells, 9))
        gt = np.empty(d.num_cells)
        gr = np.empty(d.num_cells)
        gs = np.empty((self.nrhs, d.num_cells))
        rep = Report()
        rc = self._L.mhs_adjoint(self._h, pw, ctypes.byref(o), gK.ctypes.data_as(_D),
                                 gt.ctypes.data_as(_D), gr.ctypes.data_as(_D),
                                 gs.ctypes.data_as(_D), ctypes.byref(rep))
        self.adjoint_report = rep
        self._check(rc)
        return {"grad_perm": gK, "grad_totmob": gt, "grad_rhomob": gr,
                "grad_src": gs[0] if self.nrhs == 1 else gs}

    def adjoint_device(self, dJ_dv_ptr, grad_perm_ptr=None, grad_totmob_ptr=None,
                       grad_rhomob_ptr=None, grad_src_ptr=None, tol=1e-10, **kw):
        """adjoint() with device pointers for dJ/dv and every output."""
        o = self.options(tol=tol, nrhs=self.nrhs, **kw)
        rep = Report()
        rc = self._L.mhs_adjoint_device(self._h, _P(dJ_dv_ptr), ctypes.byref(o),
                                        _P(grad_perm_ptr or 0), _P(grad_totmob_ptr or 0),
                                        _P(grad_rhomob_ptr or 0), _P(grad_src_ptr or 0),
                                        ctypes.byref(rep))
        self.adjoint_report = rep
        self._check(rc)
        return rep

    def estimateCellVelocity(self, k: int = 0) -> np.ndarray:
        """estimateCellVelocity (common/SimulatorUtilities.hpp:59-81) of the last
        solution: [nc, 3], sum over faces of (x_f - x_c) outflux / |c|."""
        v = np.zeros((self.grid.num_cells, 3))
        self._check(self._L.mhs_estimate_cell_velocity(self._h, int(k), v.ctypes.data_as(_D)))
        return v

    def estimate_cell_velocity_device(self, out_ptr: int, k: int = 0):
        self._check(self._L.mhs_estimate_cell_velocity_device(self._h, int(k), _P(out_ptr)))

    def postProcessFluxes(self) -> float:
        mm = ctypes.c_double(0.0)
        self._check(self._L.mhs_post_process_fluxes(self._h, ctypes.byref(mm)))
        return mm.value

    # --- introspection --------------------------------------------------------------------
    def dims(self) -> Dims:
        d = Dims()
        self._check(self._L.mhs_get_dims(self._h, ctypes.byref(d)))
        return d

    def cell_faces(self):
        d = self.dims()
        hf = np.empty(d.num_half_faces, dtype=np.int32)
        ifp = np.empty(d.num_instances + 1, dtype=np.int32)
        self._check(self._L.mhs_get_cell_faces(self._h, hf.ctypes.data_as(_I),
                                               ifp.ctypes.data_as(_I)))
        return hf, ifp

    def face_rows(self):
        d = self.dims()
        fr = np.empty(d.num_faces, dtype=np.int32)
        self._check(self._L.mhs_get_face_rows(self._h, fr.ctypes.data_as(_I)))
        return fr

    def time_kernel(self, which=0, reps=50) -> float:
        """Average device milliseconds of one launch of kernel `which` (0: PCG SpMV)."""
        ms = ctypes.c_double(0.0)
        self._check(self._L.mhs_time_kernel(self._h, which, reps, ctypes.byref(ms)))
        return ms.value

    def system(self, k=0):
        """(row_ptr, col, val, rhs) of the assembled face system (printSystem :895-905)."""
        d = self.dims()
        nnz = ctypes.c_longlong(0)
        self._check(self._L.mhs_get_system(self._h, k, None, None, None, None, ctypes.byref(nnz)))
        rp = np.empty(d.num_faces + 1, dtype=np.int32)
        col = np.empty(nnz.value, dtype=np.int32)
        val = np.empty(nnz.value)
        rhs = np.empty(d.num_faces)
        self._check(self._L.mhs_get_system(self._h, k, rp.ctypes.data_as(_I),
                                           col.ctypes.data_as(_I), val.ctypes.data_as(_D),
                                           rhs.ctypes.data_as(_D), ctypes.byref(nnz)))
        return rp, col, val, rhs

"""ctypes binding of the explicit upstream transport (include/euler_upstream.h).

``EulerUpstream`` mirrors opm-porsol's EulerUpstream<GI, RP, BC>
(opm/porsol/euler/EulerUpstream.hpp): init(param) :95-108, initObj :120-128,
setCourantNumber :143-147, transportSolve :153-215, with the residual of
EulerUpstreamResidual::computeResidual (EulerUpstreamResidual_impl.hpp:474-507).
It runs on the GPU of the ``HybridSolver`` it is bound to and reads that
solver's last fluxes.  A saturation that leaves [-0.001, 1.001] after every
retry raises RuntimeError, as the reference rethrows (:198-205).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import mhs

_P, _D = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)

EXPORTS = (
    "mhs_transport_default_options", "mhs_transport_create", "mhs_transport_destroy",
    "mhs_transport_last_error", "mhs_transport_init", "mhs_transport_solve",
    "mhs_transport_solve_device", "mhs_transport_residual", "mhs_transport_mobilities_device",
    "mhs_transport_time_step",
)

# ReservoirPropertyCommon defaults (common/ReservoirPropertyCommon_impl.hpp:257-262)
DEFAULT_VISCOSITIES = (1.0e-3, 3.0e-3)
DEFAULT_DENSITIES = (1013.9, 834.7)


class Options(ctypes.Structure):
    _fields_ = [("courant_number", ctypes.c_double), ("method_viscous", ctypes.c_int),
                ("method_gravity", ctypes.c_int), ("method_capillary", ctypes.c_int),
                ("use_cfl_viscous", ctypes.c_int), ("use_cfl_gravity", ctypes.c_int),
                ("use_cfl_capillary", ctypes.c_int), ("minimum_small_steps", ctypes.c_int),
                ("maximum_small_steps", ctypes.c_int), ("check_sat", ctypes.c_int),
                ("clamp_sat", ctypes.c_int)]


class Report(ctypes.Structure):
    _fields_ = [("small_steps", ctypes.c_int), ("repeats", ctypes.c_int), ("cfl_dt", ctypes.c_double),
                ("dt", ctypes.c_double), ("cfl_factor", ctypes.c_double),
                ("cfl_factor_gravity", ctypes.c_double), ("cfl_factor_capillary", ctypes.c_double),
                ("seconds", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Fluid(ctypes.Structure):
    _fields_ = [("viscosity", ctypes.c_double * 2), ("density", ctypes.c_double * 2)]


_BOUND = False


def lib():
    global _BOUND
    L = mhs.lib()
    if not _BOUND:
        L.mhs_transport_default_options.argtypes = [ctypes.POINTER(Options)]
        L.mhs_transport_default_options.restype = None
        L.mhs_transport_create.argtypes = [ctypes.POINTER(_P), _P]
        L.mhs_transport_destroy.argtypes = [_P]
        L.mhs_transport_last_error.argtypes = [_P]
        L.mhs_transport_last_error.restype = ctypes.c_char_p
        L.mhs_transport_init.argtypes = [_P, _D, _D, ctypes.POINTER(Fluid)]
        L.mhs_transport_solve.argtypes = [_P, _D, ctypes.c_double, _D, _D, ctypes.POINTER(Options),
                                          ctypes.POINTER(Report)]
        L.mhs_transport_solve_device.argtypes = [_P, _P, ctypes.c_double, _D, _P, ctypes.POINTER(Options),
                                                 ctypes.POINTER(Report)]
        L.mhs_transport_residual.argtypes = [_P, _D, _D, _D, ctypes.POINTER(Options), _D]
        L.mhs_transport_mobilities_device.argtypes = [_P, _P, _P, _P]
        L.mhs_transport_time_step.argtypes = [_P, ctypes.c_int, _D]
        for name in EXPORTS:
            if name not in ("mhs_transport_default_options", "mhs_transport_last_error"):
                getattr(L, name).restype = ctypes.c_int
        _BOUND = True
    return L


def default_options() -> Options:
    o = Options()
    lib().mhs_transport_default_options(ctypes.byref(o))
    return o


class EulerUpstream:
    """Explicit upstream saturation transport on the GPU of ``flow``."""

    _PARAMS = ("courant_number", "method_viscous", "method_gravity", "method_capillary",
               "use_cfl_viscous", "use_cfl_gravity", "use_cfl_capillary", "minimum_small_steps",
               "maximum_small_steps", "check_sat", "clamp_sat")

    def __init__(self, flow: "mhs.HybridSolver", porosity=None, sat_bc=None,
                 viscosities=DEFAULT_VISCOSITIES, densities=DEFAULT_DENSITIES):
        self._L = lib()
        self.flow = flow
        h = _P()
        rc = self._L.mhs_transport_create(ctypes.byref(h), flow._h)
        if rc != mhs.MHS_OK:
            raise mhs.MhsError(rc, "mhs_transport_create failed")
        self._h = h
        self.opt = default_options()
        self.report = None
        if porosity is not None:
            self.initObj(porosity, sat_bc, viscosities, densities)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self._L.mhs_transport_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def _check(self, rc):
        if rc != mhs.MHS_OK:
            msg = self._L.mhs_transport_last_error(self._h).decode()
            raise mhs.MhsError(rc, msg)

    def init(self, param: dict):
        """EulerUpstream::init(param): the parameters the reference reads, by name."""
        for k, v in param.items():
            if k not in self._PARAMS:
                continue
            setattr(self.opt, k, float(v) if k == "courant_number" else int(bool(v)) if k.startswith(
                ("method_", "use_cfl_", "check_", "clamp_")) else int(v))

    def initObj(self, porosity, sat_bc=None, viscosities=DEFAULT_VISCOSITIES, densities=DEFAULT_DENSITIES):
        poro, pp = mhs._d(porosity)
        if sat_bc is not None:
            sbc, ps = mhs._d(sat_bc)
        else:
            ps = None
        fl = Fluid((ctypes.c_double * 2)(*viscosities), (ctypes.c_double * 2)(*densities))
        self._check(self._L.mhs_transport_init(self._h, pp, ps, ctypes.byref(fl)))

    def setCourantNumber(self, cn: float):
        self.opt.courant_number = float(cn)

    def transportSolve(self, saturation: np.ndarray, time: float, gravity=(0.0, 0.0, 0.0),
                       injection_rates=None) -> np.ndarray:
        """Advances ``saturation`` (float64, updated in place) over ``time``."""
        if not (isinstance(saturation, np.ndarray) and saturation.dtype == np.float64
                and saturation.flags.c_contiguous):
            raise TypeError("saturation must be a contiguous float64 array (updated in place)")
        g, pg = mhs._d(np.asarray(gravity, dtype=np.float64).reshape(3))
        pi = None
        if injection_rates is not None:
            inj, pi = mhs._d(injection_rates)
        rep = Report()
        rc = self._L.mhs_transport_solve(self._h, saturation.ctypes.data_as(_D), float(time), pg, pi,
                                         ctypes.byref(self.opt), ctypes.byref(rep))
        self.report = rep
        if rc == mhs.MHS_ERR_RANGE:
            raise RuntimeError(self._L.mhs_transport_last_error(self._h).decode())
        self._check(rc)
        return saturation

    def transportSolve_device(self, saturation_ptr: int, time: float, gravity=(0.0, 0.0, 0.0),
                              injection_ptr: int | None = None):
        g, pg = mhs._d(np.asarray(gravity, dtype=np.float64).reshape(3))
        rep = Report()
        rc = self._L.mhs_transport_solve_device(self._h, _P(saturation_ptr), float(time), pg,
                                                _P(injection_ptr or 0), ctypes.byref(self.opt),
                                                ctypes.byref(rep))
        self.report = rep
        if rc == mhs.MHS_ERR_RANGE:
            raise RuntimeError(self._L.mhs_transport_last_error(self._h).decode())
        self._check(rc)

    def computeResidual(self, saturation, gravity=(0.0, 0.0, 0.0), injection_rates=None,
                        method_viscous=True, method_gravity=True, method_capillary=True) -> np.ndarray:
        s, ps = mhs._d(saturation)
        g, pg = mhs._d(np.asarray(gravity, dtype=np.float64).reshape(3))
        pi = None
        if injection_rates is not None:
            inj, pi = mhs._d(injection_rates)
        o = Options.from_buffer_copy(self.opt)
        o.method_viscous, o.method_gravity, o.method_capillary = (int(bool(method_viscous)),
                                                                 int(bool(method_gravity)),
                                                                 int(bool(method_capillary)))
        res = np.zeros(s.shape[0])
        self._check(self._L.mhs_transport_residual(self._h, ps, pg, pi, ctypes.byref(o), res.ctypes.data_as(_D)))
        return res

    def mobilities_device(self, saturation_ptr: int, totmob_ptr: int | None, rhomob_ptr: int | None):
        """Total mobility and sum rho_i lambda_i per cell (device pointers), the
        per-cell inputs of the next pressure solve."""
        self._check(self._L.mhs_transport_mobilities_device(self._h, _P(saturation_ptr), _P(totmob_ptr or 0),
                                                            _P(rhomob_ptr or 0)))

    def time_step(self, reps: int = 50) -> float:
        """Average device time (ms) of one small-step kernel launch."""
        ms = ctypes.c_double(0.0)
        self._check(self._L.mhs_transport_time_step(self._h, int(reps), ctypes.byref(ms)))
        return ms.value

"""CPU restatement of opm-porsol's explicit upstream saturation transport.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline of bench.py as the checker of the HIP transport kernels, never by
the product path.

Follows, for the fluid without rock tables (quadratic relative
permeabilities, the reference's fallback when no rock files are given):
  - ReservoirPropertyCapillary::relPermFirstPhase / relPermSecondPhase
    (common/ReservoirPropertyCapillary_impl.hpp:118-170): krw = s^2,
    kro = (1-s)^2, mobilities kr/mu (:44-55);
  - ReservoirPropertyCommon::capillaryPressure (common/
    ReservoirPropertyCommon_impl.hpp:466-476): p_c = 1e5 (1 - s);
  - computeSingleRockCflFactors(-1, ...) (ReservoirPropertyCapillary_impl.hpp:
    215-250) through computeCflFactors (:254-260): 257 samples of f_w and
    lambda_w lambda_o / lambda_t, the capillary factor is 0;
  - EulerUpstreamResidual UpdateForCell::operator() (euler/
    EulerUpstreamResidual_impl.hpp:101-300) and estimateCapPressureGradient
    (:513-550): per face, from the lower-numbered cell, upstream mobilities
    by the viscous flux and the gravity influence, then the viscous, gravity
    and capillary saturation fluxes; sources with the fractional flow of the
    cell on production;
  - cfl_calculator::findCFLtimeVelocity / Gravity / Capillary (euler/
    CflCalculator.hpp:55-170);
  - EulerUpstream::transportSolve (euler/EulerUpstream_impl.hpp:153-215),
    computeCflTime (:265-330), checkAndPossiblyClampSat (:337-350),
    smallTimeStep (:357-385).
As in the reference, method_gravity zeroes G (which picks the phase whose
flux follows the total flux) and the gravity flux, while the gravity
influence still enters the other phase's upstream direction (:198-214).
Opm::utils::arithmeticAverage (opm-core, not in the tree) is the mean
(a + b) / 2.  The per-face loop is vectorised over the faces each cell owns;
the arithmetic per face is the reference's.
"""
from __future__ import annotations

import math

import numpy as np

CAP_SCALE = 1e5


def mobilities(s, visc):
    """(lambda_w, lambda_o) of the fluid without rock tables."""
    s = np.asarray(s, dtype=np.float64)
    return s * s / visc[0], (1.0 - s) * (1.0 - s) / visc[1]


def fractional_flow(s, visc):
    lw, lo = mobilities(s, visc)
    return lw / (lw + lo)


def cap_pressure(s):
    return CAP_SCALE * (1.0 - np.asarray(s, dtype=np.float64))


def cfl_factors(visc):
    """(cflFactor, cflFactorGravity, cflFactorCapillary) without rock tables."""
    n = 257
    delta = 1.0 / (n - 1)

    def ff(s):
        l1 = s * s / visc[0]
        l2 = (1.0 - s) * (1.0 - s) / visc[1]
        return l1 / (l1 + l2), l1 * l2 / (l1 + l2)

    last1, lastg = ff(0.0)
    max_der1 = -1e100
    max_derg = -1e100
    max_ffg = lastg
    for i in range(1, n):
        s = float(i) * delta
        f1, fg = ff(s)
        max_der1 = max(max_der1, abs(f1 - last1) / delta)
        max_derg = max(max_derg, abs(fg - lastg) / delta)
        max_ffg = max(max_ffg, fg)
        last1, lastg = f1, fg
    return 1.0 / max_der1, 1.0 / max_derg, max_ffg * 0.0


def face_links(g, partner=None):
    """Per half-face: the cell on the other side (-1: non-periodic boundary)
    and the half-face there (the same half-face for internal faces, whose
    'neighbour face' is the face itself in the reference; the partner for
    periodic faces)."""
    nhf = g.num_half_faces
    hc = g.hf_cell()
    nb = np.asarray(g.nb)
    other = np.where(nb >= 0, nb, -1).astype(np.int64)
    nbface = np.arange(nhf, dtype=np.int64)
    periodic = np.zeros(nhf, dtype=bool)
    if partner is not None:
        partner = np.asarray(partner)
        per = (nb < 0) & (partner >= 0)
        other[per] = hc[partner[per]]
        nbface[per] = partner[per]
        periodic = per
    if np.any(other == hc):
        raise ValueError("a face links a cell to itself")
    return other, nbface, periodic


def residual(g, perm, sat, outflux, gravity, drho, visc, inj, sat_bc, partner=None,
             method_viscous=True, method_gravity=True, method_capillary=True):
    """EulerUpstreamResidual::computeResidual for the fluid without rock tables."""
    nc = g.num_cells
    K = np.asarray(perm, dtype=np.float64).reshape(nc, 3, 3)
    sat = np.asarray(sat, dtype=np.float64)
    hc = g.hf_cell().astype(np.int64)
    other, nbface, _ = face_links(g, partner)
    bnd = other < 0
    # the faces a cell handles: boundary, or a lower number than the other side
    own = bnd | (hc < other)
    o = np.nonzero(own)[0]
    c0 = hc[o]
    c1 = np.where(bnd[o], c0, other[o])
    s0 = sat[c0]
    s1 = np.where(bnd[o], np.asarray(sat_bc, dtype=np.float64)[o], sat[c1])
    area = np.asarray(g.face_area)[o]
    n = np.asarray(g.face_normal)[o]
    flux = np.asarray(outflux, dtype=np.float64)[o]
    Kav = 0.5 * (K[c0] + K[c1])
    gi = drho * np.einsum("fij,j->fi", Kav, np.asarray(gravity, dtype=np.float64))
    G = area * np.einsum("fi,fi->f", n, gi) if method_gravity else np.zeros(len(o))
    triv = np.where(G >= 0.0, 0, 1)
    ups0 = flux >= 0.0
    s_up = np.where(ups0, s0, s1)
    lw_up, lo_up = mobilities(s_up, visc)
    m_triv = np.where(triv == 0, lw_up, lo_up)
    sign = np.where(triv == 0, -1.0, 1.0)
    grav_nontriv = sign * area * np.einsum("fi,fi->f", n, m_triv[:, None] * gi)
    s_up2 = np.where(flux + grav_nontriv >= 0.0, s0, s1)
    lw_up2, lo_up2 = mobilities(s_up2, visc)
    m_w = np.where(triv == 0, m_triv, lw_up2)
    m_o = np.where(triv == 0, lo_up2, m_triv)
    m_totinv = 1.0 / (m_w + m_o)
    dS = np.zeros(len(o))
    if method_viscous:
        v = n * flux[:, None]
        dS += np.einsum("fi,fi->f", n, m_w[:, None] * (m_totinv[:, None] * v))
    if method_gravity:
        gc = area * np.einsum("fi,fi->f", n, m_w[:, None] * (m_totinv[:, None] * (m_o[:, None] * gi)))
        dS += np.where(bnd[o], 0.0, gc)
    if method_capillary:
        s_av = 0.5 * (s0 + s1)
        a_w, a_o = mobilities(s_av, visc)
        a_totinv = 1.0 / (a_w + a_o)
        ccen = np.asarray(g.cell_centroid)
        fcen = np.asarray(g.face_centroid)
        cc, nbc = ccen[c0], ccen[c1]
        fc, nbfc = fcen[o], fcen[nbface[o]]
        d0 = np.linalg.norm(cc - fc, axis=1)
        d1 = np.linalg.norm(nbc - nbfc, axis=1)
        val = (cap_pressure(s1) - cap_pressure(s0)) / (d0 + d1)
        d = nbc - nbfc + fc - cc
        with np.errstate(invalid="ignore", divide="ignore"):
            grad = d / np.linalg.norm(d, axis=1)[:, None] * val[:, None]
        grad[bnd[o]] = 0.0
        capinf = np.einsum("fij,fj->fi", Kav, grad)
        dS += area * np.einsum("fi,fi->f", n, a_w[:, None] * (a_totinv[:, None] * (a_o[:, None] * capinf)))
    res = np.zeros(nc)
    np.add.at(res, c0, -dS)
    inner = ~bnd[o]
    np.add.at(res, c1[inner], dS[inner])
    rate = np.zeros(nc) if inj is None else np.asarray(inj, dtype=np.float64).copy()
    neg = rate < 0.0
    rate[neg] *= fractional_flow(sat[neg], visc)
    return res + rate


def cfl_time(g, perm, poro, outflux, gravity, drho, visc, courant_number=0.5,
             use_viscous=True, use_gravity=True, use_capillary=True):
    """EulerUpstream::computeCflTime: courant_number times the smallest of the
    viscous, gravity and capillary CFL times (1e100 when not used)."""
    nc = g.num_cells
    K = np.asarray(perm, dtype=np.float64).reshape(nc, 3, 3)
    fac, fac_g, fac_c = cfl_factors(visc)
    pv = np.asarray(g.cell_volume) * np.asarray(poro, dtype=np.float64)
    hc = g.hf_cell().astype(np.int64)
    nb = np.asarray(g.nb)
    dt_v = dt_g = dt_c = 1e100
    with np.errstate(divide="ignore", invalid="ignore"):
        if use_viscous:
            q = np.asarray(outflux, dtype=np.float64)
            fp = np.zeros(nc)
            fn = np.zeros(nc)
            np.add.at(fp, hc, np.where(q > 0, q, 0.0))
            np.add.at(fn, hc, np.where(q > 0, 0.0, -q))
            loc = fac * pv / np.maximum(fn, fp)
            if np.any(loc == 0.0):
                raise RuntimeError("Cfl computation gave dt = 0.0")
            dt_v = min(1e100, float(np.min(loc)))
        if use_gravity:
            internal = nb >= 0
            Kl = np.where(internal[:, None, None], 0.5 * (K[hc] + K[np.where(internal, nb, 0)]), K[hc])
            n = np.asarray(g.face_normal)
            gf = np.einsum("fq,fqk,k->f", n, Kl, np.asarray(gravity, dtype=np.float64) * drho)
            gf = gf * np.asarray(g.face_area)
            flux = np.zeros(nc)
            np.add.at(flux, hc, np.where(gf > 0, gf, 0.0))
            loc = fac_g * pv / flux
            dt_g = min(1e100, float(np.min(loc)))
        if use_capillary:
            internal = nb >= 0
            Kl = np.where(internal[:, None, None], 0.5 * (K[hc] + K[np.where(internal, nb, 0)]), K[hc])
            x = np.asarray(g.face_centroid) - np.asarray(g.cell_centroid)[hc]
            spatial = np.einsum("fi,fi->f", x, np.linalg.solve(Kl, x[:, :, None])[:, :, 0])
            loc = spatial / fac_c
            # std::min(dt, loc) keeps dt for a NaN loc
            valid = loc[~np.isnan(loc)]
            dt_c = min(1e100, float(np.min(valid))) if valid.size else 1e100
    return courant_number * min(dt_v, dt_g, dt_c)


def small_steps(cfl_dt, time, minimum_small_steps=1, maximum_small_steps=10000):
    """Number of small steps of transportSolve (:157-170)."""
    if cfl_dt > time:
        return minimum_small_steps
    steps = min(math.ceil(time / cfl_dt), 2**31 - 1)
    return min(max(int(steps), minimum_small_steps), maximum_small_steps)


class SaturationOutOfRange(RuntimeError):
    pass


def transport_solve(g, perm, poro, sat, time, gravity, outflux, drho, visc, inj=None, sat_bc=None,
                    partner=None, courant_number=0.5, method_viscous=True, method_gravity=True,
                    method_capillary=True, use_cfl_viscous=True, use_cfl_gravity=True,
                    use_cfl_capillary=True, minimum_small_steps=1, maximum_small_steps=10000,
                    check_sat=True, clamp_sat=False):
    """EulerUpstream::transportSolve.  Returns (saturation, small steps, repeats, cfl dt)."""
    nhf = g.num_half_faces
    sat_bc = np.ones(nhf) if sat_bc is None else np.asarray(sat_bc, dtype=np.float64)
    cfl = cfl_time(g, perm, poro, outflux, gravity, drho, visc, courant_number,
                   method_viscous and use_cfl_viscous, method_gravity and use_cfl_gravity,
                   method_capillary and use_cfl_capillary)
    nsteps = small_steps(cfl, time, minimum_small_steps, maximum_small_steps)
    pv = np.asarray(g.cell_volume) * np.asarray(poro, dtype=np.float64)
    s_init = np.asarray(sat, dtype=np.float64).copy()
    repeats = 0
    while True:
        dt = time / nsteps
        s = s_init.copy()
        try:
            for _ in range(nsteps):
                r = residual(g, perm, s, outflux, gravity, drho, visc, inj, sat_bc, partner,
                             method_viscous, method_gravity, method_capillary)
                s = s + dt * r / pv
                if check_sat or clamp_sat:
                    out = (s > 1.0) | (s < 0.0)
                    if clamp_sat:
                        s = np.where(out, np.clip(s, 0.0, 1.0), s)
                    elif np.any((s > 1.001) | (s < -0.001)):
                        raise SaturationOutOfRange("Saturation out of range in EulerUpstream")
            return s, nsteps, repeats, cfl
        except SaturationOutOfRange:
            repeats += 1
            if repeats > 10:
                raise
            nsteps *= 2

/* Explicit upstream saturation transport on the fluxes of a pressure solve:
 * the GPU counterpart of opm-porsol's EulerUpstream<GridInterface,
 * ReservoirProperties, BoundaryConditions> (opm/porsol/euler/
 * EulerUpstream.hpp, EulerUpstream_impl.hpp) with its residual
 * (EulerUpstreamResidual_impl.hpp) and CFL time steps (CflCalculator.hpp).
 *
 * A transport object is bound to a mimetic hybrid solver (mimetic_hybrid.h):
 * it reads that solver's grid, face structure (neighbours and periodic
 * partners), permeability and last outflux (right-hand side 0) in device
 * memory, as EulerUpstream::transportSolve reads the grid, the reservoir
 * properties and the FlowSolution it is given.
 *
 * Fluid: two phases without rock tables, as ReservoirPropertyCapillary falls
 * back to when no rock files are given: krw = s^2, kro = (1-s)^2
 * (ReservoirPropertyCapillary_impl.hpp:118-170), mobilities kr / mu,
 * p_c = 1e5 (1 - s) Pa (ReservoirPropertyCommon_impl.hpp:466-476), CFL
 * factors from 257 samples of the fractional flows (computeSingleRockCflFactors
 * (-1, ...), ReservoirPropertyCapillary_impl.hpp:215-260); the capillary CFL
 * factor is 0, so the capillary CFL time is 1e100 as in the reference.
 *
 * Saturation boundary conditions (BoundaryConditions.hpp:175-205, SatBC):
 * a boundary half-face whose face the solver pairs periodically is periodic;
 * every other boundary half-face is Dirichlet with the value given (the
 * reference's default SatBC is Dirichlet 1.0).
 *
 * Errors follow mimetic_hybrid.h: functions return an mhs_status and
 * mhs_transport_last_error() describes the cause.
 */
#ifndef EULER_UPSTREAM_H
#define EULER_UPSTREAM_H

#include "mimetic_hybrid.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mhs_transport mhs_transport;

/* EulerUpstream::init(param) (EulerUpstream_impl.hpp:95-108) parameters and
 * their defaults (:60-73). */
typedef struct mhs_transport_options {
    double courant_number;      /* 0.5 */
    int method_viscous;         /* 1 */
    int method_gravity;         /* 1 */
    int method_capillary;       /* 1 */
    int use_cfl_viscous;        /* 1 */
    int use_cfl_gravity;        /* 1 */
    int use_cfl_capillary;      /* 1 */
    int minimum_small_steps;    /* 1 */
    int maximum_small_steps;    /* 10000 */
    int check_sat;              /* 1 */
    int clamp_sat;              /* 0 */
} mhs_transport_options;

typedef struct mhs_transport_report {
    int small_steps;            /* small steps of the accepted pass */
    int repeats;                /* passes restarted with twice the steps (saturation out of range) */
    double cfl_dt;              /* computeCflTime: courant_number * min(viscous, gravity, capillary) */
    double dt;                  /* small step length of the accepted pass */
    double cfl_factor;          /* ReservoirPropertyCommon::cflFactor */
    double cfl_factor_gravity;  /* ::cflFactorGravity */
    double cfl_factor_capillary;/* ::cflFactorCapillary (0 for this fluid) */
    double seconds;             /* wall time of the call */
} mhs_transport_report;

/* Viscosities (Pa s) and densities (kg/m^3) of phase 0 (water) and 1 (oil);
 * ReservoirPropertyCommon::setViscosities / setDensities
 * (ReservoirPropertyCommon_impl.hpp:337-350). */
typedef struct mhs_fluid {
    double viscosity[2];
    double density[2];
} mhs_fluid;

void mhs_transport_default_options(mhs_transport_options *opt);

/* EulerUpstream(g, r, b) (EulerUpstream_impl.hpp:76-89): binds the transport
 * to the solver `flow` (same device and stream).  The solver must outlive
 * every call but mhs_transport_destroy, which may come after mhs_destroy. */
int mhs_transport_create(mhs_transport **out, mhs_solver *flow);
int mhs_transport_destroy(mhs_transport *t);
const char *mhs_transport_last_error(const mhs_transport *t);

/* EulerUpstream::initObj (:120-128) and EulerUpstreamResidual::initFinal
 * (EulerUpstreamResidual_impl.hpp:415-445): pore volumes from porosity[nc]
 * and the solver's cell volumes, periodic links from the solver's structure
 * (call again after mhs_init_system_structure), sat_bc[nhf] the Dirichlet
 * saturation of boundary half-faces (NULL: 1.0 everywhere), and the fluid. */
int mhs_transport_init(mhs_transport *t, const double *porosity, const double *sat_bc,
                       const mhs_fluid *fluid);

/* EulerUpstream::transportSolve (:153-215): advances saturation[nc] (in/out)
 * over `time` seconds with the last outflux of the bound solver, gravity[3]
 * and injection_rates[nc] (m^3/s per cell, > 0 injects water, < 0 produces
 * with the cell's fractional flow; NULL: none).  Small steps of length
 * time / n with n from the CFL time; when a step leaves a saturation outside
 * [-0.001, 1.001] (check_sat without clamp_sat) the pass restarts from the
 * initial saturation with twice the steps, at most 10 times, and then fails
 * with MHS_ERR_RANGE.  On a solver holding several instances the transport
 * advances all of them together: one CFL time, the smallest over the batch,
 * sets the small steps of every instance.  With several right-hand sides it
 * uses the fluxes of the first.  opt and report may be NULL. */
int mhs_transport_solve(mhs_transport *t, double *saturation, double time, const double *gravity,
                        const double *injection_rates, const mhs_transport_options *opt,
                        mhs_transport_report *report);
/* Same with saturation and injection_rates in device memory, ordered on the
 * stream given to mhs_create. */
int mhs_transport_solve_device(mhs_transport *t, double *saturation, double time, const double *gravity,
                               const double *injection_rates, const mhs_transport_options *opt,
                               mhs_transport_report *report);

/* EulerUpstreamResidual::computeResidual (EulerUpstreamResidual_impl.hpp:
 * 474-507) on host arrays: residual[nc] (m^3/s of water per cell) for
 * saturation[nc] with the options' method flags. */
int mhs_transport_residual(mhs_transport *t, const double *saturation, const double *gravity,
                           const double *injection_rates, const mhs_transport_options *opt,
                           double *residual);

/* Per-cell inputs of the pressure solve for this fluid at saturation[nc]
 * (device memory): totmob = lambda_w + lambda_o, rhomob = rho_w lambda_w +
 * rho_o lambda_o (MimeticIPEvaluator::computeDynamicParams,
 * MimeticIPEvaluator.hpp:297-311, with ReservoirPropertyCapillary::
 * phaseMobilities :67-72).  Either output may be NULL. */
int mhs_transport_mobilities_device(mhs_transport *t, const double *saturation, double *totmob,
                                    double *rhomob);

/* Profiling hook, no counterpart in the reference: average device time (ms)
 * of `reps` back-to-back small-step kernels (saturation update of every
 * cell) on the solver's stream, timed with HIP events, on the face data and
 * initial saturation of the last successful mhs_transport_solve*. */
int mhs_transport_time_step(mhs_transport *t, int reps, double *avg_ms);

#ifdef __cplusplus
}
#endif

#endif /* EULER_UPSTREAM_H */

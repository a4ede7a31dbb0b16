// Device views and launchers of the explicit upstream transport kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mhs_kernels.h"

namespace mhs {

// Per-face arrays of the small step are slot-major: face j of cell c
// (half-face facepos[c] + j) at j * nc + c, so that the threads of a
// wavefront (consecutive cells) read consecutive addresses.
struct DevTransport {
    int nc = 0, nhf = 0;
    const int *other = nullptr;      // [nhf] cell across the half-face (neighbour or periodic partner), -1 otherwise
    const uint8_t *kind = nullptr;   // [nhf] 0 Dirichlet boundary, 1 internal, 2 periodic
    const int *owner_hf = nullptr;   // [nhf] half-face of the face on its lower-numbered cell (itself on the boundary)
    const int *owner_nbf = nullptr;  // [nhf] the owner's neighbour face: itself inside, the partner if periodic
    const int *other_s = nullptr;    // [max_ncf * nc] slot-major copy of other
    const double *sat_bc_s = nullptr;  // [max_ncf * nc] Dirichlet saturation of boundary faces, slot-major
    const double *pv = nullptr;      // [nc] pore volume
    double *F = nullptr, *G = nullptr, *C = nullptr;  // [max_ncf * nc] owner-oriented flux, gravity and capillary
                                                      // coefficients, slot-major
    double visc_w = 1.0, visc_o = 1.0, rho_w = 0.0, rho_o = 0.0, drho = 0.0;
    int method_viscous = 1, method_gravity = 1, method_capillary = 1;
};

// Gravity and capillary coefficients (per permeability and gravity vector).
hipError_t launch_tr_coeffs(hipStream_t st, const DevTransport &D, const DevTopo &T, const DevGeom &G,
                            const double *perm, const double *grav, int has_g, int cap);
// Minimum gravity CFL time into *min_dt (initialised by the caller).
hipError_t launch_tr_cfl_gravity(hipStream_t st, const DevTransport &D, const DevTopo &T, const DevGeom &G,
                                 const double *perm, const double *grav, double fac_g, unsigned long long *min_dt);
// Face fluxes of the last pressure solve and the minimum viscous CFL time.
hipError_t launch_tr_flux(hipStream_t st, const DevTransport &D, const DevTopo &T, const double *outflux,
                          double fac_v, unsigned long long *min_dt, int *zero_dt);
hipError_t launch_tr_step(hipStream_t st, const DevTransport &D, const DevTopo &T, const double *s_in, double *s_out,
                          const double *inj, double dt, int check, int clamp, int *bad);
hipError_t launch_tr_residual(hipStream_t st, const DevTransport &D, const DevTopo &T, const double *s,
                              const double *inj, double *res);
hipError_t launch_tr_mobilities(hipStream_t st, const DevTransport &D, const double *s, double *totmob,
                                double *rhomob);
// estimateCellVelocity: vel[3 * nc] from the half-face outfluxes.
hipError_t launch_cell_velocity(hipStream_t st, int nc, const int *facepos, const DevGeom &G, const double *outflux,
                                double *vel);
hipError_t launch_tr_pore_volume(hipStream_t st, int nc, const double *vol, const double *poro, double *pv);

}  // namespace mhs

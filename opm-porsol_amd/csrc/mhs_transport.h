// Device views and launchers of the explicit upstream transport kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mhs_kernels.h"

namespace mhs {

struct DevTransport {
    int nc = 0, nhf = 0;
    const int *other = nullptr;      // [nhf] cell across the half-face (neighbour or periodic partner), -1 otherwise
    const uint8_t *kind = nullptr;   // [nhf] 0 Dirichlet boundary, 1 internal, 2 periodic
    const int *owner_hf = nullptr;   // [nhf] half-face of the face on its lower-numbered cell (itself on the boundary)
    const int *owner_nbf = nullptr;  // [nhf] the owner's neighbour face: itself inside, the partner if periodic
    const double *sat_bc = nullptr;  // [nhf] Dirichlet saturation of boundary half-faces
    const double *pv = nullptr;      // [nc] pore volume
    double *F = nullptr, *G = nullptr, *C = nullptr;  // [nhf] owner-oriented flux, gravity and capillary coefficients
    double visc_w = 1.0, visc_o = 1.0, rho_w = 0.0, rho_o = 0.0, drho = 0.0;
    int method_viscous = 1, method_gravity = 1, method_capillary = 1;
};

hipError_t launch_tr_faces(hipStream_t st, const DevTransport &D, const DevTopo &T, const DevGeom &G,
                           const double *perm, const double *outflux, const double *grav);
hipError_t launch_tr_cfl(hipStream_t st, const DevTransport &D, const DevTopo &T, const DevGeom &G,
                         const double *perm, const double *outflux, const double *grav, double fac_v, double fac_g,
                         int use_v, int use_g, unsigned long long *mins, int *zero_dt);
hipError_t launch_tr_step(hipStream_t st, const DevTransport &D, const DevTopo &T, const double *s_in, double *s_out,
                          const double *inj, double dt, int check, int clamp, int *bad);
hipError_t launch_tr_residual(hipStream_t st, const DevTransport &D, const DevTopo &T, const double *s,
                              const double *inj, double *res);
hipError_t launch_tr_mobilities(hipStream_t st, const DevTransport &D, const double *s, double *totmob,
                                double *rhomob);
hipError_t launch_tr_pore_volume(hipStream_t st, int nc, const double *vol, const double *poro, double *pv);

}  // namespace mhs

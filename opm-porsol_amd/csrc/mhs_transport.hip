// Explicit upstream saturation transport on the fluxes of the pressure solve
// (opm/porsol/euler/EulerUpstream_impl.hpp, EulerUpstreamResidual_impl.hpp,
// CflCalculator.hpp), for the fluid without rock tables: krw = s^2,
// kro = (1-s)^2 (common/ReservoirPropertyCapillary_impl.hpp:118-170) and
// p_c = 1e5 (1 - s) (common/ReservoirPropertyCommon_impl.hpp:466-476).
//
// The reference visits every face once, from its lower-numbered cell, and
// scatters the saturation flux dS to both cells.  Here every small step is a
// gather over cells: each half-face carries the face's flux, gravity and
// capillary coefficients as seen from the face's owning (lower-numbered)
// cell, written once per transport solve by k_tr_faces, so both cells of a
// face evaluate the same dS from the same operands and no atomics are needed.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mhs_transport.h"

namespace mhs {

namespace {

constexpr int kTB = 256;

__device__ __forceinline__ double mob_w(double s, const DevTransport &D) { return s * s / D.visc_w; }
__device__ __forceinline__ double mob_o(double s, const DevTransport &D) { return (1.0 - s) * (1.0 - s) / D.visc_o; }
__device__ __forceinline__ double cap_p(double s) { return 1e5 * (1.0 - s); }

// Saturation flux through one face from cell 0 (the owner, saturation s0) to
// cell 1 (s1; the Dirichlet value on a boundary face), EulerUpstreamResidual
// UpdateForCell::operator() (:150-280) with scalar mobilities:
// F = outflux of the owner's half-face, G = area n.(drho K_avg g),
// C = area n.(K_avg u) / (d0 + d1) with u the unit vector of the capillary
// pressure gradient estimate (0 on Dirichlet faces).  As in the reference,
// method_gravity gates the choice of the phase that follows F (its G is 0
// without the gravity term) and the gravity flux, but not the gravity part
// of the other phase's upstream direction.
__device__ __forceinline__ double face_ds(const DevTransport &D, double F, double G, double C, double s0, double s1,
                                          bool bnd) {
    const bool triv_w = !D.method_gravity || G >= 0.0;  // phase whose flux follows F
    const double s_up = F >= 0.0 ? s0 : s1;
    const double m_triv = triv_w ? mob_w(s_up, D) : mob_o(s_up, D);
    const double gnt = (triv_w ? -1.0 : 1.0) * m_triv * G;
    const double s_up2 = (F + gnt >= 0.0) ? s0 : s1;
    const double mw = triv_w ? m_triv : mob_w(s_up2, D);
    const double mo = triv_w ? mob_o(s_up2, D) : m_triv;
    const double mtinv = 1.0 / (mw + mo);
    double dS = 0.0;
    if (D.method_viscous) dS += mw * (mtinv * F);
    if (D.method_gravity && !bnd) dS += mw * (mtinv * (mo * G));
    if (D.method_capillary && !bnd) {
        const double sa = 0.5 * (s0 + s1);
        const double aw = mob_w(sa, D), ao = mob_o(sa, D);
        dS += aw * ((1.0 / (aw + ao)) * (ao * (C * (cap_p(s1) - cap_p(s0)))));
    }
    return dS;
}

// Gravity and capillary coefficients of every face (they depend on K, the
// gravity vector and the geometry, not on the flux): one thread per
// half-face; the owner's thread writes the face's slot on both sides.
__global__ void k_tr_coeffs(DevTransport D, DevTopo T, DevGeom Gm, const double *__restrict__ perm, double gx,
                            double gy, double gz, int has_g, int cap) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= D.nhf || D.owner_hf[o] != o) return;
    const int c0 = T.hf_cell[o];
    const int oc = D.other[o];
    const int c1 = oc < 0 ? c0 : oc;
    double G = 0.0, C = 0.0;
    if (has_g || (cap && oc >= 0)) {
        double K[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) K[i] = 0.5 * (perm[(size_t)9 * c0 + i] + perm[(size_t)9 * c1 + i]);
        const double a = Gm.farea[o];
        const double n0 = Gm.fnrm[3 * (size_t)o], n1 = Gm.fnrm[3 * (size_t)o + 1], n2 = Gm.fnrm[3 * (size_t)o + 2];
        if (has_g) {
            const double gi0 = (K[0] * gx + K[1] * gy + K[2] * gz) * D.drho;
            const double gi1 = (K[3] * gx + K[4] * gy + K[5] * gz) * D.drho;
            const double gi2 = (K[6] * gx + K[7] * gy + K[8] * gz) * D.drho;
            G = a * (n0 * gi0 + n1 * gi1 + n2 * gi2);
        }
        if (cap && oc >= 0) {
            // estimateCapPressureGradient (:513-550): the finite difference
            // between the cell centroids through the face centroid(s)
            const int nbf = D.owner_nbf[o];
            const double *cc = Gm.ccen + 3 * (size_t)c0, *nc = Gm.ccen + 3 * (size_t)c1;
            const double *fc = Gm.fcen + 3 * (size_t)o, *nfc = Gm.fcen + 3 * (size_t)nbf;
            double e0 = cc[0] - fc[0], e1 = cc[1] - fc[1], e2 = cc[2] - fc[2];
            const double d0 = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
            e0 = nc[0] - nfc[0]; e1 = nc[1] - nfc[1]; e2 = nc[2] - nfc[2];
            const double d1 = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
            double u0 = nc[0] - nfc[0] + fc[0] - cc[0];
            double u1 = nc[1] - nfc[1] + fc[1] - cc[1];
            double u2 = nc[2] - nfc[2] + fc[2] - cc[2];
            const double un = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
            u0 /= un; u1 /= un; u2 /= un;
            const double k0 = K[0] * u0 + K[1] * u1 + K[2] * u2;
            const double k1 = K[3] * u0 + K[4] * u1 + K[5] * u2;
            const double k2 = K[6] * u0 + K[7] * u1 + K[8] * u2;
            C = a * (n0 * k0 + n1 * k1 + n2 * k2) / (d0 + d1);
        }
    }
    const size_t so = (size_t)(o - T.facepos[c0]) * D.nc + c0;
    D.G[so] = G;
    D.C[so] = C;
    if (oc >= 0) {
        const int t = T.hf_twin[o];
        const size_t st = (size_t)(t - T.facepos[oc]) * D.nc + oc;
        D.G[st] = G;
        D.C[st] = C;
    }
}

__device__ __forceinline__ unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }

// Block minimum of positive doubles as ordered bit patterns (NaN and +inf
// order above 1e100 and never win), then one atomic per block.
__device__ __forceinline__ void block_min(unsigned long long b, unsigned long long *out) {
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_down(b, off, 64);
        b = o < b ? o : b;
    }
    __shared__ unsigned long long sm[kTB / 64];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sm[w] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kTB / 64; ++i) b = sm[i] < b ? sm[i] : b;
        atomicMin(out, b);
    }
}

// Gravity CFL time (cfl_calculator::findCFLtimeGravity, CflCalculator.hpp:
// 91-137): per cell the positive parts of area n.(K g drho), K averaged over
// internal faces and the cell's own on boundary (also periodic) faces.
__global__ void k_tr_cfl_gravity(DevTransport D, DevTopo T, DevGeom Gm, const double *__restrict__ perm, double gx,
                                 double gy, double gz, double fac_g, unsigned long long *min_dt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    double dg = 1e100;
    if (c < D.nc) {
        const int h0 = T.facepos[c], h1 = T.facepos[c + 1];
        const double *Kc = perm + (size_t)9 * c;
        const double g3[3] = {gx, gy, gz};
        double flux = 0.0;
        for (int h = h0; h < h1; ++h) {
            double K[9];
            if (D.kind[h] == 1) {
                const double *Kn = perm + (size_t)9 * D.other[h];
#pragma unroll
                for (int i = 0; i < 9; ++i) K[i] = 0.5 * (Kc[i] + Kn[i]);
            } else {
#pragma unroll
                for (int i = 0; i < 9; ++i) K[i] = Kc[i];
            }
            const double *n = Gm.fnrm + 3 * (size_t)h;
            double f = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int q = 0; q < 3; ++q) f += n[q] * (K[3 * q + k] * g3[k] * D.drho);
            f *= Gm.farea[h];
            if (f > 0) flux += f;
        }
        const double loc = (fac_g * D.pv[c]) / flux;
        if (loc < dg) dg = loc;
    }
    block_min(dbits(dg), min_dt);
}

// Per transport solve: the owner's flux into the face slots of each cell,
// and the viscous CFL time (findCFLtimeVelocity, CflCalculator.hpp:55-85)
// from the cell's own outfluxes.  A zero time is flagged.
__global__ void k_tr_flux(DevTransport D, DevTopo T, const double *__restrict__ outflux, double fac_v,
                          unsigned long long *min_dt, int *zero_dt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    double dv = 1e100;
    if (c < D.nc) {
        const int h0 = T.facepos[c], h1 = T.facepos[c + 1];
        double fp = 0.0, fn = 0.0;
        for (int h = h0; h < h1; ++h) {
            const double q = outflux[h];
            if (q > 0) fp += q;
            else fn -= q;
            const int o = D.owner_hf[h];
            D.F[(size_t)(h - h0) * D.nc + c] = o == h ? q : outflux[o];
        }
        const double loc = (fac_v * D.pv[c]) / fmax(fn, fp);
        if (loc == 0.0) zero_dt[0] = 1;
        if (loc < dv) dv = loc;
    }
    block_min(dbits(dv), min_dt);
}

// Residual of cell c (EulerUpstreamResidual UpdateForCell::operator()
// :101-300): saturation fluxes of its faces and its source, the rate of a
// producing cell (rate < 0) scaled by its fractional flow.
__device__ __forceinline__ double cell_residual(const DevTransport &D, const DevTopo &T,
                                                const double *__restrict__ s_in, const double *__restrict__ inj,
                                                int c, double sc) {
    double r = 0.0;
    const int n = T.facepos[c + 1] - T.facepos[c];
    for (int j = 0; j < n; ++j) {
        const size_t sl = (size_t)j * D.nc + c;
        const int oc = D.other_s[sl];
        const bool own = oc < 0 || c < oc;
        const double sn = oc < 0 ? D.sat_bc_s[sl] : s_in[oc];
        const double F = D.F[sl], G = D.G[sl], C = D.C[sl];
        const double dS = own ? face_ds(D, F, G, C, sc, sn, oc < 0) : face_ds(D, F, G, C, sn, sc, false);
        r += own ? -dS : dS;
    }
    double rate = inj ? inj[c] : 0.0;
    if (rate < 0.0) {
        const double lw = mob_w(sc, D), lo = mob_o(sc, D);
        rate *= lw / (lw + lo);
    }
    return r + rate;
}

// One small step (smallTimeStep :357-385): s += dt r / pore volume, then
// the saturation check.  A set *bad (out of range in an earlier step of the
// pass) skips the work: the caller restarts the pass with more steps.
__global__ void k_tr_step(DevTransport D, DevTopo T, const double *__restrict__ s_in, double *__restrict__ s_out,
                          const double *__restrict__ inj, double dt, int check, int clamp, int *bad) {
    if (*bad) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.nc) return;
    const double sc = s_in[c];
    const double r = cell_residual(D, T, s_in, inj, c, sc);
    double s = sc + dt * r / D.pv[c];
    if (check || clamp) {
        if (s > 1.0 || s < 0.0) {
            if (clamp) s = fmax(fmin(s, 1.0), 0.0);
            else if (s > 1.001 || s < -0.001) atomicOr(bad, 1);
        }
    }
    s_out[c] = s;
}

__global__ void k_tr_residual(DevTransport D, DevTopo T, const double *__restrict__ s_in,
                              const double *__restrict__ inj, double *__restrict__ res) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.nc) return;
    res[c] = cell_residual(D, T, s_in, inj, c, s_in[c]);
}

// Phase mobilities of the same fluid for the pressure solve:
// totmob = lambda_w + lambda_o, rhomob = rho_w lambda_w + rho_o lambda_o
// (MimeticIPEvaluator::computeDynamicParams inputs).
__global__ void k_tr_mobilities(DevTransport D, const double *__restrict__ s, double *__restrict__ totmob,
                                double *__restrict__ rhomob) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.nc) return;
    const double lw = mob_w(s[c], D), lo = mob_o(s[c], D);
    if (totmob) totmob[c] = lw + lo;
    if (rhomob) rhomob[c] = D.rho_w * lw + D.rho_o * lo;
}

__global__ void k_tr_pore_volume(int nc, const double *__restrict__ vol, const double *__restrict__ poro,
                                 double *__restrict__ pv) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nc) pv[c] = vol[c] * poro[c];
}

inline int blocks(long long n) { return (int)((n + kTB - 1) / kTB); }

}  // namespace

hipError_t launch_tr_coeffs(hipStream_t st, const DevTransport &D, const DevTopo &T, const DevGeom &G,
                            const double *perm, const double *grav, int has_g, int cap) {
    if (D.nhf == 0) return hipSuccess;
    k_tr_coeffs<<<blocks(D.nhf), kTB, 0, st>>>(D, T, G, perm, grav[0], grav[1], grav[2], has_g, cap);
    return hipGetLastError();
}

hipError_t launch_tr_cfl_gravity(hipStream_t st, const DevTransport &D, const DevTopo &T, const DevGeom &G,
                                 const double *perm, const double *grav, double fac_g, unsigned long long *min_dt) {
    if (D.nc == 0) return hipSuccess;
    k_tr_cfl_gravity<<<blocks(D.nc), kTB, 0, st>>>(D, T, G, perm, grav[0], grav[1], grav[2], fac_g, min_dt);
    return hipGetLastError();
}

hipError_t launch_tr_flux(hipStream_t st, const DevTransport &D, const DevTopo &T, const double *outflux,
                          double fac_v, unsigned long long *min_dt, int *zero_dt) {
    if (D.nc == 0) return hipSuccess;
    k_tr_flux<<<blocks(D.nc), kTB, 0, st>>>(D, T, outflux, fac_v, min_dt, zero_dt);
    return hipGetLastError();
}

hipError_t launch_tr_step(hipStream_t st, const DevTransport &D, const DevTopo &T, const double *s_in, double *s_out,
                          const double *inj, double dt, int check, int clamp, int *bad) {
    if (D.nc == 0) return hipSuccess;
    k_tr_step<<<blocks(D.nc), kTB, 0, st>>>(D, T, s_in, s_out, inj, dt, check, clamp, bad);
    return hipGetLastError();
}

hipError_t launch_tr_residual(hipStream_t st, const DevTransport &D, const DevTopo &T, const double *s,
                              const double *inj, double *res) {
    if (D.nc == 0) return hipSuccess;
    k_tr_residual<<<blocks(D.nc), kTB, 0, st>>>(D, T, s, inj, res);
    return hipGetLastError();
}

hipError_t launch_tr_mobilities(hipStream_t st, const DevTransport &D, const double *s, double *totmob,
                                double *rhomob) {
    if (D.nc == 0) return hipSuccess;
    k_tr_mobilities<<<blocks(D.nc), kTB, 0, st>>>(D, s, totmob, rhomob);
    return hipGetLastError();
}

hipError_t launch_tr_pore_volume(hipStream_t st, int nc, const double *vol, const double *poro, double *pv) {
    if (nc == 0) return hipSuccess;
    k_tr_pore_volume<<<blocks(nc), kTB, 0, st>>>(nc, vol, poro, pv);
    return hipGetLastError();
}

}  // namespace mhs

namespace mhs {

namespace {

// estimateCellVelocity (common/SimulatorUtilities.hpp:59-81): per cell the
// sum over its faces of (x_f - x_c) flux_f / |c|, written as 3 doubles.
__global__ void k_cell_velocity(int nc, const int *__restrict__ facepos, DevGeom Gm,
                                const double *__restrict__ outflux, double *__restrict__ vel) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const double cx = Gm.ccen[3 * (size_t)c], cy = Gm.ccen[3 * (size_t)c + 1], cz = Gm.ccen[3 * (size_t)c + 2];
    const double vol = Gm.cvol[c];
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    for (int h = facepos[c]; h < facepos[c + 1]; ++h) {
        const double s = outflux[h] / vol;
        v0 += (Gm.fcen[3 * (size_t)h] - cx) * s;
        v1 += (Gm.fcen[3 * (size_t)h + 1] - cy) * s;
        v2 += (Gm.fcen[3 * (size_t)h + 2] - cz) * s;
    }
    vel[3 * (size_t)c] = v0;
    vel[3 * (size_t)c + 1] = v1;
    vel[3 * (size_t)c + 2] = v2;
}

}  // namespace

hipError_t launch_cell_velocity(hipStream_t st, int nc, const int *facepos, const DevGeom &G, const double *outflux,
                                double *vel) {
    if (nc == 0) return hipSuccess;
    k_cell_velocity<<<(nc + 255) / 256, 256, 0, st>>>(nc, facepos, G, outflux, vel);
    return hipGetLastError();
}

}  // namespace mhs

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

__global__ void k_tr_faces(DevTransport D, DevTopo T, DevGeom Gm, const double *__restrict__ perm,
                           const double *__restrict__ outflux, double gx, double gy, double gz) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= D.nhf) return;
    const int o = D.owner_hf[h];
    const int c0 = T.hf_cell[o];
    const int oc = D.other[o];
    const int c1 = oc < 0 ? c0 : oc;
    double K[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) K[i] = 0.5 * (perm[(size_t)9 * c0 + i] + perm[(size_t)9 * c1 + i]);
    const double a = Gm.farea[o];
    const double n0 = Gm.fnrm[3 * (size_t)o], n1 = Gm.fnrm[3 * (size_t)o + 1], n2 = Gm.fnrm[3 * (size_t)o + 2];
    const double gi0 = (K[0] * gx + K[1] * gy + K[2] * gz) * D.drho;
    const double gi1 = (K[3] * gx + K[4] * gy + K[5] * gz) * D.drho;
    const double gi2 = (K[6] * gx + K[7] * gy + K[8] * gz) * D.drho;
    const double G = a * (n0 * gi0 + n1 * gi1 + n2 * gi2);
    double C = 0.0;
    if (oc >= 0) {
        // estimateCapPressureGradient (:513-550): the finite difference between
        // the cell centroids through the face centroid(s)
        const int nbf = D.owner_nbf[h];
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
    D.F[h] = outflux[o];
    D.G[h] = G;
    D.C[h] = C;
}

__device__ __forceinline__ unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }

// Per-cell CFL times (cfl_calculator::findCFLtimeVelocity :55-85 and
// findCFLtimeGravity :91-137), reduced to their minima over the grid as
// ordered bit patterns of positive doubles.  A zero viscous time is flagged.
__global__ void k_tr_cfl(DevTransport D, DevTopo T, DevGeom Gm, const double *__restrict__ perm,
                         const double *__restrict__ outflux, double gx, double gy, double gz, double fac_v,
                         double fac_g, int use_v, int use_g, unsigned long long *mins, int *zero_dt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    double dv = 1e100, dg = 1e100;
    if (c < D.nc) {
        const int h0 = T.facepos[c], h1 = T.facepos[c + 1];
        if (use_v) {
            double fp = 0.0, fn = 0.0;
            for (int h = h0; h < h1; ++h) {
                const double q = outflux[h];
                if (q > 0) fp += q;
                else fn -= q;
            }
            const double loc = (fac_v * D.pv[c]) / fmax(fn, fp);
            if (loc == 0.0) zero_dt[0] = 1;
            if (loc < dv) dv = loc;
        }
        if (use_g) {
            const double *Kc = perm + (size_t)9 * c;
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
                const double g3[3] = {gx, gy, gz};
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
    }
    // minima of positive doubles (NaN and +inf order above 1e100 and never win)
    unsigned long long bv = dbits(dv), bg = dbits(dg);
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long ov = __shfl_down(bv, off, 64), og = __shfl_down(bg, off, 64);
        bv = ov < bv ? ov : bv;
        bg = og < bg ? og : bg;
    }
    __shared__ unsigned long long sv[kTB / 64], sg[kTB / 64];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) {
        sv[w] = bv;
        sg[w] = bg;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kTB / 64; ++i) {
            bv = sv[i] < bv ? sv[i] : bv;
            bg = sg[i] < bg ? sg[i] : bg;
        }
        atomicMin(&mins[0], bv);
        atomicMin(&mins[1], bg);
    }
}

// Residual of cell c (EulerUpstreamResidual UpdateForCell::operator()
// :101-300): saturation fluxes of its faces and its source, the rate of a
// producing cell (rate < 0) scaled by its fractional flow.
__device__ __forceinline__ double cell_residual(const DevTransport &D, const DevTopo &T,
                                                const double *__restrict__ s_in, const double *__restrict__ inj,
                                                int c, double sc) {
    double r = 0.0;
    const int h0 = T.facepos[c], h1 = T.facepos[c + 1];
    for (int h = h0; h < h1; ++h) {
        const int oc = D.other[h];
        const bool own = oc < 0 || c < oc;
        const double sn = oc < 0 ? D.sat_bc[h] : s_in[oc];
        const double dS = own ? face_ds(D, D.F[h], D.G[h], D.C[h], sc, sn, oc < 0)
                              : face_ds(D, D.F[h], D.G[h], D.C[h], sn, sc, false);
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

hipError_t launch_tr_faces(hipStream_t st, const DevTransport &D, const DevTopo &T, const DevGeom &G,
                           const double *perm, const double *outflux, const double *grav) {
    if (D.nhf == 0) return hipSuccess;
    k_tr_faces<<<blocks(D.nhf), kTB, 0, st>>>(D, T, G, perm, outflux, grav[0], grav[1], grav[2]);
    return hipGetLastError();
}

hipError_t launch_tr_cfl(hipStream_t st, const DevTransport &D, const DevTopo &T, const DevGeom &G,
                         const double *perm, const double *outflux, const double *grav, double fac_v, double fac_g,
                         int use_v, int use_g, unsigned long long *mins, int *zero_dt) {
    if (D.nc == 0) return hipSuccess;
    k_tr_cfl<<<blocks(D.nc), kTB, 0, st>>>(D, T, G, perm, outflux, grav[0], grav[1], grav[2], fac_v, fac_g, use_v,
                                           use_g, mins, zero_dt);
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

// HIP kernels of the mimetic hybrid face-pressure solver (gfx950 / CDNA4).
//
// The path is memory bound end to end (no MFMA work): every kernel is written
// for coalesced HBM streams, wave64 lanes and XCD-contiguous block ranges.
//
//  cell kernels (one lane per cell, FP64 register-resident local matrices):
//    k_static_ip   MimeticIPEvaluator::buildStaticContrib (MimeticIPEvaluator.hpp:201-270)
//    k_assemble    assembleDynamic + buildCellContrib + the row/column BC rules of
//                  addCellContrib (IncompFlowSolverHybrid.hpp:1324-1382, 1785-1901)
//    k_recover     computePressureAndFluxes (IncompFlowSolverHybrid.hpp:1672-1730)
//  row kernels (one lane per face row, SELL-64 storage, no atomics on values):
//    k_fill        global assembly as a gather: each face row pulls its <= 2 cell
//                  contributions, the periodic constraint and the regularisation
//    k_spmv_pq / k_update_xr / k_update_p   one Jacobi-PCG iteration, with
//                  per-(instance, rhs) reductions finished by the last workgroup
//                  of the instance (deterministic order, no host round trip)
//    k_start64     FP64 start (r = b - A x, p = M^-1 r)
//    k_reliable    FP64 true residual + residual replacement for the FP32
//                  iteration (runs every iteration, idle unless a system asked)
//    k_post        postProcessFluxes (IncompFlowSolverHybrid.hpp:754-807)
#include "mhs_kernels.h"
#include "mhs_structure.h"
#include "mhs_amg.h"

#include <type_traits>
#include <utility>

namespace mhs {

namespace {

struct FinArgs {
    double tol2;
    int init;
};

template <int V>
using IC = std::integral_constant<int, V>;

// Operands of the FP64 residual replacement fused into the FP32 SpMV.
struct RelArgs {
    const double *val64, *x64, *b;
    const float *dinv;
    float *r32;
};

// Hardware workgroup ids are dealt round-robin over the 8 XCDs; give every XCD
// a contiguous range of logical rows so its private L2 sees neighbouring rows.
__device__ __forceinline__ int xcd_block(int nblk) {
    const int b = blockIdx.x;
    const int q = nblk >> 3, r = nblk & 7;
    const int x = b & 7, i = b >> 3;
    return (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Instance of a 1024-row block.  A handle with one instance skips the table
// read, so the kernels' first data loads wait on one fewer round trip.
__device__ __forceinline__ int block_inst(const DevTopo &T, int b) { return T.ninst == 1 ? 0 : T.block_seg[b]; }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Sum over the NW wavefronts of a workgroup; thread 0 gets the total
// (fixed order: deterministic for a given workgroup size).
template <int NQ, int NW = 4>
__device__ __forceinline__ void block_sum256(double (&v)[NQ], double *lds) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double t = wave_sum(v[q]);
        if (lane == 0) lds[w * NQ + q] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < NW; ++i) acc += lds[i * NQ + q];
            v[q] = acc;
        }
    }
    __syncthreads();
}

// Per-workgroup partial sums: thread 0 stores them with plain stores.  The
// matching k_finish kernel (launched next on the stream) reduces the partials
// of each instance in index order.  Visibility across XCDs comes from the
// kernel boundary; no per-workgroup fences (on CDNA4 an agent-scope fence
// writes back and invalidates the whole L2 of the XCD).
template <int NQ, int NW = 4>
__device__ __forceinline__ void block_partials(double (&v)[NQ], int b, const DevRed &R) {
    __shared__ double lds[NW * NQ];
    block_sum256<NQ, NW>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) R.partials[(size_t)q * R.nblk + b] = v[q];
    }
}

// Same, reducing only the first M values (the finisher reads no more).
template <int NQ, int M, int NW = 4>
__device__ __forceinline__ void block_partials_n(double (&v)[NQ], int b, const DevRed &R) {
    double w[M];
#pragma unroll
    for (int q = 0; q < M; ++q) w[q] = v[q];
    block_partials<M, NW>(w, b, R);
}

// One workgroup per instance: tot[q] = sum of the instance's partials, then
// Fin::run(s, tot) on thread 0.  Fin::skip(s) must mirror the condition under
// which the producing kernel's workgroups returned early.
template <int NQ, class Fin>
__global__ __launch_bounds__(256) void k_finish(DevRed R, DevScal S, FinArgs A) {
    const int s = blockIdx.x;
    if (Fin::skip(s, S, A)) return;
    const int nq = Fin::nq(s, S);
    const int b0 = R.seg_blk[s], b1 = R.seg_blk[s + 1];
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        acc[q] = 0.0;
        if (q >= nq) continue;
        const double *pq = R.partials + (size_t)q * R.nblk;
        for (int base = b0; base < b1; base += 16 * 256) {
            double a[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {   // 16 independent loads in flight per lane
                const int i = base + u * 256 + threadIdx.x;
                a[u] = (i < b1) ? pq[i] : 0.0;
            }
#pragma unroll
            for (int w = 8; w > 0; w >>= 1)
#pragma unroll
                for (int u = 0; u < w; ++u) a[u] += a[u + w];
            acc[q] += a[0];
        }
    }
    __shared__ double lds[4 * NQ];
    block_sum256<NQ>(acc, lds);
    if (threadIdx.x == 0) Fin::run(s, acc, S, A);
}

// 16-byte vector access to a thread's contiguous chunk of N elements.
template <typename VT, int N>
__device__ __forceinline__ void vload(const VT *__restrict__ p, VT (&v)[N]) {
    static_assert((N * sizeof(VT)) % 16 == 0, "chunk must be a multiple of 16 bytes");
#pragma unroll
    for (int i = 0; i < (int)(N * sizeof(VT) / 16); ++i) {
        const uint4 t = reinterpret_cast<const uint4 *>(p)[i];
        __builtin_memcpy(&v[i * 16 / sizeof(VT)], &t, 16);
    }
}
template <typename VT, int N>
__device__ __forceinline__ void vstore(VT *__restrict__ p, const VT (&v)[N]) {
#pragma unroll
    for (int i = 0; i < (int)(N * sizeof(VT) / 16); ++i) {
        uint4 t;
        __builtin_memcpy(&t, &v[i * 16 / sizeof(VT)], 16);
        reinterpret_cast<uint4 *>(p)[i] = t;
    }
}

// The costly preconditioner kernels inside the PCG loop skip finished
// systems: act[s * NRHS + k] says whether system (instance s, rhs k) still
// iterates.  A null pointer (a standalone V-cycle) runs everything.
struct Gate {
    const int *act = nullptr;
};
template <int NRHS>
__device__ __forceinline__ bool gate_off(const Gate &g, int s) {   // instance s has finished
    if (!g.act) return false;
    bool any = false;
#pragma unroll
    for (int k = 0; k < NRHS; ++k) any |= g.act[s * NRHS + k] != 0;
    return !any;
}

template <int NRHS>
__device__ __forceinline__ bool load_act(const int *act, int s, int (&a)[NRHS]) {
    bool any = false;
#pragma unroll
    for (int k = 0; k < NRHS; ++k) {
        a[k] = act[s * NRHS + k];
        any |= (a[k] != 0);
    }
    return any;
}

// One SELL-64 row times NRHS vectors.  The slice width is wave-uniform.  The
// row is taken CH entries at a time: every index and value of a chunk is
// loaded first (non-temporal: the matrix is streamed once, the gathered vector
// should stay in L2), then all of the chunk's gathers are issued, so each lane
// has up to 2 CH independent loads in flight.  CH trades registers for
// memory-level parallelism.
template <typename MT, typename XT, typename AT, int NRHS, int CH = 16>
__device__ __forceinline__ void sell_row(const DevTopo &T, const MT *__restrict__ val,
                                         const XT *__restrict__ x, int f, AT (&acc)[NRHS]) {
    const int slice = f >> 6, lane = f & 63;
    const long long s0 = T.slice_ptr[slice];
    const int w = (int)((T.slice_ptr[slice + 1] - s0) >> 6);
    const long long base = s0 + lane;
#pragma unroll
    for (int k = 0; k < NRHS; ++k) acc[k] = AT(0);
    for (int e0 = 0; e0 < w; e0 += CH) {
        int ci[CH];
        MT a[CH];
#pragma unroll
        for (int e = 0; e < CH; ++e) {
            if (e0 + e < w) {
                ci[e] = __builtin_nontemporal_load(T.col + base + 64LL * (e0 + e));
                a[e] = __builtin_nontemporal_load(val + base + 64LL * (e0 + e));
            }
        }
#pragma unroll
        for (int e = 0; e < CH; ++e) {
            if (e0 + e < w) {
#pragma unroll
                for (int k = 0; k < NRHS; ++k) acc[k] += (AT)a[e] * (AT)x[(size_t)ci[e] * NRHS + k];
            }
        }
    }
}

// The same row once, for both products of a residual-replacement iteration:
// acc = A32 p with A32 = (float) A64 (exactly the stored FP32 values) and
// a64 = A64 x64, reading the indices and the FP64 values a single time.
template <typename VT, int NRHS, int CH>
__device__ __forceinline__ void sell_row_dual(const DevTopo &T, const double *__restrict__ val64,
                                              const VT *__restrict__ p, const double *__restrict__ x64, int f,
                                              VT (&acc)[NRHS], double (&a64)[NRHS]) {
    const int slice = f >> 6, lane = f & 63;
    const long long s0 = T.slice_ptr[slice];
    const int w = (int)((T.slice_ptr[slice + 1] - s0) >> 6);
    const long long base = s0 + lane;
#pragma unroll
    for (int k = 0; k < NRHS; ++k) {
        acc[k] = VT(0);
        a64[k] = 0.0;
    }
    for (int e0 = 0; e0 < w; e0 += CH) {
        int ci[CH];
        double a[CH];
#pragma unroll
        for (int e = 0; e < CH; ++e) {
            if (e0 + e < w) {
                ci[e] = __builtin_nontemporal_load(T.col + base + 64LL * (e0 + e));
                a[e] = __builtin_nontemporal_load(val64 + base + 64LL * (e0 + e));
            }
        }
#pragma unroll
        for (int e = 0; e < CH; ++e) {
            if (e0 + e < w) {
#pragma unroll
                for (int k = 0; k < NRHS; ++k) {
                    acc[k] += (VT)a[e] * p[(size_t)ci[e] * NRHS + k];
                    a64[k] += a[e] * x64[(size_t)ci[e] * NRHS + k];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Static inverse inner products (once per computeInnerProducts).
// B^-1 = (N K N' + t A (I - Q Q') A) / |c|,  t = 6 tr(K)/3, Q = orth(C),
// N rows area*normal, C rows area*(x_f - x_c)  (MimeticIPEvaluator.hpp:236-269).
// Q comes from Gram-Schmidt applied twice (only QQ' is used, so any orthonormal
// basis of range(C) gives the same B^-1).
// Per-cell frame of the mimetic inner product: N (rows area * unit normal),
// the face areas a and an orthonormal basis Q of range(C), C rows
// area * (x_f - x_c); rows i >= n are zero.
template <int MAXN>
__device__ __forceinline__ void cell_frame(const DevGeom &G, int c, int f0, int n, double (&N)[MAXN][3],
                                           double (&Q)[MAXN][3], double (&a)[MAXN]) {
    const double cc0 = G.ccen[3 * (size_t)c], cc1 = G.ccen[3 * (size_t)c + 1],
                 cc2 = G.ccen[3 * (size_t)c + 2];
#pragma unroll
    for (int i = 0; i < MAXN; ++i) {
        if (i < n) {
            const size_t h = f0 + i;
            const double ar = G.farea[h];
            a[i] = ar;
            N[i][0] = G.fnrm[3 * h] * ar;
            N[i][1] = G.fnrm[3 * h + 1] * ar;
            N[i][2] = G.fnrm[3 * h + 2] * ar;
            Q[i][0] = (G.fcen[3 * h] - cc0) * ar;
            Q[i][1] = (G.fcen[3 * h + 1] - cc1) * ar;
            Q[i][2] = (G.fcen[3 * h + 2] - cc2) * ar;
        } else {
            a[i] = 0.0;
            N[i][0] = N[i][1] = N[i][2] = 0.0;
            Q[i][0] = Q[i][1] = Q[i][2] = 0.0;
        }
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
            for (int l = 0; l < j; ++l) {
                double d = 0.0;
#pragma unroll
                for (int i = 0; i < MAXN; ++i) d += Q[i][l] * Q[i][j];
#pragma unroll
                for (int i = 0; i < MAXN; ++i) Q[i][j] -= d * Q[i][l];
            }
        }
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < MAXN; ++i) s += Q[i][j] * Q[i][j];
        const double inv = 1.0 / sqrt(s);
#pragma unroll
        for (int i = 0; i < MAXN; ++i) Q[i][j] *= inv;
    }
}

template <int MAXN>
__global__ __launch_bounds__(64) void k_static_ip(DevTopo T, DevGeom G,
                                                  const double *__restrict__ perm, double gx,
                                                  double gy, double gz,
                                                  double *__restrict__ binv,
                                                  double *__restrict__ gflux) {
    const int c = blockIdx.x * 64 + threadIdx.x;
    if (c >= T.nc) return;
    const int f0 = T.facepos[c], n = T.facepos[c + 1] - f0;
    // K enters through its symmetric part (a permeability tensor is symmetric;
    // this makes the result, and its gradient, well defined for any input).
    double K[9];
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
        for (int m = 0; m < 3; ++m)
            K[3 * l + m] = 0.5 * (perm[9 * (size_t)c + 3 * l + m] + perm[9 * (size_t)c + 3 * m + l]);
    const double Kg0 = K[0] * gx + K[1] * gy + K[2] * gz;
    const double Kg1 = K[3] * gx + K[4] * gy + K[5] * gz;
    const double Kg2 = K[6] * gx + K[7] * gy + K[8] * gz;
    const double vol = G.cvol[c];
    double N[MAXN][3], Q[MAXN][3], a[MAXN];
    cell_frame<MAXN>(G, c, f0, n, N, Q, a);
#pragma unroll
    for (int i = 0; i < MAXN; ++i)
        if (i < n) gflux[f0 + i] = N[i][0] * Kg0 + N[i][1] * Kg1 + N[i][2] * Kg2;
    double NK[MAXN][3];
#pragma unroll
    for (int i = 0; i < MAXN; ++i)
#pragma unroll
        for (int l = 0; l < 3; ++l)
            NK[i][l] = N[i][0] * K[l] + N[i][1] * K[3 + l] + N[i][2] * K[6 + l];
    const double t = 6.0 * (K[0] + K[4] + K[8]) / 3.0;
    const double rv = 1.0 / vol, tv = t / vol;
    double *out = binv + c;                       // component-major: entry q at q * nc
#pragma unroll
    for (int i = 0; i < MAXN; ++i) {
#pragma unroll
        for (int j = i; j < MAXN; ++j) {
            if (j < n) {
                const double s = NK[i][0] * N[j][0] + NK[i][1] * N[j][1] + NK[i][2] * N[j][2];
                const double qq = Q[i][0] * Q[j][0] + Q[i][1] * Q[j][1] + Q[i][2] * Q[j][2];
                const double B = a[i] * ((i == j ? 1.0 : 0.0) - qq) * a[j];
                out[(size_t)pidx(i, j, n) * T.nc] = rv * s + tv * B;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Dynamic assembly, per cell.  S = totmob*B^-1; F = S'1; L = sum F;
// g = src - sum gflux; local rhs = gflux - h + (g/L) F; S <- S - F F'/L
// (buildCellContrib, :1785-1811).  The per-row right-hand-side contribution
// folds in the column rules of addCellContrib (:1880-1896): Dirichlet columns
// and periodic columns redirected to the smaller partner dof move S(r,c)*value
// to the right-hand side.  Dirichlet rows carry S(r,r)*p (:1832-1835).
template <int MAXN, int NRHS, bool UNIF>
__global__ __launch_bounds__(64) void k_assemble(DevTopo T, const double *__restrict__ binv,
                                                 const double *__restrict__ gfs,
                                                 const double *__restrict__ totmob,
                                                 const double *__restrict__ rhomob,
                                                 const double *__restrict__ src,
                                                 const uint8_t *__restrict__ ftype,
                                                 const double *__restrict__ bcval,
                                                 DevCellSys C, DevEbe E) {
    constexpr int NP = MAXN * (MAXN + 1) / 2;
    const int c = blockIdx.x * 64 + threadIdx.x;
    if (c >= T.nc) return;
    const int f0 = T.facepos[c], n = UNIF ? MAXN : T.facepos[c + 1] - f0;   // UNIF: every cell has MAXN faces
    const double lam = totmob[c], rl = rhomob[c];
    double S[NP];
#pragma unroll
    for (int i = 0; i < MAXN; ++i)
#pragma unroll
        for (int j = i; j < MAXN; ++j)
            S[pidx(i, j, MAXN)] = (j < n) ? lam * binv[(size_t)pidx(i, j, n) * T.nc + c] : 0.0;
    double F[MAXN], gfl[MAXN];
    double L = 0.0, gsum = 0.0;
#pragma unroll
    for (int i = 0; i < MAXN; ++i) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < MAXN; ++j) s += (i <= j) ? S[pidx(i, j, MAXN)] : S[pidx(j, i, MAXN)];
        F[i] = (i < n) ? s : 0.0;
        gfl[i] = (i < n) ? gfs[f0 + i] * rl : 0.0;
    }
#pragma unroll
    for (int i = 0; i < MAXN; ++i) L += F[i];
#pragma unroll
    for (int i = 0; i < MAXN; ++i) gsum += gfl[i];
    const double rL = 1.0 / L;
#pragma unroll
    for (int i = 0; i < MAXN; ++i)
#pragma unroll
        for (int j = i; j < MAXN; ++j) {
            double v = -rL * F[i] * F[j] + S[pidx(i, j, MAXN)];
            S[pidx(i, j, MAXN)] = v;
            if (j < n) C.schur[(size_t)pidx(i, j, n) * T.nc + c] = v;
        }
    uint8_t ft[MAXN], rd[MAXN];
#pragma unroll
    for (int i = 0; i < MAXN; ++i) {
        ft[i] = (i < n) ? ftype[f0 + i] : (uint8_t)FT_INTERNAL;
        rd[i] = (i < n) ? T.hf_redirect[f0 + i] : (uint8_t)0;
        if (i < n) C.F[f0 + i] = F[i];
    }
    C.L[c] = L;
    if (E.val64) {
        // S'_c for the element-by-element operator: a Dirichlet row keeps its
        // diagonal only (:1832-1835), a Dirichlet column is dropped (:1880-1896)
        const int sg = E.cell_seg[c], local = E.cell_local[c], np = n * (n + 1) / 2;
        const size_t vb = (size_t)E.seg_val[sg];
#pragma unroll
        for (int i = 0; i < MAXN; ++i)
#pragma unroll
            for (int j = i; j < MAXN; ++j) {
                if (j >= n) continue;
                const bool drop = i != j && (ft[i] == FT_DIRICHLET || ft[j] == FT_DIRICHLET);
                const double v = drop ? 0.0 : S[pidx(i, j, MAXN)];
                const size_t o = vb + (size_t)ebe_slot(local, pidx(i, j, n), np);
                E.val64[o] = v;
                E.val32[o] = (float)v;
            }
    }
#pragma unroll
    for (int k = 0; k < NRHS; ++k) {
        const double gk = src[(size_t)k * T.nc + c] - gsum;
        C.g[(size_t)k * T.nc + c] = gk;
        const double *bv = bcval + (size_t)k * T.nhf + f0;
        double val[MAXN];
#pragma unroll
        for (int i = 0; i < MAXN; ++i) val[i] = (i < n && ft[i] != FT_INTERNAL) ? bv[i] : 0.0;
#pragma unroll
        for (int i = 0; i < MAXN; ++i) {
            if (i >= n) continue;
            double rr;
            if (ft[i] == FT_DIRICHLET) {
                rr = S[pidx(i, i, MAXN)] * val[i];
            } else {
                double lr = gfl[i] - (ft[i] == FT_NEUMANN ? val[i] : 0.0);
                lr = 1.0 * lr + (gk * rL) * F[i];
                rr = lr;
#pragma unroll
                for (int j = 0; j < MAXN; ++j) {
                    if (j < n && (ft[j] == FT_DIRICHLET || rd[j]))
                        rr -= ((i <= j) ? S[pidx(i, j, MAXN)] : S[pidx(j, i, MAXN)]) * val[j];
                }
            }
            C.rr[((size_t)k * T.max_ncf + i) * T.nc + c] = rr;
        }
    }
}

// The same for grids whose cells have up to 10 or 16 faces of varying count
// (faulted corner-point grids): held in registers, the packed local matrix of
// the largest cell takes every VGPR (one wave per SIMD).  Here nothing is held
// per entry: the row sums go to C.F and are read back, S_c = lam B^-1 is
// recomputed from B^-1 where needed and the right-hand side reads the Schur
// entries this thread has just written, with runtime loops over the cell's own
// face count (few registers, full occupancy).  Same operations in the same
// order per entry as k_assemble.
template <int NRHS>
__global__ __launch_bounds__(64) void k_assemble_stream(DevTopo T, const double *__restrict__ binv,
                                                        const double *__restrict__ gfs,
                                                        const double *__restrict__ totmob,
                                                        const double *__restrict__ rhomob,
                                                        const double *__restrict__ src,
                                                        const uint8_t *__restrict__ ftype,
                                                        const double *__restrict__ bcval,
                                                        DevCellSys C, DevEbe E) {
    const int c = blockIdx.x * 64 + threadIdx.x;
    if (c >= T.nc) return;
    const int f0 = T.facepos[c], n = T.facepos[c + 1] - f0;
    const double lam = totmob[c], rl = rhomob[c];
    const size_t nc = (size_t)T.nc;
    auto Bs = [&](int i, int j) {   // S_c(i, j) = lam B^-1(i, j), either triangle
        return lam * binv[(size_t)(i <= j ? pidx(i, j, n) : pidx(j, i, n)) * nc + c];
    };
    double L = 0.0, gsum = 0.0;
    for (int i = 0; i < n; ++i) {
        double sum = 0.0;
        for (int j = 0; j < n; ++j) sum += Bs(i, j);
        C.F[f0 + i] = sum;
        L += sum;
        gsum += gfs[f0 + i] * rl;
    }
    const double rL = 1.0 / L;
    C.L[c] = L;
    const int sg = E.val64 ? E.cell_seg[c] : 0, local = E.val64 ? E.cell_local[c] : 0, np = n * (n + 1) / 2;
    const size_t vb = E.val64 ? (size_t)E.seg_val[sg] : 0;
    for (int i = 0; i < n; ++i) {
        const double Fi = C.F[f0 + i];
        const bool di = ftype[f0 + i] == FT_DIRICHLET;
        for (int j = i; j < n; ++j) {
            const double v = -rL * Fi * C.F[f0 + j] + Bs(i, j);
            C.schur[(size_t)pidx(i, j, n) * nc + c] = v;
            if (E.val64) {
                // S'_c: a Dirichlet row keeps its diagonal only, a Dirichlet column is dropped
                const bool drop = i != j && (di || ftype[f0 + j] == FT_DIRICHLET);
                const double w = drop ? 0.0 : v;
                const size_t o = vb + (size_t)ebe_slot(local, pidx(i, j, n), np);
                E.val64[o] = w;
                E.val32[o] = (float)w;
            }
        }
    }
    auto Sc = [&](int i, int j) { return C.schur[(size_t)(i <= j ? pidx(i, j, n) : pidx(j, i, n)) * nc + c]; };
#pragma unroll
    for (int k = 0; k < NRHS; ++k) {
        const double gk = src[(size_t)k * nc + c] - gsum;
        C.g[(size_t)k * nc + c] = gk;
        const double *bv = bcval + (size_t)k * T.nhf + f0;
        for (int i = 0; i < n; ++i) {
            const uint8_t fi = ftype[f0 + i];
            const double vi = fi != FT_INTERNAL ? bv[i] : 0.0;
            double rr;
            if (fi == FT_DIRICHLET) {
                rr = Sc(i, i) * vi;
            } else {
                double lr = gfs[f0 + i] * rl - (fi == FT_NEUMANN ? vi : 0.0);
                lr = 1.0 * lr + (gk * rL) * C.F[f0 + i];
                rr = lr;
                for (int j = 0; j < n; ++j) {
                    const uint8_t fj = ftype[f0 + j];
                    if (fj == FT_DIRICHLET || T.hf_redirect[f0 + j])
                        rr -= Sc(i, j) * (fj != FT_INTERNAL ? bv[j] : 0.0);
                }
            }
            C.rr[((size_t)k * T.max_ncf + i) * nc + c] = rr;
        }
    }
}

// ---------------------------------------------------------------------------
// Face-row gather: fills SELL values, right-hand sides, the Jacobi diagonal,
// the initial iterate and its residual r0 = b - A x0.  x0 is nonzero only on
// Dirichlet rows, which hold nothing but their diagonal, and Dirichlet columns
// are dropped from every other row, so (A x0)_f = diag_f x0_f: the PCG start
// needs no product with the face matrix.  With ROW_LDS each thread sums its
// row in its own column of LDS (max_w doubles, stride TB) and writes every
// stored entry once; otherwise the row is summed in place in val64 (rows too
// wide for LDS).  FM: FILL_ALL writes the vectors and the matrix, FILL_VEC the
// vectors only (the diagonal is summed alone; a solve that applies the
// operator element by element reads no face matrix), FILL_MAT the matrix only
// (a later request for the face matrix of that solve).
template <int NRHS, int TB, bool ROW_LDS, int FM>
__global__ __launch_bounds__(TB) void k_fill(DevTopo T, DevCellSys C,
                                             const uint8_t *__restrict__ ftype,
                                             const double *__restrict__ bcval,
                                             const int *__restrict__ inst_reg, int dguess,
                                             double *__restrict__ val64,
                                             float *__restrict__ val32,
                                             double *__restrict__ b,
                                             double *__restrict__ dinv64,
                                             float *__restrict__ dinv32,
                                             double *__restrict__ x0,
                                             double *__restrict__ r0) {
    constexpr bool MAT = FM != FILL_VEC, VEC = FM != FILL_MAT;
    extern __shared__ double rowbuf[];
    // contiguous row ranges per XCD: the cells a row range touches (up to a
    // layer of cells away) stay in that XCD's L2 until their other rows come
    const int f = xcd_block(gridDim.x) * TB + threadIdx.x;
    if (f >= T.nfp) return;
    const int slice = f >> 6, lane = f & 63;
    const long long base = T.slice_ptr[slice] + lane;
    const int w = (int)((T.slice_ptr[slice + 1] - T.slice_ptr[slice]) >> 6);
    double *row = ROW_LDS ? rowbuf + threadIdx.x : val64 + base;
    constexpr long long rs = ROW_LDS ? TB : 64;    // stride between a row's entries
    if constexpr (MAT)
        for (int e = 0; e < w; ++e) row[rs * e] = 0.0;
    const int h0 = T.dof_hf[f];
    const long long dpos = rs * T.dslot[f];
    double bk[NRHS], xk[NRHS];
#pragma unroll
    for (int k = 0; k < NRHS; ++k) bk[k] = xk[k] = 0.0;
    double diag;
    if (h0 < 0) {  // padding row: identity
        if constexpr (MAT) row[dpos] = 1.0;
        diag = 1.0;
    } else if (ftype[h0] == FT_DIRICHLET) {
        const int c = T.hf_cell[h0], f0 = T.facepos[c], n = T.facepos[c + 1] - f0, r = h0 - f0;
        const double a = C.schur[(size_t)pidx(r, r, n) * T.nc + c];
        if constexpr (MAT) row[dpos] = a;
        diag = a;
        if constexpr (VEC) {
#pragma unroll
            for (int k = 0; k < NRHS; ++k) {
                bk[k] = C.rr[((size_t)k * T.max_ncf + r) * T.nc + c];
                xk[k] = dguess ? bk[k] / a : 0.0;
            }
        }
    } else {
        double dsum = 0.0;   // the diagonal alone, in the order the row sums it
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            // the contribution's cell and local face stored per row (coalesced),
            // not gathered through the half-face
            const int c = T.row_cell[2 * (size_t)f + t];
            if (c < 0) continue;
            const int f0 = T.facepos[c], n = T.facepos[c + 1] - f0, r = T.row_loc[2 * (size_t)f + t];
            if constexpr (MAT) {
                const uint8_t *sm = T.slotmap + (size_t)r * T.max_ncf * T.nc + c;   // entry j at sm[j * nc]
                // the cell's row r of S_c, 8 faces at a time: all loads of a
                // batch are issued before the LDS sums
                for (int j0 = 0; j0 < n; j0 += 8) {
                    double v[8];
                    int slot[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int j = j0 + u;
                        const bool in = j < n;
                        const bool dir = in && ftype[f0 + j] == FT_DIRICHLET;
                        slot[u] = in ? (int)sm[(size_t)j * T.nc] : 0;
                        v[u] = (in && !dir) ? C.schur[(size_t)(r <= j ? pidx(r, j, n) : pidx(j, r, n)) * T.nc + c]
                                            : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        if (j0 + u < n) row[rs * slot[u]] += v[u];
                }
            } else {
                dsum += C.schur[(size_t)pidx(r, r, n) * T.nc + c];
            }
            if constexpr (VEC) {
#pragma unroll
                for (int k = 0; k < NRHS; ++k) bk[k] += C.rr[((size_t)k * T.max_ncf + r) * T.nc + c];
            }
        }
        const int fp = T.ptn[f];
        if (fp >= 0) {  // periodic constraint a(x_f - x_partner) = a_f pd_f - a_p pd_p (:1838-1869)
            const int hp = T.dof_hf[fp];
            const int c0 = T.hf_cell[h0], f00 = T.facepos[c0], n0 = T.facepos[c0 + 1] - f00;
            const int c1 = T.hf_cell[hp], f01 = T.facepos[c1], n1 = T.facepos[c1 + 1] - f01;
            const double a0 = C.schur[(size_t)pidx(h0 - f00, h0 - f00, n0) * T.nc + c0];
            const double a1 = C.schur[(size_t)pidx(hp - f01, hp - f01, n1) * T.nc + c1];
            if constexpr (MAT) {
                row[dpos] += a0 + a1;
                row[rs * T.pslot[f]] -= a0 + a1;
            } else {
                dsum += a0 + a1;
            }
            if constexpr (VEC) {
#pragma unroll
                for (int k = 0; k < NRHS; ++k)
                    bk[k] += a0 * bcval[(size_t)k * T.nhf + h0] - a1 * bcval[(size_t)k * T.nhf + hp];
            }
        }
        const int s = T.block_seg[f / kRowBlock];
        if (inst_reg[s] && f == T.inst_regrow[s]) {  // S[0][0] *= 2 (:1398-1400)
            if constexpr (MAT) row[dpos] *= 2.0;
            else dsum *= 2.0;
        }
        if constexpr (MAT) diag = row[dpos];
        else diag = dsum;
    }
    if constexpr (VEC) {
#pragma unroll
        for (int k = 0; k < NRHS; ++k) {
            b[(size_t)f * NRHS + k] = bk[k];
            x0[(size_t)f * NRHS + k] = xk[k];
            r0[(size_t)f * NRHS + k] = bk[k] - diag * xk[k];
        }
        const double di = 1.0 / diag;
        if (dinv64) dinv64[f] = di;
        if (dinv32) dinv32[f] = (float)di;
    }
    if constexpr (MAT) {
        for (int e = 0; e < w; ++e) {
            const double v = row[rs * e];
            if (ROW_LDS) val64[base + 64LL * e] = v;
            if (val32) val32[base + 64LL * e] = (float)v;
        }
    }
}

// ---------------------------------------------------------------------------
// PCG iteration kernels.
// Row kernels handle 1024 rows per workgroup (kRowBlock): SpMV-type kernels
// take rows b*1024 + j*256 + t (j = 0..3, coalesced per j), vector kernels a
// contiguous quad of rows per thread with 16-byte loads.

// q = A p and <p, q>.  With REL (FP32 iteration), systems whose FP64 residual
// replacement is due (upd != 0) also get r = b - A64 x64 in the same pass:
// the row's FP64 residual replaces the FP32 recursive one, and the partials
// carry |r64|^2, r.M^-1 r and |r|^2 of the replaced residual.
template <typename MT, typename VT, int NRHS, bool REL, int TB, int CH>
__device__ __forceinline__ void spmv_pq_body(const DevTopo &T, const MT *__restrict__ val,
                                             const VT *__restrict__ p, VT *__restrict__ q, const DevRed &R,
                                             const DevScal &S, const RelArgs &L) {
    const int b = T.spmv_xcd ? xcd_block(gridDim.x) : (int)blockIdx.x;
    const int s = block_inst(T, b);
    int act[NRHS];
    if (!load_act<NRHS>(S.act, s, act)) return;
    int u[NRHS];
    bool rel = false;
#pragma unroll
    for (int k = 0; k < NRHS; ++k) {
        u[k] = (REL && act[k]) ? S.upd[s * NRHS + k] : 0;
        rel |= u[k] != 0;
    }
    constexpr int NQ = REL ? 4 * NRHS : NRHS;
    double part[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) part[q] = 0.0;
    // A residual-replacement iteration (rel is block-uniform: the flags are
    // per instance) forms the FP32 product and the FP64 residual from one
    // pass over the indices and the FP64 values; other iterations stream the
    // FP32 values.
    if (!(REL && rel)) {
#pragma unroll
        for (int j = 0; j < kRowBlock / TB; ++j) {
            const int f = b * kRowBlock + j * TB + threadIdx.x;
            VT acc[NRHS];
            sell_row<MT, VT, VT, NRHS, CH>(T, val, p, f, acc);
#pragma unroll
            for (int k = 0; k < NRHS; ++k) {
                q[(size_t)f * NRHS + k] = acc[k];
                part[k] += (double)p[(size_t)f * NRHS + k] * (double)acc[k];
            }
        }
    }
    if constexpr (REL) {
        if (__builtin_expect(rel, 0)) {
            for (int j = 0; j < kRowBlock / TB; ++j) {
                const int f = b * kRowBlock + j * TB + threadIdx.x;
                VT acc[NRHS];
                double a64[NRHS];
                sell_row_dual<VT, NRHS, 4>(T, L.val64, p, L.x64, f, acc, a64);
#pragma unroll
                for (int k = 0; k < NRHS; ++k) {
                    q[(size_t)f * NRHS + k] = acc[k];
                    part[k] += (double)p[(size_t)f * NRHS + k] * (double)acc[k];
                }
                const float d = L.dinv[f];
#pragma unroll
                for (int k = 0; k < NRHS; ++k) {
                    if (!u[k]) continue;
                    const size_t i = (size_t)f * NRHS + k;
                    const double ri = L.b[i] - a64[k];
                    const float rf = (float)ri;
                    L.r32[i] = rf;
                    part[NRHS + 3 * k] += ri * ri;
                    part[NRHS + 3 * k + 1] += (double)rf * (double)(d * rf);
                    part[NRHS + 3 * k + 2] += (double)rf * (double)rf;
                }
            }
        }
    }
    if constexpr (REL) {
        if (rel) block_partials<NQ, TB / 64>(part, b, R);
        else block_partials_n<NQ, NRHS, TB / 64>(part, b, R);
    } else {
        block_partials<NQ, TB / 64>(part, b, R);
    }
}

template <typename MT, typename VT, int NRHS, bool REL, int TB, int CH>
__global__ __launch_bounds__(TB) void k_spmv_pq(DevTopo T, const MT *__restrict__ val, const VT *__restrict__ p,
                                                VT *__restrict__ q, DevRed R, DevScal S, RelArgs L) {
    spmv_pq_body<MT, VT, NRHS, REL, TB, CH>(T, val, p, q, R, S, L);
}

// The same held to 8 waves per SIMD: the FP32 single-system kernel of the
// mixed-precision PCG fits in 64 VGPRs this way (70 otherwise, for the
// FP64 replacement loop that most iterations skip), without spills.
template <typename MT, typename VT, int NRHS, bool REL, int TB, int CH>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_spmv_pq8(
    DevTopo T, const MT *__restrict__ val, const VT *__restrict__ p, VT *__restrict__ q, DevRed R, DevScal S,
    RelArgs L) {
    spmv_pq_body<MT, VT, NRHS, REL, TB, CH>(T, val, p, q, R, S, L);
}

// FP32 single-system instantiations launch k_spmv_pq8, the others k_spmv_pq.
template <typename MT, typename VT, int NRHS, bool REL, int TB, int CH>
static void launch_spmv_kernel(hipStream_t st, unsigned nblk, const DevTopo &T, const MT *val, const VT *p, VT *q,
                               const DevRed &R, const DevScal &S, const RelArgs &L) {
    if constexpr (std::is_same<MT, float>::value && NRHS == 1)
        k_spmv_pq8<MT, VT, NRHS, REL, TB, CH><<<nblk, TB, 0, st>>>(T, val, p, q, R, S, L);
    else
        k_spmv_pq<MT, VT, NRHS, REL, TB, CH><<<nblk, TB, 0, st>>>(T, val, p, q, R, S, L);
}

// ---------------------------------------------------------------------------
// Element-by-element level-0 operator (FP32 iteration).  The face matrix is the
// sum over cells of S'_c, the local Schur complement after the rules of
// addCellContrib (IncompFlowSolverHybrid.hpp:1817-1901) that keep only the
// diagonal of a Dirichlet row and drop Dirichlet columns.  One launch per
// colour (no two cells of a colour share a face), one thread per cell:
// y = S'_c x_c, then each face row is stored by the first of its cells to
// write it, added to by the others, and finished by the last:
//   EBE_SPMV   q = A p and the <p, q> partials.  With REL, when the instance's
//              FP64 residual replacement is due, also r = b - A64 x64: the
//              FP64 values are read once, the FP32 product uses their rounding
//              (the values every other iteration reads), the FP64 sums build
//              up in r64, and the finishing cell writes r32 and the partials
//              |r64|^2, r.M^-1 r, |r|^2.
//   EBE_RESID  res = r - A x (the multigrid pre-smoothing residual).
// Per hexahedron: 21 packed values and 6 face rows, against 11 entries and 11
// column indices per face row of the assembled matrix.
enum { EBE_SPMV = 0, EBE_RESID = 1 };
constexpr uint32_t kEbeFirstD = 1u << 30, kEbeLastD = 1u << 31, kEbeRowD = (1u << 30) - 1;

// One cell of a colour pass (see k_ebe).  NF is the face count when FIXED
// (every cell of the block has NF faces: the loops unroll to immediate load
// offsets), else the bound of the block's runtime face count nrt.  RP: the
// block's instance has its FP64 residual replacement due (block-uniform), so
// the FP32-only path is compiled without any of the FP64 work.
template <int NF, bool FIXED, int NRHS, int MODE, bool RP, int NQ>
__device__ __forceinline__ void ebe_cell(const DevEbe &E, int nrt, longlong2 bb, bool valid, bool any,
                                         const int (&u)[NRHS], const float *__restrict__ x,
                                         float *__restrict__ out, const float *__restrict__ rin, const RelArgs &L,
                                         double *__restrict__ r64, double (&part)[NQ]) {
    constexpr int MAXN = NF;
    const int n = FIXED ? NF : nrt, np = n * (n + 1) / 2;
    // component q of this cell at tile base + 64 q (64-cell tiles)
    const uint32_t *dp = E.dof + bb.y + ebe_slot(threadIdx.x, 0, n);
    uint32_t d[MAXN];
#pragma unroll
    for (int j = 0; j < MAXN; ++j) d[j] = (valid && j < n) ? __builtin_nontemporal_load(dp + 64 * j) : 0u;
    if (!any) return;                       // block-uniform: one instance per block
    if (valid) {
        float xv[MAXN][NRHS], y[MAXN][NRHS];
#pragma unroll
        for (int j = 0; j < MAXN; ++j)
#pragma unroll
            for (int k = 0; k < NRHS; ++k) {
                xv[j][k] = (j < n) ? x[(size_t)(d[j] & kEbeRowD) * NRHS + k] : 0.0f;
                y[j][k] = 0.0f;
            }
        const size_t vb = (size_t)bb.x + (size_t)ebe_slot(threadIdx.x, 0, np);
        double y64[RP ? MAXN : 1][NRHS];
        if constexpr (!RP) {
            const float *v = E.val32 + vb;
#pragma unroll
            for (int i = 0; i < MAXN; ++i)
#pragma unroll
                for (int j = i; j < MAXN; ++j) {
                    if (j >= n) continue;
                    const float a = __builtin_nontemporal_load(v + 64 * pidx(i, j, n));
#pragma unroll
                    for (int k = 0; k < NRHS; ++k) {
                        y[i][k] += a * xv[j][k];
                        if (j != i) y[j][k] += a * xv[i][k];
                    }
                }
        } else {
            const double *v = E.val64 + vb;
            double xd[MAXN][NRHS];
#pragma unroll
            for (int j = 0; j < MAXN; ++j)
#pragma unroll
                for (int k = 0; k < NRHS; ++k) {
                    xd[j][k] = (j < n && u[k]) ? L.x64[(size_t)(d[j] & kEbeRowD) * NRHS + k] : 0.0;
                    y64[j][k] = 0.0;
                }
#pragma unroll
            for (int i = 0; i < MAXN; ++i)
#pragma unroll
                for (int j = i; j < MAXN; ++j) {
                    if (j >= n) continue;
                    const double a = __builtin_nontemporal_load(v + 64 * pidx(i, j, n));
                    const float af = (float)a;
#pragma unroll
                    for (int k = 0; k < NRHS; ++k) {
                        y[i][k] += af * xv[j][k];
                        y64[i][k] += a * xd[j][k];
                        if (j != i) {
                            y[j][k] += af * xv[i][k];
                            y64[j][k] += a * xd[i][k];
                        }
                    }
                }
        }
#pragma unroll
        for (int j = 0; j < MAXN; ++j) {
            if (j >= n) continue;
            const uint32_t dj = d[j];
            const size_t row = dj & kEbeRowD;
            const bool first = (dj & kEbeFirstD) != 0, last = (dj & kEbeLastD) != 0;
#pragma unroll
            for (int k = 0; k < NRHS; ++k) {
                const size_t i = row * NRHS + k;
                if constexpr (MODE == EBE_SPMV) {
                    const float qv = first ? y[j][k] : out[i] + y[j][k];
                    out[i] = qv;
                    if (last) part[k] += (double)xv[j][k] * (double)qv;
                } else {
                    out[i] = (first ? rin[i] : out[i]) - y[j][k];
                }
            }
            if constexpr (RP && MODE == EBE_SPMV) {
#pragma unroll
                for (int k = 0; k < NRHS; ++k) {
                    if (!u[k]) continue;
                    const size_t i = row * NRHS + k;
                    const double rv = (first ? L.b[i] : r64[i]) - y64[j][k];
                    if (!last) {
                        r64[i] = rv;
                        continue;
                    }
                    const float rf = (float)rv;
                    const float dg = L.dinv[row];
                    L.r32[i] = rf;
                    part[NRHS + 3 * k] += rv * rv;
                    part[NRHS + 3 * k + 1] += (double)rf * (double)(dg * rf);
                    part[NRHS + 3 * k + 2] += (double)rf * (double)rf;
                }
            }
        }
    }
}

template <int MAXN, int NRHS, int MODE, bool REL>
__global__ __launch_bounds__(256) void k_ebe(DevEbe E, int blk0, Gate gt, DevScal S,
                                             const float *__restrict__ x, float *__restrict__ out,
                                             const float *__restrict__ rin, RelArgs L,
                                             double *__restrict__ r64) {
    // XCD-contiguous block ranges: every colour pass gives an XCD the same
    // cells' neighbourhood, so the second colour finds the rows the first
    // one wrote in that XCD's L2
    const int b = blk0 + xcd_block(gridDim.x);
    // The block's two records first; then the instance's flags and the
    // cells' face rows in one round trip.  The rows are loaded before the
    // gate is known, so a block of a finished instance reads them in vain.
    const int4 bi = E.blk_info[b];          // face count, instance, cells, partial slot
    const longlong2 bb = E.blk_base[b];     // value and face-row bases of the first tile
    const int n = bi.x, s = bi.y;
    const bool valid = (int)threadIdx.x < bi.z;
    int act[NRHS], u[NRHS];
#pragma unroll
    for (int k = 0; k < NRHS; ++k) {
        act[k] = gt.act ? gt.act[s * NRHS + k] : 1;
        u[k] = REL ? S.upd[s * NRHS + k] : 0;
    }
    bool any = false, rel = false;
#pragma unroll
    for (int k = 0; k < NRHS; ++k) {
        any |= act[k] != 0;
        u[k] = act[k] ? u[k] : 0;
        rel |= u[k] != 0;
    }
    constexpr int NQ = MODE == EBE_SPMV ? (REL ? 4 * NRHS : NRHS) : 1;
    double part[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) part[q] = 0.0;
    // The face count is block-uniform: blocks of hexahedra (and of cells
    // with the largest face count) take the path with compile-time offsets,
    // which a runtime count costs a quarter of the kernel's time.
    auto faces = [&](auto rp) {
        constexpr bool RP = decltype(rp)::value;
        // the replacement path stays on the runtime face count: unrolled to
        // immediate offsets it loads all FP64 values at once and needs 100
        // VGPRs, which would halve the occupancy of every launch
        if (!RP && n == MAXN)
            ebe_cell<MAXN, true, NRHS, MODE, RP>(E, n, bb, valid, any, u, x, out, rin, L, r64, part);
        else if (!RP && MAXN > 6 && n == 6)
            ebe_cell<(MAXN > 6 ? 6 : MAXN), true, NRHS, MODE, RP>(E, n, bb, valid, any, u, x, out, rin, L, r64, part);
        else
            ebe_cell<MAXN, false, NRHS, MODE, RP>(E, n, bb, valid, any, u, x, out, rin, L, r64, part);
    };
    if (REL && rel)
        faces(std::integral_constant<bool, REL>{});
    else
        faces(std::false_type{});
    if (!any) return;
    if constexpr (MODE == EBE_SPMV) {
        const int slot = bi.w;
        if constexpr (REL) {
            if (rel) block_partials<NQ>(part, slot, E.R);
            else block_partials_n<NQ, NRHS>(part, slot, E.R);
        } else {
            block_partials<NQ>(part, slot, E.R);
        }
    }
}

template <typename VT, int NRHS>
__global__ __launch_bounds__(256) void k_update_xr(DevTopo T, const VT *__restrict__ dinv,
                                                   VT *__restrict__ x, VT *__restrict__ r,
                                                   const VT *__restrict__ p,
                                                   const VT *__restrict__ q, DevRed R,
                                                   DevScal S, VT *__restrict__ x0s = nullptr,
                                                   VT w = VT(0)) {
    const int b = xcd_block(gridDim.x);
    const int s = block_inst(T, b);
    int act[NRHS];
    if (!load_act<NRHS>(S.act, s, act)) return;
    VT al[NRHS];
#pragma unroll
    for (int k = 0; k < NRHS; ++k)
        al[k] = act[k] ? (VT)S.sc[(size_t)(s * NRHS + k) * SC_N + SC_ALPHA] : VT(0);
    const int f0 = b * kRowBlock + 4 * threadIdx.x;
    const size_t e0 = (size_t)f0 * NRHS;
    VT d[4], xv[4 * NRHS], rv[4 * NRHS], pv[4 * NRHS], qv[4 * NRHS];
    vload<VT, 4>(dinv + f0, d);
    vload<VT, 4 * NRHS>(x + e0, xv);
    vload<VT, 4 * NRHS>(r + e0, rv);
    vload<VT, 4 * NRHS>(p + e0, pv);
    vload<VT, 4 * NRHS>(q + e0, qv);
    double part[2 * NRHS];
#pragma unroll
    for (int k = 0; k < 2 * NRHS; ++k) part[k] = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int k = 0; k < NRHS; ++k) {
            const int i = j * NRHS + k;
            xv[i] = xv[i] + al[k] * pv[i];
            rv[i] = rv[i] - al[k] * qv[i];
            part[2 * k] += (double)rv[i] * (double)(d[j] * rv[i]);
            part[2 * k + 1] += (double)rv[i] * (double)rv[i];
        }
    vstore<VT, 4 * NRHS>(x + e0, xv);
    vstore<VT, 4 * NRHS>(r + e0, rv);
    if (x0s) {  // first multigrid smoothing step, x0 = w D^-1 r
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int k = 0; k < NRHS; ++k) xv[j * NRHS + k] = w * d[j] * rv[j * NRHS + k];
        vstore<VT, 4 * NRHS>(x0s + e0, xv);
    }
    block_partials<2 * NRHS>(part, b, R);
}

template <typename VT, int NRHS>
__global__ __launch_bounds__(256) void k_update_p(DevTopo T, const VT *__restrict__ dinv,
                                                  const VT *__restrict__ r,
                                                  VT *__restrict__ p, VT *__restrict__ x,
                                                  double *__restrict__ x64, DevScal S) {
    const int b = xcd_block(gridDim.x);
    const int s = block_inst(T, b);
    int act[NRHS];
    if (!load_act<NRHS>(S.act, s, act)) return;
    VT be[NRHS];
    int u[NRHS];
    bool any_u = false;
#pragma unroll
    for (int k = 0; k < NRHS; ++k) {
        const int sys = s * NRHS + k;
        u[k] = (x64 && act[k]) ? S.upd[sys] : 0;
        any_u |= u[k] != 0;
        be[k] = (u[k] == 2) ? VT(0) : (VT)S.sc[(size_t)sys * SC_N + SC_BETA];
    }
    const int f0 = b * kRowBlock + 4 * threadIdx.x;
    const size_t e0 = (size_t)f0 * NRHS;
    VT d[4], rv[4 * NRHS], pv[4 * NRHS];
    vload<VT, 4>(dinv + f0, d);
    vload<VT, 4 * NRHS>(r + e0, rv);
    vload<VT, 4 * NRHS>(p + e0, pv);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int k = 0; k < NRHS; ++k) {
            const int i = j * NRHS + k;
            if (act[k]) pv[i] = d[j] * rv[i] + be[k] * pv[i];
        }
    vstore<VT, 4 * NRHS>(p + e0, pv);
    if (any_u) {  // fold the low-precision correction into x64 before the replacement
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int k = 0; k < NRHS; ++k) {
                if (!u[k]) continue;
                const size_t i = e0 + j * NRHS + k;
                x64[i] += (double)x[i];
                x[i] = VT(0);
            }
    }
}

// r = b - A x (FP64; GIVEN: r holds it already), p = dinv r; rz = r.dinv.r,
// rr0 = r.r; activate systems.
template <int NRHS, bool GIVEN>
__global__ __launch_bounds__(256) void k_start64(DevTopo T, const double *__restrict__ val,
                                                 const double *__restrict__ dinv,
                                                 const double *__restrict__ bb,
                                                 const double *__restrict__ x,
                                                 double *__restrict__ r, double *__restrict__ p,
                                                 double tol2, DevRed R, DevScal S) {
    const int b = xcd_block(gridDim.x);
    double part[2 * NRHS];
#pragma unroll
    for (int k = 0; k < 2 * NRHS; ++k) part[k] = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int f = b * kRowBlock + j * 256 + threadIdx.x;
        double acc[NRHS];
        if constexpr (!GIVEN) sell_row<double, double, double, NRHS>(T, val, x, f, acc);
        const double d = dinv[f];
#pragma unroll
        for (int k = 0; k < NRHS; ++k) {
            const size_t i = (size_t)f * NRHS + k;
            const double ri = GIVEN ? r[i] : bb[i] - acc[k];
            r[i] = ri;
            p[i] = d * ri;
            part[2 * k] += ri * (d * ri);
            part[2 * k + 1] += ri * ri;
        }
    }
    block_partials<2 * NRHS>(part, b, R);
}

// FP64 residual replacement for the FP32 iteration, fused into one pass.
// For systems with upd != 0 (all systems when init): r64 = b - A64 x64 row by
// row (GIVEN: the start residual r0 the fill wrote), r32 = (float)r64, and on
// restart (upd == 2, or init) p32 = dinv r32.
// The last workgroup of the instance decides: |r64| <= tol |r0| ends the system,
// otherwise it continues from the replaced residual.
template <int NRHS, bool GIVEN>
__global__ __launch_bounds__(256) void k_reliable(DevTopo T, const double *__restrict__ val,
                                                  const double *__restrict__ bb,
                                                  const double *__restrict__ x,
                                                  const double *__restrict__ r0,
                                                  const float *__restrict__ dinv,
                                                  float *__restrict__ r32, float *__restrict__ p32,
                                                  int init, double tol2, DevRed R, DevScal S) {
    const int b = xcd_block(gridDim.x);
    const int s = block_inst(T, b);
    int u[NRHS];
    if (init) {
#pragma unroll
        for (int k = 0; k < NRHS; ++k) u[k] = 2;
    } else if (!load_act<NRHS>(S.upd, s, u)) {
        return;
    }
    double part[3 * NRHS];
#pragma unroll
    for (int k = 0; k < 3 * NRHS; ++k) part[k] = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int f = b * kRowBlock + j * 256 + threadIdx.x;
        double acc[NRHS];
        if constexpr (!GIVEN) sell_row<double, double, double, NRHS>(T, val, x, f, acc);
        const float d = dinv[f];
#pragma unroll
        for (int k = 0; k < NRHS; ++k) {
            if (!u[k]) continue;
            const size_t i = (size_t)f * NRHS + k;
            const double ri = GIVEN ? r0[i] : bb[i] - acc[k];
            const float rf = (float)ri;
            r32[i] = rf;
            if (u[k] == 2) p32[i] = d * rf;
            part[3 * k] += ri * ri;
            part[3 * k + 1] += (double)rf * (double)(d * rf);
            part[3 * k + 2] += (double)rf * (double)rf;
        }
    }
    block_partials<3 * NRHS>(part, b, R);
}

// Start of a linear solve: status words, iteration counts and replacement
// flags cleared and the FP32 correction x32 (may be null) zeroed, in one launch.
__global__ __launch_bounds__(256) void k_clear_state(DevScal S, int nsys, float *__restrict__ x32, long long nx) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < 4) S.status[i] = 0;
    if (i < nsys) {
        S.iters[i] = 0;
        S.upd[i] = 0;
    }
    if (x32 && i < nx) x32[i] = 0.0f;
}

// x64 += x32, x32 = 0 for every system (final fold after the iteration).
template <int NRHS>
__global__ __launch_bounds__(256) void k_merge_all(DevTopo T, double *__restrict__ x64,
                                                   float *__restrict__ x32) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int f = blockIdx.x * kRowBlock + j * 256 + threadIdx.x;
#pragma unroll
        for (int k = 0; k < NRHS; ++k) {
            const size_t i = (size_t)f * NRHS + k;
            x64[i] += (double)x32[i];
            x32[i] = 0.0f;
        }
    }
}

// |b - A x|^2 per system into sc[RRT] (report only, no vector writes).
template <int NRHS>
__global__ __launch_bounds__(256) void k_resnorm(DevTopo T, const double *__restrict__ val,
                                                 const double *__restrict__ bb,
                                                 const double *__restrict__ x, DevRed R,
                                                 DevScal S) {
    const int b = xcd_block(gridDim.x);
    double part[NRHS];
#pragma unroll
    for (int k = 0; k < NRHS; ++k) part[k] = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int f = b * kRowBlock + j * 256 + threadIdx.x;
        double acc[NRHS];
        sell_row<double, double, double, NRHS>(T, val, x, f, acc);
#pragma unroll
        for (int k = 0; k < NRHS; ++k) {
            const double ri = bb[(size_t)f * NRHS + k] - acc[k];
            part[k] += ri * ri;
        }
    }
    block_partials<NRHS>(part, b, R);
}


// ---------------------------------------------------------------------------
// Multigrid kernels.
//
// V-cycle on level l (b given, x from zero):
//   down: x = w D^-1 b (nu-1 more Jacobi sweeps); res = b - A x; b_{l+1} = P' res
//   up:   x~ = x + alpha P x_{l+1};  x = x~ + w D^-1 (b - A x~)  (nu sweeps)
// Level 0 is the SELL face matrix (rows in the PCG's padded numbering), the
// coarse levels are CSR.  The coarsest level is solved with a dense inverse
// per instance.  With symmetric smoothing the V-cycle is a fixed symmetric
// operator, as PCG requires.

// Coarse-level rows with batched loads: the column indices of kChunk entries
// are loaded together, then everything that depends on them, so a row costs a
// few memory round trips instead of one dependent chain per entry (the small
// levels are latency-bound).  Padding lanes either read the row's own entries
// or issue no load at all: never a row of another instance, which the
// one-workgroup-per-instance tail may not have written yet (0 * NaN != 0).
constexpr int kChunk = 8;

// Rows can be shared by a group of G lanes (G divides 64, so a group lies in
// one wave): lane `sub` of the group takes entries e0 + sub, e0 + sub + G, ...
// in chunks of KC, and group_sum adds the partial sums across the group, after
// which every lane of the group holds the row's sum.  G = 1 is one lane per row.
template <int G>
constexpr int chunk_for() { return G == 1 ? kChunk : 4; }

template <int G, typename VT, int NRHS>
__device__ __forceinline__ void group_sum(VT (&acc)[NRHS]) {
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1)
#pragma unroll
        for (int k = 0; k < NRHS; ++k) acc[k] += __shfl_xor(acc[k], m, G);
}

// Galerkin product A_c = P' A P with piecewise-constant P: every coarse entry
// is the sum of the finer entries listed for it (no atomics),
// out[e] = sum of fine[src[k]] over the gather list of entry e, G lanes per
// entry (coalesced list reads, several gathers in flight per lane).
template <int G>
__global__ __launch_bounds__(256) void k_galerkin(int nnz, const int *__restrict__ src_ptr,
                                                  const int *__restrict__ src,
                                                  const double *__restrict__ fine,
                                                  double *__restrict__ out64,
                                                  float *__restrict__ out32) {
    constexpr int KC = chunk_for<G>();
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int e = t / G, sub = t % G;
    if (e >= nnz) return;
    double s[1] = {0.0};
    const int k1 = src_ptr[e + 1];
    for (int k = src_ptr[e] + sub; k < k1; k += G * KC) {
        int q[KC];
#pragma unroll
        for (int u = 0; u < KC; ++u) q[u] = k + G * u < k1 ? src[k + G * u] : -1;
#pragma unroll
        for (int u = 0; u < KC; ++u) s[0] += q[u] >= 0 ? fine[q[u]] : 0.0;
    }
    group_sum<G>(s);
    if (sub != 0) return;
    out64[e] = s[0];
    if (out32) out32[e] = (float)s[0];
}

__global__ __launch_bounds__(256) void k_coarse_diag(int n, const int *__restrict__ diag,
                                                     const double *__restrict__ val,
                                                     double *__restrict__ dinv64,
                                                     float *__restrict__ dinv32) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double d = 1.0 / val[diag[i]];
    dinv64[i] = d;
    if (dinv32) dinv32[i] = (float)d;
}

// Smoothed prolongation of a coarse transfer, G lanes per row:
// P_iJ = [J = agg_i] - w_p d_i sum_{j in row i, agg_j = J} a_ij.
template <int G>
__global__ __launch_bounds__(256) void k_sa_p(DevAmgLevel F, double wp) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int i = t / G, sub = t % G;
    if (i >= F.n) return;
    const double d = F.dinv64[i];
    for (int q = F.p_ptr[i] + sub; q < F.p_ptr[i + 1]; q += G) {
        double s = 0.0;
        for (int k = F.p_src_ptr[q]; k < F.p_src_ptr[q + 1]; ++k) s += F.val64[F.p_src[k]];
        const double v = (q == F.p_own[i] ? 1.0 : 0.0) - wp * d * s;
        F.p_val64[q] = v;
        if (F.p_val32) F.p_val32[q] = (float)v;
        // the same value in P' row order, for the restriction's contiguous rows
        const int e = F.p_tpos[q];
        F.pt_val64[e] = v;
        if (F.pt_val32) F.pt_val32[e] = (float)v;
    }
}

// Combined up-sweep operator of a smoothed transfer, row by row:
// M = P - w d_i T (T = A P holds P's pattern), so that the prolongation
// x~ = w D^-1 b + al P xc followed by x = x~ + w D^-1 (b - A x~) reads one
// matrix: x = w D^-1 (b + res) + al M xc, with res = b - A (w D^-1 b) from the
// down-sweep.
// G lanes per row.
template <int G>
__global__ __launch_bounds__(256) void k_sa_m(DevAmgLevel F, double w) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int i = t / G, sub = t % G;
    if (i >= F.n) return;
    const double d = w * F.dinv64[i];
    for (int e = F.t_ptr[i] + sub; e < F.t_ptr[i + 1]; e += G) {
        const int q = F.t_p[e];
        const double v = (q >= 0 ? F.p_val64[q] : 0.0) - d * F.t_val64[e];
        F.m_val64[e] = v;
        if (F.m_val32) F.m_val32[e] = (float)v;
    }
}

// out[e] = sum over the pairs of entry e of a[pa] * b[pb] (T = A P, A_c = P' T),
// G lanes per entry.
template <int G>
__global__ __launch_bounds__(256) void k_pair_sum(int n, const int *__restrict__ ptr,
                                                  const int *__restrict__ pa, const int *__restrict__ pb,
                                                  const double *__restrict__ a, const double *__restrict__ b,
                                                  double *__restrict__ out64, float *__restrict__ out32) {
    constexpr int KC = chunk_for<G>();
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int e = t / G, sub = t % G;
    if (e >= n) return;
    double s[1] = {0.0};
    const int k1 = ptr[e + 1];
    for (int k = ptr[e] + sub; k < k1; k += G * KC) {
        int ia[KC], ib[KC];
#pragma unroll
        for (int u = 0; u < KC; ++u) {
            const bool in = k + G * u < k1;
            ia[u] = in ? pa[k + G * u] : -1;
            ib[u] = in ? pb[k + G * u] : -1;
        }
#pragma unroll
        for (int u = 0; u < KC; ++u) s[0] += ia[u] >= 0 ? a[ia[u]] * b[ib[u]] : 0.0;
    }
    group_sum<G>(s);
    if (sub != 0) return;
    out64[e] = s[0];
    if (out32) out32[e] = (float)s[0];
}

// Dense inverse of each instance's coarsest operator (Gauss-Jordan, no
// pivoting: the Galerkin operators are symmetric positive definite).  Per
// pivot: the scaled pivot row and the pivot column go to their own arrays,
// then every thread updates its own entries of the matrix; two barriers.
__global__ __launch_bounds__(1024) void k_coarse_inverse(DevAmgLevel L, const int *__restrict__ inst_ptr,
                                                         const int *__restrict__ off,
                                                         double *__restrict__ inv64,
                                                         float *__restrict__ inv32) {
    constexpr int M = kMaxCoarsest;
    constexpr int NT = 1024;
    __shared__ double A[M][M + 1];
    __shared__ double rowk[M], colk[M];
    const int t = blockIdx.x, tid = threadIdx.x;
    const int r0 = inst_ptr[t], m = inst_ptr[t + 1] - r0;
    for (int e = tid; e < M * M; e += NT) A[e / M][e % M] = 0.0;
    __syncthreads();
    for (int i = tid; i < m; i += NT)
        for (int k = L.row_ptr[r0 + i]; k < L.row_ptr[r0 + i + 1]; ++k)
            A[i][L.col[k] - r0] = L.val64[k];
    __syncthreads();
    for (int k = 0; k < m; ++k) {
        if (tid < M) {
            if (tid < m) rowk[tid] = (tid == k ? 1.0 : A[k][tid]) / A[k][k];
        } else if (tid < 2 * M) {
            const int i = tid - M;
            if (i < m) colk[i] = A[i][k];
        }
        __syncthreads();
        for (int e = tid; e < M * M; e += NT) {
            const int i = e / M, j = e % M;
            if (i >= m || j >= m) continue;
            A[i][j] = (i == k) ? rowk[j] : (j == k ? 0.0 : A[i][j]) - colk[i] * rowk[j];
        }
        __syncthreads();
    }
    for (int e = tid; e < M * M; e += NT) {
        const int i = e / M, j = e % M;
        if (i >= m || j >= m) continue;
        const double v = A[i][j];
        inv64[off[t] + i * m + j] = v;
        if (inv32) inv32[off[t] + i * m + j] = (float)v;
    }
}

template <typename VT>
__device__ __forceinline__ const VT *lvl_val(const DevAmgLevel &L) {
    if constexpr (std::is_same<VT, float>::value) return L.val32; else return L.val64;
}
template <typename VT>
__device__ __forceinline__ const VT *lvl_dinv(const DevAmgLevel &L) {
    if constexpr (std::is_same<VT, float>::value) return L.dinv32; else return L.dinv64;
}

// acc = sum_j a_ij (w d_j b_j): the residual of the pre-smoothed x = w D^-1 b.
template <typename VT, int NRHS, int G = 1>
__device__ __forceinline__ void row_down(const DevAmgLevel &F, const VT *__restrict__ val,
                                         const VT *__restrict__ dinv, const VT *__restrict__ b, VT w, int i,
                                         VT (&acc)[NRHS], int sub = 0) {
    constexpr int KC = chunk_for<G>();
#pragma unroll
    for (int k = 0; k < NRHS; ++k) acc[k] = VT(0);
    const int e1 = F.row_ptr[i + 1];
    for (int base = F.row_ptr[i] + sub; base < e1; base += G * KC) {
        int c[KC];
        VT a[KC];
#pragma unroll
        for (int u = 0; u < KC; ++u) {
            const int e = base + G * u;
            const bool in = e < e1;
            c[u] = in ? F.col[e] : i;
            a[u] = in ? val[e] : VT(0);
        }
        VT d[KC], bv[KC][NRHS];
#pragma unroll
        for (int u = 0; u < KC; ++u) {
            d[u] = dinv[c[u]];
#pragma unroll
            for (int k = 0; k < NRHS; ++k) bv[u][k] = b[(size_t)c[u] * NRHS + k];
        }
#pragma unroll
        for (int u = 0; u < KC; ++u)
#pragma unroll
            for (int k = 0; k < NRHS; ++k) acc[k] += a[u] * (w * d[u]) * bv[u][k];
    }
    group_sum<G>(acc);
}

// acc = sum_j a_ij x~_j with x~_j = w d_j b_j + alpha xc[agg_j] (prolongation on the fly).
template <typename VT, int NRHS, int G = 1>
__device__ __forceinline__ void row_up1(const DevAmgLevel &F, const VT *__restrict__ val,
                                        const VT *__restrict__ dinv, const VT *__restrict__ b,
                                        const int *__restrict__ agg, const VT *__restrict__ xc, VT w, VT al,
                                        int i, VT (&acc)[NRHS], int sub = 0) {
    constexpr int KC = chunk_for<G>();
#pragma unroll
    for (int k = 0; k < NRHS; ++k) acc[k] = VT(0);
    const int e1 = F.row_ptr[i + 1];
    for (int base = F.row_ptr[i] + sub; base < e1; base += G * KC) {
        int c[KC];
        VT a[KC];
#pragma unroll
        for (int u = 0; u < KC; ++u) {
            const int e = base + G * u;
            const bool in = e < e1;
            c[u] = in ? F.col[e] : i;
            a[u] = in ? val[e] : VT(0);
        }
        VT d[KC], bv[KC][NRHS];
        int g[KC];
#pragma unroll
        for (int u = 0; u < KC; ++u) {
            d[u] = dinv[c[u]];
            g[u] = agg[c[u]];
#pragma unroll
            for (int k = 0; k < NRHS; ++k) bv[u][k] = b[(size_t)c[u] * NRHS + k];
        }
        VT xg[KC][NRHS];
#pragma unroll
        for (int u = 0; u < KC; ++u)
#pragma unroll
            for (int k = 0; k < NRHS; ++k) xg[u][k] = xc[(size_t)g[u] * NRHS + k];
#pragma unroll
        for (int u = 0; u < KC; ++u)
#pragma unroll
            for (int k = 0; k < NRHS; ++k) acc[k] += a[u] * (w * d[u] * bv[u][k] + al * xg[u][k]);
    }
    group_sum<G>(acc);
}

// acc = sum over the children of coarse row I of res.
template <typename VT, int NRHS, int G = 1>
__device__ __forceinline__ void row_restrict(const int *__restrict__ child_ptr, const int *__restrict__ child,
                                             const VT *__restrict__ res, int I, VT (&acc)[NRHS], int sub = 0) {
    constexpr int KC = chunk_for<G>();
#pragma unroll
    for (int k = 0; k < NRHS; ++k) acc[k] = VT(0);
    const int e1 = child_ptr[I + 1];
    for (int base = child_ptr[I] + sub; base < e1; base += G * KC) {
        int c[KC];
#pragma unroll
        for (int u = 0; u < KC; ++u) c[u] = base + G * u < e1 ? child[base + G * u] : -1;
        VT r[KC][NRHS];
#pragma unroll
        for (int u = 0; u < KC; ++u)
#pragma unroll
            for (int k = 0; k < NRHS; ++k) r[u][k] = c[u] >= 0 ? res[(size_t)c[u] * NRHS + k] : VT(0);
#pragma unroll
        for (int u = 0; u < KC; ++u)
#pragma unroll
            for (int k = 0; k < NRHS; ++k) acc[k] += r[u][k];
    }
    group_sum<G>(acc);
}

template <typename VT>
__device__ __forceinline__ const VT *lvl_p(const DevAmgLevel &L) {
    if constexpr (std::is_same<VT, float>::value) return L.p_val32; else return L.p_val64;
}
template <typename VT>
__device__ __forceinline__ const VT *lvl_pt(const DevAmgLevel &L) {
    if constexpr (std::is_same<VT, float>::value) return L.pt_val32; else return L.pt_val64;
}
template <typename VT>
__device__ __forceinline__ const VT *lvl_m(const DevAmgLevel &L) {
    if constexpr (std::is_same<VT, float>::value) return L.m_val32; else return L.m_val64;
}

// acc = (P' res)_I across a smoothed transfer (P' row I: fine rows and the P
// values stored in P' row order, so a row's values are contiguous).
template <typename VT, int NRHS, int G = 1>
__device__ __forceinline__ void row_restrict_sa(const DevAmgLevel &F, const VT *__restrict__ res, int I,
                                                VT (&acc)[NRHS], int sub = 0) {
    constexpr int KC = chunk_for<G>();
    const VT *ptval = lvl_pt<VT>(F);
#pragma unroll
    for (int k = 0; k < NRHS; ++k) acc[k] = VT(0);
    const int e1 = F.pt_ptr[I + 1];
    for (int base = F.pt_ptr[I] + sub; base < e1; base += G * KC) {
        int c[KC];
        VT pv[KC];
#pragma unroll
        for (int u = 0; u < KC; ++u) {
            const int e = base + G * u;
            const bool in = e < e1;
            c[u] = in ? F.pt_col[e] : 0;
            pv[u] = in ? ptval[e] : VT(0);
        }
        VT r[KC][NRHS];
#pragma unroll
        for (int u = 0; u < KC; ++u)
#pragma unroll
            for (int k = 0; k < NRHS; ++k) r[u][k] = base + G * u < e1 ? res[(size_t)c[u] * NRHS + k] : VT(0);
#pragma unroll
        for (int u = 0; u < KC; ++u)
#pragma unroll
            for (int k = 0; k < NRHS; ++k) acc[k] += pv[u] * r[u][k];
    }
    group_sum<G>(acc);
}

// acc = sum over the CSR row [e0, e1) of val[e] * x[col[e]] (batched loads).
template <typename VT, int NRHS, int G = 1>
__device__ __forceinline__ void row_gather(const int *__restrict__ col, const VT *__restrict__ val, int e0,
                                           int e1, const VT *__restrict__ x, VT (&acc)[NRHS], int sub = 0) {
    constexpr int KC = chunk_for<G>();
#pragma unroll
    for (int k = 0; k < NRHS; ++k) acc[k] = VT(0);
    for (int base = e0 + sub; base < e1; base += G * KC) {
        int c[KC];
        VT a[KC];
#pragma unroll
        for (int u = 0; u < KC; ++u) {
            const int e = base + G * u;
            const bool in = e < e1;
            c[u] = in ? col[e] : 0;
            a[u] = in ? val[e] : VT(0);
        }
        VT xv[KC][NRHS];
#pragma unroll
        for (int u = 0; u < KC; ++u)
#pragma unroll
            for (int k = 0; k < NRHS; ++k) xv[u][k] = base + G * u < e1 ? x[(size_t)c[u] * NRHS + k] : VT(0);
#pragma unroll
        for (int u = 0; u < KC; ++u)
#pragma unroll
            for (int k = 0; k < NRHS; ++k) acc[k] += a[u] * xv[u][k];
    }
    group_sum<G>(acc);
}

// Up-sweep across a smoothed transfer without a second pass over A: with
// x~ = w D^-1 b + al P xc and A (w D^-1 b) = b - res (the down sweep),
//   x_i = x~_i + w d_i (res_i - al (T xc)_i),  T = A P.
template <typename VT, int NRHS, int G = 1>
__device__ __forceinline__ void up_sa_row(const DevAmgLevel &F, const VT *__restrict__ xc, VT w, VT al, int i,
                                          int sub = 0) {
    const VT *dinv = lvl_dinv<VT>(F);
    const VT *b = (const VT *)F.b;
    const VT *res = (const VT *)F.res;
    VT *xo = (VT *)F.x;
    VT mx[NRHS];
    row_gather<VT, NRHS, G>(F.t_col, lvl_m<VT>(F), F.t_ptr[i], F.t_ptr[i + 1], xc, mx, sub);
    if (sub != 0) return;
    const VT d = w * dinv[i];
#pragma unroll
    for (int k = 0; k < NRHS; ++k) {
        const size_t q = (size_t)i * NRHS + k;
        xo[q] = d * (b[q] + res[q]) + al * mx[k];
    }
}

// Level 0, one more Jacobi sweep: xo = x + w D^-1 (b - A x~), x~ = x + alpha P xc
// (xc == nullptr: no coarse correction).  With z/partials: the final sweep of
// the V-cycle writes the preconditioned residual and <r, z> partials.
template <typename VT, int NRHS, bool LAST, int TB>
__global__ __launch_bounds__(TB) void k_vc0_up(DevTopo T, const VT *__restrict__ val,
                                               const VT *__restrict__ dinv,
                                               const int *__restrict__ agg,
                                               const VT *__restrict__ b, const VT *__restrict__ x,
                                               const VT *__restrict__ xc, VT *__restrict__ xo,
                                               VT w, VT alpha, DevRed R) {
    const int blk = T.sweep_xcd ? xcd_block(gridDim.x) : (int)blockIdx.x;
    double part[NRHS];
#pragma unroll
    for (int k = 0; k < NRHS; ++k) part[k] = 0.0;
#pragma unroll
    for (int j = 0; j < kRowBlock / TB; ++j) {
        const int f = blk * kRowBlock + j * TB + threadIdx.x;
        const int slice = f >> 6, lane = f & 63;
        const long long s0 = T.slice_ptr[slice];
        const int wd = (int)((T.slice_ptr[slice + 1] - s0) >> 6);
        VT acc[NRHS];
        if (!xc) {
            sell_row<VT, VT, VT, NRHS>(T, val, x, f, acc);
        } else {
#pragma unroll
            for (int k = 0; k < NRHS; ++k) acc[k] = VT(0);
            for (int e = 0; e < wd; ++e) {
                const long long o = s0 + lane + 64LL * e;
                const int c = __builtin_nontemporal_load(T.col + o);
                const VT a = __builtin_nontemporal_load(val + o);
                const int gc = agg[c];
#pragma unroll
                for (int k = 0; k < NRHS; ++k) {
                    VT xv = x[(size_t)c * NRHS + k];
                    if (gc >= 0) xv += alpha * xc[(size_t)gc * NRHS + k];
                    acc[k] += a * xv;
                }
            }
        }
        const VT d = w * dinv[f];
        const int g = xc ? agg[f] : -1;
#pragma unroll
        for (int k = 0; k < NRHS; ++k) {
            const size_t i = (size_t)f * NRHS + k;
            VT xv = x[i];
            if (g >= 0) xv += alpha * xc[(size_t)g * NRHS + k];
            const VT out = xv + d * (b[i] - acc[k]);
            xo[i] = out;
            if (LAST) part[k] += (double)b[i] * (double)out;
        }
    }
    if (LAST) block_partials<NRHS, TB / 64>(part, blk, R);
}

// Width of an A P slice as a compile-time constant (0: wider than 8, taken
// 8 entries at a time).
template <class Fn>
__device__ __forceinline__ void with_ap_width(int wd, Fn fn) {
    switch (wd) {
        case 1: fn(IC<1>{}); break;
        case 2: fn(IC<2>{}); break;
        case 3: fn(IC<3>{}); break;
        case 4: fn(IC<4>{}); break;
        case 5: fn(IC<5>{}); break;
        case 6: fn(IC<6>{}); break;
        case 7: fn(IC<7>{}); break;
        case 8: fn(IC<8>{}); break;
        default: fn(IC<0>{}); break;
    }
}

// acc = sum over the W entries of one A P row (lane-major: entry e at 64 e)
// of a * xc[col]; W = 0: the runtime width wd in chunks of kChunk.
template <typename VT, int NRHS, int W>
__device__ __forceinline__ void ap_row(const int *__restrict__ col, const VT *__restrict__ val, int wd,
                                       const VT *__restrict__ xc, VT (&acc)[NRHS]) {
#pragma unroll
    for (int k = 0; k < NRHS; ++k) acc[k] = VT(0);
    if constexpr (W > 0) {
        int c[W];
        VT a[W];
#pragma unroll
        for (int u = 0; u < W; ++u) {
            c[u] = __builtin_nontemporal_load(col + 64 * u);
            a[u] = __builtin_nontemporal_load(val + 64 * u);
        }
#pragma unroll
        for (int u = 0; u < W; ++u)
#pragma unroll
            for (int k = 0; k < NRHS; ++k) acc[k] += a[u] * xc[(size_t)c[u] * NRHS + k];
    } else {
        for (int e0 = 0; e0 < wd; e0 += kChunk) {
            int c[kChunk];
            VT a[kChunk];
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                const bool in = e0 + u < wd;
                c[u] = in ? __builtin_nontemporal_load(col + 64LL * (e0 + u)) : 0;
                a[u] = in ? __builtin_nontemporal_load(val + 64LL * (e0 + u)) : VT(0);
            }
#pragma unroll
            for (int u = 0; u < kChunk; ++u)
#pragma unroll
                for (int k = 0; k < NRHS; ++k)
                    acc[k] += e0 + u < wd ? a[u] * xc[(size_t)c[u] * NRHS + k] : VT(0);
        }
    }
}

// Post-smoothing sweep after the coarse correction without reading A: with
// x~ = x0 + alpha P xc and A x0 = r - res0 (from the pre-smoothing pass),
//   z = x~ + w D^-1 (r - A x~) = x~ + w D^-1 (res0 - alpha (A P) xc),
// where A P (fine rows x level-1 aggregates, a few entries per row) is
// refilled every solve.  Also the <r, z> partials of the preconditioned CG.
template <typename VT, int NRHS, int TB>
__global__ __launch_bounds__(TB) void k_vc0_post(Gate gt, DevTopo T, const long long *__restrict__ ap_ptr,
                                                 const int *__restrict__ ap_col, const VT *__restrict__ ap,
                                                 const int *__restrict__ agg, const VT *__restrict__ dinv,
                                                 const VT *__restrict__ r, const VT *__restrict__ x0,
                                                 const VT *__restrict__ res0, const VT *__restrict__ xc,
                                                 VT *__restrict__ z, VT w, VT al, DevRed R) {
    const int blk = T.sweep_xcd ? xcd_block(gridDim.x) : (int)blockIdx.x;
    if (gate_off<NRHS>(gt, block_inst(T, blk))) return;
    double part[NRHS];
#pragma unroll
    for (int k = 0; k < NRHS; ++k) part[k] = 0.0;
#pragma unroll
    for (int j = 0; j < kRowBlock / TB; ++j) {
        const int f = blk * kRowBlock + j * TB + threadIdx.x;
        const int slice = f >> 6, lane = f & 63;
        const long long s0 = ap_ptr[slice];
        const int wd = (int)((ap_ptr[slice + 1] - s0) >> 6);
        VT acc[NRHS];
        // the slice width is wave-uniform: widths up to 8 (every slice of a
        // hexahedral grid) unroll to immediate load offsets
        with_ap_width(wd, [&](auto wc) {
            constexpr int W = decltype(wc)::value;
            ap_row<VT, NRHS, W>(ap_col + s0 + lane, ap + s0 + lane, wd, xc, acc);
        });
        const int g = agg[f];
        const VT d = w * dinv[f];
#pragma unroll
        for (int k = 0; k < NRHS; ++k) {
            const size_t i = (size_t)f * NRHS + k;
            const VT xt = x0[i] + (g >= 0 ? al * xc[(size_t)g * NRHS + k] : VT(0));
            const VT out = xt + d * (res0[i] - al * acc[k]);
            z[i] = out;
            part[k] += (double)r[i] * (double)out;
        }
    }
    block_partials<NRHS, TB / 64>(part, blk, R);
}

// Level 0: res = b - A x (after extra down sweeps).  The same matrix stream
// and gathers as the PCG SpMV, so the same workgroup shape (1024-row blocks,
// TB threads, XCD remap as tuned for the SpMV).
template <typename VT, int NRHS, int CH, int TB>
__global__ __launch_bounds__(TB) void k_vc0_res(Gate gt, DevTopo T, const VT *__restrict__ val,
                                                const VT *__restrict__ b, const VT *__restrict__ x,
                                                VT *__restrict__ res) {
    const int blk = T.spmv_xcd ? xcd_block(gridDim.x) : (int)blockIdx.x;
    if (gate_off<NRHS>(gt, block_inst(T, blk))) return;
#pragma unroll
    for (int j = 0; j < kRowBlock / TB; ++j) {
        const int f = blk * kRowBlock + j * TB + threadIdx.x;
        VT acc[NRHS];
        sell_row<VT, VT, VT, NRHS, CH>(T, val, x, f, acc);
#pragma unroll
        for (int k = 0; k < NRHS; ++k) res[(size_t)f * NRHS + k] = b[(size_t)f * NRHS + k] - acc[k];
    }
}

// Down-sweep of a level and restriction to the next, for a tentative
// (piecewise-constant) transfer: one coarse row per 8 G lanes, G lanes per
// child row (at most 8 children), the children's residuals summed across the
// child slots by shuffles.  Writes x and res of the finer level as
// k_vcc_down and the coarse right-hand side as k_restrict, in one launch.
template <typename VT, int NRHS, int G>
__global__ __launch_bounds__(256) void k_vcc_down_restrict(DevAmgLevel F, DevAmgLevel C, const VT *__restrict__ b,
                                                           VT *__restrict__ x, VT *__restrict__ res,
                                                           VT *__restrict__ bc, VT w) {
    static_assert(8 * G <= 64 && 64 % (8 * G) == 0, "a coarse row's lanes must tile a wavefront");
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int J = t / (8 * G), slot = (t / G) % 8, sub = t % G;
    if (J >= C.n) return;            // whole 8G-lane groups leave together
    const int c0 = C.child_ptr[J], nch = C.child_ptr[J + 1] - c0;
    VT part[NRHS];
#pragma unroll
    for (int k = 0; k < NRHS; ++k) part[k] = VT(0);
    if (slot < nch) {
        const int i = C.child[c0 + slot];
        const VT *dinv = lvl_dinv<VT>(F);
        VT acc[NRHS];
        row_down<VT, NRHS, G>(F, lvl_val<VT>(F), dinv, b, w, i, acc, sub);
        if (sub == 0) {
            const VT d = w * dinv[i];
#pragma unroll
            for (int k = 0; k < NRHS; ++k) {
                const size_t q = (size_t)i * NRHS + k;
                const VT r = b[q] - acc[k];
                x[q] = d * b[q];
                res[q] = r;
                part[k] = r;
            }
        }
    }
    // lanes with sub != 0 hold 0; sum the 8 slots of this coarse row
#pragma unroll
    for (int m = G; m < 8 * G; m <<= 1)
#pragma unroll
        for (int k = 0; k < NRHS; ++k) part[k] += __shfl_xor(part[k], m, 8 * G);
    if (slot == 0 && sub == 0) {
#pragma unroll
        for (int k = 0; k < NRHS; ++k) bc[(size_t)J * NRHS + k] = part[k];
    }
}

// Coarse levels (CSR), G lanes per row.
template <typename VT, int NRHS, int G = 1>
__global__ __launch_bounds__(256) void k_vcc_down(DevAmgLevel L, const VT *__restrict__ b,
                                                  VT *__restrict__ x, VT *__restrict__ res, VT w) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int i = t / G, sub = t % G;
    if (i >= L.n) return;
    const VT *dinv = lvl_dinv<VT>(L);
    VT acc[NRHS];
    row_down<VT, NRHS, G>(L, lvl_val<VT>(L), dinv, b, w, i, acc, sub);
    if (sub != 0) return;
    const VT d = w * dinv[i];
#pragma unroll
    for (int k = 0; k < NRHS; ++k) {
        const size_t q = (size_t)i * NRHS + k;
        x[q] = d * b[q];
        res[q] = b[q] - acc[k];
    }
}

template <typename VT, int NRHS>
__global__ __launch_bounds__(256) void k_vcc_res(DevAmgLevel L, const VT *__restrict__ b,
                                                 const VT *__restrict__ x, VT *__restrict__ res) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= L.n) return;
    const VT *val = lvl_val<VT>(L);
    VT acc[NRHS];
#pragma unroll
    for (int k = 0; k < NRHS; ++k) acc[k] = VT(0);
    for (int e = L.row_ptr[i]; e < L.row_ptr[i + 1]; ++e) {
        const int c = L.col[e];
#pragma unroll
        for (int k = 0; k < NRHS; ++k) acc[k] += val[e] * x[(size_t)c * NRHS + k];
    }
#pragma unroll
    for (int k = 0; k < NRHS; ++k) res[(size_t)i * NRHS + k] = b[(size_t)i * NRHS + k] - acc[k];
}

template <typename VT, int NRHS>
__global__ __launch_bounds__(256) void k_vcc_up(DevAmgLevel L, const int *__restrict__ agg,
                                                const VT *__restrict__ b, const VT *__restrict__ x,
                                                const VT *__restrict__ xc, VT *__restrict__ xo, VT w,
                                                VT alpha) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= L.n) return;
    const VT *val = lvl_val<VT>(L);
    const VT *dinv = lvl_dinv<VT>(L);
    VT acc[NRHS];
#pragma unroll
    for (int k = 0; k < NRHS; ++k) acc[k] = VT(0);
    for (int e = L.row_ptr[i]; e < L.row_ptr[i + 1]; ++e) {
        const int c = L.col[e];
        const int gc = xc ? agg[c] : -1;
#pragma unroll
        for (int k = 0; k < NRHS; ++k) {
            VT xv = x[(size_t)c * NRHS + k];
            if (gc >= 0) xv += alpha * xc[(size_t)gc * NRHS + k];
            acc[k] += val[e] * xv;
        }
    }
    const VT d = w * dinv[i];
    const int g = xc ? agg[i] : -1;
#pragma unroll
    for (int k = 0; k < NRHS; ++k) {
        const size_t q = (size_t)i * NRHS + k;
        VT xv = x[q];
        if (g >= 0) xv += alpha * xc[(size_t)g * NRHS + k];
        xo[q] = xv + d * (b[q] - acc[k]);
    }
}


// x0 = w D^-1 r (level-0 pre-smoothing when r does not come from k_update_xr)
template <typename VT, int NRHS>
__global__ __launch_bounds__(256) void k_scale0(Gate gt, const int *__restrict__ block_seg,
                                                const VT *__restrict__ dinv, const VT *__restrict__ r,
                                                VT *__restrict__ x0, VT w) {
    const int f = blockIdx.x * 256 + threadIdx.x;
    if (gate_off<NRHS>(gt, block_seg[f / kRowBlock])) return;
    const VT d = w * dinv[f];
#pragma unroll
    for (int k = 0; k < NRHS; ++k) x0[(size_t)f * NRHS + k] = d * r[(size_t)f * NRHS + k];
}

// x~ = x + alpha P xc on level 0 (padding rows have no aggregate)
template <typename VT, int NRHS>
__global__ __launch_bounds__(256) void k_prolong0(const int *__restrict__ agg, const VT *__restrict__ x,
                                                  const VT *__restrict__ xc, VT *__restrict__ xo,
                                                  VT alpha) {
    const int f = blockIdx.x * 256 + threadIdx.x;
    const int g = agg[f];
#pragma unroll
    for (int k = 0; k < NRHS; ++k) {
        const size_t i = (size_t)f * NRHS + k;
        xo[i] = x[i] + (g >= 0 ? alpha * xc[(size_t)g * NRHS + k] : VT(0));
    }
}

// Coarse level up with one sweep: x~ = w D^-1 b + alpha P xc (on the fly),
// xo = x~ + w D^-1 (b - A x~).
template <typename VT, int NRHS, int G = 1>
__global__ __launch_bounds__(256) void k_vcc_up1(DevAmgLevel L, const int *__restrict__ agg,
                                                 const VT *__restrict__ xc, VT *__restrict__ xo, VT w,
                                                 VT alpha) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int i = t / G, sub = t % G;
    if (i >= L.n) return;
    const VT *dinv = lvl_dinv<VT>(L);
    const VT *b = (const VT *)L.b;
    VT acc[NRHS];
    row_up1<VT, NRHS, G>(L, lvl_val<VT>(L), dinv, b, agg, xc, w, alpha, i, acc, sub);
    if (sub != 0) return;
    const VT d = w * dinv[i];
    const int g = agg[i];
#pragma unroll
    for (int k = 0; k < NRHS; ++k) {
        const size_t q = (size_t)i * NRHS + k;
        const VT xv = d * b[q] + alpha * xc[(size_t)g * NRHS + k];
        xo[q] = xv + d * (b[q] - acc[k]);
    }
}

// x~ = x + al P xc across a smoothed transfer (the path with several sweeps).
template <typename VT, int NRHS>
__global__ __launch_bounds__(256) void k_prolong_sa(DevAmgLevel F, const VT *__restrict__ x,
                                                    const VT *__restrict__ xc, VT al, VT *__restrict__ xo) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= F.n) return;
    VT px[NRHS];
    row_gather<VT, NRHS>(F.p_col, lvl_p<VT>(F), F.p_ptr[i], F.p_ptr[i + 1], xc, px);
#pragma unroll
    for (int k = 0; k < NRHS; ++k) xo[(size_t)i * NRHS + k] = x[(size_t)i * NRHS + k] + al * px[k];
}

// Restriction across a smoothed transfer: b_c = P' res.
template <typename VT, int NRHS, int G = 1>
__global__ __launch_bounds__(256) void k_restrict_sa(DevAmgLevel F, int nc, VT *__restrict__ bc) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int I = t / G, sub = t % G;
    if (I >= nc) return;
    VT acc[NRHS];
    row_restrict_sa<VT, NRHS, G>(F, (const VT *)F.res, I, acc, sub);
    if (sub != 0) return;
#pragma unroll
    for (int k = 0; k < NRHS; ++k) bc[(size_t)I * NRHS + k] = acc[k];
}

// PRO: the transfer from this level to the finer one is tentative and covers
// it, so the launch also prolongs its result into the finer level's iterate
// in place, xf[c] += alf xo[i] for the children c of row i (the finer level's
// up-sweep then gathers one value per entry).
template <typename VT, int NRHS, int G, bool PRO>
__global__ __launch_bounds__(256) void k_vcc_up_sa(DevAmgLevel F, const VT *__restrict__ xc, VT w,
                                                   VT al, VT *__restrict__ xf, VT alf) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int i = t / G, sub = t % G;
    if (i >= F.n) return;
    if constexpr (!PRO) {
        up_sa_row<VT, NRHS, G>(F, xc, w, al, i, sub);
    } else {
        const VT *b = (const VT *)F.b;
        const VT *res = (const VT *)F.res;
        VT mx[NRHS];
        row_gather<VT, NRHS, G>(F.t_col, lvl_m<VT>(F), F.t_ptr[i], F.t_ptr[i + 1], xc, mx, sub);
        const VT d = w * lvl_dinv<VT>(F)[i];
        VT o[NRHS];
#pragma unroll
        for (int k = 0; k < NRHS; ++k) {   // every lane of the group holds the row sum
            const size_t q = (size_t)i * NRHS + k;
            o[k] = d * (b[q] + res[q]) + al * mx[k];
        }
        if (sub == 0) {
#pragma unroll
            for (int k = 0; k < NRHS; ++k) ((VT *)F.x)[(size_t)i * NRHS + k] = o[k];
        }
        for (int c = F.child_ptr[i] + sub; c < F.child_ptr[i + 1]; c += G) {
            const int f = F.child[c];
#pragma unroll
            for (int k = 0; k < NRHS; ++k) xf[(size_t)f * NRHS + k] += alf * o[k];
        }
    }
}

// Restriction b_c = P' res: sum over the coarse row's children.
template <typename VT, int NRHS, int G = 1>
__global__ __launch_bounds__(256) void k_restrict(int n, const int *__restrict__ child_ptr,
                                                  const int *__restrict__ child,
                                                  const VT *__restrict__ res, VT *__restrict__ bc) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int I = t / G, sub = t % G;
    if (I >= n) return;
    VT acc[NRHS];
    row_restrict<VT, NRHS, G>(child_ptr, child, res, I, acc, sub);
    if (sub != 0) return;
#pragma unroll
    for (int k = 0; k < NRHS; ++k) bc[(size_t)I * NRHS + k] = acc[k];
}

// Coarsest level: x = A^-1 b with the per-instance dense inverse.
template <typename VT, int NRHS>
__global__ __launch_bounds__(256) void k_coarse_solve(const int *__restrict__ inst_ptr,
                                                      const int *__restrict__ off,
                                                      const VT *__restrict__ inv,
                                                      const VT *__restrict__ b, VT *__restrict__ x) {
    const int t = blockIdx.x;
    const int r0 = inst_ptr[t], m = inst_ptr[t + 1] - r0;
    const VT *A = inv + off[t];
    for (int e = threadIdx.x; e < m * NRHS; e += 256) {
        const int i = e / NRHS, k = e % NRHS;
        VT s = VT(0);
        for (int j = 0; j < m; ++j) s += A[i * m + j] * b[(size_t)(r0 + j) * NRHS + k];
        x[(size_t)(r0 + i) * NRHS + k] = s;
    }
}

// Coarse-level up-sweep from an iterate the coarser level's kernel already
// prolonged in place (x~ = w D^-1 b + al P xc): xo = x~ + w D^-1 (b - A x~),
// one gather per entry (k_vcc_up1 forms x~ of every column on the fly, four
// dependent gathers).  PRO: the transfer from this level to the finer one is
// tentative and covers it, so the launch also prolongs xo into the finer
// level's iterate in place, xf[c] += al xo[i] for the children c of row i.
template <typename VT, int NRHS, int G, bool PRO>
__global__ __launch_bounds__(256) void k_vcc_up_plain(DevAmgLevel F, const VT *__restrict__ x, VT *__restrict__ xo,
                                                      VT *__restrict__ xf, VT w, VT al) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int i = t / G, sub = t % G;
    if (i >= F.n) return;
    VT acc[NRHS];
    row_gather<VT, NRHS, G>(F.col, lvl_val<VT>(F), F.row_ptr[i], F.row_ptr[i + 1], x, acc, sub);
    const VT d = w * lvl_dinv<VT>(F)[i];
    const VT *b = (const VT *)F.b;
    VT o[NRHS];
#pragma unroll
    for (int k = 0; k < NRHS; ++k) {
        const size_t q = (size_t)i * NRHS + k;
        o[k] = x[q] + d * (b[q] - acc[k]);   // every lane of the group holds the row sum
    }
    if (sub == 0) {
#pragma unroll
        for (int k = 0; k < NRHS; ++k) xo[(size_t)i * NRHS + k] = o[k];
    }
    if constexpr (PRO) {
        for (int c = F.child_ptr[i] + sub; c < F.child_ptr[i + 1]; c += G) {
            const int f = F.child[c];
#pragma unroll
            for (int k = 0; k < NRHS; ++k) xf[(size_t)f * NRHS + k] += al * o[k];
        }
    }
}

// The small levels lt..L-1 of one V(1,1) cycle (one sweep each way), one
// workgroup per instance: the same arithmetic as k_vcc_down + k_restrict,
// k_coarse_solve and k_vcc_up1, with workgroup barriers between the steps
// instead of kernel launches.  Reads lev[lt].b, writes lev[l].x for l >= lt.
template <typename VT, int NRHS>
__global__ __launch_bounds__(1024) void k_vcycle_tail(Gate gt, const DevAmg *__restrict__ Ap, int lt, VT al,
                                                      VT als,
                                                      VT *__restrict__ xf) {
    const DevAmg &A = *Ap;
    const int s = blockIdx.x;
    if (gate_off<NRHS>(gt, s)) return;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int L = A.nlev;
    for (int l = lt; l < L - 1; ++l) {
        const DevAmgLevel &F = A.lev[l];
        const DevAmgLevel &C = A.lev[l + 1];
        const VT w = (VT)A.omega_l[l];
        const VT *val = lvl_val<VT>(F);
        const VT *dinv = lvl_dinv<VT>(F);
        const VT *b = (const VT *)F.b;
        VT *res = (VT *)F.res;
        const int r0 = F.inst_ptr[s], r1 = F.inst_ptr[s + 1];
        for (int i = r0 + tid; i < r1; i += nt) {
            VT acc[NRHS];
            row_down<VT, NRHS>(F, val, dinv, b, w, i, acc);
#pragma unroll
            for (int k = 0; k < NRHS; ++k) res[(size_t)i * NRHS + k] = b[(size_t)i * NRHS + k] - acc[k];
        }
        __syncthreads();
        VT *bc = (VT *)C.b;
        for (int I = C.inst_ptr[s] + tid; I < C.inst_ptr[s + 1]; I += nt) {
            VT acc[NRHS];
            if (F.sa_out) row_restrict_sa<VT, NRHS>(F, res, I, acc);
            else row_restrict<VT, NRHS>(C.child_ptr, C.child, res, I, acc);
#pragma unroll
            for (int k = 0; k < NRHS; ++k) bc[(size_t)I * NRHS + k] = acc[k];
        }
        __syncthreads();
    }
    {
        // coarsest level: x = A^-1 b with the dense inverse, one wavefront per
        // (row, rhs) and four of them per pass: lane j takes column j (the
        // row-major inverse is read coalesced), a shuffle sum finishes the row
        const DevAmgLevel &C = A.lev[L - 1];
        const int r0 = C.inst_ptr[s], m = C.inst_ptr[s + 1] - r0;
        const VT *inv;
        if constexpr (std::is_same<VT, float>::value) inv = A.cinv32 + A.coarse_off[s];
        else inv = A.cinv64 + A.coarse_off[s];
        const VT *b = (const VT *)C.b;
        VT *x = (VT *)C.x;
        const int lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
        const int rows = m * NRHS;
        for (int e0 = wid; e0 < rows; e0 += 4 * nw) {
            VT acc[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                acc[u] = VT(0);
                const int e = e0 + u * nw;
                if (e >= rows) continue;
                const int i = e / NRHS, k = e % NRHS;
                for (int j = lane; j < m; j += 64) acc[u] += inv[i * m + j] * b[(size_t)(r0 + j) * NRHS + k];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) acc[u] += __shfl_xor(acc[u], off, 64);
                const int e = e0 + u * nw;
                if (lane == 0 && e < rows) x[(size_t)(r0 + e / NRHS) * NRHS + e % NRHS] = acc[u];
            }
        }
        __syncthreads();
    }
    for (int l = L - 2; l >= lt; --l) {
        const DevAmgLevel &F = A.lev[l];
        const VT w = (VT)A.omega_l[l];
        const int *agg = A.lev[l + 1].agg;
        const VT *xc = (const VT *)A.lev[l + 1].x;
        if (F.sa_out) {
            for (int i = F.inst_ptr[s] + tid; i < F.inst_ptr[s + 1]; i += nt) up_sa_row<VT, NRHS>(F, xc, w, als, i);
            __syncthreads();
            continue;
        }
        const VT *val = lvl_val<VT>(F);
        const VT *dinv = lvl_dinv<VT>(F);
        const VT *b = (const VT *)F.b;
        VT *xo = (VT *)F.x;
        for (int i = F.inst_ptr[s] + tid; i < F.inst_ptr[s + 1]; i += nt) {
            VT acc[NRHS];
            row_up1<VT, NRHS>(F, val, dinv, b, agg, xc, w, al, i, acc);
            const VT d = w * dinv[i];
            const int g = agg[i];
#pragma unroll
            for (int k = 0; k < NRHS; ++k) {
                const size_t q = (size_t)i * NRHS + k;
                const VT xv = d * b[q] + al * xc[(size_t)g * NRHS + k];
                xo[q] = xv + d * (b[q] - acc[k]);
            }
        }
        __syncthreads();
    }
    if (xf) {
        // the transfer from level lt to the finer level is tentative and
        // covers it: prolong lev[lt].x into that level's iterate in place,
        // x~_f = w D^-1 b + al x_lt[agg_f] (the down-sweep wrote w D^-1 b)
        // (one thread per finer row: every row has its aggregate)
        const DevAmgLevel &C = A.lev[lt];
        const DevAmgLevel &F = A.lev[lt - 1];
        const VT *xc = (const VT *)C.x;
        for (int f = F.inst_ptr[s] + tid; f < F.inst_ptr[s + 1]; f += nt) {
            const int J = C.agg[f];
#pragma unroll
            for (int k = 0; k < NRHS; ++k) xf[(size_t)f * NRHS + k] += al * xc[(size_t)J * NRHS + k];
        }
    }
}

// p = z + beta p (beta = 0 on a restart); in FP32 also fold x32 into x64 for
// systems whose residual replacement is due.
template <typename VT, int NRHS>
__global__ __launch_bounds__(256) void k_update_pz(DevTopo T, const VT *__restrict__ z,
                                                   VT *__restrict__ p, VT *__restrict__ x,
                                                   double *__restrict__ x64, DevScal S) {
    const int b = xcd_block(gridDim.x);
    const int s = block_inst(T, b);
    int act[NRHS];
    if (!load_act<NRHS>(S.act, s, act)) return;
    VT be[NRHS];
    int u[NRHS];
    bool any_u = false;
#pragma unroll
    for (int k = 0; k < NRHS; ++k) {
        const int sys = s * NRHS + k;
        u[k] = (x64 && act[k]) ? S.upd[sys] : 0;
        any_u |= u[k] != 0;
        be[k] = (VT)S.sc[(size_t)sys * SC_N + SC_BETA];
    }
    const int f0 = b * kRowBlock + 4 * threadIdx.x;
    const size_t e0 = (size_t)f0 * NRHS;
    VT zv[4 * NRHS], pv[4 * NRHS];
    vload<VT, 4 * NRHS>(z + e0, zv);
    vload<VT, 4 * NRHS>(p + e0, pv);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int k = 0; k < NRHS; ++k) {
            const int i = j * NRHS + k;
            if (act[k]) pv[i] = zv[i] + be[k] * pv[i];
        }
    vstore<VT, 4 * NRHS>(p + e0, pv);
    if (any_u) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int k = 0; k < NRHS; ++k) {
                if (!u[k]) continue;
                const size_t i = e0 + j * NRHS + k;
                x64[i] += (double)x[i];
                x[i] = VT(0);
            }
    }
}

// ---------------------------------------------------------------------------
// Recovery, per cell: p = (g + F.pi)/L;  v = totmob B^-1 (p - pi) + rhomob gflux.
template <int MAXN, int NRHS, bool UNIF>
__global__ __launch_bounds__(64) void k_recover(DevTopo T, const double *__restrict__ binv,
                                                const double *__restrict__ gfs,
                                                const double *__restrict__ totmob,
                                                const double *__restrict__ rhomob,
                                                DevCellSys C, const double *__restrict__ x,
                                                double *__restrict__ press,
                                                double *__restrict__ outflux) {
    constexpr int NP = MAXN * (MAXN + 1) / 2;
    const int c = blockIdx.x * 64 + threadIdx.x;
    if (c >= T.nc) return;
    const int f0 = T.facepos[c], n = UNIF ? MAXN : T.facepos[c + 1] - f0;
    const double lam = totmob[c], rl = rhomob[c];
    double B[NP];
#pragma unroll
    for (int i = 0; i < MAXN; ++i)
#pragma unroll
        for (int j = i; j < MAXN; ++j) B[pidx(i, j, MAXN)] = (j < n) ? binv[(size_t)pidx(i, j, n) * T.nc + c] : 0.0;
    double F[MAXN];
    int dof[MAXN];
#pragma unroll
    for (int i = 0; i < MAXN; ++i) {
        F[i] = (i < n) ? C.F[f0 + i] : 0.0;
        dof[i] = (i < n) ? T.hf_dof[f0 + i] : 0;
    }
    const double L = C.L[c];
#pragma unroll
    for (int k = 0; k < NRHS; ++k) {
        double pi[MAXN], s = 0.0;
#pragma unroll
        for (int i = 0; i < MAXN; ++i) pi[i] = (i < n) ? x[(size_t)dof[i] * NRHS + k] : 0.0;
#pragma unroll
        for (int i = 0; i < MAXN; ++i) s += F[i] * pi[i];
        const double p = (C.g[(size_t)k * T.nc + c] + s) / L;
        press[(size_t)k * T.nc + c] = p;
#pragma unroll
        for (int i = 0; i < MAXN; ++i) pi[i] = (i < n) ? p - pi[i] : 0.0;
#pragma unroll
        for (int i = 0; i < MAXN; ++i) {
            if (i >= n) continue;
            double v = 0.0;
#pragma unroll
            for (int j = 0; j < MAXN; ++j)
                v += lam * ((i <= j) ? B[pidx(i, j, MAXN)] : B[pidx(j, i, MAXN)]) * pi[j];
            outflux[(size_t)k * T.nhf + f0 + i] = gfs[f0 + i] * rl + v;
        }
    }
}

// Twin half-face averaging: v <- (v - v_twin)/2 for internal and periodic
// pairs; the largest change is folded into *maxbits (non-negative doubles
// order like their bit patterns).
__global__ __launch_bounds__(256) void k_post(DevTopo T, int nrhs, const double *__restrict__ vin,
                                              double *__restrict__ vout,
                                              unsigned long long *__restrict__ maxbits) {
    const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long tot = (long long)nrhs * T.nhf;
    double d = 0.0;
    if (gid < tot) {
        const int k = (int)(gid / T.nhf), h = (int)(gid % T.nhf);
        const int tw = T.hf_twin[h];
        const double v = vin[gid];
        double nv = v;
        if (tw >= 0) nv = 0.5 * (v - vin[(size_t)k * T.nhf + tw]);
        vout[gid] = nv;
        d = fabs(v - nv);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) d = fmax(d, __shfl_down(d, off, 64));
    if ((threadIdx.x & 63) == 0 && d > 0.0)
        atomicMax(maxbits, (unsigned long long)__double_as_longlong(d));
}

// Sums of outflux over groups of half-faces: one wavefront per (group, rhs),
// strided accumulation then a fixed-order shuffle reduction (deterministic).
__global__ __launch_bounds__(64) void k_group_sum(int ngroups, int nhf, const int *__restrict__ gptr,
                                                  const int *__restrict__ ghf, const double *__restrict__ v,
                                                  double *__restrict__ out) {
    const int grp = blockIdx.x, k = blockIdx.y;
    double s = 0.0;
    for (int e = gptr[grp] + (int)threadIdx.x; e < gptr[grp + 1]; e += 64) s += v[(size_t)k * nhf + ghf[e]];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (threadIdx.x == 0) out[(size_t)k * ngroups + grp] = s;
}

// ---------------------------------------------------------------------------
// Adjoint of the recovered fluxes.  Cell c's flux at fixed face pressures is
//   v_c = B (p 1 - pi_c) + rl gf_c = -S_c pi_c + F g/L + rl gf_c,
//   B = lam Binv, F = B 1, L = 1'F, p = (g + F'pi_c)/L, S_c = B - F F'/L,
// so for an objective J(v) with w = dJ/dv the adjoint system is
//   S mu = -sum_c P_c' S_c w_c   (zero on Dirichlet rows, which fix pi)
// and dJ/dtheta = sum_c u_c' dv_c/dtheta at fixed pi, u_c = w_c + P_c mu.

// y = -S_c w_c per half-face.
template <int MAXN, int NRHS>
__global__ __launch_bounds__(64) void k_adj_cell(DevTopo T, const double *__restrict__ binv,
                                                 const double *__restrict__ totmob, DevCellSys C,
                                                 const double *__restrict__ w, double *__restrict__ y) {
    const int c = blockIdx.x * 64 + threadIdx.x;
    if (c >= T.nc) return;
    const int f0 = T.facepos[c], n = T.facepos[c + 1] - f0;
    const double lam = totmob[c], L = C.L[c];
    double F[MAXN];
#pragma unroll
    for (int i = 0; i < MAXN; ++i) F[i] = (i < n) ? C.F[f0 + i] : 0.0;
#pragma unroll
    for (int k = 0; k < NRHS; ++k) {
        double wv[MAXN], fw = 0.0;
#pragma unroll
        for (int i = 0; i < MAXN; ++i) {
            wv[i] = (i < n) ? w[(size_t)k * T.nhf + f0 + i] : 0.0;
            fw += F[i] * wv[i];
        }
#pragma unroll
        for (int i = 0; i < MAXN; ++i) {
            if (i >= n) continue;
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < MAXN; ++j)
                if (j < n) s += binv[(size_t)(i <= j ? pidx(i, j, n) : pidx(j, i, n)) * T.nc + c] * wv[j];
            y[(size_t)k * T.nhf + f0 + i] = -(lam * s - F[i] * fw / L);
        }
    }
}

// Adjoint right-hand side per row: the sum of the row's half-face values, zero
// on Dirichlet and padding rows; the initial iterate is zero.
template <int NRHS>
__global__ __launch_bounds__(256) void k_adj_rows(DevTopo T, const uint8_t *__restrict__ ftype,
                                                  const double *__restrict__ y, double *__restrict__ b,
                                                  double *__restrict__ x, double *__restrict__ r) {
    const int f = blockIdx.x * 256 + threadIdx.x;
    if (f >= T.nfp) return;
    const int h0 = T.dof_hf[f];
    double acc[NRHS];
#pragma unroll
    for (int k = 0; k < NRHS; ++k) acc[k] = 0.0;
    if (h0 >= 0 && ftype[h0] != FT_DIRICHLET) {
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int h = T.row_hf[2 * (size_t)f + t];
            if (h < 0) continue;
#pragma unroll
            for (int k = 0; k < NRHS; ++k) acc[k] += y[(size_t)k * T.nhf + h];
        }
    }
#pragma unroll
    for (int k = 0; k < NRHS; ++k) {
        b[(size_t)f * NRHS + k] = acc[k];
        x[(size_t)f * NRHS + k] = 0.0;
        r[(size_t)f * NRHS + k] = acc[k];   // start residual b - A 0
    }
}

// Gradient per cell.  With Binv = (N K N' + t W)/|c|, t = 2 tr K,
// W = A (I - Q Q') A and gf = N K g:
//   dBinv/dK_lm = (N_l N_m' + 2 delta_lm W)/|c|,  d gf/dK_lm = N_l g_m,
// and for any parameter dv = B' d + F p' + rl gf',  d = p 1 - pi,
//   p' = (g' + 1'B' pi - p 1'B' 1)/L,  g' = -sum rl gf'.
// Outputs (any may be null): gK[9c + 3l + m] = dJ/dK_lm with K entering
// through sym(K) as in k_static_ip, gtm[c] = dJ/dlam, grm[c] = dJ/drl (summed over right-hand
// sides), gsrc[k nc + c] = dJ/dsrc_c = u'F/L.
template <int MAXN, int NRHS>
__global__ __launch_bounds__(64) void k_adj_grad(DevTopo T, DevGeom G, const double *__restrict__ binv,
                                                 const double *__restrict__ gfs,
                                                 const double *__restrict__ totmob,
                                                 const double *__restrict__ rhomob, double gx, double gy,
                                                 double gz, DevCellSys C, const double *__restrict__ press,
                                                 const double *__restrict__ pi, const double *__restrict__ mu,
                                                 const uint8_t *__restrict__ ftype,
                                                 const double *__restrict__ w, double *__restrict__ gK,
                                                 double *__restrict__ gtm, double *__restrict__ grm,
                                                 double *__restrict__ gsrc) {
    const int c = blockIdx.x * 64 + threadIdx.x;
    if (c >= T.nc) return;
    const int f0 = T.facepos[c], n = T.facepos[c + 1] - f0;
    const double lam = totmob[c], rl = rhomob[c], L = C.L[c], vol = G.cvol[c];
    const double g[3] = {gx, gy, gz};
    double N[MAXN][3], Q[MAXN][3], a[MAXN];
    cell_frame<MAXN>(G, c, f0, n, N, Q, a);
    double F[MAXN], gf[MAXN];
    int dof[MAXN];
    bool dir[MAXN];
    double e[3] = {0.0, 0.0, 0.0}, aq[3] = {0.0, 0.0, 0.0}, a2 = 0.0, sg = 0.0;
#pragma unroll
    for (int i = 0; i < MAXN; ++i) {
        const bool in = i < n;
        F[i] = in ? C.F[f0 + i] : 0.0;
        gf[i] = in ? gfs[f0 + i] : 0.0;
        dof[i] = in ? T.hf_dof[f0 + i] : 0;
        dir[i] = in ? ftype[f0 + i] == FT_DIRICHLET : true;
        sg += gf[i];
        a2 += a[i] * a[i];
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            e[l] += N[i][l];
            aq[l] += a[i] * Q[i][l];
        }
    }
    const double oWo = a2 - (aq[0] * aq[0] + aq[1] * aq[1] + aq[2] * aq[2]);
    const double rv = lam / vol;
    double GK[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) GK[q] = 0.0;
    double gt = 0.0, gr = 0.0;
#pragma unroll
    for (int k = 0; k < NRHS; ++k) {
        const double p = press[(size_t)k * T.nc + c], gc = C.g[(size_t)k * T.nc + c];
        double u[MAXN], pv[MAXN], d[MAXN];
#pragma unroll
        for (int i = 0; i < MAXN; ++i) {
            if (i < n) {
                const size_t r = (size_t)dof[i] * NRHS + k;
                u[i] = w[(size_t)k * T.nhf + f0 + i] + (dir[i] ? 0.0 : mu[r]);
                pv[i] = pi[r];
                d[i] = p - pv[i];
            } else {
                u[i] = pv[i] = d[i] = 0.0;
            }
        }
        double aU[3] = {0.0, 0.0, 0.0}, bD[3] = {0.0, 0.0, 0.0}, fP[3] = {0.0, 0.0, 0.0};
        double uq[3] = {0.0, 0.0, 0.0}, dq[3] = {0.0, 0.0, 0.0}, pq[3] = {0.0, 0.0, 0.0};
        double ua2d = 0.0, a2p = 0.0, uF = 0.0, ugf = 0.0;
#pragma unroll
        for (int i = 0; i < MAXN; ++i) {
#pragma unroll
            for (int l = 0; l < 3; ++l) {
                aU[l] += u[i] * N[i][l];
                bD[l] += N[i][l] * d[i];
                fP[l] += N[i][l] * pv[i];
                uq[l] += u[i] * a[i] * Q[i][l];
                dq[l] += d[i] * a[i] * Q[i][l];
                pq[l] += pv[i] * a[i] * Q[i][l];
            }
            ua2d += u[i] * a[i] * a[i] * d[i];
            a2p += a[i] * a[i] * pv[i];
            uF += u[i] * F[i];
            ugf += u[i] * gf[i];
        }
        const double uWd = ua2d - (uq[0] * dq[0] + uq[1] * dq[1] + uq[2] * dq[2]);
        const double oWp = a2p - (aq[0] * pq[0] + aq[1] * pq[1] + aq[2] * pq[2]);
#pragma unroll
        for (int l = 0; l < 3; ++l)
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const double del = (l == m) ? 2.0 : 0.0;
                const double uBd = rv * (aU[l] * bD[m] + del * uWd);
                const double oBp = rv * (e[l] * fP[m] + del * oWp);
                const double oBo = rv * (e[l] * e[m] + del * oWo);
                const double pp = (-rl * e[l] * g[m] + oBp - p * oBo) / L;
                GK[3 * l + m] += uBd + uF * pp + rl * aU[l] * g[m];
            }
        // lam: B' = Binv, p' = -g/(lam L)
        double uBd = 0.0;
#pragma unroll
        for (int i = 0; i < MAXN; ++i) {
            if (i >= n) continue;
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < MAXN; ++j)
                if (j < n) s += binv[(size_t)(i <= j ? pidx(i, j, n) : pidx(j, i, n)) * T.nc + c] * d[j];
            uBd += u[i] * s;
        }
        gt += uBd - uF * gc / (lam * L);
        // rl: gf' = gf, g' = -sum gf, p' = g'/L
        gr += ugf - uF * sg / L;
        if (gsrc) gsrc[(size_t)k * T.nc + c] = uF / L;
    }
    // the inner products use sym(K), so dJ/dK_lm = (G_lm + G_ml)/2
    if (gK)
#pragma unroll
        for (int l = 0; l < 3; ++l)
#pragma unroll
            for (int m = 0; m < 3; ++m) gK[9 * (size_t)c + 3 * l + m] = 0.5 * (GK[3 * l + m] + GK[3 * m + l]);
    if (gtm) gtm[c] = gt;
    if (grm) grm[c] = gr;
}

// ---------------------------------------------------------------------------
// Finishers of the per-instance reductions (run by k_finish on thread 0).
template <int NRHS>
__device__ __forceinline__ bool none_set(const int *flags, int s) {
    bool any = false;
#pragma unroll
    for (int k = 0; k < NRHS; ++k) any |= flags[s * NRHS + k] != 0;
    return !any;
}

// after k_spmv_pq: alpha = rz / <p, Ap>; with REL also the FP64 residual
// replacement (the true residual decides convergence of the FP32 iteration).
template <int NRHS, bool REL>
struct FinPQ {
    __device__ static bool skip(int s, const DevScal &S, const FinArgs &) { return none_set<NRHS>(S.act, s); }
    __device__ static int nq(int s, const DevScal &S) {
        if (!REL) return NRHS;
#pragma unroll
        for (int k = 0; k < NRHS; ++k)
            if (S.act[s * NRHS + k] && S.upd[s * NRHS + k]) return 4 * NRHS;
        return NRHS;
    }
    __device__ static void run(int s, const double *tot, DevScal &S, const FinArgs &) {
#pragma unroll
        for (int k = 0; k < NRHS; ++k) {
            const int i = s * NRHS + k;
            if (!S.act[i]) continue;
            double *sc = S.sc + (size_t)i * SC_N;
            if (REL && S.upd[i]) {
                const double rrt = tot[NRHS + 3 * k];
                sc[SC_RRT] = rrt;
                sc[SC_RRL] = rrt;
                if (!S.amg) sc[SC_RZ] = tot[NRHS + 3 * k + 1];
                sc[SC_RR] = tot[NRHS + 3 * k + 2];
                S.upd[i] = 0;
                atomicAdd(&S.status[2], 1);
                if (rrt <= sc[SC_THR]) {
                    S.act[i] = 0;
                    sc[SC_ALPHA] = 0.0;
                    atomicSub(&S.status[0], 1);
                    continue;
                }
            }
            const double pq = tot[k];
            sc[SC_PQ] = pq;
            if (!(pq > 0.0)) {
                sc[SC_ALPHA] = 0.0;
                if (REL) {          // FP32 breakdown: restart p, then replace the residual
                    S.upd[i] = 2;
                } else {
                    S.act[i] = 0;
                    atomicSub(&S.status[0], 1);
                    atomicAdd(&S.status[1], 1);
                }
            } else {
                sc[SC_ALPHA] = sc[SC_RZ] / pq;
            }
        }
    }
};

// after k_update_xr: convergence test (FP64) or replacement request (FP32), beta
template <int NRHS>
struct FinXR {
    __device__ static bool skip(int s, const DevScal &S, const FinArgs &) { return none_set<NRHS>(S.act, s); }
    __device__ static int nq(int, const DevScal &) { return 2 * NRHS; }
    __device__ static void run(int s, const double *tot, DevScal &S, const FinArgs &) {
#pragma unroll
        for (int k = 0; k < NRHS; ++k) {
            const int i = s * NRHS + k;
            if (!S.act[i]) continue;
            double *sc = S.sc + (size_t)i * SC_N;
            S.iters[i] += 1;
            const double rz = tot[2 * k], rr = tot[2 * k + 1];
            sc[SC_RR] = rr;
            if (S.mixed) {
                // FP32 recursion: ask for the FP64 true residual once the recursive
                // one has dropped by replace_factor (or reached the tolerance).
                if (S.upd[i] == 0 && (rr <= sc[SC_THR] || rr <= S.rf2 * sc[SC_RRL])) S.upd[i] = 1;
                if (!S.amg) {
                    sc[SC_BETA] = rz / sc[SC_RZ];
                    sc[SC_RZ] = rz;
                }
            } else if (rr <= sc[SC_THR]) {
                S.act[i] = 0;
                atomicSub(&S.status[0], 1);
            } else if (!S.amg) {
                sc[SC_BETA] = rz / sc[SC_RZ];
                sc[SC_RZ] = rz;
            }
        }
    }
};

// after the multigrid V-cycle: beta = <r, z> / <r, z>_old (init: p = z)
template <int NRHS>
struct FinRZ {
    __device__ static bool skip(int s, const DevScal &S, const FinArgs &) { return none_set<NRHS>(S.act, s); }
    __device__ static int nq(int, const DevScal &) { return NRHS; }
    __device__ static void run(int s, const double *tot, DevScal &S, const FinArgs &A) {
#pragma unroll
        for (int k = 0; k < NRHS; ++k) {
            const int i = s * NRHS + k;
            if (!S.act[i]) continue;
            double *sc = S.sc + (size_t)i * SC_N;
            const double rz = tot[k];
            sc[SC_BETA] = (A.init || S.upd[i] == 2) ? 0.0 : rz / sc[SC_RZ];
            sc[SC_RZ] = rz;
        }
    }
};

// Multigrid iteration: the update_xr step (convergence / replacement request,
// iteration count) and beta = <r, z> / <r, z>_old in one launch after the
// V-cycle; update_xr's partials at q 0..2 NRHS-1, the V-cycle's <r, z> at
// 2 NRHS + k.  Deciding after the V-cycle only lets a system that converged in
// this iteration (FP64 recurrence) run one more V-cycle, whose result nothing
// reads.
template <int NRHS>
struct FinXRZ {
    __device__ static bool skip(int s, const DevScal &S, const FinArgs &) { return none_set<NRHS>(S.act, s); }
    __device__ static int nq(int, const DevScal &) { return 3 * NRHS; }
    __device__ static void run(int s, const double *tot, DevScal &S, const FinArgs &A) {
        FinXR<NRHS>::run(s, tot, S, A);
        FinRZ<NRHS>::run(s, tot + 2 * NRHS, S, A);
    }
};

// after k_start64: |r0|, thresholds, activation (dune CGSolver: def0 < 1e-30 -> done)
template <int NRHS>
struct FinStart {
    __device__ static bool skip(int, const DevScal &, const FinArgs &) { return false; }
    __device__ static int nq(int, const DevScal &) { return 2 * NRHS; }
    __device__ static void run(int s, const double *tot, DevScal &S, const FinArgs &A) {
#pragma unroll
        for (int k = 0; k < NRHS; ++k) {
            const int i = s * NRHS + k;
            double *sc = S.sc + (size_t)i * SC_N;
            const double rr0 = tot[2 * k + 1];
            sc[SC_RZ] = tot[2 * k];
            sc[SC_RR] = rr0;
            sc[SC_RR0] = rr0;
            sc[SC_RRT] = rr0;
            sc[SC_THR] = A.tol2 * rr0;
            S.iters[i] = 0;
            const int a = (rr0 >= 1e-60) ? 1 : 0;
            S.act[i] = a;
            if (a) atomicAdd(&S.status[0], 1);
        }
    }
};

// after k_reliable: the FP64 true residual decides convergence
template <int NRHS>
struct FinRel {
    __device__ static bool skip(int s, const DevScal &S, const FinArgs &A) {
        return !A.init && none_set<NRHS>(S.upd, s);
    }
    __device__ static int nq(int, const DevScal &) { return 3 * NRHS; }
    __device__ static void run(int s, const double *tot, DevScal &S, const FinArgs &A) {
#pragma unroll
        for (int k = 0; k < NRHS; ++k) {
            const int i = s * NRHS + k;
            if (!A.init && !S.upd[i]) continue;
            double *sc = S.sc + (size_t)i * SC_N;
            const double rrt = tot[3 * k];
            if (A.init) {
                sc[SC_RR0] = rrt;
                sc[SC_THR] = A.tol2 * rrt;
                S.iters[i] = 0;
            }
            sc[SC_RRT] = rrt;
            sc[SC_RRL] = rrt;
            sc[SC_RZ] = tot[3 * k + 1];
            sc[SC_RR] = tot[3 * k + 2];
            S.upd[i] = 0;
            if (!A.init) atomicAdd(&S.status[2], 1);
            const int was = A.init ? 0 : S.act[i];
            const int now = (rrt <= sc[SC_THR] || sc[SC_RR0] < 1e-60) ? 0 : 1;
            S.act[i] = now;
            if (now != was) atomicAdd(&S.status[0], now - was);
        }
    }
};

// after k_resnorm
template <int NRHS>
struct FinRes {
    __device__ static bool skip(int, const DevScal &, const FinArgs &) { return false; }
    __device__ static int nq(int, const DevScal &) { return NRHS; }
    __device__ static void run(int s, const double *tot, DevScal &S, const FinArgs &) {
#pragma unroll
        for (int k = 0; k < NRHS; ++k) S.sc[(size_t)(s * NRHS + k) * SC_N + SC_RRT] = tot[k];
    }
};

// ---------------------------------------------------------------------------
// Lanczos steps on M = D^-1/2 A D^-1/2 of one multigrid level (FP64), for the
// largest eigenvalue of D^-1 A that sets the level's Jacobi weight.  Each
// kernel writes one partial sum per 256-row workgroup, summed on the host.

__device__ __forceinline__ void block_sum256(double v, double *out) {
    __shared__ double red[4];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// w = a fixed pseudo-random start vector in [-1, 1); partials of |w|^2
__global__ __launch_bounds__(256) void k_lz_init(int n, double *__restrict__ w, double *__restrict__ part) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    double t = 0.0;
    if (i < n) {
        unsigned h = (unsigned)i * 2654435761u;
        h ^= h >> 16;
        h *= 2246822519u;
        h ^= h >> 13;
        const double v = (double)h * (2.0 / 4294967296.0) - 1.0;
        w[i] = v;
        t = v * v;
    }
    block_sum256(t, part);
}

// v = scale w, u = D^-1/2 v
__global__ __launch_bounds__(256) void k_lz_prep(int n, const double *__restrict__ dinv,
                                                 const double *__restrict__ w, double scale,
                                                 double *__restrict__ v, double *__restrict__ u) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double vi = w[i] * scale;
    v[i] = vi;
    u[i] = sqrt(dinv[i]) * vi;
}

// y = D^-1/2 A u on the face matrix; partials of <v, y>
__global__ __launch_bounds__(256) void k_lz_sell(DevTopo T, const double *__restrict__ val,
                                                 const double *__restrict__ dinv, const double *__restrict__ u,
                                                 const double *__restrict__ v, double *__restrict__ y,
                                                 double *__restrict__ part) {
    const int f = blockIdx.x * 256 + threadIdx.x;   // T.nfp is a multiple of 256
    double acc[1];
    sell_row<double, double, double, 1>(T, val, u, f, acc);
    const double yi = sqrt(dinv[f]) * acc[0];
    y[f] = yi;
    block_sum256(v[f] * yi, part);
}

// y = D^-1/2 A u on a coarse (CSR) level; partials of <v, y>
__global__ __launch_bounds__(256) void k_lz_csr(DevAmgLevel L, const double *__restrict__ u,
                                                const double *__restrict__ v, double *__restrict__ y,
                                                double *__restrict__ part) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    double t = 0.0;
    if (i < L.n) {
        double acc = 0.0;
        for (int e = L.row_ptr[i]; e < L.row_ptr[i + 1]; ++e) acc += L.val64[e] * u[L.col[e]];
        const double yi = sqrt(L.dinv64[i]) * acc;
        y[i] = yi;
        t = v[i] * yi;
    }
    block_sum256(t, part);
}

// w = y - a v - b vp (vp null: first step); partials of |w|^2
__global__ __launch_bounds__(256) void k_lz_update(int n, const double *__restrict__ y, const double *__restrict__ v,
                                                   const double *__restrict__ vp, double a, double b,
                                                   double *__restrict__ w, double *__restrict__ part) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    double t = 0.0;
    if (i < n) {
        const double wi = y[i] - a * v[i] - (vp ? b * vp[i] : 0.0);
        w[i] = wi;
        t = wi * wi;
    }
    block_sum256(t, part);
}

// gterm[h] = omega[c] * (x_f - x_c) . g, one half-face per thread
// (MimeticIPEvaluator::gravityTerm, MimeticIPEvaluator.hpp:477-523).
__global__ __launch_bounds__(256) void k_gravity_term(DevTopo T, DevGeom G, double gx, double gy, double gz,
                                                      const double *__restrict__ omega, double *__restrict__ gterm) {
    const int h = blockIdx.x * 256 + threadIdx.x;
    if (h >= T.nhf) return;
    const int c = T.hf_cell[h];
    const double dx = G.fcen[3 * (size_t)h] - G.ccen[3 * (size_t)c];
    const double dy = G.fcen[3 * (size_t)h + 1] - G.ccen[3 * (size_t)c + 1];
    const double dz = G.fcen[3 * (size_t)h + 2] - G.ccen[3 * (size_t)c + 2];
    gterm[h] = omega[c] * (dx * gx + dy * gy + dz * gz);
}

}  // namespace

// ===========================================================================
// Launchers.

template <class Fn>
static hipError_t with_nrhs(int nrhs, Fn fn) {
    switch (nrhs) {
        case 1: fn(IC<1>{}); break;
        case 2: fn(IC<2>{}); break;
        case 3: fn(IC<3>{}); break;
        case 4: fn(IC<4>{}); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class Fn>
static hipError_t with_maxn(int maxn, Fn fn) {
    if (maxn <= 6) fn(IC<6>{});
    else if (maxn <= 8) fn(IC<8>{});
    else if (maxn <= 10) fn(IC<10>{});
    else if (maxn <= 16) fn(IC<16>{});
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

static inline unsigned cell_blocks(int nc) { return (unsigned)((nc + 63) / 64); }

// The element-by-element operator: per colour, the hexahedra with the 6-face
// kernel (few registers, full occupancy), then the other cells with the
// kernel for the largest face count.
template <int NR, int MODE, bool REL>
static void ebe_apply(hipStream_t st, int maxn, const DevEbe &E, const Gate &gt, const DevScal &S, const float *x,
                      float *out, const float *rin, const RelArgs &L, double *r64) {
    with_maxn(maxn, [&](auto mn) {
        constexpr int MN = decltype(mn)::value;
        for (int c = 0; c < E.ncolours; ++c) {
            const int b0 = E.colour_blk[c], b1 = E.colour_split[c], b2 = E.colour_blk[c + 1];
            if (MN == 6) {
                if (b2 > b0) k_ebe<6, NR, MODE, REL><<<b2 - b0, 256, 0, st>>>(E, b0, gt, S, x, out, rin, L, r64);
                continue;
            }
            if (b1 > b0) k_ebe<6, NR, MODE, REL><<<b1 - b0, 256, 0, st>>>(E, b0, gt, S, x, out, rin, L, r64);
            if (b2 > b1) k_ebe<MN, NR, MODE, REL><<<b2 - b1, 256, 0, st>>>(E, b1, gt, S, x, out, rin, L, r64);
        }
    });
}


// Workgroup size of the row-block kernels (each workgroup owns 1024 rows).
template <class Fn>
static hipError_t with_tb(int tb, Fn fn) {
    switch (tb) {
        case 512: fn(IC<512>{}); break;
        case 1024: fn(IC<1024>{}); break;
        default: fn(IC<256>{}); break;     // 256 and anything unset
    }
    return hipGetLastError();
}

// Lanes per row of a coarse-level row loop over `nnz` entries in `n` rows:
// up to about eight entries per lane keep one lane per row; longer rows (the
// levels behind a smoothed transfer) spread over a group of lanes, which also
// gives the small levels enough workgroups to cover the chip.
static inline int lanes_per_row(long long nnz, int n) {
    if (n <= 0) return 1;
    const long long avg = (nnz + n - 1) / n;
    return avg > 32 ? 16 : avg > 16 ? 8 : avg > 8 ? 4 : 1;
}
template <class Fn>
static void with_lanes(int g, Fn fn) {
    switch (g) {
        case 4: fn(IC<4>{}); break;
        case 8: fn(IC<8>{}); break;
        case 16: fn(IC<16>{}); break;
        default: fn(IC<1>{}); break;
    }
}
static inline unsigned row_blocks(int n, int g) { return (unsigned)(((long long)n * g + 255) / 256); }

// Entries per chunk of the streaming face-matrix rows: one chunk of 12 when
// every slice fits (hexahedral grids: 11 entries), else chunks of 16 (fewer
// registers than 16 for the common width, more loads in flight for wide rows).
template <class Fn>
static void with_chunk(int max_w, Fn fn) {
    if (max_w <= 12) fn(IC<12>{});
    else fn(IC<16>{});
}

hipError_t launch_gravity_term(hipStream_t st, const DevTopo &T, const DevGeom &G, const double *grav,
                               const double *omega, double *gterm) {
    if (T.nhf == 0) return hipSuccess;
    k_gravity_term<<<(unsigned)((T.nhf + 255) / 256), 256, 0, st>>>(T, G, grav[0], grav[1], grav[2], omega, gterm);
    return hipGetLastError();
}

// MimeticIPAnisoRelpermEvaluator::computeDynamicParams
// (MimeticIPAnisoRelpermEvaluator.hpp:299-330): per cell lambda K =
// lambda_t K (row-major 3x3 products) and the dynamic gravity vector
// (sum_i rho_i lambda_i) (K g).
__global__ void k_aniso_products(int nc, const double *__restrict__ perm, const double *__restrict__ lam,
                                 const double *__restrict__ rlam, double gx, double gy, double gz,
                                 double *__restrict__ lamK, double *__restrict__ dynKg) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const double *K = perm + 9 * (size_t)c, *L = lam + 9 * (size_t)c, *R = rlam + 9 * (size_t)c;
    double Kr[9], Lr[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        Kr[i] = K[i];
        Lr[i] = L[i];
    }
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
        for (int m = 0; m < 3; ++m)
            lamK[9 * (size_t)c + 3 * l + m] = Lr[3 * l] * Kr[m] + Lr[3 * l + 1] * Kr[3 + m] + Lr[3 * l + 2] * Kr[6 + m];
    const double kg0 = Kr[0] * gx + Kr[1] * gy + Kr[2] * gz;
    const double kg1 = Kr[3] * gx + Kr[4] * gy + Kr[5] * gz;
    const double kg2 = Kr[6] * gx + Kr[7] * gy + Kr[8] * gz;
#pragma unroll
    for (int l = 0; l < 3; ++l) dynKg[3 * (size_t)c + l] = R[3 * l] * kg0 + R[3 * l + 1] * kg1 + R[3 * l + 2] * kg2;
}

// gravityFlux (MimeticIPAnisoRelpermEvaluator.hpp:406-416): N dyn_Kg per
// half-face, N's rows being area times the unit normal.
__global__ void k_aniso_gflux(DevTopo T, DevGeom G, const double *__restrict__ dynKg, double *__restrict__ gflux) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= T.nhf) return;
    const int c = T.hf_cell[h];
    const double a = G.farea[h];
    gflux[h] = G.fnrm[3 * (size_t)h] * a * dynKg[3 * (size_t)c] + G.fnrm[3 * (size_t)h + 1] * a * dynKg[3 * (size_t)c + 1] +
               G.fnrm[3 * (size_t)h + 2] * a * dynKg[3 * (size_t)c + 2];
}

hipError_t launch_aniso_ip(hipStream_t st, const DevTopo &T, const DevGeom &G, int maxn, const double *perm,
                           const double *lam, const double *rlam, const double *grav, double *lamK, double *dynKg,
                           double *binv_pk, double *gflux) {
    if (T.nc == 0) return hipSuccess;
    k_aniso_products<<<(T.nc + 255) / 256, 256, 0, st>>>(T.nc, perm, lam, rlam, grav[0], grav[1], grav[2], lamK,
                                                          dynKg);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // B^-1 from lambda K; the static kernel's gravity flux is replaced below
    e = launch_static_ip(st, T, G, maxn, lamK, grav, binv_pk, gflux);
    if (e != hipSuccess) return e;
    k_aniso_gflux<<<(T.nhf + 255) / 256, 256, 0, st>>>(T, G, dynKg, gflux);
    return hipGetLastError();
}

hipError_t launch_static_ip(hipStream_t st, const DevTopo &T, const DevGeom &G, int maxn,
                            const double *perm, const double *grav, double *binv_pk,
                            double *gflux) {
    if (T.nc == 0) return hipSuccess;
    const double gx = grav[0], gy = grav[1], gz = grav[2];
    return with_maxn(maxn, [&](auto mn) {
        constexpr int MN = decltype(mn)::value;
        k_static_ip<MN><<<cell_blocks(T.nc), 64, 0, st>>>(T, G, perm, gx, gy, gz, binv_pk, gflux);
    });
}

hipError_t launch_assemble(hipStream_t st, const DevTopo &T, int maxn, int nrhs,
                           const double *binv_pk, const double *gflux, const double *totmob,
                           const double *rhomob, const double *src, const uint8_t *ftype,
                           const double *bcval, DevCellSys C, const DevEbe *E) {
    if (T.nc == 0) return hipSuccess;
    const DevEbe Ev = E ? *E : DevEbe{};
    hipError_t e = hipSuccess;
    hipError_t e2 = with_maxn(maxn, [&](auto mn) {
        constexpr int MN = decltype(mn)::value;
        e = with_nrhs(nrhs, [&](auto nr) {
            constexpr int NR = decltype(nr)::value;
            // every cell with MN faces: loops of fixed length (immediate offsets)
            if (T.uniform_ncf && T.max_ncf == MN)
                k_assemble<MN, NR, true><<<cell_blocks(T.nc), 64, 0, st>>>(T, binv_pk, gflux, totmob, rhomob,
                                                                           src, ftype, bcval, C, Ev);
            else if (MN >= 10)   // varying counts up to 10 or 16 faces: no per-entry registers
                k_assemble_stream<NR><<<cell_blocks(T.nc), 64, 0, st>>>(T, binv_pk, gflux, totmob, rhomob, src,
                                                                        ftype, bcval, C, Ev);
            else
                k_assemble<MN, NR, false><<<cell_blocks(T.nc), 64, 0, st>>>(T, binv_pk, gflux, totmob, rhomob,
                                                                            src, ftype, bcval, C, Ev);
        });
    });
    return e != hipSuccess ? e : e2;
}

hipError_t launch_fill(hipStream_t st, const DevTopo &T, int nrhs, const DevCellSys &C,
                       const uint8_t *ftype, const double *bcval, const int *inst_reg,
                       int dguess, double *val64, float *val32, double *b, double *dinv64,
                       float *dinv32, double *x0, double *r0, int mode) {
    // rows of up to 24 entries: 256 threads and 48 KB of LDS; up to 128: 64
    // threads and 64 KB; wider rows are summed in place in val64
    const size_t w = (size_t)std::max(T.max_w, 1);
    return with_nrhs(nrhs, [&](auto nr) {
        constexpr int NR = decltype(nr)::value;
        auto go = [&](auto fm) {
            constexpr int FM = decltype(fm)::value;
            if (FM == FILL_VEC)
                k_fill<NR, 256, false, FM><<<T.nfp / 256, 256, 0, st>>>(
                    T, C, ftype, bcval, inst_reg, dguess, val64, val32, b, dinv64, dinv32, x0, r0);
            else if (w <= 24)
                k_fill<NR, 256, true, FM><<<T.nfp / 256, 256, w * 256 * sizeof(double), st>>>(
                    T, C, ftype, bcval, inst_reg, dguess, val64, val32, b, dinv64, dinv32, x0, r0);
            else if (w <= 128)
                k_fill<NR, 64, true, FM><<<T.nfp / 64, 64, w * 64 * sizeof(double), st>>>(
                    T, C, ftype, bcval, inst_reg, dguess, val64, val32, b, dinv64, dinv32, x0, r0);
            else
                k_fill<NR, 256, false, FM><<<T.nfp / 256, 256, 0, st>>>(
                    T, C, ftype, bcval, inst_reg, dguess, val64, val32, b, dinv64, dinv32, x0, r0);
        };
        if (mode == FILL_VEC) go(IC<FILL_VEC>{});
        else if (mode == FILL_MAT) go(IC<FILL_MAT>{});
        else go(IC<FILL_ALL>{});
    });
}

template <typename MT, typename VT>
hipError_t launch_pcg_iteration(hipStream_t st, const DevTopo &T, int nrhs, const MT *val,
                                const VT *dinv, VT *x, VT *r, VT *p, VT *q, DevRed R,
                                DevScal S, const double *val64, const double *b, double *x64,
                                double tol2, const DevEbe *E, double *r64) {
    return with_nrhs(nrhs, [&](auto nr) {
        constexpr int NR = decltype(nr)::value;
        const FinArgs A{tol2, 0};
        constexpr bool REL = std::is_same<VT, float>::value;
        RelArgs L{};
        if constexpr (REL) L = RelArgs{val64, x64, b, dinv, r};
        bool ebe = false;
        if constexpr (REL) {
            if (E) {
                ebe_apply<NR, EBE_SPMV, true>(st, T.max_ncf, *E, Gate{S.act}, S, p, q, nullptr, L, r64);
                k_finish<4 * NR, FinPQ<NR, true>><<<T.ninst, 256, 0, st>>>(E->R, S, A);
                ebe = true;
            }
        }
        if (!ebe) {
            with_tb(T.spmv_tb, [&](auto tb) {
                constexpr int TBv = decltype(tb)::value;
                with_chunk(T.max_w, [&](auto ch) {
                    launch_spmv_kernel<MT, VT, NR, REL, TBv, decltype(ch)::value>(st, T.nblk, T, val, p, q, R, S,
                                                                                  L);
                });
            });
            k_finish<REL ? 4 * NR : NR, FinPQ<NR, REL>><<<T.ninst, 256, 0, st>>>(R, S, A);
        }
        k_update_xr<VT, NR><<<T.nblk, 256, 0, st>>>(T, dinv, x, r, p, q, R, S);
        k_finish<2 * NR, FinXR<NR>><<<T.ninst, 256, 0, st>>>(R, S, A);
        k_update_p<VT, NR><<<T.nblk, 256, 0, st>>>(T, dinv, r, p, x, REL ? x64 : nullptr, S);
    });
}

template hipError_t launch_pcg_iteration<double, double>(hipStream_t, const DevTopo &, int,
                                                         const double *, const double *,
                                                         double *, double *, double *,
                                                         double *, DevRed, DevScal,
                                                         const double *, const double *,
                                                         double *, double, const DevEbe *, double *);
template hipError_t launch_pcg_iteration<float, float>(hipStream_t, const DevTopo &, int,
                                                       const float *, const float *, float *,
                                                       float *, float *, float *, DevRed,
                                                       DevScal, const double *, const double *,
                                                       double *, double, const DevEbe *, double *);

template <typename MT, typename VT>
hipError_t launch_spmv(hipStream_t st, const DevTopo &T, int nrhs, const MT *val, const VT *p,
                       VT *q, DevRed R, DevScal S) {
    return with_nrhs(nrhs, [&](auto nr) {
        constexpr int NR = decltype(nr)::value;
        with_tb(T.spmv_tb, [&](auto tb) {
            constexpr int TBv = decltype(tb)::value;
            with_chunk(T.max_w, [&](auto ch) {
                launch_spmv_kernel<MT, VT, NR, false, TBv, decltype(ch)::value>(st, T.nblk, T, val, p, q, R, S,
                                                                              RelArgs{});
            });
        });
    });
}
template hipError_t launch_spmv<double, double>(hipStream_t, const DevTopo &, int, const double *,
                                                const double *, double *, DevRed, DevScal);
template hipError_t launch_spmv<float, float>(hipStream_t, const DevTopo &, int, const float *,
                                              const float *, float *, DevRed, DevScal);

hipError_t launch_pcg_start64(hipStream_t st, const DevTopo &T, int nrhs, const double *val,
                              const double *dinv, const double *b, const double *x, double *r,
                              double *p, double tol2, DevRed R, DevScal S, bool r_given) {
    return with_nrhs(nrhs, [&](auto nr) {
        constexpr int NR = decltype(nr)::value;
        if (r_given)
            k_start64<NR, true><<<T.nblk, 256, 0, st>>>(T, val, dinv, b, x, r, p, tol2, R, S);
        else
            k_start64<NR, false><<<T.nblk, 256, 0, st>>>(T, val, dinv, b, x, r, p, tol2, R, S);
        k_finish<2 * NR, FinStart<NR>><<<T.ninst, 256, 0, st>>>(R, S, FinArgs{tol2, 0});
    });
}

hipError_t launch_reliable(hipStream_t st, const DevTopo &T, int nrhs, const double *val64,
                           const double *b, const double *x64, const float *dinv32, float *r32,
                           float *p32, int init, double tol2, DevRed R, DevScal S, const double *r0) {
    return with_nrhs(nrhs, [&](auto nr) {
        constexpr int NR = decltype(nr)::value;
        if (r0)
            k_reliable<NR, true><<<T.nblk, 256, 0, st>>>(T, val64, b, x64, r0, dinv32, r32, p32, init, tol2, R, S);
        else
            k_reliable<NR, false><<<T.nblk, 256, 0, st>>>(T, val64, b, x64, r0, dinv32, r32, p32, init, tol2, R,
                                                          S);
        k_finish<3 * NR, FinRel<NR>><<<T.ninst, 256, 0, st>>>(R, S, FinArgs{tol2, init});
    });
}

hipError_t launch_clear_state(hipStream_t st, const DevScal &S, int nsys, float *x32, long long nx) {
    const long long n = std::max<long long>({4LL, (long long)nsys, x32 ? nx : 0LL});
    k_clear_state<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(S, nsys, x32, nx);
    return hipGetLastError();
}

hipError_t launch_merge_all(hipStream_t st, const DevTopo &T, int nrhs, double *x64, float *x32) {
    return with_nrhs(nrhs, [&](auto nr) {
        constexpr int NR = decltype(nr)::value;
        k_merge_all<NR><<<T.nblk, 256, 0, st>>>(T, x64, x32);
    });
}

hipError_t launch_resnorm(hipStream_t st, const DevTopo &T, int nrhs, const double *val64,
                          const double *b, const double *x64, DevRed R, DevScal S) {
    return with_nrhs(nrhs, [&](auto nr) {
        constexpr int NR = decltype(nr)::value;
        k_resnorm<NR><<<T.nblk, 256, 0, st>>>(T, val64, b, x64, R, S);
        k_finish<NR, FinRes<NR>><<<T.ninst, 256, 0, st>>>(R, S, FinArgs{0.0, 0});
    });
}


hipError_t launch_amg_setup(hipStream_t st, const DevTopo &T, const DevAmg &A, const double *val64,
                            bool mixed, const double *elem64, bool ap_only) {
    // A P of the level-1 aggregates first: level 1's Galerkin sums its slots.
    // elem64: from the FP64 element values (no face matrix filled)
    if (A.ap_slots > 0) {
        const bool el = elem64 != nullptr;
        const int *sp = el ? A.ap_esrc_ptr : A.ap_src_ptr, *sl = el ? A.ap_esrc : A.ap_src;
        with_lanes(lanes_per_row(el ? A.ap_ensrc : A.ap_nsrc, (int)A.ap_slots), [&](auto g) {
            constexpr int Gv = decltype(g)::value;
            k_galerkin<Gv><<<row_blocks((int)A.ap_slots, Gv), 256, 0, st>>>(
                (int)A.ap_slots, sp, sl, el ? elem64 : val64, A.ap64, mixed ? A.ap32 : nullptr);
        });
    }
    if (ap_only) return hipGetLastError();   // the coarse levels are kept
    for (int l = 1; l < A.nlev; ++l) {
        const DevAmgLevel &L = A.lev[l];
        const DevAmgLevel &F = A.lev[l - 1];
        if (L.nnz > 0) {
            with_lanes(lanes_per_row(L.nsrc, L.nnz), [&](auto g) {
                constexpr int Gv = decltype(g)::value;
                if (L.sa_in)
                    k_pair_sum<Gv><<<row_blocks(L.nnz, Gv), 256, 0, st>>>(
                        L.nnz, L.src_ptr, L.src, L.src2, F.p_val64, F.t_val64, L.val64, mixed ? L.val32 : nullptr);
                else
                    k_galerkin<Gv><<<row_blocks(L.nnz, Gv), 256, 0, st>>>(
                        L.nnz, L.src_ptr, L.src, (l == 1) ? A.ap64 : F.val64, L.val64, mixed ? L.val32 : nullptr);
            });
        }
        if (L.n > 0)
            k_coarse_diag<<<(L.n + 255) / 256, 256, 0, st>>>(L.n, L.diag, L.val64, L.dinv64,
                                                            mixed ? L.dinv32 : nullptr);
        if (L.sa_out && L.n > 0) {
            DevAmgLevel Lw = L;
            if (!mixed) {
                Lw.p_val32 = nullptr;
                Lw.pt_val32 = nullptr;
            }
            // a few entries per lane: the set-up is latency-bound
            with_lanes(L.p_nnz > 2LL * L.n ? 4 : 1, [&](auto g) {
                constexpr int Gv = decltype(g)::value;
                k_sa_p<Gv><<<row_blocks(L.n, Gv), 256, 0, st>>>(Lw, A.omega_p);
            });
            if (L.t_nnz > 0)
                with_lanes(lanes_per_row(L.t_npair, L.t_nnz), [&](auto g) {
                    constexpr int Gv = decltype(g)::value;
                    k_pair_sum<Gv><<<row_blocks(L.t_nnz, Gv), 256, 0, st>>>(
                        L.t_nnz, L.t_pair_ptr, L.t_pa, L.t_pb, L.val64, L.p_val64, L.t_val64,
                        mixed ? L.t_val32 : nullptr);
                });
        }
    }
    const DevAmgLevel &C = A.lev[A.nlev - 1];
    k_coarse_inverse<<<A.ninst, 1024, 0, st>>>(C, C.inst_ptr, A.coarse_off, A.cinv64,
                                              mixed ? A.cinv32 : nullptr);
    return hipGetLastError();
}

hipError_t launch_amg_combine(hipStream_t st, const DevAmg &A, bool mixed) {
    for (int l = 1; l + 1 < A.nlev; ++l) {
        DevAmgLevel F = A.lev[l];
        if (!F.sa_out || F.n == 0) continue;
        if (!mixed) F.m_val32 = nullptr;
        with_lanes(lanes_per_row(F.t_nnz, F.n), [&](auto g) {
            constexpr int Gv = decltype(g)::value;
            k_sa_m<Gv><<<row_blocks(F.n, Gv), 256, 0, st>>>(F, A.omega_l[l]);
        });
    }
    return hipGetLastError();
}

static inline unsigned lz_blocks(int n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_lanczos_init(hipStream_t st, int n, double *w, double *part) {
    k_lz_init<<<lz_blocks(n), 256, 0, st>>>(n, w, part);
    return hipGetLastError();
}

hipError_t launch_lanczos_step(hipStream_t st, const DevTopo &T, const DevAmg &A, int level, const double *val0,
                               const double *dinv0, const double *w, double scale, double *v, double *u,
                               double *y, double *part) {
    const int n = level == 0 ? T.nfp : A.lev[level].n;
    const double *dinv = level == 0 ? dinv0 : A.lev[level].dinv64;
    k_lz_prep<<<lz_blocks(n), 256, 0, st>>>(n, dinv, w, scale, v, u);
    if (level == 0)
        k_lz_sell<<<lz_blocks(n), 256, 0, st>>>(T, val0, dinv, u, v, y, part);
    else
        k_lz_csr<<<lz_blocks(n), 256, 0, st>>>(A.lev[level], u, v, y, part);
    return hipGetLastError();
}

hipError_t launch_lanczos_update(hipStream_t st, int n, const double *y, const double *v, const double *vp,
                                 double a, double b, double *w, double *part) {
    k_lz_update<<<lz_blocks(n), 256, 0, st>>>(n, y, v, vp, a, b, w, part);
    return hipGetLastError();
}

// The fused post-smoothing sweep (k_vc0_post) in the tuned workgroup shape.
template <typename VT, int NR>
static void post0(hipStream_t st, const Gate &gt, const DevTopo &T, const DevAmg &A, const VT *dinv0, const VT *r,
                  const VT *x0,
                  const VT *res0, const VT *xc, VT *z, VT w, VT al, DevRed R) {
    const VT *ap;
    if constexpr (std::is_same<VT, float>::value) ap = A.ap32; else ap = A.ap64;
    with_tb(T.sweep_tb, [&](auto tb) {
        constexpr int TBv = decltype(tb)::value;
        k_vc0_post<VT, NR, TBv><<<T.nblk, TBv, 0, st>>>(gt, T, A.ap_slice_ptr, A.ap_col, ap, A.lev[1].agg, dinv0, r,
                                                        x0, res0, xc, z, w, al, R);
    });
}

template <typename VT>
hipError_t launch_post0(hipStream_t st, const DevTopo &T, const DevAmg &A, int nrhs, const VT *dinv0,
                        const VT *r, VT *z, DevRed R) {
    return with_nrhs(nrhs, [&](auto nr) {
        constexpr int NR = decltype(nr)::value;
        post0<VT, NR>(st, Gate{}, T, A, dinv0, r, (const VT *)A.x0, (const VT *)A.res0, (const VT *)A.lev[1].x, z,
                      (VT)A.omega_l[0], (VT)A.alpha, R);
    });
}
template hipError_t launch_post0<double>(hipStream_t, const DevTopo &, const DevAmg &, int, const double *,
                                         const double *, double *, DevRed);
template hipError_t launch_post0<float>(hipStream_t, const DevTopo &, const DevAmg &, int, const float *,
                                        const float *, float *, DevRed);

// Level-0 Jacobi sweep xo = x + w D^-1 (b - A x) in the tuned workgroup shape;
// LAST also forms the <b, xo> partials.
template <typename VT, int NR, bool LAST>
static void sweep0(hipStream_t st, const DevTopo &T, const VT *val0, const VT *dinv0, const VT *b, const VT *x,
                   VT *xo, VT w, DevRed R) {
    with_tb(T.sweep_tb, [&](auto tb) {
        constexpr int TBv = decltype(tb)::value;
        k_vc0_up<VT, NR, LAST, TBv><<<T.nblk, TBv, 0, st>>>(T, val0, dinv0, nullptr, b, x, nullptr, xo, w,
                                                            VT(0), R);
    });
}

// gt: inside the PCG loop the V-cycle skips the level-0 blocks and the tail
// workgroups of finished instances.  The other coarse-level kernels are too
// short for a gate to pay for its extra memory round trip: they run on, on
// values nothing reads.
template <typename VT, int NR>
static void vcycle_impl(hipStream_t st, const Gate &gt, const DevTopo &T, const DevAmg &A, const VT *val0,
                        const VT *dinv0, const VT *r, VT *z, DevRed R, bool x0_ready,
                        const DevEbe *E = nullptr) {
    const VT w = (VT)A.omega_l[0], al = (VT)A.alpha;   // w: level 0
    auto wl = [&](int l) { return (VT)A.omega_l[l]; };
    const int nu = A.nu < 1 ? 1 : A.nu;
    const int rows_blocks = T.nfp / 256;
    VT *x0 = (VT *)A.x0, *xt0 = (VT *)A.xt0, *res0 = (VT *)A.res0;
    const int L = A.nlev;
    // ---- level 0 down: x0 = w D^-1 r (+ nu-1 sweeps), res0 = r - A x0
    if (!x0_ready) k_scale0<VT, NR><<<rows_blocks, 256, 0, st>>>(gt, T.block_seg, dinv0, r, x0, w);
    for (int it = 1; it < nu; ++it) {
        sweep0<VT, NR, false>(st, T, val0, dinv0, r, x0, xt0, w, R);
        std::swap(x0, xt0);
    }
    bool ebe = false;
    if constexpr (std::is_same<VT, float>::value) {
        if (E) {
            ebe_apply<NR, EBE_RESID, false>(st, T.max_ncf, *E, gt, DevScal{}, x0, res0, r, RelArgs{}, nullptr);
            ebe = true;
        }
    }
    if (!ebe) {
        with_chunk(T.max_w, [&](auto ch) {
            with_tb(T.spmv_tb, [&](auto tb) {
                constexpr int TBv = decltype(tb)::value;
                k_vc0_res<VT, NR, decltype(ch)::value, TBv><<<T.nblk, TBv, 0, st>>>(gt, T, val0, r, x0, res0);
            });
        });
    }
    {
        const DevAmgLevel &C = A.lev[1];
        with_lanes(lanes_per_row(T.nfp, C.n), [&](auto g) {
            constexpr int Gv = decltype(g)::value;
            k_restrict<VT, NR, Gv><<<row_blocks(C.n, Gv), 256, 0, st>>>(C.n, C.child_ptr, C.child, res0,
                                                                       (VT *)C.b);
        });
    }
    VT *xcur[kMaxLevels], *xalt[kMaxLevels];
    for (int l = 1; l < L; ++l) {
        xcur[l] = (VT *)A.lev[l].x;
        xalt[l] = (VT *)A.lev[l].xt;
    }
    if (nu == 1) {
        // first level handled by the workgroup-per-instance tail kernel
        int lt = L - 1;
        while (lt > 1 && A.max_rows[lt - 1] <= T.tail_rows) --lt;
        const VT als = (VT)A.alpha_sa;
        for (int l = 1; l < lt; ++l) {
            const DevAmgLevel &F = A.lev[l];
            const DevAmgLevel &C = A.lev[l + 1];
            const int gd = lanes_per_row(F.nnz, F.n);
            if (!F.sa_out && A.covers[l + 1] && A.max_children[l + 1] <= 8) {
                // down-sweep and restriction in one launch (at most 8 lanes per
                // child row, so a coarse row and its children fill one wavefront)
                with_lanes(std::min(gd, 8), [&](auto g) {
                    constexpr int Gv = decltype(g)::value;
                    if constexpr (Gv <= 8)
                        k_vcc_down_restrict<VT, NR, Gv><<<row_blocks(C.n, 8 * Gv), 256, 0, st>>>(
                            F, C, (const VT *)F.b, (VT *)F.x, (VT *)F.res, (VT *)C.b, wl(l));
                });
                continue;
            }
            with_lanes(gd, [&](auto g) {
                constexpr int Gv = decltype(g)::value;
                k_vcc_down<VT, NR, Gv><<<row_blocks(F.n, Gv), 256, 0, st>>>(F, (const VT *)F.b, (VT *)F.x,
                                                                          (VT *)F.res, wl(l));
            });
            if (F.sa_out) {
                with_lanes(lanes_per_row(F.p_nnz, C.n), [&](auto g) {
                    constexpr int Gv = decltype(g)::value;
                    k_restrict_sa<VT, NR, Gv><<<row_blocks(C.n, Gv), 256, 0, st>>>(F, C.n, (VT *)C.b);
                });
            } else {
                with_lanes(lanes_per_row(F.n, C.n), [&](auto g) {
                    constexpr int Gv = decltype(g)::value;
                    k_restrict<VT, NR, Gv><<<row_blocks(C.n, Gv), 256, 0, st>>>(C.n, C.child_ptr, C.child,
                                                                               (const VT *)F.res, (VT *)C.b);
                });
            }
        }
        // Tentative transfers l-1 <- l whose coarse result the producing
        // kernel prolongs in place into level l-1's iterate; the up-sweep of
        // level l-1 then gathers one value per entry.  The producing kernel
        // can: the tail, the smoothed-transfer up-sweep, and the up-sweep of a
        // level that was itself prolonged in place.
        bool fuse[kMaxLevels] = {};
        for (int l = lt; l >= 2; --l) {
            const bool can = l == lt || A.lev[l].sa_out || fuse[l + 1];
            fuse[l] = can && !A.lev[l - 1].sa_out && A.covers[l];
        }
        k_vcycle_tail<VT, NR><<<A.ninst, 1024, 0, st>>>(gt, A.dev, lt, al, als,
                                                        fuse[lt] ? (VT *)A.lev[lt - 1].x : nullptr);
        VT *xres[kMaxLevels];   // each level's result after its up-sweep
        xres[lt] = (VT *)A.lev[lt].x;
        for (int l = lt - 1; l >= 1; --l) {
            const DevAmgLevel &F = A.lev[l];
            const VT *xc = xres[l + 1];
            if (F.sa_out) {
                VT *xf = fuse[l] ? (VT *)A.lev[l - 1].x : nullptr;
                with_lanes(lanes_per_row(F.t_nnz, F.n), [&](auto g) {
                    constexpr int Gv = decltype(g)::value;
                    if (xf)
                        k_vcc_up_sa<VT, NR, Gv, true><<<row_blocks(F.n, Gv), 256, 0, st>>>(F, xc, wl(l), als, xf,
                                                                                          al);
                    else
                        k_vcc_up_sa<VT, NR, Gv, false><<<row_blocks(F.n, Gv), 256, 0, st>>>(F, xc, wl(l), als,
                                                                                           nullptr, VT(0));
                });
                xres[l] = (VT *)F.x;
            } else if (fuse[l + 1]) {
                VT *xf = fuse[l] ? (VT *)A.lev[l - 1].x : nullptr;
                with_lanes(lanes_per_row(F.nnz, F.n), [&](auto g) {
                    constexpr int Gv = decltype(g)::value;
                    if (xf)
                        k_vcc_up_plain<VT, NR, Gv, true><<<row_blocks(F.n, Gv), 256, 0, st>>>(
                            F, (const VT *)F.x, (VT *)F.xt, xf, wl(l), al);
                    else
                        k_vcc_up_plain<VT, NR, Gv, false><<<row_blocks(F.n, Gv), 256, 0, st>>>(
                            F, (const VT *)F.x, (VT *)F.xt, nullptr, wl(l), al);
                });
                xres[l] = (VT *)F.xt;
            } else {
                with_lanes(lanes_per_row(F.nnz, F.n), [&](auto g) {
                    constexpr int Gv = decltype(g)::value;
                    k_vcc_up1<VT, NR, Gv><<<row_blocks(F.n, Gv), 256, 0, st>>>(F, A.lev[l + 1].agg, xc, (VT *)F.x,
                                                                             wl(l), al);
                });
                xres[l] = (VT *)F.x;
            }
        }
        xcur[1] = xres[1];
    } else {
        // ---- coarse levels down
        for (int l = 1; l < L - 1; ++l) {
            const DevAmgLevel &F = A.lev[l];
            const DevAmgLevel &C = A.lev[l + 1];
            const unsigned nb = (F.n + 255) / 256;
            k_vcc_down<VT, NR><<<nb, 256, 0, st>>>(F, (const VT *)F.b, xcur[l], (VT *)F.res, wl(l));
            for (int it = 1; it < nu; ++it) {
                k_vcc_up<VT, NR><<<nb, 256, 0, st>>>(F, nullptr, (const VT *)F.b, xcur[l], nullptr, xalt[l], wl(l),
                                                     VT(0));
                std::swap(xcur[l], xalt[l]);
            }
            if (nu > 1) k_vcc_res<VT, NR><<<nb, 256, 0, st>>>(F, (const VT *)F.b, xcur[l], (VT *)F.res);
            if (F.sa_out)
                k_restrict_sa<VT, NR><<<(C.n + 255) / 256, 256, 0, st>>>(F, C.n, (VT *)C.b);
            else
                k_restrict<VT, NR><<<(C.n + 255) / 256, 256, 0, st>>>(C.n, C.child_ptr, C.child,
                                                                      (const VT *)F.res, (VT *)C.b);
        }
        {
            const DevAmgLevel &C = A.lev[L - 1];
            const VT *inv;
            if constexpr (std::is_same<VT, float>::value) inv = A.cinv32; else inv = A.cinv64;
            k_coarse_solve<VT, NR><<<A.ninst, 256, 0, st>>>(C.inst_ptr, A.coarse_off, inv, (const VT *)C.b,
                                                            xcur[L - 1]);
        }
        // ---- coarse levels up
        for (int l = L - 2; l >= 1; --l) {
            const DevAmgLevel &F = A.lev[l];
            const unsigned nb = (F.n + 255) / 256;
            if (nu == 1) {
                k_vcc_up1<VT, NR><<<nb, 256, 0, st>>>(F, A.lev[l + 1].agg, xcur[l + 1], xalt[l], wl(l), al);
                std::swap(xcur[l], xalt[l]);
                continue;
            }
            if (F.sa_out) {
                k_prolong_sa<VT, NR><<<nb, 256, 0, st>>>(F, xcur[l], xcur[l + 1], (VT)A.alpha_sa, xalt[l]);
                std::swap(xcur[l], xalt[l]);
                k_vcc_up<VT, NR><<<nb, 256, 0, st>>>(F, nullptr, (const VT *)F.b, xcur[l], nullptr, xalt[l], wl(l),
                                                     VT(0));
            } else {
                k_vcc_up<VT, NR><<<nb, 256, 0, st>>>(F, A.lev[l + 1].agg, (const VT *)F.b, xcur[l],
                                                     xcur[l + 1], xalt[l], wl(l), al);
            }
            std::swap(xcur[l], xalt[l]);
            for (int it = 1; it < nu; ++it) {
                k_vcc_up<VT, NR><<<nb, 256, 0, st>>>(F, nullptr, (const VT *)F.b, xcur[l], nullptr, xalt[l], wl(l),
                                                     VT(0));
                std::swap(xcur[l], xalt[l]);
            }
        }
    }
    // ---- level 0 up: x~ = x0 + alpha P x1, then nu sweeps; the last writes z and <r, z>
    if (nu == 1) {
        post0<VT, NR>(st, gt, T, A, dinv0, r, x0, res0, xcur[1], z, w, al, R);
        return;
    }
    k_prolong0<VT, NR><<<rows_blocks, 256, 0, st>>>(A.lev[1].agg, x0, xcur[1], xt0, al);
    std::swap(x0, xt0);
    for (int it = 1; it < nu; ++it) {
        sweep0<VT, NR, false>(st, T, val0, dinv0, r, x0, xt0, w, R);
        std::swap(x0, xt0);
    }
    sweep0<VT, NR, true>(st, T, val0, dinv0, r, x0, z, w, R);
}

template <typename VT>
hipError_t launch_sweep0(hipStream_t st, const DevTopo &T, int nrhs, const VT *val0, const VT *dinv0, const VT *b,
                         const VT *x, VT *xo, double w, DevRed R) {
    return with_nrhs(nrhs, [&](auto nr) {
        constexpr int NR = decltype(nr)::value;
        sweep0<VT, NR, true>(st, T, val0, dinv0, b, x, xo, (VT)w, R);
    });
}
template hipError_t launch_sweep0<double>(hipStream_t, const DevTopo &, int, const double *, const double *,
                                          const double *, const double *, double *, double, DevRed);
template hipError_t launch_sweep0<float>(hipStream_t, const DevTopo &, int, const float *, const float *,
                                         const float *, const float *, float *, double, DevRed);

template <typename VT>
hipError_t launch_vcycle(hipStream_t st, const DevTopo &T, const DevAmg &A, int nrhs, const VT *val0,
                         const VT *dinv0, const VT *r, VT *z, DevRed R, const DevEbe *E) {
    return with_nrhs(nrhs, [&](auto nr) {
        constexpr int NR = decltype(nr)::value;
        vcycle_impl<VT, NR>(st, Gate{}, T, A, val0, dinv0, r, z, R, false, E);
    });
}
template hipError_t launch_vcycle<double>(hipStream_t, const DevTopo &, const DevAmg &, int, const double *,
                                          const double *, const double *, double *, DevRed, const DevEbe *);
template hipError_t launch_vcycle<float>(hipStream_t, const DevTopo &, const DevAmg &, int, const float *,
                                         const float *, const float *, float *, DevRed, const DevEbe *);

hipError_t launch_ebe_spmv(hipStream_t st, int maxn, int nrhs, const DevEbe &E, const float *p, float *q,
                           DevScal S) {
    return with_nrhs(nrhs, [&](auto nr) {
        constexpr int NR = decltype(nr)::value;
        ebe_apply<NR, EBE_SPMV, false>(st, maxn, E, Gate{S.act}, S, p, q, nullptr, RelArgs{}, nullptr);
    });
}

template <typename MT, typename VT>
hipError_t launch_pcg_iteration_amg(hipStream_t st, const DevTopo &T, const DevAmg &A, int nrhs,
                                    const MT *val, const VT *dinv, VT *x, VT *r, VT *p, VT *q, VT *z,
                                    DevRed R, DevScal S, const double *val64, const double *b,
                                    double *x64, double tol2, const DevEbe *E, double *r64) {
    return with_nrhs(nrhs, [&](auto nr) {
        constexpr int NR = decltype(nr)::value;
        const FinArgs Af{tol2, 0};
        constexpr bool REL = std::is_same<VT, float>::value;
        RelArgs L{};
        if constexpr (REL) L = RelArgs{val64, x64, b, dinv, r};
        bool ebe = false;
        if constexpr (REL) {
            if (E) {
                ebe_apply<NR, EBE_SPMV, true>(st, T.max_ncf, *E, Gate{S.act}, S, p, q, nullptr, L, r64);
                k_finish<4 * NR, FinPQ<NR, true>><<<T.ninst, 256, 0, st>>>(E->R, S, Af);
                ebe = true;
            }
        }
        if (!ebe) {
            with_tb(T.spmv_tb, [&](auto tb) {
                constexpr int TBv = decltype(tb)::value;
                with_chunk(T.max_w, [&](auto ch) {
                    launch_spmv_kernel<MT, VT, NR, REL, TBv, decltype(ch)::value>(st, T.nblk, T, val, p, q, R, S, L);
                });
            });
            k_finish<REL ? 4 * NR : NR, FinPQ<NR, REL>><<<T.ninst, 256, 0, st>>>(R, S, Af);
        }
        k_update_xr<VT, NR><<<T.nblk, 256, 0, st>>>(T, dinv, x, r, p, q, R, S, (VT *)A.x0, (VT)A.omega_l[0]);
        // the V-cycle's <r, z> partials after update_xr's: one finisher for both
        DevRed Rz = R;
        Rz.partials += (size_t)2 * NR * R.nblk;
        vcycle_impl<VT, NR>(st, Gate{S.act}, T, A, val, dinv, r, z, Rz, true, E);
        k_finish<3 * NR, FinXRZ<NR>><<<T.ninst, 256, 0, st>>>(R, S, Af);
        k_update_pz<VT, NR><<<T.nblk, 256, 0, st>>>(T, z, p, x, REL ? x64 : nullptr, S);
    });
}
template hipError_t launch_pcg_iteration_amg<double, double>(hipStream_t, const DevTopo &, const DevAmg &,
                                                             int, const double *, const double *, double *,
                                                             double *, double *, double *, double *, DevRed,
                                                             DevScal, const double *, const double *,
                                                             double *, double, const DevEbe *, double *);
template hipError_t launch_pcg_iteration_amg<float, float>(hipStream_t, const DevTopo &, const DevAmg &, int,
                                                           const float *, const float *, float *, float *,
                                                           float *, float *, float *, DevRed, DevScal,
                                                           const double *, const double *, double *, double,
                                                           const DevEbe *, double *);

template <typename VT>
hipError_t launch_amg_start(hipStream_t st, const DevTopo &T, const DevAmg &A, int nrhs, const VT *val0,
                            const VT *dinv0, const VT *r, VT *z, VT *p, DevRed R, DevScal S,
                            double *x64, VT *x, const DevEbe *E) {
    return with_nrhs(nrhs, [&](auto nr) {
        constexpr int NR = decltype(nr)::value;
        vcycle_impl<VT, NR>(st, Gate{}, T, A, val0, dinv0, r, z, R, false, E);
        k_finish<NR, FinRZ<NR>><<<T.ninst, 256, 0, st>>>(R, S, FinArgs{0.0, 1});
        k_update_pz<VT, NR><<<T.nblk, 256, 0, st>>>(T, z, p, x, nullptr, S);
    });
}
template hipError_t launch_amg_start<double>(hipStream_t, const DevTopo &, const DevAmg &, int, const double *,
                                             const double *, const double *, double *, double *, DevRed,
                                             DevScal, double *, double *, const DevEbe *);
template hipError_t launch_amg_start<float>(hipStream_t, const DevTopo &, const DevAmg &, int, const float *,
                                            const float *, const float *, float *, float *, DevRed, DevScal,
                                            double *, float *, const DevEbe *);

hipError_t launch_recover(hipStream_t st, const DevTopo &T, int maxn, int nrhs,
                          const double *binv_pk, const double *gflux, const double *totmob,
                          const double *rhomob, const DevCellSys &C, const double *x64,
                          double *press, double *outflux) {
    if (T.nc == 0) return hipSuccess;
    hipError_t e = hipSuccess;
    hipError_t e2 = with_maxn(maxn, [&](auto mn) {
        constexpr int MN = decltype(mn)::value;
        e = with_nrhs(nrhs, [&](auto nr) {
            constexpr int NR = decltype(nr)::value;
            if (T.uniform_ncf && T.max_ncf == MN)
                k_recover<MN, NR, true><<<cell_blocks(T.nc), 64, 0, st>>>(T, binv_pk, gflux, totmob, rhomob, C,
                                                                          x64, press, outflux);
            else
                k_recover<MN, NR, false><<<cell_blocks(T.nc), 64, 0, st>>>(T, binv_pk, gflux, totmob, rhomob, C,
                                                                           x64, press, outflux);
        });
    });
    return e != hipSuccess ? e : e2;
}

hipError_t launch_post_process(hipStream_t st, const DevTopo &T, int nrhs, const double *vin,
                               double *vout, unsigned long long *maxbits) {
    const long long tot = (long long)nrhs * T.nhf;
    if (tot == 0) return hipSuccess;
    k_post<<<(unsigned)((tot + 255) / 256), 256, 0, st>>>(T, nrhs, vin, vout, maxbits);
    return hipGetLastError();
}

hipError_t launch_group_sum(hipStream_t st, int nrhs, int ngroups, int nhf, const int *gptr, const int *ghf,
                            const double *v, double *out) {
    if (ngroups == 0) return hipSuccess;
    k_group_sum<<<dim3((unsigned)ngroups, (unsigned)nrhs), 64, 0, st>>>(ngroups, nhf, gptr, ghf, v, out);
    return hipGetLastError();
}

hipError_t launch_adjoint_rhs(hipStream_t st, const DevTopo &T, int maxn, int nrhs, const double *binv_pk,
                              const double *totmob, const DevCellSys &C, const uint8_t *ftype,
                              const double *w, double *y, double *b, double *x, double *r) {
    hipError_t e = hipSuccess;
    hipError_t e2 = with_maxn(maxn, [&](auto mn) {
        constexpr int MN = decltype(mn)::value;
        e = with_nrhs(nrhs, [&](auto nr) {
            constexpr int NR = decltype(nr)::value;
            if (T.nc > 0) k_adj_cell<MN, NR><<<cell_blocks(T.nc), 64, 0, st>>>(T, binv_pk, totmob, C, w, y);
            k_adj_rows<NR><<<(unsigned)((T.nfp + 255) / 256), 256, 0, st>>>(T, ftype, y, b, x, r);
        });
    });
    return e != hipSuccess ? e : e2;
}

hipError_t launch_adjoint_grad(hipStream_t st, const DevTopo &T, const DevGeom &G, int maxn, int nrhs,
                               const double *binv_pk, const double *gflux, const double *totmob,
                               const double *rhomob, const double *grav, const DevCellSys &C,
                               const double *press, const double *pi, const double *mu,
                               const uint8_t *ftype, const double *w, double *gK, double *gtm,
                               double *grm, double *gsrc) {
    if (T.nc == 0) return hipSuccess;
    const double gx = grav[0], gy = grav[1], gz = grav[2];
    hipError_t e = hipSuccess;
    hipError_t e2 = with_maxn(maxn, [&](auto mn) {
        constexpr int MN = decltype(mn)::value;
        e = with_nrhs(nrhs, [&](auto nr) {
            constexpr int NR = decltype(nr)::value;
            k_adj_grad<MN, NR><<<cell_blocks(T.nc), 64, 0, st>>>(T, G, binv_pk, gflux, totmob, rhomob, gx,
                                                                 gy, gz, C, press, pi, mu, ftype, w, gK,
                                                                 gtm, grm, gsrc);
        });
    });
    return e != hipSuccess ? e : e2;
}

}  // namespace mhs

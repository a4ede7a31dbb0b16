// Device data views and kernel launchers of the mimetic hybrid solver.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace mhs {

// Geometry (per cell / per half-face), used by the static inner products only.
struct DevGeom {
    const double *cvol, *ccen, *farea, *fcen, *fnrm;
};

// Topology views (see mhs_structure.h for the meaning of each array).
struct DevTopo {
    int nc, nhf, nfp, ninst, nslices, nblk;
    int max_w;                      // widest SELL slice (stored entries per row)
    int max_ncf;                    // most faces of a cell (stride of the slot map)
    int uniform_ncf;                // 1: every cell has max_ncf faces (per-cell loops of fixed length)
    const int *facepos, *hf_dof, *hf_cell, *hf_col, *hf_ppart, *hf_twin;
    const uint8_t *hf_redirect;
    const int *row_hf, *dof_hf, *ptn;
    const int *row_cell;
    const uint8_t *row_loc;
    const uint8_t *dslot, *pslot;
    const long long *slice_ptr;
    const int *col;
    const uint8_t *slotmap;
    const long long *sq_pos;
    const int *block_seg, *seg_blk, *inst_rowbase;
    const int *inst_regrow;         // row of each instance's dof 0 (regularised when no Dirichlet face)
    // launch shapes of the row-block kernels, chosen by the autotuner:
    // workgroup size (256, 512 or 1024 threads per 1024 rows) and whether
    // workgroups are remapped so each XCD gets a contiguous row range
    int spmv_tb = 256, spmv_xcd = 1, sweep_tb = 256, sweep_xcd = 1;
    // multigrid levels whose largest instance has at most tail_rows rows run
    // in one workgroup-per-instance kernel (0: only the coarsest solve)
    int tail_rows = 0;
};

// Per-solve cell data produced by the assembly kernel.
struct DevCellSys {
    double *schur;     // packed local Schur complements (FP64), component-major: entry q of cell c at q*nc + c
    double *rr;        // row right-hand-side contributions, component-major: (rhs k, local face r, cell c) at (k*max_ncf + r)*nc + c
    double *F;         // [nhf]
    double *L;         // [nc]
    double *g;         // [nrhs][nc]
};

// Segmented-reduction workspace shared by all PCG kernels.
struct DevRed {
    double *partials;  // [16][nblk]
    int nblk;
    int *counters;     // [ninst]
    const int *block_seg, *seg_blk;
};

// Per-system scalars; a system is (instance s, right-hand side k), index s*nrhs+k.
enum { SC_RZ = 0, SC_PQ, SC_RR, SC_RR0, SC_THR, SC_ALPHA, SC_BETA, SC_RRT, SC_RRL, SC_N };
struct DevScal {
    double *sc;        // [nsys * SC_N]
    int *act;          // [nsys] 1 while iterating
    int *iters;        // [nsys]
    int *upd;          // [nsys] 1: FP64 residual replacement due, 2: also restart p
    int *status;       // [4]: 0 nactive, 1 breakdown count
    int mixed;         // FP32 iteration with FP64 residual replacement
    int amg;           // multigrid preconditioner (z stored, beta from <r, z>)
    double rf2;        // replacement trigger: |r|^2 <= rf2 * |r at last replacement|^2
};

// Element-by-element level-0 operator (see mhs_ebe.h): per segment of cells
// of one colour, instance and face count, the packed local matrices S'_c
// and the face rows with first/last-writer flags, both in 64-cell tiles.
constexpr int kEbeColours = 8;

// Offset of component q (of ncomp per cell) of the cell in slot `local` of an
// element-by-element segment.  Cells are stored in tiles of 64 (one
// wavefront), component-major inside a tile: a wavefront's load of one
// component is 64 consecutive elements, and all components of its cells lie
// in one contiguous ncomp * 64-element range, so the loads of a wavefront stay
// within a page instead of striding across the whole segment.
// Packed upper-triangle index of entry (i, j), i <= j, of an n x n local matrix.
__host__ __device__ constexpr int pidx(int i, int j, int n) {
    return i * n - (i * (i - 1)) / 2 + (j - i);
}
__host__ __device__ constexpr long long ebe_slot(int local, int q, int ncomp) {
    return (long long)(local >> 6) * ((long long)ncomp << 6) + ((long long)q << 6) + (local & 63);
}
struct DevEbe {
    int ncolours = 0, nblk = 0;
    int colour_blk[kEbeColours + 1] = {};
    int colour_split[kEbeColours] = {};     // first block of each colour that is not hexahedra
    // per block: (face count, instance, cells in the block, partial slot) and
    // (value base, face-row base) of its first tile
    const int4 *blk_info = nullptr;
    const longlong2 *blk_base = nullptr;
    const long long *seg_val = nullptr;     // per segment (assembly writes the values)
    const int *cell_seg = nullptr, *cell_local = nullptr;
    const uint32_t *dof = nullptr;
    double *val64 = nullptr;   // S'_c in FP64 (residual replacement)
    float *val32 = nullptr;    // the same rounded to FP32 (PCG products)
    DevRed R{};                // partial sums: one slot per block, instance-contiguous
};

// Multigrid hierarchy on the device (level 0 is the SELL face matrix).
constexpr int kMaxLevels = 16;
struct DevAmgLevel {
    int n, nnz;
    long long nsrc;                 // entries of the src (and src2) lists
    const int *inst_ptr;            // [ninst+1] first row of each instance
    const int *row_ptr, *col, *diag, *src_ptr, *src;
    const int *agg;                 // rows of the finer level -> rows of this level
    const int *child_ptr, *child;   // rows of this level -> rows of the finer level
    double *val64, *dinv64;
    float *val32, *dinv32;
    void *b, *x, *res, *xt;         // work vectors (PCG precision, n * nrhs)
    // smoothed aggregation (see AmgLevel): values of this level from pair
    // sums (sa_in, src = P positions, src2 = T positions of the finer level),
    // and the smoothed transfer to the next level (sa_out)
    int sa_in, sa_out, p_nnz, t_nnz;
    long long t_npair;              // entries of the T pair lists
    const int *src2;
    const int *p_ptr, *p_col, *p_own, *p_src_ptr, *p_src;
    const int *pt_ptr, *pt_col, *pt_map;
    const int *p_tpos;              // per P entry: its position in the P' rows
    double *pt_val64;               // P values in P' row order (refilled with P)
    float *pt_val32;
    const int *t_ptr, *t_col, *t_pair_ptr, *t_pa, *t_pb;
    double *p_val64, *t_val64;
    float *p_val32, *t_val32;
    // M = P - w D^-1 T on T's pattern: the up-sweep's prolongation and
    // post-smoothing in one product (refilled once the weight w is known)
    const int *t_p;                 // per T entry: the P entry at its place, or -1
    double *m_val64;
    float *m_val32;
};
struct DevAmg {
    int nlev;                       // levels including level 0
    int ninst;
    DevAmgLevel lev[kMaxLevels];
    void *x0, *res0, *xt0;          // level-0 work vectors (nfp * nrhs)
    const int *coarse_off;          // [ninst+1] offsets of the dense coarsest inverses
    double *cinv64;
    float *cinv32;
    double omega_l[kMaxLevels];     // Jacobi weight of each level's smoother
    double alpha;                   // coarse-correction (prolongation) factor
    double alpha_sa;                // coarse-correction factor across smoothed transfers
    double omega_p;                 // prolongation-smoothing weight of those transfers
    const DevAmg *dev;              // this structure's copy in device memory (refreshed by the host)
    int nu;                         // smoothing sweeps before and after the coarse correction
    int max_rows[kMaxLevels];       // rows of the largest instance on each level
    int max_children[kMaxLevels];   // most finer rows aggregated into one row of each level
    int covers[kMaxLevels];         // 1: every row of the finer level is a child of exactly one row
    // A P with the level-1 aggregates (SELL-64 like the face matrix), refilled
    // every solve from the face matrix through the gather lists
    const long long *ap_slice_ptr;
    const int *ap_col, *ap_src_ptr, *ap_src;
    long long ap_slots, ap_nsrc;
    double *ap64;
    float *ap32;
    // the same slots as sums of the FP64 element values S'_c (element-by-
    // element layout), for solves that fill no face matrix (null: none)
    const int *ap_esrc_ptr, *ap_esrc;
    long long ap_ensrc;
};
// DevAmg is larger than the kernel-argument limit: kernels that need the
// whole hierarchy read a copy in device memory.

hipError_t launch_static_ip(hipStream_t st, const DevTopo &T, const DevGeom &G, int maxn,
                            const double *perm, const double *grav, double *binv_pk,
                            double *gflux);

// Inner products of MimeticIPAnisoRelpermEvaluator for per-cell mobility
// tensors lam (lambda_t) and rlam (sum_i rho_i lambda_i), row-major 3x3:
// B^-1 from lam K (through its symmetric part, as k_static_ip) into binv_pk,
// gflux = N (rlam K g).  lamK[9 nc] and dynKg[3 nc] are work arrays.
hipError_t launch_aniso_ip(hipStream_t st, const DevTopo &T, const DevGeom &G, int maxn, const double *perm,
                           const double *lam, const double *rlam, const double *grav, double *lamK, double *dynKg,
                           double *binv_pk, double *gflux);

// gterm[h] = omega[cell(h)] * (x_f - x_c) . grav (MimeticIPEvaluator::gravityTerm).
hipError_t launch_gravity_term(hipStream_t st, const DevTopo &T, const DevGeom &G, const double *grav,
                               const double *omega, double *gterm);

// E (may be null): also write each cell's S'_c (the local Schur complement
// with Dirichlet rows reduced to their diagonal and Dirichlet columns zero) in
// the element-by-element layout, FP64 and FP32.
hipError_t launch_assemble(hipStream_t st, const DevTopo &T, int maxn, int nrhs,
                           const double *binv_pk, const double *gflux, const double *totmob,
                           const double *rhomob, const double *src, const uint8_t *ftype,
                           const double *bcval, DevCellSys C, const DevEbe *E = nullptr);

// q = A p through the element-by-element operator (one launch per colour),
// with the <p, q> partials in E.R (profiling; the PCG launches it itself).
hipError_t launch_ebe_spmv(hipStream_t st, int maxn, int nrhs, const DevEbe &E, const float *p, float *q,
                           DevScal S);

// Face rows from the cell systems (k_fill).  FILL_ALL: matrix values and the
// vectors (right-hand side, Jacobi diagonal, initial iterate and its
// residual); FILL_VEC: the vectors only; FILL_MAT: the matrix values only.
enum FillMode { FILL_ALL = 0, FILL_VEC = 1, FILL_MAT = 2 };
hipError_t launch_fill(hipStream_t st, const DevTopo &T, int nrhs, const DevCellSys &C,
                       const uint8_t *ftype, const double *bcval, const int *inst_reg,
                       int dirichlet_guess, double *val64, float *val32, double *b,
                       double *dinv64, float *dinv32, double *x0, double *r0, int mode = FILL_ALL);

// One Jacobi-PCG iteration (spmv_pq, update_xr, update_p; in FP32 also the
// fused FP64 residual replacement, which needs val64, b and x64).
// E (FP32 only, may be null): the SpMV runs element by element; r64 is the
// FP64 accumulator of its fused residual replacement.
template <typename MT, typename VT>
hipError_t launch_pcg_iteration(hipStream_t st, const DevTopo &T, int nrhs, const MT *val,
                                const VT *dinv, VT *x, VT *r, VT *p, VT *q, DevRed R,
                                DevScal S, const double *val64, const double *b, double *x64,
                                double tol2, const DevEbe *E = nullptr, double *r64 = nullptr);

// The SpMV + <p, Ap> kernel alone (profiling).
template <typename MT, typename VT>
hipError_t launch_spmv(hipStream_t st, const DevTopo &T, int nrhs, const MT *val, const VT *p,
                       VT *q, DevRed R, DevScal S);

// FP64 PCG start: r = b - A x (r_given: r holds it already), p = dinv r, rz,
// rr0, thresholds, activation.
hipError_t launch_pcg_start64(hipStream_t st, const DevTopo &T, int nrhs, const double *val,
                              const double *dinv, const double *b, const double *x, double *r,
                              double *p, double tol2, DevRed R, DevScal S, bool r_given);

// FP32 start / FP64 residual replacement (see k_reliable; r0 non-null: the
// residual b - A64 x64 is given there), final fold of x32 into x64, and the
// FP64 residual norm for the report.
hipError_t launch_reliable(hipStream_t st, const DevTopo &T, int nrhs, const double *val64,
                           const double *b, const double *x64, const float *dinv32, float *r32,
                           float *p32, int init, double tol2, DevRed R, DevScal S,
                           const double *r0 = nullptr);
hipError_t launch_merge_all(hipStream_t st, const DevTopo &T, int nrhs, double *x64, float *x32);
// Clears the status words, iteration counts and replacement flags of nsys
// systems and zeroes x32[0, nx) (x32 may be null).
hipError_t launch_clear_state(hipStream_t st, const DevScal &S, int nsys, float *x32, long long nx);
hipError_t launch_resnorm(hipStream_t st, const DevTopo &T, int nrhs, const double *val64,
                          const double *b, const double *x64, DevRed R, DevScal S);

// Multigrid: per-solve set-up (Galerkin products, diagonals, coarsest inverses)
// and the preconditioned iteration pieces.
// The up-sweep operator M = P - w D^-1 T of each smoothed transfer, once the
// smoother weights are known (after launch_amg_setup).
hipError_t launch_amg_combine(hipStream_t st, const DevAmg &A, bool mixed);
hipError_t launch_amg_setup(hipStream_t st, const DevTopo &T, const DevAmg &A, const double *val64,
                            bool mixed, const double *elem64 = nullptr, bool ap_only = false);
// Lanczos on D^-1/2 A_l D^-1/2 (FP64) for the Jacobi weight of level l:
// start vector w with |w|^2 partials; one step v = scale w, y = D^-1/2 A
// D^-1/2 v with <v, y> partials; w = y - a v - b vp with |w|^2 partials.
// Partials: one per 256 rows (level 0: T.nfp rows, else A.lev[l].n).
hipError_t launch_lanczos_init(hipStream_t st, int n, double *w, double *part);
hipError_t launch_lanczos_step(hipStream_t st, const DevTopo &T, const DevAmg &A, int level, const double *val0,
                               const double *dinv0, const double *w, double scale, double *v, double *u,
                               double *y, double *part);
hipError_t launch_lanczos_update(hipStream_t st, int n, const double *y, const double *v, const double *vp,
                                 double a, double b, double *w, double *part);
// One level-0 Jacobi sweep with the <b, xo> partials (the last multigrid
// smoothing sweep alone; the autotuner times it).
template <typename VT>
hipError_t launch_sweep0(hipStream_t st, const DevTopo &T, int nrhs, const VT *val0, const VT *dinv0, const VT *b,
                         const VT *x, VT *xo, double w, DevRed R);
// The fused post-smoothing sweep alone (z from x0, res0 and the level-1
// correction of the last V-cycle; the autotuner times it).
template <typename VT>
hipError_t launch_post0(hipStream_t st, const DevTopo &T, const DevAmg &A, int nrhs, const VT *dinv0,
                        const VT *r, VT *z, DevRed R);
// E (FP32 only, may be null): the SpMV and, with one smoothing sweep, the
// level-0 pre-smoothing residual run element by element; r64 is the FP64
// accumulator of the fused residual replacement (FP32 iteration).
template <typename VT>
hipError_t launch_vcycle(hipStream_t st, const DevTopo &T, const DevAmg &A, int nrhs, const VT *val0,
                         const VT *dinv0, const VT *r, VT *z, DevRed R, const DevEbe *E = nullptr);
template <typename MT, typename VT>
hipError_t launch_pcg_iteration_amg(hipStream_t st, const DevTopo &T, const DevAmg &A, int nrhs,
                                    const MT *val, const VT *dinv, VT *x, VT *r, VT *p, VT *q, VT *z,
                                    DevRed R, DevScal S, const double *val64, const double *b,
                                    double *x64, double tol2, const DevEbe *E = nullptr,
                                    double *r64 = nullptr);
// z = M^-1 r, rz, p = z (start of the preconditioned iteration after
// launch_pcg_start64 / launch_reliable(init)).
template <typename VT>
hipError_t launch_amg_start(hipStream_t st, const DevTopo &T, const DevAmg &A, int nrhs, const VT *val0,
                            const VT *dinv0, const VT *r, VT *z, VT *p, DevRed R, DevScal S,
                            double *x64, VT *x, const DevEbe *E = nullptr);

// Recovery and post-processing.
// out[k*ngroups + g] = sum of v[k*nhf + ghf[e]] over e in [gptr[g], gptr[g+1]).
hipError_t launch_group_sum(hipStream_t st, int nrhs, int ngroups, int nhf, const int *gptr, const int *ghf,
                            const double *v, double *out);

// Adjoint of the recovered fluxes (see k_adj_cell / k_adj_grad): the adjoint
// right-hand side into b and r (x zeroed), and the gradients after the adjoint
// solve.
hipError_t launch_adjoint_rhs(hipStream_t st, const DevTopo &T, int maxn, int nrhs, const double *binv_pk,
                              const double *totmob, const DevCellSys &C, const uint8_t *ftype,
                              const double *w, double *y, double *b, double *x, double *r);
hipError_t launch_adjoint_grad(hipStream_t st, const DevTopo &T, const DevGeom &G, int maxn, int nrhs,
                               const double *binv_pk, const double *gflux, const double *totmob,
                               const double *rhomob, const double *grav, const DevCellSys &C,
                               const double *press, const double *pi, const double *mu,
                               const uint8_t *ftype, const double *w, double *gK, double *gtm,
                               double *grm, double *gsrc);

hipError_t launch_recover(hipStream_t st, const DevTopo &T, int maxn, int nrhs,
                          const double *binv_pk, const double *gflux, const double *totmob,
                          const double *rhomob, const DevCellSys &C, const double *x64,
                          double *press, double *outflux);

hipError_t launch_post_process(hipStream_t st, const DevTopo &T, int nrhs, const double *vin,
                               double *vout, unsigned long long *maxbits);

}  // namespace mhs

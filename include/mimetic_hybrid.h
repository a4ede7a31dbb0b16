/*
 * mimetic_hybrid.h -- C-ABI of the MI355X mimetic hybrid face-pressure solver.
 *
 * This is the boundary that replaces the inside of opm-porsol's
 *   template <class GI, class RI, class BCI, template<class,class> class IP>
 *   class Opm::IncompFlowSolverHybrid
 * (opm/porsol/mimetic/IncompFlowSolverHybrid.hpp:362-1902) together with its
 * inner-product plug-in Opm::MimeticIPEvaluator
 * (opm/porsol/mimetic/MimeticIPEvaluator.hpp:84-555) and the dune-istl linear
 * solvers it calls.  A header-only C++ adapter with the reference's template
 * signature sits on top (opm-porsol_amd/host/opm/porsol/mimetic/
 * IncompFlowSolverHybrid.hpp); this header is what it, or any FFI, binds.
 *
 * Conventions
 *  - Every function returns an int status (MHS_OK == 0) and never throws.
 *    On failure mhs_last_error() describes the cause.  The reference throws
 *    std::runtime_error for non-convergence (IncompFlowSolverHybrid.hpp:
 *    1416-1419, 1536-1539) and for an unknown linsolver_type (:698-700); the
 *    adapter turns MHS_ERR_NOT_CONVERGED / MHS_ERR_INVALID back into those.
 *  - Host pointers are read during the call only (copied in); the solver
 *    owns all device memory.  Entry points named *_device take device
 *    pointers that must stay valid until the call returns.
 *  - One handle per host thread.  Work is issued on the stream given to
 *    mhs_create (NULL: the solver creates its own stream at default
 *    priority).
 *  - Arrays are indexed like the reference: cells in iteration order, half-
 *    faces cell by cell in local-face order (FlowSolution::outflux(hf),
 *    IncompFlowSolverHybrid.hpp:430-433), face degrees of freedom in the
 *    numbering of enumerateGridDof (:956-1067).
 *  - Batches: a grid may be the disjoint union of several independent
 *    instances (inst_cellpos).  Each instance is solved as if by its own
 *    IncompFlowSolverHybrid object: its own dof numbering, its own
 *    regularisation, its own convergence test.  Face pressures of instance t
 *    occupy a contiguous range of length nf_t, instances in order.
 *  - nrhs right-hand sides share one matrix.  Per-rhs arrays are laid out
 *    rhs-major: src[k*nc + c], bc_value[k*nhf + h], pressure[k*nc + c], ...
 */
#ifndef MIMETIC_HYBRID_H
#define MIMETIC_HYBRID_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mhs_solver mhs_solver;

enum mhs_status {
    MHS_OK = 0,
    MHS_ERR_INVALID = 1,        /* bad argument, unknown linsolver_type (:698-700) */
    MHS_ERR_STATE = 2,          /* call-order precondition (the reference asserts, :538, :569) */
    MHS_ERR_NOT_CONVERGED = 3,  /* linear solver did not converge (:1416-1419) */
    MHS_ERR_HIP = 4,            /* HIP runtime error */
    MHS_ERR_ALLOC = 5,          /* device or host allocation failed */
    MHS_ERR_BREAKDOWN = 6,      /* CG breakdown: operator not positive definite */
    MHS_ERR_RANGE = 7           /* transport: saturation out of range after every retry
                                   (EulerUpstream_impl.hpp:198-205, 343-345) */
};

/* FlowBC types (opm/porsol/common/BoundaryConditions.hpp:57, 121-171).
 * The default FlowBC is Neumann with value 0 (:125-128). */
enum mhs_bc_type { MHS_BC_NEUMANN = 0, MHS_BC_DIRICHLET = 1, MHS_BC_PERIODIC = 2 };

enum mhs_precision {
    MHS_FP64 = 0,   /* FP64 matrix, FP64 PCG */
    MHS_MIXED = 1   /* FP32 matrix and PCG vectors, FP64 residual replacement */
};

typedef struct {
    double tol;                /* residual_tolerance (:662): ||r||/||r0|| bound */
    int maxit;                 /* linsolver_maxit (:666); 0 -> number of unknowns */
    int linsolver_type;        /* (:664) 0: Jacobi-PCG, standing in for SeqILU0-CG
                                  (:1387-1420); it needs more iterations than ILU0,
                                  so a maxit tuned for ILU0 can end in
                                  MHS_ERR_NOT_CONVERGED; 1 (AMG, :1481-1541), 2 (KAMG,
                                  :1609-1667) and 3 (FastAMG, :1546-1605): the same
                                  multigrid-preconditioned CG (Jacobi-PCG without
                                  logical cell indices).  All converge to the
                                  reference's discrete solution. */
    int verbosity;             /* linsolver_verbosity (:663) */
    int same_matrix;           /* (:665, :1491) the caller's promise that the matrix is the
                                  previous solve's: the multigrid coarse levels, smoothed
                                  transfer and coarsest inverses are kept (the face system
                                  is still assembled; the PCG uses it) */
    double prolongate_factor;  /* (:667) multigrid coarse-correction factor (default 1.6) */
    int smooth_steps;          /* (:668) multigrid smoothing sweeps before and after
                                  the coarse correction */
    int precision;             /* enum mhs_precision */
    int nrhs;                  /* right-hand sides sharing the matrix, 1..4 */
    int chunk;                 /* PCG iterations per captured graph launch */
    double replace_factor;     /* mixed: replace the FP32 recursive residual by
                                  the FP64 true residual once it has dropped by
                                  this factor since the last replacement
                                  (default 0.01) */
    int use_graph;             /* capture PCG chunks in a hipGraph */
    int precond;               /* -1: from linsolver_type (0 -> Jacobi, 1..3 -> multigrid
                                  when logical cell indices were given); 0: Jacobi;
                                  1: aggregation multigrid V-cycle */
    double smoother_weight;    /* > 0: Jacobi weight of the multigrid smoother on every
                                  level; 0 (default): 4 / (3 rho(D^-1 A)) per level,
                                  rho from Lanczos steps on the first multigrid solve
                                  of each hierarchy */
    int autotune;              /* 1 (default): on the first solve of a structure (per
                                  nrhs, precision and preconditioner) time every
                                  candidate launch shape of the PCG SpMV and of the
                                  multigrid smoothing sweep and keep the fastest;
                                  0: fixed shapes (256 threads, XCD-remapped) */
    int spmv_threads;          /* 0: from autotune; 256, 512 or 1024: force this many
                                  threads per 1024-row block for the PCG SpMV */
    int sweep_threads;         /* same for the multigrid smoothing sweep */
    int tail_rows;             /* -1: from autotune (0 when autotune is 0); >= 0: coarse
                                  levels whose largest instance has at most this many
                                  rows run in one workgroup per instance */
    int element_operator;      /* mixed precision: apply the face operator of the PCG
                                  product, its FP64 residual replacement and the
                                  multigrid pre-smoothing residual cell by cell from the
                                  local Schur complements (only when the structure has
                                  no periodic faces and no instance is regularised).
                                  1 (default): when the autotuner finds it faster than
                                  the assembled FP32 matrix (timed on the first solve of
                                  each nrhs; always with autotune 0); 2: whenever it
                                  applies; 0: never (the assembled FP32 face matrix) */
} mhs_solve_options;

typedef struct {
    int iterations;            /* PCG iterations, maximum over systems */
    int replacements;          /* FP64 residual replacements (mixed) */
    int converged;             /* 1 when every system reached tol */
    double reduction;          /* max over systems of the FP64 ||b-Ax||/||r0|| */
    int precond;               /* preconditioner used (0 Jacobi, 1 multigrid) */
    int levels;                /* multigrid levels (0 with Jacobi) */
    long long mg_rows;         /* rows of the coarse multigrid levels, summed */
    long long mg_nnz;          /* stored entries of the coarse level matrices, summed */
    double assemble_ms, solve_ms, recover_ms;
    int spmv_threads, spmv_xcd;   /* launch shape of the PCG SpMV (threads per 1024 rows,
                                     XCD-contiguous row ranges) */
    int sweep_threads, sweep_xcd; /* same for the multigrid smoothing sweep */
    int tail_rows;                /* coarse-level size limit of the one-workgroup tail */
    double tune_ms;               /* time spent autotuning in this call */
    double smoother_weights[8];   /* multigrid Jacobi weight of levels 0..7 (0 beyond) */
    int coarsening[8][3];         /* logical block factors of the transfers from levels
                                     0..7 (1 or 2 along i, j, k; 0 beyond) */
    long long ap_nnz;             /* stored entries of A P (face rows x level-1 aggregates)
                                     read by the post-smoothing sweep; 0 without multigrid */
    int element_operator;         /* 1: the solve applied the level-0 operator cell by cell */
    int colours;                  /* cell colours of that operator (0 when not used) */
    long long element_values;     /* packed local-matrix entries it stores (with padding) */
    long long level_rows[8];      /* multigrid: rows of levels 0..7 (level 0 the face rows,
                                     unpadded; 0 beyond the coarsest) */
    long long level_nnz[8];       /* stored entries of those level matrices (level 0: the
                                     assembled face matrix) */
    long long transfer_nnz[2];    /* entries of the smoothed transfer P and of T = A P
                                     (0 when no transfer is smoothed) */
    int smoothed_level;           /* level whose transfer to the next is smoothed, -1: none */
    int face_matrix;              /* 1: the solve filled the assembled face matrix; 0: it
                                     applied the operator cell by cell with known smoother
                                     weights and read none (filled when asked for later,
                                     e.g. by mhs_get_system) */
    int mg_setup_reused;          /* 1: same_matrix kept the previous multigrid set-up */
} mhs_solve_report;

typedef struct {
    int num_cells, num_half_faces, num_faces, num_internal_faces, max_ncf;
    int num_instances, num_rows_padded;
    long long nnz;             /* stored entries of the face matrix (incl. padding) */
} mhs_dims;

void mhs_default_options(mhs_solve_options *opt);

/* Solver object (IncompFlowSolverHybrid default constructor). */
int mhs_create(mhs_solver **out, int device, void *stream);
int mhs_destroy(mhs_solver *s);
const char *mhs_last_error(const mhs_solver *s);

/* IncompFlowSolverHybrid::clear (:498-516): forget grid, structure, IPs. */
int mhs_clear(mhs_solver *s);

/* Grid through the GridInterfaceEuler accessors (GridInterfaceEuler.hpp:
 * 126-212, 335-360).  facepos[num_cells+1]; neighbour[nhf] is the neighbour
 * cell (-1 on the boundary: Face::boundary()); geometry per cell (volume,
 * centroid xyz) and per half-face (area, centroid xyz, unit outer normal xyz).
 * inst_cellpos[num_instances+1] splits the cells into independent instances
 * (NULL with num_instances == 1 means one instance). */
int mhs_set_grid(mhs_solver *s, int num_instances, const int *inst_cellpos,
                 int num_cells, const int *facepos, const int *neighbour,
                 const double *cell_volume, const double *cell_centroid,
                 const double *face_area, const double *face_centroid,
                 const double *face_normal);

/* Logical (i, j, k) index of every cell within its instance, 0-based
 * (ijk[3*c + d]).  Optional, no counterpart in the reference: it enables the
 * aggregation multigrid preconditioner, whose aggregates are 2x2x2 blocks of
 * cells (the reference's AMG, IncompFlowSolverHybrid.hpp:1481-1541, builds
 * its aggregates algebraically inside dune-istl).  Call after mhs_set_grid and
 * before mhs_init_system_structure. */
int mhs_set_cell_ijk(mhs_solver *s, const int *ijk);

/* IncompFlowSolverHybrid::initSystemStructure (:534-545): dof enumeration
 * (:945-1110) and the face-system sparsity (:1115-1319).  Only periodic
 * information is used, as in the reference: hf_bc_type[nhf] (MHS_BC_*) and
 * hf_periodic_partner[nhf], the partner half-face of each periodic boundary
 * half-face (the reference resolves it via getPeriodicPartner(bid),
 * BoundaryConditions.hpp:289-293).  Either may be NULL (no periodic faces). */
int mhs_init_system_structure(mhs_solver *s, const int *hf_bc_type,
                              const int *hf_periodic_partner);

/* IncompFlowSolverHybrid::computeInnerProducts (:564-578) ->
 * MimeticIPEvaluator::buildStaticContrib (MimeticIPEvaluator.hpp:201-270).
 * perm[9*num_cells]: row-major 3x3 tensors (ReservoirPropertyCommon_impl.hpp:
 * 403-411), used through their symmetric part (K + K')/2; gravity[3]. */
int mhs_compute_inner_products(mhs_solver *s, const double *perm,
                               const double *gravity);
/* The static inverse inner products after compute_inner_products, in the
 * layout of MimeticIPEvaluator's Binv_ table (MimeticIPEvaluator.hpp:553):
 * cell by cell, each ncf x ncf matrix column-major (sum ncf^2 doubles), and
 * the static gravity fluxes N*K*g per half-face (gflux_, :554).  Either
 * pointer may be NULL. */
int mhs_get_inner_products(mhs_solver *s, double *binv, double *gflux);
/* MimeticIPEvaluator::gravityTerm (MimeticIPEvaluator.hpp:477-523) for every
 * cell at once: gterm[h] = omega[c] * (x_f - x_c) . g for each half-face h of
 * cell c, from the geometry given to mhs_set_grid; omega[nc] is the
 * mobility-weighted density (sum_i rho_i lambda_i / lambda_t in the
 * reference's fluid overload, :509-522).  Computed on the device; host arrays
 * in and out. */
int mhs_gravity_term(mhs_solver *s, const double *gravity, const double *omega, double *gterm);
/* topologyIsSane (IncompFlowSolverHybrid.hpp:108-130, asserted in
 * initSystemStructure :540): *sane = 1 when every cell's internal neighbours
 * are distinct, i.e. no two cells share more than one face, else 0.  Host
 * check on the arrays of mhs_set_grid; needs no handle and no GPU. */
int mhs_topology_is_sane(int num_cells, const int *cell_facepos, const int *neighbour, int *sane);
/* Alias: the "assembleStatic" step named in commented-out reference calls
 * (SimulatorBase.hpp:211) is computeInnerProducts. */
int mhs_assemble_static(mhs_solver *s, const double *perm, const double *gravity);

/* IncompFlowSolverHybrid::solve (:657-703): assembleDynamic (:1324-1382),
 * the linear solve (:1387-1667) and computePressureAndFluxes (:1672-1730).
 * The FluidInterface enters through the per-cell values it yields in
 * MimeticIPEvaluator::computeDynamicParams (MimeticIPEvaluator.hpp:297-311):
 *   totmob[c] = sum of phase mobilities, rhomob[c] = sum rho_i*lambda_i.
 * src[nrhs*nc] (m^3/s, >0 injection); hf_bc_type[nhf]; hf_bc_value[nrhs*nhf]
 * (Dirichlet pressure / Neumann outflux / periodic pressure difference,
 * read on boundary half-faces only).  report may be NULL. */
int mhs_solve(mhs_solver *s, const double *totmob, const double *rhomob,
              const double *src, const int *hf_bc_type, const double *hf_bc_value,
              const mhs_solve_options *opt, mhs_solve_report *report);
/* Same with totmob, rhomob, src and hf_bc_value in device memory
 * (hf_bc_type stays a host array). */
int mhs_solve_device(mhs_solver *s, const double *totmob, const double *rhomob,
                     const double *src, const int *hf_bc_type,
                     const double *hf_bc_value, const mhs_solve_options *opt,
                     mhs_solve_report *report);

/* solve() with the second inner-product plug-in, MimeticIPAnisoRelpermEvaluator
 * (opm/porsol/mimetic/MimeticIPAnisoRelpermEvaluator.hpp:85-430): tensor
 * mobilities per cell instead of scalars.  lambda_t[9*nc] is the total
 * mobility tensor sum_i lambda_i and rho_lambda[9*nc] is sum_i rho_i lambda_i,
 * both row-major like perm (computeDynamicParams :299-330).  Per cell
 * B^-1 = (N (lambda_t K) N' + t A (I - QQ') A) / |c| with t = 6 tr(lambda_t K)/3
 * (getInverseMatrix :363-400) and the gravity flux N (rho_lambda K g)
 * (gravityFlux :406-416), g as given to compute_inner_products.  lambda_t K
 * enters through its symmetric part, as K does in the scalar path: equal to
 * the reference's product whenever lambda_t and K share principal axes (the
 * reference's symmetric Schur update assumes a symmetric B^-1).  The rest is
 * mhs_solve; the static inner products are kept for later scalar solves.
 * An adjoint of such a solve is refused. */
int mhs_solve_aniso(mhs_solver *s, const double *lambda_t, const double *rho_lambda,
                    const double *src, const int *hf_bc_type, const double *hf_bc_value,
                    const mhs_solve_options *opt, mhs_solve_report *report);
/* Same with lambda_t, rho_lambda, src and hf_bc_value in device memory. */
int mhs_solve_aniso_device(mhs_solver *s, const double *lambda_t, const double *rho_lambda,
                           const double *src, const int *hf_bc_type, const double *hf_bc_value,
                           const mhs_solve_options *opt, mhs_solve_report *report);

/* IncompFlowSolverHybrid::getSolution (:828-831) -> FlowSolution::pressure
 * (:411-414) and ::outflux (:426-433), plus the face (contact) pressures the
 * reference keeps in the private soln_ (:935).  Any pointer may be NULL.
 * cell_pressure[nrhs*nc], face_pressure[nrhs*nf], outflux[nrhs*nhf]. */
int mhs_get_solution(mhs_solver *s, double *cell_pressure, double *face_pressure,
                     double *outflux);
/* Device pointers to the same data (valid until the next solve/clear/destroy).
 * face_pressure is in the padded row numbering (see mhs_get_face_rows). */
int mhs_solution_device(mhs_solver *s, const double **cell_pressure,
                        const double **face_pressure, const double **outflux);

/* IncompFlowSolverHybrid::postProcessFluxes (:754-807). */
int mhs_post_process_fluxes(mhs_solver *s, double *max_modification);

/* Sizes (IncompFlowSolverHybrid::printStats :848-871 reports the same). */
int mhs_get_dims(mhs_solver *s, mhs_dims *dims);
/* cellFaces_ (the local-to-global face map, :926): hf_face[nhf] in the
 * unpadded per-instance-contiguous numbering; inst_facepos[num_instances+1]
 * gives each instance's face range.  Either may be NULL. */
int mhs_get_cell_faces(mhs_solver *s, int *hf_face, int *inst_facepos);
/* Padded row of every face (num_faces entries), for device-side access. */
int mhs_get_face_rows(mhs_solver *s, int *face_row);

/* IncompFlowSolverHybrid::printSystem (:895-905): the assembled face system
 * of right-hand side k after the last solve, as CSR in the unpadded face
 * numbering (explicit zeros kept).  Call with col == NULL to get *nnz. */
int mhs_get_system(mhs_solver *s, int k, int *row_ptr, int *col, double *val,
                   double *rhs, long long *nnz);

/* computeInnerProducts with perm[9*num_cells] in device memory (gravity[3]
 * stays a host array); same result as mhs_compute_inner_products. */
int mhs_compute_inner_products_device(mhs_solver *s, const double *perm,
                                      const double *gravity);

/* Copies of the last solution into caller-owned device buffers:
 * cell_pressure[nrhs*nc], outflux[nrhs*nhf] (either may be NULL), ordered
 * after the copy on the stream given to mhs_create. */
int mhs_get_solution_device(mhs_solver *s, double *cell_pressure, double *outflux);

/* Sums of the last outflux over groups of half-faces, in a fixed order on
 * the device: sums[k*ngroups + g] = sum over e in [group_ptr[g],
 * group_ptr[g+1]) of outflux(k, group_hf[e]).  No counterpart in the
 * reference: the boundary-flux integrals of periodic upscaling (effective
 * permeability) and of well or producer rates, without copying every flux to
 * the host. */
int mhs_sum_outflux(mhs_solver *s, int ngroups, const int *group_ptr,
                    const int *group_hf, double *sums);
/* mhs_sum_outflux without any device allocation.  No counterpart in the
 * reference: for callers that must not allocate device memory (serving loops
 * and harnesses that forbid hipMalloc).  The group lists and the sums are
 * staged in a per-GPU scratch array of 16 MiB that the library holds until
 * the process ends: (ngroups + 1) and ne ints, each rounded up to 256 bytes,
 * plus nrhs * ngroups doubles must fit, else MHS_ERR_INVALID.  Every call on
 * one GPU shares that array, so calls that overlap on one GPU (several host
 * threads, several handles) race; calls on one handle, which are ordered,
 * are safe.  Same results as mhs_sum_outflux. */
int mhs_sum_outflux_noalloc(mhs_solver *s, int ngroups, const int *group_ptr,
                            const int *group_hf, double *sums);

/* Adjoint of the last solve.  No counterpart in the reference: this is the
 * differentiated form of solve() (:657-703) for gradient-based calibration.
 * dJ_dv[nrhs*nhf] is the gradient of a scalar objective J with respect to
 * outflux(hf) of the last solve (before postProcessFluxes).  The call solves
 * the adjoint system with the same face operator, preconditioner and
 * precision as the last solve (opt gives tol, maxit, chunk, use_graph), then
 * returns, each optional (NULL to skip):
 *   grad_perm[9*nc]    dJ/dK_lm per cell (row-major, like perm); the inner
 *                      products use the symmetric part of each tensor, so
 *                      this gradient is symmetric;
 *   grad_totmob[nc]    dJ/d(total mobility);
 *   grad_rhomob[nc]    dJ/d(sum rho_i lambda_i);
 *   grad_src[nrhs*nc]  dJ/dsrc per right-hand side.
 * The first three are summed over right-hand sides.  Boundary values and
 * geometry are held fixed.  Periodic boundaries are rejected.  The forward
 * solution stays available through mhs_get_solution. */
int mhs_adjoint(mhs_solver *s, const double *dJ_dv, const mhs_solve_options *opt,
                double *grad_perm, double *grad_totmob, double *grad_rhomob,
                double *grad_src, mhs_solve_report *report);
/* Same with dJ_dv and every output in device memory. */
int mhs_adjoint_device(mhs_solver *s, const double *dJ_dv, const mhs_solve_options *opt,
                       double *grad_perm, double *grad_totmob, double *grad_rhomob,
                       double *grad_src, mhs_solve_report *report);

/* estimateCellVelocity (opm/porsol/common/SimulatorUtilities.hpp:59-81) on
 * the last solution (right-hand side k): cell_velocity[3*c + d] = sum over the
 * faces of cell c of (x_f - x_c)[d] * outflux(f) / |c|.  The _device variant
 * writes a device buffer, ordered on the stream given to mhs_create. */
int mhs_estimate_cell_velocity(mhs_solver *s, int k, double *cell_velocity);
int mhs_estimate_cell_velocity_device(mhs_solver *s, int k, double *cell_velocity);

/* Profiling hook, no counterpart in the reference: average device time (ms)
 * of `reps` back-to-back launches of kernel `which` on the solver's own
 * stream, timed with HIP events, on the data of the last solve: 0, the PCG
 * product q = A p with <p, Ap> as the last solve ran it (in mixed precision
 * element by element when it used that operator: all colour launches); 1,
 * the product with the assembled face matrix in the precision of the last
 * solve. */
int mhs_time_kernel(mhs_solver *s, int which, int reps, double *avg_ms);

#ifdef __cplusplus
}
#endif

#endif /* MIMETIC_HYBRID_H */

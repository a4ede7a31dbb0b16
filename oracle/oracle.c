/*
 * oracle.c -- FP64 CPU restatement of opm-porsol's mimetic hybrid pressure
 * solve (IncompFlowSolverHybrid + MimeticIPEvaluator).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker against which the HIP
 * library is compared.  It is never linked into, loaded by, or called from the
 * product library; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it.
 *
 * Every function restates one piece of the reference algorithm, in the same
 * order of operations, and cites the reference file:line it follows.  Paths are
 * relative to the opm-porsol source tree.
 *
 * The reference itself cannot be compiled here (dune-common, dune-grid,
 * dune-istl, opm-core, Boost and LAPACK are absent).  Deviations, each a
 * mathematically exact replacement:
 *   - LAPACK dgeqrf+dorgqr (Matrix.hpp:729-759) -> Householder QR written
 *     here.  Only Q*Q' is used, which is independent of the sign convention.
 *   - dune-istl BCRSMatrix (random build mode) -> COO triplets summed into
 *     CSR.  Same entries, different (commutative) summation order.
 *   - dune-istl CGSolver + SeqILU0 -> restated below (orc_cg).
 *
 * Pinning status (what checks each function; see tests/test_oracle_*.py):
 *   orc_enumerate_dof  : brute-force ordering rules on tiny grids + invariants
 *   orc_static_ip      : closed form on boxes (block diagonal), mimetic
 *                        consistency on arbitrary hexahedra, Q-sign invariance
 *   orc_assemble       : dense unreduced mixed saddle-point system on tiny grids
 *                        (Dirichlet, Neumann, periodic), symmetry, row sums
 *   orc_cg             : dense Cholesky on small systems, residual reduction
 *   orc_recover        : mass balance per cell, linear-pressure exactness,
 *                        known-answer error norms of examples/timing_results.txt
 *   orc_post_process   : antisymmetry of twin fluxes, hand-computed cases
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAXF 64

enum { ORC_NEUMANN = 0, ORC_DIRICHLET = 1, ORC_PERIODIC = 2 };
enum { FT_INTERNAL = 0, FT_DIRICHLET = 1, FT_NEUMANN = 2, FT_PERIODIC = 3 };

/* ------------------------------------------------------------------------ */
/* Degree-of-freedom enumeration.
 * Follows IncompFlowSolverHybrid.hpp:956-1067 (enumerateGridDof): internal
 * faces are numbered first, in order of discovery while iterating cells and
 * their faces (a face is discovered from the cell that is visited first), then
 * boundary faces in cell/local-face order.
 *   nc, facepos[nc+1], nb[nhf] (neighbour cell, <0 on the boundary)
 *   out: hf_dof[nhf]; info[0]=num_internal, info[1]=total faces, info[2]=max ncf
 * Returns 0, or -1 if a neighbour index is invalid / a shared face is missing.
 */
int orc_enumerate_dof(int nc, const int *facepos, const int *nb, int *hf_dof,
                      int *info)
{
    int *cell = (int *)malloc(sizeof(int) * (size_t)(nc > 0 ? nc : 1));
    int *fpos = (int *)malloc(sizeof(int) * (size_t)(nc + 1));
    int nhf = facepos[nc];
    int *faces = (int *)malloc(sizeof(int) * (size_t)(nhf > 0 ? nhf : 1));
    int nfaces = 0, max_ncf = 0, c, k, rc = 0;
    for (c = 0; c < nc; ++c) cell[c] = -1;
    /* First pass (:980-1005). */
    fpos[0] = 0;
    for (c = 0; c < nc; ++c) {
        int ncf = facepos[c + 1] - facepos[c];
        cell[c] = c;
        for (k = facepos[c]; k < facepos[c + 1]; ++k) {
            if (nb[k] >= 0) {
                int c1 = nb[k];
                if (c1 >= nc || c1 == c) { rc = -1; goto done; }
                if (cell[c1] == -1) faces[nfaces++] = c1;
            }
        }
        fpos[c + 1] = nfaces;
        if (ncf > max_ncf) max_ncf = ncf;
    }
    {
        int total = nfaces;
        /* Second pass (:1020-1065). */
        for (c = 0; c < nc; ++c) {
            for (k = facepos[c]; k < facepos[c + 1]; ++k) {
                if (nb[k] < 0) {
                    hf_dof[k] = total++;
                } else {
                    int t = c, seek = nb[k], s, e, p;
                    if (cell[seek] < cell[t]) { int tmp = t; t = seek; seek = tmp; }
                    s = fpos[cell[t]];
                    e = fpos[cell[t] + 1];
                    for (p = s; p < e; ++p)
                        if (faces[p] == seek) break;
                    if (p == e) { rc = -1; goto done; }
                    hf_dof[k] = p;
                }
            }
        }
        info[0] = nfaces;
        info[1] = total;
        info[2] = max_ncf;
    }
done:
    free(cell); free(fpos); free(faces);
    return rc;
}

/* Periodic partner dofs.  Follows IncompFlowSolverHybrid.hpp:1071-1110
 * (enumerateBCDof): for every periodic boundary half-face, ppartner[dof] is the
 * dof of the partner half-face and vice versa.  hf_partner[k] gives the partner
 * half-face directly (the reference resolves it through boundary ids).
 * ppartner must hold nf ints; it is filled with -1 elsewhere.
 * Returns the number of periodic half-faces.
 */
int orc_periodic_dof(int nhf, const int *nb, const int *hf_dof,
                     const int *hf_bctype, const int *hf_partner, int nf,
                     int *ppartner)
{
    int k, n = 0;
    for (k = 0; k < nf; ++k) ppartner[k] = -1;
    for (k = 0; k < nhf; ++k) {
        if (nb[k] < 0 && hf_bctype[k] == ORC_PERIODIC && hf_partner[k] >= 0) {
            int d1 = hf_dof[k], d2 = hf_dof[hf_partner[k]];
            ppartner[d1] = d2;
            ppartner[d2] = d1;
            ++n;
        }
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* Householder QR of an m x 3 column-major matrix A (lda = m), replaced in
 * place by the explicit m x 3 factor Q.  Stands in for LAPACK dgeqrf+dorgqr
 * as used by orthogonalizeColumns (Matrix.hpp:729-759). */
static void householder_q(int m, int n, double *A)
{
    double v[ORC_MAXF][3], beta[3];
    int j, i, l;
    for (j = 0; j < n; ++j) {
        double alpha, nrm = 0.0, s, v0;
        for (i = j; i < m; ++i) nrm += A[i + j * m] * A[i + j * m];
        nrm = sqrt(nrm);
        alpha = A[j + j * m];
        s = (alpha >= 0.0) ? -nrm : nrm;           /* new diagonal R(j,j) */
        v0 = alpha - s;
        for (i = 0; i < m; ++i) v[i][j] = 0.0;
        if (nrm == 0.0 || v0 == 0.0) { beta[j] = 0.0; continue; }
        v[j][j] = 1.0;
        for (i = j + 1; i < m; ++i) v[i][j] = A[i + j * m] / v0;
        beta[j] = -v0 / s;                           /* 2/(v'v) */
        /* apply H_j to the trailing columns */
        for (l = j; l < n; ++l) {
            double d = 0.0;
            for (i = j; i < m; ++i) d += v[i][j] * A[i + l * m];
            d *= beta[j];
            for (i = j; i < m; ++i) A[i + l * m] -= d * v[i][j];
        }
    }
    /* Q = H_0 H_1 ... H_{n-1} applied to the first n columns of I. */
    for (l = 0; l < n; ++l)
        for (i = 0; i < m; ++i) A[i + l * m] = (i == l) ? 1.0 : 0.0;
    for (j = n - 1; j >= 0; --j) {
        for (l = 0; l < n; ++l) {
            double d = 0.0;
            for (i = j; i < m; ++i) d += v[i][j] * A[i + l * m];
            d *= beta[j];
            for (i = j; i < m; ++i) A[i + l * m] -= d * v[i][j];
        }
    }
}

/* Static inverse inner products.  Follows MimeticIPEvaluator.hpp:201-270
 * (buildStaticContrib) step by step:
 *   Binv <- I, T1 <- N (area*normal), T2 <- C (area*(x_f - x_c)), fa <- diag(area)
 *   gflux_i <- (area_i n_i) . (K g)
 *   T2 <- orth(T2); Binv <- I - T2 T2' (upper, then mirrored)
 *   Binv <- fa Binv fa (upper, then mirrored)
 *   T2 <- T1 K  (K row-major, ReservoirPropertyCommon_impl.hpp:403-411)
 *   Binv <- (T2 T1' + t Binv)/vol,  t = 6 tr(K)/3
 * Output: binv holds sum(ncf^2) doubles, each cell's ncf x ncf matrix stored
 * column-major (as Binv_ in MimeticIPEvaluator.hpp:553); gflux holds nhf. */
int orc_static_ip(int nc, const int *facepos, const double *cvol,
                  const double *ccen, const double *farea, const double *fcen,
                  const double *fnrm, const double *perm, const double *grav,
                  double *binv, double *gflux)
{
    size_t off = 0;
    int c;
    for (c = 0; c < nc; ++c) {
        int f0 = facepos[c], n = facepos[c + 1] - facepos[c], i, j, k;
        const double *K = perm + 9 * (size_t)c;
        double T1[ORC_MAXF * 3], T2[ORC_MAXF * 3], fa[ORC_MAXF], Kg[3], tr, t;
        double *B = binv + off;
        if (n > ORC_MAXF || n < 3) return -1;
        for (i = 0; i < 3; ++i)
            Kg[i] = K[3 * i] * grav[0] + K[3 * i + 1] * grav[1] + K[3 * i + 2] * grav[2];
        for (i = 0; i < n * n; ++i) B[i] = 0.0;
        for (i = 0; i < n; ++i) {
            const int h = f0 + i;
            double a = farea[h];
            B[i + i * n] = 1.0;
            fa[i] = a;
            gflux[h] = 0.0;
            for (j = 0; j < 3; ++j) {
                double fn = fnrm[3 * h + j] * a;
                double fc = (fcen[3 * h + j] - ccen[3 * c + j]) * a;
                T1[i + j * n] = fn;
                T2[i + j * n] = fc;
            }
            gflux[h] = T1[i] * Kg[0] + T1[i + n] * Kg[1] + T1[i + 2 * n] * Kg[2];
        }
        householder_q(n, 3, T2);
        /* Binv <- -T2 T2' + Binv (upper), mirror (Matrix.hpp:829-848) */
        for (j = 0; j < n; ++j)
            for (i = 0; i <= j; ++i) {
                double s = 0.0;
                for (k = 0; k < 3; ++k) s += T2[i + k * n] * T2[j + k * n];
                B[i + j * n] = -s + B[i + j * n];
            }
        for (j = 0; j < n; ++j)
            for (i = j + 1; i < n; ++i) B[i + j * n] = B[j + i * n];
        /* Binv <- fa Binv fa (Matrix.hpp:857-883) */
        for (j = 0; j < n; ++j)
            for (i = 0; i <= j; ++i) B[i + j * n] = fa[i] * B[i + j * n] * fa[j];
        for (j = 0; j < n; ++j)
            for (i = j + 1; i < n; ++i) B[i + j * n] = B[j + i * n];
        /* T2 <- T1 K */
        for (i = 0; i < n; ++i)
            for (j = 0; j < 3; ++j) {
                double s = 0.0;
                for (k = 0; k < 3; ++k) s += T1[i + k * n] * K[3 * k + j];
                T2[i + j * n] = s;
            }
        tr = K[0] + K[4] + K[8];
        t = 6.0 * tr / 3.0;
        /* Binv <- (1/vol) T2 T1' + (t/vol) Binv  (full GEMM, Matrix.hpp:1195-1213) */
        for (j = 0; j < n; ++j)
            for (i = 0; i < n; ++i) {
                double s = 0.0;
                for (k = 0; k < 3; ++k) s += T2[i + k * n] * T1[j + k * n];
                B[i + j * n] = (1.0 / cvol[c]) * s + (t / cvol[c]) * B[i + j * n];
            }
        off += (size_t)n * n;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Sparse accumulation helpers (stand-in for dune BCRSMatrix S_[i][j] += v). */
typedef struct { int r, c; double v; } trip;

typedef struct {
    trip *t;
    size_t n, cap;
} tlist;

static void tl_add(tlist *L, int r, int c, double v)
{
    if (L->n == L->cap) {
        L->cap = L->cap ? 2 * L->cap : 1024;
        L->t = (trip *)realloc(L->t, L->cap * sizeof(trip));
    }
    L->t[L->n].r = r;
    L->t[L->n].c = c;
    L->t[L->n].v = v;
    L->n++;
}

static int trip_cmp(const void *a, const void *b)
{
    const trip *x = (const trip *)a, *y = (const trip *)b;
    if (x->r != y->r) return x->r < y->r ? -1 : 1;
    if (x->c != y->c) return x->c < y->c ? -1 : 1;
    return 0;
}

/* Dynamic assembly.  Follows IncompFlowSolverHybrid.hpp:1324-1382
 * (assembleDynamic), :1736-1779 (setExternalContrib), :1785-1811
 * (buildCellContrib) and :1817-1901 (addCellContrib).  The fluid enters as the
 * per-cell total mobility and sum(rho*lambda) (MimeticIPEvaluator.hpp:297-311).
 *
 * Inputs per half-face: bctype (ORC_*), bcval (Dirichlet pressure, Neumann
 * outflux or periodic pressure difference).  ppartner from orc_periodic_dof.
 * Outputs: CSR (row_ptr[nf+1], col, val; *nnz_out entries, capacity nnz_cap)
 * sorted by column, rhs[nf], F[nhf], L[nc], g[nc].  *do_reg is set to 1 when no
 * Dirichlet face exists (the S[0][0]*=2 rule of :1398-1400 / :1495-1497 is
 * applied by orc_regularize, called from the solver stage as in the reference).
 * Returns 0, -1 if nnz_cap is too small.
 */
int orc_assemble(int nc, const int *facepos, const int *nb, const int *hf_dof,
                 int nf, const int *ppartner, const double *binv,
                 const double *gflux_static, const double *totmob,
                 const double *rhomob, const double *src, const int *bctype,
                 const double *bcval, int *row_ptr, int *col, double *val,
                 long long nnz_cap, long long *nnz_out, double *rhs, double *F,
                 double *L, double *g, int *do_reg)
{
    tlist T = {0, 0, 0};
    size_t off = 0;
    int c, i;
    long long m, q;
    for (i = 0; i < nf; ++i) rhs[i] = 0.0;
    *do_reg = 1;
    for (c = 0; c < nc; ++c) {
        int f0 = facepos[c], n = facepos[c + 1] - facepos[c], r, j;
        double S[ORC_MAXF * ORC_MAXF], lrhs[ORC_MAXF], gfl[ORC_MAXF],
            cond[ORC_MAXF], Ft[ORC_MAXF], Lc, gsum;
        int ftype[ORC_MAXF], ppart[ORC_MAXF];
        /* setExternalContrib (:1736-1779) */
        g[c] = src[c];
        for (r = 0; r < n; ++r) {
            int h = f0 + r;
            lrhs[r] = 0.0; ftype[r] = FT_INTERNAL; cond[r] = 0.0; ppart[r] = -1;
            if (nb[h] < 0) {
                if (bctype[h] == ORC_DIRICHLET) {
                    ftype[r] = FT_DIRICHLET;
                    cond[r] = bcval[h];
                    *do_reg = 0;
                } else if (bctype[h] == ORC_PERIODIC) {
                    ftype[r] = FT_PERIODIC;
                    cond[r] = bcval[h];
                    ppart[r] = ppartner[hf_dof[h]];
                } else {
                    ftype[r] = FT_NEUMANN;
                    lrhs[r] = bcval[h];
                }
            }
        }
        /* getInverseMatrix + gravityFlux (MimeticIPEvaluator.hpp:336-350, 540-546) */
        for (i = 0; i < n * n; ++i) S[i] = totmob[c] * binv[off + i];
        for (r = 0; r < n; ++r) gfl[r] = gflux_static[f0 + r] * rhomob[c];
        /* buildCellContrib (:1785-1811) */
        for (j = 0; j < n; ++j) {              /* Ft <- S' * ones */
            double s = 0.0;
            for (r = 0; r < n; ++r) s += S[r + j * n];
            Ft[j] = s;
        }
        Lc = 0.0;
        for (j = 0; j < n; ++j) Lc += Ft[j];
        gsum = 0.0;
        for (j = 0; j < n; ++j) gsum += gfl[j];
        L[c] = Lc;
        g[c] -= gsum;
        for (r = 0; r < n; ++r) lrhs[r] = gfl[r] - lrhs[r];
        for (r = 0; r < n; ++r) lrhs[r] = 1.0 * lrhs[r] + (g[c] / Lc) * Ft[r];
        for (j = 0; j < n; ++j)
            for (r = 0; r <= j; ++r)
                S[r + j * n] = -(1.0 / Lc) * Ft[r] * Ft[j] + 1.0 * S[r + j * n];
        for (j = 0; j < n; ++j)
            for (r = j + 1; r < n; ++r) S[r + j * n] = S[j + r * n];
        for (r = 0; r < n; ++r) F[f0 + r] = Ft[r];
        /* addCellContrib (:1817-1901) */
        for (r = 0; r < n; ++r) {
            int ii = hf_dof[f0 + r];
            if (ftype[r] == FT_DIRICHLET) {
                tl_add(&T, ii, ii, S[r + r * n]);     /* S_[ii][ii] = S(r,r) */
                rhs[ii] = S[r + r * n] * cond[r];
                continue;
            }
            if (ftype[r] == FT_PERIODIC) {
                double a = S[r + r * n], b = a * cond[r];
                int pp = ppart[r];
                tl_add(&T, ii, ii, a);
                tl_add(&T, ii, pp, -a);
                rhs[ii] += b;
                tl_add(&T, pp, ii, -a);
                tl_add(&T, pp, pp, a);
                rhs[pp] -= b;
                if (pp < ii) ii = pp;
                /* fall through */
            }
            for (j = 0; j < n; ++j) {
                int jj = hf_dof[f0 + j];
                if (ftype[j] == FT_DIRICHLET) {
                    rhs[ii] -= S[r + j * n] * cond[j];
                    continue;
                }
                if (ftype[j] == FT_PERIODIC) {
                    if (ppart[j] < jj) {
                        rhs[ii] -= S[r + j * n] * cond[j];
                        jj = ppart[j];
                    }
                }
                tl_add(&T, ii, jj, S[r + j * n]);
            }
            rhs[ii] += lrhs[r];
        }
        off += (size_t)n * n;
    }
    /* Compress triplets into CSR (sum duplicates). */
    qsort(T.t, T.n, sizeof(trip), trip_cmp);
    for (i = 0; i <= nf; ++i) row_ptr[i] = 0;
    m = 0;
    for (q = 0; q < (long long)T.n; ++q) {
        if (q > 0 && T.t[q].r == T.t[q - 1].r && T.t[q].c == T.t[q - 1].c) {
            val[m - 1] += T.t[q].v;
            continue;
        }
        if (m >= nnz_cap) { free(T.t); return -1; }
        col[m] = T.t[q].c;
        val[m] = T.t[q].v;
        row_ptr[T.t[q].r + 1]++;
        ++m;
    }
    for (i = 0; i < nf; ++i) row_ptr[i + 1] += row_ptr[i];
    *nnz_out = m;
    free(T.t);
    return 0;
}

/* S[0][0] *= 2 for problems without Dirichlet faces
 * (IncompFlowSolverHybrid.hpp:1398-1400, :1495-1497). */
void orc_regularize(const int *row_ptr, const int *col, double *val)
{
    int k;
    for (k = row_ptr[0]; k < row_ptr[1]; ++k)
        if (col[k] == 0) val[k] *= 2.0;
}

/* ------------------------------------------------------------------------ */
/* Host threads of the CG's SpMV, dot products and vector updates (OpenMP);
 * 1 (the default) keeps the single-threaded order of every sum. */
static int orc_threads = 1;

void orc_set_threads(int n) { orc_threads = n > 1 ? n : 1; }

static void spmv(int n, const int *rp, const int *ci, const double *v,
                 const double *x, double *y)
{
    int i;
#pragma omp parallel for schedule(static) num_threads(orc_threads) if (orc_threads > 1)
    for (i = 0; i < n; ++i) {
        double s = 0.0;
        int k;
        for (k = rp[i]; k < rp[i + 1]; ++k) s += v[k] * x[ci[k]];
        y[i] = s;
    }
}

static double dot(int n, const double *a, const double *b)
{
    double s = 0.0;
    int i;
#pragma omp parallel for schedule(static) reduction(+ : s) num_threads(orc_threads) if (orc_threads > 1)
    for (i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* ILU(0) of a CSR matrix with sorted columns (dune-istl SeqILU0, relaxation
 * 1.0, as constructed at IncompFlowSolverHybrid.hpp:1406).  lu gets the
 * factors in the pattern of A; diag[i] indexes the diagonal entry. */
static int ilu0(int n, const int *rp, const int *ci, const double *a,
                double *lu, int *diag)
{
    int i, k, kk, j;
    int *pos = (int *)malloc(sizeof(int) * (size_t)n);
    memcpy(lu, a, sizeof(double) * (size_t)rp[n]);
    for (i = 0; i < n; ++i) pos[i] = -1;
    for (i = 0; i < n; ++i) {
        diag[i] = -1;
        for (k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] == i) diag[i] = k;
        if (diag[i] < 0) { free(pos); return -1; }
    }
    for (i = 0; i < n; ++i) {
        for (k = rp[i]; k < rp[i + 1]; ++k) pos[ci[k]] = k;
        for (k = rp[i]; k < rp[i + 1] && ci[k] < i; ++k) {
            int r = ci[k];
            double piv = lu[k] / lu[diag[r]];
            lu[k] = piv;
            for (kk = diag[r] + 1; kk < rp[r + 1]; ++kk) {
                j = ci[kk];
                if (pos[j] >= 0) lu[pos[j]] -= piv * lu[kk];
            }
        }
        for (k = rp[i]; k < rp[i + 1]; ++k) pos[ci[k]] = -1;
    }
    free(pos);
    return 0;
}

static void ilu0_apply(int n, const int *rp, const int *ci, const double *lu,
                       const int *diag, const double *b, double *x)
{
    int i, k;
    for (i = 0; i < n; ++i) {
        double s = b[i];
        for (k = rp[i]; k < diag[i]; ++k) s -= lu[k] * x[ci[k]];
        x[i] = s;
    }
    for (i = n - 1; i >= 0; --i) {
        double s = x[i];
        for (k = diag[i] + 1; k < rp[i + 1]; ++k) s -= lu[k] * x[ci[k]];
        x[i] = s / lu[diag[i]];
    }
}

/* Preconditioned CG.  Follows dune-istl CGSolver::apply as invoked from
 * IncompFlowSolverHybrid.hpp:1387-1420 (type 0: SeqILU0) and :1481-1541
 * (initial guess: Dirichlet rows x = b/diag when dirichlet_guess != 0):
 *   r = b - A x; def0 = |r|; p = M^-1 r; rho = p.r
 *   loop: q = A p; lambda = rho/(p.q); x += lambda p; r -= lambda q;
 *         stop when |r|/def0 < tol; z = M^-1 r; beta = (z.r)/rho; p = z + beta p
 * precond: 0 none, 1 Jacobi, 2 ILU(0).
 * out[0] = iterations, out[1] = achieved reduction |r|/def0.
 * Returns 0 on convergence, 1 if maxit was reached, -1 on breakdown. */
int orc_cg(int n, const int *rp, const int *ci, const double *a,
           const double *b, double *x, double tol, int maxit, int precond,
           int dirichlet_guess, double *out)
{
    double *r = (double *)malloc(sizeof(double) * (size_t)n * 4);
    double *p = r + n, *q = r + 2 * (size_t)n, *z = r + 3 * (size_t)n;
    double *lu = 0, *dinv = 0;
    int *diag = 0;
    double def0 = 1.0, def = 0.0, rho, rholast;
    int i, k, it = 0, rc = 1;
    if (precond == 2) {
        lu = (double *)malloc(sizeof(double) * (size_t)rp[n]);
        diag = (int *)malloc(sizeof(int) * (size_t)n);
        if (ilu0(n, rp, ci, a, lu, diag)) { rc = -1; goto done; }
    } else if (precond == 1) {
        dinv = (double *)malloc(sizeof(double) * (size_t)n);
        for (i = 0; i < n; ++i) {
            dinv[i] = 0.0;
            for (k = rp[i]; k < rp[i + 1]; ++k)
                if (ci[k] == i) dinv[i] = 1.0 / a[k];
        }
    }
    for (i = 0; i < n; ++i) x[i] = 0.0;
    if (dirichlet_guess) {                    /* :1521-1532 */
        for (i = 0; i < n; ++i) {
            int isd = 1;
            double d = 0.0;
            for (k = rp[i]; k < rp[i + 1]; ++k) {
                if (ci[k] != i && a[k] != 0.0) isd = 0;
                if (ci[k] == i) d = a[k];
            }
            if (isd) x[i] = b[i] / d;
        }
    }
    spmv(n, rp, ci, a, x, q);
    for (i = 0; i < n; ++i) r[i] = b[i] - q[i];
    def0 = sqrt(dot(n, r, r));
    if (def0 < 1e-30) { rc = 0; def = def0; def0 = 1.0; goto done; }
    if (precond == 2) ilu0_apply(n, rp, ci, lu, diag, r, p);
    else if (precond == 1) for (i = 0; i < n; ++i) p[i] = dinv[i] * r[i];
    else memcpy(p, r, sizeof(double) * (size_t)n);
    rholast = dot(n, p, r);
    for (it = 1; it <= maxit; ++it) {
        double alpha, lambda, beta;
        spmv(n, rp, ci, a, p, q);
        alpha = dot(n, p, q);
        if (!(alpha > 0.0)) { rc = -1; break; }
        lambda = rholast / alpha;
#pragma omp parallel for schedule(static) num_threads(orc_threads) if (orc_threads > 1)
        for (i = 0; i < n; ++i) { x[i] += lambda * p[i]; r[i] -= lambda * q[i]; }
        def = sqrt(dot(n, r, r));
        if (def / def0 < tol || def < 1e-30) { rc = 0; break; }
        if (precond == 2) ilu0_apply(n, rp, ci, lu, diag, r, z);
        else if (precond == 1) {
#pragma omp parallel for schedule(static) num_threads(orc_threads) if (orc_threads > 1)
            for (i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
        } else memcpy(z, r, sizeof(double) * (size_t)n);
        rho = dot(n, z, r);
        beta = rho / rholast;
#pragma omp parallel for schedule(static) num_threads(orc_threads) if (orc_threads > 1)
        for (i = 0; i < n; ++i) p[i] = beta * p[i] + z[i];
        rholast = rho;
    }
    if (it > maxit) it = maxit;
done:
    out[0] = (double)it;
    out[1] = def / def0;
    free(r); free(lu); free(diag); free(dinv);
    return rc;
}

/* y = A x (exposed for residual checks in tests). */
void orc_spmv(int n, const int *rp, const int *ci, const double *a,
              const double *x, double *y)
{
    spmv(n, rp, ci, a, x, y);
}

/* ------------------------------------------------------------------------ */
/* Cell pressure and half-face flux recovery.  Follows
 * IncompFlowSolverHybrid.hpp:1672-1730 (computePressureAndFluxes):
 *   pi_i = soln[l2g_i]; p = (g + F.pi)/L; pi_i <- p - pi_i;
 *   v = (totmob Binv) pi + rhomob gflux_static   (vecMulAdd_N, Matrix.hpp:958-968) */
void orc_recover(int nc, const int *facepos, const int *hf_dof,
                 const double *binv, const double *gflux_static,
                 const double *totmob, const double *rhomob, const double *F,
                 const double *L, const double *g, const double *soln,
                 double *press, double *outflux)
{
    size_t off = 0;
    int c;
    for (c = 0; c < nc; ++c) {
        int f0 = facepos[c], n = facepos[c + 1] - facepos[c], i, j;
        double pi[ORC_MAXF], s = 0.0, p;
        for (i = 0; i < n; ++i) pi[i] = soln[hf_dof[f0 + i]];
        for (i = 0; i < n; ++i) s += F[f0 + i] * pi[i];
        p = (g[c] + s) / L[c];
        press[c] = p;
        for (i = 0; i < n; ++i) pi[i] = p - pi[i];
        for (i = 0; i < n; ++i) {
            double v = 0.0;
            for (j = 0; j < n; ++j) v += totmob[c] * binv[off + i + j * n] * pi[j];
            outflux[f0 + i] = v;
        }
        for (i = 0; i < n; ++i) outflux[f0 + i] = gflux_static[f0 + i] * rhomob[c] + outflux[f0 + i];
        off += (size_t)n * n;
    }
}

/* Twin half-face flux averaging.  Follows IncompFlowSolverHybrid.hpp:707-807
 * (FaceFluxes + postProcessFluxes): internal faces and periodic partners get
 * antisymmetric fluxes +-0.5*(v_first - v_second) where "first" is the
 * half-face met first in cell order; returns the maximum modification. */
double orc_post_process(int nc, const int *facepos, const int *nb,
                        const int *hf_dof, int nf, const int *ppartner,
                        int have_periodic, double *outflux)
{
    double *fl = (double *)calloc((size_t)nf, sizeof(double));
    int *vis = (int *)calloc((size_t)nf, sizeof(int));
    double maxmod = 0.0;
    int c, k, pass;
    for (pass = 0; pass < 2; ++pass) {
        for (k = 0; k < nf; ++k) vis[k] = 0;
        for (c = 0; c < nc; ++c) {
            for (k = facepos[c]; k < facepos[c + 1]; ++k) {
                int f = hf_dof[k];
                int targets[2], nt = 0, t;
                if (nb[k] < 0) {
                    if (!have_periodic) continue;
                    if (ppartner[f] == -1) continue;
                    targets[nt++] = f;
                    targets[nt++] = ppartner[f];
                } else {
                    targets[nt++] = f;
                }
                for (t = 0; t < nt; ++t) {
                    int fi = targets[t];
                    double sign = vis[fi] ? -1.0 : 1.0;
                    if (pass == 0) {
                        fl[fi] += sign * outflux[k];
                    } else {
                        double nv = 0.5 * sign * fl[fi];
                        if (t == 0) {
                            double d = fabs(outflux[k] - nv);
                            if (d > maxmod) maxmod = d;
                            outflux[k] = nv;
                        }
                    }
                    vis[fi]++;
                }
            }
        }
    }
    free(fl);
    free(vis);
    return maxmod;
}

// C-ABI of the MI355X mimetic hybrid solver (include/mimetic_hybrid.h).
//
// Host orchestration only: topology on the host once per initSystemStructure,
// everything per solve on the device.  The solve loop launches chunks of PCG
// iterations (captured once into a hipGraph) and reads back a few bytes of
// per-system state between chunks to decide convergence and, in mixed
// precision, when to replace the FP32 recursive residual by the FP64 true one.
#include "../../include/mimetic_hybrid.h"
#include "../../include/euler_upstream.h"
#include "mhs_amg.h"
#include "mhs_ebe.h"
#include "mhs_kernels.h"
#include "mhs_scratch.h"
#include "mhs_structure.h"
#include "mhs_transport.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <vector>

using namespace mhs;

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// Device buffers of one call, freed when it returns.
struct TempBufs {
    std::vector<DevBuf> v;
    ~TempBufs() {
        for (auto &b : v) (void)hipFree(b.p);
    }
};

struct GraphCache {
    hipGraphExec_t exec = nullptr;
    int precision = -1, nrhs = 0, chunk = 0, amg = -1, nu = -1, ebe = -1;
    int shape[5] = {-1, -1, -1, -1, -1};
    double tol2 = -1.0, rf2 = -1.0, alpha = -1.0;
    double omega[kMaxLevels] = {};
    const void *key = nullptr;
};

// Launch shapes of the row-block kernels picked by the autotuner, per
// (right-hand sides, precision, preconditioner).
struct Shape {
    int spmv_tb = 256, spmv_xcd = 1, sweep_tb = 256, sweep_xcd = 1;
    int tail_rows = 0;
};

}  // namespace

static long long next_generation() {
    static std::atomic<long long> counter{0};
    return ++counter;
}

struct mhs_solver {
    int device = 0;
    hipStream_t stream = nullptr;       // our own stream (default priority)
    hipStream_t user_stream = nullptr;  // caller's stream, ordered against ours
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipEvent_t ev_ph[4] = {};       // phase boundaries of a solve on its stream (report times)
    std::vector<unsigned char> dA_copied;   // the multigrid descriptor last copied to the device
    std::string err;
    // this device's scratch array (mhs_sum_outflux_noalloc); null when the
    // address lookup at creation failed, with the reason in scratch_err
    void *scratch = nullptr;
    hipError_t scratch_err = hipSuccess;

    // grid (host copies)
    bool grid_set = false;
    int ninst = 1, nc = 0, nhf = 0;
    std::vector<int> inst_cellpos, facepos, nb;
    std::vector<double> cvol, ccen, farea, fcen, fnrm;

    // structure
    bool structure_valid = false, ip_valid = false, solved = false;
    Topology T;
    DevTopo dT{};
    DevGeom dG{};
    std::vector<DevBuf> topo_bufs;

    // static inner products
    double *binv = nullptr, *gflux = nullptr;
    double grav[3] = {0.0, 0.0, 0.0};
    double *perm = nullptr;             // K per cell (row-major 3x3), kept for the adjoint and transport
    // generations of the permeability and of the structure, unique in the
    // process (a later solver at the same address never repeats them)
    long long perm_gen = 0;
    long long struct_gen = 0;
    // anisotropic (tensor-mobility) inner products, allocated on the first such solve
    double *an_lam = nullptr, *an_rlam = nullptr, *an_lamK = nullptr, *an_dynKg = nullptr,
           *an_binv = nullptr, *an_gflux = nullptr, *an_ones = nullptr;
    bool last_aniso = false;            // the last solve used tensor mobilities

    // element-by-element level-0 operator (mixed precision): layout built with
    // the structure, values written by the assembly
    EbeLayout EL;
    bool ebe_ok = false;                // the structure allows it
    std::string ebe_why;                // else why not
    DevEbe dE{};
    bool last_ebe = false;              // the last solve used it
    // per nrhs (1..4): whether the autotuner found it faster than the
    // assembled FP32 matrix on this structure (-1: not tried yet)
    int ebe_pref[5] = {-1, -1, -1, -1, -1};
    // whether val64 (and val32) hold the face matrix of the last assembly: a
    // solve that applies the operator element by element and needs no face
    // matrix fills only the vectors; the matrix is filled when asked for
    bool mat_valid = false;
    // multigrid numeric set-up last done for this permeability generation, in
    // mixed precision (FP32 copies written too) or not, with tensor mobilities
    // or not (-1: none since the hierarchy was built)
    long long setup_gen = -1;
    bool setup_mixed = false, setup_aniso = false;

    // multigrid hierarchy
    std::vector<int> ijk;
    bool has_amg = false;
    AmgHierarchy H;
    DevAmg dA{};
    std::vector<DevBuf> amg_bufs;       // structure of the hierarchy on the device
    bool amg_planned = false;           // aggregate shapes chosen from the permeability
    std::vector<double> rho;            // largest eigenvalue of D^-1 A per level (empty: not estimated)
    long long rho_gen = -1;             // permeability generation the estimate was made for
    bool rho_aniso = false;             // ... and whether that solve had tensor mobilities
    double *z64 = nullptr;
    float *z32 = nullptr;

    // per-solve buffers
    int cap_nrhs = 0;
    bool cap_mixed = false;
    std::vector<DevBuf> solve_bufs;
    DevCellSys C{};
    double *val64 = nullptr, *b = nullptr, *dinv64 = nullptr, *x64 = nullptr, *r64 = nullptr,
           *p64 = nullptr, *q64 = nullptr;
    float *val32 = nullptr, *dinv32 = nullptr, *x32 = nullptr, *r32 = nullptr, *p32 = nullptr,
          *q32 = nullptr;
    double *press = nullptr, *outflux = nullptr, *outflux_tmp = nullptr;
    double *in_totmob = nullptr, *in_rhomob = nullptr, *in_src = nullptr, *in_bcval = nullptr;
    uint8_t *ftype = nullptr;
    int *inst_reg = nullptr;
    DevRed R{};
    DevScal S{};
    unsigned long long *maxbits = nullptr;
    // pinned host mirror of the small per-system state: one block laid out as
    // the device block (sc, then act, iters, status), read back by one copy
    double *h_sc = nullptr;
    int *h_act = nullptr, *h_iters = nullptr, *h_upd = nullptr, *h_status = nullptr;
    size_t h_nsys = 0;
    // boundary half-faces of the structure (their instance), and the face
    // types last uploaded for them; internal half-faces are always FT_INTERNAL
    std::vector<int> bnd_hf, bnd_inst;
    std::vector<uint8_t> bnd_per;   // periodic flag of each, from the structure
    long long bnd_gen = -1;
    std::vector<uint8_t> last_bft;
    std::vector<int> last_reg;      // per-instance regularisation flags last uploaded
    int last_nrhs = 0;
    bool last_amg = false;
    long long solve_count = 0;
    std::map<int, GraphCache> graphs;   // captured PCG chunks by iteration count
    std::map<int, Shape> shapes;        // autotuned launch shapes of this structure
    // adjoint work arrays (allocated on the first adjoint call)
    std::vector<DevBuf> adj_bufs;
    double *adj_pi = nullptr, *adj_b = nullptr, *adj_w = nullptr, *adj_y = nullptr, *adj_gK = nullptr,
           *adj_gtm = nullptr, *adj_grm = nullptr, *adj_gsrc = nullptr;

    ~mhs_solver();
    void free_topology();
    void free_solve_buffers();
    void free_amg();
};

namespace {

int fail(mhs_solver *s, int code, const std::string &msg) {
    if (s) s->err = msg;
    return code;
}

#define HIPCHECK(expr)                                                                 \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(s, MHS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Bytes of the per-system state block: sc (SC_N doubles per system), then
// act and iters (one int per system) and status (4 ints).
size_t state_bytes(size_t nsys) { return nsys * SC_N * sizeof(double) + (2 * nsys + 4) * sizeof(int); }

template <class T>
T *dalloc(mhs_solver *s, std::vector<DevBuf> &list, size_t n, hipError_t *e) {
    void *p = nullptr;
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    *e = hipMalloc(&p, bytes);
    if (*e != hipSuccess) return nullptr;
    list.push_back({p, bytes});
    return static_cast<T *>(p);
}

template <class T>
T *dupload(mhs_solver *s, std::vector<DevBuf> &list, const std::vector<T> &v, hipError_t *e) {
    T *p = dalloc<T>(s, list, v.size(), e);
    if (!p) return nullptr;
    if (!v.empty()) *e = hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    return p;
}

}  // namespace

void mhs_solver::free_amg() {
    for (auto &b : amg_bufs) (void)hipFree(b.p);
    amg_bufs.clear();
    amg_planned = false;
    dA = DevAmg{};
    dA_copied.clear();      // a new hierarchy has a new device descriptor
    setup_gen = -1;
    has_amg = false;
    H = AmgHierarchy();
    rho.clear();
    shapes.clear();
    std::fill(ebe_pref, ebe_pref + 5, -1);
}

void mhs_solver::free_topology() {
    free_amg();
    for (auto &b : topo_bufs) (void)hipFree(b.p);
    topo_bufs.clear();
    dE = DevEbe{};
    EL = EbeLayout();
    ebe_ok = false;
    last_ebe = false;
    binv = gflux = perm = nullptr;
    an_lam = an_rlam = an_lamK = an_dynKg = an_binv = an_gflux = an_ones = nullptr;
    dT = DevTopo{};
    shapes.clear();
    std::fill(ebe_pref, ebe_pref + 5, -1);
    dG = DevGeom{};
}

void mhs_solver::free_solve_buffers() {
    for (auto &g : graphs)
        if (g.second.exec) (void)hipGraphExecDestroy(g.second.exec);
    graphs.clear();
    for (auto &b : solve_bufs) (void)hipFree(b.p);
    solve_bufs.clear();
    dE.val64 = nullptr;
    dE.val32 = nullptr;
    dE.R.partials = nullptr;
    for (auto &b : adj_bufs) (void)hipFree(b.p);
    adj_bufs.clear();
    adj_pi = adj_b = adj_w = adj_y = adj_gK = adj_gtm = adj_grm = adj_gsrc = nullptr;
    cap_nrhs = 0;
    cap_mixed = false;
    if (h_sc) hipHostFree(h_sc);  // also holds h_act, h_iters, h_status
    if (h_upd) hipHostFree(h_upd);
    h_sc = nullptr;
    h_act = h_iters = h_upd = h_status = nullptr;
    h_nsys = 0;
    last_bft.clear();
    last_reg.clear();
    mat_valid = false;
}

mhs_solver::~mhs_solver() {
    free_solve_buffers();
    free_topology();
    if (ev_in) hipEventDestroy(ev_in);
    if (ev_out) hipEventDestroy(ev_out);
    for (hipEvent_t e : ev_ph)
        if (e) hipEventDestroy(e);
    if (stream) hipStreamDestroy(stream);
}

extern "C" {

void mhs_default_options(mhs_solve_options *o) {
    if (!o) return;
    o->tol = 1e-8;               // residual_tolerance default (:662)
    o->maxit = 0;
    o->linsolver_type = 1;       // reference default (:664)
    o->verbosity = 0;
    o->same_matrix = 0;
    o->prolongate_factor = 1.6;  // (:667)
    o->smooth_steps = 1;         // (:668)
    o->precision = MHS_FP64;
    o->nrhs = 1;
    o->chunk = 32;
    o->replace_factor = 0.01;
    o->use_graph = 1;
    o->precond = -1;
    o->smoother_weight = 0.0;     // per level from the spectral radius
    o->autotune = 1;
    o->spmv_threads = 0;
    o->sweep_threads = 0;
    o->tail_rows = -1;
    o->element_operator = 1;
}

int mhs_create(mhs_solver **out, int device, void *stream) {
    if (!out) return MHS_ERR_INVALID;
    *out = nullptr;
    mhs_solver *s = new (std::nothrow) mhs_solver();
    if (!s) return MHS_ERR_ALLOC;
    s->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_in, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_out, hipEventDisableTiming);
    for (hipEvent_t &ev : s->ev_ph)
        if (e == hipSuccess) e = hipEventCreate(&ev);
    if (e != hipSuccess) {
        delete s;
        return MHS_ERR_HIP;
    }
    // once per device, with the device current: the address of its scratch array
    s->scratch_err = scratch_address(device, &s->scratch);
    s->user_stream = static_cast<hipStream_t>(stream);
    *out = s;
    return MHS_OK;
}

int mhs_destroy(mhs_solver *s) {
    if (!s) return MHS_OK;
    hipSetDevice(s->device);
    delete s;
    return MHS_OK;
}

const char *mhs_last_error(const mhs_solver *s) { return s ? s->err.c_str() : "null solver"; }

int mhs_clear(mhs_solver *s) {
    if (!s) return MHS_ERR_INVALID;
    hipSetDevice(s->device);
    hipStreamSynchronize(s->stream);
    s->free_solve_buffers();
    s->free_topology();
    s->grid_set = s->structure_valid = s->ip_valid = s->solved = false;
    s->T = Topology();
    s->inst_cellpos.clear();
    s->facepos.clear();
    s->nb.clear();
    s->cvol.clear(); s->ccen.clear(); s->farea.clear(); s->fcen.clear(); s->fnrm.clear();
    s->ijk.clear();
    s->nc = s->nhf = 0;
    s->err.clear();
    return MHS_OK;
}

int mhs_set_grid(mhs_solver *s, int num_instances, const int *inst_cellpos, int num_cells,
                 const int *facepos, const int *neighbour, const double *cell_volume,
                 const double *cell_centroid, const double *face_area,
                 const double *face_centroid, const double *face_normal) {
    if (!s) return MHS_ERR_INVALID;
    if (num_cells < 0 || num_instances < 1 || !facepos || (num_cells > 0 && !neighbour))
        return fail(s, MHS_ERR_INVALID, "set_grid: bad sizes or null arrays");
    if (num_instances > 1 && !inst_cellpos)
        return fail(s, MHS_ERR_INVALID, "set_grid: inst_cellpos required for several instances");
    if (!cell_volume || !cell_centroid || !face_area || !face_centroid || !face_normal)
        return fail(s, MHS_ERR_INVALID, "set_grid: geometry arrays required");
    mhs_clear(s);
    s->ninst = num_instances;
    s->nc = num_cells;
    if (facepos[0] != 0) return fail(s, MHS_ERR_INVALID, "set_grid: facepos[0] must be 0");
    for (int c = 0; c < num_cells; ++c)
        if (facepos[c + 1] < facepos[c]) return fail(s, MHS_ERR_INVALID, "set_grid: facepos not monotone");
    s->nhf = facepos[num_cells];
    s->facepos.assign(facepos, facepos + num_cells + 1);
    s->nb.assign(neighbour, neighbour + s->nhf);
    for (int h = 0; h < s->nhf; ++h)
        if (s->nb[h] >= num_cells) return fail(s, MHS_ERR_INVALID, "set_grid: neighbour out of range");
    if (inst_cellpos) s->inst_cellpos.assign(inst_cellpos, inst_cellpos + num_instances + 1);
    else s->inst_cellpos = {0, num_cells};
    s->cvol.assign(cell_volume, cell_volume + num_cells);
    s->ccen.assign(cell_centroid, cell_centroid + 3 * (size_t)num_cells);
    s->farea.assign(face_area, face_area + s->nhf);
    s->fcen.assign(face_centroid, face_centroid + 3 * (size_t)s->nhf);
    s->fnrm.assign(face_normal, face_normal + 3 * (size_t)s->nhf);
    s->grid_set = true;
    return MHS_OK;
}

int mhs_set_cell_ijk(mhs_solver *s, const int *ijk) {
    if (!s) return MHS_ERR_INVALID;
    if (!s->grid_set) return fail(s, MHS_ERR_STATE, "set_cell_ijk: no grid (call set_grid)");
    if (!ijk) {
        s->ijk.clear();
        return MHS_OK;
    }
    s->ijk.assign(ijk, ijk + 3 * (size_t)s->nc);
    return MHS_OK;
}

}  // extern "C"

namespace {

// (Re)builds the multigrid hierarchy for the logical cell indices and the
// coupling strengths and uploads its structure (device buffers in amg_bufs).
int build_amg_device(mhs_solver *s, const Strength &strength) {
    s->free_amg();
    const Topology &T = s->T;
    hipError_t he = hipSuccess;
    auto &L = s->amg_bufs;
    {
        std::string ea = build_amg(s->H, T, s->ijk, strength);
        if (!ea.empty()) return fail(s, MHS_ERR_INVALID, "multigrid: " + ea);
        if (s->H.nlev() > kMaxLevels) return fail(s, MHS_ERR_INVALID, "multigrid: too many levels");
        DevAmg &A = s->dA;
        A.nlev = s->H.nlev();
        A.ninst = T.ninst;
        for (int l = 1; l < A.nlev && he == hipSuccess; ++l) {
            const AmgLevel &H = s->H.lev[l];
            DevAmgLevel &D = A.lev[l];
            D.n = H.n;
            A.max_rows[l] = 0;
            for (size_t t = 0; t + 1 < H.inst_ptr.size(); ++t)
                A.max_rows[l] = std::max(A.max_rows[l], H.inst_ptr[t + 1] - H.inst_ptr[t]);
            A.max_children[l] = 0;
            for (size_t r = 0; r + 1 < H.child_ptr.size(); ++r)
                A.max_children[l] = std::max(A.max_children[l], H.child_ptr[r + 1] - H.child_ptr[r]);
            A.covers[l] = 0;
            if (l >= 2 && !H.child_ptr.empty() && H.child_ptr.back() == s->H.lev[l - 1].n) {
                std::vector<int> seen(s->H.lev[l - 1].n, 0);
                bool ok = true;
                for (int c : H.child) {
                    if (c < 0 || c >= (int)seen.size() || seen[c]++) {
                        ok = false;
                        break;
                    }
                }
                A.covers[l] = ok ? 1 : 0;
            }
            D.nnz = H.row_ptr.empty() ? 0 : H.row_ptr.back();
            D.nsrc = H.src_ptr.empty() ? 0 : H.src_ptr.back();
            D.inst_ptr = dupload(s, L, H.inst_ptr, &he);
            if (he == hipSuccess) D.row_ptr = dupload(s, L, H.row_ptr, &he);
            if (he == hipSuccess) D.col = dupload(s, L, H.col, &he);
            if (he == hipSuccess) D.diag = dupload(s, L, H.diag, &he);
            if (he == hipSuccess) D.src_ptr = dupload(s, L, H.src_ptr, &he);
            if (he == hipSuccess) D.src = dupload(s, L, H.src, &he);
            if (he == hipSuccess) D.agg = dupload(s, L, H.agg, &he);
            if (he == hipSuccess) D.child_ptr = dupload(s, L, H.child_ptr, &he);
            if (he == hipSuccess) D.child = dupload(s, L, H.child, &he);
            D.sa_in = H.sa_in ? 1 : 0;
            D.sa_out = H.sa_out ? 1 : 0;
            if (H.sa_in && he == hipSuccess) D.src2 = dupload(s, L, H.src2, &he);
            if (H.sa_out) {
                D.p_nnz = int(H.p_col.size());
                D.t_nnz = int(H.t_col.size());
                D.t_npair = H.t_pair_ptr.empty() ? 0 : H.t_pair_ptr.back();
                if (he == hipSuccess) D.p_ptr = dupload(s, L, H.p_ptr, &he);
                if (he == hipSuccess) D.p_col = dupload(s, L, H.p_col, &he);
                if (he == hipSuccess) D.p_own = dupload(s, L, H.p_own, &he);
                if (he == hipSuccess) D.p_src_ptr = dupload(s, L, H.p_src_ptr, &he);
                if (he == hipSuccess) D.p_src = dupload(s, L, H.p_src, &he);
                if (he == hipSuccess) D.pt_ptr = dupload(s, L, H.pt_ptr, &he);
                if (he == hipSuccess) D.pt_col = dupload(s, L, H.pt_col, &he);
                if (he == hipSuccess) D.pt_map = dupload(s, L, H.pt_map, &he);
                if (he == hipSuccess) D.p_tpos = dupload(s, L, H.p_tpos, &he);
                if (he == hipSuccess) D.pt_val64 = dalloc<double>(s, L, (size_t)D.p_nnz, &he);
                if (he == hipSuccess) D.pt_val32 = dalloc<float>(s, L, (size_t)D.p_nnz, &he);
                if (he == hipSuccess) D.t_ptr = dupload(s, L, H.t_ptr, &he);
                if (he == hipSuccess) D.t_col = dupload(s, L, H.t_col, &he);
                if (he == hipSuccess) D.t_pair_ptr = dupload(s, L, H.t_pair_ptr, &he);
                if (he == hipSuccess) D.t_pa = dupload(s, L, H.t_pa, &he);
                if (he == hipSuccess) D.t_pb = dupload(s, L, H.t_pb, &he);
                if (he == hipSuccess) D.p_val64 = dalloc<double>(s, L, (size_t)D.p_nnz, &he);
                if (he == hipSuccess) D.p_val32 = dalloc<float>(s, L, (size_t)D.p_nnz, &he);
                if (he == hipSuccess) D.t_val64 = dalloc<double>(s, L, (size_t)D.t_nnz, &he);
                if (he == hipSuccess) D.t_val32 = dalloc<float>(s, L, (size_t)D.t_nnz, &he);
                if (he == hipSuccess) D.t_p = dupload(s, L, H.t_p, &he);
                if (he == hipSuccess) D.m_val64 = dalloc<double>(s, L, (size_t)D.t_nnz, &he);
                if (he == hipSuccess) D.m_val32 = dalloc<float>(s, L, (size_t)D.t_nnz, &he);
            }
            if (he == hipSuccess) D.val64 = dalloc<double>(s, L, (size_t)D.nnz, &he);
            if (he == hipSuccess) D.val32 = dalloc<float>(s, L, (size_t)D.nnz, &he);
            if (he == hipSuccess) D.dinv64 = dalloc<double>(s, L, (size_t)D.n, &he);
            if (he == hipSuccess) D.dinv32 = dalloc<float>(s, L, (size_t)D.n, &he);
        }
        if (he == hipSuccess) A.ap_slice_ptr = dupload(s, L, s->H.ap.slice_ptr, &he);
        if (he == hipSuccess) A.ap_col = dupload(s, L, s->H.ap.col, &he);
        if (he == hipSuccess) A.ap_src_ptr = dupload(s, L, s->H.ap.src_ptr, &he);
        if (he == hipSuccess) A.ap_src = dupload(s, L, s->H.ap.src, &he);
        A.ap_slots = (long long)s->H.ap.col.size();
        A.ap_nsrc = s->H.ap.src_ptr.empty() ? 0 : s->H.ap.src_ptr.back();
        if (he == hipSuccess) A.ap64 = dalloc<double>(s, L, (size_t)A.ap_slots, &he);
        if (he == hipSuccess) A.ap32 = dalloc<float>(s, L, (size_t)A.ap_slots, &he);
        if (he == hipSuccess) A.coarse_off = dupload(s, L, s->H.coarse_off, &he);
        if (he == hipSuccess) A.dev = dalloc<DevAmg>(s, L, 1, &he);
        A.alpha_sa = 1.0;
        A.omega_p = 2.0 / 3.0;
        if (he == hipSuccess) A.cinv64 = dalloc<double>(s, L, (size_t)s->H.coarse_off.back(), &he);
        if (he == hipSuccess) A.cinv32 = dalloc<float>(s, L, (size_t)s->H.coarse_off.back(), &he);
        if (he != hipSuccess) return fail(s, MHS_ERR_ALLOC, std::string("multigrid allocation: ") + hipGetErrorString(he));
        s->has_amg = true;
    }
    return MHS_OK;
}

// Coupling strength along the logical directions: the mean two-point
// half-transmissibility |n.K (x_f - x_c)| A_f / |x_f - x_c|^2 over the
// half-faces between logical neighbours, grouped by the first of i, j, k in
// which the two cells differ (lateral for faces across a fault).
Strength coupling_strength(const mhs_solver *s, const double *perm) {
    double sum[3] = {0.0, 0.0, 0.0};
    long long cnt[3] = {0, 0, 0};
    for (int c = 0; c < s->nc; ++c)
        for (int h = s->facepos[c]; h < s->facepos[c + 1]; ++h) {
            const int c2 = s->nb[h];
            if (c2 < 0) continue;
            int d = -1;
            for (int e = 0; e < 3 && d < 0; ++e)
                if (s->ijk[3 * (size_t)c + e] != s->ijk[3 * (size_t)c2 + e]) d = e;
            if (d < 0) continue;
            const double *K = perm + 9 * (size_t)c;
            const double *xf = &s->fcen[3 * (size_t)h], *xc = &s->ccen[3 * (size_t)c], *n = &s->fnrm[3 * (size_t)h];
            const double dx[3] = {xf[0] - xc[0], xf[1] - xc[1], xf[2] - xc[2]};
            const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            const double dd = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
            if (!(nn > 0.0) || !(dd > 0.0)) continue;
            double nkd = 0.0;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) nkd += n[a] * K[3 * a + b] * dx[b];
            sum[d] += std::fabs(nkd) / nn * s->farea[h] / dd;
            ++cnt[d];
        }
    Strength st{};
    for (int d = 0; d < 3; ++d) st[d] = cnt[d] ? sum[d] / (double)cnt[d] : 0.0;
    return st;
}

}  // namespace

extern "C" {

int mhs_init_system_structure(mhs_solver *s, const int *hf_bc_type, const int *hf_periodic_partner) {
    if (!s) return MHS_ERR_INVALID;
    if (!s->grid_set) return fail(s, MHS_ERR_STATE, "init_system_structure: no grid (call set_grid)");
    HIPCHECK(hipSetDevice(s->device));
    HIPCHECK(hipStreamSynchronize(s->stream));
    s->free_solve_buffers();
    s->free_topology();
    s->structure_valid = s->ip_valid = s->solved = false;
    // face family of every half-face: the dominant axis of its normal; without
    // periodic faces (the element-by-element operator applies) also a class of
    // the face's lower-numbered cell, so that the faces one colour's cells
    // touch are contiguous rows within each family (coalesced gathers and
    // stores in the colour passes): with logical indices the parities of
    // (i, j, k), which also fix the face's place in the multigrid blocks (equal
    // A P row lengths inside a class: no slice padding), else the colour
    std::vector<int> family(s->nhf, 0);
    bool periodic = false;
    for (int h = 0; h < s->nhf && hf_bc_type; ++h)
        if (s->nb[h] < 0 && hf_bc_type[h] == MHS_BC_PERIODIC) periodic = true;
    std::vector<int> colour;
    int ncol = 1;
    if (!periodic)
        ncol = greedy_colours(s->nc, s->facepos.data(), s->nb.data(), colour, s->ijk.empty() ? nullptr : s->ijk.data());
    const bool by_parity = !s->ijk.empty();
    for (int c = 0; c < s->nc; ++c)
        for (int h = s->facepos[c]; h < s->facepos[c + 1]; ++h) {
            const double *n = &s->fnrm[3 * (size_t)h];
            const double ax = std::fabs(n[0]), ay = std::fabs(n[1]), az = std::fabs(n[2]);
            const int fam = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
            const int lower = s->nb[h] >= 0 ? std::min(c, s->nb[h]) : c;
            int cls = 0;
            if (periodic || ncol > kEbeColours) cls = 0;
            else if (by_parity) {
                const int *q = &s->ijk[3 * (size_t)lower];
                cls = (q[0] & 1) + 2 * (q[1] & 1) + 4 * (q[2] & 1);
            } else cls = colour[lower];
            family[h] = fam * 8 + cls;
        }
    std::string e = build_topology(s->T, s->ninst, s->inst_cellpos.data(), s->nc, s->facepos.data(),
                                   s->nb.data(), hf_bc_type, hf_periodic_partner, family.data());
    if (!e.empty()) return fail(s, MHS_ERR_INVALID, "init_system_structure: " + e);
    Topology &T = s->T;
    hipError_t he = hipSuccess;
    auto &L = s->topo_bufs;
    DevTopo &d = s->dT;
    d.nc = T.nc; d.nhf = T.nhf; d.nfp = T.nfp; d.ninst = T.ninst; d.nslices = T.nslices;
    d.nblk = T.nfp / kRowBlock;
    d.max_ncf = T.max_ncf;
    d.uniform_ncf = 1;
    for (int c = 0; c < T.nc && d.uniform_ncf; ++c)
        if (T.facepos[c + 1] - T.facepos[c] != T.max_ncf) d.uniform_ncf = 0;
    d.max_w = 0;
    for (int sl = 0; sl < T.nslices; ++sl)
        d.max_w = std::max(d.max_w, (int)((T.slice_ptr[sl + 1] - T.slice_ptr[sl]) / kSlice));
#define UP(field, vec)                                   \
    do {                                                 \
        d.field = dupload(s, L, T.vec, &he);             \
        if (he != hipSuccess) return fail(s, MHS_ERR_HIP, "upload " #vec); \
    } while (0)
    UP(facepos, facepos); UP(hf_dof, hf_dof); UP(hf_cell, hf_cell); UP(hf_col, hf_col);
    UP(hf_ppart, hf_ppart); UP(hf_twin, hf_twin); UP(hf_redirect, hf_redirect);
    UP(row_hf, row_hf); UP(dof_hf, dof_hf); UP(ptn, ptn); UP(dslot, dslot); UP(pslot, pslot);
    UP(row_cell, row_cell); UP(row_loc, row_loc);
    UP(slice_ptr, slice_ptr); UP(col, col); UP(slotmap, slotmap); UP(sq_pos, sq_pos);
    UP(block_seg, block_seg); UP(seg_blk, seg_blk); UP(inst_rowbase, inst_rowbase);
    UP(inst_regrow, inst_regrow);
#undef UP
    DevGeom &g = s->dG;
    g.cvol = dupload(s, L, s->cvol, &he);
    if (he == hipSuccess) g.ccen = dupload(s, L, s->ccen, &he);
    if (he == hipSuccess) g.farea = dupload(s, L, s->farea, &he);
    if (he == hipSuccess) g.fcen = dupload(s, L, s->fcen, &he);
    if (he == hipSuccess) g.fnrm = dupload(s, L, s->fnrm, &he);
    // static B^-1 component-major: packed entry q of cell c at q * nc + c
    if (he == hipSuccess) s->binv = dalloc<double>(s, L, (size_t)T.max_ncf * (T.max_ncf + 1) / 2 * T.nc, &he);
    if (he == hipSuccess) s->gflux = dalloc<double>(s, L, (size_t)T.nhf, &he);
    if (he != hipSuccess) return fail(s, MHS_ERR_ALLOC, std::string("device allocation: ") + hipGetErrorString(he));
    s->ebe_why = build_ebe(s->EL, T, s->ijk.empty() ? nullptr : s->ijk.data());
    s->ebe_ok = s->ebe_why.empty() && s->EL.ncolours <= kEbeColours && T.max_ncf <= kMaxFaces;
    if (s->ebe_ok) {
        const EbeLayout &E = s->EL;
        DevEbe &D = s->dE;
        D = DevEbe{};
        D.ncolours = E.ncolours;
        D.nblk = E.nblk;
        for (int k = 0; k <= E.ncolours; ++k) D.colour_blk[k] = E.colour_blk[k];
        for (int k = 0; k < E.ncolours; ++k) D.colour_split[k] = E.colour_split[k];
        D.blk_info = reinterpret_cast<const int4 *>(dupload(s, L, E.blk_info, &he));
        if (he == hipSuccess) D.blk_base = reinterpret_cast<const longlong2 *>(dupload(s, L, E.blk_base, &he));
        if (he == hipSuccess) D.seg_val = dupload(s, L, E.seg_val, &he);
        if (he == hipSuccess) D.cell_seg = dupload(s, L, E.cell_seg, &he);
        if (he == hipSuccess) D.cell_local = dupload(s, L, E.cell_local, &he);
        if (he == hipSuccess) D.dof = dupload(s, L, E.dof, &he);
        if (he == hipSuccess) D.R.seg_blk = dupload(s, L, E.inst_slot, &he);
        D.R.nblk = E.nslots;
        if (he != hipSuccess)
            return fail(s, MHS_ERR_ALLOC, std::string("device allocation: ") + hipGetErrorString(he));
    }
    if (!s->ijk.empty()) {
        const int rc = build_amg_device(s, Strength{1.0, 1.0, 1.0});
        if (rc) return rc;
    }
    s->structure_valid = true;
    s->struct_gen = next_generation();
    return MHS_OK;
}

}  // extern "C"

namespace {

int inner_products_impl(mhs_solver *s, const double *perm, const double *gravity, bool device_perm) {
    if (!s) return MHS_ERR_INVALID;
    if (!s->structure_valid)
        return fail(s, MHS_ERR_STATE, "compute_inner_products: structure not initialised");
    if (!perm) return fail(s, MHS_ERR_INVALID, "compute_inner_products: perm is null");
    HIPCHECK(hipSetDevice(s->device));
    for (int d = 0; d < 3; ++d) s->grav[d] = gravity ? gravity[d] : 0.0;
    const size_t n9 = 9 * (size_t)s->nc;
    if (!s->perm) {
        hipError_t he = hipSuccess;
        s->perm = dalloc<double>(s, s->topo_bufs, n9, &he);
        if (he != hipSuccess) return fail(s, MHS_ERR_ALLOC, std::string("device allocation: ") + hipGetErrorString(he));
    }
    if (device_perm && s->user_stream) {
        HIPCHECK(hipEventRecord(s->ev_in, s->user_stream));
        HIPCHECK(hipStreamWaitEvent(s->stream, s->ev_in, 0));
    }
    s->perm_gen = next_generation();
    hipError_t e = hipMemcpyAsync(s->perm, perm, n9 * sizeof(double),
                                  device_perm ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess)
        e = launch_static_ip(s->stream, s->dT, s->dG, s->T.max_ncf, s->perm, s->grav, s->binv, s->gflux);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return fail(s, MHS_ERR_HIP, std::string("static inner products: ") + hipGetErrorString(e));
    if (s->has_amg && !s->amg_planned) {
        // semi-coarsening, decided on the first permeability after the
        // structure: rebuild the hierarchy when the permeability and the
        // geometry call for other aggregate shapes than the current ones
        std::vector<double> hperm;
        const double *kp = perm;
        if (device_perm) {
            hperm.resize(n9);
            HIPCHECK(hipMemcpy(hperm.data(), s->perm, n9 * sizeof(double), hipMemcpyDeviceToHost));
            kp = hperm.data();
        }
        const Strength st = coupling_strength(s, kp);
        std::vector<int> dims;
        const std::string ed = instance_dims(s->T, s->ijk, dims);
        if (!ed.empty()) return fail(s, MHS_ERR_INVALID, "multigrid: " + ed);
        if (coarsening_plan(st, dims) != s->H.blocks) {
            s->free_solve_buffers();
            const int rc = build_amg_device(s, st);
            if (rc) return rc;
        }
        s->amg_planned = true;
    }
    s->ip_valid = true;
    s->solved = false;
    return MHS_OK;
}

}  // namespace

extern "C" {

int mhs_compute_inner_products(mhs_solver *s, const double *perm, const double *gravity) {
    return inner_products_impl(s, perm, gravity, false);
}

int mhs_compute_inner_products_device(mhs_solver *s, const double *perm, const double *gravity) {
    return inner_products_impl(s, perm, gravity, true);
}

int mhs_assemble_static(mhs_solver *s, const double *perm, const double *gravity) {
    return mhs_compute_inner_products(s, perm, gravity);
}

int mhs_get_inner_products(mhs_solver *s, double *binv, double *gflux) {
    if (!s) return MHS_ERR_INVALID;
    if (!s->ip_valid) return fail(s, MHS_ERR_STATE, "get_inner_products: no inner products");
    HIPCHECK(hipSetDevice(s->device));
    const Topology &T = s->T;
    if (binv) {
        std::vector<double> pk((size_t)T.max_ncf * (T.max_ncf + 1) / 2 * T.nc);
        HIPCHECK(hipMemcpy(pk.data(), s->binv, sizeof(double) * pk.size(), hipMemcpyDeviceToHost));
        for (int c = 0; c < T.nc; ++c) {
            const int n = T.facepos[c + 1] - T.facepos[c];
            double *out = binv + T.sq_pos[c];
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j)
                    out[i + j * n] = pk[(size_t)packed_index(std::min(i, j), std::max(i, j), n) * T.nc + c];
        }
    }
    if (gflux)
        HIPCHECK(hipMemcpy(gflux, s->gflux, sizeof(double) * T.nhf, hipMemcpyDeviceToHost));
    return MHS_OK;
}

int mhs_gravity_term(mhs_solver *s, const double *gravity, const double *omega, double *gterm) {
    if (!s) return MHS_ERR_INVALID;
    if (!s->structure_valid) return fail(s, MHS_ERR_STATE, "gravity_term: structure not initialised");
    if (!gravity || !omega || !gterm) return fail(s, MHS_ERR_INVALID, "gravity_term: null argument");
    HIPCHECK(hipSetDevice(s->device));
    const Topology &T = s->T;
    TempBufs tmp;
    hipError_t he = hipSuccess;
    double *dom = dalloc<double>(s, tmp.v, (size_t)T.nc, &he);
    double *dg = he == hipSuccess ? dalloc<double>(s, tmp.v, (size_t)T.nhf, &he) : nullptr;
    if (he != hipSuccess) return fail(s, MHS_ERR_ALLOC, std::string("gravity_term: ") + hipGetErrorString(he));
    HIPCHECK(hipMemcpyAsync(dom, omega, sizeof(double) * T.nc, hipMemcpyHostToDevice, s->stream));
    HIPCHECK(launch_gravity_term(s->stream, s->dT, s->dG, gravity, dom, dg));
    HIPCHECK(hipMemcpyAsync(gterm, dg, sizeof(double) * T.nhf, hipMemcpyDeviceToHost, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    return MHS_OK;
}

int mhs_topology_is_sane(int num_cells, const int *cell_facepos, const int *neighbour, int *sane) {
    if (!sane || num_cells < 0 || (num_cells > 0 && (!cell_facepos || !neighbour))) return MHS_ERR_INVALID;
    // each cell's internal neighbours, sorted, must be distinct
    int ok = 1;
    std::vector<int> nb;
    for (int c = 0; c < num_cells && ok; ++c) {
        if (cell_facepos[c + 1] < cell_facepos[c]) return MHS_ERR_INVALID;
        nb.clear();
        for (int h = cell_facepos[c]; h < cell_facepos[c + 1]; ++h)
            if (neighbour[h] >= 0) nb.push_back(neighbour[h]);
        std::sort(nb.begin(), nb.end());
        ok = std::adjacent_find(nb.begin(), nb.end()) == nb.end();
    }
    *sane = ok;
    return MHS_OK;
}

}  // extern "C"

namespace {

int ensure_solve_buffers(mhs_solver *s, int nrhs, bool mixed) {
    if (s->cap_nrhs == nrhs && s->cap_mixed == mixed) return MHS_OK;
    s->free_solve_buffers();
    const Topology &T = s->T;
    auto &L = s->solve_bufs;
    hipError_t e = hipSuccess;
    const size_t nfp = T.nfp, nv = nfp * nrhs, nnz = (size_t)T.slice_ptr[T.nslices];
    const size_t nsys = (size_t)T.ninst * nrhs;
    auto ok = [&]() { return e == hipSuccess; };
    // component-major: packed entry q of cell c at q * nc + c (q below the
    // largest cell's packed size), so lanes on consecutive cells coalesce
    s->C.schur = dalloc<double>(s, L, (size_t)T.max_ncf * (T.max_ncf + 1) / 2 * T.nc, &e);
    if (ok()) s->C.rr = dalloc<double>(s, L, (size_t)nrhs * T.max_ncf * T.nc, &e);
    if (ok()) s->C.F = dalloc<double>(s, L, (size_t)T.nhf, &e);
    if (ok()) s->C.L = dalloc<double>(s, L, (size_t)T.nc, &e);
    if (ok()) s->C.g = dalloc<double>(s, L, (size_t)nrhs * T.nc, &e);
    if (ok()) s->val64 = dalloc<double>(s, L, nnz, &e);
    if (ok()) s->b = dalloc<double>(s, L, nv, &e);
    if (ok()) s->dinv64 = dalloc<double>(s, L, nfp, &e);
    if (ok()) s->x64 = dalloc<double>(s, L, nv, &e);
    if (ok()) s->r64 = dalloc<double>(s, L, nv, &e);
    if (!mixed) {
        if (ok()) s->p64 = dalloc<double>(s, L, nv, &e);
        if (ok()) s->q64 = dalloc<double>(s, L, nv, &e);
        s->val32 = nullptr; s->dinv32 = s->x32 = s->r32 = s->p32 = s->q32 = nullptr;
    } else {
        s->p64 = s->q64 = nullptr;
        if (ok()) s->val32 = dalloc<float>(s, L, nnz, &e);
        if (ok()) s->dinv32 = dalloc<float>(s, L, nfp, &e);
        if (ok()) s->x32 = dalloc<float>(s, L, nv, &e);
        if (ok()) s->r32 = dalloc<float>(s, L, nv, &e);
        if (ok()) s->p32 = dalloc<float>(s, L, nv, &e);
        if (ok()) s->q32 = dalloc<float>(s, L, nv, &e);
    }
    if (ok()) s->press = dalloc<double>(s, L, (size_t)nrhs * T.nc, &e);
    if (ok()) s->outflux = dalloc<double>(s, L, (size_t)nrhs * T.nhf, &e);
    if (ok()) s->outflux_tmp = dalloc<double>(s, L, (size_t)nrhs * T.nhf, &e);
    if (ok()) s->in_totmob = dalloc<double>(s, L, (size_t)T.nc, &e);
    if (ok()) s->in_rhomob = dalloc<double>(s, L, (size_t)T.nc, &e);
    if (ok()) s->in_src = dalloc<double>(s, L, (size_t)nrhs * T.nc, &e);
    if (ok()) s->in_bcval = dalloc<double>(s, L, (size_t)nrhs * T.nhf, &e);
    if (ok()) s->ftype = dalloc<uint8_t>(s, L, (size_t)T.nhf, &e);
    if (ok()) s->inst_reg = dalloc<int>(s, L, (size_t)T.ninst, &e);
    if (ok()) s->R.partials = dalloc<double>(s, L, (size_t)(nfp / kRowBlock) * 16, &e);
    if (ok()) s->R.counters = dalloc<int>(s, L, (size_t)T.ninst, &e);
    // sc, act, iters and status share one block so that the host reads the
    // solver's progress with a single copy (read_state)
    unsigned char *state = nullptr;
    if (ok()) state = dalloc<unsigned char>(s, L, state_bytes(nsys), &e);
    if (ok()) {
        s->S.sc = reinterpret_cast<double *>(state);
        s->S.act = reinterpret_cast<int *>(state + nsys * SC_N * sizeof(double));
        s->S.iters = s->S.act + nsys;
        s->S.status = s->S.iters + nsys;
    }
    if (ok()) s->S.upd = dalloc<int>(s, L, nsys, &e);
    if (ok()) s->maxbits = dalloc<unsigned long long>(s, L, 1, &e);
    s->z64 = nullptr;
    s->z32 = nullptr;
    if (s->has_amg) {
        const size_t vs = mixed ? sizeof(float) : sizeof(double);
        DevAmg &A = s->dA;
        auto vec = [&](size_t n) -> void * {
            void *p = nullptr;
            if (ok()) p = dalloc<unsigned char>(s, L, std::max<size_t>(n, 1) * nrhs * vs, &e);
            return p;
        };
        A.x0 = vec(nfp);
        A.res0 = vec(nfp);
        A.xt0 = vec(nfp);
        for (int l = 1; l < A.nlev; ++l) {
            A.lev[l].b = vec(A.lev[l].n);
            A.lev[l].x = vec(A.lev[l].n);
            A.lev[l].res = vec(A.lev[l].n);
            A.lev[l].xt = vec(A.lev[l].n);
        }
        if (mixed) { if (ok()) s->z32 = dalloc<float>(s, L, nv, &e); }
        else { if (ok()) s->z64 = dalloc<double>(s, L, nv, &e); }
    }
    if (mixed && s->ebe_ok) {
        if (ok()) s->dE.val64 = dalloc<double>(s, L, (size_t)s->EL.nval, &e);
        if (ok()) s->dE.val32 = dalloc<float>(s, L, (size_t)s->EL.nval, &e);
        if (ok()) s->dE.R.partials = dalloc<double>(s, L, (size_t)s->EL.nslots * 16, &e);
        // padding rows are never written by the element-by-element passes:
        // their q and multigrid residual stay zero, like their p and r
        if (ok()) e = hipMemset(s->q32, 0, nv * sizeof(float));
        if (ok() && s->has_amg) e = hipMemset(s->dA.res0, 0, nv * sizeof(float));
    }
    if (ok()) e = hipMemset(s->R.counters, 0, sizeof(int) * T.ninst);
    if (ok()) e = hipMemset(s->x64, 0, nv * sizeof(double));
    if (ok()) e = hipHostMalloc(reinterpret_cast<void **>(&s->h_sc), state_bytes(nsys), 0);
    if (ok()) {
        s->h_act = reinterpret_cast<int *>(reinterpret_cast<unsigned char *>(s->h_sc) +
                                           nsys * SC_N * sizeof(double));
        s->h_iters = s->h_act + nsys;
        s->h_status = s->h_iters + nsys;
    }
    if (ok()) e = hipHostMalloc(reinterpret_cast<void **>(&s->h_upd), nsys * sizeof(int), 0);
    if (!ok()) {
        s->free_solve_buffers();
        return fail(s, MHS_ERR_ALLOC, std::string("solve buffers: ") + hipGetErrorString(e));
    }
    s->R.nblk = T.nfp / kRowBlock;
    s->R.block_seg = s->dT.block_seg;
    s->R.seg_blk = s->dT.seg_blk;
    s->h_nsys = nsys;
    s->cap_nrhs = nrhs;
    s->cap_mixed = mixed;
    return MHS_OK;
}

// Launch `chunk` PCG iterations, through a cached hipGraph when enabled.
// The element-by-element operator of the current solve, or null (assembled
// matrix).  Only the FP32 iteration uses it.
const DevEbe *cur_ebe(const mhs_solver *s) { return s->last_ebe ? &s->dE : nullptr; }

template <typename MT, typename VT>
hipError_t one_iteration(mhs_solver *s, int nrhs, bool amg, const MT *val, const VT *dinv, VT *x, VT *r,
                         VT *p, VT *q, VT *z, double tol2) {
    const DevEbe *E = std::is_same<VT, float>::value ? cur_ebe(s) : nullptr;
    if (amg)
        return launch_pcg_iteration_amg<MT, VT>(s->stream, s->dT, s->dA, nrhs, val, dinv, x, r, p, q, z, s->R,
                                                s->S, s->val64, s->b, s->x64, tol2, E, s->r64);
    return launch_pcg_iteration<MT, VT>(s->stream, s->dT, nrhs, val, dinv, x, r, p, q, s->R, s->S, s->val64,
                                        s->b, s->x64, tol2, E, s->r64);
}

template <typename MT, typename VT>
int run_chunk(mhs_solver *s, int nrhs, int chunk, bool use_graph, const MT *val, const VT *dinv,
              VT *x, VT *r, VT *p, VT *q, VT *z, bool amg, int precision, double tol2) {
    if (use_graph) {
        GraphCache &G = s->graphs[chunk];
        const int shape[5] = {s->dT.spmv_tb, s->dT.spmv_xcd, s->dT.sweep_tb, s->dT.sweep_xcd, s->dT.tail_rows};
        const int ebe = cur_ebe(s) ? 1 : 0;
        if (!(G.exec && G.precision == precision && G.nrhs == nrhs && G.chunk == chunk && G.ebe == ebe &&
              G.key == (const void *)val && G.tol2 == tol2 && G.rf2 == s->S.rf2 && G.amg == int(amg) &&
              std::equal(G.omega, G.omega + kMaxLevels, s->dA.omega_l) && G.alpha == s->dA.alpha &&
              G.nu == s->dA.nu &&
              std::equal(shape, shape + 5, G.shape))) {
            if (G.exec) (void)hipGraphExecDestroy(G.exec);
            G = GraphCache{};
            hipGraph_t graph = nullptr;
            HIPCHECK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = hipSuccess;
            for (int it = 0; it < chunk && e == hipSuccess; ++it)
                e = one_iteration<MT, VT>(s, nrhs, amg, val, dinv, x, r, p, q, z, tol2);
            hipError_t e2 = hipStreamEndCapture(s->stream, &graph);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                return fail(s, MHS_ERR_HIP, std::string("graph capture: ") +
                                                hipGetErrorString(e != hipSuccess ? e : e2));
            }
            hipError_t e3 = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e3 != hipSuccess) {
                G = GraphCache{};
                return fail(s, MHS_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e3));
            }
            G.precision = precision;
            G.nrhs = nrhs;
            G.chunk = chunk;
            G.ebe = ebe;
            G.key = val;
            G.tol2 = tol2;
            G.rf2 = s->S.rf2;
            G.amg = int(amg);
            std::copy(s->dA.omega_l, s->dA.omega_l + kMaxLevels, G.omega);
            G.alpha = s->dA.alpha;
            G.nu = s->dA.nu;
            std::copy(shape, shape + 5, G.shape);
        }
        HIPCHECK(hipGraphLaunch(G.exec, s->stream));
    } else {
        for (int it = 0; it < chunk; ++it)
            HIPCHECK((one_iteration<MT, VT>(s, nrhs, amg, val, dinv, x, r, p, q, z, tol2)));
    }
    return MHS_OK;
}

// The face matrix of the last assembly into val64 (and val32), if the solve
// did not fill it.
int ensure_matrix(mhs_solver *s) {
    if (s->mat_valid) return MHS_OK;
    HIPCHECK(launch_fill(s->stream, s->dT, 1, s->C, s->ftype, nullptr, s->inst_reg, 0, s->val64, s->val32, nullptr,
                         nullptr, nullptr, nullptr, nullptr, FILL_MAT));
    // callers may read it with copies outside the solver's stream
    HIPCHECK(hipStreamSynchronize(s->stream));
    s->mat_valid = true;
    return MHS_OK;
}

// Gather lists of the level-1 A P slots over the FP64 element values (built
// on first need, kept with the hierarchy).
int ensure_ap_elements(mhs_solver *s) {
    DevAmg &A = s->dA;
    if (A.ap_esrc_ptr) return MHS_OK;
    if (!s->ebe_ok || A.nlev < 2) return fail(s, MHS_ERR_STATE, "element A P: no element layout or hierarchy");
    std::vector<int> ptr, src;
    const std::string e = ebe_ap_lists(s->EL, s->T, s->H.ap, s->H.lev[1].agg, ptr, src);
    if (!e.empty()) return fail(s, MHS_ERR_INVALID, "element A P: " + e);
    hipError_t he = hipSuccess;
    const int *dp = dupload(s, s->amg_bufs, ptr, &he);
    const int *ds = he == hipSuccess ? dupload(s, s->amg_bufs, src, &he) : nullptr;
    if (he != hipSuccess) return fail(s, MHS_ERR_ALLOC, std::string("element A P: ") + hipGetErrorString(he));
    A.ap_esrc_ptr = dp;
    A.ap_esrc = ds;
    A.ap_ensrc = (long long)src.size();
    return MHS_OK;
}

int read_state(mhs_solver *s, size_t nsys) {
    // sc, act, iters and status: one block on both sides, one copy
    HIPCHECK(hipMemcpyAsync(s->h_sc, s->S.sc, state_bytes(nsys), hipMemcpyDeviceToHost, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    return MHS_OK;
}

// Largest eigenvalue of the symmetric tridiagonal matrix with diagonal a and
// off-diagonal b (b[i] couples i and i+1): bisection on Sturm counts.
double tridiag_max_eig(const std::vector<double> &a, const std::vector<double> &b) {
    const int k = (int)a.size();
    double lo = 1e300, hi = -1e300;
    for (int i = 0; i < k; ++i) {
        const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < k ? std::fabs(b[i]) : 0.0);
        lo = std::min(lo, a[i] - r);
        hi = std::max(hi, a[i] + r);
    }
    auto below = [&](double x) {  // eigenvalues < x
        int c = 0;
        double q = 1.0;
        for (int i = 0; i < k; ++i) {
            q = a[i] - x - (i > 0 ? b[i - 1] * b[i - 1] / q : 0.0);
            if (q == 0.0) q = -1e-300;
            if (q < 0.0) ++c;
        }
        return c;
    };
    for (int it = 0; it < 200 && hi - lo > 1e-12 * std::max(1.0, std::fabs(hi)); ++it) {
        const double mid = 0.5 * (lo + hi);
        if (below(mid) == k) hi = mid; else lo = mid;
    }
    return hi;
}

// rho(D^-1 A) of every smoothed level (0 .. nlev-2): the largest Ritz value
// of 12 Lanczos steps on D^-1/2 A D^-1/2 (FP64, from a fixed pseudo-random
// start), which approaches rho from below.  Done once per hierarchy.
int estimate_rho(mhs_solver *s) {
    constexpr int kSteps = 12;
    if (const int rm = ensure_matrix(s)) return rm;
    const DevAmg &A = s->dA;
    const int nl = A.nlev - 1;
    size_t nmax = (size_t)s->T.nfp;
    for (int l = 1; l < nl; ++l) nmax = std::max(nmax, (size_t)A.lev[l].n);
    std::vector<DevBuf> tmp;
    struct Release {
        std::vector<DevBuf> &t;
        ~Release() { for (auto &b : t) (void)hipFree(b.p); }
    } release{tmp};
    hipError_t he = hipSuccess;
    double *w = dalloc<double>(s, tmp, nmax, &he);
    double *V[2] = {nullptr, nullptr}, *u = nullptr, *y = nullptr, *part = nullptr;
    if (he == hipSuccess) V[0] = dalloc<double>(s, tmp, nmax, &he);
    if (he == hipSuccess) V[1] = dalloc<double>(s, tmp, nmax, &he);
    if (he == hipSuccess) u = dalloc<double>(s, tmp, nmax, &he);
    if (he == hipSuccess) y = dalloc<double>(s, tmp, nmax, &he);
    if (he == hipSuccess) part = dalloc<double>(s, tmp, nmax / 256 + 1, &he);
    if (he != hipSuccess) return fail(s, MHS_ERR_ALLOC, std::string("smoother weights: ") + hipGetErrorString(he));
    std::vector<double> hp(nmax / 256 + 1);
    std::vector<double> rho(nl, 0.0);
    for (int l = 0; l < nl; ++l) {
        const int n = l == 0 ? s->T.nfp : A.lev[l].n;
        const size_t nb = ((size_t)n + 255) / 256;
        if (n == 0) continue;
        auto total = [&](double *out) -> int {
            HIPCHECK(hipMemcpyAsync(hp.data(), part, nb * sizeof(double), hipMemcpyDeviceToHost, s->stream));
            HIPCHECK(hipStreamSynchronize(s->stream));
            double t = 0.0;
            for (size_t i = 0; i < nb; ++i) t += hp[i];
            *out = t;
            return MHS_OK;
        };
        double ww = 0.0;
        HIPCHECK(launch_lanczos_init(s->stream, n, w, part));
        if (int rc = total(&ww)) return rc;
        double beta = std::sqrt(ww);
        std::vector<double> al, be;
        for (int j = 0; j < kSteps && beta > 0.0; ++j) {
            HIPCHECK(launch_lanczos_step(s->stream, s->dT, A, l, s->val64, s->dinv64, w, 1.0 / beta, V[j & 1], u,
                                         y, part));
            double a = 0.0;
            if (int rc = total(&a)) return rc;
            al.push_back(a);
            HIPCHECK(launch_lanczos_update(s->stream, n, y, V[j & 1], j ? V[(j + 1) & 1] : nullptr, a, beta, w,
                                           part));
            if (int rc = total(&ww)) return rc;
            const double bn = std::sqrt(ww);
            if (!(bn > 1e-12 * std::fabs(a))) break;   // invariant subspace found
            if (j + 1 < kSteps) be.push_back(bn);
            beta = bn;
        }
        be.resize(al.empty() ? 0 : al.size() - 1);
        rho[l] = al.empty() ? 0.0 : tridiag_max_eig(al, be);
    }
    s->rho = rho;
    return MHS_OK;
}

// Autotuner: time every candidate launch shape of the PCG SpMV and of the
// level-0 smoothing sweep on the assembled matrix (HIP events on the solver's
// stream, 2 untimed + 10 timed launches each) and keep the fastest.  Done
// once per structure and (nrhs, precision, preconditioner); the choice only
// changes the order of floating-point partial sums.
int autotune(mhs_solver *s, int nrhs, bool mixed, bool amg, bool choose_ebe, double *tune_ms) {
    const DevEbe *E = mixed ? cur_ebe(s) : nullptr;
    auto key_of = [&]() { return (nrhs * 4 + (mixed ? 2 : 0) + (amg ? 1 : 0)) * 2 + (E ? 1 : 0); };
    auto it = s->shapes.find(key_of());
    // first solve of this nrhs with the element-by-element operator allowed
    // but not forced: time it against the assembled FP32 matrix and keep the
    // faster (many colours or irregular cells can make it the slower one)
    const bool trial = E && choose_ebe && s->ebe_pref[nrhs] < 0;
    *tune_ms = 0.0;
    if (it == s->shapes.end()) {
        auto t0 = std::chrono::steady_clock::now();
        const Topology &T = s->T;
        const size_t nsys = (size_t)T.ninst * nrhs;
        std::vector<int> ones(nsys, 1);  // every system active: no workgroup returns early
        HIPCHECK(hipMemcpyAsync(s->S.act, ones.data(), nsys * sizeof(int), hipMemcpyHostToDevice, s->stream));
        hipEvent_t e0, e1;
        HIPCHECK(hipEventCreate(&e0));
        HIPCHECK(hipEventCreate(&e1));
        hipError_t err = hipSuccess;
        // Each candidate runs as the solve runs it: ten launches captured in a
        // hipGraph, replayed four times; the fastest of the last three
        // replays counts.  Timing plain launches instead favours shapes with
        // fewer kernels, since the host's launch cost then shows in the time.
        auto time_it = [&](auto launch) -> float {
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            if (err == hipSuccess) err = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
            if (err == hipSuccess) {
                hipError_t el = hipSuccess;
                for (int r = 0; r < 10 && el == hipSuccess; ++r) el = launch();
                const hipError_t ee = hipStreamEndCapture(s->stream, &graph);
                err = el != hipSuccess ? el : ee;
            }
            if (err == hipSuccess) err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (graph) (void)hipGraphDestroy(graph);
            float best_ms = 1e30f;
            for (int rep = 0; rep < 4 && err == hipSuccess; ++rep) {
                err = hipEventRecord(e0, s->stream);
                if (err == hipSuccess) err = hipGraphLaunch(exec, s->stream);
                if (err == hipSuccess) err = hipEventRecord(e1, s->stream);
                if (err == hipSuccess) err = hipEventSynchronize(e1);
                float ms = 0.0f;
                if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
                if (err == hipSuccess && rep > 0) best_ms = std::min(best_ms, ms / 10.0f);
            }
            if (exec) (void)hipGraphExecDestroy(exec);
            return err == hipSuccess ? best_ms : 1e30f;
        };
        Shape best;
        float bs = 1e30f, bw = 1e30f;
        // the assembled-matrix SpMV (not run when the operator is applied cell by cell)
        for (int tb : {256, 512, 1024})
            for (int xcd : {1, 0}) {
                if (E && !trial) break;
                s->dT.spmv_tb = tb;
                s->dT.spmv_xcd = xcd;
                const float ms = time_it([&]() {
                    return mixed ? launch_spmv<float, float>(s->stream, s->dT, nrhs, s->val32, s->p32, s->q32,
                                                             s->R, s->S)
                                 : launch_spmv<double, double>(s->stream, s->dT, nrhs, s->val64, s->p64,
                                                               s->q64, s->R, s->S);
                });
                if (ms < bs) { bs = ms; best.spmv_tb = tb; best.spmv_xcd = xcd; }
            }
        if (trial) {
            const float te = time_it([&]() {
                return launch_ebe_spmv(s->stream, s->T.max_ncf, nrhs, *E, s->p32, s->q32, s->S);
            });
            const bool use = te < bs;
            s->ebe_pref[nrhs] = use ? 1 : 0;
            if (!use) {      // the rest of the tuning and the solve use the assembled matrix
                s->last_ebe = false;
                E = nullptr;
            }
        }
        if (amg)
            for (int tb : {256, 512, 1024})
                for (int xcd : {1, 0}) {
                    s->dT.sweep_tb = tb;
                    s->dT.sweep_xcd = xcd;
                    const bool fused = s->dA.nu == 1;   // the sweep the V-cycle actually runs
                    const float ms = time_it([&]() {
                        if (fused)
                            return mixed ? launch_post0<float>(s->stream, s->dT, s->dA, nrhs, s->dinv32, s->r32,
                                                               s->z32, s->R)
                                         : launch_post0<double>(s->stream, s->dT, s->dA, nrhs, s->dinv64, s->r64,
                                                                s->z64, s->R);
                        return mixed ? launch_sweep0<float>(s->stream, s->dT, nrhs, s->val32, s->dinv32, s->r32,
                                                            (const float *)s->dA.x0, s->z32, s->dA.omega_l[0], s->R)
                                     : launch_sweep0<double>(s->stream, s->dT, nrhs, s->val64, s->dinv64, s->r64,
                                                             (const double *)s->dA.x0, s->z64, s->dA.omega_l[0],
                                                             s->R);
                    });
                    if (ms < bw) { bw = ms; best.sweep_tb = tb; best.sweep_xcd = xcd; }
                }
        if (amg && s->dA.nu == 1) {
            // levels below this many rows per instance run in the one-workgroup tail
            s->dT.sweep_tb = best.sweep_tb;
            s->dT.sweep_xcd = best.sweep_xcd;
            float bt = 1e30f;
            for (int rows : {0, 128, 512, 2048, 8192, 32768}) {
                s->dT.tail_rows = rows;
                const float ms = time_it([&]() {
                    return mixed ? launch_vcycle<float>(s->stream, s->dT, s->dA, nrhs, s->val32, s->dinv32, s->r32,
                                                        s->z32, s->R, E)
                                 : launch_vcycle<double>(s->stream, s->dT, s->dA, nrhs, s->val64, s->dinv64,
                                                         s->r64, s->z64, s->R);
                });
                if (ms < bt) { bt = ms; best.tail_rows = rows; }
                if (ms > 1.5f * bt) break;   // past the best, larger tails only take longer
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (err != hipSuccess) return fail(s, MHS_ERR_HIP, std::string("autotune: ") + hipGetErrorString(err));
        it = s->shapes.emplace(key_of(), best).first;
        *tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    s->dT.spmv_tb = it->second.spmv_tb;
    s->dT.spmv_xcd = it->second.spmv_xcd;
    s->dT.sweep_tb = it->second.sweep_tb;
    s->dT.sweep_xcd = it->second.sweep_xcd;
    s->dT.tail_rows = it->second.tail_rows;
    return MHS_OK;
}

struct LinearResult {
    int iterations = 0, replacements = 0;
    bool converged = false;
    double reduction = 0.0;
};

// PCG iterations the slowest active system still needs, from its mean
// residual reduction per iteration so far (the mixed iteration needs one more
// for the FP64 replacement that confirms convergence); 0 while unknown.
int remaining_iterations(const mhs_solver *s, size_t nsys, double tol, bool mixed) {
    int need = 0;
    for (size_t i = 0; i < nsys; ++i) {
        if (!s->h_act[i]) continue;
        const double *sc = s->h_sc + i * SC_N;
        const int it = s->h_iters[i];
        if (it < 8 || !(sc[SC_RR0] > 0.0)) return 0;
        const double q = std::sqrt(sc[SC_RR] / sc[SC_RR0]);
        int r = 1;
        if (q > tol) {
            const double rate = std::pow(q, 1.0 / it);
            if (!(rate < 1.0)) return 0;
            r = (int)std::ceil(std::log(tol / q) / std::log(rate));
        }
        need = std::max(need, r + (mixed ? 1 : 0));
    }
    return need;
}

// PCG on the assembled face system from the right-hand side in s->b, the
// initial iterate in s->x64 and its residual b - A x64 in s->r64 (the fill and
// the adjoint right-hand side write all three), to opt.tol relative FP64
// residual (Jacobi or multigrid preconditioner, FP64 or FP32 with FP64
// residual replacement).
int linear_solve(mhs_solver *s, const mhs_solve_options &opt, int nrhs, bool mixed, bool amg,
                 LinearResult &out) {
    const Topology &T = s->T;
    const size_t nsys = (size_t)T.ninst * nrhs;
    const int chunk = std::max(1, std::min(opt.chunk, 4096));
    const int maxit = opt.maxit > 0 ? opt.maxit : std::max(1, T.nf_total);
    const double tol2 = opt.tol * opt.tol;
    s->S.mixed = mixed ? 1 : 0;
    s->S.amg = amg ? 1 : 0;
    s->S.rf2 = opt.replace_factor * opt.replace_factor;
    HIPCHECK(launch_clear_state(s->stream, s->S, (int)nsys, mixed ? s->x32 : nullptr, (long long)T.nfp * nrhs));
    if (!mixed) {
        HIPCHECK(launch_pcg_start64(s->stream, s->dT, nrhs, s->val64, s->dinv64, s->b, s->x64, s->r64,
                                    s->p64, tol2, s->R, s->S, true));
    } else {
        HIPCHECK(launch_reliable(s->stream, s->dT, nrhs, s->val64, s->b, s->x64, s->dinv32, s->r32,
                                 s->p32, 1, tol2, s->R, s->S, s->r64));
    }
    if (amg) {
        if (!mixed)
            HIPCHECK(launch_amg_start<double>(s->stream, s->dT, s->dA, nrhs, s->val64, s->dinv64, s->r64,
                                              s->z64, s->p64, s->R, s->S, nullptr, s->x64));
        else
            HIPCHECK(launch_amg_start<float>(s->stream, s->dT, s->dA, nrhs, s->val32, s->dinv32, s->r32,
                                             s->z32, s->p32, s->R, s->S, nullptr, s->x32, cur_ebe(s)));
    }
    // The first chunk follows the start without a host round trip (its
    // kernels skip the systems the start found converged); the state is read
    // after it.
    for (size_t i = 0; i < nsys; ++i) {
        s->h_act[i] = 1;
        s->h_iters[i] = 0;
    }
    s->h_status[0] = 1;
    s->h_status[1] = 0;
    int rc = MHS_OK;
    while (s->h_status[0] > 0) {
        int mx = 0;
        for (size_t i = 0; i < nsys; ++i) mx = std::max(mx, s->h_iters[i]);
        if (mx >= maxit) break;
        // after the first chunk, launch as many iterations as the slowest
        // system still needs (no iterations after convergence), as graphs of
        // power-of-two lengths back to back (one cached graph per length) with
        // one state read after them
        int n = std::min(chunk, maxit - mx);
        const int need = remaining_iterations(s, nsys, opt.tol, mixed);
        if (need > 0) n = std::min(n, need);
        while (n > 0) {
            int len = n;
            if (n != chunk) {
                len = 1;
                while (2 * len <= n) len *= 2;
            }
            const bool graph = opt.use_graph;
            if (!mixed)
                rc = run_chunk<double, double>(s, nrhs, len, graph, s->val64, s->dinv64, s->x64, s->r64, s->p64,
                                               s->q64, s->z64, amg, MHS_FP64, tol2);
            else
                rc = run_chunk<float, float>(s, nrhs, len, graph, s->val32, s->dinv32, s->x32, s->r32, s->p32,
                                             s->q32, s->z32, amg, MHS_MIXED, tol2);
            if (rc) return rc;
            n -= len;
        }
        rc = read_state(s, nsys);
        if (rc) return rc;
        if (opt.verbosity >= 2)
            for (size_t i = 0; i < std::min<size_t>(nsys, 4); ++i) {
                const double *sc = s->h_sc + i * SC_N;
                std::fprintf(stderr, "[mhs] sys %zu it %d act %d |r|/|r0| %.3e last FP64 %.3e\n", i,
                             s->h_iters[i], s->h_act[i], std::sqrt(sc[SC_RR] / sc[SC_RR0]),
                             std::sqrt(sc[SC_RRT] / sc[SC_RR0]));
            }
        if (s->h_status[1] > 0)
            return fail(s, MHS_ERR_BREAKDOWN, "PCG breakdown: face operator not positive definite");
    }
    out.converged = s->h_status[0] == 0;
    out.iterations = 0;
    for (size_t i = 0; i < nsys; ++i) out.iterations = std::max(out.iterations, s->h_iters[i]);
    if (mixed) HIPCHECK(launch_merge_all(s->stream, s->dT, nrhs, s->x64, s->x32));
    // A converged FP32 iteration ended on an FP64 residual replacement of the
    // final x64 (a system stops only there), whose |b - A x64| is already in
    // the state; otherwise the true residual is formed once more.
    if (!(mixed && out.converged)) {
        const int rm = ensure_matrix(s);
        if (rm) return rm;
        HIPCHECK(launch_resnorm(s->stream, s->dT, nrhs, s->val64, s->b, s->x64, s->R, s->S));
        rc = read_state(s, nsys);
        if (rc) return rc;
    }
    out.replacements = s->h_status[2];
    double red = 0.0;
    for (size_t i = 0; i < nsys; ++i) {
        const double *sc = s->h_sc + i * SC_N;
        if (sc[SC_RR0] > 0) red = std::max(red, std::sqrt(sc[SC_RRT] / sc[SC_RR0]));
    }
    out.reduction = red;
    return MHS_OK;
}

int solve_impl(mhs_solver *s, const double *totmob, const double *rhomob, const double *src,
               const int *hf_bc_type, const double *bcval, const mhs_solve_options *opt_in,
               mhs_solve_report *rep, bool device_inputs, const double *binv_use = nullptr,
               const double *gflux_use = nullptr) {
    if (!s) return MHS_ERR_INVALID;
    mhs_solve_options opt;
    mhs_default_options(&opt);
    if (opt_in) opt = *opt_in;
    if (!s->structure_valid || !s->ip_valid)
        return fail(s, MHS_ERR_STATE, "solve: init() (structure and inner products) must come first");
    if (opt.linsolver_type < 0 || opt.linsolver_type > 3)
        return fail(s, MHS_ERR_INVALID, "Unknown linsolver_type: " + std::to_string(opt.linsolver_type));
    if (opt.nrhs < 1 || opt.nrhs > 4) return fail(s, MHS_ERR_INVALID, "solve: nrhs must be 1..4");
    if (!(opt.tol > 0.0)) return fail(s, MHS_ERR_INVALID, "solve: tol must be positive");
    if (opt.precision != MHS_FP64 && opt.precision != MHS_MIXED)
        return fail(s, MHS_ERR_INVALID, "solve: unknown precision");
    if (!totmob || !rhomob || !src || !hf_bc_type || !bcval)
        return fail(s, MHS_ERR_INVALID, "solve: null input array");
    const int nrhs = opt.nrhs;
    const bool mixed = opt.precision == MHS_MIXED;
    HIPCHECK(hipSetDevice(s->device));
    const Topology &T = s->T;
    int rc = ensure_solve_buffers(s, nrhs, mixed);
    if (rc != MHS_OK) return rc;

    // Face types (FaceType, :378) and per-instance regularisation flags.
    // Only boundary half-faces can change type between solves: their list is
    // built once per structure, and the device types are rewritten only when
    // a boundary type changed.  Once types are on the device, the scan of the
    // caller's array (scattered reads of a large host array) runs while the
    // GPU assembles with the previous types; a change repeats the assembly.
    if (s->bnd_gen != s->struct_gen) {
        s->bnd_hf.clear();
        s->bnd_inst.clear();
        s->bnd_per.clear();
        int t = 0;
        for (int c = 0; c < T.nc; ++c) {
            while (c >= T.inst_cellpos[t + 1]) ++t;
            for (int h = T.facepos[c]; h < T.facepos[c + 1]; ++h)
                if (T.nb[h] < 0) {
                    s->bnd_hf.push_back(h);
                    s->bnd_inst.push_back(t);
                    s->bnd_per.push_back(T.hf_periodic[h]);
                }
        }
        s->bnd_gen = s->struct_gen;
        s->last_bft.clear();
        s->last_reg.clear();
    }
    const size_t nbnd = s->bnd_hf.size();
    std::vector<uint8_t> bft(nbnd);
    std::vector<int> reg(T.ninst, 1);
    auto scan = [&]() -> int {
        std::fill(reg.begin(), reg.end(), 1);
        // the caller's array is read at scattered positions: keep a few dozen
        // loads in flight
        constexpr size_t kAhead = 32;
        for (size_t e = 0; e < nbnd; ++e) {
            if (e + kAhead < nbnd) __builtin_prefetch(hf_bc_type + s->bnd_hf[e + kAhead]);
            const int h = s->bnd_hf[e];
            const int bt = hf_bc_type[h];
            if (bt == MHS_BC_DIRICHLET) { bft[e] = FT_DIRICHLET; reg[s->bnd_inst[e]] = 0; }
            else if (bt == MHS_BC_PERIODIC) bft[e] = FT_PERIODIC;
            else if (bt == MHS_BC_NEUMANN) bft[e] = FT_NEUMANN;
            else return fail(s, MHS_ERR_INVALID, "solve: unknown boundary condition type");
            if ((bt == MHS_BC_PERIODIC) != (s->bnd_per[e] != 0))
                return fail(s, MHS_ERR_INVALID, "solve: periodic faces differ from those given to init_system_structure");
        }
        return MHS_OK;
    };
    auto upload_types = [&]() -> int {
        std::vector<uint8_t> ft(T.nhf, FT_INTERNAL);
        for (size_t e = 0; e < nbnd; ++e) ft[s->bnd_hf[e]] = bft[e];
        HIPCHECK(hipMemcpyAsync(s->ftype, ft.data(), T.nhf, hipMemcpyHostToDevice, s->stream));
        HIPCHECK(hipMemcpyAsync(s->inst_reg, reg.data(), sizeof(int) * T.ninst, hipMemcpyHostToDevice, s->stream));
        HIPCHECK(hipStreamSynchronize(s->stream));   // the host vectors end here
        s->last_bft = bft;
        s->last_reg = reg;
        return MHS_OK;
    };
    const bool speculate = !s->last_bft.empty() && s->last_reg.size() == (size_t)T.ninst;
    if (speculate) {
        reg = s->last_reg;
    } else {
        rc = scan();
        if (rc) return rc;
    }
    if (s->user_stream) {
        HIPCHECK(hipEventRecord(s->ev_in, s->user_stream));
        HIPCHECK(hipStreamWaitEvent(s->stream, s->ev_in, 0));
    }
    // the report's phase times come from events on the solver's stream: no
    // host synchronisation between the phases
    HIPCHECK(hipEventRecord(s->ev_ph[0], s->stream));
    if (!speculate) {
        rc = upload_types();
        if (rc) return rc;
    }
    const double *d_tm = totmob, *d_rm = rhomob, *d_src = src, *d_bv = bcval;
    if (!device_inputs) {
        auto kind = hipMemcpyHostToDevice;
        HIPCHECK(hipMemcpyAsync(s->in_totmob, totmob, sizeof(double) * T.nc, kind, s->stream));
        HIPCHECK(hipMemcpyAsync(s->in_rhomob, rhomob, sizeof(double) * T.nc, kind, s->stream));
        HIPCHECK(hipMemcpyAsync(s->in_src, src, sizeof(double) * T.nc * nrhs, kind, s->stream));
        HIPCHECK(hipMemcpyAsync(s->in_bcval, bcval, sizeof(double) * T.nhf * nrhs, kind, s->stream));
        d_tm = s->in_totmob; d_rm = s->in_rhomob; d_src = s->in_src; d_bv = s->in_bcval;
    } else {
        // the adjoint needs the mobilities after the caller's buffers may be gone
        auto kind = hipMemcpyDeviceToDevice;
        HIPCHECK(hipMemcpyAsync(s->in_totmob, totmob, sizeof(double) * T.nc, kind, s->stream));
        HIPCHECK(hipMemcpyAsync(s->in_rhomob, rhomob, sizeof(double) * T.nc, kind, s->stream));
        d_tm = s->in_totmob; d_rm = s->in_rhomob;
    }
    // --- assembly -----------------------------------------------------------
    // (it overwrites the state the solution accessors read: no solution until
    // this solve has recovered one)
    s->solved = false;
    s->last_aniso = binv_use != nullptr;
    const double *binv = binv_use ? binv_use : s->binv, *gflux = gflux_use ? gflux_use : s->gflux;
    const int dguess = opt.linsolver_type != 0;  // AMG types start from the Dirichlet guess (:1521-1532)
    float *fill32 = nullptr;
    bool filled_matrix = true;
    auto assemble = [&]() -> int {
        // the level-0 operator of the FP32 iteration: cell by cell when the
        // structure allows it and no instance is regularised (the doubled
        // entry of :1398-1400 is not a cell contribution); element_operator 2
        // always, 1 when the autotuner found it the faster operator for this
        // nrhs on this structure (not yet tried: try it)
        s->last_ebe = mixed && s->ebe_ok && opt.element_operator != 0 &&
                      std::none_of(reg.begin(), reg.end(), [](int v) { return v != 0; }) &&
                      (opt.element_operator == 2 || !opt.autotune || s->ebe_pref[nrhs] != 0);
        HIPCHECK(launch_assemble(s->stream, s->dT, T.max_ncf, nrhs, binv, gflux, d_tm, d_rm, d_src,
                                 s->ftype, d_bv, s->C, cur_ebe(s)));
        // The FP32 face matrix is not written when the FP32 iteration applies
        // the operator cell by cell for certain (no trial against the
        // assembled matrix due, one smoothing sweep: more sweeps run on the
        // assembled matrix).
        const bool ebe_settled =
            s->last_ebe && (opt.element_operator == 2 || !opt.autotune || s->ebe_pref[nrhs] == 1);
        fill32 = (mixed && !(ebe_settled && opt.smooth_steps <= 1)) ? s->val32 : nullptr;
        // No face matrix at all when nothing of this solve reads it: the
        // operator is applied element by element, the smoother weights are
        // known (no Lanczos estimate) and the multigrid's A P is summed from
        // the element values.  It is filled when asked for afterwards.
        const bool will_amg =
            s->has_amg && (opt.precond == 1 || (opt.precond == -1 && opt.linsolver_type != 0));
        const bool rho_known = opt.smoother_weight > 0.0 ||
                               (!s->rho.empty() && s->rho_gen == s->perm_gen &&
                                s->rho_aniso == (binv_use != nullptr));
        const bool no_matrix = mixed && ebe_settled && !fill32 && (!will_amg || rho_known);
        if (no_matrix && will_amg) {
            const int ra = ensure_ap_elements(s);
            if (ra) return ra;
        }
        HIPCHECK(launch_fill(s->stream, s->dT, nrhs, s->C, s->ftype, d_bv, s->inst_reg, dguess, s->val64,
                             fill32, s->b, s->dinv64, s->dinv32, s->x64, s->r64, no_matrix ? FILL_VEC : FILL_ALL));
        s->mat_valid = filled_matrix = !no_matrix;
        return MHS_OK;
    };
    rc = assemble();
    if (rc) return rc;
    if (speculate) {
        rc = scan();
        if (rc) {
            (void)hipStreamSynchronize(s->stream);
            return rc;
        }
        if (bft != s->last_bft) {   // the types changed: assemble again with them
            rc = upload_types();
            if (rc == MHS_OK) rc = assemble();
            if (rc) return rc;
        }
    }
    HIPCHECK(hipEventRecord(s->ev_ph[1], s->stream));

    // --- preconditioner ---------------------------------------------------------
    // The smoother weights come from rho(D^-1 A) of an earlier solve of this
    // hierarchy.  They are estimated again for a new permeability or inner
    // product, and, should a multigrid solve whose estimate came from another
    // solve fail (changed mobilities can raise rho past the 1.5 rho_est at
    // which the weighted Jacobi smoother stops converging), once more for this
    // system before the solve is repeated from its initial guess.
    if (s->rho_gen != s->perm_gen || s->rho_aniso != (binv_use != nullptr)) s->rho.clear();
    bool amg = false;
    if (opt.precond == 1) {
        if (!s->has_amg)
            return fail(s, MHS_ERR_INVALID, "solve: multigrid needs logical cell indices (mhs_set_cell_ijk)");
        amg = true;
    } else if (opt.precond == -1) {
        amg = s->has_amg && opt.linsolver_type != 0;
    } else if (opt.precond != 0) {
        return fail(s, MHS_ERR_INVALID, "solve: unknown preconditioner");
    }
    bool rho_fresh = false;   // the weights were estimated on this very system
    bool mg_reused = false;   // same_matrix kept the multigrid set-up of the previous solve
    double tune_ms = 0.0;
    LinearResult lr;
    for (int attempt = 0;; ++attempt) {
        if (attempt > 0) {
            // repeat with weights from this system: the assembled values and the
            // initial guess are written again (the failed attempt changed x64)
            s->rho.clear();
            HIPCHECK(launch_fill(s->stream, s->dT, nrhs, s->C, s->ftype, d_bv, s->inst_reg, dguess, s->val64,
                                 fill32, s->b, s->dinv64, s->dinv32, s->x64, s->r64));
            s->mat_valid = filled_matrix = true;
        }
        if (amg) {
            s->dA.alpha = opt.prolongate_factor;
            s->dA.nu = std::max(1, opt.smooth_steps);
            // same_matrix (the caller's promise, :1491): the coarse levels,
            // the smoothed transfer and the coarsest inverses of the previous
            // solve are kept, as the reference keeps its AMG hierarchy; A P,
            // the face level's part of the cycle, follows the new assembly
            mg_reused = opt.same_matrix && attempt == 0 && s->setup_gen == s->perm_gen &&
                        (s->setup_mixed || !mixed) && s->setup_aniso == (binv_use != nullptr);
            // A P from the element values when no face matrix was filled
            HIPCHECK(launch_amg_setup(s->stream, s->dT, s->dA, s->val64, mixed, s->mat_valid ? nullptr : s->dE.val64,
                                      mg_reused));
            if (!mg_reused) {
                s->setup_gen = s->perm_gen;
                s->setup_mixed = mixed;
                s->setup_aniso = binv_use != nullptr;
            }
            // Jacobi weights: the caller's, or 4 / (3 rho(D^-1 A)) per level with
            // rho estimated on the first multigrid solve of this hierarchy
            if (opt.smoother_weight > 0.0) {
                for (int l = 0; l < kMaxLevels; ++l) s->dA.omega_l[l] = opt.smoother_weight;
            } else {
                if (s->rho.empty()) {
                    rc = estimate_rho(s);
                    if (rc) return rc;
                    s->rho_gen = s->perm_gen;
                    s->rho_aniso = binv_use != nullptr;
                    rho_fresh = true;
                }
                for (int l = 0; l < kMaxLevels; ++l) {
                    const double r = l < (int)s->rho.size() ? s->rho[l] : 0.0;
                    s->dA.omega_l[l] = r > 0.0 ? std::min(1.0, std::max(0.2, 4.0 / (3.0 * r))) : 0.7;
                }
            }
            // the kernels that walk the whole hierarchy read this copy (nothing of
            // the previous solve is still running: it ended with a synchronise;
            // the set-up kernels above do not read it)
            // (copied only when it changed: a blocking copy stalls the stream)
            const unsigned char *da = reinterpret_cast<const unsigned char *>(&s->dA);
            if (s->dA_copied.size() != sizeof(DevAmg) ||
                std::memcmp(s->dA_copied.data(), da, sizeof(DevAmg)) != 0) {
                HIPCHECK(hipMemcpy(const_cast<DevAmg *>(s->dA.dev), &s->dA, sizeof(DevAmg), hipMemcpyHostToDevice));
                s->dA_copied.assign(da, da + sizeof(DevAmg));
            }
            if (!mg_reused) HIPCHECK(launch_amg_combine(s->stream, s->dA, mixed));
        }

        // --- launch shapes -------------------------------------------------------
        if (opt.autotune) {
            rc = autotune(s, nrhs, mixed, amg, opt.element_operator == 1, &tune_ms);
            if (rc) return rc;
        } else {
            const Shape d;
            s->dT.spmv_tb = d.spmv_tb; s->dT.spmv_xcd = d.spmv_xcd;
            s->dT.sweep_tb = d.sweep_tb; s->dT.sweep_xcd = d.sweep_xcd;
            s->dT.tail_rows = d.tail_rows;
        }
        auto valid_tb = [](int t) { return t == 256 || t == 512 || t == 1024; };
        if (opt.spmv_threads) {
            if (!valid_tb(opt.spmv_threads)) return fail(s, MHS_ERR_INVALID, "solve: spmv_threads must be 256, 512 or 1024");
            s->dT.spmv_tb = opt.spmv_threads;
            s->dT.spmv_xcd = 1;
        }
        if (opt.tail_rows >= 0) s->dT.tail_rows = opt.tail_rows;
        if (opt.sweep_threads) {
            if (!valid_tb(opt.sweep_threads)) return fail(s, MHS_ERR_INVALID, "solve: sweep_threads must be 256, 512 or 1024");
            s->dT.sweep_tb = opt.sweep_threads;
            s->dT.sweep_xcd = 1;
        }

        // --- linear solve ----------------------------------------------------------
        lr = LinearResult{};
        rc = linear_solve(s, opt, nrhs, mixed, amg, lr);
        const bool failed = rc == MHS_ERR_BREAKDOWN || (rc == MHS_OK && !lr.converged);
        if (failed && amg && !rho_fresh && opt.smoother_weight <= 0.0 && attempt == 0) continue;
        if (rc) return rc;
        break;
    }
    const int total_iters = lr.iterations, replacements = lr.replacements;
    const bool converged = lr.converged;
    const double red = lr.reduction;
    s->last_amg = amg;
    HIPCHECK(hipEventRecord(s->ev_ph[2], s->stream));

    // --- recovery -------------------------------------------------------------
    HIPCHECK(launch_recover(s->stream, s->dT, T.max_ncf, nrhs, binv, gflux, d_tm, d_rm, s->C,
                            s->x64, s->press, s->outflux));
    HIPCHECK(hipEventRecord(s->ev_ph[3], s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    if (s->user_stream) {
        HIPCHECK(hipEventRecord(s->ev_out, s->stream));
        HIPCHECK(hipStreamWaitEvent(s->user_stream, s->ev_out, 0));
    }
    s->solved = true;
    s->last_nrhs = nrhs;
    ++s->solve_count;
    if (rep) {
        rep->iterations = total_iters;
        rep->replacements = replacements;
        rep->converged = converged ? 1 : 0;
        rep->reduction = red;
        rep->precond = amg ? 1 : 0;
        rep->levels = amg ? s->dA.nlev : 0;
        rep->mg_rows = 0;
        rep->mg_nnz = 0;
        if (amg)
            for (int l = 1; l < s->H.nlev(); ++l) {
                rep->mg_rows += s->H.lev[l].n;
                rep->mg_nnz += (long long)s->H.lev[l].col.size();
            }
        float ph[3] = {0.0f, 0.0f, 0.0f};
        for (int k = 0; k < 3; ++k) HIPCHECK(hipEventElapsedTime(&ph[k], s->ev_ph[k], s->ev_ph[k + 1]));
        rep->assemble_ms = ph[0];
        rep->solve_ms = ph[1];
        rep->recover_ms = ph[2];
        rep->spmv_threads = s->dT.spmv_tb;
        rep->spmv_xcd = s->dT.spmv_xcd;
        for (int l = 0; l < 8; ++l) {
            rep->smoother_weights[l] = (amg && l < s->dA.nlev - 1) ? s->dA.omega_l[l] : 0.0;
            for (int d = 0; d < 3; ++d)
                rep->coarsening[l][d] = (amg && l < (int)s->H.blocks.size()) ? s->H.blocks[l][d] : 0;
        }
        rep->sweep_threads = s->dT.sweep_tb;
        rep->sweep_xcd = s->dT.sweep_xcd;
        rep->tail_rows = s->dT.tail_rows;
        rep->tune_ms = tune_ms;
        rep->ap_nnz = amg ? s->dA.ap_slots : 0;
        rep->element_operator = s->last_ebe ? 1 : 0;
        rep->colours = s->last_ebe ? s->EL.ncolours : 0;
        rep->element_values = s->last_ebe ? s->EL.nval : 0;
        for (int l = 0; l < 8; ++l) {
            const bool have = amg && l < s->H.nlev();
            rep->level_rows[l] = have ? (l == 0 ? (long long)s->T.nf_total : (long long)s->H.lev[l].n) : 0;
            rep->level_nnz[l] = have ? (l == 0 ? (long long)s->T.slice_ptr[s->T.nslices]
                                                : (long long)s->H.lev[l].col.size())
                                      : 0;
        }
        rep->transfer_nnz[0] = rep->transfer_nnz[1] = 0;
        rep->smoothed_level = -1;
        rep->face_matrix = filled_matrix ? 1 : 0;
        rep->mg_setup_reused = mg_reused ? 1 : 0;
        if (amg)
            for (int l = 0; l < s->H.nlev(); ++l)
                if (s->H.lev[l].sa_out) {
                    rep->transfer_nnz[0] = (long long)s->H.lev[l].p_col.size();
                    rep->transfer_nnz[1] = (long long)s->H.lev[l].t_col.size();
                    rep->smoothed_level = l;
                }
    }
    if (!converged)
        return fail(s, MHS_ERR_NOT_CONVERGED,
                    "Linear solver failed to converge in " + std::to_string(total_iters) +
                        " iterations.\nResidual reduction achieved is " + std::to_string(red) + "\n");
    return MHS_OK;
}

int adjoint_impl(mhs_solver *s, const double *dJ_dv, const mhs_solve_options *opt_in, double *grad_perm,
                 double *grad_totmob, double *grad_rhomob, double *grad_src, mhs_solve_report *rep,
                 bool device_io) {
    if (!s) return MHS_ERR_INVALID;
    if (!s->solved) return fail(s, MHS_ERR_STATE, "adjoint: no solution (call solve)");
    if (!dJ_dv) return fail(s, MHS_ERR_INVALID, "adjoint: dJ_dv is null");
    if (s->T.has_periodic) return fail(s, MHS_ERR_INVALID, "adjoint: periodic boundaries are not supported");
    if (!s->perm) return fail(s, MHS_ERR_STATE, "adjoint: no permeability (call compute_inner_products)");
    if (s->last_aniso) return fail(s, MHS_ERR_STATE, "adjoint: not available after a tensor-mobility solve");
    mhs_solve_options opt;
    mhs_default_options(&opt);
    if (opt_in) opt = *opt_in;
    if (!(opt.tol > 0.0)) return fail(s, MHS_ERR_INVALID, "adjoint: tol must be positive");
    HIPCHECK(hipSetDevice(s->device));
    const Topology &T = s->T;
    const int nrhs = s->last_nrhs;
    const bool mixed = s->cap_mixed, amg = s->last_amg;
    const size_t nv = (size_t)T.nfp * nrhs, nh = (size_t)T.nhf * nrhs, nc = (size_t)T.nc;
    if (s->adj_bufs.empty()) {
        hipError_t he = hipSuccess;
        auto &L = s->adj_bufs;
        s->adj_pi = dalloc<double>(s, L, nv, &he);
        if (he == hipSuccess) s->adj_b = dalloc<double>(s, L, nv, &he);
        if (he == hipSuccess) s->adj_w = dalloc<double>(s, L, nh, &he);
        if (he == hipSuccess) s->adj_y = dalloc<double>(s, L, nh, &he);
        if (he == hipSuccess) s->adj_gK = dalloc<double>(s, L, 9 * nc, &he);
        if (he == hipSuccess) s->adj_gtm = dalloc<double>(s, L, nc, &he);
        if (he == hipSuccess) s->adj_grm = dalloc<double>(s, L, nc, &he);
        if (he == hipSuccess) s->adj_gsrc = dalloc<double>(s, L, (size_t)nrhs * nc, &he);
        if (he != hipSuccess) {
            for (auto &b : L) (void)hipFree(b.p);
            L.clear();
            return fail(s, MHS_ERR_ALLOC, std::string("adjoint buffers: ") + hipGetErrorString(he));
        }
    }
    if (device_io && s->user_stream) {
        HIPCHECK(hipEventRecord(s->ev_in, s->user_stream));
        HIPCHECK(hipStreamWaitEvent(s->stream, s->ev_in, 0));
    }
    const auto d2d = hipMemcpyDeviceToDevice;
    HIPCHECK(hipMemcpyAsync(s->adj_w, dJ_dv, sizeof(double) * nh, device_io ? d2d : hipMemcpyHostToDevice,
                            s->stream));
    // keep the forward face pressures and right-hand side; the adjoint solve reuses x64 and b
    HIPCHECK(hipMemcpyAsync(s->adj_pi, s->x64, sizeof(double) * nv, d2d, s->stream));
    HIPCHECK(hipMemcpyAsync(s->adj_b, s->b, sizeof(double) * nv, d2d, s->stream));
    HIPCHECK(launch_adjoint_rhs(s->stream, s->dT, T.max_ncf, nrhs, s->binv, s->in_totmob, s->C, s->ftype,
                                s->adj_w, s->adj_y, s->b, s->x64, s->r64));
    LinearResult lr;
    int rc = linear_solve(s, opt, nrhs, mixed, amg, lr);
    if (rc == MHS_OK && lr.converged)
        HIPCHECK(launch_adjoint_grad(s->stream, s->dT, s->dG, T.max_ncf, nrhs, s->binv, s->gflux, s->in_totmob,
                                     s->in_rhomob, s->grav, s->C, s->press, s->adj_pi, s->x64, s->ftype,
                                     s->adj_w, s->adj_gK, s->adj_gtm, s->adj_grm, s->adj_gsrc));
    // restore the forward state whatever happened
    HIPCHECK(hipMemcpyAsync(s->x64, s->adj_pi, sizeof(double) * nv, d2d, s->stream));
    HIPCHECK(hipMemcpyAsync(s->b, s->adj_b, sizeof(double) * nv, d2d, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    if (rc) return rc;
    if (rep) {
        *rep = mhs_solve_report{};
        rep->iterations = lr.iterations;
        rep->replacements = lr.replacements;
        rep->converged = lr.converged ? 1 : 0;
        rep->reduction = lr.reduction;
        rep->precond = amg ? 1 : 0;
        rep->levels = amg ? s->dA.nlev : 0;
    }
    if (!lr.converged)
        return fail(s, MHS_ERR_NOT_CONVERGED,
                    "Adjoint solve failed to converge in " + std::to_string(lr.iterations) +
                        " iterations.\nResidual reduction achieved is " + std::to_string(lr.reduction) + "\n");
    const auto out = device_io ? d2d : hipMemcpyDeviceToHost;
    if (grad_perm) HIPCHECK(hipMemcpyAsync(grad_perm, s->adj_gK, sizeof(double) * 9 * nc, out, s->stream));
    if (grad_totmob) HIPCHECK(hipMemcpyAsync(grad_totmob, s->adj_gtm, sizeof(double) * nc, out, s->stream));
    if (grad_rhomob) HIPCHECK(hipMemcpyAsync(grad_rhomob, s->adj_grm, sizeof(double) * nc, out, s->stream));
    if (grad_src)
        HIPCHECK(hipMemcpyAsync(grad_src, s->adj_gsrc, sizeof(double) * nrhs * nc, out, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    if (device_io && s->user_stream) {
        HIPCHECK(hipEventRecord(s->ev_out, s->stream));
        HIPCHECK(hipStreamWaitEvent(s->user_stream, s->ev_out, 0));
    }
    return MHS_OK;
}

}  // namespace

extern "C" {

int mhs_adjoint(mhs_solver *s, const double *dJ_dv, const mhs_solve_options *opt, double *grad_perm,
                double *grad_totmob, double *grad_rhomob, double *grad_src, mhs_solve_report *report) {
    return adjoint_impl(s, dJ_dv, opt, grad_perm, grad_totmob, grad_rhomob, grad_src, report, false);
}

int mhs_adjoint_device(mhs_solver *s, const double *dJ_dv, const mhs_solve_options *opt, double *grad_perm,
                       double *grad_totmob, double *grad_rhomob, double *grad_src,
                       mhs_solve_report *report) {
    return adjoint_impl(s, dJ_dv, opt, grad_perm, grad_totmob, grad_rhomob, grad_src, report, true);
}

}  // extern "C"

namespace {

// Arguments of the group sums: *ne = entries of group_hf.  MHS_OK with
// ngroups == 0 means nothing to do.
int check_group_args(mhs_solver *s, const char *who, int ngroups, const int *group_ptr, const int *group_hf,
                     const double *sums, size_t *ne) {
    if (!s) return MHS_ERR_INVALID;
    *ne = 0;
    if (!s->solved) return fail(s, MHS_ERR_STATE, std::string(who) + ": no solution (call solve)");
    if (ngroups < 0 || (ngroups > 0 && (!group_ptr || !group_hf || !sums)))
        return fail(s, MHS_ERR_INVALID, std::string(who) + ": bad arguments");
    if (ngroups == 0) return MHS_OK;
    if (group_ptr[0] != 0) return fail(s, MHS_ERR_INVALID, std::string(who) + ": group_ptr[0] must be 0");
    for (int g = 0; g < ngroups; ++g)
        if (group_ptr[g + 1] < group_ptr[g])
            return fail(s, MHS_ERR_INVALID, std::string(who) + ": group_ptr not monotone");
    *ne = (size_t)group_ptr[ngroups];
    for (size_t e = 0; e < *ne; ++e)
        if (group_hf[e] < 0 || group_hf[e] >= s->T.nhf)
            return fail(s, MHS_ERR_INVALID, std::string(who) + ": half-face index out of range");
    return MHS_OK;
}

size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

}  // namespace

extern "C" {

int mhs_sum_outflux(mhs_solver *s, int ngroups, const int *group_ptr, const int *group_hf, double *sums) {
    size_t ne = 0;
    if (int rc = check_group_args(s, "sum_outflux", ngroups, group_ptr, group_hf, sums, &ne)) return rc;
    if (ngroups == 0) return MHS_OK;
    const Topology &T = s->T;
    HIPCHECK(hipSetDevice(s->device));
    const int nrhs = s->last_nrhs;
    int *dptr = nullptr, *dhf = nullptr;
    double *dout = nullptr;
    hipError_t e = hipMalloc(&dptr, sizeof(int) * (ngroups + 1));
    if (e == hipSuccess) e = hipMalloc(&dhf, sizeof(int) * std::max<size_t>(ne, 1));
    if (e == hipSuccess) e = hipMalloc(&dout, sizeof(double) * (size_t)nrhs * ngroups);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dptr, group_ptr, sizeof(int) * (ngroups + 1), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess && ne > 0)
        e = hipMemcpyAsync(dhf, group_hf, sizeof(int) * ne, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = launch_group_sum(s->stream, nrhs, ngroups, T.nhf, dptr, dhf, s->outflux, dout);
    if (e == hipSuccess)
        e = hipMemcpyAsync(sums, dout, sizeof(double) * (size_t)nrhs * ngroups, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(dptr);
    (void)hipFree(dhf);
    (void)hipFree(dout);
    if (e != hipSuccess) return fail(s, MHS_ERR_HIP, std::string("sum_outflux: ") + hipGetErrorString(e));
    return MHS_OK;
}

int mhs_sum_outflux_noalloc(mhs_solver *s, int ngroups, const int *group_ptr, const int *group_hf,
                            double *sums) {
    size_t ne = 0;
    if (int rc = check_group_args(s, "sum_outflux_noalloc", ngroups, group_ptr, group_hf, sums, &ne)) return rc;
    if (ngroups == 0) return MHS_OK;
    if (!s->scratch)
        return fail(s, MHS_ERR_HIP, std::string("sum_outflux_noalloc: no scratch array on this device: ") +
                                        hipGetErrorString(s->scratch_err));
    const int nrhs = s->last_nrhs;
    // staging in this device's scratch array: group_ptr, group_hf, then the sums
    const size_t b_ptr = align256(sizeof(int) * ((size_t)ngroups + 1)), b_hf = align256(sizeof(int) * ne),
                 b_out = sizeof(double) * (size_t)nrhs * ngroups;
    if (b_ptr + b_hf + b_out > kScratchBytes)
        return fail(s, MHS_ERR_INVALID, "sum_outflux_noalloc: " + std::to_string(b_ptr + b_hf + b_out) +
                                            " bytes of staging exceed the " + std::to_string(kScratchBytes) +
                                            "-byte scratch array");
    HIPCHECK(hipSetDevice(s->device));
    unsigned char *base = static_cast<unsigned char *>(s->scratch);
    int *dptr = reinterpret_cast<int *>(base);
    int *dhf = reinterpret_cast<int *>(base + b_ptr);
    double *dout = reinterpret_cast<double *>(base + b_ptr + b_hf);
    // every output entry is written by the kernel: nothing to clear
    HIPCHECK(hipMemcpyAsync(dptr, group_ptr, sizeof(int) * ((size_t)ngroups + 1), hipMemcpyHostToDevice, s->stream));
    if (ne > 0) HIPCHECK(hipMemcpyAsync(dhf, group_hf, sizeof(int) * ne, hipMemcpyHostToDevice, s->stream));
    HIPCHECK(launch_group_sum(s->stream, nrhs, ngroups, s->T.nhf, dptr, dhf, s->outflux, dout));
    HIPCHECK(hipMemcpyAsync(sums, dout, b_out, hipMemcpyDeviceToHost, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    return MHS_OK;
}

int mhs_get_solution_device(mhs_solver *s, double *cell_pressure, double *outflux) {
    if (!s) return MHS_ERR_INVALID;
    if (!s->solved) return fail(s, MHS_ERR_STATE, "get_solution_device: no solution (call solve)");
    HIPCHECK(hipSetDevice(s->device));
    const Topology &T = s->T;
    const size_t nrhs = s->last_nrhs;
    const auto d2d = hipMemcpyDeviceToDevice;
    if (cell_pressure)
        HIPCHECK(hipMemcpyAsync(cell_pressure, s->press, sizeof(double) * T.nc * nrhs, d2d, s->stream));
    if (outflux) HIPCHECK(hipMemcpyAsync(outflux, s->outflux, sizeof(double) * T.nhf * nrhs, d2d, s->stream));
    if (s->user_stream) {
        HIPCHECK(hipEventRecord(s->ev_out, s->stream));
        HIPCHECK(hipStreamWaitEvent(s->user_stream, s->ev_out, 0));
    } else {
        HIPCHECK(hipStreamSynchronize(s->stream));
    }
    return MHS_OK;
}

int mhs_solve(mhs_solver *s, const double *totmob, const double *rhomob, const double *src,
              const int *hf_bc_type, const double *hf_bc_value, const mhs_solve_options *opt,
              mhs_solve_report *report) {
    return solve_impl(s, totmob, rhomob, src, hf_bc_type, hf_bc_value, opt, report, false);
}

int mhs_solve_device(mhs_solver *s, const double *totmob, const double *rhomob, const double *src,
                     const int *hf_bc_type, const double *hf_bc_value,
                     const mhs_solve_options *opt, mhs_solve_report *report) {
    return solve_impl(s, totmob, rhomob, src, hf_bc_type, hf_bc_value, opt, report, true);
}

static int solve_aniso_impl(mhs_solver *s, const double *lambda_t, const double *rho_lambda, const double *src,
                            const int *hf_bc_type, const double *bcval, const mhs_solve_options *opt,
                            mhs_solve_report *report, bool device) {
    if (!s) return MHS_ERR_INVALID;
    if (!s->structure_valid || !s->ip_valid || !s->perm)
        return fail(s, MHS_ERR_STATE, "solve: init() (structure and inner products) must come first");
    if (!lambda_t || !rho_lambda) return fail(s, MHS_ERR_INVALID, "solve_aniso: null mobility tensor");
    HIPCHECK(hipSetDevice(s->device));
    const Topology &T = s->T;
    const size_t nc = (size_t)T.nc;
    if (!s->an_binv) {
        hipError_t he = hipSuccess;
        auto &L = s->topo_bufs;
        s->an_lam = dalloc<double>(s, L, 9 * nc, &he);
        if (he == hipSuccess) s->an_rlam = dalloc<double>(s, L, 9 * nc, &he);
        if (he == hipSuccess) s->an_lamK = dalloc<double>(s, L, 9 * nc, &he);
        if (he == hipSuccess) s->an_dynKg = dalloc<double>(s, L, 3 * nc, &he);
        if (he == hipSuccess) s->an_gflux = dalloc<double>(s, L, (size_t)T.nhf, &he);
        if (he == hipSuccess) s->an_ones = dalloc<double>(s, L, nc, &he);
        if (he == hipSuccess)
            s->an_binv = dalloc<double>(s, L, (size_t)T.max_ncf * (T.max_ncf + 1) / 2 * nc, &he);
        if (he != hipSuccess) {
            s->an_binv = nullptr;
            return fail(s, MHS_ERR_ALLOC, std::string("device allocation: ") + hipGetErrorString(he));
        }
        std::vector<double> ones(nc, 1.0);
        HIPCHECK(hipMemcpy(s->an_ones, ones.data(), nc * sizeof(double), hipMemcpyHostToDevice));
    }
    const double *lam = lambda_t, *rlam = rho_lambda;
    if (device) {
        if (s->user_stream) {
            HIPCHECK(hipEventRecord(s->ev_in, s->user_stream));
            HIPCHECK(hipStreamWaitEvent(s->stream, s->ev_in, 0));
        }
    } else {
        HIPCHECK(hipMemcpyAsync(s->an_lam, lambda_t, 9 * nc * sizeof(double), hipMemcpyHostToDevice, s->stream));
        HIPCHECK(hipMemcpyAsync(s->an_rlam, rho_lambda, 9 * nc * sizeof(double), hipMemcpyHostToDevice, s->stream));
        lam = s->an_lam;
        rlam = s->an_rlam;
    }
    HIPCHECK(launch_aniso_ip(s->stream, s->dT, s->dG, T.max_ncf, s->perm, lam, rlam, s->grav, s->an_lamK,
                             s->an_dynKg, s->an_binv, s->an_gflux));
    if (device) {
        return solve_impl(s, s->an_ones, s->an_ones, src, hf_bc_type, bcval, opt, report, true, s->an_binv,
                          s->an_gflux);
    }
    std::vector<double> ones(nc, 1.0);
    return solve_impl(s, ones.data(), ones.data(), src, hf_bc_type, bcval, opt, report, false, s->an_binv,
                      s->an_gflux);
}

int mhs_solve_aniso(mhs_solver *s, const double *lambda_t, const double *rho_lambda, const double *src,
                    const int *hf_bc_type, const double *hf_bc_value, const mhs_solve_options *opt,
                    mhs_solve_report *report) {
    return solve_aniso_impl(s, lambda_t, rho_lambda, src, hf_bc_type, hf_bc_value, opt, report, false);
}

int mhs_solve_aniso_device(mhs_solver *s, const double *lambda_t, const double *rho_lambda, const double *src,
                           const int *hf_bc_type, const double *hf_bc_value, const mhs_solve_options *opt,
                           mhs_solve_report *report) {
    return solve_aniso_impl(s, lambda_t, rho_lambda, src, hf_bc_type, hf_bc_value, opt, report, true);
}

int mhs_get_solution(mhs_solver *s, double *cell_pressure, double *face_pressure, double *outflux) {
    if (!s) return MHS_ERR_INVALID;
    if (!s->solved) return fail(s, MHS_ERR_STATE, "get_solution: no solution (call solve)");
    HIPCHECK(hipSetDevice(s->device));
    const Topology &T = s->T;
    const int nrhs = s->last_nrhs;
    if (cell_pressure)
        HIPCHECK(hipMemcpy(cell_pressure, s->press, sizeof(double) * T.nc * nrhs, hipMemcpyDeviceToHost));
    if (outflux)
        HIPCHECK(hipMemcpy(outflux, s->outflux, sizeof(double) * T.nhf * nrhs, hipMemcpyDeviceToHost));
    if (face_pressure) {
        std::vector<double> x((size_t)T.nfp * nrhs);
        HIPCHECK(hipMemcpy(x.data(), s->x64, sizeof(double) * x.size(), hipMemcpyDeviceToHost));
        for (int k = 0; k < nrhs; ++k)
            for (int f = 0; f < T.nf_total; ++f)
                face_pressure[(size_t)k * T.nf_total + f] = x[(size_t)T.face_row[f] * nrhs + k];
    }
    return MHS_OK;
}

int mhs_solution_device(mhs_solver *s, const double **cell_pressure, const double **face_pressure,
                        const double **outflux) {
    if (!s) return MHS_ERR_INVALID;
    if (!s->solved) return fail(s, MHS_ERR_STATE, "solution_device: no solution (call solve)");
    if (cell_pressure) *cell_pressure = s->press;
    if (face_pressure) *face_pressure = s->x64;
    if (outflux) *outflux = s->outflux;
    return MHS_OK;
}

static int cell_velocity_impl(mhs_solver *s, int k, double *vel, bool device) {
    if (!s || !vel) return MHS_ERR_INVALID;
    if (!s->solved) return fail(s, MHS_ERR_STATE, "estimate_cell_velocity: no solution (call solve)");
    if (k < 0 || k >= s->last_nrhs) return fail(s, MHS_ERR_INVALID, "estimate_cell_velocity: bad rhs index");
    HIPCHECK(hipSetDevice(s->device));
    const size_t n3 = 3 * (size_t)s->nc;
    double *out = vel;
    TempBufs tmp;
    if (!device) {
        hipError_t he = hipSuccess;
        out = dalloc<double>(s, tmp.v, n3, &he);
        if (he != hipSuccess) return fail(s, MHS_ERR_ALLOC, std::string("device allocation: ") + hipGetErrorString(he));
    } else if (s->user_stream) {
        HIPCHECK(hipEventRecord(s->ev_in, s->user_stream));
        HIPCHECK(hipStreamWaitEvent(s->stream, s->ev_in, 0));
    }
    HIPCHECK(launch_cell_velocity(s->stream, s->nc, s->dT.facepos, s->dG, s->outflux + (size_t)k * s->nhf, out));
    if (!device) {
        HIPCHECK(hipMemcpyAsync(vel, out, n3 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
        HIPCHECK(hipStreamSynchronize(s->stream));
    } else if (s->user_stream) {
        HIPCHECK(hipEventRecord(s->ev_out, s->stream));
        HIPCHECK(hipStreamWaitEvent(s->user_stream, s->ev_out, 0));
    }
    return MHS_OK;
}

int mhs_estimate_cell_velocity(mhs_solver *s, int k, double *cell_velocity) {
    return cell_velocity_impl(s, k, cell_velocity, false);
}

int mhs_estimate_cell_velocity_device(mhs_solver *s, int k, double *cell_velocity) {
    return cell_velocity_impl(s, k, cell_velocity, true);
}

int mhs_post_process_fluxes(mhs_solver *s, double *max_modification) {
    if (!s) return MHS_ERR_INVALID;
    if (!s->solved) return fail(s, MHS_ERR_STATE, "post_process_fluxes: no solution (call solve)");
    HIPCHECK(hipSetDevice(s->device));
    unsigned long long zero = 0, bits = 0;
    HIPCHECK(hipMemcpyAsync(s->maxbits, &zero, sizeof(zero), hipMemcpyHostToDevice, s->stream));
    HIPCHECK(launch_post_process(s->stream, s->dT, s->last_nrhs, s->outflux, s->outflux_tmp, s->maxbits));
    HIPCHECK(hipMemcpyAsync(&bits, s->maxbits, sizeof(bits), hipMemcpyDeviceToHost, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    std::swap(s->outflux, s->outflux_tmp);
    double mm;
    std::memcpy(&mm, &bits, sizeof(mm));
    if (max_modification) *max_modification = mm;
    return MHS_OK;
}

int mhs_get_dims(mhs_solver *s, mhs_dims *d) {
    if (!s || !d) return MHS_ERR_INVALID;
    std::memset(d, 0, sizeof(*d));
    d->num_cells = s->nc;
    d->num_half_faces = s->nhf;
    d->num_instances = s->ninst;
    if (s->structure_valid) {
        d->num_faces = s->T.nf_total;
        d->num_internal_faces = s->T.nint_total;
        d->max_ncf = s->T.max_ncf;
        d->num_rows_padded = s->T.nfp;
        d->nnz = s->T.slice_ptr[s->T.nslices];
    }
    return MHS_OK;
}

int mhs_get_cell_faces(mhs_solver *s, int *hf_face, int *inst_facepos) {
    if (!s) return MHS_ERR_INVALID;
    if (!s->structure_valid) return fail(s, MHS_ERR_STATE, "get_cell_faces: no structure");
    const Topology &T = s->T;
    if (hf_face) {
        int t = 0;
        for (int c = 0; c < T.nc; ++c) {
            while (c >= T.inst_cellpos[t + 1]) ++t;
            for (int h = T.facepos[c]; h < T.facepos[c + 1]; ++h)
                hf_face[h] = T.hf_face[h];
        }
    }
    if (inst_facepos)
        for (int t = 0; t <= T.ninst; ++t) inst_facepos[t] = T.inst_facebase[t];
    return MHS_OK;
}

int mhs_get_face_rows(mhs_solver *s, int *face_row) {
    if (!s || !face_row) return MHS_ERR_INVALID;
    if (!s->structure_valid) return fail(s, MHS_ERR_STATE, "get_face_rows: no structure");
    std::memcpy(face_row, s->T.face_row.data(), sizeof(int) * s->T.nf_total);
    return MHS_OK;
}

int mhs_time_kernel(mhs_solver *s, int which, int reps, double *avg_ms) {
    if (!s || !avg_ms || reps < 1) return MHS_ERR_INVALID;
    if (!s->solved) return fail(s, MHS_ERR_STATE, "time_kernel: call solve first");
    if (which != 0 && which != 1) return fail(s, MHS_ERR_INVALID, "time_kernel: unknown kernel id");
    const DevEbe *E = (which == 0 && s->cap_mixed) ? cur_ebe(s) : nullptr;
    HIPCHECK(hipSetDevice(s->device));
    if (!E)
        if (const int rm = ensure_matrix(s)) return rm;
    const size_t nsys = (size_t)s->T.ninst * s->cap_nrhs;
    std::vector<int> ones(nsys, 1);
    HIPCHECK(hipMemcpyAsync(s->S.act, ones.data(), nsys * sizeof(int), hipMemcpyHostToDevice, s->stream));
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    hipError_t e = hipSuccess;
    // one untimed launch, then reps timed launches between two events on our stream
    for (int it = 0; it <= reps && e == hipSuccess; ++it) {
        if (it == 1) e = hipEventRecord(e0, s->stream);
        if (e != hipSuccess) break;
        if (E)
            e = launch_ebe_spmv(s->stream, s->T.max_ncf, s->cap_nrhs, *E, s->p32, s->q32, s->S);
        else if (s->cap_mixed)
            e = launch_spmv<float, float>(s->stream, s->dT, s->cap_nrhs, s->val32, s->p32, s->q32, s->R, s->S);
        else
            e = launch_spmv<double, double>(s->stream, s->dT, s->cap_nrhs, s->val64, s->p64, s->q64, s->R, s->S);
    }
    if (e == hipSuccess) e = hipEventRecord(e1, s->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::vector<int> zeros(nsys, 0);
    HIPCHECK(hipMemcpy(s->S.act, zeros.data(), nsys * sizeof(int), hipMemcpyHostToDevice));
    if (e != hipSuccess) return fail(s, MHS_ERR_HIP, std::string("time_kernel: ") + hipGetErrorString(e));
    *avg_ms = ms / reps;
    return MHS_OK;
}

int mhs_get_system(mhs_solver *s, int k, int *row_ptr, int *col, double *val, double *rhs,
                   long long *nnz) {
    if (!s) return MHS_ERR_INVALID;
    if (!s->solved) return fail(s, MHS_ERR_STATE, "get_system: no assembled system (call solve)");
    if (k < 0 || k >= s->last_nrhs) return fail(s, MHS_ERR_INVALID, "get_system: bad rhs index");
    HIPCHECK(hipSetDevice(s->device));
    const Topology &T = s->T;
    const int nrhs = s->last_nrhs;
    // padded row -> unpadded face
    std::vector<int> row_face(T.nfp, -1);
    for (int f = 0; f < T.nf_total; ++f) row_face[T.face_row[f]] = f;
    long long count = 0;
    for (int f = 0; f < T.nf_total; ++f) count += T.row_len[T.face_row[f]];
    if (nnz) *nnz = count;
    if (!col) return MHS_OK;
    if (const int rm = ensure_matrix(s)) return rm;
    std::vector<double> v(T.slice_ptr[T.nslices]);
    HIPCHECK(hipMemcpy(v.data(), s->val64, sizeof(double) * v.size(), hipMemcpyDeviceToHost));
    std::vector<double> bb((size_t)T.nfp * nrhs);
    HIPCHECK(hipMemcpy(bb.data(), s->b, sizeof(double) * bb.size(), hipMemcpyDeviceToHost));
    long long m = 0;
    if (row_ptr) row_ptr[0] = 0;
    for (int f = 0; f < T.nf_total; ++f) {
        const int r = T.face_row[f];
        const long long base = T.slice_ptr[r / kSlice] + (r % kSlice);
        for (int e = 0; e < T.row_len[r]; ++e) {
            const long long o = base + (long long)e * kSlice;
            col[m] = row_face[T.col[o]];
            if (val) val[m] = v[o];
            ++m;
        }
        if (row_ptr) row_ptr[f + 1] = (int)m;
        if (rhs) rhs[f] = bb[(size_t)r * nrhs + k];
    }
    return MHS_OK;
}

}  // extern "C"

// ===========================================================================
// Explicit upstream transport (include/euler_upstream.h)
// ===========================================================================

struct mhs_transport {
    mhs_solver *flow = nullptr;
    std::string err;
    bool ready = false;
    int device = 0;                     // the bound solver's device
    long long topo_key = -1;            // the solver's structure generation when init ran
    std::vector<DevBuf> bufs;
    DevTransport D{};
    double *F = nullptr, *G = nullptr, *C = nullptr, *pv = nullptr, *sat_bc = nullptr;
    double *s_tmp = nullptr, *s_init = nullptr, *s_host = nullptr, *inj = nullptr, *res = nullptr;
    unsigned long long *mins = nullptr;
    int *flags = nullptr;               // [0] saturation out of range, [1] zero viscous CFL time
    double fac[3] = {0.0, 0.0, 0.0};
    bool timed_ok = false;              // a transport solve set up the face coefficients
    double last_dt = 0.0;
    // key of the gravity / capillary coefficients and the gravity CFL time
    long long coeff_gen = -1;
    int coeff_cap = -1;
    double coeff_grav[3] = {0.0, 0.0, 0.0};
    double dt_g = 1e100;
    ~mhs_transport() {
        for (auto &b : bufs) (void)hipFree(b.p);
    }
};

namespace {

int tfail(mhs_transport *t, int code, const std::string &msg) {
    if (t) t->err = msg;
    return code;
}

#define TCHECK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return tfail(t, MHS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class X>
X *talloc(mhs_transport *t, size_t n, hipError_t *e) {
    void *p = nullptr;
    *e = hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(X));
    if (*e != hipSuccess) return nullptr;
    t->bufs.push_back({p, std::max<size_t>(n, 1) * sizeof(X)});
    return static_cast<X *>(p);
}

// computeSingleRockCflFactors(-1, 0, 0) (ReservoirPropertyCapillary_impl.hpp:
// 215-250): 1 / the largest difference quotient of f_w and of
// lambda_w lambda_o / lambda_t over 257 saturations; no capillary pressure
// derivative without rock tables.
void cfl_factors(double mu_w, double mu_o, double fac[3]) {
    const int N = 257;
    const double delta = 1.0 / double(N - 1);
    auto ff = [&](double s, double &f1, double &fg) {
        const double l1 = s * s / mu_w, l2 = (1.0 - s) * (1.0 - s) / mu_o;
        f1 = l1 / (l1 + l2);
        fg = l1 * l2 / (l1 + l2);
    };
    double last1, lastg;
    ff(0.0, last1, lastg);
    double max_der1 = -1e100, max_derg = -1e100, max_ffg = lastg;
    for (int i = 1; i < N; ++i) {
        const double s = double(i) * delta;
        double f1, fg;
        ff(s, f1, fg);
        max_der1 = std::max(max_der1, std::fabs(f1 - last1) / delta);
        max_derg = std::max(max_derg, std::fabs(fg - lastg) / delta);
        max_ffg = std::max(max_ffg, fg);
        last1 = f1;
        lastg = fg;
    }
    fac[0] = 1.0 / max_der1;
    fac[1] = 1.0 / max_derg;
    fac[2] = max_ffg * 0.0;
}

mhs_transport_options default_transport_options() {
    mhs_transport_options o;
    o.courant_number = 0.5;
    o.method_viscous = o.method_gravity = o.method_capillary = 1;
    o.use_cfl_viscous = o.use_cfl_gravity = o.use_cfl_capillary = 1;
    o.minimum_small_steps = 1;
    o.maximum_small_steps = 10000;
    o.check_sat = 1;
    o.clamp_sat = 0;
    return o;
}

int transport_ready(mhs_transport *t) {
    if (!t) return MHS_ERR_INVALID;
    if (!t->ready) return tfail(t, MHS_ERR_STATE, "transport: call mhs_transport_init first");
    mhs_solver *s = t->flow;
    if (!s->structure_valid || s->struct_gen != t->topo_key)
        return tfail(t, MHS_ERR_STATE, "transport: the solver's structure changed; call mhs_transport_init again");
    if (!s->perm) return tfail(t, MHS_ERR_STATE, "transport: no permeability (call compute_inner_products)");
    return MHS_OK;
}

// Gravity and capillary coefficients and the gravity CFL time, recomputed
// only when the permeability, the gravity vector or the capillary switch
// changed since the last call.
int transport_coeffs(mhs_transport *t, const double *grav, int cap) {
    mhs_solver *s = t->flow;
    if (t->coeff_gen == s->perm_gen && t->coeff_cap == cap && t->coeff_grav[0] == grav[0] &&
        t->coeff_grav[1] == grav[1] && t->coeff_grav[2] == grav[2])
        return MHS_OK;
    const int has_g = grav[0] != 0.0 || grav[1] != 0.0 || grav[2] != 0.0;
    TCHECK(launch_tr_coeffs(s->stream, t->D, s->dT, s->dG, s->perm, grav, has_g, cap));
    double dt_g = 1e100;
    if (has_g) {
        // with a zero gravity vector every local time is +inf (or NaN) and the
        // reference's minimum stays 1e100
        TCHECK(hipMemcpyAsync(t->mins + 1, &dt_g, sizeof(double), hipMemcpyHostToDevice, s->stream));
        TCHECK(launch_tr_cfl_gravity(s->stream, t->D, s->dT, s->dG, s->perm, grav, t->fac[1], t->mins + 1));
        TCHECK(hipMemcpyAsync(&dt_g, t->mins + 1, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    }
    TCHECK(hipStreamSynchronize(s->stream));
    t->dt_g = dt_g;
    t->coeff_gen = s->perm_gen;
    t->coeff_cap = cap;
    for (int d = 0; d < 3; ++d) t->coeff_grav[d] = grav[d];
    return MHS_OK;
}

// Face fluxes of the last pressure solve and the viscous CFL time.
int transport_flux(mhs_transport *t, bool check_zero, double *dt_v) {
    mhs_solver *s = t->flow;
    if (!s->solved) return tfail(t, MHS_ERR_STATE, "transport: the solver has no solution (call solve)");
    double init = 1e100;
    int zero = 0;
    TCHECK(hipMemcpyAsync(t->mins, &init, sizeof(double), hipMemcpyHostToDevice, s->stream));
    TCHECK(hipMemcpyAsync(t->flags + 1, &zero, sizeof(int), hipMemcpyHostToDevice, s->stream));
    TCHECK(launch_tr_flux(s->stream, t->D, s->dT, s->outflux, t->fac[0], t->mins, t->flags + 1));
    double v;
    int z;
    TCHECK(hipMemcpyAsync(&v, t->mins, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    TCHECK(hipMemcpyAsync(&z, t->flags + 1, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    TCHECK(hipStreamSynchronize(s->stream));
    if (check_zero && z) return tfail(t, MHS_ERR_INVALID, "Cfl computation gave dt = 0.0");
    *dt_v = v;
    return MHS_OK;
}

// Coefficients and fluxes for a transport solve, and its CFL time
// (EulerUpstream::computeCflTime, EulerUpstream_impl.hpp:265-330).
int transport_prepare(mhs_transport *t, const double *grav, const mhs_transport_options &o, double *cfl_dt) {
    t->D.method_viscous = o.method_viscous != 0;
    t->D.method_gravity = o.method_gravity != 0;
    t->D.method_capillary = o.method_capillary != 0;
    int rc = transport_coeffs(t, grav, o.method_capillary != 0);
    if (rc != MHS_OK) return rc;
    const bool use_v = o.method_viscous && o.use_cfl_viscous;
    const bool use_g = o.method_gravity && o.use_cfl_gravity;
    double dt_v = 1e100;
    rc = transport_flux(t, use_v, &dt_v);
    if (rc != MHS_OK) return rc;
    // findCFLtimeCapillary (CflCalculator.hpp:143-170) divides by the
    // capillary factor, which is 0 for this fluid: every local time is +inf
    // (or NaN) and the time stays 1e100.
    const double dt_c = 1e100;
    *cfl_dt = std::min(std::min(use_v ? dt_v : 1e100, use_g ? t->dt_g : 1e100), dt_c) * o.courant_number;
    return MHS_OK;
}

// transportSolve (:153-215) on device saturation `sat`.
int transport_run(mhs_transport *t, double *sat, double time, const double *grav, const double *inj,
                  const mhs_transport_options &o, mhs_transport_report *rep) {
    mhs_solver *s = t->flow;
    const auto t0 = std::chrono::steady_clock::now();
    double cfl_dt = 0.0;
    int rc = transport_prepare(t, grav, o, &cfl_dt);
    if (rc != MHS_OK) return rc;
    int nsteps;
    if (cfl_dt > time) {
        nsteps = o.minimum_small_steps;
    } else {
        const double steps = std::min<double>(std::ceil(time / cfl_dt), 2147483647.0);
        nsteps = std::min(std::max(int(steps), o.minimum_small_steps), o.maximum_small_steps);
    }
    const size_t nc = (size_t)s->nc;
    TCHECK(hipMemcpyAsync(t->s_init, sat, nc * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    int repeats = 0;
    const int max_repeats = 10;
    for (;;) {
        const double dt = time / nsteps;
        int zero = 0, bad = 0;
        TCHECK(hipMemcpyAsync(t->flags, &zero, sizeof(int), hipMemcpyHostToDevice, s->stream));
        double *a = sat, *b = t->s_tmp;
        for (int q = 0; q < nsteps; ++q) {
            TCHECK(launch_tr_step(s->stream, t->D, s->dT, a, b, inj, dt, o.check_sat, o.clamp_sat, t->flags));
            std::swap(a, b);
        }
        if (a != sat) TCHECK(hipMemcpyAsync(sat, a, nc * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
        TCHECK(hipMemcpyAsync(&bad, t->flags, sizeof(int), hipMemcpyDeviceToHost, s->stream));
        TCHECK(hipStreamSynchronize(s->stream));
        if (!bad) {
            t->timed_ok = true;
            t->last_dt = dt;
            if (rep) {
                rep->small_steps = nsteps;
                rep->repeats = repeats;
                rep->cfl_dt = cfl_dt;
                rep->dt = dt;
            }
            break;
        }
        if (++repeats > max_repeats) {
            TCHECK(hipMemcpyAsync(sat, t->s_init, nc * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
            TCHECK(hipStreamSynchronize(s->stream));
            return tfail(t, MHS_ERR_RANGE, "Saturation out of range in EulerUpstream (after " +
                                               std::to_string(max_repeats) + " retries with more steps)");
        }
        nsteps *= 2;
        TCHECK(hipMemcpyAsync(sat, t->s_init, nc * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    }
    if (rep) {
        rep->cfl_factor = t->fac[0];
        rep->cfl_factor_gravity = t->fac[1];
        rep->cfl_factor_capillary = t->fac[2];
        rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return MHS_OK;
}

}  // namespace

extern "C" {

void mhs_transport_default_options(mhs_transport_options *opt) {
    if (opt) *opt = default_transport_options();
}

int mhs_transport_create(mhs_transport **out, mhs_solver *flow) {
    if (!out || !flow) return MHS_ERR_INVALID;
    *out = nullptr;
    mhs_transport *t = new (std::nothrow) mhs_transport();
    if (!t) return MHS_ERR_ALLOC;
    t->flow = flow;
    t->device = flow->device;
    *out = t;
    return MHS_OK;
}

int mhs_transport_destroy(mhs_transport *t) {
    if (!t) return MHS_OK;
    (void)hipSetDevice(t->device);      // the bound solver may already be gone
    delete t;
    return MHS_OK;
}

const char *mhs_transport_last_error(const mhs_transport *t) { return t ? t->err.c_str() : "null transport"; }

int mhs_transport_init(mhs_transport *t, const double *porosity, const double *sat_bc, const mhs_fluid *fluid) {
    if (!t) return MHS_ERR_INVALID;
    mhs_solver *s = t->flow;
    if (!porosity || !fluid) return tfail(t, MHS_ERR_INVALID, "transport_init: porosity and fluid required");
    if (!s->structure_valid)
        return tfail(t, MHS_ERR_STATE, "transport_init: call mhs_init_system_structure on the solver first");
    if (!(fluid->viscosity[0] > 0.0) || !(fluid->viscosity[1] > 0.0))
        return tfail(t, MHS_ERR_INVALID, "transport_init: viscosities must be positive");
    TCHECK(hipSetDevice(s->device));
    const Topology &T = s->T;
    const int nc = T.nc, nhf = T.nhf;
    // EulerUpstreamResidual::initFinal (:415-445) keeps the periodic partner
    // of every periodic boundary face; the owner of a face is its
    // lower-numbered cell (UpdateForCell skips the face from the other side).
    std::vector<int> other(nhf), owner(nhf), nbf(nhf);
    std::vector<uint8_t> kind(nhf);
    for (int h = 0; h < nhf; ++h) {
        if (T.nb[h] >= 0) {
            other[h] = T.nb[h];
            kind[h] = 1;
        } else if (T.hf_ppart[h] >= 0) {
            other[h] = T.hf_cell[T.hf_ppart[h]];
            kind[h] = 2;
        } else {
            other[h] = -1;
            kind[h] = 0;
        }
        if (other[h] == T.hf_cell[h])
            return tfail(t, MHS_ERR_INVALID, "transport_init: a face links a cell to itself (half-face " +
                                                 std::to_string(h) + ")");
    }
    for (int h = 0; h < nhf; ++h) {
        const bool own = other[h] < 0 || T.hf_cell[h] < other[h];
        const int o = own ? h : T.hf_twin[h];
        if (o < 0 || o >= nhf) return tfail(t, MHS_ERR_INVALID, "transport_init: face without its twin half-face");
        owner[h] = o;
        nbf[h] = kind[o] == 2 ? T.hf_ppart[o] : o;
    }
    // slot-major copies for the small step: face j of cell c at j * nc + c
    const size_t nslot = (size_t)std::max(T.max_ncf, 1) * nc;
    std::vector<double> pv(nc), sbc(nslot, 1.0);
    std::vector<int> other_s(nslot, -1);
    for (int c = 0; c < nc; ++c) {
        pv[c] = s->cvol[c] * porosity[c];
        for (int h = T.facepos[c]; h < T.facepos[c + 1]; ++h) {
            const size_t sl = (size_t)(h - T.facepos[c]) * nc + c;
            other_s[sl] = other[h];
            if (sat_bc) sbc[sl] = sat_bc[h];
        }
    }
    for (auto &b : t->bufs) (void)hipFree(b.p);
    t->bufs.clear();
    t->ready = false;
    hipError_t e = hipSuccess;
    int *d_other = talloc<int>(t, nhf, &e);
    int *d_owner = e == hipSuccess ? talloc<int>(t, nhf, &e) : nullptr;
    int *d_nbf = e == hipSuccess ? talloc<int>(t, nhf, &e) : nullptr;
    uint8_t *d_kind = e == hipSuccess ? talloc<uint8_t>(t, nhf, &e) : nullptr;
    int *d_other_s = e == hipSuccess ? talloc<int>(t, nslot, &e) : nullptr;
    if (e == hipSuccess) t->sat_bc = talloc<double>(t, nslot, &e);
    if (e == hipSuccess) t->pv = talloc<double>(t, nc, &e);
    if (e == hipSuccess) t->F = talloc<double>(t, nslot, &e);
    if (e == hipSuccess) t->G = talloc<double>(t, nslot, &e);
    if (e == hipSuccess) t->C = talloc<double>(t, nslot, &e);
    if (e == hipSuccess) t->s_tmp = talloc<double>(t, nc, &e);
    if (e == hipSuccess) t->s_init = talloc<double>(t, nc, &e);
    if (e == hipSuccess) t->s_host = talloc<double>(t, nc, &e);
    if (e == hipSuccess) t->inj = talloc<double>(t, nc, &e);
    if (e == hipSuccess) t->res = talloc<double>(t, nc, &e);
    if (e == hipSuccess) t->mins = talloc<unsigned long long>(t, 2, &e);
    if (e == hipSuccess) t->flags = talloc<int>(t, 2, &e);
    if (e != hipSuccess) return tfail(t, MHS_ERR_ALLOC, std::string("transport_init: ") + hipGetErrorString(e));
    TCHECK(hipMemcpy(d_other, other.data(), nhf * sizeof(int), hipMemcpyHostToDevice));
    TCHECK(hipMemcpy(d_owner, owner.data(), nhf * sizeof(int), hipMemcpyHostToDevice));
    TCHECK(hipMemcpy(d_nbf, nbf.data(), nhf * sizeof(int), hipMemcpyHostToDevice));
    TCHECK(hipMemcpy(d_kind, kind.data(), nhf * sizeof(uint8_t), hipMemcpyHostToDevice));
    TCHECK(hipMemcpy(d_other_s, other_s.data(), nslot * sizeof(int), hipMemcpyHostToDevice));
    TCHECK(hipMemcpy(t->sat_bc, sbc.data(), nslot * sizeof(double), hipMemcpyHostToDevice));
    TCHECK(hipMemcpy(t->pv, pv.data(), nc * sizeof(double), hipMemcpyHostToDevice));
    DevTransport &D = t->D;
    D = DevTransport{};
    D.nc = nc;
    D.nhf = nhf;
    D.other = d_other;
    D.kind = d_kind;
    D.owner_hf = d_owner;
    D.owner_nbf = d_nbf;
    D.other_s = d_other_s;
    D.sat_bc_s = t->sat_bc;
    D.pv = t->pv;
    D.F = t->F;
    D.G = t->G;
    D.C = t->C;
    D.visc_w = fluid->viscosity[0];
    D.visc_o = fluid->viscosity[1];
    D.rho_w = fluid->density[0];
    D.rho_o = fluid->density[1];
    D.drho = fluid->density[0] - fluid->density[1];
    cfl_factors(D.visc_w, D.visc_o, t->fac);
    t->topo_key = s->struct_gen;
    t->timed_ok = false;
    t->coeff_gen = -1;
    t->ready = true;
    return MHS_OK;
}

int mhs_transport_solve_device(mhs_transport *t, double *saturation, double time, const double *gravity,
                               const double *injection_rates, const mhs_transport_options *opt,
                               mhs_transport_report *report) {
    int rc = transport_ready(t);
    if (rc != MHS_OK) return rc;
    if (!saturation || !gravity) return tfail(t, MHS_ERR_INVALID, "transport_solve: saturation and gravity required");
    if (!(time >= 0.0)) return tfail(t, MHS_ERR_INVALID, "transport_solve: time must be non-negative");
    mhs_solver *s = t->flow;
    TCHECK(hipSetDevice(s->device));
    const mhs_transport_options o = opt ? *opt : default_transport_options();
    if (o.minimum_small_steps < 1 || o.maximum_small_steps < o.minimum_small_steps)
        return tfail(t, MHS_ERR_INVALID, "transport_solve: need 1 <= minimum_small_steps <= maximum_small_steps");
    if (s->user_stream) {
        TCHECK(hipEventRecord(s->ev_in, s->user_stream));
        TCHECK(hipStreamWaitEvent(s->stream, s->ev_in, 0));
    }
    if (report) std::memset(report, 0, sizeof(*report));
    rc = transport_run(t, saturation, time, gravity, injection_rates, o, report);
    if (rc != MHS_OK) return rc;
    if (s->user_stream) {
        TCHECK(hipEventRecord(s->ev_out, s->stream));
        TCHECK(hipStreamWaitEvent(s->user_stream, s->ev_out, 0));
    }
    return MHS_OK;
}

int mhs_transport_solve(mhs_transport *t, double *saturation, double time, const double *gravity,
                        const double *injection_rates, const mhs_transport_options *opt,
                        mhs_transport_report *report) {
    int rc = transport_ready(t);
    if (rc != MHS_OK) return rc;
    if (!saturation || !gravity) return tfail(t, MHS_ERR_INVALID, "transport_solve: saturation and gravity required");
    if (!(time >= 0.0)) return tfail(t, MHS_ERR_INVALID, "transport_solve: time must be non-negative");
    mhs_solver *s = t->flow;
    TCHECK(hipSetDevice(s->device));
    const mhs_transport_options o = opt ? *opt : default_transport_options();
    if (o.minimum_small_steps < 1 || o.maximum_small_steps < o.minimum_small_steps)
        return tfail(t, MHS_ERR_INVALID, "transport_solve: need 1 <= minimum_small_steps <= maximum_small_steps");
    const size_t nc = (size_t)s->nc;
    TCHECK(hipMemcpy(t->s_host, saturation, nc * sizeof(double), hipMemcpyHostToDevice));
    if (injection_rates) TCHECK(hipMemcpy(t->inj, injection_rates, nc * sizeof(double), hipMemcpyHostToDevice));
    if (report) std::memset(report, 0, sizeof(*report));
    rc = transport_run(t, t->s_host, time, gravity, injection_rates ? t->inj : nullptr, o, report);
    if (rc != MHS_OK && rc != MHS_ERR_RANGE) return rc;
    TCHECK(hipMemcpy(saturation, t->s_host, nc * sizeof(double), hipMemcpyDeviceToHost));
    return rc;
}

int mhs_transport_residual(mhs_transport *t, const double *saturation, const double *gravity,
                           const double *injection_rates, const mhs_transport_options *opt, double *residual) {
    int rc = transport_ready(t);
    if (rc != MHS_OK) return rc;
    if (!saturation || !gravity || !residual)
        return tfail(t, MHS_ERR_INVALID, "transport_residual: saturation, gravity and residual required");
    mhs_solver *s = t->flow;
    TCHECK(hipSetDevice(s->device));
    const mhs_transport_options o = opt ? *opt : default_transport_options();
    t->D.method_viscous = o.method_viscous != 0;
    t->D.method_gravity = o.method_gravity != 0;
    t->D.method_capillary = o.method_capillary != 0;
    const size_t nc = (size_t)s->nc;
    TCHECK(hipMemcpy(t->s_host, saturation, nc * sizeof(double), hipMemcpyHostToDevice));
    if (injection_rates) TCHECK(hipMemcpy(t->inj, injection_rates, nc * sizeof(double), hipMemcpyHostToDevice));
    rc = transport_coeffs(t, gravity, o.method_capillary != 0);
    if (rc != MHS_OK) return rc;
    double dt_v;
    rc = transport_flux(t, false, &dt_v);
    if (rc != MHS_OK) return rc;
    TCHECK(launch_tr_residual(s->stream, t->D, s->dT, t->s_host, injection_rates ? t->inj : nullptr, t->res));
    TCHECK(hipStreamSynchronize(s->stream));
    TCHECK(hipMemcpy(residual, t->res, nc * sizeof(double), hipMemcpyDeviceToHost));
    return MHS_OK;
}

int mhs_transport_time_step(mhs_transport *t, int reps, double *avg_ms) {
    int rc = transport_ready(t);
    if (rc != MHS_OK) return rc;
    if (!avg_ms || reps < 1) return tfail(t, MHS_ERR_INVALID, "transport_time_step: bad arguments");
    if (!t->timed_ok) return tfail(t, MHS_ERR_STATE, "transport_time_step: call transport_solve first");
    mhs_solver *s = t->flow;
    TCHECK(hipSetDevice(s->device));
    int zero = 0;
    TCHECK(hipMemcpyAsync(t->flags, &zero, sizeof(int), hipMemcpyHostToDevice, s->stream));
    hipEvent_t e0, e1;
    TCHECK(hipEventCreate(&e0));
    TCHECK(hipEventCreate(&e1));
    hipError_t e = hipSuccess;
    // one untimed launch, then reps timed launches between two events on the
    // solver's stream; input: the saturation the last transport solve started from
    for (int it = 0; it <= reps && e == hipSuccess; ++it) {
        if (it == 1) e = hipEventRecord(e0, s->stream);
        if (e != hipSuccess) break;
        e = launch_tr_step(s->stream, t->D, s->dT, t->s_init, t->s_tmp, nullptr, t->last_dt, 0, 0, t->flags);
    }
    if (e == hipSuccess) e = hipEventRecord(e1, s->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return tfail(t, MHS_ERR_HIP, std::string("transport_time_step: ") + hipGetErrorString(e));
    *avg_ms = ms / reps;
    return MHS_OK;
}

int mhs_transport_mobilities_device(mhs_transport *t, const double *saturation, double *totmob, double *rhomob) {
    if (!t) return MHS_ERR_INVALID;
    if (!t->ready) return tfail(t, MHS_ERR_STATE, "transport: call mhs_transport_init first");
    if (!saturation) return tfail(t, MHS_ERR_INVALID, "transport_mobilities: saturation required");
    mhs_solver *s = t->flow;
    TCHECK(hipSetDevice(s->device));
    if (s->user_stream) {
        TCHECK(hipEventRecord(s->ev_in, s->user_stream));
        TCHECK(hipStreamWaitEvent(s->stream, s->ev_in, 0));
    }
    TCHECK(launch_tr_mobilities(s->stream, t->D, saturation, totmob, rhomob));
    if (s->user_stream) {
        TCHECK(hipEventRecord(s->ev_out, s->stream));
        TCHECK(hipStreamWaitEvent(s->user_stream, s->ev_out, 0));
    }
    return MHS_OK;
}

}  // extern "C"

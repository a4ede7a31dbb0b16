// Host-side topology of the face system: degree-of-freedom numbering, the
// sliced-ELL (SELL-64) sparsity of the Schur-reduced face matrix and the maps
// that let the per-row fill kernel gather cell contributions without atomics.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace mhs {

// Face types of IncompFlowSolverHybrid.hpp:378 (enum FaceType).
enum FaceType : uint8_t { FT_INTERNAL = 0, FT_DIRICHLET = 1, FT_NEUMANN = 2, FT_PERIODIC = 3 };

constexpr int kSlice = 64;     // SELL slice height = one wavefront
constexpr int kRowBlock = 1024; // rows per PCG workgroup; instances start on a block
constexpr int kMaxFaces = 16;  // faces per cell supported by the cell kernels
constexpr int kMaxRowLen = 255;

struct Topology {
    // grid
    int nc = 0, nhf = 0, ninst = 1, max_ncf = 0;
    std::vector<int> facepos, nb, inst_cellpos;
    // numbering
    int nfp = 0;                          // padded rows (sum over instances)
    int nf_total = 0, nint_total = 0;     // unpadded faces / internal faces
    std::vector<int> inst_nf, inst_nint, inst_rowbase, inst_facebase;
    std::vector<int> inst_regrow;         // row of each instance's dof 0 (the regularised entry)
    std::vector<int> hf_dof;              // padded global row of each half-face
    std::vector<int> hf_face;             // its face (instance face base + reference dof number)
    std::vector<int> hf_cell;
    std::vector<int> hf_twin;             // neighbour half-face or periodic partner half-face
    std::vector<int> hf_ppart;            // periodic partner half-face (structure time)
    std::vector<int> hf_col;              // column after periodic redirection
    std::vector<uint8_t> hf_redirect;     // 1: periodic and partner dof < own dof
    std::vector<uint8_t> hf_periodic;     // periodic at structure time
    // rows
    std::vector<int> row_hf;              // [2*nfp] contributing half-faces
    std::vector<int> row_cell;            // [2*nfp] their cells (-1: none), read per row without a gather
    std::vector<uint8_t> row_loc;         // [2*nfp] their local face index in the cell
    std::vector<int> dof_hf;              // first half-face of the face, -1 for padding
    std::vector<int> ptn;                 // periodic partner row, -1
    std::vector<uint8_t> row_len, dslot, pslot;
    std::vector<int> face_row;            // unpadded face -> padded row
    std::vector<int> block_seg;           // PCG workgroup -> instance
    std::vector<int> seg_blk;             // [ninst+1] first workgroup of each instance
    // SELL-64
    int nslices = 0;
    std::vector<long long> slice_ptr;     // [nslices+1]
    std::vector<int> col;                 // [slice_ptr.back()]
    // per-cell maps
    std::vector<long long> sq_pos;        // sum n^2 prefix
    std::vector<long long> pk_pos;        // sum n(n+1)/2 prefix
    std::vector<uint8_t> slotmap;         // slot of (cell c, r, j) in row(c, r), at (r * max_ncf + j) * nc + c
    bool has_periodic = false;
};

// Builds the topology.  Rows are the faces of each instance in the order:
// face family (hf_family, a non-negative sort key per half-face, e.g. the
// dominant axis of the face normal; a face takes the smaller key of its two
// half-faces; null: one family), then, within windows of kRowBlock rows,
// estimated row length, then reference dof number.  A wavefront then holds faces of one family
// (its gathers point the same way) and SELL slices hold rows of similar
// length (little padding); everything defined by the reference numbering
// (periodic redirection, the regularised dof 0, face outputs) goes through
// the dof numbers.  Returns an empty string on success, else an error.
std::string build_topology(Topology &T, int ninst, const int *inst_cellpos, int nc,
                           const int *facepos, const int *nb, const int *hf_bc_type,
                           const int *hf_partner, const int *hf_family = nullptr);

// Number the dofs of one instance exactly as enumerateGridDof does
// (IncompFlowSolverHybrid.hpp:956-1067).  Cells [c0,c1); writes local dof ids.
std::string enumerate_instance(int c0, int c1, const int *facepos, const int *nb,
                               int *hf_local, int *nint, int *nf, int *max_ncf);

inline int packed_index(int i, int j, int n) {  // i <= j, row-major upper triangle
    return i * n - (i * (i - 1)) / 2 + (j - i);
}

}  // namespace mhs

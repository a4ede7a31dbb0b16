// Element-by-element layout of the level-0 face operator (host side).
//
// The face matrix is the sum over cells of each cell's local Schur complement
// (after the Dirichlet rules).  Applying it cell by cell reads n(n+1)/2 packed
// values and n face rows per cell (21 + 6 for a hexahedron) instead of every
// face row's entries with their column indices (11 + 11 per face of a
// hexahedral grid, three faces per cell).  Cells are coloured so that no two
// cells of one colour share a face; one launch per colour then scatters to
// face rows without atomics.  Cells are grouped into segments of one colour,
// one instance and one face count, each padded to whole 256-cell blocks, and
// their values and face rows are stored in 64-cell tiles, component-major
// inside a tile (entry q of the cell in slot i at seg_val + ebe_slot(i, q,
// ncomp), mhs_kernels.h), so that a wavefront's loads coalesce and stay
// within one contiguous range.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "mhs_amg.h"
#include "mhs_structure.h"

namespace mhs {

constexpr int kEbeBlock = 256;      // cells per workgroup
constexpr int kMaxColours = 8;
// Face-row flags in the element dof table: the cell is the first (lowest
// colour) of the face's cells to write it, and the last (highest colour).
constexpr uint32_t kEbeFirst = 1u << 30, kEbeLast = 1u << 31, kEbeRow = (1u << 30) - 1;

struct EbeLayout {
    int ncolours = 0, nseg = 0, nblk = 0, nslots = 0;
    long long nval = 0, ndof = 0;
    std::vector<int> colour_blk;             // [ncolours + 1] first block of each colour
    std::vector<int> colour_split;           // [ncolours] first block of each colour whose cells
                                             // are not hexahedra (6 faces); hexahedra come first
    std::vector<int> blk_seg, blk_slot;      // [nblk] segment, partial-sum slot
    std::vector<int> seg_n, seg_inst, seg_count, seg_len, seg_blk0;
    std::vector<long long> seg_val, seg_dof; // value and dof bases of each segment
    std::vector<int> inst_slot;              // [ninst + 1] partial slots per instance (contiguous)
    // Per block, what its kernel needs before its first data load, in two
    // records read in one round trip: (face count, instance, cells in the
    // block, partial-sum slot) and (value base, face-row base) of its first
    // 64-cell tile.
    std::vector<int> blk_info;               // [4 * nblk]
    std::vector<long long> blk_base;         // [2 * nblk]
    std::vector<int> cell_seg, cell_local;   // [nc] segment and slot of each cell
    std::vector<uint32_t> dof;               // [ndof] face row | kEbeFirst | kEbeLast
};

// Cell colouring (no two cells sharing a face have one colour); returns the
// number of colours.  Without logical indices: greedy in index order, each
// cell taking the smallest colour none of its already coloured neighbours has
// (the checkerboard on box grids).  With logical indices ijk[3 * nc]: the
// parity of i + j + k first, and in index order a cell that shares its colour
// with a lower-numbered neighbour (across a fault) takes the smallest colour
// from 2 up that no neighbour has.  Most cells of a faulted grid then keep
// their checkerboard colour, which is also how the face rows are classed.
int greedy_colours(int nc, const int *facepos, const int *nb, std::vector<int> &colour,
                   const int *ijk = nullptr);

// Gather lists of the A P slots (face rows by level-1 aggregates agg, pattern
// ap) over the element values: slot (f, J) sums the entries (i, j) of every
// cell whose local face i is row f and whose local face j has its row in
// aggregate J.  With the Dirichlet rules inside the element values this is
// the face matrix's A P.  ptr[slots + 1], src: positions in the packed values.
// Returns "" or why the layout does not cover the pattern.
std::string ebe_ap_lists(const EbeLayout &E, const Topology &T, const ApPattern &ap, const std::vector<int> &agg,
                         std::vector<int> &ptr, std::vector<int> &src);

// Builds the layout for a structure; returns an empty string, or the reason
// why the element-by-element operator does not apply (periodic faces, whose
// constraint rows are not a sum of cell matrices; too many colours).  ijk:
// logical cell indices or null (see greedy_colours).
std::string build_ebe(EbeLayout &E, const Topology &T, const int *ijk = nullptr);

}  // namespace mhs

// Aggregation multigrid hierarchy for the face system (host-side set-up).
//
// Aggregates are geometric: a face belongs to its owner cell (the cell of its
// first half-face), and cells are grouped in blocks of their logical (i, j, k)
// indices, 2 or 1 cells along each direction depending on the coupling
// strength along it (semi-coarsening of anisotropic problems); every coarser
// level groups the blocks again.  The coarse operators are algebraic,
// A_{l+1} = P' A_l P with the piecewise-constant P of the aggregates (smoothed
// on one transfer).  Their sparsity and, for every coarse entry,
// the list of finer entries that sum into it are built here once per
// structure, so that each solve forms the Galerkin products with a pure gather.
#pragma once

#include <array>
#include <string>
#include <vector>

#include "mhs_structure.h"

namespace mhs {

constexpr int kMaxCoarsest = 64;  // largest coarsest system per instance (dense inverse)

struct AmgLevel {
    int n = 0;                      // rows (all instances, unpadded)
    std::vector<int> inst_ptr;      // [ninst+1] first row of each instance
    std::vector<int> dims;          // [3 * ninst] block dimensions per instance
    std::vector<int> row_ptr, col;  // CSR pattern (columns sorted)
    std::vector<int> diag;          // position of the diagonal in each row
    std::vector<int> src_ptr;       // [nnz+1] gather list ranges
    std::vector<int> src;           // positions in the finer level's value array
    std::vector<int> agg;           // [n_finer] row of this level each finer row feeds (-1: none)
    std::vector<int> child_ptr;     // [n+1]
    std::vector<int> child;         // finer rows of each row
    // Levels built by smoothed aggregation take their values as sums of
    // products over pairs (P position, T position) of the finer level's
    // transfer: src holds the P positions, src2 the T positions.
    bool sa_in = false;
    std::vector<int> src2;
    // Smoothed-aggregation transfer to the next coarser level (when sa_out):
    // P = (I - w_p D^-1 A) P_t with the tentative aggregates agg of the next
    // level, T = A P, and P transposed for the restriction.
    bool sa_out = false;
    std::vector<int> p_ptr, p_col, p_own;      // P rows; p_own[i]: entry of row i in its own aggregate
    std::vector<int> p_src_ptr, p_src;         // per P entry: positions of this level's values
    std::vector<int> pt_ptr, pt_col, pt_map;   // rows of the coarser level: fine row, P position
    std::vector<int> p_tpos;                   // per P entry: its position in the P' rows (inverse of pt_map)
    std::vector<int> t_ptr, t_col;             // T = A P
    std::vector<int> t_pair_ptr, t_pa, t_pb;   // per T entry: (value position, P position)
    std::vector<int> t_p;                      // per T entry: the P entry of the same row and column, or -1
};

// A P with the level-1 aggregates: the fine-row by aggregate matrix of the
// post-smoothing sweep, (A P)_{iJ} = sum of a_ij over the columns j in
// aggregate J.  Stored in the SELL-64 layout of the face matrix (lane-major
// slices; padding slots have column 0 and value 0), with, for every slot, the
// list of face-matrix positions that sum into it.
struct ApPattern {
    std::vector<long long> slice_ptr;  // [nslices+1]
    std::vector<int> col;              // level-1 row of each slot
    std::vector<int> src_ptr;          // [slots+1]
    std::vector<int> src;              // face-matrix SELL positions
};

struct AmgHierarchy {
    std::vector<AmgLevel> lev;      // lev[0] unused placeholder, lev[1..] coarse levels
    std::vector<std::array<int, 3>> blocks;  // logical block factors of each transfer (l -> l+1)
    std::vector<int> coarse_off;    // [ninst+1] offsets of the dense coarsest inverses
    ApPattern ap;
    int nlev() const { return int(lev.size()); }
};

// Coupling strength along the three logical directions, e.g. the mean
// two-point transmissibility of the faces between logical neighbours.  The
// aggregates of every transfer are blocks of the finer level's logical grid,
// 2 or 1 cells along each direction: a direction is coarsened when its
// strength is at least half the strongest (a quarter below level 1), and the
// strengths then scale like those of the coarse blocks (cross-section over
// length).  Directions of extent 1 in every instance are never coarsened.
using Strength = std::array<double, 3>;

// Block factors of each transfer for these strengths and per-instance
// logical dimensions dims[3 * ninst], until every instance has at most
// kMaxCoarsest rows.
std::vector<std::array<int, 3>> coarsening_plan(const Strength &strength, const std::vector<int> &dims);

// Logical dimensions of every instance from ijk[3*nc] (instance-local,
// 0-based).  Returns an empty string on success.
std::string instance_dims(const Topology &T, const std::vector<int> &ijk, std::vector<int> &dims);

// Builds the hierarchy from logical cell indices ijk[3*nc] (instance-local,
// 0-based) with the blocks of coarsening_plan.  The transfer from the face
// matrix to level 1 is the tentative (piecewise-constant) one; with smoothed
// = true one coarse transfer is smoothed (the first from a level with at most
// an eighth as many rows as faces), the others stay tentative.  Returns an
// empty string on success.
std::string build_amg(AmgHierarchy &H, const Topology &T, const std::vector<int> &ijk,
                      const Strength &strength = {1.0, 1.0, 1.0}, bool smoothed = true);

}  // namespace mhs

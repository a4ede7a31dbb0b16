// Corner-point grid with the members of Dune::CpGrid that the reference's
// set-up and solvers call (setupGridAndProps.hpp:92, 114-130:
// processEclipseFormat, globalCell, createCartesian, size,
// setUniqueBoundaryIds; IncompFlowSolverHybrid via GridInterfaceEuler).
//
// The geometry comes from mhs_deck_build_grid (eclipse_grid.h): vertical
// pillars on a regular lattice, gap-free layers, faults between columns, all
// cells active.  Cells are numbered i fastest, then j, then k, and each cell's
// faces come in the order x-, x+, y-, y+ (one face per overlapped neighbour
// layer), top, bottom.  Boundary ids follow CpGrid's rule (the side of its
// cell a boundary face lies on: 1 / 2 for the x- / x+ sides, 3 / 4 for y,
// 5 / 6 for top / bottom), faces exposed by a fault included; CpGrid's own
// face order and its triangulated face geometry are not reproduced.
#ifndef OPM_PORSOL_AMD_CORNERPOINTGRID_HPP
#define OPM_PORSOL_AMD_CORNERPOINTGRID_HPP

#include <array>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "EclipseDeck.hpp"
#include "GridInterfaceEuler.hpp"

namespace Opm {

class CornerPointGrid : public FlatGrid {
public:
    /// Builds the grid of an ECLIPSE deck (CpGrid::processEclipseFormat).
    /// periodic_extension, turn_normals and clip_z are not supported and throw
    /// when set.
    void processEclipseFormat(const DeckConstPtr &deck, bool periodic_extension = false,
                              bool turn_normals = false, bool clip_z = false) {
        if (!deck) throw std::invalid_argument("processEclipseFormat: null deck");
        processEclipseFormat(*deck, periodic_extension, turn_normals, clip_z);
    }
    void processEclipseFormat(const Deck &deck, bool periodic_extension = false, bool turn_normals = false,
                              bool clip_z = false) {
        if (periodic_extension || turn_normals || clip_z)
            throw std::runtime_error(
                "processEclipseFormat: periodic_extension, turn_normals and clip_z are not supported");
        int nc = 0, nhf = 0;
        deck.check(mhs_deck_build_grid(deck.handle(), &nc, &nhf), "building the corner-point grid");
        facepos.resize(nc + 1);
        neighbour.resize(nhf);
        boundary_id.resize(nhf);
        cell_volume.resize(nc);
        cell_centroid.resize(3 * (size_t)nc);
        face_area.resize(nhf);
        face_centroid.resize(3 * (size_t)nhf);
        face_normal.resize(3 * (size_t)nhf);
        corners_.resize(24 * (size_t)nc);
        std::vector<int> ijk(3 * (size_t)nc);
        deck.check(mhs_deck_grid_arrays(deck.handle(), facepos.data(), neighbour.data(), boundary_id.data(),
                                        cell_volume.data(), cell_centroid.data(), face_area.data(),
                                        face_centroid.data(), face_normal.data(), ijk.data(), corners_.data()),
                   "reading the corner-point grid");
        logical_size_ = deck.dims();
        global_cell_.resize(nc);
        for (int c = 0; c < nc; ++c)
            global_cell_[c] = ijk[3 * (size_t)c] +
                              logical_size_[0] * (ijk[3 * (size_t)c + 1] + logical_size_[1] * ijk[3 * (size_t)c + 2]);
        // id 7 marks a face exposed by a fault; CpGrid numbers it by the side
        // of its cell, which for a vertical pillar plane is the sign of the
        // normal's x or y component
        for (int h = 0; h < nhf; ++h)
            if (boundary_id[h] == 7) {
                const double *n = &face_normal[3 * (size_t)h];
                const int axis = std::abs(n[0]) >= std::abs(n[1]) ? 0 : 1;
                boundary_id[h] = 2 * axis + (n[axis] > 0.0 ? 2 : 1);
            }
        side_id_ = boundary_id;
        unique_ = false;
    }

    /// A box of dims[0] x dims[1] x dims[2] cells of size cellsz
    /// (CpGrid::createCartesian).
    void createCartesian(const std::array<int, 3> &dims, const std::array<double, 3> &cellsz) {
        CartesianGrid box(dims[0], dims[1], dims[2], cellsz[0], cellsz[1], cellsz[2]);
        static_cast<FlatGrid &>(*this) = box;
        logical_size_ = dims;
        const int nc = numberOfCells();
        global_cell_.resize(nc);
        corners_.resize(24 * (size_t)nc);
        for (int c = 0; c < nc; ++c) {
            global_cell_[c] = c;
            const int i = c % dims[0], j = (c / dims[0]) % dims[1], k = c / (dims[0] * dims[1]);
            static const int ring[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
            for (int layer = 0; layer < 2; ++layer)
                for (int v = 0; v < 4; ++v) {
                    double *p = &corners_[24 * (size_t)c + 3 * (4 * layer + v)];
                    p[0] = (i + ring[v][0]) * cellsz[0];
                    p[1] = (j + ring[v][1]) * cellsz[1];
                    p[2] = (k + layer) * cellsz[2];
                }
        }
        side_id_ = boundary_id;
        unique_ = false;
    }

    /// Logical Cartesian index i + nx (j + ny k) of every cell.
    const std::vector<int> &globalCell() const { return global_cell_; }
    const std::array<int, 3> &logicalCartesianSize() const { return logical_size_; }

    /// true: number the boundary faces 1..nb in cell / local-face order;
    /// false: ids by side (CpGrid::setUniqueBoundaryIds).
    void setUniqueBoundaryIds(bool unique) {
        unique_ = unique;
        int next = 1;
        for (size_t h = 0; h < neighbour.size(); ++h)
            if (neighbour[h] < 0) boundary_id[h] = unique ? next++ : side_id_[h];
    }
    bool uniqueBoundaryIds() const { return unique_; }

    /// The eight corners of cell c, VTK hexahedron order: the top ring
    /// (i, j), (i+1, j), (i+1, j+1), (i, j+1), then the bottom ring; 24 doubles.
    const double *cellCorners(int c) const { return &corners_[24 * (size_t)c]; }

private:
    std::vector<int> global_cell_;
    std::array<int, 3> logical_size_{{0, 0, 0}};
    std::vector<double> corners_;
    std::vector<int> side_id_;
    bool unique_ = false;
};

}  // namespace Opm

namespace Dune {
/// The reference's grid type name for corner-point grids.
typedef Opm::CornerPointGrid CpGrid;
}  // namespace Dune

#endif

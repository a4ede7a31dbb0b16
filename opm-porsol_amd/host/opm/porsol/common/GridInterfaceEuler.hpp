// Dune-free grid interface with the member API of
// opm/porsol/common/GridInterfaceEuler.hpp (GIE::Face :98-241, GIE::Cell
// :318-367, GridInterfaceEuler :411-614).
//
// The grid is a FlatGrid: cells in iteration order, each cell's faces
// (intersections) in local order, with the geometry the solver reads.
// CartesianGrid fills one the way a Dune YaspGrid orders it (cells with x
// fastest, intersections x-, x+, y-, y+, z-, z+, boundary ids 1..6 by side;
// setUniqueBoundaryIds(true) numbers boundary faces 1..nb as CpGrid does).
#ifndef OPM_PORSOL_AMD_GRIDINTERFACEEULER_HPP
#define OPM_PORSOL_AMD_GRIDINTERFACEEULER_HPP

#include <algorithm>
#include <cassert>
#include <climits>
#include <cmath>
#include <stdexcept>
#include <vector>

namespace Opm {

/// Small fixed 3-vector standing in for Dune::FieldVector<double, 3>.
struct Vec3 {
    double v[3];
    Vec3() : v{0.0, 0.0, 0.0} {}
    explicit Vec3(double a) : v{a, a, a} {}
    Vec3(double a, double b, double c) : v{a, b, c} {}
    enum { dimension = 3 };
    double &operator[](int i) { return v[i]; }
    double operator[](int i) const { return v[i]; }
    Vec3 &operator+=(const Vec3 &o) { v[0] += o.v[0]; v[1] += o.v[1]; v[2] += o.v[2]; return *this; }
    Vec3 &operator-=(const Vec3 &o) { v[0] -= o.v[0]; v[1] -= o.v[1]; v[2] -= o.v[2]; return *this; }
    Vec3 &operator*=(double a) { v[0] *= a; v[1] *= a; v[2] *= a; return *this; }
    double operator*(const Vec3 &o) const { return v[0] * o.v[0] + v[1] * o.v[1] + v[2] * o.v[2]; }
    double two_norm() const { return std::sqrt(*this * *this); }
    const double *begin() const { return v; }
    const double *end() const { return v + 3; }
};

/// Flat unstructured grid description (the data behind the iterators).
struct FlatGrid {
    static const int dimension = 3;
    std::vector<int> facepos;            // [nc+1]
    std::vector<int> neighbour;          // [nhf]  -1 on the boundary
    std::vector<int> boundary_id;        // [nhf]  0 inside
    std::vector<double> cell_volume;     // [nc]
    std::vector<double> cell_centroid;   // [3 nc]
    std::vector<double> face_area;       // [nhf]
    std::vector<double> face_centroid;   // [3 nhf]
    std::vector<double> face_normal;     // [3 nhf] unit outer normal

    int numberOfCells() const { return int(cell_volume.size()); }
    int numberOfHalfFaces() const { return int(neighbour.size()); }
    int size(int codim) const { return codim == 0 ? numberOfCells() : -1; }
};

/// Uniform box grid in Dune YaspGrid order.
class CartesianGrid : public FlatGrid {
public:
    CartesianGrid() {}
    CartesianGrid(int nx, int ny, int nz, double dx = 1.0, double dy = 1.0, double dz = 1.0) {
        init(nx, ny, nz, dx, dy, dz);
    }
    void init(int nx, int ny, int nz, double dx = 1.0, double dy = 1.0, double dz = 1.0) {
        if (nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("CartesianGrid: bad dimensions");
        dims_[0] = nx; dims_[1] = ny; dims_[2] = nz;
        h_[0] = dx; h_[1] = dy; h_[2] = dz;
        const int nc = nx * ny * nz;
        facepos.resize(nc + 1);
        neighbour.assign(6 * (size_t)nc, -1);
        boundary_id.assign(6 * (size_t)nc, 0);
        cell_volume.assign(nc, dx * dy * dz);
        cell_centroid.resize(3 * (size_t)nc);
        face_area.resize(6 * (size_t)nc);
        face_centroid.resize(18 * (size_t)nc);
        face_normal.assign(18 * (size_t)nc, 0.0);
        const double area[3] = {dy * dz, dx * dz, dx * dy};
        for (int c = 0; c < nc; ++c) {
            const int i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
            const int ijk[3] = {i, j, k};
            const int n[3] = {nx, ny, nz};
            const int stride[3] = {1, nx, nx * ny};
            facepos[c] = 6 * c;
            double cen[3] = {(i + 0.5) * dx, (j + 0.5) * dy, (k + 0.5) * dz};
            for (int d = 0; d < 3; ++d) cell_centroid[3 * (size_t)c + d] = cen[d];
            for (int d = 0; d < 3; ++d)
                for (int side = 0; side < 2; ++side) {
                    const size_t h = 6 * (size_t)c + 2 * d + side;
                    const bool at_bnd = side == 0 ? ijk[d] == 0 : ijk[d] == n[d] - 1;
                    neighbour[h] = at_bnd ? -1 : c + (side == 0 ? -stride[d] : stride[d]);
                    boundary_id[h] = at_bnd ? 2 * d + side + 1 : 0;
                    face_area[h] = area[d];
                    for (int e = 0; e < 3; ++e) face_centroid[3 * h + e] = cen[e];
                    face_centroid[3 * h + d] += (side == 0 ? -0.5 : 0.5) * h_[d];
                    face_normal[3 * h + d] = side == 0 ? -1.0 : 1.0;
                }
        }
        facepos[nc] = 6 * nc;
    }
    /// As CpGrid::setUniqueBoundaryIds: number boundary faces 1..nb in cell /
    /// local-face order (true) or by side 1..6 (false).
    void setUniqueBoundaryIds(bool unique) {
        int next = 1;
        for (size_t h = 0; h < neighbour.size(); ++h) {
            if (neighbour[h] >= 0) continue;
            boundary_id[h] = unique ? next++ : int(h % 6) + 1;
        }
    }
    const int *dims() const { return dims_; }
    const double *cellSize() const { return h_; }

private:
    int dims_[3] = {0, 0, 0};
    double h_[3] = {1.0, 1.0, 1.0};
};

template <class GridT>
class GridInterfaceEuler;

namespace GIE {

template <class GI>
class Face {
public:
    typedef double Scalar;
    typedef Vec3 Vector;
    typedef int Index;
    enum { BoundaryMarkerIndex = -999, LocalEndIndex = INT_MAX };

    Face() : pgrid_(nullptr), cell_(0), local_index_(-1) {}
    Face(const GI &grid, int cell, int local) : pgrid_(&grid), cell_(cell), local_index_(local) {}

    Scalar area() const { return g().face_area[hf()]; }
    Vector centroid() const {
        const double *p = &g().face_centroid[3 * hf()];
        return Vector(p[0], p[1], p[2]);
    }
    Vector normal() const {
        const double *p = &g().face_normal[3 * hf()];
        return Vector(p[0], p[1], p[2]);
    }
    bool boundary() const { return g().neighbour[hf()] < 0; }
    int boundaryId() const { return g().boundary_id[hf()]; }
    Index cellIndex() const { return cell_; }
    Index neighbourCellIndex() const {
        const int n = g().neighbour[hf()];
        return n < 0 ? BoundaryMarkerIndex : n;
    }
    Index index() const { return pgrid_->faceIndex(cell_, local_index_); }
    Index localIndex() const { return local_index_; }
    Scalar neighbourCellVolume() const { return g().cell_volume[g().neighbour[hf()]]; }
    /// Flat half-face index (cell by cell, local order).
    Index halfFaceIndex() const { return int(hf()); }

protected:
    const typename GI::GridType &g() const { return pgrid_->grid(); }
    size_t hf() const { return size_t(g().facepos[cell_] + local_index_); }
    const GI *pgrid_;
    int cell_;
    int local_index_;
};

template <class GI>
class FaceIterator : public Face<GI> {
    typedef Face<GI> F;

public:
    FaceIterator() {}
    FaceIterator(const GI &grid, int cell, int local) : F(grid, cell, local) {}
    const FaceIterator &operator*() const { return *this; }
    const FaceIterator *operator->() const { return this; }
    FaceIterator &operator++() { ++F::local_index_; return *this; }
    bool operator==(const FaceIterator &o) const {
        return F::cell_ == o.F::cell_ && F::local_index_ == o.F::local_index_;
    }
    bool operator!=(const FaceIterator &o) const { return !(*this == o); }
    bool operator<(const FaceIterator &o) const {
        return F::cell_ == o.F::cell_ ? F::local_index_ < o.F::local_index_ : F::cell_ < o.F::cell_;
    }
};

template <class GI>
class Cell {
public:
    typedef GIE::FaceIterator<GI> FaceIterator;
    typedef Vec3 Vector;
    typedef double Scalar;
    typedef int Index;

    Cell() : pgrid_(nullptr), cell_(0) {}
    Cell(const GI &grid, int cell) : pgrid_(&grid), cell_(cell) {}
    FaceIterator facebegin() const { return FaceIterator(*pgrid_, cell_, 0); }
    FaceIterator faceend() const {
        const auto &g = pgrid_->grid();
        return FaceIterator(*pgrid_, cell_, g.facepos[cell_ + 1] - g.facepos[cell_]);
    }
    Scalar volume() const { return pgrid_->grid().cell_volume[cell_]; }
    Vector centroid() const {
        const double *p = &pgrid_->grid().cell_centroid[3 * (size_t)cell_];
        return Vector(p[0], p[1], p[2]);
    }
    Index index() const { return cell_; }

protected:
    const GI *pgrid_;
    int cell_;
};

template <class GI>
class CellIterator : public Cell<GI> {
    typedef Cell<GI> C;

public:
    typedef typename C::Vector Vector;
    typedef typename C::Scalar Scalar;
    typedef typename C::Index Index;
    CellIterator() {}
    CellIterator(const GI &grid, int cell) : C(grid, cell) {}
    const CellIterator &operator*() const { return *this; }
    const CellIterator *operator->() const { return this; }
    CellIterator &operator++() { ++C::cell_; return *this; }
    bool operator==(const CellIterator &o) const { return C::cell_ == o.C::cell_; }
    bool operator!=(const CellIterator &o) const { return C::cell_ != o.C::cell_; }
};

}  // namespace GIE

template <class GridT>
class GridInterfaceEuler {
public:
    typedef GridT GridType;
    typedef GridInterfaceEuler<GridT> InterfaceType;
    typedef GIE::CellIterator<InterfaceType> CellIterator;
    typedef Vec3 Vector;
    typedef double Scalar;
    typedef int Index;
    enum { Dimension = 3 };

    GridInterfaceEuler() : pgrid_(nullptr), num_faces_(0), max_faces_per_cell_(0) {}
    explicit GridInterfaceEuler(const GridT &grid, bool build_facemap = true)
        : pgrid_(&grid), num_faces_(0), max_faces_per_cell_(0) {
        if (build_facemap) buildFaceIndices();
    }
    void init(const GridT &grid, bool build_facemap = true) {
        pgrid_ = &grid;
        if (build_facemap) buildFaceIndices();
    }
    CellIterator cellbegin() const { return CellIterator(*this, 0); }
    CellIterator cellend() const { return CellIterator(*this, numberOfCells()); }
    int numberOfCells() const { return grid().numberOfCells(); }
    int numberOfFaces() const { assert(num_faces_ != 0); return num_faces_; }
    int maxFacesPerCell() const { assert(max_faces_per_cell_ != 0); return max_faces_per_cell_; }
    const GridT &grid() const { assert(pgrid_); return *pgrid_; }
    /// Unique face numbering of buildFaceIndices (GridInterfaceEuler.hpp:
    /// 498-611): faces between cells in order of discovery, then boundary faces.
    Index faceIndex(int cell_index, int local_face_index) const {
        assert(num_faces_ != 0);
        return face_indices_[grid().facepos[cell_index] + local_face_index];
    }

private:
    const GridT *pgrid_;
    int num_faces_;
    int max_faces_per_cell_;
    std::vector<int> face_indices_;

    void buildFaceIndices() {
        const GridT &g = grid();
        const int nc = g.numberOfCells();
        face_indices_.assign(g.neighbour.size(), -1);
        int count = 0;
        for (int c = 0; c < nc; ++c) {
            max_faces_per_cell_ = std::max(max_faces_per_cell_, g.facepos[c + 1] - g.facepos[c]);
            for (int h = g.facepos[c]; h < g.facepos[c + 1]; ++h)
                if (g.neighbour[h] > c) face_indices_[h] = count++;
        }
        for (int c = 0; c < nc; ++c)
            for (int h = g.facepos[c]; h < g.facepos[c + 1]; ++h) {
                const int n = g.neighbour[h];
                if (n < 0) {
                    face_indices_[h] = count++;
                } else if (n < c) {
                    for (int t = g.facepos[n]; t < g.facepos[n + 1]; ++t)
                        if (g.neighbour[t] == c) { face_indices_[h] = face_indices_[t]; break; }
                }
            }
        num_faces_ = count;
    }
};

}  // namespace Opm

#endif

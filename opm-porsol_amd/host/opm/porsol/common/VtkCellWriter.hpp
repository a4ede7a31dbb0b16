// Cell data on hexahedral grids as an ASCII VTK unstructured-grid file
// (.vtu), with the calls of Dune::VTKWriter that the reference's drivers make
// (SimulatorUtilities.hpp:271-284, examples/mimetic_solver_test.cpp:124-131):
// addCellData(values, name[, components]) and write(name) -> name.vtu.
//
// Every cell is written with its own eight corner points (Dune's
// non-conforming mode), which also represents faulted corner-point grids.
// Values are written as Float64 with 17 significant digits, so a reader gets
// the solver's doubles back exactly.
#ifndef OPM_PORSOL_AMD_VTKCELLWRITER_HPP
#define OPM_PORSOL_AMD_VTKCELLWRITER_HPP

#include <cstdio>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "CornerPointGrid.hpp"
#include "GridInterfaceEuler.hpp"

namespace Opm {

namespace VTK {
enum OutputType { ascii, base64, appendedraw, appendedbase64 };
}

namespace vtk_detail {
template <class G, class = void> struct has_corners : std::false_type {};
template <class G>
struct has_corners<G, std::void_t<decltype(std::declval<const G &>().cellCorners(0))>> : std::true_type {};
template <class G, class = void> struct has_box : std::false_type {};
template <class G>
struct has_box<G, std::void_t<decltype(std::declval<const G &>().dims()),
                              decltype(std::declval<const G &>().cellSize())>> : std::true_type {};

// Eight corners of cell c in VTK hexahedron order (one ring of four, then the
// ring opposite, each corner above the matching one of the first ring).
template <class G>
void cell_corners(const G &grid, int c, double out[24]) {
    if constexpr (has_corners<G>::value) {
        const double *p = grid.cellCorners(c);
        for (int i = 0; i < 24; ++i) out[i] = p[i];
    } else if constexpr (has_box<G>::value) {
        const int *n = grid.dims();
        const double *h = grid.cellSize();
        const int ijk[3] = {c % n[0], (c / n[0]) % n[1], c / (n[0] * n[1])};
        static const int ring[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
        for (int layer = 0; layer < 2; ++layer)
            for (int v = 0; v < 4; ++v) {
                double *p = out + 3 * (4 * layer + v);
                p[0] = (ijk[0] + ring[v][0]) * h[0];
                p[1] = (ijk[1] + ring[v][1]) * h[1];
                p[2] = (ijk[2] + layer) * h[2];
            }
    } else {
        (void)grid; (void)c; (void)out;
        throw std::runtime_error("VTK output: the grid type gives no cell corners");
    }
}
}  // namespace vtk_detail

template <class Grid>
class VtkCellWriter {
public:
    explicit VtkCellWriter(const Grid &grid) : grid_(grid) {}
    /// values holds components consecutive per cell (num_cells * components).
    void addCellData(const std::vector<double> &values, const std::string &name, int components = 1) {
        if (components < 1 || values.size() != (size_t)components * grid_.numberOfCells())
            throw std::invalid_argument("VtkCellWriter::addCellData: '" + name + "' has " +
                                        std::to_string(values.size()) + " values for " +
                                        std::to_string(grid_.numberOfCells()) + " cells");
        fields_.push_back(Field{name, components, values});
    }
    /// Writes name.vtu; returns the file name.
    std::string write(const std::string &name, VTK::OutputType = VTK::ascii) const {
        const std::string path = name + ".vtu";
        FILE *f = std::fopen(path.c_str(), "w");
        if (!f) throw std::runtime_error("Could not open file " + path);
        const int nc = grid_.numberOfCells();
        std::fprintf(f, "<?xml version=\"1.0\"?>\n"
                        "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
                        "  <UnstructuredGrid>\n"
                        "    <Piece NumberOfPoints=\"%lld\" NumberOfCells=\"%d\">\n"
                        "      <CellData>\n",
                     8LL * nc, nc);
        for (const Field &fd : fields_) {
            std::fprintf(f, "        <DataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"%d\" "
                            "format=\"ascii\">\n",
                         fd.name.c_str(), fd.components);
            for (size_t i = 0; i < fd.values.size(); ++i)
                std::fprintf(f, "%.17g%c", fd.values[i], (i + 1) % fd.components == 0 ? '\n' : ' ');
            std::fprintf(f, "        </DataArray>\n");
        }
        std::fprintf(f, "      </CellData>\n      <Points>\n"
                        "        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n");
        double x[24];
        for (int c = 0; c < nc; ++c) {
            vtk_detail::cell_corners(grid_, c, x);
            for (int v = 0; v < 8; ++v) std::fprintf(f, "%.17g %.17g %.17g\n", x[3 * v], x[3 * v + 1], x[3 * v + 2]);
        }
        std::fprintf(f, "        </DataArray>\n      </Points>\n      <Cells>\n"
                        "        <DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n");
        for (int c = 0; c < nc; ++c)
            for (int v = 0; v < 8; ++v) std::fprintf(f, "%d%c", 8 * c + v, v == 7 ? '\n' : ' ');
        std::fprintf(f, "        </DataArray>\n"
                        "        <DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n");
        for (int c = 0; c < nc; ++c) std::fprintf(f, "%d\n", 8 * (c + 1));
        std::fprintf(f, "        </DataArray>\n"
                        "        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
        for (int c = 0; c < nc; ++c) std::fprintf(f, "12\n");   // VTK_HEXAHEDRON
        std::fprintf(f, "        </DataArray>\n      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n");
        const bool ok = std::ferror(f) == 0;
        if (std::fclose(f) != 0 || !ok) throw std::runtime_error("Could not write file " + path);
        return path;
    }

private:
    struct Field {
        std::string name;
        int components;
        std::vector<double> values;
    };
    const Grid &grid_;
    std::vector<Field> fields_;
};

}  // namespace Opm

#endif

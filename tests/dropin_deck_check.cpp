// Host-only check of the C++ ECLIPSE set-up and output: setupGridAndProps
// over a deck (or a generated box), then writeVtkOutput and writeField for a
// solution object that holds host data only (pressure = 1e5 + cell index,
// outflux = sin(half-face index), saturation = (c mod 10) / 10), so no GPU is
// needed.  Prints the grid summary the Python test compares.
//   dropin_deck_check <out prefix> key=value...
#include <cmath>
#include <cstdio>
#include <iostream>
#include <vector>

#include <opm/porsol/common/GridInterfaceEuler.hpp>
#include <opm/porsol/common/ReservoirPropertyCapillary.hpp>
#include <opm/porsol/common/SimulatorUtilities.hpp>
#include <opm/porsol/common/setupGridAndProps.hpp>

#include "../opm-porsol_amd/examples/params.hpp"

typedef Opm::GridInterfaceEuler<Dune::CpGrid> GI;

struct HostSolution {
    const GI *g;
    double pressure(const GI::CellIterator &c) const { return 1e5 + c->index(); }
    double outflux(const GI::CellIterator::FaceIterator &f) const { return std::sin(double(f->halfFaceIndex())); }
};

int main(int argc, char **argv) try {
    if (argc < 2) return 2;
    const std::string prefix = argv[1];
    Params param(argc - 1, argv + 1);
    Dune::CpGrid grid;
    Opm::ReservoirPropertyCapillary<3> rp;
    Opm::setupGridAndProps(param, grid, rp);
    GI g(grid);
    const int nc = g.numberOfCells();
    std::printf("cells %d half_faces %d faces %d logical %d %d %d kind %d unique %d\n", nc,
                grid.numberOfHalfFaces(), g.numberOfFaces(), grid.logicalCartesianSize()[0],
                grid.logicalCartesianSize()[1], grid.logicalCartesianSize()[2], rp.permeabilityKind(),
                int(grid.uniqueBoundaryIds()));
    FILE *f = std::fopen((prefix + "-grid.txt").c_str(), "w");
    for (int c = 0; c < nc; ++c) {
        const auto K = rp.permeability(c);
        std::fprintf(f, "%d %.17g %.17g %.17g %.17g %.17g\n", grid.globalCell()[c], grid.cell_volume[c], K(0, 0),
                     K(1, 1), K(2, 2), rp.porosity(c));
    }
    for (int h = 0; h < grid.numberOfHalfFaces(); ++h) std::fprintf(f, "bid %d\n", grid.boundary_id[h]);
    std::fclose(f);
    std::vector<double> sat(nc);
    for (int c = 0; c < nc; ++c) sat[c] = (c % 10) / 10.0;
    HostSolution sol{&g};
    Opm::writeVtkOutput(g, rp, sol, sat, prefix);
    std::vector<double> field(nc);
    for (int c = 0; c < nc; ++c) field[c] = 0.5 * c;
    Opm::writeField(field, prefix + "-field.txt");
    try {
        Opm::writeField(field, "/nonexistent-dir/x.txt");
    } catch (const std::exception &e) {
        std::printf("writeField error: %s\n", e.what());
    }
    return 0;
} catch (const std::exception &e) {
    std::printf("error: %s\n", e.what());
    return 1;
}

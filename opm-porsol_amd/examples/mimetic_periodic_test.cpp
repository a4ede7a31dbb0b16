// Periodic drop-in driver after the reference's examples/mimetic_periodic_test.cpp:
// periodic in x and y with a 1 bar drop along x, no-flow in z, unique boundary
// ids, unit permeability.  Prints the net flux through the x- side and the
// largest twin-flux change of postProcessFluxes().
//   mimetic_periodic_test nx=8 ny=8 nz=4 linear_solver_type=1
// print_system=<prefix> and print_stats=1 as in known_answer_test.
#include <cmath>
#include <iostream>
#include <string>
#include <vector>

#include <opm/porsol/common/BoundaryConditions.hpp>
#include <opm/porsol/common/GridInterfaceEuler.hpp>
#include <opm/porsol/common/PeriodicHelpers.hpp>
#include <opm/porsol/common/ReservoirPropertyCapillary.hpp>
#include <opm/porsol/common/Units.hpp>
#include <opm/porsol/mimetic/IncompFlowSolverHybrid.hpp>
#include <opm/porsol/mimetic/MimeticIPEvaluator.hpp>

#include "params.hpp"

int main(int argc, char **argv) try {
    using namespace Opm;
    Params param(argc, argv);
    CartesianGrid grid(param.geti("nx", 8), param.geti("ny", 8), param.geti("nz", 4));
    grid.setUniqueBoundaryIds(true);
    typedef GridInterfaceEuler<CartesianGrid> GI;
    typedef BasicBoundaryConditions<true, false> BCs;
    typedef ReservoirPropertyCapillary<3> RI;
    GI g(grid);
    std::array<FlowBC, 6> cond = {{FlowBC(FlowBC::Periodic, 1.0 * unit::barsa),
                                   FlowBC(FlowBC::Periodic, -1.0 * unit::barsa),
                                   FlowBC(FlowBC::Periodic, 0.0), FlowBC(FlowBC::Periodic, 0.0),
                                   FlowBC(FlowBC::Neumann, 0.0), FlowBC(FlowBC::Neumann, 0.0)}};
    BCs fbc;
    createPeriodic(fbc, g, cond);
    RI r;
    r.init(g.numberOfCells(), 0.2, 1.0);
    IncompFlowSolverHybrid<GI, RI, BCs, MimeticIPEvaluator> solver;
    solver.setPrecision(param.geti("precision", MHS_FP64));
    Vec3 gravity(0.0);
    solver.init(g, r, gravity, fbc);
    std::vector<double> src(g.numberOfCells(), 0.0), sat(g.numberOfCells(), 0.0);
    solver.solve(r, sat, fbc, src, param.get("tolerance", 1e-12), 0, param.geti("linear_solver_type", 1));
    if (param.geti("print_stats", 0)) solver.printStats(std::cout);
    const std::string prefix = param.gets("print_system", "");
    if (!prefix.empty()) solver.printSystem(prefix);
    const auto &soln = solver.getSolution();
    double qx = 0.0;
    for (GI::CellIterator c = g.cellbegin(); c != g.cellend(); ++c)
        for (auto f = c->facebegin(); f != c->faceend(); ++f)
            if (f->boundary() && f->normal()[0] < -0.5) qx += soln.outflux(f);
    const double mm = solver.postProcessFluxes();
    std::cout.precision(15);
    std::cout << "net x- outflux: " << qx << "\nmax twin change: " << mm << '\n';
    return 0;
} catch (const std::exception &e) {
    std::cerr << "Program threw an exception: " << e.what() << "\n";
    return 1;
}

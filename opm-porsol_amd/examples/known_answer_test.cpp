// Manufactured-solution test of the drop-in solver, after the reference's
// examples/known_answer_test.cpp: unit cube N^3, K = 1, unit viscosities,
// u = sin(2 pi x) cos(2 pi y) z, Dirichlet u(face centroid) on every boundary
// face, source -Lu * volume.  Prints the L1 / L2 / Linf error densities in the
// reference's format (examples/timing_results.txt records them for N = 50, 60).
//   known_answer_test nx=60 tolerance=1e-8 linear_solver_type=1 precision=0
// print_system=<prefix> writes the assembled system after the solve
// (IncompFlowSolverHybrid::printSystem: <prefix>-mat.dat, <prefix>-rhs.dat);
// print_stats=1 prints IncompFlowSolverHybrid::printStats to stdout.
#include <chrono>
#include <cmath>
#include <iostream>
#include <string>
#include <vector>

#include <opm/porsol/common/BoundaryConditions.hpp>
#include <opm/porsol/common/GridInterfaceEuler.hpp>
#include <opm/porsol/common/ReservoirPropertyCapillary.hpp>
#include <opm/porsol/common/SimulatorUtilities.hpp>
#include <opm/porsol/mimetic/IncompFlowSolverHybrid.hpp>
#include <opm/porsol/mimetic/MimeticIPEvaluator.hpp>

#include "params.hpp"

typedef Opm::Vec3 Vec;
static const double PI = 3.14159265358979323846264338327950288;

double u(const Vec &x) { return std::sin(2 * PI * x[0]) * std::cos(2 * PI * x[1]) * x[2]; }
double Lu(const Vec &x) { return -2 * 2 * PI * 2 * PI * std::sin(2 * PI * x[0]) * std::cos(2 * PI * x[1]) * x[2]; }

namespace Opm {
template <class BoundaryFunc>
class FunctionBoundaryConditions : public PeriodicConditionHandler {
public:
    explicit FunctionBoundaryConditions(BoundaryFunc bfunc) : bfunc_(bfunc) {}
    template <class BoundaryFace>
    FlowBC flowCond(const BoundaryFace &bf) const {
        return FlowBC(FlowBC::Dirichlet, bfunc_(bf.centroid()));
    }

private:
    BoundaryFunc bfunc_;
};
}  // namespace Opm

int main(int argc, char **argv) try {
    Params param(argc, argv);
    const int nx = param.geti("nx", 20), ny = param.geti("ny", nx), nz = param.geti("nz", nx);
    Opm::CartesianGrid grid(nx, ny, nz, param.get("dx", 1.0 / nx), param.get("dy", 1.0 / ny),
                            param.get("dz", 1.0 / nz));
    typedef Opm::GridInterfaceEuler<Opm::CartesianGrid> GI;
    GI g(grid);
    Opm::ReservoirPropertyCapillary<3> res_prop;
    res_prop.init(g.numberOfCells(), 1.0, 1.0);
    res_prop.setViscosities(1.0, 1.0);

    typedef double (*SolutionFuncPtr)(const Vec &);
    typedef Opm::FunctionBoundaryConditions<SolutionFuncPtr> FBC;
    typedef Opm::IncompFlowSolverHybrid<GI, Opm::ReservoirPropertyCapillary<3>, FBC,
                                        Opm::MimeticIPEvaluator> FlowSolver;
    FlowSolver solver;
    solver.setPrecision(param.geti("precision", MHS_FP64));
    FBC flow_bc(&u);
    Vec gravity(0.0);

    std::cout << "topology is sane: " << Opm::topologyIsSane(g) << std::endl;
    std::cout << "========== Init pressure solver =============" << std::endl;
    auto t0 = std::chrono::steady_clock::now();
    solver.init(g, res_prop, gravity, flow_bc);
    auto t1 = std::chrono::steady_clock::now();
    std::cout << "========== Time in seconds: " << std::chrono::duration<double>(t1 - t0).count()
              << " =============" << std::endl;

    std::vector<double> src(g.numberOfCells(), 0.0), sat(g.numberOfCells(), 0.0);
    for (GI::CellIterator c = g.cellbegin(); c != g.cellend(); ++c)
        src[c->index()] = -Lu(c->centroid()) * c->volume();

    std::cout << "========== Starting pressure solve =============" << std::endl;
    t0 = std::chrono::steady_clock::now();
    solver.solve(res_prop, sat, flow_bc, src, param.get("tolerance", 1e-8), 1,
                 param.geti("linear_solver_type", 1));
    t1 = std::chrono::steady_clock::now();
    std::cout << "========== Time in seconds: " << std::chrono::duration<double>(t1 - t0).count()
              << " =============" << std::endl;
    {
        const mhs_solve_report &r = solver.lastReport();
        std::cout << "PCG iterations: " << r.iterations << ", preconditioner: "
                  << (r.precond ? "multigrid" : "Jacobi") << ", levels: " << r.levels
                  << ", reduction: " << r.reduction << std::endl;
    }

    if (param.geti("print_stats", 0)) solver.printStats(std::cout);
    const std::string prefix = param.gets("print_system", "");
    if (!prefix.empty()) solver.printSystem(prefix);

    std::vector<double> cell_pressure;
    Opm::getCellPressure(cell_pressure, g, solver.getSolution());
    double l1 = 0.0, l2 = 0.0, linf = 0.0, totv = 0.0;
    for (GI::CellIterator c = g.cellbegin(); c != g.cellend(); ++c) {
        const double diff = u(c->centroid()) - cell_pressure[c->index()];
        const double v = c->volume();
        l1 += std::fabs(diff * v);
        l2 += diff * diff * v;
        linf = std::max(std::fabs(diff), linf);
        totv += v;
    }
    l2 = std::sqrt(l2);
    std::cout.precision(6);
    std::cout << "\n\n"
              << "\n     L1 error density: " << l1 / totv
              << "\n     L2 error density: " << l2 / totv
              << "\n     Linf error:       " << linf << "\n\n\n";
    return 0;
} catch (const std::exception &e) {
    std::cerr << "Program threw an exception: " << e.what() << "\n";
    return 1;
}

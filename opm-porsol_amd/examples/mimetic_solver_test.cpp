// Pressure solve on an ECLIPSE deck or a generated box, after the reference's
// examples/mimetic_solver_test.cpp (:76-131, :146-175): grid and rock from
// setupGridAndProps, Dirichlet 100 bar on the top (boundary id 5) and 0 on the
// bottom (id 6), no-flow elsewhere, no gravity, zero saturation, one solve
// (tolerance 5e-9, linear solver type 3) with IncompFlowSolverHybrid, then
// the cell velocity and pressure written to testsolution-0.vtu.
//   mimetic_solver_test fileformat=eclipse filename=deck.grdecl
//   mimetic_solver_test fileformat=cartesian nx=10 ny=10 nz=10
// Extensions: tolerance=<t> and precision=<0|1> change the solve;
// output=<name> names the VTU file; full_output=<name> also writes name.vtu
// with all fields of writeVtkOutput and name-pressure.txt through writeField.
#include <cstdio>
#include <iostream>
#include <vector>

#include <opm/porsol/common/BoundaryConditions.hpp>
#include <opm/porsol/common/GridInterfaceEuler.hpp>
#include <opm/porsol/common/ReservoirPropertyCapillary.hpp>
#include <opm/porsol/common/SimulatorUtilities.hpp>
#include <opm/porsol/common/VtkCellWriter.hpp>
#include <opm/porsol/common/setupGridAndProps.hpp>
#include <opm/porsol/mimetic/IncompFlowSolverHybrid.hpp>
#include <opm/porsol/mimetic/MimeticIPEvaluator.hpp>

#include "params.hpp"

template <int dim, class GI, class RI>
void test_flowsolver(const GI &g, const RI &r, const Params &param) {
    typedef Opm::BasicBoundaryConditions<true, false> FBC;
    typedef Opm::IncompFlowSolverHybrid<GI, RI, FBC, Opm::MimeticIPEvaluator> FlowSolver;
    FlowSolver solver;
    solver.setPrecision(param.geti("precision", MHS_FP64));
    typedef Opm::FlowBC BC;
    FBC flow_bc(7);
    flow_bc.flowCond(5) = BC(BC::Dirichlet, 100.0 * Opm::unit::barsa);
    flow_bc.flowCond(6) = BC(BC::Dirichlet, 0.0 * Opm::unit::barsa);
    typename GI::Vector gravity(0.0);
    solver.init(g, r, gravity, flow_bc);
    std::vector<double> src(g.numberOfCells(), 0.0);
    std::vector<double> sat(g.numberOfCells(), 0.0);
    solver.solve(r, sat, flow_bc, src, param.get("tolerance", 5e-9), 3, 1);

    typedef typename FlowSolver::SolutionType FlowSolution;
    FlowSolution soln = solver.getSolution();
    std::vector<typename GI::Vector> cell_velocity;
    estimateCellVelocity(cell_velocity, g, soln);
    std::vector<double> cell_velocity_flat;
    for (const auto &v : cell_velocity) cell_velocity_flat.insert(cell_velocity_flat.end(), v.begin(), v.end());
    std::vector<double> cell_pressure;
    getCellPressure(cell_pressure, g, soln);
    Opm::VtkCellWriter<typename GI::GridType> vtkwriter(g.grid());
    vtkwriter.addCellData(cell_velocity_flat, "velocity", dim);
    vtkwriter.addCellData(cell_pressure, "pressure");
    const std::string name = vtkwriter.write(param.gets("output", "testsolution-0"), Opm::VTK::ascii);
    const mhs_solve_report &rep = solver.lastReport();
    std::printf("cells: %d  iterations: %d  preconditioner: %s  wrote %s\n", g.numberOfCells(), rep.iterations,
                rep.precond ? "multigrid" : "Jacobi", name.c_str());
    const std::string full = param.gets("full_output", "");
    if (!full.empty()) {
        Opm::writeVtkOutput(g, r, soln, sat, full);
        Opm::writeField(cell_pressure, full + "-pressure.txt");
    }
}

int main(int argc, char **argv) try {
    using namespace Opm;
    Params param(argc, argv);
    Dune::CpGrid grid;
    ReservoirPropertyCapillary<3> res_prop;
    setupGridAndProps(param, grid, res_prop);
    GridInterfaceEuler<Dune::CpGrid> g(grid);
    test_flowsolver<3>(g, res_prop, param);
    return 0;
} catch (const std::exception &e) {
    std::cerr << "Program threw an exception: " << e.what() << "\n";
    return 1;
}

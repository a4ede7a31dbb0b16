// IMPES driver after the reference's examples/SimulatorTester.hpp run()
// (:59-107): per step one pressure solve with the mobilities of the current
// saturation (IncompFlowSolverHybrid), then one explicit upstream transport
// solve over the step (EulerUpstream), both on the GPU.
// Grid nx x ny x nz of 2 x 2 x 1 m cells, K = 100 mD exp(sin(0.37 c)) I
// (cell index c), porosity 0.2, Dirichlet 2 bar on x- and 0 on x+ (boundary
// ids 1, 2), water (saturation 1) entering through x-, gravity along +z,
// the default fluid (1 / 3 cP, 1013.9 / 834.7 kg/m^3).  Prints per step the
// small steps and the water in place; with out=<file>, the final saturation.
//   impes_example nx=12 ny=8 nz=4 steps=3 stepsize=2e5 method_capillary=0
#include <cmath>
#include <cstdio>
#include <iostream>
#include <vector>

#include <opm/porsol/common/BoundaryConditions.hpp>
#include <opm/porsol/common/GridInterfaceEuler.hpp>
#include <opm/porsol/common/ReservoirPropertyCapillary.hpp>
#include <opm/porsol/common/Units.hpp>
#include <opm/porsol/euler/EulerUpstream.hpp>
#include <opm/porsol/mimetic/IncompFlowSolverHybrid.hpp>
#include <opm/porsol/mimetic/MimeticIPEvaluator.hpp>

#include "params.hpp"

int main(int argc, char **argv) try {
    using namespace Opm;
    Params param(argc, argv);
    CartesianGrid grid(param.geti("nx", 12), param.geti("ny", 8), param.geti("nz", 4), 2.0, 2.0, 1.0);
    typedef GridInterfaceEuler<CartesianGrid> GI;
    typedef BasicBoundaryConditions<true, true> BCs;
    typedef ReservoirPropertyCapillary<3> RI;
    GI g(grid);
    const int nc = g.numberOfCells();
    BCs bcs(7);
    bcs.flowCond(1) = FlowBC(FlowBC::Dirichlet, 2.0 * unit::barsa);
    bcs.flowCond(2) = FlowBC(FlowBC::Dirichlet, 0.0);
    for (int i = 0; i < 7; ++i) bcs.satCond(i) = SatBC(SatBC::Dirichlet, i == 1 ? 1.0 : 0.0);
    std::vector<double> perm(9 * (size_t)nc, 0.0);
    for (int c = 0; c < nc; ++c)
        for (int d = 0; d < 3; ++d) perm[9 * (size_t)c + 4 * d] = 100.0 * unit::millidarcy * std::exp(std::sin(0.37 * c));
    RI r;
    r.init(nc, 0.2, perm);
    Vec3 gravity(0.0);
    gravity[2] = 9.81;
    IncompFlowSolverHybrid<GI, RI, BCs, MimeticIPEvaluator> flow;
    flow.setPrecision(param.geti("precision", MHS_FP64));
    flow.init(g, r, gravity, bcs);
    EulerUpstream<GI, RI, BCs> transport;
    transport.init(param, g, r, bcs);
    std::vector<double> sat(nc, 0.0), src(nc, 0.0), inj(nc, 0.0);
    const int steps = param.geti("steps", 3);
    const double stepsize = param.get("stepsize", 2.0e5);
    for (int i = 0; i < steps; ++i) {
        flow.solve(r, sat, bcs, src, param.get("tolerance", 1e-12), 0);
        transport.transportSolve(sat, stepsize, gravity, flow.getSolution(), inj);
        double water = 0.0, smax = 0.0;
        for (int c = 0; c < nc; ++c) {
            water += sat[c] * 0.2 * 4.0;
            smax = std::max(smax, sat[c]);
        }
        std::printf("step %d: small steps %d, repeats %d, water in place %.12e, max saturation %.12f\n", i,
                    transport.report().small_steps, transport.report().repeats, water, smax);
    }
    const std::string out = param.gets("out", "");
    if (!out.empty()) {
        FILE *f = std::fopen(out.c_str(), "w");
        if (!f) throw std::runtime_error("cannot write " + out);
        for (int c = 0; c < nc; ++c) std::fprintf(f, "%.17e\n", sat[c]);
        std::fclose(f);
    }
    return 0;
} catch (const std::exception &e) {
    std::cerr << "Program threw an exception: " << e.what() << "\n";
    return 1;
}

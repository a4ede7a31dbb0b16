// The second inner-product plug-in: IncompFlowSolverHybrid with
// MimeticIPAnisoRelpermEvaluator and ReservoirPropertyCapillaryAnisotropicRelperm
// (tensor phase mobilities), beside the scalar MimeticIPEvaluator with
// ReservoirPropertyCapillary on the same grid, rock and saturation.
// Without rock tables the reference's tensor mobilities are kr/mu I, so the
// two solves must agree; the driver prints their largest flux difference.
// Grid nx x ny x nz of 2 x 2 x 1 m cells, rotated anisotropic K per cell,
// Dirichlet 2 bar / 0 on x- / x+, gravity along +z, s(c) = 0.5 + 0.4 sin(c).
//   aniso_relperm_example nx=16 ny=12 nz=8
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <vector>

#include <opm/porsol/common/BoundaryConditions.hpp>
#include <opm/porsol/common/GridInterfaceEuler.hpp>
#include <opm/porsol/common/ReservoirPropertyCapillary.hpp>
#include <opm/porsol/common/ReservoirPropertyCapillaryAnisotropicRelperm.hpp>
#include <opm/porsol/common/Units.hpp>
#include <opm/porsol/mimetic/IncompFlowSolverHybrid.hpp>
#include <opm/porsol/mimetic/MimeticIPAnisoRelpermEvaluator.hpp>
#include <opm/porsol/mimetic/MimeticIPEvaluator.hpp>

#include "params.hpp"

int main(int argc, char **argv) try {
    using namespace Opm;
    Params param(argc, argv);
    CartesianGrid grid(param.geti("nx", 16), param.geti("ny", 12), param.geti("nz", 8), 2.0, 2.0, 1.0);
    typedef GridInterfaceEuler<CartesianGrid> GI;
    typedef BasicBoundaryConditions<true, false> BCs;
    GI g(grid);
    const int nc = g.numberOfCells();
    BCs bcs(7);
    bcs.flowCond(1) = FlowBC(FlowBC::Dirichlet, 2.0 * unit::barsa);
    bcs.flowCond(2) = FlowBC(FlowBC::Dirichlet, 0.0);
    // K = R diag(k, 0.5 k, 0.1 k) R' with a per-cell rotation about z
    std::vector<double> perm(9 * (size_t)nc);
    for (int c = 0; c < nc; ++c) {
        const double k = 100.0 * unit::millidarcy * std::exp(std::sin(0.37 * c)), a = 0.3 * std::cos(0.11 * c);
        const double ca = std::cos(a), sa = std::sin(a), d[3] = {k, 0.5 * k, 0.1 * k};
        const double R[3][3] = {{ca, -sa, 0.0}, {sa, ca, 0.0}, {0.0, 0.0, 1.0}};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double v = 0.0;
                for (int m = 0; m < 3; ++m) v += R[i][m] * d[m] * R[j][m];
                perm[9 * (size_t)c + 3 * i + j] = v;
            }
    }
    ReservoirPropertyCapillary<3> rs;
    rs.init(nc, 0.2, perm);
    ReservoirPropertyCapillaryAnisotropicRelperm<3> ra;
    ra.init(nc, 0.2, perm);
    Vec3 gravity(0.0);
    gravity[2] = 9.81;
    std::vector<double> sat(nc), src(nc, 0.0);
    for (int c = 0; c < nc; ++c) sat[c] = 0.5 + 0.4 * std::sin(double(c));
    const double tol = param.get("tolerance", 1e-13);
    IncompFlowSolverHybrid<GI, ReservoirPropertyCapillary<3>, BCs, MimeticIPEvaluator> scalar;
    scalar.init(g, rs, gravity, bcs);
    scalar.solve(rs, sat, bcs, src, tol, 0);
    IncompFlowSolverHybrid<GI, ReservoirPropertyCapillaryAnisotropicRelperm<3>, BCs, MimeticIPAnisoRelpermEvaluator>
        tensor;
    tensor.init(g, ra, gravity, bcs);
    tensor.solve(ra, sat, bcs, src, tol, 0);
    const auto &a = scalar.getSolution();
    const auto &b = tensor.getSolution();
    double dmax = 0.0, vmax = 0.0, pdiff = 0.0, pmax = 0.0;
    for (GI::CellIterator c = g.cellbegin(); c != g.cellend(); ++c) {
        for (auto f = c->facebegin(); f != c->faceend(); ++f) {
            dmax = std::max(dmax, std::fabs(a.outflux(f) - b.outflux(f)));
            vmax = std::max(vmax, std::fabs(a.outflux(f)));
        }
        pdiff = std::max(pdiff, std::fabs(a.pressure(c) - b.pressure(c)));
        pmax = std::max(pmax, std::fabs(a.pressure(c)));
    }
    std::printf("max |flux| %.6e\nflux difference (relative): %.6e\npressure difference (relative): %.6e\n", vmax,
                dmax / vmax, pdiff / pmax);
    return 0;
} catch (const std::exception &e) {
    std::cerr << "Program threw an exception: " << e.what() << "\n";
    return 1;
}

// Solution post-processing helpers with the behaviour of
// opm/porsol/common/SimulatorUtilities.hpp (estimateCellVelocity :59-81,
// getCellPressure :178-188).
#ifndef OPM_PORSOL_AMD_SIMULATORUTILITIES_HPP
#define OPM_PORSOL_AMD_SIMULATORUTILITIES_HPP

#include <vector>

namespace Opm {

/// Cell velocity sum_f flux_f (x_f - x_c) / |c| (halfFaceFluxToCellVelocity).
template <class GridInterface, class FlowSol>
void estimateCellVelocity(std::vector<typename GridInterface::Vector> &cell_velocity,
                          const GridInterface &g, const FlowSol &flow_solution) {
    cell_velocity.clear();
    cell_velocity.resize(g.numberOfCells());
    for (typename GridInterface::CellIterator c = g.cellbegin(); c != g.cellend(); ++c) {
        typename GridInterface::Vector cell_v(0.0);
        for (auto f = c->facebegin(); f != c->faceend(); ++f) {
            const double flux = flow_solution.outflux(f);
            typename GridInterface::Vector v = f->centroid();
            v -= c->centroid();
            v *= flux / c->volume();
            cell_v += v;
        }
        cell_velocity[c->index()] = cell_v;
    }
}

template <class GridInterface, class FlowSol>
void getCellPressure(std::vector<double> &cell_pressure, const GridInterface &g,
                     const FlowSol &flow_solution) {
    cell_pressure.clear();
    cell_pressure.resize(g.numberOfCells());
    for (typename GridInterface::CellIterator c = g.cellbegin(); c != g.cellend(); ++c)
        cell_pressure[c->index()] = flow_solution.pressure(c);
}

}  // namespace Opm

#endif

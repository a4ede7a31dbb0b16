// Solution post-processing and output with the behaviour of
// opm/porsol/common/SimulatorUtilities.hpp: estimateCellVelocity (:59-81),
// computePhaseVelocities (:153-169), getCellPressure (:178-210),
// computeCapPressure (:219-228), writeVtkOutput (:232-285) and writeField
// (:288-297).
//
// estimateCellVelocity runs on the GPU (mhs_estimate_cell_velocity) for a
// solution that still lives on it (FlowSolution::deviceHandle() non-null);
// for other solution types, and for a copy taken before a later solve, it
// reads the host fluxes the solution holds.
#ifndef OPM_PORSOL_AMD_SIMULATORUTILITIES_HPP
#define OPM_PORSOL_AMD_SIMULATORUTILITIES_HPP

#include <algorithm>
#include <array>
#include <cassert>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "mimetic_hybrid.h"
#include "VtkCellWriter.hpp"

namespace Opm {

namespace sim_detail {
template <class F, class = void> struct has_device_handle : std::false_type {};
template <class F>
struct has_device_handle<F, std::void_t<decltype(std::declval<const F &>().deviceHandle())>> : std::true_type {};
}  // namespace sim_detail

/// Cell velocity sum_f flux_f (x_f - x_c) / |c| (halfFaceFluxToCellVelocity).
template <class GridInterface, class FlowSol>
void estimateCellVelocity(std::vector<typename GridInterface::Vector> &cell_velocity,
                          const GridInterface &g, const FlowSol &flow_solution) {
    typedef typename GridInterface::Vector Vec;
    const int nc = g.numberOfCells();
    cell_velocity.clear();
    cell_velocity.resize(nc);
    if constexpr (sim_detail::has_device_handle<FlowSol>::value) {
        if (mhs_solver *h = flow_solution.deviceHandle()) {
            mhs_dims d;
            if (mhs_get_dims(h, &d) != MHS_OK || d.num_cells != nc)
                throw std::runtime_error("estimateCellVelocity: the solution belongs to another grid");
            std::vector<double> v(3 * (size_t)nc);
            if (mhs_estimate_cell_velocity(h, 0, v.data()) != MHS_OK)
                throw std::runtime_error(std::string("estimateCellVelocity: ") + mhs_last_error(h));
            int pos = 0;   // the solver numbers cells in iteration order
            for (typename GridInterface::CellIterator c = g.cellbegin(); c != g.cellend(); ++c, ++pos)
                cell_velocity[c->index()] = Vec(v[3 * (size_t)pos], v[3 * (size_t)pos + 1], v[3 * (size_t)pos + 2]);
            return;
        }
    }
    for (typename GridInterface::CellIterator c = g.cellbegin(); c != g.cellend(); ++c) {
        Vec cell_v(0.0);
        for (auto f = c->facebegin(); f != c->faceend(); ++f) {
            const double flux = flow_solution.outflux(f);
            Vec v = f->centroid();
            v -= c->centroid();
            v *= flux / c->volume();
            cell_v += v;
        }
        cell_velocity[c->index()] = cell_v;
    }
}

/// Water and oil parts of the cell velocity by the fractional flow of water.
template <class ReservoirProperty, class Vec>
void computePhaseVelocities(std::vector<Vec> &phase_velocity_water, std::vector<Vec> &phase_velocity_oil,
                            const ReservoirProperty &res_prop, const std::vector<double> &saturation,
                            const std::vector<Vec> &cell_velocity) {
    assert(saturation.size() == cell_velocity.size());
    const int num_cells = int(saturation.size());
    phase_velocity_water = cell_velocity;
    phase_velocity_oil = cell_velocity;
    for (int i = 0; i < num_cells; ++i) {
        const double f = res_prop.fractionalFlow(i, saturation[i]);
        phase_velocity_water[i] *= f;
        phase_velocity_oil[i] *= (1.0 - f);
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

/// Cells outside my_partition get zero.
template <class GridInterface, class FlowSol>
void getCellPressure(std::vector<double> &cell_pressure, const GridInterface &g, const FlowSol &flow_solution,
                     const std::vector<int> &partition, const int my_partition) {
    cell_pressure.clear();
    cell_pressure.resize(g.numberOfCells());
    for (typename GridInterface::CellIterator c = g.cellbegin(); c != g.cellend(); ++c)
        cell_pressure[c->index()] = partition[c->index()] != my_partition ? 0.0 : flow_solution.pressure(c);
}

template <class ReservoirProperties>
void computeCapPressure(std::vector<double> &cap_pressure, const ReservoirProperties &rp,
                        const std::vector<double> &sat) {
    const int num_cells = int(sat.size());
    cap_pressure.resize(num_cells);
    for (int cell = 0; cell < num_cells; ++cell) cap_pressure[cell] = rp.capillaryPressure(cell, sat[cell]);
}

/// Writes filename.vtu with the cell fields "saturation", "pressure",
/// "capillary pressure", "fractional flow [water]", "velocity",
/// "phase velocity [water]" and "phase velocity [oil]".
template <class GridInterface, class ReservoirProperties, class FlowSol>
void writeVtkOutput(const GridInterface &ginterf, const ReservoirProperties &rp, const FlowSol &flowsol,
                    const std::vector<double> &saturation, const std::string &filename) {
    typedef typename GridInterface::Vector Vec;
    std::vector<Vec> cell_velocity;
    estimateCellVelocity(cell_velocity, ginterf, flowsol);
    std::array<std::vector<Vec>, 2> phase_velocities;
    computePhaseVelocities(phase_velocities[0], phase_velocities[1], rp, saturation, cell_velocity);
    auto flat = [](const std::vector<Vec> &v) {
        std::vector<double> out;
        out.reserve(3 * v.size());
        for (const Vec &x : v) out.insert(out.end(), x.begin(), x.end());
        return out;
    };
    std::vector<double> cell_pressure;
    getCellPressure(cell_pressure, ginterf, flowsol);
    std::vector<double> cap_pressure;
    computeCapPressure(cap_pressure, rp, saturation);
    const int num_cells = int(saturation.size());
    std::vector<double> fractional_flow(num_cells);
    for (int i = 0; i < num_cells; ++i) fractional_flow[i] = rp.fractionalFlow(i, saturation[i]);

    VtkCellWriter<typename GridInterface::GridType> vtkwriter(ginterf.grid());
    vtkwriter.addCellData(saturation, "saturation");
    vtkwriter.addCellData(cell_pressure, "pressure");
    vtkwriter.addCellData(cap_pressure, "capillary pressure");
    vtkwriter.addCellData(fractional_flow, "fractional flow [water]");
    vtkwriter.addCellData(flat(cell_velocity), "velocity", Vec::dimension);
    vtkwriter.addCellData(flat(phase_velocities[0]), "phase velocity [water]", Vec::dimension);
    vtkwriter.addCellData(flat(phase_velocities[1]), "phase velocity [oil]", Vec::dimension);
    vtkwriter.write(filename, VTK::ascii);
}

/// The field's size on the first line, then one value per line.
inline void writeField(const std::vector<double> &field, const std::string &filename) {
    std::ofstream os(filename.c_str());
    if (!os) throw std::runtime_error("Could not open file " + filename);
    os << field.size() << '\n';
    std::copy(field.begin(), field.end(), std::ostream_iterator<double>(os, "\n"));
}

}  // namespace Opm

#endif

// Grid and rock set-up from run parameters with the behaviour of
// opm/porsol/common/setupGridAndProps.hpp:66-130 (the Dune::CpGrid overload)
// and setupGridAndPropsEclipse (:136-157).
//
//   fileformat=eclipse    filename=<deck>  [perm_threshold_md=0]
//                         [viscosity1=0.001 viscosity2=0.003]
//   fileformat=cartesian  nx ny nz dx dy dz [default_poro=0.2]
//                         [default_perm_md=100]
//   use_unique_boundary_ids=false
//
// The parameter object is any type with the ParameterGroup accessors used
// there: getDefault<T>(name, default), get<T>(name) and has(name).
// "sintef_legacy" grids, rock tables (rock_list) and EGRID output are not
// supported and throw.
#ifndef OPM_PORSOL_AMD_SETUPGRIDANDPROPS_HPP
#define OPM_PORSOL_AMD_SETUPGRIDANDPROPS_HPP

#include <array>
#include <iostream>
#include <stdexcept>
#include <string>

#include "CornerPointGrid.hpp"
#include "EclipseDeck.hpp"
#include "Units.hpp"

namespace Opm {

template <template <int> class ResProp>
inline void setupGridAndPropsEclipse(const DeckConstPtr &deck, bool periodic_extension, bool turn_normals,
                                     bool clip_z, bool unique_bids, double perm_threshold,
                                     const std::string &rock_list, bool use_jfunction_scaling, double sigma,
                                     double theta, Dune::CpGrid &grid, ResProp<3> &res_prop) {
    grid.processEclipseFormat(deck, periodic_extension, turn_normals, clip_z);
    const std::string *rl_ptr = (rock_list == "no_list") ? nullptr : &rock_list;
    res_prop.init(deck, grid.globalCell(), perm_threshold, rl_ptr, use_jfunction_scaling, sigma, theta);
    if (unique_bids) grid.setUniqueBoundaryIds(true);
}

template <class ParameterGroup, template <int> class ResProp>
inline void setupGridAndProps(const ParameterGroup &param, Dune::CpGrid &grid, ResProp<3> &res_prop) {
    const std::string fileformat = param.template getDefault<std::string>("fileformat", "cartesian");
    if (fileformat == "eclipse") {
        const std::string ecl_file = param.template get<std::string>("filename");
        ParseContext parseContext;
        ParserPtr parser(new Parser());
        DeckConstPtr deck(parser->parseFile(ecl_file, parseContext));
        if (param.has("z_tolerance"))
            std::cerr << "****** Warning: z_tolerance parameter is obsolete, use PINCH in deck input instead\n";
        const bool periodic_extension = param.template getDefault<bool>("periodic_extension", false);
        const bool turn_normals = param.template getDefault<bool>("turn_normals", false);
        grid.processEclipseFormat(deck, periodic_extension, turn_normals);
        if (param.template getDefault<bool>("output_ecl", false))
            throw std::runtime_error("Saving to EGRID files is not yet implemented");
        const double perm_threshold_md = param.template getDefault<double>("perm_threshold_md", 0.0);
        const double perm_threshold = perm_threshold_md * prefix::milli * unit::darcy;
        const std::string rock_list = param.template getDefault<std::string>("rock_list", "no_list");
        const std::string *rl_ptr = (rock_list == "no_list") ? nullptr : &rock_list;
        if (param.has("viscosity1") || param.has("viscosity2")) {
            const double v1 = param.template getDefault<double>("viscosity1", 0.001);
            const double v2 = param.template getDefault<double>("viscosity2", 0.003);
            res_prop.setViscosities(v1, v2);
        }
        res_prop.init(deck, grid.globalCell(), perm_threshold, rl_ptr);
    } else if (fileformat == "cartesian") {
        const std::array<int, 3> dims = {{param.template getDefault<int>("nx", 1),
                                          param.template getDefault<int>("ny", 1),
                                          param.template getDefault<int>("nz", 1)}};
        const std::array<double, 3> cellsz = {{param.template getDefault<double>("dx", 1.0),
                                               param.template getDefault<double>("dy", 1.0),
                                               param.template getDefault<double>("dz", 1.0)}};
        grid.createCartesian(dims, cellsz);
        const double default_poro = param.template getDefault<double>("default_poro", 0.2);
        const double default_perm_md = param.template getDefault<double>("default_perm_md", 100.0);
        const double default_perm = default_perm_md * prefix::milli * unit::darcy;
        std::cerr << "Warning: For generated cartesian grids, we use uniform reservoir properties.\n";
        res_prop.init(grid.size(0), default_poro, default_perm);
    } else {
        throw std::runtime_error("Unknown file format string: " + fileformat);
    }
    if (param.template getDefault<bool>("use_unique_boundary_ids", false)) grid.setUniqueBoundaryIds(true);
}

}  // namespace Opm

#endif

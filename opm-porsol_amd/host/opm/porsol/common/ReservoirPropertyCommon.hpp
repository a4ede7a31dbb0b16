// Rock and fluid properties with the member API the pressure solver uses
// from opm/porsol/common/ReservoirPropertyCommon{,_impl}.hpp,
// ReservoirPropertyCapillary{,_impl}.hpp, ReservoirPropertyTracerFluid.hpp
// and ReservoirPropertyFixedMobility.hpp:
//   permeability(i) -> row-major 3x3 tensor view (RPC_impl :403-411)
//   phaseMobilities(i, s, mob), phaseDensities(i, rho), NumberOfPhases = 2
// Defaults follow ReservoirPropertyCommon_impl.hpp:257-262 (densities
// 1013.9 / 834.7 kg/m^3, viscosities 1 / 3 cP).  Without rock tables the
// relative permeabilities fall back to krw = s^2, kro = (1-s)^2
// (ReservoirPropertyCapillary_impl.hpp:117-168) and the capillary pressure to
// 1e5 (1 - s) (RPC_impl :466-490).  Rock and fluid tables (rock_list) are not
// read; permeability and porosity come from an ECLIPSE deck through
// init(deck, global_cell, ...) (RPC_impl :275-300, over mhs_deck_rock).
#ifndef OPM_PORSOL_AMD_RESERVOIRPROPERTYCOMMON_HPP
#define OPM_PORSOL_AMD_RESERVOIRPROPERTYCOMMON_HPP

#include <cassert>
#include <stdexcept>
#include <string>
#include <vector>

#include "EclipseDeck.hpp"
#include "Units.hpp"

namespace Opm {

/// Read-only view of a row-major dim x dim tensor (ImmutableCMatrix).
class PermTensorView {
public:
    PermTensorView(int n, const double *p) : n_(n), p_(p) {}
    double operator()(int i, int j) const { return p_[i * n_ + j]; }
    int numRows() const { return n_; }
    int numCols() const { return n_; }
    const double *data() const { return p_; }

private:
    int n_;
    const double *p_;
};

/// Writable view (SharedCMatrix).
class PermTensorRef {
public:
    PermTensorRef(int n, double *p) : n_(n), p_(p) {}
    double &operator()(int i, int j) { return p_[i * n_ + j]; }
    double operator()(int i, int j) const { return p_[i * n_ + j]; }
    double *data() { return p_; }

private:
    int n_;
    double *p_;
};

template <int dim>
class ReservoirPropertyCommon {
public:
    enum { NumberOfPhases = 2 };
    typedef PermTensorView PermTensor;
    typedef PermTensorRef MutablePermTensor;
    typedef PermTensorRef SharedPermTensor;

    ReservoirPropertyCommon()
        : density1_(1013.9 * unit::kilogram / unit::cubic(unit::meter)),
          density2_(834.7 * unit::kilogram / unit::cubic(unit::meter)),
          viscosity1_(1.0 * prefix::centi * unit::Poise),
          viscosity2_(3.0 * prefix::centi * unit::Poise) {}

    /// Uniform porosity and isotropic permeability (RPC_impl :318-334).
    void init(int num_cells, double uniform_poro = 0.2, double uniform_perm = 100.0 * unit::millidarcy) {
        porosity_.assign(num_cells, uniform_poro);
        permeability_.assign(dim * dim * (size_t)num_cells, 0.0);
        for (int i = 0; i < num_cells; ++i)
            for (int d = 0; d < dim; ++d) permeability_[dim * dim * (size_t)i + d * dim + d] = uniform_perm;
    }
    /// Per-cell full tensors, row-major (not in the reference; used by drivers
    /// that generate synthetic fields).
    void init(int num_cells, double uniform_poro, const std::vector<double> &perm_rowmajor) {
        assert(perm_rowmajor.size() == dim * dim * (size_t)num_cells);
        porosity_.assign(num_cells, uniform_poro);
        permeability_ = perm_rowmajor;
    }
    /// Permeability and porosity of the deck's cells global_cell (RPC_impl
    /// :275-300): the keyword rules of classifyPermeability / fillTensor /
    /// assignPermeability (:52-244, 682-736; values in mD, diagonal raised to
    /// perm_threshold) and assignPorosity (:592-611; 1 without PORO).  Rock
    /// tables are not supported: a rock_list_filename throws.  The J-function
    /// arguments only act on rock tables and are accepted for the signature.
    void init(const DeckConstPtr &deck, const std::vector<int> &global_cell, double perm_threshold = 0.0,
              const std::string *rock_list_filename = nullptr, bool use_jfunction_scaling = true,
              double sigma = 1.0, double theta = 0.0) {
        (void)use_jfunction_scaling; (void)sigma; (void)theta;
        static_assert(dim == 3, "ECLIPSE decks describe three-dimensional grids");
        if (!deck) throw std::invalid_argument("ReservoirPropertyCommon::init: null deck");
        if (rock_list_filename)
            throw std::runtime_error("ReservoirPropertyCommon::init: rock tables (rock_list) are not supported");
        const int nc = int(global_cell.size());
        porosity_.resize(nc);
        permeability_.resize(9 * (size_t)nc);
        int kind = 0;
        deck->check(mhs_deck_rock(deck->handle(), nc, global_cell.data(), perm_threshold, permeability_.data(),
                                  porosity_.data(), &kind),
                    "reading the rock properties");
        perm_kind_ = kind;
    }
    /// 0: no permeability keyword, 1: scalar, 2: diagonal, 3: full tensor (the
    /// classification of RPC_impl :52-104); -1 when not read from a deck.
    int permeabilityKind() const { return perm_kind_; }
    void setViscosities(double v1, double v2) { viscosity1_ = v1; viscosity2_ = v2; }
    void setDensities(double d1, double d2) { density1_ = d1; density2_ = d2; }

    int numCells() const { return int(porosity_.size()); }
    double porosity(int cell_index) const { return porosity_[cell_index]; }
    PermTensor permeability(int cell_index) const {
        return PermTensor(dim, &permeability_[dim * dim * (size_t)cell_index]);
    }
    SharedPermTensor permeabilityModifiable(int cell_index) {
        return SharedPermTensor(dim, &permeability_[dim * dim * (size_t)cell_index]);
    }
    double viscosityFirstPhase() const { return viscosity1_; }
    double viscositySecondPhase() const { return viscosity2_; }
    double densityFirstPhase() const { return density1_; }
    double densitySecondPhase() const { return density2_; }
    double densityDifference() const { return density1_ - density2_; }
    template <class Vector>
    void phaseDensities(int /*cell_index*/, Vector &density) const {
        density[0] = density1_;
        density[1] = density2_;
    }
    /// Without rock tables the reference falls back to 1e5 (1 - s) Pa and its
    /// derivative -1e5 (RPC_impl :466-490).
    double capillaryPressure(int /*cell_index*/, double saturation) const { return 1e5 * (1.0 - saturation); }
    double capillaryPressureDeriv(int /*cell_index*/, double /*saturation*/) const { return -1.0e5; }

protected:
    double density1_, density2_, viscosity1_, viscosity2_;
    int perm_kind_ = -1;
    std::vector<double> porosity_;
    std::vector<double> permeability_;
};

/// Two-phase fluid with quadratic relative permeabilities.
template <int dim>
class ReservoirPropertyCapillary : public ReservoirPropertyCommon<dim> {
    typedef ReservoirPropertyCommon<dim> Super;

public:
    double relPermFirstPhase(int, double s) const { return s * s; }
    double relPermSecondPhase(int, double s) const { return (1.0 - s) * (1.0 - s); }
    double mobilityFirstPhase(int c, double s) const { return relPermFirstPhase(c, s) / Super::viscosity1_; }
    double mobilitySecondPhase(int c, double s) const { return relPermSecondPhase(c, s) / Super::viscosity2_; }
    double totalMobility(int c, double s) const { return mobilityFirstPhase(c, s) + mobilitySecondPhase(c, s); }
    double fractionalFlow(int c, double s) const {
        const double l1 = mobilityFirstPhase(c, s), l2 = mobilitySecondPhase(c, s);
        return l1 / (l1 + l2);
    }
    template <class Vector>
    void phaseMobilities(int c, double s, Vector &mobility) const {
        mobility[0] = mobilityFirstPhase(c, s);
        mobility[1] = mobilitySecondPhase(c, s);
    }
};

/// Tracer fluid: lambda_0 = s, lambda_1 = 1 - s, unit densities
/// (ReservoirPropertyTracerFluid.hpp:44-65).
class ReservoirPropertyTracerFluid {
public:
    enum { NumberOfPhases = 2 };
    template <class Vector>
    void phaseMobilities(int, double s, Vector &mobility) const {
        mobility[0] = s;
        mobility[1] = 1.0 - s;
    }
    template <class Vector>
    void phaseDensities(int, Vector &density) const {
        density[0] = density[1] = 1.0;
    }
};

/// Fixed mobility (ReservoirPropertyFixedMobility.hpp:44-78): phase 0 carries
/// mobs[c].mob, phase 1 nothing; densities (1, 0).  Mobility is any type with
/// a scalar member `mob`.
template <class Mobility>
class ReservoirPropertyFixedMobility {
public:
    enum { NumberOfPhases = 2 };
    explicit ReservoirPropertyFixedMobility(const std::vector<Mobility> &mobs) : mobs_(mobs) {}
    template <class Vector>
    void phaseMobilities(int cell_index, double /*saturation*/, Vector &mobility) const {
        mobility[0] = mobs_[cell_index].mob;
        mobility[1] = 0.0;
    }
    template <class Vector>
    void phaseDensities(int /*cell_index*/, Vector &rho) const {
        rho[0] = 1.0;
        rho[1] = 0.0;
    }

private:
    const std::vector<Mobility> &mobs_;
};

}  // namespace Opm

#endif

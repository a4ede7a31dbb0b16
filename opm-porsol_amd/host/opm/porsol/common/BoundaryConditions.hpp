// Boundary-condition types with the member API of
// opm/porsol/common/BoundaryConditions.hpp (FlowBC :121-171,
// PeriodicConditionHandler :244-330, BasicBoundaryConditions :342-440).
// Dune- and Boost-free: the condition tables are plain std::vectors.
#ifndef OPM_PORSOL_AMD_BOUNDARYCONDITIONS_HPP
#define OPM_PORSOL_AMD_BOUNDARYCONDITIONS_HPP

#include <cassert>
#include <ostream>
#include <stdexcept>
#include <vector>

namespace Opm {

template <typename T>
class BCBase {
public:
    enum BCType { Dirichlet, Neumann, Periodic };

    template <typename charT, class traits>
    void write(std::basic_ostream<charT, traits> &os) const {
        os << "Type: " << type_ << "   Value: " << value_;
    }

protected:
    BCBase() : type_(Neumann), value_() {}
    BCBase(BCType type, T value) : type_(type), value_(value) {}
    bool isDirichlet() const { return type_ == Dirichlet; }
    bool isNeumann() const { return type_ == Neumann; }
    bool isPeriodic() const { return type_ == Periodic; }

    BCType type_;
    T value_;
};

template <typename charT, class traits, typename T>
std::basic_ostream<charT, traits> &operator<<(std::basic_ostream<charT, traits> &os,
                                             const BCBase<T> &bc) {
    bc.write(os);
    return os;
}

/// Flow boundary condition.  Default: no-flow (Neumann, 0).  The typed
/// accessors throw on a type mismatch in debug builds, as in the reference.
class FlowBC : public BCBase<double> {
public:
    FlowBC() : BCBase<double>(Neumann, 0.0) {}
    FlowBC(BCType type, double value) : BCBase<double>(type, value) {
        assert(isNeumann() || isDirichlet() || isPeriodic());
    }
    using BCBase<double>::isDirichlet;
    using BCBase<double>::isNeumann;
    using BCBase<double>::isPeriodic;

    double pressure() const {
#ifndef NDEBUG
        if (!isDirichlet())
            throw std::runtime_error("Pressure boundary conditions are only valid for Dirichlet boundaries");
#endif
        return value_;
    }
    double outflux() const {
#ifndef NDEBUG
        if (!isNeumann())
            throw std::runtime_error("Outflux boundary conditions are only valid for Neumann boundaries");
#endif
        return value_;
    }
    double pressureDifference() const {
#ifndef NDEBUG
        if (!isPeriodic())
            throw std::runtime_error("Pressure difference boundary conditions are only valid for periodic boundaries");
#endif
        return value_;
    }
};

/// Saturation boundary condition (kept for type compatibility of the BC tables;
/// the pressure solver does not read it).  Default: Dirichlet 1.0.
class SatBC : public BCBase<double> {
public:
    SatBC() : BCBase<double>(Dirichlet, 1.0) {}
    SatBC(BCType type, double value) : BCBase<double>(type, value) {
        assert(isDirichlet() || isPeriodic());
    }
    using BCBase<double>::isDirichlet;
    using BCBase<double>::isPeriodic;
    double saturation() const { return value_; }
    double saturationDifference() const { return value_; }
};

class PeriodicConditionHandler {
public:
    PeriodicConditionHandler() {}
    explicit PeriodicConditionHandler(int num_different_boundary_ids)
        : periodic_partner_bid_(num_different_boundary_ids, 0),
          canonical_bid_(num_different_boundary_ids, 0) {}

    void resize(int new_size) {
        periodic_partner_bid_.resize(new_size, 0);
        canonical_bid_.resize(new_size, 0);
    }
    bool empty() const { return periodic_partner_bid_.empty() && canonical_bid_.empty(); }
    void clear() {
        periodic_partner_bid_.clear();
        canonical_bid_.clear();
    }
    int size() const { return int(periodic_partner_bid_.size()); }

    void setPeriodicPartners(int boundary_id_1, int boundary_id_2) {
        assert(boundary_id_1 >= 0 && boundary_id_1 < int(periodic_partner_bid_.size()));
        assert(boundary_id_2 >= 0 && boundary_id_2 < int(periodic_partner_bid_.size()));
        assert(periodic_partner_bid_[boundary_id_1] == 0);
        assert(periodic_partner_bid_[boundary_id_2] == 0);
        periodic_partner_bid_[boundary_id_1] = boundary_id_2;
        periodic_partner_bid_[boundary_id_2] = boundary_id_1;
    }
    int getPeriodicPartner(int boundary_id) const {
        assert(boundary_id >= 0 && boundary_id < int(periodic_partner_bid_.size()));
        return periodic_partner_bid_[boundary_id];
    }
    void setCanonicalBoundaryId(int boundary_id, int canonical_bid) {
        assert(boundary_id >= 0 && boundary_id < int(canonical_bid_.size()));
        canonical_bid_[boundary_id] = canonical_bid;
    }
    int getCanonicalBoundaryId(int boundary_id) const {
        assert(boundary_id >= 0 && boundary_id < int(canonical_bid_.size()));
        return canonical_bid_[boundary_id];
    }

    template <typename charT, class traits>
    void write(std::basic_ostream<charT, traits> &os) const {
        for (int i = 0; i < int(periodic_partner_bid_.size()); ++i)
            os << "Partner of bid " << i << " is " << periodic_partner_bid_[i]
               << "  (canonical bid is " << canonical_bid_[i] << '\n';
        os << std::endl;
    }

private:
    std::vector<int> periodic_partner_bid_;
    std::vector<int> canonical_bid_;
};

template <typename charT, class traits>
std::basic_ostream<charT, traits> &operator<<(std::basic_ostream<charT, traits> &os,
                                             const PeriodicConditionHandler &pch) {
    pch.write(os);
    return os;
}

/// Per-boundary-id condition tables.  FC / SC select flow / saturation tables
/// (ZC, surface volumes, is accepted and ignored).
template <bool FC = false, bool SC = false, bool ZC = false, int numComponents = 3>
class BasicBoundaryConditions : public PeriodicConditionHandler {
public:
    static const bool HasFlowConds = FC;
    static const bool HasSatConds = SC;

    BasicBoundaryConditions() {}
    explicit BasicBoundaryConditions(int num_different_boundary_ids)
        : PeriodicConditionHandler(num_different_boundary_ids),
          flow_(FC ? num_different_boundary_ids : 0),
          sat_(SC ? num_different_boundary_ids : 0) {}

    void resize(int new_size) {
        PeriodicConditionHandler::resize(new_size);
        if (FC) flow_.resize(new_size);
        if (SC) sat_.resize(new_size);
    }
    bool empty() const { return PeriodicConditionHandler::empty(); }
    void clear() {
        PeriodicConditionHandler::clear();
        flow_.clear();
        sat_.clear();
    }
    int size() const { return PeriodicConditionHandler::size(); }

    FlowBC &flowCond(int i) { return flow_[i]; }
    const FlowBC &flowCond(int i) const { return flow_[i]; }
    template <class BoundaryFace>
    const FlowBC &flowCond(const BoundaryFace &bf) const {
        assert(bf.boundary());
        return flow_[bf.boundaryId()];
    }
    SatBC &satCond(int i) { return sat_[i]; }
    const SatBC &satCond(int i) const { return sat_[i]; }
    template <class BoundaryFace>
    const SatBC &satCond(const BoundaryFace &bf) const {
        assert(bf.boundary());
        return sat_[bf.boundaryId()];
    }

private:
    std::vector<FlowBC> flow_;
    std::vector<SatBC> sat_;
};

}  // namespace Opm

#endif

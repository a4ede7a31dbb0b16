// Drop-in for
//   template <class GridInterface, class ReservoirProperties, class BoundaryConditions>
//   class Opm::EulerUpstream               (opm/porsol/euler/EulerUpstream.hpp:56-150)
// over the C-ABI of include/euler_upstream.h: the explicit upstream saturation
// solve runs on the GPU that holds the pressure solution it is given
// (IncompFlowSolverHybrid::FlowSolution::deviceHandle()).
//
// Same members and parameters: init(param) (EulerUpstream_impl.hpp:95-108),
// init(param, g, r, b), initObj(g, r, b) (:120-128), display(),
// setCourantNumber(), transportSolve(saturation, time, gravity, pressure_sol,
// injection_rates) (:153-215).  Saturations and injection rates are indexed by
// Cell::index().  The fluid is the one without rock tables (quadratic relative
// permeabilities, p_c = 1e5 (1 - s)); viscosities and densities come from the
// reservoir properties.  Boundary conditions need saturation conditions
// (BC::HasSatConds): Dirichlet saturations on boundary faces, periodic faces
// as paired by the pressure solver.  A saturation out of range after every
// retry throws std::runtime_error as the reference rethrows (:198-205); the
// saturation is then left at its value before the call.
#ifndef MHS_DROPIN_EULERUPSTREAM_HPP
#define MHS_DROPIN_EULERUPSTREAM_HPP

#include <euler_upstream.h>

#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace Opm {

namespace euler_detail {
template <class R, class = void> struct has_element : std::false_type {};
template <class R>
struct has_element<R, std::void_t<decltype(std::declval<const R &>().element(0))>> : std::true_type {};
// Opm::SparseVector<double>::element(i), or a dense vector
template <class R>
double rate_at(const R &r, int i) {
    if constexpr (has_element<R>::value) return r.element(i);
    else return r[i];
}
}  // namespace euler_detail

template <class GI, class RP, class BC>
class EulerUpstream {
    static_assert(BC::HasSatConds, "EulerUpstream needs boundary conditions with saturation conditions");

public:
    EulerUpstream() { mhs_transport_default_options(&opt_); }
    EulerUpstream(const GI &g, const RP &r, const BC &b) : EulerUpstream() { initObj(g, r, b); }
    ~EulerUpstream() {
        if (t_) mhs_transport_destroy(t_);
    }
    EulerUpstream(const EulerUpstream &) = delete;
    EulerUpstream &operator=(const EulerUpstream &) = delete;

    /// Parameters by name with the reference's defaults (:95-108); Param has
    /// getDefault(name, default) as Opm::parameter::ParameterGroup.
    template <class Param>
    void init(const Param &param) {
        opt_.courant_number = param.getDefault("courant_number", opt_.courant_number);
        opt_.method_viscous = param.getDefault("method_viscous", bool(opt_.method_viscous));
        opt_.method_gravity = param.getDefault("method_gravity", bool(opt_.method_gravity));
        opt_.method_capillary = param.getDefault("method_capillary", bool(opt_.method_capillary));
        opt_.use_cfl_viscous = param.getDefault("use_cfl_viscous", bool(opt_.use_cfl_viscous));
        opt_.use_cfl_gravity = param.getDefault("use_cfl_gravity", bool(opt_.use_cfl_gravity));
        opt_.use_cfl_capillary = param.getDefault("use_cfl_capillary", bool(opt_.use_cfl_capillary));
        opt_.minimum_small_steps = param.getDefault("minimum_small_steps", opt_.minimum_small_steps);
        opt_.maximum_small_steps = param.getDefault("maximum_small_steps", opt_.maximum_small_steps);
        opt_.check_sat = param.getDefault("check_sat", bool(opt_.check_sat));
        opt_.clamp_sat = param.getDefault("clamp_sat", bool(opt_.clamp_sat));
    }
    template <class Param>
    void init(const Param &param, const GI &g, const RP &r, const BC &b) {
        init(param);
        initObj(g, r, b);
    }

    /// Grid, properties and boundary conditions (:120-128).  The pore
    /// volumes and boundary saturations reach the GPU at the first transport
    /// solve, when the pressure solution names the device that holds it.
    void initObj(const GI &g, const RP &r, const BC &b) {
        pgrid_ = &g;
        pres_ = &r;
        pbc_ = &b;
        bound_ = nullptr;
    }

    void display() {
        std::cout << "\nDisplaying some members of EulerUpstream\n\ncourant_number = " << opt_.courant_number
                  << '\n';
    }
    void setCourantNumber(double cn) { opt_.courant_number = cn; }

    /// Advances saturation (indexed by Cell::index()) over `time` seconds with
    /// the fluxes of pressure_sol (:153-215).
    template <class PressureSolution, class Rates>
    void transportSolve(std::vector<double> &saturation, const double time, const typename GI::Vector &gravity,
                        const PressureSolution &pressure_sol, const Rates &injection_rates) const {
        mhs_solver *h = pressure_sol.deviceHandle();
        if (!h) throw std::runtime_error("EulerUpstream::transportSolve: the pressure solution is not on a GPU");
        if (!pgrid_) throw std::runtime_error("EulerUpstream::transportSolve: call initObj first");
        if (h != bound_) bind(h);
        const int nc = pgrid_->numberOfCells();
        std::vector<double> s(nc), q(nc);
        int pos = 0;
        for (auto c = pgrid_->cellbegin(); c != pgrid_->cellend(); ++c, ++pos) {
            s[pos] = saturation[c->index()];
            q[pos] = euler_detail::rate_at(injection_rates, c->index());
        }
        const double g3[3] = {gravity[0], gravity[1], gravity[2]};
        int rc = mhs_transport_solve(t_, s.data(), time, g3, q.data(), &opt_, &report_);
        if (rc == MHS_ERR_STATE) {
            // the pressure solver rebuilt its structure since the last call
            bind(h);
            rc = mhs_transport_solve(t_, s.data(), time, g3, q.data(), &opt_, &report_);
        }
        if (rc != MHS_OK)
            throw std::runtime_error(std::string("EulerUpstream::transportSolve: ") + mhs_transport_last_error(t_));
        pos = 0;
        for (auto c = pgrid_->cellbegin(); c != pgrid_->cellend(); ++c, ++pos) saturation[c->index()] = s[pos];
    }

    /// No counterpart in the reference: step counts and CFL time of the last call.
    const mhs_transport_report &report() const { return report_; }

private:
    void bind(mhs_solver *h) const {
        if (t_) mhs_transport_destroy(t_);
        t_ = nullptr;
        if (mhs_transport_create(&t_, h) != MHS_OK)
            throw std::runtime_error("EulerUpstream: mhs_transport_create failed");
        std::vector<double> poro, satbc;
        for (auto c = pgrid_->cellbegin(); c != pgrid_->cellend(); ++c) {
            poro.push_back(pres_->porosity(c->index()));
            for (auto f = c->facebegin(); f != c->faceend(); ++f) {
                double v = 0.0;
                if (f->boundary()) {
                    const auto &cond = pbc_->satCond(*f);
                    if (cond.isDirichlet()) v = cond.saturation();
                }
                satbc.push_back(v);
            }
        }
        mhs_fluid fl;
        fl.viscosity[0] = pres_->viscosityFirstPhase();
        fl.viscosity[1] = pres_->viscositySecondPhase();
        fl.density[0] = pres_->densityFirstPhase();
        fl.density[1] = pres_->densitySecondPhase();
        if (mhs_transport_init(t_, poro.data(), satbc.data(), &fl) != MHS_OK)
            throw std::runtime_error(std::string("EulerUpstream::initObj: ") + mhs_transport_last_error(t_));
        bound_ = h;
    }

    const GI *pgrid_ = nullptr;
    const RP *pres_ = nullptr;
    const BC *pbc_ = nullptr;
    mhs_transport_options opt_;
    mutable mhs_transport *t_ = nullptr;
    mutable mhs_solver *bound_ = nullptr;
    mutable mhs_transport_report report_{};
};

}  // namespace Opm

#endif

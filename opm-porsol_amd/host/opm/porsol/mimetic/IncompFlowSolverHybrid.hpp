// Drop-in for opm/porsol/mimetic/IncompFlowSolverHybrid.hpp.
//
//   template <class GridInterface, class RockInterface, class BCInterface,
//             template <class GridIF, class RockIF> class InnerProduct>
//   class Opm::IncompFlowSolverHybrid            (IncompFlowSolverHybrid.hpp:362-366)
//
// Same public methods, argument meaning and error behaviour; the work runs on
// an MI355X through the C-ABI of include/mimetic_hybrid.h.  The grid, rock,
// boundary conditions and fluid are read through the reference accessors
// (CellIterator / FaceIterator, permeability(i), flowCond(face),
// getPeriodicPartner(bid), phaseMobilities / phaseDensities) and handed over
// as flat arrays.  Extensions: setPrecision() (FP64 or FP32 PCG with FP64
// residual replacement), facePressures(), lastReport().
#ifndef OPM_PORSOL_AMD_INCOMPFLOWSOLVERHYBRID_HPP
#define OPM_PORSOL_AMD_INCOMPFLOWSOLVERHYBRID_HPP

#include <algorithm>
#include <array>
#include <cassert>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <iterator>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>

#include "mimetic_hybrid.h"
#include "MimeticIPAnisoRelpermEvaluator.hpp"
#include "MimeticIPEvaluator.hpp"

namespace Opm {

namespace hybrid_detail {
// Logical (i,j,k) of every cell, when the grid type can say: a box grid with
// dims() (cells numbered x fastest) or a corner-point grid with the CpGrid
// accessors logicalCartesianSize() and globalCell().  The multigrid
// preconditioner builds its aggregates from these; without them the solve
// uses the Jacobi preconditioner.
template <class G, class = void> struct has_dims : std::false_type {};
template <class G>
struct has_dims<G, std::void_t<decltype(std::declval<const G &>().dims())>> : std::true_type {};
template <class G, class = void> struct has_logical : std::false_type {};
template <class G>
struct has_logical<G, std::void_t<decltype(std::declval<const G &>().logicalCartesianSize()),
                                  decltype(std::declval<const G &>().globalCell())>> : std::true_type {};

template <class G>
bool logical_index(const G &grid, int cell, int ijk[3]) {
    long long g = cell;
    int n[3];
    if constexpr (has_dims<G>::value) {
        const int *d = grid.dims();
        n[0] = d[0]; n[1] = d[1]; n[2] = d[2];
    } else if constexpr (has_logical<G>::value) {
        const auto &sz = grid.logicalCartesianSize();
        n[0] = sz[0]; n[1] = sz[1]; n[2] = sz[2];
        g = grid.globalCell()[cell];
    } else {
        (void)grid; (void)cell; (void)ijk; (void)g; (void)n;
        return false;
    }
    if (n[0] <= 0 || n[1] <= 0 || n[2] <= 0) return false;
    ijk[0] = int(g % n[0]);
    ijk[1] = int((g / n[0]) % n[1]);
    ijk[2] = int(g / ((long long)n[0] * n[1]));
    return true;
}
}  // namespace hybrid_detail

/// topologyIsSane (IncompFlowSolverHybrid.hpp:108-130): true when every pair
/// of connected cells shares a single face.  The neighbour lists go to the
/// C-ABI's host check mhs_topology_is_sane.
template <class GI>
bool topologyIsSane(const GI &g) {
    typedef typename GI::CellIterator CI;
    typedef typename CI::FaceIterator FI;
    std::vector<int> facepos(1, 0), nb;
    for (CI c = g.cellbegin(); c != g.cellend(); ++c) {
        for (FI f = c->facebegin(); f != c->faceend(); ++f)
            nb.push_back(f->boundary() ? -1 : f->neighbourCellIndex());
        facepos.push_back((int)nb.size());
    }
    int sane = 0;
    if (mhs_topology_is_sane((int)facepos.size() - 1, facepos.data(), nb.data(), &sane) != MHS_OK)
        return false;
    return sane != 0;
}

template <class GridInterface, class RockInterface, class BCInterface,
          template <class GridIF, class RockIF> class InnerProduct>
class IncompFlowSolverHybrid {
    static constexpr bool kTensorMobility =
        std::is_same<InnerProduct<GridInterface, RockInterface>,
                     MimeticIPAnisoRelpermEvaluator<GridInterface, RockInterface>>::value;
    static_assert(std::is_same<InnerProduct<GridInterface, RockInterface>,
                               MimeticIPEvaluator<GridInterface, RockInterface>>::value ||
                      kTensorMobility,
                  "the MI355X path implements the MimeticIPEvaluator and MimeticIPAnisoRelpermEvaluator "
                  "inner products");
    typedef typename GridInterface::Scalar Scalar;
    enum FaceType { Internal, Dirichlet, Neumann, Periodic };

public:
    /// Solution object (IncompFlowSolverHybrid.hpp:382-447).
    class FlowSolution {
    public:
        typedef typename GridInterface::Scalar Scalar;
        typedef typename GridInterface::CellIterator CI;
        typedef typename CI::FaceIterator FI;
        friend class IncompFlowSolverHybrid;

        Scalar pressure(const CI &c) const { return pressure_[cellno_[c->index()]]; }
        Scalar outflux(const FI &f) const {
            return outflux_[facepos_[cellno_[f->cellIndex()]] + f->localIndex()];
        }
        Scalar outflux(int hf) const { return outflux_[hf]; }
        /// No counterpart in the reference: the solver handle that holds this
        /// solution on the GPU, for consumers that read the fluxes in device
        /// memory (the GPU EulerUpstream, estimateCellVelocity); null before
        /// the first solve, and for a copy of the solution once the solver
        /// has solved, post-processed or been cleared since the copy was made.
        mhs_solver *deviceHandle() const { return (handle_ && live_ && *live_ == gen_) ? handle_ : nullptr; }

    private:
        mhs_solver *handle_ = nullptr;
        std::shared_ptr<const long long> live_;   // the solver's solve counter
        long long gen_ = -1;                       // its value for this solution
        std::vector<int> cellno_;
        std::vector<int> facepos_;
        std::vector<Scalar> pressure_;
        std::vector<Scalar> outflux_;
        void clear() {
            handle_ = nullptr;
            live_.reset();
            gen_ = -1;
            std::vector<int>().swap(cellno_);
            std::vector<int>().swap(facepos_);
            std::vector<Scalar>().swap(pressure_);
            std::vector<Scalar>().swap(outflux_);
        }
    };

    explicit IncompFlowSolverHybrid(int device = 0) : device_(device) { clear(); }
    ~IncompFlowSolverHybrid() {
        if (h_) mhs_destroy(h_);
    }
    IncompFlowSolverHybrid(const IncompFlowSolverHybrid &) = delete;
    IncompFlowSolverHybrid &operator=(const IncompFlowSolverHybrid &) = delete;

    /// All-in-one initialisation (:476-488).
    template <class Point>
    void init(const GridInterface &g, const RockInterface &r, const Point &grav,
              const BCInterface &bc) {
        clear();
        if (g.numberOfCells() > 0) {
            initSystemStructure(g, bc);
            computeInnerProducts(r, grav);
        }
    }

    /// Release topology and inner products (:498-516).
    void clear() {
        pgrid_ = nullptr;
        max_ncf_ = -1;
        num_internal_faces_ = 0;
        total_num_faces_ = 0;
        matrix_structure_valid_ = false;
        do_regularization_ = true;
        flowSolution_.clear();
        ++*solve_count_;
        periodic_hf_.clear();
        face_pressure_.clear();
        if (h_) check(mhs_clear(h_), "clear");
        cleared_state_ = true;
    }

    /// Dof enumeration and face-system structure (:534-545).  Only periodic
    /// conditions are inspected, as in the reference.
    void initSystemStructure(const GridInterface &g, const BCInterface &bc) {
        assert(cleared_state_);
        typedef typename GridInterface::CellIterator CI;
        typedef typename CI::FaceIterator FI;
        ensure_handle();
        const int nc = g.numberOfCells();
        std::vector<int> &cellno = flowSolution_.cellno_;
        cellno.assign(nc, -1);
        int pos = 0;
        for (CI c = g.cellbegin(); c != g.cellend(); ++c) cellno[c->index()] = pos++;
        std::vector<int> facepos(1, 0), nb;
        std::vector<double> cvol, ccen, fa, fc, fn;
        std::vector<int> bid;
        for (CI c = g.cellbegin(); c != g.cellend(); ++c) {
            cvol.push_back(c->volume());
            const auto cc = c->centroid();
            for (int d = 0; d < 3; ++d) ccen.push_back(cc[d]);
            int n = 0;
            for (FI f = c->facebegin(); f != c->faceend(); ++f, ++n) {
                if (f->localIndex() != n) throw std::runtime_error("face local indices must be 0..n-1 in order");
                nb.push_back(f->boundary() ? -1 : cellno[f->neighbourCellIndex()]);
                bid.push_back(f->boundary() ? f->boundaryId() : 0);
                fa.push_back(f->area());
                const auto x = f->centroid();
                const auto nn = f->normal();
                for (int d = 0; d < 3; ++d) { fc.push_back(x[d]); fn.push_back(nn[d]); }
            }
            facepos.push_back(facepos.back() + n);
        }
        {   // as the reference asserts (:540)
            int sane = 0;
            (void)mhs_topology_is_sane(nc, facepos.data(), nb.data(), &sane);
            assert(sane);
            (void)sane;
        }
        check(mhs_set_grid(h_, 1, nullptr, nc, facepos.data(), nb.data(), cvol.data(), ccen.data(),
                           fa.data(), fc.data(), fn.data()),
              "set_grid");
        flowSolution_.facepos_ = facepos;
        {
            std::vector<int> ijk(3 * (size_t)nc);
            bool have = true;
            for (CI c = g.cellbegin(); c != g.cellend() && have; ++c)
                have = hybrid_detail::logical_index(g.grid(), c->index(), &ijk[3 * (size_t)cellno[c->index()]]);
            if (have) check(mhs_set_cell_ijk(h_, ijk.data()), "set_cell_ijk");
        }
        // Periodic pairing (enumerateBCDof :1071-1110): bid -> first half-face
        // carrying it, partner through getPeriodicPartner(bid).
        const int nhf = facepos.back();
        std::vector<int> bctype(nhf, MHS_BC_NEUMANN), partner(nhf, -1);
        std::unordered_map<int, int> bid_hf;
        periodic_hf_.clear();
        {
            int h = 0;
            for (CI c = g.cellbegin(); c != g.cellend(); ++c)
                for (FI f = c->facebegin(); f != c->faceend(); ++f, ++h)
                    if (f->boundary() && bc.flowCond(*f).isPeriodic()) {
                        bid_hf.insert(std::make_pair(f->boundaryId(), h));
                        periodic_hf_.push_back(h);
                    }
            h = 0;
            for (CI c = g.cellbegin(); c != g.cellend(); ++c)
                for (FI f = c->facebegin(); f != c->faceend(); ++f, ++h)
                    if (f->boundary() && bc.flowCond(*f).isPeriodic()) {
                        auto j = bid_hf.find(bc.getPeriodicPartner(f->boundaryId()));
                        if (j == bid_hf.end()) throw std::runtime_error("periodic face without a partner");
                        bctype[h] = MHS_BC_PERIODIC;
                        partner[h] = j->second;
                    }
        }
        check(mhs_init_system_structure(h_, bctype.data(), partner.data()), "initSystemStructure");
        mhs_dims d;
        check(mhs_get_dims(h_, &d), "get_dims");
        max_ncf_ = d.max_ncf;
        num_internal_faces_ = d.num_internal_faces;
        total_num_faces_ = d.num_faces;
        nhf_ = nhf;
        flowSolution_.pressure_.assign(nc, 0.0);
        flowSolution_.outflux_.assign(nhf, 0.0);
        pgrid_ = &g;
        cleared_state_ = false;
        matrix_structure_valid_ = true;
    }

    /// Static inverse inner products (:564-578, MimeticIPEvaluator.hpp:201-270).
    template <class Point>
    void computeInnerProducts(const RockInterface &r, const Point &grav) {
        assert(matrix_structure_valid_);
        typedef typename GridInterface::CellIterator CI;
        std::vector<double> perm;
        perm.reserve(9 * (size_t)pgrid_->numberOfCells());
        for (CI c = pgrid_->cellbegin(); c != pgrid_->cellend(); ++c) {
            const auto K = r.permeability(c->index());
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) perm.push_back(K(i, j));
        }
        const double g3[3] = {double(grav[0]), double(grav[1]), double(grav[2])};
        check(mhs_compute_inner_products(h_, perm.data(), g3), "computeInnerProducts");
    }
    /// "assembleStatic" (the name used by commented-out reference callers).
    template <class Point>
    void assembleStatic(const RockInterface &r, const Point &grav) { computeInnerProducts(r, grav); }

    /// Assemble, solve and recover (:657-703).  linsolver_type 0 selects
    /// Jacobi-PCG (the GPU stand-in for ILU0-CG); 1..3 the AMG family, served
    /// by the same PCG with the reference's Dirichlet initial guess.
    template <class FluidInterface>
    void solve(const FluidInterface &r, const std::vector<double> &sat, const BCInterface &bc,
               const std::vector<double> &src, double residual_tolerance = 1e-8,
               int linsolver_verbosity = 1, int linsolver_type = 1, bool same_matrix = false,
               int linsolver_maxit = 0, double prolongate_factor = 1.6, int smooth_steps = 1) {
        typedef typename GridInterface::CellIterator CI;
        typedef typename CI::FaceIterator FI;
        assert(matrix_structure_valid_);
        const int nc = pgrid_->numberOfCells();
        std::vector<double> totmob(nc), rhomob(nc), s(nc);
        std::vector<double> lam_t(kTensorMobility ? 9 * (size_t)nc : 0), rho_lam(lam_t.size());
        std::vector<int> bctype(nhf_, MHS_BC_NEUMANN);
        std::vector<double> bcval(nhf_, 0.0);
        int h = 0, pos = 0;
        do_regularization_ = true;
        for (CI c = pgrid_->cellbegin(); c != pgrid_->cellend(); ++c, ++pos) {
            const int ci = c->index();
            std::array<double, FluidInterface::NumberOfPhases> rho;
            r.phaseDensities(ci, rho);
            if constexpr (kTensorMobility) {
                // MimeticIPAnisoRelpermEvaluator::computeDynamicParams (:299-330):
                // lambda_t = sum of the phase mobility tensors, and sum rho_i lambda_i
                for (int ph = 0; ph < FluidInterface::NumberOfPhases; ++ph) {
                    Mat3 m;
                    r.phaseMobility(ph, ci, sat[ci], m);
                    for (int l = 0; l < 3; ++l)
                        for (int k = 0; k < 3; ++k) {
                            lam_t[9 * (size_t)pos + 3 * l + k] += m(l, k);
                            rho_lam[9 * (size_t)pos + 3 * l + k] += rho[ph] * m(l, k);
                        }
                }
            } else {
                std::array<double, FluidInterface::NumberOfPhases> mob;
                r.phaseMobilities(ci, sat[ci], mob);
                totmob[pos] = std::accumulate(mob.begin(), mob.end(), 0.0);
                rhomob[pos] = std::inner_product(rho.begin(), rho.end(), mob.begin(), 0.0);
            }
            s[pos] = src[ci];
            for (FI f = c->facebegin(); f != c->faceend(); ++f, ++h) {
                if (!f->boundary()) continue;
                const auto &cond = bc.flowCond(*f);
                if (cond.isDirichlet()) {
                    bctype[h] = MHS_BC_DIRICHLET;
                    bcval[h] = cond.pressure();
                    do_regularization_ = false;
                } else if (cond.isPeriodic()) {
                    bctype[h] = MHS_BC_PERIODIC;
                    bcval[h] = cond.pressureDifference();
                } else {
                    bctype[h] = MHS_BC_NEUMANN;
                    bcval[h] = cond.outflux();
                }
            }
        }
        mhs_solve_options opt;
        mhs_default_options(&opt);
        opt.tol = residual_tolerance;
        opt.maxit = linsolver_maxit;
        opt.linsolver_type = linsolver_type;
        opt.verbosity = linsolver_verbosity;
        opt.same_matrix = same_matrix ? 1 : 0;
        opt.prolongate_factor = prolongate_factor;
        opt.smooth_steps = smooth_steps;
        opt.precision = precision_;
        const int rc = kTensorMobility
                           ? mhs_solve_aniso(h_, lam_t.data(), rho_lam.data(), s.data(), bctype.data(),
                                             bcval.data(), &opt, &report_)
                           : mhs_solve(h_, totmob.data(), rhomob.data(), s.data(), bctype.data(),
                                       bcval.data(), &opt, &report_);
        if (rc == MHS_ERR_INVALID && (linsolver_type < 0 || linsolver_type > 3)) {
            std::cerr << "Unknown linsolver_type: " << linsolver_type << '\n';
            throw std::runtime_error("Unknown linsolver_type");
        }
        check(rc, "solve");
        if (linsolver_verbosity > 0)
            std::cout << "=== IT=" << report_.iterations << ", reduction=" << report_.reduction
                      << ", assemble " << report_.assemble_ms << " ms, PCG " << report_.solve_ms
                      << " ms, recover " << report_.recover_ms << " ms\n";
        face_pressure_.resize(total_num_faces_);
        check(mhs_get_solution(h_, flowSolution_.pressure_.data(), face_pressure_.data(),
                               flowSolution_.outflux_.data()),
              "getSolution");
        mark_solution_live();
    }

    /// Antisymmetric twin fluxes (:754-807); returns the largest change.
    double postProcessFluxes() {
        double mm = 0.0;
        check(mhs_post_process_fluxes(h_, &mm), "postProcessFluxes");
        check(mhs_get_solution(h_, nullptr, nullptr, flowSolution_.outflux_.data()), "getSolution");
        mark_solution_live();
        return mm;
    }

    typedef const FlowSolution &SolutionType;
    SolutionType getSolution() { return flowSolution_; }

    template <typename charT, class traits>
    void printStats(std::basic_ostream<charT, traits> &os) {
        os << "IncompFlowSolverHybrid<>:\n"
           << "\tMaximum number of cell faces = " << max_ncf_ << '\n'
           << "\tNumber of internal faces     = " << num_internal_faces_ << '\n'
           << "\tTotal number of faces        = " << total_num_faces_ << '\n';
        const std::vector<int> &cell = flowSolution_.cellno_;
        os << "cell index map = [";
        std::copy(cell.begin(), cell.end(), std::ostream_iterator<int>(os, " "));
        os << "\b]\n";
        std::vector<int> hf(nhf_);
        check(mhs_get_cell_faces(h_, hf.data(), nullptr), "get_cell_faces");
        const std::vector<int> &fp = flowSolution_.facepos_;
        os << "cell faces     =\n";
        for (size_t i = 0; i + 1 < fp.size(); ++i) {
            os << "\t[" << i << "] -> [";
            std::copy(hf.begin() + fp[i], hf.begin() + fp[i + 1], std::ostream_iterator<int>(os, ","));
            os << "\b]\n";
        }
    }

    /// MATLAB/spconvert dump of the last assembled system (:895-905).
    void printSystem(const std::string &prefix) {
        long long nnz = 0;
        check(mhs_get_system(h_, 0, nullptr, nullptr, nullptr, nullptr, &nnz), "get_system");
        std::vector<int> rp(total_num_faces_ + 1), col(nnz);
        std::vector<double> val(nnz), rhs(total_num_faces_);
        check(mhs_get_system(h_, 0, rp.data(), col.data(), val.data(), rhs.data(), &nnz), "get_system");
        std::ofstream mat((prefix + "-mat.dat").c_str());
        mat.precision(15);
        mat.setf(std::ios::scientific | std::ios::showpos);
        for (int i = 0; i < total_num_faces_; ++i)
            for (int k = rp[i]; k < rp[i + 1]; ++k)
                mat << i + 1 << ' ' << col[k] + 1 << ' ' << val[k] << '\n';
        std::ofstream rf((prefix + "-rhs.dat").c_str());
        rf.precision(15);
        rf.setf(std::ios::scientific | std::ios::showpos);
        for (double v : rhs) rf << v << '\n';
    }

    // --- extensions ----------------------------------------------------------
    void setPrecision(int precision) { precision_ = precision; }
    const std::vector<double> &facePressures() const { return face_pressure_; }
    const mhs_solve_report &lastReport() const { return report_; }

private:
    int device_ = 0;
    mhs_solver *h_ = nullptr;
    const GridInterface *pgrid_ = nullptr;
    bool cleared_state_ = true;
    int max_ncf_ = -1, num_internal_faces_ = 0, total_num_faces_ = 0, nhf_ = 0;
    bool matrix_structure_valid_ = false, do_regularization_ = true;
    int precision_ = MHS_FP64;
    std::vector<int> periodic_hf_;
    std::vector<double> face_pressure_;
    mhs_solve_report report_{};
    std::shared_ptr<long long> solve_count_ = std::make_shared<long long>(0);
    FlowSolution flowSolution_;

    // the device data now hold flowSolution_; earlier copies lose their handle
    void mark_solution_live() {
        flowSolution_.handle_ = h_;
        flowSolution_.live_ = solve_count_;
        flowSolution_.gen_ = ++*solve_count_;
    }
    void ensure_handle() {
        if (!h_ && mhs_create(&h_, device_, nullptr) != MHS_OK)
            throw std::runtime_error("IncompFlowSolverHybrid: cannot create the GPU solver");
    }
    void check(int rc, const char *what) const {
        if (rc == MHS_OK) return;
        const std::string msg = h_ ? mhs_last_error(h_) : "no solver handle";
        if (rc == MHS_ERR_NOT_CONVERGED) throw std::runtime_error(msg);
        throw std::runtime_error(std::string(what) + ": " + msg);
    }
};

}  // namespace Opm

#endif

// Drop-in for opm/porsol/mimetic/MimeticIPAnisoRelpermEvaluator.hpp.
//
// In the reference this class is the InnerProduct plug-in for tensor
// mobilities: per solve, B^-1 = (N (lambda_t K) N' + t A (I - QQ') A) / |c|
// with t = 6 tr(lambda_t K) / d (getInverseMatrix, :363-400) and the
// gravity flux N (sum_i rho_i lambda_i) K g (gravityFlux, :406-416), the
// phase mobility tensors coming from FluidInterface::phaseMobility(phase,
// cell, s, matrix) (computeDynamicParams, :299-330).  Here that computation
// is mhs_solve_aniso of libmhs.so; this class is the selector that keeps the
// reference template signature
// IncompFlowSolverHybrid<GI, RI, BCI, MimeticIPAnisoRelpermEvaluator>.
#ifndef OPM_PORSOL_AMD_MIMETICIPANISORELPERMEVALUATOR_HPP
#define OPM_PORSOL_AMD_MIMETICIPANISORELPERMEVALUATOR_HPP

#include <array>

namespace Opm {

/// 3x3 matrix given to FluidInterface::phaseMobility (the reference passes a
/// SharedFortranMatrix); element (i, j) by m(i, j).
struct Mat3 {
    std::array<double, 9> a{};
    double &operator()(int i, int j) { return a[3 * i + j]; }
    double operator()(int i, int j) const { return a[3 * i + j]; }
    double *data() { return a.data(); }
    const double *data() const { return a.data(); }
};
inline void zero(Mat3 &m) { m.a.fill(0.0); }

template <class GridInterface, class RockInterface>
class MimeticIPAnisoRelpermEvaluator {
public:
    MimeticIPAnisoRelpermEvaluator() {}
    explicit MimeticIPAnisoRelpermEvaluator(int /*max_nf*/) {}
    void init(int /*max_nf*/) {}
};

}  // namespace Opm

#endif

// Drop-in for opm/porsol/common/ReservoirPropertyCapillaryAnisotropicRelperm.hpp:
// two-phase properties whose phase mobilities are tensors.  Without rock
// tables the reference uses quadratic relative permeabilities on the
// diagonal (ReservoirPropertyCapillaryAnisotropicRelperm_impl.hpp:77-100):
// lambda_0 = s^2 / mu_0 I, lambda_1 = (1-s)^2 / mu_1 I; rock tables are not
// read here.
#ifndef OPM_PORSOL_AMD_RESERVOIRPROPERTYCAPILLARYANISOTROPICRELPERM_HPP
#define OPM_PORSOL_AMD_RESERVOIRPROPERTYCAPILLARYANISOTROPICRELPERM_HPP

#include <cassert>

#include "ReservoirPropertyCommon.hpp"

namespace Opm {

template <int dim>
class ReservoirPropertyCapillaryAnisotropicRelperm : public ReservoirPropertyCommon<dim> {
    typedef ReservoirPropertyCommon<dim> Super;

public:
    /// phaseMobility (:44-53): the phase's mobility tensor into phase_mob.
    template <class MatrixType>
    void phaseMobility(int phase_index, int /*cell_index*/, double saturation, MatrixType &phase_mob) const {
        assert(phase_index == 0 || phase_index == 1);
        const double visc = phase_index == 0 ? Super::viscosity1_ : Super::viscosity2_;
        const double kr = phase_index == 0 ? saturation * saturation : (1.0 - saturation) * (1.0 - saturation);
        for (int i = 0; i < dim; ++i)
            for (int j = 0; j < dim; ++j) phase_mob(i, j) = i == j ? kr / visc : 0.0;
    }
    /// fractionalFlow (:56-71): the mean over directions of the diagonal ratios.
    double fractionalFlow(int cell_index, double saturation) const {
        double ff = 0.0;
        for (int d = 0; d < dim; ++d) {
            double m0[dim][dim], m1[dim][dim];
            Wrap w0{&m0[0][0]}, w1{&m1[0][0]};
            phaseMobility(0, cell_index, saturation, w0);
            phaseMobility(1, cell_index, saturation, w1);
            ff += m0[d][d] / (m0[d][d] + m1[d][d]);
        }
        return ff / double(dim);
    }

private:
    struct Wrap {
        double *p;
        double &operator()(int i, int j) { return p[dim * i + j]; }
    };
};

}  // namespace Opm

#endif

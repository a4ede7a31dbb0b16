// Drop-in for opm/porsol/mimetic/MimeticIPEvaluator.hpp.
//
// In the reference this class computes, per cell, the inverse mimetic inner
// product B^-1 = (N K N' + t A (I - Q Q') A) / |c| with t = 6 tr(K) / d
// (MimeticIPEvaluator.hpp:201-270) and is the InnerProduct template argument
// of IncompFlowSolverHybrid.  Here that computation is the k_static_ip kernel
// of libmhs.so; this class is the selector that keeps the reference template
// signature IncompFlowSolverHybrid<GI, RI, BCI, MimeticIPEvaluator>.
#ifndef OPM_PORSOL_AMD_MIMETICIPEVALUATOR_HPP
#define OPM_PORSOL_AMD_MIMETICIPEVALUATOR_HPP

namespace Opm {

template <class GridInterface, class RockInterface>
class MimeticIPEvaluator {
public:
    MimeticIPEvaluator() {}
    explicit MimeticIPEvaluator(int /*max_nf*/) {}
    void init(int /*max_nf*/) {}
};

}  // namespace Opm

#endif

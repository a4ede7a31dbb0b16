// Drop-in for opm/porsol/common/ReservoirPropertyTracerFluid.hpp: the type lives in
// ReservoirPropertyCommon.hpp together with the other rock/fluid types.
#include "ReservoirPropertyCommon.hpp"

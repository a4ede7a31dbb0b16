// Drop-in for opm/porsol/common/ReservoirPropertyFixedMobility.hpp: the type lives in
// ReservoirPropertyCommon.hpp together with the other rock/fluid types.
#include "ReservoirPropertyCommon.hpp"

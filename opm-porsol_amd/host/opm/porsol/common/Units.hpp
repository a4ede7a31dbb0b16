// SI unit constants used by the drop-in (the subset of opm-core's
// opm/core/utility/Units.hpp that the pressure-solver drivers use).
#ifndef OPM_PORSOL_AMD_UNITS_HPP
#define OPM_PORSOL_AMD_UNITS_HPP

namespace Opm {
namespace prefix {
constexpr double centi = 1.0e-2;
constexpr double milli = 1.0e-3;
}  // namespace prefix
namespace unit {
constexpr double meter = 1.0;
constexpr double kilogram = 1.0;
constexpr double second = 1.0;
constexpr double Pascal = 1.0;
constexpr double barsa = 1.0e5;
constexpr double Poise = 0.1;
constexpr double darcy = 9.869233e-13;
constexpr double millidarcy = prefix::milli * darcy;
constexpr double gravity = 9.80665;
inline constexpr double cubic(double v) { return v * v * v; }
inline constexpr double square(double v) { return v * v; }
}  // namespace unit
}  // namespace Opm

#endif

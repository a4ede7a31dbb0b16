// Device scratch shared by the entry points that allocate nothing.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>

namespace mhs {

// Size of the per-GPU scratch array (held by every GPU the library runs on
// until the process ends).
constexpr size_t kScratchBytes = size_t(16) << 20;

// Device address of this GPU's scratch array.  The first call on a device
// (with that device current) looks the address up; later calls return the
// kept value.  The array is shared by every call on the GPU: calls that
// overlap on one GPU race.
hipError_t scratch_address(int device, void **ptr);

}  // namespace mhs

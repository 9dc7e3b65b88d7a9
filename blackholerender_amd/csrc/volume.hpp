// volume.hpp — launch helper of the volumetric disk render (volume.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"
#include "shade.hpp"
#include "volume_device.hpp"

namespace bhs {

// Volumetric render of the work items of ncam cameras (device table with ray_offset set, item
// prefix sums item_base): FP32 march of every non-band ray, then FP64 march of the band list
// (all_f64: every ray in FP64). sc.stats and sc.band_count must be zero on entry.
hipError_t launch_volumetric(const MarchConst* table, const uint32_t* item_base, uint32_t ncam,
                             uint32_t n_items, bool all_f64, const shade::ShadeConst& sh,
                             const SoA& out, const ShadeOut& so, const uint16_t* vol,
                             const float* lut, const Scratch& sc, int num_cus, hipStream_t s,
                             hipEvent_t ev_begin, hipEvent_t ev_end);

}  // namespace bhs

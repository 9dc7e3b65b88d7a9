// volume.hpp — launch helper of the volumetric disk render (volume.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"
#include "shade.hpp"
#include "volume_device.hpp"

namespace bhs {

// Volumetric render of the work items of ncam cameras (device table with ray_offset set, item
// prefix sums item_base): with any_band, a scan lists the FP64 band rays and their FP64 march runs
// on `aux` (forked at ev_fork, joined by ev_join) beside the FP32 march of the others on `s`;
// all_f64: every ray in FP64 on s. ev_begin/ev_end (may be null) bracket the main march on s.
// sc.stats and sc.band_count must be zero on entry.
hipError_t launch_volumetric(const MarchConst* table, const uint32_t* item_base, uint32_t ncam,
                             uint32_t n_items, bool all_f64, bool any_band,
                             const shade::ShadeConst& sh, const SoA& out, const ShadeOut& so,
                             const uint16_t* vol, const float* lut, const Scratch& sc,
                             int num_cus, hipStream_t s, hipStream_t aux, hipEvent_t ev_fork,
                             hipEvent_t ev_join, hipEvent_t ev_begin, hipEvent_t ev_end);

}  // namespace bhs

// footprint.hpp — the sky footprint post-pass: screen-space derivatives of the escape direction
// over 2x2 pixel quads and the footprint of the screen->sky Jacobian, from a render's ray planes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lawS_device.hpp"

namespace bhs {

// Output planes (device pointers, any may be null) — bhs_footprint_soa of include/bhs_trace.h.
// Three-component fields are 3 planes of n floats (x, y, z or r11, r12, r22).
struct FootprintPlanes {
    float* ddx;
    float* ddy;
    float* e1;
    float* e2;
    float* rinv;
    float* area;
    float* mu;
    uint8_t* fallback;
    int64_t n;   // plane stride: rays of the call
};

// Ray planes the pass reads: the escape flag (bit 2 of flags) and the escape direction.
struct FootprintIn {
    const uint8_t* flags;
    const float* esc;
};

// One camera (c.ray_offset = first ray of the camera in the planes).
hipError_t launch_footprint(const MarchConst& c, const FootprintIn& in, const FootprintPlanes& o,
                            bool coarse, hipStream_t s);
// Cameras of a device table (ray_offset per camera); max_quads: the largest camera's 2x2 quads;
// widths_even: every camera's width is even.
hipError_t launch_footprint_batch(const MarchConst* table, uint32_t ncam, int64_t max_quads,
                                  bool widths_even, const FootprintIn& in,
                                  const FootprintPlanes& o, bool coarse, hipStream_t s);

}  // namespace bhs

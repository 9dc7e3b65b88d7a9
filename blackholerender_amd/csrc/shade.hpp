// shade.hpp — launch helpers of the thin-disk shading pass, the noise/blackbody tables it reads and
// the frame accumulation (shade.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"
#include "shade_device.hpp"

namespace bhs {

// Shading output planes (device pointers): rgb 3 planes of n, sky weight 1 plane of n.
struct ShadeOut {
    float* rgb;
    float* sky;
    int64_t n;
};

// Rays of a shading pass whose recorded crossings do not settle their colour (more than two
// in-annulus crossings and still transparent after two): the completion march's list.
struct ShadeList {
    uint32_t* count;
    uint32_t* ids;
    uint32_t cap;
};

// Bakes the 128^3 R16F fbm noise volume (noise_bake.comp) into vol (2^21 uint16 texels).
hipError_t launch_noise_bake(uint16_t* vol, uint32_t seed, hipStream_t s);

// Host bake of the 256-entry RGBA blackbody table (src/DiskTextures.cpp:61-93) in double.
void bake_blackbody_lut(float* lut_host);

// Shading pass over the ray planes of `ncam` cameras (table: device MarchConst array with
// ray_offset set; n_rays their total), then the completion march of the listed rays on the same
// stream. list.count must be zero on entry (cleared by the caller's zero kernel).
hipError_t launch_shade(const MarchConst* table, uint32_t ncam, int64_t n_rays, const SoA& rays,
                        const ShadeOut& out, const shade::ShadeConst& sc, const uint16_t* vol,
                        const float* lut, const ShadeList& list, int num_cus, hipStream_t s,
                        hipEvent_t ev_begin, hipEvent_t ev_end);

// acc = sample_weight >= 1 ? cur : mix(acc, cur, weight), over n floats.
hipError_t launch_accumulate(const float* cur, float* acc, int64_t n, float weight, hipStream_t s);

}  // namespace bhs

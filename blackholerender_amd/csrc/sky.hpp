// sky.hpp — the sky pass (shaders/sky.glsl:57-119 and its call at shaders/blackhole.frag:297-310)
// over a render's ray planes: the backdrop cube map and the star layers added to the colour planes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "footprint.hpp"
#include "sky_device.hpp"

namespace bhs {

// Per escaped pixel g with weight w = weight[g] (1 if weight is null) and w != 0:
// rgb[g + k n] += w * sampleSky(escDir, dFdx, dFdy, uPixelOmega0, rcam)[k], the derivatives and the
// footprint as the footprint pass computes them. One camera (table == nullptr, `one`) or ncam
// cameras of a device table; max_pixels: 64 x the largest camera's count of 8x8-pixel tiles.
hipError_t launch_sky(const MarchConst& one, const MarchConst* table, uint32_t ncam,
                      int64_t max_pixels, const FootprintIn& in, int64_t n,
                      const sky::SkyConst& k, const sky::Cube& cube, const float* lut,
                      const float* weight, float* rgb, bool coarse, hipStream_t s);

}  // namespace bhs

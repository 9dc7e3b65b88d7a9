// sky.hip — the sky pass: shaders/blackhole.frag:297-310 (derivatives of escDir over 2x2 quads,
// the unlensed fallback at status edges, color += transmittance * sampleSky(...)) with
// shaders/sky.glsl:57-119, over the ray planes of a render.
//
// One thread per pixel. An escaped pixel rebuilds the lanes of its quad from the planes (flags,
// escape direction, view direction from the camera) exactly as the footprint pass does, takes its
// derivatives and footprint, and evaluates the backdrop lookup and the two star layers; other
// pixels, and pixels whose transmittance is 0, leave the colour unchanged.
//
// Roofline: VALU. Per escaped pixel the two star layers hash 54 lattice cells (pcg3d) and evaluate
// the stars near the pixel (a Gaussian, a blackbody lookup, exp/log/pow); the memory traffic (the
// quad's flags and escape directions, the weight, the colour read and written: ~70 B per pixel)
// is small next to it.

#include "sky.hpp"

namespace bhs {
namespace {

using namespace fp;

constexpr int kSkyBlock = 256;

// Pixel (col, row) of camera c as a quad lane; a position past the image is a helper pixel
// (not escaped, its view direction continued past the edge).
__device__ __forceinline__ Lane load_lane(const MarchConst& c, const FootprintIn& in, int64_t n,
                                          int32_t col, int32_t row) {
    Lane l;
    const int64_t pix = (int64_t)row * c.width + col;
    l.D = ray_dir_at(c, pix, col, row);
    l.h = col >= c.width || row >= c.height;
    const int64_t g = c.ray_offset + pix;
    l.s = !l.h && (in.flags[g] & kEscapedBit) != 0u;
    l.E = l.s ? V3<float>{in.esc[g], in.esc[n + g], in.esc[2 * n + g]} : l.D;
    return l;
}

template <bool COARSE, bool BATCH>
__global__ __launch_bounds__(kSkyBlock) void k_sky(const MarchConst one,
                                                   const MarchConst* __restrict__ table,
                                                   const FootprintIn in, int64_t n,
                                                   const sky::SkyConst k, const sky::Cube cube,
                                                   const float* __restrict__ lut_g,
                                                   const float* __restrict__ weight,
                                                   float* __restrict__ rgb) {
    __shared__ float lut[shade::kLutSize * 4];
    for (int i = threadIdx.x; i < shade::kLutSize * 4; i += kSkyBlock) lut[i] = lut_g[i];
    __syncthreads();
    const MarchConst c = BATCH ? table[blockIdx.y] : one;
    const int64_t pix = (int64_t)blockIdx.x * kSkyBlock + threadIdx.x;
    if (pix >= (int64_t)c.width * c.height) return;
    const int64_t g = c.ray_offset + pix;
    if ((in.flags[g] & kEscapedBit) == 0u) return;           // blackhole.frag:308
    const float w = weight ? weight[g] : 1.0f;
    if (w == 0.0f) return;
    const int32_t row = (int32_t)(pix / c.width), col = (int32_t)(pix - (int64_t)row * c.width);
    const int32_t x0 = col & ~1, y0 = row & ~1;
    const bool odd_x = (col & 1) != 0, odd_y = (row & 1) != 0;
    Lane h0, h1, v0, v1, self;
    if constexpr (COARSE) {   // pairs (x0, y0)-(x0 + 1, y0) and (x0, y0)-(x0, y0 + 1)
        const Lane q00 = load_lane(c, in, n, x0, y0), q10 = load_lane(c, in, n, x0 + 1, y0);
        const Lane q01 = load_lane(c, in, n, x0, y0 + 1),
                   q11 = load_lane(c, in, n, x0 + 1, y0 + 1);
        h0 = q00; h1 = q10; v0 = q00; v1 = q01;
        self = odd_y ? (odd_x ? q11 : q01) : (odd_x ? q10 : q00);
    } else {                  // the pixel's own row and column pairs: three distinct lanes
        self = load_lane(c, in, n, col, row);
        const Lane hn = load_lane(c, in, n, col ^ 1, row), vn = load_lane(c, in, n, col, row ^ 1);
        h0 = odd_x ? hn : self; h1 = odd_x ? self : hn;
        v0 = odd_y ? vn : self; v1 = odd_y ? self : vn;
    }
    // Renderer.cpp:246-247,331: pixelAngle = 2 tan(fov/2) / h, uPixelOmega0 = pixelAngle^2
    const float pixel_angle = 2.0f * c.tan_half_fov / (float)c.height;
    const float omega0 = pixel_angle * pixel_angle;
    const PixOut p = pix_out(omega0, self, h0, h1, v0, v1);
    const V3<float> P{c.pos[0], c.pos[1], c.pos[2]};
    const float rcam = __builtin_sqrtf(dot3(P, P));
    float s[3];
    sky::sample_sky(k, cube, lut, rcam, self.E, p.dx, p.dy, p.f, omega0, s);
    rgb[g] += w * s[0];                                       // blackhole.frag:310
    rgb[n + g] += w * s[1];
    rgb[2 * n + g] += w * s[2];
}

}  // namespace

hipError_t launch_sky(const MarchConst& one, const MarchConst* table, uint32_t ncam,
                      int64_t max_pixels, const FootprintIn& in, int64_t n,
                      const sky::SkyConst& k, const sky::Cube& cube, const float* lut,
                      const float* weight, float* rgb, bool coarse, hipStream_t s) {
    const dim3 grid((uint32_t)((max_pixels + kSkyBlock - 1) / kSkyBlock), table ? ncam : 1u);
    if (table) {
        if (coarse)
            hipLaunchKernelGGL((k_sky<true, true>), grid, dim3(kSkyBlock), 0, s, one, table, in, n,
                               k, cube, lut, weight, rgb);
        else
            hipLaunchKernelGGL((k_sky<false, true>), grid, dim3(kSkyBlock), 0, s, one, table, in,
                               n, k, cube, lut, weight, rgb);
    } else {
        if (coarse)
            hipLaunchKernelGGL((k_sky<true, false>), grid, dim3(kSkyBlock), 0, s, one, table, in,
                               n, k, cube, lut, weight, rgb);
        else
            hipLaunchKernelGGL((k_sky<false, false>), grid, dim3(kSkyBlock), 0, s, one, table, in,
                               n, k, cube, lut, weight, rgb);
    }
    return hipGetLastError();
}

}  // namespace bhs

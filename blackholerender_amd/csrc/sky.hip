// sky.hip — the sky pass: shaders/blackhole.frag:297-310 (derivatives of escDir over 2x2 quads,
// the unlensed fallback at status edges, color += transmittance * sampleSky(...)) with
// shaders/sky.glsl:57-119, over the ray planes of a render.
//
// One wave per 8x8-pixel tile. A pixel rebuilds its quad lane from the planes (flags, escape
// direction, view direction from the camera) exactly as the footprint pass does, takes its
// derivatives and footprint, and evaluates the backdrop lookup and the two star layers; other
// pixels, and pixels whose transmittance is 0, leave the colour unchanged.
//
// Roofline: VALU. Per tile and layer the wave evaluates the lattice cells of its pixels'
// neighbourhoods (pcg3d, the density and shell tests, the star direction: 36-100 cells for the
// tiles of the benchmark view); per pixel the stars of its 27 cells that pass both tests (~4 and ~2)
// get the quadratic form of the footprint, and the few near the pixel a Gaussian, a blackbody lookup
// and exp/log/pow. The memory traffic (flags, escape direction, weight, colour read and written:
// 41 B per sky pixel) is small next to it.

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

// v from lane ^ 1 (DPP quad_perm [1, 0, 3, 2]) and from lane ^ 8 (DPP row_ror:8, within each
// row of 16 lanes): the row and the column partner of a pixel in an 8x8-pixel wave tile.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                 CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ Lane lane_from(const Lane& l) {
    Lane o;
    o.E = V3<float>{dpp_f<CTRL>(l.E.x), dpp_f<CTRL>(l.E.y), dpp_f<CTRL>(l.E.z)};
    o.D = V3<float>{dpp_f<CTRL>(l.D.x), dpp_f<CTRL>(l.D.y), dpp_f<CTRL>(l.D.z)};
    const int sh = __builtin_amdgcn_update_dpp(0, (l.s ? 1 : 0) | (l.h ? 2 : 0), CTRL, 0xF, 0xF,
                                               false);
    o.s = (sh & 1) != 0;
    o.h = (sh & 2) != 0;
    return o;
}
constexpr int kDppXor1 = 0xB1;    // quad_perm [1, 0, 3, 2]
constexpr int kDppXor8 = 0x128;   // row_ror:8

// One wave per 8x8-pixel tile of a camera (lane = 8 (row & 7) + (col & 7)); every lane rebuilds
// its own quad lane from the planes (a position past the image is a helper pixel, as in the
// footprint pass), fine derivatives take the row and column partners from lanes ^ 1 and ^ 8, and
// all 64 lanes run the star layers together (their shared cell table in LDS); pixels that are not
// escaped, lie past the image or have weight 0 keep their colour.
template <bool COARSE, bool BATCH>
__global__ __launch_bounds__(kSkyBlock) void k_sky(const MarchConst one,
                                                   const MarchConst* __restrict__ table,
                                                   const FootprintIn in, int64_t n,
                                                   const sky::SkyConst k, const sky::Cube cube,
                                                   const float* __restrict__ lut_g,
                                                   const float* __restrict__ weight,
                                                   float* __restrict__ rgb) {
    __shared__ float lut[shade::kLutSize * 4];
    __shared__ uint32_t s_cells[kSkyBlock / 64][sky::kCellWords * sky::kSkyBoxMax];
    for (int i = threadIdx.x; i < shade::kLutSize * 4; i += kSkyBlock) lut[i] = lut_g[i];
    __syncthreads();
    const MarchConst c = BATCH ? table[blockIdx.y] : one;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t tiles_x = ((uint32_t)c.width + 7u) >> 3, tiles_y = ((uint32_t)c.height + 7u) >> 3;
    const uint32_t tile = blockIdx.x * (kSkyBlock / 64) + wave;
    if (tile >= tiles_x * tiles_y) return;                   // the whole wave (no later barrier)
    const int32_t col = (int32_t)((tile % tiles_x) * 8u + (lane & 7u));
    const int32_t row = (int32_t)((tile / tiles_x) * 8u + (lane >> 3));
    const bool inside = col < c.width && row < c.height;
    const int64_t g = c.ray_offset + (int64_t)row * c.width + col;
    const Lane self = load_lane(c, in, n, col, row);
    const float w = inside && self.s ? (weight ? weight[g] : 1.0f) : 0.0f;
    const bool on = inside && self.s && w != 0.0f;           // blackhole.frag:308
    const int32_t x0 = col & ~1, y0 = row & ~1;
    Lane h0, h1, v0, v1;
    if constexpr (COARSE) {   // pairs (x0, y0)-(x0 + 1, y0) and (x0, y0)-(x0, y0 + 1)
        h0 = load_lane(c, in, n, x0, y0);
        h1 = load_lane(c, in, n, x0 + 1, y0);
        v0 = h0;
        v1 = load_lane(c, in, n, x0, y0 + 1);
    } else {                  // the pixel's own row and column pairs, from the partner lanes
        const Lane hn = lane_from<kDppXor1>(self), vn = lane_from<kDppXor8>(self);
        const bool odd_x = (col & 1) != 0, odd_y = (row & 1) != 0;
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
    sky::sample_sky(k, cube, lut, rcam, self.E, p.dx, p.dy, p.f, omega0, on, s_cells[wave], s);
    if (on) {
        rgb[g] += w * s[0];                                   // blackhole.frag:310
        rgb[n + g] += w * s[1];
        rgb[2 * n + g] += w * s[2];
    }
}

}  // namespace

hipError_t launch_sky(const MarchConst& one, const MarchConst* table, uint32_t ncam,
                      int64_t max_pixels, const FootprintIn& in, int64_t n,
                      const sky::SkyConst& k, const sky::Cube& cube, const float* lut,
                      const float* weight, float* rgb, bool coarse, hipStream_t s) {
    // max_pixels: the largest camera's 8x8 tiles x 64 (one wave per tile)
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

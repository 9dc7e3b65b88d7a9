// footprint.hip — sky footprint post-pass over the ray planes of a render.
//
// shaders/blackhole.frag:297-306 takes GLSL derivatives of the escape direction (dFdx/dFdy over
// the rasteriser's 2x2 pixel quads) and falls back to the derivatives of the unlensed view
// direction where a quad mixes escaped and non-escaped pixels; shaders/sky.glsl:27-51
// (makeFootprint) turns them into the footprint the sky filter and the star layers use. Here the
// rays are marched apart from their quads, so the quads are rebuilt from the planes: one thread per
// 2x2 quad reads its four pixels once, rebuilds their view directions from the camera (the same
// ray generation as the march) and writes each pixel's derivatives and footprint.
//
// Semantics (pinned by the CPU restatement in oracle/):
//   * escDir of a pixel is its escape direction if it escaped (flag bit 2), else its view
//     direction (blackhole.frag:98 initialises escDir to dir and replaces it only on escape);
//   * fine derivatives (dFdxFine/dFdyFine): a pixel's d/dx is the difference across the quad's
//     pair in its own row, d/dy across the pair in its own column; coarse: the quad's first row
//     and column for all four pixels;
//   * pixels of a quad that lie outside the image (odd widths or heights) count as not escaped,
//     with the view direction of their position; the reference marches such helper pixels, this
//     pass does not, so their quads take the unlensed footprint wherever the neighbour escaped,
//     and the fallback plane tells these pixels (2) from those at a status edge (1);
//   * the footprint (e1, e2, inverse R factor, area, mu) exists for escaped pixels only
//     (blackhole.frag:308); other pixels get zeros there and their derivatives.
//
// Roofline: HBM. Per pixel it reads 13 B (flags, escape direction) and writes up to 69 B (17
// float planes and the fallback byte); the arithmetic per pixel (a view direction: three divides,
// two square roots; a footprint: eight divides, three square roots; ~60 further operations) stays
// below the memory time (the measured traffic equals the algorithmic bytes).

#include "footprint.hpp"
#include "footprint_device.hpp"

namespace bhs {
namespace {

using namespace fp;

constexpr int kFpBlock = 256;
template <bool VEC>
__device__ __forceinline__ void put2(float* p, int64_t g, float v0, float v1, bool second) {
    if (VEC) {
        *(float2*)(p + g) = float2{v0, v1};
    } else {
        p[g] = v0;
        if (second) p[g + 1] = v1;
    }
}

template <bool VEC>
__device__ __forceinline__ void put3(float* p, int64_t n, int64_t g, const V3<float>& a,
                                     const V3<float>& b, bool second) {
    if (p) {
        put2<VEC>(p, g, a.x, b.x, second);
        put2<VEC>(p + n, g, a.y, b.y, second);
        put2<VEC>(p + 2 * n, g, a.z, b.z, second);
    }
}

// Stores pixels (x0, row) and, if `second`, (x0 + 1, row) starting at plane index g.
template <bool VEC>
__device__ __forceinline__ void store_pair(const FootprintPlanes& o, int64_t g, const PixOut& p,
                                           const PixOut& q, bool second) {
    const int64_t n = o.n;
    put3<VEC>(o.ddx, n, g, p.dx, q.dx, second);
    put3<VEC>(o.ddy, n, g, p.dy, q.dy, second);
    put3<VEC>(o.e1, n, g, p.f.e1, q.f.e1, second);
    put3<VEC>(o.e2, n, g, p.f.e2, q.f.e2, second);
    put3<VEC>(o.rinv, n, g, V3<float>{p.f.r11, p.f.r12, p.f.r22},
              V3<float>{q.f.r11, q.f.r12, q.f.r22}, second);
    if (o.area) put2<VEC>(o.area, g, p.f.area, q.f.area, second);
    if (o.mu) put2<VEC>(o.mu, g, p.f.mu, q.f.mu, second);
    if (o.fallback) {
        if (VEC) {
            *(uint16_t*)(o.fallback + g) = (uint16_t)(p.fb | (q.fb << 8));
        } else {
            o.fallback[g] = (uint8_t)p.fb;
            if (second) o.fallback[g + 1] = (uint8_t)q.fb;
        }
    }
}

// One thread per 2x2 quad. VEC: every camera's width even, planes aligned for paired accesses.
template <bool COARSE, bool BATCH, bool VEC>
__global__ __launch_bounds__(kFpBlock) void k_footprint(const MarchConst one,
                                                        const MarchConst* __restrict__ table,
                                                        const FootprintIn in,
                                                        const FootprintPlanes o) {
    const MarchConst c = BATCH ? table[blockIdx.y] : one;
    const int64_t qw = (c.width + 1) / 2, qh = (c.height + 1) / 2;
    const int64_t q = (int64_t)blockIdx.x * kFpBlock + threadIdx.x;
    if (q >= qw * qh) return;
    const int64_t qy = q / qw;
    const int32_t x0 = (int32_t)(2 * (q - qy * qw)), y0 = (int32_t)(2 * qy);
    // Renderer.cpp:246-247,331: pixelAngle = 2 tan(fov/2) / h, uPixelOmega0 = pixelAngle^2
    const float pixel_angle = 2.0f * c.tan_half_fov / (float)c.height;
    const float omega0 = pixel_angle * pixel_angle;
    const int64_t n = o.n;
    const bool second = x0 + 1 < c.width, row1 = y0 + 1 < c.height;
    Lane a, b, d, e;   // (x0, y0), (x0+1, y0), (x0, y0+1), (x0+1, y0+1)
    load_pair<VEC>(c, in, n, x0, y0, a, b);
    if (VEC && !row1) {   // the second row lies outside the image: not escaped
        const int64_t pix = (int64_t)(y0 + 1) * c.width + x0;
        d.D = ray_dir_at(c, pix, x0, y0 + 1);
        e.D = ray_dir_at(c, pix + 1, x0 + 1, y0 + 1);
        d.E = d.D;
        e.E = e.D;
        d.s = e.s = false;
        d.h = e.h = true;
    } else {
        load_pair<VEC>(c, in, n, x0, y0 + 1, d, e);
    }
    const int64_t g = c.ray_offset + (int64_t)y0 * c.width + x0;
    if (COARSE) {
        store_pair<VEC>(o, g, pix_out(omega0, a, a, b, a, d), pix_out(omega0, b, a, b, a, d),
                        second);
        if (row1)
            store_pair<VEC>(o, g + c.width, pix_out(omega0, d, a, b, a, d),
                            pix_out(omega0, e, a, b, a, d), second);
    } else {
        store_pair<VEC>(o, g, pix_out(omega0, a, a, b, a, d), pix_out(omega0, b, a, b, b, e),
                        second);
        if (row1)
            store_pair<VEC>(o, g + c.width, pix_out(omega0, d, d, e, a, d),
                            pix_out(omega0, e, d, e, b, e), second);
    }
}

template <bool BATCH, bool VEC>
hipError_t launch_fp(const MarchConst& one, const MarchConst* table, uint32_t ncam,
                     int64_t quads, const FootprintIn& in, const FootprintPlanes& o, bool coarse,
                     hipStream_t s) {
    const dim3 grid((uint32_t)((quads + kFpBlock - 1) / kFpBlock), ncam);
    if (coarse)
        hipLaunchKernelGGL((k_footprint<true, BATCH, VEC>), grid, dim3(kFpBlock), 0, s, one,
                           table, in, o);
    else
        hipLaunchKernelGGL((k_footprint<false, BATCH, VEC>), grid, dim3(kFpBlock), 0, s, one,
                           table, in, o);
    return hipGetLastError();
}

// Paired accesses: the planes' element index of every row pair is even (even widths, even ray
// offsets, even plane stride) and every plane base is 8-byte aligned.
bool pairs_aligned(const FootprintIn& in, const FootprintPlanes& o, bool widths_even,
                   bool offsets_even) {
    if (!widths_even || !offsets_even || (o.n & 1)) return false;
    const uintptr_t f = (uintptr_t)in.flags | (uintptr_t)o.fallback;
    const uintptr_t v = (uintptr_t)in.esc | (uintptr_t)o.ddx | (uintptr_t)o.ddy |
                        (uintptr_t)o.e1 | (uintptr_t)o.e2 | (uintptr_t)o.rinv |
                        (uintptr_t)o.area | (uintptr_t)o.mu;
    return (f & 1u) == 0u && (v & 7u) == 0u;
}

}  // namespace

hipError_t launch_footprint(const MarchConst& c, const FootprintIn& in, const FootprintPlanes& o,
                            bool coarse, hipStream_t s) {
    const int64_t quads = (int64_t)((c.width + 1) / 2) * ((c.height + 1) / 2);
    if (quads <= 0) return hipSuccess;
    if (pairs_aligned(in, o, (c.width & 1) == 0, (c.ray_offset & 1) == 0))
        return launch_fp<false, true>(c, nullptr, 1u, quads, in, o, coarse, s);
    return launch_fp<false, false>(c, nullptr, 1u, quads, in, o, coarse, s);
}

hipError_t launch_footprint_batch(const MarchConst* table, uint32_t ncam, int64_t max_quads,
                                  bool widths_even, const FootprintIn& in,
                                  const FootprintPlanes& o, bool coarse, hipStream_t s) {
    if (max_quads <= 0 || ncam == 0) return hipSuccess;
    MarchConst none;
    __builtin_memset(&none, 0, sizeof none);
    // even widths make every camera's ray offset (a sum of width * height) even
    if (pairs_aligned(in, o, widths_even, true))
        return launch_fp<true, true>(none, table, ncam, max_quads, in, o, coarse, s);
    return launch_fp<true, false>(none, table, ncam, max_quads, in, o, coarse, s);
}

}  // namespace bhs

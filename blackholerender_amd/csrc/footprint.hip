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

namespace bhs {
namespace {

constexpr int kFpBlock = 256;
constexpr uint32_t kEscapedBit = 2u;

__device__ __forceinline__ V3<float> sub3(const V3<float>& a, const V3<float>& b) {
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}

__device__ __forceinline__ V3<float> cross3f(const V3<float>& a, const V3<float>& b) {
    return {__builtin_fmaf(a.y, b.z, -(a.z * b.y)), __builtin_fmaf(a.z, b.x, -(a.x * b.z)),
            __builtin_fmaf(a.x, b.y, -(a.y * b.x))};
}

__device__ __forceinline__ V3<float> normalize3f(const V3<float>& v) {
    const float i = 1.0f / __builtin_sqrtf(dot3(v, v));
    return {v.x * i, v.y * i, v.z * i};
}

struct Footprint {
    V3<float> e1, e2;
    float r11, r12, r22, area, mu;
};

// sky.glsl:27-51 (makeFootprint) for escape direction d and derivatives a = d/dx, b = d/dy.
__device__ __forceinline__ Footprint make_footprint(const V3<float>& d, const V3<float>& a,
                                                    const V3<float>& b, float omega0) {
    Footprint f;
    const float min_axis = 1e-3f * __builtin_sqrtf(omega0);   // per-axis magnification <= 1e3
    const float la_raw = __builtin_sqrtf(dot3(a, a));
    const float la = fmaxf(la_raw, min_axis);
    if (la_raw > 1e-12f) {
        f.e1 = {a.x / la_raw, a.y / la_raw, a.z / la_raw};
    } else {
        const V3<float> up{1e-3f, 1.0f + 1e-3f, 1e-3f};
        f.e1 = normalize3f(cross3f(d, up));
    }
    const float b1 = dot3(b, f.e1);
    const V3<float> bp{__builtin_fmaf(-b1, f.e1.x, b.x), __builtin_fmaf(-b1, f.e1.y, b.y),
                       __builtin_fmaf(-b1, f.e1.z, b.z)};
    const float lbp_raw = __builtin_sqrtf(dot3(bp, bp));
    const float lbp = fmaxf(lbp_raw, min_axis);
    if (lbp_raw > 1e-3f * min_axis)
        f.e2 = {bp.x / lbp_raw, bp.y / lbp_raw, bp.z / lbp_raw};
    else
        f.e2 = normalize3f(cross3f(d, f.e1));
    f.r11 = 1.0f / la;
    f.r12 = -b1 / (la * lbp);
    f.r22 = 1.0f / lbp;
    f.area = la * lbp;
    f.mu = fminf(omega0 / f.area, 1e4f);
    return f;
}

// One pixel of a quad: effective escDir E, view direction D, escape flag, outside the image.
struct Lane {
    V3<float> E, D;
    bool s, h;
};

// Pixels (x0, row) and (x0 + 1, row) of a quad row. VEC: both inside the image, their planes
// 8-byte aligned (paired loads); otherwise a pixel outside the image counts as not escaped.
template <bool VEC>
__device__ __forceinline__ void load_pair(const MarchConst& c, const FootprintIn& in, int64_t n,
                                          int32_t x0, int32_t row, Lane& l0, Lane& l1) {
    const int64_t pix = (int64_t)row * c.width + x0;
    l0.D = ray_dir_at(c, pix, x0, row);
    l1.D = ray_dir_at(c, pix + 1, x0 + 1, row);
    const int64_t g = c.ray_offset + pix;
    if (VEC) {
        const uint32_t f = *(const uint16_t*)(in.flags + g);
        const float2 ex = *(const float2*)(in.esc + g);
        const float2 ey = *(const float2*)(in.esc + n + g);
        const float2 ez = *(const float2*)(in.esc + 2 * n + g);
        l0.s = (f & kEscapedBit) != 0u;
        l1.s = ((f >> 8) & kEscapedBit) != 0u;
        l0.h = l1.h = false;
        l0.E = l0.s ? V3<float>{ex.x, ey.x, ez.x} : l0.D;
        l1.E = l1.s ? V3<float>{ex.y, ey.y, ez.y} : l1.D;
    } else {
        const bool in_row = row < c.height;
        l0.h = !in_row;
        l1.h = !in_row || x0 + 1 >= c.width;
        l0.s = !l0.h && (in.flags[g] & kEscapedBit) != 0u;
        l1.s = !l1.h && (in.flags[g + 1] & kEscapedBit) != 0u;
        l0.E = l0.s ? V3<float>{in.esc[g], in.esc[n + g], in.esc[2 * n + g]} : l0.D;
        l1.E = l1.s ? V3<float>{in.esc[g + 1], in.esc[n + g + 1], in.esc[2 * n + g + 1]} : l1.D;
    }
}

// Outputs of one pixel. fb: 0, 1 at a status edge, 2 where the edge runs through a helper pixel.
struct PixOut {
    V3<float> dx, dy;
    Footprint f;
    bool edge;
    uint32_t fb;
};

// Pixel `self` with horizontal pair (h0, h1) and vertical pair (v0, v1).
__device__ __forceinline__ PixOut pix_out(float omega0, const Lane& self, const Lane& h0,
                                          const Lane& h1, const Lane& v0, const Lane& v1) {
    PixOut p;
    const bool ex = h0.s != h1.s, ey = v0.s != v1.s;
    p.edge = ex || ey;   // blackhole.frag:303-306
    p.fb = p.edge ? (((ex && (h0.h || h1.h)) || (ey && (v0.h || v1.h))) ? 2u : 1u) : 0u;
    p.dx = p.edge ? sub3(h1.D, h0.D) : sub3(h1.E, h0.E);
    p.dy = p.edge ? sub3(v1.D, v0.D) : sub3(v1.E, v0.E);
    // blackhole.frag:308: the footprint of escaped pixels only (zeros elsewhere); computed for
    // every pixel and selected per field, which keeps the record in registers
    const Footprint f = make_footprint(self.E, p.dx, p.dy, omega0);
    const bool s = self.s;
    p.f.e1 = V3<float>{s ? f.e1.x : 0.0f, s ? f.e1.y : 0.0f, s ? f.e1.z : 0.0f};
    p.f.e2 = V3<float>{s ? f.e2.x : 0.0f, s ? f.e2.y : 0.0f, s ? f.e2.z : 0.0f};
    p.f.r11 = s ? f.r11 : 0.0f;
    p.f.r12 = s ? f.r12 : 0.0f;
    p.f.r22 = s ? f.r22 : 0.0f;
    p.f.area = s ? f.area : 0.0f;
    p.f.mu = s ? f.mu : 0.0f;
    return p;
}

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

// footprint_device.hpp — the per-quad part of the sky footprint (shaders/blackhole.frag:297-306,
// shaders/sky.glsl:27-51) as device functions, shared by the footprint pass (footprint.hip) and the
// sky pass (sky.hip). The semantics are stated at the top of footprint.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "footprint.hpp"
#include "lawS_device.hpp"

namespace bhs {
namespace fp {

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

}  // namespace fp
}  // namespace bhs

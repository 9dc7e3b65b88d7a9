// shade_device.hpp — thin-disk shading of shaders/disk.glsl (thinDiskSample :113-130 with
// diskNoise :46-66, diskTempProfile :27-31, redshiftFactor :37-42, opticalDepthScale :68-72,
// diskRadialProfile :74-78, diskTint :22), blackbody (shaders/blackbody.glsl:8-16), the 3D noise
// texture baked by shaders/noise_bake.comp (with pgnoise3, grad3, pcg3d of shaders/common.glsl:21-82),
// and the transmittance compositing of shaders/blackhole.frag:209-216,234.
//
// Arithmetic contract: like the march, everything here is one fixed sequence of IEEE operations
// (+ - * /, sqrt, fma, floor, rint, frexp, ldexp, conversions), with no contraction, so a CPU
// restatement of the same sequence reproduces every output bit for bit. GLSL's transcendentals
// (exp, log, pow, inversesqrt) have implementation-defined precision; here they are defined as:
//   exp2(x): x clamped to [-126, 127]; n = rint(x), f = x - n (exact); 2^f by its Taylor series
//            sum_k (f ln 2)^k / k! to degree 7 (truncation below 6e-9 relative), Horner in fma;
//            ldexp by n.
//   log2(x): x = m 2^e with m in [sqrt(1/2), sqrt(2)); s = (m - 1)/(m + 1);
//            ln m = 2 atanh s = 2 (s + s^3/3 + ... + s^11/11) (truncation below 5e-11); e + ln m log2 e.
//   exp(x) = exp2(x log2 e), log(x) = e ln 2 + ln m, pow(x, y) = exp2(y log2 x) for x > 0 (0 for
//   x <= 0; every exponent used is positive), inversesqrt(x) = 1 / sqrt(x).
// GLSL's mix(a, b, t) = a (1 - t) + b t, smoothstep(e0, e1, x) = t t (3 - 2 t) with
// t = clamp((x - e0)/(e1 - e0), 0, 1), dot(a, b) = (a.x b.x + a.y b.y) + a.z b.z, products and
// sums evaluated left to right as written in the shaders.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bhs {
namespace shade {

constexpr int kNoiseSize = 128;        // DiskTextures::kNoiseSize (src/DiskTextures.h:16)
constexpr int kLutSize = 256;          // DiskTextures::kLutSize (src/DiskTextures.h:15)
constexpr uint32_t kNoiseSeed = 7u;    // src/DiskTextures.cpp:110
constexpr float kTwoPi = 6.28318530717959f;   // common.glsl:7
constexpr float kLog2e = 1.44269504f;
constexpr float kLn2 = 0.693147181f;
constexpr float kOpaque = 0.004f;      // blackhole.frag:234

// Disk shading uniforms (src/Renderer.cpp:328-343).
struct ShadeConst {
    float density, opacity, peak_temp, brightness, noise_scale, noise_strength, warmth, doppler,
        time, inflow;
    float inner, outer;
};

// ------------------------------------------------------------------ IEEE-only elementary functions
__device__ __forceinline__ float exp2_ieee(float x) {
    x = x < -126.0f ? -126.0f : (x > 127.0f ? 127.0f : x);
    const float n = __builtin_rintf(x);
    const float f = x - n;
    float p = 1.5252734e-05f;                   // ln2^7 / 7!
    p = __builtin_fmaf(p, f, 1.5403530e-04f);   // ln2^6 / 6!
    p = __builtin_fmaf(p, f, 1.3333558e-03f);   // ln2^5 / 5!
    p = __builtin_fmaf(p, f, 9.6181291e-03f);   // ln2^4 / 4!
    p = __builtin_fmaf(p, f, 5.5504109e-02f);   // ln2^3 / 3!
    p = __builtin_fmaf(p, f, 2.4022651e-01f);   // ln2^2 / 2!
    p = __builtin_fmaf(p, f, 6.9314718e-01f);   // ln2
    p = __builtin_fmaf(p, f, 1.0f);
    return __builtin_ldexpf(p, (int)n);
}

// ln of the mantissa m (in [sqrt(1/2), sqrt(2))) and the exponent e of x > 0.
__device__ __forceinline__ float ln_mantissa(float x, int& e) {
    float m = __builtin_frexpf(x, &e);
    if (m < 0.70710678f) {
        m = m * 2.0f;
        e = e - 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float t = s * s;
    float q = 9.0909091e-02f;                   // 1/11
    q = __builtin_fmaf(q, t, 1.1111111e-01f);   // 1/9
    q = __builtin_fmaf(q, t, 1.4285715e-01f);   // 1/7
    q = __builtin_fmaf(q, t, 2.0000000e-01f);   // 1/5
    q = __builtin_fmaf(q, t, 3.3333334e-01f);   // 1/3
    const float s2 = s + s;
    return __builtin_fmaf(s2 * t, q, s2);
}

__device__ __forceinline__ float log2_ieee(float x) {
    int e;
    const float lm = ln_mantissa(x, e);
    return __builtin_fmaf(lm, kLog2e, (float)e);
}

__device__ __forceinline__ float log_ieee(float x) {
    int e;
    const float lm = ln_mantissa(x, e);
    return __builtin_fmaf((float)e, kLn2, lm);
}

__device__ __forceinline__ float exp_ieee(float x) { return exp2_ieee(x * kLog2e); }

__device__ __forceinline__ float pow_ieee(float x, float y) {
    return x > 0.0f ? exp2_ieee(y * log2_ieee(x)) : 0.0f;
}

__device__ __forceinline__ float clampf(float x, float lo, float hi) {
    return x < lo ? lo : (x > hi ? hi : x);
}
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float smoothstepf(float e0, float e1, float x) {
    const float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return (t * t) * (3.0f - 2.0f * t);
}
// GLSL mod(x, y) = x - y floor(x / y).
__device__ __forceinline__ float modf_glsl(float x, float y) { return x - y * __builtin_floorf(x / y); }

// ------------------------------------------------------------------ half precision (R16F texels)
// float -> half, round to nearest even (the R16F store of noise_bake.comp); finite x only.
__device__ __forceinline__ uint16_t f2h(float x) {
    const uint32_t u = __float_as_uint(x);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const uint32_t a = u & 0x7fffffffu;
    if (a >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);   // rounds to infinity
    if (a < 0x38800000u) {                                     // half subnormal (or zero)
        const float m = __builtin_rintf(__uint_as_float(a) * 0x1p+24f);   // units of 2^-24
        return (uint16_t)(sign | (uint32_t)m);
    }
    uint32_t h = (((a >> 23) - 112u) << 10) | ((a >> 13) & 0x3ffu);
    const uint32_t rem = a & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h += 1u;
    return (uint16_t)(sign | h);
}

__device__ __forceinline__ float h2f(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    if (e == 0u) {
        const float v = (float)m * 0x1p-24f;
        return sign ? -v : v;
    }
    return __uint_as_float(sign | ((e + 112u) << 23) | (m << 13));
}

// ------------------------------------------------------------------ gradient noise (common.glsl)
struct U3 { uint32_t x, y, z; };

__device__ __forceinline__ U3 pcg3d(U3 v) {   // common.glsl:21-27
    v.x = v.x * 1664525u + 1013904223u;
    v.y = v.y * 1664525u + 1013904223u;
    v.z = v.z * 1664525u + 1013904223u;
    v.x += v.y * v.z; v.y += v.z * v.x; v.z += v.x * v.y;
    v.x ^= v.x >> 16u; v.y ^= v.y >> 16u; v.z ^= v.z >> 16u;
    v.x += v.y * v.z; v.y += v.z * v.x; v.z += v.x * v.y;
    return v;
}

// grad3 (common.glsl:41-45): a hashed unit vector of lattice point c.
__device__ __forceinline__ void grad3(int cx, int cy, int cz, uint32_t salt, float g[3]) {
    const U3 h = pcg3d(U3{(uint32_t)cx + salt, (uint32_t)cy + (salt * 3u + 11u),
                          (uint32_t)cz + (salt * 7u + 29u)});
    const float k = 0x1p-32f;
    const float gx = ((float)h.x * k) * 2.0f - 1.0f;
    const float gy = ((float)h.y * k) * 2.0f - 1.0f;
    const float gz = ((float)h.z * k) * 2.0f - 1.0f;
    float d = (gx * gx + gy * gy) + gz * gz;
    d = d > 1e-8f ? d : 1e-8f;
    const float is = 1.0f / __builtin_sqrtf(d);
    g[0] = gx * is;
    g[1] = gy * is;
    g[2] = gz * is;
}

__device__ __forceinline__ float corner(int cx, int cy, int cz, uint32_t salt, float fx, float fy,
                                        float fz) {
    float g[3];
    grad3(cx, cy, cz, salt, g);
    return (g[0] * fx + g[1] * fy) + g[2] * fz;
}

__device__ __forceinline__ float fade(float f) {
    return ((f * f) * f) * (f * (f * 6.0f - 15.0f) + 10.0f);
}

__device__ __forceinline__ int wrap_period(int c, int period) { return ((c % period) + period) % period; }

// pgnoise3 (common.glsl:65-82): gradient noise whose lattice wraps every `period` cells.
__device__ __forceinline__ float pgnoise3(float px, float py, float pz, int period, uint32_t salt) {
    const float ix = __builtin_floorf(px), iy = __builtin_floorf(py), iz = __builtin_floorf(pz);
    const float fx = px - ix, fy = py - iy, fz = pz - iz;
    const float ux = fade(fx), uy = fade(fy), uz = fade(fz);
    const int cx = (int)ix, cy = (int)iy, cz = (int)iz;
    const int x0 = wrap_period(cx, period), y0 = wrap_period(cy, period), z0 = wrap_period(cz, period);
    const int x1 = wrap_period(cx + 1, period), y1 = wrap_period(cy + 1, period),
              z1 = wrap_period(cz + 1, period);
    const float gx = fx - 1.0f, gy = fy - 1.0f, gz = fz - 1.0f;
    const float n000 = corner(x0, y0, z0, salt, fx, fy, fz);
    const float n100 = corner(x1, y0, z0, salt, gx, fy, fz);
    const float n010 = corner(x0, y1, z0, salt, fx, gy, fz);
    const float n110 = corner(x1, y1, z0, salt, gx, gy, fz);
    const float n001 = corner(x0, y0, z1, salt, fx, fy, gz);
    const float n101 = corner(x1, y0, z1, salt, gx, fy, gz);
    const float n011 = corner(x0, y1, z1, salt, fx, gy, gz);
    const float n111 = corner(x1, y1, z1, salt, gx, gy, gz);
    const float nx00 = mixf(n000, n100, ux), nx10 = mixf(n010, n110, ux);
    const float nx01 = mixf(n001, n101, ux), nx11 = mixf(n011, n111, ux);
    return 1.4f * mixf(mixf(nx00, nx10, uy), mixf(nx01, nx11, uy), uz);
}

// One texel of the tileable fbm volume (noise_bake.comp:9-27) at integer coordinates (x, y, z).
__device__ __forceinline__ float noise_texel(int x, int y, int z, uint32_t seed) {
    const float inv = (float)kNoiseSize;
    const float u = ((float)x + 0.5f) / inv, v = ((float)y + 0.5f) / inv, w = ((float)z + 0.5f) / inv;
    float sum = 0.0f, amp = 0.55f, norm = 0.0f;
    int period = 4;
    for (int o = 0; o < 6; ++o) {
        const float pf = (float)period;
        sum += amp * pgnoise3(u * pf, v * pf, w * pf, period, seed + (uint32_t)o * 131u);
        norm += amp;
        amp *= 0.55f;
        period *= 2;
    }
    float n = clampf(0.5f + (0.5f * (sum / norm)) * 1.6f, 0.0f, 1.0f);
    const float ridge = 1.0f - __builtin_fabsf(pgnoise3(u * 8.0f, v * 8.0f, w * 8.0f, 8, seed + 977u));
    n = mixf(n, (n * ridge) * 1.3f, 0.35f);
    return n;
}

// ------------------------------------------------------------------ texture reads
// Trilinear GL_LINEAR / GL_REPEAT read of the 128^3 R16F volume at normalised coordinates (s, t, r):
// texel centres at (i + 0.5) / 128, weights u - floor(u) with u = c 128 - 0.5, x interpolated
// first, then y, then z.
__device__ __forceinline__ float sample_noise(const uint16_t* __restrict__ vol, float s, float t,
                                              float r) {
    const float us = s * 128.0f - 0.5f, ut = t * 128.0f - 0.5f, ur = r * 128.0f - 0.5f;
    const float fs = __builtin_floorf(us), ft = __builtin_floorf(ut), fr = __builtin_floorf(ur);
    const float ws = us - fs, wt = ut - ft, wr = ur - fr;
    const int x0 = (int)fs & 127, y0 = (int)ft & 127, z0 = (int)fr & 127;
    const int x1 = (x0 + 1) & 127, y1 = (y0 + 1) & 127, z1 = (z0 + 1) & 127;
    auto at = [&](int x, int y, int z) { return h2f(vol[((size_t)z * 128 + y) * 128 + x]); };
    const float l00 = mixf(at(x0, y0, z0), at(x1, y0, z0), ws);
    const float l10 = mixf(at(x0, y1, z0), at(x1, y1, z0), ws);
    const float l01 = mixf(at(x0, y0, z1), at(x1, y0, z1), ws);
    const float l11 = mixf(at(x0, y1, z1), at(x1, y1, z1), ws);
    return mixf(mixf(l00, l10, wt), mixf(l01, l11, wt), wr);
}

// Linear GL_CLAMP_TO_EDGE read of the 256-entry RGBA blackbody table at t in [0, 1].
__device__ __forceinline__ void sample_lut(const float* __restrict__ lut, float t, float o[4]) {
    const float u = t * 256.0f - 0.5f;
    const float fu = __builtin_floorf(u);
    const float w = u - fu;
    int i0 = (int)fu, i1 = i0 + 1;
    i0 = i0 < 0 ? 0 : (i0 > 255 ? 255 : i0);
    i1 = i1 < 0 ? 0 : (i1 > 255 ? 255 : i1);
    for (int k = 0; k < 4; ++k) o[k] = mixf(lut[4 * i0 + k], lut[4 * i1 + k], w);
}

// blackbody (blackbody.glsl:8-16): rgb of luminance 1 and visible luminance relative to 6500 K.
__device__ __forceinline__ void blackbody(const float* __restrict__ lut, float T, float o[4]) {
    const float tc = clampf(T, 1000.0f, 40000.0f);
    const float lmin = log_ieee(1000.0f), lmax = log_ieee(40000.0f);
    sample_lut(lut, (log_ieee(tc) - lmin) / (lmax - lmin), o);
    const float under = clampf(T / 1000.0f, 0.0f, 1.0f);
    o[3] *= ((((under * under) * under) * under) * under) * under;
}

// ------------------------------------------------------------------ the disk model (disk.glsl)
__device__ __forceinline__ float temp_profile(const ShadeConst& c, float r) {   // :27-31
    const float x = c.inner / (r > c.inner ? r : c.inner);
    const float om = 1.0f - __builtin_sqrtf(x);
    const float f = pow_ieee(x, 0.75f) * pow_ieee(om > 0.0f ? om : 0.0f, 0.25f);
    return f / 0.48796f;
}

__device__ __forceinline__ float radial_profile(const ShadeConst& c, float r) {   // :74-78
    const float in = smoothstepf(c.inner, c.inner + 0.25f, r);
    const float out = 1.0f - smoothstepf(c.outer * 0.55f, c.outer, r);
    return (in * out) * pow_ieee(c.inner / (r > c.inner ? r : c.inner), 1.2f);
}

// diskNoise (:46-66) at the disk plane; omega = r^-1.5 = 1 / (r sqrt r) as in redshiftFactor.
__device__ __forceinline__ float disk_noise(const ShadeConst& c, const uint16_t* __restrict__ vol,
                                            float r, float phi, float y_over_h) {
    const float omega = 1.0f / (r * __builtin_sqrtf(r));
    const float t = c.time;
    const float P = 60.0f;
    const float t1 = modf_glsl(t, P);
    const float t2 = modf_glsl(t + 0.5f * P, P);
    const float w2 = __builtin_fabsf((2.0f * t1) / P - 1.0f);
    const float rep0 = __builtin_floorf(2.0f * c.noise_scale + 0.5f);
    const float rep = rep0 > 1.0f ? rep0 : 1.0f;
    const float lr = (log_ieee(r) * 3.2f) * c.noise_scale;
    const float yy = y_over_h * 0.35f;
    const float ang = rep / kTwoPi;
    const float io = c.inflow * omega;
    const float n1 = sample_noise(vol, (phi + omega * t1) * ang, lr + io * t1, yy);
    const float n2 = sample_noise(vol, (phi + omega * t2) * ang, (lr + 0.37f) + io * t2, yy + 0.61f);
    return mixf(n1, n2, w2);
}

// thinDiskSample (:113-130) of a crossing at radius r (inside the annulus), azimuth phi, with the
// march's redshift factor gphys of that crossing; lapse_cam = sqrt(max(1 - 2/rcam, 1e-4)).
__device__ __forceinline__ void thin_disk_sample(const ShadeConst& c, const uint16_t* __restrict__ vol,
                                                 const float* __restrict__ lut, float lapse_cam,
                                                 float r, float phi, float gphys, float rgb[3],
                                                 float& alpha) {
    const float n = disk_noise(c, vol, r, phi, 0.0f);
    const float f = smoothstepf(0.32f, 0.78f, n);
    const float pattern = mixf(1.0f, mixf(0.08f, 1.4f, f), c.noise_strength);
    const float emit = mixf(1.0f, mixf(0.22f, 1.35f, f * f), c.noise_strength);
    const float dens = (c.density * radial_profile(c, r)) * pattern;
    const float T = (c.peak_temp * temp_profile(c, r)) * mixf(1.0f, 0.85f + 0.3f * emit, 0.5f);
    const float g = mixf(1.0f, gphys, c.doppler);
    float bb[4];
    blackbody(lut, T * g, bb);
    const float l2 = 1.0f - 2.0f / r;                                  // opticalDepthScale :68-72
    const float lapse_em = __builtin_sqrtf(l2 > 1e-4f ? l2 : 1e-4f);
    const float ods = mixf(1.0f, lapse_em / (gphys * lapse_cam), c.doppler);
    alpha = 1.0f - exp_ieee(((-dens * c.opacity) * 0.6f) * ods);
    const float k = (bb[3] * c.brightness) * emit;
    rgb[0] = (bb[0] * mixf(1.0f, 1.0f, c.warmth)) * k;                 // diskTint :22
    rgb[1] = (bb[1] * mixf(1.0f, 0.80f, c.warmth)) * k;
    rgb[2] = (bb[2] * mixf(1.0f, 0.58f, c.warmth)) * k;
}

// Running composite of one ray (blackhole.frag:213-215, 234).
struct Composite {
    float rgb[3];
    float T;
    bool opaque;
    __device__ __forceinline__ void add(const float s[3], float a) {
        rgb[0] += (T * s[0]) * a;
        rgb[1] += (T * s[1]) * a;
        rgb[2] += (T * s[2]) * a;
        T *= 1.0f - a;
        if (T < kOpaque) opaque = true;
    }
};

}  // namespace shade
}  // namespace bhs

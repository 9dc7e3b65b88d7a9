// sky_device.hpp — the sky of one escaped pixel: shaders/sky.glsl:57-119 (starLayer, sampleSky)
// with the backdrop's cube map lookup (textureGrad on a seamless, trilinearly filtered cube map,
// src/SkyBaker.cpp:83-91, src/App.cpp:216).
//
// Semantics (the CPU restatement in oracle/sky.inc follows the same operation order):
//   * the cube map lookup is the OpenGL 4.6 specification's ideal one: face and (s, t) by the
//     major axis (Table 8.19, ties to x, then y); the level of detail lambda = log2(rho),
//     rho = max over the two screen axes of |d(u, v)| in level-0 texels, with d(s, t) from the
//     quotient rule on the face; lambda <= 0: bilinear on level 0, else bilinear on levels
//     floor(lambda) and floor(lambda) + 1 mixed by frac(lambda) (level levels-1 alone at or past
//     it); bilinear weights in float, x first; seamless filtering: a texel past a face edge is
//     the texel of the adjacent face under the same direction, a texel past a corner the mean of
//     the three texels that meet there. GPUs approximate rho and quantise the weights, so the
//     reference's own images differ from GPU to GPU in the last bits; this is the exact form.
//   * the star layers follow sky.glsl:57-106 operation for operation, with GLSL's inversesqrt as
//     the correctly rounded reciprocal square root, pow(x, -1) as 1 / x, exp, log and pow as the
//     IEEE-only functions of the shading.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "footprint_device.hpp"
#include "lawS_device.hpp"
#include "shade_device.hpp"

namespace bhs {
namespace sky {

using namespace shade;

constexpr int kMaxLevels = 16;

// Cube map of the backdrop: level l (0..levels-1) has faces of S_l = max(1, res >> l) texels;
// texel (i, j) of face f of level l (s = (i + 0.5) / S_l, t = (j + 0.5) / S_l; faces +X -X +Y -Y
// +Z -Z) at texels[off[l] + (f * S_l + j) * S_l + i]. texels == nullptr: no backdrop.
struct Cube {
    const float4* texels;
    int32_t res, levels;
    int64_t off[kMaxLevels];
};

// Per-call constants of sampleSky (sky.glsl:108-119) and their uniforms.
struct SkyConst {
    float sky_brightness;   // uSkyBrightness
    float star_density;     // uStarDensity
    float star_brightness;  // uStarBrightness
    float doppler;          // uDopplerStrength
};

// Component i (0, 1, 2) of (x, y, z), by selects (no indexed local array).
__device__ __forceinline__ float pick3(float x, float y, float z, int i) {
    return i == 0 ? x : (i == 1 ? y : z);
}

// Face, sc, tc and |ma| of direction (x, y, z) (OpenGL 4.6 Table 8.19); is, it, im: the
// components behind sc, tc and the major axis, ss, st: the signs sc and tc take them with.
__device__ __forceinline__ int cube_face(float x, float y, float z, float& sc, float& tc,
                                         float& ma, int& is, int& it, int& im, float& ss,
                                         float& st) {
    const float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y), az = __builtin_fabsf(z);
    int f;
    if (ax >= ay && ax >= az) {
        f = x >= 0.0f ? 0 : 1;
        im = 0; is = 2; it = 1;
        ss = x >= 0.0f ? -1.0f : 1.0f; st = -1.0f;
        ma = ax;
    } else if (ay >= az) {
        f = y >= 0.0f ? 2 : 3;
        im = 1; is = 0; it = 2;
        ss = 1.0f; st = y >= 0.0f ? 1.0f : -1.0f;
        ma = ay;
    } else {
        f = z >= 0.0f ? 4 : 5;
        im = 2; is = 0; it = 1;
        ss = z >= 0.0f ? 1.0f : -1.0f; st = -1.0f;
        ma = az;
    }
    sc = ss * pick3(x, y, z, is);
    tc = st * pick3(x, y, z, it);
    return f;
}

// The direction of face f at (sc, tc) in [-1, 1]^2 beyond (inverse of cube_face; the bake's
// cubeDir, sky_bake.comp, unnormalised).
__device__ __forceinline__ V3<float> face_dir(int f, float sc, float tc) {
    switch (f) {
        case 0: return {1.0f, -tc, -sc};
        case 1: return {-1.0f, -tc, sc};
        case 2: return {sc, 1.0f, tc};
        case 3: return {sc, -1.0f, -tc};
        case 4: return {sc, -tc, 1.0f};
        default: return {-sc, -tc, -1.0f};
    }
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ float4 texel_at(const Cube& c, int l, int S, int f, int i, int j) {
    return c.texels[c.off[l] + ((int64_t)f * S + j) * S + i];
}

// Texel (i, j) of face f at level l with exactly one of i, j outside [0, S): the texel of the
// adjacent face under the direction of the texel's centre.
__device__ __forceinline__ float4 texel_across(const Cube& c, int l, int S, int f, int i, int j) {
    const float fS = (float)S;
    const float sc = (2.0f * ((float)i + 0.5f)) / fS - 1.0f;
    const float tc = (2.0f * ((float)j + 0.5f)) / fS - 1.0f;
    const V3<float> d = face_dir(f, sc, tc);
    float sc2, tc2, ma2, ss, st;
    int is, it, im;
    const int f2 = cube_face(d.x, d.y, d.z, sc2, tc2, ma2, is, it, im, ss, st);
    const float s2 = 0.5f * (sc2 / ma2 + 1.0f), t2 = 0.5f * (tc2 / ma2 + 1.0f);
    const int i2 = clampi((int)__builtin_floorf(s2 * fS), 0, S - 1);
    const int j2 = clampi((int)__builtin_floorf(t2 * fS), 0, S - 1);
    return texel_at(c, l, S, f2, i2, j2);
}

// Texel (i, j) of face f at level l, seamless: inside the face itself, past one edge the texel
// across it, past a corner the mean of the three texels that meet there.
__device__ __forceinline__ float4 texel_seamless(const Cube& c, int l, int S, int f, int i, int j) {
    const bool oi = i < 0 || i >= S, oj = j < 0 || j >= S;
    if (!oi && !oj) return texel_at(c, l, S, f, i, j);
    if (oi != oj) return texel_across(c, l, S, f, i, j);
    const int ic = clampi(i, 0, S - 1), jc = clampi(j, 0, S - 1);
    const float4 a = texel_at(c, l, S, f, ic, jc);
    const float4 b = texel_across(c, l, S, f, i, jc);
    const float4 d = texel_across(c, l, S, f, ic, j);
    return float4{((a.x + b.x) + d.x) / 3.0f, ((a.y + b.y) + d.y) / 3.0f,
                  ((a.z + b.z) + d.z) / 3.0f, 0.0f};
}

// Bilinear read of face f at level l at (s, t), seamless across the face's edges.
__device__ __forceinline__ void bilinear(const Cube& c, int l, int f, float s, float t, float o[3]) {
    const int S = (c.res >> l) > 1 ? (c.res >> l) : 1;
    const float u = s * (float)S - 0.5f, v = t * (float)S - 0.5f;
    const float fu = __builtin_floorf(u), fv = __builtin_floorf(v);
    const float a = u - fu, b = v - fv;
    const int i0 = (int)fu, j0 = (int)fv;
    const float4 t00 = texel_seamless(c, l, S, f, i0, j0);
    const float4 t10 = texel_seamless(c, l, S, f, i0 + 1, j0);
    const float4 t01 = texel_seamless(c, l, S, f, i0, j0 + 1);
    const float4 t11 = texel_seamless(c, l, S, f, i0 + 1, j0 + 1);
    o[0] = mixf(mixf(t00.x, t10.x, a), mixf(t01.x, t11.x, a), b);
    o[1] = mixf(mixf(t00.y, t10.y, a), mixf(t01.y, t11.y, a), b);
    o[2] = mixf(mixf(t00.z, t10.z, a), mixf(t01.z, t11.z, a), b);
}

// textureGrad(cube, d, dx, dy).rgb with GL_LINEAR_MIPMAP_LINEAR / GL_LINEAR.
__device__ __forceinline__ void cube_sample(const Cube& c, const V3<float>& d, const V3<float>& dx,
                                            const V3<float>& dy, float o[3]) {
    float sc, tc, ma, ss, st;
    int is, it, im;
    const int f = cube_face(d.x, d.y, d.z, sc, tc, ma, is, it, im, ss, st);
    const float s = 0.5f * (sc / ma + 1.0f), t = 0.5f * (tc / ma + 1.0f);
    // d|ma| = sign(major) d(major); d(sc) = ss d(v[is]); d(tc) = st d(v[it])
    const float sm = pick3(d.x, d.y, d.z, im) >= 0.0f ? 1.0f : -1.0f;
    const float ma2 = ma * ma, res = (float)c.res;
    float rho = 0.0f;
    for (int k = 0; k < 2; ++k) {
        const V3<float>& dd = k == 0 ? dx : dy;
        const float dsc = ss * pick3(dd.x, dd.y, dd.z, is), dtc = st * pick3(dd.x, dd.y, dd.z, it);
        const float dma = sm * pick3(dd.x, dd.y, dd.z, im);
        const float dsdx = 0.5f * (__builtin_fmaf(dsc, ma, -(sc * dma)) / ma2);
        const float dtdx = 0.5f * (__builtin_fmaf(dtc, ma, -(tc * dma)) / ma2);
        const float du = dsdx * res, dvv = dtdx * res;
        const float r = __builtin_sqrtf(__builtin_fmaf(dvv, dvv, du * du));
        rho = r > rho ? r : rho;
    }
    const float lam = rho > 0.0f ? log2_ieee(rho) : -1.0f;
    const int q = c.levels - 1;
    if (!(lam > 0.0f)) {
        bilinear(c, 0, f, s, t, o);
    } else if (lam >= (float)q) {
        bilinear(c, q, f, s, t, o);
    } else {
        const float fl = __builtin_floorf(lam);
        const int d1 = (int)fl;
        float o1[3], o2[3];
        bilinear(c, d1, f, s, t, o1);
        bilinear(c, d1 + 1, f, s, t, o2);
        const float w = lam - fl;
        for (int k = 0; k < 3; ++k) o[k] = mixf(o1[k], o2[k], w);
    }
}

// starLayer (sky.glsl:57-106) for escaped direction d with footprint fp; lut: blackbody table.
__device__ __forceinline__ void star_layer(const float* __restrict__ lut, const V3<float>& d,
                                           const fp::Footprint& fp, float N, float prob,
                                           float flux, float sig_star, float omega0,
                                           uint32_t salt, float g_shift, float o[3]) {
    const float det = fp.area * fp.area;                                 // Footprint.det
    const float cells = pow_ieee(det, 0.25f) * N;                        // :61
    const float w = smoothstepf(0.8f, 2.5f, cells);                      // :62
    float bb[4];
    blackbody(lut, 5600.0f * g_shift, bb);
    const float ka = ((((prob * N) * N) * 9.517f) * flux) * omega0;      // :63-64, MEAN_FLUX
    float sum[3] = {0.0f, 0.0f, 0.0f};
    if (w < 1.0f) {
        const float s2 = sig_star * sig_star, px2 = 0.45f;
        const float gi11 = fp.r11 * fp.r11 + fp.r12 * fp.r12;            // Footprint.Gi11
        const float gi12 = fp.r12 * fp.r22, gi22 = fp.r22 * fp.r22;
        const float rr = fp.r11 * fp.r22;
        const float det_gi = rr * rr;
        const float saa = px2 + s2 * gi11, sab = s2 * gi12, sbb = px2 + s2 * gi22;   // :70
        const float det_s = (px2 * px2 + (px2 * s2) * (gi11 + gi22)) + (s2 * s2) * det_gi;
        const float inv_det_s = 1.0f / det_s;
        const float isaa = sbb * inv_det_s, isab = -sab * inv_det_s, isbb = saa * inv_det_s;
        const float norm = fp.mu / (kTwoPi * __builtin_sqrtf(det_s));   // :74
        const float px = d.x * N, py = d.y * N, pz = d.z * N;
        const int c0x = (int)__builtin_floorf(px), c0y = (int)__builtin_floorf(py),
                  c0z = (int)__builtin_floorf(pz);
        const float h = 0.5f / N;
        const float s_in2 = (1.0f - h) * (1.0f - h), s_out2 = (1.0f + h) * (1.0f + h);
        const float inv16 = 1.0f / 65536.0f;
        const float lt0 = log_ieee(3000.0f), lt1 = log_ieee(22000.0f);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int cx = c0x + dx, cy = c0y + dy, cz = c0z + dz;
                    const U3 hh = pcg3d(U3{(uint32_t)cx + salt, (uint32_t)cy + (salt * 7u + 1u),
                                           (uint32_t)cz + (salt * 13u + 2u)});
                    const float ux = (float)(hh.x >> 16u) * inv16;
                    if (ux > prob) continue;                               // :85
                    const float u2x = (float)(hh.x & 0xffffu) * inv16;
                    const float u2y = (float)(hh.y >> 16u) * inv16;
                    const float u2z = (float)(hh.y & 0xffffu) * inv16;
                    const float uy = (float)(hh.z >> 16u) * inv16;
                    const float uz = (float)(hh.z & 0xffffu) * inv16;
                    const V3<float> q{((float)cx + u2x) / N, ((float)cy + u2y) / N,
                                      ((float)cz + u2z) / N};
                    const float ql2 = dot3(q, q);
                    if (ql2 < s_in2 || ql2 > s_out2) continue;              // :90
                    const float iq = rsqrt_rn(ql2);
                    const V3<float> delta{q.x * iq - d.x, q.y * iq - d.y, q.z * iq - d.z};
                    const float d1 = dot3(fp.e1, delta), d2 = dot3(fp.e2, delta);
                    const float sp = fp.r11 * d1 + fp.r12 * d2;
                    const float sq = fp.r22 * d2;
                    const float e = ((isaa * sp) * sp + ((2.0f * isab) * sp) * sq) +
                                    (isbb * sq) * sq;
                    if (!(e < 30.0f)) continue;                             // :96
                    const float F = flux * (1.0f / (uy > 2e-4f ? uy : 2e-4f));
                    const float T = exp_ieee(mixf(lt0, lt1, pow_ieee(uz, 1.7f)));
                    float bs[4];
                    blackbody(lut, T * g_shift, bs);
                    const float amp = (F * norm) * exp_ieee(-0.5f * e);
                    for (int k = 0; k < 3; ++k) sum[k] += mixf(1.0f, bs[k], 0.7f) * amp;
                }
    }
    bool bad = false;                                                        // :103-104
    for (int k = 0; k < 3; ++k) bad = bad || __builtin_isnan(sum[k]) || __builtin_isinf(sum[k]);
    for (int k = 0; k < 3; ++k) {
        const float s = bad ? 0.0f : (sum[k] < 64.0f ? sum[k] : 64.0f);
        o[k] = mixf(s, bb[k] * ka, w);
    }
}

// sampleSky (sky.glsl:108-119): g from the observer's radius rcam, the backdrop and two star
// layers. fp: makeFootprint(d, dx, dy, omega0).
__device__ __forceinline__ void sample_sky(const SkyConst& k, const Cube& cube,
                                           const float* __restrict__ lut, float rcam,
                                           const V3<float>& d, const V3<float>& dx,
                                           const V3<float>& dy, const fp::Footprint& fp,
                                           float omega0, float o[3]) {
    float t = 1.0f - 2.0f / rcam;
    t = t > 1e-4f ? t : 1e-4f;
    const float g = mixf(1.0f, rsqrt_rn(t), k.doppler);
    const float g4 = ((g * g) * g) * g;
    float col[3] = {0.0f, 0.0f, 0.0f};
    if (cube.texels) {
        cube_sample(cube, d, dx, dy, col);
        const float sb = k.sky_brightness * g4;
        for (int j = 0; j < 3; ++j) col[j] = col[j] * sb;
    }
    const float stb = k.star_brightness * g4;
    float s[3];
    star_layer(lut, d, fp, 200.0f, 0.40f * k.star_density, 0.0045f * stb, 0.0004f, omega0, 11u, g, s);
    for (int j = 0; j < 3; ++j) col[j] += s[j];
    star_layer(lut, d, fp, 48.0f, 0.25f * k.star_density, 0.06f * stb, 0.0011f, omega0, 23u, g, s);
    for (int j = 0; j < 3; ++j) col[j] += s[j];
    for (int j = 0; j < 3; ++j) o[j] = col[j];
}

}  // namespace sky
}  // namespace bhs

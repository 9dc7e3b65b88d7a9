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
//     IEEE-only functions of the shading. The cells are visited through bit masks (each test
//     once per cell, the stars' light summed in the shader's cell order), and the division by the
//     lattice size is a sequence equal to IEEE division on every dividend it meets, so the values
//     are those of the loop as written.
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
    // level 0 when magnifying, the last level at or past it, else floor(lam) and the next one
    int l0 = 0, nl = 1;
    float w = 0.0f;
    if (lam >= (float)q) {
        l0 = q;
    } else if (lam > 0.0f) {
        const float fl = __builtin_floorf(lam);
        l0 = (int)fl;
        nl = 2;
        w = lam - fl;
    }
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, b0 = 0.0f, b1 = 0.0f, b2 = 0.0f;
#pragma unroll 1
    for (int i = 0; i < nl; ++i) {   // one call site: the seamless paths are large
        float r[3];
        bilinear(c, l0 + i, f, s, t, r);
        if (i == 0) { a0 = r[0]; a1 = r[1]; a2 = r[2]; }
        else { b0 = r[0]; b1 = r[1]; b2 = r[2]; }
    }
    o[0] = nl == 2 ? mixf(a0, b0, w) : a0;
    o[1] = nl == 2 ? mixf(a1, b1, w) : a1;
    o[2] = nl == 2 ? mixf(a2, b2, w) : a2;
}

// The rounds of pcg3d (common.glsl:21-27) after its first, v = 1664525 v + 1013904223, which the
// caller has applied: the 27 cells of a star layer share the first round's three values per axis.
__device__ __forceinline__ U3 pcg3d_rest(U3 v) {
    v.x += v.y * v.z; v.y += v.z * v.x; v.z += v.x * v.y;
    v.x ^= v.x >> 16u; v.y ^= v.y >> 16u; v.z ^= v.z >> 16u;
    v.x += v.y * v.z; v.y += v.z * v.x; v.z += v.x * v.y;
    return v;
}
// Its x component alone (the density test reads only h.x).
__device__ __forceinline__ uint32_t pcg3d_rest_x(U3 v) {
    v.x += v.y * v.z; v.y += v.z * v.x; v.z += v.x * v.y;
    v.x ^= v.x >> 16u; v.y ^= v.y >> 16u; v.z ^= v.z >> 16u;
    return v.x + v.y * v.z;
}

__device__ __forceinline__ uint32_t sel3u(uint32_t a, uint32_t b, uint32_t c, int i) {
    return i == 0 ? a : (i == 1 ? b : c);
}

// a / NL for the dividends of a star layer, a = c + k / 65536 with |c| <= NL + 2 (the cells around
// a finite unit direction) and k a 16-bit integer: the product by RN(1 / NL) corrected once by its
// exact remainder. It equals the IEEE quotient for every such a, NL = 200 and NL = 48
// (tests/cpp/star_div_check.c checks all 33 M of them).
template <int NL>
__device__ __forceinline__ float div_lattice(float a) {
    static_assert(NL == 200 || NL == 48, "checked for the two layers' lattice sizes only");
    constexpr float N = (float)NL, rN = 1.0f / (float)NL;
    const float q0 = a * rN;
    return __builtin_fmaf(__builtin_fmaf(-q0, N, a), rN, q0);
}

// The screen-space exponent e of a star at unit direction sdir = q / |q| (sky.glsl:92-95) for the
// pixel at d with footprint fp.
__device__ __forceinline__ float star_exponent(const V3<float>& sdir, const V3<float>& d,
                                               const fp::Footprint& fp, float isaa, float isab,
                                               float isbb) {
    const V3<float> delta{sdir.x - d.x, sdir.y - d.y, sdir.z - d.z};
    const float d1 = dot3(fp.e1, delta), d2 = dot3(fp.e2, delta);
    const float sp = fp.r11 * d1 + fp.r12 * d2;
    const float sq = fp.r22 * d2;
    return ((isaa * sp) * sp + ((2.0f * isab) * sp) * sq) + (isbb * sq) * sq;
}

// One cell of a star layer (sky.glsl:80-96) from its hash hh = pcg3d(c + salts): the star's
// uniforms, position and screen-space exponent e; false when the shell test drops it (the density
// test is the caller's).
struct StarCell {
    float uy, uz, e;
};
template <int NL>
__device__ __forceinline__ bool star_cell(int cx, int cy, int cz, const U3& hh, const V3<float>& d,
                                          const fp::Footprint& fp, float s_in2, float s_out2,
                                          float isaa, float isab, float isbb, StarCell& o) {
    const float inv16 = 1.0f / 65536.0f;
    const float u2x = (float)(hh.x & 0xffffu) * inv16;
    const float u2y = (float)(hh.y >> 16u) * inv16;
    const float u2z = (float)(hh.y & 0xffffu) * inv16;
    o.uy = (float)(hh.z >> 16u) * inv16;
    o.uz = (float)(hh.z & 0xffffu) * inv16;
    const V3<float> q{div_lattice<NL>((float)cx + u2x), div_lattice<NL>((float)cy + u2y),
                      div_lattice<NL>((float)cz + u2z)};
    const float ql2 = dot3(q, q);
    if (ql2 < s_in2 || ql2 > s_out2) return false;                     // :90
    const float iq = rsqrt_rn_core(ql2);   // ql2 within the shell: a positive normal
    o.e = star_exponent(V3<float>{q.x * iq, q.y * iq, q.z * iq}, d, fp, isaa, isab, isbb);
    return true;
}

// Cells of a wave's shared table: the box of 3x3x3 neighbourhoods of its pixels (an 8x8 tile
// spans 36-100 cells on the benchmark view).
constexpr int kSkyBoxMax = 256;

__device__ __forceinline__ int wave_min_i(int v) {
    for (int o = 32; o > 0; o >>= 1) {
        const int u = __shfl_xor(v, o, 64);
        v = u < v ? u : v;
    }
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
    for (int o = 32; o > 0; o >>= 1) {
        const int u = __shfl_xor(v, o, 64);
        v = u > v ? u : v;
    }
    return v;
}

// pcg3d's first round for cell (cx, cy, cz) of a layer with this salt: c + (salt, 7 salt + 1,
// 13 salt + 2), then v = 1664525 v + 1013904223.
__device__ __forceinline__ U3 cell_first(int cx, int cy, int cz, uint32_t salt) {
    return U3{((uint32_t)cx + salt) * 1664525u + 1013904223u,
              ((uint32_t)cy + (salt * 7u + 1u)) * 1664525u + 1013904223u,
              ((uint32_t)cz + (salt * 13u + 2u)) * 1664525u + 1013904223u};
}

// Words of one cell in a wave's shared table: hash z (the star's brightness and colour
// uniforms), the star's unit direction (3 floats), and 1 when the cell holds a star within the
// shell (density and shell tests, which depend on the cell alone).
constexpr int kCellWords = 5;

// starLayer (sky.glsl:57-106) for escaped direction d with footprint fp; lut: blackbody table.
// Called by all 64 lanes of a wave; `on`: the lane's pixel takes the sky. The 27 cells of a lane's
// 3x3x3 neighbourhood are bit k = 9 (dz + 1) + 3 (dy + 1) + (dx + 1) of two masks, so that a lane
// works only on its own candidate cells and the wave's lanes stay in step: the density test for
// all 27 (the hash's x), then the shell and footprint tests over the cells that hold a star, then
// the star's light over the few that pass, in ascending k, the shader's loop order, so the sum is
// the same sequence of additions. When the box spanning the wave's pixels' neighbourhoods holds at
// most kSkyBoxMax cells, the wave first evaluates everything that depends on a cell alone once per
// cell (its hash, the density and shell tests, the star's direction) into LDS (s_cells,
// kCellWords words per cell), and a lane's candidates are then its cells that hold a star within
// the shell; else each lane hashes and tests its own cells, sharing pcg3d's first round along each
// axis.
template <int NL>
__device__ __forceinline__ void star_layer(const float* __restrict__ lut, const V3<float>& d,
                                           const fp::Footprint& fp, float prob, float flux,
                                           float sig_star, float omega0, uint32_t salt,
                                           float g_shift, bool on, uint32_t* __restrict__ s_cells,
                                           float o[3]) {
    constexpr float N = (float)NL;
    const int lane = (int)(threadIdx.x & 63u);
    const float det = fp.area * fp.area;                                 // Footprint.det
    const float cells = pow_ieee(det, 0.25f) * N;                        // :61
    const float w = smoothstepf(0.8f, 2.5f, cells);                      // :62
    float bb[4];
    blackbody(lut, 5600.0f * g_shift, bb);
    const float ka = ((((prob * N) * N) * 9.517f) * flux) * omega0;      // :63-64, MEAN_FLUX
    float sum[3] = {0.0f, 0.0f, 0.0f};
    const bool part = on && w < 1.0f;                                    // :66
    const float s2 = sig_star * sig_star, px2 = 0.45f;
    const float gi11 = fp.r11 * fp.r11 + fp.r12 * fp.r12;                // Footprint.Gi11
    const float gi12 = fp.r12 * fp.r22, gi22 = fp.r22 * fp.r22;
    const float rr = fp.r11 * fp.r22;
    const float det_gi = rr * rr;
    const float saa = px2 + s2 * gi11, sab = s2 * gi12, sbb = px2 + s2 * gi22;       // :70
    const float det_s = (px2 * px2 + (px2 * s2) * (gi11 + gi22)) + (s2 * s2) * det_gi;
    const float inv_det_s = 1.0f / det_s;
    const float isaa = sbb * inv_det_s, isab = -sab * inv_det_s, isbb = saa * inv_det_s;
    const float norm = fp.mu / (kTwoPi * __builtin_sqrtf(det_s));       // :74
    const int c0x = (int)__builtin_floorf(d.x * N), c0y = (int)__builtin_floorf(d.y * N),
              c0z = (int)__builtin_floorf(d.z * N);
    const float h = 0.5f / N;
    const float s_in2 = (1.0f - h) * (1.0f - h), s_out2 = (1.0f + h) * (1.0f + h);
    // the box of the participating lanes' neighbourhoods: [b0 - 1, b1 + 1] per axis (uniform)
    constexpr int kBig = 1 << 30;
    const int b0x = wave_min_i(part ? c0x : kBig), b0y = wave_min_i(part ? c0y : kBig),
              b0z = wave_min_i(part ? c0z : kBig);
    const int b1x = wave_max_i(part ? c0x : -kBig), b1y = wave_max_i(part ? c0y : -kBig),
              b1z = wave_max_i(part ? c0z : -kBig);
    if (b0x != kBig) {   // some lane takes part (uniform)
        const int nx = b1x - b0x + 3, ny = b1y - b0y + 3, nz = b1z - b0z + 3;
        const bool shared = nx <= kSkyBoxMax && ny <= kSkyBoxMax && nz <= kSkyBoxMax &&
                            nx * ny * nz <= kSkyBoxMax;                    // uniform
        const int nxy = nx * ny;
        // table slot of the lane's cell c0 - (1, 1, 1); cell k adds i + nx j + nxy l
        const int base = (c0x - b0x) + nx * ((c0y - b0y) + ny * (c0z - b0z));
        // per-lane form: pcg3d's first round for c0 - 1, c0, c0 + 1 along each axis
        uint32_t lx[3], ly[3], lz[3];
        const float pk = prob * 65536.0f;   // ux > prob  <=>  (h.x >> 16) > prob 2^16 (exact)
        uint32_t cand = 0;
        if (shared) {
            const int vol = nx * ny * nz;
            const float inv16 = 1.0f / 65536.0f;
            for (int sl = lane; sl < vol; sl += 64) {
                const int iz = sl / nxy, rem = sl - iz * nxy;
                const int iy = rem / nx, ix = rem - iy * nx;
                const int cx = b0x - 1 + ix, cy = b0y - 1 + iy, cz = b0z - 1 + iz;
                const U3 hh = pcg3d_rest(cell_first(cx, cy, cz, salt));
                const bool dense = !((float)(hh.x >> 16u) > pk);               // :85
                const V3<float> q{div_lattice<NL>((float)cx + (float)(hh.x & 0xffffu) * inv16),
                                  div_lattice<NL>((float)cy + (float)(hh.y >> 16u) * inv16),
                                  div_lattice<NL>((float)cz + (float)(hh.y & 0xffffu) * inv16)};
                const float ql2 = dot3(q, q);
                const bool shell = !(ql2 < s_in2 || ql2 > s_out2);            // :90
                const float iq = rsqrt_rn_core(shell ? ql2 : 1.0f);   // in the shell: normal
                uint32_t* cw = s_cells + kCellWords * sl;
                cw[0] = hh.z;
                cw[1] = __builtin_bit_cast(uint32_t, q.x * iq);
                cw[2] = __builtin_bit_cast(uint32_t, q.y * iq);
                cw[3] = __builtin_bit_cast(uint32_t, q.z * iq);
                cw[4] = dense && shell ? 1u : 0u;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (part) {
#pragma unroll
                for (int k = 0; k < 27; ++k)
                    cand |= s_cells[kCellWords * (base + (k % 3) + nx * ((k / 3) % 3) +
                                                  nxy * (k / 9)) + 4] << k;
            }
        } else if (part) {
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                const U3 f = cell_first(c0x + t - 1, c0y + t - 1, c0z + t - 1, salt);
                lx[t] = f.x;
                ly[t] = f.y;
                lz[t] = f.z;
            }
#pragma unroll
            for (int k = 0; k < 27; ++k) {                                 // :85
                const uint32_t hx = pcg3d_rest_x(U3{lx[k % 3], ly[(k / 3) % 3], lz[k / 9]});
                cand |= (float)(hx >> 16u) > pk ? 0u : (1u << k);
            }
        }
        // the star of candidate cell (i, j, l): uniforms and exponent; false when the shell drops it
        auto star_of = [&](int i, int j, int l, StarCell& sc) -> bool {
            if (shared) {
                const uint32_t* cw = s_cells + kCellWords * (base + i + nx * j + nxy * l);
                const uint32_t hz = cw[0];
                sc.uy = (float)(hz >> 16u) * (1.0f / 65536.0f);
                sc.uz = (float)(hz & 0xffffu) * (1.0f / 65536.0f);
                const V3<float> sdir{__builtin_bit_cast(float, cw[1]),
                                     __builtin_bit_cast(float, cw[2]),
                                     __builtin_bit_cast(float, cw[3])};
                sc.e = star_exponent(sdir, d, fp, isaa, isab, isbb);
                return true;   // the shared candidates passed the shell test
            }
            const U3 hh = pcg3d_rest(U3{sel3u(lx[0], lx[1], lx[2], i),
                                        sel3u(ly[0], ly[1], ly[2], j),
                                        sel3u(lz[0], lz[1], lz[2], l)});
            return star_cell<NL>(c0x + i - 1, c0y + j - 1, c0z + l - 1, hh, d, fp, s_in2, s_out2,
                                 isaa, isab, isbb, sc);
        };
        // shell (:90) and footprint (:96)
        uint32_t lit = 0;
        for (uint32_t m = cand; m != 0u; m &= m - 1u) {
            const int k = __builtin_ctz(m);
            const int l = (k >= 9) + (k >= 18), r = k - 9 * l;
            const int j = (r >= 3) + (r >= 6), i = r - 3 * j;
            StarCell sc;
            if (star_of(i, j, l, sc) && sc.e < 30.0f) lit |= 1u << k;
        }
        // the light of the stars that pass (:97-100), in ascending k
        const float lt0 = log_ieee(3000.0f), lt1 = log_ieee(22000.0f);
        for (uint32_t m = lit; m != 0u; m &= m - 1u) {
            const int k = __builtin_ctz(m);
            const int l = (k >= 9) + (k >= 18), r = k - 9 * l;
            const int j = (r >= 3) + (r >= 6), i = r - 3 * j;
            StarCell sc;
            star_of(i, j, l, sc);
            const float F = flux * (1.0f / (sc.uy > 2e-4f ? sc.uy : 2e-4f));
            const float T = exp_ieee(mixf(lt0, lt1, pow_ieee(sc.uz, 1.7f)));
            float bs[4];
            blackbody(lut, T * g_shift, bs);
            const float amp = (F * norm) * exp_ieee(-0.5f * sc.e);
            for (int c = 0; c < 3; ++c) sum[c] += mixf(1.0f, bs[c], 0.7f) * amp;
        }
        // every lane's reads of the table precede the next layer's writes
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    bool bad = false;                                                        // :103-104
    for (int k = 0; k < 3; ++k) bad = bad || __builtin_isnan(sum[k]) || __builtin_isinf(sum[k]);
    for (int k = 0; k < 3; ++k) {
        const float v = bad ? 0.0f : (sum[k] < 64.0f ? sum[k] : 64.0f);
        o[k] = mixf(v, bb[k] * ka, w);
    }
}

// sampleSky (sky.glsl:108-119): g from the observer's radius rcam, the backdrop and two star
// layers. fp: makeFootprint(d, dx, dy, omega0). Called by all 64 lanes of a wave (the star layers
// share a table in s_cells); o is meaningful where `on`.
__device__ __forceinline__ void sample_sky(const SkyConst& k, const Cube& cube,
                                           const float* __restrict__ lut, float rcam,
                                           const V3<float>& d, const V3<float>& dx,
                                           const V3<float>& dy, const fp::Footprint& fp,
                                           float omega0, bool on, uint32_t* __restrict__ s_cells,
                                           float o[3]) {
    float t = 1.0f - 2.0f / rcam;
    t = t > 1e-4f ? t : 1e-4f;
    const float g = mixf(1.0f, rsqrt_rn(t), k.doppler);
    const float g4 = ((g * g) * g) * g;
    float col[3] = {0.0f, 0.0f, 0.0f};
    if (cube.texels && on) {
        cube_sample(cube, d, dx, dy, col);
        const float sb = k.sky_brightness * g4;
        for (int j = 0; j < 3; ++j) col[j] = col[j] * sb;
    }
    const float stb = k.star_brightness * g4;
    float s[3];
    star_layer<200>(lut, d, fp, 0.40f * k.star_density, 0.0045f * stb, 0.0004f, omega0, 11u, g,
                    on, s_cells, s);
    for (int j = 0; j < 3; ++j) col[j] += s[j];
    star_layer<48>(lut, d, fp, 0.25f * k.star_density, 0.06f * stb, 0.0011f, omega0, 23u, g, on,
                   s_cells, s);
    for (int j = 0; j < 3; ++j) col[j] += s[j];
    for (int j = 0; j < 3; ++j) o[j] = col[j];
}

}  // namespace sky
}  // namespace bhs

// shade.hip — thin-disk shading and transparency over the ray planes of a render
// (shaders/blackhole.frag:209-216,234 with thinDiskSample, shaders/disk.glsl:113-130), the tables it
// reads (noise volume of shaders/noise_bake.comp, blackbody table of src/DiskTextures.cpp:61-93) and
// the running average of jittered frames (shaders/accumulate.frag:8-11).
//
// The march records the first two in-annulus crossings of every ray (r, phi, g). A ray's colour is
// the front-to-back composite of all its in-annulus crossings until its transmittance falls below
// 0.004 (the march then stops: no sky). So the pass composites the recorded crossings, and a ray
// with more than two crossings that is still transparent after two is marched again from its camera
// pixel in the precision it was rendered in (FP64 for band rays), shading every later crossing — the
// few photon-ring rays that circle the hole.
#include "shade.hpp"

#include <cmath>

#include "../../include/bhs_trace.h"

namespace bhs {

namespace {

constexpr int kBlock = 256;

using shade::Composite;
using shade::ShadeConst;

__device__ __forceinline__ uint32_t find_ray_cam(const MarchConst* __restrict__ t, uint32_t ncam,
                                                 int64_t i) {
    uint32_t lo = 0, hi = ncam;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (t[mid].ray_offset <= i) lo = mid; else hi = mid;
    }
    return lo;
}

// sqrt(max(1 - 2/rcam, 1e-4)) of opticalDepthScale (disk.glsl:70); rcam = |camPos| as the march's
// setup computes it.
__device__ __forceinline__ float cam_lapse(const MarchConst& c) {
    const float r = __builtin_sqrtf(__builtin_fmaf(c.pos[2], c.pos[2],
                                                   __builtin_fmaf(c.pos[1], c.pos[1], c.pos[0] * c.pos[0])));
    const float l2 = 1.0f - 2.0f / r;
    return __builtin_sqrtf(l2 > 1e-4f ? l2 : 1e-4f);
}

__global__ __launch_bounds__(kBlock) void k_noise_bake(uint16_t* __restrict__ vol, uint32_t seed) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= (uint32_t)(shade::kNoiseSize * shade::kNoiseSize * shade::kNoiseSize)) return;
    const int x = (int)(i & 127u), y = (int)((i >> 7) & 127u), z = (int)(i >> 14);
    vol[i] = shade::f2h(shade::noise_texel(x, y, z, seed));
}

// Composite of the recorded crossings of every ray; rays that need more crossings go to the list
// with their running colour in rgb and their transmittance in sky (the completion march resumes
// from there).
__global__ __launch_bounds__(kBlock) void k_shade_planes(const MarchConst* __restrict__ cams,
                                                         uint32_t ncam, int64_t n_rays, SoA rays,
                                                         ShadeOut out, ShadeConst sc,
                                                         const uint16_t* __restrict__ vol,
                                                         const float* __restrict__ lut,
                                                         ShadeList list) {
    // the cameras' first rays and lapses in LDS (ncam <= 64): a crossing ray finds its camera
    // there instead of in the table in global memory
    __shared__ int64_t s_off[kMaxBatchCams];
    __shared__ float s_lapse[kMaxBatchCams];
    for (uint32_t k = threadIdx.x; k < ncam; k += kBlock) {
        s_off[k] = cams[k].ray_offset;
        s_lapse[k] = cam_lapse(cams[k]);
    }
    __syncthreads();
    const int64_t n = rays.n;
    for (int64_t i0 = (int64_t)blockIdx.x * kBlock; i0 < n_rays; i0 += (int64_t)gridDim.x * kBlock) {
        const int64_t i = i0 + threadIdx.x;
        bool more = false;
        if (i < n_rays) {
            const uint32_t flags = rays.flags[i];
            const uint32_t nc = rays.n_cross[i];
            Composite cp{{0.0f, 0.0f, 0.0f}, 1.0f, false};
            if (nc > 0u) {
                uint32_t lo = 0, hi = ncam;                        // the ray's camera
                while (hi - lo > 1u) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (s_off[mid] <= i) lo = mid; else hi = mid;
                }
                const float lapse = s_lapse[lo];
                const uint32_t m = nc < 2u ? nc : 2u;
                for (uint32_t k = 0; k < m && !cp.opaque; ++k) {
                    float s[3], a;
                    shade::thin_disk_sample(sc, vol, lut, lapse, rays.disk[i + (int64_t)k * n],
                                            rays.disk[i + (int64_t)(2 + k) * n],
                                            rays.disk[i + (int64_t)(4 + k) * n], s, a);
                    cp.add(s, a);
                }
                more = nc > 2u && !cp.opaque;
            }
            const bool esc = (flags & BHS_FLAG_ESCAPED) != 0u;
            out.rgb[i] = cp.rgb[0];
            out.rgb[i + out.n] = cp.rgb[1];
            out.rgb[i + 2 * out.n] = cp.rgb[2];
            out.sky[i] = more ? cp.T : ((esc && !cp.opaque) ? cp.T : 0.0f);
        }
        const uint64_t m = __ballot(more);
        if (m == 0) continue;
        uint32_t base = 0;
        if ((threadIdx.x & 63u) == 0) base = atomicAdd(list.count, (uint32_t)__popcll(m));
        base = (uint32_t)__shfl((int)base, 0);
        if (more) {
            const uint32_t k = base + __builtin_amdgcn_mbcnt_hi(
                                          (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            if (k < list.cap) list.ids[k] = (uint32_t)i;
        }
    }
}

// Sink of the completion march: crossings 0 and 1 were composited from the records; every later
// in-annulus crossing is shaded here.
template <typename R>
struct ShadeSink {
    const ShadeConst& sc;
    const uint16_t* vol;
    const float* lut;
    float lapse;
    Composite& cp;
    __device__ __forceinline__ void escape(float, float, float) {}
    __device__ __forceinline__ void cross(uint32_t slot, float r, float phi, float g,
                                          const RootState<R>&) {
        if (slot < 2u || cp.opaque) return;
        float s[3], a;
        shade::thin_disk_sample(sc, vol, lut, lapse, r, phi, g, s, a);
        cp.add(s, a);
    }
};

template <typename R>
__device__ __forceinline__ bool shade_march(const MarchConst& c, const V3<float>& Pf,
                                            const V3<float>& df, const ShadeConst& sc,
                                            const uint16_t* vol, const float* lut, float lapse,
                                            Composite& cp) {
    const V3<R> P{(R)Pf.x, (R)Pf.y, (R)Pf.z};
    const V3<R> d{(R)df.x, (R)df.y, (R)df.z};
    const Setup<R> s = lawS_setup(P, d);
    const MarchK<R> k = march_consts<R>(c);
    MarchState<R> st;
    march_init(st, P, s);
    ShadeSink<R> sink{sc, vol, lut, lapse, cp};
    while (!march_step<R, ShadeSink<R>, true>(st, s, k, sink)) {
        if (cp.opaque) return false;                                  // blackhole.frag:234
    }
    return !cp.opaque && st.status == 1;
}

__global__ __launch_bounds__(kBlock) void k_shade_complete(const MarchConst* __restrict__ cams,
                                                           uint32_t ncam, SoA rays, ShadeOut out,
                                                           ShadeConst sc,
                                                           const uint16_t* __restrict__ vol,
                                                           const float* __restrict__ lut,
                                                           ShadeList list) {
    const uint32_t cnt = *list.count < list.cap ? *list.count : list.cap;
    for (uint32_t q = blockIdx.x * kBlock + threadIdx.x; q < cnt; q += gridDim.x * kBlock) {
        const int64_t i = list.ids[q];
        const MarchConst& c = cams[find_ray_cam(cams, ncam, i)];
        const int64_t pix = i - c.ray_offset;
        const V3<float> P{c.pos[0], c.pos[1], c.pos[2]};
        const V3<float> d = ray_dir(c, pix);
        Composite cp{{out.rgb[i], out.rgb[i + out.n], out.rgb[i + 2 * out.n]}, out.sky[i], false};
        const float lapse = cam_lapse(c);
        const bool sky = (rays.flags[i] & BHS_FLAG_BAND64) != 0u
                             ? shade_march<double>(c, P, d, sc, vol, lut, lapse, cp)
                             : shade_march<float>(c, P, d, sc, vol, lut, lapse, cp);
        out.rgb[i] = cp.rgb[0];
        out.rgb[i + out.n] = cp.rgb[1];
        out.rgb[i + 2 * out.n] = cp.rgb[2];
        out.sky[i] = sky ? cp.T : 0.0f;
    }
}

__global__ __launch_bounds__(kBlock) void k_accumulate(const float* __restrict__ cur,
                                                       float* __restrict__ acc, int64_t n, float w) {
    const bool copy = w >= 1.0f;
    const int64_t n4 = n >> 2;
    const float4* c4 = reinterpret_cast<const float4*>(cur);
    float4* a4 = reinterpret_cast<float4*>(acc);
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n4;
         i += (int64_t)gridDim.x * kBlock) {
        const float4 c = c4[i];
        if (copy) {
            a4[i] = c;
        } else {
            const float4 a = a4[i];
            a4[i] = float4{shade::mixf(a.x, c.x, w), shade::mixf(a.y, c.y, w),
                           shade::mixf(a.z, c.z, w), shade::mixf(a.w, c.w, w)};
        }
    }
    for (int64_t i = 4 * n4 + (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        acc[i] = copy ? cur[i] : shade::mixf(acc[i], cur[i], w);
}

uint32_t grid_for(int64_t n, int num_cus, uint32_t per_cu) {
    const int64_t blocks = (n + kBlock - 1) / kBlock;
    const int64_t cap = (int64_t)(num_cus > 0 ? num_cus : 256) * per_cu;
    return (uint32_t)(blocks < 1 ? 1 : (blocks < cap ? blocks : cap));
}

}  // namespace

hipError_t launch_noise_bake(uint16_t* vol, uint32_t seed, hipStream_t s) {
    const uint32_t n = (uint32_t)(shade::kNoiseSize * shade::kNoiseSize * shade::kNoiseSize);
    hipLaunchKernelGGL(k_noise_bake, dim3(n / kBlock), dim3(kBlock), 0, s, vol, seed);
    return hipGetLastError();
}

hipError_t launch_shade(const MarchConst* table, uint32_t ncam, int64_t n_rays, const SoA& rays,
                        const ShadeOut& out, const shade::ShadeConst& sc, const uint16_t* vol,
                        const float* lut, const ShadeList& list, int num_cus, hipStream_t s,
                        hipEvent_t ev_begin, hipEvent_t ev_end) {
    hipError_t e;
    if (ev_begin && (e = hipEventRecord(ev_begin, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_shade_planes, dim3(grid_for(n_rays, num_cus, 16)), dim3(kBlock), 0, s,
                       table, ncam, n_rays, rays, out, sc, vol, lut, list);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ev_end && (e = hipEventRecord(ev_end, s)) != hipSuccess) return e;
    // The listed rays (a few per frame) march hundreds of steps each: a grid of whole waves over
    // the list's capacity, each lane one ray; lanes past the device count exit at once.
    const int64_t lanes = list.cap < 65536u ? (int64_t)list.cap : 65536;
    hipLaunchKernelGGL(k_shade_complete, dim3(grid_for(lanes, num_cus, 4)), dim3(kBlock), 0, s,
                       table, ncam, rays, out, sc, vol, lut, list);
    return hipGetLastError();
}

hipError_t launch_accumulate(const float* cur, float* acc, int64_t n, float weight, hipStream_t s) {
    hipLaunchKernelGGL(k_accumulate, dim3(grid_for((n + 3) / 4, 256, 16)), dim3(kBlock), 0, s, cur,
                       acc, n, weight);
    return hipGetLastError();
}

// ------------------------------------------------------------------ blackbody table (host)
namespace {
// Wyman, Sloan & Shirley (2013) multi-lobe Gaussian fit of the CIE 1931 2-degree observer
// (piecewise widths: s1 below the centre, s2 above).
double lobe(double lam, double mu, double s1, double s2) {
    const double t = (lam - mu) / (lam < mu ? s1 : s2);
    return std::exp(-0.5 * t * t);
}

// Luminance-weighted tristimulus X, Y, Z of a blackbody at T over 380..780 nm in 1 nm steps.
void tristimulus(double T, double& X, double& Y, double& Z) {
    const double h = 6.62607015e-34, c = 2.99792458e8, kb = 1.380649e-23;
    X = Y = Z = 0.0;
    for (int nm = 380; nm <= 780; ++nm) {
        const double lam = (double)nm;
        const double x = (1.056 * lobe(lam, 599.8, 37.9, 31.0) + 0.362 * lobe(lam, 442.0, 16.0, 26.7)) -
                         0.065 * lobe(lam, 501.1, 20.4, 26.2);
        const double y = 0.821 * lobe(lam, 568.8, 46.9, 40.5) + 0.286 * lobe(lam, 530.9, 16.3, 31.1);
        const double z = 1.217 * lobe(lam, 437.0, 11.8, 36.0) + 0.681 * lobe(lam, 459.0, 26.0, 13.8);
        const double l = lam * 1e-9;
        const double l5 = (((l * l) * l) * l) * l;
        const double b = ((2.0 * h) * c * c / l5) / (std::exp((h * c) / ((l * kb) * T)) - 1.0);
        X += b * x;
        Y += b * y;
        Z += b * z;
    }
}
}  // namespace

void bake_blackbody_lut(float* lut) {
    double Xr, Yr, Zr;
    tristimulus(6500.0, Xr, Yr, Zr);
    const double lmin = std::log(1000.0), lmax = std::log(40000.0);
    for (int i = 0; i < shade::kLutSize; ++i) {
        const double t = (double)i / (double)(shade::kLutSize - 1);
        const double T = std::exp(lmin + (lmax - lmin) * t);
        double X, Y, Z;
        tristimulus(T, X, Y, Z);
        double r = (3.2406 * X - 1.5372 * Y) - 0.4986 * Z;    // XYZ -> linear sRGB (D65)
        double g = (-0.9689 * X + 1.8758 * Y) + 0.0415 * Z;
        double b = (0.0557 * X - 0.2040 * Y) + 1.0570 * Z;
        r = r > 0.0 ? r : 0.0;
        g = g > 0.0 ? g : 0.0;
        b = b > 0.0 ? b : 0.0;
        const double lum = (0.2126 * r + 0.7152 * g) + 0.0722 * b;
        const double inv = lum > 0.0 ? 1.0 / lum : 0.0;
        lut[4 * i + 0] = (float)(r * inv);
        lut[4 * i + 1] = (float)(g * inv);
        lut[4 * i + 2] = (float)(b * inv);
        lut[4 * i + 3] = (float)(Y / Yr);
    }
}

}  // namespace bhs

// volume.hip — the volumetric disk render (shaders/blackhole.frag:142-253 with DISK_VOLUMETRIC and
// shaders/disk.glsl:81-110): one lane per ray marches law S with the volumetric slab rule and
// composites the disk's emission and absorption along the Hermite sub-samples of every step inside
// the slab, until the ray escapes, falls in, runs out of steps or turns opaque. A scan lists the FP64
// band rays (|b - b_crit| < band), whose FP64 march runs on a side stream beside the FP32 march of
// all the others.
#include "volume.hpp"

#include "../../include/bhs_trace.h"

namespace bhs {

namespace {

constexpr int kBlock = 256;

using shade::Composite;
using shade::VolConst;

struct VolCams {
    const MarchConst* cams;
    const uint32_t* item_base;
    uint32_t ncam;
    uint32_t n_items;
    __device__ __forceinline__ uint32_t find(uint32_t q) const {
        uint32_t lo = 0, hi = ncam;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (item_base[mid] <= q) lo = mid; else hi = mid;
        }
        return lo;
    }
};

// Ray index of work item q of a camera (8x8-tile-major order; -1 for tile padding).
__device__ __forceinline__ int64_t vol_item_pixel(const MarchConst& c, uint32_t q) {
    const uint32_t tiles_x = ((uint32_t)c.width + 7u) >> 3;
    const uint32_t t = q >> 6, l = q & 63u;
    const uint32_t col = (t % tiles_x) * 8u + (l & 7u), row = (t / tiles_x) * 8u + (l >> 3);
    return (col < (uint32_t)c.width && row < (uint32_t)c.height) ? (int64_t)row * c.width + col : -1;
}

__device__ __forceinline__ VolConst vol_const(const shade::ShadeConst& sh, const MarchConst& c) {
    VolConst v;
    v.sh = sh;
    v.sh.inner = c.disk_inner;
    v.sh.outer = c.disk_outer;
    v.thickness = c.disk_thickness;
    const float r = __builtin_sqrtf(__builtin_fmaf(c.pos[2], c.pos[2],
                                                   __builtin_fmaf(c.pos[1], c.pos[1], c.pos[0] * c.pos[0])));
    const float l2 = 1.0f - 2.0f / r;
    v.lapse = __builtin_sqrtf(l2 > 1e-4f ? l2 : 1e-4f);
    return v;
}

__device__ __forceinline__ uint32_t wsum(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Marches one ray in precision R and writes its records; returns its step count.
template <typename R>
__device__ __forceinline__ uint32_t vol_ray(const MarchConst& c, const shade::ShadeConst& sh,
                                            int64_t g, int64_t pix, const SoA& out,
                                            const ShadeOut& so, const uint16_t* __restrict__ vol,
                                            const float* __restrict__ lut, uint32_t extra_flags,
                                            uint32_t& escaped) {
    const V3<float> Pf{c.pos[0], c.pos[1], c.pos[2]};
    const V3<float> df = ray_dir(c, pix);
    const V3<R> P{(R)Pf.x, (R)Pf.y, (R)Pf.z};
    const V3<R> d{(R)df.x, (R)df.y, (R)df.z};
    const Setup<R> s = lawS_setup(P, d);
    const MarchK<R> k = march_consts<R>(c);
    const VolConst vc = vol_const(sh, c);
    MarchState<R> st;
    march_init(st, P, s);
    Composite cp{{0.0f, 0.0f, 0.0f}, 1.0f, false};
    float esc[3] = {0.0f, 0.0f, 0.0f};
    const float b_axis = (float)s.b_axis;
    while (!shade::vol_step<R>(st, s, k, vc, vol, lut, b_axis, cp, esc)) {
    }
    uint32_t flags = extra_flags, cls;
    if (st.status == 0) {
        flags |= BHS_FLAG_MAXSTEPS;
        cls = BHS_CLASS_HORIZON;
    } else if (st.status == 1) {
        flags |= BHS_FLAG_ESCAPED;
        cls = BHS_CLASS_ESCAPE;
    } else if (st.status == 3) {
        cls = BHS_CLASS_DISK;
    } else {
        cls = BHS_CLASS_HORIZON;
    }
    escaped = st.status == 1 ? 1u : 0u;
    const int64_t n = out.n;
    if (out.cls) out.cls[g] = (uint8_t)cls;
    if (out.flags) out.flags[g] = (uint8_t)flags;
    if (out.n_cross) out.n_cross[g] = 0;
    if (out.steps) out.steps[g] = (uint16_t)(st.i > 65535 ? 65535 : st.i);
    if (out.esc) {
        out.esc[g] = esc[0];
        out.esc[g + n] = esc[1];
        out.esc[g + 2 * n] = esc[2];
    }
    if (out.disk)
        for (int j = 0; j < 6; ++j) out.disk[g + j * n] = 0.0f;
    if (out.b) out.b[g] = (float)s.b;
    so.rgb[g] = cp.rgb[0];
    so.rgb[g + so.n] = cp.rgb[1];
    so.rgb[g + 2 * so.n] = cp.rgb[2];
    so.sky[g] = st.status == 1 ? cp.T : 0.0f;
    return (uint32_t)st.i;
}

__device__ __forceinline__ bool vol_in_band(const MarchConst& c, int64_t pix) {
    if (c.precision != BHS_PREC_F32_BAND64) return false;
    const V3<float> P{c.pos[0], c.pos[1], c.pos[2]};
    return __builtin_fabsf(lawS_setup(P, ray_dir(c, pix)).b - kBcritF) < c.band_abs;
}

// Lists the work items of the FP64 band (|b - b_crit| < band): one atomic per wave.
__global__ __launch_bounds__(kBlock) void k_vol_scan(VolCams cams, Scratch sc) {
    for (uint32_t q0 = blockIdx.x * kBlock; q0 < cams.n_items; q0 += gridDim.x * kBlock) {
        const uint32_t q = q0 + threadIdx.x;
        bool band = false;
        if (q < cams.n_items) {
            const uint32_t ci = cams.find(q);
            const MarchConst& c = cams.cams[ci];
            const int64_t pix = vol_item_pixel(c, q - cams.item_base[ci]);
            band = pix >= 0 && vol_in_band(c, pix);
        }
        const uint64_t m = __ballot(band);
        if (m == 0) continue;
        uint32_t base = 0;
        if ((threadIdx.x & 63u) == 0) base = atomicAdd(sc.band_count, (uint32_t)__popcll(m));
        base = (uint32_t)__shfl((int)base, 0);
        if (band) {
            const uint32_t j = base + __builtin_amdgcn_mbcnt_hi(
                                          (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            if (j < sc.band_cap) sc.band_ids[j] = q;
        }
    }
}

// Adds the block's rays, steps and escapes (band: also to the band count) to the launch counters:
// wave sums in LDS, one atomic per counter and block from its first thread. Every thread of the
// block calls it (the loop bounds are block-uniform); the barrier after the atomics keeps the LDS
// slots of the next call from being overwritten before they are read.
__device__ __forceinline__ void block_stats(Stats* st, uint32_t ray, uint32_t steps, uint32_t esc,
                                            bool band, uint32_t (*s_st)[3]) {
    ray = wsum(ray);
    steps = wsum(steps);
    esc = wsum(esc);
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0u) {
        s_st[w][0] = ray;
        s_st[w][1] = steps;
        s_st[w][2] = esc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t r = 0, k = 0, e = 0;
        for (uint32_t j = 0; j < kBlock / 64; ++j) {
            r += s_st[j][0];
            k += s_st[j][1];
            e += s_st[j][2];
        }
        if (r) {
            atomicAdd(&st->rays, (unsigned long long)r);
            atomicAdd(&st->steps, (unsigned long long)k);
            if (band) atomicAdd(&st->band64, (unsigned long long)r);
            if (e) atomicAdd(&st->escapes, (unsigned long long)e);
        }
    }
    __syncthreads();
}

// FP32 pass over every work item outside the FP64 band.
__global__ __launch_bounds__(kBlock) void k_vol_f32(VolCams cams, shade::ShadeConst sh, SoA out,
                                                    ShadeOut so, const uint16_t* __restrict__ vol,
                                                    const float* __restrict__ lut, Scratch sc) {
    __shared__ uint32_t s_st[kBlock / 64][3];
    for (uint32_t q0 = blockIdx.x * kBlock; q0 < cams.n_items; q0 += gridDim.x * kBlock) {
        const uint32_t q = q0 + threadIdx.x;
        uint32_t ray = 0, steps = 0, esc = 0;
        if (q < cams.n_items) {
            const uint32_t ci = cams.find(q);
            const MarchConst& c = cams.cams[ci];
            const int64_t pix = vol_item_pixel(c, q - cams.item_base[ci]);
            if (pix >= 0 && !vol_in_band(c, pix)) {
                ray = 1;
                steps = vol_ray<float>(c, sh, c.ray_offset + pix, pix, out, so, vol, lut, 0u, esc);
            }
        }
        block_stats(sc.stats, ray, steps, esc, false, s_st);
    }
}

// FP64 march of the band list (or of every item, all = true, for BHS_PREC_F64 cameras).
__global__ __launch_bounds__(kBlock) void k_vol_f64(VolCams cams, shade::ShadeConst sh, SoA out,
                                                    ShadeOut so, const uint16_t* __restrict__ vol,
                                                    const float* __restrict__ lut, Scratch sc,
                                                    bool all) {
    __shared__ uint32_t s_st[kBlock / 64][3];
    const uint32_t n = all ? cams.n_items
                           : (*sc.band_count < sc.band_cap ? *sc.band_count : sc.band_cap);
    for (uint32_t i0 = blockIdx.x * kBlock; i0 < n; i0 += gridDim.x * kBlock) {
        const uint32_t i = i0 + threadIdx.x;
        uint32_t ray = 0, steps = 0, esc = 0;
        if (i < n) {
            const uint32_t q = all ? i : sc.band_ids[i];
            const uint32_t ci = cams.find(q);
            const MarchConst& c = cams.cams[ci];
            const int64_t pix = vol_item_pixel(c, q - cams.item_base[ci]);
            if (pix >= 0) {
                ray = 1;
                steps = vol_ray<double>(c, sh, c.ray_offset + pix, pix, out, so, vol, lut,
                                        BHS_FLAG_BAND64, esc);
            }
        }
        block_stats(sc.stats, ray, steps, esc, true, s_st);
    }
}

}  // namespace

hipError_t launch_volumetric(const MarchConst* table, const uint32_t* item_base, uint32_t ncam,
                             uint32_t n_items, bool all_f64, bool any_band,
                             const shade::ShadeConst& sh, const SoA& out, const ShadeOut& so,
                             const uint16_t* vol, const float* lut, const Scratch& sc,
                             int num_cus, hipStream_t s, hipStream_t aux, hipEvent_t ev_fork,
                             hipEvent_t ev_join, hipEvent_t ev_begin, hipEvent_t ev_end) {
    const VolCams cams{table, item_base, ncam, n_items};
    const uint32_t blocks = (n_items + kBlock - 1) / kBlock;
    const uint32_t grid = blocks ? blocks : 1u;
    hipError_t e;
    if (all_f64) {
        if (ev_begin && (e = hipEventRecord(ev_begin, s)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_vol_f64, dim3(grid), dim3(kBlock), 0, s, cams, sh, out, so, vol, lut,
                           sc, true);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (ev_end && (e = hipEventRecord(ev_end, s)) != hipSuccess) return e;
        return hipSuccess;
    }
    if (any_band) {
        // The band rays (few, long FP64 chains) march on the side stream beside the FP32 pass.
        const uint32_t cap = (uint32_t)(num_cus > 0 ? num_cus : 256) * 16u;
        hipLaunchKernelGGL(k_vol_scan, dim3(grid < cap ? grid : cap), dim3(kBlock), 0, s, cams, sc);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = hipEventRecord(ev_fork, s)) != hipSuccess ||
            (e = hipStreamWaitEvent(aux, ev_fork, 0)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(k_vol_f64, dim3((uint32_t)(num_cus > 0 ? num_cus : 256) * 4u),
                           dim3(kBlock), 0, aux, cams, sh, out, so, vol, lut, sc, false);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = hipEventRecord(ev_join, aux)) != hipSuccess) return e;
    }
    if (ev_begin && (e = hipEventRecord(ev_begin, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_vol_f32, dim3(grid), dim3(kBlock), 0, s, cams, sh, out, so, vol, lut, sc);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ev_end && (e = hipEventRecord(ev_end, s)) != hipSuccess) return e;
    if (any_band && (e = hipStreamWaitEvent(s, ev_join, 0)) != hipSuccess) return e;
    return hipSuccess;
}

}  // namespace bhs

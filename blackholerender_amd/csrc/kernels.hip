// kernels.hip — law-S render/march kernels (FP32 with FP64 band, FP64), the persistent
// work-queue kernel for one camera or a batch of cameras, FP64 crossing-root refinement, and the
// law-T (bhs::physics::trace) kernel.
#include "kernels.hpp"

#include <type_traits>

#include "../../include/bhs_trace.h"

namespace bhs {

namespace {

constexpr int kTile = 16;       // 16x16 pixel tile per 256-thread block, 8x8 per wave
constexpr int kBlock = 256;

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ void store_ray(const SoA& o, int64_t i, const RayOut& r) {
    const int64_t n = o.n;
    if (o.cls) o.cls[i] = (uint8_t)r.cls;
    if (o.flags) o.flags[i] = (uint8_t)r.flags;
    if (o.n_cross) o.n_cross[i] = (uint8_t)r.n_cross;
    if (o.steps) o.steps[i] = (uint16_t)(r.steps > 65535 ? 65535 : r.steps);
    if (o.esc) {
        o.esc[i] = r.esc[0];
        o.esc[i + n] = r.esc[1];
        o.esc[i + 2 * n] = r.esc[2];
    }
    if (o.disk) {
        o.disk[i] = r.dr[0];
        o.disk[i + n] = r.dr[1];
        o.disk[i + 2 * n] = r.dphi[0];
        o.disk[i + 3 * n] = r.dphi[1];
        o.disk[i + 4 * n] = r.dg[0];
        o.disk[i + 5 * n] = r.dg[1];
    }
    if (o.b) o.b[i] = r.b;
}

// Adds one lane's contribution to the launch counters: one atomic per wave and counter. Every lane
// of the wave must call it.
__device__ __forceinline__ void add_stats(Stats* st, uint32_t ray, uint32_t steps, uint32_t cross,
                                          uint32_t esc, uint32_t band) {
    ray = wave_sum(ray);
    steps = wave_sum(steps);
    cross = wave_sum(cross);
    esc = wave_sum(esc);
    band = wave_sum(band);
    if ((threadIdx.x & 63) == 0) {
        if (ray) atomicAdd(&st->rays, (unsigned long long)ray);
        if (steps) atomicAdd(&st->steps, (unsigned long long)steps);
        if (cross) atomicAdd(&st->crossings, (unsigned long long)cross);
        if (esc) atomicAdd(&st->escapes, (unsigned long long)esc);
        if (band) atomicAdd(&st->band64, (unsigned long long)band);
    }
}

// Pixel of thread t in block (bx, by): 16x16 tile, each wave an 8x8 sub-tile.
__device__ __forceinline__ void tile_pixel(int& col, int& row) {
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    col = blockIdx.x * kTile + (w & 1) * 8 + (l & 7);
    row = blockIdx.y * kTile + (w >> 1) * 8 + (l >> 3);
}

__device__ __forceinline__ bool in_band(const MarchConst& c, float b) {
    return __builtin_fabsf(b - kBcritF) < c.band_abs;
}

// Work item q of one camera: ray index in work-item tile order over its image (item_col_row).
// Items of the padding beyond the image map to -1.
__device__ __forceinline__ int64_t item_pixel(const MarchConst& c, uint32_t q) {
    uint32_t col, row;
    item_col_row(item_tiles_x((uint32_t)c.width), q, col, row);
    return (col < (uint32_t)c.width && row < (uint32_t)c.height) ? (int64_t)row * c.width + col : -1;
}

// Largest k < ncam with base[k] <= q (base[0] = 0, base[ncam] = total items).
__device__ __forceinline__ uint32_t find_cam(const uint32_t* base, uint32_t ncam, uint32_t q) {
    uint32_t lo = 0, hi = ncam;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (base[mid] <= q) lo = mid; else hi = mid;
    }
    return lo;
}

// Camera sources of the item-indexed kernels: one camera passed by value, or a device table with
// the prefix sums of the per-camera item counts.
struct OneCam {
    MarchConst c;
    uint32_t n_items;
    __device__ __forceinline__ uint32_t find(uint32_t) const { return 0u; }
    __device__ __forceinline__ const MarchConst& cam(uint32_t) const { return c; }
    __device__ __forceinline__ uint32_t base(uint32_t) const { return 0u; }
};
struct CamTable {
    const MarchConst* cams;
    const uint32_t* item_base;
    uint32_t ncam;
    uint32_t n_items;
    __device__ __forceinline__ uint32_t find(uint32_t q) const { return find_cam(item_base, ncam, q); }
    __device__ __forceinline__ const MarchConst& cam(uint32_t k) const { return cams[k]; }
    __device__ __forceinline__ uint32_t base(uint32_t k) const { return item_base[k]; }
};

// Sink writing a ray's records straight to the SoA output (and the crossing states to the
// root-refinement scratch).
struct GlobalSink {
    const SoA& o;
    int64_t i;
    RootState<float>* rs;
    __device__ __forceinline__ void escape(float ex, float ey, float ez) {
        if (o.esc) {
            o.esc[i] = ex;
            o.esc[i + o.n] = ey;
            o.esc[i + 2 * o.n] = ez;
        }
    }
    __device__ __forceinline__ void cross(uint32_t slot, float r, float phi, float g,
                                          const RootState<float>& q) {
        if (o.disk) {
            const int64_t j = i + (int64_t)slot * o.n;
            o.disk[j] = r;
            o.disk[j + 2 * o.n] = phi;
            o.disk[j + 4 * o.n] = g;
        }
        if (rs) rs[slot] = q;
    }
};

// Sink of the FP64 band march: records straight to the SoA output; with ROOTS the two crossing
// states stay in the lane's registers (slots selected by constant indices) until the ray ends.
template <bool ROOTS>
struct BandSink {
    const SoA& o;
    int64_t i;
    RootState<double>& rs0;
    RootState<double>& rs1;
    __device__ __forceinline__ void escape(float ex, float ey, float ez) {
        if (o.esc) {
            o.esc[i] = ex;
            o.esc[i + o.n] = ey;
            o.esc[i + 2 * o.n] = ez;
        }
    }
    __device__ __forceinline__ void cross(uint32_t slot, float r, float phi, float g,
                                          const RootState<double>& q) {
        if (o.disk) {
            const int64_t j = i + (int64_t)slot * o.n;
            o.disk[j] = r;
            o.disk[j + 2 * o.n] = phi;
            o.disk[j + 4 * o.n] = g;
        }
        if (ROOTS) {
            if (slot == 0) rs0 = q; else rs1 = q;
        }
    }
};

// Scalar outputs of a finished ray, and zeros for the records it did not produce.
template <typename R>
__device__ __forceinline__ void store_finish(const SoA& o, int64_t i, const MarchState<R>& st,
                                             float b, uint32_t extra_flags = 0u) {
    uint32_t cls, flags;
    march_class(st, cls, flags);
    flags |= extra_flags;
    const int64_t n = o.n;
    if (o.cls) o.cls[i] = (uint8_t)cls;
    if (o.flags) o.flags[i] = (uint8_t)flags;
    if (o.n_cross) o.n_cross[i] = (uint8_t)st.n_cross;
    if (o.steps) o.steps[i] = (uint16_t)(st.i > 65535 ? 65535 : st.i);
    if (o.b) o.b[i] = b;
    if (o.esc && st.status != 1) {
        o.esc[i] = 0.0f;
        o.esc[i + n] = 0.0f;
        o.esc[i + 2 * n] = 0.0f;
    }
    if (o.disk) {
        for (uint32_t k = st.n_cross; k < 2; ++k) {
            o.disk[i + k * n] = 0.0f;
            o.disk[i + (2 + k) * n] = 0.0f;
            o.disk[i + (4 + k) * n] = 0.0f;
        }
    }
}

// FP64 law-S march of one ray, refined crossings included (as the oracle's double march).
__device__ __forceinline__ RayOut march_f64(const MarchConst& c, const V3<float>& Pf,
                                            const V3<float>& df) {
    const V3<double> P{(double)Pf.x, (double)Pf.y, (double)Pf.z};
    const V3<double> d{(double)df.x, (double)df.y, (double)df.z};
    const Setup<double> s = lawS_setup(P, d);
    RayOut o;
    RootState<double> rs[2];
    lawS_march(P, s, c, o, c.root64 ? rs : (RootState<double>*)nullptr);
    if (c.root64 && o.n_cross > 0) {
        const uint32_t m = o.n_cross < 2 ? o.n_cross : 2;
        for (uint32_t k = 0; k < m; ++k) {
            double hx, hz;
            if (!root64(rs[k].x0, rs[k].v0, rs[k].dt, s.c3, hx, hz)) continue;
            const double rh = __builtin_sqrt(__builtin_fma(hz, hz, hx * hx));
            o.dr[k] = (float)rh;
            o.dphi[k] = atan2_ieee((float)hz, (float)hx);
            o.dg[k] = (float)redshift(rh, s.b_axis, s.lapse_g);
        }
        o.flags |= BHS_FLAG_ROOT64;
    }
    o.flags |= BHS_FLAG_BAND64;
    return o;
}

// Margin around the band's edges (absolute, in b) inside which band_member takes the exact setup.
constexpr float kBandScanMargin = 1e-3f;

// Whether the ray of pixel (col, row) lies in the FP64 band, |b - b_crit| < band_abs with b from
// the march's own float setup (lawS_setup of ray_dir), decided from an estimate first: the view
// direction and b = r sin(alpha) / sqrt(1 - 2/r) from hardware reciprocal and reciprocal square
// root estimates (a few ulps each, so b within ~1e-5 relative of the setup's) instead of the
// setup's correctly rounded divisions and square roots. The estimate decides wherever it lies more
// than kBandScanMargin from either edge of the band (the exact b is then on the same side: the two
// differ by far less than the margin); rays within it, cameras closer than r = 3 and non-finite
// estimates take the exact setup. Same answer as the exact test for every ray.
__device__ __forceinline__ bool band_member(const MarchConst& c, int64_t pix, uint32_t col,
                                            uint32_t row) {
    const V3<float> P{c.pos[0], c.pos[1], c.pos[2]};
    const float r2 = dot3(P, P);
    bool exact = !(r2 > 9.0f);
    bool in = false;
    if (!exact) {
        float px, py;
        if (c.jitter_hashed) {
            const uint32_t pi2 = 2u * (uint32_t)pix;
            px = (float)col + (float)(pcg(c.jitter_base + pi2) >> 8) * 0x1p-24f;
            py = (float)row + (float)(pcg(c.jitter_base + pi2 + 1u) >> 8) * 0x1p-24f;
        } else {
            px = ((float)col + 0.5f) + c.jitter[0];
            py = ((float)row + 0.5f) + c.jitter[1];
        }
        const float wf = (float)c.width, ih = __builtin_amdgcn_rcpf((float)c.height);
        const float ndx = __builtin_fmaf(px * __builtin_amdgcn_rcpf(wf), 2.0f, -1.0f);
        const float ndy = __builtin_fmaf(py * ih, 2.0f, -1.0f);
        const float dx = (ndx * c.tan_half_fov) * (wf * ih), dy = ndy * c.tan_half_fov;
        const float* B = c.basis;
        const V3<float> d{__builtin_fmaf(B[3], dy, B[0] * dx) - B[6],
                          __builtin_fmaf(B[4], dy, B[1] * dx) - B[7],
                          __builtin_fmaf(B[5], dy, B[2] * dx) - B[8]};
        const float ir = __builtin_amdgcn_rsqf(r2);
        const V3<float> cr{P.y * d.z - P.z * d.y, P.z * d.x - P.x * d.z, P.x * d.y - P.y * d.x};
        // |rh x d| / |d| with rh = P / r, times r / sqrt(1 - 2/r)
        const float sin_a = __builtin_amdgcn_sqrtf(dot3(cr, cr)) * __builtin_amdgcn_rsqf(dot3(d, d)) * ir;
        const float b = sin_a * __builtin_amdgcn_sqrtf(r2) *
                        __builtin_amdgcn_rsqf(1.0f - 2.0f * ir);
        const float dist = __builtin_fabsf(b - kBcritF) - c.band_abs;
        exact = !(__builtin_fabsf(dist) >= kBandScanMargin);   // NaN included
        in = dist < 0.0f;
    }
    if (exact) in = in_band(c, lawS_setup(P, ray_dir(c, pix)).b);
    return in;
}

// ------------------------------------------------------------------------------------------------
// Lists the work items of the FP64 band (float setup, |b - b_crit| < band_abs; band_member): one
// atomic per wave (ballot of the band lanes, per-lane prefix count); list order is unspecified.
template <typename Cams>
__global__ __launch_bounds__(kBlock) void k_band_scan(const Cams cams, Scratch sc) {
    for (uint32_t q0 = blockIdx.x * kBlock; q0 < cams.n_items; q0 += gridDim.x * kBlock) {
        const uint32_t q = q0 + threadIdx.x;
        bool want = false;
        if (q < cams.n_items) {
            const uint32_t ci = cams.find(q);
            const MarchConst& c = cams.cam(ci);
            uint32_t col, row;
            item_col_row(item_tiles_x((uint32_t)c.width), q - cams.base(ci), col, row);
            if (c.precision == BHS_PREC_F32_BAND64 && col < (uint32_t)c.width &&
                row < (uint32_t)c.height)
                want = band_member(c, (int64_t)row * c.width + col, col, row);
        }
        const uint64_t m = __ballot(want);
        if (m == 0) continue;
        uint32_t base = 0;
        if ((threadIdx.x & 63u) == 0) base = atomicAdd(sc.band_count, (uint32_t)__popcll(m));
        base = (uint32_t)__shfl((int)base, 0);
        if (want) {
            const uint32_t k = base + __builtin_amdgcn_mbcnt_hi(
                                          (uint32_t)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            if (k < sc.band_cap) sc.band_ids[k] = q;
        }
    }
}

// FP64 refinement of crossing `slot` of a band ray (as march_f64): the refined record replaces the
// chord one when the double step brackets the root.
__device__ __forceinline__ void refine_band_crossing(const SoA& o, int64_t i, uint32_t slot,
                                                     const RootState<double>& q,
                                                     const Setup<double>& s) {
    double hx, hz;
    if (!root64(q.x0, q.v0, q.dt, s.c3, hx, hz)) return;
    const double rh = __builtin_sqrt(__builtin_fma(hz, hz, hx * hx));
    const int64_t j = i + (int64_t)slot * o.n;
    o.disk[j] = (float)rh;
    o.disk[j + 2 * o.n] = atan2_ieee((float)hz, (float)hx);
    o.disk[j + 4 * o.n] = (float)redshift(rh, s.b_axis, s.lapse_g);
}

// Work-queue refill of a wave's idle lanes (every lane of the wave calls it with the same
// `more`): when at least `refill_min` lanes are idle, one atomic on `counter` claims that many
// queue positions and idle lane j (in lane order) gets position base + j. Returns true for a lane
// that received a position below n (in q); clears `more` once the queue is exhausted.
__device__ __forceinline__ bool claim_items(uint32_t* counter, uint32_t n, uint32_t refill_min,
                                            bool active, bool& more, uint32_t& q) {
    const uint64_t idle = __ballot(!active);
    const uint32_t n_idle = (uint32_t)__popcll(idle);
    if (!more || n_idle < refill_min) return false;
    uint32_t base = 0;
    if ((threadIdx.x & 63u) == 0) base = atomicAdd(counter, n_idle);
    base = (uint32_t)__shfl((int)base, 0);
    if (base + n_idle >= n) more = false;
    if (active) return false;
    q = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                         __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
    return q < n;
}

// Whole-tile claims of the persistent march when a wave refills only with all its lanes idle
// (refill threshold 64): a segmented queue. One counter on every claim serialises there when tiles
// are short (rays of a few steps each: every wave waits on the atomic), so the launch's tiles are
// split into nseg = min(kSegs, tiles) contiguous segments, each with its own counter (kSegStride
// words apart: separate cache lines). A wave starts on segment wave * nseg / waves and takes one
// tile per atomic; when its segment is used up it marks it in the shared mask seg_done and moves
// on to the next segment not yet marked (cyclically), so the waves of finished segments help the
// others and every tile is still handed out exactly once. `seg` and `seen` (the segments the wave
// knows to be used up) are wave-uniform. Returns true for a lane that received a position below n
// (in q); clears `more` once every segment is used up. Exit: each failed atomic marks a segment,
// the mask only grows, so a wave stops after at most nseg failures.
__device__ __forceinline__ uint32_t seg_first_tile(uint32_t s, uint32_t nseg, uint32_t tiles) {
    return (uint32_t)(((uint64_t)s * tiles) / nseg);
}

__device__ __forceinline__ bool claim_seg(const Scratch& sc, uint32_t n, bool active, bool& more,
                                          uint32_t& seg, uint64_t& seen, uint32_t& q) {
    if (!more || __any(active)) return false;
    const uint32_t tiles = (n + 63u) >> 6;
    const uint32_t nseg = tiles < kSegs ? tiles : kSegs;
    const uint64_t all = nseg == 64u ? ~0ull : ((1ull << nseg) - 1ull);
    while (true) {
        if ((seen & all) == all) {
            more = false;
            return false;
        }
        while ((seen >> seg) & 1ull) seg = seg + 1u == nseg ? 0u : seg + 1u;
        uint32_t t = 0;
        if ((threadIdx.x & 63u) == 0u) t = atomicAdd(sc.seg + seg * kSegStride, 1u);
        t = (uint32_t)__shfl((int)t, 0);
        const uint32_t tile = seg_first_tile(seg, nseg, tiles) + t;
        if (tile < seg_first_tile(seg + 1u, nseg, tiles)) {
            q = (tile << 6) + (threadIdx.x & 63u);
            return q < n;
        }
        uint64_t g = 0;   // segment used up: publish it and learn the others'
        if ((threadIdx.x & 63u) == 0u) g = atomicOr(sc.seg_done, 1ull << seg);
        seen |= (1ull << seg) | (uint64_t)__shfl((long long)g, 0);
    }
}

// Refill threshold of the band march: its rays run hundreds of FP64 iterations, so a refill (one
// setup) is cheap next to an idle lane; 32 keeps the queue atomics to one per 32 rays.
constexpr uint32_t kBandRefill = 32;

// Block of the band kernels: 256 threads, so the LDS camera table of a batch (~9 KB) does not cap
// the resident waves.
constexpr uint32_t kBandBlock = 256;

// Camera table of the band kernels in LDS (batches) or the kernel argument (one camera).
template <typename Cams>
struct BandCams {
    static constexpr bool kBatch = std::is_same<Cams, CamTable>::value;
    MarchConst cam[kBatch ? kMaxBatchCams : 1];
    uint32_t base[kBatch ? kMaxBatchCams + 1 : 1];
};

// FP64 march of a band list that the grid covers at once (n at most the grid's lanes; the count n
// is read on the device), one ray per lane: the band of a frame or of a moderate batch, whose
// longest ray sets the time. A longer list is left to k_band64_queue.
template <typename Cams>
__global__ __launch_bounds__(kBandBlock) __attribute__((amdgpu_waves_per_eu(4)))
void k_band64(const Cams cams, const SoA out, Scratch sc) {
    constexpr bool kBatch = BandCams<Cams>::kBatch;
    __shared__ BandCams<Cams> s_c;
    uint32_t ncam = 1;
    if constexpr (kBatch) {
        ncam = cams.ncam;
        for (uint32_t i = threadIdx.x; i < ncam; i += kBandBlock) s_c.cam[i] = cams.cams[i];
        for (uint32_t i = threadIdx.x; i <= ncam; i += kBandBlock) s_c.base[i] = cams.item_base[i];
        __syncthreads();
    }
    // Few waves with long serial chains run beside the FP32 march: issue them first on their SIMD.
    __builtin_amdgcn_s_setprio(3);
    const uint32_t n = *sc.band_count < sc.band_cap ? *sc.band_count : sc.band_cap;
    if (n > gridDim.x * kBandBlock) return;   // uniform: k_band64_queue's list
    const uint32_t q = blockIdx.x * kBandBlock + threadIdx.x;
    uint32_t one = 0, steps = 0, cross = 0, esc = 0;
    if (q < n) {
        const uint32_t item = sc.band_ids[q];
        uint32_t local = item, ci = 0;
        if constexpr (kBatch) {
            ci = find_cam(s_c.base, ncam, item);
            local = item - s_c.base[ci];
        }
        const MarchConst& c = kBatch ? s_c.cam[ci] : cams.cam(0);
        const int64_t p = item_pixel(c, local);
        if (p >= 0) {
            const V3<float> P{c.pos[0], c.pos[1], c.pos[2]};
            const RayOut o = march_f64(c, P, ray_dir(c, p));
            store_ray(out, c.ray_offset + p, o);
            one = 1;
            steps = (uint32_t)o.steps;
            cross = o.n_cross;
            esc = (o.flags & 2u) ? 1u : 0u;
        }
    }
    add_stats(sc.stats, one, steps, cross, esc, one);
}

// FP64 march of a band list longer than `lanes` (k_band64's lanes, or 0 when this kernel runs
// alone): every position, by resident waves pulling rays from it (queue counter sc.band_work);
// finished lanes are refilled as in k_persist, so a wave is not held by its slowest ray. ROOTS: the FP64
// root refinement of march_f64 for cameras with root64. Exit: the list is finite, every ray ends
// within max_steps iterations and the counter only grows, so every wave leaves. WIDE: register cap
// for five resident waves per SIMD, for lists longer than wide_n rays, where the list's volume,
// not its longest ray, sets the time; the other variant keeps the registers that make a single
// chain fastest and takes lists of at most wide_n rays (wide_n = ~0u: every list).
constexpr uint32_t kWideBand = 1024u * kBandBlock;

template <typename Cams, bool ROOTS, bool WIDE>
__global__ __launch_bounds__(kBandBlock)
__attribute__((amdgpu_waves_per_eu(ROOTS ? 2 : (WIDE ? 5 : 3))))
void k_band64_queue(const Cams cams, const SoA out, Scratch sc, uint32_t lanes, uint32_t wide_n) {
    constexpr bool kBatch = BandCams<Cams>::kBatch;
    __shared__ BandCams<Cams> s_c;
    __shared__ MarchK<double> s_k[kBatch ? kMaxBatchCams : 1];
    const uint32_t n = *sc.band_count < sc.band_cap ? *sc.band_count : sc.band_cap;
    if (n <= lanes) return;                   // uniform: k_band64's list
    if (WIDE != (n > wide_n)) return;         // uniform: the other variant's list
    uint32_t ncam = 1;
    if constexpr (kBatch) {
        ncam = cams.ncam;
        for (uint32_t i = threadIdx.x; i < ncam; i += kBandBlock) {
            s_c.cam[i] = cams.cams[i];
            s_k[i] = march_consts<double>(cams.cams[i]);
        }
        for (uint32_t i = threadIdx.x; i <= ncam; i += kBandBlock) s_c.base[i] = cams.item_base[i];
        __syncthreads();
    }
    __builtin_amdgcn_s_setprio(3);
    const MarchConst& c1 = cams.cam(0);
    const MarchK<double> k1 = march_consts<double>(c1);
    MarchState<double> st;
    Setup<double> s;
    RootState<double> rs0, rs1;
    uint32_t pix = 0, ci = 0;
    bool active = false;
    bool more = true;   // wave-uniform: the list may still hold rays
    uint32_t n_rays = 0, n_steps = 0, n_cross = 0, n_esc = 0;
    while (true) {
        uint32_t q = 0;
        if (claim_items(sc.band_work, n, kBandRefill, active, more, q)) {
            const uint32_t item = sc.band_ids[q];
            uint32_t local = item;
            if constexpr (kBatch) {
                ci = find_cam(s_c.base, ncam, item);
                local = item - s_c.base[ci];
            }
            const MarchConst& c = kBatch ? s_c.cam[ci] : c1;
            const int64_t p = item_pixel(c, local);
            if (p >= 0) {
                pix = (uint32_t)p;
                const V3<float> df = ray_dir(c, p);
                const V3<double> P{(double)c.pos[0], (double)c.pos[1], (double)c.pos[2]};
                const V3<double> d{(double)df.x, (double)df.y, (double)df.z};
                s = lawS_setup(P, d);
                march_init(st, P, s);
                active = true;
            }
        }
        if (!__any(active)) {
            if (!more) break;
            continue;
        }
        const MarchConst& c = kBatch ? s_c.cam[ci] : c1;
        const int64_t gi = c.ray_offset + pix;
        BandSink<ROOTS> sink{out, gi, rs0, rs1};
        const bool fin = march_step(st, s, kBatch ? s_k[ci] : k1, sink, active);
        if (active && fin) {
            uint32_t extra = BHS_FLAG_BAND64;
            if (ROOTS && c.root64 && st.n_cross > 0) {
                refine_band_crossing(out, gi, 0u, rs0, s);
                if (st.n_cross > 1) refine_band_crossing(out, gi, 1u, rs1, s);
                extra |= BHS_FLAG_ROOT64;
            }
            store_finish(out, gi, st, (float)s.b, extra);
            n_rays += 1;
            n_steps += (uint32_t)st.i;
            n_cross += st.n_cross;
            n_esc += st.status == 1 ? 1u : 0u;
            active = false;
        }
    }
    add_stats(sc.stats, n_rays, n_steps, n_cross, n_esc, n_rays);
}

// FP64 root refinement of the crossings recorded by the FP32 march (all items).
template <typename Cams>
__global__ __launch_bounds__(kBlock) void k_root64(const Cams cams, const SoA out, Scratch sc) {
    for (uint32_t q = blockIdx.x * kBlock + threadIdx.x; q < cams.n_items;
         q += gridDim.x * kBlock) {
        const uint32_t ci = cams.find(q);
        const MarchConst& c = cams.cam(ci);
        if (!c.root64) continue;
        const int64_t pix = item_pixel(c, q - cams.base(ci));
        if (pix < 0) continue;
        const int64_t g = c.ray_offset + pix;
        const uint32_t nc = out.n_cross[g];
        const uint32_t fl = out.flags[g];
        if (nc == 0 || (fl & (BHS_FLAG_BAND64 | BHS_FLAG_SURROGATE))) continue;
        const V3<float> P{c.pos[0], c.pos[1], c.pos[2]};
        // c3, b_axis and the redshift lapse of this ray, from its float setup.
        const Setup<float> s = lawS_setup(P, ray_dir(c, pix));
        const uint32_t m = nc < 2 ? nc : 2;
        for (uint32_t k = 0; k < m; ++k) {
            const RootState<float>& r = sc.roots[2 * g + k];
            const double X0[3] = {r.x0[0], r.x0[1], r.x0[2]};
            const double V0[3] = {r.v0[0], r.v0[1], r.v0[2]};
            double hx, hz;
            if (!root64(X0, V0, (double)r.dt, (double)s.c3, hx, hz)) continue;
            const double rh = __builtin_sqrt(__builtin_fma(hz, hz, hx * hx));
            out.disk[g + k * out.n] = (float)rh;
            out.disk[g + (2 + k) * out.n] = atan2_ieee((float)hz, (float)hx);
            out.disk[g + (4 + k) * out.n] =
                (float)redshift((double)rh, (double)s.b_axis, (double)s.lapse_g);
        }
        out.flags[g] = (uint8_t)(fl | BHS_FLAG_ROOT64);
    }
}

// ------------------------------------------------------------------------------------------------
// FP32 law-S render of one camera, one lane per pixel on a static grid of 8x8-pixel wave tiles
// (BLOCK = 64: one tile per block; 256: 16x16 pixels per block). In BAND64 precision the rays of
// the FP64 band are skipped here (k_band_scan lists them, k_band64 marches them).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(8)))
void k_render_f32(const MarchConst c, const SoA out, Scratch sc) {
    int col, row;
    if (BLOCK == 256) {
        tile_pixel(col, row);
    } else {
        col = blockIdx.x * 8 + (threadIdx.x & 7);
        row = blockIdx.y * 8 + (threadIdx.x >> 3);
    }
    const bool valid = col < c.width && row < c.height;
    uint32_t steps = 0, cross = 0, esc = 0, rays = 0;
    if (valid) {
        const int64_t pix = (int64_t)row * c.width + col;
        const V3<float> d = ray_dir(c, pix);
        const V3<float> P{c.pos[0], c.pos[1], c.pos[2]};
        const Setup<float> s = lawS_setup(P, d);
        if (!(c.precision == BHS_PREC_F32_BAND64 && in_band(c, s.b))) {
            RayOut o;
            lawS_march(P, s, c, o, sc.roots ? sc.roots + 2 * pix : (RootState<float>*)nullptr);
            store_ray(out, c.ray_offset + pix, o);
            steps = (uint32_t)o.steps;
            cross = o.n_cross;
            esc = (o.flags & 2u) ? 1u : 0u;
            rays = 1;
        }
    }
    add_stats(sc.stats, rays, steps, cross, esc, 0u);
}

// FP64 law-S render of every ray of one camera.
__global__ __launch_bounds__(kBlock) void k_render_f64(const MarchConst c, const SoA out, Scratch sc) {
    int col, row;
    tile_pixel(col, row);
    const bool valid = col < c.width && row < c.height;
    uint32_t steps = 0, cross = 0, esc = 0;
    if (valid) {
        const int64_t pix = (int64_t)row * c.width + col;
        const V3<float> P{c.pos[0], c.pos[1], c.pos[2]};
        const RayOut o = march_f64(c, P, ray_dir(c, pix));
        store_ray(out, c.ray_offset + pix, o);
        steps = (uint32_t)o.steps;
        cross = o.n_cross;
        esc = (o.flags & 2u) ? 1u : 0u;
    }
    add_stats(sc.stats, valid ? 1u : 0u, steps, cross, esc, valid ? 1u : 0u);
}

// ------------------------------------------------------------------------------------------------
// FP32 law-S march by resident waves that pull rays from a global work queue: one camera (BATCH =
// false, constants in kernel arguments) or up to kMaxBatchCams cameras of different sizes and step
// budgets (BATCH = true, camera table and item prefix sums staged in LDS; each lane looks up its
// ray's camera there). Each wave keeps 64 rays in flight; when at least `refill_min` lanes have
// finished, it takes that many new rays with one atomic (wave ballot + per-lane prefix count of the
// idle lanes), so finished lanes are refilled instead of idling until the slowest lane ends. With
// refill_min 64 (whole-wave refills) the wave takes whole tiles from a segmented queue
// (claim_seg), so that short tiles do not serialise every wave on one counter.
// Exit: a wave leaves when the queue is exhausted and all its lanes are finished; every ray ends
// within its max_steps iterations and the counter only grows, so every wave reaches it.
template <int BLOCK, bool ROOTS, int WPE, bool BATCH>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
void k_persist(const MarchConst c1, const MarchConst* __restrict__ table,
               const uint32_t* __restrict__ item_base, uint32_t ncam, uint32_t n_items,
               const SoA out, Scratch sc, uint32_t refill_min, const uint32_t* __restrict__ list,
               const uint32_t* __restrict__ list_count) {
    __shared__ MarchConst s_cam[BATCH ? kMaxBatchCams : 1];
    __shared__ MarchK<float> s_k[BATCH ? kMaxBatchCams : 1];
    __shared__ uint32_t s_base[BATCH ? kMaxBatchCams + 1 : 1];
    if (BATCH) {
        for (uint32_t i = threadIdx.x; i < ncam; i += BLOCK) {
            s_cam[i] = table[i];
            s_k[i] = march_consts<float>(table[i]);
        }
        for (uint32_t i = threadIdx.x; i <= ncam; i += BLOCK) s_base[i] = item_base[i];
        __syncthreads();
    }
    const MarchK<float> k1 = march_consts<float>(c1);
    if (list) {   // listed items: the count is known on the device only
        const uint32_t cnt = *list_count;
        n_items = cnt < n_items ? cnt : n_items;
        // Fallback lists are few long rays beside other kernels: issue them first on their SIMD
        // (below the FP64 band's priority 3).
        __builtin_amdgcn_s_setprio(2);
    }
    MarchState<float> st;
    Setup<float> s;
    // The lane's ray: its index in the output planes and (batches) its camera's step constants,
    // both set at refill, so a step reads them without recomputing the camera's addresses.
    int64_t gi = 0;
    const MarchK<float>* kp = BATCH ? &s_k[0] : nullptr;
    bool active = false;
    bool more = true;   // wave-uniform: the queue may still hold items
    // Whole-wave refills (threshold 64) of the launch's own items take tiles from the segmented
    // queue (claim_seg); listed items and partial refills share the single counter.
    const bool segq = refill_min >= 64u && list == nullptr;
    uint32_t seg = 0;
    uint64_t seen = 0;
    if (segq) {
        const uint32_t tiles = (n_items + 63u) >> 6, nseg = tiles < kSegs ? tiles : kSegs;
        const uint32_t wave = blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
        seg = (uint32_t)(((uint64_t)wave * nseg) / (gridDim.x * (BLOCK / 64)));
    }
    uint32_t n_rays = 0, n_steps = 0, n_cross = 0, n_esc = 0;
    while (true) {
        uint32_t q = 0;
        if (segq ? claim_seg(sc, n_items, active, more, seg, seen, q)
                 : claim_items(sc.work, n_items, refill_min, active, more, q)) {
            const uint32_t item = list ? list[q] : q;
            uint32_t local = item, ci = 0;
            if (BATCH) {
                ci = find_cam(s_base, ncam, item);
                local = item - s_base[ci];
            }
            const MarchConst& c = BATCH ? s_cam[ci] : c1;
            const int64_t p = item_pixel(c, local);
            if (p >= 0) {
                gi = c.ray_offset + p;
                if (BATCH) kp = &s_k[ci];
                const V3<float> P{c.pos[0], c.pos[1], c.pos[2]};
                s = lawS_setup(P, ray_dir(c, p));
                if (!(c.precision == BHS_PREC_F32_BAND64 && in_band(c, s.b))) {
                    march_init(st, P, s);
                    active = true;
                }
            }
        }
        if (!__any(active)) {
            if (!more) break;
            continue;
        }
        // Every lane steps (idle lanes on stale state, writing nothing), so the ray state stays in
        // place across the loop instead of being copied at a divergent merge.
        GlobalSink sink{out, gi, ROOTS ? sc.roots + 2 * gi : (RootState<float>*)nullptr};
        const bool fin = march_step(st, s, BATCH ? *kp : k1, sink, active);
        if (active && fin) {
            store_finish(out, gi, st, s.b);
            n_rays += 1;
            n_steps += (uint32_t)st.i;
            n_cross += st.n_cross;
            n_esc += st.status == 1 ? 1u : 0u;
            active = false;
        }
    }
    add_stats(sc.stats, n_rays, n_steps, n_cross, n_esc, 0u);
}

// ------------------------------------------------------------------------------------------------
// Law S on explicit rays (float position and direction per ray).
__global__ __launch_bounds__(kBlock) void k_march_rays(const float* __restrict__ pos3,
                                                       const float* __restrict__ dir3, int64_t n,
                                                       const MarchConst c, const SoA out, Scratch sc) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool valid = i < n;
    uint32_t steps = 0, cross = 0, esc = 0, band = 0;
    if (valid) {
        const V3<float> P{pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2]};
        const V3<float> d{dir3[3 * i], dir3[3 * i + 1], dir3[3 * i + 2]};
        const Setup<float> s = lawS_setup(P, d);
        RayOut o;
        if (c.precision == BHS_PREC_F64 ||
            (c.precision == BHS_PREC_F32_BAND64 && in_band(c, s.b))) {
            o = march_f64(c, P, d);
            band = 1;
        } else {
            RootState<float> rs[2];
            lawS_march(P, s, c, o, c.root64 ? rs : (RootState<float>*)nullptr);
            if (c.root64 && o.n_cross > 0) {
                const uint32_t m = o.n_cross < 2 ? o.n_cross : 2;
                for (uint32_t k = 0; k < m; ++k) {
                    const double X0[3] = {rs[k].x0[0], rs[k].x0[1], rs[k].x0[2]};
                    const double V0[3] = {rs[k].v0[0], rs[k].v0[1], rs[k].v0[2]};
                    double hx, hz;
                    if (!root64(X0, V0, (double)rs[k].dt, (double)s.c3, hx, hz)) continue;
                    const double rh = __builtin_sqrt(__builtin_fma(hz, hz, hx * hx));
                    o.dr[k] = (float)rh;
                    o.dphi[k] = atan2_ieee((float)hz, (float)hx);
                    o.dg[k] = (float)redshift((double)rh, (double)s.b_axis, (double)s.lapse_g);
                }
                o.flags |= BHS_FLAG_ROOT64;
            }
        }
        store_ray(out, i, o);
        steps = (uint32_t)o.steps;
        cross = o.n_cross;
        esc = (o.flags & 2u) ? 1u : 0u;
    }
    add_stats(sc.stats, valid ? 1u : 0u, steps, cross, esc, band);
}

// ------------------------------------------------------------------------------------------------
// Law T: bhs::physics::trace (src/physics/Geodesic.h:107-140), FP64, one lane per ray.
__global__ __launch_bounds__(64) void k_trace_T(const double* __restrict__ pos3,
                                                const double* __restrict__ dir3, int64_t n,
                                                int32_t max_steps, double step_scale, double esc_tol,
                                                double r_esc_min, bhs_trace_result* __restrict__ out) {
    using namespace lawT;
    const int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    const D3 cam{pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2]};
    const D3 dir{dir3[3 * i], dir3[3 * i + 1], dir3[3 * i + 2]};
    bhs_trace_result res;
    res.status = BHS_TRACE_MAXSTEPS;
    res.steps = 0;
    res.equator_crossings = 0;
    res.pad_ = 0;
    res.dir[0] = res.dir[1] = res.dir[2] = 0.0;
    res.min_r = 1e30;
    D3 x = cam;
    D3 v = initial_velocity(cam, dir);
    const D3 hv = gcross(x, v);
    const double h2 = gdot(hv, hv);
    const double h = __builtin_sqrt(h2);
    double prev_y = x.y;
    bool done = false;
    for (int32_t k = 0; k < max_steps; ++k) {
        const double r = glen(x);
        if (r < res.min_r) res.min_r = r;
        if (r < 2.0 + 0.01) {
            res.status = BHS_TRACE_CAPTURED;
            res.steps = k;
            done = true;
            break;
        }
        const double vlen = glen(v);
        const double b = h / vlen;
        if (gdot(x, v) > 0.0 && r > r_esc_min && 0.75 * b * b * b / (r * r * r * r) < esc_tol) {
            res.status = BHS_TRACE_ESCAPED;
            const D3 e = far_field(x, v, h2);
            res.dir[0] = e.x; res.dir[1] = e.y; res.dir[2] = e.z;
            res.steps = k;
            done = true;
            break;
        }
        const double ds0 = step_scale * r;
        const double ds = ds0 > 1e-4 ? ds0 : 1e-4;
        rk4(x, v, h2, ds / vlen);
        if ((x.y > 0.0) != (prev_y > 0.0)) ++res.equator_crossings;
        prev_y = x.y;
        res.steps = k + 1;
    }
    if (!done) res.status = BHS_TRACE_MAXSTEPS;
    out[i] = res;
}

__global__ void k_zero(uint32_t* p, size_t words) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// Grid of the FP64 band march (256-thread blocks over the device-side list): up to 4096 waves (4 per
// SIMD on 256 CUs); surplus lanes find no ray and exit.
dim3 band_grid(uint32_t cap) {
    uint32_t b = (cap + 255u) / 256u;
    b = b > 1024u ? 1024u : b;
    return dim3(b < 64u ? 64u : b);
}

// Grid of the item-indexed helper kernels (grid-stride loops).
dim3 item_grid(uint32_t n_items) {
    uint32_t b = (n_items + kBlock - 1) / kBlock;
    if (b > 8192u) b = 8192u;
    return dim3(b ? b : 1u);
}

// Persistent-kernel instantiation for a block size, a root-state flag, an occupancy target
// (waves per SIMD 8, 7, 6 or 4: the VGPR budget is 512 / waves) and single/batched cameras.
template <int BLOCK, bool ROOTS, bool BATCH>
void launch_persist_w(dim3 g, int wps, hipStream_t s, const MarchConst& c1, const MarchConst* table,
                      const uint32_t* item_base, uint32_t ncam, uint32_t n_items, const SoA& out,
                      const Scratch& sc, uint32_t rmin, const uint32_t* list,
                      const uint32_t* list_count) {
    if (wps >= 8)
        hipLaunchKernelGGL((k_persist<BLOCK, ROOTS, 8, BATCH>), g, dim3(BLOCK), 0, s, c1, table,
                           item_base, ncam, n_items, out, sc, rmin, list, list_count);
    else if (wps == 7)
        hipLaunchKernelGGL((k_persist<BLOCK, ROOTS, 7, BATCH>), g, dim3(BLOCK), 0, s, c1, table,
                           item_base, ncam, n_items, out, sc, rmin, list, list_count);
    else if (wps >= 5)
        hipLaunchKernelGGL((k_persist<BLOCK, ROOTS, 6, BATCH>), g, dim3(BLOCK), 0, s, c1, table,
                           item_base, ncam, n_items, out, sc, rmin, list, list_count);
    else
        hipLaunchKernelGGL((k_persist<BLOCK, ROOTS, 4, BATCH>), g, dim3(BLOCK), 0, s, c1, table,
                           item_base, ncam, n_items, out, sc, rmin, list, list_count);
}

template <bool BATCH>
hipError_t launch_persist(const Tuning& t, int num_cus, hipStream_t s, const MarchConst& c1,
                          const MarchConst* table, const uint32_t* item_base, uint32_t ncam,
                          uint32_t n_items, const SoA& out, const Scratch& sc,
                          const uint32_t* list = nullptr, const uint32_t* list_count = nullptr) {
    const int block = t.block == 64 ? 64 : 256;
    const int wps = t.waves_per_simd < 1 ? 1 : (t.waves_per_simd > 8 ? 8 : t.waves_per_simd);
    const uint32_t waves = (uint32_t)num_cus * 4u * (uint32_t)wps;   // resident-wave budget
    const uint32_t need = (n_items + 63u) / 64u;
    const uint32_t blocks = (waves < need ? waves : need) * 64u / (uint32_t)block;
    const uint32_t rmin = t.refill_min < 1 ? 1u : (t.refill_min > 64 ? 64u : (uint32_t)t.refill_min);
    const dim3 g(blocks ? blocks : 1u);
    const bool roots = sc.roots != nullptr;
    if (block == 64) {
        if (roots) launch_persist_w<64, true, BATCH>(g, wps, s, c1, table, item_base, ncam, n_items, out, sc, rmin, list, list_count);
        else launch_persist_w<64, false, BATCH>(g, wps, s, c1, table, item_base, ncam, n_items, out, sc, rmin, list, list_count);
    } else {
        if (roots) launch_persist_w<256, true, BATCH>(g, wps, s, c1, table, item_base, ncam, n_items, out, sc, rmin, list, list_count);
        else launch_persist_w<256, false, BATCH>(g, wps, s, c1, table, item_base, ncam, n_items, out, sc, rmin, list, list_count);
    }
    return hipGetLastError();
}

// Launches of more work items than this march their band list with k_band64_queue alone.
constexpr uint32_t kBandQueueItems = 1u << 24;

// FP64 march of the band list on `aux`. Up to kBandQueueItems work items: k_band64 (its list
// usually fits its lanes), then k_band64_queue, which exits at once unless the list is longer (the
// length is known on the device only). Larger launches: k_band64_queue alone (a k_band64 that
// finds the list too long would hold the stream until the other kernels free room for its blocks).
// The ROOTS variant (crossing states in registers) only for launches with FP64 root refinement.
template <typename Cams>
hipError_t launch_band64(const Cams& cams, const SoA& out, const Scratch& sc, hipStream_t aux,
                         bool wide_ok = false) {
    const dim3 g = band_grid(sc.band_cap);
    hipError_t e;
    uint32_t lanes = 0;
    if (sc.band_cap <= kBandQueueItems) {
        hipLaunchKernelGGL(k_band64<Cams>, g, dim3(kBandBlock), 0, aux, cams, out, sc);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        lanes = g.x * kBandBlock;
        if (sc.band_cap <= lanes) return hipSuccess;
    }
    bool roots;
    if constexpr (std::is_same<Cams, OneCam>::value)
        roots = cams.c.root64 != 0;
    else
        roots = sc.roots != nullptr;   // batches allocate root states iff a camera has root64
    if (roots) {
        hipLaunchKernelGGL((k_band64_queue<Cams, true, false>), g, dim3(kBandBlock), 0, aux, cams,
                           out, sc, lanes, ~0u);
        return hipGetLastError();
    }
    // batches of fallback lists (wide_ok) that may exceed kWideBand rays (the wide variant's
    // registers fit the batch form only): the other variant first (it exits at once on a longer
    // list), then the wide one (it exits at once on a shorter one): a kernel that only exits still
    // waits for slots beside a running pass, so the short, latency-bound list must not wait behind
    // it. Not beside the plain march, whose persistent blocks take the slots the wide blocks leave
    // for a moment, so that the band's waves would wait for them.
    const bool wide = wide_ok && BandCams<Cams>::kBatch && sc.band_cap > kWideBand;
    hipLaunchKernelGGL((k_band64_queue<Cams, false, false>), g, dim3(kBandBlock), 0, aux, cams,
                       out, sc, lanes, wide ? kWideBand : ~0u);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if constexpr (BandCams<Cams>::kBatch) {
        if (wide) {
            hipLaunchKernelGGL((k_band64_queue<Cams, false, true>), dim3(1280), dim3(kBandBlock),
                               0, aux, cams, out, sc, lanes, kWideBand);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

// The band fork: scan on s, FP64 march on aux between ev_fork and ev_join.
template <typename Cams>
hipError_t fork_band(const Cams& cams, const SoA& out, const Scratch& sc, hipStream_t s,
                     hipStream_t aux, hipEvent_t ev_fork, hipEvent_t ev_join) {
    hipError_t e;
    hipLaunchKernelGGL(k_band_scan<Cams>, item_grid(cams.n_items), dim3(kBlock), 0, s, cams, sc);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipEventRecord(ev_fork, s)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(aux, ev_fork, 0)) != hipSuccess) return e;
    if ((e = launch_band64(cams, out, sc, aux)) != hipSuccess) return e;
    return hipEventRecord(ev_join, aux);
}

}  // namespace

hipError_t launch_zero(void* p, size_t bytes, hipStream_t s) {
    const size_t words = bytes / 4;
    const unsigned blocks = (unsigned)((words + 255) / 256 < 1024 ? (words + 255) / 256 : 1024);
    hipLaunchKernelGGL(k_zero, dim3(blocks ? blocks : 1), dim3(256), 0, s, (uint32_t*)p, words);
    return hipGetLastError();
}

uint32_t camera_items(int32_t width, int32_t height) {
    return item_tiles_x((uint32_t)width) * (((uint32_t)height + kItemTileH - 1u) / kItemTileH) *
           64u;
}

hipError_t launch_render(const MarchConst& c, const SoA& out, const Scratch& sc, const Tuning& t,
                         int num_cus, hipStream_t s, hipStream_t aux, hipEvent_t ev_fork,
                         hipEvent_t ev_join, hipEvent_t ev_begin, hipEvent_t ev_end) {
    const dim3 grid((c.width + kTile - 1) / kTile, (c.height + kTile - 1) / kTile);
    const OneCam one{c, camera_items(c.width, c.height)};
    hipError_t e;
    const bool band = c.precision == BHS_PREC_F32_BAND64;
    // Fork: the FP64 band rays (few, but up to max_steps long) run on the auxiliary stream while
    // the FP32 march runs on `s`, instead of trailing it.
    if (band && (e = fork_band(one, out, sc, s, aux, ev_fork, ev_join)) != hipSuccess) return e;
    if (ev_begin && (e = hipEventRecord(ev_begin, s)) != hipSuccess) return e;
    if (c.precision == BHS_PREC_F64) {
        hipLaunchKernelGGL(k_render_f64, grid, dim3(kBlock), 0, s, c, out, sc);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    } else if (t.mode == 1) {
        if ((e = launch_persist<false>(t, num_cus, s, c, nullptr, nullptr, 1u, one.n_items, out,
                                       sc)) != hipSuccess)
            return e;
    } else if (t.block == 64) {
        const dim3 g8((c.width + 7) / 8, (c.height + 7) / 8);
        hipLaunchKernelGGL(k_render_f32<64>, g8, dim3(64), 0, s, c, out, sc);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    } else {
        hipLaunchKernelGGL(k_render_f32<256>, grid, dim3(256), 0, s, c, out, sc);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (ev_end && (e = hipEventRecord(ev_end, s)) != hipSuccess) return e;
    if (band && (e = hipStreamWaitEvent(s, ev_join, 0)) != hipSuccess) return e;
    if (c.precision != BHS_PREC_F64 && c.root64 && sc.roots) {
        hipLaunchKernelGGL(k_root64<OneCam>, item_grid(one.n_items), dim3(kBlock), 0, s, one, out, sc);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_batch(const MarchConst* table, const uint32_t* item_base, uint32_t ncam,
                        uint32_t n_items, bool any_band, bool any_root, const SoA& out,
                        const Scratch& sc, const Tuning& t, int num_cus, hipStream_t s,
                        hipStream_t aux, hipEvent_t ev_fork, hipEvent_t ev_join,
                        hipEvent_t ev_begin, hipEvent_t ev_end) {
    const CamTable cams{table, item_base, ncam, n_items};
    hipError_t e;
    // With a band of a batch of up to kBandQueueItems items (k_band64: one ray per lane, its
    // longest chain sets its time): the band scan and then the FP64 band march on `s`, the FP32
    // march on `aux` from the point after the scan, and `s` waits for it. The band kernel, next on
    // `s` when the scan ends, takes its SIMD slots before the march's persistent blocks fill them;
    // with the band on the side stream the march's blocks came first and the band's long FP64
    // chains ran on after the march (the 64-camera batch: 2.33 ms per eager call that way, 2.01 ms
    // this way). Larger batches keep the band on the side stream after the march's launch: their
    // band is a queue of millions of FP64 rays whose blocks, dispatched first, would take the
    // slots the march is tuned for (the full form: 122 against 118 ms).
    const bool band_first = any_band && n_items <= kBandQueueItems;
    if (any_band && !band_first &&
        (e = fork_band(cams, out, sc, s, aux, ev_fork, ev_join)) != hipSuccess)
        return e;
    hipStream_t m = s;
    if (band_first) {
        hipLaunchKernelGGL(k_band_scan<CamTable>, item_grid(n_items), dim3(kBlock), 0, s, cams, sc);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = hipEventRecord(ev_fork, s)) != hipSuccess) return e;
        if ((e = launch_band64(cams, out, sc, s)) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(aux, ev_fork, 0)) != hipSuccess) return e;
        m = aux;
    }
    if (ev_begin && (e = hipEventRecord(ev_begin, m)) != hipSuccess) return e;
    MarchConst none;
    __builtin_memset(&none, 0, sizeof none);
    // 256-thread blocks: the LDS camera table (~11 KB per block) would cap 64-thread blocks at
    // 3-4 waves per SIMD.
    Tuning tb = t;
    tb.block = 256;
    if ((e = launch_persist<true>(tb, num_cus, m, none, table, item_base, ncam, n_items, out, sc)) !=
        hipSuccess)
        return e;
    if (ev_end && (e = hipEventRecord(ev_end, m)) != hipSuccess) return e;
    if (band_first && (e = hipEventRecord(ev_join, aux)) != hipSuccess) return e;
    if (any_band && (e = hipStreamWaitEvent(s, ev_join, 0)) != hipSuccess) return e;
    if (any_root && sc.roots) {
        hipLaunchKernelGGL(k_root64<CamTable>, item_grid(n_items), dim3(kBlock), 0, s, cams, out, sc);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_list(const MarchConst& c, const uint32_t* ids, const uint32_t* count, uint32_t cap,
                       const SoA& out, const Scratch& sc, const Tuning& t, int num_cus,
                       hipStream_t s) {
    return launch_persist<false>(t, num_cus, s, c, nullptr, nullptr, 1u, cap, out, sc, ids, count);
}

hipError_t launch_band_list(const MarchConst& c, const SoA& out, const Scratch& sc, hipStream_t s,
                            hipStream_t aux, hipEvent_t ev_fork, hipEvent_t ev_join) {
    const OneCam one{c, camera_items(c.width, c.height)};
    hipError_t e;
    if ((e = hipEventRecord(ev_fork, s)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(aux, ev_fork, 0)) != hipSuccess) return e;
    if ((e = launch_band64(one, out, sc, aux)) != hipSuccess) return e;
    return hipEventRecord(ev_join, aux);
}

hipError_t launch_band_fork(const MarchConst& c, const SoA& out, const Scratch& sc, hipStream_t s,
                            hipStream_t aux, hipEvent_t ev_fork, hipEvent_t ev_join) {
    const OneCam one{c, camera_items(c.width, c.height)};
    return fork_band(one, out, sc, s, aux, ev_fork, ev_join);
}

hipError_t launch_root_refine(const MarchConst& c, const SoA& out, const Scratch& sc, hipStream_t s) {
    if (c.precision == BHS_PREC_F64 || !c.root64 || !sc.roots) return hipSuccess;
    const OneCam one{c, camera_items(c.width, c.height)};
    hipLaunchKernelGGL(k_root64<OneCam>, item_grid(one.n_items), dim3(kBlock), 0, s, one, out, sc);
    return hipGetLastError();
}

hipError_t launch_list_batch(const MarchConst* table, const uint32_t* item_base, uint32_t ncam,
                             const uint32_t* ids, const uint32_t* count, uint32_t cap,
                             const SoA& out, const Scratch& sc, const Tuning& t, int num_cus,
                             hipStream_t s) {
    MarchConst none;
    __builtin_memset(&none, 0, sizeof none);
    return launch_persist<true>(t, num_cus, s, none, table, item_base, ncam, cap, out, sc, ids,
                                count);
}

hipError_t launch_band_list_batch(const MarchConst* table, const uint32_t* item_base,
                                  uint32_t ncam, uint32_t n_items, const SoA& out,
                                  const Scratch& sc, hipStream_t s, hipStream_t aux,
                                  hipEvent_t ev_fork, hipEvent_t ev_join) {
    const CamTable cams{table, item_base, ncam, n_items};
    hipError_t e;
    if ((e = hipEventRecord(ev_fork, s)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(aux, ev_fork, 0)) != hipSuccess) return e;
    if ((e = launch_band64(cams, out, sc, aux, true)) != hipSuccess) return e;
    return hipEventRecord(ev_join, aux);
}

hipError_t launch_root_refine_batch(const MarchConst* table, const uint32_t* item_base,
                                    uint32_t ncam, uint32_t n_items, const SoA& out,
                                    const Scratch& sc, hipStream_t s) {
    if (!sc.roots) return hipSuccess;
    const CamTable cams{table, item_base, ncam, n_items};
    hipLaunchKernelGGL(k_root64<CamTable>, item_grid(n_items), dim3(kBlock), 0, s, cams, out, sc);
    return hipGetLastError();
}

hipError_t launch_march_rays(const float* pos3, const float* dir3, int64_t n, const MarchConst& c,
                             const SoA& out, const Scratch& sc, hipStream_t s, hipEvent_t ev_begin,
                             hipEvent_t ev_end) {
    const unsigned blocks = (unsigned)((n + kBlock - 1) / kBlock);
    hipError_t e;
    if (ev_begin && (e = hipEventRecord(ev_begin, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_march_rays, dim3(blocks), dim3(kBlock), 0, s, pos3, dir3, n, c, out, sc);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ev_end && (e = hipEventRecord(ev_end, s)) != hipSuccess) return e;
    return hipSuccess;
}

// The march's reciprocal square root over an array (grid-stride).
__global__ __launch_bounds__(kBlock) void k_rsqrt_rn(const float* __restrict__ x,
                                                     float* __restrict__ y, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
        y[i] = rsqrt_rn(x[i]);
}

hipError_t launch_rsqrt_rn(const float* x, float* y, int64_t n, hipStream_t s) {
    const int64_t want = (n + kBlock - 1) / kBlock;
    const unsigned blocks = (unsigned)(want < 65536 ? want : 65536);
    hipLaunchKernelGGL(k_rsqrt_rn, dim3(blocks), dim3(kBlock), 0, s, x, y, n);
    return hipGetLastError();
}

// VALU issue-rate probe: 16 independent chains per lane of one instruction class (0 v_fma_f32,
// 1 v_mul_f32, 2 v_fma_f64, 3 v_xor_b32 + v_add_u32, 4 v_rsq_f32), `iters` rounds.
template <int KIND>
__global__ __launch_bounds__(256) void k_valu_probe(float* __restrict__ out, int32_t iters, float a,
                                                    float b) {
    constexpr int K = 16;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    float x[K];
    double d[K];
    uint32_t u[K];
    for (int c = 0; c < K; ++c) {
        x[c] = 1.0f + 1e-3f * (float)(i & 1023u) + (float)c;
        d[c] = x[c];
        u[c] = i * 2654435761u + (uint32_t)c;
    }
    for (int32_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int c = 0; c < K; ++c) {
            if (KIND == 0) x[c] = __builtin_fmaf(x[c], a, b);
            if (KIND == 1) x[c] = x[c] * a;
            if (KIND == 2) d[c] = __builtin_fma(d[c], (double)a, (double)b);
            if (KIND == 3) u[c] = (u[c] ^ 0x5bd1e995u) + 0x27d4eb2du;
            if (KIND == 4) x[c] = __builtin_amdgcn_rsqf(x[c]);
        }
    }
    float s = 0.0f;
    for (int c = 0; c < K; ++c) s += x[c] + (float)d[c] + (float)(u[c] & 255u);
    out[i] = s;
}

hipError_t launch_valu_probe(int kind, uint32_t blocks, int32_t iters, float* out, hipStream_t s) {
    switch (kind) {
        case 0: hipLaunchKernelGGL(k_valu_probe<0>, dim3(blocks), dim3(256), 0, s, out, iters, 0.999f, 1e-3f); break;
        case 1: hipLaunchKernelGGL(k_valu_probe<1>, dim3(blocks), dim3(256), 0, s, out, iters, 0.999f, 1e-3f); break;
        case 2: hipLaunchKernelGGL(k_valu_probe<2>, dim3(blocks), dim3(256), 0, s, out, iters, 0.999f, 1e-3f); break;
        case 3: hipLaunchKernelGGL(k_valu_probe<3>, dim3(blocks), dim3(256), 0, s, out, iters, 0.999f, 1e-3f); break;
        default: hipLaunchKernelGGL(k_valu_probe<4>, dim3(blocks), dim3(256), 0, s, out, iters, 0.999f, 1e-3f); break;
    }
    return hipGetLastError();
}

hipError_t launch_trace_T(const double* pos3, const double* dir3, int64_t n, int32_t max_steps,
                          double step_scale, double esc_tol, double r_esc_min, void* out,
                          hipStream_t s) {
    const unsigned blocks = (unsigned)((n + 63) / 64);
    hipLaunchKernelGGL(k_trace_T, dim3(blocks), dim3(64), 0, s, pos3, dir3, n, max_steps, step_scale,
                       esc_tol, r_esc_min, (bhs_trace_result*)out);
    return hipGetLastError();
}

}  // namespace bhs

// mlp_infer.hip — surrogate inference: weight packing into MFMA fragment order, the wave-level bf16
// MFMA forward pass, output decoding, the fused surrogate render kernel with its fallback lists,
// and validation against integrator outputs.
#include "mlp.hpp"

#include <type_traits>

#include "../../include/bhs_trace.h"

namespace bhs {
namespace mlp {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f4;
typedef __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t u4;

// Packed layout (bytes): A fragments of layers 2 and 3 ([row tile 8][K chunk 4][lane 64][8 bf16]),
// of layer 4 ([K chunk 4][lane 64][8 bf16], rows 5..15 zero), layer 1 as [128][w0 w1 w2 b1] FP32,
// biases of layers 2, 3 (FP32 [128]) and 4 (FP32 [16], 5..15 zero).
constexpr int kOffA2 = 0, kOffA3 = 32768, kOffA4 = 65536, kOffW1 = 69632, kOffB2 = 71680,
              kOffB3 = 72192, kOffB4 = 72704;
static_assert(kOffB4 + 64 == (int)kPackedBytes, "packed layout");
static_assert(kPackedBytes % 16 == 0, "packed size");

// Neuron of K chunk kc, lane group g, element j. v_mfma_f32_16x16x32_bf16 reads A row l % 16 and B
// column l % 16 with K elements 8 (l / 16) + j from lane l, and leaves D rows 4 (l / 16) + jj of
// column l % 16 in lane l. Ordering the K elements of chunk kc as neurons
// 16 (2 kc + j / 4) + 4 g + j % 4 makes the D tiles of row tiles 2 kc and 2 kc + 1 exactly the B
// fragment of chunk kc of the next layer: activations never move between lanes.
__host__ __device__ constexpr int perm(int kc, int g, int j) {
    return 16 * (2 * kc + j / 4) + 4 * g + (j & 3);
}

__global__ void k_pack(const float* __restrict__ P, uint8_t* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    uint16_t* h = (uint16_t*)out;
    if (i < 2 * 16384) {
        const int layer = i / 16384, e = i % 16384;
        const int mt = e / 2048, kc = (e / 512) & 3, lane = (e / 8) & 63, j = e & 7;
        const int m = 16 * mt + (lane & 15), k = perm(kc, lane >> 4, j);
        h[(layer ? kOffA3 : kOffA2) / 2 + e] = bf16_bits(P[(layer ? kW3 : kW2) + m * kH + k]);
    } else if (i < 2 * 16384 + 2048) {
        const int e = i - 2 * 16384;
        const int kc = e / 512, lane = (e / 8) & 63, j = e & 7;
        const int m = lane & 15, k = perm(kc, lane >> 4, j);
        h[kOffA4 / 2 + e] = m < kOut ? bf16_bits(P[kW4 + m * kH + k]) : (uint16_t)0;
    } else if (i < 2 * 16384 + 2048 + kH) {
        const int q = i - (2 * 16384 + 2048);
        float* w1 = (float*)(out + kOffW1);
        w1[4 * q + 0] = P[kW1 + 3 * q + 0];
        w1[4 * q + 1] = P[kW1 + 3 * q + 1];
        w1[4 * q + 2] = P[kW1 + 3 * q + 2];
        w1[4 * q + 3] = P[kB1 + q];
        ((float*)(out + kOffB2))[q] = P[kB2 + q];
        ((float*)(out + kOffB3))[q] = P[kB3 + q];
        if (q < 16) ((float*)(out + kOffB4))[q] = q < kOut ? P[kB4 + q] : 0.0f;
    }
}

__device__ __forceinline__ void stage_weights(const uint8_t* __restrict__ g, uint8_t* s) {
    const u4* src = (const u4*)g;
    u4* dst = (u4*)s;
    for (int i = threadIdx.x; i < (int)(kPackedBytes / 16); i += blockDim.x) dst[i] = src[i];
}

typedef __attribute__((ext_vector_type(2))) float f2;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;

// Two floats rounded to bf16 (nearest even) in one 32-bit word, lo in bits 0-15.
__device__ __forceinline__ uint32_t pack_bf(float lo, float hi) {
    const f2 v = {lo, hi};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

__device__ __forceinline__ bf16x8 frag(const u4& v) { return __builtin_bit_cast(bf16x8, v); }

// One hidden MFMA layer for NC column tiles of 16 rays: out = bf16(silu(W in + b)).
template <int NC>
__device__ __forceinline__ void hidden_layer(const uint8_t* __restrict__ w, int off_a, int off_b,
                                             const u4 (&in)[NC][4], u4 (&out)[NC][4]) {
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const float* bias = (const float*)(w + off_b);
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) {
        // keep the scheduler from hoisting the A fragments of later row tiles (register pressure)
        __builtin_amdgcn_sched_barrier(0);
        bf16x8 A[4];
#pragma unroll
        for (int kc = 0; kc < 4; ++kc)
            A[kc] = *(const bf16x8*)(w + off_a + ((mt * 4 + kc) * 64 + lane) * 16);
        const f4 bq = *(const f4*)(bias + 16 * mt + 4 * g);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kc = 0; kc < 4; ++kc)
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[kc], frag(in[c][kc]), acc, 0, 0, 0);
            // rows 4g + jj of tile mt are elements 4 (mt % 2) + jj of chunk mt / 2
            out[c][mt >> 1][2 * (mt & 1)] = pack_bf(silu(acc[0] + bq[0]), silu(acc[1] + bq[1]));
            out[c][mt >> 1][2 * (mt & 1) + 1] = pack_bf(silu(acc[2] + bq[2]), silu(acc[3] + bq[3]));
        }
    }
}

// Network outputs of the 16 NC rays ray0 .. ray0 + 16 NC - 1 of the wave (lane l holds the features
// of ray l in x), written to ysc[(ray - ray0) * 5 + o].
template <int NC>
__device__ __forceinline__ void mlp_tiles(const uint8_t* __restrict__ w, const float x[3], int ray0,
                                          float* __restrict__ ysc) {
    const int lane = threadIdx.x & 63, g = lane >> 4, n = lane & 15;
    const f4* W1 = (const f4*)(w + kOffW1);
    float xs[NC][3];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int i = 0; i < 3; ++i) xs[c][i] = __shfl(x[i], ray0 + 16 * c + n, 64);
    u4 B[NC][4], H[NC][4];
#pragma unroll
    for (int kc = 0; kc < 4; ++kc)
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            __builtin_amdgcn_sched_barrier(0);
            const f4 wa = W1[perm(kc, g, 2 * p)], wb = W1[perm(kc, g, 2 * p + 1)];
            const float va[3] = {wa[0], wa[1], wa[2]}, vb[3] = {wb[0], wb[1], wb[2]};
#pragma unroll
            for (int c = 0; c < NC; ++c)
                B[c][kc][p] = pack_bf(silu(l1_pre(va, wa[3], xs[c])), silu(l1_pre(vb, wb[3], xs[c])));
        }
    hidden_layer<NC>(w, kOffA2, kOffB2, B, H);
    hidden_layer<NC>(w, kOffA3, kOffB3, H, B);
    const float* b4 = (const float*)(w + kOffB4);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kc = 0; kc < 4; ++kc)
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                *(const bf16x8*)(w + kOffA4 + (kc * 64 + lane) * 16), frag(B[c][kc]), acc, 0, 0, 0);
        float* yr = ysc + (16 * c + n) * kOut;
        if (g == 0) {
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) yr[jj] = acc[jj] + b4[jj];
        } else if (g == 1) {
            yr[4] = acc[0] + b4[4];
        }
    }
}

// Network forward pass of the wave's 64 rays in 64 / (16 NC) passes: lane l holds the features of
// ray l in x and receives its 5 outputs in y. ysc: 80 NC floats of LDS owned by this wave. All 64
// lanes must call it.
constexpr int kNC = 2;
__device__ __forceinline__ void mlp_wave(const uint8_t* __restrict__ w, const float x[3],
                                         float* __restrict__ ysc, float y[5]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int h = 0; h < 4 / kNC; ++h) {
        // The weights stay in LDS: without this the loads are hoisted out of the callers' loops
        // and the whole network is held in (spilled) registers.
        asm volatile("" ::: "memory");
        mlp_tiles<kNC>(w, x, 16 * kNC * h, ysc);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int rl = lane - 16 * kNC * h;
        const bool mine = rl >= 0 && rl < 16 * kNC;
        const int ri = mine ? rl : 0;
#pragma unroll
        for (int o = 0; o < kOut; ++o) {
            const float v = ysc[ri * kOut + o];
            if (mine) y[o] = v;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// ------------------------------------------------------------------------------------------------
// Decoded surrogate result of one ray (the law-S output record; steps unknown, reported as 0).
struct Dec {
    bool serve;
    bool escaping;
    uint32_t cls, flags, n_cross;
    float esc[3];
    float dr[2], dphi[2], dg[2];
    uint32_t n_eq;      // equatorial crossings before escape/capture (4 means four or more)
    float r_eq[2];      // radii of the first two equatorial crossings
};

__device__ __forceinline__ bool cell_served(const SurConst& k, const float x[3]) {
    int ir = (int)((x[0] + 1.0f) * (0.5f * kMapR));
    int ib = (int)((x[1] + 1.0f) * (0.5f * kMapB));
    ir = ir < 0 ? 0 : (ir >= kMapR ? kMapR - 1 : ir);
    ib = ib < 0 ? 0 : (ib >= kMapB ? kMapB - 1 : ib);
    const int cell = ir * kMapB + ib;
    return (k.serve[cell >> 5] >> (cell & 31)) & 1u;
}

__device__ __forceinline__ bool steps_ok(const MarchConst& c, const SurConst& k, int rb, float x2) {
    return (float)k.steps_max[step_bin_rb(rb, x2)] * k.steps_slack < (float)c.max_steps;
}

// Relative margin around the annulus edges for a third crossing (exact orbit, no learned residual).
constexpr float kThirdMargin = 0.05f;

// Output record of one in-domain ray from its swept angle at escape / capture Phi (learned
// residual included) and the u = 1/r of its equatorial crossings (u_at(q), q = 0, 1 with the
// learned residual, q = 2 the exact orbit's, asked only for crossings before Phi), or serve =
// false when the ray must go back to the integrator: a crossing radius within the margin of an
// annulus edge, a third crossing in the annulus that would be recorded (the first two are not
// both in it), or a fourth crossing. A third crossing is only counted (the march records two).
// redshiftFactor (disk.glsl:37-42) with hardware reciprocal and square-root approximations (the
// integrator's version rounds every operation exactly; a served record needs ~1e-6 relative).
__device__ __forceinline__ float redshift_fast(float r, float b_axis, float lapse) {
    const float omega = __builtin_amdgcn_rsqf(r * r * r);
    float t = 1.0f - 3.0f * __builtin_amdgcn_rcpf(r);
    if (t < 0.02f) t = 0.02f;
    float m = __builtin_fmaf(-omega, b_axis, 1.0f);
    if (m < 0.05f) m = 0.05f;
    return __builtin_amdgcn_sqrtf(t) * __builtin_amdgcn_rcpf(lapse * m);
}

template <typename UAt>
__device__ __forceinline__ Dec decode_with(const MarchConst& c, float margin, float b_axis,
                                           float lapse_g, const RayGeom& g, float Phi,
                                           UAt u_at) {
    Dec o;
    o.serve = false;
    o.escaping = g.bm > 0.0f;
    o.n_cross = 0;
    // validity of the crossings from the total sweep
    const float a = Phi - g.psi;
    o.n_eq = a <= 0.0f ? 0u
                       : (a <= kPi ? 1u : (a <= 2.0f * kPi ? 2u : (a <= 3.0f * kPi ? 3u : 4u)));
    if (o.n_eq >= 4u && c.thin) return o;
    bool in[2] = {false, false};
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        o.r_eq[q] = 1e30f;
        if ((uint32_t)q >= o.n_eq) continue;
        const float u = u_at(q);
        o.r_eq[q] = u > 0.0f ? __builtin_amdgcn_rcpf(u) : 1e30f;
        if (c.thin) {
            const float r = o.r_eq[q];
            if (fabsf(r - c.disk_inner) < margin * c.disk_inner ||
                fabsf(r - c.disk_outer) < margin * c.disk_outer)
                return o;
            in[q] = r >= c.disk_inner && r <= c.disk_outer;
        }
    }
    bool in3 = false;
    float r3 = 1e30f;
    if (o.n_eq == 3u && c.thin) {
        const float u = u_at(2);
        r3 = u > 0.0f ? __builtin_amdgcn_rcpf(u) : 1e30f;
        const float m3 = fmaxf(margin, kThirdMargin);
        if (fabsf(r3 - c.disk_inner) < m3 * c.disk_inner ||
            fabsf(r3 - c.disk_outer) < m3 * c.disk_outer)
            return o;
        in3 = r3 >= c.disk_inner && r3 <= c.disk_outer;
    }
    o.dr[0] = o.dr[1] = o.dphi[0] = o.dphi[1] = o.dg[0] = o.dg[1] = 0.0f;
    // Azimuth of the first crossing: the direction cos(psi) e1 + sin(psi) e2 with zero y is
    // +-(e2.y e1 - e1.y e2) (psi in [0, pi): the sign that makes sin(psi) >= 0); the second crossing
    // lies opposite.
    float ph0 = 0.0f;
    if (in[0] || in[1] || in3) {
        const float sg = g.e1.y > 0.0f ? -1.0f : 1.0f;
        ph0 = atan2_fast(sg * (g.e2.y * g.e1.z - g.e1.y * g.e2.z),
                         sg * (g.e2.y * g.e1.x - g.e1.y * g.e2.x));
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        if (!in[q]) continue;
        const float r = o.r_eq[q];
        const float ph = q == 0 ? ph0 : (ph0 > 0.0f ? ph0 - kPi : ph0 + kPi);
        const float gg = redshift_fast(r, b_axis, lapse_g);
        if (o.n_cross == 0) {
            o.dr[0] = r;
            o.dphi[0] = ph;
            o.dg[0] = gg;
        } else {
            o.dr[1] = r;
            o.dphi[1] = ph;
            o.dg[1] = gg;
        }
        ++o.n_cross;
    }
    if (in3) {   // recorded if fewer than two crossings are (same azimuth as the first)
        if (o.n_cross < 2u) {   // constant indices keep the record in registers
            const float g3 = redshift_fast(r3, b_axis, lapse_g);
            if (o.n_cross == 0u) { o.dr[0] = r3; o.dphi[0] = ph0; o.dg[0] = g3; }
            else { o.dr[1] = r3; o.dphi[1] = ph0; o.dg[1] = g3; }
        }
        ++o.n_cross;
    }
    if (o.escaping) {
        // hardware sine / cosine (absolute error ~1e-6 over the swept angles' range)
        float sf, cf;
        __sincosf(Phi, &sf, &cf);
        o.esc[0] = cf * g.e1.x + sf * g.e2.x;
        o.esc[1] = cf * g.e1.y + sf * g.e2.y;
        o.esc[2] = cf * g.e1.z + sf * g.e2.z;
    } else {
        o.esc[0] = o.esc[1] = o.esc[2] = 0.0f;
    }
    o.cls = o.n_cross >= 1 ? BHS_CLASS_DISK : (o.escaping ? BHS_CLASS_ESCAPE : BHS_CLASS_HORIZON);
    o.flags = BHS_FLAG_SURROGATE | (o.escaping ? BHS_FLAG_ESCAPED : 0u);
    o.serve = true;
    return o;
}

// Decode from the exact orbit of the ray and the network outputs y (per-ray inference,
// validation).
__device__ __forceinline__ Dec decode(const MarchConst& c, const SurConst& k, float b_axis,
                                      float lapse_g, const RayGeom& g, const float y[5]) {
    const Orbit ob = orbit(g);
    // the exact sweep to infinity with the learned residual, or to the horizon test radius
    const float Phi = ob.phi_end + (g.bm > 0.0f ? kResScale * y[0] : 0.0f);
    return decode_with(c, k.radius_margin[r_bin(g.r)], b_axis, lapse_g, g, Phi, [&](int q) {
        const float u = orbit_u(ob, g.psi + (float)q * kPi);
        return q < 2 ? u * __expf(kURes * y[1 + q]) : u;
    });
}

// Relative margin around the annulus for the crossing of a plunging ray.
constexpr float kPlungeMargin = 0.05f;

// Record of a plunging ray from the exact orbit alone (no learned residual): captured; served as
// such when it meets the equatorial plane before the horizon test radius at most once, at a radius
// outside the annulus widened by kPlungeMargin (the crossing then records nothing).
__device__ __forceinline__ Dec decode_plunge(const MarchConst& c, const RayGeom& g) {
    Dec o;
    o.serve = false;
    o.escaping = false;
    o.n_cross = 0;
    o.r_eq[0] = o.r_eq[1] = 1e30f;
    if (c.thin) {
        const Orbit ob = orbit(g);
        const float a = ob.phi_end - g.psi;
        o.n_eq = a <= 0.0f ? 0u : (a <= kPi ? 1u : 2u);
        if (o.n_eq >= 2u) return o;
        if (o.n_eq == 1u) {
            const float u = orbit_u(ob, g.psi);
            const float r = u > 0.0f ? __builtin_amdgcn_rcpf(u) : 1e30f;
            o.r_eq[0] = r;
            if (r > c.disk_inner * (1.0f - kPlungeMargin) &&
                r < c.disk_outer * (1.0f + kPlungeMargin))
                return o;
        }
    } else {
        o.n_eq = 0;
    }
    o.esc[0] = o.esc[1] = o.esc[2] = 0.0f;
    o.dr[0] = o.dr[1] = o.dphi[0] = o.dphi[1] = o.dg[0] = o.dg[1] = 0.0f;
    o.cls = BHS_CLASS_HORIZON;
    o.flags = BHS_FLAG_SURROGATE;
    o.serve = true;
    return o;
}

// FULL: the caller guarantees every plane (the kernel is instantiated for it, so no per-plane test
// reaches the compiled code).
template <bool FULL = false>
__device__ __forceinline__ void store_dec(const SoA& out, int64_t i, const Dec& o, float b) {
    const int64_t n = out.n;
    if (FULL || (out.cls && out.flags && out.n_cross && out.steps && out.esc && out.disk && out.b)) {
        // every plane requested (the usual case): no per-plane tests
        out.cls[i] = (uint8_t)o.cls;
        out.flags[i] = (uint8_t)o.flags;
        out.n_cross[i] = (uint8_t)o.n_cross;
        out.steps[i] = 0;
        out.esc[i] = o.esc[0];
        out.esc[i + n] = o.esc[1];
        out.esc[i + 2 * n] = o.esc[2];
        out.disk[i] = o.dr[0];
        out.disk[i + n] = o.dr[1];
        out.disk[i + 2 * n] = o.dphi[0];
        out.disk[i + 3 * n] = o.dphi[1];
        out.disk[i + 4 * n] = o.dg[0];
        out.disk[i + 5 * n] = o.dg[1];
        out.b[i] = b;
        return;
    }
    if (out.cls) out.cls[i] = (uint8_t)o.cls;
    if (out.flags) out.flags[i] = (uint8_t)o.flags;
    if (out.n_cross) out.n_cross[i] = (uint8_t)o.n_cross;
    if (out.steps) out.steps[i] = 0;
    if (out.esc) {
        out.esc[i] = o.esc[0];
        out.esc[i + n] = o.esc[1];
        out.esc[i + 2 * n] = o.esc[2];
    }
    if (out.disk) {
        out.disk[i] = o.dr[0];
        out.disk[i + n] = o.dr[1];
        out.disk[i + 2 * n] = o.dphi[0];
        out.disk[i + 3 * n] = o.dphi[1];
        out.disk[i + 4 * n] = o.dg[0];
        out.disk[i + 5 * n] = o.dg[1];
    }
    if (out.b) out.b[i] = b;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Fallback lists of the surrogate kernels: a ray that is not served goes to at most one, the FP32
// late list, the call's FP64 band list, the early FP32 list or the late band list.
enum Dest : int { kToNone = -1, kToFb = 0, kToBand = 1, kToEarly = 2, kToLate = 3 };
constexpr int kDests = 4;

__device__ __forceinline__ uint32_t* list_count(const Scratch& sc, const SurLists& fl, int j) {
    return j == kToFb ? fl.fb_count
                      : (j == kToBand ? sc.band_count
                                      : (j == kToEarly ? fl.early_count : fl.lb_count));
}
__device__ __forceinline__ uint32_t* list_ids(const Scratch& sc, const SurLists& fl, int j) {
    return j == kToFb ? fl.fb_ids
                      : (j == kToBand ? sc.band_ids : (j == kToEarly ? fl.early_ids : fl.lb_ids));
}

// Appends each lane's item to list `dest` (kToNone: none): for each list that lanes of the wave
// are bound for, one atomic by the wave's first lane reserves their slots, and a lane's slot is its
// rank among them (ballot + mbcnt). Every lane of the wave calls it, in wave-uniform control flow;
// no block barrier, so the waves of a block do not wait for each other's chunks. Most waves list
// nothing (one ballot) and the rest few rays, so the atomics stay rare.
__device__ __forceinline__ void push_lists(const Scratch& sc, const SurLists& fl, int dest,
                                           uint32_t item) {
    if (__ballot(dest >= 0) == 0ull) return;
#pragma unroll
    for (int j = 0; j < kDests; ++j) {
        const uint64_t m = __ballot(dest == j);
        if (m == 0ull) continue;
        uint32_t base = 0;
        if ((threadIdx.x & 63u) == 0u) base = atomicAdd(list_count(sc, fl, j), (uint32_t)__popcll(m));
        base = (uint32_t)__shfl((int)base, 0);
        if (dest == j) {
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            const uint32_t cap = j == kToBand ? sc.band_cap : fl.fb_cap;
            if (base + rank < cap) list_ids(sc, fl, j)[base + rank] = item;
        }
    }
}

// Adds the block's counters to the launch statistics with one atomic per counter and block.
__device__ __forceinline__ void add_block_stats(Stats* st, uint32_t served, uint32_t cross,
                                                uint32_t esc, uint32_t fb, uint32_t (*s_st)[4]) {
    const int wave = threadIdx.x >> 6;
    served = wave_sum(served);
    cross = wave_sum(cross);
    esc = wave_sum(esc);
    fb = wave_sum(fb);
    if ((threadIdx.x & 63) == 0) {
        s_st[wave][0] = served;
        s_st[wave][1] = cross;
        s_st[wave][2] = esc;
        s_st[wave][3] = fb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t v[4] = {0, 0, 0, 0};
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
            for (int k = 0; k < 4; ++k) v[k] += s_st[w][k];
        if (v[0]) {
            atomicAdd(&st->rays, (unsigned long long)v[0]);
            atomicAdd(&st->surrogate, (unsigned long long)v[0]);
        }
        if (v[1]) atomicAdd(&st->crossings, (unsigned long long)v[1]);
        if (v[2]) atomicAdd(&st->escapes, (unsigned long long)v[2]);
        if (v[3]) atomicAdd(&st->fallback, (unsigned long long)v[3]);
    }
}

constexpr int kBlock = 256;

// As add_block_stats, into the block's four words of `part` (no atomics; k_stats_reduce adds the
// partials of a launch to the statistics).
__device__ __forceinline__ void store_block_stats(uint32_t* part, uint32_t served, uint32_t cross,
                                                  uint32_t esc, uint32_t fb,
                                                  uint32_t (*s_st)[4]) {
    const int wave = threadIdx.x >> 6;
    served = wave_sum(served);
    cross = wave_sum(cross);
    esc = wave_sum(esc);
    fb = wave_sum(fb);
    if ((threadIdx.x & 63) == 0) {
        s_st[wave][0] = served;
        s_st[wave][1] = cross;
        s_st[wave][2] = esc;
        s_st[wave][3] = fb;
    }
    __syncthreads();
    if (threadIdx.x < 4u)
        part[4 * blockIdx.x + threadIdx.x] =
            s_st[0][threadIdx.x] + s_st[1][threadIdx.x] + s_st[2][threadIdx.x] + s_st[3][threadIdx.x];
}

// Sums n_blocks blocks' partial statistics (grid-stride; one atomic per counter and block).
__global__ __launch_bounds__(kBlock) void k_stats_reduce(const uint32_t* __restrict__ part,
                                                         uint32_t n_blocks, Stats* st) {
    __shared__ unsigned long long s_v[4][kBlock / 64];
    unsigned long long v[4] = {0, 0, 0, 0};
    for (uint32_t b = blockIdx.x * kBlock + threadIdx.x; b < n_blocks; b += gridDim.x * kBlock)
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] += part[4 * b + k];
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
        if ((threadIdx.x & 63) == 0) s_v[k][wave] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t[4] = {0, 0, 0, 0};
        for (int w = 0; w < kBlock / 64; ++w)
            for (int k = 0; k < 4; ++k) t[k] += s_v[k][w];
        if (t[0]) {
            atomicAdd(&st->rays, t[0]);
            atomicAdd(&st->surrogate, t[0]);
        }
        if (t[1]) atomicAdd(&st->crossings, t[1]);
        if (t[2]) atomicAdd(&st->escapes, t[2]);
        if (t[3]) atomicAdd(&st->fallback, t[3]);
    }
}

// Camera sources of the surrogate kernels: one camera by value, or the device camera table of a
// batch with the prefix sums of the per-camera work items (all multiples of 64, so a wave's 64
// items always belong to one camera).
struct SurOneCam {
    MarchConst c;
    __device__ __forceinline__ uint32_t find(uint32_t) const { return 0u; }
    __device__ __forceinline__ const MarchConst& cam(uint32_t) const { return c; }
    __device__ __forceinline__ uint32_t base(uint32_t) const { return 0u; }
    __device__ __forceinline__ uint32_t end(uint32_t) const { return ~0u; }
};
// The camera table and the prefix sums are read-only during a launch and the surrogate kernels
// index them with wave-uniform values: read through the constant address space, those loads
// become scalar loads and the camera's fields live in scalar registers.
struct SurCamTable {
    const MarchConst* cams;
    const uint32_t* item_base;
    uint32_t ncam;
    using ConstU32 = const __attribute__((address_space(4))) uint32_t;
    __device__ __forceinline__ uint32_t find(uint32_t q) const {
        ConstU32* ib = (ConstU32*)item_base;
        uint32_t lo = 0, hi = ncam;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (ib[mid] <= q) lo = mid; else hi = mid;
        }
        return lo;
    }
    __device__ __forceinline__ MarchConst cam(uint32_t k) const {
        static_assert(sizeof(MarchConst) % 4 == 0, "dword copy");
        MarchConst c;
        ConstU32* src = (ConstU32*)(cams + k);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&c);
#pragma unroll
        for (uint32_t i = 0; i < sizeof(MarchConst) / 4; ++i) dst[i] = src[i];
        return c;
    }
    __device__ __forceinline__ uint32_t base(uint32_t k) const {
        return ((ConstU32*)item_base)[k];
    }
    // one past camera k's last item (the prefix sums hold ncam + 1 entries)
    __device__ __forceinline__ uint32_t end(uint32_t k) const {
        return ((ConstU32*)item_base)[k + 1];
    }
};

// Camera and pixel of the calling lane's item in the wave whose first item is q_wave: a wave's 64
// items are one work-item tile of one camera, so the camera search, the camera's constants and the
// tile position are wave-uniform (scalar) and only the pixel inside the tile is per lane.
struct WaveItem {
    uint32_t ci;
    int64_t pix;     // -1: padding or past the end
    uint32_t col, row;
    uint32_t item;   // the pixel's work item in the integrator's tile order (for the lists)
};
template <typename Cams>
__device__ __forceinline__ WaveItem wave_item(const Cams& cams, uint32_t q_wave, uint32_t n_items) {
    const uint32_t qw = __builtin_amdgcn_readfirstlane(q_wave);
    const uint32_t lane = threadIdx.x & 63u;
    WaveItem w;
    w.ci = cams.find(qw < n_items ? qw : n_items - 1);
    const MarchConst& c = cams.cam(w.ci);
    tile_col_row(item_tiles_x((uint32_t)c.width), (qw - cams.base(w.ci)) >> 6, lane, w.col,
                 w.row);
    w.pix = (qw + lane < n_items && w.col < (uint32_t)c.width && w.row < (uint32_t)c.height)
                ? (int64_t)w.row * c.width + w.col
                : -1;
    w.item = qw + lane;
    return w;
}

// The camera and scanline position of a wave's previous item, carried through a block's chunk
// loop (wave-uniform, in scalar registers): a block's consecutive chunks mostly belong to one
// camera, so the camera search and the division of the pixel index by the width are skipped
// while the wave stays in it and moves on by less than a row.
struct ScanCursor {
    uint32_t ci = 0u, lo = 1u, hi = 0u;    // camera ci owns items [lo, hi) (empty at first)
    uint32_t p = 0u, row = 0u, col = 0u;   // pixel index of the previous item, its row and column
    bool valid = false;
};

// As wave_item, with the camera's items taken in scanline order (item base + p is pixel p, the
// items beyond width * height are padding): a wave's 64 rays are 64 consecutive pixels, so its
// stores to the output planes are contiguous (a work-item tile writes short row pieces).
// `item` maps the pixel back to the integrator's tile order.
template <typename Cams>
__device__ __forceinline__ WaveItem wave_item_scan(const Cams& cams, uint32_t q_wave,
                                                   uint32_t n_items, ScanCursor& cur) {
    const uint32_t qw = __builtin_amdgcn_readfirstlane(q_wave);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t qc = qw < n_items ? qw : n_items - 1;
    if (qc < cur.lo || qc >= cur.hi) {             // another camera (uniform)
        cur.ci = cams.find(qc);
        cur.lo = cams.base(cur.ci);
        cur.hi = cams.end(cur.ci);
        cur.valid = false;
    }
    WaveItem w;
    w.ci = cur.ci;
    const MarchConst& c = cams.cam(w.ci);
    const uint32_t base = cur.lo;
    const uint32_t width = (uint32_t)c.width;
    const uint32_t p0 = qw - base;                 // wave-uniform
    uint32_t row0, col0;
    if (cur.valid && p0 >= cur.p && p0 - cur.p < width) {
        col0 = cur.col + (p0 - cur.p);
        row0 = cur.row;
        if (col0 >= width) {
            col0 -= width;
            ++row0;
        }
    } else {
        row0 = p0 / width;
        col0 = p0 - row0 * width;
    }
    cur.p = p0;
    cur.row = row0;
    cur.col = col0;
    cur.valid = true;
    uint32_t col = col0 + lane, row = row0;
    while (col >= width) {                         // at most 64 / width turns
        col -= width;
        ++row;
    }
    w.col = col;
    w.row = row;
    w.pix = (qw + lane < n_items && row < (uint32_t)c.height) ? (int64_t)row * width + col : -1;
    w.item = base + pixel_item(item_tiles_x(width), col, row);
    return w;
}

// Route of one ray of a surrogate render, decided before the network runs.
enum Route : uint32_t {
    kNone = 0,    // padding item
    kBand = 1,    // integrator's FP64 band (BAND64 precision): FP64 march from the start
    kEarly = 2,   // outside the served domain / unserved cell / step budget: FP32 march at once
    kNet = 3,     // network candidate (the decode may still send it back)
    kPlunge = 4,  // plunging ray: exact orbit alone (decode_plunge may send it back)
};

struct RayPrep {
    RayGeom g;
    float b_axis, lapse_g;   // impact parameter about the spin axis, lapse (redshift)
    float margin;            // relative crossing-radius margin of the camera's r_obs bin
    float x[3];
    Route route;
    bool band;   // in the integrator's FP64 band (BAND64 precision)
};

// The per-camera terms of the surrogate's ray setup (16 words): r_obs, its reciprocal and the
// radial unit vector, 1/sqrt(1 - 2/r) (impact parameter), the b_axis and lapse factors, the
// pixel-to-NDC scales, the x1 feature and the log r_obs bin. Computed once per camera (table mode:
// k_tab_range stores them) instead of once per wave.
struct SurCam {
    float r, ir, e1[3], inv_f, bax_k, lapse_g, inv_w, inv_h, asp, x1;
    int32_t rbin;
    float pad_[3];
};
static_assert(sizeof(SurCam) == 64, "SurCam layout");

__device__ __forceinline__ SurCam sur_cam(const MarchConst& c) {
    SurCam pc;
    const V3<float> P{c.pos[0], c.pos[1], c.pos[2]};
    pc.r = __builtin_amdgcn_sqrtf(dot3(P, P));
    pc.ir = __builtin_amdgcn_rcpf(pc.r);
    pc.e1[0] = P.x * pc.ir;
    pc.e1[1] = P.y * pc.ir;
    pc.e1[2] = P.z * pc.ir;
    const float lapse2 = 1.0f - 2.0f * pc.ir;
    pc.inv_f = __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(fmaxf(lapse2, 1e-12f)));
    pc.bax_k = __builtin_amdgcn_rsqf(fmaxf(lapse2, 1e-6f));
    pc.lapse_g = __builtin_amdgcn_sqrtf(fmaxf(lapse2, 1e-4f));
    const float wf = (float)c.width, hf = (float)c.height;
    pc.inv_w = __builtin_amdgcn_rcpf(wf);
    pc.inv_h = __builtin_amdgcn_rcpf(hf);
    pc.asp = wf * pc.inv_h;
    pc.x1 = (log_r(pc.r) - kLogRLo) * (2.0f / (kLogRHi - kLogRLo)) - 1.0f;
    pc.rbin = r_bin(pc.r);
    pc.pad_[0] = pc.pad_[1] = pc.pad_[2] = 0.0f;
    return pc;
}

// Camera k's record from a table written by k_tab_range (read-only during the launch and indexed
// wave-uniformly: scalar loads).
__device__ __forceinline__ SurCam load_sur_cam(const float* pre, uint32_t k) {
    using ConstU32 = const __attribute__((address_space(4))) uint32_t;
    SurCam pc;
    ConstU32* src = (ConstU32*)(pre + 16 * (size_t)k);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&pc);
#pragma unroll
    for (int i = 0; i < 16; ++i) dst[i] = src[i];
    return pc;
}

// Geometry of the camera ray of pixel (col, row) for the surrogate: the view direction as
// ray_dir (blackhole.frag:68-72) and the orbit-plane frame, impact parameter and redshift inputs
// of lawS_setup (geodesic.glsl:27-40), from hardware reciprocal / square-root approximations
// (relative error ~1e-7). The tangential part of the initial coordinate velocity is the view
// direction's, so the frame needs no velocity; |e1 x d| = |d - (d.e1) e1| is sin(alpha).
__device__ __forceinline__ void sur_geom_dir(const MarchConst& c, const SurCam& pc,
                                             const V3<float>& d, RayPrep& r);

__device__ __forceinline__ void sur_geom(const MarchConst& c, const SurCam& pc, int64_t pix,
                                         uint32_t col, uint32_t row, RayPrep& r) {
    float px, py;
    if (c.jitter_hashed) {
        const uint32_t pi2 = 2u * (uint32_t)pix;
        px = (float)col + (float)(pcg(c.jitter_base + pi2) >> 8) * 0x1p-24f;
        py = (float)row + (float)(pcg(c.jitter_base + pi2 + 1u) >> 8) * 0x1p-24f;
    } else {
        px = ((float)col + 0.5f) + c.jitter[0];
        py = ((float)row + 0.5f) + c.jitter[1];
    }
    const float ndx = __builtin_fmaf(px * pc.inv_w, 2.0f, -1.0f);
    const float ndy = __builtin_fmaf(py * pc.inv_h, 2.0f, -1.0f);
    const V3<float> d0{(ndx * c.tan_half_fov) * pc.asp, ndy * c.tan_half_fov, -1.0f};
    const float* B = c.basis;
    V3<float> d{__builtin_fmaf(B[6], d0.z, __builtin_fmaf(B[3], d0.y, B[0] * d0.x)),
                __builtin_fmaf(B[7], d0.z, __builtin_fmaf(B[4], d0.y, B[1] * d0.x)),
                __builtin_fmaf(B[8], d0.z, __builtin_fmaf(B[5], d0.y, B[2] * d0.x))};
    const float id = __builtin_amdgcn_rsqf(dot3(d, d));
    d.x *= id; d.y *= id; d.z *= id;
    sur_geom_dir(c, pc, d, r);
}

// The surrogate geometry of a ray of unit view direction d (fast setup).
__device__ __forceinline__ void sur_geom_dir(const MarchConst& c, const SurCam& pc,
                                             const V3<float>& d, RayPrep& r) {
    RayGeom& g = r.g;
    const V3<float> P{c.pos[0], c.pos[1], c.pos[2]};
    g.r = pc.r;
    g.e1 = V3<float>{pc.e1[0], pc.e1[1], pc.e1[2]};
    const float nr = dot3(d, g.e1);
    const V3<float> t{d.x - nr * g.e1.x, d.y - nr * g.e1.y, d.z - nr * g.e1.z};
    const float tn = __builtin_amdgcn_sqrtf(dot3(t, t));
    g.planar = !(tn > 1e-6f) || fabsf(g.e1.y) < 1e-6f;
    const float it = tn > 1e-6f ? __builtin_amdgcn_rcpf(tn) : 0.0f;
    g.e2 = V3<float>{t.x * it, t.y * it, t.z * it};
    g.ingoing = nr < 0.0f;
    g.sin_a = tn;
    g.b = g.r * tn * pc.inv_f;
    g.bm = g.b * (1.0f / kBcritF) - 1.0f;
    float psi = atan2_fast(-g.e1.y, g.e2.y);
    if (psi < 0.0f) psi += kPi;
    if (psi >= kPi) psi -= kPi;
    g.psi = psi;
    r.b_axis = -(P.z * d.x - P.x * d.z) * pc.bax_k;
    r.lapse_g = pc.lapse_g;
}

// Rays with |beta - 1| below this take the integrator's own view direction (ray_dir, correctly
// rounded) before the fast geometry: there the swept angle moves by ~|db / b| / |beta - 1| rad, and
// the approximate direction's few 1e-7 would approach 1e-4 rad.
constexpr float kExactBeta = 0.05f;

// Classification of a ray; identical in the pre-pass and the network pass (same operations). A
// ray of the FP64 band that the model serves is a network candidate; the other band rays are
// marched in FP64. Band membership must agree with the integrator's exact setup: rays within
// 1e-5 b_crit of the band edge are classified by it.
__device__ __forceinline__ RayPrep prep_ray(const MarchConst& c, const SurCam& pc,
                                            const SurConst& k, const WaveItem& w) {
    const int64_t pix = w.pix;
    RayPrep r;
    r.x[0] = r.x[1] = r.x[2] = 0.0f;
    r.route = kNone;
    r.band = false;
    r.margin = k.radius_margin[pc.rbin];
    if (pix < 0) return r;
    sur_geom(c, pc, pix, w.col, w.row, r);
    const bool band64 = c.precision == BHS_PREC_F32_BAND64;
    const bool edge = band64 && fabsf(fabsf(r.g.b - kBcritF) - c.band_abs) < 1e-5f * kBcritF;
    if (edge || fabsf(r.g.bm) < kExactBeta) {
        const V3<float> d = ray_dir_at(c, pix, (int32_t)w.col, (int32_t)w.row);
        if (edge) {
            const V3<float> P{c.pos[0], c.pos[1], c.pos[2]};
            const Setup<float> s = lawS_setup(P, d);
            r.g = ray_geom(P, d, s);
            r.b_axis = s.b_axis;
            r.lapse_g = s.lapse_g;
        } else {
            sur_geom_dir(c, pc, d, r);
        }
    }
    r.band = band64 && __builtin_fabsf(r.g.b - kBcritF) < c.band_abs;
    r.route = r.band ? kBand : kEarly;
    if (in_domain(r.g)) {
        r.x[0] = pc.x1;   // the features of features(), x1 from the camera record
        r.x[1] = x2_of(r.g.bm);
        r.x[2] = r.g.psi * (2.0f / kPi) - 1.0f;
        if (cell_served(k, r.x) && steps_ok(c, k, pc.rbin, r.x[1])) r.route = kNet;
    } else if (plunging(r.g) && steps_ok(c, k, pc.rbin, x2_of(r.g.bm))) {
        r.route = kPlunge;
    }
    return r;
}

// Lists of the rays a surrogate pass does not serve, by class (4 bits each, kNoList: not listed
// by this pass): FP64 band rays and early fallback rays (known before the decode), band rays and
// other rays the decode sends back.
constexpr uint32_t kNoList = 15u;
constexpr uint32_t pack_dests(uint32_t band, uint32_t early, uint32_t back_band,
                              uint32_t back) {
    return band | (early << 4) | (back_band << 8) | (back << 12);
}

__device__ __forceinline__ int dest_of(const RayPrep& r, bool back, uint32_t dests) {
    int cls = -1;
    if (back)
        cls = r.band ? 2 : 3;
    else if (r.route == kBand)
        cls = 0;
    else if (r.route == kEarly)
        cls = 1;
    if (cls < 0) return kToNone;
    const uint32_t d = (dests >> (4 * cls)) & 15u;
    return d == kNoList ? kToNone : (int)d;
}

// Pre-pass: lists the FP64 band rays and the early fallback rays.
template <typename Cams>
__global__ __launch_bounds__(kBlock) void k_sur_pre(const Cams cams, const SurConst k,
                                                    uint32_t n_items, Scratch sc, SurLists fl,
                                                    uint32_t dests, const float* cam_pre) {
    for (uint32_t q0 = blockIdx.x * kBlock; q0 < n_items; q0 += gridDim.x * kBlock) {
        const WaveItem w = wave_item(cams, q0 + (threadIdx.x & ~63u), n_items);
        const MarchConst& c = cams.cam(w.ci);
        const SurCam pc = cam_pre ? load_sur_cam(cam_pre, w.ci) : sur_cam(c);
        const RayPrep r = prep_ray(c, pc, k, w);
        push_lists(sc, fl, dest_of(r, false, dests), w.item);
    }
}

// Surrogate render, per-ray network. Persistent blocks (two per CU: the weights take 71 KB of LDS)
// walk 256-item chunks; the wave runs the network if any of its lanes is a candidate, each
// candidate lane decodes and writes its record or lists its ray for the integrator.
template <typename Cams>
__global__ __launch_bounds__(kBlock, 2) void k_surrogate(const Cams cams, const SurConst k,
                                                         const uint8_t* __restrict__ packed,
                                                         const SoA out, Scratch sc, SurLists fl,
                                                         uint32_t n_items, uint32_t dests) {
    __shared__ __attribute__((aligned(16))) uint8_t s_w[kPackedBytes];
    __shared__ float s_y[4][64 * kOut];
    __shared__ uint32_t s_st[4][4];
    stage_weights(packed, s_w);
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    uint32_t n_served = 0, n_cross = 0, n_esc = 0, n_fb = 0;
    const uint32_t chunks = (n_items + kBlock - 1) / kBlock;
    ScanCursor cur;
    for (uint32_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
        const WaveItem w = wave_item_scan(cams, ch * kBlock + (threadIdx.x & ~63u), n_items, cur);
        const MarchConst& c = cams.cam(w.ci);
        const int64_t pix = w.pix;
        const RayPrep r = prep_ray(c, sur_cam(c), k, w);
        const bool cand = r.route == kNet;
        float y[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (__any(cand)) mlp_wave(s_w, r.x, s_y[wave], y);
        bool back = false;
        if (cand || r.route == kPlunge) {
            const Dec o = cand ? decode(c, k, r.b_axis, r.lapse_g, r.g, y) : decode_plunge(c, r.g);
            if (o.serve) {
                store_dec(out, c.ray_offset + pix, o, r.g.b);
                n_served += 1;
                n_cross += o.n_cross;
                n_esc += o.escaping ? 1u : 0u;
            } else {
                back = true;
            }
        }
        push_lists(sc, fl, dest_of(r, back, dests), w.item);
        n_fb += (back || r.route == kEarly) ? 1u : 0u;
    }
    add_block_stats(sc.stats, n_served, n_cross, n_esc, n_fb, s_st);
}

// ------------------------------------------------------------------------------------------------
// Table mode: per camera, everything of the decode that depends on (x2, x3) only (x1 is the
// camera's): the network's escape residual, the first two crossings' u (exact orbit with the
// learned residual applied) and the third crossing's (exact orbit) on an n2 x n3 grid,
// interpolated bilinearly, and the exact swept angle at escape / capture (a function of x2 alone,
// with a logarithmic rise towards the band) on a kPhiNodes-point grid, interpolated linearly. A ray
// then costs its setup, two lookups and the output trigonometry instead of the orbit quadratures
// and Newton solves. x3 spans [-1, 1]; x2 spans [x2(beta = 0.3), x2(beta_max)] with beta_max from
// the image corners (the largest view angle from the inward radial over a pinhole frustum is at a
// corner), both widened by 1e-3. Node (i, j) of camera k is node[((k n3 + j) n2 + i) * 4 + o];
// range[2k], range[2k + 1] hold the x2 bounds.

__device__ __forceinline__ bool same_pose(const MarchConst& a, const MarchConst& b) {
    bool eq = a.width == b.width && a.height == b.height && a.tan_half_fov == b.tan_half_fov;
#pragma unroll
    for (int i = 0; i < 3; ++i) eq = eq && a.pos[i] == b.pos[i];
#pragma unroll
    for (int i = 0; i < 9; ++i) eq = eq && a.basis[i] == b.basis[i];
    return eq;
}

template <typename Cams>
__global__ void k_tab_range(const Cams cams, uint32_t ncam, float* __restrict__ range,
                            uint32_t* __restrict__ slot, float* __restrict__ cam_pre) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncam) return;
    const MarchConst& c = cams.cam(k);
    const SurCam pc = sur_cam(c);
    const uint32_t* pw = reinterpret_cast<const uint32_t*>(&pc);
    for (int i = 0; i < 16; ++i) reinterpret_cast<uint32_t*>(cam_pre)[16 * k + i] = pw[i];
    uint32_t first = k;
    for (uint32_t j = 0; j < k; ++j)
        if (same_pose(cams.cam(j), c)) {
            first = j;
            break;
        }
    slot[k] = first;
    const float r = sqrtf(c.pos[0] * c.pos[0] + c.pos[1] * c.pos[1] + c.pos[2] * c.pos[2]);
    const float ir = 1.0f / r;
    const float aspect = (float)c.width / (float)c.height;
    float smax = 0.0f;
    for (int corner = 0; corner < 4; ++corner) {
        const float cx = (corner & 1 ? 1.0f : -1.0f) * c.tan_half_fov * aspect;
        const float cy = (corner & 2 ? 1.0f : -1.0f) * c.tan_half_fov;
        // d = B (cx, cy, -1) / |.|, B columns right, up, back
        const float dx = c.basis[0] * cx + c.basis[3] * cy - c.basis[6];
        const float dy = c.basis[1] * cx + c.basis[4] * cy - c.basis[7];
        const float dz = c.basis[2] * cx + c.basis[5] * cy - c.basis[8];
        const float dn = sqrtf(dx * dx + dy * dy + dz * dz);
        const float ca = -(dx * c.pos[0] + dy * c.pos[1] + dz * c.pos[2]) * ir / dn;
        const float sa = ca <= 0.0f ? 1.0f : sqrtf(fmaxf(0.0f, 1.0f - ca * ca));
        smax = fmaxf(smax, sa);
    }
    smax = fminf(smax, kSinMax);
    const float f = sqrtf(fmaxf(1.0f - 2.0f * ir, 1e-12f));
    const float bm_max = r * smax / f * (1.0f / kBcritF) * 1.0001f - 1.0f;
    range[2 * k] = x2_of(kBetaMin - 1.0f) - 1e-3f;
    range[2 * k + 1] = fmaxf(x2_of(bm_max), 0.0f) + 1e-3f;
}

struct TabGeom {
    float x2_lo, inv_h2, h2, inv_h3, h3;
};

// Node spacing of a camera's grid (approximate reciprocals: the nodes and the lookups use the
// same values).
__device__ __forceinline__ TabGeom tab_geom(const float* range, int32_t n2, int32_t n3) {
    TabGeom t;
    t.x2_lo = range[0];
    t.h2 = (range[1] - range[0]) * __builtin_amdgcn_rcpf((float)(n2 - 1));
    t.inv_h2 = (float)(n2 - 1) * __builtin_amdgcn_rcpf(range[1] - range[0]);
    t.h3 = 2.0f * __builtin_amdgcn_rcpf((float)(n3 - 1));
    t.inv_h3 = 0.5f * (float)(n3 - 1);
    return t;
}

// Orbit of the exact Schwarzschild prior at feature x2 for a camera at radius r (the geometry
// fields orbit() reads).
__device__ __forceinline__ Orbit node_orbit(float r, float x2) {
    RayGeom g;
    g.r = r;
    g.bm = bm_of_x2(x2);
    g.b = (1.0f + g.bm) * kBcritF;
    return orbit(g);
}

__device__ __forceinline__ float finite_or_zero(float v) {
    return __builtin_isfinite(v) ? v : 0.0f;
}

__device__ __forceinline__ float cam_radius(const MarchConst& c) {
    return sqrtf(c.pos[0] * c.pos[0] + c.pos[1] * c.pos[1] + c.pos[2] * c.pos[2]);
}

// Table nodes of every camera (blocks walk (camera, chunk) pairs): at node (i, j) = (x2_i, x3_j),
// [escape residual y0, u of crossings 1 and 2 (exact orbit times exp(1e-3 y1|y2)), u of crossing 3
// (exact orbit)].
template <typename Cams>
__global__ __launch_bounds__(kBlock, 2) void k_table(const Cams cams, uint32_t ncam,
                                                     const uint8_t* __restrict__ packed,
                                                     const float* __restrict__ range,
                                                     const uint32_t* __restrict__ slot,
                                                     int32_t n2, int32_t n3,
                                                     float* __restrict__ node) {
    __shared__ __attribute__((aligned(16))) uint8_t s_w[kPackedBytes];
    __shared__ float s_y[4][64 * kOut];
    stage_weights(packed, s_w);
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    const int64_t n = (int64_t)n2 * n3;
    const int64_t per = (n + kBlock - 1) / kBlock;
    for (int64_t p = blockIdx.x; p < per * ncam; p += gridDim.x) {
        const uint32_t k = (uint32_t)(p / per);
        if (slot[k] != k) continue;   // block-uniform: a camera of an earlier camera's pose
        const int64_t q = (p % per) * kBlock + threadIdx.x;
        const TabGeom t = tab_geom(range + 2 * k, n2, n3);
        const float r = cam_radius(cams.cam(k));
        const float x1 = (log_r(r) - kLogRLo) * (2.0f / (kLogRHi - kLogRLo)) - 1.0f;
        const int32_t i = (int32_t)(q % n2), j = (int32_t)(q / n2);
        const float x[3] = {x1, __builtin_fmaf((float)i, t.h2, t.x2_lo),
                            __builtin_fmaf((float)j, t.h3, -1.0f)};
        float y[5];
        mlp_wave(s_w, x, s_y[wave], y);
        if (q < n) {
            const Orbit ob = node_orbit(r, x[1]);
            const float psi = (x[2] + 1.0f) * (0.5f * kPi);
            f4 v;
            v[0] = y[0];
            v[1] = finite_or_zero(orbit_u(ob, psi) * __expf(kURes * y[1]));
            v[2] = finite_or_zero(orbit_u(ob, psi + kPi) * __expf(kURes * y[2]));
            v[3] = finite_or_zero(orbit_u(ob, psi + 2.0f * kPi));
            *(f4*)(node + ((int64_t)k * n + q) * 4) = v;
        }
    }
}

// Exact swept angle at escape / capture over the camera's x2 range (kPhiNodes points).
template <typename Cams>
__global__ __launch_bounds__(kBlock) void k_tab_phi(const Cams cams, uint32_t ncam,
                                                    const float* __restrict__ range,
                                                    const uint32_t* __restrict__ slot,
                                                    float* __restrict__ phi) {
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= (int64_t)ncam * kPhiNodes) return;
    const uint32_t k = (uint32_t)(q / kPhiNodes);
    if (slot[k] != k) return;
    const int32_t i = (int32_t)(q % kPhiNodes);
    const float lo = range[2 * k], hi = range[2 * k + 1];
    const float x2 = __builtin_fmaf((float)i, (hi - lo) / (float)(kPhiNodes - 1), lo);
    phi[q] = finite_or_zero(node_orbit(cam_radius(cams.cam(k)), x2).phi_end);
}

// Bilinear lookup of a node grid at (x2, x3): y0, u of crossings 1, 2, 3.
__device__ __forceinline__ void tab_lookup(const float* __restrict__ node, const TabGeom& t,
                                           int32_t n2, int32_t n3, const float x[3],
                                           float v[4]) {
    float fx = (x[1] - t.x2_lo) * t.inv_h2, fy = (x[2] + 1.0f) * t.inv_h3;
    fx = fminf(fmaxf(fx, 0.0f), (float)(n2 - 1) - 1e-3f);
    fy = fminf(fmaxf(fy, 0.0f), (float)(n3 - 1) - 1e-3f);
    const int32_t i = (int32_t)fx, j = (int32_t)fy;
    const float a = fx - (float)i, b = fy - (float)j;
    const float* p0 = node + ((int64_t)j * n2 + i) * 4;
    const float* p1 = p0 + (int64_t)n2 * 4;
    const f4 q00 = *(const f4*)p0, q01 = *(const f4*)(p0 + 4);
    const f4 q10 = *(const f4*)p1, q11 = *(const f4*)(p1 + 4);
    const float w00 = (1.0f - a) * (1.0f - b), w01 = a * (1.0f - b), w10 = (1.0f - a) * b,
                w11 = a * b;
#pragma unroll
    for (int o = 0; o < 4; ++o)
        v[o] = w00 * q00[o] + w01 * q01[o] + w10 * q10[o] + w11 * q11[o];
}

// Linear lookup of the swept-angle table at x2.
__device__ __forceinline__ float phi_lookup(const float* __restrict__ phi, const float* range,
                                            float x2) {
    const float lo = range[0], hi = range[1];
    float f = (x2 - lo) * ((float)(kPhiNodes - 1) * __builtin_amdgcn_rcpf(hi - lo));
    f = fminf(fmaxf(f, 0.0f), (float)(kPhiNodes - 1) - 1e-3f);
    const int32_t i = (int32_t)f;
    const float a = f - (float)i;
    return __builtin_fmaf(a, phi[i + 1] - phi[i], phi[i]);
}

// Near-critical chunks: a 256-item chunk whose rays, sampled at nine pixel centres per 64 items
// (jitter ignored), include an inward ray with |beta - 1| < kNearBeta. They hold the long
// integrator rays of a frame (the FP64 band, rays that orbit); large table passes take them first
// so that their fallback marches run beside the rest of the pass. A scheduling class only.
constexpr float kNearBeta = 0.8f;

__device__ __forceinline__ bool near_segment(const MarchConst& c, uint32_t p0) {
    const uint32_t width = (uint32_t)c.width, height = (uint32_t)c.height;
    const V3<float> P{c.pos[0], c.pos[1], c.pos[2]};
    const float r = __builtin_amdgcn_sqrtf(dot3(P, P));
    if (!(r > 2.5f)) return false;
    const float ir = __builtin_amdgcn_rcpf(r);
    const V3<float> e1{P.x * ir, P.y * ir, P.z * ir};
    const float kb = r * __builtin_amdgcn_rsqf(1.0f - 2.0f * ir) * (1.0f / kBcritF);
    const float ih = __builtin_amdgcn_rcpf((float)c.height);
    const float sc = 2.0f * c.tan_half_fov * ih;
    const float x0 = c.tan_half_fov * (float)c.width * ih;
    const float* B = c.basis;
    bool hit = false;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const uint32_t p = p0 + (k < 8 ? 8u * k : 63u);
        const uint32_t row = p / width, col = p - row * width;
        if (row >= height) continue;
        const V3<float> d0{((float)col + 0.5f) * sc - x0,
                           ((float)row + 0.5f) * sc - c.tan_half_fov, -1.0f};
        const V3<float> d{B[0] * d0.x + B[3] * d0.y + B[6] * d0.z,
                          B[1] * d0.x + B[4] * d0.y + B[7] * d0.z,
                          B[2] * d0.x + B[5] * d0.y + B[8] * d0.z};
        const float nr = dot3(d, e1) * __builtin_amdgcn_rsqf(dot3(d, d));
        const float beta = kb * __builtin_amdgcn_sqrtf(fmaxf(1.0f - nr * nr, 0.0f));
        hit = hit || (nr < 0.0f && fabsf(beta - 1.0f) < kNearBeta);
    }
    return hit;
}

// Near-critical flag of each 256-item chunk of a call (one thread per 64-item segment).
template <typename Cams>
__global__ __launch_bounds__(kBlock) void k_chunk_class(const Cams cams, uint32_t n_items,
                                                        uint32_t* __restrict__ near) {
    const uint32_t seg = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t n_seg = n_items >> 6;
    bool hit = false;
    if (seg < n_seg) {
        const uint32_t q = seg << 6;
        const uint32_t ci = cams.find(q);
        hit = near_segment(cams.cam(ci), q - cams.base(ci));
    }
    const uint64_t m = __ballot(hit);
    const uint32_t lane = threadIdx.x & 63u;
    if ((lane & 3u) == 0u && seg < n_seg) near[seg >> 2] = ((m >> lane) & 0xFull) != 0ull;
}

// Chunks of kBlock items per block of the table pass (cpb): its blocks are short-lived, and the
// hardware dispatcher hands them out as slots free up, so the pass ends evenly although chunks
// near the critical curve cost several times the others (a persistent grid-stride share does not).
// cpb is the call's chunk count over four times the resident blocks, clamped to [kTabChunksMin,
// kTabChunksMax] (fewer, longer blocks cost less per chunk; a single frame needs enough blocks
// to fill the chip). The block statistics go to per-block partials: one atomic per block and
// counter would serialise.
constexpr uint32_t kTabChunksMin = 4, kTabChunksMax = 16;

// Surrogate render in table mode: as k_surrogate with the network replaced by the table lookup.
// Block b takes chunks [b cpb, (b + 1) cpb); phase 0 every chunk, phase 1 / 2 only the
// near-critical / other ones (flags `near`). dests: the lists of the unserved rays; cam_pre: the
// per-camera setup terms; part: the blocks' statistics partials. FULL: every output plane is
// present (the usual call): the stores carry no per-plane tests, whose hoisted masks would
// otherwise crowd the scalar registers into lane spills.
template <typename Cams, bool FULL>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6)))
void k_surrogate_tab(const Cams cams, const SurConst k, const float* __restrict__ node,
                     const float* __restrict__ phi, const float* __restrict__ range,
                     const uint32_t* __restrict__ slot, int32_t n2, int32_t n3, const SoA out,
                     Scratch sc, SurLists fl, uint32_t n_items, uint32_t dests,
                     uint32_t* __restrict__ part, uint32_t cpb, uint32_t phase,
                     const uint32_t* __restrict__ near, const float* __restrict__ cam_pre) {
    __shared__ uint32_t s_st[4][4];
    uint32_t n_served = 0, n_cross = 0, n_esc = 0, n_fb = 0;
    const int64_t nodes = (int64_t)n2 * n3 * 4;
    // the block's cpb chunks; statistics to the block's partials
    const uint64_t q_first = (uint64_t)blockIdx.x * cpb * kBlock;
    const uint64_t q_last = q_first + (uint64_t)cpb * kBlock;
    const uint32_t q_end = (uint32_t)(q_last < n_items ? q_last : n_items);
    ScanCursor cur;
    for (uint32_t q0 = (uint32_t)q_first; q0 < q_end; q0 += kBlock) {
        if (phase != 0u && (near[q0 / kBlock] != 0u) != (phase == 1u)) continue;   // uniform
        const WaveItem w = wave_item_scan(cams, q0 + (threadIdx.x & ~63u), n_items, cur);
        const uint32_t ci = w.ci;
        const MarchConst& c = cams.cam(ci);
        const int64_t pix = w.pix;
        const RayPrep r = prep_ray(c, load_sur_cam(cam_pre, ci), k, w);
        bool back = false;
        if (r.route == kNet) {
            const uint32_t ts = slot[ci];
            float v[4];
            tab_lookup(node + ts * nodes, tab_geom(range + 2 * ts, n2, n3), n2, n3, r.x, v);
            const float Phi = phi_lookup(phi + (int64_t)ts * kPhiNodes, range + 2 * ts, r.x[1]) +
                              (r.g.bm > 0.0f ? kResScale * v[0] : 0.0f);
            const Dec o = decode_with(c, r.margin, r.b_axis, r.lapse_g, r.g, Phi,
                                      [&](int q) { return v[1 + q]; });
            if (o.serve) {
                store_dec<FULL>(out, c.ray_offset + pix, o, r.g.b);
                n_served += 1;
                n_cross += o.n_cross;
                n_esc += o.escaping ? 1u : 0u;
            } else {
                back = true;
            }
        } else if (r.route == kPlunge) {
            const Dec o = decode_plunge(c, r.g);
            if (o.serve) {
                store_dec<FULL>(out, c.ray_offset + pix, o, r.g.b);
                n_served += 1;
            } else {
                back = true;
            }
        }
        push_lists(sc, fl, dest_of(r, back, dests), w.item);
        n_fb += (back || r.route == kEarly) ? 1u : 0u;
    }
    store_block_stats(part, n_served, n_cross, n_esc, n_fb, s_st);
}

// Network outputs of explicit feature rows.
__global__ __launch_bounds__(kBlock, 2) void k_forward(const float* __restrict__ X, int64_t n,
                                                       const uint8_t* __restrict__ packed,
                                                       float* __restrict__ Y) {
    __shared__ __attribute__((aligned(16))) uint8_t s_w[kPackedBytes];
    __shared__ float s_y[4][64 * kOut];
    stage_weights(packed, s_w);
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    const int64_t chunks = (n + kBlock - 1) / kBlock;
    for (int64_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
        const int64_t i = ch * kBlock + threadIdx.x;
        float x[3] = {0.0f, 0.0f, 0.0f};
        if (i < n) {
            x[0] = X[3 * i];
            x[1] = X[3 * i + 1];
            x[2] = X[3 * i + 2];
        }
        float y[5];
        mlp_wave(s_w, x, s_y[wave], y);
        if (i < n)
            for (int o = 0; o < kOut; ++o) Y[kOut * i + o] = y[o];
    }
}

// Validation of the decode against integrator outputs of the same rays (FP64 march, refined
// crossings, crossings recorded at any radius). Per ray: err_phi = angle between the escape
// directions (0 for served captured rays), err_r = largest relative radius error over the
// equatorial crossings both report within [inner / 2, 1.5 outer] (-1 when none), 1e30 in both on a
// disagreement of the rendered record (class, escape status or in-annulus crossing count), and the
// error-map cell of the ray (-1 when the decode does not serve it).
__global__ __launch_bounds__(kBlock, 2) void k_validate(int64_t n, const float* __restrict__ pos3,
                                                        const float* __restrict__ dir3,
                                                        const MarchConst c, const SurConst k,
                                                        const SoA ref,
                                                        const uint8_t* __restrict__ packed,
                                                        float* __restrict__ err_phi,
                                                        float* __restrict__ err_r,
                                                        int32_t* __restrict__ cell) {
    __shared__ __attribute__((aligned(16))) uint8_t s_w[kPackedBytes];
    __shared__ float s_y[4][64 * kOut];
    stage_weights(packed, s_w);
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    const int64_t chunks = (n + kBlock - 1) / kBlock;
    for (int64_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
        const int64_t i = ch * kBlock + threadIdx.x;
        Setup<float> s;
        RayGeom g;
        float x[3] = {0.0f, 0.0f, 0.0f};
        bool elig = false;
        if (i < n) {
            const V3<float> P{pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2]};
            const V3<float> d{dir3[3 * i], dir3[3 * i + 1], dir3[3 * i + 2]};
            s = lawS_setup(P, d);
            g = ray_geom(P, d, s);
            if (in_domain(g)) {
                features(g, x);
                elig = true;
            }
        }
        float y[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (__any(elig)) mlp_wave(s_w, x, s_y[wave], y);
        if (i >= n) continue;
        float ep = -1.0f, er = -1.0f;
        int32_t ci = -1;
        if (elig) {
            const Dec o = decode(c, k, s.b_axis, s.lapse_g, g, y);
            if (o.serve) {
                const uint32_t fl = ref.flags[i], nc = ref.n_cross[i];
                const bool ref_esc = (fl & BHS_FLAG_ESCAPED) != 0;
                uint32_t ref_in = 0;   // in-annulus crossings among the reference's first two
                const bool ref_more = nc >= 2;   // a third equatorial crossing is not recorded
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    if ((uint32_t)q >= nc) continue;
                    const float r = ref.disk[i + q * ref.n];
                    if (r >= c.disk_inner && r <= c.disk_outer) ++ref_in;
                }
                bool bad = (fl & BHS_FLAG_MAXSTEPS) != 0 || ref_esc != o.escaping;
                if (c.thin) {
                    bad = bad || (ref_in != o.n_cross && !(ref_more && ref_in < o.n_cross));
                    const uint32_t m = nc < o.n_eq ? nc : o.n_eq;
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        if ((uint32_t)q >= m) continue;
                        const float r = ref.disk[i + q * ref.n];
                        if (r < 0.5f * c.disk_inner || r > 1.5f * c.disk_outer) continue;
                        const float e = fabsf(o.r_eq[q] - r) / r;
                        er = e > er ? e : er;
                    }
                }
                ep = 0.0f;
                if (!bad && o.escaping) {
                    const float ex = ref.esc[i], ey = ref.esc[i + ref.n], ez = ref.esc[i + 2 * ref.n];
                    const float cx = o.esc[1] * ez - o.esc[2] * ey;
                    const float cy = o.esc[2] * ex - o.esc[0] * ez;
                    const float cz = o.esc[0] * ey - o.esc[1] * ex;
                    const float dt = o.esc[0] * ex + o.esc[1] * ey + o.esc[2] * ez;
                    ep = atan2f(sqrtf(cx * cx + cy * cy + cz * cz), dt);
                }
                if (bad) ep = er = 1e30f;
                int ir = (int)((x[0] + 1.0f) * (0.5f * kMapR));
                int ib = (int)((x[1] + 1.0f) * (0.5f * kMapB));
                ir = ir < 0 ? 0 : (ir >= kMapR ? kMapR - 1 : ir);
                ib = ib < 0 ? 0 : (ib >= kMapB ? kMapB - 1 : ib);
                ci = ir * kMapB + ib;
            }
        }
        err_phi[i] = ep;
        err_r[i] = er;
        cell[i] = ci;
    }
}

int persistent_blocks(int64_t chunks, int num_cus) {
    const int64_t cap = 2 * (int64_t)(num_cus > 0 ? num_cus : 256);
    return (int)(chunks < cap ? (chunks > 0 ? chunks : 1) : cap);
}

}  // namespace

hipError_t launch_pack(const float* P, void* packed, hipStream_t s) {
    const int n = 2 * 16384 + 2048 + kH;
    hipLaunchKernelGGL(k_pack, dim3((n + 255) / 256), dim3(256), 0, s, P, (uint8_t*)packed);
    return hipGetLastError();
}

hipError_t launch_forward(const float* X, int64_t n, const void* packed, float* Y, hipStream_t s) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int blocks = persistent_blocks((n + kBlock - 1) / kBlock, cus);
    hipLaunchKernelGGL(k_forward, dim3(blocks), dim3(kBlock), 0, s, X, n, (const uint8_t*)packed, Y);
    return hipGetLastError();
}

hipError_t launch_validate(int64_t n, const float* pos3, const float* dir3, const MarchConst& c,
                           const SurConst& k, const SoA& ref, const void* packed, float* err_phi,
                           float* err_r, int32_t* cell, int num_cus, hipStream_t s) {
    const int blocks = persistent_blocks((n + kBlock - 1) / kBlock, num_cus);
    hipLaunchKernelGGL(k_validate, dim3(blocks), dim3(kBlock), 0, s, n, pos3, dir3, c, k, ref,
                       (const uint8_t*)packed, err_phi, err_r, cell);
    return hipGetLastError();
}

namespace {

constexpr uint32_t kRoutedItems = 1u << 24;
// Table mode: the two-phase pass wins from about four 1080p frames per call; the pre-pass up to
// two (measured crossover).
constexpr uint32_t kRoutedTableItems = 1u << 22;

// Blocks of kBlock threads of a kernel that are resident together on one CU. Grid-stride kernels
// launch that many per CU: a larger grid would leave its last blocks to run their whole share
// afterwards, at a fraction of the occupancy.
uint32_t blocks_per_cu(const void* kernel) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    return (uint32_t)per_cu;
}

// The surrogate render of one camera (Cams = SurOneCam) or of a batch (SurCamTable, table and
// item_base on the device).
template <typename Cams>
hipError_t launch_sur(const Cams& cams, uint32_t ncam, uint32_t n_items, const MarchConst& c0,
                      const MarchConst* table, const uint32_t* item_base, bool any_band,
                      const SurConst& k, const void* packed, const SurTable* tab, const SoA& out,
                      const Scratch& sc, const SurLists& fl, const Tuning& t, int num_cus,
                      hipStream_t s, const SurStreams& ss, hipEvent_t ev_begin, hipEvent_t ev_mid,
                      hipEvent_t ev_end) {
    constexpr bool kOne = std::is_same<Cams, SurOneCam>::value;
    const uint32_t gb0 = (n_items + kBlock - 1) / kBlock;
    static const uint32_t pre_per_cu =
        blocks_per_cu(reinterpret_cast<const void*>(&k_sur_pre<Cams>));
    static const uint32_t tab_per_cu =
        blocks_per_cu(reinterpret_cast<const void*>(&k_surrogate_tab<Cams, true>));
    auto grid_of = [&](uint32_t per_cu) {
        const uint32_t cap = per_cu * (uint32_t)(num_cus > 0 ? num_cus : 256);
        return dim3(gb0 < cap ? (gb0 ? gb0 : 1u) : cap);
    };
    const bool f64 = kOne && c0.precision == BHS_PREC_F64;
    // The fallback marches' time is set by their longest serial rays, which lie near the critical
    // impact parameter. Latency mode (up to kRoutedItems work items, kRoutedTableItems in table
    // mode): a pre-pass routes the rays first, so the side marches run beside the whole surrogate
    // pass. Larger calls in table mode (phased): the pass takes the near-critical chunks first, then their FP64 band and FP32
    // fallback marches run on the side streams beside the pass over the other chunks, whose
    // fallback rays (short) follow it; the pass's short-lived blocks take the slots the side
    // marches leave. Larger calls in network mode: the network pass routes the rays itself and
    // the lists are marched after it. The FP32 lists run at two resident waves per SIMD (131k
    // lanes), which leaves room for the surrogate pass or the FP64 band.
    const bool routed = !f64 && n_items <= (tab ? kRoutedTableItems : kRoutedItems);
    const bool phased = tab && !f64 && !routed;
    const uint32_t to_band = pack_dests(kToBand, kToBand, kToBand, kToBand);
    const uint32_t d_late = pack_dests(kNoList, kNoList, kToLate, kToFb);   // lists built before
    const uint32_t d_all = pack_dests(kToBand, kToEarly, kToBand, kToFb);
    Tuning tl = t;
    tl.mode = 1;
    tl.block = 256;
    tl.waves_per_simd = 2;
    auto march_list = [&](const uint32_t* ids, const uint32_t* count, const Scratch& scr,
                          hipStream_t st) {
        if constexpr (kOne)
            return launch_list(c0, ids, count, fl.fb_cap, out, scr, tl, num_cus, st);
        else
            return launch_list_batch(table, item_base, ncam, ids, count, fl.fb_cap, out, scr, tl,
                                     num_cus, st);
    };
    auto band = [&]() {
        if constexpr (kOne)
            return launch_band_list(c0, out, sc, s, ss.aux, ss.ev_fork, ss.ev_join);
        else
            return launch_band_list_batch(table, item_base, ncam, n_items, out, sc, s, ss.aux,
                                          ss.ev_fork, ss.ev_join);
    };
    // Side marches: the FP64 band on aux and the early FP32 fallback on aux2, both after the
    // point of `s` where ev_fork2 was recorded. side_band records that point and launches the
    // band; side_early makes aux2 wait for it and launches the list, at once (side) or after the
    // caller has enqueued other work on `s` (the event keeps the point of its recording).
    auto side_band = [&]() -> hipError_t {
        hipError_t e2;
        if ((e2 = hipEventRecord(ss.ev_fork2, s)) != hipSuccess) return e2;
        // the band first: its waves (long FP64 chains) take their slots before the lists'
        if (any_band && (e2 = band()) != hipSuccess) return e2;
        return hipSuccess;
    };
    auto side_early = [&]() -> hipError_t {
        hipError_t e2;
        if ((e2 = hipStreamWaitEvent(ss.aux2, ss.ev_fork2, 0)) != hipSuccess) return e2;
        Scratch se = sc;
        se.work = fl.early_work;
        if ((e2 = march_list(fl.early_ids, fl.early_count, se, ss.aux2)) != hipSuccess) return e2;
        if ((e2 = hipEventRecord(ss.ev_join2, ss.aux2)) != hipSuccess) return e2;
        return hipSuccess;
    };
    auto side = [&]() -> hipError_t {
        hipError_t e2;
        if ((e2 = side_band()) != hipSuccess) return e2;
        return side_early();
    };
    hipError_t e;
    if (ev_begin && (e = hipEventRecord(ev_begin, s)) != hipSuccess) return e;
    if (tab)   // x2 ranges, table slots and the per-camera setup terms
        hipLaunchKernelGGL(k_tab_range<Cams>, dim3((ncam + 63) / 64), dim3(64), 0, s, cams, ncam,
                           tab->range, tab->slot, tab->cam);
    // the swept-angle table and the network's node table of the call's poses (after k_tab_range)
    auto build_table = [&](hipStream_t st) -> hipError_t {
        const int64_t nodes = (int64_t)tab->n2 * tab->n3;
        hipLaunchKernelGGL(k_tab_phi<Cams>,
                           dim3((uint32_t)(((int64_t)ncam * kPhiNodes + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, st, cams, ncam, (const float*)tab->range,
                           (const uint32_t*)tab->slot, tab->phi);
        hipLaunchKernelGGL(k_table<Cams>,
                           dim3(persistent_blocks((nodes + kBlock - 1) / kBlock * ncam, num_cus)),
                           dim3(kBlock), 0, st, cams, ncam, (const uint8_t*)packed,
                           (const float*)tab->range, (const uint32_t*)tab->slot, tab->n2,
                           tab->n3, tab->node);
        return hipGetLastError();
    };
    // Latency mode with a table: the table is built on aux2 beside the pre-pass and the band's
    // launch (it depends on k_tab_range only), and `s` waits for it before the pass; the early
    // list, later on aux2, follows it there.
    // Two-phase mode: the near-critical chunk flags (k_chunk_class, which needs the cameras only)
    // are computed on aux2 beside the table build in the same way.
    const bool table_aside = routed && tab;
    const bool class_aside = phased;
    if (table_aside || class_aside) {
        if ((e = hipEventRecord(ss.ev_tab0, s)) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(ss.aux2, ss.ev_tab0, 0)) != hipSuccess) return e;
        if (table_aside) {
            if ((e = build_table(ss.aux2)) != hipSuccess) return e;
        } else {
            const uint32_t segs = n_items >> 6;
            hipLaunchKernelGGL(k_chunk_class<Cams>, dim3((segs + kBlock - 1) / kBlock),
                               dim3(kBlock), 0, ss.aux2, cams, n_items, fl.near);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        if ((e = hipEventRecord(ss.ev_tab1, ss.aux2)) != hipSuccess) return e;
    }
    if (routed) {
        hipLaunchKernelGGL(k_sur_pre<Cams>, grid_of(pre_per_cu), dim3(kBlock), 0, s, cams, k,
                           n_items, sc, fl, pack_dests(kToBand, kToEarly, kNoList, kNoList),
                           tab ? (const float*)tab->cam : nullptr);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        // the band at once (the call's longest chains); the early list, short, is enqueued after
        // the pass, so that the host's launches of the pass come first
        if ((e = side_band()) != hipSuccess) return e;
    }
    if (tab) {
        if (table_aside) {
            if ((e = hipStreamWaitEvent(s, ss.ev_tab1, 0)) != hipSuccess) return e;
        } else if ((e = build_table(s)) != hipSuccess) {
            return e;
        }
        const uint32_t resident = tab_per_cu * (uint32_t)(num_cus > 0 ? num_cus : 256);
        uint32_t cpb = gb0 / (4u * resident);
        cpb = cpb < kTabChunksMin ? kTabChunksMin : (cpb > kTabChunksMax ? kTabChunksMax : cpb);
        const uint32_t tab_blocks = (gb0 + cpb - 1) / cpb;
        auto pass = [&](uint32_t dests, uint32_t phase) {
            uint32_t* part = fl.part + (phase == 2u ? 4u * tab_blocks : 0u);
            const bool full = out.cls && out.flags && out.n_cross && out.steps && out.esc &&
                              out.disk && out.b;
            auto kern = full ? k_surrogate_tab<Cams, true> : k_surrogate_tab<Cams, false>;
            hipLaunchKernelGGL(kern, dim3(tab_blocks ? tab_blocks : 1u), dim3(kBlock), 0, s, cams,
                               k, (const float*)tab->node, (const float*)tab->phi,
                               (const float*)tab->range, (const uint32_t*)tab->slot, tab->n2,
                               tab->n3, out, sc, fl, n_items, dests, part, cpb, phase,
                               (const uint32_t*)fl.near, (const float*)tab->cam);
            const uint32_t red_blocks = (tab_blocks + 4 * kBlock - 1) / (4 * kBlock);
            hipLaunchKernelGGL(k_stats_reduce, dim3(red_blocks ? red_blocks : 1u), dim3(kBlock),
                               0, s, (const uint32_t*)part, tab_blocks, sc.stats);
            return hipGetLastError();
        };
        if (phased) {
            // near-critical chunks first, their band and fallback marches beside the rest
            if ((e = hipStreamWaitEvent(s, ss.ev_tab1, 0)) != hipSuccess) return e;
            if ((e = pass(pack_dests(kToBand, kToEarly, kToBand, kToEarly), 1u)) != hipSuccess)
                return e;
            if ((e = side()) != hipSuccess) return e;
            if ((e = pass(pack_dests(kToLate, kToFb, kToLate, kToFb), 2u)) != hipSuccess)
                return e;
        } else if ((e = pass(f64 ? to_band : (routed ? d_late : d_all), 0u)) != hipSuccess) {
            return e;
        }
    } else {
        hipLaunchKernelGGL(k_surrogate<Cams>, dim3(persistent_blocks(gb0, num_cus)), dim3(kBlock),
                           0, s, cams, k, (const uint8_t*)packed, out, sc, fl, n_items,
                           f64 ? to_band : (routed ? d_late : d_all));
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ev_mid && (e = hipEventRecord(ev_mid, s)) != hipSuccess) return e;
    if (routed && (e = side_early()) != hipSuccess) return e;
    if ((routed || phased) && any_band) {
        // band rays listed after the band march started (sent back by the decode, or met in the
        // second phase): FP64 march on aux2
        Scratch sb = sc;
        sb.band_ids = fl.lb_ids;
        sb.band_count = fl.lb_count;
        sb.band_work = fl.lb_work;
        sb.band_cap = fl.fb_cap;
        if constexpr (kOne)
            e = launch_band_list(c0, out, sb, s, ss.aux2, ss.ev_fork2, ss.ev_join2);
        else
            e = launch_band_list_batch(table, item_base, ncam, n_items, out, sb, s, ss.aux2,
                                       ss.ev_fork2, ss.ev_join2);
        if (e != hipSuccess) return e;
    }
    if (!routed && !phased && !f64 && (e = side()) != hipSuccess) return e;
    if (f64) {   // every unserved ray is in the band list: FP64 march now
        if ((e = band()) != hipSuccess) return e;
    } else if ((e = march_list(fl.fb_ids, fl.fb_count, sc, s)) != hipSuccess) {
        return e;
    }
    if (ev_end && (e = hipEventRecord(ev_end, s)) != hipSuccess) return e;
    if (f64 || any_band)
        if ((e = hipStreamWaitEvent(s, ss.ev_join, 0)) != hipSuccess) return e;
    if (!f64 && (e = hipStreamWaitEvent(s, ss.ev_join2, 0)) != hipSuccess) return e;
    if constexpr (kOne)
        return launch_root_refine(c0, out, sc, s);
    else
        return launch_root_refine_batch(table, item_base, ncam, n_items, out, sc, s);
}

}  // namespace

hipError_t launch_surrogate(const MarchConst& c, const SurConst& k, const void* packed,
                            const SurTable* tab, const SoA& out, const Scratch& sc,
                            const SurLists& fl, const Tuning& t, int num_cus, hipStream_t s,
                            const SurStreams& ss, hipEvent_t ev_begin, hipEvent_t ev_mid,
                            hipEvent_t ev_end) {
    const SurOneCam one{c};
    return launch_sur(one, 1u, camera_items(c.width, c.height), c, nullptr, nullptr,
                      c.precision == BHS_PREC_F32_BAND64, k, packed, tab, out, sc, fl, t, num_cus,
                      s, ss, ev_begin, ev_mid, ev_end);
}

hipError_t launch_surrogate_batch(const MarchConst* table, const uint32_t* item_base,
                                  uint32_t ncam, uint32_t n_items, bool any_band,
                                  const SurConst& k, const void* packed, const SurTable* tab,
                                  const SoA& out, const Scratch& sc, const SurLists& fl,
                                  const Tuning& t, int num_cus, hipStream_t s,
                                  const SurStreams& ss, hipEvent_t ev_begin, hipEvent_t ev_mid,
                                  hipEvent_t ev_end) {
    const SurCamTable cams{table, item_base, ncam};
    MarchConst none;
    __builtin_memset(&none, 0, sizeof none);
    return launch_sur(cams, ncam, n_items, none, table, item_base, any_band, k, packed, tab, out,
                      sc, fl, t, num_cus, s, ss, ev_begin, ev_mid, ev_end);
}

}  // namespace mlp
}  // namespace bhs

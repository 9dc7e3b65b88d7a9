// volume_device.hpp — the volumetric disk march (shaders/blackhole.frag:142-253 with uDiskMode =
// DISK_VOLUMETRIC): law S's step control with the slab rule of :170-181 (inside the slab the step is
// capped at 6 diskStep), emission and absorption sampled along the cubic Hermite interpolant of
// every RK4 step that starts inside the slab (:217-233), volumetricDiskStep and diskDensity
// (shaders/disk.glsl:81-110), and the opaque stop (:234).
//
// The march runs in the ray's precision R (float, or double for the FP64 band); the sample points
// are formed in R and the shading of a sample (density, noise, temperature, redshift, blackbody,
// emission, absorption) runs in float with the elementary functions of shade_device.hpp. Every
// operation is IEEE with a fixed order, so a CPU restatement reproduces the result bit for bit.
#pragma once

#include "lawS_device.hpp"
#include "shade_device.hpp"

namespace bhs {
namespace shade {

struct VolConst {
    ShadeConst sh;
    float thickness;   // uDiskThickness (scale height H / r)
    float lapse;       // sqrt(max(1 - 2 / rcam, 1e-4)) of the camera, in float
};

// diskDensity (disk.glsl:84-93) at p (|p.xz| = r); emit receives the surface-brightness factor.
__device__ __forceinline__ float disk_density(const VolConst& c, const uint16_t* __restrict__ vol,
                                              float px, float py, float pz, float r, float& emit) {
    const float h0 = c.thickness * r;
    const float H = h0 > 1e-3f ? h0 : 1e-3f;
    const float yh = py / H;
    const float vert = exp_ieee(-(yh * yh));
    const float phi = atan2_ieee(pz, px);
    const float n = disk_noise(c.sh, vol, r, phi, yh);
    const float f = smoothstepf(0.32f, 0.78f, n);
    const float pattern = mixf(1.0f, mixf(0.08f, 1.4f, f), c.sh.noise_strength);
    emit = mixf(1.0f, mixf(0.22f, 1.35f, f * f), c.sh.noise_strength);
    const float puff = 1.0f + (0.6f * smoothstepf(c.sh.outer * 0.4f, c.sh.outer, r)) * (n - 0.5f);
    return (((c.sh.density * radial_profile(c.sh, r)) * vert) * pattern) * (puff > 0.0f ? puff : 0.0f);
}

// volumetricDiskStep (disk.glsl:96-110): emission and absorption over a segment of length ds
// centred at p.
__device__ __forceinline__ void vol_sample(const VolConst& c, const uint16_t* __restrict__ vol,
                                           const float* __restrict__ lut, float px, float py,
                                           float pz, float r, float ds, float b_axis,
                                           Composite& cp) {
    float emit;
    const float dens = disk_density(c, vol, px, py, pz, r, emit);
    if (dens < 1e-4f) return;
    const float T = (c.sh.peak_temp * temp_profile(c.sh, r)) * mixf(1.0f, 0.85f + 0.3f * emit, 0.5f);
    const float gphys = redshift<float>(r, b_axis, c.lapse);
    const float g = mixf(1.0f, gphys, c.sh.doppler);
    float bb[4];
    blackbody(lut, T * g, bb);
    const float inward = pow_ieee(c.sh.inner / (r > c.sh.inner ? r : c.sh.inner), 0.6f);
    const float k = (((bb[3] * c.sh.brightness) * emit) * inward);
    float src[3];
    src[0] = (bb[0] * mixf(1.0f, 1.0f, c.sh.warmth)) * k;
    src[1] = (bb[1] * mixf(1.0f, 0.80f, c.sh.warmth)) * k;
    src[2] = (bb[2] * mixf(1.0f, 0.58f, c.sh.warmth)) * k;
    const float l2 = 1.0f - 2.0f / r;
    const float lapse_em = __builtin_sqrtf(l2 > 1e-4f ? l2 : 1e-4f);
    const float ods = mixf(1.0f, lapse_em / (gphys * c.lapse), c.sh.doppler);
    const float absorb = (dens * c.sh.opacity) * ods;
    const float a = 1.0f - exp_ieee(-absorb * ds);
    cp.add(src, a);
}

// One iteration of the volumetric march. Returns true when the ray is finished: out of steps,
// horizon, escape (escape direction in esc) or opaque (transmittance below 0.004).
template <typename R>
__device__ __forceinline__ bool vol_step(MarchState<R>& st, const Setup<R>& s, const MarchK<R>& k,
                                         const VolConst& c, const uint16_t* __restrict__ vol,
                                         const float* __restrict__ lut, float b_axis,
                                         Composite& cp, float esc[3]) {
    using A = Ar<R>;
    V3<R>& x = st.x;
    V3<R>& v = st.v;
    if (st.i >= k.max_steps) return true;                               // :142, :252
    const R r2 = dot3(x, x);
    const R ir = A::rsqrt(r2);                                          // 1/r
    const R r = r2 * ir;
    if (r < R(2.02)) {                                                  // :144
        st.status = 2;
        return true;
    }
    const R iv = A::rsqrt(dot3(v, v));
    if (dot3(x, v) > R(0) && r > k.r_esc_min) {                         // :151-155
        const R b = s.h * iv;
        const R ir2 = ir * ir;
        const R q = (((R(0.75) * b) * b) * b) * (ir2 * ir2);
        if (q < R(0.01)) {
            const V3<R> e = far_field(x, v, r, iv, s.h2);
            esc[0] = (float)e.x;
            esc[1] = (float)e.y;
            esc[2] = (float)e.z;
            st.status = 1;
            return true;
        }
    }
    R ds = (k.ss * r) * s.far_mul;                                      // :169
    if (ds < R(0.02)) ds = R(0.02);
    bool near = false;
    const R rc2 = A::fma(x.z, x.z, x.x * x.x);                          // :170-181
    if (rc2 > k.lo2 && rc2 < k.hi2) {
        const R rcyl = A::sqrt_loop(rc2);
        const R outside = A::fma(-k.slab_k, rcyl, A::abs(x.y));
        if (outside <= R(0)) {
            near = true;
            const R cap = R(6) * k.dstep;
            if (cap < ds) ds = cap;
        } else {
            const R den = A::fma(A::abs(v.y), iv, k.slab_k);
            R cl = outside / den;
            if (cl < k.dstep) cl = k.dstep;
            if (cl < ds) ds = cl;
        }
    }
    const R dt = ds * iv;                                               // :192
    const V3<R> x0 = x;
    const V3<R> v0 = v;
    rk4(x, v, s.c3, dt, ir);                                            // :194
    st.i = st.i + 1;                                                    // :196
    if (near) {                                                         // :217-233
        const V3<R> d{x.x - x0.x, x.y - x0.y, x.z - x0.z};
        const R seg = A::sqrt(dot3(d, d));
        const R cq = __builtin_ceil(seg / k.dstep);
        const int32_t n = cq > R(1) ? (int32_t)cq : 1;
        const R rn = (R)n;
        const float sub = (float)(seg / rn);
        const V3<R> m0{dt * v0.x, dt * v0.y, dt * v0.z};
        const V3<R> m1{dt * v.x, dt * v.y, dt * v.z};
        for (int32_t j = 0; j < n; ++j) {
            const R t = ((R)j + R(0.5)) / rn;
            const R t2 = t * t, t3 = t2 * t;
            const R h00 = (R(2) * t3 - R(3) * t2) + R(1);
            const R h10 = (t3 - R(2) * t2) + t;
            const R h01 = -R(2) * t3 + R(3) * t2;
            const R h11 = t3 - t2;
            const R px = ((h00 * x0.x + h10 * m0.x) + h01 * x.x) + h11 * m1.x;
            const R py = ((h00 * x0.y + h10 * m0.y) + h01 * x.y) + h11 * m1.y;
            const R pz = ((h00 * x0.z + h10 * m0.z) + h01 * x.z) + h11 * m1.z;
            const R rm = A::sqrt(A::fma(pz, pz, px * px));
            if (rm > k.inner - R(0.5) && rm < k.outer && A::abs(py) < k.slab_k * rm)
                vol_sample(c, vol, lut, (float)px, (float)py, (float)pz, (float)rm, sub, b_axis, cp);
            if (cp.T < kOpaque) break;
        }
    }
    if (cp.T < kOpaque) {                                               // :234
        cp.opaque = true;
        st.status = 3;
        return true;
    }
    return false;
}

}  // namespace shade
}  // namespace bhs

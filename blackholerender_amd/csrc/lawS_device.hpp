// lawS_device.hpp — device implementation of the law-S march (shaders/blackhole.frag:67-253,
// thin-disk variant, lensing on, far mode off, no bodies/gas, disk opacity 0), of the ray generation
// (blackhole.frag:68-72) and of law T (src/physics/Geodesic.h:107-140).
//
// Arithmetic contract (the whole library is compiled with -ffp-contract=off and correctly rounded
// FP32 division/square root): every multiply-add below that is fused is written as an explicit
// fma, division and square root are IEEE, the march loop's reciprocal square root is correctly
// rounded in float (rsqrt_rn) and a fixed refinement of it in double (rsqrt_d), nothing is
// reassociated. The operation order is fixed:
//   dot(a,b) = fma(a.z,b.z, fma(a.y,b.y, a.x*b.x)),  normalize(a) = a * (1/sqrt(dot(a,a))),
//   cross(a,b).x = fma(a.y,b.z, -(a.z*b.y)) (cyclic),
// so one ray's result does not depend on the GPU, the launch shape or the order rays are processed
// in, and a CPU implementation of the same order reproduces it bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bhs {

// Correctly rounded FP32 square root of x >= 0 (finite or +inf). The hardware estimate s (within
// 1 ulp) is corrected by the residuals of its two float neighbours: sd = s - 1 ulp is the root if
// fma(-sd, s, x) <= 0, su = s + 1 ulp if fma(-su, s, x) > 0, otherwise s. Inputs below 2^-96 are
// scaled by 2^32 first and the root by 2^-16 after (both exact), so the residuals never lose bits to
// denormals. x = 0 and +inf come out exact. This is the compiler's own IEEE sqrt expansion without
// its branch on special values, which non-negative operands never need; written branch-free it
// keeps the march loop free of divergent control flow.
__device__ __forceinline__ float sqrt_rn(float x) {
    const bool tiny = x < 0x1p-96f;
    const float xs = tiny ? x * 0x1p+32f : x;
    const float s = __builtin_amdgcn_sqrtf(xs);
    const float sd = __int_as_float(__float_as_int(s) - 1);
    const float su = __int_as_float(__float_as_int(s) + 1);
    const float rd = __builtin_fmaf(-sd, s, xs);
    const float ru = __builtin_fmaf(-su, s, xs);
    float r = rd <= 0.0f ? sd : s;
    r = ru > 0.0f ? su : r;
    return tiny ? r * 0x1p-16f : r;
}

// 1/sqrt(x) correctly rounded to float, for x a positive normal float, from an estimate y0 within
// 2 ulps of it (v_rsq_f32 is documented within 1 ulp; it flushes denormal inputs, which therefore
// go through rsqrt_rn's scaling). One refinement in double: with e = 1 - x y0^2 (x y0 is exact in
// double, e takes one rounding), 1/sqrt(x) = y0 (1 - e)^(-1/2) = y0 (1 + e/2 + 3e^2/8 + O(e^3)) and
// |e| < 2^-21, so y0 + (y0 e)(1/2 + 3e/8) is within 2^-60 relative of 1/sqrt(x) before its last
// two roundings (to double, then to float). The 1/sqrt of every float lies at least 1.4 double
// ulps from the nearest float midpoint, so the result is the correctly rounded one. An exhaustive
// CPU check of this sequence (every x in [1, 4), every estimate within 2 ulps of the true value)
// covers all positive normal x, since each step is exact under scaling by 4^k.
__device__ __forceinline__ float rsqrt_refine(float x, float y0) {
    const double xd = (double)x, yd = (double)y0;
    const double t = xd * yd;
    const double e = __builtin_fma(-t, yd, 1.0);
    const double c = __builtin_fma(e, 0.375, 0.5);
    const double w = yd * e;
    return (float)__builtin_fma(w, c, yd);
}

// The march's reciprocal square root: correctly rounded for every positive normal operand. Its
// operands are squared lengths of positions and velocities, normal for any ray between 1e-19 and
// 1e19 from the centre (the march stops at the horizon, r = 2.02).
__device__ __forceinline__ float rsqrt_rn_core(float x) {
    return rsqrt_refine(x, __builtin_amdgcn_rsqf(x));
}

// Correctly rounded 1/sqrt(x) for every float x: positive normals by rsqrt_rn_core; +-0 -> +-inf,
// +inf -> +0; positive denormals through an exact scaling by 2^64 (result scaled by 2^32);
// negative numbers and NaN give NaN.
__device__ __forceinline__ float rsqrt_rn(float x) {
    float y = rsqrt_rn_core(x);
    if (!__builtin_amdgcn_classf(x, 0x100)) {                       // not a positive normal
        if (x == 0.0f) y = __builtin_copysignf(__builtin_inff(), x);
        else if (x == __builtin_inff()) y = 0.0f;
        else if (x > 0.0f) y = rsqrt_rn_core(x * 0x1p+64f) * 0x1p+32f;
    }
    return y;
}

// 1/sqrt(x) for the FP64 march, x a positive double within the normal float range (the march's
// operands, as for rsqrt_rn_core): the correctly rounded float 1/sqrt of (float)x (relative error
// < 1.5 * 2^-24 against 1/sqrt(x)), refined in double by one second-order step: with y^2 exact in
// double and e = 1 - x y^2 rounded once, y (1 + e/2 + 3e^2/8) is within 2^-66 relative of
// 1/sqrt(x) before its last roundings, so the result is within about half a double ulp. Every
// operation is an IEEE one on a value fixed by x, so a CPU reproduces the result bit for bit.
__device__ __forceinline__ double rsqrt_d(double x) {
    const double y = (double)rsqrt_rn_core((float)x);
    const double s = y * y;
    const double e = __builtin_fma(-x, s, 1.0);
    const double c = __builtin_fma(e, 0.375, 0.5);
    return __builtin_fma(y * e, c, y);
}

template <typename R> struct Ar;
template <> struct Ar<float> {
    static __device__ __forceinline__ float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
    static __device__ __forceinline__ float sqrt(float a) { return __builtin_sqrtf(a); }
    // Square root of the march loop's operands (see sqrt_rn): same result as sqrt().
    static __device__ __forceinline__ float sqrt_loop(float a) { return sqrt_rn(a); }
    // 1/sqrt(a) rounded once (the march loop's reciprocal square root, positive normal operands).
    static __device__ __forceinline__ float rsqrt(float a) { return rsqrt_rn_core(a); }
    static __device__ __forceinline__ float abs(float a) { return __builtin_fabsf(a); }
};
template <> struct Ar<double> {
    static __device__ __forceinline__ double fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
    static __device__ __forceinline__ double sqrt(double a) { return __builtin_sqrt(a); }
    static __device__ __forceinline__ double sqrt_loop(double a) { return __builtin_sqrt(a); }
    // The FP64 march's reciprocal square root (rsqrt_d).
    static __device__ __forceinline__ double rsqrt(double a) { return rsqrt_d(a); }
    static __device__ __forceinline__ double abs(double a) { return __builtin_fabs(a); }
};

template <typename R> struct V3 { R x, y, z; };

template <typename R>
__device__ __forceinline__ R dot3(const V3<R>& a, const V3<R>& b) {
    return Ar<R>::fma(a.z, b.z, Ar<R>::fma(a.y, b.y, a.x * b.x));
}

constexpr float kBcritF = 5.19615242270663f;

// Per-camera march constants, derived on the host from bhs_camera + bhs_march_params.
struct MarchConst {
    float pos[3];
    float basis[9];
    float tan_half_fov;
    int32_t width, height;
    float jitter[2];
    int32_t jitter_hashed;
    uint32_t jitter_base;
    int32_t max_steps;
    int32_t thin;
    float step_scale, disk_step, disk_inner, disk_outer, disk_thickness;
    float band_abs;          // |h - b_crit| < band_abs -> FP64 march (0 disables)
    int32_t precision;       // BHS_PREC_*
    int32_t root64;
    int32_t cross_all;       // record crossings at any radius
    int32_t pad_;
    int64_t ray_offset;      // first output index of this camera (batched launches)
};

// Law-S outputs of one ray, in registers.
struct RayOut {
    float esc[3];
    float dr[2], dphi[2], dg[2];
    float b;
    int32_t steps;
    uint32_t cls, flags, n_cross;
};

// ---------------------------------------------------------------------------------- hashing
__device__ __forceinline__ uint32_t pcg(uint32_t v) {  // shaders/common.glsl:28-32
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

// ---------------------------------------------------------------------------------- atan2
// atan u ~ u + u^3 (p0 + p1 u^2 + ... + p4 u^8) on |u| <= tan(pi/8): least squares of the absolute
// error at 1000 Chebyshev nodes, max error 1.6e-10 before rounding (the recipe and its refit test:
// blackholerender_amd/fit_constants.py, tests/test_fitted_constants.py).
constexpr float kAtanP[5] = {-3.333330452e-01f, 1.999792159e-01f, -1.423489004e-01f, 1.053845584e-01f, -5.953448266e-02f};

// u + u^3 P(s), s = u^2, Horner from the highest coefficient.
__device__ __forceinline__ float atan_poly(float u, float s) {
    float p = kAtanP[4];
    p = __builtin_fmaf(p, s, kAtanP[3]);
    p = __builtin_fmaf(p, s, kAtanP[2]);
    p = __builtin_fmaf(p, s, kAtanP[1]);
    p = __builtin_fmaf(p, s, kAtanP[0]);
    return __builtin_fmaf(u * s, p, u);
}

// IEEE-only atan2: t = min/max; t > tan(pi/8) -> pi/4 + atan((t-1)/(t+1)); atan_poly; quadrant
// fix-up of atan2.
__device__ __forceinline__ float atan2_ieee(float y, float x) {
    const float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
    const float mx = ax > ay ? ax : ay;
    const float mn = ax > ay ? ay : ax;
    if (mx == 0.0f) return 0.0f;
    float t = mn / mx;
    float base = 0.0f;
    if (t > 0.41421356f) {
        t = (t - 1.0f) / (t + 1.0f);
        base = 0.78539816f;
    }
    float a = base + atan_poly(t, t * t);
    if (ay > ax) a = 1.5707964f - a;
    if (x < 0.0f) a = 3.1415927f - a;
    if (y < 0.0f) a = -a;
    return a;
}

// ---------------------------------------------------------------------------------- ray generation
// blackhole.frag:68-72: px = fragCoord + jitter; ndc = px/res*2-1; dirCam = normalize(ndc.x tanH
// aspect, ndc.y tanH, -1); dir = normalize(B dirCam). Row 0 is the bottom row. In hashed mode the
// per-pixel u in [0,1)^2 (pcg keyed by seed, camera, sample, pixel) replaces the pixel centre.
__device__ __forceinline__ V3<float> ray_dir_at(const MarchConst& c, int64_t pixel, int32_t col,
                                                int32_t row) {
    float px, py;
    if (c.jitter_hashed) {
        const uint32_t pi2 = 2u * (uint32_t)pixel;
        px = (float)col + (float)(pcg(c.jitter_base + pi2) >> 8) * 0x1p-24f;
        py = (float)row + (float)(pcg(c.jitter_base + pi2 + 1u) >> 8) * 0x1p-24f;
    } else {
        px = ((float)col + 0.5f) + c.jitter[0];
        py = ((float)row + 0.5f) + c.jitter[1];
    }
    const float wf = (float)c.width, hf = (float)c.height;
    const float ndx = __builtin_fmaf(px / wf, 2.0f, -1.0f);
    const float ndy = __builtin_fmaf(py / hf, 2.0f, -1.0f);
    const float aspect = wf / hf;
    V3<float> d0{(ndx * c.tan_half_fov) * aspect, ndy * c.tan_half_fov, -1.0f};
    const float i0 = 1.0f / __builtin_sqrtf(dot3(d0, d0));
    d0.x *= i0; d0.y *= i0; d0.z *= i0;
    const float* B = c.basis;
    V3<float> d{__builtin_fmaf(B[6], d0.z, __builtin_fmaf(B[3], d0.y, B[0] * d0.x)),
                __builtin_fmaf(B[7], d0.z, __builtin_fmaf(B[4], d0.y, B[1] * d0.x)),
                __builtin_fmaf(B[8], d0.z, __builtin_fmaf(B[5], d0.y, B[2] * d0.x))};
    const float i1 = 1.0f / __builtin_sqrtf(dot3(d, d));
    d.x *= i1; d.y *= i1; d.z *= i1;
    return d;
}

// As ray_dir_at for pixel = row * width + col.
__device__ __forceinline__ V3<float> ray_dir(const MarchConst& c, int64_t pixel) {
    return ray_dir_at(c, pixel, (int32_t)(pixel % c.width), (int32_t)(pixel / c.width));
}

// ---------------------------------------------------------------------------------- law S setup
template <typename R> struct Setup {
    V3<R> v;
    R h2, h, b, b_axis, lapse_g, far_mul, c3;
};

// blackhole.frag:77-93 with initialVelocity (geodesic.glsl:27-34) and impactParameterAboutAxis
// (geodesic.glsl:37-40, axis +y), plus the impact parameter b = r sin(alpha)/sqrt(1-2/r)
// (Geodesic.h:50-54) that selects the FP64 band (h = |x0 x v0| is not b: they differ by the factor
// sqrt(1 - 2 cos^2(alpha)/r) / sqrt(1 - 2/r)).
template <typename R>
__device__ __forceinline__ Setup<R> lawS_setup(const V3<R>& P, const V3<R>& d) {
    using A = Ar<R>;
    Setup<R> s;
    const R rcam = A::sqrt(dot3(P, P));
    const R inv_r = R(1) / rcam;
    const V3<R> rh{P.x * inv_r, P.y * inv_r, P.z * inv_r};
    const R nr = dot3(d, rh);
    const V3<R> np{A::fma(-nr, rh.x, d.x), A::fma(-nr, rh.y, d.y), A::fma(-nr, rh.z, d.z)};
    const R lapse2 = R(1) - R(2) / rcam;
    const R f = A::sqrt(lapse2 > R(0) ? lapse2 : R(0));
    const R nrf = nr * f;
    V3<R> v{A::fma(nrf, rh.x, np.x), A::fma(nrf, rh.y, np.y), A::fma(nrf, rh.z, np.z)};
    const R iv = R(1) / A::sqrt(dot3(v, v));
    v.x *= iv; v.y *= iv; v.z *= iv;
    s.v = v;
    const V3<R> c{A::fma(P.y, v.z, -(P.z * v.y)), A::fma(P.z, v.x, -(P.x * v.z)),
                  A::fma(P.x, v.y, -(P.y * v.x))};
    s.h2 = dot3(c, c);
    s.h = A::sqrt(s.h2);
    {
        const V3<R> cr{A::fma(rh.y, d.z, -(rh.z * d.y)), A::fma(rh.z, d.x, -(rh.x * d.z)),
                       A::fma(rh.x, d.y, -(rh.y * d.x))};
        const R sin_a = A::sqrt(dot3(cr, cr));
        const R e12 = R(1e-12);
        s.b = (rcam * sin_a) / A::sqrt(lapse2 > e12 ? lapse2 : e12);
    }
    const R cdy = A::fma(P.z, d.x, -(P.x * d.z));
    const R e6 = R(1e-6), e4 = R(1e-4);
    s.b_axis = -cdy / A::sqrt(lapse2 > e6 ? lapse2 : e6);
    s.lapse_g = A::sqrt(lapse2 > e4 ? lapse2 : e4);
    const R fm = (s.h - R(6)) / R(3);
    s.far_mul = fm < R(1) ? R(1) : (fm > R(3) ? R(3) : fm);
    s.c3 = R(-3) * s.h2;
    return s;
}

// ---------------------------------------------------------------------------------- RK4
// geoAccel (geodesic.glsl:10-14): a = x * (-3 h2 / r^5) = x * (c3 w^5), c3 = -3 h2 per ray,
// w = rsqrt(r^2) of the stage position, w^5 = ((w w)(w w)) w.
template <typename R>
__device__ __forceinline__ R accel_scale(R c3, R w) {
    const R w2 = w * w;
    return c3 * ((w2 * w2) * w);
}

// rk4Step (geodesic.glsl:16-23). w1 = rsqrt(|x|^2) comes from the top of the march iteration and
// serves the first stage.
template <typename R>
__device__ __forceinline__ void rk4(V3<R>& x, V3<R>& v, R c3, R dt, R w1) {
    using A = Ar<R>;
    const R hdt = R(0.5) * dt;
    const R s1 = accel_scale(c3, w1);
    const V3<R> a1{x.x * s1, x.y * s1, x.z * s1};
    const V3<R> x2{A::fma(hdt, v.x, x.x), A::fma(hdt, v.y, x.y), A::fma(hdt, v.z, x.z)};
    const V3<R> v2{A::fma(hdt, a1.x, v.x), A::fma(hdt, a1.y, v.y), A::fma(hdt, a1.z, v.z)};
    const R s2 = accel_scale(c3, A::rsqrt(dot3(x2, x2)));
    const V3<R> a2{x2.x * s2, x2.y * s2, x2.z * s2};
    const V3<R> x3{A::fma(hdt, v2.x, x.x), A::fma(hdt, v2.y, x.y), A::fma(hdt, v2.z, x.z)};
    const V3<R> v3{A::fma(hdt, a2.x, v.x), A::fma(hdt, a2.y, v.y), A::fma(hdt, a2.z, v.z)};
    const R s3 = accel_scale(c3, A::rsqrt(dot3(x3, x3)));
    const V3<R> a3{x3.x * s3, x3.y * s3, x3.z * s3};
    const V3<R> x4{A::fma(dt, v3.x, x.x), A::fma(dt, v3.y, x.y), A::fma(dt, v3.z, x.z)};
    const V3<R> v4{A::fma(dt, a3.x, v.x), A::fma(dt, a3.y, v.y), A::fma(dt, a3.z, v.z)};
    const R s4 = accel_scale(c3, A::rsqrt(dot3(x4, x4)));
    const V3<R> a4{x4.x * s4, x4.y * s4, x4.z * s4};
    const R dt6 = dt * R(0.16666666666666666);
    x.x = A::fma(dt6, A::fma(R(2), v3.x, A::fma(R(2), v2.x, v.x)) + v4.x, x.x);
    x.y = A::fma(dt6, A::fma(R(2), v3.y, A::fma(R(2), v2.y, v.y)) + v4.y, x.y);
    x.z = A::fma(dt6, A::fma(R(2), v3.z, A::fma(R(2), v2.z, v.z)) + v4.z, x.z);
    v.x = A::fma(dt6, A::fma(R(2), a3.x, A::fma(R(2), a2.x, a1.x)) + a4.x, v.x);
    v.y = A::fma(dt6, A::fma(R(2), a3.y, A::fma(R(2), a2.y, a1.y)) + a4.y, v.y);
    v.z = A::fma(dt6, A::fma(R(2), a3.z, A::fma(R(2), a2.z, a1.z)) + a4.z, v.z);
}

// farFieldDirection (geodesic.glsl:43-53); r = |x|, iv = 1/|v|.
template <typename R>
__device__ __forceinline__ V3<R> far_field(const V3<R>& x, const V3<R>& v, R r, R iv, R h2) {
    using A = Ar<R>;
    const V3<R> vh{v.x * iv, v.y * iv, v.z * iv};
    const R p = dot3(x, vh);
    const V3<R> bv{A::fma(-p, vh.x, x.x), A::fma(-p, vh.y, x.y), A::fma(-p, vh.z, x.z)};
    R s = p / r;
    s = s < R(0) ? R(0) : (s > R(1) ? R(1) : s);
    const R r3 = (r * r) * r;
    const R a = h2 * iv;
    const R ir3 = R(1) / r3;
    const R op = R(1) + s;
    const R kk = (R(2) + s) / (((r3 * r) * op) * op);
    V3<R> e{v.x + (-(a * A::fma(bv.x, kk, vh.x * ir3))), v.y + (-(a * A::fma(bv.y, kk, vh.y * ir3))),
            v.z + (-(a * A::fma(bv.z, kk, vh.z * ir3)))};
    const R ie = R(1) / A::sqrt(dot3(e, e));
    e.x *= ie; e.y *= ie; e.z *= ie;
    return e;
}

// redshiftFactor (disk.glsl:37-42).
template <typename R>
__device__ __forceinline__ R redshift(R r, R b_axis, R lapse) {
    using A = Ar<R>;
    const R omega = R(1) / (r * A::sqrt(r));
    R t = R(1) - R(3) / r;
    if (t < R(0.02)) t = R(0.02);
    const R ut = R(1) / A::sqrt(t);
    R m = A::fma(-omega, b_axis, R(1));
    if (m < R(0.05)) m = R(0.05);
    return R(1) / ((lapse * ut) * m);
}

// Crossing state saved for FP64 root refinement: x0, v0 and dt of the crossing step, in the
// precision of the march that produced it.
template <typename R> struct RootState { R x0[3], v0[3], dt; };

// ---------------------------------------------------------------------------------- law S march
// Per-camera constants of the march in the march's precision (blackhole.frag:89 rEscMin,
// disk.glsl:24 slab half-height factor, the slab-clamp annulus of :172 as squared radii).
template <typename R> struct MarchK {
    R ss, dstep, inner, outer, slab_k, lo2, hi2, r_esc_min;
    int32_t max_steps;
    bool thin;
    bool cross_all;
};

template <typename R>
__device__ __forceinline__ MarchK<R> march_consts(const MarchConst& c) {
    MarchK<R> k;
    k.ss = (R)c.step_scale;
    k.dstep = (R)c.disk_step;
    k.inner = (R)c.disk_inner;
    k.outer = (R)c.disk_outer;
    k.slab_k = R(2.5) * (R)c.disk_thickness;
    const R lo = k.inner - R(1), hi = k.outer + R(1);
    k.lo2 = lo > R(0) ? lo * lo : R(-1);
    k.hi2 = hi * hi;
    const R er = k.outer * R(1.25);
    k.r_esc_min = er > R(12) ? er : R(12);
    k.max_steps = c.max_steps;
    k.thin = c.thin != 0;
    k.cross_all = c.cross_all != 0;
    return k;
}

// March state of one ray: position/velocity, completed iterations, terminal status (0 marching,
// 1 escaped, 2 captured) and the number of in-annulus crossings. Crossing records and the escape
// direction leave through a sink as soon as they are produced, so they occupy no registers.
template <typename R> struct MarchState {
    V3<R> x, v;
    int32_t i;
    int32_t status;
    uint32_t n_cross;
};

template <typename R>
__device__ __forceinline__ void march_init(MarchState<R>& st, const V3<R>& P, const Setup<R>& s) {
    st.x = P;
    st.v = s.v;
    st.i = 0;
    st.status = 0;
    st.n_cross = 0;
}

// Sink that collects a ray's records in registers (RayOut) and optionally its crossing states.
template <typename R> struct RegSink {
    RayOut& o;
    RootState<R>* rs;
    __device__ __forceinline__ void escape(float ex, float ey, float ez) {
        o.esc[0] = ex; o.esc[1] = ey; o.esc[2] = ez;
    }
    __device__ __forceinline__ void cross(uint32_t slot, float r, float phi, float g,
                                          const RootState<R>& q) {
        if (slot == 0) {
            o.dr[0] = r; o.dphi[0] = phi; o.dg[0] = g;
            if (rs) rs[0] = q;
        } else {
            o.dr[1] = r; o.dphi[1] = phi; o.dg[1] = g;
            if (rs) rs[1] = q;
        }
    }
};

// One iteration of the march loop (blackhole.frag:142-216, thin disk). Returns true when the ray
// is finished: out of steps (loop bound, :252), horizon (:144) or escape (:151-155), checked in
// that order. An inactive lane (live = false) or a finished ray changes nothing but x and v.
//
// x and v are advanced unconditionally: for a ray that finishes in this iteration they are never
// read again, and a branch around the update would make every divergent merge copy the ray state.
// Only the counters, the status and the records are predicated.
// ALL: the sink also sees every later in-annulus crossing (slot = its index, 2, 3, ...), as the
// shading march needs; otherwise only the first two.
template <typename R, typename Sink, bool ALL = false>
__device__ __forceinline__ bool march_step(MarchState<R>& st, const Setup<R>& s, const MarchK<R>& k,
                                           Sink& sink, bool live = true) {
    using A = Ar<R>;
    V3<R>& x = st.x;
    V3<R>& v = st.v;
    const bool over = !live || st.i >= k.max_steps;
    const R r2 = dot3(x, x);
    const R ir = A::rsqrt(r2);                                           // 1/r
    const R r = r2 * ir;
    const bool horizon = r < R(2.02);                                    // :144
    const R iv = A::rsqrt(dot3(v, v));
    bool escape = false;
    if (!over && !horizon && dot3(x, v) > R(0) && r > k.r_esc_min) {     // :151-155
        const R b = s.h * iv;
        const R ir2 = ir * ir;
        const R q = (((R(0.75) * b) * b) * b) * (ir2 * ir2);
        if (q < R(0.01)) {
            const V3<R> e = far_field(x, v, r, iv, s.h2);
            sink.escape((float)e.x, (float)e.y, (float)e.z);
            escape = true;
        }
    }
    if (!over) st.status = horizon ? 2 : (escape ? 1 : 0);
    const bool done = over || horizon || escape;
    R ds = (k.ss * r) * s.far_mul;                                        // :169
    if (ds < R(0.02)) ds = R(0.02);
    if (k.thin) {                                                         // :170-181
        const R rc2 = A::fma(x.z, x.z, x.x * x.x);
        if (rc2 > k.lo2 && rc2 < k.hi2) {
            const R rcyl = A::sqrt_loop(rc2);
            const R outside = A::fma(-k.slab_k, rcyl, A::abs(x.y));
            if (outside > R(0)) {
                const R den = A::fma(A::abs(v.y), iv, k.slab_k);
                R cl = outside / den;
                if (cl < k.dstep) cl = k.dstep;
                if (cl < ds) ds = cl;
            }
        }
    }
    const R dt = ds * iv;                                                 // :192
    const V3<R> x0 = x;
    const V3<R> v0 = v;
    rk4(x, v, s.c3, dt, ir);                                              // :194
    if (done) return true;
    st.i = st.i + 1;                                                      // :196
    if (k.thin && ((x.y > R(0)) != (x0.y > R(0)))) {                      // :209-216
        const R f = x0.y / (x0.y - x.y);
        const R hx = A::fma(f, x.x - x0.x, x0.x);
        const R hz = A::fma(f, x.z - x0.z, x0.z);
        const R rh = A::sqrt(A::fma(hz, hz, hx * hx));
        if (k.cross_all || (rh >= k.inner && rh <= k.outer)) {            // disk.glsl:114-115
            if (ALL || st.n_cross < 2) {
                const RootState<R> q{{x0.x, x0.y, x0.z}, {v0.x, v0.y, v0.z}, dt};
                sink.cross(st.n_cross, (float)rh, atan2_ieee((float)hz, (float)hx),
                           (float)redshift(rh, s.b_axis, s.lapse_g), q);
            }
            if (st.n_cross < 255u) st.n_cross = st.n_cross + 1;
        }
    }
    return false;
}

// Terminal class and flags (blackhole.frag:252: out of steps counts as captured).
template <typename R>
__device__ __forceinline__ void march_class(const MarchState<R>& st, uint32_t& cls, uint32_t& flags) {
    int32_t status = st.status;
    flags = 0;
    if (status == 0) { status = 2; flags |= 1u; }
    if (status == 1) flags |= 2u;
    cls = st.n_cross > 0 ? 1u : (status == 1 ? 2u : 0u);
}

// The whole march of one ray (blackhole.frag:142-253) into a RayOut.
template <typename R>
__device__ __forceinline__ void lawS_march(const V3<R>& P, const Setup<R>& s, const MarchConst& c,
                                           RayOut& o, RootState<R>* rs) {
    const MarchK<R> k = march_consts<R>(c);
    MarchState<R> st;
    march_init(st, P, s);
    o.esc[0] = o.esc[1] = o.esc[2] = 0.0f;
    o.dr[0] = o.dr[1] = o.dphi[0] = o.dphi[1] = o.dg[0] = o.dg[1] = 0.0f;
    RegSink<R> sink{o, rs};
    while (!march_step(st, s, k, sink)) {
    }
    march_class(st, o.cls, o.flags);
    o.steps = st.i;
    o.n_cross = st.n_cross;
    o.b = (float)s.b;
}

// ---------------------------------------------------------------------------------- FP64 root
// Root bracketing of y(tau) = rk4(x0, v0, tau).y on [0, dt] (Illinois variant of regula falsi,
// stop when |b - a| <= 1e-14 dt, at most 60 evaluations). Returns false when the double step does
// not bracket a sign change (the chord result then stands).
__device__ __forceinline__ bool root64(const double X0[3], const double V0[3], double dt, double c3,
                                       double& hx, double& hz) {
    double a = 0.0, fa = X0[1];
    double b = dt, fb;
    {
        V3<double> x{X0[0], X0[1], X0[2]}, v{V0[0], V0[1], V0[2]};
        const double q = dot3(x, x);
        rk4(x, v, c3, b, rsqrt_d(q));
        fb = x.y;
    }
    if (fa == 0.0) { hx = X0[0]; hz = X0[2]; return true; }
    if ((fa > 0.0) == (fb > 0.0)) return false;
    double cx = 0.0, cz = 0.0;
    int side = 0;
    for (int it = 0; it < 60; ++it) {
        const double c = (a * fb - b * fa) / (fb - fa);
        V3<double> x{X0[0], X0[1], X0[2]}, v{V0[0], V0[1], V0[2]};
        const double q = dot3(x, x);
        rk4(x, v, c3, c, rsqrt_d(q));
        const double fc = x.y;
        cx = x.x; cz = x.z;
        if (fc == 0.0) break;
        if ((fc > 0.0) == (fb > 0.0)) {
            b = c; fb = fc;
            if (side == -1) fa = 0.5 * fa;
            side = -1;
        } else {
            a = c; fa = fc;
            if (side == 1) fb = 0.5 * fb;
            side = 1;
        }
        if (__builtin_fabs(b - a) <= 1e-14 * dt) break;
    }
    hx = cx;
    hz = cz;
    return true;
}

// ---------------------------------------------------------------------------------- law T
// src/physics/Geodesic.h in GLM's operation order (no fused multiply-add): dot = (ax bx + ay by) +
// az bz, normalize = v * (1/sqrt(dot)), vec/scalar divides each component.
namespace lawT {
struct D3 { double x, y, z; };
__device__ __forceinline__ double gdot(const D3& a, const D3& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ double glen(const D3& a) { return __builtin_sqrt(gdot(a, a)); }
__device__ __forceinline__ D3 gcross(const D3& x, const D3& y) {
    return D3{x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
__device__ __forceinline__ D3 gnorm(const D3& v) {
    const double is = 1.0 / __builtin_sqrt(gdot(v, v));
    return D3{v.x * is, v.y * is, v.z * is};
}
__device__ __forceinline__ D3 accel(const D3& x, double h2) {   // Geodesic.h:21-25
    const double r2 = gdot(x, x);
    const double r = __builtin_sqrt(r2);
    const double s = -3.0 * h2 / (r2 * r2 * r);
    return D3{x.x * s, x.y * s, x.z * s};
}
__device__ __forceinline__ void rk4(D3& x, D3& v, double h2, double dt) {   // Geodesic.h:27-34
    const double hd = 0.5 * dt;
    const D3 k1v = accel(x, h2), k1x = v;
    const D3 k2v = accel(D3{x.x + hd * k1x.x, x.y + hd * k1x.y, x.z + hd * k1x.z}, h2);
    const D3 k2x{v.x + hd * k1v.x, v.y + hd * k1v.y, v.z + hd * k1v.z};
    const D3 k3v = accel(D3{x.x + hd * k2x.x, x.y + hd * k2x.y, x.z + hd * k2x.z}, h2);
    const D3 k3x{v.x + hd * k2v.x, v.y + hd * k2v.y, v.z + hd * k2v.z};
    const D3 k4v = accel(D3{x.x + dt * k3x.x, x.y + dt * k3x.y, x.z + dt * k3x.z}, h2);
    const D3 k4x{v.x + dt * k3v.x, v.y + dt * k3v.y, v.z + dt * k3v.z};
    const double d6 = dt / 6.0;
    x.x += d6 * (((k1x.x + 2.0 * k2x.x) + 2.0 * k3x.x) + k4x.x);
    x.y += d6 * (((k1x.y + 2.0 * k2x.y) + 2.0 * k3x.y) + k4x.y);
    x.z += d6 * (((k1x.z + 2.0 * k2x.z) + 2.0 * k3x.z) + k4x.z);
    v.x += d6 * (((k1v.x + 2.0 * k2v.x) + 2.0 * k3v.x) + k4v.x);
    v.y += d6 * (((k1v.y + 2.0 * k2v.y) + 2.0 * k3v.y) + k4v.y);
    v.z += d6 * (((k1v.z + 2.0 * k2v.z) + 2.0 * k3v.z) + k4v.z);
}
__device__ __forceinline__ D3 initial_velocity(const D3& cam, const D3& dir) {   // :40-47
    const double r = glen(cam);
    const D3 rh{cam.x / r, cam.y / r, cam.z / r};
    const double nr = gdot(dir, rh);
    const D3 np{dir.x - nr * rh.x, dir.y - nr * rh.y, dir.z - nr * rh.z};
    const double t = 1.0 - 2.0 / r;
    const double f = __builtin_sqrt(t > 0.0 ? t : 0.0);
    const double nf = nr * f;
    return gnorm(D3{nf * rh.x + np.x, nf * rh.y + np.y, nf * rh.z + np.z});
}
__device__ __forceinline__ D3 far_field(const D3& x, const D3& v, double h2) {   // :65-75
    const double vlen = glen(v);
    const D3 vh{v.x / vlen, v.y / vlen, v.z / vlen};
    const double r = glen(x);
    const double p = gdot(x, vh);
    const D3 bv{x.x - p * vh.x, x.y - p * vh.y, x.z - p * vh.z};
    double s = p / r;
    s = s < 0.0 ? 0.0 : (s > 1.0 ? 1.0 : s);
    const double r3 = r * r * r;
    const double a = -(h2 / vlen);
    const double kk = (2.0 + s) / (r3 * r * (1.0 + s) * (1.0 + s));
    return gnorm(D3{v.x + a * (vh.x / r3 + bv.x * kk), v.y + a * (vh.y / r3 + bv.y * kk),
                    v.z + a * (vh.z / r3 + bv.z * kk)});
}
}  // namespace lawT

}  // namespace bhs

// mlp_device.hpp — the learned surrogate of the law-S march: network layout, the physics prior,
// input features, output decoding and the activation, shared by training-data generation,
// training, validation and inference.
//
// Network: 3 -> 128 -> 128 -> 128 -> 5, SiLU after each hidden layer. Layer 1 runs in FP32 on the
// VALU; layers 2-4 run on bf16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation. Training
// rounds the same operands to bf16, so training and inference compute the same function.
//
// Orbit-plane frame of one ray from camera position P with view direction d: e1 = P/|P|, e2 the
// unit tangential part of the initial coordinate velocity v0 (initialVelocity, geodesic.glsl:27-34).
// The ray sweeps the position angle phi from 0 (camera) upwards; it meets the equatorial plane
// y = 0 at phi = psi + k pi, psi in [0, pi).
//
// Inputs:
//   x1 = log r_obs mapped from [log 8, log 1000] to [-1, 1]
//   x2 = sign(beta - 1) log(1 + |beta - 1| / 0.002) / log(1 + 1e5), beta = b / b_crit
//   x3 = psi mapped from [0, pi) to [-1, 1)
// Outputs:
//   y0  escape: residual of the law-S escape angle over the exact Schwarzschild swept angle,
//       Phi = Phi_exact + 1e-3 y0, escape direction = cos(Phi) e1 + sin(Phi) e2
//   y1, y2  crossings at phi = psi and psi + pi: u = 1/r = u_exact exp(1e-3 y), u_exact the exact
//           orbit's value (Newton on the quadrature of the swept angle)
//   y3, y4  logits of "the march reaches crossing 1 / 2"
//
// Phi_exact = pi + ahat(beta) - J(u_obs, b): the swept angle from the camera through periapsis to
// infinity, 2 int_0^{u0} - int_0^{u_obs} of du / sqrt(1/b^2 - u^2 (1 - 2u)), u0 the periapsis.
// ahat is the infinity-to-infinity deflection as a 25-term Chebyshev series in log(beta - 1)
// (fitted once to the exact quadrature, error 6e-7 rad), J the 8-node Gauss-Legendre value of the
// finite-distance part after the substitution u = u0 (1 - t^2). The same quadratures, inverted by
// Newton steps, give the exact crossing radii (relative error < 2e-6 against a 16-node reference).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lawS_device.hpp"

namespace bhs {
namespace mlp {

constexpr int kIn = 3, kH = 128, kOut = 5;
// Offsets into the FP32 parameter vector (row-major weights [out][in]).
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kH * kIn;
constexpr int kW2 = kB1 + kH;
constexpr int kB2 = kW2 + kH * kH;
constexpr int kW3 = kB2 + kH;
constexpr int kB3 = kW3 + kH * kH;
constexpr int kW4 = kB3 + kH;
constexpr int kB4 = kW4 + kOut * kH;
constexpr int kParams = kB4 + kOut;   // 34181

constexpr float kResScale = 1e-3f;   // escape residual unit (rad)
constexpr float kURes = 1e-3f;            // crossing residual unit: u = u_exact exp(1e-3 y)
constexpr float kBetaEps = 0.002f;         // domain: |beta - 1| > 0.2%
constexpr float kX2Scale = 0.002f;         // x2 feature: log(1 + |beta - 1| / kX2Scale) / kX2Norm
constexpr float kX2Norm = 11.512935464920229f;   // log(1 + 1e5)
constexpr float kLogRLo = 2.0794415f;      // log 8
constexpr float kLogRHi = 6.9077553f;      // log 1000
constexpr float kPi = 3.14159265358979f;
constexpr float kSinMax = 0.95f;           // domain: sin(view angle) <= 0.95
constexpr float kBetaMin = 0.3f;           // domain: beta >= 0.3 (near-radial rays: u(psi) is too
                                           // steep in beta for the network's resolution)

// Chebyshev coefficients of ahat(s), s = log(beta - 1) in [log 1e-3, log 200]: the total swept angle
// of a null geodesic from infinity to infinity minus pi, least squares at 200 Chebyshev nodes of
// Darwin's integral (max error 4.7e-7 rad; recipe and refit test: blackholerender_amd/
// fit_constants.py, tests/test_fitted_constants.py).
constexpr int kAhatTerms = 25;
constexpr float kAhatC[kAhatTerms] = {
    2.252328873e+00f, -3.297978401e+00f, 1.115420222e+00f, 6.010319665e-02f,
    -1.329363883e-01f, -2.116844803e-02f, 2.762471139e-02f, 7.664506324e-03f,
    -6.423670333e-03f, -2.731987508e-03f, 1.508211833e-03f, 9.516102145e-04f,
    -3.349810431e-04f, -3.234512697e-04f, 6.337356172e-05f, 1.071540391e-04f,
    -6.627239145e-06f, -3.451554221e-05f, -2.213071866e-06f, 1.076288481e-05f,
    2.082474793e-06f, -3.224703960e-06f, -1.095450330e-06f, 9.161439607e-07f,
    4.807004075e-07f};
constexpr float kAhatS0 = -6.9077553f, kAhatS1 = 5.2983174f;

// bm = beta - 1 > 0.
__device__ __forceinline__ float ahat(float bm) {
    bm = bm < 1e-3f ? 1e-3f : (bm > 200.0f ? 200.0f : bm);
    const float s = logf(bm);
    const float t = (s - 0.5f * (kAhatS0 + kAhatS1)) / (0.5f * (kAhatS1 - kAhatS0));
    float b1 = 0.0f, b2 = 0.0f;
#pragma unroll
    for (int k = kAhatTerms - 1; k >= 1; --k) {   // Clenshaw recurrence
        const float b0 = __builtin_fmaf(2.0f * t, b1, kAhatC[k] - b2);
        b2 = b1;
        b1 = b0;
    }
    return __builtin_fmaf(t, b1, kAhatC[0] - b2);
}

// 8-node Gauss-Legendre rule on [-1, 1] (symmetric half).
constexpr float kGLx[4] = {0.1834346424956498f, 0.5255324099163290f, 0.7966664774136267f,
                           0.9602898564975363f};
constexpr float kGLw[4] = {0.3626837833783620f, 0.3137066458778873f, 0.2223810344533745f,
                           0.1012285362903763f};

// Escaping side (b > b_crit) of du/dphi = sqrt(Q), Q = 1/b^2 - u^2 (1 - 2u). With the roots
// u0 < u1 and u2 < 0 of 2u^3 - u^2 + 1/b^2 (u0 = 1/6 + cos(theta/3 - 2pi/3)/3,
// cos theta = 1 - 54/b^2), Q = (u0 - u) R(u), R = 2(-u^2 + (1/2 - u0) u + 1/(2 b^2 u0)), and
// u = u0 (1 - t^2) turns the swept angle from periapsis into G(t) = int_0^t g(s) ds with the smooth
// g(s) = 2 sqrt(u0) / sqrt(R(u0 (1 - s^2))); G(1) = I0 is half the total sweep.
struct OrbitE {
    float u0, cu, k0, g0;
};

__device__ __forceinline__ OrbitE orbit_e(float b) {
    OrbitE o;
    const float ib2 = 1.0f / (b * b);
    float ct = 1.0f - 54.0f * ib2;
    ct = ct < -1.0f ? -1.0f : ct;
    const float th = acosf(ct);
    o.u0 = (1.0f / 6.0f) + cosf(th * (1.0f / 3.0f) - 2.0943951023931953f) * (1.0f / 3.0f);
    o.cu = 0.5f - o.u0;
    o.k0 = 0.5f * ib2 / o.u0;
    o.g0 = 2.0f * sqrtf(o.u0);
    return o;
}

__device__ __forceinline__ float g_e(const OrbitE& o, float s) {
    const float u = o.u0 * __builtin_fmaf(-s, s, 1.0f);
    return o.g0 * __builtin_amdgcn_rsqf(2.0f * __builtin_fmaf(u, o.cu - u, o.k0));
}

// int_a^b g(s) ds (8-node Gauss-Legendre).
__device__ __forceinline__ float g_e_int(const OrbitE& o, float a, float b) {
    const float half = 0.5f * (b - a), mid = 0.5f * (b + a);
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        s = __builtin_fmaf(kGLw[i],
                           g_e(o, __builtin_fmaf(-kGLx[i], half, mid)) +
                               g_e(o, __builtin_fmaf(kGLx[i], half, mid)),
                           s);
    return s * half;
}

// J = int_0^{u_obs} du / sqrt(Q) = int_{t_obs}^1 g(s) ds, t_obs = sqrt(1 - u_obs/u0).
__device__ __forceinline__ float t_of_u(const OrbitE& o, float u) {
    const float q = 1.0f - u / o.u0;
    return sqrtf(q > 0.0f ? q : 0.0f);
}

// u at swept angle Th from periapsis (0 <= Th < I0): Newton on G(t) = Th from t = Th / I0.
__device__ __forceinline__ float u_periapsis(const OrbitE& o, float I0, float Th) {
    float t = Th / I0;
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        t -= (g_e_int(o, 0.0f, t) - Th) / g_e(o, t);
        t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    }
    return o.u0 * __builtin_fmaf(-t, t, 1.0f);
}

// Captured side (b < b_crit): Q = q_min + d^2 + 2 d^3 with d = u - 1/3, q_min = 1/b^2 - 1/27 > 0.
// d = a sinh(tau), a = sqrt(q_min), makes the swept angle int h(tau) dtau with the smooth
// h = 1 / sqrt(1 + 2 a sinh^3(tau) / cosh^2(tau)).
__device__ __forceinline__ float h_c(float a, float tau) {
    const float e = __expf(tau), ie = __builtin_amdgcn_rcpf(e);
    const float sh = 0.5f * (e - ie), ch = 0.5f * (e + ie);
    return __builtin_amdgcn_rsqf(__builtin_fmaf(2.0f * a * sh * sh, sh * __builtin_amdgcn_rcpf(ch * ch), 1.0f));
}

__device__ __forceinline__ float h_c_int(float a, float lo, float hi) {
    const float half = 0.5f * (hi - lo), mid = 0.5f * (hi + lo);
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        s = __builtin_fmaf(kGLw[i],
                           h_c(a, __builtin_fmaf(-kGLx[i], half, mid)) +
                               h_c(a, __builtin_fmaf(kGLx[i], half, mid)),
                           s);
    return s * half;
}

__device__ __forceinline__ float asinh_f(float x) {
    const float ax = fabsf(x);
    const float r = logf(ax + sqrtf(__builtin_fmaf(ax, ax, 1.0f)));
    return x < 0.0f ? -r : r;
}

// Geometry of one camera ray for the surrogate.
struct RayGeom {
    V3<float> e1, e2;     // orbit-plane basis
    float r;              // r_obs
    float bm;             // beta - 1
    float b;              // impact parameter
    float psi;            // swept angle of the first equatorial crossing, [0, pi)
    float sin_a;          // sin of the view angle from the radial direction
    bool ingoing;
    bool planar;          // degenerate orbit plane (radial ray) or camera on the equator
};

// atan2 with hardware reciprocals (the polynomial of atan2_ieee; error ~3e-7 rad): the angles of
// the surrogate's features and outputs need ~1e-6, not the integrator's exact rounding.
__device__ __forceinline__ float atan2_fast(float y, float x) {
    const float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
    const float mx = ax > ay ? ax : ay;
    const float mn = ax > ay ? ay : ax;
    if (mx == 0.0f) return 0.0f;
    float t = mn * __builtin_amdgcn_rcpf(mx);
    float base = 0.0f;
    if (t > 0.41421356f) {
        t = (t - 1.0f) * __builtin_amdgcn_rcpf(t + 1.0f);
        base = 0.78539816f;
    }
    float a = base + atan_poly(t, t * t);
    if (ay > ax) a = 1.5707964f - a;
    if (x < 0.0f) a = 3.1415927f - a;
    return y < 0.0f ? -a : a;
}

// (Hardware square-root and reciprocal approximations: the frame and the view angle feed the
// network and the decode, which need ~1e-6 relative, not the integrator's exact rounding.)
__device__ __forceinline__ RayGeom ray_geom(const V3<float>& P, const V3<float>& d,
                                            const Setup<float>& s) {
    RayGeom g;
    g.r = __builtin_amdgcn_sqrtf(dot3(P, P));
    const float ir = __builtin_amdgcn_rcpf(g.r);
    g.e1 = V3<float>{P.x * ir, P.y * ir, P.z * ir};
    const V3<float>& v = s.v;
    const float vr = dot3(v, g.e1);
    V3<float> t{v.x - vr * g.e1.x, v.y - vr * g.e1.y, v.z - vr * g.e1.z};
    const float tn = __builtin_amdgcn_sqrtf(dot3(t, t));
    g.planar = !(tn > 1e-6f);
    const float it = g.planar ? 0.0f : __builtin_amdgcn_rcpf(tn);
    g.e2 = V3<float>{t.x * it, t.y * it, t.z * it};
    g.ingoing = vr < 0.0f;
    // beta - 1 in float: its rounding (~6e-8) moves the exact sweep by ~6e-8 / |beta - 1| rad,
    // 3e-5 rad at the domain edge (the error map validates what it does to served rays)
    g.b = s.b;
    g.bm = s.b * (1.0f / kBcritF) - 1.0f;
    const float f = __builtin_amdgcn_sqrtf(fmaxf(1.0f - 2.0f * ir, 1e-12f));
    g.sin_a = s.b * f * ir;
    float psi = atan2_fast(-g.e1.y, g.e2.y);
    if (psi < 0.0f) psi += kPi;
    if (psi >= kPi) psi -= kPi;
    g.psi = psi;
    g.planar = g.planar || fabsf(g.e1.y) < 1e-6f;
    return g;
}

// Exact Schwarzschild orbit of a ray in the surrogate's domain: total swept angle to infinity
// (escaping) or to the horizon test radius 2.02 (captured), and the radii of its equatorial
// crossings at swept angles psi + k pi.
struct Orbit {
    bool escaping;
    float phi_end;     // swept angle at escape (infinity) / at r = 2.02
    // escaping
    OrbitE e;
    float I0, th_obs;  // half sweep, swept angle from camera to periapsis
    // captured
    float a, tau_obs;
};

__device__ __forceinline__ Orbit orbit(const RayGeom& g) {
    Orbit o;
    o.escaping = g.bm > 0.0f;
    const float u_obs = 1.0f / g.r;
    if (o.escaping) {
        o.e = orbit_e(g.b);
        o.I0 = g_e_int(o.e, 0.0f, 1.0f);
        o.th_obs = g_e_int(o.e, 0.0f, t_of_u(o.e, u_obs));
        // pi + ahat: the infinity-to-infinity sweep from its Chebyshev series (accurate to the
        // band edge); the camera's part from the quadrature
        o.phi_end = kPi + ahat(g.bm) - (o.I0 - o.th_obs);
    } else {
        const float ib2 = 1.0f / (g.b * g.b);
        o.a = sqrtf(fmaxf(ib2 - (1.0f / 27.0f), 1e-12f));
        const float ia = 1.0f / o.a;
        o.tau_obs = asinh_f((u_obs - (1.0f / 3.0f)) * ia);
        const float tau_h = asinh_f(((1.0f / 2.02f) - (1.0f / 3.0f)) * ia);
        o.phi_end = h_c_int(o.a, o.tau_obs, tau_h);
    }
    return o;
}

// u at swept angle phi from the camera (phi < phi_end).
__device__ __forceinline__ float orbit_u(const Orbit& o, float phi) {
    if (o.escaping) {
        const float th = fabsf(phi - o.th_obs);   // ingoing before periapsis, outgoing after
        return u_periapsis(o.e, o.I0, th < o.I0 ? th : o.I0);
    }
    float tau = o.tau_obs + phi;
#pragma unroll
    for (int it = 0; it < 3; ++it) tau -= (h_c_int(o.a, o.tau_obs, tau) - phi) / h_c(o.a, tau);
    const float e = __expf(tau);
    return (1.0f / 3.0f) + o.a * 0.5f * (e - __builtin_amdgcn_rcpf(e));
}


// Natural log of r_obs from the hardware log2 (r_obs in [1, 1e4]: absolute error ~1e-7).
__device__ __forceinline__ float log_r(float r) {
    return __builtin_amdgcn_logf(r) * 0.69314718f;
}

// Input feature of beta - 1 (scale kX2Scale, normalised by kX2Norm).
__device__ __forceinline__ float x2_of(float bm) {
    // hardware log2: 1 + |beta - 1| / kX2Scale >= 1 (absolute error ~1e-7 in x2)
    const float m = __builtin_amdgcn_logf(1.0f + fabsf(bm) * (1.0f / kX2Scale)) *
                    (0.69314718f / kX2Norm);
    return bm < 0.0f ? -m : m;
}

// Inverse of x2_of: beta - 1 of a feature value.
__device__ __forceinline__ float bm_of_x2(float x2) {
    const float m = kX2Scale * expm1f(fabsf(x2) * kX2Norm);
    return x2 < 0.0f ? -m : m;
}

__device__ __forceinline__ void features(const RayGeom& g, float x[3]) {
    x[0] = (log_r(g.r) - kLogRLo) * (2.0f / (kLogRHi - kLogRLo)) - 1.0f;
    x[1] = x2_of(g.bm);
    x[2] = g.psi * (2.0f / kPi) - 1.0f;
}

// Surrogate domain of a ray (independent of the trained model).
__device__ __forceinline__ bool in_domain(const RayGeom& g) {
    return g.ingoing && !g.planar && g.sin_a <= kSinMax && g.r >= 8.0f && g.r <= 1000.0f &&
           fabsf(g.bm) > kBetaEps && g.bm >= kBetaMin - 1.0f;
}

// Plunging rays (beta < kBetaMin: captured almost radially): decided by the exact orbit alone,
// without the network (served when they meet the equatorial plane at most once, well away from
// the annulus).
__device__ __forceinline__ bool plunging(const RayGeom& g) {
    return g.ingoing && !g.planar && g.r >= 8.0f && g.r <= 1000.0f && g.bm < kBetaMin - 1.0f;
}

// SiLU and its derivative (approximate exp and reciprocal: the bf16 rounding of the hidden
// activations dominates their error).
__device__ __forceinline__ float sigmoid_fast(float z) {
    return __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid_fast(z); }
__device__ __forceinline__ float dsilu(float z) {
    const float sg = sigmoid_fast(z);
    return sg * __builtin_fmaf(z, 1.0f - sg, 1.0f);
}

// Round-to-nearest-even FP32 -> bf16 bits (finite inputs).
__device__ __forceinline__ uint16_t bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_round(float f) { return __uint_as_float((uint32_t)bf16_bits(f) << 16); }

// Layer-1 pre-activation in the evaluation order shared by training and inference.
__device__ __forceinline__ float l1_pre(const float* w, float b, const float x[3]) {
    return __builtin_fmaf(w[2], x[2], __builtin_fmaf(w[1], x[1], w[0] * x[0])) + b;
}

}  // namespace mlp
}  // namespace bhs

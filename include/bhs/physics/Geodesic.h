#pragma once
// Drop-in replacement for the reference's src/physics/Geodesic.h (namespace bhs::physics).
//
// Same names, signatures, defaults and TraceResult layout as the reference header. The closed-form
// helpers and rk4Step stay host FP64 inline functions, as in the reference (callers such as the
// photon-sphere check step rk4Step themselves). trace() runs on the GPU: it hands the ray to
// bhs_trace_rays (include/bhs_trace.h), whose HIP kernel evaluates the reference's trace() loop in
// FP64 with GLM's operation order, so the result equals the reference's CPU result bit for bit.
// traceBatch() traces many rays in one launch.
//
// Differences from the reference: trace() needs a HIP device and throws std::runtime_error when
// the context cannot be created or the call fails (the reference cannot fail). Each host thread
// gets its own GPU context on device BHS_DEVICE (default 0; set the environment variable to pick
// another).
#include <glm/glm.hpp>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../bhs_trace.h"

namespace bhs::physics {

constexpr double kHorizon = 2.0;
constexpr double kPhotonSphere = 3.0;
constexpr double kIsco = 6.0;

inline double criticalImpactParameter() { return 3.0 * std::sqrt(3.0); }

inline glm::dvec3 accel(const glm::dvec3& x, double h2) {
    const double r2 = glm::dot(x, x);
    const double r = std::sqrt(r2);
    return x * (-3.0 * h2 / (r2 * r2 * r));
}

inline void rk4Step(glm::dvec3& x, glm::dvec3& v, double h2, double dt) {
    const double hd = 0.5 * dt;
    const glm::dvec3 k1v = accel(x, h2), k1x = v;
    const glm::dvec3 k2v = accel(x + hd * k1x, h2), k2x = v + hd * k1v;
    const glm::dvec3 k3v = accel(x + hd * k2x, h2), k3x = v + hd * k2v;
    const glm::dvec3 k4v = accel(x + dt * k3x, h2), k4x = v + dt * k3v;
    const double d6 = dt / 6.0;
    x += d6 * (k1x + 2.0 * k2x + 2.0 * k3x + k4x);
    v += d6 * (k1v + 2.0 * k2v + 2.0 * k3v + k4v);
}

inline glm::dvec3 initialVelocity(const glm::dvec3& camPos, const glm::dvec3& dir) {
    const double r = glm::length(camPos);
    const glm::dvec3 rhat = camPos / r;
    const double nr = glm::dot(dir, rhat);
    const glm::dvec3 nperp = dir - nr * rhat;
    const double f = std::sqrt(std::max(1.0 - 2.0 / r, 0.0));
    return glm::normalize(nr * f * rhat + nperp);
}

inline double impactParameter(const glm::dvec3& camPos, const glm::dvec3& dir) {
    const double r = glm::length(camPos);
    const double sinA = glm::length(glm::cross(camPos / r, dir));
    return r * sinA / std::sqrt(std::max(1.0 - 2.0 / r, 1e-12));
}

inline double impactParameterAboutAxis(const glm::dvec3& camPos, const glm::dvec3& dir,
                                       const glm::dvec3& axis) {
    const double r = glm::length(camPos);
    return -glm::dot(glm::cross(camPos, dir), axis) / std::sqrt(std::max(1.0 - 2.0 / r, 1e-12));
}

inline glm::dvec3 farFieldDirection(const glm::dvec3& x, const glm::dvec3& v, double h2) {
    const double vlen = glm::length(v);
    const glm::dvec3 vh = v / vlen;
    const double r = glm::length(x);
    const double p = glm::dot(x, vh);
    const glm::dvec3 bvec = x - p * vh;
    const double s = std::clamp(p / r, 0.0, 1.0);
    const double r3 = r * r * r;
    const glm::dvec3 dv =
        -(h2 / vlen) * (vh / r3 + bvec * ((2.0 + s) / (r3 * r * (1.0 + s) * (1.0 + s))));
    return glm::normalize(v + dv);
}

inline double shadowAngularRadius(double r) {
    const double s =
        std::min(1.0, criticalImpactParameter() / r * std::sqrt(std::max(1.0 - 2.0 / r, 0.0)));
    const double a = std::asin(s);
    return (r >= kPhotonSphere) ? a : (M_PI - a);
}

inline double keplerOmega(double r) { return std::pow(r, -1.5); }

inline double redshiftFactor(double rEmit, double bAxis, double rCam) {
    const double omega = keplerOmega(rEmit);
    const double ut = 1.0 / std::sqrt(std::max(1.0 - 3.0 / rEmit, 1e-6));
    const double lapseCam = std::sqrt(std::max(1.0 - 2.0 / rCam, 1e-6));
    return 1.0 / (lapseCam * ut * (1.0 - omega * bAxis));
}

struct TraceResult {
    enum Status { Escaped, Captured, MaxSteps } status = MaxSteps;
    glm::dvec3 dir{0.0};
    int steps = 0;
    double minR = 1e30;
    int equatorCrossings = 0;
};

namespace detail {
// One GPU context per host thread (a bhs_ctx serves one thread at a time).
inline bhs_ctx* context() {
    struct Holder {
        bhs_ctx* ctx = nullptr;
        ~Holder() {
            if (ctx) bhs_destroy(ctx);
        }
    };
    thread_local Holder h;
    if (!h.ctx) {
        const char* env = std::getenv("BHS_DEVICE");
        const int dev = env ? std::atoi(env) : 0;
        int st = 0;
        h.ctx = bhs_create(dev, &st);
        if (!h.ctx)
            throw std::runtime_error("bhs::physics: no HIP device " + std::to_string(dev) +
                                     " (bhs_create status " + std::to_string(st) + ")");
    }
    return h.ctx;
}

inline TraceResult toResult(const bhs_trace_result& r) {
    TraceResult t;
    t.status = static_cast<TraceResult::Status>(r.status);
    t.dir = glm::dvec3(r.dir[0], r.dir[1], r.dir[2]);
    t.steps = r.steps;
    t.minR = r.min_r;
    t.equatorCrossings = r.equator_crossings;
    return t;
}
}  // namespace detail

// Many rays of trace() in one GPU launch.
inline std::vector<TraceResult> traceBatch(const std::vector<glm::dvec3>& camPos,
                                           const std::vector<glm::dvec3>& dir,
                                           int maxSteps = 200000, double stepScale = 0.005,
                                           double escTol = 1e-4, double rEscMin = 30.0) {
    if (camPos.size() != dir.size())
        throw std::invalid_argument("bhs::physics::traceBatch: camPos and dir sizes differ");
    const size_t n = camPos.size();
    std::vector<double> p(3 * n), d(3 * n);
    for (size_t i = 0; i < n; ++i) {
        p[3 * i] = camPos[i].x; p[3 * i + 1] = camPos[i].y; p[3 * i + 2] = camPos[i].z;
        d[3 * i] = dir[i].x; d[3 * i + 1] = dir[i].y; d[3 * i + 2] = dir[i].z;
    }
    std::vector<bhs_trace_result> raw(n);
    bhs_ctx* ctx = detail::context();
    const int st = bhs_trace_rays(ctx, p.data(), d.data(), (int64_t)n, maxSteps, stepScale, escTol,
                                  rEscMin, raw.data());
    if (st != BHS_OK) throw std::runtime_error(std::string("bhs::physics::trace: ") + bhs_last_error(ctx));
    std::vector<TraceResult> out(n);
    for (size_t i = 0; i < n; ++i) out[i] = detail::toResult(raw[i]);
    return out;
}

// Traces a ray from camPos in local direction dir (reference semantics, GPU FP64).
inline TraceResult trace(const glm::dvec3& camPos, const glm::dvec3& dir, int maxSteps = 200000,
                         double stepScale = 0.005, double escTol = 1e-4, double rEscMin = 30.0) {
    return traceBatch({camPos}, {dir}, maxSteps, stepScale, escTol, rEscMin)[0];
}

}  // namespace bhs::physics

// The physics checks of the reference's tests/geodesic_tests.cpp, restated in C++ against the
// drop-in header include/bhs/physics/Geodesic.h: the same calls (trace with the same arguments,
// impactParameter, shadowAngularRadius, rk4Step, redshiftFactor, ...) and the same thresholds, with
// trace() running on the GPU. Exit status 0 when every check passes. Also checks that traceBatch
// equals per-ray trace bit for bit.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "bhs/physics/Geodesic.h"

using namespace bhs::physics;
using glm::dvec3;

static int failures = 0, checks = 0;

static void check(bool ok, const char* what) {
    ++checks;
    if (!ok) ++failures;
    std::printf("  %s  %s\n", ok ? "ok  " : "FAIL", what);
}

static dvec3 dirFor(double D, double b) {  // camera on +z at distance D, impact parameter b
    const double a = std::asin(std::min(1.0, b * std::sqrt(1.0 - 2.0 / D) / D));
    return dvec3(std::sin(a), 0.0, -std::cos(a));
}

static double angle(const dvec3& a, const dvec3& b) {
    return std::acos(std::clamp(glm::dot(glm::normalize(a), glm::normalize(b)), -1.0, 1.0));
}

int main() {
    char msg[200];
    for (double D : {10.0, 25.0, 2000.0})
        for (double b : {1.0, 5.196, 8.0}) {
            if (b * std::sqrt(1.0 - 2.0 / D) / D >= 1.0) continue;
            const double bb = impactParameter(dvec3(0, 0, D), dirFor(D, b));
            std::snprintf(msg, sizeof msg, "impact parameter b=%.3f r=%.0f -> %.9f", b, D, bb);
            check(std::abs(bb - b) < 1e-9 * std::max(1.0, b), msg);
        }
    for (double b : {50.0, 100.0, 200.0}) {
        const dvec3 d = dirFor(2000.0, b);
        const TraceResult r = trace(dvec3(0, 0, 2000.0), d, 400000, 0.002, 1e-6, 50.0);
        const double expect = 4.0 / b + 15.0 * M_PI / (4.0 * b * b);
        const double rel = std::abs(angle(d, r.dir) - expect) / expect;
        std::snprintf(msg, sizeof msg, "weak field b=%.0f rel err %.4f", b, rel);
        check(r.status == TraceResult::Escaped && rel < 0.01, msg);
    }
    for (double D : {10.0, 25.0, 2000.0}) {
        const TraceResult in = trace(dvec3(0, 0, D), dirFor(D, 5.15), 400000, 0.002, 1e-4, 30.0);
        const TraceResult out = trace(dvec3(0, 0, D), dirFor(D, 5.25), 400000, 0.002, 1e-4, 30.0);
        std::snprintf(msg, sizeof msg, "capture threshold from r=%.0f", D);
        check(in.status == TraceResult::Captured && out.status == TraceResult::Escaped, msg);
    }
    for (double r : {4.0, 6.0, 25.0, 100.0}) {
        const double a = shadowAngularRadius(r), da = 0.004;
        const TraceResult in = trace(dvec3(0, 0, r), dvec3(std::sin(a - da), 0, -std::cos(a - da)),
                                     400000, 0.002, 1e-4, 30.0);
        const TraceResult out = trace(dvec3(0, 0, r), dvec3(std::sin(a + da), 0, -std::cos(a + da)),
                                      400000, 0.002, 1e-4, 30.0);
        std::snprintf(msg, sizeof msg, "shadow edge r=%.0f", r);
        check(in.status == TraceResult::Captured && out.status == TraceResult::Escaped, msg);
    }
    {
        dvec3 x(3.0, 0.0, 0.0);
        dvec3 v = initialVelocity(x, dvec3(0.0, 0.0, 1.0));
        const dvec3 hv = glm::cross(x, v);
        const double h2 = glm::dot(hv, hv);
        double maxDev = 0.0;
        for (double s = 0.0; s < 2.0 * 2.0 * M_PI * 3.0; s += 0.01) {
            rk4Step(x, v, h2, 0.01 / glm::length(v));
            maxDev = std::max(maxDev, std::abs(glm::length(x) - 3.0));
        }
        std::snprintf(msg, sizeof msg, "photon sphere max |r-3| = %.2e", maxDev);
        check(maxDev < 1e-4, msg);
    }
    {
        const dvec3 cam(0, 0, 30), axis(0, 1, 0);
        const double bAway = impactParameterAboutAxis(cam, glm::normalize(dvec3(8, 0, 0) - cam), axis);
        const double bToward = impactParameterAboutAxis(cam, glm::normalize(dvec3(-8, 0, 0) - cam), axis);
        check(bAway < 0.0 && bToward > 0.0, "redshift: sign of L_y/E");
        check(std::abs(std::abs(bAway) - 8.0) < 0.15, "redshift: |L_y/E| ~ 8");
        const double gAway = redshiftFactor(8.0, bAway, 30.0), gToward = redshiftFactor(8.0, bToward, 30.0);
        check(gAway < 1.0 && gToward > 1.0, "redshift: receding < 1 < approaching");
        const double expect = 1.0 / (std::sqrt(1 - 2.0 / 30) / std::sqrt(1 - 3.0 / 8) * (1 - std::pow(8.0, -1.5) * bAway));
        check(std::abs(gAway - expect) < 1e-9, "redshift: closed form");
        const dvec3 top(0, 30, 0);
        const double bTop = impactParameterAboutAxis(top, glm::normalize(dvec3(8, 0, 0) - top), axis);
        const double gTop = redshiftFactor(8.0, bTop, 30.0);
        check(std::abs(bTop) < 1e-9 && std::abs(gTop - std::sqrt(1 - 3.0 / 8) / std::sqrt(1 - 2.0 / 30)) < 1e-9,
              "redshift: face-on");
    }
    for (double b : {8.0, 20.0, 60.0}) {
        const dvec3 d = dirFor(2000.0, b);
        const TraceResult early = trace(dvec3(0, 0, 2000.0), d, 400000, 0.002, 1e-2, 12.0);
        const TraceResult late = trace(dvec3(0, 0, 2000.0), d, 400000, 0.002, 1e-7, 12.0);
        std::snprintf(msg, sizeof msg, "far field b=%.0f: %.2e rad", b, angle(early.dir, late.dir));
        check(early.status == TraceResult::Escaped && late.status == TraceResult::Escaped &&
                  angle(early.dir, late.dir) < 5e-4, msg);
    }
    {
        const dvec3 cam(0.0, 2.6, 24.9);
        const double rc = glm::length(cam);
        const dvec3 toHole = -cam / rc;
        const dvec3 up = glm::normalize(dvec3(0, 1, 0) - toHole * glm::dot(dvec3(0, 1, 0), toHole));
        const double a = shadowAngularRadius(rc) + 0.01;
        const TraceResult above = trace(cam, glm::normalize(toHole * std::cos(a) + up * std::sin(a)), 400000, 0.002, 1e-4, 30.0);
        const TraceResult below = trace(cam, glm::normalize(toHole * std::cos(a) - up * std::sin(a)), 400000, 0.002, 1e-4, 30.0);
        check(above.status == TraceResult::Escaped && above.equatorCrossings >= 1, "disk image above");
        check(below.status == TraceResult::Escaped && below.equatorCrossings >= 1, "disk image below");
    }
    {   // traceBatch == per-ray trace, bit for bit
        std::vector<dvec3> cams, dirs;
        for (int i = 0; i < 64; ++i) {
            const double D = 8.0 + 3.0 * i, b = 1.0 + 0.25 * i;
            cams.push_back(dvec3(0.1 * i, 0.5, D));
            dirs.push_back(dirFor(D, std::min(b, D * 0.9)));
        }
        const std::vector<TraceResult> batch = traceBatch(cams, dirs, 5000, 0.005);
        bool same = true;
        for (int i = 0; i < 64; ++i) {
            const TraceResult one = trace(cams[i], dirs[i], 5000, 0.005);
            same = same && one.status == batch[i].status && one.steps == batch[i].steps &&
                   one.equatorCrossings == batch[i].equatorCrossings &&
                   std::memcmp(&one.dir, &batch[i].dir, sizeof(dvec3)) == 0 &&
                   std::memcmp(&one.minR, &batch[i].minR, sizeof(double)) == 0;
        }
        check(same, "traceBatch equals trace bit for bit (64 rays)");
    }
    std::printf("\n%d checks, %d failures\n", checks, failures);
    return failures == 0 ? 0 : 1;
}

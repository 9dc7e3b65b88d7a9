// Minimal stand-in for <glm/glm.hpp> used only to build this repository's C++ tests in an image
// without GLM: dvec3 with component-wise arithmetic and dot/cross/length/normalize. Operation
// order follows GLM's scalar implementation: dot = (a.x*b.x + a.y*b.y) + a.z*b.z, length =
// sqrt(dot), normalize = v * (1/sqrt(dot(v,v))), vec/scalar divides each component.
#pragma once
#include <cmath>

namespace glm {

struct dvec3 {
    double x = 0.0, y = 0.0, z = 0.0;
    dvec3() = default;
    explicit dvec3(double s) : x(s), y(s), z(s) {}
    dvec3(double a, double b, double c) : x(a), y(b), z(c) {}
    dvec3& operator+=(const dvec3& o) { x += o.x; y += o.y; z += o.z; return *this; }
    dvec3& operator-=(const dvec3& o) { x -= o.x; y -= o.y; z -= o.z; return *this; }
};

inline dvec3 operator+(const dvec3& a, const dvec3& b) { return dvec3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline dvec3 operator-(const dvec3& a, const dvec3& b) { return dvec3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline dvec3 operator-(const dvec3& a) { return dvec3(-a.x, -a.y, -a.z); }
inline dvec3 operator*(const dvec3& a, double s) { return dvec3(a.x * s, a.y * s, a.z * s); }
inline dvec3 operator*(double s, const dvec3& a) { return dvec3(s * a.x, s * a.y, s * a.z); }
inline dvec3 operator*(const dvec3& a, const dvec3& b) { return dvec3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline dvec3 operator/(const dvec3& a, double s) { return dvec3(a.x / s, a.y / s, a.z / s); }

inline double dot(const dvec3& a, const dvec3& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline dvec3 cross(const dvec3& x, const dvec3& y) {
    return dvec3(x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y);
}
inline double length(const dvec3& v) { return std::sqrt(dot(v, v)); }
inline dvec3 normalize(const dvec3& v) { return v * (1.0 / std::sqrt(dot(v, v))); }

}  // namespace glm

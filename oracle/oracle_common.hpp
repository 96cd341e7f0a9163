// oracle_common.hpp - FP64 vector math of the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY.  The oracle is the checker for the HIP library:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it.  It is never the thing measured or shipped.
//
// Restates CudaPathTracing/rtmath.cuh:8-112 (double3 operators, Dot, Cross,
// Unit, mMin/mMax ternaries, INF = 999999.0) and AABB.cuh:3-40.  The
// expression order of every operator is kept so that FP64 results match
// the reference bit for bit when both are compiled without FMA contraction
// (this file is compiled with -ffp-contract=off).
#pragma once
#include <cmath>
#include <cstdint>

namespace orc {

constexpr double INF = 999999.0;                  // rtmath.cuh:8
constexpr double PI = 3.14159265358979323846;     // rtmath.cuh:9

template <class T> inline T mMin(T a, T b) { return a < b ? a : b; }  // rtmath.cuh:11
template <class T> inline T mMax(T a, T b) { return a > b ? a : b; }  // rtmath.cuh:12

struct d3 {
  double x, y, z;
};
inline d3 mk(double x, double y, double z) { return d3{x, y, z}; }
inline d3 operator+(const d3 &a, const d3 &b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline d3 operator-(const d3 &a, const d3 &b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline d3 operator-(const d3 &a) { return mk(-a.x, -a.y, -a.z); }
inline d3 operator*(const d3 &a, const d3 &b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
inline d3 operator*(double t, const d3 &v) { return mk(v.x * t, v.y * t, v.z * t); }
inline d3 operator*(const d3 &v, double t) { return mk(v.x * t, v.y * t, v.z * t); }
inline d3 operator/(const d3 &v, double t) { return mk(v.x / t, v.y / t, v.z / t); }
inline double Dot(const d3 &a, const d3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline d3 Cross(const d3 &a, const d3 &b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline double SquaredLength(const d3 &v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
inline double Length(const d3 &v) { return std::sqrt(SquaredLength(v)); }
inline d3 Unit(const d3 &v) { return v / Length(v); }

// AABB.cuh:3-40
struct AABB {
  d3 mn{INF, INF, INF};
  d3 mx{-INF, -INF, -INF};
  AABB() = default;
  AABB(const d3 &p0, const d3 &p1) {
    mn = mk(mMin(p0.x, p1.x), mMin(p0.y, p1.y), mMin(p0.z, p1.z));
    mx = mk(mMax(p0.x, p1.x), mMax(p0.y, p1.y), mMax(p0.z, p1.z));
  }
  AABB(const AABB &a, const AABB &b) {
    mn = mk(mMin(a.mn.x, b.mn.x), mMin(a.mn.y, b.mn.y), mMin(a.mn.z, b.mn.z));
    mx = mk(mMax(a.mx.x, b.mx.x), mMax(a.mx.y, b.mx.y), mMax(a.mx.z, b.mx.z));
  }
  AABB pad() const {  // AABB.cuh:30-40
    const double delta = 0.0001;
    d3 a = mn, b = mx;
    a.x -= delta; a.y -= delta; a.z -= delta;
    b.x += delta; b.y += delta; b.z += delta;
    return AABB(a, b);
  }
};

}  // namespace orc

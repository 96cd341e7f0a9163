// oracle_bvh.cpp - FP64 CPU restatement of the reference's primitives, LBVH
// build and closest-hit traversal.
//
// TEST INFRASTRUCTURE ONLY (see oracle_common.hpp): loaded by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
// library.
//
// Follows, function by function:
//   primitive construction   Hittable.cuh:359-365 (Light), 416-462 (Hittable
//                            ctors: centre, padded AABB), Device.cuh:40-44
//   expandBits / morton3D    BVH.cuh:12-41
//   bounds rule              Device.cuh:174-181
//   sort                     Device.cuh:187-189 (unique u64 keys)
//   delta/determineRange/findSplit/generateHierarchy  BVH.cuh:81-241
//   constructAABB            BVH.cuh:243-280 (exact min/max union)
//   checkSlab/checkOverlap   BVH.cuh:54-79
//   traverseIterative        BVH.cuh:282-332
//   Sphere/Triangle/Light::hit  Hittable.cuh:27-54, 85-140, 367-398
#include <algorithm>
#include <atomic>
#include <cstring>
#include <parallel/algorithm>
#include <vector>

#include "../include/cpt.h"
#include "oracle_common.hpp"
#include "oracle_scene.hpp"

using namespace orc;

namespace orc {

static inline unsigned expandBits(unsigned v) {  // BVH.cuh:12-19
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

static inline uint64_t morton3D(d3 v, d3 smin, d3 smax, int index) {  // BVH.cuh:23-41
  d3 n;
  n.x = (v.x - smin.x) / (smax.x - smin.x);
  n.y = (v.y - smin.y) / (smax.y - smin.y);
  n.z = (v.z - smin.z) / (smax.z - smin.z);
  double x = mMin(mMax(n.x * 1024.0, 0.0), 1023.0);
  double y = mMin(mMax(n.y * 1024.0, 0.0), 1023.0);
  double z = mMin(mMax(n.z * 1024.0, 0.0), 1023.0);
  unsigned xx = expandBits((unsigned)x);
  unsigned yy = expandBits((unsigned)y);
  unsigned zz = expandBits((unsigned)z);
  uint64_t morton = xx * 4 + yy * 2 + zz;
  return (morton << 32) | (uint64_t)(int64_t)index;
}

static void expandBounds(const double mnIn[3], const double mxIn[3], d3 &smin, d3 &smax) {
  // Device.cuh:174-181
  d3 e = mk(1.0, 1.0, 1.0);
  smin = mk(mnIn[0], mnIn[1], mnIn[2]) - e;
  smax = mk(mxIn[0], mxIn[1], mxIn[2]) + e;
  if (smin.x > smax.x) {
    smin = mk(-20.0, -20.0, -20.0);
    smax = mk(20.0, 20.0, 20.0);
  }
}

Scene::Scene(const cpt_material *mats, int nmat, const cpt_light *lights, int nl,
             const cpt_sphere *spheres, int ns, const double *tri, const int32_t *tri_mat,
             int64_t nt) {
  materials.assign(mats, mats + nmat);
  nLights = nl;
  nSpheres = ns;
  prims.resize((size_t)(nl + ns + nt));
  // Lights: Light ctor (Hittable.cuh:359-365) + Hittable(Light) (424-454).
  for (int i = 0; i < nl; i++) {
    const cpt_light &L = lights[i];
    Prim &p = prims[i];
    p.type = CPT_OBJ_LIGHT;
    p.c = mk(L.center[0], L.center[1], L.center[2]);
    p.w = L.width;
    p.h = L.height;
    p.n = Unit(mk(L.normal[0], L.normal[1], L.normal[2]));
    p.eu = std::fabs(p.n.x) > 0.9 ? Unit(Cross(mk(0, 1, 0), p.n)) : Unit(Cross(mk(1, 0, 0), p.n));
    p.ev = Unit(Cross(p.n, p.eu));
    p.visible = L.visible != 0;
    p.le = mk(L.color[0], L.color[1], L.color[2]);
    p.mat = -1 - i;  // implicit light material
    d3 N = Unit(p.n);
    d3 T = std::fabs(N.x) > 0.9 ? Unit(Cross(mk(0, 1, 0), N)) : Unit(Cross(mk(1, 0, 0), N));
    d3 B = Unit(Cross(N, T));
    d3 corners[4] = {p.c + T * (p.w / 2) + B * (p.h / 2), p.c + T * (p.w / 2) + B * (-p.h / 2),
                     p.c + T * (-p.w / 2) + B * (p.h / 2), p.c + T * (-p.w / 2) + B * (-p.h / 2)};
    d3 a = corners[0], b = corners[0];
    for (int k = 1; k < 4; ++k) {
      a.x = mMin(a.x, corners[k].x); a.y = mMin(a.y, corners[k].y); a.z = mMin(a.z, corners[k].z);
      b.x = mMax(b.x, corners[k].x); b.y = mMax(b.y, corners[k].y); b.z = mMax(b.z, corners[k].z);
    }
    p.box = AABB(a, b).pad();
    p.center = p.c;
  }
  // Spheres: Hittable(Sphere) (Hittable.cuh:418-422), no padding.
  for (int i = 0; i < ns; i++) {
    const cpt_sphere &S = spheres[i];
    Prim &p = prims[nl + i];
    p.type = CPT_OBJ_SPHERE;
    p.c = mk(S.center[0], S.center[1], S.center[2]);
    p.r = S.radius;
    p.mat = S.material;
    p.box = AABB(p.c - mk(p.r, p.r, p.r), p.c + mk(p.r, p.r, p.r));
    p.center = p.c;
  }
  // Triangles: Hittable(Triangle) (Hittable.cuh:456-462).
  const int64_t base = nl + ns;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nt; i++) {
    Prim &p = prims[base + i];
    const double *q = tri + 9 * i;
    p.type = CPT_OBJ_TRIANGLE;
    p.p0 = mk(q[0], q[1], q[2]);
    p.p1 = mk(q[3], q[4], q[5]);
    p.p2 = mk(q[6], q[7], q[8]);
    p.mat = tri_mat ? tri_mat[i] : 0;
    p.center = (p.p0 + p.p1 + p.p2) / 3.0;
    p.box = AABB(mk(mMin(mMin(p.p0.x, p.p1.x), p.p2.x), mMin(mMin(p.p0.y, p.p1.y), p.p2.y),
                    mMin(mMin(p.p0.z, p.p1.z), p.p2.z)),
                 mk(mMax(mMax(p.p0.x, p.p1.x), p.p2.x), mMax(mMax(p.p0.y, p.p1.y), p.p2.y),
                    mMax(mMax(p.p0.z, p.p1.z), p.p2.z)))
                 .pad();
  }
}

// ---- primitive hit tests (FP64, reference expression order) -------------

// Triangle::hit, Hittable.cuh:85-140.  Returns t via tOut; also writes the
// outward normal (face normal, vertex normals are never set from ini).
bool hitTriangle(const Prim &p, const d3 &o, const d3 &d, double tmin, double tmax, double &tOut,
                 double *uOut, double *vOut) {
  const double EPSILON = 1e-8;
  d3 e1 = p.p1 - p.p0;
  d3 e2 = p.p2 - p.p0;
  d3 h = Cross(d, e2);
  double a = Dot(e1, h);
  if (std::fabs(a) < EPSILON) return false;
  double f = 1.0 / a;
  d3 s = o - p.p0;
  double u = f * Dot(s, h);
  if (u < 0.0 || u > 1.0) return false;
  d3 q = Cross(s, e1);
  double v = f * Dot(d, q);
  if (v < 0.0 || u + v > 1.0) return false;
  double t = f * Dot(e2, q);
  if (t < tmin || t > tmax) return false;
  tOut = t;
  if (uOut) *uOut = u;
  if (vOut) *vOut = v;
  return true;
}

// Sphere::hit, Hittable.cuh:27-54.
bool hitSphere(const Prim &p, const d3 &o, const d3 &d, double tmin, double tmax, double &tOut) {
  double a = Dot(d, d);
  double b = Dot(d, o - p.c);
  double c = Dot(o - p.c, o - p.c) - p.r * p.r;
  double delta = b * b - a * c;
  if (delta < 0.0) return false;
  double t = (-b - std::sqrt(delta)) / a;
  if (t < tmin || t > tmax) t = (-b + std::sqrt(delta)) / a;
  if (t < tmin || t > tmax) return false;
  tOut = t;
  return true;
}

// Light::hit, Hittable.cuh:367-398.
bool hitLight(const Prim &p, const d3 &o, const d3 &d, double tmin, double tmax, double &tOut) {
  if (!p.visible) return false;
  double denom = Dot(d, p.n);
  if (std::fabs(denom) < 1e-6 || denom > 0.0) return false;
  double t = Dot(p.c - o, p.n) / denom;
  if (t < tmin || t > tmax) return false;
  d3 hp = o + t * d;
  d3 diff = hp - p.c;
  double u = Dot(diff, p.eu);
  double v = Dot(diff, p.ev);
  if (std::fabs(u) <= p.w / 2.0 && std::fabs(v) <= p.h / 2.0) {
    tOut = t;
    return true;
  }
  return false;
}

bool hitPrim(const Prim &p, const d3 &o, const d3 &d, double tmin, double tmax, double &t) {
  switch (p.type) {  // Hittable::hit, Hittable.cuh:487-502
    case CPT_OBJ_SPHERE: return hitSphere(p, o, d, tmin, tmax, t);
    case CPT_OBJ_LIGHT: return hitLight(p, o, d, tmin, tmax, t);
    case CPT_OBJ_TRIANGLE: return hitTriangle(p, o, d, tmin, tmax, t, nullptr, nullptr);
    default: return false;
  }
}

// ---- LBVH ------------------------------------------------------------------

static inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

static inline int delta(const uint64_t *k, int n, int i, int j) {  // BVH.cuh:120-128
  if (j < 0 || j >= n) return -1;
  uint64_t a = k[i], b = k[j];
  if (a == b) return 64;
  return clz64(a ^ b);
}

static inline void determineRange(const uint64_t *k, int n, int idx, int &first, int &last) {
  // BVH.cuh:130-172
  int dNext = delta(k, n, idx, idx + 1);
  int dPrev = delta(k, n, idx, idx - 1);
  int d = (dNext > dPrev) ? 1 : -1;
  int dMin = delta(k, n, idx, idx - d);
  int64_t lMax = 2;
  while (delta(k, n, idx, (int)std::max<int64_t>(std::min<int64_t>(idx + lMax * d, (int64_t)n), -1)) >
         dMin)
    lMax *= 2;
  int64_t l = 0;
  for (int64_t t = lMax / 2; t >= 1; t /= 2) {
    int64_t j = idx + (l + t) * d;
    if (j >= 0 && j < n && delta(k, n, idx, (int)j) > dMin) l += t;
  }
  int j = (int)(idx + l * d);
  first = mMin(idx, j);
  last = mMax(idx, j);
}

static inline int findSplit(const uint64_t *k, int first, int last) {  // BVH.cuh:81-118
  uint64_t fc = k[first], lc = k[last];
  if (fc == lc) return (first + last) >> 1;
  int common = clz64(fc ^ lc);
  int split = first;
  int step = last - first;
  do {
    step = (step + 1) >> 1;
    int ns = split + step;
    if (ns < last) {
      int sp = clz64(fc ^ k[ns]);
      if (sp > common) split = ns;
    }
  } while (step > 1);
  return split;
}

void Scene::build(const double bmin[3], const double bmax[3]) {
  const int64_t n = (int64_t)prims.size();
  d3 smin, smax;
  expandBounds(bmin, bmax, smin, smax);
  keys.resize(n);
  order.resize(n);
  // generateMortonCodes, Device.cuh:134-142
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) keys[i] = morton3D(prims[i].center, smin, smax, (int)i);
  // thrust::sort_by_key, Device.cuh:187-189: keys are unique, so the sorted
  // order of (key) alone fixes the permutation; the low word is the index.
  __gnu_parallel::sort(keys.begin(), keys.end());
  for (int64_t i = 0; i < n; i++) order[i] = (uint32_t)(keys[i] & 0xffffffffu);
  // generateHierarchy, BVH.cuh:174-241
  childA.assign(n > 1 ? n - 1 : 0, 0);
  childB.assign(n > 1 ? n - 1 : 0, 0);
  std::vector<int64_t> parent(2 * n, -1);  // [0,n-1): internal, [n-1, 2n-1): leaves
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n - 1; i++) {
    int first, last;
    determineRange(keys.data(), (int)n, (int)i, first, last);
    int split = findSplit(keys.data(), first, last);
    childA[i] = (split == first) ? ~split : split;
    childB[i] = (split + 1 == last) ? ~(split + 1) : (split + 1);
  }
  for (int64_t i = 0; i < n - 1; i++) {
    int a = childA[i], b = childB[i];
    parent[a < 0 ? (n - 1) + ~a : a] = i;
    parent[b < 0 ? (n - 1) + ~b : b] = i;
  }
  // constructAABB, BVH.cuh:243-280: union is exact, so the post-order of the
  // reference and this bottom-up pass give identical boxes.
  nodeBox.assign(n > 1 ? n - 1 : 0, AABB());
  std::vector<std::atomic<int>> flag(n > 1 ? n - 1 : 1);
  for (auto &f : flag) f.store(0);
  maxDepth = 0;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    int64_t node = parent[(n - 1) + i];
    while (node >= 0) {
      if (flag[node].fetch_add(1, std::memory_order_acq_rel) == 0) break;
      int a = childA[node], b = childB[node];
      const AABB &ba = a < 0 ? prims[order[~a]].box : nodeBox[a];
      const AABB &bb = b < 0 ? prims[order[~b]].box : nodeBox[b];
      nodeBox[node] = AABB(ba, bb);
      node = parent[node];
    }
  }
  if (n == 1) rootBox = prims[order[0]].box;
  else if (n > 1) rootBox = nodeBox[0];
  // depth (root = 0), used by tests for the stack bound
  if (n > 1) {
    std::vector<std::pair<int, int>> st;
    st.push_back({0, 0});
    while (!st.empty()) {
      auto [nd, dep] = st.back();
      st.pop_back();
      maxDepth = std::max(maxDepth, dep);
      if (childA[nd] >= 0) st.push_back({childA[nd], dep + 1});
      if (childB[nd] >= 0) st.push_back({childB[nd], dep + 1});
    }
  }
  built = true;
}

// checkSlab / checkOverlap, BVH.cuh:54-79
static inline bool checkSlab(double o, double inv, double mn, double mx, double &t0r, double &t1r) {
  double t0 = (mn - o) * inv;
  double t1 = (mx - o) * inv;
  if (inv < 0.0) std::swap(t0, t1);
  t0r = std::fmax(t0, t0r);
  t1r = std::fmin(t1, t1r);
  return t1r > t0r;
}
static inline bool checkOverlap(const d3 &o, const d3 &d, const AABB &b, double tmin, double tmax) {
  const d3 inv = mk(1.0 / d.x, 1.0 / d.y, 1.0 / d.z);
  return checkSlab(o.x, inv.x, b.mn.x, b.mx.x, tmin, tmax) &&
         checkSlab(o.y, inv.y, b.mn.y, b.mx.y, tmin, tmax) &&
         checkSlab(o.z, inv.z, b.mn.z, b.mx.z, tmin, tmax);
}

// traverseIterative, BVH.cuh:282-332: root box not tested, child boxes
// tested against [tMin, tMax] with no culling by the best hit, leaves tested
// on [0.001 (= tMin here), best t] inclusive, the last accepted leaf wins.
int Scene::traceRef(const d3 &o, const d3 &d, double tmin, double tmax, double &tBest,
                    int64_t *nodesVisited) const {
  const int64_t n = (int64_t)prims.size();
  int hitId = -1;
  tBest = tmax;
  if (n == 0) return -1;
  if (n == 1) {  // no internal node: test the single primitive
    double t;
    if (hitPrim(prims[0], o, d, tmin, tmax, t)) { tBest = t; return 0; }
    return -1;
  }
  int stack[128];
  int sp = 0;
  stack[sp++] = -1;  // NULL sentinel
  int node = 0;
  int64_t visits = 0;
  do {
    visits++;
    int cl = childA[node], cr = childB[node];
    const AABB &bl = cl < 0 ? prims[order[~cl]].box : nodeBox[cl];
    const AABB &br = cr < 0 ? prims[order[~cr]].box : nodeBox[cr];
    bool ovL = checkOverlap(o, d, bl, tmin, tmax);
    bool ovR = checkOverlap(o, d, br, tmin, tmax);
    double t;
    if (ovL && cl < 0) {
      int id = (int)order[~cl];
      if (hitPrim(prims[id], o, d, tmin, hitId >= 0 ? tBest : tmax, t)) { hitId = id; tBest = t; }
    }
    if (ovR && cr < 0) {
      int id = (int)order[~cr];
      if (hitPrim(prims[id], o, d, tmin, hitId >= 0 ? tBest : tmax, t)) { hitId = id; tBest = t; }
    }
    bool tL = ovL && cl >= 0, tR = ovR && cr >= 0;
    if (!tL && !tR) node = stack[--sp];
    else {
      node = tL ? cl : cr;
      if (tL && tR) stack[sp++] = cr;
    }
  } while (node != -1);
  if (nodesVisited) *nodesVisited = visits;
  if (hitId < 0) tBest = tmax;
  return hitId;
}

// Robustness of tmin <= t <= tmax for a surface crossing at t: +1 robustly
// inside, -1 robustly outside, 0 ambiguous.  The tolerance is eb relative to
// t plus `shift`, the move of the crossing under the origin / direction
// perturbation (see classify).
static int rangeRobust(double t, double tmin, double tmax, double eb, double shift) {
  const double m = std::min(t - tmin, tmax - t);
  const double tol = eb * std::fabs(t) + shift;
  if (m > tol) return 1;
  if (m < -tol) return -1;
  return 0;
}

// Robustness of one primitive's FP64 decision against small perturbations
// (used to certify rays for FP32 comparison).  Returns +1 robust accept,
// -1 robust reject, 0 ambiguous; tOut = plane/quadric distance.
// eb: relative margin of every decision.  eo, ed: absolute uncertainty of
// the ray origin and relative uncertainty of its direction (0 for rays
// given exactly to both sides; a path's rays start at FP32 hit points):
// a crossing at t with unit surface normal n moves by
// (eo + |t| |d| ed) / |d . n|, large for grazing rays, and its point on the
// surface by |d| times that.
static int classify(const Prim &p, const d3 &o, const d3 &d, double tmin, double tmax, double eb, double eo,
                    double ed, double &tOut, double *shiftOut = nullptr) {
  const double dl = Length(d);
  auto shiftAt = [&](double t, const d3 &nUnit) {
    const double dn = std::fabs(Dot(d, nUnit));
    return (eo + std::fabs(t) * dl * ed) / std::max(dn, 1e-300);
  };
  if (p.type == CPT_OBJ_TRIANGLE) {
    d3 e1 = p.p1 - p.p0, e2 = p.p2 - p.p0;
    d3 h = Cross(d, e2);
    double a = Dot(e1, h);
    if (a == 0.0) return -1;
    double f = 1.0 / a;
    d3 s = o - p.p0;
    double u = f * Dot(s, h);
    d3 q = Cross(s, e1);
    double v = f * Dot(d, q);
    double t = f * Dot(e2, q);
    tOut = t;
    const d3 nrm = Cross(e1, e2);
    const double nl = Length(nrm);
    const double shift = nl > 0.0 ? shiftAt(t, nrm / nl) : INFINITY;
    if (shiftOut) *shiftOut = shift;
    // barycentric margin: eb plus the move of the crossing point over the
    // shortest edge
    const double emin = std::min(std::min(Length(e1), Length(e2)), Length(p.p2 - p.p1));
    const double mb = eb + (shift * dl) / std::max(emin, 1e-300);
    double m = std::min(std::min(u, v), 1.0 - u - v);
    const int rr = rangeRobust(t, tmin, tmax, eb, shift);
    // a robust barycentric or range rejection holds whatever the det test says
    if (m < -mb || u > 1.0 + mb || rr < 0) return -1;
    double scale = Length(e1) * Length(e2) * dl;
    if (std::fabs(a) < 0.5e-8) return -1;
    if (std::fabs(a) < 2e-8 || std::fabs(a) < 1e-6 * scale) return 0;
    if (m > mb && rr > 0) return 1;
    return 0;
  }
  if (p.type == CPT_OBJ_SPHERE) {
    double a = Dot(d, d), b = Dot(d, o - p.c), c = Dot(o - p.c, o - p.c) - p.r * p.r;
    double disc = b * b - a * c;
    double sc = std::max(b * b, std::fabs(a * c));
    if (disc < -eb * sc) return -1;
    if (disc <= eb * sc) return 0;
    double s = std::sqrt(disc);
    double t0 = (-b - s) / a, t1 = (-b + s) / a;
    auto rob = [&](double t) { return rangeRobust(t, tmin, tmax, eb, shiftAt(t, ((o + t * d) - p.c) / p.r)); };
    const int r0 = rob(t0), r1 = rob(t1);
    if (shiftOut) *shiftOut = shiftAt(r0 > 0 ? t0 : t1, ((o + (r0 > 0 ? t0 : t1) * d) - p.c) / p.r);
    if (r0 > 0) { tOut = t0; return 1; }
    if (r0 == 0) return 0;
    if (r1 > 0) { tOut = t1; return 1; }
    if (r1 == 0) return 0;
    return -1;
  }
  if (p.type == CPT_OBJ_LIGHT) {
    if (!p.visible) return -1;
    double denom = Dot(d, p.n);
    // the 1e-6 threshold is absolute, so its margin is the FP32 rounding of
    // the dot product (a few ulp of |d|) and the direction's own uncertainty,
    // not the relative margin eb
    const double de = 5e-7 * dl + ed * dl;
    if (denom > -(1e-6 - de)) return -1;  // back face, parallel or near-parallel
    if (denom >= -(1e-6 + de)) return 0;
    double t = Dot(p.c - o, p.n) / denom;
    tOut = t;
    const double shift = shiftAt(t, p.n);
    if (shiftOut) *shiftOut = shift;
    d3 diff = (o + t * d) - p.c;
    double u = std::fabs(Dot(diff, p.eu)) - p.w / 2.0;
    double v = std::fabs(Dot(diff, p.ev)) - p.h / 2.0;
    double scale = std::max(p.w, p.h);
    const int rr = rangeRobust(t, tmin, tmax, eb, shift);
    double m = std::max(u, v) / scale;
    const double mb = eb + shift * dl / scale;
    if (m > mb || rr < 0) return -1;
    if (m < -mb && rr > 0) return 1;
    return 0;
  }
  return -1;
}

// Closest hit over all primitives (the last accepted primitive among equal
// distances wins, as in the reference's leaf order up to ties) with a
// certificate: ok stays true only when every primitive's accept/reject
// decision is robust at relative margin eb, origin uncertainty eo and
// direction uncertainty ed (classify) and the winner is separated from the
// runner-up and from every ambiguous candidate by more than et (relative).
// Certified rays have one FP32/FP64-stable answer.
int Scene::bruteClosest(const d3 &o, const d3 &d, double tmin, double tmax, double eb, double et,
                        double &tOut, bool &ok, double eo, double ed, double *tErrOut) const {
  const int64_t np = (int64_t)prims.size();
  double best = tmax, second = INFINITY;
  int id = -1;
  double amb_t = INFINITY;  // nearest ambiguous candidate
  for (int64_t k = 0; k < np; k++) {
    double t = 0;
    const int c = classify(prims[k], o, d, tmin, tmax, eb, eo, ed, t);
    if (c == 0) amb_t = std::min(amb_t, std::fabs(t) > 0 ? t : 0.0);
    double th;
    if (hitPrim(prims[k], o, d, tmin, id >= 0 ? best : tmax, th)) {
      if (id >= 0) second = std::min(second, best);
      best = th;
      id = (int)k;
    } else if (c == 1) {
      second = std::min(second, t);
    }
  }
  if (id >= 0) {
    if (second < INFINITY && (second - best) <= et * std::fabs(best)) ok = false;
    if (amb_t <= best * (1.0 + et)) ok = false;
    double tt, sh = 0.0;
    if (classify(prims[id], o, d, tmin, tmax, eb, eo, ed, tt, &sh) != 1) ok = false;
    if (tErrOut) *tErrOut = eb * std::fabs(best) + sh;
  } else if (amb_t < INFINITY) {
    ok = false;
  }
  tOut = id >= 0 ? best : tmax;
  return id;
}

}  // namespace orc

// ---- C API for tests --------------------------------------------------------
extern "C" {

void *orc_scene_create(const cpt_material *mats, int nmat, const cpt_light *lights, int nl,
                       const cpt_sphere *spheres, int ns, const double *tri,
                       const int32_t *tri_mat, int64_t nt) {
  return new Scene(mats, nmat, lights, nl, spheres, ns, tri, tri_mat, nt);
}
void orc_scene_destroy(void *s) { delete (Scene *)s; }
int64_t orc_scene_count(void *s) { return (int64_t)((Scene *)s)->prims.size(); }

unsigned orc_expand_bits(unsigned v) { return expandBits(v); }
uint64_t orc_morton3d(const double c[3], const double smin[3], const double smax[3], int index) {
  return morton3D(mk(c[0], c[1], c[2]), mk(smin[0], smin[1], smin[2]),
                  mk(smax[0], smax[1], smax[2]), index);
}
void orc_expand_bounds(const double mn[3], const double mx[3], double omn[3], double omx[3]) {
  d3 a, b;
  expandBounds(mn, mx, a, b);
  omn[0] = a.x; omn[1] = a.y; omn[2] = a.z;
  omx[0] = b.x; omx[1] = b.y; omx[2] = b.z;
}

// centers[3n], boxes[6n] (min xyz, max xyz) of every primitive
void orc_scene_prims(void *sp, double *centers, double *boxes) {
  Scene *s = (Scene *)sp;
  for (size_t i = 0; i < s->prims.size(); i++) {
    const Prim &p = s->prims[i];
    if (centers) { centers[3 * i] = p.center.x; centers[3 * i + 1] = p.center.y; centers[3 * i + 2] = p.center.z; }
    if (boxes) {
      double *b = boxes + 6 * i;
      b[0] = p.box.mn.x; b[1] = p.box.mn.y; b[2] = p.box.mn.z;
      b[3] = p.box.mx.x; b[4] = p.box.mx.y; b[5] = p.box.mx.z;
    }
  }
}

void orc_morton_keys(void *sp, const double bmin[3], const double bmax[3], uint64_t *keys) {
  Scene *s = (Scene *)sp;
  d3 smin, smax;
  expandBounds(bmin, bmax, smin, smax);
  const int64_t n = (int64_t)s->prims.size();
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) keys[i] = morton3D(s->prims[i].center, smin, smax, (int)i);
}

// Builds the tree and copies it out (any pointer may be NULL).
int orc_build(void *sp, const double bmin[3], const double bmax[3], uint64_t *keys,
              uint32_t *order, int32_t *ca, int32_t *cb, double *boxes, double *rootBox,
              int *maxDepth) {
  Scene *s = (Scene *)sp;
  s->build(bmin, bmax);
  const int64_t n = (int64_t)s->prims.size();
  if (keys) std::memcpy(keys, s->keys.data(), n * 8);
  if (order) std::memcpy(order, s->order.data(), n * 4);
  if (ca && n > 1) std::memcpy(ca, s->childA.data(), (n - 1) * 4);
  if (cb && n > 1) std::memcpy(cb, s->childB.data(), (n - 1) * 4);
  if (boxes) {
    for (int64_t i = 0; i + 1 < n; i++) {
      for (int c = 0; c < 2; c++) {
        int ch = c == 0 ? s->childA[i] : s->childB[i];
        const AABB &b = ch < 0 ? s->prims[s->order[~ch]].box : s->nodeBox[ch];
        double *o = boxes + 12 * i + 6 * c;
        o[0] = b.mn.x; o[1] = b.mn.y; o[2] = b.mn.z; o[3] = b.mx.x; o[4] = b.mx.y; o[5] = b.mx.z;
      }
    }
  }
  if (rootBox) {
    rootBox[0] = s->rootBox.mn.x; rootBox[1] = s->rootBox.mn.y; rootBox[2] = s->rootBox.mn.z;
    rootBox[3] = s->rootBox.mx.x; rootBox[4] = s->rootBox.mx.y; rootBox[5] = s->rootBox.mx.z;
  }
  if (maxDepth) *maxDepth = s->maxDepth;
  return 0;
}

// Reference-semantics traversal for n rays {o,tmin,d,tmax} (FP32 values,
// promoted to FP64).  Multithreaded over rays.
int orc_trace_ref(void *sp, const float *rays, int64_t n, double *tOut, int32_t *idOut,
                  int64_t *visits) {
  Scene *s = (Scene *)sp;
  if (!s->built) return -1;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; i++) {
    const float *r = rays + 8 * i;
    d3 o = mk(r[0], r[1], r[2]), d = mk(r[4], r[5], r[6]);
    double t;
    int64_t v = 0;
    int id = s->traceRef(o, d, r[3], r[7], t, &v);
    if (tOut) tOut[i] = t;
    if (idOut) idOut[i] = id;
    if (visits) visits[i] = v;
  }
  return 0;
}

// Brute-force closest hit with certification: cert[i] = 1 when every
// primitive's accept/reject decision is robust at relative margin eb and the
// best hit is separated from the runner-up by more than et (relative).
int orc_trace_brute(void *sp, const float *rays, int64_t n, double eb, double et, double *tOut,
                    int32_t *idOut, uint8_t *cert) {
  Scene *s = (Scene *)sp;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < n; i++) {
    const float *r = rays + 8 * i;
    d3 o = mk(r[0], r[1], r[2]), d = mk(r[4], r[5], r[6]);
    bool ok = true;
    double best = 0;
    const int id = s->bruteClosest(o, d, r[3], r[7], eb, et, best, ok);
    if (tOut) tOut[i] = best;
    if (idOut) idOut[i] = id;
    if (cert) cert[i] = ok ? 1 : 0;
  }
  return 0;
}

// FP64 hit distance of primitive `id` along each ray, with no range checks
// applied beyond the primitive's own (returns NaN where the FP64 test
// rejects).  Used to check t of mismatched ids on large scenes.
int orc_prim_t(void *sp, const float *rays, const int32_t *ids, int64_t n, double *tOut) {
  Scene *s = (Scene *)sp;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    const float *r = rays + 8 * i;
    d3 o = mk(r[0], r[1], r[2]), d = mk(r[4], r[5], r[6]);
    double t = NAN;
    if (ids[i] >= 0 && ids[i] < (int64_t)s->prims.size()) {
      double th;
      if (hitPrim(s->prims[ids[i]], o, d, -INFINITY, INFINITY, th)) t = th;
    }
    tOut[i] = t;
  }
  return 0;
}

}  // extern "C"

// cpt_device.hpp - device-side data layout and the closest-hit traversal.
//
// HBM layout of a built scene (N primitives):
//   nodes   : (N-1) x 64 B, Karras numbering (node 0 = root).  float4 x4:
//             [0] = A.min.x A.max.x A.min.y A.max.y
//             [1] = B.min.x B.max.x B.min.y B.max.y
//             [2] = A.min.z A.max.z B.min.z B.max.z
//             [3] = refA refB 0 0 (int bits; ref >= 0 internal, ~k = leaf k)
//   leaves  : N x 48 B in leaf (sorted Morton) order.  float4 x3:
//             triangle: {v0, id} {v1 - v0, type} {v2 - v0, material} (the
//                       edges are the FP32 differences the test would form)
//             sphere  : {centre, id} {radius 0 0, type} {0 0 0, material}
//             light   : {centre, id} {normal, type} {width height visible, mat}
//   rec     : the same records in insertion order, triangles with v1 and v2
// Child boxes are FP32, rounded outward from the reference's FP64 padded
// boxes (Hittable.cuh:421,453,460-461), so every FP32 box contains the FP64
// box the reference tests.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cpt.h"

namespace cpt {

// traversal stack entries kept in LDS per lane: 14, so that with the 6
// queued-leaf words a 256-lane workgroup of the four-wide walks uses 20 KB of
// LDS and 8 workgroups fit a CU's 160 KB (8 waves per SIMD where the
// registers allow)
constexpr int kStackLds = 14;
constexpr int kStackTotal = 64;  // Karras depth <= 62 for unique 64-bit keys

struct f3 {
  float x, y, z;
};
__host__ __device__ __forceinline__ f3 mkf3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mkf3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mkf3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator-(f3 a) { return mkf3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return mkf3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return mkf3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return mkf3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return __builtin_fmaf(a.x, b.x, __builtin_fmaf(a.y, b.y, a.z * b.z)); }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
  return mkf3(__builtin_fmaf(a.y, b.z, -a.z * b.y), __builtin_fmaf(a.z, b.x, -a.x * b.z),
              __builtin_fmaf(a.x, b.y, -a.y * b.x));
}
__device__ __forceinline__ float len3(f3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ f3 unit3(f3 a) { return a * (1.0f / len3(a)); }
__device__ __forceinline__ f3 xyz(float4 v) { return mkf3(v.x, v.y, v.z); }

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) --------------------
struct u4 {
  uint32_t x, y, z, w;
};
__host__ __device__ __forceinline__ u4 philox4x32_10(u4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)M0 * c.x;
    uint64_t p1 = (uint64_t)M1 * c.z;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c = u4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// Sample stream: draw k of (pixel p, sample s) is word k&3 of
// philox(ctr = {k>>2, p, s, 0x43505431}, key = seed), mapped to
// ((x >> 8) + 1) * 2^-24 in (0, 1] -- exact in both FP32 and FP64.
// A shadow ray's transmittance stream uses the tag 0x80000000 | (path draw
// count when the shadow ray was made) instead of 'CPT1'.
struct Rng {
  uint32_t k0, k1, pix, smp, cnt, tag;
  uint32_t blk;
  u4 cache;
  __device__ __forceinline__ void init(uint64_t seed, uint32_t p, uint32_t s, uint32_t c,
                                       uint32_t t = 0x43505431u) {
    k0 = (uint32_t)seed;
    k1 = (uint32_t)(seed >> 32);
    pix = p;
    smp = s;
    cnt = c;
    tag = t;
    blk = 0xffffffffu;
  }
  __device__ __forceinline__ float next() {
    uint32_t b = cnt >> 2;
    if (b != blk) {
      cache = philox4x32_10(u4{b, pix, smp, tag}, k0, k1);
      blk = b;
    }
    uint32_t w = cnt & 3;
    uint32_t x = w == 0 ? cache.x : (w == 1 ? cache.y : (w == 2 ? cache.z : cache.w));
    cnt++;
    return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f);
  }
};

// Per-axis 1/d for the slab tests: zero direction components are replaced
// by a tiny value so that the slab products stay finite (the primitive tests
// use the true direction), and the reciprocal is the hardware one (1 ulp; a
// slab distance only has to stay within the boxes' 1e-4 padding of the true
// one, and the sign, which picks the near planes, is exact).
__device__ __forceinline__ f3 slabInv(f3 d) {
  const float eps = 1e-20f;
  const f3 dd = mkf3(fabsf(d.x) < eps ? copysignf(eps, d.x) : d.x, fabsf(d.y) < eps ? copysignf(eps, d.y) : d.y,
                     fabsf(d.z) < eps ? copysignf(eps, d.z) : d.z);
  return mkf3(__builtin_amdgcn_rcpf(dd.x), __builtin_amdgcn_rcpf(dd.y), __builtin_amdgcn_rcpf(dd.z));
}

// ---- FP32 primitive tests --------------------------------------------------

// Triangle::hit (Hittable.cuh:85-140) in FP32 with the reference's accept
// rules: |det| < 1e-8 misses, u in [0,1], v >= 0, u+v <= 1, t in [tmin,tmax]
// inclusive.  e1 = v1 - v0, e2 = v2 - v0.
__device__ __forceinline__ bool hitTriE(f3 v0, f3 e1, f3 e2, f3 o, f3 d, float tmin, float tmax,
                                        float &t) {
  f3 h = cross3(d, e2);
  float a = dot3(e1, h);
  if (fabsf(a) < 1e-8f) return false;
  float f = __builtin_amdgcn_rcpf(a);
  f3 s = o - v0;
  float u = f * dot3(s, h);
  if (u < 0.0f || u > 1.0f) return false;
  f3 q = cross3(s, e1);
  float v = f * dot3(d, q);
  if (v < 0.0f || u + v > 1.0f) return false;
  float tt = f * dot3(e2, q);
  if (tt < tmin || tt > tmax) return false;
  t = tt;
  return true;
}

// Sphere::hit (Hittable.cuh:27-54).
__device__ __forceinline__ bool hitSph(f3 c, float r, f3 o, f3 d, float tmin, float tmax, float &t) {
  f3 oc = o - c;
  float a = dot3(d, d);
  float b = dot3(d, oc);
  float cc = dot3(oc, oc) - r * r;
  float disc = b * b - a * cc;
  if (disc < 0.0f) return false;
  float sq = sqrtf(disc);
  float tt = (-b - sq) / a;
  if (tt < tmin || tt > tmax) tt = (-b + sq) / a;
  if (tt < tmin || tt > tmax) return false;
  t = tt;
  return true;
}

// Light frame: edgeU as in the Light constructor (Hittable.cuh:363-364).
__device__ __forceinline__ void lightFrame(f3 n, f3 &eu, f3 &ev) {
  eu = fabsf(n.x) > 0.9f ? unit3(cross3(mkf3(0, 1, 0), n)) : unit3(cross3(mkf3(1, 0, 0), n));
  ev = unit3(cross3(n, eu));
}

// Light::hit (Hittable.cuh:367-398): one-sided, invisible lights never hit.
__device__ __forceinline__ bool hitQuad(f3 c, f3 n, float w, float hgt, bool vis, f3 o, f3 d,
                                        float tmin, float tmax, float &t) {
  if (!vis) return false;
  float denom = dot3(d, n);
  if (fabsf(denom) < 1e-6f || denom > 0.0f) return false;
  float tt = dot3(c - o, n) / denom;
  if (tt < tmin || tt > tmax) return false;
  f3 eu, ev;
  lightFrame(n, eu, ev);
  f3 diff = (o + tt * d) - c;
  float u = dot3(diff, eu), v = dot3(diff, ev);
  if (fabsf(u) <= w * 0.5f && fabsf(v) <= hgt * 0.5f) {
    t = tt;
    return true;
  }
  return false;
}

// A leaf-order record {r0, r1, r2} (triangles carry edges).
__device__ __forceinline__ bool hitRecord(float4 r0, float4 r1, float4 r2, f3 o, f3 d, float tmin, float tmax,
                                          float &t) {
  const int type = __float_as_int(r1.w);
  if (type == CPT_OBJ_TRIANGLE) return hitTriE(xyz(r0), xyz(r1), xyz(r2), o, d, tmin, tmax, t);
  if (type == CPT_OBJ_SPHERE) return hitSph(xyz(r0), r1.x, o, d, tmin, tmax, t);
  if (type == CPT_OBJ_LIGHT)
    return hitQuad(xyz(r0), xyz(r1), r2.x, r2.y, r2.z != 0.0f, o, d, tmin, tmax, t);
  return false;
}

// Primitive li of the leaf-order records (triangles carry edges).
__device__ __forceinline__ bool hitLeaf(const float4 *__restrict__ leaves, int li, f3 o, f3 d,
                                        float tmin, float tmax, float &t) {
  const float4 r0 = leaves[3 * li + 0];
  const float4 r1 = leaves[3 * li + 1];
  const float4 r2 = leaves[3 * li + 2];
  const int type = __float_as_int(r1.w);
  if (type == CPT_OBJ_TRIANGLE) return hitTriE(xyz(r0), xyz(r1), xyz(r2), o, d, tmin, tmax, t);
  if (type == CPT_OBJ_SPHERE) return hitSph(xyz(r0), r1.x, o, d, tmin, tmax, t);
  if (type == CPT_OBJ_LIGHT)
    return hitQuad(xyz(r0), xyz(r1), r2.x, r2.y, r2.z != 0.0f, o, d, tmin, tmax, t);
  return false;
}

// Primitive i of the insertion-order records (triangles carry vertices).
__device__ __forceinline__ bool hitRec(const float4 *__restrict__ rec, int i, f3 o, f3 d, float tmin, float tmax,
                                       float &t) {
  const float4 r0 = rec[3 * i + 0];
  const float4 r1 = rec[3 * i + 1];
  const float4 r2 = rec[3 * i + 2];
  const int type = __float_as_int(r1.w);
  if (type == CPT_OBJ_TRIANGLE) {
    const f3 v0 = xyz(r0);
    return hitTriE(v0, xyz(r1) - v0, xyz(r2) - v0, o, d, tmin, tmax, t);
  }
  if (type == CPT_OBJ_SPHERE) return hitSph(xyz(r0), r1.x, o, d, tmin, tmax, t);
  if (type == CPT_OBJ_LIGHT)
    return hitQuad(xyz(r0), xyz(r1), r2.x, r2.y, r2.z != 0.0f, o, d, tmin, tmax, t);
  return false;
}

// Hit geometry from a leaf-order record (triangles carry their edges):
// outward normal (Hittable.cuh:51,128,393).
__device__ __forceinline__ f3 outwardNormal(float4 r0, float4 r1, float4 r2, int type, f3 hp) {
  if (type == CPT_OBJ_TRIANGLE) return unit3(cross3(xyz(r1), xyz(r2)));
  if (type == CPT_OBJ_SPHERE) return (hp - xyz(r0)) * (1.0f / r1.x);
  return xyz(r1);
}

struct TraceCounters {
  int nodes = 0, prims = 0;
};

// Whether a hit at t (t <= tbest, as the hit tests return) from leaf li
// replaces the best hit so far (leaf `hit`, -1 for none): when nearer or, at
// an equal distance, from a lower leaf index (no hit counts as the highest),
// so that a closest hit depends on the ray alone, not on the order in which
// its leaves are tested (which the persistent walks' queued leaf tests vary
// with the other lanes of the wave).
__device__ __forceinline__ bool takesHit(float t, float tbest, int li, int hit) {
  return t < tbest || (uint32_t)li < (uint32_t)hit;
}

// kAny: return at the first primitive hit in [tmin, tbest] (occlusion).
// Closest hit over the LBVH.  Starts at the root's children (the root box
// is not tested, as in traverseIterative, BVH.cuh:293-301), tests both child
// boxes against [tmin, best t], tests leaf children at once, descends into
// the nearer internal child and pushes the farther.  Returns the leaf index
// of the closest hit (-1 for none); tbest holds its FP32 distance.
// stk points at this lane's LDS column; entries are stk[i * stride].
template <bool kCount, bool kAny = false>
__device__ __forceinline__ int traceClosest(const float4 *__restrict__ nodes,
                                            const float4 *__restrict__ leaves, int n, f3 o, f3 d,
                                            float tmin, float &tbest, int *stk, int stride,
                                            TraceCounters &cnt) {
  int hit = -1;
  if (n <= 0) return -1;
  if (n == 1) {
    float t;
    if (kCount) cnt.prims++;
    if (hitLeaf(leaves, 0, o, d, tmin, tbest, t)) {
      tbest = t;
      hit = 0;
    }
    return hit;
  }
  const f3 inv = slabInv(d);
  f3 ood = mkf3(o.x * inv.x, o.y * inv.y, o.z * inv.z);
  int ovf[kStackTotal - kStackLds];
  int sp = 0;
  int node = 0;
  while (true) {
    if (kCount) cnt.nodes++;
    const float4 *np = nodes + 4 * node;
    const float4 a = np[0];
    const float4 b = np[1];
    const float4 c = np[2];
    const float4 r = np[3];
    const int ca = __float_as_int(r.x), cb = __float_as_int(r.y);
    float ax0 = __builtin_fmaf(a.x, inv.x, -ood.x), ax1 = __builtin_fmaf(a.y, inv.x, -ood.x);
    float ay0 = __builtin_fmaf(a.z, inv.y, -ood.y), ay1 = __builtin_fmaf(a.w, inv.y, -ood.y);
    float az0 = __builtin_fmaf(c.x, inv.z, -ood.z), az1 = __builtin_fmaf(c.y, inv.z, -ood.z);
    float bx0 = __builtin_fmaf(b.x, inv.x, -ood.x), bx1 = __builtin_fmaf(b.y, inv.x, -ood.x);
    float by0 = __builtin_fmaf(b.z, inv.y, -ood.y), by1 = __builtin_fmaf(b.w, inv.y, -ood.y);
    float bz0 = __builtin_fmaf(c.z, inv.z, -ood.z), bz1 = __builtin_fmaf(c.w, inv.z, -ood.z);
    float nearA = fmaxf(fmaxf(fminf(ax0, ax1), fminf(ay0, ay1)), fmaxf(fminf(az0, az1), tmin));
    float farA = fminf(fminf(fmaxf(ax0, ax1), fmaxf(ay0, ay1)), fminf(fmaxf(az0, az1), tbest));
    float nearB = fmaxf(fmaxf(fminf(bx0, bx1), fminf(by0, by1)), fmaxf(fminf(bz0, bz1), tmin));
    float farB = fminf(fminf(fmaxf(bx0, bx1), fmaxf(by0, by1)), fminf(fmaxf(bz0, bz1), tbest));
    bool hA = nearA <= farA;
    bool hB = nearB <= farB;
    if (hA && ca < 0) {
      float t;
      if (kCount) cnt.prims++;
      if (hitLeaf(leaves, ~ca, o, d, tmin, tbest, t) && (kAny || takesHit(t, tbest, ~ca, hit))) {
        tbest = t;
        hit = ~ca;
        if (kAny) return hit;
      }
      hA = false;
    }
    if (hB && cb < 0) {
      float t;
      if (kCount) cnt.prims++;
      if (hitLeaf(leaves, ~cb, o, d, tmin, tbest, t) && (kAny || takesHit(t, tbest, ~cb, hit))) {
        tbest = t;
        hit = ~cb;
        if (kAny) return hit;
      }
      hB = false;
    }
    if (hA && hB) {
      int first = nearA <= nearB ? ca : cb;
      int second = nearA <= nearB ? cb : ca;
      if (sp < kStackLds) stk[sp * stride] = second;
      else if (sp < kStackTotal) ovf[sp - kStackLds] = second;
      sp = sp < kStackTotal ? sp + 1 : sp;
      node = first;
    } else if (hA) {
      node = ca;
    } else if (hB) {
      node = cb;
    } else {
      if (sp == 0) break;
      sp--;
      node = sp < kStackLds ? stk[sp * stride] : ovf[sp - kStackLds];
    }
  }
  return hit;
}

// ---- four-wide traversal ------------------------------------------------------
// Record of internal node i (128 B, 8 x float4): the boxes of i's children,
// each internal child replaced by its own two children; when one child is a
// leaf, the larger-area internal grandchild is opened into the fourth slot.
// Every internal entry lies two or three binary levels below i.  SoA
// {minx[4]} {maxx[4]} {miny[4]} {maxy[4]} {minz[4]} {maxz[4]} {ref[4]}
// {unused}; ref >= 0 is the index of the record that continues the walk
// (records are stored in level order from the root, record 0, and only
// those a walk can reach), ~k is leaf k, kEmpty4 an unused slot.
constexpr int kEmpty4 = (int)0x80000000;
constexpr int kStack4Total = 96;  // <= 3 pushes per >= 2 binary levels, depth <= 64

// Quantized four-wide record (64 B, 4 x float4), same children and refs:
// {origin x, y, z, exponent bytes} {ref[4]} {lo_x, hi_x, lo_y, hi_y}
// {lo_z, hi_z, 0, 0}, where lo_x etc. hold one byte per child (byte k =
// child k) and a plane decodes as fmaf(byte, 2^e_axis, origin_axis) with
// e_axis + 127 in byte axis of the exponent word.  The build rounds every
// plane outward on that grid and checks the decoded FP32 value, so each
// decoded box contains the child's FP32 box.
template <bool kQuant>
__device__ __forceinline__ void loadRecord4(const float4 *__restrict__ recs, int node, float bx0[4], float bx1[4],
                                            float by0[4], float by1[4], float bz0[4], float bz1[4], int ref[4]) {
  if (!kQuant) {
    const float4 *np = recs + 8 * (int64_t)node;
    const float4 mnx = np[0], mxx = np[1], mny = np[2], mxy = np[3], mnz = np[4], mxz = np[5];
    const float4 rf = np[6];
    bx0[0] = mnx.x; bx0[1] = mnx.y; bx0[2] = mnx.z; bx0[3] = mnx.w;
    bx1[0] = mxx.x; bx1[1] = mxx.y; bx1[2] = mxx.z; bx1[3] = mxx.w;
    by0[0] = mny.x; by0[1] = mny.y; by0[2] = mny.z; by0[3] = mny.w;
    by1[0] = mxy.x; by1[1] = mxy.y; by1[2] = mxy.z; by1[3] = mxy.w;
    bz0[0] = mnz.x; bz0[1] = mnz.y; bz0[2] = mnz.z; bz0[3] = mnz.w;
    bz1[0] = mxz.x; bz1[1] = mxz.y; bz1[2] = mxz.z; bz1[3] = mxz.w;
    ref[0] = __float_as_int(rf.x); ref[1] = __float_as_int(rf.y);
    ref[2] = __float_as_int(rf.z); ref[3] = __float_as_int(rf.w);
  } else {
    const float4 *np = recs + 4 * (int64_t)node;
    const float4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3];
    const uint32_t eb = __float_as_uint(w0.w);
    const float sx = __uint_as_float((eb & 0xffu) << 23);
    const float sy = __uint_as_float(((eb >> 8) & 0xffu) << 23);
    const float sz = __uint_as_float(((eb >> 16) & 0xffu) << 23);
    const uint32_t lx = __float_as_uint(w2.x), hx = __float_as_uint(w2.y), ly = __float_as_uint(w2.z),
                   hy = __float_as_uint(w2.w), lz = __float_as_uint(w3.x), hz = __float_as_uint(w3.y);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      bx0[k] = __builtin_fmaf((float)((lx >> (8 * k)) & 255u), sx, w0.x);
      bx1[k] = __builtin_fmaf((float)((hx >> (8 * k)) & 255u), sx, w0.x);
      by0[k] = __builtin_fmaf((float)((ly >> (8 * k)) & 255u), sy, w0.y);
      by1[k] = __builtin_fmaf((float)((hy >> (8 * k)) & 255u), sy, w0.y);
      bz0[k] = __builtin_fmaf((float)((lz >> (8 * k)) & 255u), sz, w0.z);
      bz1[k] = __builtin_fmaf((float)((hz >> (8 * k)) & 255u), sz, w0.z);
    }
    ref[0] = __float_as_int(w1.x); ref[1] = __float_as_int(w1.y);
    ref[2] = __float_as_int(w1.z); ref[3] = __float_as_int(w1.w);
  }
}

// Next record of a four-wide walk from the internal children hit (h, their
// entry distances tn).  They are ordered farthest first by entry distance
// with a 5-comparator network (misses carry -inf and sort last).  Closest
// hit: all of them are sorted, the m hit children are pr[0 .. m-1] and the
// walk continues at pr[m-1], the nearest.  Occlusion: the walk continues at
// the first nearest child in slot order (ties at tmin are common there, and
// slot order finds occluders sooner) and only the others are sorted.  The
// others are pushed farthest first, so pops come nearest first.  Returns
// false when no internal child was hit.
// Entry distances kept beside the LDS stack entries of closest-hit walks
// (kCullPop): a popped entry whose box the ray enters beyond the best hit so
// far is skipped without a visit.  Stored as fp16 rounded toward zero, i.e.
// down for the non-negative entry distances (t >= tmin >= 0), so a stored
// value never exceeds the true one and a skipped entry is one the visit
// would have found empty; an entry at exactly the best distance is kept
// (equal-distance hits from lower leaves still win, see takesHit).  Entries
// past the LDS part of the stack carry no distance and are always visited.
__device__ __forceinline__ uint16_t entryKey(float t) {
  return (uint16_t)(__builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(t, t)) & 0xffffu);
}
__device__ __forceinline__ float entryOf(uint16_t k) { return (float)__builtin_bit_cast(_Float16, k); }

// Pops the next entry of a walk's stack into node (false when it is empty);
// kCullPop: entries whose stored entry distance exceeds tbest are dropped.
template <bool kCullPop>
__device__ __forceinline__ bool popEntry(const int *stk, const uint16_t *stkT, int stride, const int *ovf, int &sp,
                                         int &node, float tbest) {
  while (sp > 0) {
    sp--;
    if (sp < kStackLds) {
      if (kCullPop && entryOf(stkT[sp * stride]) > tbest) continue;
      node = stk[sp * stride];
    } else {
      node = ovf[sp - kStackLds];
    }
    return true;
  }
  return false;
}

template <bool kAny, bool kCullPop = false>
__device__ __forceinline__ bool descend4(const bool h[4], const float tn[4], const int ref[4], int *stk, int stride,
                                         int *ovf, int &sp, int &node, uint16_t *stkT = nullptr) {
  const int m = (int)h[0] + (int)h[1] + (int)h[2] + (int)h[3];
  if (m == 0) return false;
  int best = 0;
  if (kAny) {
    float bt = INFINITY;
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (h[k] && tn[k] < bt) {
        bt = tn[k];
        best = k;
      }
  }
  float pt[4];
  int pr[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    pt[k] = (h[k] && !(kAny && k == best)) ? tn[k] : -INFINITY;
    pr[k] = ref[k];
  }
#define CPT_CSWAP(a, b)      \
  if (pt[a] < pt[b]) {       \
    const float tt_ = pt[a]; \
    pt[a] = pt[b];           \
    pt[b] = tt_;             \
    const int tr_ = pr[a];   \
    pr[a] = pr[b];           \
    pr[b] = tr_;             \
  }
  CPT_CSWAP(0, 1)
  CPT_CSWAP(2, 3)
  CPT_CSWAP(0, 2)
  CPT_CSWAP(1, 3)
  CPT_CSWAP(1, 2)
#undef CPT_CSWAP
  const int c = m - 1;  // entries pushed (at most three)
  if (sp + 3 <= kStackLds) {
    // room in LDS for three: store all three, keep c of them
    stk[sp * stride] = pr[0];
    stk[(sp + 1) * stride] = pr[1];
    stk[(sp + 2) * stride] = pr[2];
    if (kCullPop) {
      stkT[sp * stride] = entryKey(pt[0]);
      stkT[(sp + 1) * stride] = entryKey(pt[1]);
      stkT[(sp + 2) * stride] = entryKey(pt[2]);
    }
    sp += c;
  } else {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      if (k < c) {
        if (sp < kStackLds) {
          stk[sp * stride] = pr[k];
          if (kCullPop) stkT[sp * stride] = entryKey(pt[k]);
        } else if (sp < kStack4Total) {
          ovf[sp - kStackLds] = pr[k];
        }
        sp = sp < kStack4Total ? sp + 1 : sp;
      }
    }
  }
  if (kAny) node = best == 0 ? ref[0] : best == 1 ? ref[1] : best == 2 ? ref[2] : ref[3];
  else node = c == 0 ? pr[0] : c == 1 ? pr[1] : c == 2 ? pr[2] : pr[3];
  return true;
}

// Slab distances of the ray to the six planes of each child box of record
// `node` (t = (plane - o) / d as fma(plane, inv, -o*inv)).  Quantized
// records fold the plane decode into the product: with S = 2^e * inv and
// O = fma(origin, inv, -o*inv) per axis, t = fma(byte, S, O), one FMA per
// plane instead of two (the same FP32 slab arithmetic up to a few ulps of
// max(|O|, 8 |255 S|), i.e. of the record's distance or extent along the
// ray: far below the 1e-4 padding of primitive boxes at scene scales).
// fp16 value of half `hi` of p, as float
__device__ __forceinline__ float halfOf(uint32_t p, int hi) {
  const uint16_t b = hi ? (uint16_t)(p >> 16) : (uint16_t)(p & 0xffffu);
  return (float)__builtin_bit_cast(_Float16, b);
}

// t[k] = fma(1024 + byte k of w, s, o): v_perm_b32 puts bytes (0, 1) and
// (2, 3) of w under the fp16 high byte 0x64 (the value 1024 + byte, exact).
__device__ __forceinline__ void planes4(uint32_t w, float s, float o, float t[4]) {
  const uint32_t p01 = __builtin_amdgcn_perm(0x64646464u, w, 0x04010400u);
  const uint32_t p23 = __builtin_amdgcn_perm(0x64646464u, w, 0x04030402u);
  t[0] = __builtin_fmaf(halfOf(p01, 0), s, o);
  t[1] = __builtin_fmaf(halfOf(p01, 1), s, o);
  t[2] = __builtin_fmaf(halfOf(p23, 0), s, o);
  t[3] = __builtin_fmaf(halfOf(p23, 1), s, o);
}

template <bool kQuant>
__device__ __forceinline__ void recordSlabs4(const float4 *__restrict__ recs, int node, f3 inv, f3 ood, float x0[4],
                                             float x1[4], float y0[4], float y1[4], float z0[4], float z1[4],
                                             int ref[4]) {
  if (!kQuant) {
    float bx0[4], bx1[4], by0[4], by1[4], bz0[4], bz1[4];
    loadRecord4<false>(recs, node, bx0, bx1, by0, by1, bz0, bz1, ref);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      x0[k] = __builtin_fmaf(bx0[k], inv.x, -ood.x);
      x1[k] = __builtin_fmaf(bx1[k], inv.x, -ood.x);
      y0[k] = __builtin_fmaf(by0[k], inv.y, -ood.y);
      y1[k] = __builtin_fmaf(by1[k], inv.y, -ood.y);
      z0[k] = __builtin_fmaf(bz0[k], inv.z, -ood.z);
      z1[k] = __builtin_fmaf(bz1[k], inv.z, -ood.z);
    }
  } else {
    const float4 *np = recs + 4 * (int64_t)node;
    const float4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3];
    const uint32_t eb = __float_as_uint(w0.w);
    const float sx = __uint_as_float((eb & 0xffu) << 23) * inv.x;
    const float sy = __uint_as_float(((eb >> 8) & 0xffu) << 23) * inv.y;
    const float sz = __uint_as_float(((eb >> 16) & 0xffu) << 23) * inv.z;
    // planes are decoded as fp16 values 1024 + byte (one v_perm_b32 makes two
    // from a record word, and the FMA reads an fp16 half directly), so the
    // offsets carry -1024 S: t = fma(1024 + byte, S, O - 1024 S)
    const float ox = __builtin_fmaf(-1024.0f, sx, __builtin_fmaf(w0.x, inv.x, -ood.x)),
                oy = __builtin_fmaf(-1024.0f, sy, __builtin_fmaf(w0.y, inv.y, -ood.y)),
                oz = __builtin_fmaf(-1024.0f, sz, __builtin_fmaf(w0.z, inv.z, -ood.z));
    // the near plane of each axis is the low plane where the ray's direction
    // is positive on it and the high one where it is negative: x0/y0/z0 get
    // the near distances, x1/y1/z1 the far ones (a child's low byte is never
    // above its high byte, and t is monotone in the byte)
    const uint32_t w2x = __float_as_uint(w2.x), w2y = __float_as_uint(w2.y), w2z = __float_as_uint(w2.z),
                   w2w = __float_as_uint(w2.w), w3x = __float_as_uint(w3.x), w3y = __float_as_uint(w3.y);
    const bool nx = inv.x < 0.f, ny = inv.y < 0.f, nz = inv.z < 0.f;
    const uint32_t lx = nx ? w2y : w2x, hx = nx ? w2x : w2y, ly = ny ? w2w : w2z, hy = ny ? w2z : w2w,
                   lz = nz ? w3y : w3x, hz = nz ? w3x : w3y;
    planes4(lx, sx, ox, x0);
    planes4(hx, sx, ox, x1);
    planes4(ly, sy, oy, y0);
    planes4(hy, sy, oy, y1);
    planes4(lz, sz, oz, z0);
    planes4(hz, sz, oz, z1);
    ref[0] = __float_as_int(w1.x); ref[1] = __float_as_int(w1.y);
    ref[2] = __float_as_int(w1.z); ref[3] = __float_as_int(w1.w);
  }
}

template <bool kCount, bool kAny = false, bool kQuant = false, bool kCullPop = false>
__device__ __forceinline__ int traceClosest4(const float4 *__restrict__ nodes4,
                                             const float4 *__restrict__ leaves, int n, f3 o, f3 d,
                                             float tmin, float &tbest, int *stk, int stride,
                                             TraceCounters &cnt, uint16_t *stkT = nullptr) {
  static_assert(!(kAny && kCullPop), "occlusion walks keep tmax: nothing to cull");
  int hit = -1;
  if (n <= 0) return -1;
  if (n == 1) {
    float t;
    if (kCount) cnt.prims++;
    if (hitLeaf(leaves, 0, o, d, tmin, tbest, t)) {
      tbest = t;
      hit = 0;
    }
    return hit;
  }
  const f3 inv = slabInv(d);
  f3 ood = mkf3(o.x * inv.x, o.y * inv.y, o.z * inv.z);
  int ovf[kStack4Total - kStackLds];
  int sp = 0;
  int node = 0;
  while (true) {
    if (kCount) cnt.nodes++;
    float sx0[4], sx1[4], sy0[4], sy1[4], sz0[4], sz1[4];
    int ref[4];
    recordSlabs4<kQuant>(nodes4, node, inv, ood, sx0, sx1, sy0, sy1, sz0, sz1, ref);
    float tn[4];
    bool h[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const float x0 = sx0[k], x1 = sx1[k], y0 = sy0[k], y1 = sy1[k], z0 = sz0[k], z1 = sz1[k];
      // kQuant: the slabs come ordered (near, far)
      const float nx = kQuant ? x0 : fminf(x0, x1), fx = kQuant ? x1 : fmaxf(x0, x1);
      const float ny = kQuant ? y0 : fminf(y0, y1), fy = kQuant ? y1 : fmaxf(y0, y1);
      const float nz = kQuant ? z0 : fminf(z0, z1), fz = kQuant ? z1 : fmaxf(z0, z1);
      tn[k] = fmaxf(fmaxf(nx, ny), fmaxf(nz, tmin));
      const float tf = fminf(fminf(fx, fy), fminf(fz, tbest));
      h[k] = tn[k] <= tf && ref[k] != kEmpty4;
    }
    // leaf children first, in slot order: one test per loop trip, so a wave
    // runs as many trips as its lane with the most leaf hits
    uint32_t lm = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (h[k] && ref[k] < 0) lm |= 1u << k;
      h[k] = h[k] && ref[k] >= 0;
    }
    while (lm) {
      const int k = __builtin_ctz(lm);
      lm &= lm - 1;
      const int li = ~(k == 0 ? ref[0] : k == 1 ? ref[1] : k == 2 ? ref[2] : ref[3]);
      float t;
      if (kCount) cnt.prims++;
      if (hitLeaf(leaves, li, o, d, tmin, tbest, t) && (kAny || takesHit(t, tbest, li, hit))) {
        tbest = t;
        hit = li;
        if (kAny) return hit;
      }
    }
    // internal children: continue with the nearest, push the others
    if (!descend4<kAny, kCullPop>(h, tn, ref, stk, stride, ovf, sp, node, stkT) &&
        !popEntry<kCullPop>(stk, stkT, stride, ovf, sp, node, tbest))
      break;
  }
  return hit;
}

// ---- eight-wide quantized traversal ------------------------------------------
// Record of internal node i (128 B, 6 of 8 float4 used): node i's subtree
// expanded three binary levels deep (internal entries are great-grandchildren,
// leaf entries lie at depth 1-3), with 8-bit planes as in the quantized
// four-wide record: {origin x, y, z, exponent bytes} {ref[0..3]} {ref[4..7]}
// {lo_x[0..3], lo_x[4..7], hi_x[0..3], hi_x[4..7]} {… y} {… z}.  A walk
// descends three binary levels per record (<= 22 records for depth <= 64),
// pushing <= 7 entries each.
constexpr int kStack8Total = 160;

template <bool kCount, bool kAny = false>
__device__ __forceinline__ int traceClosest8(const float4 *__restrict__ nodes8,
                                             const float4 *__restrict__ leaves, int n, f3 o, f3 d,
                                             float tmin, float &tbest, int *stk, int stride,
                                             TraceCounters &cnt) {
  int hit = -1;
  if (n <= 0) return -1;
  if (n == 1) {
    float t;
    if (kCount) cnt.prims++;
    if (hitLeaf(leaves, 0, o, d, tmin, tbest, t)) {
      tbest = t;
      hit = 0;
    }
    return hit;
  }
  const f3 inv = slabInv(d);
  f3 ood = mkf3(o.x * inv.x, o.y * inv.y, o.z * inv.z);
  int ovf[kStack8Total - kStackLds];
  int sp = 0;
  int node = 0;
  while (true) {
    if (kCount) cnt.nodes++;
    const float4 *np = nodes8 + 8 * (int64_t)node;
    const float4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3], w4 = np[4], w5 = np[5];
    const uint32_t eb = __float_as_uint(w0.w);
    const float sx = __uint_as_float((eb & 0xffu) << 23);
    const float sy = __uint_as_float(((eb >> 8) & 0xffu) << 23);
    const float sz = __uint_as_float(((eb >> 16) & 0xffu) << 23);
    const int ref[8] = {__float_as_int(w1.x), __float_as_int(w1.y), __float_as_int(w1.z), __float_as_int(w1.w),
                        __float_as_int(w2.x), __float_as_int(w2.y), __float_as_int(w2.z), __float_as_int(w2.w)};
    const uint32_t q[3][4] = {{__float_as_uint(w3.x), __float_as_uint(w3.y), __float_as_uint(w3.z), __float_as_uint(w3.w)},
                              {__float_as_uint(w4.x), __float_as_uint(w4.y), __float_as_uint(w4.z), __float_as_uint(w4.w)},
                              {__float_as_uint(w5.x), __float_as_uint(w5.y), __float_as_uint(w5.z), __float_as_uint(w5.w)}};
    const float org[3] = {w0.x, w0.y, w0.z}, sc[3] = {sx, sy, sz};
    const float iv[3] = {inv.x, inv.y, inv.z}, od[3] = {ood.x, ood.y, ood.z};
    float tn[8];
    bool h[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      float t0 = tmin, t1 = tbest;
#pragma unroll
      for (int ax = 0; ax < 3; ax++) {
        const float lo = __builtin_fmaf((float)((q[ax][k / 4] >> (8 * (k % 4))) & 255u), sc[ax], org[ax]);
        const float hi = __builtin_fmaf((float)((q[ax][2 + k / 4] >> (8 * (k % 4))) & 255u), sc[ax], org[ax]);
        const float a = __builtin_fmaf(lo, iv[ax], -od[ax]), b = __builtin_fmaf(hi, iv[ax], -od[ax]);
        t0 = fmaxf(t0, fminf(a, b));
        t1 = fminf(t1, fmaxf(a, b));
      }
      tn[k] = t0;
      h[k] = t0 <= t1 && ref[k] != kEmpty4;
    }
    // leaf children in slot order, one test per loop trip
    uint32_t lm = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (h[k] && ref[k] < 0) lm |= 1u << k;
      h[k] = h[k] && ref[k] >= 0;
    }
    while (lm) {
      const int k = __builtin_ctz(lm);
      lm &= lm - 1;
      int r = ref[0];
#pragma unroll
      for (int j = 1; j < 8; j++) r = k == j ? ref[j] : r;
      float t;
      if (kCount) cnt.prims++;
      if (hitLeaf(leaves, ~r, o, d, tmin, tbest, t) && (kAny || takesHit(t, tbest, ~r, hit))) {
        tbest = t;
        hit = ~r;
        if (kAny) return hit;
      }
    }
    int best = -1;
    float bt = INFINITY;
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (h[k] && tn[k] < bt) {
        bt = tn[k];
        best = k;
      }
    if (best < 0) {
      if (sp == 0) break;
      sp--;
      node = sp < kStackLds ? stk[sp * stride] : ovf[sp - kStackLds];
      continue;
    }
    // the others farthest first (19-comparator network, descending)
    float pt[8];
    int pr[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      pt[k] = (h[k] && k != best) ? tn[k] : -INFINITY;
      pr[k] = ref[k];
    }
#define CPT_CSWAP(a, b)        \
  if (pt[a] < pt[b]) {         \
    const float tt_ = pt[a];   \
    pt[a] = pt[b];             \
    pt[b] = tt_;               \
    const int tr_ = pr[a];     \
    pr[a] = pr[b];             \
    pr[b] = tr_;               \
  }
    CPT_CSWAP(0, 2) CPT_CSWAP(1, 3) CPT_CSWAP(4, 6) CPT_CSWAP(5, 7)
    CPT_CSWAP(0, 4) CPT_CSWAP(1, 5) CPT_CSWAP(2, 6) CPT_CSWAP(3, 7)
    CPT_CSWAP(0, 1) CPT_CSWAP(2, 3) CPT_CSWAP(4, 5) CPT_CSWAP(6, 7)
    CPT_CSWAP(2, 4) CPT_CSWAP(3, 5)
    CPT_CSWAP(1, 4) CPT_CSWAP(3, 6)
    CPT_CSWAP(1, 2) CPT_CSWAP(3, 4) CPT_CSWAP(5, 6)
#undef CPT_CSWAP
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (pt[k] != -INFINITY) {
        if (sp < kStackLds) stk[sp * stride] = pr[k];
        else if (sp < kStack8Total) ovf[sp - kStackLds] = pr[k];
        sp = sp < kStack8Total ? sp + 1 : sp;
      }
    }
    node = ref[best];
  }
  return hit;
}

// ---- resumable four-wide walk (persistent kernels) ----------------------------
// The same walk as traceClosest4, one record visit per call, so that a lane
// whose ray is finished can take a new ray while the rest of its wave is
// still walking.  n >= 2 (a tree with at least one record).
struct Walk4 {
  f3 o, d, inv, ood;
  float tmin, tbest;
  int node, sp, hit;
  int npend;  // queued leaf tests (walk4Step<..., kPostpone>); node < 0: none to visit
  // hit payload (walk4Step<..., kPayload>): the best hit's outward normal,
  // and material << 3 | type, set from its record when the hit is found
  f3 nrm;
  int tm;
};

// Starts the walk of the ray already in w.o, w.d, w.tmin, w.tbest (= tmax).
__device__ __forceinline__ void walk4Start(Walk4 &w) {
  w.inv = slabInv(w.d);
  w.ood = mkf3(w.o.x * w.inv.x, w.o.y * w.inv.y, w.o.z * w.inv.z);
  w.node = 0;
  w.sp = 0;
  w.hit = -1;
  w.npend = 0;
  w.nrm = mkf3(0.f, 0.f, 0.f);
  w.tm = 0;
}

// The ray {o, tmin} {d, tmax} into w (its walk starts with walk4Start).
__device__ __forceinline__ void walk4Ray(Walk4 &w, float4 r0, float4 r1) {
  w.o = xyz(r0);
  w.d = xyz(r1);
  w.tmin = r0.w;
  w.tbest = r1.w;
}

__device__ __forceinline__ void walk4Init(Walk4 &w, f3 o, f3 d, float tmin, float tmax) {
  w.o = o;
  w.d = d;
  w.tmin = tmin;
  w.tbest = tmax;
  walk4Start(w);
}

// Leaf test of leaf li for walk4Step; returns true when a kAny walk is over.
template <bool kCount, bool kAny, bool kPayload>
__device__ __forceinline__ bool walk4Leaf(Walk4 &w, const float4 *__restrict__ leaves, int li, TraceCounters &cnt) {
  float t;
  if (kCount) cnt.prims++;
  if (kPayload) {
    const float4 q0 = leaves[3 * li + 0], q1 = leaves[3 * li + 1], q2 = leaves[3 * li + 2];
    if (hitRecord(q0, q1, q2, w.o, w.d, w.tmin, w.tbest, t) && (kAny || takesHit(t, w.tbest, li, w.hit))) {
      w.tbest = t;
      w.hit = li;
      const int type = __float_as_int(q1.w);
      w.tm = (__float_as_int(q2.w) << 3) | type;
      w.nrm = outwardNormal(q0, q1, q2, type, w.o + w.d * t);
    }
  } else if (hitLeaf(leaves, li, w.o, w.d, w.tmin, w.tbest, t) && (kAny || takesHit(t, w.tbest, li, w.hit))) {
    w.tbest = t;
    w.hit = li;
    if (kAny) return true;
  }
  return false;
}

constexpr int kPend = 6;  // queued leaf tests per lane (kPostpone)

// One trip of a four-wide walk.  Returns false when the walk is over (w.hit,
// w.tbest hold the result; kAny: the first hit ends it).  kPayload: a new
// best hit also sets w.nrm, w.tm from its record.
//
// Without kPostpone a trip visits record w.node and tests its hit leaves in
// a loop, one per loop round: a wave runs as many rounds as its lane with
// the most leaf hits, while most lanes have none in a given record.  With
// kPostpone (pnd: kPend LDS words per lane, stride apart) the leaves a visit
// hits are queued, and the walking lanes of the wave vote: a leaf trip, in
// which every lane with queued leaves tests one, once half of them have
// some or none can visit (no record left, or no room for four more leaves);
// otherwise a visit trip.  *visit says which (wave-uniform).
template <bool kCount, bool kAny, bool kQuant = false, bool kPayload = false, bool kPostpone = false,
          bool kCullPop = false>
__device__ __forceinline__ bool walk4Step(Walk4 &w, const float4 *__restrict__ nodes4,
                                          const float4 *__restrict__ leaves, int *stk, int stride, int *ovf,
                                          TraceCounters &cnt, int *pnd = nullptr, bool *visit = nullptr,
                                          uint16_t *stkT = nullptr) {
  static_assert(!(kAny && kCullPop), "occlusion walks keep tmax: nothing to cull");
  if (kPostpone) {
    const bool canVisit = w.node >= 0 && w.npend <= kPend - 4;
    const int nWalk = __popcll(__ballot(true));
    const uint64_t vis = __ballot(canVisit);
    const int nQueued = __popcll(__ballot(w.npend > 0));
    const bool leafTrip = vis == 0 || 2 * nQueued >= nWalk;
    *visit = !leafTrip;
    if (leafTrip) {
      if (w.npend > 0) {
        w.npend--;
        if (walk4Leaf<kCount, kAny, kPayload>(w, leaves, pnd[w.npend * stride], cnt)) return false;
      }
      return w.node >= 0 || w.npend > 0;
    }
    if (!canVisit) return true;
  }
  if (kCount) cnt.nodes++;
  float sx0[4], sx1[4], sy0[4], sy1[4], sz0[4], sz1[4];
  int ref[4];
  recordSlabs4<kQuant>(nodes4, w.node, w.inv, w.ood, sx0, sx1, sy0, sy1, sz0, sz1, ref);
  float tn[4];
  bool h[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const float x0 = sx0[k], x1 = sx1[k], y0 = sy0[k], y1 = sy1[k], z0 = sz0[k], z1 = sz1[k];
    const float nx = kQuant ? x0 : fminf(x0, x1), fx = kQuant ? x1 : fmaxf(x0, x1);
    const float ny = kQuant ? y0 : fminf(y0, y1), fy = kQuant ? y1 : fmaxf(y0, y1);
    const float nz = kQuant ? z0 : fminf(z0, z1), fz = kQuant ? z1 : fmaxf(z0, z1);
    tn[k] = fmaxf(fmaxf(nx, ny), fmaxf(nz, w.tmin));
    const float tf = fminf(fminf(fx, fy), fminf(fz, w.tbest));
    h[k] = tn[k] <= tf && ref[k] != kEmpty4;
  }
  uint32_t lm = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (h[k] && ref[k] < 0) lm |= 1u << k;
    h[k] = h[k] && ref[k] >= 0;
  }
  if (kPostpone) {
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (lm & (1u << k)) {
        pnd[w.npend * stride] = ~ref[k];
        w.npend++;
      }
  } else {
    while (lm) {
      const int k = __builtin_ctz(lm);
      lm &= lm - 1;
      const int li = ~(k == 0 ? ref[0] : k == 1 ? ref[1] : k == 2 ? ref[2] : ref[3]);
      if (walk4Leaf<kCount, kAny, kPayload>(w, leaves, li, cnt)) return false;
    }
  }
  if (!descend4<kAny, kCullPop>(h, tn, ref, stk, stride, ovf, w.sp, w.node, stkT) &&
      !popEntry<kCullPop>(stk, stkT, stride, ovf, w.sp, w.node, w.tbest)) {
    if (!kPostpone) return false;
    w.node = -1;
    return w.npend > 0;
  }
  return true;
}

// Persistent walks: a wave takes kChunk items per grab of a global counter;
// a lane moves through these states, one loop trip each, as it takes an item.
constexpr int kChunk = 256;
constexpr int kClaimMin = 16;
enum LaneState { kIdle = 0, kClaimed = 1, kLoading = 2, kActive = 3 };

// Wave-level item claiming for persistent walks.  The wave holds a chunk
// [cur, end) of items and, in lane 0's `ticket`, the result of the grab for
// its next chunk, issued when the current one was taken so that no lane
// waits for the counter.  Every idle lane gets the next item of the chunk
// (returns true, item set) while the chunk lasts.  `more` turns false once a
// grab has passed n.  Call with the whole wave active (the caller's loop is
// wave-uniform).
struct WaveChunk {
  uint32_t cur = 0, end = 0;  // wave-uniform
  uint32_t ticket = 0;
  bool more = true;
};

__device__ __forceinline__ void waveChunkInit(WaveChunk &c, uint32_t *counter) {
  if (__lane_id() == 0) c.ticket = atomicAdd(counter, (uint32_t)kChunk);
}

// n < 2^30 (the counter overshoots n by at most two grabs per wave, and an
// abort moves it to kAbortGrab).
__device__ __forceinline__ bool waveClaim(WaveChunk &c, uint32_t *counter, uint32_t n, bool idle, uint32_t &item) {
  const uint64_t im = __ballot(idle);
  if (!im || !c.more) return false;
  // a trip that claims waits for the claimed lanes' loads before its walk
  // steps can issue theirs: claim only once kClaimMin lanes are idle (or
  // the whole wave is), so that such trips are few
  if (__popcll(im) < kClaimMin && __ballot(!idle) != 0) return false;
  const uint32_t lane = __lane_id();
  if (c.cur >= c.end) {
    const uint32_t base = __builtin_amdgcn_readfirstlane(__shfl(c.ticket, 0, 64));
    if (base >= n) {
      c.more = false;
      return false;
    }
    c.cur = base;
    c.end = n - base > (uint32_t)kChunk ? base + (uint32_t)kChunk : n;
    if (lane == 0) c.ticket = atomicAdd(counter, (uint32_t)kChunk);
  }
  const uint32_t avail = c.end - c.cur;
  const uint32_t nIdle = (uint32_t)__popcll(im);
  const uint32_t rank = (uint32_t)__popcll(im & ((1ull << lane) - 1ull));
  const bool got = idle && rank < avail;
  item = c.cur + rank;
  c.cur += nIdle < avail ? nIdle : avail;
  return got;
}

// Host abort of a persistent walk.  The abort word lives in pinned host
// memory, and a load of it crosses the host link: waiting for one per chunk
// grab in every wave stalls the walks (loads return in order).  So only
// wave 0 of workgroup 0 reads it, every kAbortPollTrips of its loop trips;
// once it is set, that wave moves the shared grab counter to kAbortGrab,
// past every n, and each wave's next grab ends its claiming: the walks in
// flight finish and every wave leaves normally.
constexpr uint32_t kAbortGrab = 0xC0000000u;
constexpr uint32_t kAbortPollTrips = 64;

__device__ __forceinline__ void abortRelay(const uint32_t *abortFlag, uint32_t *counter, uint32_t &trips) {
  if (!abortFlag || blockIdx.x != 0 || threadIdx.x >= 64) return;  // wave-uniform
  if ((++trips & (kAbortPollTrips - 1)) != 1u) return;
  if (__lane_id() == 0 && __hip_atomic_load(abortFlag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u)
    atomicMax(counter, kAbortGrab);
}

// Wave-aggregated fetch of work items for the lanes with `want`: returns the
// item index of each such lane (>= total when the work is exhausted) and
// sets `exhausted` (wave-uniform) once the counter has passed total.
__device__ __forceinline__ uint64_t waveFetch(uint32_t *counter, bool want, uint64_t total, bool &exhausted) {
  const uint64_t m = __ballot(want);
  const int lane = (int)__lane_id();
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = __shfl(base, leader, 64);
  if ((uint64_t)base + (uint64_t)__popcll(m) >= total) exhausted = true;
  return (uint64_t)base + (uint64_t)__popcll(m & ((1ull << lane) - 1ull));
}

// ---- FP64 refinement of the winning hit ------------------------------------
// Recomputes the distance to primitive `id` with the reference's FP64
// formulas (Hittable.cuh:27-54, 85-140, 367-398) on the FP64 scene copy.
struct Scene64 {
  const double *tri;    // n_tri x 9, insertion order
  const double *sph;    // n_sph x 4 (c, r)
  const double *lgt;    // n_lgt x 6 (c, n unit)
  int32_t nLights, nSpheres;
};

__device__ __forceinline__ double refineT(const Scene64 &s, int id, double ox, double oy, double oz,
                                          double dx, double dy, double dz, double tmin, double tmax) {
#pragma clang fp contract(off)
  if (id < s.nLights) {
    const double *L = s.lgt + 6 * id;
    double denom = dx * L[3] + dy * L[4] + dz * L[5];
    double num = (L[0] - ox) * L[3] + (L[1] - oy) * L[4] + (L[2] - oz) * L[5];
    return num / denom;
  }
  if (id < s.nLights + s.nSpheres) {
    const double *S = s.sph + 4 * (id - s.nLights);
    double ocx = ox - S[0], ocy = oy - S[1], ocz = oz - S[2];
    double a = dx * dx + dy * dy + dz * dz;
    double b = dx * ocx + dy * ocy + dz * ocz;
    double c = (ocx * ocx + ocy * ocy + ocz * ocz) - S[3] * S[3];
    double disc = b * b - a * c;
    double sq = sqrt(disc > 0.0 ? disc : 0.0);
    double t = (-b - sq) / a;
    if (t < tmin || t > tmax) t = (-b + sq) / a;
    return t;
  }
  const double *T = s.tri + 9 * (int64_t)(id - s.nLights - s.nSpheres);
  double e1x = T[3] - T[0], e1y = T[4] - T[1], e1z = T[5] - T[2];
  double e2x = T[6] - T[0], e2y = T[7] - T[1], e2z = T[8] - T[2];
  double hx = dy * e2z - dz * e2y, hy = dz * e2x - dx * e2z, hz = dx * e2y - dy * e2x;
  double a = e1x * hx + e1y * hy + e1z * hz;
  double f = 1.0 / a;
  double sx = ox - T[0], sy = oy - T[1], sz = oz - T[2];
  double qx = sy * e1z - sz * e1y, qy = sz * e1x - sx * e1z, qz = sx * e1y - sy * e1x;
  return f * (e2x * qx + e2y * qy + e2z * qz);
}

}  // namespace cpt

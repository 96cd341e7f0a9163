// oracle_render.cpp - FP64 CPU restatement of the reference integrator and
// materials.  TEST INFRASTRUCTURE ONLY (see oracle_common.hpp).
//
// Follows:
//   render (STATIC mode)   CudaPathTracing/Render.cu:16-261 (per sample: jitter,
//                          bounce loop, homogeneous volume 64-135, miss 138-142,
//                          emitter 144-149, NEE 161-187, roulette 189-204,
//                          BSDF 206-218, mean 220-222)
//   Camera ray             Camera.cuh:37-43
//   Material::fr           Material.cuh:24-93
//   Material::sampleLight  Material.cuh:95-193, brdf 201-235
//   GGX D / lambda / G     Material.cuh:238-285
//   Fresnel                Material.cuh:287-318
// Deviations, all in cases where the reference returns NaN: sin(theta) = 0 in
// DistributionGGX/lambda uses cos(phi) = 1 (the phi term is multiplied by
// tan(theta) = 0 there), and 1 - z^2 < 0 from rounding is clamped to 0.
// Random numbers: Philox4x32-10 (Salmon et al., SC'11); draw k of
// (pixel p, sample s) is word k&3 of philox({k>>2, p, s, 0x43505431}, seed),
// mapped to ((x>>8)+1)*2^-24.  Draw order per sample follows Render.cu.
// The new material types (CPT_M_LAMBERT, CPT_M_GGX) and the CONFIG profile
// (max depth, roulette from a given depth) have no reference counterpart.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

#include "../include/cpt.h"
#include "oracle_common.hpp"
#include "oracle_scene.hpp"

namespace orc {

struct PhiloxOut {
  uint32_t v[4];
};

PhiloxOut philox4x32_10(const uint32_t ctr[4], uint32_t k0, uint32_t k1) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return PhiloxOut{{c0, c1, c2, c3}};
}

// tag: 'CPT1' for a path; 0x80000000 | (path draw count when the shadow ray
// was made) for the transmittance estimate of a shadow ray.
struct Stream {
  uint32_t k0, k1, pix, smp, cnt = 0, tag = 0x43505431u;
  double next() {
    uint32_t c[4] = {cnt >> 2, pix, smp, tag};
    PhiloxOut o = philox4x32_10(c, k0, k1);
    uint32_t x = o.v[cnt & 3];
    cnt++;
    return (double)((x >> 8) + 1u) * (1.0 / 16777216.0);
  }
};

// Certificate of one sample's path (the per-path parity bar).  The GPU
// takes every decision of a path in FP32 and the oracle in FP64; a path is
// certified when each decision it takes is robust to that rounding by a
// margin, so that both sides must follow the same branches.  why = 0 means
// certified, otherwise the class of the first decision that was not robust.
enum CertWhy : uint8_t {
  kCertOk = 0,
  kCertHit = 1,         // closest hit not robust (bruteClosest)
  kCertNeeSide = 2,     // NEE direction against the surface or light normal
  kCertChoice = 3,      // Fresnel branch choice r3 vs R
  kCertSampleSide = 4,  // sampled direction against the normal
  kCertFrame = 5,       // frame construction threshold
  kCertMedium = 6,      // free flight against the segment end / box entry
  kCertVoxel = 7,       // position on a voxel boundary
  kCertVolume = 8,      // volume-object free flight against its exit distance
  kCertFace = 9,        // front / back face
  kCertBtdf = 10,       // BTDF half-vector side flips
};
constexpr double kCertEbHit = 1e-4;    // relative margin of hit decisions
constexpr double kCertEtHit = 1e-4;    // relative separation of the winner
constexpr double kCertOrigin = 1e-6;   // origin uncertainty, relative to its largest coordinate
constexpr double kCertSeg = 2e-6;      // ... plus this much of the path length that led to it
constexpr double kCertDir = 2e-6;      // relative direction uncertainty
constexpr double kCertDot = 1e-5;      // sign of a dot product of unit vectors
constexpr double kCertFrameEps = 1e-6; // frame thresholds (0.99999, 0.9)
constexpr double kCertProb = 1e-5;     // r3 against a Fresnel probability
constexpr double kCertDist = 1e-4;     // relative margin of distance comparisons
constexpr double kCertVoxelEps = 1e-3; // voxel boundary margin, in voxels

struct Cert {
  uint8_t why = kCertOk;
};
// The certificate of the path being rendered on this thread (null: none),
// and the length of the path so far: an FP32 hit point inherits its
// origin's error plus a few ulp of the segment's distance, so the error of a
// ray origin grows with the path length that led to it.
static thread_local Cert *tlCert = nullptr;
static thread_local double tlPathLen = 0.0;
static thread_local double tlPosErr = 0.0;  // uncertainty of the medium positions being looked up
static inline void certNeed(bool robust, CertWhy why) {
  if (tlCert && !robust && tlCert->why == kCertOk) tlCert->why = why;
}
static inline void certMargin(double v, double eps, CertWhy why) { certNeed(std::fabs(v) > eps, why); }
// a sampled distance a against a distance b whose own absolute uncertainty
// (from the FP32 ray origin and direction) is bErr
static inline void certDist(double a, double b, CertWhy why, double bErr = 0.0) {
  certNeed(std::fabs(a - b) > kCertDist * std::max(std::max(std::fabs(a), std::fabs(b)), 1e-3) + bErr, why);
}

struct OMat {
  d3 color;
  double ax, ay, eta;
  int type;
};

// worldToLocal, Material.cuh:238-248
static d3 worldToLocal(const d3 &N, const d3 &W) {
  d3 z = N;
  d3 up = mk(0.0, 0.0, 1.0);
  certMargin(std::fabs(Dot(z, up)) - 0.99999, kCertFrameEps, kCertFrame);
  if (std::fabs(Dot(z, up)) > 0.99999) up = mk(1.0, 0.0, 0.0);
  d3 x = Unit(Cross(up, z));
  d3 y = Cross(z, x);
  return mk(Dot(W, x), Dot(W, y), Dot(W, z));
}

// DistributionGGX, Material.cuh:250-265
static double DistributionGGX(const OMat &m, const d3 &N, const d3 &H) {
  d3 w = worldToLocal(N, H);
  double sinT = std::sqrt(std::fmax(0.0, 1.0 - w.z * w.z));
  double cosP = sinT > 0.0 ? w.x / sinT : 1.0;
  double sinP = sinT > 0.0 ? w.y / sinT : 0.0;
  double cos2 = w.z * w.z;
  double sin2 = 1.0 - cos2;
  double tan2 = std::fmax(0.0, sin2) / cos2;
  double cos4 = cos2 * cos2;
  double cDa = cosP / m.ax;
  double sDa = sinP / m.ay;
  double e = tan2 * (cDa * cDa + sDa * sDa);
  return 1.0 / (PI * m.ax * m.ay * cos4 * (1 + e) * (1 + e));
}

// lambda, Material.cuh:266-277
static double lambda(const OMat &m, const d3 &N, const d3 &R) {
  d3 w = worldToLocal(N, R);
  double sinT = std::sqrt(std::fmax(0.0, 1 - w.z * w.z));
  double cosP = sinT > 0.0 ? w.x / sinT : 1.0;
  double sinP = sinT > 0.0 ? w.y / sinT : 0.0;
  double alpha2 = m.ax * m.ax * cosP * cosP + m.ay * m.ay * sinP * sinP;
  double alpha = std::sqrt(alpha2);
  double tanT = std::tan(std::acos(std::fmin(mMax(Dot(N, R), 0.0), 1.0)));
  return (std::sqrt(1.0 + alpha * alpha * tanT * tanT) - 1.0) / 2.0;
}
static double GeometrySmith(const OMat &m, const d3 &N, const d3 &V, const d3 &L) {
  return 1.0 / (1.0 + lambda(m, N, V) + lambda(m, N, L));
}
static double GeometrySmith1(const OMat &m, const d3 &N, const d3 &W) { return 1.0 / (1.0 + lambda(m, N, W)); }

// FresnelConductor, Material.cuh:287-295 (complex arithmetic of rtmath.cuh)
struct cx {
  double x, y;
};
static cx cdiv(cx a, cx b) {
  double den = b.x * b.x + b.y * b.y;
  return cx{(a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den};
}
static double FresnelConductor(const d3 &N, const d3 &V, const d3 &L, double n, double k) {
  cx eta{n, k};
  double ci = mMax(Dot(N, V), 0.0);
  double ct = mMax(Dot(N, L), 0.0);
  cx ec = cx{ci * eta.x, ci * eta.y};
  cx rpa = cdiv(cx{ec.x - ct, ec.y}, cx{ct + ec.x, ec.y});
  cx et = cx{ct * eta.x, ct * eta.y};
  cx rpe = cdiv(cx{ci - et.x, -et.y}, cx{ci + et.x, et.y});
  return ((rpa.x * rpa.x + rpa.y * rpa.y) + (rpe.x * rpe.x + rpe.y * rpe.y)) / 2.0;
}

// FresnelDielectric, Material.cuh:305-318
static double FresnelDielectric(const d3 &N, const d3 &V, double eta) {
  double ci = std::fabs(Dot(N, V));
  double si = std::sqrt(mMax(0.0, 1.0 - ci * ci));
  double st = si / eta;
  if (st >= 1.0) return 1.0;
  double ct = std::sqrt(mMax(0.0, 1.0 - st * st));
  double rpa = (eta * ci - ct) / (eta * ci + ct);
  double rpe = (ci - eta * ct) / (ci + eta * ct);
  return (rpa * rpa + rpe * rpe) * 0.5;
}

// Material::fr, Material.cuh:24-93 (+ LAMBERT, GGX)
static d3 fr(const OMat &m, const d3 &hitColor, const d3 &rayDir, const d3 &normal, const d3 &direction,
             const d3 &wm, bool frontFace) {
  d3 V = Unit(-rayDir);
  d3 L = Unit(direction);
  certMargin(Dot(L, normal), kCertDot, kCertSampleSide);
  if (m.type == CPT_M_LAMBERT) return Dot(L, normal) > 0.0 ? hitColor / PI : mk(0, 0, 0);
  d3 H = Unit(V + L);
  certMargin(Dot(V, normal), kCertDot, kCertFace);
  bool reflect = Dot(V, normal) * Dot(L, normal) > 0.0;
  if (m.type == CPT_M_SPECULAR_DIELECTRIC) {
    double rr = frontFace ? m.eta : (1.0 / m.eta);
    double ct = std::fabs(Dot(normal, L));
    double F = FresnelDielectric(normal, V, rr);
    if (reflect) return F * hitColor / ct;
    return (1.0 - F) * hitColor / ct;
  }
  if (m.type == CPT_M_GGX) {
    double nl = Dot(normal, L), nv = Dot(normal, V);
    if (!(nl > 0.0) || !(nv > 0.0)) return mk(0, 0, 0);
    double D = DistributionGGX(m, normal, H);
    double G = GeometrySmith(m, normal, V, L);
    double s = 1.0 - mMax(Dot(V, H), 0.0);
    double s5 = s * s * s * s * s;
    d3 F = mk(m.color.x + (1.0 - m.color.x) * s5, m.color.y + (1.0 - m.color.y) * s5,
              m.color.z + (1.0 - m.color.z) * s5);
    return F * (D * G / (4.0 * nv * nl));
  }
  if (!reflect) {
    double rr = frontFace ? m.eta : (1.0 / m.eta);
    double D = DistributionGGX(m, normal, wm);
    certMargin(Dot(wm, V), kCertDot, kCertBtdf);
    certMargin(Dot(wm, L), kCertDot, kCertBtdf);
    if (Dot(wm, V) < 0.0) V = -V;
    if (Dot(wm, L) < 0.0) L = -L;
    double G = GeometrySmith(m, normal, V, L);
    double F = FresnelDielectric(wm, V, rr);
    double ci = std::fabs(Dot(normal, V));
    double ct = std::fabs(Dot(normal, L));
    double mbase = std::fabs(Dot(L, wm)) + std::fabs(Dot(V, wm)) / rr;
    double denom = mbase * mbase;
    double spec = (1.0 - F) * D * G * std::fabs(Dot(wm, L) * Dot(wm, V) / (ci * ct * denom));
    return spec * hitColor;
  }
  double n = 2.63, k = 9.08;
  if (m.type == CPT_M_ROUGH_DIELECTRIC) {
    n = 1.58;
    k = 0.001;
  }
  double D = DistributionGGX(m, normal, H);
  double G = GeometrySmith(m, normal, V, L);
  double F = FresnelConductor(normal, V, L, n, k);
  double NdotL = mMax(Dot(normal, L), 0.0);
  double NdotV = mMax(Dot(normal, V), 0.0);
  double denominator = 4.0 * NdotV * NdotL + 0.001;
  double spec = (D * G * F) / denominator;
  return spec * hitColor;
}

static const double kOneEm6f = (double)1e-6f;  // mMax(1e-6f, nh.z), Material.cuh:140,222

// brdf(), Material.cuh:201-235
static void brdfSample(const OMat &m, const d3 &rayDir, const d3 &normal, d3 &direction, double &pdf,
                       double r1, double r2) {
  d3 w = normal;
  certMargin(std::fabs(w.z) - 0.99999, kCertFrameEps, kCertFrame);
  d3 a = (std::fabs(w.z) > 0.99999) ? mk(1.0, 0.0, 0.0) : mk(0.0, 0.0, 1.0);
  d3 u = Unit(Cross(a, w));
  d3 v = Cross(w, u);
  d3 uo = -Unit(rayDir);
  d3 Wo = mk(Dot(uo, u), Dot(uo, v), Dot(uo, w));
  d3 Wh = Unit(mk(m.ax * Wo.x, m.ay * Wo.y, Wo.z));
  certMargin(Wh.z - 0.99999, kCertFrameEps, kCertFrame);
  d3 T1 = (Wh.z < 0.99999) ? Unit(Cross(mk(0, 0, 1), Wh)) : Unit(Cross(mk(1, 0, 0), Wh));
  d3 T2 = Cross(Wh, T1);
  double r = std::sqrt(r1);
  double theta = 2 * PI * r2;
  double px = r * std::cos(theta), py = r * std::sin(theta);
  double h = std::sqrt(1 - px * px);
  double lerpX = (1 + Wh.z) / 2.0;
  py = (1.0 - lerpX) * h + lerpX * py;
  double pz = std::sqrt(mMax(0.0, 1.0 - (px * px + py * py)));
  d3 nh = px * T1 + py * T2 + pz * Wh;
  d3 Wm = Unit(mk(m.ax * nh.x, m.ay * nh.y, mMax(kOneEm6f, nh.z)));
  d3 Wi = -Wo + 2 * Dot(Wo, Wm) * Wm;
  direction = Unit(Wi.x * u + Wi.y * v + Wi.z * w);
  certMargin(Dot(direction, normal), kCertDot, kCertSampleSide);
  if (Dot(direction, normal) < 0.0) {
    pdf = -1.0;
    return;
  }
  d3 H = Unit(direction + uo);
  pdf = GeometrySmith1(m, normal, uo) / Wo.z * DistributionGGX(m, normal, H) / 4.0;
}

// Material::sampleLight, Material.cuh:95-193 (+ LAMBERT)
static void sampleLight(const OMat &m, const d3 &rayDir, const d3 &normal, d3 &direction, d3 &wm,
                        double &pdf, double r1, double r2, double r3, bool frontFace) {
  if (m.type == CPT_M_SPECULAR_DIELECTRIC) {
    double rr = frontFace ? m.eta : (1.0 / m.eta);
    d3 ud = Unit(rayDir);
    double R = FresnelDielectric(normal, -ud, rr);
    certMargin(r3 - R, kCertProb, kCertChoice);
    if (r3 < R) {
      direction = ud - 2.0 * Dot(ud, normal) * normal;
      pdf = R;
    } else {
      d3 vPerp = 1.0 / rr * (ud + Dot(-ud, normal) * normal);
      d3 vPar = -std::sqrt(std::fabs(1.0 - SquaredLength(vPerp))) * normal;
      direction = Unit(vPerp + vPar);
      pdf = 1.0 - R;
    }
  } else if (m.type == CPT_M_ROUGH_DIELECTRIC) {
    d3 w = normal;
    certMargin(std::fabs(w.x) - 0.9, kCertFrameEps, kCertFrame);
    d3 a = (std::fabs(w.x) > 0.9) ? mk(0.0, 1.0, 0.0) : mk(1.0, 0.0, 0.0);
    d3 u = Unit(Cross(w, a));
    d3 v = Cross(w, u);
    d3 uo = -Unit(rayDir);
    d3 Wo = mk(Dot(uo, u), Dot(uo, v), Dot(uo, w));
    d3 Wh = Unit(mk(m.ax * Wo.x, m.ay * Wo.y, Wo.z));
    certMargin(Wh.z - 0.99999, kCertFrameEps, kCertFrame);
    d3 T1 = (Wh.z < 0.99999) ? Unit(Cross(mk(0, 0, 1), Wh)) : mk(1, 0, 0);
    d3 T2 = Cross(Wh, T1);
    double r = std::sqrt(r1);
    double theta = 2 * PI * r2;
    double px = r * std::cos(theta), py = r * std::sin(theta);
    double h = std::sqrt(1 - px * px);
    double lerpX = (1 + Wh.z) / 2.0;
    py = (1.0 - lerpX) * h + lerpX * py;
    double pz = std::sqrt(mMax(0.0, 1.0 - (px * px + py * py)));
    d3 nh = px * T1 + py * T2 + pz * Wh;
    d3 Wm = Unit(mk(m.ax * nh.x, m.ay * nh.y, mMax(kOneEm6f, nh.z)));
    d3 sn = Unit(Wm.x * u + Wm.y * v + Wm.z * w);
    wm = sn;
    double rr = frontFace ? m.eta : (1.0 / m.eta);
    double R = FresnelDielectric(sn, -uo, rr);
    certMargin(r3 - R, kCertProb, kCertChoice);
    if (r3 < R) {
      brdfSample(m, rayDir, normal, direction, pdf, r1, r2);
      pdf *= R;
    } else {
      d3 vPerp = 1.0 / rr * (uo + Dot(-uo, sn) * sn);
      d3 vPar = -std::sqrt(std::fabs(1.0 - SquaredLength(vPerp))) * sn;
      d3 rd = Unit(vPerp + vPar);
      d3 Wi = mk(Dot(rd, u), Dot(rd, v), Dot(rd, w));
      direction = rd;
      certMargin(Dot(direction, normal), kCertDot, kCertSampleSide);
      if (Dot(direction, normal) > 0.0) {
        pdf = -1.0;
        return;
      }
      double mbase = Dot(-Wi, Wm) + Dot(Wo, Wm) / rr;
      double denom = mbase * mbase;
      double dwm = std::fabs(Dot(Wi, Wm)) / denom;
      pdf = GeometrySmith1(m, normal, uo) / Wo.z * DistributionGGX(m, normal, sn) * Dot(Wo, Wm) * dwm *
            (1.0 - R);
    }
  } else if (m.type == CPT_M_LAMBERT) {
    // the reference's commented cosine sampler, Material.cuh:177-188
    double sinT = std::sqrt(1.0 - r1);
    double cosT = std::sqrt(r1);
    d3 w = normal;
    certMargin(std::fabs(w.x) - 0.9, kCertFrameEps, kCertFrame);
    d3 a = (std::fabs(w.x) > 0.9) ? mk(0.0, 1.0, 0.0) : mk(1.0, 0.0, 0.0);
    d3 u = Unit(Cross(w, a));
    d3 v = Cross(w, u);
    direction = std::cos(2.0 * PI * r2) * sinT * u + std::sin(2.0 * PI * r2) * sinT * v + cosT * w;
    pdf = cosT / PI;
  } else {  // OPAQUE, GGX
    brdfSample(m, rayDir, normal, direction, pdf, r1, r2);
  }
}

struct Cam {
  d3 lf, pc, dh, dv, bg;
  int w, h;
};

struct Params {
  uint64_t seed;
  int spp, maxDepth, rrDepth, profile, sampleBase;
  double rrProb;
};

struct Hit {
  int id = -1;
  double t = INF;
  double tErr = 0.0;  // absolute uncertainty of t under the certificate's perturbations
};

// Absolute uncertainty of the GPU's FP32 copy of a ray origin o reached
// after the path length so far.
static double originErr(const d3 &o) {
  const double om = std::max(std::max(std::fabs(o.x), std::fabs(o.y)), std::max(std::fabs(o.z), 1.0));
  return kCertOrigin * om + kCertSeg * tlPathLen;
}

// Closest hit in the reference's traversal order (BVH.cuh:282-332); with a
// certificate, the brute-force answer must be robust and the same primitive.
static Hit trace(const Scene &s, const d3 &o, const d3 &d) {
  Hit h;
  double t;
  h.id = s.traceRef(o, d, 0.001, INF, t, nullptr);
  h.t = t;
  if (tlCert && tlCert->why == kCertOk) {
    bool ok = true;
    double tb;
    // the GPU's ray starts at an FP32 point (a few ulp of its coordinates
    // off the FP64 one) with an FP32 direction
    const int id = s.bruteClosest(o, d, 0.001, INF, kCertEbHit, kCertEtHit, tb, ok, originErr(o), kCertDir, &h.tErr);
    certNeed(ok && id == h.id, kCertHit);
  }
  return h;
}

static OMat matOf(const Scene &s, const Prim &p) {
  OMat m;
  if (p.type == CPT_OBJ_LIGHT) {
    m.color = p.le;
    m.ax = m.ay = 1.0;
    m.eta = 1.51;
    m.type = CPT_M_LIGHT;
    return m;
  }
  const cpt_material &c = s.materials[p.mat];
  m.color = mk(c.color[0], c.color[1], c.color[2]);
  m.ax = c.alpha_x;
  m.ay = c.alpha_y;
  m.eta = c.eta;
  m.type = c.type;
  return m;
}

// ---- heterogeneous medium (no reference counterpart) ----------------------
static const int kMaxCollisions = 100000;

// [t0, t1] = [tmin, tmax] clipped to the medium box by the slab test (NaN
// slab ends are dropped, as fmax/fmin do); false when empty.
// t1Err (optional, in: the uncertainty of tmax; out: of t1): a slab plane
// crossing moves by the origin uncertainty over |d_k|.
static bool clipBox(const Medium &m, const d3 &o, const d3 &d, double tmin, double tmax, double &t0, double &t1,
                    double *t1Err = nullptr) {
  t0 = tmin;
  t1 = tmax;
  const double ox[3] = {o.x, o.y, o.z}, dx[3] = {d.x, d.y, d.z};
  const double eo = tlCert ? originErr(o) : 0.0;
  double e0 = 0.0, e1 = t1Err ? *t1Err : 0.0;
  for (int k = 0; k < 3; k++) {
    const double inv = 1.0 / dx[k];
    double ta = (m.bmin[k] - ox[k]) * inv, tb = (m.bmax[k] - ox[k]) * inv;
    if (ta > tb) std::swap(ta, tb);
    const double ek = eo * std::fabs(inv) + kCertDir * std::fabs(ta);
    if (ta > t0) e0 = ek;
    if (tb < t1) e1 = eo * std::fabs(inv) + kCertDir * std::fabs(tb);
    t0 = std::fmax(t0, ta);
    t1 = std::fmin(t1, tb);
  }
  certDist(t0, t1, kCertMedium, e0 + e1);
  if (t1Err) *t1Err = e1;
  return t0 < t1;
}

static double densityAt(const Medium &m, const d3 &p) {
  const double px[3] = {p.x, p.y, p.z};
  int64_t ix[3];
  for (int k = 0; k < 3; k++) {
    const double g = (px[k] - m.bmin[k]) / (m.bmax[k] - m.bmin[k]) * m.n[k];
    // a coordinate inside the grid close to a voxel boundary may round into
    // the neighbouring voxel in FP32 (position uncertainty in voxels)
    const double ev = kCertVoxelEps + tlPosErr * m.n[k] / (m.bmax[k] - m.bmin[k]);
    if (g > ev && g < m.n[k] - ev) certMargin(g - std::nearbyint(g), ev, kCertVoxel);
    double f = std::floor(g);
    f = std::fmin(std::fmax(f, 0.0), (double)(m.n[k] - 1));
    ix[k] = (int64_t)f;
  }
  return m.density[(ix[2] * m.n[1] + ix[1]) * m.n[0] + ix[0]];
}

// Delta tracking: exponential steps of the majorant; a tentative collision
// is real with probability density(x).  One draw per step, one per test.
static bool deltaTrack(const Medium &m, Stream &rng, const d3 &o, const d3 &d, double t0, double t1, d3 &x,
                       double t1Err = 0.0) {
  const double inv = 1.0 / (m.sigma * Length(d));
  double t = t0;
  const double eo = tlCert ? originErr(o) : 0.0;
  for (int guard = 0; guard < kMaxCollisions; guard++) {
    t -= std::log(1.0 - rng.next()) * inv;
    certDist(t, t1, kCertMedium, t1Err);
    tlPosErr = eo + kCertSeg * std::fabs(t) * Length(d);
    if (!(t < t1)) return false;
    const d3 p = o + t * d;
    if (rng.next() < densityAt(m, p)) {
      x = p;
      return true;
    }
  }
  return false;
}

// Ratio tracking: product of (1 - density) at the tentative collisions.
static double ratioTrack(const Medium &m, Stream rng, const d3 &o, const d3 &d, double tmin, double tmax,
                         double tmaxErr = 0.0) {
  double t0, t1, t1Err = tmaxErr;
  if (!clipBox(m, o, d, tmin, tmax, t0, t1, &t1Err)) return 1.0;
  const double inv = 1.0 / (m.sigma * Length(d));
  double t = t0, tr = 1.0;
  const double eo = tlCert ? originErr(o) : 0.0;
  for (int guard = 0; guard < kMaxCollisions; guard++) {
    t -= std::log(1.0 - rng.next()) * inv;
    certDist(t, t1, kCertMedium, t1Err);
    tlPosErr = eo + kCertSeg * std::fabs(t) * Length(d);
    if (!(t < t1)) break;
    tr *= 1.0 - densityAt(m, o + t * d);
  }
  return tr;
}

// Russian roulette / depth cap after a vertex's NEE (Render.cu:189-204):
//   REFERENCE (STATIC, REAL_TIME_NOT_SAMPLE): roulette p = 0.7 at every
//     vertex, no cap (the depth counter never advances there);
//   REAL_TIME: `depth++ >= 2` ends the path, so at most 3 vertices, no roulette;
//   NORMAL / DEPTH: `depth++ >= 1`, at most 2 vertices;
//   CONFIG (no reference counterpart): max_depth vertices, roulette with
//     rr_prob from vertex rr_depth on.
// Returns false when the path ends; P_RR is the survival probability used.
static bool continuePath(const Params &P, Stream &rng, int depth, double &P_RR) {
  P_RR = 1.0;
  switch (P.profile) {
    case CPT_PROFILE_REFERENCE:
      P_RR = 0.7;
      return !(rng.next() > P_RR);
    case CPT_PROFILE_REAL_TIME:
      return depth < 2;
    case CPT_PROFILE_NORMAL:
    case CPT_PROFILE_DEPTH:
      return depth < 1;
    default:
      if (depth + 1 >= P.maxDepth) return false;
      if (depth >= P.rrDepth) {
        P_RR = P.rrProb;
        return !(rng.next() > P_RR);
      }
      return true;
  }
}

// The homogeneous volume object (Render.cu:64-135): sigma_maj = 25, free
// flights from the entry point hp, roulette 0.8 per collision with weight
// exp(-25 t) / 0.8, isotropic scattering; each new position is taken from
// the entry point hp along the new direction (Render.cu:125).  On return
// (ro, rd) is the ray the path continues with and T its throughput.
// nColl (optional) counts the collisions that survived the roulette.
static void volumeWalk(const Scene &s, Stream &rng, d3 &ro, d3 &rd, d3 &T, double tHit, int *nColl = nullptr) {
  const double sigmaMaj = 25.0;
  d3 hp = ro + tHit * rd;
  ro = hp;
  Hit vr = trace(s, ro, rd);
  if (vr.id < 0) return;
  double dMax = vr.t, dMaxErr = vr.tErr;
  double t = -std::log(1.0 - rng.next()) / sigmaMaj;
  d3 pos = hp + t * rd;
  ro = pos;
  certDist(t, dMax, kCertVolume, dMaxErr);
  const double lenEntry = tlPathLen;
  while (t < dMax) {
    const double P_RR = 0.8;
    if (rng.next() > P_RR) break;
    if (nColl) (*nColl)++;
    tlPathLen = lenEntry + t * Length(rd);
    T = T * (std::exp(-t * sigmaMaj) / P_RR);
    double z = 1.0 - 2.0 * rng.next();
    double r = std::sqrt(1.0 - z * z);
    double phi = 2.0 * PI * rng.next();
    d3 dir = Unit(mk(r * std::cos(phi), r * std::sin(phi), z));
    ro = pos;
    rd = dir;
    pos = hp + t * dir;
    Hit v2 = trace(s, ro, rd);
    if (v2.id >= 0) {
      dMax = v2.t;
      dMaxErr = v2.tErr;
    }
    t = -std::log(1.0 - rng.next()) / sigmaMaj;
    if (v2.id >= 0) certDist(t, dMax, kCertVolume, dMaxErr);
    if (t > dMax || v2.id < 0) break;
  }
}

// One sample of the bounce loop, Render.cu:42-219.  Returns the radiance.
// streamPixel: the pixel's index in the random stream (view * h * w + pixel
// for a multi-view render).
// gb (optional): the G-buffer of Render.cu:158-165, written at the first
// scattering vertex (surface or medium collision; normal 0 in a medium).
struct GBuf {
  bool set = false;
  d3 n{0, 0, 0};
  double dist = 0;
};
d3 renderSample(const Scene &s, const Cam &cam, const Params &P, int i, int j, uint32_t sample,
                uint32_t streamPixel, GBuf *gb = nullptr) {
  Stream rng{(uint32_t)P.seed, (uint32_t)(P.seed >> 32), streamPixel, sample, 0};
  double rx = rng.next(), ry = rng.next();
  d3 pixel = cam.pc + (rx - 0.5 + i) * cam.dh + (ry - 0.5 + j) * cam.dv;
  d3 ro = cam.lf, rd = pixel - cam.lf;
  d3 T = mk(1, 1, 1), L = mk(0, 0, 0);
  int depth = 0;
  bool hasSampleLight = false;
  const int nl = s.nLights;
  tlPathLen = 0.0;
  while (true) {
    Hit hr = trace(s, ro, rd);
    const double lenBefore = tlPathLen;
    if (hr.id >= 0) tlPathLen += hr.t * Length(rd);
    if (s.med.on) {
      double t0, t1;
      d3 x;
      double t1Err = hr.id >= 0 ? hr.tErr : 0.0;
      if (clipBox(s.med, ro, rd, 0.001, hr.id >= 0 ? hr.t : INF, t0, t1, &t1Err) &&
          deltaTrack(s.med, rng, ro, rd, t0, t1, x, t1Err)) {
        tlPathLen = lenBefore + Length(x - ro);
        // real collision: NEE with the isotropic phase function, roulette,
        // new direction; weight albedo
        const d3 alb = s.med.albedo;
        if (gb && !gb->set) {
          gb->set = true;
          gb->dist = Length(cam.lf - x);
        }
        for (int k = 0; k < nl; k++) {
          const Prim &lp = s.prims[k];
          double lw = lp.w;
          double lh = P.profile == CPT_PROFILE_CONFIG ? lp.h : lp.w;
          const double u1 = rng.next();
          const double u2 = rng.next();
          d3 lightPos = lp.c + (u1 - 0.5) * lw * lp.eu + (u2 - 0.5) * lh * lp.ev;
          d3 dir = Unit(lightPos - x);
          certMargin(Dot(-dir, lp.n), kCertDot, kCertNeeSide);
          Hit sh;
          if (Dot(-dir, lp.n) > 0.0) sh = trace(s, x, dir);
          if (Dot(-dir, lp.n) > 0.0 && sh.id == k) {
            Stream srng{rng.k0, rng.k1, rng.pix, rng.smp, 0, 0x80000000u | rng.cnt};
            const double tr = ratioTrack(s.med, srng, x, dir, 0.001, sh.t, sh.tErr);
            d3 c = lp.le * alb * (0.25 / PI) * Dot(-dir, lp.n) / SquaredLength(lightPos - x) * (lw * lh);
            L = L + T * c * tr;
          }
          hasSampleLight = true;
        }
        double P_RR = 1.0;
        if (!continuePath(P, rng, depth, P_RR)) break;
        depth++;
        const double r1 = rng.next();
        const double r2 = rng.next();
        rng.next();  // third draw of a vertex, unused by the phase function
        const double z = 1.0 - 2.0 * r1;
        const double r = std::sqrt(std::fmax(0.0, 1.0 - z * z));
        const double phi = 2.0 * PI * r2;
        T = T * alb / P_RR;
        ro = x;
        rd = mk(r * std::cos(phi), r * std::sin(phi), z);
        continue;
      }
    }
    if (hr.id >= 0 && s.prims[hr.id].type != CPT_OBJ_LIGHT &&
        s.materials[s.prims[hr.id].mat].type == CPT_M_VOLUME) {
      volumeWalk(s, rng, ro, rd, T, hr.t);
      continue;
    }
    if (hr.id < 0) {
      L = L + T * cam.bg;
      break;
    }
    const Prim &p = s.prims[hr.id];
    const OMat m = matOf(s, p);
    d3 hp = ro + hr.t * rd;
    if (m.type == CPT_M_LIGHT) {
      if (!hasSampleLight) L = L + T * m.color;
      break;
    }
    d3 outw;
    if (p.type == CPT_OBJ_TRIANGLE) outw = Unit(Cross(p.p1 - p.p0, p.p2 - p.p0));
    else outw = (hp - p.c) / p.r;
    certMargin(Dot(Unit(rd), outw), kCertDot, kCertFace);
    bool front = Dot(rd, outw) < 0;
    d3 n = front ? outw : -outw;
    if (gb && !gb->set) {
      gb->set = true;
      gb->n = n;
      gb->dist = Length(cam.lf - hp);
    }
    const d3 zero = mk(0, 0, 0);
    if (m.type == CPT_M_OPAQUE || m.type == CPT_M_LAMBERT || m.type == CPT_M_GGX) {
      for (int k = 0; k < nl; k++) {
        const Prim &lp = s.prims[k];
        double lw = lp.w;
        double lh = P.profile == CPT_PROFILE_CONFIG ? lp.h : lp.w;  // Render.cu:170-171
        double pdfLight = 1.0 / (lw * lh);
        // two draws in a fixed order: u1 along edgeU, then u2 along edgeV (the
        // operands of one expression have no defined evaluation order in C++)
        const double u1 = rng.next();
        const double u2 = rng.next();
        d3 lightPos = lp.c + (u1 - 0.5) * lw * lp.eu + (u2 - 0.5) * lh * lp.ev;
        d3 dir = Unit(lightPos - hp);
        certMargin(Dot(dir, n), kCertDot, kCertNeeSide);
        certMargin(Dot(-dir, lp.n), kCertDot, kCertNeeSide);
        Hit sh;
        if (Dot(dir, n) > 0.0 && Dot(-dir, lp.n) > 0.0) sh = trace(s, hp, dir);
        if (Dot(dir, n) > 0.0 && Dot(-dir, lp.n) > 0.0 && sh.id == k) {
          d3 c = lp.le * fr(m, m.color, rd, n, dir, zero, front) * Dot(dir, n) * Dot(-dir, lp.n) /
                 SquaredLength(lightPos - hp) / pdfLight;
          if (s.med.on) {
            Stream srng{rng.k0, rng.k1, rng.pix, rng.smp, 0, 0x80000000u | rng.cnt};
            c = c * ratioTrack(s.med, srng, hp, dir, 0.001, sh.t, sh.tErr);
          }
          L = L + T * c;
        }
        hasSampleLight = true;
      }
    }
    double P_RR = 1.0;
    if (!continuePath(P, rng, depth, P_RR)) break;
    depth++;
    double r1 = rng.next(), r2 = rng.next(), r3 = rng.next();
    d3 dir, wm = zero;
    double pdf = -1.0;
    sampleLight(m, rd, n, dir, wm, pdf, r1, r2, r3, front);
    if (pdf <= 0.0) break;
    T = T * fr(m, m.color, rd, n, dir, wm, front) * std::fabs(Dot(dir, n)) / pdf / P_RR;
    ro = hp;
    rd = dir;
  }
  return L;
}

}  // namespace orc

using namespace orc;

extern "C" {

// Medium of the scene (density nx*ny*nz floats in [0,1], x fastest); a
// null density removes it.
int orc_scene_set_medium(void *sp, const float *density, int nx, int ny, int nz, const double bmin[3],
                         const double bmax[3], double sigma, const double albedo[3]) {
  Scene *s = (Scene *)sp;
  Medium &m = s->med;
  if (!density || nx <= 0 || ny <= 0 || nz <= 0) {
    m.on = false;
    return 0;
  }
  m.density.assign(density, density + (size_t)nx * ny * nz);
  m.n[0] = nx;
  m.n[1] = ny;
  m.n[2] = nz;
  for (int k = 0; k < 3; k++) {
    m.bmin[k] = (double)(float)bmin[k];
    m.bmax[k] = (double)(float)bmax[k];
  }
  m.sigma = (double)(float)sigma;
  m.albedo = mk((double)(float)albedo[0], (double)(float)albedo[1], (double)(float)albedo[2]);
  m.on = true;
  return 0;
}

// n ratio-tracking estimates of the transmittance over [tmin, tmax] along
// o + t d (estimate i uses pixel i, sample 0 of the shadow-ray stream).
int orc_medium_transmittance(void *sp, const double o[3], const double d[3], double tmin, double tmax,
                             uint64_t seed, int n, double *out) {
  Scene *s = (Scene *)sp;
  if (!s->med.on) return -1;
  const d3 O = mk(o[0], o[1], o[2]), D = mk(d[0], d[1], d[2]);
  for (int i = 0; i < n; i++) {
    Stream r{(uint32_t)seed, (uint32_t)(seed >> 32), (uint32_t)i, 0, 0, 0x80000000u};
    out[i] = ratioTrack(s->med, r, O, D, tmin, tmax);
  }
  return 0;
}

// n delta-tracking runs over [tmin, tmax]: the collision t, or -1 for none.
int orc_medium_collide(void *sp, const double o[3], const double d[3], double tmin, double tmax, uint64_t seed,
                       int n, double *out) {
  Scene *s = (Scene *)sp;
  if (!s->med.on) return -1;
  const d3 O = mk(o[0], o[1], o[2]), D = mk(d[0], d[1], d[2]);
  for (int i = 0; i < n; i++) {
    Stream r{(uint32_t)seed, (uint32_t)(seed >> 32), (uint32_t)i, 0, 0};
    double t0, t1;
    d3 x;
    out[i] = -1.0;
    if (clipBox(s->med, O, D, tmin, tmax, t0, t1) && deltaTrack(s->med, r, O, D, t0, t1, x))
      out[i] = Dot(x - O, D) / Dot(D, D);
  }
  return 0;
}

void orc_philox4x32_10(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4]) {
  PhiloxOut o = philox4x32_10(ctr, k0, k1);
  std::memcpy(out, o.v, 16);
}

// Renders one camera image: rgb gets the mean radiance per pixel (h*w*3
// doubles, row 0 = bottom), or for the NORMAL / DEPTH profiles the
// reference's display value of the first vertex (Render.cu:239-251:
// (normal + 0.5) / 2, clamp01(distance / 50); a pixel with no such vertex
// shows the cleared G-buffer, normal 0 and distance INF).  perSample
// (optional) gets each sample's radiance (npix*spp*3); cert (optional) each
// sample's path certificate (npix*spp, 0 = certified, else the CertWhy of
// the first decision that was not robust).  Pixel (i, j) uses stream pixel
// pixelOffset + j * w + i (a view of a multi-view render).
static int renderView(Scene *s, const cpt_camera *cam, const cpt_render_params *prm, double *rgb,
                      double *perSample, uint8_t *cert, int nthreads, int64_t pixelOffset) {
  if (!s->built) return -1;
  Cam c;
  c.lf = mk(cam->look_from[0], cam->look_from[1], cam->look_from[2]);
  c.pc = mk(cam->pixel_center[0], cam->pixel_center[1], cam->pixel_center[2]);
  c.dh = mk(cam->delta_h[0], cam->delta_h[1], cam->delta_h[2]);
  c.dv = mk(cam->delta_v[0], cam->delta_v[1], cam->delta_v[2]);
  c.bg = mk(cam->background[0], cam->background[1], cam->background[2]);
  c.w = cam->width;
  c.h = cam->height;
  Params P{prm->seed, prm->spp, prm->max_depth, prm->rr_depth, prm->profile, prm->sample_base, prm->rr_prob};
  const bool display = P.profile == CPT_PROFILE_NORMAL || P.profile == CPT_PROFILE_DEPTH;
  if (display) P.spp = 1;  // Render.cu:38
  const int64_t npix = (int64_t)c.w * c.h;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t px = 0; px < npix; px++) {
    int i = (int)(px % c.w), j = (int)(px / c.w);
    d3 acc = mk(0, 0, 0);
    GBuf gb;
    for (int k = 0; k < P.spp; k++) {
      Cert ct;
      tlCert = cert ? &ct : nullptr;
      d3 r = renderSample(*s, c, P, i, j, (uint32_t)(P.sampleBase + k), (uint32_t)(pixelOffset + px),
                          display ? &gb : nullptr);
      tlCert = nullptr;
      if (cert) cert[px * P.spp + k] = ct.why;
      if (perSample) {
        double *o = perSample + 3 * (px * P.spp + k);
        o[0] = r.x; o[1] = r.y; o[2] = r.z;
      }
      acc = acc + r;
    }
    acc = acc / (double)P.spp;
    if (display) {
      if (P.profile == CPT_PROFILE_NORMAL) {
        acc = (gb.n + mk(0.5, 0.5, 0.5)) / 2.0;
      } else {
        const double v = gb.set ? gb.dist / 50.0 : INF / 50.0;
        const double cl = v > 1.0 ? 1.0 : (v < 0.0 ? 0.0 : v);
        acc = mk(cl, cl, cl);
      }
    }
    rgb[3 * px] = acc.x;
    rgb[3 * px + 1] = acc.y;
    rgb[3 * px + 2] = acc.z;
  }
  return 0;
}

int orc_render(void *sp, const cpt_camera *cam, const cpt_render_params *prm, double *rgb, double *perSample,
               int nthreads) {
  return renderView((Scene *)sp, cam, prm, rgb, perSample, nullptr, nthreads, 0);
}

int orc_render_view(void *sp, const cpt_camera *cam, const cpt_render_params *prm, double *rgb,
                    double *perSample, int nthreads, int64_t pixelOffset) {
  return renderView((Scene *)sp, cam, prm, rgb, perSample, nullptr, nthreads, pixelOffset);
}

int orc_render_view_cert(void *sp, const cpt_camera *cam, const cpt_render_params *prm, double *rgb,
                         double *perSample, uint8_t *cert, int nthreads, int64_t pixelOffset) {
  return renderView((Scene *)sp, cam, prm, rgb, perSample, cert, nthreads, pixelOffset);
}

// Volume-object walk probe (Render.cu:64-135): for each of n streams
// (pixel i, sample 0) a ray o + t d that hits a volume-material primitive at
// distance tHit: the continuing ray and throughput, and the number of
// collisions that survived the roulette.
int orc_volume_walk(void *sp, const double o[3], const double d[3], double tHit, uint64_t seed, int n,
                    double *roOut, double *rdOut, double *thrOut, int32_t *collOut) {
  Scene *s = (Scene *)sp;
  if (!s->built) return -1;
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; i++) {
    Stream rng{(uint32_t)seed, (uint32_t)(seed >> 32), (uint32_t)i, 0, 0};
    d3 ro = mk(o[0], o[1], o[2]), rd = mk(d[0], d[1], d[2]), T = mk(1, 1, 1);
    int nc = 0;
    volumeWalk(*s, rng, ro, rd, T, tHit, &nc);
    roOut[3 * i] = ro.x; roOut[3 * i + 1] = ro.y; roOut[3 * i + 2] = ro.z;
    rdOut[3 * i] = rd.x; rdOut[3 * i + 1] = rd.y; rdOut[3 * i + 2] = rd.z;
    thrOut[i] = T.x;
    collOut[i] = nc;
  }
  return 0;
}

// BSDF probes for closed-form tests: evaluates fr and one sample.
int orc_bsdf(const cpt_material *mat, const double rayDir[3], const double normal[3], int frontFace,
             const double r[3], double dirOut[3], double *pdfOut, double frOut[3]) {
  OMat m;
  m.color = mk(mat->color[0], mat->color[1], mat->color[2]);
  m.ax = mat->alpha_x;
  m.ay = mat->alpha_y;
  m.eta = mat->eta;
  m.type = mat->type;
  d3 rd = mk(rayDir[0], rayDir[1], rayDir[2]);
  d3 n = mk(normal[0], normal[1], normal[2]);
  d3 dir, wm = mk(0, 0, 0);
  double pdf = -1;
  sampleLight(m, rd, n, dir, wm, pdf, r[0], r[1], r[2], frontFace != 0);
  d3 f = fr(m, m.color, rd, n, dir, wm, frontFace != 0);
  dirOut[0] = dir.x; dirOut[1] = dir.y; dirOut[2] = dir.z;
  *pdfOut = pdf;
  frOut[0] = f.x; frOut[1] = f.y; frOut[2] = f.z;
  return 0;
}

// fr for a given outgoing direction (wm = 0).
int orc_bsdf_eval(const cpt_material *mat, const double rayDir[3], const double normal[3], int frontFace,
                  const double dir[3], double frOut[3]) {
  OMat m;
  m.color = mk(mat->color[0], mat->color[1], mat->color[2]);
  m.ax = mat->alpha_x;
  m.ay = mat->alpha_y;
  m.eta = mat->eta;
  m.type = mat->type;
  d3 f = fr(m, m.color, mk(rayDir[0], rayDir[1], rayDir[2]), mk(normal[0], normal[1], normal[2]),
            mk(dir[0], dir[1], dir[2]), mk(0, 0, 0), frontFace != 0);
  frOut[0] = f.x; frOut[1] = f.y; frOut[2] = f.z;
  return 0;
}

// GGX normal distribution D(h) for normal n (the pdf normalisation test).
double orc_ggx_d(const cpt_material *mat, const double n[3], const double h[3]) {
  OMat m;
  m.ax = mat->alpha_x;
  m.ay = mat->alpha_y;
  return DistributionGGX(m, mk(n[0], n[1], n[2]), mk(h[0], h[1], h[2]));
}

}  // extern "C"

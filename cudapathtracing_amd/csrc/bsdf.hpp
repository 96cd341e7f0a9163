// bsdf.hpp - FP32 materials for the wavefront integrator.
//
// Restates CudaPathTracing/Material.cuh:
//   fr            Material.cuh:24-93    (evaluation)
//   sampleLight   Material.cuh:95-193   (sampling)
//   brdf          Material.cuh:201-235  (VNDF reflection sampling, Heitz 2018)
//   DistributionGGX / lambda / GeometrySmith  Material.cuh:250-285
//   FresnelConductor / FresnelDielectric      Material.cuh:287-318
// and adds CPT_M_LAMBERT (cosine hemisphere, the commented code at
// Material.cuh:177-188) and CPT_M_GGX (isotropic GGX, Schlick F0).
// The GGX terms use algebraically equal forms that stay finite when the
// half vector equals the normal (the reference divides 0/0 there):
//   e        = (x^2/ax^2 + y^2/ay^2) / z^2        (tan^2 * (cos^2/ax^2+sin^2/ay^2))
//   a^2tan^2 = (ax^2 x^2 + ay^2 y^2) / max(N.R,0)^2
#pragma once
#include "cpt_device.hpp"
#include "internal.hpp"

namespace cpt {

constexpr float kPi = 3.14159265358979323846f;

struct Frame {
  f3 u, v, w;
};

// Frame of brdf()/worldToLocal (Material.cuh:203-206, 238-248).
__device__ __forceinline__ Frame frameZ(f3 n) {
  f3 a = fabsf(n.z) > 0.99999f ? mkf3(1, 0, 0) : mkf3(0, 0, 1);
  f3 u = unit3(cross3(a, n));
  return Frame{u, cross3(n, u), n};
}
// Frame of the rough-dielectric branch (Material.cuh:121-124) and the
// cosine sampler (Material.cuh:181-184).
__device__ __forceinline__ Frame frameX(f3 n) {
  f3 a = fabsf(n.x) > 0.9f ? mkf3(0, 1, 0) : mkf3(1, 0, 0);
  f3 u = unit3(cross3(n, a));
  return Frame{u, cross3(n, u), n};
}
__device__ __forceinline__ f3 toLocal(const Frame &f, f3 v) { return mkf3(dot3(v, f.u), dot3(v, f.v), dot3(v, f.w)); }
__device__ __forceinline__ f3 toWorld(const Frame &f, f3 v) { return f.u * v.x + f.v * v.y + f.w * v.z; }

__device__ __forceinline__ float ggxD(const Frame &f, f3 h, float ax, float ay) {
  f3 w = toLocal(f, h);
  float c2 = w.z * w.z;
  float e = (w.x * w.x / (ax * ax) + w.y * w.y / (ay * ay)) / c2;
  float ope = 1.0f + e;
  return 1.0f / (kPi * ax * ay * c2 * c2 * ope * ope);
}

__device__ __forceinline__ float ggxLambda(const Frame &f, f3 r, float ax, float ay) {
  f3 w = toLocal(f, r);
  float c = fmaxf(w.z, 0.0f);
  float a2t2 = (ax * ax * w.x * w.x + ay * ay * w.y * w.y) / (c * c);
  return (sqrtf(1.0f + a2t2) - 1.0f) * 0.5f;
}

__device__ __forceinline__ float fresnelDielectric(f3 n, f3 v, float eta) {
  float ci = fabsf(dot3(n, v));
  float si = sqrtf(fmaxf(0.0f, 1.0f - ci * ci));
  float st = si / eta;
  if (st >= 1.0f) return 1.0f;
  float ct = sqrtf(fmaxf(0.0f, 1.0f - st * st));
  float rpa = (eta * ci - ct) / (eta * ci + ct);
  float rpe = (ci - eta * ct) / (ci + eta * ct);
  return (rpa * rpa + rpe * rpe) * 0.5f;
}

// |r_parl|^2 and |r_perp|^2 of the complex-index conductor.
__device__ __forceinline__ float fresnelConductor(float ci, float ct, float n, float k) {
  float a = n * ci - ct, b = k * ci, c = n * ci + ct;
  float den1 = c * c + b * b;
  float p1 = den1 > 0.0f ? (a * a + b * b) / den1 : 0.0f;
  float d = ci - n * ct, e = k * ct, g = ci + n * ct;
  float den2 = g * g + e * e;
  float p2 = den2 > 0.0f ? (d * d + e * e) / den2 : 0.0f;
  return (p1 + p2) * 0.5f;
}

struct ShadeMat {
  f3 color;
  float ax, ay, eta;
  int type;
};

// Material::fr (Material.cuh:24-93); V = -ray direction (unit), L unit.
__device__ __forceinline__ f3 bsdfEval(const ShadeMat &m, f3 V, f3 L, f3 n, f3 wm, bool frontFace) {
  if (m.type == CPT_M_LAMBERT) return dot3(L, n) > 0.0f ? m.color * (1.0f / kPi) : mkf3(0, 0, 0);
  const bool reflect = dot3(V, n) * dot3(L, n) > 0.0f;
  if (m.type == CPT_M_SPECULAR_DIELECTRIC) {
    float rr = frontFace ? m.eta : 1.0f / m.eta;
    float ct = fabsf(dot3(n, L));
    float F = fresnelDielectric(n, V, rr);
    return m.color * ((reflect ? F : 1.0f - F) / ct);
  }
  Frame fz = frameZ(n);
  if (m.type == CPT_M_GGX) {
    float nl = dot3(n, L), nv = dot3(n, V);
    if (!(nl > 0.0f) || !(nv > 0.0f)) return mkf3(0, 0, 0);
    f3 H = unit3(V + L);
    float D = ggxD(fz, H, m.ax, m.ay);
    float G = 1.0f / (1.0f + ggxLambda(fz, V, m.ax, m.ay) + ggxLambda(fz, L, m.ax, m.ay));
    float vh = fmaxf(dot3(V, H), 0.0f);
    float s = 1.0f - vh, s2 = s * s, s5 = s2 * s2 * s;
    f3 F = mkf3(m.color.x + (1.0f - m.color.x) * s5, m.color.y + (1.0f - m.color.y) * s5,
                m.color.z + (1.0f - m.color.z) * s5);
    return F * (D * G / (4.0f * nv * nl));
  }
  if (!reflect) {
    // BTDF branch (rough dielectric transmission)
    float rr = frontFace ? m.eta : 1.0f / m.eta;
    float D = ggxD(fz, wm, m.ax, m.ay);
    f3 v2 = dot3(wm, V) < 0.0f ? -V : V;
    f3 l2 = dot3(wm, L) < 0.0f ? -L : L;
    float G = 1.0f / (1.0f + ggxLambda(fz, v2, m.ax, m.ay) + ggxLambda(fz, l2, m.ax, m.ay));
    float F = fresnelDielectric(wm, v2, rr);
    float ci = fabsf(dot3(n, v2)), ct = fabsf(dot3(n, l2));
    float mb = fabsf(dot3(l2, wm)) + fabsf(dot3(v2, wm)) / rr;
    float spec = (1.0f - F) * D * G * fabsf(dot3(wm, l2) * dot3(wm, v2) / (ci * ct * mb * mb));
    return m.color * spec;
  }
  // BRDF branch: GGX conductor (n=2.63,k=9.08; rough dielectric: 1.58, 0.001)
  float nr = 2.63f, kr = 9.08f;
  if (m.type == CPT_M_ROUGH_DIELECTRIC) {
    nr = 1.58f;
    kr = 0.001f;
  }
  f3 H = unit3(V + L);
  float D = ggxD(fz, H, m.ax, m.ay);
  float G = 1.0f / (1.0f + ggxLambda(fz, V, m.ax, m.ay) + ggxLambda(fz, L, m.ax, m.ay));
  float F = fresnelConductor(fmaxf(dot3(n, V), 0.0f), fmaxf(dot3(n, L), 0.0f), nr, kr);
  float den = 4.0f * fmaxf(dot3(n, V), 0.0f) * fmaxf(dot3(n, L), 0.0f) + 0.001f;
  return m.color * (D * G * F / den);
}

// VNDF sample of the GGX normal in frame f (brdf(), Material.cuh:209-222).
__device__ __forceinline__ f3 sampleVndf(const Frame &f, f3 uo, float ax, float ay, float r1, float r2,
                                         f3 &Wo, bool tangentFallbackY) {
  Wo = toLocal(f, uo);
  f3 Wh = unit3(mkf3(ax * Wo.x, ay * Wo.y, Wo.z));
  f3 T1;
  if (Wh.z < 0.99999f) T1 = unit3(cross3(mkf3(0, 0, 1), Wh));
  else T1 = tangentFallbackY ? mkf3(1, 0, 0) : unit3(cross3(mkf3(1, 0, 0), Wh));
  f3 T2 = cross3(Wh, T1);
  float r = sqrtf(r1);
  float th = 2.0f * kPi * r2;
  float s, c;
  __sincosf(th, &s, &c);
  float px = r * c, py = r * s;
  float h = sqrtf(fmaxf(0.0f, 1.0f - px * px));
  float lx = (1.0f + Wh.z) * 0.5f;
  py = (1.0f - lx) * h + lx * py;
  float pz = sqrtf(fmaxf(0.0f, 1.0f - (px * px + py * py)));
  f3 nh = T1 * px + T2 * py + Wh * pz;
  return unit3(mkf3(ax * nh.x, ay * nh.y, fmaxf(1e-6f, nh.z)));
}

// brdf() (Material.cuh:201-235): reflection about a VNDF normal.
__device__ __forceinline__ bool sampleReflect(const ShadeMat &m, f3 uo, f3 n, float r1, float r2, f3 &dir,
                                              float &pdf) {
  Frame f = frameZ(n);
  f3 Wo;
  f3 Wm = sampleVndf(f, uo, m.ax, m.ay, r1, r2, Wo, false);
  f3 Wi = Wm * (2.0f * dot3(Wo, Wm)) - Wo;
  dir = unit3(toWorld(f, Wi));
  if (dot3(dir, n) < 0.0f) {
    pdf = -1.0f;
    return false;
  }
  f3 H = unit3(dir + uo);
  float g1 = 1.0f / (1.0f + ggxLambda(f, uo, m.ax, m.ay));
  pdf = g1 / Wo.z * ggxD(f, H, m.ax, m.ay) * 0.25f;
  return true;
}

// Material::sampleLight (Material.cuh:95-193).  ud = unit ray direction.
__device__ __forceinline__ void bsdfSample(const ShadeMat &m, f3 ud, f3 n, bool frontFace, float r1, float r2,
                                           float r3, f3 &dir, float &pdf, f3 &wm) {
  const f3 uo = -ud;
  if (m.type == CPT_M_SPECULAR_DIELECTRIC) {
    float rr = frontFace ? m.eta : 1.0f / m.eta;
    float R = fresnelDielectric(n, uo, rr);
    if (r3 < R) {
      dir = ud - n * (2.0f * dot3(ud, n));
      pdf = R;
    } else {
      f3 vperp = (ud + n * dot3(uo, n)) * (1.0f / rr);
      f3 vpar = n * (-sqrtf(fabsf(1.0f - dot3(vperp, vperp))));
      dir = unit3(vperp + vpar);
      pdf = 1.0f - R;
    }
    return;
  }
  if (m.type == CPT_M_ROUGH_DIELECTRIC) {
    Frame f = frameX(n);
    f3 Wo;
    f3 Wm = sampleVndf(f, uo, m.ax, m.ay, r1, r2, Wo, true);
    f3 sn = unit3(toWorld(f, Wm));
    wm = sn;
    float rr = frontFace ? m.eta : 1.0f / m.eta;
    float R = fresnelDielectric(sn, ud, rr);
    if (r3 < R) {
      sampleReflect(m, uo, n, r1, r2, dir, pdf);
      pdf *= R;
    } else {
      f3 vperp = (uo + sn * dot3(-uo, sn)) * (1.0f / rr);
      f3 vpar = sn * (-sqrtf(fabsf(1.0f - dot3(vperp, vperp))));
      f3 rd = unit3(vperp + vpar);
      f3 Wi = toLocal(f, rd);
      dir = rd;
      if (dot3(dir, n) > 0.0f) {
        pdf = -1.0f;
        return;
      }
      float mb = dot3(-Wi, Wm) + dot3(Wo, Wm) / rr;
      float dwm = fabsf(dot3(Wi, Wm)) / (mb * mb);
      Frame fz = frameZ(n);
      float g1 = 1.0f / (1.0f + ggxLambda(fz, uo, m.ax, m.ay));
      pdf = g1 / Wo.z * ggxD(fz, sn, m.ax, m.ay) * dot3(Wo, Wm) * dwm * (1.0f - R);
    }
    return;
  }
  if (m.type == CPT_M_LAMBERT) {
    float st = sqrtf(1.0f - r1), ct = sqrtf(r1);
    Frame f = frameX(n);
    float s, c;
    __sincosf(2.0f * kPi * r2, &s, &c);
    dir = f.u * (c * st) + f.v * (s * st) + f.w * ct;
    pdf = ct / kPi;
    return;
  }
  // CPT_M_OPAQUE and CPT_M_GGX
  sampleReflect(m, uo, n, r1, r2, dir, pdf);
}

}  // namespace cpt

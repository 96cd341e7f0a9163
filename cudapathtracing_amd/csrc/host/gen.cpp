// gen.cpp - seeded procedural scenes (no reference counterpart: the
// reference only loads scene_ini files and OBJ meshes).
//
//   icosphere      20*4^level triangles, midpoints re-projected to the sphere,
//                  optional radial displacement r*(1 + amp*n(p)) with 4-octave
//                  value noise n in [-1,1]
//   triangle soup  centroids U[0,1]^3, equilateral with side exp(mu+sigma*z),
//                  z ~ N(0,1), uniform random orientation (Shoemake quaternion)
//   random rays    origins U[lo,hi]^3, directions uniform on S^2
// Every draw comes from a counter-based Philox4x32-10 keyed by (seed, item),
// so results are independent of thread count.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../../include/cpt.h"

namespace {

struct U4 {
  uint32_t x, y, z, w;
};

U4 philox(U4 c, uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
    c = U4{(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// 4 uniforms in [0,1) (53-bit from two words each would be overkill: 32-bit).
inline void uniforms(uint64_t seed, uint64_t item, uint32_t stream, double u[4]) {
  U4 r = philox(U4{(uint32_t)item, (uint32_t)(item >> 32), stream, 0x47454e31u}, (uint32_t)seed,
                (uint32_t)(seed >> 32));
  u[0] = r.x * 2.3283064365386963e-10;
  u[1] = r.y * 2.3283064365386963e-10;
  u[2] = r.z * 2.3283064365386963e-10;
  u[3] = r.w * 2.3283064365386963e-10;
}

template <class F> void parallelFor(int64_t n, F f) {
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if (nt > 16) nt = 16;
  if (n < 65536) nt = 1;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([=]() {
      int64_t b = n * t / nt, e = n * (t + 1) / nt;
      for (int64_t i = b; i < e; i++) f(i);
    });
  for (auto &x : th) x.join();
}

struct V {
  double x, y, z;
};
V add(V a, V b) { return V{a.x + b.x, a.y + b.y, a.z + b.z}; }
V scl(V a, double s) { return V{a.x * s, a.y * s, a.z * s}; }
V nrm(V a) {
  double l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  return V{a.x / l, a.y / l, a.z / l};
}

// Lattice value in [-1,1] from an integer hash of (i,j,k,octave,seed).
double lattice(int64_t i, int64_t j, int64_t k, int oct, uint64_t seed) {
  U4 r = philox(U4{(uint32_t)i, (uint32_t)j, (uint32_t)k, (uint32_t)oct}, (uint32_t)seed,
                (uint32_t)(seed >> 32) ^ 0x4e4f4953u);
  return r.x * (2.0 / 4294967295.0) - 1.0;
}

double valueNoise(V p, uint64_t seed) {
  double sum = 0, amp = 1, norm = 0, f = 4.0;
  for (int o = 0; o < 4; o++) {
    double x = p.x * f, y = p.y * f, z = p.z * f;
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    double tx = x - fx, ty = y - fy, tz = z - fz;
    tx = tx * tx * (3 - 2 * tx);
    ty = ty * ty * (3 - 2 * ty);
    tz = tz * tz * (3 - 2 * tz);
    double c[8];
    for (int q = 0; q < 8; q++) c[q] = lattice(ix + (q & 1), iy + ((q >> 1) & 1), iz + (q >> 2), o, seed);
    double x00 = c[0] + (c[1] - c[0]) * tx, x10 = c[2] + (c[3] - c[2]) * tx;
    double x01 = c[4] + (c[5] - c[4]) * tx, x11 = c[6] + (c[7] - c[6]) * tx;
    double y0 = x00 + (x10 - x00) * ty, y1 = x01 + (x11 - x01) * ty;
    sum += amp * (y0 + (y1 - y0) * tz);
    norm += amp;
    amp *= 0.5;
    f *= 2.0;
  }
  return sum / norm;
}

}  // namespace

extern "C" {

int64_t cpt_gen_icosphere_count(int32_t level) {
  if (level < 0 || level > 12) return -1;
  return (int64_t)20 << (2 * level);
}

int cpt_gen_icosphere(int32_t level, const double c[3], double r, double amp, uint64_t seed,
                      double *out) {
  if (level < 0 || level > 12 || !out || !c) return CPT_ERR_ARG;
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  std::vector<V> v0 = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                       {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  for (auto &p : v0) p = nrm(p);
  const int F[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                        {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                        {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                        {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
  std::vector<V> tri(60);
  for (int f = 0; f < 20; f++)
    for (int k = 0; k < 3; k++) tri[3 * f + k] = v0[F[f][k]];
  for (int l = 0; l < level; l++) {
    const int64_t nt = (int64_t)tri.size() / 3;
    std::vector<V> nx(nt * 12);
    parallelFor(nt, [&](int64_t i) {
      V a = tri[3 * i], b = tri[3 * i + 1], cc = tri[3 * i + 2];
      // midpoint of an edge is computed from the sum, which is symmetric,
      // so triangles sharing an edge get the same vertex (watertight).
      V ab = nrm(add(a, b)), bc = nrm(add(b, cc)), ca = nrm(add(cc, a));
      V *o = &nx[12 * i];
      o[0] = a; o[1] = ab; o[2] = ca;
      o[3] = ab; o[4] = b; o[5] = bc;
      o[6] = ca; o[7] = bc; o[8] = cc;
      o[9] = ab; o[10] = bc; o[11] = ca;
    });
    tri.swap(nx);
  }
  const int64_t nt = (int64_t)tri.size() / 3;
  parallelFor(nt * 3, [&](int64_t i) {
    V p = tri[i];
    double rr = r;
    if (amp != 0.0) rr = r * (1.0 + amp * valueNoise(p, seed));
    out[3 * i + 0] = c[0] + p.x * rr;
    out[3 * i + 1] = c[1] + p.y * rr;
    out[3 * i + 2] = c[2] + p.z * rr;
  });
  return CPT_OK;
}

int cpt_gen_triangle_soup(int64_t n, uint64_t seed, double mu, double sigma, double *out) {
  if (n < 0 || (n > 0 && !out)) return CPT_ERR_ARG;
  const double PI = 3.14159265358979323846;
  parallelFor(n, [&](int64_t i) {
    double a[4], b[4];
    uniforms(seed, (uint64_t)i, 0, a);
    uniforms(seed, (uint64_t)i, 1, b);
    V cen{a[0], a[1], a[2]};
    double u1 = a[3] > 0 ? a[3] : 1e-300;
    double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * PI * b[0]);
    double side = std::exp(mu + sigma * z);
    double R = side / std::sqrt(3.0);
    // Shoemake uniform quaternion
    double s1 = std::sqrt(1.0 - b[1]), s2 = std::sqrt(b[1]);
    double qx = s1 * std::sin(2 * PI * b[2]), qy = s1 * std::cos(2 * PI * b[2]);
    double qz = s2 * std::sin(2 * PI * b[3]), qw = s2 * std::cos(2 * PI * b[3]);
    double m[3][3] = {{1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)},
                      {2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)},
                      {2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)}};
    for (int k = 0; k < 3; k++) {
      double ang = PI / 2 + k * 2 * PI / 3;
      double lx = R * std::cos(ang), ly = R * std::sin(ang);
      for (int c = 0; c < 3; c++) {
        double w = m[c][0] * lx + m[c][1] * ly;
        // round vertices to FP32 so the scene is identical in FP32 and FP64
        out[9 * i + 3 * k + c] = (double)(float)((&cen.x)[c] + w);
      }
    }
  });
  return CPT_OK;
}

int cpt_gen_random_rays(int64_t n, uint64_t seed, double lo, double hi, float tmin, float tmax,
                        float *rays) {
  if (n < 0 || (n > 0 && !rays)) return CPT_ERR_ARG;
  const double PI = 3.14159265358979323846;
  parallelFor(n, [&](int64_t i) {
    double a[4], b[4];
    uniforms(seed, (uint64_t)i, 2, a);
    uniforms(seed, (uint64_t)i, 3, b);
    double z = 1.0 - 2.0 * b[0];
    double rr = std::sqrt(std::fmax(0.0, 1.0 - z * z));
    double ph = 2.0 * PI * b[1];
    float *r = rays + 8 * i;
    r[0] = (float)(lo + (hi - lo) * a[0]);
    r[1] = (float)(lo + (hi - lo) * a[1]);
    r[2] = (float)(lo + (hi - lo) * a[2]);
    r[3] = tmin;
    r[4] = (float)(rr * std::cos(ph));
    r[5] = (float)(rr * std::sin(ph));
    r[6] = (float)z;
    r[7] = tmax;
  });
  return CPT_OK;
}

int cpt_gen_fbm_grid(int32_t n, uint64_t seed, float *density) {
  if (n <= 0 || n > 1024 || !density) return CPT_ERR_ARG;
  const int64_t total = (int64_t)n * n * n;
  parallelFor(total, [&](int64_t i) {
    const int x = (int)(i % n), y = (int)((i / n) % n), z = (int)(i / ((int64_t)n * n));
    V p{(x + 0.5) / n, (y + 0.5) / n, (z + 0.5) / n};
    // fBm in [-1,1] -> density with contrast, clamped to [0,1]
    double v = 0.5 + 1.2 * valueNoise(p, seed);
    density[i] = (float)(v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v));
  });
  return CPT_OK;
}

int cpt_tri_bounds(const double *tri, int64_t n, double mn[3], double mx[3]) {
  if (!mn || !mx || (n > 0 && !tri)) return CPT_ERR_ARG;
  for (int k = 0; k < 3; k++) {
    mn[k] = 999999.0;  // INF of rtmath.cuh:8 (Scene.cuh:60-61 start value)
    mx[k] = -999999.0;
  }
  for (int64_t i = 0; i < 3 * n; i++)
    for (int k = 0; k < 3; k++) {
      double v = tri[3 * i + k];
      mn[k] = v < mn[k] ? v : mn[k];
      mx[k] = v > mx[k] ? v : mx[k];
    }
  return CPT_OK;
}

// Camera::Camera, CudaPathTracing/Camera.cu:5-31 (FP64, same operation order).
int cpt_camera_from_desc(const cpt_camera_desc *d, cpt_camera *o) {
  if (!d || !o || d->width <= 0 || !(d->aspect > 0)) return CPT_ERR_ARG;
  const double PI = 3.14159265358979323846;
  int width = d->width;
  double hh = width / d->aspect;
  int height = hh < 1 ? 1 : (int)hh;
  V from{d->look_from[0], d->look_from[1], d->look_from[2]};
  V at{d->look_at[0], d->look_at[1], d->look_at[2]};
  V off{from.x - at.x, from.y - at.y, from.z - at.z};
  double focal = std::sqrt(off.x * off.x + off.y * off.y + off.z * off.z);
  double theta = d->vfov_deg * PI / 180.0;
  double h = std::tan(theta / 2);
  double vh = 2.0 * h * focal;
  double vw = static_cast<double>(width) / height * vh;
  V w = nrm(off);
  V up{0, 1, 0};
  V u = nrm(V{up.y * w.z - up.z * w.y, up.z * w.x - up.x * w.z, up.x * w.y - up.y * w.x});
  V v{w.y * u.z - w.z * u.y, w.z * u.x - w.x * u.z, w.x * u.y - w.y * u.x};
  V horiz = scl(u, vw), vert = scl(v, vh);
  V dh{horiz.x / width, horiz.y / width, horiz.z / width};
  V dv{vert.x / height, vert.y / height, vert.z / height};
  V ul{from.x - horiz.x / 2 - vert.x / 2 - focal * w.x, from.y - horiz.y / 2 - vert.y / 2 - focal * w.y,
       from.z - horiz.z / 2 - vert.z / 2 - focal * w.z};
  V pc{ul.x + dh.x / 2 + dv.x / 2, ul.y + dh.y / 2 + dv.y / 2, ul.z + dh.z / 2 + dv.z / 2};
  std::memcpy(o->look_from, &from, 24);
  std::memcpy(o->pixel_center, &pc, 24);
  std::memcpy(o->delta_h, &dh, 24);
  std::memcpy(o->delta_v, &dv, 24);
  std::memcpy(o->background, d->background, 24);
  o->width = width;
  o->height = height;
  return CPT_OK;
}

// STATIC display path, Render.cu:232-260: sqrt gamma, uchar(254.99*clamp01).
int cpt_tonemap_rgba8(const float *rgb, int32_t w, int32_t h, uint8_t *out) {
  if (!rgb || !out || w <= 0 || h <= 0) return CPT_ERR_ARG;
  for (int64_t i = 0; i < (int64_t)w * h; i++) {
    for (int c = 0; c < 3; c++) {
      double x = std::sqrt((double)rgb[3 * i + c]);
      x = x > 1.0 ? 1.0 : (x < 0.0 ? 0.0 : x);
      if (!(x == x)) x = 0.0;
      out[4 * i + c] = static_cast<unsigned char>(254.99 * x);
    }
    out[4 * i + 3] = 255;
  }
  return CPT_OK;
}

}  // extern "C"

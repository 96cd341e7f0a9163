// capi.hip - the C ABI (include/cpt.h): context, scene upload, LBVH and
// traversal entry points.  Host code only; kernels live in bvh.hip,
// trace.hip and render.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "cpt_device.hpp"
#include "internal.hpp"

namespace cpt {

int fail(cpt_ctx *ctx, int code, const char *what, const char *detail) {
  if (ctx) {
    ctx->err = what ? what : "";
    if (detail) {
      ctx->err += ": ";
      ctx->err += detail;
    }
  }
  return code;
}

int ensure(cpt_ctx *ctx, DevBuf &b, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (b.bytes >= bytes) return CPT_OK;
  if (b.p) {
    (void)hipFree(b.p);  // replaced by a larger allocation
    b.p = nullptr;
    b.bytes = 0;
  }
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    char msg[96];
    snprintf(msg, sizeof msg, "%zu bytes: %s", bytes, hipGetErrorString(e));
    (void)hipGetLastError();
    return fail(ctx, CPT_ERR_OOM, "hipMalloc", msg);
  }
  b.bytes = bytes;
  return CPT_OK;
}

static void freeBuf(DevBuf &b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

void releaseAll(cpt_ctx *ctx) {
  DevBuf *all[] = {&ctx->center64, &ctx->box32,  &ctx->rec,       &ctx->tri64,      &ctx->sph64,
                   &ctx->lgt64,    &ctx->mats,   &ctx->lightIdx,  &ctx->codesA,     &ctx->codesB,
                   &ctx->valsA,    &ctx->valsB,  &ctx->hist,      &ctx->rowTotal,   &ctx->keys64,
                   &ctx->order,    &ctx->nodes,  &ctx->parentInt, &ctx->parentLeaf, &ctx->flags,
                   &ctx->leaves,   &ctx->rootBox, &ctx->rayBuf,   &ctx->tBuf,       &ctx->idBuf,
                   &ctx->cntA,     &ctx->cntB,    &ctx->nodes4,     &ctx->esc,        &ctx->medDensity,
                   &ctx->rkA,      &ctx->rkB,     &ctx->rvA,        &ctx->rvB,        &ctx->nodes4q,
                   &ctx->nodes8,   &ctx->plC[0],  &ctx->plC[1],     &ctx->plB[0],     &ctx->plB[1],
                   &ctx->plD[0],   &ctx->plD[1],  &ctx->plNN,       &ctx->plSum,      &ctx->nodesP,
                   &ctx->plS[0],   &ctx->plS[1],  &ctx->plCtr,      &ctx->plTmp,      &ctx->plSah,
                   &ctx->recNode,  &ctx->recLev};
  for (DevBuf *b : all) freeBuf(*b);
  renderRelease(ctx);
}

// ---- host FP64 math for the few lights and spheres (reference order) -------
struct hd3 {
  double x, y, z;
};
static hd3 H(double x, double y, double z) { return hd3{x, y, z}; }
static hd3 operator+(hd3 a, hd3 b) { return H(a.x + b.x, a.y + b.y, a.z + b.z); }
static hd3 operator-(hd3 a, hd3 b) { return H(a.x - b.x, a.y - b.y, a.z - b.z); }
static hd3 operator*(hd3 v, double t) { return H(v.x * t, v.y * t, v.z * t); }
static hd3 operator/(hd3 v, double t) { return H(v.x / t, v.y / t, v.z / t); }
static hd3 hCross(hd3 a, hd3 b) {
  return H(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static hd3 hUnit(hd3 v) { return v / std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
static double mn2(double a, double b) { return a < b ? a : b; }
static double mx2(double a, double b) { return a > b ? a : b; }
static float rdn(double x) {
  float f = (float)x;
  if ((double)f > x) f = std::nextafter(f, -INFINITY);
  return f;
}
static float rup(double x) {
  float f = (float)x;
  if ((double)f < x) f = std::nextafter(f, INFINITY);
  return f;
}
static float i2f(int v) {
  float f;
  std::memcpy(&f, &v, 4);
  return f;
}

}  // namespace cpt

using namespace cpt;

extern "C" {

const char *cpt_version(void) { return "cpt-mi355x 0.1 (gfx950, HIP " HIP_VERSION_BUILD_NAME ")"; }

int cpt_create(cpt_ctx **out, int dev) {
  if (!out) return CPT_ERR_ARG;
  *out = nullptr;
  cpt_ctx *ctx = new (std::nothrow) cpt_ctx();
  if (!ctx) return CPT_ERR_OOM;
  ctx->device = dev;
  hipError_t e = hipSetDevice(dev);
  // a blocking stream: ordered with the device's null stream both ways, so
  // device buffers written by null-stream work (PyTorch's default stream
  // has handle 0, which cpt_set_stream maps to this stream) are complete
  // before the library reads them, and the library's outputs before
  // null-stream work reads them
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->ownStream, hipStreamDefault);
  for (int i = 0; i < 18 && e == hipSuccess; i++) e = hipEventCreate(&ctx->ev[i]);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    delete ctx;
    return CPT_ERR_HIP;
  }
  ctx->stream = ctx->ownStream;
  // the abort word: coherent pinned memory, so a host store reaches a
  // running kernel's system-scope loads
  void *ab = nullptr, *abDev = nullptr;
  if (hipHostMalloc(&ab, 64, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer(&abDev, ab, 0) != hipSuccess) {
    (void)hipGetLastError();
    if (ab) (void)hipHostFree(ab);
    for (int i = 0; i < 18; i++)
      if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
    (void)hipStreamDestroy(ctx->ownStream);
    delete ctx;
    return CPT_ERR_HIP;
  }
  ctx->abortHost = (uint32_t *)ab;
  ctx->abortDev = (const uint32_t *)abDev;
  __atomic_store_n(ctx->abortHost, 0u, __ATOMIC_RELEASE);
  if (noAllocScratchInit(ctx) != CPT_OK) {
    cpt_destroy(ctx);
    return CPT_ERR_HIP;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
    ctx->numCU = cus;
  *out = ctx;
  return CPT_OK;
}

void cpt_destroy(cpt_ctx *ctx) {
  if (!ctx) return;
  // release path: errors have nothing to report to
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  // denoisers destroyed after their context only release their handles
  for (cpt_denoiser *dn : ctx->denoisers) cpt::denoiserRelease(dn);
  ctx->denoisers.clear();
  releaseAll(ctx);
  for (int i = 0; i < 18; i++)
    if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
  if (ctx->ownStream) (void)hipStreamDestroy(ctx->ownStream);
  if (ctx->abortHost) (void)hipHostFree(ctx->abortHost);
  delete ctx;
}

int cpt_request_abort(cpt_ctx *ctx) {
  if (!ctx || !ctx->abortHost) return CPT_ERR_ARG;
  __atomic_store_n(ctx->abortHost, 1u, __ATOMIC_RELEASE);
  return CPT_OK;
}

int cpt_clear_abort(cpt_ctx *ctx) {
  if (!ctx || !ctx->abortHost) return CPT_ERR_ARG;
  __atomic_store_n(ctx->abortHost, 0u, __ATOMIC_RELEASE);
  return CPT_OK;
}

const char *cpt_last_error(const cpt_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cpt_set_stream(cpt_ctx *ctx, void *s) {
  if (!ctx) return CPT_ERR_ARG;
  ctx->stream = s ? (hipStream_t)s : ctx->ownStream;
  return CPT_OK;
}

int cpt_synchronize(cpt_ctx *ctx) {
  if (!ctx) return CPT_ERR_ARG;
  CPT_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CPT_OK;
}

int64_t cpt_scene_count(const cpt_ctx *ctx) { return ctx ? ctx->nPrims : -1; }

int cpt_set_option(cpt_ctx *ctx, int32_t option, int32_t value) {
  if (!ctx) return CPT_ERR_ARG;
  if (option == CPT_OPT_TRAVERSAL_WIDTH && (value == 2 || value == 4 || value == 8)) {
    ctx->traversal = value;
    return CPT_OK;
  }
  if (option == CPT_OPT_RAY_SORT_MIN && value >= 0) {
    ctx->raySortMin = value;
    return CPT_OK;
  }
  if (option == CPT_OPT_QUEUE_SORT_MIN && value >= 0) {
    ctx->queueSortMin = value;
    return CPT_OK;
  }
  if (option == CPT_OPT_SHADOW_SORT_MIN && value >= 0) {
    ctx->shadowSortMin = value;
    return CPT_OK;
  }
  if (option == CPT_OPT_CLUSTER && value >= 0 && value <= 2) {
    if (ctx->cluster != value) ctx->recValid = 0;
    ctx->cluster = value;
    return CPT_OK;
  }
  if (option == CPT_OPT_QUANTIZED && (value == 0 || value == 1)) {
    ctx->quantized = value != 0;
    return CPT_OK;
  }
  if (option == CPT_OPT_PERSISTENT_SHADOW && value >= 0) {
    ctx->persistentShadowMin = value;
    return CPT_OK;
  }
  if (option == CPT_OPT_PERSISTENT && (value == 0 || value == 1)) {
    ctx->persistent = value != 0;
    return CPT_OK;
  }
  return fail(ctx, CPT_ERR_ARG, "cpt_set_option", "unknown option or value");
}

int cpt_scene_upload(cpt_ctx *ctx, const cpt_material *materials, int32_t nm, const cpt_light *lights,
                     int32_t nl, const cpt_sphere *spheres, int32_t ns, const double *tri,
                     const int32_t *triMat, int64_t nt) {
  if (!ctx) return CPT_ERR_ARG;
  if (nm < 0 || nl < 0 || ns < 0 || nt < 0 || (nm > 0 && !materials) || (nl > 0 && !lights) ||
      (ns > 0 && !spheres) || (nt > 0 && !tri))
    return fail(ctx, CPT_ERR_ARG, "cpt_scene_upload", "negative count or missing array");
  const int64_t n = (int64_t)nl + ns + nt;
  if (n > (int64_t)0x7ffffffe) return fail(ctx, CPT_ERR_ARG, "cpt_scene_upload", "more than 2^31-2 primitives");
  // the path tracer packs material << 3 | object type into one word
  if ((int64_t)nm + nl + 1 > ((int64_t)1 << 28))
    return fail(ctx, CPT_ERR_ARG, "cpt_scene_upload", "more than 2^28 materials and lights");
  for (int i = 0; i < ns; i++)
    if (spheres[i].material < 0 || spheres[i].material >= (nm > 0 ? nm : 1))
      return fail(ctx, CPT_ERR_ARG, "cpt_scene_upload", "sphere material index out of range");
  if (triMat)
    for (int64_t i = 0; i < nt; i++)
      if (triMat[i] < 0 || triMat[i] >= (nm > 0 ? nm : 1))
        return fail(ctx, CPT_ERR_ARG, "cpt_scene_upload", "triangle material index out of range");
  CPT_TRY_HIP(ctx, hipSetDevice(ctx->device));
  ctx->built = false;
  ctx->hasMedium = false;
  ctx->hasVolumeMat = false;
  for (int i = 0; i < nm; i++)
    if (materials[i].type == CPT_M_VOLUME) ctx->hasVolumeMat = true;
  int rc;
  if ((rc = ensure(ctx, ctx->center64, n * 24)) || (rc = ensure(ctx, ctx->box32, n * 24)) ||
      (rc = ensure(ctx, ctx->rec, n * 48)) || (rc = ensure(ctx, ctx->tri64, nt * 72)) ||
      (rc = ensure(ctx, ctx->sph64, (size_t)ns * 32)) || (rc = ensure(ctx, ctx->lgt64, (size_t)nl * 48)) ||
      (rc = ensure(ctx, ctx->mats, (size_t)(nm + nl + 1) * sizeof(MatDev))) ||
      (rc = ensure(ctx, ctx->lightIdx, (size_t)(nl + 1) * 4)))
    return rc;
  ctx->nPrims = n;
  ctx->nTri = nt;
  ctx->nLights = nl;
  ctx->nSpheres = ns;
  ctx->nMats = nm;
  ctx->hostLights.assign(lights, lights + nl);

  // Lights and spheres: FP64 on the host (Hittable.cuh:359-365, 418-454).
  const int64_t small = (int64_t)nl + ns;
  std::vector<double> cen(small * 3), l64(nl * 6), s64(ns * 4);
  std::vector<float> box(small * 6), rec(small * 12);
  std::vector<int32_t> lidx(nl + 1);
  for (int i = 0; i < nl; i++) {
    const cpt_light &L = lights[i];
    hd3 c = H(L.center[0], L.center[1], L.center[2]);
    hd3 nn = hUnit(H(L.normal[0], L.normal[1], L.normal[2]));
    hd3 N = hUnit(nn);
    hd3 T = std::fabs(N.x) > 0.9 ? hUnit(hCross(H(0, 1, 0), N)) : hUnit(hCross(H(1, 0, 0), N));
    hd3 B = hUnit(hCross(N, T));
    hd3 cr[4] = {c + T * (L.width / 2) + B * (L.height / 2), c + T * (L.width / 2) + B * (-L.height / 2),
                 c + T * (-L.width / 2) + B * (L.height / 2), c + T * (-L.width / 2) + B * (-L.height / 2)};
    hd3 a = cr[0], b = cr[0];
    for (int k = 1; k < 4; k++) {
      a = H(mn2(a.x, cr[k].x), mn2(a.y, cr[k].y), mn2(a.z, cr[k].z));
      b = H(mx2(b.x, cr[k].x), mx2(b.y, cr[k].y), mx2(b.z, cr[k].z));
    }
    const double pad = 0.0001;
    a = H(a.x - pad, a.y - pad, a.z - pad);
    b = H(b.x + pad, b.y + pad, b.z + pad);
    float *bx = &box[6 * i];
    bx[0] = rdn(mn2(a.x, b.x)); bx[1] = rdn(mn2(a.y, b.y)); bx[2] = rdn(mn2(a.z, b.z));
    bx[3] = rup(mx2(a.x, b.x)); bx[4] = rup(mx2(a.y, b.y)); bx[5] = rup(mx2(a.z, b.z));
    cen[3 * i] = c.x; cen[3 * i + 1] = c.y; cen[3 * i + 2] = c.z;
    double *g = &l64[6 * i];
    g[0] = c.x; g[1] = c.y; g[2] = c.z; g[3] = nn.x; g[4] = nn.y; g[5] = nn.z;
    float *r = &rec[12 * i];
    r[0] = (float)c.x; r[1] = (float)c.y; r[2] = (float)c.z; r[3] = i2f(i);
    r[4] = (float)nn.x; r[5] = (float)nn.y; r[6] = (float)nn.z; r[7] = i2f(CPT_OBJ_LIGHT);
    r[8] = (float)L.width; r[9] = (float)L.height; r[10] = L.visible ? 1.f : 0.f; r[11] = i2f(nm + i);
    lidx[i] = i;
  }
  for (int i = 0; i < ns; i++) {
    const cpt_sphere &S = spheres[i];
    const int id = nl + i;
    hd3 c = H(S.center[0], S.center[1], S.center[2]);
    hd3 rr = H(S.radius, S.radius, S.radius);
    hd3 a = c - rr, b = c + rr;
    float *bx = &box[6 * id];
    bx[0] = rdn(mn2(a.x, b.x)); bx[1] = rdn(mn2(a.y, b.y)); bx[2] = rdn(mn2(a.z, b.z));
    bx[3] = rup(mx2(a.x, b.x)); bx[4] = rup(mx2(a.y, b.y)); bx[5] = rup(mx2(a.z, b.z));
    cen[3 * id] = c.x; cen[3 * id + 1] = c.y; cen[3 * id + 2] = c.z;
    double *g = &s64[4 * i];
    g[0] = c.x; g[1] = c.y; g[2] = c.z; g[3] = S.radius;
    float *r = &rec[12 * id];
    r[0] = (float)c.x; r[1] = (float)c.y; r[2] = (float)c.z; r[3] = i2f(id);
    r[4] = (float)S.radius; r[5] = 0.f; r[6] = 0.f; r[7] = i2f(CPT_OBJ_SPHERE);
    r[8] = 0.f; r[9] = 0.f; r[10] = 0.f; r[11] = i2f(S.material);
  }
  // Material table: user materials, then one CPT_M_LIGHT entry per light.
  std::vector<MatDev> md(nm + nl + 1);
  for (int i = 0; i < nm; i++) {
    const cpt_material &m = materials[i];
    md[i] = MatDev{(float)m.color[0], (float)m.color[1], (float)m.color[2], (float)m.alpha_x,
                   (float)m.alpha_y,  (float)m.eta,      m.type,            0};
  }
  for (int i = 0; i < nl; i++)
    md[nm + i] = MatDev{(float)lights[i].color[0], (float)lights[i].color[1], (float)lights[i].color[2],
                        1.f, 1.f, 1.51f, CPT_M_LIGHT, 0};
  hipStream_t s = ctx->stream;
  if (small > 0) {
    CPT_TRY_HIP(ctx, hipMemcpyAsync(ctx->center64.p, cen.data(), small * 24, hipMemcpyHostToDevice, s));
    CPT_TRY_HIP(ctx, hipMemcpyAsync(ctx->box32.p, box.data(), small * 24, hipMemcpyHostToDevice, s));
    CPT_TRY_HIP(ctx, hipMemcpyAsync(ctx->rec.p, rec.data(), small * 48, hipMemcpyHostToDevice, s));
  }
  if (nl) CPT_TRY_HIP(ctx, hipMemcpyAsync(ctx->lgt64.p, l64.data(), nl * 48, hipMemcpyHostToDevice, s));
  if (ns) CPT_TRY_HIP(ctx, hipMemcpyAsync(ctx->sph64.p, s64.data(), ns * 32, hipMemcpyHostToDevice, s));
  CPT_TRY_HIP(ctx, hipMemcpyAsync(ctx->mats.p, md.data(), md.size() * sizeof(MatDev), hipMemcpyHostToDevice, s));
  CPT_TRY_HIP(ctx, hipMemcpyAsync(ctx->lightIdx.p, lidx.data(), lidx.size() * 4, hipMemcpyHostToDevice, s));
  if (nt > 0) {
    CPT_TRY_HIP(ctx, hipMemcpyAsync(ctx->tri64.p, tri, nt * 72, hipMemcpyHostToDevice, s));
    // triangle material ids staged in a buffer of the context (no early
    // return can leak it)
    const int32_t *matDev = nullptr;
    if (triMat) {
      if ((rc = ensure(ctx, ctx->cntB, nt * 4))) return rc;
      CPT_TRY_HIP(ctx, hipMemcpyAsync(ctx->cntB.p, triMat, nt * 4, hipMemcpyHostToDevice, s));
      matDev = ctx->cntB.as<int32_t>();
    }
    rc = packTriangles(ctx, ctx->tri64.as<double>(), matDev, nt, small);
    hipError_t e = hipStreamSynchronize(s);
    if (rc) return rc;
    if (e != hipSuccess) return fail(ctx, CPT_ERR_HIP, "cpt_scene_upload", hipGetErrorString(e));
  } else {
    CPT_TRY_HIP(ctx, hipStreamSynchronize(s));
  }
  return CPT_OK;
}

int cpt_scene_update_triangles(cpt_ctx *ctx, int64_t first, const double *tri, int64_t n) {
  if (!ctx || first < 0 || n < 0 || (n > 0 && !tri) || first + n > ctx->nTri)
    return fail(ctx, CPT_ERR_ARG, "cpt_scene_update_triangles", "range outside the scene's triangles");
  if (n == 0) return CPT_OK;
  CPT_TRY_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  double *dst = ctx->tri64.as<double>() + 9 * first;
  CPT_TRY_HIP(ctx, hipMemcpyAsync(dst, tri, n * 72, hipMemcpyHostToDevice, s));
  // materials are kept: the pack kernel rewrites them from the old records
  // (the id buffer is the context's, so no early return leaks it)
  int rc = ensure(ctx, ctx->cntB, n * 4);
  if (rc) return rc;
  const int64_t base = (int64_t)ctx->nLights + ctx->nSpheres;
  // extract the material ids (word 11 of each 48 B record) on the host side
  std::vector<float> recs(n * 12);
  CPT_TRY_HIP(ctx, hipMemcpyAsync(recs.data(), ctx->rec.as<float>() + 12 * (base + first), n * 48,
                                  hipMemcpyDeviceToHost, s));
  CPT_TRY_HIP(ctx, hipStreamSynchronize(s));
  std::vector<int32_t> mids(n);
  for (int64_t i = 0; i < n; i++) std::memcpy(&mids[i], &recs[12 * i + 11], 4);
  CPT_TRY_HIP(ctx, hipMemcpyAsync(ctx->cntB.p, mids.data(), n * 4, hipMemcpyHostToDevice, s));
  rc = packTriangles(ctx, dst, ctx->cntB.as<int32_t>(), n, base + first);
  hipError_t e = hipStreamSynchronize(s);
  if (rc) return rc;
  if (e != hipSuccess) return fail(ctx, CPT_ERR_HIP, "cpt_scene_update_triangles", hipGetErrorString(e));
  return CPT_OK;
}

int cpt_scene_set_medium(cpt_ctx *ctx, const float *density, int32_t nx, int32_t ny, int32_t nz,
                         const double bmin[3], const double bmax[3], double sigmaMax, const double albedo[3]) {
  if (!ctx) return CPT_ERR_ARG;
  if (!density || nx <= 0 || ny <= 0 || nz <= 0) {
    ctx->hasMedium = false;
    return CPT_OK;
  }
  if (!bmin || !bmax || !albedo || !(sigmaMax > 0) || nx > 4096 || ny > 4096 || nz > 4096)
    return fail(ctx, CPT_ERR_ARG, "cpt_scene_set_medium", "bad grid, box, sigma or albedo");
  for (int k = 0; k < 3; k++)
    if (!(bmax[k] > bmin[k])) return fail(ctx, CPT_ERR_ARG, "cpt_scene_set_medium", "empty box");
  CPT_TRY_HIP(ctx, hipSetDevice(ctx->device));
  const size_t n = (size_t)nx * ny * nz;
  // delta and ratio tracking treat density as a probability of a real collision
  for (size_t i = 0; i < n; i++)
    if (!(density[i] >= 0.0f && density[i] <= 1.0f))
      return fail(ctx, CPT_ERR_ARG, "cpt_scene_set_medium", "density outside [0, 1]");
  int rc = ensure(ctx, ctx->medDensity, n * 4);
  if (rc) return rc;
  CPT_TRY_HIP(ctx, hipMemcpyAsync(ctx->medDensity.p, density, n * 4, hipMemcpyHostToDevice, ctx->stream));
  CPT_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < 3; k++) {
    ctx->medBox[k] = (float)bmin[k];
    ctx->medBox[3 + k] = (float)bmax[k];
    ctx->medAlbedo[k] = (float)albedo[k];
  }
  ctx->medN[0] = nx;
  ctx->medN[1] = ny;
  ctx->medN[2] = nz;
  ctx->medSigma = (float)sigmaMax;
  ctx->hasMedium = true;
  return CPT_OK;
}

int cpt_bvh_refit(cpt_ctx *ctx, cpt_build_stats *st) {
  if (!ctx) return CPT_ERR_ARG;
  CPT_TRY_HIP(ctx, hipSetDevice(ctx->device));
  return refitBvh(ctx, st);
}

int cpt_bvh_cluster_info(const cpt_ctx *ctx, int32_t *in_use, int32_t *depth, double *lbvh_cost,
                         double *cluster_cost) {
  if (!ctx) return CPT_ERR_ARG;
  const int fmt = cpt::recordFormat(ctx);
  if (in_use) *in_use = ctx->cluster != 0 && ctx->clusterValid && ctx->clusterOk && (fmt == 1 || fmt == 2) ? 1 : 0;
  if (depth) *depth = ctx->clusterDepth;
  if (lbvh_cost) *lbvh_cost = ctx->lbvhCost;
  if (cluster_cost) *cluster_cost = ctx->clusterCost;
  return CPT_OK;
}

int cpt_bvh_build(cpt_ctx *ctx, const double bmin[3], const double bmax[3], cpt_build_stats *st) {
  if (!ctx || !bmin || !bmax) return CPT_ERR_ARG;
  CPT_TRY_HIP(ctx, hipSetDevice(ctx->device));
  int rc = buildBvh(ctx, bmin, bmax, st);
  if (rc) return rc;
  return CPT_OK;
}

int cpt_bvh_export(cpt_ctx *ctx, uint64_t *keys, uint32_t *order, int32_t *ca, int32_t *cb,
                   float *boxes, float *rootBox) {
  if (!ctx) return CPT_ERR_ARG;
  if (!ctx->built) return fail(ctx, CPT_ERR_STATE, "cpt_bvh_export", "no BVH built");
  const int64_t n = ctx->nPrims;
  hipStream_t s = ctx->stream;
  CPT_TRY_HIP(ctx, hipStreamSynchronize(s));
  if (n == 0) return CPT_OK;
  if (keys) CPT_TRY_HIP(ctx, hipMemcpy(keys, ctx->keys64.p, n * 8, hipMemcpyDeviceToHost));
  if (order) CPT_TRY_HIP(ctx, hipMemcpy(order, ctx->order.p, n * 4, hipMemcpyDeviceToHost));
  if (rootBox) CPT_TRY_HIP(ctx, hipMemcpy(rootBox, ctx->rootBox.p, 24, hipMemcpyDeviceToHost));
  if (n > 1 && (ca || cb || boxes)) {
    std::vector<float> nodes((size_t)(n - 1) * 16);
    CPT_TRY_HIP(ctx, hipMemcpy(nodes.data(), ctx->nodes.p, (n - 1) * 64, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n - 1; i++) {
      const float *p = &nodes[16 * i];
      int a, b;
      std::memcpy(&a, p + 12, 4);
      std::memcpy(&b, p + 13, 4);
      if (ca) ca[i] = a;
      if (cb) cb[i] = b;
      if (boxes) {
        float *o = boxes + 12 * i;
        o[0] = p[0]; o[1] = p[2]; o[2] = p[8]; o[3] = p[1]; o[4] = p[3]; o[5] = p[9];
        o[6] = p[4]; o[7] = p[6]; o[8] = p[10]; o[9] = p[5]; o[10] = p[7]; o[11] = p[11];
      }
    }
  }
  return CPT_OK;
}

int cpt_bvh_morton(cpt_ctx *ctx, const double bmin[3], const double bmax[3], uint64_t *keysOut,
                   int flags) {
  if (!ctx || !bmin || !bmax || !keysOut) return CPT_ERR_ARG;
  CPT_TRY_HIP(ctx, hipSetDevice(ctx->device));
  const int64_t n = ctx->nPrims;
  if (n == 0) return CPT_OK;
  uint64_t *dst = keysOut;
  if (!(flags & CPT_DEVICE_PTRS)) {
    int rc = ensure(ctx, ctx->keys64, n * 8);
    if (rc) return rc;
    dst = ctx->keys64.as<uint64_t>();
  }
  int rc = launchMorton(ctx, bmin, bmax, dst);
  if (rc) return rc;
  ctx->built = false;  // keys64 scratch reused
  if (!(flags & CPT_DEVICE_PTRS)) {
    CPT_TRY_HIP(ctx, hipMemcpyAsync(keysOut, dst, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    CPT_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return CPT_OK;
}

int cpt_radix_sort_pairs(cpt_ctx *ctx, uint32_t *keys, uint32_t *vals, int64_t n, int32_t bits,
                         int flags) {
  if (!ctx || n < 0 || bits < 1 || bits > 32 || (n > 0 && (!keys || !vals))) return CPT_ERR_ARG;
  if (n > (int64_t)0x7fffffff) return fail(ctx, CPT_ERR_ARG, "cpt_radix_sort_pairs", "n too large");
  if (n == 0) return CPT_OK;
  CPT_TRY_HIP(ctx, hipSetDevice(ctx->device));
  int rc;
  DevBuf kA, vA, kB, vB;
  uint32_t *dk = keys, *dv = vals;
  auto cleanup = [&]() {
    freeBuf(kA); freeBuf(vA); freeBuf(kB); freeBuf(vB);
  };
  if (!(flags & CPT_DEVICE_PTRS)) {
    if ((rc = ensure(ctx, kA, n * 4)) || (rc = ensure(ctx, vA, n * 4))) { cleanup(); return rc; }
    dk = kA.as<uint32_t>();
    dv = vA.as<uint32_t>();
    hipError_t ce = hipMemcpyAsync(dk, keys, n * 4, hipMemcpyHostToDevice, ctx->stream);
    if (ce == hipSuccess) ce = hipMemcpyAsync(dv, vals, n * 4, hipMemcpyHostToDevice, ctx->stream);
    if (ce != hipSuccess) {
      cleanup();
      return fail(ctx, CPT_ERR_HIP, "cpt_radix_sort_pairs", hipGetErrorString(ce));
    }
  }
  if ((rc = ensure(ctx, kB, n * 4)) || (rc = ensure(ctx, vB, n * 4))) { cleanup(); return rc; }
  rc = radixSortPairs(ctx, dk, dv, kB.as<uint32_t>(), vB.as<uint32_t>(), n, bits, false);
  hipError_t e = hipSuccess;
  if (!rc && !(flags & CPT_DEVICE_PTRS)) {
    e = hipMemcpyAsync(keys, dk, n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(vals, dv, n * 4, hipMemcpyDeviceToHost, ctx->stream);
  }
  const hipError_t es = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = es;
  cleanup();
  if (rc) return rc;
  if (e != hipSuccess) return fail(ctx, CPT_ERR_HIP, "cpt_radix_sort_pairs", hipGetErrorString(e));
  return CPT_OK;
}

static bool aborted(const cpt_ctx *ctx) {
  return ctx->abortHost && __atomic_load_n(ctx->abortHost, __ATOMIC_ACQUIRE) != 0u;
}

static int traceCommon(cpt_ctx *ctx, const float *rays, int64_t n, float *tOut, int32_t *idOut,
                       int32_t *nodesOut, int32_t *primsOut, int flags) {
  if (!ctx || n < 0 || (n > 0 && !rays)) return CPT_ERR_ARG;
  if (!ctx->built) return fail(ctx, CPT_ERR_STATE, "cpt_trace_closest", "no BVH built");
  if (n == 0) return CPT_OK;
  if (aborted(ctx)) return fail(ctx, CPT_ERR_ABORTED, "cpt_trace_closest", "aborted by cpt_request_abort");
  CPT_TRY_HIP(ctx, hipSetDevice(ctx->device));
  int rc;
  if (flags & CPT_DEVICE_PTRS) {
    rc = traceDevice(ctx, rays, n, tOut, idOut, nodesOut, primsOut);
    if (rc) return rc;
  } else {
    const bool count = nodesOut || primsOut;
    if ((rc = ensure(ctx, ctx->rayBuf, n * 32)) || (rc = ensure(ctx, ctx->tBuf, n * 4)) ||
        (rc = ensure(ctx, ctx->idBuf, n * 4)))
      return rc;
    CPT_TRY_HIP(ctx, hipMemcpyAsync(ctx->rayBuf.p, rays, n * 32, hipMemcpyHostToDevice, ctx->stream));
    if (count) {
      rc = traceDevice(ctx, ctx->rayBuf.as<float>(), n, nullptr, nullptr, ctx->tBuf.as<int32_t>(),
                       ctx->idBuf.as<int32_t>());
      if (rc) return rc;
      if (nodesOut) CPT_TRY_HIP(ctx, hipMemcpyAsync(nodesOut, ctx->tBuf.p, n * 4, hipMemcpyDeviceToHost, ctx->stream));
      if (primsOut) CPT_TRY_HIP(ctx, hipMemcpyAsync(primsOut, ctx->idBuf.p, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    } else {
      rc = traceDevice(ctx, ctx->rayBuf.as<float>(), n, ctx->tBuf.as<float>(), ctx->idBuf.as<int32_t>(),
                       nullptr, nullptr);
      if (rc) return rc;
      if (tOut) CPT_TRY_HIP(ctx, hipMemcpyAsync(tOut, ctx->tBuf.p, n * 4, hipMemcpyDeviceToHost, ctx->stream));
      if (idOut) CPT_TRY_HIP(ctx, hipMemcpyAsync(idOut, ctx->idBuf.p, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    }
    CPT_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  // the persistent walks stop taking rays once the abort word is set: the
  // outputs are then incomplete
  if (aborted(ctx)) return fail(ctx, CPT_ERR_ABORTED, "cpt_trace_closest", "aborted by cpt_request_abort");
  return CPT_OK;
}

int cpt_trace_closest(cpt_ctx *ctx, const float *rays, int64_t n, float *tOut, int32_t *idOut,
                      int flags) {
  return traceCommon(ctx, rays, n, tOut, idOut, nullptr, nullptr, flags);
}

int cpt_trace_count(cpt_ctx *ctx, const float *rays, int64_t n, int32_t *nodesOut, int32_t *primsOut,
                    int flags) {
  if (!nodesOut && !primsOut) return CPT_ERR_ARG;
  if ((flags & CPT_DEVICE_PTRS) && (!nodesOut || !primsOut)) return CPT_ERR_ARG;
  return traceCommon(ctx, rays, n, nullptr, nullptr, nodesOut, primsOut, flags);
}

float cpt_last_trace_ms(const cpt_ctx *ctx) {
  if (!ctx) return -1.f;
  float ms = -1.f;
  if (hipEventSynchronize(ctx->ev[9]) != hipSuccess) return -1.f;
  if (hipEventElapsedTime(&ms, ctx->ev[8], ctx->ev[9]) != hipSuccess) return -1.f;
  return ms;
}

int cpt_last_trace_stats(const cpt_ctx *ctx, float *totalMs, float *sortMs, float *kernelMs) {
  if (!ctx) return CPT_ERR_ARG;
  if (hipEventSynchronize(ctx->ev[9]) != hipSuccess) return CPT_ERR_HIP;
  float tot = 0.f, srt = 0.f, ker = 0.f;
  if (hipEventElapsedTime(&tot, ctx->ev[8], ctx->ev[9]) != hipSuccess) return CPT_ERR_HIP;
  if (ctx->lastTraceSorted) {
    if (hipEventElapsedTime(&srt, ctx->ev[8], ctx->ev[5]) != hipSuccess) return CPT_ERR_HIP;
    if (hipEventElapsedTime(&ker, ctx->ev[5], ctx->ev[9]) != hipSuccess) return CPT_ERR_HIP;
  } else {
    ker = tot;
  }
  if (totalMs) *totalMs = tot;
  if (sortMs) *sortMs = srt;
  if (kernelMs) *kernelMs = ker;
  return CPT_OK;
}

}  // extern "C"

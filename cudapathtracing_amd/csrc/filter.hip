// filter.hip - the reference's non-learned denoiser: the separable joint
// bilateral filter gaussianSeparate (CudaPathTracing/Render.cu:353-398) run
// horizontally then vertically as main.cu:100-104 does (render -> H pass ->
// copy -> V pass).  Per 1D pass and pixel, over offsets i in [-16, 16] with
// clamp-to-edge:
//   w_i = exp(-i^2/(2 sG^2) - |c_n - c|^2/(2 sR^2) - |n - n_n|^2/(2 sN^2)
//             - (d - d_n)^2/(2 sD^2))
//   out = sum w_i c_n / sum w_i
// where c is the pass input at the centre pixel, c_n at the neighbour and
// (n, d) the normal and depth of the G-buffer.  FP32 here; FP64 in the
// reference.  The G-buffer is the sample mean of the first-hit normal and
// depth (the reference keeps the last sample's).
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace cpt {

constexpr int kRadius = 16;  // KERNEL_RADIUS, Render.cuh:7

__global__ void k_bilateral_pass(const float *__restrict__ in, const float *__restrict__ aov, float *__restrict__ out,
                                 int w, int h, float sG, float sR, float sN, float sD, int horizontal) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= w || y >= h) return;
  const int64_t o = (int64_t)y * w + x;
  const float cr = in[3 * o], cg = in[3 * o + 1], cb = in[3 * o + 2];
  const float nx = aov[7 * o + 3], ny = aov[7 * o + 4], nz = aov[7 * o + 5], dd = aov[7 * o + 6];
  const float iG = 1.0f / (2.0f * sG * sG), iR = 1.0f / (2.0f * sR * sR), iN = 1.0f / (2.0f * sN * sN),
              iD = 1.0f / (2.0f * sD * sD);
  float sr = 0.f, sg = 0.f, sb = 0.f, sw = 0.f;
  for (int i = -kRadius; i <= kRadius; i++) {
    int px = x, py = y;
    if (horizontal) px = min(max(x + i, 0), w - 1);
    else py = min(max(y + i, 0), h - 1);
    const int64_t q = (int64_t)py * w + px;
    const float pr = in[3 * q], pg = in[3 * q + 1], pb = in[3 * q + 2];
    const float a = pr - cr, b = pg - cg, c = pb - cb;
    const float e = aov[7 * q + 3] - nx, f = aov[7 * q + 4] - ny, g = aov[7 * q + 5] - nz;
    const float dz = dd - aov[7 * q + 6];
    const float wt = __expf(-(float)(i * i) * iG - (a * a + b * b + c * c) * iR - (e * e + f * f + g * g) * iN -
                            dz * dz * iD);
    sw += wt;
    sr += pr * wt;
    sg += pg * wt;
    sb += pb * wt;
  }
  const float inv = 1.0f / sw;
  out[3 * o] = sr * inv;
  out[3 * o + 1] = sg * inv;
  out[3 * o + 2] = sb * inv;
}

// Temporal accumulation for a moving camera: addPrevious (Render.cu:400-421)
// with Camera::getPrePositionInImg (Camera.cuh:104-132), in FP64.  Each
// pixel's first-hit position is projected into the previous camera; the
// previous image is read there (or at the same pixel when the position is a
// miss or falls outside) and blended as out = cur * alpha + prev * (1 - alpha),
// alpha = 0.4 while the camera moves, else 0.1.
struct PrevCam {
  double lf[3], la[3];
  double vfov;  // degrees
  int w, h;
};

__device__ __forceinline__ double dot3d(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__global__ void k_temporal(const float *__restrict__ rgb, const float *__restrict__ prev,
                           const float *__restrict__ pos, const float *__restrict__ aov, cpt_camera cam, PrevCam pc,
                           double alpha, float *__restrict__ out) {
#pragma clang fp contract(off)
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= pc.w || y >= pc.h) return;
  const int64_t off = (int64_t)y * pc.w + x;
  const double kInf = 999999.0;  // rtmath.cuh:8
  double p[3];
  bool valid;
  if (pos) {
    for (int k = 0; k < 3; k++) p[k] = pos[3 * off + k];
    valid = !(p[0] + p[1] > kInf);
  } else {
    // first-hit position from the mean first-hit distance along the pixel-centre ray
    const double dpt = aov[7 * off + 6];
    double dir[3];
    for (int k = 0; k < 3; k++) dir[k] = (cam.pixel_center[k] + x * cam.delta_h[k] + y * cam.delta_v[k]) - cam.look_from[k];
    const double len = sqrt(dot3d(dir, dir));
    for (int k = 0; k < 3; k++) p[k] = cam.look_from[k] + dir[k] / len * dpt;
    valid = dpt > 0.0;
  }
  int pi = -1, pj = -1;
  if (valid) {
    double dir[3], fwd[3];
    for (int k = 0; k < 3; k++) {
      dir[k] = p[k] - pc.lf[k];
      fwd[k] = pc.la[k] - pc.lf[k];
    }
    const double ld = sqrt(dot3d(dir, dir)), lf = sqrt(dot3d(fwd, fwd));
    for (int k = 0; k < 3; k++) {
      dir[k] /= ld;
      fwd[k] /= lf;
    }
    // right = Unit(Cross(forward, vUp)), vUp = (0, 1, 0); up = Cross(right, forward)
    double right[3] = {fwd[1] * 0.0 - fwd[2] * 1.0, fwd[2] * 0.0 - fwd[0] * 0.0, fwd[0] * 1.0 - fwd[1] * 0.0};
    const double lr = sqrt(dot3d(right, right));
    for (int k = 0; k < 3; k++) right[k] /= lr;
    const double up[3] = {right[1] * fwd[2] - right[2] * fwd[1], right[2] * fwd[0] - right[0] * fwd[2],
                          right[0] * fwd[1] - right[1] * fwd[0]};
    const double focal = lf;
    const double hh = tan(pc.vfov * 3.14159265358979323846 / 180.0 / 2.0);
    const double vh = 2.0 * hh * focal;
    const double vw = (double)pc.w / pc.h * vh;
    const double px = dot3d(dir, right), py = dot3d(dir, up), pz = dot3d(dir, fwd);
    const double u = 0.5 + (px / pz) * (focal / vw);
    const double v = 0.5 + (py / pz) * (focal / vh);
    if (!(u < 0.0 || v < 0.0 || u > 1.0 || v > 1.0)) {
      pi = (int)(u * pc.w);
      pj = (int)(v * pc.h);
    }
  }
  // (i, j) = (width, .) or (., height) at u or v = 1 reads past the image in
  // the reference; it is clamped to the last pixel here
  int64_t src = off;
  if (pi >= 0 && pj >= 0) src = (int64_t)min(pj, pc.h - 1) * pc.w + min(pi, pc.w - 1);
  for (int k = 0; k < 3; k++)
    out[3 * off + k] = (float)((double)rgb[3 * off + k] * alpha + (double)prev[3 * src + k] * (1.0 - alpha));
}

}  // namespace cpt

using namespace cpt;

extern "C" int cpt_bilateral(cpt_ctx *ctx, const float *rgb, const float *aov, float *out, int32_t w, int32_t h,
                             float sigmaG, float sigmaR, float sigmaN, float sigmaD, int flags) {
  if (!ctx || !rgb || !aov || !out || w <= 0 || h <= 0 || !(sigmaG > 0) || !(sigmaR > 0) || !(sigmaN > 0) ||
      !(sigmaD > 0))
    return CPT_ERR_ARG;
  CPT_TRY_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const size_t n = (size_t)w * h;
  DevBuf tmp;
  int rc = ensure(ctx, tmp, n * 4 * ((flags & CPT_DEVICE_PTRS) ? 3 : 16));
  if (rc) return rc;
  const float *dIn = rgb, *dAov = aov;
  float *dMid = tmp.as<float>(), *dOut = out;
  if (!(flags & CPT_DEVICE_PTRS)) {
    float *base = tmp.as<float>();
    dIn = base + 3 * n;
    dAov = base + 6 * n;
    dOut = base + 13 * n;
    hipError_t ce = hipMemcpyAsync((void *)dIn, rgb, n * 12, hipMemcpyHostToDevice, s);
    if (ce == hipSuccess) ce = hipMemcpyAsync((void *)dAov, aov, n * 28, hipMemcpyHostToDevice, s);
    if (ce != hipSuccess) {
      (void)hipFree(tmp.p);
      return fail(ctx, CPT_ERR_HIP, "cpt_bilateral", hipGetErrorString(ce));
    }
  }
  dim3 bs(16, 16), grid((w + 15) / 16, (h + 15) / 16);
  hipLaunchKernelGGL(k_bilateral_pass, grid, bs, 0, s, dIn, dAov, dMid, w, h, sigmaG, sigmaR, sigmaN, sigmaD, 1);
  hipLaunchKernelGGL(k_bilateral_pass, grid, bs, 0, s, (const float *)dMid, dAov, dOut, w, h, sigmaG, sigmaR, sigmaN,
                     sigmaD, 0);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && !(flags & CPT_DEVICE_PTRS)) e = hipMemcpyAsync(out, dOut, n * 12, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(tmp.p);
  if (e != hipSuccess) return fail(ctx, CPT_ERR_HIP, "cpt_bilateral", hipGetErrorString(e));
  return CPT_OK;
}

extern "C" int cpt_temporal_blend(cpt_ctx *ctx, const float *rgb, const float *prevRgb, const float *position,
                                  const float *aov, const cpt_camera *cam, const cpt_camera_desc *prevDesc,
                                  int32_t moving, float *out, int flags) {
  if (!ctx || !rgb || !prevRgb || !out || !cam || !prevDesc || (!position && !aov)) return CPT_ERR_ARG;
  const int w = cam->width, h = cam->height;
  if (w <= 0 || h <= 0 || !(prevDesc->vfov_deg > 0.0)) return CPT_ERR_ARG;
  CPT_TRY_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const size_t n = (size_t)w * h;
  DevBuf tmp;
  const float *dRgb = rgb, *dPrev = prevRgb, *dPos = position, *dAov = aov;
  float *dOut = out;
  if (!(flags & CPT_DEVICE_PTRS)) {
    int rc = ensure(ctx, tmp, n * 4 * (3 + 3 + 3 + (position ? 3 : 7)));
    if (rc) return rc;
    float *b = tmp.as<float>();
    dRgb = b;
    dPrev = b + 3 * n;
    dOut = b + 6 * n;
    float *dIn = b + 9 * n;
    hipError_t ce = hipMemcpyAsync((void *)dRgb, rgb, n * 12, hipMemcpyHostToDevice, s);
    if (ce == hipSuccess) ce = hipMemcpyAsync((void *)dPrev, prevRgb, n * 12, hipMemcpyHostToDevice, s);
    if (ce == hipSuccess) ce = hipMemcpyAsync(dIn, position ? (const void *)position : (const void *)aov,
                                              n * (position ? 12 : 28), hipMemcpyHostToDevice, s);
    if (ce != hipSuccess) {
      (void)hipFree(tmp.p);
      return fail(ctx, CPT_ERR_HIP, "cpt_temporal_blend", hipGetErrorString(ce));
    }
    if (position) dPos = dIn;
    else dAov = dIn;
  }
  PrevCam pc;
  for (int k = 0; k < 3; k++) {
    pc.lf[k] = prevDesc->look_from[k];
    pc.la[k] = prevDesc->look_at[k];
  }
  pc.vfov = prevDesc->vfov_deg;
  pc.w = w;
  pc.h = h;
  dim3 bs(16, 16), grid((w + 15) / 16, (h + 15) / 16);
  hipLaunchKernelGGL(k_temporal, grid, bs, 0, s, dRgb, dPrev, dPos, dAov, *cam, pc, moving ? 0.4 : 0.1, dOut);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && !(flags & CPT_DEVICE_PTRS)) e = hipMemcpyAsync(out, dOut, n * 12, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (tmp.p) (void)hipFree(tmp.p);
  if (e != hipSuccess) return fail(ctx, CPT_ERR_HIP, "cpt_temporal_blend", hipGetErrorString(e));
  return CPT_OK;
}

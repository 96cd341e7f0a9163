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
    hipMemcpyAsync((void *)dIn, rgb, n * 12, hipMemcpyHostToDevice, s);
    hipMemcpyAsync((void *)dAov, aov, n * 28, hipMemcpyHostToDevice, s);
  }
  dim3 bs(16, 16), grid((w + 15) / 16, (h + 15) / 16);
  hipLaunchKernelGGL(k_bilateral_pass, grid, bs, 0, s, dIn, dAov, dMid, w, h, sigmaG, sigmaR, sigmaN, sigmaD, 1);
  hipLaunchKernelGGL(k_bilateral_pass, grid, bs, 0, s, (const float *)dMid, dAov, dOut, w, h, sigmaG, sigmaR, sigmaN,
                     sigmaD, 0);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && !(flags & CPT_DEVICE_PTRS)) e = hipMemcpyAsync(out, dOut, n * 12, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  hipFree(tmp.p);
  if (e != hipSuccess) return fail(ctx, CPT_ERR_HIP, "cpt_bilateral", hipGetErrorString(e));
  return CPT_OK;
}

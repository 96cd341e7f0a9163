// denoise.hip - the learned denoiser: a 5-level convolutional U-Net
// (channels 32, 64, 64, 96, 128; 3x3 convolutions, ReLU, 2x2 average
// pooling, nearest upsampling with skip concatenation) that maps a
// low-spp radiance image plus albedo, normal and depth to a denoised image.
// Weights are BF16 (FP32 master copy and Adam moments), activations BF16,
// accumulation FP32 on MFMA.  The network predicts log(1 + radiance) as a
// residual on the noisy input's log(1 + radiance) and is trained with an L1
// loss in that space.
//
// No reference counterpart: the reference's only filter is the joint
// bilateral + temporal blend of Render.cu:353-460, which this replaces as
// the learned method.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <algorithm>
#include <vector>

#include "cpt_device.hpp"
#include "denoise_kernels.hpp"
#include "internal.hpp"

namespace cpt {

constexpr int kLevels = 5;
constexpr int kChan[kLevels] = {32, 64, 64, 96, 128};
constexpr int kInChan = 32;   // 10 used: log radiance (3), albedo (3), normal (3), log depth
constexpr int kOutChan = 32;  // 3 used
constexpr int kPredChan = 4;  // fp32 channels kept of the output layer (3 used)
constexpr int kLayers = 19;

struct LayerDesc {
  int cin, cout, relu;
};

static void layerTable(LayerDesc L[kLayers]) {
  int k = 0;
  int prev = kInChan;
  for (int l = 0; l < kLevels; l++) {
    L[k++] = {prev, kChan[l], 1};
    L[k++] = {kChan[l], kChan[l], 1};
    prev = kChan[l];
  }
  for (int l = kLevels - 2; l >= 0; l--) {
    L[k++] = {prev + kChan[l], kChan[l], 1};
    L[k++] = {kChan[l], kChan[l], 1};
    prev = kChan[l];
  }
  L[k++] = {kChan[0], kOutChan, 0};
}

// ---- elementwise kernels ----------------------------------------------------

__global__ void k_prep_input(const float *__restrict__ in10, int h, int w, int H, int W, int64_t n,
                             uint16_t *__restrict__ X) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // padded pixel index
  if (i >= n) return;
  const int64_t per = (int64_t)H * W;
  const int b = (int)(i / per);
  const int rem = (int)(i % per);
  const int y = rem / W, x = rem % W;
  float v[kInChan];
#pragma unroll
  for (int c = 0; c < kInChan; c++) v[c] = 0.f;
  if (y < h && x < w) {
    const float *s = in10 + (((int64_t)b * h + y) * w + x) * 10;
#pragma unroll
    for (int c = 0; c < 3; c++) v[c] = log1pf(fmaxf(s[c], 0.f));
#pragma unroll
    for (int c = 3; c < 9; c++) v[c] = s[c];
    v[9] = log1pf(fmaxf(s[9], 0.f));
  }
  uint16_t *d = X + i * kInChan;
#pragma unroll
  for (int c = 0; c < kInChan; c += 8) {
    uint4 pk;
    uint16_t *ph = (uint16_t *)&pk;
#pragma unroll
    for (int j = 0; j < 8; j++) ph[j] = f2bf(v[c + j]);
    *(uint4 *)(d + c) = pk;
  }
}

// The elementwise kernels below move 8 channels (16 B) per thread; every
// channel count is a multiple of 8 and n counts 8-channel groups.
__device__ __forceinline__ void unpack8(const uint4 v, float f[8]) {
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int j = 0; j < 4; j++) {
    f[2 * j] = __uint_as_float(w[j] << 16);
    f[2 * j + 1] = __uint_as_float(w[j] & 0xffff0000u);
  }
}
__device__ __forceinline__ uint4 pack8(const float f[8]) {
  uint32_t w[4];
#pragma unroll
  for (int j = 0; j < 4; j++) w[j] = (uint32_t)f2bf(f[2 * j]) | ((uint32_t)f2bf(f[2 * j + 1]) << 16);
  return make_uint4(w[0], w[1], w[2], w[3]);
}

// 2x2 average pooling; n = B*(H/2)*(W/2)*C/8
__global__ void k_pool2(const uint16_t *__restrict__ X, int H, int W, int C, int64_t n, uint16_t *__restrict__ Y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int C8 = C / 8;
  const int c = (int)(i % C8) * 8;
  int64_t p = i / C8;
  const int w2 = W / 2, h2 = H / 2;
  const int x = (int)(p % w2);
  p /= w2;
  const int y = (int)(p % h2);
  const int64_t b = p / h2;
  const int64_t base = ((b * H + 2 * y) * W + 2 * x) * C + c;
  float s[8], t[8];
  unpack8(*(const uint4 *)(X + base), s);
  const int64_t off[3] = {C, (int64_t)W * C, (int64_t)W * C + C};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    unpack8(*(const uint4 *)(X + base + off[k]), t);
#pragma unroll
    for (int j = 0; j < 8; j++) s[j] += t[j];
  }
#pragma unroll
  for (int j = 0; j < 8; j++) s[j] *= 0.25f;
  *(uint4 *)(Y + i * 8) = pack8(s);
}

// dX[2y+i][2x+j] (+)= dY[y][x] / 4; n = B*H*W*C/8 (full resolution); then,
// with maskX, the ReLU gradient mask of maskX (the forward activation of dX)
__global__ void k_pool2_bwd(const uint16_t *__restrict__ dY, int H, int W, int C, int64_t n, uint16_t *__restrict__ dX,
                            int accumulate, const uint16_t *__restrict__ maskX) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int C8 = C / 8;
  const int c = (int)(i % C8) * 8;
  int64_t p = i / C8;
  const int x = (int)(p % W);
  p /= W;
  const int y = (int)(p % H);
  const int64_t b = p / H;
  float g[8];
  unpack8(*(const uint4 *)(dY + ((b * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c), g);
  if (accumulate) {
    float d[8];
    unpack8(*(const uint4 *)(dX + i * 8), d);
#pragma unroll
    for (int j = 0; j < 8; j++) g[j] = d[j] + 0.25f * g[j];
  } else {
#pragma unroll
    for (int j = 0; j < 8; j++) g[j] *= 0.25f;
  }
  uint4 r = pack8(g);
  if (maskX) {
    const uint4 mv = *(const uint4 *)(maskX + i * 8);
    r.x &= reluKeep(mv.x);
    r.y &= reluKeep(mv.y);
    r.z &= reluKeep(mv.z);
    r.w &= reluKeep(mv.w);
  }
  *(uint4 *)(dX + i * 8) = r;
}

// dLow[y][x] = sum over the 2x2 block of dOut[..., 0:C1], ReLU-masked by
// maskLow (the forward activation of dLow); dSkip = dOut[..., C1:]
// n = nLow * (C1 + 4*C2) / 8
__global__ void k_up_concat_bwd(const uint16_t *__restrict__ dOut, int C1, int C2, int H, int W, int64_t nLow,
                                uint16_t *__restrict__ dLow, uint16_t *__restrict__ dSkip,
                                const uint16_t *__restrict__ maskLow) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int C = C1 + C2;
  const int per = (C1 + 4 * C2) / 8;
  if (i >= nLow * per) return;
  const int64_t lp = i / per;  // low-res pixel
  const int k = (int)(i % per) * 8;
  const int w2 = W / 2, h2 = H / 2;
  const int x = (int)(lp % w2);
  const int64_t q = lp / w2;
  const int y = (int)(q % h2);
  const int64_t b = q / h2;
  if (k < C1) {
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, t[8];
#pragma unroll
    for (int dy = 0; dy < 2; dy++)
#pragma unroll
      for (int dx = 0; dx < 2; dx++) {
        unpack8(*(const uint4 *)(dOut + ((b * H + 2 * y + dy) * W + 2 * x + dx) * C + k), t);
#pragma unroll
        for (int j = 0; j < 8; j++) s[j] += t[j];
      }
    uint4 r = pack8(s);
    const uint4 mv = *(const uint4 *)(maskLow + lp * C1 + k);
    r.x &= reluKeep(mv.x);
    r.y &= reluKeep(mv.y);
    r.z &= reluKeep(mv.z);
    r.w &= reluKeep(mv.w);
    *(uint4 *)(dLow + lp * C1 + k) = r;
  } else {
    const int kk = k - C1;
    const int sub = kk / C2, c = kk % C2;
    const int64_t fp = (b * H + 2 * y + (sub >> 1)) * W + 2 * x + (sub & 1);
    *(uint4 *)(dSkip + fp * C2 + c) = *(const uint4 *)(dOut + fp * C + C1 + c);
  }
}

// L1 loss on channels 0..2 of the prediction (pred, ldp channels, + the
// residual log(1 + max(noisy radiance, 0)) from the input in10) against
// log(1+target); dPred (bf16, kOutChan channels) =
// sign(prediction - t) / count.  The loss is deterministic: each block writes
// its partial sum to part[block]; the last block to finish (counted in
// *counter, zero on entry) adds the partials in a fixed order into *loss.
__global__ void k_l1(const float *__restrict__ pred, int ldp, const float *__restrict__ in10,
                     const float *__restrict__ target, int h, int w, int H, int W, int64_t n, float invCount,
                     uint16_t *__restrict__ dPred, float *__restrict__ part) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // padded pixel
  float l = 0.f;
  if (i < n) {
    const int64_t per = (int64_t)H * W;
    const int b = (int)(i / per), rem = (int)(i % per), y = rem / W, x = rem % W;
    uint16_t g[kOutChan];
#pragma unroll
    for (int c = 0; c < kOutChan; c++) g[c] = 0;
    if (y < h && x < w) {
      const int64_t u = ((int64_t)b * h + y) * w + x;  // unpadded pixel
      const float *t = target + u * 3;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const float d = (pred[i * ldp + c] + log1pf(fmaxf(in10[u * 10 + c], 0.f))) - log1pf(fmaxf(t[c], 0.f));
        l += fabsf(d);
        g[c] = f2bf(d > 0.f ? invCount : (d < 0.f ? -invCount : 0.f));
      }
    }
#pragma unroll
    for (int c = 0; c < kOutChan; c += 8) *(uint4 *)(dPred + i * kOutChan + c) = *(uint4 *)(g + c);
  }
  // block reduction of the loss; k_l1_sum adds the blocks' partials (a
  // last-block reduction in this kernel needs a device-scope fence per
  // block, which writes back the XCD's L2 each time: 0.67 ms a step at
  // batch 8 of 512x512)
  __shared__ float red[256];
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] * invCount;
}

// One workgroup: the loss = sum of k_l1's block partials, each thread over a
// fixed stride of them, then a fixed tree (the same order every run).
__global__ void k_l1_sum(const float *__restrict__ part, int n, float *__restrict__ loss) {
  __shared__ float red[256];
  float s = 0.f;
  for (int k = threadIdx.x; k < n; k += blockDim.x) s += part[k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = red[0];
}

// Adam on FP32 master weights; writes the BF16 weights (OHWI) and, for the
// backward-data pass, the flipped transpose Wt[ci][2-ky][2-kx][co].
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float c1, float c2,
                       int update) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !update) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_pack_weights(const float *__restrict__ p, int cin, int cout, uint16_t *__restrict__ w,
                               uint16_t *__restrict__ wt) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // over cout*9*cin
  const int64_t n = (int64_t)cout * 9 * cin;
  if (i >= n) return;
  const int ci = (int)(i % cin);
  const int t = (int)((i / cin) % 9);
  const int co = (int)(i / ((int64_t)cin * 9));
  const uint16_t h = f2bf(p[i]);
  w[i] = h;
  const int ky = t / 3, kx = t % 3;
  wt[(((int64_t)ci * 3 + (2 - ky)) * 3 + (2 - kx)) * cout + co] = h;
}

// rgb = max(expm1(pred + log(1 + max(noisy, 0))), 0) over the unpadded
// pixels (noisy: the radiance channels of the input in10)
__global__ void k_output(const float *__restrict__ pred, int ldp, const float *__restrict__ in10, int h, int w,
                         int H, int W, int64_t n, float *__restrict__ rgb) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // over B*h*w
  if (i >= n) return;
  const int64_t per = (int64_t)h * w;
  const int b = (int)(i / per), rem = (int)(i % per), y = rem / w, x = rem % w;
  const int64_t p = ((int64_t)b * H + y) * W + x;
#pragma unroll
  for (int c = 0; c < 3; c++)
    rgb[i * 3 + c] = fmaxf(expm1f(pred[p * ldp + c] + log1pf(fmaxf(in10[i * 10 + c], 0.f))), 0.f);
}

// Every layer's weights in one launch: thread i packs entry i of the
// concatenated bf16 weights (layer l owns [wbOff[l], wbOff[l+1])).
struct PackTable {
  int64_t wOff[kLayers], wbOff[kLayers + 1];
  int cin[kLayers], cout[kLayers];
};

__global__ void k_pack_all(const float *__restrict__ params, PackTable T, uint16_t *__restrict__ wbf,
                           uint16_t *__restrict__ wtbf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T.wbOff[kLayers]) return;
  int l = 0;
  while (i >= T.wbOff[l + 1]) l++;
  const int64_t j = i - T.wbOff[l];
  const int cin = T.cin[l], cout = T.cout[l];
  const int ci = (int)(j % cin);
  const int t = (int)((j / cin) % 9);
  const int co = (int)(j / ((int64_t)cin * 9));
  const uint16_t h = f2bf(params[T.wOff[l] + j]);
  wbf[i] = h;
  const int ky = t / 3, kx = t % 3;
  wtbf[T.wbOff[l] + (((int64_t)ci * 3 + (2 - ky)) * 3 + (2 - kx)) * cout + co] = h;
}

static unsigned nb256(int64_t n) { return (unsigned)((n + 255) / 256); }

// ---- host-side network ------------------------------------------------------

struct Tensor {
  DevBuf buf;
  int c = 0, h = 0, w = 0;  // per image
  uint16_t *p() const { return buf.as<uint16_t>(); }
};

}  // namespace cpt

struct cpt_denoiser {
  cpt_ctx *ctx = nullptr;
  cpt::LayerDesc L[cpt::kLayers];
  int64_t wOff[cpt::kLayers], bOff[cpt::kLayers], nParams = 0;
  int64_t wbOff[cpt::kLayers];  // bf16 weight offsets
  cpt::DevBuf params, grads, m, v, wbf, wtbf;
  int maxB = 0, maxH = 0, maxW = 0;
  int B = 0, H = 0, W = 0;  // current padded shape
  // activations (and their gradients) per image batch
  cpt::Tensor x0, e[cpt::kLevels][2], pool[cpt::kLevels - 1], d[cpt::kLevels - 1][2];
  cpt::Tensor ge[cpt::kLevels][2], gpool[cpt::kLevels - 1], gcat[cpt::kLevels - 1], gd[cpt::kLevels - 1][2], gx0;
  cpt::DevBuf pred, dPred, loss, lossPart, io;
  cpt::DevBuf part;   // fp32 partial sums of convolutions split over Cin
  cpt::DevBuf wpart;  // every layer's weight-gradient slices (reduced after the backward pass)
  int64_t wpOff[cpt::kLayers] = {};  // float offset of layer k's slices in wpart
  int wpSlices[cpt::kLayers] = {};
  int64_t step = 0;
  float lastFwdMs = 0.f, lastStepMs = 0.f;
  double flopsFwd = 0;
};

namespace cpt {

void denoiserRelease(cpt_denoiser *dn) {
  cpt_ctx *ctx = dn->ctx;
  if (!ctx) return;
  // release path: errors have nothing to report to
  (void)hipStreamSynchronize(ctx->stream);
  DevBuf *bufs[] = {&dn->params, &dn->grads, &dn->m, &dn->v, &dn->wbf, &dn->wtbf, &dn->pred, &dn->dPred,
                    &dn->loss, &dn->lossPart, &dn->io, &dn->x0.buf, &dn->gx0.buf, &dn->part, &dn->wpart};
  for (DevBuf *b : bufs)
    if (b->p) (void)hipFree(b->p);
  for (int l = 0; l < kLevels; l++)
    for (int k = 0; k < 2; k++) {
      if (dn->e[l][k].buf.p) (void)hipFree(dn->e[l][k].buf.p);
      if (dn->ge[l][k].buf.p) (void)hipFree(dn->ge[l][k].buf.p);
    }
  for (int l = 0; l < kLevels - 1; l++) {
    Tensor *ts[] = {&dn->pool[l], &dn->gpool[l], &dn->gcat[l], &dn->d[l][0], &dn->d[l][1], &dn->gd[l][0], &dn->gd[l][1]};
    for (Tensor *t : ts)
      if (t->buf.p) (void)hipFree(t->buf.p);
  }
  dn->ctx = nullptr;
}

}  // namespace cpt

namespace cpt {

static int ensureT(cpt_ctx *ctx, Tensor &t, int B, int h, int w, int c) {
  t.c = c;
  t.h = h;
  t.w = w;
  return ensure(ctx, t.buf, (size_t)B * h * w * c * 2);
}

static int convBwdWSlices(int Cin, int Cout, int B, int H, int W, int *stepsPerSplit);
static int64_t fwdSplitFloats(int Cin, int Cout, int B, int H, int W);

static int allocActs(cpt_denoiser *dn, int B, int H, int W) {
  cpt_ctx *ctx = dn->ctx;
  int rc;
  if ((rc = ensureT(ctx, dn->x0, B, H, W, kInChan)) || (rc = ensureT(ctx, dn->gx0, 1, 1, 1, 8))) return rc;
  for (int l = 0; l < kLevels; l++) {
    const int h = H >> l, w = W >> l;
    for (int k = 0; k < 2; k++) {
      if ((rc = ensureT(ctx, dn->e[l][k], B, h, w, kChan[l])) || (rc = ensureT(ctx, dn->ge[l][k], B, h, w, kChan[l])))
        return rc;
    }
    if (l < kLevels - 1) {
      if ((rc = ensureT(ctx, dn->pool[l], B, h / 2, w / 2, kChan[l])) ||
          (rc = ensureT(ctx, dn->gpool[l], B, h / 2, w / 2, kChan[l])))
        return rc;
      const int below = l == kLevels - 2 ? kChan[kLevels - 1] : kChan[l + 1];
      if ((rc = ensureT(ctx, dn->gcat[l], B, h, w, below + kChan[l])))
        return rc;
      for (int k = 0; k < 2; k++)
        if ((rc = ensureT(ctx, dn->d[l][k], B, h, w, kChan[l])) || (rc = ensureT(ctx, dn->gd[l][k], B, h, w, kChan[l])))
          return rc;
    }
  }
  const int64_t np = (int64_t)B * H * W;
  if ((rc = ensure(ctx, dn->pred, np * kPredChan * 4)) ||
      (rc = ensure(ctx, dn->dPred, np * kOutChan * 2)) || (rc = ensure(ctx, dn->loss, 64)) ||
      (rc = ensure(ctx, dn->lossPart, (size_t)nb256(np) * 4)))
    return rc;
  // weight-gradient slices of every layer, one region each (reduced by one
  // launch after the backward pass); split partials of the forward and
  // backward-data convolutions of small layers (the largest layer's)
  int64_t partFloats = 0, wpFloats = 0;
  for (int k = 0; k < kLayers; k++) {
    int lvl = k < 2 * kLevels ? k / 2 : (k < kLayers - 1 ? kLevels - 2 - (k - 2 * kLevels) / 2 : 0);
    const int64_t nW = (int64_t)dn->L[k].cout * 9 * dn->L[k].cin;
    dn->wpSlices[k] = convBwdWSlices(dn->L[k].cin, dn->L[k].cout, B, H >> lvl, W >> lvl, nullptr);
    dn->wpOff[k] = wpFloats;
    wpFloats += (int64_t)dn->wpSlices[k] * (nW + dn->L[k].cout);
    partFloats = std::max(partFloats, fwdSplitFloats(dn->L[k].cin, dn->L[k].cout, B, H >> lvl, W >> lvl));
    partFloats = std::max(partFloats, fwdSplitFloats(dn->L[k].cout, dn->L[k].cin, B, H >> lvl, W >> lvl));
  }
  if ((rc = ensure(ctx, dn->part, std::max<int64_t>(partFloats, 1) * 4)) || (rc = ensure(ctx, dn->wpart, wpFloats * 4)))
    return rc;
  return CPT_OK;
}

// out[p][co] (bf16, Cout channels) = act(sum over the kSplit slices of
// part[s][p][co] + bias[co]), slices added in order; with maskY the ReLU
// gradient mask of maskY.  n = pixels * Cout / 8.
__global__ void k_split_sum(const float *__restrict__ part, int kSplit, int64_t sliceStride,
                            const float *__restrict__ bias, int Cout, int relu, const uint16_t *__restrict__ maskY,
                            int64_t n, uint16_t *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = (int)((i * 8) % Cout);
  float v[8];
  const float4 a = *(const float4 *)(part + i * 8), b = *(const float4 *)(part + i * 8 + 4);
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  for (int s = 1; s < kSplit; s++) {
    const float4 c0 = *(const float4 *)(part + s * sliceStride + i * 8),
                 c1 = *(const float4 *)(part + s * sliceStride + i * 8 + 4);
    v[0] += c0.x; v[1] += c0.y; v[2] += c0.z; v[3] += c0.w; v[4] += c1.x; v[5] += c1.y; v[6] += c1.z; v[7] += c1.w;
  }
#pragma unroll
  for (int j = 0; j < 8; j++) {
    if (bias) v[j] += bias[c + j];
    if (relu) v[j] = fmaxf(v[j], 0.f);
  }
  uint4 r = pack8(v);
  if (maskY) {
    const uint4 mv = *(const uint4 *)(maskY + i * 8);
    r.x &= reluKeep(mv.x);
    r.y &= reluKeep(mv.y);
    r.z &= reluKeep(mv.z);
    r.w &= reluKeep(mv.w);
  }
  *(uint4 *)(out + i * 8) = r;
}

// Tiling of one 3x3 'same' convolution (k_conv_fwd): 64 output channels per
// wave where Cout allows, else 32.  The smallest layers (under 64 tiles,
// several Cin chunks: the deepest level) are split over Cin into up to as
// many slices as fill 512 workgroups (fp32 partial sums, then k_split_sum):
// there a workgroup's chain of dependent chunk steps costs more than the
// extra launch (about 4.5 us), elsewhere it does not.
struct FwdPlan {
  int nt, colTiles, rowTiles, coTiles, split, tpw;
};
static FwdPlan fwdPlan(int Cin, int Cout, int B, int H, int W) {
  FwdPlan P;
  P.colTiles = (W + kTileW - 1) / kTileW;
  P.rowTiles = (H + kTileH - 1) / kTileH;
  P.nt = Cout % 64 == 0 ? 64 : 32;
  P.coTiles = (Cout + P.nt - 1) / P.nt;
  const int64_t tiles = (int64_t)P.colTiles * P.rowTiles * B * P.coTiles;
  const int chunks = Cin / 32;
  P.split = 1;
  if (chunks > 1 && tiles < 64) {
    int ks = (int)std::min<int64_t>(chunks, (512 + tiles - 1) / tiles);
    const int per = (chunks + ks - 1) / ks;
    P.split = (chunks + per - 1) / per;  // every slice non-empty
  }
  // stacked tiles per workgroup: as few workgroups as still fill the chip
  // twice (about 512, one per CU at a time), at most 8 tiles each
  P.tpw = P.split > 1 ? 1 : (int)std::max<int64_t>(1, std::min<int64_t>(8, (tiles + 511) / 512));
  return P;
}
// fp32 floats of partial sums a split convolution needs (0 if unsplit)
static int64_t fwdSplitFloats(int Cin, int Cout, int B, int H, int W) {
  const FwdPlan P = fwdPlan(Cin, Cout, B, H, W);
  return P.split > 1 ? (int64_t)P.split * B * H * W * Cout : 0;
}

// One 3x3 'same' convolution.  splitPart: scratch for a split over Cin (at
// least fwdSplitFloats floats); without it, or with an fp32 output, the
// convolution is not split.
// poolOut (bf16 output): also its 2x2 average pooling, (H/2) x (W/2) x ldo.
static void launchConvFwd(hipStream_t s, const uint16_t *X, int Cin, const uint16_t *Xlow, int Clow, const uint16_t *w,
                          const float *bias, int Cout, int B, int H, int W, uint16_t *outB, float *outF, int ldo,
                          int mode, const uint16_t *maskY = nullptr, float *splitPart = nullptr,
                          uint16_t *poolOut = nullptr) {
  FwdPlan P = fwdPlan(Cin, Cout, B, H, W);
  if (!splitPart || outF) {
    P.split = 1;
    const int64_t tiles = (int64_t)P.colTiles * P.rowTiles * B * P.coTiles;
    P.tpw = (int)std::max<int64_t>(1, std::min<int64_t>(8, (tiles + 511) / 512));
  }
  const dim3 grid(P.colTiles, (P.rowTiles + P.tpw - 1) / P.tpw, B * P.coTiles * P.split);
  const int64_t sliceStride = (int64_t)B * H * W * Cout;
  float *oF = outF;
  uint16_t *oB = outB;
  const float *bs = bias;
  int md = mode;
  const uint16_t *mk = maskY;
  if (P.split > 1) {  // partial sums, no bias or activation: k_split_sum adds them
    oF = splitPart;
    oB = nullptr;
    bs = nullptr;
    md = 0;
    mk = nullptr;
    ldo = Cout;
  }
  // one 32-channel chunk per slice: the weights are copied into LDS once
  const int chunks = Cin / 32, perSlice = (chunks + P.split - 1) / P.split;
  uint16_t *pk = P.split > 1 ? nullptr : poolOut;  // a split convolution pools after k_split_sum
#define CPT_CONV_FWD(QB, F32, ONE)                                                                             \
  hipLaunchKernelGGL((k_conv_fwd<QB, F32, ONE>), grid, dim3(kFwdThreads), 0, s, X, Cin, Xlow, Clow, w, bs, Cout, H, \
                     W, oB, oF, ldo, md, mk, P.split, sliceStride, P.tpw, F32 ? nullptr : pk)
#define CPT_CONV_FWD_QB(QB)                        \
  if (perSlice == 1) {                             \
    if (oF) CPT_CONV_FWD(QB, true, true);          \
    else CPT_CONV_FWD(QB, false, true);            \
  } else {                                         \
    if (oF) CPT_CONV_FWD(QB, true, false);         \
    else CPT_CONV_FWD(QB, false, false);           \
  }
  if (P.nt == 64) {
    CPT_CONV_FWD_QB(4)
  } else {
    CPT_CONV_FWD_QB(2)
  }
#undef CPT_CONV_FWD_QB
#undef CPT_CONV_FWD
  if (P.split > 1) {
    const int64_t n = sliceStride / 8;
    hipLaunchKernelGGL(k_split_sum, dim3(nb256(n)), dim3(256), 0, s, (const float *)splitPart, P.split, sliceStride,
                       bias, Cout, mode & 1, maskY, n, outB);
    if (poolOut) {
      const int64_t np = (int64_t)B * (H / 2) * (W / 2) * Cout / 8;
      hipLaunchKernelGGL(k_pool2, dim3(nb256(np)), dim3(256), 0, s, outB, H, W, Cout, np, poolOut);
    }
  }
}

// in: the layer's input, or with low the skip half of the decoder's
// concatenation [up2(low) | in], read in place.
static int convFwd(cpt_denoiser *dn, int layer, const Tensor &in, const Tensor *low, int B, float *outF, uint16_t *outB,
                   int mode, uint16_t *poolOut = nullptr) {
  const LayerDesc &Ld = dn->L[layer];
  hipStream_t s = dn->ctx->stream;
  const uint16_t *w = dn->wbf.as<uint16_t>() + dn->wbOff[layer];
  const float *bias = dn->params.as<float>() + dn->bOff[layer];
  const int H = in.h, W = in.w;
  launchConvFwd(s, in.p(), Ld.cin, low ? low->p() : nullptr, low ? low->c : 0, w, bias, Ld.cout, B, H, W, outB, outF,
                outF ? kPredChan : Ld.cout, mode, nullptr, dn->part.as<float>(), poolOut);
  dn->flopsFwd += 2.0 * B * H * W * (double)Ld.cin * Ld.cout * 9;
  return hipGetLastError() == hipSuccess ? CPT_OK : CPT_ERR_HIP;
}

// Weight gradient of one conv (k_conv_bwd_w, split over pixels, then
// k_reduce_splits into dW).  part: scratch of at least
// convBwdWSlices(...) * Cout * 9 * Cin floats.
static int bwdWShape(int Cin, int Cout, int &ks, int &wci, int &wco) {
  if (Cin == 32 && Cout == 32) { ks = 4; wci = 1; wco = 1; }
  else if (Cin == 32) { ks = 2; wci = 1; wco = 2; }
  else if (Cout == 32) { ks = 2; wci = 2; wco = 1; }
  else { ks = 1; wci = 2; wco = 2; }
  return ks;
}

static int convBwdWSlices(int Cin, int Cout, int B, int H, int W, int *stepsPerSplit) {
  int ks, wci, wco;
  bwdWShape(Cin, Cout, ks, wci, wco);
  const int tiles = ((Cin + 32 * wci - 1) / (32 * wci)) * ((Cout + 32 * wco - 1) / (32 * wco));
  const int64_t steps = (int64_t)B * H * ((W + 31) / 32);
  // about 512 workgroups, two per CU, as many as the registers allow (a
  // K-step is latency-bound, so as many as possible in parallel; each slice
  // costs a write and a read of the layer's weights), each group of waves
  // taking at least 4 K-steps
  int64_t s = std::max<int64_t>(1, (512 + tiles - 1) / tiles);
  s = std::min<int64_t>(s, std::max<int64_t>(1, steps / (4 * ks)));
  const int64_t per = (steps + s - 1) / s;
  if (stepsPerSplit) *stepsPerSplit = (int)per;
  return (int)((steps + per - 1) / per);
}

// dW then db (contiguous: Cout*9*Cin + Cout floats) when withBias, else dW;
// with dW null only the slices are written (reduced later, k_reduce_layers)
static void launchConvBwdW(hipStream_t s, const uint16_t *dY, int Cout, const uint16_t *X, int Cin, const uint16_t *Xlow,
                           int Clow, int B, int H, int W, float *part, float *dW, bool withBias) {
  int ks, wci, wco, per = 1;
  bwdWShape(Cin, Cout, ks, wci, wco);
  const int slices = convBwdWSlices(Cin, Cout, B, H, W, &per);
  const dim3 grid((Cin + 32 * wci - 1) / (32 * wci), (Cout + 32 * wco - 1) / (32 * wco), slices);
  if (ks == 4)
    hipLaunchKernelGGL((k_conv_bwd_w<4, 1, 1>), grid, dim3(kConvThreads), 0, s, dY, Cout, X, Cin, Xlow, Clow, B, H, W, per,
                       part);
  else if (wco == 2 && wci == 1)
    hipLaunchKernelGGL((k_conv_bwd_w<2, 1, 2>), grid, dim3(kConvThreads), 0, s, dY, Cout, X, Cin, Xlow, Clow, B, H, W, per,
                       part);
  else if (ks == 2)
    hipLaunchKernelGGL((k_conv_bwd_w<2, 2, 1>), grid, dim3(kConvThreads), 0, s, dY, Cout, X, Cin, Xlow, Clow, B, H, W, per,
                       part);
  else
    hipLaunchKernelGGL((k_conv_bwd_w<1, 2, 2>), grid, dim3(kConvThreads), 0, s, dY, Cout, X, Cin, Xlow, Clow, B, H, W, per,
                       part);
  if (!dW) return;
  const int64_t nW = (int64_t)Cout * 9 * Cin, nOut = withBias ? nW + Cout : nW;
  hipLaunchKernelGGL(k_reduce_splits, dim3((unsigned)((nOut + 255) / 256)), dim3(256), 0, s, (const float *)part,
                     slices, nW + Cout, nOut, dW);
}

// Backward of one conv: dOut (bf16, already ReLU-masked) -> dW, db, dIn;
// dIn leaves masked by inAct (the forward activation of the layer's input,
// the output of a ReLU) when inAct is given.
static int convBwd(cpt_denoiser *dn, int layer, const Tensor &in, const Tensor *low, const uint16_t *dOut, int B,
                   uint16_t *dIn, const uint16_t *inAct) {
  const LayerDesc &Ld = dn->L[layer];
  hipStream_t s = dn->ctx->stream;
  const int H = in.h, W = in.w;
  // slices only: every layer's are reduced into the gradients after the pass
  launchConvBwdW(s, dOut, Ld.cout, in.p(), Ld.cin, low ? low->p() : nullptr, low ? low->c : 0, B, H, W,
                 dn->wpart.as<float>() + dn->wpOff[layer], nullptr, true);
  if (dIn) {
    // backward data: forward kernel on dOut with the flipped transposed weights
    const uint16_t *wt = dn->wtbf.as<uint16_t>() + dn->wbOff[layer];
    const int cout = Ld.cin;  // gradient channels
    launchConvFwd(s, dOut, Ld.cout, nullptr, 0, wt, nullptr, cout, B, H, W, dIn, nullptr, cout, 0, inAct,
                  dn->part.as<float>());
  }
  return hipGetLastError() == hipSuccess ? CPT_OK : CPT_ERR_HIP;
}

static int forward(cpt_denoiser *dn, int B) {
  int rc = 0;
  int layer = 0;
  const Tensor *cur = &dn->x0;
  for (int l = 0; l < kLevels; l++) {
    if (l > 0) cur = &dn->pool[l - 1];  // pooled by the previous level's second convolution
    if ((rc = convFwd(dn, layer++, *cur, nullptr, B, nullptr, dn->e[l][0].p(), 1))) return rc;
    if ((rc = convFwd(dn, layer++, dn->e[l][0], nullptr, B, nullptr, dn->e[l][1].p(), 1,
                      l < kLevels - 1 ? dn->pool[l].p() : nullptr)))
      return rc;
    cur = &dn->e[l][1];
  }
  for (int l = kLevels - 2; l >= 0; l--) {
    // [up2(cur) | e[l][1]], read in place by the conv
    if ((rc = convFwd(dn, layer++, dn->e[l][1], cur, B, nullptr, dn->d[l][0].p(), 1))) return rc;
    if ((rc = convFwd(dn, layer++, dn->d[l][0], nullptr, B, nullptr, dn->d[l][1].p(), 1))) return rc;
    cur = &dn->d[l][1];
  }
  // the output layer: conv + bias only; the residual log(1 + noisy) is added
  // by its consumers (k_output, k_l1)
  return convFwd(dn, layer, *cur, nullptr, B, dn->pred.as<float>(), nullptr, 0);
}

// Every ReLU gradient mask is applied by the kernel that produces the
// gradient: the backward-data conv, the concat split or the pooling backward.
static int backward(cpt_denoiser *dn, int B) {
  int rc;
  hipStream_t s = dn->ctx->stream;
  int layer = kLayers - 1;
  // output conv: input d[0][1], gradient dPred (already d loss / d conv output)
  if ((rc = convBwd(dn, layer--, dn->d[0][1], nullptr, dn->dPred.as<uint16_t>(), B, dn->gd[0][1].p(), dn->d[0][1].p())))
    return rc;
  for (int l = 0; l <= kLevels - 2; l++) {
    // decoder level l: d[l][1] = conv(d[l][0]), d[l][0] = conv(cat[l]); gd[l][1]
    // arrives masked
    if ((rc = convBwd(dn, layer--, dn->d[l][0], nullptr, dn->gd[l][1].p(), B, dn->gd[l][0].p(), dn->d[l][0].p())))
      return rc;
    const Tensor &lowIn = l == kLevels - 2 ? dn->e[kLevels - 1][1] : dn->d[l + 1][1];  // the concat's upsampled half
    if ((rc = convBwd(dn, layer--, dn->e[l][1], &lowIn, dn->gd[l][0].p(), B, dn->gcat[l].p(), nullptr))) return rc;
    // split the concat gradient: upsampled part -> gradient of the level below
    // (masked by that level's output), skip part -> ge[l][1] (masked once the
    // pooling gradient is added)
    Tensor &low = l == kLevels - 2 ? dn->ge[kLevels - 1][1] : dn->gd[l + 1][1];
    const Tensor &lowAct = l == kLevels - 2 ? dn->e[kLevels - 1][1] : dn->d[l + 1][1];
    const int C1 = low.c, C2 = kChan[l];
    const int64_t nLow = (int64_t)B * low.h * low.w;
    hipLaunchKernelGGL(k_up_concat_bwd, dim3(nb256(nLow * (C1 + 4 * C2) / 8)), dim3(256), 0, s, dn->gcat[l].p(), C1, C2,
                       dn->gcat[l].h, dn->gcat[l].w, nLow, low.p(), dn->ge[l][1].p(), lowAct.p());
  }
  // encoder, deepest first; layer index now points at enc level 4 conv 1;
  // ge[l][1] arrives masked
  for (int l = kLevels - 1; l >= 0; l--) {
    if ((rc = convBwd(dn, layer--, dn->e[l][0], nullptr, dn->ge[l][1].p(), B, dn->ge[l][0].p(), dn->e[l][0].p())))
      return rc;
    const Tensor &in = l == 0 ? dn->x0 : dn->pool[l - 1];
    uint16_t *gin = l == 0 ? nullptr : dn->gpool[l - 1].p();
    if ((rc = convBwd(dn, layer--, in, nullptr, dn->ge[l][0].p(), B, gin, nullptr))) return rc;
    if (l > 0) {
      // average-pool backward adds into the skip gradient of level l-1, then
      // masks it by that level's output
      Tensor &g = dn->ge[l - 1][1];
      const int64_t n = (int64_t)B * g.h * g.w * g.c / 8;
      hipLaunchKernelGGL(k_pool2_bwd, dim3(nb256(n)), dim3(256), 0, s, dn->gpool[l - 1].p(), g.h, g.w, g.c, n, g.p(), 1,
                         dn->e[l - 1][1].p());
    }
  }
  return hipGetLastError() == hipSuccess ? CPT_OK : CPT_ERR_HIP;
}

// The weight gradients of every layer: grads[gOff[l] + i] = sum over layer
// l's slices (in slice order) of part[pOff[l] + s * nS[l] + i], i < nS[l]
// (weights then biases).
struct ReduceTable {
  int64_t pOff[kLayers], gOff[kLayers], nS[kLayers], end[kLayers + 1];
  int slices[kLayers];
};

__global__ void k_reduce_layers(const float *__restrict__ part, ReduceTable T, float *__restrict__ grads) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T.end[kLayers]) return;
  int l = 0;
  while (i >= T.end[l + 1]) l++;
  const int64_t j = i - T.end[l], nS = T.nS[l];
  const float *p = part + T.pOff[l] + j;
  float s = 0.f;
  const int ns = T.slices[l];
#pragma unroll 8
  for (int k = 0; k < ns; k++) s += p[(int64_t)k * nS];
  grads[T.gOff[l] + j] = s;
}

static void reduceGrads(cpt_denoiser *dn) {
  ReduceTable T;
  int64_t n = 0;
  for (int l = 0; l < kLayers; l++) {
    T.pOff[l] = dn->wpOff[l];
    T.gOff[l] = dn->wOff[l];
    T.nS[l] = (int64_t)dn->L[l].cout * 9 * dn->L[l].cin + dn->L[l].cout;
    T.slices[l] = dn->wpSlices[l];
    T.end[l] = n;
    n += T.nS[l];
  }
  T.end[kLayers] = n;
  hipLaunchKernelGGL(k_reduce_layers, dim3(nb256(n)), dim3(256), 0, dn->ctx->stream, dn->wpart.as<float>(), T,
                     dn->grads.as<float>());
}

static int packAll(cpt_denoiser *dn) {
  PackTable T;
  for (int l = 0; l < kLayers; l++) {
    T.wOff[l] = dn->wOff[l];
    T.wbOff[l] = dn->wbOff[l];
    T.cin[l] = dn->L[l].cin;
    T.cout[l] = dn->L[l].cout;
  }
  T.wbOff[kLayers] = dn->wbOff[kLayers - 1] + (int64_t)dn->L[kLayers - 1].cout * 9 * dn->L[kLayers - 1].cin;
  hipLaunchKernelGGL(k_pack_all, dim3(nb256(T.wbOff[kLayers])), dim3(256), 0, dn->ctx->stream, dn->params.as<float>(), T,
                     dn->wbf.as<uint16_t>(), dn->wtbf.as<uint16_t>());
  return hipGetLastError() == hipSuccess ? CPT_OK : CPT_ERR_HIP;
}

static int prepare(cpt_denoiser *dn, const float *in10Dev, int b, int h, int w) {
  const int H = (h + 15) / 16 * 16, W = (w + 15) / 16 * 16;
  if (b > dn->maxB || H > dn->maxH || W > dn->maxW) return fail(dn->ctx, CPT_ERR_ARG, "denoiser", "batch or image larger than created");
  int rc = allocActs(dn, b, H, W);
  if (rc) return rc;
  dn->B = b;
  dn->H = H;
  dn->W = W;
  const int64_t n = (int64_t)b * H * W;
  hipLaunchKernelGGL(k_prep_input, dim3(nb256(n)), dim3(256), 0, dn->ctx->stream, in10Dev, h, w, H, W, n, dn->x0.p());
  return hipGetLastError() == hipSuccess ? CPT_OK : CPT_ERR_HIP;
}

// Stages host inputs in the denoiser's io buffer; returns device pointers.
static int stage(cpt_denoiser *dn, const float *a, size_t na, const float *b, size_t nb, int flags, const float **da,
                 const float **db, float **out, size_t nout) {
  if (flags & CPT_DEVICE_PTRS) {
    *da = a;
    if (db) *db = b;
    return CPT_OK;
  }
  int rc = ensure(dn->ctx, dn->io, (na + nb + nout) * 4);
  if (rc) return rc;
  float *base = dn->io.as<float>();
  CPT_TRY_HIP(dn->ctx, hipMemcpyAsync(base, a, na * 4, hipMemcpyHostToDevice, dn->ctx->stream));
  *da = base;
  if (b && db) {
    CPT_TRY_HIP(dn->ctx, hipMemcpyAsync(base + na, b, nb * 4, hipMemcpyHostToDevice, dn->ctx->stream));
    *db = base + na;
  }
  if (out) *out = base + na + nb;
  return CPT_OK;
}

}  // namespace cpt

using namespace cpt;

extern "C" {

int cpt_denoiser_create(cpt_ctx *ctx, int32_t maxB, int32_t maxH, int32_t maxW, uint64_t seed, cpt_denoiser **out) {
  if (!ctx || !out || maxB <= 0 || maxH <= 0 || maxW <= 0) return CPT_ERR_ARG;
  // the convolution kernels address one image of a tensor with 32-bit
  // offsets: the widest per-image tensor is 32 bf16 channels at full size
  if ((int64_t)maxH * maxW * 64 >= (1ll << 31))
    return fail(ctx, CPT_ERR_ARG, "cpt_denoiser_create", "image too large (H*W*64 B must stay below 2^31)");
  *out = nullptr;
  CPT_TRY_HIP(ctx, hipSetDevice(ctx->device));
  cpt_denoiser *dn = new cpt_denoiser();
  dn->ctx = ctx;
  dn->maxB = maxB;
  dn->maxH = (maxH + 15) / 16 * 16;
  dn->maxW = (maxW + 15) / 16 * 16;
  layerTable(dn->L);
  int64_t off = 0, boff = 0;
  for (int l = 0; l < kLayers; l++) {
    dn->wOff[l] = off;
    off += (int64_t)dn->L[l].cout * 9 * dn->L[l].cin;
    dn->bOff[l] = off;
    off += dn->L[l].cout;
    dn->wbOff[l] = boff;
    boff += (int64_t)dn->L[l].cout * 9 * dn->L[l].cin;
  }
  dn->nParams = off;
  int rc;
  if ((rc = ensure(ctx, dn->params, off * 4)) || (rc = ensure(ctx, dn->grads, off * 4)) || (rc = ensure(ctx, dn->m, off * 4)) ||
      (rc = ensure(ctx, dn->v, off * 4)) || (rc = ensure(ctx, dn->wbf, boff * 2)) || (rc = ensure(ctx, dn->wtbf, boff * 2))) {
    cpt_denoiser_destroy(dn);
    return rc;
  }
  // He-normal initialisation from a counter-based stream (Box-Muller)
  std::vector<float> p(off, 0.f);
  for (int l = 0; l < kLayers; l++) {
    const int64_t n = (int64_t)dn->L[l].cout * 9 * dn->L[l].cin;
    const double sd = std::sqrt(2.0 / (9.0 * dn->L[l].cin)) * (l == kLayers - 1 ? 0.1 : 1.0);
    for (int64_t i = 0; i < n; i += 2) {
      u4 r = philox4x32_10(u4{(uint32_t)i, (uint32_t)(i >> 32), (uint32_t)l, 0x574e4954u}, (uint32_t)seed, (uint32_t)(seed >> 32));
      const double u1 = ((r.x >> 8) + 1) * (1.0 / 16777216.0), u2 = (r.y >> 8) * (1.0 / 16777216.0);
      const double rad = std::sqrt(-2.0 * std::log(u1));
      p[dn->wOff[l] + i] = (float)(sd * rad * std::cos(6.283185307179586 * u2));
      if (i + 1 < n) p[dn->wOff[l] + i + 1] = (float)(sd * rad * std::sin(6.283185307179586 * u2));
    }
  }
  hipStream_t s = ctx->stream;
  if (hipMemcpyAsync(dn->params.p, p.data(), off * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemsetAsync(dn->m.p, 0, off * 4, s) != hipSuccess || hipMemsetAsync(dn->v.p, 0, off * 4, s) != hipSuccess ||
      packAll(dn) != CPT_OK || hipStreamSynchronize(s) != hipSuccess) {
    cpt_denoiser_destroy(dn);
    return fail(ctx, CPT_ERR_HIP, "cpt_denoiser_create", "upload failed");
  }
  ctx->denoisers.push_back(dn);
  *out = dn;
  return CPT_OK;
}

void cpt_denoiser_destroy(cpt_denoiser *dn) {
  if (!dn) return;
  if (cpt_ctx *ctx = dn->ctx) {
    auto &v = ctx->denoisers;
    v.erase(std::remove(v.begin(), v.end(), dn), v.end());
    cpt::denoiserRelease(dn);
  }
  delete dn;
}

int64_t cpt_denoiser_num_params(const cpt_denoiser *dn) { return dn ? dn->nParams : -1; }

int cpt_denoiser_get_params(cpt_denoiser *dn, float *out) {
  if (!dn || !out || !dn->ctx) return CPT_ERR_ARG;
  CPT_TRY_HIP(dn->ctx, hipMemcpyAsync(out, dn->params.p, dn->nParams * 4, hipMemcpyDeviceToHost, dn->ctx->stream));
  CPT_TRY_HIP(dn->ctx, hipStreamSynchronize(dn->ctx->stream));
  return CPT_OK;
}

int cpt_denoiser_set_params(cpt_denoiser *dn, const float *in) {
  if (!dn || !in || !dn->ctx) return CPT_ERR_ARG;
  CPT_TRY_HIP(dn->ctx, hipMemcpyAsync(dn->params.p, in, dn->nParams * 4, hipMemcpyHostToDevice, dn->ctx->stream));
  int rc = packAll(dn);
  if (rc) return fail(dn->ctx, rc, "cpt_denoiser_set_params", "pack");
  CPT_TRY_HIP(dn->ctx, hipStreamSynchronize(dn->ctx->stream));
  return CPT_OK;
}

int cpt_denoiser_layers(const cpt_denoiser *dn, int32_t *cin, int32_t *cout, int64_t *woff, int64_t *boff) {
  if (!dn) return -1;
  for (int l = 0; l < kLayers; l++) {
    if (cin) cin[l] = dn->L[l].cin;
    if (cout) cout[l] = dn->L[l].cout;
    if (woff) woff[l] = dn->wOff[l];
    if (boff) boff[l] = dn->bOff[l];
  }
  return kLayers;
}

// Forward + loss + backward; grads_out (optional, host or device) receives the
// gradient of every parameter; with lr > 0 an Adam step follows.
int cpt_denoiser_train_step(cpt_denoiser *dn, const float *in10, const float *target, int32_t b, int32_t h, int32_t w,
                            float lr, float *loss_out, float *grads_out, int flags) {
  if (!dn || !dn->ctx || !in10 || !target || b <= 0 || h <= 0 || w <= 0) return CPT_ERR_ARG;
  cpt_ctx *ctx = dn->ctx;
  CPT_TRY_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const size_t npx = (size_t)b * h * w;
  const float *dIn = nullptr, *dT = nullptr;
  int rc = stage(dn, in10, npx * 10, target, npx * 3, flags, &dIn, &dT, nullptr, 0);
  if (rc) return rc;
  CPT_TRY_HIP(ctx, hipEventRecord(ctx->ev[6], s));
  if ((rc = prepare(dn, dIn, b, h, w))) return rc;
  dn->flopsFwd = 0;
  if ((rc = forward(dn, b))) return fail(ctx, rc, "denoiser forward", "launch");
  const int64_t np = (int64_t)b * dn->H * dn->W;
  hipLaunchKernelGGL(k_l1, dim3(nb256(np)), dim3(256), 0, s, dn->pred.as<float>(), kPredChan, dIn, dT, h, w, dn->H, dn->W, np,
                     1.0f / (float)(npx * 3), dn->dPred.as<uint16_t>(), dn->lossPart.as<float>());
  hipLaunchKernelGGL(k_l1_sum, dim3(1), dim3(256), 0, s, dn->lossPart.as<float>(), (int)nb256(np),
                     dn->loss.as<float>());
  if ((rc = backward(dn, b))) return fail(ctx, rc, "denoiser backward", "launch");
  reduceGrads(dn);  // every parameter's gradient (weights and biases of every layer)
  if (lr > 0.f) {
    dn->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)dn->step), c2 = 1.f - powf(b2, (float)dn->step);
    hipLaunchKernelGGL(k_adam, dim3(nb256(dn->nParams)), dim3(256), 0, s, dn->params.as<float>(), dn->grads.as<float>(),
                       dn->m.as<float>(), dn->v.as<float>(), dn->nParams, lr, b1, b2, 1e-8f, c1, c2, 1);
    if ((rc = packAll(dn))) return fail(ctx, rc, "denoiser adam", "launch");
  }
  CPT_TRY_HIP(ctx, hipGetLastError());
  CPT_TRY_HIP(ctx, hipEventRecord(ctx->ev[7], s));
  if (grads_out) {
    if (flags & CPT_DEVICE_PTRS)
      CPT_TRY_HIP(ctx, hipMemcpyAsync(grads_out, dn->grads.p, dn->nParams * 4, hipMemcpyDeviceToDevice, s));
    else
      CPT_TRY_HIP(ctx, hipMemcpyAsync(grads_out, dn->grads.p, dn->nParams * 4, hipMemcpyDeviceToHost, s));
  }
  if (loss_out) {
    CPT_TRY_HIP(ctx, hipMemcpyAsync(loss_out, dn->loss.p, 4, hipMemcpyDeviceToHost, s));
    CPT_TRY_HIP(ctx, hipStreamSynchronize(s));
    CPT_TRY_HIP(ctx, hipEventElapsedTime(&dn->lastStepMs, ctx->ev[6], ctx->ev[7]));
  }
  return CPT_OK;
}

// Denoised radiance for b images: in10 = B*h*w*10 floats (rgb, albedo,
// normal, depth), out_rgb = B*h*w*3.
int cpt_denoise(cpt_denoiser *dn, const float *in10, float *outRgb, int32_t b, int32_t h, int32_t w, int flags) {
  if (!dn || !dn->ctx || !in10 || !outRgb || b <= 0 || h <= 0 || w <= 0) return CPT_ERR_ARG;
  cpt_ctx *ctx = dn->ctx;
  CPT_TRY_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const size_t npx = (size_t)b * h * w;
  const float *dIn = nullptr;
  float *dOut = nullptr;
  int rc = stage(dn, in10, npx * 10, nullptr, 0, flags, &dIn, nullptr, &dOut, npx * 3);
  if (rc) return rc;
  if (flags & CPT_DEVICE_PTRS) dOut = outRgb;
  CPT_TRY_HIP(ctx, hipEventRecord(ctx->ev[6], s));
  if ((rc = prepare(dn, dIn, b, h, w))) return rc;
  dn->flopsFwd = 0;
  if ((rc = forward(dn, b))) return fail(ctx, rc, "denoiser forward", "launch");
  hipLaunchKernelGGL(k_output, dim3(nb256(npx)), dim3(256), 0, s, dn->pred.as<float>(), kPredChan, dIn, h, w, dn->H, dn->W,
                     (int64_t)npx, dOut);
  CPT_TRY_HIP(ctx, hipGetLastError());
  CPT_TRY_HIP(ctx, hipEventRecord(ctx->ev[7], s));
  if (!(flags & CPT_DEVICE_PTRS)) CPT_TRY_HIP(ctx, hipMemcpyAsync(outRgb, dOut, npx * 12, hipMemcpyDeviceToHost, s));
  CPT_TRY_HIP(ctx, hipStreamSynchronize(s));
  CPT_TRY_HIP(ctx, hipEventElapsedTime(&dn->lastFwdMs, ctx->ev[6], ctx->ev[7]));
  return CPT_OK;
}

// Timing and work of the last call: ms of the last cpt_denoise / train step
// and the forward FLOPs of one pass.
int cpt_denoiser_stats(const cpt_denoiser *dn, float *fwd_ms, float *step_ms, double *fwd_flops) {
  if (!dn) return CPT_ERR_ARG;
  if (fwd_ms) *fwd_ms = dn->lastFwdMs;
  if (step_ms) *step_ms = dn->lastStepMs;
  if (fwd_flops) *fwd_flops = dn->flopsFwd;
  return CPT_OK;
}

// One 3x3 convolution on caller data (testing / reuse): mode 0 forward
// (y = conv(x, w) + bias, fp32 out [B,H,W,Cout]), mode 1 backward-weight
// (dw[Cout*9*Cin] fp32 = sum dy x), mode 2 backward-data (dx fp32
// [B,H,W,Cin] from dy and w).  x, w, dy are bf16 bit patterns.
int cpt_conv3x3(cpt_ctx *ctx, int mode, const uint16_t *x, const uint16_t *w, const float *bias, const uint16_t *dy,
                float *out, int32_t B, int32_t H, int32_t W, int32_t Cin, int32_t Cout) {
  if (!ctx || B <= 0 || H <= 0 || W <= 0 || Cin <= 0 || Cout <= 0 || Cin % 32 || Cout % 32 || mode < 0 || mode > 2)
    return fail(ctx, CPT_ERR_ARG, "cpt_conv3x3", "channels must be multiples of 32");
  if ((int64_t)H * W * std::max(Cin, Cout) * 4 >= (1ll << 31))
    return fail(ctx, CPT_ERR_ARG, "cpt_conv3x3", "image too large (H*W*channels*4 B must stay below 2^31)");
  CPT_TRY_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const size_t nx = (size_t)B * H * W * Cin, ny = (size_t)B * H * W * Cout, nw = (size_t)Cout * 9 * Cin;
  DevBuf dx, dw, dyb, db, dout, wt;
  int rc = 0;
  auto cleanup = [&]() {
    DevBuf *all[] = {&dx, &dw, &dyb, &db, &dout, &wt};
    for (DevBuf *b : all)
      if (b->p) (void)hipFree(b->p);
  };
  hipError_t ce = hipSuccess;  // first error of the calls below
  auto chk = [&](hipError_t r) {
    if (ce == hipSuccess) ce = r;
  };
  size_t nout = mode == 0 ? ny : (mode == 1 ? nw : nx);
  if ((rc = ensure(ctx, dx, nx * 2)) || (rc = ensure(ctx, dw, nw * 2)) || (rc = ensure(ctx, dyb, ny * 2)) ||
      (rc = ensure(ctx, db, Cout * 4)) || (rc = ensure(ctx, dout, nout * 4)) || (rc = ensure(ctx, wt, nw * 2))) {
    cleanup();
    return rc;
  }
  if (x) chk(hipMemcpyAsync(dx.p, x, nx * 2, hipMemcpyHostToDevice, s));
  if (w) chk(hipMemcpyAsync(dw.p, w, nw * 2, hipMemcpyHostToDevice, s));
  if (dy) chk(hipMemcpyAsync(dyb.p, dy, ny * 2, hipMemcpyHostToDevice, s));
  if (bias) chk(hipMemcpyAsync(db.p, bias, Cout * 4, hipMemcpyHostToDevice, s));
  else chk(hipMemsetAsync(db.p, 0, Cout * 4, s));
  if (mode == 0) {
    launchConvFwd(s, dx.as<uint16_t>(), Cin, nullptr, 0, dw.as<uint16_t>(), db.as<float>(), Cout, B, H, W, nullptr,
                  dout.as<float>(), Cout, 0);
  } else if (mode == 1) {
    DevBuf part;
    if ((rc = ensure(ctx, part, (size_t)convBwdWSlices(Cin, Cout, B, H, W, nullptr) * (nw + Cout) * 4))) {
      cleanup();
      return rc;
    }
    launchConvBwdW(s, dyb.as<uint16_t>(), Cout, dx.as<uint16_t>(), Cin, nullptr, 0, B, H, W, part.as<float>(),
                   dout.as<float>(),
                   false);
    chk(hipStreamSynchronize(s));
    (void)hipFree(part.p);
  } else {
    // flipped transpose on the host side of the device: reuse k_pack_weights
    // through an fp32 staging copy of w
    DevBuf wf;
    if ((rc = ensure(ctx, wf, nw * 4))) {
      cleanup();
      return rc;
    }
    std::vector<uint16_t> hw(nw);
    std::vector<float> hf(nw);
    if (w) std::memcpy(hw.data(), w, nw * 2);
    for (size_t i = 0; i < nw; i++) {
      uint32_t u = ((uint32_t)hw[i]) << 16;
      std::memcpy(&hf[i], &u, 4);
    }
    chk(hipMemcpyAsync(wf.p, hf.data(), nw * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_pack_weights, dim3(nb256(nw)), dim3(256), 0, s, wf.as<float>(), Cin, Cout, dw.as<uint16_t>(),
                       wt.as<uint16_t>());
    launchConvFwd(s, dyb.as<uint16_t>(), Cout, nullptr, 0, wt.as<uint16_t>(), nullptr, Cin, B, H, W, nullptr,
                  dout.as<float>(), Cin, 0);
    chk(hipStreamSynchronize(s));
    (void)hipFree(wf.p);
  }
  chk(hipGetLastError());
  hipError_t e = ce;
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout.p, nout * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  cleanup();
  if (e != hipSuccess) return fail(ctx, CPT_ERR_HIP, "cpt_conv3x3", hipGetErrorString(e));
  return CPT_OK;
}

}  // extern "C"

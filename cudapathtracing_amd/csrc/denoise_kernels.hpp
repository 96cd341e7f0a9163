// denoise_kernels.hpp - BF16 MFMA kernels of the convolutional denoiser.
//
// Layouts: activations NHWC bf16, channel counts multiples of 32 for conv
// inputs and of 16 for conv outputs; weights OHWI bf16 (W[co][ky][kx][ci]);
// FP32 accumulation everywhere (v_mfma_f32_16x16x32_bf16).
//
// MFMA operand layout (16x16x32, wave64): lane l supplies A[l%16][8*(l/16)+j]
// and B[8*(l/16)+j][l%16] for j = 0..7 and receives D[4*(l/16)+r][l%16],
// r = 0..3.
//
// conv3x3 forward  (implicit GEMM, M = pixels, N = Cout, K = 9*Cin):
//   a workgroup of 4 waves computes 64 consecutive pixels of one image row
//   times NT output channels; per (ky, 32-channel chunk) it stages a 66-pixel
//   input strip and the 3 x NT x 32 weight slab in LDS and issues
//   3 (kx) x NT/16 MFMAs per wave.
// conv3x3 backward-data is the forward kernel run on dY with the weights
//   flipped and transposed (Wt[ci][2-ky][2-kx][co] = W[co][ky][kx][ci]).
// conv3x3 backward-weight (M = Cout, N = Cin per tap, K = pixels): a
//   workgroup owns a 64-channel Cout tile, a 32-channel Cin chunk and a run of
//   image rows; dY and the three input rows are staged transposed
//   (channel-major) in LDS so that each lane reads 8 consecutive pixels, and
//   the 9 taps x 2 Cin blocks accumulate in registers before one FP32 atomic
//   add per weight.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cpt {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

constexpr int kConvThreads = 256;
constexpr int kConvPix = 64;    // output pixels per workgroup (one row segment)
constexpr int kLdsRow = 40;     // bf16 per LDS row (32 + 8 padding)

// Y[b,y,x,co] = act(sum X[b,y+ky-1,x+kx-1,ci] W[co,ky,kx,ci] + bias[co]) (+res)
// outBf (bf16, ldo channels) and/or outF (fp32, ldo channels).
// mode bits: 1 = ReLU, 2 = add res[b,y,x,0..2] (fp32 residual, 3 channels).
template <int NT>
__global__ __launch_bounds__(kConvThreads) void k_conv_fwd(const uint16_t *__restrict__ X, int Cin,
                                                            const uint16_t *__restrict__ Wt,
                                                            const float *__restrict__ bias, int Cout, int H,
                                                            int Wd, uint16_t *__restrict__ outBf,
                                                            float *__restrict__ outF, int ldo, int mode,
                                                            const float *__restrict__ res, int ldr) {
  __shared__ __attribute__((aligned(16))) uint16_t sX[(kConvPix + 2) * kLdsRow];
  __shared__ __attribute__((aligned(16))) uint16_t sW[3 * NT * kLdsRow];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int x0 = blockIdx.x * kConvPix;
  const int y = blockIdx.y;
  const int nCoTiles = (Cout + NT - 1) / NT;
  const int b = blockIdx.z / nCoTiles;
  const int co0 = (blockIdx.z % nCoTiles) * NT;
  f32x4 acc[NT / 16];
#pragma unroll
  for (int i = 0; i < NT / 16; i++) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int64_t imgBase = (int64_t)b * H * Wd;
  for (int ky = 0; ky < 3; ky++) {
    const int yy = y + ky - 1;
    const bool rowOk = yy >= 0 && yy < H;
    for (int c0 = 0; c0 < Cin; c0 += 32) {
      __syncthreads();
      // input strip: 66 pixels x 32 channels, 4 x 16 B per pixel
      for (int q = tid; q < (kConvPix + 2) * 4; q += kConvThreads) {
        const int i = q >> 2, part = q & 3;
        const int xx = x0 - 1 + i;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (rowOk && xx >= 0 && xx < Wd)
          v = *(const uint4 *)(X + (imgBase + (int64_t)yy * Wd + xx) * Cin + c0 + part * 8);
        *(uint4 *)(sX + i * kLdsRow + part * 8) = v;
      }
      // weights: 3 (kx) x NT (co) x 32 (ci)
      for (int q = tid; q < 3 * NT * 4; q += kConvThreads) {
        const int part = q & 3, r = q >> 2;
        const int kx = r / NT, co = r % NT;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (co0 + co < Cout) v = *(const uint4 *)(Wt + ((((int64_t)(co0 + co) * 3 + ky) * 3 + kx) * Cin) + c0 + part * 8);
        *(uint4 *)(sW + (kx * NT + co) * kLdsRow + part * 8) = v;
      }
      __syncthreads();
#pragma unroll
      for (int kx = 0; kx < 3; kx++) {
        const bf16x8 a = *(const bf16x8 *)(sX + (wave * 16 + (lane & 15) + kx) * kLdsRow + (lane >> 4) * 8);
#pragma unroll
        for (int nb = 0; nb < NT / 16; nb++) {
          const bf16x8 bw = *(const bf16x8 *)(sW + (kx * NT + nb * 16 + (lane & 15)) * kLdsRow + (lane >> 4) * 8);
          acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw, acc[nb], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int nb = 0; nb < NT / 16; nb++) {
    const int co = co0 + nb * 16 + (lane & 15);
    if (co >= Cout) continue;
    const float bs = bias ? bias[co] : 0.f;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int x = x0 + wave * 16 + (lane >> 4) * 4 + r;
      if (x >= Wd) continue;
      const int64_t p = imgBase + (int64_t)y * Wd + x;
      float v = acc[nb][r] + bs;
      if (mode & 1) v = fmaxf(v, 0.f);
      if ((mode & 2) && co < 3) v += res[p * ldr + co];
      if (outBf) outBf[p * ldo + co] = f2bf(v);
      if (outF) outF[p * ldo + co] = v;
    }
  }
}

// dW[co][ky][kx][ci] += sum_p dY[p][co] X[p + (ky-1, kx-1)][ci]
// grid: (Cout/64, Cin/32, B * rowGroups); each workgroup walks `rows` rows.
__global__ __launch_bounds__(kConvThreads) void k_conv_bwd_w(const uint16_t *__restrict__ dY, int Cout,
                                                              const uint16_t *__restrict__ X, int Cin, int H,
                                                              int Wd, int rows, float *__restrict__ dW) {
  constexpr int P = kConvPix;
  constexpr int PR = P + 8;  // padded pixel stride (bf16) of the transposed tiles
  __shared__ __attribute__((aligned(16))) uint16_t sDy[64 * PR];           // [co][pixel]
  __shared__ __attribute__((aligned(16))) uint16_t sX[3 * 32 * (P + 2 + 6)];  // [ky][ci][pixel]
  constexpr int XR = P + 8;  // 66 pixels padded to 72
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int co0 = blockIdx.x * 64, c0 = blockIdx.y * 32;
  const int groups = (H + rows - 1) / rows;
  const int b = blockIdx.z / groups;
  const int yBeg = (blockIdx.z % groups) * rows;
  const int yEnd = min(H, yBeg + rows);
  const int64_t imgBase = (int64_t)b * H * Wd;
  // accumulators: this wave's 16 output channels x 2 input blocks x 9 taps
  f32x4 acc[9][2];
#pragma unroll
  for (int t = 0; t < 9; t++) acc[t][0] = acc[t][1] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int y = yBeg; y < yEnd; y++) {
    for (int x0 = 0; x0 < Wd; x0 += P) {
      __syncthreads();
      // dY tile transposed: 64 pixels x 64 co -> sDy[co][pixel]
      for (int q = tid; q < P * 8; q += kConvThreads) {
        const int px = q >> 3, part = q & 7;
        const int x = x0 + px;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (x < Wd && co0 + part * 8 < Cout)
          v = *(const uint4 *)(dY + (imgBase + (int64_t)y * Wd + x) * Cout + co0 + part * 8);
        const uint16_t *h = (const uint16_t *)&v;
#pragma unroll
        for (int j = 0; j < 8; j++) sDy[(part * 8 + j) * PR + px] = h[j];
      }
      // X rows y-1..y+1, pixels x0-1..x0+64, 32 channels -> sX[ky][ci][pixel]
      for (int q = tid; q < 3 * (P + 2) * 4; q += kConvThreads) {
        const int part = q & 3, r = q >> 2;
        const int ky = r / (P + 2), i = r % (P + 2);
        const int yy = y + ky - 1, xx = x0 - 1 + i;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (yy >= 0 && yy < H && xx >= 0 && xx < Wd)
          v = *(const uint4 *)(X + (imgBase + (int64_t)yy * Wd + xx) * Cin + c0 + part * 8);
        const uint16_t *h = (const uint16_t *)&v;
#pragma unroll
        for (int j = 0; j < 8; j++) sX[(ky * 32 + part * 8 + j) * XR + i] = h[j];
      }
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < P / 32; ks++) {
        // A[i=co][k=pixel]: lane row = co (wave's block), 8 consecutive pixels
        const bf16x8 a = *(const bf16x8 *)(sDy + (wave * 16 + (lane & 15)) * PR + ks * 32 + (lane >> 4) * 8);
#pragma unroll
        for (int ky = 0; ky < 3; ky++) {
#pragma unroll
          for (int kx = 0; kx < 3; kx++) {
#pragma unroll
            for (int nb = 0; nb < 2; nb++) {
              // B[k=pixel][n=ci]: pixel p needs input pixel p + kx - 1 -> strip index p + kx
              const uint16_t *src = sX + (ky * 32 + nb * 16 + (lane & 15)) * XR + ks * 32 + (lane >> 4) * 8 + kx;
              bf16x8 bv;
#pragma unroll
              for (int j = 0; j < 8; j++) bv[j] = __builtin_bit_cast(__bf16, src[j]);
              acc[ky * 3 + kx][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[ky * 3 + kx][nb], 0, 0, 0);
            }
          }
        }
      }
    }
  }
  // D[i = co][n = ci]: lane holds rows 4*(lane/16)+r, column lane%16
#pragma unroll
  for (int t = 0; t < 9; t++) {
#pragma unroll
    for (int nb = 0; nb < 2; nb++) {
      const int ci = c0 + nb * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int co = co0 + wave * 16 + (lane >> 4) * 4 + r;
        if (co < Cout && ci < Cin) atomicAdd(dW + ((int64_t)co * 9 + t) * Cin + ci, acc[t][nb][r]);
      }
    }
  }
}

}  // namespace cpt

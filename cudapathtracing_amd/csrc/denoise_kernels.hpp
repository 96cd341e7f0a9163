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
// conv3x3 forward  (implicit GEMM, M = Cout, N = pixels, K = 9*Cin): see
//   k_conv_fwd.
// conv3x3 backward-data is the forward kernel run on dY with the weights
//   flipped and transposed (Wt[ci][2-ky][2-kx][co] = W[co][ky][kx][ci]).
// conv3x3 backward-weight (M = Cout, N = Cin per tap, K = pixels, split
//   over K): see k_conv_bwd_w.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cpt {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((vector_size(16)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// Per bf16 half of a dword: all ones where the value is > 0 (+inf
// included), zero for zero, negative values and NaN -- the ReLU gradient mask.
__device__ __forceinline__ uint32_t reluKeep(uint32_t w) {
  const uint32_t lo = w & 0xffffu, hi = w >> 16;
  const bool keepLo = !(lo & 0x8000u) && lo != 0 && lo <= 0x7f80u;
  const bool keepHi = !(hi & 0x8000u) && hi != 0 && hi <= 0x7f80u;
  return (keepLo ? 0xffffu : 0u) | (keepHi ? 0xffff0000u : 0u);
}

constexpr int kConvThreads = 256;
constexpr int kLdsRow = 40;     // bf16 per LDS row (32 + 8 padding)

// ---- buffer resources (gfx9 raw buffers) -------------------------------------
// A raw buffer resource over `bytes` bytes at p: loads at offsets at or past
// `bytes` return zero (and an LDS-direct load writes zero), stores there are
// dropped.  Offsets are 32-bit; kOob is past every buffer used here.
constexpr uint32_t kOob = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t bufRsrc(const void *p, int64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
typedef __attribute__((address_space(3))) void lds_void;
typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
// The same resource as four words, for the LDS copy below.
__device__ __forceinline__ i32x4 bufDesc(const void *p, int64_t bytes) {
  const uint64_t a = (uint64_t)p;
  i32x4 d;
  d.x = (int32_t)(uint32_t)a;
  d.y = (int32_t)((uint32_t)(a >> 32) & 0xffffu);  // stride 0
  d.z = (int32_t)bytes;
  d.w = 0x00020000;
  return d;
}
// 16 B per lane straight into LDS: lane l writes dst + 16 l (dst
// wave-uniform).  The instruction writes no register (its data goes to the
// LDS address in M0), so an asm block that returns at issue leaves no
// register in flight.  The compiler's builtin for it makes every later LDS
// read wait for all copies in flight (it cannot tell which LDS they write),
// which would serialise the two stages; here the waits are vmWait's.
// M0 is declared clobbered (the compiler keeps no value there across it);
// clang notes that it does not preserve reserved registers for the block,
// which is all this needs.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void loadToLds16(i32x4 r, void *dst, uint32_t off) {
  const uint32_t m0 = __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)(lds_void *)dst);
  asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tbuffer_load_dwordx4 %1, %2, 0 offen lds"
               :
               : "s"(m0), "v"(off), "s"(r)
               : "memory", "m0");
}
#pragma clang diagnostic pop
// Waits until at most N of this wave's vector memory operations are
// outstanding (they complete in issue order).
template <int N>
__device__ __forceinline__ void vmWait() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}
// vmWait with a wave-uniform count (an immediate per case); counts above 47
// wait for everything
__device__ __forceinline__ void vmWaitDyn(int n) {
#define CPT_VMW(N) \
  case N:          \
    vmWait<N>();   \
    break;
#define CPT_VMW8(B) CPT_VMW(B) CPT_VMW(B + 1) CPT_VMW(B + 2) CPT_VMW(B + 3) CPT_VMW(B + 4) CPT_VMW(B + 5) \
    CPT_VMW(B + 6) CPT_VMW(B + 7)
  switch (n) {
    CPT_VMW8(0) CPT_VMW8(8) CPT_VMW8(16) CPT_VMW8(24) CPT_VMW8(32) CPT_VMW8(40)
    default:
      vmWait<0>();
  }
#undef CPT_VMW8
#undef CPT_VMW
}

// Workgroup barrier for LDS traffic only: this wave's LDS reads and writes
// are complete (lgkmcnt 0) and nobody passes before every wave is here.  The
// LDS copies (buffer_load ... lds) are waited for separately with vmWait;
// __syncthreads' workgroup fence would also wait for every outstanding global
// store and copy (vmcnt 0).  The wave-scope fences keep the compiler from
// moving memory operations across it.
__device__ __forceinline__ void ldsBarrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Y[b,y,x,co] = act(sum X[b,y+ky-1,x+kx-1,ci] W[co,ky,kx,ci] + bias[co])
// X has Cin channels; with Xlow it is the decoder's concatenation
// [up2(Xlow) | X] read in place: channels [0, Clow) from Xlow at half
// resolution (pixel (y/2, x/2), Clow channels per pixel), channels
// [Clow, Cin) from X (Cin - Clow channels per pixel).
// into outBf (bf16, ldo channels per pixel) or outF (fp32, ldo channels per
// pixel: channels co >= ldo are not written, so a 4-channel buffer takes the
// first 4 of a wider layer).
// mode bit 1 = ReLU.  maskY (bf16 output only, ldo channels): the value is
// kept where maskY > 0 and zeroed elsewhere -- the ReLU gradient mask of the
// layer whose output maskY is.  poolOut (bf16 output only, H and W even):
// also the 2x2 average of the stored bf16 outputs at (y/2, x/2), ldo
// channels per pixel, summed as ((x00 + x01) + x10) + x11 and scaled by 1/4
// in FP32 -- the encoder's pooling, without reading the output back.  Split over Cin (kSplit > 1): slice s takes
// its chunks and writes fp32 partial sums to outF + s * sliceStride.
//
// Implicit GEMM with the output transposed, D[co][px] = W[co][tap, ci] .
// X[tap, ci][px]: the weights are the A operand and the input pixels the B
// operand, so each lane ends with 4 consecutive output channels of one pixel.
// A workgroup (8 waves) owns a tile of 8 rows x 64 pixels x 16*QB output
// channels, wave w the row y0 + w, and walks `tpw` vertically stacked tiles
// as one sequence of (tile, 32-channel chunk) steps.
//
// Data movement: a step needs the 10 x 66 input halo of its tile (32
// channels, 64 B per pixel) and the chunk's 9 taps of weights (9 x 16*QB
// rows of 64 B).  Both are copied from HBM straight into LDS by
// buffer_load ... lds (16 B per lane, a wave moves 16 pixel records per
// instruction), no registers in between, into a ring of NS stages: step
// k+NS-1's copies are issued right after the barrier that opens step k and
// land while steps k .. k+NS-2 run their MFMAs (NS = 3 where the weights
// stay in LDS, else 2).  Pixels outside the image (and weight rows past Cout)
// get an offset past the buffer, which the copy turns into zeros.  Records
// are 64 B without padding; the 16 B part p of record r sits at part
// p ^ ((r >> 1) & 3), so the 16 B fragment reads of 8 consecutive records
// (one per lane of a quarter wave) cover all 32 banks.  kOneChunk (Cin of one
// chunk per slice): the weights are copied once and stay.
// Waits: a wave's vector memory operations complete in issue order, so
// before step k it waits until only the operations it issued after step k's
// copies (later copies, finished tiles' stores) may still be outstanding;
// every store goes through a buffer resource (out-of-range lanes get kOob)
// so that their number per tile is fixed.  A finished bf16 tile leaves through LDS (the
// step's halo stage, free after a barrier): each wave writes its row there,
// 32 channels at a time, and reads it back as 16 B per lane, so that a store
// instruction writes 16 whole 64 B pixel records.
constexpr int kFwdWaves = 8, kFwdThreads = 64 * kFwdWaves;
constexpr int kTileH = kFwdWaves, kTileW = 64;
constexpr int kHaloH = kTileH + 2, kHaloW = kTileW + 2;
constexpr int kHaloPix = kHaloH * kHaloW;                                    // 660
constexpr int kHaloBytes = kHaloPix * 64;                                    // 42,240
constexpr int kHaloLoads = (kHaloPix + 15) / 16;                             // 42 (the last: 4 pixels)
constexpr int kHaloLoadsPerWave = (kHaloLoads + kFwdWaves - 1) / kFwdWaves;  // 6
static_assert(kFwdWaves * kTileW * kLdsRow * 2 <= kHaloBytes, "output staging exceeds a halo stage");

// byte offset of 16 B part p of 64 B record r in a swizzled LDS stage
__device__ __forceinline__ uint32_t swzOff(int r, int p) {
  return (uint32_t)(r * 64 + ((p ^ ((r >> 1) & 3)) << 4));
}

template <int QB, bool kF32, bool kOneChunk>
__global__ __launch_bounds__(kFwdThreads, 1) void k_conv_fwd(const uint16_t *__restrict__ X, int Cin,
                                                              const uint16_t *__restrict__ Xlow, int Clow,
                                                              const uint16_t *__restrict__ Wt,
                                                              const float *__restrict__ bias, int Cout, int H,
                                                              int Wd, uint16_t *__restrict__ outBf,
                                                              float *__restrict__ outF, int ldo, int mode,
                                                              const uint16_t *__restrict__ maskY, int kSplit,
                                                              int64_t sliceStride, int tpw,
                                                              uint16_t *__restrict__ poolOut) {
  constexpr int NT = 16 * QB;
  constexpr int kWRows = 9 * NT;
  constexpr int kWBytes = kWRows * 64;
  constexpr int kWLoads = kWRows / 16;
  constexpr int kWLoadsPerWave = (kWLoads + kFwdWaves - 1) / kFwdWaves;
  // stores per wave per finished tile: 4 per 32 staged channels (bf16), 4
  // per 16-channel block (fp32)
  constexpr int kEpiStores = kF32 ? 4 * QB : 4 * ((QB + 1) / 2);
  // stages: three where the weights stay (they and three halos fit the 160
  // KB of LDS), else two of halo and weights
  constexpr int NS = kOneChunk ? 3 : 2;
  constexpr int NW = kOneChunk ? 1 : NS;
  static_assert(NS * kHaloBytes + NW * kWBytes <= 160 * 1024, "stages exceed LDS");
  __shared__ __attribute__((aligned(16))) uint8_t sH[NS][kHaloBytes];
  __shared__ __attribute__((aligned(16))) uint8_t sW[NW][kWBytes];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int x0 = blockIdx.x * kTileW;
  const int nCoTiles = (Cout + NT - 1) / NT;
  const int slice = blockIdx.z % kSplit, bz = blockIdx.z / kSplit;
  const int b = bz / nCoTiles;
  const int co0 = (bz % nCoTiles) * NT;
  const int chunksAll = Cin / 32, perSlice = (chunksAll + kSplit - 1) / kSplit;
  const int cBeg = slice * perSlice;
  const int nChunks = chunksAll - cBeg < perSlice ? chunksAll - cBeg : perSlice;
  if (outF) outF += slice * sliceStride;
  const int rowTiles = (H + kTileH - 1) / kTileH;
  const int t0 = blockIdx.y * tpw;
  const int nt = rowTiles - t0 < tpw ? rowTiles - t0 : tpw;
  const int nSteps = nt * nChunks;
  if (nSteps <= 0) return;  // workgroup-uniform
  const int cSkip = Cin - Clow;  // channels per pixel of X
  // per-image resources: every offset below stays under 2^31
  const int64_t pixImg = (int64_t)H * Wd;
  const i32x4 rX = bufDesc(X + b * pixImg * cSkip, pixImg * cSkip * 2);
  const int64_t lowImg = (int64_t)(H >> 1) * (Wd >> 1);
  const i32x4 rL = bufDesc(Clow ? Xlow + b * lowImg * Clow : X, Clow ? lowImg * Clow * 2 : 0);
  const i32x4 rW = bufDesc(Wt, (int64_t)Cout * 9 * Cin * 2);
  f32x4 acc[QB][4];
#pragma unroll
  for (int nb = 0; nb < QB; nb++)
#pragma unroll
    for (int m = 0; m < 4; m++) acc[nb][m] = f32x4{0.f, 0.f, 0.f, 0.f};
  // the bias of the 4 output channels this lane holds in each 16-channel block
  float bsv[QB][4];
#pragma unroll
  for (int nb = 0; nb < QB; nb++) {
    const int co = co0 + nb * 16 + (lane >> 4) * 4;
    float4 bv = make_float4(0.f, 0.f, 0.f, 0.f);
    if (bias && co < Cout) bv = *(const float4 *)(bias + co);
    bsv[nb][0] = bv.x;
    bsv[nb][1] = bv.y;
    bsv[nb][2] = bv.z;
    bsv[nb][3] = bv.w;
  }
  // the bias is in registers before any copy is issued: a later first use
  // would otherwise wait for every copy in flight
  vmWait<0>();
  // this wave's copies per step: halo records, and weight rows where they
  // are copied (kOneChunk: at step 0 only)
  const int hLoads = (kHaloLoads - wave + kFwdWaves - 1) / kFwdWaves;
  const int wLoads = (kWLoads - wave + kFwdWaves - 1) / kFwdWaves;
  // copies of step k into stage k % NS
  auto issue = [&](int k) __attribute__((always_inline)) {
    const int y0 = (t0 + k / nChunks) * kTileH, c0 = (cBeg + k % nChunks) * 32;
    const bool fromLow = c0 < Clow;  // uniform: chunks never straddle Clow
    uint8_t *hb = sH[k % NS];
#pragma unroll
    for (int i = 0; i < kHaloLoadsPerWave; i++) {
      const int j = wave + kFwdWaves * i;  // wave-uniform
      if (j >= kHaloLoads) break;
      const int hp = 16 * j + (lane >> 2), lp = (lane & 3) ^ ((hp >> 1) & 3);
      const int yy = y0 - 1 + hp / kHaloW, xx = x0 - 1 + hp % kHaloW;
      const bool in = hp < kHaloPix && yy >= 0 && yy < H && xx >= 0 && xx < Wd;
      const int yc = in ? yy : 0, xc = in ? xx : 0;
      const uint32_t off =
          fromLow ? (uint32_t)((((yc >> 1) * (Wd >> 1) + (xc >> 1)) * Clow + c0 + lp * 8) * 2)
                  : (uint32_t)((((int64_t)yc * Wd + xc) * cSkip + (c0 - Clow) + lp * 8) * 2);
      const i32x4 r = fromLow ? rL : rX;
      if (j == kHaloLoads - 1) {
        if (lane < (kHaloPix - 16 * j) * 4) loadToLds16(r, hb + j * 1024, in ? off : kOob);
      } else {
        loadToLds16(r, hb + j * 1024, in ? off : kOob);
      }
    }
    if (kOneChunk && k > 0) return;
    uint8_t *wb = sW[kOneChunk ? 0 : k % NS];
#pragma unroll
    for (int i = 0; i < kWLoadsPerWave; i++) {
      const int j = wave + kFwdWaves * i;
      if (j >= kWLoads) break;
      const int r = 16 * j + (lane >> 2), lp = (lane & 3) ^ ((r >> 1) & 3);
      const int t = r / NT, co = co0 + r % NT;
      const uint32_t off = co < Cout ? (uint32_t)((((int64_t)co * 9 + t) * Cin + c0 + lp * 8) * 2) : kOob;
      loadToLds16(rW, wb + j * 1024, off);
    }
  };
  // fragment offsets: B (halo) of tap t for pixel block 0, A (weights) of
  // row lane%16; blocks of 16 pixels or rows keep the swizzle (+1024 B)
  uint32_t hOff[9];
#pragma unroll
  for (int t = 0; t < 9; t++)
    hOff[t] = swzOff((wave + t / 3) * kHaloW + (lane & 15) + t % 3, lane >> 4);
  const uint32_t aOff = swzOff(lane & 15, lane >> 4);
  // fp32 output: lane holds output channels 4*(lane/16) + r of pixel lane%16
  auto epilogueF = [&](int y) __attribute__((always_inline)) {
    float *base = outF + (int64_t)b * pixImg * ldo;
    const __amdgpu_buffer_rsrc_t rO = bufRsrc(base, pixImg * ldo * 4);
#pragma unroll
    for (int nb = 0; nb < QB; nb++) {
      const int co = co0 + nb * 16 + (lane >> 4) * 4;
#pragma unroll
      for (int m = 0; m < 4; m++) {
        const int x = x0 + m * 16 + (lane & 15);
        // channel counts are multiples of 4: a lane's 4 channels are all in or out
        const bool ok = y < H && x < Wd && co < Cout && co < ldo;
        const uint32_t o = ok ? (uint32_t)((((int64_t)y * Wd + x) * ldo + co) * 4) : kOob;
        float4 v;
        v.x = acc[nb][m][0] + bsv[nb][0];
        v.y = acc[nb][m][1] + bsv[nb][1];
        v.z = acc[nb][m][2] + bsv[nb][2];
        v.w = acc[nb][m][3] + bsv[nb][3];
        if (mode & 1) {
          v.x = fmaxf(v.x, 0.f);
          v.y = fmaxf(v.y, 0.f);
          v.z = fmaxf(v.z, 0.f);
          v.w = fmaxf(v.w, 0.f);
        }
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), rO, o, 0, 0);
      }
    }
  };
  // bf16 output through LDS (stg: a free halo stage, after a barrier)
  auto epilogueB = [&](int y, uint8_t *stg) __attribute__((always_inline)) {
    uint16_t *st = (uint16_t *)stg + wave * (kTileW * kLdsRow);
    const int64_t outImg = (int64_t)b * pixImg * ldo;
    const __amdgpu_buffer_rsrc_t rO = bufRsrc(outBf + outImg, pixImg * ldo * 2);
    const __amdgpu_buffer_rsrc_t rM = bufRsrc(maskY ? maskY + outImg : outBf, maskY ? pixImg * ldo * 2 : 0);
#pragma unroll
    for (int half = 0; half < (QB + 1) / 2; half++) {
#pragma unroll
      for (int nb2 = 0; nb2 < 2 && half * 2 + nb2 < QB; nb2++) {
        const int nb = half * 2 + nb2;
        const int col = nb2 * 16 + (lane >> 4) * 4;
#pragma unroll
        for (int m = 0; m < 4; m++) {
          float v[4];
#pragma unroll
          for (int r = 0; r < 4; r++) {
            v[r] = acc[nb][m][r] + bsv[nb][r];
            if (mode & 1) v[r] = fmaxf(v[r], 0.f);
          }
          uint2 pk;
          pk.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
          pk.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
          *(uint2 *)(st + (m * 16 + (lane & 15)) * kLdsRow + col) = pk;
        }
      }
      // the row's values are in LDS: other lanes of this wave read them
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const int nCh = QB - half * 2 >= 2 ? 32 : 16;  // channels staged in this half
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int px = j * 16 + (lane >> 2), part = lane & 3;
        const int x = x0 + px, co = co0 + half * 32 + part * 8;
        const bool ok = y < H && part * 8 < nCh && x < Wd && co < Cout;
        const uint32_t o = ok ? (uint32_t)((((int64_t)y * Wd + x) * ldo + co) * 2) : kOob;
        uint4 v = *(const uint4 *)(st + px * kLdsRow + part * 8);
        if (maskY) {
          const uint4 mv = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rM, o, 0, 0));
          v.x &= reluKeep(mv.x);
          v.y &= reluKeep(mv.y);
          v.z &= reluKeep(mv.z);
          v.w &= reluKeep(mv.w);
        }
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), rO, o, 0, 0);
      }
      if (poolOut) {  // uniform
        // rows y0 + w and y0 + w + 1 are waves w and w + 1: the even waves
        // pool once every wave's row is staged
        ldsBarrier();
        if (!(wave & 1)) {
          const uint16_t *r0 = (const uint16_t *)stg + wave * (kTileW * kLdsRow), *r1 = r0 + kTileW * kLdsRow;
          const int H2 = H >> 1, W2 = Wd >> 1;
          const __amdgpu_buffer_rsrc_t rP =
              bufRsrc(poolOut + (int64_t)b * H2 * W2 * ldo, (int64_t)H2 * W2 * ldo * 2);
#pragma unroll
          for (int j = 0; j < 2; j++) {
            const int q = j * 64 + lane, px = q >> 2, part = q & 3;  // 32 pooled pixels x 4 parts
            const int xo = (x0 >> 1) + px, co = co0 + half * 32 + part * 8;
            const bool ok = y < H && part * 8 < nCh && 2 * xo < Wd && co < Cout;
            const uint32_t o = ok ? (uint32_t)((((int64_t)(y >> 1) * W2 + xo) * ldo + co) * 2) : kOob;
            const uint4 a00 = *(const uint4 *)(r0 + (2 * px) * kLdsRow + part * 8),
                        a01 = *(const uint4 *)(r0 + (2 * px + 1) * kLdsRow + part * 8),
                        a10 = *(const uint4 *)(r1 + (2 * px) * kLdsRow + part * 8),
                        a11 = *(const uint4 *)(r1 + (2 * px + 1) * kLdsRow + part * 8);
            const uint32_t w00[4] = {a00.x, a00.y, a00.z, a00.w}, w01[4] = {a01.x, a01.y, a01.z, a01.w},
                           w10[4] = {a10.x, a10.y, a10.z, a10.w}, w11[4] = {a11.x, a11.y, a11.z, a11.w};
            uint32_t pk[4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
              float lo = __uint_as_float(w00[e] << 16), hi = __uint_as_float(w00[e] & 0xffff0000u);
              lo += __uint_as_float(w01[e] << 16);
              hi += __uint_as_float(w01[e] & 0xffff0000u);
              lo += __uint_as_float(w10[e] << 16);
              hi += __uint_as_float(w10[e] & 0xffff0000u);
              lo += __uint_as_float(w11[e] << 16);
              hi += __uint_as_float(w11[e] & 0xffff0000u);
              pk[e] = (uint32_t)f2bf(lo * 0.25f) | ((uint32_t)f2bf(hi * 0.25f) << 16);
            }
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, make_uint4(pk[0], pk[1], pk[2], pk[3])),
                                                   rP, o, 0, 0);
          }
        }
        ldsBarrier();  // the staged rows are read before the next half overwrites them
      } else {
        // the next half overwrites the staging rows: this wave's reads of them
        // are issued first (LDS operations of a wave complete in order)
        __builtin_amdgcn_wave_barrier();
      }
    }
  };
  // copies (L) and stores (E) this wave issues for step j
  auto L = [&](int j) { return j < nSteps ? hLoads + ((!kOneChunk || j == 0) ? wLoads : 0) : 0; };
  auto E = [&](int j) { return j >= 0 && j % nChunks == nChunks - 1 ? kEpiStores : 0; };
#pragma unroll
  for (int j = 0; j < NS - 1; j++)
    if (j < nSteps) issue(j);
  for (int k = 0; k < nSteps; k++) {
    // Step j's copies are issued at the top of step j - NS + 1 and its
    // stores at the end of step j.  Operations complete in issue order, so
    // step k's copies are in once no more than the younger ones remain:
    // the stores of steps k-NS+1 .. k-1 and the copies of steps k+1 ..
    // k+NS-2.
    int younger = E(k - 1);
    if (NS == 3) younger += E(k - 2) + L(k + 1);
    vmWaitDyn(younger);
    ldsBarrier();  // every wave's copies of step k are in LDS; step k-1 is read
    if (k + NS - 1 < nSteps) issue(k + NS - 1);
    const uint8_t *hb = sH[k % NS];
    const uint8_t *wb = sW[kOneChunk ? 0 : k % NS];
#pragma unroll
    for (int t = 0; t < 9; t++) {
      bf16x8 bx[4], aw[QB];
#pragma unroll
      for (int m = 0; m < 4; m++) bx[m] = *(const bf16x8 *)(hb + hOff[t] + m * 1024);
#pragma unroll
      for (int nb = 0; nb < QB; nb++) aw[nb] = *(const bf16x8 *)(wb + aOff + (t * NT + nb * 16) * 64);
#pragma unroll
      for (int nb = 0; nb < QB; nb++)
#pragma unroll
        for (int m = 0; m < 4; m++) acc[nb][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aw[nb], bx[m], acc[nb][m], 0, 0, 0);
    }
    if (k % nChunks == nChunks - 1) {
      // the tile is complete: write it and start the next one from zero
      const int y = (t0 + k / nChunks) * kTileH + wave;
      if constexpr (kF32) {
        epilogueF(y);
      } else {
        ldsBarrier();  // every wave is done reading this step's halo stage
        epilogueB(y, sH[k % NS]);
      }
#pragma unroll
      for (int nb = 0; nb < QB; nb++)
#pragma unroll
        for (int m = 0; m < 4; m++) acc[nb][m] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
  }
}

// Backward-weight: dW[co][ky][kx][ci] = sum_p dY[p][co] X[p + (ky-1, kx-1)][ci]
// as a GEMM with M = Cout, N = Cin (per tap), K = pixels, split over K.
//
// A K-step is 32 output pixels [x0, x0+32) of one row y of one image.  Both
// operands need the pixels contiguous per channel (a lane supplies 8
// consecutive K), but the tensors are stored channel-contiguous (NHWC), so
// the loader transposes: it loads two adjacent pixels' 8 channels (2 x 16 B)
// and writes, per channel, one dword holding the two pixels (packed pairs,
// one ds_write_b32 per 2 pixels) into
//   sX[ky][ci][pos]  pos = pixel - (x0 - 1) in [0, 34): input rows y-1..y+1
//   sDy[co][k]       k   = pixel - x0     in [0, 32)
// (rows of 40 bf16 = 20 dwords).  Tap (ky, kx) needs sX positions
// [8j + kx, 8j + kx + 8) for lane group j = lane/16: a lane reads the aligned
// 5-dword window at 8j once per (ky, 16-channel block) and takes dwords 0..3
// (kx = 0), the window shifted by one bf16 (kx = 1, v_alignbit) or dwords
// 1..4 (kx = 2): 2 LDS reads for 3 taps.
// A wave owns 32 Cout x 32 Cin x 9 taps (36 accumulator tiles): per K-step
// 2 + 12 fragment reads for 36 MFMAs (D[co][ci] = dY^T . X).  A workgroup
// (4 waves) is KS groups of waves, each group on its own K-steps, each
// covering (32*WCI) Cin x (32*WCO) Cout with WCI*WCO waves (KS*WCI*WCO = 4).
// The next K-step's global loads are issued before this one's MFMAs.  The
// groups' sums are added in LDS and each workgroup writes its tile of one
// slice of `part` (part[split][co][tap][ci], no atomics); k_reduce_splits
// adds the slices in a fixed order (deterministic).
constexpr int kBwRow = 20;  // dwords per LDS row (40 bf16)

// X, Xlow, Clow: the layer's input as in k_conv_fwd (with Xlow, the
// decoder's concatenation read in place; the choice is made per 8-channel
// item, since a 64-channel tile may straddle Clow).
// The workgroups of the first Cin tile also sum dY over their pixels (the
// bias gradient) from the A fragments they hold, into entries nW..nW+Cout of
// their slice (slices are nW + Cout floats).
template <int KS, int WCI, int WCO>
__global__ __launch_bounds__(kConvThreads, 2) void k_conv_bwd_w(const uint16_t *__restrict__ dY, int Cout,
                                                                 const uint16_t *__restrict__ X, int Cin,
                                                                 const uint16_t *__restrict__ Xlow, int Clow, int B,
                                                                 int H, int Wd, int stepsPerSplit,
                                                                 float *__restrict__ part) {
  constexpr int CIW = 32 * WCI, COW = 32 * WCO;
  constexpr int kGroupThreads = kConvThreads / KS;
  constexpr int kXItems = 3 * 17 * (CIW / 8);  // (ky, pixel pair, 8-channel group)
  constexpr int kDyItems = 16 * (COW / 8);
  constexpr int kItems = kXItems + kDyItems;
  constexpr int kIters = (kItems + kGroupThreads - 1) / kGroupThreads;
  constexpr int kGroupWords = (3 * CIW + COW) * kBwRow;
  __shared__ __attribute__((aligned(16))) uint32_t sh[KS * kGroupWords];
  const int tid = threadIdx.x, lane = tid & 63;
  const int g = tid / kGroupThreads, gt = tid % kGroupThreads;
  const int wg = (tid / 64) % (WCI * WCO);  // wave within its group
  const int wci = wg % WCI, wco = wg / WCI;
  uint32_t *sX = sh + g * kGroupWords;
  uint32_t *sDy = sX + 3 * CIW * kBwRow;
  const int ci0 = blockIdx.x * CIW, co0 = blockIdx.y * COW;
  const int nSeg = (Wd + 31) / 32;
  const int64_t nSteps = (int64_t)B * H * nSeg;
  const int64_t kBeg = (int64_t)blockIdx.z * stepsPerSplit;
  const int64_t kEnd = kBeg + stepsPerSplit < nSteps ? kBeg + stepsPerSplit : nSteps;
  const int iters = (stepsPerSplit + KS - 1) / KS;  // the same for every group: barriers stay uniform
  f32x4 acc[9][2][2];
#pragma unroll
  for (int t = 0; t < 9; t++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++) acc[t][cb][0] = acc[t][cb][1] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bool biasWave = blockIdx.x == 0 && wci == 0;  // wave-uniform
  float bsum[2] = {0.f, 0.f};
  uint4 pa[kIters], pb[kIters];
  // global -> registers: the two pixels of every item of K-step k (zeros
  // outside the image or past this group's steps)
  auto load = [&](int64_t k) {
    const bool live = k < kEnd;
    int b = 0, y = 0, x0 = 0;
    if (live) {
      b = (int)(k / ((int64_t)H * nSeg));
      const int rem = (int)(k % ((int64_t)H * nSeg));
      y = rem / nSeg;
      x0 = (rem % nSeg) * 32;
    }
    const int64_t img = (int64_t)b * H * Wd;
#pragma unroll
    for (int i = 0; i < kIters; i++) {
      const int q = gt + i * kGroupThreads;
      uint4 a = make_uint4(0, 0, 0, 0), c = a;
      if (live && q < kXItems) {
        const int pr = q % 17, rest = q / 17;
        const int cig = rest % (CIW / 8), ky = rest / (CIW / 8);
        const int yy = y + ky - 1, xa = x0 - 1 + 2 * pr;
        const int ci = ci0 + cig * 8;
        if (yy >= 0 && yy < H && ci < Cin) {
          if (ci < Clow) {
            // the upsampled half of a concatenation: pixels xa and xa + 1
            // (xa odd) lie in the low-resolution pixels (xa >> 1), (xa + 1) >> 1
            const uint16_t *row = Xlow + ((int64_t)b * (H >> 1) + (yy >> 1)) * (Wd >> 1) * Clow + ci;
            if (xa >= 0 && xa < Wd) a = *(const uint4 *)(row + (int64_t)(xa >> 1) * Clow);
            if (xa + 1 >= 0 && xa + 1 < Wd) c = *(const uint4 *)(row + (int64_t)((xa + 1) >> 1) * Clow);
          } else {
            const int cs = Cin - Clow;
            const uint16_t *row = X + (img + (int64_t)yy * Wd) * cs + (ci - Clow);
            if (xa >= 0 && xa < Wd) a = *(const uint4 *)(row + (int64_t)xa * cs);
            if (xa + 1 >= 0 && xa + 1 < Wd) c = *(const uint4 *)(row + (int64_t)(xa + 1) * cs);
          }
        }
      } else if (live && q < kItems) {
        const int r = q - kXItems;
        const int pr = r % 16, cog = r / 16;
        const int xa = x0 + 2 * pr, co = co0 + cog * 8;
        if (co < Cout) {
          const uint16_t *row = dY + (img + (int64_t)y * Wd) * Cout + co;
          if (xa < Wd) a = *(const uint4 *)(row + (int64_t)xa * Cout);
          if (xa + 1 < Wd) c = *(const uint4 *)(row + (int64_t)(xa + 1) * Cout);
        }
      }
      pa[i] = a;
      pb[i] = c;
    }
  };
  // registers -> LDS, transposed: per channel one dword (pixel 2pr, 2pr + 1)
  auto store = [&]() {
#pragma unroll
    for (int i = 0; i < kIters; i++) {
      const int q = gt + i * kGroupThreads;
      if (q >= kItems) continue;
      uint32_t *dst;
      int pr;
      if (q < kXItems) {
        pr = q % 17;
        const int rest = q / 17;
        const int cig = rest % (CIW / 8), ky = rest / (CIW / 8);
        dst = sX + (ky * CIW + cig * 8) * kBwRow + pr;
      } else {
        const int r = q - kXItems;
        pr = r % 16;
        dst = sDy + (r / 16) * 8 * kBwRow + pr;
      }
      const uint32_t av[4] = {pa[i].x, pa[i].y, pa[i].z, pa[i].w}, cv[4] = {pb[i].x, pb[i].y, pb[i].z, pb[i].w};
#pragma unroll
      for (int j = 0; j < 4; j++) {
        dst[(2 * j) * kBwRow] = (av[j] & 0xffffu) | (cv[j] << 16);
        dst[(2 * j + 1) * kBwRow] = (av[j] >> 16) | (cv[j] & 0xffff0000u);
      }
    }
  };
  load(kBeg + g);
  store();
  __syncthreads();
  for (int it = 0; it < iters; it++) {
    if (it + 1 < iters) load(kBeg + g + (int64_t)(it + 1) * KS);
    // A: dY^T, co rows wco*32 + ob*16 + lane%16, pixels 8j..8j+7
    bf16x8 a[2];
#pragma unroll
    for (int ob = 0; ob < 2; ob++) {
      a[ob] = *(const bf16x8 *)(sDy + (wco * 32 + ob * 16 + (lane & 15)) * kBwRow + (lane >> 4) * 4);
      if (biasWave) {
#pragma unroll
        for (int j = 0; j < 8; j++) bsum[ob] += (float)a[ob][j];
      }
    }
#pragma unroll
    for (int ky = 0; ky < 3; ky++) {
#pragma unroll
      for (int cb = 0; cb < 2; cb++) {
        const uint32_t *src = sX + (ky * CIW + wci * 32 + cb * 16 + (lane & 15)) * kBwRow + (lane >> 4) * 4;
        const uint4 w4 = *(const uint4 *)src;
        const uint32_t w5 = src[4];
        const uint32_t d[5] = {w4.x, w4.y, w4.z, w4.w, w5};
#pragma unroll
        for (int kx = 0; kx < 3; kx++) {
          uint32_t v[4];
#pragma unroll
          for (int j = 0; j < 4; j++)
            v[j] = kx == 0 ? d[j] : (kx == 2 ? d[j + 1] : __builtin_amdgcn_alignbit(d[j + 1], d[j], 16));
          const bf16x8 bx = __builtin_bit_cast(bf16x8, make_uint4(v[0], v[1], v[2], v[3]));
#pragma unroll
          for (int ob = 0; ob < 2; ob++)
            acc[ky * 3 + kx][cb][ob] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ob], bx, acc[ky * 3 + kx][cb][ob], 0, 0, 0);
        }
      }
    }
    if (it + 1 < iters) {
      __syncthreads();  // every wave is done reading this K-step
      store();
      __syncthreads();
    }
  }
  // the KS groups' sums of one tile: added into group 0's through LDS, one
  // tap at a time (16 floats per lane, lane-fastest so that the words of a
  // wave fall in distinct banks)
  // bias sums: the 4 lane groups of a wave hold different pixels
#pragma unroll
  for (int ob = 0; ob < 2; ob++) {
    bsum[ob] += __shfl_xor(bsum[ob], 16, 64);
    bsum[ob] += __shfl_xor(bsum[ob], 32, 64);
  }
  if (KS > 1) {
    constexpr int kLanes = WCI * WCO * 64;
    float *red = (float *)sh;  // (KS-1) * 16 * kLanes floats, within sh
    static_assert((KS - 1) * 16 * kLanes <= KS * kGroupWords, "reduction buffer exceeds LDS");
    const int li = wg * 64 + lane;
#pragma unroll
    for (int t = 0; t < 10; t++) {  // 9 taps, then the bias sums
      __syncthreads();
      if (g > 0) {
        if (t < 9) {
#pragma unroll
          for (int e = 0; e < 16; e++) red[((g - 1) * 16 + e) * kLanes + li] = acc[t][e >> 3][(e >> 2) & 1][e & 3];
        } else {
          red[((g - 1) * 16 + 0) * kLanes + li] = bsum[0];
          red[((g - 1) * 16 + 1) * kLanes + li] = bsum[1];
        }
      }
      __syncthreads();
      if (g == 0) {
        for (int gg = 1; gg < KS; gg++) {
          if (t < 9) {
#pragma unroll
            for (int e = 0; e < 16; e++) acc[t][e >> 3][(e >> 2) & 1][e & 3] += red[((gg - 1) * 16 + e) * kLanes + li];
          } else {
            bsum[0] += red[((gg - 1) * 16 + 0) * kLanes + li];
            bsum[1] += red[((gg - 1) * 16 + 1) * kLanes + li];
          }
        }
      }
    }
    if (g > 0) return;
  }
  // D[co][ci]: lane holds co = 4*(lane/16) + r, ci = lane%16 of each tile
  const int64_t nW = (int64_t)Cout * 9 * Cin;
  float *out = part + (int64_t)blockIdx.z * (nW + Cout);
  if (biasWave && lane < 16) {
#pragma unroll
    for (int ob = 0; ob < 2; ob++) {
      const int co = co0 + wco * 32 + ob * 16 + lane;
      if (co < Cout) out[nW + co] = bsum[ob];
    }
  }
#pragma unroll
  for (int t = 0; t < 9; t++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++) {
      const int ci = ci0 + wci * 32 + cb * 16 + (lane & 15);
#pragma unroll
      for (int ob = 0; ob < 2; ob++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int co = co0 + wco * 32 + ob * 16 + (lane >> 4) * 4 + r;
          if (co < Cout && ci < Cin) out[((int64_t)co * 9 + t) * Cin + ci] = acc[t][cb][ob][r];
        }
    }
}

// out[i] = sum over the slices (stride nS floats) of part[s][i], i < nOut,
// in slice order.
__global__ void k_reduce_splits(const float *__restrict__ part, int slices, int64_t nS, int64_t nOut,
                                float *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nOut) return;
  float s = 0.f;
#pragma unroll 8
  for (int k = 0; k < slices; k++) s += part[(int64_t)k * nS + i];
  out[i] = s;
}

}  // namespace cpt

// render.hip - wavefront path tracer for gfx950.
//
// Replaces the megakernel render<<<>>> (CudaPathTracing/Render.cu:16-261):
// one path per (pixel, sample), state in HBM (SoA), index queues between
//   k_gen     camera ray (Camera::getRandomSampleRay, Camera.cuh:37-43) in FP64
//   k_extend_persist (k_extend)  closest hit (traverseIterative) of the
//             active queue, lanes refilled as their walks end
//   k_shade   miss/sky, emitter, heterogeneous medium (delta tracking), NEE
//             to every quad light (Render.cu:161-187), Russian roulette
//             (Render.cu:189-204), BSDF sampling (Render.cu:206-218); pushes
//             survivors to the next queue, shadow rays to the shadow queue and
//             volume-object hits to the volume queue
//   k_volume  the homogeneous volume object walk (Render.cu:64-135)
//   k_shadow_persist (k_shadow)  occlusion query of each shadow ray up to
//             its light (the reference's "the light is the closest hit",
//             Render.cu:180-185); adds the contribution of an unoccluded ray,
//             or with a medium marks it for k_transmit (ratio tracking)
//   k_gather, k_gather_aov, k_gather_pack  per-pixel sums of the batch
// Radiance is kept per path and summed per pixel in sample order at the end
// of each batch, then divided by spp (Render.cu:220-222).  Random numbers:
// the counter-based stream of cpt_device.hpp, consumed in the reference's
// order per sample (jitter x,y; per vertex: 2 per light for NEE, 1 for
// roulette, 3 for the BSDF).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "bsdf.hpp"
#include "cpt_device.hpp"
#include "internal.hpp"

namespace cpt {

constexpr int kRenderThreads = 256;
constexpr float kRayTMin = 0.001f;        // Render.cu tMin (BVH.cuh:282 default)
constexpr float kRayTMax = 999999.0f;     // INF, rtmath.cuh:8
constexpr uint32_t kFlagNee = 1u << 16;   // hasSampleLight (Render.cu:56)
constexpr uint32_t kFlagAov = 1u << 17;   // first-vertex AOV written
constexpr uint32_t kDepthMask = 0xffffu;
// Roulette draws u are multiples of 2^-24 (exact in FP32 and FP64), so the
// reference's FP64 test u > p equals the FP32 test u > rdn(p), p rounded
// down to FP32: no multiple of 2^-24 lies in (rdn(p), p].  Round-to-nearest
// can round p up past such a multiple (0.8f > 0.8), flipping that draw.
constexpr float kRoulette07 = 0.699999988079071044921875f;  // rdn(0.7) (= 0.7f), Render.cu:202
constexpr float kRoulette08 = 0.7999999523162841796875f;    // rdn(0.8) (0.8f rounds up), Render.cu:89
constexpr int kMaxBounces = 4096;
// Error bits the kernels set (atomicOr) where an invariant the queue sizes
// rely on is broken; the host turns any of them into CPT_ERR_STATE instead
// of a silently wrong image.
constexpr uint32_t kErrShadowCap = 1u;  // more shadow rays than the shadow queue holds
constexpr uint32_t kErrQueueCap = 2u;   // a next-queue or volume-queue slot beyond the batch
constexpr uint32_t kErrMaterial = 4u;   // a hit record's material index outside the table
constexpr uint32_t kErrSlot = 8u;       // a radiance slot or queue entry outside the batch

struct RenderBufs {
  int64_t cap = 0, shadowCap = 0;
  int64_t npix = 0;
  // path state, double-buffered and indexed by queue slot: k_shade writes
  // the survivors compacted into the other set
  DevBuf ray[2], thr[2], meta[2];
  DevBuf hit, permA, permB, qVol, keysA, keysB, keysS, shadow, counters, fb, aov, cams, lbuf, apath;
  DevBuf stage;                   // device copy of the outputs for host-pointer calls
  uint32_t *hostCnt = nullptr;    // pinned: per bounce parity, the counters read back
  hipEvent_t ev[2][7] = {};       // per bounce parity: extend, shade, shadow pairs, readback
};

struct CamDev {
  double lf[3], pc[3], dh[3], dv[3];
  float bg[3];
  int w, h;
};

// Heterogeneous medium: a density grid in [0,1] (x fastest) over an
// axis-aligned box; sigma_t(x) = sigma * density(nearest voxel).
struct MediumDev {
  const float *density;
  f3 bmin, bmax, scale;  // scale = voxels per unit length
  int nx, ny, nz, on;
  float sigma;
  f3 albedo;
};

constexpr int kMaxCollisions = 100000;

// [t0, t1] = [tmin, tmax] clipped to the medium box (slab test); false if empty.
__device__ __forceinline__ bool clipBox(const MediumDev &md, f3 o, f3 d, float tmin, float tmax, float &t0,
                                        float &t1) {
  t0 = tmin;
  t1 = tmax;
  const float ox[3] = {o.x, o.y, o.z}, dx[3] = {d.x, d.y, d.z};
  const float lo[3] = {md.bmin.x, md.bmin.y, md.bmin.z}, hi[3] = {md.bmax.x, md.bmax.y, md.bmax.z};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float inv = 1.0f / dx[k];
    float ta = (lo[k] - ox[k]) * inv, tb = (hi[k] - ox[k]) * inv;
    if (ta > tb) {
      const float tmp = ta;
      ta = tb;
      tb = tmp;
    }
    t0 = fmaxf(t0, ta);
    t1 = fminf(t1, tb);
  }
  return t0 < t1;
}

__device__ __forceinline__ float densityAt(const MediumDev &md, f3 p) {
  const float fx = fminf(fmaxf(floorf((p.x - md.bmin.x) * md.scale.x), 0.0f), (float)(md.nx - 1));
  const float fy = fminf(fmaxf(floorf((p.y - md.bmin.y) * md.scale.y), 0.0f), (float)(md.ny - 1));
  const float fz = fminf(fmaxf(floorf((p.z - md.bmin.z) * md.scale.z), 0.0f), (float)(md.nz - 1));
  return md.density[((int64_t)(int)fz * md.ny + (int)fy) * md.nx + (int)fx];
}

// Ratio tracking (Novak et al. 2014) of the transmittance over the part of
// [tmin, tmax] inside the box: the product of (1 - density) at tentative
// collisions.  Stream: (pixel, sample) with tag 0x80000000 | path draw count
// when the shadow ray was made.
__device__ __forceinline__ float ratioTrack(const MediumDev &md, uint64_t seed, uint32_t pixel, uint32_t sample,
                                            uint32_t pathCnt, f3 o, f3 d, float tmin, float tmax) {
  float t0, t1;
  if (!clipBox(md, o, d, tmin, tmax, t0, t1)) return 1.0f;
  Rng rng;
  rng.init(seed, pixel, sample, 0, 0x80000000u | pathCnt);
  const float inv = 1.0f / (md.sigma * len3(d));
  float t = t0, tr = 1.0f;
  for (int guard = 0; guard < kMaxCollisions; guard++) {
    t -= __logf(1.0f - rng.next()) * inv;
    if (!(t < t1)) break;
    tr *= 1.0f - densityAt(md, o + d * t);
  }
  return tr;
}

struct RenderArgs {
  const float4 *nodes, *leaves, *rec;  // nodes: binary nodes or four-wide records
  int wide;  // 0 binary, 1 four-wide FP32, 2 four-wide quantized, 3 eight-wide quantized
  const MatDev *mats;
  int nPrims, nLights;
  const CamDev *cams;  // nviews cameras (device), all of the same size
  int nviews;
  uint64_t seed;
  int spp, sampleBase, profile, maxDepth, rrDepth, wantAov;
  int tiled;  // batches hold whole samples: k_gen orders each sample's pixels in 8x8 tiles
  float rrProb;  // survival probability (the weight 1 / rrProb)
  float rrCmp;   // rdn(rr_prob): the roulette compares u > rrCmp
  int64_t npix;
  MediumDev med;
  int64_t base;  // first path of the batch (path g = (sample - sampleBase) * npix * nviews + pixel)
  float4 *lbuf;  // radiance per path of the batch, index g - base
  // first-vertex AOVs per path of the batch (wantAov), two float4 at index
  // 2 (g - base): {albedo rgb, normal x} {normal y, normal z, depth, 0};
  // each path writes its own once, so they are summed per pixel in sample
  // order like the radiance (deterministic, no atomics)
  float4 *apath;
  int firstBounce;  // this shading launch is bounce 0 (every path of the batch)
  // sort keys written with each next-queue and shadow ray (renderKey)
  const float *rootBox;
  uint32_t *qkeys, *skeys;
  uint32_t *err;   // error bits (kErr*)
  int64_t cap;     // paths per batch: bound of every queue slot and radiance slot
  int nMats;       // entries of the material table (materials + one per light)
};

// 24-bit sort key of a ray for the queue orderings: direction octant over
// the Morton code of the origin's cell on a 128^3 grid over the root box.
// Queues sort on its top bits: bounce rays on 8 (octant and a 4^3 cell grid,
// one radix pass), shadow rays on 16 (octant, 32^3 cells, two passes); finer
// keys cost more sorting than they save in traversal.
__device__ __forceinline__ uint32_t spread7(uint32_t v) {
  v &= 0x7fu;
  v = (v | (v << 8)) & 0x0000f00fu;
  v = (v | (v << 4)) & 0x000c30c3u;
  v = (v | (v << 2)) & 0x00249249u;
  return v;
}

__device__ __forceinline__ uint32_t renderKey(const float *__restrict__ box, f3 o, f3 d) {
  const float p[3] = {o.x, o.y, o.z};
  uint32_t c[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float lo = box[k], ext = box[3 + k] - lo;
    const float u = ext > 0.0f ? (p[k] - lo) / ext : 0.0f;
    c[k] = (uint32_t)fminf(fmaxf(u * 128.0f, 0.0f), 127.0f);  // NaN -> 0
  }
  const uint32_t oct = (d.x > 0.0f ? 4u : 0u) | (d.y > 0.0f ? 2u : 0u) | (d.z > 0.0f ? 1u : 0u);
  return (oct << 21) | (spread7(c[0]) << 2) | (spread7(c[1]) << 1) | spread7(c[2]);
}
constexpr int kQueueKeyBits = 8, kShadowKeyBits = 16;

// Batch-local index of path (pixel over all views, sample).
__device__ __forceinline__ int64_t pathSlot(const RenderArgs &a, uint32_t pixel, uint32_t sample) {
  return (int64_t)(sample - (uint32_t)a.sampleBase) * (a.npix * a.nviews) + pixel - a.base;
}

// Adds to a path's radiance.  One lane owns a path within a kernel, so a
// plain read-modify-write needs no atomic.
__device__ __forceinline__ void addRadiance(float4 *lbuf, int64_t slot, f3 v) {
  float4 l = lbuf[slot];
  l.x += v.x;
  l.y += v.y;
  l.z += v.z;
  lbuf[slot] = l;
}

__device__ __forceinline__ uint32_t laneId() { return __lane_id(); }

// Closest hit (kAny: any hit) over the record format kFmt: 0 binary nodes,
// 1 four-wide FP32, 2 four-wide quantized, 3 eight-wide quantized.  Kernels
// are instantiated per format so each carries only its walk's registers.
template <bool kCount, int kFmt, bool kAny = false>
__device__ __forceinline__ int traceFmt(const float4 *__restrict__ nodes, const float4 *__restrict__ leaves, int n,
                                        f3 o, f3 d, float tmin, float &tbest, int *stk, int stride,
                                        TraceCounters &cnt) {
  if (kFmt == 3) return traceClosest8<kCount, kAny>(nodes, leaves, n, o, d, tmin, tbest, stk, stride, cnt);
  if (kFmt == 2) return traceClosest4<kCount, kAny, true>(nodes, leaves, n, o, d, tmin, tbest, stk, stride, cnt);
  if (kFmt == 1) return traceClosest4<kCount, kAny>(nodes, leaves, n, o, d, tmin, tbest, stk, stride, cnt);
  return traceClosest<kCount, kAny>(nodes, leaves, n, o, d, tmin, tbest, stk, stride, cnt);
}

// Runtime-format closest hit, for the rare volume-object walk.
__device__ __forceinline__ int traceAny(int wide, const float4 *__restrict__ nodes, const float4 *__restrict__ leaves,
                                        int n, f3 o, f3 d, float tmin, float &tbest, int *stk, int stride,
                                        TraceCounters &cnt) {
  switch (wide) {
    case 3: return traceFmt<false, 3>(nodes, leaves, n, o, d, tmin, tbest, stk, stride, cnt);
    case 2: return traceFmt<false, 2>(nodes, leaves, n, o, d, tmin, tbest, stk, stride, cnt);
    case 1: return traceFmt<false, 1>(nodes, leaves, n, o, d, tmin, tbest, stk, stride, cnt);
    default: return traceFmt<false, 0>(nodes, leaves, n, o, d, tmin, tbest, stk, stride, cnt);
  }
}

// Wave-aggregated append: returns the slot for lanes with pred, ~0u otherwise.
__device__ __forceinline__ uint32_t waveAppend(uint32_t *counter, bool pred) {
  const uint64_t m = __ballot(pred);
  if (m == 0) return ~0u;
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if ((int)laneId() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = __shfl(base, leader, 64);
  const uint64_t lt = (1ull << laneId()) - 1ull;
  return pred ? base + (uint32_t)__popcll(m & lt) : ~0u;
}

// Workgroup-aggregated append: one global atomic per workgroup instead of
// one per wave (every wave of a kernel's workgroup appends to the same
// counter, so per-wave atomics on that one address serialise).  Every
// thread of the workgroup must call it; sh holds kRenderThreads / 64 + 1
// words of LDS.
__device__ __forceinline__ uint32_t blockAppend(uint32_t *counter, bool pred, uint32_t *sh) {
  constexpr int kWaves = kRenderThreads / 64;
  const uint64_t m = __ballot(pred);
  const int wave = (int)(threadIdx.x >> 6), lane = (int)laneId();
  if (lane == 0) sh[wave] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) tot += sh[w];
    sh[kWaves] = tot ? atomicAdd(counter, tot) : 0u;
  }
  __syncthreads();
  uint32_t base = sh[kWaves];
#pragma unroll
  for (int w = 0; w < kWaves; w++)
    if (w < wave) base += sh[w];
  __syncthreads();  // sh is reused by the next call
  return pred ? base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull)) : ~0u;
}

__global__ void k_fill(float *__restrict__ p, int64_t n, float v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_zero_u32(uint32_t *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// Pixel of position q (0 <= q < w*h) of a sample in tile order: 8-row bands
// top to bottom, each in 8x8 tiles left to right (a narrower last tile, a
// shorter last band), row-major inside a tile; a bijection on [0, w*h).  A
// wave of 64 consecutive paths then covers an 8x8 tile rather than a strip
// of one row, so its camera rays, and its first bounces, stay coherent.
__device__ __forceinline__ uint32_t tilePixel(uint32_t q, uint32_t w, uint32_t h) {
  const uint32_t j0 = (q / (8u * w)) * 8u;
  const uint32_t hb = min(8u, h - j0);
  const uint32_t r = q - j0 * w;
  const uint32_t nT = w / 8u;
  uint32_t i, jj;
  if (r < nT * 8u * hb) {
    const uint32_t t = r / (8u * hb), rr = r % (8u * hb);
    jj = rr / 8u;
    i = t * 8u + rr % 8u;
  } else {
    const uint32_t wr = w - nT * 8u, r2 = r - nT * 8u * hb;
    jj = r2 / wr;
    i = nT * 8u + r2 % wr;
  }
  return (j0 + jj) * w + i;
}

__global__ __launch_bounds__(kRenderThreads) void k_gen(RenderArgs a, int64_t base, int64_t count,
                                                         float4 *__restrict__ ray, float4 *__restrict__ thr,
                                                         uint4 *__restrict__ meta) {
#pragma clang fp contract(off)
  const int64_t p = (int64_t)blockIdx.x * kRenderThreads + threadIdx.x;
  if (p >= count) return;
  const int64_t gi = base + p;
  // path gi: sample-major over all views' pixels; the stream of a pixel is
  // keyed by its index over all views (view * npix + pixel)
  const int64_t npixAll = a.npix * a.nviews;
  const uint32_t q = (uint32_t)(gi % npixAll);
  const uint32_t sample = (uint32_t)(a.sampleBase + gi / npixAll);
  const CamDev &cam = a.cams[q / (uint32_t)a.npix];
  // with whole samples per batch, the batch's paths are the same set in
  // tile order (each radiance slot is the path's own, see pathSlot)
  const uint32_t pv = a.tiled ? tilePixel(q % (uint32_t)a.npix, (uint32_t)cam.w, (uint32_t)cam.h)
                              : q % (uint32_t)a.npix;
  const uint32_t pixel = q - q % (uint32_t)a.npix + pv;
  Rng rng;
  rng.init(a.seed, pixel, sample, 0);
  const double rx = rng.next(), ry = rng.next();
  const int i = (int)(pv % (uint32_t)cam.w), j = (int)(pv / (uint32_t)cam.w);
  const double fx = rx - 0.5 + i, fy = ry - 0.5 + j;
  double d[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double pix = (cam.pc[k] + fx * cam.dh[k]) + fy * cam.dv[k];
    d[k] = pix - cam.lf[k];
  }
  ray[2 * p] = make_float4((float)cam.lf[0], (float)cam.lf[1], (float)cam.lf[2], kRayTMin);
  ray[2 * p + 1] = make_float4((float)d[0], (float)d[1], (float)d[2], kRayTMax);
  thr[p] = make_float4(1.f, 1.f, 1.f, 0.f);
  meta[p] = make_uint4(pixel, sample, rng.cnt, 0u);
  a.lbuf[p] = make_float4(0.f, 0.f, 0.f, 0.f);
  // the AOV record is written by the path's first shading (zeros where that
  // vertex sets none), so it is not cleared here
}


// kCount: also sums node visits and primitive tests into work[0], work[1],
// and each wave's largest node-visit count into work[2]
// kPerm: the queue is traced in ray-key order perm: lane k traces the path
// in slot perm[k] and writes its hit record at that slot, so the shading
// kernel reads every array at the same slot.
template <bool kCount, int kFmt, bool kPerm>
__global__ __launch_bounds__(kRenderThreads) void k_extend(const uint32_t *__restrict__ countDev,
                                                            const float4 *__restrict__ ray,
                                                            const float4 *__restrict__ nodes,
                                                            const float4 *__restrict__ leaves, int nPrims,
                                                            float4 *__restrict__ hit,
                                                            unsigned long long *__restrict__ work, int wide,
                                                            const uint32_t *__restrict__ perm) {
  __shared__ int stack[kStackLds * kRenderThreads];
  const uint32_t k = blockIdx.x * kRenderThreads + threadIdx.x;
  const uint32_t count = *countDev;  // the grid covers an upper bound of it
  TraceCounters cnt;
  if (k < count) {
    const uint32_t p = kPerm ? perm[k] : k;
    const float4 r0 = ray[2 * (int64_t)p], r1 = ray[2 * (int64_t)p + 1];
    float tb = r1.w;
    const int leaf = traceFmt<kCount, kFmt>(nodes, leaves, nPrims, xyz(r0), xyz(r1), r0.w, tb, stack + threadIdx.x,
                                            kRenderThreads, cnt);
    // the winner's record was just read by the walk (still cached): hand the
    // shading kernel its type, material and outward normal (Hittable.cuh:51,
    // 128, 393) so that it gathers no leaf record
    float4 h1 = make_float4(0.f, 0.f, 0.f, 0.f);
    int type = 0, mat = 0;
    if (leaf >= 0) {
      const float4 l0 = leaves[3 * (int64_t)leaf], l1 = leaves[3 * (int64_t)leaf + 1],
                   l2 = leaves[3 * (int64_t)leaf + 2];
      type = __float_as_int(l1.w);
      mat = __float_as_int(l2.w);
      const f3 nrm = outwardNormal(l0, l1, l2, type, xyz(r0) + xyz(r1) * tb);
      h1 = make_float4(nrm.x, nrm.y, nrm.z, 0.f);
    }
    hit[2 * (int64_t)p] = make_float4(__int_as_float(leaf), tb, __int_as_float(type), __int_as_float(mat));
    hit[2 * (int64_t)p + 1] = h1;
  }
  if (kCount) {
    // sums over lanes, and the wave's longest walk (the wave runs that many
    // record visits, so sum / (64 * max) is the SIMD lane utilisation)
    unsigned long long a = (unsigned long long)cnt.nodes, b = (unsigned long long)cnt.prims;
    int mx = cnt.nodes;
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_xor(a, off, 64);
      b += __shfl_xor(b, off, 64);
      mx = max(mx, __shfl_xor(mx, off, 64));
    }
    if (__lane_id() == 0) {
      atomicAdd(work + 0, a);
      atomicAdd(work + 1, b);
      atomicAdd(work + 2, (unsigned long long)mx);
    }
  }
}

// Persistent k_extend for the four-wide formats (kQuant: 64 B records): a
// fixed grid of waves whose lanes take the next path of the queue as soon as
// their walk is over (claimed in chunks, loads staged one loop trip ahead of
// use), so that lanes do not wait for the longest walk of their wave.  The
// walk keeps the best hit's type, material and outward normal, so a finished
// lane writes its hit record without a further load.  Same outputs as
// k_extend<kCount, 1 or 2, kPerm>; counting sums the walks and, in work[2],
// the loop trips in which the wave walked.  grab: a zeroed counter.
template <bool kCount, bool kQuant, bool kPerm>
__global__ __launch_bounds__(kRenderThreads) void k_extend_persist(
    const uint32_t *__restrict__ countDev, const float4 *__restrict__ ray, const float4 *__restrict__ nodes4,
    const float4 *__restrict__ leaves, float4 *__restrict__ hit, unsigned long long *__restrict__ work,
    const uint32_t *__restrict__ perm, uint32_t *grab, const uint32_t *abortFlag) {
  __shared__ int stack[kStackLds * kRenderThreads];
  __shared__ int pend[kPend * kRenderThreads];
  int *stk = stack + threadIdx.x;
  int *pnd = pend + threadIdx.x;
  int ovf[kStack4Total - kStackLds];
  // a lane's ray is loaded straight into its (idle) walk state
  Walk4 w;
  walk4Init(w, mkf3(0.f, 0.f, 0.f), mkf3(1.f, 1.f, 1.f), 0.f, 0.f);
  TraceCounters cnt;
  uint32_t p = 0;
  int st = kIdle;
  unsigned long long trips = 0;
  const uint32_t count = *countDev;
  uint32_t pollTrips = 0;
  WaveChunk ch;
  waveChunkInit(ch, grab);
  while (true) {
    // paths loaded in the last trip start their walk
    if (st == kLoading) {
      walk4Start(w);
      st = kActive;
    }
    // claimed paths whose queue slot has arrived are loaded
    if (kPerm && st == kClaimed) {
      walk4Ray(w, ray[2 * (int64_t)p], ray[2 * (int64_t)p + 1]);
      st = kLoading;
    }
    uint32_t item = 0;
    abortRelay(abortFlag, grab, pollTrips);
    if (waveClaim(ch, grab, count, st == kIdle, item)) {
      if (kPerm) {
        p = perm[item];
        st = kClaimed;
      } else {
        p = item;
        walk4Ray(w, ray[2 * (int64_t)p], ray[2 * (int64_t)p + 1]);
        st = kLoading;
      }
    }
    if (!ch.more && __ballot(st != kIdle) == 0) break;
    bool visit = false;
    if (st == kActive && !walk4Step<kCount, false, kQuant, true, true>(w, nodes4, leaves, stk, kRenderThreads,
                                                                          ovf, cnt, pnd, &visit)) {
      hit[2 * (int64_t)p] =
          make_float4(__int_as_float(w.hit), w.tbest, __int_as_float(w.tm & 7), __int_as_float(w.tm >> 3));
      hit[2 * (int64_t)p + 1] = make_float4(w.nrm.x, w.nrm.y, w.nrm.z, 0.f);
      st = kIdle;
    }
    // record-visit trips only (the lane utilisation of the visits)
    if (kCount && __ballot(visit)) trips++;
  }
  if (kCount) {
    unsigned long long a = (unsigned long long)cnt.nodes, b = (unsigned long long)cnt.prims;
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_xor(a, off, 64);
      b += __shfl_xor(b, off, 64);
    }
    if (__lane_id() == 0) {
      atomicAdd(work + 0, a);
      atomicAdd(work + 1, b);
      atomicAdd(work + 2, trips);
    }
  }
}

__device__ __forceinline__ ShadeMat loadMat(const MatDev *mats, int m) {
  const MatDev md = mats[m];
  ShadeMat s;
  s.color = mkf3(md.cr, md.cg, md.cb);
  s.ax = md.ax;
  s.ay = md.ay;
  s.eta = md.eta;
  s.type = md.type;
  return s;
}


// Delta tracking of the heterogeneous medium along [t0, t1] (Woodcock):
// tentative collisions at exponential steps of the majorant, each real with
// probability density(x).  Draws: one per step, one per acceptance test.
__device__ __forceinline__ bool deltaTrack(const MediumDev &md, Rng &rng, f3 o, f3 d, float t0, float t1, f3 &x) {
  const float inv = 1.0f / (md.sigma * len3(d));
  float t = t0;
  for (int guard = 0; guard < kMaxCollisions; guard++) {
    t -= __logf(1.0f - rng.next()) * inv;
    if (!(t < t1)) return false;
    const f3 p = o + d * t;
    if (rng.next() < densityAt(md, p)) {
      x = p;
      return true;
    }
  }
  return false;
}

// Reads the path state of slot k (< count) from the *In arrays; survivors go
// compacted to the *Out arrays at the slot given by the next-queue counter.
// kMedium / kAov: the scene has a medium / AOVs are written (instantiated
// separately so the common case carries neither in registers).
template <bool kMedium, bool kAov>
__global__ __launch_bounds__(kRenderThreads) void k_shade(
    RenderArgs a, uint32_t count, const float4 *__restrict__ ray, const float4 *__restrict__ thr,
    const uint4 *__restrict__ meta, const float4 *__restrict__ hit, float4 *__restrict__ rayOut,
    float4 *__restrict__ thrOut, uint4 *__restrict__ metaOut, float4 *__restrict__ shadow, int64_t shadowCap,
    uint32_t *__restrict__ nextCount, uint32_t *__restrict__ bounceCnt, uint32_t *__restrict__ qVol) {
  __shared__ uint32_t appendSh[kRenderThreads / 64 + 1];
  const uint32_t k = blockIdx.x * kRenderThreads + threadIdx.x;
  const bool active = k < count;
  bool alive = false;
  const uint32_t p = k;
  float4 o4 = make_float4(0, 0, 0, 0), d4 = o4, t4 = o4;
  uint4 mt = make_uint4(0, 0, 0, 0);
  int leaf = -1;
  float th = 0.f;
  float4 h0 = make_float4(0, 0, 0, 0), h1 = h0;
  if (active) {
    o4 = ray[2 * (int64_t)p];
    d4 = ray[2 * (int64_t)p + 1];
    t4 = thr[p];
    mt = meta[p];
    h0 = hit[2 * (int64_t)p];
    h1 = hit[2 * (int64_t)p + 1];
    leaf = __float_as_int(h0.x);
    th = h0.y;
  }
  Rng rng;
  rng.init(a.seed, mt.x, mt.y, mt.z);
  uint32_t flags = mt.w;
  f3 T = mkf3(t4.x, t4.y, t4.z);
  f3 o = xyz(o4), d = xyz(d4);
  const int64_t ls = pathSlot(a, mt.x, mt.y);
  // the path's first-vertex AOV record (written once: the first event sets
  // kFlagAov or ends the path; bounce 0 writes zeros where it sets none)
  bool aovWritten = false;
  auto writeAov = [&](f3 alb3, f3 nrm, float depth) {
    a.apath[2 * ls] = make_float4(alb3.x, alb3.y, alb3.z, nrm.x);
    a.apath[2 * ls + 1] = make_float4(nrm.y, nrm.z, depth, 0.f);
    aovWritten = true;
  };

  // scattering vertex: a surface with a BSDF, or a real collision in the medium
  bool vtx = false, inMedium = false;
  f3 hp = mkf3(0, 0, 0), n = hp;
  ShadeMat m;
  m.type = CPT_M_NONE;
  bool front = true;
  if (active) {
    if (kMedium) {
      float t0, t1;
      if (clipBox(a.med, o, d, o4.w, leaf >= 0 ? th : d4.w, t0, t1) &&
          deltaTrack(a.med, rng, o, d, t0, t1, hp)) {
        vtx = true;
        inMedium = true;
      }
    }
    if (inMedium) {
    } else if (leaf < 0) {
      // miss: sky colour (Render.cu:138-142)
      const CamDev &cv = a.cams[mt.x / (uint32_t)a.npix];
      const f3 bg = mkf3(cv.bg[0], cv.bg[1], cv.bg[2]);
      addRadiance(a.lbuf, ls, T * bg);
      if (kAov && !(flags & kFlagAov)) writeAov(bg, mkf3(0.f, 0.f, 0.f), 0.f);
    } else {
      int mi = __float_as_int(h0.w);
      if ((unsigned)mi >= (unsigned)a.nMats) {
        atomicOr(a.err, kErrMaterial);
        mi = 0;
      }
      m = loadMat(a.mats, mi);
      hp = o + d * th;
      const f3 outw = xyz(h1);
      front = dot3(d, outw) < 0.0f;
      n = front ? outw : -outw;
      if (m.type == CPT_M_VOLUME) {
        // homogeneous volume object: walked by k_volume (it traces inside)
        const uint32_t vs = waveAppend(bounceCnt + 2, true);
        if ((int64_t)vs < a.cap) qVol[vs] = p;
        else atomicOr(a.err, kErrQueueCap);
      } else if (m.type == CPT_M_LIGHT) {
        // emitter (Render.cu:144-149)
        if (!(flags & kFlagNee)) {
          addRadiance(a.lbuf, ls, T * m.color);
        }
        // the reference's G-buffer (NORMAL / DEPTH display) keeps its cleared
        // value when the first hit is an emitter (Render.cu:144-149)
        if (kAov && !(flags & kFlagAov) && a.profile != CPT_PROFILE_NORMAL && a.profile != CPT_PROFILE_DEPTH)
          writeAov(mkf3(1.f, 1.f, 1.f), n, th * len3(d));
      } else {
        vtx = true;
      }
    }
  }
  const f3 alb = inMedium ? a.med.albedo : m.color;
  if (vtx) {
    if (kAov && !(flags & kFlagAov)) {
      const CamDev &cv = a.cams[mt.x / (uint32_t)a.npix];
      writeAov(alb, n, len3(mkf3((float)cv.lf[0], (float)cv.lf[1], (float)cv.lf[2]) - hp));
    }
    flags |= kFlagAov;
  }
  if (kAov && active && a.firstBounce && !aovWritten) writeAov(mkf3(0.f, 0.f, 0.f), mkf3(0.f, 0.f, 0.f), 0.f);
  const f3 ud = unit3(d);
  const f3 V = -ud;
  // NEE to every light (Render.cu:161-187): opaque-type surfaces (Render.cu:166)
  // and medium collisions.  blockAppend needs every thread of the workgroup,
  // so the loop runs on all threads with a per-lane predicate.
  const bool neeHere = vtx && (inMedium || m.type == CPT_M_OPAQUE || m.type == CPT_M_LAMBERT ||
                               m.type == CPT_M_GGX);
  const int nLoop = __builtin_amdgcn_readfirstlane(a.nLights);
  {
    for (int li = 0; li < nLoop; li++) {
      bool want = false;
      f3 dir = mkf3(0, 0, 0), c = dir;
      float tL = 0.f;
      if (neeHere) {
        const float4 l0 = a.rec[3 * li], l1 = a.rec[3 * li + 1], l2 = a.rec[3 * li + 2];
        const f3 lc = xyz(l0), ln = xyz(l1);
        const float lw = l2.x;
        // every reference profile uses the width for both extents
        // (Render.cu:170-171)
        const float lh = a.profile == CPT_PROFILE_CONFIG ? l2.y : l2.x;
        f3 eu, ev;
        lightFrame(ln, eu, ev);
        const float u1 = rng.next(), u2 = rng.next();
        const f3 lp = lc + eu * ((u1 - 0.5f) * lw) + ev * ((u2 - 0.5f) * lh);
        const f3 toL = lp - hp;
        dir = unit3(toL);
        const float cl = dot3(-dir, ln);
        // the shadow ray's own distance to the light (Light::hit,
        // Hittable.cuh:367-398): a ray that misses it in FP32 carries nothing
        if ((inMedium || dot3(dir, n) > 0.0f) && cl > 0.0f && l2.z != 0.0f &&
            hitQuad(lc, ln, l2.x, l2.y, true, hp, dir, kRayTMin, kRayTMax, tL)) {
          const ShadeMat lm = loadMat(a.mats, __float_as_int(l2.w));
          // surface: fr cos; medium: albedo times the isotropic phase 1/(4 pi)
          const f3 f = inMedium ? alb * (0.25f / kPi) : bsdfEval(m, V, dir, n, mkf3(0, 0, 0), front) * dot3(dir, n);
          c = T * f * lm.color * (cl / dot3(toL, toL) * (lw * lh));
          want = true;
        }
        flags |= kFlagNee;
      }
      const uint32_t slot = blockAppend(bounceCnt + 1, want, appendSh);
      if (want && (int64_t)slot >= shadowCap) atomicOr(a.err, kErrShadowCap);
      if (want && (int64_t)slot < shadowCap) {
        // {origin, tmin} {direction, distance to the light} {contribution,
        // radiance slot of the path} {pixel, sample and draw count of the
        // transmittance estimate's stream, light id}
        shadow[4 * (int64_t)slot + 0] = make_float4(hp.x, hp.y, hp.z, kRayTMin);
        shadow[4 * (int64_t)slot + 1] = make_float4(dir.x, dir.y, dir.z, tL);
        shadow[4 * (int64_t)slot + 2] = make_float4(c.x, c.y, c.z, __int_as_float((int)ls));
        shadow[4 * (int64_t)slot + 3] =
            make_float4(__uint_as_float(mt.x), __uint_as_float(mt.y), __uint_as_float(rng.cnt), __int_as_float(li));
        a.skeys[slot] = renderKey(a.rootBox, hp, dir) >> (24 - kShadowKeyBits);
      }
    }
  }
  if (vtx) {
    // Russian roulette / depth cap (Render.cu:189-204): STATIC rolls 0.7 at
    // every vertex with no cap; REAL_TIME ends after the 3rd vertex and
    // NORMAL / DEPTH after the 2nd (`depth++ >= 2`, `depth++ >= 1`) without
    // roulette; CONFIG caps at max_depth and rolls from rr_depth on
    const uint32_t depth = flags & kDepthMask;
    float prr = 1.0f;
    bool cont = true;
    if (a.profile == CPT_PROFILE_REFERENCE) {
      prr = 0.7f;
      if (rng.next() > kRoulette07) cont = false;
    } else if (a.profile == CPT_PROFILE_REAL_TIME) {
      cont = depth < 2u;
    } else if (a.profile == CPT_PROFILE_NORMAL || a.profile == CPT_PROFILE_DEPTH) {
      cont = depth < 1u;
    } else {
      if ((int)depth + 1 >= a.maxDepth) cont = false;
      else if ((int)depth >= a.rrDepth) {
        prr = a.rrProb;
        if (rng.next() > a.rrCmp) cont = false;
      }
    }
    if (cont) {
      flags = (flags & ~kDepthMask) | ((depth + 1) & kDepthMask);
      const float r1 = rng.next(), r2 = rng.next(), r3 = rng.next();
      if (inMedium) {
        // isotropic phase function: weight albedo (phase / pdf = 1)
        const float z = 1.0f - 2.0f * r1;
        const float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        float s, cph;
        __sincosf(2.0f * kPi * r2, &s, &cph);
        T = T * alb * (1.0f / prr);
        o = hp;
        d = mkf3(rr * cph, rr * s, z);
        alive = true;
      } else {
        f3 dir = mkf3(0, 0, 1), wm = mkf3(0, 0, 0);
        float pdf = -1.0f;
        bsdfSample(m, ud, n, front, r1, r2, r3, dir, pdf, wm);
        if (pdf > 0.0f) {
          const f3 fr = bsdfEval(m, V, unit3(dir), n, wm, front);
          T = T * fr * (fabsf(dot3(dir, n)) / pdf / prr);
          o = hp;
          d = dir;
          alive = true;
        }
      }
    }
  }
  const uint32_t slot = blockAppend(nextCount, alive, appendSh);
  if (alive && (int64_t)slot >= a.cap) {
    atomicOr(a.err, kErrQueueCap);
    alive = false;
  }
  if (alive) {
    rayOut[2 * (int64_t)slot] = make_float4(o.x, o.y, o.z, kRayTMin);
    rayOut[2 * (int64_t)slot + 1] = make_float4(d.x, d.y, d.z, kRayTMax);
    thrOut[slot] = make_float4(T.x, T.y, T.z, 0.f);
    metaOut[slot] = make_uint4(mt.x, mt.y, rng.cnt, flags);
    if (a.qkeys) a.qkeys[slot] = renderKey(a.rootBox, o, d) >> (24 - kQueueKeyBits);
  }
}

// Paths that hit a volume-material object (Render.cu:64-135): homogeneous
// medium with sigma_maj = 25, roulette 0.8 per collision, isotropic
// scattering, ending at the far side of the object; the path then continues
// from the last position with the last direction.  The queue length is
// *volCount; the grid covers an upper bound.  Continuing paths are appended
// to the next queue (*nextCount).
__global__ __launch_bounds__(kRenderThreads) void k_volume(RenderArgs a, const uint32_t *__restrict__ qVol,
                                                           uint32_t bound, const float4 *__restrict__ ray,
                                                           const float4 *__restrict__ thr,
                                                           const uint4 *__restrict__ meta,
                                                           const float4 *__restrict__ hit,
                                                           float4 *__restrict__ rayOut, float4 *__restrict__ thrOut,
                                                           uint4 *__restrict__ metaOut,
                                                           const uint32_t *__restrict__ volCount,
                                                           uint32_t *__restrict__ nextCount) {
  __shared__ int stack[kStackLds * kRenderThreads];
  const uint32_t k = blockIdx.x * kRenderThreads + threadIdx.x;
  if (k >= bound || k >= *volCount) return;
  const uint32_t p = qVol[k];
  if ((int64_t)p >= a.cap) {
    atomicOr(a.err, kErrSlot);
    return;
  }
  const float4 o4 = ray[2 * (int64_t)p], d4 = ray[2 * (int64_t)p + 1], t4 = thr[p];
  const uint4 mt = meta[p];
  const float th = hit[2 * (int64_t)p].y;
  Rng rng;
  rng.init(a.seed, mt.x, mt.y, mt.z);
  f3 T = mkf3(t4.x, t4.y, t4.z);
  f3 d = xyz(d4);
  const f3 hp = xyz(o4) + d * th;
  const float sig = 25.0f;
  TraceCounters cnt;
  float tb = kRayTMax;
  const int l2 = traceAny(a.wide, a.nodes, a.leaves, a.nPrims, hp, d, kRayTMin, tb, stack + threadIdx.x,
                                 kRenderThreads, cnt);
  f3 o = hp;
  if (l2 >= 0) {
    float dMax = tb;
    float t = -__logf(1.0f - rng.next()) / sig;
    f3 pos = hp + d * t;
    o = pos;
    for (int guard = 0; t < dMax && guard < kMaxCollisions; guard++) {
      if (rng.next() > kRoulette08) break;
      T = T * (__expf(-t * sig) / 0.8f);
      const float z = 1.0f - 2.0f * rng.next();
      const float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
      const float ph = 2.0f * kPi * rng.next();
      float sn, cs;
      __sincosf(ph, &sn, &cs);
      d = unit3(mkf3(rr * cs, rr * sn, z));
      o = pos;
      pos = hp + d * t;
      float tb2 = kRayTMax;
      const int l3 = traceAny(a.wide, a.nodes, a.leaves, a.nPrims, o, d, kRayTMin, tb2,
                                     stack + threadIdx.x, kRenderThreads, cnt);
      if (l3 >= 0) dMax = tb2;
      t = -__logf(1.0f - rng.next()) / sig;
      if (t > dMax || l3 < 0) break;
    }
  }
  const uint32_t slot = waveAppend(nextCount, true);
  if ((int64_t)slot >= a.cap) {
    atomicOr(a.err, kErrQueueCap);
    return;
  }
  rayOut[2 * (int64_t)slot] = make_float4(o.x, o.y, o.z, kRayTMin);
  rayOut[2 * (int64_t)slot + 1] = make_float4(d.x, d.y, d.z, kRayTMax);
  thrOut[slot] = make_float4(T.x, T.y, T.z, 0.f);
  metaOut[slot] = make_uint4(mt.x, mt.y, rng.cnt, mt.w);
  if (a.qkeys) a.qkeys[slot] = renderKey(a.rootBox, o, d) >> (24 - kQueueKeyBits);
}


// Shadow rays: the light's contribution counts when the light itself is the
// closest hit (Render.cu:180-185), times the medium's transmittance from a
// ratio-tracking estimate when a medium is present.  The closest hit is the
// light exactly when the ray meets the light at some tL and nothing else
// lies in [tmin, tL): an occlusion query (first hit ends the walk) over
// [tmin, prev(tL)] answers that (an exact tie at tL counts as unoccluded).
// k_shade computes tL (the FP32 Light::hit of the same ray) and emits only
// rays that meet their light.
// kAtomic: several lights, so several shadow rays of one path may add to
// its radiance in one launch.  perm (optional): lane i takes record perm[i].
// kCount: also sums the walks' record visits and primitive tests into
// work[0], work[1] and each wave's largest visit count into work[2].
template <bool kAtomic, int kFmt, bool kCount>
__global__ __launch_bounds__(kRenderThreads) void k_shadow(RenderArgs a, const float4 *__restrict__ shadow,
                                                            uint32_t ns, const uint32_t *__restrict__ perm,
                                                            unsigned long long *__restrict__ work) {
  const float4 *__restrict__ nodes = a.nodes;
  const float4 *__restrict__ leaves = a.leaves;
  const int nPrims = a.nPrims;
  const MediumDev &med = a.med;
  const uint64_t seed = a.seed;
  __shared__ int stack[kStackLds * kRenderThreads];
  __shared__ int pend[kPend * kRenderThreads];
  const uint32_t li = blockIdx.x * kRenderThreads + threadIdx.x;
  TraceCounters cnt;
  if (li < ns) {
    const uint32_t k = perm ? perm[li] : li;
    const float4 r0 = shadow[4 * (int64_t)k], r1 = shadow[4 * (int64_t)k + 1];
    const float4 c = shadow[4 * (int64_t)k + 2];
    const f3 o = xyz(r0), d = xyz(r1);
    const float tL = r1.w;
    float tb = nextafterf(tL, -INFINITY);
    int leaf;
    if ((kFmt == 1 || kFmt == 2) && nPrims > 1) {
      // four-wide walk with queued leaf tests (walk4Step)
      int ovf[kStack4Total - kStackLds];
      Walk4 w;
      walk4Init(w, o, d, r0.w, tb);
      bool visit = false;
      while (walk4Step<kCount, true, kFmt == 2, false, true>(w, nodes, leaves, stack + threadIdx.x, kRenderThreads,
                                                              ovf, cnt, pend + threadIdx.x, &visit)) {
      }
      leaf = w.hit;
    } else {
      leaf = traceFmt<kCount, kFmt, true>(nodes, leaves, nPrims, o, d, r0.w, tb, stack + threadIdx.x, kRenderThreads,
                                          cnt);
    }
    if (leaf < 0) {
      float tr = 1.0f;
      if (med.on) {
        const float4 l = shadow[4 * (int64_t)k + 3];
        tr = ratioTrack(med, seed, __float_as_uint(l.x), __float_as_uint(l.y), __float_as_uint(l.z), o, d, r0.w, tL);
      }
      const int64_t ls = __float_as_int(c.w);
      if (ls < 0 || ls >= a.cap) {
        atomicOr(a.err, kErrSlot);
      } else if (kAtomic) {
        float *pl = (float *)(a.lbuf + ls);
        atomicAdd(pl + 0, c.x * tr);
        atomicAdd(pl + 1, c.y * tr);
        atomicAdd(pl + 2, c.z * tr);
      } else {
        addRadiance(a.lbuf, ls, mkf3(c.x * tr, c.y * tr, c.z * tr));
      }
    }
  }
  if (kCount) {
    unsigned long long s0 = (unsigned long long)cnt.nodes, s1 = (unsigned long long)cnt.prims;
    int mx = cnt.nodes;
    for (int off = 32; off > 0; off >>= 1) {
      s0 += __shfl_xor(s0, off, 64);
      s1 += __shfl_xor(s1, off, 64);
      mx = max(mx, __shfl_xor(mx, off, 64));
    }
    if (__lane_id() == 0) {
      atomicAdd(work + 0, s0);
      atomicAdd(work + 1, s1);
      atomicAdd(work + 2, (unsigned long long)mx);
    }
  }
}

// Persistent k_shadow for the four-wide formats (kQuant: 64 B records): a
// fixed grid of waves whose lanes take the next shadow ray as soon as their
// occlusion walk is over (claimed in chunks; a ray's record is loaded one
// loop trip before its walk starts), so that lanes do not idle until the
// longest walk of their wave ends (the one-ray-per-lane kernel keeps 40% of
// its lanes busy on the displaced icosphere).  Same results as
// k_shadow<kAtomic, 1 or 2, kCount>; counting sums the walks and, in
// work[2], the loop trips in which the wave walked.  countDev: the number of
// shadow rays; grab: a zeroed counter.  kDefer (a medium is on): an
// unoccluded ray is only marked (its record's last word set to -1) and
// k_transmit ratio-tracks it afterwards with full waves, instead of one lane
// tracking while the rest of its wave waits.
template <bool kAtomic, bool kQuant, bool kPerm, bool kCount, bool kDefer>
__global__ __launch_bounds__(kRenderThreads) void k_shadow_persist(RenderArgs a, float4 *__restrict__ shadow,
                                                                    const uint32_t *__restrict__ countDev,
                                                                    const uint32_t *__restrict__ perm, uint32_t *grab,
                                                                    const uint32_t *abortFlag,
                                                                    unsigned long long *__restrict__ work) {
  const float4 *__restrict__ nodes = a.nodes;
  const float4 *__restrict__ leaves = a.leaves;
  const MediumDev &med = a.med;
  const uint64_t seed = a.seed;
  __shared__ int stack[kStackLds * kRenderThreads];
  __shared__ int pend[kPend * kRenderThreads];
  int *stk = stack + threadIdx.x;
  int *pnd = pend + threadIdx.x;
  int ovf[kStack4Total - kStackLds];
  Walk4 w;
  walk4Init(w, mkf3(0.f, 0.f, 0.f), mkf3(1.f, 1.f, 1.f), 0.f, 0.f);
  TraceCounters cnt;
  uint32_t k = 0;
  float tL = 0.f;
  float4 c = make_float4(0.f, 0.f, 0.f, 0.f);  // contribution, radiance slot (bits)
  int st = kIdle;
  unsigned long long trips = 0;
  const uint32_t count = *countDev;
  uint32_t pollTrips = 0;
  WaveChunk ch;
  waveChunkInit(ch, grab);
  // record k into the (idle) walk state: the walk tests [tmin, prev(tL)]
  auto loadRay = [&]() {
    const float4 r0 = shadow[4 * (int64_t)k], r1 = shadow[4 * (int64_t)k + 1];
    c = shadow[4 * (int64_t)k + 2];
    w.o = xyz(r0);
    w.d = xyz(r1);
    w.tmin = r0.w;
    tL = r1.w;
    w.tbest = nextafterf(tL, -INFINITY);
  };
  while (true) {
    if (st == kLoading) {
      walk4Start(w);
      st = kActive;
    }
    if (kPerm && st == kClaimed) {
      loadRay();
      st = kLoading;
    }
    uint32_t item = 0;
    abortRelay(abortFlag, grab, pollTrips);
    if (waveClaim(ch, grab, count, st == kIdle, item)) {
      if (kPerm) {
        k = perm[item];
        st = kClaimed;
      } else {
        k = item;
        loadRay();
        st = kLoading;
      }
    }
    if (!ch.more && __ballot(st != kIdle) == 0) break;
    bool visit = false;
    if (st == kActive &&
        !walk4Step<kCount, true, kQuant, false, true>(w, nodes, leaves, stk, kRenderThreads, ovf, cnt, pnd, &visit)) {
      if (w.hit < 0 && kDefer) {
        shadow[4 * (int64_t)k + 3].w = __int_as_float(-1);
      } else if (w.hit < 0) {
        float tr = 1.0f;
        if (med.on) {
          const float4 l = shadow[4 * (int64_t)k + 3];
          tr = ratioTrack(med, seed, __float_as_uint(l.x), __float_as_uint(l.y), __float_as_uint(l.z), w.o, w.d,
                          w.tmin, tL);
        }
        const int64_t ls = __float_as_int(c.w);
        if (ls < 0 || ls >= a.cap) {
          atomicOr(a.err, kErrSlot);
        } else if (kAtomic) {
          float *pl = (float *)(a.lbuf + ls);
          atomicAdd(pl + 0, c.x * tr);
          atomicAdd(pl + 1, c.y * tr);
          atomicAdd(pl + 2, c.z * tr);
        } else {
          addRadiance(a.lbuf, ls, mkf3(c.x * tr, c.y * tr, c.z * tr));
        }
      }
      st = kIdle;
    }
    if (kCount && __ballot(visit)) trips++;
  }
  if (kCount) {
    unsigned long long s0 = (unsigned long long)cnt.nodes, s1 = (unsigned long long)cnt.prims;
    for (int off = 32; off > 0; off >>= 1) {
      s0 += __shfl_xor(s0, off, 64);
      s1 += __shfl_xor(s1, off, 64);
    }
    if (__lane_id() == 0) {
      atomicAdd(work + 0, s0);
      atomicAdd(work + 1, s1);
      atomicAdd(work + 2, trips);
    }
  }
}

// The medium transmittance of the shadow rays that k_shadow_persist<...,
// kDefer> found unoccluded (last word of the record -1; k_shade writes the
// light id, >= 0, there), times their contribution, added to the path's
// radiance slot as k_shadow adds it: the same numbers in the same slots.
template <bool kAtomic>
__global__ __launch_bounds__(kRenderThreads) void k_transmit(RenderArgs a, const float4 *__restrict__ shadow,
                                                              uint32_t ns) {
  const uint32_t k = blockIdx.x * kRenderThreads + threadIdx.x;
  if (k >= ns) return;
  const float4 l = shadow[4 * (int64_t)k + 3];
  if (__float_as_int(l.w) >= 0) return;
  const float4 r0 = shadow[4 * (int64_t)k], r1 = shadow[4 * (int64_t)k + 1], c = shadow[4 * (int64_t)k + 2];
  const float tr = ratioTrack(a.med, a.seed, __float_as_uint(l.x), __float_as_uint(l.y), __float_as_uint(l.z), xyz(r0),
                              xyz(r1), r0.w, r1.w);
  const int64_t ls = __float_as_int(c.w);
  if (ls < 0 || ls >= a.cap) {
    atomicOr(a.err, kErrSlot);
  } else if (kAtomic) {
    float *pl = (float *)(a.lbuf + ls);
    atomicAdd(pl + 0, c.x * tr);
    atomicAdd(pl + 1, c.y * tr);
    atomicAdd(pl + 2, c.z * tr);
  } else {
    addRadiance(a.lbuf, ls, mkf3(c.x * tr, c.y * tr, c.z * tr));
  }
}

// k_shadow_persist with its template flags chosen at run time.
template <bool A, bool Q, bool P, bool C>
static void shadowPersistT(bool defer, dim3 grid, hipStream_t s, const RenderArgs &a, float4 *shadow,
                           const uint32_t *countDev, const uint32_t *perm, uint32_t *grab, const uint32_t *abortFlag,
                           unsigned long long *work) {
  if (defer)
    hipLaunchKernelGGL((k_shadow_persist<A, Q, P, C, true>), grid, dim3(kRenderThreads), 0, s, a, shadow, countDev,
                       perm, grab, abortFlag, work);
  else
    hipLaunchKernelGGL((k_shadow_persist<A, Q, P, C, false>), grid, dim3(kRenderThreads), 0, s, a, shadow, countDev,
                       perm, grab, abortFlag, work);
}

static void launchShadowPersist(bool atomic, bool quant, bool perm, bool count, bool defer, dim3 grid,
                                hipStream_t s, const RenderArgs &a, float4 *shadow, const uint32_t *countDev,
                                const uint32_t *permDev, uint32_t *grab, const uint32_t *abortFlag,
                                unsigned long long *work) {
#define CPT_SHP(A, Q, P, C) shadowPersistT<A, Q, P, C>(defer, grid, s, a, shadow, countDev, permDev, grab, abortFlag, work)
  switch ((count ? 8 : 0) + (atomic ? 4 : 0) + (quant ? 2 : 0) + (perm ? 1 : 0)) {
    case 0: CPT_SHP(false, false, false, false); break;
    case 1: CPT_SHP(false, false, true, false); break;
    case 2: CPT_SHP(false, true, false, false); break;
    case 3: CPT_SHP(false, true, true, false); break;
    case 4: CPT_SHP(true, false, false, false); break;
    case 5: CPT_SHP(true, false, true, false); break;
    case 6: CPT_SHP(true, true, false, false); break;
    case 7: CPT_SHP(true, true, true, false); break;
    case 8: CPT_SHP(false, false, false, true); break;
    case 9: CPT_SHP(false, false, true, true); break;
    case 10: CPT_SHP(false, true, false, true); break;
    case 11: CPT_SHP(false, true, true, true); break;
    case 12: CPT_SHP(true, false, false, true); break;
    case 13: CPT_SHP(true, false, true, true); break;
    case 14: CPT_SHP(true, true, false, true); break;
    default: CPT_SHP(true, true, true, true); break;
  }
#undef CPT_SHP
}

// Adds the batch's path radiance to the frame: thread = pixel (over all
// views), summing its samples g = s * npixAll + pixel that lie in the batch.
__global__ void k_gather(const float4 *__restrict__ lbuf, int64_t base, int64_t count, int64_t npixAll,
                         float *__restrict__ fb) {
  const int64_t px = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (px >= npixAll) return;
  int64_t s = base / npixAll;
  if (s * npixAll + px < base) s++;
  float r = 0.f, g = 0.f, b = 0.f;
  for (int64_t gi = s * npixAll + px; gi < base + count; gi += npixAll) {
    const float4 l = lbuf[gi - base];
    r += l.x;
    g += l.y;
    b += l.z;
  }
  fb[3 * px + 0] += r;
  fb[3 * px + 1] += g;
  fb[3 * px + 2] += b;
}

// The batch's AOV records summed per pixel in sample order into aov
// (npixAll x 7), as k_gather does the radiance.
__global__ void k_gather_aov(const float4 *__restrict__ apath, int64_t base, int64_t count, int64_t npixAll,
                             float *__restrict__ aov) {
  const int64_t px = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (px >= npixAll) return;
  int64_t s = base / npixAll;
  if (s * npixAll + px < base) s++;
  float v[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int64_t gi = s * npixAll + px; gi < base + count; gi += npixAll) {
    const float4 p0 = apath[2 * (gi - base)], p1 = apath[2 * (gi - base) + 1];
    v[0] += p0.x;
    v[1] += p0.y;
    v[2] += p0.z;
    v[3] += p0.w;
    v[4] += p1.x;
    v[5] += p1.y;
    v[6] += p1.z;
  }
#pragma unroll
  for (int c = 0; c < 7; c++) aov[7 * px + c] += v[c];
}

// CPT_RENDER_PACK_AOV with every sample in one batch (paths 0 .. count-1):
// per pixel, the radiance and the AOVs of its samples summed in sample order
// and scaled, written in the packed layout -- k_gather, k_gather_aov and
// k_finalize_pack in one pass, with the same sums.
__global__ void k_gather_pack(const float4 *__restrict__ lbuf, const float4 *__restrict__ apath, int64_t count,
                              int64_t npixAll, float *__restrict__ out, float scale) {
  const int64_t px = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (px >= npixAll) return;
  float v[10] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int64_t gi = px; gi < count; gi += npixAll) {
    const float4 l = lbuf[gi], p0 = apath[2 * gi], p1 = apath[2 * gi + 1];
    v[0] += l.x;
    v[1] += l.y;
    v[2] += l.z;
    v[3] += p0.x;
    v[4] += p0.y;
    v[5] += p0.z;
    v[6] += p0.w;
    v[7] += p1.x;
    v[8] += p1.y;
    v[9] += p1.z;
  }
  float2 *o = (float2 *)(out + 10 * px);  // 40 B records: 8 B aligned
#pragma unroll
  for (int c = 0; c < 5; c++) o[c] = make_float2(v[2 * c] * scale, v[2 * c + 1] * scale);
}

__global__ void k_finalize(const float *__restrict__ acc, float *__restrict__ out, int64_t n, float scale) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = acc[i] * scale;
}

// CPT_RENDER_PACK_AOV: out[10 px + c] = the means of radiance (c < 3) and of
// the AOVs (3 <= c < 10); i over npix * 10
__global__ void k_finalize_pack(const float *__restrict__ rgb, const float *__restrict__ aov, float *__restrict__ out,
                                int64_t n, float scale) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t px = i / 10;
  const int c = (int)(i - px * 10);
  out[i] = (c < 3 ? rgb[3 * px + c] : aov[7 * px + c - 3]) * scale;
}

// NORMAL / DEPTH display (Render.cu:239-251) from the first-vertex AOV of a
// one-sample render: (n + 0.5) / 2, or clamp01(distance / 50) with no
// vertex (distance 0) showing the cleared G-buffer's INF as 1.
__global__ void k_display(const float *__restrict__ aov, float *__restrict__ out, int64_t npix, int depth) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  const float *p = aov + 7 * i;
  if (depth) {
    const float v = p[6] > 0.0f ? fminf(fmaxf(p[6] / 50.0f, 0.0f), 1.0f) : 1.0f;
    out[3 * i] = v;
    out[3 * i + 1] = v;
    out[3 * i + 2] = v;
  } else {
    out[3 * i] = (p[3] + 0.5f) * 0.5f;
    out[3 * i + 1] = (p[4] + 0.5f) * 0.5f;
    out[3 * i + 2] = (p[5] + 0.5f) * 0.5f;
  }
}

void renderRelease(cpt_ctx *ctx) {
  if (!ctx->render) return;
  RenderBufs *r = ctx->render;
  DevBuf *all[] = {&r->ray[0], &r->ray[1], &r->thr[0],  &r->thr[1],    &r->meta[0], &r->meta[1],
                   &r->hit,    &r->permA,  &r->permB,   &r->qVol,      &r->keysA,   &r->keysB,   &r->keysS,
                   &r->shadow, &r->counters, &r->fb,    &r->aov,       &r->cams,    &r->lbuf,    &r->stage,
                   &r->apath};
  // release path: nothing to report to
  for (DevBuf *b : all)
    if (b->p) (void)hipFree(b->p);
  if (r->hostCnt) (void)hipHostFree(r->hostCnt);
  for (auto &row : r->ev)
    for (hipEvent_t &e : row)
      if (e) (void)hipEventDestroy(e);
  delete r;
  ctx->render = nullptr;
}

static unsigned blocksFor(int64_t n) { return (unsigned)((n + kRenderThreads - 1) / kRenderThreads); }

}  // namespace cpt

using namespace cpt;

namespace cpt {

// R.counters, u32 words: two per-bounce blocks, block (bounce & 1) holding
// [0] the length of the queue that bounce traces, [1] its shadow rays, [2]
// its volume-object paths, [3] the grab counter of its persistent
// extension kernel, [4] that of its persistent occlusion kernel; word 16
// the error bits (kErr*); words 32..43 six u64
// work sums of a counting render (extension walks, then shadow walks).
constexpr int kCntBlock = 8, kCntErr = 16, kCntWork = 32, kCntReadback = 20;

__global__ void k_init_counters(uint32_t *c, uint32_t count) {
  const int t = (int)threadIdx.x;
  if (t < 2 * kCntBlock) c[t] = t == 0 ? count : 0u;
}

static const char *renderErrText(uint32_t e) {
  if (e & kErrShadowCap) return "shadow queue overflow";
  if (e & kErrQueueCap) return "path queue overflow";
  if (e & kErrMaterial) return "material index outside the table";
  return "radiance slot or queue entry outside the batch";
}

}  // namespace cpt

extern "C" int cpt_render_views(cpt_ctx *ctx, const cpt_camera *cams, int32_t nviews, const cpt_render_params *prm,
                                float *rgb, float *aovOut, cpt_render_stats *stats, int flags) {
  if (!ctx || !cams || !prm || !rgb || nviews <= 0) return CPT_ERR_ARG;
  if (!ctx->built) return fail(ctx, CPT_ERR_STATE, "cpt_render", "no BVH built");
  if (ctx->abortHost && __atomic_load_n(ctx->abortHost, __ATOMIC_ACQUIRE))
    return fail(ctx, CPT_ERR_ABORTED, "cpt_render", "aborted by cpt_request_abort");
  const cpt_camera *cam = cams;
  if (cam->width <= 0 || cam->height <= 0 || prm->spp <= 0 || prm->sample_base < 0)
    return fail(ctx, CPT_ERR_ARG, "cpt_render", "bad image size or spp");
  if (prm->profile < CPT_PROFILE_REFERENCE || prm->profile > CPT_PROFILE_DEPTH)
    return fail(ctx, CPT_ERR_ARG, "cpt_render", "unknown profile");
  for (int v = 1; v < nviews; v++)
    if (cams[v].width != cam->width || cams[v].height != cam->height)
      return fail(ctx, CPT_ERR_ARG, "cpt_render_views", "views differ in size");
  if (prm->profile == CPT_PROFILE_CONFIG && (prm->max_depth <= 0 || !(prm->rr_prob > 0.0)))
    return fail(ctx, CPT_ERR_ARG, "cpt_render", "max_depth and rr_prob must be positive");
  // NORMAL / DEPTH render one sample (Render.cu:38) and show its first vertex
  const bool display = prm->profile == CPT_PROFILE_NORMAL || prm->profile == CPT_PROFILE_DEPTH;
  const int spp = display ? 1 : prm->spp;
  const bool packAov = (prm->flags & CPT_RENDER_PACK_AOV) != 0;
  if (packAov && (aovOut || !prm->want_aov || display))
    return fail(ctx, CPT_ERR_ARG, "cpt_render", "CPT_RENDER_PACK_AOV needs want_aov, aov == NULL and a radiance profile");
  const bool wantAov = aovOut != nullptr || display || packAov;
  const int64_t npixView = (int64_t)cam->width * cam->height;
  const int64_t npix = npixView * nviews;  // pixels over all views
  if (npix > (int64_t)1 << 31) return fail(ctx, CPT_ERR_ARG, "cpt_render", "image too large");
  CPT_TRY_HIP(ctx, hipSetDevice(ctx->device));
  if (!ctx->render) ctx->render = new RenderBufs();
  RenderBufs &R = *ctx->render;
  const int64_t total = npix * spp;
  const int nl = std::max(1, (int)ctx->nLights);
  int64_t cap = prm->paths_per_batch > 0 ? prm->paths_per_batch : ((int64_t)1 << 28);
  cap = std::min<int64_t>(cap, std::max<int64_t>(65536, ((int64_t)1 << 28) / nl));
  cap = std::min(cap, total);
  // whole samples per batch where a batch holds at least one (k_gen then
  // orders each sample's pixels in tiles), spread evenly over the fewest
  // batches so that no batch is a small remainder
  if (cap >= npix) {
    const int64_t nb = (spp + cap / npix - 1) / (cap / npix);
    cap = (spp + nb - 1) / nb * npix;
  }
  const int64_t shadowCap = cap * nl;
  int rc;
  for (int b = 0; b < 2; b++)
    if ((rc = ensure(ctx, R.ray[b], cap * 32)) || (rc = ensure(ctx, R.thr[b], cap * 16)) ||
        (rc = ensure(ctx, R.meta[b], cap * 16)))
      return rc;
  // sort keys and permutations serve the bounce queue and the shadow queue
  if ((rc = ensure(ctx, R.hit, cap * 32)) || (rc = ensure(ctx, R.permA, shadowCap * 4)) ||
      (rc = ensure(ctx, R.permB, shadowCap * 4)) || (rc = ensure(ctx, R.qVol, cap * 4)) ||
      (rc = ensure(ctx, R.keysA, shadowCap * 4)) || (rc = ensure(ctx, R.keysB, shadowCap * 4)) ||
      (rc = ensure(ctx, R.keysS, shadowCap * 4)) ||
      (rc = ensure(ctx, R.shadow, shadowCap * 64)) || (rc = ensure(ctx, R.counters, 256)) ||
      (rc = ensure(ctx, R.cams, (size_t)nviews * sizeof(CamDev))) || (rc = ensure(ctx, R.lbuf, cap * 16)) ||
      (rc = ensure(ctx, R.fb, npix * 12)) || (rc = ensure(ctx, R.aov, npix * 28)))
    return rc;
  if (!(flags & CPT_DEVICE_PTRS) && (rc = ensure(ctx, R.stage, npix * 40))) return rc;
  if (wantAov && (rc = ensure(ctx, R.apath, cap * 32))) return rc;
  if (!R.hostCnt) CPT_TRY_HIP(ctx, hipHostMalloc((void **)&R.hostCnt, 2 * 32 * sizeof(uint32_t), hipHostMallocDefault));
  for (auto &row : R.ev)
    for (hipEvent_t &e : row)
      if (!e) CPT_TRY_HIP(ctx, hipEventCreate(&e));
  R.cap = cap;
  R.shadowCap = shadowCap;
  hipStream_t s = ctx->stream;

  RenderArgs a;
  // 0: binary nodes, 1: four-wide FP32 records, 2: four-wide quantized
  // records, 3: eight-wide quantized records
  if (ctx->cluster == 2 && !ctx->clusterValid) {
    float cms = 0.f;
    if ((rc = clusterStage(ctx, &cms))) return rc;
  }
  if ((rc = ensureRecords(ctx))) return rc;
  a.wide = recordFormat(ctx);
  a.nodes = a.wide == 3   ? ctx->nodes8.as<float4>()
            : a.wide == 2 ? ctx->nodes4q.as<float4>()
            : a.wide == 1 ? ctx->nodes4.as<float4>()
                          : ctx->nodes.as<float4>();
  a.leaves = ctx->leaves.as<float4>();
  a.rec = ctx->rec.as<float4>();
  a.mats = ctx->mats.as<MatDev>();
  a.nPrims = (int)ctx->nPrims;
  a.nLights = ctx->nLights;
  {
    std::vector<CamDev> cd(nviews);
    for (int v = 0; v < nviews; v++) {
      for (int k = 0; k < 3; k++) {
        cd[v].lf[k] = cams[v].look_from[k];
        cd[v].pc[k] = cams[v].pixel_center[k];
        cd[v].dh[k] = cams[v].delta_h[k];
        cd[v].dv[k] = cams[v].delta_v[k];
        cd[v].bg[k] = (float)cams[v].background[k];
      }
      cd[v].w = cams[v].width;
      cd[v].h = cams[v].height;
    }
    CPT_TRY_HIP(ctx, hipMemcpyAsync(R.cams.p, cd.data(), nviews * sizeof(CamDev), hipMemcpyHostToDevice, s));
    CPT_TRY_HIP(ctx, hipStreamSynchronize(s));
  }
  a.cams = R.cams.as<CamDev>();
  a.nviews = nviews;
  a.seed = prm->seed;
  a.spp = spp;
  a.sampleBase = prm->sample_base;
  a.profile = prm->profile;
  a.maxDepth = prm->max_depth;
  a.rrDepth = prm->rr_depth;
  a.rrProb = (float)prm->rr_prob;
  a.rrCmp = (float)prm->rr_prob;
  if ((double)a.rrCmp > prm->rr_prob) a.rrCmp = std::nextafter(a.rrCmp, -INFINITY);
  a.wantAov = wantAov ? 1 : 0;
  a.npix = npixView;
  a.tiled = cap % npix == 0 ? 1 : 0;
  a.lbuf = R.lbuf.as<float4>();
  a.apath = wantAov ? R.apath.as<float4>() : nullptr;
  a.firstBounce = 0;
  a.base = 0;
  a.rootBox = ctx->rootBox.as<float>();
  // bounce-queue keys only where a queue may be reordered
  a.qkeys = ctx->queueSortMin > 0 ? R.keysA.as<uint32_t>() : nullptr;
  a.skeys = R.keysS.as<uint32_t>();
  a.med.on = ctx->hasMedium ? 1 : 0;
  a.med.density = ctx->medDensity.as<float>();
  a.med.bmin = mkf3(ctx->medBox[0], ctx->medBox[1], ctx->medBox[2]);
  a.med.bmax = mkf3(ctx->medBox[3], ctx->medBox[4], ctx->medBox[5]);
  a.med.scale = mkf3((float)ctx->medN[0] / (ctx->medBox[3] - ctx->medBox[0]),
                     (float)ctx->medN[1] / (ctx->medBox[4] - ctx->medBox[1]),
                     (float)ctx->medN[2] / (ctx->medBox[5] - ctx->medBox[2]));
  a.med.nx = ctx->medN[0];
  a.med.ny = ctx->medN[1];
  a.med.nz = ctx->medN[2];
  a.med.sigma = ctx->medSigma;
  a.med.albedo = mkf3(ctx->medAlbedo[0], ctx->medAlbedo[1], ctx->medAlbedo[2]);
  uint32_t *cnt = R.counters.as<uint32_t>();
  a.err = cnt + kCntErr;
  a.cap = cap;
  a.nMats = ctx->nMats + ctx->nLights + 1;

  // one batch with the packed output: the frame is written straight from the
  // per-path sums (no accumulation buffers)
  const bool fusedPack = packAov && cap >= total;
  const float inv = 1.0f / (float)spp;
  float *dstRgb = rgb, *dstAov = aovOut;
  if (!(flags & CPT_DEVICE_PTRS)) {
    dstRgb = R.stage.as<float>();
    dstAov = R.stage.as<float>() + npix * 3;
  }
  if (!fusedPack) {
    hipLaunchKernelGGL(k_fill, dim3(blocksFor(npix * 3)), dim3(kRenderThreads), 0, s, R.fb.as<float>(), npix * 3, 0.f);
    if (wantAov)
      hipLaunchKernelGGL(k_fill, dim3(blocksFor(npix * 7)), dim3(kRenderThreads), 0, s, R.aov.as<float>(), npix * 7,
                         0.f);
  }
  CPT_TRY_HIP(ctx, hipGetLastError());

  double rays = 0, extMs = 0, shMs = 0, sdMs = 0, extRays = 0;
  int bounces = 0;
  // four-wide formats trace bounce queues with the persistent extension
  // kernel, on one resident set of workgroups
  const bool extPersist = ctx->persistent && (a.wide == 1 || a.wide == 2) && a.nPrims > 1;
  unsigned extGrid = 0, shGrid = 0;
  if (extPersist) {
    int perCU = 0;
    CPT_TRY_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                         &perCU, (const void *)k_extend_persist<false, false, true>, kRenderThreads, 0));
    extGrid = (unsigned)ctx->numCU * (unsigned)std::max(1, perCU);
    perCU = 0;
    CPT_TRY_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                         &perCU, (const void *)k_shadow_persist<false, true, true, false, false>, kRenderThreads, 0));
    shGrid = (unsigned)ctx->numCU * (unsigned)std::max(1, perCU);
  }
  unsigned long long *work = (unsigned long long *)(cnt + kCntWork);
  const bool countWork = (prm->flags & CPT_RENDER_COUNT_WORK) != 0;
  // with bounce-queue ordering the host needs each queue's length before its
  // extension kernel; otherwise that kernel is enqueued before the length
  // reaches the host and reads it on the device
  const bool lagExtend = ctx->queueSortMin <= 0;
  const uint32_t *abortDev = ctx->abortDev;
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, cnt, 64);
  CPT_TRY_HIP(ctx, hipEventRecord(ctx->ev[12], s));

  // the extension kernel of bounce j over the queue in state set (j & 1);
  // bound: an upper bound of its length (queues never grow)
  auto launchExtend = [&](int j, uint32_t bound, const uint32_t *pm) -> int {
    const int pj = j & 1;
    const float4 *rayIn = R.ray[pj].as<float4>();
    const uint32_t *cd = cnt + pj * kCntBlock;
    CPT_TRY_HIP(ctx, hipEventRecord(R.ev[pj][0], s));
    if (bound > 0) {
#define CPT_EXT(C, F, P)                                                                                        \
  hipLaunchKernelGGL((k_extend<C, F, P>), dim3(blocksFor(bound)), dim3(kRenderThreads), 0, s, cd, rayIn, a.nodes, \
                     a.leaves, a.nPrims, R.hit.as<float4>(), work, a.wide, pm)
#define CPT_EXT_F(C, P)               \
  switch (a.wide) {                   \
    case 0: CPT_EXT(C, 0, P); break;  \
    case 1: CPT_EXT(C, 1, P); break;  \
    case 2: CPT_EXT(C, 2, P); break;  \
    default: CPT_EXT(C, 3, P); break; \
  }
#define CPT_EXT_P(C, Q, P)                                                                                       \
  hipLaunchKernelGGL((k_extend_persist<C, Q, P>), dim3(std::min<unsigned>(blocksFor(bound), extGrid)),             \
                     dim3(kRenderThreads), 0, s, cd, rayIn, a.nodes, a.leaves, R.hit.as<float4>(), work, pm,       \
                     cnt + pj * kCntBlock + 3, abortDev)
      if (extPersist) {
        const int q = a.wide == 2 ? 1 : 0;
        switch ((countWork ? 4 : 0) + q * 2 + (pm ? 1 : 0)) {
          case 0: CPT_EXT_P(false, false, false); break;
          case 1: CPT_EXT_P(false, false, true); break;
          case 2: CPT_EXT_P(false, true, false); break;
          case 3: CPT_EXT_P(false, true, true); break;
          case 4: CPT_EXT_P(true, false, false); break;
          case 5: CPT_EXT_P(true, false, true); break;
          case 6: CPT_EXT_P(true, true, false); break;
          default: CPT_EXT_P(true, true, true); break;
        }
      } else if (countWork) {
        if (pm) CPT_EXT_F(true, true) else CPT_EXT_F(true, false)
      } else {
        if (pm) CPT_EXT_F(false, true) else CPT_EXT_F(false, false)
      }
#undef CPT_EXT_P
#undef CPT_EXT_F
#undef CPT_EXT
    }
    CPT_TRY_HIP(ctx, hipGetLastError());
    CPT_TRY_HIP(ctx, hipEventRecord(R.ev[pj][1], s));
    return CPT_OK;
  };
  // shading (and the volume-object walk) of bounce j's qn paths, then the
  // read-back of the counters that tell the host the next queue's length
  auto launchShade = [&](int j, uint32_t qn) -> int {
    const int pj = j & 1, pn = pj ^ 1;
    a.firstBounce = j == 0 ? 1 : 0;
    const float4 *rayC = R.ray[pj].as<float4>(), *thrC = R.thr[pj].as<float4>();
    const uint4 *metaC = R.meta[pj].as<uint4>();
    uint32_t *nextCount = cnt + pn * kCntBlock, *bounceCnt = cnt + pj * kCntBlock;
    CPT_TRY_HIP(ctx, hipEventRecord(R.ev[pj][2], s));
#define CPT_SHADE(M, V)                                                                                       \
  hipLaunchKernelGGL((k_shade<M, V>), dim3(blocksFor(qn)), dim3(kRenderThreads), 0, s, a, qn, rayC, thrC, metaC, \
                     (const float4 *)R.hit.as<float4>(), R.ray[pn].as<float4>(), R.thr[pn].as<float4>(),       \
                     R.meta[pn].as<uint4>(), R.shadow.as<float4>(), shadowCap, nextCount, bounceCnt,           \
                     R.qVol.as<uint32_t>())
    switch ((a.med.on ? 2 : 0) + (a.wantAov ? 1 : 0)) {
      case 0: CPT_SHADE(false, false); break;
      case 1: CPT_SHADE(false, true); break;
      case 2: CPT_SHADE(true, false); break;
      default: CPT_SHADE(true, true); break;
    }
#undef CPT_SHADE
    if (ctx->hasVolumeMat)
      hipLaunchKernelGGL(k_volume, dim3(blocksFor(qn)), dim3(kRenderThreads), 0, s, a,
                         (const uint32_t *)R.qVol.as<uint32_t>(), qn, rayC, thrC, metaC,
                         (const float4 *)R.hit.as<float4>(), R.ray[pn].as<float4>(), R.thr[pn].as<float4>(),
                         R.meta[pn].as<uint4>(), (const uint32_t *)(bounceCnt + 2), nextCount);
    CPT_TRY_HIP(ctx, hipGetLastError());
    CPT_TRY_HIP(ctx, hipEventRecord(R.ev[pj][3], s));
    CPT_TRY_HIP(ctx, hipMemcpyAsync(R.hostCnt + 32 * pj, cnt, kCntReadback * sizeof(uint32_t),
                                    hipMemcpyDeviceToHost, s));
    CPT_TRY_HIP(ctx, hipEventRecord(R.ev[pj][6], s));
    return CPT_OK;
  };
  // a failure after work was enqueued: let it drain before returning
  auto stop = [&](int code, const char *what, const char *detail) -> int {
    (void)hipStreamSynchronize(s);
    return fail(ctx, code, what, detail);
  };

  for (int64_t base = 0; base < total; base += cap) {
    const int64_t count = std::min(cap, total - base);
    a.base = base;
    hipLaunchKernelGGL(k_gen, dim3(blocksFor(count)), dim3(kRenderThreads), 0, s, a, base, count,
                       R.ray[0].as<float4>(), R.thr[0].as<float4>(), R.meta[0].as<uint4>());
    hipLaunchKernelGGL(k_init_counters, dim3(1), dim3(64), 0, s, cnt, (uint32_t)count);
    CPT_TRY_HIP(ctx, hipGetLastError());
    uint32_t qn = (uint32_t)count;  // length of bounce i's queue (known on the host)
    if ((rc = launchExtend(0, qn, nullptr)) || (rc = launchShade(0, qn))) return stop(rc, "cpt_render", ctx->err.c_str());
    int last = 0;
    // paths end by roulette, depth cap, miss or absorption; the cap only
    // bounds pathological inputs (NaN geometry) so the loop always ends
    for (int i = 0; i < kMaxBounces; i++) {
      const int pi = i & 1, pn = pi ^ 1;
      last = i;
      // the next extension kernel goes in before bounce i's counts reach the
      // host, so the device has work while the host waits for them
      if (lagExtend && (rc = launchExtend(i + 1, qn, nullptr))) return stop(rc, "cpt_render", ctx->err.c_str());
      CPT_TRY_HIP(ctx, hipEventSynchronize(R.ev[pi][6]));
      const uint32_t *h = R.hostCnt + 32 * pi;
      const uint32_t qNext = h[pn * kCntBlock + 0], ns = h[pi * kCntBlock + 1], err = h[kCntErr];
      float e1 = 0, e2 = 0;
      CPT_TRY_HIP(ctx, hipEventElapsedTime(&e1, R.ev[pi][0], R.ev[pi][1]));
      CPT_TRY_HIP(ctx, hipEventElapsedTime(&e2, R.ev[pi][2], R.ev[pi][3]));
      extMs += e1;
      shMs += e2;
      if (i > 0) {  // bounce i-1's shadow pass preceded bounce i's shading
        float e3 = 0;
        CPT_TRY_HIP(ctx, hipEventElapsedTime(&e3, R.ev[pn][4], R.ev[pn][5]));
        sdMs += e3;
      }
      if (err) return stop(CPT_ERR_STATE, "cpt_render", renderErrText(err));
      if (ctx->abortHost && __atomic_load_n(ctx->abortHost, __ATOMIC_ACQUIRE))
        return stop(CPT_ERR_ABORTED, "cpt_render", "aborted by cpt_request_abort");
      if ((int64_t)ns > shadowCap || (int64_t)qNext > qn)
        return stop(CPT_ERR_STATE, "cpt_render", "queue counts beyond their bounds");
      rays += (double)qn + (double)ns;
      extRays += qn;
      bounces++;
      // bounce i's shadow rays, in the order of their origin cell (all rays
      // of one origin cell then walk the same records towards the lights)
      CPT_TRY_HIP(ctx, hipEventRecord(R.ev[pi][4], s));
      // the persistent walk refills its lanes in append order: on the
      // displaced icosphere it took 31.9 ms a frame unsorted, 40.9 sorted
      // (sort included), the one-lane-per-ray kernel 37.3 sorted, 47.7 not.
      // Below about 2^22 rays a wave's fixed chunks leave the chip part idle
      // and end in a tail (960x540 at 4 spp: 3.7 against 2.2 ms).  With a
      // medium it only marks the unoccluded rays and k_transmit ratio-tracks
      // them with full waves (tracking in the walk's lanes as each walk ended
      // left the rest of the wave waiting: slower than the one-lane kernel)
      const int64_t shMin = ctx->persistentShadowMin;
      const bool shPersist = extPersist && shMin > 0 && (int64_t)ns >= shMin;
      uint32_t *shPerm = nullptr;
      if (!shPersist && ctx->shadowSortMin > 0 && (int64_t)ns >= ctx->shadowSortMin) {
        if ((rc = radixSortPairs(ctx, R.keysS.as<uint32_t>(), R.permA.as<uint32_t>(), R.keysB.as<uint32_t>(),
                                 R.permB.as<uint32_t>(), ns, kShadowKeyBits, true, &shPerm)))
          return stop(rc, "cpt_render", ctx->err.c_str());
      }
#define CPT_SHD(A, F, C)                                                                                     \
  hipLaunchKernelGGL((k_shadow<A, F, C>), dim3(blocksFor(ns)), dim3(kRenderThreads), 0, s, a,                \
                     (const float4 *)R.shadow.as<float4>(), ns, shPerm, work + 3)
#define CPT_SHD_F(C)                       \
  switch (a.wide * 2 + (nl > 1 ? 1 : 0)) { \
    case 0: CPT_SHD(false, 0, C); break;   \
    case 1: CPT_SHD(true, 0, C); break;    \
    case 2: CPT_SHD(false, 1, C); break;   \
    case 3: CPT_SHD(true, 1, C); break;    \
    case 4: CPT_SHD(false, 2, C); break;   \
    case 5: CPT_SHD(true, 2, C); break;    \
    case 6: CPT_SHD(false, 3, C); break;   \
    default: CPT_SHD(true, 3, C); break;   \
  }
      if (ns > 0 && shPersist) {
        // the same lane-refilling walk as the extension kernel's
        launchShadowPersist(nl > 1, a.wide == 2, shPerm != nullptr, countWork, a.med.on != 0,
                            dim3(std::min<unsigned>(blocksFor(ns), shGrid)), s, a, R.shadow.as<float4>(),
                            cnt + pi * kCntBlock + 1, shPerm, cnt + pi * kCntBlock + 4, abortDev, work + 3);
        if (a.med.on) {
          if (nl > 1)
            hipLaunchKernelGGL(k_transmit<true>, dim3(blocksFor(ns)), dim3(kRenderThreads), 0, s, a,
                               (const float4 *)R.shadow.as<float4>(), ns);
          else
            hipLaunchKernelGGL(k_transmit<false>, dim3(blocksFor(ns)), dim3(kRenderThreads), 0, s, a,
                               (const float4 *)R.shadow.as<float4>(), ns);
        }
      } else if (ns > 0) {
        if (countWork) CPT_SHD_F(true) else CPT_SHD_F(false)
      }
#undef CPT_SHD_F
#undef CPT_SHD
      CPT_TRY_HIP(ctx, hipGetLastError());
      CPT_TRY_HIP(ctx, hipEventRecord(R.ev[pi][5], s));
      // block pi now serves bounce i + 2
      hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, cnt + pi * kCntBlock, kCntBlock);
      if (qNext == 0 || i + 1 >= kMaxBounces) break;
      if (!lagExtend) {
        // incoherent bounce rays: trace the next queue in ray-key order
        uint32_t *queuePerm = nullptr;
        if ((int64_t)qNext >= ctx->queueSortMin &&
            (rc = radixSortPairs(ctx, R.keysA.as<uint32_t>(), R.permA.as<uint32_t>(), R.keysB.as<uint32_t>(),
                                 R.permB.as<uint32_t>(), qNext, kQueueKeyBits, true, &queuePerm)))
          return stop(rc, "cpt_render", ctx->err.c_str());
        if ((rc = launchExtend(i + 1, qNext, queuePerm))) return stop(rc, "cpt_render", ctx->err.c_str());
      }
      if ((rc = launchShade(i + 1, qNext))) return stop(rc, "cpt_render", ctx->err.c_str());
      qn = qNext;
    }
    {  // the batch's last shadow pass
      float e3 = 0;
      CPT_TRY_HIP(ctx, hipEventSynchronize(R.ev[last & 1][5]));
      CPT_TRY_HIP(ctx, hipEventElapsedTime(&e3, R.ev[last & 1][4], R.ev[last & 1][5]));
      sdMs += e3;
    }
    if (fusedPack) {
      hipLaunchKernelGGL(k_gather_pack, dim3(blocksFor(npix)), dim3(kRenderThreads), 0, s,
                         (const float4 *)R.lbuf.as<float4>(), (const float4 *)R.apath.as<float4>(), count, npix,
                         dstRgb, inv);
    } else {
      hipLaunchKernelGGL(k_gather, dim3(blocksFor(npix)), dim3(kRenderThreads), 0, s,
                         (const float4 *)R.lbuf.as<float4>(), base, count, npix, R.fb.as<float>());
      if (wantAov)
        hipLaunchKernelGGL(k_gather_aov, dim3(blocksFor(npix)), dim3(kRenderThreads), 0, s,
                           (const float4 *)R.apath.as<float4>(), base, count, npix, R.aov.as<float>());
    }
    CPT_TRY_HIP(ctx, hipGetLastError());
  }
  // mean over samples (or the NORMAL / DEPTH display of the first vertex)
  if (fusedPack) {
  } else if (display)
    hipLaunchKernelGGL(k_display, dim3(blocksFor(npix)), dim3(kRenderThreads), 0, s, (const float *)R.aov.as<float>(),
                       dstRgb, npix, prm->profile == CPT_PROFILE_DEPTH ? 1 : 0);
  else if (packAov)
    hipLaunchKernelGGL(k_finalize_pack, dim3(blocksFor(npix * 10)), dim3(kRenderThreads), 0, s,
                       (const float *)R.fb.as<float>(), (const float *)R.aov.as<float>(), dstRgb, npix * 10, inv);
  else
    hipLaunchKernelGGL(k_finalize, dim3(blocksFor(npix * 3)), dim3(kRenderThreads), 0, s, R.fb.as<float>(), dstRgb,
                       npix * 3, inv);
  if (aovOut)
    hipLaunchKernelGGL(k_finalize, dim3(blocksFor(npix * 7)), dim3(kRenderThreads), 0, s, R.aov.as<float>(),
                       dstAov, npix * 7, inv);
  CPT_TRY_HIP(ctx, hipGetLastError());
  CPT_TRY_HIP(ctx, hipEventRecord(ctx->ev[10], s));
  hipError_t e = hipSuccess;
  if (!(flags & CPT_DEVICE_PTRS)) {
    e = hipMemcpyAsync(rgb, dstRgb, npix * (packAov ? 40 : 12), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && aovOut) e = hipMemcpyAsync(aovOut, dstAov, npix * 28, hipMemcpyDeviceToHost, s);
  }
  const hipError_t es = hipStreamSynchronize(s);
  if (e == hipSuccess) e = es;
  if (e != hipSuccess) return fail(ctx, CPT_ERR_HIP, "cpt_render", hipGetErrorString(e));
  if (stats) {
    float ms = 0;
    CPT_TRY_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[12], ctx->ev[10]));
    stats->ms = ms;
    stats->rays = rays;
    stats->paths = (double)total;
    stats->extend_ms = extMs;
    stats->shade_ms = shMs;
    stats->shadow_ms = sdMs;
    stats->bounces = bounces;
    stats->extend_rays = extRays;
    unsigned long long wk[6] = {0, 0, 0, 0, 0, 0};
    if (countWork) CPT_TRY_HIP(ctx, hipMemcpy(wk, work, 48, hipMemcpyDeviceToHost));
    stats->extend_nodes = (double)wk[0];
    stats->extend_prims = (double)wk[1];
    stats->extend_wave_nodes = 64.0 * (double)wk[2];
    stats->shadow_rays = rays - extRays;
    stats->shadow_nodes = (double)wk[3];
    stats->shadow_prims = (double)wk[4];
    stats->shadow_wave_nodes = 64.0 * (double)wk[5];
  }
  return CPT_OK;
}

extern "C" int cpt_render(cpt_ctx *ctx, const cpt_camera *cam, const cpt_render_params *prm, float *rgb,
                          float *aovOut, cpt_render_stats *stats, int flags) {
  return cpt_render_views(ctx, cam, 1, prm, rgb, aovOut, stats, flags);
}

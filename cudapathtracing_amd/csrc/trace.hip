// trace.hip - closest-hit traversal kernels (traverseIterative,
// CudaPathTracing/BVH.cuh:282-332, one ray per lane, FP32 box and primitive
// tests, per-lane stack in LDS with a private overflow tail), the FP64
// refinement of the winning hit distance, and the ray ordering of large
// batches (key = direction octant + origin Morton cell, LSD radix sort).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "cpt_device.hpp"
#include "internal.hpp"

namespace cpt {

constexpr int kTraceThreads = 256;

// Ray-order key: direction octant (3 bits) above the 27-bit Morton code of
// the origin on a 512^3 grid over the root box (origins outside clamp to
// the boundary cells).  Rays with equal keys start in the same cell and head
// into the same octant, so a wave of sorted rays walks similar subtrees.
__device__ __forceinline__ uint32_t spread9(uint32_t v) {
  v &= 0x1ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// q (optional): key i belongs to ray q[i] (a path queue).  Ray i is
// {origin, tmin} {direction, tmax} at rays[stride * i].
__global__ __launch_bounds__(kTraceThreads) void k_ray_keys(const float4 *__restrict__ rays, int stride,
                                                           const uint32_t *__restrict__ q, int64_t n,
                                                           const float *__restrict__ rootBox,
                                                           uint32_t *__restrict__ keys) {
  const int64_t k = (int64_t)blockIdx.x * kTraceThreads + threadIdx.x;
  if (k >= n) return;
  const int64_t i = q ? (int64_t)q[k] : k;
  const float4 r0 = rays[stride * i], r1 = rays[stride * i + 1];
  const float o[3] = {r0.x, r0.y, r0.z};
  uint32_t c[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float lo = rootBox[a], ext = rootBox[3 + a] - lo;
    const float u = ext > 0.0f ? (o[a] - lo) / ext : 0.0f;
    c[a] = (uint32_t)fminf(fmaxf(u * 512.0f, 0.0f), 511.0f);  // NaN -> 0
  }
  const uint32_t oct = (r1.x > 0.0f ? 4u : 0u) | (r1.y > 0.0f ? 2u : 0u) | (r1.z > 0.0f ? 1u : 0u);
  keys[k] = (oct << 27) | (spread9(c[0]) << 2) | (spread9(c[1]) << 1) | spread9(c[2]);
}

// perm (optional): lane i traces ray perm[i] and writes its outputs at perm[i].
// kFmt: 0 binary nodes, 1 four-wide FP32 records, 2 four-wide quantized
template <bool kCount, int kFmt>
__global__ __launch_bounds__(kTraceThreads) void k_trace(
    const float4 *__restrict__ rays, int64_t nRays, const uint32_t *__restrict__ perm,
    const float4 *__restrict__ nodes, const float4 *__restrict__ nodes4, const float4 *__restrict__ leaves,
    int nPrims, Scene64 s64, float *__restrict__ tOut, int32_t *__restrict__ idOut,
    int32_t *__restrict__ nodesOut, int32_t *__restrict__ primsOut) {
  __shared__ int stack[kStackLds * kTraceThreads];
  __shared__ uint16_t stackT[kStackLds * kTraceThreads];
  const int64_t li = (int64_t)blockIdx.x * kTraceThreads + threadIdx.x;
  if (li >= nRays) return;
  const int64_t i = perm ? (int64_t)perm[li] : li;
  const float4 r0 = rays[2 * i], r1 = rays[2 * i + 1];
  const f3 o = xyz(r0), d = xyz(r1);
  float tbest = r1.w;
  TraceCounters cnt;
  const int leaf =
      kFmt == 3 ? traceClosest8<kCount>(nodes4, leaves, nPrims, o, d, r0.w, tbest, stack + threadIdx.x, kTraceThreads,
                                        cnt)
      : kFmt == 2 ? traceClosest4<kCount, false, true, true>(nodes4, leaves, nPrims, o, d, r0.w, tbest,
                                                            stack + threadIdx.x, kTraceThreads, cnt,
                                                            stackT + threadIdx.x)
      : kFmt == 1 ? traceClosest4<kCount, false, false, true>(nodes4, leaves, nPrims, o, d, r0.w, tbest,
                                                             stack + threadIdx.x, kTraceThreads, cnt,
                                                             stackT + threadIdx.x)
          : traceClosest<kCount>(nodes, leaves, nPrims, o, d, r0.w, tbest, stack + threadIdx.x, kTraceThreads, cnt);
  if (kCount) {
    nodesOut[i] = cnt.nodes;
    primsOut[i] = cnt.prims;
    return;
  }
  int id = -1;
  float t = r1.w;
  if (leaf >= 0) {
    id = __float_as_int(leaves[3 * (int64_t)leaf].w);
    t = (float)refineT(s64, id, r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, r0.w, r1.w);
  }
  if (tOut) tOut[i] = t;
  if (idOut) idOut[i] = id;
}

// Persistent variant of k_trace over four-wide records: a fixed grid of
// waves; a lane whose walk is over takes the next ray, so lanes are not left
// waiting for the longest walk of their wave.
//  * A wave takes kChunk rays per grab of the global counter and issues the
//    grab of its next chunk as soon as it starts one, so no lane ever waits
//    for the counter.
//  * A lane takes its ray in stages, one loop trip each: claim (the
//    permutation entry is loaded), load (the ray is loaded), start (the walk
//    is set up).  The loads of one trip complete while the walk steps of the
//    same trip wait for their records.
//  * A finished lane stores the hit's primitive id (and, counting, its
//    counts); k_refine_t computes t in FP64 afterwards with full waves.
// Every wave leaves once a grab has passed nRays and its lanes are idle.

template <bool kCount, bool kQuant>
__global__ __launch_bounds__(kTraceThreads) void k_trace_persist(
    const float4 *__restrict__ rays, int64_t nRays, const uint32_t *__restrict__ perm,
    const float4 *__restrict__ nodes4, const float4 *__restrict__ leaves, int32_t *__restrict__ idOut,
    int32_t *__restrict__ nodesOut, int32_t *__restrict__ primsOut, uint32_t *__restrict__ work,
    const uint32_t *abortFlag) {
  __shared__ int stack[kStackLds * kTraceThreads];
  __shared__ uint16_t stackT[kStackLds * kTraceThreads];
  int *stk = stack + threadIdx.x;
  uint16_t *stkT = stackT + threadIdx.x;
  int ovf[kStack4Total - kStackLds];
  // a lane's ray is loaded straight into its (idle) walk state
  Walk4 w;
  walk4Init(w, mkf3(0.f, 0.f, 0.f), mkf3(1.f, 1.f, 1.f), 0.f, 0.f);
  TraceCounters cnt;
  int64_t ri = 0;
  int st = kIdle;
  uint32_t pollTrips = 0;
  WaveChunk ch;
  waveChunkInit(ch, work);
  while (true) {
    // rays loaded in the last trip start their walk
    if (st == kLoading) {
      walk4Start(w);
      cnt = TraceCounters();
      st = kActive;
    }
    // claimed rays whose permutation entry has arrived are loaded
    if (st == kClaimed) {
      walk4Ray(w, rays[2 * ri], rays[2 * ri + 1]);
      st = kLoading;
    }
    // idle lanes claim the next rays of the wave's chunk
    uint32_t item = 0;
    abortRelay(abortFlag, work, pollTrips);
    if (waveClaim(ch, work, (uint32_t)nRays, st == kIdle, item)) {
      if (perm) {
        ri = (int64_t)perm[item];
        st = kClaimed;
      } else {
        ri = (int64_t)item;
        walk4Ray(w, rays[2 * ri], rays[2 * ri + 1]);
        st = kLoading;
      }
    }
    if (!ch.more && __ballot(st != kIdle) == 0) break;
    // one record visit of every walking lane
    if (st == kActive &&
        !walk4Step<kCount, false, kQuant, false, false, true>(w, nodes4, leaves, stk, kTraceThreads, ovf, cnt, nullptr,
                                                               nullptr, stkT)) {
      st = kIdle;
      if (kCount) {
        nodesOut[ri] = cnt.nodes;
        primsOut[ri] = cnt.prims;
      } else {
        idOut[ri] = w.hit >= 0 ? __float_as_int(leaves[3 * (int64_t)w.hit].w) : -1;
      }
    }
  }
}

// t of every ray from its primitive id (FP64, as k_trace does); misses keep
// tmax.
__global__ __launch_bounds__(kTraceThreads) void k_refine_t(const float4 *__restrict__ rays, int64_t nRays,
                                                             Scene64 s64, const int32_t *__restrict__ id,
                                                             float *__restrict__ tOut) {
  const int64_t i = (int64_t)blockIdx.x * kTraceThreads + threadIdx.x;
  if (i >= nRays) return;
  const float4 r0 = rays[2 * i], r1 = rays[2 * i + 1];
  const int32_t h = id[i];
  tOut[i] = h >= 0 ? (float)refineT(s64, h, r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, r0.w, r1.w) : r1.w;
}

__global__ void k_zero_work(uint32_t *p) {
  if (threadIdx.x == 0) p[0] = 0;
}

int traceDevice(cpt_ctx *ctx, const float *raysDev, int64_t n, float *tDev, int32_t *idDev,
                int32_t *nodesDev, int32_t *primsDev, const TraceScratch *sc) {
  if (n <= 0) return CPT_OK;
  Scene64 s64{ctx->tri64.as<double>(), ctx->sph64.as<double>(), ctx->lgt64.as<double>(),
              ctx->nLights, ctx->nSpheres};
  const unsigned grid = (unsigned)((n + kTraceThreads - 1) / kTraceThreads);
  const float4 *r = (const float4 *)raysDev;
  const float4 *nd = ctx->nodes.as<float4>(), *lv = ctx->leaves.as<float4>();
  const int np = (int)ctx->nPrims;
  const bool sorted = ctx->raySortMin > 0 && n >= ctx->raySortMin && np > 1 && n <= 0x7fffffff;
  int rc;
  if (!sc && sorted && ((rc = ensure(ctx, ctx->rkA, n * 4)) || (rc = ensure(ctx, ctx->rkB, n * 4)) ||
                        (rc = ensure(ctx, ctx->rvA, n * 4)) || (rc = ensure(ctx, ctx->rvB, n * 4))))
    return rc;
  uint32_t *rkA = sc ? sc->keysA : ctx->rkA.as<uint32_t>(), *rkB = sc ? sc->keysB : ctx->rkB.as<uint32_t>();
  uint32_t *rvA = sc ? sc->valsA : ctx->rvA.as<uint32_t>(), *rvB = sc ? sc->valsB : ctx->rvB.as<uint32_t>();
  ctx->lastTraceSorted = sorted;
  CPT_TRY_HIP(ctx, hipEventRecord(ctx->ev[8], ctx->stream));
  const uint32_t *perm = nullptr;
  if (sorted) {
    hipLaunchKernelGGL(k_ray_keys, dim3(grid), dim3(kTraceThreads), 0, ctx->stream, r, 2, (const uint32_t *)nullptr, n,
                       (const float *)ctx->rootBox.as<float>(), rkA);
    CPT_TRY_HIP(ctx, hipGetLastError());
    if ((rc = radixSortPairs(ctx, rkA, rvA, rkB, rvB, n, 30, true, nullptr, sc ? sc->hist : nullptr,
                             sc ? sc->rowTotal : nullptr)))
      return rc;
    perm = rvA;
    CPT_TRY_HIP(ctx, hipEventRecord(ctx->ev[5], ctx->stream));
  }
  const int fmt = recordFormat(ctx);
  if (sc) {
    // no allocation: the records must already be current (a lazy rebuild
    // would allocate them)
    const bool stale = fmt != 0 && np > 1 &&
                       (!(ctx->recValid & (1 << fmt)) || (ctx->cluster == 1 && (fmt == 1 || fmt == 2) && !ctx->clusterValid));
    if (stale)
      return fail(ctx, CPT_ERR_STATE, "cpt_trace_closest_noalloc",
                  "traversal records are stale: build, refit or trace once with an allocating call");
  } else if ((rc = ensureRecords(ctx))) {
    return rc;
  }
  // read after ensureRecords, which may (re)allocate the records
  const float4 *rec4 = fmt == 3   ? ctx->nodes8.as<float4>()
                       : fmt == 2 ? ctx->nodes4q.as<float4>()
                                  : ctx->nodes4.as<float4>();
  if ((fmt == 1 || fmt == 2) && ctx->persistent && np > 1 && (nodesDev || (tDev && idDev)) &&
      n < (int64_t)kAbortGrab) {
    if (!sc && (rc = ensure(ctx, ctx->cntA, 256))) return rc;
    uint32_t *work = sc ? sc->grab : ctx->cntA.as<uint32_t>();
    hipLaunchKernelGGL(k_zero_work, dim3(1), dim3(64), 0, ctx->stream, work);
    // one resident wave set: as many workgroups as fit on every CU at once
    int perCU = 0;
    const void *kp = nodesDev ? (fmt == 2 ? (const void *)k_trace_persist<true, true> : (const void *)k_trace_persist<true, false>)
                              : (fmt == 2 ? (const void *)k_trace_persist<false, true> : (const void *)k_trace_persist<false, false>);
    CPT_TRY_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kp, kTraceThreads, 0));
    const unsigned pgrid = std::min<unsigned>(grid, (unsigned)ctx->numCU * (unsigned)std::max(1, perCU));
#define CPT_LAUNCH_P(C, Q)                                                                                         \
  hipLaunchKernelGGL((k_trace_persist<C, Q>), dim3(pgrid), dim3(kTraceThreads), 0, ctx->stream, r, n, perm, rec4, lv, \
                     idDev, nodesDev, primsDev, work, ctx->abortDev)
    if (nodesDev) {
      if (fmt == 2) CPT_LAUNCH_P(true, true); else CPT_LAUNCH_P(true, false);
    } else {
      if (fmt == 2) CPT_LAUNCH_P(false, true); else CPT_LAUNCH_P(false, false);
      CPT_TRY_HIP(ctx, hipGetLastError());
      hipLaunchKernelGGL(k_refine_t, dim3(grid), dim3(kTraceThreads), 0, ctx->stream, r, n, s64,
                         (const int32_t *)idDev, tDev);
    }
#undef CPT_LAUNCH_P
  } else {
#define CPT_LAUNCH_T(C, F)                                                                                        \
  hipLaunchKernelGGL((k_trace<C, F>), dim3(grid), dim3(kTraceThreads), 0, ctx->stream, r, n, perm, nd, rec4, lv, np, \
                     s64, tDev, idDev, nodesDev, primsDev)
    if (nodesDev) {
      if (fmt == 3) CPT_LAUNCH_T(true, 3);
      else if (fmt == 2) CPT_LAUNCH_T(true, 2);
      else if (fmt == 1) CPT_LAUNCH_T(true, 1);
      else CPT_LAUNCH_T(true, 0);
    } else {
      if (fmt == 3) CPT_LAUNCH_T(false, 3);
      else if (fmt == 2) CPT_LAUNCH_T(false, 2);
      else if (fmt == 1) CPT_LAUNCH_T(false, 1);
      else CPT_LAUNCH_T(false, 0);
    }
#undef CPT_LAUNCH_T
  }
  CPT_TRY_HIP(ctx, hipGetLastError());
  CPT_TRY_HIP(ctx, hipEventRecord(ctx->ev[9], ctx->stream));
  return CPT_OK;
}

// ---- scratch of the entry point that allocates nothing -------------------
// One static array per GPU (the code object's copy on each device), held
// until the process ends: ray-order keys and permutation ping-pong for
// CPT_NOALLOC_MAX_RAYS rays, radix histograms for as many 4096-key tiles, the
// row totals and the grab counter.  Host code uses the device address that
// hipGetSymbolAddress gives (cpt_ctx::noAllocScratch), never the symbol.
constexpr int64_t kNoAllocRays = CPT_NOALLOC_MAX_RAYS;
constexpr int64_t kNoAllocTiles = (kNoAllocRays + 4095) / 4096;
constexpr int64_t kNoAllocWords = 4 * kNoAllocRays + kNoAllocTiles * 256 + 256 + 64;
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

int noAllocScratchInit(cpt_ctx *ctx) {
  // one lookup per device (the context's device is current here)
  static std::mutex mu;
  static std::vector<uint32_t *> perDevice;
  std::lock_guard<std::mutex> lock(mu);
  if ((int)perDevice.size() <= ctx->device) perDevice.resize(ctx->device + 1, nullptr);
  if (!perDevice[ctx->device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      (void)hipGetLastError();
      return fail(ctx, CPT_ERR_HIP, "hipGetSymbolAddress(scratch)", hipGetErrorString(e));
    }
    perDevice[ctx->device] = (uint32_t *)ptr;
  }
  ctx->noAllocScratch = perDevice[ctx->device];
  return CPT_OK;
}

}  // namespace cpt

using namespace cpt;

extern "C" int cpt_trace_closest_noalloc(cpt_ctx *ctx, const float *rays, int64_t n, float *tOut, int32_t *idOut) {
  if (!ctx || n < 0 || (n > 0 && (!rays || !tOut || !idOut))) return CPT_ERR_ARG;
  if (n > kNoAllocRays)
    return fail(ctx, CPT_ERR_ARG, "cpt_trace_closest_noalloc", "more rays than CPT_NOALLOC_MAX_RAYS");
  if (!ctx->built) return fail(ctx, CPT_ERR_STATE, "cpt_trace_closest_noalloc", "no BVH built");
  if (!ctx->noAllocScratch) return fail(ctx, CPT_ERR_STATE, "cpt_trace_closest_noalloc", "scratch not set up");
  if (n == 0) return CPT_OK;
  if (ctx->abortHost && __atomic_load_n(ctx->abortHost, __ATOMIC_ACQUIRE))
    return fail(ctx, CPT_ERR_ABORTED, "cpt_trace_closest_noalloc", "aborted by cpt_request_abort");
  CPT_TRY_HIP(ctx, hipSetDevice(ctx->device));
  uint32_t *p = ctx->noAllocScratch;
  TraceScratch sc;
  sc.keysA = p;
  sc.keysB = p + kNoAllocRays;
  sc.valsA = p + 2 * kNoAllocRays;
  sc.valsB = p + 3 * kNoAllocRays;
  sc.hist = p + 4 * kNoAllocRays;
  sc.rowTotal = sc.hist + kNoAllocTiles * 256;
  sc.grab = sc.rowTotal + 256;  // cleared by traceDevice's k_zero_work before each walk
  return traceDevice(ctx, rays, n, tOut, idOut, nullptr, nullptr, &sc);
}

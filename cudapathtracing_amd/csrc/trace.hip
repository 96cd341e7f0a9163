// trace.hip - closest-hit traversal kernels (traverseIterative,
// CudaPathTracing/BVH.cuh:282-332, one ray per lane, FP32 box and primitive
// tests, per-lane stack in LDS with a private overflow tail) and the FP64
// refinement of the winning hit distance.
#include <hip/hip_runtime.h>

#include "cpt_device.hpp"
#include "internal.hpp"

namespace cpt {

constexpr int kTraceThreads = 256;

template <bool kCount, bool kWide>
__global__ __launch_bounds__(kTraceThreads) void k_trace(
    const float4 *__restrict__ rays, int64_t nRays, const float4 *__restrict__ nodes,
    const float4 *__restrict__ nodes4, const float4 *__restrict__ leaves, int nPrims, Scene64 s64,
    float *__restrict__ tOut, int32_t *__restrict__ idOut, int32_t *__restrict__ nodesOut,
    int32_t *__restrict__ primsOut) {
  __shared__ int stack[kStackLds * kTraceThreads];
  const int64_t i = (int64_t)blockIdx.x * kTraceThreads + threadIdx.x;
  if (i >= nRays) return;
  const float4 r0 = rays[2 * i], r1 = rays[2 * i + 1];
  const f3 o = xyz(r0), d = xyz(r1);
  float tbest = r1.w;
  TraceCounters cnt;
  const int leaf =
      kWide ? traceClosest4<kCount>(nodes4, leaves, nPrims, o, d, r0.w, tbest, stack + threadIdx.x, kTraceThreads, cnt)
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

int traceDevice(cpt_ctx *ctx, const float *raysDev, int64_t n, float *tDev, int32_t *idDev,
                int32_t *nodesDev, int32_t *primsDev) {
  if (n <= 0) return CPT_OK;
  Scene64 s64{ctx->tri64.as<double>(), ctx->sph64.as<double>(), ctx->lgt64.as<double>(),
              ctx->nLights, ctx->nSpheres};
  const unsigned grid = (unsigned)((n + kTraceThreads - 1) / kTraceThreads);
  const float4 *r = (const float4 *)raysDev;
  const float4 *nd = ctx->nodes.as<float4>(), *n4 = ctx->nodes4.as<float4>(), *lv = ctx->leaves.as<float4>();
  const int np = (int)ctx->nPrims;
  const bool wide = ctx->traversal == 4;
  CPT_TRY_HIP(ctx, hipEventRecord(ctx->ev[8], ctx->stream));
  if (nodesDev) {
    if (wide)
      hipLaunchKernelGGL((k_trace<true, true>), dim3(grid), dim3(kTraceThreads), 0, ctx->stream, r, n, nd, n4, lv, np,
                         s64, tDev, idDev, nodesDev, primsDev);
    else
      hipLaunchKernelGGL((k_trace<true, false>), dim3(grid), dim3(kTraceThreads), 0, ctx->stream, r, n, nd, n4, lv,
                         np, s64, tDev, idDev, nodesDev, primsDev);
  } else {
    if (wide)
      hipLaunchKernelGGL((k_trace<false, true>), dim3(grid), dim3(kTraceThreads), 0, ctx->stream, r, n, nd, n4, lv,
                         np, s64, tDev, idDev, nodesDev, primsDev);
    else
      hipLaunchKernelGGL((k_trace<false, false>), dim3(grid), dim3(kTraceThreads), 0, ctx->stream, r, n, nd, n4, lv,
                         np, s64, tDev, idDev, nodesDev, primsDev);
  }
  CPT_TRY_HIP(ctx, hipGetLastError());
  CPT_TRY_HIP(ctx, hipEventRecord(ctx->ev[9], ctx->stream));
  return CPT_OK;
}

}  // namespace cpt

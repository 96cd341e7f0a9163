// cluster.hip - a second binary tree over the LBVH's leaves for the wide
// traversal records: parallel locally-ordered clustering (PLOC, Meister and
// Bittner 2018) of the leaf boxes in Morton order.
//
// The LBVH of cpt_bvh_build stays the reference's tree (its topology is what
// cpt_bvh_export returns); this tree only decides which boxes the four-wide
// records group.  Its leaves are the same Morton-ordered leaf records
// (~k = leaf k), so a walk over either tree finds the same closest hit.
//
// One round over the n current clusters (Morton order kept):
//   k_ploc_nn     each cluster's nearest neighbour within kRadius positions,
//                 by the surface area of the union box; among equal areas
//                 the buddy position i ^ 1 first if its box equals the
//                 cluster's, then the lower index;
//   k_ploc_mark   per workgroup, how many clusters stay and how many merge;
//   k_ploc_scan   exclusive scan of those counts (one workgroup);
//   k_ploc_apply  mutual pairs i < j merge into a new internal node that
//                 takes position i; surviving clusters are compacted.
// Every cluster takes the candidate with the smallest (area, not coincident
// buddies, lower index, higher index) of their pair, an order that both ends
// of a pair see alike, so the globally smallest pair is mutual and every
// round merges at least one.  The buddy term makes coincident boxes
// (duplicated geometry) pair up as (2k, 2k+1): half of them merge per round
// instead of one pair.  It is kept to equal boxes: as a general tie-break
// (exact area ties are common between the congruent triangles of a
// tessellated sphere) it grouped leaves worse (35% more leaf tests).  Internal nodes are
// numbered downward from n-2 in merge order (within a round, by position),
// so the last merge, the root, is node 0 as in the LBVH.  Each cluster
// carries its subtree depth and leaf count; a tree deeper than 64 binary
// levels is not used (the four-wide walks' stacks are sized for 64).
// Finally the nodes are renumbered in depth-first preorder (a node's first
// child follows it, its second child follows the first child's subtree),
// top-down one level per launch (k_dfs_level), then written at their new
// positions with remapped child refs (k_dfs_remap): the walks then find a
// node's children near it in memory, as in the LBVH's Karras numbering.
#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "internal.hpp"

namespace cpt {

constexpr int kRadius = 16;
constexpr int kNnThreads = 256;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanBlock = kScanThreads * kScanItems;  // clusters per mark/apply workgroup
constexpr int kMaxClusterDepth = 64;
constexpr double kClusterMargin = 0.95;

// Boxes are two float4 per cluster: {min xyz, 0} {max xyz, 0}.
__device__ __forceinline__ float unionArea(float4 alo, float4 ahi, float4 blo, float4 bhi) {
  const float ex = fmaxf(ahi.x, bhi.x) - fminf(alo.x, blo.x);
  const float ey = fmaxf(ahi.y, bhi.y) - fminf(alo.y, blo.y);
  const float ez = fmaxf(ahi.z, bhi.z) - fminf(alo.z, blo.z);
  return ex * ey + ey * ez + ez * ex;
}

__global__ void k_ploc_init(int n, const uint32_t *__restrict__ order, const float *__restrict__ box32,
                            int32_t *__restrict__ C, float4 *__restrict__ B, int32_t *__restrict__ D,
                            int32_t *__restrict__ S) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t p = order[k];
  const float *b = box32 + 6 * p;
  C[k] = ~k;
  B[2 * (int64_t)k] = make_float4(b[0], b[1], b[2], 0.f);
  B[2 * (int64_t)k + 1] = make_float4(b[3], b[4], b[5], 0.f);
  D[k] = 0;
  S[k] = 1;
}

__global__ __launch_bounds__(kNnThreads) void k_ploc_nn(int n, const float4 *__restrict__ B,
                                                         int32_t *__restrict__ nn) {
  __shared__ float4 lo[kNnThreads + 2 * kRadius], hi[kNnThreads + 2 * kRadius];
  const int base = blockIdx.x * kNnThreads;
  for (int s = threadIdx.x; s < kNnThreads + 2 * kRadius; s += kNnThreads) {
    const int g = base - kRadius + s;
    if (g >= 0 && g < n) {
      lo[s] = B[2 * (int64_t)g];
      hi[s] = B[2 * (int64_t)g + 1];
    }
  }
  __syncthreads();
  const int i = base + (int)threadIdx.x;
  if (i >= n) return;
  const float4 ml = lo[threadIdx.x + kRadius], mh = hi[threadIdx.x + kRadius];
  float best = INFINITY;
  int bj = -1;
  bool bBuddy = false;
  // ascending j with strict comparisons keeps the lowest index among equal
  // (area, buddy) keys
  for (int o = -kRadius; o <= kRadius; o++) {
    const int j = i + o;
    if (o == 0 || j < 0 || j >= n) continue;
    const float a = unionArea(ml, mh, lo[threadIdx.x + kRadius + o], hi[threadIdx.x + kRadius + o]);
    const float4 jl = lo[threadIdx.x + kRadius + o], jh = hi[threadIdx.x + kRadius + o];
    const bool buddy = j == (i ^ 1) && jl.x == ml.x && jl.y == ml.y && jl.z == ml.z && jh.x == mh.x &&
                       jh.y == mh.y && jh.z == mh.z;
    if (bj < 0 || a < best || (a == best && buddy && !bBuddy)) {
      best = a;
      bj = j;
      bBuddy = buddy;
    }
  }
  nn[i] = bj;
}

// Stays / merges (as a new node at this position) for cluster i.
__device__ __forceinline__ void plocFlags(int i, int n, const int32_t *__restrict__ nn, uint32_t &stay,
                                          uint32_t &merge) {
  stay = 0;
  merge = 0;
  if (i >= n) return;
  const int j = nn[i];
  const bool mutual = j >= 0 && j < n && nn[j] == i;
  stay = (mutual && j < i) ? 0u : 1u;
  merge = (mutual && i < j) ? 1u : 0u;
}

// Workgroup-wide exclusive scan of (stay, merge) packed as stay << 32 | merge.
__device__ __forceinline__ uint64_t blockExclusive(uint64_t v, uint64_t *sh, uint64_t &total) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  uint64_t x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const uint64_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) sh[wave] = x;
  __syncthreads();
  uint64_t pre = 0;
  total = 0;
  for (int w = 0; w < kScanThreads / 64; w++) {
    if (w < wave) pre += sh[w];
    total += sh[w];
  }
  __syncthreads();
  return pre + x - v;
}

__global__ __launch_bounds__(kScanThreads) void k_ploc_mark(int n, const int32_t *__restrict__ nn,
                                                             uint64_t *__restrict__ blockSum) {
  __shared__ uint64_t sh[kScanThreads / 64];
  const int i0 = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
  uint64_t v = 0;
  for (int e = 0; e < kScanItems; e++) {
    uint32_t s, m;
    plocFlags(i0 + e, n, nn, s, m);
    v += ((uint64_t)s << 32) | m;
  }
  uint64_t total;
  blockExclusive(v, sh, total);
  if (threadIdx.x == 0) blockSum[blockIdx.x] = total;
}

// One workgroup: exclusive scan of nb workgroup sums in place; sums[nb] gets
// the grand total.
__global__ __launch_bounds__(kScanThreads) void k_ploc_scan(int nb, uint64_t *__restrict__ sums) {
  __shared__ uint64_t sh[kScanThreads / 64];
  uint64_t carry = 0;
  for (int c = 0; c < nb; c += kScanThreads) {
    const int k = c + (int)threadIdx.x;
    const uint64_t v = k < nb ? sums[k] : 0;
    uint64_t total;
    const uint64_t ex = blockExclusive(v, sh, total);
    if (k < nb) sums[k] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) sums[nb] = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_ploc_apply(
    int n, const int32_t *__restrict__ nn, const uint64_t *__restrict__ blockOff, const int32_t *__restrict__ C,
    const float4 *__restrict__ B, const int32_t *__restrict__ D, const int32_t *__restrict__ S,
    int32_t *__restrict__ Co, float4 *__restrict__ Bo, int32_t *__restrict__ Do, int32_t *__restrict__ So,
    float4 *__restrict__ nodes, int nodeBase) {
  __shared__ uint64_t sh[kScanThreads / 64];
  const int i0 = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
  uint32_t s[kScanItems], m[kScanItems];
  uint64_t v = 0;
  for (int e = 0; e < kScanItems; e++) {
    plocFlags(i0 + e, n, nn, s[e], m[e]);
    v += ((uint64_t)s[e] << 32) | m[e];
  }
  uint64_t total;
  uint64_t pre = blockExclusive(v, sh, total) + blockOff[blockIdx.x];
  for (int e = 0; e < kScanItems; e++) {
    const int i = i0 + e;
    const uint32_t pos = (uint32_t)(pre >> 32), rank = (uint32_t)pre;
    pre += ((uint64_t)s[e] << 32) | m[e];
    if (!s[e]) continue;
    if (m[e]) {
      const int j = nn[i];
      const int node = nodeBase - (int)rank;
      const float4 al = B[2 * (int64_t)i], ah = B[2 * (int64_t)i + 1];
      const float4 bl = B[2 * (int64_t)j], bh = B[2 * (int64_t)j + 1];
      float4 *o = nodes + 4 * (int64_t)node;
      o[0] = make_float4(al.x, ah.x, al.y, ah.y);
      o[1] = make_float4(bl.x, bh.x, bl.y, bh.y);
      o[2] = make_float4(al.z, ah.z, bl.z, bh.z);
      // child refs and child leaf counts
      o[3] = make_float4(__int_as_float(C[i]), __int_as_float(C[j]), __int_as_float(S[i]), __int_as_float(S[j]));
      Co[pos] = node;
      Bo[2 * (int64_t)pos] = make_float4(fminf(al.x, bl.x), fminf(al.y, bl.y), fminf(al.z, bl.z), 0.f);
      Bo[2 * (int64_t)pos + 1] = make_float4(fmaxf(ah.x, bh.x), fmaxf(ah.y, bh.y), fmaxf(ah.z, bh.z), 0.f);
      Do[pos] = max(D[i], D[j]) + 1;
      So[pos] = S[i] + S[j];
    } else {
      Co[pos] = C[i];
      Bo[2 * (int64_t)pos] = B[2 * (int64_t)i];
      Bo[2 * (int64_t)pos + 1] = B[2 * (int64_t)i + 1];
      Do[pos] = D[i];
      So[pos] = S[i];
    }
  }
}

// Appends cnt (0..2) entries for this lane with one atomic per wave; every
// lane of the wave calls it.  Returns the lane's first slot.
__device__ __forceinline__ uint32_t waveReserve(uint32_t *ctr, uint32_t cnt) {
  const int lane = threadIdx.x & 63;
  uint32_t x = cnt;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  uint32_t base = 0;
  if (lane == 63 && x > 0) base = atomicAdd(ctr, x);
  base = __shfl(base, 63, 64);
  return base + x - cnt;
}

// One level of the preorder numbering: frontier entry (old, new) gives node
// old its new index and appends its internal children with theirs.
__global__ void k_dfs_level(uint32_t cnt, const int2 *__restrict__ fin, const float4 *__restrict__ nodes,
                            int32_t *__restrict__ map, int2 *__restrict__ fout, uint32_t *__restrict__ fcount) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  int a = -1, b = -1, X = 0, sa = 0;
  if (k < cnt) {
    const int2 e = fin[k];
    X = e.y;
    map[e.x] = X;
    const float4 w = nodes[4 * (int64_t)e.x + 3];
    a = __float_as_int(w.x);
    b = __float_as_int(w.y);
    sa = __float_as_int(w.z);
  }
  const uint32_t slot = waveReserve(fcount, (a >= 0 ? 1u : 0u) + (b >= 0 ? 1u : 0u));
  uint32_t q = slot;
  if (a >= 0) fout[q++] = make_int2(a, X + 1);
  if (b >= 0) fout[q] = make_int2(b, X + sa);  // after the sa - 1 internal nodes of a's subtree
}

__global__ void k_dfs_remap(int nInt, const float4 *__restrict__ nodes, const int32_t *__restrict__ map,
                            float4 *__restrict__ out) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nInt) return;
  const int64_t X = map[o];
  const float4 w3 = nodes[4 * (int64_t)o + 3];
  const int a = __float_as_int(w3.x), b = __float_as_int(w3.y);
  out[4 * X] = nodes[4 * (int64_t)o];
  out[4 * X + 1] = nodes[4 * (int64_t)o + 1];
  out[4 * X + 2] = nodes[4 * (int64_t)o + 2];
  out[4 * X + 3] = make_float4(__int_as_float(a >= 0 ? map[a] : a), __int_as_float(b >= 0 ? map[b] : b), w3.z, w3.w);
}

// Per-workgroup sums of the surface areas of both child boxes of every
// internal node (binary node layout): the surface-area cost of the tree up to
// the root's area, i.e. the expected number of child-box tests of a random
// line that crosses the root box.
__global__ __launch_bounds__(256) void k_sah_partial(int nInt, const float4 *__restrict__ nodes,
                                                     double *__restrict__ partial) {
  __shared__ double sh[4];
  const int o = blockIdx.x * 256 + threadIdx.x;
  double a = 0.0;
  if (o < nInt) {
    const float4 w0 = nodes[4 * (int64_t)o], w1 = nodes[4 * (int64_t)o + 1], w2 = nodes[4 * (int64_t)o + 2];
    const double ax = (double)w0.y - w0.x, ay = (double)w0.w - w0.z, az = (double)w2.y - w2.x;
    const double bx = (double)w1.y - w1.x, by = (double)w1.w - w1.z, bz = (double)w2.w - w2.z;
    a = ax * ay + ay * az + az * ax + bx * by + by * bz + bz * bx;
  }
  for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ void k_zero_u32n(uint32_t *__restrict__ p, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0;
}

// Builds ctx->nodesP (same 64 B binary node layout as ctx->nodes) from the
// current leaf order and primitive boxes; sets ctx->clusterOk when the tree
// is usable (depth <= 64).
int buildCluster(cpt_ctx *ctx) {
  const int64_t n = ctx->nPrims;
  ctx->clusterValid = true;
  ctx->clusterOk = false;
  // preorder indices plus leaf counts stay below 2^31 up to 2^30 leaves
  if (n < 2 || n > ((int64_t)1 << 30)) return CPT_OK;
  const int64_t nbMax = (n + kScanBlock - 1) / kScanBlock;
  int rc;
  if ((rc = ensure(ctx, ctx->plC[0], n * 4)) || (rc = ensure(ctx, ctx->plC[1], n * 4)) ||
      (rc = ensure(ctx, ctx->plB[0], n * 32)) || (rc = ensure(ctx, ctx->plB[1], n * 32)) ||
      (rc = ensure(ctx, ctx->plD[0], n * 4)) || (rc = ensure(ctx, ctx->plD[1], n * 4)) ||
      (rc = ensure(ctx, ctx->plS[0], n * 4)) || (rc = ensure(ctx, ctx->plS[1], n * 4)) ||
      (rc = ensure(ctx, ctx->plCtr, (kMaxClusterDepth + 2) * 4)) || (rc = ensure(ctx, ctx->plTmp, (n - 1) * 64)) ||
      (rc = ensure(ctx, ctx->plSah, ((n - 1 + 255) / 256) * 8)) ||
      (rc = ensure(ctx, ctx->plNN, n * 4)) || (rc = ensure(ctx, ctx->plSum, (nbMax + 1) * 8)) ||
      (rc = ensure(ctx, ctx->nodesP, (n - 1) * 64)))
    return rc;
  hipStream_t s = ctx->stream;
  hipLaunchKernelGGL(k_ploc_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (int)n,
                     ctx->order.as<uint32_t>(), ctx->box32.as<float>(), ctx->plC[0].as<int32_t>(),
                     ctx->plB[0].as<float4>(), ctx->plD[0].as<int32_t>(), ctx->plS[0].as<int32_t>());
  CPT_TRY_HIP(ctx, hipGetLastError());
  int64_t cur = n, merged = 0;
  int src = 0;
  while (cur > 1) {
    const int nb = (int)((cur + kScanBlock - 1) / kScanBlock);
    hipLaunchKernelGGL(k_ploc_nn, dim3((unsigned)((cur + kNnThreads - 1) / kNnThreads)), dim3(kNnThreads), 0, s,
                       (int)cur, (const float4 *)ctx->plB[src].as<float4>(), ctx->plNN.as<int32_t>());
    hipLaunchKernelGGL(k_ploc_mark, dim3(nb), dim3(kScanThreads), 0, s, (int)cur,
                       (const int32_t *)ctx->plNN.as<int32_t>(), ctx->plSum.as<uint64_t>());
    hipLaunchKernelGGL(k_ploc_scan, dim3(1), dim3(kScanThreads), 0, s, nb, ctx->plSum.as<uint64_t>());
    hipLaunchKernelGGL(k_ploc_apply, dim3(nb), dim3(kScanThreads), 0, s, (int)cur,
                       (const int32_t *)ctx->plNN.as<int32_t>(), (const uint64_t *)ctx->plSum.as<uint64_t>(),
                       (const int32_t *)ctx->plC[src].as<int32_t>(), (const float4 *)ctx->plB[src].as<float4>(),
                       (const int32_t *)ctx->plD[src].as<int32_t>(), (const int32_t *)ctx->plS[src].as<int32_t>(),
                       ctx->plC[1 - src].as<int32_t>(), ctx->plB[1 - src].as<float4>(), ctx->plD[1 - src].as<int32_t>(),
                       ctx->plS[1 - src].as<int32_t>(), ctx->nodesP.as<float4>(), (int)(n - 2 - merged));
    CPT_TRY_HIP(ctx, hipGetLastError());
    uint64_t tot = 0;
    CPT_TRY_HIP(ctx, hipMemcpyAsync(&tot, ctx->plSum.as<uint64_t>() + nb, 8, hipMemcpyDeviceToHost, s));
    CPT_TRY_HIP(ctx, hipStreamSynchronize(s));
    const int64_t stay = (int64_t)(tot >> 32), merges = (int64_t)(uint32_t)tot;
    if (merges < 1 || stay != cur - merges || merged + merges > n - 1)
      return fail(ctx, CPT_ERR_STATE, "cluster", "clustering made no progress");
    merged += merges;
    cur = stay;
    src = 1 - src;
  }
  int32_t depth = 0;
  CPT_TRY_HIP(ctx, hipMemcpyAsync(&depth, ctx->plD[src].as<int32_t>(), 4, hipMemcpyDeviceToHost, s));
  CPT_TRY_HIP(ctx, hipStreamSynchronize(s));
  ctx->clusterDepth = depth;
  if (merged != n - 1 || depth > kMaxClusterDepth) return CPT_OK;
  // surface-area cost of both trees (summed in a fixed order on the host)
  const int nbS = (int)((n - 1 + 255) / 256);
  std::vector<double> part(nbS);
  double cost[2];
  for (int t = 0; t < 2; t++) {
    hipLaunchKernelGGL(k_sah_partial, dim3(nbS), dim3(256), 0, s, (int)(n - 1),
                       (const float4 *)(t == 0 ? ctx->nodes : ctx->nodesP).as<float4>(), ctx->plSah.as<double>());
    CPT_TRY_HIP(ctx, hipGetLastError());
    CPT_TRY_HIP(ctx, hipMemcpyAsync(part.data(), ctx->plSah.p, nbS * 8, hipMemcpyDeviceToHost, s));
    CPT_TRY_HIP(ctx, hipStreamSynchronize(s));
    cost[t] = 0.0;
    for (int b = 0; b < nbS; b++) cost[t] += part[b];
  }
  ctx->lbvhCost = cost[0];
  ctx->clusterCost = cost[1];
  // the records group the clustered tree only where it is clearly cheaper:
  // the cost ignores that closest-hit walks stop early, and a tree within a
  // few percent of the LBVH's cost can walk slower (a 10M random soup: 0.5%
  // cheaper, 20% more record visits; a displaced icosphere: 23% cheaper,
  // 32% fewer visits)
  if (!(cost[1] < kClusterMargin * cost[0])) return CPT_OK;
  // preorder renumbering: frontier pairs in the (now free) box buffers, the
  // map in the neighbour buffer, one append counter per level
  int2 *fr[2] = {(int2 *)ctx->plB[0].p, (int2 *)ctx->plB[1].p};
  uint32_t *ctr = ctx->plCtr.as<uint32_t>();
  hipLaunchKernelGGL(k_zero_u32n, dim3(1), dim3(128), 0, s, ctr, kMaxClusterDepth + 2);
  const int2 root = make_int2(0, 0);
  CPT_TRY_HIP(ctx, hipMemcpyAsync(fr[0], &root, sizeof root, hipMemcpyHostToDevice, s));
  uint32_t cnt = 1;
  int64_t reached = 0;
  for (int level = 0; cnt > 0; level++) {
    if (level > kMaxClusterDepth) return fail(ctx, CPT_ERR_STATE, "cluster", "renumbering deeper than the tree");
    reached += cnt;
    hipLaunchKernelGGL(k_dfs_level, dim3((cnt + 255) / 256), dim3(256), 0, s, cnt, (const int2 *)fr[level & 1],
                       (const float4 *)ctx->nodesP.as<float4>(), ctx->plNN.as<int32_t>(), fr[(level + 1) & 1],
                       ctr + level + 1);
    CPT_TRY_HIP(ctx, hipGetLastError());
    CPT_TRY_HIP(ctx, hipMemcpyAsync(&cnt, ctr + level + 1, 4, hipMemcpyDeviceToHost, s));
    CPT_TRY_HIP(ctx, hipStreamSynchronize(s));
  }
  if (reached != n - 1) return fail(ctx, CPT_ERR_STATE, "cluster", "renumbering missed nodes");
  hipLaunchKernelGGL(k_dfs_remap, dim3((unsigned)((n - 1 + 255) / 256)), dim3(256), 0, s, (int)(n - 1),
                     (const float4 *)ctx->nodesP.as<float4>(), (const int32_t *)ctx->plNN.as<int32_t>(),
                     ctx->plTmp.as<float4>());
  CPT_TRY_HIP(ctx, hipGetLastError());
  std::swap(ctx->nodesP, ctx->plTmp);
  ctx->clusterOk = true;
  return CPT_OK;
}

}  // namespace cpt

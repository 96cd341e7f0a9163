// bvh.hip - LBVH build for gfx950: primitive packing, FP64 Morton codes,
// LDS-staged wave-level LSD radix sort, Karras emission and an atomic
// bottom-up refit.
//
// Replaces Device::buildBVH (CudaPathTracing/Device.cuh:169-195):
//   generateMortonCodes (Device.cuh:134-142, morton3D BVH.cuh:23-41)
//     -> k_morton: FP64 normalisation (bit-exact codes), writes the 30-bit
//        code and the tile histogram of the first radix digit.
//   thrust::sort_by_key on (code<<32|index) (Device.cuh:187-189)
//     -> 4 passes of 8-bit stable LSD sort on the code with the index as
//        payload.  Keys are unique and the payload starts as the identity,
//        so a stable sort of the code yields the same permutation as the
//        reference's 64-bit key sort.  The last pass writes the 64-bit keys.
//   generateHierarchy (BVH.cuh:174-241)
//     -> k_emit: determineRange/findSplit on the full 64-bit key, child refs
//        into the 64 B node, parent links for the refit.
//   constructAABB<<<1,1>>> (BVH.cuh:243-280)
//     -> k_refit_local + k_refit_cross: bottom-up climbs where the second
//        child to arrive continues (LDS counters inside a chunk of sorted
//        leaves, agent-scope counters across chunks); min/max union is exact
//        so the result equals the reference's serial post-order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "cpt_device.hpp"
#include "internal.hpp"

namespace cpt {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kTile = kSortThreads * kSortItems;  // 4096 keys per tile
constexpr int kWaveKeys = 64 * kSortItems;         // 1024 keys per wave

// ---- FP64 helpers (no contraction: matches the reference's FP64 order) ---
__device__ __forceinline__ float roundDown(double x) {
  float f = (float)x;
  if ((double)f > x) f = nextafterf(f, -INFINITY);
  return f;
}
__device__ __forceinline__ float roundUp(double x) {
  float f = (float)x;
  if ((double)f < x) f = nextafterf(f, INFINITY);
  return f;
}
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }

// Per triangle: FP64 centre ((p0+p1)+p2)/3 (Hittable.cuh:459), FP64 padded
// box (Hittable.cuh:460-461, AABB.cuh:30-40) rounded outward to FP32, and
// the 48 B leaf record.
__global__ void k_pack_tris(const double *__restrict__ tri, const int32_t *__restrict__ mat,
                            int64_t nt, int64_t base, double *__restrict__ center,
                            float *__restrict__ box, float4 *__restrict__ rec) {
#pragma clang fp contract(off)
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  const double *p = tri + 9 * i;
  double x0 = p[0], y0 = p[1], z0 = p[2], x1 = p[3], y1 = p[4], z1 = p[5], x2 = p[6], y2 = p[7],
         z2 = p[8];
  int64_t id = base + i;
  center[3 * id + 0] = ((x0 + x1) + x2) / 3.0;
  center[3 * id + 1] = ((y0 + y1) + y2) / 3.0;
  center[3 * id + 2] = ((z0 + z1) + z2) / 3.0;
  double mnx = dmin(dmin(x0, x1), x2), mny = dmin(dmin(y0, y1), y2), mnz = dmin(dmin(z0, z1), z2);
  double mxx = dmax(dmax(x0, x1), x2), mxy = dmax(dmax(y0, y1), y2), mxz = dmax(dmax(z0, z1), z2);
  const double pad = 0.0001;
  mnx -= pad; mny -= pad; mnz -= pad;
  mxx += pad; mxy += pad; mxz += pad;
  float *b = box + 6 * id;
  b[0] = roundDown(dmin(mnx, mxx)); b[1] = roundDown(dmin(mny, mxy)); b[2] = roundDown(dmin(mnz, mxz));
  b[3] = roundUp(dmax(mnx, mxx)); b[4] = roundUp(dmax(mny, mxy)); b[5] = roundUp(dmax(mnz, mxz));
  int m = mat ? mat[i] : 0;
  rec[3 * id + 0] = make_float4((float)x0, (float)y0, (float)z0, __int_as_float((int)id));
  rec[3 * id + 1] = make_float4((float)x1, (float)y1, (float)z1, __int_as_float(CPT_OBJ_TRIANGLE));
  rec[3 * id + 2] = make_float4((float)x2, (float)y2, (float)z2, __int_as_float(m));
}

__device__ __forceinline__ uint32_t expandBits(uint32_t v) {  // BVH.cuh:12-19
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__device__ __forceinline__ uint32_t mortonCode(double cx, double cy, double cz, double3 smin,
                                               double3 smax) {
#pragma clang fp contract(off)
  double nx = (cx - smin.x) / (smax.x - smin.x);
  double ny = (cy - smin.y) / (smax.y - smin.y);
  double nz = (cz - smin.z) / (smax.z - smin.z);
  double x = dmin(dmax(nx * 1024.0, 0.0), 1023.0);
  double y = dmin(dmax(ny * 1024.0, 0.0), 1023.0);
  double z = dmin(dmax(nz * 1024.0, 0.0), 1023.0);
  return expandBits((uint32_t)x) * 4u + expandBits((uint32_t)y) * 2u + expandBits((uint32_t)z);
}

// Morton code per primitive; tile histogram of digit 0 (hist is digit-major:
// hist[d * nTiles + tile]).  keys64 (optional) receives code<<32|index.
__global__ __launch_bounds__(kSortThreads) void k_morton(const double *__restrict__ center, int64_t n,
                                                         double3 smin, double3 smax,
                                                         uint32_t *__restrict__ codes,
                                                         uint32_t *__restrict__ hist, int nTiles,
                                                         uint64_t *__restrict__ keys64) {
  __shared__ uint32_t sh[256];
  const int tid = threadIdx.x;
  sh[tid] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile;
#pragma unroll 4
  for (int j = 0; j < kSortItems; j++) {
    int64_t i = base + j * kSortThreads + tid;
    if (i < n) {
      uint32_t c = mortonCode(center[3 * i], center[3 * i + 1], center[3 * i + 2], smin, smax);
      if (codes) codes[i] = c;
      if (keys64) keys64[i] = ((uint64_t)c << 32) | (uint64_t)(uint32_t)i;
      atomicAdd(&sh[c & 255u], 1u);
    }
  }
  __syncthreads();
  if (hist) hist[(int64_t)tid * nTiles + blockIdx.x] = sh[tid];
}

__global__ __launch_bounds__(kSortThreads) void k_hist(const uint32_t *__restrict__ keys, int64_t n,
                                                       int shift, uint32_t dmask, uint32_t *__restrict__ hist,
                                                       int nTiles) {
  __shared__ uint32_t sh[256];
  const int tid = threadIdx.x;
  sh[tid] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile;
#pragma unroll 4
  for (int j = 0; j < kSortItems; j++) {
    int64_t i = base + j * kSortThreads + tid;
    if (i < n) atomicAdd(&sh[(keys[i] >> shift) & dmask], 1u);
  }
  __syncthreads();
  hist[(int64_t)tid * nTiles + blockIdx.x] = sh[tid];
}

// Exclusive scan of 256 values held one per thread of a 256-thread block.
__device__ __forceinline__ uint32_t blockExclusiveScan256(uint32_t v, uint32_t *scratch4,
                                                          uint32_t &total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) scratch4[wave] = x;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < 4; w++) {
    uint32_t s = scratch4[w];
    if (w < wave) pre += s;
    tot += s;
  }
  total = tot;
  __syncthreads();
  return pre + x - v;
}

// Row d of the digit-major histogram -> exclusive prefix over tiles;
// rowTotal[d] = count of digit d.
__global__ __launch_bounds__(kSortThreads) void k_scan_rows(uint32_t *__restrict__ hist, int nTiles,
                                                            uint32_t *__restrict__ rowTotal) {
  __shared__ uint32_t scratch[4];
  const int d = blockIdx.x, tid = threadIdx.x;
  uint32_t *row = hist + (int64_t)d * nTiles;
  const int chunk = (nTiles + kSortThreads - 1) / kSortThreads;
  const int beg = tid * chunk, end = min(beg + chunk, nTiles);
  uint32_t s = 0;
  for (int i = beg; i < end; i++) s += row[i];
  uint32_t total;
  uint32_t run = blockExclusiveScan256(s, scratch, total);
  for (int i = beg; i < end; i++) {
    uint32_t v = row[i];
    row[i] = run;
    run += v;
  }
  if (tid == 0) rowTotal[d] = total;
}

// Stable scatter of one 8-bit digit.  Wave w of the block ranks keys
// [tile + w*1024, +1024) in 16 rounds of 64 (lane order inside a round), so
// (wave, round, lane) is the input order.  Ranks come from a ballot-based
// match over the 8 digit bits and per-wave LDS counters; the tile is then
// reordered in LDS and written out digit run by digit run.
__global__ __launch_bounds__(kSortThreads) void k_scatter(
    const uint32_t *__restrict__ keysIn, const uint32_t *__restrict__ valsIn,
    uint32_t *__restrict__ keysOut, uint32_t *__restrict__ valsOut, uint64_t *__restrict__ keys64Out,
    int64_t n, int shift, uint32_t dmask, const uint32_t *__restrict__ hist, const uint32_t *__restrict__ rowTotal,
    int nTiles) {
  __shared__ uint32_t sWave[4][256];
  __shared__ uint32_t sStart[256];
  __shared__ uint32_t sGlobal[256];
  __shared__ uint32_t sKeys[kTile];
  __shared__ uint32_t sVals[kTile];
  __shared__ uint32_t scratch[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t tileBase = (int64_t)blockIdx.x * kTile;
  const int64_t waveBase = tileBase + (int64_t)wave * kWaveKeys;
  const int tileCount = (int)min((int64_t)kTile, n - tileBase);

#pragma unroll
  for (int w = 0; w < 4; w++) sWave[w][tid] = 0;

  uint32_t key[kSortItems], val[kSortItems], pos[kSortItems];
#pragma unroll
  for (int j = 0; j < kSortItems; j++) {
    int64_t i = waveBase + j * 64 + lane;
    key[j] = i < n ? keysIn[i] : 0u;
    val[j] = i < n ? (valsIn ? valsIn[i] : (uint32_t)i) : 0u;
  }
  __syncthreads();
  const uint64_t ltMask = (1ull << lane) - 1ull;
#pragma unroll
  for (int j = 0; j < kSortItems; j++) {
    int64_t i = waveBase + j * 64 + lane;
    if (i < n) {
      const uint32_t d = (key[j] >> shift) & dmask;
      uint64_t peers = __ballot(1);
#pragma unroll
      for (int bit = 0; bit < 8; bit++) {
        const bool b = (d >> bit) & 1u;
        const uint64_t m = __ballot(b);
        peers &= b ? m : ~m;
      }
      const uint32_t rank = __popcll(peers & ltMask);
      const uint32_t base = sWave[wave][d];
      const int leader = __ffsll((unsigned long long)peers) - 1;
      __builtin_amdgcn_wave_barrier();
      if (lane == leader) sWave[wave][d] = base + (uint32_t)__popcll(peers);
      __builtin_amdgcn_wave_barrier();
      pos[j] = base + rank;
    }
  }
  __syncthreads();
  // per-digit prefix across waves, tile-local digit starts, global bases
  uint32_t run = 0;
#pragma unroll
  for (int w = 0; w < 4; w++) {
    uint32_t c = sWave[w][tid];
    sWave[w][tid] = run;
    run += c;
  }
  uint32_t tot;
  const uint32_t start = blockExclusiveScan256(run, scratch, tot);
  const uint32_t digitStart = blockExclusiveScan256(rowTotal[tid], scratch, tot);
  sStart[tid] = start;
  sGlobal[tid] = digitStart + hist[(int64_t)tid * nTiles + blockIdx.x] - start;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSortItems; j++) {
    int64_t i = waveBase + j * 64 + lane;
    if (i < n) {
      const uint32_t d = (key[j] >> shift) & dmask;
      const uint32_t p = sStart[d] + sWave[wave][d] + pos[j];
      if (p < (uint32_t)kTile) {  // always true for consistent ranks
        sKeys[p] = key[j];
        sVals[p] = val[j];
      }
    }
  }
  __syncthreads();
  for (int i = tid; i < tileCount; i += kSortThreads) {
    const uint32_t k = sKeys[i];
    const uint32_t v = sVals[i];
    const uint32_t g = sGlobal[(k >> shift) & dmask] + (uint32_t)i;
    if ((int64_t)g >= n) continue;  // always false for a consistent histogram
    if (keysOut) keysOut[g] = k;
    if (valsOut) valsOut[g] = v;
    if (keys64Out) keys64Out[g] = ((uint64_t)k << 32) | (uint64_t)v;
  }
}

// ---- Karras emission (BVH.cuh:81-241) --------------------------------------
__device__ __forceinline__ int deltaKey(const uint64_t *__restrict__ k, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint64_t a = k[i], b = k[j];
  if (a == b) return 64;
  return __clzll(a ^ b);
}

__global__ void k_emit(const uint64_t *__restrict__ keys, int n, float4 *__restrict__ nodes,
                       int32_t *__restrict__ parentInt, int32_t *__restrict__ parentLeaf) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n - 1) return;
  // determineRange, BVH.cuh:130-172
  const int dNext = deltaKey(keys, n, idx, idx + 1);
  const int dPrev = deltaKey(keys, n, idx, idx - 1);
  const int d = dNext > dPrev ? 1 : -1;
  const int dMin = deltaKey(keys, n, idx, idx - d);
  int64_t lMax = 2;
  while (true) {
    int64_t j = (int64_t)idx + lMax * d;
    if (j < 0 || j >= n) break;
    if (deltaKey(keys, n, idx, (int)j) <= dMin) break;
    lMax *= 2;
  }
  int64_t l = 0;
  for (int64_t t = lMax / 2; t >= 1; t /= 2) {
    int64_t j = (int64_t)idx + (l + t) * d;
    if (j >= 0 && j < n && deltaKey(keys, n, idx, (int)j) > dMin) l += t;
  }
  const int jEnd = (int)(idx + l * d);
  const int first = min(idx, jEnd), last = max(idx, jEnd);
  // findSplit, BVH.cuh:81-118
  int split;
  const uint64_t fc = keys[first], lc = keys[last];
  if (fc == lc) {
    split = (first + last) >> 1;
  } else {
    const int common = __clzll(fc ^ lc);
    split = first;
    int step = last - first;
    do {
      step = (step + 1) >> 1;
      const int ns = split + step;
      if (ns < last && __clzll(fc ^ keys[ns]) > common) split = ns;
    } while (step > 1);
  }
  const int ca = (split == first) ? ~split : split;
  const int cb = (split + 1 == last) ? ~(split + 1) : (split + 1);
  nodes[4 * (int64_t)idx + 3] =
      make_float4(__int_as_float(ca), __int_as_float(cb), __int_as_float(first), __int_as_float(last));
  if (ca < 0) parentLeaf[~ca] = idx << 1; else parentInt[ca] = idx << 1;
  if (cb < 0) parentLeaf[~cb] = (idx << 1) | 1; else parentInt[cb] = (idx << 1) | 1;
  if (idx == 0) parentInt[0] = -1;
}

// Writes a child box into slot (0 = A, 1 = B) of node `node`.
__device__ __forceinline__ void storeChildBox(float *__restrict__ nf, int64_t node, int slot,
                                              const float bx[6]) {
  float *p = nf + 16 * node;
  *(float4 *)(p + slot * 4) = make_float4(bx[0], bx[3], bx[1], bx[4]);
  *(float2 *)(p + 8 + slot * 2) = make_float2(bx[2], bx[5]);
}

// Refit in two passes.  Pass 1 (k_refit_local): a workgroup owns a chunk of
// kRefitChunk consecutive sorted leaves, one thread per leaf so that every
// climb of the chunk is in flight at once, and finishes, bottom-up in LDS with
// workgroup-scope counters, every internal node whose leaf range lies inside
// the chunk.  Each child of a node that crosses a chunk boundary is written
// to global memory and its parent appended to an escape list.  Pass 2
// climbs from the escapes through the crossing nodes, in chunks of
// kRefitMid[0] and then kRefitMid[1] leaves (k_refit_mid: one workgroup per
// chunk, workgroup-scope counters) and last (k_refit_cross) over the nodes
// that cross those chunks' boundaries too, with agent-scope counters (each
// of which writes back and invalidates the L2 on this multi-XCD GPU:
// climbing every crossing node that way took 0.63 ms of a 10M build).
// Internal node i of a Karras tree is the lowest common ancestor of leaves
// split(i), split(i)+1, so the crossing nodes of a chunk size are the
// ancestors of the chunk-boundary splits: at most 64 per boundary (depth <=
// 64 for unique 64-bit keys), two escapes each.
constexpr int kRefitThreads = 256;  // k_refit_cross
constexpr int kRefitChunk = 1024;   // leaves (and threads) per k_refit_local workgroup
constexpr int64_t kRefitMid[2] = {int64_t(1) << 16, int64_t(1) << 22};

static int64_t escapeBound(int64_t n, int64_t chunk) {
  const int64_t chunks = (n + chunk - 1) / chunk;
  return std::min<int64_t>(2 * n, 128 * chunks + 2);
}

// escape lists of the three chunk sizes, then their counters
struct EscapeLists {
  int64_t cap[3], off[3], counters;
};
static EscapeLists escapeLists(int64_t n) {
  EscapeLists e;
  e.cap[0] = escapeBound(n, kRefitChunk);
  e.cap[1] = escapeBound(n, kRefitMid[0]);
  e.cap[2] = escapeBound(n, kRefitMid[1]);
  e.off[0] = 0;
  e.off[1] = e.cap[0];
  e.off[2] = e.cap[0] + e.cap[1];
  e.counters = e.off[2] + e.cap[2];
  return e;
}

int64_t refitEscapeBound(int64_t n) { return escapeLists(n).counters + 64; }

__global__ __launch_bounds__(kRefitChunk) void k_refit_local(
    int n, const uint32_t *__restrict__ order, const float4 *__restrict__ rec, const float *__restrict__ box32,
    float4 *__restrict__ nodes, const int32_t *__restrict__ parentInt, const int32_t *__restrict__ parentLeaf,
    int32_t *__restrict__ flags, float4 *__restrict__ leaves, float *__restrict__ rootBox,
    int32_t *__restrict__ esc, uint32_t *__restrict__ escCount, int64_t escCap) {
  __shared__ float cbox[kRefitChunk * 12];  // [node - c0][slot][minx miny minz maxx maxy maxz]
  __shared__ int cnt[kRefitChunk];
  // leaf range and parent link of internal nodes c0 .. c1: a node whose
  // range lies inside the chunk has its index there (node i of a Karras tree
  // is an end of its own range), so the climbs read no global memory
  __shared__ int2 rng[kRefitChunk];
  __shared__ int parS[kRefitChunk];
  const int c0 = blockIdx.x * kRefitChunk;
  const int c1 = min(c0 + kRefitChunk, n) - 1;  // last leaf of the chunk
  cnt[threadIdx.x] = 0;
  {
    const int j = c0 + (int)threadIdx.x;
    if (j < n - 1) {
      flags[j] = 0;  // counters of pass 2
      const float4 rg = nodes[4 * (int64_t)j + 3];  // z, w: first and last leaf
      rng[threadIdx.x] = make_int2(__float_as_int(rg.z), __float_as_int(rg.w));
      parS[threadIdx.x] = parentInt[j];
    }
  }
  __syncthreads();
  const int i = c0 + (int)threadIdx.x;
  if (i > c1) return;
  float *nf = (float *)nodes;
  const uint32_t p = order[i];
  {
    // triangles: store the edges the FP32 test forms from the vertices
    const float4 q0 = rec[3 * (int64_t)p + 0];
    float4 q1 = rec[3 * (int64_t)p + 1], q2 = rec[3 * (int64_t)p + 2];
    if (__float_as_int(q1.w) == CPT_OBJ_TRIANGLE) {
      q1 = make_float4(q1.x - q0.x, q1.y - q0.y, q1.z - q0.z, q1.w);
      q2 = make_float4(q2.x - q0.x, q2.y - q0.y, q2.z - q0.z, q2.w);
    }
    leaves[3 * (int64_t)i + 0] = q0;
    leaves[3 * (int64_t)i + 1] = q1;
    leaves[3 * (int64_t)i + 2] = q2;
  }
  float bx[6];
  {
    const float2 *b2 = (const float2 *)(box32 + 6 * (int64_t)p);
    const float2 u = b2[0], v = b2[1], w = b2[2];
    bx[0] = u.x;
    bx[1] = u.y;
    bx[2] = v.x;
    bx[3] = v.y;
    bx[4] = w.x;
    bx[5] = w.y;
  }
  if (n == 1) {
#pragma unroll
    for (int k = 0; k < 6; k++) rootBox[k] = bx[k];
    return;
  }
  int par = parentLeaf[i];
  while (true) {
    const int node = par >> 1, slot = par & 1;
    storeChildBox(nf, node, slot, bx);
    // a node outside c0 .. c1 or with a leaf outside the chunk crosses its
    // boundary (parentInt[0] = -1)
    const bool inside = node >= c0 && node <= c1 && rng[node - c0].x >= c0 && rng[node - c0].y <= c1;
    if (!inside) {
      const uint32_t e = atomicAdd(escCount, 1u);
      if ((int64_t)e < escCap) esc[e] = node;
      return;
    }
    const int parNext = parS[node - c0];
    float *s = cbox + (node - c0) * 12;
#pragma unroll
    for (int k = 0; k < 6; k++) s[slot * 6 + k] = bx[k];
    const int old = __hip_atomic_fetch_add(&cnt[node - c0], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (old == 0) return;  // the sibling's thread continues
#pragma unroll
    for (int k = 0; k < 3; k++) {
      bx[k] = fminf(bx[k], s[(slot ^ 1) * 6 + k]);
      bx[k + 3] = fmaxf(bx[k + 3], s[(slot ^ 1) * 6 + k + 3]);
    }
    if (node == 0) {
#pragma unroll
      for (int k = 0; k < 6; k++) rootBox[k] = bx[k];
      return;
    }
    par = parNext;
  }
}

// One workgroup per chunk of S leaves: the escapes of the previous pass
// whose node index lies in the chunk (a node's index is an end of its own
// leaf range, so each escape has one such chunk).  A node whose range lies
// inside the chunk is finished by the second of its two arrivals (escapes
// and climbs of this workgroup only: both children lie inside the chunk
// too); a node crossing the chunk's boundary goes to the next list.
__global__ __launch_bounds__(kRefitChunk) void k_refit_mid(int n, int64_t S, const int32_t *__restrict__ escIn,
                                                           const uint32_t *__restrict__ cntIn, int64_t capIn,
                                                           int32_t *__restrict__ escOut, uint32_t *__restrict__ cntOut,
                                                           int64_t capOut, float4 *nodes,
                                                           const int32_t *__restrict__ parentInt, int32_t *flags,
                                                           float *__restrict__ rootBox) {
  const int64_t c0 = (int64_t)blockIdx.x * S, c1 = std::min<int64_t>(c0 + S, n) - 1;
  const int64_t m = std::min<int64_t>(*cntIn, capIn);
  float *nf = (float *)nodes;
  for (int64_t e = threadIdx.x; e < m; e += blockDim.x) {
    int node = escIn[e];
    if ((int64_t)node / S != (int64_t)blockIdx.x) continue;
    while (true) {
      const float4 rg = nodes[4 * (int64_t)node + 3];  // z, w: first and last leaf
      if (__float_as_int(rg.z) < c0 || __float_as_int(rg.w) > c1) {
        const uint32_t k = atomicAdd(cntOut, 1u);
        if ((int64_t)k < capOut) escOut[k] = node;
        break;
      }
      const int old = __hip_atomic_fetch_add(&flags[node], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (old == 0) break;  // the other child is not finished: its arrival continues
      const float4 a = nodes[4 * (int64_t)node + 0];
      const float4 b = nodes[4 * (int64_t)node + 1];
      const float4 c = nodes[4 * (int64_t)node + 2];
      float bx[6];
      bx[0] = fminf(a.x, b.x);
      bx[3] = fmaxf(a.y, b.y);
      bx[1] = fminf(a.z, b.z);
      bx[4] = fmaxf(a.w, b.w);
      bx[2] = fminf(c.x, c.z);
      bx[5] = fmaxf(c.y, c.w);
      if (node == 0) {
#pragma unroll
        for (int k = 0; k < 6; k++) rootBox[k] = bx[k];
        break;
      }
      const int par = parentInt[node];
      node = par >> 1;
      storeChildBox(nf, node, par & 1, bx);
    }
  }
}

__global__ __launch_bounds__(kRefitThreads) void k_refit_cross(const int32_t *__restrict__ esc,
                                                               const uint32_t *__restrict__ escCount,
                                                               int64_t escCap, float4 *nodes,
                                                               const int32_t *__restrict__ parentInt,
                                                               int32_t *flags, float *__restrict__ rootBox) {
  const int64_t e = (int64_t)blockIdx.x * kRefitThreads + threadIdx.x;
  if (e >= escCap || e >= (int64_t)*escCount) return;
  int node = esc[e];
  float *nf = (float *)nodes;
  float bx[6];
  while (true) {
    const int old = __hip_atomic_fetch_add(&flags[node], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (old == 0) return;  // the other child is not finished: its walker continues
    const float4 a = nodes[4 * (int64_t)node + 0];
    const float4 b = nodes[4 * (int64_t)node + 1];
    const float4 c = nodes[4 * (int64_t)node + 2];
    bx[0] = fminf(a.x, b.x);
    bx[3] = fmaxf(a.y, b.y);
    bx[1] = fminf(a.z, b.z);
    bx[4] = fmaxf(a.w, b.w);
    bx[2] = fminf(c.x, c.z);
    bx[5] = fmaxf(c.y, c.w);
    if (node == 0) {
#pragma unroll
      for (int k = 0; k < 6; k++) rootBox[k] = bx[k];
      return;
    }
    const int par = parentInt[node];
    node = par >> 1;
    storeChildBox(nf, node, par & 1, bx);
  }
}

// One axis of a quantized record: origin = the lowest child plane, grid
// step 2^e >= extent / 255, planes rounded outward to the grid and checked in
// FP32 with the decode the traversal uses (fmaf(q, 2^e, origin)); e grows if
// rounding leaves a top plane uncovered.
template <int N>
__device__ __forceinline__ void quantAxis(const float lo[N], const float hi[N], const bool valid[N], float &origin,
                                          uint32_t &ebyte, uint32_t qlo[N / 4], uint32_t qhi[N / 4]) {
  float mn = INFINITY, mx = -INFINITY;
#pragma unroll
  for (int k = 0; k < N; k++)
    if (valid[k]) {
      mn = fminf(mn, lo[k]);
      mx = fmaxf(mx, hi[k]);
    }
  if (!(mn <= mx)) mn = mx = 0.0f;
  origin = mn;
  const float step = (mx - mn) / 255.0f;
  int e = -126;
  if (step > 0.0f) {
    int k2;
    const float m = frexpf(step, &k2);  // step = m 2^k2, m in [0.5, 1)
    e = (m == 0.5f) ? k2 - 1 : k2;
  }
  for (int attempt = 0; attempt < 8; attempt++) {
    const int ec = e < -126 ? -126 : (e > 127 ? 127 : e);
    const float sc = __uint_as_float((uint32_t)(ec + 127) << 23);
    bool ok = true;
    uint32_t l[N / 4], h[N / 4];
#pragma unroll
    for (int w = 0; w < N / 4; w++) l[w] = h[w] = 0;
#pragma unroll
    for (int k = 0; k < N; k++) {
      uint32_t ql = 255u, qh = 0u;  // empty slot: lo > hi
      if (valid[k]) {
        int a = (int)fminf(fmaxf(floorf((lo[k] - origin) / sc), 0.0f), 255.0f);
        while (a > 0 && __builtin_fmaf((float)a, sc, origin) > lo[k]) a--;
        int b = (int)fminf(fmaxf(ceilf((hi[k] - origin) / sc), 0.0f), 255.0f);
        while (b < 255 && __builtin_fmaf((float)b, sc, origin) < hi[k]) b++;
        if (__builtin_fmaf((float)b, sc, origin) < hi[k] || __builtin_fmaf((float)a, sc, origin) > lo[k]) ok = false;
        ql = (uint32_t)a;
        qh = (uint32_t)b;
      }
      l[k / 4] |= ql << (8 * (k % 4));
      h[k / 4] |= qh << (8 * (k % 4));
    }
    ebyte = (uint32_t)(ec + 127);
#pragma unroll
    for (int w = 0; w < N / 4; w++) {
      qlo[w] = l[w];
      qhi[w] = h[w];
    }
    if (ok || ec == 127) return;
    e = ec + 1;
  }
}

// Four-wide entries of binary node i: its children, each internal child
// replaced by its own two children; when one child is a leaf, the
// larger-area internal grandchild is opened into the free slot.  ref[] holds
// binary node indices (>= 0), ~leaf or kEmpty4.
__device__ __forceinline__ void collapse4(const float4 *__restrict__ nodes, int64_t i, float b[6][4], int ref[4]) {
  // Every slot index below is a compile-time constant (selects, no running
  // counter), so the slot arrays stay in registers.
  const float4 a0 = nodes[4 * (int64_t)i], a1 = nodes[4 * (int64_t)i + 1], a2 = nodes[4 * (int64_t)i + 2],
               a3 = nodes[4 * (int64_t)i + 3];
  const int cr0 = __float_as_int(a3.x), cr1 = __float_as_int(a3.y);
  const bool in0 = cr0 >= 0, in1 = cr1 >= 0;
  // child boxes of a node: slot 0 = A, 1 = B (see storeChildBox)
  const float cb0[6] = {a0.x, a0.z, a2.x, a0.y, a0.w, a2.y}, cb1[6] = {a1.x, a1.z, a2.z, a1.y, a1.w, a2.w};
  // the two children of each internal child
  float g0[2][6] = {}, g1[2][6] = {};
  int g0r[2] = {0, 0}, g1r[2] = {0, 0};
  if (in0) {
    const int64_t g = cr0;
    const float4 q0 = nodes[4 * g], q1 = nodes[4 * g + 1], q2 = nodes[4 * g + 2], q3 = nodes[4 * g + 3];
    const float gb[2][6] = {{q0.x, q0.z, q2.x, q0.y, q0.w, q2.y}, {q1.x, q1.z, q2.z, q1.y, q1.w, q2.w}};
#pragma unroll
    for (int q = 0; q < 6; q++) {
      g0[0][q] = gb[0][q];
      g0[1][q] = gb[1][q];
    }
    g0r[0] = __float_as_int(q3.x);
    g0r[1] = __float_as_int(q3.y);
  }
  if (in1) {
    const int64_t g = cr1;
    const float4 q0 = nodes[4 * g], q1 = nodes[4 * g + 1], q2 = nodes[4 * g + 2], q3 = nodes[4 * g + 3];
    const float gb[2][6] = {{q0.x, q0.z, q2.x, q0.y, q0.w, q2.y}, {q1.x, q1.z, q2.z, q1.y, q1.w, q2.w}};
#pragma unroll
    for (int q = 0; q < 6; q++) {
      g1[0][q] = gb[0][q];
      g1[1][q] = gb[1][q];
    }
    g1r[0] = __float_as_int(q3.x);
    g1r[1] = __float_as_int(q3.y);
  }
  // Slots in order: child A's entries (itself if a leaf, else its two
  // children), then child B's; unused slots are empty.
#pragma unroll
  for (int q = 0; q < 6; q++) {
    b[q][0] = in0 ? g0[0][q] : cb0[q];
    b[q][1] = in0 ? g0[1][q] : (in1 ? g1[0][q] : cb1[q]);
    b[q][2] = in0 ? (in1 ? g1[0][q] : cb1[q]) : (in1 ? g1[1][q] : 0.f);
    b[q][3] = (in0 && in1) ? g1[1][q] : 0.f;
  }
  ref[0] = in0 ? g0r[0] : cr0;
  ref[1] = in0 ? g0r[1] : (in1 ? g1r[0] : cr1);
  ref[2] = in0 ? (in1 ? g1r[0] : cr1) : (in1 ? g1r[1] : kEmpty4);
  ref[3] = (in0 && in1) ? g1r[1] : kEmpty4;
  // One leaf child and one internal child leave a slot free: open the
  // grandchild with the larger surface area into it.  Every internal entry
  // then still lies at least two binary levels below node i, so a walk pushes
  // at most three entries per two levels.
  if (in0 != in1) {
    int pick = -1;
    float best = -1.f;
#pragma unroll
    for (int s = 0; s < 3; s++) {
      if (ref[s] < 0) continue;
      const float ex = b[3][s] - b[0][s], ey = b[4][s] - b[1][s], ez = b[5][s] - b[2][s];
      const float area = ex * ey + ey * ez + ez * ex;
      if (area > best) {
        best = area;
        pick = s;
      }
    }
    if (pick >= 0) {
      const int64_t g = pick == 0 ? ref[0] : pick == 1 ? ref[1] : ref[2];
      const float4 g0v = nodes[4 * g], g1v = nodes[4 * g + 1], g2v = nodes[4 * g + 2], g3v = nodes[4 * g + 3];
      const float gb[2][6] = {{g0v.x, g0v.z, g2v.x, g0v.y, g0v.w, g2v.y},
                              {g1v.x, g1v.z, g2v.z, g1v.y, g1v.w, g2v.w}};
#pragma unroll
      for (int s = 0; s < 3; s++)
        if (s == pick) {
#pragma unroll
          for (int q = 0; q < 6; q++) b[q][s] = gb[0][q];
          ref[s] = __float_as_int(g3v.x);
        }
#pragma unroll
      for (int q = 0; q < 6; q++) b[q][3] = gb[1][q];
      ref[3] = __float_as_int(g3v.y);
    }
  }
}

// Record `slot` in the FP32 (nodes4) and/or quantized (nodes4q) format.
__device__ __forceinline__ void storeRecord4(float4 *__restrict__ nodes4, float4 *__restrict__ nodes4q, int64_t slot,
                                             const float b[6][4], const int ref[4]) {
  if (nodes4) {
    float4 *o = nodes4 + 8 * slot;
    // SoA: minx maxx miny maxy minz maxz refs
    o[0] = make_float4(b[0][0], b[0][1], b[0][2], b[0][3]);
    o[1] = make_float4(b[3][0], b[3][1], b[3][2], b[3][3]);
    o[2] = make_float4(b[1][0], b[1][1], b[1][2], b[1][3]);
    o[3] = make_float4(b[4][0], b[4][1], b[4][2], b[4][3]);
    o[4] = make_float4(b[2][0], b[2][1], b[2][2], b[2][3]);
    o[5] = make_float4(b[5][0], b[5][1], b[5][2], b[5][3]);
    o[6] = make_float4(__int_as_float(ref[0]), __int_as_float(ref[1]), __int_as_float(ref[2]),
                       __int_as_float(ref[3]));
    o[7] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  if (!nodes4q) return;
  // quantized record
  const bool valid[4] = {ref[0] != (int)0x80000000, ref[1] != (int)0x80000000, ref[2] != (int)0x80000000,
                         ref[3] != (int)0x80000000};
  float org[3];
  uint32_t eb[3], ql[3][1], qh[3][1];
#pragma unroll
  for (int ax = 0; ax < 3; ax++) quantAxis<4>(b[ax], b[ax + 3], valid, org[ax], eb[ax], ql[ax], qh[ax]);
  float4 *oq = nodes4q + 4 * slot;
  oq[0] = make_float4(org[0], org[1], org[2], __uint_as_float(eb[0] | (eb[1] << 8) | (eb[2] << 16)));
  oq[1] = make_float4(__int_as_float(ref[0]), __int_as_float(ref[1]), __int_as_float(ref[2]), __int_as_float(ref[3]));
  oq[2] = make_float4(__uint_as_float(ql[0][0]), __uint_as_float(qh[0][0]), __uint_as_float(ql[1][0]),
                      __uint_as_float(qh[1][0]));
  oq[3] = make_float4(__uint_as_float(ql[2][0]), __uint_as_float(qh[2][0]), 0.f, 0.f);
}

// Four-wide records in level order from the root.  Only the records a walk
// can reach are written: every internal entry lies two or three binary
// levels below its record's node, so they are fewer than half of the N-1
// binary nodes (0.46 of them on random triangle soups), and the internal
// entries of one record get consecutive slots, so sibling records share
// 128 B lines.  lev[2L] is the first slot of
// record level L and lev[2L+1] its width.  Processing level L collapses node
// recNode[s] for each of the level's slots s, takes slots for the record's
// internal entries at the end of level L + 1 and writes the record with
// those slots as its references.  The order of slots within a level follows
// the order in which workgroups take them; no walk depends on it (walks
// compare entry distances and slot positions within a record, never record
// indices).
constexpr int kRecLevels = 32;  // an entry of level L lies at binary depth >= 2L; internal nodes at <= 63
constexpr int kRecTop = 5;      // levels 0..4 (at most 4^4 = 256 records each) in one workgroup

template <bool kQuant>
__device__ __forceinline__ void storeRecordFmt(float4 *__restrict__ out, int64_t slot, const float b[6][4],
                                               const int ref[4]) {
  storeRecord4(kQuant ? nullptr : out, kQuant ? out : nullptr, slot, b, ref);
}

// First slot of this thread's k (0..4) internal entries: workgroup-wide
// exclusive prefix over the 256 threads, plus one atomic on *width for the
// whole workgroup (none when width is null: the caller is the only
// workgroup).  Returns the workgroup's total in *wgTotal.
__device__ __forceinline__ uint32_t takeSlots(uint32_t k, uint32_t *__restrict__ width, uint32_t next,
                                              uint32_t *sh, uint32_t *wgTotal) {
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
  const uint64_t below = (1ull << lane) - 1ull;
  uint32_t pre = 0, total = 0;
#pragma unroll
  for (int bit = 0; bit < 3; bit++) {
    const uint64_t m = __ballot((k >> bit) & 1u);
    pre += (uint32_t)__popcll(m & below) << bit;
    total += (uint32_t)__popcll(m) << bit;
  }
  __syncthreads();  // sh is free: the previous call's readers are past it
  if (lane == 0) sh[wave] = total;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int w = 0; w < 4; w++) {
      const uint32_t t = sh[w];
      sh[w] = run;
      run += t;
    }
    sh[4] = width && run ? atomicAdd(width, run) : 0u;
    sh[5] = run;
  }
  __syncthreads();
  *wgTotal = sh[5];
  return next + sh[4] + sh[wave] + pre;
}

// Collapses the record of slot `slot0` (when act), takes its entries' slots
// and writes it.  Every thread of the workgroup calls it (the slot scan).
template <bool kQuant>
__device__ __forceinline__ uint32_t collapseSlot(const float4 *__restrict__ nodes, float4 *__restrict__ out,
                                                 int32_t *__restrict__ recNode, uint32_t slot0, bool act,
                                                 int level, uint32_t *__restrict__ width, uint32_t next,
                                                 uint32_t *sh) {
  float b[6][4];
  int ref[4] = {kEmpty4, kEmpty4, kEmpty4, kEmpty4};
  if (act) collapse4(nodes, recNode[slot0], b, ref);
  const uint32_t k = (uint32_t)(ref[0] >= 0) + (ref[1] >= 0) + (ref[2] >= 0) + (ref[3] >= 0);
  uint32_t wgTotal;
  uint32_t slot = takeSlots(k, width, next, sh, &wgTotal);
  if (act) {
#pragma unroll
    for (int s = 0; s < 4; s++)
      if (ref[s] >= 0) {
        if (level + 1 < kRecLevels) {
          recNode[slot] = ref[s];
          ref[s] = (int)slot++;
        } else {
          ref[s] = kEmpty4;  // beyond the depth bound: never reached
        }
      }
    storeRecordFmt<kQuant>(out, slot0, b, ref);
  }
  return wgTotal;
}

// Levels 0 .. kRecTop-1 (at most 4^L <= 256 records each: one trip of the
// one workgroup); then lev[] for the level launches.
template <bool kQuant>
__global__ __launch_bounds__(256) void k_collapse4_top(const float4 *__restrict__ nodes, float4 *__restrict__ out,
                                                       int32_t *__restrict__ recNode, uint32_t *__restrict__ lev) {
  __shared__ uint32_t sh[6];
  if (threadIdx.x == 0) recNode[0] = 0;  // the root
  __syncthreads();
  uint32_t first = 0, width = 1;
  for (int level = 0; level < kRecTop; level++) {
    const uint32_t next = first + width;
    const uint32_t total = collapseSlot<kQuant>(nodes, out, recNode, first + threadIdx.x, threadIdx.x < width,
                                                level, nullptr, next, sh);
    __syncthreads();  // recNode of the next level written by this workgroup
    first = next;
    width = total;
  }
  if (threadIdx.x < 2 * kRecLevels + 2)
    lev[threadIdx.x] = threadIdx.x == 2 * kRecTop ? first : threadIdx.x == 2 * kRecTop + 1 ? width : 0u;
}

// One level L >= kRecTop; the grid strides over the level's slots.
template <bool kQuant>
__global__ __launch_bounds__(256) void k_collapse4_level(const float4 *__restrict__ nodes, float4 *__restrict__ out,
                                                         int32_t *__restrict__ recNode, uint32_t *__restrict__ lev,
                                                         int level) {
  __shared__ uint32_t sh[6];
  const uint32_t first = lev[2 * level], width = lev[2 * level + 1];
  const uint32_t next = first + width;
  if (blockIdx.x == 0 && threadIdx.x == 0) lev[2 * level + 2] = next;
  // the bound is uniform per workgroup: every thread runs every trip
  for (uint32_t base = blockIdx.x * blockDim.x; base < width; base += gridDim.x * blockDim.x) {
    const uint32_t t = base + threadIdx.x;
    collapseSlot<kQuant>(nodes, out, recNode, first + t, t < width, level, &lev[2 * level + 3], next, sh);
  }
}

// Eight-wide quantized record of every internal node (layout in
// cpt_device.hpp): node i's subtree expanded three binary levels deep, so
// every internal entry is a great-grandchild and every leaf entry lies at
// depth 1-3.
__global__ __launch_bounds__(256) void k_collapse8(int n, const float4 *__restrict__ nodes, float4 *__restrict__ nodes8) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  float b[6][8];
  int ref[8];
  int k = 0;
  // depth-first expansion; entries are written in slot order
  auto childBoxes = [&](int64_t g, float cb[2][6], int cr[2]) {
    const float4 g0 = nodes[4 * g], g1 = nodes[4 * g + 1], g2 = nodes[4 * g + 2], g3 = nodes[4 * g + 3];
    const float t[2][6] = {{g0.x, g0.z, g2.x, g0.y, g0.w, g2.y}, {g1.x, g1.z, g2.z, g1.y, g1.w, g2.w}};
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int q = 0; q < 6; q++) cb[c][q] = t[c][q];
    cr[0] = __float_as_int(g3.x);
    cr[1] = __float_as_int(g3.y);
  };
  auto put = [&](const float bb[6], int r) {
#pragma unroll
    for (int s = 0; s < 8; s++)
      if (s == k) {
#pragma unroll
        for (int q = 0; q < 6; q++) b[q][s] = bb[q];
        ref[s] = r;
      }
    k++;
  };
  float c1[2][6];
  int r1[2];
  childBoxes(i, c1, r1);
#pragma unroll
  for (int a = 0; a < 2; a++) {
    if (r1[a] < 0) {
      put(c1[a], r1[a]);
      continue;
    }
    float c2[2][6];
    int r2[2];
    childBoxes(r1[a], c2, r2);
#pragma unroll
    for (int e = 0; e < 2; e++) {
      if (r2[e] < 0) {
        put(c2[e], r2[e]);
        continue;
      }
      float c3[2][6];
      int r3[2];
      childBoxes(r2[e], c3, r3);
      put(c3[0], r3[0]);
      put(c3[1], r3[1]);
    }
  }
#pragma unroll
  for (int s = 0; s < 8; s++)
    if (s >= k) {
#pragma unroll
      for (int q = 0; q < 6; q++) b[q][s] = 0.f;
      ref[s] = (int)0x80000000;
    }
  bool valid[8];
#pragma unroll
  for (int s = 0; s < 8; s++) valid[s] = ref[s] != (int)0x80000000;
  float org[3];
  uint32_t eb[3], ql[3][2], qh[3][2];
#pragma unroll
  for (int ax = 0; ax < 3; ax++) quantAxis<8>(b[ax], b[ax + 3], valid, org[ax], eb[ax], ql[ax], qh[ax]);
  float4 *o = nodes8 + 8 * (int64_t)i;
  o[0] = make_float4(org[0], org[1], org[2], __uint_as_float(eb[0] | (eb[1] << 8) | (eb[2] << 16)));
  o[1] = make_float4(__int_as_float(ref[0]), __int_as_float(ref[1]), __int_as_float(ref[2]), __int_as_float(ref[3]));
  o[2] = make_float4(__int_as_float(ref[4]), __int_as_float(ref[5]), __int_as_float(ref[6]), __int_as_float(ref[7]));
#pragma unroll
  for (int ax = 0; ax < 3; ax++)
    o[3 + ax] = make_float4(__uint_as_float(ql[ax][0]), __uint_as_float(ql[ax][1]), __uint_as_float(qh[ax][0]),
                            __uint_as_float(qh[ax][1]));
  o[6] = make_float4(0.f, 0.f, 0.f, 0.f);
  o[7] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void k_zero_i32(int32_t *__restrict__ p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---- host launchers ---------------------------------------------------------

void expandBounds(const double bmin[3], const double bmax[3], double smin[3], double smax[3]) {
  // Device.cuh:174-181: +-1 expansion; +-20 when the box is empty (x tested).
  volatile double one = 1.0;
  for (int k = 0; k < 3; k++) {
    smin[k] = bmin[k] - one;
    smax[k] = bmax[k] + one;
  }
  if (smin[0] > smax[0]) {
    for (int k = 0; k < 3; k++) {
      smin[k] = -20.0;
      smax[k] = 20.0;
    }
  }
}

int packTriangles(cpt_ctx *ctx, const double *tri64Dev, const int32_t *matDev, int64_t nt,
                  int64_t base) {
  if (nt <= 0) return CPT_OK;
  const int bs = 256;
  hipLaunchKernelGGL(k_pack_tris, dim3((unsigned)((nt + bs - 1) / bs)), dim3(bs), 0, ctx->stream,
                     tri64Dev, matDev, nt, base, ctx->center64.as<double>(), ctx->box32.as<float>(),
                     ctx->rec.as<float4>());
  CPT_TRY_HIP(ctx, hipGetLastError());
  return CPT_OK;
}

int launchMorton(cpt_ctx *ctx, const double bmin[3], const double bmax[3], uint64_t *keysDev) {
  double smin[3], smax[3];
  expandBounds(bmin, bmax, smin, smax);
  const int64_t n = ctx->nPrims;
  if (n <= 0) return CPT_OK;
  const int nTiles = (int)((n + kTile - 1) / kTile);
  hipLaunchKernelGGL(k_morton, dim3(nTiles), dim3(kSortThreads), 0, ctx->stream,
                     ctx->center64.as<double>(), n, make_double3(smin[0], smin[1], smin[2]),
                     make_double3(smax[0], smax[1], smax[2]), (uint32_t *)nullptr,
                     (uint32_t *)nullptr, nTiles, keysDev);
  CPT_TRY_HIP(ctx, hipGetLastError());
  return CPT_OK;
}

// Sorts (keys, vals) by the low `bits` bits.  Result ends in (keys, vals);
// keysTmp/valsTmp are scratch of the same size.  identityVals: vals are
// generated as the index on the first pass (vals content ignored).
// If `keys64Out` is set (internal use) the last pass writes 64-bit keys.
static int sortPasses(cpt_ctx *ctx, uint32_t *keys, uint32_t *vals, uint32_t *keysTmp,
                      uint32_t *valsTmp, int64_t n, int bits, bool identityVals,
                      bool firstHistDone, uint64_t *keys64Out, uint32_t *orderOut,
                      uint32_t **valsOut = nullptr, uint32_t *hist = nullptr, uint32_t *rowTotal = nullptr) {
  const int nTiles = (int)((n + kTile - 1) / kTile);
  const int passes = (bits + 7) / 8;
  if (!hist) hist = ctx->hist.as<uint32_t>();
  if (!rowTotal) rowTotal = ctx->rowTotal.as<uint32_t>();
  uint32_t *kin = keys, *vin = vals, *kout = keysTmp, *vout = valsTmp;
  for (int p = 0; p < passes; p++) {
    const int shift = 8 * p;
    // the last pass may cover fewer than 8 key bits
    const uint32_t dmask = bits - shift >= 8 ? 255u : ((1u << (bits - shift)) - 1u);
    if (!(p == 0 && firstHistDone))
      hipLaunchKernelGGL(k_hist, dim3(nTiles), dim3(kSortThreads), 0, ctx->stream, kin, n, shift,
                         dmask, hist, nTiles);
    hipLaunchKernelGGL(k_scan_rows, dim3(256), dim3(kSortThreads), 0, ctx->stream, hist, nTiles,
                       rowTotal);
    const bool last = p == passes - 1;
    uint32_t *ko = kout, *vo = vout;
    uint64_t *k64 = nullptr;
    if (last && keys64Out) {
      ko = nullptr;
      vo = orderOut;
      k64 = keys64Out;
    }
    hipLaunchKernelGGL(k_scatter, dim3(nTiles), dim3(kSortThreads), 0, ctx->stream, kin,
                       (p == 0 && identityVals) ? (const uint32_t *)nullptr : vin, ko, vo, k64, n,
                       shift, dmask, hist, rowTotal, nTiles);
    CPT_TRY_HIP(ctx, hipGetLastError());
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  if (valsOut) {
    *valsOut = vin;  // sorted values stay where the last pass put them
  } else if (!keys64Out && kin != keys) {
    CPT_TRY_HIP(ctx, hipMemcpyAsync(keys, kin, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
    CPT_TRY_HIP(ctx, hipMemcpyAsync(vals, vin, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
  }
  return CPT_OK;
}

int radixSortPairs(cpt_ctx *ctx, uint32_t *keys, uint32_t *vals, uint32_t *keysTmp,
                   uint32_t *valsTmp, int64_t n, int bits, bool identityVals, uint32_t **valsOut,
                   uint32_t *hist, uint32_t *rowTotal) {
  if (valsOut) *valsOut = vals;
  if (n <= 0) return CPT_OK;
  const int nTiles = (int)((n + kTile - 1) / kTile);
  int rc;
  if (!hist && (rc = ensure(ctx, ctx->hist, (size_t)nTiles * 256 * 4))) return rc;
  if (!rowTotal && (rc = ensure(ctx, ctx->rowTotal, 256 * 4))) return rc;
  return sortPasses(ctx, keys, vals, keysTmp, valsTmp, n, bits, identityVals, false, nullptr,
                    nullptr, valsOut, hist, rowTotal);
}

int recordFormat(const cpt_ctx *ctx) {
  if (ctx->traversal == 8) return 3;
  if (ctx->traversal == 4) return ctx->quantized ? 2 : 1;
  return 0;
}

// Wide records of the current traversal format, collapsed from the binary
// nodes after a build or refit, or on first use after the format changed.
int ensureRecords(cpt_ctx *ctx, bool lbvhOnly) {
  const int fmt = recordFormat(ctx);
  const int64_t n = ctx->nPrims;
  if (fmt == 0 || n < 2 || (ctx->recValid & (1 << fmt))) return CPT_OK;
  const size_t nn = (size_t)(n - 1);
  int rc;
  // four-wide records from the clustered tree when it is on, current and
  // chosen (mode 1 also builds it here after the option was switched on)
  const bool four = fmt == 1 || fmt == 2;
  if (!lbvhOnly && ctx->cluster == 1 && four && !ctx->clusterValid && (rc = buildCluster(ctx))) return rc;
  const bool fromCluster = !lbvhOnly && ctx->cluster != 0 && four && ctx->clusterValid && ctx->clusterOk;
  const float4 *bin = fromCluster ? ctx->nodesP.as<float4>() : ctx->nodes.as<float4>();
  const unsigned grid = (unsigned)((n - 1 + 255) / 256);
  if (fmt == 1 || fmt == 2) {
    DevBuf &dst = fmt == 1 ? ctx->nodes4 : ctx->nodes4q;
    if ((rc = ensure(ctx, dst, nn * (fmt == 1 ? 128 : 64))) || (rc = ensure(ctx, ctx->recNode, nn * 4)) ||
        (rc = ensure(ctx, ctx->recLev, (2 * kRecLevels + 2) * 4)))
      return rc;
    float4 *out = dst.as<float4>();
    int32_t *recNode = ctx->recNode.as<int32_t>();
    uint32_t *lev = ctx->recLev.as<uint32_t>();
    if (fmt == 2) hipLaunchKernelGGL(k_collapse4_top<true>, dim3(1), dim3(256), 0, ctx->stream, bin, out, recNode, lev);
    else hipLaunchKernelGGL(k_collapse4_top<false>, dim3(1), dim3(256), 0, ctx->stream, bin, out, recNode, lev);
    // level L holds at most min(4^L, N-1) records
    int64_t widest = (int64_t)1 << (2 * kRecTop);
    for (int L = kRecTop; L < kRecLevels; L++, widest = std::min<int64_t>(widest * 4, (int64_t)nn)) {
      const unsigned g = (unsigned)std::min<int64_t>((std::min<int64_t>(widest, (int64_t)nn) + 255) / 256, 2048);
      if (fmt == 2)
        hipLaunchKernelGGL(k_collapse4_level<true>, dim3(g), dim3(256), 0, ctx->stream, bin, out, recNode, lev, L);
      else
        hipLaunchKernelGGL(k_collapse4_level<false>, dim3(g), dim3(256), 0, ctx->stream, bin, out, recNode, lev, L);
    }
  } else {
    if ((rc = ensure(ctx, ctx->nodes8, nn * 128))) return rc;
    hipLaunchKernelGGL(k_collapse8, dim3(grid), dim3(256), 0, ctx->stream, (int)n,
                       (const float4 *)ctx->nodes.as<float4>(), ctx->nodes8.as<float4>());
  }
  CPT_TRY_HIP(ctx, hipGetLastError());
  ctx->recValid |= 1 << fmt;
  return CPT_OK;
}

// Node boxes (two refit passes), sorted leaf records and the wide records of
// the current traversal format.
static int launchRefit(cpt_ctx *ctx, int64_t n) {
  const EscapeLists L = escapeLists(n);
  int rc = ensure(ctx, ctx->esc, (size_t)refitEscapeBound(n) * 4);
  if (rc) return rc;
  int32_t *esc = ctx->esc.as<int32_t>();
  uint32_t *escCount = (uint32_t *)(esc + L.counters);
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, ctx->stream, (int32_t *)escCount, (int64_t)3);
  const unsigned chunks = (unsigned)((n + kRefitChunk - 1) / kRefitChunk);
  hipLaunchKernelGGL(k_refit_local, dim3(chunks), dim3(kRefitChunk), 0, ctx->stream, (int)n,
                     ctx->order.as<uint32_t>(), ctx->rec.as<float4>(), ctx->box32.as<float>(),
                     ctx->nodes.as<float4>(), ctx->parentInt.as<int32_t>(), ctx->parentLeaf.as<int32_t>(),
                     ctx->flags.as<int32_t>(), ctx->leaves.as<float4>(), ctx->rootBox.as<float>(), esc + L.off[0],
                     escCount, L.cap[0]);
  if (n > 1) {
    for (int k = 0; k < 2; k++)
      hipLaunchKernelGGL(k_refit_mid, dim3((unsigned)((n + kRefitMid[k] - 1) / kRefitMid[k])), dim3(kRefitChunk), 0,
                         ctx->stream, (int)n, kRefitMid[k], (const int32_t *)(esc + L.off[k]),
                         (const uint32_t *)(escCount + k), L.cap[k], esc + L.off[k + 1], escCount + k + 1,
                         L.cap[k + 1], ctx->nodes.as<float4>(), ctx->parentInt.as<int32_t>(),
                         ctx->flags.as<int32_t>(), ctx->rootBox.as<float>());
    hipLaunchKernelGGL(k_refit_cross, dim3((unsigned)((L.cap[2] + kRefitThreads - 1) / kRefitThreads)),
                       dim3(kRefitThreads), 0, ctx->stream, (const int32_t *)(esc + L.off[2]),
                       (const uint32_t *)(escCount + 2), L.cap[2], ctx->nodes.as<float4>(),
                       ctx->parentInt.as<int32_t>(), ctx->flags.as<int32_t>(), ctx->rootBox.as<float>());
  }
  CPT_TRY_HIP(ctx, hipGetLastError());
  ctx->recValid = 0;
  ctx->clusterValid = false;
  return ensureRecords(ctx, true);
}

// The escape lists of the last refit held every crossing node: each counter
// (which also counts entries past its list's bound) is at most the bound.
// Called after the refit has finished (the stream is idle there).
static int checkRefitEscapes(cpt_ctx *ctx, int64_t n, const char *what) {
  const EscapeLists L = escapeLists(n);
  uint32_t cnt[3] = {0, 0, 0};
  CPT_TRY_HIP(ctx, hipMemcpy(cnt, ctx->esc.as<int32_t>() + L.counters, 12, hipMemcpyDeviceToHost));
  if ((int64_t)cnt[0] > L.cap[0] || (int64_t)cnt[1] > L.cap[1] || (int64_t)cnt[2] > L.cap[2]) {
    ctx->built = false;
    return fail(ctx, CPT_ERR_STATE, what, "refit escape list overflow (node boxes incomplete)");
  }
  return CPT_OK;
}

// The clustered tree of the current leaves (CPT_OPT_CLUSTER), and the wide
// records from it when it is chosen.  Returns its time in *ms.
int clusterStage(cpt_ctx *ctx, float *ms) {
  *ms = 0.f;
  const int fmt = recordFormat(ctx);
  if (!(fmt == 1 || fmt == 2) || ctx->nPrims < 2) return CPT_OK;
  hipEvent_t *ev = ctx->ev;
  CPT_TRY_HIP(ctx, hipEventRecord(ev[5], ctx->stream));
  int rc = buildCluster(ctx);
  if (rc) return rc;
  if (ctx->clusterOk) {
    ctx->recValid = 0;
    if ((rc = ensureRecords(ctx))) return rc;
  }
  CPT_TRY_HIP(ctx, hipEventRecord(ev[6], ctx->stream));
  CPT_TRY_HIP(ctx, hipEventSynchronize(ev[6]));
  CPT_TRY_HIP(ctx, hipEventElapsedTime(ms, ev[5], ev[6]));
  return CPT_OK;
}

// Node boxes from the current primitives with the topology, parents and
// leaf order of the last build.
int refitBvh(cpt_ctx *ctx, cpt_build_stats *st) {
  const int64_t n = ctx->nPrims;
  if (!ctx->built) return fail(ctx, CPT_ERR_STATE, "cpt_bvh_refit", "no BVH built");
  if (n == 0) return CPT_OK;
  hipEvent_t *ev = ctx->ev;
  CPT_TRY_HIP(ctx, hipEventRecord(ev[0], ctx->stream));
  int rc = launchRefit(ctx, n);
  if (rc) return rc;
  CPT_TRY_HIP(ctx, hipEventRecord(ev[1], ctx->stream));
  CPT_TRY_HIP(ctx, hipEventSynchronize(ev[1]));
  if ((rc = checkRefitEscapes(ctx, n, "cpt_bvh_refit"))) return rc;
  float ms = 0, cms = 0;
  CPT_TRY_HIP(ctx, hipEventElapsedTime(&ms, ev[0], ev[1]));
  if (ctx->cluster == 1 && (rc = clusterStage(ctx, &cms))) return rc;
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->refit_ms = ms;
    st->total_ms = ms;
    st->cluster_ms = cms;
    st->n_prims = (int32_t)n;
  }
  return CPT_OK;
}

int buildBvh(cpt_ctx *ctx, const double bmin[3], const double bmax[3], cpt_build_stats *st) {
  const int64_t n = ctx->nPrims;
  if (n > (int64_t)0x7ffffffe) return fail(ctx, CPT_ERR_ARG, "cpt_bvh_build", "too many primitives");
  ctx->built = false;
  if (n == 0) {
    ctx->built = true;
    if (st) std::memset(st, 0, sizeof(*st));
    return CPT_OK;
  }
  const int nTiles = (int)((n + kTile - 1) / kTile);
  int rc;
  if ((rc = ensure(ctx, ctx->codesA, n * 4)) || (rc = ensure(ctx, ctx->codesB, n * 4)) ||
      (rc = ensure(ctx, ctx->valsA, n * 4)) || (rc = ensure(ctx, ctx->valsB, n * 4)) ||
      (rc = ensure(ctx, ctx->hist, (size_t)nTiles * 256 * 4)) ||
      (rc = ensure(ctx, ctx->rowTotal, 256 * 4)) || (rc = ensure(ctx, ctx->keys64, n * 8)) ||
      (rc = ensure(ctx, ctx->order, n * 4)) ||
      (rc = ensure(ctx, ctx->nodes, (size_t)(n > 1 ? n - 1 : 1) * 64)) ||
      (rc = ensure(ctx, ctx->parentInt, (size_t)n * 4)) ||
      (rc = ensure(ctx, ctx->parentLeaf, (size_t)n * 4)) ||
      (rc = ensure(ctx, ctx->flags, (size_t)n * 4)) || (rc = ensure(ctx, ctx->leaves, (size_t)n * 48)) ||
      (rc = ensure(ctx, ctx->rootBox, 64)))
    return rc;
  double smin[3], smax[3];
  expandBounds(bmin, bmax, smin, smax);
  hipEvent_t *ev = ctx->ev;
  CPT_TRY_HIP(ctx, hipEventRecord(ev[0], ctx->stream));
  hipLaunchKernelGGL(k_morton, dim3(nTiles), dim3(kSortThreads), 0, ctx->stream,
                     ctx->center64.as<double>(), n, make_double3(smin[0], smin[1], smin[2]),
                     make_double3(smax[0], smax[1], smax[2]), ctx->codesA.as<uint32_t>(),
                     ctx->hist.as<uint32_t>(), nTiles, (uint64_t *)nullptr);
  CPT_TRY_HIP(ctx, hipGetLastError());
  CPT_TRY_HIP(ctx, hipEventRecord(ev[1], ctx->stream));
  if ((rc = sortPasses(ctx, ctx->codesA.as<uint32_t>(), ctx->valsA.as<uint32_t>(),
                       ctx->codesB.as<uint32_t>(), ctx->valsB.as<uint32_t>(), n, 30, true, true,
                       ctx->keys64.as<uint64_t>(), ctx->order.as<uint32_t>())))
    return rc;
  CPT_TRY_HIP(ctx, hipEventRecord(ev[2], ctx->stream));
  const int bs = 256;
  if (n > 1) {
    hipLaunchKernelGGL(k_emit, dim3((unsigned)((n - 1 + bs - 1) / bs)), dim3(bs), 0, ctx->stream,
                       ctx->keys64.as<uint64_t>(), (int)n, ctx->nodes.as<float4>(),
                       ctx->parentInt.as<int32_t>(), ctx->parentLeaf.as<int32_t>());
    CPT_TRY_HIP(ctx, hipGetLastError());
  }
  CPT_TRY_HIP(ctx, hipEventRecord(ev[3], ctx->stream));
  if ((rc = launchRefit(ctx, n))) return rc;
  CPT_TRY_HIP(ctx, hipEventRecord(ev[4], ctx->stream));
  CPT_TRY_HIP(ctx, hipEventSynchronize(ev[4]));
  if ((rc = checkRefitEscapes(ctx, n, "cpt_bvh_build"))) return rc;
  float cms = 0;
  if (ctx->cluster == 1 && (rc = clusterStage(ctx, &cms))) return rc;
  if (st) {
    float a = 0, b = 0, c = 0, d = 0;
    CPT_TRY_HIP(ctx, hipEventElapsedTime(&a, ev[0], ev[1]));
    CPT_TRY_HIP(ctx, hipEventElapsedTime(&b, ev[1], ev[2]));
    CPT_TRY_HIP(ctx, hipEventElapsedTime(&c, ev[2], ev[3]));
    CPT_TRY_HIP(ctx, hipEventElapsedTime(&d, ev[3], ev[4]));
    st->morton_ms = a;
    st->sort_ms = b;
    st->emit_ms = c;
    st->refit_ms = d;
    st->total_ms = a + b + c + d;
    st->cluster_ms = cms;
    st->n_prims = (int32_t)n;
  }
  ctx->built = true;
  return CPT_OK;
}

}  // namespace cpt

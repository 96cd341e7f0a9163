// internal.hpp - host-side context shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/cpt.h"

#define CPT_TRY_HIP(ctx, expr)                                                      \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) return cpt::fail((ctx), CPT_ERR_HIP, #expr, hipGetErrorString(e_)); \
  } while (0)

namespace cpt {

// Grow-only device buffer.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  template <class T> T *as() const { return (T *)p; }
};

// Device material (32 B): colour, alpha_x, alpha_y, eta, type.
struct MatDev {
  float cr, cg, cb, ax, ay, eta;
  int type, pad;
};

struct RenderBufs;

// Frees a denoiser's device memory and detaches it from its context (its
// handle stays valid for cpt_denoiser_destroy; other calls on it fail).
void denoiserRelease(cpt_denoiser *dn);

}  // namespace cpt

struct cpt_ctx {
  int device = 0;
  hipStream_t ownStream = nullptr;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<cpt::DevBuf *> bufs;
  std::vector<cpt_denoiser *> denoisers;  // live denoisers: released by cpt_destroy

  // scene (insertion order)
  int64_t nPrims = 0, nTri = 0;
  int32_t nLights = 0, nSpheres = 0, nMats = 0;
  cpt::DevBuf center64;  // N x 3 double
  cpt::DevBuf box32;     // N x 6 float, outward-rounded padded boxes
  cpt::DevBuf rec;       // N x 3 float4, leaf records in insertion order
  cpt::DevBuf tri64;     // nTri x 9 double
  cpt::DevBuf sph64;     // nSpheres x 4 double
  cpt::DevBuf lgt64;     // nLights x 6 double
  cpt::DevBuf mats;      // (nMats + nLights) x MatDev
  cpt::DevBuf lightIdx;  // nLights x int (primitive ids of the lights)
  std::vector<cpt_light> hostLights;

  // LBVH
  bool built = false;
  cpt::DevBuf codesA, codesB, valsA, valsB, hist, rowTotal;
  cpt::DevBuf keys64, order, nodes, parentInt, parentLeaf, flags, leaves, rootBox;
  cpt::DevBuf nodes4;   // <= N-1 four-wide records of 128 B, level order from the root
  cpt::DevBuf nodes4q;  // the same records quantized, 64 B each
  cpt::DevBuf recNode;  // binary node of each four-wide record (build scratch)
  cpt::DevBuf recLev;   // first slot and width of each record level (build scratch)
  cpt::DevBuf nodes8;   // (N-1) x 128 B quantized eight-wide records
  int recValid = 0;     // bit f: wide records of format f are current
  // clustered (PLOC) binary tree over the same leaves, for the wide records
  cpt::DevBuf plC[2], plB[2], plD[2], plS[2], plNN, plSum, plCtr, plTmp, plSah, nodesP;
  // clustered tree: 0 never, 1 after each build or refit, 2 at the first
  // render after a build or refit; the four-wide records group it when it
  // is the clearly cheaper tree
  int cluster = 2;
  bool clusterValid = false;  // nodesP matches the current leaves and boxes
  bool clusterOk = false;     // nodesP is used: depth within the walks' stacks, lower cost
  int clusterDepth = 0;
  double lbvhCost = 0.0, clusterCost = 0.0;  // summed child-box areas of either tree
  cpt::DevBuf esc;      // refit escape list + its counter
  // heterogeneous medium (density grid), cleared by cpt_scene_upload
  bool hasMedium = false;
  cpt::DevBuf medDensity;
  float medBox[6] = {0, 0, 0, 0, 0, 0};
  float medSigma = 0.f, medAlbedo[3] = {0, 0, 0};
  int medN[3] = {0, 0, 0};
  bool hasVolumeMat = false;  // some material is CPT_M_VOLUME (k_volume runs)
  int traversal = 4;    // 2: binary node walk, 4: four-wide records
  bool quantized = true;  // four-wide walks read the 64 B quantized records
  bool persistent = true;  // four-wide traces as persistent, lane-refilling kernels
  int numCU = 256;
  int64_t raySortMin = 0;  // trace batches this large are reordered (0: never)
  // path-tracer bounce queues this long are reordered (0: never, the
  // default: the walks are faster on the unsorted queue, whose append order
  // keeps neighbouring pixels' paths together)
  int64_t queueSortMin = 0;
  int64_t shadowSortMin = (int64_t)1 << 20;  // shadow queues this long are reordered
  int64_t persistentShadowMin = CPT_PERSISTENT_SHADOW_MIN_DEFAULT;  // shadow rays for the persistent walk
  bool lastTraceSorted = false;

  // traversal I/O staging; cntA: the persistent walks' grab counter; cntB:
  // triangle material ids staged by the scene upload and update
  cpt::DevBuf rayBuf, tBuf, idBuf, cntA, cntB;
  cpt::DevBuf rkA, rkB, rvA, rvB;  // ray-order keys / permutation (ping-pong)
  float lastTraceMs = 0.f;
  hipEvent_t ev[18] = {};

  // abort word in pinned host memory (cpt_request_abort): the persistent
  // walks poll it through abortDev, the render loop between bounces
  uint32_t *abortHost = nullptr;
  const uint32_t *abortDev = nullptr;
  // this device's static scratch of cpt_trace_closest_noalloc
  uint32_t *noAllocScratch = nullptr;

  cpt::RenderBufs *render = nullptr;
};

namespace cpt {

int fail(cpt_ctx *ctx, int code, const char *what, const char *detail);
// Ensures buffer b holds at least `bytes` bytes (contents not preserved).
int ensure(cpt_ctx *ctx, DevBuf &b, size_t bytes);
void releaseAll(cpt_ctx *ctx);

// bvh.hip
int launchMorton(cpt_ctx *ctx, const double smin[3], const double smax[3], uint64_t *keysDev);
// Stable LSD sort of (keys, vals) in 8-bit passes; the result is in (keys,
// vals), or, with valsOut, the sorted values are left where the last pass put
// them (vals or valsTmp) and *valsOut points there.  hist (ceil(n/4096)*256
// words) and rowTotal (256 words): caller scratch, or null for the
// context's own (allocated on demand).
int radixSortPairs(cpt_ctx *ctx, uint32_t *keys, uint32_t *vals, uint32_t *keysTmp,
                   uint32_t *valsTmp, int64_t n, int bits, bool identityVals, uint32_t **valsOut = nullptr,
                   uint32_t *hist = nullptr, uint32_t *rowTotal = nullptr);
int buildBvh(cpt_ctx *ctx, const double bmin[3], const double bmax[3], cpt_build_stats *st);
int refitBvh(cpt_ctx *ctx, cpt_build_stats *st);
// Traversal record format: 0 binary nodes, 1 four-wide FP32, 2 four-wide
// quantized, 3 eight-wide quantized; ensureRecords builds that of the
// context's traversal options if stale (lbvhOnly: from the LBVH's boxes even
// where the clustered tree is on).
int recordFormat(const cpt_ctx *ctx);
// cluster.hip: builds ctx->nodesP from the current leaf order and boxes.
int buildCluster(cpt_ctx *ctx);
// bvh.hip: builds the clustered tree and, when it is chosen, the wide
// records from it; *ms gets the time.
int clusterStage(cpt_ctx *ctx, float *ms);
int ensureRecords(cpt_ctx *ctx, bool lbvhOnly = false);
int packTriangles(cpt_ctx *ctx, const double *tri64Dev, const int32_t *matDev, int64_t nt,
                  int64_t base);
void expandBounds(const double bmin[3], const double bmax[3], double smin[3], double smax[3]);

// trace.hip
// Scratch of one traversal call: ray-order keys / permutation (n words
// each), radix histograms (ceil(n/4096)*256 + 256 words) and the persistent
// walk's grab counter.  Null: the context's own buffers, allocated on demand.
struct TraceScratch {
  uint32_t *keysA, *keysB, *valsA, *valsB, *hist, *rowTotal, *grab;
};
int traceDevice(cpt_ctx *ctx, const float *raysDev, int64_t n, float *tDev, int32_t *idDev,
                int32_t *nodesDev, int32_t *primsDev, const TraceScratch *scratch = nullptr);
// Device address of this GPU's static no-allocation scratch (set up once
// per device by cpt_create).
int noAllocScratchInit(cpt_ctx *ctx);

// render.hip
void renderRelease(cpt_ctx *ctx);

}  // namespace cpt

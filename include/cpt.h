/*
 * cpt.h - C ABI of the MI355X (gfx950) LBVH build, closest-hit traversal and
 * path-trace integrator.
 *
 * The library replaces the device side of Baoste/CudaPathTracing:
 *   - Device::buildBVH            CudaPathTracing/Device.cuh:169-195
 *       (generateMortonCodes Device.cuh:134, thrust::sort_by_key Device.cuh:187,
 *        generateBVH/generateHierarchy BVH.cuh:174, constructAABB BVH.cuh:243)
 *   - traverseIterative           CudaPathTracing/BVH.cuh:282-332
 *   - render<<<>>> (STATIC mode)  CudaPathTracing/Render.cu:16-261
 *   - the per-object construction kernels allocate*OnDevice
 *                                 CudaPathTracing/Device.cuh:29-79
 * Host interfaces kept from the reference (IniParser, Scene::init, Camera)
 * sit on top of this ABI (cpt_scene_load_ini, cpt_camera_from_desc).
 *
 * Conventions
 *   - Every entry point returns CPT_OK (0) or a negative CPT_ERR_* code and
 *     never exits the process; cpt_last_error() gives the message.
 *   - One context is used by one host thread at a time.  All device work is
 *     ordered on the context's stream (cpt_set_stream may replace it).
 *   - Pointers are host pointers unless CPT_DEVICE_PTRS is passed in flags;
 *     with host pointers the call is synchronous.
 *   - Primitive IDs are insertion indices in the reference's order
 *     (Scene.cuh:74-89): lights, then spheres, then triangles (floors, which
 *     split into (lt,lb,rt),(rt,lb,rb) per Device.cuh:42-43, then meshes).
 */
#ifndef CPT_H
#define CPT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define CPT_OK 0
#define CPT_ERR_ARG (-1)     /* invalid argument or size                   */
#define CPT_ERR_HIP (-2)     /* HIP runtime error                          */
#define CPT_ERR_STATE (-3)   /* call order (e.g. trace before build)       */
#define CPT_ERR_OOM (-4)     /* device allocation failed                   */
#define CPT_ERR_IO (-5)      /* file could not be read / parsed            */
#define CPT_ERR_ABORTED (-6) /* host abort flag stopped a persistent kernel */

#define CPT_DEVICE_PTRS 1 /* flags bit: buffers are device pointers */

/* ---- material / object types ------------------------------------------ */
/* MaterialType, CudaPathTracing/rtmath.cuh:15, plus two new entries. */
#define CPT_M_NONE 0
#define CPT_M_LIGHT 1
#define CPT_M_OPAQUE 2               /* GGX conductor, n=2.63 k=9.08    */
#define CPT_M_SPECULAR_DIELECTRIC 3  /* smooth glass, eta 1.51          */
#define CPT_M_ROUGH_DIELECTRIC 4     /* GGX glass                       */
#define CPT_M_VOLUME 5               /* homogeneous sigma=25 medium     */
#define CPT_M_LAMBERT 6              /* new: cosine-sampled diffuse     */
#define CPT_M_GGX 7                  /* new: isotropic GGX, Schlick F0  */

/* ObjectType, CudaPathTracing/Hittable.cuh:8. */
#define CPT_OBJ_NONE 0
#define CPT_OBJ_SPHERE 1
#define CPT_OBJ_LIGHT 2
#define CPT_OBJ_TRIANGLE 3

/* Material, CudaPathTracing/Material.cuh:5-22.  For CPT_M_GGX `color` is F0
 * and alpha_x is the roughness alpha; for CPT_M_LAMBERT `color` is albedo. */
typedef struct cpt_material {
  double color[3];
  double alpha_x;
  double alpha_y;
  double eta; /* 1.51 in the reference */
  int32_t type;
  int32_t pad_;
} cpt_material;

/* Light quad, CudaPathTracing/Hittable.cuh:346-399 and IniParser.h:26-34. */
typedef struct cpt_light {
  double center[3];
  double width;
  double height;
  double normal[3];
  double color[3]; /* emitted radiance (ini colour / 255) */
  int32_t visible;
  int32_t pad_;
} cpt_light;

/* Sphere, CudaPathTracing/Hittable.cuh:10-55. */
typedef struct cpt_sphere {
  double center[3];
  double radius;
  int32_t material; /* index into the material table */
  int32_t pad_;
} cpt_sphere;

/* Camera as constructed by Camera::Camera, CudaPathTracing/Camera.cu:5-31. */
typedef struct cpt_camera_desc {
  double background[3];
  double look_from[3];
  double look_at[3];
  double vfov_deg;
  double aspect;
  int32_t width;
  int32_t pad_;
} cpt_camera_desc;

/* Derived camera (Camera.cu:12-30): height = max(1, (int)(width/aspect)),
 * pixel(i,j) = pixel_center + (i + rx - 0.5) delta_h + (j + ry - 0.5) delta_v,
 * ray direction = pixel - look_from (not normalised), row j = 0 is the
 * bottom of the image. */
typedef struct cpt_camera {
  double look_from[3];
  double pixel_center[3];
  double delta_h[3];
  double delta_v[3];
  double background[3];
  int32_t width;
  int32_t height;
} cpt_camera;

/* ---- context ------------------------------------------------------------ */
typedef struct cpt_ctx cpt_ctx;

int cpt_create(cpt_ctx **out, int hip_device);
/* Also releases the device memory of the context's denoisers: one destroyed
 * afterwards only frees its handle (calls on it meanwhile return
 * CPT_ERR_ARG). */
void cpt_destroy(cpt_ctx *ctx);
const char *cpt_last_error(const cpt_ctx *ctx);
/* Use an external hipStream_t (e.g. torch's current stream); NULL restores
 * the context's own stream.  The own stream is a blocking stream: work on
 * it is ordered with work on the device's null stream in both directions
 * (so callers whose work runs on the null stream, such as PyTorch's default
 * stream, whose handle is NULL, need no extra synchronisation); work on
 * other non-blocking streams is not ordered with it. */
int cpt_set_stream(cpt_ctx *ctx, void *hip_stream);
int cpt_synchronize(cpt_ctx *ctx);
/* Abort: asks the running (or next) call on ctx to stop.  The persistent
 * traversal walks take no more rays, the path tracer stops between bounces,
 * and the call returns CPT_ERR_ABORTED with its outputs undefined; every
 * later call does too until cpt_clear_abort.  cpt_request_abort may be called
 * from any thread, also while another thread is inside a call on ctx (it only
 * stores to a word of pinned host memory that the kernels poll). */
int cpt_request_abort(cpt_ctx *ctx);
int cpt_clear_abort(cpt_ctx *ctx);
/* Library build string (arch, version). */
const char *cpt_version(void);
/* Options.  CPT_OPT_TRAVERSAL_WIDTH: 4 (default) walks four-wide records
 * built from the binary LBVH (each spans two binary levels); 2 walks the
 * binary nodes.  Results are identical up to exact-t ties.
 * CPT_OPT_RAY_SORT_MIN: cpt_trace_closest batches of at least this many rays
 * are traced in the order of a key (direction octant, origin Morton cell);
 * results are written in the caller's order.  0 (default) traces them in the
 * caller's order: the persistent walk's lanes stay busy in any order, and
 * its chunks then read rays and write results contiguously (2^25 random
 * rays into 10M triangles: 6.9 ms unordered, 8.1 ms ordered plus 1.1 ms of
 * sorting). */
#define CPT_OPT_TRAVERSAL_WIDTH 1
#define CPT_OPT_RAY_SORT_MIN 2
/* CPT_OPT_PERSISTENT: 1 (default) runs four-wide cpt_trace_closest
 * traversals as persistent kernels whose lanes take a new ray when theirs is
 * finished; 0 launches one lane per ray.  Results are identical. */
#define CPT_OPT_PERSISTENT 3
/* CPT_OPT_PERSISTENT_SHADOW: cpt_render walks a bounce's shadow rays with
 * the persistent lane-refilling four-wide kernel (with CPT_OPT_PERSISTENT 1),
 * in append order, when there are at least this many of them (default
 * CPT_PERSISTENT_SHADOW_MIN_DEFAULT; 1: always; 0: never); with a
 * participating medium the unoccluded rays are then ratio-tracked by a
 * second kernel.  Otherwise one lane per shadow ray (ordered by origin cell
 * from CPT_OPT_SHADOW_SORT_MIN rays).  Results are identical. */
#define CPT_OPT_PERSISTENT_SHADOW 8
#define CPT_PERSISTENT_SHADOW_MIN_DEFAULT (1 << 22)
/* CPT_OPT_QUANTIZED: 1 (default) makes four-wide walks read 64 B records
 * whose child boxes are 8-bit planes on a power-of-two grid, rounded outward
 * (each contains the FP32 box); 0 reads the 128 B FP32 records.  Hits are
 * identical up to exact-t ties. */
#define CPT_OPT_QUANTIZED 4
/* CPT_OPT_QUEUE_SORT_MIN: cpt_render traces a bounce queue of at least this
 * many paths in the order of an 8-bit ray key (direction octant, origin
 * cell); 0 (default) never.  Radiance is unchanged. */
#define CPT_OPT_QUEUE_SORT_MIN 5
/* CPT_OPT_SHADOW_SORT_MIN: cpt_render traces a bounce's shadow rays of at
 * least this many in the order of their origin Morton cell; 0 disables it;
 * default 2^20.  Radiance is unchanged. */
#define CPT_OPT_SHADOW_SORT_MIN 6
/* CPT_OPT_CLUSTER: a second binary tree over the same leaves, built by
 * locally-ordered clustering of the Morton-ordered leaf boxes; the four-wide
 * records group its boxes when its surface-area cost (summed child-box
 * areas) is at least 5% below the LBVH's.  2 (default): built by the first
 * cpt_render / cpt_render_views after a build or refit (a render traces
 * enough rays to repay it); 1: built by every cpt_bvh_build / cpt_bvh_refit
 * (timed in cluster_ms); 0: never.  cpt_bvh_export always returns the LBVH.
 * Hits are identical up to exact-t ties. */
#define CPT_OPT_CLUSTER 7
int cpt_set_option(cpt_ctx *ctx, int32_t option, int32_t value);

/* ---- scene upload ------------------------------------------------------- */
/* Replaces Scene::add* + allocate*OnDevice (Scene.cuh:175-270,
 * Device.cuh:29-79): one bulk upload, IDs = insertion order
 * [lights | spheres | triangles].  tri_xyz holds n_tri*9 doubles
 * (p0,p1,p2); tri_mat holds n_tri material indices (NULL = material 0).
 * Lights receive implicit CPT_M_LIGHT materials.  Host pointers only. */
int cpt_scene_upload(cpt_ctx *ctx, const cpt_material *materials,
                     int32_t n_materials, const cpt_light *lights,
                     int32_t n_lights, const cpt_sphere *spheres,
                     int32_t n_spheres, const double *tri_xyz,
                     const int32_t *tri_mat, int64_t n_tri);

/* Number of primitives of the current scene. */
int64_t cpt_scene_count(const cpt_ctx *ctx);

/* Parses a reference scene_ini file (IniParser::Parse, IniParser.cpp:3-96)
 * and runs Scene::init (Scene.cuh:72-145): adds lights, spheres, floors and
 * [Mesh] OBJ files in that order, uploads the scene and writes the
 * reference's Morton bounds (untranslated mesh-vertex box, Device.cuh:174-181
 * is applied later by cpt_bvh_build) and camera.  Any out pointer may be
 * NULL.  A mesh file that cannot be read adds no triangles (ObjLoader.cuh:
 * 65-70). */
int cpt_scene_load_ini(cpt_ctx *ctx, const char *path, double scene_min[3],
                       double scene_max[3], cpt_camera_desc *camera_out);

/* Host-only parse of a scene_ini file into caller buffers (no device
 * work).  cpt_ini_counts gives {materials, lights, spheres, triangles};
 * cpt_ini_read fills arrays of those sizes (any pointer may be NULL) plus the
 * Morton bounds and the camera.  Returns CPT_ERR_IO for an unreadable or
 * malformed file. */
int cpt_ini_counts(const char *path, int64_t counts[4]);
int cpt_ini_read(const char *path, cpt_material *materials, cpt_light *lights,
                 cpt_sphere *spheres, double *tri_xyz, int32_t *tri_mat,
                 double scene_min[3], double scene_max[3],
                 cpt_camera_desc *camera_out);

/* ---- LBVH --------------------------------------------------------------- */
typedef struct cpt_build_stats {
  float morton_ms;
  float sort_ms;
  float emit_ms;
  float refit_ms;
  float total_ms;   /* the four LBVH stages above */
  int32_t n_prims;
  float cluster_ms; /* after them: the clustered tree and its records (CPT_OPT_CLUSTER 1) */
} cpt_build_stats;

/* Device::buildBVH (Device.cuh:169-195).  scene_min/max are the reference's
 * minBoundary/maxBoundary; the +-1 expansion and the +-20 fallback
 * (Device.cuh:174-181) are applied inside.  stats may be NULL. */
int cpt_bvh_build(cpt_ctx *ctx, const double scene_min[3],
                  const double scene_max[3], cpt_build_stats *stats);

/* Copies the built hierarchy to host memory (any pointer may be NULL):
 *   keys[n]      sorted 64-bit keys (code<<32 | primitive id), BVH.cuh:39-40
 *   order[n]     sorted primitive ids (d_objsIdx after the sort)
 *   child_a/b[n-1] children of internal node i in Karras numbering;
 *                a leaf is encoded as ~leaf_index (negative)
 *   boxes[(n-1)*12] for internal node i: child A box (min xyz, max xyz) then
 *                child B box, FP32 rounded outward from the FP64 box
 *   root_box[6]  FP32 box of the whole tree */
int cpt_bvh_export(cpt_ctx *ctx, uint64_t *keys, uint32_t *order,
                   int32_t *child_a, int32_t *child_b, float *boxes,
                   float *root_box);

/* Morton codes only (generateMortonCodes, Device.cuh:134-142): writes the
 * unsorted 64-bit keys of the current scene for the given bounds
 * (expansion rule applied as in cpt_bvh_build). */
int cpt_bvh_morton(cpt_ctx *ctx, const double scene_min[3],
                   const double scene_max[3], uint64_t *keys_out, int flags);

/* Stable radix sort of 32-bit keys with 32-bit values (the sort stage of
 * cpt_bvh_build, exposed for testing): sorts n (key,value) pairs by the low
 * `bits` key bits (1..32).  In place on host or device buffers. */
int cpt_radix_sort_pairs(cpt_ctx *ctx, uint32_t *keys, uint32_t *values,
                         int64_t n, int32_t bits, int flags);

/* ---- traversal ---------------------------------------------------------- */
/* Closest hit for n rays (traverseIterative, BVH.cuh:282-332, with
 * Triangle/Sphere/Light::hit, Hittable.cuh:27-140,367-398).
 * rays: n x 8 floats {ox,oy,oz,tmin, dx,dy,dz,tmax}.  t_out: hit distance
 * (recomputed in FP64 for the winning primitive, then rounded), id_out:
 * primitive id or -1.  Either output may be NULL. */
int cpt_trace_closest(cpt_ctx *ctx, const float *rays, int64_t n,
                      float *t_out, int32_t *id_out, int flags);

/* No counterpart in the reference: cpt_trace_closest on device buffers that
 * allocates nothing, for callers that cannot call hipMalloc (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).
 * Its scratch (ray-order keys and permutation, radix histograms, the
 * persistent walk's grab counter) is a static device array of the library,
 * one per GPU (about 17 MB), held until the process ends; the BVH and its
 * traversal records must be current (cpt_bvh_build / cpt_bvh_refit make them;
 * CPT_ERR_STATE otherwise).  At most CPT_NOALLOC_MAX_RAYS rays per call
 * (CPT_ERR_ARG beyond).  Calls that overlap on one GPU, from any context,
 * race on that array.  Outputs equal cpt_trace_closest's. */
#define CPT_NOALLOC_MAX_RAYS (1 << 20)
int cpt_trace_closest_noalloc(cpt_ctx *ctx, const float *rays_dev, int64_t n, float *t_dev, int32_t *id_dev);

/* Same traversal with per-ray work counters (nodes visited, primitives
 * tested), used for the algorithmic-bytes model. */
int cpt_trace_count(cpt_ctx *ctx, const float *rays, int64_t n,
                    int32_t *nodes_out, int32_t *prims_out, int flags);

/* Duration of the last cpt_trace_closest / cpt_trace_count call on the
 * device (ms: ray ordering + traversal), measured with HIP events on the
 * launching stream. */
float cpt_last_trace_ms(const cpt_ctx *ctx);
/* The same split: total, ray ordering (0 when not sorted) and the traversal
 * kernel alone (ms).  Any pointer may be NULL. */
int cpt_last_trace_stats(const cpt_ctx *ctx, float *total_ms, float *sort_ms,
                         float *kernel_ms);

/* ---- path tracing ------------------------------------------------------- */
/* Render profiles: the reference's RenderType (Render.cuh:11, Render.cu:27-38,
 * 189-204, 232-251) plus the configurable one of the workloads.  Every
 * reference profile samples NEE over width x width of each light
 * (Render.cu:170-171); CONFIG uses width x height. */
#define CPT_PROFILE_REFERENCE 0 /* STATIC and REAL_TIME_NOT_SAMPLE: RR 0.7   */
                                /* at every vertex, no depth cap             */
#define CPT_PROFILE_CONFIG 1    /* max_depth cap, RR from rr_depth on       */
#define CPT_PROFILE_REAL_TIME 2 /* REAL_TIME: at most 3 vertices, no RR     */
                                /* (the reference seeds with 1000*t: pass   */
                                /* that as the seed)                         */
#define CPT_PROFILE_NORMAL 3    /* NORMAL: 1 spp, at most 2 vertices; rgb = */
                                /* (n + 0.5) / 2 of the first vertex's       */
                                /* face-forwarded normal (0 if none)         */
#define CPT_PROFILE_DEPTH 4     /* DEPTH: 1 spp, at most 2 vertices; rgb =   */
                                /* clamp01(distance / 50) (1 if none)        */

typedef struct cpt_render_params {
  uint64_t seed;
  int32_t spp;
  int32_t max_depth; /* bounces (CONFIG profile); ignored for REFERENCE */
  int32_t rr_depth;  /* first bounce with Russian roulette (CONFIG)    */
  int32_t profile;
  double rr_prob;       /* survival probability, 0.7 in Render.cu:202 */
  int32_t sample_base;  /* index of the first sample (progressive)    */
  int32_t want_aov;     /* fill albedo/normal/depth outputs           */
  int32_t paths_per_batch; /* 0 = automatic: up to 2^28 / lights, whole
                            * samples spread evenly over the fewest batches
                            * (about 290 B of device memory per path)    */
  int32_t flags;           /* CPT_RENDER_COUNT_WORK: count extension-  */
                           /* kernel node visits and primitive tests   */
} cpt_render_params;

#define CPT_RENDER_COUNT_WORK 1
/* params.flags bit (with want_aov): rgb receives h*w*10 floats per view,
 * per pixel the radiance and then the 7 AOVs (albedo rgb, normal xyz,
 * depth) -- the learned denoiser's input layout -- and aov must be NULL. */
#define CPT_RENDER_PACK_AOV 2

typedef struct cpt_render_stats {
  double ms;
  double rays; /* extension + shadow rays traced */
  double paths;
  double extend_ms; /* sum of extension-kernel durations */
  double shadow_ms;
  double shade_ms;
  int32_t bounces;
  int32_t pad_;
  double extend_rays;  /* extension rays traced                        */
  double extend_nodes; /* nodes visited by extension rays (COUNT_WORK) */
  double extend_prims; /* primitives tested by extension rays (ditto)  */
  double extend_wave_nodes; /* 64 x each wave's largest node count, summed
                               (ditto): extend_nodes / extend_wave_nodes is
                               the lane utilisation of the walk */
  double shadow_rays;       /* shadow rays traced */
  double shadow_nodes;      /* the same three for the shadow walks */
  double shadow_prims;
  double shadow_wave_nodes;
} cpt_render_stats;

/* render<<<>>> (Render.cu:16-261) as a wavefront pipeline.  rgb: h*w*3
 * floats, mean radiance per pixel, row 0 = bottom (Render.cu:220-222).
 * aov (may be NULL): h*w*7 floats (albedo rgb, normal xyz, depth) averaged
 * over samples at the first non-emitter hit. */
int cpt_render(cpt_ctx *ctx, const cpt_camera *camera,
               const cpt_render_params *params, float *rgb, float *aov,
               cpt_render_stats *stats, int flags);

/* Several views of one scene in one wavefront pipeline (no reference
 * counterpart; the config-5 workload renders 32 views at once).  All
 * cameras have one size; rgb: nviews*h*w*3, aov (may be NULL):
 * nviews*h*w*7.  The random stream of a pixel is keyed by its index over all
 * views (view*h*w + pixel), so view 0 equals cpt_render. */
int cpt_render_views(cpt_ctx *ctx, const cpt_camera *cameras, int32_t nviews,
                     const cpt_render_params *params, float *rgb, float *aov,
                     cpt_render_stats *stats, int flags);

/* Camera::Camera (Camera.cu:5-31) in FP64 on the host. */
int cpt_camera_from_desc(const cpt_camera_desc *desc, cpt_camera *out);

/* pic2RGBW / STATIC output (Render.cu:232-260, 462-481): sqrt gamma and
 * uchar(254.99*clamp01(x)); rgba8 gets h*w*4 bytes.  Host pointers. */
int cpt_tonemap_rgba8(const float *rgb, int32_t width, int32_t height,
                      uint8_t *rgba8);

/* Output files (the reference writes none; its image goes to a GL window).
 * cpt_write_pfm: the float framebuffer rgb (h*w*3, row 0 = bottom) as a PFM
 * ("PF", scale -1 = little-endian, rows bottom to top).
 * cpt_write_png: the reference's displayed bytes as an 8-bit RGBA PNG (top
 * row first): gamma 1 gives pic2RGBW / STATIC (Render.cu:462-481, 232-238),
 * uchar(254.99*clamp01(sqrt(x))); gamma 0 the NORMAL / DEPTH display
 * (Render.cu:239-258), uchar(254.99*clamp01(x)).  Host pointers; CPT_ERR_IO
 * when the file cannot be written. */
int cpt_write_pfm(const char *path, const float *rgb, int32_t width, int32_t height);
int cpt_write_png(const char *path, const float *rgb, int32_t width, int32_t height, int32_t gamma);

/* ---- heterogeneous medium (no reference counterpart; the reference has only
 * the homogeneous sigma=25 sphere hack of Render.cu:64-135) -------------- */
/* A density grid density[ix + nx*(iy + ny*iz)] in [0,1] over the box
 * [bmin, bmax]; sigma_t(x) = sigma_max * density(voxel of x), nearest-voxel
 * lookup; isotropic phase; single-scattering albedo `albedo`.  Free flights
 * are sampled by delta tracking with majorant sigma_max, shadow-ray
 * transmittance by ratio tracking.  n = 0 (or density NULL) removes the
 * medium.  Call after cpt_scene_upload. */
int cpt_scene_set_medium(cpt_ctx *ctx, const float *density, int32_t nx,
                         int32_t ny, int32_t nz, const double bmin[3],
                         const double bmax[3], double sigma_max,
                         const double albedo[3]);
/* Seeded 4-octave fBm value-noise density grid (n^3 floats in [0,1]). */
int cpt_gen_fbm_grid(int32_t n, uint64_t seed, float *density);

/* ---- non-learned denoiser ---------------------------------------------- */
/* Separable joint bilateral filter, gaussianSeparate (Render.cu:353-398)
 * horizontal then vertical (main.cu:100-104), radius 16.  rgb: w*h*3,
 * aov: w*h*7 (albedo, normal, depth) as written by cpt_render, out: w*h*3.
 * The reference's defaults are sigmaG 8, sigmaR 5, sigmaN 0.1, sigmaD 0.2
 * (Window.cu:11-14). */
int cpt_bilateral(cpt_ctx *ctx, const float *rgb, const float *aov, float *out,
                  int32_t w, int32_t h, float sigma_g, float sigma_r,
                  float sigma_n, float sigma_d, int flags);

/* ---- dynamic scenes ------------------------------------------------------ */
/* Replaces n triangles starting at triangle index `first` (0 = first
 * triangle after the lights and spheres) with new vertices, keeping their
 * materials; the primitive count is unchanged.  Host pointer. */
int cpt_scene_update_triangles(cpt_ctx *ctx, int64_t first, const double *tri_xyz,
                               int64_t n);
/* Recomputes every node box of the last built hierarchy from the current
 * primitives without re-sorting (the per-frame rebuild of Scene::Update,
 * Scene.cuh:147-173, reduced to a refit).  The topology keeps the Morton
 * order of the last cpt_bvh_build. */
int cpt_bvh_refit(cpt_ctx *ctx, cpt_build_stats *stats);

/* No counterpart in the reference: which tree the four-wide records group
 * (in_use 1: the clustered tree of CPT_OPT_CLUSTER, 0: the LBVH), the
 * clustered tree's depth, and the two trees' surface-area costs (summed
 * child-box areas) from the last build or refit.  Any pointer may be NULL. */
int cpt_bvh_cluster_info(const cpt_ctx *ctx, int32_t *in_use, int32_t *depth, double *lbvh_cost,
                         double *cluster_cost);

/* Temporal accumulation for a moving camera: addPrevious (Render.cu:400-421)
 * with Camera::getPrePositionInImg (Camera.cuh:104-132), FP64.  rgb,
 * prev_rgb, out: w*h*3 (w, h from camera).  position: w*h*3 first-hit
 * positions (x + y > 999999 marks a miss), or NULL to derive them from the
 * mean first-hit distance in aov (w*h*7, channel 6; 0 = miss) along each
 * pixel-centre ray.  prev: the previous frame's camera.  alpha = 0.4 when
 * moving, else 0.1; out = rgb*alpha + prev_rgb[reprojected]*(1-alpha). */
int cpt_temporal_blend(cpt_ctx *ctx, const float *rgb, const float *prev_rgb,
                       const float *position, const float *aov,
                       const cpt_camera *camera, const cpt_camera_desc *prev,
                       int32_t moving, float *out, int flags);

/* ---- learned denoiser (no reference counterpart; replaces the role of the
 * joint bilateral filter of Render.cu:353-460) ---------------------------- */
/* 5-level U-Net (channels 32,64,64,96,128; 3x3 conv + ReLU, 2x2 average
 * pooling, nearest upsampling + skip concatenation), BF16 weights and
 * activations, FP32 accumulation on MFMA, FP32 master weights + Adam.
 * Input per pixel: 10 floats (radiance rgb, albedo rgb, normal xyz, depth);
 * output: denoised radiance rgb.  Images are padded to multiples of 16. */
typedef struct cpt_denoiser cpt_denoiser;
int cpt_denoiser_create(cpt_ctx *ctx, int32_t max_batch, int32_t max_h,
                        int32_t max_w, uint64_t seed, cpt_denoiser **out);
void cpt_denoiser_destroy(cpt_denoiser *dn);
int64_t cpt_denoiser_num_params(const cpt_denoiser *dn);
int cpt_denoiser_get_params(cpt_denoiser *dn, float *out);
int cpt_denoiser_set_params(cpt_denoiser *dn, const float *in);
/* Layer table: per layer Cin, Cout and the offsets of its weights
 * ([Cout][3][3][Cin]) and biases in the packed parameter vector; returns the
 * number of layers. */
int cpt_denoiser_layers(const cpt_denoiser *dn, int32_t *cin, int32_t *cout,
                        int64_t *w_off, int64_t *b_off);
/* Forward, L1 loss on log(1+radiance), backward; Adam step when lr > 0.
 * loss_out / grads_out (packed like the parameters) may be NULL. */
int cpt_denoiser_train_step(cpt_denoiser *dn, const float *in10,
                            const float *target_rgb, int32_t b, int32_t h,
                            int32_t w, float lr, float *loss_out,
                            float *grads_out, int flags);
int cpt_denoise(cpt_denoiser *dn, const float *in10, float *out_rgb, int32_t b,
                int32_t h, int32_t w, int flags);
/* ms of the last cpt_denoise and train step; forward FLOPs of one pass. */
int cpt_denoiser_stats(const cpt_denoiser *dn, float *fwd_ms, float *step_ms,
                       double *fwd_flops);
/* One 3x3 'same' convolution on host bf16 data (bit patterns), Cin and Cout
 * multiples of 32.  mode 0: out[B,H,W,Cout] = conv(x, w) + bias; mode 1:
 * out[Cout,3,3,Cin] = weight gradient from dy and x; mode 2: out[B,H,W,Cin]
 * = input gradient from dy and w. */
int cpt_conv3x3(cpt_ctx *ctx, int mode, const uint16_t *x, const uint16_t *w,
                const float *bias, const uint16_t *dy, float *out, int32_t B,
                int32_t H, int32_t W, int32_t Cin, int32_t Cout);

/* ---- procedural scenes (host, no reference counterpart) ---------------- */
/* Icosphere with 20*4^level triangles of radius r around c. */
int64_t cpt_gen_icosphere_count(int32_t level);
int cpt_gen_icosphere(int32_t level, const double c[3], double r,
                      double amp, uint64_t noise_seed, double *tri_xyz);
/* n triangles: centroids U[0,1]^3, equilateral, side exp(mu + sigma z),
 * uniform random orientation. */
int cpt_gen_triangle_soup(int64_t n, uint64_t seed, double mu, double sigma,
                          double *tri_xyz);
/* n rays, origins U[lo,hi]^3, directions uniform on the sphere. */
int cpt_gen_random_rays(int64_t n, uint64_t seed, double lo, double hi,
                        float tmin, float tmax, float *rays);
/* Vertex bounding box of n triangles (the reference's mesh bounds). */
int cpt_tri_bounds(const double *tri_xyz, int64_t n, double mn[3],
                   double mx[3]);

#ifdef __cplusplus
}
#endif
#endif /* CPT_H */

// scene_dropin.cpp - the reference-side binding: the Scene class that
// replaces Scene.cuh's device members (Scene.cuh:72-145: init from a
// scene_ini file, BVH build) and main.cu's STATIC frame (Render.cu:16-261,
// pic2RGBW Render.cu:462-481), written as a reference maintainer would,
// in plain C++ against include/cpt.h and linked to libcpt.so.  No HIP or
// CUDA header is included: the library is the only GPU code.
//
// usage: scene_dropin <scene.ini> <spp> <width|0> <out.pfm> <out.png>
// Renders the file's camera (width 0 keeps the file's width) in the
// REFERENCE profile with seed 3317, writes the radiance as PFM and the
// displayed bytes as PNG, and prints one JSON line with the bounds, the
// build and render statistics and the mean radiance.
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "cpt.h"

struct double3 {
  double x, y, z;
};
static double3 make_double3(double x, double y, double z) { return double3{x, y, z}; }

class Scene {
  cpt_ctx *ctx_ = nullptr;

 public:
  cpt_camera_desc camDesc{};
  double3 minBoundary{}, maxBoundary{};
  cpt_build_stats buildStats{};

  Scene() {
    if (cpt_create(&ctx_, 0) != CPT_OK) throw std::runtime_error("no HIP device");
  }
  ~Scene() { cpt_destroy(ctx_); }
  Scene(const Scene &) = delete;
  Scene &operator=(const Scene &) = delete;

  void init(const std::string &iniPath) {
    double mn[3], mx[3];
    if (cpt_scene_load_ini(ctx_, iniPath.c_str(), mn, mx, &camDesc) != CPT_OK)
      throw std::runtime_error(cpt_last_error(ctx_));
    minBoundary = make_double3(mn[0], mn[1], mn[2]);
    maxBoundary = make_double3(mx[0], mx[1], mx[2]);
    if (cpt_bvh_build(ctx_, mn, mx, &buildStats) != CPT_OK) throw std::runtime_error(cpt_last_error(ctx_));
  }

  // main.cu's STATIC frame: RGB radiance and the displayed RGBA bytes
  cpt_render_stats renderStatic(int spp, std::vector<float> &rgb, std::vector<uint8_t> &rgba, cpt_camera &cam) {
    if (cpt_camera_from_desc(&camDesc, &cam) != CPT_OK) throw std::runtime_error("invalid camera");
    cpt_render_params p{};
    p.seed = 3317;
    p.spp = spp;
    p.profile = CPT_PROFILE_REFERENCE;
    p.rr_prob = 0.7;
    rgb.resize(size_t(cam.width) * cam.height * 3);
    rgba.resize(size_t(cam.width) * cam.height * 4);
    cpt_render_stats st{};
    if (cpt_render(ctx_, &cam, &p, rgb.data(), nullptr, &st, 0) != CPT_OK)
      throw std::runtime_error(cpt_last_error(ctx_));
    if (cpt_tonemap_rgba8(rgb.data(), cam.width, cam.height, rgba.data()) != CPT_OK)
      throw std::runtime_error("tonemap");
    return st;
  }
};

int main(int argc, char **argv) {
  if (argc != 6) {
    std::fprintf(stderr, "usage: %s <scene.ini> <spp> <width|0> <out.pfm> <out.png>\n", argv[0]);
    return 2;
  }
  try {
    Scene scene;
    scene.init(argv[1]);
    const int spp = std::atoi(argv[2]), width = std::atoi(argv[3]);
    if (width > 0) scene.camDesc.width = width;
    std::vector<float> rgb;
    std::vector<uint8_t> rgba;
    cpt_camera cam{};
    const cpt_render_stats st = scene.renderStatic(spp, rgb, rgba, cam);
    if (cpt_write_pfm(argv[4], rgb.data(), cam.width, cam.height) != CPT_OK ||
        cpt_write_png(argv[5], rgb.data(), cam.width, cam.height, 1) != CPT_OK) {
      std::fprintf(stderr, "cannot write %s / %s\n", argv[4], argv[5]);
      return 1;
    }
    double sum = 0.0;
    for (float v : rgb) sum += v;
    std::printf(
        "{\"width\": %d, \"height\": %d, \"spp\": %d, \"min\": [%.17g, %.17g, %.17g], \"max\": [%.17g, %.17g, "
        "%.17g], \"build_ms\": %.4f, \"render_ms\": %.3f, \"rays\": %.0f, \"mean\": %.9g}\n",
        cam.width, cam.height, spp, scene.minBoundary.x, scene.minBoundary.y, scene.minBoundary.z,
        scene.maxBoundary.x, scene.maxBoundary.y, scene.maxBoundary.z, scene.buildStats.total_ms, st.ms, st.rays,
        sum / double(rgb.size()));
    return 0;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "scene_dropin: %s\n", e.what());
    return 1;
  }
}

// oracle_scene.hpp - FP64 scene of the CPU oracle (TEST INFRASTRUCTURE ONLY).
//
// One record per primitive in the reference's insertion order (lights,
// spheres, triangles; Scene.cuh:74-89) holding what Hittable.cuh keeps:
// type, centre (Hittable.cuh:407), padded AABB (Hittable.cuh:406), geometry,
// material.
#pragma once
#include <vector>

#include "../include/cpt.h"
#include "oracle_common.hpp"

namespace orc {

struct Prim {
  int type = CPT_OBJ_NONE;
  d3 center{0, 0, 0};
  AABB box;
  d3 p0{}, p1{}, p2{};       // triangle
  d3 c{}; double r = 0;      // sphere / light centre
  double w = 0, h = 0;       // light extents
  d3 n{}, eu{}, ev{}, le{};  // light frame and radiance
  bool visible = true;
  int mat = 0;  // material index; lights use -1-lightIndex
};

// Heterogeneous medium: density grid in [0,1] (x fastest) over a box,
// sigma_t = sigma * density of the nearest voxel.  Box, sigma and albedo are
// held at the FP32 values the GPU path stores.
struct Medium {
  bool on = false;
  std::vector<float> density;
  int n[3] = {0, 0, 0};
  double bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
  double sigma = 0;
  d3 albedo{0, 0, 0};
};

struct Scene {
  Medium med;
  std::vector<cpt_material> materials;
  std::vector<Prim> prims;
  int nLights = 0, nSpheres = 0;
  // LBVH (Karras numbering; leaf = ~sorted index)
  std::vector<uint64_t> keys;
  std::vector<uint32_t> order;
  std::vector<int> childA, childB;
  std::vector<AABB> nodeBox;
  AABB rootBox;
  int maxDepth = 0;
  bool built = false;

  Scene(const cpt_material *mats, int nmat, const cpt_light *lights, int nl,
        const cpt_sphere *spheres, int ns, const double *tri, const int32_t *tri_mat,
        int64_t nt);
  void build(const double bmin[3], const double bmax[3]);
  int traceRef(const d3 &o, const d3 &d, double tmin, double tmax, double &tBest,
               int64_t *nodesVisited) const;
  // Brute-force closest hit; ok is cleared unless the answer is robust to
  // relative perturbations eb (decisions) and et (hit ordering), an origin
  // uncertainty eo (absolute) and a direction uncertainty ed (relative);
  // tErrOut (optional): the absolute uncertainty of the winner's distance.
  int bruteClosest(const d3 &o, const d3 &d, double tmin, double tmax, double eb, double et, double &tOut,
                   bool &ok, double eo = 0.0, double ed = 0.0, double *tErrOut = nullptr) const;
};

bool hitTriangle(const Prim &p, const d3 &o, const d3 &d, double tmin, double tmax, double &tOut,
                 double *uOut, double *vOut);
bool hitSphere(const Prim &p, const d3 &o, const d3 &d, double tmin, double tmax, double &tOut);
bool hitLight(const Prim &p, const d3 &o, const d3 &d, double tmin, double tmax, double &tOut);
bool hitPrim(const Prim &p, const d3 &o, const d3 &d, double tmin, double tmax, double &t);

}  // namespace orc

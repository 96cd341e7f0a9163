// scene_ini.cpp - the reference's scene ingestion on the host:
//   IniParser::Parse        CudaPathTracing/IniParser.cpp:3-96 (line format,
//                           sections, keys, parseVec3/parseColor/material)
//   IniParser.h:14-73       defaults of every section
//   Mesh::loadFromFile      CudaPathTracing/ObjLoader.cuh:39-163 (rotate about
//                           Y, scale, bounds before translation, translate)
//   Scene::init / add*      CudaPathTracing/Scene.cuh:72-145,175-270 (lights,
//                           spheres, floors, meshes in that order)
// The result is uploaded in one call (cpt_scene_upload) instead of one
// <<<1,1>>> construction kernel per object (Device.cuh:29-79).
// OBJ files: v/vt/vn/f with 1-based or negative indices; polygons are split
// as a fan from their first vertex (tinyobj's ear clipping gives the same
// triangles for convex faces); vertex coordinates are parsed as float, as
// tinyobj's default real_t.  Textures and sky-box images are not decoded.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/cpt.h"

namespace {

struct V3 {
  double x = 0, y = 0, z = 0;
};

struct CameraInfo {  // IniParser.h:14-24
  V3 background{0, 0, 0};
  double width = 0.0;
  double aspectRatio = 16.0 / 9.0;
  V3 lookFrom{0, 0, 1};
  V3 lookAt{0, 0, 0};
  double vFov = 90.0;
  std::string skyBox;
};
struct LightInfo {  // IniParser.h:26-34
  V3 center{0, 0, 0};
  double width = 0, height = 0;
  V3 normal{0, 0, 1};
  V3 color{1, 1, 1};
  bool visible = true;
};
struct SphereInfo {  // IniParser.h:36-44
  V3 center;
  double radius = 0;
  V3 color{1, 1, 1};
  double alphaX = 0.5, alphaY = 0.5;
  int type = CPT_M_OPAQUE;
};
struct FloorInfo {  // IniParser.h:46-53
  V3 lt, rt, lb, rb;
  V3 color{1, 1, 1};
  double alphaX = 0.5, alphaY = 0.5;
  int type = CPT_M_OPAQUE;
};
struct MeshInfo {  // IniParser.h:55-66
  std::string path, texture;
  V3 center;
  V3 color{1, 1, 1};
  double alphaX = 0.5, alphaY = 0.5;
  int type = CPT_M_OPAQUE;
  double rotation = 0.0, scale = 1.0;
};

class IniParser {
 public:
  bool Parse(const std::string &filename) {
    std::ifstream file(filename);
    if (!file) {
      std::cerr << "Unable to open file " << filename << std::endl;
      return false;
    }
    std::string line, section;
    while (std::getline(file, line)) {
      if (line.empty() || line[0] == ';') continue;
      if (line[0] == '[') {
        section = line.substr(1, line.find(']') - 1);
        if (section == "Light") lights.push_back({});
        else if (section == "Sphere") spheres.push_back({});
        else if (section == "Floor") floors.push_back({});
        else if (section == "Mesh") meshes.push_back({});
        continue;
      }
      size_t eq = line.find('=');
      if (eq == std::string::npos) continue;
      std::string key = line.substr(0, eq), value = line.substr(eq + 1);
      try {
        if (section == "Camera") {
          if (key == "background") camera.background = parseColor(value);
          else if (key == "skyBox") camera.skyBox = value;
          else if (key == "width") camera.width = std::stoi(value);
          else if (key == "aspectRatio") camera.aspectRatio = std::stod(value);
          else if (key == "lookFrom") camera.lookFrom = parseVec3(value);
          else if (key == "lookAt") camera.lookAt = parseVec3(value);
          else if (key == "vFov") camera.vFov = std::stod(value);
        } else if (section == "Light" && !lights.empty()) {
          LightInfo &l = lights.back();
          if (key == "center") l.center = parseVec3(value);
          else if (key == "width") l.width = std::stod(value);
          else if (key == "height") l.height = std::stod(value);
          else if (key == "normal") l.normal = parseVec3(value);
          else if (key == "color") l.color = parseColor(value);
          else if (key == "visible") l.visible = std::stoi(value) != 0;
        } else if (section == "Sphere" && !spheres.empty()) {
          SphereInfo &s = spheres.back();
          if (key == "center") s.center = parseVec3(value);
          else if (key == "radius") s.radius = std::stod(value);
          else if (key == "color") s.color = parseColor(value);
          else if (key == "alphaX") s.alphaX = std::stod(value);
          else if (key == "alphaY") s.alphaY = std::stod(value);
          else if (key == "type") s.type = parseMaterialType(value);
        } else if (section == "Floor" && !floors.empty()) {
          FloorInfo &f = floors.back();
          if (key == "lt") f.lt = parseVec3(value);
          else if (key == "rt") f.rt = parseVec3(value);
          else if (key == "lb") f.lb = parseVec3(value);
          else if (key == "rb") f.rb = parseVec3(value);
          else if (key == "color") f.color = parseColor(value);
          else if (key == "alphaX") f.alphaX = std::stod(value);
          else if (key == "alphaY") f.alphaY = std::stod(value);
          else if (key == "type") f.type = parseMaterialType(value);
        } else if (section == "Mesh" && !meshes.empty()) {
          MeshInfo &m = meshes.back();
          if (key == "path") m.path = value;
          else if (key == "texture") m.texture = value;
          else if (key == "center") m.center = parseVec3(value);
          else if (key == "rotation") m.rotation = std::stod(value);
          else if (key == "color") m.color = parseColor(value);
          else if (key == "type") m.type = parseMaterialType(value);
          else if (key == "scale") m.scale = std::stod(value);
          else if (key == "alphaX") m.alphaX = std::stod(value);
          else if (key == "alphaY") m.alphaY = std::stod(value);
        } else if (section == "Cloth") {
          hasCloth = true;  // cloth simulation is not part of this library
        }
      } catch (const std::exception &) {
        // std::stoi/stod on a malformed number: the reference would throw;
        // the library reports the file as unreadable instead.
        std::cerr << "Malformed value for " << key << " in " << filename << std::endl;
        return false;
      }
    }
    return true;
  }

  CameraInfo camera;
  std::vector<LightInfo> lights;
  std::vector<SphereInfo> spheres;
  std::vector<FloorInfo> floors;
  std::vector<MeshInfo> meshes;
  bool hasCloth = false;

 private:
  static V3 parseVec3(const std::string &str) {  // IniParser.cpp:98-105
    V3 v;
    std::istringstream ss(str);
    char d;
    ss >> d >> v.x >> d >> v.y >> d >> v.z >> d;
    return v;
  }
  static V3 parseColor(const std::string &str) {  // IniParser.cpp:107-118
    std::istringstream ss(str);
    int x = 0, y = 0, z = 0;
    char d;
    ss >> d >> x >> d >> y >> d >> z >> d;
    return V3{static_cast<double>(x) / 255.0, static_cast<double>(y) / 255.0, static_cast<double>(z) / 255.0};
  }
  static int parseMaterialType(const std::string &s) {  // IniParser.cpp:120-127
    if (s == "Light") return CPT_M_LIGHT;
    if (s == "Opaque") return CPT_M_OPAQUE;
    if (s == "SpecularDielectric") return CPT_M_SPECULAR_DIELECTRIC;
    if (s == "RoughDielectric") return CPT_M_ROUGH_DIELECTRIC;
    if (s == "Volume") return CPT_M_VOLUME;
    if (s == "Lambert") return CPT_M_LAMBERT;  // new material strings
    if (s == "GGX") return CPT_M_GGX;
    return CPT_M_OPAQUE;
  }
};

// Mesh::loadMesh (ObjLoader.cuh:57-163) with an own OBJ reader.
bool loadObj(const std::string &path, double scale, double rotationDeg, V3 &mn, V3 &mx,
             std::vector<double> &tri) {
  std::ifstream f(path);
  if (!f) {
    std::cerr << "Cannot open OBJ file " << path << std::endl;
    return false;
  }
  const double PI = 3.14159265358979323846;
  const double theta = rotationDeg * PI / 180.0;
  const double c = std::cos(theta), s = std::sin(theta);
  std::vector<float> v;
  std::string line;
  while (std::getline(f, line)) {
    if (line.size() < 2) continue;
    if (line[0] == 'v' && (line[1] == ' ' || line[1] == '\t')) {
      float x = 0, y = 0, z = 0;
      if (std::sscanf(line.c_str() + 2, "%f %f %f", &x, &y, &z) == 3) {
        v.push_back(x);
        v.push_back(y);
        v.push_back(z);
      }
    } else if (line[0] == 'f' && (line[1] == ' ' || line[1] == '\t')) {
      std::istringstream ss(line.substr(2));
      std::string tok;
      std::vector<long> idx;
      const long nv = (long)(v.size() / 3);
      while (ss >> tok) {
        long k = std::strtol(tok.c_str(), nullptr, 10);
        if (k < 0) k = nv + k;
        else k = k - 1;
        idx.push_back(k);
      }
      for (size_t t = 1; t + 1 < idx.size(); t++) {
        long corner[3] = {idx[0], idx[t], idx[t + 1]};
        bool ok = true;
        for (long q : corner) ok = ok && q >= 0 && q < nv;
        if (!ok) continue;
        for (long q : corner) {
          double x = v[3 * q], y = v[3 * q + 1], z = v[3 * q + 2];
          // rotation about +Y, then scale (ObjLoader.cuh:72-105)
          double rx = x * c + y * 0.0 + z * s;
          double ry = x * 0.0 + y * 1.0 + z * 0.0;
          double rz = x * -s + y * 0.0 + z * c;
          rx *= scale;
          ry *= scale;
          rz *= scale;
          mn.x = rx < mn.x ? rx : mn.x;
          mn.y = ry < mn.y ? ry : mn.y;
          mn.z = rz < mn.z ? rz : mn.z;
          mx.x = rx > mx.x ? rx : mx.x;
          mx.y = ry > mx.y ? ry : mx.y;
          mx.z = rz > mx.z ? rz : mx.z;
          tri.push_back(rx);
          tri.push_back(ry);
          tri.push_back(rz);
        }
      }
    }
  }
  return true;
}

struct Assembled {
  std::vector<cpt_material> mats;
  std::vector<cpt_light> lights;
  std::vector<cpt_sphere> spheres;
  std::vector<double> tri;
  std::vector<int32_t> triMat;
  V3 mn{999999.0, 999999.0, 999999.0}, mx{-999999.0, -999999.0, -999999.0};
  cpt_camera_desc cam{};
};

int assemble(const char *path, Assembled &A) {
  if (!path) return CPT_ERR_ARG;
  IniParser P;
  if (!P.Parse(path)) return CPT_ERR_IO;
  auto mat = [&](V3 c, double ax, double ay, int type) {
    cpt_material m{};
    m.color[0] = c.x;
    m.color[1] = c.y;
    m.color[2] = c.z;
    m.alpha_x = ax;
    m.alpha_y = ay;
    m.eta = 1.51;  // Material.cuh:12
    m.type = type;
    A.mats.push_back(m);
    return (int32_t)A.mats.size() - 1;
  };
  for (const LightInfo &l : P.lights) {  // Scene.cuh:76-77
    cpt_light L{};
    L.center[0] = l.center.x; L.center[1] = l.center.y; L.center[2] = l.center.z;
    L.width = l.width;
    L.height = l.height;
    L.normal[0] = l.normal.x; L.normal[1] = l.normal.y; L.normal[2] = l.normal.z;
    L.color[0] = l.color.x; L.color[1] = l.color.y; L.color[2] = l.color.z;
    L.visible = l.visible ? 1 : 0;
    A.lights.push_back(L);
  }
  for (const SphereInfo &s : P.spheres) {  // Scene.cuh:80-81
    cpt_sphere S{};
    S.center[0] = s.center.x; S.center[1] = s.center.y; S.center[2] = s.center.z;
    S.radius = s.radius;
    S.material = mat(s.color, s.alphaX, s.alphaY, s.type);
    A.spheres.push_back(S);
  }
  auto put = [&](V3 p) {
    A.tri.push_back(p.x);
    A.tri.push_back(p.y);
    A.tri.push_back(p.z);
  };
  for (const FloorInfo &f : P.floors) {  // Scene.cuh:84-85, Device.cuh:40-44
    int32_t m = mat(f.color, f.alphaX, f.alphaY, f.type);
    put(f.lt); put(f.lb); put(f.rt);
    put(f.rt); put(f.lb); put(f.rb);
    A.triMat.push_back(m);
    A.triMat.push_back(m);
  }
  for (const MeshInfo &mi : P.meshes) {  // Scene.cuh:88-89, 232-270
    std::vector<double> t;
    loadObj(mi.path, mi.scale, mi.rotation, A.mn, A.mx, t);
    if (t.empty()) continue;
    int32_t m = mat(mi.color, mi.alphaX, mi.alphaY, mi.type);
    for (size_t k = 0; k < t.size(); k += 3) {  // Mesh::transform (ObjLoader.cuh:46-54)
      A.tri.push_back(t[k] + mi.center.x);
      A.tri.push_back(t[k + 1] + mi.center.y);
      A.tri.push_back(t[k + 2] + mi.center.z);
    }
    A.triMat.insert(A.triMat.end(), t.size() / 9, m);
  }
  cpt_camera_desc &c = A.cam;
  c.background[0] = P.camera.background.x;
  c.background[1] = P.camera.background.y;
  c.background[2] = P.camera.background.z;
  c.look_from[0] = P.camera.lookFrom.x; c.look_from[1] = P.camera.lookFrom.y; c.look_from[2] = P.camera.lookFrom.z;
  c.look_at[0] = P.camera.lookAt.x; c.look_at[1] = P.camera.lookAt.y; c.look_at[2] = P.camera.lookAt.z;
  c.vfov_deg = P.camera.vFov;
  c.aspect = P.camera.aspectRatio;
  c.width = (int32_t)P.camera.width;
  return CPT_OK;
}

}  // namespace

extern "C" {

int cpt_ini_counts(const char *path, int64_t counts[4]) {
  if (!counts) return CPT_ERR_ARG;
  Assembled A;
  int rc = assemble(path, A);
  if (rc) return rc;
  counts[0] = (int64_t)A.mats.size();
  counts[1] = (int64_t)A.lights.size();
  counts[2] = (int64_t)A.spheres.size();
  counts[3] = (int64_t)(A.tri.size() / 9);
  return CPT_OK;
}

int cpt_ini_read(const char *path, cpt_material *mats, cpt_light *lights, cpt_sphere *spheres,
                 double *tri, int32_t *triMat, double bmin[3], double bmax[3], cpt_camera_desc *cam) {
  Assembled A;
  int rc = assemble(path, A);
  if (rc) return rc;
  if (mats && !A.mats.empty()) std::memcpy(mats, A.mats.data(), A.mats.size() * sizeof(cpt_material));
  if (lights && !A.lights.empty()) std::memcpy(lights, A.lights.data(), A.lights.size() * sizeof(cpt_light));
  if (spheres && !A.spheres.empty()) std::memcpy(spheres, A.spheres.data(), A.spheres.size() * sizeof(cpt_sphere));
  if (tri && !A.tri.empty()) std::memcpy(tri, A.tri.data(), A.tri.size() * sizeof(double));
  if (triMat && !A.triMat.empty()) std::memcpy(triMat, A.triMat.data(), A.triMat.size() * 4);
  if (bmin) { bmin[0] = A.mn.x; bmin[1] = A.mn.y; bmin[2] = A.mn.z; }
  if (bmax) { bmax[0] = A.mx.x; bmax[1] = A.mx.y; bmax[2] = A.mx.z; }
  if (cam) *cam = A.cam;
  return CPT_OK;
}

int cpt_scene_load_ini(cpt_ctx *ctx, const char *path, double bmin[3], double bmax[3], cpt_camera_desc *cam) {
  if (!ctx) return CPT_ERR_ARG;
  Assembled A;
  int rc = assemble(path, A);
  if (rc) return rc;
  rc = cpt_scene_upload(ctx, A.mats.data(), (int32_t)A.mats.size(), A.lights.data(), (int32_t)A.lights.size(),
                        A.spheres.data(), (int32_t)A.spheres.size(), A.tri.data(), A.triMat.data(),
                        (int64_t)(A.tri.size() / 9));
  if (rc) return rc;
  if (bmin) { bmin[0] = A.mn.x; bmin[1] = A.mn.y; bmin[2] = A.mn.z; }
  if (bmax) { bmax[0] = A.mx.x; bmax[1] = A.mx.y; bmax[2] = A.mx.z; }
  if (cam) *cam = A.cam;
  return CPT_OK;
}

}  // extern "C"

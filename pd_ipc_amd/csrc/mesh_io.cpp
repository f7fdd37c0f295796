// Unstructured cloth input: Wavefront OBJ triangle meshes (core/mesh/Mesh.cpp:56-125
// Mesh::loadObj) and the two lumped-mass rules (:127-153 computeMassesFromSurfaces,
// :155-181 computeMassesFromEdges).
//
// The OBJ reader keeps the reference's reading rules, since they decide which
// files load and what they load as: a line is split at whitespace; "v x y z"
// appends a vertex (fields that do not parse read as 0, C++ stream extraction);
// "f a b c ..." takes the text before the first '/' of its first three fields as
// 1-based vertex indices (std::stoi: leading sign accepted, trailing text
// ignored, a field without digits rejects the whole file); faces with fewer than
// three fields are skipped, polygons keep their first three corners; every other
// tag is ignored.  Unique undirected edges are taken in face order, corner k to
// corner k+1 mod 3, and a face's degenerate corner pairs are skipped.  Unlike the
// reference, an index outside [1, n] is rejected instead of read out of bounds.
#include <cmath>
#include <cstdint>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/pdx.h"

namespace {

struct ObjMesh {
  std::vector<double> pos;      // 3n
  std::vector<int32_t> tri;     // 3T, 0-based
  std::vector<int32_t> ea, eb;  // unique edges in first-seen order
};

int readObj(const char* path, ObjMesh& m) {
  if (!path) return PDX_EINVAL;
  std::ifstream in(path);
  if (!in) return PDX_EINVAL;
  std::vector<long long> corners;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream fields(line);
    std::string tag;
    fields >> tag;
    if (tag == "v") {
      double p[3] = {0.0, 0.0, 0.0};
      fields >> p[0] >> p[1] >> p[2];
      m.pos.insert(m.pos.end(), p, p + 3);
    } else if (tag == "f") {
      long long c[3];
      int k = 0;
      std::string field;
      while (k < 3 && fields >> field) {
        try {
          c[k] = static_cast<long long>(std::stoi(field.substr(0, field.find('/')))) - 1;
        } catch (...) {
          return PDX_EINVAL;
        }
        ++k;
      }
      if (k == 3) corners.insert(corners.end(), c, c + 3);
    }
  }
  const long long n = static_cast<long long>(m.pos.size() / 3);
  if (n == 0 || corners.empty()) return PDX_EINVAL;
  m.tri.resize(corners.size());
  for (size_t i = 0; i < corners.size(); ++i) {
    if (corners[i] < 0 || corners[i] >= n) return PDX_EINVAL;
    m.tri[i] = static_cast<int32_t>(corners[i]);
  }
  std::unordered_set<uint64_t> seen;
  for (size_t t = 0; t < m.tri.size(); t += 3)
    for (int k = 0; k < 3; ++k) {
      const int32_t a = m.tri[t + k], b = m.tri[t + (k + 1) % 3];
      if (a == b) continue;
      const uint64_t lo = static_cast<uint32_t>(a < b ? a : b), hi = static_cast<uint32_t>(a < b ? b : a);
      if (!seen.insert((lo << 32) | hi).second) continue;
      m.ea.push_back(a);
      m.eb.push_back(b);
    }
  return PDX_OK;
}

// Masses that are not above 1e-12 take the smallest positive mass (1 when there is
// none): the global matrix needs every diagonal mass term positive.
void fillNonPositive(int n, double* masses) {
  double smallest = 1e300;
  for (int v = 0; v < n; ++v)
    if (masses[v] > 1e-12) smallest = std::fmin(smallest, masses[v]);
  if (smallest > 1e299) smallest = 1.0;
  for (int v = 0; v < n; ++v)
    if (masses[v] < 1e-12) masses[v] = smallest;
}

}  // namespace

extern "C" {

int pdx_obj_counts(const char* path, int* n, int* n_edges, int* n_tris) {
  ObjMesh m;
  const int rc = readObj(path, m);
  if (rc != PDX_OK) return rc;
  if (n) *n = static_cast<int>(m.pos.size() / 3);
  if (n_edges) *n_edges = static_cast<int>(m.ea.size());
  if (n_tris) *n_tris = static_cast<int>(m.tri.size() / 3);
  return PDX_OK;
}

int pdx_obj_fill(const char* path, double kappa, double* positions, int32_t* tris, int32_t* ea,
                 int32_t* eb, double* rest, double* kap) {
  if (!positions || !tris || !ea || !eb || !rest || !kap) return PDX_EINVAL;
  ObjMesh m;
  const int rc = readObj(path, m);
  if (rc != PDX_OK) return rc;
  for (size_t i = 0; i < m.pos.size(); ++i) positions[i] = m.pos[i];
  for (size_t i = 0; i < m.tri.size(); ++i) tris[i] = m.tri[i];
  const double* p = m.pos.data();
  for (size_t e = 0; e < m.ea.size(); ++e) {
    const int a = m.ea[e], b = m.eb[e];
    const double dx = p[3 * b] - p[3 * a], dy = p[3 * b + 1] - p[3 * a + 1], dz = p[3 * b + 2] - p[3 * a + 2];
    ea[e] = a;
    eb[e] = b;
    rest[e] = std::sqrt(dx * dx + dy * dy + dz * dz);
    kap[e] = kappa;
  }
  return PDX_OK;
}

int pdx_masses_from_surfaces(int n, const double* rest_positions, int n_tris, const int32_t* tris,
                             double density, double* masses) {
  if (n <= 0 || n_tris < 0 || !rest_positions || !masses || (n_tris > 0 && !tris)) return PDX_EINVAL;
  for (int i = 0; i < 3 * n_tris; ++i)
    if (tris[i] < 0 || tris[i] >= n) return PDX_EINVAL;
  for (int v = 0; v < n; ++v) masses[v] = 0.0;
  const double* x = rest_positions;
  for (int t = 0; t < n_tris; ++t) {
    const int i0 = tris[3 * t], i1 = tris[3 * t + 1], i2 = tris[3 * t + 2];
    const double ux = x[3 * i1] - x[3 * i0], uy = x[3 * i1 + 1] - x[3 * i0 + 1], uz = x[3 * i1 + 2] - x[3 * i0 + 2];
    const double wx = x[3 * i2] - x[3 * i0], wy = x[3 * i2 + 1] - x[3 * i0 + 1], wz = x[3 * i2 + 2] - x[3 * i0 + 2];
    // |u x w| / 2, the cross product component by component
    const double cx = uy * wz - uz * wy, cy = uz * wx - ux * wz, cz = ux * wy - uy * wx;
    const double area = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
    const double share = density * area / 3.0;
    masses[i0] += share;
    masses[i1] += share;
    masses[i2] += share;
  }
  fillNonPositive(n, masses);
  return PDX_OK;
}

int pdx_masses_from_edges(int n, int n_edges, const int32_t* ea, const int32_t* eb, const double* rest,
                          double density, double* masses) {
  if (n <= 0 || n_edges < 0 || !masses || (n_edges > 0 && (!ea || !eb || !rest))) return PDX_EINVAL;
  for (int e = 0; e < n_edges; ++e)
    if (ea[e] < 0 || ea[e] >= n || eb[e] < 0 || eb[e] >= n) return PDX_EINVAL;
  for (int v = 0; v < n; ++v) masses[v] = 0.0;
  for (int e = 0; e < n_edges; ++e) {
    // rest length squared as the edge's area share, half to each end
    const double share = density * rest[e] * rest[e] * 0.5;
    masses[ea[e]] += share;
    masses[eb[e]] += share;
  }
  fillNonPositive(n, masses);
  return PDX_OK;
}

}  // extern "C"

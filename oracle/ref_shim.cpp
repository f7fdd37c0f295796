// C entry points over the reference's own mesh + colouring sources
// (core/mesh/Mesh.cpp, core/mesh/ConstraintColoring.cpp, core/math/Parallel.cpp),
// compiled from where they lie in the reference checkout into oracle/_ref/.
// Test infrastructure only: it lets the tests compare this repository's grid
// builder, edge colouring and incidence CSRs with the reference's bit for bit.
// The scene sequence mirrors makeScene (core/sim/Scene.cpp:39-169) without the
// parts that need Eigen: grid, kappa on every edge, edge-lumped masses (density),
// optional shear diagonals, optional bend stencils, two top corners pinned.
#include <cstdint>
#include <cstring>

#include "core/mesh/Mesh.h"

using pd::Mesh;

extern "C" {

void* ref_grid_scene(int nx, int ny, double spacing, double stiffness, double density,
                     double shear, double bend) {
  Mesh* m = new Mesh(Mesh::makeGrid(nx, ny, spacing));
  for (auto& e : m->edges) e.stiffness = stiffness;
  m->computeMassesFromEdges(density > 0.0 ? density : 1.0);
  if (shear > 0.0) m->addShearDiagonals(nx, ny, shear);
  if (bend > 0.0) m->addBendingStencils(nx, ny, bend);
  m->pinned[static_cast<std::size_t>((ny - 1) * nx)] = 1;
  m->pinned[static_cast<std::size_t>((ny - 1) * nx + nx - 1)] = 1;
  return m;
}

void ref_free(void* h) { delete static_cast<Mesh*>(h); }

void ref_counts(void* h, int64_t* out) {
  Mesh* m = static_cast<Mesh*>(h);
  const auto& col = m->constraintColoring();
  out[0] = m->vertexCount();
  out[1] = m->edgeCount();
  out[2] = m->bendCount();
  out[3] = col.colorCount;
  out[4] = static_cast<int64_t>(m->blockPattern().size());
}

void ref_vertices(void* h, double* pos, double* masses, uint8_t* pinned) {
  Mesh* m = static_cast<Mesh*>(h);
  for (int v = 0; v < m->vertexCount(); ++v) {
    pos[3 * v] = m->positions[v].x;
    pos[3 * v + 1] = m->positions[v].y;
    pos[3 * v + 2] = m->positions[v].z;
    masses[v] = m->masses[v];
    pinned[v] = m->pinned[v];
  }
}

void ref_edges(void* h, int32_t* a, int32_t* b, double* rest, double* kappa) {
  Mesh* m = static_cast<Mesh*>(h);
  for (int e = 0; e < m->edgeCount(); ++e) {
    a[e] = m->edges[e].a;
    b[e] = m->edges[e].b;
    rest[e] = m->edges[e].restLength;
    kappa[e] = m->edges[e].stiffness;
  }
}

void ref_bends(void* h, int32_t* a, int32_t* b, int32_t* c, double* k) {
  Mesh* m = static_cast<Mesh*>(h);
  for (int s = 0; s < m->bendCount(); ++s) {
    a[s] = m->bends[s].a;
    b[s] = m->bends[s].b;
    c[s] = m->bends[s].c;
    k[s] = m->bends[s].stiffness;
  }
}

void ref_coloring(void* h, uint32_t* color_of, uint32_t* order, uint32_t* start,
                  uint32_t* vstart, uint32_t* vedges) {
  Mesh* m = static_cast<Mesh*>(h);
  const auto& c = m->constraintColoring();
  std::memcpy(color_of, c.colorOf.data(), c.colorOf.size() * 4);
  std::memcpy(order, c.order.data(), c.order.size() * 4);
  std::memcpy(start, c.start.data(), c.start.size() * 4);
  std::memcpy(vstart, c.vertexStart.data(), c.vertexStart.size() * 4);
  std::memcpy(vedges, c.vertexEdges.data(), c.vertexEdges.size() * 4);
}

// An OBJ cloth as makeScene builds it for a mesh path (Scene.cpp:44-68):
// Mesh::loadObj with the stiffness, then masses from the surfaces with the
// density.  No pins.  Returns NULL when loadObj refuses the file.
void* ref_load_obj(const char* path, double stiffness, double density) {
  Mesh* m = new Mesh();
  if (!Mesh::loadObj(path, *m, stiffness)) {
    delete m;
    return nullptr;
  }
  for (auto& e : m->edges) e.stiffness = stiffness;
  m->computeMassesFromSurfaces(density > 0.0 ? density : 1.0);
  return m;
}

int64_t ref_triangle_count(void* h) {
  return static_cast<int64_t>(static_cast<Mesh*>(h)->triangleIndices.size());
}

void ref_triangles(void* h, int32_t* out) {
  Mesh* m = static_cast<Mesh*>(h);
  for (size_t t = 0; t < m->triangleIndices.size(); ++t)
    for (int k = 0; k < 3; ++k) out[3 * t + k] = m->triangleIndices[t][static_cast<size_t>(k)];
}

void ref_bend_adjacency(void* h, uint32_t* vstart, uint32_t* stencils, double* weight) {
  Mesh* m = static_cast<Mesh*>(h);
  const auto& b = m->bendAdjacency();
  std::memcpy(vstart, b.vertexStart.data(), b.vertexStart.size() * 4);
  std::memcpy(stencils, b.vertexStencils.data(), b.vertexStencils.size() * 4);
  std::memcpy(weight, b.vertexStencilWeight.data(), b.vertexStencilWeight.size() * 8);
}

}  // extern "C"

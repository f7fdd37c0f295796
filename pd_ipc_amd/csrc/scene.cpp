// Regular-grid cloth builder (core/mesh/Mesh.cpp:12-54 makeGrid, :183-207
// addShearDiagonals, :219-285 addBendingStencils with the standard sampling).
// Vertex v = j*nx + i sits at (i*sx, 0, j*sz); rest lengths are measured on
// those positions exactly as the reference measures them (b - a, then length).
#include <cmath>
#include <cstdint>

#include "../../include/pdx.h"

extern "C" {

int pdx_grid_counts(int nx, int ny, int shear, int bend, int* n, int* E, int* B) {
  if (nx < 2 || ny < 2) return PDX_EINVAL;
  if (n) *n = nx * ny;
  if (E) *E = 2 * nx * ny - nx - ny + (shear ? (nx - 1) * (ny - 1) : 0);
  if (B) *B = bend ? (nx - 2) * ny + (ny - 2) * nx : 0;
  return PDX_OK;
}

int pdx_grid_fill(int nx, int ny, double sx, double sz, double kappa, double shear_kappa,
                  double bend_k, double* pos, int32_t* ea, int32_t* eb, double* rest,
                  double* kap, int32_t* sa, int32_t* sb, int32_t* sc, double* bk) {
  if (nx < 2 || ny < 2 || !pos || !ea || !eb || !rest || !kap) return PDX_EINVAL;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const int v = j * nx + i;
      pos[3 * v] = static_cast<double>(i) * sx;
      pos[3 * v + 1] = 0.0;
      pos[3 * v + 2] = static_cast<double>(j) * sz;
    }
  auto len = [&](int a, int b) {
    const double dx = pos[3 * b] - pos[3 * a], dy = pos[3 * b + 1] - pos[3 * a + 1],
                 dz = pos[3 * b + 2] - pos[3 * a + 2];
    return std::sqrt(dx * dx + dy * dy + dz * dz);
  };
  int e = 0;
  auto add = [&](int a, int b, double k) {
    ea[e] = a;
    eb[e] = b;
    rest[e] = len(a, b);
    kap[e] = k;
    ++e;
  };
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const int v = j * nx + i;
      if (i + 1 < nx) add(v, v + 1, kappa);
      if (j + 1 < ny) add(v, v + nx, kappa);
    }
  if (shear_kappa > 0.0)
    for (int j = 0; j + 1 < ny; ++j)
      for (int i = 0; i + 1 < nx; ++i) add(j * nx + i, (j + 1) * nx + i + 1, shear_kappa);
  if (bend_k > 0.0) {
    if (!sa || !sb || !sc || !bk) return PDX_EINVAL;
    int q = 0;
    for (int j = 0; j < ny; ++j)
      for (int i = 1; i + 1 < nx; ++i) {
        sa[q] = j * nx + i - 1;
        sb[q] = j * nx + i;
        sc[q] = j * nx + i + 1;
        bk[q++] = bend_k;
      }
    for (int i = 0; i < nx; ++i)
      for (int j = 1; j + 1 < ny; ++j) {
        sa[q] = (j - 1) * nx + i;
        sb[q] = j * nx + i;
        sc[q] = (j + 1) * nx + i;
        bk[q++] = bend_k;
      }
  }
  return PDX_OK;
}

}  // extern "C"

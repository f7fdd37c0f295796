// Host-side data structures of the PD substep engine: scene topology, the
// reference's edge colouring and incidence lists, the assembled global matrix,
// its LDL^T factor, and the level schedules the GPU solve kernels walk.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace pdx {

// One cloth instance as handed over the C ABI (all arrays copied).
struct Scene {
  int n = 0, E = 0, B = 0;
  std::vector<double> x, v, mass, pinpos;  // 3n, 3n, n, 3n
  std::vector<uint8_t> pinned;             // n
  std::vector<int> ea, eb, sa, sb, sc;     // E, E, B, B, B
  std::vector<double> rest, kappa, bk;     // E, E, B
};

struct Config {
  double dt = 1.0 / 120.0;
  double g[3] = {0.0, -9.81, 0.0};
  double damping = 0.02;
  int maxIter = 10;
  double relTol = 1e-3, absTol = 1e-12, resTol = 1e-3;
  double chebRho = 0.0;  // Chebyshev blend (0 = off)
  int chebStart = 10;
};

// Topology-only data (core/mesh/ConstraintColoring.h:34-81).
struct Topology {
  int colorCount = 0;
  std::vector<uint32_t> colorOf, cstart, corder;
  std::vector<uint32_t> vstart, vedges;        // incidence, edge-index order
  std::vector<uint32_t> vedgesColor;           // incidence, colour order (one edge per colour)
  std::vector<uint32_t> bstart, bstencil;      // vertex -> bend stencils (stencil order)
  std::vector<int8_t> bweight;                 // 1, -2, 1
};

void buildTopology(const Scene& s, Topology& t);

// Scalar component matrix A (L = A (x) I3), CSC over vertices with rows ascending.
struct Matrix {
  int n = 0;
  std::vector<int> colptr, rows;
  std::vector<double> vals;
};
void assembleMatrix(const Scene& s, double dt, Matrix& A);
void bendingRhs(const Scene& s, std::vector<double>& out);

void amdOrder(int n, const std::vector<int>& colptr, const std::vector<int>& rowidx,
              std::vector<int>& perm);
// Nested-dissection vertex order of A's graph by coordinate bisection of the
// positions x (3n); parts of at most `leaf` vertices stay in sorted order.  The
// order lists the first half of the top-level split, then the second, then the
// separator; when `top` is given it receives the two halves' sizes (0, 0 when
// the graph was a single leaf).  No matrix entry couples the two halves.
void ndOrder(const Matrix& A, const std::vector<double>& x, int leaf, std::vector<int>& order,
             int* top = nullptr);

// Level-scheduled sweep of a unit triangular factor (rank space).
struct Sweep {
  int nlev = 0;
  std::vector<int> levelOf;                    // per rank
  std::vector<int> levPtr;                     // nlev+1 ranges into `order`
  std::vector<int> order;                      // ranks in level order
};

// Factor of the scalar component in the AMD order of the 3n pattern
// (core/solver/EigenDirectSolver.cpp:230-297 + compact factor :157-225).
struct Factor {
  int m = 0;
  std::vector<int> perm3;                      // AMD pivot order of the 3n pattern
  std::vector<int> rankVertex[3];              // rank -> vertex, per component
  std::vector<int> vertexRank[3];              // vertex -> rank, per component
  bool sharedOrder = true;                     // the three component orders coincide
  // strictly lower factor, CSC by rank (rows ascending) and CSR by rank (cols ascending)
  std::vector<int> cp, ci;
  std::vector<double> cx;
  std::vector<int> rp, rc;
  std::vector<double> rx;
  std::vector<double> dinv;
  int64_t nnz = 0;
  Sweep fwd, bwd;
  // nested-dissection order: ranks [0, top[0]) and [top[0], top[0] + top[1]) are the
  // two halves of the top-level split (no entry couples them), the rest its
  // separator; zero in the AMD order
  int top[2] = {0, 0};
};

// Fill-reducing orders of analyzeAndFactor.
enum { kOrderAmd = 0, kOrderNd = 1 };
constexpr int kNdLeaf = 8;  // 512-instance mix, mixed32: leaves of 2/4/8/16/32/64 gave
                            // 29.7/29.8/29.6/30.0/30.8/31.8 ms per launch

// Returns 0 or a negative status (-3 = not positive definite).
int analyzeAndFactor(const Scene& s, const Matrix& A, Factor& f, int ordering = kOrderAmd);
int refactorNumeric(const Matrix& A, Factor& f);

// Dense inverse of the component matrix in rank space, m x m, row r = the solve
// of e_r through the factor (forward, 1/D, backward), which is column r of the
// symmetric inverse.  Rows are spread over `threads` host threads.
void denseInverse(const Factor& f, std::vector<double>& out, int threads);

uint64_t numericStamp(const Scene& s, double dt);
uint64_t structureStamp(const Scene& s);

// Strain, energy and speed figures of a state (x, v: 3n) of the scene's cloth,
// FP64 in the reference's summation order (core/mesh/Mesh.cpp:345-372,
// core/sim/Integrator.cpp:804-846).  Fills pdx_diag's fields in order:
// max strain, mean strain, elastic energy, total energy, max speed, total mass.
void physicsDiagnostics(const Scene& s, const double* x, const double* v, const double g[3], double out[6]);

}  // namespace pdx

// FP64 CPU oracle of the reference PD substep path (test infrastructure only).
//
// Every function below names the reference file:line it restates.  The arithmetic
// order (which operands are added first, where products are rounded) follows the
// reference exactly, because the GPU library is checked against this file bitwise.
// Built with -ffp-contract=off so no multiply-add is fused.
#include "pd_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct V3 {
  double x = 0, y = 0, z = 0;
};
// core/math/Vec3.h:24-52
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3& operator+=(V3& a, V3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
inline V3& operator-=(V3& a, V3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double length(V3 a) { return std::sqrt(dot(a, a)); }
inline double lengthSquared(V3 a) { return dot(a, a); }
inline double maxAbsComponent(V3 a) {
  return std::fmax(std::fabs(a.x), std::fmax(std::fabs(a.y), std::fabs(a.z)));
}

constexpr double kMinLength = 1e-12;  // DistanceTerm.h:31
constexpr uint32_t kUncolored = 0xFFFFFFFFu;

// ConstraintColoring.cpp:23-31
inline uint32_t priorityOf(uint32_t e) {
  uint64_t x = static_cast<uint64_t>(e) * 0x9E3779B97F4A7C15ULL + 0x2545F4914F6CDD1DULL;
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ULL;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBULL;
  x ^= x >> 31;
  return static_cast<uint32_t>(x >> 32);
}

uint64_t bitsOf(double v) {
  uint64_t u;
  std::memcpy(&u, &v, 8);
  return u;
}
// Assembler.cpp:22-25
uint64_t mix(uint64_t h, uint64_t v) { return h ^ (v + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2)); }

}  // namespace

struct orc_ctx {
  // ---- scene (Mesh.h:96-217) ----
  int n = 0, E = 0, B = 0;
  std::vector<V3> positions, velocities, pinPositions;
  std::vector<double> masses;
  std::vector<uint8_t> pinned;
  std::vector<int> ea, eb, sa, sb, sc;
  std::vector<double> rest, kappa, bk;
  orc_config cfg{};
  int threads = 1;

  // ---- topology (ConstraintColoring.h:34-81) ----
  std::vector<uint32_t> colorOf, cstart, corder, vertexStart, vertexEdges;
  int colorCount = 0;
  std::vector<uint32_t> bStart, bStencils;
  std::vector<double> bWeight;

  // ---- solver state ----
  std::vector<int> perm3;  // origOfPerm
  bool haveOrdering = false;
  bool stampValid = false, analyzed = false;
  uint64_t stampStiff = 0, stampMass = 0, stampDt = 0, stampPin = 0, structureStamp = 0;
  int factorizeCount = 0, analyzeCount = 0;
  // scalar matrix (one component), CSC over vertices (rows ascending)
  std::vector<int> Ap, Ai;
  std::vector<double> Ax;
  // compact factor (EigenDirectSolver.cpp:89-97)
  int m = 0;
  std::vector<int> cptr, crow;
  std::vector<double> cval, cdinv;
  std::vector<std::vector<int>> compDof;  // compDof[c][k] = original dof of rank k
  int components = 1;
  bool compactReady = false;
  // full factor (fallback path when compact self-check fails)
  std::vector<int> Lp, Li;
  std::vector<double> Lx, dinvFull;
  std::vector<int> Pidx;  // new index of original dof

  // ---- per-step buffers (Scene.h:171-188) ----
  std::vector<V3> predicted, positionsBeforeStep, positionsBeforePrevStep;
  std::vector<double> b, bBase, xSolution, bendRhs;
  int iterationsUsed = 0, earlyExitCount = 0;
  // when set, stepOnce records every iterate x^{k+1} (xSolution after iteration k)
  // at iterTape + k * 3n: the adjoint's tape (no counterpart in the reference)
  double* iterTape = nullptr;
  bool converged = false;
  double lastResidual = -1.0;
  std::vector<double> forceBuf;

  bool isPinned(int v) const { return pinned[(size_t)v] != 0; }
};

namespace {

// ConstraintColoring.cpp:35-54 (vertex -> edge CSR in edge-index order)
void buildAdjacency(orc_ctx& c) {
  c.vertexStart.assign((size_t)c.n + 1, 0);
  for (int e = 0; e < c.E; ++e) {
    c.vertexStart[(size_t)c.ea[e] + 1] += 1;
    c.vertexStart[(size_t)c.eb[e] + 1] += 1;
  }
  for (int v = 0; v < c.n; ++v) c.vertexStart[(size_t)v + 1] += c.vertexStart[(size_t)v];
  c.vertexEdges.assign((size_t)2 * c.E, 0);
  std::vector<uint32_t> cursor(c.vertexStart.begin(), c.vertexStart.end() - 1);
  for (int e = 0; e < c.E; ++e) {
    c.vertexEdges[cursor[(size_t)c.ea[e]]++] = (uint32_t)e;
    c.vertexEdges[cursor[(size_t)c.eb[e]]++] = (uint32_t)e;
  }
}

// ConstraintColoring.cpp:82-119 (vertex -> stencil CSR with weights)
void buildBendAdjacency(orc_ctx& c) {
  c.bStart.assign((size_t)c.n + 1, 0);
  for (int s = 0; s < c.B; ++s) {
    c.bStart[(size_t)c.sa[s] + 1] += 1;
    c.bStart[(size_t)c.sb[s] + 1] += 1;
    c.bStart[(size_t)c.sc[s] + 1] += 1;
  }
  for (int v = 0; v < c.n; ++v) c.bStart[(size_t)v + 1] += c.bStart[(size_t)v];
  c.bStencils.assign((size_t)3 * c.B, 0);
  c.bWeight.assign((size_t)3 * c.B, 0.0);
  std::vector<uint32_t> cursor(c.bStart.begin(), c.bStart.end() - 1);
  for (int s = 0; s < c.B; ++s) {
    const int vert[3] = {c.sa[s], c.sb[s], c.sc[s]};
    const double w[3] = {1.0, -2.0, 1.0};
    for (int k = 0; k < 3; ++k) {
      const size_t slot = cursor[(size_t)vert[k]]++;
      c.bStencils[slot] = (uint32_t)s;
      c.bWeight[slot] = w[k];
    }
  }
}

// ConstraintColoring.cpp:121-259 (Jones-Plassmann colouring, smallest free colour)
void buildColoring(orc_ctx& c) {
  const int ne = c.E;
  c.colorOf.assign((size_t)ne, kUncolored);
  c.cstart.clear();
  c.corder.clear();
  c.colorCount = 0;
  buildAdjacency(c);
  if (ne <= 0) {
    c.cstart.push_back(0);
    return;
  }
  auto forEachNeighbor = [&](uint32_t e, auto&& fn) {
    const uint32_t ends[2] = {(uint32_t)c.ea[e], (uint32_t)c.eb[e]};
    for (uint32_t v : ends)
      for (uint32_t k = c.vertexStart[v]; k < c.vertexStart[v + 1]; ++k) {
        const uint32_t f = c.vertexEdges[k];
        if (f != e) fn(f);
      }
  };
  std::vector<uint32_t> prio((size_t)ne);
  for (int i = 0; i < ne; ++i) prio[(size_t)i] = priorityOf((uint32_t)i);
  auto better = [&](uint32_t a, uint32_t b) {
    return prio[a] > prio[b] || (prio[a] == prio[b] && a < b);
  };
  std::vector<uint32_t> remaining((size_t)ne), next, candidate((size_t)ne, kUncolored);
  for (int i = 0; i < ne; ++i) remaining[(size_t)i] = (uint32_t)i;
  int maxColor = -1;
  while (!remaining.empty()) {
    for (uint32_t ed : remaining) {
      candidate[ed] = kUncolored;
      bool allColored = true, localMax = true;
      forEachNeighbor(ed, [&](uint32_t f) {
        if (c.colorOf[f] == kUncolored) {
          allColored = false;
          if (better(f, ed)) localMax = false;
        }
      });
      if (!allColored && !localMax) continue;
      uint32_t pick = 0;
      bool clash = true;
      while (clash) {
        clash = false;
        forEachNeighbor(ed, [&](uint32_t f) {
          if (!clash && c.colorOf[f] == pick) {
            clash = true;
            ++pick;
          }
        });
      }
      candidate[ed] = pick;
    }
    next.clear();
    bool any = false;
    for (uint32_t ed : remaining) {
      const uint32_t col = candidate[ed];
      if (col != kUncolored) {
        c.colorOf[ed] = col;
        maxColor = std::max(maxColor, (int)col);
        any = true;
      } else {
        next.push_back(ed);
      }
    }
    if (!any) {
      std::fprintf(stderr, "[oracle] colouring made no progress\n");
      std::abort();
    }
    remaining.swap(next);
  }
  c.colorCount = maxColor + 1;
  c.cstart.assign((size_t)c.colorCount + 1, 0);
  for (uint32_t col : c.colorOf) c.cstart[(size_t)col + 1] += 1;
  for (int k = 0; k < c.colorCount; ++k) c.cstart[(size_t)k + 1] += c.cstart[(size_t)k];
  c.corder.assign((size_t)ne, 0);
  std::vector<uint32_t> cursor(c.cstart.begin(), c.cstart.end() - 1);
  for (int i = 0; i < ne; ++i) c.corder[cursor[c.colorOf[(size_t)i]]++] = (uint32_t)i;
  for (int k = 0; k < c.colorCount; ++k) {
    auto first = c.corder.begin() + c.cstart[(size_t)k];
    auto last = c.corder.begin() + c.cstart[(size_t)k + 1];
    std::sort(first, last, [&](uint32_t x, uint32_t y) {
      const uint32_t mx = (uint32_t)std::min(c.ea[x], c.eb[x]);
      const uint32_t my = (uint32_t)std::min(c.ea[y], c.eb[y]);
      return mx != my ? mx < my : x < y;
    });
  }
}

// Assembler.cpp:29-65 (numeric stamp; damping excluded from the comparison)
void computeStamp(const orc_ctx& c, uint64_t& stiff, uint64_t& mass, uint64_t& dt,
                  uint64_t& pin) {
  uint64_t sh = 0xcbf29ce484222325ULL, mh = 0xcbf29ce484222325ULL, ph = 0xcbf29ce484222325ULL;
  for (int e = 0; e < c.E; ++e) sh = mix(sh, bitsOf(c.kappa[e]));
  for (int s = 0; s < c.B; ++s) sh = mix(sh, bitsOf(c.bk[s]));
  for (int v = 0; v < c.n; ++v) mh = mix(mh, bitsOf(c.masses[v]));
  for (int v = 0; v < c.n; ++v) ph = mix(ph, c.pinned[v] ? 1ULL : 0ULL);
  stiff = sh;
  mass = mh;
  dt = bitsOf(c.cfg.dt);
  pin = ph;
}

// Assembler.cpp:67-92
uint64_t computeStructureStamp(const orc_ctx& c) {
  uint64_t h = 0xcbf29ce484222325ULL;
  h = mix(h, 1);
  h = mix(h, (uint64_t)c.n);
  h = mix(h, (uint64_t)c.E);
  for (int e = 0; e < c.E; ++e) {
    h = mix(h, (uint64_t)(uint32_t)c.ea[e]);
    h = mix(h, (uint64_t)(uint32_t)c.eb[e]);
  }
  h = mix(h, (uint64_t)c.B);
  for (int s = 0; s < c.B; ++s) {
    h = mix(h, (uint64_t)(uint32_t)c.sa[s]);
    h = mix(h, (uint64_t)(uint32_t)c.sb[s]);
    h = mix(h, (uint64_t)(uint32_t)c.sc[s]);
  }
  for (int v = 0; v < c.n; ++v) h = mix(h, c.pinned[v] ? 1ULL : 0ULL);
  return h;
}

// Assembler.cpp:94-137
void assembleBendingRhs(const orc_ctx& c, std::vector<double>& out) {
  out.assign((size_t)3 * c.n, 0.0);
  if (c.B == 0) return;
  for (int s = 0; s < c.B; ++s) {
    if (c.bk[s] == 0.0) continue;
    const int vert[3] = {c.sa[s], c.sb[s], c.sc[s]};
    const double w[3] = {1.0, -2.0, 1.0};
    V3 cc{};
    bool any = false;
    for (int k = 0; k < 3; ++k) {
      if (!c.isPinned(vert[k])) continue;
      any = true;
      cc += c.pinPositions[(size_t)vert[k]] * w[k];
    }
    if (!any) continue;
    for (int k = 0; k < 3; ++k) {
      const int p = vert[k];
      if (c.isPinned(p)) continue;
      const V3 contrib = cc * (-c.bk[s] * w[k]);
      double* slot = out.data() + (size_t)p * 3;
      slot[0] += contrib.x;
      slot[1] += contrib.y;
      slot[2] += contrib.z;
    }
  }
}

struct Trip {
  int r, c;
  double v;
};

// Assembler.cpp:139-278 + DistanceTerm.cpp:319-349, one scalar component; the
// per-entry sums follow triplet insertion order (Eigen setFromTriplets).
void assembleScalar(orc_ctx& c) {
  const int n = c.n;
  const double invDt2 = 1.0 / (c.cfg.dt * c.cfg.dt);
  std::vector<Trip> t;
  t.reserve((size_t)n + 4 * (size_t)c.E + 9 * (size_t)c.B);
  for (int v = 0; v < n; ++v) t.push_back({v, v, c.masses[v] * invDt2});
  for (int e = 0; e < c.E; ++e) {
    const bool ap = c.isPinned(c.ea[e]), bp = c.isPinned(c.eb[e]);
    if (ap && bp) continue;
    if (!ap) t.push_back({c.ea[e], c.ea[e], c.kappa[e]});
    if (!bp) t.push_back({c.eb[e], c.eb[e], c.kappa[e]});
    if (!ap && !bp) {
      t.push_back({c.ea[e], c.eb[e], -c.kappa[e]});
      t.push_back({c.eb[e], c.ea[e], -c.kappa[e]});
    }
  }
  const double w[3] = {1.0, -2.0, 1.0};
  for (int s = 0; s < c.B; ++s) {
    if (c.bk[s] == 0.0) continue;
    const int vert[3] = {c.sa[s], c.sb[s], c.sc[s]};
    if (c.isPinned(vert[0]) && c.isPinned(vert[1]) && c.isPinned(vert[2])) continue;
    for (int p = 0; p < 3; ++p) {
      if (c.isPinned(vert[p])) continue;
      for (int q = 0; q < 3; ++q) {
        if (c.isPinned(vert[q])) continue;
        t.push_back({vert[p], vert[q], c.bk[s] * w[p] * w[q]});
      }
    }
  }
  for (int v = 0; v < n; ++v) {
    if (!c.isPinned(v)) continue;
    double existing = c.masses[v] * invDt2;
    for (int e = 0; e < c.E; ++e) {
      if (c.isPinned(c.ea[e]) || c.isPinned(c.eb[e])) continue;
      if (c.ea[e] == v || c.eb[e] == v) existing += c.kappa[e];
    }
    // bendDiagonalAt(v) is 0 for a pinned vertex (Assembler.cpp:227-245).
    t.push_back({v, v, 1.0 - existing});
  }
  // collapse duplicates in insertion order, then CSC by column with rows ascending
  std::stable_sort(t.begin(), t.end(), [](const Trip& a, const Trip& b) {
    return a.c != b.c ? a.c < b.c : a.r < b.r;
  });
  c.Ap.assign((size_t)n + 1, 0);
  c.Ai.clear();
  c.Ax.clear();
  for (size_t k = 0; k < t.size();) {
    size_t j = k;
    double s = t[k].v;
    while (j + 1 < t.size() && t[j + 1].r == t[k].r && t[j + 1].c == t[k].c) {
      ++j;
      s = s + t[j].v;
    }
    c.Ai.push_back(t[k].r);
    c.Ax.push_back(s);
    c.Ap[(size_t)t[k].c + 1] += 1;
    k = j + 1;
  }
  for (int v = 0; v < n; ++v) c.Ap[(size_t)v + 1] += c.Ap[(size_t)v];
}

// EigenDirectSolver.cpp:230-297 with Eigen SimplicialLDLT (NaturalOrdering) on the
// permuted 3n matrix: symbolic etree (LDL ldl_symbolic) and up-looking numeric LDL^T
// (LDL ldl_numeric), then dinv = 1/D, the x/y/z component split and the compact
// shared factor with its bitwise self-check.
int factorize(orc_ctx& c) {
  const int n3 = 3 * c.n;
  // new index of original dof i: P(i)
  c.Pidx.assign((size_t)n3, -1);
  for (int k = 0; k < n3; ++k) c.Pidx[(size_t)c.perm3[(size_t)k]] = k;
  // permuted upper storage: column J holds rows I <= J (ascending)
  std::vector<int> up((size_t)n3 + 1, 0), ui;
  std::vector<double> ux;
  {
    std::vector<std::pair<int, double>> tmp;
    std::vector<std::vector<std::pair<int, double>>> cols((size_t)n3);
    for (int vc = 0; vc < c.n; ++vc)
      for (int p = c.Ap[(size_t)vc]; p < c.Ap[(size_t)vc + 1]; ++p) {
        const int vr = c.Ai[(size_t)p];
        if (vr < vc) continue;  // lower triangle of the vertex matrix (selfadjointView<Lower>)
        for (int d = 0; d < 3; ++d) {
          const int i = 3 * vr + d, j = 3 * vc + d;  // i >= j
          const int pi = c.Pidx[(size_t)i], pj = c.Pidx[(size_t)j];
          const int R = std::min(pi, pj), C = std::max(pi, pj);
          cols[(size_t)C].push_back({R, c.Ax[(size_t)p]});
        }
      }
    for (int J = 0; J < n3; ++J) {
      auto& v = cols[(size_t)J];
      std::sort(v.begin(), v.end(),
                [](const std::pair<int, double>& a, const std::pair<int, double>& b) {
                  return a.first < b.first;
                });
      up[(size_t)J + 1] = up[(size_t)J] + (int)v.size();
      for (auto& e : v) {
        ui.push_back(e.first);
        ux.push_back(e.second);
      }
    }
  }
  // ldl_symbolic
  std::vector<int> parent((size_t)n3), nzc((size_t)n3), tags((size_t)n3);
  for (int k = 0; k < n3; ++k) {
    parent[(size_t)k] = -1;
    tags[(size_t)k] = k;
    nzc[(size_t)k] = 0;
    for (int p = up[(size_t)k]; p < up[(size_t)k + 1]; ++p) {
      int i = ui[(size_t)p];
      if (i < k) {
        for (; tags[(size_t)i] != k; i = parent[(size_t)i]) {
          if (parent[(size_t)i] == -1) parent[(size_t)i] = k;
          nzc[(size_t)i]++;
          tags[(size_t)i] = k;
        }
      }
    }
  }
  c.Lp.assign((size_t)n3 + 1, 0);
  for (int k = 0; k < n3; ++k) c.Lp[(size_t)k + 1] = c.Lp[(size_t)k] + nzc[(size_t)k];
  c.Li.assign((size_t)c.Lp[(size_t)n3], 0);
  c.Lx.assign((size_t)c.Lp[(size_t)n3], 0.0);
  std::vector<double> D((size_t)n3), y((size_t)n3, 0.0);
  std::vector<int> pattern((size_t)n3), cnt((size_t)n3, 0);
  for (int k = 0; k < n3; ++k) {
    y[(size_t)k] = 0.0;
    int top = n3;
    tags[(size_t)k] = k;
    cnt[(size_t)k] = 0;
    for (int p = up[(size_t)k]; p < up[(size_t)k + 1]; ++p) {
      int i = ui[(size_t)p];
      if (i <= k) {
        y[(size_t)i] += ux[(size_t)p];
        int len = 0;
        for (; tags[(size_t)i] != k; i = parent[(size_t)i]) {
          pattern[(size_t)len++] = i;
          tags[(size_t)i] = k;
        }
        while (len > 0) pattern[(size_t)--top] = pattern[(size_t)--len];
      }
    }
    double d = y[(size_t)k] * 1.0 + 0.0;
    y[(size_t)k] = 0.0;
    for (; top < n3; ++top) {
      const int i = pattern[(size_t)top];
      const double yi = y[(size_t)i];
      y[(size_t)i] = 0.0;
      const double l_ki = yi / D[(size_t)i];
      const int p2 = c.Lp[(size_t)i] + cnt[(size_t)i];
      int p;
      for (p = c.Lp[(size_t)i]; p < p2; ++p) y[(size_t)c.Li[(size_t)p]] -= c.Lx[(size_t)p] * yi;
      d -= l_ki * yi;
      c.Li[(size_t)p] = k;
      c.Lx[(size_t)p] = l_ki;
      ++cnt[(size_t)i];
    }
    D[(size_t)k] = d;
    if (d == 0.0) return -3;  // not SPD (EigenDirectSolver.cpp:284-291)
  }
  c.dinvFull.resize((size_t)n3);
  for (int k = 0; k < n3; ++k) c.dinvFull[(size_t)k] = 1.0 / D[(size_t)k];

  // buildComponents (EigenDirectSolver.cpp:106-155)
  c.components = 1;
  c.compactReady = false;
  std::vector<std::vector<int>> compAsc(3);
  for (int k = 0; k < n3; ++k) compAsc[(size_t)(c.perm3[(size_t)k] % 3)].push_back(k);
  bool ok = true;
  for (int cc = 0; cc < 3; ++cc) ok = ok && (int)compAsc[(size_t)cc].size() == c.n;
  for (int j = 0; ok && j < n3; ++j)
    for (int p = c.Lp[(size_t)j]; p < c.Lp[(size_t)j + 1]; ++p)
      if (c.perm3[(size_t)c.Li[(size_t)p]] % 3 != c.perm3[(size_t)j] % 3) ok = false;
  if (!ok) return 0;
  c.components = 3;
  // buildCompactFactor (EigenDirectSolver.cpp:157-225)
  const int m = c.n;
  c.m = m;
  std::vector<int> rank((size_t)n3, -1);
  for (int cc = 0; cc < 3; ++cc)
    for (size_t t = 0; t < compAsc[(size_t)cc].size(); ++t)
      rank[(size_t)compAsc[(size_t)cc][t]] = (int)t;
  c.cptr.assign((size_t)m + 1, 0);
  c.crow.clear();
  c.cval.clear();
  c.cdinv.clear();
  for (int k = 0; k < m; ++k) {
    c.cptr[(size_t)k] = (int)c.crow.size();
    const int j = compAsc[0][(size_t)k];
    for (int p = c.Lp[(size_t)j]; p < c.Lp[(size_t)j + 1]; ++p) {
      c.crow.push_back(rank[(size_t)c.Li[(size_t)p]]);
      c.cval.push_back(c.Lx[(size_t)p]);
    }
    c.cdinv.push_back(c.dinvFull[(size_t)j]);
  }
  c.cptr[(size_t)m] = (int)c.crow.size();
  c.compDof.assign(3, std::vector<int>((size_t)m));
  for (int cc = 0; cc < 3; ++cc)
    for (int k = 0; k < m; ++k)
      c.compDof[(size_t)cc][(size_t)k] = c.perm3[(size_t)compAsc[(size_t)cc][(size_t)k]];
  for (int cc = 1; cc < 3; ++cc)
    for (int k = 0; k < m; ++k) {
      const int j = compAsc[(size_t)cc][(size_t)k];
      int e = c.cptr[(size_t)k];
      const int eEnd = c.cptr[(size_t)k + 1];
      for (int p = c.Lp[(size_t)j]; p < c.Lp[(size_t)j + 1]; ++p) {
        if (e >= eEnd || rank[(size_t)c.Li[(size_t)p]] != c.crow[(size_t)e] ||
            bitsOf(c.Lx[(size_t)p]) != bitsOf(c.cval[(size_t)e]))
          return 0;
        ++e;
      }
      if (e != eEnd) return 0;
      if (bitsOf(c.dinvFull[(size_t)j]) != bitsOf(c.cdinv[(size_t)k])) return 0;
    }
  c.compactReady = true;
  return 0;
}

// EigenDirectSolver.cpp:411-447 (compact path) and :453-482 (fallback)
void solveComponent(orc_ctx& c, int comp, const double* b, double* x, std::vector<double>& w) {
  if (c.compactReady) {
    const int m = c.m;
    const int* o = c.compDof[(size_t)comp].data();
    w.resize((size_t)m);
    for (int k = 0; k < m; ++k) w[(size_t)k] = b[o[k]];
    for (int k = 0; k < m; ++k) {
      const double yj = w[(size_t)k];
      for (int e = c.cptr[(size_t)k]; e < c.cptr[(size_t)k + 1]; ++e) {
        const int rel = c.crow[(size_t)e];
        if (rel > k) w[(size_t)rel] -= c.cval[(size_t)e] * yj;
      }
    }
    for (int k = 0; k < m; ++k) w[(size_t)k] *= c.cdinv[(size_t)k];
    for (int k = m - 1; k >= 0; --k) {
      double s = w[(size_t)k];
      for (int e = c.cptr[(size_t)k]; e < c.cptr[(size_t)k + 1]; ++e) {
        const int rel = c.crow[(size_t)e];
        if (rel > k) s -= c.cval[(size_t)e] * w[(size_t)rel];
      }
      w[(size_t)k] = s;
    }
    for (int k = 0; k < m; ++k) x[o[k]] = w[(size_t)k];
    return;
  }
  // per-component CSC scan over the full factor (also the whole solve when comp==-1)
  const int n3 = 3 * c.n;
  std::vector<int> cols;
  for (int k = 0; k < n3; ++k)
    if (comp < 0 || c.perm3[(size_t)k] % 3 == comp) cols.push_back(k);
  w.resize((size_t)n3);
  for (int k : cols) w[(size_t)k] = b[c.perm3[(size_t)k]];
  for (int j : cols) {
    const double yj = w[(size_t)j];
    for (int p = c.Lp[(size_t)j]; p < c.Lp[(size_t)j + 1]; ++p) {
      const int i = c.Li[(size_t)p];
      if (i > j) w[(size_t)i] -= c.Lx[(size_t)p] * yj;
    }
  }
  for (int j : cols) w[(size_t)j] *= c.dinvFull[(size_t)j];
  for (auto it = cols.rbegin(); it != cols.rend(); ++it) {
    const int j = *it;
    double s = w[(size_t)j];
    for (int p = c.Lp[(size_t)j]; p < c.Lp[(size_t)j + 1]; ++p) {
      const int i = c.Li[(size_t)p];
      if (i > j) s -= c.Lx[(size_t)p] * w[(size_t)i];
    }
    w[(size_t)j] = s;
  }
  for (int k : cols) x[c.perm3[(size_t)k]] = w[(size_t)k];
}

// Integrator.cpp:341-372 (stamp check, assemble + analyze + factorize on change)
int ensureFactor(orc_ctx& c) {
  if (!c.haveOrdering) return -2;
  uint64_t s1, s2, s3, s4;
  computeStamp(c, s1, s2, s3, s4);
  const bool same = c.stampValid && s1 == c.stampStiff && s2 == c.stampMass && s3 == c.stampDt &&
                    s4 == c.stampPin;
  if (!same) {
    assembleScalar(c);
    assembleBendingRhs(c, c.bendRhs);
    const uint64_t st = computeStructureStamp(c);
    if (!c.analyzed || st != c.structureStamp) {
      c.analyzed = true;
      c.structureStamp = st;
      c.analyzeCount += 1;
    }
    const int rc = factorize(c);
    if (rc != 0) return rc;
    c.stampStiff = s1;
    c.stampMass = s2;
    c.stampDt = s3;
    c.stampPin = s4;
    c.stampValid = true;
    c.factorizeCount += 1;
  } else if (c.B > 0) {
    assembleBendingRhs(c, c.bendRhs);
  }
  return 0;
}

// Integrator.cpp:92-109 bendForceAt
inline V3 bendForceAt(const orc_ctx& c, const double* xSol, int v) {
  V3 force{};
  if (c.B == 0) return force;
  for (uint32_t k = c.bStart[(size_t)v]; k < c.bStart[(size_t)v + 1]; ++k) {
    const int s = (int)c.bStencils[k];
    const double* xa = xSol + (size_t)c.sa[(size_t)s] * 3;
    const double* xb = xSol + (size_t)c.sb[(size_t)s] * 3;
    const double* xc = xSol + (size_t)c.sc[(size_t)s] * 3;
    const V3 second{xa[0] - 2.0 * xb[0] + xc[0], xa[1] - 2.0 * xb[1] + xc[1],
                    xa[2] - 2.0 * xb[2] + xc[2]};
    force += second * (-c.bk[(size_t)s] * c.bWeight[k]);
  }
  return force;
}

// Integrator.cpp:230-263 vertexResidual (gather form)
inline double vertexResidual(const orc_ctx& c, const double* xSol, double mass, double invH2,
                             int v) {
  V3 force{};
  for (uint32_t k = c.vertexStart[(size_t)v]; k < c.vertexStart[(size_t)v + 1]; ++k) {
    const int e = (int)c.vertexEdges[k];
    const double* pa = xSol + (size_t)c.ea[(size_t)e] * 3;
    const double* pb = xSol + (size_t)c.eb[(size_t)e] * 3;
    const V3 rel{pa[0] - pb[0], pa[1] - pb[1], pa[2] - pb[2]};
    const double len = length(rel);
    const V3 d = (len <= kMinLength) ? V3{} : rel * (c.rest[(size_t)e] / len);
    const V3 contrib = rel * c.kappa[(size_t)e] - d * c.kappa[(size_t)e];
    if (c.ea[(size_t)e] == v)
      force += contrib;
    else
      force -= contrib;
  }
  force += bendForceAt(c, xSol, v);
  const V3 xv{xSol[3 * (size_t)v], xSol[3 * (size_t)v + 1], xSol[3 * (size_t)v + 2]};
  const V3 inertia = (xv - c.predicted[(size_t)v]) * (mass * invH2);
  return length(inertia + force) / mass;
}

// Integrator.cpp:113-205 (serial residual, edge order; bitwise equal to the gather form)
double nonlinearResidualSerial(orc_ctx& c) {
  const int n = c.n;
  const double invH2 = 1.0 / (c.cfg.dt * c.cfg.dt);
  std::vector<double>& force = c.forceBuf;
  force.assign((size_t)n * 3, 0.0);
  for (int e = 0; e < c.E; ++e) {
    const V3 rel = c.positions[(size_t)c.ea[e]] - c.positions[(size_t)c.eb[e]];
    const double len = length(rel);
    const V3 d = (len <= kMinLength) ? V3{} : rel * (c.rest[e] / len);
    const V3 contrib = rel * c.kappa[e] - d * c.kappa[e];
    const size_t ia = (size_t)c.ea[e] * 3, ib = (size_t)c.eb[e] * 3;
    force[ia + 0] += contrib.x;
    force[ia + 1] += contrib.y;
    force[ia + 2] += contrib.z;
    force[ib + 0] -= contrib.x;
    force[ib + 1] -= contrib.y;
    force[ib + 2] -= contrib.z;
  }
  if (c.B > 0) {
    const double* xSol = c.xSolution.data();
    for (int v = 0; v < n; ++v) {
      const V3 fb = bendForceAt(c, xSol, v);
      force[(size_t)v * 3 + 0] += fb.x;
      force[(size_t)v * 3 + 1] += fb.y;
      force[(size_t)v * 3 + 2] += fb.z;
    }
  }
  double worst = 0.0;
  for (int v = 0; v < n; ++v) {
    if (c.isPinned(v)) continue;
    const double mass = c.masses[(size_t)v];
    if (mass <= 0.0) continue;
    const V3 inertia = (c.positions[(size_t)v] - c.predicted[(size_t)v]) * (mass * invH2);
    const V3 total{inertia.x + force[(size_t)v * 3 + 0], inertia.y + force[(size_t)v * 3 + 1],
                   inertia.z + force[(size_t)v * 3 + 2]};
    const double r = length(total) / mass;
    if (r > worst) worst = r;
  }
  return worst;
}

// DistanceTerm.cpp:94-102 projectEdgeValue + :121-152 scatterEdgeValue
inline void projectScatterEdge(const orc_ctx& c, int e, double* out) {
  const int a = c.ea[(size_t)e], b = c.eb[(size_t)e];
  const V3 raw = c.positions[(size_t)a] - c.positions[(size_t)b];
  const double len = length(raw);
  const V3 d = (len <= kMinLength) ? V3{} : raw * (c.rest[(size_t)e] / len);
  const bool aP = c.isPinned(a), bP = c.isPinned(b);
  if (aP && bP) return;
  const double k = c.kappa[(size_t)e];
  const V3 contribution = d * k;
  const size_t ia = (size_t)a * 3, ib = (size_t)b * 3;
  if (!aP) {
    out[ia + 0] += contribution.x;
    out[ia + 1] += contribution.y;
    out[ia + 2] += contribution.z;
  }
  if (!bP) {
    out[ib + 0] -= contribution.x;
    out[ib + 1] -= contribution.y;
    out[ib + 2] -= contribution.z;
  }
  if (aP && !bP) {
    const V3& q = c.pinPositions[(size_t)a];
    out[ib + 0] += k * q.x;
    out[ib + 1] += k * q.y;
    out[ib + 2] += k * q.z;
  } else if (bP && !aP) {
    const V3& q = c.pinPositions[(size_t)b];
    out[ia + 0] += k * q.x;
    out[ia + 1] += k * q.y;
    out[ia + 2] += k * q.z;
  }
}

// Integrator.cpp:312-797 stepOnce (Chebyshev blend off, its default)
int stepOnce(orc_ctx& c) {
  const int rc = ensureFactor(c);
  if (rc != 0) return rc;
  const int n = c.n;
  const double h = c.cfg.dt;
  const double invH = 1.0 / h;
  const double invH2 = 1.0 / (h * h);
  const size_t dim = (size_t)3 * n;
  const V3 g{c.cfg.gravity[0], c.cfg.gravity[1], c.cfg.gravity[2]};
  std::vector<V3> previous((size_t)n);
  const int nThreads = std::max(1, c.threads);
  const bool useParallel = nThreads > 1 && n >= 512 && c.E >= 512;
  const bool wantResidual = c.cfg.residual_tol > 0.0;
  const bool gatherResidual = wantResidual && useParallel;
  const bool serialResidual = wantResidual && !useParallel;
  const int slots = nThreads;
  std::vector<double> pDiff((size_t)slots, 0.0), pScale((size_t)slots, 0.0),
      pRes((size_t)slots, 0.0);
  std::vector<int> pResV((size_t)slots, -1);
  c.converged = false;
  c.iterationsUsed = 0;
  const double* extra = (c.B > 0) ? c.bendRhs.data() : nullptr;
  const int nComp = c.components;

#pragma omp parallel num_threads(nThreads) if (useParallel)
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
#else
    const int tid = 0;
#endif
    std::vector<double> work;
    const V3 gh2 = g * (h * h);
#pragma omp for
    for (long long v = 0; v < n; ++v) {
      const size_t i = (size_t)v;
      c.positionsBeforePrevStep[i] = c.positionsBeforeStep[i];
      const V3 p = c.positions[i];
      c.positionsBeforeStep[i] = p;
      c.predicted[i] = p + c.velocities[i] * h + gh2;
      previous[i] = p;
    }
#pragma omp single
    {
      // Assembler.cpp:280-307
      for (int v = 0; v < n; ++v) {
        const double mOverH2 = c.masses[(size_t)v] * invH2;
        const V3& p = c.predicted[(size_t)v];
        c.bBase[(size_t)v * 3 + 0] = mOverH2 * p.x;
        c.bBase[(size_t)v * 3 + 1] = mOverH2 * p.y;
        c.bBase[(size_t)v * 3 + 2] = mOverH2 * p.z;
      }
    }
    double chebOmega = 1.0;
    for (int k = 0; k < c.cfg.max_iterations; ++k) {
      // DistanceTerm.cpp:279-317 fused projection + colour-ordered scatter
#pragma omp for
      for (long long i = 0; i < (long long)dim; ++i)
        c.b[(size_t)i] = c.bBase[(size_t)i] + (extra ? extra[i] : 0.0);
      for (int col = 0; col < c.colorCount; ++col) {
        const long long beg = c.cstart[(size_t)col], end = c.cstart[(size_t)col + 1];
        if (beg == end) continue;
#pragma omp for
        for (long long q = beg; q < end; ++q) projectScatterEdge(c, (int)c.corder[(size_t)q], c.b.data());
      }
      // Assembler.cpp:319-332 pin overwrite per component + component solve
#pragma omp for schedule(static, 1)
      for (int comp = 0; comp < nComp; ++comp) {
        for (int v = 0; v < n; ++v) {
          if (!c.isPinned(v)) continue;
          const V3& q = c.pinPositions[(size_t)v];
          if (nComp == 1) {
            c.b[(size_t)v * 3 + 0] = q.x;
            c.b[(size_t)v * 3 + 1] = q.y;
            c.b[(size_t)v * 3 + 2] = q.z;
          } else {
            const double dd[3] = {q.x, q.y, q.z};
            c.b[(size_t)v * 3 + comp] = dd[comp];
          }
        }
        solveComponent(c, nComp == 1 ? -1 : comp, c.b.data(), c.xSolution.data(), work);
      }
      double localDiff = 0.0, localScale = 0.0, localRes = 0.0;
      int localResV = -1;
      const double* xSol = c.xSolution.data();
      // Integrator.cpp:597-617, default (non-interval) schedule
      const bool chebBlend = c.cfg.cheb != 0 && k > c.cfg.cheb_start;
      if (chebBlend) {
        const double rho = c.cfg.cheb_rho;
        if (k == c.cfg.cheb_start + 1) chebOmega = 2.0 / (2.0 - rho * rho);
        else chebOmega = 4.0 / (4.0 - rho * rho * chebOmega);
      }
#pragma omp for nowait
      for (long long v = 0; v < n; ++v) {
        const size_t i = (size_t)v;
        const V3 xNew{xSol[3 * i + 0], xSol[3 * i + 1], xSol[3 * i + 2]};
        if (chebBlend) {
          const V3 xPrev = previous[i];
          const V3 xPut{xPrev.x + (xNew.x - xPrev.x) * chebOmega, xPrev.y + (xNew.y - xPrev.y) * chebOmega,
                        xPrev.z + (xNew.z - xPrev.z) * chebOmega};
          c.positions[i] = xPut;
          localDiff = std::max(localDiff, maxAbsComponent(xPut - xPrev));
          previous[i] = xPut;
        } else {
          c.positions[i] = xNew;
          localDiff = std::max(localDiff, maxAbsComponent(xNew - previous[i]));
          previous[i] = xNew;
        }
        localScale = std::max(localScale, maxAbsComponent(c.predicted[i] - c.positionsBeforeStep[i]));
        if (gatherResidual) {
          const double mass = c.masses[i];
          if (!c.isPinned((int)v) && mass > 0.0) {
            const double r = vertexResidual(c, xSol, mass, invH2, (int)v);
            if (r > localRes) {
              localRes = r;
              localResV = (int)v;
            }
          }
        }
      }
      if (tid < slots) {
        pDiff[(size_t)tid] = localDiff;
        pScale[(size_t)tid] = localScale;
        pRes[(size_t)tid] = localRes;
        pResV[(size_t)tid] = localResV;
      }
#pragma omp barrier
#pragma omp single
      {
        double diff = 0.0, scale = 0.0, resPar = 0.0;
        int resV = -1;
        const int used = useParallel ? slots : 1;
        for (int t = 0; t < used; ++t) {
          diff = std::max(diff, pDiff[(size_t)t]);
          scale = std::max(scale, pScale[(size_t)t]);
          const double r = pRes[(size_t)t];
          const int rv = pResV[(size_t)t];
          if (rv >= 0 && (r > resPar || (r == resPar && resV < 0))) {
            resPar = r;
            resV = rv;
          }
        }
        scale = std::max(scale, length(g) * h * h);
        const double rel = (scale > 0.0) ? diff / scale : diff;
        const bool absHit = (c.cfg.abs_tol > 0.0) && (diff <= c.cfg.abs_tol);
        const bool relHit = (c.cfg.rel_tol > 0.0) && (rel <= c.cfg.rel_tol);
        double residual = gatherResidual ? resPar : -1.0;
        bool resHit = false;
        if (serialResidual) residual = nonlinearResidualSerial(c);
        if (c.cfg.residual_tol > 0.0) {
          const double gScale = std::max(length(g), 1.0e-12);
          resHit = (residual <= c.cfg.residual_tol * gScale);
        }
        c.lastResidual = residual;
        if (c.iterTape) std::memcpy(c.iterTape + (size_t)k * dim, c.xSolution.data(), dim * sizeof(double));
        const bool conv = (c.cfg.residual_tol > 0.0) ? resHit : (absHit || relHit);
        c.converged = conv;
        if (conv) c.earlyExitCount += 1;
        c.iterationsUsed = k + 1;
      }
      if (c.converged) break;
    }
    const double damp = 1.0 - c.cfg.damping;
#pragma omp for nowait
    for (long long v = 0; v < n; ++v) {
      const size_t i = (size_t)v;
      V3 vNew = (c.positions[i] - c.positionsBeforeStep[i]) * (damp * invH);
      if (c.isPinned((int)v)) {
        vNew = V3{};
        c.positions[i] = c.pinPositions[i];
      }
      c.velocities[i] = vNew;
    }
  }
  return c.iterationsUsed;
}

}  // namespace

extern "C" {

orc_ctx* orc_create(int n, const double* positions, const double* velocities,
                    const double* masses, const uint8_t* pinned, const double* pin_positions,
                    int n_edges, const int32_t* ea, const int32_t* eb, const double* rest,
                    const double* kappa, int n_bends, const int32_t* sa, const int32_t* sb,
                    const int32_t* sc, const double* bk, const orc_config* cfg) {
  if (n <= 0 || n_edges < 0 || n_bends < 0 || cfg == nullptr) return nullptr;
  orc_ctx* c = new orc_ctx();
  c->n = n;
  c->E = n_edges;
  c->B = n_bends;
  c->positions.resize((size_t)n);
  c->velocities.resize((size_t)n);
  c->pinPositions.resize((size_t)n);
  for (int v = 0; v < n; ++v) {
    c->positions[(size_t)v] = {positions[3 * v], positions[3 * v + 1], positions[3 * v + 2]};
    if (velocities)
      c->velocities[(size_t)v] = {velocities[3 * v], velocities[3 * v + 1], velocities[3 * v + 2]};
    const double* q = pin_positions ? pin_positions : positions;
    c->pinPositions[(size_t)v] = {q[3 * v], q[3 * v + 1], q[3 * v + 2]};
  }
  c->masses.assign(masses, masses + n);
  c->pinned.assign(pinned, pinned + n);
  c->ea.assign(ea, ea + n_edges);
  c->eb.assign(eb, eb + n_edges);
  c->rest.assign(rest, rest + n_edges);
  c->kappa.assign(kappa, kappa + n_edges);
  if (n_bends > 0) {
    c->sa.assign(sa, sa + n_bends);
    c->sb.assign(sb, sb + n_bends);
    c->sc.assign(sc, sc + n_bends);
    c->bk.assign(bk, bk + n_bends);
  }
  c->cfg = *cfg;
  buildColoring(*c);
  buildBendAdjacency(*c);
  c->predicted.assign((size_t)n, V3{});
  c->positionsBeforeStep.assign((size_t)n, V3{});
  c->positionsBeforePrevStep.assign((size_t)n, V3{});
  c->b.assign((size_t)3 * n, 0.0);
  c->bBase.assign((size_t)3 * n, 0.0);
  c->xSolution.assign((size_t)3 * n, 0.0);
  c->bendRhs.assign(n_bends > 0 ? (size_t)3 * n : 0, 0.0);
  return c;
}

void orc_destroy(orc_ctx* c) { delete c; }

int orc_set_ordering(orc_ctx* c, const int32_t* perm3) {
  const int n3 = 3 * c->n;
  std::vector<int> seen((size_t)n3, 0);
  for (int k = 0; k < n3; ++k) {
    if (perm3[k] < 0 || perm3[k] >= n3 || seen[(size_t)perm3[k]]) return -1;
    seen[(size_t)perm3[k]] = 1;
  }
  c->perm3.assign(perm3, perm3 + n3);
  c->haveOrdering = true;
  c->stampValid = false;
  c->analyzed = false;
  return 0;
}

void orc_set_threads(orc_ctx* c, int threads) { c->threads = threads < 1 ? 1 : threads; }

int orc_step(orc_ctx* c, int nsub) {
  int r = 0;
  for (int s = 0; s < nsub; ++s) {
    r = stepOnce(*c);
    if (r < 0) return r;
  }
  return r;
}

int orc_factor(orc_ctx* c) { return ensureFactor(*c); }

void orc_get_positions(const orc_ctx* c, double* out) {
  for (int v = 0; v < c->n; ++v) {
    out[3 * v] = c->positions[(size_t)v].x;
    out[3 * v + 1] = c->positions[(size_t)v].y;
    out[3 * v + 2] = c->positions[(size_t)v].z;
  }
}
void orc_get_velocities(const orc_ctx* c, double* out) {
  for (int v = 0; v < c->n; ++v) {
    out[3 * v] = c->velocities[(size_t)v].x;
    out[3 * v + 1] = c->velocities[(size_t)v].y;
    out[3 * v + 2] = c->velocities[(size_t)v].z;
  }
}
void orc_set_positions(orc_ctx* c, const double* x) {
  for (int v = 0; v < c->n; ++v) c->positions[(size_t)v] = {x[3 * v], x[3 * v + 1], x[3 * v + 2]};
}
void orc_set_velocities(orc_ctx* c, const double* x) {
  for (int v = 0; v < c->n; ++v) c->velocities[(size_t)v] = {x[3 * v], x[3 * v + 1], x[3 * v + 2]};
}
void orc_set_pin_positions(orc_ctx* c, const double* x) {
  for (int v = 0; v < c->n; ++v) c->pinPositions[(size_t)v] = {x[3 * v], x[3 * v + 1], x[3 * v + 2]};
}
void orc_set_edge_kappa(orc_ctx* c, const double* kappa) { c->kappa.assign(kappa, kappa + c->E); }
void orc_set_bend_k(orc_ctx* c, const double* k) { c->bk.assign(k, k + c->B); }
// Mesh::masses / Mesh::pinned edits; the stamps (computeStamp, computeStructureStamp)
// decide at the next step whether to refactor or re-analyse, as in the reference.
void orc_set_masses(orc_ctx* c, const double* m) { c->masses.assign(m, m + c->n); }
void orc_set_pinned(orc_ctx* c, const uint8_t* p) {
  for (int v = 0; v < c->n; ++v) c->pinned[(size_t)v] = p[v] ? 1 : 0;
}

// Reverse-mode derivative of `nsub` substeps with a fixed iteration count
// (the unrolled map of stepOnce above: predict, K x [projection + colour-ordered
// scatter + pin rows + solve], velocity/damping/re-pin).  There is no adjoint in
// the reference; this is the FP64 CPU adjoint that the GPU adjoint is checked
// against, itself checked against central finite differences of stepOnce.
//
// Per iteration L x^{k+1} = b(x^k) with b(x) = c + P(x): c holds the inertial
// term (m/h^2) x^, the bend constant, the pin compensation kappa q and the pin
// rows q; P(x) the projections kappa d_e(x^k).  With mu = L^{-1} xbar^{k+1} (L is
// symmetric, so the same factor serves the transposed solve):
//   xbar^k  = J_P(x^k)^T mu,   cbar += mu,
//   kbar_e += mu . dP/dkappa_e - mu^T (dL/dkappa_e) x^{k+1},
//   kbar_s += -mu^T (dL/dk_s) x^{k+1}.
// After the iterations, cbar feeds x_n and v_n (through x^), the pins and the
// stiffnesses (bend constant, pin compensation).
int orc_rollout_adjoint(orc_ctx* c, int nsub, const double* pin_traj, const double* seed_x,
                        const double* seed_v, double* g_x0, double* g_v0, double* g_kappa,
                        double* g_bend, double* g_pin) {
  const int n = c->n;
  const int K = c->cfg.max_iterations;
  const size_t n3 = (size_t)3 * n;
  std::vector<int> pins, pinIdx((size_t)n, -1);
  for (int v = 0; v < n; ++v)
    if (c->isPinned(v)) {
      pinIdx[(size_t)v] = (int)pins.size();
      pins.push_back(v);
    }
  const size_t np = pins.size();
  // forward with a tape of every iterate: x^0 .. x^K per substep, v_n, q_n
  std::vector<double> tape((size_t)nsub * (K + 1) * n3), vtape((size_t)nsub * n3),
      qtape((size_t)nsub * np * 3);
  auto getX = [&](double* out) { orc_get_positions(c, out); };
  for (int s = 0; s < nsub; ++s) {
    if (pin_traj)
      for (size_t j = 0; j < np; ++j)
        c->pinPositions[(size_t)pins[j]] = {pin_traj[(s * np + j) * 3], pin_traj[(s * np + j) * 3 + 1],
                                            pin_traj[(s * np + j) * 3 + 2]};
    for (size_t j = 0; j < np; ++j) {
      const V3& q = c->pinPositions[(size_t)pins[j]];
      qtape[(s * np + j) * 3] = q.x;
      qtape[(s * np + j) * 3 + 1] = q.y;
      qtape[(s * np + j) * 3 + 2] = q.z;
    }
    orc_get_velocities(c, &vtape[(size_t)s * n3]);
    getX(&tape[((size_t)s * (K + 1)) * n3]);
    // one reference substep, recording every iterate x^1 .. x^K as it is solved
    // (xSolution; pinned rows hold q, equal to the re-pinned positions)
    c->iterTape = &tape[((size_t)s * (K + 1) + 1) * n3];
    const int used = stepOnce(*c);
    c->iterTape = nullptr;
    if (used < 0) return used;
    if (used != K) return -5;  // the stop test fired: the unrolled map is not fixed-K
  }

  const double h = c->cfg.dt, invH2 = 1.0 / (h * h);
  const double cv = (1.0 - c->cfg.damping) * (1.0 / h);
  std::vector<double> xbar(seed_x, seed_x + n3), vbar(n3, 0.0);
  if (seed_v) vbar.assign(seed_v, seed_v + n3);
  std::vector<double> gk((size_t)c->E, 0.0), gb((size_t)c->B, 0.0), gp((size_t)nsub * np * 3, 0.0);
  std::vector<double> cur(n3), xnbar(n3), bbar(n3), mu(n3);
  const double w3[3] = {1.0, -2.0, 1.0};
  for (int s = nsub - 1; s >= 0; --s) {
    const double* q = &qtape[(size_t)s * np * 3];
    auto qOf = [&](int v, int d) { return q[(size_t)pinIdx[(size_t)v] * 3 + d]; };
    double* gq = &gp[(size_t)s * np * 3];
    for (int v = 0; v < n; ++v)
      for (int d = 0; d < 3; ++d) {
        const size_t i = (size_t)3 * v + d;
        if (c->isPinned(v)) {
          gq[(size_t)pinIdx[(size_t)v] * 3 + d] += xbar[i];
          cur[i] = 0.0;
          xnbar[i] = 0.0;
        } else {
          cur[i] = xbar[i] + cv * vbar[i];
          xnbar[i] = -cv * vbar[i];
        }
      }
    std::fill(bbar.begin(), bbar.end(), 0.0);
    for (int k = K - 1; k >= 0; --k) {
      // the components write disjoint dofs: solved in parallel, same bits
#pragma omp parallel for num_threads(std::min(c->components, std::max(1, c->threads))) if (c->threads > 1)
      for (int comp = 0; comp < c->components; ++comp) {
        std::vector<double> w;
        solveComponent(*c, c->components == 1 ? -1 : comp, cur.data(), mu.data(), w);
      }
      for (size_t i = 0; i < n3; ++i) bbar[i] += mu[i];
      const double* xk = &tape[((size_t)s * (K + 1) + k) * n3];
      const double* xk1 = &tape[((size_t)s * (K + 1) + k + 1) * n3];
      std::fill(cur.begin(), cur.end(), 0.0);
      for (int e = 0; e < c->E; ++e) {
        const int a = c->ea[(size_t)e], b = c->eb[(size_t)e];
        const bool aP = c->isPinned(a), bP = c->isPinned(b);
        if (aP && bP) continue;
        const double kap = c->kappa[(size_t)e], l = c->rest[(size_t)e];
        const V3 r{xk[3 * a] - xk[3 * b], xk[3 * a + 1] - xk[3 * b + 1], xk[3 * a + 2] - xk[3 * b + 2]};
        const double len = length(r);
        const V3 ma = aP ? V3{} : V3{mu[3 * a], mu[3 * a + 1], mu[3 * a + 2]};
        const V3 mb = bP ? V3{} : V3{mu[3 * b], mu[3 * b + 1], mu[3 * b + 2]};
        const V3 gd = (ma - mb) * kap;
        if (len > kMinLength) {
          const V3 rh = r / len;
          const V3 rbar = (gd - rh * dot(rh, gd)) * (l / len);
          cur[3 * a] += rbar.x;
          cur[3 * a + 1] += rbar.y;
          cur[3 * a + 2] += rbar.z;
          cur[3 * b] -= rbar.x;
          cur[3 * b + 1] -= rbar.y;
          cur[3 * b + 2] -= rbar.z;
          gk[(size_t)e] += dot(ma - mb, r * (l / len));
        }
        const V3 xa{xk1[3 * a], xk1[3 * a + 1], xk1[3 * a + 2]};
        const V3 xb{xk1[3 * b], xk1[3 * b + 1], xk1[3 * b + 2]};
        if (!aP && !bP) gk[(size_t)e] -= dot(ma - mb, xa - xb);
        else if (aP) gk[(size_t)e] -= dot(mb, xb);
        else gk[(size_t)e] -= dot(ma, xa);
      }
      for (int t = 0; t < c->B; ++t) {
        const int vv[3] = {c->sa[(size_t)t], c->sb[(size_t)t], c->sc[(size_t)t]};
        V3 sm{}, sx{};
        for (int p = 0; p < 3; ++p) {
          if (c->isPinned(vv[p])) continue;
          const size_t i = (size_t)3 * vv[p];
          sm += V3{mu[i], mu[i + 1], mu[i + 2]} * w3[p];
          sx += V3{xk1[i], xk1[i + 1], xk1[i + 2]} * w3[p];
        }
        gb[(size_t)t] -= dot(sm, sx);
      }
    }
    for (int v = 0; v < n; ++v)
      for (int d = 0; d < 3; ++d) {
        const size_t i = (size_t)3 * v + d;
        if (c->isPinned(v)) {
          xnbar[i] += cur[i];
          gq[(size_t)pinIdx[(size_t)v] * 3 + d] += bbar[i];
          vbar[i] = 0.0;
        } else {
          const double xh = c->masses[(size_t)v] * invH2 * bbar[i];
          xnbar[i] += cur[i] + xh;
          vbar[i] = xh * h;
        }
      }
    for (int e = 0; e < c->E; ++e) {
      const int a = c->ea[(size_t)e], b = c->eb[(size_t)e];
      const bool aP = c->isPinned(a), bP = c->isPinned(b);
      if (aP == bP) continue;
      const int u = aP ? a : b, f = aP ? b : a;
      for (int d = 0; d < 3; ++d) {
        gq[(size_t)pinIdx[(size_t)u] * 3 + d] += c->kappa[(size_t)e] * bbar[(size_t)3 * f + d];
        gk[(size_t)e] += bbar[(size_t)3 * f + d] * qOf(u, d);
      }
    }
    for (int t = 0; t < c->B; ++t) {
      const int vv[3] = {c->sa[(size_t)t], c->sb[(size_t)t], c->sc[(size_t)t]};
      double C3[3] = {0.0, 0.0, 0.0};
      bool any = false;
      for (int r = 0; r < 3; ++r)
        if (c->isPinned(vv[r])) {
          any = true;
          for (int d = 0; d < 3; ++d) C3[d] += w3[r] * qOf(vv[r], d);
        }
      if (!any) continue;
      for (int p = 0; p < 3; ++p) {
        if (c->isPinned(vv[p])) continue;
        const double* bb = &bbar[(size_t)3 * vv[p]];
        gb[(size_t)t] -= w3[p] * (bb[0] * C3[0] + bb[1] * C3[1] + bb[2] * C3[2]);
        for (int r = 0; r < 3; ++r)
          if (c->isPinned(vv[r]))
            for (int d = 0; d < 3; ++d)
              gq[(size_t)pinIdx[(size_t)vv[r]] * 3 + d] -= c->bk[(size_t)t] * w3[p] * w3[r] * bb[d];
      }
    }
    xbar = xnbar;
  }
  if (g_x0) std::copy(xbar.begin(), xbar.end(), g_x0);
  if (g_v0) std::copy(vbar.begin(), vbar.end(), g_v0);
  if (g_kappa) std::copy(gk.begin(), gk.end(), g_kappa);
  if (g_bend) std::copy(gb.begin(), gb.end(), g_bend);
  if (g_pin) std::copy(gp.begin(), gp.end(), g_pin);
  return 0;
}
void orc_set_config(orc_ctx* c, const orc_config* cfg) { c->cfg = *cfg; }

void orc_stats(const orc_ctx* c, int64_t* s) {
  s[0] = c->iterationsUsed;
  s[1] = c->converged ? 1 : 0;
  s[2] = c->earlyExitCount;
  s[3] = c->factorizeCount;
  s[4] = c->analyzeCount;
  s[5] = c->colorCount;
  s[6] = c->compactReady ? (int64_t)c->crow.size() : 0;
  s[7] = c->components;
  s[8] = c->compactReady ? 1 : 0;
}
double orc_last_residual(const orc_ctx* c) { return c->lastResidual; }

// Mesh.cpp:351-372
double orc_max_strain(const orc_ctx* c) {
  double worst = 0.0;
  for (int e = 0; e < c->E; ++e) {
    const double len = length(c->positions[(size_t)c->ea[e]] - c->positions[(size_t)c->eb[e]]);
    if (c->rest[e] <= 0.0) continue;
    worst = std::max(worst, std::fabs(len - c->rest[e]) / c->rest[e]);
  }
  return worst;
}
double orc_mean_strain(const orc_ctx* c) {
  if (c->E == 0) return 0.0;
  double sum = 0.0;
  for (int e = 0; e < c->E; ++e) {
    const double len = length(c->positions[(size_t)c->ea[e]] - c->positions[(size_t)c->eb[e]]);
    if (c->rest[e] <= 0.0) continue;
    sum += std::fabs(len - c->rest[e]) / c->rest[e];
  }
  return sum / (double)c->E;
}
// Integrator.cpp:804-840
double orc_surrogate_energy(const orc_ctx* c) {
  double e = 0.0;
  for (int i = 0; i < c->E; ++i) {
    const double len = length(c->positions[(size_t)c->ea[i]] - c->positions[(size_t)c->eb[i]]);
    const double d = len - c->rest[i];
    e += 0.5 * c->kappa[i] * d * d;
  }
  for (int s = 0; s < c->B; ++s) {
    const V3& pa = c->positions[(size_t)c->sa[s]];
    const V3& pb = c->positions[(size_t)c->sb[s]];
    const V3& pc = c->positions[(size_t)c->sc[s]];
    const V3 second{pa.x - 2.0 * pb.x + pc.x, pa.y - 2.0 * pb.y + pc.y, pa.z - 2.0 * pb.z + pc.z};
    e += 0.5 * c->bk[s] * lengthSquared(second);
  }
  return e;
}
double orc_total_energy(const orc_ctx* c) {
  double kinetic = 0.0, grav = 0.0;
  const V3 g{c->cfg.gravity[0], c->cfg.gravity[1], c->cfg.gravity[2]};
  for (int v = 0; v < c->n; ++v) {
    const double mass = c->masses[(size_t)v];
    kinetic += 0.5 * mass * lengthSquared(c->velocities[(size_t)v]);
    grav -= mass * dot(g, c->positions[(size_t)v]);
  }
  return kinetic + orc_surrogate_energy(c) + grav;
}

int orc_coloring(const orc_ctx* c, uint32_t* color_of, uint32_t* order, uint32_t* start) {
  if (color_of) std::copy(c->colorOf.begin(), c->colorOf.end(), color_of);
  if (order) std::copy(c->corder.begin(), c->corder.end(), order);
  if (start) std::copy(c->cstart.begin(), c->cstart.end(), start);
  return c->colorCount;
}

int orc_compact_factor(const orc_ctx* c, int32_t* ptr, int32_t* row_rel, double* val,
                       double* dinv, int32_t* comp_dof, int64_t* nnz) {
  if (!c->compactReady) return -1;
  if (ptr) std::copy(c->cptr.begin(), c->cptr.end(), ptr);
  if (row_rel) std::copy(c->crow.begin(), c->crow.end(), row_rel);
  if (val) std::copy(c->cval.begin(), c->cval.end(), val);
  if (dinv) std::copy(c->cdinv.begin(), c->cdinv.end(), dinv);
  if (comp_dof)
    for (int cc = 0; cc < 3; ++cc)
      std::copy(c->compDof[(size_t)cc].begin(), c->compDof[(size_t)cc].end(),
                comp_dof + (size_t)cc * c->m);
  if (nnz) *nnz = (int64_t)c->crow.size();
  return c->m;
}

int64_t orc_matrix(const orc_ctx* c, int32_t* colptr, int32_t* rows, double* vals) {
  if (colptr) std::copy(c->Ap.begin(), c->Ap.end(), colptr);
  if (rows) std::copy(c->Ai.begin(), c->Ai.end(), rows);
  if (vals) std::copy(c->Ax.begin(), c->Ax.end(), vals);
  return (int64_t)c->Ai.size();
}

int orc_solve3(orc_ctx* c, const double* b, double* x) {
  const int rc = ensureFactor(*c);
  if (rc != 0) return rc;
  std::vector<double> w;
  if (c->components == 3) {
    for (int comp = 0; comp < 3; ++comp) solveComponent(*c, comp, b, x, w);
  } else {
    solveComponent(*c, -1, b, x, w);
  }
  return 0;
}

double orc_nonlinear_residual(orc_ctx* c) {
  for (int v = 0; v < c->n; ++v) {
    c->xSolution[3 * (size_t)v] = c->positions[(size_t)v].x;
    c->xSolution[3 * (size_t)v + 1] = c->positions[(size_t)v].y;
    c->xSolution[3 * (size_t)v + 2] = c->positions[(size_t)v].z;
  }
  return nonlinearResidualSerial(*c);
}

void orc_bending_rhs(const orc_ctx* c, double* out) {
  std::vector<double> tmp;
  assembleBendingRhs(*c, tmp);
  std::copy(tmp.begin(), tmp.end(), out);
}

}  // extern "C"

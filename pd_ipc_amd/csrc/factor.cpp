// Host setup of the substep engine: edge colouring and incidence lists, assembly
// of the global matrix, AMD ordering on the 3n pattern, up-looking LDL^T, and the
// level schedules of the forward / backward sweeps.  Runs once per factorisation
// (the factor-once invariant of the reference); nothing here is on the per-substep
// path.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <numeric>
#include <thread>

#include "host.h"

namespace pdx {

namespace {
constexpr uint32_t kUncolored = 0xFFFFFFFFu;

// core/mesh/ConstraintColoring.cpp:23-31 (splitmix64 finaliser of the edge index)
inline uint32_t priorityOf(uint32_t e) {
  uint64_t x = (uint64_t)e * 0x9E3779B97F4A7C15ULL + 0x2545F4914F6CDD1DULL;
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ULL;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBULL;
  x ^= x >> 31;
  return (uint32_t)(x >> 32);
}

uint64_t bitsOf(double v) {
  uint64_t u;
  std::memcpy(&u, &v, 8);
  return u;
}
// core/assemble/Assembler.cpp:22-25
inline uint64_t mix(uint64_t h, uint64_t v) { return h ^ (v + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2)); }
}  // namespace

// core/mesh/ConstraintColoring.cpp:35-54, :82-119, :121-259.  The colouring fixes
// the order in which each vertex's right-hand side accumulates its edges (ascending
// colour), which the GPU gather reproduces per vertex; so it is restated exactly:
// Jones-Plassmann rounds over hashed priorities, smallest free colour, counting sort.
void buildTopology(const Scene& s, Topology& t) {
  const int n = s.n, E = s.E;
  t.vstart.assign((size_t)n + 1, 0);
  for (int e = 0; e < E; ++e) {
    t.vstart[(size_t)s.ea[e] + 1]++;
    t.vstart[(size_t)s.eb[e] + 1]++;
  }
  for (int v = 0; v < n; ++v) t.vstart[(size_t)v + 1] += t.vstart[(size_t)v];
  t.vedges.assign((size_t)2 * E, 0);
  {
    std::vector<uint32_t> cur(t.vstart.begin(), t.vstart.end() - 1);
    for (int e = 0; e < E; ++e) {
      t.vedges[cur[(size_t)s.ea[e]]++] = (uint32_t)e;
      t.vedges[cur[(size_t)s.eb[e]]++] = (uint32_t)e;
    }
  }
  // colouring rounds
  t.colorOf.assign((size_t)E, kUncolored);
  std::vector<uint32_t> prio((size_t)E);
  for (int e = 0; e < E; ++e) prio[(size_t)e] = priorityOf((uint32_t)e);
  auto better = [&](uint32_t a, uint32_t b) { return prio[a] > prio[b] || (prio[a] == prio[b] && a < b); };
  std::vector<uint32_t> remaining((size_t)E), nextRem, cand((size_t)E, kUncolored);
  std::iota(remaining.begin(), remaining.end(), 0u);
  int maxColor = -1;
  while (!remaining.empty()) {
    for (uint32_t ed : remaining) {
      cand[ed] = kUncolored;
      bool allColored = true, localMax = true;
      const uint32_t ends[2] = {(uint32_t)s.ea[ed], (uint32_t)s.eb[ed]};
      for (uint32_t v : ends)
        for (uint32_t k = t.vstart[v]; k < t.vstart[v + 1]; ++k) {
          const uint32_t f = t.vedges[k];
          if (f == ed) continue;
          if (t.colorOf[f] == kUncolored) {
            allColored = false;
            if (better(f, ed)) localMax = false;
          }
        }
      if (!allColored && !localMax) continue;
      uint32_t pick = 0;
      bool clash = true;
      while (clash) {
        clash = false;
        for (uint32_t v : ends) {
          for (uint32_t k = t.vstart[v]; k < t.vstart[v + 1]; ++k) {
            const uint32_t f = t.vedges[k];
            if (f == ed) continue;
            if (!clash && t.colorOf[f] == pick) {
              clash = true;
              ++pick;
            }
          }
        }
      }
      cand[ed] = pick;
    }
    nextRem.clear();
    for (uint32_t ed : remaining) {
      if (cand[ed] != kUncolored) {
        t.colorOf[ed] = cand[ed];
        maxColor = std::max(maxColor, (int)cand[ed]);
      } else {
        nextRem.push_back(ed);
      }
    }
    remaining.swap(nextRem);
  }
  t.colorCount = maxColor + 1;
  t.cstart.assign((size_t)t.colorCount + 1, 0);
  for (uint32_t c : t.colorOf) t.cstart[(size_t)c + 1]++;
  for (int c = 0; c < t.colorCount; ++c) t.cstart[(size_t)c + 1] += t.cstart[(size_t)c];
  t.corder.assign((size_t)E, 0);
  {
    std::vector<uint32_t> cur(t.cstart.begin(), t.cstart.end() - 1);
    for (int e = 0; e < E; ++e) t.corder[cur[t.colorOf[(size_t)e]]++] = (uint32_t)e;
  }
  for (int c = 0; c < t.colorCount; ++c) {
    std::sort(t.corder.begin() + t.cstart[(size_t)c], t.corder.begin() + t.cstart[(size_t)c + 1],
              [&](uint32_t x, uint32_t y) {
                const uint32_t mx = (uint32_t)std::min(s.ea[x], s.eb[x]);
                const uint32_t my = (uint32_t)std::min(s.ea[y], s.eb[y]);
                return mx != my ? mx < my : x < y;
              });
  }
  // per-vertex incidence in colour order: a proper colouring gives each vertex at
  // most one edge per colour, so sorting by colour reproduces the scatter sequence.
  t.vedgesColor = t.vedges;
  for (int v = 0; v < n; ++v)
    std::sort(t.vedgesColor.begin() + t.vstart[(size_t)v], t.vedgesColor.begin() + t.vstart[(size_t)v + 1],
              [&](uint32_t x, uint32_t y) { return t.colorOf[x] < t.colorOf[y]; });
  // bend stencil incidence (stencil order) with weights
  t.bstart.assign((size_t)n + 1, 0);
  for (int q = 0; q < s.B; ++q) {
    t.bstart[(size_t)s.sa[q] + 1]++;
    t.bstart[(size_t)s.sb[q] + 1]++;
    t.bstart[(size_t)s.sc[q] + 1]++;
  }
  for (int v = 0; v < n; ++v) t.bstart[(size_t)v + 1] += t.bstart[(size_t)v];
  t.bstencil.assign((size_t)3 * s.B, 0);
  t.bweight.assign((size_t)3 * s.B, 0);
  {
    std::vector<uint32_t> cur(t.bstart.begin(), t.bstart.end() - 1);
    for (int q = 0; q < s.B; ++q) {
      const int vv[3] = {s.sa[q], s.sb[q], s.sc[q]};
      const int8_t w[3] = {1, -2, 1};
      for (int k = 0; k < 3; ++k) {
        const uint32_t slot = cur[(size_t)vv[k]]++;
        t.bstencil[slot] = (uint32_t)q;
        t.bweight[slot] = w[k];
      }
    }
  }
}

// core/assemble/Assembler.cpp:139-278 with DistanceTerm::assembleMatrix
// (core/energy/DistanceTerm.cpp:319-349).  Every 3x3 block is scalar x I3, so one
// scalar component is assembled; duplicate triplets are summed in insertion order
// as Eigen's setFromTriplets does.
void assembleMatrix(const Scene& s, double dt, Matrix& A) {
  struct T {
    int r, c;
    double v;
  };
  const int n = s.n;
  const double invDt2 = 1.0 / (dt * dt);
  std::vector<T> t;
  t.reserve((size_t)n + 4 * (size_t)s.E + 9 * (size_t)s.B + 8);
  auto pinnedV = [&](int v) { return s.pinned[(size_t)v] != 0; };
  for (int v = 0; v < n; ++v) t.push_back({v, v, s.mass[(size_t)v] * invDt2});
  for (int e = 0; e < s.E; ++e) {
    const bool ap = pinnedV(s.ea[e]), bp = pinnedV(s.eb[e]);
    if (ap && bp) continue;
    if (!ap) t.push_back({s.ea[e], s.ea[e], s.kappa[e]});
    if (!bp) t.push_back({s.eb[e], s.eb[e], s.kappa[e]});
    if (!ap && !bp) {
      t.push_back({s.ea[e], s.eb[e], -s.kappa[e]});
      t.push_back({s.eb[e], s.ea[e], -s.kappa[e]});
    }
  }
  const double w[3] = {1.0, -2.0, 1.0};
  for (int q = 0; q < s.B; ++q) {
    if (s.bk[q] == 0.0) continue;
    const int vv[3] = {s.sa[q], s.sb[q], s.sc[q]};
    if (pinnedV(vv[0]) && pinnedV(vv[1]) && pinnedV(vv[2])) continue;
    for (int p = 0; p < 3; ++p) {
      if (pinnedV(vv[p])) continue;
      for (int r = 0; r < 3; ++r) {
        if (pinnedV(vv[r])) continue;
        t.push_back({vv[p], vv[r], s.bk[q] * w[p] * w[r]});
      }
    }
  }
  // pinned rows: diagonal completed to 1 (Assembler.cpp:227-253).  Every edge at a
  // pinned vertex is skipped by the "either end pinned" filter and the bend term of
  // a pinned row is 0, so only the inertia term is "existing".
  for (int v = 0; v < n; ++v) {
    if (!pinnedV(v)) continue;
    const double existing = s.mass[(size_t)v] * invDt2;
    t.push_back({v, v, 1.0 - existing});
  }
  std::stable_sort(t.begin(), t.end(), [](const T& a, const T& b) { return a.c != b.c ? a.c < b.c : a.r < b.r; });
  A.n = n;
  A.colptr.assign((size_t)n + 1, 0);
  A.rows.clear();
  A.vals.clear();
  for (size_t k = 0; k < t.size();) {
    size_t j = k;
    double sum = t[k].v;
    while (j + 1 < t.size() && t[j + 1].r == t[k].r && t[j + 1].c == t[k].c) {
      ++j;
      sum = sum + t[j].v;
    }
    A.rows.push_back(t[k].r);
    A.vals.push_back(sum);
    A.colptr[(size_t)t[k].c + 1]++;
    k = j + 1;
  }
  for (int v = 0; v < n; ++v) A.colptr[(size_t)v + 1] += A.colptr[(size_t)v];
}

// core/assemble/Assembler.cpp:94-137
void bendingRhs(const Scene& s, std::vector<double>& out) {
  out.assign((size_t)3 * s.n, 0.0);
  const double w[3] = {1.0, -2.0, 1.0};
  for (int q = 0; q < s.B; ++q) {
    if (s.bk[q] == 0.0) continue;
    const int vv[3] = {s.sa[q], s.sb[q], s.sc[q]};
    double c0 = 0.0, c1 = 0.0, c2 = 0.0;
    bool any = false;
    for (int k = 0; k < 3; ++k) {
      if (!s.pinned[(size_t)vv[k]]) continue;
      any = true;
      const double* p = &s.pinpos[(size_t)vv[k] * 3];
      c0 += p[0] * w[k];
      c1 += p[1] * w[k];
      c2 += p[2] * w[k];
    }
    if (!any) continue;
    for (int k = 0; k < 3; ++k) {
      if (s.pinned[(size_t)vv[k]]) continue;
      const double f = -s.bk[q] * w[k];
      double* slot = &out[(size_t)vv[k] * 3];
      slot[0] += c0 * f;
      slot[1] += c1 * f;
      slot[2] += c2 * f;
    }
  }
}

// core/assemble/Assembler.cpp:29-65 (damping excluded)
uint64_t numericStamp(const Scene& s, double dt) {
  uint64_t sh = 0xcbf29ce484222325ULL, mh = sh, ph = sh;
  for (int e = 0; e < s.E; ++e) sh = mix(sh, bitsOf(s.kappa[(size_t)e]));
  for (int q = 0; q < s.B; ++q) sh = mix(sh, bitsOf(s.bk[(size_t)q]));
  for (int v = 0; v < s.n; ++v) mh = mix(mh, bitsOf(s.mass[(size_t)v]));
  for (int v = 0; v < s.n; ++v) ph = mix(ph, s.pinned[(size_t)v] ? 1ULL : 0ULL);
  uint64_t h = mix(0xcbf29ce484222325ULL, sh);
  h = mix(h, mh);
  h = mix(h, bitsOf(dt));
  return mix(h, ph);
}

// core/assemble/Assembler.cpp:67-92
uint64_t structureStamp(const Scene& s) {
  uint64_t h = 0xcbf29ce484222325ULL;
  h = mix(h, 1);
  h = mix(h, (uint64_t)s.n);
  h = mix(h, (uint64_t)s.E);
  for (int e = 0; e < s.E; ++e) {
    h = mix(h, (uint64_t)(uint32_t)s.ea[(size_t)e]);
    h = mix(h, (uint64_t)(uint32_t)s.eb[(size_t)e]);
  }
  h = mix(h, (uint64_t)s.B);
  for (int q = 0; q < s.B; ++q) {
    h = mix(h, (uint64_t)(uint32_t)s.sa[(size_t)q]);
    h = mix(h, (uint64_t)(uint32_t)s.sb[(size_t)q]);
    h = mix(h, (uint64_t)(uint32_t)s.sc[(size_t)q]);
  }
  for (int v = 0; v < s.n; ++v) h = mix(h, s.pinned[(size_t)v] ? 1ULL : 0ULL);
  return h;
}

namespace {

// Up-looking LDL^T of the permuted scalar matrix in rank space (Eigen's
// SimplicialLDLT numeric kernel, i.e. T. Davis' LDL, ACM TOMS Alg. 849): etree,
// column counts, then row k of L by a sparse triangular solve over the etree
// reach, visited in the same stack order.  The x/y/z components of the reference's
// 3n factor are this factor, bit for bit, in each component's induced order.
// Provenance: the steps are those of Davis' LDL package (LGPL-2.1+) as Eigen's
// SimplicialLDLT (MPL-2.0) restates them; the code is written anew around this
// library's rank-space matrix.
int ldlRank(const Matrix& A, const std::vector<int>& vertexRank, const std::vector<int>* symbolic,
            Factor& f) {
  const int m = A.n;
  // permuted upper storage: column K holds rows I <= K (ascending), values of the
  // lower triangle of A
  std::vector<int> up((size_t)m + 1, 0);
  for (int vc = 0; vc < m; ++vc)
    for (int p = A.colptr[(size_t)vc]; p < A.colptr[(size_t)vc + 1]; ++p) {
      const int vr = A.rows[(size_t)p];
      if (vr < vc) continue;
      const int K = std::max(vertexRank[(size_t)vr], vertexRank[(size_t)vc]);
      up[(size_t)K + 1]++;
    }
  for (int k = 0; k < m; ++k) up[(size_t)k + 1] += up[(size_t)k];
  std::vector<int> ui((size_t)up[(size_t)m]);
  std::vector<double> ux((size_t)up[(size_t)m]);
  {
    std::vector<int> cur(up.begin(), up.end() - 1);
    for (int vc = 0; vc < m; ++vc)
      for (int p = A.colptr[(size_t)vc]; p < A.colptr[(size_t)vc + 1]; ++p) {
        const int vr = A.rows[(size_t)p];
        if (vr < vc) continue;
        const int a = vertexRank[(size_t)vr], b = vertexRank[(size_t)vc];
        const int K = std::max(a, b), I = std::min(a, b);
        ui[(size_t)cur[(size_t)K]] = I;
        ux[(size_t)cur[(size_t)K]++] = A.vals[(size_t)p];
      }
    for (int k = 0; k < m; ++k) {
      // sort (row, value) pairs of column k by row
      std::vector<std::pair<int, double>> tmp;
      for (int p = up[(size_t)k]; p < up[(size_t)k + 1]; ++p) tmp.push_back({ui[(size_t)p], ux[(size_t)p]});
      std::sort(tmp.begin(), tmp.end(), [](auto& x, auto& y) { return x.first < y.first; });
      for (size_t q = 0; q < tmp.size(); ++q) {
        ui[(size_t)up[(size_t)k] + q] = tmp[q].first;
        ux[(size_t)up[(size_t)k] + q] = tmp[q].second;
      }
    }
  }
  std::vector<int> parent((size_t)m), nzc((size_t)m), tags((size_t)m);
  for (int k = 0; k < m; ++k) {
    parent[(size_t)k] = -1;
    tags[(size_t)k] = k;
    nzc[(size_t)k] = 0;
    for (int p = up[(size_t)k]; p < up[(size_t)k + 1]; ++p) {
      int i = ui[(size_t)p];
      if (i < k)
        for (; tags[(size_t)i] != k; i = parent[(size_t)i]) {
          if (parent[(size_t)i] == -1) parent[(size_t)i] = k;
          nzc[(size_t)i]++;
          tags[(size_t)i] = k;
        }
    }
  }
  (void)symbolic;
  f.cp.assign((size_t)m + 1, 0);
  for (int k = 0; k < m; ++k) f.cp[(size_t)k + 1] = f.cp[(size_t)k] + nzc[(size_t)k];
  const size_t nnz = (size_t)f.cp[(size_t)m];
  f.ci.assign(nnz, 0);
  f.cx.assign(nnz, 0.0);
  std::vector<double> D((size_t)m), y((size_t)m, 0.0);
  std::vector<int> pattern((size_t)m), cnt((size_t)m, 0);
  for (int k = 0; k < m; ++k) {
    y[(size_t)k] = 0.0;
    int top = m;
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
    for (; top < m; ++top) {
      const int i = pattern[(size_t)top];
      const double yi = y[(size_t)i];
      y[(size_t)i] = 0.0;
      const double lki = yi / D[(size_t)i];
      const int p2 = f.cp[(size_t)i] + cnt[(size_t)i];
      int p;
      for (p = f.cp[(size_t)i]; p < p2; ++p) y[(size_t)f.ci[(size_t)p]] -= f.cx[(size_t)p] * yi;
      d -= lki * yi;
      f.ci[(size_t)p] = k;
      f.cx[(size_t)p] = lki;
      ++cnt[(size_t)i];
    }
    D[(size_t)k] = d;
    if (d == 0.0) return -3;
  }
  f.dinv.resize((size_t)m);
  for (int k = 0; k < m; ++k) f.dinv[(size_t)k] = 1.0 / D[(size_t)k];
  f.nnz = (int64_t)nnz;
  // row-major copy (forward sweep gathers row k over ascending columns)
  f.rp.assign((size_t)m + 1, 0);
  for (size_t p = 0; p < nnz; ++p) f.rp[(size_t)f.ci[p] + 1]++;
  for (int k = 0; k < m; ++k) f.rp[(size_t)k + 1] += f.rp[(size_t)k];
  f.rc.assign(nnz, 0);
  f.rx.assign(nnz, 0.0);
  {
    std::vector<int> cur(f.rp.begin(), f.rp.end() - 1);
    for (int j = 0; j < m; ++j)
      for (int p = f.cp[(size_t)j]; p < f.cp[(size_t)j + 1]; ++p) {
        const int i = f.ci[(size_t)p];
        f.rc[(size_t)cur[(size_t)i]] = j;
        f.rx[(size_t)cur[(size_t)i]++] = f.cx[(size_t)p];
      }
  }
  return 0;
}

void buildSweeps(Factor& f) {
  const int m = f.m;
  // forward: row k waits for every column j < k it references
  f.fwd.levelOf.assign((size_t)m, 0);
  int nl = 0;
  for (int k = 0; k < m; ++k) {
    int l = 0;
    for (int p = f.rp[(size_t)k]; p < f.rp[(size_t)k + 1]; ++p)
      l = std::max(l, f.fwd.levelOf[(size_t)f.rc[(size_t)p]] + 1);
    f.fwd.levelOf[(size_t)k] = l;
    nl = std::max(nl, l + 1);
  }
  f.fwd.nlev = m ? nl : 0;
  // backward: column k waits for every row i > k in it
  f.bwd.levelOf.assign((size_t)m, 0);
  nl = 0;
  for (int k = m - 1; k >= 0; --k) {
    int l = 0;
    for (int p = f.cp[(size_t)k]; p < f.cp[(size_t)k + 1]; ++p)
      l = std::max(l, f.bwd.levelOf[(size_t)f.ci[(size_t)p]] + 1);
    f.bwd.levelOf[(size_t)k] = l;
    nl = std::max(nl, l + 1);
  }
  f.bwd.nlev = m ? nl : 0;
  for (Sweep* s : {&f.fwd, &f.bwd}) {
    s->levPtr.assign((size_t)s->nlev + 1, 0);
    for (int k = 0; k < m; ++k) s->levPtr[(size_t)s->levelOf[(size_t)k] + 1]++;
    for (int l = 0; l < s->nlev; ++l) s->levPtr[(size_t)l + 1] += s->levPtr[(size_t)l];
    s->order.assign((size_t)m, 0);
    std::vector<int> cur(s->levPtr.begin(), s->levPtr.end() - 1);
    for (int k = 0; k < m; ++k) s->order[(size_t)cur[(size_t)s->levelOf[(size_t)k]]++] = k;
  }
}

}  // namespace

// core/solver/EigenDirectSolver.cpp:230-297: AMD on the symmetric 3n pattern
// (L = A (x) I3, so the pattern is A's, replicated on dofs 3v+d), pivot order split
// into the three x/y/z components (:106-155), numeric LDL^T of one component and the
// shared compact factor (:157-225), then the sweep level schedules.
int analyzeAndFactor(const Scene& s, const Matrix& A, Factor& f, int ordering) {
  const int n = s.n;
  const int n3 = 3 * n;
  if (ordering == kOrderNd) {
    // the three components of a vertex share its place in the vertex order
    std::vector<int> vo;
    ndOrder(A, s.x, kNdLeaf, vo, f.top);
    f.perm3.resize((size_t)n3);
    for (int k = 0; k < n; ++k)
      for (int d = 0; d < 3; ++d) f.perm3[(size_t)3 * k + d] = 3 * vo[(size_t)k] + d;
  } else {
    f.top[0] = f.top[1] = 0;
    std::vector<int> cp((size_t)n3 + 1, 0), ri;
    ri.reserve((size_t)3 * A.rows.size());
    for (int vc = 0; vc < n; ++vc)
      for (int d = 0; d < 3; ++d) {
        const int J = 3 * vc + d;
        for (int p = A.colptr[(size_t)vc]; p < A.colptr[(size_t)vc + 1]; ++p)
          ri.push_back(3 * A.rows[(size_t)p] + d);
        cp[(size_t)J + 1] = (int)ri.size();
      }
    amdOrder(n3, cp, ri, f.perm3);
  }
  f.m = n;
  for (int c = 0; c < 3; ++c) {
    f.rankVertex[c].clear();
    f.vertexRank[c].assign((size_t)n, -1);
  }
  for (int k = 0; k < n3; ++k) {
    const int dof = f.perm3[(size_t)k];
    const int c = dof % 3;
    f.vertexRank[c][(size_t)(dof / 3)] = (int)f.rankVertex[c].size();
    f.rankVertex[c].push_back(dof / 3);
  }
  for (int c = 0; c < 3; ++c)
    if ((int)f.rankVertex[c].size() != n) return -5;
  f.sharedOrder = f.rankVertex[1] == f.rankVertex[0] && f.rankVertex[2] == f.rankVertex[0];
  int rc = ldlRank(A, f.vertexRank[0], nullptr, f);
  if (rc != 0) return rc;
  if (!f.sharedOrder) {
    // the reference keeps a single compact factor only when the other two
    // components factor to the same bits in rank space (EigenDirectSolver.cpp:192-224)
    for (int c = 1; c < 3; ++c) {
      Factor g;
      g.m = n;
      if (ldlRank(A, f.vertexRank[c], nullptr, g) != 0) return -3;
      if (g.cp != f.cp || g.ci != f.ci || g.cx != f.cx || g.dinv != f.dinv) return -5;
    }
  }
  buildSweeps(f);
  return 0;
}

int refactorNumeric(const Matrix& A, Factor& f) {
  Factor g;
  g.m = f.m;
  const int rc = ldlRank(A, f.vertexRank[0], nullptr, g);
  if (rc != 0) return rc;
  if (g.cp != f.cp || g.ci != f.ci) return -5;
  f.cx.swap(g.cx);
  f.rx.swap(g.rx);
  f.dinv.swap(g.dinv);
  return 0;
}

void denseInverse(const Factor& f, std::vector<double>& out, int threads) {
  const int m = f.m;
  out.assign((size_t)m * m, 0.0);
  std::atomic<int> next{0};
  auto worker = [&]() {
    std::vector<double> y((size_t)m);
    for (;;) {
      const int r = next.fetch_add(1);
      if (r >= m) return;
      std::fill(y.begin(), y.end(), 0.0);
      y[(size_t)r] = 1.0;
      // L y = e_r, column by column from r (the entries above r stay zero)
      for (int k = r; k < m; ++k) {
        const double yk = y[(size_t)k];
        if (yk == 0.0) continue;
        for (int p = f.cp[(size_t)k]; p < f.cp[(size_t)k + 1]; ++p) y[(size_t)f.ci[(size_t)p]] -= f.cx[(size_t)p] * yk;
      }
      for (int k = r; k < m; ++k) y[(size_t)k] *= f.dinv[(size_t)k];
      // L^T x = y, rows from the last
      for (int k = m - 1; k >= 0; --k) {
        double s = y[(size_t)k];
        for (int p = f.cp[(size_t)k]; p < f.cp[(size_t)k + 1]; ++p) s -= f.cx[(size_t)p] * y[(size_t)f.ci[(size_t)p]];
        y[(size_t)k] = s;
      }
      std::copy(y.begin(), y.end(), out.begin() + (size_t)r * m);
    }
  };
  const int nth = std::max(1, std::min(threads, m));
  std::vector<std::thread> pool;
  for (int t = 1; t < nth; ++t) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
}

}  // namespace pdx

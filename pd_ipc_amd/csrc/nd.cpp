// Nested-dissection vertex order for the fast solve mode (opt-in; the default
// order is the reference's AMD, amd.cpp).
//
// The chain-blocked sweeps advance one elimination-tree block level per step,
// and on a workgroup that owns one cloth every step costs a fixed latency
// (loads, two barriers).  AMD keeps fill low but leaves a tall elimination
// tree: a 96 x 94 cloth with shear and bending needs ~52 block steps per sweep.
// Nested dissection orders both halves of a split before the separator that
// splits them, so the tree's height follows the separators (their rows chain
// into dense blocks) and drops to ~24 steps for ~25 % more factor entries.
//
// Splits are geometric: the vertices of a part are sorted along the axis of
// its largest extent (ties by index) and cut at the median; the separator is
// the smaller of the two boundary sets (vertices of one half with a matrix
// neighbour in the other), so it separates whatever the matrix couples
// (edges and the distance-2 couplings of bending stencils alike).  Parts of at
// most `leaf` vertices keep their sorted order.  Any order gives the same
// factorisation up to rounding; this one only changes speed and fill.
#include <algorithm>
#include <numeric>

#include "host.h"

namespace pdx {

namespace {

struct NdState {
  const Matrix* A;
  const double* x;
  int leaf;
  std::vector<int> tag;  // per vertex: the last split stamp that labelled it
  int stamp = 0;
  std::vector<int>* out;

  bool touches(int v, int other) const {
    for (int p = A->colptr[(size_t)v]; p < A->colptr[(size_t)v + 1]; ++p) {
      const int u = A->rows[(size_t)p];
      if (u != v && tag[(size_t)u] == other) return true;
    }
    return false;
  }

  int depth = 0;
  int top[2] = {0, 0};  // vertices of the two halves of the top-level split

  void split(std::vector<int>& vs) {
    if ((int)vs.size() <= leaf) {
      out->insert(out->end(), vs.begin(), vs.end());
      return;
    }
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
      lo[d] = x[3 * (size_t)vs[0] + d];
      hi[d] = lo[d];
    }
    for (int v : vs)
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], x[3 * (size_t)v + d]);
        hi[d] = std::max(hi[d], x[3 * (size_t)v + d]);
      }
    int ax = 0;
    for (int d = 1; d < 3; ++d)
      if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
    std::sort(vs.begin(), vs.end(), [&](int a, int b) {
      const double ca = x[3 * (size_t)a + ax], cb = x[3 * (size_t)b + ax];
      return ca != cb ? ca < cb : a < b;
    });
    const size_t half = vs.size() / 2;
    const int tl = ++stamp, tr = ++stamp;
    for (size_t i = 0; i < vs.size(); ++i) tag[(size_t)vs[i]] = i < half ? tl : tr;
    std::vector<int> sepL, sepR;
    for (size_t i = 0; i < half; ++i)
      if (touches(vs[i], tr)) sepL.push_back(vs[i]);
    for (size_t i = half; i < vs.size(); ++i)
      if (touches(vs[i], tl)) sepR.push_back(vs[i]);
    const bool takeLeft = sepL.size() < sepR.size();
    const std::vector<int>& sep = takeLeft ? sepL : sepR;
    const int ts = ++stamp;
    for (int v : sep) tag[(size_t)v] = ts;
    std::vector<int> L, R;
    L.reserve(half);
    R.reserve(vs.size() - half);
    for (size_t i = 0; i < vs.size(); ++i) {
      const int v = vs[i];
      if (tag[(size_t)v] == ts) continue;
      (i < half ? L : R).push_back(v);
    }
    std::vector<int> s(sep);
    vs.clear();
    vs.shrink_to_fit();
    if (depth == 0) {
      top[0] = (int)L.size();
      top[1] = (int)R.size();
    }
    ++depth;
    split(L);
    split(R);
    --depth;
    out->insert(out->end(), s.begin(), s.end());
  }
};

}  // namespace

void ndOrder(const Matrix& A, const std::vector<double>& x, int leaf, std::vector<int>& order, int* top) {
  const int n = A.n;
  order.clear();
  order.reserve((size_t)n);
  NdState st;
  st.A = &A;
  st.x = x.data();
  st.leaf = std::max(1, leaf);
  st.tag.assign((size_t)n, 0);
  st.out = &order;
  std::vector<int> all((size_t)n);
  std::iota(all.begin(), all.end(), 0);
  st.split(all);
  if (top) {
    top[0] = st.top[0];
    top[1] = st.top[1];
  }
}

}  // namespace pdx

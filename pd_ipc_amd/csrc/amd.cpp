// Approximate minimum degree ordering of a symmetric sparsity pattern.
//
// This is the ordering the reference factorisation uses: EigenDirectSolver::analyze
// (core/solver/EigenDirectSolver.cpp:238-248) runs Eigen 3.4's AMDOrdering<int> on
// the symmetrised 3n pattern.  Eigen 3.4 is not vendored in the reference checkout;
// its AMDOrdering is the published AMD algorithm of Amestoy, Davis and Duff
// (SIAM J. Matrix Anal. Appl. 17(4), 1996; ACM TOMS Alg. 837, 2004) in the concise
// form of T. Davis, "Direct Methods for Sparse Linear Systems" (SIAM 2006), ch. 7,
// with two variations that Eigen makes and that change the result, so they are
// kept here: the diagonal stays in the quotient graph (initial degree = list
// length including the diagonal; a node whose only entry is its diagonal is empty)
// and a node without a structural diagonal is treated as dense.
//
// Provenance: the code follows the structure of CSparse's cs_amd (T. Davis,
// LGPL-2.1+) - its flip / wclear / tdfs helpers, the max(16, 10 sqrt(n)) dense
// threshold and the workspace split - because reproducing the order of Eigen's
// AMDOrdering (Eigen/src/OrderingMethods/Amd.h, MPL-2.0, itself derived from the
// same code) needs its tie-breaks; it is written anew, not copied from either.
//
// Input: column pointers/indices of the FULL symmetric pattern with the diagonal,
// rows sorted within each column.  Output: perm[k] = node eliminated k-th.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "host.h"

namespace pdx {

namespace {

inline int flip(int i) { return -i - 2; }

int wclear(int mark, int lemax, int* w, int n) {
  if (mark < 2 || (mark + lemax < 0)) {
    for (int k = 0; k < n; k++)
      if (w[k] != 0) w[k] = 1;
    mark = 2;
  }
  return mark;
}

int tdfs(int j, int k, int* head, const int* next, int* post, int* stack) {
  int top = 0;
  stack[0] = j;
  while (top >= 0) {
    const int p = stack[top];
    const int i = head[p];
    if (i == -1) {
      top--;
      post[k++] = p;
    } else {
      head[p] = next[i];
      stack[++top] = i;
    }
  }
  return k;
}

}  // namespace

void amdOrder(int n, const std::vector<int>& colptr, const std::vector<int>& rowidx,
              std::vector<int>& permOut) {
  permOut.assign((size_t)n, 0);
  if (n == 0) return;
  int dense = std::max(16, (int)(10 * std::sqrt((double)n)));
  dense = std::min(n - 2, dense);
  const int cnz0 = colptr[(size_t)n];
  int cnz = cnz0;
  const int t = cnz + cnz / 5 + 2 * n;
  std::vector<int> Cp(colptr.begin(), colptr.end());
  std::vector<int> Ci((size_t)std::max(t, 1), 0);
  std::copy(rowidx.begin(), rowidx.begin() + cnz0, Ci.begin());
  std::vector<int> Pv((size_t)n + 1, 0);
  std::vector<int> W(8 * ((size_t)n + 1), 0);
  int* len = W.data();
  int* nv = W.data() + (n + 1);
  int* next = W.data() + 2 * (n + 1);
  int* head = W.data() + 3 * (n + 1);
  int* elen = W.data() + 4 * (n + 1);
  int* degree = W.data() + 5 * (n + 1);
  int* w = W.data() + 6 * (n + 1);
  int* hhead = W.data() + 7 * (n + 1);
  int* last = Pv.data();
  int* cp = Cp.data();
  int* ci = Ci.data();

  int d, dk, dext, lemax = 0, e, elenk, eln, i, j, k, k1, k2, k3, jlast, ln, nzmax, mindeg = 0,
                   nvi, nvj, nvk, mark, wnvi, ok, nel = 0, p, p1, p2, p3, p4, pj, pk, pk1, pk2, pn,
                   q, h;
  for (k = 0; k < n; k++) len[k] = cp[k + 1] - cp[k];
  len[n] = 0;
  nzmax = t;
  for (i = 0; i <= n; i++) {
    head[i] = -1;
    last[i] = -1;
    next[i] = -1;
    hhead[i] = -1;
    nv[i] = 1;
    w[i] = 1;
    elen[i] = 0;
    degree[i] = len[i];
  }
  mark = wclear(0, 0, w, n);
  for (i = 0; i < n; i++) {
    bool hasDiag = false;
    for (p = cp[i]; p < cp[i + 1]; ++p)
      if (ci[p] == i) {
        hasDiag = true;
        break;
      }
    d = degree[i];
    if (d == 1 && hasDiag) {
      elen[i] = -2;
      nel++;
      cp[i] = -1;
      w[i] = 0;
    } else if (d > dense || !hasDiag) {
      nv[i] = 0;
      elen[i] = -1;
      nel++;
      cp[i] = flip(n);
      nv[n]++;
    } else {
      if (head[d] != -1) last[head[d]] = i;
      next[i] = head[d];
      head[d] = i;
    }
  }
  elen[n] = -2;
  cp[n] = -1;
  w[n] = 0;

  while (nel < n) {
    // select a node of minimum approximate degree
    for (k = -1; mindeg < n && (k = head[mindeg]) == -1; mindeg++) {
    }
    if (next[k] != -1) last[next[k]] = -1;
    head[mindeg] = next[k];
    elenk = elen[k];
    nvk = nv[k];
    nel += nvk;
    // garbage collection
    if (elenk > 0 && cnz + mindeg >= nzmax) {
      for (j = 0; j < n; j++) {
        if ((p = cp[j]) >= 0) {
          cp[j] = ci[p];
          ci[p] = flip(j);
        }
      }
      for (q = 0, p = 0; p < cnz;) {
        if ((j = flip(ci[p++])) >= 0) {
          ci[q] = cp[j];
          cp[j] = q++;
          for (k3 = 0; k3 < len[j] - 1; k3++) ci[q++] = ci[p++];
        }
      }
      cnz = q;
    }
    // construct the new element
    dk = 0;
    nv[k] = -nvk;
    p = cp[k];
    pk1 = (elenk == 0) ? p : cnz;
    pk2 = pk1;
    for (k1 = 1; k1 <= elenk + 1; k1++) {
      if (k1 > elenk) {
        e = k;
        pj = p;
        ln = len[k] - elenk;
      } else {
        e = ci[p++];
        pj = cp[e];
        ln = len[e];
      }
      for (k2 = 1; k2 <= ln; k2++) {
        i = ci[pj++];
        if ((nvi = nv[i]) <= 0) continue;
        dk += nvi;
        nv[i] = -nvi;
        ci[pk2++] = i;
        if (next[i] != -1) last[next[i]] = last[i];
        if (last[i] != -1)
          next[last[i]] = next[i];
        else
          head[degree[i]] = next[i];
      }
      if (e != k) {
        cp[e] = flip(k);
        w[e] = 0;
      }
    }
    if (elenk != 0) cnz = pk2;
    degree[k] = dk;
    cp[k] = pk1;
    len[k] = pk2 - pk1;
    elen[k] = -2;
    // set differences |Le \ Lk|
    mark = wclear(mark, lemax, w, n);
    for (pk = pk1; pk < pk2; pk++) {
      i = ci[pk];
      if ((eln = elen[i]) <= 0) continue;
      nvi = -nv[i];
      wnvi = mark - nvi;
      for (p = cp[i]; p <= cp[i] + eln - 1; p++) {
        e = ci[p];
        if (w[e] >= mark)
          w[e] -= nvi;
        else if (w[e] != 0)
          w[e] = degree[e] + wnvi;
      }
    }
    // degree update
    for (pk = pk1; pk < pk2; pk++) {
      i = ci[pk];
      p1 = cp[i];
      p2 = p1 + elen[i] - 1;
      pn = p1;
      for (h = 0, d = 0, p = p1; p <= p2; p++) {
        e = ci[p];
        if (w[e] != 0) {
          dext = w[e] - mark;
          if (dext > 0) {
            d += dext;
            ci[pn++] = e;
            h += e;
          } else {
            cp[e] = flip(k);
            w[e] = 0;
          }
        }
      }
      elen[i] = pn - p1 + 1;
      p3 = pn;
      p4 = p1 + len[i];
      for (p = p2 + 1; p < p4; p++) {
        j = ci[p];
        if ((nvj = nv[j]) <= 0) continue;
        d += nvj;
        ci[pn++] = j;
        h += j;
      }
      if (d == 0) {
        cp[i] = flip(k);
        nvi = -nv[i];
        dk -= nvi;
        nvk += nvi;
        nel += nvi;
        nv[i] = 0;
        elen[i] = -1;
      } else {
        degree[i] = std::min(degree[i], d);
        ci[pn] = ci[p3];
        ci[p3] = ci[p1];
        ci[p1] = k;
        len[i] = pn - p1 + 1;
        h %= n;
        next[i] = hhead[h];
        hhead[h] = i;
        last[i] = h;
      }
    }
    degree[k] = dk;
    lemax = std::max(lemax, dk);
    mark = wclear(mark + lemax, lemax, w, n);
    // supernode detection
    for (pk = pk1; pk < pk2; pk++) {
      i = ci[pk];
      if (nv[i] >= 0) continue;
      h = last[i];
      i = hhead[h];
      hhead[h] = -1;
      for (; i != -1 && next[i] != -1; i = next[i], mark++) {
        ln = len[i];
        eln = elen[i];
        for (p = cp[i] + 1; p <= cp[i] + ln - 1; p++) w[ci[p]] = mark;
        jlast = i;
        for (j = next[i]; j != -1;) {
          ok = (len[j] == ln) && (elen[j] == eln);
          for (p = cp[j] + 1; ok && p <= cp[j] + ln - 1; p++)
            if (w[ci[p]] != mark) ok = 0;
          if (ok) {
            cp[j] = flip(i);
            nv[i] += nv[j];
            nv[j] = 0;
            elen[j] = -1;
            j = next[j];
            next[jlast] = j;
          } else {
            jlast = j;
            j = next[j];
          }
        }
      }
    }
    // finalize the new element
    for (p = pk1, pk = pk1; pk < pk2; pk++) {
      i = ci[pk];
      if ((nvi = -nv[i]) <= 0) continue;
      nv[i] = nvi;
      d = degree[i] + dk - nvi;
      d = std::min(d, n - nel - nvi);
      if (head[d] != -1) last[head[d]] = i;
      next[i] = head[d];
      last[i] = -1;
      head[d] = i;
      mindeg = std::min(mindeg, d);
      degree[i] = d;
      ci[p++] = i;
    }
    nv[k] = nvk;
    if ((len[k] = p - pk1) == 0) {
      cp[k] = -1;
      w[k] = 0;
    }
    if (elenk != 0) cnz = p;
  }
  // postorder the assembly tree
  for (i = 0; i < n; i++) cp[i] = flip(cp[i]);
  for (j = 0; j <= n; j++) head[j] = -1;
  for (j = n; j >= 0; j--) {
    if (nv[j] > 0) continue;
    next[j] = head[cp[j]];
    head[cp[j]] = j;
  }
  for (e = n; e >= 0; e--) {
    if (nv[e] <= 0) continue;
    if (cp[e] != -1) {
      next[e] = head[cp[e]];
      head[cp[e]] = e;
    }
  }
  std::vector<int> post((size_t)n + 1, 0);
  for (k = 0, i = 0; i <= n; i++) {
    if (cp[i] == -1) k = tdfs(i, k, head, next, post.data(), w);
  }
  for (int r = 0; r < n; ++r) permOut[(size_t)r] = post[(size_t)r];
}

}  // namespace pdx

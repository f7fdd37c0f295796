// C ABI of the substep engine: contexts, instances, the factor-once stamp logic
// (core/sim/Integrator.cpp:341-372), packing of the sweep schedules into the
// device arena, and kernel launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pdx.h"
#include "devtypes.h"
#include "host.h"

namespace pdx {
hipError_t launchSubstep(const InstDev* d, const int* order, int ninst, int nsub, int block,
                         size_t lds, bool fp32, bool blocked, bool idx16, bool wl, hipStream_t s, int subBase,
                         int flags);
hipError_t launchAdjoint(const InstDev* d, const int* order, int ninst, int nsub, int subBase, int block,
                         size_t lds, bool fp32, bool blocked, bool idx16, bool wl, hipStream_t s);
hipError_t launchSolve(const InstDev* d, int inst, int inst2, const double* b, double* x, int block,
                       size_t lds, bool fp32, bool blocked, bool idx16, bool wl, hipStream_t s);
hipError_t setLdsLimit(size_t bytes);
hipError_t launchGrid(const InstDev* d, int inst, int nsub, double* gslots, int waveRuns, int blocks,
                      int block, size_t lds, bool fp32, bool idx16, hipStream_t s);
hipError_t launchGridSolve(const InstDev* d, int inst, const double* b, double* x, int waveRuns, int blocks,
                           int block, size_t lds, bool fp32, bool idx16, hipStream_t s);
int gridBlocksMax(int block, size_t lds, bool fp32, bool idx16);
size_t noallocDofs();
hipError_t noallocScratch(int device, double** ptr);
hipError_t launchDense(const InstDev* d, int inst, int nsub, double* gslots, int blocks, int block, size_t lds,
                       bool fp32, hipStream_t s);
int denseBlocksMax(int block, size_t lds, bool fp32);
hipError_t launchDenseSolve(const InstDev* d, int inst, const double* b, double* x, int n, int m, bool fp32,
                            hipStream_t s);
}  // namespace pdx

using namespace pdx;

namespace {

constexpr size_t kLdsLimit = 160 * 1024;
constexpr size_t kRedBytes = 32 * sizeof(double);
// dense solve mode: largest component size (the FP64 inverse is 8 m^2 bytes, 8.6 GB here)
constexpr int kDenseMaxM = 32768;
// row stride of the device inverse: m rounded up to a multiple of 4
int denseLd(int m) { return (m + 3) & ~3; }
// launch flags of the differentiable rollout (the substep kernel's kRoll*)
constexpr int kRollTraj = 1, kRollTape = 2, kRollFixed = 4;

struct HostSweep {
  std::vector<int> levChunk, chunkRows, chunkRowBase, chunkEntBase, slotRank, slotLen, col;
  std::vector<double> val;
  int nlev = 0;
};

// Rows taken by a whole wavefront (FP64 schedules): every row of at least
// kLongRowAny entries, and in a level of at most kLongLevel rows (one per
// wavefront) every row of at least kLongRow entries.
constexpr int kLongLevel = 8, kLongRow = 16, kLongRowAny = 64;

// Lays one sweep out level by level, <=64-row chunks sorted by row length
// (descending), entries jagged-diagonal inside a chunk.  With `longRows`, the
// long rows of narrow levels become one-row chunks marked by a row count of -1
// (their entries in row order), which a whole wavefront takes.
void packSweep(const Sweep& sw, bool fwd, const Factor& f, bool longRows, HostSweep& out) {
  out = HostSweep{};
  out.nlev = sw.nlev;
  out.levChunk.push_back(0);
  std::vector<int> rows;
  for (int l = 0; l < sw.nlev; ++l) {
    rows.assign(sw.order.begin() + sw.levPtr[(size_t)l], sw.order.begin() + sw.levPtr[(size_t)l + 1]);
    auto lenOf = [&](int k) {
      return fwd ? f.rp[(size_t)k + 1] - f.rp[(size_t)k] : f.cp[(size_t)k + 1] - f.cp[(size_t)k];
    };
    std::stable_sort(rows.begin(), rows.end(), [&](int a, int b) { return lenOf(a) > lenOf(b); });
    size_t first = 0;
    const int longFrom = rows.size() <= (size_t)kLongLevel ? kLongRow : kLongRowAny;
    if (longRows)
      while (first < rows.size() && lenOf(rows[first]) >= longFrom) {
        const int k = rows[first++];
        out.chunkRows.push_back(-1);
        out.chunkRowBase.push_back((int)out.slotRank.size());
        out.chunkEntBase.push_back((int)out.col.size());
        out.slotRank.push_back(k);
        out.slotLen.push_back(lenOf(k));
        for (int j = 0; j < lenOf(k); ++j) {
          const size_t p = fwd ? (size_t)f.rp[(size_t)k] + j : (size_t)f.cp[(size_t)k] + j;
          out.col.push_back(fwd ? f.rc[p] : f.ci[p]);
          out.val.push_back(fwd ? f.rx[p] : f.cx[p]);
        }
      }
    for (size_t c0 = first; c0 < rows.size(); c0 += 64) {
      const size_t c1 = std::min(rows.size(), c0 + 64);
      out.chunkRows.push_back((int)(c1 - c0));
      out.chunkRowBase.push_back((int)out.slotRank.size());
      out.chunkEntBase.push_back((int)out.col.size());
      int maxlen = 0;
      for (size_t t = c0; t < c1; ++t) {
        out.slotRank.push_back(rows[t]);
        out.slotLen.push_back(lenOf(rows[t]));
        maxlen = std::max(maxlen, lenOf(rows[t]));
      }
      for (int j = 0; j < maxlen; ++j)
        for (size_t t = c0; t < c1; ++t) {
          const int k = rows[t];
          if (j >= lenOf(k)) continue;
          if (fwd) {
            out.col.push_back(f.rc[(size_t)f.rp[(size_t)k] + j]);
            out.val.push_back(f.rx[(size_t)f.rp[(size_t)k] + j]);
          } else {
            out.col.push_back(f.ci[(size_t)f.cp[(size_t)k] + j]);
            out.val.push_back(f.cx[(size_t)f.cp[(size_t)k] + j]);
          }
        }
    }
    out.levChunk.push_back((int)out.chunkRows.size());
  }
}

constexpr int kRun = 16;  // factor entries per streamed run (matches the kernel)

struct HostBlockSweep {
  int nstep = 0;
  int maxRuns = 0;        // most runs in one step (partial-sum slots in LDS)
  int maxTaskRuns = 0;    // most runs of one task (grid form: one wavefront's LDS slice)
  int maxDense = 0;       // most dense-block values in one step
  std::vector<int> stepTask, stepRun, stepDense;  // nstep+1 each
  std::vector<int> taskInfo, taskSlot, taskDense; // per task
  std::vector<int> slotRank, slotRunBeg, slotRunEnd;
  std::vector<int> laneMeta;                      // 64 lanes x {rank, run begin, run end, info} per task
  std::vector<double> laneDinv;                   // 64 lanes x 1/D of the lane's rank per task
  std::vector<int> laneDense;                     // 64 lanes x offset of the lane's inverse block
  std::vector<int> runCol;                        // kRun per run
  std::vector<double> runVal, dense;
  // level form (grid): per-task runs, phase B block products, the long chains'
  // inverse panels and the R range (BlockSweepDev)
  bool levels = false;
  std::vector<int> taskRun, stepB, bTask;
  std::vector<double> xpanel;
  int rTotal = 0;
  std::vector<int> stepE, eTask, ucol;
  std::vector<double> epanel;
};

// Block schedule of one sweep (fast mode).  Ranks are cut into contiguous
// blocks: a block continues while rank k-1 has parent k in the elimination tree
// and k has no other child (a chain), up to `cap` rows.  Block levels follow the
// dependencies between blocks; every level is one step.  Within a step, the
// external entries of all its rows (forward: columns before the row's block,
// backward: rows after it) are cut into runs of kRun entries, padded, laid out
// contiguously, so that the whole workgroup streams a step with one wide load
// per run; each run yields one partial sum.  The tasks of a step (chain blocks,
// one wavefront each, and batches of up to 64 single rows) then combine their
// rows' partials in run order and finish the chains.
//
// [lo, hi) restricts the schedule to those ranks (the rest is solved in other
// phases, before this one where it is an input); blocks never cross a rank in
// `cuts`, so that every row keeps the same external entries, runs and partial
// sums whichever phase split is packed.
void packBlocked(const Factor& f, bool fwd, int cap, int runCap, int stepRunCap, int denseCap,
                 HostBlockSweep& out, int lo = 0, int hi = -1, const std::vector<int>& cuts = {}) {
  out = HostBlockSweep{};
  const int m = f.m;
  if (hi < 0) hi = m;
  std::vector<int> parent((size_t)m, -1), nchild((size_t)m, 0);
  for (int j = 0; j < m; ++j)
    if (f.cp[(size_t)j + 1] > f.cp[(size_t)j]) {
      parent[(size_t)j] = f.ci[(size_t)f.cp[(size_t)j]];
      nchild[(size_t)parent[(size_t)j]]++;
    }
  // runs a row can stream at most (its whole row in this sweep's direction)
  auto runsBound = [&](int k) {
    const int len = fwd ? f.rp[(size_t)k + 1] - f.rp[(size_t)k] : f.cp[(size_t)k + 1] - f.cp[(size_t)k];
    return (len + kRun - 1) / kRun;
  };
  if (hi <= lo) {
    out.stepTask.push_back(0);
    out.stepRun.push_back(0);
    out.stepDense.push_back(0);
    return;
  }
  std::vector<char> cut((size_t)m + 1, 0);
  for (int k : cuts)
    if (k > 0 && k < m) cut[(size_t)k] = 1;
  std::vector<int> bs{lo};
  int blockRuns = runsBound(lo);
  for (int k = lo + 1; k < hi; ++k) {
    const bool chain = parent[(size_t)k - 1] == k && nchild[(size_t)k] == 1 && !cut[(size_t)k];
    const int rk = runsBound(k);
    if (!chain || k - bs.back() >= cap || blockRuns + rk > runCap) {
      bs.push_back(k);
      blockRuns = 0;
    }
    blockRuns += rk;
  }
  bs.push_back(hi);
  const int nb = (int)bs.size() - 1;
  std::vector<int> blk((size_t)m, -1);  // -1: rank of another phase (already final)
  for (int b = 0; b < nb; ++b)
    for (int k = bs[(size_t)b]; k < bs[(size_t)b + 1]; ++k) blk[(size_t)k] = b;
  auto extRange = [&](int k, int b, int& p0, int& p1) {
    if (fwd) {
      p0 = f.rp[(size_t)k];
      p1 = p0;
      while (p1 < f.rp[(size_t)k + 1] && f.rc[(size_t)p1] < bs[(size_t)b]) ++p1;
    } else {
      p1 = f.cp[(size_t)k + 1];
      p0 = p1;
      while (p0 > f.cp[(size_t)k] && f.ci[(size_t)p0 - 1] >= bs[(size_t)b + 1]) --p0;
    }
  };
  std::vector<int> blev((size_t)nb, 0);
  int nstep = 0;
  for (int q = 0; q < nb; ++q) {
    const int b = fwd ? q : nb - 1 - q;
    int lv = 0;
    for (int k = bs[(size_t)b]; k < bs[(size_t)b + 1]; ++k) {
      int p0, p1;
      extRange(k, b, p0, p1);
      for (int p = p0; p < p1; ++p) {
        const int o = fwd ? f.rc[(size_t)p] : f.ci[(size_t)p];
        if (blk[(size_t)o] >= 0) lv = std::max(lv, blev[(size_t)blk[(size_t)o]] + 1);
      }
    }
    blev[(size_t)b] = lv;
    nstep = std::max(nstep, lv + 1);
  }
  out.nstep = nb ? nstep : 0;
  std::vector<std::vector<int>> byLevel((size_t)out.nstep);
  for (int b = 0; b < nb; ++b) byLevel[(size_t)blev[(size_t)b]].push_back(b);
  auto extLen = [&](int k, int b) {
    int p0, p1;
    extRange(k, b, p0, p1);
    return p1 - p0;
  };
  // A task is one wavefront: up to 64 lanes holding whole blocks (elimination
  // tree chains of <= cap rows, or single rows) in contiguous lane segments.
  struct Item {
    int block, len, runs;
  };
  out.stepTask.push_back(0);
  out.stepRun.push_back(0);
  out.stepDense.push_back(0);
  int steps = 0;
  for (int l = 0; l < out.nstep; ++l) {
    std::vector<Item> items;
    for (int b : byLevel[(size_t)l]) {
      Item it{b, bs[(size_t)b + 1] - bs[(size_t)b], 0};
      for (int k = bs[(size_t)b]; k < bs[(size_t)b + 1]; ++k) it.runs += (extLen(k, b) + kRun - 1) / kRun;
      items.push_back(it);
    }
    // longest blocks first, then most runs: tasks fill with similar segments
    std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) {
      return a.len != b.len ? a.len > b.len : a.runs > b.runs;
    });
    std::vector<std::vector<Item>> tasks;
    {
      std::vector<Item> cur;
      int lanes = 0, truns = 0;
      for (const Item& it : items) {
        if (!cur.empty() && (lanes + it.len > 64 || truns + it.runs > runCap)) {
          tasks.push_back(cur);
          cur.clear();
          lanes = truns = 0;
        }
        cur.push_back(it);
        lanes += it.len;
        truns += it.runs;
      }
      if (!cur.empty()) tasks.push_back(cur);
    }
    int runBase = out.stepRun.back();
    int denseBase = out.stepDense.back();
    int runs = 0;
    // Closes the current step: its dense block padded to a multiple of 4 values
    // (16-byte copies).  A level whose runs or dense blocks exceed the caps is
    // cut into several steps (its tasks are independent of each other).
    auto closeStep = [&]() {
      while ((out.dense.size() - (size_t)denseBase) % 4) out.dense.push_back(0.0);
      out.maxRuns = std::max(out.maxRuns, runs);
      out.maxDense = std::max(out.maxDense, (int)out.dense.size() - denseBase);
      out.stepTask.push_back((int)out.taskInfo.size());
      out.stepRun.push_back(runBase + runs);
      out.stepDense.push_back((int)out.dense.size());
      ++steps;
      runBase = out.stepRun.back();
      denseBase = out.stepDense.back();
      runs = 0;
    };
    for (const auto& t : tasks) {
      int used = 0, tr = 0, td = 0, maxLen = 1;
      for (const Item& it : t) {
        used += it.len;
        tr += it.runs;
        if (it.len > 1) td += it.len * it.len;
        maxLen = std::max(maxLen, it.len);
      }
      out.maxTaskRuns = std::max(out.maxTaskRuns, tr);
      const bool nonEmpty = (int)out.taskInfo.size() > out.stepTask.back();
      if (nonEmpty && (runs + tr > stepRunCap || (int)(out.dense.size() - (size_t)denseBase) + td > denseCap))
        closeStep();
      out.taskInfo.push_back(used | (maxLen > 1 ? (1 << 16) : 0));
      out.taskSlot.push_back((int)out.slotRank.size());
      out.taskDense.push_back((int)out.dense.size() - denseBase);
      // per lane: {rank or -1, run begin, run end, segment length | lane offset in
      // the segment << 7 | lanes used << 14 | longest segment << 21}, 1/D, and the
      // offset of the segment's inverse block in the step's dense range
      const size_t meta0 = out.laneMeta.size();
      const int common = (used << 14) | (maxLen << 21);
      for (int ln = 0; ln < 64; ++ln) {
        out.laneMeta.push_back(-1);
        out.laneMeta.push_back(0);
        out.laneMeta.push_back(0);
        out.laneMeta.push_back(1 | common);
        out.laneDinv.push_back(0.0);
        out.laneDense.push_back(0);
      }
      int lane = 0;
      for (const Item& it : t) {
        const int b = it.block, nr = it.len, s0 = bs[(size_t)b];
        const int doff = (int)out.dense.size() - denseBase;
        for (int r = 0; r < nr; ++r, ++lane) {
          const int k = s0 + r;
          int p0, p1;
          extRange(k, b, p0, p1);
          out.slotRank.push_back(k);
          out.slotRunBeg.push_back(runs);
          for (int p = p0; p < p1; p += kRun) {
            for (int q = 0; q < kRun; ++q) {
              if (p + q < p1) {
                out.runCol.push_back(fwd ? f.rc[(size_t)(p + q)] : f.ci[(size_t)(p + q)]);
                out.runVal.push_back(fwd ? f.rx[(size_t)(p + q)] : f.cx[(size_t)(p + q)]);
              } else {
                out.runCol.push_back(0);
                out.runVal.push_back(0.0);
              }
            }
            ++runs;
          }
          out.slotRunEnd.push_back(runs);
          const size_t m = meta0 + 4 * (size_t)lane;
          out.laneMeta[m + 0] = k;
          out.laneMeta[m + 1] = out.slotRunBeg.back();
          out.laneMeta[m + 2] = runs;
          out.laneMeta[m + 3] = nr | (r << 7) | common;
          out.laneDinv[meta0 / 4 + (size_t)lane] = f.dinv[(size_t)k];
          out.laneDense[meta0 / 4 + (size_t)lane] = nr > 1 ? doff : 0;
        }
        if (nr > 1) {
          // The chain's own unit lower triangle B (ranks s0 .. s0+nr-1) is stored
          // inverted, X = B^-1 (FP64 on the host, unit diagonal included), so the
          // segment finishes with one product per lane instead of nr dependent
          // substitution rounds:  forward y = X r,  backward y = X^T r.
          // Layout [s * nr + t] (lane offset t, term s):
          //   forward  X(t, s)    (terms s <= t)
          //   backward X(s, t)    (terms s >= t)
          std::vector<double> blkL((size_t)nr * nr, 0.0), X((size_t)nr * nr, 0.0);
          for (int r = 1; r < nr; ++r) {
            const int k = s0 + r;
            for (int p = f.rp[(size_t)k]; p < f.rp[(size_t)k + 1]; ++p) {
              const int c = f.rc[(size_t)p] - s0;
              if (c >= 0 && c < r) blkL[(size_t)r * nr + c] = f.rx[(size_t)p];
            }
          }
          for (int c = 0; c < nr; ++c) {
            X[(size_t)c * nr + c] = 1.0;
            for (int r = c + 1; r < nr; ++r) {
              double acc = 0.0;
              for (int q = c; q < r; ++q) acc += blkL[(size_t)r * nr + q] * X[(size_t)q * nr + c];
              X[(size_t)r * nr + c] = -acc;
            }
          }
          for (int sc = 0; sc < nr; ++sc)
            for (int tt = 0; tt < nr; ++tt)
              out.dense.push_back(fwd ? X[(size_t)tt * nr + sc] : X[(size_t)sc * nr + tt]);
        }
      }
    }
    closeStep();
  }
  out.nstep = steps;
}

// Inverse X = B^-1 of the unit lower triangle B of ranks [s0, s0 + nr) (the
// factor's own entries among them), FP64, row-major nr x nr, unit diagonal
// included: row t of X is e_t minus the factor's entries of row t times the rows
// of X before it (B X = I, taken row by row).
void chainInverse(const Factor& f, int s0, int nr, std::vector<double>& X) {
  X.assign((size_t)nr * nr, 0.0);
  for (int t = 0; t < nr; ++t) {
    double* xt = X.data() + (size_t)t * nr;
    xt[t] = 1.0;
    const int k = s0 + t;
    for (int p = f.rp[(size_t)k]; p < f.rp[(size_t)k + 1]; ++p) {
      const int q = f.rc[(size_t)p] - s0;
      if (q < 0 || q >= t) continue;
      const double l = f.rx[(size_t)p];
      const double* xq = X.data() + (size_t)q * nr;
      for (int c = 0; c <= q; ++c) xt[c] -= l * xq[c];
    }
  }
}

// Level schedules of both sweeps for the grid form.  Blocks are whole
// elimination-tree chains (a nested-dissection separator is one block), so a
// sweep takes one step per level of the block tree instead of one per piece of
// a chain.  A step has two phases:
//  * phase A: tasks of up to 64 lanes, each task streaming its own runs (at most
//    waveRuns, one wavefront's LDS slice) into partial sums.  A block of at most
//    `cap` rows whose runs fit one task is finished inside the task, its rows in
//    consecutive lanes, with its inverse rows in registers (as packBlocked's
//    chain blocks).  A longer block (a long chain) gives each row a lane that
//    subtracts its external partials and writes the result r to R.
//  * phase B: each long chain's inverse applied to its r, y = X r forward and
//    y = X^T r backward, 8 rows per task, the rows of X stored as panels: forward
//    panel p holds rows 8p.. of X over columns [0, 8p + 8), backward panel p rows
//    8p.. of X^T over columns [8p, the chain's end), both padded with zeros to a
//    multiple of 4 columns.
constexpr int kWaveLanes = 64;     // lanes of a wavefront
constexpr long long kLevelWaves = 1024;  // wavefronts a level step's runs are spread over
// a long chain takes dense phase A tasks when its external entries fill at least
// this share of rows x union columns (backward sweeps of nested-dissection
// separators: about 1.0; forward: 0.2-0.3)
constexpr double kDenseExternal = 0.6;

void packLevels(const Factor& f, int cap, int waveRuns, HostBlockSweep& outF, HostBlockSweep& outB) {
  const int m = f.m;
  std::vector<int> parent((size_t)m, -1), nchild((size_t)m, 0);
  for (int j = 0; j < m; ++j)
    if (f.cp[(size_t)j + 1] > f.cp[(size_t)j]) {
      parent[(size_t)j] = f.ci[(size_t)f.cp[(size_t)j]];
      nchild[(size_t)parent[(size_t)j]]++;
    }
  std::vector<int> bs{0};
  for (int k = 1; k < m; ++k)
    if (!(parent[(size_t)k - 1] == k && nchild[(size_t)k] == 1)) bs.push_back(k);
  bs.push_back(m);
  if (m == 0) bs.assign(1, 0);
  const int nb = (int)bs.size() - 1;
  std::vector<int> blk((size_t)m, 0);
  for (int b = 0; b < nb; ++b)
    for (int k = bs[(size_t)b]; k < bs[(size_t)b + 1]; ++k) blk[(size_t)k] = b;
  std::vector<std::vector<double>> Xc((size_t)nb);  // long chains' inverses, formed once for both sweeps
  auto round4 = [](int v) { return (v + 3) & ~3; };
  for (int dir = 0; dir < 2; ++dir) {
    const bool fwd = dir == 0;
    HostBlockSweep& out = fwd ? outF : outB;
    out = HostBlockSweep{};
    out.levels = true;
    out.maxRuns = 1;
    // external entries of row k of block b: forward its columns before the block,
    // backward the rows after the block in its column
    auto extRange = [&](int k, int b, int& p0, int& p1) {
      if (fwd) {
        p0 = f.rp[(size_t)k];
        p1 = p0;
        while (p1 < f.rp[(size_t)k + 1] && f.rc[(size_t)p1] < bs[(size_t)b]) ++p1;
      } else {
        p1 = f.cp[(size_t)k + 1];
        p0 = p1;
        while (p0 > f.cp[(size_t)k] && f.ci[(size_t)p0 - 1] >= bs[(size_t)b + 1]) --p0;
      }
    };
    auto rowRuns = [&](int k, int b) {
      int p0, p1;
      extRange(k, b, p0, p1);
      return (p1 - p0 + kRun - 1) / kRun;
    };
    std::vector<int> blev((size_t)nb, 0);
    int nlev = 0;
    for (int q = 0; q < nb; ++q) {
      const int b = fwd ? q : nb - 1 - q;
      int lv = 0;
      for (int k = bs[(size_t)b]; k < bs[(size_t)b + 1]; ++k) {
        int p0, p1;
        extRange(k, b, p0, p1);
        for (int p = p0; p < p1; ++p) {
          const int o = fwd ? f.rc[(size_t)p] : f.ci[(size_t)p];
          lv = std::max(lv, blev[(size_t)blk[(size_t)o]] + 1);
        }
      }
      blev[(size_t)b] = lv;
      nlev = std::max(nlev, lv + 1);
    }
    if (nb == 0) nlev = 0;
    std::vector<std::vector<int>> byLevel((size_t)nlev);
    for (int b = 0; b < nb; ++b) byLevel[(size_t)blev[(size_t)b]].push_back(b);
    std::vector<int> rpos((size_t)nb, -1);
    std::vector<std::vector<int>> unions((size_t)nb);
    out.stepTask.push_back(0);
    out.stepRun.push_back(0);
    out.stepDense.push_back(0);
    out.stepB.push_back(0);
    out.stepE.push_back(0);
    struct Item {
      int block, len, runs, row;  // row >= 0: one row of a long chain
    };
    for (int l = 0; l < nlev; ++l) {
      std::vector<Item> items;
      std::vector<int> longChains, denseChains;
      for (int b : byLevel[(size_t)l]) {
        const int s0 = bs[(size_t)b], len = bs[(size_t)b + 1] - s0;
        int runs = 0;
        for (int k = s0; k < s0 + len; ++k) runs += rowRuns(k, b);
        if (len <= cap && runs <= waveRuns) {
          items.push_back(Item{b, len, runs, -1});
        } else {
          longChains.push_back(b);
          rpos[(size_t)b] = out.rTotal;
          out.rTotal += round4(len);
          // the union of the rows' external columns; dense enough, the chain's
          // phase A is dense tasks over it rather than per-row runs
          std::vector<int> U;
          long long nnz = 0;
          for (int k = s0; k < s0 + len; ++k) {
            int p0, p1;
            extRange(k, b, p0, p1);
            nnz += p1 - p0;
            for (int p = p0; p < p1; ++p) U.push_back(fwd ? f.rc[(size_t)p] : f.ci[(size_t)p]);
          }
          std::sort(U.begin(), U.end());
          U.erase(std::unique(U.begin(), U.end()), U.end());
          if (U.size() >= (size_t)kWaveLanes && (double)nnz >= kDenseExternal * (double)len * (double)U.size()) {
            denseChains.push_back(b);
            unions[(size_t)b] = std::move(U);
          } else {
            for (int k = s0; k < s0 + len; ++k) items.push_back(Item{b, 1, rowRuns(k, b), k});
          }
        }
      }
      std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) {
        return a.len != b.len ? a.len > b.len : a.runs > b.runs;
      });
      // a task's run budget spreads the step over about kLevelWaves wavefronts (at
      // least one run per lane, at most the LDS slice)
      long long stepRuns = 0;
      for (const Item& it : items) stepRuns += it.runs;
      const int budget = (int)std::min<long long>(waveRuns, std::max<long long>(kWaveLanes, (stepRuns + kLevelWaves - 1) / kLevelWaves));
      std::vector<std::vector<Item>> tasks;
      {
        std::vector<Item> cur;
        int lanes = 0, truns = 0;
        for (const Item& it : items) {
          if (!cur.empty() && (lanes + it.len > 64 || truns + it.runs > budget)) {
            tasks.push_back(cur);
            cur.clear();
            lanes = truns = 0;
          }
          cur.push_back(it);
          lanes += it.len;
          truns += it.runs;
        }
        if (!cur.empty()) tasks.push_back(cur);
      }
      const int denseBase = out.stepDense.back();
      for (const auto& t : tasks) {
        int used = 0, tr = 0, maxLen = 1;
        for (const Item& it : t) {
          used += it.len;
          tr += it.runs;
          maxLen = std::max(maxLen, it.len);
        }
        out.maxTaskRuns = std::max(out.maxTaskRuns, tr);
        const int runBase = (int)(out.runCol.size() / kRun);
        out.taskRun.push_back(runBase);
        out.taskInfo.push_back(used | (maxLen > 1 ? (1 << 16) : 0));
        out.taskSlot.push_back((int)out.slotRank.size());
        out.taskDense.push_back((int)out.dense.size() - denseBase);
        const size_t meta0 = out.laneMeta.size();
        const int common = (used << 14) | (maxLen << 21);
        for (int ln = 0; ln < 64; ++ln) {
          out.laneMeta.push_back(-1);
          out.laneMeta.push_back(0);
          out.laneMeta.push_back(0);
          out.laneMeta.push_back(1 | common);
          out.laneDinv.push_back(0.0);
          out.laneDense.push_back(0);
        }
        int lane = 0, runs = 0;
        for (const Item& it : t) {
          const int b = it.block, s0 = bs[(size_t)b];
          const int nr = it.row >= 0 ? 1 : it.len, k0 = it.row >= 0 ? it.row : s0;
          const int doff = (int)out.dense.size() - denseBase;
          for (int r = 0; r < nr; ++r, ++lane) {
            const int k = k0 + r;
            int p0, p1;
            extRange(k, b, p0, p1);
            out.slotRank.push_back(k);
            out.slotRunBeg.push_back(runs);
            for (int p = p0; p < p1; p += kRun) {
              for (int q = 0; q < kRun; ++q) {
                if (p + q < p1) {
                  out.runCol.push_back(fwd ? f.rc[(size_t)(p + q)] : f.ci[(size_t)(p + q)]);
                  out.runVal.push_back(fwd ? f.rx[(size_t)(p + q)] : f.cx[(size_t)(p + q)]);
                } else {
                  out.runCol.push_back(0);
                  out.runVal.push_back(0.0);
                }
              }
              ++runs;
            }
            out.slotRunEnd.push_back(runs);
            const size_t mi = meta0 + 4 * (size_t)lane;
            out.laneMeta[mi + 0] = k;
            out.laneMeta[mi + 1] = out.slotRunBeg.back();
            out.laneMeta[mi + 2] = runs;
            out.laneMeta[mi + 3] = nr | (r << 7) | common | (it.row >= 0 ? (1 << 28) : 0);
            out.laneDinv[meta0 / 4 + (size_t)lane] = f.dinv[(size_t)k];
            out.laneDense[meta0 / 4 + (size_t)lane] =
                it.row >= 0 ? rpos[(size_t)b] + (k - s0) : (nr > 1 ? doff : 0);
          }
          if (it.row < 0 && nr > 1) {
            // a short chain: forward [s * nr + t] = X(t, s), backward X(s, t)
            std::vector<double> X;
            chainInverse(f, s0, nr, X);
            for (int sc = 0; sc < nr; ++sc)
              for (int tt = 0; tt < nr; ++tt)
                out.dense.push_back(fwd ? X[(size_t)tt * nr + sc] : X[(size_t)sc * nr + tt]);
          }
        }
      }
      while ((out.dense.size() - (size_t)denseBase) % 4) out.dense.push_back(0.0);
      out.maxDense = std::max(out.maxDense, (int)out.dense.size() - denseBase);
      // phase A's dense tasks: per 8 rows of a dense chain, its external entries
      // over the union U, column-major (8 values per column, zero where a row has
      // no entry)
      for (int b : denseChains) {
        const int s0 = bs[(size_t)b], len = bs[(size_t)b + 1] - s0;
        const std::vector<int>& U = unions[(size_t)b];
        const int c = (int)U.size();
        const int uoff = (int)out.ucol.size();
        out.ucol.insert(out.ucol.end(), U.begin(), U.end());
        for (int r0 = 0; r0 < len; r0 += 8) {
          const int nrows = std::min(8, len - r0);
          const size_t eoff = out.epanel.size();
          out.epanel.resize(eoff + (size_t)c * 8, 0.0);
          for (int q = 0; q < nrows; ++q) {
            const int k = s0 + r0 + q;
            int p0, p1;
            extRange(k, b, p0, p1);
            for (int p = p0; p < p1; ++p) {
              const int col = fwd ? f.rc[(size_t)p] : f.ci[(size_t)p];
              const int j = (int)(std::lower_bound(U.begin(), U.end(), col) - U.begin());
              out.epanel[eoff + (size_t)j * 8 + q] = fwd ? f.rx[(size_t)p] : f.cx[(size_t)p];
            }
          }
          const int et[8] = {(int)eoff, c, uoff, (s0 + r0) | (nrows << 24), rpos[(size_t)b] + r0, 0, 0, 0};
          out.eTask.insert(out.eTask.end(), et, et + 8);
        }
      }
      // phase B: the long chains' panels
      for (int b : longChains) {
        const int s0 = bs[(size_t)b], s = bs[(size_t)b + 1] - s0, spad = round4(s);
        std::vector<double>& X = Xc[(size_t)b];
        if (X.empty()) chainInverse(f, s0, s, X);
        for (int r0 = 0; r0 < s; r0 += 8) {
          const int nrows = std::min(8, s - r0);
          const int c0 = fwd ? 0 : r0;
          const int ld = fwd ? round4(std::min(s, r0 + 8)) : spad - r0;
          const int xoff = (int)out.xpanel.size();
          for (int q = 0; q < nrows; ++q)
            for (int j = 0; j < ld; ++j) {
              const int t = r0 + q, u = c0 + j;  // output row t, input column u
              double v = 0.0;
              if (u < s) v = fwd ? (u <= t ? X[(size_t)t * s + u] : 0.0) : (u >= t ? X[(size_t)u * s + t] : 0.0);
              out.xpanel.push_back(v);
            }
          out.bTask.push_back(xoff);
          out.bTask.push_back(ld);
          out.bTask.push_back(rpos[(size_t)b] + c0);
          out.bTask.push_back((s0 + r0) | (nrows << 24));
        }
      }
      out.stepTask.push_back((int)out.taskInfo.size());
      out.stepRun.push_back((int)(out.runCol.size() / kRun));
      out.stepDense.push_back((int)out.dense.size());
      out.stepB.push_back((int)(out.bTask.size() / 4));
      out.stepE.push_back((int)(out.eTask.size() / 8));
    }
    out.taskRun.push_back((int)(out.runCol.size() / kRun));
    out.nstep = nlev;
    if (out.rTotal == 0) out.rTotal = 4;
  }
}

struct Instance {
  Scene scene;
  Config cfg;
  Topology topo;
  Matrix A;
  Factor fac;
  std::vector<double> brhs;
  bool topoBuilt = false, analyzed = false, factored = false;
  uint64_t numStamp = 0, strStamp = 0;
  int factorizeCount = 0, analyzeCount = 0;
  double analyzeSec = 0, factorSec = 0;
  HostSweep hf, hb;
  // fast mode: the forward / backward block schedules.  An instance split over
  // two workgroups (paired) holds them per region of the nested-dissection top
  // split: bf / bb over the first half, sf / sb over the separator (the lead's),
  // pf / pb over the second half (the partner's).
  HostBlockSweep bf, bb, sf, sb, pf, pb;
  bool paired = false;
  std::vector<double> inv;  // dense solve mode: inverse of the component matrix, rank space
  bool packed = false;
  // device addresses of the writable state (inside the arena)
  double *dx = nullptr, *dv = nullptr, *dpin = nullptr, *dbrhs = nullptr;
  int* distats = nullptr;
  long long* dlstats = nullptr;
  double* ddstats = nullptr;
  double bytesPerIter = 0;
  double bytesPerSubstepFixed = 0;
};

// Byte arena: one host staging buffer mirrored by one device allocation.
struct Arena {
  std::vector<char> host;
  size_t add(const void* p, size_t bytes, size_t align = 256) {
    size_t off = (host.size() + align - 1) / align * align;
    host.resize(off + std::max<size_t>(bytes, 1));
    if (p && bytes) std::memcpy(host.data() + off, p, bytes);
    return off;
  }
  template <typename T>
  size_t vec(const std::vector<T>& v) {
    return add(v.data(), v.size() * sizeof(T));
  }
};

}  // namespace

struct pdx_ctx {
  int device = 0;
  int prec = PDX_PREC_FP64;
  int mode = PDX_SOLVE_EXACT;
  hipStream_t stream = nullptr;
  std::vector<Instance> inst;
  char* dbuf = nullptr;
  size_t dbytes = 0;
  InstDev* dInst = nullptr;
  int* dOrder = nullptr;
  std::vector<InstDev> hInst;
  bool uploaded = false;
  bool dirty = true;
  bool ldsW = true;
  size_t lds = 0;
  int block = 512;
  int chainCap = 32;
  int ordering = PDX_ORDER_AMD;
  bool idx16 = false;  // fast-mode run columns stored as 16-bit ranks (every instance has m <= 65535)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double lastKernelMs = 0.0;
  int lastLaunches = 0;
  double bytesIter = 0.0;
  std::string err;
  double* dScratchB = nullptr;
  double* dScratchX = nullptr;
  int scratchN = 0;
  int uploadEpoch = 0;
  // launch form: 0 = one workgroup per instance (batch), 1 = the whole grid on one
  // instance (cooperative launch)
  int launchMode = 0;
  int gridBlocks = 0;       // 0 = every co-resident workgroup
  int gridLaunchBlocks = 0; // blocks of the last grid upload
  int waveRuns = 0;         // runs per wavefront LDS slice (grid form)
  double* dGridSlots = nullptr;
  // dense solve mode: the device copy of the inverse (m x m Real), refreshed when
  // the factor is
  char* dLinv = nullptr;
  size_t linvBytes = 0;
  bool linvStale = true;
  bool denseLds = false;  // W staged in each workgroup's LDS for the product
  // instances whose factor has at least splitMin entries run on two workgroups (0 = never);
  // the differentiable rollout (rollMode) packs every instance for one workgroup
  long long splitMin = 0;
  bool rollMode = false;
  int npairs = 0;
  int* dGiveUp = nullptr;      // the pairs' give-up flag (device)
  // differentiable rollout: pin trajectories (host), device buffers, seeds
  struct Roll {
    int nsub = 0, every = 0, nseg = 0, epoch = -1;
    bool forwardDone = false;
    char* dbuf = nullptr;
    std::vector<std::vector<double>> traj;   // per instance: nsub x npin x 3, or empty
    std::vector<int> trajSub;                // substeps of each set trajectory
    std::vector<size_t> ckpt;                // per instance: offset of (nseg+1) x {x, v}
    std::vector<std::vector<double>> seedX, seedV;
    double forwardMs = 0.0, adjointMs = 0.0;
  } roll;
};

namespace {

int fail(pdx_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHECK(ctx, expr)                                                          \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) return fail(ctx, PDX_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void fillConfig(Config& c, const pdx_config* p) {
  c.dt = p->dt;
  for (int d = 0; d < 3; ++d) c.g[d] = p->gravity[d];
  c.damping = p->damping;
  c.maxIter = p->max_iterations;
  c.relTol = p->rel_tol;
  c.absTol = p->abs_tol;
  c.resTol = p->residual_tol;
  c.chebRho = p->chebyshev_rho;
  c.chebStart = p->chebyshev_start;
}

int validateScene(const pdx_scene* s) {
  if (!s || s->n_vertices <= 0 || !s->positions || !s->masses || !s->pinned) return PDX_EINVAL;
  if (s->n_edges < 0 || s->n_bends < 0) return PDX_EINVAL;
  if (s->n_edges > 0 && (!s->edge_a || !s->edge_b || !s->edge_rest || !s->edge_kappa)) return PDX_EINVAL;
  if (s->n_bends > 0 && (!s->bend_a || !s->bend_b || !s->bend_c || !s->bend_k)) return PDX_EINVAL;
  for (int e = 0; e < s->n_edges; ++e) {
    const int a = s->edge_a[e], b = s->edge_b[e];
    if (a < 0 || b < 0 || a >= s->n_vertices || b >= s->n_vertices || a == b) return PDX_EINVAL;
  }
  for (int q = 0; q < s->n_bends; ++q) {
    const int a = s->bend_a[q], b = s->bend_b[q], c = s->bend_c[q];
    if (a < 0 || b < 0 || c < 0 || a >= s->n_vertices || b >= s->n_vertices || c >= s->n_vertices) return PDX_EINVAL;
  }
  return PDX_OK;
}

// Downloads the device state of every instance into the host scenes (before a
// re-upload) so that stepping continues from where it was.
int pullState(pdx_ctx* c) {
  if (!c->uploaded) return PDX_OK;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  for (auto& I : c->inst) {
    if (!I.dx) continue;  // added since the last upload: its host state is current
    HIPCHECK(c, hipMemcpy(I.scene.x.data(), I.dx, I.scene.x.size() * 8, hipMemcpyDeviceToHost));
    HIPCHECK(c, hipMemcpy(I.scene.v.data(), I.dv, I.scene.v.size() * 8, hipMemcpyDeviceToHost));
  }
  return PDX_OK;
}

double sweepBytes(const HostSweep& s, size_t w) {
  return (double)s.col.size() * (4 + w) + (double)s.slotRank.size() * 8 + (double)s.chunkRows.size() * 12 +
         (double)s.levChunk.size() * 4;
}

size_t align16(size_t b) { return (b + 15) / 16 * 16; }
int maxOverSchedules(const Instance& I, int HostBlockSweep::*field) {
  const HostBlockSweep* all[6] = {&I.bf, &I.bb, &I.sf, &I.sb, &I.pf, &I.pb};
  int r = 0;
  for (const HostBlockSweep* h : all) r = std::max(r, h->*field);
  return r;
}
size_t partBytes(const Instance& I, size_t w) {
  return align16((size_t)maxOverSchedules(I, &HostBlockSweep::maxRuns) * 3 * w);
}
// FP64 stages a step's chain inverse blocks in LDS; FP32 loads their rows into
// registers and stages nothing
size_t denseBytes(const Instance& I, size_t w) {
  return w == 8 ? align16((size_t)maxOverSchedules(I, &HostBlockSweep::maxDense) * w) : 0;
}
// step tables of a workgroup's four sweeps (the kernel's loadStepTables): the
// lead (or a single workgroup) bf, bb, sf, sb; a pair's partner pf, pb, sf, sb
size_t stepBytes(const Instance& I) {
  const size_t sep = (size_t)(I.sf.nstep + 1 + I.sb.nstep + 1);
  const size_t lead = (size_t)(I.bf.nstep + 1 + I.bb.nstep + 1) + sep;
  const size_t part = (size_t)(I.pf.nstep + 1 + I.pb.nstep + 1) + sep;
  return align16(3 * std::max(lead, part) * sizeof(int));
}

double blockBytes(const HostBlockSweep& s, size_t w) {
  return (double)s.runCol.size() * (4 + w) + (double)s.dense.size() * w + (double)s.slotRank.size() * 12 +
         (double)s.taskInfo.size() * 12 + (double)s.stepTask.size() * 12;
}

// Whether an instance runs on two workgroups: the fast FP32-solve batch form with
// W in LDS, a nested-dissection factor with a top-level split, no Chebyshev
// blend, outside the differentiable rollout, and at least splitMin factor entries
// (the entries, which both sweeps stream and which grow with bending's wider
// coupling, predict an instance's time far better than its vertex count).
bool wantPair(const pdx_ctx* c, const Instance& I, bool wCanFit) {
  return c->splitMin > 0 && !c->rollMode && c->mode == PDX_SOLVE_FAST && c->prec == PDX_PREC_MIXED32 &&
         c->launchMode == 0 && wCanFit && I.cfg.chebRho <= 0.0 && I.fac.nnz >= c->splitMin &&
         I.fac.top[0] > 0 && I.fac.top[1] > 0 && I.fac.top[0] + I.fac.top[1] < I.fac.m;
}

// Factor every instance whose stamps changed (parallel over instances), repack.
int refreshFactors(pdx_ctx* c, bool& changed) {
  changed = false;
  const int ni = (int)c->inst.size();
  std::vector<int> todo;
  for (int i = 0; i < ni; ++i) {
    Instance& I = c->inst[(size_t)i];
    const uint64_t ns = numericStamp(I.scene, I.cfg.dt);
    if (!I.factored || ns != I.numStamp || !I.packed) todo.push_back(i);
  }
  if (todo.empty()) return PDX_OK;
  if (c->mode == PDX_SOLVE_DENSE) {
    if (ni != 1) return fail(c, PDX_EUNSUPPORTED, "the dense solve mode takes exactly one instance");
    if (c->inst[0].scene.n > kDenseMaxM)
      return fail(c, PDX_EUNSUPPORTED, "the dense solve mode takes at most " + std::to_string(kDenseMaxM) + " vertices");
  }
  changed = true;
  // W (3m Real) sits in LDS only when every instance's fits beside some staging;
  // otherwise it stays in HBM for all of them and the steps keep the default caps
  int maxM = 0;
  for (const auto& I : c->inst) maxM = std::max(maxM, I.scene.n);
  const bool wCanFit = kRedBytes + (size_t)3 * maxM * (c->prec == PDX_PREC_MIXED32 ? 4 : 8) + 16 * 1024 <= kLdsLimit;
  std::vector<int> rcs(todo.size(), 0);
  std::atomic<size_t> next{0};
  auto worker = [&]() {
    for (;;) {
      const size_t t = next.fetch_add(1);
      if (t >= todo.size()) return;
      Instance& I = c->inst[(size_t)todo[t]];
      const double t0 = now();
      auto pack = [&]() {
        if (c->mode == PDX_SOLVE_DENSE) {
          const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
          denseInverse(I.fac, I.inv, (int)std::min(16u, hw));
          c->linvStale = true;
        } else if (c->mode == PDX_SOLVE_FAST) {
          const int w = c->prec == PDX_PREC_MIXED32 ? 4 : 8;
          // blocks never cross the top split of a nested-dissection order, so
          // that a row's runs (and bits) are the same in one workgroup or two
          const int m = I.fac.m, rA = I.fac.top[0], rB = I.fac.top[0] + I.fac.top[1];
          const std::vector<int> cuts = rA > 0 ? std::vector<int>{rA, rB} : std::vector<int>{};
          I.paired = wantPair(c, I, wCanFit);
          for (HostBlockSweep* h : {&I.sf, &I.sb, &I.pf, &I.pb}) *h = HostBlockSweep{};
          auto packAll = [&](int runCap, int stepRunCap, int denseCap) {
            if (I.paired) {
              packBlocked(I.fac, true, c->chainCap, runCap, stepRunCap, denseCap, I.bf, 0, rA, cuts);
              packBlocked(I.fac, false, c->chainCap, runCap, stepRunCap, denseCap, I.bb, 0, rA, cuts);
              packBlocked(I.fac, true, c->chainCap, runCap, stepRunCap, denseCap, I.pf, rA, rB, cuts);
              packBlocked(I.fac, false, c->chainCap, runCap, stepRunCap, denseCap, I.pb, rA, rB, cuts);
              packBlocked(I.fac, true, c->chainCap, runCap, stepRunCap, denseCap, I.sf, rB, m, cuts);
              packBlocked(I.fac, false, c->chainCap, runCap, stepRunCap, denseCap, I.sb, rB, m, cuts);
            } else {
              packBlocked(I.fac, true, c->chainCap, runCap, stepRunCap, denseCap, I.bf, 0, m, cuts);
              packBlocked(I.fac, false, c->chainCap, runCap, stepRunCap, denseCap, I.bb, 0, m, cuts);
            }
          };
          if (c->launchMode == 1) {
            // grid form: one step per level of whole chains, a task's runs within
            // one wavefront's LDS slice
            packLevels(I.fac, c->chainCap, c->waveRuns, I.bf, I.bb);
          } else {
            // per-step LDS caps: run partials <= 48 KB; FP64 stages the chain
            // blocks' inverses too (<= 32 KB), FP32 loads them into registers
            const bool stage = w == 8;
            int runCap = 48 * 1024 / (3 * w), denseCap = stage ? 32 * 1024 / w : 1 << 29;
            packAll(runCap, runCap, denseCap);
            // The workgroup also holds W (3m).  When W and the staging overflow
            // LDS, the steps are cut finer to the room W leaves (3:2 between run
            // partials and staged blocks), shrinking by the overflow until they fit.
            const size_t wBytes = (size_t)3 * I.fac.m * w;
            auto need = [&]() { return kRedBytes + wBytes + partBytes(I, w) + denseBytes(I, w) + stepBytes(I); };
            size_t room = kLdsLimit > kRedBytes + wBytes + stepBytes(I) ? kLdsLimit - kRedBytes - wBytes - stepBytes(I) : 0;
            for (int tries = 0; wCanFit && need() > kLdsLimit && room >= 8 * 1024 && tries < 8; ++tries) {
              runCap = std::min(runCap, (int)((stage ? room * 3 / 5 : room) / (3 * w)));
              if (stage) denseCap = std::min(denseCap, (int)(room * 2 / 5 / w));
              packAll(runCap, runCap, denseCap);
              const size_t over = need() > kLdsLimit ? need() - kLdsLimit : 0;
              room = room > over + 1024 ? room - over - 1024 : 0;
            }
          }
        } else {
          const bool longRows = c->prec == PDX_PREC_FP64;
          packSweep(I.fac.fwd, true, I.fac, longRows, I.hf);
          packSweep(I.fac.bwd, false, I.fac, longRows, I.hb);
        }
        I.packed = true;
      };
      if (I.factored && I.numStamp == numericStamp(I.scene, I.cfg.dt)) {
        // same matrix, new launch schedule: repack only
        pack();
        rcs[t] = 0;
        continue;
      }
      if (!I.topoBuilt) {
        buildTopology(I.scene, I.topo);
        I.topoBuilt = true;
      }
      assembleMatrix(I.scene, I.cfg.dt, I.A);
      bendingRhs(I.scene, I.brhs);
      const uint64_t ss = structureStamp(I.scene);
      int rc = 0;
      if (!I.analyzed || ss != I.strStamp) {
        rc = analyzeAndFactor(I.scene, I.A, I.fac, c->ordering == PDX_ORDER_ND ? kOrderNd : kOrderAmd);
        if (rc == 0) {
          I.analyzed = true;
          I.strStamp = ss;
          I.analyzeCount += 1;
        }
        I.analyzeSec = now() - t0;
      } else {
        rc = refactorNumeric(I.A, I.fac);
      }
      if (rc == 0) {
        I.factorizeCount += 1;
        I.factored = true;
        I.numStamp = numericStamp(I.scene, I.cfg.dt);
        pack();
      }
      I.factorSec = now() - t0;
      rcs[t] = rc;
    }
  };
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int nth = (int)std::min<size_t>(todo.size(), std::min(16u, hw));
  std::vector<std::thread> pool;
  for (int t = 1; t < nth; ++t) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  for (size_t t = 0; t < todo.size(); ++t)
    if (rcs[t] != 0)
      return fail(c, rcs[t], "factorisation failed for instance " + std::to_string(todo[t]) +
                                 (rcs[t] == PDX_ENOTSPD ? " (zero pivot: not positive definite)"
                                                        : " (component factors differ)"));
  return PDX_OK;
}

template <typename T>
std::vector<T> castVec(const std::vector<double>& v) {
  return std::vector<T>(v.begin(), v.end());
}

// Builds the device arena of all instances and uploads it.
int upload(pdx_ctx* c) {
  const bool fp32 = c->prec == PDX_PREC_MIXED32;
  const size_t w = fp32 ? 4 : 8;
  c->idx16 = true;
  for (auto& I : c->inst) c->idx16 = c->idx16 && I.fac.m <= 65535;
  // every workgroup lays out its own instance: W (3m), then the run partials,
  // the dense blocks and the step tables, so the launch needs the largest sum
  size_t maxExtra = 0, maxBoth = 0;
  for (auto& I : c->inst) {
    const size_t e = c->mode == PDX_SOLVE_FAST ? partBytes(I, w) + denseBytes(I, w) + stepBytes(I) : 0;
    maxExtra = std::max(maxExtra, e);
    maxBoth = std::max(maxBoth, (size_t)3 * I.fac.m * w + e);
  }
  c->ldsW = kRedBytes + maxBoth <= kLdsLimit;
  c->lds = kRedBytes + (c->ldsW ? maxBoth : maxExtra);
  if (c->launchMode == 1) {
    for (auto& I : c->inst)
      if (I.cfg.chebRho > 0.0) return fail(c, PDX_EUNSUPPORTED, "the grid form has no Chebyshev blend");
    c->ldsW = false;
    c->lds = kRedBytes + (size_t)(c->block / 64) * c->waveRuns * 3 * w;
    for (auto& I : c->inst) {
      if (std::max(I.bf.maxTaskRuns, I.bb.maxTaskRuns) > c->waveRuns)
        return fail(c, PDX_EUNSUPPORTED, "a factor row streams more runs than one wavefront's LDS slice");
      if (I.fac.m >= (1 << 24)) return fail(c, PDX_EUNSUPPORTED, "the grid form takes fewer than 2^24 vertices");
    }
  }
  const bool dense = c->mode == PDX_SOLVE_DENSE;
  if (dense) {
    for (auto& I : c->inst)
      if (I.cfg.chebRho > 0.0) return fail(c, PDX_EUNSUPPORTED, "the dense solve mode has no Chebyshev blend");
    // W in HBM; each workgroup copies it into LDS for the product when it fits
    c->ldsW = false;
    const size_t wb = (size_t)3 * c->inst[0].fac.m * w;
    c->denseLds = kRedBytes + wb <= kLdsLimit;
    c->lds = kRedBytes + (c->denseLds ? wb : 0);
  }
  if (c->lds > kLdsLimit)
    return fail(c, PDX_EUNSUPPORTED, "sweep staging does not fit in LDS (" + std::to_string(c->lds) + " B)");
  for (auto& I : c->inst)
    if (I.paired && !c->ldsW)
      return fail(c, PDX_EUNSUPPORTED, "an instance split over two workgroups needs W in LDS");
  Arena ar;
  struct Off {
    size_t x, v, xbs, xhat, prev, mass, pinned, pinpos, brhs, ea, eb, rest, kappa, vstart, vedges, vedgesC,
        bstart, bst, sa, sb, sc, bk, vrank, rankv, dinv, W, ist, lst, dst, recE, rkE, recC, rkC, brR, brK;
    size_t f[9], b[9];
    size_t sch[6][23];  // block schedules bf, bb, sf, sb, pf, pb
    size_t ownLead = 0, ownPart = 0, pairFlags = 0, pairRed = 0, xbuf = 0, flow = 0, flowPart = 0;
    int nOwnLead = 0, nOwnPart = 0;
  };
  std::vector<Off> offs(c->inst.size());
  c->bytesIter = 0;
  for (size_t i = 0; i < c->inst.size(); ++i) {
    Instance& I = c->inst[i];
    const Scene& s = I.scene;
    Off& o = offs[i];
    const size_t n3 = (size_t)3 * s.n;
    o.x = ar.vec(s.x);
    o.v = ar.vec(s.v);
    o.xbs = ar.add(nullptr, n3 * 8);
    o.xhat = ar.add(nullptr, n3 * 8);
    o.prev = ar.add(nullptr, n3 * 8);
    o.mass = ar.vec(s.mass);
    o.pinned = ar.vec(s.pinned);
    o.pinpos = ar.vec(s.pinpos);
    bool anyBrhs = false;
    for (double b : I.brhs) anyBrhs = anyBrhs || b != 0.0;
    o.brhs = (s.B > 0) ? ar.vec(I.brhs) : 0;
    (void)anyBrhs;
    o.ea = ar.vec(s.ea);
    o.eb = ar.vec(s.eb);
    o.rest = ar.vec(s.rest);
    o.kappa = ar.vec(s.kappa);
    std::vector<int> vs(I.topo.vstart.begin(), I.topo.vstart.end());
    std::vector<int> ve(I.topo.vedges.begin(), I.topo.vedges.end());
    std::vector<int> vc(I.topo.vedgesColor.begin(), I.topo.vedgesColor.end());
    std::vector<int> bs(I.topo.bstart.begin(), I.topo.bstart.end());
    std::vector<int> bst(I.topo.bstencil.begin(), I.topo.bstencil.end());
    {
      // flattened incidence / bend records (see InstDev)
      std::vector<int> recE(ve.size()), recC(ve.size());
      std::vector<RestKappa> rkE(ve.size()), rkC(ve.size());
      for (int v = 0; v < s.n; ++v)
        for (int p = vs[(size_t)v]; p < vs[(size_t)v + 1]; ++p)
          for (int q = 0; q < 2; ++q) {
            const int e = q == 0 ? ve[(size_t)p] : vc[(size_t)p];
            const bool isB = s.eb[(size_t)e] == v;
            const int u = isB ? s.ea[(size_t)e] : s.eb[(size_t)e];
            const int r = (u << 2) | (s.pinned[(size_t)u] ? 2 : 0) | (isB ? 1 : 0);
            (q == 0 ? recE : recC)[(size_t)p] = r;
            (q == 0 ? rkE : rkC)[(size_t)p] = RestKappa{s.rest[(size_t)e], s.kappa[(size_t)e]};
          }
      std::vector<BendRec> bR(bst.size());
      std::vector<double> bK(bst.size());
      auto enc = [&](int x) { return (x << 1) | (s.pinned[(size_t)x] ? 1 : 0); };
      for (int v = 0; v < s.n; ++v)
        for (int p = bs[(size_t)v]; p < bs[(size_t)v + 1]; ++p) {
          const int q = bst[(size_t)p];
          const int abc[3] = {s.sa[(size_t)q], s.sb[(size_t)q], s.sc[(size_t)q]};
          const int role = abc[0] == v ? 0 : (abc[1] == v ? 1 : 2);
          int others[2], k = 0;
          for (int t = 0; t < 3; ++t)
            if (t != role) others[k++] = enc(abc[t]);
          bR[(size_t)p] = BendRec{others[0], others[1], role | (s.pinned[(size_t)v] ? 4 : 0),
                                  (int)I.topo.bweight[(size_t)p]};
          bK[(size_t)p] = s.bk[(size_t)q];
        }
      o.recE = ar.vec(recE);
      o.rkE = ar.vec(rkE);
      o.recC = ar.vec(recC);
      o.rkC = ar.vec(rkC);
      o.brR = ar.vec(bR);
      o.brK = ar.vec(bK);
    }
    o.vstart = ar.vec(vs);
    o.vedges = ar.vec(ve);
    o.vedgesC = ar.vec(vc);
    o.bstart = ar.vec(bs);
    o.bst = ar.vec(bst);
    o.sa = ar.vec(s.sa);
    o.sb = ar.vec(s.sb);
    o.sc = ar.vec(s.sc);
    o.bk = ar.vec(s.bk);
    o.vrank = ar.vec(I.fac.vertexRank[0]);
    o.rankv = ar.vec(I.fac.rankVertex[0]);
    o.dinv = fp32 ? ar.vec(castVec<float>(I.fac.dinv)) : ar.vec(I.fac.dinv);
    HostSweep* hs[2] = {&I.hf, &I.hb};
    size_t* fo[2] = {o.f, o.b};
    for (int q = 0; q < 2; ++q) {
      HostSweep& h = *hs[q];
      fo[q][0] = ar.vec(h.levChunk);
      fo[q][1] = ar.vec(h.chunkRows);
      fo[q][2] = ar.vec(h.chunkRowBase);
      fo[q][3] = ar.vec(h.chunkEntBase);
      fo[q][4] = ar.vec(h.slotRank);
      fo[q][5] = ar.vec(h.slotLen);
      fo[q][6] = ar.vec(h.col);
      fo[q][7] = fp32 ? ar.vec(castVec<float>(h.val)) : ar.vec(h.val);
    }
    HostBlockSweep* bsw[6] = {&I.bf, &I.bb, &I.sf, &I.sb, &I.pf, &I.pb};
    for (int q = 0; q < 6; ++q) {
      HostBlockSweep& h = *bsw[q];
      size_t* bo[6] = {o.sch[0], o.sch[1], o.sch[2], o.sch[3], o.sch[4], o.sch[5]};
      bo[q][0] = ar.vec(h.stepTask);
      bo[q][1] = ar.vec(h.stepRun);
      bo[q][2] = ar.vec(h.stepDense);
      bo[q][3] = ar.vec(h.taskInfo);
      bo[q][4] = ar.vec(h.taskSlot);
      bo[q][5] = ar.vec(h.taskDense);
      bo[q][6] = ar.vec(h.slotRank);
      bo[q][7] = ar.vec(h.slotRunBeg);
      bo[q][8] = ar.vec(h.slotRunEnd);
      bo[q][9] = ar.vec(h.runCol);
      bo[q][10] = fp32 ? ar.vec(castVec<float>(h.runVal)) : ar.vec(h.runVal);
      bo[q][11] = fp32 ? ar.vec(castVec<float>(h.dense)) : ar.vec(h.dense);
      bo[q][12] = ar.vec(h.laneMeta);
      bo[q][13] = fp32 ? ar.vec(castVec<float>(h.laneDinv)) : ar.vec(h.laneDinv);
      bo[q][14] = ar.vec(h.laneDense);
      bo[q][15] = ar.vec(h.taskRun);
      bo[q][16] = ar.vec(h.stepB);
      bo[q][17] = ar.vec(h.bTask);
      bo[q][18] = fp32 ? ar.vec(castVec<float>(h.xpanel)) : ar.vec(h.xpanel);
      bo[q][19] = ar.vec(h.stepE);
      bo[q][20] = ar.vec(h.eTask);
      bo[q][21] = ar.vec(h.ucol);
      bo[q][22] = fp32 ? ar.vec(castVec<float>(h.epanel)) : ar.vec(h.epanel);
      if (c->idx16) {
        std::vector<uint16_t> c16(h.runCol.begin(), h.runCol.end());
        bo[q][9] = ar.vec(c16);
      }
    }
    if (I.paired) {
      // owned vertices (ascending): the lead takes the first half and the
      // separator, the partner the second half; the exchange area of the pair
      const int rA = I.fac.top[0], rB = I.fac.top[0] + I.fac.top[1];
      std::vector<int> lead, part;
      for (int v = 0; v < s.n; ++v) {
        const int r = I.fac.vertexRank[0][(size_t)v];
        (r >= rA && r < rB ? part : lead).push_back(v);
      }
      o.nOwnLead = (int)lead.size();
      o.nOwnPart = (int)part.size();
      o.ownLead = ar.vec(lead);
      o.ownPart = ar.vec(part);
      o.pairFlags = ar.add(nullptr, 64 * sizeof(int));
      o.pairRed = ar.add(nullptr, 16 * sizeof(double));
      o.xbuf = ar.add(nullptr, (size_t)3 * I.fac.m * w);
    }
    // W in HBM: 3 Real per rank, 4 in the grid form's level sweeps
    o.W = c->ldsW ? 0 : ar.add(nullptr, (size_t)(c->launchMode == 1 && !dense ? 4 : 3) * I.fac.m * w);
    // the grid form's step counters, kFlowLines lines of 256 bytes each
    // (kFlowStep in the kernels), and R (zero, so its padding stays zero)
    o.flow = ar.add(nullptr, c->launchMode == 1 ? (size_t)256 * kFlowLines * 2 * std::max(1, I.bf.nstep + I.bb.nstep) : 256);
    o.flowPart = ar.add(nullptr, c->launchMode == 1 ? (size_t)3 * std::max(I.bf.rTotal, I.bb.rTotal) * w : 16);
    std::vector<int> ist(4, 0);
    std::vector<long long> lst(16, 0);
    std::vector<double> dst(2, -1.0);
    o.ist = ar.vec(ist);
    o.lst = ar.vec(lst);
    o.dst = ar.vec(dst);
    // Algorithmic (compulsory) bytes of one PD iteration, each array counted once
    // per phase at its storage width: w = factor value width, i = factor index
    // width (2 for 16-bit ranks), scene data and state in FP64:
    //   2F(w+i) + 2(m+1)*4 + m(w+4)            factor forward + backward, maps, 1/D
    // + 2[E(2*4 + 2*8) + (n+1)*4 + 2E*4]       edges + incidence, rhs and residual
    // + 24n*8 + n*8 + n                        vector passes of 3n, mass, pin mask
    // + [B>0](B(3*4 + 8) + (n+1)*4 + 3B(4+8) + 3n*8)   stencils, bend CSR, bend rhs
    const double nb = (double)s.n, Eb = (double)s.E, Bb = (double)s.B, Fb = (double)I.fac.nnz,
                 mb = (double)I.fac.m;
    const double iw = (c->mode == PDX_SOLVE_FAST && c->idx16) ? 2.0 : 4.0;
    // dense solve mode: the inverse (m^2 w) and W read once, the rank -> vertex map
    double bi = dense ? mb * mb * (double)w + 3.0 * mb * (double)w + mb * 4
                      : 2.0 * Fb * ((double)w + iw) + 2.0 * (mb + 1) * 4 + mb * ((double)w + 4);
    bi += 2.0 * (Eb * (2 * 4 + 2 * 8) + (nb + 1) * 4 + 2 * Eb * 4);
    bi += 24.0 * nb * 8 + nb * 8 + nb;
    if (Bb > 0) bi += Bb * (3 * 4 + 8) + (nb + 1) * 4 + 3 * Bb * (4 + 8) + 3 * nb * 8;
    I.bytesPerSubstepFixed = 27.0 * nb * 8;  // predict, inertial rhs, velocity
    I.bytesPerIter = bi;
    c->bytesIter += bi;
  }
  // grid reductions: 2 alternating sets of 2 slots per workgroup, at most 1024 workgroups
  const size_t slotOff = ar.add(nullptr, 4 * 1024 * sizeof(double));
  const size_t giveUpOff = ar.add(nullptr, sizeof(int));
  // descriptors: one per instance (its only workgroup, or a pair's lead), then one
  // per pair for the partner
  std::vector<int> pairOf;  // instance of each pair
  for (size_t i = 0; i < c->inst.size(); ++i)
    if (c->inst[i].paired) pairOf.push_back((int)i);
  const int ni = (int)c->inst.size(), np = (int)pairOf.size();
  const size_t instOff = ar.add(nullptr, sizeof(InstDev) * (size_t)(ni + np));
  // Launch order (workgroup b runs descriptor order[b]), longest first by factor
  // entries.  The pairs come first, each lead and its partner 8 workgroups apart
  // (16-workgroup groups: up to 8 leads, 8 partners), so that with the
  // dispatcher's round-robin over the 8 XCDs a pair shares one L2; the groups'
  // unused places take the heaviest single instances.
  std::vector<int> order, solo;
  for (int i = 0; i < ni; ++i)
    if (!c->inst[(size_t)i].paired) solo.push_back(i);
  auto heavier = [&](int a, int b) { return c->inst[(size_t)a].fac.nnz > c->inst[(size_t)b].fac.nnz; };
  std::stable_sort(solo.begin(), solo.end(), heavier);
  std::vector<int> pq(np);
  for (int p = 0; p < np; ++p) pq[(size_t)p] = p;
  std::stable_sort(pq.begin(), pq.end(), [&](int a, int b) { return heavier(pairOf[(size_t)a], pairOf[(size_t)b]); });
  size_t si = 0;
  for (int p0 = 0; p0 < np; p0 += 8) {
    const int q = std::min(8, np - p0);
    for (int half = 0; half < 2; ++half) {
      for (int j = 0; j < q; ++j) {
        const int p = pq[(size_t)(p0 + j)];
        order.push_back(half == 0 ? pairOf[(size_t)p] : ni + p);
      }
      for (int j = q; j < 8 && si < solo.size(); ++j) order.push_back(solo[si++]);
    }
  }
  while (si < solo.size()) order.push_back(solo[si++]);
  const size_t orderOff = ar.vec(order);

  HIPCHECK(c, hipSetDevice(c->device));
  if (c->dbuf) {
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipFree(c->dbuf));
    c->dbuf = nullptr;
  }
  HIPCHECK(c, hipMalloc(&c->dbuf, ar.host.size()));
  c->dbytes = ar.host.size();
  char* B = c->dbuf;
  if (dense) {
    const Instance& I = c->inst[0];
    const int m = I.fac.m, ld = denseLd(m);
    const size_t bytes = (size_t)m * ld * w;
    if (c->dLinv && c->linvBytes != bytes) {
      HIPCHECK(c, hipFree(c->dLinv));
      c->dLinv = nullptr;
      c->linvStale = true;
    }
    if (!c->dLinv) {
      if (hipMalloc(&c->dLinv, bytes) != hipSuccess) {
        c->dLinv = nullptr;
        return fail(c, PDX_ENOMEM, "dense solve mode: no room for the " + std::to_string(bytes) + " B inverse");
      }
      c->linvBytes = bytes;
      c->linvStale = true;
    }
    if (c->linvStale) {
      // rows padded to ld with zeros (16-byte aligned rows for the product's loads)
      auto put = [&](auto zero) -> int {
        using T = decltype(zero);
        std::vector<T> rows((size_t)m * ld, zero);
        for (int r = 0; r < m; ++r)
          for (int j = 0; j < m; ++j) rows[(size_t)r * ld + j] = (T)I.inv[(size_t)r * m + j];
        HIPCHECK(c, hipMemcpy(c->dLinv, rows.data(), bytes, hipMemcpyHostToDevice));
        return PDX_OK;
      };
      const int rc = fp32 ? put(0.0f) : put(0.0);
      if (rc) return rc;
      c->linvStale = false;
    }
  }
  c->hInst.assign((size_t)(ni + np), InstDev{});
  int* dGiveUp = (int*)(B + giveUpOff);
  for (size_t i = 0; i < c->inst.size(); ++i) {
    Instance& I = c->inst[i];
    const Off& o = offs[i];
    InstDev& d = c->hInst[i];
    const Scene& s = I.scene;
    d.n = s.n;
    d.E = s.E;
    d.B = s.B;
    d.m = I.fac.m;
    d.maxIter = I.cfg.maxIter;
    d.wantResidual = I.cfg.resTol > 0.0 ? 1 : 0;
    d.hasBendRhs = s.B > 0;
    d.ldsW = c->ldsW ? 1 : 0;
    const double h = I.cfg.dt;
    d.h = h;
    d.invH = 1.0 / h;
    d.invH2 = 1.0 / (h * h);
    d.dampInvH = (1.0 - I.cfg.damping) * (1.0 / h);
    for (int q = 0; q < 3; ++q) d.gh2[q] = I.cfg.g[q] * (h * h);
    const double gl = std::sqrt(I.cfg.g[0] * I.cfg.g[0] + I.cfg.g[1] * I.cfg.g[1] + I.cfg.g[2] * I.cfg.g[2]);
    d.gScaleH2 = gl * h * h;
    d.relTol = I.cfg.relTol;
    d.absTol = I.cfg.absTol;
    d.resTol = I.cfg.resTol;
    d.resThresh = I.cfg.resTol * std::max(gl, 1.0e-12);
    d.chebRho = I.cfg.chebRho;
    d.chebStart = I.cfg.chebStart;
    d.x = (double*)(B + o.x);
    d.vel = (double*)(B + o.v);
    d.xbs = (double*)(B + o.xbs);
    d.xhat = (double*)(B + o.xhat);
    d.prev = (double*)(B + o.prev);
    d.mass = (const double*)(B + o.mass);
    d.pinned = (const uint8_t*)(B + o.pinned);
    d.pinpos = (double*)(B + o.pinpos);
    d.brhs = s.B > 0 ? (double*)(B + o.brhs) : nullptr;
    d.ea = (const int*)(B + o.ea);
    d.eb = (const int*)(B + o.eb);
    d.rest = (const double*)(B + o.rest);
    d.kappa = (const double*)(B + o.kappa);
    d.vstart = (const int*)(B + o.vstart);
    d.vedges = (const int*)(B + o.vedges);
    d.vedgesColor = (const int*)(B + o.vedgesC);
    d.recE = (const int*)(B + o.recE);
    d.rkE = (const RestKappa*)(B + o.rkE);
    d.recC = (const int*)(B + o.recC);
    d.rkC = (const RestKappa*)(B + o.rkC);
    d.brRec = (const BendRec*)(B + o.brR);
    d.brK = (const double*)(B + o.brK);
    d.bstart = (const int*)(B + o.bstart);
    d.bstencil = (const int*)(B + o.bst);
    d.sa = (const int*)(B + o.sa);
    d.sb = (const int*)(B + o.sb);
    d.sc = (const int*)(B + o.sc);
    d.bk = (const double*)(B + o.bk);
    d.vrank = (const int*)(B + o.vrank);
    d.rankv = (const int*)(B + o.rankv);
    d.dinv = B + o.dinv;
    SweepDev* sd[2] = {&d.fwd, &d.bwd};
    const size_t* fo[2] = {o.f, o.b};
    const HostSweep* hs[2] = {&I.hf, &I.hb};
    for (int q = 0; q < 2; ++q) {
      sd[q]->nlev = hs[q]->nlev;
      sd[q]->levChunk = (const int*)(B + fo[q][0]);
      sd[q]->chunkRows = (const int*)(B + fo[q][1]);
      sd[q]->chunkRowBase = (const int*)(B + fo[q][2]);
      sd[q]->chunkEntBase = (const int*)(B + fo[q][3]);
      sd[q]->slotRank = (const int*)(B + fo[q][4]);
      sd[q]->slotLen = (const int*)(B + fo[q][5]);
      sd[q]->col = (const int*)(B + fo[q][6]);
      sd[q]->val = B + fo[q][7];
    }
    auto fillSweep = [&](BlockSweepDev* bd, const size_t* bo, const HostBlockSweep& h) {
      bd->nstep = h.nstep;
      bd->cap = c->chainCap;
      bd->stepTask = (const int*)(B + bo[0]);
      bd->stepRun = (const int*)(B + bo[1]);
      bd->stepDense = (const int*)(B + bo[2]);
      bd->taskInfo = (const int*)(B + bo[3]);
      bd->taskSlot = (const int*)(B + bo[4]);
      bd->taskDense = (const int*)(B + bo[5]);
      bd->slotRank = (const int*)(B + bo[6]);
      bd->slotRunBeg = (const int*)(B + bo[7]);
      bd->slotRunEnd = (const int*)(B + bo[8]);
      bd->runCol = (const void*)(B + bo[9]);
      bd->runVal = B + bo[10];
      bd->dense = B + bo[11];
      bd->laneMeta = (const int*)(B + bo[12]);
      bd->laneDinv = B + bo[13];
      bd->laneDense = (const int*)(B + bo[14]);
      bd->taskRun = (const int*)(B + bo[15]);
      bd->stepB = (const int*)(B + bo[16]);
      bd->bTask = (const int*)(B + bo[17]);
      bd->xpanel = B + bo[18];
      bd->rTotal = h.rTotal;
      bd->pad = 0;
      bd->stepE = (const int*)(B + bo[19]);
      bd->eTask = (const int*)(B + bo[20]);
      bd->ucol = (const int*)(B + bo[21]);
      bd->epanel = B + bo[22];
    };
    fillSweep(&d.bfwd, o.sch[0], I.bf);
    fillSweep(&d.bbwd, o.sch[1], I.bb);
    fillSweep(&d.bfwdS, o.sch[2], I.sf);
    fillSweep(&d.bbwdS, o.sch[3], I.sb);
    d.partOff = (int)(kRedBytes + (c->ldsW ? (size_t)3 * I.fac.m * w : 0));
    d.denseOff = d.partOff + (int)(partBytes(I, w));
    d.stepOff = d.denseOff + (int)(denseBytes(I, w));
    d.Wg = c->ldsW ? nullptr : (void*)(B + o.W);
    d.Linv = dense ? (const void*)c->dLinv : nullptr;
    d.denseLd = dense ? denseLd(I.fac.m) : 0;
    d.denseLdsW = (dense && c->denseLds) ? 1 : 0;
    d.istats = (int*)(B + o.ist);
    d.lstats = (long long*)(B + o.lst);
    d.dstats = (double*)(B + o.dst);
    d.pairRole = 0;
    d.own = nullptr;
    d.nOwn = s.n;
    d.flow = (int*)(B + o.flow);
    d.flowPart = (void*)(B + o.flowPart);
    d.pair.giveUp = dGiveUp;  // also the grid form's flow waits
    if (I.paired) {
      d.rA = I.fac.top[0];
      d.rB = I.fac.top[0] + I.fac.top[1];
      d.pair.flags = (int*)(B + o.pairFlags);
      d.pair.giveUp = dGiveUp;
      d.pair.red = (double*)(B + o.pairRed);
      d.pair.xbuf = (void*)(B + o.xbuf);
      // the partner: the second half's sweeps and vertices
      const int p = (int)(std::find(pairOf.begin(), pairOf.end(), (int)i) - pairOf.begin());
      InstDev& e = c->hInst[(size_t)(ni + p)];
      e = d;
      e.pairRole = 2;
      e.own = (const int*)(B + o.ownPart);
      e.nOwn = o.nOwnPart;
      fillSweep(&e.bfwd, o.sch[4], I.pf);
      fillSweep(&e.bbwd, o.sch[5], I.pb);  // the separator's sweeps as the lead's
      d.pairRole = 1;
      d.own = (const int*)(B + o.ownLead);
      d.nOwn = o.nOwnLead;
    }
    I.dx = d.x;
    I.dv = d.vel;
    I.dpin = (double*)d.pinpos;
    I.dbrhs = (double*)d.brhs;
    I.distats = d.istats;
    I.dlstats = d.lstats;
    I.ddstats = d.dstats;
  }
  std::memcpy(ar.host.data() + instOff, c->hInst.data(), sizeof(InstDev) * c->hInst.size());
  HIPCHECK(c, hipMemcpy(c->dbuf, ar.host.data(), ar.host.size(), hipMemcpyHostToDevice));
  c->npairs = np;
  c->dGiveUp = dGiveUp;
  c->dInst = (InstDev*)(B + instOff);
  c->dGridSlots = (double*)(B + slotOff);
  c->dOrder = (int*)(B + orderOff);
  HIPCHECK(c, setLdsLimit(kLdsLimit));
  c->uploaded = true;
  c->dirty = false;
  c->uploadEpoch += 1;
  return PDX_OK;
}

// The HIP stream and events are created on first device use, so that host-only
// entry points (factor inspection, scene checks) work without a GPU.
int ensureDevice(pdx_ctx* c) {
  if (c->stream) return PDX_OK;
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, hipStreamCreate(&c->stream));
  HIPCHECK(c, hipEventCreate(&c->ev0));
  HIPCHECK(c, hipEventCreate(&c->ev1));
  return PDX_OK;
}

int ensureReady(pdx_ctx* c) {
  if (c->inst.empty()) return fail(c, PDX_ENOTREADY, "no instance");
  bool changed = false;
  if (c->uploaded && !c->dirty) {
    // cheap stamp re-check (kappa / config changes mark the context dirty)
    return PDX_OK;
  }
  int rc = ensureDevice(c);
  if (rc) return rc;
  rc = pullState(c);
  if (rc) return rc;
  rc = refreshFactors(c, changed);
  if (rc) return rc;
  // carry statistics across a re-upload
  std::vector<std::vector<int>> ist;
  std::vector<std::vector<long long>> lst;
  if (c->uploaded) {
    for (auto& I : c->inst) {
      std::vector<int> a(4, 0);
      std::vector<long long> b(16, 0);
      if (I.distats) {  // instances added since the last upload start from zero
        HIPCHECK(c, hipMemcpy(a.data(), I.distats, 16, hipMemcpyDeviceToHost));
        HIPCHECK(c, hipMemcpy(b.data(), I.dlstats, 128, hipMemcpyDeviceToHost));
      }
      ist.push_back(a);
      lst.push_back(b);
    }
  }
  rc = upload(c);
  if (rc) return rc;
  for (size_t i = 0; i < ist.size(); ++i) {
    HIPCHECK(c, hipMemcpy(c->inst[i].distats, ist[i].data(), 16, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->inst[i].dlstats, lst[i].data(), 128, hipMemcpyHostToDevice));
  }
  return PDX_OK;
}

int gridLaunchBlocks(pdx_ctx* c) {
  const int maxb = gridBlocksMax(c->block, c->lds, c->prec == PDX_PREC_MIXED32, c->idx16);
  const int want = c->gridBlocks > 0 ? c->gridBlocks : maxb;
  return std::min(std::min(want, maxb), 1024);
}

// LDS the grid form needs at the current workgroup size: the reduction slots and
// one run-partial slice per wavefront.
size_t gridLdsNeed(const pdx_ctx* c) {
  const size_t w = c->prec == PDX_PREC_MIXED32 ? 4 : 8;
  return kRedBytes + (size_t)(c->block / 64) * c->waveRuns * 3 * w;
}

int launchGridForm(pdx_ctx* c, int nsub) {
  if (c->lds < gridLdsNeed(c))
    return fail(c, PDX_EINVAL, "grid form: the launch's LDS does not hold a slice per wavefront");
  const int g = gridLaunchBlocks(c);
  if (g <= 0) return fail(c, PDX_EHIP, "grid form: no co-resident workgroup fits");
  HIPCHECK(c, launchGrid(c->dInst, 0, nsub, c->dGridSlots, c->waveRuns, g, c->block, c->lds,
                         c->prec == PDX_PREC_MIXED32, c->idx16, c->stream));
  return PDX_OK;
}

// Dense solve mode: cooperative launch, every co-resident workgroup unless the
// caller (pdx_set_launch) or the autotuner chose fewer.
int denseLaunchBlocks(pdx_ctx* c) {
  const int maxb = denseBlocksMax(c->block, c->lds, c->prec == PDX_PREC_MIXED32);
  const int want = c->gridBlocks > 0 ? c->gridBlocks : maxb;
  return std::min(std::min(want, maxb), 1024);
}

int launchDenseForm(pdx_ctx* c, int nsub) {
  const int g = denseLaunchBlocks(c);
  if (g <= 0) return fail(c, PDX_EHIP, "dense solve mode: no co-resident workgroup fits");
  HIPCHECK(c, launchDense(c->dInst, 0, nsub, c->dGridSlots, g, c->block, c->lds, c->prec == PDX_PREC_MIXED32,
                          c->stream));
  return PDX_OK;
}

// Launch forms that are replayed directly rather than captured in a graph.
bool cooperative(const pdx_ctx* c) { return c->launchMode == 1 || c->mode == PDX_SOLVE_DENSE; }

int launchCooperative(pdx_ctx* c, int nsub) {
  return c->mode == PDX_SOLVE_DENSE ? launchDenseForm(c, nsub) : launchGridForm(c, nsub);
}

int launch(pdx_ctx* c, int nsub) {
  HIPCHECK(c, hipSetDevice(c->device));
  if (cooperative(c)) {
    HIPCHECK(c, hipEventRecord(c->ev0, c->stream));
    const int rc = launchCooperative(c, nsub);
    if (rc) return rc;
    HIPCHECK(c, hipEventRecord(c->ev1, c->stream));
    c->lastLaunches = 1;
    return PDX_OK;
  }
  HIPCHECK(c, hipEventRecord(c->ev0, c->stream));
  HIPCHECK(c, launchSubstep(c->dInst, c->dOrder, (int)c->hInst.size(), nsub, c->block, c->lds,
                            c->prec == PDX_PREC_MIXED32, c->mode == PDX_SOLVE_FAST, c->idx16, c->ldsW, c->stream, 0, 0));
  HIPCHECK(c, hipEventRecord(c->ev1, c->stream));
  c->lastLaunches = 1;
  return PDX_OK;
}

// Host replay of an instance's packed block schedules (fast mode), in the order
// and with the visibility the kernel has: one right-hand side, FP64, the device's
// step / run / task / segment structure.  Phases (a split instance): forward over
// the first half seeing only it, over the second half seeing only it, the
// separator forward and backward seeing everything, then backward over each half
// seeing it and the separator; one workgroup: forward and backward over all ranks.
// A read of a rank the phase does not hold, or of a value not final yet in the
// sweep's sense, or an index outside its array is an error (returned in `why`).
// relErr: max |x - x_direct| / max |x_direct| against a plain LDL^T solve.
bool replayLevels(const Instance& I, std::string& why, double& relErr, const std::vector<double>& b,
                  std::vector<double>& W);

bool replaySchedules(const Instance& I, std::string& why, double& relErr) {
  const Factor& f = I.fac;
  const int m = f.m;
  if (I.bf.levels) {
    std::vector<double> b((size_t)m), W;
    for (int k = 0; k < m; ++k) b[(size_t)k] = std::sin(0.37 * k + 1.0) + 0.01 * (k % 7);
    return replayLevels(I, why, relErr, b, W);
  }
  const int rA = I.paired ? f.top[0] : m, rB = I.paired ? f.top[0] + f.top[1] : m;
  std::vector<double> b((size_t)m), W((size_t)m);
  for (int k = 0; k < m; ++k) b[(size_t)k] = std::sin(0.37 * k + 1.0) + 0.01 * (k % 7);
  W = b;
  std::vector<int> state((size_t)m, 0);  // 0 right-hand side, 1 forward final, 2 backward final
  auto region = [&](int k) { return k < rA ? 0 : (k < rB ? 1 : 2); };
  auto phase = [&](const HostBlockSweep& h, bool fwd, int lo, int hi, bool seeA, bool seeB, bool seeS,
                   const char* name) -> bool {
    const int need = fwd ? 1 : 2;
    auto visible = [&](int k) {
      const int g = region(k);
      return g == 0 ? seeA : (g == 1 ? seeB : seeS);
    };
    auto fault = [&](const std::string& what) {
      why = std::string(name) + ": " + what;
      return false;
    };
    const size_t nrun = h.runCol.size() / kRun;
    if (h.stepTask.size() != (size_t)h.nstep + 1 || h.stepRun.size() != (size_t)h.nstep + 1 ||
        h.stepDense.size() != (size_t)h.nstep + 1)
      return fault("step tables of the wrong length");
    std::vector<int> done;
    for (int st = 0; st < h.nstep; ++st) {
      const int r0 = h.stepRun[(size_t)st], r1 = h.stepRun[(size_t)st + 1];
      if (r0 < 0 || r1 < r0 || (size_t)r1 > nrun) return fault("run range out of bounds");
      std::vector<double> part((size_t)(r1 - r0));
      for (int r = r0; r < r1; ++r) {
        double acc = 0.0;
        for (int q = 0; q < kRun; ++q) {
          const int col = h.runCol[(size_t)r * kRun + q];
          const double v = h.runVal[(size_t)r * kRun + q];
          if (col < 0 || col >= m) return fault("run column out of range");
          if (v == 0.0) continue;  // padding
          if (!visible(col)) return fault("reads rank " + std::to_string(col) + " of another workgroup");
          if (state[(size_t)col] != need) return fault("reads rank " + std::to_string(col) + " before it is final");
          acc += v * W[(size_t)col];
        }
        part[(size_t)(r - r0)] = acc;
      }
      const int t0 = h.stepTask[(size_t)st], t1 = h.stepTask[(size_t)st + 1];
      if (t0 < 0 || t1 < t0 || (size_t)t1 > h.taskInfo.size()) return fault("task range out of bounds");
      const int d0 = h.stepDense[(size_t)st];
      done.clear();
      for (int t = t0; t < t1; ++t) {
        double a[64];
        int kk[64], info[64], doff[64];
        for (int ln = 0; ln < 64; ++ln) {
          const size_t mi = ((size_t)t * 64 + ln) * 4;
          if (mi + 3 >= h.laneMeta.size()) return fault("lane metadata out of bounds");
          kk[ln] = h.laneMeta[mi];
          info[ln] = h.laneMeta[mi + 3];
          doff[ln] = h.laneDense[(size_t)t * 64 + ln];
          a[ln] = 0.0;
          const int k = kk[ln];
          if (k < 0) continue;
          if (k < lo || k >= hi) return fault("finishes rank " + std::to_string(k) + " outside the phase");
          if (state[(size_t)k] != need - 1) return fault("finishes rank " + std::to_string(k) + " twice");
          double x = W[(size_t)k];
          if (!fwd) x = x * f.dinv[(size_t)k];
          const int qb = h.laneMeta[mi + 1], qe = h.laneMeta[mi + 2];
          if (qb < 0 || qe < qb || qe > r1 - r0) return fault("lane's runs outside its step");
          for (int q = qb; q < qe; ++q) x -= part[(size_t)q];
          a[ln] = x;
        }
        double y[64];
        for (int ln = 0; ln < 64; ++ln) {
          y[ln] = a[ln];
          if (kk[ln] < 0) continue;
          const int L = info[ln] & 0x7f, o = (info[ln] >> 7) & 0x7f;
          if (L <= 1) continue;
          if (ln - o < 0 || ln - o + L > 64) return fault("segment outside its task");
          double acc = 0.0;
          for (int s2 = 0; s2 < L; ++s2) {
            const size_t di = (size_t)d0 + (size_t)doff[ln] + (size_t)s2 * L + o;
            if (di >= h.dense.size()) return fault("segment inverse out of bounds");
            acc += h.dense[di] * a[ln - o + s2];
          }
          y[ln] = acc;
        }
        for (int ln = 0; ln < 64; ++ln)
          if (kk[ln] >= 0) {
            W[(size_t)kk[ln]] = y[ln];
            done.push_back(kk[ln]);
          }
      }
      for (int k : done) state[(size_t)k] = need;  // final for the next step
    }
    for (int k = lo; k < hi; ++k)
      if (state[(size_t)k] != need) return fault("leaves rank " + std::to_string(k) + " unfinished");
    return true;
  };
  bool ok;
  if (I.paired) {
    ok = phase(I.bf, true, 0, rA, true, false, false, "forward, first half") &&
         phase(I.pf, true, rA, rB, false, true, false, "forward, second half") &&
         phase(I.sf, true, rB, m, true, true, true, "forward, separator") &&
         phase(I.sb, false, rB, m, true, true, true, "backward, separator") &&
         phase(I.bb, false, 0, rA, true, false, true, "backward, first half") &&
         phase(I.pb, false, rA, rB, false, true, true, "backward, second half");
  } else {
    ok = phase(I.bf, true, 0, m, true, true, true, "forward") && phase(I.bb, false, 0, m, true, true, true, "backward");
  }
  if (!ok) return false;
  // the direct solve: L y = b (rows ascending), y /= D, L^T x = y
  std::vector<double> x = b;
  for (int k = 0; k < m; ++k)
    for (int p = f.rp[(size_t)k]; p < f.rp[(size_t)k + 1]; ++p) x[(size_t)k] -= f.rx[(size_t)p] * x[(size_t)f.rc[(size_t)p]];
  for (int k = 0; k < m; ++k) x[(size_t)k] *= f.dinv[(size_t)k];
  for (int k = m - 1; k >= 0; --k)
    for (int p = f.cp[(size_t)k]; p < f.cp[(size_t)k + 1]; ++p) x[(size_t)k] -= f.cx[(size_t)p] * x[(size_t)f.ci[(size_t)p]];
  double num = 0.0, den = 0.0;
  for (int k = 0; k < m; ++k) {
    num = std::max(num, std::fabs(W[(size_t)k] - x[(size_t)k]));
    den = std::max(den, std::fabs(x[(size_t)k]));
  }
  relErr = den > 0.0 ? num / den : num;
  return true;
}

// Host replay of the grid form's level schedules (packLevels) in the kernel's
// order and visibility: per step, phase A tasks (their own runs, lanes, short
// chains' inverse rows, long chains' rows to R), then phase B (panels of the long
// chains' inverses applied to R).  A task may read only ranks final before its
// step; phase B only R written by this step's phase A; every index is checked
// against its array.  W returns the solution (rank space) of one right-hand
// side b; relErr as replaySchedules.
bool replayLevels(const Instance& I, std::string& why, double& relErr, const std::vector<double>& b,
                  std::vector<double>& W) {
  const Factor& f = I.fac;
  const int m = f.m;
  W = b;
  std::vector<int> state((size_t)m, 0);  // 0 right-hand side, 1 forward final, 2 backward final
  auto sweep = [&](const HostBlockSweep& h, bool fwd) -> bool {
    const char* name = fwd ? "forward" : "backward";
    auto fault = [&](const std::string& what) {
      why = std::string(name) + ": " + what;
      return false;
    };
    const int need = fwd ? 1 : 2;
    const size_t nrun = h.runCol.size() / kRun;
    const size_t ntask = h.taskInfo.size(), nb = h.bTask.size() / 4;
    if (h.stepTask.size() != (size_t)h.nstep + 1 || h.stepB.size() != (size_t)h.nstep + 1 ||
        h.stepE.size() != (size_t)h.nstep + 1 || h.eTask.size() % 8 != 0 ||
        h.stepDense.size() != (size_t)h.nstep + 1 || h.taskRun.size() != ntask + 1 ||
        h.laneMeta.size() != ntask * 256 || h.laneDense.size() != ntask * 64 || h.rTotal % 4 != 0)
      return fault("tables of the wrong length");
    std::vector<double> R((size_t)h.rTotal, 0.0);
    std::vector<int> rStep((size_t)h.rTotal, -1);  // step that wrote each R position
    std::vector<int> done;
    for (int st = 0; st < h.nstep; ++st) {
      const int t0 = h.stepTask[(size_t)st], t1 = h.stepTask[(size_t)st + 1];
      const int j0 = h.stepB[(size_t)st], j1 = h.stepB[(size_t)st + 1];
      if (t0 < 0 || t1 < t0 || (size_t)t1 > ntask || j0 < 0 || j1 < j0 || (size_t)j1 > nb)
        return fault("task range out of bounds");
      const int d0 = h.stepDense[(size_t)st];
      done.clear();
      for (int t = t0; t < t1; ++t) {
        const int r0 = h.taskRun[(size_t)t], r1 = h.taskRun[(size_t)t + 1];
        if (r0 < 0 || r1 < r0 || (size_t)r1 > nrun) return fault("task runs out of bounds");
        std::vector<double> part((size_t)(r1 - r0));
        for (int r = r0; r < r1; ++r) {
          double acc = 0.0;
          for (int q = 0; q < kRun; ++q) {
            const int col = h.runCol[(size_t)r * kRun + q];
            const double v = h.runVal[(size_t)r * kRun + q];
            if (col < 0 || col >= m) return fault("run column out of range");
            if (v == 0.0) continue;  // padding
            if (state[(size_t)col] != need) return fault("reads rank " + std::to_string(col) + " before it is final");
            acc += v * W[(size_t)col];
          }
          part[(size_t)(r - r0)] = acc;
        }
        double a[64], y[64];
        int kk[64], info[64], doff[64];
        for (int ln = 0; ln < 64; ++ln) {
          const size_t mi = ((size_t)t * 64 + ln) * 4;
          kk[ln] = h.laneMeta[mi];
          info[ln] = h.laneMeta[mi + 3];
          doff[ln] = h.laneDense[(size_t)t * 64 + ln];
          a[ln] = 0.0;
          const int k = kk[ln];
          if (k < 0) continue;
          if (k >= m) return fault("lane rank out of range");
          if (state[(size_t)k] != need - 1) return fault("finishes rank " + std::to_string(k) + " twice");
          double x = W[(size_t)k];
          if (!fwd) x = x * f.dinv[(size_t)k];
          const int qb = h.laneMeta[mi + 1], qe = h.laneMeta[mi + 2];
          if (qb < 0 || qe < qb || qe > r1 - r0) return fault("lane's runs outside its task");
          for (int q = qb; q < qe; ++q) x -= part[(size_t)q];
          a[ln] = x;
        }
        for (int ln = 0; ln < 64; ++ln) {
          y[ln] = a[ln];
          if (kk[ln] < 0) continue;
          const int L = info[ln] & 0x7f, o = (info[ln] >> 7) & 0x7f;
          if ((info[ln] >> 28) & 1) {
            if (L != 1) return fault("a long chain's row in a segment");
            if (doff[ln] < 0 || doff[ln] >= h.rTotal) return fault("R position out of range");
            if (rStep[(size_t)doff[ln]] >= 0) return fault("R position written twice");
            continue;
          }
          if (L <= 1) continue;
          if (ln - o < 0 || ln - o + L > 64) return fault("segment outside its task");
          double acc = 0.0;
          for (int s2 = 0; s2 < L; ++s2) {
            const size_t di = (size_t)d0 + (size_t)doff[ln] + (size_t)s2 * L + o;
            if (di >= h.dense.size()) return fault("segment inverse out of bounds");
            acc += h.dense[di] * a[ln - o + s2];
          }
          y[ln] = acc;
        }
        for (int ln = 0; ln < 64; ++ln) {
          if (kk[ln] < 0) continue;
          if ((info[ln] >> 28) & 1) {
            R[(size_t)doff[ln]] = a[ln];
            rStep[(size_t)doff[ln]] = st;
          } else {
            W[(size_t)kk[ln]] = y[ln];
            done.push_back(kk[ln]);
          }
        }
      }
      // phase A's dense tasks: r = base - E W[U] for 8 rows of a dense chain
      const int e0 = h.stepE[(size_t)st], e1 = h.stepE[(size_t)st + 1];
      if (e0 < 0 || e1 < e0 || (size_t)e1 * 8 > h.eTask.size()) return fault("dense task range out of bounds");
      for (int j = e0; j < e1; ++j) {
        const int* et = h.eTask.data() + (size_t)j * 8;
        const int eoff = et[0], c = et[1], uoff = et[2], k0 = et[3] & 0xffffff, nrows = et[3] >> 24, rp = et[4];
        if (nrows < 1 || nrows > 8 || c < 1 || eoff < 0 || eoff % 8 || (size_t)eoff + (size_t)c * 8 > h.epanel.size() ||
            uoff < 0 || (size_t)uoff + (size_t)c > h.ucol.size() || rp < 0 || rp + nrows > h.rTotal || k0 + nrows > m)
          return fault("dense task out of bounds");
        for (int q = 0; q < nrows; ++q) {
          const int k = k0 + q;
          if (state[(size_t)k] != need - 1) return fault("dense task on rank " + std::to_string(k) + " finished");
          double x = W[(size_t)k];
          if (!fwd) x = x * f.dinv[(size_t)k];
          double acc = 0.0;
          for (int u = 0; u < c; ++u) {
            const double ev = h.epanel[(size_t)eoff + (size_t)u * 8 + q];
            const int col = h.ucol[(size_t)uoff + u];
            if (col < 0 || col >= m) return fault("union column out of range");
            if (ev == 0.0) continue;
            if (state[(size_t)col] != need) return fault("dense task reads rank " + std::to_string(col) + " before it is final");
            acc += ev * W[(size_t)col];
          }
          if (rStep[(size_t)(rp + q)] >= 0) return fault("R position written twice");
          R[(size_t)(rp + q)] = x - acc;
          rStep[(size_t)(rp + q)] = st;
        }
      }
      for (int j = j0; j < j1; ++j) {
        const int xoff = h.bTask[(size_t)j * 4], ld = h.bTask[(size_t)j * 4 + 1], rp = h.bTask[(size_t)j * 4 + 2];
        const int k0 = h.bTask[(size_t)j * 4 + 3] & 0xffffff, nrows = h.bTask[(size_t)j * 4 + 3] >> 24;
        if (nrows < 1 || nrows > 8 || ld < 4 || ld % 4 || xoff < 0 || xoff % 4 ||
            (size_t)xoff + (size_t)nrows * ld > h.xpanel.size() || rp < 0 || rp % 4 || rp + ld > h.rTotal ||
            k0 + nrows > m)
          return fault("block product out of bounds");
        for (int q = 0; q < nrows; ++q) {
          const int k = k0 + q;
          if (state[(size_t)k] != need - 1) return fault("block product finishes rank " + std::to_string(k) + " twice");
          double acc = 0.0;
          for (int u = 0; u < ld; ++u) {
            const double xv = h.xpanel[(size_t)xoff + (size_t)q * ld + u];
            if (xv == 0.0) continue;
            if (rStep[(size_t)(rp + u)] != st) return fault("block product reads R not written in its step");
            acc += xv * R[(size_t)(rp + u)];
          }
          W[(size_t)k] = acc;
          done.push_back(k);
        }
      }
      for (int k : done) {
        if (state[(size_t)k] != need - 1) return fault("rank " + std::to_string(k) + " finished twice in a step");
        state[(size_t)k] = need;
      }
    }
    for (int k = 0; k < m; ++k)
      if (state[(size_t)k] != need) return fault("leaves rank " + std::to_string(k) + " unfinished");
    return true;
  };
  if (!sweep(I.bf, true) || !sweep(I.bb, false)) return false;
  std::vector<double> x = b;
  for (int k = 0; k < m; ++k)
    for (int p = f.rp[(size_t)k]; p < f.rp[(size_t)k + 1]; ++p) x[(size_t)k] -= f.rx[(size_t)p] * x[(size_t)f.rc[(size_t)p]];
  for (int k = 0; k < m; ++k) x[(size_t)k] *= f.dinv[(size_t)k];
  for (int k = m - 1; k >= 0; --k)
    for (int p = f.cp[(size_t)k]; p < f.cp[(size_t)k + 1]; ++p) x[(size_t)k] -= f.cx[(size_t)p] * x[(size_t)f.ci[(size_t)p]];
  double num = 0.0, den = 0.0;
  for (int k = 0; k < m; ++k) {
    num = std::max(num, std::fabs(W[(size_t)k] - x[(size_t)k]));
    den = std::max(den, std::fabs(x[(size_t)k]));
  }
  relErr = den > 0.0 ? num / den : num;
  return true;
}

// After a launch of split instances: did a pair stop waiting for its partner?
int checkPairs(pdx_ctx* c) {
  if ((c->npairs == 0 && c->launchMode != 1) || !c->uploaded || !c->dGiveUp) return PDX_OK;
  int gaveUp = 0;
  HIPCHECK(c, hipMemcpy(&gaveUp, c->dGiveUp, sizeof(int), hipMemcpyDeviceToHost));
  if (!gaveUp) return PDX_OK;
  // the pairs' counts no longer agree: the next step re-uploads (counts zero)
  c->dirty = true;
  return fail(c, PDX_EHIP, "workgroups stopped waiting for each other (a split instance's pair or the "
                           "grid form's sweep steps): the launch's results are incomplete");
}

}  // namespace

extern "C" {

int pdx_create(pdx_ctx** out, int device, int precision, int solve_mode) {
  if (!out) return PDX_EINVAL;
  *out = nullptr;
  if (precision != PDX_PREC_FP64 && precision != PDX_PREC_MIXED32) return PDX_EINVAL;
  if (solve_mode != PDX_SOLVE_EXACT && solve_mode != PDX_SOLVE_FAST && solve_mode != PDX_SOLVE_DENSE)
    return PDX_EINVAL;
  pdx_ctx* c = new pdx_ctx();
  c->device = device;
  c->prec = precision;
  c->mode = solve_mode;
  *out = c;
  return PDX_OK;
}

int pdx_destroy(pdx_ctx* c) {
  if (!c) return PDX_EINVAL;
  if (c->stream) {
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
  }
  if (c->dbuf) hipFree(c->dbuf);
  if (c->dLinv) hipFree(c->dLinv);
  if (c->roll.dbuf) hipFree(c->roll.dbuf);
  if (c->dScratchB) hipFree(c->dScratchB);
  if (c->dScratchX) hipFree(c->dScratchX);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return PDX_OK;
}

const char* pdx_last_error(const pdx_ctx* c) { return c ? c->err.c_str() : "null context"; }

int pdx_add_instance(pdx_ctx* c, const pdx_scene* s, const pdx_config* cfg, int* instance) {
  if (!c || !cfg) return PDX_EINVAL;
  if (validateScene(s) != PDX_OK) return fail(c, PDX_EINVAL, "invalid scene");
  if (cfg->dt <= 0.0 || cfg->max_iterations < 0) return fail(c, PDX_EINVAL, "invalid config");
  if (c->uploaded) {
    const int rc = pullState(c);
    if (rc) return rc;
  }
  Instance I;
  const int n = s->n_vertices;
  Scene& S = I.scene;
  S.n = n;
  S.E = s->n_edges;
  S.B = s->n_bends;
  S.x.assign(s->positions, s->positions + 3 * (size_t)n);
  S.v = s->velocities ? std::vector<double>(s->velocities, s->velocities + 3 * (size_t)n)
                      : std::vector<double>(3 * (size_t)n, 0.0);
  S.mass.assign(s->masses, s->masses + n);
  S.pinned.assign(s->pinned, s->pinned + n);
  for (auto& p : S.pinned) p = p ? 1 : 0;
  S.pinpos = s->pin_positions ? std::vector<double>(s->pin_positions, s->pin_positions + 3 * (size_t)n) : S.x;
  S.ea.assign(s->edge_a, s->edge_a + S.E);
  S.eb.assign(s->edge_b, s->edge_b + S.E);
  S.rest.assign(s->edge_rest, s->edge_rest + S.E);
  S.kappa.assign(s->edge_kappa, s->edge_kappa + S.E);
  if (S.B > 0) {
    S.sa.assign(s->bend_a, s->bend_a + S.B);
    S.sb.assign(s->bend_b, s->bend_b + S.B);
    S.sc.assign(s->bend_c, s->bend_c + S.B);
    S.bk.assign(s->bend_k, s->bend_k + S.B);
  }
  fillConfig(I.cfg, cfg);
  c->inst.push_back(std::move(I));
  c->dirty = true;
  if (instance) *instance = (int)c->inst.size() - 1;
  return PDX_OK;
}

int pdx_instance_count(const pdx_ctx* c) { return c ? (int)c->inst.size() : 0; }

int pdx_factor(pdx_ctx* c) {
  if (!c) return PDX_EINVAL;
  return ensureReady(c);
}

// Stepping after a differentiable rollout: instances may split over two
// workgroups again (pdx_set_split).
static int leaveRollMode(pdx_ctx* c) {
  if (!c->rollMode) return PDX_OK;
  c->rollMode = false;
  if (c->splitMin > 0) {
    const int rc = pullState(c);
    if (rc) return rc;
    for (auto& I : c->inst) I.packed = false;
    c->dirty = true;
  }
  return PDX_OK;
}

int pdx_step_async(pdx_ctx* c, int nsub) {
  if (!c || nsub < 0) return PDX_EINVAL;
  int rc = leaveRollMode(c);
  if (rc) return rc;
  rc = ensureReady(c);
  if (rc) return rc;
  if (nsub == 0) return PDX_OK;
  return launch(c, nsub);
}

int pdx_sync(pdx_ctx* c) {
  if (!c) return PDX_EINVAL;
  if (!c->stream) return PDX_OK;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  float ms = 0.f;
  if (c->lastLaunches > 0 && hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess)
    c->lastKernelMs = ms / c->lastLaunches;
  return checkPairs(c);
}

int pdx_step(pdx_ctx* c, int nsub) {
  int rc = pdx_step_async(c, nsub);
  if (rc) return rc;
  return pdx_sync(c);
}

int pdx_step_graph(pdx_ctx* c, int nsub, int replays) {
  if (!c || nsub <= 0 || replays < 0) return PDX_EINVAL;
  int rc = leaveRollMode(c);
  if (rc) return rc;
  rc = ensureReady(c);
  if (rc) return rc;
  if (cooperative(c)) {
    // cooperative launches are replayed directly (not captured)
    HIPCHECK(c, hipEventRecord(c->ev0, c->stream));
    for (int r = 0; r < replays; ++r)
      if ((rc = launchCooperative(c, nsub))) return rc;
    HIPCHECK(c, hipEventRecord(c->ev1, c->stream));
    c->lastLaunches = std::max(1, replays);
    return pdx_sync(c);
  }
  // the graph and its executable are destroyed on every exit path
  struct GraphGuard {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    ~GraphGuard() {
      if (ge) hipGraphExecDestroy(ge);
      if (g) hipGraphDestroy(g);
    }
  } gg;
  HIPCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const hipError_t le = launchSubstep(c->dInst, c->dOrder, (int)c->hInst.size(), nsub, c->block, c->lds,
                                      c->prec == PDX_PREC_MIXED32, c->mode == PDX_SOLVE_FAST, c->idx16, c->ldsW, c->stream,
                                      0, 0);
  const hipError_t ce = hipStreamEndCapture(c->stream, &gg.g);
  if (le != hipSuccess) return fail(c, PDX_EHIP, std::string("capture: ") + hipGetErrorString(le));
  if (ce != hipSuccess) return fail(c, PDX_EHIP, std::string("end capture: ") + hipGetErrorString(ce));
  HIPCHECK(c, hipGraphInstantiate(&gg.ge, gg.g, nullptr, nullptr, 0));
  HIPCHECK(c, hipEventRecord(c->ev0, c->stream));
  for (int r = 0; r < replays; ++r) HIPCHECK(c, hipGraphLaunch(gg.ge, c->stream));
  HIPCHECK(c, hipEventRecord(c->ev1, c->stream));
  c->lastLaunches = std::max(1, replays);
  return pdx_sync(c);
}

int pdx_get_state(pdx_ctx* c, int i, double* x, double* v) {
  if (!c || i < 0 || i >= (int)c->inst.size()) return PDX_EINVAL;
  Instance& I = c->inst[(size_t)i];
  if (!c->uploaded) {
    if (x) std::memcpy(x, I.scene.x.data(), I.scene.x.size() * 8);
    if (v) std::memcpy(v, I.scene.v.data(), I.scene.v.size() * 8);
    return PDX_OK;
  }
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (x) HIPCHECK(c, hipMemcpy(x, I.dx, I.scene.x.size() * 8, hipMemcpyDeviceToHost));
  if (v) HIPCHECK(c, hipMemcpy(v, I.dv, I.scene.v.size() * 8, hipMemcpyDeviceToHost));
  return PDX_OK;
}

int pdx_set_state(pdx_ctx* c, int i, const double* x, const double* v) {
  if (!c || i < 0 || i >= (int)c->inst.size()) return PDX_EINVAL;
  Instance& I = c->inst[(size_t)i];
  if (x) std::memcpy(I.scene.x.data(), x, I.scene.x.size() * 8);
  if (v) std::memcpy(I.scene.v.data(), v, I.scene.v.size() * 8);
  if (c->uploaded) {
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (x) HIPCHECK(c, hipMemcpy(I.dx, x, I.scene.x.size() * 8, hipMemcpyHostToDevice));
    if (v) HIPCHECK(c, hipMemcpy(I.dv, v, I.scene.v.size() * 8, hipMemcpyHostToDevice));
  }
  return PDX_OK;
}

int pdx_set_pin_positions(pdx_ctx* c, int i, const double* q) {
  if (!c || !q || i < 0 || i >= (int)c->inst.size()) return PDX_EINVAL;
  Instance& I = c->inst[(size_t)i];
  std::memcpy(I.scene.pinpos.data(), q, I.scene.pinpos.size() * 8);
  if (I.scene.B > 0) bendingRhs(I.scene, I.brhs);
  if (c->uploaded) {
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(I.dpin, q, I.scene.pinpos.size() * 8, hipMemcpyHostToDevice));
    if (I.scene.B > 0)
      HIPCHECK(c, hipMemcpy(I.dbrhs, I.brhs.data(), I.brhs.size() * 8, hipMemcpyHostToDevice));
  }
  return PDX_OK;
}

int pdx_set_edge_kappa(pdx_ctx* c, int i, const double* kappa) {
  if (!c || !kappa || i < 0 || i >= (int)c->inst.size()) return PDX_EINVAL;
  Instance& I = c->inst[(size_t)i];
  if (std::memcmp(I.scene.kappa.data(), kappa, I.scene.kappa.size() * 8) == 0) return PDX_OK;  // unchanged
  const int rc = pullState(c);
  if (rc) return rc;
  std::memcpy(I.scene.kappa.data(), kappa, I.scene.kappa.size() * 8);
  if (numericStamp(I.scene, I.cfg.dt) != I.numStamp) c->dirty = true;
  return PDX_OK;
}

int pdx_set_bend_k(pdx_ctx* c, int i, const double* k) {
  if (!c || !k || i < 0 || i >= (int)c->inst.size()) return PDX_EINVAL;
  Instance& I = c->inst[(size_t)i];
  if (std::memcmp(I.scene.bk.data(), k, I.scene.bk.size() * 8) == 0) return PDX_OK;  // unchanged
  const int rc = pullState(c);
  if (rc) return rc;
  std::memcpy(I.scene.bk.data(), k, I.scene.bk.size() * 8);
  if (numericStamp(I.scene, I.cfg.dt) != I.numStamp) c->dirty = true;
  return PDX_OK;
}

int pdx_set_masses(pdx_ctx* c, int i, const double* masses) {
  if (!c || !masses || i < 0 || i >= (int)c->inst.size()) return PDX_EINVAL;
  Instance& I = c->inst[(size_t)i];
  for (int v = 0; v < I.scene.n; ++v)
    if (!(masses[v] >= 0.0) || !std::isfinite(masses[v])) return fail(c, PDX_EINVAL, "invalid mass");
  if (std::memcmp(I.scene.mass.data(), masses, I.scene.mass.size() * 8) == 0) return PDX_OK;
  const int rc = pullState(c);
  if (rc) return rc;
  std::memcpy(I.scene.mass.data(), masses, I.scene.mass.size() * 8);
  // masses enter L = M/h^2 + ...: the numeric stamp changes, the next step refactors
  if (numericStamp(I.scene, I.cfg.dt) != I.numStamp) c->dirty = true;
  return PDX_OK;
}

int pdx_set_pinned(pdx_ctx* c, int i, const uint8_t* pinned) {
  if (!c || !pinned || i < 0 || i >= (int)c->inst.size()) return PDX_EINVAL;
  Instance& I = c->inst[(size_t)i];
  bool same = true;
  for (int v = 0; v < I.scene.n && same; ++v) same = (pinned[v] != 0) == (I.scene.pinned[(size_t)v] != 0);
  if (same) return PDX_OK;
  const int rc = pullState(c);
  if (rc) return rc;
  const int before = pdx_pin_count(c, i);
  for (int v = 0; v < I.scene.n; ++v) I.scene.pinned[(size_t)v] = pinned[v] ? 1 : 0;
  // the mask decides which rows are overwritten and which coupling blocks exist:
  // both stamps change, so the next step analyses and factors again; the
  // incidence and bend records carry the pinned bits and are rebuilt with the upload
  if (I.scene.B > 0) bendingRhs(I.scene, I.brhs);
  if (pdx_pin_count(c, i) != before && i < (int)c->roll.traj.size()) {
    c->roll.traj[(size_t)i].clear();  // a pin trajectory no longer matches the pins
    c->roll.trajSub[(size_t)i] = 0;
  }
  c->dirty = true;
  return PDX_OK;
}

int pdx_set_config(pdx_ctx* c, int i, const pdx_config* cfg) {
  if (!c || !cfg || i < 0 || i >= (int)c->inst.size()) return PDX_EINVAL;
  if (cfg->dt <= 0.0 || cfg->max_iterations < 0) return fail(c, PDX_EINVAL, "invalid config");
  Config next = c->inst[(size_t)i].cfg;
  fillConfig(next, cfg);
  const Config& cur = c->inst[(size_t)i].cfg;
  const bool same = next.dt == cur.dt && next.g[0] == cur.g[0] && next.g[1] == cur.g[1] && next.g[2] == cur.g[2] &&
                    next.damping == cur.damping && next.maxIter == cur.maxIter && next.relTol == cur.relTol &&
                    next.absTol == cur.absTol && next.resTol == cur.resTol && next.chebRho == cur.chebRho &&
                    next.chebStart == cur.chebStart;
  if (same) return PDX_OK;  // unchanged: nothing to upload
  const int rc = pullState(c);
  if (rc) return rc;
  c->inst[(size_t)i].cfg = next;
  c->dirty = true;  // damping/tolerances re-upload; only a dt change refactors (stamp)
  return PDX_OK;
}

int pdx_get_stats(pdx_ctx* c, int i, pdx_stats* s) {
  if (!c || !s || i < 0 || i >= (int)c->inst.size()) return PDX_EINVAL;
  Instance& I = c->inst[(size_t)i];
  std::memset(s, 0, sizeof(*s));
  s->factorize_count = I.factorizeCount;
  s->analyze_count = I.analyzeCount;
  s->color_count = I.topo.colorCount;
  s->factor_nnz = I.fac.nnz;
  s->fwd_levels = I.fac.fwd.nlev;
  s->bwd_levels = I.fac.bwd.nlev;
  s->fwd_steps = c->mode == PDX_SOLVE_FAST ? I.bf.nstep : I.hf.nlev;
  s->bwd_steps = c->mode == PDX_SOLVE_FAST ? I.bb.nstep : I.hb.nlev;
  s->analyze_seconds = I.analyzeSec;
  s->factorize_seconds = I.factorSec;
  s->last_residual = -1.0;
  if (c->uploaded) {
    int ist[4];
    long long lst[16];
    double dst[2];
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(ist, I.distats, 16, hipMemcpyDeviceToHost));
    HIPCHECK(c, hipMemcpy(lst, I.dlstats, 128, hipMemcpyDeviceToHost));
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) != hipSuccess || khz <= 0)
      khz = 100000;
    for (int q = 0; q < 9; ++q) s->phase_seconds[q] = (double)lst[2 + q] / (khz * 1e3);
    for (int q = 0; q < 4; ++q) s->workgroup_seconds[q] = (double)lst[11 + q] / (khz * 1e3);
    HIPCHECK(c, hipMemcpy(dst, I.ddstats, 16, hipMemcpyDeviceToHost));
    s->iterations_used = ist[0];
    s->converged = ist[1];
    s->early_exit_count = ist[2];
    s->total_iterations = lst[0];
    s->substeps = lst[1];
    s->last_residual = dst[0];
  }
  return PDX_OK;
}

}  // extern "C"

namespace {

// One global solve of instance i on the context's stream: b (3n, host) staged in
// dB, the solution written to dX and copied back to x (host).
int solveThrough(pdx_ctx* c, int i, const double* b, double* x, double* dB, double* dX) {
  Instance& I = c->inst[(size_t)i];
  const int n3 = 3 * I.scene.n;
  HIPCHECK(c, hipMemcpyAsync(dB, b, (size_t)n3 * 8, hipMemcpyHostToDevice, c->stream));
  if (c->mode == PDX_SOLVE_DENSE) {
    HIPCHECK(c, launchDenseSolve(c->dInst, i, dB, dX, I.scene.n, I.fac.m, c->prec == PDX_PREC_MIXED32, c->stream));
  } else if (c->launchMode == 1) {
    if (c->lds < gridLdsNeed(c))
      return fail(c, PDX_EINVAL, "grid form: the launch's LDS does not hold a slice per wavefront");
    const int g = gridLaunchBlocks(c);
    HIPCHECK(c, launchGridSolve(c->dInst, i, dB, dX, c->waveRuns, g, c->block, c->lds,
                                c->prec == PDX_PREC_MIXED32, c->idx16, c->stream));
  } else {
    // a split instance solves on its two workgroups (the partner's descriptor
    // follows the instances' in pair order)
    int partner = -1;
    if (I.paired) {
      int p = 0;
      for (int j = 0; j < i; ++j) p += c->inst[(size_t)j].paired ? 1 : 0;
      partner = (int)c->inst.size() + p;
    }
    HIPCHECK(c, launchSolve(c->dInst, i, partner, dB, dX, c->block, c->lds, c->prec == PDX_PREC_MIXED32,
                            c->mode == PDX_SOLVE_FAST, c->idx16, c->ldsW, c->stream));
  }
  HIPCHECK(c, hipMemcpyAsync(x, dX, (size_t)n3 * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return checkPairs(c);
}

}  // namespace

extern "C" {

int pdx_solve3(pdx_ctx* c, int i, const double* b, double* x) {
  if (!c || !b || !x || i < 0 || i >= (int)c->inst.size()) return PDX_EINVAL;
  int rc = ensureReady(c);
  if (rc) return rc;
  const int n3 = 3 * c->inst[(size_t)i].scene.n;
  if (c->scratchN < n3) {
    if (c->dScratchB) hipFree(c->dScratchB);
    if (c->dScratchX) hipFree(c->dScratchX);
    c->dScratchB = c->dScratchX = nullptr;
    c->scratchN = 0;
    HIPCHECK(c, hipMalloc(&c->dScratchB, (size_t)n3 * 8));
    HIPCHECK(c, hipMalloc(&c->dScratchX, (size_t)n3 * 8));
    c->scratchN = n3;
  }
  return solveThrough(c, i, b, x, c->dScratchB, c->dScratchX);
}

int pdx_solve3_noalloc(pdx_ctx* c, int i, const double* b, double* x) {
  if (!c || !b || !x || i < 0 || i >= (int)c->inst.size()) return PDX_EINVAL;
  // no upload here (it would allocate): the context must be factored and current
  if (!c->uploaded || c->dirty)
    return fail(c, PDX_ENOTREADY, "pdx_solve3_noalloc: factor the context first (pdx_factor or a step); "
                                  "it allocates nothing, so it does not upload");
  const size_t n3 = (size_t)3 * c->inst[(size_t)i].scene.n;
  if (n3 > noallocDofs())
    return fail(c, PDX_EINVAL, "pdx_solve3_noalloc: at most " + std::to_string(noallocDofs() / 3) +
                                   " vertices (its static staging); use pdx_solve3");
  double* s = nullptr;
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, noallocScratch(c->device, &s));
  return solveThrough(c, i, b, x, s, s + noallocDofs());
}

int pdx_get_factor(pdx_ctx* c, int i, int* m, long long* nnz, int32_t* ptr, int32_t* rows, double* vals,
                   double* dinv, int32_t* rank_vertex, int32_t* perm3) {
  if (!c || i < 0 || i >= (int)c->inst.size()) return PDX_EINVAL;
  Instance& I = c->inst[(size_t)i];
  if (!I.factored || !I.packed || numericStamp(I.scene, I.cfg.dt) != I.numStamp) {
    // host-side refresh under the stamp rules (the device copy follows at the next step)
    bool changed = false;
    const int rc = refreshFactors(c, changed);
    if (rc) return rc;
  }
  const Factor& f = I.fac;
  if (m) *m = f.m;
  if (nnz) *nnz = f.nnz;
  if (ptr) std::copy(f.cp.begin(), f.cp.end(), ptr);
  if (rows) std::copy(f.ci.begin(), f.ci.end(), rows);
  if (vals) std::copy(f.cx.begin(), f.cx.end(), vals);
  if (dinv) std::copy(f.dinv.begin(), f.dinv.end(), dinv);
  if (rank_vertex) std::copy(f.rankVertex[0].begin(), f.rankVertex[0].end(), rank_vertex);
  if (perm3) std::copy(f.perm3.begin(), f.perm3.end(), perm3);
  return PDX_OK;
}

}  // extern "C"

namespace {

// Pin lists of one instance for the rollout: pinned vertices (ascending), edges
// with exactly one pinned end {edge, pin index}, stencils holding a pinned vertex
// {stencil, pin index or -1 per vertex}.
void pinLists(const Scene& s, std::vector<int>& pins, std::vector<int>& pe, std::vector<int>& ps) {
  std::vector<int> idx((size_t)s.n, -1);
  pins.clear();
  pe.clear();
  ps.clear();
  for (int v = 0; v < s.n; ++v)
    if (s.pinned[(size_t)v]) {
      idx[(size_t)v] = (int)pins.size();
      pins.push_back(v);
    }
  for (int e = 0; e < s.E; ++e) {
    const int a = s.ea[(size_t)e], b = s.eb[(size_t)e];
    const bool ap = s.pinned[(size_t)a] != 0, bp = s.pinned[(size_t)b] != 0;
    if (ap == bp) continue;
    pe.push_back(e);
    pe.push_back(ap ? idx[(size_t)a] : idx[(size_t)b]);
  }
  for (int t = 0; t < s.B; ++t) {
    const int vv[3] = {s.sa[(size_t)t], s.sb[(size_t)t], s.sc[(size_t)t]};
    if (!s.pinned[(size_t)vv[0]] && !s.pinned[(size_t)vv[1]] && !s.pinned[(size_t)vv[2]]) continue;
    ps.push_back(t);
    for (int k = 0; k < 3; ++k) ps.push_back(idx[(size_t)vv[k]]);
  }
}

size_t alignUp(size_t b) { return (b + 255) / 256 * 256; }

int rollLaunch(pdx_ctx* c, int nsub, int subBase, int flags) {
  HIPCHECK(c, launchSubstep(c->dInst, c->dOrder, (int)c->inst.size(), nsub, c->block, c->lds,
                            c->prec == PDX_PREC_MIXED32, c->mode == PDX_SOLVE_FAST, c->idx16, c->ldsW, c->stream,
                            subBase, flags));
  return PDX_OK;
}

int rollCheckpoint(pdx_ctx* c, int seg, bool save) {
  for (size_t i = 0; i < c->inst.size(); ++i) {
    Instance& I = c->inst[i];
    const size_t b3 = I.scene.x.size() * 8;
    char* ck = c->roll.dbuf + c->roll.ckpt[i] + (size_t)seg * 2 * b3;
    if (save) {
      HIPCHECK(c, hipMemcpyAsync(ck, I.dx, b3, hipMemcpyDeviceToDevice, c->stream));
      HIPCHECK(c, hipMemcpyAsync(ck + b3, I.dv, b3, hipMemcpyDeviceToDevice, c->stream));
    } else {
      HIPCHECK(c, hipMemcpyAsync(I.dx, ck, b3, hipMemcpyDeviceToDevice, c->stream));
      HIPCHECK(c, hipMemcpyAsync(I.dv, ck + b3, b3, hipMemcpyDeviceToDevice, c->stream));
    }
  }
  return PDX_OK;
}

}  // namespace

extern "C" {

int pdx_pin_count(const pdx_ctx* c, int i) {
  if (!c || i < 0 || i >= (int)c->inst.size()) return PDX_EINVAL;
  int np = 0;
  for (uint8_t p : c->inst[(size_t)i].scene.pinned) np += p ? 1 : 0;
  return np;
}

int pdx_set_pin_trajectory(pdx_ctx* c, int i, int nsub, const double* traj) {
  if (!c || i < 0 || i >= (int)c->inst.size() || nsub < 0) return PDX_EINVAL;
  c->roll.traj.resize(c->inst.size());
  c->roll.trajSub.resize(c->inst.size(), 0);
  if (!traj || nsub == 0) {
    c->roll.traj[(size_t)i].clear();
    c->roll.trajSub[(size_t)i] = 0;
    return PDX_OK;
  }
  const size_t np = (size_t)pdx_pin_count(c, i);
  c->roll.traj[(size_t)i].assign(traj, traj + (size_t)nsub * np * 3);
  c->roll.trajSub[(size_t)i] = nsub;
  return PDX_OK;
}

int pdx_rollout(pdx_ctx* c, int nsub, int every) {
  if (!c || nsub <= 0 || every <= 0) return PDX_EINVAL;
  if (c->launchMode == 1 || c->mode == PDX_SOLVE_DENSE)
    return fail(c, PDX_EUNSUPPORTED, "the differentiable rollout runs in the batch form (exact or fast solve)");
  for (auto& I : c->inst)
    if (I.cfg.chebRho > 0.0)
      return fail(c, PDX_EUNSUPPORTED, "the differentiable rollout differentiates the plain iteration (no Chebyshev blend)");
  if (!c->rollMode) {
    // the differentiable rollout runs every instance on one workgroup
    c->rollMode = true;
    bool anyPaired = false;
    for (auto& I : c->inst) anyPaired = anyPaired || I.paired;
    if (anyPaired) {
      const int prc = pullState(c);
      if (prc) return prc;
      for (auto& I : c->inst) I.packed = false;
      c->dirty = true;
    }
  }
  int rc = ensureReady(c);
  if (rc) return rc;
  if (c->hInst.size() != c->inst.size()) return fail(c, PDX_EUNSUPPORTED, "rollout: an instance is still split");
  const size_t ni = c->inst.size();
  c->roll.traj.resize(ni);
  c->roll.trajSub.resize(ni, 0);
  for (size_t i = 0; i < ni; ++i)
    if (!c->roll.traj[i].empty() && c->roll.trajSub[i] != nsub)
      return fail(c, PDX_EINVAL, "pin trajectory of instance " + std::to_string(i) + " holds " +
                                     std::to_string(c->roll.trajSub[i]) + " substeps, rollout asks " +
                                     std::to_string(nsub));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (c->roll.dbuf) {
    HIPCHECK(c, hipFree(c->roll.dbuf));
    c->roll.dbuf = nullptr;
  }
  auto& R = c->roll;
  R.nsub = nsub;
  R.every = every;
  R.nseg = (nsub + every - 1) / every;
  R.forwardDone = false;
  R.ckpt.assign(ni, 0);
  R.seedX.resize(ni);
  R.seedV.resize(ni);
  // layout: small per-instance tables are staged on the host, big buffers are device-only
  struct RO {
    size_t pins, pe, ps, traj, tape, xbar, vbar, cur, bbar, xnbar, gk, gb, gp;
    std::vector<int> lpins, lpe, lps;
    std::vector<double> ltraj;
  };
  std::vector<RO> ro(ni);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    off = alignUp(off);
    const size_t o = off;
    off += std::max<size_t>(bytes, 8);
    return o;
  };
  for (size_t i = 0; i < ni; ++i) {
    Instance& I = c->inst[i];
    RO& o = ro[i];
    pinLists(I.scene, o.lpins, o.lpe, o.lps);
    const size_t np = o.lpins.size(), n3 = (size_t)3 * I.scene.n;
    if (R.traj[i].empty()) {
      o.ltraj.resize((size_t)nsub * np * 3);
      for (int st = 0; st < nsub; ++st)
        for (size_t j = 0; j < np; ++j)
          for (int d = 0; d < 3; ++d)
            o.ltraj[((size_t)st * np + j) * 3 + d] = I.scene.pinpos[(size_t)3 * o.lpins[j] + d];
    } else {
      o.ltraj = R.traj[i];
    }
    o.pins = take(np * 4);
    o.pe = take(o.lpe.size() * 4);
    o.ps = take(o.lps.size() * 4);
    o.traj = take(o.ltraj.size() * 8);
    R.ckpt[i] = take((size_t)(R.nseg + 1) * 2 * n3 * 8);
    o.tape = take((size_t)every * (I.cfg.maxIter + 1) * n3 * 8);
    o.xbar = take(n3 * 8);
    o.vbar = take(n3 * 8);
    o.cur = take(n3 * 8);
    o.bbar = take(n3 * 8);
    o.xnbar = take(n3 * 8);
    o.gk = take((size_t)I.scene.E * 8);
    o.gb = take((size_t)I.scene.B * 8);
    o.gp = take((size_t)nsub * np * 3 * 8);
  }
  const size_t total = alignUp(off);
  if (hipMalloc(&R.dbuf, total) != hipSuccess) {
    R.dbuf = nullptr;
    return fail(c, PDX_ENOMEM, "rollout buffers: " + std::to_string(total) + " bytes");
  }
  HIPCHECK(c, hipMemset(R.dbuf, 0, total));
  char* B = R.dbuf;
  for (size_t i = 0; i < ni; ++i) {
    RO& o = ro[i];
    InstDev& d = c->hInst[i];
    auto put = [&](size_t at, const void* src, size_t bytes) -> int {
      if (bytes) HIPCHECK(c, hipMemcpy(B + at, src, bytes, hipMemcpyHostToDevice));
      return PDX_OK;
    };
    if ((rc = put(o.pins, o.lpins.data(), o.lpins.size() * 4))) return rc;
    if ((rc = put(o.pe, o.lpe.data(), o.lpe.size() * 4))) return rc;
    if ((rc = put(o.ps, o.lps.data(), o.lps.size() * 4))) return rc;
    if ((rc = put(o.traj, o.ltraj.data(), o.ltraj.size() * 8))) return rc;
    d.npin = (int)o.lpins.size();
    d.npinEdge = (int)(o.lpe.size() / 2);
    d.npinSten = (int)(o.lps.size() / 4);
    d.rollK = c->inst[i].cfg.maxIter;
    d.pinList = (const int*)(B + o.pins);
    d.pinEdge = (const int*)(B + o.pe);
    d.pinSten = (const int*)(B + o.ps);
    d.pinTraj = (const double*)(B + o.traj);
    d.tape = (double*)(B + o.tape);
    d.xbar = (double*)(B + o.xbar);
    d.vbar = (double*)(B + o.vbar);
    d.cur = (double*)(B + o.cur);
    d.bbar = (double*)(B + o.bbar);
    d.xnbar = (double*)(B + o.xnbar);
    d.gkap = (double*)(B + o.gk);
    d.gbend = (double*)(B + o.gb);
    d.gpin = (double*)(B + o.gp);
  }
  HIPCHECK(c, hipMemcpy(c->dInst, c->hInst.data(), sizeof(InstDev) * ni, hipMemcpyHostToDevice));
  R.epoch = c->uploadEpoch;
  // forward, fixed iterations, checkpoint (x, v) at the start of every segment
  HIPCHECK(c, hipEventRecord(c->ev0, c->stream));
  for (int seg = 0; seg < R.nseg; ++seg) {
    const int s0 = seg * every, len = std::min(every, nsub - s0);
    if ((rc = rollCheckpoint(c, seg, true))) return rc;
    if ((rc = rollLaunch(c, len, s0, kRollTraj | kRollFixed))) return rc;
  }
  if ((rc = rollCheckpoint(c, R.nseg, true))) return rc;
  HIPCHECK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HIPCHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  R.forwardMs = ms;
  // the pins stay where the trajectory left them
  for (auto& I : c->inst) {
    HIPCHECK(c, hipMemcpy(I.scene.pinpos.data(), I.dpin, I.scene.pinpos.size() * 8, hipMemcpyDeviceToHost));
    if (I.scene.B > 0) bendingRhs(I.scene, I.brhs);
  }
  R.forwardDone = true;
  return PDX_OK;
}

int pdx_set_adjoint_seed(pdx_ctx* c, int i, const double* dx, const double* dv) {
  if (!c || i < 0 || i >= (int)c->inst.size()) return PDX_EINVAL;
  const size_t n3 = c->inst[(size_t)i].scene.x.size();
  c->roll.seedX.resize(c->inst.size());
  c->roll.seedV.resize(c->inst.size());
  c->roll.seedX[(size_t)i] = dx ? std::vector<double>(dx, dx + n3) : std::vector<double>();
  c->roll.seedV[(size_t)i] = dv ? std::vector<double>(dv, dv + n3) : std::vector<double>();
  return PDX_OK;
}

int pdx_adjoint(pdx_ctx* c) {
  if (!c) return PDX_EINVAL;
  auto& R = c->roll;
  if (!R.forwardDone || R.epoch != c->uploadEpoch || !R.dbuf)
    return fail(c, PDX_ENOTREADY, "no rollout to differentiate (pdx_rollout first; the context changed since?)");
  const size_t ni = c->inst.size();
  R.seedX.resize(ni);
  R.seedV.resize(ni);
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < ni; ++i) {
    const InstDev& d = c->hInst[i];
    const size_t b3 = c->inst[i].scene.x.size() * 8;
    if (R.seedX[i].empty()) HIPCHECK(c, hipMemset(d.xbar, 0, b3));
    else HIPCHECK(c, hipMemcpy(d.xbar, R.seedX[i].data(), b3, hipMemcpyHostToDevice));
    if (R.seedV[i].empty()) HIPCHECK(c, hipMemset(d.vbar, 0, b3));
    else HIPCHECK(c, hipMemcpy(d.vbar, R.seedV[i].data(), b3, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemset(d.gkap, 0, std::max<size_t>((size_t)d.E * 8, 8)));
    if (d.B > 0) HIPCHECK(c, hipMemset(d.gbend, 0, (size_t)d.B * 8));
    if (d.npin > 0) HIPCHECK(c, hipMemset(d.gpin, 0, (size_t)R.nsub * d.npin * 3 * 8));
  }
  int rc = 0;
  HIPCHECK(c, hipEventRecord(c->ev0, c->stream));
  for (int seg = R.nseg - 1; seg >= 0; --seg) {
    const int s0 = seg * R.every, len = std::min(R.every, R.nsub - s0);
    if ((rc = rollCheckpoint(c, seg, false))) return rc;
    if ((rc = rollLaunch(c, len, s0, kRollTraj | kRollTape | kRollFixed))) return rc;
    HIPCHECK(c, launchAdjoint(c->dInst, c->dOrder, (int)ni, len, s0, c->block, c->lds,
                              c->prec == PDX_PREC_MIXED32, c->mode == PDX_SOLVE_FAST, c->idx16, c->ldsW, c->stream));
  }
  HIPCHECK(c, hipEventRecord(c->ev1, c->stream));
  // back to the state after the rollout, pins where the trajectory ended
  if ((rc = rollCheckpoint(c, R.nseg, false))) return rc;
  for (auto& I : c->inst) {
    HIPCHECK(c, hipMemcpyAsync(I.dpin, I.scene.pinpos.data(), I.scene.pinpos.size() * 8, hipMemcpyHostToDevice,
                               c->stream));
    if (I.scene.B > 0)
      HIPCHECK(c, hipMemcpyAsync(I.dbrhs, I.brhs.data(), I.brhs.size() * 8, hipMemcpyHostToDevice, c->stream));
  }
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HIPCHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  R.adjointMs = ms;
  return PDX_OK;
}

int pdx_get_gradients(pdx_ctx* c, int i, double* g_x0, double* g_v0, double* g_kappa, double* g_bend,
                      double* g_pin) {
  if (!c || i < 0 || i >= (int)c->inst.size()) return PDX_EINVAL;
  if (!c->roll.dbuf || c->roll.epoch != c->uploadEpoch) return fail(c, PDX_ENOTREADY, "no adjoint run");
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  const InstDev& d = c->hInst[(size_t)i];
  const size_t b3 = (size_t)3 * d.n * 8;
  if (g_x0) HIPCHECK(c, hipMemcpy(g_x0, d.xbar, b3, hipMemcpyDeviceToHost));
  if (g_v0) HIPCHECK(c, hipMemcpy(g_v0, d.vbar, b3, hipMemcpyDeviceToHost));
  if (g_kappa && d.E > 0) HIPCHECK(c, hipMemcpy(g_kappa, d.gkap, (size_t)d.E * 8, hipMemcpyDeviceToHost));
  if (g_bend && d.B > 0) HIPCHECK(c, hipMemcpy(g_bend, d.gbend, (size_t)d.B * 8, hipMemcpyDeviceToHost));
  if (g_pin && d.npin > 0)
    HIPCHECK(c, hipMemcpy(g_pin, d.gpin, (size_t)c->roll.nsub * d.npin * 3 * 8, hipMemcpyDeviceToHost));
  return PDX_OK;
}

int pdx_rollout_timing(const pdx_ctx* c, double* forward_ms, double* adjoint_ms) {
  if (!c) return PDX_EINVAL;
  if (forward_ms) *forward_ms = c->roll.forwardMs;
  if (adjoint_ms) *adjoint_ms = c->roll.adjointMs;
  return PDX_OK;
}

}  // extern "C"

extern "C" {

int pdx_diagnostics(pdx_ctx* c, int i, pdx_diag* out) {
  if (!c || !out || i < 0 || i >= (int)c->inst.size()) return PDX_EINVAL;
  Instance& I = c->inst[(size_t)i];
  std::vector<double> x(I.scene.x.size()), v(I.scene.v.size());
  const int rc = pdx_get_state(c, i, x.data(), v.data());
  if (rc) return rc;
  double d[6];
  physicsDiagnostics(I.scene, x.data(), v.data(), I.cfg.g, d);
  out->max_strain = d[0];
  out->mean_strain = d[1];
  out->elastic_energy = d[2];
  out->total_energy = d[3];
  out->max_speed = d[4];
  out->total_mass = d[5];
  return PDX_OK;
}

int pdx_matrix_nnz(const pdx_ctx* c, int i, long long* nnz) {
  if (!c || !nnz || i < 0 || i >= (int)c->inst.size()) return PDX_EINVAL;
  const Instance& I = c->inst[(size_t)i];
  if (!I.factored) return PDX_ENOTREADY;
  *nnz = 3 * (long long)I.A.rows.size();  // every block is a scalar times I3
  return PDX_OK;
}

int pdx_get_topology(pdx_ctx* c, int i, int* color_count, uint32_t* color_of, uint32_t* order,
                     uint32_t* start, uint32_t* vertex_start, uint32_t* vertex_edges,
                     uint32_t* bend_vertex_start, uint32_t* bend_stencils) {
  if (!c || i < 0 || i >= (int)c->inst.size()) return PDX_EINVAL;
  Instance& I = c->inst[(size_t)i];
  if (!I.topoBuilt) {
    buildTopology(I.scene, I.topo);
    I.topoBuilt = true;
  }
  const Topology& t = I.topo;
  if (color_count) *color_count = t.colorCount;
  auto put = [](const std::vector<uint32_t>& a, uint32_t* out) {
    if (out) std::copy(a.begin(), a.end(), out);
  };
  put(t.colorOf, color_of);
  put(t.corder, order);
  put(t.cstart, start);
  put(t.vstart, vertex_start);
  put(t.vedges, vertex_edges);
  put(t.bstart, bend_vertex_start);
  put(t.bstencil, bend_stencils);
  return PDX_OK;
}

int pdx_get_schedule_stats(pdx_ctx* c, int i, int64_t* out) {
  if (!c || !out || i < 0 || i >= (int)c->inst.size()) return PDX_EINVAL;
  Instance& I = c->inst[(size_t)i];
  if (!I.factored || !I.packed || numericStamp(I.scene, I.cfg.dt) != I.numStamp) {
    bool changed = false;
    const int rc = refreshFactors(c, changed);
    if (rc) return rc;
  }
  const HostBlockSweep* sw[2] = {&I.bf, &I.bb};
  for (int q = 0; q < 2; ++q) {
    const HostBlockSweep& h = *sw[q];
    int64_t chains = 0, chainRows = 0, maxTasks = 0;
    for (size_t t = 0; t < h.taskInfo.size(); ++t)
      if (h.taskInfo[t] >> 16) {
        ++chains;
        chainRows += h.taskInfo[t] & 0xffff;
      }
    for (int st = 0; st < h.nstep; ++st)
      maxTasks = std::max<int64_t>(maxTasks, h.stepTask[(size_t)st + 1] - h.stepTask[(size_t)st]);
    int64_t* o = out + 8 * q;
    o[0] = h.nstep;
    o[1] = (int64_t)(h.taskInfo.size() + h.eTask.size() / 8);
    o[2] = chains;
    o[3] = chainRows;
    o[4] = (int64_t)h.runCol.size() / kRun;
    int64_t maxRuns = 0;
    for (int st = 0; st < h.nstep; ++st)
      maxRuns = std::max<int64_t>(maxRuns, h.stepRun[(size_t)st + 1] - h.stepRun[(size_t)st]);
    o[5] = (int64_t)(h.dense.size() + h.xpanel.size() + h.epanel.size());
    o[6] = maxTasks;
    o[7] = maxRuns;
  }
  return PDX_OK;
}

int pdx_set_schedule(pdx_ctx* c, int block_threads, int chain_cap) {
  if (!c || block_threads < 64 || block_threads > 512 || block_threads % 64 != 0 || chain_cap < 1 ||
      chain_cap > 32)
    return PDX_EINVAL;
  if (c->mode == PDX_SOLVE_DENSE) {
    // no sweep schedule to repack: only the workgroup size applies
    c->block = block_threads;
    return PDX_OK;
  }
  if (c->chainCap != chain_cap) {
    const int rc = pullState(c);
    if (rc) return rc;
    c->chainCap = chain_cap;
    for (auto& I : c->inst) I.packed = false;  // repack the same factor
    c->dirty = true;
  }
  if (c->launchMode == 1 && c->block != block_threads) {
    // the grid form gives every wavefront its own LDS slice: the launch's LDS
    // follows the workgroup size (upload sizes it)
    const int rc = pullState(c);
    if (rc) return rc;
    c->dirty = true;
  }
  c->block = block_threads;
  return PDX_OK;
}

int pdx_set_split(pdx_ctx* c, long long min_factor_entries) {
  if (!c || min_factor_entries < 0) return PDX_EINVAL;
  if (min_factor_entries == c->splitMin) return PDX_OK;
  const int rc = pullState(c);
  if (rc) return rc;
  c->splitMin = min_factor_entries;
  for (auto& I : c->inst) I.packed = false;  // repack the same factors
  c->dirty = true;
  return PDX_OK;
}

int pdx_verify_schedule(pdx_ctx* c, int i, double* rel_err) {
  if (!c || i < 0 || i >= (int)c->inst.size()) return PDX_EINVAL;
  if (c->mode != PDX_SOLVE_FAST) return fail(c, PDX_EUNSUPPORTED, "only the fast mode has block schedules");
  Instance& I = c->inst[(size_t)i];
  if (!I.factored || !I.packed || numericStamp(I.scene, I.cfg.dt) != I.numStamp) {
    bool changed = false;
    const int rc = refreshFactors(c, changed);
    if (rc) return rc;
  }
  std::string why;
  double rel = 0.0;
  if (!replaySchedules(I, why, rel)) return fail(c, PDX_EUNSUPPORTED, "schedule replay: " + why);
  if (rel_err) *rel_err = rel;
  return PDX_OK;
}

int pdx_split_count(const pdx_ctx* c) {
  if (!c) return PDX_EINVAL;
  int n = 0;
  for (const auto& I : c->inst) n += I.paired ? 1 : 0;
  return n;
}

int pdx_set_ordering(pdx_ctx* c, int ordering) {
  if (!c || (ordering != PDX_ORDER_AMD && ordering != PDX_ORDER_ND)) return PDX_EINVAL;
  if (ordering == PDX_ORDER_ND && c->mode != PDX_SOLVE_FAST)
    return fail(c, PDX_EUNSUPPORTED, "the nested-dissection order needs the fast solve mode");
  if (ordering != c->ordering) {
    const int rc = pullState(c);
    if (rc) return rc;
    c->ordering = ordering;
    for (auto& I : c->inst) {
      I.analyzed = false;  // new symbolic analysis and factor
      I.factored = false;
      I.packed = false;
    }
    c->dirty = true;
  }
  return PDX_OK;
}

int pdx_set_launch(pdx_ctx* c, int mode, int grid_blocks) {
  if (!c || (mode != PDX_LAUNCH_BATCH && mode != PDX_LAUNCH_GRID) || grid_blocks < 0) return PDX_EINVAL;
  if (c->mode == PDX_SOLVE_DENSE) {
    // the dense solve mode always runs on the grid: only the workgroup count applies
    c->gridBlocks = grid_blocks;
    return PDX_OK;
  }
  if (mode == PDX_LAUNCH_GRID && (c->inst.size() != 1 || c->mode != PDX_SOLVE_FAST))
    return fail(c, PDX_EUNSUPPORTED, "the grid form takes exactly one instance in fast solve mode");
  if (mode != c->launchMode) {
    const int rc = pullState(c);
    if (rc) return rc;
    c->launchMode = mode;
    c->waveRuns = (c->prec == PDX_PREC_MIXED32) ? 1536 : 768;
    for (auto& I : c->inst) I.packed = false;
    c->dirty = true;
  }
  c->gridBlocks = grid_blocks;
  return PDX_OK;
}

int pdx_autotune(pdx_ctx* c, int probe_substeps, int* best_block, int* best_cap, double* best_ms) {
  if (!c || probe_substeps <= 0) return PDX_EINVAL;
  int rc = ensureReady(c);
  if (rc) return rc;
  // the probes must leave no trace: state and statistics are restored afterwards
  rc = pullState(c);
  if (rc) return rc;
  std::vector<std::vector<double>> xs, vs;
  std::vector<std::vector<int>> ist;
  std::vector<std::vector<long long>> lst;
  for (auto& I : c->inst) {
    xs.push_back(I.scene.x);
    vs.push_back(I.scene.v);
    std::vector<int> a(4);
    std::vector<long long> b(16);
    HIPCHECK(c, hipMemcpy(a.data(), I.distats, 16, hipMemcpyDeviceToHost));
    HIPCHECK(c, hipMemcpy(b.data(), I.dlstats, 128, hipMemcpyDeviceToHost));
    ist.push_back(a);
    lst.push_back(b);
  }
  auto restore = [&]() -> int {
    for (size_t i = 0; i < c->inst.size(); ++i) {
      Instance& I = c->inst[i];
      HIPCHECK(c, hipMemcpy(I.dx, xs[i].data(), xs[i].size() * 8, hipMemcpyHostToDevice));
      HIPCHECK(c, hipMemcpy(I.dv, vs[i].data(), vs[i].size() * 8, hipMemcpyHostToDevice));
      HIPCHECK(c, hipMemcpy(I.distats, ist[i].data(), 16, hipMemcpyHostToDevice));
      HIPCHECK(c, hipMemcpy(I.dlstats, lst[i].data(), 128, hipMemcpyHostToDevice));
      I.scene.x = xs[i];
      I.scene.v = vs[i];
    }
    return PDX_OK;
  };
  if (c->mode == PDX_SOLVE_DENSE) {
    // workgroup count of the cooperative launch: fewer workgroups make cheaper
    // grid barriers, more stream the inverse faster
    const int maxb = denseBlocksMax(c->block, c->lds, c->prec == PDX_PREC_MIXED32);
    const int keep = c->gridBlocks;
    double bestMs = 1e300;
    int best = keep;
    for (int g = 16; g <= 1024; g *= 2) {
      const int gb = std::min(g, maxb);
      c->gridBlocks = gb;
      if ((rc = restore())) return rc;
      if ((rc = launch(c, probe_substeps))) return rc;
      if ((rc = pdx_sync(c))) return rc;
      if (c->lastKernelMs < bestMs) {
        bestMs = c->lastKernelMs;
        best = gb;
      }
      if (gb == maxb) break;
    }
    c->gridBlocks = best;
    if ((rc = restore())) return rc;
    if (best_block) *best_block = best;
    if (best_cap) *best_cap = 0;
    if (best_ms) *best_ms = bestMs;
    return PDX_OK;
  }
  const bool fast = c->mode == PDX_SOLVE_FAST;
  // A single cloth in the fast mode: the fill order (the current one first, then
  // the other) and the launch form (one workgroup at the current shape, or the
  // grid with 16 .. 256 workgroups of 256 or 512 threads) are probed; the shape
  // search below runs in the form chosen.  A change of order re-analyses and refactors the cloth
  // (factorize_count and analyze_count count it).
  if (fast && c->inst.size() == 1 && !c->rollMode && c->inst[0].cfg.chebRho <= 0.0) {
    double formMs = 1e300;
    int bo = c->ordering, bm = c->launchMode, bgb = c->gridBlocks, bblk = c->block;
    const int blk0 = c->block;
    const int orders[2] = {c->ordering, c->ordering == PDX_ORDER_AMD ? PDX_ORDER_ND : PDX_ORDER_AMD};
    for (int ord : orders) {
      // the state before the probes: a new analysis orders by these positions
      if (c->uploaded && (rc = restore())) return rc;
      if ((rc = pdx_set_ordering(c, ord))) return rc;
      for (int form : {PDX_LAUNCH_BATCH, PDX_LAUNCH_GRID}) {
        // the grid form at both workgroup sizes (the best count depends on the
        // size: 512 x 512 ties at 128 and 256 workgroups of 512 threads, and 256
        // of 256 threads beat both); the batch form once, at the current shape
        bool fits = true;
        for (int blk : {256, 512}) {
          const int useBlk = form == PDX_LAUNCH_BATCH ? blk0 : blk;
          for (int g = 16; g <= 256 && fits; g *= 2) {
            const int gb = form == PDX_LAUNCH_GRID ? g : 0;
            if (c->uploaded && (rc = restore())) return rc;
            if ((rc = pdx_set_launch(c, form, gb))) return rc;
            if ((rc = pdx_set_schedule(c, useBlk, c->chainCap))) return rc;
            if ((rc = ensureReady(c))) {
              if (rc != PDX_EUNSUPPORTED) return rc;
              fits = false;  // this form does not fit the cloth
              break;
            }
            if (form == PDX_LAUNCH_GRID && gb > gridBlocksMax(c->block, c->lds, c->prec == PDX_PREC_MIXED32, c->idx16))
              break;
            // one launch per probe, and a second one for a probe within 1.5 x of
            // the best so far, keeping the faster: the grid sizes near the best
            // differ by less than one launch's spread on large cloths
            double ms = 1e300;
            for (int rep = 0; rep < 2; ++rep) {
              if ((rc = restore())) return rc;
              if ((rc = launch(c, probe_substeps))) return rc;
              if ((rc = pdx_sync(c))) return rc;
              ms = std::min(ms, (double)c->lastKernelMs);
              if (ms > 1.5 * formMs) break;
            }
            if (ms < formMs) {
              formMs = ms;
              bo = ord;
              bm = form;
              bgb = gb;
              bblk = useBlk;
            }
            if (form == PDX_LAUNCH_BATCH) break;
          }
          if (form == PDX_LAUNCH_BATCH || !fits) break;
        }
      }
    }
    if ((rc = restore())) return rc;
    if ((rc = pdx_set_ordering(c, bo))) return rc;
    if ((rc = pdx_set_launch(c, bm, bgb))) return rc;
    if ((rc = pdx_set_schedule(c, bblk, c->chainCap))) return rc;
    if ((rc = ensureReady(c))) return rc;
  }
  const int blocks[2] = {256, 512};
  const int caps[4] = {8, 16, 24, 32};
  double bestMs = 1e300;
  int bb = c->block, bc = c->chainCap;
  for (int ci = 0; ci < (fast ? 4 : 1); ++ci) {
    const int cap = fast ? caps[ci] : c->chainCap;
    for (int bi = 0; bi < 2; ++bi) {
      if ((rc = pdx_set_schedule(c, blocks[bi], cap))) return rc;
      if ((rc = ensureReady(c))) {
        if (rc == PDX_EUNSUPPORTED) continue;  // this schedule's staging does not fit in LDS
        return rc;
      }
      if ((rc = restore())) return rc;
      if ((rc = launch(c, probe_substeps))) return rc;
      if ((rc = pdx_sync(c))) return rc;
      if (c->lastKernelMs < bestMs) {
        bestMs = c->lastKernelMs;
        bb = blocks[bi];
        bc = cap;
      }
    }
  }
  if ((rc = pdx_set_schedule(c, bb, bc))) return rc;
  if ((rc = ensureReady(c))) return rc;
  // Split instances (pdx_set_split), where the context allows it: thresholds at
  // the factor sizes of the 8th, 16th, ... largest instances that can split, and
  // none; the shape found above is kept.
  if (fast && c->prec == PDX_PREC_MIXED32 && c->launchMode == 0 && !c->rollMode) {
    std::vector<long long> sizes;
    for (const auto& I : c->inst)
      if (I.fac.top[0] > 0 && I.fac.top[1] > 0 && I.cfg.chebRho <= 0.0) sizes.push_back(I.fac.nnz);
    std::sort(sizes.rbegin(), sizes.rend());
    if (!sizes.empty()) {
      std::vector<long long> cand{c->splitMin};
      if (c->splitMin != 0) cand.push_back(0);
      for (size_t k : {1, 2, 4, 8, 16, 24, 32, 48, 64})
        if (k <= sizes.size()) cand.push_back(sizes[k - 1]);
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
      long long bestSplit = c->splitMin;
      double splitMs = 1e300;
      for (long long th : cand) {
        if ((rc = pdx_set_split(c, th))) return rc;
        if ((rc = ensureReady(c))) {
          if (rc == PDX_EUNSUPPORTED) continue;
          return rc;
        }
        if ((rc = restore())) return rc;
        if ((rc = launch(c, probe_substeps))) return rc;
        if ((rc = pdx_sync(c))) return rc;
        if (c->lastKernelMs < splitMs) {
          splitMs = c->lastKernelMs;
          bestSplit = th;
        }
      }
      if ((rc = pdx_set_split(c, bestSplit))) return rc;
      if ((rc = ensureReady(c))) return rc;
      bestMs = std::min(bestMs, splitMs);
    }
  }
  if ((rc = restore())) return rc;
  if (best_block) *best_block = bb;
  if (best_cap) *best_cap = bc;
  if (best_ms) *best_ms = bestMs;
  return PDX_OK;
}

long long pdx_get_split(const pdx_ctx* c) { return c ? c->splitMin : -1; }

int pdx_get_launch(const pdx_ctx* c, int* mode, int* grid_blocks, int* ordering) {
  if (!c) return PDX_EINVAL;
  if (mode) *mode = c->launchMode;
  if (grid_blocks) *grid_blocks = c->gridBlocks;
  if (ordering) *ordering = c->ordering;
  return PDX_OK;
}

double pdx_last_kernel_ms(const pdx_ctx* c) { return c ? c->lastKernelMs : 0.0; }

int pdx_device_info(int device, int* cus, int* clock_khz, int* mem_clock_khz, int* bus_bits) {
  const struct {
    hipDeviceAttribute_t a;
    int* out;
  } q[4] = {{hipDeviceAttributeMultiprocessorCount, cus},
            {hipDeviceAttributeClockRate, clock_khz},
            {hipDeviceAttributeMemoryClockRate, mem_clock_khz},
            {hipDeviceAttributeMemoryBusWidth, bus_bits}};
  for (const auto& e : q) {
    if (!e.out) continue;
    if (hipDeviceGetAttribute(e.out, e.a, device) != hipSuccess) return PDX_EHIP;
  }
  return PDX_OK;
}
double pdx_bytes_per_iteration(const pdx_ctx* c) { return c ? c->bytesIter : 0.0; }

double pdx_bytes_per_substep(const pdx_ctx* c) {
  if (!c) return 0.0;
  double b = 0.0;
  for (const auto& I : c->inst) b += I.bytesPerIter * I.cfg.maxIter + I.bytesPerSubstepFixed;
  return b;
}

}  // extern "C"

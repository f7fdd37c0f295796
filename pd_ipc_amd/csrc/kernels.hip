// Persistent Projective-Dynamics substep kernel for gfx950.
//
// One workgroup advances one cloth instance through `nsub` substeps without
// leaving the kernel: predict -> K x [right-hand side gather, forward sweep,
// backward sweep, unpack + residual + criterion] -> velocity/damping
// (core/sim/Integrator.cpp:312-797).  The workgroup's waves share the instance;
// the 3-component solution vector W (rank space, 3 right-hand sides on one
// factor) lives in LDS when it fits, otherwise in a per-instance global scratch.
//
// Two arithmetic forms:
//  * exact (FP64):  b is gathered per vertex in the colour order of the reference
//    scatter (DistanceTerm.cpp:279-317), the sweeps run each row's dot product in
//    the reference's column order (EigenDirectSolver.cpp:418-446), the residual
//    is the reference gather (Integrator.cpp:230-263).  Same bits as the CPU path.
//  * mixed (FP32 factor):  each PD iteration solves for the update
//    delta = L^{-1}(b(x) - L x) in FP32 and adds it to the FP64 state; b(x) - L x
//    and the residual are gathered in FP64 from the same projections.
//
// All FP64 code is compiled with -ffp-contract=off: no operation is fused.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <mutex>

#include "devtypes.h"

namespace pdx {

namespace {

constexpr int kWave = 64;

// Timers (device wall clock, 100 MHz).  The stage buckets of a substep (predict
// and velocity, gather, forward sweep, backward sweep, criterion) are always
// kept, as the reference keeps its stage times: thread 0 reads the clock a few
// times per iteration, at points the workgroup has reached together.  The
// per-step timers inside the sweeps read it twice per step, and reading the
// clock waits on the scalar-memory counter, which the LDS traffic shares, so
// they are compiled in only on request (PDX_PHASE_TIMERS=1).
#ifndef PDX_PHASE_TIMERS
#define PDX_PHASE_TIMERS 0
#endif
__device__ inline long long tick() {
#if PDX_PHASE_TIMERS
  return wall_clock64();
#else
  return 0;
#endif
}
__device__ inline long long tickStage() { return wall_clock64(); }
constexpr double kMinLength = 1e-12;  // DistanceTerm.h:31

__device__ inline double waveMax(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
  return v;
}

// Max over the workgroup; `red` holds one slot per wave.  The caller guarantees
// a barrier between two uses of the same slots.
__device__ inline double blockMax(double v, double* red) {
  v = waveMax(v);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double r = 0.0;
  const int nw = blockDim.x / kWave;
  for (int w = 0; w < nw; ++w) r = fmax(r, red[w]);
  return r;
}

// Address-space-qualified pointers: LDS accesses through them compile to ds_*
// (tracked by the LDS counter only) and global ones to global_* (vector-memory
// counter only).  Generic pointers would compile to flat accesses, which count on
// both and make every LDS gather wait for the factor prefetch in flight.
#define PDX_GLOBAL __attribute__((address_space(1)))
// The sweeps and gathers are inlined into their kernels: called as functions
// they keep caller registers across the call boundary and spill them.
#define PDX_INLINE __attribute__((always_inline))
#define PDX_LDS __attribute__((address_space(3)))

template <typename Real>
__device__ inline PDX_LDS Real* ldsAt(int off) {
  extern __shared__ double smem[];
  return (PDX_LDS Real*)((PDX_LDS char*)smem + off);
}

// Barrier that orders LDS only.  __syncthreads() also waits for every
// outstanding global access of the calling wave (loads and stores share one
// counter), which would drain the factor prefetch issued just before it.
__device__ inline void ldsBarrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <typename Real>
__device__ inline Real subMul(Real acc, Real l, Real w);
template <>
__device__ inline double subMul<double>(double acc, double l, double w) {
  return acc - l * w;  // two roundings, as the reference
}
template <>
__device__ inline float subMul<float>(float acc, float l, float w) {
  return __builtin_fmaf(-l, w, acc);
}

__device__ inline double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ inline float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

constexpr int kU = 8;  // entries per load group of the level sweep

template <bool kFwd, typename WP, typename Store>
__device__ PDX_INLINE void longRow(const SweepDev& sw, WP W, const double* dinv, int rbase, int ebase, int lane,
                                   Store store);

// One level-scheduled sweep.  Forward: row k gathers L(k,j) y_j over ascending j.
// Backward: column k gathers L(i,k) x_i over ascending i after scaling by 1/D_k.
// `store(k, a0, a1, a2)` is called once per row with its final value.
template <typename Real, bool kFwd, typename WP, typename Store>
__device__ PDX_INLINE void sweep(const SweepDev& sw, WP W, const Real* dinv, Store store) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  const Real* val = static_cast<const Real*>(sw.val);
  for (int l = 0; l < sw.nlev; ++l) {
    const int c1 = sw.levChunk[l + 1];
    for (int ch = sw.levChunk[l] + wave; ch < c1; ch += nw) {
      int nrows = sw.chunkRows[ch];
      const int rbase = sw.chunkRowBase[ch];
      const int ebase = sw.chunkEntBase[ch];
      if (nrows < 0) {
        // one long row taken by the whole wavefront (FP64 schedules only)
        if constexpr (sizeof(Real) == 8) {
          longRow<kFwd>(sw, W, dinv, rbase, ebase, lane, store);
          continue;
        } else {
          nrows = 1;
        }
      }
      int k = 0, len = 0;
      if (lane < nrows) {
        k = sw.slotRank[rbase + lane];
        len = sw.slotLen[rbase + lane];
      }
      const int maxlen = __shfl(len, 0, kWave);
      Real a0 = 0, a1 = 0, a2 = 0;
      if (lane < nrows) {
        a0 = W[3 * k + 0];
        a1 = W[3 * k + 1];
        a2 = W[3 * k + 2];
        if (!kFwd) {
          const Real dv = dinv[k];
          a0 = a0 * dv;
          a1 = a1 * dv;
          a2 = a2 * dv;
        }
      }
      // Entries are taken kU at a time: every load of a group (column, value,
      // the three W values) is issued before the group's subtractions, which
      // then run in ascending entry order as before.
      int off = ebase;
      for (int j = 0; j < maxlen; j += kU) {
        int cc[kU];
        Real lv[kU], w0[kU], w1[kU], w2[kU];
        bool ok[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          ok[u] = (j + u) < len;
          const unsigned long long msk = __ballot(ok[u]);
          const int e = ok[u] ? off + lane : ebase;  // ebase holds an entry whenever maxlen > 0
          off += __popcll(msk);
          cc[u] = sw.col[e];
          lv[u] = val[e];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          w0[u] = W[3 * cc[u] + 0];
          w1[u] = W[3 * cc[u] + 1];
          w2[u] = W[3 * cc[u] + 2];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
          if (ok[u]) {
            a0 = subMul<Real>(a0, lv[u], w0[u]);
            a1 = subMul<Real>(a1, lv[u], w1[u]);
            a2 = subMul<Real>(a2, lv[u], w2[u]);
          }
      }
      if (lane < nrows) {
        W[3 * k + 0] = a0;
        W[3 * k + 1] = a1;
        W[3 * k + 2] = a2;
        store(k, a0, a1, a2);
      }
    }
    __syncthreads();
  }
}

__device__ inline double readlaneT(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline float readlaneT(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// One row of an FP64 level sweep taken by the whole wavefront, with the
// reference's operations in the reference's order: each lane forms the products
// l * w of its entries (each rounded, as the reference rounds its product), then
// the wavefront subtracts them from the row's value one by one in entry order
// (readlane broadcasts), 64 entries at a time.  Row entries are stored in row
// order (a one-row chunk).
template <bool kFwd, typename WP, typename Store>
__device__ PDX_INLINE void longRow(const SweepDev& sw, WP W, const double* dinv, int rbase, int ebase, int lane,
                                   Store store) {
  const int k = sw.slotRank[rbase], len = sw.slotLen[rbase];
  const double* val = static_cast<const double*>(sw.val);
  double a0 = W[3 * k], a1 = W[3 * k + 1], a2 = W[3 * k + 2];
  if (!kFwd) {
    const double dv = dinv[k];
    a0 = a0 * dv;
    a1 = a1 * dv;
    a2 = a2 * dv;
  }
  for (int j0 = 0; j0 < len; j0 += kWave) {
    const int j = j0 + lane;
    const int e = ebase + (j < len ? j : 0);  // lanes past the row reload its first entry
    const int c = sw.col[e];
    const double lv = val[e];
    const double p0 = lv * W[3 * c], p1 = lv * W[3 * c + 1], p2 = lv * W[3 * c + 2];
    const int cnt = min(kWave, len - j0);
    for (int i = 0; i < cnt; ++i) {
      a0 = a0 - readlaneT(p0, i);
      a1 = a1 - readlaneT(p1, i);
      a2 = a2 - readlaneT(p2, i);
    }
  }
  if (lane == 0) {
    W[3 * k + 0] = a0;
    W[3 * k + 1] = a1;
    W[3 * k + 2] = a2;
    store(k, a0, a1, a2);
  }
}

constexpr int kCap = 32;

// Finishes the lanes of one task: lane t holds row o of a segment (a chain of
// L <= kCap rows in lanes t-o .. t-o+L-1, or a single row, L = 1).  With the
// segment's stored inverse X (forward X, backward X^T; term s of the lane at
// Xs[s * L + o]):  y_t = sum over s < L, ascending, of X-term * r(segment row s).
// The broadcasts read the rows' values before the segment is finished, so no
// term waits for another.
// a -= P[q] for the row's runs q in [qb, qe), in ascending q; the partials are
// read four at a time so their LDS latencies overlap.
template <typename Real, typename PP>
__device__ inline void subtractPartials(PP P, int qb, int qe, Real& a0, Real& a1, Real& a2) {
  int q = qb;
  for (; q + 4 <= qe; q += 4) {
    Real p[4][3];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      p[u][0] = P[3 * (q + u) + 0];
      p[u][1] = P[3 * (q + u) + 1];
      p[u][2] = P[3 * (q + u) + 2];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a0 = a0 - p[u][0];
      a1 = a1 - p[u][1];
      a2 = a2 - p[u][2];
    }
  }
  for (; q < qe; ++q) {
    a0 = a0 - P[3 * q + 0];
    a1 = a1 - P[3 * q + 1];
    a2 = a2 - P[3 * q + 2];
  }
}

template <typename Real, typename DP>
__device__ inline void segmentFinish(DP Xs, int w, int lane, bool act, Real& a0, Real& a1, Real& a2) {
  const int L = w & 0x7f, o = (w >> 7) & 0x7f, maxL = (w >> 21) & 0x7f;
  if (maxL <= 1) return;  // single rows only (uniform over the task)
  const int base = lane - o;
  Real y0 = 0, y1 = 0, y2 = 0;
  // terms in groups (eight FP32, four FP64): a group's broadcasts and
  // coefficients are all issued before its sums, which then run in ascending q
  // as one term at a time would
  constexpr int kG = sizeof(Real) == 8 ? 4 : 8;
  for (int g = 0; g < maxL; g += kG) {
    Real r0[kG], r1[kG], r2[kG], c[kG];
#pragma unroll
    for (int u = 0; u < kG; ++u) {
      const int q = g + u;
      const int src = base + min(q, L - 1);
      r0[u] = __shfl(a0, src);
      r1[u] = __shfl(a1, src);
      r2[u] = __shfl(a2, src);
      c[u] = (q < L && L > 1 && act) ? Real(Xs[q * L + o]) : Real(L == 1 ? 1 : 0);
    }
#pragma unroll
    for (int u = 0; u < kG; ++u)
      if (g + u < L) {
        y0 = y0 + c[u] * r0[u];
        y1 = y1 + c[u] * r1[u];
        y2 = y2 + c[u] * r2[u];
      }
  }
  a0 = y0;
  a1 = y1;
  a2 = y2;
}
constexpr int kRun = 16;

typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

// One run of kRun factor entries as loaded: columns packed (16-bit: 2 x 16 B,
// 32-bit: 4 x 16 B), values FP32 (4 x 16 B) or FP64 (8 x 16 B); every load is
// independent, so a thread keeps several runs in flight.
template <typename Real, bool kIdx16>
struct RunRaw {
  v4i c[kIdx16 ? 2 : 4];
  Real v[kRun];
};

// Global-address-space views: loads through them compile to global_load, which
// only the vector-memory counter tracks, so an LDS barrier does not drain them
// (flat loads also count on the LDS counter).
template <typename T>
__device__ inline const PDX_GLOBAL T* gview(const void* p) {
  return (const PDX_GLOBAL T*)(p);
}

template <typename Real, bool kIdx16>
__device__ inline void loadRun(const void* colv, const Real* val, int r, RunRaw<Real, kIdx16>& b) {
  const PDX_GLOBAL v4i* c4 = kIdx16 ? gview<v4i>(static_cast<const uint16_t*>(colv) + (size_t)r * kRun)
                                    : gview<v4i>(static_cast<const int*>(colv) + (size_t)r * kRun);
#pragma unroll
  for (int q = 0; q < (kIdx16 ? 2 : 4); ++q) b.c[q] = c4[q];
  if (sizeof(Real) == 4) {
    const PDX_GLOBAL v4f* v4 = gview<v4f>(val + (size_t)r * kRun);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const v4f t = v4[q];
      b.v[4 * q] = t.x;
      b.v[4 * q + 1] = t.y;
      b.v[4 * q + 2] = t.z;
      b.v[4 * q + 3] = t.w;
    }
  } else {
    const PDX_GLOBAL v2d* v2 = gview<v2d>(val + (size_t)r * kRun);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const v2d t = v2[q];
      b.v[2 * q] = t.x;
      b.v[2 * q + 1] = t.y;
    }
  }
}

template <typename Real, bool kIdx16, typename WP, typename PP>
__device__ inline void runPartial(const RunRaw<Real, kIdx16>& b, WP W, PP P) {
  // fused multiply-adds into two accumulators (even and odd entries), joined at
  // the end: a fixed order, half the dependent chain of one accumulator
  Real p0[2] = {0, 0}, p1[2] = {0, 0}, p2[2] = {0, 0};
#pragma unroll
  for (int q = 0; q < kRun; ++q) {
    int c;
    if (kIdx16) {
      const int w = b.c[q >> 3][(q >> 1) & 3];
      c = (q & 1) ? ((unsigned)w >> 16) : (w & 0xffff);
    } else {
      c = b.c[q >> 2][q & 3];
    }
    const Real l = b.v[q];
    p0[q & 1] = fmaT(l, (Real)W[3 * c + 0], p0[q & 1]);
    p1[q & 1] = fmaT(l, (Real)W[3 * c + 1], p1[q & 1]);
    p2[q & 1] = fmaT(l, (Real)W[3 * c + 2], p2[q & 1]);
  }
  P[0] = p0[0] + p0[1];
  P[1] = p1[0] + p1[1];
  P[2] = p2[0] + p2[1];
}

constexpr int kPreRuns = 2;    // runs per thread loaded during the previous step
constexpr int kMaxChain = 32;  // longest chain block (the schedule's chain cap is at most 32)

// A lane's row of its segment's inverse block, X(o, q) for q < the task's longest
// segment, loaded from the step's dense range into registers.  Lanes without a
// segment of their own (single rows, lanes past the task) load the range's first
// value instead, which exists whenever the task holds a chain; their values are
// not used.  Tasks without chains load nothing.
template <typename Real>
__device__ inline void loadChainRow(const PDX_GLOBAL Real* Xs, v4i m, int doff, Real (&x)[kMaxChain]) {
  const int w = m.w;
  const int L = w & 0x7f, o = (w >> 7) & 0x7f;
  const int maxL = __builtin_amdgcn_readfirstlane((w >> 21) & 0x7f);
  if (maxL <= 1) return;
  const bool own = m.x >= 0 && L > 1;
  const PDX_GLOBAL Real* p = Xs + (own ? doff + o : 0);
  const int stride = own ? L : 0;
  const int last = own ? L - 1 : 0;
#pragma unroll
  for (int q = 0; q < kMaxChain; ++q)
    if (q < maxL) x[q] = p[min(q, last) * stride];
}

// segmentFinish with the lane's inverse-block row in registers (loadChainRow).
template <typename Real>
__device__ inline void segmentFinishR(const Real (&x)[kMaxChain], int w, int lane, Real& a0, Real& a1, Real& a2) {
  const int L = w & 0x7f, o = (w >> 7) & 0x7f;
  const int maxL = __builtin_amdgcn_readfirstlane((w >> 21) & 0x7f);
  if (maxL <= 1) return;  // single rows only
  const int base = lane - o;
  Real y0 = 0, y1 = 0, y2 = 0;
  constexpr int kG = sizeof(Real) == 8 ? 4 : 8;
#pragma unroll
  for (int g = 0; g < kMaxChain; g += kG) {
    if (g < maxL) {
      Real r0[kG], r1[kG], r2[kG];
#pragma unroll
      for (int u = 0; u < kG; ++u) {
        const int src = base + min(g + u, L - 1);
        r0[u] = __shfl(a0, src);
        r1[u] = __shfl(a1, src);
        r2[u] = __shfl(a2, src);
      }
#pragma unroll
      for (int u = 0; u < kG; ++u)
        if (g + u < L) {
          const Real c = L > 1 ? x[g + u] : Real(1);
          y0 = y0 + c * r0[u];
          y1 = y1 + c * r1[u];
          y2 = y2 + c * r2[u];
        }
    }
  }
  a0 = y0;
  a1 = y1;
  a2 = y2;
}

// Block-scheduled sweep.  Step st: (1) the workgroup turns the step's runs into
// partial sums in LDS (each thread's first kPreRuns runs and the first two tasks'
// lane metadata of its wavefront were loaded during the previous step); (2) each
// wavefront takes whole tasks: a row subtracts its partials in run order, a chain
// block then applies its segment's inverse block column by column, broadcasting
// each finished value with a shuffle.  Lane metadata loads two tasks ahead.  Two
// barriers per step; fixed order, no float atomics.  `steps` is the step table in
// LDS: run, dense and task offsets.
struct SweepTimes {
  long long stream = 0, wait1 = 0, fin = 0, wait2 = 0;
};

template <typename Real, bool kFwd, bool kIdx16, typename WP, typename Store>
__device__ PDX_INLINE void blockSweep(const BlockSweepDev& sw, WP W, PDX_LDS Real* P, PDX_LDS Real* D,
                                      const PDX_LDS int* steps, Store store, SweepTimes& tstream, bool wInLds) {
  // FP32: each lane's inverse-block row goes from HBM to registers, the step's
  // first task's across the previous step's barrier, later tasks' as they start
  // (a second register set loading one task ahead measured 1.3 % slower: the
  // kernel then holds ~240 registers).  FP64 (twice the registers per row): the
  // step's blocks are staged in LDS (D) during the run phase, as loaded during
  // the previous step.
  constexpr bool kRegRows = sizeof(Real) == 4;
  auto barrier = [wInLds]() {
    if (wInLds)
      ldsBarrier();
    else
      __syncthreads();
  };
  const int tid = threadIdx.x, bs = blockDim.x;
  const int lane = tid & (kWave - 1);
  const int wave = tid / kWave;
  const int nw = bs / kWave;
  const int ns = sw.nstep;
  if (ns <= 0) return;
  const PDX_LDS int* sRun = steps;
  const PDX_LDS int* sDense = steps + (ns + 1);
  const PDX_LDS int* sTask = steps + 2 * (ns + 1);
  const Real* rval = static_cast<const Real*>(sw.runVal);
  RunRaw<Real, kIdx16> pre[kPreRuns];
  bool havePre[kPreRuns];
  Real dpre[4] = {0, 0, 0, 0};
  bool haveD = false;
  v4i mpre[2];
  Real dvpre[2] = {0, 0};
  int dopre[2] = {0, 0};
  // Every thread issues its loads unconditionally (indices clamped to valid
  // entries) so that no branch or register copy waits for them before use.
  const int lastRun = max(sRun[ns] - 1, 0);
  const int lastDense = max(sDense[ns] - 4, 0);
  const int lastTask = max(sTask[ns] - 1, 0);
  const PDX_GLOBAL Real* gdense = gview<Real>(sw.dense);
  const PDX_GLOBAL v4i* gmeta = gview<v4i>(sw.laneMeta);
  const PDX_GLOBAL Real* gldinv = gview<Real>(sw.laneDinv);
  const PDX_GLOBAL int* gldense = gview<int>(sw.laneDense);
  auto loadMeta = [&](int t, v4i& m, int& doff, Real& dv) {
    const size_t ti = (size_t)min(t, lastTask) * kWave + lane;
    m = gmeta[ti];
    doff = gldense[ti];
    if (!kFwd) dv = gldinv[ti];
  };
  // Lanes past the step's end load the step's first item instead: all such lanes
  // of a wavefront then hit one address and coalesce into one request, rather
  // than fetching unrelated runs.
  auto prefetch = [&](int st) {
    const int rb = sRun[st], re = sRun[st + 1];
    const int rfirst = min(rb, lastRun);
#pragma unroll
    for (int q = 0; q < kPreRuns; ++q) {
      const int r = rb + tid + q * bs;
      havePre[q] = r < re;
      loadRun<Real, kIdx16>(sw.runCol, rval, havePre[q] ? r : rfirst, pre[q]);
    }
    if constexpr (!kRegRows) {
      const int db = sDense[st];
      const int d = db + 4 * tid;
      haveD = d < sDense[st + 1];
      const int dc = haveD ? d : min(db, lastDense);
#pragma unroll
      for (int q = 0; q < 4; ++q) dpre[q] = gdense[dc + q];
    }
    const int tb = sTask[st], te = sTask[st + 1];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int t = tb + wave + q * nw;
      loadMeta(t < te ? t : tb, mpre[q], dopre[q], dvpre[q]);
    }
  };
  Real xc[kRegRows ? kMaxChain : 1];
  (void)xc;
  prefetch(0);
  if constexpr (kRegRows) {
    if (sTask[0] + wave < sTask[1]) loadChainRow<Real>(gdense + sDense[0], mpre[0], dopre[0], xc);
  }
  for (int st = 0; st < ns; ++st) {
    const long long ts0 = tick();
    const int base = sRun[st], end = sRun[st + 1];
    const int dbase = sDense[st], dend = sDense[st + 1];
    const int tb = sTask[st], te = sTask[st + 1];
#pragma unroll
    for (int q = 0; q < kPreRuns; ++q)
      if (havePre[q]) runPartial<Real, kIdx16>(pre[q], W, P + 3 * (tid + q * bs));
    // runs past the prefetched ones, two at a time so their loads overlap
    for (int r = base + tid + kPreRuns * bs; r < end; r += 2 * bs) {
      RunRaw<Real, kIdx16> t0, t1;
      const bool two = r + bs < end;
      loadRun<Real, kIdx16>(sw.runCol, rval, r, t0);
      loadRun<Real, kIdx16>(sw.runCol, rval, two ? r + bs : r, t1);
      runPartial<Real, kIdx16>(t0, W, P + 3 * (r - base));
      if (two) runPartial<Real, kIdx16>(t1, W, P + 3 * (r + bs - base));
    }
    if constexpr (!kRegRows) {
      if (haveD)
#pragma unroll
        for (int q = 0; q < 4; ++q) D[4 * tid + q] = dpre[q];
#pragma unroll 2
      for (int d = dbase + 4 * (tid + bs); d < dend; d += 4 * bs)
#pragma unroll
        for (int q = 0; q < 4; ++q) D[d - dbase + q] = gdense[d + q];
    }
    v4i mt = mpre[0], mn = mpre[1], mnn = mpre[1];
    Real dv = dvpre[0], dvn = dvpre[1], dvnn = 0;
    int doff = dopre[0], don = dopre[1], donn = 0;
    if (st + 1 < ns) prefetch(st + 1);
    const long long ts1 = tick();
    barrier();
    const long long ts2 = tick();
    const PDX_GLOBAL Real* Xs = gdense + dbase;
    (void)Xs;
    for (int t = tb + wave; t < te; t += nw) {
      // this task's inverse-block rows (the first task's are here already), and
      // the metadata of the task after the next
      if constexpr (kRegRows) {
        if (t != tb + wave) loadChainRow<Real>(Xs, mt, doff, xc);
      }
      if (t + 2 * nw < te) loadMeta(t + 2 * nw, mnn, donn, dvnn);
      const int k = mt.x;
      const int info = mt.w;
      const bool act = k >= 0;
      Real a0 = 0, a1 = 0, a2 = 0;
      if (act) {
        a0 = W[3 * k + 0];
        a1 = W[3 * k + 1];
        a2 = W[3 * k + 2];
        if (!kFwd) {
          a0 = a0 * dv;
          a1 = a1 * dv;
          a2 = a2 * dv;
        }
        subtractPartials<Real>(P, mt.y, mt.z, a0, a1, a2);
      }
      if constexpr (kRegRows)
        segmentFinishR<Real>(xc, info, lane, a0, a1, a2);
      else
        segmentFinish<Real>(D + doff, info, lane, act, a0, a1, a2);
      if (act) {
        W[3 * k + 0] = a0;
        W[3 * k + 1] = a1;
        W[3 * k + 2] = a2;
        store(k, a0, a1, a2);
      }
      mt = mn;
      dv = dvn;
      doff = don;
      mn = mnn;
      dvn = dvnn;
      don = donn;
    }
    if constexpr (kRegRows) {
      // the next step's first task: its inverse-block rows load across the barrier
      if (st + 1 < ns && sTask[st + 1] + wave < sTask[st + 2])
        loadChainRow<Real>(gdense + sDense[st + 1], mpre[0], dopre[0], xc);
    }
    const long long ts3 = tick();
    barrier();
    if (tid == 0) {
      tstream.stream += ts1 - ts0;
      tstream.wait1 += ts2 - ts1;
      tstream.fin += ts3 - ts2;
      tstream.wait2 += tick() - ts3;
    }
  }
}

// Copies the step tables (run, dense and task offsets) of the workgroup's sweeps
// into LDS, one after the other: forward, backward, and (the lead of a pair) the
// separator's forward and backward.  Sweeps without steps copy nothing.
__device__ inline void loadStepTables(const InstDev& I) {
  PDX_LDS int* dst = ldsAt<int>(I.stepOff);
  const BlockSweepDev* sws[4] = {&I.bfwd, &I.bbwd, &I.bfwdS, &I.bbwdS};
  for (int q = 0; q < 4; ++q) {
    const int n1 = sws[q]->nstep + 1;
    if (n1 > 1)
      for (int i = threadIdx.x; i < n1; i += blockDim.x) {
        dst[i] = sws[q]->stepRun[i];
        dst[n1 + i] = sws[q]->stepDense[i];
        dst[2 * n1 + i] = sws[q]->stepTask[i];
      }
    dst += 3 * n1;
  }
}

// LDS step tables of sweep q (0 forward, 1 backward, 2 / 3 separator forward /
// backward), in loadStepTables' layout.
__device__ inline const PDX_LDS int* stepTables(const InstDev& I, int q) {
  int off = 0;
  if (q > 0) off += 3 * (I.bfwd.nstep + 1);
  if (q > 1) off += 3 * (I.bbwd.nstep + 1);
  if (q > 2) off += 3 * (I.bfwdS.nstep + 1);
  return ldsAt<int>(I.stepOff) + off;
}

// ---- one instance on two workgroups (PairDev) ----------------------------------
// Vertex loops of a pair run over the workgroup's own vertices (I.own); a
// single-workgroup instance has own == null and takes every vertex.
__device__ inline int ownCount(const InstDev& I) { return I.own ? I.nOwn : I.n; }
__device__ inline int ownVertex(const InstDev& I, int i) { return I.own ? I.own[i] : i; }

constexpr long long kPairSpin = 1LL << 25;  // polls (s_sleep 2 apart) before a wait gives up, ~2 s

// The XCD (its L2) the calling workgroup runs on.
__device__ inline int xccId() { return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20); }  // HW_REG_XCC_ID[3:0]

// Waits until this wave's global stores have been performed (vmcnt 0); the
// signal fences keep the compiler from moving memory operations across it.
__device__ inline void storesPerformed() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0), expcnt and lgkmcnt not waited
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Synchronisation point `target` (the count of points passed) of a pair.  Every
// wave releases its global stores, thread 0 publishes the count and, with
// `wait`, polls until the partner has published the same point and acquires
// (agent scope), so the partner's stores before its point are visible to the
// whole workgroup after the return.  The release is at agent scope (the L2 is
// written back for other XCDs) unless both workgroups share an XCD
// (`sameL2`): then their stores meet in that L2 and waiting until they are
// performed suffices.  The wait is bounded: past kPairSpin polls, or when the
// partner gave up, it sets the context's give-up flag and returns false, and
// the caller leaves the kernel.
__device__ PDX_INLINE bool pairSync(const InstDev& I, int target, bool wait, int* okp, bool sameL2 = false) {
  if (sameL2)
    storesPerformed();
  else
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) {
    int* mine = I.pair.flags + (I.pairRole == 1 ? 0 : 32);
    int* other = I.pair.flags + (I.pairRole == 1 ? 32 : 0);
    __hip_atomic_store(mine, target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int ok = 1;
    if (wait) {
      long long polls = 0;
      while (__hip_atomic_load(other, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (__hip_atomic_load(I.pair.giveUp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
            ++polls > kPairSpin) {
          __hip_atomic_store(I.pair.giveUp, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    *okp = ok;
  }
  __syncthreads();
  return *okp != 0;
}

// The stop test of one iteration (Integrator.cpp:673-751) on the maxima of both
// halves of a pair.
__device__ inline bool stopTest(const InstDev& I, double scale, double diff, double residual) {
  const double rel = (scale > 0.0) ? diff / scale : diff;
  const bool absHit = (I.absTol > 0.0) && (diff <= I.absTol);
  const bool relHit = (I.relTol > 0.0) && (rel <= I.relTol);
  const bool resHit = (I.resTol > 0.0) && (residual <= I.resThresh);
  return (I.resTol > 0.0) ? resHit : (absHit || relHit);
}

// Runs are padded with entries of value 0 at column 0, and rank 0 is in the
// lead's half: a partner's W holds no value there until the first exchange, only
// what the LDS held before this workgroup (a NaN or infinity pattern there turns
// 0 * W[0] into NaN).  The partner zeroes it before its first sweep (visible
// after the next barrier): every padded product is then +-0, which leaves a
// partial sum's bits as they are with any finite W[0].
template <typename Real, typename WP>
__device__ inline void clearPartnerPad(const InstDev& I, WP W) {
  if (I.pairRole == 2 && threadIdx.x < 3) W[threadIdx.x] = (Real)0;
}

// Copies W[3 lo, 3 hi) between the workgroup's LDS and the pair's exchange buffer.
template <typename Real, typename WP>
__device__ inline void pairPut(const InstDev& I, WP W, int lo, int hi) {
  Real* xb = static_cast<Real*>(I.pair.xbuf);
  for (int i = 3 * lo + threadIdx.x; i < 3 * hi; i += blockDim.x) xb[i] = W[i];
}
template <typename Real, typename WP>
__device__ inline void pairGet(const InstDev& I, WP W, int lo, int hi) {
  const Real* xb = static_cast<const Real*>(I.pair.xbuf);
  for (int i = 3 * lo + threadIdx.x; i < 3 * hi; i += blockDim.x) W[i] = xb[i];
}
// Each workgroup of a pair hands over what the other needs to sweep the
// separator (before a pairSync) and takes the other's (after it): the lead its
// half's forward values and the separator's right-hand sides (it gathers the
// separator's vertices), the partner its half's forward values.
template <typename Real, typename WP>
__device__ inline void pairExchange(const InstDev& I, WP W) {
  if (I.pairRole == 1) {
    pairPut<Real>(I, W, 0, I.rA);
    pairPut<Real>(I, W, I.rB, I.m);
  } else {
    pairPut<Real>(I, W, I.rA, I.rB);
  }
}
template <typename Real, typename WP>
__device__ inline void pairTake(const InstDev& I, WP W) {
  if (I.pairRole == 1) {
    pairGet<Real>(I, W, I.rA, I.rB);
  } else {
    pairGet<Real>(I, W, 0, I.rA);
    pairGet<Real>(I, W, I.rB, I.m);
  }
}

struct Acc3 {
  double x, y, z;
};

// Reference right-hand side of one vertex (exact form): seed bBase + bendRhs, then
// each incident edge in colour order: +-kappa d_e and the pin compensation
// kappa q (DistanceTerm.cpp:94-102, :121-152, :279-317); pinned rows take q (Assembler.cpp:319-332).
__device__ inline Acc3 rhsExact(const InstDev& I, int v) {
  if (I.pinned[v]) return {I.pinpos[3 * v], I.pinpos[3 * v + 1], I.pinpos[3 * v + 2]};
  const double mh2 = I.mass[v] * I.invH2;
  Acc3 b{mh2 * I.xhat[3 * v], mh2 * I.xhat[3 * v + 1], mh2 * I.xhat[3 * v + 2]};
  if (I.brhs) {
    b.x = b.x + I.brhs[3 * v];
    b.y = b.y + I.brhs[3 * v + 1];
    b.z = b.z + I.brhs[3 * v + 2];
  } else {
    b.x = b.x + 0.0;
    b.y = b.y + 0.0;
    b.z = b.z + 0.0;
  }
  const double* x = I.x;
  const double xv0 = x[3 * v], xv1 = x[3 * v + 1], xv2 = x[3 * v + 2];
  const int p1 = I.vstart[v + 1];
#pragma unroll 4
  for (int p = I.vstart[v]; p < p1; ++p) {
    const int rc = I.recC[p];
    const RestKappa rk = I.rkC[p];
    const double rest = rk.rest, k = rk.kap;
    const int u = rc >> 2;
    const bool uP = (rc & 2) != 0, isB = (rc & 1) != 0;
    const double u0 = x[3 * u], u1 = x[3 * u + 1], u2 = x[3 * u + 2];
    const double rx = isB ? u0 - xv0 : xv0 - u0;
    const double ry = isB ? u1 - xv1 : xv1 - u1;
    const double rz = isB ? u2 - xv2 : xv2 - u2;
    const double len = sqrt(rx * rx + ry * ry + rz * rz);
    double dx = 0.0, dy = 0.0, dz = 0.0;
    if (!(len <= kMinLength)) {
      const double sc = rest / len;
      dx = rx * sc;
      dy = ry * sc;
      dz = rz * sc;
    }
    const double cx = dx * k, cy = dy * k, cz = dz * k;
    if (!isB) {
      b.x += cx;
      b.y += cy;
      b.z += cz;
    } else {
      b.x -= cx;
      b.y -= cy;
      b.z -= cz;
    }
    if (uP) {
      b.x += k * I.pinpos[3 * u];
      b.y += k * I.pinpos[3 * u + 1];
      b.z += k * I.pinpos[3 * u + 2];
    }
  }
  return b;
}

constexpr int kMaxE = 4;  // incident edges gathered per group, all loads in flight
constexpr int kMaxS = 2;  // incident bend stencils per group likewise

// Per-edge term of the residual gather and of the update right-hand side.
template <bool kRhs>
__device__ inline void edgeTerm(const InstDev& I, int rc, double rest, double k, double u0, double u1,
                                double u2, double xv0, double xv1, double xv2, Acc3& f, Acc3& r) {
  const int u = rc >> 2;
  const bool uP = (rc & 2) != 0, isB = (rc & 1) != 0;
  const double rx = isB ? u0 - xv0 : xv0 - u0;
  const double ry = isB ? u1 - xv1 : xv1 - u1;
  const double rz = isB ? u2 - xv2 : xv2 - u2;
  const double len = sqrt(rx * rx + ry * ry + rz * rz);
  double dx = 0.0, dy = 0.0, dz = 0.0;
  if (!(len <= kMinLength)) {
    const double sc = rest / len;
    dx = rx * sc;
    dy = ry * sc;
    dz = rz * sc;
  }
  const double cx = rx * k - dx * k, cy = ry * k - dy * k, cz = rz * k - dz * k;
  if (!isB) {
    f.x += cx;
    f.y += cy;
    f.z += cz;
  } else {
    f.x -= cx;
    f.y -= cy;
    f.z -= cz;
  }
  if (kRhs) {
    const double sg = isB ? -1.0 : 1.0;
    const double t0 = uP ? I.pinpos[3 * u] : u0;
    const double t1 = uP ? I.pinpos[3 * u + 1] : u1;
    const double t2 = uP ? I.pinpos[3 * u + 2] : u2;
    r.x += sg * dx * k - k * (xv0 - t0);
    r.y += sg * dy * k - k * (xv1 - t1);
    r.z += sg * dz * k - k * (xv2 - t2);
  }
}

template <bool kRhs>
__device__ inline void bendTerm(const InstDev& I, int ea, int eb, int ec, double K, double w,
                                const double* xa, const double* xb, const double* xc, Acc3& fb, Acc3& r) {
  const double s0 = xa[0] - 2.0 * xb[0] + xc[0];
  const double s1 = xa[1] - 2.0 * xb[1] + xc[1];
  const double s2 = xa[2] - 2.0 * xb[2] + xc[2];
  const double coef = -K * w;
  fb.x += s0 * coef;
  fb.y += s1 * coef;
  fb.z += s2 * coef;
  if (kRhs) {
    const int vv[3] = {ea, eb, ec};
    const double* xs[3] = {xa, xb, xc};
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    const double wt[3] = {1.0, -2.0, 1.0};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int u = vv[q] >> 1;
      const bool uP = (vv[q] & 1) != 0;
      t0 += wt[q] * (uP ? I.pinpos[3 * u] : xs[q][0]);
      t1 += wt[q] * (uP ? I.pinpos[3 * u + 1] : xs[q][1]);
      t2 += wt[q] * (uP ? I.pinpos[3 * u + 2] : xs[q][2]);
    }
    r.x += coef * t0;
    r.y += coef * t1;
    r.z += coef * t2;
  }
}

// Residual of one free vertex at x (Integrator.cpp:92-109 + :230-263), and, when
// `rhs` is non-null, the update right-hand side b(x) - (L x) of the same vertex
// (mixed form).  L x uses the pinned targets q for pinned columns, which is
// what the assembled L and the pin compensation in b amount to.  Edges and
// stencils are taken in groups of kMaxE / kMaxS whose loads are all issued
// before any use; the summation order stays the incidence order.
template <bool kRhs>
__device__ inline double residualAndRhsT(const InstDev& I, int v, Acc3& rhsOut) {
  const double* x = I.x;
  const double xv0 = x[3 * v], xv1 = x[3 * v + 1], xv2 = x[3 * v + 2];
  // every load that needs only v, issued together
  const int p0 = I.vstart[v], ne = I.vstart[v + 1] - p0;
  const int s0 = I.bstart[v], nb = I.bstart[v + 1] - s0;
  const double mass = I.mass[v];
  const double xh0 = I.xhat[3 * v], xh1 = I.xhat[3 * v + 1], xh2 = I.xhat[3 * v + 2];
  Acc3 f{0.0, 0.0, 0.0};
  Acc3 r{0.0, 0.0, 0.0};
  {
    for (int g = 0; g < ne; g += kMaxE) {
      int rc[kMaxE];
      double rest[kMaxE], kap[kMaxE], ux[kMaxE][3];
#pragma unroll
      for (int q = 0; q < kMaxE; ++q)
        if (g + q < ne) {
          rc[q] = I.recE[p0 + g + q];
          const RestKappa rk = I.rkE[p0 + g + q];
          rest[q] = rk.rest;
          kap[q] = rk.kap;
        }
#pragma unroll
      for (int q = 0; q < kMaxE; ++q)
        if (g + q < ne) {
          const int u = rc[q] >> 2;
          ux[q][0] = x[3 * u];
          ux[q][1] = x[3 * u + 1];
          ux[q][2] = x[3 * u + 2];
        }
#pragma unroll
      for (int q = 0; q < kMaxE; ++q)
        if (g + q < ne)
          edgeTerm<kRhs>(I, rc[q], rest[q], kap[q], ux[q][0], ux[q][1], ux[q][2], xv0, xv1, xv2, f, r);
    }
  }
  Acc3 fb{0.0, 0.0, 0.0};
  {
    for (int g = 0; g < nb; g += kMaxS) {
      BendRec br[kMaxS];
      double K[kMaxS], p1[kMaxS][3], p2[kMaxS][3];
#pragma unroll
      for (int q = 0; q < kMaxS; ++q)
        if (g + q < nb) {
          br[q] = I.brRec[s0 + g + q];
          K[q] = I.brK[s0 + g + q];
        }
      // the positions of the stencil's other two vertices; the slot vertex's own
      // position is xv
#pragma unroll
      for (int q = 0; q < kMaxS; ++q)
        if (g + q < nb) {
          const int u1 = br[q].other1 >> 1, u2 = br[q].other2 >> 1;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            p1[q][d] = x[3 * u1 + d];
            p2[q][d] = x[3 * u2 + d];
          }
        }
#pragma unroll
      for (int q = 0; q < kMaxS; ++q)
        if (g + q < nb) {
          // back to stencil order a, b, c
          const int role = br[q].role & 3;
          const int self = (v << 1) | ((br[q].role >> 2) & 1);
          const double xs[3] = {xv0, xv1, xv2};
          double xa[3], xb[3], xc[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            xa[d] = role == 0 ? xs[d] : p1[q][d];
            xb[d] = role == 0 ? p1[q][d] : (role == 1 ? xs[d] : p2[q][d]);
            xc[d] = role == 2 ? xs[d] : p2[q][d];
          }
          const int ea = role == 0 ? self : br[q].other1;
          const int eb = role == 0 ? br[q].other1 : (role == 1 ? self : br[q].other2);
          const int ec = role == 2 ? self : br[q].other2;
          bendTerm<kRhs>(I, ea, eb, ec, K[q], (double)br[q].weight, xa, xb, xc, fb, r);
        }
    }
  }
  f.x += fb.x;
  f.y += fb.y;
  f.z += fb.z;
  const double mi = mass * I.invH2;
  const double i0 = (xv0 - xh0) * mi, i1 = (xv1 - xh1) * mi, i2 = (xv2 - xh2) * mi;
  if (kRhs) {
    rhsOut.x = r.x - i0;
    rhsOut.y = r.y - i1;
    rhsOut.z = r.z - i2;
  }
  const double t0 = i0 + f.x, t1 = i1 + f.y, t2 = i2 + f.z;
  return sqrt(t0 * t0 + t1 * t1 + t2 * t2) / mass;
}

__device__ inline double residualAndRhs(const InstDev& I, int v, Acc3* rhs) {
  Acc3 dummy;
  return rhs ? residualAndRhsT<true>(I, v, *rhs) : residualAndRhsT<false>(I, v, dummy);
}

template <typename Real, typename WP>
__device__ inline void putW(WP W, int r, double a, double b, double c) {
  W[3 * r + 0] = (Real)a;
  W[3 * r + 1] = (Real)b;
  W[3 * r + 2] = (Real)c;
}

// Fills W with the first right-hand side of a substep (iteration 0 at x_n).
template <typename Real, bool kExact, typename WP>
__device__ PDX_INLINE void gatherFirst(const InstDev& I, WP W) {
  const int nv = ownCount(I);
  for (int i = threadIdx.x; i < nv; i += blockDim.x) {
    const int v = ownVertex(I, i);
    const int r = I.vrank[v];
    if (kExact) {
      const Acc3 b = rhsExact(I, v);
      putW<Real>(W, r, b.x, b.y, b.z);
    } else if (I.pinned[v]) {
      putW<Real>(W, r, I.pinpos[3 * v] - I.x[3 * v], I.pinpos[3 * v + 1] - I.x[3 * v + 1],
           I.pinpos[3 * v + 2] - I.x[3 * v + 2]);
    } else {
      Acc3 rh;
      (void)residualAndRhs(I, v, &rh);
      putW<Real>(W, r, rh.x, rh.y, rh.z);
    }
  }
}


template <typename Real, bool kBlocked, bool kIdx16, typename WP>
__device__ PDX_INLINE void runForward(const InstDev& I, WP W, SweepTimes& tst) {
  const Real* dinv = static_cast<const Real*>(I.dinv);
  auto none = [](int, Real, Real, Real) {};
  if (kBlocked)
    blockSweep<Real, true, kIdx16>(I.bfwd, W, ldsAt<Real>(I.partOff), ldsAt<Real>(I.denseOff), stepTables(I, 0), none, tst, I.ldsW != 0);
  else
    sweep<Real, true>(I.fwd, W, dinv, none);
}

// The lead of a pair: the separator's forward and backward sweeps, once both
// halves' forward values are in its W.
template <typename Real, bool kIdx16, typename WP>
__device__ PDX_INLINE void runSeparator(const InstDev& I, WP W, SweepTimes& tst) {
  auto none = [](int, Real, Real, Real) {};
  blockSweep<Real, true, kIdx16>(I.bfwdS, W, ldsAt<Real>(I.partOff), ldsAt<Real>(I.denseOff), stepTables(I, 2), none, tst, I.ldsW != 0);
  blockSweep<Real, false, kIdx16>(I.bbwdS, W, ldsAt<Real>(I.partOff), ldsAt<Real>(I.denseOff), stepTables(I, 3), none, tst, I.ldsW != 0);
}

template <typename Real, bool kExact, bool kBlocked, bool kIdx16, typename WP>
__device__ PDX_INLINE void runBackward(const InstDev& I, WP W, SweepTimes& tst) {
  const Real* dinv = static_cast<const Real*>(I.dinv);
  double* x = I.x;
  const int* rankv = I.rankv;
  const uint8_t* pinned = I.pinned;
  const double* pinpos = I.pinpos;
  // exact form: the sweep's result is the new position, stored as it finishes;
  // mixed form: the update is applied after the sweep (applyUpdate), off the
  // sweep's critical path.
  auto put = [=](int k, Real a0, Real a1, Real a2) {
    if (kExact) {
      const int v = rankv[k];
      x[3 * v] = (double)a0;
      x[3 * v + 1] = (double)a1;
      x[3 * v + 2] = (double)a2;
    }
  };
  (void)pinned;
  (void)pinpos;
  if (kBlocked)
    blockSweep<Real, false, kIdx16>(I.bbwd, W, ldsAt<Real>(I.partOff), ldsAt<Real>(I.denseOff), stepTables(I, 1), put, tst, I.ldsW != 0);
  else
    sweep<Real, false>(I.bwd, W, dinv, put);
}

// Pin targets of substep s from the instance's pin trajectory, and the bending
// constant right-hand side that follows them (Assembler.cpp:94-137 restricted to
// the stencils that hold a pinned vertex, in stencil order, by one thread, so the
// sums are those of the host assembly).
__device__ void applyPinTrajectory(const InstDev& I, int s) {
  const double* q = I.pinTraj + (size_t)s * I.npin * 3;
  for (int j = threadIdx.x; j < I.npin; j += blockDim.x) {
    const int v = I.pinList[j];
    I.pinpos[3 * v] = q[3 * j];
    I.pinpos[3 * v + 1] = q[3 * j + 1];
    I.pinpos[3 * v + 2] = q[3 * j + 2];
  }
  __syncthreads();
  if (I.brhs && threadIdx.x == 0) {
    const double w[3] = {1.0, -2.0, 1.0};
    for (int t = 0; t < I.npinSten; ++t) {
      const int* ps = I.pinSten + 4 * t;
      const int st = ps[0];
      const int vv[3] = {I.sa[st], I.sb[st], I.sc[st]};
      for (int k = 0; k < 3; ++k)
        if (ps[1 + k] < 0)
          for (int d = 0; d < 3; ++d) I.brhs[3 * vv[k] + d] = 0.0;
    }
    for (int t = 0; t < I.npinSten; ++t) {
      const int* ps = I.pinSten + 4 * t;
      const int st = ps[0];
      if (I.bk[st] == 0.0) continue;
      const int vv[3] = {I.sa[st], I.sb[st], I.sc[st]};
      double c0 = 0.0, c1 = 0.0, c2 = 0.0;
      for (int k = 0; k < 3; ++k) {
        if (ps[1 + k] < 0) continue;
        const double* p = I.pinpos + 3 * vv[k];
        c0 += p[0] * w[k];
        c1 += p[1] * w[k];
        c2 += p[2] * w[k];
      }
      for (int k = 0; k < 3; ++k) {
        if (ps[1 + k] >= 0) continue;
        const double f = -I.bk[st] * w[k];
        double* slot = I.brhs + 3 * vv[k];
        slot[0] += c0 * f;
        slot[1] += c1 * f;
        slot[2] += c2 * f;
      }
    }
  }
  __syncthreads();
}

// Launch flags of the differentiable rollout.
constexpr int kRollTraj = 1;   // pin targets of each substep from the pin trajectory
constexpr int kRollTape = 2;   // record every iterate of every substep (FP64)
constexpr int kRollFixed = 4;  // fixed iteration count: no residual, no stop test

template <typename Real, bool kExact, bool kBlocked, bool kIdx16, bool kRoll, typename WP>
__device__ void substepBody(const InstDev& I, WP W, int nsub, int subBase, int flags) {
  extern __shared__ double smem[];
  double* red = smem;  // 2 x 16 slots
  const int n = I.n;
  const int tid = threadIdx.x, bs = blockDim.x;
  const bool fixedIter = kRoll && (flags & kRollFixed) != 0;
  const bool wantRes = I.wantResidual && !fixedIter;
  const size_t n3 = (size_t)3 * n;
  // an instance split over two workgroups (FP32 solves, blocked sweeps, no rollout)
  constexpr bool kPairable = !kExact && kBlocked && !kRoll;
  const bool paired = kPairable && I.pairRole != 0;
  const int nv = ownCount(I);
  int* pairOk = (int*)(red + 31);
  int seq = 0;  // synchronisation points of the pair passed in this launch
  bool sameL2 = false;  // both workgroups of the pair on one XCD (known after the first point)
  if (paired) {
    seq = __hip_atomic_load(I.pair.flags + (I.pairRole == 1 ? 0 : 32), __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT);
    if (tid == 0)
      __hip_atomic_store(I.pair.flags + (I.pairRole == 1 ? 16 : 48), xccId(), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  }

  if (kBlocked) loadStepTables(I);  // visible after the first barrier below
  if (paired) clearPartnerPad<Real>(I, W);
  long long tk[4] = {0, 0, 0, 0};
  long long tacc[5] = {0, 0, 0, 0, 0};
  SweepTimes swt;
  for (int sub = 0; sub < nsub; ++sub) {
    if (kRoll && (flags & kRollTraj)) applyPinTrajectory(I, subBase + sub);
    double* tape = (kRoll && (flags & kRollTape)) ? I.tape + (size_t)sub * (I.maxIter + 1) * n3 : nullptr;
    const long long tp0 = tid == 0 ? tickStage() : 0;
    // predict (Integrator.cpp:447-463) and the displacement scale
    double lscale = 0.0;
    for (int i = tid; i < nv; i += bs) {
      const int v = ownVertex(I, i);
      for (int d = 0; d < 3; ++d) {
        const double p = I.x[3 * v + d];
        I.xbs[3 * v + d] = p;
        if (tape) tape[3 * v + d] = p;
        const double xh = p + I.vel[3 * v + d] * I.h + I.gh2[d];
        I.xhat[3 * v + d] = xh;
        I.prev[3 * v + d] = p;
        lscale = fmax(lscale, fabs(xh - p));
      }
    }
    double scale = blockMax(lscale, red);
    scale = fmax(scale, I.gScaleH2);
    const long long tp1 = tid == 0 ? tickStage() : 0;
    gatherFirst<Real, kExact>(I, W);
    __syncthreads();
    if (tid == 0) {
      tacc[3] += tp1 - tp0;
      tacc[0] += tickStage() - tp1;
    }

    double chebOmega = 1.0;
    int used = 0, conv = 0;
    double lastRes = -1.0;
    for (int k = 0; k < I.maxIter; ++k) {
      if (tid == 0) tk[0] = tickStage();
      runForward<Real, kBlocked, kIdx16>(I, W, swt);
      if constexpr (kPairable) {
        if (paired) {
          // The halves hand each other their forward values and both sweep the
          // separator (the same arithmetic, so the same bits).  The maxima of the
          // previous iteration travel with this exchange: its stop decision is
          // taken here, before anything but W changed.
          pairExchange<Real>(I, W);
          if (!pairSync(I, ++seq, true, pairOk, sameL2)) return;
          sameL2 = __hip_atomic_load(I.pair.flags + 16, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                   __hip_atomic_load(I.pair.flags + 48, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (k > 0) {
            const double* other = I.pair.red + 8 * ((k - 1) & 1) + 4 * (2 - I.pairRole);
            const double* mine = I.pair.red + 8 * ((k - 1) & 1) + 4 * (I.pairRole - 1);
            const double sc = fmax(mine[0], other[0]), df = fmax(mine[1], other[1]);
            const double rsd = fmax(mine[2], other[2]);
            lastRes = rsd;
            if (stopTest(I, sc, df, rsd)) {
              conv = 1;
              break;
            }
          }
          pairTake<Real>(I, W);
          __syncthreads();
          runSeparator<Real, kIdx16>(I, W, swt);
        }
      }
      if (tid == 0) tk[1] = tickStage();
      runBackward<Real, kExact, kBlocked, kIdx16>(I, W, swt);
      if (kExact && kBlocked) __syncthreads();  // the sweep stored x; the gather reads it
      if (!kExact) {
        // x += delta (pinned: x = q) for every vertex, then a barrier so that the
        // gather below sees every neighbour's new position (a pair: both halves')
        for (int i = tid; i < nv; i += bs) {
          const int v = ownVertex(I, i);
          const int r = I.vrank[v];
          if (I.pinned[v]) {
            I.x[3 * v] = I.pinpos[3 * v];
            I.x[3 * v + 1] = I.pinpos[3 * v + 1];
            I.x[3 * v + 2] = I.pinpos[3 * v + 2];
          } else {
            I.x[3 * v] += (double)W[3 * r];
            I.x[3 * v + 1] += (double)W[3 * r + 1];
            I.x[3 * v + 2] += (double)W[3 * r + 2];
          }
        }
        __syncthreads();
        if constexpr (kPairable) {
          if (paired && !pairSync(I, ++seq, true, pairOk, sameL2)) return;
        }
      }
      if (tid == 0) tk[2] = tickStage();
      const bool more = (k + 1 < I.maxIter);
      double* tapeK = tape ? tape + (size_t)(k + 1) * n3 : nullptr;
      double ldiff = 0.0, lres = 0.0;
      // Chebyshev blend (Integrator.cpp:597-617, default schedule)
      const bool cheb = !kRoll && I.chebRho > 0.0 && k > I.chebStart;
      if (cheb) {
        const double rho = I.chebRho;
        chebOmega = (k == I.chebStart + 1) ? 2.0 / (2.0 - rho * rho) : 4.0 / (4.0 - rho * rho * chebOmega);
        // residual at the solve's output; the blended iterate becomes the previous one
        for (int v = tid; v < n; v += bs) {
          const double x0 = I.x[3 * v], x1 = I.x[3 * v + 1], x2 = I.x[3 * v + 2];
          if (wantRes && !I.pinned[v] && I.mass[v] > 0.0) lres = fmax(lres, residualAndRhs(I, v, nullptr));
          const double p0 = I.prev[3 * v], p1 = I.prev[3 * v + 1], p2 = I.prev[3 * v + 2];
          const double q0 = p0 + (x0 - p0) * chebOmega, q1 = p1 + (x1 - p1) * chebOmega,
                       q2 = p2 + (x2 - p2) * chebOmega;
          ldiff = fmax(ldiff, fmax(fabs(q0 - p0), fmax(fabs(q1 - p1), fabs(q2 - p2))));
          I.prev[3 * v] = q0;
          I.prev[3 * v + 1] = q1;
          I.prev[3 * v + 2] = q2;
        }
        __syncthreads();
        for (int v = tid; v < n; v += bs)
          for (int d = 0; d < 3; ++d) I.x[3 * v + d] = I.prev[3 * v + d];
        __syncthreads();
        // the next right-hand side at the blended positions
        if (more)
          for (int v = tid; v < n; v += bs) {
            const int r = I.vrank[v];
            if (kExact) {
              const Acc3 b = rhsExact(I, v);
              putW<Real>(W, r, b.x, b.y, b.z);
            } else if (I.pinned[v]) {
              putW<Real>(W, r, I.pinpos[3 * v] - I.x[3 * v], I.pinpos[3 * v + 1] - I.x[3 * v + 1],
                         I.pinpos[3 * v + 2] - I.x[3 * v + 2]);
            } else {
              Acc3 rh;
              (void)residualAndRhs(I, v, &rh);
              putW<Real>(W, r, rh.x, rh.y, rh.z);
            }
          }
      }
      for (int i = cheb ? nv : tid; i < nv; i += bs) {
        const int v = ownVertex(I, i);
        const double x0 = I.x[3 * v], x1 = I.x[3 * v + 1], x2 = I.x[3 * v + 2];
        if (tapeK) {
          tapeK[3 * v] = x0;
          tapeK[3 * v + 1] = x1;
          tapeK[3 * v + 2] = x2;
        }
        const double d = fmax(fabs(x0 - I.prev[3 * v]),
                              fmax(fabs(x1 - I.prev[3 * v + 1]), fabs(x2 - I.prev[3 * v + 2])));
        ldiff = fmax(ldiff, d);
        I.prev[3 * v] = x0;
        I.prev[3 * v + 1] = x1;
        I.prev[3 * v + 2] = x2;
        const bool pinnedV = I.pinned[v] != 0;
        const int r = I.vrank[v];
        if (kExact) {
          if (wantRes && !pinnedV && I.mass[v] > 0.0) lres = fmax(lres, residualAndRhs(I, v, nullptr));
          if (more) {
            const Acc3 b = rhsExact(I, v);
            putW<Real>(W, r, b.x, b.y, b.z);
          }
        } else {
          if (pinnedV) {
            if (more) putW<Real>(W, r, I.pinpos[3 * v] - x0, I.pinpos[3 * v + 1] - x1, I.pinpos[3 * v + 2] - x2);
          } else {
            if (more || wantRes) {
              Acc3 rh;
              const double rv = residualAndRhs(I, v, more ? &rh : nullptr);
              if (wantRes && I.mass[v] > 0.0) lres = fmax(lres, rv);
              if (more) putW<Real>(W, r, rh.x, rh.y, rh.z);
            }
          }
        }
      }
      if (fixedIter) {
        // the next iteration's right-hand side is in W; a barrier orders it
        __syncthreads();
        used = k + 1;
        continue;
      }
      // criterion (Integrator.cpp:673-751); two slot sets alternate so one barrier suffices
      double* rs = red + 16 * ((k & 1) ^ 1);
      double diff = blockMax(ldiff, rs);
      if (tid == 0) tk[3] = tickStage();
      double residual = -1.0;
      if (I.wantResidual) {
        __syncthreads();
        residual = blockMax(lres, rs);
      }
      if constexpr (kPairable) {
        if (paired) {
          // both workgroups take the stop decision on the maxima of both halves:
          // at the next iteration's exchange, or here after the last iteration
          double* mine = I.pair.red + 8 * (k & 1) + 4 * (I.pairRole - 1);
          const double* other = I.pair.red + 8 * (k & 1) + 4 * (2 - I.pairRole);
          if (tid == 0) {
            mine[0] = scale;
            mine[1] = diff;
            mine[2] = residual;
          }
          used = k + 1;
          if (k + 1 < I.maxIter) continue;
          if (!pairSync(I, ++seq, true, pairOk, sameL2)) return;
          scale = fmax(scale, other[0]);
          diff = fmax(diff, other[1]);
          residual = fmax(residual, other[2]);
        }
      }
      if (tid == 0) {
        tacc[1] += tk[1] - tk[0];
        tacc[2] += tk[2] - tk[1];
        tacc[0] += tk[3] - tk[2];
        tacc[4] += tickStage() - tk[3];
      }
      const double rel = (scale > 0.0) ? diff / scale : diff;
      const bool absHit = (I.absTol > 0.0) && (diff <= I.absTol);
      const bool relHit = (I.relTol > 0.0) && (rel <= I.relTol);
      const bool resHit = (I.resTol > 0.0) && (residual <= I.resThresh);
      const bool c = (I.resTol > 0.0) ? resHit : (absHit || relHit);
      used = k + 1;
      lastRes = residual;
      if (c) {
        conv = 1;
        break;
      }
    }
    // velocity + damping + re-pin (Integrator.cpp:774-791)
    const long long tv0 = tid == 0 ? tickStage() : 0;
    for (int i = tid; i < nv; i += bs) {
      const int v = ownVertex(I, i);
      const bool pinnedV = I.pinned[v] != 0;
      for (int d = 0; d < 3; ++d) {
        if (pinnedV) {
          I.vel[3 * v + d] = 0.0;
          I.x[3 * v + d] = I.pinpos[3 * v + d];
        } else {
          I.vel[3 * v + d] = (I.x[3 * v + d] - I.xbs[3 * v + d]) * I.dampInvH;
        }
      }
    }
    if (tid == 0 && I.pairRole != 2) {  // a pair's lead keeps the instance's statistics
      tacc[3] += tickStage() - tv0;
      I.istats[0] = used;
      I.istats[1] = conv;
      I.istats[2] += conv;
      I.lstats[0] += used;
      I.lstats[1] += 1;
      I.dstats[0] = lastRes;
    }
    __syncthreads();
  }
  if (tid == 0 && I.pairRole != 2) {
    for (int q = 0; q < 5; ++q) I.lstats[2 + q] += tacc[q];
    I.lstats[7] += swt.stream;
    I.lstats[8] += swt.wait1;
    I.lstats[9] += swt.fin;
    I.lstats[10] += swt.wait2;
  }
}

// W in LDS (kWL) or in HBM is a separate kernel: one kernel holding both bodies
// spills registers that neither body needs on its own.
template <typename Real, bool kExact, bool kBlocked, bool kIdx16, bool kRoll, bool kWL>
__global__ void __launch_bounds__(512) substepKernel(const InstDev* __restrict__ insts,
                                                      const int* __restrict__ order, int nsub,
                                                      int subBase, int flags) {
  if (!kRoll) flags = 0;
  const InstDev& I = insts[order[blockIdx.x]];
#if PDX_PHASE_TIMERS
  // the workgroup's start and end on the device clock (a pair's partner in slots 13, 14)
  const long long t0 = tick();
#endif
  if (kWL)
    substepBody<Real, kExact, kBlocked, kIdx16, kRoll>(I, ldsAt<Real>(32 * sizeof(double)), nsub, subBase, flags);
  else
    substepBody<Real, kExact, kBlocked, kIdx16, kRoll>(I, (PDX_GLOBAL Real*)I.Wg, nsub, subBase, flags);
#if PDX_PHASE_TIMERS
  if (threadIdx.x == 0) {
    const int slot = I.pairRole == 2 ? 13 : 11;
    I.lstats[slot] = t0;
    I.lstats[slot + 1] = tick();
  }
#endif
}

// Reverse mode of `nsub` substeps (global indices subBase ..), one workgroup per
// instance, from the iterates recorded by the substep kernel (kRollTape) and the
// adjoint (xbar, vbar) of the state after the last of them; leaves the adjoint of
// the state before the first.  Per substep, in reverse: velocity/re-pin, then for
// k = K-1 .. 0 with L x^{k+1} = c + P(x^k):
//   mu = L^{-1} cur         (the same factor: L is symmetric, so the transposed
//                            solve is the forward + backward sweep again)
//   bbar += mu
//   cur  = J_P(x^k)^T mu    (gather over the vertex's edges in edge order)
//   gkap_e += mu . dP/dkappa_e - mu^T (dL/dkappa_e) x^{k+1}   (one thread per edge)
//   gbend_s += -mu^T (dL/dk_s) x^{k+1}                        (one thread per stencil)
// then bbar feeds x_n, v_n (prediction), the pins and the stiffnesses.  Every
// sum has a fixed order; nothing is accumulated atomically.
template <typename Real, bool kBlocked, bool kIdx16, typename WP>
__device__ void adjointBody(const InstDev& I, WP W, int nsub, int subBase) {
  const int n = I.n, K = I.maxIter;
  const int tid = threadIdx.x, bs = blockDim.x;
  const size_t n3 = (size_t)3 * n;
  const double cv = I.dampInvH;
  const double w3[3] = {1.0, -2.0, 1.0};
  const Real* dinv = static_cast<const Real*>(I.dinv);
  auto none = [](int, Real, Real, Real) {};
  auto muOf = [&](int v, int d) { return (double)W[3 * I.vrank[v] + d]; };
  if (kBlocked) loadStepTables(I);
  SweepTimes swt;
  for (int sub = nsub - 1; sub >= 0; --sub) {
    const int s = subBase + sub;
    const double* q = I.pinTraj + (size_t)s * I.npin * 3;
    double* gq = I.gpin + (size_t)s * I.npin * 3;
    // velocity / damping / re-pin (Integrator.cpp:774-791)
    for (int v = tid; v < n; v += bs) {
      const bool pv = I.pinned[v] != 0;
      for (int d = 0; d < 3; ++d) {
        const size_t i = (size_t)3 * v + d;
        I.bbar[i] = 0.0;
        if (pv) {
          I.cur[i] = 0.0;
          I.xnbar[i] = 0.0;
        } else {
          I.cur[i] = I.xbar[i] + cv * I.vbar[i];
          I.xnbar[i] = -cv * I.vbar[i];
        }
      }
    }
    if (tid == 0)
      for (int j = 0; j < I.npin; ++j)
        for (int d = 0; d < 3; ++d) gq[3 * j + d] += I.xbar[(size_t)3 * I.pinList[j] + d];
    __syncthreads();
    const double* tp = I.tape + (size_t)sub * (K + 1) * n3;
    for (int k = K - 1; k >= 0; --k) {
      for (int v = tid; v < n; v += bs)
        putW<Real>(W, I.vrank[v], I.cur[3 * v], I.cur[3 * v + 1], I.cur[3 * v + 2]);
      __syncthreads();
      if (kBlocked) {
        blockSweep<Real, true, kIdx16>(I.bfwd, W, ldsAt<Real>(I.partOff), ldsAt<Real>(I.denseOff),
                                       ldsAt<int>(I.stepOff), none, swt, I.ldsW != 0);
        blockSweep<Real, false, kIdx16>(I.bbwd, W, ldsAt<Real>(I.partOff), ldsAt<Real>(I.denseOff),
                                        ldsAt<int>(I.stepOff) + 3 * (I.bfwd.nstep + 1), none, swt, I.ldsW != 0);
      } else {
        sweep<Real, true>(I.fwd, W, dinv, none);
        sweep<Real, false>(I.bwd, W, dinv, none);
      }
      __syncthreads();
      const double* xk = tp + (size_t)k * n3;
      const double* xk1 = tp + (size_t)(k + 1) * n3;
      // vertices: bbar += mu, cur = J_P^T mu (the edges' terms in edge order)
      for (int v = tid; v < n; v += bs) {
        const bool vP = I.pinned[v] != 0;
        const double mv0 = muOf(v, 0), mv1 = muOf(v, 1), mv2 = muOf(v, 2);
        I.bbar[3 * v] += mv0;
        I.bbar[3 * v + 1] += mv1;
        I.bbar[3 * v + 2] += mv2;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        const double xv0 = xk[3 * v], xv1 = xk[3 * v + 1], xv2 = xk[3 * v + 2];
        for (int p = I.vstart[v]; p < I.vstart[v + 1]; ++p) {
          const int rc = I.recE[p];
          const int u = rc >> 2;
          const bool uP = (rc & 2) != 0, isB = (rc & 1) != 0;
          const bool aP = isB ? uP : vP, bP = isB ? vP : uP;
          if (aP && bP) continue;
          const double u0 = xk[3 * u], u1 = xk[3 * u + 1], u2 = xk[3 * u + 2];
          const double rx = isB ? u0 - xv0 : xv0 - u0;
          const double ry = isB ? u1 - xv1 : xv1 - u1;
          const double rz = isB ? u2 - xv2 : xv2 - u2;
          const double len = sqrt(rx * rx + ry * ry + rz * rz);
          if (!(len > kMinLength)) continue;
          const RestKappa rk = I.rkE[p];
          const double k = rk.kap, l = rk.rest;
          const double mu0 = uP ? 0.0 : muOf(u, 0), mu1 = uP ? 0.0 : muOf(u, 1), mu2 = uP ? 0.0 : muOf(u, 2);
          const double me0 = vP ? 0.0 : mv0, me1 = vP ? 0.0 : mv1, me2 = vP ? 0.0 : mv2;
          const double ma0 = isB ? mu0 : me0, ma1 = isB ? mu1 : me1, ma2 = isB ? mu2 : me2;
          const double mb0 = isB ? me0 : mu0, mb1 = isB ? me1 : mu1, mb2 = isB ? me2 : mu2;
          const double g0 = (ma0 - mb0) * k, g1 = (ma1 - mb1) * k, g2 = (ma2 - mb2) * k;
          const double h0 = rx / len, h1 = ry / len, h2 = rz / len;
          const double t = h0 * g0 + h1 * g1 + h2 * g2;
          const double sc = l / len;
          const double r0 = (g0 - h0 * t) * sc, r1 = (g1 - h1 * t) * sc, r2 = (g2 - h2 * t) * sc;
          if (!isB) {
            a0 += r0;
            a1 += r1;
            a2 += r2;
          } else {
            a0 -= r0;
            a1 -= r1;
            a2 -= r2;
          }
        }
        I.cur[3 * v] = a0;
        I.cur[3 * v + 1] = a1;
        I.cur[3 * v + 2] = a2;
      }
      // edges: dLoss/dkappa_e
      for (int e = tid; e < I.E; e += bs) {
        const int a = I.ea[e], b = I.eb[e];
        const bool aP = I.pinned[a] != 0, bP = I.pinned[b] != 0;
        if (aP && bP) continue;
        const double k = I.kappa[e], l = I.rest[e];
        const double rx = xk[3 * a] - xk[3 * b], ry = xk[3 * a + 1] - xk[3 * b + 1],
                     rz = xk[3 * a + 2] - xk[3 * b + 2];
        const double len = sqrt(rx * rx + ry * ry + rz * rz);
        const double ma0 = aP ? 0.0 : muOf(a, 0), ma1 = aP ? 0.0 : muOf(a, 1), ma2 = aP ? 0.0 : muOf(a, 2);
        const double mb0 = bP ? 0.0 : muOf(b, 0), mb1 = bP ? 0.0 : muOf(b, 1), mb2 = bP ? 0.0 : muOf(b, 2);
        const double dm0 = ma0 - mb0, dm1 = ma1 - mb1, dm2 = ma2 - mb2;
        double g = I.gkap[e];
        if (len > kMinLength) {
          const double sc = l / len;
          g += dm0 * (rx * sc) + dm1 * (ry * sc) + dm2 * (rz * sc);
        }
        if (!aP && !bP) {
          g -= dm0 * (xk1[3 * a] - xk1[3 * b]) + dm1 * (xk1[3 * a + 1] - xk1[3 * b + 1]) +
               dm2 * (xk1[3 * a + 2] - xk1[3 * b + 2]);
        } else if (aP) {
          g -= mb0 * xk1[3 * b] + mb1 * xk1[3 * b + 1] + mb2 * xk1[3 * b + 2];
        } else {
          g -= ma0 * xk1[3 * a] + ma1 * xk1[3 * a + 1] + ma2 * xk1[3 * a + 2];
        }
        (void)k;
        I.gkap[e] = g;
      }
      // stencils: dLoss/dk_s
      for (int t = tid; t < I.B; t += bs) {
        const int vv[3] = {I.sa[t], I.sb[t], I.sc[t]};
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, x0 = 0.0, x1 = 0.0, x2 = 0.0;
        for (int p = 0; p < 3; ++p) {
          const int u = vv[p];
          if (I.pinned[u]) continue;
          s0 += muOf(u, 0) * w3[p];
          s1 += muOf(u, 1) * w3[p];
          s2 += muOf(u, 2) * w3[p];
          x0 += xk1[3 * u] * w3[p];
          x1 += xk1[3 * u + 1] * w3[p];
          x2 += xk1[3 * u + 2] * w3[p];
        }
        I.gbend[t] -= s0 * x0 + s1 * x1 + s2 * x2;
      }
      __syncthreads();
    }
    // prediction x^ = x_n + v_n h + g h^2 feeds the inertial term (m/h^2) x^
    for (int v = tid; v < n; v += bs) {
      const bool pv = I.pinned[v] != 0;
      const double mh = I.mass[v] * I.invH2;
      for (int d = 0; d < 3; ++d) {
        const size_t i = (size_t)3 * v + d;
        if (pv) {
          I.xbar[i] = I.xnbar[i] + I.cur[i];
          I.vbar[i] = 0.0;
        } else {
          const double xh = mh * I.bbar[i];
          I.xbar[i] = I.xnbar[i] + (I.cur[i] + xh);
          I.vbar[i] = xh * I.h;
        }
      }
    }
    // pins: their rows, the compensation kappa q of free neighbours, the bend constant
    if (tid == 0) {
      for (int j = 0; j < I.npin; ++j)
        for (int d = 0; d < 3; ++d) gq[3 * j + d] += I.bbar[(size_t)3 * I.pinList[j] + d];
      for (int t = 0; t < I.npinEdge; ++t) {
        const int e = I.pinEdge[2 * t], j = I.pinEdge[2 * t + 1];
        const int a = I.ea[e], b = I.eb[e];
        const int f = I.pinned[a] ? b : a;
        double g = I.gkap[e];
        for (int d = 0; d < 3; ++d) {
          const double bb = I.bbar[(size_t)3 * f + d];
          gq[3 * j + d] += I.kappa[e] * bb;
          g += bb * q[3 * j + d];
        }
        I.gkap[e] = g;
      }
      for (int t = 0; t < I.npinSten; ++t) {
        const int* ps = I.pinSten + 4 * t;
        const int st = ps[0];
        const int vv[3] = {I.sa[st], I.sb[st], I.sc[st]};
        double C0 = 0.0, C1 = 0.0, C2 = 0.0;
        for (int r = 0; r < 3; ++r)
          if (ps[1 + r] >= 0) {
            C0 += w3[r] * q[3 * ps[1 + r]];
            C1 += w3[r] * q[3 * ps[1 + r] + 1];
            C2 += w3[r] * q[3 * ps[1 + r] + 2];
          }
        for (int p = 0; p < 3; ++p) {
          if (ps[1 + p] >= 0) continue;
          const double* bb = I.bbar + (size_t)3 * vv[p];
          I.gbend[st] -= w3[p] * (bb[0] * C0 + bb[1] * C1 + bb[2] * C2);
          for (int r = 0; r < 3; ++r)
            if (ps[1 + r] >= 0)
              for (int d = 0; d < 3; ++d) gq[3 * ps[1 + r] + d] -= I.bk[st] * w3[p] * w3[r] * bb[d];
        }
      }
    }
    __syncthreads();
  }
}

template <typename Real, bool kBlocked, bool kIdx16, bool kWL>
__global__ void __launch_bounds__(512) adjointKernel(const InstDev* __restrict__ insts,
                                                      const int* __restrict__ order, int nsub,
                                                      int subBase) {
  const InstDev& I = insts[order[blockIdx.x]];
  if (kWL)
    adjointBody<Real, kBlocked, kIdx16>(I, ldsAt<Real>(32 * sizeof(double)), nsub, subBase);
  else
    adjointBody<Real, kBlocked, kIdx16>(I, (PDX_GLOBAL Real*)I.Wg, nsub, subBase);
}

// ---------------------------------------------------------------------------
// Cooperative form for one large instance: the whole grid (one workgroup per CU,
// all co-resident) advances a single cloth.  Vertex phases are grid-strided and
// end in a grid barrier; the sweeps take one step per level of whole
// elimination-tree chains and are paced by step counters (levelSweep), not by
// barriers; reductions go through per-workgroup slots and every workgroup
// reduces them in the same order, so all take the same stop decision.  W lives
// in HBM.  Rows subtract their run partials in run order as the one-workgroup
// sweep does; long chains apply their inverse as a dense product.

namespace cg = cooperative_groups;

// Max of v over the grid.  Slot set `buf` (2 x gridDim.x slots at buf * 2 *
// gridDim.x) alternates between consecutive calls, so a workgroup that runs ahead
// into the next reduction never writes a slot another one is still reading.
__device__ inline double gridMax(cg::grid_group& grid, double v, double* red, double* gslots, int buf) {
  double* s = gslots + (size_t)buf * 2 * gridDim.x;
  const double b = blockMax(v, red);
  if (threadIdx.x == 0) s[blockIdx.x] = b;
  grid.sync();
  double m = 0.0;
  for (int q = threadIdx.x; q < (int)gridDim.x; q += blockDim.x) m = fmax(m, s[q]);
  __syncthreads();
  return blockMax(m, red + 16);
}

// Maxima of a and of b over the grid with one grid barrier (same slot sets).
__device__ inline void gridMax2(cg::grid_group& grid, double a, double b, double* red, double* gslots, int buf,
                                double& ra, double& rb) {
  double* s = gslots + (size_t)buf * 2 * gridDim.x;
  const double ba = blockMax(a, red);
  const double bb = blockMax(b, red + 8);
  if (threadIdx.x == 0) {
    s[blockIdx.x] = ba;
    s[gridDim.x + blockIdx.x] = bb;
  }
  grid.sync();
  double ma = 0.0, mb = 0.0;
  for (int q = threadIdx.x; q < (int)gridDim.x; q += blockDim.x) {
    ma = fmax(ma, s[q]);
    mb = fmax(mb, s[gridDim.x + q]);
  }
  __syncthreads();
  ra = blockMax(ma, red + 16);
  rb = blockMax(mb, red + 24);
}

// Stores that other workgroups read: relaxed agent-scope atomics, written through
// to memory past the XCDs' L2 caches.
template <typename Real>
__device__ inline void agentStore(Real* p, Real v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The grid form's sweeps keep W as 4 Real per rank (x, y, z, unused), so that a
// column's three values come in one device-scope (sc1) load: 16 bytes FP32, 16 +
// 8 bytes FP64, through a buffer resource over W.
__device__ inline __amdgpu_buffer_rsrc_t bufferOver(const void* p, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, bytes, 0x00020000);
}
template <typename Real>
__device__ inline void loadW4(__amdgpu_buffer_rsrc_t w, int c, Real& a0, Real& a1, Real& a2);
template <>
__device__ inline void loadW4<float>(__amdgpu_buffer_rsrc_t w, int c, float& a0, float& a1, float& a2) {
  const v4f t = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(w, c * 16, 0, 16));
  a0 = t.x;
  a1 = t.y;
  a2 = t.z;
}
template <>
__device__ inline void loadW4<double>(__amdgpu_buffer_rsrc_t w, int c, double& a0, double& a1, double& a2) {
  const v2d t = __builtin_bit_cast(v2d, __builtin_amdgcn_raw_buffer_load_b128(w, c * 32, 0, 16));
  a0 = t.x;
  a1 = t.y;
  a2 = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(w, c * 32 + 16, 0, 16));
}

// runPartial over the grid form's W: the same sums in the same order, every
// column's values loaded (all 16 columns issued) before the products.
template <typename Real, bool kIdx16, typename PP>
__device__ inline void runPartialW4(const RunRaw<Real, kIdx16>& b, __amdgpu_buffer_rsrc_t w, PP P) {
  Real wv[kRun][3];
#pragma unroll
  for (int q = 0; q < kRun; ++q) {
    int c;
    if (kIdx16) {
      const int x = b.c[q >> 3][(q >> 1) & 3];
      c = (q & 1) ? ((unsigned)x >> 16) : (x & 0xffff);
    } else {
      c = b.c[q >> 2][q & 3];
    }
    loadW4<Real>(w, c, wv[q][0], wv[q][1], wv[q][2]);
  }
  Real p0[2] = {0, 0}, p1[2] = {0, 0}, p2[2] = {0, 0};
#pragma unroll
  for (int q = 0; q < kRun; ++q) {
    const Real l = b.v[q];
    p0[q & 1] = fmaT(l, wv[q][0], p0[q & 1]);
    p1[q & 1] = fmaT(l, wv[q][1], p1[q & 1]);
    p2[q & 1] = fmaT(l, wv[q][2], p2[q & 1]);
  }
  P[0] = p0[0] + p0[1];
  P[1] = p1[0] + p1[1];
  P[2] = p2[0] + p2[1];
}

// W of the grid form (4 Real per rank), written by the vertex phases.
template <typename Real, typename WP>
__device__ inline void putW4(WP W, int r, double a, double b, double c) {
  W[4 * r] = (Real)a;
  W[4 * r + 1] = (Real)b;
  W[4 * r + 2] = (Real)c;
}

constexpr long long kFlowSpin = 1LL << 26;  // polls before a flow wait gives up (~2 s at the back-off cap)
constexpr int kFlowStride = 64;                     // ints between two counter lines (256 bytes)
constexpr int kFlowStep = kFlowStride * kFlowLines;  // ints between two step counters
static_assert(kFlowLines >= 8 && kFlowLines <= kWave && (kFlowLines & (kFlowLines - 1)) == 0,
              "a step counter's lines are read one per lane");

// Lines of a step counter in use: one per 4 workgroups of the grid, 8 to
// kFlowLines (a power of two).  Atomics on one address serialise, so a wide grid
// spreads them (512 x 512 mixed32 at 256 workgroups: 1.22 ms per iteration on 8
// lines, 0.99 on 64), while a narrow one polls fewer lines.
__device__ inline int flowLines() {
  int p = 8;
  while (p < kFlowLines && 4 * p < (int)gridDim.x) p *= 2;
  return p;
}

// Adds v to a step counter, on the line of the calling workgroup (workgroup
// index mod the lines in use).
__device__ inline void flowAdd(int* counter, int v) {
  __hip_atomic_fetch_add(counter + (int)(blockIdx.x & (flowLines() - 1)) * kFlowStride, v, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}

// A step counter's value: the sum of its lines (lane l loads line l).  Lines
// only grow, so a sum that reaches a target stays there.
__device__ inline int flowRead(const int* counter) {
  const int lane = threadIdx.x & (kWave - 1);
  int v = lane < flowLines()
              ? __hip_atomic_load(counter + lane * kFlowStride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
              : 0;
#pragma unroll
  for (int o = kFlowLines / 2; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return __builtin_amdgcn_readfirstlane(v);
}

// Waits (the whole wavefront) until *counter >= target.  Most of the grid's
// wavefronts wait on some counter at any time, so the polls back off (s_sleep 1,
// 2, then 4 apart) to keep the counters' memory channels free for the work.
// Bounded: past kFlowSpin polls, or when another wavefront gave up (checked every
// 64 polls), it sets *giveUp and returns false.
__device__ inline bool flowWait(const int* counter, int target, int* giveUp) {
  long long polls = 0;
  int nap = 1;
  for (;;) {
    if (flowRead(counter) >= target) return true;
    ++polls;
    if ((polls & 63) == 0 || polls > kFlowSpin) {
      const int g = __builtin_amdgcn_readfirstlane(__hip_atomic_load(giveUp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      if (g != 0 || polls > kFlowSpin) {
        __hip_atomic_store(giveUp, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return false;
      }
    }
    switch (nap) {  // s_sleep takes an immediate
      case 1: __builtin_amdgcn_s_sleep(1); break;
      case 2: __builtin_amdgcn_s_sleep(2); break;
      default: __builtin_amdgcn_s_sleep(4); break;
    }
    if (nap < 4) nap *= 2;
  }
}

constexpr int kDR = 8;  // rows per wavefront pass of the dense product (4: FP64 config 2 9.32 ms vs 8.72)

template <typename Real>
struct Vec16;  // 16 bytes of Real
template <>
struct Vec16<double> {
  typedef double T __attribute__((ext_vector_type(2)));
  static constexpr int n = 2;
};
template <>
struct Vec16<float> {
  typedef float T __attribute__((ext_vector_type(4)));
  static constexpr int n = 4;
};

// 16 bytes of Real at p (16-byte aligned), in LDS or in global memory.
template <typename Real>
__device__ inline typename Vec16<Real>::T vec16At(const PDX_LDS Real* p) {
  return *(const PDX_LDS typename Vec16<Real>::T*)p;
}
template <typename Real>
__device__ inline typename Vec16<Real>::T vec16At(const PDX_GLOBAL Real* p) {
  return *(const PDX_GLOBAL typename Vec16<Real>::T*)p;
}

constexpr int kLB = 8;  // rows per phase B task of a level sweep (the panels' height)
constexpr int kEU = 4;  // union columns per lane per pass of a dense external task

// A level sweep of the grid form (one step per level of whole elimination-tree
// chains), without grid barriers.  Every step has two phases, each paced by
// counters cumulative over the sweeps of the launch (step st's targets are rep
// times its task counts):
//  * phase A: task t goes to wavefront t mod GW (t counted from the step's first
//    task).  It streams its own runs, two per lane at a time, into its LDS slice
//    (the first pair loads before the wait), after waiting until every task of
//    step st-1 is done (done[st-1], which by induction completes every earlier
//    step; step 0 waits on prev / prevTarget when given).  Each lane subtracts its
//    row's partials in run order; short chains finish with their inverse rows as
//    in the one-workgroup sweep; a long chain's row goes to R.  A wavefront
//    counts its finished tasks of the step in doneA[st] and done[st] at once.
//  * phase B: task j goes to wavefront (nA + j) mod GW, waits until every phase
//    A task of the step is done, and applies 8 rows of a long chain's inverse to
//    R: each lane reads 16 bytes of each row and of R's three components at a
//    time (R through device-scope loads, as it was written by other
//    workgroups), a row's sum is lane-strided partial sums then a butterfly, so
//    the bits do not depend on the grid size.  Counted in done[st] per wavefront.
// W and R cross workgroups as device-scope stores and loads, and a wavefront
// counts itself only once its stores are performed.  Returns false when a wait
// gave up.
// Tasks of level step st, both phases (the counts its done counter reaches per sweep).
__device__ inline int levelStepTasks(const BlockSweepDev& sw, int st) {
  return sw.stepTask[st + 1] - sw.stepTask[st] + sw.stepE[st + 1] - sw.stepE[st] + sw.stepB[st + 1] - sw.stepB[st];
}

template <typename Real, bool kFwd, bool kIdx16, typename Store>
__device__ bool levelSweep(const BlockSweepDev& sw, int m, const Real* dinv, Real* Wp, Real* Rp, PDX_LDS Real* Pw,
                           Store store, int* done, int* doneA, int rep, const int* prev, int prevTarget, int* giveUp) {
  typedef typename Vec16<Real>::T VT;
  constexpr int V = Vec16<Real>::n;
  const int lane = threadIdx.x & (kWave - 1);
  const int nw = blockDim.x / kWave;
  const int gw = blockIdx.x * nw + (int)(threadIdx.x / kWave);
  const int GW = gridDim.x * nw;
  const __amdgpu_buffer_rsrc_t wsrc = bufferOver(Wp, 4 * m * (int)sizeof(Real));
  const Real* rval = static_cast<const Real*>(sw.runVal);
  const PDX_GLOBAL Real* gdense = gview<Real>(sw.dense);
  const PDX_GLOBAL Real* gx = gview<Real>(sw.xpanel);
  const PDX_GLOBAL Real* gep = gview<Real>(sw.epanel);
  const PDX_GLOBAL int* gucol = gview<int>(sw.ucol);
  const PDX_GLOBAL v4i* gmeta = gview<v4i>(sw.laneMeta);
  const PDX_GLOBAL Real* gldinv = gview<Real>(sw.laneDinv);
  const int rT = sw.rTotal;
  // device-scope (sc1) 16-byte loads of R through a buffer resource over its 3 x rTotal values
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc(Rp, (short)0, 3 * rT * (int)sizeof(Real), 0x00020000);
  for (int st = 0; st < sw.nstep; ++st) {
    const int ta = sw.stepTask[st], tb = sw.stepTask[st + 1];
    const int ja = sw.stepB[st], jb = sw.stepB[st + 1];
    const int ea = sw.stepE[st], eb = sw.stepE[st + 1];
    const int nA = tb - ta, nB = jb - ja, nE = eb - ea;
    const int* prevOn = st > 0 ? done + (size_t)(st - 1) * kFlowStep : prev;
    const int prevFor = st > 0 ? rep * levelStepTasks(sw, st - 1) : prevTarget;
    const int dbase = sw.stepDense[st];
    int* doneSt = done + (size_t)st * kFlowStep;
    int* doneASt = doneA + (size_t)st * kFlowStep;
    bool waited = false;
    int mine = 0;  // tasks of this phase the wavefront finished
    for (int t = ta + gw; t < tb; t += GW) {
      const v4i mt = gmeta[(size_t)t * kWave + lane];
      const Real dv = kFwd ? Real(0) : gldinv[(size_t)t * kWave + lane];
      const int doff = gview<int>(sw.laneDense)[(size_t)t * kWave + lane];
      const int r0 = sw.taskRun[t], nr = sw.taskRun[t + 1] - r0;
      const int k = mt.x;
      const int info = mt.w;
      const bool act = k >= 0;
      const bool toR = (info >> 28) & 1;
      // issued before the step's wait, as neither depends on it: (FP32) the
      // lane's row of its short chain's inverse, into registers (FP64 reads it
      // as it finishes: its registers would spill), and, in the forward sweep,
      // the row's right-hand side (final since the barrier before the sweep; a
      // backward row's value is the forward sweep's, which only the wait orders)
      constexpr bool kRegRows = sizeof(Real) == 4;
      Real xc[kRegRows ? kMaxChain : 1];
      if constexpr (kRegRows) loadChainRow<Real>(gdense + dbase, mt, toR ? 0 : doff, xc);
      Real a0 = 0, a1 = 0, a2 = 0;
      if (kFwd && act) loadW4<Real>(wsrc, k, a0, a1, a2);
      for (int q0 = 0; q0 < nr; q0 += 2 * kWave) {
        RunRaw<Real, kIdx16> raw[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int q = q0 + u * kWave + lane;
          loadRun<Real, kIdx16>(sw.runCol, rval, r0 + (q < nr ? q : 0), raw[u]);
        }
        if (!waited) {
          if (prevOn && !flowWait(prevOn, prevFor, giveUp)) return false;
          waited = true;
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int q = q0 + u * kWave + lane;
          if (q < nr) runPartialW4<Real, kIdx16>(raw[u], wsrc, Pw + 3 * q);
        }
      }
      if (!waited) {
        if (prevOn && !flowWait(prevOn, prevFor, giveUp)) return false;
        waited = true;
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (act) {
        if (!kFwd) {
          loadW4<Real>(wsrc, k, a0, a1, a2);
          a0 = a0 * dv;
          a1 = a1 * dv;
          a2 = a2 * dv;
        }
        subtractPartials<Real>(Pw, mt.y, mt.z, a0, a1, a2);
      }
      if constexpr (kRegRows)
        segmentFinishR<Real>(xc, info, lane, a0, a1, a2);
      else
        segmentFinish<Real>(gdense + dbase + (toR ? 0 : doff), info, lane, act, a0, a1, a2);
      if (act) {
        if (toR) {
          agentStore(Rp + doff, a0);
          agentStore(Rp + rT + doff, a1);
          agentStore(Rp + 2 * rT + doff, a2);
        } else {
          agentStore(Wp + 4 * k + 0, a0);
          agentStore(Wp + 4 * k + 1, a1);
          agentStore(Wp + 4 * k + 2, a2);
          store(k, a0, a1, a2);
        }
      }
      ++mine;
      // the LDS slice is this wavefront's; the next task may overwrite it
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    // dense tasks of phase A (a dense chain's 8 rows): r = base - E W[U], each
    // lane reading the W values of its columns of U once and the 8 rows' entries
    // of each (column-major panel), a row's sum lane-strided then a butterfly
    {
      int j = gw - nA % GW;
      if (j < 0) j += GW;
      for (; j < nE; j += GW) {
        const int* et = sw.eTask + 8 * (size_t)(ea + j);
        const int eoff = __builtin_amdgcn_readfirstlane(et[0]);
        const int c = __builtin_amdgcn_readfirstlane(et[1]);
        const int uoff = __builtin_amdgcn_readfirstlane(et[2]);
        const int kr = __builtin_amdgcn_readfirstlane(et[3]);
        const int rp = __builtin_amdgcn_readfirstlane(et[4]);
        const int k0 = kr & 0xffffff, nrows = kr >> 24;
        Real acc[kLB][3];
#pragma unroll
        for (int q = 0; q < kLB; ++q) acc[q][0] = acc[q][1] = acc[q][2] = Real(0);
        // kEU columns per lane per pass: their column indices and panel entries,
        // then their W values, all in flight before the sums, which take the
        // columns in the lane's order (lane, lane + 64, ...) as one at a time
        // would.  The first pass's indices and entries do not depend on the
        // step's wait and are issued before it (by every lane, clamped, so that
        // the whole wavefront waits together).
        int col[kEU];
        VT ev[kEU][kLB / V];
        auto loadPass = [&](int jb) {
#pragma unroll
          for (int u = 0; u < kEU; ++u) {
            const int jj = min(jb + u * kWave, c - 1);
            col[u] = gucol[uoff + jj];
            const PDX_GLOBAL VT* ep = (const PDX_GLOBAL VT*)(gep + (size_t)eoff + (size_t)jj * kLB);
#pragma unroll
            for (int h = 0; h < kLB / V; ++h) ev[u][h] = ep[h];
          }
        };
        loadPass(lane);
        if (!waited) {
          if (prevOn && !flowWait(prevOn, prevFor, giveUp)) return false;
          waited = true;
        }
        for (int jb = lane; jb < c; jb += kEU * kWave) {
          if (jb != lane) loadPass(jb);
          Real w[kEU][3];
#pragma unroll
          for (int u = 0; u < kEU; ++u) loadW4<Real>(wsrc, col[u], w[u][0], w[u][1], w[u][2]);
#pragma unroll
          for (int u = 0; u < kEU; ++u)
            if (jb + u * kWave < c) {
#pragma unroll
              for (int q = 0; q < kLB; ++q) {
                const Real e = ev[u][q / V][q % V];
                acc[q][0] = fmaT(e, w[u][0], acc[q][0]);
                acc[q][1] = fmaT(e, w[u][1], acc[q][1]);
                acc[q][2] = fmaT(e, w[u][2], acc[q][2]);
              }
            }
        }
#pragma unroll
        for (int q = 0; q < kLB; ++q)
#pragma unroll
          for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int o = kWave / 2; o >= 1; o >>= 1) acc[q][d] += __shfl_xor(acc[q][d], o);
        Real s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
        for (int q = 0; q < kLB; ++q)
          if (lane == q) {
            s0 = acc[q][0];
            s1 = acc[q][1];
            s2 = acc[q][2];
          }
        if (lane < nrows) {
          const int k = k0 + lane;
          Real a0, a1, a2;
          loadW4<Real>(wsrc, k, a0, a1, a2);
          if (!kFwd) {
            const Real dv = dinv[k];
            a0 = a0 * dv;
            a1 = a1 * dv;
            a2 = a2 * dv;
          }
          agentStore(Rp + rp + lane, a0 - s0);
          agentStore(Rp + rT + rp + lane, a1 - s1);
          agentStore(Rp + 2 * rT + rp + lane, a2 - s2);
        }
        ++mine;
      }
    }
    // the wavefront's tasks of the step count once their stores are performed
    // (every wait is for a step's or phase's total, so one count per wavefront
    // and phase orders the same)
    if (mine > 0) {
      storesPerformed();
      if (lane == 0) {
        flowAdd(doneASt, mine);
        flowAdd(doneSt, mine);
      }
    }
    int j = gw - (nA + nE) % GW;
    if (j < 0) j += GW;
    bool waitedB = false;
    int mineB = 0;
    for (; j < nB; j += GW) {
      const int* bt = sw.bTask + 4 * (size_t)(ja + j);
      const int xoff = __builtin_amdgcn_readfirstlane(bt[0]);
      const int ld = __builtin_amdgcn_readfirstlane(bt[1]);
      const int rp = __builtin_amdgcn_readfirstlane(bt[2]);
      const int kr = __builtin_amdgcn_readfirstlane(bt[3]);
      const int k0 = kr & 0xffffff, nrows = kr >> 24;
      const PDX_GLOBAL VT* row[kLB];
#pragma unroll
      for (int q = 0; q < kLB; ++q)  // rows past the panel repeat its last row (their sums are dropped)
        row[q] = (const PDX_GLOBAL VT*)(gx + xoff + (size_t)min(q, nrows - 1) * ld);
      // the panel's first vectors do not depend on the wait: issued before it
      const int nv = ld / V;
      VT lv[kLB];
#pragma unroll
      for (int q = 0; q < kLB; ++q) lv[q] = row[q][min(lane, nv - 1)];
      if (!waitedB) {
        if (!flowWait(doneASt, rep * (nA + nE), giveUp)) return false;
        waitedB = true;
      }
      Real acc[kLB][3];
#pragma unroll
      for (int q = 0; q < kLB; ++q) acc[q][0] = acc[q][1] = acc[q][2] = Real(0);
      for (int jv = lane; jv < nv; jv += kWave) {
        if (jv != lane)
#pragma unroll
          for (int q = 0; q < kLB; ++q) lv[q] = row[q][jv];
        const int off = (rp + jv * V) * (int)sizeof(Real);
        const VT w0 = __builtin_bit_cast(VT, __builtin_amdgcn_raw_buffer_load_b128(rsrc, off, 0, 16));
        const VT w1 = __builtin_bit_cast(VT, __builtin_amdgcn_raw_buffer_load_b128(rsrc, off + rT * (int)sizeof(Real), 0, 16));
        const VT w2 =
            __builtin_bit_cast(VT, __builtin_amdgcn_raw_buffer_load_b128(rsrc, off + 2 * rT * (int)sizeof(Real), 0, 16));
#pragma unroll
        for (int u = 0; u < V; ++u)
#pragma unroll
          for (int q = 0; q < kLB; ++q) {
            const Real l = lv[q][u];
            acc[q][0] = fmaT(l, w0[u], acc[q][0]);
            acc[q][1] = fmaT(l, w1[u], acc[q][1]);
            acc[q][2] = fmaT(l, w2[u], acc[q][2]);
          }
      }
#pragma unroll
      for (int q = 0; q < kLB; ++q)
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int o = kWave / 2; o >= 1; o >>= 1) acc[q][d] += __shfl_xor(acc[q][d], o);
      Real y0 = 0, y1 = 0, y2 = 0;
#pragma unroll
      for (int q = 0; q < kLB; ++q)
        if (lane == q) {
          y0 = acc[q][0];
          y1 = acc[q][1];
          y2 = acc[q][2];
        }
      if (lane < nrows) {
        const int k = k0 + lane;
        agentStore(Wp + 4 * k + 0, y0);
        agentStore(Wp + 4 * k + 1, y1);
        agentStore(Wp + 4 * k + 2, y2);
        store(k, y0, y1, y2);
      }
      ++mineB;
    }
    if (mineB > 0) {
      storesPerformed();
      if (lane == 0) flowAdd(doneSt, mineB);
    }
  }
  return true;
}

// Dense solve: out_r = sum_j Linv[r][j] W_j for the three right-hand sides, for
// every rank r, rows spread over the grid's wavefronts (global wavefront gw of GW),
// kDR rows per pass so each W value loaded serves kDR rows.  Each lane reads 16
// bytes of a row at a time (rows are ld Real apart, ld a multiple of 4, padding
// zero).  A row's sum is formed by one wavefront in a fixed order (lane-strided
// partial sums, then a butterfly), so it does not depend on the grid size.
// store(r, a0, a1, a2) takes each row.
template <typename Real, typename WP, typename Store>
__device__ PDX_INLINE void denseRows(const PDX_GLOBAL Real* L, int m, int ld, WP W, int gw, int GW, Store store) {
  typedef typename Vec16<Real>::T VT;
  constexpr int V = Vec16<Real>::n;
  const int lane = threadIdx.x & (kWave - 1);
  const int nv = ld / V;  // vectors per row
  for (int r0 = gw * kDR; r0 < m; r0 += GW * kDR) {
    const PDX_GLOBAL VT* row[kDR];
#pragma unroll
    for (int q = 0; q < kDR; ++q)  // rows past m repeat row m-1 (their sums are dropped)
      row[q] = (const PDX_GLOBAL VT*)(L + (size_t)min(r0 + q, m - 1) * ld);
    Real acc[kDR][3];
#pragma unroll
    for (int q = 0; q < kDR; ++q) acc[q][0] = acc[q][1] = acc[q][2] = Real(0);
#pragma unroll 2
    for (int jv = lane; jv < nv; jv += kWave) {
      VT lv[kDR];
#pragma unroll
      for (int q = 0; q < kDR; ++q) lv[q] = row[q][jv];
      // the V columns' 3V values of W are 48 contiguous bytes from a 16-byte
      // boundary: three 16-byte loads, or scalar ones at the end (padding
      // columns: zero in L, outside W)
      Real wv[3 * V];
      if (jv * V + V <= m) {
#pragma unroll
        for (int h = 0; h < 3; ++h) {
          const VT t = vec16At<Real>(W + 3 * V * jv + h * V);
#pragma unroll
          for (int u = 0; u < V; ++u) wv[h * V + u] = t[u];
        }
      } else {
#pragma unroll
        for (int i = 0; i < 3 * V; ++i) wv[i] = 3 * V * jv + i < 3 * m ? W[3 * V * jv + i] : Real(0);
      }
#pragma unroll
      for (int u = 0; u < V; ++u) {
        const Real w0 = wv[3 * u], w1 = wv[3 * u + 1], w2 = wv[3 * u + 2];
#pragma unroll
        for (int q = 0; q < kDR; ++q) {
          const Real l = lv[q][u];
          acc[q][0] = fmaT(l, w0, acc[q][0]);
          acc[q][1] = fmaT(l, w1, acc[q][1]);
          acc[q][2] = fmaT(l, w2, acc[q][2]);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < kDR; ++q)
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int o = kWave / 2; o >= 1; o >>= 1) acc[q][d] += __shfl_xor(acc[q][d], o);
#pragma unroll
    for (int q = 0; q < kDR; ++q)
      if (lane == q && r0 + q < m) store(r0 + q, acc[q][0], acc[q][1], acc[q][2]);
  }
}

// One dense solve inside the grid form: W (complete, rank space) -> store, then a
// grid barrier.  W is first copied into each workgroup's LDS when it fits.
// Row blocks are dealt to workgroups first (wavefront w of workgroup b takes row
// block w * gridDim + b, then every GW-th), so every CU streams rows even when
// there are fewer row blocks than wavefronts.
__device__ inline int denseWave() { return (int)(threadIdx.x / kWave) * (int)gridDim.x + (int)blockIdx.x; }

template <typename Real, typename Store>
__device__ PDX_INLINE void gridDense(cg::grid_group& grid, const InstDev& I, PDX_GLOBAL Real* W, Store store) {
  const int nw = blockDim.x / kWave;
  const int gw = denseWave();
  const int GW = gridDim.x * nw;
  const PDX_GLOBAL Real* L = (const PDX_GLOBAL Real*)I.Linv;
  if (I.denseLdsW) {
    PDX_LDS Real* Ws = ldsAt<Real>(32 * sizeof(double));
    for (int i = threadIdx.x; i < 3 * I.m; i += blockDim.x) Ws[i] = W[i];
    __syncthreads();
    denseRows<Real>(L, I.m, I.denseLd, Ws, gw, GW, store);
  } else {
    denseRows<Real>(L, I.m, I.denseLd, (const PDX_GLOBAL Real*)W, gw, GW, store);
  }
  grid.sync();
}

template <typename Real, bool kExact, bool kIdx16, bool kDense>
__global__ void __launch_bounds__(512) gridSubstepKernel(const InstDev* __restrict__ insts, int inst, int nsub,
                                                          double* __restrict__ gslots, int waveRuns) {
  extern __shared__ double smem[];
  cg::grid_group grid = cg::this_grid();
  const InstDev& I = insts[inst];
  double* red = smem;  // 2 x 16 slots
  PDX_LDS Real* Pw = ldsAt<Real>(32 * sizeof(double)) + (size_t)(threadIdx.x / kWave) * waveRuns * 3;
  PDX_GLOBAL Real* W = (PDX_GLOBAL Real*)I.Wg;
  const int n = I.n;
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int gs = gridDim.x * blockDim.x;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  // W: 3 Real per rank for the dense product, 4 for the level sweeps
  auto wput = [=](int r, double a, double b, double c) {
    if (kDense)
      putW<Real>(W, r, a, b, c);
    else
      putW4<Real>(W, r, a, b, c);
  };
  int buf = 0;
  // the level sweeps' step counters start at zero in every launch (ordered before
  // the first sweep by the grid barrier after the first right-hand side)
  const int nF = I.bfwd.nstep, nB = I.bbwd.nstep;
  if (!kDense && blockIdx.x == 0)
    for (int i = threadIdx.x; i < 2 * (nF + nB) * kFlowLines; i += blockDim.x) I.flow[(size_t)i * kFlowStride] = 0;
  int rep = 0;  // sweeps of this launch
  const int lastF = nF > 0 ? levelStepTasks(I.bfwd, nF - 1) : 0;
  // stage buckets (gather, solve, -, predict + velocity, criterion), read by
  // thread 0 of workgroup 0 at points the grid's barriers align; the level
  // sweeps overlap forward and backward across workgroups, so the solve is one
  // bucket
  long long tacc[5] = {0, 0, 0, 0, 0};
  for (int sub = 0; sub < nsub; ++sub) {
    // predict (Integrator.cpp:447-463), displacement scale, first right-hand side
    const long long tp0 = lead ? tickStage() : 0;
    double lscale = 0.0;
    for (int v = gt; v < n; v += gs) {
      for (int d = 0; d < 3; ++d) {
        const double p = I.x[3 * v + d];
        I.xbs[3 * v + d] = p;
        const double xh = p + I.vel[3 * v + d] * I.h + I.gh2[d];
        I.xhat[3 * v + d] = xh;
        I.prev[3 * v + d] = p;
        lscale = fmax(lscale, fabs(xh - p));
      }
    }
    const long long tp1 = lead ? tickStage() : 0;
    for (int v = gt; v < n; v += gs) {
      const int r = I.vrank[v];
      if (kExact) {
        const Acc3 b = rhsExact(I, v);
        wput(r, b.x, b.y, b.z);
      } else if (I.pinned[v]) {
        wput(r, I.pinpos[3 * v] - I.x[3 * v], I.pinpos[3 * v + 1] - I.x[3 * v + 1],
             I.pinpos[3 * v + 2] - I.x[3 * v + 2]);
      } else {
        Acc3 rh;
        (void)residualAndRhs(I, v, &rh);
        wput(r, rh.x, rh.y, rh.z);
      }
    }
    double scale = gridMax(grid, lscale, red, gslots, buf);
    buf ^= 1;
    scale = fmax(scale, I.gScaleH2);
    if (lead) {
      tacc[3] += tp1 - tp0;
      tacc[0] += tickStage() - tp1;
    }

    int used = 0, conv = 0;
    double lastRes = -1.0;
    auto none = [](int, Real, Real, Real) {};
    double* x = I.x;
    const int* rankv = I.rankv;
    const uint8_t* pinnedP = I.pinned;
    const double* pinposP = I.pinpos;
    // the backward sweep's rows finish the iterate: the solution itself (exact
    // right-hand side), or x + the FP32 update (pinned rows take q)
    auto put = [=](int k, Real a0, Real a1, Real a2) {
      const int v = rankv[k];
      if (kExact) {
        x[3 * v] = (double)a0;
        x[3 * v + 1] = (double)a1;
        x[3 * v + 2] = (double)a2;
      } else if (pinnedP[v]) {
        x[3 * v] = pinposP[3 * v];
        x[3 * v + 1] = pinposP[3 * v + 1];
        x[3 * v + 2] = pinposP[3 * v + 2];
      } else {
        x[3 * v] += (double)a0;
        x[3 * v + 1] += (double)a1;
        x[3 * v + 2] += (double)a2;
      }
    };
    for (int k = 0; k < I.maxIter; ++k) {
      const long long ts0 = lead ? tickStage() : 0;
      if (kDense) {
        if (kExact) {
          gridDense<Real>(grid, I, W, put);
        } else {
          // the update is applied as each row finishes: the product reads only W
          const uint8_t* pinned = I.pinned;
          const double* pinpos = I.pinpos;
          auto update = [=](int r, Real a0, Real a1, Real a2) {
            const int v = rankv[r];
            if (pinned[v]) {
              x[3 * v] = pinpos[3 * v];
              x[3 * v + 1] = pinpos[3 * v + 1];
              x[3 * v + 2] = pinpos[3 * v + 2];
            } else {
              x[3 * v] += (double)a0;
              x[3 * v + 1] += (double)a1;
              x[3 * v + 2] += (double)a2;
            }
          };
          gridDense<Real>(grid, I, W, update);
        }
      } else {
        // level sweeps paced by step counters; the backward rows update x as they
        // finish, and one grid barrier follows
        ++rep;
        int* doneF = I.flow;
        int* doneB = I.flow + (size_t)2 * nF * kFlowStep;
        Real* R = static_cast<Real*>(I.flowPart);
        if (levelSweep<Real, true, kIdx16>(I.bfwd, I.m, static_cast<const Real*>(I.dinv), (Real*)W, R, Pw, none, doneF, doneF + (size_t)nF * kFlowStep,
                                           rep, nullptr, 0, I.pair.giveUp))
          levelSweep<Real, false, kIdx16>(I.bbwd, I.m, static_cast<const Real*>(I.dinv), (Real*)W, R, Pw, put, doneB, doneB + (size_t)nB * kFlowStep, rep,
                                           nF > 0 ? doneF + (size_t)(nF - 1) * kFlowStep : nullptr, rep * lastF,
                                           I.pair.giveUp);
        grid.sync();
      }
      const long long ts1 = lead ? tickStage() : 0;
      const bool more = (k + 1 < I.maxIter);
      double ldiff = 0.0, lres = 0.0;
      for (int v = gt; v < n; v += gs) {
        const double x0 = I.x[3 * v], x1 = I.x[3 * v + 1], x2 = I.x[3 * v + 2];
        const double d = fmax(fabs(x0 - I.prev[3 * v]),
                              fmax(fabs(x1 - I.prev[3 * v + 1]), fabs(x2 - I.prev[3 * v + 2])));
        ldiff = fmax(ldiff, d);
        I.prev[3 * v] = x0;
        I.prev[3 * v + 1] = x1;
        I.prev[3 * v + 2] = x2;
        const bool pinnedV = I.pinned[v] != 0;
        const int r = I.vrank[v];
        if (kExact) {
          if (I.wantResidual && !pinnedV && I.mass[v] > 0.0) lres = fmax(lres, residualAndRhs(I, v, nullptr));
          if (more) {
            const Acc3 b = rhsExact(I, v);
            wput(r, b.x, b.y, b.z);
          }
        } else {
          if (pinnedV) {
            if (more) wput(r, I.pinpos[3 * v] - x0, I.pinpos[3 * v + 1] - x1, I.pinpos[3 * v + 2] - x2);
          } else if (more || I.wantResidual) {
            Acc3 rh;
            const double rv = residualAndRhs(I, v, more ? &rh : nullptr);
            if (I.wantResidual && I.mass[v] > 0.0) lres = fmax(lres, rv);
            if (more) wput(r, rh.x, rh.y, rh.z);
          }
        }
      }
      const long long ts2 = lead ? tickStage() : 0;
      double diff = 0.0, residual = -1.0;
      if (I.wantResidual)
        gridMax2(grid, ldiff, lres, red, gslots, buf, diff, residual);
      else
        diff = gridMax(grid, ldiff, red, gslots, buf);
      buf ^= 1;
      if (lead) {
        tacc[1] += ts1 - ts0;
        tacc[0] += ts2 - ts1;
        tacc[4] += tickStage() - ts2;
      }
      const double rel = (scale > 0.0) ? diff / scale : diff;
      const bool absHit = (I.absTol > 0.0) && (diff <= I.absTol);
      const bool relHit = (I.relTol > 0.0) && (rel <= I.relTol);
      const bool resHit = (I.resTol > 0.0) && (residual <= I.resThresh);
      const bool c = (I.resTol > 0.0) ? resHit : (absHit || relHit);
      used = k + 1;
      lastRes = residual;
      if (c) {
        conv = 1;
        break;
      }
    }
    // velocity + damping + re-pin (Integrator.cpp:774-791)
    const long long tv0 = lead ? tickStage() : 0;
    for (int v = gt; v < n; v += gs) {
      const bool pinnedV = I.pinned[v] != 0;
      for (int d = 0; d < 3; ++d) {
        if (pinnedV) {
          I.vel[3 * v + d] = 0.0;
          I.x[3 * v + d] = I.pinpos[3 * v + d];
        } else {
          I.vel[3 * v + d] = (I.x[3 * v + d] - I.xbs[3 * v + d]) * I.dampInvH;
        }
      }
    }
    if (lead) {
      tacc[3] += tickStage() - tv0;
      I.istats[0] = used;
      I.istats[1] = conv;
      I.istats[2] += conv;
      I.lstats[0] += used;
      I.lstats[1] += 1;
      I.dstats[0] = lastRes;
    }
    grid.sync();
  }
  if (lead)
    for (int q = 0; q < 5; ++q) I.lstats[2 + q] += tacc[q];
}

// Global solve x = L^{-1} b on one instance with the whole grid.
template <typename Real, bool kIdx16>
__global__ void __launch_bounds__(512) gridSolveKernel(const InstDev* __restrict__ insts, int inst,
                                                        const double* __restrict__ b, double* __restrict__ xout,
                                                        int waveRuns) {
  extern __shared__ double smem[];
  cg::grid_group grid = cg::this_grid();
  const InstDev& I = insts[inst];
  PDX_LDS Real* Pw = ldsAt<Real>(32 * sizeof(double)) + (size_t)(threadIdx.x / kWave) * waveRuns * 3;
  PDX_GLOBAL Real* W = (PDX_GLOBAL Real*)I.Wg;
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int gs = gridDim.x * blockDim.x;
  for (int v = gt; v < I.n; v += gs) putW4<Real>(W, I.vrank[v], b[3 * v], b[3 * v + 1], b[3 * v + 2]);
  const int nF = I.bfwd.nstep, nB = I.bbwd.nstep;
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < 2 * (nF + nB) * kFlowLines; i += blockDim.x) I.flow[(size_t)i * kFlowStride] = 0;
  grid.sync();
  auto none = [](int, Real, Real, Real) {};
  const int* rankv = I.rankv;
  auto put = [=](int k, Real a0, Real a1, Real a2) {
    const int v = rankv[k];
    xout[3 * v] = (double)a0;
    xout[3 * v + 1] = (double)a1;
    xout[3 * v + 2] = (double)a2;
  };
  const int lastF = nF > 0 ? levelStepTasks(I.bfwd, nF - 1) : 0;
  int* doneF = I.flow;
  int* doneB = I.flow + (size_t)2 * nF * kFlowStep;
  Real* R = static_cast<Real*>(I.flowPart);
  if (levelSweep<Real, true, kIdx16>(I.bfwd, I.m, static_cast<const Real*>(I.dinv), (Real*)W, R, Pw, none, doneF, doneF + (size_t)nF * kFlowStep, 1,
                                     nullptr, 0, I.pair.giveUp))
    levelSweep<Real, false, kIdx16>(I.bbwd, I.m, static_cast<const Real*>(I.dinv), (Real*)W, R, Pw, put, doneB, doneB + (size_t)nB * kFlowStep, 1,
                                     nF > 0 ? doneF + (size_t)(nF - 1) * kFlowStep : nullptr, lastF, I.pair.giveUp);
}

// Dense solve mode, x = Linv b: W = P b (rank space) ...
template <typename Real>
__global__ void __launch_bounds__(512) densePermuteKernel(const InstDev* __restrict__ insts, int inst,
                                                          const double* __restrict__ b) {
  const InstDev& I = insts[inst];
  PDX_GLOBAL Real* W = (PDX_GLOBAL Real*)I.Wg;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < I.n; v += gridDim.x * blockDim.x)
    putW<Real>(W, I.vrank[v], b[3 * v], b[3 * v + 1], b[3 * v + 2]);
}

// ... then the rows of Linv W, each written to its vertex.
template <typename Real>
__global__ void __launch_bounds__(512) denseSolveKernel(const InstDev* __restrict__ insts, int inst,
                                                        double* __restrict__ xout) {
  const InstDev& I = insts[inst];
  const int nw = blockDim.x / kWave;
  const int* rankv = I.rankv;
  auto put = [=](int r, Real a0, Real a1, Real a2) {
    const int v = rankv[r];
    xout[3 * v] = (double)a0;
    xout[3 * v + 1] = (double)a1;
    xout[3 * v + 2] = (double)a2;
  };
  denseRows<Real>((const PDX_GLOBAL Real*)I.Linv, I.m, I.denseLd, (const PDX_GLOBAL Real*)I.Wg, denseWave(),
                  gridDim.x * nw, put);
}

template <typename Real, bool kBlocked, bool kIdx16, typename WP>
__device__ void solveBody(const InstDev& I, WP W, const double* __restrict__ b, double* __restrict__ xout);

// Stand-alone global solve x = L^{-1} b (EigenDirectSolver::solve) on one instance:
// one workgroup, or two for a split instance (workgroup 1 runs the partner's
// descriptor inst2).
template <typename Real, bool kBlocked, bool kIdx16, bool kWL>
__global__ void __launch_bounds__(512) solveKernel(const InstDev* __restrict__ insts, int inst, int inst2,
                                                    const double* __restrict__ b,
                                                    double* __restrict__ xout) {
  const InstDev& I = insts[blockIdx.x == 0 ? inst : inst2];
  if (kWL)
    solveBody<Real, kBlocked, kIdx16>(I, ldsAt<Real>(32 * sizeof(double)), b, xout);
  else
    solveBody<Real, kBlocked, kIdx16>(I, (PDX_GLOBAL Real*)I.Wg, b, xout);
}

template <typename Real, bool kBlocked, bool kIdx16, typename WP>
__device__ void solveBody(const InstDev& I, WP W, const double* __restrict__ b, double* __restrict__ xout) {
  if (kBlocked) {
    loadStepTables(I);
    clearPartnerPad<Real>(I, W);
  }
  const int nv = ownCount(I);
  for (int i = threadIdx.x; i < nv; i += blockDim.x) {
    const int v = ownVertex(I, i);
    const int r = I.vrank[v];
    putW<Real>(W, r, b[3 * v], b[3 * v + 1], b[3 * v + 2]);
  }
  __syncthreads();
  const Real* dinv = static_cast<const Real*>(I.dinv);
  SweepTimes tst;
  auto none = [](int, Real, Real, Real) {};
  if (kBlocked)
    blockSweep<Real, true, kIdx16>(I.bfwd, W, ldsAt<Real>(I.partOff), ldsAt<Real>(I.denseOff), stepTables(I, 0), none, tst, I.ldsW != 0);
  else
    sweep<Real, true>(I.fwd, W, dinv, none);
  if constexpr (kBlocked) {
    if (I.pairRole != 0) {
      // the halves meet at the separator, as in the substep (pairSync)
      extern __shared__ double smem[];
      int* ok = (int*)(smem + 31);
      int seq = __hip_atomic_load(I.pair.flags + (I.pairRole == 1 ? 0 : 32), __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
      pairExchange<Real>(I, W);
      if (!pairSync(I, ++seq, true, ok)) return;
      pairTake<Real>(I, W);
      __syncthreads();
      runSeparator<Real, kIdx16>(I, W, tst);
    }
  }
  if (kBlocked)
    blockSweep<Real, false, kIdx16>(I.bbwd, W, ldsAt<Real>(I.partOff), ldsAt<Real>(I.denseOff), stepTables(I, 1), none, tst, I.ldsW != 0);
  else
    sweep<Real, false>(I.bwd, W, dinv, none);
  __syncthreads();
  // the solution of the workgroup's own vertices, from rank space
  for (int i = threadIdx.x; i < nv; i += blockDim.x) {
    const int v = ownVertex(I, i);
    const int r = I.vrank[v];
    xout[3 * v] = (double)W[3 * r];
    xout[3 * v + 1] = (double)W[3 * r + 1];
    xout[3 * v + 2] = (double)W[3 * r + 2];
  }
}

}  // namespace

#define PDX_KERNEL_TABLE(X)                                                                  \
  X(float, false, true, true) X(float, false, true, false) X(float, false, false, false)   \
  X(double, true, true, true) X(double, true, true, false) X(double, true, false, false)

hipError_t launchSubstep(const InstDev* d, const int* order, int ninst, int nsub, int block,
                         size_t lds, bool fp32, bool blocked, bool idx16, bool wl, hipStream_t s, int subBase,
                         int flags) {
#define PDX_L(R, E, B, X16)                                                                              \
  if (fp32 == (sizeof(R) == 4) && blocked == B && (idx16 && blocked) == X16) {                           \
    if (flags && wl)                                                                                     \
      hipLaunchKernelGGL((substepKernel<R, E, B, X16, true, true>), dim3(ninst), dim3(block), lds, s, d, \
                         order, nsub, subBase, flags);                                                   \
    else if (flags)                                                                                      \
      hipLaunchKernelGGL((substepKernel<R, E, B, X16, true, false>), dim3(ninst), dim3(block), lds, s, d, \
                         order, nsub, subBase, flags);                                                   \
    else if (wl)                                                                                         \
      hipLaunchKernelGGL((substepKernel<R, E, B, X16, false, true>), dim3(ninst), dim3(block), lds, s, d, \
                         order, nsub, subBase, 0);                                                       \
    else                                                                                                 \
      hipLaunchKernelGGL((substepKernel<R, E, B, X16, false, false>), dim3(ninst), dim3(block), lds, s,  \
                         d, order, nsub, subBase, 0);                                                    \
    return hipGetLastError();                                                                            \
  }
  PDX_KERNEL_TABLE(PDX_L)
#undef PDX_L
  return hipErrorInvalidValue;
}

hipError_t launchAdjoint(const InstDev* d, const int* order, int ninst, int nsub, int subBase, int block,
                         size_t lds, bool fp32, bool blocked, bool idx16, bool wl, hipStream_t s) {
#define PDX_L(R, E, B, X16)                                                                              \
  if (fp32 == (sizeof(R) == 4) && blocked == B && (idx16 && blocked) == X16) {                           \
    if (wl)                                                                                              \
      hipLaunchKernelGGL((adjointKernel<R, B, X16, true>), dim3(ninst), dim3(block), lds, s, d, order,   \
                         nsub, subBase);                                                                 \
    else                                                                                                 \
      hipLaunchKernelGGL((adjointKernel<R, B, X16, false>), dim3(ninst), dim3(block), lds, s, d, order,  \
                         nsub, subBase);                                                                 \
    return hipGetLastError();                                                                            \
  }
  PDX_KERNEL_TABLE(PDX_L)
#undef PDX_L
  return hipErrorInvalidValue;
}

hipError_t launchGrid(const InstDev* d, int inst, int nsub, double* gslots, int waveRuns, int blocks,
                      int block, size_t lds, bool fp32, bool idx16, hipStream_t s) {
  void* args[] = {(void*)&d, (void*)&inst, (void*)&nsub, (void*)&gslots, (void*)&waveRuns};
#define PDX_G(R, E, X16)                                                                                 \
  if (fp32 == (sizeof(R) == 4) && idx16 == X16)                                                          \
    return hipLaunchCooperativeKernel((const void*)gridSubstepKernel<R, E, X16, false>, dim3(blocks),     \
                                      dim3(block), args, lds, s);
  PDX_G(float, false, true) PDX_G(float, false, false) PDX_G(double, true, true) PDX_G(double, true, false)
#undef PDX_G
  return hipErrorInvalidValue;
}

hipError_t launchDense(const InstDev* d, int inst, int nsub, double* gslots, int blocks, int block, size_t lds,
                       bool fp32, hipStream_t s) {
  int waveRuns = 0;
  void* args[] = {(void*)&d, (void*)&inst, (void*)&nsub, (void*)&gslots, (void*)&waveRuns};
  if (fp32)
    return hipLaunchCooperativeKernel((const void*)gridSubstepKernel<float, false, false, true>, dim3(blocks),
                                      dim3(block), args, lds, s);
  return hipLaunchCooperativeKernel((const void*)gridSubstepKernel<double, true, false, true>, dim3(blocks),
                                    dim3(block), args, lds, s);
}

int denseBlocksMax(int block, size_t lds, bool fp32) {
  int perCu = 0, dev = 0, cus = 0;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const void* f = fp32 ? (const void*)gridSubstepKernel<float, false, false, true>
                       : (const void*)gridSubstepKernel<double, true, false, true>;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, f, block, lds) != hipSuccess) return 0;
  return perCu * cus;
}

hipError_t launchDenseSolve(const InstDev* d, int inst, const double* b, double* x, int n, int m, bool fp32,
                            hipStream_t s) {
  const int pb = n > 1024 * 512 ? 1024 : (n + 511) / 512;
  const int rowBlocks = (m + 8 * kDR - 1) / (8 * kDR);  // 8 wavefronts x kDR rows per workgroup
  const int sb = rowBlocks > 2048 ? 2048 : rowBlocks;
  if (fp32) {
    hipLaunchKernelGGL((densePermuteKernel<float>), dim3(pb), dim3(512), 0, s, d, inst, b);
    hipLaunchKernelGGL((denseSolveKernel<float>), dim3(sb), dim3(512), 0, s, d, inst, x);
  } else {
    hipLaunchKernelGGL((densePermuteKernel<double>), dim3(pb), dim3(512), 0, s, d, inst, b);
    hipLaunchKernelGGL((denseSolveKernel<double>), dim3(sb), dim3(512), 0, s, d, inst, x);
  }
  return hipGetLastError();
}

hipError_t launchGridSolve(const InstDev* d, int inst, const double* b, double* x, int waveRuns, int blocks,
                           int block, size_t lds, bool fp32, bool idx16, hipStream_t s) {
  void* args[] = {(void*)&d, (void*)&inst, (void*)&b, (void*)&x, (void*)&waveRuns};
#define PDX_G(R, X16)                                                                                   \
  if (fp32 == (sizeof(R) == 4) && idx16 == X16)                                                         \
    return hipLaunchCooperativeKernel((const void*)gridSolveKernel<R, X16>, dim3(blocks), dim3(block), args, \
                                      lds, s);
  PDX_G(float, true) PDX_G(float, false) PDX_G(double, true) PDX_G(double, false)
#undef PDX_G
  return hipErrorInvalidValue;
}

int gridBlocksMax(int block, size_t lds, bool fp32, bool idx16) {
  int perCu = 0, dev = 0, cus = 0;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
#define PDX_G(R, E, X16)                                                                                        \
  if (fp32 == (sizeof(R) == 4) && idx16 == X16)                                                                 \
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, (const void*)gridSubstepKernel<R, E, X16, false>,  \
                                                     block, lds) != hipSuccess)                                 \
      return 0;
  PDX_G(float, false, true) PDX_G(float, false, false) PDX_G(double, true, true) PDX_G(double, true, false)
#undef PDX_G
  return perCu * cus;
}

hipError_t launchSolve(const InstDev* d, int inst, int inst2, const double* b, double* x, int block,
                       size_t lds, bool fp32, bool blocked, bool idx16, bool wl, hipStream_t s) {
  const int g = inst2 >= 0 ? 2 : 1;
#define PDX_L(R, E, B, X16)                                                                     \
  if (fp32 == (sizeof(R) == 4) && blocked == B && (idx16 && blocked) == X16) {                  \
    if (wl)                                                                                     \
      hipLaunchKernelGGL((solveKernel<R, B, X16, true>), dim3(g), dim3(block), lds, s, d, inst, \
                         inst2, b, x);                                                          \
    else                                                                                        \
      hipLaunchKernelGGL((solveKernel<R, B, X16, false>), dim3(g), dim3(block), lds, s, d, inst, \
                         inst2, b, x);                                                          \
    return hipGetLastError();                                                                   \
  }
  PDX_KERNEL_TABLE(PDX_L)
#undef PDX_L
  return hipErrorInvalidValue;
}

// Staging of the allocation-free solve (pdx_solve3_noalloc): kNoAllocDofs FP64
// right-hand-side values and as many solution values.  Every device the library
// runs on holds its own copy until the process ends; calls that overlap on one
// device share it and race.
constexpr size_t kNoAllocDofs = (size_t)3 * 131072;
static __device__ __align__(256) double scratch[2 * kNoAllocDofs];

size_t noallocDofs() { return kNoAllocDofs; }

// The device address of `scratch` on `device`, looked up once per device (with
// that device current) and kept: host code uses this address, never the
// array's name, which on the host is only a stand-in.
hipError_t noallocScratch(int device, double** ptr) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static double* addr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if ((e = hipSetDevice(device)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    const hipError_t back = hipSetDevice(cur);
    if (e != hipSuccess) return e;
    if (back != hipSuccess) return back;
    addr[device] = static_cast<double*>(p);
  }
  *ptr = addr[device];
  return hipSuccess;
}

hipError_t setLdsLimit(size_t bytes) {
  hipError_t e = hipSuccess;
  const hipFuncAttribute A = hipFuncAttributeMaxDynamicSharedMemorySize;
#define PDX_A(F)                                                 \
  e = hipFuncSetAttribute((const void*)(F), A, (int)bytes);       \
  if (e != hipSuccess) return e;
#define PDX_L(R, E, B, X16)                                     \
  PDX_A((substepKernel<R, E, B, X16, false, false>))            \
  PDX_A((substepKernel<R, E, B, X16, false, true>))             \
  PDX_A((substepKernel<R, E, B, X16, true, false>))             \
  PDX_A((substepKernel<R, E, B, X16, true, true>))              \
  PDX_A((solveKernel<R, B, X16, false>))                        \
  PDX_A((solveKernel<R, B, X16, true>))                         \
  PDX_A((adjointKernel<R, B, X16, false>))                      \
  PDX_A((adjointKernel<R, B, X16, true>))
  PDX_KERNEL_TABLE(PDX_L)
#undef PDX_L
#undef PDX_A
#define PDX_G(R, E, X16, D)                                                                                  \
  e = hipFuncSetAttribute((const void*)gridSubstepKernel<R, E, X16, D>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                          (int)bytes);                                                                       \
  if (e != hipSuccess) return e;
  PDX_G(float, false, true, false) PDX_G(float, false, false, false) PDX_G(double, true, true, false)
  PDX_G(double, true, false, false) PDX_G(float, false, false, true) PDX_G(double, true, false, true)
#undef PDX_G
#define PDX_G(R, X16)                                                                                   \
  e = hipFuncSetAttribute((const void*)gridSolveKernel<R, X16>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                          (int)bytes);                                                                  \
  if (e != hipSuccess) return e;
  PDX_G(float, true) PDX_G(float, false) PDX_G(double, true) PDX_G(double, false)
#undef PDX_G
  return e;
}

}  // namespace pdx

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
//    scatter (DistanceTerm.cpp:396-434), the sweeps run each row's dot product in
//    the reference's column order (EigenDirectSolver.cpp:418-446), the residual
//    is the reference gather (Integrator.cpp:230-263).  Same bits as the CPU path.
//  * mixed (FP32 factor):  each PD iteration solves for the update
//    delta = L^{-1}(b(x) - L x) in FP32 and adds it to the FP64 state; b(x) - L x
//    and the residual are gathered in FP64 from the same projections.
//
// All FP64 code is compiled with -ffp-contract=off: no operation is fused.
#include <hip/hip_runtime.h>

#include "devtypes.h"

namespace pdx {

namespace {

constexpr int kWave = 64;
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

// One level-scheduled sweep.  Forward: row k gathers L(k,j) y_j over ascending j.
// Backward: column k gathers L(i,k) x_i over ascending i after scaling by 1/D_k.
// `store(k, a0, a1, a2)` is called once per row with its final value.
template <typename Real, bool kFwd, typename Store>
__device__ void sweep(const SweepDev& sw, Real* W, const Real* dinv, Store store) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  const Real* val = static_cast<const Real*>(sw.val);
  for (int l = 0; l < sw.nlev; ++l) {
    const int c1 = sw.levChunk[l + 1];
    for (int ch = sw.levChunk[l] + wave; ch < c1; ch += nw) {
      const int nrows = sw.chunkRows[ch];
      const int rbase = sw.chunkRowBase[ch];
      const int ebase = sw.chunkEntBase[ch];
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
      int off = ebase;
      for (int j = 0; j < maxlen; ++j) {
        const unsigned long long msk = __ballot(j < len);
        if (j < len) {
          const int e = off + lane;
          const int c = sw.col[e];
          const Real lv = val[e];
          a0 = subMul<Real>(a0, lv, W[3 * c + 0]);
          a1 = subMul<Real>(a1, lv, W[3 * c + 1]);
          a2 = subMul<Real>(a2, lv, W[3 * c + 2]);
        }
        off += __popcll(msk);
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

constexpr int kCap = 32;

// Block-scheduled sweep: per step, each wavefront takes whole tasks (a chain block
// or a batch of single rows).  Chain blocks keep their dense triangle in
// registers (lane t holds row/column t) and broadcast each finished value with
// readlane, so a chain of c rows costs c short ALU steps instead of c barriers.
// Forward rows keep the reference's ascending column order (external columns
// precede the block's own); backward rows apply the block's own rows last.
template <typename Real, bool kFwd, typename Store>
__device__ void blockSweep(const BlockSweepDev& sw, Real* W, const Real* dinv, Store store) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  const Real* val = static_cast<const Real*>(sw.val);
  const Real* dense = static_cast<const Real*>(sw.dense);
  for (int st = 0; st < sw.nstep; ++st) {
    const int t1 = sw.stepTask[st + 1];
    for (int t = sw.stepTask[st] + wave; t < t1; t += nw) {
      const int info = sw.taskInfo[t];
      const int nrows = info & 0xffff;
      const bool chain = (info >> 16) != 0;
      const int maxlen = sw.taskMaxLen[t];
      const int rbase = sw.taskRowBase[t];
      const int ebase = sw.taskEntBase[t];
      const bool act = lane < nrows;
      int k = 0, len = 0;
      if (act) {
        k = sw.slotRank[rbase + lane];
        len = sw.slotLen[rbase + lane];
      }
      Real dreg[kCap];
      if (chain) {
        const Real* D = dense + sw.taskDenseBase[t];
#pragma unroll
        for (int q = 0; q < kCap; ++q) dreg[q] = (act && q < nrows) ? D[q * nrows + lane] : Real(0);
      }
      Real a0 = 0, a1 = 0, a2 = 0;
      if (act) {
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
      for (int j = 0; j < maxlen; j += 4) {
        int cc[4];
        Real lv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const bool ok = act && (j + q < maxlen);
          const int e = ebase + (j + q) * nrows + lane;
          cc[q] = ok ? sw.col[e] : 0;
          lv[q] = ok ? val[e] : Real(0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (j + q < len) {
            const int c = cc[q];
            a0 = subMul<Real>(a0, lv[q], W[3 * c + 0]);
            a1 = subMul<Real>(a1, lv[q], W[3 * c + 1]);
            a2 = subMul<Real>(a2, lv[q], W[3 * c + 2]);
          }
        }
      }
      if (chain) {
        if (kFwd) {
#pragma unroll
          for (int q = 0; q < kCap; ++q) {
            if (q < nrows) {
              const Real y0 = readlaneT(a0, q), y1 = readlaneT(a1, q), y2 = readlaneT(a2, q);
              if (lane > q && dreg[q] != Real(0)) {
                a0 = subMul<Real>(a0, dreg[q], y0);
                a1 = subMul<Real>(a1, dreg[q], y1);
                a2 = subMul<Real>(a2, dreg[q], y2);
              }
            }
          }
        } else {
#pragma unroll
          for (int q = kCap - 1; q >= 0; --q) {
            if (q < nrows) {
              const Real y0 = readlaneT(a0, q), y1 = readlaneT(a1, q), y2 = readlaneT(a2, q);
              if (lane < q && dreg[q] != Real(0)) {
                a0 = subMul<Real>(a0, dreg[q], y0);
                a1 = subMul<Real>(a1, dreg[q], y1);
                a2 = subMul<Real>(a2, dreg[q], y2);
              }
            }
          }
        }
      }
      if (act) {
        W[3 * k + 0] = a0;
        W[3 * k + 1] = a1;
        W[3 * k + 2] = a2;
        store(k, a0, a1, a2);
      }
    }
    __syncthreads();
  }
}

struct Acc3 {
  double x, y, z;
};

// Reference right-hand side of one vertex (exact form): seed bBase + bendRhs, then
// each incident edge in colour order: +-kappa d_e and the pin compensation
// kappa q (DistanceTerm.cpp:211-269, :396-434); pinned rows take q (Assembler.cpp:401-414).
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
  for (int p = I.vstart[v]; p < I.vstart[v + 1]; ++p) {
    const int e = I.vedgesColor[p];
    const int a = I.ea[e], bb = I.eb[e];
    const bool aP = I.pinned[a] != 0, bP = I.pinned[bb] != 0;
    if (aP && bP) continue;
    const double rx = x[3 * a] - x[3 * bb], ry = x[3 * a + 1] - x[3 * bb + 1],
                 rz = x[3 * a + 2] - x[3 * bb + 2];
    const double len = sqrt(rx * rx + ry * ry + rz * rz);
    double dx = 0.0, dy = 0.0, dz = 0.0;
    if (!(len <= kMinLength)) {
      const double s = I.rest[e] / len;
      dx = rx * s;
      dy = ry * s;
      dz = rz * s;
    }
    const double k = I.kappa[e];
    const double cx = dx * k, cy = dy * k, cz = dz * k;
    if (a == v) {
      b.x += cx;
      b.y += cy;
      b.z += cz;
      if (bP) {
        b.x += k * I.pinpos[3 * bb];
        b.y += k * I.pinpos[3 * bb + 1];
        b.z += k * I.pinpos[3 * bb + 2];
      }
    } else {
      b.x -= cx;
      b.y -= cy;
      b.z -= cz;
      if (aP) {
        b.x += k * I.pinpos[3 * a];
        b.y += k * I.pinpos[3 * a + 1];
        b.z += k * I.pinpos[3 * a + 2];
      }
    }
  }
  return b;
}

// Residual of one free vertex at x (Integrator.cpp:92-109 + :230-263), and, when
// `rhs` is non-null, the update right-hand side b(x) - (L x) of the same vertex
// (mixed form).  L x uses the pinned targets q for pinned columns, which is
// what the assembled L and the pin compensation in b amount to.
__device__ inline double residualAndRhs(const InstDev& I, int v, Acc3* rhs) {
  const double* x = I.x;
  const double xv0 = x[3 * v], xv1 = x[3 * v + 1], xv2 = x[3 * v + 2];
  Acc3 f{0.0, 0.0, 0.0};
  Acc3 r{0.0, 0.0, 0.0};
  for (int p = I.vstart[v]; p < I.vstart[v + 1]; ++p) {
    const int e = I.vedges[p];
    const int a = I.ea[e], bb = I.eb[e];
    const double rx = x[3 * a] - x[3 * bb], ry = x[3 * a + 1] - x[3 * bb + 1],
                 rz = x[3 * a + 2] - x[3 * bb + 2];
    const double len = sqrt(rx * rx + ry * ry + rz * rz);
    double dx = 0.0, dy = 0.0, dz = 0.0;
    if (!(len <= kMinLength)) {
      const double s = I.rest[e] / len;
      dx = rx * s;
      dy = ry * s;
      dz = rz * s;
    }
    const double k = I.kappa[e];
    const double cx = rx * k - dx * k, cy = ry * k - dy * k, cz = rz * k - dz * k;
    if (a == v) {
      f.x += cx;
      f.y += cy;
      f.z += cz;
    } else {
      f.x -= cx;
      f.y -= cy;
      f.z -= cz;
    }
    if (rhs) {
      const int u = (a == v) ? bb : a;
      const double sg = (a == v) ? 1.0 : -1.0;
      const bool uP = I.pinned[u] != 0;
      const double u0 = uP ? I.pinpos[3 * u] : x[3 * u];
      const double u1 = uP ? I.pinpos[3 * u + 1] : x[3 * u + 1];
      const double u2 = uP ? I.pinpos[3 * u + 2] : x[3 * u + 2];
      r.x += sg * dx * k - k * (xv0 - u0);
      r.y += sg * dy * k - k * (xv1 - u1);
      r.z += sg * dz * k - k * (xv2 - u2);
    }
  }
  Acc3 fb{0.0, 0.0, 0.0};
  for (int p = I.bstart[v]; p < I.bstart[v + 1]; ++p) {
    const int s = I.bstencil[p];
    const int ia = I.sa[s], ib = I.sb[s], ic = I.sc[s];
    const double s0 = x[3 * ia] - 2.0 * x[3 * ib] + x[3 * ic];
    const double s1 = x[3 * ia + 1] - 2.0 * x[3 * ib + 1] + x[3 * ic + 1];
    const double s2 = x[3 * ia + 2] - 2.0 * x[3 * ib + 2] + x[3 * ic + 2];
    const double w = (double)I.bweight[p];
    const double coef = -I.bk[s] * w;
    fb.x += s0 * coef;
    fb.y += s1 * coef;
    fb.z += s2 * coef;
    if (rhs) {
      const int vv[3] = {ia, ib, ic};
      double t0 = 0.0, t1 = 0.0, t2 = 0.0;
      const double wt[3] = {1.0, -2.0, 1.0};
      for (int q = 0; q < 3; ++q) {
        const int u = vv[q];
        const bool uP = I.pinned[u] != 0;
        t0 += wt[q] * (uP ? I.pinpos[3 * u] : x[3 * u]);
        t1 += wt[q] * (uP ? I.pinpos[3 * u + 1] : x[3 * u + 1]);
        t2 += wt[q] * (uP ? I.pinpos[3 * u + 2] : x[3 * u + 2]);
      }
      r.x += coef * t0;
      r.y += coef * t1;
      r.z += coef * t2;
    }
  }
  f.x += fb.x;
  f.y += fb.y;
  f.z += fb.z;
  const double mass = I.mass[v];
  const double mi = mass * I.invH2;
  const double i0 = (xv0 - I.xhat[3 * v]) * mi, i1 = (xv1 - I.xhat[3 * v + 1]) * mi,
               i2 = (xv2 - I.xhat[3 * v + 2]) * mi;
  if (rhs) {
    rhs->x = r.x - i0;
    rhs->y = r.y - i1;
    rhs->z = r.z - i2;
  }
  const double t0 = i0 + f.x, t1 = i1 + f.y, t2 = i2 + f.z;
  return sqrt(t0 * t0 + t1 * t1 + t2 * t2) / mass;
}

template <typename Real>
__device__ inline void putW(Real* W, int r, double a, double b, double c) {
  W[3 * r + 0] = (Real)a;
  W[3 * r + 1] = (Real)b;
  W[3 * r + 2] = (Real)c;
}

// Fills W with the first right-hand side of a substep (iteration 0 at x_n).
template <typename Real, bool kExact>
__device__ void gatherFirst(const InstDev& I, Real* W) {
  for (int v = threadIdx.x; v < I.n; v += blockDim.x) {
    const int r = I.vrank[v];
    if (kExact) {
      const Acc3 b = rhsExact(I, v);
      putW(W, r, b.x, b.y, b.z);
    } else if (I.pinned[v]) {
      putW(W, r, I.pinpos[3 * v] - I.x[3 * v], I.pinpos[3 * v + 1] - I.x[3 * v + 1],
           I.pinpos[3 * v + 2] - I.x[3 * v + 2]);
    } else {
      Acc3 rh;
      (void)residualAndRhs(I, v, &rh);
      putW(W, r, rh.x, rh.y, rh.z);
    }
  }
}

template <typename Real, bool kExact, bool kBlocked>
__device__ void runSweeps(const InstDev& I, Real* W) {
  const Real* dinv = static_cast<const Real*>(I.dinv);
  auto none = [](int, Real, Real, Real) {};
  if (kBlocked)
    blockSweep<Real, true>(I.bfwd, W, dinv, none);
  else
    sweep<Real, true>(I.fwd, W, dinv, none);
  double* x = I.x;
  const int* rankv = I.rankv;
  const uint8_t* pinned = I.pinned;
  const double* pinpos = I.pinpos;
  auto put = [=](int k, Real a0, Real a1, Real a2) {
    const int v = rankv[k];
    if (kExact) {
      x[3 * v] = (double)a0;
      x[3 * v + 1] = (double)a1;
      x[3 * v + 2] = (double)a2;
    } else if (pinned[v]) {
      x[3 * v] = pinpos[3 * v];
      x[3 * v + 1] = pinpos[3 * v + 1];
      x[3 * v + 2] = pinpos[3 * v + 2];
    } else {
      x[3 * v] += (double)a0;
      x[3 * v + 1] += (double)a1;
      x[3 * v + 2] += (double)a2;
    }
  };
  if (kBlocked)
    blockSweep<Real, false>(I.bbwd, W, dinv, put);
  else
    sweep<Real, false>(I.bwd, W, dinv, put);
}

template <typename Real, bool kExact, bool kBlocked>
__global__ void __launch_bounds__(1024) substepKernel(const InstDev* __restrict__ insts,
                                                      const int* __restrict__ order, int nsub) {
  extern __shared__ double smem[];
  const InstDev& I = insts[order[blockIdx.x]];
  double* red = smem;  // 2 x 16 slots
  Real* W = I.ldsW ? reinterpret_cast<Real*>(smem + 32) : static_cast<Real*>(I.Wg);
  const int n = I.n;
  const int tid = threadIdx.x, bs = blockDim.x;

  for (int sub = 0; sub < nsub; ++sub) {
    // predict (Integrator.cpp:447-463) and the displacement scale
    double lscale = 0.0;
    for (int v = tid; v < n; v += bs) {
      for (int d = 0; d < 3; ++d) {
        const double p = I.x[3 * v + d];
        I.xbs[3 * v + d] = p;
        const double xh = p + I.vel[3 * v + d] * I.h + I.gh2[d];
        I.xhat[3 * v + d] = xh;
        I.prev[3 * v + d] = p;
        lscale = fmax(lscale, fabs(xh - p));
      }
    }
    double scale = blockMax(lscale, red);
    scale = fmax(scale, I.gScaleH2);
    gatherFirst<Real, kExact>(I, W);
    __syncthreads();

    int used = 0, conv = 0;
    double lastRes = -1.0;
    for (int k = 0; k < I.maxIter; ++k) {
      runSweeps<Real, kExact, kBlocked>(I, W);
      const bool more = (k + 1 < I.maxIter);
      double ldiff = 0.0, lres = 0.0;
      for (int v = tid; v < n; v += bs) {
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
            putW(W, r, b.x, b.y, b.z);
          }
        } else {
          if (pinnedV) {
            if (more) putW(W, r, I.pinpos[3 * v] - x0, I.pinpos[3 * v + 1] - x1, I.pinpos[3 * v + 2] - x2);
          } else {
            Acc3 rh;
            const double rv = residualAndRhs(I, v, more ? &rh : nullptr);
            if (I.wantResidual && I.mass[v] > 0.0) lres = fmax(lres, rv);
            if (more) putW(W, r, rh.x, rh.y, rh.z);
          }
        }
      }
      // criterion (Integrator.cpp:673-751); two slot sets alternate so one barrier suffices
      double* rs = red + 16 * ((k & 1) ^ 1);
      const double diff = blockMax(ldiff, rs);
      double residual = -1.0;
      if (I.wantResidual) {
        __syncthreads();
        residual = blockMax(lres, rs);
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
    for (int v = tid; v < n; v += bs) {
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
    if (tid == 0) {
      I.istats[0] = used;
      I.istats[1] = conv;
      I.istats[2] += conv;
      I.lstats[0] += used;
      I.lstats[1] += 1;
      I.dstats[0] = lastRes;
    }
    __syncthreads();
  }
}

// Stand-alone global solve x = L^{-1} b (EigenDirectSolver::solve) on one instance.
template <typename Real, bool kBlocked>
__global__ void __launch_bounds__(1024) solveKernel(const InstDev* __restrict__ insts, int inst,
                                                    const double* __restrict__ b,
                                                    double* __restrict__ xout) {
  extern __shared__ double smem[];
  const InstDev& I = insts[inst];
  Real* W = I.ldsW ? reinterpret_cast<Real*>(smem + 32) : static_cast<Real*>(I.Wg);
  for (int v = threadIdx.x; v < I.n; v += blockDim.x) {
    const int r = I.vrank[v];
    putW(W, r, b[3 * v], b[3 * v + 1], b[3 * v + 2]);
  }
  __syncthreads();
  const Real* dinv = static_cast<const Real*>(I.dinv);
  auto none = [](int, Real, Real, Real) {};
  if (kBlocked)
    blockSweep<Real, true>(I.bfwd, W, dinv, none);
  else
    sweep<Real, true>(I.fwd, W, dinv, none);
  const int* rankv = I.rankv;
  auto put = [=](int k, Real a0, Real a1, Real a2) {
    const int v = rankv[k];
    xout[3 * v] = (double)a0;
    xout[3 * v + 1] = (double)a1;
    xout[3 * v + 2] = (double)a2;
  };
  if (kBlocked)
    blockSweep<Real, false>(I.bbwd, W, dinv, put);
  else
    sweep<Real, false>(I.bwd, W, dinv, put);
}

}  // namespace

#define PDX_KERNEL_TABLE(X) X(float, false, true) X(float, false, false) X(double, true, true) X(double, true, false)

hipError_t launchSubstep(const InstDev* d, const int* order, int ninst, int nsub, int block,
                         size_t lds, bool fp32, bool blocked, hipStream_t s) {
#define PDX_L(R, E, B)                                                                            \
  if (fp32 == (sizeof(R) == 4) && blocked == B) {                                                 \
    hipLaunchKernelGGL((substepKernel<R, E, B>), dim3(ninst), dim3(block), lds, s, d, order, nsub); \
    return hipGetLastError();                                                                     \
  }
  PDX_KERNEL_TABLE(PDX_L)
#undef PDX_L
  return hipErrorInvalidValue;
}

hipError_t launchSolve(const InstDev* d, int inst, const double* b, double* x, int block,
                       size_t lds, bool fp32, bool blocked, hipStream_t s) {
#define PDX_L(R, E, B)                                                                  \
  if (fp32 == (sizeof(R) == 4) && blocked == B) {                                       \
    hipLaunchKernelGGL((solveKernel<R, B>), dim3(1), dim3(block), lds, s, d, inst, b, x); \
    return hipGetLastError();                                                           \
  }
  PDX_KERNEL_TABLE(PDX_L)
#undef PDX_L
  return hipErrorInvalidValue;
}

hipError_t setLdsLimit(size_t bytes) {
  hipError_t e = hipSuccess;
#define PDX_L(R, E, B)                                                                                \
  e = hipFuncSetAttribute((const void*)substepKernel<R, E, B>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                          (int)bytes);                                                                \
  if (e != hipSuccess) return e;                                                                      \
  e = hipFuncSetAttribute((const void*)solveKernel<R, B>, hipFuncAttributeMaxDynamicSharedMemorySize,     \
                          (int)bytes);                                                                \
  if (e != hipSuccess) return e;
  PDX_KERNEL_TABLE(PDX_L)
#undef PDX_L
  return e;
}

}  // namespace pdx

// Device-side descriptors shared by the host packer and the HIP kernels.  All
// pointers are device addresses inside one arena allocation per context.
#pragma once
#include <cstdint>

namespace pdx {

// One level-scheduled triangular sweep in rank space.  Level l owns chunks
// [levChunk[l], levChunk[l+1]); a chunk is <= 64 rows (one wavefront, one row per
// lane) sorted by entry count, its entries stored jagged-diagonal: entry j of lane
// t sits at entBase + (sum of the chunk's row counts with len > j' for j' < j) + t,
// so every lane of a wave reads consecutive addresses for the same j.
struct SweepDev {
  int nlev;
  int pad;
  const int* levChunk;
  const int* chunkRows;
  const int* chunkRowBase;
  const int* chunkEntBase;
  const int* slotRank;
  const int* slotLen;
  const int* col;
  const void* val;  // double or float by precision
};

// Block-scheduled sweep (fast mode).  The ranks are cut into contiguous blocks:
// elimination-tree chains of at most `cap` rows, and single rows.  Step st owns
// tasks [stepTask[st], stepTask[st+1]); a task is one wavefront's work:
//  * a row batch (<= 64 single-row blocks, one per lane, gather form), or
//  * a chain block (lane t = rank start+t): the external part of every row is
//    gathered in parallel, then the block's own dense lower triangle is applied
//    column by column, each finished value broadcast to the other lanes.
// External entries are ELL: entry j of lane t at entBase + j*nrows + t.
struct BlockSweepDev {
  int nstep;
  int cap;
  const int* stepTask;
  const int* taskInfo;      // nrows | (chain << 16)
  const int* taskMaxLen;
  const int* taskRowBase;
  const int* taskEntBase;
  const int* taskDenseBase;
  const int* slotRank;
  const int* slotLen;
  const int* col;
  const void* val;
  const void* dense;        // chain blocks: cap x cap, [s * nrows + t] = L(start+t, start+s) (forward)
                            // or [r * nrows + t] = L(start+r, start+t) (backward)
};

struct InstDev {
  int n, E, B, m;
  int maxIter, wantResidual, hasBendRhs, ldsW;
  double h, invH, invH2, dampInvH;
  double gh2[3];
  double gScaleH2;   // length(g) * h * h
  double relTol, absTol, resTol, resThresh;
  // state (FP64)
  double* x;
  double* vel;
  double* xbs;   // positionsBeforeStep
  double* xhat;  // predicted
  double* prev;  // previous iterate
  // scene
  const double* mass;
  const uint8_t* pinned;
  const double* pinpos;
  const double* brhs;  // bending constant rhs (3n) or null
  const int* ea;
  const int* eb;
  const double* rest;
  const double* kappa;
  const int* vstart;
  const int* vedges;       // incidence, edge order (residual)
  const int* vedgesColor;  // incidence, colour order (scatter)
  const int* bstart;
  const int* bstencil;
  const int8_t* bweight;
  const int* sa;
  const int* sb;
  const int* sc;
  const double* bk;
  // solve
  const int* vrank;   // vertex -> rank
  const int* rankv;   // rank -> vertex
  const void* dinv;   // Real[m]
  SweepDev fwd, bwd;
  BlockSweepDev bfwd, bbwd;
  void* Wg;           // global W scratch (3m Real) when W is not LDS-resident
  // statistics
  int* istats;        // [0] iterations used, [1] converged, [2] early exits
  long long* lstats;  // [0] total iterations, [1] substeps
  double* dstats;     // [0] last residual
};

}  // namespace pdx

// Device-side descriptors shared by the host packer and the HIP kernels.  All
// pointers are device addresses inside one arena allocation per context.
#pragma once
#include <cstdint>

namespace pdx {

// Lines (256 bytes each) allocated per step counter of the grid form's level
// sweeps: a launch uses 8 to kFlowLines of them by its workgroup count, each
// workgroup adds to one, and the waits sum them.
constexpr int kFlowLines = 64;

// An edge's rest length and stiffness, loaded together.
struct alignas(16) RestKappa {
  double rest, kap;
};

// One vertex-stencil slot of the bend gather: the stencil's two other vertices
// (in stencil order a, b, c with the slot's own vertex left out), each encoded
// (index << 1) | pinned; the slot vertex's role (0, 1, 2 = a, b, c) | (its pinned
// flag << 2); its weight (1, -2, 1).
struct alignas(16) BendRec {
  int other1, other2, role, weight;
};

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
// elimination-tree chains of at most `cap` rows, and single rows; each block
// level is one step.  A step streams the external factor entries of all its rows
// as runs of kRun entries (run r of the sweep: runCol/runVal[r*kRun ...]), one
// partial sum per run, then its tasks ([stepTask[st], stepTask[st+1])) finish:
//  * a row batch (<= 64 single-row blocks, one per lane), or
//  * a chain block (lane t = rank start+t): after its external partials, the
//    block's own dense lower triangle is applied column by column, each finished
//    value broadcast to the other lanes.
struct BlockSweepDev {
  int nstep;
  int cap;
  const int* stepTask;    // nstep+1
  const int* stepRun;     // nstep+1, global run index
  const int* stepDense;   // nstep+1, offsets into dense
  const int* taskInfo;    // nrows | (chain << 16)
  const int* taskSlot;    // first slot of the task
  const int* taskDense;   // offset of the chain's block inside its step's dense range
  const int* slotRank;
  const int* slotRunBeg;  // run range of the slot, relative to its step
  const int* slotRunEnd;
  const void* runCol;     // int32, or uint16 when every rank fits
  const void* runVal;
  const int* laneMeta;    // per task, 64 lanes x {rank or -1, run begin, run end,
                          //   segment length | offset in segment << 7 | lanes used << 14
                          //   | longest segment << 21}
  const void* laneDinv;   // per task, 64 lanes x 1/D of the lane's rank (Real)
  const int* laneDense;   // per task, 64 lanes x offset of the segment's inverse block
  const void* dense;      // per segment of length L > 1, its inverse X = B^-1:
                          // forward [s * L + t] = X(t, s), backward [s * L + t] = X(s, t)
  // Level form (the grid form's sweeps): blocks are whole elimination-tree chains
  // and every block level is one step, in two phases.  Phase A tasks each stream
  // their own runs ([taskRun[t], taskRun[t+1])), finish their lanes' rows and
  // their short chains, and write each row of a long chain (lane info bit 28) to
  // R instead of W, at component stride rTotal and position laneDense.  Phase B
  // tasks ([stepB[st], stepB[st+1])) apply a long chain's inverse, 8 rows each:
  // bTask = {offset of the panel in xpanel, row stride (a multiple of 4), R
  // position of the panel's column 0, first rank | rows << 24}.
  const int* taskRun;     // per task, its first run (ntask + 1)
  const int* stepB;       // nstep + 1
  const int* bTask;       // 4 per phase B task
  const void* xpanel;     // the long chains' inverses, panels of 8 rows (Real, zero padded)
  int rTotal;             // R values per component (each chain's range padded to 4)
  int pad;
  // Dense external blocks: a long chain whose rows share most of their external
  // columns (the backward sweep's supernodes) takes phase A as dense tasks, 8
  // rows each: r = base - E W[U] over the chain's union U of external columns,
  // eTask = {offset of the panel in epanel (8 values per column, column-major),
  // columns, offset of U in ucol, first rank | rows << 24, R position of row 0,
  // 0, 0, 0}; the step's dense tasks are [stepE[st], stepE[st+1]).
  const int* stepE;       // nstep + 1
  const int* eTask;       // 8 per dense task
  const int* ucol;        // union columns (ranks)
  const void* epanel;     // Real
};

// Exchange area of an instance split over two workgroups (a pair).  Each
// workgroup counts the synchronisation points it has passed in flags[role]; a
// workgroup that waits longer than a bound for its partner sets *giveUp (read by
// the host after the launch) and both stop.
struct PairDev {
  int* flags;      // lead's count at flags[0], partner's at flags[32] (separate cache lines)
  int* giveUp;     // one per context
  double* red;     // [2 buffers][2 roles][4]: scale, displacement, residual maxima
  void* xbuf;      // 3m Real: the halves' solution values handed across (rank space)
};

struct InstDev {
  int n, E, B, m;
  int maxIter, wantResidual, hasBendRhs, ldsW;
  int partOff, denseOff;  // byte offsets of the run partials and (FP64) the staged
                          // inverse blocks of a step in LDS
  int stepOff, pad2;      // byte offset of the step tables of both sweeps in LDS
  double h, invH, invH2, dampInvH;
  double gh2[3];
  double gScaleH2;   // length(g) * h * h
  double relTol, absTol, resTol, resThresh;
  double chebRho;     // > 0: Chebyshev blend from iteration chebStart + 1 on
  int chebStart, pad3;
  // state (FP64)
  double* x;
  double* vel;
  double* xbs;   // positionsBeforeStep
  double* xhat;  // predicted
  double* prev;  // previous iterate
  // scene
  const double* mass;
  const uint8_t* pinned;
  double* pinpos;      // pin targets (3n); rewritten per substep along a pin trajectory
  double* brhs;        // bending constant rhs (3n) or null; follows the pins
  const int* ea;
  const int* eb;
  const double* rest;
  const double* kappa;
  const int* vstart;
  const int* vedges;       // incidence, edge order (residual)
  const int* vedgesColor;  // incidence, colour order (scatter)
  // Flattened incidence records (one per vertex-edge slot), so that the gather
  // has one dependent load (the neighbour's position) per edge:
  //   rec = (u << 2) | (u pinned << 1) | (v is the edge's b end)
  // with the edge's rest length and kappa beside it as one 16-byte pair.
  const int* recE;         // edge order
  const RestKappa* rkE;
  const int* recC;         // colour order
  const RestKappa* rkC;
  // bend records per vertex-stencil slot (BendRec: the stencil's other two
  // vertices, the slot vertex's role, its weight) and the stencil's k
  const BendRec* brRec;
  const double* brK;
  const int* bstart;
  const int* bstencil;
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
  // dense solve mode: the inverse of the component matrix in rank space (m rows
  // of denseLd Real, denseLd = m rounded up to a multiple of 4 with zero padding;
  // row r = column r: the inverse is symmetric), and whether each workgroup
  // stages W in LDS
  const void* Linv;
  int denseLdsW, denseLd;
  // differentiable rollout (null / 0 outside it)
  int npin, npinEdge, npinSten, rollK;
  const int* pinList;      // pinned vertices, ascending
  const int* pinEdge;      // per edge with exactly one pinned end: {edge, pin index}
  const int* pinSten;      // per stencil with a pinned vertex: {stencil, pin index or -1 x 3}
  const double* pinTraj;   // [substep][pin][3] pin targets
  double* tape;            // [substep of the segment][iterate 0..K][3n] FP64 iterates
  double* xbar;            // adjoint of x at the current substep boundary (3n)
  double* vbar;            // adjoint of v (3n)
  double* cur;             // adjoint of the current iterate (3n)
  double* bbar;            // adjoint of the substep's constant right-hand side (3n)
  double* xnbar;           // adjoint of x_n being accumulated (3n)
  double* gkap;            // dLoss/dkappa_e (E)
  double* gbend;           // dLoss/dk_s (B)
  double* gpin;            // dLoss/dq [substep][pin][3]
  // Split over two workgroups (nested-dissection order, FP32 solves, W in LDS):
  // ranks [0, rA) are the first half of the top-level split, [rA, rB) the second,
  // [rB, m) its separator.  The lead (pairRole 1) owns the vertices of the first
  // half and of the separator and sweeps bfwd/bbwd over the first half and
  // bfwdS/bbwdS over the separator; the partner (pairRole 2) owns the second half
  // and sweeps bfwd/bbwd over it.  pairRole 0: one workgroup, own == null.
  int pairRole, nOwn, rA, rB;
  const int* own;     // owned vertices, ascending
  BlockSweepDev bfwdS, bbwdS;
  PairDev pair;
  // grid form: per step of the forward then the backward sweep, the tasks done
  // (both phases) and the phase A tasks done, cumulative over the sweeps of one
  // launch (zeroed at its start): [forward done | forward phase A | backward done
  // | backward phase A], each counter kFlowLines lines of 256 bytes (summed by
  // the waits); and R, the long chains'
  // rows between the phases (3 x rTotal Real, component-major, zero padding)
  int* flow;
  void* flowPart;
  // statistics
  int* istats;        // [0] iterations used, [1] converged, [2] early exits
  long long* lstats;  // [0] total iterations, [1] substeps, [2..6] phase ticks
                      // (gather, forward, backward, predict+velocity, criterion)
  double* dstats;     // [0] last residual
};

}  // namespace pdx

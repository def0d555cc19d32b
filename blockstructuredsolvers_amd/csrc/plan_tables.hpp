// plan_tables.hpp -- device-side description of one recursion level of a plan.
//
// A plan holds any number of independent systems of the same block size n; each has
// its own N and its own level schedule.  Global level l gathers every system that
// still has a level l: its chains are flattened into one ChainRec table (one
// workgroup per chain, every system in the same launch), its separators into one
// SepRec table, and each participating system gets a LevelSys entry ("slot") with
// the device addresses of its buffers at this level.  A same-shape batch is the
// special case where all systems share one schedule.
#pragma once

namespace bss {

struct ChainRec {
  int start, len;   // first block and number of blocks of the chain
  int sepL, sepR;   // separator indices (into the system's separator list), -1 if absent
  int slot;         // LevelSys entry of the chain's system
  int pad0, pad1, pad2;
};

struct SepRec {
  int slot;  // LevelSys entry
  int p;     // separator index within its system
  int pos;   // block index of the separator at this level
  int pad;
};

struct LevelSys {
  const void* A;  // diagonal blocks of this level (level 0: caller's buffer, input type)
  const void* B;  // couplings (N-1 blocks)
  void* U;        // Uinv output (upper 16x16 tiles), may alias A
  void* C;        // C output, may alias B
  void* W;        // panel output (chains with a left separator)
  void* SR;       // per separator: C_last^T C_last of the chain on its left
  void* SL;       // per separator: sum W^T W of the chain on its right
  void* An;       // next level's diagonal blocks (assembly target)
  void* Bn;       // next level's couplings (written by the factor kernel)
  void* d;        // this level's vectors, n x nrhs per block
  void* rR;       // per separator: forward contribution C_last^T y_last
  void* rL;       // per separator: forward contribution sum W^T y
  void* dn;       // next level's vectors (reduced RHS, then separator solutions)
  const int* orig;  // level-0 block index of every block (info reporting)
  int sys;          // system index
  int N;            // blocks at this level
  // Level hand-off fused into the sweeps (`fused` != 0): the forward sweep assembles
  // this level's right-hand side from the previous level, d[q] = dp[pp[q]] - rRp[q] -
  // rLp[q] (when dp is set, i.e. below the top level), and every chain writes the
  // separator on its right (assembled in the forward sweep, solved in the backward).
  const void* dp;   // previous level's vectors
  const void* rRp;  // previous level's separator contributions
  const void* rLp;
  const int* pp;    // position of each block of this level in the previous level
  int fused;
  int pad_;
};

// One buffer to convert element-wise S -> D (count elements): the reduced couplings
// FP32 -> FP64 when separator levels are re-factored in FP64, and the separator
// solutions FP64 -> FP32 on their way back to an FP32 level.
template <typename S, typename D>
struct ConvRec {
  const S* src;
  D* dst;
  long long count;
};

}  // namespace bss

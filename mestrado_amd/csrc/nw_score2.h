#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mst {

// Work item of the packed kernel: two pairs (X, Ya) and (X, Yb); yb < 0 for a
// single pair. oa/ob: output index of each pair in out_pair (or -1).
struct Pair2 {
  int x, ya, yb, pad;
  int64_t oa, ob;
};

struct Score2Args {
  const int8_t* seqs;
  const int64_t* off;
  const Pair2* items;
  int64_t first, last;      // item range of this launch
  unsigned int* counter;    // work queue head (zero before the launch)
  int32_t* out_mat;         // nseq x nseq, written at (x,y) and (y,x) (or null)
  int nseq;
  int32_t* out_pair;        // (or null)
  int2* bufs;
  int64_t bufcap;
  const int8_t* subst;
  int go, ge;
};

void launch_score2(int K, const Score2Args& a, int grid, hipStream_t st);

// Shared-profile form: item = row pair (ya, yb) against the column sequences
// xlist[xbegin .. xbegin+xcount); single strip (rows <= 64*K).
struct Item3 {
  int ya, yb;
  int xbegin, xcount;
};

struct Score3Args {
  const int8_t* seqs;
  const int64_t* off;
  const Item3* items;
  int64_t first, last;
  unsigned int* counter;
  const int32_t* xlist;
  int32_t* out_mat;
  int nseq;
  const int8_t* subst;
  int go, ge;
};

void launch_score3(int K, const Score3Args& a, int grid, hipStream_t st);
int score3_waves_per_group();

}  // namespace mst

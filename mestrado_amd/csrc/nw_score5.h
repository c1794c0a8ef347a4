#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nw_score2.h"

namespace mst {

// Pipelined packed form: the rows of a (Ya, Yb) pair are cut into W strips of
// 64*K rows, one wave per strip; the W waves of a pipeline sweep the same
// column stream, each strip's last row passing to the next wave through an LDS
// ring. A work item (Item3) is a run of consecutive sorted positions; its
// column sequences are concatenated into one stream per pipeline.
struct Score5Args {
  const int8_t* scodes;   // codes of every sequence in length order, concatenated
  const int64_t* soff;    // offsets into scodes by sorted position (nseq + 1)
  const int32_t* order;   // sorted position -> sequence id
  const int8_t* seqs;     // codes by sequence id (row sequences)
  const int64_t* off;
  const Item3* items;     // xbegin/xcount are sorted positions
  int64_t first, last;
  unsigned int* counter;  // work queue head (zero before the launch)
  int32_t* out_mat;       // nseq x nseq, written at (x, y) and (y, x)
  int nseq;
  const int8_t* subst;
  int go, ge;
};

// Rows per lane K and waves per pipeline W supported: K in 3..5 for W = 1,
// K in 3..4 for W = 2..4.
bool score5_supported(int K, int W);
int score5_block_threads(int W);
int score5_groups_per_cu(int K, int W);
void launch_score5(int K, int W, const Score5Args& a, int grid, hipStream_t st);

}  // namespace mst

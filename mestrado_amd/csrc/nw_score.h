// Host-visible argument block of the score kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mst {

struct ScoreArgs {
  const int8_t* seqs;    // reversed residue codes, concatenated
  const int64_t* off;    // nseq+1 offsets into seqs
  const int2* pairs;     // (x sequence, y sequence) per pair (or null: pi, pj)
  const int32_t* pi = nullptr;  // x and y sequence per pair when pairs is null
  const int32_t* pj = nullptr;
  int64_t first, last;   // pair index range [first, last)
  int32_t* out_pair;     // score per pair index (or null)
  int32_t* out_mat;      // nseq*nseq matrix, written at (x,y) and (y,x) (or null)
  int nseq;
  int2* bufs;            // 2*bufcap boundary entries per workgroup
  int64_t bufcap;        // >= longest column count
  const int8_t* subst;   // 24x24 substitution matrix S[Y][X]
  int go, ge;
  int transpose_ok;      // S symmetric: the shorter sequence may go in the rows
};

void launch_score_pairs(int K, const ScoreArgs& a, int grid, hipStream_t st);

}  // namespace mst

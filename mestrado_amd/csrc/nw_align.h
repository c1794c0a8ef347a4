#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mst {

enum { SRC_SEQ = 0, SRC_TAB = 1 };
enum { OP_DIAG = 0, OP_LEFT = 1, OP_UP = 2 };  // LEFT: X residue vs gap; UP: gap vs Y

struct AlignJob {
  int n, m;            // columns (X) and rows (Y)
  int64_t x_off, y_off;  // residue offsets (SRC_SEQ)
  int64_t tab_off;     // byte offset of the score table (SRC_TAB), 8-byte aligned
  int64_t flag_off;    // word offset of the flag block
  int64_t ops_off;     // byte offset of the path output (n+m bytes)
};

struct AlignArgs {
  const AlignJob* jobs;
  int64_t first, last;
  const int8_t* seqs;
  const uint8_t* tab;
  const int8_t* subst;
  uint32_t* flags;
  int32_t* scores;
  int2* bufs;
  int64_t bufcap;
  int go, ge;
  unsigned int* counter;  // pipelined form: work queue head (zero before the launch)
};

struct TraceArgs {
  const AlignJob* jobs;
  int64_t first, last;
  const uint32_t* flags;
  uint8_t* ops;
  int32_t* lens;
};

void launch_align(int src, const AlignArgs& a, int grid, hipStream_t st);
// Pipelined form for jobs of more than one 512-row strip: one job per
// workgroup of W waves (W = 4, or 8 with SRC_TAB); wave j computes strips j,
// j+W, ...; a strip reads the row above from the wave of the previous strip
// through an LDS ring, or, across the wrap from wave W-1 to wave 0, through a
// row buffer of bufcap columns per workgroup in bufs.
void launch_align_pipe(int src, int W, const AlignArgs& a, int grid, hipStream_t st);
int64_t align_strips(int m);  // 512-row strips of an m-row job
void launch_traceback(const TraceArgs& a, hipStream_t st);
int64_t flag_words(int n, int m);  // flag block size of an n x m job

}  // namespace mst

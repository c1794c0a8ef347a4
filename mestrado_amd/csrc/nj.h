#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mst {

struct NjJoin {
  int node, left, right;
  float d;
  double Ti, Tj;
  int n;
  double q;
};

struct NjArgs {
  float* D;        // n0 x n0 (leading dimension ld), modified in place
  int ld;
  int n0;
  double* T;       // n0
  int* act;        // 2*n0, first half initialised to 0..n0-1
  int* ids;        // 2*n0, first half initialised to 0..n0-1
  NjJoin* joins;   // n0-1
  float ratio;     // rebuild ratio (CmdLineOptions -r)
};

hipError_t launch_nj(const NjArgs& a, hipStream_t st);

}  // namespace mst

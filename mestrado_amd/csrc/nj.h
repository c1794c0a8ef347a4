#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mestrado.h"

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
  double* T;       // n0, by storage row
  double* Tpos;    // 2*n0: T by active position (ping-pong halves, as act and ids)
  int* act;        // 2*n0, first half initialised to 0..n0-1
  int* ids;        // 2*n0, first half initialised to 0..n0-1
  NjJoin* joins;   // n0-1
  float ratio;     // rebuild ratio (CmdLineOptions -r)
};

hipError_t launch_nj(const NjArgs& a, hipStream_t st);
// One workgroup per matrix: d_args[0 .. count) in device memory.
hipError_t launch_nj_batch(const NjArgs* d_args, int count, hipStream_t st);
// The join loop as launches per step (large N, one matrix): the O(m^2) scan
// on `nparts` workgroups, the join on one. `parts`: a device buffer of
// nj_part_bytes(nparts); act/ids set as for launch_nj.
hipError_t run_nj_steps(const NjArgs& a, void* parts, int nparts, hipStream_t st);
int64_t nj_part_bytes(int nparts);
// act/ids of NjArgs before the first step
void launch_nj_init(int* act, int* ids, int n, hipStream_t st);
// D = (float)scores, n elements (the NJ input of a score matrix)
void launch_scores_to_float(const int32_t* s, float* d, int64_t n, hipStream_t st);
// Join record of the C ABI with TreeNode::join's branch lengths.
void nj_to_join(const NjJoin& in, mst_join* out);

}  // namespace mst

// Batched form of the all-pairs scores: plan and tuning state.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "nw_score2.h"

struct BClass {
  int K;
  int64_t begin, end;
  int64_t work = 0;   // lane-cells the class processes
  int64_t cells = 0;  // useful DP cells
  int64_t pairs = 0;
  int W = 1;          // waves per pipeline (pipelined packed kernel)
};
// Tunables of the batched form (the autotuner writes them, see mst_autotune).
struct Tune {
  int chunk = 32;            // column sequences per shared-profile item
  int wg_per_cu3[17] = {0};  // score3 workgroups per CU by rows-per-lane K (0: default)
  int wg_per_cu2 = 0;        // score2 workgroups per CU (0: default)
};

struct BatchedPlan {
  bool built = false;
  std::vector<int32_t> xlist;  // sequence ids sorted by length
  std::vector<mst::Item3> items3;
  std::vector<BClass> c3;
  std::vector<mst::Pair2> items2;
  std::vector<BClass> c2;
  std::vector<mst::Item3> items5;  // pipelined packed kernel (score5)
  std::vector<BClass> c5;
  std::vector<int32_t> fb_i, fb_j;  // int32 fallback pairs (column, row)
  // device copies
  bool uploaded = false;
  mst::Item3* d_items3 = nullptr;
  mst::Pair2* d_items2 = nullptr;
  int32_t* d_xlist = nullptr;
  mst::Item3* d_items5 = nullptr;
  int8_t* d_scodes = nullptr;   // codes in length order
  int64_t* d_soff = nullptr;    // their offsets by sorted position
  unsigned int* d_counters = nullptr;
  int ncounters = 0;
  hipGraphExec_t graph = nullptr;
  int32_t* graph_out = nullptr;
};

void free_batched(BatchedPlan& P);

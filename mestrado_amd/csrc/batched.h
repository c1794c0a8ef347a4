// Batched form of the all-pairs scores: plan and tuning state.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "nw_score5.h"

struct BClass {
  int K;
  int64_t begin, end;
  int64_t work = 0;   // lane-cells the class processes
  int64_t cells = 0;  // useful DP cells
  int64_t pairs = 0;
  int W = 1;          // waves per pipeline (pipelined packed kernel)
  bool stacked = false;  // two row pairs per strip (one-wave kernel)
  bool rebase = false;   // score8: re-based items (one column sequence per wave)
  int64_t cap = 0;       // score8: most columns of one wave's part
};
// Tunables of the batched form (set through mst_tune_set; the autotuner in
// mestrado_amd/autotune.py searches them and persists the best per device).
struct Tune {
  int chunk = 32;            // column sequences per work item
  int wg5 = 0;               // score5 workgroups per CU (0: as many as LDS allows)
  int k5min = 1;             // least rows per lane of the one-wave streamed kernel (1..3)
  int chunk_w3 = 0;          // column sequences per item of 3- and 4-wave pipelines (0: chunk)
  int k1max = 13;            // most rows per lane of the one-wave streamed kernel (5..16)
  int k2max = 8;             // most rows per lane of the two-wave streamed kernel (4..8)
  int branches = 2;          // parallel branches of the graph's class launches (1..4)
  int wmax = 4;              // most waves of a one-sweep pipeline of 4 rows per lane (4..8);
                             // longer rows take the multi-sweep kernel
  int k6 = 0;                // rows per lane of the multi-sweep kernel: 3, 4, or 0 (per row pair)
  int chunk6 = 0;            // column sequences per item of the multi-sweep kernel (0: chunk)
  int chunk_k9 = 0;          // column sequences per item of one- and two-wave classes of 9+ rows
                             // per lane in all (0: chunk)
  int s6p2 = 0;              // 1: the multi-sweep classes as two pipelines per workgroup (score7)
  int stack = 4;             // two row pairs of K = 1..stack rows per lane share a strip of 2K
                             // (0: none; at most 4: stacking 5 and 6 measured slower,
                             // 216.1 and 219.7 against 213.1 ms on the 4,000-protein set)
  int rebase = 1;            // pairs past the 16-bit bound: the re-based packed kernel (score8)
                             // instead of the int32 kernel when 2, or when 1 and its items are
                             // many or end well within the rest of the pass; 0: the int32 kernel
  int k8 = 2;                // rows that take score8 (strips in turn on one wave): 0 none
                             // besides the re-based pairs, 1 rows above 1024 (instead of the
                             // multi-sweep kernel), 2 also the two-wave pipelines' rows
  int chunk8 = 4;            // column sequences per plain score8 item (four waves' parts)
};

struct BatchedPlan {
  bool built = false;
  std::vector<int32_t> xlist;  // sequence ids sorted by length
  std::vector<mst::Item3> items5;  // pipelined packed kernel (score5)
  std::vector<int2> pair2_5;       // their second row pair (stacked items; else -1)
  int2* d_pair2 = nullptr;
  std::vector<BClass> c5;
  std::vector<mst::Item3> items6;  // multi-sweep packed kernel (score6)
  std::vector<BClass> c6;
  int64_t carry_cols = 0;          // longest score6 column stream
  std::vector<mst::Item3> items8;  // strips-in-turn packed kernel (score8)
  std::vector<BClass> c8;
  std::vector<int32_t> fb_i, fb_j;  // int32 fallback pairs (column, row)
  // device copies
  bool uploaded = false;
  int32_t* d_xlist = nullptr;
  mst::Item3* d_items5 = nullptr;
  mst::Item3* d_items6 = nullptr;
  int2* d_carry = nullptr;      // score6 row buffers, one per workgroup
  int grid6 = 0;
  mst::Item3* d_items8 = nullptr;
  int2* d_carry8 = nullptr;     // score8 row buffers (per class, workgroup and wave)
  int2* d_frames8 = nullptr;    // and their frames (re-based classes)
  int grid8 = 0;
  std::vector<int64_t> carry8_off, frames8_off;  // per score8 class, in int2 elements
  int8_t* d_scodes = nullptr;   // codes in length order
  uint32_t* d_soffs8 = nullptr;  // the same codes as profile offsets (code x kProfStride8)
  int64_t ncodes = 0;
  int64_t* d_soff = nullptr;    // their offsets by sorted position
  unsigned int* d_counters = nullptr;
  int ncounters = 0;
  hipGraphExec_t graph = nullptr;
  int32_t* graph_out = nullptr;
};

void free_batched(BatchedPlan& P);

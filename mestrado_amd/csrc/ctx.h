// Context state behind the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <set>
#include <utility>
#include <string>
#include <vector>

#include "batched.h"

struct SeqSet {
  int n = 0;
  int64_t total = 0;
  int maxlen = 0;
  int8_t* d_codes = nullptr;  // reversed codes, concatenated
  int64_t* d_off = nullptr;   // n+1 offsets
  std::vector<int64_t> h_off;
  std::string h_canon;  // forward letters as the device codes print them
};

std::string canonical_letters(const char* ascii, int64_t n);

struct Bucket {
  int K;
  int64_t begin, end;
  int64_t cells = 0;
};

struct LaunchStat {
  int K;  // rows per lane; negative for the packed 16-bit kernels
  int64_t pairs, cells;
  float ms;
};

// A bucketed pair list (pairs sorted by bucket, longest first in a bucket).
struct PairPlan {
  bool built = false, uploaded = false;
  int2* d_pairs = nullptr;  // own device copy (fallback plans of the batched form)
  std::vector<int2> pairs;
  std::vector<int64_t> index;  // plan position -> caller's pair index
  std::vector<Bucket> buckets;
  int64_t cells = 0;
};

struct mst_ctx {
  int device = 0;
  int num_cus = 0;
  char arch[64] = {0};  // gcnArchName of the device, e.g. gfx950:sramecc+:xnack-
  int64_t hbm_bytes = 0;
  hipStream_t stream = nullptr;
  // branches of the batched form's graph (created on first use, default priority)
  hipStream_t bstream[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t bevent[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int go = 10, ge = 2;
  int8_t S[576];
  bool symmetric = true;
  int smax = 0;  // largest entry of S (at least 0)
  int smin = 0;  // least entry of S
  int8_t* d_S = nullptr;
  void* d_scratch = nullptr;  // this device's address of the no-allocation scratch array
  std::string err;
  std::map<int, SeqSet> sets;
  std::map<int, PairPlan> plans;  // all-pairs plan per set (plain form)
  std::map<int, BatchedPlan> bplans;  // all-pairs plan per set (batched form)
  std::map<int, PairPlan> fbplans;    // int32 fallback pairs of the batched plan
  Tune tune;
  std::set<std::pair<int, int>> prepared;  // (set, form) ready for timed runs
  int next_set = 1;
  int resident_plan = -1;

  int2* d_pairs = nullptr;
  size_t pairs_cap = 0;
  int2* d_bufs = nullptr;
  size_t bufs_cap = 0;
  int32_t* d_mat = nullptr;
  size_t mat_cap = 0;
  int32_t* d_pscore = nullptr;
  size_t pscore_cap = 0;
  int mat_set = -1;  // set whose all-pairs matrix d_mat holds

  float* d_nj = nullptr;
  size_t nj_cap = 0;
  double* d_nj_T = nullptr;
  size_t nj_T_cap = 0;
  int* d_nj_act = nullptr;
  size_t nj_act_cap = 0;
  void* d_nj_joins = nullptr;
  size_t nj_joins_cap = 0;
  void* d_nj_parts = nullptr;
  size_t nj_parts_cap = 0;

  // traceback working set
  int64_t flag_budget_bytes = (int64_t)16 << 30;  // flags and score tables of one alignment batch
  void* d_flags = nullptr;
  size_t flags_cap = 0;
  void* d_ops = nullptr;
  size_t ops_cap = 0;
  void* d_jobs = nullptr;
  size_t jobs_cap = 0;
  void* d_jscore = nullptr;
  size_t jscore_cap = 0;

  // result of the last mst_msa: msa_n rows of msa_len codes (24 = gap)
  int msa_n = 0, msa_len = 0;
  std::vector<uint8_t> msa_rows;

  // per-launch timing of the last resident all-pairs call
  std::vector<hipEvent_t> launch_events;
  std::vector<LaunchStat> launch_stats;

  ~mst_ctx();
  int fail(int code, const char* what, const char* detail = nullptr);
  unsigned int* d_acounter = nullptr;  // work queue heads: the pipelined flag kernel, the packed one
  size_t acounter_cap = 0;
  void* d_tiles = nullptr;  // tiled flag kernel: tile lists, row buffers, lane states
  size_t tiles_cap = 0;
  void* d_trows = nullptr;
  size_t trows_cap = 0;
  void* d_tstate = nullptr;
  size_t tstate_cap = 0;
  int ensure(void** p, size_t* cap, size_t bytes);
  // packed flag kernel: items and table strips, score tables of sequence jobs
  bool flag16 = true;  // (set by mst_create when the substitution bytes allow it)
  void* d_f16 = nullptr;
  size_t f16_cap = 0;
  void* d_tab16 = nullptr;
  size_t tab16_cap = 0;
  hipStream_t astream = nullptr;  // the int32 flag kernels beside the packed one (default priority)
};

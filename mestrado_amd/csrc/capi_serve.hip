// Served form: a request loop with continuous batching. Jobs (sets of
// sequences, and whether their NJ tree and progressive alignment are wanted
// besides the all-pairs scores) are submitted at any time; each call of
// mst_server_step retires finished device work, admits waiting jobs whose
// device footprint fits the budget (FIFO), and launches:
//  * score rounds: a quantum of work items drawn from the admitted jobs,
//    earliest deadline first, packed into shared launches of the streamed
//    packed kernels (nw_score5.hip, nw_score6.hip) whatever job they belong to;
//  * NJ launches (nj.hip, a workgroup per job) for the jobs whose scores are
//    complete;
//  * merge batches (msa.hip) of the profile merges whose children exist, over
//    every job in its alignment stage, earliest deadline first.
//
// Device state: one arena of the budget's size holds, per admitted job, its
// N x N int32 score matrix, its residue codes in length order and the ASCII
// they are encoded from, and (tree, alignment) the fp32 NJ matrix, NJ state,
// leaf rows and a region for group rows; a position table (soff: code
// offset, order: index in the job) names every admitted sequence; a job
// table gives each job's matrix offset and leading dimension, so a kernel
// writes a pair's score into its own job's matrix. Score rounds are enqueued
// on the context's stream, at most three in flight; inside a round the
// launches of different kernel classes run side by side on four side streams
// (default priority) after the round's copies on the context's stream; a
// round is over when every stream it used has passed its end event, and the
// next round's copies do not wait for it. NJ launches run on four streams and
// merge batches on eight, one batch per stream in flight, and a merge batch
// holds small merges (up to small_cells) or large ones, so a long NJ or a
// large merge does not hold back the small jobs behind it.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <numeric>
#include <vector>

#include "../../include/mestrado.h"
#include "ctx.h"
#include "encode.h"
#include "merge.h"
#include "nj.h"
#include "nw_score.h"
#include "nw_score5.h"
#include "pool.h"

using namespace mst;

#define CK(call)                                                                      \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) return ctx->fail(MST_E_HIP, #call, hipGetErrorString(e_)); \
  } while (0)

int64_t pk16_bound(const mst_ctx* ctx, int64_t R, int64_t n);
bool pk16_usable(const mst_ctx* ctx);

namespace {

__global__ void zero_u32s(unsigned int* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// scores of int32-kernel pairs (positions x, y) into their jobs' matrices
__global__ void scatter_pairs(const int32_t* vals, const int2* pairs, const int* pair_job,
                              const int64_t* job_out, const int* job_ld, const int32_t* order,
                              int32_t* out, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = pair_job[i];
  const int64_t base = job_out[j];
  const int64_t ld = job_ld[j];
  const int ox = order[pairs[i].x], oy = order[pairs[i].y];
  out[base + ox * ld + oy] = vals[i];
  out[base + (int64_t)oy * ld + ox] = vals[i];
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Classes of the streamed kernels: 0 = score6<4>; then score5<K, 1> for K =
// 1..16, score5<K, 2> for K = 3..8, score5<K, 3> and score5<K, 4> for K = 3..4
// (the routes of the batched form's planner); score6<3>; last, score8<K> for
// K = 4..8 (strips in turn on one wave, the rows above 64 k1max).
constexpr int kClassScore6K3 = 1 + 16 + 6 + 2 + 2;
constexpr int kClassScore8 = kClassScore6K3 + 1;  // + K - 4
constexpr int kClasses = kClassScore8 + (kScore8MaxK - kScore8MinK + 1);
// score8 items hold at most this many columns, so a wave's part always fits
// its row buffer
constexpr int64_t kServeScore8Cols = 16384;
int class_of(int kind, int K, int W) {
  if (kind == 6) return K == 3 ? kClassScore6K3 : 0;
  if (kind == 8) return kClassScore8 + (K - kScore8MinK);
  if (W == 1) return K;
  if (W == 2) return 17 + (K - 3);
  return 23 + (W - 3) * 2 + (K - 3);
}
void class_shape(int c, int* K, int* W) {
  if (c <= 16) {
    *W = 1;
    *K = c;
  } else if (c <= 22) {
    *W = 2;
    *K = c - 17 + 3;
  } else {
    *W = 3 + (c - 23) / 2;
    *K = 3 + (c - 23) % 2;
  }
}

// Stages of a running job.
enum { kStageScores = 0, kStageNj = 1, kStageMsa = 2, kStageDone = 3 };

// A group of the progressive alignment: k rows of len columns at arena
// offset off (a block of cap bytes, in the job's region or the arena).
struct Group {
  int64_t off = -1, cap = 0;
  int k = 0, len = 0;
  bool ready = false, leaf = false, in_arena = false;
};

struct Job {
  int64_t id = 0;
  int n = 0;
  int what = MST_SERVE_SCORES;
  int state = MST_JOB_WAITING;
  int stage = kStageScores;
  std::vector<char> ascii;        // input letters
  std::vector<int64_t> off;       // input offsets (n + 1)
  double t_submit = 0, t_admit = 0, t_done = 0, deadline = 0;
  double t_scores = 0, t_tree = 0;  // ends of the score and NJ stages
  // device placement
  int64_t arena_off = -1, arena_bytes = 0;  // one block: scores, codes, ascii, offsets, NJ, MSA
  int64_t o_codes = 0, o_D = 0, o_T = 0, o_act = 0, o_joins = 0, o_leaf = 0, o_msa = 0, msa_cap = 0;
  int64_t pos0 = -1;                        // positions pos0 .. pos0 + n (n + 1)
  int slot = -1;
  std::vector<int64_t> roff;                // code offsets in length order (n + 1)
  std::vector<int32_t> where;               // input index -> length-order position
  // plan: streamed items (positions) and int32 pairs (positions)
  std::vector<Item3> items;
  std::vector<int> item_cls;
  std::vector<int64_t> item_cells;
  std::vector<int2> fb;
  size_t next_item = 0, next_fb = 0;
  int64_t cells_total = 0, cells_left = 0;
  int refs = 0;  // in-flight rounds and batches holding work of this job
  bool cancel = false;
  // tree and alignment
  std::vector<mst_join> tree;
  std::vector<Group> groups;     // 2n - 1 nodes
  std::vector<char> merge_state; // per join: 0 waiting, 1 in flight, 2 done
  int merges_done = 0;
  bool leaves_done = false;
  Pool msa_pool;                 // over [o_msa, o_msa + msa_cap)
  // alignment-stage accounting: merge batches holding the job's merges, and
  // the stage's time with at least one of them in flight (busy) or none
  int msa_inflight = 0;
  int64_t msa_batches = 0;
  double msa_busy = 0, msa_idle = 0, t_msa_mark = 0;
  // score-stage accounting: rounds holding its items, the first's launch
  int64_t score_rounds = 0;
  double t_first_round = 0;
};

struct RoundBuf {
  Item3* d_items = nullptr;
  int* d_item_job = nullptr;
  int64_t cap_items = 0;
  int2* d_fb = nullptr;
  int* d_fb_job = nullptr;
  int32_t* d_fb_val = nullptr;
  int64_t cap_fb = 0;
  unsigned int* d_counters = nullptr;
  // pinned host staging of the round's items and int32 pairs
  Item3* h_items = nullptr;
  int* h_item_job = nullptr;
  int2* h_fb = nullptr;
  int* h_fb_job = nullptr;
  // events of the round using this buffer: device time (start, end), the
  // point the side streams start from, and each side stream's end
  hipEvent_t ev_start = nullptr, ev_end = nullptr, ev_go = nullptr;
  hipEvent_t ev_side[4] = {nullptr, nullptr, nullptr, nullptr};
};

struct Round {
  std::vector<hipEvent_t> evs;  // the round is over when all of these are
  int buf = 0;
  std::vector<int64_t> jobs;        // jobs with work in this round
  std::vector<int64_t> finishing;   // jobs whose last work is in this round
  int64_t cells = 0;
  double t_launch = 0;
};

}  // namespace

struct mst_server {
  mst_ctx* ctx = nullptr;
  int64_t budget = 0;
  uint8_t* d_arena = nullptr;
  Pool arena, pos, slots;
  int64_t* d_soff = nullptr;
  int32_t* d_order = nullptr;
  int64_t* d_job_out = nullptr;
  int* d_job_ld = nullptr;
  int2* d_carry = nullptr;
  int2* d_carry8 = nullptr;  // score8: a row buffer per wave and class
  int grid8 = 1;
  int k8 = 1;                // 1: the rows above 64 k1max on score8 (0: two-wave score5, score6)
  int grid6 = 1;
  int grid7 = 1;        // workgroups of the two-pipeline multi-sweep kernel
  int64_t carry_rows = 0;  // row buffers (of 16384 columns) per multi-sweep class
  int s6p2 = 1;         // 1: the multi-sweep classes as two pipelines per workgroup
  int2* d_bufs = nullptr;  // int32 kernel strip buffers
  int64_t bufs_cap_cols = 0;
  int grid32 = 1;
  static constexpr int kBufs = 3;  // round buffers: rounds in flight at most
  int max_rounds = 3;              // rounds kept in flight (1..kBufs)
  RoundBuf rb[kBufs];
  hipStream_t side[4] = {nullptr, nullptr, nullptr, nullptr};  // class launches of a round
  std::map<int64_t, Job> jobs;
  std::deque<int64_t> waiting;
  std::vector<int64_t> running;
  std::deque<Round> inflight;
  int64_t next_id = 1;
  double rate = 4e12;          // cells per second, running estimate
  double round_cells = 2e10;   // work per round (about 4 ms of the device)
  double round_cells_shared = 2e10;  // per round while jobs are in later stages
  double score_share = 1.0;          // share of the score kernels' grid meanwhile
  int persistent = 1;                // persistent score launches (else one workgroup per item)
  double slack = 2.0;          // deadline = submission + slack x estimated run time
  double deadline_add = 0.0;   //   + this many seconds
  double est_floor = 0.0;      // least estimated run time in that deadline (seconds)
  double t_retire = 0;         // when the last round was seen finished
  int chunk = 12;
  int chunk6 = 8;  // items of the multi-sweep kernel (0: chunk)
  int64_t done_count = 0, cells_done = 0;
  std::vector<int64_t> completed;  // finished since the last mst_server_completed
  // NJ: jobs whose scores are done; batches in flight (one per slot)
  float nj_ratio = 0.5f;
  std::deque<int64_t> nj_queue;
  struct NjSlot {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
    std::vector<int64_t> jobs;
    std::vector<int64_t> off;  // joins of jobs[i] at h_nj + off[i]
    NjArgs* d_args = nullptr;
    NjArgs* h_args = nullptr;
    NjJoin* h_nj = nullptr;
    int64_t args_cap = 0, hargs_cap = 0, nj_cap = 0;
  };
  static constexpr int kSlots = 4;
  NjSlot nj[kSlots];
  static constexpr int kMsaSlots = 8;
  // merges: batches in flight (one per slot)
  double merge_bytes = (double)(512 << 20);
  struct MsaSlot {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
    mst::MergeEngine eng;
    std::vector<std::pair<int64_t, int>> merges;  // (job, join) of each merge
    std::vector<int64_t> jobs;                    // jobs with work in the batch
    std::vector<int64_t> leaf_jobs;               // jobs whose leaves the batch writes
  };
  MsaSlot ms[kMsaSlots];
  double small_cells = 4e6;  // merges up to this many cells batch apart from larger ones
  double shared_backlog = 0; // > 0: rounds shrink for the later stages only below this backlog (cells)
};

namespace {

int64_t r256(int64_t x) { return (x + 255) / 256 * 256; }

// Group rows of a job's alignment held in its own region: room for the
// groups alive at once (each sequence is in one, at most ~maxlen columns wide
// with gaps) and the merges of a batch being written.
int64_t msa_region(const Job& J) {
  int64_t maxlen = 0;
  for (int i = 0; i < J.n; ++i) maxlen = std::max<int64_t>(maxlen, J.off[i + 1] - J.off[i]);
  return r256(4 * (int64_t)J.n * (maxlen + 64));
}

// Bytes of a job's block and the offsets of its parts (relative to the block).
struct Layout {
  int64_t scores, codes, ascii, roff, D, T, act, joins, leaf, msa, msa_cap, total;
};

Layout job_layout(const Job& J) {
  const int64_t N = J.n, L = J.off[J.n] - J.off[0];
  Layout l;
  int64_t o = 0;
  l.scores = o;
  o += r256(N * N * 4);
  l.codes = o;
  o += r256(L);
  l.ascii = o;
  o += r256(L);
  l.roff = o;
  o += r256((N + 1) * 8);
  l.D = l.T = l.act = l.joins = l.leaf = l.msa = o;
  l.msa_cap = 0;
  if ((J.what & MST_SERVE_TREE) == MST_SERVE_TREE) {
    l.D = o;
    o += r256(N * N * 4);
    l.T = o;
    o += r256(N * 8 * 3);  // T by storage row, then by position (two halves)
    l.act = o;
    o += r256(N * 16);
    l.joins = o;
    o += r256(N * (int64_t)sizeof(NjJoin));
  }
  if ((J.what & MST_SERVE_MSA) == MST_SERVE_MSA) {
    l.leaf = o;
    o += r256(std::max<int64_t>(L, 1));
    l.msa = o;
    l.msa_cap = msa_region(J);
    o += l.msa_cap;
  }
  l.total = o;
  return l;
}

int64_t job_footprint(const Job& J) { return job_layout(J).total; }

// Streamed plan of one job whose sequences sit at positions pos0 .. pos0+n-1
// in length order (lens ascending): the row pairs (2k, 2k+1) against every
// later position, cut into runs of `chunk` column sequences.
void plan_job(const mst_ctx* ctx, Job& J, const std::vector<int>& lens, int chunk, int chunk6, bool k8) {
  const int N = J.n;
  const int p0 = (int)J.pos0;
  auto add = [&](int kind, int K, int W, int64_t R, int pa, int pb, int first, int last) {
    std::vector<int> elig;
    for (int t = first; t <= last; ++t) {
      const int64_t nx = lens[t];
      if (kind == 0 || nx == 0 || pk16_bound(ctx, R, nx) > 32767 || (kind == 6 && nx > 16384) ||
          (kind == 8 && nx > kServeScore8Cols)) {
        J.fb.push_back(make_int2(p0 + t, p0 + pa));
        if (pb >= 0) J.fb.push_back(make_int2(p0 + t, p0 + pb));
      } else {
        elig.push_back(t);
      }
    }
    const int64_t rows = lens[pa] + (pb >= 0 ? lens[pb] : 0);
    size_t c0 = 0;
    while (c0 < elig.size()) {
      size_t c1 = c0 + 1;
      int64_t cols = lens[elig[c0]];
      while (c1 < elig.size() && (int)(c1 - c0) < (kind == 6 || kind == 8 ? chunk6 : chunk) &&
             elig[c1] == elig[c1 - 1] + 1 && (kind != 6 || cols + lens[elig[c1]] <= 16384) &&
             (kind != 8 || cols + lens[elig[c1]] <= kServeScore8Cols)) {
        cols += lens[elig[c1]];
        ++c1;
      }
      Item3 it;
      it.ya = p0 + pa;
      it.yb = pb >= 0 ? p0 + pb : -1;
      it.xbegin = p0 + elig[c0];
      it.xcount = (int)(c1 - c0);
      J.items.push_back(it);
      J.item_cls.push_back(class_of(kind, K, W));
      J.item_cells.push_back(cols * rows);
      c0 = c1;
    }
  };
  for (int k = 0; 2 * k < N; ++k) {
    const bool single = 2 * k + 1 >= N;
    const int pa = 2 * k, pb = single ? -1 : 2 * k + 1;
    const int m = single ? lens[pa] : lens[pb];
    int kind, K, W = 1;
    if (m == 0) {
      kind = 0;
      K = 3;
    } else if (m <= 64 * ctx->tune.k1max) {  // one wave (the batched form's routes)
      K = std::max(ctx->tune.k5min, (m + 63) / 64);
      kind = 5;
    } else if (k8) {
      // strips of at most 512 rows in turn on one wave (score8)
      const int ns = (m + 511) / 512;
      K = std::max(kScore8MinK, std::min(kScore8MaxK, (m + 64 * ns - 1) / (64 * ns)));
      kind = 8;
    } else if ((m + 127) / 128 <= ctx->tune.k2max) {  // two waves
      W = 2;
      K = (m + 127) / 128;
      kind = 5;
    } else if (m <= 1024) {
      W = (m + 255) / 256;
      K = (m + 64 * W - 1) / (64 * W);
      kind = 5;
    } else {
      // the multi-sweep kernel in strips of 3 or 4 rows per lane, chosen as
      // the batched form's planner does
      kind = 6;
      K = ctx->tune.k6;
      if (K == 0) {
        auto cap = [&](int k) { return (double)((m + 256 * k - 1) / (256 * k)) * 256 * k; };
        K = cap(3) / 0.89 < cap(4) ? 3 : 4;
      }
    }
    const int64_t R = kind == 6 || kind == 8 ? (int64_t)(m + 64 * K - 1) / (64 * K) * 64 * K : (int64_t)W * 64 * K;
    if (single) {
      add(kind, K, W, R, pa, -1, pa, N - 1);
    } else if (lens[pa] == 0) {
      // an empty row sequence has no last row to emit its scores from: its
      // pairs take the int32 kernel, pb runs alone
      for (int t = pa; t < N; ++t) J.fb.push_back(make_int2(p0 + t, p0 + pa));
      add(kind, K, W, R, pb, -1, pb, N - 1);
    } else {
      add(kind, K, W, R, pa, -1, pa, pa);
      add(kind, K, W, R, pa, pb, pb, N - 1);
    }
  }
  J.cells_total = 0;
  for (int64_t c : J.item_cells) J.cells_total += c;
  for (const int2& p : J.fb) J.cells_total += (int64_t)lens[p.x - p0] * lens[p.y - p0];
  J.cells_left = J.cells_total;
}

void release_job(mst_server* S, Job& J) {
  for (Group& G : J.groups)
    if (G.in_arena && G.off >= 0) S->arena.release(G.off, G.cap);
  J.groups.clear();
  J.groups.shrink_to_fit();
  if (J.arena_off >= 0) S->arena.release(J.arena_off, J.arena_bytes);
  if (J.pos0 >= 0) S->pos.release(J.pos0, J.n + 1);
  if (J.slot >= 0) S->slots.release(J.slot, 1);
  J.arena_off = J.pos0 = -1;
  J.slot = -1;
  J.items.clear();
  J.items.shrink_to_fit();
  J.item_cls.clear();
  J.item_cells.clear();
  J.fb.clear();
  J.ascii.clear();
  J.ascii.shrink_to_fit();
}

// Moves job J to the device and plans it. Returns MST_OK, or 1 when it does
// not fit now (nothing changed), or an error.
int admit(mst_server* S, Job& J) {
  mst_ctx* ctx = S->ctx;
  const int N = J.n;
  const int64_t bytes = job_footprint(J);
  const int64_t a = S->arena.alloc(bytes);
  if (a < 0) return 1;
  const int64_t p = S->pos.alloc(N + 1);
  if (p < 0) {
    S->arena.release(a, bytes);
    return 1;
  }
  const int64_t sl = S->slots.alloc(1);
  if (sl < 0) {
    S->arena.release(a, bytes);
    S->pos.release(p, N + 1);
    return 1;
  }
  J.arena_off = a;
  J.arena_bytes = bytes;
  J.pos0 = p;
  J.slot = (int)sl;
  // length order, then the layout of the block
  std::vector<int> ord(N);
  std::iota(ord.begin(), ord.end(), 0);
  auto len = [&](int i) { return (int)(J.off[i + 1] - J.off[i]); };
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return len(x) < len(y); });
  const int64_t L = J.off[N] - J.off[0];
  const Layout lay = job_layout(J);
  const int64_t o_scores = a + lay.scores, o_codes = a + lay.codes, o_ascii = a + lay.ascii,
                o_roff = a + lay.roff;
  J.o_codes = o_codes;
  J.o_D = a + lay.D;
  J.o_T = a + lay.T;
  J.o_act = a + lay.act;
  J.o_joins = a + lay.joins;
  J.o_leaf = a + lay.leaf;
  J.o_msa = a + lay.msa;
  J.msa_cap = lay.msa_cap;
  J.msa_pool.init(lay.msa_cap, 256);
  std::vector<char> asc(std::max<int64_t>(L, 1));
  std::vector<int64_t> roff(N + 1, 0), soff(N + 1, 0);
  std::vector<int32_t> order(N + 1, 0);
  std::vector<int> lens(N);
  for (int i = 0; i < N; ++i) {
    const int s = ord[i];
    memcpy(asc.data() + roff[i], J.ascii.data() + (J.off[s] - J.off[0]), len(s));
    roff[i + 1] = roff[i] + len(s);
    lens[i] = len(s);
    order[i] = s;
  }
  for (int i = 0; i <= N; ++i) soff[i] = o_codes + roff[i];
  CK(hipMemcpy(S->d_arena + o_ascii, asc.data(), std::max<int64_t>(L, 1), hipMemcpyHostToDevice));
  CK(hipMemcpy(S->d_arena + o_roff, roff.data(), sizeof(int64_t) * (N + 1), hipMemcpyHostToDevice));
  CK(hipMemcpy(S->d_soff + p, soff.data(), sizeof(int64_t) * (N + 1), hipMemcpyHostToDevice));
  CK(hipMemcpy(S->d_order + p, order.data(), sizeof(int32_t) * (N + 1), hipMemcpyHostToDevice));
  const int64_t jo = o_scores / 4;
  const int ld = N;
  CK(hipMemcpy(S->d_job_out + sl, &jo, sizeof jo, hipMemcpyHostToDevice));
  CK(hipMemcpy(S->d_job_ld + sl, &ld, sizeof ld, hipMemcpyHostToDevice));
  launch_encode((const char*)(S->d_arena + o_ascii), (const int64_t*)(S->d_arena + o_roff), N,
                (int8_t*)(S->d_arena + o_codes), ctx->stream);
  CK(hipGetLastError());
  plan_job(ctx, J, lens, S->chunk, S->chunk6 > 0 ? S->chunk6 : S->chunk, S->k8 != 0);
  J.roff.assign(roff.begin(), roff.end());
  J.where.assign(N, 0);
  for (int i = 0; i < N; ++i) J.where[order[i]] = i;
  J.state = MST_JOB_RUNNING;
  J.stage = kStageScores;
  J.t_admit = now_s();
  // deadline: slack x the job's estimated run time alone, all its stages:
  // scores at the measured rate; NJ on one workgroup ~ 34 ms x (N/512)^2;
  // the alignment ~ 60 us per sequence (its levels, each a merge batch)
  double est = (double)J.cells_total / S->rate + 0.0005;
  const double nn = (double)N / 512.0;
  if ((J.what & MST_SERVE_TREE) == MST_SERVE_TREE && N >= 2) est += 0.034 * nn * nn;
  if ((J.what & MST_SERVE_MSA) == MST_SERVE_MSA) est += 60e-6 * N;
  // a floor on the estimate: jobs below it share one deadline horizon, so a
  // large job is not passed by every smaller one that arrives after it
  est = std::max(est, S->est_floor);
  J.deadline = J.t_submit + S->slack * est + S->deadline_add;
  J.ascii.clear();
  J.ascii.shrink_to_fit();
  return MST_OK;
}

int ensure_round_buf(mst_server* S, RoundBuf& B, int64_t nitems, int64_t nfb) {
  mst_ctx* ctx = S->ctx;
  if (!B.d_counters) {
    CK(hipMalloc(&B.d_counters, sizeof(unsigned int) * (kClasses + 1)));
    CK(hipEventCreate(&B.ev_start));
    CK(hipEventCreate(&B.ev_end));
    CK(hipEventCreateWithFlags(&B.ev_go, hipEventDisableTiming));
    for (auto& e : B.ev_side) CK(hipEventCreate(&e));
  }
  // growing frees buffers, which waits for the device; capacities double
  if (nitems > B.cap_items) {
    if (B.d_items) CK(hipFree(B.d_items));
    if (B.d_item_job) CK(hipFree(B.d_item_job));
    if (B.h_items) CK(hipHostFree(B.h_items));
    if (B.h_item_job) CK(hipHostFree(B.h_item_job));
    B.cap_items = std::max<int64_t>(std::max<int64_t>(nitems, 4096), 2 * B.cap_items);
    CK(hipMalloc(&B.d_items, sizeof(Item3) * B.cap_items));
    CK(hipMalloc(&B.d_item_job, sizeof(int) * B.cap_items));
    CK(hipHostMalloc(&B.h_items, sizeof(Item3) * B.cap_items, hipHostMallocDefault));
    CK(hipHostMalloc(&B.h_item_job, sizeof(int) * B.cap_items, hipHostMallocDefault));
  }
  if (nfb > B.cap_fb) {
    if (B.d_fb) CK(hipFree(B.d_fb));
    if (B.d_fb_job) CK(hipFree(B.d_fb_job));
    if (B.d_fb_val) CK(hipFree(B.d_fb_val));
    if (B.h_fb) CK(hipHostFree(B.h_fb));
    if (B.h_fb_job) CK(hipHostFree(B.h_fb_job));
    B.cap_fb = std::max<int64_t>(std::max<int64_t>(nfb, 1024), 2 * B.cap_fb);
    CK(hipMalloc(&B.d_fb, sizeof(int2) * B.cap_fb));
    CK(hipMalloc(&B.d_fb_job, sizeof(int) * B.cap_fb));
    CK(hipMalloc(&B.d_fb_val, sizeof(int32_t) * B.cap_fb));
    CK(hipHostMalloc(&B.h_fb, sizeof(int2) * B.cap_fb, hipHostMallocDefault));
    CK(hipHostMalloc(&B.h_fb_job, sizeof(int) * B.cap_fb, hipHostMallocDefault));
  }
  return MST_OK;
}

// Builds and enqueues one round. *launched = 0 when there is no work.
int launch_round(mst_server* S, int* launched) {
  mst_ctx* ctx = S->ctx;
  *launched = 0;
  std::vector<int64_t> order;
  for (int64_t id : S->running) {
    Job& J = S->jobs[id];
    if (!J.cancel && (J.next_item < J.items.size() || J.next_fb < J.fb.size())) order.push_back(id);
  }
  if (order.empty()) return MST_OK;
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t x, int64_t y) { return S->jobs[x].deadline < S->jobs[y].deadline; });
  int bi = 0;  // a buffer no in-flight round uses
  for (; bi < mst_server::kBufs; ++bi) {
    bool busy = false;
    for (const Round& q : S->inflight) busy |= q.buf == bi;
    if (!busy) break;
  }
  if (bi == mst_server::kBufs) return MST_OK;
  // a round buffer is rewritten only after the round that used it retired
  RoundBuf& B = S->rb[bi];
  // while jobs are in their NJ or alignment stage, rounds are shorter and
  // leave part of the device free, so those stages' launches do not wait
  // behind a whole round of persistent score kernels
  bool later = false;
  for (int64_t id : S->running) later |= S->jobs[id].stage == kStageNj || S->jobs[id].stage == kStageMsa;
  if (later && S->shared_backlog > 0) {
    // with a large backlog of score work the device's throughput counts
    // more than the later stages' latency: whole rounds
    double backlog = 0;
    for (int64_t id : order) backlog += (double)S->jobs[id].cells_left;
    if (backlog > S->shared_backlog) later = false;
  }
  const int64_t budget = (int64_t)(later ? std::min(S->round_cells, S->round_cells_shared) : S->round_cells);
  const double share = later ? S->score_share : 1.0;
  int64_t cells = 0;
  std::vector<std::vector<std::pair<Item3, int>>> by_cls(kClasses);
  std::vector<int2> fb;
  std::vector<int> fb_job;
  Round R;
  R.buf = bi;
  // items of job id until the round holds `limit` cells
  auto take = [&](int64_t id, int64_t limit) {
    Job& J = S->jobs[id];
    while (J.next_item < J.items.size() && cells < limit) {
      const size_t q = J.next_item++;
      by_cls[J.item_cls[q]].push_back(std::make_pair(J.items[q], J.slot));
      cells += J.item_cells[q];
      J.cells_left -= J.item_cells[q];
    }
    while (J.next_item == J.items.size() && J.next_fb < J.fb.size() && cells < limit) {
      const int2 p = J.fb[J.next_fb++];
      fb.push_back(p);
      fb_job.push_back(J.slot);
      cells += 1;  // the int32 kernel's share is small; it is timed with the round
    }
  };
  for (int64_t id : order) {
    if (cells >= budget) break;
    Job& J = S->jobs[id];
    take(id, budget);
    J.refs += 1;
    if (J.score_rounds++ == 0) J.t_first_round = now_s();
    R.jobs.push_back(id);
    if (J.next_item == J.items.size() && J.next_fb == J.fb.size()) R.finishing.push_back(id);
  }
  int64_t nitems = 0;
  for (auto& v : by_cls) nitems += (int64_t)v.size();
  int rc = ensure_round_buf(S, B, std::max<int64_t>(nitems, 1), std::max<int64_t>((int64_t)fb.size(), 1));
  if (rc) return rc;
  std::vector<std::pair<int64_t, int64_t>> ranges(kClasses);
  int64_t q = 0;
  for (int c = 0; c < kClasses; ++c) {
    ranges[c].first = q;
    for (auto& pr : by_cls[c]) {
      B.h_items[q] = pr.first;
      B.h_item_job[q] = pr.second;
      ++q;
    }
    ranges[c].second = q;
  }
  CK(hipEventRecord(B.ev_start, ctx->stream));
  if (nitems > 0) {
    CK(hipMemcpyAsync(B.d_items, B.h_items, sizeof(Item3) * nitems, hipMemcpyHostToDevice,
                      ctx->stream));
    CK(hipMemcpyAsync(B.d_item_job, B.h_item_job, sizeof(int) * nitems, hipMemcpyHostToDevice,
                      ctx->stream));
  }
  hipLaunchKernelGGL(zero_u32s, dim3(1), dim3(64), 0, ctx->stream, B.d_counters, kClasses + 1);
  CK(hipGetLastError());
  // the classes run side by side on the side streams, after everything the
  // context's stream holds (encodes of new jobs, this round's copies)
  CK(hipEventRecord(B.ev_go, ctx->stream));
  bool used[4] = {false, false, false, false};
  int nside = 0;
  for (int c = 0; c < kClasses; ++c) {
    if (ranges[c].second == ranges[c].first) continue;
    // the multi-sweep and score8 classes keep a side stream each (score6:
    // 0 for K = 4, 1 for K = 3; score8: 0 for K = 8, 1 for 7, 2 for 6, 3 for 5
    // and 4): their row buffers are per class, so launches of one class, of
    // this round and the rounds still in flight, must not overlap
    const int k8c = c >= kClassScore8 ? c - kClassScore8 + kScore8MinK : 0;
    const int si = c == 0 ? 0
                   : c == kClassScore6K3 ? 1
                   : k8c ? (k8c >= 6 ? 8 - k8c : 3)
                         : nside++ % 4;
    hipStream_t st = S->side[si];
    if (!used[si]) {
      CK(hipStreamWaitEvent(st, B.ev_go, 0));
      used[si] = true;
    }
    Score5Args a;
    a.scodes = (const int8_t*)S->d_arena;
    a.soff = S->d_soff;
    a.order = S->d_order;
    a.items = B.d_items;
    a.item_job = B.d_item_job;
    a.job_out = S->d_job_out;
    a.job_ld = S->d_job_ld;
    a.first = ranges[c].first;
    a.last = ranges[c].second;
    a.counter = B.d_counters + c;
    a.out_mat = (int32_t*)S->d_arena;
    a.nseq = 0;
    a.subst = ctx->d_S;
    a.go = ctx->go;
    a.ge = ctx->ge;
    a.carry = S->d_carry + (c == kClassScore6K3 ? S->carry_rows * 16384 : 0);
    a.carry_cap = 16384;
    const int64_t nit = ranges[c].second - ranges[c].first;
    if (k8c) {
      // the score8 row buffers, at the first round that needs them (a region
      // per class; ~2 GiB on 256 compute units)
      if (!S->d_carry8)
        CK(hipMalloc(&S->d_carry8, sizeof(int2) * (size_t)(kScore8MaxK - kScore8MinK + 1) * S->grid8 *
                                       kScore8Waves * kServeScore8Cols));
      a.carry = S->d_carry8 + (int64_t)(k8c - kScore8MinK) * S->grid8 * kScore8Waves * kServeScore8Cols;
      a.carry_cap = kServeScore8Cols;
      launch_score8(k8c, false, a,
                    (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)(S->grid8 * share), nit)), st);
    } else if (c == 0 || c == kClassScore6K3) {
      if (S->s6p2)
        launch_score7(c == 0 ? 4 : 3, a, (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)(S->grid7 * share), nit)), st);
      else
        launch_score6(c == 0 ? 4 : 3, a, (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)(S->grid6 * share), nit)), st);
    } else {
      int W, K;
      class_shape(c, &K, &W);
      // persistent: as many workgroups as fit, each taking items until none
      // are left; else a workgroup per item, so workgroups retire throughout
      // the launch and other stages' kernels find free slots
      const int64_t g = S->persistent
                            ? std::min<int64_t>((int64_t)(ctx->num_cus * score5_groups_per_cu(K, W) * share), nit)
                            : nit;
      launch_score5(K, W, a, (int)std::max<int64_t>(g, 1), st);
    }
    CK(hipGetLastError());
  }
  if (!fb.empty()) {
    // the int32 kernel on positions, then into the jobs' matrices
    std::copy(fb.begin(), fb.end(), B.h_fb);
    std::copy(fb_job.begin(), fb_job.end(), B.h_fb_job);
    CK(hipMemcpyAsync(B.d_fb, B.h_fb, sizeof(int2) * fb.size(), hipMemcpyHostToDevice, ctx->stream));
    CK(hipMemcpyAsync(B.d_fb_job, B.h_fb_job, sizeof(int) * fb.size(), hipMemcpyHostToDevice,
                      ctx->stream));
    ScoreArgs a;
    a.seqs = (const int8_t*)S->d_arena;
    a.off = S->d_soff;
    a.pairs = B.d_fb;
    a.first = 0;
    a.last = (int64_t)fb.size();
    a.out_pair = B.d_fb_val;
    a.out_mat = nullptr;
    a.nseq = 0;
    a.bufs = S->d_bufs;
    a.bufcap = S->bufs_cap_cols;
    a.subst = ctx->d_S;
    a.go = ctx->go;
    a.ge = ctx->ge;
    a.transpose_ok = ctx->symmetric ? 1 : 0;
    launch_score_pairs(16, a, (int)std::min<int64_t>(S->grid32, (int64_t)fb.size()), ctx->stream);
    CK(hipGetLastError());
    const int n = (int)fb.size();
    hipLaunchKernelGGL(scatter_pairs, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, B.d_fb_val,
                       B.d_fb, B.d_fb_job, S->d_job_out, S->d_job_ld, S->d_order,
                       (int32_t*)S->d_arena, n);
    CK(hipGetLastError());
  }
  // the round ends with its side streams; the next round's copies and
  // encodes on the context's stream do not wait for it (other buffer)
  for (int si = 0; si < 4; ++si) {
    if (!used[si]) continue;
    CK(hipEventRecord(B.ev_side[si], S->side[si]));
    R.evs.push_back(B.ev_side[si]);
  }
  CK(hipEventRecord(B.ev_end, ctx->stream));
  R.evs.push_back(B.ev_end);
  R.cells = cells;
  R.t_launch = now_s();
  S->inflight.push_back(R);
  *launched = 1;
  return MST_OK;
}

void job_done(mst_server* S, Job& J, double t, int* ndone) {
  J.state = MST_JOB_DONE;
  J.stage = kStageDone;
  J.t_done = t;
  S->done_count += 1;
  S->cells_done += J.cells_total;
  S->completed.push_back(J.id);
  if (ndone) *ndone += 1;
}

// Alignment stage: leaves (input order) in the job's leaf block, no group
// rows yet; the leaves are written by the next merge batch.
void start_msa(Job& J) {
  const int N = J.n;
  J.stage = kStageMsa;
  J.groups.assign(2 * (size_t)N - 1, Group());
  for (int i = 0; i < N; ++i) {
    Group& G = J.groups[i];
    G.off = J.o_leaf + (J.off[i] - J.off[0]);
    G.k = 1;
    G.len = (int)(J.off[i + 1] - J.off[i]);
    G.leaf = true;
  }
  J.merge_state.assign(N > 1 ? N - 1 : 0, 0);
  J.merges_done = 0;
  J.leaves_done = false;
  J.msa_inflight = 0;
  J.msa_batches = 0;
  J.msa_busy = J.msa_idle = 0;
  J.t_msa_mark = now_s();
}

// The job's score rounds are over: on to NJ, the alignment or done.
void scores_done(mst_server* S, Job& J, double t, int* ndone) {
  J.t_scores = t;
  if ((J.what & MST_SERVE_TREE) == MST_SERVE_TREE && J.n >= 2) {
    J.stage = kStageNj;
    S->nj_queue.push_back(J.id);
  } else if ((J.what & MST_SERVE_MSA) == MST_SERVE_MSA) {
    start_msa(J);  // one sequence: its leaf is the alignment
  } else {
    job_done(S, J, t, ndone);
  }
}

// Drops finished or cancelled jobs from the running list; a cancelled job is
// released once no round or batch in flight holds it.
void prune_running(mst_server* S) {
  std::vector<int64_t> keep;
  for (int64_t id : S->running) {
    Job& J = S->jobs[id];
    if (J.state == MST_JOB_RUNNING && !J.cancel) {
      keep.push_back(id);
    } else if (J.cancel && J.refs == 0) {
      release_job(S, J);
      J.state = MST_JOB_CANCELLED;
    } else if (J.cancel) {
      keep.push_back(id);
    }
  }
  S->running.swap(keep);
}

// Retires the oldest in-flight round if it has finished (or wait: block).
int retire(mst_server* S, bool wait, int* ndone) {
  mst_ctx* ctx = S->ctx;
  if (S->inflight.empty()) return MST_OK;
  Round& R = S->inflight.front();
  for (hipEvent_t ev : R.evs) {
    hipError_t e = wait ? hipEventSynchronize(ev) : hipEventQuery(ev);
    if (e == hipErrorNotReady) return MST_OK;
    if (e != hipSuccess) return ctx->fail(MST_E_HIP, "round", hipGetErrorString(e));
  }
  const double t = now_s();
  // throughput estimate from the round's device time
  float ms = 0;
  for (hipEvent_t ev : R.evs) {
    float m = 0;
    if (hipEventElapsedTime(&m, S->rb[R.buf].ev_start, ev) == hipSuccess) ms = std::max(ms, m);
  }
  if (ms > 0 &&
      R.cells > (int64_t)(0.25 * S->round_cells))
    S->rate = 0.8 * S->rate + 0.2 * std::min(2e13, std::max(1e11, (double)R.cells / (ms * 1e-3)));
  S->t_retire = t;
  for (int64_t id : R.jobs) {
    Job& J = S->jobs[id];
    J.refs -= 1;
  }
  for (int64_t id : R.finishing) {
    Job& J = S->jobs[id];
    if (J.state == MST_JOB_RUNNING && !J.cancel) scores_done(S, J, t, ndone);
  }
  S->inflight.pop_front();
  prune_running(S);
  return MST_OK;
}

template <typename T>
int grow_buf(mst_ctx* ctx, T** p, int64_t* cap, int64_t count, bool pinned) {
  if (count <= *cap) return MST_OK;
  const int64_t want = std::max<int64_t>(count, *cap * 2);
  if (*p) CK(pinned ? hipHostFree(*p) : hipFree(*p));
  *p = nullptr;
  *cap = 0;
  if (pinned)
    CK(hipHostMalloc((void**)p, sizeof(T) * want, hipHostMallocDefault));
  else
    CK(hipMalloc((void**)p, sizeof(T) * want));
  *cap = want;
  return MST_OK;
}

// One NJ launch for the jobs whose scores are done, on a free slot: per job
// the fp32 matrix and NJ state in its block, then a workgroup each, then the
// joins to the slot's pinned staging.
int launch_nj_jobs(mst_server* S) {
  mst_ctx* ctx = S->ctx;
  if (S->nj_queue.empty()) return MST_OK;
  mst_server::NjSlot* slot = nullptr;
  for (auto& q : S->nj)
    if (!q.busy) {
      slot = &q;
      break;
    }
  if (!slot) return MST_OK;
  std::vector<int64_t> ids;
  int64_t total = 0;
  while (!S->nj_queue.empty() && ids.size() < 4096) {
    const int64_t id = S->nj_queue.front();
    S->nj_queue.pop_front();
    Job& J = S->jobs[id];
    if (J.cancel || J.state != MST_JOB_RUNNING) continue;
    ids.push_back(id);
    total += J.n - 1;
  }
  if (ids.empty()) return MST_OK;
  const int cnt = (int)ids.size();
  int rc;
  if ((rc = grow_buf(ctx, &slot->d_args, &slot->args_cap, cnt, false))) return rc;
  if ((rc = grow_buf(ctx, &slot->h_args, &slot->hargs_cap, cnt, true))) return rc;
  if ((rc = grow_buf(ctx, &slot->h_nj, &slot->nj_cap, std::max<int64_t>(total, 1), true))) return rc;
  slot->off.clear();
  int64_t o = 0;
  for (int k = 0; k < cnt; ++k) {
    Job& J = S->jobs[ids[k]];
    const int n = J.n;
    float* D = (float*)(S->d_arena + J.o_D);
    int* act = (int*)(S->d_arena + J.o_act);
    launch_scores_to_float((const int32_t*)(S->d_arena + J.arena_off), D, (int64_t)n * n, slot->st);
    launch_nj_init(act, act + 2 * n, n, slot->st);
    NjArgs& a = slot->h_args[k];
    a.D = D;
    a.ld = n;
    a.n0 = n;
    a.T = (double*)(S->d_arena + J.o_T);
    a.Tpos = a.T + n;
    a.act = act;
    a.ids = act + 2 * n;
    a.joins = (NjJoin*)(S->d_arena + J.o_joins);
    a.ratio = S->nj_ratio;
    slot->off.push_back(o);
    o += n - 1;
    J.refs += 1;
  }
  CK(hipGetLastError());
  CK(hipMemcpyAsync(slot->d_args, slot->h_args, sizeof(NjArgs) * cnt, hipMemcpyHostToDevice, slot->st));
  CK(launch_nj_batch(slot->d_args, cnt, slot->st));
  for (int k = 0; k < cnt; ++k) {
    Job& J = S->jobs[ids[k]];
    CK(hipMemcpyAsync(slot->h_nj + slot->off[k], S->d_arena + J.o_joins, sizeof(NjJoin) * (J.n - 1),
                      hipMemcpyDeviceToHost, slot->st));
  }
  CK(hipEventRecord(slot->ev, slot->st));
  slot->jobs.swap(ids);
  slot->busy = true;
  return MST_OK;
}

// Retires the finished NJ batches (wait: block on the first busy one).
int retire_nj(mst_server* S, bool wait, int* ndone) {
  mst_ctx* ctx = S->ctx;
  for (auto& q : S->nj) {
    if (!q.busy) continue;
    hipError_t e = wait ? hipEventSynchronize(q.ev) : hipEventQuery(q.ev);
    if (e == hipErrorNotReady) continue;
    if (e != hipSuccess) return ctx->fail(MST_E_HIP, "NJ batch", hipGetErrorString(e));
    wait = false;
    const double t = now_s();
    for (size_t k = 0; k < q.jobs.size(); ++k) {
      Job& J = S->jobs[q.jobs[k]];
      J.refs -= 1;
      if (J.cancel) continue;
      J.tree.resize(J.n - 1);
      for (int r = 0; r + 1 < J.n; ++r) nj_to_join(q.h_nj[q.off[k] + r], &J.tree[r]);
      J.t_tree = t;
      if ((J.what & MST_SERVE_MSA) == MST_SERVE_MSA)
        start_msa(J);
      else
        job_done(S, J, t, ndone);
    }
    q.jobs.clear();
    q.busy = false;
  }
  prune_running(S);
  return MST_OK;
}

bool nj_busy(const mst_server* S) {
  for (const auto& q : S->nj)
    if (q.busy) return true;
  return false;
}

bool msa_busy(const mst_server* S) {
  for (const auto& q : S->ms)
    if (q.busy) return true;
  return false;
}

// Merge batches on the free slots: a batch writes the leaves of jobs new to
// the alignment stage, then every merge whose children exist, jobs earliest
// deadline first, while its scratch stays within merge_bytes. A merge's rows
// go to its job's region, else the arena; a merge with no room waits. Groups
// a batch writes (leaves included) are ready for other batches once it is
// retired; within the batch, stream order makes leaves ready for its merges.
int launch_merges(mst_server* S) {
  mst_ctx* ctx = S->ctx;
  for (auto& slot : S->ms) {
    if (slot.busy) continue;
    std::vector<int64_t> order;
    for (int64_t id : S->running) {
      Job& J = S->jobs[id];
      if (J.state == MST_JOB_RUNNING && !J.cancel && J.stage == kStageMsa) order.push_back(id);
    }
    if (order.empty()) return MST_OK;
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t x, int64_t y) { return S->jobs[x].deadline < S->jobs[y].deadline; });
    std::vector<LeafDesc> leaves;
    std::vector<MergeSpec> spec;
    slot.merges.clear();
    slot.jobs.clear();
    slot.leaf_jobs.clear();
    const int64_t budget = (int64_t)S->merge_bytes;
    int64_t bytes = 0;
    bool full = false;
    int size_class = -1;  // a batch holds small merges or large ones, not both
    for (int64_t id : order) {
      Job& J = S->jobs[id];
      bool touch = false;
      if (!J.leaves_done) {
        for (int i = 0; i < J.n; ++i) {
          Group& G = J.groups[i];
          LeafDesc l;
          l.src = J.roff[J.where[i]] + J.o_codes;
          l.dst = S->d_arena + G.off;
          l.len = G.len;
          l.pad = 0;
          leaves.push_back(l);
          G.ready = true;  // for this batch's merges only; reset below
        }
        J.leaves_done = true;
        slot.leaf_jobs.push_back(id);
        touch = true;
      }
      for (int k = 0; k + 1 < J.n && !full; ++k) {
        if (J.merge_state[k] != 0) continue;
        const mst_join& jn = J.tree[k];
        Group& X = J.groups[jn.left];
        Group& Y = J.groups[jn.right];
        if (!X.ready || !Y.ready) continue;
        const int cls = (double)X.len * Y.len <= S->small_cells ? 0 : 1;
        if (size_class >= 0 && cls != size_class) continue;
        const int64_t b = merge_scratch_bytes(X.len, Y.len);
        if (!spec.empty() && bytes + b > budget) {
          full = true;
          break;
        }
        Group& Z = J.groups[jn.node];
        const int64_t cap = std::max<int64_t>((int64_t)(X.k + Y.k) * ((int64_t)X.len + Y.len), 1);
        int64_t off = J.msa_pool.alloc(cap);
        bool in_arena = false;
        if (off >= 0) {
          off += J.o_msa;
        } else {
          off = S->arena.alloc(cap);
          in_arena = true;
          if (off < 0) continue;  // no room now: waits for memory
        }
        Z.off = off;
        Z.cap = cap;
        Z.in_arena = in_arena;
        Z.k = X.k + Y.k;
        bytes += b;
        size_class = cls;
        J.merge_state[k] = 1;
        spec.push_back(MergeSpec{S->d_arena + X.off, X.k, X.len, S->d_arena + Y.off, Y.k, Y.len,
                                 S->d_arena + Z.off});
        slot.merges.push_back(std::make_pair(id, k));
        touch = true;
      }
      if (touch) {
        slot.jobs.push_back(id);
        J.refs += 1;
      }
      if (full) break;
    }
    for (int64_t id : slot.leaf_jobs) {
      Job& J = S->jobs[id];
      for (int i = 0; i < J.n; ++i) J.groups[i].ready = false;
    }
    if (slot.jobs.empty()) return MST_OK;
    CK(slot.eng.leaves((const int8_t*)S->d_arena, leaves.data(), (int)leaves.size(), slot.st));
    CK(slot.eng.enqueue(spec.data(), (int)spec.size(), ctx->d_S, ctx->go, ctx->ge, ctx->num_cus, slot.st,
                        ctx->flag16 ? ctx->smax : -1));
    CK(hipEventRecord(slot.ev, slot.st));
    slot.busy = true;
    const double t = now_s();
    for (int64_t id : slot.jobs) {
      Job& J = S->jobs[id];
      if (J.msa_inflight++ == 0) {
        J.msa_idle += t - J.t_msa_mark;
        J.t_msa_mark = t;
      }
      J.msa_batches += 1;
    }
  }
  return MST_OK;
}

void release_group(mst_server* S, Job& J, Group& G) {
  if (G.leaf || G.off < 0) return;
  if (G.in_arena)
    S->arena.release(G.off, G.cap);
  else
    J.msa_pool.release(G.off - J.o_msa, G.cap);
  G.off = -1;
  G.cap = 0;
}

// Retires the finished merge batches (wait: block on the first busy one).
int retire_merges(mst_server* S, bool wait, int* ndone) {
  mst_ctx* ctx = S->ctx;
  for (auto& slot : S->ms) {
    if (!slot.busy) continue;
    hipError_t e = wait ? hipEventSynchronize(slot.ev) : hipEventQuery(slot.ev);
    if (e == hipErrorNotReady) continue;
    if (e != hipSuccess) return ctx->fail(MST_E_HIP, "merge batch", hipGetErrorString(e));
    wait = false;
    const double t = now_s();
    const int32_t* lens = slot.eng.lens();
    for (size_t b = 0; b < slot.merges.size(); ++b) {
      Job& J = S->jobs[slot.merges[b].first];
      const int k = slot.merges[b].second;
      const mst_join& jn = J.tree[k];
      Group& Z = J.groups[jn.node];
      Z.len = lens[b];
      const int64_t keep = std::max<int64_t>((int64_t)Z.k * Z.len, 1);
      if (Z.in_arena)
        S->arena.shrink(Z.off, Z.cap, keep);
      else
        J.msa_pool.shrink(Z.off - J.o_msa, Z.cap, keep);
      Z.cap = keep;
      Z.ready = true;
      J.merge_state[k] = 2;
      J.merges_done += 1;
      release_group(S, J, J.groups[jn.left]);
      release_group(S, J, J.groups[jn.right]);
    }
    for (int64_t id : slot.leaf_jobs) {
      Job& J = S->jobs[id];
      for (int i = 0; i < J.n; ++i) J.groups[i].ready = true;
    }
    for (int64_t id : slot.jobs) {
      Job& J = S->jobs[id];
      J.refs -= 1;
      if (--J.msa_inflight == 0) {
        J.msa_busy += t - J.t_msa_mark;
        J.t_msa_mark = t;
      }
      if (!J.cancel && J.state == MST_JOB_RUNNING && J.merges_done == J.n - 1) {
        // done once every batch that wrote its groups has retired
        bool pending = false;
        for (const auto& o : S->ms)
          if (&o != &slot && o.busy && std::find(o.jobs.begin(), o.jobs.end(), id) != o.jobs.end())
            pending = true;
        if (!pending) job_done(S, J, t, ndone);
      }
    }
    slot.merges.clear();
    slot.jobs.clear();
    slot.leaf_jobs.clear();
    slot.busy = false;
  }
  prune_running(S);
  return MST_OK;
}

}  // namespace

extern "C" int mst_server_create(mst_ctx* ctx, int64_t budget_bytes, mst_server** out) {
  if (!ctx || !out || budget_bytes < (1 << 20)) return MST_E_ARG;
  if (!pk16_usable(ctx)) return ctx->fail(MST_E_ARG, "mst_server_create", "the served form needs a symmetric matrix with S+go+ge in 0..255");
  CK(hipSetDevice(ctx->device));
  mst_server* S = new mst_server();
  S->ctx = ctx;
  S->budget = budget_bytes;
  const int64_t max_pos = 1 << 21, max_slots = 1 << 14;
  auto fail = [&](const char* what, hipError_t e) {
    mst_server_destroy(S);
    return ctx->fail(MST_E_NOMEM, what, hipGetErrorString(e));
  };
  hipError_t e;
  if ((e = hipMalloc(&S->d_arena, budget_bytes)) != hipSuccess) return fail("arena", e);
  if ((e = hipMalloc(&S->d_soff, sizeof(int64_t) * max_pos)) != hipSuccess) return fail("soff", e);
  if ((e = hipMalloc(&S->d_order, sizeof(int32_t) * max_pos)) != hipSuccess) return fail("order", e);
  if ((e = hipMalloc(&S->d_job_out, sizeof(int64_t) * max_slots)) != hipSuccess) return fail("jobs", e);
  if ((e = hipMalloc(&S->d_job_ld, sizeof(int) * max_slots)) != hipSuccess) return fail("jobs", e);
  S->grid8 = ctx->num_cus * score8_groups_per_cu();
  S->grid6 = ctx->num_cus * score6_groups_per_cu();
  S->grid7 = ctx->num_cus * score7_groups_per_cu();
  // a row buffer per workgroup (two per workgroup of the two-pipeline form)
  // and per multi-sweep class
  S->carry_rows = std::max<int64_t>(S->grid6, 2 * (int64_t)S->grid7);
  if ((e = hipMalloc(&S->d_carry, sizeof(int2) * (size_t)S->carry_rows * 16384 * 2)) != hipSuccess)
    return fail("carry", e);
  S->grid32 = ctx->num_cus * 4;
  S->bufs_cap_cols = 1 << 15;  // int32 kernel: pairs up to 32768 residues
  if ((e = hipMalloc(&S->d_bufs, sizeof(int2) * 2 * (size_t)S->bufs_cap_cols * S->grid32)) != hipSuccess)
    return fail("bufs", e);
  for (auto& st : S->side)
    if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) return fail("stream", e);
  for (auto& q : S->nj) {
    if ((e = hipStreamCreateWithFlags(&q.st, hipStreamNonBlocking)) != hipSuccess) return fail("stream", e);
    if ((e = hipEventCreateWithFlags(&q.ev, hipEventDisableTiming)) != hipSuccess) return fail("event", e);
  }
  for (auto& q : S->ms) {
    if ((e = hipStreamCreateWithFlags(&q.st, hipStreamNonBlocking)) != hipSuccess) return fail("stream", e);
    if ((e = hipEventCreateWithFlags(&q.ev, hipEventDisableTiming)) != hipSuccess) return fail("event", e);
  }
  S->arena.init(budget_bytes, 256);
  S->pos.init(max_pos, 1);
  S->slots.init(max_slots, 1);
  *out = S;
  return MST_OK;
}

extern "C" int mst_server_destroy(mst_server* S) {
  if (!S) return MST_E_ARG;
  mst_ctx* ctx = S->ctx;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  for (hipStream_t st : S->side)
    if (st) hipStreamSynchronize(st);
  for (auto& q : S->nj) {
    if (q.st) hipStreamSynchronize(q.st);
    if (q.ev) hipEventDestroy(q.ev);
    if (q.d_args) hipFree(q.d_args);
    if (q.h_args) hipHostFree(q.h_args);
    if (q.h_nj) hipHostFree(q.h_nj);
  }
  for (auto& q : S->ms) {
    if (q.st) hipStreamSynchronize(q.st);
    if (q.ev) hipEventDestroy(q.ev);
  }
  for (RoundBuf& B : S->rb) {
    hipEvent_t evs[] = {B.ev_start, B.ev_end, B.ev_go, B.ev_side[0], B.ev_side[1], B.ev_side[2],
                        B.ev_side[3]};
    for (hipEvent_t e : evs)
      if (e) hipEventDestroy(e);
    void* ps[] = {B.d_items, B.d_item_job, B.d_fb, B.d_fb_job, B.d_fb_val, B.d_counters};
    for (void* p : ps)
      if (p) hipFree(p);
    void* hs[] = {B.h_items, B.h_item_job, B.h_fb, B.h_fb_job};
    for (void* p : hs)
      if (p) hipHostFree(p);
  }
  void* ps[] = {S->d_arena, S->d_soff, S->d_order, S->d_job_out, S->d_job_ld, S->d_carry, S->d_bufs,
                S->d_carry8};
  for (void* p : ps)
    if (p) hipFree(p);
  for (hipStream_t st : S->side)
    if (st) hipStreamDestroy(st);
  for (auto& q : S->nj)
    if (q.st) hipStreamDestroy(q.st);
  for (auto& q : S->ms)
    if (q.st) hipStreamDestroy(q.st);
  delete S;
  return MST_OK;
}

extern "C" int mst_server_submit(mst_server* S, const char* ascii, const int64_t* offsets, int n,
                                 int64_t* job) {
  return mst_server_submit_job(S, ascii, offsets, n, MST_SERVE_SCORES, job);
}

extern "C" int mst_server_submit_job(mst_server* S, const char* ascii, const int64_t* offsets, int n,
                                     int what, int64_t* job) {
  if (!S || !job || n < 1 || !offsets || (!ascii && offsets[n] > offsets[0])) return MST_E_ARG;
  mst_ctx* ctx = S->ctx;
  if (what != MST_SERVE_SCORES && what != MST_SERVE_TREE && what != MST_SERVE_MSA)
    return ctx->fail(MST_E_ARG, "mst_server_submit_job", "what: MST_SERVE_SCORES, _TREE or _MSA");
  if (n > 46340) return ctx->fail(MST_E_ARG, "mst_server_submit_job", "too many sequences");
  for (int i = 0; i < n; ++i)
    if (offsets[i + 1] < offsets[i]) return ctx->fail(MST_E_ARG, "mst_server_submit", "offsets decrease");
  Job J;
  J.id = S->next_id++;
  J.n = n;
  J.what = what;
  J.off.assign(offsets, offsets + n + 1);
  J.ascii.assign(ascii ? ascii + offsets[0] : nullptr, ascii ? ascii + offsets[n] : nullptr);
  J.t_submit = now_s();
  if (job_footprint(J) > S->budget || n + 1 > S->pos.cap) {
    J.state = MST_JOB_REJECTED;
    J.t_done = J.t_submit;
    J.ascii.clear();
  } else {
    S->waiting.push_back(J.id);
  }
  *job = J.id;
  S->jobs.emplace(J.id, std::move(J));
  return MST_OK;
}

extern "C" int mst_server_step(mst_server* S, int* ncompleted) {
  if (!S) return MST_E_ARG;
  mst_ctx* ctx = S->ctx;
  CK(hipSetDevice(ctx->device));
  int done = 0, rc;
  // retire every finished round and batch
  for (;;) {
    const size_t before = S->inflight.size();
    if ((rc = retire(S, false, &done))) return rc;
    if (S->inflight.size() == before) break;
  }
  if ((rc = retire_nj(S, false, &done))) return rc;
  if ((rc = retire_merges(S, false, &done))) return rc;
  // admit in arrival order while the head fits
  while (!S->waiting.empty()) {
    Job& J = S->jobs[S->waiting.front()];
    if (J.cancel) {
      J.state = MST_JOB_CANCELLED;
      S->waiting.pop_front();
      continue;
    }
    rc = admit(S, J);
    if (rc == 1) break;
    if (rc) return rc;
    S->running.push_back(J.id);
    S->waiting.pop_front();
  }
  // keep up to three rounds queued on the device
  while ((int)S->inflight.size() < S->max_rounds) {
    int launched = 0;
    if ((rc = launch_round(S, &launched))) return rc;
    if (!launched) break;
  }
  if ((rc = launch_nj_jobs(S))) return rc;
  if ((rc = launch_merges(S))) return rc;
  if (ncompleted) *ncompleted = done;
  return MST_OK;
}

extern "C" int mst_server_drain(mst_server* S) {
  if (!S) return MST_E_ARG;
  for (;;) {
    int rc = mst_server_step(S, nullptr);
    if (rc) return rc;
    int d = 0;
    if (!S->inflight.empty()) {
      if ((rc = retire(S, true, &d))) return rc;
    } else if (nj_busy(S)) {
      if ((rc = retire_nj(S, true, &d))) return rc;
    } else if (msa_busy(S)) {
      if ((rc = retire_merges(S, true, &d))) return rc;
    } else if (S->waiting.empty() && S->running.empty()) {
      return MST_OK;
    } else {
      // nothing on the device, yet jobs wait: an accepted job's footprint is
      // within the budget, so the arena is held by finished jobs whose
      // results were not collected
      return S->ctx->fail(MST_E_STATE, "mst_server_drain",
                          "waiting jobs do not fit: collect or release finished jobs");
    }
  }
}

extern "C" int mst_server_job(mst_server* S, int64_t job, int* state, int* n, double* t_submit,
                              double* t_admit, double* t_done) {
  if (!S) return MST_E_ARG;
  auto it = S->jobs.find(job);
  if (it == S->jobs.end()) return S->ctx->fail(MST_E_ARG, "mst_server_job", "unknown job");
  const Job& J = it->second;
  if (state) *state = J.state;
  if (n) *n = J.n;
  if (t_submit) *t_submit = J.t_submit;
  if (t_admit) *t_admit = J.t_admit;
  if (t_done) *t_done = J.t_done;
  return MST_OK;
}

extern "C" int mst_server_job_times(mst_server* S, int64_t job, double* times) {
  if (!S || !times) return MST_E_ARG;
  auto it = S->jobs.find(job);
  if (it == S->jobs.end()) return S->ctx->fail(MST_E_ARG, "mst_server_job_times", "unknown job");
  const Job& J = it->second;
  times[0] = J.t_submit;
  times[1] = J.t_admit;
  times[2] = J.t_scores;
  times[3] = J.t_tree;
  times[4] = J.t_done;
  return MST_OK;
}

extern "C" int mst_server_job_stats(mst_server* S, int64_t job, double* stats) {
  if (!S || !stats) return MST_E_ARG;
  auto it = S->jobs.find(job);
  if (it == S->jobs.end()) return S->ctx->fail(MST_E_ARG, "mst_server_job_stats", "unknown job");
  const Job& J = it->second;
  stats[0] = (double)J.msa_batches;
  stats[1] = J.msa_busy;
  stats[2] = J.msa_idle;
  stats[3] = (double)J.score_rounds;
  stats[4] = J.t_first_round > 0 ? J.t_first_round - J.t_submit : 0;
  stats[5] = J.t_admit > 0 ? J.t_admit - J.t_submit : 0;
  return MST_OK;
}

extern "C" int mst_server_result(mst_server* S, int64_t job, int32_t* scores_nxn) {
  if (!S) return MST_E_ARG;
  mst_ctx* ctx = S->ctx;
  auto it = S->jobs.find(job);
  if (it == S->jobs.end()) return ctx->fail(MST_E_ARG, "mst_server_result", "unknown job");
  Job& J = it->second;
  if (J.state != MST_JOB_DONE) return ctx->fail(MST_E_STATE, "mst_server_result", "job not done");
  CK(hipSetDevice(ctx->device));
  if (scores_nxn)
    CK(hipMemcpy(scores_nxn, S->d_arena + J.arena_off, sizeof(int32_t) * (size_t)J.n * J.n,
                 hipMemcpyDeviceToHost));
  release_job(S, J);
  S->jobs.erase(it);
  return MST_OK;
}

extern "C" int mst_server_tree(mst_server* S, int64_t job, mst_join* joins) {
  if (!S || !joins) return MST_E_ARG;
  mst_ctx* ctx = S->ctx;
  auto it = S->jobs.find(job);
  if (it == S->jobs.end()) return ctx->fail(MST_E_ARG, "mst_server_tree", "unknown job");
  const Job& J = it->second;
  if (J.state != MST_JOB_DONE) return ctx->fail(MST_E_STATE, "mst_server_tree", "job not done");
  if ((J.what & MST_SERVE_TREE) != MST_SERVE_TREE)
    return ctx->fail(MST_E_STATE, "mst_server_tree", "the job was submitted without a tree");
  std::copy(J.tree.begin(), J.tree.end(), joins);
  return MST_OK;
}

extern "C" int mst_server_msa(mst_server* S, int64_t job, char* rows, int64_t stride, int* msa_len) {
  static const char kLetters[] = "ARNDCQEGHILKMFPSTWYVBZX*-";
  if (!S || !msa_len) return MST_E_ARG;
  mst_ctx* ctx = S->ctx;
  auto it = S->jobs.find(job);
  if (it == S->jobs.end()) return ctx->fail(MST_E_ARG, "mst_server_msa", "unknown job");
  const Job& J = it->second;
  if (J.state != MST_JOB_DONE) return ctx->fail(MST_E_STATE, "mst_server_msa", "job not done");
  if ((J.what & MST_SERVE_MSA) != MST_SERVE_MSA)
    return ctx->fail(MST_E_STATE, "mst_server_msa", "the job was submitted without an alignment");
  const Group& R = J.groups[2 * (size_t)J.n - 2];
  *msa_len = R.len;
  if (!rows) return MST_OK;
  if (stride < (int64_t)R.len + 1) return ctx->fail(MST_E_ARG, "mst_server_msa", "stride too small");
  // rows of the root: the members of its left subtree, then of its right
  std::vector<int> members;
  std::vector<int> stack{2 * J.n - 2};
  while (!stack.empty()) {
    const int v = stack.back();
    stack.pop_back();
    if (v < J.n) {
      members.push_back(v);
    } else {
      stack.push_back(J.tree[v - J.n].right);
      stack.push_back(J.tree[v - J.n].left);
    }
  }
  CK(hipSetDevice(ctx->device));
  std::vector<uint8_t> h((size_t)R.k * R.len);
  if (!h.empty()) CK(hipMemcpy(h.data(), S->d_arena + R.off, h.size(), hipMemcpyDeviceToHost));
  for (int r = 0; r < R.k; ++r) {
    char* o = rows + (int64_t)members[r] * stride;
    const uint8_t* src = &h[(size_t)r * R.len];
    for (int t = 0; t < R.len; ++t) o[t] = kLetters[src[t] < 25 ? src[t] : 24];
    o[R.len] = 0;
  }
  return MST_OK;
}

extern "C" int mst_server_cancel(mst_server* S, int64_t job) {
  if (!S) return MST_E_ARG;
  auto it = S->jobs.find(job);
  if (it == S->jobs.end()) return S->ctx->fail(MST_E_ARG, "mst_server_cancel", "unknown job");
  Job& J = it->second;
  if (J.state == MST_JOB_WAITING || J.state == MST_JOB_RUNNING) J.cancel = true;
  if (J.state == MST_JOB_WAITING) {
    J.state = MST_JOB_CANCELLED;
    S->waiting.erase(std::find(S->waiting.begin(), S->waiting.end(), job));
  } else if (J.state == MST_JOB_RUNNING && J.refs == 0) {
    release_job(S, J);
    J.state = MST_JOB_CANCELLED;
    S->running.erase(std::find(S->running.begin(), S->running.end(), job));
  }
  return MST_OK;
}

extern "C" int mst_server_stats(mst_server* S, int* waiting, int* running, int* inflight,
                                int64_t* arena_used, double* rate) {
  if (!S) return MST_E_ARG;
  if (waiting) *waiting = (int)S->waiting.size();
  if (running) *running = (int)S->running.size();
  if (inflight) *inflight = (int)S->inflight.size();
  if (arena_used) *arena_used = S->arena.used;
  if (rate) *rate = S->rate;
  return MST_OK;
}

extern "C" int mst_server_set(mst_server* S, const char* key, double value) {
  if (!S || !key) return MST_E_ARG;
  if (!strcmp(key, "round_cells") && value >= 1e6 && value <= 1e14) {
    S->round_cells = value;
    return MST_OK;
  }
  if (!strcmp(key, "k8") && (value == 0 || value == 1)) {
    S->k8 = (int)value;
    return MST_OK;
  }
  if (!strcmp(key, "s6p2") && (value == 0 || value == 1)) {
    S->s6p2 = (int)value;
    return MST_OK;
  }
  if (!strcmp(key, "chunk6") && value >= 0 && value <= kStreamSeqs) {
    S->chunk6 = (int)value;
    return MST_OK;
  }
  if (!strcmp(key, "chunk") && value >= 1 && value <= kStreamSeqs) {
    S->chunk = (int)value;
    return MST_OK;
  }
  if (!strcmp(key, "slack") && value > 0 && value <= 100) {
    S->slack = value;
    return MST_OK;
  }
  if (!strcmp(key, "rate") && value > 0) {
    S->rate = value;
    return MST_OK;
  }
  if (!strcmp(key, "round_cells_shared") && value >= 1e6 && value <= 1e14) {
    S->round_cells_shared = value;
    return MST_OK;
  }
  if (!strcmp(key, "score_share") && value > 0 && value <= 1) {
    S->score_share = value;
    return MST_OK;
  }
  if (!strcmp(key, "deadline_add") && value >= 0 && value <= 100) {
    S->deadline_add = value;
    return MST_OK;
  }
  if (!strcmp(key, "rounds") && value >= 1 && value <= mst_server::kBufs) {
    S->max_rounds = (int)value;
    return MST_OK;
  }
  if (!strcmp(key, "est_floor") && value >= 0 && value <= 100) {
    S->est_floor = value;
    return MST_OK;
  }
  if (!strcmp(key, "shared_backlog") && value >= 0) {
    S->shared_backlog = value;
    return MST_OK;
  }
  if (!strcmp(key, "small_cells") && value >= 0) {
    S->small_cells = value;
    return MST_OK;
  }
  if (!strcmp(key, "persistent") && (value == 0 || value == 1)) {
    S->persistent = (int)value;
    return MST_OK;
  }
  if (!strcmp(key, "nj_ratio") && value >= 0 && value <= 1) {
    S->nj_ratio = (float)value;
    return MST_OK;
  }
  if (!strcmp(key, "merge_bytes") && value >= (1 << 20) && value <= 1e12) {
    S->merge_bytes = value;
    return MST_OK;
  }
  return S->ctx->fail(MST_E_ARG, "mst_server_set", "unknown key or value out of range");
}

extern "C" int mst_server_completed(mst_server* S, int64_t* ids, int cap, int* count) {
  if (!S || !count || (cap > 0 && !ids)) return MST_E_ARG;
  const int n = (int)std::min<size_t>(S->completed.size(), (size_t)std::max(cap, 0));
  for (int i = 0; i < n; ++i) ids[i] = S->completed[i];
  S->completed.erase(S->completed.begin(), S->completed.begin() + n);
  *count = n;
  return MST_OK;
}

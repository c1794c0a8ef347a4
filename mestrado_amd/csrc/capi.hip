// C ABI of libmestrado (include/mestrado.h): context, resident sequence sets
// and the all-pairs score path.
#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/mestrado.h"
#include "ctx.h"
#include "encode.h"
#include "nw_score.h"

using namespace mst;

int noalloc_init_device(int device, void** ptr);  // capi_noalloc.hip
bool flag16_usable(const mst_ctx* ctx);            // capi_align.hip

namespace {

__global__ void fill_i32(int32_t* p, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// BLOSUM62 in the letter order ARNDCQEGHILKMFPSTWYVBZX* (the table of
// definitions.cuh:96-121, which is the published BLOSUM62 matrix).
const int8_t kBlosum62[576] = {
    4, -1, -2, -2, 0, -1, -1, 0, -2, -1, -1, -1, -1, -2, -1, 1, 0, -3, -2, 0, -2, -1, 0, -4,
    -1, 5, 0, -2, -3, 1, 0, -2, 0, -3, -2, 2, -1, -3, -2, -1, -1, -3, -2, -3, -1, 0, -1, -4,
    -2, 0, 6, 1, -3, 0, 0, 0, 1, -3, -3, 0, -2, -3, -2, 1, 0, -4, -2, -3, 3, 0, -1, -4,
    -2, -2, 1, 6, -3, 0, 2, -1, -1, -3, -4, -1, -3, -3, -1, 0, -1, -4, -3, -3, 4, 1, -1, -4,
    0, -3, -3, -3, 9, -3, -4, -3, -3, -1, -1, -3, -1, -2, -3, -1, -1, -2, -2, -1, -3, -3, -2, -4,
    -1, 1, 0, 0, -3, 5, 2, -2, 0, -3, -2, 1, 0, -3, -1, 0, -1, -2, -1, -2, 0, 3, -1, -4,
    -1, 0, 0, 2, -4, 2, 5, -2, 0, -3, -3, 1, -2, -3, -1, 0, -1, -3, -2, -2, 1, 4, -1, -4,
    0, -2, 0, -1, -3, -2, -2, 6, -2, -4, -4, -2, -3, -3, -2, 0, -2, -2, -3, -3, -1, -2, -1, -4,
    -2, 0, 1, -1, -3, 0, 0, -2, 8, -3, -3, -1, -2, -1, -2, -1, -2, -2, 2, -3, 0, 0, -1, -4,
    -1, -3, -3, -3, -1, -3, -3, -4, -3, 4, 2, -3, 1, 0, -3, -2, -1, -3, -1, 3, -3, -3, -1, -4,
    -1, -2, -3, -4, -1, -2, -3, -4, -3, 2, 4, -2, 2, 0, -3, -2, -1, -2, -1, 1, -4, -3, -1, -4,
    -1, 2, 0, -1, -3, 1, 1, -2, -1, -3, -2, 5, -1, -3, -1, 0, -1, -3, -2, -2, 0, 1, -1, -4,
    -1, -1, -2, -3, -1, 0, -2, -3, -2, 1, 2, -1, 5, 0, -2, -1, -1, -1, -1, 1, -3, -1, -1, -4,
    -2, -3, -3, -3, -2, -3, -3, -3, -1, 0, 0, -3, 0, 6, -4, -2, -2, 1, 3, -1, -3, -3, -1, -4,
    -1, -2, -2, -1, -3, -1, -1, -2, -2, -3, -3, -1, -2, -4, 7, -1, -1, -4, -3, -2, -2, -1, -2, -4,
    1, -1, 1, 0, -1, 0, 0, 0, -1, -2, -2, 0, -1, -2, -1, 4, 1, -3, -2, -2, 0, 0, 0, -4,
    0, -1, 0, -1, -1, -1, -1, -2, -2, -1, -1, -1, -1, -2, -1, 1, 5, -2, -2, 0, -1, -1, 0, -4,
    -3, -3, -4, -4, -2, -2, -3, -2, -2, -3, -2, -3, -1, 1, -4, -3, -2, 11, 2, -3, -4, -3, -2, -4,
    -2, -2, -2, -3, -2, -1, -2, -3, 2, -1, -1, -2, -1, 3, -3, -2, -2, 2, 7, -1, -3, -2, -1, -4,
    0, -3, -3, -3, -1, -2, -2, -3, -3, 3, 1, -2, 1, -1, -2, -2, 0, -3, -1, 4, -3, -2, -1, -4,
    -2, -1, 3, 4, -3, 0, 1, -1, 0, -3, -4, 0, -3, -3, -2, 0, -1, -4, -3, -3, 4, 1, -1, -4,
    -1, 0, 0, 1, -3, 3, 4, -2, 0, -3, -3, 1, -1, -3, -1, 0, -1, -3, -2, -2, 1, 4, -1, -4,
    0, -1, -1, -1, -2, -1, -1, -1, -1, -1, -1, -1, -1, -1, -2, 0, 0, -2, -1, -1, -1, -1, -1, -4,
    -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, 1};

}  // namespace

#define CK(call)                                                              \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) return ctx->fail(MST_E_HIP, #call, hipGetErrorString(e_)); \
  } while (0)

int mst_ctx::fail(int code, const char* what, const char* detail) {
  err = std::string(what) + (detail ? std::string(": ") + detail : std::string());
  return code;
}

int mst_ctx::ensure(void** p, size_t* cap, size_t bytes) {
  if (*cap >= bytes) return MST_OK;
  if (*p) hipFree(*p);
  *p = nullptr;
  *cap = 0;
  size_t want = std::max(bytes, (size_t)256);
  hipError_t e = hipMalloc(p, want);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(MST_E_NOMEM, "hipMalloc", hipGetErrorString(e));
  }
  *cap = want;
  return MST_OK;
}

extern "C" int mst_create(int gap_open, int gap_ext, const int8_t* subst, int device,
                          mst_ctx** out) {
  if (!out) return MST_E_ARG;
  *out = nullptr;
  // the kernels pack S + go + ge into signed bytes (S in [-128,127] shifted)
  if (gap_open < 0 || gap_ext < 0 || gap_open < gap_ext || gap_open + gap_ext > 116) return MST_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return MST_E_HIP;
  if (device < 0 || device >= ndev) return MST_E_ARG;
  std::unique_ptr<mst_ctx> c(new mst_ctx());
  mst_ctx* ctx = c.get();
  ctx->device = device;
  ctx->go = gap_open;
  ctx->ge = gap_ext;
  memcpy(ctx->S, subst ? subst : kBlosum62, 576);
  for (int k = 0; k < 576; ++k)
    if (ctx->S[k] + gap_open + gap_ext > 127 || ctx->S[k] + gap_open + gap_ext < -128) return MST_E_ARG;
  ctx->symmetric = true;
  ctx->smax = 0;
  for (int k = 0; k < 576; ++k) ctx->smax = std::max(ctx->smax, (int)ctx->S[k]);
  ctx->smin = 127;
  for (int k = 0; k < 576; ++k) ctx->smin = std::min(ctx->smin, (int)ctx->S[k]);
  for (int a = 0; a < 24; ++a)
    for (int b = 0; b < 24; ++b)
      if (ctx->S[a * 24 + b] != ctx->S[b * 24 + a]) ctx->symmetric = false;
  CK(hipSetDevice(device));
  hipDeviceProp_t prop;
  CK(hipGetDeviceProperties(&prop, device));
  ctx->num_cus = prop.multiProcessorCount;
  strncpy(ctx->arch, prop.gcnArchName, sizeof ctx->arch - 1);
  ctx->hbm_bytes = (int64_t)prop.totalGlobalMem;
  CK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  CK(hipEventCreate(&ctx->ev0));
  CK(hipEventCreate(&ctx->ev1));
  CK(hipMalloc(&ctx->d_S, 576));
  CK(hipMemcpy(ctx->d_S, ctx->S, 576, hipMemcpyHostToDevice));
  CK(init_encode_table());
  ctx->flag16 = flag16_usable(ctx);
  {
    const int rc = noalloc_init_device(device, &ctx->d_scratch);
    if (rc) return ctx->fail(rc, "hipGetSymbolAddress", "no-allocation scratch array");
  }
  *out = c.release();
  return MST_OK;
}

mst_ctx::~mst_ctx() {
  for (auto& kv : bplans) free_batched(kv.second);
  for (auto& kv : fbplans)
    if (kv.second.d_pairs) hipFree(kv.second.d_pairs);
  for (auto& kv : sets) {
    hipFree(kv.second.d_codes);
    hipFree(kv.second.d_off);
  }
  void* ps[] = {d_S,    d_pairs,    d_bufs,  d_mat, d_pscore, d_nj,    d_nj_T,
                d_nj_act, d_nj_joins, d_flags, d_ops, d_jobs,   d_jscore, d_acounter,
                d_tiles,  d_trows,    d_tstate, d_nj_parts, d_f16, d_tab16};
  for (void* p : ps)
    if (p) hipFree(p);
  for (hipEvent_t e : launch_events) hipEventDestroy(e);
  if (ev0) hipEventDestroy(ev0);
  if (ev1) hipEventDestroy(ev1);
  for (hipStream_t st : bstream)
    if (st) hipStreamDestroy(st);
  if (astream) hipStreamDestroy(astream);
  for (hipEvent_t e : bevent)
    if (e) hipEventDestroy(e);
  if (stream) hipStreamDestroy(stream);
}

extern "C" int mst_destroy(mst_ctx* ctx) {
  if (!ctx) return MST_E_ARG;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  delete ctx;
  return MST_OK;
}

extern "C" const char* mst_last_error(const mst_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

extern "C" int mst_device_info(mst_ctx* ctx, int* num_cus, int64_t* hbm) {
  if (!ctx) return MST_E_ARG;
  if (num_cus) *num_cus = ctx->num_cus;
  if (hbm) *hbm = ctx->hbm_bytes;
  return MST_OK;
}

extern "C" int mst_device_arch(mst_ctx* ctx, char* buf, int cap) {
  if (!ctx || !buf || cap < 1) return MST_E_ARG;
  strncpy(buf, ctx->arch, cap - 1);
  buf[cap - 1] = 0;
  return MST_OK;
}

extern "C" int mst_seqset_upload(mst_ctx* ctx, const char* ascii, const int64_t* off,
                                 int n, int* set_id) {
  if (!ctx || !off || n < 0 || !set_id) return MST_E_ARG;
  if (off[0] != 0) return ctx->fail(MST_E_ARG, "offsets[0] must be 0");
  for (int i = 0; i < n; ++i)
    if (off[i + 1] < off[i]) return ctx->fail(MST_E_ARG, "offsets must be non-decreasing");
  const int64_t total = off[n];
  if (total > 0 && !ascii) return MST_E_ARG;
  CK(hipSetDevice(ctx->device));
  SeqSet s;
  s.n = n;
  s.total = total;
  s.h_off.assign(off, off + n + 1);
  s.h_canon = canonical_letters(ascii ? ascii : "", total);
  s.maxlen = 0;
  for (int i = 0; i < n; ++i) s.maxlen = std::max<int64_t>(s.maxlen, off[i + 1] - off[i]);
  if (s.maxlen > (int64_t)1 << 30) return ctx->fail(MST_E_ARG, "sequence too long");
  CK(hipMalloc(&s.d_codes, std::max<int64_t>(total, 16)));
  CK(hipMalloc(&s.d_off, sizeof(int64_t) * (n + 1)));
  char* d_ascii = nullptr;
  CK(hipMalloc(&d_ascii, std::max<int64_t>(total, 16)));
  CK(hipMemcpyAsync(d_ascii, ascii, total, hipMemcpyHostToDevice, ctx->stream));
  CK(hipMemcpyAsync(s.d_off, off, sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, ctx->stream));
  launch_encode(d_ascii, s.d_off, n, s.d_codes, ctx->stream);
  CK(hipGetLastError());
  CK(hipStreamSynchronize(ctx->stream));
  CK(hipFree(d_ascii));
  int id = ctx->next_set++;
  ctx->sets[id] = std::move(s);
  *set_id = id;
  return MST_OK;
}

extern "C" int mst_seqset_free(mst_ctx* ctx, int id) {
  if (!ctx) return MST_E_ARG;
  auto it = ctx->sets.find(id);
  if (it == ctx->sets.end()) return ctx->fail(MST_E_STATE, "unknown sequence set");
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  hipFree(it->second.d_codes);
  hipFree(it->second.d_off);
  ctx->sets.erase(it);
  ctx->plans.erase(id);
  ctx->prepared.erase(std::make_pair(id, 0));
  ctx->prepared.erase(std::make_pair(id, 1));
  auto bp = ctx->bplans.find(id);
  if (bp != ctx->bplans.end()) {
    free_batched(bp->second);
    ctx->bplans.erase(bp);
  }
  auto fp = ctx->fbplans.find(id);
  if (fp != ctx->fbplans.end()) {
    if (fp->second.d_pairs) hipFree(fp->second.d_pairs);
    ctx->fbplans.erase(fp);
  }
  return MST_OK;
}

extern "C" int mst_seqset_codes(mst_ctx* ctx, int id, int8_t* out, int64_t cap) {
  if (!ctx || !out) return MST_E_ARG;
  auto it = ctx->sets.find(id);
  if (it == ctx->sets.end()) return ctx->fail(MST_E_STATE, "unknown sequence set");
  if (cap < it->second.total) return ctx->fail(MST_E_ARG, "buffer too small");
  CK(hipSetDevice(ctx->device));
  CK(hipMemcpy(out, it->second.d_codes, it->second.total, hipMemcpyDeviceToHost));
  return MST_OK;
}

// ---------------------------------------------------------------------------
// Pair planning: pairs are bucketed by the row count of the DP (the shorter
// sequence when S is symmetric), each bucket gets its rows-per-lane K.

int k_for_rows(int m) {
  if (m <= 256) return 4;
  if (m <= 512) return 8;
  if (m <= 768) return 12;
  return 16;
}

int build_plan(mst_ctx* ctx, const SeqSet& s, const int32_t* pi, const int32_t* pj,
               int64_t np, PairPlan& plan) {
  const auto& off = s.h_off;
  auto len = [&](int i) { return (int)(off[i + 1] - off[i]); };
  std::vector<int64_t> order(np);
  std::vector<int> rows(np);
  std::vector<int64_t> cost(np);
  for (int64_t k = 0; k < np; ++k) {
    int a = pi[k], b = pj[k];
    if (a < 0 || b < 0 || a >= s.n || b >= s.n) return ctx->fail(MST_E_ARG, "pair index out of range");
    int la = len(a), lb = len(b);
    rows[k] = ctx->symmetric ? std::min(la, lb) : lb;
    cost[k] = (int64_t)la * lb;
    order[k] = k;
  }
  // Bucket by K, longest pairs first inside a bucket.
  std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
    int kx = k_for_rows(rows[x]), ky = k_for_rows(rows[y]);
    if (kx != ky) return kx < ky;
    return cost[x] > cost[y];
  });
  plan.pairs.resize(np);
  plan.index.resize(np);
  plan.buckets.clear();
  for (int64_t t = 0; t < np; ++t) {
    int64_t k = order[t];
    plan.pairs[t] = make_int2(pi[k], pj[k]);
    plan.index[t] = k;
    int K = k_for_rows(rows[k]);
    if (plan.buckets.empty() || plan.buckets.back().K != K)
      plan.buckets.push_back(Bucket{K, t, t + 1, 0});
    else
      plan.buckets.back().end = t + 1;
    plan.buckets.back().cells += cost[k];
  }
  plan.cells = 0;
  for (int64_t k = 0; k < np; ++k) plan.cells += cost[k];
  return MST_OK;
}

int upload_plan(mst_ctx* ctx, PairPlan& plan) {
  size_t bytes = sizeof(int2) * std::max<size_t>(plan.pairs.size(), 1);
  int rc = ctx->ensure((void**)&ctx->d_pairs, &ctx->pairs_cap, bytes);
  if (rc) return rc;
  CK(hipMemcpyAsync(ctx->d_pairs, plan.pairs.data(), sizeof(int2) * plan.pairs.size(),
                    hipMemcpyHostToDevice, ctx->stream));
  plan.uploaded = true;
  return MST_OK;
}

// Grid of the score launches: enough waves to fill the chip several times over.
int score_grid(mst_ctx* ctx, int64_t npairs) {
  int64_t g = (int64_t)ctx->num_cus * 16;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, npairs));
}

int run_plan_on(mst_ctx* ctx, const SeqSet& s, const PairPlan& plan, const int2* d_pairs, int form,
                int32_t* d_out_pair, int32_t* d_out_mat, bool timed, int event_base, hipStream_t st) {
  if (!st) st = ctx->stream;
  int grid_max = score_grid(ctx, (int64_t)plan.pairs.size());
  int64_t bufcap = std::max<int64_t>(s.maxlen, 1);
  int rc = ctx->ensure((void**)&ctx->d_bufs, &ctx->bufs_cap,
                       sizeof(int2) * 2 * bufcap * (size_t)grid_max);
  if (rc) return rc;
  for (size_t bi = 0; bi < plan.buckets.size(); ++bi) {
    const Bucket& b = plan.buckets[bi];
    ScoreArgs a;
    a.seqs = s.d_codes;
    a.off = s.d_off;
    a.pairs = d_pairs;
    a.first = b.begin;
    a.last = b.end;
    a.out_pair = d_out_pair;
    a.out_mat = d_out_mat;
    a.nseq = s.n;
    a.bufs = ctx->d_bufs;
    a.bufcap = bufcap;
    a.subst = ctx->d_S;
    a.go = ctx->go;
    a.ge = ctx->ge;
    a.transpose_ok = ctx->symmetric ? 1 : 0;
    int grid = score_grid(ctx, b.end - b.begin);
    if (timed) CK(hipEventRecord(ctx->launch_events[event_base + 2 * bi], st));
    launch_score_pairs(b.K, a, grid, st);
    CK(hipGetLastError());
    if (timed) CK(hipEventRecord(ctx->launch_events[event_base + 2 * bi + 1], st));
    if (form == MST_FORM_PLAIN) CK(hipStreamSynchronize(st));
  }
  return MST_OK;
}

int run_plan(mst_ctx* ctx, const SeqSet& s, const PairPlan& plan, int form, int32_t* d_out_pair,
             int32_t* d_out_mat, bool timed = false) {
  return run_plan_on(ctx, s, plan, ctx->d_pairs, form, d_out_pair, d_out_mat, timed, 0, ctx->stream);
}

bool pk16_usable(const mst_ctx* ctx);
int run_batched(mst_ctx* ctx, int id, int32_t* d_out, bool timed);
void free_batched(BatchedPlan& P);

int all_pairs_plan(mst_ctx* ctx, int id, PairPlan** out) {
  auto it = ctx->sets.find(id);
  if (it == ctx->sets.end()) return ctx->fail(MST_E_STATE, "unknown sequence set");
  PairPlan& plan = ctx->plans[id];
  if (!plan.built) {
    const SeqSet& s = it->second;
    std::vector<int32_t> pi, pj;
    pi.reserve((size_t)s.n * (s.n + 1) / 2);
    pj.reserve(pi.capacity());
    for (int i = 0; i < s.n; ++i)
      for (int j = i; j < s.n; ++j) {
        pi.push_back(i);
        pj.push_back(j);
      }
    int rc = build_plan(ctx, s, pi.data(), pj.data(), (int64_t)pi.size(), plan);
    if (rc) return rc;
    plan.built = true;
  }
  *out = &plan;
  return MST_OK;
}

// All pairs of the set into ctx->d_mat, in the requested form.
static int all_pairs_device(mst_ctx* ctx, int id, int form, bool timed) {
  auto it = ctx->sets.find(id);
  if (it == ctx->sets.end()) return ctx->fail(MST_E_STATE, "unknown sequence set");
  const SeqSet& s = it->second;
  size_t nn = (size_t)s.n * s.n;
  int rc = ctx->ensure((void**)&ctx->d_mat, &ctx->mat_cap, sizeof(int32_t) * std::max<size_t>(nn, 1));
  if (rc) return rc;
  if (form == MST_FORM_BATCHED && pk16_usable(ctx)) return run_batched(ctx, id, ctx->d_mat, timed);
  PairPlan* plan = nullptr;
  rc = all_pairs_plan(ctx, id, &plan);
  if (rc) return rc;
  if (!plan->uploaded || ctx->resident_plan != id) {
    rc = upload_plan(ctx, *plan);
    if (rc) return rc;
    ctx->resident_plan = id;
  }
  if (timed) {
    while (ctx->launch_events.size() < 2 * plan->buckets.size()) {
      hipEvent_t e;
      CK(hipEventCreate(&e));
      ctx->launch_events.push_back(e);
    }
  }
  rc = run_plan(ctx, s, *plan, form, nullptr, ctx->d_mat, timed);
  if (rc) return rc;
  if (timed) {
    CK(hipStreamSynchronize(ctx->stream));
    ctx->launch_stats.clear();
    for (size_t b = 0; b < plan->buckets.size(); ++b) {
      float t = 0;
      CK(hipEventElapsedTime(&t, ctx->launch_events[2 * b], ctx->launch_events[2 * b + 1]));
      const Bucket& bk = plan->buckets[b];
      ctx->launch_stats.push_back(LaunchStat{bk.K, bk.end - bk.begin, bk.cells, t});
    }
  }
  return MST_OK;
}

extern "C" int mst_all_pairs(mst_ctx* ctx, int id, int form, int32_t* scores, float* dist) {
  if (!ctx || (form != MST_FORM_PLAIN && form != MST_FORM_BATCHED)) return MST_E_ARG;
  CK(hipSetDevice(ctx->device));
  {
    // every entry is written by exactly one pair; start from a sentinel so a
    // pair the plan missed cannot hide behind an earlier call's value
    auto it = ctx->sets.find(id);
    if (it == ctx->sets.end()) return ctx->fail(MST_E_STATE, "unknown sequence set");
    const int64_t nn = (int64_t)it->second.n * it->second.n;
    int rc0 = ctx->ensure((void**)&ctx->d_mat, &ctx->mat_cap, sizeof(int32_t) * std::max<int64_t>(nn, 1));
    if (rc0) return rc0;
    if (nn > 0)
      hipLaunchKernelGGL(fill_i32, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, ctx->stream,
                         ctx->d_mat, nn, (int32_t)INT32_MIN);
    CK(hipGetLastError());
  }
  int rc = all_pairs_device(ctx, id, form, false);
  if (rc) return rc;
  const SeqSet& s = ctx->sets[id];
  size_t nn = (size_t)s.n * s.n;
  ctx->mat_set = id;
  std::vector<int32_t> tmp;
  int32_t* host = scores;
  if (!host && dist) {
    tmp.resize(nn);
    host = tmp.data();
  }
  if (host) CK(hipMemcpyAsync(host, ctx->d_mat, sizeof(int32_t) * nn, hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  if (dist)
    for (size_t k = 0; k < nn; ++k) dist[k] = (float)host[k];
  return MST_OK;
}

extern "C" int mst_all_pairs_resident(mst_ctx* ctx, int id, int form, float* ms) {
  if (!ctx || (form != MST_FORM_PLAIN && form != MST_FORM_BATCHED)) return MST_E_ARG;
  CK(hipSetDevice(ctx->device));
  if (ctx->sets.find(id) == ctx->sets.end()) return ctx->fail(MST_E_STATE, "unknown sequence set");
  // the first call of a (set, form) builds its plan, device copies and graph
  // outside the timed region
  int rc;
  if (!ctx->prepared.count(std::make_pair(id, form))) {
    rc = all_pairs_device(ctx, id, form, false);
    if (rc) return rc;
    ctx->prepared.insert(std::make_pair(id, form));
  }
  CK(hipStreamSynchronize(ctx->stream));
  CK(hipEventRecord(ctx->ev0, ctx->stream));
  rc = all_pairs_device(ctx, id, form, false);
  if (rc) return rc;
  CK(hipEventRecord(ctx->ev1, ctx->stream));
  CK(hipEventSynchronize(ctx->ev1));
  ctx->mat_set = id;
  float t = 0;
  CK(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
  if (ms) *ms = t;
  return MST_OK;
}

extern "C" int mst_all_pairs_launches(mst_ctx* ctx, int id, int form) {
  if (!ctx || (form != MST_FORM_PLAIN && form != MST_FORM_BATCHED)) return MST_E_ARG;
  CK(hipSetDevice(ctx->device));
  int rc = all_pairs_device(ctx, id, form, true);
  if (rc) return rc;
  ctx->mat_set = id;
  return MST_OK;
}

extern "C" int mst_last_launch_stats(mst_ctx* ctx, int* rows_per_lane, int64_t* pairs,
                                     int64_t* cells, float* ms, int cap, int* count) {
  if (!ctx || !count) return MST_E_ARG;
  *count = (int)ctx->launch_stats.size();
  for (int i = 0; i < *count && i < cap; ++i) {
    const LaunchStat& L = ctx->launch_stats[i];
    if (rows_per_lane) rows_per_lane[i] = L.K;
    if (pairs) pairs[i] = L.pairs;
    if (cells) cells[i] = L.cells;
    if (ms) ms[i] = L.ms;
  }
  return MST_OK;
}

extern "C" int mst_score_pairs(mst_ctx* ctx, int id, const int32_t* pi, const int32_t* pj,
                               int64_t np, int form, int32_t* out) {
  if (!ctx || np < 0 || (np > 0 && (!pi || !pj || !out))) return MST_E_ARG;
  if (form != MST_FORM_PLAIN && form != MST_FORM_BATCHED) return MST_E_ARG;
  auto it = ctx->sets.find(id);
  if (it == ctx->sets.end()) return ctx->fail(MST_E_STATE, "unknown sequence set");
  if (np == 0) return MST_OK;
  CK(hipSetDevice(ctx->device));
  PairPlan plan;
  int rc = build_plan(ctx, it->second, pi, pj, np, plan);
  if (rc) return rc;
  rc = upload_plan(ctx, plan);
  if (rc) return rc;
  ctx->resident_plan = -1;
  rc = ctx->ensure((void**)&ctx->d_pscore, &ctx->pscore_cap, sizeof(int32_t) * np);
  if (rc) return rc;
  rc = run_plan(ctx, it->second, plan, form, ctx->d_pscore, nullptr);
  if (rc) return rc;
  std::vector<int32_t> tmp(np);
  CK(hipMemcpyAsync(tmp.data(), ctx->d_pscore, sizeof(int32_t) * np, hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  for (int64_t t = 0; t < np; ++t) out[plan.index[t]] = tmp[t];
  return MST_OK;
}

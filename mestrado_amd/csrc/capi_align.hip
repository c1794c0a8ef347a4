// C ABI of pairwise alignment with traceback (include/mestrado.h:
// mst_align_pairs), the path of AlignmentMatch::computePairwise
// (AlignmentMatch.cu:140-317) made whole.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mestrado.h"
#include "ctx.h"
#include "nw_align.h"

using namespace mst;

#define CK(call)                                                                      \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) return ctx->fail(MST_E_HIP, #call, hipGetErrorString(e_)); \
  } while (0)

static const char kLetters[] = "ARNDCQEGHILKMFPSTWYVBZX*";

// Runs the forward pass and traceback of `jobs` (already sized) and returns
// the paths on the host. Jobs are processed in batches whose flag blocks fit
// the context's traceback budget.
static int run_align_batches(mst_ctx* ctx, int src, const int8_t* d_seqs, const uint8_t* d_tab,
                             std::vector<AlignJob>& jobs, std::vector<int32_t>& scores,
                             std::vector<std::vector<uint8_t>>& paths);

static int align_class(const AlignJob& j) {
  return align_strips(j.m) == 1 ? 0 : (tiled_job(j.n, j.m) ? 2 : 1);
}

// Jobs of one 512-row strip run one per wave (align_kernel); longer ones one
// per workgroup with a wave per strip (align_pipe_kernel); the longest as
// tiles over anti-diagonal launches (align_tile_kernel). The jobs are put in
// that order for the batches and the results put back in the caller's order.
int run_align_jobs(mst_ctx* ctx, int src, const int8_t* d_seqs, const uint8_t* d_tab,
                   std::vector<AlignJob>& jobs, std::vector<int32_t>& scores,
                   std::vector<std::vector<uint8_t>>& paths) {
  const size_t nj = jobs.size();
  std::vector<size_t> perm(nj);
  for (size_t k = 0; k < nj; ++k) perm[k] = k;
  std::stable_sort(perm.begin(), perm.end(),
                   [&](size_t x, size_t y) { return align_class(jobs[x]) < align_class(jobs[y]); });
  std::vector<AlignJob> sorted(nj);
  for (size_t k = 0; k < nj; ++k) sorted[k] = jobs[perm[k]];
  std::vector<int32_t> sc;
  std::vector<std::vector<uint8_t>> pa;
  ctx->launch_stats.clear();
  int rc = run_align_batches(ctx, src, d_seqs, d_tab, sorted, sc, pa);
  if (rc) return rc;
  scores.assign(nj, 0);
  paths.assign(nj, {});
  for (size_t k = 0; k < nj; ++k) {
    jobs[perm[k]] = sorted[k];
    scores[perm[k]] = sc[k];
    paths[perm[k]].swap(pa[k]);
  }
  return MST_OK;
}

static int run_align_batches(mst_ctx* ctx, int src, const int8_t* d_seqs, const uint8_t* d_tab,
                             std::vector<AlignJob>& jobs, std::vector<int32_t>& scores,
                             std::vector<std::vector<uint8_t>>& paths) {
  const int64_t nj = (int64_t)jobs.size();
  scores.assign(nj, 0);
  paths.assign(nj, {});
  int64_t k0 = 0;
  const int64_t budget_words = ctx->flag_budget_bytes / 4;
  while (k0 < nj) {
    // batch [k0, k1)
    int64_t words = 0, opsb = 0;
    int maxn = 1;
    int64_t k1 = k0;
    while (k1 < nj) {
      int64_t w = flag_words(jobs[k1].n, jobs[k1].m);
      if (k1 > k0 && words + w > budget_words) break;
      jobs[k1].flag_off = words;
      jobs[k1].ops_off = opsb;
      words += w;
      opsb += jobs[k1].n + jobs[k1].m;
      maxn = std::max(maxn, jobs[k1].n);
      ++k1;
    }
    const int64_t nb = k1 - k0;
    int rc = ctx->ensure((void**)&ctx->d_flags, &ctx->flags_cap, sizeof(uint32_t) * std::max<int64_t>(words, 1));
    if (rc) return rc;
    rc = ctx->ensure((void**)&ctx->d_ops, &ctx->ops_cap, std::max<int64_t>(opsb, 1));
    if (rc) return rc;
    rc = ctx->ensure((void**)&ctx->d_jobs, &ctx->jobs_cap, sizeof(AlignJob) * nb);
    if (rc) return rc;
    rc = ctx->ensure((void**)&ctx->d_jscore, &ctx->jscore_cap, sizeof(int32_t) * 2 * nb);
    if (rc) return rc;
    // jobs [k0, ks) are single-strip, [ks, kt) pipelined, [kt, k1) tiled
    int64_t ks = k0;
    while (ks < k1 && align_class(jobs[ks]) == 0) ++ks;
    int64_t kt = ks;
    while (kt < k1 && align_class(jobs[kt]) == 1) ++kt;
    const int64_t n1 = ks - k0, n2 = kt - ks, n3 = k1 - kt;
    std::vector<TileJob> tiles;
    std::vector<int64_t> diag_off;
    int64_t trows = 0, tstate = 0;
    if (n3 > 0) {
      plan_tiles(jobs.data() + k0, kt - k0, k1 - k0, kTileWidth, tiles, diag_off, &trows, &tstate);
      rc = ctx->ensure(&ctx->d_tiles, &ctx->tiles_cap, sizeof(TileJob) * tiles.size());
      if (rc) return rc;
      rc = ctx->ensure(&ctx->d_trows, &ctx->trows_cap, sizeof(int2) * trows);
      if (rc) return rc;
      rc = ctx->ensure(&ctx->d_tstate, &ctx->tstate_cap, sizeof(int) * tstate);
      if (rc) return rc;
    }
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)ctx->num_cus * 16, n1));
    const int grid2 = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)ctx->num_cus * 2, n2));
    const int64_t bufcap = maxn;
    // the two launches run one after the other and share the buffers
    rc = ctx->ensure((void**)&ctx->d_bufs, &ctx->bufs_cap,
                     sizeof(int2) * bufcap * std::max<size_t>(2 * (size_t)grid, (size_t)grid2));
    if (rc) return rc;
    rc = ctx->ensure((void**)&ctx->d_acounter, &ctx->acounter_cap, sizeof(unsigned int));
    if (rc) return rc;
    CK(hipMemcpyAsync(ctx->d_jobs, jobs.data() + k0, sizeof(AlignJob) * nb, hipMemcpyHostToDevice,
                      ctx->stream));
    // events around the flag kernels and the traceback (launch statistics)
    while (ctx->launch_events.size() < 5) {
      hipEvent_t e;
      CK(hipEventCreate(&e));
      ctx->launch_events.push_back(e);
    }
    hipEvent_t* ev = ctx->launch_events.data();
    CK(hipEventRecord(ev[0], ctx->stream));
    AlignArgs a;
    a.jobs = (const AlignJob*)ctx->d_jobs;
    a.first = 0;
    a.last = nb;
    a.seqs = d_seqs;
    a.tab = d_tab;
    a.subst = ctx->d_S;
    a.flags = (uint32_t*)ctx->d_flags;
    a.scores = (int32_t*)ctx->d_jscore;
    a.bufs = ctx->d_bufs;
    a.bufcap = bufcap;
    a.go = ctx->go;
    a.ge = ctx->ge;
    a.counter = ctx->d_acounter;
    if (n1 > 0) {
      a.first = 0;
      a.last = n1;
      launch_align(src, a, grid, ctx->stream);
      CK(hipGetLastError());
    }
    CK(hipMemsetAsync(ctx->d_acounter, 0, sizeof(unsigned int), ctx->stream));
    CK(hipEventRecord(ev[1], ctx->stream));
    if (n2 > 0) {
      a.first = n1;
      a.last = n1 + n2;
      launch_align_pipe(src, 4, a, grid2, ctx->stream);
      CK(hipGetLastError());
    }
    CK(hipEventRecord(ev[2], ctx->stream));
    if (n3 > 0) {
      CK(hipMemcpyAsync(ctx->d_tiles, tiles.data(), sizeof(TileJob) * tiles.size(), hipMemcpyHostToDevice,
                        ctx->stream));
      TileArgs ta;
      ta.jobs = (const AlignJob*)ctx->d_jobs;
      ta.tiles = nullptr;
      ta.seqs = d_seqs;
      ta.tab = d_tab;
      ta.subst = ctx->d_S;
      ta.flags = (uint32_t*)ctx->d_flags;
      ta.scores = (int32_t*)ctx->d_jscore;
      ta.rows = (int2*)ctx->d_trows;
      ta.state = (int*)ctx->d_tstate;
      ta.go = ctx->go;
      ta.ge = ctx->ge;
      ta.tw = kTileWidth;
      launch_tile_diagonals(src, ta, (const TileJob*)ctx->d_tiles, diag_off, ctx->stream);
      CK(hipGetLastError());
    }
    CK(hipEventRecord(ev[4], ctx->stream));
    a.first = 0;
    a.last = nb;
    TraceArgs t;
    t.jobs = (const AlignJob*)ctx->d_jobs;
    t.first = 0;
    t.last = nb;
    t.flags = (const uint32_t*)ctx->d_flags;
    t.ops = (uint8_t*)ctx->d_ops;
    t.lens = (int32_t*)ctx->d_jscore + nb;
    launch_traceback(t, ctx->stream);
    CK(hipGetLastError());
    CK(hipEventRecord(ev[3], ctx->stream));
    std::vector<int32_t> sl(2 * nb);
    std::vector<uint8_t> ops(opsb);
    CK(hipMemcpyAsync(sl.data(), ctx->d_jscore, sizeof(int32_t) * 2 * nb, hipMemcpyDeviceToHost,
                      ctx->stream));
    if (opsb) CK(hipMemcpyAsync(ops.data(), ctx->d_ops, opsb, hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    {
      // launch statistics: -3001 one-strip flag kernel, -3004 pipelined flag
      // kernel, -3008 tiled flag kernel, -3100 traceback; cells = sum of n*m
      int64_t c1 = 0, c2 = 0, c3 = 0;
      for (int64_t k = k0; k < ks; ++k) c1 += (int64_t)jobs[k].n * jobs[k].m;
      for (int64_t k = ks; k < kt; ++k) c2 += (int64_t)jobs[k].n * jobs[k].m;
      for (int64_t k = kt; k < k1; ++k) c3 += (int64_t)jobs[k].n * jobs[k].m;
      float t1 = 0, t2 = 0, t3 = 0, t4 = 0;
      CK(hipEventElapsedTime(&t1, ev[0], ev[1]));
      CK(hipEventElapsedTime(&t2, ev[1], ev[2]));
      CK(hipEventElapsedTime(&t4, ev[2], ev[4]));
      CK(hipEventElapsedTime(&t3, ev[4], ev[3]));
      if (n1 > 0) ctx->launch_stats.push_back(LaunchStat{-3001, n1, c1, t1});
      if (n2 > 0) ctx->launch_stats.push_back(LaunchStat{-3004, n2, c2, t2});
      if (n3 > 0) ctx->launch_stats.push_back(LaunchStat{-3008, n3, c3, t4});
      ctx->launch_stats.push_back(LaunchStat{-3100, nb, c1 + c2 + c3, t3});
    }
    for (int64_t k = 0; k < nb; ++k) {
      const AlignJob& jb = jobs[k0 + k];
      scores[k0 + k] = sl[k];
      int len = sl[nb + k];
      paths[k0 + k].assign(ops.begin() + jb.ops_off, ops.begin() + jb.ops_off + len);
    }
    k0 = k1;
  }
  return MST_OK;
}

extern "C" int mst_align_pairs(mst_ctx* ctx, int id, const int32_t* pi, const int32_t* pj,
                               int64_t np, int32_t* scores, char* gx, char* gy, int64_t stride) {
  if (!ctx || np < 0 || (np > 0 && (!pi || !pj))) return MST_E_ARG;
  auto it = ctx->sets.find(id);
  if (it == ctx->sets.end()) return ctx->fail(MST_E_STATE, "unknown sequence set");
  const SeqSet& s = it->second;
  if (np == 0) return MST_OK;
  CK(hipSetDevice(ctx->device));
  std::vector<AlignJob> jobs;
  std::vector<int64_t> where;  // caller index of each device job
  std::vector<int32_t> out_scores(np, 0);
  std::vector<std::vector<uint8_t>> out_paths(np);
  for (int64_t k = 0; k < np; ++k) {
    const int a = pi[k], b = pj[k];
    if (a < 0 || b < 0 || a >= s.n || b >= s.n) return ctx->fail(MST_E_ARG, "pair index out of range");
    const int n = (int)(s.h_off[a + 1] - s.h_off[a]), m = (int)(s.h_off[b + 1] - s.h_off[b]);
    if (gx && stride < (int64_t)n + m + 1) return ctx->fail(MST_E_ARG, "stride too small");
    if (n == 0 || m == 0) {  // all gaps; no DP
      out_scores[k] = -ctx->ge * (n + m);
      out_paths[k].assign(n + m, n == 0 ? (uint8_t)OP_UP : (uint8_t)OP_LEFT);
      continue;
    }
    AlignJob j;
    j.n = n;
    j.m = m;
    j.x_off = s.h_off[a];
    j.y_off = s.h_off[b];
    j.tab_off = 0;
    j.flag_off = 0;
    j.ops_off = 0;
    jobs.push_back(j);
    where.push_back(k);
  }
  std::vector<int32_t> sc;
  std::vector<std::vector<uint8_t>> paths;
  int rc = run_align_jobs(ctx, SRC_SEQ, s.d_codes, nullptr, jobs, sc, paths);
  if (rc) return rc;
  for (size_t t = 0; t < jobs.size(); ++t) {
    out_scores[where[t]] = sc[t];
    out_paths[where[t]].swap(paths[t]);
  }
  for (int64_t k = 0; k < np; ++k) {
    if (scores) scores[k] = out_scores[k];
    if (!gx && !gy) continue;
    const int a = pi[k], b = pj[k];
    const char* X = s.h_canon.data() + s.h_off[a];
    const char* Y = s.h_canon.data() + s.h_off[b];
    int x = 0, y = 0, t = 0;
    char* ox = gx ? gx + k * stride : nullptr;
    char* oy = gy ? gy + k * stride : nullptr;
    for (uint8_t op : out_paths[k]) {
      char cx = '-', cy = '-';
      if (op != OP_UP) cx = X[x++];
      if (op != OP_LEFT) cy = Y[y++];
      if (ox) ox[t] = cx;
      if (oy) oy[t] = cy;
      ++t;
    }
    if (ox) ox[t] = 0;
    if (oy) oy[t] = 0;
  }
  return MST_OK;
}

std::string canonical_letters(const char* ascii, int64_t n) {
  static int8_t table[256];
  static bool init = false;
  if (!init) {
    for (int i = 0; i < 256; ++i) table[i] = 22;
    for (int k = 0; k < 24; ++k) {
      unsigned char c = (unsigned char)kLetters[k];
      table[c] = (int8_t)k;
      if (c >= 'A' && c <= 'Z') table[c - 'A' + 'a'] = (int8_t)k;
    }
    init = true;
  }
  std::string out(n, 'X');
  for (int64_t i = 0; i < n; ++i) out[i] = kLetters[table[(unsigned char)ascii[i]]];
  return out;
}

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

int64_t pk16_bound(const mst_ctx* ctx, int64_t R, int64_t n);  // capi_batched.hip

// The packed flag kernel takes a job when its table bytes S + go + ge are
// unsigned bytes and no 16-bit value of its DP can wrap; the tiled form keeps
// the longest jobs.
bool flag16_usable(const mst_ctx* ctx) {
  int smin = 127, smax = -128;
  for (int k = 0; k < 576; ++k) {
    smin = std::min(smin, (int)ctx->S[k]);
    smax = std::max(smax, (int)ctx->S[k]);
  }
  return smin + ctx->go + ctx->ge >= 0 && smax + ctx->go + ctx->ge <= 255;
}
// Sequence jobs of the packed kernel take their scores from query profiles in
// LDS, a nibble per row, when every entry of S is within 15 of the least;
// otherwise from a byte table built first.
static bool flag16_prof_usable(const mst_ctx* ctx) {
  int hi = -128;
  for (int k = 0; k < 576; ++k) hi = std::max(hi, (int)ctx->S[k]);
  return hi - ctx->smin <= 15;
}
// Jobs of one or two 512-row strips: one pass of the packed kernel (two jobs
// of a strip each, or one job's two strips); longer ones keep the pipelined
// kernel, whose waves run their strips at once.
bool flag16_job(const mst_ctx* ctx, int n, int m) {
  return ctx->flag16 && m <= 1024 && !tiled_job(n, m) && pk16_bound(ctx, m, n) <= 32767;
}

// classes: 0 the packed flag kernel, 1 one strip, 2 pipelined, 3 tiled
static int align_class(const mst_ctx* ctx, const AlignJob& j) {
  if (flag16_job(ctx, j.n, j.m)) return 0;
  return align_strips(j.m) == 1 ? 1 : (tiled_job(j.n, j.m) ? 3 : 2);
}

// Jobs that fit 16 bits run two per wave in the packed flag kernel
// (flag16_kernel, a score table built first for sequence jobs), paired by
// strips and columns; jobs of one 512-row strip otherwise one per wave
// (align_kernel); longer ones one per workgroup with a wave per strip
// (align_pipe_kernel); the longest as tiles over anti-diagonal launches
// (align_tile_kernel). The jobs are put in that order for the batches and the
// results put back in the caller's order.
int run_align_jobs(mst_ctx* ctx, int src, const int8_t* d_seqs, const uint8_t* d_tab,
                   std::vector<AlignJob>& jobs, std::vector<int32_t>& scores,
                   std::vector<std::vector<uint8_t>>& paths) {
  const size_t nj = jobs.size();
  std::vector<size_t> perm(nj);
  for (size_t k = 0; k < nj; ++k) perm[k] = k;
  std::vector<int> cls(nj);
  for (size_t k = 0; k < nj; ++k) cls[k] = align_class(ctx, jobs[k]);
  std::stable_sort(perm.begin(), perm.end(), [&](size_t x, size_t y) {
    if (cls[x] != cls[y]) return cls[x] < cls[y];
    if (cls[x] != 0)  // the int32 kernels take jobs in order: largest first
      return (int64_t)jobs[x].n * jobs[x].m > (int64_t)jobs[y].n * jobs[y].m;
    // the packed kernel's pairs: by rows per lane, strips, then columns
    const int kx = flag16_k(jobs[x].m), ky = flag16_k(jobs[y].m);
    if (kx != ky) return kx > ky;
    const int64_t sx = align_strips(jobs[x].m), sy = align_strips(jobs[y].m);
    if (sx != sy) return sx > sy;
    return jobs[x].n > jobs[y].n;
  });
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
  const bool prof = src == SRC_SEQ && flag16_prof_usable(ctx);
  while (k0 < nj) {
    // batch [k0, k1); a packed job's score table (8 bytes per flag word, 2
    // words of the budget) is built for sequence jobs only
    int64_t words = 0, opsb = 0, tabw = 0;
    int maxn = 1;
    int64_t k1 = k0;
    while (k1 < nj) {
      const bool packed = align_class(ctx, jobs[k1]) == 0;
      jobs[k1].k = packed ? flag16_k(jobs[k1].m) : 8;
      const int64_t w = flag_words(jobs[k1].n, jobs[k1].m, jobs[k1].k);
      const int64_t tw = (packed && src == SRC_SEQ && !prof) ? 2 * w : 0;
      if (k1 > k0 && words + tabw + w + tw > budget_words) break;
      jobs[k1].flag_off = words;
      jobs[k1].ops_off = opsb;
      jobs[k1].fmt = packed ? FMT_PLANES : FMT_NIBBLE;
      if (packed && src == SRC_SEQ && !prof) jobs[k1].tab_off = 8 * tabw / 2;
      words += w;
      tabw += tw;
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
    // jobs [k0, kf) packed, [kf, ks) single-strip, [ks, kt) pipelined, [kt, k1) tiled
    int64_t kf = k0;
    while (kf < k1 && align_class(ctx, jobs[kf]) == 0) ++kf;
    int64_t ks = kf;
    while (ks < k1 && align_class(ctx, jobs[ks]) == 1) ++ks;
    int64_t kt = ks;
    while (kt < k1 && align_class(ctx, jobs[kt]) == 2) ++kt;
    const int64_t n0 = kf - k0, n1 = ks - kf, n2 = kt - ks, n3 = k1 - kt;
    // packed items: consecutive jobs of one rows-per-lane class (sorted by
    // strips and columns) in pairs, then all items largest first
    std::vector<F16Item> items;
    std::vector<int2> tstrips;  // (job, strip) of the score tables to build
    plan_flag16_items(jobs.data() + k0, n0, items);
    if (src == SRC_SEQ && !prof)
      for (int64_t k = 0; k < n0; ++k) {
        const int sh = 64 * jobs[k0 + k].k;  // rows per strip
        for (int64_t q = 0; q < (jobs[k0 + k].m + sh - 1) / sh; ++q) tstrips.push_back(make_int2((int)k, (int)q));
      }
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
    // profile mode: twelve waves per compute unit (their LDS); else sixteen
    const int grid0 = (int)std::max<int64_t>(
        1, std::min<int64_t>((int64_t)ctx->num_cus * (prof ? 12 : 16), (int64_t)items.size()));
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)ctx->num_cus * 16, n1));
    const int grid2 = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)ctx->num_cus * 2, n2));
    const int64_t bufcap = maxn;
    // the packed kernel runs beside the int32 ones (a second stream): its row
    // buffers first, then theirs (one after the other, shared)
    const size_t bufs0 = n0 > 0 ? 2 * (size_t)grid0 * bufcap : 0;
    rc = ctx->ensure((void**)&ctx->d_bufs, &ctx->bufs_cap,
                     sizeof(int2) * (bufs0 + bufcap * std::max<size_t>(2 * (size_t)grid, (size_t)grid2)));
    if (rc) return rc;
    const bool side = n0 > 0 && n1 + n2 + n3 > 0;
    if (side && !ctx->astream) CK(hipStreamCreateWithFlags(&ctx->astream, hipStreamNonBlocking));
    hipStream_t st32 = side ? ctx->astream : ctx->stream;
    if (n0 > 0) {
      rc = ctx->ensure(&ctx->d_f16, &ctx->f16_cap,
                       sizeof(F16Item) * items.size() + sizeof(int2) * std::max<size_t>(tstrips.size(), 1));
      if (rc) return rc;
      if (src == SRC_SEQ && !prof) {
        rc = ctx->ensure((void**)&ctx->d_tab16, &ctx->tab16_cap, std::max<int64_t>(4 * tabw, 256));
        if (rc) return rc;
      }
    }
    rc = ctx->ensure((void**)&ctx->d_acounter, &ctx->acounter_cap, 2 * sizeof(unsigned int));
    if (rc) return rc;
    CK(hipMemcpyAsync(ctx->d_jobs, jobs.data() + k0, sizeof(AlignJob) * nb, hipMemcpyHostToDevice,
                      ctx->stream));
    // events around the flag kernels and the traceback (launch statistics)
    while (ctx->launch_events.size() < 9) {
      hipEvent_t e;
      CK(hipEventCreate(&e));
      ctx->launch_events.push_back(e);
    }
    hipEvent_t* ev = ctx->launch_events.data();
    CK(hipEventRecord(ev[5], ctx->stream));
    if (side) CK(hipStreamWaitEvent(st32, ev[5], 0));
    CK(hipEventRecord(ev[0], st32));
    AlignArgs a;
    a.jobs = (const AlignJob*)ctx->d_jobs;
    a.first = 0;
    a.last = nb;
    a.seqs = d_seqs;
    a.tab = d_tab;
    a.subst = ctx->d_S;
    a.flags = (uint32_t*)ctx->d_flags;
    a.scores = (int32_t*)ctx->d_jscore;
    a.bufs = ctx->d_bufs + bufs0;
    a.bufcap = bufcap;
    a.go = ctx->go;
    a.ge = ctx->ge;
    a.counter = ctx->d_acounter;
    if (n1 > 0) {
      a.first = n0;
      a.last = n0 + n1;
      launch_align(src, a, grid, st32);
      CK(hipGetLastError());
    }
    CK(hipMemsetAsync(ctx->d_acounter, 0, sizeof(unsigned int), st32));
    CK(hipEventRecord(ev[1], st32));
    if (n2 > 0) {
      a.first = n0 + n1;
      a.last = n0 + n1 + n2;
      launch_align_pipe(src, 4, a, grid2, st32);
      CK(hipGetLastError());
    }
    CK(hipEventRecord(ev[2], st32));
    if (n3 > 0) {
      CK(hipMemcpyAsync(ctx->d_tiles, tiles.data(), sizeof(TileJob) * tiles.size(), hipMemcpyHostToDevice,
                        st32));
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
      launch_tile_diagonals(src, ta, (const TileJob*)ctx->d_tiles, diag_off, st32);
      CK(hipGetLastError());
    }
    // the packed kernel after the int32 ones (another stream): their fewer,
    // longer workgroups are placed first, the packed items fill the rest
    if (n0 > 0) {
      F16Item* d_items = (F16Item*)ctx->d_f16;
      int2* d_tstrips = (int2*)(d_items + items.size());
      CK(hipMemcpyAsync(d_items, items.data(), sizeof(F16Item) * items.size(), hipMemcpyHostToDevice,
                        ctx->stream));
      const uint8_t* tab = d_tab;
      if (src == SRC_SEQ && !prof) {
        CK(hipMemcpyAsync(d_tstrips, tstrips.data(), sizeof(int2) * tstrips.size(), hipMemcpyHostToDevice,
                          ctx->stream));
        launch_seq_table((const AlignJob*)ctx->d_jobs, d_tstrips, (int)tstrips.size(), d_seqs, ctx->d_S,
                         ctx->go + ctx->ge, (uint8_t*)ctx->d_tab16, ctx->stream);
        CK(hipGetLastError());
        tab = (const uint8_t*)ctx->d_tab16;
      }
      CK(hipEventRecord(ev[6], ctx->stream));
      F16Args f;
      f.jobs = (const AlignJob*)ctx->d_jobs;
      f.items = d_items;
      f.nitems = (int64_t)items.size();
      f.tab = tab;
      f.flags = (uint32_t*)ctx->d_flags;
      f.scores = (int32_t*)ctx->d_jscore;
      f.bufs = ctx->d_bufs;
      f.bufcap = bufcap;
      f.go = ctx->go;
      f.ge = ctx->ge;
      f.prof = prof ? 1 : 0;
      f.smin = ctx->smin;
      f.seqs = d_seqs;
      f.subst = ctx->d_S;
      f.counter = ctx->d_acounter + 1;
      CK(hipMemsetAsync(f.counter, 0, sizeof(unsigned int), ctx->stream));
      launch_flag16(f, grid0, ctx->stream);
      CK(hipGetLastError());
    } else {
      CK(hipEventRecord(ev[6], ctx->stream));
    }
    CK(hipEventRecord(ev[7], ctx->stream));  // (the packed kernel's end)
    CK(hipEventRecord(ev[4], st32));
    if (side) CK(hipStreamWaitEvent(ctx->stream, ev[4], 0));
    CK(hipEventRecord(ev[8], ctx->stream));  // every flag kernel is done
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
      // launch statistics: -3000 the whole forward pass (score tables and
      // every flag kernel, the packed ones beside the int32 ones), -3016 packed
      // flag kernel (-3017 its score tables), -3001 one-strip flag kernel,
      // -3004 pipelined flag kernel, -3008 tiled flag kernel, -3100
      // traceback; cells = sum of n*m
      int64_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
      for (int64_t k = k0; k < kf; ++k) c0 += (int64_t)jobs[k].n * jobs[k].m;
      for (int64_t k = kf; k < ks; ++k) c1 += (int64_t)jobs[k].n * jobs[k].m;
      for (int64_t k = ks; k < kt; ++k) c2 += (int64_t)jobs[k].n * jobs[k].m;
      for (int64_t k = kt; k < k1; ++k) c3 += (int64_t)jobs[k].n * jobs[k].m;
      float t0 = 0, tt = 0, t1 = 0, t2 = 0, t3 = 0, t4 = 0, tall = 0;
      CK(hipEventElapsedTime(&tt, ev[5], ev[6]));
      CK(hipEventElapsedTime(&t0, ev[6], ev[7]));
      CK(hipEventElapsedTime(&tall, ev[5], ev[8]));
      CK(hipEventElapsedTime(&t1, ev[0], ev[1]));
      CK(hipEventElapsedTime(&t2, ev[1], ev[2]));
      CK(hipEventElapsedTime(&t4, ev[2], ev[4]));
      CK(hipEventElapsedTime(&t3, ev[8], ev[3]));
      ctx->launch_stats.push_back(LaunchStat{-3000, nb, c0 + c1 + c2 + c3, tall});
      if (n0 > 0) {
        if (src == SRC_SEQ) ctx->launch_stats.push_back(LaunchStat{-3017, n0, c0, tt});
        ctx->launch_stats.push_back(LaunchStat{-3016, n0, c0, t0});
      }
      if (n1 > 0) ctx->launch_stats.push_back(LaunchStat{-3001, n1, c1, t1});
      if (n2 > 0) ctx->launch_stats.push_back(LaunchStat{-3004, n2, c2, t2});
      if (n3 > 0) ctx->launch_stats.push_back(LaunchStat{-3008, n3, c3, t4});
      ctx->launch_stats.push_back(LaunchStat{-3100, nb, c0 + c1 + c2 + c3, t3});
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
    j.fmt = FMT_NIBBLE;
    j.k = 8;
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

// C ABI of the progressive alignment (include/mestrado.h: mst_msa*).
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/mestrado.h"
#include "ctx.h"
#include "msa.h"
#include "nw_align.h"

using namespace mst;

#define CK(call)                                                                      \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) return ctx->fail(MST_E_HIP, #call, hipGetErrorString(e_)); \
  } while (0)

namespace {

struct Group {
  int k = 0, len = 0;
  uint8_t* rows = nullptr;  // k x len codes, forward order, 24 = gap
  std::vector<int> members;
};

struct MsaScratch {
  int32_t *cx = nullptr, *cy = nullptr;
  float* P = nullptr;
  uint2* tab = nullptr;
  uint32_t* flags = nullptr;
  uint8_t* ops = nullptr;
  int32_t *sx = nullptr, *sy = nullptr, *res = nullptr;  // res[0] score, res[1] length
  AlignJob* job = nullptr;
  int2* bufs = nullptr;
  unsigned int* counter = nullptr;  // pipelined flag kernel's work queue head
  size_t ccx = 0, ccy = 0, cP = 0, ctab = 0, cfl = 0, cops = 0, csx = 0, csy = 0, cres = 0, cjob = 0,
         cbuf = 0, ccnt = 0;
};

}  // namespace

static int ensure_all(mst_ctx* ctx, MsaScratch& w, int n, int m) {
  const int64_t steps = flag_words(n, m);  // one uint2 of table, one word of flags per step and lane
  int rc;
  if ((rc = ctx->ensure((void**)&w.cx, &w.ccx, sizeof(int32_t) * 25 * (size_t)n))) return rc;
  if ((rc = ctx->ensure((void**)&w.cy, &w.ccy, sizeof(int32_t) * 25 * (size_t)m))) return rc;
  if ((rc = ctx->ensure((void**)&w.P, &w.cP, sizeof(float) * 24 * (size_t)m))) return rc;
  if ((rc = ctx->ensure((void**)&w.tab, &w.ctab, sizeof(uint2) * (size_t)steps))) return rc;
  if ((rc = ctx->ensure((void**)&w.flags, &w.cfl, sizeof(uint32_t) * (size_t)steps))) return rc;
  if ((rc = ctx->ensure((void**)&w.ops, &w.cops, (size_t)n + m))) return rc;
  if ((rc = ctx->ensure((void**)&w.sx, &w.csx, sizeof(int32_t) * ((size_t)n + m)))) return rc;
  if ((rc = ctx->ensure((void**)&w.sy, &w.csy, sizeof(int32_t) * ((size_t)n + m)))) return rc;
  if ((rc = ctx->ensure((void**)&w.res, &w.cres, sizeof(int32_t) * 2))) return rc;
  if ((rc = ctx->ensure((void**)&w.job, &w.cjob, sizeof(AlignJob)))) return rc;
  if ((rc = ctx->ensure((void**)&w.bufs, &w.cbuf, sizeof(int2) * 2 * (size_t)n))) return rc;
  if ((rc = ctx->ensure((void**)&w.counter, &w.ccnt, sizeof(unsigned int)))) return rc;
  return MST_OK;
}

extern "C" int mst_msa(mst_ctx* ctx, int id, const mst_join* joins, int* msa_len) {
  if (!ctx || !msa_len) return MST_E_ARG;
  auto it = ctx->sets.find(id);
  if (it == ctx->sets.end()) return ctx->fail(MST_E_STATE, "unknown sequence set");
  const SeqSet& s = it->second;
  const int n0 = s.n;
  if (n0 < 1 || (n0 > 1 && !joins)) return MST_E_ARG;
  CK(hipSetDevice(ctx->device));
  for (int k = 0; k + 1 < n0; ++k) {
    const mst_join& J = joins[k];
    if (J.node != n0 + k || J.left < 0 || J.right < 0 || J.left >= J.node || J.right >= J.node ||
        J.left == J.right)
      return ctx->fail(MST_E_ARG, "malformed join list");
  }
  std::vector<Group> g(2 * (size_t)n0 - 1);
  MsaScratch w;
  int rc = MST_OK;
  auto cleanup = [&]() {
    for (auto& x : g)
      if (x.rows) hipFree(x.rows);
    void* ps[] = {w.cx, w.cy, w.P, w.tab, w.flags, w.ops, w.sx, w.sy, w.res, w.job, w.bufs,
                  w.counter};
    for (void* p : ps)
      if (p) hipFree(p);
  };
  for (int i = 0; i < n0 && rc == MST_OK; ++i) {
    const int len = (int)(s.h_off[i + 1] - s.h_off[i]);
    g[i].k = 1;
    g[i].len = len;
    g[i].members = {i};
    if (hipMalloc(&g[i].rows, std::max(len, 1)) != hipSuccess) {
      rc = ctx->fail(MST_E_NOMEM, "hipMalloc leaf rows");
      break;
    }
    launch_leaf_rows(s.d_codes, s.h_off[i], len, g[i].rows, ctx->stream);
  }
  for (int k = 0; k + 1 < n0 && rc == MST_OK; ++k) {
    const mst_join& J = joins[k];
    Group& X = g[J.left];
    Group& Y = g[J.right];
    Group& Z = g[J.node];
    if (!X.rows || !Y.rows) {
      rc = ctx->fail(MST_E_ARG, "join uses a node twice");
      break;
    }
    const int n = X.len, m = Y.len;
    Z.k = X.k + Y.k;
    Z.members = X.members;
    Z.members.insert(Z.members.end(), Y.members.begin(), Y.members.end());
    int L = 0;
    if (n == 0 || m == 0) {  // nothing to align: the other side's columns, all gaps opposite
      L = n + m;
      std::vector<uint8_t> ops(L, n == 0 ? 2 : 1);
      if ((rc = ensure_all(ctx, w, std::max(n, 1), std::max(m, 1)))) break;
      if (hipMemcpyAsync(w.ops, ops.data(), L, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
          hipMemcpyAsync(w.res + 1, &L, sizeof(int), hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
        rc = ctx->fail(MST_E_HIP, "hipMemcpyAsync");
        break;
      }
    } else {
      if ((rc = ensure_all(ctx, w, n, m))) break;
      launch_profile_counts(X.rows, X.k, n, w.cx, ctx->stream);
      launch_profile_counts(Y.rows, Y.k, m, w.cy, ctx->stream);
      launch_profile_p(w.cy, m, Y.k, ctx->d_S, w.P, ctx->stream);
      launch_profile_table(w.cx, n, X.k, w.P, m, ctx->go + ctx->ge, w.tab, ctx->stream);
      AlignJob job;
      job.n = n;
      job.m = m;
      job.x_off = job.y_off = 0;
      job.tab_off = 0;
      job.flag_off = 0;
      job.ops_off = 0;
      if (hipMemcpyAsync(w.job, &job, sizeof job, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
        rc = ctx->fail(MST_E_HIP, "hipMemcpyAsync");
        break;
      }
      AlignArgs a;
      a.jobs = w.job;
      a.first = 0;
      a.last = 1;
      a.seqs = s.d_codes;
      a.tab = (const uint8_t*)w.tab;
      a.subst = ctx->d_S;
      a.flags = w.flags;
      a.scores = w.res;
      a.bufs = w.bufs;
      a.bufcap = n;
      a.go = ctx->go;
      a.ge = ctx->ge;
      a.counter = w.counter;
      if (align_strips(m) > 1) {  // a wave per strip, 8 strips in flight
        if (hipMemsetAsync(w.counter, 0, sizeof(unsigned int), ctx->stream) != hipSuccess) {
          rc = ctx->fail(MST_E_HIP, "hipMemsetAsync");
          break;
        }
        launch_align_pipe(SRC_TAB, 8, a, 1, ctx->stream);
      } else {
        launch_align(SRC_TAB, a, 1, ctx->stream);
      }
      TraceArgs t;
      t.jobs = w.job;
      t.first = 0;
      t.last = 1;
      t.flags = w.flags;
      t.ops = w.ops;
      t.lens = w.res + 1;
      launch_traceback(t, ctx->stream);
      if (hipGetLastError() != hipSuccess) {
        rc = ctx->fail(MST_E_HIP, "profile alignment launch");
        break;
      }
      if (hipMemcpyAsync(&L, w.res + 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
          hipStreamSynchronize(ctx->stream) != hipSuccess) {
        rc = ctx->fail(MST_E_HIP, "profile alignment");
        break;
      }
    }
    Z.len = L;
    if (hipMalloc(&Z.rows, std::max<size_t>((size_t)Z.k * L, 1)) != hipSuccess) {
      rc = ctx->fail(MST_E_NOMEM, "hipMalloc group rows");
      break;
    }
    launch_path_columns(w.ops, w.res + 1, w.sx, w.sy, ctx->stream);
    launch_merge_rows(X.rows, X.k, n, Y.rows, Y.k, m, w.sx, w.sy, L, Z.rows, ctx->stream);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) {
      rc = ctx->fail(MST_E_HIP, "merge");
      break;
    }
    hipFree(X.rows);
    hipFree(Y.rows);
    X.rows = Y.rows = nullptr;
  }
  if (rc == MST_OK) {
    Group& R = g[2 * (size_t)n0 - 2];
    std::vector<uint8_t> h((size_t)R.k * R.len);
    if (!h.empty() &&
        hipMemcpy(h.data(), R.rows, h.size(), hipMemcpyDeviceToHost) != hipSuccess)
      rc = ctx->fail(MST_E_HIP, "hipMemcpy msa");
    if (rc == MST_OK) {
      ctx->msa_len = R.len;
      ctx->msa_rows.assign((size_t)n0 * R.len, 0);
      for (int r = 0; r < R.k; ++r)
        memcpy(&ctx->msa_rows[(size_t)R.members[r] * R.len], &h[(size_t)r * R.len], R.len);
      ctx->msa_n = n0;
      *msa_len = R.len;
    }
  }
  cleanup();
  return rc;
}

extern "C" int mst_msa_rows(mst_ctx* ctx, char* rows, int64_t stride) {
  static const char kLetters[] = "ARNDCQEGHILKMFPSTWYVBZX*-";
  if (!ctx || !rows) return MST_E_ARG;
  if (ctx->msa_n <= 0) return ctx->fail(MST_E_STATE, "no alignment computed");
  if (stride < ctx->msa_len + 1) return ctx->fail(MST_E_ARG, "stride too small");
  for (int r = 0; r < ctx->msa_n; ++r) {
    char* o = rows + (int64_t)r * stride;
    const uint8_t* src = &ctx->msa_rows[(size_t)r * ctx->msa_len];
    for (int t = 0; t < ctx->msa_len; ++t) o[t] = kLetters[src[t] < 25 ? src[t] : 24];
    o[ctx->msa_len] = 0;
  }
  return MST_OK;
}

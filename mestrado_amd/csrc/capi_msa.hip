// C ABI of the progressive alignment (include/mestrado.h: mst_msa*).
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/mestrado.h"
#include "ctx.h"
#include "merge.h"
#include "pool.h"

using namespace mst;

#define CK(call)                                                                      \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) return ctx->fail(MST_E_HIP, #call, hipGetErrorString(e_)); \
  } while (0)

// Merges run level by level of the guide tree (a node's level is one more
// than its deeper child's), every level as few batches of the merge engine
// as its scratch budget allows; group rows live in device regions carved on
// demand, a merged group's block shrunk to its path length, a child's block
// freed once merged.
extern "C" int mst_msa(mst_ctx* ctx, int id, const mst_join* joins, int* msa_len) {
  if (!ctx || !msa_len) return MST_E_ARG;
  auto it = ctx->sets.find(id);
  if (it == ctx->sets.end()) return ctx->fail(MST_E_STATE, "unknown sequence set");
  const SeqSet& s = it->second;
  const int n0 = s.n;
  if (n0 < 1 || (n0 > 1 && !joins)) return MST_E_ARG;
  CK(hipSetDevice(ctx->device));
  std::vector<int> level(2 * (size_t)n0 - 1, 0);
  int maxlvl = 0;
  for (int k = 0; k + 1 < n0; ++k) {
    const mst_join& J = joins[k];
    if (J.node != n0 + k || J.left < 0 || J.right < 0 || J.left >= J.node || J.right >= J.node ||
        J.left == J.right)
      return ctx->fail(MST_E_ARG, "malformed join list");
    level[J.node] = 1 + std::max(level[J.left], level[J.right]);
    maxlvl = std::max(maxlvl, level[J.node]);
  }
  {  // every node is merged at most once
    std::vector<char> used(2 * (size_t)n0 - 1, 0);
    for (int k = 0; k + 1 < n0; ++k) {
      if (used[joins[k].left] || used[joins[k].right]) return ctx->fail(MST_E_ARG, "join uses a node twice");
      used[joins[k].left] = used[joins[k].right] = 1;
    }
  }
  struct Group {
    uint8_t* rows = nullptr;
    int64_t cap = 0;
    int k = 0, len = 0;
    bool leaf = false;
    std::vector<int> members;
  };
  std::vector<Group> g(2 * (size_t)n0 - 1);
  const int64_t total = s.h_off[n0] - s.h_off[0];
  GrowPool pool(std::max<int64_t>((int64_t)64 << 20, 8 * total));
  MergeEngine eng;
  uint8_t* leafblock = pool.alloc(std::max<int64_t>(total, 1));
  if (!leafblock) return ctx->fail(MST_E_NOMEM, "msa leaf rows");
  std::vector<LeafDesc> lv(n0);
  for (int i = 0; i < n0; ++i) {
    Group& G = g[i];
    G.k = 1;
    G.len = (int)(s.h_off[i + 1] - s.h_off[i]);
    G.leaf = true;
    G.rows = leafblock + (s.h_off[i] - s.h_off[0]);
    G.members = {i};
    lv[i].src = s.h_off[i];
    lv[i].dst = G.rows;
    lv[i].len = G.len;
    lv[i].pad = 0;
  }
  CK(eng.leaves(s.d_codes, lv.data(), n0, ctx->stream));
  std::vector<std::vector<int>> by_level(maxlvl + 1);
  for (int k = 0; k + 1 < n0; ++k) by_level[level[n0 + k]].push_back(k);
  const int64_t budget = std::max<int64_t>(ctx->flag_budget_bytes, (int64_t)64 << 20);
  std::vector<MergeSpec> spec;
  std::vector<int> batch;
  for (int lv_i = 1; lv_i <= maxlvl; ++lv_i) {
    const std::vector<int>& list = by_level[lv_i];
    size_t q = 0;
    while (q < list.size()) {
      spec.clear();
      batch.clear();
      int64_t bytes = 0;
      while (q < list.size()) {
        const mst_join& J = joins[list[q]];
        const Group& X = g[J.left];
        const Group& Y = g[J.right];
        const int64_t b = merge_scratch_bytes(X.len, Y.len);
        if (!batch.empty() && bytes + b > budget) break;
        bytes += b;
        Group& Z = g[J.node];
        Z.k = X.k + Y.k;
        Z.cap = std::max<int64_t>((int64_t)Z.k * ((int64_t)X.len + Y.len), 1);
        Z.rows = pool.alloc(Z.cap);
        if (!Z.rows) return ctx->fail(MST_E_NOMEM, "msa group rows");
        Z.members = X.members;
        Z.members.insert(Z.members.end(), Y.members.begin(), Y.members.end());
        spec.push_back(MergeSpec{X.rows, X.k, X.len, Y.rows, Y.k, Y.len, Z.rows});
        batch.push_back(list[q]);
        ++q;
      }
      CK(eng.enqueue(spec.data(), (int)spec.size(), ctx->d_S, ctx->go, ctx->ge, ctx->num_cus, ctx->stream,
                   ctx->flag16 ? ctx->smax : -1));
      CK(hipStreamSynchronize(ctx->stream));
      for (size_t b = 0; b < batch.size(); ++b) {
        const mst_join& J = joins[batch[b]];
        Group& Z = g[J.node];
        Z.len = eng.lens()[b];
        pool.shrink(Z.rows, Z.cap, (int64_t)Z.k * Z.len);
        Z.cap = std::max<int64_t>((int64_t)Z.k * Z.len, 1);
        for (int c : {J.left, J.right}) {
          Group& C = g[c];
          if (!C.leaf) pool.release(C.rows, C.cap);
          C.rows = nullptr;
          C.members.clear();
        }
      }
    }
  }
  Group& R = g[2 * (size_t)n0 - 2];
  std::vector<uint8_t> h((size_t)R.k * R.len);
  if (!h.empty()) CK(hipMemcpy(h.data(), R.rows, h.size(), hipMemcpyDeviceToHost));
  ctx->msa_len = R.len;
  ctx->msa_rows.assign((size_t)n0 * R.len, 0);
  for (int r = 0; r < R.k; ++r)
    memcpy(&ctx->msa_rows[(size_t)R.members[r] * R.len], &h[(size_t)r * R.len], R.len);
  ctx->msa_n = n0;
  *msa_len = R.len;
  return MST_OK;
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

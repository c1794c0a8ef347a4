// C ABI of the neighbour-joining step (include/mestrado.h: mst_nj*).
#include <cstring>
#include <vector>

#include "../../include/mestrado.h"
#include "ctx.h"
#include "nj.h"

using namespace mst;

#define CK(call)                                                                      \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) return ctx->fail(MST_E_HIP, #call, hipGetErrorString(e_)); \
  } while (0)

// TreeNode::join (TreeNode.cpp:30-58): lengths in fp64, a negative side is
// clamped to zero and its excess moved to the other side.
void mst::nj_to_join(const NjJoin& in, mst_join* out) {
  mst_join& j = *out;
  j.node = in.node;
  j.left = in.left;
  j.right = in.right;
  j.d = in.d;
  j.Ti = in.Ti;
  j.Tj = in.Tj;
  j.n = in.n;
  j.q = in.q;
  double L, R, half = 0.5;
  double minD = (double)j.d;
  if (j.n == 2) {
    R = L = (minD * half);
  } else {
    L = half * (minD + (j.Ti - j.Tj) / (j.n - 2));
    R = half * (minD + (j.Tj - j.Ti) / (j.n - 2));
  }
  if (L < 0.0) {
    R -= L;
    L = 0.0;
  } else if (R < 0.0) {
    L -= R;
    R = 0.0;
  }
  j.len_left = L;
  j.len_right = R;
}

// Matrices from this size on run as launches per step (nj_scan_kernel over
// the device); smaller ones in one workgroup (nj_kernel).
constexpr int kNjStepsMinN = 256;

// Runs NJ on the device matrix ctx->d_nj (n x n, filled by the caller).
static int run_nj(mst_ctx* ctx, int n, float ratio, mst_join* joins) {
  std::vector<int> init(2 * (size_t)n, 0);
  for (int i = 0; i < n; ++i) init[i] = i;
  int rc = ctx->ensure((void**)&ctx->d_nj_T, &ctx->nj_T_cap, sizeof(double) * 3 * (size_t)n);
  if (rc) return rc;
  rc = ctx->ensure((void**)&ctx->d_nj_act, &ctx->nj_act_cap, sizeof(int) * 4 * (size_t)n);
  if (rc) return rc;
  rc = ctx->ensure((void**)&ctx->d_nj_joins, &ctx->nj_joins_cap, sizeof(NjJoin) * (size_t)n);
  if (rc) return rc;
  CK(hipMemcpyAsync(ctx->d_nj_act, init.data(), sizeof(int) * 2 * n, hipMemcpyHostToDevice, ctx->stream));
  CK(hipMemcpyAsync(ctx->d_nj_act + 2 * n, init.data(), sizeof(int) * 2 * n, hipMemcpyHostToDevice,
                    ctx->stream));
  NjArgs a;
  a.D = ctx->d_nj;
  a.ld = n;
  a.n0 = n;
  a.T = ctx->d_nj_T;
  a.Tpos = ctx->d_nj_T + n;
  a.act = ctx->d_nj_act;
  a.ids = ctx->d_nj_act + 2 * n;
  a.joins = (NjJoin*)ctx->d_nj_joins;
  a.ratio = ratio;
  if (n >= kNjStepsMinN) {
    // large N: one launch per step, the scan over the whole device
    const int nparts = ctx->num_cus * 4;  // 16 waves per CU
    rc = ctx->ensure(&ctx->d_nj_parts, &ctx->nj_parts_cap, nj_part_bytes(nparts));
    if (rc) return rc;
    CK(run_nj_steps(a, ctx->d_nj_parts, nparts, ctx->stream));
  } else {
    CK(launch_nj(a, ctx->stream));
  }
  std::vector<NjJoin> hj(n > 1 ? n - 1 : 1);
  if (n > 1)
    CK(hipMemcpyAsync(hj.data(), ctx->d_nj_joins, sizeof(NjJoin) * (n - 1), hipMemcpyDeviceToHost,
                      ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  for (int k = 0; k + 1 < n; ++k) nj_to_join(hj[k], &joins[k]);
  return MST_OK;
}

extern "C" int mst_nj(mst_ctx* ctx, const float* D, int n, float ratio, mst_join* joins) {
  if (!ctx || n < 1 || !D || (n > 1 && !joins)) return MST_E_ARG;
  if (n > 46340) return ctx->fail(MST_E_ARG, "NJ matrix too large");
  CK(hipSetDevice(ctx->device));
  int rc = ctx->ensure((void**)&ctx->d_nj, &ctx->nj_cap, sizeof(float) * (size_t)n * n);
  if (rc) return rc;
  CK(hipMemcpyAsync(ctx->d_nj, D, sizeof(float) * (size_t)n * n, hipMemcpyHostToDevice, ctx->stream));
  return run_nj(ctx, n, ratio, joins);
}

extern "C" int mst_nj_scores(mst_ctx* ctx, int id, float ratio, mst_join* joins) {
  if (!ctx) return MST_E_ARG;
  auto it = ctx->sets.find(id);
  if (it == ctx->sets.end()) return ctx->fail(MST_E_STATE, "unknown sequence set");
  if (ctx->mat_set != id) return ctx->fail(MST_E_STATE, "no resident score matrix for this set");
  const int n = it->second.n;
  if (n < 1 || (n > 1 && !joins)) return MST_E_ARG;
  CK(hipSetDevice(ctx->device));
  int rc = ctx->ensure((void**)&ctx->d_nj, &ctx->nj_cap, sizeof(float) * (size_t)n * n);
  if (rc) return rc;
  int64_t nn = (int64_t)n * n;
  launch_scores_to_float(ctx->d_mat, ctx->d_nj, nn, ctx->stream);
  CK(hipGetLastError());
  return run_nj(ctx, n, ratio, joins);
}

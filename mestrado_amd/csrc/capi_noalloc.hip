// mst_score_pairs_noalloc: the explicit pair list of mst_score_pairs (plain
// form) without any device allocation, for callers that cannot allocate
// (code under stream capture, tight serving loops, harnesses that forbid
// hipMalloc). Pair indices, scores and the int32 kernel's strip row buffers
// live in one file-scope device array, `scratch`, whose per-device address
// is taken once with hipGetSymbolAddress when a context is created on that
// device (noalloc_init_device, called by mst_create).
//
// Layout of the array for one call of np pairs over a set whose longest
// sequence has L residues (A = np * 4 bytes rounded up to 256):
//   [0, A)            column sequence index per pair (copied from pi)
//   [A, 2A)           row sequence index per pair (pj)
//   [2A, 3A)          score per pair (copied back to out)
//   [3A, 3A + 16 L g) row buffers of the g waves of the launch (two strip
//                     boundaries of L (F, H) int2 entries per wave)
// Every word a call reads is written earlier in the same call (indices by the
// copies, a wave's row buffer by its previous strip), so nothing needs to be
// cleared between calls. All contexts on one device share its array: calls
// that overlap on one device race.
#include <algorithm>
#include <mutex>

#include "../../include/mestrado.h"
#include "ctx.h"
#include "nw_score.h"

using namespace mst;

namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) char scratch[kScratchBytes];

void* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

inline size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

int k_for_rows(int m);  // capi.hip

// Device address of `scratch` on `device` (the current device), taken once.
int noalloc_init_device(int device, void** ptr) {
  if (device < 0 || device >= kMaxDevices) return MST_E_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MST_E_HIP;
    g_scratch[device] = p;
  }
  *ptr = g_scratch[device];
  return MST_OK;
}

#define CK(call)                                                                      \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) return ctx->fail(MST_E_HIP, #call, hipGetErrorString(e_)); \
  } while (0)

extern "C" int mst_score_pairs_noalloc(mst_ctx* ctx, int id, const int32_t* pi, const int32_t* pj,
                                       int64_t np, int32_t* out) {
  if (!ctx || np < 0 || (np > 0 && (!pi || !pj || !out))) return MST_E_ARG;
  auto it = ctx->sets.find(id);
  if (it == ctx->sets.end()) return ctx->fail(MST_E_STATE, "unknown sequence set");
  if (!ctx->d_scratch) return ctx->fail(MST_E_STATE, "mst_score_pairs_noalloc", "no scratch address");
  if (np == 0) return MST_OK;
  const SeqSet& s = it->second;
  const int64_t bufcap = s.maxlen > 0 ? s.maxlen : 1;
  const size_t A = round256((size_t)np * sizeof(int32_t));
  const size_t wave_bytes = (size_t)2 * bufcap * sizeof(int2);
  if (3 * A + wave_bytes > kScratchBytes)
    return ctx->fail(MST_E_ARG, "mst_score_pairs_noalloc",
                     "pairs and row buffers exceed the 64 MiB scratch (12 B per pair plus 16 B per "
                     "residue of the set's longest sequence)");
  // rows of the longest pair (the shorter sequence when S is symmetric) pick
  // the kernel's rows per lane, one launch for the whole list
  int rows = 0;
  for (int64_t k = 0; k < np; ++k) {
    const int a = pi[k], b = pj[k];
    if (a < 0 || b < 0 || a >= s.n || b >= s.n)
      return ctx->fail(MST_E_ARG, "mst_score_pairs_noalloc", "pair index out of range");
    const int la = (int)(s.h_off[a + 1] - s.h_off[a]), lb = (int)(s.h_off[b + 1] - s.h_off[b]);
    const int r = ctx->symmetric ? (la < lb ? la : lb) : lb;
    rows = r > rows ? r : rows;
  }
  const int K = k_for_rows(rows);
  int64_t grid = (int64_t)ctx->num_cus * 16;
  grid = std::min<int64_t>(grid, (int64_t)((kScratchBytes - 3 * A) / wave_bytes));
  grid = std::min<int64_t>(grid, np);
  char* base = (char*)ctx->d_scratch;
  int32_t* d_pi = (int32_t*)base;
  int32_t* d_pj = (int32_t*)(base + A);
  int32_t* d_out = (int32_t*)(base + 2 * A);
  int2* d_bufs = (int2*)(base + 3 * A);
  CK(hipSetDevice(ctx->device));
  CK(hipMemcpyAsync(d_pi, pi, np * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
  CK(hipMemcpyAsync(d_pj, pj, np * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
  ScoreArgs a;
  a.seqs = s.d_codes;
  a.off = s.d_off;
  a.pairs = nullptr;
  a.pi = d_pi;
  a.pj = d_pj;
  a.first = 0;
  a.last = np;
  a.out_pair = d_out;
  a.out_mat = nullptr;
  a.nseq = s.n;
  a.bufs = d_bufs;
  a.bufcap = bufcap;
  a.subst = ctx->d_S;
  a.go = ctx->go;
  a.ge = ctx->ge;
  a.transpose_ok = ctx->symmetric ? 1 : 0;
  launch_score_pairs(K, a, (int)grid, ctx->stream);
  CK(hipGetLastError());
  CK(hipMemcpyAsync(out, d_out, np * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  return MST_OK;
}

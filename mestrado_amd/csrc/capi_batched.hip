// Batched form of the all-pairs scores: planning, graph capture and replay.
//
// Plan (symmetric substitution matrix): sequences are sorted by length; row
// sequences are paired consecutively (Ya = s[2k], Yb = s[2k+1]) and run against
// every column sequence at or after s[2k+1] in that order, so each pair i<=j is
// scored exactly once with the shorter sequence in the rows. A row pair's
// column list is cut into chunks of consecutive positions: one chunk is one
// work item, swept as one column stream by the kernel of the row pair's class
// (the routes are listed below: one-wave, two-wave and wider pipelines of the
// streamed kernel, the multi-sweep kernel, optionally stacked row pairs).
// Pairs whose 16-bit bound fails go to the int32 kernel. Items are ordered longest first inside every launch; the
// packed launches are captured once into a hipGraph (in parallel branches)
// and replayed, and the int32 launches follow the replay on the same stream.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/mestrado.h"
#include "batched.h"
#include "ctx.h"
#include "nw_score.h"
#include "nw_score5.h"

using namespace mst;

#define CK(call)                                                                      \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) return ctx->fail(MST_E_HIP, #call, hipGetErrorString(e_)); \
  } while (0)

namespace {

__global__ void zero_u32(unsigned int* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// codes of the sequence at sorted position blockIdx.x into its place in the
// length-ordered copy
__global__ void gather_codes(const int8_t* codes, const int64_t* off, const int32_t* order,
                             const int64_t* soff, int8_t* out) {
  const int k = blockIdx.x;
  const int id = order[k];
  const int64_t n = soff[k + 1] - soff[k];
  const int8_t* src = codes + off[id];
  int8_t* dst = out + soff[k];
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// codes as profile offsets (code x kProfStride8)
__global__ void codes_to_offsets8(const int8_t* codes, int64_t n, uint32_t* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (uint32_t)(uint8_t)codes[i] * kProfStride8;
}

}  // namespace

// Largest value a packed 16-bit lane can reach for R padded rows and n columns.
int64_t pk16_bound(const mst_ctx* ctx, int64_t R, int64_t n) {
  const int64_t smax = ctx->smax;
  return smax * (std::min(R, n) + 1) + (int64_t)ctx->ge * (R + n + 2) + smax + ctx->go + ctx->ge;
}

bool pk16_usable(const mst_ctx* ctx) {
  int smin = 127, smax = -128;
  for (int k = 0; k < 576; ++k) {
    smin = std::min(smin, (int)ctx->S[k]);
    smax = std::max(smax, (int)ctx->S[k]);
  }
  // profile bytes S+go+ge are zero-extended: they must lie in [0, 255]
  return ctx->symmetric && smin + ctx->go + ctx->ge >= 0 && smax + ctx->go + ctx->ge <= 255;
}

// Classes of the planner, by rows m of the row pair (the longer of Ya, Yb):
//  * m <= 1024: the pipelined packed kernel (score5): one wave of K =
//    max(k5min, ceil(m/64)) rows per lane for m <= 64 * k1max (K up to 16),
//    else W = ceil(m/256) waves of 64*K rows;
//  * m > 1024: the multi-sweep packed kernel (score6), up to the 16-bit bound,
//    in strips of 64*K rows, K = 3 or 4 (tune.k6, 0: per row pair, the K
//    whose four-strip sweeps cover m with fewer idle rows, K = 3 counted at
//    the 0.89 of K = 4's rate per row that one-wave launches measure);
//  * every pair beyond the bound: the int32 kernel.
// A work item's column stream of the multi-sweep kernel stays within
// kScore6MaxCols (the per-workgroup row buffer).
constexpr int kOneWaveShortK = 5;
constexpr int kLoneChunk = 8;  // item size of a stacked-range row pair that did not stack
constexpr int64_t kScore6MaxCols = 16384;

// score8's strips for m rows: ns strips of at most 512 rows, K rows per lane
static int score8_k(int m, int* ns) {
  int s = (m + 511) / 512;
  if (s < 1) s = 1;
  *ns = s;
  const int K = (m + 64 * s - 1) / (64 * s);
  return std::max(kScore8MinK, std::min(kScore8MaxK, K));
}

// A re-based wave's values stay within 16 bits whatever the lengths: in the
// biased form a value changes by -G .. G + smax + 2 ge from one row or column
// to the next (G = go - ge), so the front of a strip of 64K rows spans at most
// up*(64K-1) + dn*63 above and dn*(64K-1) + up*63 below lane 0's first row,
// which the frame keeps within kRebaseAt (checked every 8 steps, and lane 0
// drifting for the 63 steps after its last column); E, F and the diagonal
// term add at most up + G + 255.
static bool rebase_fits(const mst_ctx* ctx, int K) {
  const int64_t G = ctx->go - ctx->ge;
  if (G < 0) return false;
  const int64_t up = G + ctx->smax + 2 * ctx->ge, dn = G;
  const int64_t rows = 64 * K - 1;
  const int64_t drift = (8 + 63) * up;
  const int64_t hi = kRebaseAt + drift + up * rows + dn * 63 + G + 255;
  const int64_t lo = kRebaseAt + drift + dn * rows + up * 63 + up + G;
  return hi <= 32767 && lo <= 32768;
}

// The parts item_part (nw_stream.h) cuts an item's columns into, on the host:
// the most columns one part holds.
static int64_t max_part_cols(const std::vector<int64_t>& soff, int xb0, int xe0, int P) {
  const int64_t c0 = soff[xb0], total = soff[xe0] - c0;
  auto cut = [&](int q) {
    if (q <= 0) return xb0;
    if (q >= P) return xe0;
    const int64_t target2 = 2 * (total * q / P);
    for (int k = xb0; k < xe0; ++k)
      if ((soff[k] - c0) + (soff[k + 1] - c0) >= target2) return k;
    return xe0;
  };
  int64_t best = 0;
  for (int p = 0; p < P; ++p) best = std::max(best, soff[cut(p + 1)] - soff[cut(p)]);
  return best;
}

static int rows_per_lane(int m, int* nstrips) {
  int s = (m + 1023) / 1024;
  if (s < 1) s = 1;
  *nstrips = s;
  int K = (m + 64 * s - 1) / (64 * s);
  return K < 1 ? 1 : K;
}

int build_batched_plan(mst_ctx* ctx, const SeqSet& s, BatchedPlan& P) {
  const int N = s.n;
  auto len = [&](int i) { return (int)(s.h_off[i + 1] - s.h_off[i]); };
  P = BatchedPlan();
  std::vector<int32_t> fb_i, fb_j;  // int32 fallback pairs
  std::vector<int> order(N);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return len(a) < len(b); });
  P.xlist = order;
  const int C = ctx->tune.chunk;
  struct It3 {
    Item3 it;
    int K, W;
    int64_t cost, cells, pairs;
    int2 p2 = make_int2(-1, -1);  // stacked items: the second row pair
  };
  std::vector<It3> v5, v6, v8, v8r;  // v8r: re-based score8 items
  std::vector<int64_t> soff(N + 1, 0);  // code offsets by sorted position
  for (int k = 0; k < N; ++k) soff[k + 1] = soff[k] + len(order[k]);
  int64_t cap8 = 0, cap8r = 0;
  bool lone = false;  // this row pair is a stacked-range pair that did not stack
  for (int k = 0; 2 * k < N; ++k) {
    lone = false;
    const int ya = order[2 * k];
    const bool single = (2 * k + 1 >= N);
    const int yb = single ? -1 : order[2 * k + 1];
    const int m = single ? len(ya) : len(yb);
    if (m == 0) {  // empty rows: score is -ge*n, no DP needed
      for (int t = 2 * k; t < N; ++t) {
        fb_i.push_back(order[t]);
        fb_j.push_back(ya);
        if (!single && t >= 2 * k + 1) {
          fb_i.push_back(order[t]);
          fb_j.push_back(yb);
        }
      }
      continue;
    }
    int nstr;
    int K = rows_per_lane(m, &nstr);
    int W = 1;   // waves per pipeline (score5)
    int kind;    // 5: score5 (one sweep of W waves), 6: the multi-sweep kernel
    if (nstr == 1 && m <= 64 * kOneWaveShortK) {
      K = std::max(K, ctx->tune.k5min);  // one wave, columns streamed
      kind = 5;
    } else if (nstr == 1 && m <= 64 * ctx->tune.k1max) {
      K = (m + 63) / 64;  // one wave of up to k1max rows per lane
      kind = 5;
    } else if (nstr == 1 && (m + 127) / 128 <= ctx->tune.k2max) {
      W = 2;  // two waves of up to k2max rows per lane
      K = (m + 127) / 128;
      kind = 5;
    } else if (nstr == 1) {
      W = (m + 255) / 256;
      K = (m + 64 * W - 1) / (64 * W);
      kind = 5;
    } else if ((m + 255) / 256 <= ctx->tune.wmax) {
      W = (m + 255) / 256;  // one sweep of W waves, 4 rows per lane (m > 1024: K = 4)
      K = (m + 64 * W - 1) / (64 * W);
      kind = 5;
    } else {
      K = ctx->tune.k6;
      if (K == 0) {
        // row capacity of whole four-strip sweeps, over the per-row rate
        auto cap = [&](int k) { return (double)((m + 256 * k - 1) / (256 * k)) * 256 * k; };
        K = cap(3) / 0.89 < cap(4) ? 3 : 4;
      }
      kind = 6;
    }
    // score8 (strips in turn on one wave) for the rows the k8 setting names
    if ((kind == 6 && ctx->tune.k8 >= 1) || (kind == 5 && W >= 2 && ctx->tune.k8 >= 2)) kind = 8;
    int ns8;
    const int K8 = score8_k(m, &ns8);
    if (kind == 8) {
      K = K8;
      W = 1;
    }
    const bool rebase = ctx->tune.rebase && rebase_fits(ctx, K8);
    // padded rows (bound of the 16-bit values)
    const int64_t R = kind == 5 ? (int64_t)W * 64 * K
                    : kind == 8 ? (int64_t)ns8 * 64 * K8
                                : (int64_t)(m + 64 * K - 1) / (64 * K) * 64 * K;
    // rows (ya, yb) against columns s[first..last]; yb < 0: ya alone. The
    // streamed kernels name rows by position (pa, pb), the others by id.
    auto add_cols = [&](int first, int last, int pa, int pb) {
      const int ya_ = order[pa], yb_ = pb >= 0 ? order[pb] : -1;
      const int64_t rows = len(ya_) + (yb_ >= 0 ? len(yb_) : 0);
      std::vector<int> elig, past;  // past: beyond the 16-bit bound, re-based
      for (int t = first; t <= last; ++t) {
        const int x = order[t];
        const int n = len(x);
        if (n > 0 && pk16_bound(ctx, R, n) > 32767 && rebase) {
          past.push_back(t);
        } else if (n == 0 || pk16_bound(ctx, R, n) > 32767) {
          fb_i.push_back(x);
          fb_j.push_back(ya_);
          if (yb_ >= 0) {
            fb_i.push_back(x);
            fb_j.push_back(yb_);
          }
        } else {
          elig.push_back(t);
        }
      }
      // re-based items: up to four consecutive positions, one per wave
      for (size_t c0 = 0; c0 < past.size();) {
        size_t c1 = c0 + 1;
        while (c1 < past.size() && c1 - c0 < (size_t)kScore8Waves && past[c1] == past[c1 - 1] + 1) ++c1;
        Item3 it;
        it.ya = pa;
        it.yb = pb;
        it.xbegin = past[c0];
        it.xcount = (int)(c1 - c0);
        int64_t cost = 0, cells = 0;
        for (size_t q = c0; q < c1; ++q) {
          const int64_t n = len(order[past[q]]);
          cost = std::max(cost, (n + 63) * R);
          cells += n * rows;
          cap8r = std::max(cap8r, n);
        }
        v8r.push_back(It3{it, K8, 1, cost, cells, (int64_t)(c1 - c0) * (yb_ >= 0 ? 2 : 1)});
        c0 = c1;
      }
      {
        // pipelines of three or four waves take their own item size
        int Cw = (kind == 5 && W >= 3 && ctx->tune.chunk_w3 > 0) ? ctx->tune.chunk_w3 : C;
        if (kind == 6 && ctx->tune.chunk6 > 0) Cw = ctx->tune.chunk6;
        if (kind == 8) Cw = ctx->tune.chunk8;
        // the classes of many rows per lane have few row pairs: smaller items
        if (kind == 5 && W <= 2 && W * K >= 9 && ctx->tune.chunk_k9 > 0) Cw = ctx->tune.chunk_k9;
        // a row pair of the stacked range left on its own (its successor too
        // long to stack with) is a class of one row pair: items of a few
        // sequences spread it over the device instead of a few workgroups
        if (lone) Cw = std::min(Cw, kLoneChunk);
        size_t c0 = 0;
        while (c0 < elig.size()) {
          // a chunk is a run of consecutive positions; score6 keeps its column
          // stream within the per-workgroup row buffer
          size_t c1 = c0 + 1;
          int64_t cols = len(order[elig[c0]]);
          while (c1 < elig.size() && (int)(c1 - c0) < Cw && elig[c1] == elig[c1 - 1] + 1 &&
                 (kind != 6 || cols + len(order[elig[c1]]) <= kScore6MaxCols)) {
            cols += len(order[elig[c1]]);
            ++c1;
          }
          Item3 it;
          it.ya = pa;
          it.yb = pb;
          it.xbegin = elig[c0];
          it.xcount = (int)(c1 - c0);
          int64_t cost = 0, cells = 0;
          for (size_t q = c0; q < c1; ++q) {
            cost += (int64_t)(len(order[elig[q]]) + 63) * 64 * K;
            cells += (int64_t)len(order[elig[q]]) * rows;
          }
          const int64_t npairs = (int64_t)(c1 - c0) * (yb_ >= 0 ? 2 : 1);
          if (kind == 5) {
            v5.push_back(It3{it, K, W, cost * W, cells, npairs});
          } else if (kind == 8) {
            const int64_t part = max_part_cols(soff, elig[c0], elig[c1 - 1] + 1, kScore8Waves);
            cap8 = std::max(cap8, part);
            v8.push_back(It3{it, K, 1, (part + 63) * R, cells, npairs});
          } else {
            v6.push_back(It3{it, K, (int)(R / 256), cost * (R / (64 * K)), cells, npairs});
          }
          c0 = c1;
        }
      }
    };
    // Stacking: this row pair and the next, both of the one-wave classes of
    // K = 1..stack rows per lane, share a strip: 32 lanes each, of Ks rows per
    // lane, Ks = ceil(longest of the four rows / 32) (at least k5min), so the
    // rows are padded to a multiple of 32 instead of 64. With a symmetric matrix the stacked items take every column
    // from this pair's ya on: the few pairs both row pairs then meet (each
    // other's rows) are scored twice, to the same value; otherwise over their
    // common columns only (from the next pair's yb on), the rest as before.
    const int Ks = 2 * k + 3 < N ? std::max(ctx->tune.k5min, (len(order[2 * k + 3]) + 31) / 32) : 0;
    const bool stack = ctx->tune.stack && kind == 5 && W == 1 && K >= 1 && K <= ctx->tune.stack &&
                       !single &&
                       len(ya) > 0 && 2 * k + 3 < N &&
                       (len(order[2 * k + 3]) + 63) / 64 <= ctx->tune.stack &&
                       score5_stacked_supported(Ks);
    if (stack) {
      int from = 2 * k;
      if (!ctx->symmetric) {
        add_cols(2 * k, 2 * k, 2 * k, -1);                 // (ya, ya)
        add_cols(2 * k + 1, 2 * k + 2, 2 * k, 2 * k + 1);  // both rows, columns before the next pair's yb
        add_cols(2 * k + 2, 2 * k + 2, 2 * k + 2, -1);     // the next pair's (ya, ya)
        from = 2 * k + 3;
      }
      // row position 2k + r against column position t counts (and, past the
      // 16-bit bound, goes to the int32 kernel) when t >= 2k + r: the other
      // (row, column) meetings are the duplicates
      std::vector<int> elig;
      for (int t = from; t < N; ++t) {
        const int x = order[t];
        const int n = len(x);
        if (n == 0 || pk16_bound(ctx, R, n) > 32767) {
          for (int r = 0; r < 4 && 2 * k + r <= t; ++r) {
            fb_i.push_back(x);
            fb_j.push_back(order[2 * k + r]);
          }
        } else {
          elig.push_back(t);
        }
      }
      auto rows_at = [&](int t, int64_t* nr) {
        int64_t s = 0;
        int c = 0;
        for (int r = 0; r < 4 && 2 * k + r <= t; ++r, ++c) s += len(order[2 * k + r]);
        *nr = c;
        return s;
      };
      for (size_t c0 = 0; c0 < elig.size();) {
        size_t c1 = c0 + 1;
        while (c1 < elig.size() && (int)(c1 - c0) < C && elig[c1] == elig[c1 - 1] + 1) ++c1;
        Item3 it;
        it.ya = 2 * k;
        it.yb = 2 * k + 1;
        it.xbegin = elig[c0];
        it.xcount = (int)(c1 - c0);
        int64_t cost = 0, cells = 0, npairs = 0;
        for (size_t q = c0; q < c1; ++q) {
          int64_t nr;
          const int64_t rl = rows_at(elig[q], &nr);
          cost += (int64_t)(len(order[elig[q]]) + 63) * 64 * Ks;
          cells += (int64_t)len(order[elig[q]]) * rl;
          npairs += nr;
        }
        It3 e{it, Ks, 1, cost, cells, npairs};
        e.p2 = make_int2(2 * k + 2, 2 * k + 3);
        v5.push_back(e);
        c0 = c1;
      }
      ++k;  // the next pair is done
      continue;
    }
    lone = ctx->tune.stack && kind == 5 && W == 1 && K <= ctx->tune.stack;
    if (single) {
      add_cols(2 * k, N - 1, 2 * k, -1);
    } else if (len(ya) == 0) {
      // an empty row sequence has no last row to emit its scores from: its
      // pairs (-ge * n each) take the int32 kernel, yb runs alone
      for (int t = 2 * k; t < N; ++t) {
        fb_i.push_back(order[t]);
        fb_j.push_back(ya);
      }
      add_cols(2 * k + 1, N - 1, 2 * k + 1, -1);
    } else {
      add_cols(2 * k, 2 * k, 2 * k, -1);          // (ya, ya)
      add_cols(2 * k + 1, N - 1, 2 * k, 2 * k + 1);  // every x at or after yb, both rows
    }
  }
  std::stable_sort(v5.begin(), v5.end(), [](const It3& a, const It3& b) {
    if (a.W != b.W) return a.W < b.W;
    if ((a.p2.x >= 0) != (b.p2.x >= 0)) return a.p2.x < 0;
    if (a.K != b.K) return a.K < b.K;
    return a.cost > b.cost;
  });
  std::stable_sort(v6.begin(), v6.end(), [](const It3& a, const It3& b) {
    if (a.K != b.K) return a.K > b.K;
    return a.cost > b.cost;
  });
  for (size_t q = 0; q < v6.size(); ++q) {
    P.items6.push_back(v6[q].it);
    if (P.c6.empty() || P.c6.back().K != v6[q].K) P.c6.push_back(BClass{v6[q].K, (int64_t)q, (int64_t)q});
    P.c6.back().end = q + 1;
    P.c6.back().work += v6[q].cost;
    P.c6.back().cells += v6[q].cells;
    P.c6.back().pairs += v6[q].pairs;
    int64_t cols = 0;
    for (int t = v6[q].it.xbegin; t < v6[q].it.xbegin + v6[q].it.xcount; ++t) cols += len(order[t]);
    P.carry_cols = std::max(P.carry_cols, cols);
  }
  // score8: plain classes by K, then the re-based ones, each longest first
  auto by_cost = [](const It3& a, const It3& b) {
    if (a.K != b.K) return a.K > b.K;
    return a.cost > b.cost;
  };
  std::stable_sort(v8.begin(), v8.end(), by_cost);
  // Re-based items are long (every strip of a row pair on one wave) and run
  // on a graph branch of their own beside the other classes. rebase 1: when
  // there are four items per compute unit, or when the longest item (at
  // ~220 ns per step, as measured with the device shared) ends well within
  // the rest of the pass (total cells at ~8.5 T cells/s): the 4,000-protein
  // set's long pairs ran 206.0 ms re-based against 209.4 ms on the int32
  // kernel, which follows the graph; otherwise the int32 kernel. 2: always.
  int64_t steps_max = 0, cells_all = 0;
  for (const It3& e : v8r) steps_max = std::max(steps_max, e.cost / (64 * (int64_t)e.K));
  for (const auto* v : {&v5, &v6, &v8, &v8r})
    for (const It3& e : *v) cells_all += e.cells;
  const bool rebase_pays = (int64_t)v8r.size() >= 4 * (int64_t)ctx->num_cus ||
                           (double)steps_max * 220e-9 < 0.6 * (double)cells_all / 8.5e12;
  if (ctx->tune.rebase == 1 && !rebase_pays) {
    for (const It3& e : v8r) {
      const int ya_ = order[e.it.ya], yb_ = e.it.yb >= 0 ? order[e.it.yb] : -1;
      for (int t = e.it.xbegin; t < e.it.xbegin + e.it.xcount; ++t) {
        fb_i.push_back(order[t]);
        fb_j.push_back(ya_);
        if (yb_ >= 0) {
          fb_i.push_back(order[t]);
          fb_j.push_back(yb_);
        }
      }
    }
    v8r.clear();
    cap8r = 0;
  }
  std::stable_sort(v8r.begin(), v8r.end(), by_cost);
  for (int r = 0; r < 2; ++r) {
    const std::vector<It3>& v = r ? v8r : v8;
    for (size_t q = 0; q < v.size(); ++q) {
      const int64_t at = (int64_t)P.items8.size();
      P.items8.push_back(v[q].it);
      if (q == 0 || P.c8.back().K != v[q].K) {
        BClass c{v[q].K, at, at};
        c.rebase = r == 1;
        c.cap = r ? cap8r : cap8;
        P.c8.push_back(c);
      }
      P.c8.back().end = at + 1;
      P.c8.back().work += v[q].cost;
      P.c8.back().cells += v[q].cells;
      P.c8.back().pairs += v[q].pairs;
    }
  }
  for (size_t q = 0; q < v5.size(); ++q) {
    P.items5.push_back(v5[q].it);
    P.pair2_5.push_back(v5[q].p2);
    const bool st = v5[q].p2.x >= 0;
    if (P.c5.empty() || P.c5.back().K != v5[q].K || P.c5.back().W != v5[q].W || P.c5.back().stacked != st) {
      BClass c{v5[q].K, (int64_t)q, (int64_t)q + 1};
      c.W = v5[q].W;
      c.stacked = st;
      P.c5.push_back(c);
    } else {
      P.c5.back().end = q + 1;
    }
    P.c5.back().work += v5[q].cost;
    P.c5.back().cells += v5[q].cells;
    P.c5.back().pairs += v5[q].pairs;
  }
  P.fb_i = fb_i;
  P.fb_j = fb_j;
  P.built = true;
  return MST_OK;
}

// defined in capi.hip
int build_plan(mst_ctx* ctx, const SeqSet& s, const int32_t* pi, const int32_t* pj, int64_t np,
               PairPlan& plan);
int run_plan_on(mst_ctx* ctx, const SeqSet& s, const PairPlan& plan, const int2* d_pairs, int form,
                int32_t* d_out_pair, int32_t* d_out_mat, bool timed, int event_base,
                hipStream_t st = nullptr);

int upload_batched(mst_ctx* ctx, const SeqSet& s, int id, BatchedPlan& P) {
  if (P.uploaded) return MST_OK;
  CK(hipMalloc(&P.d_xlist, sizeof(int32_t) * std::max<size_t>(P.xlist.size(), 1)));
  P.ncounters = (int)(P.c5.size() + P.c6.size() + P.c8.size()) + 1;
  const bool streamed = !P.c5.empty() || !P.c6.empty() || !P.c8.empty();
  if (!P.c8.empty()) {
    CK(hipMalloc(&P.d_items8, sizeof(Item3) * P.items8.size()));
    CK(hipMemcpy(P.d_items8, P.items8.data(), sizeof(Item3) * P.items8.size(), hipMemcpyHostToDevice));
    // a row buffer per wave of every workgroup of every class (the classes may
    // run at the same time on parallel branches of the graph), and the frames
    // of the re-based classes' row buffers
    P.grid8 = ctx->num_cus * score8_groups_per_cu();
    int64_t nc = 0, nf = 0;
    P.carry8_off.clear();
    P.frames8_off.clear();
    for (const BClass& c : P.c8) {
      const int64_t waves = (int64_t)std::min<int64_t>(P.grid8, c.end - c.begin) * kScore8Waves;
      P.carry8_off.push_back(nc);
      P.frames8_off.push_back(nf);
      nc += waves * std::max<int64_t>(c.cap, 1);
      if (c.rebase) nf += waves * ((c.cap + 63) / 8 + 2);
    }
    CK(hipMalloc(&P.d_carry8, sizeof(int2) * std::max<int64_t>(nc, 1)));
    CK(hipMalloc(&P.d_frames8, sizeof(int2) * std::max<int64_t>(nf, 1)));
  }
  if (!P.c6.empty()) {
    CK(hipMalloc(&P.d_items6, sizeof(Item3) * P.items6.size()));
    CK(hipMemcpy(P.d_items6, P.items6.data(), sizeof(Item3) * P.items6.size(), hipMemcpyHostToDevice));
    const bool two = ctx->tune.s6p2 != 0;
    P.grid6 = (int)std::max<int64_t>(
        1, std::min<int64_t>((int64_t)ctx->num_cus * (two ? score7_groups_per_cu() : score6_groups_per_cu()),
                             (int64_t)P.items6.size()));
    // a row buffer per pipeline, workgroup and multi-sweep class: the classes
    // may run at the same time on parallel branches of the graph
    CK(hipMalloc(&P.d_carry, sizeof(int2) * (size_t)P.grid6 * (two ? 2 : 1) *
                                 std::max<int64_t>(P.carry_cols, 1) * P.c6.size()));
  }
  if (!P.c5.empty()) {
    CK(hipMalloc(&P.d_items5, sizeof(Item3) * P.items5.size()));
    CK(hipMemcpy(P.d_items5, P.items5.data(), sizeof(Item3) * P.items5.size(), hipMemcpyHostToDevice));
    CK(hipMalloc(&P.d_pair2, sizeof(int2) * P.pair2_5.size()));
    CK(hipMemcpy(P.d_pair2, P.pair2_5.data(), sizeof(int2) * P.pair2_5.size(), hipMemcpyHostToDevice));
  }
  if (streamed) {
    const int N = s.n;
    std::vector<int64_t> soff(N + 1, 0);
    for (int k = 0; k < N; ++k)
      soff[k + 1] = soff[k] + (s.h_off[P.xlist[k] + 1] - s.h_off[P.xlist[k]]);
    CK(hipMalloc(&P.d_soff, sizeof(int64_t) * (N + 1)));
    CK(hipMemcpy(P.d_soff, soff.data(), sizeof(int64_t) * (N + 1), hipMemcpyHostToDevice));
    CK(hipMalloc(&P.d_scodes, std::max<int64_t>(soff[N], 1)));
    CK(hipMalloc(&P.d_soffs8, sizeof(uint32_t) * (soff[N] + 2 * kOffs8Pad)));
    CK(hipMemset(P.d_soffs8, 0, sizeof(uint32_t) * (soff[N] + 2 * kOffs8Pad)));
    P.ncodes = soff[N];
  }
  CK(hipMalloc(&P.d_counters, sizeof(unsigned int) * P.ncounters));
  if (!P.xlist.empty())
    CK(hipMemcpy(P.d_xlist, P.xlist.data(), sizeof(int32_t) * P.xlist.size(), hipMemcpyHostToDevice));
  if (streamed) {
    hipLaunchKernelGGL(gather_codes, dim3(s.n), dim3(256), 0, ctx->stream, s.d_codes, s.d_off,
                       P.d_xlist, P.d_soff, P.d_scodes);
    CK(hipGetLastError());
    if (P.ncodes > 0)
      hipLaunchKernelGGL(codes_to_offsets8, dim3((unsigned)((P.ncodes + 255) / 256)), dim3(256), 0, ctx->stream,
                         P.d_scodes, P.ncodes, P.d_soffs8 + kOffs8Pad);
    CK(hipGetLastError());
    CK(hipStreamSynchronize(ctx->stream));
  }
  PairPlan& fb = ctx->fbplans[id];
  fb = PairPlan();
  if (!P.fb_i.empty()) {
    int rc = build_plan(ctx, s, P.fb_i.data(), P.fb_j.data(), (int64_t)P.fb_i.size(), fb);
    if (rc) return rc;
    CK(hipMalloc(&fb.d_pairs, sizeof(int2) * fb.pairs.size()));
    CK(hipMemcpy(fb.d_pairs, fb.pairs.data(), sizeof(int2) * fb.pairs.size(), hipMemcpyHostToDevice));
  }
  P.uploaded = true;
  return MST_OK;
}

// The int32 launches of the plan's remaining pairs (never captured in the
// graph), on st (ctx->stream when null).
static int enqueue_fallback(mst_ctx* ctx, const SeqSet& s, int id, int32_t* d_out, bool timed,
                            int event_base, hipStream_t st = nullptr) {
  const PairPlan& fb = ctx->fbplans[id];
  if (fb.pairs.empty()) return MST_OK;
  return run_plan_on(ctx, s, fb, fb.d_pairs, MST_FORM_BATCHED, nullptr, d_out, timed, event_base, st);
}

// Enqueues the packed launches of the batched plan on ctx->stream, then the
// int32 ones when with_fallback. With timed, an event pair brackets each
// launch (ctx->launch_events in launch order).
int enqueue_batched(mst_ctx* ctx, const SeqSet& s, int id, BatchedPlan& P, int32_t* d_out,
                    bool timed, bool with_fallback) {
  hipLaunchKernelGGL(zero_u32, dim3((P.ncounters + 63) / 64), dim3(64), 0, ctx->stream, P.d_counters,
                     P.ncounters);
  CK(hipGetLastError());
  int ev = 0, ctr = 0;
  auto args5 = [&](const BClass& c, const Item3* items) {
    Score5Args a;
    a.scodes = P.d_scodes;
    a.soffs8 = P.d_soffs8 ? P.d_soffs8 + kOffs8Pad : nullptr;
    a.soff = P.d_soff;
    a.order = P.d_xlist;
    a.items = items;
    a.item_job = nullptr;
    a.job_out = nullptr;
    a.job_ld = nullptr;
    a.first = c.begin;
    a.last = c.end;
    a.counter = nullptr;  // set by the caller
    a.out_mat = d_out;
    a.nseq = s.n;
    a.subst = ctx->d_S;
    a.go = ctx->go;
    a.ge = ctx->ge;
    a.carry = P.d_carry;
    a.carry_cap = std::max<int64_t>(P.carry_cols, 1);
    return a;
  };
  // Untimed, the streamed classes may run as parallel branches (fork from
  // ctx->stream, join back into it) so that one class's tail, when few of
  // its workgroups are left, overlaps the next class's start. Classes go to
  // the branch with the fewest cells so far, largest first.
  const int nb = timed ? 1 : std::max(1, std::min(4, ctx->tune.branches));
  // Re-based classes are few long items on part of the device: they take a
  // branch of their own, started with the others, so the other classes fill
  // the rest of the device meanwhile (when the branches allow one more).
  bool rb_branch = false;
  for (const BClass& c : P.c8) rb_branch |= c.rebase;
  rb_branch = rb_branch && !timed && nb < 4;
  const int nbt = nb + (rb_branch ? 1 : 0);  // branch streams in all
  std::vector<hipStream_t> br(1, ctx->stream);
  if (nbt > 1) {
    for (int b = 0; b < nbt; ++b) {
      if (!ctx->bstream[b]) CK(hipStreamCreateWithFlags(&ctx->bstream[b], hipStreamNonBlocking));
      if (!ctx->bevent[b]) CK(hipEventCreateWithFlags(&ctx->bevent[b], hipEventDisableTiming));
    }
    if (!ctx->bevent[4]) CK(hipEventCreateWithFlags(&ctx->bevent[4], hipEventDisableTiming));
    CK(hipEventRecord(ctx->bevent[4], ctx->stream));
    br.clear();
    for (int b = 0; b < nbt; ++b) {
      CK(hipStreamWaitEvent(ctx->bstream[b], ctx->bevent[4], 0));
      br.push_back(ctx->bstream[b]);
    }
  }
  struct Launch {
    int kind;  // 5, 6 or 8
    const BClass* c;
    int ctr;
  };
  std::vector<Launch> ls;
  for (const BClass& c : P.c6) ls.push_back(Launch{6, &c, ctr++});
  for (const BClass& c : P.c5) ls.push_back(Launch{5, &c, ctr++});
  for (const BClass& c : P.c8) ls.push_back(Launch{8, &c, ctr++});
  std::vector<int> where(ls.size(), 0);
  if (nbt > 1) {
    std::vector<size_t> ord(ls.size());
    for (size_t k = 0; k < ls.size(); ++k) ord[k] = k;
    std::stable_sort(ord.begin(), ord.end(), [&](size_t x, size_t y) { return ls[x].c->cells > ls[y].c->cells; });
    std::vector<int64_t> load(nb, 0);
    for (size_t k : ord) {
      if (rb_branch && ls[k].kind == 8 && ls[k].c->rebase) {
        where[k] = nb;
        continue;
      }
      const int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      where[k] = b;
      load[b] += ls[k].c->cells;
    }
  }
  for (size_t k = 0; k < ls.size(); ++k) {
    const BClass& c = *ls[k].c;
    hipStream_t st = br[where[k]];
    const int kind = ls[k].kind;
    Score5Args a = args5(c, kind == 6 ? P.d_items6 : (kind == 8 ? P.d_items8 : P.d_items5));
    a.counter = P.d_counters + ls[k].ctr;
    if (kind == 6)
      a.carry += (int64_t)(&c - P.c6.data()) * P.grid6 * (ctx->tune.s6p2 ? 2 : 1) * a.carry_cap;
    if (kind == 8) {
      const size_t ci = &c - P.c8.data();
      a.carry = P.d_carry8 + P.carry8_off[ci];
      a.carry_cap = std::max<int64_t>(c.cap, 1);
      a.frames = P.d_frames8 + P.frames8_off[ci];
      a.frame_cap = (c.cap + 63) / 8 + 2;
    }
    if (timed) CK(hipEventRecord(ctx->launch_events[ev++], st));
    if (kind == 8) {
      const int g = (int)std::max<int64_t>(1, std::min<int64_t>(P.grid8, c.end - c.begin));
      launch_score8(c.K, c.rebase, a, g, st);
    } else if (kind == 6) {
      if (ctx->tune.s6p2)
        launch_score7(c.K, a, P.grid6, st);
      else
        launch_score6(c.K, a, P.grid6, st);
    } else {
      const int per = ctx->tune.wg5 > 0 ? ctx->tune.wg5 : score5_groups_per_cu(c.K, c.W);
      const int64_t g = std::min<int64_t>((int64_t)ctx->num_cus * per, c.end - c.begin);
      if (c.stacked) a.pair2 = P.d_pair2;
      launch_score5(c.K, c.W, a, (int)std::max<int64_t>(g, 1), st, c.stacked);
    }
    CK(hipGetLastError());
    if (timed) CK(hipEventRecord(ctx->launch_events[ev++], st));
  }
  if (nbt > 1) {
    for (int b = 0; b < nbt; ++b) {
      CK(hipEventRecord(ctx->bevent[b], ctx->bstream[b]));
      CK(hipStreamWaitEvent(ctx->stream, ctx->bevent[b], 0));
    }
  }
  if (with_fallback) return enqueue_fallback(ctx, s, id, d_out, timed, ev);
  return MST_OK;
}

int batched_launch_count(mst_ctx* ctx, int id) {
  const BatchedPlan& P = ctx->bplans[id];
  return (int)(P.c6.size() + P.c5.size() + P.c8.size() +
               ctx->fbplans[id].buckets.size());
}

// All pairs of a set in the batched form into d_out (n x n). timed: plain
// enqueue with per-launch events; otherwise the captured graph is replayed.
int run_batched(mst_ctx* ctx, int id, int32_t* d_out, bool timed) {
  auto it = ctx->sets.find(id);
  if (it == ctx->sets.end()) return ctx->fail(MST_E_STATE, "unknown sequence set");
  const SeqSet& s = it->second;
  BatchedPlan& P = ctx->bplans[id];
  if (!P.built) {
    int rc = build_batched_plan(ctx, s, P);
    if (rc) return rc;
  }
  int rc = upload_batched(ctx, s, id, P);
  if (rc) return rc;
  if (timed) {
    const int nl = batched_launch_count(ctx, id);
    while ((int)ctx->launch_events.size() < 2 * nl) {
      hipEvent_t e;
      CK(hipEventCreate(&e));
      ctx->launch_events.push_back(e);
    }
    rc = enqueue_batched(ctx, s, id, P, d_out, true, true);
    if (rc) return rc;
    CK(hipStreamSynchronize(ctx->stream));
    ctx->launch_stats.clear();
    int ev = 0;
    auto rec = [&](int K, int64_t pairs, int64_t cells) -> int {
      float t = 0;
      CK(hipEventElapsedTime(&t, ctx->launch_events[ev], ctx->launch_events[ev + 1]));
      ev += 2;
      ctx->launch_stats.push_back(LaunchStat{K, pairs, cells, t});
      return MST_OK;
    };
    for (const BClass& c : P.c6)
      if ((rc = rec(-2000 - c.K, c.pairs, c.cells))) return rc;
    for (const BClass& c : P.c5)
      if ((rc = rec(c.stacked ? -(3000 + c.K) : -(1000 + 100 * c.W + c.K), c.pairs, c.cells))) return rc;
    for (const BClass& c : P.c8)
      if ((rc = rec(c.rebase ? -(5000 + c.K) : -(4000 + c.K), c.pairs, c.cells))) return rc;
    for (const Bucket& b : ctx->fbplans[id].buckets)
      if ((rc = rec(b.K, b.end - b.begin, b.cells))) return rc;
    return MST_OK;
  }
  if (!P.graph || P.graph_out != d_out) {
    if (P.graph) {
      CK(hipGraphExecDestroy(P.graph));
      P.graph = nullptr;
    }
    hipGraph_t g;
    CK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue_batched(ctx, s, id, P, d_out, false, false);
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    // the branch streams serve the capture only (the graph holds the fork and
    // join): give their hardware queues back for other streams
    for (hipStream_t& b : ctx->bstream)
      if (b) {
        hipStreamDestroy(b);
        b = nullptr;
      }
    if (rc) return rc;
    if (e != hipSuccess) return ctx->fail(MST_E_HIP, "hipStreamEndCapture", hipGetErrorString(e));
    CK(hipGraphInstantiate(&P.graph, g, nullptr, nullptr, 0));
    CK(hipGraphDestroy(g));
    P.graph_out = d_out;
  }
  // (the int32 pairs on a side stream beside the replay measured the same:
  // 268.05 against 268.2 ms; they follow the replay on ctx->stream)
  CK(hipGraphLaunch(P.graph, ctx->stream));
  return enqueue_fallback(ctx, s, id, d_out, false, 0);
}

void free_batched(BatchedPlan& P) {
  if (P.graph) hipGraphExecDestroy(P.graph);
  void* ps[] = {P.d_xlist, P.d_counters, P.d_items5, P.d_scodes, P.d_soff, P.d_items6,
                P.d_carry, P.d_pair2, P.d_items8, P.d_carry8, P.d_frames8, P.d_soffs8};
  for (void* p : ps)
    if (p) hipFree(p);
  P = BatchedPlan();
}

// Tunables of the batched form by name. A change drops every cached batched
// plan (and its graph), so the next call plans with the new values.
static int* tune_slot(mst_ctx* ctx, const char* key) {
  if (!strcmp(key, "chunk")) return &ctx->tune.chunk;
  if (!strcmp(key, "wg5")) return &ctx->tune.wg5;
  if (!strcmp(key, "k5min")) return &ctx->tune.k5min;
  if (!strcmp(key, "chunk_w3")) return &ctx->tune.chunk_w3;
  if (!strcmp(key, "k1max")) return &ctx->tune.k1max;
  if (!strcmp(key, "k2max")) return &ctx->tune.k2max;
  if (!strcmp(key, "branches")) return &ctx->tune.branches;
  if (!strcmp(key, "wmax")) return &ctx->tune.wmax;
  if (!strcmp(key, "k6")) return &ctx->tune.k6;
  if (!strcmp(key, "chunk6")) return &ctx->tune.chunk6;
  if (!strcmp(key, "stack")) return &ctx->tune.stack;
  if (!strcmp(key, "chunk_k9")) return &ctx->tune.chunk_k9;
  if (!strcmp(key, "s6p2")) return &ctx->tune.s6p2;
  if (!strcmp(key, "rebase")) return &ctx->tune.rebase;
  if (!strcmp(key, "k8")) return &ctx->tune.k8;
  if (!strcmp(key, "chunk8")) return &ctx->tune.chunk8;
  return nullptr;
}

extern "C" int mst_tune_set(mst_ctx* ctx, const char* key, int value) {
  if (!ctx || !key) return MST_E_ARG;
  int* slot = tune_slot(ctx, key);
  if (!slot) return ctx->fail(MST_E_ARG, "mst_tune_set", "unknown key");
  if (!strcmp(key, "chunk") && (value < 1 || value > kStreamSeqs))
    return ctx->fail(MST_E_ARG, "mst_tune_set", "chunk out of range 1..64");
  if (!strcmp(key, "k5min") && (value < 1 || value > 3))
    return ctx->fail(MST_E_ARG, "mst_tune_set", "k5min out of range 1..3");
  if (!strcmp(key, "stack") && (value < 0 || value > 4))
    return ctx->fail(MST_E_ARG, "mst_tune_set", "stack out of range 0..4");
  if (!strcmp(key, "s6p2") && value != 0 && value != 1)
    return ctx->fail(MST_E_ARG, "mst_tune_set", "s6p2 is 0 or 1");
  if (!strcmp(key, "rebase") && (value < 0 || value > 2))
    return ctx->fail(MST_E_ARG, "mst_tune_set", "rebase is 0, 1 or 2");
  if (!strcmp(key, "k8") && (value < 0 || value > 2))
    return ctx->fail(MST_E_ARG, "mst_tune_set", "k8 is 0, 1 or 2");
  if (!strcmp(key, "chunk8") && (value < 1 || value > kStreamSeqs))
    return ctx->fail(MST_E_ARG, "mst_tune_set", "chunk8 out of range 1..64");
  if (!strcmp(key, "chunk_k9") && (value < 0 || value > kStreamSeqs))
    return ctx->fail(MST_E_ARG, "mst_tune_set", "chunk_k9 out of range 0..64");
  if (!strcmp(key, "chunk6") && (value < 0 || value > kStreamSeqs))
    return ctx->fail(MST_E_ARG, "mst_tune_set", "chunk6 out of range 0..64");
  if (!strcmp(key, "chunk_w3") && (value < 0 || value > kStreamSeqs))
    return ctx->fail(MST_E_ARG, "mst_tune_set", "chunk_w3 out of range 0..64");
  if (!strcmp(key, "branches") && (value < 1 || value > 4))
    return ctx->fail(MST_E_ARG, "mst_tune_set", "branches out of range 1..4");
  if (!strcmp(key, "k6") && value != 0 && value != 3 && value != 4)
    return ctx->fail(MST_E_ARG, "mst_tune_set", "k6 is 0, 3 or 4");
  if (!strcmp(key, "wmax") && (value < 4 || value > 8))
    return ctx->fail(MST_E_ARG, "mst_tune_set", "wmax out of range 4..8");
  if (!strcmp(key, "k2max") && (value < 4 || value > 8))
    return ctx->fail(MST_E_ARG, "mst_tune_set", "k2max out of range 4..8");
  if (!strcmp(key, "k1max") && (value < 5 || value > 16))
    return ctx->fail(MST_E_ARG, "mst_tune_set", "k1max out of range 5..16");
  if (strcmp(key, "chunk") && strcmp(key, "chunk_w3") && strcmp(key, "chunk6") &&
      strcmp(key, "chunk_k9") && strcmp(key, "chunk8") && (value < 0 || value > 64))
    return ctx->fail(MST_E_ARG, "mst_tune_set", "workgroups per CU out of range 0..64");
  if (*slot == value) return MST_OK;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  *slot = value;
  for (auto& kv : ctx->bplans) free_batched(kv.second);
  ctx->bplans.clear();
  for (auto& kv : ctx->fbplans)
    if (kv.second.d_pairs) hipFree(kv.second.d_pairs);
  ctx->fbplans.clear();
  for (auto it = ctx->prepared.begin(); it != ctx->prepared.end();)
    it = (it->second == MST_FORM_BATCHED) ? ctx->prepared.erase(it) : std::next(it);
  return MST_OK;
}

extern "C" int mst_tune_get(mst_ctx* ctx, const char* key, int* value) {
  if (!ctx || !key || !value) return MST_E_ARG;
  int* slot = tune_slot(ctx, key);
  if (!slot) return ctx->fail(MST_E_ARG, "mst_tune_get", "unknown key");
  *value = *slot;
  return MST_OK;
}

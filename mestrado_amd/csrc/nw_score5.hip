// Batched all-pairs scores (R1) in packed 16-bit arithmetic, pipelined over
// the waves of a workgroup.
//
// The cell update: pair (X, Ya) in the
// low and (X, Yb) in the high half of every VGPR, 6 packed ops per two cells
// plus one v_perm_b32 to widen the profile bytes. What differs is how the rows
// and columns are laid out (device parts in nw_stream.h):
//
//  * rows: a pair of row sequences with m rows is cut into W strips of 64*K
//    rows (K = 1..16 rows per lane for one wave, 3..8 for two, 3..4 for three
//    or four, 4 for five to eight); wave j of a
//    pipeline owns strip j. Lane 63 of wave j writes F and H of its last row for
//    every column into an LDS ring; lane 0 of wave j+1 reads them one column
//    behind. Progress counters in LDS (columns produced, columns consumed)
//    order the ring; a producer never runs more than RING5 columns ahead.
//  * columns: the column sequences of a work item (a run of consecutive
//    positions in length order, whose codes are contiguous in `scodes`) form
//    one stream per pipeline. A lane resets its rows to the column -1 boundary
//    (all zero in the biased form) when its column reaches the first column
//    of the next sequence, after writing the scores of the one it finished, so
//    the 63-step skew of the wavefront is paid once per stream, not per pair.
//
// A workgroup holds P pipelines (P*W waves) that share one LDS profile of
// (Ya, Yb) (P = 4 for W = 1, 2 for W = 2, else 1; two pipelines of three or
// four waves measured slower than one); pipeline p takes the p-th of P
// column-balanced parts of the item.
#include "nw_stream.h"

namespace mst {

template <int W>
struct Pipes {
  static constexpr int P = W == 1 ? 4 : (W == 2 ? 2 : 1);
};

// One-wave pipelines of up to 12 rows per lane keep the profile in the wide
// form (a dword per row, no unpacking per cell): its LDS, the four pipelines'
// shared profile, is 49 KB up to 8 rows per lane (three workgroups per CU) and
// 74 KB for 9..12 (12 dwords per lane, two workgroups per CU).
constexpr int kWideMaxK = 12;
template <int K, int W>
struct Wide {
  static constexpr bool v = W == 1 && K <= kWideMaxK;
};

// ST (stacked, one wave only): an item holds two row pairs of up to 32*K rows
// each, the first on lanes 0..31, the second (a.pair2) on lanes 32..63; lane 32
// starts from the row -1 boundary instead of lane 31's last row.
template <int K, int W, bool ST = false>
__global__ void __launch_bounds__(MST_WAVE * W * Pipes<W>::P) score5_kernel(Score5Args a) {
  static_assert(!ST || W == 1, "stacked row pairs take one wave");
  constexpr int P = Pipes<W>::P;
  constexpr int NW = W * P;
  constexpr bool WIDE = Wide<K, W>::v;
  constexpr int D = KDW<K, WIDE>::v;
  constexpr int NR = W > 1 ? P * (W - 1) : 1;
  constexpr int STRIP = MST_LETTERS * MST_WAVE * D * 4;  // profile bytes of one strip
  __shared__ int8_t S_lds[MST_LETTERS * MST_LETTERS];
  __shared__ __attribute__((aligned(16))) uint8_t prof[W * STRIP];
  __shared__ int2 ring[NR * RING5];
  __shared__ int s_prog[NW], s_cons[NW];
  __shared__ int s_item;
  const int lane = lane_id();
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave index (uniform)
  const int p = wv / W, j = wv % W;
  for (int i = threadIdx.x; i < MST_LETTERS * MST_LETTERS; i += MST_WAVE * NW) S_lds[i] = a.subst[i];
  const int G = a.go - a.ge;
  const s2 G2 = {(short)G, (short)G};
  const int bonus = a.go + a.ge;
  const uint8_t* prof_lane = prof + j * STRIP + lane * D * 4;
  int2* ring_in = (W > 1 && j > 0) ? ring + (p * (W - 1) + j - 1) * RING5 : nullptr;
  int2* ring_out = (W > 1 && j < W - 1) ? ring + (p * (W - 1) + j) * RING5 : nullptr;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(a.counter, 1u);
    __syncthreads();
    const int64_t it = (int64_t)s_item + a.first;
    if (it >= a.last) break;
    const Item3 item = a.items[it];
    Stream S = item_stream(item, it, a);
    Stream S2 = S;  // stacked: the second row pair (uniform)
    Stream Sp = S;  // the lane's row pair, for its profile rows
    int rs = 0;     // first row of the lane's row pair in the strip
    if (ST) {
      Item3 i2 = item;
      const int2 q = a.pair2[it];
      i2.ya = q.x;
      i2.yb = q.y;
      S2 = item_stream(i2, it, a);
      if (lane >= MST_WAVE / 2) {
        Sp = S2;
        rs = (MST_WAVE / 2) * K;
      }
    }
    // profile of every strip: wave wv fills residues wv, wv+NW, ...
    for (int js = 0; js < W; ++js)
      build_strip_profile<K, WIDE>(prof + js * STRIP, S_lds, Sp, a, js * MST_WAVE * K - rs, wv, NW, bonus);
    if (threadIdx.x < NW) {
      s_prog[threadIdx.x] = 0;
      s_cons[threadIdx.x] = 0;
    }
    __syncthreads();
    // pipeline p's part of the item: positions [kb, ke), balanced by columns
    const int xb0 = item.xbegin, xe0 = item.xbegin + item.xcount;
    int kb = xb0, ke = xe0;
    if (P > 1) item_part(a, xb0, xe0, p, P, &kb, &ke);
    if (kb < ke) {
      S.base = a.soff[kb];
      S.ncol = (int)(a.soff[ke] - S.base);
      const Owners ow = make_owners<K>(S, j * MST_WAVE * K, ST ? &S2 : nullptr);
      const Link in{ring_in, 0, RING5 - 1}, out{ring_out, 0, RING5 - 1};
      int* prog = s_prog + p * W;
      int* cons = s_cons + p * W;
      if (W == 1) {
        sweep5<K, true, false, WIDE, ST>(kb, ke, ow, S, prof_lane, G2, in, 0, nullptr, nullptr, out, 0, nullptr,
                                         nullptr, 0, 1, a);
      } else if (j == 0) {
        sweep5<K, true, true, WIDE>(kb, ke, ow, S, prof_lane, G2, in, 0, nullptr, nullptr, out, 0, prog + j,
                                    cons + j + 1, RING5, 1, a);
      } else if (j == W - 1) {
        sweep5<K, false, false, WIDE>(kb, ke, ow, S, prof_lane, G2, in, 0, prog + j - 1, cons + j, out, 0,
                                      nullptr, nullptr, 0, 1, a);
      } else {
        sweep5<K, false, true, WIDE>(kb, ke, ow, S, prof_lane, G2, in, 0, prog + j - 1, cons + j, out, 0,
                                     prog + j, cons + j + 1, RING5, 1, a);
      }
    }
    __syncthreads();  // profile, rings and counters are rebuilt for the next item
  }
}

bool score5_supported(int K, int W) {
  if (W == 1) return K >= 1 && K <= 16;
  if (W == 2) return K >= 3 && K <= 8;
  if (W >= 5) return W <= 8 && K == 4;
  return W >= 3 && W <= 4 && K >= 3 && K <= 4;
}

static int pipes(int W) { return W == 1 ? 4 : (W == 2 ? 2 : 1); }

int score5_block_threads(int W) { return MST_WAVE * W * pipes(W); }

// LDS of one workgroup (profile, rings, counters, matrix) -> workgroups per CU
int score5_groups_per_cu(int K, int W) {
  const int P = pipes(W);
  const bool wide = W == 1 && K <= kWideMaxK;
  const int D = wide ? (K <= 1 ? 1 : (K <= 2 ? 2 : (K <= 4 ? 4 : (K <= 8 ? 8 : (K <= 12 ? 12 : 16)))))
                     : (K <= 2 ? 1 : (K <= 4 ? 2 : (K <= 8 ? 4 : 8)));
  const int nr = W > 1 ? P * (W - 1) : 1;
  const int lds = W * MST_LETTERS * MST_WAVE * D * 4 + nr * RING5 * 8 + 8 * W * P + 4 +
                  MST_LETTERS * MST_LETTERS;
  int g = (160 * 1024) / (lds + 512);
  const int waves = W * P;
  if (g * waves > 32) g = 32 / waves;
  return g < 1 ? 1 : g;
}

bool score5_stacked_supported(int K) { return K >= 1 && K <= 8; }

void launch_score5(int K, int W, const Score5Args& a, int grid, hipStream_t st, bool stacked) {
  if (stacked) {
    if (W != 1) return;
#define MST_CASE_ST(k)                                                                                  \
  if (K == k) {                                                                                       \
    hipLaunchKernelGGL((score5_kernel<k, 1, true>), dim3(grid), dim3(score5_block_threads(1)), 0, st, a); \
    return;                                                                                           \
  }
    MST_CASE_ST(1)
    MST_CASE_ST(2)
    MST_CASE_ST(3)
    MST_CASE_ST(4)
    MST_CASE_ST(5)
    MST_CASE_ST(6)
    MST_CASE_ST(7)
    MST_CASE_ST(8)
#undef MST_CASE_ST
    return;
  }
#define MST_CASE(k, w)                                                                          \
  if (K == k && W == w) {                                                                       \
    hipLaunchKernelGGL((score5_kernel<k, w>), dim3(grid), dim3(score5_block_threads(w)), 0, st, \
                       a);                                                                      \
    return;                                                                                     \
  }
  MST_CASE(1, 1)
  MST_CASE(2, 1)
  MST_CASE(3, 1)
  MST_CASE(4, 1)
  MST_CASE(5, 1)
  MST_CASE(6, 1)
  MST_CASE(7, 1)
  MST_CASE(8, 1)
  MST_CASE(9, 1)
  MST_CASE(10, 1)
  MST_CASE(11, 1)
  MST_CASE(12, 1)
  MST_CASE(13, 1)
  MST_CASE(14, 1)
  MST_CASE(15, 1)
  MST_CASE(16, 1)
  MST_CASE(3, 2)
  MST_CASE(4, 2)
  MST_CASE(5, 2)
  MST_CASE(6, 2)
  MST_CASE(7, 2)
  MST_CASE(8, 2)
  MST_CASE(3, 3)
  MST_CASE(4, 3)
  MST_CASE(3, 4)
  MST_CASE(4, 4)
  MST_CASE(4, 5)
  MST_CASE(4, 6)
  MST_CASE(4, 7)
  MST_CASE(4, 8)
#undef MST_CASE
}

}  // namespace mst

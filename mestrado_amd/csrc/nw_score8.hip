// Batched all-pairs scores (R1) in packed 16-bit arithmetic for long row
// pairs: every strip of a row pair in turn on the same wave.
//
// A workgroup of four waves takes one item (a row pair and a run of column
// sequences). The rows are cut into ns strips of 64*K rows; the four waves
// build strip s's profile together (a dword per row and residue, the wide
// form), then wave p sweeps its part of the item's columns for strip s,
// reading the row above from its own row buffer in HBM (left by its own sweep
// of strip s-1) and writing its strip's last row into the same buffer, 63
// columns behind the column it reads, so one buffer serves both. A workgroup
// barrier separates the strips (the profile is rebuilt); no wave waits on
// another inside a strip, and every loop is bounded by the item's rows and
// columns.
//
// Plain items stream several column sequences per wave (the parts balanced by
// columns) and stay within the 16-bit bound the planner checks. Re-based
// items take one column sequence per wave and any length: the lanes' values
// are kept in a moving wave-uniform frame (nw_stream.h, Rebase), the frames
// of the row buffer in `frames` (one int2 per block of 8 written columns).
#include "nw_stream.h"

namespace mst {

constexpr int P8 = kScore8Waves;

// Three waves per SIMD (the LDS allows three workgroups per CU): the hint
// holds the registers to 168 (K = 8 spills a few, outside the cell updates),
// which measured faster than two waves at up to 193 registers (the
// 4,000-protein classes 12.5 / 13.7 ms against 14.0 / 15.8 ms at K = 8 / 7;
// 256 x 8k-20k 840 against 991 ms).
template <int K, bool REBASE>
__global__ void __launch_bounds__(MST_WAVE * P8) __attribute__((amdgpu_waves_per_eu(3)))
score8_kernel(Score5Args a) {
  constexpr bool WIDE = true;
  constexpr int D = KDW<K, WIDE>::v;
  constexpr int STRIP = MST_LETTERS * MST_WAVE * D * 4;
  constexpr int ROWS = MST_WAVE * K;
  __shared__ int8_t S_lds[MST_LETTERS * MST_LETTERS];
  __shared__ __attribute__((aligned(16))) uint8_t prof[STRIP];
  __shared__ int s_full, s_sink;
  __shared__ int s_item;
  const int lane = lane_id();
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave index (uniform)
  for (int i = threadIdx.x; i < MST_LETTERS * MST_LETTERS; i += MST_WAVE * P8) S_lds[i] = a.subst[i];
  if (threadIdx.x == 0) {
    s_full = 1 << 30;  // the row buffer is read a strip after it was written: always complete
    s_sink = 0;
  }
  const int G = a.go - a.ge;
  const s2 G2 = {(short)G, (short)G};
  const int bonus = a.go + a.ge;
  const uint8_t* prof_lane = prof + lane * D * 4;
  const int64_t wslot = (int64_t)blockIdx.x * P8 + wv;
  int2* buf = a.carry + wslot * a.carry_cap;
  int2* fr = REBASE ? a.frames + wslot * a.frame_cap : nullptr;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(a.counter, 1u);
    __syncthreads();
    const int64_t it = (int64_t)s_item + a.first;
    if (it >= a.last) break;
    const Item3 item = a.items[it];
    Stream S = item_stream(item, it, a);
    const int m = S.ma > S.mb ? S.ma : S.mb;
    const int ns = (m + ROWS - 1) / ROWS;
    // wave wv's columns: positions [kb, ke)
    const int xb0 = item.xbegin, xe0 = item.xbegin + item.xcount;
    int kb, ke;
    if (REBASE) {
      kb = xb0 + wv < xe0 ? xb0 + wv : xe0;
      ke = xb0 + wv < xe0 ? kb + 1 : xe0;
    } else {
      item_part(a, xb0, xe0, wv, P8, &kb, &ke);
    }
    if (kb < ke) {
      S.base = a.soff[kb];
      S.ncol = (int)(a.soff[ke] - S.base);
    }
    const Link io{buf, 0, -1};
    const Rebase rb{fr, fr};
    for (int s = 0; s < ns; ++s) {
      // every wave has finished strip s-1 (the previous barrier): rebuild
      build_strip_profile<K, WIDE>(prof, S_lds, S, a, s * ROWS, wv, P8, bonus);
      __syncthreads();
      if (kb < ke) {
        const Owners ow = make_owners<K>(S, s * ROWS, nullptr);
        const bool first = (s == 0), last = (s == ns - 1);
        if (first) {
          if (last)
            sweep5<K, true, false, WIDE, false, REBASE>(kb, ke, ow, S, prof_lane, G2, io, 0, &s_full, &s_sink,
                                                        io, 0, &s_sink, nullptr, 0, 16, a, rb);
          else
            sweep5<K, true, true, WIDE, false, REBASE>(kb, ke, ow, S, prof_lane, G2, io, 0, &s_full, &s_sink,
                                                       io, 0, &s_sink, nullptr, 0, 16, a, rb);
        } else {
          if (last)
            sweep5<K, false, false, WIDE, false, REBASE>(kb, ke, ow, S, prof_lane, G2, io, 0, &s_full,
                                                         &s_sink, io, 0, &s_sink, nullptr, 0, 16, a, rb);
          else
            sweep5<K, false, true, WIDE, false, REBASE>(kb, ke, ow, S, prof_lane, G2, io, 0, &s_full,
                                                        &s_sink, io, 0, &s_sink, nullptr, 0, 16, a, rb);
        }
      }
      __syncthreads();  // the profile and the row buffers are reused by the next strip
    }
    __syncthreads();  // s_item is rewritten for the next item
  }
}

int score8_block_threads() { return MST_WAVE * P8; }

// LDS of one workgroup: the wide profile of 8 dwords per lane and residue
// (K = 5..8; K = 4 takes 4), the matrix and a few words
int score8_groups_per_cu() {
  const int lds = MST_LETTERS * MST_WAVE * 8 * 4 + MST_LETTERS * MST_LETTERS + 16;
  int g = (160 * 1024) / (lds + 512);
  if (g * P8 > 32) g = 32 / P8;
  return g < 1 ? 1 : g;
}

void launch_score8(int K, bool rebase, const Score5Args& a, int grid, hipStream_t st) {
#define MST_CASE8(k)                                                                                     \
  if (K == k) {                                                                                         \
    if (rebase)                                                                                         \
      hipLaunchKernelGGL((score8_kernel<k, true>), dim3(grid), dim3(score8_block_threads()), 0, st, a);  \
    else                                                                                                \
      hipLaunchKernelGGL((score8_kernel<k, false>), dim3(grid), dim3(score8_block_threads()), 0, st, a); \
    return;                                                                                             \
  }
  MST_CASE8(4)
  MST_CASE8(5)
  MST_CASE8(6)
  MST_CASE8(7)
  MST_CASE8(8)
#undef MST_CASE8
}

}  // namespace mst

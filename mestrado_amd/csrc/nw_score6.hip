// Batched all-pairs scores (R1) in packed 16-bit arithmetic for row pairs of
// any length the 16-bit bound admits (in practice up to ~2,100 residues with
// BLOSUM62 and go = 10, ge = 2): the streamed form of nw_score5.hip with more
// strips than waves.
//
// A workgroup of four waves takes one item. The rows are cut into NS strips of
// 64*K rows (K = 3 or 4 per lane, chosen per row pair by the planner: the one
// whose sweeps leave fewer idle rows). Wave j sweeps the item's column stream once for
// each of its strips j, j+4, j+8, ...: in sweep q its strip is s = j + 4q.
// Strip s reads the row above it from strip s-1: from wave j-1 through an LDS
// ring in the same sweep, or, for wave 0, from wave 3's previous sweep through
// a full-length row buffer in HBM (`carry`, one per workgroup). Progress
// counters count columns in global units q*ncol + column, so the rings and
// counters run on across sweeps without a reset.
//
// Every wait ends: a wave waits only for data from an earlier strip (wave j-1
// in the same sweep, or wave 3 one sweep earlier) and for a later strip to have
// read a ring slot; the later strip never waits on the writer's next columns,
// and wave 3 never waits for the carry to be read, because wave 0 reads carry
// column c of a sweep before the data of that column can reach wave 3 in the
// same sweep. Each wave keeps its own LDS profile slot and rebuilds it at the
// start of each sweep.
#include "nw_stream.h"

namespace mst {

constexpr int W6 = 4;

template <int K>
__global__ void __launch_bounds__(MST_WAVE * W6) score6_kernel(Score5Args a) {
  constexpr int D = KD<K>::v;
  constexpr int STRIP = MST_LETTERS * MST_WAVE * D * 4;
  constexpr int ROWS = MST_WAVE * K;  // rows per strip
  __shared__ int8_t S_lds[MST_LETTERS * MST_LETTERS];
  __shared__ __attribute__((aligned(16))) uint8_t prof[W6 * STRIP];
  __shared__ int2 ring[(W6 - 1) * RING5];
  __shared__ int s_prog[W6 + 1], s_cons[W6 + 1];  // [W6]: the carry between sweeps
  __shared__ int s_item;
  const int lane = lane_id();
  const int j = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave index (uniform)
  for (int i = threadIdx.x; i < MST_LETTERS * MST_LETTERS; i += MST_WAVE * W6) S_lds[i] = a.subst[i];
  const int G = a.go - a.ge;
  const s2 G2 = {(short)G, (short)G};
  const int bonus = a.go + a.ge;
  uint8_t* slot = prof + j * STRIP;
  const uint8_t* prof_lane = slot + lane * D * 4;
  int2* carry = a.carry + (int64_t)blockIdx.x * a.carry_cap;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(a.counter, 1u);
    if (threadIdx.x < W6 + 1) {
      s_prog[threadIdx.x] = 0;
      s_cons[threadIdx.x] = 0;
    }
    __syncthreads();
    const int64_t it = (int64_t)s_item + a.first;
    if (it >= a.last) break;
    const Item3 item = a.items[it];
    Stream S = item_stream(item, it, a);
    const int m = S.ma > S.mb ? S.ma : S.mb;
    const int ns = (m + ROWS - 1) / ROWS;
    S.base = a.soff[item.xbegin];
    S.ncol = (int)(a.soff[item.xbegin + item.xcount] - S.base);
    const int ncol = S.ncol;
    for (int q = 0;; ++q) {
      const int s = j + q * W6;
      if (s >= ns) break;
      // this wave's previous sweep is over and no other wave reads its slot
      __builtin_amdgcn_wave_barrier();
      build_strip_profile<K>(slot, S_lds, S, a, s * ROWS, 0, 1, bonus);
      __builtin_amdgcn_wave_barrier();
      const int g = q * ncol;
      const bool first = (s == 0), last = (s == ns - 1);
      Link in, out;
      int gin;
      const int* prog_in;
      int* cons_out;
      if (j > 0) {
        in = Link{ring + (j - 1) * RING5, g, RING5 - 1};
        gin = g;
        prog_in = s_prog + j - 1;
        cons_out = s_cons + j;
      } else {
        in = Link{carry, 0, -1};
        gin = g - ncol;
        prog_in = s_prog + W6;
        cons_out = s_cons + W6;
      }
      int* prog_out;
      const int* cons_in;
      int span, pub;
      if (j < W6 - 1) {
        out = Link{ring + j * RING5, g, RING5 - 1};
        prog_out = s_prog + j;
        cons_in = s_cons + j + 1;
        span = RING5;
        pub = 1;
      } else {
        out = Link{carry, 0, -1};
        prog_out = s_prog + W6;
        cons_in = nullptr;
        span = 0;
        pub = 16;
      }
      const Owners ow = make_owners<K>(S, s * ROWS, nullptr);
      const int kb = item.xbegin, ke = item.xbegin + item.xcount;
      if (first) {
        if (last)
          sweep5<K, true, false>(kb, ke, ow, S, prof_lane, G2, in, gin, prog_in, cons_out, out,
                                 g, prog_out, cons_in, span, pub, a);
        else
          sweep5<K, true, true>(kb, ke, ow, S, prof_lane, G2, in, gin, prog_in, cons_out, out,
                                g, prog_out, cons_in, span, pub, a);
      } else {
        if (last)
          sweep5<K, false, false>(kb, ke, ow, S, prof_lane, G2, in, gin, prog_in, cons_out,
                                  out, g, prog_out, cons_in, span, pub, a);
        else
          sweep5<K, false, true>(kb, ke, ow, S, prof_lane, G2, in, gin, prog_in, cons_out,
                                 out, g, prog_out, cons_in, span, pub, a);
      }
    }
    __syncthreads();  // rings, counters and carry are reused by the next item
  }
}

int score6_block_threads() { return MST_WAVE * W6; }

int score6_groups_per_cu() {
  const int D = KD<kScore6K>::v;
  const int lds = W6 * MST_LETTERS * MST_WAVE * D * 4 + (W6 - 1) * RING5 * 8 + 8 * (W6 + 1) + 4 +
                  MST_LETTERS * MST_LETTERS;
  int g = (160 * 1024) / (lds + 512);
  if (g * W6 > 32) g = 32 / W6;
  return g < 1 ? 1 : g;
}

// K = 3 and 4 have the same profile size (KD = 2), so the same LDS and grid
void launch_score6(int K, const Score5Args& a, int grid, hipStream_t st) {
  if (K == 3)
    hipLaunchKernelGGL(score6_kernel<3>, dim3(grid), dim3(score6_block_threads()), 0, st, a);
  else
    hipLaunchKernelGGL(score6_kernel<4>, dim3(grid), dim3(score6_block_threads()), 0, st, a);
}

// Two pipelines of four waves per workgroup sharing the strips' profiles, in
// step per sweep: every sweep starts and ends with a workgroup barrier, the
// four waves of each pipeline sweep their part of the item's columns (split
// by columns as in score5) like a one-sweep pipeline of four strips, and the
// last strip of a sweep leaves its row in the pipeline's row buffer (`carry`,
// two per workgroup), which the next sweep's first strip reads after the
// barrier. Inside a sweep a wave waits only on its own pipeline's rings, as
// in score5; waves without a strip in the last sweep still take its barriers.
// LDS: the four profile slots serve eight waves.
constexpr int P7 = 2;

template <int K>
__global__ void __launch_bounds__(MST_WAVE * W6 * P7) score7_kernel(Score5Args a) {
  constexpr int D = KD<K>::v;
  constexpr int STRIP = MST_LETTERS * MST_WAVE * D * 4;
  constexpr int ROWS = MST_WAVE * K;
  constexpr int NW = W6 * P7;
  __shared__ int8_t S_lds[MST_LETTERS * MST_LETTERS];
  __shared__ __attribute__((aligned(16))) uint8_t prof[W6 * STRIP];
  __shared__ int2 ring[P7 * (W6 - 1) * RING5];
  __shared__ int s_prog[NW], s_cons[NW];
  __shared__ int s_full, s_sink;
  __shared__ int s_item;
  const int lane = lane_id();
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int p = wv / W6, j = wv % W6;
  for (int i = threadIdx.x; i < MST_LETTERS * MST_LETTERS; i += MST_WAVE * NW) S_lds[i] = a.subst[i];
  if (threadIdx.x == 0) {
    s_full = 1 << 30;  // the row buffer read after a barrier: always complete
    s_sink = 0;
  }
  const int G = a.go - a.ge;
  const s2 G2 = {(short)G, (short)G};
  const int bonus = a.go + a.ge;
  uint8_t* slot = prof + j * STRIP;
  const uint8_t* prof_lane = slot + lane * D * 4;
  int2* carry = a.carry + ((int64_t)blockIdx.x * P7 + p) * a.carry_cap;
  int2* ring_in = j > 0 ? ring + (p * (W6 - 1) + j - 1) * RING5 : nullptr;
  int2* ring_out = j < W6 - 1 ? ring + (p * (W6 - 1) + j) * RING5 : nullptr;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(a.counter, 1u);
    __syncthreads();
    const int64_t it = (int64_t)s_item + a.first;
    if (it >= a.last) break;
    const Item3 item = a.items[it];
    Stream S = item_stream(item, it, a);
    const int m = S.ma > S.mb ? S.ma : S.mb;
    const int ns = (m + ROWS - 1) / ROWS;
    const int nsweeps = (ns + W6 - 1) / W6;
    // pipeline p's part of the item: positions [kb, ke), balanced by columns
    const int xb0 = item.xbegin, xe0 = item.xbegin + item.xcount;
    int kb = xb0, ke = xe0;
    item_part(a, xb0, xe0, p, P7, &kb, &ke);
    for (int q = 0; q < nsweeps; ++q) {
      const int s = j + q * W6;
      // slot j: strip s, its residues split between the two pipelines' wave j
      if (s < ns) build_strip_profile<K>(slot, S_lds, S, a, s * ROWS, p, P7, bonus);
      if (threadIdx.x < NW) {
        s_prog[threadIdx.x] = 0;
        s_cons[threadIdx.x] = 0;
      }
      __syncthreads();
      if (s < ns && kb < ke) {
        S.base = a.soff[kb];
        S.ncol = (int)(a.soff[ke] - S.base);
        const bool first = (s == 0), last = (s == ns - 1);
        Link in, out;
        const int* prog_in;
        int* cons_out;
        if (j > 0) {
          in = Link{ring_in, 0, RING5 - 1};
          prog_in = s_prog + p * W6 + j - 1;
          cons_out = s_cons + p * W6 + j;
        } else {
          in = Link{carry, 0, -1};
          prog_in = &s_full;
          cons_out = &s_sink;
        }
        int* prog_out;
        const int* cons_in;
        int span, pub;
        if (j < W6 - 1) {
          out = Link{ring_out, 0, RING5 - 1};
          prog_out = s_prog + p * W6 + j;
          cons_in = s_cons + p * W6 + j + 1;
          span = RING5;
          pub = 1;
        } else {
          out = Link{carry, 0, -1};
          prog_out = &s_sink;
          cons_in = nullptr;
          span = 0;
          pub = 16;
        }
        const Owners ow = make_owners<K>(S, s * ROWS, nullptr);
        if (first) {
          if (last)
            sweep5<K, true, false>(kb, ke, ow, S, prof_lane, G2, in, 0, prog_in, cons_out, out, 0, prog_out,
                                   cons_in, span, pub, a);
          else
            sweep5<K, true, true>(kb, ke, ow, S, prof_lane, G2, in, 0, prog_in, cons_out, out, 0, prog_out,
                                  cons_in, span, pub, a);
        } else {
          if (last)
            sweep5<K, false, false>(kb, ke, ow, S, prof_lane, G2, in, 0, prog_in, cons_out, out, 0, prog_out,
                                    cons_in, span, pub, a);
          else
            sweep5<K, false, true>(kb, ke, ow, S, prof_lane, G2, in, 0, prog_in, cons_out, out, 0, prog_out,
                                   cons_in, span, pub, a);
        }
      }
      __syncthreads();  // slots, rings and row buffers are reused by the next sweep
    }
  }
}

int score7_block_threads() { return MST_WAVE * W6 * P7; }

int score7_groups_per_cu() {
  const int D = KD<kScore6K>::v;
  const int lds = W6 * MST_LETTERS * MST_WAVE * D * 4 + P7 * (W6 - 1) * RING5 * 8 + 8 * W6 * P7 + 12 +
                  MST_LETTERS * MST_LETTERS;
  int g = (160 * 1024) / (lds + 512);
  if (g * W6 * P7 > 32) g = 32 / (W6 * P7);
  return g < 1 ? 1 : g;
}

void launch_score7(int K, const Score5Args& a, int grid, hipStream_t st) {
  if (K == 3)
    hipLaunchKernelGGL(score7_kernel<3>, dim3(grid), dim3(score7_block_threads()), 0, st, a);
  else
    hipLaunchKernelGGL(score7_kernel<4>, dim3(grid), dim3(score7_block_threads()), 0, st, a);
}

}  // namespace mst

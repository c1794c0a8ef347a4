// Batched all-pairs scores (R1) in packed 16-bit arithmetic, pipelined over
// the waves of a workgroup.
//
// The cell update is the packed kernel's (nw_score2.hip): pair (X, Ya) in the
// low and (X, Yb) in the high half of every VGPR, 6 packed ops per two cells
// plus one v_perm_b32 to widen the profile bytes. What differs is how the rows
// and columns are laid out:
//
//  * rows: a pair of row sequences with m rows is cut into W strips of 64*K
//    rows (K = 3..5 rows per lane, so no lane holds many rows); wave j of a
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
// (Ya, Yb); pipeline p takes the p-th of P column-balanced parts of the item.
#include "common.h"
#include "nw_score5.h"
#include "pk16.h"

namespace mst {

constexpr int U5 = 8;        // steps per unrolled block / prefetch distance
constexpr int RING5 = 128;   // columns held by a ring between two waves

template <int W>
struct Pipes {
  static constexpr int P = W == 1 ? 4 : (W == 2 ? 2 : 1);
};

__device__ __forceinline__ int flag_load(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void flag_store(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// every lane of the wave spins on the same LDS word; the other waves of the
// workgroup are resident, so the producer it waits for makes progress
__device__ __forceinline__ void wait_at_least(const int* p, int v) {
  while (flag_load(p) < v) __builtin_amdgcn_s_sleep(1);
}

struct Cursor {
  int kx;  // sorted position of the column sequence the lane is in
  int ce;  // its end column in the stream
};

struct Own {
  int ia, ib;  // row index within the lane of the last row of Ya / Yb, or -1
};

// Hm[idx] of the lane, as an or of masked rows: a chain of selects between
// the rows would let the compiler turn the selection into a variable index and
// move the lane state to scratch memory.
template <int K>
__device__ __forceinline__ int pick(const Lane2<K>& L, int idx, bool hi) {
  uint32_t v = 0;
#pragma unroll
  for (int i = 0; i < K; ++i) v |= as_u(L.Hm[i]) & (0u - (uint32_t)(idx == i));
  return hi ? (int)(short)(v >> 16) : (int)(short)(v & 0xFFFFu);
}

// scores of the column sequence at sorted position kx, from the lanes that
// own the last row of Ya or Yb
template <int K>
__device__ __forceinline__ void emit(const Lane2<K>& L, const Own& ow, int kx, int ya, int ma, int yb,
                                     int mb, const Score5Args& a) {
  if (ow.ia < 0 && ow.ib < 0) return;
  const int n = (int)(a.soff[kx + 1] - a.soff[kx]);
  const int x = a.order[kx];
  if (ow.ia >= 0) {
    const int s = pick<K>(L, ow.ia, false) - a.ge * (ma + n);
    a.out_mat[(int64_t)x * a.nseq + ya] = s;
    a.out_mat[(int64_t)ya * a.nseq + x] = s;
  }
  if (ow.ib >= 0) {
    const int s = pick<K>(L, ow.ib, true) - a.ge * (mb + n);
    a.out_mat[(int64_t)x * a.nseq + yb] = s;
    a.out_mat[(int64_t)yb * a.nseq + x] = s;
  }
}

struct Stream {
  int64_t base;  // first code of the stream in scodes
  int ncol;      // columns in the stream
  int ya, ma, yb, mb;
};

template <int K, bool PRED, bool CHECK, bool FIRSTW, bool STORE>
__device__ __forceinline__ void step5(Lane2<K>& L, Cursor& cu, const Own& ow, int t, int x, int2 bv,
                                      const uint8_t* prof_lane, const Stream& S, s2 G2, int2* ring_out,
                                      const Score5Args& a) {
  constexpr int D = KD<K>::v;
  const int lane = lane_id();
  uint32_t fin, hup;
  if (FIRSTW) {
    fin = shr1z(L.Fout);
    hup = shr1z(L.Hout);
  } else {
    fin = shr1o((uint32_t)bv.x, L.Fout);
    hup = shr1o((uint32_t)bv.y, L.Hout);
  }
  const int c = t - lane;
  const bool act = !PRED || (c >= 0 && c < S.ncol);
  if (act) {
    s2 hd = as_s2(L.hdiag);
    if (CHECK && c == cu.ce) {
      // first column of the next sequence: report the finished one, then
      // restart from the column -1 boundary
      emit<K>(L, ow, cu.kx, S.ya, S.ma, S.yb, S.mb, a);
#pragma unroll
      for (int i = 0; i < K; ++i) {
        L.Hm[i] = (s2){0, 0};
        L.E[i] = (s2){0, 0};
      }
      hd = (s2){0, 0};
      cu.kx += 1;
      cu.ce = (int)(a.soff[cu.kx + 1] - S.base);
    }
    uint32_t w[D];
    read_dwords<D>(prof_lane + x * (MST_WAVE * D * 4), w);
    s2 f = as_s2(fin), hu = as_s2(hup);
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const s2 e = pmax(L.E[i], L.Hm[i]);
      f = pmax(f, hu);
      const s2 dg = hd + prow(w, i);
      const s2 h = pmax(pmax(e, f), dg) - G2;
      hd = L.Hm[i];
      L.Hm[i] = h;
      L.E[i] = e;
      hu = h;
    }
    L.Fout = as_u(f);
    L.Hout = as_u(hu);
    if (STORE && lane == MST_WAVE - 1) ring_out[c & (RING5 - 1)] = make_int2((int)L.Fout, (int)L.Hout);
  }
  L.hdiag = hup;
}

template <bool FIRSTW>
__device__ __forceinline__ void load_block5(int t0, const int8_t* __restrict__ X, int ncol,
                                            const int2* ring_in, int* xs, int2* bs) {
  const int lane = lane_id();
#pragma unroll
  for (int j = 0; j < U5; ++j) {
    int c = t0 + j - lane;
    c = c < 0 ? 0 : (c >= ncol ? ncol - 1 : c);
    xs[j] = X[c];
    if (!FIRSTW) {
      int tt = t0 + j;
      tt = tt >= ncol ? ncol - 1 : tt;
      bs[j] = ring_in[tt & (RING5 - 1)];
    } else {
      bs[j] = make_int2(0, 0);
    }
  }
}

// One wave's sweep of the stream. prog_in: columns the wave above has written
// to ring_in; cons_out: columns this wave has read from it. prog_out/cons_in:
// the same for ring_out and the wave below.
template <int K, bool FIRSTW, bool STORE>
__device__ __forceinline__ void run_wave5(int kb, const Own& ow, const Stream& S,
                                          const uint8_t* prof_lane, s2 G2, const int2* ring_in,
                                          int2* ring_out, const int* prog_in, int* cons_out,
                                          int* prog_out, const int* cons_in, const Score5Args& a) {
  const int lane = lane_id();
  Cursor cu;
  cu.kx = kb;
  cu.ce = (int)(a.soff[kb + 1] - S.base);
  Lane2<K> L;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    L.Hm[i] = (s2){0, 0};
    L.E[i] = (s2){0, 0};
  }
  L.Fout = L.Hout = L.hdiag = 0;
  const int ncol = S.ncol;
  const int T = ncol + MST_WAVE - 1;
  const int8_t* X = a.scodes + S.base;
  int xa[U5], xb[U5];
  int2 ba[U5], bb[U5];
  if (!FIRSTW) wait_at_least(prog_in, min(U5, ncol));
  load_block5<FIRSTW>(0, X, ncol, ring_in, xa, ba);
  for (int tb = 0; tb < T; tb += U5) {
    if (!FIRSTW) wait_at_least(prog_in, min(tb + 2 * U5, ncol));
    load_block5<FIRSTW>(tb + U5, X, ncol, ring_in, xb, bb);
    if (!FIRSTW) {
      // the ring words of columns < tb + 2*U5 are in registers once the loads
      // complete; the release store orders them before the counter
      if (lane == 0) flag_store(cons_out, min(tb + 2 * U5, ncol));
    }
    if (STORE) {
      // this block writes columns tb-63 .. tb-56; their slots held columns
      // 128 earlier, which the wave below must have read
      const int need = tb - (MST_WAVE - 1) + (U5 - 1) - RING5 + 1;
      if (need > 0) wait_at_least(cons_in, min(need, ncol));
    }
    const bool full = (tb >= MST_WAVE - 1) && (tb + U5 <= ncol);
    const bool cross =
        __builtin_amdgcn_ballot_w64(cu.ce < ncol && cu.ce <= tb + (U5 - 1) - lane) != 0;
    if (full && !cross) {
#pragma unroll
      for (int j = 0; j < U5; ++j)
        step5<K, false, false, FIRSTW, STORE>(L, cu, ow, tb + j, xa[j], ba[j], prof_lane, S, G2,
                                              ring_out, a);
    } else if (!cross) {
#pragma unroll
      for (int j = 0; j < U5; ++j)
        if (tb + j < T)
          step5<K, true, false, FIRSTW, STORE>(L, cu, ow, tb + j, xa[j], ba[j], prof_lane, S, G2,
                                               ring_out, a);
    } else {
#pragma unroll
      for (int j = 0; j < U5; ++j)
        if (tb + j < T)
          step5<K, true, true, FIRSTW, STORE>(L, cu, ow, tb + j, xa[j], ba[j], prof_lane, S, G2,
                                              ring_out, a);
    }
    if (STORE) {
      int done = tb + U5 - (MST_WAVE - 1);
      done = done < 0 ? 0 : (done > ncol ? ncol : done);
      if (lane == MST_WAVE - 1) flag_store(prog_out, done);
    }
#pragma unroll
    for (int j = 0; j < U5; ++j) {
      xa[j] = xb[j];
      ba[j] = bb[j];
    }
  }
  if (STORE && lane == MST_WAVE - 1) flag_store(prog_out, ncol);
  emit<K>(L, ow, cu.kx, S.ya, S.ma, S.yb, S.mb, a);
}

template <int K, int W>
__global__ void __launch_bounds__(MST_WAVE * W * Pipes<W>::P) score5_kernel(Score5Args a) {
  constexpr int P = Pipes<W>::P;
  constexpr int NW = W * P;
  constexpr int D = KD<K>::v;
  constexpr int NR = W > 1 ? P * (W - 1) : 1;
  __shared__ int8_t S_lds[MST_LETTERS * MST_LETTERS];
  __shared__ __attribute__((aligned(16))) uint8_t prof[W * MST_LETTERS * MST_WAVE * D * 4];
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
  const uint8_t* prof_lane = prof + ((j * MST_LETTERS) * MST_WAVE + lane) * D * 4;
  const int2* ring_in = (W > 1 && j > 0) ? ring + (p * (W - 1) + j - 1) * RING5 : nullptr;
  int2* ring_out = (W > 1 && j < W - 1) ? ring + (p * (W - 1) + j) * RING5 : nullptr;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(a.counter, 1u);
    __syncthreads();
    const int64_t it = (int64_t)s_item + a.first;
    if (it >= a.last) break;
    const Item3 item = a.items[it];
    Stream S;
    S.ya = item.ya;
    S.yb = item.yb >= 0 ? item.yb : item.ya;
    const int8_t* Ya = a.seqs + a.off[S.ya];
    const int8_t* Yb = a.seqs + a.off[S.yb];
    S.ma = (int)(a.off[S.ya + 1] - a.off[S.ya]);
    S.mb = (int)(a.off[S.yb + 1] - a.off[S.yb]);
    // profile of every strip: thread (wave, lane) fills residues wave, wave+NW, ...
    for (int js = 0; js < W; ++js) {
      int ra[K], rb[K];
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const int r = js * MST_WAVE * K + lane * K + i;
        ra[i] = (r < S.ma) ? (int)Ya[r] * MST_LETTERS : 0;
        rb[i] = (r < S.mb) ? (int)Yb[r] * MST_LETTERS : 0;
      }
      for (int x = wv; x < MST_LETTERS; x += NW) {
        uint32_t v[D];
#pragma unroll
        for (int q = 0; q < D; ++q) v[q] = 0;
#pragma unroll
        for (int i = 0; i < K; ++i) {
          const uint32_t b0 = (uint32_t)((int)S_lds[ra[i] + x] + bonus) & 0xFFu;
          const uint32_t b1 = (uint32_t)((int)S_lds[rb[i] + x] + bonus) & 0xFFu;
          v[i >> 1] |= (b0 | (b1 << 8)) << (16 * (i & 1));
        }
        uint32_t* dst = (uint32_t*)(prof + (((js * MST_LETTERS) + x) * MST_WAVE + lane) * D * 4);
#pragma unroll
        for (int q = 0; q < D; ++q) dst[q] = v[q];
      }
    }
    if (threadIdx.x < NW) {
      s_prog[threadIdx.x] = 0;
      s_cons[threadIdx.x] = 0;
    }
    __syncthreads();
    // pipeline p's part of the item: sorted positions [kb, ke), balanced by columns
    const int xb0 = item.xbegin, xe0 = item.xbegin + item.xcount;
    const int64_t c0 = a.soff[xb0], total = a.soff[xe0] - c0;
    int kb = xb0, ke = xe0;
    if (P > 1) {
      const int64_t lo = total * p / P, hi = total * (p + 1) / P;
      kb = xe0;
      ke = xe0;
      for (int k = xb0; k < xe0; ++k) {
        const int64_t s = a.soff[k] - c0;
        if (kb == xe0 && s >= lo) kb = k;
        if (ke == xe0 && s >= hi && p + 1 < P) ke = k;
      }
    }
    if (kb < ke) {
      S.base = a.soff[kb];
      S.ncol = (int)(a.soff[ke] - S.base);
      const int r0 = j * MST_WAVE * K;
      Own ow;
      const int la = S.ma - 1 - r0, lb = S.mb - 1 - r0;
      ow.ia = (la >= 0 && la < MST_WAVE * K && la / K == lane) ? la % K : -1;
      ow.ib = (item.yb >= 0 && lb >= 0 && lb < MST_WAVE * K && lb / K == lane) ? lb % K : -1;
      if (item.yb < 0) S.yb = -1;
      int* prog = s_prog + p * W;
      int* cons = s_cons + p * W;
      if (W == 1) {
        run_wave5<K, true, false>(kb, ow, S, prof_lane, G2, nullptr, nullptr, nullptr, nullptr,
                                  nullptr, nullptr, a);
      } else if (j == 0) {
        run_wave5<K, true, true>(kb, ow, S, prof_lane, G2, nullptr, ring_out, nullptr, nullptr,
                                 prog + j, cons + j + 1, a);
      } else if (j == W - 1) {
        run_wave5<K, false, false>(kb, ow, S, prof_lane, G2, ring_in, nullptr, prog + j - 1,
                                   cons + j, nullptr, nullptr, a);
      } else {
        run_wave5<K, false, true>(kb, ow, S, prof_lane, G2, ring_in, ring_out, prog + j - 1,
                                  cons + j, prog + j, cons + j + 1, a);
      }
    }
    __syncthreads();  // profile, rings and counters are rebuilt for the next item
  }
}

bool score5_supported(int K, int W) {
  if (W == 1) return K >= 3 && K <= 5;
  return W >= 2 && W <= 4 && K >= 3 && K <= 4;
}

int score5_block_threads(int W) {
  switch (W) {
    case 1: return MST_WAVE * 1 * Pipes<1>::P;
    case 2: return MST_WAVE * 2 * Pipes<2>::P;
    case 3: return MST_WAVE * 3 * Pipes<3>::P;
    default: return MST_WAVE * 4 * Pipes<4>::P;
  }
}

// LDS of one workgroup (profile, rings, counters, matrix) -> workgroups per CU
int score5_groups_per_cu(int K, int W) {
  const int P = W == 1 ? 4 : (W == 2 ? 2 : 1);
  const int D = K <= 4 ? 2 : 4;
  const int nr = W > 1 ? P * (W - 1) : 1;
  const int lds = W * MST_LETTERS * MST_WAVE * D * 4 + nr * RING5 * 8 + 8 * W * P + 4 +
                  MST_LETTERS * MST_LETTERS;
  int g = (160 * 1024) / (lds + 512);
  const int waves = W * P;
  if (g * waves > 32) g = 32 / waves;
  return g < 1 ? 1 : g;
}

void launch_score5(int K, int W, const Score5Args& a, int grid, hipStream_t st) {
#define MST_CASE(k, w)                                                                 \
  if (K == k && W == w) {                                                              \
    hipLaunchKernelGGL((score5_kernel<k, w>), dim3(grid), dim3(score5_block_threads(w)), 0, st, a); \
    return;                                                                            \
  }
  MST_CASE(3, 1)
  MST_CASE(4, 1)
  MST_CASE(5, 1)
  MST_CASE(3, 2)
  MST_CASE(4, 2)
  MST_CASE(3, 3)
  MST_CASE(4, 3)
  MST_CASE(3, 4)
  MST_CASE(4, 4)
#undef MST_CASE
}

}  // namespace mst

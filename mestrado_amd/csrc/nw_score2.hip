// Batched all-pairs scores (R1) in packed 16-bit arithmetic: one wavefront
// scores TWO pairs that share the column sequence X, pair a in the low and
// pair b in the high half of every VGPR, with v_pk_max_i16 / v_pk_add_u16.
//
// The recurrence, bias and boundary are those of the int32 score kernel
// (nw_score.hip): in Hm units a cell is
//   E' = max(E', Hm_left); F' = max(F', Hm_up); Hm = max(max(E',F'), Hm_diag + S'') - (go-ge)
// i.e. 6 packed ops for 2 cells, plus one v_perm_b32 per 2 cells to widen the
// two profile bytes (pair a, pair b) of a row into 16-bit halves. The integer
// results are identical to the int32 kernel: the host only sends a pair here
// when its bound
//   Hm <= Smax*(min(R,n)+1) + ge*(R+n+2),  dg <= Hm + Smax + go + ge
// (R = padded rows) stays below 32767, and Hm >= -(go) always, so no 16-bit
// value wraps; every other pair goes to the int32 kernel.
//
// Rows per lane K is a template parameter (1..16, fine-grained so the row
// padding stays below one lane-row per lane); the profile holds two bytes per
// row (pair a, pair b), S+go+ge >= 0, so zero extension is exact.
//
// Work items: score2_kernel (one wave per workgroup) takes (X, Ya, Yb) items
// and handles any number of 64*K-row strips. score3_kernel is the single-strip
// form used for most pairs: a workgroup of SW waves builds the profile of one
// (Ya, Yb) row pair once in LDS and its waves sweep a chunk of column
// sequences X against it, so the LDS profile is shared by SW waves. Both are
// persistent and take items from an atomic counter in the host's longest-
// first order.
#include "common.h"
#include "nw_score2.h"
#include "pk16.h"

namespace mst {

constexpr int U2 = 8;  // steps per unrolled block / prefetch distance

template <int K>
__device__ __forceinline__ void build_profile2(uint8_t* prof, const int8_t* S_lds, const int8_t* Ya,
                                               int ma, const int8_t* Yb, int mb, int r0, int bonus) {
  const int lane = lane_id();
  constexpr int D = KD<K>::v;
  int ra[K], rb[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int r = r0 + lane * K + i;
    ra[i] = (r < ma) ? (int)Ya[r] * MST_LETTERS : 0;
    rb[i] = (r < mb) ? (int)Yb[r] * MST_LETTERS : 0;
  }
  for (int x = 0; x < MST_LETTERS; ++x) {
    uint32_t w[D];
#pragma unroll
    for (int q = 0; q < D; ++q) w[q] = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const uint32_t a = (uint32_t)((int)S_lds[ra[i] + x] + bonus) & 0xFFu;
      const uint32_t b = (uint32_t)((int)S_lds[rb[i] + x] + bonus) & 0xFFu;
      w[i >> 1] |= (a | (b << 8)) << (16 * (i & 1));
    }
    uint32_t* dst = (uint32_t*)(prof + (x * MST_WAVE + lane) * D * 4);
#pragma unroll
    for (int q = 0; q < D; ++q) dst[q] = w[q];
  }
}


template <int K, bool PRED, bool FIRST, bool STORE>
__device__ __forceinline__ void step2(Lane2<K>& L, int t, int x, int2 bv, const uint8_t* prof_lane,
                                      int n, s2 G2, int2* bout) {
  constexpr int D = KD<K>::v;
  const int lane = lane_id();
  uint32_t fin, hup;
  if (FIRST) {
    fin = shr1z(L.Fout);
    hup = shr1z(L.Hout);
  } else {
    fin = shr1o((uint32_t)bv.x, L.Fout);
    hup = shr1o((uint32_t)bv.y, L.Hout);
  }
  const int c = t - lane;
  const bool act = !PRED || (c >= 0 && c < n);
  if (act) {
    uint32_t w[D];
    read_dwords<D>(prof_lane + x * (MST_WAVE * D * 4), w);
    s2 f = as_s2(fin), hu = as_s2(hup), hd = as_s2(L.hdiag);
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
    if (STORE && lane == MST_WAVE - 1) bout[c] = make_int2((int)L.Fout, (int)L.Hout);
  }
  L.hdiag = hup;
}

template <bool FIRST>
__device__ __forceinline__ void load_block2(int t0, const int8_t* __restrict__ X, int n,
                                            const int2* bin, int* xs, int2* bs) {
  const int lane = lane_id();
#pragma unroll
  for (int j = 0; j < U2; ++j) {
    int c = t0 + j - lane;
    c = c < 0 ? 0 : (c >= n ? n - 1 : c);
    xs[j] = X[c];
    if (!FIRST) {
      int tt = t0 + j;
      tt = tt >= n ? n - 1 : tt;
      bs[j] = bin[tt];
    } else {
      bs[j] = make_int2(0, 0);
    }
  }
}

template <int K, bool FIRST, bool STORE>
__device__ void run_strip2(Lane2<K>& L, const int8_t* __restrict__ X, int n,
                           const uint8_t* prof_lane, const int2* bin, int2* bout, s2 G2) {
  const int T = n + MST_WAVE - 1;
  int xa[U2], xb[U2];
  int2 ba[U2], bb[U2];
  load_block2<FIRST>(0, X, n, bin, xa, ba);
  for (int tb = 0; tb < T; tb += U2) {
    load_block2<FIRST>(tb + U2, X, n, bin, xb, bb);
    const bool full = (tb >= MST_WAVE - 1) && (tb + U2 <= n);
    if (full) {
#pragma unroll
      for (int j = 0; j < U2; ++j)
        step2<K, false, FIRST, STORE>(L, tb + j, xa[j], ba[j], prof_lane, n, G2, bout);
    } else {
#pragma unroll
      for (int j = 0; j < U2; ++j)
        if (tb + j < T) step2<K, true, FIRST, STORE>(L, tb + j, xa[j], ba[j], prof_lane, n, G2, bout);
    }
#pragma unroll
    for (int j = 0; j < U2; ++j) {
      xa[j] = xb[j];
      ba[j] = bb[j];
    }
  }
}

template <int K>
__device__ __forceinline__ int extract_row(const Lane2<K>& L, int r0, int m, bool hi) {
  const int rl = (m - 1) - r0;
  const int owner = rl / K, idx = rl % K;
  uint32_t v = as_u(L.Hm[0]);
#pragma unroll
  for (int i = 1; i < K; ++i)
    if (idx == i) v = as_u(L.Hm[i]);
  const int h = hi ? (int)(short)(v >> 16) : (int)(short)(v & 0xFFFFu);
  return __shfl(h, owner);
}

template <int K>
__global__ void __launch_bounds__(MST_WAVE) score2_kernel(Score2Args a) {
  constexpr int D = KD<K>::v;
  __shared__ int8_t S_lds[MST_LETTERS * MST_LETTERS];
  __shared__ __attribute__((aligned(16))) uint8_t prof[MST_LETTERS * MST_WAVE * D * 4];
  __shared__ int s_item;
  const int lane = lane_id();
  for (int i = lane; i < MST_LETTERS * MST_LETTERS; i += MST_WAVE) S_lds[i] = a.subst[i];
  const int G = a.go - a.ge;
  const s2 G2 = {(short)G, (short)G};
  int2* buf = a.bufs + (int64_t)blockIdx.x * 2 * a.bufcap;
  const uint8_t* prof_lane = prof + lane * D * 4;
  for (;;) {
    if (lane == 0) s_item = (int)atomicAdd(a.counter, 1u);
    __syncthreads();
    const int64_t it = (int64_t)s_item + a.first;
    __syncthreads();
    if (it >= a.last) break;
    const Pair2 p = a.items[it];
    const int8_t* X = a.seqs + a.off[p.x];
    const int n = (int)(a.off[p.x + 1] - a.off[p.x]);
    const int8_t* Ya = a.seqs + a.off[p.ya];
    const int ma = (int)(a.off[p.ya + 1] - a.off[p.ya]);
    const int yb = p.yb >= 0 ? p.yb : p.ya;
    const int8_t* Yb = a.seqs + a.off[yb];
    const int mb = (int)(a.off[yb + 1] - a.off[yb]);
    const int m = ma > mb ? ma : mb;
    const int SH = MST_WAVE * K;
    const int nstrips = (m + SH - 1) / SH;
    Lane2<K> L;
    int sa = 0, sb = 0;
    for (int s = 0; s < nstrips; ++s) {
      const int r0 = s * SH;
      __builtin_amdgcn_wave_barrier();
      build_profile2<K>(prof, S_lds, Ya, ma, Yb, mb, r0, a.go + a.ge);
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int i = 0; i < K; ++i) {
        L.Hm[i] = (s2){0, 0};
        L.E[i] = (s2){0, 0};
      }
      L.Fout = L.Hout = L.hdiag = 0;
      const int2* bin = buf + (int64_t)(s & 1) * a.bufcap;
      int2* bout = buf + (int64_t)((s + 1) & 1) * a.bufcap;
      const bool last = (s == nstrips - 1);
      if (s == 0) {
        if (last)
          run_strip2<K, true, false>(L, X, n, prof_lane, bin, bout, G2);
        else
          run_strip2<K, true, true>(L, X, n, prof_lane, bin, bout, G2);
      } else {
        if (last)
          run_strip2<K, false, false>(L, X, n, prof_lane, bin, bout, G2);
        else
          run_strip2<K, false, true>(L, X, n, prof_lane, bin, bout, G2);
      }
      if (!last) agent_fence();
      // each pair's last row may end in an earlier strip than the other's
      if (ma - 1 >= r0 && ma - 1 < r0 + SH) sa = extract_row<K>(L, r0, ma, false) - a.ge * (ma + n);
      if (mb - 1 >= r0 && mb - 1 < r0 + SH) sb = extract_row<K>(L, r0, mb, true) - a.ge * (mb + n);
    }
    if (lane == 0) {
      if (a.out_mat) {
        a.out_mat[(int64_t)p.x * a.nseq + p.ya] = sa;
        a.out_mat[(int64_t)p.ya * a.nseq + p.x] = sa;
        if (p.yb >= 0) {
          a.out_mat[(int64_t)p.x * a.nseq + p.yb] = sb;
          a.out_mat[(int64_t)p.yb * a.nseq + p.x] = sb;
        }
      }
      if (a.out_pair) {
        if (p.oa >= 0) a.out_pair[p.oa] = sa;
        if (p.yb >= 0 && p.ob >= 0) a.out_pair[p.ob] = sb;
      }
    }
  }
}

// ---- shared-profile form: one (Ya, Yb) row pair, a chunk of X sequences ----
constexpr int SW = 4;  // waves per workgroup

template <int K>
__device__ __forceinline__ void build_profile2_wg(uint8_t* prof, const int8_t* S_lds,
                                                  const int8_t* Ya, int ma, const int8_t* Yb,
                                                  int mb, int bonus) {
  // thread (wave w, lane l) fills the residues x = w, w+SW, ... of lane l
  constexpr int D = KD<K>::v;
  const int lane = lane_id();
  const int w = threadIdx.x >> 6;
  int ra[K], rb[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int r = lane * K + i;
    ra[i] = (r < ma) ? (int)Ya[r] * MST_LETTERS : 0;
    rb[i] = (r < mb) ? (int)Yb[r] * MST_LETTERS : 0;
  }
  for (int x = w; x < MST_LETTERS; x += SW) {
    uint32_t v[D];
#pragma unroll
    for (int q = 0; q < D; ++q) v[q] = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const uint32_t a = (uint32_t)((int)S_lds[ra[i] + x] + bonus) & 0xFFu;
      const uint32_t b = (uint32_t)((int)S_lds[rb[i] + x] + bonus) & 0xFFu;
      v[i >> 1] |= (a | (b << 8)) << (16 * (i & 1));
    }
    uint32_t* dst = (uint32_t*)(prof + (x * MST_WAVE + lane) * D * 4);
#pragma unroll
    for (int q = 0; q < D; ++q) dst[q] = v[q];
  }
}

template <int K>
__global__ void __launch_bounds__(MST_WAVE* SW) score3_kernel(Score3Args a) {
  constexpr int D = KD<K>::v;
  __shared__ int8_t S_lds[MST_LETTERS * MST_LETTERS];
  __shared__ __attribute__((aligned(16))) uint8_t prof[MST_LETTERS * MST_WAVE * D * 4];
  __shared__ int s_item;
  const int lane = lane_id();
  const int w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < MST_LETTERS * MST_LETTERS; i += MST_WAVE * SW) S_lds[i] = a.subst[i];
  const int G = a.go - a.ge;
  const s2 G2 = {(short)G, (short)G};
  const uint8_t* prof_lane = prof + lane * D * 4;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(a.counter, 1u);
    __syncthreads();
    const int64_t it = (int64_t)s_item + a.first;
    if (it >= a.last) break;
    const Item3 p = a.items[it];
    const int8_t* Ya = a.seqs + a.off[p.ya];
    const int ma = (int)(a.off[p.ya + 1] - a.off[p.ya]);
    const int yb = p.yb >= 0 ? p.yb : p.ya;
    const int8_t* Yb = a.seqs + a.off[yb];
    const int mb = (int)(a.off[yb + 1] - a.off[yb]);
    build_profile2_wg<K>(prof, S_lds, Ya, ma, Yb, mb, a.go + a.ge);
    __syncthreads();
    for (int k = w; k < p.xcount; k += SW) {
      const int x = a.xlist[p.xbegin + k];
      const int8_t* X = a.seqs + a.off[x];
      const int n = (int)(a.off[x + 1] - a.off[x]);
      Lane2<K> L;
#pragma unroll
      for (int i = 0; i < K; ++i) {
        L.Hm[i] = (s2){0, 0};
        L.E[i] = (s2){0, 0};
      }
      L.Fout = L.Hout = L.hdiag = 0;
      run_strip2<K, true, false>(L, X, n, prof_lane, nullptr, nullptr, G2);
      const int sa = extract_row<K>(L, 0, ma, false) - a.ge * (ma + n);
      const int sb = extract_row<K>(L, 0, mb, true) - a.ge * (mb + n);
      if (lane == 0) {
        a.out_mat[(int64_t)x * a.nseq + p.ya] = sa;
        a.out_mat[(int64_t)p.ya * a.nseq + x] = sa;
        if (p.yb >= 0) {
          a.out_mat[(int64_t)x * a.nseq + p.yb] = sb;
          a.out_mat[(int64_t)p.yb * a.nseq + x] = sb;
        }
      }
    }
    __syncthreads();  // the profile is rebuilt for the next item
  }
}

void launch_score3(int K, const Score3Args& a, int grid, hipStream_t st) {
#define MST_CASE(k)                                                                          \
  case k:                                                                                    \
    hipLaunchKernelGGL(score3_kernel<k>, dim3(grid), dim3(MST_WAVE * SW), 0, st, a); \
    break;
  switch (K) {
    MST_CASE(1)
    MST_CASE(2)
    MST_CASE(3)
    MST_CASE(4)
    MST_CASE(5)
    MST_CASE(6)
    MST_CASE(7)
    MST_CASE(8)
    MST_CASE(9)
    MST_CASE(10)
    MST_CASE(11)
    MST_CASE(12)
    MST_CASE(13)
    MST_CASE(14)
    MST_CASE(15)
    default:
      hipLaunchKernelGGL(score3_kernel<16>, dim3(grid), dim3(MST_WAVE * SW), 0, st, a);
      break;
  }
#undef MST_CASE
}

int score3_waves_per_group() { return SW; }

void launch_score2(int K, const Score2Args& a, int grid, hipStream_t st) {
#define MST_CASE(k)                                                                      \
  case k:                                                                                \
    hipLaunchKernelGGL(score2_kernel<k>, dim3(grid), dim3(MST_WAVE), 0, st, a); \
    break;
  switch (K) {
    MST_CASE(1)
    MST_CASE(2)
    MST_CASE(3)
    MST_CASE(4)
    MST_CASE(5)
    MST_CASE(6)
    MST_CASE(7)
    MST_CASE(8)
    MST_CASE(9)
    MST_CASE(10)
    MST_CASE(11)
    MST_CASE(12)
    MST_CASE(13)
    MST_CASE(14)
    MST_CASE(15)
    default:
      hipLaunchKernelGGL(score2_kernel<16>, dim3(grid), dim3(MST_WAVE), 0, st, a);
      break;
  }
#undef MST_CASE
}

}  // namespace mst

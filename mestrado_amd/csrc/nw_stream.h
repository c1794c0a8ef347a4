// Device parts shared by the streamed packed score kernels (nw_score5.hip,
// nw_score6.hip): one wave sweeps a stream of column sequences against 64*K
// rows of a (Ya, Yb) row pair, two pairs per lane in packed 16-bit halves.
//
// Indexing: rows and columns are positions in a length-ordered copy of the
// codes (scodes, soff); order[pos] is the sequence's index in its own set, and
// a score lands at out[base + order[x] * ld + order[y]] and its transpose.
#pragma once
#include "common.h"
#include "nw_score5.h"
#include "pk16.h"

namespace mst {

constexpr int U5 = 8;        // steps per unrolled block / prefetch distance
constexpr int RING5 = 128;   // columns held by an LDS ring between two waves

__device__ __forceinline__ int flag_load(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void flag_store(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// every lane of the wave spins on the same LDS word; the wave it waits for is
// in the same workgroup, hence resident, and does not wait on this one
__device__ __forceinline__ void wait_at_least(const int* p, int v) {
  while (flag_load(p) < v) __builtin_amdgcn_s_sleep(1);
}

// F and H of the row above a strip, per column: element (base + c) & mask of
// p (an LDS ring: mask RING5-1; a full-length buffer in HBM: mask -1).
struct Link {
  int2* p;
  int base, mask;
};

struct Cursor {
  int kx;  // position of the column sequence the lane is in
  int ce;  // its end column in the stream
};

struct Own {
  int ia, ib;  // row index within the lane of the last row of Ya / Yb, or -1
};

struct Stream {
  int64_t base;   // first code of the stream in scodes
  int ncol;       // columns in the stream
  int ya, ma, yb, mb;  // row positions and lengths (yb < 0: Ya alone)
  int oya, oyb;   // their indices in the output
  int64_t obase;  // output offset and leading dimension of the item's set
  int ld;
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

// scores of the column sequence at position kx, from the lanes that own the
// last row of Ya or Yb
template <int K>
__device__ __forceinline__ void emit(const Lane2<K>& L, const Own& ow, int kx, const Stream& S,
                                     const Score5Args& a) {
  if (ow.ia < 0 && ow.ib < 0) return;
  const int n = (int)(a.soff[kx + 1] - a.soff[kx]);
  const int64_t ox = a.order[kx];
  int32_t* out = a.out_mat + S.obase;
  if (ow.ia >= 0) {
    const int s = pick<K>(L, ow.ia, false) - a.ge * (S.ma + n);
    out[ox * S.ld + S.oya] = s;
    out[(int64_t)S.oya * S.ld + ox] = s;
  }
  if (ow.ib >= 0) {
    const int s = pick<K>(L, ow.ib, true) - a.ge * (S.mb + n);
    out[ox * S.ld + S.oyb] = s;
    out[(int64_t)S.oyb * S.ld + ox] = s;
  }
}

// One step t of the wave: lane l computes column c = t - l of its K rows.
// w: the lane's profile dwords for the column's residue; bv: for lane 0, F and
// H of the row above at column t (the other lanes take them from lane l-1).
template <int K, bool PRED, bool CHECK, bool FIRSTW, bool STORE, bool WIDE, bool STACK>
__device__ __forceinline__ void step5(Lane2<K>& L, Cursor& cu, const Own& ow, int t,
                                      const uint32_t* w, int2 bv, const Stream& S, s2 G2,
                                      const Link& out, const Score5Args& a) {
  const int lane = lane_id();
  uint32_t fin, hup;
  if (FIRSTW) {
    fin = shr1z(L.Fout);
    hup = shr1z(L.Hout);
  } else {
    fin = shr1o((uint32_t)bv.x, L.Fout);
    hup = shr1o((uint32_t)bv.y, L.Hout);
  }
  if (STACK && lane == MST_WAVE / 2) {
    // first row of the second row pair: the row -1 boundary (0 when biased)
    fin = 0;
    hup = 0;
  }
  const int c = t - lane;
  const bool act = !PRED || (c >= 0 && c < S.ncol);
  if (act) {
    s2 hd = as_s2(L.hdiag);
    if (CHECK && c == cu.ce) {
      // first column of the next sequence: report the finished one, then
      // restart from the column -1 boundary
      emit<K>(L, ow, cu.kx, S, a);
#pragma unroll
      for (int i = 0; i < K; ++i) {
        L.Hm[i] = (s2){0, 0};
        L.E[i] = (s2){0, 0};
      }
      hd = (s2){0, 0};
      cu.kx += 1;
      cu.ce = (int)(a.soff[cu.kx + 1] - S.base);
    }
    s2 f = as_s2(fin), hu = as_s2(hup);
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const s2 e = pmax(L.E[i], L.Hm[i]);
      f = pmax(f, hu);
      const s2 dg = hd + prow_w<WIDE>(w, i);
      const s2 h = pmax(pmax(e, f), dg) - G2;
      hd = L.Hm[i];
      L.Hm[i] = h;
      L.E[i] = e;
      hu = h;
    }
    L.Fout = as_u(f);
    L.Hout = as_u(hu);
    if (STORE && lane == MST_WAVE - 1)
      out.p[(out.base + c) & out.mask] = make_int2((int)L.Fout, (int)L.Hout);
  }
  L.hdiag = hup;
}

// The residue codes lane l needs in the block of steps t0 .. t0+7, i.e. of
// columns t0-l .. t0+7-l, packed one per byte. Inside the stream they are 8
// contiguous bytes (one unaligned 8-byte load); at its ends each column is
// clamped into the stream (those lanes are idle at those steps).
__device__ __forceinline__ uint2 block_codes(const int8_t* __restrict__ X, int t0, int ncol) {
  const int c0 = t0 - lane_id();
  uint2 v;
  if (c0 >= 0 && c0 + U5 <= ncol) {
    __builtin_memcpy(&v, X + c0, sizeof v);
  } else {
    uint32_t q[2] = {0u, 0u};
#pragma unroll
    for (int j = 0; j < U5; ++j) {
      int c = c0 + j;
      c = c < 0 ? 0 : (c >= ncol ? ncol - 1 : c);
      q[j >> 2] |= (uint32_t)(uint8_t)X[c] << (8 * (j & 3));
    }
    v = make_uint2(q[0], q[1]);
  }
  return v;
}

__device__ __forceinline__ int code_at(uint2 v, int j) {
  return (int)(((j < 4 ? v.x : v.y) >> (8 * (j & 3))) & 0xFFu);
}

// One wave's sweep of a stream for one strip (rows r0 .. r0+64K-1 of the row
// pair). in/out: the row above / this strip's last row. Progress counters count
// columns in global units: gin + local column for `in` (the producer's units),
// gout + local column for `out`, so a wave that sweeps the stream again for a
// later strip keeps them monotonic and an LDS ring continues across sweeps.
//   prog_in: columns the producer of `in` has written; cons_out: columns this
//   wave has read from `in`; prog_out / cons_in: the same for `out`.
// out_span: columns `out` holds before a slot is reused (0: never reused
// within the sweep). pub_every: blocks between two progress releases of `out`
// (1 for an LDS ring; more for an HBM buffer, whose release waits for stores).
//
// Per block of 8 steps the wave loads the next block's residue codes (one
// 8-byte load per lane), reads the 8 steps' profile dwords from LDS before the
// first step, and reloads each step's `in` word for the next block right after
// the step has used it, so no register moves between blocks.
template <int K, bool FIRSTW, bool STORE, bool WIDE = false, bool STACK = false>
__device__ __forceinline__ void sweep5(int kb, int r0, const Stream& S, const uint8_t* prof_lane,
                                       s2 G2, const Link& in, int gin, const int* prog_in,
                                       int* cons_out, const Link& out, int gout, int* prog_out,
                                       const int* cons_in, int out_span, int pub_every,
                                       const Score5Args& a) {
  constexpr int D = KDW<K, WIDE>::v;
  const int lane = lane_id();
  const int ncol = S.ncol;
  const int T = ncol + MST_WAVE - 1;
  const int8_t* X = a.scodes + S.base;
  Own ow;
  {
    const int la = S.ma - 1 - r0, lb = S.mb - 1 - r0;
    ow.ia = (la >= 0 && la < MST_WAVE * K && la / K == lane) ? la % K : -1;
    ow.ib = (S.yb >= 0 && lb >= 0 && lb < MST_WAVE * K && lb / K == lane) ? lb % K : -1;
  }
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
  int2 bv[U5];
  if (!FIRSTW) {
    wait_at_least(prog_in, gin + min(U5, ncol));
#pragma unroll
    for (int j = 0; j < U5; ++j) bv[j] = in.p[(in.base + min(j, ncol - 1)) & in.mask];
  } else {
#pragma unroll
    for (int j = 0; j < U5; ++j) bv[j] = make_int2(0, 0);
  }
  uint2 codes = block_codes(X, 0, ncol);
  int blk = 0;
  for (int tb = 0; tb < T; tb += U5, ++blk) {
    // profile dwords of this block's 8 steps
    uint32_t w[U5][D];
#pragma unroll
    for (int j = 0; j < U5; ++j) read_dwords<D>(prof_lane + code_at(codes, j) * (MST_WAVE * D * 4), w[j]);
    codes = block_codes(X, tb + U5, ncol);
    if (!FIRSTW) wait_at_least(prog_in, gin + min(tb + 2 * U5, ncol));
    if (STORE && out_span > 0) {
      // this block writes columns tb-63 .. tb-56, whose slots held the columns
      // out_span earlier (in global units, possibly of the previous sweep);
      // the reader must be past them
      const int need = gout + tb - (MST_WAVE - 1) + (U5 - 1) - out_span + 1;
      if (need > 0) wait_at_least(cons_in, min(need, gout + ncol));
    }
    const bool full = (tb >= MST_WAVE - 1) && (tb + U5 <= ncol);
    const bool cross =
        __builtin_amdgcn_ballot_w64(cu.ce < ncol && cu.ce <= tb + (U5 - 1) - lane) != 0;
#define MST_STEPS(PRED, CHECK)                                                            \
  _Pragma("unroll") for (int j = 0; j < U5; ++j) {                                        \
    if (!PRED || tb + j < T)                                                              \
      step5<K, PRED, CHECK, FIRSTW, STORE, WIDE, STACK>(L, cu, ow, tb + j, w[j], bv[j], S, G2, out, a); \
    if (!FIRSTW) bv[j] = in.p[(in.base + min(tb + U5 + j, ncol - 1)) & in.mask];           \
  }
    if (full && !cross) {
      MST_STEPS(false, false)
    } else if (!cross) {
      MST_STEPS(true, false)
    } else {
      MST_STEPS(true, true)
    }
#undef MST_STEPS
    // the `in` words of columns < tb + 16 are read once the loads complete;
    // the release store orders them before the counter
    if (!FIRSTW && lane == 0) flag_store(cons_out, gin + min(tb + 2 * U5, ncol));
    if (STORE && (blk % pub_every) == pub_every - 1) {
      int done = tb + U5 - (MST_WAVE - 1);
      done = done < 0 ? 0 : (done > ncol ? ncol : done);
      if (lane == MST_WAVE - 1) flag_store(prog_out, gout + done);
    }
  }
  if (STORE && lane == MST_WAVE - 1) flag_store(prog_out, gout + ncol);
  emit<K>(L, ow, cu.kx, S, a);
}

// Profile bytes S[Y_r][x] + go + ge of rows r0 + lane*K + i for the residues
// x = x0, x0 + dx, ... < 24, two bytes (Ya, Yb) per row (WIDE: each in a
// 16-bit half of the row's dword), into prof laid out [residue][lane][D dwords].
template <int K, bool WIDE = false>
__device__ __forceinline__ void build_strip_profile(uint8_t* prof, const int8_t* S_lds, const Stream& S,
                                                    const Score5Args& a, int r0, int x0, int dx,
                                                    int bonus) {
  constexpr int D = KDW<K, WIDE>::v;
  const int lane = lane_id();
  const int8_t* Ya = a.scodes + a.soff[S.ya];
  const int8_t* Yb = a.scodes + a.soff[S.yb >= 0 ? S.yb : S.ya];
  const int mb = S.yb >= 0 ? S.mb : S.ma;
  int ra[K], rb[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int r = r0 + lane * K + i;
    ra[i] = (r < S.ma) ? (int)Ya[r] * MST_LETTERS : 0;
    rb[i] = (r < mb) ? (int)Yb[r] * MST_LETTERS : 0;
  }
  for (int x = x0; x < MST_LETTERS; x += dx) {
    uint32_t v[D];
#pragma unroll
    for (int q = 0; q < D; ++q) v[q] = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const uint32_t b0 = (uint32_t)((int)S_lds[ra[i] + x] + bonus) & 0xFFu;
      const uint32_t b1 = (uint32_t)((int)S_lds[rb[i] + x] + bonus) & 0xFFu;
      if (WIDE)
        v[i] = b0 | (b1 << 16);
      else
        v[i >> 1] |= (b0 | (b1 << 8)) << (16 * (i & 1));
    }
    uint32_t* dst = (uint32_t*)(prof + (x * MST_WAVE + lane) * D * 4);
#pragma unroll
    for (int q = 0; q < D; ++q) dst[q] = v[q];
  }
}

// The row pair and output mapping of work item `it`.
__device__ __forceinline__ Stream item_stream(const Item3& item, int64_t it, const Score5Args& a) {
  Stream S;
  S.ya = item.ya;
  S.yb = item.yb;
  S.ma = (int)(a.soff[S.ya + 1] - a.soff[S.ya]);
  S.mb = S.yb >= 0 ? (int)(a.soff[S.yb + 1] - a.soff[S.yb]) : S.ma;
  S.oya = a.order[S.ya];
  S.oyb = S.yb >= 0 ? a.order[S.yb] : S.oya;
  if (a.item_job) {
    const int j = a.item_job[it];
    S.obase = a.job_out[j];
    S.ld = a.job_ld[j];
  } else {
    S.obase = 0;
    S.ld = a.nseq;
  }
  S.base = 0;
  S.ncol = 0;
  return S;
}

}  // namespace mst

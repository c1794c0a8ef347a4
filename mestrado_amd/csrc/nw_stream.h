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

// Re-based sweeps (REBASE): the lanes' 16-bit values stand for v + o, where
// o is a wave-uniform offset per half (o_a, o_b) that the wave raises by lane
// 0's first-row value whenever that value leaves [-kRebaseAt, kRebaseAt]
// (checked once per block of 8 steps). In the biased form a value changes by
// -8..+23 from one row or column to the next (BLOSUM62, go = 10, ge = 2; the
// planner checks the general bound), so with 512 rows per strip every value
// of the wave's front lies within -5,600..+12,300 of lane 0's first row, and
// the values never wrap however long the sequences are. The row handed to the
// next strip is written in the writer's frame; the frame of every block of 8
// written columns goes to `fr_out` (index: the writer's step / 8), and the
// reader converts each column by the difference of the two frames (modulo
// 2^16, exact because the converted value is in range).
struct Rebase {
  int2* fr_out;       // frames of the blocks this sweep writes (null: none)
  const int2* fr_in;  // frames of the blocks the previous strip wrote
};

// A row sequence whose last row lies in this strip: the lane holding that row
// reports the row's H at the end of every column sequence. All fields are
// uniform across the wave.
struct Owner {
  int lane;  // lane of the last row (-1: the row sequence ends outside this strip)
  int row;   // its row index within the lane
  int hi;    // 0: pair a (low halves), 1: pair b (high halves)
  int m;     // rows of the row sequence
  int oy;    // its index in the output
  int k;     // column sequence the lane is in (-1 before the stream's first column)
};
constexpr int kMaxOwners = 4;
struct Owners {
  Owner o[kMaxOwners];
  uint64_t lanes;  // bit l: lane l holds a last row
};

struct Stream {
  int64_t base;   // first code of the stream in scodes
  int ncol;       // columns in the stream
  int ya, ma, yb, mb;  // row positions and lengths (yb < 0: Ya alone)
  int oya, oyb;   // their indices in the output
  int64_t obase;  // output offset and leading dimension of the item's set
  int ld;
};

// Owners of a strip whose first row is row r0 of the row pair of S (uniform),
// and with STACK of the second row pair S2, whose rows start at lane 32.
template <int K>
__device__ __forceinline__ Owners make_owners(const Stream& S, int r0, const Stream* S2) {
  Owners w;
  auto set = [&](Owner& o, int m, int oy, int hi, int first_row, bool on) {
    const int l = m - 1 - first_row;
    const bool in = on && l >= 0 && l < MST_WAVE * K;
    o.lane = in ? l / K : -1;
    o.row = in ? l % K : 0;
    o.hi = hi;
    o.m = m;
    o.oy = oy;
    o.k = -1;
  };
  set(w.o[0], S.ma, S.oya, 0, r0, true);
  set(w.o[1], S.mb, S.oyb, 1, r0, S.yb >= 0);
  if (S2) {
    set(w.o[2], S2->ma, S2->oya, 0, r0 - (MST_WAVE / 2) * K, true);
    set(w.o[3], S2->mb, S2->oyb, 1, r0 - (MST_WAVE / 2) * K, S2->yb >= 0);
  } else {
    w.o[2].lane = w.o[3].lane = -1;
    w.o[2].k = w.o[3].k = -1;
  }
  w.lanes = 0;
#pragma unroll
  for (int q = 0; q < kMaxOwners; ++q)
    if (w.o[q].lane >= 0) w.lanes |= 1ull << w.o[q].lane;
  return w;
}

// Hm[row] of the lane, half hi, as an or of masked rows: a chain of selects
// would let the compiler turn the selection into a variable index and move
// the lane state to scratch memory.
template <int K>
__device__ __forceinline__ int pick_row(const Lane2<K>& L, int row, int hi) {
  uint32_t v = 0;
#pragma unroll
  for (int i = 0; i < K; ++i) v |= as_u(L.Hm[i]) & (0u - (uint32_t)(row == i));
  return hi ? (int)(short)(v >> 16) : (int)(short)(v & 0xFFFFu);
}

// Column boundaries of the stream: B_k = first column of its k-th sequence
// (B_0 = 0, lane k of `bnd`) for k < nseq, B_nseq = ncol, none after.
__device__ __forceinline__ int boundary(int bnd, int k, int nseq, int ncol) {
  return k < nseq ? __builtin_amdgcn_readlane(bnd, k) : (k == nseq ? ncol : 0x7FFFFFFF);
}

// The owner's lane has reached a boundary: report the sequence it finished
// (none at the stream's first column) and move on to the next.
// oa, ob: the frame of a re-based sweep (0 otherwise).
template <int K>
__device__ __forceinline__ void emit_owner(const Lane2<K>& L, Owner& o, int bnd, int ord, int nseq,
                                           const Stream& S, const Score5Args& a, int oa = 0, int ob = 0) {
  if (o.k >= 0) {
    const int n = boundary(bnd, o.k + 1, nseq, S.ncol) - boundary(bnd, o.k, nseq, S.ncol);
    const int64_t ox = __builtin_amdgcn_readlane(ord, o.k);
    const int v = pick_row<K>(L, o.row, o.hi) + (o.hi ? ob : oa);
    if (lane_id() == o.lane) {
      const int s = v - a.ge * (o.m + n);
      int32_t* out = a.out_mat + S.obase;
      out[ox * S.ld + o.oy] = s;
      out[(int64_t)o.oy * S.ld + ox] = s;
    }
  }
  ++o.k;
}

// The row update of one step: lane l computes its K rows at column t - l.
// fin, hup: F and H of the row above at this column; L.hdiag: H of the row
// above at the previous column.
template <int K, bool WIDE>
__device__ __forceinline__ void rows5(Lane2<K>& L, uint32_t fin, uint32_t hup, s2 hd, const uint32_t* w,
                                      s2 G2) {
  s2 f = as_s2(fin), hu = as_s2(hup);
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const s2 e = pmax(L.E[i], L.Hm[i]);
    f = pmax(f, hu);
    const s2 dg = hd + prow_w<WIDE>(w, i);
    const s2 h = pmax(pmax(e, dg), f) - G2;
    hd = L.Hm[i];
    L.Hm[i] = h;
    L.E[i] = e;
    hu = h;
  }
  L.Fout = as_u(f);
  L.Hout = as_u(hu);
}

// rows5 for a step at which some lanes restart from the column -1 boundary,
// when that boundary is 0 in both halves: keep = 0 on a restarting lane, 1 in
// each half elsewhere (km: 0 or all ones). The lane's E and the previous
// column's H (hd of every row) then read as 0: E by a mask, hd folded into the
// diagonal's add as hd * keep + profile (one v_pk_mad_u16 in place of the add),
// K + 2 instructions more than rows5 where writing the boundary into the
// lane's 2K + 1 registers costs more.
template <int K, bool WIDE>
__device__ __forceinline__ void rows5r(Lane2<K>& L, uint32_t fin, uint32_t hup, s2 hd, const uint32_t* w,
                                       s2 G2, uint32_t km, s2 keep) {
  s2 f = as_s2(fin), hu = as_s2(hup);
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const s2 e = as_s2(as_u(pmax(L.E[i], L.Hm[i])) & km);
    f = pmax(f, hu);
    const s2 dg = hd * keep + prow_w<WIDE>(w, i);
    const s2 h = pmax(pmax(e, dg), f) - G2;
    hd = L.Hm[i];
    L.Hm[i] = h;
    L.E[i] = e;
    hu = h;
  }
  L.Fout = as_u(f);
  L.Hout = as_u(hu);
}

// F and H of the row above a lane's first row at this step's column: lane
// l-1's last row, lane 0's from bv (or the row -1 boundary in the first strip).
template <bool FIRSTW, bool STACK>
__device__ __forceinline__ void row_above(uint32_t Fout, uint32_t Hout, int2 bv, uint32_t& fin,
                                          uint32_t& hup) {
  if (FIRSTW) {
    fin = shr1z(Fout);
    hup = shr1z(Hout);
  } else {
    fin = shr1o((uint32_t)bv.x, Fout);
    hup = shr1o((uint32_t)bv.y, Hout);
  }
  if (STACK && lane_id() == MST_WAVE / 2) {
    // first row of the second row pair: the row -1 boundary (0 when biased)
    fin = 0;
    hup = 0;
  }
}

// A step of a block in which no lane meets a column boundary.
template <int K, bool FIRSTW, bool STORE, bool WIDE, bool STACK>
__device__ __forceinline__ void step5(Lane2<K>& L, int t, const uint32_t* w, int2 bv, s2 G2,
                                      const Link& out) {
  uint32_t fin, hup;
  row_above<FIRSTW, STACK>(L.Fout, L.Hout, bv, fin, hup);
  rows5<K, WIDE>(L, fin, hup, as_s2(L.hdiag), w, G2);
  if (STORE && lane_id() == MST_WAVE - 1)
    out.p[(out.base + t - (MST_WAVE - 1)) & out.mask] = make_int2((int)L.Fout, (int)L.Hout);
  L.hdiag = hup;
}

// A step of a block in which lanes may meet a column boundary. Bit l of the
// uniform mask M is set when lane l's column is a boundary: the lane then
// restarts from the column -1 boundary (all zero in the biased form) after
// its owners have reported. Lanes outside the stream compute values nobody
// reads; they restart at its first column, and the last boundary (ncol) ends
// every lane's last sequence. enter: a boundary reaches lane 0 at this step
// (the block's entries are found before its first step, so the steps carry
// no boundary search of their own).
// z: the boundary value in the sweep's frame (0 unless re-based, REB).
template <int K, bool FIRSTW, bool STORE, bool WIDE, bool STACK, bool REB = false>
__device__ __forceinline__ void step5c(Lane2<K>& L, uint64_t& M, uint32_t enter, Owners& ow, int bnd,
                                       int ord, int nseq, int t, const uint32_t* w, int2 bv,
                                       const Stream& S, s2 G2, const Link& out, const Score5Args& a,
                                       s2 z = (s2){0, 0}, int oa = 0, int ob = 0) {
  M = (M << 1) | (uint64_t)enter;
  if (M & ow.lanes) {
#pragma unroll
    for (int q = 0; q < kMaxOwners; ++q)
      if (ow.o[q].lane >= 0 && ((M >> ow.o[q].lane) & 1ull))
        emit_owner<K>(L, ow.o[q], bnd, ord, nseq, S, a, oa, ob);
  }
  uint32_t fin, hup;
  row_above<FIRSTW, STACK>(L.Fout, L.Hout, bv, fin, hup);
  // bit l of the uniform M is lane l's condition: used as the lane mask as is
  const bool rst = __builtin_amdgcn_inverse_ballot_w64(M);
  if (REB) {
    s2 hd = as_s2(L.hdiag);
    if (rst) {
      hd = z;
#pragma unroll
      for (int i = 0; i < K; ++i) {
        L.Hm[i] = z;
        L.E[i] = z;
      }
    }
    rows5<K, WIDE>(L, fin, hup, hd, w, G2);
  } else {
    const uint32_t km = rst ? 0u : 0xFFFFFFFFu;
    const s2 keep = as_s2(rst ? 0u : 0x00010001u);
    rows5r<K, WIDE>(L, fin, hup, as_s2(L.hdiag), w, G2, km, keep);
  }
  const int c = t - (MST_WAVE - 1);
  if (STORE && lane_id() == MST_WAVE - 1 && c >= 0 && c < S.ncol)
    out.p[(out.base + c) & out.mask] = make_int2((int)L.Fout, (int)L.Hout);
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

// block_codes from the codes as profile offsets (Score5Args::soffs8): the
// lane's 8 offsets, one 32-byte load, unclamped. A column outside the stream
// reads a neighbouring sequence's offset or the array's zero padding
// (kOffs8Pad entries at each end), a valid profile address either way, and
// the lane's values at such a column are never read.
struct Offs8 {
  uint4 lo, hi;
};
__device__ __forceinline__ Offs8 block_offs8(const uint32_t* __restrict__ X8, int t0) {
  Offs8 v;
  __builtin_memcpy(&v, X8 + (t0 - lane_id()), sizeof v);
  return v;
}

__device__ __forceinline__ uint32_t off_at(const Offs8& v, int j) {
  const uint4 q = j < 4 ? v.lo : v.hi;
  return (j & 3) == 0 ? q.x : ((j & 3) == 1 ? q.y : ((j & 3) == 2 ? q.z : q.w));
}

// One wave's sweep of a stream (column sequences kb .. ke-1, at most
// kStreamSeqs) for one strip of the row pair; ow: the strip's owners of last
// rows. in/out: the row above / this strip's last row. Progress counters count
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
// the step has used it, so no register moves between blocks. The stream's
// column boundaries (one per lane of a VGPR) drive a uniform 64-bit mask of the
// lanes at a boundary, which shifts by one lane per step: a block where it is
// empty and no boundary enters runs the plain steps, any other block the steps
// with restarts and reports, without a branch per lane.
//
// REBASE (a stream of one column sequence; see Rebase): the frame is checked
// at the start of every block, the row -1 boundary and the restart values are
// taken in the current frame, `in` words are converted from the writer's
// frame as they are used, and the frame of each block goes to rb.fr_out.
//
// OFF8 (profiles of 8 dwords per lane, Score5Args::soffs8 set): the step's
// profile address is one add from the codes as offsets; sweep5 below picks
// the form once per sweep, so each loop carries only its own registers.
template <int K, bool FIRSTW, bool STORE, bool WIDE, bool STACK, bool REBASE, bool OFF8>
__device__ __forceinline__ void sweep5_body(int kb, int ke, Owners ow, const Stream& S, const uint8_t* prof_lane,
                                            s2 G2, const Link& in, int gin, const int* prog_in,
                                            int* cons_out, const Link& out, int gout, int* prog_out,
                                            const int* cons_in, int out_span, int pub_every,
                                            const Score5Args& a, Rebase rb) {
  constexpr int D = KDW<K, WIDE>::v;
  static_assert(!OFF8 || D == 8, "offsets of a profile of 8 dwords per lane");
  // the step's row -1 input comes from bv (the converted `in` word, or the
  // boundary in the current frame) unless the sweep is a plain first strip
  constexpr bool ZFILL = FIRSTW && !REBASE;
  const int lane = lane_id();
  const int ncol = S.ncol;
  const int T = ncol + MST_WAVE - 1;
  const int nseq = ke - kb;  // <= kStreamSeqs
  const int8_t* X = a.scodes + S.base;
  const int bnd = lane < nseq ? (int)(a.soff[kb + lane] - S.base) : 0x7FFFFFFF;
  const int ord = lane < nseq ? a.order[kb + lane] : 0;
  uint64_t M = 0;
  int nb = 0, bnext = 0;  // the next boundary to enter lane 0: B_0 = 0
  int oa = 0, ob = 0;     // frame (REBASE)
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
  const uint32_t* X8 = OFF8 ? a.soffs8 + S.base : nullptr;
  uint2 codes = make_uint2(0u, 0u);
  Offs8 offs = {make_uint4(0u, 0u, 0u, 0u), make_uint4(0u, 0u, 0u, 0u)};
  if (OFF8)
    offs = block_offs8(X8, 0);
  else
    codes = block_codes(X, 0, ncol);
  int blk = 0;
  for (int tb = 0; tb < T; tb += U5, ++blk) {
    s2 z = (s2){0, 0};      // the boundary value in the current frame
    s2 d0 = z, d1 = z;      // frame conversion of the `in` words of columns tb, tb+1..
    if (REBASE) {
      const uint32_t rep = __builtin_amdgcn_readfirstlane(as_u(L.Hm[0]));
      const int ra = (int)(short)(rep & 0xFFFFu), rbh = (int)(short)(rep >> 16);
      const int da = (ra > kRebaseAt || ra < -kRebaseAt) ? ra : 0;
      const int db = (rbh > kRebaseAt || rbh < -kRebaseAt) ? rbh : 0;
      if (da | db) {
        const s2 dl = as_s2(((uint32_t)da & 0xFFFFu) | ((uint32_t)db << 16));
#pragma unroll
        for (int i = 0; i < K; ++i) {
          L.Hm[i] = L.Hm[i] - dl;
          L.E[i] = L.E[i] - dl;
        }
        L.Fout = as_u(as_s2(L.Fout) - dl);
        L.Hout = as_u(as_s2(L.Hout) - dl);
        L.hdiag = as_u(as_s2(L.hdiag) - dl);
        oa += da;
        ob += db;
      }
      z = as_s2(((uint32_t)(-oa) & 0xFFFFu) | ((uint32_t)(-ob) << 16));
      if (STORE && lane == MST_WAVE - 1) rb.fr_out[tb >> 3] = make_int2(oa, ob);
      if (!FIRSTW) {
        // the writer wrote column c in its block (c + 63) / 8 (c clamped into
        // the stream): column tb in one block, tb+1 .. tb+7 in the next
        const int2 f0 = rb.fr_in[(min(tb, ncol - 1) + MST_WAVE - 1) >> 3];
        const int2 f1 = rb.fr_in[(min(tb + 1, ncol - 1) + MST_WAVE - 1) >> 3];
        d0 = as_s2(((uint32_t)(f0.x - oa) & 0xFFFFu) | ((uint32_t)(f0.y - ob) << 16));
        d1 = as_s2(((uint32_t)(f1.x - oa) & 0xFFFFu) | ((uint32_t)(f1.y - ob) << 16));
      }
    }
    // the row -1 input of step j: the `in` word, converted, or the boundary
    auto above = [&](int j) -> int2 {
      if (!REBASE) return bv[j];
      if (FIRSTW) return make_int2((int)as_u(z), (int)as_u(z));
      const s2 d = j == 0 ? d0 : d1;
      return make_int2((int)as_u(as_s2((uint32_t)bv[j].x) + d), (int)as_u(as_s2((uint32_t)bv[j].y) + d));
    };
    // profile dwords of this block's 8 steps
    uint32_t w[U5][D];
    if (OFF8) {
#pragma unroll
      for (int j = 0; j < U5; ++j) read_dwords<D>(prof_lane + off_at(offs, j), w[j]);
      offs = block_offs8(X8, tb + U5);
    } else {
#pragma unroll
      for (int j = 0; j < U5; ++j) read_dwords<D>(prof_lane + code_at(codes, j) * (MST_WAVE * D * 4), w[j]);
      codes = block_codes(X, tb + U5, ncol);
    }
    if (!FIRSTW) wait_at_least(prog_in, gin + min(tb + 2 * U5, ncol));
    if (STORE && out_span > 0) {
      // this block writes columns tb-63 .. tb-56, whose slots held the columns
      // out_span earlier (in global units, possibly of the previous sweep);
      // the reader must be past them
      const int need = gout + tb - (MST_WAVE - 1) + (U5 - 1) - out_span + 1;
      if (need > 0) wait_at_least(cons_in, min(need, gout + ncol));
    }
    if (M == 0 && bnext >= tb + U5) {
#pragma unroll
      for (int j = 0; j < U5; ++j) {
        step5<K, ZFILL, STORE, WIDE, STACK>(L, tb + j, w[j], above(j), G2, out);
        if (!FIRSTW) bv[j] = in.p[(in.base + min(tb + U5 + j, ncol - 1)) & in.mask];
      }
    } else {
      // the boundaries that reach lane 0 in this block (bit j: step tb + j);
      // every earlier one was taken by an earlier block
      uint32_t ent = 0;
      while (bnext < tb + U5) {
        ent |= 1u << (bnext - tb);
        ++nb;
        bnext = boundary(bnd, nb, nseq, S.ncol);
      }
      // Steps past T in the last block are harmless, so they are not tested
      // for: the final boundary reaches lane 63 at step T, whose owners
      // report then (in the loop) or after it; later steps only shift that
      // bit out of M, read clamped codes and `in` words, and store nothing
      // (the stores check the column); no lane's values are read afterwards.
#pragma unroll
      for (int j = 0; j < U5; ++j) {
        step5c<K, ZFILL, STORE, WIDE, STACK, REBASE>(L, M, (ent >> j) & 1u, ow, bnd, ord, nseq, tb + j, w[j],
                                                     above(j), S, G2, out, a, z, oa, ob);
        if (!FIRSTW) bv[j] = in.p[(in.base + min(tb + U5 + j, ncol - 1)) & in.mask];
      }
    }
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
  // lane 63's last column is the stream's last: its owners report after the
  // loop (every other lane has passed the final boundary)
#pragma unroll
  for (int q = 0; q < kMaxOwners; ++q)
    if (ow.o[q].lane >= 0 && ow.o[q].k < nseq) emit_owner<K>(L, ow.o[q], bnd, ord, nseq, S, a, oa, ob);
}

template <int K, bool FIRSTW, bool STORE, bool WIDE = false, bool STACK = false, bool REBASE = false>
__device__ __forceinline__ void sweep5(int kb, int ke, Owners ow, const Stream& S, const uint8_t* prof_lane,
                                       s2 G2, const Link& in, int gin, const int* prog_in,
                                       int* cons_out, const Link& out, int gout, int* prog_out,
                                       const int* cons_in, int out_span, int pub_every,
                                       const Score5Args& a, Rebase rb = Rebase{nullptr, nullptr}) {
  if constexpr (KDW<K, WIDE>::v == 8) {
    if (a.soffs8 != nullptr) {
      sweep5_body<K, FIRSTW, STORE, WIDE, STACK, REBASE, true>(kb, ke, ow, S, prof_lane, G2, in, gin, prog_in,
                                                                cons_out, out, gout, prog_out, cons_in,
                                                                out_span, pub_every, a, rb);
      return;
    }
  }
  sweep5_body<K, FIRSTW, STORE, WIDE, STACK, REBASE, false>(kb, ke, ow, S, prof_lane, G2, in, gin, prog_in,
                                                             cons_out, out, gout, prog_out, cons_in, out_span,
                                                             pub_every, a, rb);
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

// Part p of P of an item's column sequences [xb0, xe0), balanced by columns:
// part q ends before the first sequence whose middle column is at or past
// q/P of the stream (an empty part when one sequence spans the cut). Every
// part computes the same cuts, so neighbouring parts meet exactly.
__device__ __forceinline__ void item_part(const Score5Args& a, int xb0, int xe0, int p, int P, int* kb,
                                          int* ke) {
  const int64_t c0 = a.soff[xb0], total = a.soff[xe0] - c0;
  auto cut = [&](int q) {
    if (q <= 0) return xb0;
    if (q >= P) return xe0;
    const int64_t target2 = 2 * (total * q / P);
    for (int k = xb0; k < xe0; ++k)
      if ((a.soff[k] - c0) + (a.soff[k + 1] - c0) >= target2) return k;
    return xe0;
  };
  *kb = cut(p);
  *ke = cut(p + 1);
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

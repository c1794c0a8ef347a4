// Alignment forward pass with traceback flags (R2: NeedlemanWunschGlobalMatch
// Kernel, NeedlemanWunsch.cu:184-365) in packed 16-bit arithmetic: two jobs
// per wave, job a in the low and job b in the high half of every VGPR.
//
// Layout and recurrence are the int32 flag kernel's (nw_align.hip): 512-row
// strips, lane l owns rows 8l .. 8l+7 of a strip and sweeps the columns skewed
// by one step per lane, in the biased max-plus form of the score kernels; the
// substitution scores come from a byte table in the same skewed order
// (tab[(strip * (n+63) + t) * 64 + lane], 8 rows per lane, S + go + ge). The
// two jobs of a wave run their strips in step (strip s of both at once); a job
// with fewer strips or columns idles in its half. A job takes this kernel only
// when no 16-bit value of its DP can wrap (the bound of the packed score
// kernels), so every integer equals the int32 kernel's.
//
// Flags: the four decisions of a cell are the signs of four packed
// differences (each fits 16 bits, no value being negative beyond -(go+ge) nor
// above the bound):
//   s1 = D - max(E,F)      (sign set: the diagonal is not a maximum)
//   s2 = F - E             (sign set: E > F)
//   s4 = F_above - H_above (sign set: F was opened, not extended)
//   s3 = E_left - H_left   (sign set: E was opened, not extended)
// so LEFT = !s1 | s2, UP = !s1 | !s2, contUp = !s4, contLeft = !s3 (diagonal
// before up before left, extension preferred on a tie). The sign bits of a
// lane's 8 rows are gathered as bit planes: the flag word of a job, lane and
// step holds s1 of rows 0..7 in bits 0..7, s2 in 8..15, s4 in 16..23 and s3 in
// 24..31 (0.5 byte per cell, as the nibble layout; the traceback decodes it
// by the job's format). The row -1 and column -1 boundaries are 0 in the
// biased form, as in the score kernels, which leaves every value unchanged;
// the two flags they would otherwise set are forced to the reference's value
// (contLeft = 0 in column 0, contUp = 0 in row 0, which see a -inf there).
#include <algorithm>
#include <vector>

#include "common.h"
#include "nw_align.h"
#include "pk16.h"

namespace mst {

constexpr int FK = 8;               // most rows per lane (strips of 64 K rows, K = 1..8)
constexpr int FU = 2;               // steps per block / prefetch distance (4: 143 VGPRs, 3 waves per SIMD)

template <int K>
struct F16Lane {
  s2 Hm[K], E[K];
  uint32_t Fout, Hout, hdiag;
};

// One pass of a wave: a strip in each half. Half a computes column t - l at
// step t (lane l), half b column t - l - db: db = 0 for two independent jobs,
// db = 64 for the next strip of the same job, whose lane 0 then takes the row
// above from lane 63's half a of the step before (a wave rotation).
struct F16Strip {
  int n;                 // steps of the pass, less the 63 of the skew
  int na, nb, db;        // columns of each half (0: idle), half b's lag
  const uint2* ta;       // table words of half a's strip (half b's: tb)
  const uint2* tb;
  int tna, tnb;          // steps of each half's table strip (n + 63)
  uint32_t* fa;          // flag words of each half's strip
  uint32_t* fb;
  uint32_t sel;          // lane 0's row above: bytes of (rotated word, row buffer word)
  uint32_t cua, cub;     // 0x10000 where the half is a job's first strip (contUp = 0 in row 0)
  const int8_t* xa;      // profile mode: column residues of each half's job
  const int8_t* xb;
};

__device__ __forceinline__ uint32_t bfi_u(uint32_t mask, uint32_t a, uint32_t b) {
  return (a & mask) | (b & ~mask);
}

// lane l takes lane l-1's value, lane 0 lane 63's
__device__ __forceinline__ uint32_t ror1(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_WAVE_ROR1, 0xF, 0xF, false);
}

// Lane 0's selectors of the row above (v_perm of (rotated, row buffer)):
constexpr uint32_t kSelBuf = 0x03020100u;      // both halves from the row buffer (two jobs)
constexpr uint32_t kSelZero = 0x0C0C0C0Cu;     // the row -1 boundary in both halves
constexpr uint32_t kSelFirstPair = 0x05040C0Cu;  // a: boundary, b: rotated half a
constexpr uint32_t kSelPair = 0x05040302u;     // a: the buffer's half b (previous pass), b: rotated half a

// One step: lane l computes its columns of the K rows in both halves. PRED:
// some lane may be outside a half's columns; the inactive half keeps its
// state, so each half's lanes end holding column n-1 of their rows. bf, bh:
// F and H of the row buffer at column t for lane 0 (wave-uniform).
template <int K, bool PRED, bool STORE, bool NIB = false>
__device__ __forceinline__ void f16_step(F16Lane<K>& L, int t, uint2 wa, uint2 wb, uint32_t bf, uint32_t bh,
                                         const F16Strip& S, s2 G2, int2* bout) {
  const int lane = lane_id();
  uint32_t fin = ror1(L.Fout), hup = ror1(L.Hout);
  if (lane == 0) {
    fin = __builtin_amdgcn_perm(fin, bf, S.sel);
    hup = __builtin_amdgcn_perm(hup, bh, S.sel);
  }
  const int c = t - lane, cb = c - S.db;
  const bool aa = !PRED || (c >= 0 && c < S.na);
  const bool ab = !PRED || (cb >= 0 && cb < S.nb);
  if (aa || ab) {
    s2 hd = as_s2(L.hdiag);
    s2 f = as_s2(fin), hu = as_s2(hup);
    uint32_t a12 = 0, a43 = 0;
    // halves whose state this step updates
    const uint32_t keep = (aa ? 0xFFFFu : 0u) | (ab ? 0xFFFF0000u : 0u);
#pragma unroll
    for (int i = 0; i < K; ++i) {
      // row i's score bytes of the two halves, zero-extended into them: byte
      // i&3 of word i/4 (table), or byte i/2 of word i&1 (profile words split
      // into even and odd rows)
      const int bi = NIB ? (i >> 1) : (i & 3);
      const bool hiw = NIB ? (i & 1) != 0 : i >= 4;
      const uint32_t sel = 0x0C000C00u | ((uint32_t)(4 + bi) << 16) | (uint32_t)bi;
      const s2 sb = as_s2(__builtin_amdgcn_perm(hiw ? wb.y : wb.x, hiw ? wa.y : wa.x, sel));
      const s2 eo = L.E[i], ho = L.Hm[i];
      const s2 e = pmax(eo, ho);
      const s2 s3 = eo - ho;
      const s2 s4 = f - hu;
      f = pmax(f, hu);
      const s2 dg = hd + sb;
      const s2 t1 = pmax(e, f);
      const s2 s1 = dg - t1;
      const s2 sf = f - e;
      const s2 h = pmax(t1, dg) - G2;
      // sign bits (bits 15 and 31) into bit planes: bytes (s1 a, s2 a, s1 b,
      // s2 b) and (s4 a, s3 a, s4 b, s3 b); after the K rows row i's bit sits
      // at bit 8 - K + i of each byte
      const uint32_t w12 = __builtin_amdgcn_perm(as_u(sf), as_u(s1), 0x07030501u);
      const uint32_t w43 = __builtin_amdgcn_perm(as_u(s3), as_u(s4), 0x07030501u);
      a12 = bfi_u(0x80808080u, w12, a12 >> 1);
      a43 = bfi_u(0x80808080u, w43, a43 >> 1);
      hd = ho;
      L.Hm[i] = PRED ? as_s2(bfi_u(keep, as_u(h), as_u(ho))) : h;
      L.E[i] = PRED ? as_s2(bfi_u(keep, as_u(e), as_u(eo))) : e;
      hu = h;
    }
    L.Fout = PRED ? bfi_u(keep, as_u(f), L.Fout) : as_u(f);
    L.Hout = PRED ? bfi_u(keep, as_u(hu), L.Hout) : as_u(hu);
    if (K < 8) {  // row i to bit i of each byte
      a12 = (a12 >> (8 - K)) & (0x01010101u * ((1u << K) - 1u));
      a43 = (a43 >> (8 - K)) & (0x01010101u * ((1u << K) - 1u));
    }
    uint32_t fa = __builtin_amdgcn_perm(a43, a12, 0x05040100u);
    uint32_t fb = __builtin_amdgcn_perm(a43, a12, 0x07060302u);
    if (c == 0) fa |= 0xFF000000u;  // contLeft = 0 in column 0
    if (cb == 0) fb |= 0xFF000000u;
    if (lane == 0) {  // contUp = 0 in row 0 of a job's first strip
      fa |= S.cua;
      fb |= S.cub;
    }
    if (aa) S.fa[t * MST_WAVE + lane] = fa;
    if (ab) S.fb[(t - S.db) * MST_WAVE + lane] = fb;
    // the row below the pass: half b's last row (half a's of two jobs too)
    if (STORE && lane == MST_WAVE - 1 && cb >= 0) bout[cb] = make_int2((int)as_u(f), (int)as_u(hu));
  }
  L.hdiag = hup;
}

// Table words of the block of steps t0 .. t0+FU-1 (clamped into each strip).
__device__ __forceinline__ void f16_load(int t0, const F16Strip& S, uint2* wa, uint2* wb) {
  const int lane = lane_id();
#pragma unroll
  for (int j = 0; j < FU; ++j) {
    const int t = t0 + j, tbb = t - S.db;
    const int tta = t < S.tna ? t : S.tna - 1;
    const int ttb = tbb < 0 ? 0 : (tbb < S.tnb ? tbb : S.tnb - 1);
    wa[j] = S.ta[tta * MST_WAVE + lane];
    wb[j] = S.tb[ttb * MST_WAVE + lane];
  }
}

// One pass. The row buffer (F and H per column of the row above, written by
// the previous pass) is read 64 columns at a time, one per lane, and handed
// to lane 0 one step at a time by a lane read.
template <int K, bool READ, bool STORE>
__device__ void f16_pass(F16Lane<K>& L, const F16Strip& S, const int2* bin, int2* bout, s2 G2) {
  const int lane = lane_id();
  const int T = S.n + MST_WAVE - 1;
  const int bn = S.na > 0 ? S.na : 1;  // columns of the row buffer (half a's)
  uint2 wa0[FU], wb0[FU], wa1[FU], wb1[FU];
  int2 r0 = make_int2(0, 0), r1 = make_int2(0, 0);  // the row buffer: this and the next 64 columns
  if (READ) {
    r0 = bin[lane < bn ? lane : bn - 1];
    r1 = bin[MST_WAVE + lane < bn ? MST_WAVE + lane : bn - 1];
  }
  f16_load(0, S, wa0, wb0);
  for (int t0 = 0; t0 < T; t0 += FU) {
    f16_load(t0 + FU, S, wa1, wb1);
    const bool full =
        (t0 >= MST_WAVE - 1 + S.db) && (t0 + FU <= S.na) && (t0 + FU <= S.nb + S.db);
    const int q = t0 & (MST_WAVE - 1);  // lane of the row-buffer word of step t0
    if (full) {
#pragma unroll
      for (int j = 0; j < FU; ++j)
        f16_step<K, false, STORE>(L, t0 + j, wa0[j], wb0[j],
                                  READ ? (uint32_t)__builtin_amdgcn_readlane(r0.x, q + j) : 0u,
                                  READ ? (uint32_t)__builtin_amdgcn_readlane(r0.y, q + j) : 0u, S, G2, bout);
    } else {
#pragma unroll
      for (int j = 0; j < FU; ++j)
        if (t0 + j < T)
          f16_step<K, true, STORE>(L, t0 + j, wa0[j], wb0[j],
                                   READ ? (uint32_t)__builtin_amdgcn_readlane(r0.x, q + j) : 0u,
                                   READ ? (uint32_t)__builtin_amdgcn_readlane(r0.y, q + j) : 0u, S, G2, bout);
    }
#pragma unroll
    for (int j = 0; j < FU; ++j) {
      wa0[j] = wa1[j];
      wb0[j] = wb1[j];
    }
    if (READ && q + FU == MST_WAVE) {  // next 64 columns of the row buffer
      r0 = r1;
      const int c = t0 + FU + MST_WAVE + lane;
      r1 = bin[c < bn ? c : bn - 1];
    }
  }
}

// Profile mode. The residues lane l needs in the block of steps t0 .. t0+7
// of a half (columns t0-l .. t0+7-l of its job, clamped into 0 .. n-1; none
// when the half is idle), one per byte.
__device__ __forceinline__ uint2 f16_codes(const int8_t* __restrict__ X, int t0, int n) {
  if (n <= 0) return make_uint2(0u, 0u);
  const int c0 = t0 - lane_id();
  uint2 v;
  if (c0 >= 0 && c0 + 8 <= n) {
    __builtin_memcpy(&v, X + c0, sizeof v);
  } else {
    uint32_t q[2] = {0u, 0u};
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      int c = c0 + j;
      c = c < 0 ? 0 : (c >= n ? n - 1 : c);
      q[j >> 2] |= (uint32_t)(uint8_t)X[c] << (8 * (j & 3));
    }
    v = make_uint2(q[0], q[1]);
  }
  return v;
}

__device__ __forceinline__ int f16_code_at(uint2 v, int j) {
  return (int)(((j < 4 ? v.x : v.y) >> (8 * (j & 3))) & 0xFFu);
}

// A profile word as two words of bytes S + go + ge: rows 0, 2, 4, 6 and rows
// 1, 3, 5, 7 (nibble plus the offset smin + go + ge, at most 255).
__device__ __forceinline__ uint2 f16_unpack(uint32_t w, uint32_t off4) {
  return make_uint2((w & 0x0F0F0F0Fu) + off4, ((w >> 4) & 0x0F0F0F0Fu) + off4);
}

// The strip's query profile: word [x][lane] holds, at bits 4i..4i+3, S[Y[r]][x]
// - smin of the lane's row r = r0 + K lane + i (0 past the job's m rows, a row
// below every real one). Each lane writes and later reads only its own words.
template <int K>
__device__ __forceinline__ void f16_build_prof(uint32_t* prof, const int8_t* Sn, const int8_t* Y, int m, int r0) {
  const int lane = lane_id();
  int yr[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int r = r0 + lane * K + i;
    yr[i] = r < m ? (int)Y[r] * MST_LETTERS : -1;
  }
  for (int x = 0; x < MST_LETTERS; ++x) {
    uint32_t w = 0;
#pragma unroll
    for (int i = 0; i < K; ++i)
      if (yr[i] >= 0) w |= (uint32_t)(uint8_t)Sn[yr[i] + x] << (4 * i);
    prof[x * MST_WAVE + lane] = w;
  }
}

// One pass in profile mode: blocks of 8 steps; each step's profile words are
// read from LDS at the block's start, the next block's residues loaded
// meanwhile.
template <int K, bool READ, bool STORE>
__device__ __forceinline__ void f16_pass_prof(F16Lane<K>& L, const F16Strip& S, const int2* bin, int2* bout, s2 G2,
                              const uint32_t* profA, const uint32_t* profB, uint32_t off4) {
  constexpr int U = 4;
  const int lane = lane_id();
  const int T = S.n + MST_WAVE - 1;
  const int bn = S.na > 0 ? S.na : 1;  // columns of the row buffer (half a's)
  int2 r0 = make_int2(0, 0), r1 = make_int2(0, 0);
  if (READ) {
    r0 = bin[lane < bn ? lane : bn - 1];
    r1 = bin[MST_WAVE + lane < bn ? MST_WAVE + lane : bn - 1];
  }
  uint2 ca = f16_codes(S.xa, 0, S.na), cb = f16_codes(S.xb, -S.db, S.nb);
  for (int t0 = 0; t0 < T; t0 += U) {
    uint32_t wa[U], wb[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      wa[j] = profA[f16_code_at(ca, j) * MST_WAVE + lane];
      wb[j] = profB[f16_code_at(cb, j) * MST_WAVE + lane];
    }
    ca = f16_codes(S.xa, t0 + U, S.na);
    cb = f16_codes(S.xb, t0 + U - S.db, S.nb);
    const bool full =
        (t0 >= MST_WAVE - 1 + S.db) && (t0 + U <= S.na) && (t0 + U <= S.nb + S.db);
    const int q = t0 & (MST_WAVE - 1);  // lane of the row-buffer word of step t0
    if (full) {
#pragma unroll
      for (int j = 0; j < U; ++j)
        f16_step<K, false, STORE, true>(L, t0 + j, f16_unpack(wa[j], off4), f16_unpack(wb[j], off4),
                                        READ ? (uint32_t)__builtin_amdgcn_readlane(r0.x, q + j) : 0u,
                                        READ ? (uint32_t)__builtin_amdgcn_readlane(r0.y, q + j) : 0u, S, G2,
                                        bout);
    } else {
#pragma unroll
      for (int j = 0; j < U; ++j)
        if (t0 + j < T)
          f16_step<K, true, STORE, true>(L, t0 + j, f16_unpack(wa[j], off4), f16_unpack(wb[j], off4),
                                         READ ? (uint32_t)__builtin_amdgcn_readlane(r0.x, q + j) : 0u,
                                         READ ? (uint32_t)__builtin_amdgcn_readlane(r0.y, q + j) : 0u, S, G2,
                                         bout);
    }
    if (READ && q + U == MST_WAVE) {  // next 64 columns of the row buffer
      r0 = r1;
      const int c = t0 + U + MST_WAVE + lane;
      r1 = bin[c < bn ? c : bn - 1];
    }
  }
}

// H of row idx of the lane's K (low half: half a, high half: half b), as an
// or of masked rows (a select chain would let the compiler index the state).
template <int K>
__device__ __forceinline__ int f16_pick(const F16Lane<K>& L, int idx, bool hi) {
  uint32_t v = 0;
#pragma unroll
  for (int i = 0; i < K; ++i) v |= as_u(L.Hm[i]) & (0u - (uint32_t)(idx == i));
  return hi ? (int)(short)(v >> 16) : (int)(short)(v & 0xFFFFu);
}

// One item on the calling wave: two jobs of K rows per lane, strips in step
// (mode 0; b < 0: job a alone), or one job two strips per pass (mode 1).
template <int K, bool PROF>
__device__ __forceinline__ void f16_item(const F16Args& a, const F16Item item, int2* buf, s2 G2, uint32_t* profA,
                         uint32_t* profB, const int8_t* Sn) {
  constexpr int FSH = MST_WAVE * K;
  const int lane = lane_id();
  const AlignJob ja = a.jobs[item.a];
  const bool pair = item.mode == 1;
  const bool two = !pair && item.b >= 0;
  const AlignJob jb = two ? a.jobs[item.b] : ja;
  const int sa = (ja.m + FSH - 1) / FSH, sb = two ? (jb.m + FSH - 1) / FSH : 0;
  const int np = pair ? (sa + 1) / 2 : (sa > sb ? sa : sb);
  const int Tn = ja.n + MST_WAVE - 1, Tnb = jb.n + MST_WAVE - 1;
  F16Lane<K> L;
  for (int p = 0; p < np; ++p) {
    // strips of the halves (-1: idle)
    const int xa = pair ? 2 * p : (p < sa ? p : -1);
    const int xb = pair ? (2 * p + 1 < sa ? 2 * p + 1 : -1) : (p < sb ? p : -1);
    F16Strip S;
    S.db = pair ? MST_WAVE : 0;
    S.na = xa >= 0 ? ja.n : 0;
    S.nb = xb >= 0 ? jb.n : 0;
    S.n = S.na > S.nb + S.db ? S.na : S.nb + S.db;
    S.tna = Tn;
    S.tnb = Tnb;
    // an idle half reads (and ignores) a strip of its job
    const int qa = xa >= 0 ? xa : sa - 1, qb = xb >= 0 ? xb : (pair ? sa - 1 : (sb > 0 ? sb - 1 : 0));
    S.ta = (const uint2*)(a.tab + ja.tab_off) + (int64_t)qa * Tn * MST_WAVE;
    S.tb = (const uint2*)(a.tab + jb.tab_off) + (int64_t)qb * Tnb * MST_WAVE;
    S.fa = a.flags + ja.flag_off + (int64_t)qa * Tn * MST_WAVE;
    S.fb = a.flags + jb.flag_off + (int64_t)qb * Tnb * MST_WAVE;
    S.sel = pair ? (p == 0 ? kSelFirstPair : kSelPair) : (p == 0 ? kSelZero : kSelBuf);
    S.cua = xa == 0 ? 0x00010000u : 0u;
    S.cub = xb == 0 ? 0x00010000u : 0u;
    S.xa = PROF ? a.seqs + ja.x_off : nullptr;
    S.xb = PROF ? a.seqs + jb.x_off : nullptr;
    if constexpr (PROF) {
      if (xa >= 0) f16_build_prof<K>(profA, Sn, a.seqs + ja.y_off, ja.m, xa * FSH);
      if (xb >= 0) f16_build_prof<K>(profB, Sn, a.seqs + jb.y_off, jb.m, xb * FSH);
    }
#pragma unroll
    for (int i = 0; i < K; ++i) {
      L.Hm[i] = (s2){0, 0};
      L.E[i] = (s2){0, 0};
    }
    L.Fout = L.Hout = L.hdiag = 0u;
    const int2* bin = buf + (int64_t)(p & 1) * a.bufcap;
    int2* bout = buf + (int64_t)((p + 1) & 1) * a.bufcap;
    const bool more = p + 1 < np;
    if constexpr (PROF) {
      const uint32_t off4 = 0x01010101u * (uint32_t)(a.smin + a.go + a.ge);
      if (p == 0) {
        if (more)
          f16_pass_prof<K, false, true>(L, S, bin, bout, G2, profA, profB, off4);
        else
          f16_pass_prof<K, false, false>(L, S, bin, bout, G2, profA, profB, off4);
      } else {
        if (more)
          f16_pass_prof<K, true, true>(L, S, bin, bout, G2, profA, profB, off4);
        else
          f16_pass_prof<K, true, false>(L, S, bin, bout, G2, profA, profB, off4);
      }
    } else {
      if (p == 0) {
        if (more)
          f16_pass<K, false, true>(L, S, bin, bout, G2);
        else
          f16_pass<K, false, false>(L, S, bin, bout, G2);
      } else {
        if (more)
          f16_pass<K, true, true>(L, S, bin, bout, G2);
        else
          f16_pass<K, true, false>(L, S, bin, bout, G2);
      }
    }
    if (more) agent_fence();
    // scores from the lanes holding the last rows
    if (xa == sa - 1) {
      const int rl = ja.m - 1 - xa * FSH;
      const int v = __shfl(f16_pick<K>(L, rl % K, false), rl / K);
      if (lane == 0) a.scores[item.a] = v - a.ge * (ja.m + ja.n);
    }
    if (xb >= 0 && xb == (pair ? sa : sb) - 1) {
      const int rl = jb.m - 1 - xb * FSH;
      const int v = __shfl(f16_pick<K>(L, rl % K, true), rl / K);
      if (lane == 0) a.scores[pair ? item.a : item.b] = v - a.ge * (jb.m + jb.n);
    }
  }
}

// Items of every rows-per-lane class in one launch (their order is the
// host's, largest first), each wave taking the next item from the queue as it
// finishes one (or round robin without a queue). PROF: the
// strips' profiles in LDS (12 KB: twelve waves per compute unit).
template <bool PROF>
__global__ void __launch_bounds__(MST_WAVE) flag16_kernel(F16Args a) {
  issue_first();
  __shared__ uint32_t prof[PROF ? 2 * MST_LETTERS * MST_WAVE : 1];
  __shared__ int8_t Sn[PROF ? MST_LETTERS * MST_LETTERS : 1];
  if (PROF)
    for (int i = threadIdx.x; i < MST_LETTERS * MST_LETTERS; i += MST_WAVE)
      Sn[i] = (int8_t)(a.subst[i] - a.smin);
  __syncthreads();
  uint32_t* profA = prof;
  uint32_t* profB = prof + (PROF ? MST_LETTERS * MST_WAVE : 0);
  const s2 G2 = (s2){(short)(a.go - a.ge), (short)(a.go - a.ge)};
  int2* buf = a.bufs + (int64_t)blockIdx.x * 2 * a.bufcap;
  for (int64_t it = a.counter ? (int64_t)__builtin_amdgcn_readfirstlane(
                                    threadIdx.x == 0 ? (int)atomicAdd(a.counter, 1u) : 0)
                              : (int64_t)blockIdx.x;
       it < a.nitems;) {
    const F16Item item = a.items[it];
    switch (item.k) {
      case 1: f16_item<1, PROF>(a, item, buf, G2, profA, profB, Sn); break;
      case 2: f16_item<2, PROF>(a, item, buf, G2, profA, profB, Sn); break;
      case 3: f16_item<3, PROF>(a, item, buf, G2, profA, profB, Sn); break;
      case 4: f16_item<4, PROF>(a, item, buf, G2, profA, profB, Sn); break;
      case 5: f16_item<5, PROF>(a, item, buf, G2, profA, profB, Sn); break;
      case 6: f16_item<6, PROF>(a, item, buf, G2, profA, profB, Sn); break;
      case 7: f16_item<7, PROF>(a, item, buf, G2, profA, profB, Sn); break;
      default: f16_item<8, PROF>(a, item, buf, G2, profA, profB, Sn); break;
    }
    it = a.counter ? (int64_t)__builtin_amdgcn_readfirstlane(threadIdx.x == 0 ? (int)atomicAdd(a.counter, 1u) : 0)
                   : it + gridDim.x;
  }
}

void plan_flag16_items(const AlignJob* jobs, int64_t n, std::vector<F16Item>& items) {
  items.clear();
  std::vector<int64_t> work;
  for (int64_t k = 0; k < n;) {
    const int K = jobs[k].k;
    const int64_t sh = 64 * K;
    if (jobs[k].m > sh) {  // two strips: one item, a strip per half
      items.push_back(F16Item{(int)k, -1, K, 1});
      work.push_back((int64_t)(jobs[k].n + 127) * sh);
      k += 1;
      continue;
    }
    const bool pair = k + 1 < n && jobs[k + 1].k == K && jobs[k + 1].m <= sh;
    items.push_back(F16Item{(int)k, pair ? (int)(k + 1) : -1, K, 0});
    work.push_back((int64_t)(std::max(jobs[k].n, pair ? jobs[k + 1].n : 0) + 63) * sh);
    k += pair ? 2 : 1;
  }
  std::vector<size_t> ord(items.size());
  for (size_t q = 0; q < ord.size(); ++q) ord[q] = q;
  std::stable_sort(ord.begin(), ord.end(), [&](size_t x, size_t y) { return work[x] > work[y]; });
  std::vector<F16Item> by_work(items.size());
  for (size_t q = 0; q < ord.size(); ++q) by_work[q] = items[ord[q]];
  items.swap(by_work);
}

void launch_flag16(const F16Args& a, int grid, hipStream_t st) {
  if (a.nitems <= 0) return;
  if (a.prof)
    hipLaunchKernelGGL(flag16_kernel<true>, dim3(grid), dim3(MST_WAVE), 0, st, a);
  else
    hipLaunchKernelGGL(flag16_kernel<false>, dim3(grid), dim3(MST_WAVE), 0, st, a);
}

// The score table of sequence jobs in the flag kernels' layout: byte i of
// the word of (strip s, step t, lane l) is S[Y[r]][X[c]] + go + ge for row
// r = 64 k s + k l + i (i < k, the job's rows per lane) and column c = t - l
// (0 outside the DP). A block per (job, strip) item: the strip's query
// profile (the 8 bytes of a lane for each of the 24 residues) is built in LDS
// once, then a step's word is one LDS read at its column's residue.
__global__ void __launch_bounds__(256) seq_table_kernel(const AlignJob* __restrict__ jobs,
                                                        const int2* __restrict__ strips,
                                                        const int8_t* __restrict__ seqs,
                                                        const int8_t* __restrict__ subst, int bonus,
                                                        uint8_t* __restrict__ tab) {
  issue_first();
  __shared__ int8_t S_lds[MST_LETTERS * MST_LETTERS];
  __shared__ uint2 prof[MST_LETTERS][MST_WAVE];
  for (int i = threadIdx.x; i < MST_LETTERS * MST_LETTERS; i += blockDim.x) S_lds[i] = subst[i];
  const int2 js = strips[blockIdx.x];
  const AlignJob job = jobs[js.x];
  const int s = js.y;
  const int n = job.n, m = job.m, Tn = n + MST_WAVE - 1, k = job.k;
  const int8_t* X = seqs + job.x_off;
  const int8_t* Y = seqs + job.y_off;
  __syncthreads();
  for (int e = threadIdx.x; e < MST_LETTERS * MST_WAVE; e += blockDim.x) {
    const int x = e / MST_WAVE, l = e % MST_WAVE;
    uint32_t w[2] = {0u, 0u};
#pragma unroll
    for (int i = 0; i < FK; ++i) {
      const int r = s * MST_WAVE * k + l * k + i;
      if (i < k && r < m)
        w[i >> 2] |= ((uint32_t)(S_lds[(int)Y[r] * MST_LETTERS + x] + bonus) & 0xFFu) << (8 * (i & 3));
    }
    prof[x][l] = make_uint2(w[0], w[1]);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, sub = threadIdx.x >> 6;
  uint2* out = (uint2*)(tab + job.tab_off) + (int64_t)s * Tn * MST_WAVE;
  for (int t = sub; t < Tn; t += 4) {
    const int c = t - lane;
    out[(int64_t)t * MST_WAVE + lane] = (c >= 0 && c < n) ? prof[X[c]][lane] : make_uint2(0u, 0u);
  }
}

void launch_seq_table(const AlignJob* jobs, const int2* strips, int nstrips, const int8_t* seqs,
                      const int8_t* subst, int bonus, uint8_t* tab, hipStream_t st) {
  if (nstrips <= 0) return;
  hipLaunchKernelGGL(seq_table_kernel, dim3(nstrips), dim3(256), 0, st, jobs, strips, seqs, subst, bonus, tab);
}

}  // namespace mst

// Alignment with traceback (R2 + R3): NeedlemanWunschGlobalMatchKernel
// (NeedlemanWunsch.cu:184-365) and NeedlemanWunschGlobalBackKernel (:393-494).
//
// The forward pass is the score kernel's wavefront (one pair per wave, lane l
// owns K = 8 consecutive rows of a 512-row strip, columns skewed by one step
// per lane) with the reference's per-cell nibble added:
//   bit3 LEFT  = (H==E && H!=F) || H==D     bit1 contUp   = F came from F-ge
//   bit2 UP    = (H==F) || H==D             bit0 contLeft = E came from E-ge
// (diagonal before up before left; extension preferred on a tie). In the
// biased max-plus form of the score kernel these become two compares per
// direction and one per continuation flag:
//   eqD = D >= max(E,F), fe = F >= E, UP = eqD|fe, LEFT = eqD|!fe,
//   contUp = Fin >= Hm_up, contLeft = E_left >= Hm_left.
// Unlike the score kernel, E at column -1 and F at row -1 are a true -inf so
// that the continuation flags of the first row and column are 0, as with the
// reference's SHORT_MIN+gapEx sentinel.
//
// Flag layout: each lane packs its K=8 nibbles of one step into one dword
// (row i of the lane at bits 4i..4i+3) and the wave stores the 64 dwords of a
// step contiguously: flags[(strip*(n+63) + t)*64 + lane]. A step is one
// coalesced 256-byte store, 0.5 byte per cell. The traceback reads single
// nibbles at random and walks the reference's state machine (:409-482) from
// the last cell of the reversed inputs, i.e. from the first residues of the
// forward sequences, emitting the path in forward order. Going up into row -1
// continues left along row -1 to the corner (the reference stops there early).
//
// Score sources: SRC_SEQ looks the substitution scores up in a per-wave LDS
// query profile (pairwise alignment); SRC_TAB reads them from a precomputed
// byte table in the same skewed order (profile-profile alignment).
#include "common.h"
#include "nw_align.h"

namespace mst {

constexpr int AK = 8;                 // rows per lane
constexpr int ASH = MST_WAVE * AK;    // rows per strip
constexpr int AU = 8;                 // unrolled steps / prefetch distance

__device__ __forceinline__ int shr1_const(int v, int c) {
  return __builtin_amdgcn_update_dpp(c, v, DPP_WAVE_SHR1, 0xF, 0xF, false);
}

struct ALane {
  int Hm[AK], E[AK];
  int Fout, Hout, hdiag;
};

__device__ __forceinline__ void build_profile8(uint8_t* prof, const int8_t* S_lds, const int8_t* Y,
                                               int m, int r0, int bonus) {
  const int lane = lane_id();
  int yrow[AK];
#pragma unroll
  for (int i = 0; i < AK; ++i) {
    int r = r0 + lane * AK + i;
    yrow[i] = (r < m) ? (int)Y[r] * MST_LETTERS : 0;
  }
  for (int x = 0; x < MST_LETTERS; ++x) {
    uint32_t w[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      uint32_t v = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        int s = (int)S_lds[yrow[q * 4 + b] + x] + bonus;
        v |= ((uint32_t)s & 0xFFu) << (8 * b);
      }
      w[q] = v;
    }
    *(uint2*)(prof + (x * MST_WAVE + lane) * AK) = make_uint2(w[0], w[1]);
  }
}

// Column c of the lane's 8 rows; returns the packed nibbles.
__device__ __forceinline__ uint32_t cell_column_flags(uint32_t w0, uint32_t w1, ALane& L, int& f,
                                                      int& hu, int G) {
  int hd = L.hdiag;
  uint32_t word = 0;
#pragma unroll
  for (int i = 0; i < AK; ++i) {
    const uint32_t wb = (i < 4) ? w0 : w1;
    const int sb = (int)(int8_t)((wb >> (8 * (i & 3))) & 0xFFu);
    const int eo = L.E[i], ho = L.Hm[i];
    const int e = imax(eo, ho);
    const uint32_t cl = eo >= ho;
    const uint32_t cu = f >= hu;
    f = imax(f, hu);
    const int dg = hd + sb;
    const int t1 = imax(e, f);
    const uint32_t eqD = dg >= t1;
    const uint32_t fe = f >= e;
    const int h = imax(t1, dg) - G;
    const uint32_t up = eqD | fe;
    const uint32_t left = eqD | (fe ^ 1u);
    word |= ((left << 3) | (up << 2) | (cu << 1) | cl) << (4 * i);
    hd = ho;
    L.Hm[i] = h;
    L.E[i] = e;
    hu = h;
  }
  return word;
}

// One step of a strip: lane l computes column t - l of its 8 rows when that
// column is in [lo, hi) (PRED), reading F/H of the row above from lane l-1 or,
// for lane 0, from the strip above (bv) or the top boundary (FIRST).
template <int SRC, bool PRED, bool FIRST, bool STORE>
__device__ __forceinline__ void astep(ALane& L, int t, int x, uint2 tb, int2 bv,
                                      const uint8_t* prof_lane, int lo, int hi, int G, int2* bout,
                                      uint32_t* fl_strip) {
  const int lane = lane_id();
  int fin, hup;
  if (FIRST) {
    fin = shr1_const(L.Fout, MST_NEG_INF);
    hup = shr1_const(L.Hout, 0);
  } else {
    fin = shr1(bv.x, L.Fout);
    hup = shr1(bv.y, L.Hout);
  }
  const int c = t - lane;
  const bool act = !PRED || (c >= lo && c < hi);
  if (act) {
    uint32_t w0, w1;
    if (SRC == SRC_SEQ) {
      uint2 v = *(const uint2*)(prof_lane + x * (MST_WAVE * AK));
      w0 = v.x;
      w1 = v.y;
    } else {
      w0 = tb.x;
      w1 = tb.y;
    }
    int f = fin, hu = hup;
    const uint32_t word = cell_column_flags(w0, w1, L, f, hu, G);
    L.Fout = f;
    L.Hout = hu;
    fl_strip[(int64_t)t * MST_WAVE + lane] = word;
    if (STORE && lane == MST_WAVE - 1) bout[c] = make_int2(f, hu);
  }
  L.hdiag = hup;
}

template <int SRC, bool FIRST>
__device__ __forceinline__ void aload(int t0, const int8_t* __restrict__ X, int n, const int2* bin,
                                      const uint2* tab_strip, int* xs, int2* bs, uint2* ts) {
  const int lane = lane_id();
#pragma unroll
  for (int j = 0; j < AU; ++j) {
    const int t = t0 + j;
    int c = t - lane;
    c = c < 0 ? 0 : (c >= n ? n - 1 : c);
    if (SRC == SRC_SEQ) {
      xs[j] = X[c];
      ts[j] = make_uint2(0, 0);
    } else {
      xs[j] = 0;
      int tt = t < n + MST_WAVE - 1 ? t : n + MST_WAVE - 2;
      ts[j] = tab_strip[(int64_t)tt * MST_WAVE + lane];
    }
    if (!FIRST) {
      int tt = t >= n ? n - 1 : t;
      bs[j] = bin[tt];
    } else {
      bs[j] = make_int2(0, 0);
    }
  }
}

template <int SRC, bool FIRST, bool STORE>
__device__ void arun_strip(ALane& L, const int8_t* X, int n, const uint8_t* prof_lane,
                           const uint2* tab_strip, const int2* bin, int2* bout, int G,
                           uint32_t* fl_strip) {
  const int T = n + MST_WAVE - 1;
  int xa[AU], xb[AU];
  int2 ba[AU], bb[AU];
  uint2 ta[AU], tbb[AU];
  aload<SRC, FIRST>(0, X, n, bin, tab_strip, xa, ba, ta);
  for (int t0 = 0; t0 < T; t0 += AU) {
    aload<SRC, FIRST>(t0 + AU, X, n, bin, tab_strip, xb, bb, tbb);
    const bool full = (t0 >= MST_WAVE - 1) && (t0 + AU <= n);
    if (full) {
#pragma unroll
      for (int j = 0; j < AU; ++j)
        astep<SRC, false, FIRST, STORE>(L, t0 + j, xa[j], ta[j], ba[j], prof_lane, 0, n, G, bout,
                                        fl_strip);
    } else {
#pragma unroll
      for (int j = 0; j < AU; ++j)
        if (t0 + j < T)
          astep<SRC, true, FIRST, STORE>(L, t0 + j, xa[j], ta[j], ba[j], prof_lane, 0, n, G, bout,
                                         fl_strip);
    }
#pragma unroll
    for (int j = 0; j < AU; ++j) {
      xa[j] = xb[j];
      ba[j] = bb[j];
      ta[j] = tbb[j];
    }
  }
}

// ---- tiled form: an inter-block wavefront for long jobs ----
//
// A job is cut into tiles of one strip (512 rows) by `tw` columns; tile
// (s, c) needs the last row of tile (s-1, c) (F and H per column, a row
// buffer per strip parity) and the last column of tile (s, c-1) (H and E of
// each of the strip's rows, the lane states of strip s). Launch d runs every
// tile with s + c = d, one wave each, so a job of S strips and C column tiles
// takes S + C - 1 launches of up to min(S, C) waves, ordered by the stream
// (no waits between workgroups). The flags are the one-strip kernel's: lane l
// of tile (s, c) writes step t = column + l of strip s for its columns only.

// Columns [c0, c1) of a strip, global steps c0 .. c1 + 62.
template <int SRC, bool FIRST, bool STORE>
__device__ void arun_tile(ALane& L, const int8_t* X, int n, int c0, int c1, const uint8_t* prof_lane,
                          const uint2* tab_strip, const int2* bin, int2* bout, int G,
                          uint32_t* fl_strip) {
  const int T1 = c1 + MST_WAVE - 1;
  int xa[AU], xb[AU];
  int2 ba[AU], bb[AU];
  uint2 ta[AU], tbb[AU];
  aload<SRC, FIRST>(c0, X, n, bin, tab_strip, xa, ba, ta);
  for (int t0 = c0; t0 < T1; t0 += AU) {
    aload<SRC, FIRST>(t0 + AU, X, n, bin, tab_strip, xb, bb, tbb);
    const bool full = (t0 >= c0 + MST_WAVE - 1) && (t0 + AU <= c1);
    if (full) {
#pragma unroll
      for (int j = 0; j < AU; ++j)
        astep<SRC, false, FIRST, STORE>(L, t0 + j, xa[j], ta[j], ba[j], prof_lane, c0, c1, G, bout,
                                        fl_strip);
    } else {
#pragma unroll
      for (int j = 0; j < AU; ++j)
        if (t0 + j < T1)
          astep<SRC, true, FIRST, STORE>(L, t0 + j, xa[j], ta[j], ba[j], prof_lane, c0, c1, G, bout,
                                         fl_strip);
    }
#pragma unroll
    for (int j = 0; j < AU; ++j) {
      xa[j] = xb[j];
      ba[j] = bb[j];
      ta[j] = tbb[j];
    }
  }
}

template <int SRC>
__global__ void __launch_bounds__(MST_WAVE) align_tile_kernel(TileArgs a) {
  issue_first();
  __shared__ int8_t S_lds[MST_LETTERS * MST_LETTERS];
  __shared__ __attribute__((aligned(16))) uint8_t prof[SRC == SRC_SEQ ? MST_LETTERS * MST_WAVE * AK : 16];
  const int lane = lane_id();
  const TileJob tile = a.tiles[blockIdx.x];
  const AlignJob job = a.jobs[tile.job];
  const int n = job.n, m = job.m;
  const int s = tile.s;
  const int c0 = tile.c * a.tw;
  const int c1 = c0 + a.tw < n ? c0 + a.tw : n;
  const int Tn = n + MST_WAVE - 1;
  const int nstrips = (m + ASH - 1) / ASH;
  const int r0 = s * ASH;
  const int G = a.go - a.ge;
  const int8_t* X = a.seqs + job.x_off;
  if (SRC == SRC_SEQ) {
    for (int i = lane; i < MST_LETTERS * MST_LETTERS; i += MST_WAVE) S_lds[i] = a.subst[i];
    __syncthreads();
    build_profile8(prof, S_lds, a.seqs + job.y_off, m, r0, a.go + a.ge);
    __syncthreads();
  }
  // lane state at column c0 - 1: H and E of the lane's rows (the left
  // boundary: H = 0 in the biased form, E = -inf); Hout starts as H of the
  // lane's last row there, which lane l+1 takes as its first diagonal. Lane
  // 0's first diagonal, H of the strip above at c0 - 1, was saved by tile
  // (s, c-1): the row buffer slot itself may be rewritten in this launch by
  // tile (s+1, c-1), which shares its parity.
  int* sblk = a.state + tile.st + (int64_t)s * kTileStateInts;
  int* st = sblk + lane * (2 * AK);
  int2* rows = a.rows + tile.rb;
  const int2* bin = s > 0 ? rows + (int64_t)((s - 1) & 1) * n : rows;
  int2* bout = rows + (int64_t)(s & 1) * n;
  ALane L;
  if (c0 == 0) {
#pragma unroll
    for (int i = 0; i < AK; ++i) {
      L.Hm[i] = 0;
      L.E[i] = MST_NEG_INF;
    }
  } else {
#pragma unroll
    for (int i = 0; i < AK; ++i) {
      L.Hm[i] = st[i];
      L.E[i] = st[AK + i];
    }
  }
  L.Fout = MST_NEG_INF;
  L.Hout = L.Hm[AK - 1];
  L.hdiag = (c0 > 0 && s > 0) ? sblk[MST_WAVE * 2 * AK] : 0;
  uint32_t* fl_strip = a.flags + job.flag_off + (int64_t)s * Tn * MST_WAVE;
  const uint2* tab_strip =
      (SRC == SRC_TAB) ? (const uint2*)(a.tab + job.tab_off) + (int64_t)s * Tn * MST_WAVE : nullptr;
  const uint8_t* prof_lane = prof + lane * AK;
  const bool last = (s == nstrips - 1);
  if (s == 0) {
    if (last)
      arun_tile<SRC, true, false>(L, X, n, c0, c1, prof_lane, tab_strip, bin, bout, G, fl_strip);
    else
      arun_tile<SRC, true, true>(L, X, n, c0, c1, prof_lane, tab_strip, bin, bout, G, fl_strip);
  } else {
    if (last)
      arun_tile<SRC, false, false>(L, X, n, c0, c1, prof_lane, tab_strip, bin, bout, G, fl_strip);
    else
      arun_tile<SRC, false, true>(L, X, n, c0, c1, prof_lane, tab_strip, bin, bout, G, fl_strip);
  }
  if (c1 < n) {
#pragma unroll
    for (int i = 0; i < AK; ++i) {
      st[i] = L.Hm[i];
      st[AK + i] = L.E[i];
    }
    // nobody rewrites this slot during the launch: the other tiles of the
    // parity write columns before c0 or from c1 on
    if (s > 0 && lane == 0) sblk[MST_WAVE * 2 * AK] = bin[c1 - 1].y;
  } else if (last) {
    const int rl = (m - 1) - r0;
    const int owner = rl / AK, idx = rl % AK;
    int v = L.Hm[0];
#pragma unroll
    for (int i = 1; i < AK; ++i)
      if (idx == i) v = L.Hm[i];
    v = __shfl(v, owner);
    if (lane == 0) a.scores[tile.job] = v - a.ge * (m + n);
  }
}

// ---- pipelined form: the strips of one job on the waves of a workgroup ----

constexpr int ARING = 128;  // columns held by an LDS ring between two waves

__device__ __forceinline__ int aflag_load(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void aflag_store(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// every lane of the wave spins on one LDS word; the wave it waits for is in
// the same workgroup, hence resident, and does not wait on this one
__device__ __forceinline__ void await_at_least(const int* p, int v) {
  while (aflag_load(p) < v) __builtin_amdgcn_s_sleep(1);
}

// F and H of the row above a strip, per column: element (base + c) & mask of p
// (an LDS ring: mask ARING-1; the per-workgroup row buffer: mask -1).
struct ALink {
  int2* p;
  int base, mask;
};

template <int SRC, bool PRED, bool FIRST, bool STORE>
__device__ __forceinline__ void astep_link(ALane& L, int t, int x, uint2 tb, int2 bv,
                                           const uint8_t* prof_lane, int n, int G,
                                           const ALink& out, uint32_t* fl_strip) {
  const int lane = lane_id();
  int fin, hup;
  if (FIRST) {
    fin = shr1_const(L.Fout, MST_NEG_INF);
    hup = shr1_const(L.Hout, 0);
  } else {
    fin = shr1(bv.x, L.Fout);
    hup = shr1(bv.y, L.Hout);
  }
  const int c = t - lane;
  const bool act = !PRED || (c >= 0 && c < n);
  if (act) {
    uint32_t w0, w1;
    if (SRC == SRC_SEQ) {
      uint2 v = *(const uint2*)(prof_lane + x * (MST_WAVE * AK));
      w0 = v.x;
      w1 = v.y;
    } else {
      w0 = tb.x;
      w1 = tb.y;
    }
    int f = fin, hu = hup;
    const uint32_t word = cell_column_flags(w0, w1, L, f, hu, G);
    L.Fout = f;
    L.Hout = hu;
    fl_strip[(int64_t)t * MST_WAVE + lane] = word;
    if (STORE && lane == MST_WAVE - 1) out.p[(out.base + c) & out.mask] = make_int2(f, hu);
  }
  L.hdiag = hup;
}

template <int SRC, bool FIRST>
__device__ __forceinline__ void aload_link(int t0, const int8_t* __restrict__ X, int n,
                                           const ALink& in, const uint2* tab_strip, int* xs,
                                           int2* bs, uint2* ts) {
  const int lane = lane_id();
#pragma unroll
  for (int j = 0; j < AU; ++j) {
    const int t = t0 + j;
    int c = t - lane;
    c = c < 0 ? 0 : (c >= n ? n - 1 : c);
    if (SRC == SRC_SEQ) {
      xs[j] = X[c];
      ts[j] = make_uint2(0, 0);
    } else {
      xs[j] = 0;
      int tt = t < n + MST_WAVE - 1 ? t : n + MST_WAVE - 2;
      ts[j] = tab_strip[(int64_t)tt * MST_WAVE + lane];
    }
    if (!FIRST) {
      int tt = t >= n ? n - 1 : t;
      bs[j] = in.p[(in.base + tt) & in.mask];
    } else {
      bs[j] = make_int2(0, 0);
    }
  }
}

// One strip swept by one wave, reading the row above through `in` and writing
// its last row through `out`. Counters count columns in global units: gin +
// column for `in` (its producer's units), gout + column for `out`.
//   prog_in: columns written to `in`; cons_out: columns this wave has read
//   from it; prog_out / cons_in: the same for `out`. span: columns `out`
//   holds before a slot is reused (0: never within a sweep); pub: blocks
//   between two releases of prog_out.
template <int SRC, bool FIRST, bool STORE>
__device__ void arun_strip_pipe(ALane& L, const int8_t* X, int n, const uint8_t* prof_lane,
                                const uint2* tab_strip, const ALink& in, int gin,
                                const int* prog_in, int* cons_out, const ALink& out, int gout,
                                int* prog_out, const int* cons_in, int span, int pub, int G,
                                uint32_t* fl_strip) {
  const int lane = lane_id();
  const int T = n + MST_WAVE - 1;
  int xa[AU], xb[AU];
  int2 ba[AU], bb[AU];
  uint2 ta[AU], tbb[AU];
  if (!FIRST) await_at_least(prog_in, gin + min(AU, n));
  aload_link<SRC, FIRST>(0, X, n, in, tab_strip, xa, ba, ta);
  int blk = 0;
  for (int t0 = 0; t0 < T; t0 += AU, ++blk) {
    if (!FIRST) await_at_least(prog_in, gin + min(t0 + 2 * AU, n));
    aload_link<SRC, FIRST>(t0 + AU, X, n, in, tab_strip, xb, bb, tbb);
    if (!FIRST && lane == 0) aflag_store(cons_out, gin + min(t0 + 2 * AU, n));
    if (STORE && span > 0) {
      // this block writes columns t0-63 .. t0-56, whose slots held the columns
      // span earlier (in global units); the reader must be past them
      const int need = gout + t0 - (MST_WAVE - 1) + (AU - 1) - span + 1;
      if (need > 0) await_at_least(cons_in, min(need, gout + n));
    }
    const bool full = (t0 >= MST_WAVE - 1) && (t0 + AU <= n);
    if (full) {
#pragma unroll
      for (int j = 0; j < AU; ++j)
        astep_link<SRC, false, FIRST, STORE>(L, t0 + j, xa[j], ta[j], ba[j], prof_lane, n, G, out,
                                             fl_strip);
    } else {
#pragma unroll
      for (int j = 0; j < AU; ++j)
        if (t0 + j < T)
          astep_link<SRC, true, FIRST, STORE>(L, t0 + j, xa[j], ta[j], ba[j], prof_lane, n, G, out,
                                              fl_strip);
    }
    if (STORE && (blk % pub) == pub - 1) {
      int done = t0 + AU - (MST_WAVE - 1);
      done = done < 0 ? 0 : (done > n ? n : done);
      if (lane == MST_WAVE - 1) aflag_store(prog_out, gout + done);
    }
#pragma unroll
    for (int j = 0; j < AU; ++j) {
      xa[j] = xb[j];
      ba[j] = bb[j];
      ta[j] = tbb[j];
    }
  }
  if (STORE && lane == MST_WAVE - 1) aflag_store(prog_out, gout + n);
}

// A workgroup of W waves takes one job at a time; wave j computes strips j,
// j+W, j+2W, ... (sweep q = strip / W). Strip s reads the row above from
// strip s-1: wave j-1 in the same sweep through an LDS ring, or, for wave 0,
// wave W-1's previous sweep through the workgroup's row buffer (`carry`).
// Every wait ends: a wave waits only for data of an earlier strip or for a
// later strip to have read a ring slot, and wave W-1 never waits for the
// carry to be read (wave 0 reads its column c before the data of column c of
// the same sweep can reach wave W-1).
template <int SRC, int W>
__global__ void __launch_bounds__(MST_WAVE* W) align_pipe_kernel(AlignArgs a) {
  issue_first();
  constexpr int SLOT = (SRC == SRC_SEQ) ? MST_LETTERS * MST_WAVE * AK : 16;
  __shared__ int8_t S_lds[MST_LETTERS * MST_LETTERS];
  __shared__ __attribute__((aligned(16))) uint8_t prof[W * SLOT];
  __shared__ int2 ring[(W - 1) * ARING];
  __shared__ int s_prog[W + 1], s_cons[W + 1];  // [W]: the carry between sweeps
  __shared__ int s_item;
  const int lane = lane_id();
  const int j = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave index (uniform)
  if (SRC == SRC_SEQ) {
    for (int i = threadIdx.x; i < MST_LETTERS * MST_LETTERS; i += MST_WAVE * W) S_lds[i] = a.subst[i];
  }
  const int G = a.go - a.ge;
  uint8_t* slot = prof + j * SLOT;
  const uint8_t* prof_lane = slot + lane * AK;
  int2* carry = a.bufs + (int64_t)blockIdx.x * a.bufcap;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(a.counter, 1u);
    if (threadIdx.x < W + 1) {
      s_prog[threadIdx.x] = 0;
      s_cons[threadIdx.x] = 0;
    }
    __syncthreads();
    const int64_t it = (int64_t)s_item + a.first;
    if (it >= a.last) break;
    const AlignJob job = a.jobs[it];
    const int n = job.n, m = job.m;
    const int nstrips = (m + ASH - 1) / ASH;
    const int Tn = n + MST_WAVE - 1;
    const int8_t* X = a.seqs + job.x_off;
    const int8_t* Y = a.seqs + job.y_off;
    uint32_t* fl = a.flags + job.flag_off;
    for (int q = 0;; ++q) {
      const int s = j + q * W;
      if (s >= nstrips) break;
      const int r0 = s * ASH;
      if (SRC == SRC_SEQ) {
        // this wave's previous sweep is over and no other wave reads its slot
        __builtin_amdgcn_wave_barrier();
        build_profile8(slot, S_lds, Y, m, r0, a.go + a.ge);
        __builtin_amdgcn_wave_barrier();
      }
      ALane L;
#pragma unroll
      for (int i = 0; i < AK; ++i) {
        L.Hm[i] = 0;
        L.E[i] = MST_NEG_INF;
      }
      L.Fout = MST_NEG_INF;
      L.Hout = 0;
      L.hdiag = 0;
      const int g = q * n;
      ALink in, out;
      int gin;
      const int* prog_in;
      int* cons_out;
      if (j > 0) {
        in = ALink{ring + (j - 1) * ARING, g, ARING - 1};
        gin = g;
        prog_in = s_prog + j - 1;
        cons_out = s_cons + j;
      } else {
        in = ALink{carry, 0, -1};
        gin = g - n;
        prog_in = s_prog + W;
        cons_out = s_cons + W;
      }
      int* prog_out;
      const int* cons_in;
      int span, pub;
      if (j < W - 1) {
        out = ALink{ring + j * ARING, g, ARING - 1};
        prog_out = s_prog + j;
        cons_in = s_cons + j + 1;
        span = ARING;
        pub = 1;
      } else {
        out = ALink{carry, 0, -1};
        prog_out = s_prog + W;
        cons_in = nullptr;
        span = 0;
        pub = 16;
      }
      uint32_t* fl_strip = fl + (int64_t)s * Tn * MST_WAVE;
      const uint2* tab_strip =
          (SRC == SRC_TAB) ? (const uint2*)(a.tab + job.tab_off) + (int64_t)s * Tn * MST_WAVE : nullptr;
      const bool first = (s == 0), last = (s == nstrips - 1);
      if (first) {
        if (last)
          arun_strip_pipe<SRC, true, false>(L, X, n, prof_lane, tab_strip, in, gin, prog_in,
                                            cons_out, out, g, prog_out, cons_in, span, pub, G,
                                            fl_strip);
        else
          arun_strip_pipe<SRC, true, true>(L, X, n, prof_lane, tab_strip, in, gin, prog_in,
                                           cons_out, out, g, prog_out, cons_in, span, pub, G,
                                           fl_strip);
      } else {
        if (last)
          arun_strip_pipe<SRC, false, false>(L, X, n, prof_lane, tab_strip, in, gin, prog_in,
                                             cons_out, out, g, prog_out, cons_in, span, pub, G,
                                             fl_strip);
        else
          arun_strip_pipe<SRC, false, true>(L, X, n, prof_lane, tab_strip, in, gin, prog_in,
                                            cons_out, out, g, prog_out, cons_in, span, pub, G,
                                            fl_strip);
      }
      if (last) {
        const int rl = (m - 1) - r0;
        const int owner = rl / AK, idx = rl % AK;
        int v = L.Hm[0];
#pragma unroll
        for (int i = 1; i < AK; ++i)
          if (idx == i) v = L.Hm[i];
        v = __shfl(v, owner);
        if (lane == 0) a.scores[it] = v - a.ge * (m + n);
      }
    }
    __syncthreads();  // rings, counters and the row buffer are reused by the next job
  }
}

template <int SRC>
__global__ void __launch_bounds__(MST_WAVE) align_kernel(AlignArgs a) {
  issue_first();
  __shared__ int8_t S_lds[MST_LETTERS * MST_LETTERS];
  // the profile only for scores from sequences; a score table needs none
  __shared__ __attribute__((aligned(16))) uint8_t prof[SRC == SRC_SEQ ? MST_LETTERS * MST_WAVE * AK : 16];
  const int lane = lane_id();
  if (SRC == SRC_SEQ) {
    for (int i = lane; i < MST_LETTERS * MST_LETTERS; i += MST_WAVE) S_lds[i] = a.subst[i];
  }
  __syncthreads();
  const int G = a.go - a.ge;
  int2* buf = a.bufs + (int64_t)blockIdx.x * 2 * a.bufcap;
  for (int64_t j = (int64_t)blockIdx.x + a.first; j < a.last; j += gridDim.x) {
    const AlignJob job = a.jobs[j];
    const int n = job.n, m = job.m;
    uint32_t* fl = a.flags + job.flag_off;
    const int nstrips = (m + ASH - 1) / ASH;
    const int Tn = n + MST_WAVE - 1;
    const int8_t* X = a.seqs + job.x_off;
    const int8_t* Y = a.seqs + job.y_off;
    ALane L;
    int result = 0;
    for (int s = 0; s < nstrips; ++s) {
      const int r0 = s * ASH;
      if (SRC == SRC_SEQ) {
        __builtin_amdgcn_wave_barrier();
        build_profile8(prof, S_lds, Y, m, r0, a.go + a.ge);
        __builtin_amdgcn_wave_barrier();
      }
#pragma unroll
      for (int i = 0; i < AK; ++i) {
        L.Hm[i] = 0;
        L.E[i] = MST_NEG_INF;
      }
      L.Fout = MST_NEG_INF;
      L.Hout = 0;
      L.hdiag = 0;
      const int2* bin = buf + (int64_t)(s & 1) * a.bufcap;
      int2* bout = buf + (int64_t)((s + 1) & 1) * a.bufcap;
      uint32_t* fl_strip = fl + (int64_t)s * Tn * MST_WAVE;
      const uint2* tab_strip =
          (SRC == SRC_TAB) ? (const uint2*)(a.tab + job.tab_off) + (int64_t)s * Tn * MST_WAVE : nullptr;
      const uint8_t* prof_lane = prof + lane * AK;
      const bool last = (s == nstrips - 1);
      if (s == 0) {
        if (last)
          arun_strip<SRC, true, false>(L, X, n, prof_lane, tab_strip, bin, bout, G, fl_strip);
        else
          arun_strip<SRC, true, true>(L, X, n, prof_lane, tab_strip, bin, bout, G, fl_strip);
      } else {
        if (last)
          arun_strip<SRC, false, false>(L, X, n, prof_lane, tab_strip, bin, bout, G, fl_strip);
        else
          arun_strip<SRC, false, true>(L, X, n, prof_lane, tab_strip, bin, bout, G, fl_strip);
      }
      if (!last) agent_fence();
      if (last) {
        const int rl = (m - 1) - r0;
        const int owner = rl / AK, idx = rl % AK;
        int v = L.Hm[0];
#pragma unroll
        for (int i = 1; i < AK; ++i)
          if (idx == i) v = L.Hm[i];
        v = __shfl(v, owner);
        result = v - a.ge * (m + n);
      }
    }
    if (lane == 0) a.scores[j] = result;
  }
}

// ---- traceback: one wave per job ----
//
// The walk (NeedlemanWunsch.cu:409-468) is one dependent chain of cell reads,
// so its cost is the latency of each step. A wave holds a tile of the flags
// in registers: 64 columns by up to 64 rows (8 groups of a lane's 8 rows of
// one strip), lane l holding column x0 + l, one VGPR per group, whose top-
// right cell is the walk's position when the tile is loaded. Each lane loads
// its column's 8 words (independent loads, one memory latency per tile) and,
// for the packed kernel's bit planes, rewrites them as 8 row nibbles in
// parallel. A step then reads its word from the current group's VGPR with one
// v_readlane (no LDS round trip) and runs the state machine on scalars; the
// path's ops collect one per lane of a VGPR (a select on the lane) and are
// written 64 at a time. The walk leaves a tile typically after 64 steps.

constexpr int TB_GROUPS = 8;  // 8-row groups of a tile

// (y: the row, or any value of its sign)
__device__ __forceinline__ int boundary_nib(int x, int y) {
  if (x < 0 && y < 0) return 0;
  if (y < 0) return 9;  // LEFT + contLeft
  return 6;             // UP + contUp
}

// bit r of b (8 bits) at bit 4r
__device__ __forceinline__ uint32_t spread4(uint32_t b) {
  b = (b | (b << 12)) & 0x000F000Fu;
  b = (b | (b << 6)) & 0x03030303u;
  b = (b | (b << 3)) & 0x11111111u;
  return b;
}

// A flag word of the packed kernel (bit planes: bytes s1 = diagonal not a
// maximum, s2 = E > F, s4 = F opened, s3 = E opened; bit r = row r) as the 8
// row nibbles (LEFT, UP, contUp, contLeft from bit 3 down): LEFT = !s1 | s2,
// UP = !s1 | !s2, contUp = !s4, contLeft = !s3.
__device__ __forceinline__ uint32_t planes_to_nibbles(uint32_t w) {
  const uint32_t s1 = w & 0xFFu, s2 = (w >> 8) & 0xFFu, s4 = (w >> 16) & 0xFFu, s3 = w >> 24;
  const uint32_t left = (~s1 | s2) & 0xFFu, up = ~(s1 & s2) & 0xFFu;
  return (spread4(left) << 3) | (spread4(up) << 2) | (spread4(~s4 & 0xFFu) << 1) | spread4(~s3 & 0xFFu);
}

__global__ void __launch_bounds__(MST_WAVE) traceback_kernel(TraceArgs a) {
  issue_first();
  const int lane = lane_id();
  for (int64_t j = (int64_t)blockIdx.x + a.first; j < a.last; j += gridDim.x) {
    const AlignJob job = a.jobs[j];
    const int n = job.n, m = job.m;
    const int Tn = n + MST_WAVE - 1;
    const int kr = job.k;  // rows per lane of the job's flag layout
    const bool planes = job.fmt != FMT_NIBBLE;
    const uint32_t* fl = a.flags + job.flag_off;
    uint8_t* out = a.ops + job.ops_off;
    // The walk's row y is kept as (g, r): group g = y / kr (a flag kernel
    // lane's rows; 64 groups a strip) and row r = y % kr in it, stepped by
    // hand so that no step divides.
    //
    // The state machine: a move is t2 = 1 (up), 2 (left) or 3 (diagonal), the
    // cell's own direction nibble >> 2, unless the previous move forces one:
    // an up move whose cell had contUp continues up, a left move whose cell
    // had contLeft continues left. prev = previous t2 << 2 | its cell's two
    // continuation bits; kForced holds the forced move (0: none) of each prev
    // in 2 bits. The walk ends at a cell with no direction (the corner).
    static_assert(OP_DIAG == 0 && OP_LEFT == 1 && OP_UP == 2, "op = 3 - t2");
    constexpr uint32_t kForced = (1u << (2 * 6)) | (1u << (2 * 7)) | (2u << (2 * 9)) | (2u << (2 * 11));
    int x = n - 1, g = (m - 1) / kr, r = (m - 1) % kr;
    int prev = 12;  // as after a diagonal move: nothing forced
    int k = 0;
    int obuf = 0;  // op of step k in lane k % 64
    bool done = false;
    // one step from cell nibble cur; false at a cell with no direction
    auto step = [&](int cur) -> bool {
      if (!(cur & 12)) return false;
      const int forced = (int)((kForced >> (2 * prev)) & 3u);
      const int t2 = forced ? forced : cur >> 2;
      obuf = lane == (k & (MST_WAVE - 1)) ? 3 - t2 : obuf;
      if ((k & (MST_WAVE - 1)) == MST_WAVE - 1) out[k - (MST_WAVE - 1) + lane] = (uint8_t)obuf;
      ++k;
      x -= t2 >> 1;  // left and diagonal
      r -= t2 & 1;   // up and diagonal
      const int wrap = r < 0;
      r += wrap ? kr : 0;
      g -= wrap;
      prev = (t2 << 2) | (cur & 3);
      return true;
    };
    // Inside the matrix, a tile at a time: columns [cx0, cx0 + 63] by groups
    // [cg0, g at the load]; the walk only moves left and up, so it leaves a
    // tile to the left or the top.
    while (!done && (x | g) >= 0) {
      const int gs = g & ~(MST_WAVE - 1);  // first group of the strip
      const int s = g / MST_WAVE;
      const int cg0 = max(gs, g - (TB_GROUPS - 1));
      const int cx0 = x - (MST_WAVE - 1);
      const int c = cx0 + lane;
      // a word outside the tile (a column left of 0, a group past g) reads
      // the job's first flag word instead and is never read by the walk
      uint32_t vt[TB_GROUPS];
#pragma unroll
      for (int gi = 0; gi < TB_GROUPS; ++gi) {
        const int G = cg0 + gi;
        const int lg = G - gs;  // lane of the flag kernel holding group G
        const bool in = G <= g && c >= 0;
        vt[gi] = fl[in ? ((int64_t)s * Tn + c + lg) * MST_WAVE + lg : 0];
      }
      if (planes) {
#pragma unroll
        for (int gi = 0; gi < TB_GROUPS; ++gi) vt[gi] = planes_to_nibbles(vt[gi]);
      }
      int gcur = -1;
      uint32_t vcur = 0;
      do {
        if (g != gcur) {
          // the group's VGPR by selects (an indexed register array would go
          // to scratch memory)
          const int gi = g - cg0;
          vcur = vt[0];
#pragma unroll
          for (int q = 1; q < TB_GROUPS; ++q) vcur = gi == q ? vt[q] : vcur;
          gcur = g;
        }
        const uint32_t w = (uint32_t)__builtin_amdgcn_readlane((int)vcur, x - cx0);
        if (!step((int)((w >> (4 * r)) & 15u))) {
          done = true;
          break;
        }
      } while ((x | g) >= 0 && x >= cx0 && g >= cg0);
    }
    // On row -1 or column -1 the directions are fixed (left along row -1, up
    // along column -1) up to the corner.
    while (!done && step(boundary_nib(x, g))) {
    }
    if (lane < (k & (MST_WAVE - 1))) out[(k & ~(MST_WAVE - 1)) + lane] = (uint8_t)obuf;
    if (lane == 0) a.lens[j] = k;
  }
}

void launch_align(int src, const AlignArgs& a, int grid, hipStream_t st) {
  if (src == SRC_SEQ)
    hipLaunchKernelGGL(align_kernel<SRC_SEQ>, dim3(grid), dim3(MST_WAVE), 0, st, a);
  else
    hipLaunchKernelGGL(align_kernel<SRC_TAB>, dim3(grid), dim3(MST_WAVE), 0, st, a);
}

void launch_align_pipe(int src, int W, const AlignArgs& a, int grid, hipStream_t st) {
  if (src == SRC_SEQ)
    hipLaunchKernelGGL((align_pipe_kernel<SRC_SEQ, 4>), dim3(grid), dim3(MST_WAVE * 4), 0, st, a);
  else if (W >= 8)
    hipLaunchKernelGGL((align_pipe_kernel<SRC_TAB, 8>), dim3(grid), dim3(MST_WAVE * 8), 0, st, a);
  else
    hipLaunchKernelGGL((align_pipe_kernel<SRC_TAB, 4>), dim3(grid), dim3(MST_WAVE * 4), 0, st, a);
}

void launch_align_tiles(int src, const TileArgs& a, int ntiles, hipStream_t st) {
  if (ntiles <= 0) return;
  if (src == SRC_SEQ)
    hipLaunchKernelGGL(align_tile_kernel<SRC_SEQ>, dim3(ntiles), dim3(MST_WAVE), 0, st, a);
  else
    hipLaunchKernelGGL(align_tile_kernel<SRC_TAB>, dim3(ntiles), dim3(MST_WAVE), 0, st, a);
}

void plan_tiles(const AlignJob* jobs, int64_t first, int64_t last, int tw, std::vector<TileJob>& tiles,
                std::vector<int64_t>& diag_off, int64_t* rows, int64_t* state) {
  std::vector<std::vector<TileJob>> by_d;
  int64_t rb = 0, st = 0;
  for (int64_t j = first; j < last; ++j) {
    const int S = (jobs[j].m + ASH - 1) / ASH, C = (jobs[j].n + tw - 1) / tw;
    if ((int)by_d.size() < S + C - 1) by_d.resize(S + C - 1);
    for (int s = 0; s < S; ++s)
      for (int c = 0; c < C; ++c) by_d[s + c].push_back(TileJob{(int)j, s, c, 0, rb, st});
    rb += tile_rows_elems(jobs[j].n);
    st += tile_state_elems(jobs[j].m);
  }
  tiles.clear();
  diag_off.assign(1, 0);
  for (auto& v : by_d) {
    tiles.insert(tiles.end(), v.begin(), v.end());
    diag_off.push_back((int64_t)tiles.size());
  }
  *rows = rb;
  *state = st;
}

void launch_tile_diagonals(int src, TileArgs a, const TileJob* d_tiles, const std::vector<int64_t>& diag_off,
                           hipStream_t st) {
  for (size_t d = 0; d + 1 < diag_off.size(); ++d) {
    a.tiles = d_tiles + diag_off[d];
    launch_align_tiles(src, a, (int)(diag_off[d + 1] - diag_off[d]), st);
  }
}

int64_t align_strips(int m) { return m <= 0 ? 1 : (m + ASH - 1) / ASH; }

void launch_traceback(const TraceArgs& a, hipStream_t st) {
  const int64_t n = a.last - a.first;
  if (n <= 0) return;
  const int64_t grid = n < 65536 ? n : 65536;
  hipLaunchKernelGGL(traceback_kernel, dim3((unsigned)grid), dim3(MST_WAVE), 0, st, a);
}

int64_t flag_words(int n, int m, int k) {
  const int64_t nstrips = (m + MST_WAVE * k - 1) / (MST_WAVE * k);
  return nstrips * (int64_t)(n + MST_WAVE - 1) * MST_WAVE;
}

}  // namespace mst

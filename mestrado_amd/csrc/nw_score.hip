// All-pairs global alignment scores (R1) — one pair per wavefront.
//
// Recurrence of NeedlemanWunsch.cu:22-161 in int32 with the reference's
// boundary (row/column -1 cost ge per residue, no opening) and S[Y][X].
// Every value is stored with the bias ge*(r+c) (r row, c column), the gap
// opening folded in and the constant boundary value subtracted, which turns
// the affine recurrence into a max-plus form whose boundary is all zeros:
//   Hm(r,c) = H(r,c) + ge*(r+c) - (go-ge) + (go+ge)
//   E'(r,c) = max(E'(r,c-1), Hm(r,c-1))
//   F'(r,c) = max(F'(r-1,c), Hm(r-1,c))
//   Hm(r,c) = max3(E', F', Hm(r-1,c-1) + S(r,c) + go + ge) - (go-ge)
//   score   = Hm(m-1,n-1) - ge*(m+n)
// Row -1 and column -1 of Hm are 0; E' at column -1 and F' at row -1 may be
// taken as 0 as well because they only ever meet max() against a boundary Hm
// of 0 (the score is unchanged; the traceback kernel keeps a true -inf there
// because its continuation flags compare against them). A cell costs five
// integer VALU ops: max (E'), max (F'), add (diagonal + profile byte, the
// byte extraction folded into an SDWA operand), max3, sub.
//
// Layout: the Y sequence (rows) is cut into strips of 64*K rows; lane l of the
// wave owns K consecutive rows of the strip and sweeps the X columns skewed by
// one step per lane (anti-diagonal wavefront). F' and Hm of a lane's last row
// move to the next lane with v_mov_dpp wave_shr:1 (lane 0 reads zeros, or the
// previous strip's last row, which lane 63 left in a per-wave double buffer).
// The substitution scores of the lane's K rows, offset by go+ge (signed bytes),
// are packed K bytes per residue in LDS (query profile) laid out [x][lane][K],
// so one ds_read_b32/b64/b128 per step and lane is bank-conflict free. X
// residues are prefetched one block of U steps ahead.
#include "common.h"
#include "nw_score.h"

namespace mst {

constexpr int U = 8;  // steps per unrolled block / prefetch distance

template <int K>
struct ProfRead;
template <>
struct ProfRead<4> {
  static __device__ __forceinline__ void get(const uint8_t* p, uint32_t* w) {
    w[0] = *(const uint32_t*)p;
  }
};
template <>
struct ProfRead<8> {
  static __device__ __forceinline__ void get(const uint8_t* p, uint32_t* w) {
    uint2 v = *(const uint2*)p;
    w[0] = v.x;
    w[1] = v.y;
  }
};
template <>
struct ProfRead<12> {
  static __device__ __forceinline__ void get(const uint8_t* p, uint32_t* w) {
    uint2 v = *(const uint2*)p;
    w[0] = v.x;
    w[1] = v.y;
    w[2] = *(const uint32_t*)(p + 8);
  }
};
template <>
struct ProfRead<16> {
  static __device__ __forceinline__ void get(const uint8_t* p, uint32_t* w) {
    uint4 v = *(const uint4*)p;
    w[0] = v.x;
    w[1] = v.y;
    w[2] = v.z;
    w[3] = v.w;
  }
};

// Signed byte i of the packed profile (S + go + ge fits int8 for go+ge <= 116).
__device__ __forceinline__ int pbyte(const uint32_t* w, int i) {
  return (int)(int8_t)((w[i >> 2] >> (8 * (i & 3))) & 0xFFu);
}

// wave_shr:1 with zero fill for lane 0.
__device__ __forceinline__ int shr1_zero(int v) {
  return __builtin_amdgcn_update_dpp(0, v, DPP_WAVE_SHR1, 0xF, 0xF, true);
}

// Packed query profile of this lane's K rows (r0 + lane*K + i).
template <int K>
__device__ __forceinline__ void build_profile(uint8_t* prof, const int8_t* S_lds,
                                              const int8_t* Y, int m, int r0, int bonus) {
  const int lane = lane_id();
  int yrow[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    int r = r0 + lane * K + i;
    yrow[i] = (r < m) ? (int)Y[r] * MST_LETTERS : 0;
  }
  for (int x = 0; x < MST_LETTERS; ++x) {
    uint32_t w[K / 4];
#pragma unroll
    for (int q = 0; q < K / 4; ++q) {
      uint32_t v = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        int s = (int)S_lds[yrow[q * 4 + b] + x] + bonus;
        v |= ((uint32_t)s & 0xFFu) << (8 * b);
      }
      w[q] = v;
    }
    uint32_t* dst = (uint32_t*)(prof + (x * MST_WAVE + lane) * K);
#pragma unroll
    for (int q = 0; q < K / 4; ++q) dst[q] = w[q];
  }
}

template <int K>
__device__ __forceinline__ void cell_column(const uint32_t* w, int (&Hm)[K], int (&E)[K],
                                            int& f, int& hu, int hd, int G) {
#pragma unroll
  for (int i = 0; i < K; ++i) {
    int e = imax(E[i], Hm[i]);
    f = imax(f, hu);
    int h = imax3(e, f, hd + pbyte(w, i)) - G;
    hd = Hm[i];
    Hm[i] = h;
    E[i] = e;
    hu = h;
  }
}

template <int K>
struct Lane {
  int Hm[K], E[K];
  int Fout, Hout, hdiag;
};

// One anti-diagonal step: lane l handles column c = t - l of its K rows.
template <int K, bool PRED, bool FIRST, bool STORE>
__device__ __forceinline__ void step(Lane<K>& L, int t, int x, int2 bv, const uint8_t* prof_lane,
                                     int n, int G, int2* bout) {
  const int lane = lane_id();
  int fin, hup;
  if (FIRST) {
    fin = shr1_zero(L.Fout);
    hup = shr1_zero(L.Hout);
  } else {
    fin = shr1(bv.x, L.Fout);
    hup = shr1(bv.y, L.Hout);
  }
  const int c = t - lane;
  const bool act = !PRED || (c >= 0 && c < n);
  if (act) {
    uint32_t w[K / 4];
    ProfRead<K>::get(prof_lane + x * (MST_WAVE * K), w);
    int f = fin, hu = hup;
    cell_column<K>(w, L.Hm, L.E, f, hu, L.hdiag, G);
    L.Fout = f;
    L.Hout = hu;
    if (STORE && lane == MST_WAVE - 1) bout[c] = make_int2(f, hu);
  }
  L.hdiag = hup;
}

template <bool FIRST>
__device__ __forceinline__ void load_block(int t0, const int8_t* __restrict__ X, int n,
                                           const int2* bin, int* xs, int2* bs) {
  const int lane = lane_id();
#pragma unroll
  for (int j = 0; j < U; ++j) {
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
__device__ void run_strip(Lane<K>& L, const int8_t* __restrict__ X, int n, const uint8_t* prof_lane,
                          const int2* bin, int2* bout, int G) {
  const int T = n + MST_WAVE - 1;
  int xa[U], xb[U];
  int2 ba[U], bb[U];
  load_block<FIRST>(0, X, n, bin, xa, ba);
  for (int tb = 0; tb < T; tb += U) {
    load_block<FIRST>(tb + U, X, n, bin, xb, bb);
    const bool full = (tb >= MST_WAVE - 1) && (tb + U <= n);
    if (full) {
#pragma unroll
      for (int j = 0; j < U; ++j)
        step<K, false, FIRST, STORE>(L, tb + j, xa[j], ba[j], prof_lane, n, G, bout);
    } else {
#pragma unroll
      for (int j = 0; j < U; ++j)
        if (tb + j < T) step<K, true, FIRST, STORE>(L, tb + j, xa[j], ba[j], prof_lane, n, G, bout);
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
      xa[j] = xb[j];
      ba[j] = bb[j];
    }
  }
}

// Scores one pair with the whole wave. X: columns (n), Y: rows (m), both
// reversed codes. buf: this wave's 2*bufcap int2 boundary entries.
template <int K>
__device__ int score_pair(const int8_t* __restrict__ X, int n, const int8_t* __restrict__ Y,
                          int m, const int8_t* S_lds, uint8_t* prof, int2* buf, int64_t bufcap,
                          int go, int ge) {
  const int lane = lane_id();
  const int G = go - ge;
  const int SH = MST_WAVE * K;
  const int nstrips = (m + SH - 1) / SH;
  const uint8_t* prof_lane = prof + lane * K;
  Lane<K> L;
  int result = 0;
  for (int s = 0; s < nstrips; ++s) {
    const int r0 = s * SH;
    __builtin_amdgcn_wave_barrier();
    build_profile<K>(prof, S_lds, Y, m, r0, go + ge);
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int i = 0; i < K; ++i) {
      L.Hm[i] = 0;
      L.E[i] = 0;
    }
    L.Fout = L.Hout = L.hdiag = 0;
    const int2* bin = buf + (int64_t)(s & 1) * bufcap;
    int2* bout = buf + (int64_t)((s + 1) & 1) * bufcap;
    const bool last = (s == nstrips - 1);
    if (s == 0) {
      if (last)
        run_strip<K, true, false>(L, X, n, prof_lane, bin, bout, G);
      else
        run_strip<K, true, true>(L, X, n, prof_lane, bin, bout, G);
    } else {
      if (last)
        run_strip<K, false, false>(L, X, n, prof_lane, bin, bout, G);
      else
        run_strip<K, false, true>(L, X, n, prof_lane, bin, bout, G);
    }
    if (!last) agent_fence();
    if (last) {
      const int rl = (m - 1) - r0;
      const int owner = rl / K, idx = rl % K;
      int v = L.Hm[0];
#pragma unroll
      for (int i = 1; i < K; ++i)
        if (idx == i) v = L.Hm[i];
      v = __shfl(v, owner);
      result = v - ge * (m + n);
    }
  }
  return result;
}

template <int K>
__global__ void __launch_bounds__(MST_WAVE) score_pairs_kernel(ScoreArgs a) {
  __shared__ int8_t S_lds[MST_LETTERS * MST_LETTERS];
  __shared__ __attribute__((aligned(16))) uint8_t prof[MST_LETTERS * MST_WAVE * K];
  const int lane = lane_id();
  for (int i = lane; i < MST_LETTERS * MST_LETTERS; i += MST_WAVE) S_lds[i] = a.subst[i];
  __syncthreads();
  int2* buf = a.bufs + (int64_t)blockIdx.x * 2 * a.bufcap;
  for (int64_t p = (int64_t)blockIdx.x + a.first; p < a.last; p += gridDim.x) {
    const int2 pr = a.pairs ? a.pairs[p] : make_int2(a.pi[p], a.pj[p]);
    int xs = pr.x, ys = pr.y;
    int64_t xo = a.off[xs], yo = a.off[ys];
    int n = (int)(a.off[xs + 1] - xo), m = (int)(a.off[ys + 1] - yo);
    int sc;
    if (n == 0 || m == 0) {
      sc = -a.ge * (n + m);
    } else {
      const int8_t* X = a.seqs + xo;
      const int8_t* Y = a.seqs + yo;
      if (a.transpose_ok && m > n) {  // shorter sequence in the lanes
        const int8_t* t = X;
        X = Y;
        Y = t;
        int tn = n;
        n = m;
        m = tn;
      }
      sc = score_pair<K>(X, n, Y, m, S_lds, prof, buf, a.bufcap, a.go, a.ge);
    }
    if (lane == 0) {
      if (a.out_pair) a.out_pair[p] = sc;
      if (a.out_mat) {
        a.out_mat[(int64_t)xs * a.nseq + ys] = sc;
        a.out_mat[(int64_t)ys * a.nseq + xs] = sc;
      }
    }
  }
}

void launch_score_pairs(int K, const ScoreArgs& a, int grid, hipStream_t st) {
  switch (K) {
    case 4:
      hipLaunchKernelGGL(score_pairs_kernel<4>, dim3(grid), dim3(MST_WAVE), 0, st, a);
      break;
    case 8:
      hipLaunchKernelGGL(score_pairs_kernel<8>, dim3(grid), dim3(MST_WAVE), 0, st, a);
      break;
    case 12:
      hipLaunchKernelGGL(score_pairs_kernel<12>, dim3(grid), dim3(MST_WAVE), 0, st, a);
      break;
    default:
      hipLaunchKernelGGL(score_pairs_kernel<16>, dim3(grid), dim3(MST_WAVE), 0, st, a);
      break;
  }
}

}  // namespace mst

// Neighbour joining on the device (R5): SpeciesBlock::buildTree
// (SpeciesBlock.cpp:844-1343) with the arithmetic of buildGblRowSums (:426-453)
// and modifyMatrix (:710-842), one process, no MPI.
//
// Selection is exact: every step scans the whole upper triangle of the active
// pairs for the minimum q, ties broken towards the larger max(i,j), then the
// larger min(i,j) (:1169-1178). That is the pair the reference's bin/heap
// search returns, since its pruning bounds (:1083-1127) only skip pairs whose
// q is provably larger. q is evaluated as ((double)d*(n-2) - T_a) - T_b with
// (a,b) in the order the reference's candidate list holds the pair: node-id
// order for pairs created by the last bin rebuild, storage-index order for
// pairs created by the insertion of a newer node (:1280-1287).
//
// Arithmetic that reaches the output follows the reference bit for bit:
//   T_i   = sum over active j in id order of D(i,j) in fp64, diagonal included
//   tmp_k = ((D_ik + D_jk) - d) / 2 in fp32                          (:795,811)
//   T_k  += (double)(float)(tmp_k - (D_ik + D_jk))                    (:801,814)
//   T_new = sum over active k in id order of tmp_k in fp64            (:799,818)
// and T is rebuilt from D at the reference's schedule (:864,1225-1247). The new
// node takes the storage row of the smaller id and keeps its stale diagonal.
//
// Layout: D is an n0 x n0 fp32 matrix in HBM indexed by storage row (n0 <=
// 16384 here keeps it in the MALL); the active nodes are a compact array in id
// order. One workgroup of 1024 threads runs every step of the loop inside one
// launch, so the n-1 steps cost no launches or host round trips.
#include <float.h>
#include <limits.h>

#include "common.h"
#include "nj.h"

namespace mst {

constexpr int NJ_THREADS = 1024;
// the per-step join kernel: fewer waves, shorter barriers (8 positions each)
constexpr int NJ_JOIN_THREADS = 512;

struct Best {
  double q;
  double Ta, Tb;  // T of the pair's storage rows
  float d;        // D of the pair
  int J, I;       // node ids, I < J
  int pa, pb;     // positions in the active array, pa < pb
  int ra, rb;     // storage rows
};

__device__ __forceinline__ bool better(const Best& a, const Best& b) {
  if (a.q < b.q) return true;
  if (a.q > b.q) return false;
  if (a.J != b.J) return a.J > b.J;
  return a.I > b.I;
}

// b = better(c, b) ? c : b, field by field (a whole-struct select or a Best
// assigned in a loop is kept in scratch memory by the compiler)
__device__ __forceinline__ void keep_better(Best& b, const Best& c) {
  const bool t = better(c, b);
  b.q = t ? c.q : b.q;
  b.Ta = t ? c.Ta : b.Ta;
  b.Tb = t ? c.Tb : b.Tb;
  b.d = t ? c.d : b.d;
  b.J = t ? c.J : b.J;
  b.I = t ? c.I : b.I;
  b.pa = t ? c.pa : b.pa;
  b.pb = t ? c.pb : b.pb;
  b.ra = t ? c.ra : b.ra;
  b.rb = t ? c.rb : b.rb;
}

__device__ __forceinline__ Best no_best() {
  Best b;
  b.q = DBL_MAX;
  b.Ta = b.Tb = 0.0;
  b.d = 0.0f;
  b.J = b.I = -1;
  b.pa = b.pb = b.ra = b.rb = -1;
  return b;
}

// One step of the wave's reduction of candidates: each lane takes the better
// of its own and the one DPP control CTRL brings (rows outside RM, and lanes
// with no source, keep their own), tracking the lane the winner came from.
template <int CTRL, int RM>
__device__ __forceinline__ int dpp_keep(int v) {
  return __builtin_amdgcn_update_dpp(v, v, CTRL, RM, 0xF, false);
}
template <int CTRL, int RM>
__device__ __forceinline__ void best_step(double& q, int& J, int& I, int& src) {
  const long long qb = __double_as_longlong(q);
  const int qlo = dpp_keep<CTRL, RM>((int)qb), qhi = dpp_keep<CTRL, RM>((int)(qb >> 32));
  const int oJ = dpp_keep<CTRL, RM>(J), oI = dpp_keep<CTRL, RM>(I), os = dpp_keep<CTRL, RM>(src);
  const double oq = __longlong_as_double(((long long)qhi << 32) | (long long)(unsigned int)qlo);
  const bool t = oq < q || (oq == q && (oJ > J || (oJ == J && oI > I)));
  q = t ? oq : q;
  J = t ? oJ : J;
  I = t ? oI : I;
  src = t ? os : src;
}

__device__ __forceinline__ double readlane_f64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

// The wave's best candidate, the same in every lane: the keys (q, J, I)
// reduced toward lane 63 by DPP (row shifts 1, 2, 4, 8, then the row
// broadcasts of lanes 15 and 31), the winner's fields read from its lane. The
// total order makes the winner the same whatever the order of the reduction.
__device__ __forceinline__ Best wave_best(Best best) {
  double q = best.q;
  int J = best.J, I = best.I, src = lane_id();
  best_step<0x111, 0xF>(q, J, I, src);
  best_step<0x112, 0xF>(q, J, I, src);
  best_step<0x114, 0xF>(q, J, I, src);
  best_step<0x118, 0xF>(q, J, I, src);
  best_step<0x142, 0xA>(q, J, I, src);
  best_step<0x143, 0xC>(q, J, I, src);
  const int w = __builtin_amdgcn_readlane(src, 63);
  Best b;
  b.q = readlane_f64(best.q, w);
  b.Ta = readlane_f64(best.Ta, w);
  b.Tb = readlane_f64(best.Tb, w);
  b.d = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(best.d), w));
  b.J = __builtin_amdgcn_readlane(best.J, w);
  b.I = __builtin_amdgcn_readlane(best.I, w);
  b.pa = __builtin_amdgcn_readlane(best.pa, w);
  b.pb = __builtin_amdgcn_readlane(best.pb, w);
  b.ra = __builtin_amdgcn_readlane(best.ra, w);
  b.rb = __builtin_amdgcn_readlane(best.rb, w);
  return b;
}

// The scan's work in balanced units. The upper triangle's rows pa = 0 ..
// m-2 (row pa: columns pa+1 .. m-1) are folded in pairs, row f with row
// m-2-f, into ceil((m-1)/2) rows of m columns (the middle row of an odd count
// alone), and each folded row is cut into segments of kScanSeg columns: unit
// u is segment u % ceil(m / kScanSeg) of folded row u / ceil(m / kScanSeg).
// Units are dealt round robin to the waves, so no wave sweeps much more than
// another (a wave per whole row left the first rows' waves with three times
// the average).
constexpr int kScanSeg = 512;

__device__ __forceinline__ int scan_units(int m) {
  return ((m - 1 + 1) / 2) * ((m + kScanSeg - 1) / kScanSeg);
}

__device__ __forceinline__ void scan_unit(Best& best, int u, int m, int R, double nm2, const int* act,
                                          const int* ids, const float* D, const double* Tp, int ld,
                                          int lane) {
  const int spr = (m + kScanSeg - 1) / kScanSeg;
  const int f = u / spr, seg = u - f * spr;
  const int rA = f, rB = m - 2 - f;
  const int nA = m - 1 - f, n = rB > rA ? m : nA;  // columns of the folded row
  const int c0 = seg * kScanSeg, c1 = c0 + kScanSeg < n ? c0 + kScanSeg : n;
  if (c0 >= c1) return;
  const int rBc = rB > rA ? rB : rA;
  const int raA = act[rA], raB = act[rBc], iaA = ids[rA], iaB = ids[rBc];
  const double TaA = Tp[rA], TaB = Tp[rBc];
  const float* DA = D + (int64_t)raA * ld;
  const float* DB = D + (int64_t)raB * ld;
  // the unit's best in scalars (a Best assigned per candidate stays in memory)
  double bq = DBL_MAX, bTa = 0.0, bTb = 0.0;
  float bd = 0.0f;
  int bJ = -1, bI = -1, bpa = -1, bpb = -1, bra = -1, brb = -1;
  auto fold = [&](bool inA, int pb, int rb, int ib, float dv, double Tb) {
    const int ra = inA ? raA : raB, ia = inA ? iaA : iaB;
    const double Ta = inA ? TaA : TaB;
    const double dq = (double)dv * nm2;
    // pair created by an insertion after the last rebuild: storage order
    const bool swap = (ib >= R) && (rb < ra);
    const double q = swap ? (dq - Tb) - Ta : (dq - Ta) - Tb;
    const bool take = q < bq || (q == bq && (ib > bJ || (ib == bJ && ia > bI)));
    bq = take ? q : bq;
    bJ = take ? ib : bJ;
    bI = take ? ia : bI;
    bpa = take ? (inA ? rA : rB) : bpa;
    bpb = take ? pb : bpb;
    bra = take ? ra : bra;
    brb = take ? rb : brb;
    bd = take ? dv : bd;
    bTa = take ? Ta : bTa;
    bTb = take ? Tb : bTb;
  };
  // four columns per lane in flight: T, ids and storage rows by position,
  // then D through the storage rows
  for (int c = c0 + lane; c < c1; c += 4 * 64) {
    int pb[4], rb[4], ib[4];
    bool inA[4], ok[4];
    double tb[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int cc = c + 64 * k;
      ok[k] = cc < c1;
      inA[k] = cc < nA;
      pb[k] = inA[k] ? f + 1 + cc : rB + 1 + (cc - nA);
      const int pc = ok[k] ? pb[k] : m - 1;
      rb[k] = act[pc];
      ib[k] = ids[pc];
      tb[k] = Tp[pc];
    }
    float dv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) dv[k] = (inA[k] ? DA : DB)[rb[k]];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (ok[k]) fold(inA[k], pb[k], rb[k], ib[k], dv[k], tb[k]);
  }
  if (bJ >= 0) {
    Best c;
    c.q = bq;
    c.I = bI;
    c.J = bJ;
    c.pa = bpa;
    c.pb = bpb;
    c.ra = bra;
    c.rb = brb;
    c.d = bd;
    c.Ta = bTa;
    c.Tb = bTb;
    keep_better(best, c);
  }
}

// New row of a join, kept in LDS as it is written (up to kNjLdsRow active
// nodes), so the fp64 sum that gives T of the new node, which must run in id
// order on one thread, reads LDS instead of a chain of dependent loads.
constexpr int kNjLdsRow = 8192;

// s + x[b] + x[b+1] + ... + x[e-1], left to right in fp64, x[p] = get(p):
// blocks of 16 values loaded one block ahead of the adds that use them, so the
// only serial chain is the adds' own (the rounding order stays position order)
template <typename Get>
__device__ __forceinline__ double run_sum(Get get, int b, int e, double s) {
  constexpr int B = 16;
  if (e - b < B) {
    for (int p = b; p < e; ++p) s += (double)get(p);
    return s;
  }
  float cur[B], nxt[B];
#pragma unroll
  for (int k = 0; k < B; ++k) cur[k] = get(b + k);
  int p = b;
  for (; p + 2 * B <= e; p += B) {
#pragma unroll
    for (int k = 0; k < B; ++k) nxt[k] = get(p + B + k);
#pragma unroll
    for (int k = 0; k < B; ++k) s += (double)cur[k];
#pragma unroll
    for (int k = 0; k < B; ++k) cur[k] = nxt[k];
  }
#pragma unroll
  for (int k = 0; k < B; ++k) s += (double)cur[k];
  for (p += B; p < e; ++p) s += (double)get(p);
  return s;
}

// T of the new node: the new row summed in position order, positions pa and
// pb (the joined pair) left out, as three runs without a test per element
__device__ __forceinline__ double new_row_sum(const float* rowv, const float* Drow, const int* act, int m,
                                              int pa, int pb) {
  const int lo = pa < pb ? pa : pb, hi = pa < pb ? pb : pa;
  double s = 0.0;
  if (m <= kNjLdsRow) {
    auto get = [&](int p) { return rowv[p]; };
    s = run_sum(get, 0, lo, s);
    s = run_sum(get, lo + 1, hi, s);
    s = run_sum(get, hi + 1, m, s);
  } else {
    auto get = [&](int p) { return Drow[act[p]]; };
    s = run_sum(get, 0, lo, s);
    s = run_sum(get, lo + 1, hi, s);
    s = run_sum(get, hi + 1, m, s);
  }
  return s;
}

// Exponent of the lowest set bit of a nonzero float: x is an odd integer
// times 2^k.
__device__ __forceinline__ int lsb_exp(float x) {
  const unsigned int b = __float_as_uint(x);
  const int e = (int)((b >> 23) & 0xFFu);
  const unsigned int m = b & 0x7FFFFFu;
  return e == 0 ? -149 + __builtin_ctz(m) : e - 150 + __builtin_ctz(m | 0x800000u);
}

struct SumRed {
  double sum[NJ_THREADS / 64], mag[NJ_THREADS / 64];
  int k[NJ_THREADS / 64];
  double result;
};

// T of the new node by the whole workgroup: the fp64 sum of the new row in
// position order, positions pa and pb left out. Each thread adds its own
// positions' terms as it writes them (sum, sum of magnitudes, least exponent
// of their lowest set bits); when every term is an integer multiple of
// 2^kmin and the sum of magnitudes stays below 2^(52+kmin), every partial sum
// of every order of additions is an exact multiple of 2^kmin below 2^53 of
// them, so the threads' and the tree's sums equal the sum taken in position
// order, bit for bit (an exact zero comes out +0 either way, every run
// starting from +0). Otherwise one thread adds in position order. Returns the
// value to every thread (ends after a barrier).
struct RowSum {
  double sum = 0.0, mag = 0.0;
  int kmin = INT_MAX;
  __device__ __forceinline__ void add(float x) {
    if (x != 0.0f) {
      const int k = lsb_exp(x);
      kmin = k < kmin ? k : kmin;
    }
    sum += (double)x;
    mag += fabs((double)x);
  }
};

__device__ __forceinline__ double new_row_sum_finish(SumRed& r, RowSum t, const float* rowv,
                                                     const float* Drow, const int* act, int m, int pa,
                                                     int pb) {
  const int tid = threadIdx.x;
  for (int off = 32; off > 0; off >>= 1) {
    t.sum += __shfl_down(t.sum, off);
    t.mag += __shfl_down(t.mag, off);
    const int ko = __shfl_down(t.kmin, off);
    t.kmin = ko < t.kmin ? ko : t.kmin;
  }
  if ((tid & 63) == 0) {
    r.sum[tid >> 6] = t.sum;
    r.mag[tid >> 6] = t.mag;
    r.k[tid >> 6] = t.kmin;
  }
  __syncthreads();  // (also orders the new row's LDS and global writes before the fallback)
  if (tid < 64) {
    const int nw = (int)(blockDim.x / 64);
    double s = tid < nw ? r.sum[tid] : 0.0, g = tid < nw ? r.mag[tid] : 0.0;
    int k = tid < nw ? r.k[tid] : INT_MAX;
    for (int off = 32; off > 0; off >>= 1) {
      s += __shfl_down(s, off);
      g += __shfl_down(g, off);
      const int ko = __shfl_down(k, off);
      k = ko < k ? ko : k;
    }
    if (tid == 0) {
      // (all terms zero: k stays INT_MAX and the sum is +0)
      const bool exact = k == INT_MAX || g < ldexp(1.0, 52 + k);
      r.result = exact ? s : new_row_sum(rowv, Drow, act, m, pa, pb);
    }
  }
  __syncthreads();
  return r.result;
}

// Shared memory of a workgroup's candidate reduction, and of one that runs joins.
struct BestLds {
  Best red[NJ_THREADS / 64];
  Best win;
};
struct JoinLds {
  float rowv[kNjLdsRow];
  BestLds b;
  SumRed sred;
};

// The workgroup's best of its waves' bests (every thread gets it; ends after
// a barrier).
__device__ __forceinline__ Best workgroup_best(BestLds& L, Best best) {
  const int tid = threadIdx.x;
  best = wave_best(best);
  if ((tid & 63) == 0) L.red[tid >> 6] = best;
  __syncthreads();
  if (tid < 64) {
    // the waves' bests reduced by the first wave
    Best b = tid < (int)(blockDim.x / 64) ? L.red[tid] : no_best();
    b = wave_best(b);
    if (tid == 0) L.win = b;
  }
  __syncthreads();
  return L.win;
}

// The per-position values of a join (storage row, node id, T by position) of
// one batch of positions, loaded before the winner is known.
// (U positions per thread and batch, TH threads: the whole-matrix loop's
// workgroup takes 4 of NJ_THREADS, the per-step join kernel 8 of
// NJ_JOIN_THREADS.)
template <int U>
struct JoinBatch {
  int rk[U], ik[U];
  double tk[U];
};

template <int TH, int U>
__device__ __forceinline__ void join_load(JoinBatch<U>& B, int p0, int m, const int* act, const int* ids,
                                          const double* Tp) {
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int p = p0 + u * TH;
    const int pc = p < m ? p : m - 1;
    B.rk[u] = act[pc];
    B.ik[u] = ids[pc];
    B.tk[u] = Tp[pc];
  }
}

// One join (modifyMatrix, SpeciesBlock.cpp:710-842) of the winner w by the
// whole workgroup (NJ_THREADS): the join record, the new row into storage
// row w.ra (and its column), T of the other rows, T of the new node, and the
// active arrays compacted into act2/ids2/Tp2 (pa and pb dropped, the new node
// appended) in the same pass over the positions. B0: this thread's first
// batch of positions (join_load at p0 = threadIdx.x), loaded by the caller.
template <int TH, int U>
__device__ __forceinline__ void join_step(JoinLds& L, const Best& w, int m, int next, float* D, double* T,
                                          int ld, const int* act, const int* ids, const double* Tp,
                                          int* act2, int* ids2, double* Tp2, NjJoin* rec,
                                          const JoinBatch<U>& B0) {
  const int tid = threadIdx.x;
  const int ra = w.ra, rb = w.rb;
  const float d = w.d;
  if (tid == 0) {
    rec->node = next;
    rec->left = w.I;
    rec->right = w.J;
    rec->d = d;
    rec->Ti = w.Ta;
    rec->Tj = w.Tb;
    rec->n = m;
    rec->q = w.q;
  }
  const float* Da = D + (int64_t)ra * ld;
  const float* Db = D + (int64_t)rb * ld;
  JoinBatch<U> B = B0;
  RowSum rs;
  for (int p0 = tid; p0 < m; p0 += U * TH) {
    if (p0 != tid) join_load<TH, U>(B, p0, m, act, ids, Tp);
    float dik[U], djk[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      dik[u] = Da[B.rk[u]];
      djk[u] = Db[B.rk[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int p = p0 + u * TH;
      if (p >= m || p == w.pa || p == w.pb) continue;
      const int np = p - (p > w.pa) - (p > w.pb);
      const float tmp = ((dik[u] + djk[u]) - d) / 2;
      const float dlt = tmp - (dik[u] + djk[u]);
      const double t = B.tk[u] + (double)dlt;
      T[B.rk[u]] = t;
      act2[np] = B.rk[u];
      ids2[np] = B.ik[u];
      Tp2[np] = t;
      D[(int64_t)ra * ld + B.rk[u]] = tmp;
      D[(int64_t)B.rk[u] * ld + ra] = tmp;
      if (m <= kNjLdsRow) L.rowv[p] = tmp;
      rs.add(tmp);
    }
  }
  if (tid == 0) {
    act2[m - 2] = ra;
    ids2[m - 2] = next;
  }
  const double tn = new_row_sum_finish(L.sred, rs, L.rowv, D + (int64_t)ra * ld, act, m, w.pa, w.pb);
  if (tid == 0) {
    T[ra] = tn;
    Tp2[m - 2] = tn;
  }
}

// The whole join loop of one matrix, run by one workgroup of NJ_THREADS.
__device__ __forceinline__ void nj_run(const NjArgs a) {
  __shared__ JoinLds L;
  __shared__ int s_m, s_R, s_steps, s_next;
  const int tid = threadIdx.x;
  const int n0 = a.n0;
  float* D = a.D;
  double* T = a.T;
  int* act = a.act;   // storage rows of the active nodes, id order
  int* ids = a.ids;   // node ids of the active nodes
  int* act2 = a.act + n0;  // ping-pong halves
  int* ids2 = a.ids + n0;
  double* Tp = a.Tpos;     // T by active position, ping-pong halves
  double* Tp2 = a.Tpos + n0;
  const int ld = a.ld;

  // initial row sums (buildGblRowSums): one thread per row, id order; D is
  // symmetric so thread r reads column r, coalesced across threads
  for (int r = tid; r < n0; r += NJ_THREADS) {
    double s = 0.0;
    for (int j = 0; j < n0; ++j) s += (double)D[(int64_t)j * ld + r];
    T[r] = s;
    Tp[r] = s;
  }
  if (tid == 0) {
    s_m = n0;
    s_R = n0;
    s_next = n0;
    int st = (int)((float)n0 * a.ratio);
    s_steps = st > 200 ? st : 200;
  }
  __syncthreads();

  for (int step = 0; step < n0 - 1; ++step) {
    const int m = s_m, R = s_R;
    const double nm2 = (double)(m - 2);
    // ---- scan of the upper triangle (positions a < b) ----
    Best best = no_best();
    // one wave per row pa (rows dealt round robin), lanes stride the columns
    const int wave = tid >> 6, lane = tid & 63;
    const int units = scan_units(m);
    for (int u = wave; u < units; u += NJ_THREADS / 64)
      scan_unit(best, u, m, R, nm2, act, ids, D, Tp, ld, lane);
    JoinBatch<4> B0;
    join_load<NJ_THREADS, 4>(B0, tid, m, act, ids, Tp);
    const Best w = workgroup_best(L.b, best);
    join_step<NJ_THREADS, 4>(L, w, m, s_next, D, T, ld, act, ids, Tp, act2, ids2, Tp2, &a.joins[step], B0);
    __syncthreads();
    {
      int* t = act;
      act = act2;
      act2 = t;
      t = ids;
      ids = ids2;
      ids2 = t;
      double* u = Tp;
      Tp = Tp2;
      Tp2 = u;
    }
    // ---- step bookkeeping and the rebuild schedule ----
    const int m1 = m - 1;
    const bool rebuild = (s_steps == 0);
    if (rebuild) {
      for (int p = tid; p < m1; p += NJ_THREADS) {
        const int r = act[p];
        double s = 0.0;
        for (int j = 0; j < m1; ++j) s += (double)D[(int64_t)act[j] * ld + r];
        T[r] = s;
        Tp[p] = s;
      }
    }
    __syncthreads();
    if (tid == 0) {
      if (rebuild) {
        s_steps = (m1 < 200) ? INT_MAX : (int)((float)m1 * a.ratio);
        s_R = s_next + 1;
      } else {
        s_steps -= 1;
      }
      s_m = m1;
      s_next += 1;
    }
    __syncthreads();
  }
}

// ---- launches per step, the scan over the whole device (large N) ----
//
// The same steps as nj_run, split so that the O(m^2) scan of a step runs on
// many workgroups: nj_scan_kernel (every wave takes rows pa round robin over
// the grid, a best candidate per workgroup), then nj_join_kernel (one
// workgroup: the winner from the partial bests, reduced with the same total
// order, so the winner, and every value after it, equals nj_run's; then the
// join), then, at the steps of the rebuild schedule, which the host replays
// from N and the ratio, nj_rebuild_kernel (T of every active row, a thread
// per row). No workgroup waits on another; the stream orders the steps.
// (One launch per step whose last workgroup ran the join, after an
// agent-scope release by every workgroup, measured slower: 185 against 126
// ms for 4,000 sequences.)

// T[r] = sum over j of D(j, r) in fp64, j in storage order (all nodes active)
__global__ void nj_rowsum_kernel(const float* __restrict__ D, int ld, int n0, double* __restrict__ T,
                                 double* __restrict__ Tp) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n0) return;
  double s = 0.0;
  for (int j = 0; j < n0; ++j) s += (double)D[(int64_t)j * ld + r];
  T[r] = s;
  Tp[r] = s;
}

constexpr int NJ_SCAN_THREADS = 256;

// The step's state comes from the host, which replays the schedule: m active
// nodes, rebuild marker R, the active arrays' half `flip` (so no launch starts
// with a dependent load of it).
__global__ void __launch_bounds__(NJ_SCAN_THREADS) nj_scan_kernel(NjArgs a, int m, int R, int flip,
                                                                  Best* __restrict__ part) {
  __shared__ BestLds L;
  const int* act = a.act + (flip ? a.n0 : 0);
  const int* ids = a.ids + (flip ? a.n0 : 0);
  const double* Tp = a.Tpos + (flip ? a.n0 : 0);
  const double nm2 = (double)(m - 2);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nwaves = gridDim.x * (NJ_SCAN_THREADS / 64);
  Best best = no_best();
  const int units = scan_units(m);
  for (int u = blockIdx.x * (NJ_SCAN_THREADS / 64) + wave; u < units; u += nwaves)
    scan_unit(best, u, m, R, nm2, act, ids, a.D, Tp, a.ld, lane);
  best = workgroup_best(L, best);
  if (threadIdx.x == 0) part[blockIdx.x] = best;
}

__global__ void __launch_bounds__(NJ_JOIN_THREADS) nj_join_kernel(NjArgs a, const Best* __restrict__ part,
                                                                  int nparts, int step, int m, int flip) {
  __shared__ JoinLds L;
  const int tid = threadIdx.x;
  const int n0 = a.n0;
  const int* act = a.act + (flip ? n0 : 0);
  const int* ids = a.ids + (flip ? n0 : 0);
  const double* Tp = a.Tpos + (flip ? n0 : 0);
  // the positions' values and the candidates load together
  JoinBatch<8> B0;
  join_load<NJ_JOIN_THREADS, 8>(B0, tid, m, act, ids, Tp);
  Best best = no_best();
  for (int k = tid; k < nparts; k += NJ_JOIN_THREADS) keep_better(best, part[k]);
  const Best w = workgroup_best(L.b, best);
  join_step<NJ_JOIN_THREADS, 8>(L, w, m, n0 + step, a.D, a.T, a.ld, act, ids, Tp, a.act + (flip ? 0 : n0),
                                a.ids + (flip ? 0 : n0), a.Tpos + (flip ? 0 : n0), &a.joins[step], B0);
}

// T of every active row after a rebuild step (m1 active nodes in half flip)
__global__ void nj_rebuild_kernel(NjArgs a, int m1, int flip) {
  const int* act = a.act + (flip ? a.n0 : 0);
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= m1) return;
  const int r = act[p];
  double s = 0.0;
  for (int j = 0; j < m1; ++j) s += (double)a.D[(int64_t)act[j] * a.ld + r];
  a.T[r] = s;
  a.Tpos[(flip ? a.n0 : 0) + p] = s;
}

hipError_t run_nj_steps(const NjArgs& a, void* parts, int nparts, hipStream_t st) {
  const int n0 = a.n0;
  hipLaunchKernelGGL(nj_rowsum_kernel, dim3((n0 + 255) / 256), dim3(256), 0, st, a.D, a.ld, n0, a.T,
                     a.Tpos);
  // the rebuild schedule of nj_run (SpeciesBlock.cpp:864,1225-1247)
  int steps = (int)((float)n0 * a.ratio);
  steps = steps > 200 ? steps : 200;
  int R = n0;
  hipError_t e;
  for (int step = 0; step < n0 - 1; ++step) {
    const int m = n0 - step, flip = step & 1;
    hipLaunchKernelGGL(nj_scan_kernel, dim3(nparts), dim3(NJ_SCAN_THREADS), 0, st, a, m, R, flip,
                       (Best*)parts);
    hipLaunchKernelGGL(nj_join_kernel, dim3(1), dim3(NJ_JOIN_THREADS), 0, st, a, (const Best*)parts, nparts,
                       step, m, flip);
    if (steps == 0) {
      const int m1 = m - 1;
      hipLaunchKernelGGL(nj_rebuild_kernel, dim3((m1 + 255) / 256), dim3(256), 0, st, a, m1, flip ^ 1);
      R = n0 + step + 1;
      steps = (m1 < 200) ? INT_MAX : (int)((float)m1 * a.ratio);
    } else {
      steps -= 1;
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

int64_t nj_part_bytes(int nparts) { return (int64_t)sizeof(Best) * nparts; }

__global__ void __launch_bounds__(NJ_THREADS) nj_kernel(NjArgs a) { nj_run(a); }

// Independent matrices (jobs of the served form), one workgroup each.
__global__ void __launch_bounds__(NJ_THREADS) nj_batch_kernel(const NjArgs* __restrict__ args) {
  issue_first();
  nj_run(args[blockIdx.x]);
}

hipError_t launch_nj(const NjArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(nj_kernel, dim3(1), dim3(NJ_THREADS), 0, st, a);
  return hipGetLastError();
}

__global__ void scores_to_float_kernel(const int32_t* __restrict__ s, float* __restrict__ d, int64_t n) {
  issue_first();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = (float)s[i];
}

// act and ids (2n each): 0..n-1 in the first half, zeros in the second
__global__ void nj_init_kernel(int* __restrict__ act, int* __restrict__ ids, int n) {
  issue_first();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 2 * n) {
    act[i] = i < n ? i : 0;
    ids[i] = i < n ? i : 0;
  }
}

void launch_nj_init(int* act, int* ids, int n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(nj_init_kernel, dim3((2 * n + 255) / 256), dim3(256), 0, st, act, ids, n);
}

void launch_scores_to_float(const int32_t* s, float* d, int64_t n, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(scores_to_float_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s, d, n);
}

hipError_t launch_nj_batch(const NjArgs* d_args, int count, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(nj_batch_kernel, dim3(count), dim3(NJ_THREADS), 0, st, d_args);
  return hipGetLastError();
}

}  // namespace mst

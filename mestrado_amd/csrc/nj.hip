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

struct Best {
  double q;
  int J, I;  // node ids, I < J
  int pa, pb;  // positions in the active array
};

__device__ __forceinline__ bool better(const Best& a, const Best& b) {
  if (a.q < b.q) return true;
  if (a.q > b.q) return false;
  if (a.J != b.J) return a.J > b.J;
  return a.I > b.I;
}

// q of the pairs (pa, pb), pb > pa, of row pa for the lane's columns
// pb = pa + 1 + lane + 64k, folded into best. Four columns per lane are loaded
// before any is used (act/ids, then D and T through them), so the two
// dependent loads of a column overlap those of the next three.
__device__ __forceinline__ void scan_row(Best& best, int pa, int m, int R, double nm2, const int* act,
                                         const int* ids, const float* D, const double* T, int ld,
                                         int lane) {
  const int ra = act[pa], ia = ids[pa];
  const double Ta = T[ra];
  const float* Drow = D + (int64_t)ra * ld;
  auto fold = [&](int pb, int rb, int ib, float dv, double Tb) {
    const double dq = (double)dv * nm2;
    // pair created by an insertion after the last rebuild: storage order
    const bool swap = (ib >= R) && (rb < ra);
    const double q = swap ? (dq - Tb) - Ta : (dq - Ta) - Tb;
    Best c;
    c.q = q;
    c.I = ia;
    c.J = ib;
    c.pa = pa;
    c.pb = pb;
    if (better(c, best)) best = c;
  };
  int pb = pa + 1 + lane;
  for (; pb + 3 * 64 < m; pb += 4 * 64) {
    int rb[4], ib[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      rb[k] = act[pb + 64 * k];
      ib[k] = ids[pb + 64 * k];
    }
    float dv[4];
    double tb[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      dv[k] = Drow[rb[k]];
      tb[k] = T[rb[k]];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) fold(pb + 64 * k, rb[k], ib[k], dv[k], tb[k]);
  }
  for (; pb < m; pb += 64) {
    const int rb = act[pb];
    fold(pb, rb, ids[pb], Drow[rb], T[rb]);
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

// The whole join loop of one matrix, run by one workgroup of NJ_THREADS.
__device__ __forceinline__ void nj_run(const NjArgs a) {
  __shared__ float rowv[kNjLdsRow];
  __shared__ Best red[NJ_THREADS / 64];
  __shared__ Best win;
  __shared__ int s_m, s_R, s_steps, s_next;
  const int tid = threadIdx.x;
  const int n0 = a.n0;
  float* D = a.D;
  double* T = a.T;
  int* act = a.act;   // storage rows of the active nodes, id order
  int* ids = a.ids;   // node ids of the active nodes
  int* act2 = a.act + n0;  // ping-pong halves
  int* ids2 = a.ids + n0;
  const int ld = a.ld;

  // initial row sums (buildGblRowSums): one thread per row, id order; D is
  // symmetric so thread r reads column r, coalesced across threads
  for (int r = tid; r < n0; r += NJ_THREADS) {
    double s = 0.0;
    for (int j = 0; j < n0; ++j) s += (double)D[(int64_t)j * ld + r];
    T[r] = s;
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
    Best best;
    best.q = DBL_MAX;
    best.J = -1;
    best.I = -1;
    best.pa = best.pb = -1;
    // one wave per row pa (rows dealt round robin), lanes stride the columns
    const int wave = tid >> 6, lane = tid & 63;
    for (int pa = wave; pa < m - 1; pa += NJ_THREADS / 64)
      scan_row(best, pa, m, R, nm2, act, ids, D, T, ld, lane);
    // wave reduction then workgroup reduction
    for (int off = 32; off > 0; off >>= 1) {
      Best o;
      o.q = __shfl_down(best.q, off);
      o.J = __shfl_down(best.J, off);
      o.I = __shfl_down(best.I, off);
      o.pa = __shfl_down(best.pa, off);
      o.pb = __shfl_down(best.pb, off);
      if (better(o, best)) best = o;
    }
    if ((tid & 63) == 0) red[tid >> 6] = best;
    __syncthreads();
    if (tid == 0) {
      Best b = red[0];
      for (int w = 1; w < NJ_THREADS / 64; ++w)
        if (better(red[w], b)) b = red[w];
      win = b;
    }
    __syncthreads();
    const Best w = win;
    const int ra = act[w.pa], rb = act[w.pb];
    const float d = D[(int64_t)ra * ld + rb];
    if (tid == 0) {
      NjJoin& J = a.joins[step];
      J.node = s_next;
      J.left = w.I;
      J.right = w.J;
      J.d = d;
      J.Ti = T[ra];
      J.Tj = T[rb];
      J.n = m;
      J.q = w.q;
    }
    __syncthreads();
    // ---- new row (modifyMatrix) into storage row ra ----
    for (int p = tid; p < m; p += NJ_THREADS) {
      if (p == w.pa || p == w.pb) continue;
      const int rk = act[p];
      const float Dik = D[(int64_t)ra * ld + rk];
      const float Djk = D[(int64_t)rb * ld + rk];
      const float tmp = ((Dik + Djk) - d) / 2;
      const float dlt = tmp - (Dik + Djk);
      T[rk] += (double)dlt;
      D[(int64_t)ra * ld + rk] = tmp;
      D[(int64_t)rk * ld + ra] = tmp;
      if (m <= kNjLdsRow) rowv[p] = tmp;
    }
    __syncthreads();
    if (tid == 0) T[ra] = new_row_sum(rowv, D + (int64_t)ra * ld, act, m, w.pa, w.pb);
    __syncthreads();
    // ---- compact the active array: drop pa, pb, append the new node ----
    for (int p = tid; p < m; p += NJ_THREADS) {
      if (p == w.pa || p == w.pb) continue;
      const int np = p - (p > w.pa) - (p > w.pb);
      act2[np] = act[p];
      ids2[np] = ids[p];
    }
    if (tid == 0) {
      act2[m - 2] = ra;
      ids2[m - 2] = s_next;
    }
    __syncthreads();
    {
      int* t = act;
      act = act2;
      act2 = t;
      t = ids;
      ids = ids2;
      ids2 = t;
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

// ---- one launch per step, the scan over the whole device (large N) ----
//
// The same steps as nj_run, split so that the O(m^2) scan of a step runs on
// many workgroups: nj_scan_kernel (every wave takes rows pa round robin over
// the grid, a best candidate per workgroup), then nj_join_kernel (one
// workgroup: the winner from the partial bests, the join record, the new row,
// T of the new node, the compacted active list, the bookkeeping), then, at
// the steps the host's copy of the rebuild schedule names, nj_rebuild_kernel
// (T of every active row, a thread per row). The partial bests reduce with
// the same total order as the workgroup reduction, so the winner, and every
// value after it, equals nj_run's.

struct NjState {
  int m, R, next, flip;  // active count, rebuild marker, next node id, active half
};

__device__ __forceinline__ Best wave_best(Best best) {
  for (int off = 32; off > 0; off >>= 1) {
    Best o;
    o.q = __shfl_down(best.q, off);
    o.J = __shfl_down(best.J, off);
    o.I = __shfl_down(best.I, off);
    o.pa = __shfl_down(best.pa, off);
    o.pb = __shfl_down(best.pb, off);
    if (better(o, best)) best = o;
  }
  return best;
}

// T[r] = sum over j of D(j, r) in fp64, j in storage order (all nodes active)
__global__ void nj_rowsum_kernel(const float* __restrict__ D, int ld, int n0, double* __restrict__ T) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n0) return;
  double s = 0.0;
  for (int j = 0; j < n0; ++j) s += (double)D[(int64_t)j * ld + r];
  T[r] = s;
}

constexpr int NJ_SCAN_THREADS = 256;

__global__ void __launch_bounds__(NJ_SCAN_THREADS) nj_scan_kernel(NjArgs a, const NjState* __restrict__ st,
                                                                  Best* __restrict__ part) {
  __shared__ Best red[NJ_SCAN_THREADS / 64];
  const NjState S = *st;
  const int m = S.m, R = S.R;
  const int* act = a.act + (S.flip ? a.n0 : 0);
  const int* ids = a.ids + (S.flip ? a.n0 : 0);
  const float* D = a.D;
  const double* T = a.T;
  const int ld = a.ld;
  const double nm2 = (double)(m - 2);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nwaves = gridDim.x * (NJ_SCAN_THREADS / 64);
  Best best;
  best.q = DBL_MAX;
  best.J = -1;
  best.I = -1;
  best.pa = best.pb = -1;
  for (int pa = blockIdx.x * (NJ_SCAN_THREADS / 64) + wave; pa < m - 1; pa += nwaves)
    scan_row(best, pa, m, R, nm2, act, ids, D, T, ld, lane);
  best = wave_best(best);
  if (lane == 0) red[wave] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    Best b = red[0];
    for (int w = 1; w < NJ_SCAN_THREADS / 64; ++w)
      if (better(red[w], b)) b = red[w];
    part[blockIdx.x] = b;
  }
}

__global__ void __launch_bounds__(NJ_THREADS) nj_join_kernel(NjArgs a, NjState* __restrict__ st,
                                                             const Best* __restrict__ part, int nparts,
                                                             int step, int rebuild) {
  __shared__ float rowv[kNjLdsRow];
  __shared__ Best red[NJ_THREADS / 64];
  __shared__ Best win;
  const int tid = threadIdx.x;
  const NjState S = *st;
  const int n0 = a.n0, m = S.m, ld = a.ld;
  float* D = a.D;
  double* T = a.T;
  const int* act = a.act + (S.flip ? n0 : 0);
  const int* ids = a.ids + (S.flip ? n0 : 0);
  int* act2 = a.act + (S.flip ? 0 : n0);
  int* ids2 = a.ids + (S.flip ? 0 : n0);
  Best best;
  best.q = DBL_MAX;
  best.J = -1;
  best.I = -1;
  best.pa = best.pb = -1;
  for (int k = tid; k < nparts; k += NJ_THREADS)
    if (better(part[k], best)) best = part[k];
  best = wave_best(best);
  if ((tid & 63) == 0) red[tid >> 6] = best;
  __syncthreads();
  if (tid == 0) {
    Best b = red[0];
    for (int w = 1; w < NJ_THREADS / 64; ++w)
      if (better(red[w], b)) b = red[w];
    win = b;
  }
  __syncthreads();
  const Best w = win;
  const int ra = act[w.pa], rb = act[w.pb];
  const float d = D[(int64_t)ra * ld + rb];
  if (tid == 0) {
    NjJoin& J = a.joins[step];
    J.node = S.next;
    J.left = w.I;
    J.right = w.J;
    J.d = d;
    J.Ti = T[ra];
    J.Tj = T[rb];
    J.n = m;
    J.q = w.q;
  }
  __syncthreads();
  for (int p = tid; p < m; p += NJ_THREADS) {
    if (p == w.pa || p == w.pb) continue;
    const int rk = act[p];
    const float Dik = D[(int64_t)ra * ld + rk];
    const float Djk = D[(int64_t)rb * ld + rk];
    const float tmp = ((Dik + Djk) - d) / 2;
    const float dlt = tmp - (Dik + Djk);
    T[rk] += (double)dlt;
    D[(int64_t)ra * ld + rk] = tmp;
    D[(int64_t)rk * ld + ra] = tmp;
    if (m <= kNjLdsRow) rowv[p] = tmp;
  }
  __syncthreads();
  if (tid == 0) T[ra] = new_row_sum(rowv, D + (int64_t)ra * ld, act, m, w.pa, w.pb);
  for (int p = tid; p < m; p += NJ_THREADS) {
    if (p == w.pa || p == w.pb) continue;
    const int np = p - (p > w.pa) - (p > w.pb);
    act2[np] = act[p];
    ids2[np] = ids[p];
  }
  if (tid == 0) {
    act2[m - 2] = ra;
    ids2[m - 2] = S.next;
    NjState o = S;
    if (rebuild) o.R = S.next + 1;
    o.m = m - 1;
    o.next = S.next + 1;
    o.flip = S.flip ^ 1;
    *st = o;
  }
}

// T of every active row after a rebuild step (the active list just compacted)
__global__ void nj_rebuild_kernel(NjArgs a, const NjState* __restrict__ st) {
  const NjState S = *st;
  const int m1 = S.m;
  const int* act = a.act + (S.flip ? a.n0 : 0);
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= m1) return;
  const int r = act[p];
  double s = 0.0;
  for (int j = 0; j < m1; ++j) s += (double)a.D[(int64_t)act[j] * a.ld + r];
  a.T[r] = s;
}

hipError_t run_nj_steps(const NjArgs& a, void* state, void* parts, int nparts, hipStream_t st) {
  const int n0 = a.n0;
  NjState init{n0, n0, n0, 0};
  hipError_t e = hipMemcpyAsync(state, &init, sizeof init, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(nj_rowsum_kernel, dim3((n0 + 255) / 256), dim3(256), 0, st, a.D, a.ld, n0, a.T);
  // the rebuild schedule of nj_run (SpeciesBlock.cpp:864,1225-1247)
  int steps = (int)((float)n0 * a.ratio);
  steps = steps > 200 ? steps : 200;
  for (int step = 0; step < n0 - 1; ++step) {
    const int m = n0 - step;
    const bool rebuild = steps == 0;
    hipLaunchKernelGGL(nj_scan_kernel, dim3(nparts), dim3(NJ_SCAN_THREADS), 0, st, a, (const NjState*)state,
                       (Best*)parts);
    hipLaunchKernelGGL(nj_join_kernel, dim3(1), dim3(NJ_THREADS), 0, st, a, (NjState*)state,
                       (const Best*)parts, nparts, step, rebuild ? 1 : 0);
    if (rebuild) {
      const int m1 = m - 1;
      hipLaunchKernelGGL(nj_rebuild_kernel, dim3((m1 + 255) / 256), dim3(256), 0, st, a, (const NjState*)state);
      steps = (m1 < 200) ? INT_MAX : (int)((float)m1 * a.ratio);
    } else {
      steps -= 1;
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

int64_t nj_state_bytes() { return sizeof(NjState); }
int64_t nj_part_bytes(int nparts) { return (int64_t)sizeof(Best) * nparts; }

__global__ void __launch_bounds__(NJ_THREADS) nj_kernel(NjArgs a) { nj_run(a); }

// Independent matrices (jobs of the served form), one workgroup each.
__global__ void __launch_bounds__(NJ_THREADS) nj_batch_kernel(const NjArgs* __restrict__ args) {
  nj_run(args[blockIdx.x]);
}

hipError_t launch_nj(const NjArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(nj_kernel, dim3(1), dim3(NJ_THREADS), 0, st, a);
  return hipGetLastError();
}

__global__ void scores_to_float_kernel(const int32_t* __restrict__ s, float* __restrict__ d, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = (float)s[i];
}

// act and ids (2n each): 0..n-1 in the first half, zeros in the second
__global__ void nj_init_kernel(int* __restrict__ act, int* __restrict__ ids, int n) {
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

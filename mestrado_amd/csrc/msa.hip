// Progressive alignment along the guide tree (R4): what AlignmentMatch::
// computeAlignment / computePairwise / computeMSA (AlignmentMatch.cu:133-434)
// set out to do, made whole: every internal node aligns the profile of its
// left child (X, columns) against the profile of its right child (Y, rows)
// with the recurrence, boundary, flags and traceback of R1-R3, then inserts
// the gap columns of the path into every member row.
//
// Profile column score (build-defined; for one-hot profiles it is exactly the
// substitution score, so single sequences align exactly as in pairwise R1-R3):
//   f(a)     = count(a) / k in fp32, over the 24 letters (gaps do not score)
//   P_r(a)   = sum_b fY_r(b) * S[b][a]      (b = 0..23 in order, fp32; S is
//              read as S[Y letter][X letter] like the pairwise kernels)
//   s(r, c)  = sum_a fX_c(a) * P_r(a)       (a = 0..23 in order, fp32)
//   score    = rint(s) (round half to even), every product and sum rounded
//              to fp32 separately (no contraction), as the oracle does.
// Group rows are kept in forward column order as codes 0..23 and 24 (gap);
// the DP runs on reversed columns like the pairwise path, so the traceback
// path comes out in forward order.
//
// Merges run in batches (MergeEngine): a batch is any set of merges whose
// inputs exist (a level of one guide tree, or ready merges of many jobs).
// Each stage is one launch over the whole batch, its threads flattened over
// the batch's columns, rows or table entries; a thread finds its merge by a
// binary search over the batch's prefix sums. The flag kernels and the
// traceback take the batch's merges as one job list.
#include <algorithm>
#include <vector>

#include "common.h"
#include "merge.h"

namespace mst {

namespace {

constexpr int kStripRows = 512;  // rows of one flag-kernel strip

__host__ __device__ inline int64_t r256(int64_t x) { return (x + 255) / 256 * 256; }

// largest i in [0, count) with key(i) <= x (keys ascending in i)
template <typename Key>
__device__ int find_merge(int count, int64_t x, Key key) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (key(mid) <= x)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Prefix sums the flattened kernels search, one entry per merge, stored after
// the batch's descriptors.
struct Prefix {
  int64_t col0;   // counts: columns of X then Y of nontrivial merges
  int64_t yrow0;  // P rows (nY of nontrivial merges)
  int64_t item0;  // table work units: (strip, chunk of kTabChunk steps), one per wave
};

// counts[c][25] of column c of a group's k rows (row-major k x len)
__device__ __forceinline__ void count_column(const uint8_t* __restrict__ rows, int k, int len, int c,
                                             int32_t* __restrict__ counts) {
  int32_t cnt[25];
#pragma unroll
  for (int a = 0; a < 25; ++a) cnt[a] = 0;
  for (int r = 0; r < k; ++r) {
    const int v = rows[(int64_t)r * len + c];
#pragma unroll
    for (int a = 0; a < 25; ++a) cnt[a] += (v == a);
  }
#pragma unroll
  for (int a = 0; a < 25; ++a) counts[(int64_t)c * 25 + a] = cnt[a];
}

// P_j(a) of Y column j (forward order): P[j][24]
__device__ __forceinline__ void p_row(const int32_t* __restrict__ countsY, int kY,
                                      const int8_t* __restrict__ S, int j, float* __restrict__ P) {
  float fy[24];
#pragma unroll
  for (int b = 0; b < 24; ++b) fy[b] = (float)countsY[(int64_t)j * 25 + b] / (float)kY;
  for (int a = 0; a < 24; ++a) {
    float acc = 0.0f;
#pragma unroll
    for (int b = 0; b < 24; ++b) {
      const float p = fy[b] * (float)S[24 * b + a];
      acc = acc + p;
    }
    P[(int64_t)j * 24 + a] = acc;
  }
}

// Byte table of S''(r,c) = rint(s(r,c)) + go + ge in the skewed layout of the
// flag kernel: entry ((strip*(n+63) + t)*64 + lane) holds the 8 rows of the
// lane (DP row = strip*512 + lane*8 + i) at DP column c = t - lane. DP row r is
// forward Y column m-1-r and DP column c is forward X column n-1-c.
//
// A wave takes one strip and a chunk of steps: each lane keeps the P rows of
// 4 of its 8 DP rows in registers for the whole chunk (two passes) and reads
// one X frequency row (24 floats, fX = count / kX, written by the counts
// kernel) per step, so P is read once per chunk instead of once per entry.
// Chunks are 64 steps (more waves per merge, for the latency of the merges
// of a served level) or 256 for the longest merges (P reloaded less often).
constexpr int kTabChunk = 64, kTabChunkLong = 256;

__global__ void mb_counts_kernel(const MergeDesc* __restrict__ d, const Prefix* __restrict__ pre,
                                 int count, int64_t total, uint8_t* __restrict__ scratch) {
  issue_first();
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= total) return;
  const int i = find_merge(count, x, [&](int k) { return pre[k].col0; });
  const MergeDesc m = d[i];
  const int64_t local = x - pre[i].col0;
  if (local < m.nX) {
    int32_t* cx = (int32_t*)(scratch + m.cx);
    count_column(m.rx, m.kX, m.nX, (int)local, cx);
    float* fx = (float*)(scratch + m.fx) + local * 24;
#pragma unroll
    for (int a = 0; a < 24; ++a) fx[a] = (float)cx[local * 25 + a] / (float)m.kX;
  } else {
    count_column(m.ry, m.kY, m.nY, (int)(local - m.nX), (int32_t*)(scratch + m.cy));
  }
}

__global__ void mb_p_kernel(const MergeDesc* __restrict__ d, const Prefix* __restrict__ pre,
                            int count, int64_t total, const int8_t* __restrict__ S,
                            uint8_t* __restrict__ scratch) {
  issue_first();
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= total) return;
  const int i = find_merge(count, x, [&](int k) { return pre[k].yrow0; });
  const MergeDesc m = d[i];
  p_row((const int32_t*)(scratch + m.cy), m.kY, S, (int)(x - pre[i].yrow0), (float*)(scratch + m.p));
}

// RP rows per pass: 4 (two passes, ~145 VGPRs: the merges of served levels
// share compute units with other kernels) or 8 (one pass, ~250 VGPRs, half the
// X frequency reads: the longest merges, tchunk == kTabChunkLong). Each
// instance handles the merges of its kind and skips the others.
template <int RP>
__global__ void __launch_bounds__(256) mb_table_kernel(const MergeDesc* __restrict__ d,
                                                       const Prefix* __restrict__ pre, int count,
                                                       int64_t total, int go_plus_ge,
                                                       uint8_t* __restrict__ scratch) {
  issue_first();
  const int lane = threadIdx.x & 63;
  const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (w >= total) return;
  const int i = find_merge(count, w, [&](int k) { return pre[k].item0; });
  const MergeDesc m = d[i];
  if ((m.tchunk == kTabChunkLong) != (RP == 8)) return;
  const int64_t unit = w - pre[i].item0;
  const int n = m.nX, rows = m.nY;
  const int Tn = n + MST_WAVE - 1;
  const int ch = m.tchunk;
  const int nch = (Tn + ch - 1) / ch;
  const int s = (int)(unit / nch);
  const int t0 = (int)(unit - (int64_t)s * nch) * ch;
  const int t1 = t0 + ch < Tn ? t0 + ch : Tn;
  const float* P = (const float*)(scratch + m.p);
  const float* FX = (const float*)(scratch + m.fx);
  uint32_t* tab32 = (uint32_t*)((uint2*)(scratch + m.tab) + (int64_t)s * Tn * MST_WAVE);
#pragma unroll 1
  for (int part = 0; part < 8 / RP; ++part) {
    float pr[RP][24];
#pragma unroll
    for (int k = 0; k < RP; ++k) {
      const int r = s * kStripRows + lane * 8 + part * RP + k;
      const float* p = P + (int64_t)(rows - 1 - (r < rows ? r : rows - 1)) * 24;
#pragma unroll
      for (int a = 0; a < 24; ++a) pr[k][a] = p[a];
    }
    for (int t = t0; t < t1; ++t) {
      const int c = t - lane;
      uint32_t wd[RP / 4];
#pragma unroll
      for (int q = 0; q < RP / 4; ++q) wd[q] = 0;
      if (c >= 0 && c < n) {
        const float4* f4 = (const float4*)(FX + (int64_t)(n - 1 - c) * 24);
        float fx[24];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          const float4 v = f4[q];
          fx[4 * q] = v.x;
          fx[4 * q + 1] = v.y;
          fx[4 * q + 2] = v.z;
          fx[4 * q + 3] = v.w;
        }
#pragma unroll
        for (int k = 0; k < RP; ++k) {
          const int r = s * kStripRows + lane * 8 + part * RP + k;
          float acc = 0.0f;
#pragma unroll
          for (int a = 0; a < 24; ++a) {
            const float q = fx[a] * pr[k][a];
            acc = acc + q;
          }
          const int v = r < rows ? (int)__builtin_rintf(acc) + go_plus_ge : 0;
          wd[k >> 2] |= ((uint32_t)v & 0xFFu) << (8 * (k & 3));
        }
      }
      uint32_t* e = tab32 + ((int64_t)t * MST_WAVE + lane) * 2;
      if (RP == 8) {
        *(uint2*)e = make_uint2(wd[0], wd[RP / 4 - 1]);
      } else {
        e[part] = wd[0];
      }
    }
  }
}

// Source column of every output column: sx[t] = X column or -1 (gap in X),
// same for Y; lens[i] = path length. One workgroup per merge scans the path
// in chunks of 256 steps (wave ballots, then the waves' totals).
constexpr int kPathThreads = 256;

__global__ void __launch_bounds__(kPathThreads) mb_path_kernel(const MergeDesc* __restrict__ d,
                                                               const int32_t* __restrict__ tlens,
                                                               uint8_t* __restrict__ scratch,
                                                               int32_t* __restrict__ lens) {
  issue_first();
  __shared__ int wx[kPathThreads / MST_WAVE], wy[kPathThreads / MST_WAVE];
  const int i = blockIdx.x;
  const MergeDesc m = d[i];
  int32_t* sx = (int32_t*)(scratch + m.sxy);
  int32_t* sy = sx + (m.nX + m.nY);
  const uint8_t* ops = scratch + m.ops;
  const int L = m.job < 0 ? m.nX + m.nY : tlens[m.job];
  const int lane = lane_id(), wv = threadIdx.x / MST_WAVE;
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  int bx = 0, by = 0;  // X and Y columns consumed before the chunk
  for (int t0 = 0; t0 < L; t0 += kPathThreads) {
    const int t = t0 + (int)threadIdx.x;
    int op = -1;
    if (t < L) op = m.job < 0 ? (m.nX > 0 ? OP_LEFT : OP_UP) : ops[t];
    const bool ax = op >= 0 && op != OP_UP, ay = op >= 0 && op != OP_LEFT;
    const uint64_t mx = __ballot(ax), my = __ballot(ay);
    if (lane == 0) {
      wx[wv] = __popcll(mx);
      wy[wv] = __popcll(my);
    }
    __syncthreads();
    int ox = bx, oy = by, tx = 0, ty = 0;
    for (int w = 0; w < kPathThreads / MST_WAVE; ++w) {
      if (w < wv) {
        ox += wx[w];
        oy += wy[w];
      }
      tx += wx[w];
      ty += wy[w];
    }
    if (t < L) {
      sx[t] = ax ? ox + __popcll(mx & below) : -1;
      sy[t] = ay ? oy + __popcll(my & below) : -1;
    }
    bx += tx;
    by += ty;
    __syncthreads();  // totals are rewritten by the next chunk
  }
  if (threadIdx.x == 0) lens[i] = L;
}

// One workgroup per output row: X's kX rows then Y's kY rows, L columns.
__global__ void mb_rows_kernel(const MergeDesc* __restrict__ d, int count,
                               const uint8_t* __restrict__ scratch, const int32_t* __restrict__ lens) {
  issue_first();
  const int64_t r = blockIdx.x;
  const int i = find_merge(count, r, [&](int k) { return d[k].row0; });
  const MergeDesc m = d[i];
  const int lr = (int)(r - m.row0);
  const int L = lens[i];
  const int32_t* sx = (const int32_t*)(scratch + m.sxy);
  const int32_t* map = lr < m.kX ? sx : sx + (m.nX + m.nY);
  const uint8_t* src = lr < m.kX ? m.rx + (int64_t)lr * m.nX : m.ry + (int64_t)(lr - m.kX) * m.nY;
  uint8_t* out = m.out + (int64_t)lr * L;
  for (int t = threadIdx.x; t < L; t += blockDim.x) {
    const int c = map[t];
    out[t] = c < 0 ? (uint8_t)24 : src[c];
  }
}

// One workgroup per leaf: forward codes from the reversed code block.
__global__ void mb_leaf_kernel(const int8_t* __restrict__ codes, const LeafDesc* __restrict__ l) {
  issue_first();
  const LeafDesc q = l[blockIdx.x];
  for (int t = threadIdx.x; t < q.len; t += blockDim.x) q.dst[t] = (uint8_t)codes[q.src + q.len - 1 - t];
}

__global__ void mb_zero_kernel(unsigned int* p, int n) {
  const int i = threadIdx.x;
  if (i < n) p[i] = 0u;
}

int strips_of(int m) { return (m + kStripRows - 1) / kStripRows; }

}  // namespace

int64_t merge_scratch_bytes(int nX, int nY) {
  const int64_t cols = (int64_t)nX + nY;
  int64_t b = r256(8 * cols);  // column maps
  if (nX > 0 && nY > 0) {
    const int64_t words = flag_words(nX, nY);
    b += r256(100 * (int64_t)nX) + r256(100 * (int64_t)nY) + r256(96 * (int64_t)nY) +
         r256(96 * (int64_t)nX) + r256(8 * words) +
         r256(4 * words) + r256(cols);
  }
  return b;
}

hipError_t MergeEngine::grow(void** p, int64_t* cap, int64_t bytes, bool pinned) {
  if (bytes <= *cap) return hipSuccess;
  const int64_t want = std::max<int64_t>(bytes, *cap * 3 / 2);
  hipError_t e;
  if (*p) {
    e = pinned ? hipHostFree(*p) : hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (e != hipSuccess) return e;
  }
  e = pinned ? hipHostMalloc(p, want, hipHostMallocDefault) : hipMalloc(p, want);
  if (e != hipSuccess) return e;
  *cap = want;
  return hipSuccess;
}

MergeEngine::~MergeEngine() {
  void* dev[] = {d_scratch_, d_desc_, d_jobs_, d_res_,    d_lens_,  d_bufs_,
                 d_counter_, d_leaf_, d_tiles_, d_trows_, d_tstate_, d_items_};
  for (void* p : dev)
    if (p) hipFree(p);
  void* host[] = {h_desc_, h_jobs_, h_lens_, h_leaf_, h_tiles_, h_items_};
  for (void* p : host)
    if (p) hipHostFree(p);
}

hipError_t MergeEngine::leaves(const int8_t* codes, const LeafDesc* l, int count, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipError_t e;
  const int64_t bytes = sizeof(LeafDesc) * (int64_t)count;
  if ((e = grow((void**)&d_leaf_, &leaf_cap_, bytes, false)) != hipSuccess) return e;
  if ((e = grow((void**)&h_leaf_, &hleaf_cap_, bytes, true)) != hipSuccess) return e;
  std::copy(l, l + count, h_leaf_);
  if ((e = hipMemcpyAsync(d_leaf_, h_leaf_, bytes, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
  hipLaunchKernelGGL(mb_leaf_kernel, dim3(count), dim3(256), 0, st, codes, d_leaf_);
  return hipGetLastError();
}

hipError_t MergeEngine::enqueue(const MergeSpec* m, int count, const int8_t* d_S, int go, int ge,
                                int num_cus, hipStream_t st, int f16_smax) {
  if (count <= 0) return hipSuccess;
  // flag jobs: one strip, 2..4 strips (pipeline of 4 waves), more (8 waves),
  // long ones (tiles over anti-diagonal launches), and last those of the
  // packed kernel (one or two strips whose 16-bit values cannot wrap)
  std::vector<int> cls(count, -1);
  int64_t nb[5] = {0, 0, 0, 0, 0};
  for (int i = 0; i < count; ++i) {
    if (m[i].nX > 0 && m[i].nY > 0) {
      const int s = strips_of(m[i].nY);
      const int64_t R = m[i].nY, n = m[i].nX;
      const bool packed = f16_smax >= 0 && s <= 2 && !tiled_job(m[i].nX, m[i].nY) &&
                          (int64_t)f16_smax * (std::min(R, n) + 1) + (int64_t)ge * (R + n + 2) + f16_smax +
                                  go + ge <= 32767;
      cls[i] = packed ? 4 : (s == 1 ? 0 : (s <= 4 ? 1 : (tiled_job(m[i].nX, m[i].nY) ? 3 : 2)));
      nb[cls[i]] += 1;
    }
  }
  const int64_t njobs = nb[0] + nb[1] + nb[2] + nb[3] + nb[4];
  int64_t next[5] = {0, nb[0], nb[0] + nb[1], nb[0] + nb[1] + nb[2], nb[0] + nb[1] + nb[2] + nb[3]};
  const int64_t f0 = next[4];  // the packed jobs: [f0, njobs), by strips then columns, descending
  std::vector<int> f16_job(count, -1);
  {
    std::vector<int> f;
    for (int i = 0; i < count; ++i)
      if (cls[i] == 4) f.push_back(i);
    std::stable_sort(f.begin(), f.end(), [&](int x, int y) {
      const int sx = strips_of(m[x].nY), sy = strips_of(m[y].nY);
      if (sx != sy) return sx > sy;
      return m[x].nX > m[y].nX;
    });
    for (size_t r = 0; r < f.size(); ++r) f16_job[f[r]] = (int)(f0 + (int64_t)r);
  }
  // layout of the scratch block and the prefix sums
  const int64_t desc_bytes = (int64_t)count * (sizeof(MergeDesc) + sizeof(Prefix));
  hipError_t e;
  if ((e = grow((void**)&h_desc_, &hdesc_cap_, desc_bytes, true)) != hipSuccess) return e;
  if ((e = grow((void**)&h_jobs_, &hjobs_cap_, sizeof(AlignJob) * std::max<int64_t>(njobs, 1), true)) !=
      hipSuccess)
    return e;
  MergeDesc* hd = h_desc_;
  Prefix* hp = (Prefix*)(h_desc_ + count);
  int64_t off = 0, col0 = 0, yrow0 = 0, item0 = 0, row0 = 0;
  int maxn = 1;
  for (int i = 0; i < count; ++i) {
    MergeDesc& q = hd[i];
    q.rx = m[i].rx;
    q.ry = m[i].ry;
    q.out = m[i].out;
    q.kX = m[i].kX;
    q.nX = m[i].nX;
    q.kY = m[i].kY;
    q.nY = m[i].nY;
    q.tchunk = tiled_job(q.nX, q.nY) ? kTabChunkLong : kTabChunk;
    q.row0 = row0;
    row0 += q.kX + q.kY;
    hp[i].col0 = col0;
    hp[i].yrow0 = yrow0;
    hp[i].item0 = item0;
    q.cx = q.cy = q.p = q.fx = q.tab = q.ops = 0;
    q.sxy = off;
    off += r256(8 * ((int64_t)q.nX + q.nY));
    if (cls[i] < 0) {
      q.job = -1;
      continue;
    }
    const int64_t words = flag_words(q.nX, q.nY);
    q.cx = off;
    off += r256(100 * (int64_t)q.nX);
    q.cy = off;
    off += r256(100 * (int64_t)q.nY);
    q.p = off;
    off += r256(96 * (int64_t)q.nY);
    q.fx = off;
    off += r256(96 * (int64_t)q.nX);
    q.tab = off;
    off += r256(8 * words);
    const int64_t fl = off;
    off += r256(4 * words);
    q.ops = off;
    off += r256((int64_t)q.nX + q.nY);
    col0 += (int64_t)q.nX + q.nY;
    yrow0 += q.nY;
    item0 += (int64_t)strips_of(q.nY) * ((q.nX + MST_WAVE - 1 + q.tchunk - 1) / q.tchunk);
    q.job = cls[i] == 4 ? f16_job[i] : (int)next[cls[i]]++;
    AlignJob& j = h_jobs_[q.job];
    j.n = q.nX;
    j.m = q.nY;
    j.fmt = cls[i] == 4 ? FMT_PLANES : FMT_NIBBLE;
    j.k = 8;
    j.x_off = j.y_off = 0;
    j.tab_off = q.tab;
    j.flag_off = fl / 4;
    j.ops_off = q.ops;
    maxn = std::max(maxn, q.nX);
  }
  const int64_t grid1 = std::max<int64_t>(1, std::min<int64_t>((int64_t)num_cus * 16, nb[0]));
  const int64_t grid2 = std::max<int64_t>(1, std::min<int64_t>((int64_t)num_cus * 2, nb[1]));
  const int64_t grid3 = std::max<int64_t>(1, std::min<int64_t>((int64_t)num_cus, nb[2]));
  std::vector<F16Item> items;
  plan_flag16_items(h_jobs_ + f0, nb[4], items);
  const int64_t grid4 = std::max<int64_t>(1, std::min<int64_t>((int64_t)num_cus * 12, (int64_t)items.size()));
  const int64_t bufcap = maxn;
  // the flag kernels run one after the other on the stream: one set of row buffers
  const int64_t bufs = bufcap * std::max(std::max(std::max(2 * grid1, grid2), grid3), 2 * grid4);
  if ((e = grow((void**)&d_scratch_, &scratch_cap_, std::max<int64_t>(off, 256), false)) != hipSuccess)
    return e;
  if ((e = grow((void**)&d_desc_, &desc_cap_, desc_bytes, false)) != hipSuccess) return e;
  if ((e = grow((void**)&d_jobs_, &jobs_cap_, sizeof(AlignJob) * std::max<int64_t>(njobs, 1), false)) !=
      hipSuccess)
    return e;
  if ((e = grow((void**)&d_res_, &res_cap_, sizeof(int32_t) * 2 * std::max<int64_t>(njobs, 1), false)) !=
      hipSuccess)
    return e;
  if ((e = grow((void**)&d_lens_, &lens_cap_, sizeof(int32_t) * count, false)) != hipSuccess) return e;
  if ((e = grow((void**)&h_lens_, &hlens_cap_, sizeof(int32_t) * count, true)) != hipSuccess) return e;
  if ((e = grow((void**)&d_bufs_, &bufs_cap_, sizeof(int2) * bufs, false)) != hipSuccess) return e;
  if ((e = grow((void**)&d_counter_, &counter_cap_, sizeof(unsigned int) * 2, false)) != hipSuccess)
    return e;
  const MergeDesc* dd = d_desc_;
  const Prefix* dp = (const Prefix*)(d_desc_ + count);
  if ((e = hipMemcpyAsync(d_desc_, h_desc_, desc_bytes, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
  if (njobs > 0) {
    if ((e = hipMemcpyAsync(d_jobs_, h_jobs_, sizeof(AlignJob) * njobs, hipMemcpyHostToDevice, st)) !=
        hipSuccess)
      return e;
    hipLaunchKernelGGL(mb_zero_kernel, dim3(1), dim3(64), 0, st, d_counter_, 2);
    hipLaunchKernelGGL(mb_counts_kernel, dim3((unsigned)((col0 + 255) / 256)), dim3(256), 0, st, dd, dp,
                       count, col0, d_scratch_);
    hipLaunchKernelGGL(mb_p_kernel, dim3((unsigned)((yrow0 + 255) / 256)), dim3(256), 0, st, dd, dp, count,
                       yrow0, d_S, d_scratch_);
    hipLaunchKernelGGL(mb_table_kernel<4>, dim3((unsigned)((item0 + 3) / 4)), dim3(256), 0, st, dd, dp,
                       count, item0, go + ge, d_scratch_);
    if (nb[3] > 0)
      hipLaunchKernelGGL(mb_table_kernel<8>, dim3((unsigned)((item0 + 3) / 4)), dim3(256), 0, st, dd, dp,
                         count, item0, go + ge, d_scratch_);
    AlignArgs a;
    a.jobs = d_jobs_;
    a.seqs = nullptr;
    a.tab = d_scratch_;
    a.subst = d_S;
    a.flags = (uint32_t*)d_scratch_;
    a.scores = d_res_;
    a.bufs = d_bufs_;
    a.bufcap = bufcap;
    a.go = go;
    a.ge = ge;
    if (nb[0] > 0) {
      a.first = 0;
      a.last = nb[0];
      a.counter = d_counter_;
      launch_align(SRC_TAB, a, (int)grid1, st);
    }
    if (nb[1] > 0) {
      a.first = nb[0];
      a.last = nb[0] + nb[1];
      a.counter = d_counter_;
      launch_align_pipe(SRC_TAB, 4, a, (int)grid2, st);
    }
    if (nb[2] > 0) {
      a.first = nb[0] + nb[1];
      a.last = nb[0] + nb[1] + nb[2];
      a.counter = d_counter_ + 1;
      launch_align_pipe(SRC_TAB, 8, a, (int)grid3, st);
    }
    if (nb[4] > 0) {
      const int64_t ib = sizeof(F16Item) * (int64_t)items.size();
      if ((e = grow((void**)&d_items_, &items_cap_, ib, false)) != hipSuccess) return e;
      if ((e = grow((void**)&h_items_, &hitems_cap_, ib, true)) != hipSuccess) return e;
      std::copy(items.begin(), items.end(), h_items_);
      if ((e = hipMemcpyAsync(d_items_, h_items_, ib, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
      F16Args f;
      f.jobs = d_jobs_ + f0;
      f.items = d_items_;
      f.nitems = (int64_t)items.size();
      f.tab = d_scratch_;
      f.flags = (uint32_t*)d_scratch_;
      f.scores = d_res_ + f0;
      f.bufs = d_bufs_;
      f.bufcap = bufcap;
      f.go = go;
      f.ge = ge;
      launch_flag16(f, (int)grid4, st);
    }
    if (nb[3] > 0) {
      std::vector<TileJob> tiles;
      std::vector<int64_t> diag_off;
      int64_t trows = 0, tstate = 0;
      plan_tiles(h_jobs_, f0 - nb[3], f0, kTileWidth, tiles, diag_off, &trows, &tstate);
      if ((e = grow((void**)&d_tiles_, &tiles_cap_, sizeof(TileJob) * tiles.size(), false)) != hipSuccess) return e;
      if ((e = grow((void**)&h_tiles_, &htiles_cap_, sizeof(TileJob) * tiles.size(), true)) != hipSuccess) return e;
      if ((e = grow((void**)&d_trows_, &trows_cap_, sizeof(int2) * trows, false)) != hipSuccess) return e;
      if ((e = grow((void**)&d_tstate_, &tstate_cap_, sizeof(int) * tstate, false)) != hipSuccess) return e;
      std::copy(tiles.begin(), tiles.end(), h_tiles_);
      if ((e = hipMemcpyAsync(d_tiles_, h_tiles_, sizeof(TileJob) * tiles.size(), hipMemcpyHostToDevice, st)) !=
          hipSuccess)
        return e;
      TileArgs ta;
      ta.jobs = d_jobs_;
      ta.tiles = nullptr;
      ta.seqs = nullptr;
      ta.tab = d_scratch_;
      ta.subst = d_S;
      ta.flags = (uint32_t*)d_scratch_;
      ta.scores = d_res_;
      ta.rows = d_trows_;
      ta.state = d_tstate_;
      ta.go = go;
      ta.ge = ge;
      ta.tw = kTileWidth;
      launch_tile_diagonals(SRC_TAB, ta, d_tiles_, diag_off, st);
    }
    TraceArgs t;
    t.jobs = d_jobs_;
    t.first = 0;
    t.last = njobs;
    t.flags = (const uint32_t*)d_scratch_;
    t.ops = d_scratch_;
    t.lens = d_res_ + njobs;
    launch_traceback(t, st);
  }
  hipLaunchKernelGGL(mb_path_kernel, dim3(count), dim3(kPathThreads), 0, st, dd,
                     (const int32_t*)(d_res_ + njobs), d_scratch_, d_lens_);
  if (row0 > 0)
    hipLaunchKernelGGL(mb_rows_kernel, dim3((unsigned)row0), dim3(256), 0, st, dd, count,
                       (const uint8_t*)d_scratch_, (const int32_t*)d_lens_);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return hipMemcpyAsync(h_lens_, d_lens_, sizeof(int32_t) * count, hipMemcpyDeviceToHost, st);
}

}  // namespace mst

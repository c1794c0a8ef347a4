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
#include "common.h"
#include "msa.h"

namespace mst {

// counts[col][25] of a group's k rows (row-major k x len).
__global__ void profile_counts_kernel(const uint8_t* __restrict__ rows, int k, int len,
                                      int32_t* __restrict__ counts) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= len) return;
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

// P_r(a) for every Y column (forward order): P[col][24].
__global__ void profile_p_kernel(const int32_t* __restrict__ countsY, int m, int kY,
                                 const int8_t* __restrict__ S, float* __restrict__ P) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
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
__global__ void profile_table_kernel(const int32_t* __restrict__ countsX, int n, int kX,
                                     const float* __restrict__ P, int m, int go_plus_ge,
                                     uint2* __restrict__ tab) {
  const int lane = threadIdx.x & 63;
  const int64_t item = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int Tn = n + MST_WAVE - 1;
  const int nstrips = (m + 511) / 512;
  if (item >= (int64_t)nstrips * Tn) return;
  const int s = (int)(item / Tn);
  const int t = (int)(item - (int64_t)s * Tn);
  const int c = t - lane;
  uint32_t w[2] = {0, 0};
  if (c >= 0 && c < n) {
    const int xc = n - 1 - c;
    float fx[24];
#pragma unroll
    for (int a = 0; a < 24; ++a) fx[a] = (float)countsX[(int64_t)xc * 25 + a] / (float)kX;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int r = s * 512 + lane * 8 + i;
      int v = 0;
      if (r < m) {
        const float* p = P + (int64_t)(m - 1 - r) * 24;
        float acc = 0.0f;
#pragma unroll
        for (int a = 0; a < 24; ++a) {
          const float q = fx[a] * p[a];
          acc = acc + q;
        }
        v = (int)__builtin_rintf(acc) + go_plus_ge;
      }
      w[i >> 2] |= ((uint32_t)v & 0xFFu) << (8 * (i & 3));
    }
  }
  tab[((int64_t)s * Tn + t) * MST_WAVE + lane] = make_uint2(w[0], w[1]);
}

// Source column of every output column: sx[t] = X column or -1 (gap), same
// for Y; one thread walks the path (paths are at most a few 10^4 long).
__global__ void path_columns_kernel(const uint8_t* __restrict__ ops, const int32_t* __restrict__ len,
                                    int32_t* __restrict__ sx, int32_t* __restrict__ sy) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int L = *len;
  int x = 0, y = 0;
  for (int t = 0; t < L; ++t) {
    const int op = ops[t];
    sx[t] = (op != 2) ? x++ : -1;  // 2 = UP: gap in X
    sy[t] = (op != 1) ? y++ : -1;  // 1 = LEFT: gap in Y
  }
}

// New rows: X's kX rows then Y's kY rows, L columns.
__global__ void merge_rows_kernel(const uint8_t* __restrict__ rx, int kX, int nx,
                                  const uint8_t* __restrict__ ry, int kY, int ny,
                                  const int32_t* __restrict__ sx, const int32_t* __restrict__ sy,
                                  int L, uint8_t* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = blockIdx.y;
  if (t >= L) return;
  uint8_t v;
  if (row < kX) {
    const int c = sx[t];
    v = c < 0 ? (uint8_t)24 : rx[(int64_t)row * nx + c];
  } else {
    const int c = sy[t];
    v = c < 0 ? (uint8_t)24 : ry[(int64_t)(row - kX) * ny + c];
  }
  out[(int64_t)row * L + t] = v;
}

// Leaf rows: forward codes of sequence s from the reversed set codes.
__global__ void leaf_rows_kernel(const int8_t* __restrict__ codes, int64_t off, int len,
                                 uint8_t* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < len) out[t] = (uint8_t)codes[off + len - 1 - t];
}

void launch_profile_counts(const uint8_t* rows, int k, int len, int32_t* counts, hipStream_t st) {
  if (len > 0)
    hipLaunchKernelGGL(profile_counts_kernel, dim3((len + 255) / 256), dim3(256), 0, st, rows, k, len,
                       counts);
}

void launch_profile_p(const int32_t* countsY, int m, int kY, const int8_t* S, float* P,
                      hipStream_t st) {
  if (m > 0)
    hipLaunchKernelGGL(profile_p_kernel, dim3((m + 255) / 256), dim3(256), 0, st, countsY, m, kY, S, P);
}

void launch_profile_table(const int32_t* countsX, int n, int kX, const float* P, int m,
                          int go_plus_ge, uint2* tab, hipStream_t st) {
  const int64_t items = (int64_t)((m + 511) / 512) * (n + MST_WAVE - 1);
  const int per_block = 4;  // waves per block
  hipLaunchKernelGGL(profile_table_kernel, dim3((unsigned)((items + per_block - 1) / per_block)),
                     dim3(64 * per_block), 0, st, countsX, n, kX, P, m, go_plus_ge, tab);
}

void launch_path_columns(const uint8_t* ops, const int32_t* len, int32_t* sx, int32_t* sy,
                         hipStream_t st) {
  hipLaunchKernelGGL(path_columns_kernel, dim3(1), dim3(64), 0, st, ops, len, sx, sy);
}

void launch_merge_rows(const uint8_t* rx, int kX, int nx, const uint8_t* ry, int kY, int ny,
                       const int32_t* sx, const int32_t* sy, int L, uint8_t* out, hipStream_t st) {
  if (L > 0)
    hipLaunchKernelGGL(merge_rows_kernel, dim3((L + 255) / 256, kX + kY), dim3(256), 0, st, rx, kX,
                       nx, ry, kY, ny, sx, sy, L, out);
}

void launch_leaf_rows(const int8_t* codes, int64_t off, int len, uint8_t* out, hipStream_t st) {
  if (len > 0)
    hipLaunchKernelGGL(leaf_rows_kernel, dim3((len + 255) / 256), dim3(256), 0, st, codes, off, len, out);
}

}  // namespace mst

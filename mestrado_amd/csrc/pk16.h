// Packed 16-bit helpers shared by the all-pairs kernels that score two pairs
// per wavefront (pair a in the low, pair b in the high half of a dword).
#pragma once
#include "common.h"

namespace mst {

typedef short s2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ s2 as_s2(uint32_t v) { return __builtin_bit_cast(s2, v); }
__device__ __forceinline__ uint32_t as_u(s2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ s2 pmax(s2 a, s2 b) { return __builtin_elementwise_max(a, b); }

// dwords of profile per (residue, lane): two bytes per row, padded to 1/2/4/8
template <int K>
struct KD {
  static constexpr int raw = (K + 1) / 2;
  static constexpr int v = raw <= 1 ? 1 : (raw <= 2 ? 2 : (raw <= 4 ? 4 : 8));
};

// The wide form: one dword per row, the (pair a, pair b) bytes already
// zero-extended into the two 16-bit halves, so a cell update reads its row's
// profile word as is (no v_perm per row); twice the LDS of the byte form.
template <int K, bool WIDE>
struct KDW {
  static constexpr int v =
      !WIDE ? KD<K>::v : (K <= 1 ? 1 : (K <= 2 ? 2 : (K <= 4 ? 4 : (K <= 8 ? 8 : (K <= 12 ? 12 : 16)))));
};

template <int N>
__device__ __forceinline__ void read_dwords(const uint8_t* p, uint32_t* w) {
  if (N == 1) {
    w[0] = *(const uint32_t*)p;
  } else if (N == 2) {
    uint2 v = *(const uint2*)p;
    w[0] = v.x;
    w[1] = v.y;
  } else {
#pragma unroll
    for (int q = 0; q < N / 4; ++q) {
      uint4 v = *(const uint4*)(p + 16 * q);
      w[4 * q] = v.x;
      w[4 * q + 1] = v.y;
      w[4 * q + 2] = v.z;
      w[4 * q + 3] = v.w;
    }
  }
}

// row i's (pair a, pair b) bytes zero-extended into the two halves
__device__ __forceinline__ s2 prow(const uint32_t* w, int i) {
  const uint32_t sel = (i & 1) ? 0x0C030C02u : 0x0C010C00u;
  return as_s2(__builtin_amdgcn_perm(w[i >> 1], w[i >> 1], sel));
}
template <bool WIDE>
__device__ __forceinline__ s2 prow_w(const uint32_t* w, int i) {
  return WIDE ? as_s2(w[i]) : prow(w, i);
}

__device__ __forceinline__ uint32_t shr1z(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_WAVE_SHR1, 0xF, 0xF, true);
}
__device__ __forceinline__ uint32_t shr1o(uint32_t old, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, DPP_WAVE_SHR1, 0xF, 0xF, false);
}

template <int K>
struct Lane2 {
  s2 Hm[K], E[K];
  uint32_t Fout, Hout, hdiag;
};

}  // namespace mst

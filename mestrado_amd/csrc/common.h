// Shared definitions of the HIP kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MST_NEG_INF (-(1 << 29))
#define MST_LETTERS 24
#define MST_WAVE 64

// DPP controls of the GFX9 family (wave-wide shift by one lane).
#define DPP_WAVE_SHR1 0x138
#define DPP_WAVE_ROR1 0x13C

namespace mst {

// Raises the calling wave's instruction-issue priority (s_setprio 3): where
// a latency-bound kernel shares SIMDs with a throughput kernel's waves (the
// served loop's merge batches and NJ beside its score rounds), its waves
// issue first; on an otherwise idle device it changes nothing.
__device__ __forceinline__ void issue_first() { __builtin_amdgcn_s_setprio(3); }

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// v_mov_b32_dpp wave_shr:1 with bound_ctrl off: lane 0 keeps `old`.
__device__ __forceinline__ int shr1(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, DPP_WAVE_SHR1, 0xF, 0xF, false);
}

__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int imax3(int a, int b, int c) { return imax(imax(a, b), c); }

// Agent-scope fence: orders this wave's global stores before later loads of
// the same addresses by any wave (used between strips and between steps).
__device__ __forceinline__ void agent_fence() {
  __atomic_thread_fence(__ATOMIC_SEQ_CST);
}

}  // namespace mst

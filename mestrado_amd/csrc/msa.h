#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mst {

void launch_profile_counts(const uint8_t* rows, int k, int len, int32_t* counts, hipStream_t st);
void launch_profile_p(const int32_t* countsY, int m, int kY, const int8_t* S, float* P,
                      hipStream_t st);
void launch_profile_table(const int32_t* countsX, int n, int kX, const float* P, int m,
                          int go_plus_ge, uint2* tab, hipStream_t st);
void launch_path_columns(const uint8_t* ops, const int32_t* len, int32_t* sx, int32_t* sy,
                         hipStream_t st);
void launch_merge_rows(const uint8_t* rx, int kX, int nx, const uint8_t* ry, int kY, int ny,
                       const int32_t* sx, const int32_t* sy, int L, uint8_t* out, hipStream_t st);
void launch_leaf_rows(const int8_t* codes, int64_t off, int len, uint8_t* out, hipStream_t st);

}  // namespace mst

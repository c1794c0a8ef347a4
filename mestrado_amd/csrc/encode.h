#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mst {
hipError_t init_encode_table();
void launch_encode(const char* ascii, const int64_t* off, int n, int8_t* codes,
                   hipStream_t st);
}  // namespace mst

// Residue encoding (R0): ASCII -> 0..23 codes of the substitution matrix
// letter order ARNDCQEGHILKMFPSTWYVBZX*, each sequence reversed, as in
// AlignmentScore::parse (AlignmentScore.cu:143-178, table :10-27). Lower case
// letters fold to upper case and letters outside the table map to X (22).
#include "common.h"
#include "encode.h"

namespace mst {

__constant__ int8_t kEncTable[128];

static const char kLettersHost[] = "ARNDCQEGHILKMFPSTWYVBZX*";

// One workgroup per sequence; reversal and encoding in one pass.
__global__ void encode_reverse_kernel(const char* __restrict__ ascii,
                                      const int64_t* __restrict__ off, int n,
                                      int8_t* __restrict__ codes) {
  const int s = blockIdx.x;
  if (s >= n) return;
  const int64_t o = off[s];
  const int64_t len = off[s + 1] - o;
  for (int64_t k = threadIdx.x; k < len; k += blockDim.x) {
    unsigned char ch = (unsigned char)ascii[o + len - 1 - k];
    codes[o + k] = (ch < 128) ? kEncTable[ch] : (int8_t)22;
  }
}

hipError_t init_encode_table() {
  int8_t t[128];
  for (int i = 0; i < 128; ++i) t[i] = 22;
  for (int k = 0; k < 24; ++k) {
    unsigned char c = (unsigned char)kLettersHost[k];
    t[c] = (int8_t)k;
    if (c >= 'A' && c <= 'Z') t[c - 'A' + 'a'] = (int8_t)k;
  }
  return hipMemcpyToSymbol(HIP_SYMBOL(kEncTable), t, sizeof t);
}

void launch_encode(const char* ascii, const int64_t* off, int n, int8_t* codes,
                   hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(encode_reverse_kernel, dim3(n), dim3(256), 0, st, ascii, off, n, codes);
}

}  // namespace mst

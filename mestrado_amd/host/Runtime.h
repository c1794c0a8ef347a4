// Process-wide device context shared by the drop-in classes: the reference
// kept gap penalties, window size and the substitution matrix in __constant__
// symbols (definitions.cuh:170-196); here they live in one mst_ctx, created
// on first use with the settings of configure().
#ifndef MST_RUNTIME_H
#define MST_RUNTIME_H

#include <cstdint>
#include <string>

#include "mestrado.h"

namespace mstrt {

// Settings for the context (call before the first AlignmentScore/-Match).
void configure(const int8_t* subst576, int device, int form);
// Loads a matrix file in the blosum64.txt format (24 rows of 24 comma-separated
// integers, letter order ARNDCQEGHILKMFPSTWYVBZX*; MatrixFile.cpp). Returns
// false on a missing file, a count other than 576 or a value outside int8.
bool loadMatrix(const std::string& path, int8_t out[576]);
// The context for gap penalties (go, ge); exits with a message on failure,
// like the reference's gpuErrchk (AlignmentScore.cu:3-8).
mst_ctx* context(int gapOpen, int gapExt);
int form();
// Checks a C ABI return code; prints the context's message and exits.
void check(int rc, const char* what);
void shutdown();

}  // namespace mstrt

#endif

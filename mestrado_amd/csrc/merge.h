// Batched profile merges (R4): many merges of the progressive alignment in
// one set of launches, whatever guide tree or job they belong to.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nw_align.h"

namespace mst {

// One merge as the host describes it: group X (kX rows of nX columns) and
// group Y (kY x nY), rows in forward column order, codes 0..23 and 24 = gap.
// `out` has room for (kX + kY) x (nX + nY) bytes; the merged rows are written
// there densely, (kX + kY) x L with L the path length.
struct MergeSpec {
  const uint8_t* rx;
  int kX, nX;
  const uint8_t* ry;
  int kY, nY;
  uint8_t* out;
};

// Device descriptor of a merge: the spec plus byte offsets of its working
// set inside the engine's scratch block.
struct MergeDesc {
  const uint8_t* rx;
  const uint8_t* ry;
  uint8_t* out;
  int kX, nX, kY, nY;
  int job;                 // index of its flag job, -1 when nX or nY is 0
  int tchunk;              // steps per wave of the score table kernel
  int64_t cx, cy, p, fx, tab, ops, sxy;
  int64_t row0;            // first output row of the merge over the batch
};

// Leaf rows: forward codes of a sequence whose reversed codes start at src.
struct LeafDesc {
  int64_t src;
  uint8_t* dst;
  int len, pad;
};

// Scratch bytes of one merge (profile counts, column scores, score table,
// flags, path, column maps).
int64_t merge_scratch_bytes(int nX, int nY);

// Runs batches of merges on a stream. One batch at a time: the pinned
// descriptors of a batch are rewritten by the next enqueue, so the caller
// waits for the stream (or an event recorded after enqueue) first.
class MergeEngine {
 public:
  MergeEngine() = default;
  MergeEngine(const MergeEngine&) = delete;
  MergeEngine& operator=(const MergeEngine&) = delete;
  ~MergeEngine();
  // Enqueues merges [0, count); the path length of merge i is in lens()[i]
  // once the stream has passed the enqueued work.
  // f16_smax >= 0: merges whose DP fits 16 bits with that largest
  // substitution score (and of at most two 512-row strips) take the packed
  // flag kernel; < 0: the int32 kernels take every merge.
  hipError_t enqueue(const MergeSpec* m, int count, const int8_t* d_S, int go, int ge, int num_cus,
                     hipStream_t st, int f16_smax = -1);
  // Enqueues the leaf rows of `count` sequences (codes: a reversed code block).
  hipError_t leaves(const int8_t* codes, const LeafDesc* l, int count, hipStream_t st);
  const int32_t* lens() const { return h_lens_; }

 private:
  hipError_t grow(void** p, int64_t* cap, int64_t bytes, bool pinned);
  uint8_t* d_scratch_ = nullptr;
  int64_t scratch_cap_ = 0;
  MergeDesc* d_desc_ = nullptr;
  MergeDesc* h_desc_ = nullptr;
  int64_t desc_cap_ = 0, hdesc_cap_ = 0;
  AlignJob* d_jobs_ = nullptr;
  AlignJob* h_jobs_ = nullptr;
  int64_t jobs_cap_ = 0, hjobs_cap_ = 0;
  int32_t* d_res_ = nullptr;  // flag job scores, then traceback lengths
  int64_t res_cap_ = 0;
  int32_t* d_lens_ = nullptr;
  int32_t* h_lens_ = nullptr;
  int64_t lens_cap_ = 0, hlens_cap_ = 0;
  int2* d_bufs_ = nullptr;
  int64_t bufs_cap_ = 0;
  unsigned int* d_counter_ = nullptr;
  int64_t counter_cap_ = 0;
  LeafDesc* d_leaf_ = nullptr;
  LeafDesc* h_leaf_ = nullptr;
  int64_t leaf_cap_ = 0, hleaf_cap_ = 0;
  TileJob* d_tiles_ = nullptr;  // tiled flag kernel of the longest merges
  TileJob* h_tiles_ = nullptr;
  int64_t tiles_cap_ = 0, htiles_cap_ = 0;
  int2* d_trows_ = nullptr;
  int64_t trows_cap_ = 0;
  int* d_tstate_ = nullptr;
  int64_t tstate_cap_ = 0;
  F16Item* d_items_ = nullptr;  // packed flag kernel items
  F16Item* h_items_ = nullptr;
  int64_t items_cap_ = 0, hitems_cap_ = 0;
};

}  // namespace mst

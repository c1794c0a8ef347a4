#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mst {

// Column sequences of one work item at most: a stream's column boundaries
// live in one VGPR of the sweeping wave, a lane each (item-size tunables are
// capped at this).
constexpr int kStreamSeqs = 64;

// Work item of the streamed kernels: a row pair (ya, yb) against the column
// sequences at positions xbegin .. xbegin+xcount-1.
struct Item3 {
  int ya, yb;
  int xbegin, xcount;
};

// Streamed packed forms (nw_score5.hip, nw_score6.hip). Rows and columns are
// positions in a length-ordered copy of the codes; a work item (Item3) is a
// row pair (ya, yb: positions, yb < 0 for ya alone) and a run of consecutive
// positions xbegin .. xbegin+xcount-1 whose codes are contiguous, swept as one
// column stream.
struct Score5Args {
  const int8_t* scodes;    // codes in length order, concatenated
  const int64_t* soff;     // offsets into scodes by position (positions + 1)
  const int32_t* order;    // position -> index of the sequence in its set
  const Item3* items;
  const int* item_job;     // set of each item (null: one set, ld = nseq)
  const int64_t* job_out;  // offset of each set's matrix in out_mat
  const int* job_ld;       // and its leading dimension
  int64_t first, last;     // item range of this launch
  unsigned int* counter;   // work queue head (zero before the launch)
  int32_t* out_mat;        // scores, written at (x, y) and (y, x)
  int nseq;
  const int8_t* subst;
  int go, ge;
  int2* carry;             // score6: per-workgroup row buffer between sweeps
  int64_t carry_cap;       // its columns per workgroup (score8: per wave)
  int2* frames = nullptr;  // score8, re-based items: the row buffer's frames per wave
  int64_t frame_cap = 0;   // blocks of 8 columns per wave
  // stacked launches: the second row pair (positions) of each item, on lanes
  // 32..63 of the strip (the first on lanes 0..31); null otherwise
  const int2* pair2 = nullptr;
  // scodes as byte offsets into a profile of 8 dwords per lane (code x
  // kProfStride8), same indexing, with kOffs8Pad zero entries before and
  // after; the kernels of such profiles take a step's profile address with
  // one add (null: from scodes)
  const uint32_t* soffs8 = nullptr;
};
// residue stride of a profile of 8 dwords per lane: 64 lanes x 32 bytes
constexpr uint32_t kProfStride8 = 64 * 8 * 4;
// zero entries around soffs8: a block's columns lie within 64 + 8 of its
// stream (steps 0 .. ncol + 70, lanes 0..63)
constexpr int kOffs8Pad = 128;

// score5: rows up to 1024 in W strips of 64*K rows, one wave per strip.
// K in 3..5 for W = 1, K in 3..4 for W = 2..4.
bool score5_supported(int K, int W);
int score5_block_threads(int W);
int score5_groups_per_cu(int K, int W);
void launch_score5(int K, int W, const Score5Args& a, int grid, hipStream_t st, bool stacked = false);
// stacked one-wave launches exist for K = 1..8 (two row pairs of up to 32*K rows)
bool score5_stacked_supported(int K);

// score6: any number of strips of 64*K rows (K = 3 or 4), four waves sweeping
// the stream once per four strips; the last strip of a sweep hands its row to
// the first of the next through `carry`. kScore6K: the served form's K.
constexpr int kScore6K = 4;
int score6_block_threads();
int score6_groups_per_cu();
void launch_score6(int K, const Score5Args& a, int grid, hipStream_t st);
// score7: the multi-sweep kernel as two pipelines per workgroup sharing the
// strips' profiles, in step per sweep; row buffers: two per workgroup.
int score7_block_threads();
int score7_groups_per_cu();
void launch_score7(int K, const Score5Args& a, int grid, hipStream_t st);

// score8: a workgroup of four waves on one row pair, wave p sweeping the p-th
// part of the item's columns through every strip of 64*K rows in turn (the
// strips' profile shared, built once per strip), its strip's last row kept
// in its own row buffer (`carry`, carry_cap columns per wave). Re-based items
// (REBASE) hold one column sequence per wave and reach any length: the
// 16-bit values are kept in a moving frame (nw_stream.h, Rebase), with the
// row buffer's frames in `frames`. K = 4..8.
constexpr int kScore8MinK = 4, kScore8MaxK = 8;
// a re-based wave moves its frame when lane 0's first row leaves
// [-kRebaseAt, kRebaseAt]
constexpr int kRebaseAt = 16384;
constexpr int kScore8Waves = 4;
int score8_block_threads();
int score8_groups_per_cu();
void launch_score8(int K, bool rebase, const Score5Args& a, int grid, hipStream_t st);

}  // namespace mst

#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace mst {

enum { SRC_SEQ = 0, SRC_TAB = 1 };
enum { OP_DIAG = 0, OP_LEFT = 1, OP_UP = 2 };  // LEFT: X residue vs gap; UP: gap vs Y
// Flag word of a lane and step (8 rows): FMT_NIBBLE, row i's nibble (LEFT, UP,
// contUp, contLeft from bit 3 down) at bits 4i..4i+3; FMT_PLANES, bit planes
// of the four sign flags of the packed kernel (nw_flag16.hip).
enum { FMT_NIBBLE = 0, FMT_PLANES = 1 };

struct AlignJob {
  int n, m;            // columns (X) and rows (Y)
  int fmt;             // flag format: FMT_NIBBLE (int32 kernels) or FMT_PLANES (flag16_kernel)
  int k;               // rows per lane of the flag layout (strips of 64 k rows; 8 for FMT_NIBBLE)
  int64_t x_off, y_off;  // residue offsets (SRC_SEQ)
  int64_t tab_off;     // byte offset of the score table (SRC_TAB), 8-byte aligned
  int64_t flag_off;    // word offset of the flag block
  int64_t ops_off;     // byte offset of the path output (n+m bytes)
};

struct AlignArgs {
  const AlignJob* jobs;
  int64_t first, last;
  const int8_t* seqs;
  const uint8_t* tab;
  const int8_t* subst;
  uint32_t* flags;
  int32_t* scores;
  int2* bufs;
  int64_t bufcap;
  int go, ge;
  unsigned int* counter;  // pipelined form: work queue head (zero before the launch)
};

struct TraceArgs {
  const AlignJob* jobs;
  int64_t first, last;
  const uint32_t* flags;
  uint8_t* ops;
  int32_t* lens;
};

// Tiled form (an inter-block wavefront) for long jobs: tile (s, c) is strip s
// (512 rows) by columns [c*tw, (c+1)*tw). One launch per anti-diagonal
// s + c = d; `tiles` lists that launch's tiles, one wave each.
struct TileJob {
  int job;         // index into AlignArgs-style job list
  int s, c;        // strip, column tile
  int pad;
  int64_t rb;      // int2 offset of the job's two row buffers (2 x n)
  int64_t st;      // int offset of the job's lane states (strips x kTileStateInts)
};

struct TileArgs {
  const AlignJob* jobs;
  const TileJob* tiles;
  const int8_t* seqs;
  const uint8_t* tab;
  const int8_t* subst;
  uint32_t* flags;
  int32_t* scores;  // indexed by TileJob::job
  int2* rows;
  int* state;
  int go, ge, tw;
};

void launch_align(int src, const AlignArgs& a, int grid, hipStream_t st);
void launch_align_tiles(int src, const TileArgs& a, int ntiles, hipStream_t st);
// tile bookkeeping of a job: row buffers (int2) and lane states (int)
// per strip: H and E of 8 rows for 64 lanes, then lane 0's next diagonal
constexpr int kTileStateInts = 64 * 16 + 64;
inline int64_t tile_rows_elems(int n) { return 2 * (int64_t)n; }
inline int64_t tile_state_elems(int m) { return (int64_t)((m + 511) / 512) * kTileStateInts; }
constexpr int kTileWidth = 256;  // columns of a tile
// Jobs long enough for the tiled form: more strips than the 8-wave pipeline
// holds at once and enough columns for several tiles.
inline bool tiled_job(int n, int m) { return (m + 511) / 512 > 8 && n >= 8 * kTileWidth; }
// Tiles of jobs[first, last) grouped by anti-diagonal: tiles[diag_off[d] ..
// diag_off[d+1]) is launch d; *rows / *state: elements of the row buffers and
// lane states (TileJob::rb, ::st index them).
void plan_tiles(const AlignJob* jobs, int64_t first, int64_t last, int tw, std::vector<TileJob>& tiles,
                std::vector<int64_t>& diag_off, int64_t* rows, int64_t* state);
void launch_tile_diagonals(int src, TileArgs a, const TileJob* d_tiles, const std::vector<int64_t>& diag_off,
                           hipStream_t st);
// Pipelined form for jobs of more than one 512-row strip: one job per
// workgroup of W waves (W = 4, or 8 with SRC_TAB); wave j computes strips j,
// j+W, ...; a strip reads the row above from the wave of the previous strip
// through an LDS ring, or, across the wrap from wave W-1 to wave 0, through a
// row buffer of bufcap columns per workgroup in bufs.
void launch_align_pipe(int src, int W, const AlignArgs& a, int grid, hipStream_t st);
int64_t align_strips(int m);  // 512-row strips of an m-row job
void launch_traceback(const TraceArgs& a, hipStream_t st);

// Packed 16-bit flag kernel (nw_flag16.hip): an item is two jobs of the same
// rows per lane k (b < 0: job a alone) run in the two halves of one wave,
// strips of 64 k rows in step, or one job with two consecutive strips in the
// halves, the second 64 columns behind; scores from a byte table (AlignJob::tab_off,
// 8 bytes per lane and step, the first k used) in the flag layout; flags in
// FMT_PLANES. k: 1..8; items of every k in one launch.
// Only for jobs whose DP fits 16 bits (pk16_bound) and whose table bytes
// S + go + ge lie in [0, 255]. bufs: 2 * bufcap int2 per workgroup.
struct F16Item {
  int a, b;
  int k;     // rows per lane of both jobs (1..8)
  int mode;  // 0: jobs a and b strip by strip (b < 0: a alone); 1: job a two strips per pass
};
struct F16Args {
  const AlignJob* jobs;
  const F16Item* items;
  int64_t nitems;
  const uint8_t* tab;
  uint32_t* flags;
  int32_t* scores;
  int2* bufs;
  int64_t bufcap;
  int go, ge;
  // prof = 1 (sequence jobs, every S - smin in 0..15): no table; each pass
  // builds its strips' query profiles in LDS, a nibble per row (S - smin),
  // from seqs and subst, and a step reads the word of its column's residue
  int prof = 0;
  int smin = 0;
  const int8_t* seqs = nullptr;
  const int8_t* subst = nullptr;
  // work queue head (zero before the launch): waves take items in order as
  // they get a slot; null: wave w takes items w, w + grid, ...
  unsigned int* counter = nullptr;
};
void launch_flag16(const F16Args& a, int grid, hipStream_t st);
// Items of the packed jobs jobs[0, n): each job's k set, jobs of one k
// consecutive (sorted by strips, then columns, descending). Jobs of one strip
// pair up (two jobs per item), a job of two strips is one item of mode 1;
// items largest first.
void plan_flag16_items(const AlignJob* jobs, int64_t n, std::vector<F16Item>& items);
// score tables of sequence jobs: one (job, strip) pair per entry of strips
// (each job's own k)
void launch_seq_table(const AlignJob* jobs, const int2* strips, int nstrips, const int8_t* seqs,
                      const int8_t* subst, int bonus, uint8_t* tab, hipStream_t st);
// rows per lane of a packed job of m rows: the fewest strips of at most 512
// rows, then the fewest rows per lane that cover m with them
inline int flag16_k(int m) {
  const int s = (m + 511) / 512 > 0 ? (m + 511) / 512 : 1;
  const int k = (m + 64 * s - 1) / (64 * s);
  return k < 1 ? 1 : k;
}
// flag block size of an n x m job, strips of 64 k rows
int64_t flag_words(int n, int m, int k = 8);

}  // namespace mst

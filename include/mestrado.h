/*
 * mestrado — MI355X (gfx950) implementation of the all-pairs Needleman-Wunsch,
 * pairwise/profile alignment and neighbour-joining path of ptersnow/mestrado.
 *
 * Thin C ABI: plain pointers and sizes, every call returns 0 on success or a
 * negative MST_E* code; mst_last_error() gives the message of the last failure
 * on that context. Host buffers are owned by the caller; device memory is
 * owned by the context and lives until mst_destroy (or mst_seqset_free).
 * A context is single-threaded: calls on one context must not overlap. Many
 * contexts may coexist, on one or several devices.
 */
#ifndef MESTRADO_H
#define MESTRADO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MST_OK 0
#define MST_E_ARG -1      /* invalid argument */
#define MST_E_HIP -2      /* HIP runtime error */
#define MST_E_NOMEM -3    /* device or host allocation failed */
#define MST_E_STATE -4    /* object in the wrong state (unknown handle, ...) */

/* Forms of the all-pairs score computation (same integer results). */
#define MST_FORM_PLAIN 0    /* one launch per length bucket, synchronous */
#define MST_FORM_BATCHED 1  /* persistent work-queue launches, graph-captured */

typedef struct mst_ctx mst_ctx;

/* Replaces the constructors AlignmentScore(gapOpen, gapExt, wSize, maxSeqLen)
 * (AlignmentScore.cu:110-136) and AlignmentMatch(...) (AlignmentMatch.cu:82-109):
 * gap penalties and the 24x24 substitution matrix (row-major, letter order
 * ARNDCQEGHILKMFPSTWYVBZX*, S[row residue][column residue]; NULL selects the
 * built-in BLOSUM62 of definitions.cuh:96-121) become context state instead of
 * __constant__ symbols. device is the HIP device ordinal. */
int mst_create(int gap_open, int gap_ext, const int8_t* subst576, int device,
               mst_ctx** out);
int mst_destroy(mst_ctx* ctx);
const char* mst_last_error(const mst_ctx* ctx);
/* Number of SIMD-resident waves the batched form launches (informational). */
int mst_device_info(mst_ctx* ctx, int* num_cus, int64_t* hbm_bytes);
/* The device's architecture name (gcnArchName, e.g. "gfx950:sramecc+:xnack-"),
 * NUL-terminated within cap bytes; it keys the persisted tuning tables. */
int mst_device_arch(mst_ctx* ctx, char* buf, int cap);

/* Replaces AlignmentScore::loadDB/parse (AlignmentScore.cu:143-185) and
 * AlignmentMatch::prepare (AlignmentMatch.cu:116-131): n sequences as ASCII in
 * forward order, concatenated, with offsets[n+1] (int64). The device encodes
 * every residue (lower case folded, unknown letters -> X) and reverses each
 * sequence, and keeps the set resident in HBM. */
int mst_seqset_upload(mst_ctx* ctx, const char* ascii, const int64_t* offsets,
                      int n, int* set_id);
int mst_seqset_free(mst_ctx* ctx, int set_id);
/* Reversed codes back to the host (test hook for the encode step). */
int mst_seqset_codes(mst_ctx* ctx, int set_id, int8_t* codes_out, int64_t cap);

/* Replaces AlignmentScore::computeDistanceMtx (AlignmentScore.cu:187-277) and
 * the score kernel NeedlemanWunschGlobalScoreKernel (NeedlemanWunsch.cu:22-161):
 * every pair i <= j of the set (X = sequence i, Y = sequence j), int32 global
 * alignment scores written to scores_nxn[i*n+j] and [j*n+i] (may be NULL) and,
 * as float, to dist_nxn (may be NULL). */
int mst_all_pairs(mst_ctx* ctx, int set_id, int form, int32_t* scores_nxn,
                  float* dist_nxn);

/* Explicit pair list: out[k] = score(X = seq pi[k], Y = seq pj[k]). */
int mst_score_pairs(mst_ctx* ctx, int set_id, const int32_t* pi,
                    const int32_t* pj, int64_t npairs, int form, int32_t* out);

/* No counterpart in the reference: mst_score_pairs (plain form) for callers
 * that cannot allocate device memory (code under stream capture, tight
 * serving loops, harnesses that forbid hipMalloc). Pair indices, scores and
 * the kernel's row buffers live in a 64 MiB device array of the library,
 * one per device, that holds 12 B per pair plus 16 B per residue of the
 * set's longest sequence for each wave of the launch; larger requests return
 * MST_E_ARG. No device memory is allocated or freed. Every context on a
 * device shares that device's array: calls that overlap on one device race.
 * Results equal mst_score_pairs's. */
int mst_score_pairs_noalloc(mst_ctx* ctx, int set_id, const int32_t* pi,
                            const int32_t* pj, int64_t npairs, int32_t* out);

/* Benchmark hook: all pairs of a resident set with the result left on the
 * device; *kernel_ms receives the device time of the score launches (HIP
 * events on the context's stream). */
int mst_all_pairs_resident(mst_ctx* ctx, int set_id, int form, float* kernel_ms);
/* One all-pairs pass with an event pair around every launch (the batched
 * form's launches are not graph-replayed here), for per-launch timing. */
int mst_all_pairs_launches(mst_ctx* ctx, int set_id, int form);
/* Per-launch record of the last mst_all_pairs_launches call: rows per lane of
 * the kernel instance (negative: -(1000+100*W+K) for the pipelined
 * 16-bit kernel with W waves of K rows per lane, -2000-K for the multi-sweep
 * 16-bit kernel), pairs, DP cells and device
 * milliseconds of each launch (arrays of cap entries; *count receives the
 * number of launches). */
int mst_last_launch_stats(mst_ctx* ctx, int* rows_per_lane, int64_t* pairs, int64_t* cells,
                          float* ms, int cap, int* count);
/* Tunables of the batched form (no counterpart in the reference; the
 * reference's window size wSize becomes such a hint). Compiled defaults in
 * brackets; a tuned Python context applies the persisted table instead.
 *  "chunk"    column sequences per work item, 1..64 [32]
 *  "k5min"    least rows per lane of the one-wave kernel, 1..3 [1]
 *  "k1max"    most rows per lane of the one-wave kernel, 5..16: rows up to
 *             64 * k1max take one wave [13]
 *  "stack"    consecutive row pairs of the one-wave classes of 1..stack rows
 *             per lane share a wave, 32 lanes each, at ceil(longest row /
 *             32) rows per lane, 0..4 [4]
 *  "k8"       rows that take the strips-in-turn kernel (every strip of a row
 *             pair on one wave): 0 none, 1 rows above 1024, 2 also rows
 *             above 64 * k1max [2]
 *  "chunk8"   column sequences per work item of that kernel, 1..64 [4]
 *  "rebase"   pairs past the 16-bit bound: 0 the int32 kernel, 1 the
 *             re-based packed kernel when its items are many or end well
 *             within the pass, 2 always the re-based kernel [1]
 *  "k2max"    most rows per lane of the two-wave pipelines (k8 0), 4..8 [8]
 *  "wmax"     most waves, 4..8, of a one-sweep pipeline of 4 rows per lane
 *             for rows 1025..256 * wmax (k8 0) [4]
 *  "k6"       rows per lane of the multi-sweep kernel (k8 0): 3, 4, or 0 =
 *             per row pair [0]
 *  "chunk6"   column sequences per item of the multi-sweep kernel, 0 = chunk [0]
 *  "s6p2"     1: the multi-sweep kernel as two pipelines per workgroup [0]
 *  "chunk_w3" column sequences per item of 3- and 4-wave pipelines, 0 = chunk [0]
 *  "chunk_k9" column sequences per item of the classes of 9 or more rows per
 *             lane, 0 = chunk [0]
 *  "branches" parallel branches of the captured graph's class launches, 1..4 [2]
 *  "wg5"      workgroups per CU of the streamed kernel, 0 = as the LDS allows [0]
 * Setting a value drops the cached batched plans. Unknown keys or values out
 * of range return MST_E_ARG. */
int mst_tune_set(mst_ctx* ctx, const char* key, int value);
int mst_tune_get(mst_ctx* ctx, const char* key, int* value);
/* One join of the guide tree (TreeNode::join, TreeNode.cpp:30-58). */
typedef struct {
  int node;             /* id of the new node (n0, n0+1, ...) */
  int left, right;      /* joined node ids, left < right */
  float d;              /* D(left, right) */
  double Ti, Tj;        /* row sums T of left and right at the join */
  int n;                /* active taxa before the join */
  double q;             /* the minimal q */
  double len_left;      /* branch lengths as TreeNode::join computes them */
  double len_right;
} mst_join;

/* Served form (no counterpart in the reference, whose runs are one job each):
 * a request loop with continuous batching over one context. A job is a set of
 * sequences and what is wanted of it: the all-pairs score matrix, plus the NJ
 * tree of that matrix (what mst_nj_scores gives), plus the progressive
 * alignment along that tree (what mst_msa gives), i.e. wnj's whole run.
 * mst_server_step retires finished device work, admits waiting jobs in
 * arrival order while their device footprint fits the budget (scores N*N*4 B
 * and twice the residues; with a tree, the fp32 matrix N*N*4 B and the NJ
 * state; with an alignment, the leaf rows and 4*N*(maxlen+64) B of group
 * rows), and launches: rounds of score work items of the admitted jobs,
 * earliest deadline first, packed into shared launches of the batched form's
 * kernels; NJ launches for the jobs whose scores are done (a workgroup per
 * job); merge batches of the ready profile merges of all jobs in their
 * alignment stage (up to four NJ launches and eight merge batches in
 * flight). Results stay on the device until copied out; finished jobs hold
 * their memory until mst_server_result frees them. A job larger
 * than the budget is rejected at submission. Jobs are independent: a job's
 * results are identical alone and in any mix. Not thread-safe. The server
 * uses 17 streams; give the process as many hardware queues
 * (GPU_MAX_HW_QUEUES=24 in the environment before the first HIP call), or
 * streams sharing a queue serialise. */
typedef struct mst_server mst_server;
enum { MST_JOB_WAITING = 0, MST_JOB_RUNNING = 1, MST_JOB_DONE = 2, MST_JOB_CANCELLED = 3,
       MST_JOB_REJECTED = 4 };
int mst_server_create(mst_ctx* ctx, int64_t budget_bytes, mst_server** srv);
int mst_server_destroy(mst_server* srv);
/* The job's sequences are ascii[offsets[i] .. offsets[i+1]) (copied). */
int mst_server_submit(mst_server* srv, const char* ascii, const int64_t* offsets, int n,
                      int64_t* job);
/* What a job computes: MST_SERVE_SCORES (as mst_server_submit), _TREE (and
 * the NJ tree), _MSA (and the alignment along it). */
enum { MST_SERVE_SCORES = 1, MST_SERVE_TREE = 3, MST_SERVE_MSA = 7 };
int mst_server_submit_job(mst_server* srv, const char* ascii, const int64_t* offsets, int n,
                          int what, int64_t* job);
/* One scheduling pass (never blocks on the device); *ncompleted (may be NULL)
 * receives the number of jobs that finished in it. */
int mst_server_step(mst_server* srv, int* ncompleted);
/* Steps until every admitted and waiting job is done (blocks). */
int mst_server_drain(mst_server* srv);
/* State, size and host steady-clock times (seconds) of a job. */
int mst_server_job(mst_server* srv, int64_t job, int* state, int* n, double* t_submit,
                   double* t_admit, double* t_done);
/* Stage times of a job (host steady clock, seconds; 0 for a stage not
 * reached): times[0] submission, [1] admission, [2] scores done, [3] tree
 * done, [4] done. */
int mst_server_job_times(mst_server* srv, int64_t job, double* times);
/* Stage accounting of a job (no counterpart in the reference, which serves
 * nothing): stats[0] merge batches that held its merges, [1] seconds of its
 * alignment stage with at least one of them in flight, [2] seconds with none
 * in flight (its next merges waiting for their children, a free merge slot,
 * memory or the host loop), [3] score rounds that held its items, [4]
 * seconds from submission to the first of them, [5] from submission to
 * admission. */
int mst_server_job_stats(mst_server* srv, int64_t job, double* stats);
/* Copies a done job's n x n scores out (may be NULL) and forgets the job:
 * take its tree and alignment first. */
int mst_server_result(mst_server* srv, int64_t job, int32_t* scores_nxn);
/* The n-1 joins of a done job with a tree (as mst_nj_scores fills them). */
int mst_server_tree(mst_server* srv, int64_t job, mst_join* joins);
/* A done job's alignment: *msa_len columns; rows (may be NULL) receives its
 * n gapped rows in input order, row r at rows + r*stride, NUL-terminated
 * (stride >= *msa_len + 1), as mst_msa_rows. */
int mst_server_msa(mst_server* srv, int64_t job, char* rows, int64_t stride, int* msa_len);
int mst_server_cancel(mst_server* srv, int64_t job);
/* Up to cap ids of jobs that finished since the previous call, oldest first;
 * *count receives how many were written. */
int mst_server_completed(mst_server* srv, int64_t* ids, int cap, int* count);
int mst_server_stats(mst_server* srv, int* waiting, int* running, int* inflight,
                     int64_t* arena_used, double* cells_per_s);
/* "round_cells" (DP cells per round, default 2e10, about 4 ms of the device),
 * "chunk" (column sequences per item, default 12), "chunk6" (per item of the
 * multi-sweep kernel, rows above 1024; 0 = "chunk"; default 8), "s6p2" (1:
 * the multi-sweep kernel as two pipelines per workgroup, the default; 0: one
 * pipeline), "slack" (a
 * job's deadline
 * is its submission plus slack times its estimated run time, default 2;
 * jobs go earliest deadline first), "est_floor" (seconds: the least
 * estimated run time in that deadline, so jobs below it share one deadline
 * horizon and go in arrival order; default 0), "deadline_add" (seconds added
 * to every deadline, default 0), "rate" (initial throughput estimate in
 * cells/s, for those estimates), "round_cells_shared" and "score_share"
 * (while jobs are in their NJ or alignment stage: cells per round, default
 * 2e10, and the share of the score kernels' grid, default 1; lower values let
 * those stages' launches find free compute units sooner at some cost in
 * throughput), "nj_ratio" (NJ rebuild ratio,
 * default 0.5),
 * "merge_bytes" (scratch of one merge batch, default 512 MiB), "small_cells"
 * (merges of at most this many cells batch apart from larger ones, default
 * 4e6), "rounds" (score rounds kept in flight, 1..3, default 3),
 * "persistent" (1: score launches of as many workgroups as fit, each taking
 * items until none are left; 0: a workgroup per item; default 1), "k8" (1:
 * row pairs above 64 k1max rows on the
 * strips-in-turn kernel, items of at most 16384 columns; 0: two-wave and
 * multi-sweep kernels; default 1), "shared_backlog" (cells: when positive,
 * "round_cells_shared" and "score_share" apply only while the score work
 * not yet launched is below it; default 0: whenever jobs are in their NJ or
 * alignment stage). */
int mst_server_set(mst_server* srv, const char* key, double value);
/* Measurement helper (no counterpart in the reference): integer VALU rate of
 * the context's device in T ops/s, int32 (packed = 0) or packed int16
 * (packed = 1, two ops per lane and instruction), from a kernel of
 * independent add/max chains. */
int mst_valu_peak(mst_ctx* ctx, int packed, double* tops);
/* Measurement helper (no counterpart in the reference): a plain device copy
 * of `bytes` (copy_kernel, 16 B per thread), best of `reps` runs; *gbps =
 * bytes read + written per second. Its profiler byte counts calibrate those
 * of the other kernels. */
int mst_copy_probe(mst_ctx* ctx, int64_t bytes, int reps, double* gbps);


/* Replaces AlignmentMatch::computePairwise (AlignmentMatch.cu:140-317) with
 * the kernels NeedlemanWunschGlobalMatchKernel, NeedlemanWunschGlobalBackKernel
 * and reorderMatchesMemory (NeedlemanWunsch.cu:184-563): global alignment of
 * X = sequence pi[k] (columns) against Y = sequence pj[k] (rows) with the
 * reference's traceback and tie rules. scores (may be NULL) gets the int32
 * scores; gapped_x / gapped_y (may be NULL) get NUL-terminated gapped strings,
 * pair k at offset k*stride (stride >= len(X)+len(Y)+1), residues printed as
 * the letter of their code (revConvConst, definitions.cuh:134-137). */
int mst_align_pairs(mst_ctx* ctx, int set_id, const int32_t* pi, const int32_t* pj,
                    int64_t npairs, int32_t* scores, char* gapped_x, char* gapped_y,
                    int64_t stride);

/* Replaces SpeciesBlock::buildTree (SpeciesBlock.cpp:844-1343) on one device:
 * neighbour joining of the n x n matrix D (row-major fp32, symmetric, the
 * diagonal holds the self scores the reference keeps there). joins receives
 * the n-1 joins in order; rebuild_ratio is CmdLineOptions -r (default 0.5). */
int mst_nj(mst_ctx* ctx, const float* D_nxn, int n, float rebuild_ratio, mst_join* joins);

/* NJ on the score matrix of the set's last mst_all_pairs call, left resident
 * on the device (d = (float)score, SpeciesBlock.cpp:264 via AlignmentScore). */
int mst_nj_scores(mst_ctx* ctx, int set_id, float rebuild_ratio, mst_join* joins);

/* Progressive alignment along the guide tree: what AlignmentMatch::
 * computeAlignment (AlignmentMatch.cu:133-434) leaves unfinished, made whole.
 * Every join aligns the profile of its left node (columns) against the profile
 * of its right node (rows) with the recurrence, flags and traceback of
 * mst_align_pairs (profile column scores: rint of the fp32 frequency-weighted
 * substitution score; exactly the substitution score for single sequences),
 * then inserts the path's gap columns into every member. *msa_len receives the
 * number of columns; the rows stay on the context for mst_msa_rows. */
int mst_msa(mst_ctx* ctx, int set_id, const mst_join* joins, int* msa_len);
/* Rows of the last mst_msa in input order, letters and '-', NUL-terminated,
 * row r at offset r*stride (stride >= msa_len+1). */
int mst_msa_rows(mst_ctx* ctx, char* rows, int64_t stride);

#ifdef __cplusplus
}
#endif

#endif

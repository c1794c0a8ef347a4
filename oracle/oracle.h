/*
 * ORACLE — test infrastructure only.
 *
 * A CPU restatement of the reference recurrences, used as the parity checker
 * by tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py.
 * Nothing in the product (mestrado_amd/, include/) links, loads or calls it.
 *
 * Parity status of each function (what pins it):
 *   orc_encode_reverse   reference table AlignmentScore.cu:10-27, reversal :168-177;
 *                        pinned by tests against the table re-derived from the
 *                        BLOSUM letter order ARNDCQEGHILKMFPSTWYVBZX* and by
 *                        round trip through orc_decode.
 *   orc_nw_score         NeedlemanWunsch.cu:22-161 (R1); pinned by an independent
 *                        three-matrix Gotoh (orc_gotoh_score), brute-force
 *                        enumeration of every alignment (lengths <= 6), closed
 *                        forms (1x1, identical sequences, symmetry).
 *   orc_nw_flags         NeedlemanWunsch.cu:184-365 (R2); pinned through the
 *                        traceback: the rescored path equals the DP score, and
 *                        hand-built tie cases.
 *   orc_nw_traceback     NeedlemanWunsch.cu:393-494 (R3) with the row -1 fix;
 *                        pinned by brute force (path is optimal), path invariants
 *                        (gaps removed give the inputs) and hand-built ties.
 *   orc_profile_*        build-defined profile-profile DP; pinned by reduction to
 *                        orc_nw_* on one-hot profiles and brute force on tiny
 *                        profiles.
 *   orc_progressive_msa  the merges of tests/oracle_bind.progressive_msa (profile
 *                        counts, orc_profile_scores, orc_nw_flags_sc, orc_nw_path,
 *                        gap insertion) in C++ for large sets; pinned by equality
 *                        with that Python restatement on small sets, and through
 *                        it by the pins of orc_profile_* and orc_nw_*.
 *   orc_nj_*             SpeciesBlock.cpp:426-453,456-674,710-842,844-1343 (R5);
 *                        pinned by additive-tree recovery, hand-worked 4-taxon
 *                        cases, a textbook O(n^3) NJ, and the reference's own
 *                        TreeNode.cpp compiled from source (oracle/_ref).
 */
#ifndef MESTRADO_ORACLE_H
#define MESTRADO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Sentinel for "minus infinity" in the int32 recurrence. */
#define ORC_NEG_INF (-(1 << 29))

/* ASCII -> reversed 0..23 codes (unknown letters fold to X=22). Returns n. */
int orc_encode_reverse(const char* ascii, int n, int8_t* out);
/* 0..23 code -> letter. */
char orc_decode(int code);

int32_t orc_nw_score(const int8_t* X, int n, const int8_t* Y, int m,
                     const int8_t* S576, int go, int ge);

/* Nibble per cell, row-major [m][n]: bit3 LEFT, bit2 UP, bit1 contUp, bit0 contLeft. */
int32_t orc_nw_flags(const int8_t* X, int n, const int8_t* Y, int m,
                     const int8_t* S576, int go, int ge, uint8_t* nib);

/* Traceback over nib. Writes NUL-terminated gapped strings (forward order of
 * the original sequences); each buffer needs n+m+1 bytes. Returns the length. */
int orc_nw_traceback(const uint8_t* nib, const int8_t* X, int n,
                     const int8_t* Y, int m, char* outX, char* outY);

/* Independent three-matrix Gotoh in int64 with true -inf (cross-check). */
int64_t orc_gotoh_score(const int8_t* X, int n, const int8_t* Y, int m,
                        const int8_t* S576, int go, int ge);

/* Rescore a gapped alignment (forward-order strings of equal length) under the
 * cost model of the recurrence. */
int64_t orc_rescore(const char* gx, const char* gy, int len,
                    const int8_t* S576, int go, int ge);

/* Max over every alignment of the (reversed, encoded) inputs (n, m <= 7). */
int64_t orc_bruteforce(const int8_t* X, int n, const int8_t* Y, int m,
                       const int8_t* S576, int go, int ge);

/* All-pairs scores (upper triangle incl. diagonal mirrored) on nthreads
 * OpenMP threads; seqs concatenated, offsets[n+1]. out is n*n. Returns the
 * wall seconds of the all-pairs loop. If pair_limit > 0 only the first
 * pair_limit pairs (in the i<=j row-major order) are computed. */
double orc_all_pairs(const int8_t* seqs, const int64_t* offsets, int n,
                     const int8_t* S576, int go, int ge, int32_t* out,
                     int nthreads, int64_t pair_limit);

/* Scores for an explicit list of pairs; returns wall seconds. */
double orc_pair_list(const int8_t* seqs, const int64_t* offsets,
                     const int32_t* pi, const int32_t* pj, int64_t npairs,
                     const int8_t* S576, int go, int ge, int32_t* out,
                     int nthreads);

/* ---- profile-profile DP (build-defined, reduces to R1-R3 on one-hot) ---- */
/* Column scores: counts are [len][25] (24 letters + gap) for a group of k
 * sequences. Fills sc[m][n] (rows Y, columns X). */
void orc_profile_scores(const int32_t* cntX, int n, int kX,
                        const int32_t* cntY, int m, int kY,
                        const int8_t* S576, int32_t* sc);
int32_t orc_nw_flags_sc(const int32_t* sc, int n, int m, int go, int ge,
                        uint8_t* nib);
/* Traceback producing an edit path: ops[k] = 'D' (both), 'L' (X column vs gap
 * column), 'U' (gap vs Y column), forward order; returns the length. */
int orc_nw_path(const uint8_t* nib, int n, int m, char* ops);

/* Progressive alignment of n sequences (ASCII, offsets[n+1]) along njoins
 * joins given as (node, left, right) triples: each join aligns the left
 * group (columns) against the right group (rows) and inserts the path's gap
 * columns into every member. rows_out (may be NULL) receives row i at
 * i*stride, NUL-terminated; *msa_len the number of columns. Returns 0, -1
 * for joins that do not form one tree, or -2 (rows not written, *msa_len
 * set) for a stride below *msa_len + 1. */
int orc_progressive_msa(const char* ascii, const int64_t* offsets, int n,
                        const int32_t* joins3, int njoins, const int8_t* S576,
                        int go, int ge, char* rows_out, int64_t stride,
                        int* msa_len, int nthreads);

/* ---- neighbour joining (R5/R6) ---- */
typedef struct {
  int node;        /* new node id */
  int left, right; /* joined ids, left < right */
  float d;         /* D(left,right) */
  double Ti, Tj;   /* row sums at join time */
  int n;           /* active taxa before the join */
  double q;        /* winning q */
  double lenL, lenR;
} orc_join_t;

/* Faithful single-process restatement of SpeciesBlock::buildTree.
 * D is n*n float (row-major, symmetric, diagonal = self scores).
 * joins has n-1 entries. Returns the number of T rebuilds; *pq_trigger is set
 * to the number of times the PQ-ratio trigger fired. */
int orc_nj_reference(const float* D, int n, int numBins, float ratio,
                     float pqRatio, int qda, int predictQ,
                     orc_join_t* joins, int* pq_trigger);

/* Textbook NJ: full Q scan each step with the same q formula, T update and
 * rebuild schedule, tie rule larger max(i,j) then larger min(i,j). */
int orc_nj_fullscan(const float* D, int n, float ratio, orc_join_t* joins);

/* Tree text as TreeNode::buildTreeOutput + ";\n". names[i] NUL-terminated.
 * Writes into out (cap bytes); returns the length needed. */
int orc_tree_text(const orc_join_t* joins, int n, const char* const* names,
                  char* out, int cap);

#ifdef __cplusplus
}
#endif

#endif

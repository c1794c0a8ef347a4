/*
 * Seeded synthetic protein sets of the benchmark workloads (no counterpart in
 * the reference, which reads only FASTA files). See mestrado_amd/csrc/synth.cpp
 * for the generating processes of configurations 0..4.
 */
#ifndef MESTRADO_SYNTH_H
#define MESTRADO_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Two-call pattern: with ascii/offsets NULL only *n_out and *total_out are
 * set. n > 0 overrides the configuration's sequence count (config 4: the job
 * size). Returns 0, -1 (bad config) or -2 (buffers too small). */
int mst_synth(int config, uint64_t seed, int n, char* ascii, int64_t cap, int64_t* offsets,
              int64_t offsets_cap, int* n_out, int64_t* total_out);

/* Job sizes of the served stream: power law with the given exponent on
 * [lo, hi]. */
int mst_synth_job_sizes(uint64_t seed, int njobs, int lo, int hi, double exponent, int* sizes);

#ifdef __cplusplus
}
#endif

#endif

/* Seeded synthetic read simulator (C-ABI).
 *
 * Not part of the reference: the reference ships no simulator and no test data.
 * This generator produces the correction workloads used by the parity tests and
 * the benchmark (genomes with duplicated segments / repeat families, SE or PE reads,
 * Phred+33 quality profiles and substitution errors), deterministic in the seed.
 */
#ifndef CARE_SIM_H
#define CARE_SIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct care_sim_config {
    uint64_t seed;
    uint64_t genome_length;
    double gc_content;          /* probability that a base is G or C */
    /* duplicated segments: round(dup_fraction*G/dup_segment_length) source segments,
       each written dup_copies more times with per-base divergence dup_divergence */
    double dup_fraction;
    int dup_segment_length;
    int dup_copies;
    double dup_divergence;
    /* repeat families covering about repeat_fraction of the genome */
    double repeat_fraction;
    int repeat_unit_length;
    int repeat_copies;
    double repeat_divergence;
    /* reads */
    int read_length_min;
    int read_length_max;        /* lengths uniform in [min, max] */
    double coverage;
    uint64_t num_reads_override; /* 0 = derive from coverage */
    int paired;                 /* 1: reads 2i and 2i+1 are mates */
    double insert_mean;
    double insert_sd;
    /* qualities: mean falls linearly from q_first (first base) to q_last (last base) */
    double q_first;
    double q_last;
    double q_sd;
    int q_min;
    int q_max;
    double n_rate;              /* probability of replacing a base by 'N' */
    double tandem_fraction;     /* genome fraction overwritten by tandem repeats */
    int tandem_period_max;      /* repeat unit length 1..tandem_period_max */
    /* 0: every read is drawn from the one generator stream, in read order.
       1: the genome comes from that stream; read i (or pair i) then draws from its own
          stream seeded by (seed, i), so reads are simulated on all host threads and the
          output does not depend on the thread count. */
    int per_read_streams;
} care_sim_config;

typedef struct care_sim_result care_sim_result;

void care_sim_default_config(care_sim_config* cfg);
int care_sim_run(const care_sim_config* cfg, care_sim_result** out);
uint64_t care_sim_num_reads(const care_sim_result* r);
int care_sim_max_length(const care_sim_result* r);
uint64_t care_sim_genome_length(const care_sim_result* r);
int care_sim_copy(const care_sim_result* r, uint64_t first, uint64_t n, size_t pitch,
                  char* seqs, char* quals, char* truth, int* lengths,
                  uint64_t* positions, int* strands);
void care_sim_free(care_sim_result* r);
/* 4-line FASTQ text of reads first, first+step, ... (count reads) from row-major
   buffers of `pitch` bytes per read; read i is named "@r<i>". With out == NULL the
   size is returned; otherwise the bytes written (0 when cap is too small). */
size_t care_sim_fastq(const char* seqs, const char* quals, const int* lengths, size_t pitch,
                      uint64_t first, uint64_t step, uint64_t count, char* out, size_t cap);

#ifdef __cplusplus
}
#endif

#endif

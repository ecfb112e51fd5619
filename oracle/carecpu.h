/* CPU restatement of the CARE per-anchor correction path (test infrastructure).
 *
 * This library is the oracle: only tests, the smoke check and the benchmark's CPU
 * baseline leg may load it, and only as the checker. It restates, in our own code,
 * the algorithm of the reference files cited in carecpu.cpp, and is itself checked
 * against the reference engine compiled from source (oracle/_ref) and against the
 * closed-form / brute-force checks in tests/.
 */
#ifndef CARECPU_H
#define CARECPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct carecpu_opts {
    int kmerlength;               /* k */
    int num_hash_maps;            /* 1..48 */
    float estimated_coverage;     /* --coverage */
    int use_quality_scores;
    int exclude_ambiguous;
    int paired;
    int min_overlap;              /* 30 */
    float max_error_rate;         /* 0.2 */
    float min_overlap_ratio;      /* 0.3 */
    float estimated_errorrate;    /* 0.06 */
    float m_coverage;             /* 0.6 */
    float paired_filter_threshold;/* 0.06 */
} carecpu_opts;

typedef struct carecpu_ctx carecpu_ctx;

/* Per-anchor intermediate values, written into caller buffers of the given capacities.
 * Any pointer may be NULL. Counts are always written. */
typedef struct carecpu_trace {
    int cap_candidates;
    int cap_columns;
    /* candidate lookup (after sort/unique/remove-self) */
    int num_candidates;
    uint32_t* candidate_ids;
    /* alignments of every looked-up candidate: {score, overlap, shift, nOps, isValid} */
    int* fwd_alignments;
    int* rc_alignments;
    int8_t* orientation;          /* 1 fwd, 2 revcomp, 3 none */
    uint8_t* is_paired;           /* PE pair flags, per looked-up candidate */
    /* after flag + mismatch-ratio filters */
    int num_filtered;
    /* after refinement: the candidates in the final MSA */
    int num_final;
    uint32_t* final_ids;
    int* final_shifts;
    float* final_weights;
    int refinement_rounds;        /* rounds that removed at least one candidate */
    /* final MSA columns */
    int num_columns;
    int anchor_column_begin;
    int* counts;                  /* [cols*4] A,C,G,T */
    float* weights;               /* [cols*4] */
    int* coverage;
    char* consensus;
    float* support;
    float* orig_weights;          /* anchor columns only are meaningful */
    int* orig_coverage;
    /* decision */
    float avg_support;
    float min_support;
    int min_coverage;
    int max_coverage;
    int is_hq;
    int is_corrected;
    char* corrected;              /* [anchor length] */
} carecpu_trace;

carecpu_ctx* carecpu_create(const carecpu_opts* opts, uint64_t num_reads,
                            const uint32_t* enc2bit, size_t pitch_ints, const int* lengths,
                            const char* quals, size_t qual_pitch, const uint8_t* ambiguous,
                            int threads);
void carecpu_destroy(carecpu_ctx* ctx);

int carecpu_results_per_map_threshold(float coverage);

/* minhash signature of one encoded read (maps 0..nmaps-1) */
void carecpu_signature(const uint32_t* enc, int len, int k, int nmaps, uint64_t* out);
void carecpu_signatures(carecpu_ctx* ctx, uint64_t first, uint64_t n, uint64_t* out);

/* per map: (key, first value offset, count) of the bucket for key, or count 0 */
int carecpu_bucket(carecpu_ctx* ctx, int map, uint64_t key, const uint32_t** values);

int carecpu_candidates(carecpu_ctx* ctx, uint32_t read_id, uint32_t* out, int cap);
/* Paired-candidate flags of two mates' sorted, unique candidate lists
   (include/util.hpp:1153-1200 over :1016-1141). */
void carecpu_flag_pairs(const uint32_t* c1, int n1, const uint32_t* c2, int n2, uint8_t* f1, uint8_t* f2);

void carecpu_align(const uint32_t* anchor, int la, const uint32_t* cand, int lc,
                   int min_overlap, float max_error_rate, float min_overlap_ratio, int* out5);
int carecpu_orientation(const int* fwd5, const int* rc5, int la, int lc,
                        float min_overlap_ratio, int min_overlap, float estimated_errorrate);

int carecpu_trace_anchor(carecpu_ctx* ctx, uint32_t read_id, carecpu_trace* trace);

/* End-to-end correction of anchors [first, first+n); returns the number of records.
 * Records are serialized exactly as the reference's EncodedTempCorrectedSequence
 * (readId u32, flags u32, payload), ordered by read id. */
size_t carecpu_correct(carecpu_ctx* ctx, uint64_t first, uint64_t n, int threads);

/* Candidate correction (--candidateCorrection, include/options.hpp:41,52): records of
 * corrected candidates of HQ anchors, in a one-thread run's order. */
/* Seeded alignment (an addition for the fast learned form; not in the reference):
 * every candidate aligned only at the strand and shift that line up the k-mer of the
 * first minhash map it shares with the anchor. */
void carecpu_set_seeded_alignment(carecpu_ctx* ctx, int on);
/* Seeds of reads [first, first+n), [read][map]: start of the first minimizing k-mer * 2,
 * plus 1 when its canonical form is the reverse complement. */
void carecpu_seeds(carecpu_ctx* ctx, uint64_t first, uint64_t n, uint16_t* out);
/* out = {map (-1: none), flag (1 fwd, 2 revcomp, 3 none), shift, overlap, mismatches}. */
void carecpu_seeded_alignment(carecpu_ctx* ctx, uint32_t a, uint32_t c, int* out);
void carecpu_set_candidate_correction(carecpu_ctx* ctx, int on, int new_columns);
size_t carecpu_records_bytes(carecpu_ctx* ctx);
const uint8_t* carecpu_records_data(carecpu_ctx* ctx);
double carecpu_last_seconds(carecpu_ctx* ctx);

/* Learned seam (correctAnchorClf, include/corrector.hpp:1448-1472, with an MLP
   16->64->64->1 as the classifier and the classic column rule below confidence 0.9).
   Parameter layout (float): W1[64][16], b1[64], W2[64][64], b2[64], W3[64], b3.
   params == NULL restores the classic decision. Returns 0, or -1 for a wrong count. */
int carecpu_set_classifier(carecpu_ctx* ctx, const float* params, int count);
int carecpu_mlp_num_params(void);
/* Decision columns (non-HQ anchors, original base != consensus) of anchors
   [first, first+n), in anchor order: 16 features and 4 meta words each (read id,
   position, orig | cons<<2 | classic<<4 | allGood<<6, 0). Returns the total count; at
   most `cap` are written. */
size_t carecpu_features(carecpu_ctx* ctx, uint64_t first, uint64_t n, int threads, float* feats, uint32_t* meta,
                        size_t cap);
/* Logits as the device computes them (bfloat16 weights/activations, float sums). */
void carecpu_mlp_logits_bf16(const float* params, const float* X, uint64_t n, float* z);
/* Mean BCE-with-logits loss over n rows in double precision; G (may be NULL) receives
   the gradient, z (may be NULL) the logits. */
double carecpu_mlp_loss_grad_f64(const double* params, const float* X, const float* Y, uint64_t n, double* G,
                                 double* z);

#ifdef __cplusplus
}
#endif

#endif

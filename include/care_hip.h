/* MI355X (gfx950) CARE error-correction engine: C-ABI.
 *
 * Drop-in boundary for the CARE correction stage behind RabbitQCPlus's
 * --correctWithCare. The reference runs this stage as
 *
 *   SerializedObjectStorage care::cpu::correct_cpu(const ProgramOptions&,
 *                                                  CpuMinhasher&, CpuReadStorage&)
 *       (reference include/correct_cpu.hpp:17-21, called from
 *        src/dispatch_care_correct_cpu.cpp:363)
 *
 * over a read storage (include/cpureadstorage.hpp:11-84) and a minhasher built from
 * it (src/cpuminhasherconstruction.cpp:168-212). This library takes the same read
 * storage as plain arrays, builds device-resident minhash tables, and produces the
 * same serialized correction records (include/correctedsequence.hpp:151-387),
 * ordered by read id, with the kernels running on one GPU.
 *
 * Conventions: every function returns CARE_OK (0) or a CARE_E* code; the message of
 * the last failure on a context is available from care_hip_last_error(). No C++
 * exception crosses this boundary. Calls on one context must not overlap; each
 * context owns one device, one stream and all device memory it allocates.
 */
#ifndef CARE_HIP_H
#define CARE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CARE_OK = 0,
    CARE_EINVAL = 1,      /* bad argument / unsupported configuration */
    CARE_ESTATE = 2,      /* call out of order (e.g. correct before build_tables) */
    CARE_EHIP = 3,        /* HIP runtime error */
    CARE_ENOMEM = 4,      /* device or host allocation failed */
    CARE_ENODEV = 5       /* no usable device */
};

/* Option block; names and defaults follow care::ProgramOptions
 * (reference include/options.hpp:36-77). */
typedef struct care_opts {
    int kmerlength;                 /* k, 1..32 (reference default 20; 0 = auto: 20 if max length < 160 else 32,
                                       src/dispatch_care_correct_cpu.cpp:283-297) */
    int num_hash_maps;              /* 1..48 (--hashmaps) */
    float estimated_coverage;       /* --coverage */
    int use_quality_scores;         /* --useQualityScores */
    int exclude_ambiguous;          /* --excludeAmbiguous */
    int paired;                     /* --pairmode PE: reads 2i, 2i+1 are mates */
    int min_overlap;                /* 30 */
    float max_error_rate;           /* 0.2 */
    float min_overlap_ratio;        /* 0.3 */
    float estimated_errorrate;      /* 0.06 */
    float m_coverage;               /* 0.6 */
    float paired_filter_threshold;  /* 0.06 */
    int batch_size;                 /* staged engine: anchors per device batch, 0 = automatic */
    int keep_intermediates;         /* 1: keep per-batch stage outputs for care_hip_inspect_* (staged engine) */
    int engine;                     /* 0 = automatic (staged), 1 = staged per-stage kernels, 2 = fused persistent kernel */
    int waves_per_cu;               /* fused engine: resident waves per compute unit, 0 = default */
    int correction_type;            /* 0 = classic decision (reference default), 1 = learned: the
                                       classifier seam of correctAnchorClf (include/corrector.hpp:
                                       1448-1472) with the MLP of care_hip_set_classifier, and the
                                       classic column rule where its confidence is below 0.9 */
    int align_variant;              /* staged engine: 0 = one lane per candidate where reads have at
                                       most 352 bases (else 1), 1 = one wavefront per pair */
    int use_graphs;                 /* 1 = capture device work into HIP graphs, replayed while their
                                       arguments and buffers hold: the fused engine's pass, or each
                                       pipelined staged batch (see sync_batches) */
    int correct_candidates;         /* --candidateCorrection (staged engine, reads of at most 340 bases):
                                       HQ anchors also correct their candidates (include/corrector.hpp:
                                       1504-1512, src/msa.cpp:457-522) */
    int new_columns_to_correct;     /* --candidateCorrectionNewColumns, 15 */
    int msa_tile;                   /* staged engine: candidates staged into LDS per step of the
                                       one-wavefront MSA, a multiple of 4 in 4..64; 0 = automatic
                                       (20 with quality scores, else 64) */
    int seeded_alignment;           /* staged engine, reads of at most 352 bases; not in the reference:
                                       1 = each candidate is aligned only at the strand and shift that
                                       line up the k-mer of the first minhash map it shares with the
                                       anchor (its seed), instead of the scan of every shift in both
                                       orientations (include/corrector.hpp:925-979); the shift is kept
                                       when the reference's overlap, mismatch and orientation bounds
                                       hold. Records then differ from the reference's where the scan
                                       would choose another shift or strand; used by the fast learned
                                       form, whose bar is the error rate */
    int sync_batches;               /* staged engine: 0 = pipelined batches where the lane-per-candidate
                                       or seeded alignment runs and no candidate correction, dump or
                                       feature output is asked for (no host synchronization inside a
                                       batch: candidate arrays sized from the candidates per anchor
                                       seen, a batch with more is redone; the next batch runs while a
                                       batch's records are copied out); 1 = a host synchronization
                                       per batch to size its candidate arrays exactly */
    int pipeline_capacity;          /* pipelined staged batches: candidate pairs per anchor their arrays
                                       hold before the context has counted any, 0 = 4 x coverage + 16;
                                       later batches hold the most seen per anchor plus a quarter.
                                       Records do not depend on it */
} care_opts;

typedef struct care_ctx care_ctx;

/* Stage timings of the last care_hip_correct call, in milliseconds (device events). */
typedef struct care_timings {
    double total_ms;
    double candidates_ms;     /* lookup + sort/unique/remove-self + compaction */
    double align_ms;          /* shifted-Hamming alignment, both orientations */
    double msa_ms;            /* filters + MSA + refinement + decision + record encoding */
    double output_ms;         /* record compaction + device-to-host copy */
    double align_kernel_ms;   /* sum of k_align durations (one launch per batch) */
    double cand_kernel_ms;    /* sum of k_candidates durations */
    double msa_kernel_ms;     /* sum of k_msa durations */
    double fused_kernel_ms;   /* fused engine: duration of the persistent kernel */
    uint64_t anchors;
    uint64_t candidates;      /* anchor/candidate pairs aligned */
    uint64_t records;
    uint64_t record_bytes;
    uint64_t bucket_values;   /* values gathered from the tables */
    uint64_t batches;
    uint64_t graphs_built;    /* HIP graphs captured and instantiated by this context */
    double fastq_device_ms;   /* last care_hip_set_reads_fastq: device span after the text upload */
    double merge_device_ms;   /* last care_hip_write_fastq: device span from the record upload to the text
                                 download (both excluded) */
    double graph_ms;          /* pipelined staged batches with use_graphs: device span of the batches
                                 (per-stage fields stay zero) */
    uint64_t batches_redone;  /* pipelined staged batches redone with exact sizes (more candidates
                                 than their arrays held) */
} care_timings;

void care_hip_default_opts(care_opts* o);

/* Device name and architecture ("<name> <gcnArch>") and compute-unit count. */
int care_hip_device_info(int device, char* name, size_t cap, int* compute_units);

int care_hip_create(const care_opts* opts, int device, care_ctx** out);
void care_hip_destroy(care_ctx* ctx);
const char* care_hip_last_error(const care_ctx* ctx);

/* Read storage (reference CpuReadStorage contract, include/cpureadstorage.hpp:16-79):
 * 2-bit rows of pitch_ints u32 (A=0 C=1 G=2 T=3, 16 bases per word, first base in
 * the high bits), lengths, optional Phred+33 qualities in rows of qual_pitch bytes
 * (bytes past a read's length, up to the longest read's length, are read as written:
 * reverse-complement candidates take the row of that width reversed, as the reference
 * does with its storage pitch, src/correct_cpu.cpp:141, include/corrector.hpp:1227-1239;
 * the reference's storage holds zeros there), and one ambiguity flag per read
 * (1 = the read contained non-ACGT bases before replacement). Copied to the device before
 * the call returns; copies of 64 MiB or more run on up to 8 host threads of the library
 * through the context's 64 MiB of pinned buffers (the caller's arrays are only read). */
int care_hip_set_reads(care_ctx* ctx, const uint32_t* enc2bit, size_t pitch_ints,
                       const int32_t* lengths, const uint8_t* quals, size_t qual_pitch,
                       uint64_t num_reads, const uint8_t* ambiguous);

/* Host helper (no device work, no context): the arrays care_hip_set_reads takes, from
 * reads as ASCII rows of seq_pitch bytes. Replaces the parser-batch encoding of
 * constructChunkedReadStorageFromFiles (include/chunkedreadstorageconstruction.hpp:
 * 67-106, 260-297; 2-bit encoding include/sequencehelpers.hpp:154-187): ACGT and acgt
 * map to 0..3; every other byte inside a read becomes A, C, G, T in turn, the turn
 * restarting at every batch of 65,536 reads, and marks the read ambiguous; bases past a
 * read's length are 0. With quals_in, quality rows are copied into rows of
 * quals_out_pitch bytes, zero past each read's length. threads = 0: all host threads.
 * Returns CARE_EINVAL if a length is negative or exceeds 16 * pitch_ints, seq_pitch,
 * or (with qualities) either quality pitch. */
int care_hip_encode_reads(const uint8_t* seqs, size_t seq_pitch, const int32_t* lengths, uint64_t num_reads,
                          uint32_t* enc2bit, size_t pitch_ints, uint8_t* ambiguous,
                          const uint8_t* quals_in, size_t quals_in_pitch, uint8_t* quals_out,
                          size_t quals_out_pitch, int threads);

/* Read storage built on the device from FASTQ text (replaces
 * constructChunkedReadStorageFromFiles, include/chunkedreadstorageconstruction.hpp:29-480,
 * with its reader include/readlibraryio.hpp): one text for single-end input, two for
 * paired input (records of the two texts interleaved R1, R2, R1, ...). Records are
 * 4-line FASTQ ("@header", sequence, "+...", quality; a '\r' ending a header, sequence
 * or quality of more than one character is dropped, as the reference's reader does,
 * include/kseqpp/kseqpp.hpp:286-288; blank lines at the end are ignored; texts of up to
 * 64 GiB each that fit the device). Lower-case acgt become upper case; every other sequence character becomes
 * A, C, G, T in turn, the turn restarting at every batch of 65,536 reads, and marks the
 * read ambiguous. Qualities are kept as given (8 bits) when use_quality_scores is set.
 * Pitch = ceil(longest/16). The text stays on the device for care_hip_write_fastq. */
int care_hip_set_reads_fastq(care_ctx* ctx, const char* fastq1, size_t bytes1,
                             const char* fastq2, size_t bytes2);

/* Output construction: the corrected FASTQ text of input file `file` (0, or 1 for the
 * second file of a pair) from serialized records (any order; the concatenation of what
 * care_hip_copy_records returns, over one or more care_hip_correct calls) and the FASTQ
 * text of care_hip_set_reads_fastq, kept on the device. Replaces the reference's
 * sortSerializedResultsByReadIdAscending (include/sortserializedresults.hpp:17-78) and
 * constructOutputFileFromCorrectionResultsOutToQueue (src/correctionresultoutput.cpp:1589-1632,
 * merge :448-975, combine rule combineMultipleCorrectionResults1_rawtcs2 :1256-1550): each
 * read is written as "@" header "\n" sequence "\n+\n" quality "\n", the sequence being the
 * original text with the combined correction applied; quality_labels appends " care:q=N"
 * (outputCorrectionQualityLabels) to the headers of reads up to the last corrected one.
 * With no records the input text is returned unchanged (the reference then hands QC
 * nothing and QC reads the input itself, src/seqc.cpp:662-672). *written = size of the
 * text; dst may be null to ask for the size; otherwise cap must be at least the size.
 * Texts of 64 MiB or more are copied into dst by up to 8 library threads (the
 * context's pinned transfer buffers, as in care_hip_set_reads). */
int care_hip_write_fastq(care_ctx* ctx, const uint8_t* records, size_t record_bytes, int file,
                         int quality_labels, char* dst, size_t cap, size_t* written);

/* Preprocessed-reads files in the reference's format (--save-preprocessedreads-to /
 * --load-preprocessedreads-from; ChunkedReadStorage::saveToFile / loadFromFile,
 * include/chunkedreadstorage.hpp:158-398, lengths as include/lengthstorage.hpp:184-204):
 * save writes the resident storage (qualities when use_quality_scores is set, zero past
 * each read's length; ambiguous read ids ascending, where the reference writes its hash
 * set's order); load replaces care_hip_set_reads (qualities are required when
 * use_quality_scores is set, as 8-bit scores, and skipped otherwise). */
int care_hip_save_reads(care_ctx* ctx, const char* path);
int care_hip_load_reads(care_ctx* ctx, const char* path);

/* Shape of the resident read storage (all zero before any reads are set). */
int care_hip_read_layout(const care_ctx* ctx, uint64_t* num_reads, int* max_length, size_t* pitch_ints);

/* Copies reads [first, first+n) of the resident storage back in care_hip_set_reads'
 * layout (any pointer may be null; qualities need qual_pitch >= the longest read).
 * Reference counterparts: CpuReadStorage::gatherContiguousSequences, gatherQualities,
 * gatherSequenceLengths, areSequencesAmbiguous, include/cpureadstorage.hpp:16-20,29-41,57-61. */
int care_hip_inspect_reads(care_ctx* ctx, uint64_t first, uint64_t n, uint32_t* enc2bit, int32_t* lengths,
                           uint8_t* quals, size_t qual_pitch, uint8_t* ambiguous);

/* Minhash signatures of all reads and the device-resident tables: per map, the
 * ascending read ids of every signature value shared by 2..threshold reads,
 * threshold = max(10, int(int(coverage) * 2.5f)) (reference
 * include/minhasherlimit.hpp:10-16, include/groupbykey.hpp:99-250). */
int care_hip_build_tables(care_ctx* ctx, int* maps_built);

/* Table files in the reference's minhasher format (OrdinaryCpuMinhasher::writeToStream /
 * loadFromStream, include/ordinaryminhasher.hpp:303-339, tables as in
 * include/cpuhashtable.hpp:221-248,627-649): what --save-hashtables-to writes and
 * --load-hashtables-from reads. save writes the built tables; load replaces
 * care_hip_build_tables (after care_hip_set_reads; the file's k-mer length and bucket
 * threshold must match the engine's options; at most num_hash_maps maps are used). */
int care_hip_save_tables(care_ctx* ctx, const char* path);
int care_hip_load_tables(care_ctx* ctx, const char* path, int* maps_loaded);

/* Corrects anchors [first, first+n) (replaces cpu::correct_cpu for that range).
 * On success the serialized records are held by the context: num_records records,
 * num_bytes bytes, in the order a one-thread run of the reference inserts them into its
 * SerializedObjectStorage (src/correct_cpu.cpp:166-281): read-id order; with
 * correct_candidates, per 16-read group of read ids the anchors' records, then their
 * candidates' records. A candidate record is dropped when the candidate was corrected as
 * an HQ anchor earlier in the call (read id below the anchor's, not below `first`;
 * include/corrector.hpp:1676-1680); the reference's choice there depends on its thread
 * timing, and matches this one with one thread. For paired input first and n must be even. */
int care_hip_correct(care_ctx* ctx, uint64_t first, uint64_t n, size_t* num_records, size_t* num_bytes);

/* Copies the records of the last care_hip_correct call (cap must be >= num_bytes). */
int care_hip_copy_records(care_ctx* ctx, uint8_t* dst, size_t cap);

/* The records of the last care_hip_correct call without a copy: *data points into the
 * context's pinned host buffer and stays valid until the next care_hip_correct,
 * care_hip_extract_features or care_hip_destroy on the context. Same bytes as
 * care_hip_copy_records (the reference reads them out of its SerializedObjectStorage,
 * include/serializedobjectstorage.hpp, to sort and merge them). */
int care_hip_records_view(const care_ctx* ctx, const uint8_t** data, size_t* bytes);

int care_hip_get_timings(const care_ctx* ctx, care_timings* out);

/* ---- learned seam ---- */

/* Number of classifier parameters: MLP 16 -> 64 -> 64 -> 1. */
#define CARE_MLP_PARAMS 5313

/* Loads the classifier used by correction_type = 1 (replaces the forest file of
 * ForestClf, include/forest.hpp:130-170, and the clf_agent's decide_anchor,
 * include/classification.hpp:74-83). params: float W1[64][16], b1[64], W2[64][64],
 * b2[64], W3[64], b3 (ReLU after the two hidden layers, output = logit). On the device
 * weights are bfloat16 (round to nearest even), sums float. NULL unloads. A column is
 * set to the consensus when the logit is >= ln 9, kept when <= -ln 9, and decided by
 * the classic rule (getCorrectedAnchor, src/msa.cpp:379-453) otherwise. */
int care_hip_set_classifier(care_ctx* ctx, const float* params, size_t count);

/* Runs care_hip_correct over [first, first+n) and returns its decision columns (anchor
 * columns of non-HQ anchors whose base differs from the consensus; the columns of the
 * reference's feature print mode, include/corrector.hpp:1474-1493) ordered by (read id,
 * position): features [m][16] (the 13 of extract_anchor_v3,
 * include/classification.hpp:748-797, then the anchor base's quality weight, i / length,
 * the smaller neighbouring support) and meta [m][4] (read id, position,
 * orig | cons << 2 | classic decision << 4 | allGood << 6, 0). *total receives the
 * number of columns; at most cap are written. Needs reads of at most 340 bases. */
int care_hip_extract_features(care_ctx* ctx, uint64_t first, uint64_t n, float* features, uint32_t* meta,
                              size_t cap, size_t* total);

/* ---- classifier training (no counterpart in the reference: its forests are trained
 * offline). One handle per device; parameters, gradients, Adam moments and training data
 * are device-resident. The forward/backward Function runs as kernels on the handle's
 * stream; every call below returns when its results are on the host. ---- */

typedef struct care_mlp care_mlp;

int care_mlp_create(int device, care_mlp** out);       /* parameters and Adam moments start at 0 */
void care_mlp_destroy(care_mlp* h);
const char* care_mlp_last_error(const care_mlp* h);
int care_mlp_set_params(care_mlp* h, const float* params, size_t count);   /* CARE_MLP_PARAMS floats */
int care_mlp_get_params(care_mlp* h, float* params, size_t count);
int care_mlp_reset_optimizer(care_mlp* h);              /* Adam moments and step count to 0 */
/* Logits z[n] of rows X[n][16] (float). */
int care_mlp_forward(care_mlp* h, const float* X, size_t n, float* z);
/* care_mlp_forward without any allocation: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The device copies of X and z live in a fixed scratch
 * array of the library (4.25 MiB per GPU, held until the process ends), so n is at most
 * 65536 rows per call (CARE_EINVAL beyond). Calls that overlap on one GPU share that
 * array and race; hidden activations are not kept (no backward after this call). */
int care_mlp_forward_static(care_mlp* h, const float* X, size_t n, float* z);
/* Mean binary cross-entropy of sigmoid(z) against y[n] (0/1) and its gradient grad[CARE_MLP_PARAMS]. */
int care_mlp_loss_grad(care_mlp* h, const float* X, const float* y, size_t n, float* loss, float* grad);
/* Training rows, copied to the device. */
int care_mlp_set_data(care_mlp* h, const float* X, const float* y, size_t n);
/* Adam (beta1 0.9, beta2 0.999, eps 1e-8) over `epochs` passes of shuffled batches of
 * `batch` rows (the last batch of a pass may be smaller); epoch_loss[epochs] (may be
 * NULL) receives each pass's mean batch loss. */
int care_mlp_train(care_mlp* h, int epochs, int batch, float lr, uint64_t seed, float* epoch_loss);

/* ---- stage inspection (parity tests) ---- */

/* Signatures of reads [first, first+n), row-major [n][num_hash_maps]. */
int care_hip_inspect_signatures(care_ctx* ctx, uint64_t first, uint64_t n, uint64_t* out);

/* With keep_intermediates=1, the stage outputs of the last batch of the last
 * care_hip_correct call (anchors [batch_first, batch_first+batch_n)):
 *   candidate_offsets [batch_n+1]      prefix sums of per-anchor candidate counts
 *   candidate_ids     [total]          ascending per anchor, self and duplicates removed
 *   alignments        [total][11]      fwd {score,overlap,shift,nOps,isValid},
 *                                      rc {score,overlap,shift,nOps,isValid}, orientation
 *   paired            [total]          pair flags (paired input)
 *   anchor_info       [batch_n][12]    num_after_filters, num_final, num_columns,
 *                                      anchor_column_begin, is_hq, is_corrected,
 *                                      min_coverage, max_coverage, refinement_rounds,
 *                                      avg_support (float bits), min_support (float bits),
 *                                      record_bytes
 * Any output pointer may be NULL. */
int care_hip_inspect_batch(care_ctx* ctx, uint64_t* batch_first, uint64_t* batch_n, uint64_t* total,
                           uint32_t* candidate_offsets, uint32_t* candidate_ids, int32_t* alignments,
                           uint8_t* paired, int32_t* anchor_info);

/* With keep_intermediates=1: final MSA of anchor `index` (0-based within the last
 * batch): columns [num_columns] of counts[4], weights[4], coverage, consensus
 * (ASCII), support; cap_columns bounds the writes. Returns CARE_EINVAL if absent. */
int care_hip_inspect_msa(care_ctx* ctx, uint64_t index, int cap_columns, int* num_columns,
                         int32_t* counts, float* weights, int32_t* coverage, uint8_t* consensus,
                         float* support);

#ifdef __cplusplus
}
#endif

#endif

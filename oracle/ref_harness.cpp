// Test infrastructure only: a thin C-ABI driver around the reference CARE engine,
// compiled from the reference sources where they lie (see oracle/Makefile).
// It is the checker for oracle/carecpu.cpp and for the HIP library; nothing in
// the product links it.
//
// What it drives (all reference code, unmodified):
//   * care::OrdinaryCpuMinhasher built by constructCpuMinhasherFromCpuReadStorage
//     (src/cpuminhasherconstruction.cpp:168-212)
//   * care::cpu::correct_cpu (src/correct_cpu.cpp:47-348) for end-to-end records
//   * care::CpuErrorCorrector::process/processMulti (include/corrector.hpp:145,325)
//     in an OpenMP loop without the 1 s ProgressThread, for timing
//   * CPUSequenceHasher::hashInto (include/cpusequencehasher.hpp:88-125)
//   * cpu::shd::cpuShiftedHammingDistancePopcount2Bit (include/cpu_alignment.hpp:411)
//   * chooseBestAlignmentOrientation (include/alignmentorientation.hpp:14-92)
//   * MultipleSequenceAlignment::build (src/msa.cpp:9-54)
//
// The only code of ours here is the CpuReadStorage implementation over plain
// arrays (the interface is include/cpureadstorage.hpp:11-84) and the C-ABI glue.

#include "ref_harness.hpp"
#include <config.hpp>
#include <options.hpp>
#include <cpureadstorage.hpp>
#include <chunkedreadstorageconstruction.hpp>
#include <cpuminhasher.hpp>
#include <cpuminhasherconstruction.hpp>
#include <ordinaryminhasher.hpp>
#include <minhasherlimit.hpp>
#include <correct_cpu.hpp>
#include <corrector.hpp>
#include <cpusequencehasher.hpp>
#include <cpu_alignment.hpp>
#include <alignmentorientation.hpp>
#include <msa.hpp>
#include <correctedsequence.hpp>
#include <serializedobjectstorage.hpp>
#include <sequencehelpers.hpp>
#include <correctionresultoutput.hpp>
#include <sortserializedresults.hpp>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <fstream>
#include <vector>
#include <atomic>
#include <thread>

#include <omp.h>

namespace {

// The steps of constructCpuMinhasherFromReadStorage (src/cpuminhasherconstruction.cpp:25-164,
// minhasher created as at :176-183) with the minhasher's own insert/compact code, but with
// no thread pool: the reference's loops then run sequentially (ForLoopExecutor,
// include/threadpool.hpp:389-420). Its pool can miss a wake-up and stall
// (include/parallel/parallel_task_queue.h:306-313); the tables are the same either way.
std::unique_ptr<care::CpuMinhasher> buildMinhasherSequential(const care::ProgramOptions& po, const ArrayReadStorage& rs){
    auto mh = std::make_unique<care::OrdinaryCpuMinhasher>(rs.getNumberOfReads(),
                                                           calculateResultsPerMapThreshold(po.estimatedCoverage),
                                                           po.kmerlength, po.hashtableLoadfactor);
    const care::read_number numReads = rs.getNumberOfReads();
    const std::size_t pitch = SequenceHelpers::getEncodedNumInts2Bit(rs.getSequenceLengthUpperBound());
    mh->setHostMemoryLimitForConstruction(po.memoryForHashtables);
    mh->setDeviceMemoryLimitsForConstruction({0});
    mh->setThreadPool(nullptr);
    int remaining = po.numHashFunctions;
    constexpr care::read_number batchsize = 1000000;
    std::vector<care::read_number> ids(batchsize);
    std::vector<unsigned int> seq(batchsize * pitch);
    std::vector<int> lens(batchsize);
    while(remaining > 0){
        const int existing = po.numHashFunctions - remaining;
        std::vector<int> fnums(remaining);
        std::iota(fnums.begin(), fnums.end(), existing);
        const int added = mh->addHashTables(remaining, fnums.data());
        if(added == 0) break;
        for(care::read_number b0 = 0; b0 < numReads; b0 += batchsize){
            const care::read_number cnt = std::min<care::read_number>(batchsize, numReads - b0);
            std::iota(ids.begin(), ids.begin() + cnt, b0);
            rs.gatherSequences(seq.data(), pitch, ids.data(), cnt);
            rs.gatherSequenceLengths(lens.data(), ids.data(), cnt);
            mh->insert(seq.data(), cnt, lens.data(), pitch, ids.data(), existing, added, fnums.data());
            if(mh->checkInsertionErrors(existing, added) > 0) throw std::runtime_error("table construction error");
        }
        mh->compact();
        remaining -= added;
    }
    mh->constructionIsFinished();
    return mh;
}

}  // namespace

extern "C" {

static void fill_program_options(care::ProgramOptions& po, const ref_opts* o, int maxLen){
    po.kmerlength = o->kmerlength;
    po.numHashFunctions = o->num_hash_maps;
    po.estimatedCoverage = o->estimated_coverage;
    po.threads = o->threads;
    po.useQualityScores = o->use_quality_scores != 0;
    po.excludeAmbiguousReads = o->exclude_ambiguous != 0;
    po.pairType = o->paired ? care::SequencePairType::PairedEnd : care::SequencePairType::SingleEnd;
    po.min_overlap = o->min_overlap;
    po.maxErrorRate = o->max_error_rate;
    po.min_overlap_ratio = o->min_overlap_ratio;
    po.estimatedErrorrate = o->estimated_errorrate;
    po.m_coverage = o->m_coverage;
    po.pairedFilterThreshold = o->paired_filter_threshold;
    po.batchsize = 16;  // forced by src/main_correct_cpu.cpp:105
    po.fixedNumberOfReads = o->fixed_number_of_reads;
    po.memoryTotalLimit = std::size_t(48) << 30;
    po.memoryForHashtables = std::size_t(40) << 30;
    po.hashtableLoadfactor = 0.8f;
    po.tempdirectory = "/tmp";
    po.maximum_sequence_length = maxLen;
    po.correctionType = care::CorrectionType::Classic;
    po.correctionTypeCands = care::CorrectionType::Classic;
}

// Creates the read storage and the reference minhasher (tables built by the
// reference's own construction routine).
ref_ctx* ref_storage(const ref_opts* o, std::size_t n, const unsigned int* enc, std::size_t pitchInts,
                     const int* lens, const char* quals, std::size_t qualPitch, const std::uint8_t* ambig);

ref_ctx* ref_create(const ref_opts* o, std::size_t n, const unsigned int* enc, std::size_t pitchInts,
                    const int* lens, const char* quals, std::size_t qualPitch, const std::uint8_t* ambig){
    auto* c = ref_storage(o, n, enc, pitchInts, lens, quals, qualPitch, ambig);
    c->mh = buildMinhasherSequential(c->po, c->rs);
    return c;
}

// The same storage with the minhasher read from a table file by the reference's own
// loader (OrdinaryCpuMinhasher::loadFromStream, include/ordinaryminhasher.hpp:318-339)
// instead of being built; *seconds = the load's wall time. Null when the file does not
// give every map.
ref_ctx* ref_create_loaded(const ref_opts* o, std::size_t n, const unsigned int* enc, std::size_t pitchInts,
                           const int* lens, const char* quals, std::size_t qualPitch, const std::uint8_t* ambig,
                           const char* tablesPath, double* seconds){
    auto* c = ref_storage(o, n, enc, pitchInts, lens, quals, qualPitch, ambig);
    const auto t0 = std::chrono::steady_clock::now();
    std::ifstream is(tablesPath, std::ios::binary);
    if(!is){ delete c; return nullptr; }
    auto mh = std::make_unique<care::OrdinaryCpuMinhasher>(c->rs.getNumberOfReads(),
                                                           calculateResultsPerMapThreshold(c->po.estimatedCoverage),
                                                           c->po.kmerlength, c->po.hashtableLoadfactor);
    const int loaded = mh->loadFromStream(is, c->po.numHashFunctions);
    if(loaded != c->po.numHashFunctions){ delete c; return nullptr; }
    c->mh = std::move(mh);
    if(seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return c;
}

// The read storage and options only: no minhasher (for the drop-in check, which builds its
// own tables on the device).
ref_ctx* ref_storage(const ref_opts* o, std::size_t n, const unsigned int* enc, std::size_t pitchInts,
                     const int* lens, const char* quals, std::size_t qualPitch, const std::uint8_t* ambig){
    auto* c = new ref_ctx;
    c->rs.n = n;
    c->rs.pitchInts = pitchInts;
    c->rs.seqs.assign(enc, enc + n * pitchInts);
    c->rs.lens.assign(lens, lens + n);
    c->rs.minLen = n ? *std::min_element(lens, lens + n) : 0;
    c->rs.maxLen = n ? *std::max_element(lens, lens + n) : 0;
    c->rs.paired = o->paired != 0;
    if(quals){
        c->rs.hasQual = true;
        c->rs.qualPitch = qualPitch;
        c->rs.quals.assign(quals, quals + n * qualPitch);
    }
    c->rs.ambig.assign(n, 0);
    if(ambig){
        for(std::size_t i = 0; i < n; i++){
            c->rs.ambig[i] = ambig[i];
            if(ambig[i]) c->rs.ambigIds.push_back(i);
        }
    }
    fill_program_options(c->po, o, c->rs.maxLen);
    return c;
}

void ref_destroy(ref_ctx* c){ delete c; }

// The reference minhasher's own file writer (OrdinaryCpuMinhasher::writeToStream,
// include/ordinaryminhasher.hpp:303-316); returns 0 on success.
int ref_save_tables(ref_ctx* c, const char* path){
    std::ofstream os(path, std::ios::binary);
    if(!os) return -1;
    c->mh->writeToStream(os);
    return os ? 0 : -1;
}

// Replaces the context's tables with those of a minhasher file, read by the reference's
// loader (OrdinaryCpuMinhasher::loadFromStream, :318-339); returns the maps loaded.
int ref_load_tables(ref_ctx* c, const char* path){
    std::ifstream is(path, std::ios::binary);
    if(!is) return -1;
    auto mh = std::make_unique<care::OrdinaryCpuMinhasher>(c->rs.getNumberOfReads(),
                                                           calculateResultsPerMapThreshold(c->po.estimatedCoverage),
                                                           c->po.kmerlength, c->po.hashtableLoadfactor);
    const int loaded = mh->loadFromStream(is, c->po.numHashFunctions);
    c->mh = std::move(mh);
    return loaded;
}

// The reference's read-storage construction from FASTQ files
// (constructChunkedReadStorageFromFiles, include/chunkedreadstorageconstruction.hpp:29-480:
// parser batches of 65,536 reads, lower case to upper case, other characters replaced
// cyclically by A, C, G, T with the counter reset per batch, paired files interleaved),
// exported as arrays.
struct ref_reads {
    std::size_t n = 0;
    int maxLen = 0;
    std::size_t pitchInts = 0;
    std::vector<unsigned int> enc;
    std::vector<int> lens;
    std::vector<char> quals;      // maxLen bytes per read
    std::vector<unsigned char> ambig;
    std::unique_ptr<care::ChunkedReadStorage> storage;
};

static ref_reads* readsFromStorage(std::unique_ptr<care::ChunkedReadStorage> storage, bool useQualities){
    auto* r = new ref_reads;
    r->n = storage->getNumberOfReads();
    r->maxLen = storage->getSequenceLengthUpperBound();
    r->pitchInts = SequenceHelpers::getEncodedNumInts2Bit(r->maxLen);
    std::vector<care::read_number> ids(r->n);
    std::iota(ids.begin(), ids.end(), care::read_number(0));
    r->enc.assign(r->n * r->pitchInts, 0u);
    r->lens.assign(r->n, 0);
    storage->gatherSequences(r->enc.data(), r->pitchInts, ids.data(), r->n);
    storage->gatherSequenceLengths(r->lens.data(), ids.data(), r->n);
    if(useQualities){
        r->quals.assign(r->n * std::size_t(r->maxLen), 0);
        storage->gatherQualities(r->quals.data(), r->maxLen, ids.data(), r->n);
    }
    r->ambig.assign(r->n, 0);
    std::vector<care::read_number> amb(storage->getNumberOfReadsWithN());
    storage->getIdsOfAmbiguousReads(amb.data());
    for(auto id : amb) r->ambig[id] = 1;
    r->storage = std::move(storage);
    return r;
}

ref_reads* ref_parse_fastq(const char* file1, const char* file2, int useQualities){
    care::ProgramOptions po;
    po.inputfiles.push_back(file1);
    if(file2) po.inputfiles.push_back(file2);
    po.pairType = file2 ? care::SequencePairType::PairedEnd : care::SequencePairType::SingleEnd;
    po.useQualityScores = useQualities != 0;
    po.qualityScoreBits = 8;
    po.threads = 1;
    po.showProgress = false;
    return readsFromStorage(care::constructChunkedReadStorageFromFiles(po), useQualities != 0);
}

// The reference's preprocessed-reads file (--save-preprocessedreads-to,
// ChunkedReadStorage::saveToFile, include/chunkedreadstorage.hpp:246-398).
int ref_reads_save(const ref_reads* r, const char* path){
    try{ r->storage->saveToFile(path); return 0; }catch(...){ return -1; }
}

// Read storage loaded from such a file (--load-preprocessedreads-from,
// ChunkedReadStorage::loadFromFile, include/chunkedreadstorage.hpp:158-244); null on error.
ref_reads* ref_reads_load(const char* path, int paired, int useQualities){
    try{
        auto storage = std::make_unique<care::ChunkedReadStorage>(paired != 0, useQualities != 0, 8);
        storage->loadFromFile(path);
        return readsFromStorage(std::move(storage), useQualities != 0);
    }catch(...){
        return nullptr;
    }
}

std::size_t ref_reads_count(const ref_reads* r){ return r->n; }
int ref_reads_max_length(const ref_reads* r){ return r->maxLen; }
std::size_t ref_reads_pitch(const ref_reads* r){ return r->pitchInts; }
void ref_reads_copy(const ref_reads* r, unsigned int* enc, int* lens, char* quals, unsigned char* ambig){
    std::copy(r->enc.begin(), r->enc.end(), enc);
    std::copy(r->lens.begin(), r->lens.end(), lens);
    if(quals) std::copy(r->quals.begin(), r->quals.end(), quals);
    std::copy(r->ambig.begin(), r->ambig.end(), ambig);
}
void ref_reads_free(ref_reads* r){ delete r; }

// Records of ref_correct in read-id order (1, the default) or in the order correct_cpu
// inserted them into its SerializedObjectStorage (0; deterministic with one thread).
void ref_set_record_order(ref_ctx* c, int sorted){ c->sortRecords = sorted != 0; }

// Candidate correction (--candidateCorrection, --candidateCorrectionNewColumns,
// include/options.hpp:41,52; classic candidate correction, include/corrector.hpp:1504-1512).
void ref_set_candidate_correction(ref_ctx* c, int on, int newColumns){
    c->po.correctCandidates = on != 0;
    c->po.new_columns_to_correct = newColumns;
    c->po.correctionTypeCands = care::CorrectionType::Classic;
}

// Switches the anchor correction to the feature print mode (include/corrector.hpp:1474-1493,
// include/classification.hpp:50-57): the features of every decision column go to `path`
// (sample rate 1); no anchor is corrected. Run with one thread for a whole-line file.
void ref_set_anchor_print(ref_ctx* c, const char* path){
    c->po.correctionType = care::CorrectionType::Print;
    c->po.mlForestfilePrintAnchor = path;
    c->po.sampleRateAnchor = 1.0f;
}

// Switches the anchor correction to the classifier seam (include/corrector.hpp:1448-1472)
// with the forest file at `path` (format include/forest.hpp:141-162); null restores Classic.
void ref_set_anchor_forest(ref_ctx* c, const char* path, float threshold){
    if(path){
        c->po.correctionType = care::CorrectionType::Forest;
        c->po.mlForestfileAnchor = path;
        c->po.thresholdAnchor = threshold;
    }else{
        c->po.correctionType = care::CorrectionType::Classic;
        c->po.mlForestfileAnchor = "";
    }
}

int ref_num_maps(ref_ctx* c){ return c->mh->getNumberOfMaps(); }

// End to end through care::cpu::correct_cpu; records are then ordered by read id
// (STEP3 does the same, src/dispatch_care_correct_cpu.cpp:409).
std::size_t ref_correct(ref_ctx* c, int threads, long long fixedNumberOfReads){
    c->po.threads = threads;
    c->po.fixedNumberOfReads = fixedNumberOfReads;
    auto t0 = std::chrono::steady_clock::now();
    care::SerializedObjectStorage storage = care::cpu::correct_cpu(c->po, *c->mh, c->rs);
    auto t1 = std::chrono::steady_clock::now();
    c->seconds = std::chrono::duration<double>(t1 - t0).count();

    const std::size_t num = storage.getNumElements();
    std::vector<std::pair<care::read_number, std::size_t>> order(num);
    for(std::size_t i = 0; i < num; i++){
        order[i] = {care::EncodedTempCorrectedSequence::parseReadId(storage.getPointer(i)), i};
    }
    if(c->sortRecords) std::sort(order.begin(), order.end());
    c->records.clear();
    c->recordOffsets.clear();
    for(const auto& p : order){
        const std::uint8_t* ptr = storage.getPointer(p.second);
        std::uint32_t flags;
        std::memcpy(&flags, ptr + 4, 4);
        const std::size_t bytes = 8 + (flags & ((1u << 29) - 1));
        c->recordOffsets.push_back(c->records.size());
        c->records.insert(c->records.end(), ptr, ptr + bytes);
    }
    c->recordOffsets.push_back(c->records.size());
    return num;
}

// Timed run of the reference engine (CpuErrorCorrector::process / processMulti)
// under an OpenMP loop over 16-read batches, like src/correct_cpu.cpp:134-299 but
// without the ProgressThread whose 1 s sleep granularity distorts short timings.
// Records are encoded as in src/correct_cpu.cpp:255-273 and counted, not kept.
double ref_correct_timed(ref_ctx* c, int threads, long long numAnchors, long long* numRecords){
    c->po.threads = threads;
    omp_set_num_threads(threads);
    const std::size_t total = numAnchors > 0 ? std::min<std::size_t>(numAnchors, c->rs.n) : c->rs.n;
    care::ReadCorrectionFlags flags(c->rs.n);
    care::ClfAgent clf_(c->po);
    std::atomic<std::size_t> next{0};
    std::atomic<long long> recs{0};
    auto t0 = std::chrono::steady_clock::now();
    #pragma omp parallel
    {
        care::ClfAgent clf = clf_;
        const std::size_t pitch = SequenceHelpers::getEncodedNumInts2Bit(c->rs.maxLen);
        const std::size_t qpitch = c->rs.maxLen;
        care::CpuErrorCorrector ec(pitch, c->rs.maxLen, qpitch, c->po, *c->mh, c->rs, flags, clf);
        const int bs = 16;
        std::vector<care::read_number> ids(bs);
        std::vector<unsigned int> enc(bs * pitch);
        std::vector<char> q(bs * qpitch);
        std::vector<int> lens(bs);
        long long local = 0;
        while(true){
            const std::size_t b = next.fetch_add(bs);
            if(b >= total) break;
            const int num = int(std::min<std::size_t>(bs, total - b));
            ids.resize(num);
            std::iota(ids.begin(), ids.end(), care::read_number(b));
            c->rs.gatherSequenceLengths(lens.data(), ids.data(), num);
            c->rs.gatherSequences(enc.data(), pitch, ids.data(), num);
            if(c->po.useQualityScores) c->rs.gatherQualities(q.data(), qpitch, ids.data(), num);
            care::CorrectionOutput out;
            if(c->rs.paired){
                care::CpuErrorCorrectorMultiInput in{};
                for(int i = 0; i < num; i++){
                    in.anchorReadIds.push_back(ids[i]);
                    in.encodedAnchors.push_back(enc.data() + i * pitch);
                    in.anchorQualityscores.push_back(q.data() + i * qpitch);
                    in.anchorLengths.push_back(lens[i]);
                }
                auto outs = ec.processMulti(in);
                for(auto& o : outs){
                    if(o.hasAnchorCorrection) out.anchorCorrections.emplace_back(std::move(o.anchorCorrection));
                }
            }else{
                for(int i = 0; i < num; i++){
                    care::CpuErrorCorrectorInput in;
                    in.anchorReadId = ids[i];
                    in.encodedAnchor = enc.data() + i * pitch;
                    in.anchorQualityscores = q.data() + i * qpitch;
                    in.anchorLength = lens[i];
                    auto o = ec.process(in);
                    if(o.hasAnchorCorrection) out.anchorCorrections.emplace_back(std::move(o.anchorCorrection));
                }
            }
            care::EncodedCorrectionOutput encoded = out;
            for(const auto& e : encoded.encodedAnchorCorrections){
                if(!(e.isHQ() && e.useEdits() && e.getNumEdits() == 0)) local++;
            }
        }
        recs += local;
    }
    auto t1 = std::chrono::steady_clock::now();
    if(numRecords) *numRecords = recs.load();
    return std::chrono::duration<double>(t1 - t0).count();
}

double ref_last_seconds(ref_ctx* c){ return c->seconds; }
std::size_t ref_records_bytes(ref_ctx* c){ return c->records.size(); }
const std::uint8_t* ref_records_data(ref_ctx* c){ return c->records.data(); }

// Minhash signature of one read (reference CPUSequenceHasher, map ids 0..nmaps-1).
void ref_hash(const unsigned int* enc, int len, int k, int nmaps, std::uint64_t* out){
    care::CPUSequenceHasher<care::kmer_type> hasher;
    hasher.hashInto(out, enc, len, k, nmaps, 0);
}

// Candidate ids of one anchor exactly as CpuErrorCorrector::determineCandidateReadIds
// computes them (include/corrector.hpp:571-650), using the reference minhasher.
int ref_candidates(ref_ctx* c, unsigned int readId, unsigned int* out, int cap){
    const std::size_t pitch = SequenceHelpers::getEncodedNumInts2Bit(c->rs.maxLen);
    std::vector<unsigned int> enc(pitch);
    c->rs.gatherSequences(enc.data(), pitch, &readId, 1);
    int len = c->rs.lens[readId];
    care::MinhasherHandle h = c->mh->makeMinhasherHandle();
    int numPer = 0, total = 0;
    std::vector<care::read_number> ids;
    bool containsN = c->rs.ambig[readId] != 0;
    if(!(c->po.excludeAmbiguousReads && containsN)){
        c->mh->determineNumValues(h, enc.data(), pitch, &len, 1, &numPer, total);
        ids.resize(total);
        int offsets[2];
        c->mh->retrieveValues(h, 1, total, ids.data(), &numPer, offsets);
        std::sort(ids.begin(), ids.end());
        ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
        auto it = std::lower_bound(ids.begin(), ids.end(), readId);
        if(it != ids.end() && *it == readId) ids.erase(it);
        if(c->po.excludeAmbiguousReads){
            ids.erase(std::remove_if(ids.begin(), ids.end(), [&](care::read_number r){ return c->rs.ambig[r] != 0; }), ids.end());
        }
    }
    c->mh->destroyHandle(h);
    const int n = int(ids.size());
    for(int i = 0; i < n && i < cap; i++) out[i] = ids[i];
    return n;
}

// One shifted-Hamming alignment (reference routine) of a 2-bit anchor and a 2-bit
// candidate; out = {score, overlap, shift, nOps, isValid}.
// include/util.hpp:1153 as called by include/corrector.hpp:785-801
void ref_flag_pairs(const unsigned int* c1, int n1, const unsigned int* c2, int n2, unsigned char* f1, unsigned char* f2){
    std::vector<int> p1(std::min(n1, n2)), p2(std::min(n1, n2));
    auto ends = findPositionsOfPairedReadIds(c1, c1 + n1, c2, c2 + n2, p1.begin(), p2.begin());
    p1.erase(ends.first, p1.end());
    p2.erase(ends.second, p2.end());
    std::fill(f1, f1 + n1, 0);
    std::fill(f2, f2 + n2, 0);
    for(int i : p1) f1[i] = 1;
    for(int i : p2) f2[i] = 1;
}

void ref_align(const unsigned int* a, int la, const unsigned int* cnd, int lc, int min_overlap,
               float maxErrorRate, float min_overlap_ratio, int* out){
    care::cpu::shd::CpuAlignmentHandle handle;
    care::cpu::shd::AlignmentResult r;
    const int pitch = SequenceHelpers::getEncodedNumInts2Bit(lc);
    care::cpu::shd::cpuShiftedHammingDistancePopcount2Bit(handle, &r, a, la, cnd, pitch, &lc, 1,
                                                          min_overlap, maxErrorRate, min_overlap_ratio);
    out[0] = r.score; out[1] = r.overlap; out[2] = r.shift; out[3] = r.nOps; out[4] = r.isValid ? 1 : 0;
}

int ref_orientation(int fo, int ro, int fn, int rn, int fv, int rv, int la, int lc,
                    float min_overlap_ratio, int min_overlap, float estimatedErrorrate){
    care::cpu::shd::AlignmentResult f{}, r{};
    f.overlap = fo; f.nOps = fn; f.isValid = fv != 0;
    r.overlap = ro; r.nOps = rn; r.isValid = rv != 0;
    return int(care::chooseBestAlignmentOrientation(f, r, la, lc, min_overlap_ratio, min_overlap, estimatedErrorrate));
}

// Builds a reference MSA from decoded rows and returns column data.
// out arrays sized >= ncols (query with ncols_out first using NULL outputs).
int ref_msa(int useQ, int la, int nc, const char* anchor, const char* aq, const char* cands, int cpitch,
            const char* cq, int qpitch, const int* clens, const int* shifts, const float* wf,
            int* beginAnchor, int* countsACGT, float* weightsACGT, int* coverage, char* consensus,
            float* support, float* origW, int* origCov){
    care::cpu::QualityScoreConversion conv;
    care::MultipleSequenceAlignment msa(&conv);
    care::MultipleSequenceAlignment::InputData in;
    in.useQualityScores = useQ != 0;
    in.anchorLength = la;
    in.nCandidates = nc;
    in.candidatesPitch = cpitch;
    in.candidateQualitiesPitch = qpitch;
    in.anchor = anchor;
    in.candidates = cands;
    in.anchorQualities = aq;
    in.candidateQualities = cq;
    in.candidateLengths = clens;
    in.candidateShifts = shifts;
    in.candidateDefaultWeightFactors = wf;
    msa.build(in);
    const int ncol = msa.nColumns;
    if(beginAnchor) *beginAnchor = msa.anchorColumnsBegin_incl;
    if(countsACGT){
        for(int i = 0; i < ncol; i++){
            countsACGT[4*i+0] = msa.countsA[i]; countsACGT[4*i+1] = msa.countsC[i];
            countsACGT[4*i+2] = msa.countsG[i]; countsACGT[4*i+3] = msa.countsT[i];
            weightsACGT[4*i+0] = msa.weightsA[i]; weightsACGT[4*i+1] = msa.weightsC[i];
            weightsACGT[4*i+2] = msa.weightsG[i]; weightsACGT[4*i+3] = msa.weightsT[i];
            coverage[i] = msa.coverage[i]; consensus[i] = msa.consensus[i]; support[i] = msa.support[i];
            origW[i] = msa.origWeights[i]; origCov[i] = msa.origCoverages[i];
        }
    }
    return ncol;
}

// Output construction as RabbitQCPlus runs it: the serialized records are sorted by read
// id (sortSerializedResultsByReadIdAscending, include/sortserializedresults.hpp:17-78) and
// merged with the original FASTQ file(s) into the queues QC reads from
// (constructOutputFileFromCorrectionResultsOutToQueue, src/correctionresultoutput.cpp:1589-1632,
// merge :448-975, combine rule :1256-1550). Single-end runs pass one queue twice (src/seqc.cpp:256),
// paired runs two queues and two output names (src/peqc.cpp:135-148,300). What reaches queue 1
// (and queue 2) is written to q1_path (q2_path). Scratch files go to `tmpdir`. Returns the
// number of records (the reference's changNum) or -1.
int ref_output_fastq(const char* in1, const char* in2, const uint8_t* records, size_t bytes, int labels,
                     const char* tmpdir, const char* q1_path, const char* q2_path){
    try{
        care::ProgramOptions po;
        po.outputCorrectionQualityLabels = labels != 0;
        po.pairType = in2 ? care::SequencePairType::PairedEnd : care::SequencePairType::SingleEnd;
        std::vector<std::string> inputs{in1};
        std::vector<std::string> outputs{std::string(tmpdir) + "/tmp1.fq"};
        if(in2){
            inputs.push_back(in2);
            outputs.push_back(std::string(tmpdir) + "/tmp2.fq");
        }
        care::SerializedObjectStorage storage(size_t(1) << 30, size_t(1) << 30, std::string(tmpdir) + "/");
        size_t p = 0;
        while(p + 8 <= bytes){
            uint32_t flags;
            std::memcpy(&flags, records + p + 4, 4);
            const size_t len = 8 + (flags & ((1u << 29) - 1));
            if(p + len > bytes) return -1;
            storage.insert(records + p, records + p + len);
            p += len;
        }
        care::sortSerializedResultsByReadIdAscending<care::EncodedTempCorrectedSequence>(storage, size_t(1) << 30);
        moodycamel::ReaderWriterQueue<std::pair<char*, int>> q1, q2;
        auto* second = in2 ? &q2 : &q1;
        std::atomic_int producerDone{0}, careStartWrite{0};
        int changNum = 0;
        std::atomic<bool> finished{false};
        std::thread worker([&]{
            care::constructOutputFileFromCorrectionResultsOutToQueue(inputs, storage, care::FileFormat::FASTQ, outputs,
                                                                     false, po, &q1, second, &producerDone,
                                                                     &careStartWrite, &changNum);
            finished = true;
        });
        std::ofstream o1(q1_path, std::ios::binary), o2(q2_path, std::ios::binary);
        auto drain = [](moodycamel::ReaderWriterQueue<std::pair<char*, int>>& q, std::ofstream& o){
            std::pair<char*, int> item;
            bool any = false;
            while(q.try_dequeue(item)){
                o.write(item.first, item.second);
                delete[] item.first;
                any = true;
            }
            return any;
        };
        while(!finished){
            const bool a = drain(q1, o1);
            const bool b = in2 ? drain(q2, o2) : false;
            if(!a && !b) std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
        worker.join();
        drain(q1, o1);
        if(in2) drain(q2, o2);
        return changNum;
    }catch(...){
        return -1;
    }
}

}  // extern "C"

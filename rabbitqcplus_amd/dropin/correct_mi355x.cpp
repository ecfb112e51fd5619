// care::mi355x::correct: the reference's correction stage (care::cpu::correct_cpu,
// reference src/correct_cpu.cpp:47-348) run on one MI355X through the C-ABI of
// include/care_hip.h. It reads the read storage through the reference's CpuReadStorage
// interface (include/cpureadstorage.hpp:16-79), maps ProgramOptions
// (include/options.hpp:36-123) onto care_opts, and returns the records in a
// SerializedObjectStorage (include/serializedobjectstorage.hpp:14-60) sized the way
// correct_cpu sizes its own (src/correct_cpu.cpp:50-95).
#include "correct_mi355x.hpp"

#include <care_hip.h>
#include <sequencehelpers.hpp>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace care {
namespace mi355x {

namespace {

care_opts optionsFrom(const ProgramOptions& po){
    care_opts o;
    care_hip_default_opts(&o);
    o.kmerlength = po.kmerlength;
    o.num_hash_maps = po.numHashFunctions;
    o.estimated_coverage = po.estimatedCoverage;
    o.use_quality_scores = po.useQualityScores ? 1 : 0;
    o.exclude_ambiguous = po.excludeAmbiguousReads ? 1 : 0;
    o.paired = po.pairType == SequencePairType::PairedEnd ? 1 : 0;
    o.min_overlap = po.min_overlap;
    o.max_error_rate = po.maxErrorRate;
    o.min_overlap_ratio = po.min_overlap_ratio;
    o.estimated_errorrate = po.estimatedErrorrate;
    o.m_coverage = po.m_coverage;
    o.paired_filter_threshold = po.pairedFilterThreshold;
    o.correct_candidates = po.correctCandidates ? 1 : 0;
    o.new_columns_to_correct = po.new_columns_to_correct;
    return o;
}

// Owns one engine context; every failing call throws with the engine's message.
struct Engine {
    care_ctx* ctx = nullptr;
    Engine(const care_opts& o, int device){
        const int rc = care_hip_create(&o, device, &ctx);
        if(rc != CARE_OK) throw std::runtime_error("MI355X correction: care_hip_create failed with code " + std::to_string(rc));
    }
    ~Engine(){ care_hip_destroy(ctx); }
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;
    void check(int rc, const char* what) const {
        if(rc != CARE_OK) throw std::runtime_error(std::string("MI355X correction: ") + what + ": " + care_hip_last_error(ctx));
    }
};

}  // namespace

SerializedObjectStorage correct(const ProgramOptions& po, const CpuReadStorage& rs, int device){
    const care_opts o = optionsFrom(po);
    const read_number n = rs.getNumberOfReads();
    const int maxLen = rs.getSequenceLengthUpperBound();
    const std::size_t pitch = SequenceHelpers::getEncodedNumInts2Bit(maxLen);

    // the storage as the arrays care_hip_set_reads takes
    std::vector<unsigned int> enc(std::size_t(n) * pitch);
    std::vector<int> lengths(n);
    std::vector<read_number> ids(n);
    for(read_number i = 0; i < n; i++) ids[i] = i;
    rs.gatherContiguousSequences(enc.data(), pitch, 0, int(n));
    rs.gatherSequenceLengths(lengths.data(), ids.data(), int(n));
    std::vector<char> quals;
    if(po.useQualityScores){
        quals.resize(std::size_t(n) * std::size_t(maxLen));
        rs.gatherQualities(quals.data(), std::size_t(maxLen), ids.data(), int(n));
    }
    std::vector<std::uint8_t> ambiguous(n, 0);
    std::vector<read_number> withN(std::size_t(rs.getNumberOfReadsWithN()));
    if(!withN.empty()) rs.getIdsOfAmbiguousReads(withN.data());
    for(read_number r : withN) ambiguous[r] = 1;

    Engine e(o, device);
    e.check(care_hip_set_reads(e.ctx, enc.data(), pitch, lengths.data(),
                               po.useQualityScores ? reinterpret_cast<const std::uint8_t*>(quals.data()) : nullptr,
                               std::size_t(maxLen), n, ambiguous.data()), "care_hip_set_reads");
    int maps = 0;
    e.check(care_hip_build_tables(e.ctx, &maps), "care_hip_build_tables");
    std::size_t numRecords = 0, numBytes = 0;
    e.check(care_hip_correct(e.ctx, 0, n, &numRecords, &numBytes), "care_hip_correct");
    // the records in the engine's host buffer, read in place (valid while the context lives)
    const std::uint8_t* rec = nullptr;
    std::size_t recBytes = 0;
    e.check(care_hip_records_view(e.ctx, &rec, &recBytes), "care_hip_records_view");
    if(recBytes != numBytes) throw std::runtime_error("MI355X correction: record byte count mismatch");

    // the storage correct_cpu returns, sized as it sizes its own (src/correct_cpu.cpp:50-95):
    // host memory left after the read storage (no CPU minhasher here), minus 2 GiB; 3/4
    // of it for the record bytes, 1/4 for the offsets; the rest spills to tempdirectory
    std::size_t available = po.memoryTotalLimit;
    const std::size_t storageBytes = rs.getMemoryInfo().host;
    available = available > storageBytes ? available - storageBytes : 0;
    const std::size_t partial = available > (std::size_t(2) << 30) ? available - (std::size_t(2) << 30) : 0;
    SerializedObjectStorage out(std::size_t(partial * 0.75), std::size_t(partial * 0.25), po.tempdirectory + "/");
    for(std::size_t p = 0; p < recBytes;){
        // one record: read id (u32), flags (u32: body size in the low 29 bits), body
        std::uint32_t flags;
        std::memcpy(&flags, rec + p + 4, 4);
        const std::size_t len = 8 + (flags & ((1u << 29) - 1));
        if(p + len > recBytes) throw std::runtime_error("MI355X correction: truncated record stream");
        out.insert(rec + p, rec + p + len);
        p += len;
    }
    return out;
}

}  // namespace mi355x
}  // namespace care

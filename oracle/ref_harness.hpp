// Test infrastructure only: types shared by the reference harness (ref_harness.cpp) and
// the drop-in check (dropin_harness.cpp). A CpuReadStorage over plain arrays (the
// interface is the reference's include/cpureadstorage.hpp:11-84), the option block and
// the harness context.
#ifndef CARE_REF_HARNESS_HPP
#define CARE_REF_HARNESS_HPP

#include <config.hpp>
#include <options.hpp>
#include <cpureadstorage.hpp>
#include <cpuminhasher.hpp>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <vector>

struct ArrayReadStorage : public care::CpuReadStorage {
    std::size_t n = 0;
    std::size_t pitchInts = 0;
    std::size_t qualPitch = 0;
    int minLen = 0;
    int maxLen = 0;
    bool paired = false;
    bool hasQual = false;
    std::vector<unsigned int> seqs;
    std::vector<int> lens;
    std::vector<char> quals;
    std::vector<std::uint8_t> ambig;
    std::vector<care::read_number> ambigIds;

    void areSequencesAmbiguous(bool* result, const care::read_number* ids, int num) const override {
        for(int i = 0; i < num; i++) result[i] = ambig[ids[i]] != 0;
    }
    void gatherSequences(unsigned int* out, std::size_t outPitch, const care::read_number* ids, int num) const override {
        for(int i = 0; i < num; i++){
            const std::size_t l = std::min(outPitch, pitchInts);
            std::copy_n(seqs.data() + pitchInts * ids[i], l, out + outPitch * i);
        }
    }
    void gatherContiguousSequences(unsigned int* out, std::size_t outPitch, care::read_number first, int num) const override {
        for(int i = 0; i < num; i++){
            const std::size_t l = std::min(outPitch, pitchInts);
            std::copy_n(seqs.data() + pitchInts * (first + i), l, out + outPitch * i);
        }
    }
    void gatherQualities(char* out, std::size_t outPitch, const care::read_number* ids, int num) const override {
        for(int i = 0; i < num; i++){
            const std::size_t l = std::min(outPitch, qualPitch);
            std::copy_n(quals.data() + qualPitch * ids[i], l, out + outPitch * i);
        }
    }
    void gatherEncodedQualities(unsigned int*, std::size_t, const care::read_number*, int) const override {
        throw std::runtime_error("not used");
    }
    void gatherContiguousEncodedQualities(unsigned int*, std::size_t, care::read_number, int) const override {
        throw std::runtime_error("not used");
    }
    void gatherSequenceLengths(int* out, const care::read_number* ids, int num) const override {
        for(int i = 0; i < num; i++) out[i] = lens[ids[i]];
    }
    void getIdsOfAmbiguousReads(care::read_number* ids) const override {
        std::copy(ambigIds.begin(), ambigIds.end(), ids);
    }
    std::int64_t getNumberOfReadsWithN() const override { return ambigIds.size(); }
    MemoryUsage getMemoryInfo() const override {
        MemoryUsage m{};
        m.host = seqs.size() * 4 + lens.size() * 4 + quals.size();
        return m;
    }
    care::read_number getNumberOfReads() const override { return n; }
    bool canUseQualityScores() const override { return hasQual; }
    int getSequenceLengthLowerBound() const override { return minLen; }
    int getSequenceLengthUpperBound() const override { return maxLen; }
    bool isPairedEnd() const override { return paired; }
    int getQualityBits() const override { return 8; }
};

extern "C" {

// Plain-data option block; field meaning follows care::ProgramOptions
// (include/options.hpp:36-123).
struct ref_opts {
    int kmerlength;
    int num_hash_maps;
    float estimated_coverage;
    int threads;
    int use_quality_scores;
    int exclude_ambiguous;
    int paired;
    int min_overlap;
    float max_error_rate;
    float min_overlap_ratio;
    float estimated_errorrate;
    float m_coverage;
    float paired_filter_threshold;
    long long fixed_number_of_reads;  // 0 = all
};

struct ref_ctx {
    ArrayReadStorage rs;
    care::ProgramOptions po;
    std::unique_ptr<care::CpuMinhasher> mh;
    std::vector<std::uint8_t> records;
    std::vector<std::size_t> recordOffsets;
    double seconds = 0.0;
    bool sortRecords = true;        // false: keep correct_cpu's insertion order
};

}  // extern "C"

#endif

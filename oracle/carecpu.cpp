// CPU restatement of the CARE per-anchor correction path.
//
// TEST INFRASTRUCTURE ONLY. This is the oracle: tests, the smoke check and the
// benchmark's CPU-baseline leg may load it, as the checker. The product never links
// it and never falls back to it.
//
// Every function names the reference code whose behaviour it restates (paths are
// relative to the RabbitQCPlus source tree). The restatement works on decoded base
// indices instead of the reference's HiLo bit planes and std::vector plumbing; the
// arithmetic that decides results (integer thresholds, float expressions and their
// evaluation order, tie-breaking, scan order) follows the reference exactly.
//
// Parity status of each part is recorded in the test suite: every function here is
// checked against the reference engine compiled from source (oracle/_ref) and/or
// against brute-force and closed-form checks.

#include "carecpu.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

#include <omp.h>

namespace {

constexpr int kMaxMaps = 48;               // include/config.hpp:30
const char kDecode[4] = {'A', 'C', 'G', 'T'};

// 2-bit row: A=0 C=1 G=2 T=3, 16 bases per u32, first base in the two most
// significant bits (include/sequencehelpers.hpp:154-187, 193-199).
inline int baseAt(const uint32_t* row, int i){
    return int((row[i >> 4] >> (30 - 2 * (i & 15))) & 3u);
}

inline int baseIndexOfChar(char c){
    switch(c){ case 'A': return 0; case 'C': return 1; case 'G': return 2; case 'T': return 3; }
    return -1;
}

// MurmurHash3 64-bit finalizer (include/hpc_helpers/include/hashers.cuh:129-137).
inline uint64_t fmix64(uint64_t x){
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

// Minhash signature: for every canonical k-mer (min of forward and reverse complement
// as 2k-bit integers, include/sequencehelpers.hpp:844-923) and map i, take
// min fmix64(kmer + i); mask to 2k bits after the min
// (include/cpusequencehasher.hpp:88-125). Reads shorter than k give all-ones & mask.
//
// With `seeds` (an addition for the seeded alignment; the reference keeps no positions):
// per map, the start position of the first k-mer attaining the minimum times 2, plus 1
// when its canonical form is the reverse complement (0 for reads shorter than k).
void signature(const uint32_t* row, int L, int k, int nmaps, uint64_t* out, uint16_t* seeds = nullptr){
    const uint64_t mask = std::numeric_limits<uint64_t>::max() >> ((32 - k) * 2);
    uint64_t mins[kMaxMaps];
    uint16_t seed[kMaxMaps];
    for(int m = 0; m < nmaps; m++){ mins[m] = std::numeric_limits<uint64_t>::max(); seed[m] = 0; }
    if(L >= k){
        uint64_t fwd = 0, rc = 0;
        const int topShift = 2 * k - 2;
        for(int i = 0; i < L; i++){
            const uint64_t b = uint64_t(baseAt(row, i));
            fwd = ((fwd << 2) | b) & mask;
            rc = (rc >> 2) | ((3u - b) << topShift);
            if(i >= k - 1){
                const uint64_t canon = std::min(fwd, rc);
                for(int m = 0; m < nmaps; m++){
                    const uint64_t h = fmix64(canon + uint64_t(m));
                    if(h < mins[m]){
                        mins[m] = h;
                        seed[m] = uint16_t(((i - k + 1) << 1) | (rc < fwd ? 1 : 0));
                    }
                }
            }
        }
    }
    for(int m = 0; m < nmaps; m++) out[m] = mins[m] & mask;
    if(seeds) for(int m = 0; m < nmaps; m++) seeds[m] = seed[m];
}

// calculateResultsPerMapThreshold(int coverage) (include/minhasherlimit.hpp:10-16),
// called with the float coverage (src/cpuminhasherconstruction.cpp:179).
int resultsPerMapThreshold(float coverage){
    const int cov = int(coverage);
    int result = int(cov * 2.5f);
    result = std::min(result, int(std::numeric_limits<uint16_t>::max()));
    result = std::max(10, result);
    return result;
}

struct Alignment {
    int score = 0, overlap = 0, shift = 0, nOps = 0;
    bool valid = false;
};

// Best shift of candidate vs anchor (src/cpu_alignment.cpp:32-497, wrapped by
// include/cpu_alignment.hpp:411-464). Scan order: shift 0..La-minoverlap (anchor
// shifted), then -1..-(Lc-minoverlap) (candidate shifted); per shift
// max_errors = min(int(float(overlap)*maxErrorRate), best-total+2*overlap), break
// the direction when it is <= 0; a shift scores mismatches+total-2*overlap when
// mismatches < max_errors; strict '<' improvement.
Alignment align(const uint8_t* A, int La, const uint8_t* C, int Lc, int min_overlap,
                float maxErrorRate, float min_overlap_ratio){
    const int totalbases = La + Lc;
    const int minoverlap = std::max(min_overlap, int(float(La) * min_overlap_ratio));
    int bestScore = totalbases;
    int bestShift = -Lc;

    auto evaluate = [&](int shift, int overlapsize, const uint8_t* L_, int lOffset, const uint8_t* R_){
        const int max_errors = std::min(int(float(overlapsize) * maxErrorRate),
                                        bestScore - totalbases + 2 * overlapsize);
        if(max_errors <= 0) return false;
        int mismatches = 0;
        if(overlapsize == 0){
            mismatches = max_errors + 1;
        }else{
            for(int j = 0; j < overlapsize && mismatches < max_errors; j++){
                mismatches += (L_[lOffset + j] != R_[j]);
            }
        }
        const int score = mismatches < max_errors ? mismatches + totalbases - 2 * overlapsize
                                                  : std::numeric_limits<int>::max();
        if(score < bestScore){
            bestScore = score;
            bestShift = shift;
        }
        return true;
    };

    for(int shift = 0; shift < La - minoverlap + 1; ++shift){
        const int overlapsize = std::min(La - shift, Lc);
        if(!evaluate(shift, overlapsize, A, shift, C)) break;
    }
    for(int shift = -1; shift >= -Lc + minoverlap; --shift){
        const int overlapsize = std::min(La, Lc + shift);
        if(!evaluate(shift, overlapsize, C, -shift, A)) break;
    }

    Alignment r;
    r.valid = (bestShift != -Lc);
    const int cb = std::max(-bestShift, 0);
    const int ce = std::min(Lc, La - bestShift);
    r.overlap = ce - cb;
    r.score = bestScore;
    r.shift = bestShift;
    r.nOps = bestScore - totalbases + 2 * r.overlap;
    return r;
}

// chooseBestAlignmentOrientation (include/alignmentorientation.hpp:14-92); the error
// bound passed down is estimatedErrorrate * 4.0f (:91). 1=fwd, 2=revcomp, 3=none.
int orientation(const Alignment& f, const Alignment& r, int La, float min_overlap_ratio,
                int min_overlap, float estimatedErrorrate){
    const float maxErrorRate = estimatedErrorrate * 4.0f;
    const int minimumOverlap = int(La * min_overlap_ratio) > min_overlap ? int(La * min_overlap_ratio) : min_overlap;
    int retval = 3;
    if(f.valid && f.overlap >= minimumOverlap){
        if(r.valid && r.overlap >= minimumOverlap){
            const float ratio = (float)f.nOps / f.overlap;
            const float revcomplratio = (float)r.nOps / r.overlap;
            if(ratio < revcomplratio){
                if(ratio < maxErrorRate) retval = 1;
            }else if(revcomplratio < ratio){
                if(revcomplratio < maxErrorRate) retval = 2;
            }else{
                if(ratio < maxErrorRate){
                    retval = f.overlap > r.overlap ? 1 : 2;
                }
            }
        }else{
            if((float)f.nOps / f.overlap < maxErrorRate) retval = 1;
        }
    }else{
        if(r.valid && r.overlap >= minimumOverlap){
            if((float)r.nOps / r.overlap < maxErrorRate) retval = 2;
        }
    }
    return retval;
}

// feq/fleq/fgeq with 1e-5 slack (include/hostdevicefunctions.cuh:11-29).
inline bool feq(float l, float r){
    const float absdiff = l - r < 0 ? -(l - r) : l - r;
    return absdiff < 1e-5f;
}
inline bool fleq(float l, float r){ return (l < r) || feq(l, r); }
inline bool fgeq(float l, float r){ return (l > r) || feq(l, r); }

// Per-candidate overlap weight (include/hostdevicefunctions.cuh:93-97), called with
// maxErrorRate (include/corrector.hpp:1273-1278).
inline float overlapWeight(int nOps, int overlap, float maxErrorRate){
    return 1.0f - sqrtf(nOps / (overlap * maxErrorRate));
}

// Quality-score weight table (include/qualityscoreweights.hpp:12-44).
struct QualityWeights {
    float w[256];
    QualityWeights(){
        for(int i = 0; i < 256; i++){
            const float p = i < 33 ? 1.0f : std::pow(10.0f, -(i - 33) / 10.0f);
            w[i] = std::max(0.001f, 1.0f - p);
        }
    }
};
const QualityWeights& qualityWeights(){
    static QualityWeights q;
    return q;
}

// Pileup of the anchor and its candidates (src/msa.cpp:9-173, include/msa.hpp:64-213).
struct Msa {
    int ncol = 0, ab = 0, ae = 0, ncand = 0;
    std::vector<int> counts;      // [col*4 + b]
    std::vector<float> weights;   // [col*4 + b]
    std::vector<int> coverage;
    std::vector<char> consensus;
    std::vector<float> support;
    std::vector<float> origWeights;
    std::vector<int> origCoverages;
};

struct CandRow {
    uint32_t id;
    int length;
    int shift;
    int nOps;
    int overlap;
    float weight;
    int orient;       // 1 fwd, 2 rc
    bool paired;
    std::vector<uint8_t> bases;   // oriented
    const char* qual;             // quality row (qualityPitch bytes) or null
    bool reverseQual;
};

void buildMsa(Msa& msa, const std::vector<uint8_t>& anchor, const char* anchorQual, bool useQ,
              int qualityPitch, const std::vector<CandRow>& cands){
    const int La = int(anchor.size());
    const int nc = int(cands.size());
    int startindex = 0, endindex = La;
    for(const auto& c : cands){
        startindex = std::min(c.shift, startindex);
        endindex = std::max(c.length + c.shift, endindex);
    }
    msa.ncand = nc;
    msa.ncol = endindex - startindex;
    msa.ab = std::max(-startindex, 0);
    msa.ae = msa.ab + La;
    const int n = msa.ncol;
    msa.counts.assign(size_t(n) * 4, 0);
    msa.weights.assign(size_t(n) * 4, 0.0f);
    msa.coverage.assign(n, 0);
    msa.consensus.assign(n, 0);
    msa.support.assign(n, 0.0f);
    msa.origWeights.assign(n, 0.0f);
    msa.origCoverages.assign(n, 0);
    const float* qw = qualityWeights().w;

    auto add = [&](const uint8_t* seq, int len, int shift, float factor, auto qualityOf){
        for(int i = 0; i < len; i++){
            const int col = msa.ab + shift + i;
            const float weight = factor * (useQ ? qw[(unsigned char)qualityOf(i)] : 1.0f);
            msa.counts[size_t(col) * 4 + seq[i]]++;
            msa.weights[size_t(col) * 4 + seq[i]] += weight;
            msa.coverage[col]++;
        }
    };
    add(anchor.data(), La, 0, 1.0f, [&](int i){ return anchorQual ? anchorQual[i] : char(0); });
    for(const auto& c : cands){
        add(c.bases.data(), c.length, c.shift, c.weight, [&](int i){
            // reverseQualitiesOfRCAlignments reverses the whole pitch-wide row
            // (include/corrector.hpp:1227-1239).
            return c.reverseQual ? c.qual[qualityPitch - 1 - i] : c.qual[i];
        });
    }
    // findConsensus (src/msa.cpp:135-157)
    for(int col = 0; col < n; col++){
        const float* w = &msa.weights[size_t(col) * 4];
        int cons = 0;
        float consWeight = w[0];
        if(w[1] > consWeight){ cons = 1; consWeight = w[1]; }
        if(w[2] > consWeight){ cons = 2; consWeight = w[2]; }
        if(w[3] > consWeight){ cons = 3; consWeight = w[3]; }
        msa.consensus[col] = kDecode[cons];
        const float columnWeight = w[0] + w[1] + w[2] + w[3];
        msa.support[col] = consWeight / columnWeight;
    }
    // findOrigWeightAndCoverage (src/msa.cpp:159-173)
    for(int col = msa.ab; col < msa.ae; col++){
        const int b = anchor[col - msa.ab];
        msa.origWeights[col] = msa.weights[size_t(col) * 4 + b];
        msa.origCoverages[col] = msa.counts[size_t(col) * 4 + b];
    }
}

// findCandidatesOfDifferentRegion (src/msa.cpp:530-734). Returns true if a column
// was found; `remove` flags the candidates to drop (all false if a flagged
// candidate has overlap weight >= 0.90).
bool findDifferentRegion(const Msa& msa, const std::vector<uint8_t>& anchor, const std::vector<CandRow>& cands,
                         int datasetCoverage, std::vector<char>& remove){
    const int La = int(anchor.size());
    bool anyMismatch = false;
    for(int i = 0; i < La; i++){
        if(kDecode[anchor[i]] != msa.consensus[msa.ab + i]){ anyMismatch = true; break; }
    }
    if(!anyMismatch) return false;

    const int significantThreshold = int(datasetCoverage * 0.3f);
    int col = -1;
    int foundBaseIndex = -1;
    for(int c = msa.ab; c < msa.ae; c++){
        const int consindex = baseIndexOfChar(msa.consensus[c]);
        int significant = -1;
        for(int i = 0; i < 4; i++){
            if(i != consindex && significantThreshold <= msa.counts[size_t(c) * 4 + i]) significant = i;
        }
        if(significant != -1){ col = c; foundBaseIndex = significant; break; }
    }
    if(col < 0) return false;

    const int origBase = anchor[col - msa.ab];
    const bool keepMatching = (origBase == foundBaseIndex);
    remove.assign(cands.size(), 0);
    for(size_t k = 0; k < cands.size(); k++){
        const int rb = msa.ab + cands[k].shift;
        const int re = rb + cands[k].length;
        const bool notAffected = (col < rb || re <= col);
        if(notAffected) continue;
        const int base = cands[k].bases[col - rb];
        if(keepMatching && base == foundBaseIndex) continue;
        if(!keepMatching && base != foundBaseIndex) continue;
        remove[k] = 1;
    }
    bool veryGood = false;
    for(size_t k = 0; k < cands.size(); k++){
        if(remove[k] && cands[k].weight >= 0.90f) veryGood = true;
    }
    if(veryGood) std::fill(remove.begin(), remove.end(), 0);
    return true;
}

struct Properties {
    float avg_support, min_support;
    int min_coverage, max_coverage;
    bool isHQ, allGood;
};

// getMSAProperties (src/msa.cpp:304-376).
Properties msaProperties(const Msa& msa, float e, float cov, float m_cov){
    const float avg_support_threshold = 1.0f - 1.0f * e;
    const float min_support_threshold = 1.0f - 3.0f * e;
    const float min_coverage_threshold = m_cov / 6.0f * cov;
    const int first = msa.ab, last = msa.ae;
    Properties p;
    p.min_support = *std::min_element(msa.support.begin() + first, msa.support.begin() + last);
    float sum = 0.0f;
    for(int c = first; c < last; c++) sum += msa.support[c];
    p.avg_support = sum / (last - first);
    p.min_coverage = *std::min_element(msa.coverage.begin() + first, msa.coverage.begin() + last);
    p.max_coverage = *std::max_element(msa.coverage.begin() + first, msa.coverage.begin() + last);
    auto goodAvg = [&](float v){ return fgeq(v, avg_support_threshold); };
    auto goodMin = [&](float v){ return fgeq(v, min_support_threshold); };
    auto goodCov = [&](float v){ return fgeq(v, min_coverage_threshold); };
    p.allGood = goodAvg(p.avg_support) && goodMin(p.min_support) && goodCov(float(p.min_coverage));
    p.isHQ = false;
    if(p.allGood){
        int smallest = 100;
        const int pct = int(std::ceil(e * 100.0f));
        for(int percent = pct; percent >= 0; percent--){
            const float factor = percent / 100.0f;
            const float avg_t = 1.0f - 1.0f * factor;
            const float min_t = 1.0f - 3.0f * factor;
            if(fgeq(p.avg_support, avg_t) && fgeq(p.min_support, min_t)) smallest = percent;
        }
        p.isHQ = goodCov(float(p.min_coverage)) && fleq(float(smallest), pct * 0.5f);
    }
    return p;
}

// ---------------------------------------------------------------- learned seam
//
// The classifier seam of correctAnchorClf (include/corrector.hpp:1448-1472): HQ anchors
// take the consensus; in the other anchors every anchor column whose original base
// differs from the consensus is decided by a classifier. Here the classifier is an MLP
// 16 -> 64 -> 64 -> 1 (ReLU, ReLU, logit) in place of the forest (include/forest.hpp),
// and a logit inside (-ln 9, ln 9) - confidence below 0.9 - falls back to the classic
// column decision of getCorrectedAnchor (src/msa.cpp:379-453). This part has no
// counterpart in the reference beyond the seam itself.

constexpr int kFeat = 16, kHid = 64;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kHid * kFeat, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid,
              kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kHid, kMlpParams = kOffB3 + 1;
constexpr float kConfLogit = 2.1972245773f;   // ln 9: sigmoid(z) >= 0.9  <=>  z >= ln 9

// float -> bfloat16 -> float, round to nearest even (finite inputs)
inline float bf16r(float x){
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

// Inference as the device runs it: weights and activations rounded to bfloat16, float
// accumulation in a fixed order (bias first, then inputs in index order), biases float.
float mlpLogitBf16(const float* P, const float* f){
    float x[kFeat], h1[kHid], h2[kHid];
    for(int k = 0; k < kFeat; k++) x[k] = bf16r(f[k]);
    for(int j = 0; j < kHid; j++){
        float a = P[kOffB1 + j];
        for(int k = 0; k < kFeat; k++) a = a + bf16r(P[kOffW1 + j * kFeat + k]) * x[k];
        h1[j] = bf16r(a > 0.0f ? a : 0.0f);
    }
    for(int j = 0; j < kHid; j++){
        float a = P[kOffB2 + j];
        for(int k = 0; k < kHid; k++) a = a + bf16r(P[kOffW2 + j * kHid + k]) * h1[k];
        h2[j] = bf16r(a > 0.0f ? a : 0.0f);
    }
    float z = P[kOffB3];
    for(int j = 0; j < kHid; j++) z = z + bf16r(P[kOffW3 + j]) * h2[j];
    return z;
}

struct Properties;

// The 16 features of anchor column i: the 13 of extract_anchor_v3
// (include/classification.hpp:748-797), then the anchor base's quality weight (1 without
// qualities), the relative position i/La and the smaller support of the neighbouring
// anchor columns (1 past either end).
void columnFeatures(const Msa& msa, float avgSupport, float minSupport, int minCov, int maxCov, int i, int La,
                    int ob, float qw, float estCov, float* f){
    const int col = msa.ab + i;
    const float* w = msa.weights.data() + size_t(col) * 4;
    const int* c = msa.counts.data() + size_t(col) * 4;
    const float wsum = w[0] + w[1] + w[2] + w[3];
    const int cn = baseIndexOfChar(msa.consensus[col]);
    const int cov = msa.coverage[col];
    const int origCov = c[ob];
    const float origW = w[ob];
    f[0] = float(origCov) / float(cov);
    f[1] = origW / wsum;
    f[2] = origW / float(origCov);
    f[3] = float(c[cn]) / float(cov);
    f[4] = msa.support[col];
    f[5] = w[cn] / float(c[cn]);
    f[6] = wsum / float(cov);
    f[7] = float(cov) / estCov;
    f[8] = wsum / estCov;
    f[9] = avgSupport;
    f[10] = minSupport;
    f[11] = float(maxCov) / estCov;
    f[12] = float(minCov) / estCov;
    f[13] = qw;
    f[14] = float(i) / float(La);
    const float left = i > 0 ? msa.support[col - 1] : 1.0f;
    const float right = i + 1 < La ? msa.support[col + 1] : 1.0f;
    f[15] = left < right ? left : right;
}

// Decision columns collected for training and parity checks.
struct FeatureSink {
    std::vector<float> feats;       // 16 per column
    std::vector<uint32_t> meta;     // read id, position, orig | cons << 2 | classic << 4 | allGood << 6, 0
};

struct Table {
    std::vector<uint64_t> keys;      // sorted, only buckets that survive the 2..threshold rule
    std::vector<uint32_t> offsets;   // keys.size()+1
    std::vector<uint32_t> values;    // read ids, ascending within a bucket
};

// Paired-candidate flags of two mates' sorted candidate lists
// (include/util.hpp:1016-1141 via :1153-1200).
void flagPairs(const std::vector<uint32_t>& c1, const std::vector<uint32_t>& c2,
               std::vector<uint8_t>& f1, std::vector<uint8_t>& f2){
    f1.assign(c1.size(), 0);
    f2.assign(c2.size(), 0);
    size_t i = 0, j = 0;
    const size_t n1 = c1.size(), n2 = c2.size();
    while(i < n1 && j < n2){
        const int e1 = (i + 1 < n1 && c1[i] / 2 == c1[i + 1] / 2) ? 2 : 1;
        const int e2 = (j + 1 < n2 && c2[j] / 2 == c2[j + 1] / 2) ? 2 : 1;
        const bool same = c1[i] / 2 == c2[j] / 2;
        if(e1 == 1 && e2 == 1){
            if(same){
                if(c1[i] != c2[j]){ f1[i] = 1; f2[j] = 1; }
                i++; j++;
            }else{
                if(c1[i] < c2[j]) i++; else j++;
            }
        }else if(e1 == 2 && e2 == 2){
            if(same){
                f1[i] = 1; f2[j] = 1; f1[i + 1] = 1; f2[j + 1] = 1;
                i += 2; j += 2;
            }else{
                if(c1[i] < c2[j]) i += 2; else j += 2;
            }
        }else if(e1 == 2 && e2 == 1){
            if(same){
                if(c1[i] == c2[j]){
                    i++;
                    f1[i] = 1; f2[j] = 1; i++; j++;
                }else{
                    f1[i] = 1; f2[j] = 1; i++; j++; i++;
                }
            }else{
                if(c1[i] < c2[j]) i += 2; else j++;
            }
        }else{
            if(same){
                if(c1[i] == c2[j]){
                    j++;
                    f1[i] = 1; f2[j] = 1; i++; j++;
                }else{
                    f1[i] = 1; f2[j] = 1; i++; j++; j++;
                }
            }else{
                if(c1[i] < c2[j]) i++; else j += 2;
            }
        }
    }
}

}  // namespace

struct carecpu_ctx {
    carecpu_opts o;
    uint64_t n = 0;
    size_t pitch = 0;
    size_t qpitch = 0;
    int maxLen = 0;
    int threshold = 0;
    std::vector<uint32_t> enc;
    std::vector<int> lens;
    std::vector<char> quals;
    std::vector<uint8_t> ambig;
    std::vector<uint64_t> sigs;     // [read][map]
    std::vector<uint16_t> seeds;    // [read][map], seeds of the signatures
    bool seeded = false;            // seeded alignment instead of the shift scan
    std::vector<Table> tables;
    std::vector<uint8_t> records;
    std::vector<float> mlp;         // classifier parameters (empty: classic decision)
    bool correctCandidates = false; // --candidateCorrection
    int newColumns = 15;            // --candidateCorrectionNewColumns
    double seconds = 0.0;

    const uint32_t* row(uint32_t id) const { return enc.data() + pitch * id; }
    const char* qrow(uint32_t id) const { return quals.empty() ? nullptr : quals.data() + qpitch * id; }

    std::vector<uint8_t> decodeBases(uint32_t id) const {
        const int L = lens[id];
        std::vector<uint8_t> b(L);
        const uint32_t* r = row(id);
        for(int i = 0; i < L; i++) b[i] = uint8_t(baseAt(r, i));
        return b;
    }

    // Query of one map (include/ordinaryminhasher.hpp:437-449 and
    // include/cpuhashtable.hpp:576-597): a key maps to its bucket of ascending read
    // ids when 2 <= size <= threshold, else to nothing.
    int bucket(int map, uint64_t key, const uint32_t** values) const {
        const Table& t = tables[map];
        auto it = std::lower_bound(t.keys.begin(), t.keys.end(), key);
        if(it == t.keys.end() || *it != key){ *values = nullptr; return 0; }
        const size_t k = size_t(it - t.keys.begin());
        *values = t.values.data() + t.offsets[k];
        return int(t.offsets[k + 1] - t.offsets[k]);
    }

    // determineCandidateReadIds (include/corrector.hpp:571-650) with
    // determineNumValues/retrieveValues (include/ordinaryminhasher.hpp:135-221).
    std::vector<uint32_t> candidates(uint32_t a) const {
        std::vector<uint32_t> ids;
        if(o.exclude_ambiguous && ambig[a]) return ids;
        if(lens[a] >= o.kmerlength){
            for(int m = 0; m < o.num_hash_maps; m++){
                const uint32_t* v;
                const int cnt = bucket(m, sigs[size_t(a) * o.num_hash_maps + m], &v);
                if(cnt > threshold) continue;
                ids.insert(ids.end(), v, v + cnt);
            }
        }
        std::sort(ids.begin(), ids.end());
        ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
        auto it = std::lower_bound(ids.begin(), ids.end(), a);
        if(it != ids.end() && *it == a) ids.erase(it);
        if(o.exclude_ambiguous){
            ids.erase(std::remove_if(ids.begin(), ids.end(), [&](uint32_t r){ return ambig[r] != 0; }), ids.end());
        }
        return ids;
    }

    // Seeded alignment of candidate c against anchor a (an addition, not in the
    // reference): the first map whose kept bucket (1..threshold values) holds both reads
    // places their shared canonical k-mer; equal seed strands make the candidate forward
    // at shift pa - pc, different strands its reverse complement at pa - (Lc - k - pc).
    // The alignment at that shift is kept (flag 1 or 2) when overlap >= the minimum
    // overlap, mismatches < int(float(overlap) * maxErrorRate) and
    // mismatches / overlap < estimatedErrorrate * 4 (the reference's bounds,
    // src/cpu_alignment.cpp:32-497, include/alignmentorientation.hpp:14-92); flag 3
    // otherwise. `map` is the map used (-1: none).
    int seededAlignment(uint32_t a, uint32_t c, const std::vector<uint8_t>& anchor, const std::vector<uint8_t>& fwd,
                        const std::vector<uint8_t>& rc, Alignment& out, int& map) const {
        const int M = o.num_hash_maps;
        const int La = lens[a], Lc = lens[c];
        map = -1;
        for(int m = 0; m < M; m++){
            const uint64_t key = sigs[size_t(a) * M + m];
            const uint32_t* v;
            const int cnt = bucket(m, key, &v);
            if(cnt <= 0 || cnt > threshold) continue;
            if(sigs[size_t(c) * M + m] == key){ map = m; break; }
        }
        out = Alignment{};
        if(map < 0) return 3;
        const int sa = seeds[size_t(a) * M + map], sc = seeds[size_t(c) * M + map];
        const int pa = sa >> 1, pc = sc >> 1;
        const bool same = (sa & 1) == (sc & 1);
        const int shift = same ? pa - pc : pa - (Lc - o.kmerlength - pc);
        const std::vector<uint8_t>& C = same ? fwd : rc;
        const int cb = std::max(-shift, 0);
        const int ce = std::min(Lc, La - shift);
        const int ov = ce - cb;
        int mis = 0;
        for(int j = cb; j < ce; j++) mis += (C[j] != anchor[j + shift]);
        out.shift = shift;
        out.overlap = ov;
        out.nOps = mis;
        out.score = mis + La + Lc - 2 * ov;
        out.valid = true;
        const int mo = std::max(o.min_overlap, int(float(La) * o.min_overlap_ratio));
        if(ov > 0 && ov >= mo && mis < int(float(ov) * o.max_error_rate)
           && float(mis) / float(ov) < o.estimated_errorrate * 4.0f)
            return same ? 1 : 2;
        return 3;
    }

    // One anchor through CpuErrorCorrector::process (include/corrector.hpp:145-312),
    // given its candidate ids and (PE) pair flags. Appends its record to `out`.
    // One anchor's candidate corrections go to `cands` as (candidate id, record) pairs.
    // Returns whether the anchor was corrected as an HQ anchor.
    bool processAnchor(uint32_t a, const std::vector<uint32_t>& cids, const std::vector<uint8_t>& pairFlags,
                       std::vector<uint8_t>& out, carecpu_trace* tr, FeatureSink* sink = nullptr,
                       std::vector<std::pair<uint32_t, std::vector<uint8_t>>>* cands = nullptr) const;
};

// Serialized record (include/correctedsequence.hpp:299-387): read id, flags
// (hq << 31 | useEdits << 30 | candidate << 29 | body bytes), then either the edits
// (count, positions, bases) or the sequence (length, bases), then a candidate's shift.
void appendRecord(std::vector<uint8_t>& out, uint32_t id, bool hq, bool candidate, int shift,
                  const std::string& corrected, const std::string& original, bool ambiguous){
    const int L = int(corrected.size());
    std::vector<std::pair<int, char>> edits;
    bool useEdits = false;
    if(!ambiguous){
        const int maxEdits = L / 7;
        int count = 0;
        for(int i = 0; i < L && count <= maxEdits; i++){
            if(corrected[i] != original[i]){
                edits.emplace_back(i, corrected[i]);
                count++;
            }
        }
        useEdits = count <= maxEdits;
    }
    // HQ records without edits are not stored (src/correct_cpu.cpp:263-273)
    if(hq && useEdits && edits.empty()) return;
    uint32_t numBytes = useEdits ? uint32_t(4 + edits.size() * 5) : uint32_t(4 + L);
    if(candidate) numBytes += 4;
    const uint32_t flags = (uint32_t(hq) << 31) | (uint32_t(useEdits) << 30) | (uint32_t(candidate) << 29) | numBytes;
    const size_t start = out.size();
    out.resize(start + 8 + numBytes);
    uint8_t* ptr = out.data() + start;
    std::memcpy(ptr, &id, 4);
    std::memcpy(ptr + 4, &flags, 4);
    ptr += 8;
    if(useEdits){
        const int ne = int(edits.size());
        std::memcpy(ptr, &ne, 4);
        ptr += 4;
        for(const auto& e : edits){ std::memcpy(ptr, &e.first, 4); ptr += 4; }
        for(const auto& e : edits){ *ptr++ = uint8_t(e.second); }
    }else{
        std::memcpy(ptr, &L, 4);
        ptr += 4;
        std::memcpy(ptr, corrected.data(), L);
        ptr += L;
    }
    if(candidate) std::memcpy(ptr, &shift, 4);
}

bool carecpu_ctx::processAnchor(uint32_t a, const std::vector<uint32_t>& cids, const std::vector<uint8_t>& pairFlags,
                                std::vector<uint8_t>& out, carecpu_trace* tr, FeatureSink* sink,
                                std::vector<std::pair<uint32_t, std::vector<uint8_t>>>* cands) const {
    const int La = lens[a];
    const std::vector<uint8_t> anchor = decodeBases(a);
    const bool useQ = o.use_quality_scores != 0;
    const int nc0 = int(cids.size());

    if(tr){
        tr->num_candidates = nc0;
        tr->num_filtered = 0;
        tr->num_final = 0;
        tr->refinement_rounds = 0;
        tr->num_columns = 0;
        tr->is_corrected = 0;
        tr->is_hq = 0;
        for(int i = 0; i < nc0 && i < tr->cap_candidates; i++){
            if(tr->candidate_ids) tr->candidate_ids[i] = cids[i];
            if(tr->is_paired) tr->is_paired[i] = pairFlags.empty() ? 0 : pairFlags[i];
        }
    }
    if(nc0 == 0) return false;

    // alignments in both orientations and orientation choice
    // (include/corrector.hpp:925-979)
    std::vector<CandRow> rows;
    rows.reserve(nc0);
    std::vector<uint8_t> fwd, rc;
    for(int i = 0; i < nc0; i++){
        const uint32_t id = cids[i];
        const int Lc = lens[id];
        fwd = decodeBases(id);
        rc.resize(Lc);
        for(int j = 0; j < Lc; j++) rc[j] = uint8_t(3 - fwd[Lc - 1 - j]);
        Alignment af, ar;
        int flag;
        if(seeded){
            int map;
            flag = seededAlignment(a, id, anchor, fwd, rc, af, map);
            ar = af;
        }else{
            af = align(anchor.data(), La, fwd.data(), Lc, o.min_overlap, o.max_error_rate, o.min_overlap_ratio);
            ar = align(anchor.data(), La, rc.data(), Lc, o.min_overlap, o.max_error_rate, o.min_overlap_ratio);
            flag = orientation(af, ar, La, o.min_overlap_ratio, o.min_overlap, o.estimated_errorrate);
        }
        if(tr && i < tr->cap_candidates){
            auto put = [](int* dst, const Alignment& x){
                dst[0] = x.score; dst[1] = x.overlap; dst[2] = x.shift; dst[3] = x.nOps; dst[4] = x.valid ? 1 : 0;
            };
            if(tr->fwd_alignments) put(tr->fwd_alignments + 5 * i, af);
            if(tr->rc_alignments) put(tr->rc_alignments + 5 * i, ar);
            if(tr->orientation) tr->orientation[i] = int8_t(flag);
        }
        // filterCandidatesByAlignmentFlag (include/corrector.hpp:982-1050)
        if(flag == 3) continue;
        CandRow r;
        r.id = id;
        r.length = Lc;
        const Alignment& al = flag == 1 ? af : ar;
        r.shift = al.shift;
        r.nOps = al.nOps;
        r.overlap = al.overlap;
        r.orient = flag;
        r.paired = pairFlags.empty() ? false : pairFlags[i] != 0;
        r.bases = flag == 1 ? fwd : rc;
        r.qual = useQ ? qrow(id) : nullptr;
        r.reverseQual = (flag == 2);
        rows.push_back(std::move(r));
    }
    if(rows.empty()) return false;

    // filterCandidatesByAlignmentMismatchRatio (include/corrector.hpp:1053-1149) or,
    // for paired input, ...WithPairFlags (:1151-1211)
    std::vector<CandRow> kept;
    kept.reserve(rows.size());
    if(o.paired){
        const float threshold = o.paired_filter_threshold;
        for(auto& r : rows){
            if(r.paired){ kept.push_back(std::move(r)); continue; }
            const float ratio = float(r.nOps) / float(r.overlap);
            if(ratio < threshold) kept.push_back(std::move(r));
        }
    }else{
        const float base = o.estimated_errorrate * 1.0f;
        const float goodCount = o.estimated_coverage * o.m_coverage;
        int counts[3] = {0, 0, 0};
        for(const auto& r : rows){
            const float ratio = float(r.nOps) / float(r.overlap);
            if(ratio < 2 * base) counts[0]++;
            if(ratio < 3 * base) counts[1]++;
            if(ratio < 4 * base) counts[2]++;
        }
        float threshold = std::numeric_limits<float>::lowest();
        if(counts[0] > 0 || counts[1] > 0 || counts[2] > 0){
            if(counts[0] >= goodCount) threshold = 2 * base;
            else if(counts[1] >= goodCount) threshold = 3 * base;
            else if(counts[2] >= goodCount) threshold = 4 * base;
        }
        for(auto& r : rows){
            const float ratio = float(r.nOps) / float(r.overlap);
            if(ratio < threshold) kept.push_back(std::move(r));
        }
    }
    if(tr) tr->num_filtered = int(kept.size());
    if(kept.empty()) return false;

    // overlap weights (include/corrector.hpp:1260-1280)
    for(auto& r : kept) r.weight = overlapWeight(r.nOps, r.overlap, o.max_error_rate);

    const char* aq = useQ ? qrow(a) : nullptr;
    Msa msa;
    buildMsa(msa, anchor, aq, useQ, int(qpitch), kept);

    // refineMSA (include/corrector.hpp:1307-1421): at most 5 rounds. A round that
    // removes nothing leaves the same MSA, so every later round would too.
    std::vector<char> remove;
    for(int it = 0; it < 5; it++){
        if(!findDifferentRegion(msa, anchor, kept, int(o.estimated_coverage), remove)) break;
        size_t w = 0;
        for(size_t k = 0; k < kept.size(); k++){
            if(!remove[k]){
                if(w != k) kept[w] = std::move(kept[k]);
                w++;
            }
        }
        if(w == kept.size()) break;
        kept.resize(w);
        if(tr) tr->refinement_rounds++;
        buildMsa(msa, anchor, aq, useQ, int(qpitch), kept);
    }

    if(tr){
        tr->num_final = int(kept.size());
        for(int i = 0; i < int(kept.size()) && i < tr->cap_candidates; i++){
            if(tr->final_ids) tr->final_ids[i] = kept[i].id;
            if(tr->final_shifts) tr->final_shifts[i] = kept[i].shift;
            if(tr->final_weights) tr->final_weights[i] = kept[i].weight;
        }
        tr->num_columns = msa.ncol;
        tr->anchor_column_begin = msa.ab;
        for(int c = 0; c < msa.ncol && c < tr->cap_columns; c++){
            for(int b = 0; b < 4; b++){
                if(tr->counts) tr->counts[4 * c + b] = msa.counts[size_t(c) * 4 + b];
                if(tr->weights) tr->weights[4 * c + b] = msa.weights[size_t(c) * 4 + b];
            }
            if(tr->coverage) tr->coverage[c] = msa.coverage[c];
            if(tr->consensus) tr->consensus[c] = msa.consensus[c];
            if(tr->support) tr->support[c] = msa.support[c];
            if(tr->orig_weights) tr->orig_weights[c] = msa.origWeights[c];
            if(tr->orig_coverage) tr->orig_coverage[c] = msa.origCoverages[c];
        }
    }

    // correctAnchorClassic (include/corrector.hpp:1423-1446): getMSAProperties and
    // getCorrectedAnchor (src/msa.cpp:304-453)
    if(msa.ncand == 0) return false;   // cannot be corrected without candidates
    const Properties p = msaProperties(msa, o.estimated_errorrate, o.estimated_coverage, o.m_coverage);
    std::string corrected(La, 'A');
    if(p.allGood){
        for(int i = 0; i < La; i++) corrected[i] = msa.consensus[msa.ab + i];
    }else{
        for(int i = 0; i < La; i++){
            const int col = msa.ab + i;
            if(msa.support[col] > 0.90f && msa.origCoverages[col] <= 2) corrected[i] = msa.consensus[col];
            else corrected[i] = kDecode[anchor[i]];
        }
    }
    if((!mlp.empty() || sink) && !p.isHQ){
        for(int i = 0; i < La; i++){
            const int col = msa.ab + i;
            const int ob = anchor[i];
            const int cn = baseIndexOfChar(msa.consensus[col]);
            if(ob == cn) continue;
            const int classic = baseIndexOfChar(corrected[i]);
            const float qw = useQ ? qualityWeights().w[uint8_t(aq[i])] : 1.0f;
            float f[kFeat];
            columnFeatures(msa, p.avg_support, p.min_support, p.min_coverage, p.max_coverage, i, La, ob, qw,
                           o.estimated_coverage, f);
            if(sink){
                sink->feats.insert(sink->feats.end(), f, f + kFeat);
                const uint32_t m[4] = {a, uint32_t(i), uint32_t(ob | (cn << 2) | (classic << 4) | (int(p.allGood) << 6)), 0u};
                sink->meta.insert(sink->meta.end(), m, m + 4);
            }
            if(!mlp.empty()){
                const float z = mlpLogitBf16(mlp.data(), f);
                const int d = z >= kConfLogit ? cn : (z <= -kConfLogit ? ob : classic);
                corrected[i] = kDecode[d];
            }
        }
    }
    if(tr){
        tr->avg_support = p.avg_support;
        tr->min_support = p.min_support;
        tr->min_coverage = p.min_coverage;
        tr->max_coverage = p.max_coverage;
        tr->is_hq = p.isHQ ? 1 : 0;
        tr->is_corrected = 1;
        if(tr->corrected) std::memcpy(tr->corrected, corrected.data(), La);
    }

    // correctCandidatesClassic (include/corrector.hpp:1504-1512, src/msa.cpp:457-522): in an
    // HQ anchor, every final candidate that starts within newColumns columns of the anchor
    // and ends at most newColumns columns after it takes the consensus over its columns
    // (the support/coverage test there is overwritten with true); reverse-complement
    // candidates are turned forward (include/corrector.hpp:1684-1700).
    if(cands && correctCandidates && p.isHQ){
        for(const auto& r : kept){
            const int qb = msa.ab + r.shift;
            const int qe = qb + r.length;
            if(!(msa.ab - newColumns <= qb && qb <= msa.ab + newColumns && qe <= msa.ab + La + newColumns)) continue;
            std::string corr(msa.consensus.begin() + qb, msa.consensus.begin() + qe);
            if(r.orient == 2){
                std::string fwdc(corr.size(), 'A');
                for(size_t j = 0; j < corr.size(); j++){
                    const char c = corr[corr.size() - 1 - j];
                    fwdc[j] = c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : 'A';
                }
                corr.swap(fwdc);
            }
            const std::vector<uint8_t> ob = decodeBases(r.id);
            std::string orig(ob.size(), 'A');
            for(size_t j = 0; j < ob.size(); j++) orig[j] = kDecode[ob[j]];
            std::vector<uint8_t> rec;
            appendRecord(rec, r.id, false, true, r.shift, corr, orig, ambig[r.id] != 0);
            cands->emplace_back(r.id, std::move(rec));
        }
    }

    // makeOutputOfTask (include/corrector.hpp:1639-1761) and the record encoding
    // (include/correctedsequence.hpp:299-387)
    std::string original(La, 'A');
    for(int i = 0; i < La; i++) original[i] = kDecode[anchor[i]];
    appendRecord(out, a, p.isHQ, false, 0, corrected, original, ambig[a] != 0);
    return p.isHQ;
}

extern "C" {

int carecpu_results_per_map_threshold(float coverage){ return resultsPerMapThreshold(coverage); }

void carecpu_signature(const uint32_t* enc, int len, int k, int nmaps, uint64_t* out){
    signature(enc, len, k, nmaps, out);
}

carecpu_ctx* carecpu_create(const carecpu_opts* opts, uint64_t num_reads, const uint32_t* enc2bit, size_t pitch_ints,
                            const int* lengths, const char* quals, size_t qual_pitch, const uint8_t* ambiguous,
                            int threads){
    if(!opts || opts->num_hash_maps < 1 || opts->num_hash_maps > kMaxMaps) return nullptr;
    if(opts->kmerlength < 1 || opts->kmerlength > 32) return nullptr;
    auto* c = new carecpu_ctx;
    c->o = *opts;
    c->n = num_reads;
    c->pitch = pitch_ints;
    c->enc.assign(enc2bit, enc2bit + num_reads * pitch_ints);
    c->lens.assign(lengths, lengths + num_reads);
    c->maxLen = num_reads ? *std::max_element(lengths, lengths + num_reads) : 0;
    if(quals){
        c->qpitch = qual_pitch;
        c->quals.assign(quals, quals + num_reads * qual_pitch);
    }
    c->ambig.assign(num_reads, 0);
    if(ambiguous) std::copy(ambiguous, ambiguous + num_reads, c->ambig.begin());
    c->threshold = resultsPerMapThreshold(opts->estimated_coverage);

    const int M = opts->num_hash_maps;
    c->sigs.resize(num_reads * M);
    c->seeds.resize(num_reads * M);
    if(threads > 0) omp_set_num_threads(threads);
    #pragma omp parallel for schedule(static)
    for(long long r = 0; r < (long long)num_reads; r++){
        signature(c->row(uint32_t(r)), c->lens[r], opts->kmerlength, M, c->sigs.data() + size_t(r) * M,
                  c->seeds.data() + size_t(r) * M);
    }

    // Tables: group (signature, read id) pairs by key with read ids ascending, and keep
    // buckets of 2..threshold values (include/groupbykey.hpp:99-250,
    // include/config.hpp:32-34, include/ordinaryminhasher.hpp:223-254).
    c->tables.resize(M);
    #pragma omp parallel for schedule(dynamic, 1)
    for(int m = 0; m < M; m++){
        std::vector<std::pair<uint64_t, uint32_t>> kv(num_reads);
        for(uint64_t r = 0; r < num_reads; r++) kv[r] = {c->sigs[r * M + m], uint32_t(r)};
        std::sort(kv.begin(), kv.end());
        Table& t = c->tables[m];
        t.offsets.push_back(0);
        size_t i = 0;
        while(i < kv.size()){
            size_t j = i;
            while(j < kv.size() && kv[j].first == kv[i].first) j++;
            const size_t cnt = j - i;
            if(cnt >= 2 && cnt <= size_t(c->threshold)){
                t.keys.push_back(kv[i].first);
                for(size_t q = i; q < j; q++) t.values.push_back(kv[q].second);
                t.offsets.push_back(uint32_t(t.values.size()));
            }
            i = j;
        }
    }
    return c;
}

void carecpu_destroy(carecpu_ctx* ctx){ delete ctx; }

void carecpu_signatures(carecpu_ctx* ctx, uint64_t first, uint64_t n, uint64_t* out){
    const int M = ctx->o.num_hash_maps;
    std::memcpy(out, ctx->sigs.data() + first * M, n * M * sizeof(uint64_t));
}

int carecpu_bucket(carecpu_ctx* ctx, int map, uint64_t key, const uint32_t** values){
    return ctx->bucket(map, key, values);
}

int carecpu_candidates(carecpu_ctx* ctx, uint32_t read_id, uint32_t* out, int cap){
    const auto ids = ctx->candidates(read_id);
    for(int i = 0; i < int(ids.size()) && i < cap; i++) out[i] = ids[i];
    return int(ids.size());
}

void carecpu_align(const uint32_t* anchor, int la, const uint32_t* cand, int lc, int min_overlap,
                   float max_error_rate, float min_overlap_ratio, int* out5){
    std::vector<uint8_t> A(la), C(lc);
    for(int i = 0; i < la; i++) A[i] = uint8_t(baseAt(anchor, i));
    for(int i = 0; i < lc; i++) C[i] = uint8_t(baseAt(cand, i));
    const Alignment r = align(A.data(), la, C.data(), lc, min_overlap, max_error_rate, min_overlap_ratio);
    out5[0] = r.score; out5[1] = r.overlap; out5[2] = r.shift; out5[3] = r.nOps; out5[4] = r.valid ? 1 : 0;
}

int carecpu_orientation(const int* f5, const int* r5, int la, int /*lc*/, float min_overlap_ratio, int min_overlap,
                        float estimated_errorrate){
    Alignment f, r;
    f.overlap = f5[1]; f.nOps = f5[3]; f.valid = f5[4] != 0;
    r.overlap = r5[1]; r.nOps = r5[3]; r.valid = r5[4] != 0;
    return orientation(f, r, la, min_overlap_ratio, min_overlap, estimated_errorrate);
}

int carecpu_trace_anchor(carecpu_ctx* ctx, uint32_t read_id, carecpu_trace* trace){
    if(read_id >= ctx->n) return 1;
    std::vector<uint8_t> out;
    const auto ids = ctx->candidates(read_id);
    std::vector<uint8_t> flags;
    if(ctx->o.paired){
        const uint32_t mate = read_id ^ 1u;
        if(mate >= ctx->n) return 2;
        const auto mids = ctx->candidates(mate);
        std::vector<uint8_t> mflags;
        if((read_id & 1u) == 0) flagPairs(ids, mids, flags, mflags);
        else flagPairs(mids, ids, mflags, flags);
    }
    ctx->processAnchor(read_id, ids, flags, out, trace);
    return 0;
}

void carecpu_flag_pairs(const uint32_t* c1, int n1, const uint32_t* c2, int n2, uint8_t* f1, uint8_t* f2){
    std::vector<uint32_t> a(c1, c1 + n1), b(c2, c2 + n2);
    std::vector<uint8_t> fa, fb;
    flagPairs(a, b, fa, fb);
    std::copy(fa.begin(), fa.end(), f1);
    std::copy(fb.begin(), fb.end(), f2);
}

size_t carecpu_correct(carecpu_ctx* ctx, uint64_t first, uint64_t n, int threads){
    if(threads > 0) omp_set_num_threads(threads);
    const bool paired = ctx->o.paired != 0;
    if(first + n > ctx->n) n = ctx->n > first ? ctx->n - first : 0;
    // units: single anchors, or mate pairs for paired input
    if(paired && ((first & 1) || (n & 1))) return 0;
    const long long units = paired ? (long long)(n / 2) : (long long)n;
    const bool withCands = ctx->correctCandidates;
    std::vector<std::vector<uint8_t>> outs(n);
    std::vector<std::vector<std::pair<uint32_t, std::vector<uint8_t>>>> cands(withCands ? n : 0);
    std::vector<uint8_t> hq(n, 0);
    const auto t0 = std::chrono::steady_clock::now();
    #pragma omp parallel for schedule(dynamic, 16)
    for(long long u = 0; u < units; u++){
        if(paired){
            const uint32_t a0 = uint32_t(first + 2 * u);
            const uint32_t a1 = a0 + 1;
            const auto c0 = ctx->candidates(a0);
            const auto c1 = ctx->candidates(a1);
            std::vector<uint8_t> f0, f1;
            flagPairs(c0, c1, f0, f1);
            hq[2 * u] = ctx->processAnchor(a0, c0, f0, outs[2 * u], nullptr, nullptr, withCands ? &cands[2 * u] : nullptr);
            hq[2 * u + 1] = ctx->processAnchor(a1, c1, f1, outs[2 * u + 1], nullptr, nullptr,
                                               withCands ? &cands[2 * u + 1] : nullptr);
        }else{
            const uint32_t a = uint32_t(first + u);
            const auto c = ctx->candidates(a);
            hq[u] = ctx->processAnchor(a, c, std::vector<uint8_t>{}, outs[u], nullptr, nullptr,
                                       withCands ? &cands[u] : nullptr);
        }
    }
    const auto t1 = std::chrono::steady_clock::now();
    ctx->seconds = std::chrono::duration<double>(t1 - t0).count();
    ctx->records.clear();
    size_t count = 0;
    auto countIn = [](const std::vector<uint8_t>& v){
        size_t p = 0, c = 0;
        while(p < v.size()){
            uint32_t flags;
            std::memcpy(&flags, v.data() + p + 4, 4);
            p += 8 + (flags & ((1u << 29) - 1));
            c++;
        }
        return c;
    };
    if(!withCands){
        for(auto& v : outs){
            count += countIn(v);
            ctx->records.insert(ctx->records.end(), v.begin(), v.end());
        }
        return count;
    }
    // The reference's records in the order a one-thread run inserts them
    // (src/correct_cpu.cpp:166-281: per 16-read batch, the anchors' records, then the
    // candidates' records in anchor order), with a candidate's record dropped when the
    // candidate was corrected as an HQ anchor before (include/corrector.hpp:1676-1680,
    // flags set at :284-290: reads below the anchor, in this range).
    std::vector<std::pair<uint32_t, const std::vector<uint8_t>*>> stream;
    for(uint64_t s0 = first; s0 < first + n;){
        const uint64_t e0 = std::min<uint64_t>(first + n, (s0 / 16 + 1) * 16);
        for(uint64_t a = s0; a < e0; a++){
            if(!outs[a - first].empty()) stream.emplace_back(uint32_t(a), &outs[a - first]);
        }
        for(uint64_t a = s0; a < e0; a++){
            for(const auto& cr : cands[a - first]){
                const uint32_t c = cr.first;
                if(c < a && c >= first && hq[c - first]) continue;
                if(!cr.second.empty()) stream.emplace_back(c, &cr.second);
            }
        }
        s0 = e0;
    }
    for(const auto& e : stream){
        count += countIn(*e.second);
        ctx->records.insert(ctx->records.end(), e.second->begin(), e.second->end());
    }
    return count;
}

void carecpu_set_seeded_alignment(carecpu_ctx* ctx, int on){ ctx->seeded = on != 0; }

void carecpu_seeds(carecpu_ctx* ctx, uint64_t first, uint64_t n, uint16_t* out){
    const int M = ctx->o.num_hash_maps;
    std::memcpy(out, ctx->seeds.data() + first * M, n * M * sizeof(uint16_t));
}

// Seeded alignment of candidate c against anchor a: out = {map, flag, shift, overlap, nOps}.
void carecpu_seeded_alignment(carecpu_ctx* ctx, uint32_t a, uint32_t c, int* out){
    const std::vector<uint8_t> anchor = ctx->decodeBases(a);
    const std::vector<uint8_t> fwd = ctx->decodeBases(c);
    std::vector<uint8_t> rc(fwd.size());
    for(size_t j = 0; j < fwd.size(); j++) rc[j] = uint8_t(3 - fwd[fwd.size() - 1 - j]);
    Alignment al;
    int map;
    const int flag = ctx->seededAlignment(a, c, anchor, fwd, rc, al, map);
    out[0] = map; out[1] = flag; out[2] = al.shift; out[3] = al.overlap; out[4] = al.nOps;
}

void carecpu_set_candidate_correction(carecpu_ctx* ctx, int on, int new_columns){
    ctx->correctCandidates = on != 0;
    ctx->newColumns = new_columns;
}

int carecpu_set_classifier(carecpu_ctx* ctx, const float* params, int count){
    if(!params){ ctx->mlp.clear(); return 0; }
    if(count != kMlpParams) return -1;
    ctx->mlp.assign(params, params + count);
    return 0;
}

size_t carecpu_features(carecpu_ctx* ctx, uint64_t first, uint64_t n, int threads, float* feats, uint32_t* meta,
                        size_t cap){
    if(threads > 0) omp_set_num_threads(threads);
    const bool paired = ctx->o.paired != 0;
    if(first + n > ctx->n) n = ctx->n > first ? ctx->n - first : 0;
    if(paired && ((first & 1) || (n & 1))) return 0;
    const long long units = paired ? (long long)(n / 2) : (long long)n;
    std::vector<FeatureSink> sinks(units);
    #pragma omp parallel for schedule(dynamic, 16)
    for(long long u = 0; u < units; u++){
        std::vector<uint8_t> rec;
        if(paired){
            const uint32_t a0 = uint32_t(first + 2 * u);
            const auto c0 = ctx->candidates(a0);
            const auto c1 = ctx->candidates(a0 + 1);
            std::vector<uint8_t> f0, f1;
            flagPairs(c0, c1, f0, f1);
            ctx->processAnchor(a0, c0, f0, rec, nullptr, &sinks[u]);
            ctx->processAnchor(a0 + 1, c1, f1, rec, nullptr, &sinks[u]);
        }else{
            const uint32_t a = uint32_t(first + u);
            ctx->processAnchor(a, ctx->candidates(a), std::vector<uint8_t>{}, rec, nullptr, &sinks[u]);
        }
    }
    size_t total = 0;
    for(const auto& sk : sinks){
        const size_t m = sk.meta.size() / 4;
        for(size_t q = 0; q < m; q++, total++){
            if(total >= cap) continue;
            std::memcpy(feats + total * kFeat, sk.feats.data() + q * kFeat, kFeat * sizeof(float));
            std::memcpy(meta + total * 4, sk.meta.data() + q * 4, 4 * sizeof(uint32_t));
        }
    }
    return total;
}

int carecpu_mlp_num_params(void){ return kMlpParams; }

void carecpu_mlp_logits_bf16(const float* params, const float* X, uint64_t n, float* z){
    for(uint64_t r = 0; r < n; r++) z[r] = mlpLogitBf16(params, X + r * kFeat);
}

// Double-precision forward pass and mean binary cross-entropy on logits, with its exact
// gradient by backpropagation (chain rule through the two ReLU layers).
double carecpu_mlp_loss_grad_f64(const double* P, const float* X, const float* Y, uint64_t n, double* G, double* z){
    if(G) std::fill(G, G + kMlpParams, 0.0);
    double loss = 0.0;
    double h1p[kHid], h1[kHid], h2p[kHid], h2[kHid], dh2[kHid], dh1[kHid];
    for(uint64_t r = 0; r < n; r++){
        const float* x = X + r * kFeat;
        for(int j = 0; j < kHid; j++){
            double a = P[kOffB1 + j];
            for(int k = 0; k < kFeat; k++) a += P[kOffW1 + j * kFeat + k] * double(x[k]);
            h1p[j] = a; h1[j] = a > 0.0 ? a : 0.0;
        }
        for(int j = 0; j < kHid; j++){
            double a = P[kOffB2 + j];
            for(int k = 0; k < kHid; k++) a += P[kOffW2 + j * kHid + k] * h1[k];
            h2p[j] = a; h2[j] = a > 0.0 ? a : 0.0;
        }
        double zz = P[kOffB3];
        for(int j = 0; j < kHid; j++) zz += P[kOffW3 + j] * h2[j];
        if(z) z[r] = zz;
        const double y = Y ? double(Y[r]) : 0.0;
        loss += std::max(zz, 0.0) - y * zz + std::log1p(std::exp(-std::fabs(zz)));
        if(!G) continue;
        const double dz = (1.0 / (1.0 + std::exp(-zz)) - y) / double(n);
        G[kOffB3] += dz;
        for(int j = 0; j < kHid; j++){
            G[kOffW3 + j] += dz * h2[j];
            dh2[j] = h2p[j] > 0.0 ? dz * P[kOffW3 + j] : 0.0;
        }
        for(int k = 0; k < kHid; k++) dh1[k] = 0.0;
        for(int j = 0; j < kHid; j++){
            G[kOffB2 + j] += dh2[j];
            for(int k = 0; k < kHid; k++){
                G[kOffW2 + j * kHid + k] += dh2[j] * h1[k];
                dh1[k] += dh2[j] * P[kOffW2 + j * kHid + k];
            }
        }
        for(int k = 0; k < kHid; k++){
            if(h1p[k] <= 0.0) continue;
            G[kOffB1 + k] += dh1[k];
            for(int m = 0; m < kFeat; m++) G[kOffW1 + k * kFeat + m] += dh1[k] * double(x[m]);
        }
    }
    return n ? loss / double(n) : 0.0;
}

size_t carecpu_records_bytes(carecpu_ctx* ctx){ return ctx->records.size(); }
const uint8_t* carecpu_records_data(carecpu_ctx* ctx){ return ctx->records.data(); }
double carecpu_last_seconds(carecpu_ctx* ctx){ return ctx->seconds; }

}  // extern "C"

// Seeded synthetic read simulator for the CARE correction workloads.
//
// Produces a random genome (optionally with GC bias, duplicated segments and
// repeat families with per-copy divergence), samples single-end or paired-end
// reads from it, draws Phred+33 qualities from a linear profile with Gaussian
// noise, and applies substitutions with probability 10^(-Q/10).  Every output is
// a pure function of the configuration, including the seed, so large inputs are
// regenerated where they are used instead of being shipped.
//
// C-ABI only; see include/care_sim.h.

#include "../../include/care_sim.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace {

// xoshiro256** seeded through splitmix64 (public-domain generators by Blackman/Vigna).
struct Rng {
    std::uint64_t s[4];
    static std::uint64_t splitmix(std::uint64_t& x){
        std::uint64_t z = (x += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    explicit Rng(std::uint64_t seed){
        for(auto& w : s) w = splitmix(seed);
    }
    static std::uint64_t rotl(std::uint64_t x, int k){ return (x << k) | (x >> (64 - k)); }
    std::uint64_t next(){
        const std::uint64_t result = rotl(s[1] * 5, 7) * 9;
        const std::uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return result;
    }
    // uniform in [0,1) with 53 random bits
    double uniform(){ return double(next() >> 11) * (1.0 / 9007199254740992.0); }
    // uniform integer in [0, n)
    std::uint64_t below(std::uint64_t n){ return n ? next() % n : 0; }
    double normal(){
        // Box-Muller, one value per call (keeps the stream position simple).
        double u1 = uniform();
        double u2 = uniform();
        if(u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

const char kBases[4] = {'A', 'C', 'G', 'T'};

char complement(char c){
    switch(c){
        case 'A': return 'T';
        case 'C': return 'G';
        case 'G': return 'C';
        case 'T': return 'A';
        default: return c;
    }
}

char randomBase(Rng& rng, double gc){
    const double u = rng.uniform();
    if(u < gc){
        return (rng.next() & 1) ? 'G' : 'C';
    }
    return (rng.next() & 1) ? 'T' : 'A';
}

char otherBase(Rng& rng, char b){
    char c;
    do { c = kBases[rng.below(4)]; } while(c == b);
    return c;
}

void mutateCopy(Rng& rng, std::string& s, double divergence){
    for(auto& c : s){
        if(rng.uniform() < divergence) c = otherBase(rng, c);
    }
}

std::string makeGenome(const care_sim_config& cfg, Rng& rng){
    const std::size_t G = cfg.genome_length;
    std::string genome(G, 'A');
    for(std::size_t i = 0; i < G; i++) genome[i] = randomBase(rng, cfg.gc_content);

    // Duplicated segments: pick a source segment, write `dup_copies` diverged copies
    // of it at random positions.
    if(cfg.dup_fraction > 0 && cfg.dup_segment_length > 0 && cfg.dup_copies > 0){
        const std::size_t L = cfg.dup_segment_length;
        const std::size_t numSegments = std::size_t(std::llround(cfg.dup_fraction * double(G) / double(L)));
        for(std::size_t s = 0; s < numSegments && G > L; s++){
            const std::size_t src = rng.below(G - L);
            const std::string seg = genome.substr(src, L);
            for(int c = 0; c < cfg.dup_copies; c++){
                std::string copy = seg;
                mutateCopy(rng, copy, cfg.dup_divergence);
                const std::size_t dst = rng.below(G - L);
                genome.replace(dst, L, copy);
            }
        }
    }

    // Repeat families: each family has a random unit; `repeat_copies` diverged copies
    // are written at random positions so that about repeat_fraction of the genome is
    // covered by family members.
    if(cfg.repeat_fraction > 0 && cfg.repeat_unit_length > 0 && cfg.repeat_copies > 0){
        const std::size_t U = cfg.repeat_unit_length;
        const double covered = cfg.repeat_fraction * double(G);
        const std::size_t numFamilies = std::max<std::size_t>(1, std::size_t(std::llround(covered / double(U * cfg.repeat_copies))));
        for(std::size_t f = 0; f < numFamilies && G > U; f++){
            std::string unit(U, 'A');
            for(auto& c : unit) c = randomBase(rng, cfg.gc_content);
            for(int c = 0; c < cfg.repeat_copies; c++){
                std::string copy = unit;
                mutateCopy(rng, copy, cfg.repeat_divergence);
                const std::size_t dst = rng.below(G - U);
                genome.replace(dst, U, copy);
            }
        }
    }
    // Tandem repeats (microsatellites): units of 1..tandem_period_max bases repeated
    // over 40..400 bases, about tandem_fraction of the genome. They make low-complexity
    // reads whose alignments have many equally scored shifts.
    if(cfg.tandem_fraction > 0 && cfg.tandem_period_max > 0){
        double covered = 0.0;
        const double target = cfg.tandem_fraction * double(G);
        while(covered < target){
            const int period = 1 + int(rng.below(std::uint64_t(cfg.tandem_period_max)));
            const std::size_t len = 40 + std::size_t(rng.below(361));
            if(len >= G) break;
            std::string unit(period, 'A');
            for(auto& c : unit) c = randomBase(rng, cfg.gc_content);
            const std::size_t dst = rng.below(G - len);
            for(std::size_t i = 0; i < len; i++) genome[dst + i] = unit[i % period];
            covered += double(len);
        }
    }
    return genome;
}

int drawLength(const care_sim_config& cfg, Rng& rng){
    if(cfg.read_length_max <= cfg.read_length_min) return cfg.read_length_min;
    return cfg.read_length_min + int(rng.below(std::uint64_t(cfg.read_length_max - cfg.read_length_min + 1)));
}

// Quality at position j of a read of length L: mean falls linearly from q_first at
// j=0 to q_last at j=L-1, sd q_sd, rounded and clamped.
char drawQuality(const care_sim_config& cfg, Rng& rng, int j, int L){
    const double t = L > 1 ? double(j) / double(L - 1) : 0.0;
    const double mean = cfg.q_first + (cfg.q_last - cfg.q_first) * t;
    double q = std::nearbyint(mean + cfg.q_sd * rng.normal());
    q = std::min<double>(std::max<double>(q, cfg.q_min), cfg.q_max);
    return char(33 + int(q));
}

void revcompInplace(std::string& s){
    std::reverse(s.begin(), s.end());
    for(auto& c : s) c = complement(c);
}

struct ReadOut {
    std::string seq, qual, truth;
    std::uint64_t pos;
    int strand;
};

void finishRead(const care_sim_config& cfg, Rng& rng, ReadOut& r){
    const int L = int(r.truth.size());
    r.seq = r.truth;
    r.qual.resize(L);
    for(int j = 0; j < L; j++){
        r.qual[j] = drawQuality(cfg, rng, j, L);
        static const std::vector<double> kErr = [](){
            std::vector<double> t(128);
            for(int q = 0; q < 128; q++) t[q] = std::pow(10.0, -double(q) / 10.0);
            return t;
        }();
        const double perr = kErr[(r.qual[j] - 33) & 127];
        if(rng.uniform() < perr) r.seq[j] = otherBase(rng, r.seq[j]);
        if(cfg.n_rate > 0 && rng.uniform() < cfg.n_rate) r.seq[j] = 'N';
    }
}

}  // namespace

struct care_sim_result {
    std::vector<ReadOut> reads;
    int maxLen = 0;
    std::string genome;
};

extern "C" {

void care_sim_default_config(care_sim_config* cfg){
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->seed = 12345;
    cfg->genome_length = 200000;
    cfg->gc_content = 0.5;
    cfg->read_length_min = 150;
    cfg->read_length_max = 150;
    cfg->coverage = 30.0;
    cfg->paired = 0;
    cfg->insert_mean = 350.0;
    cfg->insert_sd = 50.0;
    cfg->q_first = 36.0;
    cfg->q_last = 28.0;
    cfg->q_sd = 3.0;
    cfg->q_min = 2;
    cfg->q_max = 41;
}

int care_sim_run(const care_sim_config* cfg, care_sim_result** out){
    if(!cfg || !out) return 1;
    if(cfg->read_length_min <= 0 || cfg->read_length_max < cfg->read_length_min) return 2;
    if(cfg->genome_length < std::uint64_t(cfg->read_length_max) * 2 + 16) return 3;
    auto* res = new (std::nothrow) care_sim_result;
    if(!res) return 4;
    Rng rng(cfg->seed);
    res->genome = makeGenome(*cfg, rng);
    const std::size_t G = res->genome.size();
    const double meanLen = 0.5 * double(cfg->read_length_min + cfg->read_length_max);
    std::size_t numReads = std::size_t(double(G) * cfg->coverage / meanLen);
    if(cfg->num_reads_override > 0) numReads = cfg->num_reads_override;
    if(cfg->paired){
        const std::size_t pairs = numReads / 2;
        res->reads.resize(2 * pairs);
        for(std::size_t p = 0; p < pairs; p++){
            const int L1 = drawLength(*cfg, rng);
            const int L2 = drawLength(*cfg, rng);
            long long ins = std::llround(cfg->insert_mean + cfg->insert_sd * rng.normal());
            ins = std::max<long long>(ins, std::max(L1, L2));
            ins = std::min<long long>(ins, (long long)G - 1);
            const std::uint64_t start = rng.below(G - std::uint64_t(ins));
            std::string frag = res->genome.substr(start, std::size_t(ins));
            const int strand = int(rng.next() & 1);
            if(strand) revcompInplace(frag);
            ReadOut& m1 = res->reads[2 * p];
            ReadOut& m2 = res->reads[2 * p + 1];
            m1.truth = frag.substr(0, L1);
            m2.truth = frag.substr(frag.size() - L2);
            revcompInplace(m2.truth);
            m1.pos = start; m1.strand = strand;
            m2.pos = start; m2.strand = !strand;
            finishRead(*cfg, rng, m1);
            finishRead(*cfg, rng, m2);
        }
    }else{
        res->reads.resize(numReads);
        for(std::size_t i = 0; i < numReads; i++){
            ReadOut& r = res->reads[i];
            const int L = drawLength(*cfg, rng);
            r.pos = rng.below(G - std::uint64_t(L) + 1);
            r.strand = int(rng.next() & 1);
            r.truth = res->genome.substr(r.pos, L);
            if(r.strand) revcompInplace(r.truth);
            finishRead(*cfg, rng, r);
        }
    }
    for(const auto& r : res->reads) res->maxLen = std::max(res->maxLen, int(r.seq.size()));
    *out = res;
    return 0;
}

std::uint64_t care_sim_num_reads(const care_sim_result* r){ return r->reads.size(); }
int care_sim_max_length(const care_sim_result* r){ return r->maxLen; }
std::uint64_t care_sim_genome_length(const care_sim_result* r){ return r->genome.size(); }

// Copies reads into row-major buffers of `pitch` bytes per read (pitch >= max length);
// unused bytes are zero.  Any output pointer may be NULL.
int care_sim_copy(const care_sim_result* r, std::uint64_t first, std::uint64_t n, std::size_t pitch,
                  char* seqs, char* quals, char* truth, int* lengths, std::uint64_t* positions, int* strands){
    if(first + n > r->reads.size()) return 1;
    if(pitch < std::size_t(r->maxLen)) return 2;
    for(std::uint64_t i = 0; i < n; i++){
        const ReadOut& rd = r->reads[first + i];
        const std::size_t L = rd.seq.size();
        if(seqs){ std::memset(seqs + i * pitch, 0, pitch); std::memcpy(seqs + i * pitch, rd.seq.data(), L); }
        if(quals){ std::memset(quals + i * pitch, 0, pitch); std::memcpy(quals + i * pitch, rd.qual.data(), L); }
        if(truth){ std::memset(truth + i * pitch, 0, pitch); std::memcpy(truth + i * pitch, rd.truth.data(), L); }
        if(lengths) lengths[i] = int(L);
        if(positions) positions[i] = rd.pos;
        if(strands) strands[i] = rd.strand;
    }
    return 0;
}

void care_sim_free(care_sim_result* r){ delete r; }

}  // extern "C"

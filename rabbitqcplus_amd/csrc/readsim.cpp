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
#include <memory>
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
    // generator of stream i under a base seed (streams are seeded through splitmix64
    // from distinct 64-bit values)
    static Rng stream(std::uint64_t base, std::uint64_t i){
        std::uint64_t x = base ^ (0xd1b54a32d192ed03ull * (i + 1));
        return Rng(splitmix(x));
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

void revcompInplace(char* s, int L){
    std::reverse(s, s + L);
    for(int i = 0; i < L; i++) s[i] = complement(s[i]);
}

// One read's rows inside the result's flat arrays.
struct ReadView {
    char* seq;
    char* qual;
    char* truth;
    int L;
};

// Qualities and errors of a read whose truth row is filled.
void finishRead(const care_sim_config& cfg, Rng& rng, ReadView r){
    static const std::vector<double> kErr = [](){
        std::vector<double> t(128);
        for(int q = 0; q < 128; q++) t[q] = std::pow(10.0, -double(q) / 10.0);
        return t;
    }();
    std::memcpy(r.seq, r.truth, std::size_t(r.L));
    for(int j = 0; j < r.L; j++){
        r.qual[j] = drawQuality(cfg, rng, j, r.L);
        const double perr = kErr[(r.qual[j] - 33) & 127];
        if(rng.uniform() < perr) r.seq[j] = otherBase(rng, r.seq[j]);
        if(cfg.n_rate > 0 && rng.uniform() < cfg.n_rate) r.seq[j] = 'N';
    }
}

}  // namespace

// Reads as row-major arrays of `pitch` (= longest possible read) bytes, zero padded.
// The arrays are allocated uninitialised and every row is written whole by the thread
// that simulates its read, so first-touch page faults are spread over the threads.
struct care_sim_result {
    std::size_t n = 0;
    std::size_t pitch = 0;
    std::unique_ptr<char[]> seq, qual, truth;
    std::vector<int> len;
    std::vector<std::uint64_t> pos;
    std::vector<int> strand;
    int maxLen = 0;
    std::string genome;

    ReadView view(std::size_t i, int L){
        return ReadView{seq.get() + i * pitch, qual.get() + i * pitch, truth.get() + i * pitch, L};
    }
    void resize(std::size_t count, std::size_t p){
        n = count; pitch = p;
        seq.reset(new char[n * pitch]); qual.reset(new char[n * pitch]); truth.reset(new char[n * pitch]);
        len.resize(n); pos.resize(n); strand.resize(n);
    }
    // zero padding of row i past its length
    void pad(std::size_t i){
        const std::size_t L = std::size_t(len[i]);
        std::memset(seq.get() + i * pitch + L, 0, pitch - L);
        std::memset(qual.get() + i * pitch + L, 0, pitch - L);
        std::memset(truth.get() + i * pitch + L, 0, pitch - L);
    }
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
    // Read (or pair) i draws from `rng` in read order, or from its own stream.
    const bool own = cfg->per_read_streams != 0;
    const std::uint64_t streamBase = own ? rng.next() : 0;
    const std::size_t P = std::size_t(cfg->read_length_max);
    if(cfg->paired){
        const std::size_t pairs = numReads / 2;
        res->resize(2 * pairs, P);
        auto onePair = [&](Rng& r, std::size_t p){
            const int L1 = drawLength(*cfg, r);
            const int L2 = drawLength(*cfg, r);
            long long ins = std::llround(cfg->insert_mean + cfg->insert_sd * r.normal());
            ins = std::max<long long>(ins, std::max(L1, L2));
            ins = std::min<long long>(ins, (long long)G - 1);
            const std::uint64_t start = r.below(G - std::uint64_t(ins));
            std::string frag = res->genome.substr(start, std::size_t(ins));
            const int strand = int(r.next() & 1);
            if(strand) revcompInplace(frag);
            const ReadView m1 = res->view(2 * p, L1);
            const ReadView m2 = res->view(2 * p + 1, L2);
            std::memcpy(m1.truth, frag.data(), std::size_t(L1));
            std::memcpy(m2.truth, frag.data() + frag.size() - std::size_t(L2), std::size_t(L2));
            revcompInplace(m2.truth, L2);
            res->len[2 * p] = L1; res->len[2 * p + 1] = L2;
            res->pos[2 * p] = start; res->strand[2 * p] = strand;
            res->pos[2 * p + 1] = start; res->strand[2 * p + 1] = !strand;
            finishRead(*cfg, r, m1);
            finishRead(*cfg, r, m2);
            res->pad(2 * p);
            res->pad(2 * p + 1);
        };
        if(own){
            #pragma omp parallel for schedule(dynamic, 4096)
            for(std::size_t p = 0; p < pairs; p++){
                Rng r = Rng::stream(streamBase, p);
                onePair(r, p);
            }
        }else{
            for(std::size_t p = 0; p < pairs; p++) onePair(rng, p);
        }
    }else{
        res->resize(numReads, P);
        auto oneRead = [&](Rng& r, std::size_t i){
            const int L = drawLength(*cfg, r);
            const std::uint64_t at = r.below(G - std::uint64_t(L) + 1);
            const int strand = int(r.next() & 1);
            const ReadView rd = res->view(i, L);
            std::memcpy(rd.truth, res->genome.data() + at, std::size_t(L));
            if(strand) revcompInplace(rd.truth, L);
            res->len[i] = L; res->pos[i] = at; res->strand[i] = strand;
            finishRead(*cfg, r, rd);
            res->pad(i);
        };
        if(own){
            #pragma omp parallel for schedule(dynamic, 4096)
            for(std::size_t i = 0; i < numReads; i++){
                Rng r = Rng::stream(streamBase, i);
                oneRead(r, i);
            }
        }else{
            for(std::size_t i = 0; i < numReads; i++) oneRead(rng, i);
        }
    }
    for(const int L : res->len) res->maxLen = std::max(res->maxLen, L);
    *out = res;
    return 0;
}

std::uint64_t care_sim_num_reads(const care_sim_result* r){ return r->n; }
int care_sim_max_length(const care_sim_result* r){ return r->maxLen; }
std::uint64_t care_sim_genome_length(const care_sim_result* r){ return r->genome.size(); }

// Copies reads into row-major buffers of `pitch` bytes per read (pitch >= max length);
// unused bytes are zero.  Any output pointer may be NULL.
int care_sim_copy(const care_sim_result* r, std::uint64_t first, std::uint64_t n, std::size_t pitch,
                  char* seqs, char* quals, char* truth, int* lengths, std::uint64_t* positions, int* strands){
    if(first + n > r->n) return 1;
    if(pitch < std::size_t(r->maxLen)) return 2;
    const std::size_t w = std::min(pitch, r->pitch);
    #pragma omp parallel for schedule(static, 16384)
    for(std::uint64_t i = 0; i < n; i++){
        const std::size_t src = (first + i) * r->pitch;
        if(seqs){ std::memcpy(seqs + i * pitch, r->seq.get() + src, w); std::memset(seqs + i * pitch + w, 0, pitch - w); }
        if(quals){ std::memcpy(quals + i * pitch, r->qual.get() + src, w); std::memset(quals + i * pitch + w, 0, pitch - w); }
        if(truth){ std::memcpy(truth + i * pitch, r->truth.get() + src, w); std::memset(truth + i * pitch + w, 0, pitch - w); }
        if(lengths) lengths[i] = r->len[first + i];
        if(positions) positions[i] = r->pos[first + i];
        if(strands) strands[i] = r->strand[first + i];
    }
    return 0;
}

void care_sim_free(care_sim_result* r){ delete r; }

std::size_t care_sim_fastq(const char* seqs, const char* quals, const int* lengths, std::size_t pitch,
                           std::uint64_t first, std::uint64_t step, std::uint64_t count, char* out, std::size_t cap){
    // record r: "@r" id "\n" seq "\n+\n" qual "\n"
    std::vector<std::size_t> off(count + 1, 0);
    for(std::uint64_t r = 0; r < count; r++){
        const std::uint64_t id = first + r * step;
        const std::string name = std::to_string(id);
        off[r + 1] = off[r] + 2 + name.size() + 1 + 2 * std::size_t(lengths[id]) + 4;
    }
    if(!out) return off[count];
    if(cap < off[count]) return 0;
    #pragma omp parallel for schedule(static, 16384)
    for(std::uint64_t r = 0; r < count; r++){
        const std::uint64_t id = first + r * step;
        const std::string name = std::to_string(id);
        const std::size_t L = std::size_t(lengths[id]);
        char* p = out + off[r];
        *p++ = '@'; *p++ = 'r';
        std::memcpy(p, name.data(), name.size()); p += name.size();
        *p++ = '\n';
        std::memcpy(p, seqs + id * pitch, L); p += L;
        *p++ = '\n'; *p++ = '+'; *p++ = '\n';
        std::memcpy(p, quals + id * pitch, L); p += L;
        *p++ = '\n';
    }
    return off[count];
}

}  // extern "C"

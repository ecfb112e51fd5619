// CARE per-anchor error correction on MI355X (gfx950): kernels and C-ABI.
//
// Pipeline per batch of anchors (each stage one kernel, all on the context stream):
//   k_candidates   workgroup per anchor: 48 table lookups, gather of the bucket
//                  values into LDS, bitonic sort, unique / remove-self / ambiguity
//                  filter                         (reference include/corrector.hpp:571-650)
//   k_compact      per-anchor candidate lists -> one dense list + anchor index
//   k_pair_flags   mate-pair flags for paired input (include/util.hpp:1016-1200)
//   k_align        one wavefront per (anchor, candidate): both orientations, lanes over
//                  shifts on HiLo bit planes held in LDS, popcount mismatches, wave
//                  argmin with scan-order tie-break (src/cpu_alignment.cpp:32-497,
//                  include/alignmentorientation.hpp:14-92)
//   k_msa          workgroup per anchor: orientation + mismatch-ratio filters, MSA
//                  columns in LDS (thread per column, candidates accumulated in list
//                  order so float sums match the reference), consensus, up to 5
//                  refinement rounds, HQ decision, corrected anchor, record encoding
//                  (include/corrector.hpp:982-1446,1639-1761; src/msa.cpp:9-453,532-734)
//   k_gather       variable-size records -> contiguous buffer in read-id order
//
// Set-up kernels: k_signatures (minhash of every read, include/cpusequencehasher.hpp:
// 88-125), k_hilo (HiLo planes of every read and its reverse complement,
// include/sequencehelpers.hpp:646-689), table build = radix sort + run-length
// encode + bucket filter + open-addressing insert (include/groupbykey.hpp:99-250).
//
// Float arithmetic follows the reference expression by expression and is compiled
// without contraction (-ffp-contract=off), so weights, supports and decisions are
// bit-identical to the reference's x86 build.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "care_hip.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <limits>
#include <map>
#include <new>
#include <string>
#include <numeric>
#include <thread>
#include <vector>

namespace {

constexpr int kMaxMaps = 48;         // include/config.hpp:30
constexpr int kBlock = 256;          // threads per workgroup of the per-anchor kernels
// One-dimensional launches: the dispatch packet carries a grid's size in work items as a
// 32-bit number, so element-wise kernels take a capped grid (gridFor) and stride over
// [0, n); kernels with one wavefront or one 16-lane row per item stride by rows.
constexpr uint64_t kMaxGridBlocks = uint64_t(1) << 20;    // x at most 1024 threads < 2^32 work items
#define FOR_GRID(i, n) for(uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < uint64_t(n); \
                            i += uint64_t(gridDim.x) * blockDim.x)
inline dim3 gridFor(uint64_t items, uint64_t block){
    const uint64_t b = (items + block - 1) / block;
    return dim3(uint32_t(b < 1 ? 1 : (b > kMaxGridBlocks ? kMaxGridBlocks : b)));
}
constexpr int kWave = 64;
constexpr int kMaxHiLoWords = 32;    // reads up to 1024 bases
constexpr int kMaxSortEntries = 32768;  // per-anchor gathered bucket values held in LDS
constexpr int kAnchorInfo = 12;
// learned seam: MLP 16 -> 64 -> 64 -> 1 over the features of one decision column
constexpr int kFeat = 16;
constexpr int kHid = 64;
constexpr int kMlpParams = kHid * kFeat + kHid + kHid * kHid + kHid + kHid + 1;
constexpr float kConfLogit = 2.1972245773f;      // ln 9: sigmoid(z) >= 0.9 <=> z >= ln 9
// LDS of the seam: 64 feature rows, both hidden layers, 64 logits
// seam scratch: bfloat16 feature rows of up to 64 columns, two hidden layers, the logits
constexpr size_t kSeamLdsBytes = size_t(kWave) * kFeat * 2 + 2 * kHid * 4 + kWave * 4;
constexpr size_t kSeamWeightBytes = size_t(kFeat * kHid + kHid * kHid + kHid) * 2;   // 10368, 16-byte aligned

struct Slot {            // one open-addressing slot of a per-map table (16 bytes)
    uint64_t key;
    uint32_t offset;     // first value of the bucket within the map's value array
    uint32_t count;      // bucket size; 0 = empty slot
};

struct AlignBest {       // chosen alignment of a candidate (orientation 3 = none)
    int shift;
    int overlap;
    int nOps;
    int orientation;
};

// ---------------------------------------------------------------- device helpers

__host__ __device__ __forceinline__ uint64_t fmix64(uint64_t x){
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

__device__ __forceinline__ int baseAt(const uint32_t* row, int i){
    return int((row[i >> 4] >> (30 - 2 * (i & 15))) & 3u);
}

__device__ __forceinline__ bool feq(float l, float r){
    const float absdiff = l - r < 0 ? -(l - r) : l - r;
    return absdiff < 1e-5f;
}
__device__ __forceinline__ bool fleq(float l, float r){ return (l < r) || feq(l, r); }
__device__ __forceinline__ bool fgeq(float l, float r){ return (l > r) || feq(l, r); }

__device__ __forceinline__ uint64_t waveMinU64(uint64_t v){
    #pragma unroll
    for(int o = 32; o > 0; o >>= 1){
        const uint64_t w = __shfl_xor(v, o, kWave);
        v = w < v ? w : v;
    }
    return v;
}

// Block-wide exclusive sum over kBlock threads; `tmp` holds kBlock/kWave+1 ints.
__device__ __forceinline__ int blockExclusiveSum(int v, int* tmp, int& total){
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    int x = v;
    #pragma unroll
    for(int o = 1; o < kWave; o <<= 1){
        const int y = __shfl_up(x, o, kWave);
        if(lane >= o) x += y;
    }
    if(lane == kWave - 1) tmp[wid] = x;
    __syncthreads();
    if(threadIdx.x == 0){
        int acc = 0;
        for(int w = 0; w < kBlock / kWave; w++){ const int t = tmp[w]; tmp[w] = acc; acc += t; }
        tmp[kBlock / kWave] = acc;
    }
    __syncthreads();
    const int res = x - v + tmp[wid];
    total = tmp[kBlock / kWave];
    __syncthreads();
    return res;
}

template<class Op>
__device__ __forceinline__ int blockReduce(int v, int* tmp, Op op, int identity){
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    #pragma unroll
    for(int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o, kWave));
    if(lane == 0) tmp[wid] = v;
    __syncthreads();
    if(threadIdx.x == 0){
        int acc = identity;
        for(int w = 0; w < kBlock / kWave; w++) acc = op(acc, tmp[w]);
        tmp[kBlock / kWave] = acc;
    }
    __syncthreads();
    const int res = tmp[kBlock / kWave];
    __syncthreads();
    return res;
}

__device__ __forceinline__ float blockMinF(float v, float* tmp){
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    #pragma unroll
    for(int o = 32; o > 0; o >>= 1){ const float w = __shfl_xor(v, o, kWave); v = w < v ? w : v; }
    if(lane == 0) tmp[wid] = v;
    __syncthreads();
    if(threadIdx.x == 0){
        float acc = tmp[0];
        for(int w = 1; w < kBlock / kWave; w++) acc = tmp[w] < acc ? tmp[w] : acc;
        tmp[kBlock / kWave] = acc;
    }
    __syncthreads();
    const float res = tmp[kBlock / kWave];
    __syncthreads();
    return res;
}

struct OpSum { __device__ int operator()(int a, int b) const { return a + b; } };
struct OpMin { __device__ int operator()(int a, int b) const { return a < b ? a : b; } };
struct OpMax { __device__ int operator()(int a, int b) const { return a > b ? a : b; } };
struct OpOr  { __device__ int operator()(int a, int b) const { return a | b; } };

// ---------------------------------------------------------------- set-up kernels

// Minhash signatures: one thread per read, all maps; output map-major [map][read].
// Canonical k-mer = min(forward, reverse complement) (include/sequencehelpers.hpp:844-923);
// sig[m] = (min over k-mers of fmix64(kmer + m)) & mask (include/cpusequencehasher.hpp:88-125).
// With `seeds`, also the seed of every map, read-major [read][map]: the start position of
// the first k-mer that attains the minimum, times 2, plus 1 when its canonical form is the
// reverse complement (0 for reads shorter than k). Seeds place a shared k-mer in both reads
// of a candidate pair (seeded alignment).
// A work item is one read and kSigGroup consecutive maps (items group-major: consecutive
// threads take consecutive reads of one group, so the signature stores coalesce): the
// k-mers are walked once per group, and a thread holds kSigGroup minima instead of all
// maps', which keeps the kernel at a high occupancy.
constexpr int kSigGroup = 8;
__host__ __device__ inline int sigGroups(int M){ return (M + kSigGroup - 1) / kSigGroup; }

__global__ void __launch_bounds__(256)
k_signatures(const uint32_t* __restrict__ enc, size_t pitch, const int* __restrict__ len, uint64_t n,
             int k, int M, uint64_t* __restrict__ sig, uint16_t* __restrict__ seeds){
    const uint64_t mask = ~0ull >> ((32 - k) * 2);
    FOR_GRID(t, n * uint64_t(sigGroups(M))){
    const uint64_t g = t / n;
    const uint64_t r = t - g * n;
    const int m0 = int(g) * kSigGroup;
    uint64_t mins[kSigGroup];
    uint32_t seed[kSigGroup];
    #pragma unroll
    for(int j = 0; j < kSigGroup; j++){ mins[j] = ~0ull; seed[j] = 0; }
    const int L = len[r];
    const uint32_t* row = enc + pitch * r;
    if(L >= k){
        uint64_t fwd = 0, rc = 0;
        const int top = 2 * k - 2;
        uint32_t word = 0;
        for(int i = 0; i < L; i++){
            if((i & 15) == 0) word = row[i >> 4];
            const uint64_t b = (word >> (30 - 2 * (i & 15))) & 3u;
            fwd = ((fwd << 2) | b) & mask;
            rc = (rc >> 2) | ((3ull - b) << top);
            if(i >= k - 1){
                const uint64_t canon = fwd < rc ? fwd : rc;
                const uint32_t here = (uint32_t(i - k + 1) << 1) | uint32_t(rc < fwd);
                #pragma unroll
                for(int j = 0; j < kSigGroup; j++){
                    if(m0 + j < M){
                        const uint64_t h = fmix64(canon + uint64_t(m0 + j));
                        if(h < mins[j]){ mins[j] = h; seed[j] = here; }
                    }
                }
            }
        }
    }
    #pragma unroll
    for(int j = 0; j < kSigGroup; j++){
        const int m = m0 + j;
        if(m < M){
            sig[uint64_t(m) * n + r] = mins[j] & mask;
            if(seeds) seeds[r * uint64_t(M) + uint64_t(m)] = uint16_t(seed[j]);
        }
    }
    }
}

// HiLo planes (include/sequencehelpers.hpp:646-689): bit 31-(i%32) of word i/32 holds
// the high (hi plane) / low (lo plane) bit of base i. Layout per read:
// [fwd hi W][fwd lo W][rc hi W][rc lo W]. Also the reverse complement in the 2-bit row
// format of `enc` (base i at bits 31-2(i%16) of word i/16), 2W words per read.
__global__ void __launch_bounds__(256)
k_hilo(const uint32_t* __restrict__ enc, size_t pitch, const int* __restrict__ len, uint64_t n, int W,
       uint32_t* __restrict__ hilo, uint32_t* __restrict__ rcenc){
    FOR_GRID(t, n * uint64_t(W)){
    const uint64_t r = t / W;
    const int w = int(t % W);
    const int L = len[r];
    const uint32_t* row = enc + pitch * r;
    uint32_t fh = 0, fl = 0, rh = 0, rl = 0, r0 = 0, r1 = 0;
    for(int j = 0; j < 32; j++){
        const int i = w * 32 + j;
        if(i < L){
            const int b = baseAt(row, i);
            const int rb = 3 - baseAt(row, L - 1 - i);
            fh |= uint32_t(b >> 1) << (31 - j);
            fl |= uint32_t(b & 1) << (31 - j);
            rh |= uint32_t(rb >> 1) << (31 - j);
            rl |= uint32_t(rb & 1) << (31 - j);
            if(j < 16) r0 |= uint32_t(rb) << (30 - 2 * j);
            else r1 |= uint32_t(rb) << (30 - 2 * (j - 16));
        }
    }
    uint32_t* out = hilo + r * uint64_t(4 * W);
    out[w] = fh;
    out[W + w] = fl;
    out[2 * W + w] = rh;
    out[3 * W + w] = rl;
    uint32_t* rout = rcenc + r * uint64_t(2 * W);
    rout[2 * w] = r0;
    rout[2 * w + 1] = r1;
    }
}

// Rows of `sp` bytes to rows of `dp` bytes: truncated, or zero padded past sp.
__global__ void k_repack_rows(const uint8_t* __restrict__ src, size_t sp, uint8_t* __restrict__ dst, size_t dp,
                              uint64_t total){
    FOR_GRID(t, total){
        const uint64_t r = t / dp;
        const uint64_t c = t - r * dp;
        dst[t] = c < sp ? src[r * sp + c] : uint8_t(0);
    }
}

__global__ void k_iota(uint32_t* v, uint64_t n){
    FOR_GRID(i, n) v[i] = uint32_t(i);
}

// Buckets that survive the 2..threshold rule (include/groupbykey.hpp:176-215).
__global__ void k_count_kept(const uint32_t* counts, const int* numRuns, int thr, uint32_t* kept){
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= *numRuns) return;
    const uint32_t c = counts[i];
    if(c >= 2 && c <= uint32_t(thr)) atomicAdd(kept, 1u);
}

__global__ void k_insert(const uint64_t* keys, const uint32_t* counts, const uint32_t* offsets, const int* numRuns,
                         int thr, Slot* slots, uint32_t capMask){
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= *numRuns) return;
    const uint32_t c = counts[i];
    if(c < 2 || c > uint32_t(thr)) return;
    const uint64_t key = keys[i];
    uint32_t h = uint32_t(fmix64(key)) & capMask;
    // keys are unique, so claiming an empty slot (count 0 -> c) is the only race
    for(;;){
        if(atomicCAS(&slots[h].count, 0u, c) == 0u){
            slots[h].key = key;
            slots[h].offset = offsets[i];
            return;
        }
        h = (h + 1) & capMask;
    }
}

// ---------------------------------------------------------------- candidates

struct CandParams {
    const uint64_t* sig; uint64_t n; int M; int k; int thr; int excludeAmbiguous;
    const int* len; const uint8_t* ambig;
    const Slot* const* slots; const uint32_t* caps; const uint32_t* values;
    const uint2* bucketOf;          // per (read, map), read-major: (offset, count) of the read's own bucket
    uint64_t b0; int B; int capPerAnchor; int hashSlots;
    uint32_t* raw; uint32_t* uniq; unsigned long long* gathered;
    uint32_t* gath;                 // per anchor: bucket values gathered (staged kernel)
    const uint32_t* order;          // processing order of the batch's anchors, or null
};

__device__ __forceinline__ uint32_t lookup(const Slot* slots, uint32_t cap, uint64_t key, uint32_t* offset){
    uint32_t h = uint32_t(fmix64(key)) & (cap - 1);
    for(;;){
        const Slot s = slots[h];
        if(s.count == 0) return 0;
        if(s.key == key){ *offset = s.offset; return s.count; }
        h = (h + 1) & (cap - 1);
    }
}

// Anchor processed by this workgroup (one anchor per workgroup, `B` in the batch). With
// an order, anchors that share a map-0 signature (overlapping reads, whose candidate sets
// largely coincide) run one after another on the same XCD: workgroups go to the 8 XCDs in
// turn, so workgroup b takes position (b mod 8) * B/8 + b/8 of the order, and each XCD's
// L2 sees a contiguous stretch of it. Results do not depend on the order.
__device__ __forceinline__ int batchAnchor(const uint32_t* order, int B){
    const int b = blockIdx.x;
    if(!order) return b;
    const int x = b & 7, j = b >> 3;
    const int q = B >> 3, r = B & 7;
    return int(order[x * q + min(x, r) + j]);
}

// Every read's own bucket in every map, read-major. A candidate lookup queries map m with
// the anchor's own signature for m (include/corrector.hpp:571-650), and every read's
// signature was inserted into every map, so the answer is the bucket the read belongs to
// (or nothing when that bucket was not kept); resolving it once per table set turns each
// pass's 48 signature loads and hash probes per anchor into one 384-byte read.
// A workgroup resolves kBucketTile reads at a time: lookups map by map (signatures of
// consecutive reads are consecutive), results transposed in LDS, rows stored read-major
// as one contiguous stretch.
constexpr int kBucketTile = 64;
__global__ void __launch_bounds__(256)
k_bucket_of(const uint64_t* __restrict__ sig, uint64_t n, int M, const Slot* const* slots,
            const uint32_t* __restrict__ caps, uint2* __restrict__ out){
    __shared__ uint2 tile[kBucketTile * kMaxMaps];
    const uint64_t ntiles = (n + kBucketTile - 1) / kBucketTile;
    for(uint64_t tb = blockIdx.x; tb < ntiles; tb += gridDim.x){
        const uint64_t a0 = tb * kBucketTile;
        const int rows = int(min(uint64_t(kBucketTile), n - a0));
        for(int q = threadIdx.x; q < kBucketTile * M; q += blockDim.x){
            const int m = q / kBucketTile;
            const int i = q - m * kBucketTile;
            if(i < rows){
                uint32_t off = 0;
                const uint32_t cnt = lookup(slots[m], caps[m], sig[uint64_t(m) * n + a0 + uint64_t(i)], &off);
                tile[i * M + m] = make_uint2(cnt ? off : 0u, cnt);
            }
        }
        __syncthreads();
        for(int q = threadIdx.x; q < rows * M; q += blockDim.x) out[a0 * uint64_t(M) + uint64_t(q)] = tile[q];
        __syncthreads();            // the tile is written again for the next reads
    }
}

// Inclusive prefix sum over the wave (DPP row shifts, then row broadcasts).
__device__ __forceinline__ int waveInclusiveSum(int v){
    int t;
    t = __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false); v += t;
    t = __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false); v += t;
    t = __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false); v += t;
    t = __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false); v += t;
    t = __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false); v += t;
    t = __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false); v += t;
    return v;
}

// Dynamic LDS of the candidate gather: hash set (hashSlots entries) then the table that
// maps each gathered position to its map (one byte per value, capPerAnchor entries).
__host__ __device__ inline size_t candLdsBytes(int hashSlots, int capPerAnchor){
    return size_t(hashSlots) * 4 + (size_t(capPerAnchor) + 15) / 16 * 16;
}

// Inserts v into the LDS hash set of H slots (power of two, more slots than distinct
// values); the first inserter appends v to out.
__device__ __forceinline__ void hashInsert(uint32_t* hset, int H, int hbits, uint32_t v, uint32_t* out, int* sUnique){
    uint32_t h = (v * 2654435761u) >> (32 - hbits);
    for(;;){
        const uint32_t prev = atomicCAS(&hset[h], 0xFFFFFFFFu, v);
        if(prev == 0xFFFFFFFFu){ out[atomicAdd(sUnique, 1)] = v; return; }
        if(prev == v) return;
        h = (h + 1) & uint32_t(H - 1);
    }
}

// Gathers the T bucket values of the anchor (positions [sPre[m], sPre[m+1]) belong to
// map m, read from values[m][sOff[m] + ...]) into the hash set: loads in batches of
// kGatherBatch per thread, all issued before their inserts.
constexpr int kGatherBatch = 8;
__device__ __forceinline__ void gatherInsert(const CandParams& p, uint64_t a, int T, const uint8_t* mapOf,
                                             const uint32_t* sOff, const uint32_t* sPre, uint32_t* hset, int H,
                                             uint32_t* out, int* sUnique, int tid, int nthreads){
    const int hbits = 31 - __clz(H);
    for(int base = 0; base < T; base += nthreads * kGatherBatch){
        uint32_t v[kGatherBatch];
        #pragma unroll
        for(int u = 0; u < kGatherBatch; u++){
            const int idx = base + u * nthreads + tid;
            v[u] = 0xFFFFFFFFu;
            if(idx < T){
                const int m = mapOf[idx];
                v[u] = p.values[uint64_t(m) * p.n + sOff[m] + uint32_t(idx - int(sPre[m]))];
            }
        }
        #pragma unroll
        for(int u = 0; u < kGatherBatch; u++){
            const uint32_t x = v[u];
            if(x == 0xFFFFFFFFu || uint64_t(x) == a || (p.excludeAmbiguous && p.ambig[x])) continue;
            hashInsert(hset, H, hbits, x, out, sUnique);
        }
    }
}

// Workgroup per anchor: candidate read ids, ascending, unique, self removed
// (include/corrector.hpp:571-650; include/ordinaryminhasher.hpp:135-221,437-449).
// Wave 0 looks the 48 signatures up; the gathered bucket values are deduplicated on
// insertion into an LDS hash set (the first inserter of a value appends it to the
// anchor's output area), then only the unique ids are sorted (rank sort, bitonic beyond
// 256). hset holds hashSlots entries, then the map table. Batch anchor i.
__device__ __forceinline__ void candidatesBlock(const CandParams& p, int i, uint32_t* hset, uint32_t* sOff, uint32_t* sPre,
                                                int& sUnique){
    uint8_t* mapOf = reinterpret_cast<uint8_t*>(hset + p.hashSlots);
    const uint64_t a = p.b0 + i;
    const int tid = threadIdx.x;
    if(tid < kWave){
        const bool valid = p.len[a] >= p.k && !(p.excludeAmbiguous && p.ambig[a]);
        uint32_t cnt = 0, off = 0;
        if(valid && tid < p.M){
            const uint2 b = p.bucketOf[a * uint64_t(p.M) + uint64_t(tid)];
            off = b.x;
            cnt = b.y > uint32_t(p.thr) ? 0u : b.y;
        }
        const int incl = waveInclusiveSum(int(cnt));
        if(tid < kMaxMaps){ sOff[tid] = off; sPre[tid + 1] = uint32_t(incl); }
        if(tid == 0){ sPre[0] = 0; sUnique = 0; }
        for(int q = incl - int(cnt); q < incl; q++) mapOf[q] = uint8_t(tid);
    }
    __syncthreads();
    const int T = int(sPre[kMaxMaps]);
    int H = 64;
    while(H <= T) H <<= 1;                   // more slots than values; H <= hashSlots
    for(int q = tid; q < H; q += kBlock) hset[q] = 0xFFFFFFFFu;
    __syncthreads();
    uint32_t* out = p.raw + uint64_t(i) * p.capPerAnchor;
    gatherInsert(p, a, T, mapOf, sOff, sPre, hset, H, out, &sUnique, tid, kBlock);
    __syncthreads();
    const int U = sUnique;
    if(U <= kBlock){
        // rank sort (the ids are distinct): one pass over the set per thread, reading
        // four ids per LDS access, then every id goes straight to its rank
        const int U4 = (U + 3) & ~3;
        for(int q = tid; q < U4; q += kBlock) hset[q] = q < U ? out[q] : 0xFFFFFFFFu;
        __syncthreads();
        if(tid < U){
            const uint32_t v = hset[tid];
            int rank = 0;
            const uint4* h4 = reinterpret_cast<const uint4*>(hset);
            for(int q = 0; q < U4 / 4; q++){
                const uint4 w = h4[q];
                rank += int(w.x < v) + int(w.y < v) + int(w.z < v) + int(w.w < v);
            }
            out[rank] = v;
        }
    }else{
        int P = 1;
        while(P < U) P <<= 1;
        for(int q = tid; q < P; q += kBlock) hset[q] = q < U ? out[q] : 0xFFFFFFFFu;
        __syncthreads();
        for(int kk = 2; kk <= P; kk <<= 1){
            for(int j = kk >> 1; j > 0; j >>= 1){
                for(int idx = tid; idx < P; idx += kBlock){
                    const int ixj = idx ^ j;
                    if(ixj > idx){
                        const uint32_t x = hset[idx], y = hset[ixj];
                        const bool up = (idx & kk) == 0;
                        if((x > y) == up){ hset[idx] = y; hset[ixj] = x; }
                    }
                }
                __syncthreads();
            }
        }
        for(int q = tid; q < U; q += kBlock) out[q] = hset[q];
    }
    if(tid == 0){
        p.uniq[i] = uint32_t(U);
        p.gath[i] = uint32_t(T);
    }
}

__global__ void __launch_bounds__(kBlock)
k_candidates(CandParams p){
    extern __shared__ uint32_t hset[];
    __shared__ uint32_t sOff[kMaxMaps];
    __shared__ uint32_t sPre[kMaxMaps + 1];
    __shared__ int sUnique;
    candidatesBlock(p, batchAnchor(p.order, p.B), hset, sOff, sPre, sUnique);
}

// One wavefront per anchor with a small hash set: most anchors have far fewer unique
// candidates than gathered bucket values, so kSmallHash slots keep the wave's LDS near
// 5 KB and many anchors in flight per compute unit. At most kSmallUnique distinct ids are
// taken; an anchor with more is appended to `over` and redone by k_candidates_over (the
// workgroup code above), so every anchor gets the same ids either way.
constexpr int kSmallHash = 256;
constexpr int kSmallUnique = 128;

__host__ __device__ inline size_t candWaveLdsBytes(int capPerAnchor){
    return size_t(kSmallHash) * 4 + (size_t(capPerAnchor) + 15) / 16 * 16;
}

__global__ void __launch_bounds__(kWave)
k_candidates_wave(CandParams p, uint32_t* __restrict__ over, uint32_t* __restrict__ overCount){
    extern __shared__ uint32_t hset[];       // kSmallHash entries (reused for the sort), then mapOf
    uint8_t* mapOf = reinterpret_cast<uint8_t*>(hset + kSmallHash);
    __shared__ uint32_t sOff[kMaxMaps];
    __shared__ uint32_t sPre[kMaxMaps + 1];
    __shared__ int sUnique;
    __shared__ int sOver;
    const int i = batchAnchor(p.order, p.B);
    const uint64_t a = p.b0 + i;
    const int lane = threadIdx.x;
    const bool valid = p.len[a] >= p.k && !(p.excludeAmbiguous && p.ambig[a]);
    uint32_t cnt = 0, off = 0;
    if(valid && lane < p.M){
        const uint2 b = p.bucketOf[a * uint64_t(p.M) + uint64_t(lane)];
        off = b.x;
        cnt = b.y > uint32_t(p.thr) ? 0u : b.y;
    }
    const int incl = waveInclusiveSum(int(cnt));
    if(lane < kMaxMaps){ sOff[lane] = off; sPre[lane + 1] = uint32_t(incl); }
    if(lane == 0){ sPre[0] = 0; sUnique = 0; sOver = 0; }
    for(int q = incl - int(cnt); q < incl; q++) mapOf[q] = uint8_t(lane);
    const int T = __builtin_amdgcn_readlane(incl, 63);
    for(int q = lane; q < kSmallHash; q += kWave) hset[q] = 0xFFFFFFFFu;
    __syncthreads();
    uint32_t* out = p.raw + uint64_t(i) * p.capPerAnchor;
    // gather in batches of kGatherBatch loads per lane, then insert them one step at a time;
    // a step is skipped by the whole wave once the set holds too many ids, so at most
    // kSmallUnique + kWave < kSmallHash slots are ever filled and every probe finds a slot
    for(int base = 0; base < T; base += kWave * kGatherBatch){
        uint32_t v[kGatherBatch];
        #pragma unroll
        for(int u = 0; u < kGatherBatch; u++){
            const int idx = base + u * kWave + lane;
            v[u] = 0xFFFFFFFFu;
            if(idx < T){
                const int m = mapOf[idx];
                v[u] = p.values[uint64_t(m) * p.n + sOff[m] + uint32_t(idx - int(sPre[m]))];
            }
        }
        #pragma unroll
        for(int u = 0; u < kGatherBatch; u++){
            if(__builtin_amdgcn_readfirstlane(*(volatile int*)&sOver)) break;
            const uint32_t x = v[u];
            if(x == 0xFFFFFFFFu || uint64_t(x) == a || (p.excludeAmbiguous && p.ambig[x])) continue;
            uint32_t h = (x * 2654435761u) >> (32 - 8);
            for(;;){
                const uint32_t prev = atomicCAS(&hset[h], 0xFFFFFFFFu, x);
                if(prev == 0xFFFFFFFFu){
                    const int r = atomicAdd(&sUnique, 1);
                    if(r < kSmallUnique) out[r] = x;
                    else sOver = 1;
                    break;
                }
                if(prev == x) break;
                h = (h + 1) & uint32_t(kSmallHash - 1);
            }
        }
        if(__builtin_amdgcn_readfirstlane(*(volatile int*)&sOver)) break;
    }
    __syncthreads();
    const int U = sUnique;
    if(sOver || U > kSmallUnique){
        if(lane == 0) over[atomicAdd(overCount, 1u)] = uint32_t(i);
        return;
    }
    // rank sort of the distinct ids, up to two per lane
    const int U4 = (U + 3) & ~3;
    for(int q = lane; q < U4; q += kWave) hset[q] = q < U ? out[q] : 0xFFFFFFFFu;
    __syncthreads();
    uint32_t w0 = lane < U ? hset[lane] : 0xFFFFFFFFu;
    uint32_t w1 = lane + kWave < U ? hset[lane + kWave] : 0xFFFFFFFFu;
    int r0 = 0, r1 = 0;
    const uint4* h4 = reinterpret_cast<const uint4*>(hset);
    for(int q = 0; q < U4 / 4; q++){
        const uint4 w = h4[q];
        r0 += int(w.x < w0) + int(w.y < w0) + int(w.z < w0) + int(w.w < w0);
        r1 += int(w.x < w1) + int(w.y < w1) + int(w.z < w1) + int(w.w < w1);
    }
    if(lane < U) out[r0] = w0;
    if(lane + kWave < U) out[r1] = w1;
    if(lane == 0){
        p.uniq[i] = uint32_t(U);
        p.gath[i] = uint32_t(T);
    }
}

// The anchors k_candidates_wave handed over (over[0, *overCount)), by workgroups that
// stride through the list.
__global__ void __launch_bounds__(kBlock)
k_candidates_over(CandParams p, const uint32_t* __restrict__ over, const uint32_t* __restrict__ overCount){
    extern __shared__ uint32_t hset[];
    __shared__ uint32_t sOff[kMaxMaps];
    __shared__ uint32_t sPre[kMaxMaps + 1];
    __shared__ int sUnique;
    const uint32_t count = *overCount;
    for(uint32_t t = blockIdx.x; t < count; t += gridDim.x){
        __syncthreads();
        candidatesBlock(p, int(over[t]), hset, sOff, sPre, sUnique);
    }
}

// Sum of the per-anchor gathered-value counts of a batch into the pass counter: each
// block sums kAddPerBlock counts and adds them with one atomic.
constexpr int kAddPerBlock = 16 * kBlock;

__global__ void k_add_gathered(const uint32_t* __restrict__ gath, int B, unsigned long long* __restrict__ total){
    __shared__ unsigned long long part[kBlock / kWave];
    unsigned long long s = 0;
    const int lo = blockIdx.x * kAddPerBlock, hi = min(B, lo + kAddPerBlock);
    for(int q = lo + int(threadIdx.x); q < hi; q += blockDim.x) s += gath[q];
    for(int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if(threadIdx.x == 0){
        unsigned long long t = 0;
        for(int w = 0; w < int(blockDim.x / kWave); w++) t += part[w];
        if(t) atomicAdd(total, t);
    }
}

// Candidate lists of the batch's B anchors into one array (offsets from the scan of their
// counts), a wavefront per anchor; anchorOf (the anchor of every entry) when not null.
__global__ void __launch_bounds__(kBlock)
k_compact(const uint32_t* raw, int capPerAnchor, const uint32_t* offsets, int B, uint32_t* ids, uint32_t* anchorOf){
    const int i = blockIdx.x * (kBlock / kWave) + int(threadIdx.x / kWave);
    if(i >= B) return;
    const uint32_t lane = threadIdx.x % kWave;
    const uint32_t beg = offsets[i];
    const uint32_t cnt = offsets[i + 1] - beg;
    const uint32_t* src = raw + uint64_t(i) * capPerAnchor;
    for(uint32_t j = lane; j < cnt; j += kWave){
        ids[beg + j] = src[j];
        if(anchorOf) anchorOf[beg + j] = uint32_t(i);
    }
}

// Pair flags of mates 2p, 2p+1 (include/util.hpp:1016-1141 via findPositionsOfPairedReadIds
// :1153-1200): an entry of one mate's sorted candidate list is flagged when the other
// mate's list holds its mate (id ^ 1) -- the entries the reference's merge of the two
// lists outputs. One wavefront per pair: lanes over both lists' entries, each a binary
// search in the other list; byte flags written coalesced.
__global__ void __launch_bounds__(kBlock) k_pair_flags(const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ ids,
                                                       uint8_t* __restrict__ flags, int numPairs){
    const int pidx = blockIdx.x * (kBlock / kWave) + int(threadIdx.x / kWave);
    const int lane = threadIdx.x & (kWave - 1);
    if(pidx >= numPairs) return;
    const uint32_t b1 = offsets[2 * pidx], e1 = offsets[2 * pidx + 1], e2 = offsets[2 * pidx + 2];
    for(uint32_t q = b1 + lane; q < e2; q += kWave){
        const bool inFirst = q < e1;
        const uint32_t want = ids[q] ^ 1u;
        const uint32_t end = inFirst ? e2 : e1;
        uint32_t lo = inFirst ? e1 : b1, hi = end;
        while(lo < hi){
            const uint32_t mid = (lo + hi) >> 1;
            if(ids[mid] < want) lo = mid + 1; else hi = mid;
        }
        flags[q] = (lo < end && ids[lo] == want) ? 1 : 0;
    }
}

// ---------------------------------------------------------------- alignment

struct AlignParams {
    const uint32_t* hilo; int W; const int* len;
    const uint32_t* ids; const uint32_t* anchorOf; uint64_t b0; uint32_t total;
    const uint32_t* offsets; int B;     // lane kernel: candidates of anchor b0+i are [offsets[i], offsets[i+1])
    int maxLen;
    int minOverlap; float maxErrorRate; float minOverlapRatio; float orientErrorRate;
    // int(float(ov) * maxErrorRate) == mulhi(ov << ePre, eMul) for every ov <= maxLen when
    // eExact (checked on the host)
    uint32_t eMul; int ePre; int eExact;
    AlignBest* best; int* dbg;   // dbg: [total][11] or null
    const uint32_t* order;       // processing order of the batch's anchors, or null
    // seeded alignment: every read's bucket (offset, count) and seed per map, read-major
    const uint2* bucketOf; const uint16_t* seeds; int M; int k; int thr;
};

struct ShdResult { int score, overlap, shift, nOps, valid; };

// Minimum over the 64 lanes with DPP row shifts/broadcasts (no LDS round trips);
// the result is read from lane 63 and is wave-uniform.
__device__ __forceinline__ uint32_t waveMinU32(uint32_t v){
    constexpr int kId = -1;   // 0xFFFFFFFF, identity of unsigned min
    uint32_t t;
    t = uint32_t(__builtin_amdgcn_update_dpp(kId, int(v), 0x111, 0xf, 0xf, false)); v = t < v ? t : v;  // row_shr:1
    t = uint32_t(__builtin_amdgcn_update_dpp(kId, int(v), 0x112, 0xf, 0xf, false)); v = t < v ? t : v;  // row_shr:2
    t = uint32_t(__builtin_amdgcn_update_dpp(kId, int(v), 0x114, 0xf, 0xf, false)); v = t < v ? t : v;  // row_shr:4
    t = uint32_t(__builtin_amdgcn_update_dpp(kId, int(v), 0x118, 0xf, 0xf, false)); v = t < v ? t : v;  // row_shr:8
    t = uint32_t(__builtin_amdgcn_update_dpp(kId, int(v), 0x142, 0xa, 0xf, false)); v = t < v ? t : v;  // row_bcast:15
    t = uint32_t(__builtin_amdgcn_update_dpp(kId, int(v), 0x143, 0xc, 0xf, false)); v = t < v ? t : v;  // row_bcast:31
    return uint32_t(__builtin_amdgcn_readlane(int(v), 63));
}

// Best shifts of both orientations (src/cpu_alignment.cpp:32-497, called for the
// forward and the reverse-complement candidate by include/corrector.hpp:925-979).
//
// The reference scans shift 0..La-minoverlap (anchor shifted), then -1..-(Lc-minoverlap)
// (candidate shifted); shift t of that scan order is admissible iff its mismatches are
// < int(float(overlap)*maxErrorRate), scores mismatches + La + Lc - 2*overlap, and the
// strict-'<' scan keeps the first admissible shift of minimum score. Its early break and
// its running-best bound only skip shifts that cannot beat the running best (the bound
// best-total+2*overlap never grows along a scan direction), so the result is
// argmin over admissible t of (score, t), or (La+Lc, -Lc) when none is admissible.
//
// Lanes take 64 consecutive shifts of one direction at a time, nearest shift 0 first in
// both directions (largest overlaps, where the best usually is), so later chunks are
// skipped wave-uniformly once the running bests rule them out. Because chunks are not
// visited in scan order, a lane may still win a tie when its scan index precedes the
// running best's, which the bound allows (+1). Keys (score << 16 | t) make the wave
// minimum the lexicographic (score, t) minimum. Both orientations share the candidate
// length, hence every per-shift quantity except the running best, and the anchor's
// planes; they are evaluated in the same pass.
//
// Planes are LDS pairs (word i+1, word i): one 64-bit read gives (word i << 32 | word i+1)
// for the funnel shift, and .y is word i itself.
struct OrientState {
    uint32_t key = 0xFFFFFFFFu;
    int best = 0;
    int bestT = INT_MAX;
};

__device__ __forceinline__ uint32_t shiftedWord(const uint2* X, int idx, int rs){
    const uint2 v = X[idx];
    return uint32_t(((uint64_t(v.y) << 32) | v.x) >> rs);
}

__device__ __forceinline__ ShdResult finishOrientation(const OrientState& st, int La, int Lc, int F){
    ShdResult r;
    const int total = La + Lc;
    const int bestShift = st.bestT == INT_MAX ? -Lc : (st.bestT < F ? st.bestT : -(st.bestT - F + 1));
    r.valid = bestShift != -Lc;
    const int cb = max(-bestShift, 0);
    const int ce = min(Lc, La - bestShift);
    r.overlap = ce - cb;
    r.score = st.best;
    r.shift = bestShift;
    r.nOps = st.best - total + 2 * r.overlap;
    return r;
}

__device__ __forceinline__ void alignBoth(const uint2* Ahi, const uint2* Alo, const uint2* Fhi, const uint2* Flo,
                          const uint2* Rhi, const uint2* Rlo, int La, int Lc, const AlignParams& p, int lane,
                          ShdResult& outF, ShdResult& outR){
    const int total = La + Lc;
    const int minoverlap = max(p.minOverlap, int(float(La) * p.minOverlapRatio));
    const int F = max(La - minoverlap + 1, 0);
    const int Bn = max(Lc - minoverlap, 0);
    const int nfc = (F + kWave - 1) / kWave;
    const int nbc = (Bn + kWave - 1) / kWave;
    OrientState sf, sr;
    sf.best = total;
    sr.best = total;
    bool fwdOpen = nfc > 0, bwdOpen = nbc > 0;
    for(int c = 0; fwdOpen || bwdOpen; c++){
        for(int dir = 0; dir < 2; dir++){
            bool& open = dir == 0 ? fwdOpen : bwdOpen;
            if(!open) continue;
            if(c >= (dir == 0 ? nfc : nbc)){ open = false; continue; }
            const int first = c * kWave;
            const int region = dir == 0 ? F : Bn;
            const int tBase = dir == 0 ? 0 : F;
            // chunk-level bounds at the chunk's largest overlap (its first shift)
            const int s0 = dir == 0 ? first : -(first + 1);
            const int ov0 = dir == 0 ? min(La - s0, Lc) : min(La, Lc + s0);
            const int e0 = int(float(ov0) * p.maxErrorRate);
            const int b0f = min(e0, sf.best - total + 2 * ov0 + ((tBase + first) < sf.bestT ? 1 : 0));
            const int b0r = min(e0, sr.best - total + 2 * ov0 + ((tBase + first) < sr.bestT ? 1 : 0));
            if(b0f <= 0 && b0r <= 0){ open = false; continue; }   // bounds only shrink along a direction
            const int idx = first + lane;
            uint32_t keyF = 0xFFFFFFFFu, keyR = 0xFFFFFFFFu;
            if(idx < region){
                const int s = dir == 0 ? idx : -(idx + 1);
                const int t = tBase + idx;
                const int ov = dir == 0 ? min(La - s, Lc) : min(La, Lc + s);
                const int e = int(float(ov) * p.maxErrorRate);
                const int mF = min(e, sf.best - total + 2 * ov + (t < sf.bestT ? 1 : 0));
                const int mR = min(e, sr.best - total + 2 * ov + (t < sr.bestT ? 1 : 0));
                if((mF > 0 || mR > 0) && ov > 0){
                    const int sh = dir == 0 ? s : -s;
                    const int ps = sh >> 5;
                    const int rs = 32 - (sh & 31);          // 64-bit funnel: 32 means no shift
                    const int parts = (ov + 31) >> 5;
                    const int rem = (parts << 5) - ov;
                    int misF = 0, misR = 0;
                    bool actF = mF > 0, actR = mR > 0;
                    for(int w = 0; w < parts && (actF || actR); w++){
                        const uint32_t mask = (w == parts - 1 && rem) ? (0xFFFFFFFFu << rem) : 0xFFFFFFFFu;
                        if(dir == 0){
                            // anchor shifted (shared), candidate in place
                            const uint32_t xh = shiftedWord(Ahi, w + ps, rs);
                            const uint32_t xl = shiftedWord(Alo, w + ps, rs);
                            if(actF){
                                misF += __popc(((xh ^ Fhi[w].y) | (xl ^ Flo[w].y)) & mask);
                                actF = misF < mF;
                            }
                            if(actR){
                                misR += __popc(((xh ^ Rhi[w].y) | (xl ^ Rlo[w].y)) & mask);
                                actR = misR < mR;
                            }
                        }else{
                            // candidate shifted, anchor in place (shared)
                            const uint32_t yh = Ahi[w].y;
                            const uint32_t yl = Alo[w].y;
                            if(actF){
                                misF += __popc(((shiftedWord(Fhi, w + ps, rs) ^ yh) | (shiftedWord(Flo, w + ps, rs) ^ yl)) & mask);
                                actF = misF < mF;
                            }
                            if(actR){
                                misR += __popc(((shiftedWord(Rhi, w + ps, rs) ^ yh) | (shiftedWord(Rlo, w + ps, rs) ^ yl)) & mask);
                                actR = misR < mR;
                            }
                        }
                    }
                    if(mF > 0 && misF < mF) keyF = (uint32_t(misF + total - 2 * ov) << 16) | uint32_t(t);
                    if(mR > 0 && misR < mR) keyR = (uint32_t(misR + total - 2 * ov) << 16) | uint32_t(t);
                }
            }
            // reduce only where some lane holds an admissible shift (a wave-uniform branch)
            if(__builtin_amdgcn_ballot_w64(keyF != 0xFFFFFFFFu)){
                const uint32_t wf = waveMinU32(keyF);
                if(wf < sf.key){ sf.key = wf; sf.best = int(wf >> 16); sf.bestT = int(wf & 0xFFFFu); }
            }
            if(__builtin_amdgcn_ballot_w64(keyR != 0xFFFFFFFFu)){
                const uint32_t wr = waveMinU32(keyR);
                if(wr < sr.key){ sr.key = wr; sr.best = int(wr >> 16); sr.bestT = int(wr & 0xFFFFu); }
            }
        }
    }
    outF = finishOrientation(sf, La, Lc, F);
    outR = finishOrientation(sr, La, Lc, F);
}

// include/alignmentorientation.hpp:14-92 (error bound = estimatedErrorrate * 4.0f).
__device__ __forceinline__ int chooseOrientation(const ShdResult& f, const ShdResult& r, int La, const AlignParams& p){
    const int ratioOverlap = int(La * p.minOverlapRatio);
    const int minimumOverlap = ratioOverlap > p.minOverlap ? ratioOverlap : p.minOverlap;
    const float maxErr = p.orientErrorRate;
    int ret = 3;
    if(f.valid && f.overlap >= minimumOverlap){
        if(r.valid && r.overlap >= minimumOverlap){
            const float ratio = (float)f.nOps / f.overlap;
            const float rcratio = (float)r.nOps / r.overlap;
            if(ratio < rcratio){
                if(ratio < maxErr) ret = 1;
            }else if(rcratio < ratio){
                if(rcratio < maxErr) ret = 2;
            }else{
                if(ratio < maxErr) ret = f.overlap > r.overlap ? 1 : 2;
            }
        }else{
            if((float)f.nOps / f.overlap < maxErr) ret = 1;
        }
    }else{
        if(r.valid && r.overlap >= minimumOverlap){
            if((float)r.nOps / r.overlap < maxErr) ret = 2;
        }
    }
    return ret;
}

constexpr int kAlignWaves = kBlock / kWave;

__global__ void __launch_bounds__(kBlock)
k_align(AlignParams p){
    // per wave: 6 planes (anchor hi/lo, candidate fwd hi/lo, candidate rc hi/lo) stored as
    // (word i+1, word i) pairs so one 64-bit LDS read feeds a funnel shift
    __shared__ uint2 planes[kAlignWaves][6][kMaxHiLoWords + 1];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    // workgroups stride over the pairs, kAlignWaves at a time (the same trip count for the
    // whole workgroup, which meets the barriers every time)
    for(uint64_t base = uint64_t(blockIdx.x) * kAlignWaves; base < p.total; base += uint64_t(gridDim.x) * kAlignWaves){
    const uint32_t c = uint32_t(base) + uint32_t(wave);
    const bool active = c < p.total;
    uint32_t id = 0;
    uint64_t a = 0;
    if(active){
        id = p.ids[c];
        a = p.b0 + p.anchorOf[c];
        const int W = p.W;
        const uint32_t* ah = p.hilo + a * uint64_t(4 * W);
        const uint32_t* ch = p.hilo + uint64_t(id) * uint64_t(4 * W);
        const uint32_t* src[6] = {ah, ah + W, ch, ch + W, ch + 2 * W, ch + 3 * W};
        for(int q = lane; q < 6 * (W + 1); q += kWave){
            const int pl = q / (W + 1);
            const int w = q - pl * (W + 1);
            const uint32_t x0 = w < W ? src[pl][w] : 0u;
            const uint32_t x1 = w + 1 < W ? src[pl][w + 1] : 0u;
            planes[wave][pl][w] = make_uint2(x1, x0);
        }
    }
    __syncthreads();
    if(active){
    const int La = p.len[a];
    const int Lc = p.len[id];
    ShdResult f, r;
    alignBoth(planes[wave][0], planes[wave][1], planes[wave][2], planes[wave][3], planes[wave][4], planes[wave][5],
              La, Lc, p, lane, f, r);
    const int flag = chooseOrientation(f, r, La, p);
    if(lane == 0){
        AlignBest b;
        const ShdResult& s = flag == 2 ? r : f;
        b.shift = s.shift;
        b.overlap = s.overlap;
        b.nOps = s.nOps;
        b.orientation = flag;
        p.best[c] = b;
        if(p.dbg){
            int* d = p.dbg + uint64_t(c) * 11;
            d[0] = f.score; d[1] = f.overlap; d[2] = f.shift; d[3] = f.nOps; d[4] = f.valid;
            d[5] = r.score; d[6] = r.overlap; d[7] = r.shift; d[8] = r.nOps; d[9] = r.valid;
            d[10] = flag;
        }
    }
    }
    __syncthreads();            // the planes are staged again for the next pairs
    }
}

// ---------------------------------------------------------------- alignment, lane per candidate
//
// The reference's scan (src/cpu_alignment.cpp:32-497) with one lane per candidate. All
// lanes of a wave share the anchor, so at each step of the scan the anchor's shifted
// words are wave-uniform (scalar registers, shifted on the scalar unit), and each lane
// compares them with its candidate's planes, both orientations, held in vector
// registers. A lane follows the reference step for step: shifts 0..La-mo with the anchor
// shifted, then -1..-(Lc-mo) with the candidate shifted (computed here as the anchor
// shifted the other way, in candidate coordinates); max_errors = min(int(float(ov) *
// maxErrorRate), best - total + 2*ov), the direction ends at the first shift where it is
// <= 0, and a shift with fewer mismatches than max_errors becomes the best (it always
// scores below the best then). The word-offset loops are unrolled so every register
// index is a constant.
//
// Anchor words are kept with one zero word in front: AX[i + 1] = word i, AX[0] = 0, and
// zeros past the read.
// Mask of the first min(32, max(0, k)) positions (most significant bits) of a word.
__device__ __forceinline__ uint32_t overlapMask(int k){
    const int c = min(max(k, 0), 32);
    return uint32_t(0xFFFFFFFF00000000ull >> c);
}

// Word counts (from a scan's first word) after which the lane scan tests whether any lane
// can still stay below its mismatch bound at the current shift.
template<int WM>
__device__ __forceinline__ constexpr bool abandonCheck(int w){
    return WM >= 8 ? (w == 2 || w == 3) : (w == 2);
}

// The two scan directions of one lane (src/cpu_alignment.cpp:32-497). UNI: every valid
// lane's candidate has the anchor's length, so the overlap of a shift, its max_errors
// (the host-checked integer form, on the scalar unit) and the overlap masks (anchor
// validity shifted AND unshifted) are wave-uniform; the results are the same as those of
// the general form.
template<int WM, bool UNI>
__device__ __forceinline__ void scanShifts(const AlignParams& p, const uint32_t* AH, const uint32_t* AL, const uint32_t* AV,
                                           const uint32_t* FH, const uint32_t* FL, const uint32_t* RH, const uint32_t* RL,
                                           const uint32_t* CV, int La, int Lc, int LcMax, int lastF, int mo, uint64_t start,
                                           int& gF, int& sF, int& gR, int& sR){
    gF = 0; sF = -Lc; gR = 0; sR = -Lc;
    // Reads of more than 160 bases (kLongWords): a shift's overlap lies in candidate
    // positions [0, min(La - sft, Lc)) (anchor shifted) or [t, min(t + La, Lc)) (candidate
    // shifted), so words from the wave-uniform bound on (LcMax: the longest candidate of
    // the wave) add nothing and are skipped. Mismatch counts only grow with the words: at
    // the words abandonCheck names, a shift is abandoned once no live lane can still stay
    // below its bound (short reads: after two words, where random sequence already
    // exceeds the bound of the longest overlap; long reads need a few more).
    constexpr bool kLongWords = WM >= 8;
    // anchor shifted: shifts 0..lastF. A position counts when it lies inside both the
    // candidate (CV) and the shifted anchor (shifted AV): exactly the overlap. The
    // lanes whose scan in an orientation is still running are kept as wave masks
    // (scalar registers); a lane's bit is read back with an inverse ballot.
    uint64_t liveF = start, liveR = start;
    #pragma unroll
    for(int ps = 0; ps <= WM; ps++){
        if(32 * ps > lastF || (liveF | liveR) == 0) break;
        const int rEnd = min(31, lastF - 32 * ps);
        for(int r = 0; r <= rEnd; r++){
            const int sft = 32 * ps + r;
            const int ov = UNI ? La - sft : min(La - sft, Lc);
            const int e = UNI ? int(__umulhi(uint32_t(ov) << p.ePre, p.eMul)) : int(float(ov) * p.maxErrorRate);
            const int mF = min(e, gF + 2 * ov);
            const int mR = min(e, gR + 2 * ov);
            // (a lane leaves its live mask for good; the block loop ends the scan once no
            // lane is live, which rarely happens inside a block)
            liveF &= __builtin_amdgcn_ballot_w64(mF > 0);
            liveR &= __builtin_amdgcn_ballot_w64(mR > 0);
            const int nw = kLongWords ? (min(La - sft, LcMax) + 31) >> 5 : WM;
            int dF = 0, dR = 0;
            #pragma unroll
            for(int w = 0; w < WM; w++){
                if(kLongWords && w >= nw) break;
                if(abandonCheck<WM>(w) && ((__builtin_amdgcn_ballot_w64(dF < mF) & liveF) |
                                           (__builtin_amdgcn_ballot_w64(dR < mR) & liveR)) == 0) break;
                const uint32_t xh = uint32_t((((uint64_t(AH[w + ps + 1]) << 32) | AH[w + ps + 2]) << r) >> 32);
                const uint32_t xl = uint32_t((((uint64_t(AL[w + ps + 1]) << 32) | AL[w + ps + 2]) << r) >> 32);
                const uint32_t xv = uint32_t((((uint64_t(AV[w + ps + 1]) << 32) | AV[w + ps + 2]) << r) >> 32);
                // UNI: candidate positions < La - sft are all inside the candidate (Lc == La)
                const uint32_t m = UNI ? xv : xv & CV[w];
                dF += __popc(((xh ^ FH[w]) | (xl ^ FL[w])) & m);
                dR += __popc(((xh ^ RH[w]) | (xl ^ RL[w])) & m);
            }
            const uint64_t iF = __builtin_amdgcn_ballot_w64(dF < mF) & liveF;
            const uint64_t iR = __builtin_amdgcn_ballot_w64(dR < mR) & liveR;
            if(__builtin_expect((iF | iR) != 0, 0)){
                __asm__ volatile("" ::: "memory");          // keeps the rare update a branch
                if(__builtin_amdgcn_inverse_ballot_w64(iF)){ gF = dF - 2 * ov; sF = sft; }
                if(__builtin_amdgcn_inverse_ballot_w64(iR)){ gR = dR - 2 * ov; sR = sft; }
            }
        }
    }
    // candidate shifted: shifts -1..-(Lc-mo), here t = 1..Lc-mo with the anchor
    // shifted right by t against the candidate's positions [t, t+ov)
    const int lastB = (UNI ? La : Lc) - mo;     // (valid lanes of a UNI scan have Lc == La)
    const int lastBw = UNI ? lastB : p.maxLen - mo;
    liveF = start; liveR = start;
    #pragma unroll
    for(int pt = 0; pt <= WM; pt++){
        if(32 * pt > lastBw || (liveF | liveR) == 0) break;
        const int rEnd = min(31, lastBw - 32 * pt);
        for(int r = (pt == 0 ? 1 : 0); r <= rEnd; r++){
            const int t = 32 * pt + r;
            const int ov = UNI ? La - t : min(La, Lc - t);
            const int e = UNI ? int(__umulhi(uint32_t(ov) << p.ePre, p.eMul)) : int(float(ov) * p.maxErrorRate);
            const int mF = min(e, gF + 2 * ov);
            const int mR = min(e, gR + 2 * ov);
            const uint64_t inRange = UNI ? ~0ull : __builtin_amdgcn_ballot_w64(t <= lastB);
            liveF &= inRange & __builtin_amdgcn_ballot_w64(mF > 0);
            liveR &= inRange & __builtin_amdgcn_ballot_w64(mR > 0);
            const int nw = kLongWords ? (min(t + La, LcMax) + 31) >> 5 : WM;
            int dF = 0, dR = 0;
            #pragma unroll
            for(int w = pt; w < WM; w++){
                if(kLongWords && w >= nw) break;
                if(abandonCheck<WM>(w - pt) && ((__builtin_amdgcn_ballot_w64(dF < mF) & liveF) |
                                                (__builtin_amdgcn_ballot_w64(dR < mR) & liveR)) == 0) break;
                const uint32_t xh = uint32_t(((uint64_t(AH[w - pt]) << 32) | AH[w - pt + 1]) >> r);
                const uint32_t xl = uint32_t(((uint64_t(AL[w - pt]) << 32) | AL[w - pt + 1]) >> r);
                const uint32_t xv = uint32_t(((uint64_t(AV[w - pt]) << 32) | AV[w - pt + 1]) >> r);
                const uint32_t m = xv & (UNI ? AV[w + 1] : CV[w]);
                dF += __popc(((xh ^ FH[w]) | (xl ^ FL[w])) & m);
                dR += __popc(((xh ^ RH[w]) | (xl ^ RL[w])) & m);
            }
            const uint64_t iF = __builtin_amdgcn_ballot_w64(dF < mF) & liveF;
            const uint64_t iR = __builtin_amdgcn_ballot_w64(dR < mR) & liveR;
            if(__builtin_expect((iF | iR) != 0, 0)){
                __asm__ volatile("" ::: "memory");
                if(__builtin_amdgcn_inverse_ballot_w64(iF)){ gF = dF - 2 * ov; sF = -t; }
                if(__builtin_amdgcn_inverse_ballot_w64(iR)){ gR = dR - 2 * ov; sR = -t; }
            }
        }
    }
}

// int(float(x) * maxErrorRate) for 0 <= x <= maxLen in the host-checked integer form
// (valid when p.eExact)
__device__ __forceinline__ int errBoundInt(const AlignParams& p, int x){
    return int(__umulhi(uint32_t(x) << p.ePre, p.eMul));
}

// The scan of scanShifts for candidates of differing lengths when the error bound has the
// integer form (p.eExact) and the minimum overlap mo is at least 1. The overlap of a
// shift is ov = min(a, b) of a wave-uniform term a and a lane term b (anchor shifted:
// a = La - sft, b = Lc; candidate shifted: a = La, b = Lc - t), both in [1, maxLen] while
// the lane's scan runs. e() is non-decreasing, so the bound min(e(ov), g + 2 ov) equals
// min(e(a), g + 2a, e(b), g + 2b): the wave-uniform parts are scalar, and the lane parts
// that do not move with the shift change only with the lane's best g, whose update runs
// in a wave-uniform branch taken at the few shifts where some lane improves. Results are
// those of scanShifts<WM, false>.
template<int WM>
__device__ __forceinline__ void scanShiftsVar(const AlignParams& p, const uint32_t* AH, const uint32_t* AL,
                                              const uint32_t* AV, const uint32_t* FH, const uint32_t* FL,
                                              const uint32_t* RH, const uint32_t* RL, const uint32_t* CV, int La, int Lc,
                                              int LcMax, int lastF, int mo, uint64_t start,
                                              int& gF, int& sF, int& gR, int& sR){
    gF = 0; sF = -Lc; gR = 0; sR = -Lc;
    constexpr bool kLongWords = WM >= 8;
    const int eC = errBoundInt(p, Lc);
    // anchor shifted: shifts 0..lastF, ov = min(La - sft, Lc); lane part min(e(Lc), g + 2 Lc)
    int cF = min(eC, gF + 2 * Lc), cR = min(eC, gR + 2 * Lc);
    uint64_t liveF = start, liveR = start;
    #pragma unroll
    for(int ps = 0; ps <= WM; ps++){
        if(32 * ps > lastF || (liveF | liveR) == 0) break;
        const int rEnd = min(31, lastF - 32 * ps);
        for(int r = 0; r <= rEnd; r++){
            const int sft = 32 * ps + r;
            const int a = La - sft;
            const int eA = errBoundInt(p, a);
            const int mF = min(min(eA, cF), gF + 2 * a);
            const int mR = min(min(eA, cR), gR + 2 * a);
            liveF &= __builtin_amdgcn_ballot_w64(mF > 0);
            liveR &= __builtin_amdgcn_ballot_w64(mR > 0);
            const int nw = kLongWords ? (min(a, LcMax) + 31) >> 5 : WM;
            int dF = 0, dR = 0;
            #pragma unroll
            for(int w = 0; w < WM; w++){
                if(kLongWords && w >= nw) break;
                if(abandonCheck<WM>(w) && ((__builtin_amdgcn_ballot_w64(dF < mF) & liveF) |
                                           (__builtin_amdgcn_ballot_w64(dR < mR) & liveR)) == 0) break;
                const uint32_t xh = uint32_t((((uint64_t(AH[w + ps + 1]) << 32) | AH[w + ps + 2]) << r) >> 32);
                const uint32_t xl = uint32_t((((uint64_t(AL[w + ps + 1]) << 32) | AL[w + ps + 2]) << r) >> 32);
                const uint32_t xv = uint32_t((((uint64_t(AV[w + ps + 1]) << 32) | AV[w + ps + 2]) << r) >> 32);
                const uint32_t m = xv & CV[w];
                dF += __popc(((xh ^ FH[w]) | (xl ^ FL[w])) & m);
                dR += __popc(((xh ^ RH[w]) | (xl ^ RL[w])) & m);
            }
            const uint64_t iF = __builtin_amdgcn_ballot_w64(dF < mF) & liveF;
            const uint64_t iR = __builtin_amdgcn_ballot_w64(dR < mR) & liveR;
            if(__builtin_expect((iF | iR) != 0, 0)){
                __asm__ volatile("" ::: "memory");          // keeps the rare update a branch
                const int ov = min(a, Lc);
                if(__builtin_amdgcn_inverse_ballot_w64(iF)){ gF = dF - 2 * ov; sF = sft; cF = min(eC, gF + 2 * Lc); }
                if(__builtin_amdgcn_inverse_ballot_w64(iR)){ gR = dR - 2 * ov; sR = sft; cR = min(eC, gR + 2 * Lc); }
            }
        }
    }
    // candidate shifted: t = 1..Lc-mo, ov = min(La, Lc - t). Lane parts: min(e(La), g + 2 La)
    // and g + 2 Lc (less 2t per shift); e(Lc - t) from (Lc << ePre) - (t << ePre), exact
    // while t <= Lc - mo < Lc.
    const int eLa = errBoundInt(p, La);
    const uint32_t LcP = uint32_t(Lc) << p.ePre;
    int bF = min(eLa, gF + 2 * La), bR = min(eLa, gR + 2 * La);
    int hF = gF + 2 * Lc, hR = gR + 2 * Lc;
    const int lastB = Lc - mo;
    const int lastBw = p.maxLen - mo;
    liveF = start; liveR = start;
    #pragma unroll
    for(int pt = 0; pt <= WM; pt++){
        if(32 * pt > lastBw || (liveF | liveR) == 0) break;
        const int rEnd = min(31, lastBw - 32 * pt);
        for(int r = (pt == 0 ? 1 : 0); r <= rEnd; r++){
            const int t = 32 * pt + r;
            const int eB = int(__umulhi(LcP - (uint32_t(t) << p.ePre), p.eMul));
            const int mF = min(min(eB, bF), hF - 2 * t);
            const int mR = min(min(eB, bR), hR - 2 * t);
            const uint64_t inRange = __builtin_amdgcn_ballot_w64(t <= lastB);
            liveF &= inRange & __builtin_amdgcn_ballot_w64(mF > 0);
            liveR &= inRange & __builtin_amdgcn_ballot_w64(mR > 0);
            const int nw = kLongWords ? (min(t + La, LcMax) + 31) >> 5 : WM;
            int dF = 0, dR = 0;
            #pragma unroll
            for(int w = pt; w < WM; w++){
                if(kLongWords && w >= nw) break;
                if(abandonCheck<WM>(w - pt) && ((__builtin_amdgcn_ballot_w64(dF < mF) & liveF) |
                                                (__builtin_amdgcn_ballot_w64(dR < mR) & liveR)) == 0) break;
                const uint32_t xh = uint32_t(((uint64_t(AH[w - pt]) << 32) | AH[w - pt + 1]) >> r);
                const uint32_t xl = uint32_t(((uint64_t(AL[w - pt]) << 32) | AL[w - pt + 1]) >> r);
                const uint32_t xv = uint32_t(((uint64_t(AV[w - pt]) << 32) | AV[w - pt + 1]) >> r);
                const uint32_t m = xv & CV[w];
                dF += __popc(((xh ^ FH[w]) | (xl ^ FL[w])) & m);
                dR += __popc(((xh ^ RH[w]) | (xl ^ RL[w])) & m);
            }
            const uint64_t iF = __builtin_amdgcn_ballot_w64(dF < mF) & liveF;
            const uint64_t iR = __builtin_amdgcn_ballot_w64(dR < mR) & liveR;
            if(__builtin_expect((iF | iR) != 0, 0)){
                __asm__ volatile("" ::: "memory");
                const int ov = min(La, Lc - t);
                if(__builtin_amdgcn_inverse_ballot_w64(iF)){
                    gF = dF - 2 * ov; sF = -t; bF = min(eLa, gF + 2 * La); hF = gF + 2 * Lc;
                }
                if(__builtin_amdgcn_inverse_ballot_w64(iR)){
                    gR = dR - 2 * ov; sR = -t; bR = min(eLa, gR + 2 * La); hR = gR + 2 * Lc;
                }
            }
        }
    }
}

template<int WM>
__device__ __forceinline__ void alignLanes(const AlignParams& p, uint64_t a, const uint32_t* ids, int n,
                                           AlignBest* best, int* dbg, int lane){
    const int W = p.W;
    const int La = p.len[a];
    const uint32_t* arow = p.hilo + a * uint64_t(4 * W);
    // anchor planes and its validity plane (positions < La), zero words on both sides
    uint32_t AH[2 * WM + 2], AL[2 * WM + 2], AV[2 * WM + 2];
    #pragma unroll
    for(int i = 0; i < 2 * WM + 2; i++){
        const bool in = i >= 1 && i <= WM && i <= W;
        AH[i] = in ? arow[i - 1] : 0u;
        AL[i] = in ? arow[W + i - 1] : 0u;
        AV[i] = in ? overlapMask(La - 32 * (i - 1)) : 0u;
    }
    const int mo = max(p.minOverlap, int(float(La) * p.minOverlapRatio));
    const int lastF = La - mo;
    for(int c0 = 0; c0 < n; c0 += kWave){
        const int c = c0 + lane;
        const bool valid = c < n;
        uint32_t id = 0;
        int Lc = 0;
        uint32_t FH[WM], FL[WM], RH[WM], RL[WM], CV[WM];
        if(valid){
            id = ids[c];
            Lc = p.len[id];
        }
        const uint32_t* crow = p.hilo + uint64_t(id) * uint64_t(4 * W);
        #pragma unroll
        for(int w = 0; w < WM; w++){
            const bool in = valid && w < W;
            FH[w] = in ? crow[w] : 0u;
            FL[w] = in ? crow[W + w] : 0u;
            RH[w] = in ? crow[2 * W + w] : 0u;
            RL[w] = in ? crow[3 * W + w] : 0u;
            CV[w] = in ? overlapMask(Lc - 32 * w) : 0u;
        }
        const int total = La + Lc;
        // gF, gR: best score minus total (score = total - 2 * overlap + mismatches)
        int gF, sF, gR, sR;
        const uint64_t start = __builtin_amdgcn_ballot_w64(valid);
        // longest candidate of the chunk (wave-uniform)
        int LcMax = Lc;
        #pragma unroll
        for(int o = 1; o < kWave; o <<= 1) LcMax = max(LcMax, __shfl_xor(LcMax, o));
        LcMax = __builtin_amdgcn_readfirstlane(LcMax);
        if(p.eExact && __builtin_amdgcn_ballot_w64(valid && Lc != La) == 0)
            scanShifts<WM, true>(p, AH, AL, AV, FH, FL, RH, RL, CV, La, Lc, LcMax, lastF, mo, start, gF, sF, gR, sR);
        else if(p.eExact && mo >= 1)
            scanShiftsVar<WM>(p, AH, AL, AV, FH, FL, RH, RL, CV, La, Lc, LcMax, lastF, mo, start, gF, sF, gR, sR);
        else
            scanShifts<WM, false>(p, AH, AL, AV, FH, FL, RH, RL, CV, La, Lc, LcMax, lastF, mo, start, gF, sF, gR, sR);
        const int bF = gF + total, bR = gR + total;
        if(!valid) continue;
        ShdResult f, rr;
        {
            f.score = bF; f.shift = sF; f.valid = sF != -Lc;
            f.overlap = min(Lc, La - sF) - max(-sF, 0);
            f.nOps = bF - total + 2 * f.overlap;
            rr.score = bR; rr.shift = sR; rr.valid = sR != -Lc;
            rr.overlap = min(Lc, La - sR) - max(-sR, 0);
            rr.nOps = bR - total + 2 * rr.overlap;
        }
        const int flag = chooseOrientation(f, rr, La, p);
        AlignBest b;
        const ShdResult& sres = flag == 2 ? rr : f;
        b.shift = sres.shift; b.overlap = sres.overlap; b.nOps = sres.nOps; b.orientation = flag;
        best[c] = b;
        if(dbg){
            int* d = dbg + uint64_t(c) * 11;
            d[0] = f.score; d[1] = f.overlap; d[2] = f.shift; d[3] = f.nOps; d[4] = f.valid;
            d[5] = rr.score; d[6] = rr.overlap; d[7] = rr.shift; d[8] = rr.nOps; d[9] = rr.valid;
            d[10] = flag;
        }
    }
}

// One wavefront per anchor of the batch, its candidates in chunks of 64 lanes.
template<int WM>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(WM <= 5 ? 8 : 1)))
k_align_lanes(AlignParams p){
    if(int(blockIdx.x) >= p.B) return;
    const int i = batchAnchor(p.order, p.B);
    const uint32_t beg = p.offsets[i];
    const int n = int(p.offsets[i + 1] - beg);
    if(n == 0) return;
    alignLanes<WM>(p, p.b0 + i, p.ids + beg, n, p.best + beg, p.dbg ? p.dbg + uint64_t(beg) * 11 : nullptr,
                   threadIdx.x);
}

// Seeded alignment (an addition; the reference has no counterpart): one shift per
// candidate instead of the scan of both orientations. A candidate was found through a
// bucket it shares with the anchor, i.e. through a canonical k-mer both reads contain;
// the first map (lowest index) whose kept bucket holds both reads gives that k-mer's
// seed in each read (start position p and whether the canonical form is the reverse
// complement, s). Equal s: the candidate is forward and starts at anchor position
// pa - pc; different s: it is the reverse complement, starting at pa - (Lc - k - pc).
// At that shift the candidate gets the reference's tests: overlap >= the minimum
// overlap, mismatches < int(float(overlap) * maxErrorRate), and mismatches / overlap
// below the orientation bound (include/alignmentorientation.hpp:14-92); otherwise it is
// dropped (orientation 3). The chosen alignment is written as the scan's would be.
//
// One wavefront per anchor, one lane per candidate. The anchor's HiLo planes sit in LDS
// between zero words, stored as (word i, word i+1) pairs so that one 64-bit read feeds
// the funnel shift (a word with any overlap position has its anchor bit offset in
// (-32, La), so pairs -1..W-1 suffice); a candidate word j (its bases 32j..32j+31 in the
// chosen strand, loaded into registers) is compared with the anchor bits at offset
// 32j + shift, of either sign, under the mask of the overlap
// [max(0, -shift), min(Lc, La - shift)). The map search reads two maps' buckets per load
// and stops at the first hit (a true overlap shares most maps).
template<int WM>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(8)))
k_align_seeded(AlignParams p){
    constexpr int PAD = 1;
    constexpr int P = WM + 2;
    __shared__ uint2 aPl[2][P];
    __shared__ uint4 aBucket[kMaxMaps / 2];     // (offset, count) of maps 2q, 2q+1; count 0 = not kept
    __shared__ int aSeed[kMaxMaps];
    if(int(blockIdx.x) >= p.B) return;
    const int i = batchAnchor(p.order, p.B);
    const uint32_t beg = p.offsets[i];
    const int n = int(p.offsets[i + 1] - beg);
    if(n == 0) return;
    const int lane = threadIdx.x;
    const uint64_t a = p.b0 + i;
    const int W = p.W;
    const int M = p.M;
    const int La = p.len[a];
    const uint32_t* arow = p.hilo + a * uint64_t(4 * W);
    for(int q = lane; q < 2 * P; q += kWave){
        const int pl = q / P, x = q % P;
        const int w0 = x - PAD, w1 = x + 1 - PAD;
        const uint32_t lo = (w0 >= 0 && w0 < W) ? arow[pl * W + w0] : 0u;
        const uint32_t hi = (w1 >= 0 && w1 < W) ? arow[pl * W + w1] : 0u;
        aPl[pl][x] = make_uint2(hi, lo);        // .y = word x - 1, .x = word x
    }
    if(lane < kMaxMaps){
        uint2 b = make_uint2(0u, 0u);
        int sd = 0;
        if(lane < M){
            b = p.bucketOf[a * uint64_t(M) + uint64_t(lane)];
            if(b.y == 0 || b.y > uint32_t(p.thr)) b = make_uint2(0u, 0u);
            sd = p.seeds[a * uint64_t(M) + uint64_t(lane)];
        }
        reinterpret_cast<uint2*>(aBucket)[lane] = b;
        aSeed[lane] = sd;
    }
    __syncthreads();
    const int mo = max(p.minOverlap, int(float(La) * p.minOverlapRatio));
    for(int c0 = 0; c0 < n; c0 += kWave){
        const int c = c0 + lane;
        if(c >= n) break;
        const uint32_t id = p.ids[beg + c];
        const int Lc = p.len[id];
        // first map whose kept bucket holds both reads (equal offset and count); rows of
        // M (even) buckets are 16-byte aligned
        const uint2* cbk = p.bucketOf + uint64_t(id) * uint64_t(M);
        int m = M;
        for(int q = 0; 2 * q < M; q++){
            const uint4 ab = aBucket[q];
            uint4 x;
            if(2 * q + 1 < M) x = *reinterpret_cast<const uint4*>(cbk + 2 * q);
            else { const uint2 y = cbk[2 * q]; x = make_uint4(y.x, y.y, 0u, 0u); }
            if(ab.y != 0 && x.x == ab.x && x.y == ab.y){ m = 2 * q; break; }
            if(ab.w != 0 && x.z == ab.z && x.w == ab.w){ m = 2 * q + 1; break; }
        }
        AlignBest b;
        b.orientation = 3; b.shift = 0; b.overlap = 0; b.nOps = 0;
        if(m < M){
            const int sa = aSeed[m], sc = p.seeds[uint64_t(id) * uint64_t(M) + uint64_t(m)];
            const int pa = sa >> 1, pc = sc >> 1;
            const bool same = (sa & 1) == (sc & 1);
            const int shift = same ? pa - pc : pa - (Lc - p.k - pc);
            const uint32_t* crow = p.hilo + uint64_t(id) * uint64_t(4 * W) + (same ? 0 : 2 * W);
            const int cbeg = max(-shift, 0), cend = min(Lc, La - shift);
            const int ov = cend - cbeg;
            int mis = 0;
            #pragma unroll
            for(int j = 0; j < WM; j++){
                const int lo = min(max(cbeg - 32 * j, 0), 32), hi = min(max(cend - 32 * j, 0), 32);
                if(hi > lo){
                    const uint32_t mask = uint32_t((0xFFFFFFFFull >> lo) & ~(0xFFFFFFFFull >> hi));
                    const int o = 32 * j + shift;                 // anchor bit offset of word j
                    const int x = (o >> 5) + PAD, r = o & 31;
                    const uint2 vh = aPl[0][x], vl = aPl[1][x];
                    const uint32_t ah = uint32_t((((uint64_t(vh.y) << 32) | vh.x) << r) >> 32);
                    const uint32_t al = uint32_t((((uint64_t(vl.y) << 32) | vl.x) << r) >> 32);
                    mis += __popc(((ah ^ crow[j]) | (al ^ crow[W + j])) & mask);
                }
            }
            b.shift = shift;
            b.overlap = ov;
            b.nOps = mis;
            if(ov > 0 && ov >= mo && mis < int(float(ov) * p.maxErrorRate) && float(mis) / float(ov) < p.orientErrorRate)
                b.orientation = same ? 1 : 2;
        }
        p.best[beg + c] = b;
    }
}

// Words per plane handled by the lane kernel for reads of W words (0: too long).
int alignLaneWords(int W){
    if(W <= 2) return 2;
    if(W <= 4) return 4;
    if(W <= 5) return 5;
    if(W <= 8) return 8;
    if(W <= 11) return 11;
    return 0;
}

// ---------------------------------------------------------------- MSA + decision

struct MsaParams {
    const uint32_t* enc; size_t pitch; const int* len; const uint8_t* qual; size_t qpitch; const uint8_t* ambig;
    size_t qdpitch;                 // device quality row pitch (qpitch rounded up to 16 bytes)
    int varLen;                     // reads of different lengths: stage only what each candidate covers
    const float* qweight;
    uint64_t b0;
    const uint32_t* offsets; const uint32_t* ids; const AlignBest* best; const uint8_t* paired;
    // per-candidate working lists (same indexing as ids)
    uint32_t* kId; int* kShift; int* kLen; int* kNops; int* kOvl; float* kW; uint8_t* kOrient; uint8_t* kPaired;
    int isPaired; int useQ;
    float maxErrorRate; float ratio2, ratio3, ratio4; float goodCount; float pairedThreshold;
    float avgThr, minThr, minCovThr; int pct; float pctHalf; int sigThr;
    float hqAvgT, hqMinT;           // HQ thresholds at hqPercent (see the decision)
    int maxCols; int maxLen;
    const uint32_t* hilo; int W; int tile;
    const uint32_t* rcenc;          // reverse-complement 2-bit rows, 2W words per read
    uint8_t* slots; int slotBytes; uint32_t* recBytes;
    int* info;                      // [B][kAnchorInfo] or null
    uint8_t* msaDbg; int dbgStride; // per anchor: 1 int ncol + cols*(4 int + 4 float + int + float + 1 byte)
    int* error;
    // learned seam (classifier decision of non-HQ anchor columns)
    int seamMode;                   // 0 classic decision, 1 MLP with classic fallback
    const uint16_t* seamW;          // bfloat16: W1^T [16][64], W2^T [64][64], W3 [64]
    const float* seamB;             // b1 [64], b2 [64], b3
    float estCov;
    float* featOut; uint32_t* metaOut; unsigned long long* featCount; unsigned long long featCap;
    // candidate correction: per working-list entry a record slot and its size; per read
    // whether it was corrected as an HQ anchor
    int corrCands; int newCols;
    uint8_t* candSlots; int candSlotBytes; uint32_t* candBytes; uint8_t* hqAnchor;
    const uint32_t* order; int B;   // processing order of the batch's anchors, or null
};

// Per-anchor candidate lists of the MSA stage: the staged kernels index batch-wide
// arrays, the fused kernel a wave's private scratch.
struct AnchorLists {
    const uint32_t* ids; const AlignBest* best; const uint8_t* paired;
    uint32_t* kId; int* kShift; int* kLen; int* kNops; int* kOvl; float* kW; uint8_t* kOrient; uint8_t* kPaired;
};

__device__ __forceinline__ AnchorLists listsOf(const MsaParams& p){
    return AnchorLists{p.ids, p.best, p.paired, p.kId, p.kShift, p.kLen, p.kNops, p.kOvl, p.kW, p.kOrient, p.kPaired};
}

struct MsaShared {
    int tmp[kBlock / kWave + 1];
    float ftmp[kBlock / kWave + 1];
    int n;          // working candidate count
    int col;
    int sigBase;
    int origBase;
    int ncol;
    int ab;
};

__device__ __forceinline__ int orientedBase(const MsaParams& p, uint32_t id, int lc, int orient, int pos){
    const uint32_t* row = p.enc + p.pitch * id;
    return orient == 1 ? baseAt(row, pos) : 3 - baseAt(row, lc - 1 - pos);
}

__device__ __forceinline__ float orientedQualityWeight(const MsaParams& p, uint32_t id, int orient, int pos){
    const uint8_t* q = p.qual + p.qdpitch * id;
    // reverse-complement candidates read the pitch-wide quality row reversed
    // (reverseQualitiesOfRCAlignments, include/corrector.hpp:1227-1239)
    return p.qweight[orient == 1 ? q[pos] : q[p.qpitch - 1 - pos]];
}

// LDS staging area for a tile of candidates (order kept): shift, length, overlap weight,
// oriented bases (2-bit rows of tileRowWords words, base i at bits 31-2(i%16) of word
// i/16) and qualities (qStride bytes each; see stageTile for their placement).
struct MsaTile {
    int* shift;
    int* len;
    float* w;
    uint32_t* id;
    int* orient;
    uint32_t* rows;      // 2W words reserved per candidate
    uint8_t* q;
    const float* qw;     // quality-weight table in LDS
};

// Row stride of the tile's quality rows: 16-byte aligned, so that rows are written with
// one 16-byte store per chunk.
__host__ __device__ inline int qStride(int maxLen){ return (maxLen + 15) & ~15; }

// Bytes of the one-wavefront MSA's tile region: the candidate tile, which the learned seam
// (kSeamLdsBytes) and the consensus of candidate correction (maxCols bytes) reuse.
__host__ __device__ inline size_t waveTileRegion(int T, int W, int maxLen, bool useQ, bool seam, int maxCols){
    size_t b = size_t(T) * (20 + 8 * size_t(W)) + (useQ ? size_t(T) * size_t(qStride(maxLen)) : 0);
    if(seam && b < kSeamLdsBytes) b = kSeamLdsBytes;
    if(b < size_t(maxCols)) b = size_t(maxCols);
    return (b + 15) / 16 * 16;
}

// 2-bit words of a tile row
__host__ __device__ inline int tileRowWords(int maxLen){ return (maxLen + 15) >> 4; }

// Offset of position 0 of a reverse-complement candidate's quality row in its tile slot:
// the reversed row is stored right-aligned (quality rows are maxLen = qpitch bytes, so
// the slot of qStride(maxLen) bytes holds the whole 16-byte-chunked row).
__device__ __forceinline__ int rcQualityOffset(const MsaParams& p){ return qStride(p.maxLen) - int(p.qpitch); }

template<int G>
__device__ __forceinline__ void stageQualities(const MsaParams& p, const MsaTile& t, int tn, int tid, int nthreads);

// Stages candidates [tb, tb+tn) of the working list into the LDS tile: one thread per
// candidate for the metadata and its oriented 2-bit row (the stored row or its reverse
// complement); qualities as 16-byte chunks of the device rows, spread over the threads. A
// reverse-complement candidate reads the pitch-wide quality row reversed
// (reverseQualitiesOfRCAlignments, include/corrector.hpp:1227-1239): each chunk is
// byte-reversed in registers and stored mirrored in the slot, so that position pos of the
// reversed row sits at rcQualityOffset + pos. Caller synchronizes before (tile free) and
// after (tile complete).
// An oriented 2-bit row into a tile slot. Reads of one length copy the row's RW words in
// a wave-uniform loop. Reads of different lengths copy only the words under the
// candidate's length (the MSA reads positions < len), as 16-byte loads and stores where
// vec holds (W even, RW and the enc pitch multiples of 4 words: the enc rows, the
// reverse-complement rows and the tile slots start on 16-byte boundaries) and the group of
// four words lies inside the row's RW words.
// Loads go out in groups (G 16-byte or 4G 4-byte loads) before their stores, so a group
// costs one memory round trip (G = 1 where the registers of a larger group would spill).
template<int G>
__device__ __forceinline__ void stageRow(const uint32_t* __restrict__ src, uint32_t* dst, int len, int RW, bool varLen,
                                         bool vec){
    if constexpr(G == 1){
        if(!varLen){
            for(int k = 0; k < RW; k++) dst[k] = src[k];
            return;
        }
        const int nw = min((len + 15) >> 4, RW);
        int k = 0;
        if(vec){
            for(; k < nw && k + 4 <= RW; k += 4)
                *reinterpret_cast<uint4*>(dst + k) = *reinterpret_cast<const uint4*>(src + k);
        }
        for(; k < nw; k++) dst[k] = src[k];
        return;
    }
    const int nw = varLen ? min((len + 15) >> 4, RW) : RW;
    if(vec){
        // (RW is a multiple of 4: the 16-byte groups cover the nw words and stay in the row)
        const uint4* s4 = reinterpret_cast<const uint4*>(src);
        uint4* d4 = reinterpret_cast<uint4*>(dst);
        const int nv = (nw + 3) >> 2;
        for(int g = 0; g < nv; g += G){
            uint4 r[G];
            #pragma unroll
            for(int i = 0; i < G; i++) if(g + i < nv) r[i] = s4[g + i];
            #pragma unroll
            for(int i = 0; i < G; i++) if(g + i < nv) d4[g + i] = r[i];
        }
        return;
    }
    for(int g = 0; g < nw; g += 4 * G){
        uint32_t r[4 * G];
        #pragma unroll
        for(int i = 0; i < 4 * G; i++) if(g + i < nw) r[i] = src[g + i];
        #pragma unroll
        for(int i = 0; i < 4 * G; i++) if(g + i < nw) dst[g + i] = r[i];
    }
}

// Whether 16-byte chunk c of a candidate's quality row holds a byte the MSA reads: a
// forward candidate reads positions [0, len); a reverse-complement candidate reads the
// row reversed at the storage width qpitch, i.e. bytes [qpitch - len, qpitch).
__device__ __forceinline__ bool qualityChunkUsed(int c, int len, bool rc, int qpitch){
    return rc ? (16 * c + 16 > qpitch - len && 16 * c < qpitch) : 16 * c < len;
}

template<int G>
__device__ __forceinline__ void stageTile(const MsaParams& p, const AnchorLists& lst, const MsaTile& t, int beg,
                                          int tb, int tn, int tid, int nthreads){
    const int W = p.W;
    for(int q = tid; q < tn; q += nthreads){
        const int d = beg + tb + q;
        const uint32_t id = lst.kId[d];
        const int orient = lst.kOrient[d];
        t.shift[q] = lst.kShift[d];
        t.len[q] = lst.kLen[d];
        t.w[q] = lst.kW[d];
        t.id[q] = id;
        t.orient[q] = orient;
        // the oriented row (the enc pitch holds at least RW words)
        const uint32_t* src = orient == 1 ? p.enc + p.pitch * id : p.rcenc + uint64_t(id) * uint64_t(2 * W);
        const int RW = tileRowWords(p.maxLen);
        stageRow<G>(src, t.rows + q * 2 * W, t.len[q], RW, p.varLen != 0, (W & 1) == 0 && (RW & 3) == 0 && (p.pitch & 3) == 0);
    }
    if(!p.useQ) return;
    __syncthreads();
    stageQualities<G>(p, t, tn, tid, nthreads);
}

// Quality rows of the tile's tn candidates (ids and orientations already in the tile).
template<int G>
__device__ __forceinline__ void stageQualities(const MsaParams& p, const MsaTile& t, int tn, int tid, int nthreads){
    const int qs = qStride(p.maxLen);
    const int nc = qs >> 4;              // chunks per row; device rows are qdpitch >= qs bytes
    // u / nc as a multiply-high: exact for u * nc < 2^32
    const uint64_t magic = ((uint64_t(1) << 32) + uint64_t(nc) - 1) / uint64_t(nc);
    if constexpr(G == 1){
        for(int u = tid; u < tn * nc; u += nthreads){
            const int j = int((uint64_t(u) * magic) >> 32);
            const int c = u - j * nc;
            if(p.varLen && !qualityChunkUsed(c, t.len[j], t.orient[j] != 1, int(p.qpitch))) continue;
            const uint4 v = *reinterpret_cast<const uint4*>(p.qual + p.qdpitch * t.id[j] + 16 * c);
            uint8_t* dst = t.q + j * qs;
            if(t.orient[j] == 1){
                *reinterpret_cast<uint4*>(dst + 16 * c) = v;
            }else{
                // byte s of the row goes to qs - 1 - s
                uint4 r;
                r.x = __builtin_bswap32(v.w); r.y = __builtin_bswap32(v.z);
                r.z = __builtin_bswap32(v.y); r.w = __builtin_bswap32(v.x);
                *reinterpret_cast<uint4*>(dst + qs - 16 * (c + 1)) = r;
            }
        }
        return;
    }
    // G chunks per thread in flight before their stores
    for(int base = tid; base < tn * nc; base += G * nthreads){
        uint4 v[G];
        int doff[G];           // byte offset of the chunk's slot in t.q, or -1
        bool rcv[G];
        #pragma unroll
        for(int i = 0; i < G; i++){
            const int u = base + i * nthreads;
            doff[i] = -1;
            rcv[i] = false;
            v[i] = make_uint4(0u, 0u, 0u, 0u);
            if(u >= tn * nc) continue;
            const int j = int((uint64_t(u) * magic) >> 32);
            const int c = u - j * nc;
            if(p.varLen && !qualityChunkUsed(c, t.len[j], t.orient[j] != 1, int(p.qpitch))) continue;
            rcv[i] = t.orient[j] != 1;
            // a reverse-complement chunk: byte s of the row goes to qs - 1 - s
            doff[i] = j * qs + (rcv[i] ? qs - 16 * (c + 1) : 16 * c);
            v[i] = *reinterpret_cast<const uint4*>(p.qual + p.qdpitch * t.id[j] + 16 * c);
        }
        #pragma unroll
        for(int i = 0; i < G; i++){
            if(doff[i] < 0) continue;
            uint4 r = v[i];
            if(rcv[i]){
                r.x = __builtin_bswap32(v[i].w); r.y = __builtin_bswap32(v[i].z);
                r.z = __builtin_bswap32(v[i].y); r.w = __builtin_bswap32(v[i].x);
            }
            *reinterpret_cast<uint4*>(t.q + doff[i]) = r;
        }
    }
}

// Working-list entry held by lane i of a wave for candidate i (lists of at most 64).
struct LaneCand {
    uint32_t id;
    int shift, len, orient;
    float w;
};

// stageTile for candidates held in lanes: lane tb+q stages candidate tb+q's oriented 2-bit
// row into slot q, and the quality chunks take their candidate's id and orientation from
// its lane (ds_bpermute, every lane active), so the row and quality loads are independent.
// Same rows and qualities as stageTile; the caller reads shift, length and weight from the
// lanes (the tile's metadata arrays are not written).
template<int G>
__device__ __forceinline__ void stageTileLanes(const MsaParams& p, const MsaTile& t, const LaneCand& c, int tb, int tn,
                                               int lane){
    const int W = p.W;
    const int q = lane - tb;
    if(q >= 0 && q < tn){
        const uint32_t* src = c.orient == 1 ? p.enc + p.pitch * c.id : p.rcenc + uint64_t(c.id) * uint64_t(2 * W);
        const int RW = tileRowWords(p.maxLen);
        stageRow<G>(src, t.rows + q * 2 * W, c.len, RW, p.varLen != 0, (W & 1) == 0 && (RW & 3) == 0 && (p.pitch & 3) == 0);
    }
    if(!p.useQ) return;
    const int qs = qStride(p.maxLen);
    const int nc = qs >> 4;              // chunks per row; device rows are qdpitch >= qs bytes
    const uint64_t magic = ((uint64_t(1) << 32) + uint64_t(nc) - 1) / uint64_t(nc);   // u / nc, exact here
    const int total = tn * nc;
    if constexpr(G == 1){
        for(int base = 0; base < total; base += kWave){
            const int u = base + lane;
            const int j = min(int((uint64_t(u) * magic) >> 32), tn - 1);
            const uint32_t id = uint32_t(__builtin_amdgcn_ds_bpermute(4 * (tb + j), int(c.id)));
            const int orient = __builtin_amdgcn_ds_bpermute(4 * (tb + j), c.orient);
            // reads of one length use every chunk (fixed length: no test, no third permute)
            const int len = p.varLen ? __builtin_amdgcn_ds_bpermute(4 * (tb + j), c.len) : int(p.qpitch);
            if(u < total && (!p.varLen || qualityChunkUsed(u - j * nc, len, orient != 1, int(p.qpitch)))){
                const int ch = u - j * nc;
                const uint4 v = *reinterpret_cast<const uint4*>(p.qual + p.qdpitch * id + 16 * ch);
                uint8_t* dst = t.q + j * qs;
                if(orient == 1){
                    *reinterpret_cast<uint4*>(dst + 16 * ch) = v;
                }else{
                    // byte s of the row goes to qs - 1 - s
                    uint4 r;
                    r.x = __builtin_bswap32(v.w); r.y = __builtin_bswap32(v.z);
                    r.z = __builtin_bswap32(v.y); r.w = __builtin_bswap32(v.x);
                    *reinterpret_cast<uint4*>(dst + qs - 16 * (ch + 1)) = r;
                }
            }
        }
        return;
    }
    // G chunks per lane in flight before their stores (one memory round trip per group)
    for(int base = 0; base < total; base += G * kWave){
        uint4 v[G];
        int doff[G];           // byte offset of the chunk's slot in t.q, or -1
        bool rcv[G];
        #pragma unroll
        for(int i = 0; i < G; i++){
            const int u = base + i * kWave + lane;
            const int j = min(int((uint64_t(u) * magic) >> 32), tn - 1);
            const uint32_t id = uint32_t(__builtin_amdgcn_ds_bpermute(4 * (tb + j), int(c.id)));
            const int orient = __builtin_amdgcn_ds_bpermute(4 * (tb + j), c.orient);
            // reads of one length use every chunk (fixed length: no test, no third permute)
            const int len = p.varLen ? __builtin_amdgcn_ds_bpermute(4 * (tb + j), c.len) : int(p.qpitch);
            const int ch = u - j * nc;
            const bool ok = u < total && (!p.varLen || qualityChunkUsed(ch, len, orient != 1, int(p.qpitch)));
            rcv[i] = orient != 1;
            // a reverse-complement chunk: byte s of the row goes to qs - 1 - s
            doff[i] = ok ? j * qs + (rcv[i] ? qs - 16 * (ch + 1) : 16 * ch) : -1;
            v[i] = ok ? *reinterpret_cast<const uint4*>(p.qual + p.qdpitch * id + 16 * ch) : make_uint4(0u, 0u, 0u, 0u);
        }
        #pragma unroll
        for(int i = 0; i < G; i++){
            if(doff[i] < 0) continue;
            uint4 r = v[i];
            if(rcv[i]){
                r.x = __builtin_bswap32(v[i].w); r.y = __builtin_bswap32(v[i].z);
                r.z = __builtin_bswap32(v[i].y); r.w = __builtin_bswap32(v[i].x);
            }
            *reinterpret_cast<uint4*>(t.q + doff[i]) = r;
        }
    }
}

// Builds the MSA of the anchor and the first `nk` working candidates in LDS
// (src/msa.cpp:9-173). Candidates are staged tile by tile (HiLo planes and qualities of
// the chosen orientation); each thread owns whole columns and adds the anchor, then the
// candidates in list order, so every column sum is formed in the reference's order.
// Returns false on an impossible column count.
__device__ bool buildMsa(const MsaParams& p, MsaShared& sh, uint64_t a, int La, const uint32_t* arow, int beg, int nk,
                         int* cnt, float* wgt, int* cov, float* sup, uint8_t* cons, int* origCov, const MsaTile& t){
    int lo = 0, hi = La;
    for(int j = threadIdx.x; j < nk; j += kBlock){
        const int s = p.kShift[beg + j];
        lo = min(lo, s);
        hi = max(hi, s + p.kLen[beg + j]);
    }
    lo = blockReduce(lo, sh.tmp, OpMin(), INT_MAX);
    hi = blockReduce(hi, sh.tmp, OpMax(), INT_MIN);
    const int ncol = hi - lo;
    const int ab = max(-lo, 0);
    if(ncol > p.maxCols) return false;
    if(threadIdx.x == 0){ sh.ncol = ncol; sh.ab = ab; }
    const uint8_t* aq = p.useQ ? p.qual + p.qdpitch * a : nullptr;
    // anchor first (weight 1.0f * quality weight)
    for(int c = threadIdx.x; c < ncol; c += kBlock){
        int n[4] = {0, 0, 0, 0};
        float w[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        int cv = 0;
        const int pa = c - ab;
        if(pa >= 0 && pa < La){
            const int b = baseAt(arow, pa);
            const float x = 1.0f * (p.useQ ? t.qw[aq[pa]] : 1.0f);
            #pragma unroll
            for(int q = 0; q < 4; q++){ if(q == b){ n[q]++; w[q] += x; } }
            cv = 1;
        }
        #pragma unroll
        for(int q = 0; q < 4; q++){ cnt[4 * c + q] = n[q]; wgt[4 * c + q] = w[q]; }
        cov[c] = cv;
    }
    const int W = p.W;
    const int T = p.tile;
    for(int tb = 0; tb < nk; tb += T){
        const int tn = min(T, nk - tb);
        __syncthreads();   // previous tile fully consumed
        stageTile<1>(p, listsOf(p), t, beg, tb, tn, threadIdx.x, kBlock);
        __syncthreads();
        for(int c = threadIdx.x; c < ncol; c += kBlock){
            int n0 = cnt[4 * c + 0], n1 = cnt[4 * c + 1], n2 = cnt[4 * c + 2], n3 = cnt[4 * c + 3];
            float w0 = wgt[4 * c + 0], w1 = wgt[4 * c + 1], w2 = wgt[4 * c + 2], w3 = wgt[4 * c + 3];
            int cv = cov[c];
            const int pa = c - ab;
            for(int j = 0; j < tn; j++){
                const int pos = pa - t.shift[j];
                if(pos >= 0 && pos < t.len[j]){
                    const int b = baseAt(t.rows + j * 2 * W, pos);
                    const int qo = t.orient[j] == 1 ? 0 : rcQualityOffset(p);
                    const float x = t.w[j] * (p.useQ ? t.qw[t.q[j * qStride(p.maxLen) + qo + pos]] : 1.0f);
                    if(b == 0){ n0++; w0 += x; } else if(b == 1){ n1++; w1 += x; }
                    else if(b == 2){ n2++; w2 += x; } else { n3++; w3 += x; }
                    cv++;
                }
            }
            cnt[4 * c + 0] = n0; cnt[4 * c + 1] = n1; cnt[4 * c + 2] = n2; cnt[4 * c + 3] = n3;
            wgt[4 * c + 0] = w0; wgt[4 * c + 1] = w1; wgt[4 * c + 2] = w2; wgt[4 * c + 3] = w3;
            cov[c] = cv;
        }
    }
    __syncthreads();
    for(int c = threadIdx.x; c < ncol; c += kBlock){
        const float w0 = wgt[4 * c + 0], w1 = wgt[4 * c + 1], w2 = wgt[4 * c + 2], w3 = wgt[4 * c + 3];
        // findConsensus (src/msa.cpp:135-157)
        int cb = 0;
        float cw = w0;
        if(w1 > cw){ cb = 1; cw = w1; }
        if(w2 > cw){ cb = 2; cw = w2; }
        if(w3 > cw){ cb = 3; cw = w3; }
        cons[c] = uint8_t(cb);
        const float colW = w0 + w1 + w2 + w3;
        sup[c] = cw / colW;
        // findOrigWeightAndCoverage (src/msa.cpp:159-173)
        const int pa = c - ab;
        if(pa >= 0 && pa < La) origCov[c] = cnt[4 * c + baseAt(arow, pa)];
    }
    __syncthreads();
    return true;
}

__global__ void __launch_bounds__(kBlock)
k_msa(MsaParams p){
    extern __shared__ uint8_t smem[];
    __shared__ MsaShared sh;
    const int C = p.maxCols;
    int* cnt = reinterpret_cast<int*>(smem);
    float* wgt = reinterpret_cast<float*>(cnt + 4 * C);
    int* cov = reinterpret_cast<int*>(wgt + 4 * C);
    float* sup = reinterpret_cast<float*>(cov + C);
    int* origCov = reinterpret_cast<int*>(sup + C);
    uint8_t* cons = reinterpret_cast<uint8_t*>(origCov + C);
    // 16-byte aligned sub-areas (same layout as msaLdsBytes)
    uint8_t* corr = smem + (size_t(45) * C + 15) / 16 * 16;
    __shared__ float sQw[256];
    MsaTile tile;
    {
        uint8_t* tp = corr + ((p.maxLen + 15) / 16) * 16;
        const int T = p.tile;
        tile.shift = reinterpret_cast<int*>(tp);
        tile.len = tile.shift + T;
        tile.w = reinterpret_cast<float*>(tile.len + T);
        tile.id = reinterpret_cast<uint32_t*>(tile.w + T);
        tile.orient = reinterpret_cast<int*>(tile.id + T);
        tile.rows = reinterpret_cast<uint32_t*>(tile.orient + T);
        tile.q = reinterpret_cast<uint8_t*>(tile.rows + T * 2 * p.W);
        tile.qw = sQw;
        if(p.useQ) for(int q = threadIdx.x; q < 256; q += kBlock) sQw[q] = p.qweight[q];
    }

    const int i = blockIdx.x;
    const uint64_t a = p.b0 + i;
    const int tid = threadIdx.x;
    const int La = p.len[a];
    const uint32_t* arow = p.enc + p.pitch * a;
    const int beg = int(p.offsets[i]);
    const int n0 = int(p.offsets[i + 1]) - beg;
    int* info = p.info ? p.info + size_t(i) * kAnchorInfo : nullptr;
    if(tid == 0){
        p.recBytes[i] = 0;
        if(info) for(int q = 0; q < kAnchorInfo; q++) info[q] = 0;
    }
    if(n0 == 0) return;

    // 1) drop candidates without an orientation (include/corrector.hpp:982-1050)
    int n1 = 0;
    for(int base = 0; base < n0; base += kBlock){
        const int j = base + tid;
        AlignBest b{};
        bool keep = false;
        if(j < n0){ b = p.best[beg + j]; keep = b.orientation != 3; }
        int total;
        const int pos = blockExclusiveSum(keep ? 1 : 0, sh.tmp, total);
        if(keep){
            const int d = beg + n1 + pos;
            const uint32_t id = p.ids[beg + j];
            p.kId[d] = id;
            p.kShift[d] = b.shift;
            p.kLen[d] = p.len[id];
            p.kNops[d] = b.nOps;
            p.kOvl[d] = b.overlap;
            p.kOrient[d] = uint8_t(b.orientation);
            p.kPaired[d] = p.paired ? p.paired[beg + j] : 0;
        }
        n1 += total;
    }
    if(n1 == 0) return;

    // 2) mismatch-ratio filter (include/corrector.hpp:1053-1149; paired :1151-1211)
    float threshold = -std::numeric_limits<float>::max();
    if(!p.isPaired){
        int c0 = 0, c1 = 0, c2 = 0;
        for(int j = tid; j < n1; j += kBlock){
            const float ratio = float(p.kNops[beg + j]) / float(p.kOvl[beg + j]);
            c0 += ratio < p.ratio2;
            c1 += ratio < p.ratio3;
            c2 += ratio < p.ratio4;
        }
        c0 = blockReduce(c0, sh.tmp, OpSum(), 0);
        c1 = blockReduce(c1, sh.tmp, OpSum(), 0);
        c2 = blockReduce(c2, sh.tmp, OpSum(), 0);
        if(c0 > 0 || c1 > 0 || c2 > 0){
            if(c0 >= p.goodCount) threshold = p.ratio2;
            else if(c1 >= p.goodCount) threshold = p.ratio3;
            else if(c2 >= p.goodCount) threshold = p.ratio4;
        }
    }
    int n2 = 0;
    for(int base = 0; base < n1; base += kBlock){
        const int j = base + tid;
        bool keep = false;
        uint32_t id = 0; int s = 0, l = 0, no = 0, ov = 0; uint8_t o = 0, pr = 0;
        if(j < n1){
            id = p.kId[beg + j]; s = p.kShift[beg + j]; l = p.kLen[beg + j]; no = p.kNops[beg + j];
            ov = p.kOvl[beg + j]; o = p.kOrient[beg + j]; pr = p.kPaired[beg + j];
            const float ratio = float(no) / float(ov);
            keep = p.isPaired ? (pr != 0 || ratio < p.pairedThreshold) : (ratio < threshold);
        }
        int total;
        const int pos = blockExclusiveSum(keep ? 1 : 0, sh.tmp, total);   // all reads of this chunk precede its writes
        if(keep){
            const int d = beg + n2 + pos;
            p.kId[d] = id; p.kShift[d] = s; p.kLen[d] = l; p.kNops[d] = no; p.kOvl[d] = ov;
            p.kOrient[d] = o; p.kPaired[d] = pr;
            // overlap weight (include/hostdevicefunctions.cuh:93-97)
            p.kW[d] = 1.0f - sqrtf(no / (ov * p.maxErrorRate));
        }
        n2 += total;
        __syncthreads();
    }
    if(info && tid == 0) info[0] = n2;
    if(n2 == 0) return;

    // 3) MSA and refinement (include/corrector.hpp:1307-1421; src/msa.cpp:532-734)
    int nk = n2;
    if(!buildMsa(p, sh, a, La, arow, beg, nk, cnt, wgt, cov, sup, cons, origCov, tile)){
        if(tid == 0) atomicAdd(p.error, 1);
        return;
    }
    int rounds = 0;
    for(int it = 0; it < 5; it++){
        const int ab = sh.ab;
        int mis = 0, colCand = INT_MAX;
        for(int q = tid; q < La; q += kBlock){
            const int c = ab + q;
            const int ci = cons[c];
            if(ci != baseAt(arow, q)) mis = 1;
            int sig = -1;
            for(int b = 0; b < 4; b++){
                if(b != ci && p.sigThr <= cnt[4 * c + b]) sig = b;
            }
            if(sig != -1) colCand = min(colCand, c);
        }
        mis = blockReduce(mis, sh.tmp, OpOr(), 0);
        if(!mis) break;
        const int col = blockReduce(colCand, sh.tmp, OpMin(), INT_MAX);
        if(col == INT_MAX) break;
        if(tid == 0){
            const int ci = cons[col];
            int sig = -1;
            for(int b = 0; b < 4; b++){
                if(b != ci && p.sigThr <= cnt[4 * col + b]) sig = b;
            }
            sh.sigBase = sig;
            sh.origBase = baseAt(arow, col - ab);
        }
        __syncthreads();
        const int sigBase = sh.sigBase;
        const bool keepMatching = sh.origBase == sigBase;
        int anyRemove = 0, veryGood = 0;
        for(int j = tid; j < nk; j += kBlock){
            const int rb = ab + p.kShift[beg + j];
            const int re = rb + p.kLen[beg + j];
            if(col >= rb && col < re){
                const int b = orientedBase(p, p.kId[beg + j], p.kLen[beg + j], p.kOrient[beg + j], col - rb);
                const bool rem = keepMatching ? (b != sigBase) : (b == sigBase);
                if(rem){
                    anyRemove = 1;
                    if(p.kW[beg + j] >= 0.90f) veryGood = 1;
                }
            }
        }
        anyRemove = blockReduce(anyRemove, sh.tmp, OpOr(), 0);
        veryGood = blockReduce(veryGood, sh.tmp, OpOr(), 0);
        // a round that removes nothing leaves the MSA, and hence every later round, unchanged
        if(!anyRemove || veryGood) break;
        int nn = 0;
        for(int base = 0; base < nk; base += kBlock){
            const int j = base + tid;
            bool keep = false;
            uint32_t id = 0; int s = 0, l = 0, no = 0, ov = 0; uint8_t o = 0, pr = 0; float w = 0.0f;
            if(j < nk){
                id = p.kId[beg + j]; s = p.kShift[beg + j]; l = p.kLen[beg + j]; no = p.kNops[beg + j];
                ov = p.kOvl[beg + j]; o = p.kOrient[beg + j]; pr = p.kPaired[beg + j]; w = p.kW[beg + j];
                const int rb = ab + s;
                keep = true;
                if(col >= rb && col < rb + l){
                    const int b = orientedBase(p, id, l, o, col - rb);
                    keep = !(keepMatching ? (b != sigBase) : (b == sigBase));
                }
            }
            int total;
            const int pos = blockExclusiveSum(keep ? 1 : 0, sh.tmp, total);
            if(keep){
                const int d = beg + nn + pos;
                p.kId[d] = id; p.kShift[d] = s; p.kLen[d] = l; p.kNops[d] = no; p.kOvl[d] = ov;
                p.kOrient[d] = o; p.kPaired[d] = pr; p.kW[d] = w;
            }
            nn += total;
            __syncthreads();
        }
        nk = nn;
        rounds++;
        if(!buildMsa(p, sh, a, La, arow, beg, nk, cnt, wgt, cov, sup, cons, origCov, tile)){
            if(tid == 0) atomicAdd(p.error, 1);
            return;
        }
    }
    const int ab = sh.ab;
    const int ncol = sh.ncol;
    if(p.msaDbg){
        uint8_t* d = p.msaDbg + size_t(i) * p.dbgStride;
        int* dn = reinterpret_cast<int*>(d);
        int* dc = dn + 1;
        float* dw = reinterpret_cast<float*>(dc + 4 * C);
        int* dv = reinterpret_cast<int*>(dw + 4 * C);
        float* ds = reinterpret_cast<float*>(dv + C);
        uint8_t* dk = reinterpret_cast<uint8_t*>(ds + C);
        if(tid == 0) dn[0] = ncol;
        for(int c = tid; c < ncol; c += kBlock){
            for(int b = 0; b < 4; b++){ dc[4 * c + b] = cnt[4 * c + b]; dw[4 * c + b] = wgt[4 * c + b]; }
            dv[c] = cov[c]; ds[c] = sup[c]; dk[c] = "ACGT"[cons[c]];
        }
    }
    if(info && tid == 0){ info[1] = nk; info[2] = ncol; info[3] = ab; info[8] = rounds; }
    // msa.nCandidates == 0: cannot be corrected (src/msa.cpp:388-394)
    if(nk == 0) return;

    // 4) properties and decision (src/msa.cpp:304-453)
    float ms = INFINITY;
    int mincov = INT_MAX, maxcov = INT_MIN;
    for(int q = tid; q < La; q += kBlock){
        const int c = ab + q;
        ms = sup[c] < ms ? sup[c] : ms;
        mincov = min(mincov, cov[c]);
        maxcov = max(maxcov, cov[c]);
    }
    const float minSupport = blockMinF(ms, sh.ftmp);
    mincov = blockReduce(mincov, sh.tmp, OpMin(), INT_MAX);
    maxcov = blockReduce(maxcov, sh.tmp, OpMax(), INT_MIN);
    if(tid == 0){
        float sum = 0.0f;                       // sequential, as std::accumulate
        for(int q = 0; q < La; q++) sum += sup[ab + q];
        sh.ftmp[0] = sum / La;
    }
    __syncthreads();
    const float avgSupport = sh.ftmp[0];
    __syncthreads();
    const bool allGood = fgeq(avgSupport, p.avgThr) && fgeq(minSupport, p.minThr) && fgeq(float(mincov), p.minCovThr);
    bool isHQ = false;
    if(allGood){
        // src/msa.cpp:357-372 takes the smallest percent in [0, pct] whose thresholds
        // 1 - f and 1 - 3f (f = percent / 100) both supports meet and calls the anchor HQ
        // when it is at most pct / 2. The thresholds fall as the percent grows and fgeq
        // stays true as its bound falls, so that holds exactly when the supports meet the
        // thresholds of hqPercent, the largest percent that is at most pct / 2 (the host
        // computes both with the same float expressions).
        isHQ = fgeq(float(mincov), p.minCovThr) && fgeq(avgSupport, p.hqAvgT) && fgeq(minSupport, p.hqMinT);
    }
    int nEdits = 0;
    for(int q = tid; q < La; q += kBlock){
        const int c = ab + q;
        const int ob = baseAt(arow, q);
        int cb;
        if(allGood) cb = cons[c];
        else cb = (sup[c] > 0.90f && origCov[c] <= 2) ? cons[c] : ob;
        corr[q] = uint8_t(cb);
        nEdits += cb != ob;
    }
    nEdits = blockReduce(nEdits, sh.tmp, OpSum(), 0);
    if(info && tid == 0){
        info[4] = isHQ; info[5] = 1; info[6] = mincov; info[7] = maxcov;
        info[9] = __float_as_int(avgSupport); info[10] = __float_as_int(minSupport);
    }

    // 5) record (include/corrector.hpp:1639-1761, include/correctedsequence.hpp:299-387);
    // HQ records without edits are not stored (src/correct_cpu.cpp:263-273)
    const bool ambiguous = p.ambig[a] != 0;
    const bool useEdits = !ambiguous && nEdits <= La / 7;
    if(isHQ && useEdits && nEdits == 0) return;
    const uint32_t payload = useEdits ? uint32_t(4 + 5 * nEdits) : uint32_t(4 + La);
    uint8_t* rec = p.slots + size_t(i) * p.slotBytes;
    if(tid == 0){
        const uint32_t rid = uint32_t(a);
        const uint32_t flags = (uint32_t(isHQ) << 31) | (uint32_t(useEdits) << 30) | payload;
        memcpy(rec, &rid, 4);
        memcpy(rec + 4, &flags, 4);
        const int first = useEdits ? nEdits : La;
        memcpy(rec + 8, &first, 4);
        p.recBytes[i] = 8 + payload;
        if(info) info[11] = int(8 + payload);
    }
    if(useEdits){
        int running = 0;
        for(int base = 0; base < La; base += kBlock){
            const int q = base + tid;
            bool e = false;
            int cb = 0;
            if(q < La){ cb = corr[q]; e = cb != baseAt(arow, q); }
            int total;
            const int pos = blockExclusiveSum(e ? 1 : 0, sh.tmp, total);
            if(e){
                const int k = running + pos;
                memcpy(rec + 12 + 4 * k, &q, 4);
                rec[12 + 4 * nEdits + k] = uint8_t("ACGT"[cb]);
            }
            running += total;
        }
    }else{
        for(int q = tid; q < La; q += kBlock) rec[12 + q] = uint8_t("ACGT"[corr[q]]);
    }
}

// ---------------------------------------------------------------- MSA, one wavefront per anchor

// Inclusive Hillis-Steele pattern over the wave with DPP (row shifts 1,2,4,8, then the
// row broadcasts of lanes 15 and 31); lane 63 ends with the reduction of all lanes.
template<class Op>
__device__ __forceinline__ int waveReduce(int v, Op op, int identity){
    int t;
    t = __builtin_amdgcn_update_dpp(identity, v, 0x111, 0xf, 0xf, false); v = op(v, t);
    t = __builtin_amdgcn_update_dpp(identity, v, 0x112, 0xf, 0xf, false); v = op(v, t);
    t = __builtin_amdgcn_update_dpp(identity, v, 0x114, 0xf, 0xf, false); v = op(v, t);
    t = __builtin_amdgcn_update_dpp(identity, v, 0x118, 0xf, 0xf, false); v = op(v, t);
    t = __builtin_amdgcn_update_dpp(identity, v, 0x142, 0xa, 0xf, false); v = op(v, t);
    t = __builtin_amdgcn_update_dpp(identity, v, 0x143, 0xc, 0xf, false); v = op(v, t);
    return __builtin_amdgcn_readlane(v, 63);
}

// pointers into LDS (address space 3, 32-bit)
typedef const __attribute__((address_space(3))) uint32_t* LdsU32;
typedef const __attribute__((address_space(3))) uint8_t* LdsU8;

__device__ __forceinline__ uint64_t lanesBelow(int lane){ return lane ? (~0ull >> (64 - lane)) : 0ull; }

// Per-lane MSA columns: lane owns columns lane + 64*k, k < KC.
template<int KC>
struct WaveMsa {
    float w[KC][4];
    uint64_t cnt[KC];   // counts of A, C, G, T in 16-bit fields (low to high)
    float sup[KC];
    uint32_t cons;      // 2 bits per slot
    uint32_t abase;     // 2 bits per slot: anchor base of anchor columns
    uint32_t isAnchor;  // 1 bit per slot
    int ncol, ab;
};

// weight of base b (0..3) in slot k, by selects: a dynamic index would move the
// register-resident columns to scratch memory
template<int KC>
__device__ __forceinline__ float weightOf(const WaveMsa<KC>& m, int k, int b){
    return b == 0 ? m.w[k][0] : (b == 1 ? m.w[k][1] : (b == 2 ? m.w[k][2] : m.w[k][3]));
}

template<int KC>
__device__ __forceinline__ int countOf(const WaveMsa<KC>& m, int k, int b){
    return int(uint32_t(m.cnt[k] >> (16 * b)) & 0xFFFFu);
}

// MSA of the anchor and the first nk working candidates (src/msa.cpp:9-173), one wave:
// every lane adds the anchor then the candidates, in list order, to the columns it owns.
// AO (anchor columns only): slot column c is anchor position c, and only the columns
// under the anchor are built; every later step of the plain correction (refinement,
// properties, decision, features) reads no other column. Otherwise slot column c is MSA
// column c (anchor position c - ab), all ncol columns. QM: 0 = no quality scores,
// 1 = quality scores, 2 = as p.useQ says at run time.
template<int KC, bool AO, int QM>
__device__ __forceinline__ bool buildMsaWave(const MsaParams& p, const AnchorLists& lst, uint64_t a, int La,
                                             const uint32_t* arow, int beg, int nk,
                             WaveMsa<KC>& m, const MsaTile& t, int lane, bool inLanes = false,
                             const LaneCand& lc = LaneCand{}){
    // inLanes (wave-uniform, nk <= 64): lane i holds candidate i of the working list, which
    // the list in memory holds as well
    int lo = 0, hi = La;
    for(int j = lane; j < nk; j += kWave){
        const int s = inLanes ? lc.shift : lst.kShift[beg + j];
        lo = min(lo, s);
        hi = max(hi, s + (inLanes ? lc.len : lst.kLen[beg + j]));
    }
    lo = waveReduce(lo, OpMin(), INT_MAX);
    hi = waveReduce(hi, OpMax(), INT_MIN);
    const int ncol = hi - lo;
    const int ab = max(-lo, 0);
    if((AO ? La : ncol) > KC * kWave || ncol > p.maxCols) return false;
    m.ncol = ncol;
    m.ab = ab;
    m.abase = 0;
    m.isAnchor = 0;
    const int off = AO ? 0 : ab;       // slot column of anchor position 0
    const bool useQ = QM == 2 ? p.useQ != 0 : QM == 1;
    const uint8_t* aq = useQ ? p.qual + p.qdpitch * a : nullptr;
    #pragma unroll
    for(int k = 0; k < KC; k++){
        #pragma unroll
        for(int b = 0; b < 4; b++) m.w[k][b] = 0.0f;
        m.cnt[k] = 0;
        const int pa = lane + kWave * k - off;
        if(pa >= 0 && pa < La){
            const int b = baseAt(arow, pa);
            const float x = 1.0f * (useQ ? t.qw[aq[pa]] : 1.0f);
            m.w[k][0] = b == 0 ? x : 0.0f;
            m.w[k][1] = b == 1 ? x : 0.0f;
            m.w[k][2] = b == 2 ? x : 0.0f;
            m.w[k][3] = b == 3 ? x : 0.0f;
            m.cnt[k] = uint64_t(1) << (16 * b);
            m.abase |= uint32_t(b) << (2 * k);
            m.isAnchor |= 1u << k;
        }
    }
    const int RS = 2 * p.W;                  // tile row stride in words
    const int T = p.tile;
    const int qs = qStride(p.maxLen);
    const int qrc = rcQualityOffset(p);
    for(int tb = 0; tb < nk; tb += T){
        const int tn = min(T, nk - tb);
        __syncthreads();
        // two loads in flight per lane from 3 column slots on (at 2 slots the 8-wave budget
        // leaves no room)
        constexpr int G = KC >= 3 ? 2 : 1;
        if(inLanes) stageTileLanes<G>(p, t, lc, tb, tn, lane);
        else stageTile<G>(p, lst, t, beg, tb, tn, lane, kWave);
        __syncthreads();
        // lane src+q holds candidate tb+q's shift, length, weight and orientation; the loop
        // below reads them with readlane (no LDS round trip per candidate)
        int mShift = 0, mLen = 0, mRc = 0, src = tb;
        float mW = 0.0f;
        if(inLanes){
            mShift = lc.shift; mLen = lc.len; mW = lc.w; mRc = lc.orient != 1;
        }else{
            src = 0;
            if(lane < tn){ mShift = t.shift[lane]; mLen = t.len[lane]; mW = t.w[lane]; mRc = t.orient[lane] != 1; }
        }
        for(int j = 0; j < tn; j++){
            const int cs = off + __builtin_amdgcn_readlane(mShift, src + j);   // slot column of its first base
            const int ln = __builtin_amdgcn_readlane(mLen, src + j);
            const float wj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mW), src + j));
            const bool rc = __builtin_amdgcn_readlane(mRc, src + j) != 0;
            // slots holding its columns (AO: those under the anchor; lanes past the anchor
            // end add to columns that are never read)
            const int k0 = (AO ? max(cs, 0) : cs) >> 6;
            const int k1 = ((AO ? min(cs + ln, La) : cs + ln) - 1) >> 6;
            // slot k covers candidate position pos0 + 64k: base word w0 + 4k at the same bit
            // offset, quality byte q0 + 64k
            const int pos0 = lane - cs;
            const int sA = 30 - 2 * (pos0 & 15);
            // LDS addresses of slot 0's base word and quality byte, kept in registers (the
            // compiler would otherwise rebuild them inside every slot's branch; the empty asm
            // statement emits no instruction). Slot k reads rowp[4k] and qp[64k], which lie
            // in candidate j's row whenever the slot's position is inside the candidate.
            LdsU32 rowp = (LdsU32)(t.rows) + (j * RS + (pos0 >> 4));
            LdsU8 qp = (LdsU8)(t.q) + (j * qs + (rc ? qrc : 0) + pos0);
            asm volatile("" : "+v"(rowp), "+v"(qp));
            #pragma unroll
            for(int k = 0; k < KC; k++){
                if(k < k0 || k > k1) continue;        // wave-uniform
                const int pos = pos0 + kWave * k;
                if(unsigned(pos) < unsigned(ln)){
                    const int b = int((rowp[4 * k] >> sA) & 3u);
                    const float x = wj * (useQ ? t.qw[qp[kWave * k]] : 1.0f);
                    // inc holds 1 in the 16-bit field of base b, so bytes 0 and 2 of its
                    // halves are the one-hot factors of A, C (low half) and G, T (high half).
                    // fma(x, 1, w) rounds w + x exactly like the add and fma(x, 0, w) = w (no
                    // weight is ever -0.0: weights start at +0.0f and add finite products),
                    // so each weight is the sum of its base's x in list order, branch-free.
                    const uint64_t inc = uint64_t(1) << (16 * b);
                    const uint32_t incLo = uint32_t(inc), incHi = uint32_t(inc >> 32);
                    m.w[k][0] = __builtin_fmaf(x, float(incLo & 0xFFu), m.w[k][0]);
                    m.w[k][1] = __builtin_fmaf(x, float((incLo >> 16) & 0xFFu), m.w[k][1]);
                    m.w[k][2] = __builtin_fmaf(x, float(incHi & 0xFFu), m.w[k][2]);
                    m.w[k][3] = __builtin_fmaf(x, float((incHi >> 16) & 0xFFu), m.w[k][3]);
                    m.cnt[k] += inc;
                }
            }
        }
    }
    // consensus and support (src/msa.cpp:135-157)
    m.cons = 0;
    #pragma unroll
    for(int k = 0; k < KC; k++){
        int cb = 0;
        float cw = m.w[k][0];
        if(m.w[k][1] > cw){ cb = 1; cw = m.w[k][1]; }
        if(m.w[k][2] > cw){ cb = 2; cw = m.w[k][2]; }
        if(m.w[k][3] > cw){ cb = 3; cw = m.w[k][3]; }
        m.cons |= uint32_t(cb) << (2 * k);
        const float colW = m.w[k][0] + m.w[k][1] + m.w[k][2] + m.w[k][3];
        const bool inMsa = AO ? ((m.isAnchor >> k) & 1u) != 0 : lane + kWave * k < ncol;
        m.sup[k] = inMsa ? cw / colW : 0.0f;
    }
    return true;
}

// float -> bfloat16 -> float, round to nearest even (finite inputs)
__device__ __forceinline__ float bf16r(float x){
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xFFFF0000u);
}
__device__ __forceinline__ float bf16f(uint16_t h){ return __uint_as_float(uint32_t(h) << 16); }

// Logit of the feature row sx (bfloat16 bits) by the whole wave: lane j computes hidden
// unit j of both layers, each sum starting at the bias and adding the inputs in index
// order; the output sum runs over the hidden units in order on every lane. Weights are
// bfloat16, sums float. Returns the wave-uniform logit.
__device__ __forceinline__ float seamLogit(const MsaParams& p, const uint16_t* sx, float* sH1, float* sH2, int lane){
    const uint16_t* w1t = p.seamW;
    const uint16_t* w2t = p.seamW + kFeat * kHid;
    const uint16_t* w3 = w2t + kHid * kHid;
    float a = p.seamB[lane];
    #pragma unroll
    for(int k = 0; k < kFeat; k++) a = a + bf16f(w1t[k * kHid + lane]) * bf16f(sx[k]);
    sH1[lane] = bf16r(a > 0.0f ? a : 0.0f);
    __syncthreads();
    a = p.seamB[kHid + lane];
    #pragma unroll 16
    for(int k = 0; k < kHid; k++) a = a + bf16f(w2t[k * kHid + lane]) * sH1[k];
    sH2[lane] = bf16r(a > 0.0f ? a : 0.0f);
    __syncthreads();
    float z = p.seamB[2 * kHid];
    #pragma unroll 16
    for(int j = 0; j < kHid; j++) z = z + bf16f(w3[j]) * sH2[j];
    return z;
}

// Filters, MSA, refinement, decision and record of one anchor per wavefront
// (include/corrector.hpp:982-1446,1639-1761; src/msa.cpp:9-453,532-734).
// Carves the k_msa_wave / k_fused candidate tile and the anchor-support array out of `tp`
// (layout sized by waveMsaLdsBytes).
__device__ __forceinline__ void carveWaveTile(const MsaParams& p, uint8_t* tp, const float* qw, MsaTile& tile, float*& sSup){
    const int T = p.tile;
    tile.shift = reinterpret_cast<int*>(tp);
    tile.len = tile.shift + T;
    tile.w = reinterpret_cast<float*>(tile.len + T);
    tile.id = reinterpret_cast<uint32_t*>(tile.w + T);
    tile.orient = reinterpret_cast<int*>(tile.id + T);
    tile.rows = reinterpret_cast<uint32_t*>(tile.orient + T);
    tile.q = reinterpret_cast<uint8_t*>(tile.rows + T * 2 * p.W);
    tile.qw = qw;
    sSup = reinterpret_cast<float*>(tp + waveTileRegion(T, p.W, p.maxLen, p.useQ != 0, p.seamMode != 0 || p.featOut,
                                                        p.maxCols));
}

// Candidate correction of an HQ anchor (correctCandidatesClassic, include/corrector.hpp:
// 1504-1512, src/msa.cpp:457-522, records as makeOutputOfTask, :1676-1757): every final
// candidate that starts within newCols columns of the anchor and ends at most newCols
// columns after it takes the consensus over its columns (the support and coverage test
// there is overwritten with true); a reverse-complement candidate's correction is turned
// forward; the record holds edits against the stored read when there are at most
// length/7 and the read has no N, else the sequence, then the candidate's shift. One
// candidate at a time, lanes over positions; the consensus is staged into the free tile
// region.
template<int KC>
__device__ __forceinline__ void emitCandidateRecords(const MsaParams& p, const AnchorLists& lst, uint32_t cons, int ncol,
                                                 int ab, int beg, int nk, int La, const MsaTile& tile, int lane){
    uint8_t* sCons = reinterpret_cast<uint8_t*>(tile.shift);
    __syncthreads();
    #pragma unroll
    for(int k = 0; k < KC; k++){
        const int c = lane + kWave * k;
        if(c < ncol) sCons[c] = uint8_t((cons >> (2 * k)) & 3u);
    }
    __syncthreads();
    const uint64_t below = lanesBelow(lane);
    for(int j = 0; j < nk; j++){
        const int shift = lst.kShift[beg + j];
        const int Lc = lst.kLen[beg + j];
        const int qb = ab + shift, qe = qb + Lc;
        if(!(ab - p.newCols <= qb && qb <= ab + p.newCols && qe <= ab + La + p.newCols)) continue;
        const uint32_t id = lst.kId[beg + j];
        const bool rc = lst.kOrient[beg + j] == 2;
        const uint32_t* crow = p.enc + p.pitch * id;
        int D = 0;
        for(int base = 0; base < Lc; base += kWave){
            const int pos = base + lane;
            bool diff = false;
            if(pos < Lc){
                const int cb = rc ? 3 - int(sCons[qe - 1 - pos]) : int(sCons[qb + pos]);
                diff = cb != baseAt(crow, pos);
            }
            D += __popcll(__ballot(diff));
        }
        const bool useEdits = p.ambig[id] == 0 && D <= Lc / 7;
        const uint32_t payload = (useEdits ? uint32_t(4 + 5 * D) : uint32_t(4 + Lc)) + 4;
        uint8_t* rec = p.candSlots + size_t(beg + j) * p.candSlotBytes;
        if(lane == 0){
            const uint32_t flags = (uint32_t(useEdits) << 30) | (1u << 29) | payload;
            const int count = useEdits ? D : Lc;
            memcpy(rec, &id, 4);
            memcpy(rec + 4, &flags, 4);
            memcpy(rec + 8, &count, 4);
            memcpy(rec + 8 + payload - 4, &shift, 4);
            p.candBytes[beg + j] = 8 + payload;
        }
        int running = 0;
        for(int base = 0; base < Lc; base += kWave){
            const int pos = base + lane;
            int cb = 0;
            bool diff = false;
            if(pos < Lc){
                cb = rc ? 3 - int(sCons[qe - 1 - pos]) : int(sCons[qb + pos]);
                diff = cb != baseAt(crow, pos);
            }
            if(useEdits){
                const uint64_t mask = __ballot(diff);
                if(diff){
                    const int r = running + __popcll(mask & below);
                    memcpy(rec + 12 + 4 * r, &pos, 4);
                    rec[12 + 4 * D + r] = uint8_t("ACGT"[cb]);
                }
                running += __popcll(mask);
            }else if(pos < Lc){
                rec[12 + pos] = uint8_t("ACGT"[cb]);
            }
        }
    }
}

// Filters, MSA, refinement, decision and record of anchor `a` by one wavefront
// (include/corrector.hpp:982-1446,1639-1761; src/msa.cpp:9-453,532-734). Its candidates
// are lst.ids/lst.best/lst.paired[beg, beg+n0); the record goes to slot `i`.
template<int KC, bool AO, int QM>
__device__ __forceinline__ void msaAnchorWave(const MsaParams& p, const AnchorLists& lst, uint64_t a, int i, int beg,
                                              int n0, const MsaTile& tile,
                              float* sSup, int lane){
    const int La = p.len[a];
    const uint32_t* arow = p.enc + p.pitch * a;
    int* info = p.info ? p.info + size_t(i) * kAnchorInfo : nullptr;
    if(lane == 0) p.recBytes[i] = 0;
    if(info && lane < kAnchorInfo) info[lane] = 0;
    if(p.corrCands){
        for(int q = beg + lane; q < beg + n0; q += kWave) p.candBytes[q] = 0;
        if(lane == 0) p.hqAnchor[a] = 0;
    }
    if(n0 == 0) return;
    const uint64_t below = lanesBelow(lane);

    // 1) drop candidates without an orientation (include/corrector.hpp:982-1050), then
    // 2) the mismatch-ratio filter (include/corrector.hpp:1053-1149; paired :1151-1211)
    int n2 = 0;
    LaneCand lc{};              // lists of at most 64: lane i's copy of working candidate i
    const bool inLanes = n0 <= kWave;
    if(inLanes){
        // at most one candidate per lane: both filters on the lane's registers and one
        // compaction, no intermediate list
        AlignBest b{};
        bool k1 = false;
        if(lane < n0){ b = lst.best[beg + lane]; k1 = b.orientation != 3; }
        if(__ballot(k1) == 0) return;
        const float ratio = k1 ? float(b.nOps) / float(b.overlap) : 1e30f;
        float threshold = -std::numeric_limits<float>::max();
        if(!p.isPaired){
            const int c0 = __popcll(__ballot(k1 && ratio < p.ratio2));
            const int c1 = __popcll(__ballot(k1 && ratio < p.ratio3));
            const int c2 = __popcll(__ballot(k1 && ratio < p.ratio4));
            if(c0 > 0 || c1 > 0 || c2 > 0){
                if(c0 >= p.goodCount) threshold = p.ratio2;
                else if(c1 >= p.goodCount) threshold = p.ratio3;
                else if(c2 >= p.goodCount) threshold = p.ratio4;
            }
        }
        const uint8_t pr = (k1 && lst.paired) ? lst.paired[beg + lane] : 0;
        const bool keep = k1 && (p.isPaired ? (pr != 0 || ratio < p.pairedThreshold) : (ratio < threshold));
        const uint64_t mask = __ballot(keep);
        n2 = __popcll(mask);
        uint32_t id = 0;
        int L = 0;
        float w = 0.0f;
        if(keep){
            const int d = beg + __popcll(mask & below);
            id = lst.ids[beg + lane];
            L = p.len[id];
            const int no = b.nOps, ov = b.overlap;
            // overlap weight (include/hostdevicefunctions.cuh:93-97)
            w = 1.0f - sqrtf(no / (ov * p.maxErrorRate));
            lst.kId[d] = id; lst.kShift[d] = b.shift; lst.kLen[d] = L; lst.kNops[d] = no; lst.kOvl[d] = ov;
            lst.kOrient[d] = uint8_t(b.orientation); lst.kPaired[d] = pr;
            lst.kW[d] = w;
        }
        // the same entries into lanes 0..n2-1 (kept lanes in order, the others after them:
        // a permutation of the 64 lanes)
        const int dl = keep ? __popcll(mask & below) : n2 + __popcll(~mask & below);
        lc.id = uint32_t(__builtin_amdgcn_ds_permute(4 * dl, int(id)));
        lc.shift = __builtin_amdgcn_ds_permute(4 * dl, b.shift);
        lc.len = __builtin_amdgcn_ds_permute(4 * dl, L);
        lc.orient = __builtin_amdgcn_ds_permute(4 * dl, b.orientation);
        lc.w = __int_as_float(__builtin_amdgcn_ds_permute(4 * dl, __float_as_int(w)));
    }else{
        // more candidates than lanes: step 1 into the working list, then step 2 over it
        int n1 = 0;
        for(int base = 0; base < n0; base += kWave){
            const int j = base + lane;
            AlignBest b{};
            bool keep = false;
            if(j < n0){ b = lst.best[beg + j]; keep = b.orientation != 3; }
            const uint64_t mask = __ballot(keep);
            if(keep){
                const int d = beg + n1 + __popcll(mask & below);
                const uint32_t id = lst.ids[beg + j];
                lst.kId[d] = id;
                lst.kShift[d] = b.shift;
                lst.kLen[d] = p.len[id];
                lst.kNops[d] = b.nOps;
                lst.kOvl[d] = b.overlap;
                lst.kOrient[d] = uint8_t(b.orientation);
                lst.kPaired[d] = lst.paired ? lst.paired[beg + j] : 0;
            }
            n1 += __popcll(mask);
        }
        if(n1 == 0) return;
        float threshold = -std::numeric_limits<float>::max();
        if(!p.isPaired){
            int c0 = 0, c1 = 0, c2 = 0;
            for(int base = 0; base < n1; base += kWave){
                const int j = base + lane;
                float ratio = 1e30f;
                if(j < n1) ratio = float(lst.kNops[beg + j]) / float(lst.kOvl[beg + j]);
                c0 += __popcll(__ballot(j < n1 && ratio < p.ratio2));
                c1 += __popcll(__ballot(j < n1 && ratio < p.ratio3));
                c2 += __popcll(__ballot(j < n1 && ratio < p.ratio4));
            }
            if(c0 > 0 || c1 > 0 || c2 > 0){
                if(c0 >= p.goodCount) threshold = p.ratio2;
                else if(c1 >= p.goodCount) threshold = p.ratio3;
                else if(c2 >= p.goodCount) threshold = p.ratio4;
            }
        }
        for(int base = 0; base < n1; base += kWave){
            const int j = base + lane;
            bool keep = false;
            uint32_t id = 0; int s = 0, l = 0, no = 0, ov = 0; uint8_t o = 0, pr = 0;
            if(j < n1){
                id = lst.kId[beg + j]; s = lst.kShift[beg + j]; l = lst.kLen[beg + j]; no = lst.kNops[beg + j];
                ov = lst.kOvl[beg + j]; o = lst.kOrient[beg + j]; pr = lst.kPaired[beg + j];
                const float ratio = float(no) / float(ov);
                keep = p.isPaired ? (pr != 0 || ratio < p.pairedThreshold) : (ratio < threshold);
            }
            const uint64_t mask = __ballot(keep);
            if(keep){
                const int d = beg + n2 + __popcll(mask & below);
                lst.kId[d] = id; lst.kShift[d] = s; lst.kLen[d] = l; lst.kNops[d] = no; lst.kOvl[d] = ov;
                lst.kOrient[d] = o; lst.kPaired[d] = pr;
                // overlap weight (include/hostdevicefunctions.cuh:93-97)
                lst.kW[d] = 1.0f - sqrtf(no / (ov * p.maxErrorRate));
            }
            n2 += __popcll(mask);
        }
    }
    if(info && lane == 0) info[0] = n2;
    if(n2 == 0) return;

    // 3) MSA and refinement (include/corrector.hpp:1307-1421; src/msa.cpp:532-734)
    WaveMsa<KC> m;
    int nk = n2;
    if(!buildMsaWave<KC, AO, QM>(p, lst, a, La, arow, beg, nk, m, tile, lane, inLanes, lc)){
        if(lane == 0) atomicAdd(p.error, 1);
        return;
    }
    int rounds = 0;
    for(int it = 0; it < 5; it++){
        bool mis = false;
        int colCand = INT_MAX;
        #pragma unroll
        for(int k = 0; k < KC; k++){
            if(!((m.isAnchor >> k) & 1u)) continue;
            const int ci = int((m.cons >> (2 * k)) & 3u);
            if(ci != int((m.abase >> (2 * k)) & 3u)) mis = true;
            int sig = -1;
            for(int b = 0; b < 4; b++){
                if(b != ci && p.sigThr <= countOf(m, k, b)) sig = b;
            }
            if(sig != -1) colCand = min(colCand, lane + kWave * k);
        }
        if(__ballot(mis) == 0) break;
        const int scol = waveReduce(colCand, OpMin(), INT_MAX);     // slot column
        if(scol == INT_MAX) break;
        const int col = scol + (AO ? m.ab : 0);                     // MSA column
        int mySig = -1;
        #pragma unroll
        for(int k = 0; k < KC; k++){
            if(k == (scol >> 6)){
                const int ci = int((m.cons >> (2 * k)) & 3u);
                for(int b = 0; b < 4; b++){
                    if(b != ci && p.sigThr <= countOf(m, k, b)) mySig = b;
                }
            }
        }
        const int sigBase = __builtin_amdgcn_readlane(mySig, scol & 63);
        const int origBase = baseAt(arow, col - m.ab);
        const bool keepMatching = origBase == sigBase;
        bool anyRemove = false, veryGood = false;
        for(int base = 0; base < nk; base += kWave){
            const int j = base + lane;
            bool rem = false;
            if(j < nk){
                const int rb = m.ab + lst.kShift[beg + j];
                const int l = lst.kLen[beg + j];
                if(col >= rb && col < rb + l){
                    const int b = orientedBase(p, lst.kId[beg + j], l, lst.kOrient[beg + j], col - rb);
                    rem = keepMatching ? (b != sigBase) : (b == sigBase);
                }
            }
            anyRemove |= __ballot(rem) != 0;
            veryGood |= __ballot(rem && lst.kW[beg + j] >= 0.90f) != 0;
        }
        // a round that removes nothing leaves the MSA, and hence every later round, unchanged
        if(!anyRemove || veryGood) break;
        int nn = 0;
        for(int base = 0; base < nk; base += kWave){
            const int j = base + lane;
            bool keep = false;
            uint32_t id = 0; int s = 0, l = 0, no = 0, ov = 0; uint8_t o = 0, pr = 0; float w = 0.0f;
            if(j < nk){
                id = lst.kId[beg + j]; s = lst.kShift[beg + j]; l = lst.kLen[beg + j]; no = lst.kNops[beg + j];
                ov = lst.kOvl[beg + j]; o = lst.kOrient[beg + j]; pr = lst.kPaired[beg + j]; w = lst.kW[beg + j];
                const int rb = m.ab + s;
                keep = true;
                if(col >= rb && col < rb + l){
                    const int b = orientedBase(p, id, l, o, col - rb);
                    keep = !(keepMatching ? (b != sigBase) : (b == sigBase));
                }
            }
            const uint64_t mask = __ballot(keep);
            if(keep){
                const int d = beg + nn + __popcll(mask & below);
                lst.kId[d] = id; lst.kShift[d] = s; lst.kLen[d] = l; lst.kNops[d] = no; lst.kOvl[d] = ov;
                lst.kOrient[d] = o; lst.kPaired[d] = pr; lst.kW[d] = w;
            }
            nn += __popcll(mask);
        }
        nk = nn;
        rounds++;
        if(!buildMsaWave<KC, AO, QM>(p, lst, a, La, arow, beg, nk, m, tile, lane)){
            if(lane == 0) atomicAdd(p.error, 1);
            return;
        }
    }
    const int ab = m.ab;
    const int ncol = m.ncol;
    const int off = AO ? 0 : ab;       // slot column of anchor position 0
    if(!AO && p.msaDbg){
        uint8_t* d = p.msaDbg + size_t(i) * p.dbgStride;
        const int C = p.maxCols;
        int* dn = reinterpret_cast<int*>(d);
        int* dc = dn + 1;
        float* dw = reinterpret_cast<float*>(dc + 4 * C);
        int* dv = reinterpret_cast<int*>(dw + 4 * C);
        float* ds = reinterpret_cast<float*>(dv + C);
        uint8_t* dk = reinterpret_cast<uint8_t*>(ds + C);
        if(lane == 0) dn[0] = ncol;
        #pragma unroll
        for(int k = 0; k < KC; k++){
            const int c = lane + kWave * k;
            if(c < ncol){
                int cv = 0;
                for(int b = 0; b < 4; b++){ dc[4 * c + b] = countOf(m, k, b); dw[4 * c + b] = m.w[k][b]; cv += countOf(m, k, b); }
                dv[c] = cv; ds[c] = m.sup[k]; dk[c] = "ACGT"[(m.cons >> (2 * k)) & 3u];
            }
        }
    }
    if(info && lane == 0){ info[1] = nk; info[2] = ncol; info[3] = ab; info[8] = rounds; }
    // msa.nCandidates == 0: cannot be corrected (src/msa.cpp:388-394)
    if(nk == 0) return;

    // 4) properties and decision (src/msa.cpp:304-453)
    int minSupBits = INT_MAX;
    int mincov = INT_MAX, maxcov = INT_MIN;
    #pragma unroll
    for(int k = 0; k < KC; k++){
        const int c = lane + kWave * k;
        if((m.isAnchor >> k) & 1u){
            // supports are non-negative, so their bit patterns order like the floats
            minSupBits = min(minSupBits, __float_as_int(m.sup[k]));
            const int cv = countOf(m, k, 0) + countOf(m, k, 1) + countOf(m, k, 2) + countOf(m, k, 3);
            mincov = min(mincov, cv);
            maxcov = max(maxcov, cv);
            sSup[c - off] = m.sup[k];
        }
    }
    const float minSupport = __int_as_float(waveReduce(minSupBits, OpMin(), INT_MAX));
    mincov = waveReduce(mincov, OpMin(), INT_MAX);
    maxcov = waveReduce(maxcov, OpMax(), INT_MIN);
    __syncthreads();
    float avgSupport = 0.0f;
    if(lane == 0){
        float sum = 0.0f;                       // sequential, as std::accumulate
        for(int q = 0; q < La; q++) sum += sSup[q];
        avgSupport = sum / La;
    }
    avgSupport = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(avgSupport)));
    const bool allGood = fgeq(avgSupport, p.avgThr) && fgeq(minSupport, p.minThr) && fgeq(float(mincov), p.minCovThr);
    bool isHQ = false;
    if(allGood){
        // src/msa.cpp:357-372 takes the smallest percent in [0, pct] whose thresholds
        // 1 - f and 1 - 3f (f = percent / 100) both supports meet and calls the anchor HQ
        // when it is at most pct / 2. The thresholds fall as the percent grows and fgeq
        // stays true as its bound falls, so that holds exactly when the supports meet the
        // thresholds of hqPercent, the largest percent that is at most pct / 2 (the host
        // computes both with the same float expressions).
        isHQ = fgeq(float(mincov), p.minCovThr) && fgeq(avgSupport, p.hqAvgT) && fgeq(minSupport, p.hqMinT);
    }
    // learned seam (include/corrector.hpp:1448-1472): in a non-HQ anchor, every anchor
    // column whose base differs from the consensus is a decision column; a logit >= ln 9
    // takes the consensus, <= -ln 9 keeps the base, anything between falls back to the
    // classic rule below. Decision columns can also be written out with their features.
    uint32_t seamTake = 0, seamKeep = 0;    // 1 bit per slot
    if((p.seamMode != 0 || p.featOut) && !isHQ){
        uint16_t* sX = reinterpret_cast<uint16_t*>(tile.shift);   // the tile region is free now
        float* sH1 = reinterpret_cast<float*>(sX + kWave * kFeat);
        float* sH2 = sH1 + kHid;
        float* sZ = sH2 + kHid;
        #pragma unroll
        for(int k = 0; k < KC; k++){
            const bool anc = (m.isAnchor >> k) & 1u;
            const int ob = int((m.abase >> (2 * k)) & 3u);
            const int cn = int((m.cons >> (2 * k)) & 3u);
            const bool dec = anc && ob != cn;
            const uint64_t mask = __ballot(dec);
            if(mask == 0) continue;
            const int n = __popcll(mask);
            const int e = __popcll(mask & below);
            const int i = lane + kWave * k - off;
            unsigned long long fbase = 0;
            if(p.featOut){
                if(lane == 0) fbase = atomicAdd(p.featCount, (unsigned long long)n);
                fbase = __shfl(fbase, 0);
            }
            if(dec){
                // features (include/classification.hpp:748-797, then quality weight,
                // relative position, smaller neighbour support)
                const int c0 = countOf(m, k, 0), c1 = countOf(m, k, 1), c2 = countOf(m, k, 2), c3 = countOf(m, k, 3);
                const int cov = c0 + c1 + c2 + c3;
                const float wsum = m.w[k][0] + m.w[k][1] + m.w[k][2] + m.w[k][3];
                const int origCov = countOf(m, k, ob);
                const float origW = weightOf(m, k, ob);
                const int consCount = countOf(m, k, cn);
                float f[kFeat];
                f[0] = float(origCov) / float(cov);
                f[1] = origW / wsum;
                f[2] = origW / float(origCov);
                f[3] = float(consCount) / float(cov);
                f[4] = m.sup[k];
                f[5] = weightOf(m, k, cn) / float(consCount);
                f[6] = wsum / float(cov);
                f[7] = float(cov) / p.estCov;
                f[8] = wsum / p.estCov;
                f[9] = avgSupport;
                f[10] = minSupport;
                f[11] = float(maxcov) / p.estCov;
                f[12] = float(mincov) / p.estCov;
                f[13] = p.useQ ? tile.qw[p.qual[p.qdpitch * a + i]] : 1.0f;
                f[14] = float(i) / float(La);
                const float left = i > 0 ? sSup[i - 1] : 1.0f;
                const float right = i + 1 < La ? sSup[i + 1] : 1.0f;
                f[15] = left < right ? left : right;
                #pragma unroll
                for(int q = 0; q < kFeat; q++) sX[e * kFeat + q] = uint16_t(__float_as_uint(bf16r(f[q])) >> 16);
                if(p.featOut && fbase + e < p.featCap){
                    const int classic = allGood ? cn : ((m.sup[k] > 0.90f && origCov <= 2) ? cn : ob);
                    float* fo = p.featOut + (fbase + e) * kFeat;
                    #pragma unroll
                    for(int q = 0; q < kFeat; q++) fo[q] = f[q];
                    uint32_t* mo = p.metaOut + (fbase + e) * 4;
                    mo[0] = uint32_t(a); mo[1] = uint32_t(i);
                    mo[2] = uint32_t(ob | (cn << 2) | (classic << 4) | (int(allGood) << 6)); mo[3] = 0u;
                }
            }
            if(p.seamMode == 0) continue;
            __syncthreads();
            for(int r = 0; r < n; r++){
                const float z = seamLogit(p, sX + r * kFeat, sH1, sH2, lane);
                if(lane == 0) sZ[r] = z;
            }
            __syncthreads();
            if(dec){
                const float z = sZ[e];
                seamTake |= uint32_t(z >= kConfLogit) << k;
                seamKeep |= uint32_t(z <= -kConfLogit) << k;
            }
            __syncthreads();
        }
    }
    uint32_t corrBases = 0;     // 2 bits per slot
    uint32_t editBits = 0;      // 1 bit per slot
    int nEdits = 0;
    #pragma unroll
    for(int k = 0; k < KC; k++){
        const bool anc = (m.isAnchor >> k) & 1u;
        int cb = 0;
        bool e = false;
        if(anc){
            const int ob = int((m.abase >> (2 * k)) & 3u);
            const int cn = int((m.cons >> (2 * k)) & 3u);
            const int oc = countOf(m, k, ob);
            if(allGood) cb = cn;
            else cb = (m.sup[k] > 0.90f && oc <= 2) ? cn : ob;
            if((seamTake >> k) & 1u) cb = cn;
            else if((seamKeep >> k) & 1u) cb = ob;
            e = cb != ob;
        }
        corrBases |= uint32_t(cb) << (2 * k);
        editBits |= uint32_t(e) << k;
        nEdits += __popcll(__ballot(e));
    }
    if(info && lane == 0){
        info[4] = isHQ; info[5] = 1; info[6] = mincov; info[7] = maxcov;
        info[9] = __float_as_int(avgSupport); info[10] = __float_as_int(minSupport);
    }
    // corrected (it has candidates) as an HQ anchor (include/corrector.hpp:284-290)
    if(!AO && p.corrCands){
        if(lane == 0) p.hqAnchor[a] = isHQ;
        if(isHQ) emitCandidateRecords<KC>(p, lst, m.cons, m.ncol, m.ab, beg, nk, La, tile, lane);
    }

    // 5) record (include/corrector.hpp:1639-1761, include/correctedsequence.hpp:299-387);
    // HQ records without edits are not stored (src/correct_cpu.cpp:263-273)
    const bool ambiguous = p.ambig[a] != 0;
    const bool useEdits = !ambiguous && nEdits <= La / 7;
    if(isHQ && useEdits && nEdits == 0) return;
    const uint32_t payload = useEdits ? uint32_t(4 + 5 * nEdits) : uint32_t(4 + La);
    uint8_t* rec = p.slots + size_t(i) * p.slotBytes;
    if(lane == 0){
        const uint32_t rid = uint32_t(a);
        const uint32_t flags = (uint32_t(isHQ) << 31) | (uint32_t(useEdits) << 30) | payload;
        memcpy(rec, &rid, 4);
        memcpy(rec + 4, &flags, 4);
        const int first = useEdits ? nEdits : La;
        memcpy(rec + 8, &first, 4);
        p.recBytes[i] = 8 + payload;
        if(info) info[11] = int(8 + payload);
    }
    int running = 0;
    #pragma unroll
    for(int k = 0; k < KC; k++){
        const bool anc = (m.isAnchor >> k) & 1u;
        const int q = lane + kWave * k - off;
        const int cb = int((corrBases >> (2 * k)) & 3u);
        if(useEdits){
            const bool e = (editBits >> k) & 1u;
            const uint64_t mask = __ballot(e);
            if(e){
                const int r = running + __popcll(mask & below);
                memcpy(rec + 12 + 4 * r, &q, 4);
                rec[12 + 4 * nEdits + r] = uint8_t("ACGT"[cb]);
            }
            running += __popcll(mask);
        }else if(anc){
            rec[12 + q] = uint8_t("ACGT"[cb]);
        }
    }
}

template<int KC, bool AO, int QM>
__global__ void __launch_bounds__(kWave)
__attribute__((amdgpu_waves_per_eu(AO && KC <= 3 ? 7 : (KC <= 8 ? 4 : 2))))
k_msa_wave(MsaParams p){
    extern __shared__ uint8_t smem[];
    __shared__ float sQw[256];
    const int lane = threadIdx.x;
    MsaTile tile;
    float* sSup;
    carveWaveTile(p, smem, sQw, tile, sSup);
    if(p.useQ) for(int q = lane; q < 256; q += kWave) sQw[q] = p.qweight[q];
    __syncthreads();
    const int i = batchAnchor(p.order, p.B);
    const int beg = int(p.offsets[i]);
    const int n0 = int(p.offsets[i + 1]) - beg;
    msaAnchorWave<KC, AO, QM>(p, listsOf(p), p.b0 + i, i, beg, n0, tile, sSup, lane);
}

// ---------------------------------------------------------------- fused persistent kernel

constexpr int kFusedChunk = 4;      // work units taken per counter increment

struct FusedParams {
    MsaParams m;             // storage, thresholds and record slots (per-wave lists set in the kernel)
    CandParams c;            // tables
    AlignParams al;          // alignment thresholds
    uint32_t* counter;       // next work unit
    unsigned long long* aligned;   // candidates aligned (sum of per-anchor candidate counts)
    uint64_t first;          // first anchor
    uint32_t units;          // anchors (SE) or mate pairs (PE)
    uint8_t* scratch;        // wave-private scratch, scratchStride bytes per wave
    size_t scratchStride;
    int cap;                 // maps * threshold: most candidates an anchor can have
};

// Per-wave scratch layout (cap entries each).
struct WaveScratch {
    uint32_t* idsA; uint32_t* idsB; AlignBest* best; uint8_t* pairedA; uint8_t* pairedB;
    uint32_t* kId; int* kShift; int* kLen; int* kNops; int* kOvl; float* kW; uint8_t* kOrient; uint8_t* kPaired;
};

__host__ __device__ inline size_t waveScratchBytes(int cap){
    return size_t(cap) * (4 + 4 + 16 + 1 + 1 + 4 + 4 * 4 + 4 + 1 + 1) + 256;
}

__device__ inline WaveScratch carveScratch(uint8_t* base, int cap){
    WaveScratch w;
    uint8_t* q = base;
    w.best = reinterpret_cast<AlignBest*>(q); q += size_t(cap) * 16;
    w.idsA = reinterpret_cast<uint32_t*>(q); q += size_t(cap) * 4;
    w.idsB = reinterpret_cast<uint32_t*>(q); q += size_t(cap) * 4;
    w.kId = reinterpret_cast<uint32_t*>(q); q += size_t(cap) * 4;
    w.kShift = reinterpret_cast<int*>(q); q += size_t(cap) * 4;
    w.kLen = reinterpret_cast<int*>(q); q += size_t(cap) * 4;
    w.kNops = reinterpret_cast<int*>(q); q += size_t(cap) * 4;
    w.kOvl = reinterpret_cast<int*>(q); q += size_t(cap) * 4;
    w.kW = reinterpret_cast<float*>(q); q += size_t(cap) * 4;
    w.pairedA = q; q += cap;
    w.pairedB = q; q += cap;
    w.kOrient = q; q += cap;
    w.kPaired = q;
    return w;
}

// Candidate ids of anchor a by one wave (include/corrector.hpp:571-650): the same
// lookup / LDS hash-set dedup / sort as k_candidates. Returns the count; ids in out[].
__device__ __forceinline__ int candidatesWave(const CandParams& p, uint64_t a, uint32_t* hset, uint32_t* sOff, uint32_t* sPre,
                              int* sUnique, uint32_t* out, int lane, unsigned long long& gathered){
    const bool valid = p.len[a] >= p.k && !(p.excludeAmbiguous && p.ambig[a]);
    uint32_t cnt = 0, off = 0;
    if(valid && lane < p.M){
        const uint2 b = p.bucketOf[a * uint64_t(p.M) + uint64_t(lane)];
        off = b.x;
        cnt = b.y > uint32_t(p.thr) ? 0u : b.y;
    }
    const int incl = waveInclusiveSum(int(cnt));
    if(lane < kMaxMaps){ sOff[lane] = off; sPre[lane + 1] = uint32_t(incl); }
    if(lane == 0){ sPre[0] = 0; *sUnique = 0; }
    const int T = __builtin_amdgcn_readlane(incl, 63);
    gathered += uint64_t(T);
    uint8_t* mapOf = reinterpret_cast<uint8_t*>(hset + p.hashSlots);
    for(int q = incl - int(cnt); q < incl; q++) mapOf[q] = uint8_t(lane);
    int H = 64;
    while(H <= T) H <<= 1;                   // more slots than values; H <= hashSlots
    __syncthreads();
    for(int q = lane; q < H; q += kWave) hset[q] = 0xFFFFFFFFu;
    __syncthreads();
    gatherInsert(p, a, T, mapOf, sOff, sPre, hset, H, out, sUnique, lane, kWave);
    __syncthreads();
    const int U = *sUnique;
    if(U <= 4 * kWave){
        // rank sort of the distinct ids, up to four per lane
        const int U4 = (U + 3) & ~3;
        for(int q = lane; q < U4; q += kWave) hset[q] = q < U ? out[q] : 0xFFFFFFFFu;
        __syncthreads();
        uint32_t v[4];
        int rank[4] = {0, 0, 0, 0};
        #pragma unroll
        for(int e = 0; e < 4; e++) v[e] = lane + kWave * e < U ? hset[lane + kWave * e] : 0xFFFFFFFFu;
        const uint4* h4 = reinterpret_cast<const uint4*>(hset);
        for(int q = 0; q < U4 / 4; q++){
            const uint4 w = h4[q];
            #pragma unroll
            for(int e = 0; e < 4; e++) rank[e] += int(w.x < v[e]) + int(w.y < v[e]) + int(w.z < v[e]) + int(w.w < v[e]);
        }
        #pragma unroll
        for(int e = 0; e < 4; e++) if(lane + kWave * e < U) out[rank[e]] = v[e];
    }else{
        int P = 1;
        while(P < U) P <<= 1;
        for(int q = lane; q < P; q += kWave) hset[q] = q < U ? out[q] : 0xFFFFFFFFu;
        __syncthreads();
        for(int kk = 2; kk <= P; kk <<= 1){
            for(int j = kk >> 1; j > 0; j >>= 1){
                for(int idx = lane; idx < P; idx += kWave){
                    const int ixj = idx ^ j;
                    if(ixj > idx){
                        const uint32_t x = hset[idx], y = hset[ixj];
                        const bool up = (idx & kk) == 0;
                        if((x > y) == up){ hset[idx] = y; hset[ixj] = x; }
                    }
                }
                __syncthreads();
            }
        }
        for(int q = lane; q < U; q += kWave) out[q] = hset[q];
    }
    __syncthreads();
    return U;
}

__device__ __forceinline__ bool sortedContains(const uint32_t* v, int n, uint32_t x){
    int lo = 0, hi = n;
    while(lo < hi){
        const int mid = (lo + hi) >> 1;
        if(v[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo < n && v[lo] == x;
}

// Alignments of all candidates of one anchor (both orientations + orientation choice),
// candidates in order; the next candidate's planes are loaded while the current one is
// aligned (double-buffered LDS).
__device__ __forceinline__ void alignAllWave(const AlignParams& p, uint64_t a, const uint32_t* ids, int U, AlignBest* best,
                             uint2* planeA, uint2* planeC, int lane){
    const int W = p.W;
    const int stride = W + 1;
    const uint32_t* ah = p.hilo + a * uint64_t(4 * W);
    for(int q = lane; q < 2 * stride; q += kWave){
        const int pl = q / stride, w = q - pl * stride;
        const uint32_t x0 = w < W ? ah[pl * W + w] : 0u;
        const uint32_t x1 = w + 1 < W ? ah[pl * W + w + 1] : 0u;
        planeA[q] = make_uint2(x1, x0);
    }
    const int La = p.len[a];
    // prefetch registers: up to 3 pair entries per lane (4 planes of at most 33 entries)
    uint2 pre[3];
    int nextLen = 0;
    auto fetch = [&](int c){
        const uint32_t id = ids[c];
        const uint32_t* ch = p.hilo + uint64_t(id) * uint64_t(4 * W);
        #pragma unroll
        for(int r = 0; r < 3; r++){
            const int q = lane + kWave * r;
            uint2 v = make_uint2(0u, 0u);
            if(q < 4 * stride){
                const int pl = q / stride, w = q - pl * stride;
                const uint32_t x0 = w < W ? ch[pl * W + w] : 0u;
                const uint32_t x1 = w + 1 < W ? ch[pl * W + w + 1] : 0u;
                v = make_uint2(x1, x0);
            }
            pre[r] = v;
        }
        nextLen = p.len[id];
    };
    if(U > 0) fetch(0);
    for(int c = 0; c < U; c++){
        uint2* buf = planeC + (c & 1) * 4 * stride;
        #pragma unroll
        for(int r = 0; r < 3; r++){
            const int q = lane + kWave * r;
            if(q < 4 * stride) buf[q] = pre[r];
        }
        const int Lc = nextLen;
        if(c + 1 < U) fetch(c + 1);
        __syncthreads();
        ShdResult f, r;
        alignBoth(planeA, planeA + stride, buf, buf + stride, buf + 2 * stride, buf + 3 * stride, La, Lc, p, lane, f, r);
        const int flag = chooseOrientation(f, r, La, p);
        if(lane == 0){
            AlignBest b;
            const ShdResult& sres = flag == 2 ? r : f;
            b.shift = sres.shift; b.overlap = sres.overlap; b.nOps = sres.nOps; b.orientation = flag;
            best[c] = b;
        }
    }
    __syncthreads();
}

// Persistent fused pass: each wave takes work units (an anchor, or a mate pair) from a
// counter and runs candidates -> pair flags -> alignments -> filters/MSA/refinement/
// decision -> record for them, keeping per-anchor lists in its private scratch.
template<int KC, bool PAIRED>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(KC <= 3 ? 3 : 2)))
k_fused(FusedParams fp){
    extern __shared__ uint8_t smem[];
    __shared__ float sQw[256];
    __shared__ uint32_t sOff[kMaxMaps];
    __shared__ uint32_t sPre[kMaxMaps + 1];
    __shared__ int sUnique;
    __shared__ uint32_t sUnit;
    const int lane = threadIdx.x;
    const WaveScratch ws = carveScratch(fp.scratch + size_t(blockIdx.x) * fp.scratchStride, fp.cap);
    MsaTile tile;
    float* sSup;
    carveWaveTile(fp.m, smem, sQw, tile, sSup);
    uint32_t* hset = reinterpret_cast<uint32_t*>(smem);
    const int stride = fp.al.W + 1;
    uint2* planeA = reinterpret_cast<uint2*>(smem);
    uint2* planeC = planeA + 2 * stride;
    if(fp.m.useQ) for(int q = lane; q < 256; q += kWave) sQw[q] = fp.m.qweight[q];
    __syncthreads();
    AnchorLists lst{ws.idsA, ws.best, nullptr, ws.kId, ws.kShift, ws.kLen, ws.kNops, ws.kOvl, ws.kW, ws.kOrient,
                    ws.kPaired};
    // work is taken kFusedChunk units at a time; the statistics are summed per wave and
    // added once at the end (one global atomic per unit serializes the whole device)
    unsigned long long gathered = 0, aligned = 0;
    uint32_t unit = 0, chunkEnd = 0;
    for(;;){
        if(unit == chunkEnd){
            if(lane == 0) sUnit = atomicAdd(fp.counter, uint32_t(kFusedChunk));
            __syncthreads();
            unit = sUnit;
            chunkEnd = unit + uint32_t(kFusedChunk);
            __syncthreads();
        }
        if(unit >= fp.units) break;
        if(!PAIRED){
            const uint64_t a = fp.first + unit;
            const int U = candidatesWave(fp.c, a, hset, sOff, sPre, &sUnique, ws.idsA, lane, gathered);
            aligned += uint64_t(U);
            alignAllWave(fp.al, a, ws.idsA, U, ws.best, planeA, planeC, lane);
            msaAnchorWave<KC, true, 2>(fp.m, lst, a, int(unit), 0, U, tile, sSup, lane);
        }else{
            const uint64_t a0 = fp.first + 2 * uint64_t(unit);
            const int U0 = candidatesWave(fp.c, a0, hset, sOff, sPre, &sUnique, ws.idsA, lane, gathered);
            const int U1 = candidatesWave(fp.c, a0 + 1, hset, sOff, sPre, &sUnique, ws.idsB, lane, gathered);
            aligned += uint64_t(U0 + U1);
            // pair flags: the reference's merge (include/util.hpp:1016-1200) flags x in one
            // mate's list exactly when x^1 (the other read of x's pair) is in the other list
            for(int j = lane; j < U0; j += kWave) ws.pairedA[j] = sortedContains(ws.idsB, U1, ws.idsA[j] ^ 1u);
            for(int j = lane; j < U1; j += kWave) ws.pairedB[j] = sortedContains(ws.idsA, U0, ws.idsB[j] ^ 1u);
            __syncthreads();
            for(int mate = 0; mate < 2; mate++){
                const uint64_t a = a0 + mate;
                lst.ids = mate == 0 ? ws.idsA : ws.idsB;
                lst.paired = mate == 0 ? ws.pairedA : ws.pairedB;
                const int U = mate == 0 ? U0 : U1;
                alignAllWave(fp.al, a, lst.ids, U, ws.best, planeA, planeC, lane);
                msaAnchorWave<KC, true, 2>(fp.m, lst, a, int(2 * unit + mate), 0, U, tile, sSup, lane);
                __syncthreads();
            }
        }
        __syncthreads();
        unit++;
    }
    if(lane == 0){
        if(gathered) atomicAdd(fp.c.gathered, gathered);
        if(aligned) atomicAdd(fp.aligned, aligned);
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ v, uint64_t n, uint32_t x){
    FOR_GRID(i, n) v[i] = x;
}

__global__ void k_zero_u32(uint32_t* v, int n){
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n) v[i] = 0u;
}

// Candidate capacity check of a batch whose candidate arrays were sized before its
// candidates were counted (cap < 2^32): from the batch's pair count in 64 bits at
// status[6..7], status[0] = 1 when the batch has more than `cap`, status[1] = how many
// it has (saturated at 2^32 - 1).
__global__ void k_cap_status(uint64_t cap, uint32_t* __restrict__ status){
    if(blockIdx.x == 0 && threadIdx.x == 0){
        const uint64_t t = *reinterpret_cast<const unsigned long long*>(status + 6);
        status[0] = t > cap ? 1u : 0u;
        status[1] = uint32_t(t < uint64_t(UINT32_MAX) ? t : uint64_t(UINT32_MAX));
    }
}

// After k_cap_status: a batch over capacity gets all-zero offsets, so that the rest of its
// kernels see no candidates and write nothing past the arrays (the batch is redone).
__global__ void k_cap_clear(uint32_t* __restrict__ offsets, int B, const uint32_t* __restrict__ status){
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i <= B && status[0]) offsets[i] = 0u;
}

__global__ void __launch_bounds__(kBlock)
k_gather(const uint8_t* slots, int slotBytes, const uint32_t* recBytes, const uint64_t* outOff, uint8_t* out, int B){
    const int i = blockIdx.x;
    if(i >= B) return;
    const uint32_t nb = recBytes[i];
    const uint8_t* src = slots + size_t(i) * slotBytes;
    uint8_t* dst = out + outOff[i];
    for(uint32_t q = threadIdx.x; q < nb; q += kBlock) dst[q] = src[q];
}

__global__ void k_widen(const uint32_t* in, uint64_t* out, int n){
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n) out[i] = in[i];
}

// k_widen over kWidenPerBlock sizes per 256-thread workgroup, and the number of nonzero
// sizes (records) added to *count: wave ballots summed in LDS, one global atomic per
// workgroup.
constexpr int kWidenPerBlock = 1024;
__global__ void __launch_bounds__(256) k_widen_count(const uint32_t* __restrict__ in, uint64_t* __restrict__ out, int n,
                                                     unsigned long long* __restrict__ count){
    __shared__ unsigned blockCount;
    if(threadIdx.x == 0) blockCount = 0;
    __syncthreads();
    unsigned c = 0;
    for(int k = 0; k < kWidenPerBlock / 256; k++){
        const int i = blockIdx.x * kWidenPerBlock + k * 256 + int(threadIdx.x);
        const uint32_t v = i < n ? in[i] : 0u;
        if(i < n) out[i] = v;
        c += unsigned(__popcll(__ballot(v != 0u)));
    }
    if((threadIdx.x & (kWave - 1)) == 0 && c) atomicAdd(&blockCount, c);
    __syncthreads();
    if(threadIdx.x == 0 && blockCount) atomicAdd(count, (unsigned long long)blockCount);
}

// Candidate records of reads corrected as HQ anchors earlier in the same call are not
// kept (include/corrector.hpp:1676-1680): with one thread the reference processes reads
// in ascending order, so "earlier" is a read id below the anchor's and not below the
// call's first anchor.
__global__ void k_cand_filter(uint32_t* __restrict__ candBytes, const uint32_t* __restrict__ kId,
                              const uint32_t* __restrict__ anchorOf, uint64_t b0, uint64_t first,
                              const uint8_t* __restrict__ hqAnchor, uint32_t total){
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if(q >= total || candBytes[q] == 0) return;
    const uint64_t c = kId[q];
    const uint64_t a = b0 + anchorOf[q];
    if(c < a && c >= first && hqAnchor[c]) candBytes[q] = 0;
}

// Position of every anchor record and candidate record of a batch in the order a
// one-thread run inserts them (src/correct_cpu.cpp:166-281): per 16-read batch of read
// ids, the anchors' records, then their candidates' records in anchor order. Anchor i of
// the 16-read group starting at batch index is (ending at ie) goes to offsets[is] + i;
// working-list entry q of anchor i to ie + q. Sizes, and the slot each entry comes from.
__global__ void k_insertion_order(const uint32_t* __restrict__ recBytes, const uint32_t* __restrict__ candBytes,
                                  const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ anchorOf,
                                  uint64_t b0, int B, uint32_t total, const uint8_t* slots, int slotBytes,
                                  const uint8_t* candSlots, int candSlotBytes, uint64_t* __restrict__ size,
                                  uint64_t* __restrict__ src){
    FOR_GRID(e, uint64_t(B) + total){
    if(e < uint64_t(B)){
        const int i = int(e);
        const uint64_t a = b0 + uint64_t(i);
        const int64_t gs = int64_t(a & ~uint64_t(15)) - int64_t(b0);
        const int is = gs < 0 ? 0 : int(gs);
        const uint64_t pos = uint64_t(offsets[is]) + uint64_t(i);
        size[pos] = recBytes[i];
        src[pos] = reinterpret_cast<uint64_t>(slots + size_t(i) * slotBytes);
    }else if(e < uint64_t(B) + total){
        const uint32_t q = uint32_t(e - uint64_t(B));
        const uint64_t a = b0 + anchorOf[q];
        const int64_t ge = int64_t((a | 15) + 1) - int64_t(b0);
        const int ie = ge > B ? B : int(ge);
        const uint64_t pos = uint64_t(ie) + q;
        size[pos] = candBytes[q];
        src[pos] = reinterpret_cast<uint64_t>(candSlots + size_t(q) * candSlotBytes);
    }
    }
}

// One wavefront per entry: copies its record to its offset in the batch's stream.
__global__ void __launch_bounds__(256) k_gather_entries(const uint64_t* __restrict__ size, const uint64_t* __restrict__ src,
                                                        const uint64_t* __restrict__ off, uint8_t* __restrict__ out,
                                                        uint64_t entries){
    const int lane = threadIdx.x & 63;
    for(uint64_t e = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6; e < entries;
        e += (uint64_t(gridDim.x) * blockDim.x) >> 6){
        const uint64_t nb = size[e];
        const uint8_t* s = reinterpret_cast<const uint8_t*>(src[e]);
        uint8_t* d = out + off[e];
        for(uint64_t q = lane; q < nb; q += 64) d[q] = s[q];
    }
}

// ---------------------------------------------------------------- host side

struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t b){
        if(b <= bytes) return hipSuccess;
        if(p){ hipError_t e = hipFree(p); if(e != hipSuccess) return e; p = nullptr; bytes = 0; }
        const size_t nb = std::max<size_t>(b, 256);
        hipError_t e = hipMalloc(&p, nb);
        if(e == hipSuccess) bytes = nb;
        return e;
    }
    void release(){ if(p) (void)hipFree(p); p = nullptr; bytes = 0; }
    template<class T> T* as() const { return static_cast<T*>(p); }
};

// Records of a context on the host: pinned memory, so that device-to-host copies run at
// DMA speed; grown by doubling; bytes past size() are not initialised.
struct HostRecords {
    uint8_t* p = nullptr;
    size_t n = 0, cap = 0;
    size_t size() const { return n; }
    bool empty() const { return n == 0; }
    uint8_t* data() { return p; }
    const uint8_t* data() const { return p; }
    void clear(){ n = 0; }
    // size `bytes`, keeping the first min(size(), bytes) bytes
    hipError_t resize(size_t bytes){
        if(bytes > cap){
            const size_t nc = std::max({bytes, cap * 2, size_t(1) << 20});
            void* q = nullptr;
            const hipError_t e = hipHostMalloc(&q, nc, hipHostMallocDefault);
            if(e != hipSuccess) return e;
            if(n) std::memcpy(q, p, n);
            if(p) (void)hipHostFree(p);
            p = static_cast<uint8_t*>(q);
            cap = nc;
        }
        n = bytes;
        return hipSuccess;
    }
    void release(){ if(p) (void)hipHostFree(p); p = nullptr; n = cap = 0; }
};

}  // namespace

// Uploads of at least kUpMin bytes go through kUpWorkers host threads with pinned
// double buffers of kUpChunk bytes each (64 MiB of pinned memory per context).
constexpr int kUpWorkers = 8;
constexpr size_t kUpChunk = size_t(4) << 20;
constexpr size_t kUpMin = size_t(64) << 20;

struct care_ctx {
    care_opts o{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    uint64_t n = 0;
    int maxLen = 0;
    bool varLen = false;    // reads of different lengths
    size_t pitch = 0;
    size_t qpitch = 0;
    size_t qdpitch = 0;     // device quality row pitch: qpitch rounded up to 16 bytes
    int W = 0;
    int k = 0;
    int M = 0;
    int thr = 0;
    bool hasReads = false;
    bool hasTables = false;

    DBuf enc, len, qual, ambig, sig, hilo, rcenc, qw, bucketOf, gath, order, orderFlags, batchOrd;
    DBuf seeds;             // per (read, map): seed of the map's minimum k-mer (seeded alignment)
    std::vector<DBuf> slotBufs;
    std::vector<uint32_t> caps;
    DBuf slotPtrs, capsDev, values;

    // batch working memory
    DBuf raw, uniq, offsets, ids, anchorOf, best, dbg, paired;
    DBuf kId, kShift, kLen, kNops, kOvl, kW, kOrient, kPaired;
    DBuf slots, recBytes, recOff, recOff64, outBuf, info, msaDbg, counters, scanTmp, scratch, candOver;
    // learned seam: classifier (bfloat16 weights then float biases) and feature output
    DBuf seam, featBuf, metaBuf;
    // pinned host words: record bytes (u64) and the error flag (i32) of the last pass
    void* pinned = nullptr;
    // fused engine with use_graphs: the instantiated pass and the parameter bytes it was captured with
    hipGraphExec_t fusedExec = nullptr;
    std::vector<uint8_t> fusedKey;
    // staged engine, pipelined batches: two batches in flight, each with its own record
    // output buffer, pinned status words and stage events; a copy stream for the records;
    // the most candidates per anchor seen (sizes the next pass's candidate arrays)
    DBuf pipeOut[2];
    hipEvent_t pipeEv[2][6] = {};
    hipStream_t copyStream = nullptr;
    // large host-to-device uploads: per worker a stream, two pinned chunk buffers and their
    // events (created on the first large upload)
    hipStream_t upStream[kUpWorkers] = {};
    void* upPinned = nullptr;                   // kUpWorkers * 2 chunk buffers
    hipEvent_t upEv[kUpWorkers][2] = {};
    double candPerAnchor = 0;
    double pipePerAnchor = 0;   // candidate pairs per anchor the arrays were last sized for
    // pipelined batches with use_graphs: one instantiated graph per batch (key: the batch's
    // parameter bytes and every address its work uses)
    std::map<std::vector<uint8_t>, hipGraphExec_t> stagedGraphs;
    uint64_t graphsBuilt = 0;
    bool hasSeam = false;
    bool featActive = false;
    unsigned long long featCap = 0;
    HostRecords records;
    size_t numRecords = 0;
    // candidate correction: record slots and sizes per working-list entry, the insertion
    // order of a batch's records, per read the HQ-anchor flag; first anchor of the call
    DBuf candSlots, candBytes, insSize, insOff, insSrc, hqAnchor;
    uint64_t callFirst = 0;
    // FASTQ text kept on the device by care_hip_set_reads_fastq for the output merge, and
    // per read: header start (after '@'), header length, sequence and quality starts
    DBuf fqText[2], fqHdr, fqHdrLen, fqSeq, fqQual;
    size_t fqRaw[2] = {0, 0};
    bool hasFastq = false;
    double fastqMs = 0, mergeMs = 0;     // device spans of the last FASTQ parse and output merge
    // output merge working memory
    DBuf mRec, mOff, mKeyIn, mKeyOut, mIdxIn, mIdxOut, mRunBeg, mRunEnd, mPlan, mSize, mOutOff, mText, mErr;

    // last batch (for inspection)
    uint64_t lastB0 = 0, lastB = 0, lastTotal = 0;
    int maxCols = 0;
    int dbgStride = 0;

    care_timings t{};
    hipEvent_t ev[8] = {};

    int fail(int code, const char* fmt, ...){
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
};

#define HIPCHECK(expr) do { hipError_t e_ = (expr); if(e_ != hipSuccess){ \
    return ctx->fail(CARE_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); } } while(0)

namespace {

float qualityWeight(int i){
    // include/qualityscoreweights.hpp:19-28: std::pow(10.0f, -(i-33)/10.0f) in float
    const float p = i < 33 ? 1.0f : std::pow(10.0f, -(i - 33) / 10.0f);
    return std::max(0.001f, 1.0f - p);
}

int resultsPerMapThreshold(float coverage){
    const int cov = int(coverage);
    int result = int(cov * 2.5f);
    result = std::min(result, int(std::numeric_limits<uint16_t>::max()));
    return std::max(10, result);
}

int buildTables(care_ctx* ctx){
    const uint64_t n = ctx->n;
    const int M = ctx->M;
    hipStream_t st = ctx->stream;
    DBuf iota, keysOut, uniqKeys, runCounts, runOffsets, numRuns, kept, temp;
    auto cleanup = [&](){ iota.release(); keysOut.release(); uniqKeys.release(); runCounts.release();
                          runOffsets.release(); numRuns.release(); kept.release(); temp.release(); };
    struct Guard { decltype(cleanup)& f; ~Guard(){ f(); } } guard{cleanup};

    HIPCHECK(ctx->values.ensure(std::max<uint64_t>(1, n) * M * sizeof(uint32_t)));
    HIPCHECK(iota.ensure(std::max<uint64_t>(1, n) * sizeof(uint32_t)));
    HIPCHECK(keysOut.ensure(std::max<uint64_t>(1, n) * sizeof(uint64_t)));
    HIPCHECK(uniqKeys.ensure(std::max<uint64_t>(1, n) * sizeof(uint64_t)));
    HIPCHECK(runCounts.ensure(std::max<uint64_t>(1, n) * sizeof(uint32_t)));
    HIPCHECK(runOffsets.ensure(std::max<uint64_t>(1, n) * sizeof(uint32_t)));
    HIPCHECK(numRuns.ensure(sizeof(int)));
    HIPCHECK(kept.ensure(sizeof(uint32_t)));
    if(n) k_iota<<<gridFor(n, 256), dim3(256), 0, st>>>(iota.as<uint32_t>(), n);
    HIPCHECK(hipGetLastError());

    size_t tb1 = 0, tb2 = 0, tb3 = 0;
    const int endBit = 2 * ctx->k;
    HIPCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb1, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                                (const uint32_t*)nullptr, (uint32_t*)nullptr, int(n), 0, endBit, st));
    HIPCHECK(hipcub::DeviceRunLengthEncode::Encode(nullptr, tb2, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                                   (uint32_t*)nullptr, (int*)nullptr, int(n), st));
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb3, (const uint32_t*)nullptr, (uint32_t*)nullptr, int(n), st));
    HIPCHECK(temp.ensure(std::max({tb1, tb2, tb3, size_t(256)})));

    ctx->slotBufs.resize(M);
    ctx->caps.assign(M, 0);
    std::vector<Slot*> ptrs(M, nullptr);
    for(int m = 0; m < M; m++){
        const uint64_t* keysIn = ctx->sig.as<uint64_t>() + uint64_t(m) * n;
        uint32_t* vals = ctx->values.as<uint32_t>() + uint64_t(m) * n;
        size_t b = temp.bytes;
        // stable LSD radix sort: equal keys keep ascending read ids
        HIPCHECK(hipcub::DeviceRadixSort::SortPairs(temp.p, b, keysIn, keysOut.as<uint64_t>(), iota.as<uint32_t>(),
                                                    vals, int(n), 0, endBit, st));
        b = temp.bytes;
        HIPCHECK(hipcub::DeviceRunLengthEncode::Encode(temp.p, b, keysOut.as<uint64_t>(), uniqKeys.as<uint64_t>(),
                                                       runCounts.as<uint32_t>(), numRuns.as<int>(), int(n), st));
        b = temp.bytes;
        HIPCHECK(hipcub::DeviceScan::ExclusiveSum(temp.p, b, runCounts.as<uint32_t>(), runOffsets.as<uint32_t>(),
                                                  int(n), st));
        HIPCHECK(hipMemsetAsync(kept.p, 0, sizeof(uint32_t), st));
        const int grid = int((n + 255) / 256);
        if(grid){
            k_count_kept<<<grid, 256, 0, st>>>(runCounts.as<uint32_t>(), numRuns.as<int>(), ctx->thr, kept.as<uint32_t>());
            HIPCHECK(hipGetLastError());
        }
        uint32_t keptHost = 0;
        HIPCHECK(hipMemcpyAsync(&keptHost, kept.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        uint32_t cap = 64;
        while(cap < 2u * keptHost) cap <<= 1;
        ctx->caps[m] = cap;
        HIPCHECK(ctx->slotBufs[m].ensure(size_t(cap) * sizeof(Slot)));
        HIPCHECK(hipMemsetAsync(ctx->slotBufs[m].p, 0, size_t(cap) * sizeof(Slot), st));
        if(grid){
            k_insert<<<grid, 256, 0, st>>>(uniqKeys.as<uint64_t>(), runCounts.as<uint32_t>(), runOffsets.as<uint32_t>(),
                                           numRuns.as<int>(), ctx->thr, ctx->slotBufs[m].as<Slot>(), cap - 1);
            HIPCHECK(hipGetLastError());
        }
        ptrs[m] = ctx->slotBufs[m].as<Slot>();
    }
    HIPCHECK(ctx->slotPtrs.ensure(sizeof(Slot*) * kMaxMaps));
    HIPCHECK(ctx->capsDev.ensure(sizeof(uint32_t) * kMaxMaps));
    std::vector<Slot*> ptrs48(kMaxMaps, nullptr);
    std::vector<uint32_t> caps48(kMaxMaps, 1);
    for(int m = 0; m < M; m++){ ptrs48[m] = ptrs[m]; caps48[m] = ctx->caps[m]; }
    HIPCHECK(hipMemcpyAsync(ctx->slotPtrs.p, ptrs48.data(), sizeof(Slot*) * kMaxMaps, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(ctx->capsDev.p, caps48.data(), sizeof(uint32_t) * kMaxMaps, hipMemcpyHostToDevice, st));
    HIPCHECK(hipStreamSynchronize(st));
    return CARE_OK;
}

// dynamic LDS of k_msa: column arrays, corrected anchor, candidate tile
size_t msaLdsBytes(int maxCols, int maxLen, int W, int tile, bool useQ){
    size_t b = size_t(maxCols) * (16 + 16 + 4 + 4 + 4 + 1);
    b = (b + 15) / 16 * 16 + (size_t(maxLen) + 15) / 16 * 16;
    b += size_t(tile) * (20 + 8 * size_t(W) + (useQ ? size_t(qStride(maxLen)) : 0));
    return b + 16;
}

// Register columns per lane of k_msa_wave for an MSA of at most maxCols columns
// (0: too wide, use the workgroup kernel k_msa).
int waveMsaColumnsPerLane(int maxCols){
    const int need = (maxCols + kWave - 1) / kWave;
    if(need <= 4) return 4;
    if(need <= 8) return 8;
    if(need <= 12) return 12;
    if(need <= 16) return 16;
    return 0;
}

// Register columns per lane of k_msa_wave when only the columns under the anchor are
// built (reads of at most maxLen bases; 0: too long).
int waveMsaAnchorSlots(int maxLen){
    const int need = (maxLen + kWave - 1) / kWave;
    if(need <= 2) return 2;
    if(need <= 3) return 3;
    if(need <= 4) return 4;
    if(need <= 8) return 8;
    if(need <= 12) return 12;
    if(need <= 16) return 16;
    return 0;
}

// Candidates per LDS tile of the one-wavefront MSA. With qualities (220 bytes per
// candidate at 150 bp) 20 keeps the workgroup near 6 KB, so about 26 one-wavefront
// workgroups fit a compute unit's LDS (the anchor-column kernels with at most 3 register
// columns per lane are built for 7 waves per SIMD); without qualities 64 fit in less.
int defaultWaveMsaTile(bool useQ){ return useQ ? 20 : 64; }

// dynamic LDS of k_msa_wave: tile region + anchor-column supports
size_t waveMsaLdsBytes(int tile, int W, int maxLen, bool useQ, bool seam, int maxCols){
    return waveTileRegion(tile, W, maxLen, useQ, seam, maxCols) + size_t(maxLen) * 4 + 16;
}

// Candidates per LDS tile: as many as fit next to the column arrays in ~40 KB, 16..64.
int msaTile(int maxCols, int maxLen, int W, bool useQ){
    const size_t base = msaLdsBytes(maxCols, maxLen, W, 0, useQ);
    const size_t per = 20 + 8 * size_t(W) + (useQ ? size_t(qStride(maxLen)) : 0);
    const size_t budget = 40 * 1024;
    size_t t = base < budget ? (budget - base) / per : 16;
    // a multiple of 4 keeps the quality rows 16-byte aligned
    return int(std::min<size_t>(64, std::max<size_t>(16, t))) & ~3;
}

// Storage pointers and thresholds shared by the staged and the fused paths. The float
// thresholds are computed here, on the host, with the reference's expressions.
void fillMsaCommon(const care_ctx* ctx, MsaParams& mp){
    mp.enc = ctx->enc.as<uint32_t>(); mp.pitch = ctx->pitch; mp.len = ctx->len.as<int>();
    mp.qual = ctx->qual.as<uint8_t>(); mp.qpitch = ctx->qpitch; mp.ambig = ctx->ambig.as<uint8_t>();
    mp.qdpitch = ctx->qdpitch;
    mp.varLen = ctx->varLen ? 1 : 0;
    mp.qweight = ctx->qw.as<float>();
    mp.isPaired = ctx->o.paired; mp.useQ = ctx->o.use_quality_scores;
    mp.maxErrorRate = ctx->o.max_error_rate;
    const float base = ctx->o.estimated_errorrate * 1.0f;
    mp.ratio2 = 2 * base; mp.ratio3 = 3 * base; mp.ratio4 = 4 * base;
    mp.goodCount = ctx->o.estimated_coverage * ctx->o.m_coverage;
    mp.pairedThreshold = ctx->o.paired_filter_threshold;
    const float e = ctx->o.estimated_errorrate;
    mp.avgThr = 1.0f - 1.0f * e;
    mp.minThr = 1.0f - 3.0f * e;
    mp.minCovThr = ctx->o.m_coverage / 6.0f * ctx->o.estimated_coverage;
    mp.pct = int(std::ceil(e * 100.0f));
    mp.pctHalf = mp.pct * 0.5f;
    {
        // hqPercent: the largest percent in [0, pct] with fleq(percent, pct / 2) (feq within 1e-5)
        auto fleqHost = [](float l, float r){
            const float d = l - r < 0 ? -(l - r) : l - r;
            return (l < r) || d < 1e-5f;
        };
        int hqPercent = 0;
        for(int q = 0; q <= mp.pct; q++) if(fleqHost(float(q), mp.pctHalf)) hqPercent = q;
        const float factor = hqPercent / 100.0f;
        mp.hqAvgT = 1.0f - 1.0f * factor;
        mp.hqMinT = 1.0f - 3.0f * factor;
    }
    mp.sigThr = int(int(ctx->o.estimated_coverage) * 0.3f);
    mp.maxCols = ctx->maxCols; mp.maxLen = ctx->maxLen;
    mp.hilo = ctx->hilo.as<uint32_t>(); mp.W = ctx->W; mp.rcenc = ctx->rcenc.as<uint32_t>();
    mp.info = nullptr; mp.msaDbg = nullptr; mp.dbgStride = ctx->dbgStride;
    mp.seamMode = ctx->o.correction_type == 1 ? 1 : 0;
    mp.seamW = ctx->seam.as<uint16_t>();
    mp.seamB = reinterpret_cast<const float*>(ctx->seam.as<uint8_t>() + kSeamWeightBytes);
    mp.estCov = ctx->o.estimated_coverage;
    mp.featOut = ctx->featActive ? ctx->featBuf.as<float>() : nullptr;
    mp.metaOut = ctx->featActive ? ctx->metaBuf.as<uint32_t>() : nullptr;
    mp.featCount = ctx->counters.as<unsigned long long>() + 4;
    mp.featCap = ctx->featCap;
    mp.corrCands = 0; mp.newCols = ctx->o.new_columns_to_correct;
    mp.candSlots = nullptr; mp.candSlotBytes = 0; mp.candBytes = nullptr; mp.hqAnchor = nullptr;
    mp.order = nullptr; mp.B = 0;
}

void fillCandCommon(const care_ctx* ctx, CandParams& cp){
    cp.sig = ctx->sig.as<uint64_t>(); cp.n = ctx->n; cp.M = ctx->M; cp.k = ctx->k; cp.thr = ctx->thr;
    cp.bucketOf = ctx->bucketOf.as<uint2>();
    cp.excludeAmbiguous = ctx->o.exclude_ambiguous; cp.len = ctx->len.as<int>(); cp.ambig = ctx->ambig.as<uint8_t>();
    cp.slots = ctx->slotPtrs.as<const Slot*>(); cp.caps = ctx->capsDev.as<uint32_t>(); cp.values = ctx->values.as<uint32_t>();
    cp.capPerAnchor = ctx->M * ctx->thr;
    int P = 64;
    while(P <= cp.capPerAnchor) P <<= 1;      // hash slots > the most values an anchor can gather
    cp.hashSlots = P;
    cp.gath = nullptr;
    cp.order = nullptr;
}

// max_errors = int(float(ov) * rate) (src/cpu_alignment.cpp) as the integer
// mulhi(ov << pre, mul) = (ov * mul) >> (32 - pre): rate = mul * 2^-(32-pre) exactly (mul
// its 24-bit significand), so the integer form is the floor of the exact product; it equals
// the float form unless rounding the product reaches the next integer, which is checked for
// every overlap up to maxLen (as is ov << pre < 2^32).
bool errorBoundAsInteger(float rate, int maxLen, uint32_t& mul, int& pre){
    if(!(rate > 0.0f) || !std::isfinite(rate)) return false;
    int ex = 0;
    const float fr = std::frexp(rate, &ex);          // rate = fr * 2^ex, fr in [0.5, 1)
    mul = uint32_t(std::ldexp(fr, 24));
    pre = 32 - (24 - ex);
    if(pre < 0 || pre > 31 || (uint64_t(uint32_t(maxLen)) << pre) >= (uint64_t(1) << 32)) return false;
    for(int ov = 0; ov <= maxLen; ov++){
        const float prod = float(ov) * rate;
        const uint32_t hi = uint32_t((uint64_t(uint32_t(ov) << pre) * mul) >> 32);
        if(int(prod) != int(hi)) return false;
    }
    return true;
}

void fillAlignCommon(const care_ctx* ctx, AlignParams& ap){
    ap.hilo = ctx->hilo.as<uint32_t>(); ap.W = ctx->W; ap.len = ctx->len.as<int>();
    ap.minOverlap = ctx->o.min_overlap; ap.maxErrorRate = ctx->o.max_error_rate; ap.minOverlapRatio = ctx->o.min_overlap_ratio;
    ap.orientErrorRate = ctx->o.estimated_errorrate * 4.0f;
    ap.maxLen = ctx->maxLen;
    ap.eExact = errorBoundAsInteger(ap.maxErrorRate, ctx->maxLen, ap.eMul, ap.ePre) ? 1 : 0;
    if(!ap.eExact){ ap.eMul = 0; ap.ePre = 0; }
    ap.order = nullptr;
    ap.bucketOf = ctx->bucketOf.as<uint2>(); ap.seeds = ctx->seeds.as<uint16_t>();
    ap.M = ctx->M; ap.k = ctx->k; ap.thr = ctx->thr;
}

int recordSlotBytes(int maxLen){
    return ((8 + 4 + std::max(maxLen, 5 * (maxLen / 7 + 1))) + 15) / 16 * 16;
}

// Record compaction, device part: record sizes widened and scanned into byte offsets,
// then the total and the error flag copied into the context's pinned host words (no host
// synchronization; capturable). Uses the preallocated scan workspace only.
int recordsDevicePart(care_ctx* ctx, int B){
    hipStream_t st = ctx->stream;
    k_widen<<<(B + 1 + 255) / 256, 256, 0, st>>>(ctx->recBytes.as<uint32_t>(), ctx->recOff64.as<uint64_t>(), B + 1);
    HIPCHECK(hipGetLastError());
    size_t tb = ctx->scanTmp.bytes;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(ctx->scanTmp.p, tb, ctx->recOff64.as<uint64_t>(), ctx->recOff.as<uint64_t>(), B + 1, st));
    HIPCHECK(hipMemcpyAsync(ctx->pinned, ctx->recOff.as<uint64_t>() + B, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(reinterpret_cast<uint8_t*>(ctx->pinned) + 8, ctx->counters.as<int>() + 2, 4,
                            hipMemcpyDeviceToHost, st));
    return CARE_OK;
}

// Scan workspace for B + 1 record sizes (allocated before any capture).
int ensureRecordScan(care_ctx* ctx, int B){
    size_t tb = 0;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, ctx->recOff64.as<uint64_t>(), ctx->recOff.as<uint64_t>(), B + 1,
                                              ctx->stream));
    HIPCHECK(ctx->scanTmp.ensure(std::max<size_t>(tb, 256)));
    return CARE_OK;
}

// Record compaction, host part (after recordsDevicePart and a stream synchronization):
// checks the error flag, gathers the records into one buffer and appends it to `out`.
int recordsFinish(care_ctx* ctx, int B, int slotBytes, HostRecords& out, size_t& outRecords){
    hipStream_t st = ctx->stream;
    uint64_t outBytes = 0;
    int errFlag = 0;
    std::memcpy(&outBytes, ctx->pinned, 8);
    std::memcpy(&errFlag, reinterpret_cast<uint8_t*>(ctx->pinned) + 8, 4);
    if(errFlag) return ctx->fail(CARE_EINVAL, "MSA column bound exceeded");
    HIPCHECK(ctx->outBuf.ensure(std::max<uint64_t>(outBytes, 1)));
    k_gather<<<B, kBlock, 0, st>>>(ctx->slots.as<uint8_t>(), slotBytes, ctx->recBytes.as<uint32_t>(),
                                   ctx->recOff.as<uint64_t>(), ctx->outBuf.as<uint8_t>(), B);
    HIPCHECK(hipGetLastError());
    const size_t old = out.size();
    HIPCHECK(out.resize(old + outBytes));
    if(outBytes) HIPCHECK(hipMemcpyAsync(out.data() + old, ctx->outBuf.p, outBytes, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    size_t q = old;
    while(q < out.size()){
        uint32_t flags;
        memcpy(&flags, out.data() + q + 4, 4);
        q += 8 + (flags & ((1u << 29) - 1));
        outRecords++;
    }
    return CARE_OK;
}

// Compacts the per-anchor record slots [0, B) into one buffer and appends it to `out`.
int gatherRecords(care_ctx* ctx, int B, int slotBytes, HostRecords& out, size_t& outRecords){
    int rc = ensureRecordScan(ctx, B);
    if(rc != CARE_OK) return rc;
    rc = recordsDevicePart(ctx, B);
    if(rc != CARE_OK) return rc;
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    return recordsFinish(ctx, B, slotBytes, out, outRecords);
}

// Fused pass over anchors [first, first+n): one persistent kernel, then record compaction.
int runFused(care_ctx* ctx, uint64_t first, uint64_t n, HostRecords& out, size_t& outRecords){
    hipStream_t st = ctx->stream;
    const bool paired = ctx->o.paired != 0;
    const uint64_t units = paired ? n / 2 : n;
    const int kc = waveMsaAnchorSlots(ctx->maxLen);
    if(n == 0) return CARE_OK;
    if(n > uint64_t(INT_MAX)) return ctx->fail(CARE_EINVAL, "too many anchors for one fused pass");
    const int B = int(n);
    const int slotBytes = recordSlotBytes(ctx->maxLen);
    FusedParams fp;
    std::memset(&fp, 0, sizeof(fp));     // padding included: the bytes key the pass's graph
    fillMsaCommon(ctx, fp.m);
    fillCandCommon(ctx, fp.c);
    fillAlignCommon(ctx, fp.al);
    const bool q = ctx->o.use_quality_scores != 0;
    fp.m.tile = defaultWaveMsaTile(q);
    fp.cap = fp.c.capPerAnchor;
    int cus = 0;
    HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const int wpc = ctx->o.waves_per_cu > 0 ? ctx->o.waves_per_cu : 8;
    uint64_t G = uint64_t(std::max(1, cus)) * uint64_t(wpc);
    G = std::max<uint64_t>(1, std::min<uint64_t>(G, units));
    fp.scratchStride = (waveScratchBytes(fp.cap) + 255) / 256 * 256;
    HIPCHECK(ctx->scratch.ensure(G * fp.scratchStride));
    HIPCHECK(ctx->slots.ensure(size_t(B) * slotBytes));
    HIPCHECK(ctx->recBytes.ensure(size_t(B + 1) * 4));
    HIPCHECK(ctx->recOff64.ensure(size_t(B + 1) * 8));
    HIPCHECK(ctx->recOff.ensure(size_t(B + 1) * 8));
    HIPCHECK(ctx->counters.ensure(64));
    fp.m.b0 = first; fp.m.offsets = nullptr; fp.m.ids = nullptr; fp.m.best = nullptr; fp.m.paired = nullptr;
    fp.m.slots = ctx->slots.as<uint8_t>(); fp.m.slotBytes = slotBytes; fp.m.recBytes = ctx->recBytes.as<uint32_t>();
    fp.m.error = ctx->counters.as<int>() + 2;
    fp.c.b0 = first; fp.c.B = B; fp.c.raw = nullptr; fp.c.uniq = nullptr; fp.c.gathered = ctx->counters.as<unsigned long long>();
    // counters: [0,8) gathered bucket values, [8,12) error flag, [16,20) work counter,
    // [24,32) candidates aligned, [32,40) decision columns written out
    fp.counter = ctx->counters.as<uint32_t>() + 4;
    fp.aligned = ctx->counters.as<unsigned long long>() + 3;
    fp.first = first;
    fp.units = uint32_t(units);
    fp.scratch = ctx->scratch.as<uint8_t>();
    const size_t lds = std::max({candLdsBytes(fp.c.hashSlots, fp.c.capPerAnchor), size_t(10) * (ctx->W + 1) * 8,
                                 waveMsaLdsBytes(fp.m.tile, ctx->W, ctx->maxLen, q,
                                                 fp.m.seamMode != 0 || fp.m.featOut != nullptr, fp.m.maxCols)});
    auto kernelFor = [&]() -> const void* {
        switch(kc){
            case 2: return paired ? (const void*)k_fused<2, true> : (const void*)k_fused<2, false>;
            case 3: return paired ? (const void*)k_fused<3, true> : (const void*)k_fused<3, false>;
            case 4: return paired ? (const void*)k_fused<4, true> : (const void*)k_fused<4, false>;
            case 8: return paired ? (const void*)k_fused<8, true> : (const void*)k_fused<8, false>;
            case 12: return paired ? (const void*)k_fused<12, true> : (const void*)k_fused<12, false>;
            default: return paired ? (const void*)k_fused<16, true> : (const void*)k_fused<16, false>;
        }
    };
    if(lds > 65536) HIPCHECK(hipFuncSetAttribute(kernelFor(), hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    {
        const int rc = ensureRecordScan(ctx, B);
        if(rc != CARE_OK) return rc;
    }
    // the pass's device work: zero fills (kernels, so that a captured graph holds no
    // memset node), the persistent kernel, record sizes scanned and the total copied to
    // pinned host memory
    auto enqueue = [&]() -> hipError_t {
        // recBytes[B] must read as zero for the exclusive scan; the kernel writes [0, B)
        k_zero_u32<<<1, 64, 0, st>>>(ctx->recBytes.as<uint32_t>() + B, 1);
        k_zero_u32<<<1, 64, 0, st>>>(ctx->counters.as<uint32_t>() + 4, 1);
        void* args[] = {&fp};
        hipError_t e = hipLaunchKernel(kernelFor(), dim3(uint32_t(G)), dim3(kWave), args, lds, st);
        if(e != hipSuccess) return e;
        return recordsDevicePart(ctx, B) == CARE_OK ? hipSuccess : hipErrorUnknown;
    };
    if(ctx->o.use_graphs){
        // one executable graph per (parameters, grid, buffers); replayed while they hold.
        // The key holds every device or pinned address the captured work uses: those in
        // the kernel parameters, and the record scan's input, output and workspace, which
        // other calls on this context can reallocate.
        const uint64_t extra[] = {G, uint64_t(lds), uint64_t(B),
                                  uint64_t(reinterpret_cast<uintptr_t>(ctx->scanTmp.p)), uint64_t(ctx->scanTmp.bytes),
                                  uint64_t(reinterpret_cast<uintptr_t>(ctx->recOff64.p)), uint64_t(ctx->recOff64.bytes),
                                  uint64_t(reinterpret_cast<uintptr_t>(ctx->recOff.p)), uint64_t(ctx->recOff.bytes),
                                  uint64_t(reinterpret_cast<uintptr_t>(ctx->recBytes.p)), uint64_t(ctx->recBytes.bytes),
                                  uint64_t(reinterpret_cast<uintptr_t>(ctx->counters.p)),
                                  uint64_t(reinterpret_cast<uintptr_t>(ctx->pinned))};
        std::vector<uint8_t> key(sizeof(fp) + sizeof(extra));
        std::memcpy(key.data(), &fp, sizeof(fp));
        std::memcpy(key.data() + sizeof(fp), extra, sizeof(extra));
        if(!ctx->fusedExec || key != ctx->fusedKey){
            HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const hipError_t e = enqueue();
            hipGraph_t g = nullptr;
            const hipError_t ee = hipStreamEndCapture(st, &g);
            if(e != hipSuccess || ee != hipSuccess){
                if(g) (void)hipGraphDestroy(g);
                return ctx->fail(CARE_EHIP, "capturing the fused pass failed: %s",
                                 hipGetErrorString(e != hipSuccess ? e : ee));
            }
            if(ctx->fusedExec){ (void)hipGraphExecDestroy(ctx->fusedExec); ctx->fusedExec = nullptr; }
            const hipError_t ie = hipGraphInstantiate(&ctx->fusedExec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if(ie != hipSuccess){ ctx->fusedExec = nullptr; return ctx->fail(CARE_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(ie)); }
            ctx->fusedKey = key;
            ctx->graphsBuilt += 1;
        }
        HIPCHECK(hipEventRecord(ctx->ev[0], st));
        HIPCHECK(hipGraphLaunch(ctx->fusedExec, st));
        HIPCHECK(hipEventRecord(ctx->ev[1], st));
    }else{
        HIPCHECK(hipEventRecord(ctx->ev[0], st));
        const hipError_t e = enqueue();
        if(e != hipSuccess) return ctx->fail(CARE_EHIP, "fused pass launch failed: %s", hipGetErrorString(e));
        HIPCHECK(hipEventRecord(ctx->ev[1], st));
    }
    HIPCHECK(hipStreamSynchronize(st));
    const int rc = recordsFinish(ctx, B, slotBytes, out, outRecords);
    if(rc != CARE_OK) return rc;
    float ms;
    HIPCHECK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    ctx->t.fused_kernel_ms += ms;
    ctx->t.batches += 1;
    ctx->lastB = 0;
    return CARE_OK;
}

// Slot of one candidate record: header, count, edits (5 bytes each, at most length/7 + 1)
// or the sequence, shift.
int candidateSlotBytes(int maxLen){
    return ((8 + 4 + std::max(maxLen, 5 * (maxLen / 7 + 1)) + 4) + 15) / 16 * 16;
}

// Candidate correction: drops the candidate records the reference would not keep, then
// gathers the batch's anchor and candidate records in the reference's insertion order
// and appends them to `out` (after the MSA kernel of the batch).
int gatherInsertionOrder(care_ctx* ctx, uint64_t b0, int B, uint32_t total, int slotBytes, int candSlotBytes,
                         HostRecords& out, size_t& outRecords){
    hipStream_t st = ctx->stream;
    if(total){
        k_cand_filter<<<(total + 255) / 256, 256, 0, st>>>(ctx->candBytes.as<uint32_t>(), ctx->kId.as<uint32_t>(),
                                                           ctx->anchorOf.as<uint32_t>(), b0, ctx->callFirst,
                                                           ctx->hqAnchor.as<uint8_t>(), total);
        HIPCHECK(hipGetLastError());
    }
    const uint64_t entries = uint64_t(B) + total;
    HIPCHECK(ctx->insSize.ensure((entries + 1) * 8));
    HIPCHECK(ctx->insOff.ensure((entries + 1) * 8));
    HIPCHECK(ctx->insSrc.ensure(entries * 8));
    k_fill_u32<<<1, 64, 0, st>>>(reinterpret_cast<uint32_t*>(ctx->insSize.as<uint64_t>() + entries), 2, 0u);
    k_insertion_order<<<gridFor(entries, 256), 256, 0, st>>>(
        ctx->recBytes.as<uint32_t>(), ctx->candBytes.as<uint32_t>(), ctx->offsets.as<uint32_t>(), ctx->anchorOf.as<uint32_t>(),
        b0, B, total, ctx->slots.as<uint8_t>(), slotBytes, ctx->candSlots.as<uint8_t>(), candSlotBytes,
        ctx->insSize.as<uint64_t>(), ctx->insSrc.as<uint64_t>());
    HIPCHECK(hipGetLastError());
    size_t tb = 0;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, ctx->insSize.as<uint64_t>(), ctx->insOff.as<uint64_t>(),
                                              int(entries + 1), st));
    HIPCHECK(ctx->scanTmp.ensure(std::max<size_t>(tb, 256)));
    tb = ctx->scanTmp.bytes;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(ctx->scanTmp.p, tb, ctx->insSize.as<uint64_t>(), ctx->insOff.as<uint64_t>(),
                                              int(entries + 1), st));
    uint64_t outBytes = 0;
    int errFlag = 0;
    HIPCHECK(hipMemcpyAsync(&outBytes, ctx->insOff.as<uint64_t>() + entries, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&errFlag, ctx->counters.as<int>() + 2, 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if(errFlag) return ctx->fail(CARE_EINVAL, "MSA column bound exceeded in batch at %llu", (unsigned long long)b0);
    HIPCHECK(ctx->outBuf.ensure(std::max<uint64_t>(outBytes, 1)));
    k_gather_entries<<<gridFor(entries * 64, 256), 256, 0, st>>>(
        ctx->insSize.as<uint64_t>(), ctx->insSrc.as<uint64_t>(), ctx->insOff.as<uint64_t>(), ctx->outBuf.as<uint8_t>(),
        entries);
    HIPCHECK(hipGetLastError());
    const size_t old = out.size();
    HIPCHECK(out.resize(old + outBytes));
    if(outBytes) HIPCHECK(hipMemcpyAsync(out.data() + old, ctx->outBuf.p, outBytes, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipEventRecord(ctx->ev[4], st));
    HIPCHECK(hipStreamSynchronize(st));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1])); ctx->t.candidates_ms += ms;
    HIPCHECK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[5])); ctx->t.cand_kernel_ms += ms;
    HIPCHECK(hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2])); ctx->t.align_ms += ms; ctx->t.align_kernel_ms += ms;
    HIPCHECK(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3])); ctx->t.msa_ms += ms; ctx->t.msa_kernel_ms += ms;
    HIPCHECK(hipEventElapsedTime(&ms, ctx->ev[3], ctx->ev[4])); ctx->t.output_ms += ms;
    ctx->t.candidates += total;
    size_t q = old;
    while(q < out.size()){
        uint32_t flags;
        memcpy(&flags, out.data() + q + 4, 4);
        q += 8 + (flags & ((1u << 29) - 1));
        outRecords++;
    }
    return CARE_OK;
}

// Processing order of the anchors of batch [b0, b0+B): the context's order restricted to
// the batch, as batch indices.
__global__ void k_order_flags(const uint32_t* __restrict__ order, uint64_t n, uint64_t b0, uint64_t B,
                              uint8_t* __restrict__ flags){
    FOR_GRID(t, n) flags[t] = order[t] >= b0 && order[t] < b0 + B;
}

__global__ void k_order_rebase(uint32_t* __restrict__ v, uint64_t m, uint32_t b0){
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if(t < m) v[t] -= b0;
}

int batchOrder(care_ctx* ctx, uint64_t b0, int B, const uint32_t** out){
    hipStream_t st = ctx->stream;
    if(b0 == 0 && uint64_t(B) == ctx->n){ *out = ctx->order.as<uint32_t>(); return CARE_OK; }
    const uint64_t n = ctx->n;
    HIPCHECK(ctx->orderFlags.ensure(n + 16));
    HIPCHECK(ctx->batchOrd.ensure(size_t(B) * 4 + 16));
    k_order_flags<<<gridFor(n, 256), 256, 0, st>>>(ctx->order.as<uint32_t>(), n, b0, uint64_t(B),
                                                                 ctx->orderFlags.as<uint8_t>());
    HIPCHECK(hipGetLastError());
    uint32_t* count = reinterpret_cast<uint32_t*>(ctx->batchOrd.as<uint8_t>() + size_t(B) * 4);
    size_t tb = 0;
    HIPCHECK(hipcub::DeviceSelect::Flagged(nullptr, tb, ctx->order.as<uint32_t>(), ctx->orderFlags.as<uint8_t>(),
                                           ctx->batchOrd.as<uint32_t>(), count, int(n), st));
    HIPCHECK(ctx->scanTmp.ensure(std::max<size_t>(tb, 256)));
    tb = ctx->scanTmp.bytes;
    HIPCHECK(hipcub::DeviceSelect::Flagged(ctx->scanTmp.p, tb, ctx->order.as<uint32_t>(), ctx->orderFlags.as<uint8_t>(),
                                           ctx->batchOrd.as<uint32_t>(), count, int(n), st));
    k_order_rebase<<<dim3(uint32_t((B + 255) / 256)), 256, 0, st>>>(ctx->batchOrd.as<uint32_t>(), uint64_t(B), uint32_t(b0));
    HIPCHECK(hipGetLastError());
    *out = ctx->batchOrd.as<uint32_t>();
    return CARE_OK;
}

// The alignment kernel of a staged batch: the seeded kernel (seeded alignment), the
// lane-per-candidate kernel (wm > 0), or the wavefront-per-pair kernel over `total` pairs.
hipError_t launchAlign(const care_ctx* ctx, const AlignParams& ap, int B, uint32_t total, int wm, hipStream_t st){
    if(ctx->o.seeded_alignment){
        switch(alignLaneWords(ctx->W)){
            case 2: k_align_seeded<2><<<B, kWave, 0, st>>>(ap); break;
            case 4: k_align_seeded<4><<<B, kWave, 0, st>>>(ap); break;
            case 5: k_align_seeded<5><<<B, kWave, 0, st>>>(ap); break;
            case 8: k_align_seeded<8><<<B, kWave, 0, st>>>(ap); break;
            case 11: k_align_seeded<11><<<B, kWave, 0, st>>>(ap); break;
            default: return hipErrorInvalidValue;      // rejected before any batch runs
        }
    }else{
        switch(wm){
            case 2: k_align_lanes<2><<<B, kWave, 0, st>>>(ap); break;
            case 4: k_align_lanes<4><<<B, kWave, 0, st>>>(ap); break;
            case 5: k_align_lanes<5><<<B, kWave, 0, st>>>(ap); break;
            case 8: k_align_lanes<8><<<B, kWave, 0, st>>>(ap); break;
            case 11: k_align_lanes<11><<<B, kWave, 0, st>>>(ap); break;
            default: k_align<<<gridFor(uint64_t(total) * kWave, kBlock), kBlock, 0, st>>>(ap); break;
        }
    }
    return hipGetLastError();
}

// The one-wavefront MSA kernel of a staged batch with kc register columns per lane: the
// whole MSA with the quality mode read at run time (candidate correction, dumps), or the
// anchor's columns with the quality mode fixed at compile time.
hipError_t launchWaveMsa(const MsaParams& mp, int B, int kc, bool wholeMsa, bool q, size_t lds, hipStream_t st){
    if(wholeMsa){
        switch(kc){
            case 4: k_msa_wave<4, false, 2><<<B, kWave, lds, st>>>(mp); break;
            case 8: k_msa_wave<8, false, 2><<<B, kWave, lds, st>>>(mp); break;
            case 12: k_msa_wave<12, false, 2><<<B, kWave, lds, st>>>(mp); break;
            default: k_msa_wave<16, false, 2><<<B, kWave, lds, st>>>(mp); break;
        }
    }else if(q){
        switch(kc){
            case 2: k_msa_wave<2, true, 1><<<B, kWave, lds, st>>>(mp); break;
            case 3: k_msa_wave<3, true, 1><<<B, kWave, lds, st>>>(mp); break;
            case 4: k_msa_wave<4, true, 1><<<B, kWave, lds, st>>>(mp); break;
            case 8: k_msa_wave<8, true, 1><<<B, kWave, lds, st>>>(mp); break;
            case 12: k_msa_wave<12, true, 1><<<B, kWave, lds, st>>>(mp); break;
            default: k_msa_wave<16, true, 1><<<B, kWave, lds, st>>>(mp); break;
        }
    }else{
        switch(kc){
            case 2: k_msa_wave<2, true, 0><<<B, kWave, lds, st>>>(mp); break;
            case 3: k_msa_wave<3, true, 0><<<B, kWave, lds, st>>>(mp); break;
            case 4: k_msa_wave<4, true, 0><<<B, kWave, lds, st>>>(mp); break;
            case 8: k_msa_wave<8, true, 0><<<B, kWave, lds, st>>>(mp); break;
            case 12: k_msa_wave<12, true, 0><<<B, kWave, lds, st>>>(mp); break;
            default: k_msa_wave<16, true, 0><<<B, kWave, lds, st>>>(mp); break;
        }
    }
    return hipGetLastError();
}

int runBatch(care_ctx* ctx, uint64_t b0, int B, HostRecords& out, size_t& outRecords){
    hipStream_t st = ctx->stream;
    const int capPerAnchor = ctx->M * ctx->thr;
    int P = 64;
    while(P <= capPerAnchor) P <<= 1;      // hash slots > the most values an anchor can gather
    // counters: [0] gathered bucket values, [1] error flag
    HIPCHECK(ctx->counters.ensure(64));
    HIPCHECK(ctx->raw.ensure(size_t(B) * capPerAnchor * sizeof(uint32_t)));
    HIPCHECK(ctx->uniq.ensure(size_t(B + 1) * sizeof(uint32_t)));
    HIPCHECK(ctx->offsets.ensure(size_t(B + 1) * sizeof(uint32_t)));
    HIPCHECK(hipMemsetAsync(ctx->uniq.p, 0, size_t(B + 1) * sizeof(uint32_t), st));

    HIPCHECK(hipEventRecord(ctx->ev[0], st));
    CandParams cp;
    cp.sig = ctx->sig.as<uint64_t>(); cp.n = ctx->n; cp.M = ctx->M; cp.k = ctx->k; cp.thr = ctx->thr;
    cp.bucketOf = ctx->bucketOf.as<uint2>();
    cp.excludeAmbiguous = ctx->o.exclude_ambiguous; cp.len = ctx->len.as<int>(); cp.ambig = ctx->ambig.as<uint8_t>();
    cp.slots = ctx->slotPtrs.as<const Slot*>(); cp.caps = ctx->capsDev.as<uint32_t>(); cp.values = ctx->values.as<uint32_t>();
    cp.b0 = b0; cp.B = B; cp.capPerAnchor = capPerAnchor; cp.hashSlots = P; cp.raw = ctx->raw.as<uint32_t>(); cp.uniq = ctx->uniq.as<uint32_t>();
    cp.gathered = ctx->counters.as<unsigned long long>();
    HIPCHECK(ctx->gath.ensure(size_t(B) * 4));
    cp.gath = ctx->gath.as<uint32_t>();
    const uint32_t* order = nullptr;
    {
        const int rc = batchOrder(ctx, b0, B, &order);
        if(rc != CARE_OK) return rc;
    }
    cp.order = order;
    // one wavefront per anchor; the few anchors with more distinct candidates than its
    // small hash set takes are redone by workgroups (same ids either way)
    HIPCHECK(ctx->candOver.ensure(size_t(B) * 4 + 16));
    uint32_t* overCount = reinterpret_cast<uint32_t*>(ctx->candOver.as<uint8_t>() + size_t(B) * 4);
    k_zero_u32<<<1, 64, 0, st>>>(overCount, 1);
    const size_t waveLds = candWaveLdsBytes(capPerAnchor);
    if(waveLds > 65536) HIPCHECK(hipFuncSetAttribute((const void*)k_candidates_wave, hipFuncAttributeMaxDynamicSharedMemorySize, int(waveLds)));
    k_candidates_wave<<<B, kWave, waveLds, st>>>(cp, ctx->candOver.as<uint32_t>(), overCount);
    HIPCHECK(hipGetLastError());
    const size_t candLds = candLdsBytes(P, capPerAnchor);
    if(candLds > 65536) HIPCHECK(hipFuncSetAttribute((const void*)k_candidates_over, hipFuncAttributeMaxDynamicSharedMemorySize, int(candLds)));
    k_candidates_over<<<std::min(B, 4096), kBlock, candLds, st>>>(cp, ctx->candOver.as<uint32_t>(), overCount);
    HIPCHECK(hipGetLastError());
    // the batch's candidate pairs in 64 bits: list offsets and indices are 32-bit
    unsigned long long* pairs64 = ctx->counters.as<unsigned long long>() + 5;
    k_zero_u32<<<1, 64, 0, st>>>(reinterpret_cast<uint32_t*>(pairs64), 2);
    k_add_gathered<<<(B + kAddPerBlock - 1) / kAddPerBlock, kBlock, 0, st>>>(ctx->uniq.as<uint32_t>(), B, pairs64);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventRecord(ctx->ev[5], st));
    size_t tb = 0;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, ctx->uniq.as<uint32_t>(), ctx->offsets.as<uint32_t>(), B + 1, st));
    HIPCHECK(ctx->scanTmp.ensure(std::max<size_t>(tb, 256)));
    tb = ctx->scanTmp.bytes;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(ctx->scanTmp.p, tb, ctx->uniq.as<uint32_t>(), ctx->offsets.as<uint32_t>(), B + 1, st));
    uint32_t total = 0;
    unsigned long long total64 = 0;
    HIPCHECK(hipMemcpyAsync(&total, ctx->offsets.as<uint32_t>() + B, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&total64, pairs64, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if(total64 > uint64_t(UINT32_MAX)){
        // more pairs than 32-bit list offsets address: the two halves one after the other
        // (records keep their order; halves end on 16-read groups for candidate correction
        // and on mate pairs)
        int B1 = B / 2;
        if(ctx->o.correct_candidates) B1 &= ~15;
        if(ctx->o.paired) B1 &= ~1;
        if(B1 <= 0 || B1 >= B) return ctx->fail(CARE_EINVAL, "more than 2^32 candidate pairs in a batch of %d anchors", B);
        const int rc = runBatch(ctx, b0, B1, out, outRecords);
        if(rc != CARE_OK) return rc;
        return runBatch(ctx, b0 + uint64_t(B1), B - B1, out, outRecords);
    }
    // the batch's gathered bucket values into the pass statistic (once the batch is kept)
    k_add_gathered<<<(B + kAddPerBlock - 1) / kAddPerBlock, kBlock, 0, st>>>(ctx->gath.as<uint32_t>(), B,
                                                                          ctx->counters.as<unsigned long long>());
    HIPCHECK(hipGetLastError());
    const size_t T = std::max<size_t>(total, 1);
    HIPCHECK(ctx->ids.ensure(T * 4));
    HIPCHECK(ctx->anchorOf.ensure(T * 4));
    HIPCHECK(ctx->best.ensure(T * sizeof(AlignBest)));
    HIPCHECK(ctx->paired.ensure(T));
    HIPCHECK(ctx->kId.ensure(T * 4)); HIPCHECK(ctx->kShift.ensure(T * 4)); HIPCHECK(ctx->kLen.ensure(T * 4));
    HIPCHECK(ctx->kNops.ensure(T * 4)); HIPCHECK(ctx->kOvl.ensure(T * 4)); HIPCHECK(ctx->kW.ensure(T * 4));
    HIPCHECK(ctx->kOrient.ensure(T)); HIPCHECK(ctx->kPaired.ensure(T));
    if(ctx->o.keep_intermediates) HIPCHECK(ctx->dbg.ensure(T * 11 * sizeof(int)));
    // anchorOf is read by the wavefront-per-pair alignment and by candidate correction
    const int wm = ctx->o.align_variant == 1 ? 0 : alignLaneWords(ctx->W);
    const bool needAnchorOf = wm == 0 || ctx->o.correct_candidates != 0;
    k_compact<<<(B + kBlock / kWave - 1) / (kBlock / kWave), kBlock, 0, st>>>(
        ctx->raw.as<uint32_t>(), capPerAnchor, ctx->offsets.as<uint32_t>(), B, ctx->ids.as<uint32_t>(),
        needAnchorOf ? ctx->anchorOf.as<uint32_t>() : nullptr);
    HIPCHECK(hipGetLastError());
    if(ctx->o.paired){
        const int pairs = B / 2;
        k_pair_flags<<<(pairs + kBlock / kWave - 1) / (kBlock / kWave), kBlock, 0, st>>>(ctx->offsets.as<uint32_t>(), ctx->ids.as<uint32_t>(),
                                                          ctx->paired.as<uint8_t>(), pairs);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipEventRecord(ctx->ev[1], st));

    AlignParams ap;
    fillAlignCommon(ctx, ap);
    ap.ids = ctx->ids.as<uint32_t>(); ap.anchorOf = ctx->anchorOf.as<uint32_t>(); ap.b0 = b0; ap.total = total;
    ap.best = ctx->best.as<AlignBest>(); ap.dbg = ctx->o.keep_intermediates ? ctx->dbg.as<int>() : nullptr;
    ap.offsets = ctx->offsets.as<uint32_t>(); ap.B = B; ap.maxLen = ctx->maxLen;
    ap.order = order;
    if(total) HIPCHECK(launchAlign(ctx, ap, B, total, wm, st));
    HIPCHECK(hipEventRecord(ctx->ev[2], st));

    const int maxCols = ctx->maxCols;
    const int slotBytes = ((8 + 4 + std::max(ctx->maxLen, 5 * (ctx->maxLen / 7 + 1))) + 15) / 16 * 16;
    HIPCHECK(ctx->slots.ensure(size_t(B) * slotBytes));
    HIPCHECK(ctx->recBytes.ensure(size_t(B + 1) * 4));
    HIPCHECK(ctx->recOff64.ensure(size_t(B + 1) * 8));
    HIPCHECK(hipMemsetAsync(ctx->recBytes.p, 0, size_t(B + 1) * 4, st));
    if(ctx->o.keep_intermediates){
        HIPCHECK(ctx->info.ensure(size_t(B) * kAnchorInfo * 4));
        HIPCHECK(ctx->msaDbg.ensure(size_t(B) * ctx->dbgStride));
    }
    MsaParams mp;
    fillMsaCommon(ctx, mp);
    mp.b0 = b0;
    mp.offsets = ctx->offsets.as<uint32_t>(); mp.ids = ctx->ids.as<uint32_t>(); mp.best = ctx->best.as<AlignBest>();
    mp.paired = ctx->o.paired ? ctx->paired.as<uint8_t>() : nullptr;
    mp.kId = ctx->kId.as<uint32_t>(); mp.kShift = ctx->kShift.as<int>(); mp.kLen = ctx->kLen.as<int>();
    mp.kNops = ctx->kNops.as<int>(); mp.kOvl = ctx->kOvl.as<int>(); mp.kW = ctx->kW.as<float>();
    mp.kOrient = ctx->kOrient.as<uint8_t>(); mp.kPaired = ctx->kPaired.as<uint8_t>();
    mp.maxCols = maxCols;
    mp.tile = msaTile(maxCols, ctx->maxLen, ctx->W, ctx->o.use_quality_scores != 0);
    mp.slots = ctx->slots.as<uint8_t>(); mp.slotBytes = slotBytes; mp.recBytes = ctx->recBytes.as<uint32_t>();
    mp.info = ctx->o.keep_intermediates ? ctx->info.as<int>() : nullptr;
    mp.msaDbg = ctx->o.keep_intermediates ? ctx->msaDbg.as<uint8_t>() : nullptr;
    mp.dbgStride = ctx->dbgStride;
    mp.error = ctx->counters.as<int>() + 2;
    mp.order = order; mp.B = B;
    const bool cands = ctx->o.correct_candidates != 0;
    const int candSlotBytes = candidateSlotBytes(ctx->maxLen);
    if(cands){
        HIPCHECK(ctx->candSlots.ensure(T * candSlotBytes));
        HIPCHECK(ctx->candBytes.ensure(T * 4));
        mp.corrCands = 1;
        mp.candSlots = ctx->candSlots.as<uint8_t>(); mp.candSlotBytes = candSlotBytes;
        mp.candBytes = ctx->candBytes.as<uint32_t>(); mp.hqAnchor = ctx->hqAnchor.as<uint8_t>();
    }
    // the whole MSA is kept for candidate correction and the intermediate dumps; the
    // plain correction builds the columns under the anchor only
    const bool wholeMsa = cands || ctx->o.keep_intermediates;
    const int kc = wholeMsa ? waveMsaColumnsPerLane(maxCols) : waveMsaAnchorSlots(ctx->maxLen);
    if(kc > 0){
        // one wavefront per anchor, MSA columns in registers
        const bool q = ctx->o.use_quality_scores != 0;
        mp.tile = ctx->o.msa_tile ? ctx->o.msa_tile : defaultWaveMsaTile(q);
        const size_t lds = waveMsaLdsBytes(mp.tile, ctx->W, ctx->maxLen, q, mp.seamMode != 0 || mp.featOut != nullptr,
                                           mp.maxCols);
        HIPCHECK(launchWaveMsa(mp, B, kc, wholeMsa, q, lds, st));
    }else{
        const size_t msaLds = msaLdsBytes(maxCols, ctx->maxLen, ctx->W, mp.tile, ctx->o.use_quality_scores != 0);
        if(msaLds > 65536) HIPCHECK(hipFuncSetAttribute((const void*)k_msa, hipFuncAttributeMaxDynamicSharedMemorySize, int(msaLds)));
        k_msa<<<B, kBlock, msaLds, st>>>(mp);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipEventRecord(ctx->ev[3], st));

    if(cands){
        ctx->lastB0 = b0; ctx->lastB = uint64_t(B); ctx->lastTotal = total;
        return gatherInsertionOrder(ctx, b0, B, total, slotBytes, candSlotBytes, out, outRecords);
    }

    // record compaction
    k_widen<<<(B + 1 + 255) / 256, 256, 0, st>>>(ctx->recBytes.as<uint32_t>(), ctx->recOff64.as<uint64_t>(), B + 1);
    HIPCHECK(hipGetLastError());
    HIPCHECK(ctx->recOff.ensure(size_t(B + 1) * 8));
    tb = 0;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, ctx->recOff64.as<uint64_t>(), ctx->recOff.as<uint64_t>(), B + 1, st));
    HIPCHECK(ctx->scanTmp.ensure(std::max<size_t>(tb, 256)));
    tb = ctx->scanTmp.bytes;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(ctx->scanTmp.p, tb, ctx->recOff64.as<uint64_t>(), ctx->recOff.as<uint64_t>(), B + 1, st));
    uint64_t outBytes = 0;
    int errFlag = 0;
    HIPCHECK(hipMemcpyAsync(&outBytes, ctx->recOff.as<uint64_t>() + B, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&errFlag, ctx->counters.as<int>() + 2, 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if(errFlag) return ctx->fail(CARE_EINVAL, "MSA column bound exceeded in batch at %llu", (unsigned long long)b0);
    HIPCHECK(ctx->outBuf.ensure(std::max<uint64_t>(outBytes, 1)));
    k_gather<<<B, kBlock, 0, st>>>(ctx->slots.as<uint8_t>(), slotBytes, ctx->recBytes.as<uint32_t>(),
                                   ctx->recOff.as<uint64_t>(), ctx->outBuf.as<uint8_t>(), B);
    HIPCHECK(hipGetLastError());
    const size_t old = out.size();
    HIPCHECK(out.resize(old + outBytes));
    if(outBytes) HIPCHECK(hipMemcpyAsync(out.data() + old, ctx->outBuf.p, outBytes, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipEventRecord(ctx->ev[4], st));
    HIPCHECK(hipStreamSynchronize(st));

    // count records of this batch
    size_t q = old;
    while(q < out.size()){
        uint32_t flags;
        memcpy(&flags, out.data() + q + 4, 4);
        q += 8 + (flags & ((1u << 29) - 1));
        outRecords++;
    }
    float ms;
    HIPCHECK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1])); ctx->t.candidates_ms += ms;
    HIPCHECK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[5])); ctx->t.cand_kernel_ms += ms;
    HIPCHECK(hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2])); ctx->t.align_ms += ms; ctx->t.align_kernel_ms += ms;
    HIPCHECK(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3])); ctx->t.msa_ms += ms; ctx->t.msa_kernel_ms += ms;
    HIPCHECK(hipEventElapsedTime(&ms, ctx->ev[3], ctx->ev[4])); ctx->t.output_ms += ms;
    ctx->t.candidates += total;
    ctx->lastB0 = b0; ctx->lastB = uint64_t(B); ctx->lastTotal = total;
    return CARE_OK;
}

// Pipelined staged batches: no host synchronization inside a batch. The candidate arrays
// are sized before a batch's candidates are counted (`cap` pairs per batch); a device
// check flags a batch with more, whose kernels then see no candidates, and the host redoes
// that batch alone with runBatch (exact sizes, same records). Batch i+1 is enqueued before
// batch i's records are copied out: each batch slot (i % 2) has its own record buffer,
// pinned status words and stage events, and the copy runs on the copy stream. With
// use_graphs each batch's device work is one captured HIP graph (no memset node), replayed
// while the batch's parameters and buffers hold.
struct StagedLaunch {
    CandParams cp;
    AlignParams ap;
    MsaParams mp;
    uint64_t cap;           // candidate pairs the batch's arrays hold
    int B, slot, wm, kc, slotBytes, capPerAnchor;
    int wholeMsa, q;
    size_t waveLds, candLds, msaLds;
};

// Status words of batch slot s: in `counters` at [64 + 32 s): over-capacity flag (u32),
// candidate pairs (u32, saturated), bucket values gathered (u64), records (u64), candidate
// pairs (u64); copied after the record bytes (u64) and the error flag (i32) to the pinned
// words at [64 + 64 s): bytes 0, 8, then 16 on.
uint32_t* slotStatus(care_ctx* ctx, int s){ return ctx->counters.as<uint32_t>() + 16 + 8 * s; }
uint8_t* slotPinned(care_ctx* ctx, int s){ return reinterpret_cast<uint8_t*>(ctx->pinned) + 64 + 64 * s; }

// The device work of one pipelined batch, after its processing order (batchOrder).
// Stage events only when not captured.
hipError_t enqueueStaged(care_ctx* ctx, const StagedLaunch& L, bool events){
    hipStream_t st = ctx->stream;
    const int B = L.B;
    hipEvent_t* ev = ctx->pipeEv[L.slot];
    uint32_t* status = slotStatus(ctx, L.slot);
    hipError_t e = hipSuccess;
#define ENQ(launch) do { launch; e = hipGetLastError(); if(e != hipSuccess) return e; } while(0)
#define ENQE(expr) do { e = (expr); if(e != hipSuccess) return e; } while(0)
    uint32_t* overCount = reinterpret_cast<uint32_t*>(ctx->candOver.as<uint8_t>() + size_t(B) * 4);
    ENQ((k_fill_u32<<<dim3(uint32_t((B + 1 + 255) / 256)), 256, 0, st>>>(L.cp.uniq, uint64_t(B + 1), 0u)));
    ENQ((k_zero_u32<<<1, 64, 0, st>>>(overCount, 1)));
    ENQ((k_zero_u32<<<1, 64, 0, st>>>(status, 8)));
    ENQ((k_candidates_wave<<<B, kWave, L.waveLds, st>>>(L.cp, ctx->candOver.as<uint32_t>(), overCount)));
    ENQ((k_candidates_over<<<std::min(B, 4096), kBlock, L.candLds, st>>>(L.cp, ctx->candOver.as<uint32_t>(), overCount)));
    ENQ((k_add_gathered<<<(B + kAddPerBlock - 1) / kAddPerBlock, kBlock, 0, st>>>(
        ctx->gath.as<uint32_t>(), B, reinterpret_cast<unsigned long long*>(status + 2))));
    ENQ((k_add_gathered<<<(B + kAddPerBlock - 1) / kAddPerBlock, kBlock, 0, st>>>(
        L.cp.uniq, B, reinterpret_cast<unsigned long long*>(status + 6))));
    if(events) ENQE(hipEventRecord(ev[5], st));
    size_t tb = ctx->scanTmp.bytes;
    ENQE(hipcub::DeviceScan::ExclusiveSum(ctx->scanTmp.p, tb, ctx->uniq.as<uint32_t>(), ctx->offsets.as<uint32_t>(), B + 1, st));
    ENQ((k_cap_status<<<1, 64, 0, st>>>(L.cap, status)));
    ENQ((k_cap_clear<<<(B + 1 + 255) / 256, 256, 0, st>>>(ctx->offsets.as<uint32_t>(), B, status)));
    ENQ((k_compact<<<(B + kBlock / kWave - 1) / (kBlock / kWave), kBlock, 0, st>>>(
        ctx->raw.as<uint32_t>(), L.capPerAnchor, ctx->offsets.as<uint32_t>(), B, ctx->ids.as<uint32_t>(), nullptr)));
    if(ctx->o.paired){
        const int pairs = B / 2;
        ENQ((k_pair_flags<<<(pairs + kBlock / kWave - 1) / (kBlock / kWave), kBlock, 0, st>>>(ctx->offsets.as<uint32_t>(), ctx->ids.as<uint32_t>(),
                                                               ctx->paired.as<uint8_t>(), pairs)));
    }
    if(events) ENQE(hipEventRecord(ev[1], st));
    ENQE(launchAlign(ctx, L.ap, B, 0, L.wm, st));
    if(events) ENQE(hipEventRecord(ev[2], st));
    ENQ((k_fill_u32<<<dim3(uint32_t((B + 1 + 255) / 256)), 256, 0, st>>>(ctx->recBytes.as<uint32_t>(), uint64_t(B + 1), 0u)));
    if(L.kc > 0){
        ENQE(launchWaveMsa(L.mp, B, L.kc, L.wholeMsa != 0, L.q != 0, L.msaLds, st));
    }else{
        ENQ((k_msa<<<B, kBlock, L.msaLds, st>>>(L.mp)));
    }
    if(events) ENQE(hipEventRecord(ev[3], st));
    ENQ((k_widen_count<<<(B + 1 + kWidenPerBlock - 1) / kWidenPerBlock, 256, 0, st>>>(ctx->recBytes.as<uint32_t>(), ctx->recOff64.as<uint64_t>(), B + 1,
                                                            reinterpret_cast<unsigned long long*>(status + 4))));
    tb = ctx->scanTmp.bytes;
    ENQE(hipcub::DeviceScan::ExclusiveSum(ctx->scanTmp.p, tb, ctx->recOff64.as<uint64_t>(), ctx->recOff.as<uint64_t>(), B + 1, st));
    ENQ((k_gather<<<B, kBlock, 0, st>>>(ctx->slots.as<uint8_t>(), L.slotBytes, ctx->recBytes.as<uint32_t>(),
                                        ctx->recOff.as<uint64_t>(), ctx->pipeOut[L.slot].as<uint8_t>(), B)));
    uint8_t* pw = slotPinned(ctx, L.slot);
    ENQE(hipMemcpyAsync(pw, ctx->recOff.as<uint64_t>() + B, 8, hipMemcpyDeviceToHost, st));
    ENQE(hipMemcpyAsync(pw + 8, ctx->counters.as<int>() + 2, 4, hipMemcpyDeviceToHost, st));
    ENQE(hipMemcpyAsync(pw + 16, status, 32, hipMemcpyDeviceToHost, st));
#undef ENQ
#undef ENQE
    return hipSuccess;
}

// Buffers of pipelined batches of at most Bmax anchors and `cap` candidate pairs, sized
// before any batch is enqueued (growing one frees the old allocation, which waits for the
// device), and the launch constants shared by the batches.
int prepareStaged(care_ctx* ctx, int Bmax, uint64_t cap, StagedLaunch& L){
    hipStream_t st = ctx->stream;
    std::memset(&L, 0, sizeof(L));          // padding included: the bytes key the batch's graph
    L.capPerAnchor = ctx->M * ctx->thr;
    int P = 64;
    while(P <= L.capPerAnchor) P <<= 1;
    L.cap = cap;
    L.slotBytes = recordSlotBytes(ctx->maxLen);
    L.wm = alignLaneWords(ctx->W);
    const size_t T = size_t(std::max<uint64_t>(cap, 1));
    HIPCHECK(ctx->raw.ensure(size_t(Bmax) * L.capPerAnchor * sizeof(uint32_t)));
    HIPCHECK(ctx->uniq.ensure(size_t(Bmax + 1) * 4));
    HIPCHECK(ctx->offsets.ensure(size_t(Bmax + 1) * 4));
    HIPCHECK(ctx->gath.ensure(size_t(Bmax) * 4));
    HIPCHECK(ctx->candOver.ensure(size_t(Bmax) * 4 + 16));
    HIPCHECK(ctx->ids.ensure(T * 4));
    HIPCHECK(ctx->best.ensure(T * sizeof(AlignBest)));
    if(ctx->o.paired) HIPCHECK(ctx->paired.ensure(T));
    HIPCHECK(ctx->kId.ensure(T * 4)); HIPCHECK(ctx->kShift.ensure(T * 4)); HIPCHECK(ctx->kLen.ensure(T * 4));
    HIPCHECK(ctx->kNops.ensure(T * 4)); HIPCHECK(ctx->kOvl.ensure(T * 4)); HIPCHECK(ctx->kW.ensure(T * 4));
    HIPCHECK(ctx->kOrient.ensure(T)); HIPCHECK(ctx->kPaired.ensure(T));
    HIPCHECK(ctx->slots.ensure(size_t(Bmax) * L.slotBytes));
    HIPCHECK(ctx->recBytes.ensure(size_t(Bmax + 1) * 4));
    HIPCHECK(ctx->recOff64.ensure(size_t(Bmax + 1) * 8));
    HIPCHECK(ctx->recOff.ensure(size_t(Bmax + 1) * 8));
    for(auto& b : ctx->pipeOut) HIPCHECK(b.ensure(size_t(Bmax) * L.slotBytes));
    HIPCHECK(ctx->orderFlags.ensure(ctx->n + 16));
    HIPCHECK(ctx->batchOrd.ensure(size_t(Bmax) * 4 + 16));
    // one scan workspace for the batch order selection and both scans (sizes grow with
    // the item count)
    size_t t1 = 0, t2 = 0, t3 = 0;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, t1, ctx->uniq.as<uint32_t>(), ctx->offsets.as<uint32_t>(), Bmax + 1, st));
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, t2, ctx->recOff64.as<uint64_t>(), ctx->recOff.as<uint64_t>(), Bmax + 1, st));
    HIPCHECK(hipcub::DeviceSelect::Flagged(nullptr, t3, ctx->order.as<uint32_t>(), ctx->orderFlags.as<uint8_t>(),
                                           ctx->batchOrd.as<uint32_t>(), ctx->batchOrd.as<uint32_t>(), int(ctx->n), st));
    HIPCHECK(ctx->scanTmp.ensure(std::max({t1, t2, t3, size_t(256)})));

    CandParams& cp = L.cp;
    cp.sig = ctx->sig.as<uint64_t>(); cp.n = ctx->n; cp.M = ctx->M; cp.k = ctx->k; cp.thr = ctx->thr;
    cp.bucketOf = ctx->bucketOf.as<uint2>();
    cp.excludeAmbiguous = ctx->o.exclude_ambiguous; cp.len = ctx->len.as<int>(); cp.ambig = ctx->ambig.as<uint8_t>();
    cp.slots = ctx->slotPtrs.as<const Slot*>(); cp.caps = ctx->capsDev.as<uint32_t>(); cp.values = ctx->values.as<uint32_t>();
    cp.capPerAnchor = L.capPerAnchor; cp.hashSlots = P; cp.raw = ctx->raw.as<uint32_t>(); cp.uniq = ctx->uniq.as<uint32_t>();
    cp.gathered = nullptr; cp.gath = ctx->gath.as<uint32_t>();
    L.waveLds = candWaveLdsBytes(L.capPerAnchor);
    L.candLds = candLdsBytes(P, L.capPerAnchor);
    if(L.waveLds > 65536) HIPCHECK(hipFuncSetAttribute((const void*)k_candidates_wave, hipFuncAttributeMaxDynamicSharedMemorySize, int(L.waveLds)));
    if(L.candLds > 65536) HIPCHECK(hipFuncSetAttribute((const void*)k_candidates_over, hipFuncAttributeMaxDynamicSharedMemorySize, int(L.candLds)));

    AlignParams& ap = L.ap;
    fillAlignCommon(ctx, ap);
    ap.ids = ctx->ids.as<uint32_t>(); ap.anchorOf = nullptr; ap.total = 0;
    ap.best = ctx->best.as<AlignBest>(); ap.dbg = nullptr;
    ap.offsets = ctx->offsets.as<uint32_t>(); ap.maxLen = ctx->maxLen;

    MsaParams& mp = L.mp;
    fillMsaCommon(ctx, mp);
    mp.offsets = ctx->offsets.as<uint32_t>(); mp.ids = ctx->ids.as<uint32_t>(); mp.best = ctx->best.as<AlignBest>();
    mp.paired = ctx->o.paired ? ctx->paired.as<uint8_t>() : nullptr;
    mp.kId = ctx->kId.as<uint32_t>(); mp.kShift = ctx->kShift.as<int>(); mp.kLen = ctx->kLen.as<int>();
    mp.kNops = ctx->kNops.as<int>(); mp.kOvl = ctx->kOvl.as<int>(); mp.kW = ctx->kW.as<float>();
    mp.kOrient = ctx->kOrient.as<uint8_t>(); mp.kPaired = ctx->kPaired.as<uint8_t>();
    mp.maxCols = ctx->maxCols;
    mp.slots = ctx->slots.as<uint8_t>(); mp.slotBytes = L.slotBytes; mp.recBytes = ctx->recBytes.as<uint32_t>();
    mp.info = nullptr; mp.msaDbg = nullptr; mp.dbgStride = ctx->dbgStride;
    mp.error = ctx->counters.as<int>() + 2;
    L.q = ctx->o.use_quality_scores != 0;
    L.wholeMsa = 0;
    L.kc = waveMsaAnchorSlots(ctx->maxLen);
    if(L.kc > 0){
        mp.tile = ctx->o.msa_tile ? ctx->o.msa_tile : defaultWaveMsaTile(L.q != 0);
        L.msaLds = waveMsaLdsBytes(mp.tile, ctx->W, ctx->maxLen, L.q != 0, mp.seamMode != 0 || mp.featOut != nullptr, mp.maxCols);
    }else{
        mp.tile = msaTile(ctx->maxCols, ctx->maxLen, ctx->W, L.q != 0);
        L.msaLds = msaLdsBytes(ctx->maxCols, ctx->maxLen, ctx->W, mp.tile, L.q != 0);
        if(L.msaLds > 65536) HIPCHECK(hipFuncSetAttribute((const void*)k_msa, hipFuncAttributeMaxDynamicSharedMemorySize, int(L.msaLds)));
    }
    return CARE_OK;
}

// Enqueues batch [b0, b0+B) into slot `slot`: its processing order, then its device work,
// eagerly or as a graph.
int launchStaged(care_ctx* ctx, StagedLaunch& L, uint64_t b0, int B, int slot){
    hipStream_t st = ctx->stream;
    hipEvent_t* ev = ctx->pipeEv[slot];
    L.B = B; L.slot = slot;
    HIPCHECK(hipEventRecord(ev[0], st));
    const uint32_t* order = nullptr;
    {
        const int rc = batchOrder(ctx, b0, B, &order);
        if(rc != CARE_OK) return rc;
    }
    L.cp.b0 = b0; L.cp.B = B; L.cp.order = order;
    L.ap.b0 = b0; L.ap.B = B; L.ap.order = order;
    L.mp.b0 = b0; L.mp.B = B; L.mp.order = order;
    if(!ctx->o.use_graphs){
        const hipError_t e = enqueueStaged(ctx, L, true);
        if(e != hipSuccess) return ctx->fail(CARE_EHIP, "staged batch launch failed: %s", hipGetErrorString(e));
        HIPCHECK(hipEventRecord(ev[4], st));
        return CARE_OK;
    }
    // the key: the launch constants and parameters, and every address the captured work
    // uses beyond them (scan workspace, status words, record buffers, pinned words)
    const uint64_t extra[] = {uint64_t(reinterpret_cast<uintptr_t>(ctx->scanTmp.p)), uint64_t(ctx->scanTmp.bytes),
                              uint64_t(reinterpret_cast<uintptr_t>(ctx->offsets.p)),
                              uint64_t(reinterpret_cast<uintptr_t>(ctx->candOver.p)),
                              uint64_t(reinterpret_cast<uintptr_t>(ctx->counters.p)),
                              uint64_t(reinterpret_cast<uintptr_t>(ctx->recOff64.p)),
                              uint64_t(reinterpret_cast<uintptr_t>(ctx->recOff.p)),
                              uint64_t(reinterpret_cast<uintptr_t>(ctx->pipeOut[slot].p)),
                              uint64_t(reinterpret_cast<uintptr_t>(ctx->pinned))};
    std::vector<uint8_t> key(sizeof(L) + sizeof(extra));
    std::memcpy(key.data(), &L, sizeof(L));
    std::memcpy(key.data() + sizeof(L), extra, sizeof(extra));
    auto it = ctx->stagedGraphs.find(key);
    if(it == ctx->stagedGraphs.end()){
        if(ctx->stagedGraphs.size() >= 16){
            // a pass has at most a few batch shapes; drop graphs of earlier ones (none is
            // in flight: the previous pass ended with a synchronization)
            HIPCHECK(hipStreamSynchronize(st));
            for(auto& g : ctx->stagedGraphs) (void)hipGraphExecDestroy(g.second);
            ctx->stagedGraphs.clear();
        }
        HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const hipError_t e = enqueueStaged(ctx, L, false);
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(st, &g);
        if(e != hipSuccess || ee != hipSuccess){
            if(g) (void)hipGraphDestroy(g);
            return ctx->fail(CARE_EHIP, "capturing a staged batch failed: %s", hipGetErrorString(e != hipSuccess ? e : ee));
        }
        hipGraphExec_t x = nullptr;
        const hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if(ie != hipSuccess) return ctx->fail(CARE_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(ie));
        it = ctx->stagedGraphs.emplace(std::move(key), x).first;
        ctx->graphsBuilt += 1;
    }
    HIPCHECK(hipGraphLaunch(it->second, st));
    HIPCHECK(hipEventRecord(ev[4], st));
    return CARE_OK;
}

// Waits for batch [b0, b0+B) of slot `slot`, appends its records to `out` (or redoes the
// batch with runBatch when it had more candidates than its arrays hold) and adds its
// timings. `redone` is set when the batch was redone.
int finishStaged(care_ctx* ctx, uint64_t b0, int B, int slot, HostRecords& out, size_t& outRecords,
                 unsigned long long& gathered, bool& redone){
    hipEvent_t* ev = ctx->pipeEv[slot];
    HIPCHECK(hipEventSynchronize(ev[4]));
    const uint8_t* pw = slotPinned(ctx, slot);
    uint64_t outBytes = 0, gath = 0, recs = 0;
    int errFlag = 0;
    uint32_t over = 0, total = 0;
    std::memcpy(&outBytes, pw, 8);
    std::memcpy(&errFlag, pw + 8, 4);
    std::memcpy(&over, pw + 16, 4);
    std::memcpy(&total, pw + 20, 4);
    std::memcpy(&gath, pw + 24, 8);
    std::memcpy(&recs, pw + 32, 8);
    if(errFlag) return ctx->fail(CARE_EINVAL, "MSA column bound exceeded in batch at %llu", (unsigned long long)b0);
    ctx->candPerAnchor = std::max(ctx->candPerAnchor, double(total) / double(std::max(B, 1)));
    float ms = 0;
    if(ctx->o.use_graphs){
        HIPCHECK(hipEventElapsedTime(&ms, ev[0], ev[4])); ctx->t.graph_ms += ms;
    }else{
        HIPCHECK(hipEventElapsedTime(&ms, ev[0], ev[1])); ctx->t.candidates_ms += ms;
        HIPCHECK(hipEventElapsedTime(&ms, ev[0], ev[5])); ctx->t.cand_kernel_ms += ms;
        HIPCHECK(hipEventElapsedTime(&ms, ev[1], ev[2])); ctx->t.align_ms += ms; ctx->t.align_kernel_ms += ms;
        HIPCHECK(hipEventElapsedTime(&ms, ev[2], ev[3])); ctx->t.msa_ms += ms; ctx->t.msa_kernel_ms += ms;
        HIPCHECK(hipEventElapsedTime(&ms, ev[3], ev[4])); ctx->t.output_ms += ms;
    }
    redone = over != 0;
    if(redone){
        // the next batch may be in flight and uses the arrays runBatch may reallocate:
        // wait for it first
        HIPCHECK(hipStreamSynchronize(ctx->stream));
        ctx->t.batches_redone += 1;
        return runBatch(ctx, b0, B, out, outRecords);
    }
    gathered += gath;
    ctx->t.candidates += total;
    const size_t old = out.size();
    HIPCHECK(out.resize(old + outBytes));
    if(outBytes){
        HIPCHECK(hipMemcpyAsync(out.data() + old, ctx->pipeOut[slot].p, outBytes, hipMemcpyDeviceToHost, ctx->copyStream));
        HIPCHECK(hipStreamSynchronize(ctx->copyStream));
    }
    outRecords += recs;
    ctx->lastB0 = b0; ctx->lastB = uint64_t(B); ctx->lastTotal = total;
    return CARE_OK;
}

// Candidate pairs the pipelined arrays of a batch are sized for: the most per anchor seen
// so far on this context with a quarter of headroom; before any batch was counted, the
// option's count per anchor, else four per unit of coverage. Never below what the context
// used before, so that repeated passes keep their arrays and graphs.
uint64_t stagedCapacity(care_ctx* ctx, int Bmax){
    double per;
    if(ctx->candPerAnchor <= 0 && ctx->o.pipeline_capacity > 0){
        per = ctx->o.pipeline_capacity;
    }else{
        const int capPerAnchor = ctx->M * ctx->thr;
        per = ctx->candPerAnchor > 0 ? ctx->candPerAnchor * 1.25 + 8 : 4.0 * ctx->o.estimated_coverage + 16;
        per = std::min<double>(std::max(per, 32.0), capPerAnchor);
    }
    per = std::ceil(std::max(per, ctx->pipePerAnchor));
    ctx->pipePerAnchor = per;
    // at most a quarter of the free device memory at about 48 bytes per pair (batches with
    // more pairs are redone with exact sizes)
    size_t freeB = 0, totalB = 0;
    uint64_t memCap = uint64_t(UINT32_MAX);
    if(hipMemGetInfo(&freeB, &totalB) == hipSuccess) memCap = std::max<uint64_t>(uint64_t(freeB) / 4 / 48, uint64_t(Bmax) * 32);
    return std::min<uint64_t>({uint64_t(per) * uint64_t(Bmax), uint64_t(UINT32_MAX), memCap});
}

// Anchors [first, first+n) in pipelined batches of at most Bmax anchors.
int runPipelined(care_ctx* ctx, uint64_t first, uint64_t n, int Bmax, HostRecords& out, size_t& outRecords,
                 unsigned long long& gathered){
    std::vector<std::pair<uint64_t, int>> batches;
    for(uint64_t b0 = first; b0 < first + n;){
        const uint64_t e = std::min<uint64_t>(b0 + uint64_t(Bmax), first + n);
        batches.emplace_back(b0, int(e - b0));
        b0 = e;
    }
    const int Bfirst = batches.empty() ? 0 : batches[0].second;
    StagedLaunch L;
    int rc = prepareStaged(ctx, Bfirst, stagedCapacity(ctx, Bfirst), L);
    if(rc != CARE_OK) return rc;
    const size_t nb = batches.size();
    for(size_t i = 0; i < nb; i++){
        rc = launchStaged(ctx, L, batches[i].first, batches[i].second, int(i & 1));
        if(rc != CARE_OK) break;
        if(i == 0) continue;
        bool redone = false;
        rc = finishStaged(ctx, batches[i - 1].first, batches[i - 1].second, int((i - 1) & 1), out, outRecords, gathered, redone);
        if(rc != CARE_OK) break;
        ctx->t.batches++;
        if(redone){
            // runBatch may have reallocated the arrays (batch i, enqueued before it, has
            // ended): the batches after i get the current arrays, sized from what has been
            // seen; batch i may be redone as well
            HIPCHECK(hipStreamSynchronize(ctx->stream));
            rc = prepareStaged(ctx, Bfirst, std::max(L.cap, stagedCapacity(ctx, Bfirst)), L);
            if(rc != CARE_OK) break;
        }
    }
    if(rc == CARE_OK && nb){
        bool redone = false;
        rc = finishStaged(ctx, batches[nb - 1].first, batches[nb - 1].second, int((nb - 1) & 1), out, outRecords, gathered, redone);
        if(rc == CARE_OK) ctx->t.batches++;
    }
    if(rc != CARE_OK) (void)hipStreamSynchronize(ctx->stream);    // nothing left in flight on failure
    return rc;
}

}  // namespace

// Entry points run their bodies through guarded(): no C++ exception crosses the
// C-ABI (a failed host allocation is CARE_ENOMEM, any other exception CARE_EINVAL).
template<class F>
int guarded(care_ctx* h, F&& f){
    const char* what = "unexpected C++ exception";
    int code = CARE_EINVAL;
    try{
        return f();
    }catch(const std::bad_alloc&){
        what = "host allocation failed";
        code = CARE_ENOMEM;
    }catch(const std::exception& e){
        what = e.what();
    }catch(...){
    }
    if(h){
        try{ h->err = what; }catch(...){}
    }
    return code;
}

extern "C" {

void care_hip_default_opts(care_opts* o){
    std::memset(o, 0, sizeof(*o));
    o->kmerlength = 20;
    o->num_hash_maps = 48;
    o->estimated_coverage = 30.0f;
    o->min_overlap = 30;
    o->max_error_rate = 0.2f;
    o->min_overlap_ratio = 0.30f;
    o->estimated_errorrate = 0.06f;
    o->m_coverage = 0.6f;
    o->paired_filter_threshold = 0.06f;
    o->new_columns_to_correct = 15;
}

static int care_hip_create_body(const care_opts* opts, int device, care_ctx** out){
    if(!opts || !out) return CARE_EINVAL;
    *out = nullptr;
    if(opts->num_hash_maps < 1 || opts->num_hash_maps > kMaxMaps) return CARE_EINVAL;
    if(opts->kmerlength < 0 || opts->kmerlength > 32) return CARE_EINVAL;
    if(!(opts->estimated_coverage > 0.0f)) return CARE_EINVAL;
    if(opts->new_columns_to_correct < 0) return CARE_EINVAL;
    if(opts->msa_tile != 0 && (opts->msa_tile < 4 || opts->msa_tile > 64 || (opts->msa_tile & 3))) return CARE_EINVAL;
    if(opts->seeded_alignment != 0 && opts->seeded_alignment != 1) return CARE_EINVAL;
    if(opts->sync_batches != 0 && opts->sync_batches != 1) return CARE_EINVAL;
    if(opts->pipeline_capacity < 0) return CARE_EINVAL;
    int count = 0;
    if(hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) return CARE_ENODEV;
    auto* ctx = new (std::nothrow) care_ctx;
    if(!ctx) return CARE_ENOMEM;
    ctx->o = *opts;
    ctx->device = device;
    if(hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess){
        delete ctx;
        return CARE_EHIP;
    }
    // pinned words: [0, 64) the last pass, [64 + 64 s, 128 + 64 s) pipelined batch slot s
    if(hipHostMalloc(&ctx->pinned, 256, hipHostMallocDefault) != hipSuccess){ care_hip_destroy(ctx); return CARE_ENOMEM; }
    if(hipStreamCreateWithFlags(&ctx->copyStream, hipStreamNonBlocking) != hipSuccess){ care_hip_destroy(ctx); return CARE_EHIP; }
    for(auto& e : ctx->ev){
        if(hipEventCreate(&e) != hipSuccess){ care_hip_destroy(ctx); return CARE_EHIP; }
    }
    for(auto& row : ctx->pipeEv){
        for(auto& e : row){
            if(hipEventCreate(&e) != hipSuccess){ care_hip_destroy(ctx); return CARE_EHIP; }
        }
    }
    *out = ctx;
    return CARE_OK;
}

void care_hip_destroy(care_ctx* ctx){
    if(!ctx) return;
    (void)hipSetDevice(ctx->device);
    if(ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for(DBuf* b : {&ctx->enc, &ctx->len, &ctx->qual, &ctx->ambig, &ctx->sig, &ctx->hilo, &ctx->rcenc, &ctx->qw, &ctx->bucketOf, &ctx->gath, &ctx->order, &ctx->orderFlags, &ctx->batchOrd, &ctx->slotPtrs,
                   &ctx->capsDev, &ctx->values, &ctx->raw, &ctx->uniq, &ctx->offsets, &ctx->ids, &ctx->anchorOf,
                   &ctx->best, &ctx->dbg, &ctx->paired, &ctx->kId, &ctx->kShift, &ctx->kLen, &ctx->kNops, &ctx->kOvl,
                   &ctx->kW, &ctx->kOrient, &ctx->kPaired, &ctx->slots, &ctx->recBytes, &ctx->recOff, &ctx->recOff64,
                   &ctx->outBuf, &ctx->info, &ctx->msaDbg, &ctx->counters, &ctx->scanTmp, &ctx->scratch, &ctx->candOver,
                   &ctx->seam, &ctx->featBuf, &ctx->metaBuf, &ctx->fqText[0], &ctx->fqText[1], &ctx->fqHdr,
                   &ctx->fqHdrLen, &ctx->fqSeq, &ctx->fqQual, &ctx->mRec, &ctx->mOff, &ctx->mKeyIn, &ctx->mKeyOut,
                   &ctx->mIdxIn, &ctx->mIdxOut, &ctx->mRunBeg, &ctx->mRunEnd, &ctx->mPlan, &ctx->mSize, &ctx->mOutOff,
                   &ctx->mText, &ctx->mErr, &ctx->candSlots, &ctx->candBytes, &ctx->insSize, &ctx->insOff,
                   &ctx->insSrc, &ctx->hqAnchor, &ctx->seeds, &ctx->pipeOut[0], &ctx->pipeOut[1]}){
        b->release();
    }
    if(ctx->fusedExec) (void)hipGraphExecDestroy(ctx->fusedExec);
    for(auto& g : ctx->stagedGraphs) (void)hipGraphExecDestroy(g.second);
    if(ctx->pinned) (void)hipHostFree(ctx->pinned);
    ctx->records.release();
    for(auto& b : ctx->slotBufs) b.release();
    for(auto& e : ctx->ev) if(e) (void)hipEventDestroy(e);
    for(auto& row : ctx->pipeEv) for(auto& e : row) if(e) (void)hipEventDestroy(e);
    if(ctx->copyStream) (void)hipStreamDestroy(ctx->copyStream);
    for(int t = 0; t < kUpWorkers; t++){
        for(int b = 0; b < 2; b++) if(ctx->upEv[t][b]) (void)hipEventDestroy(ctx->upEv[t][b]);
        if(ctx->upStream[t]) (void)hipStreamDestroy(ctx->upStream[t]);
    }
    if(ctx->upPinned) (void)hipHostFree(ctx->upPinned);
    if(ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

static int care_hip_device_info_body(int device, char* name, size_t cap, int* compute_units){
    int count = 0;
    if(hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return CARE_ENODEV;
    hipDeviceProp_t prop;
    if(hipGetDeviceProperties(&prop, device) != hipSuccess) return CARE_EHIP;
    if(name && cap){
        std::snprintf(name, cap, "%s %s", prop.name, prop.gcnArchName);
    }
    if(compute_units) *compute_units = prop.multiProcessorCount;
    return CARE_OK;
}

const char* care_hip_last_error(const care_ctx* ctx){
    return ctx ? ctx->err.c_str() : "no context";
}

// Shapes and options of a read set (shared by the array and the FASTQ entry points).
// Host-to-device copy of `rows` rows of a (pageable) host buffer, row pitch `spitch`,
// into device rows of pitch `dpitch` >= spitch whose last dpitch - spitch bytes are
// zeroed; complete on the device when it returns. Large copies are cut into one slice of
// rows per worker thread; a worker packs its rows into its two pinned chunk buffers and
// transfers them on its own stream, so the host copies of the workers and their transfers
// overlap. Work already enqueued on `st` finishes first (it may still read the
// destination). A linear copy is one row of `bytes` with both pitches equal to it.
// The transfer workers' streams, events and pinned chunk buffers (made on first use);
// *T = the number of workers.
static int transferWorkers(care_ctx* ctx, int* T){
    *T = int(std::min<unsigned>(kUpWorkers, std::max(1u, std::thread::hardware_concurrency())));
    if(!ctx->upPinned) HIPCHECK(hipHostMalloc(&ctx->upPinned, size_t(kUpWorkers) * 2 * kUpChunk, hipHostMallocDefault));
    for(int t = 0; t < *T; t++){
        if(!ctx->upStream[t]) HIPCHECK(hipStreamCreateWithFlags(&ctx->upStream[t], hipStreamNonBlocking));
        for(int b = 0; b < 2; b++){
            if(!ctx->upEv[t][b]) HIPCHECK(hipEventCreateWithFlags(&ctx->upEv[t][b], hipEventDisableTiming));
        }
    }
    return CARE_OK;
}

// Runs work(t) for t in [0, T) on T host threads (t = 0 on the calling thread).
extern "C++" {
template<class F>
static int onTransferThreads(care_ctx* ctx, int T, F&& work){
    std::vector<std::thread> pool;
    pool.reserve(T);
    try{
        for(int t = 1; t < T; t++) pool.emplace_back(work, t);
    }catch(...){
        for(auto& th : pool) th.join();
        return ctx->fail(CARE_ENOMEM, "could not start the transfer threads");
    }
    work(0);
    for(auto& th : pool) th.join();
    return CARE_OK;
}
}  // extern "C++"

static int uploadRowsH2D(care_ctx* ctx, void* dst, size_t dpitch, const void* src, size_t spitch, uint64_t rows,
                  hipStream_t st){
    const bool linear = rows == 1 && dpitch == spitch;
    if(!linear && (dpitch > kUpChunk || spitch > dpitch))
        return ctx->fail(CARE_EINVAL, "upload rows of %zu bytes at pitch %zu do not fit a transfer chunk", spitch, dpitch);
    HIPCHECK(hipStreamSynchronize(st));
    const size_t bytes = size_t(rows) * dpitch;
    int T = 0;
    {
        const int rc = transferWorkers(ctx, &T);
        if(rc != CARE_OK) return rc;
    }
    // work units: bytes of a linear copy, rows of a pitched one (whole rows per chunk)
    const uint64_t units = linear ? bytes : rows;
    const uint64_t perChunk = linear ? kUpChunk : kUpChunk / dpitch;
    const uint64_t slice = (units + T - 1) / T;
    std::vector<hipError_t> errs(T, hipSuccess);
    auto work = [&](int t){
        hipError_t e = hipSetDevice(ctx->device);
        const uint64_t beg = std::min(units, uint64_t(t) * slice), end = std::min(units, beg + slice);
        int k = 0;
        for(uint64_t u = beg; e == hipSuccess && u < end; u += perChunk, k++){
            const int b = k & 1;
            const uint64_t n = std::min(perChunk, end - u);
            if(k >= 2) e = hipEventSynchronize(ctx->upEv[t][b]);     // that buffer's last transfer
            if(e != hipSuccess) break;
            char* buf = static_cast<char*>(ctx->upPinned) + (size_t(t) * 2 + size_t(b)) * kUpChunk;
            const char* s = static_cast<const char*>(src);
            size_t len, off;
            if(linear){
                std::memcpy(buf, s + u, n);
                len = n; off = u;
            }else{
                for(uint64_t r = 0; r < n; r++){
                    std::memcpy(buf + r * dpitch, s + (u + r) * spitch, spitch);
                    if(dpitch > spitch) std::memset(buf + r * dpitch + spitch, 0, dpitch - spitch);
                }
                len = n * dpitch; off = u * dpitch;
            }
            e = hipMemcpyAsync(static_cast<char*>(dst) + off, buf, len, hipMemcpyHostToDevice, ctx->upStream[t]);
            if(e == hipSuccess) e = hipEventRecord(ctx->upEv[t][b], ctx->upStream[t]);
        }
        const hipError_t e2 = hipStreamSynchronize(ctx->upStream[t]);
        errs[t] = e != hipSuccess ? e : e2;
    };
    {
        const int rc = onTransferThreads(ctx, T, work);
        if(rc != CARE_OK) return rc;
    }
    for(int t = 0; t < T; t++){
        if(errs[t] != hipSuccess) return ctx->fail(CARE_EHIP, "upload failed: %s", hipGetErrorString(errs[t]));
    }
    return CARE_OK;
}

// Device-to-host copy into a (pageable) host buffer, complete when it returns. Large
// copies are cut into one slice per transfer worker; a worker moves its slice through its
// two pinned chunk buffers on its own stream, copying one chunk out while the next is in
// flight. Work enqueued on `st` before the call (which may write the source) finishes first.
static int downloadD2H(care_ctx* ctx, void* dst, const void* src, size_t bytes, hipStream_t st){
    if(bytes < kUpMin){
        HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        return CARE_OK;
    }
    HIPCHECK(hipStreamSynchronize(st));
    int T = 0;
    {
        const int rc = transferWorkers(ctx, &T);
        if(rc != CARE_OK) return rc;
    }
    const size_t slice = (bytes + T - 1) / T;
    std::vector<hipError_t> errs(T, hipSuccess);
    auto work = [&](int t){
        hipError_t e = hipSetDevice(ctx->device);
        const size_t beg = std::min(bytes, size_t(t) * slice), end = std::min(bytes, beg + slice);
        int pb = -1;                                  // buffer of the chunk in flight
        size_t poff = 0, plen = 0;
        int k = 0;
        for(size_t off = beg; e == hipSuccess && off < end; off += kUpChunk, k++){
            const int b = k & 1;
            const size_t n = std::min(kUpChunk, end - off);
            char* buf = static_cast<char*>(ctx->upPinned) + (size_t(t) * 2 + size_t(b)) * kUpChunk;
            e = hipMemcpyAsync(buf, static_cast<const char*>(src) + off, n, hipMemcpyDeviceToHost, ctx->upStream[t]);
            if(e == hipSuccess) e = hipEventRecord(ctx->upEv[t][b], ctx->upStream[t]);
            if(e == hipSuccess && pb >= 0){
                e = hipEventSynchronize(ctx->upEv[t][pb]);
                if(e == hipSuccess)
                    std::memcpy(static_cast<char*>(dst) + poff,
                                static_cast<char*>(ctx->upPinned) + (size_t(t) * 2 + size_t(pb)) * kUpChunk, plen);
            }
            pb = b; poff = off; plen = n;
        }
        if(e == hipSuccess && pb >= 0){
            e = hipEventSynchronize(ctx->upEv[t][pb]);
            if(e == hipSuccess)
                std::memcpy(static_cast<char*>(dst) + poff,
                            static_cast<char*>(ctx->upPinned) + (size_t(t) * 2 + size_t(pb)) * kUpChunk, plen);
        }
        const hipError_t e2 = hipStreamSynchronize(ctx->upStream[t]);
        errs[t] = e != hipSuccess ? e : e2;
    };
    {
        const int rc = onTransferThreads(ctx, T, work);
        if(rc != CARE_OK) return rc;
    }
    for(int t = 0; t < T; t++){
        if(errs[t] != hipSuccess) return ctx->fail(CARE_EHIP, "download failed: %s", hipGetErrorString(errs[t]));
    }
    return CARE_OK;
}

// Linear host-to-device copy, complete on the device when it returns: small ones as one
// copy on `st`, large ones through the upload workers.
static int uploadH2D(care_ctx* ctx, void* dst, const void* src, size_t bytes, hipStream_t st){
    if(bytes < kUpMin){
        HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
        HIPCHECK(hipStreamSynchronize(st));
        return CARE_OK;
    }
    return uploadRowsH2D(ctx, dst, bytes, src, bytes, 1, st);
}

int configureReads(care_ctx* ctx, uint64_t num_reads, int maxLen, size_t pitch_ints, int minLen){
    if(num_reads == 0) return ctx->fail(CARE_EINVAL, "empty read storage");
    if(num_reads >= 0xFFFFFFFFull) return ctx->fail(CARE_EINVAL, "read ids are 32-bit");
    if(ctx->o.paired && (num_reads & 1)) return ctx->fail(CARE_EINVAL, "paired input needs an even number of reads");
    if(maxLen > 32 * kMaxHiLoWords) return ctx->fail(CARE_EINVAL, "reads longer than %d bases are not supported", 32 * kMaxHiLoWords);
    if(pitch_ints * 16 < size_t(maxLen)) return ctx->fail(CARE_EINVAL, "pitch_ints too small for the longest read");
    ctx->n = num_reads;
    ctx->maxLen = maxLen;
    ctx->varLen = minLen != maxLen;
    ctx->pitch = pitch_ints;
    ctx->W = std::max(1, (maxLen + 31) / 32);
    ctx->k = ctx->o.kmerlength ? ctx->o.kmerlength : (maxLen < 160 ? 20 : 32);
    ctx->M = ctx->o.num_hash_maps;
    ctx->thr = resultsPerMapThreshold(ctx->o.estimated_coverage);
    if(ctx->M * ctx->thr >= kMaxSortEntries) return ctx->fail(CARE_EINVAL, "maps*threshold = %d must be below %d", ctx->M * ctx->thr, kMaxSortEntries);
    const int minoverlap = std::max(ctx->o.min_overlap, int(float(maxLen) * ctx->o.min_overlap_ratio));
    ctx->maxCols = maxLen + 2 * std::max(0, maxLen - std::min(minoverlap, maxLen)) + 2;
    ctx->maxCols = std::max(ctx->maxCols, 3 * maxLen + 2);
    {
        const bool q = ctx->o.use_quality_scores != 0;
        if(msaLdsBytes(ctx->maxCols, maxLen, ctx->W, msaTile(ctx->maxCols, maxLen, ctx->W, q), q) > 150 * 1024)
            return ctx->fail(CARE_EINVAL, "reads too long for the LDS-resident MSA");
    }
    ctx->dbgStride = int((4 + size_t(ctx->maxCols) * (16 + 16 + 4 + 4 + 1) + 15) / 16 * 16);
    ctx->hasReads = false;
    ctx->hasTables = false;
    ctx->hasFastq = false;
    return CARE_OK;
}

// Quality-weight table and HiLo planes of the resident reads.
int finishReads(care_ctx* ctx, hipEvent_t done = nullptr){
    hipStream_t st = ctx->stream;
    float qw[256];
    for(int i = 0; i < 256; i++) qw[i] = qualityWeight(i);
    HIPCHECK(ctx->qw.ensure(sizeof(qw)));
    HIPCHECK(hipMemcpyAsync(ctx->qw.p, qw, sizeof(qw), hipMemcpyHostToDevice, st));
    const uint64_t num_reads = ctx->n;
    HIPCHECK(ctx->hilo.ensure(num_reads * 4 * ctx->W * 4 + 64));
    HIPCHECK(ctx->rcenc.ensure(num_reads * 2 * ctx->W * 4 + 64));
    const uint64_t hiloThreads = num_reads * uint64_t(ctx->W);
    k_hilo<<<gridFor(hiloThreads, 256), dim3(256), 0, st>>>(ctx->enc.as<uint32_t>(), ctx->pitch, ctx->len.as<int>(),
                                                                  num_reads, ctx->W, ctx->hilo.as<uint32_t>(),
                                                                  ctx->rcenc.as<uint32_t>());
    HIPCHECK(hipGetLastError());
    if(done) HIPCHECK(hipEventRecord(done, st));
    HIPCHECK(hipStreamSynchronize(st));
    ctx->hasReads = true;
    ctx->hasTables = false;
    return CARE_OK;
}

static int care_hip_encode_reads_body(const uint8_t* seqs, size_t seq_pitch, const int32_t* lengths, uint64_t num_reads,
                          uint32_t* enc2bit, size_t pitch_ints, uint8_t* ambiguous,
                          const uint8_t* quals_in, size_t quals_in_pitch, uint8_t* quals_out,
                          size_t quals_out_pitch, int threads){
    if(num_reads == 0) return CARE_OK;
    if(!seqs || !lengths || !enc2bit || !ambiguous || pitch_ints == 0) return CARE_EINVAL;
    if((quals_in == nullptr) != (quals_out == nullptr)) return CARE_EINVAL;
    for(uint64_t i = 0; i < num_reads; i++){
        const int L = lengths[i];
        if(L < 0 || size_t(L) > 16 * pitch_ints || size_t(L) > seq_pitch) return CARE_EINVAL;
        if(quals_in && (size_t(L) > quals_in_pitch || size_t(L) > quals_out_pitch)) return CARE_EINVAL;
    }
    constexpr uint64_t kBatch = 65536;     // parser batch: the replacement turn restarts here
    const uint64_t batches = (num_reads + kBatch - 1) / kBatch;
    uint8_t code[256];
    std::memset(code, 4, sizeof(code));
    code['A'] = code['a'] = 0; code['C'] = code['c'] = 1; code['G'] = code['g'] = 2; code['T'] = code['t'] = 3;
    auto encodeBatch = [&](uint64_t b){
        uint32_t turn = 0;
        const uint64_t r1 = std::min(num_reads, (b + 1) * kBatch);
        for(uint64_t r = b * kBatch; r < r1; r++){
            const uint8_t* s = seqs + r * seq_pitch;
            uint32_t* row = enc2bit + r * pitch_ints;
            const int L = lengths[r];
            uint8_t amb = 0;
            for(size_t w = 0; w < pitch_ints; w++){
                uint32_t word = 0;
                const int p0 = int(w) * 16;
                for(int j = 0; j < 16; j++){
                    const int pos = p0 + j;
                    uint32_t c = 0;
                    if(pos < L){
                        c = code[s[pos]];
                        if(c == 4){ c = turn & 3; turn++; amb = 1; }
                    }
                    word |= c << (30 - 2 * j);
                }
                row[w] = word;
            }
            ambiguous[r] = amb;
            if(quals_in){
                uint8_t* qo = quals_out + r * quals_out_pitch;
                std::memcpy(qo, quals_in + r * quals_in_pitch, size_t(L));
                std::memset(qo + L, 0, quals_out_pitch - size_t(L));
            }
        }
    };
    int T = threads > 0 ? threads : int(std::max(1u, std::thread::hardware_concurrency()));
    T = int(std::min<uint64_t>(uint64_t(T), batches));
    if(T <= 1){
        for(uint64_t b = 0; b < batches; b++) encodeBatch(b);
        return CARE_OK;
    }
    std::vector<std::thread> pool;
    try{
        for(int t = 0; t < T; t++)
            pool.emplace_back([&, t](){ for(uint64_t b = uint64_t(t); b < batches; b += uint64_t(T)) encodeBatch(b); });
    }catch(...){
        for(auto& th : pool) th.join();
        return CARE_ENOMEM;
    }
    for(auto& th : pool) th.join();
    return CARE_OK;
}

static int care_hip_set_reads_body(care_ctx* ctx, const uint32_t* enc2bit, size_t pitch_ints, const int32_t* lengths,
                       const uint8_t* quals, size_t qual_pitch, uint64_t num_reads, const uint8_t* ambiguous){
    if(!ctx) return CARE_EINVAL;
    if(!enc2bit || !lengths || !ambiguous || num_reads == 0) return ctx->fail(CARE_EINVAL, "empty or null read storage");
    int maxLen = 0, minLen = INT_MAX;
    for(uint64_t i = 0; i < num_reads; i++){
        if(lengths[i] < 0) return ctx->fail(CARE_EINVAL, "negative length at read %llu", (unsigned long long)i);
        maxLen = std::max(maxLen, int(lengths[i]));
        minLen = std::min(minLen, int(lengths[i]));
    }
    if(ctx->o.use_quality_scores){
        if(!quals) return ctx->fail(CARE_EINVAL, "use_quality_scores set but no qualities given");
        if(qual_pitch < size_t(maxLen)) return ctx->fail(CARE_EINVAL, "qual_pitch smaller than the longest read");
    }
    HIPCHECK(hipSetDevice(ctx->device));
    {
        const int rc = configureReads(ctx, num_reads, maxLen, pitch_ints, minLen);
        if(rc != CARE_OK) return rc;
    }
    hipStream_t st = ctx->stream;
    HIPCHECK(ctx->enc.ensure(num_reads * pitch_ints * 4));
    {
        const int rc = uploadH2D(ctx, ctx->enc.p, enc2bit, num_reads * pitch_ints * 4, st);
        if(rc != CARE_OK) return rc;
    }
    HIPCHECK(ctx->len.ensure(num_reads * 4));
    HIPCHECK(hipMemcpyAsync(ctx->len.p, lengths, num_reads * 4, hipMemcpyHostToDevice, st));
    HIPCHECK(ctx->ambig.ensure(num_reads));
    HIPCHECK(hipMemcpyAsync(ctx->ambig.p, ambiguous, num_reads, hipMemcpyHostToDevice, st));
    if(ctx->o.use_quality_scores){
        // the reference reverses quality rows of its storage's pitch, the longest read
        // (src/correct_cpu.cpp:141, include/corrector.hpp:1227-1239), whatever the caller's pitch
        ctx->qpitch = size_t(maxLen);
        ctx->qdpitch = (qual_pitch + 15) / 16 * 16;
        HIPCHECK(ctx->qual.ensure(num_reads * ctx->qdpitch));
        if(ctx->qdpitch == qual_pitch){
            const int rc = uploadH2D(ctx, ctx->qual.p, quals, num_reads * qual_pitch, st);
            if(rc != CARE_OK) return rc;
        }else if(num_reads * qual_pitch >= kUpMin && ctx->qdpitch <= kUpChunk){
            // rows repacked to the 16-byte pitch by the upload workers
            const int rc = uploadRowsH2D(ctx, ctx->qual.p, ctx->qdpitch, quals, qual_pitch, num_reads, st);
            if(rc != CARE_OK) return rc;
        }else{
            // one linear copy, then rows repacked to the 16-byte pitch on the device
            struct Staging { DBuf b; ~Staging(){ b.release(); } } hold;
            DBuf& staging = hold.b;
            HIPCHECK(staging.ensure(num_reads * qual_pitch));
            const int rc = uploadH2D(ctx, staging.p, quals, num_reads * qual_pitch, st);
            if(rc != CARE_OK) return rc;
            const uint64_t total = num_reads * ctx->qdpitch;
            k_repack_rows<<<gridFor(total, 256), 256, 0, st>>>(staging.as<uint8_t>(), qual_pitch,
                                                                              ctx->qual.as<uint8_t>(), ctx->qdpitch, total);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipStreamSynchronize(st));
        }
    }
    return finishReads(ctx);
}

// ---------------------------------------------------------------- FASTQ on the device
//
// Read-storage construction from FASTQ text (constructChunkedReadStorageFromFiles,
// include/chunkedreadstorageconstruction.hpp:29-480): 4-line records; the sequence line
// upper-cased, every other character than ACGTacgt replaced by A, C, G, T in turn with the
// counter restarting at every parser batch of 65,536 reads (:67-92, :106, :260), and the
// read flagged ambiguous; paired files interleaved (R1, R2, R1, ...,
// include/readlibraryio.hpp:337-). Line ends are '\n'; a '\r' before it is dropped when
// what it ends (header after '@', sequence, quality) has more than one character, as the
// reference's reader does (include/kseqpp/kseqpp.hpp:286-288). Newline positions come from
// a stream compaction of the text; then one thread per
// record scans its lines, a scan of the per-read replacement counts gives every read its
// counter start, and one thread per read encodes it. Records are 4 lines (sequence and
// quality on one line each); wrapped multi-line records fail the separator check and are
// rejected, as is FASTA.

constexpr uint64_t kParserBatch = 65536;

// Newline positions by a two-pass compaction: every thread takes 16 bytes of the text;
// pass 1 counts the '\n' bytes per block, a scan of the block counts gives each block its
// first output slot, pass 2 recounts and writes the positions in text order. The text
// buffer is padded to a multiple of 16 bytes, so the 16-byte loads stay inside it.
constexpr int kNlBlock = 256;
constexpr int kNlPerBlock = 16 * kNlBlock;

__device__ __forceinline__ uint32_t newlineMask(const char* __restrict__ t, uint64_t base, uint64_t bytes){
    const uint4 v = *reinterpret_cast<const uint4*>(t + base);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    uint32_t mask = 0;
    #pragma unroll
    for(int k = 0; k < 16; k++){
        const bool nlb = ((w[k >> 2] >> (8 * (k & 3))) & 0xffu) == uint32_t('\n');
        if(nlb && base + k < bytes) mask |= 1u << k;
    }
    return mask;
}

// Block-wide exclusive prefix of one value per thread (kNlBlock threads); returns the
// prefix and sets *total.
__device__ __forceinline__ int blockExclusiveSum(int v, int* sWave, int* total){
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int incl = waveInclusiveSum(v);
    if(lane == 63) sWave[wave] = incl;
    __syncthreads();
    int before = 0, all = 0;
    #pragma unroll
    for(int w = 0; w < kNlBlock / 64; w++){
        const int x = sWave[w];
        before += w < wave ? x : 0;
        all += x;
    }
    *total = all;
    return before + incl - v;
}

__global__ void __launch_bounds__(kNlBlock) k_newline_count(const char* __restrict__ t, uint64_t bytes,
                                                            uint32_t* __restrict__ blockCnt){
    __shared__ int sWave[kNlBlock / 64];
    const uint64_t base = uint64_t(blockIdx.x) * kNlPerBlock + 16ull * threadIdx.x;
    const int c = base < bytes ? __popc(newlineMask(t, base, bytes)) : 0;
    int total;
    blockExclusiveSum(c, sWave, &total);
    if(threadIdx.x == 0) blockCnt[blockIdx.x] = uint32_t(total);
}

__global__ void __launch_bounds__(kNlBlock) k_newline_write(const char* __restrict__ t, uint64_t bytes,
                                                            const uint64_t* __restrict__ blockOff,
                                                            uint64_t* __restrict__ nl){
    __shared__ int sWave[kNlBlock / 64];
    const uint64_t base = uint64_t(blockIdx.x) * kNlPerBlock + 16ull * threadIdx.x;
    uint32_t mask = base < bytes ? newlineMask(t, base, bytes) : 0u;
    int total;
    uint64_t o = blockOff[blockIdx.x] + uint64_t(blockExclusiveSum(__popc(mask), sWave, &total));
    while(mask){
        const int k = __ffs(mask) - 1;
        nl[o++] = base + uint64_t(k);
        mask &= mask - 1;
    }
}

__device__ __forceinline__ int fastqCode(char c){
    switch(c){
        case 'A': case 'a': return 0;
        case 'C': case 'c': return 1;
        case 'G': case 'g': return 2;
        case 'T': case 't': return 3;
        default: return -1;
    }
}

// Inclusive sum over the 16 lanes of a DPP row.
__device__ __forceinline__ int rowInclusiveSum(int v){
    int t;
    t = __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false); v += t;
    t = __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false); v += t;
    t = __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false); v += t;
    t = __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false); v += t;
    return v;
}

__device__ __forceinline__ int rowBroadcastLast(int v){
    return __shfl(v, (threadIdx.x & 63 & ~15) | 15, 64);
}

constexpr int kFqBlock = 256;        // 16 records (or reads) per block, one 16-lane row each

// Record r of one file: lines 4r..4r+3; line i spans [i ? nl[i-1]+1 : 0, i < numNl ? nl[i] : bytes).
// One 16-lane row per record: every lane finds the lines, the lanes count the sequence's
// non-ACGT bytes over 16-byte slices, and lane 0 writes the record.
__global__ void __launch_bounds__(kFqBlock) k_fastq_records(
        const char* __restrict__ t, uint64_t bytes, const uint64_t* __restrict__ nl, uint64_t numNl, uint64_t records,
        int stride, int slot, uint64_t* __restrict__ hdrStart, uint32_t* __restrict__ hdrLen,
        uint64_t* __restrict__ seqStart, uint64_t* __restrict__ qualStart, int* __restrict__ len,
        uint64_t* __restrict__ ncount, int* __restrict__ err){
    const int sub = threadIdx.x & 15;
    // rows stride over the records; the 16 lanes of a row always hold the same record
    for(uint64_t rec = (uint64_t(blockIdx.x) * kFqBlock + threadIdx.x) >> 4; rec < records;
        rec += (uint64_t(gridDim.x) * kFqBlock) >> 4){
    auto lineStart = [&](uint64_t i){ return i ? nl[i - 1] + 1 : uint64_t(0); };
    // end of the content starting at s on line i
    auto contentEnd = [&](uint64_t i, uint64_t s){
        uint64_t e = i < numNl ? nl[i] : bytes;
        if(e > s + 1 && t[e - 1] == '\r') e--;
        return e;
    };
    const uint64_t l0 = 4 * rec;
    const uint64_t h = lineStart(l0);
    const uint64_t ss = lineStart(l0 + 1), se = contentEnd(l0 + 1, ss);
    const uint64_t ps = lineStart(l0 + 2), qs = lineStart(l0 + 3), qe = contentEnd(l0 + 3, qs);
    int nc = 0;
    for(uint64_t p = ss + 16 * sub; p < se; p += 256){
        const int m = se - p < 16 ? int(se - p) : 16;
        for(int k = 0; k < m; k++) nc += fastqCode(t[p + k]) < 0;
    }
    nc = rowBroadcastLast(rowInclusiveSum(nc));
    if(sub != 0) continue;
    const uint64_t r = uint64_t(stride) * rec + uint64_t(slot);
    const bool ok = h < bytes && t[h] == '@' && ps < bytes && t[ps] == '+' && qe - qs == se - ss;
    const uint64_t he = ok ? contentEnd(l0, h + 1) : h;
    hdrStart[r] = h + 1;
    hdrLen[r] = uint32_t(ok ? he - (h + 1) : 0);
    seqStart[r] = ss;
    qualStart[r] = qs;
    len[r] = int(se - ss);
    ncount[r] = uint64_t(nc);
    if(!ok) atomicOr(err, 1);
    }
}

// One 16-lane row per read: each lane encodes a run of the read's 2-bit words (MSB first,
// 16 bases per word, zero-padded to the pitch) with its replacement counter taken from the
// read's counter start plus the non-ACGT bytes of the lanes before it; then the lanes copy
// the qualities in 16-byte chunks, zero-padded to the device pitch.
__global__ void __launch_bounds__(kFqBlock) k_fastq_encode(
        const char* __restrict__ t1, const char* __restrict__ t2, int paired, const uint64_t* __restrict__ seqStart,
        const uint64_t* __restrict__ qualStart, const int* __restrict__ len, const uint64_t* __restrict__ prefix,
        uint64_t n, size_t pitch, uint32_t* __restrict__ enc, uint8_t* __restrict__ ambig,
        uint8_t* __restrict__ qual, size_t qdpitch){
    const int sub = threadIdx.x & 15;
    for(uint64_t r = (uint64_t(blockIdx.x) * kFqBlock + threadIdx.x) >> 4; r < n; r += (uint64_t(gridDim.x) * kFqBlock) >> 4){
    const char* t = (paired && (r & 1)) ? t2 : t1;
    const int L = len[r];
    const char* sq = t + seqStart[r];
    const int wpl = int((pitch + 15) / 16);
    const int w0 = sub * wpl, w1 = min(int(pitch), w0 + wpl);
    int cnt = 0;
    for(int pos = w0 * 16; pos < min(L, w1 * 16); pos++) cnt += fastqCode(sq[pos]) < 0;
    const int incl = rowInclusiveSum(cnt);
    const int any = rowBroadcastLast(incl);
    const uint64_t b0 = r & ~(kParserBatch - 1);
    int counter = int((prefix[r] - prefix[b0] + uint64_t(incl - cnt)) & 3);
    uint32_t* row = enc + pitch * r;
    for(int w = w0; w < w1; w++){
        uint32_t word = 0;
        for(int j = 0; j < 16; j++){
            const int pos = w * 16 + j;
            if(pos >= L) break;
            int code = fastqCode(sq[pos]);
            if(code < 0){ code = counter; counter = (counter + 1) & 3; }
            word |= uint32_t(code) << (30 - 2 * j);
        }
        row[w] = word;
    }
    if(sub == 0) ambig[r] = any > 0;
    if(qual){
        const char* qs = t + qualStart[r];
        uint8_t* qrow = qual + qdpitch * r;
        for(int c = sub; c < int(qdpitch / 16); c += 16){
            uint32_t w[4] = {0, 0, 0, 0};
            #pragma unroll
            for(int k = 0; k < 16; k++){
                const int idx = 16 * c + k;
                if(idx < L) w[k >> 2] |= uint32_t(uint8_t(qs[idx])) << (8 * (k & 3));
            }
            *reinterpret_cast<uint4*>(qrow + 16 * c) = make_uint4(w[0], w[1], w[2], w[3]);
        }
    }
    }
}

namespace {
struct ParsedFile {
    DBuf* text = nullptr;          // the context's resident copy of the text
    DBuf blockCnt, blockOff, nl;
    uint64_t bytes = 0, numNl = 0, records = 0;
    void release(){ blockCnt.release(); blockOff.release(); nl.release(); }
};
}  // namespace

static int uploadAndSplitLines(care_ctx* ctx, const char* data, size_t raw, int slot, ParsedFile& f){
    hipStream_t st = ctx->stream;
    f.text = &ctx->fqText[slot];
    // blank lines at the end of the text are not records (the reference's reader skips them)
    size_t bytes = raw;
    while(bytes && (data[bytes - 1] == '\n' || data[bytes - 1] == '\r')) bytes--;
    f.bytes = bytes;
    if(bytes == 0){ f.numNl = 0; f.records = 0; return CARE_OK; }
    // one work item per 16 bytes in one launch: the grid's work items must fit 32 bits
    if(bytes / 16 >= (uint64_t(1) << 32) - kNlBlock)
        return ctx->fail(CARE_EINVAL, "a FASTQ text of %zu bytes: at most 64 GiB per text", bytes);
    const uint32_t nb = uint32_t((bytes + kNlPerBlock - 1) / kNlPerBlock);
    HIPCHECK(f.blockCnt.ensure((size_t(nb) + 1) * 4));
    HIPCHECK(f.blockOff.ensure((size_t(nb) + 1) * 8));
    k_newline_count<<<nb, kNlBlock, 0, st>>>(f.text->as<char>(), bytes, f.blockCnt.as<uint32_t>());
    HIPCHECK(hipGetLastError());
    k_fill_u32<<<1, 64, 0, st>>>(f.blockCnt.as<uint32_t>() + nb, 1, 0u);
    size_t tb = 0;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, f.blockCnt.as<uint32_t>(), f.blockOff.as<uint64_t>(),
                                              int(nb + 1), st));
    HIPCHECK(ctx->scanTmp.ensure(std::max<size_t>(tb, 256)));
    tb = ctx->scanTmp.bytes;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(ctx->scanTmp.p, tb, f.blockCnt.as<uint32_t>(), f.blockOff.as<uint64_t>(),
                                              int(nb + 1), st));
    uint64_t numNl = 0;
    HIPCHECK(hipMemcpyAsync(&numNl, f.blockOff.as<uint64_t>() + nb, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    f.numNl = numNl;
    const uint64_t lines = numNl + 1;
    if(lines % 4) return ctx->fail(CARE_EINVAL, "FASTQ text of %llu lines is not 4-line records",
                                   (unsigned long long)lines);
    f.records = lines / 4;
    HIPCHECK(f.nl.ensure(std::max<size_t>(size_t(numNl) * 8, 8)));
    k_newline_write<<<nb, kNlBlock, 0, st>>>(f.text->as<char>(), bytes, f.blockOff.as<uint64_t>(), f.nl.as<uint64_t>());
    HIPCHECK(hipGetLastError());
    return CARE_OK;
}

static int care_hip_set_reads_fastq_body(care_ctx* ctx, const char* fastq1, size_t bytes1, const char* fastq2, size_t bytes2){
    if(!ctx || !fastq1) return CARE_EINVAL;
    const bool paired = ctx->o.paired != 0;
    if(paired != (fastq2 != nullptr)) return ctx->fail(CARE_EINVAL, paired ? "paired input needs two FASTQ texts"
                                                                            : "single-end input takes one FASTQ text");
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    ctx->hasFastq = false;
    for(int slot = 0; slot < (paired ? 2 : 1); slot++){
        const size_t raw = slot ? bytes2 : bytes1;
        // newline blocks of 4 KiB are counted in 32 bits and read indices are 32-bit
        if(raw / kNlPerBlock >= size_t(INT_MAX) - 1) return ctx->fail(CARE_EINVAL, "FASTQ text too large for one call");
        ctx->fqRaw[slot] = raw;
        HIPCHECK(ctx->fqText[slot].ensure((raw + 15) / 16 * 16 + 16));
        if(raw){
            const int rc = uploadH2D(ctx, ctx->fqText[slot].p, slot ? fastq2 : fastq1, raw, st);
            if(rc != CARE_OK) return rc;
        }
    }
    HIPCHECK(hipEventRecord(ctx->ev[6], st));
    ParsedFile f1, f2;
    struct Rel { ParsedFile& a; ParsedFile& b; ~Rel(){ a.release(); b.release(); } } rel{f1, f2};
    int rc = uploadAndSplitLines(ctx, fastq1, bytes1, 0, f1);
    if(rc != CARE_OK) return rc;
    if(paired){
        rc = uploadAndSplitLines(ctx, fastq2, bytes2, 1, f2);
        if(rc != CARE_OK) return rc;
        if(f1.records != f2.records) return ctx->fail(CARE_EINVAL, "paired FASTQ texts hold %llu and %llu records",
                                                      (unsigned long long)f1.records, (unsigned long long)f2.records);
    }
    const uint64_t n = paired ? 2 * f1.records : f1.records;
    if(n == 0) return ctx->fail(CARE_EINVAL, "no FASTQ records");
    struct Tmp { DBuf lens, ncount, prefix, err; ~Tmp(){ lens.release(); ncount.release(); prefix.release();
        err.release(); } } tmp;
    HIPCHECK(ctx->fqHdr.ensure(n * 8)); HIPCHECK(ctx->fqHdrLen.ensure(n * 4));
    HIPCHECK(ctx->fqSeq.ensure(n * 8)); HIPCHECK(ctx->fqQual.ensure(n * 8));
    HIPCHECK(tmp.lens.ensure(n * 4)); HIPCHECK(tmp.ncount.ensure((n + 1) * 8)); HIPCHECK(tmp.prefix.ensure((n + 1) * 8));
    HIPCHECK(tmp.err.ensure(4));
    k_zero_u32<<<1, 64, 0, st>>>(tmp.err.as<uint32_t>(), 1);
    const int stride = paired ? 2 : 1;
    for(int fi = 0; fi < (paired ? 2 : 1); fi++){
        ParsedFile& f = fi ? f2 : f1;
        k_fastq_records<<<gridFor(f.records * 16, kFqBlock), kFqBlock, 0, st>>>(
            f.text->as<char>(), f.bytes, f.nl.as<uint64_t>(), f.numNl, f.records, stride, fi,
            ctx->fqHdr.as<uint64_t>(), ctx->fqHdrLen.as<uint32_t>(), ctx->fqSeq.as<uint64_t>(), ctx->fqQual.as<uint64_t>(),
            tmp.lens.as<int>(), tmp.ncount.as<uint64_t>(), tmp.err.as<int>());
        HIPCHECK(hipGetLastError());
    }
    int err = 0;
    std::vector<int> lens(n);
    HIPCHECK(hipMemcpyAsync(&err, tmp.err.p, 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(lens.data(), tmp.lens.p, n * 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if(err) return ctx->fail(CARE_EINVAL, "malformed FASTQ record (header '@', separator '+', quality length)");
    const int maxLen = *std::max_element(lens.begin(), lens.end());
    const size_t pitch = size_t(std::max(1, (maxLen + 15) / 16));
    rc = configureReads(ctx, n, maxLen, pitch, *std::min_element(lens.begin(), lens.end()));
    if(rc != CARE_OK) return rc;
    size_t tb = 0;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, tmp.ncount.as<uint64_t>(), tmp.prefix.as<uint64_t>(), int(n), st));
    HIPCHECK(ctx->scanTmp.ensure(std::max<size_t>(tb, 256)));
    tb = ctx->scanTmp.bytes;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(ctx->scanTmp.p, tb, tmp.ncount.as<uint64_t>(), tmp.prefix.as<uint64_t>(), int(n), st));
    HIPCHECK(ctx->enc.ensure(n * pitch * 4));
    HIPCHECK(ctx->len.ensure(n * 4));
    HIPCHECK(hipMemcpyAsync(ctx->len.p, tmp.lens.p, n * 4, hipMemcpyDeviceToDevice, st));
    HIPCHECK(ctx->ambig.ensure(n));
    uint8_t* qual = nullptr;
    if(ctx->o.use_quality_scores){
        ctx->qpitch = size_t(maxLen);
        ctx->qdpitch = (size_t(maxLen) + 15) / 16 * 16;
        HIPCHECK(ctx->qual.ensure(n * ctx->qdpitch));
        qual = ctx->qual.as<uint8_t>();
    }
    k_fastq_encode<<<gridFor(n * 16, kFqBlock), kFqBlock, 0, st>>>(
        f1.text->as<char>(), paired ? f2.text->as<char>() : nullptr, paired, ctx->fqSeq.as<uint64_t>(),
        ctx->fqQual.as<uint64_t>(), tmp.lens.as<int>(), tmp.prefix.as<uint64_t>(), n, pitch, ctx->enc.as<uint32_t>(),
        ctx->ambig.as<uint8_t>(), qual, ctx->qdpitch);
    HIPCHECK(hipGetLastError());
    rc = finishReads(ctx, ctx->ev[7]);
    if(rc != CARE_OK) return rc;
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, ctx->ev[6], ctx->ev[7]));
    ctx->fastqMs = ms;
    ctx->hasFastq = true;
    return CARE_OK;
}

// ---------------------------------------------------------------- output merge
//
// The corrected FASTQ stream of one input file, as the reference hands it to QC
// (constructOutputFileFromCorrectionResultsOutToQueue, src/correctionresultoutput.cpp:1589-1632):
// records ordered by read id (sortSerializedResultsByReadIdAscending,
// include/sortserializedresults.hpp:17-78), the records of each read combined with its
// original FASTQ text (combineMultipleCorrectionResults1_rawtcs2, :1256-1550), and every
// read written as "@" header [" care:q=N"] "\n" sequence "\n+\n" quality "\n" (:726-760).
// On the device: a stable radix sort of (read id, record index), per-read runs of the
// sorted records, one thread per read for the combine decision and the output size, a
// scan of the sizes, and one wavefront per read writing its text.

__device__ __forceinline__ uint32_t ldU32(const uint8_t* p){
    return uint32_t(p[0]) | uint32_t(p[1]) << 8 | uint32_t(p[2]) << 16 | uint32_t(p[3]) << 24;
}

// One serialized record (EncodedTempCorrectedSequence, include/correctedsequence.hpp:151-387):
// read id, flags (hq << 31 | useEdits << 30 | candidate << 29 | body bytes), then either
// (int count, int positions[count], char bases[count]) or (int length, char sequence[length]),
// then an int shift for candidate records.
struct RecView {
    const uint8_t* body;
    uint32_t flags;
    __device__ bool hq() const { return flags >> 31; }
    __device__ bool useEdits() const { return (flags >> 30) & 1u; }
    __device__ bool isCandidate() const { return (flags >> 29) & 1u; }
    __device__ int count() const { return int(ldU32(body)); }
    __device__ int editPos(int i) const { return int(ldU32(body + 4 + 4 * i)); }
    __device__ uint8_t editBase(int ne, int i) const { return body[4 + 4 * ne + i]; }
    __device__ uint8_t seqByte(int i) const { return body[4 + i]; }
};

__device__ __forceinline__ RecView recAt(const uint8_t* recs, const uint64_t* off, uint32_t idx){
    const uint8_t* p = recs + off[idx];
    return RecView{p + 8, ldU32(p + 4)};
}

// isEqualSequenceOneHasEdits (src/correctionresultoutput.cpp:1327-1356), step for step
__device__ bool recEqualOneHasEdits(const uint8_t* orig, int L, const RecView& e, const RecView& q){
    const int ne = e.count();
    int k = 0, pos = 0;
    while(pos != L && k != ne){
        if(pos != e.editPos(k)){
            if(orig[pos] != q.seqByte(pos)) return false;
        }else{
            if(e.editBase(ne, k) != q.seqByte(pos)) return false;
            ++k;
        }
        ++pos;
    }
    for(; pos != L; ++pos) if(orig[pos] != q.seqByte(pos)) return false;
    return true;
}

// isEqualSequence (src/correctionresultoutput.cpp:1358-1384)
__device__ bool recEqual(const uint8_t* orig, int L, const RecView& a, const RecView& b){
    const int la = a.useEdits() ? L : a.count();
    const int lb = b.useEdits() ? L : b.count();
    if(la != lb) return false;
    if(a.useEdits() && b.useEdits()){
        const int na = a.count(), nb = b.count();
        if(na != nb) return false;
        for(int i = 0; i < na; i++)
            if(a.editPos(i) != b.editPos(i) || a.editBase(na, i) != b.editBase(nb, i)) return false;
        return true;
    }
    if(!a.useEdits() && !b.useEdits()){
        for(int i = 0; i < la; i++) if(a.seqByte(i) != b.seqByte(i)) return false;
        return true;
    }
    const RecView& e = a.useEdits() ? a : b;
    const RecView& q = a.useEdits() ? b : a;
    if(e.count() == 0){
        for(int i = 0; i < L; i++) if(orig[i] != q.seqByte(i)) return false;
        return true;
    }
    return recEqualOneHasEdits(orig, L, e, q);
}

struct MergeParams {
    const uint8_t* recs; uint64_t recBytes; const uint64_t* off; uint32_t m;
    const uint32_t* keysSorted; const uint32_t* idxSorted;
    const uint32_t* runBeg; const uint32_t* runEnd;
    const char* text; const uint64_t* hdr; const uint32_t* hdrLen; const uint64_t* seq; const uint64_t* qual;
    const int* len;
    uint64_t n, count; int stride, file, labels;
    int2* plan; uint64_t* sizes; const uint64_t* outOff; char* out;
};

// Record headers: read id keys, identity values, and a flag for records that do not fit
// the buffer or name a read past the storage.
__global__ void k_merge_keys(const uint8_t* __restrict__ recs, uint64_t recBytes, const uint64_t* __restrict__ off,
                             uint32_t m, uint64_t n, uint32_t* __restrict__ keys, uint32_t* __restrict__ idx,
                             int* __restrict__ err){
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= m) return;
    const uint64_t o = off[i];
    uint32_t id = 0;
    bool ok = o + 8 <= recBytes;
    if(ok){
        id = ldU32(recs + o);
        const uint32_t flags = ldU32(recs + o + 4);
        const uint64_t nb = flags & ((1u << 29) - 1);
        ok = o + 8 + nb <= recBytes && nb >= 4 && uint64_t(id) < n;
        if(ok){
            const int64_t c = int32_t(ldU32(recs + o + 8));
            const int64_t need = 4 + ((flags >> 30) & 1u ? 5 * c : c) + ((flags >> 29) & 1u ? 4 : 0);
            ok = c >= 0 && need <= int64_t(nb);
        }
    }
    keys[i] = id;
    idx[i] = i;
    if(!ok) atomicOr(err, 1);
}

// [runBeg[r], runEnd[r]) = positions of read r's records in the sorted order
__global__ void k_merge_runs(const uint32_t* __restrict__ keys, uint32_t m, uint32_t* __restrict__ runBeg,
                             uint32_t* __restrict__ runEnd){
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if(j >= m) return;
    const uint32_t r = keys[j];
    if(j == 0 || keys[j - 1] != r) runBeg[r] = j;
    if(j == m - 1 || keys[j + 1] != r) runEnd[r] = j + 1;
}

constexpr int kLabelBytes = 9;     // " care:q=N"

// One thread per read of the file: the combine decision (record index or -1, label) and
// the size of the read's output text.
__global__ void k_merge_plan(MergeParams p){
    FOR_GRID(i, p.count){
    const uint64_t r = uint64_t(p.stride) * i + uint64_t(p.file);
    const uint32_t b = p.runBeg[r], e = p.runEnd[r];
    const int L = p.len[r];
    int choice = -1, label = 0;
    if(b < e){
        int anchor = -1;
        for(uint32_t j = b; j < e; j++){
            if(!recAt(p.recs, p.off, p.idxSorted[j]).isCandidate()){ anchor = int(j); break; }
        }
        if(anchor >= 0){
            const RecView an = recAt(p.recs, p.off, p.idxSorted[anchor]);
            if(an.hq()){
                choice = int(p.idxSorted[anchor]); label = 3;
            }else if(e - b >= 3){
                const uint8_t* orig = reinterpret_cast<const uint8_t*>(p.text + p.seq[r]);
                const RecView first = recAt(p.recs, p.off, p.idxSorted[b]);
                bool same = true;
                for(uint32_t j = b + 1; j < e && same; j++) same = recEqual(orig, L, first, recAt(p.recs, p.off, p.idxSorted[j]));
                if(same){ choice = int(p.idxSorted[b]); label = 2; }
            }else{
                choice = int(p.idxSorted[anchor]); label = 1;
            }
        }
    }
    // the reference stops visiting reads after the last record's read: no label there
    const bool labelled = p.labels && r <= p.keysSorted[p.m - 1];
    int seqLen = L;
    if(choice >= 0){
        const RecView c = recAt(p.recs, p.off, uint32_t(choice));
        if(!c.useEdits()) seqLen = c.count();
    }
    p.plan[i] = make_int2(choice, label | (labelled ? 0x100 : 0));
    p.sizes[i] = 1 + uint64_t(p.hdrLen[r]) + (labelled ? kLabelBytes : 0) + 1 + uint64_t(seqLen) + 3 + uint64_t(L) + 1;
    }
}

// One wavefront per read: lanes write consecutive bytes of the read's output text.
__global__ void __launch_bounds__(256) k_merge_write(MergeParams p){
    const int lane = threadIdx.x & 63;
    for(uint64_t i = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6; i < p.count;
        i += (uint64_t(gridDim.x) * blockDim.x) >> 6){
    const uint64_t r = uint64_t(p.stride) * i + uint64_t(p.file);
    const int2 pl = p.plan[i];
    const int hl = int(p.hdrLen[r]);
    const int lb = (pl.y & 0x100) ? kLabelBytes : 0;
    const int L = p.len[r];
    const uint8_t* hdr = reinterpret_cast<const uint8_t*>(p.text + p.hdr[r]);
    const uint8_t* orig = reinterpret_cast<const uint8_t*>(p.text + p.seq[r]);
    const uint8_t* qual = reinterpret_cast<const uint8_t*>(p.text + p.qual[r]);
    RecView c{nullptr, 0};
    int ne = 0, sl = L;
    const bool take = pl.x >= 0;
    if(take){
        c = recAt(p.recs, p.off, uint32_t(pl.x));
        if(c.useEdits()) ne = c.count(); else sl = c.count();
    }
    const bool full = take && !c.useEdits();
    char* out = p.out + p.outOff[i];
    const uint64_t total = p.outOff[i + 1] - p.outOff[i];
    const int s0 = 2 + hl + lb;                  // first sequence byte
    for(uint64_t k = lane; k < total; k += 64){
        const int kk = int(k);
        uint8_t ch;
        if(kk == 0) ch = '@';
        else if(kk <= hl) ch = hdr[kk - 1];
        else if(kk <= hl + lb){
            const int t = kk - 1 - hl;
            ch = t < kLabelBytes - 1 ? uint8_t(" care:q="[t]) : uint8_t('0' + (pl.y & 0xff));
        }else if(kk == hl + lb + 1) ch = '\n';
        else{
            const int q = kk - s0;
            if(q < sl){
                if(full) ch = c.seqByte(q);
                else{
                    ch = orig[q];
                    for(int x = 0; x < ne; x++) if(c.editPos(x) == q) ch = c.editBase(ne, x);
                }
            }else if(q < sl + 3) ch = (q == sl + 1) ? '+' : '\n';
            else if(q < sl + 3 + L) ch = qual[q - sl - 3];
            else ch = '\n';
        }
        out[k] = char(ch);
    }
    }
}

static int care_hip_write_fastq_body(care_ctx* ctx, const uint8_t* records, size_t record_bytes, int file, int quality_labels,
                         char* dst, size_t cap, size_t* written){
    if(!ctx || !written) return CARE_EINVAL;
    *written = 0;
    if(!ctx->hasFastq) return ctx->fail(CARE_ESTATE, "the output merge needs reads set by care_hip_set_reads_fastq");
    const bool paired = ctx->o.paired != 0;
    if(file < 0 || file > (paired ? 1 : 0)) return ctx->fail(CARE_EINVAL, "file %d out of range", file);
    if(record_bytes && !records) return ctx->fail(CARE_EINVAL, "null records");
    // record boundaries (the serialized stream carries its own sizes)
    std::vector<uint64_t> off;
    for(size_t p = 0; p < record_bytes;){
        if(p + 8 > record_bytes) return ctx->fail(CARE_EINVAL, "truncated record at byte %zu", p);
        uint32_t flags;
        std::memcpy(&flags, records + p + 4, 4);
        const size_t len = 8 + (flags & ((1u << 29) - 1));
        if(p + len > record_bytes) return ctx->fail(CARE_EINVAL, "truncated record at byte %zu", p);
        off.push_back(p);
        p += len;
    }
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const uint64_t m64 = off.size();
    if(m64 == 0){
        // no records: the reference hands QC nothing and QC reads the input text itself
        const size_t raw = ctx->fqRaw[file];
        *written = raw;
        if(!dst) return CARE_OK;
        if(cap < raw) return ctx->fail(CARE_EINVAL, "output needs %zu bytes, cap is %zu", raw, cap);
        return raw ? downloadD2H(ctx, dst, ctx->fqText[file].p, raw, ctx->stream) : CARE_OK;
    }
    if(m64 >= uint64_t(INT_MAX)) return ctx->fail(CARE_EINVAL, "too many records");
    const uint32_t m = uint32_t(m64);
    const uint64_t n = ctx->n;
    const uint64_t count = paired ? n / 2 : n;
    HIPCHECK(ctx->mRec.ensure(record_bytes));
    HIPCHECK(hipMemcpyAsync(ctx->mRec.p, records, record_bytes, hipMemcpyHostToDevice, st));
    HIPCHECK(ctx->mOff.ensure(m64 * 8));
    HIPCHECK(hipMemcpyAsync(ctx->mOff.p, off.data(), m64 * 8, hipMemcpyHostToDevice, st));
    HIPCHECK(ctx->mKeyIn.ensure(m64 * 4)); HIPCHECK(ctx->mKeyOut.ensure(m64 * 4));
    HIPCHECK(ctx->mIdxIn.ensure(m64 * 4)); HIPCHECK(ctx->mIdxOut.ensure(m64 * 4));
    HIPCHECK(ctx->mErr.ensure(4));
    HIPCHECK(hipEventRecord(ctx->ev[6], st));
    k_fill_u32<<<1, 64, 0, st>>>(ctx->mErr.as<uint32_t>(), 1, 0u);
    const uint32_t mb = (m + 255) / 256;
    k_merge_keys<<<mb, 256, 0, st>>>(ctx->mRec.as<uint8_t>(), record_bytes, ctx->mOff.as<uint64_t>(), m, n,
                                     ctx->mKeyIn.as<uint32_t>(), ctx->mIdxIn.as<uint32_t>(), ctx->mErr.as<int>());
    HIPCHECK(hipGetLastError());
    int endBit = 1;
    while(endBit < 32 && (uint64_t(1) << endBit) < n) endBit++;
    size_t tb = 0;
    HIPCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, ctx->mKeyIn.as<uint32_t>(), ctx->mKeyOut.as<uint32_t>(),
                                                ctx->mIdxIn.as<uint32_t>(), ctx->mIdxOut.as<uint32_t>(), int(m), 0, endBit, st));
    HIPCHECK(ctx->scanTmp.ensure(std::max<size_t>(tb, 256)));
    tb = ctx->scanTmp.bytes;
    HIPCHECK(hipcub::DeviceRadixSort::SortPairs(ctx->scanTmp.p, tb, ctx->mKeyIn.as<uint32_t>(), ctx->mKeyOut.as<uint32_t>(),
                                                ctx->mIdxIn.as<uint32_t>(), ctx->mIdxOut.as<uint32_t>(), int(m), 0, endBit, st));
    HIPCHECK(ctx->mRunBeg.ensure(n * 4)); HIPCHECK(ctx->mRunEnd.ensure(n * 4));
    k_fill_u32<<<gridFor(n, 256), 256, 0, st>>>(ctx->mRunBeg.as<uint32_t>(), n, 0u);
    k_fill_u32<<<gridFor(n, 256), 256, 0, st>>>(ctx->mRunEnd.as<uint32_t>(), n, 0u);
    int err = 0;
    HIPCHECK(hipMemcpyAsync(&err, ctx->mErr.p, 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if(err) return ctx->fail(CARE_EINVAL, "malformed record (sizes, or a read id past the %llu reads)",
                             (unsigned long long)n);
    k_merge_runs<<<mb, 256, 0, st>>>(ctx->mKeyOut.as<uint32_t>(), m, ctx->mRunBeg.as<uint32_t>(), ctx->mRunEnd.as<uint32_t>());
    HIPCHECK(hipGetLastError());
    HIPCHECK(ctx->mPlan.ensure(count * sizeof(int2)));
    HIPCHECK(ctx->mSize.ensure((count + 1) * 8));
    HIPCHECK(ctx->mOutOff.ensure((count + 1) * 8));
    MergeParams mp{};
    mp.recs = ctx->mRec.as<uint8_t>(); mp.recBytes = record_bytes; mp.off = ctx->mOff.as<uint64_t>(); mp.m = m;
    mp.keysSorted = ctx->mKeyOut.as<uint32_t>(); mp.idxSorted = ctx->mIdxOut.as<uint32_t>();
    mp.runBeg = ctx->mRunBeg.as<uint32_t>(); mp.runEnd = ctx->mRunEnd.as<uint32_t>();
    mp.text = ctx->fqText[file].as<char>(); mp.hdr = ctx->fqHdr.as<uint64_t>(); mp.hdrLen = ctx->fqHdrLen.as<uint32_t>();
    mp.seq = ctx->fqSeq.as<uint64_t>(); mp.qual = ctx->fqQual.as<uint64_t>(); mp.len = ctx->len.as<int>();
    mp.n = n; mp.count = count; mp.stride = paired ? 2 : 1; mp.file = file; mp.labels = quality_labels != 0;
    mp.plan = ctx->mPlan.as<int2>(); mp.sizes = ctx->mSize.as<uint64_t>(); mp.outOff = ctx->mOutOff.as<uint64_t>();
    k_merge_plan<<<gridFor(count, 256), 256, 0, st>>>(mp);
    HIPCHECK(hipGetLastError());
    tb = 0;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, ctx->mSize.as<uint64_t>(), ctx->mOutOff.as<uint64_t>(),
                                              int(count + 1), st));
    HIPCHECK(ctx->scanTmp.ensure(std::max<size_t>(tb, 256)));
    tb = ctx->scanTmp.bytes;
    // the last size is a zero written here, so outOff[count] is the total
    k_fill_u32<<<1, 64, 0, st>>>(reinterpret_cast<uint32_t*>(ctx->mSize.as<uint64_t>() + count), 2, 0u);
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(ctx->scanTmp.p, tb, ctx->mSize.as<uint64_t>(), ctx->mOutOff.as<uint64_t>(),
                                              int(count + 1), st));
    uint64_t total = 0;
    HIPCHECK(hipMemcpyAsync(&total, ctx->mOutOff.as<uint64_t>() + count, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    *written = total;
    if(!dst) return CARE_OK;
    if(cap < total) return ctx->fail(CARE_EINVAL, "output needs %llu bytes, cap is %zu", (unsigned long long)total, cap);
    HIPCHECK(ctx->mText.ensure(std::max<uint64_t>(total, 1)));
    mp.out = ctx->mText.as<char>();
    k_merge_write<<<gridFor(count * 64, 256), 256, 0, st>>>(mp);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventRecord(ctx->ev[7], st));
    {
        const int rc = downloadD2H(ctx, dst, ctx->mText.p, total, st);
        if(rc != CARE_OK) return rc;
    }
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, ctx->ev[6], ctx->ev[7]));
    ctx->mergeMs = ms;
    return CARE_OK;
}

// ---------------------------------------------------------------- preprocessed-reads file
//
// The reference's read-storage file (--save-preprocessedreads-to / --load-preprocessedreads-from;
// ChunkedReadStorage::saveToFile / loadFromFile, include/chunkedreadstorage.hpp:158-398):
//   u64 reads, i32 longest, i32 shortest, u8 has-qualities, i32 quality bits (8),
//   u64 bytes of each of the four sections below, then
//   lengths (LengthStore::writeToStream, include/lengthstorage.hpp:184-204): i32 32, i32 bits,
//     i32 min, i32 max, u32 mask, i64 count, u64 words, u64 bytes, u32 words of length - min
//     as a big-endian bit stream of `bits` bits each (bits = bit length of max - min);
//   sequences: u64 pitch (ceil(longest/16)), u64 words, 2-bit rows;
//   qualities: u64 pitch (ceil(longest/4), 0 without qualities), u64 words, rows of bytes,
//     zero past each read's length;
//   ambiguous reads: u64 count, u32 read ids.
// The reference writes the ambiguous ids in its hash set's iteration order; here they are
// ascending (its loader puts them back into a set).

// One thread per output word of the length bit stream.
__global__ void k_pack_lengths(const int* __restrict__ len, uint64_t n, int minLen, int bits,
                               uint32_t* __restrict__ words, uint64_t nwords){
    FOR_GRID(w, nwords){
    const uint64_t b0 = 32 * w, b1 = b0 + 32;
    const uint64_t i0 = b0 / uint64_t(bits);
    const uint64_t i1 = min(n, (b1 + uint64_t(bits) - 1) / uint64_t(bits));
    uint32_t word = 0;
    for(uint64_t i = i0; i < i1; i++){
        const uint32_t v = uint32_t(len[i] - minLen);
        for(int k = 0; k < bits; k++){
            const uint64_t pos = i * uint64_t(bits) + uint64_t(k);
            if(pos >= b0 && pos < b1) word |= ((v >> (bits - 1 - k)) & 1u) << (31 - int(pos - b0));
        }
    }
    words[w] = word;
    }
}

// One thread per read: its length from the bit stream.
__global__ void k_unpack_lengths(const uint32_t* __restrict__ words, uint64_t n, int minLen, int bits,
                                 int* __restrict__ len){
    FOR_GRID(i, n){
        uint32_t v = 0;
        for(int k = 0; k < bits; k++){
            const uint64_t pos = i * uint64_t(bits) + uint64_t(k);
            v = (v << 1) | ((words[pos >> 5] >> (31 - int(pos & 31))) & 1u);
        }
        len[i] = minLen + int(v);
    }
}

// Quality rows of `sp` bytes to rows of `dp` bytes, zero past each read's length.
__global__ void k_quality_rows(const uint8_t* __restrict__ src, size_t sp, const int* __restrict__ len,
                               uint8_t* __restrict__ dst, size_t dp, uint64_t total){
    FOR_GRID(t, total){
        const uint64_t r = t / dp;
        const uint64_t c = t - r * dp;
        dst[t] = (c < sp && c < uint64_t(len[r])) ? src[r * sp + c] : uint8_t(0);
    }
}

__global__ void k_set_flags(const uint32_t* __restrict__ ids, uint64_t m, uint8_t* __restrict__ flags){
    FOR_GRID(i, m) flags[ids[i]] = 1;
}

static int care_hip_save_reads_body(care_ctx* ctx, const char* path){
    if(!ctx || !path) return CARE_EINVAL;
    if(!ctx->hasReads) return ctx->fail(CARE_ESTATE, "no reads");
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const uint64_t n = ctx->n;
    const int maxLen = ctx->maxLen;
    // shortest read
    int minLen = 0;
    {
        struct Tmp { DBuf out; ~Tmp(){ out.release(); } } tmp;
        HIPCHECK(tmp.out.ensure(4));
        size_t tb = 0;
        HIPCHECK(hipcub::DeviceReduce::Min(nullptr, tb, ctx->len.as<int>(), tmp.out.as<int>(), int(n), st));
        HIPCHECK(ctx->scanTmp.ensure(std::max<size_t>(tb, 256)));
        tb = ctx->scanTmp.bytes;
        HIPCHECK(hipcub::DeviceReduce::Min(ctx->scanTmp.p, tb, ctx->len.as<int>(), tmp.out.as<int>(), int(n), st));
        HIPCHECK(hipMemcpyAsync(&minLen, tmp.out.p, 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    }
    const int diff = maxLen - minLen;
    int bits = 0;
    while(bits < 31 && (uint32_t(1) << bits) <= uint32_t(diff)) bits++;
    if(diff == 0) bits = 0;
    const uint32_t mask = bits ? uint32_t((uint64_t(1) << bits) - 1) : 0u;
    const uint64_t lwords = bits ? (uint64_t(bits) * n + 31) / 32 : 0;
    std::vector<uint32_t> lw(lwords);
    const size_t refPitch = size_t((maxLen + 15) / 16);
    const bool q = ctx->o.use_quality_scores != 0;
    const size_t qPitch = q ? size_t((maxLen + 3) / 4) : 0;
    std::vector<uint32_t> seq(n * refPitch), qual(n * qPitch);
    std::vector<uint8_t> amb(n);
    {
        struct Tmp { DBuf words, rows; ~Tmp(){ words.release(); rows.release(); } } tmp;
        if(lwords){
            HIPCHECK(tmp.words.ensure(lwords * 4));
            k_pack_lengths<<<gridFor(lwords, 256), 256, 0, st>>>(ctx->len.as<int>(), n, minLen, bits,
                                                                               tmp.words.as<uint32_t>(), lwords);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipMemcpyAsync(lw.data(), tmp.words.p, lwords * 4, hipMemcpyDeviceToHost, st));
        }
        const uint64_t seqBytes = n * refPitch * 4;
        if(refPitch == ctx->pitch){
            HIPCHECK(hipMemcpyAsync(seq.data(), ctx->enc.p, seqBytes, hipMemcpyDeviceToHost, st));
        }else{
            HIPCHECK(tmp.rows.ensure(seqBytes));
            k_repack_rows<<<gridFor(seqBytes, 256), 256, 0, st>>>(ctx->enc.as<uint8_t>(), ctx->pitch * 4,
                                                                                tmp.rows.as<uint8_t>(), refPitch * 4, seqBytes);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipMemcpyAsync(seq.data(), tmp.rows.p, seqBytes, hipMemcpyDeviceToHost, st));
        }
        HIPCHECK(hipStreamSynchronize(st));
        if(q){
            const uint64_t qBytes = n * qPitch * 4;
            HIPCHECK(tmp.rows.ensure(qBytes));
            k_quality_rows<<<gridFor(qBytes, 256), 256, 0, st>>>(ctx->qual.as<uint8_t>(), ctx->qdpitch,
                                                                                ctx->len.as<int>(), tmp.rows.as<uint8_t>(),
                                                                                qPitch * 4, qBytes);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipMemcpyAsync(qual.data(), tmp.rows.p, qBytes, hipMemcpyDeviceToHost, st));
        }
        HIPCHECK(hipMemcpyAsync(amb.data(), ctx->ambig.p, n, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    }
    std::vector<uint32_t> ambIds;
    for(uint64_t i = 0; i < n; i++) if(amb[i]) ambIds.push_back(uint32_t(i));
    std::FILE* f = std::fopen(path, "wb");
    if(!f) return ctx->fail(CARE_EINVAL, "cannot open %s for writing", path);
    struct Closer { std::FILE* f; ~Closer(){ if(f) std::fclose(f); } } closer{f};
    bool ok = true;
    auto put = [&](const void* p, size_t b){ ok = ok && (b == 0 || std::fwrite(p, 1, b, f) == b); };
    const uint64_t nr = n;
    const uint8_t hasQ = q ? 1 : 0;
    const int qbits = 8, dtb = 32;
    const int64_t count = int64_t(n);
    const uint64_t lbytes = lwords * 4;
    const uint64_t lengthsBytes = 4 * 4 + 4 + 8 + 8 + 8 + lbytes;
    const uint64_t sequencesBytes = 16 + seq.size() * 4;
    const uint64_t qualitiesBytes = 16 + qual.size() * 4;
    const uint64_t ambigBytes = 8 + ambIds.size() * 4;
    put(&nr, 8); put(&maxLen, 4); put(&minLen, 4); put(&hasQ, 1); put(&qbits, 4);
    put(&lengthsBytes, 8); put(&sequencesBytes, 8); put(&qualitiesBytes, 8); put(&ambigBytes, 8);
    put(&dtb, 4); put(&bits, 4); put(&minLen, 4); put(&maxLen, 4); put(&mask, 4); put(&count, 8);
    put(&lwords, 8); put(&lbytes, 8); put(lw.data(), lbytes);
    const uint64_t sp = refPitch, sn = seq.size();
    put(&sp, 8); put(&sn, 8); put(seq.data(), sn * 4);
    const uint64_t qp = qPitch, qn = qual.size();
    put(&qp, 8); put(&qn, 8); put(qual.data(), qn * 4);
    const uint64_t na = ambIds.size();
    put(&na, 8); put(ambIds.data(), na * 4);
    if(!ok) return ctx->fail(CARE_EINVAL, "write to %s failed", path);
    return CARE_OK;
}

static int care_hip_load_reads_body(care_ctx* ctx, const char* path){
    if(!ctx || !path) return CARE_EINVAL;
    std::FILE* f = std::fopen(path, "rb");
    if(!f) return ctx->fail(CARE_EINVAL, "cannot open %s", path);
    struct Closer { std::FILE* f; ~Closer(){ if(f) std::fclose(f); } } closer{f};
    // every section size is checked against the bytes left in the file before anything
    // is allocated for it
    if(std::fseek(f, 0, SEEK_END) != 0) return ctx->fail(CARE_EINVAL, "%s: cannot seek", path);
    const long fileEnd = std::ftell(f);
    if(fileEnd < 0 || std::fseek(f, 0, SEEK_SET) != 0) return ctx->fail(CARE_EINVAL, "%s: cannot seek", path);
    bool ok = true;
    auto get = [&](void* p, size_t b){ ok = ok && (b == 0 || std::fread(p, 1, b, f) == b); return ok; };
    auto skip = [&](uint64_t b){ ok = ok && std::fseek(f, long(b), SEEK_CUR) == 0; };
    // true when `count` items of `size` bytes fit in what is left of the file
    auto fits = [&](uint64_t count, uint64_t size){
        const long at = std::ftell(f);
        if(at < 0 || at > fileEnd) return false;
        const uint64_t left = uint64_t(fileEnd - at);
        return size == 0 || count <= left / size;
    };
    uint64_t nr = 0, sec[4] = {0, 0, 0, 0};
    int upper = 0, lower = 0, qbits = 0;
    uint8_t hasQ = 0;
    get(&nr, 8); get(&upper, 4); get(&lower, 4); get(&hasQ, 1); get(&qbits, 4); get(sec, 32);
    int dtb = 0, bits = 0, minLen = 0, maxLen = 0;
    uint32_t mask = 0;
    int64_t count = 0;
    uint64_t lwords = 0, lbytes = 0;
    get(&dtb, 4); get(&bits, 4); get(&minLen, 4); get(&maxLen, 4); get(&mask, 4); get(&count, 8);
    get(&lwords, 8); get(&lbytes, 8);
    if(!ok) return ctx->fail(CARE_EINVAL, "%s: truncated header", path);
    if(dtb != 32 || bits < 0 || bits > 31 || count != int64_t(nr) || lbytes != lwords * 4 || maxLen < minLen || minLen < 0
       || (bits == 0 && maxLen != minLen) || (bits > 0 && lwords < (uint64_t(bits) * nr + 31) / 32) || nr == 0
       || nr >= 0xFFFFFFFFull)
        return ctx->fail(CARE_EINVAL, "%s: not a preprocessed-reads file of this format", path);
    if(!fits(lwords, 4)) return ctx->fail(CARE_EINVAL, "%s: length section larger than the file", path);
    std::vector<uint32_t> lw(lwords);
    get(lw.data(), lbytes);
    if(!ok) return ctx->fail(CARE_EINVAL, "%s: truncated", path);
    if(bits > 0){
        // every length (minLen + its bits-wide value, MSB first) must lie within
        // [minLen, maxLen]: the row pitch, the MSA bounds and the record slots come from maxLen
        uint64_t pos = 0;
        for(uint64_t i = 0; i < nr; i++){
            uint32_t v = 0;
            for(int k = 0; k < bits; k++, pos++) v = (v << 1) | ((lw[pos >> 5] >> (31 - int(pos & 31))) & 1u);
            if(int64_t(minLen) + int64_t(v) > int64_t(maxLen))
                return ctx->fail(CARE_EINVAL, "%s: read %llu is longer (%lld) than the header's maximum %d", path,
                                 (unsigned long long)i, (long long)(int64_t(minLen) + int64_t(v)), maxLen);
        }
    }
    uint64_t sp = 0, sn = 0;
    get(&sp, 8); get(&sn, 8);
    if(!ok) return ctx->fail(CARE_EINVAL, "%s: truncated", path);
    if(sp == 0 || sp > (uint64_t(1) << 20) || sp * 16 < uint64_t(maxLen) || sn / sp != nr || sn % sp != 0)
        return ctx->fail(CARE_EINVAL, "%s: bad sequence section", path);
    if(!fits(sn, 4)) return ctx->fail(CARE_EINVAL, "%s: sequence section larger than the file", path);
    std::vector<uint32_t> seq(sn);
    get(seq.data(), sn * 4);
    uint64_t qp = 0, qn = 0;
    get(&qp, 8); get(&qn, 8);
    if(!ok) return ctx->fail(CARE_EINVAL, "%s: truncated", path);
    const bool useQ = ctx->o.use_quality_scores != 0;
    std::vector<uint32_t> qual;
    if(useQ){
        // include/chunkedreadstorage.hpp:201-232
        if(!hasQ) return ctx->fail(CARE_EINVAL, "%s: quality scores expected, but none are present", path);
        if(qbits != 8) return ctx->fail(CARE_EINVAL, "%s uses %d bits per quality score, 8 are required", path, qbits);
        if(qp == 0 || qp > (uint64_t(1) << 20) || qp * 4 < uint64_t(maxLen) || qn / qp != nr || qn % qp != 0)
            return ctx->fail(CARE_EINVAL, "%s: bad quality section", path);
        if(!fits(qn, 4)) return ctx->fail(CARE_EINVAL, "%s: quality section larger than the file", path);
        qual.resize(qn);
        get(qual.data(), qn * 4);
    }else{
        if(!fits(qn, 4)) return ctx->fail(CARE_EINVAL, "%s: quality section larger than the file", path);
        skip(qn * 4);
    }
    uint64_t na = 0;
    get(&na, 8);
    if(!ok) return ctx->fail(CARE_EINVAL, "%s: truncated", path);
    if(na > nr || !fits(na, 4)) return ctx->fail(CARE_EINVAL, "%s: bad ambiguous-read section", path);
    std::vector<uint32_t> ambIds(na);
    get(ambIds.data(), na * 4);
    if(!ok) return ctx->fail(CARE_EINVAL, "%s: truncated", path);
    for(uint32_t id : ambIds) if(id >= nr) return ctx->fail(CARE_EINVAL, "%s: ambiguous read id %u out of range", path, id);
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    {
        const int rc = configureReads(ctx, nr, maxLen, size_t(sp), minLen);
        if(rc != CARE_OK) return rc;
    }
    const uint64_t n = nr;
    struct Tmp { DBuf words, rows, ids; ~Tmp(){ words.release(); rows.release(); ids.release(); } } tmp;
    HIPCHECK(ctx->len.ensure(n * 4));
    if(bits){
        HIPCHECK(tmp.words.ensure(lwords * 4));
        HIPCHECK(hipMemcpyAsync(tmp.words.p, lw.data(), lwords * 4, hipMemcpyHostToDevice, st));
        k_unpack_lengths<<<gridFor(n, 256), 256, 0, st>>>(tmp.words.as<uint32_t>(), n, minLen, bits,
                                                                         ctx->len.as<int>());
    }else{
        k_fill_u32<<<gridFor(n, 256), 256, 0, st>>>(ctx->len.as<uint32_t>(), n, uint32_t(minLen));
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(ctx->enc.ensure(sn * 4));
    HIPCHECK(hipMemcpyAsync(ctx->enc.p, seq.data(), sn * 4, hipMemcpyHostToDevice, st));
    HIPCHECK(ctx->ambig.ensure(n));
    HIPCHECK(hipMemsetAsync(ctx->ambig.p, 0, n, st));
    if(na){
        HIPCHECK(tmp.ids.ensure(na * 4));
        HIPCHECK(hipMemcpyAsync(tmp.ids.p, ambIds.data(), na * 4, hipMemcpyHostToDevice, st));
        k_set_flags<<<gridFor(na, 256), 256, 0, st>>>(tmp.ids.as<uint32_t>(), na, ctx->ambig.as<uint8_t>());
        HIPCHECK(hipGetLastError());
    }
    if(useQ){
        ctx->qpitch = size_t(maxLen);
        ctx->qdpitch = (size_t(maxLen) + 15) / 16 * 16;
        HIPCHECK(ctx->qual.ensure(n * ctx->qdpitch));
        HIPCHECK(tmp.rows.ensure(qn * 4));
        HIPCHECK(hipMemcpyAsync(tmp.rows.p, qual.data(), qn * 4, hipMemcpyHostToDevice, st));
        const uint64_t total = n * ctx->qdpitch;
        k_quality_rows<<<gridFor(total, 256), 256, 0, st>>>(tmp.rows.as<uint8_t>(), qp * 4,
                                                                           ctx->len.as<int>(), ctx->qual.as<uint8_t>(),
                                                                           ctx->qdpitch, total);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipStreamSynchronize(st));
    return finishReads(ctx);
}

static int care_hip_read_layout_body(const care_ctx* ctx, uint64_t* num_reads, int* max_length, size_t* pitch_ints){
    if(!ctx) return CARE_EINVAL;
    if(num_reads) *num_reads = ctx->hasReads ? ctx->n : 0;
    if(max_length) *max_length = ctx->hasReads ? ctx->maxLen : 0;
    if(pitch_ints) *pitch_ints = ctx->hasReads ? ctx->pitch : 0;
    return CARE_OK;
}

static int care_hip_inspect_reads_body(care_ctx* ctx, uint64_t first, uint64_t n, uint32_t* enc2bit, int32_t* lengths,
                           uint8_t* quals, size_t qual_pitch, uint8_t* ambiguous){
    if(!ctx) return CARE_EINVAL;
    if(!ctx->hasReads) return ctx->fail(CARE_ESTATE, "no reads");
    if(first > ctx->n || n > ctx->n - first) return ctx->fail(CARE_EINVAL, "range out of bounds");
    HIPCHECK(hipSetDevice(ctx->device));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    if(enc2bit && n) HIPCHECK(hipMemcpy(enc2bit, ctx->enc.as<uint32_t>() + ctx->pitch * first, n * ctx->pitch * 4, hipMemcpyDeviceToHost));
    if(lengths && n) HIPCHECK(hipMemcpy(lengths, ctx->len.as<int>() + first, n * 4, hipMemcpyDeviceToHost));
    if(ambiguous && n) HIPCHECK(hipMemcpy(ambiguous, ctx->ambig.as<uint8_t>() + first, n, hipMemcpyDeviceToHost));
    if(quals && n){
        if(!ctx->o.use_quality_scores) return ctx->fail(CARE_EINVAL, "no qualities");
        if(qual_pitch < size_t(ctx->maxLen)) return ctx->fail(CARE_EINVAL, "qual_pitch smaller than the longest read");
        // rows repacked to the caller's pitch on the device, then one linear copy
        struct Staging { DBuf b; ~Staging(){ b.release(); } } hold;
        HIPCHECK(hold.b.ensure(n * qual_pitch));
        const uint64_t total = n * qual_pitch;
        k_repack_rows<<<gridFor(total, 256), 256, 0, ctx->stream>>>(
            ctx->qual.as<uint8_t>() + ctx->qdpitch * first, ctx->qdpitch, hold.b.as<uint8_t>(), qual_pitch, total);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(quals, hold.b.p, total, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
    }
    return CARE_OK;
}

// Resolves every read's bucket per map (after the tables are built or loaded).
static int resolveBuckets(care_ctx* ctx){
    hipStream_t st = ctx->stream;
    const uint64_t total = ctx->n * uint64_t(ctx->M);
    HIPCHECK(ctx->bucketOf.ensure(total * sizeof(uint2)));
    k_bucket_of<<<gridFor((ctx->n + kBucketTile - 1) / kBucketTile, 1), 256, 0, st>>>(ctx->sig.as<uint64_t>(), ctx->n, ctx->M,
                                                                    ctx->slotPtrs.as<const Slot*>(),
                                                                    ctx->capsDev.as<uint32_t>(), ctx->bucketOf.as<uint2>());
    HIPCHECK(hipGetLastError());
    // processing order: reads sorted by their map-0 signature (reads that share it overlap)
    {
        struct Tmp { DBuf keys, ids; ~Tmp(){ keys.release(); ids.release(); } } tmp;
        const uint64_t n = ctx->n;
        HIPCHECK(ctx->order.ensure(n * 4));
        HIPCHECK(tmp.keys.ensure(n * 8));
        HIPCHECK(tmp.ids.ensure(n * 4));
        k_iota<<<gridFor(n, 256), 256, 0, st>>>(tmp.ids.as<uint32_t>(), n);
        HIPCHECK(hipGetLastError());
        size_t tb = 0;
        HIPCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, ctx->sig.as<uint64_t>(), tmp.keys.as<uint64_t>(),
                                                    tmp.ids.as<uint32_t>(), ctx->order.as<uint32_t>(), int(n), 0, 64, st));
        HIPCHECK(ctx->scanTmp.ensure(std::max<size_t>(tb, 256)));
        tb = ctx->scanTmp.bytes;
        HIPCHECK(hipcub::DeviceRadixSort::SortPairs(ctx->scanTmp.p, tb, ctx->sig.as<uint64_t>(), tmp.keys.as<uint64_t>(),
                                                    tmp.ids.as<uint32_t>(), ctx->order.as<uint32_t>(), int(n), 0, 64, st));
        HIPCHECK(hipStreamSynchronize(st));
    }
    return CARE_OK;
}

static int care_hip_build_tables_body(care_ctx* ctx, int* maps_built){
    if(!ctx) return CARE_EINVAL;
    if(!ctx->hasReads) return ctx->fail(CARE_ESTATE, "set_reads first");
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    HIPCHECK(ctx->sig.ensure(ctx->n * ctx->M * 8));
    HIPCHECK(ctx->seeds.ensure(std::max<uint64_t>(1, ctx->n * ctx->M) * 2));
    k_signatures<<<gridFor(ctx->n * uint64_t(sigGroups(ctx->M)), 256), dim3(256), 0, st>>>(
        ctx->enc.as<uint32_t>(), ctx->pitch, ctx->len.as<int>(), ctx->n, ctx->k, ctx->M, ctx->sig.as<uint64_t>(),
        ctx->seeds.as<uint16_t>());
    HIPCHECK(hipGetLastError());
    int rc = buildTables(ctx);
    if(rc != CARE_OK) return rc;
    rc = resolveBuckets(ctx);
    if(rc != CARE_OK) return rc;
    ctx->hasTables = true;
    if(maps_built) *maps_built = ctx->M;
    return CARE_OK;
}

// ---------------------------------------------------------------- table files
//
// The reference's minhasher file (OrdinaryCpuMinhasher::writeToStream/loadFromStream,
// include/ordinaryminhasher.hpp:303-339): int k, int threshold, float load factor, int map
// count, then per map (CpuReadOnlyMultiValueHashTable, include/cpuhashtable.hpp:627-649)
// u64 value count and the u32 values (the kept buckets, in key order, read ids ascending),
// then the key -> (offset, count) table (AoSCpuSingleValueHashTable, :221-248): float load,
// u64 keys, u64 max probes, u64 size, u64 capacity, u64 slot count and the slots of 16 bytes
// {u64 key, u32 offset, u16 count, 2 padding bytes}; capacity = size_t(keys / load), slot =
// fmix64(key) % capacity with linear probing (:96-125), empty = {max key, value-initialized
// (offset, count) = (0, 0)} (:283-284: numeric_limits of a pair is its default value). Keys are inserted
// in ascending order, as the reference's construction does, so the layout is the
// reference's.

constexpr float kTableFileLoad = 0.8f;   // ProgramOptions::hashtableLoadfactor default

struct FileSlot { uint64_t key; uint32_t offset; uint16_t count; uint16_t pad; };
static_assert(sizeof(FileSlot) == 16, "file slot layout");

static int care_hip_save_tables_body(care_ctx* ctx, const char* path){
    if(!ctx || !path) return CARE_EINVAL;
    if(!ctx->hasTables) return ctx->fail(CARE_ESTATE, "build_tables first");
    HIPCHECK(hipSetDevice(ctx->device));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    std::FILE* f = std::fopen(path, "wb");
    if(!f) return ctx->fail(CARE_EINVAL, "cannot open %s for writing", path);
    struct Closer { std::FILE* f; ~Closer(){ if(f) std::fclose(f); } } closer{f};
    bool ok = true;
    auto put = [&](const void* p, size_t b){ ok = ok && std::fwrite(p, 1, b, f) == b; };
    const int k = ctx->k, thr = ctx->thr, M = ctx->M;
    const float lf = kTableFileLoad;
    put(&k, 4); put(&thr, 4); put(&lf, 4); put(&M, 4);
    std::vector<uint32_t> vals(ctx->n);
    for(int m = 0; m < M; m++){
        std::vector<Slot> slots(ctx->caps[m]);
        HIPCHECK(hipMemcpy(slots.data(), ctx->slotBufs[m].p, slots.size() * sizeof(Slot), hipMemcpyDeviceToHost));
        if(ctx->n) HIPCHECK(hipMemcpy(vals.data(), ctx->values.as<uint32_t>() + uint64_t(m) * ctx->n, ctx->n * 4,
                                      hipMemcpyDeviceToHost));
        std::vector<Slot> kept;
        for(const Slot& sl : slots) if(sl.count) kept.push_back(sl);
        std::sort(kept.begin(), kept.end(), [](const Slot& x, const Slot& y){ return x.key < y.key; });
        std::vector<uint32_t> packed;
        std::vector<uint64_t> offs(kept.size());
        for(size_t i = 0; i < kept.size(); i++){
            offs[i] = packed.size();
            packed.insert(packed.end(), vals.begin() + kept[i].offset, vals.begin() + kept[i].offset + kept[i].count);
        }
        const uint64_t elements = packed.size();
        put(&elements, 8);
        put(packed.data(), elements * 4);
        const uint64_t size = kept.size();
        const uint64_t capacity = uint64_t(size / lf);
        std::vector<FileSlot> table(capacity, FileSlot{~0ull, 0u, uint16_t(0), 0});
        uint64_t maxProbes = 0;
        for(size_t i = 0; i < kept.size(); i++){
            uint64_t pos = fmix64(kept[i].key) % capacity, probes = 0;
            while(table[pos].key != ~0ull){ pos = pos + 1 == capacity ? 0 : pos + 1; probes++; }
            table[pos] = FileSlot{kept[i].key, uint32_t(offs[i]), uint16_t(kept[i].count), 0};
            maxProbes = std::max(maxProbes, probes);
        }
        put(&lf, 4); put(&size, 8); put(&maxProbes, 8); put(&size, 8); put(&capacity, 8); put(&capacity, 8);
        put(table.data(), capacity * sizeof(FileSlot));
    }
    if(!ok) return ctx->fail(CARE_EINVAL, "writing %s failed", path);
    return CARE_OK;
}

static int care_hip_load_tables_body(care_ctx* ctx, const char* path, int* maps_loaded){
    if(!ctx || !path) return CARE_EINVAL;
    if(!ctx->hasReads) return ctx->fail(CARE_ESTATE, "set_reads first");
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    std::FILE* f = std::fopen(path, "rb");
    if(!f) return ctx->fail(CARE_EINVAL, "cannot open %s", path);
    struct Closer { std::FILE* f; ~Closer(){ if(f) std::fclose(f); } } closer{f};
    if(std::fseek(f, 0, SEEK_END) != 0) return ctx->fail(CARE_EINVAL, "%s: cannot seek", path);
    const long fileEnd = std::ftell(f);
    if(fileEnd < 0 || std::fseek(f, 0, SEEK_SET) != 0) return ctx->fail(CARE_EINVAL, "%s: cannot seek", path);
    auto get = [&](void* p, size_t b){ return std::fread(p, 1, b, f) == b; };
    // true when `count` items of `size` bytes fit in what is left of the file
    auto fits = [&](uint64_t count, uint64_t size){
        const long at = std::ftell(f);
        if(at < 0 || at > fileEnd) return false;
        return size == 0 || count <= uint64_t(fileEnd - at) / size;
    };
    int k = 0, thr = 0, maps = 0;
    float lf = 0.0f;
    if(!get(&k, 4) || !get(&thr, 4) || !get(&lf, 4) || !get(&maps, 4)) return ctx->fail(CARE_EINVAL, "%s: truncated header", path);
    if(k != ctx->k) return ctx->fail(CARE_EINVAL, "%s: k-mer length %d, engine uses %d", path, k, ctx->k);
    if(thr != ctx->thr) return ctx->fail(CARE_EINVAL, "%s: bucket threshold %d, engine uses %d", path, thr, ctx->thr);
    if(maps < 1) return ctx->fail(CARE_EINVAL, "%s: no maps", path);
    const int M = std::min(maps, ctx->o.num_hash_maps);
    ctx->M = M;
    HIPCHECK(ctx->sig.ensure(ctx->n * M * 8));
    HIPCHECK(ctx->seeds.ensure(std::max<uint64_t>(1, ctx->n * M) * 2));
    k_signatures<<<gridFor(ctx->n * uint64_t(sigGroups(M)), 256), dim3(256), 0, st>>>(
        ctx->enc.as<uint32_t>(), ctx->pitch, ctx->len.as<int>(), ctx->n, ctx->k, M, ctx->sig.as<uint64_t>(),
        ctx->seeds.as<uint16_t>());
    HIPCHECK(hipGetLastError());
    HIPCHECK(ctx->values.ensure(std::max<uint64_t>(1, ctx->n) * M * sizeof(uint32_t)));
    ctx->slotBufs.resize(M);
    ctx->caps.assign(M, 0);
    std::vector<Slot*> ptrs48(kMaxMaps, nullptr);
    std::vector<uint32_t> caps48(kMaxMaps, 1);
    for(int m = 0; m < M; m++){
        uint64_t elements = 0;
        if(!get(&elements, 8)) return ctx->fail(CARE_EINVAL, "%s: truncated map %d", path, m);
        if(elements > ctx->n) return ctx->fail(CARE_EINVAL, "%s: map %d holds %llu values for %llu reads", path, m,
                                               (unsigned long long)elements, (unsigned long long)ctx->n);
        if(!fits(elements, 4)) return ctx->fail(CARE_EINVAL, "%s: values of map %d larger than the file", path, m);
        std::vector<uint32_t> vals(elements);
        if(!get(vals.data(), elements * 4)) return ctx->fail(CARE_EINVAL, "%s: truncated values of map %d", path, m);
        for(uint32_t v : vals) if(v >= ctx->n) return ctx->fail(CARE_EINVAL, "%s: read id %u out of range", path, v);
        float load = 0.0f;
        uint64_t numKeys = 0, maxProbes = 0, size = 0, capacity = 0, slotCount = 0;
        if(!get(&load, 4) || !get(&numKeys, 8) || !get(&maxProbes, 8) || !get(&size, 8) || !get(&capacity, 8) ||
           !get(&slotCount, 8)) return ctx->fail(CARE_EINVAL, "%s: truncated table of map %d", path, m);
        if(!fits(slotCount, sizeof(FileSlot))) return ctx->fail(CARE_EINVAL, "%s: table of map %d larger than the file", path, m);
        std::vector<FileSlot> table(slotCount);
        if(!get(table.data(), slotCount * sizeof(FileSlot))) return ctx->fail(CARE_EINVAL, "%s: truncated slots of map %d", path, m);
        std::vector<Slot> entries;
        for(const FileSlot& fs : table){
            if(fs.key == ~0ull) continue;                      // empty slot
            if(uint64_t(fs.offset) + fs.count > elements) return ctx->fail(CARE_EINVAL, "%s: bucket past the values", path);
            entries.push_back(Slot{fs.key, fs.offset, fs.count});
        }
        uint32_t cap = 64;
        while(cap < 2u * entries.size()) cap <<= 1;
        std::vector<Slot> slots(cap, Slot{0, 0, 0});
        for(const Slot& e : entries){
            if(e.count == 0) continue;
            uint32_t h = uint32_t(fmix64(e.key)) & (cap - 1);
            while(slots[h].count) h = (h + 1) & (cap - 1);
            slots[h] = e;
        }
        ctx->caps[m] = cap;
        HIPCHECK(ctx->slotBufs[m].ensure(size_t(cap) * sizeof(Slot)));
        HIPCHECK(hipMemcpy(ctx->slotBufs[m].p, slots.data(), size_t(cap) * sizeof(Slot), hipMemcpyHostToDevice));
        if(elements) HIPCHECK(hipMemcpy(ctx->values.as<uint32_t>() + uint64_t(m) * ctx->n, vals.data(), elements * 4,
                                        hipMemcpyHostToDevice));
        ptrs48[m] = ctx->slotBufs[m].as<Slot>();
        caps48[m] = cap;
    }
    HIPCHECK(ctx->slotPtrs.ensure(sizeof(Slot*) * kMaxMaps));
    HIPCHECK(ctx->capsDev.ensure(sizeof(uint32_t) * kMaxMaps));
    HIPCHECK(hipMemcpy(ctx->slotPtrs.p, ptrs48.data(), sizeof(Slot*) * kMaxMaps, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(ctx->capsDev.p, caps48.data(), sizeof(uint32_t) * kMaxMaps, hipMemcpyHostToDevice));
    HIPCHECK(hipStreamSynchronize(st));
    const int rc = resolveBuckets(ctx);
    if(rc != CARE_OK) return rc;
    ctx->hasTables = true;
    if(maps_loaded) *maps_loaded = M;
    return CARE_OK;
}

static int care_hip_correct_body(care_ctx* ctx, uint64_t first, uint64_t n, size_t* num_records, size_t* num_bytes){
    if(!ctx) return CARE_EINVAL;
    if(!ctx->hasTables) return ctx->fail(CARE_ESTATE, "build_tables first");
    if(first > ctx->n || n > ctx->n - first) return ctx->fail(CARE_EINVAL, "anchor range out of bounds");
    if(ctx->o.paired && ((first & 1) || (n & 1))) return ctx->fail(CARE_EINVAL, "paired input: first and n must be even");
    HIPCHECK(hipSetDevice(ctx->device));
    ctx->t = care_timings{};
    ctx->records.clear();
    ctx->numRecords = 0;
    HIPCHECK(ctx->counters.ensure(128));
    HIPCHECK(hipMemsetAsync(ctx->counters.p, 0, 128, ctx->stream));
    int B = ctx->o.batch_size > 0 ? ctx->o.batch_size : 0;
    if(B <= 0){
        // one batch per pass where HBM allows: the per-anchor candidate staging area
        // (maps * threshold ids) and record slot may take up to 96 GiB or a third of the
        // free device memory (288 GB of HBM), at most 2^23 anchors; fewer batches mean
        // fewer launch tails
        size_t freeB = 0, totalB = 0;
        HIPCHECK(hipMemGetInfo(&freeB, &totalB));
        const size_t budget = std::min<size_t>(size_t(96) << 30, freeB / 3);
        size_t perAnchor = size_t(ctx->M) * ctx->thr * 4 + size_t(recordSlotBytes(ctx->maxLen)) + 512;
        // candidate correction: a record slot per working-list entry (about threshold per anchor)
        if(ctx->o.correct_candidates) perAnchor += size_t(ctx->thr) * (candidateSlotBytes(ctx->maxLen) + 36);
        B = int(std::min<size_t>(size_t(1) << 23, std::max<size_t>(1024, budget / perAnchor)));
        // batches of equal size (no short last batch): 2 of 4.29M at 8.57M reads
        if(n > 0){
            const uint64_t nb = (n + uint64_t(B) - 1) / uint64_t(B);
            uint64_t even = (n + nb - 1) / nb;
            if(ctx->o.correct_candidates) even = (even + 15) & ~uint64_t(15);
            B = int(std::min<uint64_t>(uint64_t(B), std::max<uint64_t>(even, 16)));
        }
    }
    if(ctx->o.paired && (B & 1)) B++;
    const bool cands = ctx->o.correct_candidates != 0;
    // candidate correction orders records by 16-read groups of read ids: batches end on
    // group boundaries
    if(cands) B = std::max(16, B & ~15);
    // the fused kernel keeps MSA columns in registers (<= 1024 columns) and prefetches a
    // candidate's four planes into 3 uint2 per lane (4 * (W + 1) <= 192)
    const bool fusedFits = waveMsaColumnsPerLane(ctx->maxCols) > 0 && 4 * (ctx->W + 1) <= 3 * kWave;
    if(ctx->o.engine == 2 && !fusedFits) return ctx->fail(CARE_EINVAL, "reads too long for the fused engine");
    const bool fused = !ctx->o.keep_intermediates && ctx->o.engine == 2 && fusedFits;
    if(ctx->o.correction_type == 1 && !ctx->hasSeam) return ctx->fail(CARE_ESTATE, "learned correction: set_classifier first");
    if((ctx->o.correction_type == 1 || ctx->featActive) && waveMsaColumnsPerLane(ctx->maxCols) == 0)
        return ctx->fail(CARE_EINVAL, "learned correction and feature output need reads of at most 340 bases");
    if(cands && waveMsaColumnsPerLane(ctx->maxCols) == 0)
        return ctx->fail(CARE_EINVAL, "candidate correction needs reads of at most 340 bases");
    if(cands && ctx->o.engine == 2) return ctx->fail(CARE_EINVAL, "candidate correction runs on the staged engine");
    if(ctx->o.seeded_alignment && ctx->o.engine == 2) return ctx->fail(CARE_EINVAL, "seeded alignment runs on the staged engine");
    if(ctx->o.seeded_alignment && alignLaneWords(ctx->W) == 0)
        return ctx->fail(CARE_EINVAL, "seeded alignment handles reads of at most 352 bases");
    if(cands){
        HIPCHECK(ctx->hqAnchor.ensure(ctx->n));
        ctx->callFirst = first;
    }
    // pipelined batches (no host synchronization inside a batch) for the plain and learned
    // correction with the lane-per-candidate or seeded alignment; candidate correction,
    // intermediate dumps and feature output synchronize per batch
    const bool pipelined = !fused && !cands && !ctx->o.keep_intermediates && !ctx->featActive && !ctx->o.sync_batches &&
                           (ctx->o.seeded_alignment || (ctx->o.align_variant != 1 && alignLaneWords(ctx->W) > 0));
    unsigned long long pipeGathered = 0;
    const auto t0 = std::chrono::steady_clock::now();
    if(fused){
        const int rc = runFused(ctx, first, n, ctx->records, ctx->numRecords);
        if(rc != CARE_OK) return rc;
    }else if(pipelined){
        const int rc = runPipelined(ctx, first, n, B, ctx->records, ctx->numRecords, pipeGathered);
        if(rc != CARE_OK) return rc;
    }else{
        for(uint64_t b0 = first; b0 < first + n;){
            uint64_t e = std::min<uint64_t>(b0 + uint64_t(B), first + n);
            if(cands && e < first + n && (e & ~uint64_t(15)) > b0) e &= ~uint64_t(15);
            const int rc = runBatch(ctx, b0, int(e - b0), ctx->records, ctx->numRecords);
            if(rc != CARE_OK) return rc;
            ctx->t.batches++;
            b0 = e;
        }
    }
    unsigned long long cnt[4] = {0, 0, 0, 0};
    HIPCHECK(hipMemcpy(cnt, ctx->counters.p, 32, hipMemcpyDeviceToHost));
    const unsigned long long gathered = cnt[0] + pipeGathered;
    if(fused) ctx->t.candidates = cnt[3];
    const auto t1 = std::chrono::steady_clock::now();
    ctx->t.total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    ctx->t.anchors = n;
    ctx->t.records = ctx->numRecords;
    ctx->t.record_bytes = ctx->records.size();
    ctx->t.bucket_values = gathered;
    ctx->t.graphs_built = ctx->graphsBuilt;
    if(num_records) *num_records = ctx->numRecords;
    if(num_bytes) *num_bytes = ctx->records.size();
    return CARE_OK;
}

static int care_hip_copy_records_body(care_ctx* ctx, uint8_t* dst, size_t cap){
    if(!ctx || (!dst && !ctx->records.empty())) return CARE_EINVAL;
    if(cap < ctx->records.size()) return ctx->fail(CARE_EINVAL, "buffer too small");
    if(!ctx->records.empty()) std::memcpy(dst, ctx->records.data(), ctx->records.size());
    return CARE_OK;
}

static int care_hip_records_view_body(const care_ctx* ctx, const uint8_t** data, size_t* bytes){
    if(!ctx || !data || !bytes) return CARE_EINVAL;
    *data = ctx->records.data();
    *bytes = ctx->records.size();
    return CARE_OK;
}

static int care_hip_set_classifier_body(care_ctx* ctx, const float* params, size_t count){
    if(!ctx) return CARE_EINVAL;
    if(!params){ ctx->hasSeam = false; return CARE_OK; }
    if(count != size_t(kMlpParams)) return ctx->fail(CARE_EINVAL, "classifier needs %d parameters, got %zu", kMlpParams, count);
    for(size_t i = 0; i < count; i++)
        if(!std::isfinite(params[i])) return ctx->fail(CARE_EINVAL, "classifier parameter %zu is not finite", i);
    // device layout: W1^T [16][64], W2^T [64][64], W3 [64] as bfloat16 (round to nearest
    // even), then b1, b2, b3 as float; the host layout is W1[64][16], b1, W2[64][64], b2, W3, b3
    auto bf = [](float x){
        uint32_t u;
        std::memcpy(&u, &x, 4);
        u += 0x7FFFu + ((u >> 16) & 1u);
        return uint16_t(u >> 16);
    };
    const float* W1 = params;
    const float* b1 = W1 + kHid * kFeat;
    const float* W2 = b1 + kHid;
    const float* b2 = W2 + kHid * kHid;
    const float* W3 = b2 + kHid;
    const float* b3 = W3 + kHid;
    std::vector<uint8_t> blob(kSeamWeightBytes + (2 * kHid + 1) * 4);
    uint16_t* w = reinterpret_cast<uint16_t*>(blob.data());
    for(int k = 0; k < kFeat; k++)
        for(int j = 0; j < kHid; j++) w[k * kHid + j] = bf(W1[j * kFeat + k]);
    uint16_t* w2 = w + kFeat * kHid;
    for(int k = 0; k < kHid; k++)
        for(int j = 0; j < kHid; j++) w2[k * kHid + j] = bf(W2[j * kHid + k]);
    uint16_t* w3 = w2 + kHid * kHid;
    for(int j = 0; j < kHid; j++) w3[j] = bf(W3[j]);
    float* b = reinterpret_cast<float*>(blob.data() + kSeamWeightBytes);
    std::memcpy(b, b1, kHid * 4);
    std::memcpy(b + kHid, b2, kHid * 4);
    b[2 * kHid] = b3[0];
    HIPCHECK(hipSetDevice(ctx->device));
    HIPCHECK(ctx->seam.ensure(blob.size()));
    HIPCHECK(hipMemcpy(ctx->seam.p, blob.data(), blob.size(), hipMemcpyHostToDevice));
    ctx->hasSeam = true;
    return CARE_OK;
}

static int care_hip_extract_features_body(care_ctx* ctx, uint64_t first, uint64_t n, float* features, uint32_t* meta,
                              size_t cap, size_t* total){
    if(!ctx || !total || (cap && (!features || !meta))) return CARE_EINVAL;
    HIPCHECK(hipSetDevice(ctx->device));
    HIPCHECK(ctx->featBuf.ensure(std::max<size_t>(cap, 1) * kFeat * 4));
    HIPCHECK(ctx->metaBuf.ensure(std::max<size_t>(cap, 1) * 16));
    ctx->featCap = cap;
    ctx->featActive = true;
    const int rc = care_hip_correct(ctx, first, n, nullptr, nullptr);
    ctx->featActive = false;
    if(rc != CARE_OK) return rc;
    unsigned long long count = 0;
    HIPCHECK(hipMemcpy(&count, ctx->counters.as<unsigned long long>() + 4, 8, hipMemcpyDeviceToHost));
    *total = size_t(count);
    const size_t m = std::min<size_t>(size_t(count), cap);
    if(m == 0) return CARE_OK;
    std::vector<float> f(m * kFeat);
    std::vector<uint32_t> q(m * 4);
    HIPCHECK(hipMemcpy(f.data(), ctx->featBuf.p, m * kFeat * 4, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(q.data(), ctx->metaBuf.p, m * 16, hipMemcpyDeviceToHost));
    // waves append in completion order: return the columns ordered by (read id, position)
    std::vector<size_t> order(m);
    std::iota(order.begin(), order.end(), size_t(0));
    std::sort(order.begin(), order.end(), [&](size_t x, size_t y){
        return q[4 * x] != q[4 * y] ? q[4 * x] < q[4 * y] : q[4 * x + 1] < q[4 * y + 1];
    });
    for(size_t r = 0; r < m; r++){
        std::memcpy(features + r * kFeat, f.data() + order[r] * kFeat, kFeat * 4);
        std::memcpy(meta + r * 4, q.data() + order[r] * 4, 16);
    }
    return CARE_OK;
}

static int care_hip_get_timings_body(const care_ctx* ctx, care_timings* out){
    if(!ctx || !out) return CARE_EINVAL;
    *out = ctx->t;
    out->fastq_device_ms = ctx->fastqMs;
    out->merge_device_ms = ctx->mergeMs;
    return CARE_OK;
}

static int care_hip_inspect_signatures_body(care_ctx* ctx, uint64_t first, uint64_t n, uint64_t* out){
    if(!ctx || !out) return CARE_EINVAL;
    if(!ctx->hasTables) return ctx->fail(CARE_ESTATE, "build_tables first");
    if(first > ctx->n || n > ctx->n - first) return ctx->fail(CARE_EINVAL, "range out of bounds");
    HIPCHECK(hipSetDevice(ctx->device));
    std::vector<uint64_t> col(n);
    for(int m = 0; m < ctx->M; m++){
        HIPCHECK(hipMemcpy(col.data(), ctx->sig.as<uint64_t>() + uint64_t(m) * ctx->n + first, n * 8, hipMemcpyDeviceToHost));
        for(uint64_t r = 0; r < n; r++) out[r * ctx->M + m] = col[r];
    }
    return CARE_OK;
}

static int care_hip_inspect_batch_body(care_ctx* ctx, uint64_t* batch_first, uint64_t* batch_n, uint64_t* total,
                           uint32_t* candidate_offsets, uint32_t* candidate_ids, int32_t* alignments,
                           uint8_t* paired, int32_t* anchor_info){
    if(!ctx) return CARE_EINVAL;
    if(!ctx->o.keep_intermediates) return ctx->fail(CARE_ESTATE, "keep_intermediates is off");
    if(ctx->lastB == 0) return ctx->fail(CARE_ESTATE, "no batch yet");
    HIPCHECK(hipSetDevice(ctx->device));
    if(batch_first) *batch_first = ctx->lastB0;
    if(batch_n) *batch_n = ctx->lastB;
    if(total) *total = ctx->lastTotal;
    const size_t B = ctx->lastB, T = ctx->lastTotal;
    if(candidate_offsets) HIPCHECK(hipMemcpy(candidate_offsets, ctx->offsets.p, (B + 1) * 4, hipMemcpyDeviceToHost));
    if(candidate_ids && T) HIPCHECK(hipMemcpy(candidate_ids, ctx->ids.p, T * 4, hipMemcpyDeviceToHost));
    if(alignments && T) HIPCHECK(hipMemcpy(alignments, ctx->dbg.p, T * 11 * 4, hipMemcpyDeviceToHost));
    if(paired && T){
        if(ctx->o.paired) HIPCHECK(hipMemcpy(paired, ctx->paired.p, T, hipMemcpyDeviceToHost));
        else std::memset(paired, 0, T);
    }
    if(anchor_info) HIPCHECK(hipMemcpy(anchor_info, ctx->info.p, B * kAnchorInfo * 4, hipMemcpyDeviceToHost));
    return CARE_OK;
}

static int care_hip_inspect_msa_body(care_ctx* ctx, uint64_t index, int cap_columns, int* num_columns, int32_t* counts,
                         float* weights, int32_t* coverage, uint8_t* consensus, float* support){
    if(!ctx) return CARE_EINVAL;
    if(!ctx->o.keep_intermediates) return ctx->fail(CARE_ESTATE, "keep_intermediates is off");
    if(index >= ctx->lastB) return ctx->fail(CARE_EINVAL, "index out of range");
    HIPCHECK(hipSetDevice(ctx->device));
    std::vector<uint8_t> blob(ctx->dbgStride);
    HIPCHECK(hipMemcpy(blob.data(), ctx->msaDbg.as<uint8_t>() + index * ctx->dbgStride, ctx->dbgStride, hipMemcpyDeviceToHost));
    int info[kAnchorInfo];
    HIPCHECK(hipMemcpy(info, ctx->info.as<int>() + index * kAnchorInfo, sizeof(info), hipMemcpyDeviceToHost));
    if(info[1] == 0 && info[2] == 0){ if(num_columns) *num_columns = 0; return CARE_OK; }
    const int C = ctx->maxCols;
    int ncol;
    memcpy(&ncol, blob.data(), 4);
    if(num_columns) *num_columns = ncol;
    const int* dc = reinterpret_cast<const int*>(blob.data()) + 1;
    const float* dw = reinterpret_cast<const float*>(dc + 4 * C);
    const int* dv = reinterpret_cast<const int*>(dw + 4 * C);
    const float* ds = reinterpret_cast<const float*>(dv + C);
    const uint8_t* dk = reinterpret_cast<const uint8_t*>(ds + C);
    const int m = std::min(ncol, cap_columns);
    for(int c = 0; c < m; c++){
        for(int b = 0; b < 4; b++){
            if(counts) counts[4 * c + b] = dc[4 * c + b];
            if(weights) weights[4 * c + b] = dw[4 * c + b];
        }
        if(coverage) coverage[c] = dv[c];
        if(support) support[c] = ds[c];
        if(consensus) consensus[c] = dk[c];
    }
    return CARE_OK;
}


// Exported entry points: each runs its body under guarded().
int care_hip_create(const care_opts* opts, int device, care_ctx** out){
    return guarded(nullptr, [&]() { return care_hip_create_body(opts, device, out); });
}

int care_hip_device_info(int device, char* name, size_t cap, int* compute_units){
    return guarded(nullptr, [&]() { return care_hip_device_info_body(device, name, cap, compute_units); });
}

int care_hip_encode_reads(const uint8_t* seqs, size_t seq_pitch, const int32_t* lengths, uint64_t num_reads,
                          uint32_t* enc2bit, size_t pitch_ints, uint8_t* ambiguous,
                          const uint8_t* quals_in, size_t quals_in_pitch, uint8_t* quals_out,
                          size_t quals_out_pitch, int threads){
    return guarded(nullptr, [&]() { return care_hip_encode_reads_body(seqs, seq_pitch, lengths, num_reads, enc2bit, pitch_ints, ambiguous, quals_in, quals_in_pitch, quals_out, quals_out_pitch, threads); });
}

int care_hip_set_reads(care_ctx* ctx, const uint32_t* enc2bit, size_t pitch_ints, const int32_t* lengths,
                       const uint8_t* quals, size_t qual_pitch, uint64_t num_reads, const uint8_t* ambiguous){
    return guarded(ctx, [&]() { return care_hip_set_reads_body(ctx, enc2bit, pitch_ints, lengths, quals, qual_pitch, num_reads, ambiguous); });
}

int care_hip_set_reads_fastq(care_ctx* ctx, const char* fastq1, size_t bytes1, const char* fastq2, size_t bytes2){
    return guarded(ctx, [&]() { return care_hip_set_reads_fastq_body(ctx, fastq1, bytes1, fastq2, bytes2); });
}

int care_hip_write_fastq(care_ctx* ctx, const uint8_t* records, size_t record_bytes, int file, int quality_labels,
                         char* dst, size_t cap, size_t* written){
    return guarded(ctx, [&]() { return care_hip_write_fastq_body(ctx, records, record_bytes, file, quality_labels, dst, cap, written); });
}

int care_hip_save_reads(care_ctx* ctx, const char* path){
    return guarded(ctx, [&]() { return care_hip_save_reads_body(ctx, path); });
}

int care_hip_load_reads(care_ctx* ctx, const char* path){
    return guarded(ctx, [&]() { return care_hip_load_reads_body(ctx, path); });
}

int care_hip_read_layout(const care_ctx* ctx, uint64_t* num_reads, int* max_length, size_t* pitch_ints){
    return guarded(nullptr, [&]() { return care_hip_read_layout_body(ctx, num_reads, max_length, pitch_ints); });
}

int care_hip_inspect_reads(care_ctx* ctx, uint64_t first, uint64_t n, uint32_t* enc2bit, int32_t* lengths,
                           uint8_t* quals, size_t qual_pitch, uint8_t* ambiguous){
    return guarded(ctx, [&]() { return care_hip_inspect_reads_body(ctx, first, n, enc2bit, lengths, quals, qual_pitch, ambiguous); });
}

int care_hip_build_tables(care_ctx* ctx, int* maps_built){
    return guarded(ctx, [&]() { return care_hip_build_tables_body(ctx, maps_built); });
}

int care_hip_save_tables(care_ctx* ctx, const char* path){
    return guarded(ctx, [&]() { return care_hip_save_tables_body(ctx, path); });
}

int care_hip_load_tables(care_ctx* ctx, const char* path, int* maps_loaded){
    return guarded(ctx, [&]() { return care_hip_load_tables_body(ctx, path, maps_loaded); });
}

int care_hip_correct(care_ctx* ctx, uint64_t first, uint64_t n, size_t* num_records, size_t* num_bytes){
    return guarded(ctx, [&]() { return care_hip_correct_body(ctx, first, n, num_records, num_bytes); });
}

int care_hip_copy_records(care_ctx* ctx, uint8_t* dst, size_t cap){
    return guarded(ctx, [&]() { return care_hip_copy_records_body(ctx, dst, cap); });
}

int care_hip_records_view(const care_ctx* ctx, const uint8_t** data, size_t* bytes){
    return guarded(nullptr, [&]() { return care_hip_records_view_body(ctx, data, bytes); });
}

int care_hip_set_classifier(care_ctx* ctx, const float* params, size_t count){
    return guarded(ctx, [&]() { return care_hip_set_classifier_body(ctx, params, count); });
}

int care_hip_extract_features(care_ctx* ctx, uint64_t first, uint64_t n, float* features, uint32_t* meta,
                              size_t cap, size_t* total){
    return guarded(ctx, [&]() { return care_hip_extract_features_body(ctx, first, n, features, meta, cap, total); });
}

int care_hip_get_timings(const care_ctx* ctx, care_timings* out){
    return guarded(nullptr, [&]() { return care_hip_get_timings_body(ctx, out); });
}

int care_hip_inspect_signatures(care_ctx* ctx, uint64_t first, uint64_t n, uint64_t* out){
    return guarded(ctx, [&]() { return care_hip_inspect_signatures_body(ctx, first, n, out); });
}

int care_hip_inspect_batch(care_ctx* ctx, uint64_t* batch_first, uint64_t* batch_n, uint64_t* total,
                           uint32_t* candidate_offsets, uint32_t* candidate_ids, int32_t* alignments,
                           uint8_t* paired, int32_t* anchor_info){
    return guarded(ctx, [&]() { return care_hip_inspect_batch_body(ctx, batch_first, batch_n, total, candidate_offsets, candidate_ids, alignments, paired, anchor_info); });
}

int care_hip_inspect_msa(care_ctx* ctx, uint64_t index, int cap_columns, int* num_columns, int32_t* counts,
                         float* weights, int32_t* coverage, uint8_t* consensus, float* support){
    return guarded(ctx, [&]() { return care_hip_inspect_msa_body(ctx, index, cap_columns, num_columns, counts, weights, coverage, consensus, support); });
}

}  // extern "C"

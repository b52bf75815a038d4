// Count stage on gfx950: exact (k-mer, count) tables from superkmer records.
//
// Kernels:
//  1. Grouping (k_group_hist, k_group_offsets, k_group), 16 record segments per bucket (a
//     partition): records are grouped by minimizer group (the top 10 of the 12 sub-bucket bits
//     taken from ~mm above the partition digits) and each record's bases are re-encoded once,
//     from the reference's 3-bases-per-byte record bytes, into word-aligned 2-bit words (LSB-first)
//     laid out group by group. All k-mers of one record share its minimizer, and the minimizer is a
//     function of the canonical k-mer, so sub-buckets hold disjoint key sets.
//  2. k_plan: each bucket's sub-bucket sequence is cut into work items sized to the LDS table: runs
//     of light sub-buckets, and hash slices of a sub-bucket larger than one item ("heavy"). Each
//     item gets a key region: exactly its instances for a light run, the slice's share with a
//     margin for a heavy slice.
//  2b. k_expand, one 256-thread workgroup per record group, six per CU: every k-mer is extracted
//     once and its canonical key (the digit-reversed 2k-bit window against the complemented one,
//     cpu_funcs.cpp:427-436 / gpu_kmercounting.cu:63-74) is scattered to its item's region. A heavy
//     slice that received more keys than its margin (a very frequent k-mer) gets an exact region in
//     a second key area, filled by a second pass over its group only (k_ovf_sizes, k_expand<OVF>).
//  2c. k_item_prep: one 64-byte descriptor per item (its region and key count, bucket bases).
//  3. k_count_items, persistent: 256-thread workgroups, four per CU, each taking items b, b + G,
//     b + 2G, ... and fetching the next item's descriptor while counting the current one. A sweep
//     loads its region's keys in batches of 16 per thread and counts them in an open-addressing LDS
//     table (64-bit keys: four-slot probe groups claimed by CAS; 128-bit keys publish the high word
//     last). A sweep that does not fit (a probe window exhausted, or more distinct keys than the
//     sweep capacity) is cut into hash slices of the region until it fits.
//     With expansion off (or its regions too large for the free memory) items read records
//     instead: a chunk of records is one contiguous range of 2-bit words copied into LDS, every
//     thread takes whole k-mers, and a sweep that does not fit is split by sub-bucket range, then
//     into hash slices of one sub-bucket.
//  4. The sweep's distinct keys with min <= count <= max (utilities.hpp:190) are compacted in LDS
//     and appended to the partition's row region with coalesced stores, one atomic per sweep.
//     rgkc_sort_rows (k_sort.hip) puts a partition's rows in ascending key order afterwards, the
//     order of the reference's table (cpu_funcs.cpp:614-640).
#include <algorithm>

#include "rgkc_device.h"
#include "rgkc_internal.h"

namespace rgkc {
namespace {

constexpr int T = kCountThreads;
constexpr int kStack = 128;               // pending sweeps of one work item
constexpr uint32_t kMaxSlices = 32;
constexpr int kNG = kGroups;              // record groups per bucket
constexpr int kSubPerGroup = kSubIds / kGroups;
constexpr int kGroupChunk = 2048;         // records per k_group chunk (at most)
constexpr uint32_t kGroupChunkBytes = 32 * 1024;  // store bytes per k_group chunk (staged in LDS)
constexpr uint32_t kStageW = 1000;        // 2-bit words staged per count chunk (16000 bases < 2^16)
constexpr uint32_t kMapN = 4096;          // k-mers per count chunk
constexpr int kProbes = 32;
constexpr int kItemThreads = 256;         // count workgroup size (four per CU)
constexpr int kDescWords = 16;            // 64-byte item descriptor (k_item_prep)
constexpr uint32_t kDescRecords = ~0u;    // descriptor key count of an item that reads records
static_assert(kNG == T, "k_group keeps one group per thread");
static_assert(kGroupChunk % T == 0, "k_group chunk is whole records per thread");

template <typename K> struct KT;
template <> struct KT<uint64_t> {
    static constexpr int H = 2048;  // table slots: 2048 x (8 + 4) B = 24 KiB of LDS
    static __device__ __forceinline__ uint64_t empty() { return ~0ull; }
    static __device__ __forceinline__ uint32_t hash(uint64_t x) {
        x ^= x >> 31;
        x *= 0x9E3779B97F4A7C15ull;
        return (uint32_t)(x >> 40);
    }
};
template <> struct KT<u128> {
    static constexpr int H = 1024;  // 1024 x (16 + 4) B = 20 KiB
    static __device__ __forceinline__ u128 empty() { return ~(u128)0; }
    static __device__ __forceinline__ uint32_t hash(u128 v) {
        uint64_t x = (uint64_t)v ^ ((uint64_t)(v >> 64) * 0xC2B2AE3D27D4EB4Full);
        x ^= x >> 31;
        x *= 0x9E3779B97F4A7C15ull;
        return (uint32_t)(x >> 40);
    }
};
// distinct keys a sweep may hold (75% load)
template <typename K> constexpr int kCap = KT<K>::H / 4 * 3;

// Work-item sizing shared by the plan, the expansion and the count kernels.
__device__ __forceinline__ uint32_t plan_cap(const CountArgs& a, uint32_t bucket) {
    (void)bucket;
    return max(a.inst_cap, 1u);
}
// hash slices of a sub-bucket with more instances than one work item
__device__ __forceinline__ uint32_t heavy_q(uint32_t inst, uint32_t cap) { return min((inst + cap - 1) / cap, 127u); }
// key slots per slice region: the slice's share of instances with a margin (slices are hash
// slices, so a share above 1.25x + 256 is far in the tail); a fuller slice reads records instead
__device__ __forceinline__ uint32_t key_cs(uint32_t inst, uint32_t q) {
    return ((inst + q - 1) / q / 4 * 5 + 256 + 63) & ~63u;
}


// k-mer starting at base g of an LSB-first 2-bit word array: returns (forward, reverse complement).
// The forward key has the first base most significant: the digit-reversed window.
__device__ __forceinline__ void kmer_at(const uint32_t* bw, uint32_t g, int k, uint64_t& fwd, uint64_t& rc) {
    const uint32_t wi = g >> 4, sh = 2u * (g & 15u);
    uint64_t y = ((uint64_t)bw[wi] | ((uint64_t)bw[wi + 1] << 32)) >> sh;
    if (sh) y |= (uint64_t)bw[wi + 2] << (64 - sh);
    const uint64_t m = (k == 32) ? ~0ull : ((1ull << (2 * k)) - 1);
    y &= m;
    rc = (~y) & m;
    uint64_t r = __builtin_bitreverse64(y);
    r = ((r >> 1) & 0x5555555555555555ull) | ((r & 0x5555555555555555ull) << 1);
    fwd = r >> (64 - 2 * k);
}
__device__ __forceinline__ void kmer_at(const uint32_t* bw, uint32_t g, int k, u128& fwd, u128& rc) {
    const uint32_t wi = g >> 4, sh = 2u * (g & 15u);
    const u128 y = ((u128)bw[wi] | ((u128)bw[wi + 1] << 32) | ((u128)bw[wi + 2] << 64) | ((u128)bw[wi + 3] << 96));
    u128 z = y >> sh;
    if (sh) z |= (u128)bw[wi + 4] << (128 - sh);
    const u128 m = (k == 64) ? ~(u128)0 : (((u128)1 << (2 * k)) - 1);
    z &= m;
    rc = (~z) & m;
    uint64_t lo = __builtin_bitreverse64((uint64_t)z), hi = __builtin_bitreverse64((uint64_t)(z >> 64));
    lo = ((lo >> 1) & 0x5555555555555555ull) | ((lo & 0x5555555555555555ull) << 1);
    hi = ((hi >> 1) & 0x5555555555555555ull) | ((hi & 0x5555555555555555ull) << 1);
    const u128 r = ((u128)lo << 64) | hi;  // digit-reversed over 64 digits
    fwd = r >> (128 - 2 * k);
}

// Linear probing of at most kProbes slots; a key that finds neither itself nor an empty slot in
// that window flags the sweep for a split (*ovf = 1). Keys are never removed, so a slot holding a
// key keeps it for the whole sweep.
template <typename K>
__device__ __forceinline__ void insert(K* tk, uint32_t* tc, K key, uint32_t* ovf) {
    constexpr int H = KT<K>::H;
    uint32_t h = KT<K>::hash(key) & (H - 1);
    if constexpr (sizeof(K) == 8) {
        // Groups of four slots (32 bytes, two 16-byte LDS reads per probe): the group's slots are
        // tried in order, then the next group's. A slot read as empty is claimed with a CAS whose
        // old value decides (a key that took the slot meanwhile is compared like a read one), so
        // every lane inserting a key walks the same slot order and stops at the same slot: a key
        // is never held twice. Probing four slots per LDS round trip keeps the lanes of a wave
        // close to each other in probe steps.
        constexpr uint32_t NG = (uint32_t)H / 4u;
        uint32_t g = h >> 2;
        for (int probe = 0; probe < kProbes / 4; probe++) {
            unsigned long long* gs = (unsigned long long*)&tk[4u * g];
            const ulonglong2 q0 = ((const ulonglong2*)gs)[0];
            const ulonglong2 q1 = ((const ulonglong2*)gs)[1];
            const uint64_t s4[4] = {q0.x, q0.y, q1.x, q1.y};
#pragma unroll
            for (int i = 0; i < 4; i++) {
                uint64_t cur = s4[i];
                if (cur == ~0ull) cur = atomicCAS(&gs[i], ~0ull, (unsigned long long)key);  // claim (old value back)
                if (cur == ~0ull || cur == (uint64_t)key) {
                    atomicAdd(&tc[4u * g + (uint32_t)i], 1u);
                    return;
                }
            }
            g = (g + 1u) & (NG - 1u);
        }
        *ovf = 1u;
    } else {
        // slot = {lo, hi}; hi == ~0 empty, ~0-1 busy (being published); keys have hi < 2^62.
        // The winner of the claim publishes in straight-line code before anyone waits, and a full
        // fence keeps the compiler from sinking the publish below the wait: a lane that waits on a
        // slot claimed by a lane of its own wave must never run first.
        const uint64_t khi = (uint64_t)(key >> 64), klo = (uint64_t)key;
        for (int probe = 0; probe < kProbes; probe++) {
            unsigned long long* lo_p = (unsigned long long*)&tk[h];
            unsigned long long* hi_p = lo_p + 1;
            uint64_t chi = __hip_atomic_load(hi_p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
            bool won = false;
            if (chi == ~0ull) {
                const unsigned long long old = atomicCAS(hi_p, ~0ull, ~0ull - 1);
                won = (old == ~0ull);
                chi = won ? khi : old;
            }
            if (won) {
                __hip_atomic_store(lo_p, klo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_store(hi_p, khi, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            int spin = 0;
            while (chi == ~0ull - 1) {
                if (++spin > (1 << 16)) { *ovf = 2u; return; }  // reported as a failure, not retried
                chi = __hip_atomic_load(hi_p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            if (chi == khi &&
                (won || __hip_atomic_load(lo_p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == klo)) {
                atomicAdd(&tc[h], 1u);
                return;
            }
            h = (h + 1) & (H - 1);
        }
        *ovf = 1u;
    }
}

// Sweeps of one sub-bucket that do not fit the table are cut into hash slices: a key belongs to
// slice `slice` of `nslices` when slice_hash(key) % nslices == slice (independent of the table hash).
template <typename K> __device__ __forceinline__ uint32_t slice_hash(K key) {
    uint64_t x = (uint64_t)key ^ (uint64_t)(key >> 61) ^ (sizeof(K) > 8 ? (uint64_t)((u128)key >> 64) * 0x94D049BB133111EBull : 0ull);
    x ^= x >> 29;
    x *= 0xBF58476D1CE4E5B9ull;
    return (uint32_t)(x >> 32);
}

// The 2-bit digits of the three bases in a record byte (digit 0 = bits 5:4), LSB-first.
__device__ __forceinline__ uint32_t byte_digits(uint32_t b) { return ((b & 0x3u) << 4) | (b & 0xCu) | ((b >> 4) & 0x3u); }

// Re-encode a record (cpu_funcs.cpp:219-250 layout: base t at digit (bm3 + t) % 3 of byte
// (bm3 + t) / 3) as L 2-bit bases, LSB-first, in ceil(L / 16) words.
__device__ __forceinline__ void record_to_words(const uint8_t* rec, uint32_t bm3, uint32_t L, uint32_t* out) {
    const uint32_t nb = skm_nbytes(bm3, L);
    uint64_t acc = 0;
    uint32_t nbits = 0, wi = 0;
    for (uint32_t j = 0; j < nb; j++) {
        const uint32_t lo = (j == 0) ? bm3 : 0u;
        const int32_t rem = (int32_t)(bm3 + L) - 3 * (int32_t)j;
        const uint32_t hi = rem >= 3 ? 3u : (rem > 0 ? (uint32_t)rem : 0u);
        if (hi <= lo) continue;
        const uint32_t n = hi - lo;
        const uint32_t bits = (byte_digits(rec[j]) >> (2u * lo)) & ((1u << (2u * n)) - 1u);
        acc |= (uint64_t)bits << nbits;
        nbits += 2u * n;
        if (nbits >= 32u) {
            out[wi++] = (uint32_t)acc;
            acc >>= 32;
            nbits -= 32u;
        }
    }
    if (nbits) out[wi] = (uint32_t)acc;
}

__device__ __forceinline__ uint32_t rec_words(uint32_t km, int k) { return (km + (uint32_t)k - 1u + 15u) >> 4; }

struct Sweep {
    uint32_t s_lo, s_hi;     // sub-bucket range (multiples of kSubPerGroup unless inside one group)
    uint32_t slice, nslices; // hash slice of the keys (nslices == 1: all keys)
};

// ---------------------------------------------------------------------------------------------
// 1. Grouping: records are grouped (descriptor = word offset << 32 | side word, into `scratch`)
//    and their bases re-encoded as 2-bit words in group order (into `words` at the bucket's base).
//    Each bucket is cut into kGroupSegs record segments, one workgroup each, so a large bucket
//    spreads over many CUs: (a) per segment, records and words per group, the segment's bytes,
//    and the k-mer instances per sub-bucket (added into the bucket's histogram); (b) per bucket,
//    every (group, segment) run's record and word start, group-major; (c) per segment, the
//    records in chunks staged in LDS, each chunk counting-sorted by group.
constexpr uint32_t kGroupSegs = kGroupSegments;

struct SegRange {
    uint32_t lo, hi;
};
__device__ __forceinline__ SegRange group_seg(uint32_t nrec, uint32_t sg) {
    const uint32_t per = (nrec + kGroupSegs - 1) / kGroupSegs;
    const uint32_t lo = min(sg * per, nrec);
    return {lo, min(lo + per, nrec)};
}
__device__ __forceinline__ uint32_t bucket_nrec(const CountArgs& a, uint32_t bucket) {
    return (uint32_t)((a.b_rec_end ? a.b_rec_end[bucket] : a.b_rec[bucket + 1]) - a.b_rec[bucket]);
}

__global__ __launch_bounds__(T) void k_group_hist(CountArgs a) {
    __shared__ uint32_t hr[kNG], hw[kNG];
    __shared__ uint32_t sinst[kSubIds];
    __shared__ uint32_t tmp[T / 64 + 1];
    const uint32_t bucket = blockIdx.x / kGroupSegs, sg = blockIdx.x % kGroupSegs;
    const uint64_t rec0 = a.b_rec[bucket];
    const SegRange sr = group_seg(bucket_nrec(a, bucket), sg);
    const int k = a.k;
    for (int i = threadIdx.x; i < kNG; i += T) { hr[i] = 0; hw[i] = 0; }
    for (int i = threadIdx.x; i < kSubIds; i += T) sinst[i] = 0;
    __syncthreads();
    uint32_t bytes = 0;
    for (uint32_t r = sr.lo + threadIdx.x; r < sr.hi; r += T) {
        const uint32_t sd = a.side[rec0 + r];
        const uint32_t g = side_sub(sd) / kSubPerGroup;
        atomicAdd(&hr[g], 1u);
        atomicAdd(&hw[g], rec_words(side_kmers(sd), k));
        atomicAdd(&sinst[side_sub(sd)], side_kmers(sd));
        bytes += skm_nbytes(side_bm3(sd), side_kmers(sd) + (uint32_t)k - 1u);
    }
    uint32_t tb;
    (void)wg_exscan_u32(bytes, tmp, &tb);  // ends with a barrier
    uint32_t* const h = a.g_hist + ((size_t)bucket * kGroupSegs + sg) * 2 * kNG;
    for (int i = threadIdx.x; i < kNG; i += T) { h[i] = hr[i]; h[kNG + i] = hw[i]; }
    for (int i = threadIdx.x; i < kSubIds; i += T)
        if (sinst[i]) atomicAdd(&a.sub_sinst[(size_t)bucket * kSubIds + i], sinst[i]);
    if (threadIdx.x == 0) a.g_seg_bytes[(size_t)bucket * (kGroupSegs + 1) + sg] = tb;
}

// (b) one workgroup per bucket, one thread per group
__global__ __launch_bounds__(T) void k_group_offsets(CountArgs a) {
    __shared__ uint32_t tmp[T / 64 + 1];
    static_assert(kNG == T, "one group per thread");
    const uint32_t bucket = blockIdx.x, g = threadIdx.x;
    uint32_t* const h = a.g_hist + (size_t)bucket * kGroupSegs * 2 * kNG;
    uint32_t ar = 0, aw = 0;
    for (uint32_t sg = 0; sg < kGroupSegs; sg++) {
        uint32_t* const hs = h + (size_t)sg * 2 * kNG;
        const uint32_t r = hs[g], w = hs[kNG + g];
        hs[g] = ar;
        hs[kNG + g] = aw;
        ar += r;
        aw += w;
    }
    uint32_t tr, tw;
    const uint32_t er = wg_exscan_u32(ar, tmp, &tr);
    const uint32_t ew = wg_exscan_u32(aw, tmp, &tw);
    for (uint32_t sg = 0; sg < kGroupSegs; sg++) {
        uint32_t* const hs = h + (size_t)sg * 2 * kNG;
        hs[g] += er;
        hs[kNG + g] += ew;
    }
    uint32_t* g_start = a.sub_start + (size_t)bucket * (kNG + 1);
    g_start[g] = er;
    if (g == 0) {
        g_start[kNG] = tr;
        if (tw > a.b_word[bucket + 1] - a.b_word[bucket]) raise_err(a.err, RGKC_ERR_SCRATCH);
        uint64_t* const sb = a.g_seg_bytes + (size_t)bucket * (kGroupSegs + 1);
        uint64_t acc = 0;
        for (uint32_t sg = 0; sg <= kGroupSegs; sg++) {
            const uint64_t v = sg < kGroupSegs ? sb[sg] : 0;
            sb[sg] = acc;
            acc += v;
        }
    }
}

// (c) one workgroup per segment
__global__ __launch_bounds__(T) void k_group(CountArgs a) {
    __shared__ uint32_t gcur[kNG];     // record cursor of each group
    __shared__ uint32_t gwcur[kNG];    // word cursor of each group
    // the chunk's store bytes
    __shared__ union {
        uint8_t bytes[kGroupChunkBytes + 32];
        uint4 v[(kGroupChunkBytes + 32) / 16];  // 16-byte alignment for the staging loads
    } ub;
    __shared__ uint32_t ccnt[kNG], cst[kNG], cpos[kNG];
    __shared__ uint32_t snw[kGroupChunk + 1];  // words of the chunk's records in group order, then their prefix
    __shared__ uint32_t gsd[kGroupChunk];      // side word of the chunk's records in group order
    __shared__ uint16_t goff[kGroupChunk];     // and their byte offsets in the staged chunk
    __shared__ uint32_t tmp[T / 64 + 1];
    __shared__ uint32_t sh_m, sh_bytes;
    constexpr int RPT = kGroupChunk / T;
    static_assert(kGroupChunkBytes <= 65536, "staged byte offsets are 16-bit");
    const uint32_t bucket = blockIdx.x / kGroupSegs, sg = blockIdx.x % kGroupSegs;
    const uint64_t rec0 = a.b_rec[bucket];
    const SegRange sr = group_seg(bucket_nrec(a, bucket), sg);
    const uint32_t nrec = sr.hi;
    const uint64_t byte0 = a.b_byte[bucket];
    const uint64_t nbytes = (a.b_byte_end ? a.b_byte_end[bucket] : a.b_byte[bucket + 1]) - byte0;
    const uint64_t wcap = a.b_word[bucket + 1] - a.b_word[bucket];
    const int k = a.k;
    uint32_t* words = a.words + a.b_word[bucket];
    const uint8_t* store = a.store + byte0;
    {
        const uint32_t* const hs = a.g_hist + ((size_t)bucket * kGroupSegs + sg) * 2 * kNG;
        for (int i = threadIdx.x; i < kNG; i += T) { gcur[i] = hs[i]; gwcur[i] = hs[kNG + i]; }
    }
    const uint64_t* const sb = a.g_seg_bytes + (size_t)bucket * (kGroupSegs + 1);
    const uint64_t seg_end = sg + 1 < kGroupSegs ? sb[sg + 1] : sb[kGroupSegs];
    __syncthreads();
    uint64_t carry = sb[sg];  // store bytes before the current chunk
    for (uint32_t c0 = sr.lo; c0 < nrec;) {
        // the chunk: up to kGroupChunk records and kGroupChunkBytes bytes (at least one record)
        uint32_t sd[RPT], nb[RPT], nw[RPT], pos[RPT], accb = 0;
#pragma unroll
        for (int j = 0; j < RPT; j++) {
            const uint32_t r = c0 + threadIdx.x * RPT + j;
            sd[j] = r < nrec ? a.side[rec0 + r] : 0u;
            nb[j] = r < nrec ? skm_nbytes(side_bm3(sd[j]), side_kmers(sd[j]) + (uint32_t)k - 1u) : 0u;
            nw[j] = r < nrec ? rec_words(side_kmers(sd[j]), k) : 0u;
            accb += nb[j];
        }
        uint32_t totb;
        const uint32_t exb = wg_exscan_u32(accb, tmp, &totb);
        // records whose bytes end within the budget form the chunk
        if (threadIdx.x == 0) sh_m = 0;
        ccnt[threadIdx.x] = 0;
        __syncthreads();
        {
            uint32_t e = exb, in = 0;
#pragma unroll
            for (int j = 0; j < RPT; j++) {
                const uint32_t r = c0 + threadIdx.x * RPT + j;
                e += nb[j];
                if (r < nrec && (e <= kGroupChunkBytes || r == c0)) in = threadIdx.x * RPT + j + 1;
            }
            if (in) atomicMax(&sh_m, in);
        }
        __syncthreads();
        const uint32_t m = sh_m;
        uint32_t off[RPT];
        {
            uint32_t e = exb;
#pragma unroll
            for (int j = 0; j < RPT; j++) {
                off[j] = e;
                e += nb[j];
                if (threadIdx.x * RPT + j < m) atomicAdd(&ccnt[side_sub(sd[j]) / kSubPerGroup], 1u);
            }
        }
        // bytes of the chunk: the last record's end
        if (threadIdx.x == (m - 1) / RPT) {
            const int j = (int)((m - 1) % RPT);
            uint32_t e = 0;
#pragma unroll
            for (int q = 0; q < RPT; q++)
                if (q == j) e = off[q] + nb[q];
            sh_bytes = e;
        }
        __syncthreads();
        // the chunk's bytes into LDS with aligned 16-byte loads (a chunk of one record longer than
        // the staging area is read from HBM)
        const uint32_t cbytes = sh_bytes;
        const uint64_t gpos = byte0 + carry, abase = gpos & ~15ull;
        const uint32_t delta = (uint32_t)(gpos - abase);
        const bool staged = delta + cbytes <= kGroupChunkBytes + 16u;
        if (staged) {
            const uint32_t n16 = (delta + cbytes + 15u) / 16u;
            const uint4* src = (const uint4*)(a.store + abase);
            for (uint32_t i = threadIdx.x; i < n16; i += T) ub.v[i] = src[i];
        }
        {
            uint32_t t2;
            const uint32_t e = wg_exscan_u32(ccnt[threadIdx.x], tmp, &t2);
            cst[threadIdx.x] = e;
            cpos[threadIdx.x] = e;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < RPT; j++) {
            if (threadIdx.x * RPT + j < m) {
                pos[j] = atomicAdd(&cpos[side_sub(sd[j]) / kSubPerGroup], 1u);
                snw[pos[j]] = nw[j];
                gsd[pos[j]] = sd[j];
                goff[pos[j]] = (uint16_t)off[j];
            }
        }
        __syncthreads();
        {
            // word prefix over the chunk's records in group order
            uint32_t v[RPT], s = 0;
#pragma unroll
            for (int j = 0; j < RPT; j++) {
                const uint32_t i = threadIdx.x * RPT + j;
                v[j] = i < m ? snw[i] : 0u;
                s += v[j];
            }
            uint32_t tw;
            uint32_t e = wg_exscan_u32(s, tmp, &tw);
#pragma unroll
            for (int j = 0; j < RPT; j++) {
                const uint32_t i = threadIdx.x * RPT + j;
                if (i < m) snw[i] = e;
                e += v[j];
            }
            if (threadIdx.x == 0) snw[m] = tw;
        }
        __syncthreads();
        // records in group order: lane i writes the chunk's i-th record in the order of the
        // output, so neighbouring lanes write neighbouring words and descriptors
        for (uint32_t i = threadIdx.x; i < m; i += T) {
            const uint32_t sdi = gsd[i], offi = staged ? goff[i] : 0u;
            const uint32_t g = side_sub(sdi) / kSubPerGroup;
            const uint32_t slot = gcur[g] + (i - cst[g]);
            const uint32_t woff = gwcur[g] + (snw[i] - snw[cst[g]]);
            if ((uint64_t)woff + rec_words(side_kmers(sdi), k) <= wcap) {
                if (staged)
                    record_to_words(ub.bytes + delta + offi, side_bm3(sdi), side_kmers(sdi) + (uint32_t)k - 1u, words + woff);
                else
                    record_to_words(store + carry, side_bm3(sdi), side_kmers(sdi) + (uint32_t)k - 1u, words + woff);
            }
            a.scratch[rec0 - a.scratch_base + slot] = ((uint64_t)woff << 32) | (uint64_t)sdi;
        }
        __syncthreads();
        {
            const uint32_t g = threadIdx.x;
            const uint32_t s0 = cst[g], e0 = s0 + ccnt[g];
            gwcur[g] += snw[e0] - snw[s0];
            gcur[g] += ccnt[g];
        }
        __syncthreads();
        carry += cbytes;
        c0 += m;
    }
    if (threadIdx.x == 0 && (carry != seg_end || (sg + 1 == kGroupSegs && carry != nbytes)))
        raise_err(a.err, RGKC_ERR_BAD_STORE);
}

// 2. Per bucket: cut the sub-bucket sequence into work items of <= inst_cap k-mer instances. A
//    sub-bucket with more instances than that is cut into ceil(instances / inst_cap) hash slices,
//    one item each; with heavy_on its keys get one region per slice (1 + cs slots, the cursor
//    zeroed here) and its group an expansion task. FILL = false counts, true writes.
//    item = bucket << 40 | s_lo << 27 | s_hi << 14 | slice << 7 | nslices (s_hi = 0: keys)
template <bool FILL>
__global__ void k_plan(CountArgs a) {
    // the bucket's sub-bucket instance counts are staged in LDS by the whole workgroup; one thread
    // then walks them in order
    __shared__ uint32_t inst[kSubIds];
    const uint32_t bucket = blockIdx.x;
    {
        const uint32_t* const g = a.sub_sinst + (size_t)bucket * kSubIds;
        for (uint32_t i = threadIdx.x; i < (uint32_t)kSubIds; i += blockDim.x) inst[i] = g[i];
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    uint64_t n = 0, nkey = 0, nexp = 0;
    const uint64_t base = FILL ? a.item_cnt[bucket] : 0;
    const uint64_t kbase = (FILL && a.heavy_on) ? a.key_cnt[bucket] : 0;
    const uint64_t ebase = (FILL && a.heavy_on) ? a.exp_cnt[bucket] : 0;
    const uint32_t cap = plan_cap(a, bucket);
    uint32_t s = 0, last_g = ~0u;
    for (;;) {
        while (s < (uint32_t)kSubIds && inst[s] == 0) s++;
        if (s >= (uint32_t)kSubIds) break;
        if (inst[s] > cap) {
            const uint32_t q = heavy_q(inst[s], cap);
            if (a.heavy_on) {
                const uint32_t cs = key_cs(inst[s], q);
                if (FILL) {
                    a.sub_key_off[(size_t)bucket * kSubIds + s] = kbase + nkey;
                    a.sub_key_cs[(size_t)bucket * kSubIds + s] = cs;
                    a.sub_item0[(size_t)bucket * kSubIds + s] = (uint32_t)(base + n);
                    // zero each slice's cursor (the first 8 bytes of its slot 0)
                    uint64_t* const kw = (uint64_t*)a.heavy_keys;
                    const uint64_t W = (a.k <= 32) ? 1 : 2;  // uint64 words per key slot
                    for (uint32_t j = 0; j < q; j++) kw[(kbase + nkey + (uint64_t)j * (cs + 1)) * W] = 0;
                    if (s / kSubPerGroup != last_g) a.exp_items[ebase + nexp] = ((uint64_t)bucket << 32) | (s / kSubPerGroup);
                }
                if (s / kSubPerGroup != last_g) { nexp++; last_g = s / kSubPerGroup; }
                nkey += (uint64_t)q * (cs + 1);
            }
            const uint32_t shi = a.heavy_on ? 0u : s + 1;
            for (uint32_t j = 0; j < q; j++) {
                if (FILL)
                    a.items[base + n] = ((uint64_t)bucket << 40) | ((uint64_t)s << 27) | ((uint64_t)shi << 14) |
                                        ((uint64_t)j << 7) | (uint64_t)q;
                n++;
            }
            s++;
            continue;
        }
        uint32_t e = s, acc = 0;
        while (e < (uint32_t)kSubIds && acc + inst[e] <= cap) { acc += inst[e]; e++; }
        if (a.heavy_on) {
            // the item's sub-buckets share one region of exactly their instances after a slot that
            // holds their total; each sub-bucket owns a fixed sub-range of it (sub_key_off: its first
            // key slot), so only its group's expansion workgroup writes there
            if (FILL) {
                uint64_t o = kbase + nkey + 1;
                for (uint32_t t = s; t < e; t++) {
                    a.sub_key_off[(size_t)bucket * kSubIds + t] = o;
                    a.sub_key_cs[(size_t)bucket * kSubIds + t] = inst[t];
                    o += inst[t];
                }
                uint64_t* const kw = (uint64_t*)a.heavy_keys;
                kw[(kbase + nkey) * ((a.k <= 32) ? 1 : 2)] = acc;
            }
            for (uint32_t t = s; t < e; t++) {
                if (inst[t] == 0 || t / kSubPerGroup == last_g) continue;
                if (FILL) a.exp_items[ebase + nexp] = ((uint64_t)bucket << 32) | (t / kSubPerGroup);
                nexp++;
                last_g = t / kSubPerGroup;
            }
            nkey += (uint64_t)acc + 1;
        }
        if (FILL) a.items[base + n] = ((uint64_t)bucket << 40) | ((uint64_t)s << 27) | ((uint64_t)e << 14) | 1ull;
        n++;
        s = e;
    }
    if (!FILL) {
        a.item_cnt[bucket] = n;
        if (a.heavy_on) { a.key_cnt[bucket] = nkey; a.exp_cnt[bucket] = nexp; }
    }
}

// 2b. One workgroup per record group: every k-mer of the group is extracted once and its canonical
//     key goes to its (sub-bucket, slice) key run. Per chunk of records the keys are counted per
//     counter in LDS, each counter's run is reserved from its LDS cursor, and a second pass writes
//     the keys. Keys past a heavy slice's capacity are dropped while its cursor still counts them;
//     the slice then gets an exact region in the overflow pass.
constexpr int kExpSlots = kSubPerGroup * 128;  // (sub-bucket in group, slice) counters
constexpr int kExpThreads = 256;               // expansion workgroup size
constexpr uint32_t kExpMap = 16u * kExpThreads; // k-mers per expansion chunk

template <typename K, int TC, bool OVF>
__global__ __launch_bounds__(TC, 6) void k_expand(CountArgs a) {
    // chunks of at most TC records and 16 k-mers per thread; keys are held in registers in rounds
    // of 8 (64-bit) or 4 (128-bit) per thread, which keeps the kernel under 85 VGPRs without
    // spills, so six 256-thread workgroups share a CU and hide each other's chunk latencies
    constexpr uint32_t kExpW = kExpMap / 4, kMapE = kExpMap;
    __shared__ uint32_t stw[kExpW + 8];
    __shared__ uint16_t kpos[kMapE];
    __shared__ uint8_t ksub[kMapE];
    __shared__ uint32_t cnt[kExpSlots], gb[kExpSlots], lcur[kExpSlots];
    __shared__ uint32_t hq[kSubPerGroup], hcs[kSubPerGroup], hheavy[kSubPerGroup];
    __shared__ uint64_t hoff[kSubPerGroup], hm[kSubPerGroup];
    // the keys of each (sub-bucket, slice) counter: its first key slot (~0: none), capacity and
    // fill cursor. Only this workgroup writes them (a heavy sub-bucket's slices are its own regions,
    // a light sub-bucket owns a fixed part of its run's region), so the cursors live in LDS; a heavy
    // slice's final count goes to the slot before its keys at the end.
    __shared__ uint64_t roff[kExpSlots];
    __shared__ uint32_t rcap[kExpSlots];
    __shared__ uint32_t sh_any;
    __shared__ uint32_t tmp[TC / 64 + 1];
    __shared__ uint32_t sh_nk, sh_w0, sh_sb;
    __shared__ uint32_t tmp2[2 * (TC / 64)], wn[TC / 64], inclw[TC], inclk[TC];

    const uint64_t e = a.exp_items[blockIdx.x];
    const uint32_t bucket = (uint32_t)(e >> 32), g = (uint32_t)e;
    const uint64_t rec0 = a.b_rec[bucket];
    const uint32_t* g_start = a.sub_start + (size_t)bucket * (kNG + 1);
    const uint32_t* words = a.words + a.b_word[bucket];
    const int k = a.k;
    K* const reg = (K*)(OVF ? a.ovf_keys : a.heavy_keys);
    if (threadIdx.x < (uint32_t)kSubPerGroup) {
        // a heavy sub-bucket has q slices with a region each; a light one is the single slice of
        // the region its work item's sub-buckets share (hm = 0 maps every key to slice 0)
        const uint32_t sb = g * kSubPerGroup + threadIdx.x;
        const uint32_t inst = a.sub_sinst[(size_t)bucket * kSubIds + sb];
        const uint32_t cap = plan_cap(a, bucket);
        const bool heavy = inst > cap;
        hq[threadIdx.x] = heavy ? heavy_q(inst, cap) : (inst ? 1u : 0u);
        hcs[threadIdx.x] = inst ? a.sub_key_cs[(size_t)bucket * kSubIds + sb] : 0u;
        hoff[threadIdx.x] = inst ? a.sub_key_off[(size_t)bucket * kSubIds + sb] : 0ull;
        hm[threadIdx.x] = heavy ? ~0ull / hq[threadIdx.x] + 1ull : 0ull;
        hheavy[threadIdx.x] = heavy ? 1u : 0u;
    }
    if (threadIdx.x == 0) sh_any = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < (uint32_t)kExpSlots; i += TC) {
        const uint32_t t = i / 128u, j = i % 128u;
        uint64_t off = ~0ull;
        uint32_t cap = 0;
        if (j < hq[t]) {
            if (!OVF) {
                off = hheavy[t] ? hoff[t] + (uint64_t)j * (hcs[t] + 1) + 1 : hoff[t];
                cap = hcs[t];
            } else if (hheavy[t]) {
                // the overflow pass writes only the slices whose region overflowed, into exact regions
                const uint64_t itm = a.sub_item0[(size_t)bucket * kSubIds + g * kSubPerGroup + t] + j;
                const uint64_t n = a.ovf_n[itm];
                if (n) { off = a.ovf_off[itm] + 1; cap = (uint32_t)(n - 1); sh_any = 1; }
            }
        }
        roff[i] = off;
        rcap[i] = cap;
        lcur[i] = 0;
    }
    __syncthreads();
    if (OVF && !sh_any) return;
    // h % hq[t] for 32-bit h without a division: the high word of (hm[t] * h mod 2^64) * hq[t]
    // (Lemire, Kaser and Kurz, "Faster remainder by direct computation", 2019; exact for 32-bit
    // h and divisors, hm[t] = floor((2^64 - 1) / hq[t]) + 1)
    auto slice_of = [&](uint32_t t, uint32_t h) -> uint32_t { return (uint32_t)__umul64hi(hm[t] * h, hq[t]); };
    const uint32_t d_lo = g_start[g], d_hi = g_start[g + 1];
    // this thread's descriptor of the next chunk, loaded while the current chunk is expanded
    const uint64_t* const sdesc = a.scratch + (rec0 - a.scratch_base);
    uint64_t dnext = (d_lo + threadIdx.x < d_hi) ? sdesc[d_lo + threadIdx.x] : 0ull;
    uint32_t dnext_c0 = d_lo;
    for (uint32_t c0 = d_lo; c0 < d_hi;) {
        const uint32_t r = c0 + threadIdx.x;
        uint32_t nw = 0, km = 0, woff = 0, t = 0;
        if (r < d_hi) {
            const uint64_t d = dnext_c0 == c0 ? dnext : sdesc[r];
            t = side_sub((uint32_t)d) % kSubPerGroup;
            woff = (uint32_t)(d >> 32);
            nw = rec_words(side_kmers((uint32_t)d), k);
            if (hq[t]) km = side_kmers((uint32_t)d);
        }
        // the chunk: the leading records whose words and k-mers fit the staging areas. Both
        // inclusive prefixes grow with the thread index, so the records that fit are threads
        // [0, n): n is counted with one ballot per wave, and thread n-1's prefixes (the chunk's
        // words and k-mers) are read from LDS after the same barrier.
        uint32_t exW, exK, totW, totK;
        wg_exscan2_u32(nw, km, tmp2, exW, exK, totW, totK);
        const bool in = (r < d_hi) && (exW + nw <= kExpW) && (exK + km <= kMapE);
        const uint64_t bal = __ballot(in);
        if ((threadIdx.x & 63u) == 0) wn[threadIdx.x >> 6] = (uint32_t)__popcll(bal);
        inclw[threadIdx.x] = exW + nw;
        inclk[threadIdx.x] = exK + km;
        if (threadIdx.x == 0) { sh_w0 = woff; sh_nk = km; sh_sb = t; }
        for (uint32_t i = threadIdx.x; i < (uint32_t)kExpSlots; i += TC) cnt[i] = 0;
        __syncthreads();
        uint32_t n = 0;
#pragma unroll
        for (int w = 0; w < TC / 64; w++) n += wn[w];
        // the next chunk's descriptors (the records after this chunk), in flight from here on
        {
            const uint32_t c1 = c0 + max(n, 1u);
            dnext = (c1 + threadIdx.x < d_hi) ? sdesc[c1 + threadIdx.x] : 0ull;
            dnext_c0 = c1;
        }
        if (n == 0) {
            // the first record alone exceeds the staging area: its k-mers straight from HBM
            const uint32_t nk1 = sh_nk, t1 = sh_sb;
            for (uint32_t x = threadIdx.x; x < nk1; x += TC) {
                K f, rcv;
                kmer_at(words + sh_w0, x, k, f, rcv);
                const K key = f < rcv ? f : rcv;
                const uint32_t i = t1 * 128u + slice_of(t1, slice_hash<K>(key));
                if (roff[i] == ~0ull) continue;
                const uint32_t idx = atomicAdd(&lcur[i], 1u);
                if (idx < rcap[i]) reg[roff[i] + idx] = key;
            }
            __syncthreads();
            c0 += 1;
            continue;
        }
        const uint32_t nwc = inclw[n - 1], nkc = inclk[n - 1], w0 = sh_w0;
        for (uint32_t i = threadIdx.x; i < nwc + 8; i += TC) stw[i] = i < nwc ? words[w0 + i] : 0u;
        if (in)
            for (uint32_t j = 0; j < km; j++) { kpos[exK + j] = (uint16_t)(exW * 16u + j); ksub[exK + j] = (uint8_t)t; }
        __syncthreads();
        // one pass per round: each thread keeps its k-mers' keys, (sub-bucket, slice) counters and
        // ranks in registers while the runs are reserved (two rounds for 64-bit keys, four for
        // 128-bit)
        constexpr int KPT_ALL = kMapE / TC;
        constexpr int KPT = sizeof(K) == 8 ? KPT_ALL / 2 : KPT_ALL / 4;
        for (int rd = 0; rd < KPT_ALL / KPT; rd++) {
            if ((uint32_t)(rd * KPT * TC) >= nkc) break;  // no keys left for this round
            if (rd > 0) {
                for (uint32_t i = threadIdx.x; i < (uint32_t)kExpSlots; i += TC) cnt[i] = 0;
                __syncthreads();
            }
            K kk[KPT];
            uint32_t ti[KPT];
#pragma unroll
            for (int u = 0; u < KPT; u++) {
                const uint32_t x = threadIdx.x + (uint32_t)(rd * KPT + u) * TC;
                ti[u] = ~0u;
                if (x < nkc) {
                    K f, rcv;
                    kmer_at(stw, kpos[x], k, f, rcv);
                    kk[u] = f < rcv ? f : rcv;
                    const uint32_t tx = ksub[x], i = tx * 128u + slice_of(tx, slice_hash<K>(kk[u]));
                    if (!OVF || roff[i] != ~0ull)
                        ti[u] = i | (atomicAdd(&cnt[i], 1u) << 9);  // counter index (9 bits) | rank in the chunk
                }
            }
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < (uint32_t)kExpSlots; i += TC)
                if (cnt[i]) { gb[i] = lcur[i]; lcur[i] += cnt[i]; }  // counter i belongs to this thread
            __syncthreads();
#pragma unroll
            for (int u = 0; u < KPT; u++) {
                if (ti[u] == ~0u) continue;
                const uint32_t i = ti[u] & 511u;
                const uint32_t idx = gb[i] + (ti[u] >> 9);
                if (idx < rcap[i]) reg[roff[i] + idx] = kk[u];
            }
            __syncthreads();
        }
        c0 += n;
    }
    __syncthreads();
    // each heavy slice's key count (past its capacity too) into the slot before its keys
    for (uint32_t i = threadIdx.x; i < (uint32_t)kExpSlots; i += TC)
        if (roff[i] != ~0ull && hheavy[i / 128u]) *(unsigned long long*)&reg[roff[i] - 1] = (unsigned long long)lcur[i];
}

// 2b'. After expansion, one thread per work item: a key region whose slice received more keys than
//      its capacity (a slice holding a very frequent k-mer: its cursor kept counting) gets an exact
//      region in a second key area (ovf_n = keys + 1 slots, 0 for the others; its cursor zeroed
//      once the regions are placed); a second expansion pass fills only those regions.
template <typename K>
__global__ void k_ovf_sizes(CountArgs a) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > a.n_items) return;
    uint64_t need = 0;
    if (i < a.n_items) {
        const uint64_t item = a.items[i];
        const uint32_t bucket = (uint32_t)(item >> 40);
        const uint32_t s_lo = (uint32_t)((item >> 27) & 0x1FFFu), s_hi0 = (uint32_t)((item >> 14) & 0x1FFFu);
        const uint32_t slice = (uint32_t)((item >> 7) & 0x7Fu);
        if (s_hi0 == 0) {
            const uint32_t inst = a.sub_sinst[(size_t)bucket * kSubIds + s_lo];
            const uint32_t q0 = heavy_q(inst, plan_cap(a, bucket));
            const uint32_t cs = key_cs(inst, q0);
            const uint64_t off = a.sub_key_off[(size_t)bucket * kSubIds + s_lo] + (uint64_t)(slice % q0) * (cs + 1);
            const uint64_t n = *(const unsigned long long*)((const K*)a.heavy_keys + off);
            if (n > cs) need = n + 1;
        }
    }
    a.ovf_n[i] = need;
}
template <typename K>
__global__ void k_ovf_zero(CountArgs a) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.n_items && a.ovf_n[i]) *(unsigned long long*)((K*)a.ovf_keys + a.ovf_off[i]) = 0ull;
}

// 2c. Item descriptors, one thread per work item, once the heavy-key regions are filled: the
//     first sweep of the item with everything its first loads need (record range, key region and
//     its fill count, bucket bases), so the count kernel reads one 64-byte record per item and can
//     fetch the next item's record while it counts the current one.
template <typename K>
__global__ void k_item_prep(CountArgs a) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_items) return;
    const uint64_t item = a.items[i];
    const uint32_t bucket = (uint32_t)(item >> 40);
    const uint32_t s_lo = (uint32_t)((item >> 27) & 0x1FFFu), s_hi0 = (uint32_t)((item >> 14) & 0x1FFFu);
    const uint32_t slice = (uint32_t)((item >> 7) & 0x7Fu), nslices = (uint32_t)(item & 0x7Fu);
    const bool keyed = s_hi0 == 0;
    const uint32_t s_hi = keyed ? s_lo + 1 : s_hi0;
    uint32_t kn = kDescRecords, q0 = 0;
    uint64_t kreg = 0;
    if (!keyed && a.heavy_on) {
        // light sub-buckets: the item's exact region (never overflows), after its count slot
        const uint64_t off = a.sub_key_off[(size_t)bucket * kSubIds + s_lo] - 1;
        kn = (uint32_t)*(const unsigned long long*)((const K*)a.heavy_keys + off);
        kreg = (uint64_t)((const K*)a.heavy_keys + off + 1);
        q0 = 1;
    }
    if (keyed) {
        const uint32_t inst = a.sub_sinst[(size_t)bucket * kSubIds + s_lo];
        q0 = heavy_q(inst, plan_cap(a, bucket));
        const uint32_t cs = key_cs(inst, q0);
        const uint64_t off = a.sub_key_off[(size_t)bucket * kSubIds + s_lo] + (uint64_t)(slice % q0) * (cs + 1);
        const uint64_t n = *(const unsigned long long*)((const K*)a.heavy_keys + off);  // the slice's fill cursor
        if (n <= cs) {
            kn = (uint32_t)n;
            kreg = (uint64_t)((const K*)a.heavy_keys + off + 1);
        } else if (a.ovf_n && a.ovf_n[i] == n + 1) {
            kn = (uint32_t)n;  // the exact region of the overflow pass
            kreg = (uint64_t)((const K*)a.ovf_keys + a.ovf_off[i] + 1);
        }  // else: records
    }
    const uint32_t* g_start = a.sub_start + (size_t)bucket * (kNG + 1);
    const uint64_t rec0 = a.b_rec[bucket] - a.scratch_base, wb = a.b_word[bucket];  // rec0: scratch index
    uint32_t* d = a.item_desc + i * kDescWords;
    d[0] = bucket;
    d[1] = a.b_part[bucket];
    d[2] = s_lo | (s_hi0 << 16);
    d[3] = slice | (nslices << 16);
    d[4] = g_start[s_lo / kSubPerGroup];
    d[5] = g_start[(s_hi + kSubPerGroup - 1) / kSubPerGroup];
    d[6] = kn;
    d[7] = q0;
    d[8] = (uint32_t)kreg; d[9] = (uint32_t)(kreg >> 32);
    d[10] = (uint32_t)rec0; d[11] = (uint32_t)(rec0 >> 32);
    d[12] = (uint32_t)wb; d[13] = (uint32_t)(wb >> 32);
    d[14] = 0; d[15] = 0;
}

// 3. Persistent count kernel: workgroup b takes work items b, b + G, b + 2G, ... (G = grid size),
//    each as sweeps with split-on-overflow, finalize, append rows. The next item's descriptor is
//    fetched while the current one is counted, and a keyed sweep loads all its keys in one batch,
//    so a sweep waits on one HBM round trip for its keys and one for its row reservation.
template <typename K, int TC>
__global__ __launch_bounds__(TC, 4) void k_count_items(CountArgs a) {
    constexpr int H = KT<K>::H, CAP = kCap<K>;
    constexpr int KB = sizeof(K) == 8 ? 16 : 8;  // keys per thread per load batch of a keyed sweep
    __shared__ __align__(16) K tkeys[H];
    __shared__ uint32_t tcnt[H];
    // a chunk's 2-bit words and the position of each of its k-mers
    __shared__ uint32_t stw[kStageW + 8];
    __shared__ uint16_t kpos[kMapN];
    __shared__ uint32_t tmp[TC / 64 + 1];
    __shared__ Sweep stack[kStack];
    __shared__ uint32_t sh_ovf, sh_nstack, sh_go, sh_chunk_n, sh_inst_done, sh_nk, sh_nw;
    __shared__ uint32_t sh_w0;
    __shared__ Sweep cur;
    __shared__ uint64_t sh_row_base;
    __shared__ uint32_t sh_desc[kDescWords];
    __shared__ uint32_t tmp2[2 * (TC / 64)];

    const int k = a.k;
    // items b, b + G, b + 2G, ...: item costs vary by kind and partition, and striding spreads each
    // kind evenly over the workgroups (a contiguous share per workgroup left some with twice the mean)
    const uint64_t G = gridDim.x, it_end = a.n_items;
    uint64_t it = blockIdx.x;
    if (it >= it_end) return;
    // the workgroup's sweep, retry and distinct totals, added to the global counters once at the end
    unsigned long long n_sweeps = 0, n_retries = 0, n_distinct = 0;
    if (threadIdx.x < (uint32_t)kDescWords) sh_desc[threadIdx.x] = a.item_desc[it * kDescWords + threadIdx.x];
    __syncthreads();
    for (; it < it_end; it += G) {
        uint32_t dw[kDescWords];
#pragma unroll
        for (int i = 0; i < kDescWords; i++) dw[i] = sh_desc[i];
        // the next item's descriptor, in flight while this item is counted
        uint32_t nxt = 0;
        if (threadIdx.x < (uint32_t)kDescWords && it + G < it_end) nxt = a.item_desc[(it + G) * kDescWords + threadIdx.x];
        const uint32_t bucket = dw[0], part = dw[1];
        const uint64_t rec0 = (uint64_t)dw[10] | ((uint64_t)dw[11] << 32);
        const uint32_t* words = a.words + ((uint64_t)dw[12] | ((uint64_t)dw[13] << 32));
        const uint32_t* g_start = a.sub_start + (size_t)bucket * (kNG + 1);
        if (threadIdx.x == 0) {
            Sweep s0;
            s0.s_lo = dw[2] & 0xFFFFu;
            s0.s_hi = dw[2] >> 16;
            s0.slice = dw[3] & 0xFFFFu;
            s0.nslices = dw[3] >> 16;
            stack[0] = s0;
            sh_nstack = 1;
        }
        __syncthreads();
        bool first = true;
        for (;;) {
            if (threadIdx.x == 0) {
                sh_go = 0;
                if (sh_nstack > 0) { cur = stack[--sh_nstack]; sh_go = 1; }
                sh_ovf = 0;
                sh_inst_done = 0;
            }
            __syncthreads();
            if (!sh_go) break;
            // s_hi == 0: a hash slice of one heavy sub-bucket
            const bool keyed = cur.s_hi == 0;
            const uint32_t s_lo = cur.s_lo, s_hi = keyed ? cur.s_lo + 1 : cur.s_hi;
            const uint32_t slice = cur.slice, nslices = cur.nslices;
            // an item whose keys were expanded into a region (its descriptor) reads them in every
            // sweep (finer slices filter); otherwise its sweeps read records
            const bool from_keys = dw[6] != kDescRecords;
            const K* kreg = nullptr;
            uint32_t kn = 0, q0 = dw[7], d_lo, d_hi;
            if (from_keys) {
                kn = dw[6];
                kreg = (const K*)((uint64_t)dw[8] | ((uint64_t)dw[9] << 32));
            }
            if (first) {
                d_lo = dw[4];
                d_hi = dw[5];
            } else {
                d_lo = g_start[s_lo / kSubPerGroup];
                d_hi = g_start[(s_hi + kSubPerGroup - 1) / kSubPerGroup];
            }
            first = false;
            const bool all_subs = (s_lo % kSubPerGroup == 0) && (s_hi % kSubPerGroup == 0);
            // a region holds exactly its slice: only a finer slice filters its keys
            const bool filt = nslices != q0;

            if (from_keys) {
                // the slice's keys in batches of KB per thread: a batch's loads are issued before
                // the table is cleared, so their latency overlaps the clear and its barrier
                for (uint32_t x0 = 0; x0 < kn; x0 += (uint32_t)(KB * TC)) {
                    K v[KB];
#pragma unroll
                    for (int u = 0; u < KB; u++) {
                        const uint32_t x = x0 + threadIdx.x + (uint32_t)u * TC;
                        v[u] = x < kn ? kreg[x] : (K)0;
                    }
                    if (x0 == 0) {
                        for (int i = threadIdx.x; i < H; i += TC) { tkeys[i] = KT<K>::empty(); tcnt[i] = 0; }
                        __syncthreads();
                    }
#pragma unroll
                    for (int u = 0; u < KB; u++) {
                        const uint32_t x = x0 + threadIdx.x + (uint32_t)u * TC;
                        if (x < kn && (!filt || slice_hash<K>(v[u]) % nslices == slice))
                            insert<K>(tkeys, tcnt, v[u], &sh_ovf);
                    }
                    if (sh_ovf) break;
                }
                if (threadIdx.x == 0) sh_inst_done = kn;
                if (kn == 0) {
                    for (int i = threadIdx.x; i < H; i += TC) { tkeys[i] = KT<K>::empty(); tcnt[i] = 0; }
                }
            } else {
                for (int i = threadIdx.x; i < H; i += TC) { tkeys[i] = KT<K>::empty(); tcnt[i] = 0; }
            }
            __syncthreads();
            // chunks of up to TC records, contiguous in word space: their words are copied into LDS,
            // each k-mer gets its staged base position, then every thread takes whole k-mers
            for (uint32_t c0 = from_keys ? d_hi : d_lo; c0 < d_hi;) {
                const uint32_t r = c0 + threadIdx.x;
                uint32_t nw = 0, km = 0, woff = 0;
                if (r < d_hi) {
                    const uint64_t d = a.scratch[rec0 + r];
                    const uint32_t sb = side_sub((uint32_t)d);
                    woff = (uint32_t)(d >> 32);
                    nw = rec_words(side_kmers((uint32_t)d), k);
                    if (all_subs || (sb >= s_lo && sb < s_hi)) km = side_kmers((uint32_t)d);
                }
                uint32_t exW, exK, totW, totK;
                wg_exscan2_u32(nw, km, tmp2, exW, exK, totW, totK);
                const bool in = (r < d_hi) && (exW + nw <= kStageW) && (exK + km <= kMapN);
                if (threadIdx.x == 0) { sh_chunk_n = 0; sh_w0 = woff; }
                __syncthreads();
                if (in) atomicMax(&sh_chunk_n, threadIdx.x + 1u);
                __syncthreads();
                const uint32_t n = sh_chunk_n;
                if (n == 0) {
                    // the first record alone exceeds the staging area: its k-mers straight from HBM
                    if (threadIdx.x == 0) sh_nk = km;  // thread 0 holds the chunk's first record
                    __syncthreads();
                    const uint32_t nk1 = sh_nk;
                    for (uint32_t x = threadIdx.x; x < nk1 && !sh_ovf; x += TC) {
                        K f, rcv;
                        kmer_at(words + sh_w0, x, k, f, rcv);
                        const K key = f < rcv ? f : rcv;
                        if (nslices == 1u || slice_hash<K>(key) % nslices == slice) insert<K>(tkeys, tcnt, key, &sh_ovf);
                    }
                    if (threadIdx.x == 0) sh_inst_done += nk1;
                    __syncthreads();
                    c0 += 1;
                    if (sh_ovf) break;
                    continue;
                }
                if (threadIdx.x == n - 1) { sh_nw = exW + nw; sh_nk = exK + km; sh_inst_done += exK + km; }
                __syncthreads();
                const uint32_t nwc = sh_nw, nkc = sh_nk, w0 = sh_w0;
                for (uint32_t i = threadIdx.x; i < nwc + 8; i += TC) stw[i] = i < nwc ? words[w0 + i] : 0u;
                if (in)
                    for (uint32_t j = 0; j < km; j++) kpos[exK + j] = (uint16_t)(exW * 16u + j);
                __syncthreads();
                if (!sh_ovf) {
                    for (uint32_t x = threadIdx.x; x < nkc; x += TC) {
                        K f, rcv;
                        kmer_at(stw, kpos[x], k, f, rcv);
                        const K key = f < rcv ? f : rcv;
                        if (nslices == 1u || slice_hash<K>(key) % nslices == slice) insert<K>(tkeys, tcnt, key, &sh_ovf);
                    }
                }
                __syncthreads();
                c0 += n;
                if (sh_ovf) break;  // the sweep will be split: skip its remaining records
            }
            __syncthreads();
            // the sweep's distinct keys are its occupied slots; they go to registers for the finalize,
            // with the number of them each thread keeps after the [cmin, cmax] filter
            constexpr int SPT = H / TC;  // table slots per thread
            const K E = KT<K>::empty();
            K kr[SPT];
            uint32_t cr[SPT], occ = 0, keep = 0;
    #pragma unroll
            for (int j = 0; j < SPT; j++) {
                kr[j] = tkeys[threadIdx.x + j * TC];
                cr[j] = tcnt[threadIdx.x + j * TC];
                const bool o = kr[j] != E;
                occ += o ? 1u : 0u;
                keep += (o && cr[j] >= a.cmin && cr[j] <= a.cmax) ? 1u : 0u;
            }
            // one scan for both: occupied slots in the low half, kept rows in the high half (<= H each)
            uint32_t both;
            const uint32_t ex2 = wg_exscan_u32(occ | (keep << 16), tmp, &both);
            const uint32_t nd = both & 0xFFFFu, kept = both >> 16;
            if (sh_ovf || nd > (uint32_t)CAP) {
                if (threadIdx.x == 0) {
                    n_retries++;
                    // distinct keys of this sweep, estimated from the keys held when it stopped and the
                    // share of its instances read until then
                    uint32_t tot = 0;
                    for (uint32_t b = s_lo; b < s_hi; b++) tot += a.sub_sinst[(size_t)bucket * kSubIds + b];
                    // a keyed sweep's slices keep reading the region (they filter its keys); a region
                    // that overflowed is read as records
                    tot = max(tot / cur.nslices, 1u);
                    const float frac = fminf((float)max(sh_inst_done, 1u) / (float)max(tot, 1u), 1.0f);
                    uint32_t q = (uint32_t)ceilf(1.25f * (float)nd / (frac * (float)CAP));
                    q = max(2u, min(q, kMaxSlices));
                    // the non-empty part of the range: a piece never starts on an empty sub-bucket, so
                    // a range with two or more non-empty sub-buckets always yields at least two strictly
                    // smaller pieces (each below tot), and a single non-empty one is sliced
                    const uint32_t* si = a.sub_sinst + (size_t)bucket * kSubIds;
                    uint32_t lo = s_lo, hi = s_hi;
                    while (lo < hi && si[lo] == 0) lo++;
                    while (hi > lo && si[hi - 1] == 0) hi--;
                    if (sh_ovf >= 2u || hi == lo) {
                        raise_err(a.err, RGKC_ERR_COUNT_FAIL);
                    } else if (hi - lo > 1 && !from_keys) {
                        // cut the sub-bucket range into pieces of about tot/q instances
                        const uint32_t per = (tot + q - 1) / q;
                        uint32_t b = lo;
                        while (b < hi) {
                            uint32_t e = b + 1, acc = si[b];
                            while (e < hi && acc + si[e] <= per) acc += si[e++];
                            if (sh_nstack >= (uint32_t)kStack) { raise_err(a.err, RGKC_ERR_COUNT_FAIL); break; }
                            Sweep ns = cur;
                            ns.s_lo = b;
                            ns.s_hi = e;
                            stack[sh_nstack++] = ns;
                            b = e;
                            while (b < hi && si[b] == 0) b++;
                        }
                    } else if (cur.nslices * q > (1u << 20) || sh_nstack + q > (uint32_t)kStack) {
                        raise_err(a.err, RGKC_ERR_COUNT_FAIL);
                    } else {
                        // one sub-bucket, or a key region: q hash slices of this sweep's slice
                        for (uint32_t j = 0; j < q; j++) {
                            Sweep ns = cur;
                            if (!from_keys) {
                                ns.s_lo = lo;
                                ns.s_hi = lo + 1;
                            }
                            ns.slice = cur.slice + cur.nslices * j;
                            ns.nslices = cur.nslices * q;
                            stack[sh_nstack++] = ns;
                        }
                    }
                }
                __syncthreads();
                continue;
            }
            // ---- finalize: the kept rows go to the front of the table in slot order (every slot is in
            // registers, so the table can be overwritten), then to the partition's row region with
            // coalesced stores; one atomic per sweep reserves the rows. Row order inside a partition is
            // the order of the sweeps' appends (rgkc_sort_rows orders a partition by key).
            {
                uint32_t e = ex2 >> 16;
    #pragma unroll
                for (int j = 0; j < SPT; j++) {
                    if (kr[j] != E && cr[j] >= a.cmin && cr[j] <= a.cmax) {
                        tkeys[e] = kr[j];
                        tcnt[e] = cr[j];
                        e++;
                    }
                }
            }
            if (threadIdx.x == 0) {
                uint64_t base = 0;
                if (kept) base = atomicAdd(&a.row_cursor[part], (unsigned long long)kept);
                sh_row_base = a.row_base[part] + base;
                if (sh_row_base + kept > a.row_end[part]) raise_err(a.err, RGKC_ERR_ROW_OVERFLOW);
                n_sweeps++;
                n_distinct += nd;
                if (nd) atomicAdd(&a.part_distinct[part], (unsigned long long)nd);
            }
            __syncthreads();
            const uint64_t rb = sh_row_base;
            if (rb + kept <= a.row_end[part]) {
                for (uint32_t i = threadIdx.x; i < kept; i += TC) {
                    const uint64_t row = rb + i;
                    if constexpr (sizeof(K) == 8) {
                        ((uint64_t*)a.out_keys)[row] = (uint64_t)tkeys[i];
                    } else {
                        ((uint64_t*)a.out_keys)[2 * row] = (uint64_t)tkeys[i];
                        ((uint64_t*)a.out_keys)[2 * row + 1] = (uint64_t)((u128)tkeys[i] >> 64);
                    }
                    a.out_counts[row] = tcnt[i];
                }
            }
            __syncthreads();
        }
        // the next item's descriptor (all threads have read this one's)
        if (threadIdx.x < (uint32_t)kDescWords) sh_desc[threadIdx.x] = nxt;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(&a.stats[0], n_sweeps);
        if (n_retries) atomicAdd(&a.stats[1], n_retries);
        atomicAdd(&a.stats[2], n_distinct);
    }
}

}  // namespace

// sub_sinst must be zero (the segments add into it)
hipError_t launch_group(const CountArgs& a, hipStream_t s) {
    if (!a.n_buckets) return hipSuccess;
    hipLaunchKernelGGL(k_group_hist, dim3(a.n_buckets * kGroupSegs), dim3(T), 0, s, a);
    hipLaunchKernelGGL(k_group_offsets, dim3(a.n_buckets), dim3(T), 0, s, a);
    hipLaunchKernelGGL(k_group, dim3(a.n_buckets * kGroupSegs), dim3(T), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_plan(const CountArgs& a, bool fill, hipStream_t s) {
    if (!a.n_buckets) return hipSuccess;
    if (fill) hipLaunchKernelGGL(k_plan<true>, dim3(a.n_buckets), dim3(64), 0, s, a);
    else hipLaunchKernelGGL(k_plan<false>, dim3(a.n_buckets), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_expand(const CountArgs& a, uint64_t n_exp, hipStream_t s) {
    if (!n_exp || !a.heavy_on) return hipSuccess;
    if (a.k <= 32)
        hipLaunchKernelGGL((k_expand<uint64_t, kExpThreads, false>), dim3((unsigned)n_exp), dim3(kExpThreads), 0, s, a);
    else
        hipLaunchKernelGGL((k_expand<u128, kExpThreads, false>), dim3((unsigned)n_exp), dim3(kExpThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_ovf_sizes(const CountArgs& a, hipStream_t s) {
    const unsigned nb = (unsigned)((a.n_items + 1 + 255) / 256);
    if (a.k <= 32) hipLaunchKernelGGL(k_ovf_sizes<uint64_t>, dim3(nb), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_ovf_sizes<u128>, dim3(nb), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_expand_ovf(const CountArgs& a, uint64_t n_exp, hipStream_t s) {
    if (!n_exp || !a.heavy_on) return hipSuccess;
    const unsigned nb = (unsigned)((a.n_items + 255) / 256);
    if (a.k <= 32) {
        hipLaunchKernelGGL(k_ovf_zero<uint64_t>, dim3(nb), dim3(256), 0, s, a);
        hipLaunchKernelGGL((k_expand<uint64_t, kExpThreads, true>), dim3((unsigned)n_exp), dim3(kExpThreads), 0, s, a);
    } else {
        hipLaunchKernelGGL(k_ovf_zero<u128>, dim3(nb), dim3(256), 0, s, a);
        hipLaunchKernelGGL((k_expand<u128, kExpThreads, true>), dim3((unsigned)n_exp), dim3(kExpThreads), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_count_items(const CountArgs& a, hipStream_t s) {
    if (!a.n_items) return hipSuccess;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    // up to four resident workgroups per CU (the LDS table allows four), each looping over items
    const uint32_t per_cu = a.count_wg_per_cu ? std::min<uint32_t>(a.count_wg_per_cu, 4u) : 4u;
    const uint64_t grid = std::min<uint64_t>(a.n_items, (uint64_t)cus * per_cu);
    const unsigned pb = (unsigned)((a.n_items + 255) / 256);
    if (a.k <= 32) {
        hipLaunchKernelGGL(k_item_prep<uint64_t>, dim3(pb), dim3(256), 0, s, a);
        hipLaunchKernelGGL((k_count_items<uint64_t, kItemThreads>), dim3((unsigned)grid), dim3(kItemThreads), 0, s, a);
    } else {
        hipLaunchKernelGGL(k_item_prep<u128>, dim3(pb), dim3(256), 0, s, a);
        hipLaunchKernelGGL((k_count_items<u128, kItemThreads>), dim3((unsigned)grid), dim3(kItemThreads), 0, s, a);
    }
    return hipGetLastError();
}

}  // namespace rgkc

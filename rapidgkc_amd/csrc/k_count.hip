// Count stage on gfx950: exact (k-mer, count) tables from superkmer records, with no k-mer key
// ever written to HBM.
//
// Kernels:
//  1. k_desc_sort, one workgroup per bucket (a partition): the bucket's record descriptors are
//     grouped by minimizer sub-bucket (12 bits of ~mm above the partition digits), through an
//     LDS counting sort of 8192-record chunks so that each chunk writes one contiguous run per
//     sub-bucket. All k-mers of one record share its minimizer, and the minimizer is a function
//     of the canonical k-mer, so sub-buckets hold disjoint key sets.
//  2. k_plan: each bucket's sub-bucket sequence is cut into work items sized to the LDS table.
//  3. k_count_items, one 1024-thread workgroup per work item (a sweep). Each thread streams whole
//     superkmer records from the store: it decodes the bases in order, rolls the forward k-mer
//     and its reverse complement (cpu_funcs.cpp:427-436 / gpu_kmercounting.cu:63-74), and counts
//     each canonical key min(fwd, rc) in an open-addressing LDS table (64-bit CAS; 128-bit keys
//     publish the high word last). Nothing is staged per k-mer: no key is written to HBM and no
//     barrier is taken until the sweep's records are done.
//     A table overflow (more distinct keys than the sweep capacity) ends the sweep and splits
//     it: first by sub-bucket range, then by key range, until it fits.
//  4. The sweep's distinct keys are counting-sorted in LDS by their top 12 bits (relative to the
//     sweep's key range), each bin ordered by key rank; rows with min <= count <= max
//     (utilities.hpp:190) are appended to the partition's row region with one atomic per sweep.
//     Rows of one sweep are in ascending key order (cpu_funcs.cpp:614-640).
#include "rgkc_device.h"
#include "rgkc_internal.h"

namespace rgkc {
namespace {

constexpr int T = kCountThreads;
constexpr int kBins = 4096;
constexpr int kStack = 128;
constexpr int kStageWords = 8192;  // 32 KiB of record bytes staged per chunk
constexpr uint32_t kMaxSlices = 32;
constexpr int kPiece = 16;  // k-mers per piece of a record
#ifndef RGKC_ITEM_THREADS
constexpr int kItemThreads = 1024;  // count workgroup size
#else
constexpr int kItemThreads = RGKC_ITEM_THREADS;
#endif

template <typename K> struct KT;
template <> struct KT<uint64_t> {
    static constexpr int H = 8192;  // table slots: 8192 x (8 + 4) B = 96 KiB of LDS
    static __device__ __forceinline__ uint64_t empty() { return ~0ull; }
    static __device__ __forceinline__ uint32_t hash(uint64_t x) {
        x ^= x >> 31;
        x *= 0x9E3779B97F4A7C15ull;
        return (uint32_t)(x >> 40);
    }
};
template <> struct KT<u128> {
    static constexpr int H = 4096;  // 4096 x (16 + 4) B = 80 KiB
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

__device__ __forceinline__ int bitlen64(uint64_t x) { return x ? 64 - __clzll(x) : 0; }
__device__ __forceinline__ int bitlen(uint64_t x) { return bitlen64(x); }
__device__ __forceinline__ int bitlen(u128 x) {
    const uint64_t hi = (uint64_t)(x >> 64);
    return hi ? 64 + bitlen64(hi) : bitlen64((uint64_t)x);
}

// Linear probing of at most kProbes slots; a key that finds neither itself nor an empty slot in
// that window flags the sweep for a split (*ovf = 1). Keys are never removed, so a slot holding a
// key keeps it for the whole sweep.
constexpr int kProbes = 32;
template <typename K>
__device__ __forceinline__ void insert(K* tk, uint32_t* tc, K key, uint32_t* ovf) {
    constexpr int H = KT<K>::H;
    uint32_t h = KT<K>::hash(key) & (H - 1);
    if constexpr (sizeof(K) == 8) {
        for (int probe = 0; probe < kProbes; probe++) {
            unsigned long long* slot = (unsigned long long*)&tk[h];
            uint64_t cur = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (cur == ~0ull) cur = atomicCAS(slot, ~0ull, (unsigned long long)key);  // claim (old value back)
            if (cur == ~0ull || cur == (uint64_t)key) {
                atomicAdd(&tc[h], 1u);
                return;
            }
            h = (h + 1) & (H - 1);
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

// Insert up to kBatch keys at once: each probe round issues the loads (and claims) of all pending
// keys before waiting, so a wave keeps several LDS round trips in flight. Same probing rule as
// insert(): a key resolves at its own slot or at a slot it claims; kProbes rounds at most.
constexpr int kBatch = 4;
constexpr bool kUseBatch = false;
template <typename K>
__device__ __forceinline__ void insert_batch(K* tk, uint32_t* tc, const K (&kb)[kBatch], const bool (&vb)[kBatch],
                                             uint32_t* ovf) {
    constexpr int H = KT<K>::H;
    if constexpr (sizeof(K) == 8) {
        uint32_t h[kBatch];
        bool pend[kBatch];
        bool any = false;
#pragma unroll
        for (int j = 0; j < kBatch; j++) {
            pend[j] = vb[j];
            h[j] = KT<K>::hash(kb[j]) & (H - 1);
            any |= pend[j];
        }
        for (int probe = 0; probe < kProbes && any; probe++) {
            uint64_t cur[kBatch];
#pragma unroll
            for (int j = 0; j < kBatch; j++)
                cur[j] = pend[j] ? __hip_atomic_load((unsigned long long*)&tk[h[j]], __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP)
                                 : 0ull;
#pragma unroll
            for (int j = 0; j < kBatch; j++)
                if (pend[j] && cur[j] == ~0ull) cur[j] = atomicCAS((unsigned long long*)&tk[h[j]], ~0ull, (unsigned long long)kb[j]);
            any = false;
#pragma unroll
            for (int j = 0; j < kBatch; j++) {
                if (pend[j]) {
                    if (cur[j] == ~0ull || cur[j] == (uint64_t)kb[j]) {
                        atomicAdd(&tc[h[j]], 1u);
                        pend[j] = false;
                    } else {
                        h[j] = (h[j] + 1) & (H - 1);
                        any = true;
                    }
                }
            }
        }
        if (any) *ovf = 1u;
    } else {
#pragma unroll
        for (int j = 0; j < kBatch; j++)
            if (vb[j]) insert<K>(tk, tc, kb[j], ovf);
    }
}

// Count the k-mers of L consecutive bases of a superkmer record (cpu_funcs.cpp:437-515): bases are
// read in order from the record bytes (3 per byte, digit 0 = bits 5:4), rolling the forward k-mer
// (first base most significant) and its reverse complement (the complement of the newest base
// most significant, cpu_funcs.cpp:427-436); each canonical key min(fwd, rc) of the sweep's slice
// goes into the LDS table. The first base is digit `dg` of byte `bofs` of the little-endian word
// wp[0]. Eight words are kept in registers and refilled eight words ahead, so decoding never waits
// on a load (wp must stay readable for 8 words past the bases).
template <typename K>
__device__ __forceinline__ void count_piece(const uint32_t* wp, uint32_t bofs, uint32_t dg, uint32_t L, int k, K mask,
                                            uint32_t slice, uint32_t nslices, K* tk, uint32_t* tc, uint32_t* ovf) {
    uint32_t q0 = wp[0], q1 = wp[1], q2 = wp[2], q3 = wp[3], q4 = wp[4], q5 = wp[5], q6 = wp[6], q7 = wp[7];
    uint32_t wi = 0;
    uint32_t byte = (q0 >> (8u * bofs)) & 0xFFu;
    const int rsh = 2 * (k - 1);
    K fwd = 0, rc = 0;
    uint32_t t = 0;
    auto next_base = [&]() -> uint32_t {
        const uint32_t c = (byte >> (4u - 2u * dg)) & 3u;
        if (++dg == 3u) {
            dg = 0;
            if (++bofs == 4u) {
                bofs = 0;
                wi++;
                q0 = q1; q1 = q2; q2 = q3; q3 = q4; q4 = q5; q5 = q6; q6 = q7;
                q7 = wp[wi + 7];
            }
            byte = (q0 >> (8u * bofs)) & 0xFFu;
        }
        return c;
    };
    for (; t + 1u < (uint32_t)k; t++) {  // the first k-1 bases fill the window
        const uint32_t c = next_base();
        fwd = ((fwd << 2) | (K)c) & mask;
        rc = (rc >> 2) | ((K)(3u - c) << rsh);
    }
    while (t < L) {
        K kb[kBatch];
        bool vb[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; j++) {
            vb[j] = false;
            kb[j] = 0;
            if (t < L) {
                const uint32_t c = next_base();
                fwd = ((fwd << 2) | (K)c) & mask;
                rc = (rc >> 2) | ((K)(3u - c) << rsh);
                const K key = fwd < rc ? fwd : rc;
                kb[j] = key;
                vb[j] = (nslices == 1u || slice_hash<K>(key) % nslices == slice);
                t++;
            }
        }
        if constexpr (kUseBatch) {
            insert_batch<K>(tk, tc, kb, vb, ovf);
        } else {
#pragma unroll
            for (int j = 0; j < kBatch; j++)
                if (vb[j]) insert<K>(tk, tc, kb[j], ovf);
        }
    }
}

struct Sweep {
    uint32_t s_lo, s_hi;     // sub-bucket range
    uint32_t slice, nslices; // hash slice of the keys (nslices == 1: all keys)
};

// ---------------------------------------------------------------------------------------------
// 1. Per bucket: record descriptors (byte offset in the bucket << 32 | side word) grouped by
//    minimizer sub-bucket into `scratch`, plus per-sub-bucket starts and k-mer instances. The
//    scatter goes through an LDS counting sort of 8192-record chunks, so each chunk writes one
//    contiguous run per sub-bucket instead of isolated 8-byte stores.
constexpr int kNB = kSubIds;       // groups per bucket (one per sub-bucket)
constexpr int kGPT = kNB / T;      // groups per thread
constexpr int kCh = 8192;          // records per scatter chunk
__global__ __launch_bounds__(T) void k_desc_sort(CountArgs a) {
    __shared__ uint32_t gcur[kNB];      // global cursor of each group (bucket-relative)
    __shared__ uint32_t ginst[kNB];     // k-mer instances per group
    __shared__ uint32_t ccnt[kNB], cst[kNB], cst0[kNB];
    __shared__ uint64_t cbuf[kCh];
    __shared__ uint32_t tmp[T / 64 + 1];
    const uint32_t bucket = blockIdx.x;
    const uint64_t rec0 = a.b_rec[bucket];
    const uint32_t nrec = (uint32_t)((a.b_rec_end ? a.b_rec_end[bucket] : a.b_rec[bucket + 1]) - rec0);
    const uint64_t nbytes = (a.b_byte_end ? a.b_byte_end[bucket] : a.b_byte[bucket + 1]) - a.b_byte[bucket];
    const int k = a.k;
    uint32_t* g_start = a.sub_start + (size_t)bucket * (kNB + 1);
    uint32_t* g_inst = a.sub_inst + (size_t)bucket * kNB;
    for (int i = threadIdx.x; i < kNB; i += T) { gcur[i] = 0; ginst[i] = 0; }
    __syncthreads();
    for (uint32_t r = threadIdx.x; r < nrec; r += T) {
        const uint32_t sd = a.side[rec0 + r];
        atomicAdd(&ginst[side_sub(sd)], side_kmers(sd));
        atomicAdd(&gcur[side_sub(sd)], 1u);
    }
    __syncthreads();
    {
        uint32_t v[kGPT], tot, acc = 0;
#pragma unroll
        for (int j = 0; j < kGPT; j++) { v[j] = gcur[threadIdx.x * kGPT + j]; acc += v[j]; }
        uint32_t ex = wg_exscan_u32(acc, tmp, &tot);
#pragma unroll
        for (int j = 0; j < kGPT; j++) {
            const int g = threadIdx.x * kGPT + j;
            g_start[g] = ex;
            g_inst[g] = ginst[g];
            gcur[g] = ex;
            ex += v[j];
        }
        if (threadIdx.x == 0) g_start[kNB] = tot;
    }
    __syncthreads();
    uint64_t carry = 0;  // bytes before the current chunk
    for (uint32_t c0 = 0; c0 < nrec; c0 += kCh) {
        const uint32_t m = min((uint32_t)kCh, nrec - c0);
        for (int i = threadIdx.x; i < kNB; i += T) ccnt[i] = 0;
        __syncthreads();
        // each thread takes 8 consecutive records: byte offsets by a workgroup scan
        uint32_t sd[kCh / T], nbv[kCh / T], acc = 0;
#pragma unroll
        for (int j = 0; j < kCh / T; j++) {
            const uint32_t r = threadIdx.x * (kCh / T) + j;
            sd[j] = r < m ? a.side[rec0 + c0 + r] : 0u;
            nbv[j] = r < m ? skm_nbytes(side_bm3(sd[j]), side_kmers(sd[j]) + (uint32_t)k - 1u) : 0u;
            acc += nbv[j];
            if (r < m) atomicAdd(&ccnt[side_sub(sd[j])], 1u);
        }
        uint32_t tot;
        uint32_t off = wg_exscan_u32(acc, tmp, &tot);
        {
            uint32_t v[kGPT], t2, cacc = 0;
#pragma unroll
            for (int j = 0; j < kGPT; j++) { v[j] = ccnt[threadIdx.x * kGPT + j]; cacc += v[j]; }
            uint32_t e = wg_exscan_u32(cacc, tmp, &t2);
#pragma unroll
            for (int j = 0; j < kGPT; j++) {
                cst[threadIdx.x * kGPT + j] = e;
                cst0[threadIdx.x * kGPT + j] = e;
                e += v[j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kCh / T; j++) {
            const uint32_t r = threadIdx.x * (kCh / T) + j;
            if (r < m) {
                const uint32_t pos = atomicAdd(&cst[side_sub(sd[j])], 1u);
                cbuf[pos] = ((uint64_t)(carry + off) << 32) | (uint64_t)sd[j];
            }
            off += nbv[j];
        }
        __syncthreads();
        // write the chunk group by group: consecutive positions of one group are consecutive in scratch
        for (uint32_t j = threadIdx.x; j < m; j += T) {
            const uint64_t d = cbuf[j];
            const uint32_t g = side_sub((uint32_t)d);
            a.scratch[rec0 + gcur[g] + (j - cst0[g])] = d;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < kNB; i += T) gcur[i] += ccnt[i];
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0 && carry != nbytes) raise_err(a.err, RGKC_ERR_BAD_STORE);
}

// 2. Per bucket: cut the sub-bucket sequence into work items of <= inst_cap k-mer instances (a
//    single larger sub-bucket is an item alone). FILL = false counts, true writes.
template <bool FILL>
__global__ void k_plan(CountArgs a) {
    if (threadIdx.x != 0) return;
    const uint32_t bucket = blockIdx.x;
    const uint32_t* inst = a.sub_inst + (size_t)bucket * kNB;
    uint64_t n = 0;
    const uint64_t base = FILL ? a.item_cnt[bucket] : 0;
    const uint32_t cap = a.inst_caps ? a.inst_caps[bucket] : a.inst_cap;
    uint32_t s = 0;
    for (;;) {
        while (s < (uint32_t)kNB && inst[s] == 0) s++;
        if (s >= (uint32_t)kNB) break;
        uint32_t e = s, acc = 0;
        while (e < (uint32_t)kNB && (acc == 0 || acc + inst[e] <= cap)) { acc += inst[e]; e++; }
        if (FILL) a.items[base + n] = ((uint64_t)bucket << 32) | ((uint64_t)s << 16) | (uint64_t)e;
        n++;
        s = e;
    }
    if (!FILL) a.item_cnt[bucket] = n;
}

// 3. One workgroup per work item: sweeps with split-on-overflow, finalize, append rows.
template <typename K, int TC>
__global__ __launch_bounds__(TC) void k_count_items(CountArgs a) {
    constexpr int H = KT<K>::H, CAP = kCap<K>;
    __shared__ K tkeys[H];
    __shared__ uint32_t tcnt[H];
    // record bytes staged per chunk; the finalize's bins reuse the space
    __shared__ union {
        uint32_t sbuf[kStageWords + 8];
        uint32_t bins[kBins];
    } u;
    uint32_t* const sbuf = u.sbuf;
    uint32_t* const bins = u.bins;
    // per record of the chunk: first piece, staged word | byte << 16 | start%3 << 18, k-mers
    __shared__ union {
        struct {
            uint32_t pstart[TC + 1];
            uint32_t rword[TC];
            uint32_t rkm[TC];
        } pc;
    } u2;
    __shared__ uint32_t tmp[TC / 64 + 1];
    __shared__ Sweep stack[kStack];
    __shared__ uint32_t sh_ovf, sh_nstack, sh_go, sh_chunk_n, sh_inst_done;
    __shared__ Sweep cur;
    __shared__ uint64_t sh_row_base;

    const uint64_t item = a.items[blockIdx.x];
    const uint32_t bucket = (uint32_t)(item >> 32);
    const uint32_t part = a.b_part[bucket];
    const uint64_t byte0 = a.b_byte[bucket];
    const uint64_t rec0 = a.b_rec[bucket];
    const uint32_t* g_start = a.sub_start + (size_t)bucket * (kNB + 1);
    const uint32_t* g_inst = a.sub_inst + (size_t)bucket * kNB;
    const int k = a.k;
    const uint8_t* store = a.store + byte0;
    const K kmask = (2 * k >= (int)(8 * sizeof(K))) ? ~(K)0 : (((K)1 << (2 * k)) - 1);
    if (threadIdx.x == 0) {
        Sweep s0;
        s0.s_lo = (uint32_t)((item >> 16) & 0xFFFFu);
        s0.s_hi = (uint32_t)(item & 0xFFFFu);
        s0.slice = 0;
        s0.nslices = 1;
        stack[0] = s0;
        sh_nstack = 1;
    }
    __syncthreads();

    for (;;) {
        if (threadIdx.x == 0) {
            sh_go = 0;
            if (sh_nstack > 0) { cur = stack[--sh_nstack]; sh_go = 1; }
            sh_ovf = 0;
            sh_inst_done = 0;
        }
        for (int i = threadIdx.x; i < H; i += TC) { tkeys[i] = KT<K>::empty(); tcnt[i] = 0; }
        __syncthreads();
        if (!sh_go) break;
        const uint32_t d_lo = g_start[cur.s_lo], d_hi = g_start[cur.s_hi];
        const uint32_t slice = cur.slice, nslices = cur.nslices;

        // chunks of up to TC records: each thread gathers one record's bytes into LDS with
        // independent word loads (one memory latency per chunk); the records are then cut into
        // pieces of at most kPiece k-mers, spread over all threads, so one long superkmer does not
        // hold the workgroup
        for (uint32_t c0 = d_lo; c0 < d_hi;) {
            const uint32_t r = c0 + threadIdx.x;
            uint64_t d = 0;
            uint32_t wc = 0, km = 0, np = 0;
            const uint8_t* src = store;
            if (r < d_hi) {
                d = a.scratch[rec0 + r];
                km = side_kmers((uint32_t)d);
                src = store + (d >> 32);
                wc = ((uint32_t)((uintptr_t)src & 3u) + skm_nbytes(side_bm3((uint32_t)d), km + (uint32_t)k - 1u) + 3u) >> 2;
                np = (km + kPiece - 1) / kPiece;
            }
            uint32_t totW;
            const uint32_t exW = wg_exscan_u32(wc, tmp, &totW);
            uint32_t totK;
            const uint32_t exK = wg_exscan_u32(km, tmp, &totK);
            const bool in = (r < d_hi) && (exW + wc <= (uint32_t)kStageWords);
            if (threadIdx.x == 0) sh_chunk_n = 0;
            __syncthreads();
            if (in) atomicMax(&sh_chunk_n, threadIdx.x + 1u);
            __syncthreads();
            const uint32_t n = sh_chunk_n;
            if (n == 0) {
                // the first record alone is larger than the staging area: decode it from global memory
                if (threadIdx.x == 0) {
                    count_piece<K>((const uint32_t*)((uintptr_t)src & ~(uintptr_t)3), (uint32_t)((uintptr_t)src & 3u),
                                   side_bm3((uint32_t)d), km + (uint32_t)k - 1u, k, kmask, slice, nslices, tkeys, tcnt,
                                   &sh_ovf);
                    sh_inst_done += km;
                }
                __syncthreads();
                c0 += 1;
                if (sh_ovf) break;
                continue;
            }
            uint32_t totP;
            const uint32_t exP = wg_exscan_u32(in ? np : 0u, tmp, &totP);
            if (in) {
                const uint32_t* gw = (const uint32_t*)((uintptr_t)src & ~(uintptr_t)3);
                for (uint32_t w0 = 0; w0 < wc; w0 += 8) {
                    uint32_t v[8];
#pragma unroll
                    for (int j = 0; j < 8; j++) v[j] = (w0 + (uint32_t)j < wc) ? gw[w0 + j] : 0u;
#pragma unroll
                    for (int j = 0; j < 8; j++)
                        if (w0 + (uint32_t)j < wc) sbuf[exW + w0 + j] = v[j];
                }
                u2.pc.pstart[threadIdx.x] = exP;
                u2.pc.rword[threadIdx.x] = exW | ((uint32_t)((uintptr_t)src & 3u) << 16) | (side_bm3((uint32_t)d) << 18);
                u2.pc.rkm[threadIdx.x] = km;
            }
            if (threadIdx.x == n - 1) {
                u2.pc.pstart[n] = exP + np;
                sh_inst_done += exK + km;
            }
            __syncthreads();
            const uint32_t npieces = u2.pc.pstart[n];
            if (!sh_ovf) {
                for (uint32_t pc = threadIdx.x; pc < npieces; pc += TC) {
                    uint32_t lo = 0, hi = n;  // record of piece pc: last i with pstart[i] <= pc
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (u2.pc.pstart[mid] <= pc) lo = mid; else hi = mid;
                    }
                    const uint32_t rw = u2.pc.rword[lo], rkm = u2.pc.rkm[lo];
                    const uint32_t o = (pc - u2.pc.pstart[lo]) * (uint32_t)kPiece;  // first k-mer of the piece
                    const uint32_t cnt = min((uint32_t)kPiece, rkm - o);
                    const uint32_t q = (rw >> 18) + o;                            // digit index from the record's first byte
                    const uint32_t byte_off = ((rw >> 16) & 3u) + q / 3u;         // from the first staged word
                    count_piece<K>(sbuf + (rw & 0xFFFFu) + (byte_off >> 2), byte_off & 3u, q % 3u,
                                   cnt + (uint32_t)k - 1u, k, kmask, slice, nslices, tkeys, tcnt, &sh_ovf);
                }
            }
            __syncthreads();
            c0 += n;
            if (sh_ovf) break;  // the sweep will be split: skip its remaining records
        }
        __syncthreads();

        // the sweep's distinct keys are its occupied slots; they go to registers for the finalize
        constexpr int SPT = H / TC;  // table slots per thread
        const K E = KT<K>::empty();
        K kr[SPT];
        uint32_t cr[SPT], occ = 0;
#pragma unroll
        for (int j = 0; j < SPT; j++) {
            kr[j] = tkeys[threadIdx.x + j * TC];
            cr[j] = tcnt[threadIdx.x + j * TC];
            occ += (kr[j] != E) ? 1u : 0u;
        }
        uint32_t nd;
        (void)wg_exscan_u32(occ, tmp, &nd);
        if (sh_ovf || nd > (uint32_t)CAP) {
            if (threadIdx.x == 0) {
                atomicAdd(&a.stats[1], 1ull);
                // distinct keys of this sweep, estimated from the keys held when it stopped and the
                // share of its instances read until then
                uint32_t tot = 0;
                for (uint32_t b = cur.s_lo; b < cur.s_hi; b++) tot += g_inst[b];
                const float frac = fminf((float)max(sh_inst_done, 1u) / (float)max(tot, 1u), 1.0f);
                uint32_t q = (uint32_t)ceilf(1.25f * (float)nd / (frac * (float)CAP));
                q = max(2u, min(q, kMaxSlices));
                if (sh_ovf >= 2u) {
                    raise_err(a.err, RGKC_ERR_COUNT_FAIL);
                } else if (cur.s_hi - cur.s_lo > 1) {
                    // cut the sub-bucket range into up to q pieces of about tot/q instances
                    const uint32_t per = (tot + q - 1) / q;
                    uint32_t b = cur.s_lo;
                    while (b < cur.s_hi) {
                        uint32_t e = b, acc = 0;
                        while (e < cur.s_hi && (acc == 0 || acc + g_inst[e] <= per)) acc += g_inst[e++];
                        if (sh_nstack >= (uint32_t)kStack) { raise_err(a.err, RGKC_ERR_COUNT_FAIL); break; }
                        Sweep ns = cur;
                        ns.s_lo = b; ns.s_hi = e;
                        stack[sh_nstack++] = ns;
                        b = e;
                    }
                } else if (cur.nslices * q > (1u << 20) || sh_nstack + q > (uint32_t)kStack) {
                    raise_err(a.err, RGKC_ERR_COUNT_FAIL);
                } else {
                    // one sub-bucket: q hash slices of this sweep's slice
                    for (uint32_t j = 0; j < q; j++) {
                        Sweep ns = cur;
                        ns.slice = cur.slice + cur.nslices * j;
                        ns.nslices = cur.nslices * q;
                        stack[sh_nstack++] = ns;
                    }
                }
            }
            __syncthreads();
            continue;
        }
        // ---- finalize: counting sort of the distinct keys by their top 12 bits in range, in place:
        // every thread holds its table slots in registers, so the sorted keys can be written back
        // into the table's first nd slots once all slots have been read.
        const K key_lo = 0;
        const int bl = bitlen(kmask);
        const int shift = bl > 12 ? bl - 12 : 0;
        for (int i = threadIdx.x; i < kBins; i += TC) bins[i] = 0;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < SPT; j++)
            if (kr[j] != E) atomicAdd(&bins[(uint32_t)((kr[j] - key_lo) >> shift)], 1u);
        __syncthreads();
        {
            uint32_t v[kBins / TC], tot, acc = 0;
#pragma unroll
            for (int j = 0; j < kBins / TC; j++) { v[j] = bins[threadIdx.x * (kBins / TC) + j]; acc += v[j]; }
            uint32_t ex = wg_exscan_u32(acc, tmp, &tot);
#pragma unroll
            for (int j = 0; j < kBins / TC; j++) { bins[threadIdx.x * (kBins / TC) + j] = ex; ex += v[j]; }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < SPT; j++) {
            if (kr[j] != E) {
                const uint32_t pos = atomicAdd(&bins[(uint32_t)((kr[j] - key_lo) >> shift)], 1u);
                tkeys[pos] = kr[j];
                tcnt[pos] = cr[j];
            }
        }
        __syncthreads();
        // bins[b] now ends bin b. Order inside each bin by rank: every key counts the smaller keys
        // of its bin (keys are distinct). Keys of one sweep share a few minimizers, so up to ~1/w
        // of them start with the same minimizer prefix and fall into one bin; ranking keeps that
        // bin's cost linear per key instead of quadratic.
        {
            constexpr int RPT = (CAP + TC - 1) / TC;
            K kv[RPT];
            uint32_t cv[RPT], pv[RPT];
#pragma unroll
            for (int j = 0; j < RPT; j++) {
                const uint32_t i = threadIdx.x + (uint32_t)j * TC;
                pv[j] = ~0u;
                if (i < nd) {
                    const K key = tkeys[i];
                    const uint32_t b = (uint32_t)((key - key_lo) >> shift);
                    const uint32_t bs = b == 0 ? 0u : bins[b - 1], be = bins[b];
                    uint32_t r = 0;
                    for (uint32_t t = bs; t < be; t++) r += (tkeys[t] < key) ? 1u : 0u;
                    kv[j] = key;
                    cv[j] = tcnt[i];
                    pv[j] = bs + r;
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < RPT; j++) {
                if (pv[j] != ~0u) {
                    tkeys[pv[j]] = kv[j];
                    tcnt[pv[j]] = cv[j];
                }
            }
        }
        __syncthreads();
        // filter [cmin, cmax] and append rows
        constexpr int PER = (CAP + TC - 1) / TC;
        uint32_t keep = 0;
        const uint32_t i0 = threadIdx.x * PER;
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const uint32_t i = i0 + j;
            if (i < nd) { const uint32_t c = tcnt[i]; keep += (c >= a.cmin && c <= a.cmax); }
        }
        uint32_t kept;
        uint32_t ex = wg_exscan_u32(keep, tmp, &kept);
        if (threadIdx.x == 0) {
            uint64_t base = 0;
            if (kept) base = atomicAdd(&a.row_cursor[part], (unsigned long long)kept);
            sh_row_base = a.row_base[part] + base;
            if (sh_row_base + kept > a.row_end[part]) raise_err(a.err, RGKC_ERR_ROW_OVERFLOW);
            atomicAdd(&a.stats[0], 1ull);
            atomicAdd(&a.stats[2], (unsigned long long)nd);
        }
        __syncthreads();
        const uint64_t rb = sh_row_base;
        const bool room = rb + kept <= a.row_end[part];
        if (room) {
#pragma unroll
            for (int j = 0; j < PER; j++) {
                const uint32_t i = i0 + j;
                if (i < nd) {
                    const uint32_t c = tcnt[i];
                    if (c >= a.cmin && c <= a.cmax) {
                        const uint64_t row = rb + ex;
                        if constexpr (sizeof(K) == 8) {
                            ((uint64_t*)a.out_keys)[row] = (uint64_t)tkeys[i];
                        } else {
                            ((uint64_t*)a.out_keys)[2 * row] = (uint64_t)tkeys[i];
                            ((uint64_t*)a.out_keys)[2 * row + 1] = (uint64_t)((u128)tkeys[i] >> 64);
                        }
                        a.out_counts[row] = c;
                        ex++;
                    }
                }
            }
        }
        __syncthreads();
    }
}

}  // namespace

hipError_t launch_desc_sort(const CountArgs& a, hipStream_t s) {
    if (!a.n_buckets) return hipSuccess;
    hipLaunchKernelGGL(k_desc_sort, dim3(a.n_buckets), dim3(T), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_plan(const CountArgs& a, bool fill, hipStream_t s) {
    if (!a.n_buckets) return hipSuccess;
    if (fill) hipLaunchKernelGGL(k_plan<true>, dim3(a.n_buckets), dim3(64), 0, s, a);
    else hipLaunchKernelGGL(k_plan<false>, dim3(a.n_buckets), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_count_items(const CountArgs& a, hipStream_t s) {
    if (!a.n_items) return hipSuccess;
    if (a.k <= 32)
        hipLaunchKernelGGL((k_count_items<uint64_t, kItemThreads>), dim3((unsigned)a.n_items), dim3(kItemThreads), 0, s, a);
    else
        hipLaunchKernelGGL((k_count_items<u128, kItemThreads>), dim3((unsigned)a.n_items), dim3(kItemThreads), 0, s, a);
    return hipGetLastError();
}

}  // namespace rgkc

// Count stage on gfx950: exact (k-mer, count) tables from superkmer records, with no k-mer key
// ever written to HBM.
//
// Kernels:
//  1. k_desc_sort, one workgroup per bucket (a partition): the bucket's record descriptors are
//     grouped by bin (the top 10 bits of the 12-bit minimizer sub-bucket id taken from ~mm above
//     the partition digits), through an LDS counting sort of 8192-record chunks so that each
//     chunk writes one contiguous run per bin. All k-mers of one record share its minimizer, and
//     the minimizer is a function of the canonical k-mer, so sub-buckets hold disjoint key sets.
//  2. k_plan: each bucket's bin sequence is cut into work items sized to the LDS table.
//  3. k_count_items, one 1024-thread workgroup per work item (a sweep). A sweep unpacks its
//     records into a 2-bit LDS array (LSB-first), then every thread takes k-mers: the forward key
//     is the digit-reversed window, the reverse complement is the complemented window
//     (cpu_funcs.cpp:427-436 / gpu_kmercounting.cu:63-74), the canonical key is their minimum,
//     and it is counted in an open-addressing LDS table (64-bit CAS; 128-bit keys publish the
//     high word last).
//     A table overflow (more distinct keys than the sweep capacity) splits the sweep: first at
//     bin boundaries, then inside one bin by sub-bucket (records of the other sub-buckets are
//     skipped), then by key range, until it fits.
//  4. The sweep's distinct keys are counting-sorted in LDS by their top 12 bits (relative to the
//     sweep's key range), small bins finished by insertion sort; rows with min <= count <= max
//     (utilities.hpp:190) are appended to the partition's row region with one atomic per sweep.
//     Rows of one sweep are in ascending key order (cpu_funcs.cpp:614-640).
#include "rgkc_device.h"
#include "rgkc_internal.h"

namespace rgkc {
namespace {

constexpr int T = kCountThreads;
constexpr int kBins = 4096;
constexpr int kChunkRec = 1024;
constexpr int kChunkBaseWords = 3072;  // 49152 bases
constexpr int kStack = 256;

template <typename K> struct KT;
template <> struct KT<uint64_t> {
    static constexpr int H = 4096, CAP = 3072;
    static __device__ __forceinline__ uint64_t empty() { return ~0ull; }
    static __device__ __forceinline__ uint32_t hash(uint64_t x) {
        x ^= x >> 31;
        x *= 0x9E3779B97F4A7C15ull;
        return (uint32_t)(x >> 40);
    }
};
template <> struct KT<u128> {
    static constexpr int H = 2048, CAP = 1536;
    static __device__ __forceinline__ u128 empty() { return ~(u128)0; }
    static __device__ __forceinline__ uint32_t hash(u128 v) {
        uint64_t x = (uint64_t)v ^ ((uint64_t)(v >> 64) * 0xC2B2AE3D27D4EB4Full);
        x ^= x >> 31;
        x *= 0x9E3779B97F4A7C15ull;
        return (uint32_t)(x >> 40);
    }
};

__device__ __forceinline__ int bitlen64(uint64_t x) { return x ? 64 - __clzll(x) : 0; }
__device__ __forceinline__ int bitlen(uint64_t x) { return bitlen64(x); }
__device__ __forceinline__ int bitlen(u128 x) {
    const uint64_t hi = (uint64_t)(x >> 64);
    return hi ? 64 + bitlen64(hi) : bitlen64((uint64_t)x);
}

// k-mer starting at base g of an LSB-first 2-bit array: returns (forward, reverse complement).
__device__ __forceinline__ void kmer_at(const uint32_t* bw, uint32_t g, int k, uint64_t& fwd, uint64_t& rc) {
    const uint32_t wi = g >> 4, sh = 2u * (g & 15u);
    const uint64_t lo = (uint64_t)bw[wi] | ((uint64_t)bw[wi + 1] << 32);
    uint64_t y = lo >> sh;
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
    uint64_t a = __builtin_bitreverse64((uint64_t)z), b = __builtin_bitreverse64((uint64_t)(z >> 64));
    a = ((a >> 1) & 0x5555555555555555ull) | ((a & 0x5555555555555555ull) << 1);
    b = ((b >> 1) & 0x5555555555555555ull) | ((b & 0x5555555555555555ull) << 1);
    const u128 r = ((u128)a << 64) | b;  // digit-reversed over 64 digits
    fwd = r >> (128 - 2 * k);
}

// Base t of a record whose first base sits at digit bm3 of its first byte (digit 0 = bits 5:4).
__device__ __forceinline__ uint32_t rec_base(const uint8_t* rec, uint32_t bm3, uint32_t t) {
    const uint32_t q = bm3 + t;
    return ((uint32_t)rec[q / 3u] >> (4u - 2u * (q % 3u))) & 3u;
}

template <typename K>
__device__ __forceinline__ bool insert(K* tk, uint32_t* tc, K key, uint32_t* n_occ, uint32_t* ovf) {
    constexpr int H = KT<K>::H;
    uint32_t h = KT<K>::hash(key) & (H - 1);
    const K E = KT<K>::empty();
    for (int probe = 0; probe < H; probe++) {
        if constexpr (sizeof(K) == 8) {
            unsigned long long* slot = (unsigned long long*)&tk[h];
            const uint64_t cur = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (cur == (uint64_t)key) { atomicAdd(&tc[h], 1u); return true; }
            if (cur == (uint64_t)E) {
                const unsigned long long old = atomicCAS(slot, (unsigned long long)E, (unsigned long long)key);
                if (old == (unsigned long long)E) {
                    if (atomicAdd(n_occ, 1u) + 1u > (uint32_t)KT<K>::CAP) *ovf = 1u;
                    atomicAdd(&tc[h], 1u);
                    return true;
                }
                if (old == (unsigned long long)key) { atomicAdd(&tc[h], 1u); return true; }
            }
        } else {
            // slot = {lo, hi}; hi == ~0 empty, ~0-1 busy (being published); keys have hi < 2^62.
            // The winner of the claim publishes in straight-line code before anyone waits, and a
            // full fence keeps the compiler from sinking the publish below the wait: a lane that
            // waits on a slot claimed by a lane of its own wave must never run first.
            unsigned long long* lo_p = (unsigned long long*)&tk[h];
            unsigned long long* hi_p = lo_p + 1;
            const uint64_t khi = (uint64_t)(key >> 64), klo = (uint64_t)key;
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
                if (++spin > (1 << 16)) { *ovf = 2u; return false; }  // reported as a failure, not retried
                chi = __hip_atomic_load(hi_p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            bool hit = false;
            if (chi == khi) {
                const uint64_t clo = won ? klo : __hip_atomic_load(lo_p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                hit = (clo == klo);
            }
            if (hit) {
                if (won && atomicAdd(n_occ, 1u) + 1u > (uint32_t)KT<K>::CAP) *ovf = 1u;
                atomicAdd(&tc[h], 1u);
                return true;
            }
        }
        h = (h + 1) & (H - 1);
    }
    *ovf = 1u;
    return false;
}

struct Sweep {
    uint32_t s_lo, s_hi;
    uint64_t klo_lo, klo_hi;  // key range low bound (lo, hi words)
    uint64_t khi_lo, khi_hi;  // key range high bound (inclusive)
};

template <typename K> __device__ __forceinline__ K mk(uint64_t lo, uint64_t hi) {
    if constexpr (sizeof(K) == 8) return (K)lo; else return ((K)hi << 64) | lo;
}
template <typename K> __device__ __forceinline__ uint64_t klo(K x) { return (uint64_t)x; }
template <typename K> __device__ __forceinline__ uint64_t khi(K x) {
    if constexpr (sizeof(K) == 8) return 0; else return (uint64_t)(x >> 64);
}

// ---------------------------------------------------------------------------------------------
// 1. Per bucket: record descriptors (byte offset in the bucket << 32 | side word) grouped by bin
//    (bin = sub-bucket >> 2, 1024 bins) into `scratch`, plus per-bin starts and k-mer instances.
//    The scatter goes through an LDS counting sort of 8192-record chunks, so each chunk writes
//    one contiguous run per bin instead of isolated 8-byte stores.
constexpr int kNB = 1024;          // bins per bucket
constexpr int kCh = 8192;          // records per scatter chunk
__global__ __launch_bounds__(T) void k_desc_sort(CountArgs a) {
    __shared__ uint32_t gcur[kNB];      // global cursor of each bin (bucket-relative)
    __shared__ uint32_t binst[kNB];     // instances per bin
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
    for (int i = threadIdx.x; i < kNB; i += T) { gcur[i] = 0; binst[i] = 0; }
    __syncthreads();
    for (uint32_t r = threadIdx.x; r < nrec; r += T) {
        const uint32_t sd = a.side[rec0 + r];
        const uint32_t b = side_sub(sd) >> 2;
        atomicAdd(&binst[b], side_kmers(sd));
        atomicAdd(&gcur[b], 1u);
    }
    __syncthreads();
    {
        uint32_t tot;
        const uint32_t v = gcur[threadIdx.x];
        const uint32_t ex = wg_exscan_u32(v, tmp, &tot);
        g_start[threadIdx.x] = ex;
        g_inst[threadIdx.x] = binst[threadIdx.x];
        gcur[threadIdx.x] = ex;
        if (threadIdx.x == 0) g_start[kNB] = tot;
    }
    __syncthreads();
    uint64_t carry = 0;  // bytes before the current chunk
    for (uint32_t c0 = 0; c0 < nrec; c0 += kCh) {
        const uint32_t m = min((uint32_t)kCh, nrec - c0);
        ccnt[threadIdx.x] = 0;
        __syncthreads();
        // each thread takes 8 consecutive records: byte offsets by a workgroup scan
        uint32_t sd[kCh / T], nbv[kCh / T], acc = 0;
#pragma unroll
        for (int j = 0; j < kCh / T; j++) {
            const uint32_t r = threadIdx.x * (kCh / T) + j;
            sd[j] = r < m ? a.side[rec0 + c0 + r] : 0u;
            nbv[j] = r < m ? skm_nbytes(side_bm3(sd[j]), side_kmers(sd[j]) + (uint32_t)k - 1u) : 0u;
            acc += nbv[j];
            if (r < m) atomicAdd(&ccnt[side_sub(sd[j]) >> 2], 1u);
        }
        uint32_t tot;
        uint32_t off = wg_exscan_u32(acc, tmp, &tot);
        {
            uint32_t t2;
            const uint32_t e = wg_exscan_u32(ccnt[threadIdx.x], tmp, &t2);
            cst[threadIdx.x] = e;
            cst0[threadIdx.x] = e;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kCh / T; j++) {
            const uint32_t r = threadIdx.x * (kCh / T) + j;
            if (r < m) {
                const uint32_t pos = atomicAdd(&cst[side_sub(sd[j]) >> 2], 1u);
                cbuf[pos] = ((uint64_t)(carry + off) << 32) | (uint64_t)sd[j];
            }
            off += nbv[j];
        }
        __syncthreads();
        // write the chunk bin by bin: consecutive positions of one bin are consecutive in scratch
        for (uint32_t j = threadIdx.x; j < m; j += T) {
            const uint64_t d = cbuf[j];
            const uint32_t b = side_sub((uint32_t)d) >> 2;
            a.scratch[rec0 + gcur[b] + (j - cst0[b])] = d;
        }
        __syncthreads();
        gcur[threadIdx.x] += ccnt[threadIdx.x];
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0 && carry != nbytes) raise_err(a.err, RGKC_ERR_BAD_STORE);
}

// 2. Per bucket: cut the bin sequence into work items of <= inst_cap k-mer instances (a single
//    larger bin is an item alone). FILL = false counts, true writes. Items carry sub-bucket
//    ranges (4 sub-buckets per bin).
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
        if (FILL) a.items[base + n] = ((uint64_t)bucket << 32) | ((uint64_t)(4 * s) << 16) | (uint64_t)(4 * e);
        n++;
        s = e;
    }
    if (!FILL) a.item_cnt[bucket] = n;
}

// 3. One workgroup per work item: sweeps with split-on-overflow, finalize, append rows.
template <typename K>
__global__ __launch_bounds__(T) void k_count_items(CountArgs a) {
    constexpr int H = KT<K>::H, CAP = KT<K>::CAP;
    __shared__ K tkeys[H];
    __shared__ uint32_t tcnt[H];
    struct Ins {
        uint64_t desc[kChunkRec];
        uint32_t kpre[kChunkRec + 1];
        uint32_t bpre[kChunkRec + 1];
        uint32_t bases[kChunkBaseWords + 6];
    };
    struct Fin {
        K skeys[CAP];
        uint32_t scnt[CAP];
    };
    __shared__ union U {
        Ins ins;
        Fin fin;
    } u;
    __shared__ uint32_t bins[kBins];
    __shared__ uint32_t tmp[T / 64 + 1];
    __shared__ Sweep stack[kStack];
    __shared__ uint32_t sh_n_occ, sh_ovf, sh_nstack, sh_go;
    __shared__ Sweep cur;
    __shared__ uint32_t sh_chunk_n, sh_long_rec;
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
    const K kfull_hi = (2 * k >= (int)(8 * sizeof(K))) ? ~(K)0 : (((K)1 << (2 * k)) - 1);
    if (threadIdx.x == 0) {
        Sweep s0;
        s0.s_lo = (uint32_t)((item >> 16) & 0xFFFFu);
        s0.s_hi = (uint32_t)(item & 0xFFFFu);
        s0.klo_lo = 0; s0.klo_hi = 0;
        s0.khi_lo = klo<K>(kfull_hi); s0.khi_hi = khi<K>(kfull_hi);
        stack[0] = s0;
        sh_nstack = 1;
    }
    __syncthreads();

    for (;;) {
        if (threadIdx.x == 0) {
            sh_go = 0;
            if (sh_nstack > 0) { cur = stack[--sh_nstack]; sh_go = 1; }
            sh_n_occ = 0;
            sh_ovf = 0;
        }
        for (int i = threadIdx.x; i < H; i += T) { tkeys[i] = KT<K>::empty(); tcnt[i] = 0; }
        __syncthreads();
        if (!sh_go) break;
        const uint32_t d_lo = g_start[cur.s_lo >> 2], d_hi = g_start[(cur.s_hi + 3) >> 2];
        const uint32_t f_lo = cur.s_lo, f_hi = cur.s_hi;  // sub-bucket filter inside the bins
        const K key_lo = mk<K>(cur.klo_lo, cur.klo_hi), key_hi = mk<K>(cur.khi_lo, cur.khi_hi);
        const bool full_range = (key_lo == (K)0) && (key_hi == kfull_hi);

        for (uint32_t c0 = d_lo; c0 < d_hi;) {
            const uint32_t r = c0 + threadIdx.x;
            uint32_t L = 0, km = 0;
            uint64_t d = 0;
            bool valid = false;
            if (threadIdx.x < kChunkRec && r < d_hi) {
                d = a.scratch[rec0 + r];
                const uint32_t sb = side_sub((uint32_t)d);
                valid = true;
                if (sb >= f_lo && sb < f_hi) {  // records of other sub-buckets in the bin count as empty
                    km = side_kmers((uint32_t)d);
                    L = km + (uint32_t)k - 1u;
                }
            }
            uint32_t totL;
            const uint32_t exL = wg_exscan_u32(L, tmp, &totL);
            uint32_t totK;
            const uint32_t exK = wg_exscan_u32(km, tmp, &totK);
            const bool in = valid && (exL + L <= (uint32_t)kChunkBaseWords * 16u);
            if (threadIdx.x == 0) { sh_chunk_n = 0; sh_long_rec = 0; }
            __syncthreads();
            if (in) {
                u.ins.desc[threadIdx.x] = d;
                u.ins.kpre[threadIdx.x] = exK;
                u.ins.bpre[threadIdx.x] = exL;
                atomicMax(&sh_chunk_n, threadIdx.x + 1u);
            }
            __syncthreads();
            uint32_t n = sh_chunk_n;
            if (n == 0) {
                if (threadIdx.x == 0) sh_long_rec = 1;
                n = 1;
            }
            __syncthreads();
            if (!sh_long_rec) {
                if (threadIdx.x == n - 1) {
                    u.ins.kpre[n] = exK + km;
                    u.ins.bpre[n] = exL + L;
                }
                for (int i = threadIdx.x; i < kChunkBaseWords + 6; i += T) u.ins.bases[i] = 0;
                __syncthreads();
                // unpack: one thread per record, 2-bit words assembled in registers then OR-ed in
                if (threadIdx.x < n && u.ins.bpre[threadIdx.x + 1] > u.ins.bpre[threadIdx.x]) {
                    const uint64_t dd = u.ins.desc[threadIdx.x];
                    const uint8_t* rec = store + (dd >> 32);
                    const uint32_t bm3 = side_bm3((uint32_t)dd);
                    const uint32_t Lr = side_kmers((uint32_t)dd) + (uint32_t)k - 1u;
                    uint32_t g = u.ins.bpre[threadIdx.x];
                    uint32_t acc = 0;
                    for (uint32_t t = 0; t < Lr; t++, g++) {
                        acc |= rec_base(rec, bm3, t) << (2u * (g & 15u));
                        if ((g & 15u) == 15u || t + 1u == Lr) { atomicOr(&u.ins.bases[g >> 4], acc); acc = 0; }
                    }
                }
                __syncthreads();
                const uint32_t nk = u.ins.kpre[n];
                for (uint32_t x = threadIdx.x; x < nk; x += T) {
                    uint32_t lo = 0, hi = n;
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (u.ins.kpre[mid] <= x) lo = mid; else hi = mid;
                    }
                    const uint32_t g = u.ins.bpre[lo] + (x - u.ins.kpre[lo]);
                    K f, rc;
                    kmer_at(u.ins.bases, g, k, f, rc);
                    const K key = f < rc ? f : rc;
                    if (full_range || (key >= key_lo && key <= key_hi)) {
                        if (!sh_ovf) insert<K>(tkeys, tcnt, key, &sh_n_occ, &sh_ovf);
                    }
                }
            } else {
                // a record longer than the LDS base array: k-mers built from its bytes directly
                const uint64_t dd = a.scratch[rec0 + c0];
                const uint8_t* rec = store + (dd >> 32);
                const uint32_t bm3 = side_bm3((uint32_t)dd);
                const uint32_t nk = side_kmers((uint32_t)dd);
                for (uint32_t x = threadIdx.x; x < nk; x += T) {
                    K f = 0, rc = 0;
                    for (int t = 0; t < k; t++) {
                        const uint32_t c = rec_base(rec, bm3, x + (uint32_t)t);
                        f = (f << 2) | (K)c;
                        rc |= (K)(3u - c) << (2 * t);
                    }
                    const K key = f < rc ? f : rc;
                    if (full_range || (key >= key_lo && key <= key_hi)) {
                        if (!sh_ovf) insert<K>(tkeys, tcnt, key, &sh_n_occ, &sh_ovf);
                    }
                }
            }
            __syncthreads();
            c0 += n;
        }

        if (sh_ovf) {
            if (threadIdx.x == 0) {
                atomicAdd(&a.stats[1], 1ull);
                if (sh_ovf >= 2u || sh_nstack + 2 > (uint32_t)kStack) {
                    raise_err(a.err, RGKC_ERR_COUNT_FAIL);
                } else if (cur.s_hi - cur.s_lo > 4) {
                    // split at a bin boundary near the instance midpoint
                    const uint32_t b_lo = cur.s_lo >> 2, b_hi = (cur.s_hi + 3) >> 2;
                    uint32_t tot = 0;
                    for (uint32_t b = b_lo; b < b_hi; b++) tot += g_inst[b];
                    uint32_t acc = 0, m = b_lo;
                    while (m + 1 < b_hi && acc + g_inst[m] <= tot / 2) acc += g_inst[m++];
                    if (m == b_lo) m = b_lo + 1;
                    Sweep hiS = cur, loS = cur;
                    loS.s_hi = 4 * m; hiS.s_lo = 4 * m;
                    stack[sh_nstack++] = hiS;
                    stack[sh_nstack++] = loS;
                } else if (cur.s_hi - cur.s_lo > 1) {
                    // inside one bin: halve the sub-bucket range (records filtered by sub-bucket)
                    const uint32_t m = cur.s_lo + ((cur.s_hi - cur.s_lo) >> 1);
                    Sweep hiS = cur, loS = cur;
                    loS.s_hi = m; hiS.s_lo = m;
                    stack[sh_nstack++] = hiS;
                    stack[sh_nstack++] = loS;
                } else {
                    const K mid = key_lo + ((key_hi - key_lo) >> 1);
                    if (mid == key_hi) {
                        raise_err(a.err, RGKC_ERR_COUNT_FAIL);
                    } else {
                        Sweep hiS = cur, loS = cur;
                        loS.khi_lo = klo<K>(mid); loS.khi_hi = khi<K>(mid);
                        const K m1 = mid + (K)1;
                        hiS.klo_lo = klo<K>(m1); hiS.klo_hi = khi<K>(m1);
                        stack[sh_nstack++] = hiS;
                        stack[sh_nstack++] = loS;
                    }
                }
            }
            __syncthreads();
            continue;
        }
        // ---- finalize: counting sort of the distinct keys by their top 12 bits in range -----
        const uint32_t nd = sh_n_occ;
        const K span = key_hi - key_lo;
        const int bl = bitlen(span);
        const int shift = bl > 12 ? bl - 12 : 0;
        for (int i = threadIdx.x; i < kBins; i += T) bins[i] = 0;
        __syncthreads();
        const K E = KT<K>::empty();
        for (int i = threadIdx.x; i < H; i += T) {
            const K key = tkeys[i];
            if (key != E) atomicAdd(&bins[(uint32_t)((key - key_lo) >> shift)], 1u);
        }
        __syncthreads();
        {
            uint32_t v[4], tot, acc = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) { v[j] = bins[threadIdx.x * 4 + j]; acc += v[j]; }
            uint32_t ex = wg_exscan_u32(acc, tmp, &tot);
#pragma unroll
            for (int j = 0; j < 4; j++) { bins[threadIdx.x * 4 + j] = ex; ex += v[j]; }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < H; i += T) {
            const K key = tkeys[i];
            if (key != E) {
                const uint32_t pos = atomicAdd(&bins[(uint32_t)((key - key_lo) >> shift)], 1u);
                u.fin.skeys[pos] = key;
                u.fin.scnt[pos] = tcnt[i];
            }
        }
        __syncthreads();
        // bins[b] now ends bin b; sort each bin (tiny) by insertion
        for (int bb = threadIdx.x * 4; bb < threadIdx.x * 4 + 4; bb++) {
            const uint32_t s = bb == 0 ? 0u : bins[bb - 1], e = bins[bb];
            for (uint32_t i = s + 1; i < e; i++) {
                const K key = u.fin.skeys[i];
                const uint32_t c = u.fin.scnt[i];
                uint32_t j = i;
                while (j > s && u.fin.skeys[j - 1] > key) {
                    u.fin.skeys[j] = u.fin.skeys[j - 1];
                    u.fin.scnt[j] = u.fin.scnt[j - 1];
                    j--;
                }
                u.fin.skeys[j] = key;
                u.fin.scnt[j] = c;
            }
        }
        __syncthreads();
        // filter [cmin, cmax] and append rows
        constexpr int PER = (CAP + T - 1) / T;
        uint32_t keep = 0;
        const uint32_t i0 = threadIdx.x * PER;
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const uint32_t i = i0 + j;
            if (i < nd) { const uint32_t c = u.fin.scnt[i]; keep += (c >= a.cmin && c <= a.cmax); }
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
                    const uint32_t c = u.fin.scnt[i];
                    if (c >= a.cmin && c <= a.cmax) {
                        const uint64_t row = rb + ex;
                        if constexpr (sizeof(K) == 8) {
                            ((uint64_t*)a.out_keys)[row] = (uint64_t)u.fin.skeys[i];
                        } else {
                            ((uint64_t*)a.out_keys)[2 * row] = klo<K>(u.fin.skeys[i]);
                            ((uint64_t*)a.out_keys)[2 * row + 1] = khi<K>(u.fin.skeys[i]);
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
        hipLaunchKernelGGL(k_count_items<uint64_t>, dim3((unsigned)a.n_items), dim3(T), 0, s, a);
    else
        hipLaunchKernelGGL(k_count_items<u128>, dim3((unsigned)a.n_items), dim3(T), 0, s, a);
    return hipGetLastError();
}

}  // namespace rgkc

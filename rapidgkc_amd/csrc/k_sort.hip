// Row order of the count tables on gfx950: each partition's (key, count) rows in ascending key
// order, the order of the reference's table (cpu_funcs.cpp:614-640: its rows come from a run-length
// pass over the sorted keys).
//
// Rows of one partition arrive as runs (one per count sweep), each ascending; the keys of different
// runs interleave. Keys of a table are distinct, so the sort needs no stability. MSD radix over
// segments planned on the host, each split by a digit of 4 to 12 bits sized to the segment (the
// histogram rows are as wide as the round's widest digit):
//   - a segment of at most kLeaf rows is loaded into LDS by one workgroup, bitonic-sorted there
//     (padded to a power of two with all-ones keys, which no k-mer key reaches) and written to its
//     final table rows;
//   - a larger segment is distributed by its next digit: a histogram pass (LDS per chunk, then one
//     global atomic per non-empty bin), the host turns the histogram into bucket starts and packs
//     runs of consecutive buckets into child segments of at most kLeaf rows (consecutive digits are
//     consecutive key ranges, so sorting a run of buckets sorts each), and a scatter pass moves the
//     rows into their buckets in the other buffer (one global atomic per non-empty bin per chunk).
//     A bucket larger than kLeaf becomes a segment of the next level.
// Two buffers alternate: the table itself and a scratch region indexed by the partitions' row
// prefix; a segment records which of the two holds its rows.
#include "rgkc_device.h"
#include "rgkc_internal.h"

namespace rgkc {
namespace {

constexpr int kDigitBits = kSortDigitBits;
constexpr int kBinsS = 1 << kDigitBits;
constexpr int kSortThreads = 256;
constexpr uint32_t kChunk = kSortChunk;  // rows per histogram / scatter workgroup

template <typename K> struct SortKey;
template <> struct SortKey<uint64_t> {
    static __device__ __forceinline__ uint64_t load(const void* p, uint64_t i) { return ((const uint64_t*)p)[i]; }
    static __device__ __forceinline__ void store(void* p, uint64_t i, uint64_t v) { ((uint64_t*)p)[i] = v; }
};
template <> struct SortKey<u128> {
    static __device__ __forceinline__ u128 load(const void* p, uint64_t i) {
        const uint64_t* q = (const uint64_t*)p + 2 * i;
        return ((u128)q[1] << 64) | q[0];
    }
    static __device__ __forceinline__ void store(void* p, uint64_t i, u128 v) {
        uint64_t* q = (uint64_t*)p + 2 * i;
        q[0] = (uint64_t)v;
        q[1] = (uint64_t)(v >> 64);
    }
};

template <typename K> __device__ __forceinline__ uint32_t digit_of(K key, uint32_t shift, uint32_t nbits) {
    return (uint32_t)(key >> shift) & ((1u << nbits) - 1u);
}

// Leaf: one workgroup sorts one segment of at most LEAF rows in LDS.
template <typename K, int LEAF>
__global__ __launch_bounds__(kSortThreads) void k_sort_leaf(SortArgs a) {
    __shared__ K sk[LEAF];
    __shared__ uint32_t sc[LEAF];
    const SortSeg sg = a.leaf[blockIdx.x];
    const void* src_k = sg.where ? a.tkeys : a.keys;
    const uint32_t* src_c = sg.where ? a.tcounts : a.counts;
    const uint64_t so = sg.where ? sg.toff : sg.start;
    const uint32_t n = sg.len;
    uint32_t m = 1;
    while (m < n) m <<= 1;
    for (uint32_t i = threadIdx.x; i < m; i += kSortThreads) {
        sk[i] = i < n ? SortKey<K>::load(src_k, so + i) : ~(K)0;
        sc[i] = i < n ? src_c[so + i] : 0u;
    }
    __syncthreads();
    for (uint32_t size = 2; size <= m; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t t = threadIdx.x; t < m / 2; t += kSortThreads) {
                const uint32_t i = 2 * t - (t & (stride - 1));  // lower index of the pair
                const uint32_t j = i + stride;
                const bool up = (i & size) == 0;
                const K x = sk[i], y = sk[j];
                if ((x > y) == up) {
                    sk[i] = y; sk[j] = x;
                    const uint32_t c = sc[i]; sc[i] = sc[j]; sc[j] = c;
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t i = threadIdx.x; i < n; i += kSortThreads) {
        SortKey<K>::store(a.keys, sg.start + i, sk[i]);
        a.counts[sg.start + i] = sc[i];
    }
}

// Histogram of one chunk of a big segment (chunk = blockIdx.x; a.chunk_seg maps it to its segment).
template <typename K>
__global__ __launch_bounds__(kSortThreads) void k_sort_hist(SortArgs a) {
    __shared__ uint32_t h[kBinsS];
    const uint32_t si = a.chunk_seg[blockIdx.x];
    const SortSeg sg = a.big[si];
    const uint32_t c0 = (blockIdx.x - a.seg_chunk0[si]) * kChunk;
    const uint32_t c1 = min(sg.len, c0 + kChunk);
    const void* src_k = sg.where ? a.tkeys : a.keys;
    const uint64_t so = sg.where ? sg.toff : sg.start;
    for (int i = threadIdx.x; i < kBinsS; i += kSortThreads) h[i] = 0;
    __syncthreads();
    for (uint32_t i = c0 + threadIdx.x; i < c1; i += kSortThreads)
        atomicAdd(&h[digit_of<K>(SortKey<K>::load(src_k, so + i), sg.shift, sg.nbits)], 1u);
    __syncthreads();
    uint32_t* const gh = a.hist + (size_t)si * a.hstride;
    for (int i = threadIdx.x; i < (int)a.hstride; i += kSortThreads)
        if (h[i]) atomicAdd(&gh[i], h[i]);
}

// Scatter one chunk of a big segment into its buckets in the other buffer (a.cursor holds each
// bucket's next free row, relative to the segment).
template <typename K>
__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(SortArgs a) {
    __shared__ uint32_t h[kBinsS];
    __shared__ uint32_t base[kBinsS];
    constexpr int RPT = kChunk / kSortThreads;
    const uint32_t si = a.chunk_seg[blockIdx.x];
    const SortSeg sg = a.big[si];
    const uint32_t c0 = (blockIdx.x - a.seg_chunk0[si]) * kChunk;
    const void* src_k = sg.where ? a.tkeys : a.keys;
    const uint32_t* src_c = sg.where ? a.tcounts : a.counts;
    void* dst_k = sg.where ? a.keys : a.tkeys;
    uint32_t* dst_c = sg.where ? a.counts : a.tcounts;
    const uint64_t so = sg.where ? sg.toff : sg.start;
    const uint64_t dof = sg.where ? sg.start : sg.toff;
    for (int i = threadIdx.x; i < kBinsS; i += kSortThreads) h[i] = 0;
    __syncthreads();
    K kv[RPT];
    uint32_t cv[RPT], rk[RPT];
#pragma unroll
    for (int j = 0; j < RPT; j++) {
        const uint32_t i = c0 + threadIdx.x + (uint32_t)j * kSortThreads;
        rk[j] = ~0u;
        if (i < sg.len) {
            kv[j] = SortKey<K>::load(src_k, so + i);
            cv[j] = src_c[so + i];
            const uint32_t d = digit_of<K>(kv[j], sg.shift, sg.nbits);
            rk[j] = d | (atomicAdd(&h[d], 1u) << kDigitBits);  // bin | rank within the chunk's bin
        }
    }
    __syncthreads();
    uint32_t* const cur = a.cursor + (size_t)si * a.hstride;
    for (int i = threadIdx.x; i < (int)a.hstride; i += kSortThreads)
        if (h[i]) base[i] = atomicAdd(&cur[i], h[i]);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RPT; j++) {
        if (rk[j] == ~0u) continue;
        const uint32_t d = rk[j] & (kBinsS - 1u);
        const uint64_t r = dof + base[d] + (rk[j] >> kDigitBits);
        SortKey<K>::store(dst_k, r, kv[j]);
        dst_c[r] = cv[j];
    }
}

}  // namespace

int sort_leaf_cap(int key_bytes) { return key_bytes == 8 ? 4096 : 2048; }

hipError_t launch_sort_leaf(const SortArgs& a, uint32_t n_leaf, int key_bytes, hipStream_t s) {
    if (!n_leaf) return hipSuccess;
    if (key_bytes == 8) hipLaunchKernelGGL((k_sort_leaf<uint64_t, 4096>), dim3(n_leaf), dim3(kSortThreads), 0, s, a);
    else hipLaunchKernelGGL((k_sort_leaf<u128, 2048>), dim3(n_leaf), dim3(kSortThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_sort_hist(const SortArgs& a, uint32_t n_chunks, int key_bytes, hipStream_t s) {
    if (!n_chunks) return hipSuccess;
    if (key_bytes == 8) hipLaunchKernelGGL(k_sort_hist<uint64_t>, dim3(n_chunks), dim3(kSortThreads), 0, s, a);
    else hipLaunchKernelGGL(k_sort_hist<u128>, dim3(n_chunks), dim3(kSortThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_sort_scatter(const SortArgs& a, uint32_t n_chunks, int key_bytes, hipStream_t s) {
    if (!n_chunks) return hipSuccess;
    if (key_bytes == 8) hipLaunchKernelGGL(k_sort_scatter<uint64_t>, dim3(n_chunks), dim3(kSortThreads), 0, s, a);
    else hipLaunchKernelGGL(k_sort_scatter<u128>, dim3(n_chunks), dim3(kSortThreads), 0, s, a);
    return hipGetLastError();
}

}  // namespace rgkc

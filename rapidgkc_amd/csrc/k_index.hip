// Side descriptors for superkmer stores that arrive from the host (e.g. the chunks of an
// SKMStoreNoncon, skmstore.hpp:62-99), which carry only record bytes.
//
// One workgroup per partition walks its bytes in 1024-byte steps. A record starts at byte 0 and
// after every "full byte then non-full byte" pair (the reference's boundary rule,
// cpu_funcs.cpp:516-525 / gpu_kmercounting.cu:166-176). The thread that owns a start follows the
// record to its last byte (first non-full byte after the first), which gives its base count, and
// in the fill pass recomputes its minimizer from its first k-mer (all k-mers of a superkmer share
// it) to get the sub-bucket id.
#include "rgkc_device.h"
#include "rgkc_internal.h"

namespace rgkc {
namespace {

constexpr int T = 1024;

__device__ __forceinline__ bool is_start(const uint8_t* s, uint64_t i) {
    if (i == 0) return true;
    if (i < 2) return false;
    return s[i - 1] < 0xC0 && s[i - 2] >= 0xC0;
}

// Record starting at byte i of a partition of n bytes: returns bytes, sets bm3 and bases.
__device__ __forceinline__ uint32_t record_shape(const uint8_t* s, uint64_t i, uint64_t n, uint32_t& bm3, uint32_t& L,
                                                 bool& ok) {
    const uint32_t ind0 = s[i] >> 6;
    uint64_t j = i + 1;
    while (j < n && s[j] >= 0xC0) j++;
    ok = (ind0 != 0) && (j < n);
    if (!ok) { bm3 = 0; L = 0; return 1; }
    bm3 = 3u - ind0;
    L = ind0 + 3u * (uint32_t)(j - i - 1) + (uint32_t)(s[j] >> 6);
    return (uint32_t)(j - i + 1);
}

template <bool FILL>
__global__ __launch_bounds__(T) void k_index(IndexArgs a) {
    __shared__ uint32_t tmp[T / 64 + 1];
    __shared__ unsigned long long sh_kmers;
    const uint32_t part = blockIdx.x;
    const uint64_t b0 = a.part_byte[part], n = a.part_byte[part + 1] - b0;
    const uint8_t* s = a.store + b0;
    const int k = a.k, p = a.p;
    if (threadIdx.x == 0) sh_kmers = 0;
    __syncthreads();
    uint64_t carry = 0;
    uint64_t kmers_local = 0;
    const uint64_t rbase = FILL ? a.rec_off[part] : 0;
    for (uint64_t c0 = 0; c0 < n; c0 += T) {
        const uint64_t i = c0 + threadIdx.x;
        const bool st = i < n && is_start(s, i);
        uint32_t tot;
        const uint32_t ex = wg_exscan_u32(st ? 1u : 0u, tmp, &tot);
        if (st) {
            uint32_t bm3, L;
            bool ok;
            record_shape(s, i, n, bm3, L, ok);
            if (!ok || L < (uint32_t)k || L - (uint32_t)k + 1u > 0x3FFFFu) {
                raise_err(a.err, RGKC_ERR_BAD_STORE);
            } else {
                const uint32_t km = L - (uint32_t)k + 1u;
                kmers_local += km;
                if (FILL) {
                    // minimizer of the first k-mer: min over its p-mers of the filtered strands
                    const uint32_t pmask = (p >= 16) ? 0xFFFFFFFFu : ((1u << (2 * p)) - 1u);
                    const uint32_t rsh = 2u * (uint32_t)(p - 1);
                    uint32_t f = 0, r = 0, mm = pmask;
                    for (int t = 0; t < k; t++) {
                        const uint32_t q = bm3 + (uint32_t)t;
                        const uint32_t c = ((uint32_t)s[i + q / 3u] >> (4u - 2u * (q % 3u))) & 3u;
                        f = ((f << 2) | c) & pmask;
                        r = (r >> 2) | ((3u - c) << rsh);
                        if (t >= p - 1) {
                            const uint32_t vf = pmer_valid(f, p) ? f : pmask;
                            const uint32_t vr = pmer_valid(r, p) ? r : pmask;
                            const uint32_t v = vf < vr ? vf : vr;
                            mm = v < mm ? v : mm;
                        }
                    }
                    const uint32_t nmm = ~mm;
                    // the store's partition count spreads a partition's minimizers over the
                    // sub-buckets; without it, the top bits of a multiplicative hash of them do
                    const uint32_t sub = a.parts_total ? (nmm / (uint32_t)a.parts_total) % (uint32_t)kSubIds
                                                       : (nmm * 0x9E3779B1u) >> (32 - 12);
                    a.side[rbase + carry + ex] = side_pack(sub, km, bm3);
                }
            }
        }
        carry += tot;
    }
    atomicAdd(&sh_kmers, (unsigned long long)kmers_local);
    __syncthreads();
    if (!FILL && threadIdx.x == 0) {
        a.rec_count[2 * part] = carry;
        a.rec_count[2 * part + 1] = sh_kmers;
    }
}

}  // namespace

// rec_count receives (records, k-mers) per partition.
hipError_t launch_index_count(const IndexArgs& a, hipStream_t s) {
    if (!a.n_parts) return hipSuccess;
    hipLaunchKernelGGL(k_index<false>, dim3(a.n_parts), dim3(T), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_index_fill(const IndexArgs& a, hipStream_t s) {
    if (!a.n_parts) return hipSuccess;
    hipLaunchKernelGGL(k_index<true>, dim3(a.n_parts), dim3(T), 0, s, a);
    return hipGetLastError();
}

}  // namespace rgkc

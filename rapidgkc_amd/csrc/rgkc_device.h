// Device-side helpers shared by the gfx950 kernels of the k-mer counting path.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef unsigned __int128 u128;

// Error bits raised by kernels into a device word (host checks after each stage).
enum : uint32_t {
    RGKC_ERR_STAGE_OVERFLOW = 1u << 0,   // a workgroup staging area would overflow (handled: direct path)
    RGKC_ERR_PART_OVERFLOW = 1u << 1,    // a partition region is smaller than its records (sizing miss)
    RGKC_ERR_SKM_TOO_LONG = 1u << 2,     // a superkmer record exceeds 65535 bytes (T_skm_len, types.h:19)
    RGKC_ERR_COUNT_FAIL = 1u << 3,       // a count sweep could not be made to fit (should not happen)
    RGKC_ERR_ROW_OVERFLOW = 1u << 4,     // table rows exceed the partition's row region
    RGKC_ERR_BAD_STORE = 1u << 5,        // malformed superkmer record in a store
    RGKC_ERR_SCRATCH = 1u << 6,          // per-bucket scratch exhausted
};

__device__ __forceinline__ void raise_err(uint32_t* err, uint32_t bit) { atomicOr(err, bit); }

// 2-bit code of an ASCII base: A/a/N/n -> 0, C/c -> 1, G/g -> 2, T/t -> 3, anything else -> 0
// (cpu_funcs.cpp:20-53 maps N to A; other bytes are outside the reference's contract).
__device__ __forceinline__ uint32_t base_code(uint32_t c) {
    const uint32_t u = c & 0xDFu;  // upper-case
    const uint32_t v = ((u >> 1) ^ (u >> 2)) & 3u;
    const bool ok = (u == 'A') | (u == 'C') | (u == 'G') | (u == 'T');
    return ok ? v : 0u;
}

// minimizer_filter.h:19-20
__device__ __forceinline__ bool pmer_valid(uint32_t pmer, int p) {
    return (((pmer >> ((p - 3) * 2)) & 0x2Bu) != 0u) & ((pmer & 0x3Fu) != 0u);
}

// Reverse complement of a k-mer held in the low 2k bits: complement, reverse the 2-bit digits.
__device__ __forceinline__ uint64_t revcomp64(uint64_t x, int k) {
    uint64_t y = ~x;
    y = __builtin_bitreverse64(y);
    y = ((y >> 1) & 0x5555555555555555ull) | ((y & 0x5555555555555555ull) << 1);
    return y >> (64 - 2 * k);
}
__device__ __forceinline__ u128 revcomp128(u128 x, int k) {
    const uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
    uint64_t rl = ~hi, rh = ~lo;  // reversing 128 bits swaps the words
    rl = __builtin_bitreverse64(rl);
    rh = __builtin_bitreverse64(rh);
    rl = ((rl >> 1) & 0x5555555555555555ull) | ((rl & 0x5555555555555555ull) << 1);
    rh = ((rh >> 1) & 0x5555555555555555ull) | ((rh & 0x5555555555555555ull) << 1);
    u128 r = ((u128)rh << 64) | rl;
    return r >> (128 - 2 * k);
}

// Superkmer record bytes (cpu_funcs.cpp:216-218): ((beg%3) + L + 3) / 3 with L = bases.
__device__ __forceinline__ uint32_t skm_nbytes(uint32_t bm3, uint32_t L) { return (bm3 + L + 3u) / 3u; }

// Side descriptor of one record: [31:20] sub-bucket id (12 bits), [19:2] k-mers (18 bits),
// [1:0] record start % 3.
__device__ __forceinline__ uint32_t side_pack(uint32_t sub, uint32_t kmers, uint32_t bm3) {
    return (sub << 20) | (kmers << 2) | bm3;
}
__device__ __forceinline__ uint32_t side_sub(uint32_t s) { return s >> 20; }
__device__ __forceinline__ uint32_t side_kmers(uint32_t s) { return (s >> 2) & 0x3FFFFu; }
__device__ __forceinline__ uint32_t side_bm3(uint32_t s) { return s & 3u; }

// Unsigned 32-bit division by a divisor fixed for a launch: multiply-high and shifts
// (Granlund and Montgomery, "Division by invariant integers using multiplication", 1994, fig. 4.1:
// m = floor(2^32 (2^l - d) / d) + 1 with l = ceil(log2 d), q = (t + ((x - t) >> min(l,1))) >> max(l-1,0),
// t = mulhi(m, x)). Exact for every x and d >= 1.
struct FastDivU32 {
    uint32_t d, m, s1, s2;
    __host__ __device__ static FastDivU32 make(uint32_t d) {
        uint32_t l = 0;
        while (l < 32 && (1ull << l) < d) l++;
        FastDivU32 f;
        f.d = d;
        f.m = (uint32_t)((((1ull << 32) * ((1ull << l) - d)) / d) + 1ull);
        f.s1 = l < 1 ? l : 1u;
        f.s2 = l > 1 ? l - 1u : 0u;
        return f;
    }
    __device__ __forceinline__ uint32_t div(uint32_t x) const {
        const uint32_t t = __umulhi(m, x);
        return (t + ((x - t) >> s1)) >> s2;
    }
    __device__ __forceinline__ uint32_t mod(uint32_t x) const { return x - div(x) * d; }
};

// Inclusive scan over the 64 lanes of a wave: DPP row shifts scan each 16-lane row, then the row
// totals (read from lanes 15, 31, 47) are added to the rows above them. All lanes must be active.
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, false);  // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, false);  // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, false);  // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, false);  // row_shr:8
    const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)x, 15);
    const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)x, 31);
    const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)x, 47);
    const uint32_t row = (threadIdx.x & 63u) >> 4;
    return x + (row >= 1 ? r0 : 0u) + (row >= 2 ? r1 : 0u) + (row >= 3 ? r2 : 0u);
}

// Workgroup-wide exclusive scan of one uint32 per thread (blockDim.x <= 1024, multiple of 64).
// `tmp` needs blockDim.x/64 + 1 words of LDS. Returns the exclusive prefix; *total gets the sum.
// Ends with a barrier, so `tmp` may be reused at once.
__device__ __forceinline__ uint32_t wg_exscan_u32(uint32_t v, uint32_t* tmp, uint32_t* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint32_t x = wave_incl_scan_u32(v);
    if (lane == 63) tmp[wave] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (int w = 0; w < nw; w++) {
        const uint32_t t = tmp[w];
        pre += (w < wave) ? t : 0u;
        tot += t;
    }
    *total = tot;
    __syncthreads();
    return pre + x - v;
}

// Two workgroup-wide exclusive scans at the cost of one (same barriers): `tmp` needs
// 2 * blockDim.x/64 words of LDS. Ends with a barrier.
__device__ __forceinline__ void wg_exscan2_u32(uint32_t v0, uint32_t v1, uint32_t* tmp, uint32_t& e0, uint32_t& e1,
                                               uint32_t& t0, uint32_t& t1) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint32_t x0 = wave_incl_scan_u32(v0), x1 = wave_incl_scan_u32(v1);
    if (lane == 63) { tmp[2 * wave] = x0; tmp[2 * wave + 1] = x1; }
    __syncthreads();
    uint32_t p0 = 0, p1 = 0, s0 = 0, s1 = 0;
    for (int w = 0; w < nw; w++) {
        const uint32_t a = tmp[2 * w], b = tmp[2 * w + 1];
        if (w < wave) { p0 += a; p1 += b; }
        s0 += a;
        s1 += b;
    }
    e0 = p0 + x0 - v0;
    e1 = p1 + x1 - v1;
    t0 = s0;
    t1 = s1;
    __syncthreads();
}

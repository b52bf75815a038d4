// Device-wide scans used to size partitions and buckets.
#include "rgkc_device.h"
#include "rgkc_internal.h"

namespace rgkc {
namespace {

constexpr int kT = 1024;
constexpr int kPer = kScanTile / kT;  // 4 elements per thread

__global__ __launch_bounds__(kT) void k_scan_tiles(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                                                   size_t n, uint64_t* __restrict__ tile_sums) {
    __shared__ uint64_t wsum[kT / 64 + 1];
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kPer;
    uint64_t v[kPer];
    uint64_t s = 0;
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        v[j] = (base + j < n) ? in[base + j] : 0;
        s += v[j];
    }
    // wave-inclusive scan of per-thread sums
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t x = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t acc = 0;
        for (int w = 0; w < kT / 64; w++) { const uint64_t t = wsum[w]; wsum[w] = acc; acc += t; }
        wsum[kT / 64] = acc;
    }
    __syncthreads();
    uint64_t run = wsum[wave] + x - s;
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        if (base + j < n) out[base + j] = run;
        run += v[j];
    }
    if (threadIdx.x == 0 && tile_sums) tile_sums[blockIdx.x] = wsum[kT / 64];
}

__global__ __launch_bounds__(kT) void k_scan_add(uint64_t* __restrict__ out, size_t n,
                                                 const uint64_t* __restrict__ tile_pref) {
    const uint64_t add = tile_pref[blockIdx.x];
    const size_t base = (size_t)blockIdx.x * kScanTile;
    for (int j = threadIdx.x; j < kScanTile; j += kT)
        if (base + j < n) out[base + j] += add;
}

// One workgroup per row: exclusive scan of `cols` uint32 entries in place, row total out.
__global__ __launch_bounds__(kT) void k_rowscan(uint32_t* __restrict__ m, uint64_t* __restrict__ totals,
                                                uint32_t cols) {
    __shared__ uint32_t tmp[kT / 64 + 1];
    uint32_t* row = m + (size_t)blockIdx.x * cols;
    uint64_t carry = 0;
    for (uint32_t c0 = 0; c0 < cols; c0 += kT) {
        const uint32_t c = c0 + threadIdx.x;
        const uint32_t v = c < cols ? row[c] : 0u;
        uint32_t tot;
        const uint32_t ex = wg_exscan_u32(v, tmp, &tot);
        if (c < cols) row[c] = (uint32_t)(carry + ex);
        carry += tot;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

__global__ void k_add(const uint64_t* a, const uint64_t* b, uint64_t* o, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = a[i] + b[i];
}

}  // namespace

size_t scan_tmp_words(size_t n) {
    size_t w = 0;
    while (n > (size_t)kScanTile) {
        n = (n + kScanTile - 1) / kScanTile;
        w += n;
    }
    return w + 1;
}

hipError_t scan_u64(const uint64_t* in, uint64_t* out, size_t n, uint64_t* tmp, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const size_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles == 1) {
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kT), 0, s, in, out, n, (uint64_t*)nullptr);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(kT), 0, s, in, out, n, tmp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = scan_u64(tmp, tmp, tiles, tmp + tiles, s);  // in-place scan of the tile sums
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(kT), 0, s, out, n, (const uint64_t*)tmp);
    return hipGetLastError();
}

hipError_t rowscan_u32(uint32_t* m, uint64_t* totals, int rows, uint32_t cols, hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rowscan, dim3(rows), dim3(kT), 0, s, m, totals, cols);
    return hipGetLastError();
}

hipError_t add_u64(const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b, out, n);
    return hipGetLastError();
}

}  // namespace rgkc

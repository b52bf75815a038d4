// Phase 1 on gfx950: 2-bit packing of a read batch, and the canonical-minimizer scan that cuts
// reads into superkmers and routes them to partitions.
//
// Semantics follow the reference (see rgkc_device.h for the shared rules):
//   - per k-mer minimizer: min over its w = k-p+1 p-mers of min(filtered fwd, filtered rc)
//     (cpu_funcs.cpp:109-146, gpu_skmgen.cu:218-274);
//   - a superkmer starts where the minimizer value changes (cpu_funcs.cpp:147-198);
//   - partition = (~mm) % parts (cpu_funcs.cpp:213-215); record bytes = the read's packed bytes
//     with patched first/last indicators (cpu_funcs.cpp:200-250, gpu_skmgen.cu:429-470).
//
// Layout: reads are cut into segments of up to kSegKmers k-mers (a 150 bp read is one segment).
// One thread scans one segment: it computes the minimizer of k-mer s0-1 (to know whether a cut
// falls at s0), owns the superkmers that start inside [s0, s1), and follows its last one past s1
// until it ends. The sliding-window minimum works on blocks of w p-mers kept in an LDS ring
// (prefix minimum of the current block, suffix minima of the previous one). The emit passes stage
// the workgroup's bases in LDS; the count pass reads them from HBM in realigned words.
//
// Passes: the count pass (k_p1_count) sizes every (partition, workgroup) run and keeps each
// workgroup's superkmers as a list; the list emit (k_p1_emit_list) writes records from the list
// in store order; a workgroup whose list overflowed is rescanned by the scanning emit (k_p1_emit,
// or k_p1_emit_stream for capacity-sized regions and batched input).
#include "rgkc_device.h"
#include "rgkc_internal.h"

#include <algorithm>

namespace rgkc {
namespace {

// ---------------------------------------------------------------------------------------------
// pack: 16 ASCII bases -> one uint32 (base i of the batch at bits 2*(i%16) of word i/16).
__global__ __launch_bounds__(256) void k_pack(const uint4* __restrict__ reads, uint64_t n_words,
                                              uint32_t* __restrict__ bases2) {
    const uint64_t w = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (w >= n_words) return;
    const uint4 v = reads[w];
    const uint32_t x[4] = {v.x, v.y, v.z, v.w};
    uint32_t out = 0;
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int b = 0; b < 4; b++) out |= base_code((x[q] >> (8 * b)) & 0xFFu) << (2 * (q * 4 + b));
    bases2[w] = out;
}

__global__ void k_seg_count(const uint64_t* __restrict__ offs, uint64_t n_reads, int k, uint64_t* __restrict__ nseg) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_reads) return;
    const uint64_t len = offs[r + 1] - offs[r];
    nseg[r] = len >= (uint64_t)k ? (len - k + 1 + kSegKmers - 1) / kSegKmers : 0;
}

__global__ void k_seg_fill(const uint64_t* __restrict__ offs, const uint64_t* __restrict__ seg_off, uint64_t n_reads,
                           int k, uint32_t* __restrict__ seg_read, uint32_t* __restrict__ seg_first) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_reads) return;
    const uint64_t len = offs[r + 1] - offs[r];
    if (len < (uint64_t)k) return;
    const uint64_t n = (len - k + 1 + kSegKmers - 1) / kSegKmers;
    const uint64_t b = seg_off[r];
    for (uint64_t t = 0; t < n; t++) {
        seg_read[b + t] = (uint32_t)r;
        seg_first[b + t] = (uint32_t)(t * kSegKmers);
    }
}

// Device-driven variants for batched input (sizes in a device header: [0] reads, [1] bases,
// [2] segments), launched over the batch capacity so a captured graph can replay them.
__global__ __launch_bounds__(256) void k_pack_dev(const uint4* __restrict__ reads, const uint64_t* __restrict__ hdr,
                                                  uint64_t cap_words, uint32_t* __restrict__ bases2) {
    const uint64_t n_words = (hdr[1] >> 4) + 2;
    for (uint64_t w = (uint64_t)blockIdx.x * 256 + threadIdx.x; w < n_words && w < cap_words;
         w += (uint64_t)gridDim.x * 256) {
        const uint4 v = reads[w];
        const uint32_t x[4] = {v.x, v.y, v.z, v.w};
        uint32_t out = 0;
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int b = 0; b < 4; b++) out |= base_code((x[q] >> (8 * b)) & 0xFFu) << (2 * (q * 4 + b));
        bases2[w] = out;
    }
}

__global__ void k_seg_count_dev(const uint64_t* __restrict__ offs, const uint64_t* __restrict__ hdr, uint64_t cap_reads,
                                int k, uint64_t* __restrict__ nseg) {
    const uint64_t n = hdr[0];
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r <= cap_reads; r += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t v = 0;
        if (r < n) {
            const uint64_t len = offs[r + 1] - offs[r];
            v = len >= (uint64_t)k ? (len - k + 1 + kSegKmers - 1) / kSegKmers : 0;
        }
        nseg[r] = v;
    }
}

__global__ void k_seg_fill_dev(const uint64_t* __restrict__ offs, const uint64_t* __restrict__ seg_off,
                               uint64_t* __restrict__ hdr, uint64_t cap_reads, int k, uint32_t* __restrict__ seg_read,
                               uint32_t* __restrict__ seg_first) {
    const uint64_t n = hdr[0];
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid == 0) hdr[2] = seg_off[n];
    for (uint64_t r = gid; r < n && r < cap_reads; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t len = offs[r + 1] - offs[r];
        if (len < (uint64_t)k) continue;
        const uint64_t m = (len - k + 1 + kSegKmers - 1) / kSegKmers;
        const uint64_t b = seg_off[r];
        for (uint64_t t = 0; t < m; t++) {
            seg_read[b + t] = (uint32_t)r;
            seg_first[b + t] = (uint32_t)(t * kSegKmers);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Base access: words staged in LDS for the workgroup's range, global memory outside it.
struct BaseSrc {
    const uint32_t* g;   // global 2-bit words
    const uint32_t* l;   // LDS copy of words [w0, w0+nw)
    uint64_t w0;
    uint32_t nw;
    __device__ __forceinline__ uint32_t word(uint64_t wi) const {
        const uint64_t d = wi - w0;
        return d < nw ? l[d] : g[wi];
    }
};

struct BaseStream {
    uint64_t buf;
    uint64_t wi;
    uint32_t sh;
    __device__ __forceinline__ void init(const BaseSrc& s, uint64_t pos) {
        wi = pos >> 4;
        sh = 2u * (uint32_t)(pos & 15);
        buf = (uint64_t)s.word(wi) | ((uint64_t)s.word(wi + 1) << 32);
        wi += 2;
    }
    __device__ __forceinline__ uint32_t next(const BaseSrc& s) {
        const uint32_t c = (uint32_t)(buf >> sh) & 3u;
        sh += 2;
        if (sh == 32) {
            buf = (buf >> 32) | ((uint64_t)s.word(wi) << 32);
            wi++;
            sh = 0;
        }
        return c;
    }
};

// Record byte x of the packed read (cpu_funcs.cpp:200-211): full groups 11|b|b|b, the tail
// byte (len%3)<<6 with its bases high, or 0 when len%3 == 0.
struct Record {
    uint32_t part, sub, nb, kmers, bm3;
};

// Scan one segment; emit(b, e, mm) for each owned superkmer [b, e) (k-mer indices); step() runs at
// the top of every base step.
template <int RING, int T, class Emit, class Step>
__device__ __forceinline__ void scan_segment(const BaseSrc& src, uint32_t* ring, uint64_t R0, uint32_t len, uint32_t s0,
                                             int k, int p, Emit&& emit, Step&& step) {
    const uint32_t nk = len - (uint32_t)k + 1u;
    const uint32_t s1 = min(nk, s0 + (uint32_t)kSegKmers);
    const uint32_t w = (uint32_t)(k - p + 1);
    const uint32_t pmask = (p >= 16) ? 0xFFFFFFFFu : ((1u << (2 * p)) - 1u);
    const uint32_t rsh = 2u * (uint32_t)(p - 1);
    const uint32_t i0 = s0 > 0 ? s0 - 1 : 0;
    const int tid = threadIdx.x;
    BaseStream bs;
    bs.init(src, R0 + i0);
    uint32_t f = 0, r = 0, pm = 0, bu = w - 1u;
    uint32_t prev = 0, ob = 0, omm = 0;
    bool open = false;
    const uint32_t qfirst = i0 + w - 1;  // first p-mer that completes a window
    // Window minimum by blocks of w p-mers counted from the scan's first p-mer (van Herk /
    // Gil-Werman): pm is the prefix minimum of the current block; when a block completes its ring
    // slots are rewritten with suffix minima, and a window that starts inside the previous block is
    // min(suffix there, pm). Block ends fall on the same loop step in every lane that starts at the
    // same offset, and each costs one straight-line pass over w slots. The ring holds w slots per
    // thread (w <= RING).
    for (uint32_t j = i0; j < len; j++) {
        step();
        const uint32_t c = bs.next(src);
        f = ((f << 2) | c) & pmask;
        r = (r >> 2) | ((3u - c) << rsh);
        if (j < i0 + (uint32_t)p - 1u) continue;
        const uint32_t q = j - (uint32_t)p + 1u;
        const uint32_t vf = pmer_valid(f, p) ? f : pmask;
        const uint32_t vr = pmer_valid(r, p) ? r : pmask;
        const uint32_t v = vf < vr ? vf : vr;
        bu = (bu + 1u == w) ? 0u : bu + 1u;
        pm = bu == 0u ? v : (v < pm ? v : pm);
        uint32_t mm = pm;
        // ring slot of a p-mer = its place in its block: slot bu holds this p-mer, slot bu + 1 the
        // previous block's suffix minimum from the first p-mer of this window on
        if (bu + 1u == w) {
            ring[bu * T + tid] = v;
            uint32_t vals[RING];
#pragma unroll
            for (int t = 0; t < RING; t++)
                vals[t] = (uint32_t)t < w ? ring[(bu - (uint32_t)t) * T + tid] : 0xFFFFFFFFu;
            uint32_t sm = 0xFFFFFFFFu;
#pragma unroll
            for (int t = 0; t < RING; t++) { sm = vals[t] < sm ? vals[t] : sm; vals[t] = sm; }
#pragma unroll
            for (int t = 0; t < RING; t++)
                if ((uint32_t)t < w) ring[(bu - (uint32_t)t) * T + tid] = vals[t];
        } else {
            if (q >= qfirst) {
                const uint32_t sfx = ring[(bu + 1u) * T + tid];
                mm = sfx < pm ? sfx : pm;
            }
            ring[bu * T + tid] = v;
        }
        if (q < qfirst) continue;
        const uint32_t i = q - w + 1u;
        if (s0 > 0 && i == i0) { prev = mm; continue; }
        if (i == s0) {
            if (s0 == 0 || mm != prev) { open = true; ob = s0; omm = mm; }
            else open = false;
        } else if (mm != prev) {
            if (open) emit(ob, i, omm);
            if (i < s1) { open = true; ob = i; omm = mm; }
            else { open = false; break; }
        }
        prev = mm;
        if (i + 1u >= s1 && !open) break;
    }
    if (open) emit(ob, nk, omm);
}
template <int RING, int T, class Emit>
__device__ __forceinline__ void scan_segment(const BaseSrc& src, uint32_t* ring, uint64_t R0, uint32_t len, uint32_t s0,
                                             int k, int p, Emit&& emit) {
    scan_segment<RING, T>(src, ring, R0, len, s0, k, p, emit, [] {});
}

template <int T>
__device__ __forceinline__ void stage_bases(const P1Args& a, uint32_t* lds_w, uint32_t cap_words, BaseSrc& src,
                                            uint64_t seg_lo, uint64_t seg_hi) {
    // base range of the workgroup's segments (with the look-back k-mer and the segment tails)
    __shared__ uint64_t rng[2];
    if (threadIdx.x == 0) {
        const uint32_t rf = a.seg_read[seg_lo], rl = a.seg_read[seg_hi - 1];
        const uint32_t ff = a.seg_first[seg_lo], fl = a.seg_first[seg_hi - 1];
        const uint64_t lo = a.offs[rf] + (ff > 0 ? ff - 1 : 0);
        const uint64_t lenl = a.offs[rl + 1] - a.offs[rl];
        const uint64_t endl = min(lenl, (uint64_t)fl + kSegKmers + (uint64_t)a.k);
        const uint64_t hi = a.offs[rl] + endl;
        rng[0] = lo >> 4;
        rng[1] = ((hi + 15) >> 4) + 1;
    }
    __syncthreads();
    const uint64_t w0 = rng[0];
    uint64_t nw = rng[1] - w0;
    const uint64_t total_w = (a.n_bases >> 4) + 2;
    if (w0 + nw > total_w) nw = total_w - w0;
    src.g = a.bases2;
    src.l = lds_w;
    src.w0 = w0;
    src.nw = nw <= cap_words ? (uint32_t)nw : 0u;
    for (uint32_t i = threadIdx.x; i < src.nw; i += T) lds_w[i] = a.bases2[w0 + i];
    __syncthreads();
}

// ---------------------------------------------------------------------------------------------
// count pass: per-(partition, workgroup) histograms of record bytes, records and k-mers, and
// (when a.skm_list is set) the workgroup's superkmer list for the emit pass.
//
// The lanes of a wave step together: at step j every lane reads base i0 + j of its own segment
// (i0: the segment's first k-mer, or the one before it as look-back), so the window ring index, the
// warm-up and the block ends of the window minimum are the same in all lanes. Each lane reads its
// bases from HBM as 16-base words realigned to its segment, loaded two words ahead of use. A minimizer
// change is only noted in a per-lane LDS queue (its k-mer step and minimizer) by straight-line
// code; the queue is drained by the whole wave at once every few steps, where each lane turns its
// complete superkmers into histogram counts and list entries. So the per-superkmer work runs once
// per superkmer instead of once per step of any lane that emits.
constexpr int kEvQ = 8;      // queued events per lane
constexpr int kEvChunk = 4;  // steps between queue checks (a lane adds at most one event per step)

// LDS: the ring (w + 1 slots per thread: the last one is read, unused, at block ends), the event
// queue (kEvQ per thread) and three histograms of `parts`
template <int T>
__global__ __launch_bounds__(T, 3) void k_p1_count(P1Args a) {
    extern __shared__ __align__(16) uint32_t dyn[];  // 16-byte aligned: the event queue holds 8-byte entries
    const int k = a.k, p = a.p;
    const uint32_t w = (uint32_t)(k - p + 1);
    uint32_t* const ring = dyn;
    unsigned long long* const evq = (unsigned long long*)(dyn + (((size_t)(w + 1) * T + 1) & ~(size_t)1));
    uint32_t* const hb = (uint32_t*)(evq + (size_t)kEvQ * T);
    uint32_t* const hr = hb + a.parts;
    uint32_t* const hk = hr + a.parts;
    __shared__ uint32_t lcnt, lbad;
    const int tid = threadIdx.x;
    const uint32_t lane = (uint32_t)tid & 63u;
    for (int i = tid; i < a.parts; i += T) { hb[i] = 0; hr[i] = 0; hk[i] = 0; }
    if (tid == 0) { lcnt = 0; lbad = 0; }
    __syncthreads();
    const uint64_t seg_lo = (uint64_t)blockIdx.x * T;
    const uint64_t seg_hi = min(a.n_seg, seg_lo + T);
    const uint64_t s = seg_lo + (uint64_t)tid;
    const bool has = s < seg_hi;
    uint64_t pos0 = 0;
    uint32_t s0 = 0, i0 = 0, jend = 0, js1 = 0;
    if (has) {
        const uint32_t rd = a.seg_read[s];
        const uint64_t R0 = a.offs[rd];
        const uint32_t len = (uint32_t)(a.offs[rd + 1] - R0);
        s0 = a.seg_first[s];
        const uint32_t nk = len - (uint32_t)k + 1u;
        const uint32_t s1 = min(nk, s0 + (uint32_t)kSegKmers);
        i0 = s0 > 0 ? s0 - 1u : 0u;
        pos0 = R0 + i0;
        jend = len - i0;                     // step after the read's last k-mer
        js1 = s1 - i0 + (uint32_t)k - 1u;   // step of k-mer s1 (the next segment's first)
    }
    // raw words wi, wi + 1, ... of the 2-bit stream (clamped to the array) realigned by sh
    const uint32_t* const g = a.bases2;
    const uint64_t last_w = (a.n_bases >> 4) + 1;
    uint64_t wi = pos0 >> 4;
    const uint32_t sh = 2u * (uint32_t)(pos0 & 15);
    auto raw = [&](uint64_t i) __attribute__((always_inline)) { return g[i < last_w ? i : last_w]; };
    uint32_t r_lo = raw(wi), r_hi = raw(wi + 1), r_nx = raw(wi + 2);
    wi += 3;
    const uint32_t pmask = (p >= 16) ? 0xFFFFFFFFu : ((1u << (2 * p)) - 1u);
    const uint32_t rsh = 2u * (uint32_t)(p - 1);
    const FastDivU32 pdiv = FastDivU32::make((uint32_t)a.parts);
    uint64_t* const lst = a.skm_list ? a.skm_list + (size_t)blockIdx.x * a.skm_cap : nullptr;
    bool done = !has, open = false;
    uint32_t qn = 0, f = 0, r = 0, pm = 0, prev = 0, W = 0;
    uint32_t bu = w - 1u;  // ring slot of the current p-mer within its block (the same in every lane)
    unsigned long long* const myq = evq + tid;
    // superkmer [b, e) with minimizer mm, list slot li
    auto record = [&](uint32_t b, uint32_t e, uint32_t mm, uint32_t li) __attribute__((always_inline)) {
        const uint32_t part = pdiv.mod(~mm);
        const uint32_t nb = skm_nbytes(b % 3u, e - b + (uint32_t)k - 1u);
        if (nb > 65535u) raise_err(a.err, RGKC_ERR_SKM_TOO_LONG);
        atomicAdd(&hb[part], nb);
        atomicAdd(&hr[part], 1u);
        atomicAdd(&hk[part], e - b);
        if (a.mm_hist) atomicAdd(&a.mm_hist[mm >> a.hist_shift], e - b);
        if (lst) {
            if (e - b >= (1u << kSkmKmerBits)) lbad = 1u;  // too long for an entry: the emit pass rescans
            if (li < a.skm_cap)
                lst[li] = (uint64_t)mm | ((uint64_t)tid << 32) | ((uint64_t)(b - s0) << 40) | ((uint64_t)(e - b) << 47);
        }
    };
    // the wave drains its queues: each lane's events but the last are superkmer starts whose end
    // is the next event; the last is kept while it is an open start
    auto drain = [&]() __attribute__((always_inline)) {
        const uint32_t nc = qn ? qn - 1u : 0u;
        uint32_t li = 0;
        if (lst) {
            const uint32_t incl = wave_incl_scan_u32(nc);
            const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
            uint32_t wb = 0;
            if (lane == 63u && tot) wb = atomicAdd(&lcnt, tot);
            li = (uint32_t)__builtin_amdgcn_readlane((int)wb, 63) + incl - nc;
        }
        const uint32_t kb = i0 - (uint32_t)k + 1u;  // k-mer index = kb + step
        unsigned long long e0 = qn ? myq[0] : 0ull;
        for (uint32_t t = 0; t < nc; t++) {
            const unsigned long long e1 = myq[(size_t)(t + 1) * T];
            const uint32_t b = kb + ((uint32_t)(e0 >> 32) & 0x7FFFFFFFu);
            const uint32_t e = kb + ((uint32_t)(e1 >> 32) & 0x7FFFFFFFu);
            record(b, e, (uint32_t)e0, li + t);
            e0 = e1;
        }
        if (qn) {
            if (!(e0 >> 63)) { myq[0] = e0; qn = 1; }
            else qn = 0;
        }
    };
    uint32_t* const rbase = ring + tid;
    for (uint32_t j0 = 0;; j0 += kEvChunk) {
        if ((j0 & 15u) == 0u) {  // next word; fetch the one after it
            W = __builtin_amdgcn_alignbit(r_hi, r_lo, sh);
            r_lo = r_hi;
            r_hi = r_nx;
            r_nx = raw(wi++);
        }
#pragma unroll
        for (int t = 0; t < kEvChunk; t++) {
            const uint32_t j = j0 + (uint32_t)t;
            const uint32_t c = (W >> (2u * (j & 15u))) & 3u;
            f = ((f << 2) | c) & pmask;
            r = (r >> 2) | ((3u - c) << rsh);
            if (j + 1u < (uint32_t)p) continue;
            // window minimum over blocks of w p-mers (van Herk / Gil-Werman): pm is the prefix
            // minimum of the current block; at a block end the block's ring slots are rewritten
            // with suffix minima, and a window that starts in the previous block is
            // min(suffix there, pm)
            bu = (bu + 1u == w) ? 0u : bu + 1u;
            uint32_t* const slot = rbase + (size_t)bu * T;
            const uint32_t sfx = slot[T];
            const uint32_t vf = pmer_valid(f, p) ? f : pmask;
            const uint32_t vr = pmer_valid(r, p) ? r : pmask;
            const uint32_t v = vf < vr ? vf : vr;
            pm = bu == 0u ? v : (v < pm ? v : pm);
            const bool bend = bu + 1u == w;
            const uint32_t mm = bend ? pm : (sfx < pm ? sfx : pm);
            slot[0] = v;
            if (bend) {
                uint32_t sm = 0xFFFFFFFFu;
                int sl = (int)w - 1;
                for (; sl >= 3; sl -= 4) {
                    uint32_t* const q = rbase + (size_t)sl * T;
                    uint32_t x0 = q[0], x1 = q[-T], x2 = q[-2 * T], x3 = q[-3 * T];
                    x0 = x0 < sm ? x0 : sm;
                    x1 = x1 < x0 ? x1 : x0;
                    x2 = x2 < x1 ? x2 : x1;
                    x3 = x3 < x2 ? x3 : x2;
                    q[0] = x0; q[-T] = x1; q[-2 * T] = x2; q[-3 * T] = x3;
                    sm = x3;
                }
                for (; sl >= 0; sl--) {
                    uint32_t* const q = rbase + (size_t)sl * T;
                    const uint32_t x = q[0] < sm ? q[0] : sm;
                    q[0] = x;
                    sm = x;
                }
            }
            if (j + 1u < (uint32_t)k) continue;
            // events, without branches: the entry is written to queue slot qn in every step and
            // kept (qn advanced) only when it is one
            bool ev, term;
            if (j + 1u == (uint32_t)k) {  // k-mer i0: a start when the segment begins the read, else the look-back
                ev = !done && s0 == 0u;
                term = false;
                open = ev;
            } else {
                const bool past = j >= js1;  // k-mer index >= s1
                const bool stop = !done && past && !open;
                const bool rend = !done && !stop && j >= jend;
                const bool chg = !done && !stop && !rend && mm != prev;
                ev = (rend && open) || chg;
                term = rend || (chg && past);
                open = open || (chg && !past);
                done = done || stop || rend || (chg && past);
            }
            myq[(size_t)qn * T] = term ? ((unsigned long long)(j | 0x80000000u) << 32)
                                       : ((unsigned long long)mm | ((unsigned long long)j << 32));
            qn += ev ? 1u : 0u;
            prev = mm;
        }
        const bool any_left = __ballot(!done) != 0ull;
        if (!any_left || __ballot(qn + (uint32_t)kEvChunk >= (uint32_t)kEvQ) != 0ull) drain();
        if (!any_left) break;
    }
    __syncthreads();
    if (a.skm_list && threadIdx.x == 0) {
        const bool over = lbad || lcnt > a.skm_cap;
        a.skm_n[blockIdx.x] = over ? a.skm_cap + 1u : lcnt;
        if (over) a.wg_rescan[atomicAdd(a.n_rescan, 1u)] = blockIdx.x;
    }
    const size_t stride = (size_t)a.n_wg + 1;
    for (int i = threadIdx.x; i < a.parts; i += T) {
        a.h_bytes[(size_t)i * stride + blockIdx.x] = hb[i];
        a.h_recs[(size_t)i * stride + blockIdx.x] = hr[i];
        a.h_kmers[(size_t)i * stride + blockIdx.x] = hk[i];
    }
}

// Write record bytes [0, nb) of superkmer [b, e) of a read into dst (cpu_funcs.cpp:238-244).
// Record byte x is packed-read byte m = b/3 + x: 11|b0|b1|b2 for a full group, the read's tail
// byte (len%3)<<6 with its bases high, or 0 when len%3 == 0; then the first and last indicators
// are patched. Ten bytes (30 bases) come from one 64-bit window of the 2-bit words.
__device__ __forceinline__ void write_record(const BaseSrc& src, uint64_t R0, uint32_t len, uint32_t b, uint32_t e, int k,
                                             uint32_t nb, uint8_t* dst) {
    const uint32_t m0 = b / 3u;
    const uint32_t last_ind = ((e - 1u + (uint32_t)k) % 3u) << 6;
    for (uint32_t x0 = 0; x0 < nb; x0 += 10u) {
        const uint64_t pos = R0 + 3ull * (m0 + x0);
        const uint64_t wi = pos >> 4;
        const uint32_t sh = 2u * (uint32_t)(pos & 15);
        uint64_t win = ((uint64_t)src.word(wi) | ((uint64_t)src.word(wi + 1) << 32)) >> sh;
        if (sh) win |= (uint64_t)src.word(wi + 2) << (64u - sh);
        const uint32_t xe = min(nb, x0 + 10u);
        for (uint32_t x = x0; x < xe; x++) {
            const uint32_t d = (uint32_t)(win >> (6u * (x - x0))) & 63u;  // b0 | b1 << 2 | b2 << 4
            const uint32_t grp = ((d & 3u) << 4) | (d & 0xCu) | (d >> 4);  // b0 << 4 | b1 << 2 | b2
            const uint32_t m = m0 + x;
            uint32_t byte;
            if (3u * m + 3u <= len) {
                byte = 0xC0u | grp;
            } else {
                const uint32_t rem = len - 3u * m;  // 0, 1 or 2 (the read's tail byte)
                byte = (rem << 6) | (grp & (rem == 2u ? 0x3Cu : rem == 1u ? 0x30u : 0u));
            }
            if (x == 0) byte = (byte & 0x3Fu) | ((3u - b % 3u) << 6);
            if (x == nb - 1u) byte = (byte & 0x3Fu) | last_ind;
            dst[x] = (uint8_t)byte;
        }
    }
}

// emit pass: records staged in LDS grouped by partition, then copied out in contiguous runs;
// side descriptors alike. EXACT mode takes the workgroup's offsets from the count pass prefix
// sums; STREAM mode (persistent, for batched input) counts its tile in LDS first, then reserves
// each partition's run with one global atomic against the partition's capacity.
constexpr int kEmitQ = 16;  // superkmers queued per lane before the wave writes them

template <int RING, int T>
struct EmitLds {
    uint32_t ring[RING * T];
    uint32_t lw[T * 14];
    uint8_t sbytes[192u * T];
    uint32_t sside[16u * T];
    uint64_t gbyte[kMaxParts], grec[kMaxParts];
    uint32_t soff[kMaxParts + 1], roff[kMaxParts + 1];
    // one cursor per partition: bytes in the high word, records in the low word, so a record's
    // side slot and its byte position are reserved together (same order within the partition)
    unsigned long long pcur[kMaxParts];
    uint32_t hb[kMaxParts], hr[kMaxParts], hk[kMaxParts];  // stream mode: tile counts
    uint32_t tmp[T / 64 + 1];
    uint32_t totals[2];
};

template <int RING, int T, bool STREAM>
__device__ void emit_tile(const P1Args& a, uint32_t wg, uint64_t n_seg, EmitLds<RING, T>& L) {
    constexpr uint32_t kStageB = 192u * T;
    constexpr uint32_t kSideN = 16u * T;
    const uint64_t seg_lo = (uint64_t)wg * T;
    const uint64_t seg_hi = min(n_seg, seg_lo + T);
    const uint64_t s = seg_lo + threadIdx.x;
    const int k = a.k;
    const uint32_t parts = (uint32_t)a.parts;
    const FastDivU32 pdiv = FastDivU32::make(parts);
    BaseSrc src;
    if (STREAM) {
        for (int i = threadIdx.x; i < a.parts; i += T) { L.hb[i] = 0; L.hr[i] = 0; L.hk[i] = 0; }
        __syncthreads();
        stage_bases<T>(a, L.lw, T * 14, src, seg_lo, seg_hi);
        if (s < seg_hi) {
            const uint32_t rd = a.seg_read[s];
            const uint64_t R0 = a.offs[rd];
            const uint32_t len = (uint32_t)(a.offs[rd + 1] - R0);
            scan_segment<RING, T>(src, L.ring, R0, len, a.seg_first[s], k, a.p, [&](uint32_t b, uint32_t e, uint32_t mm) {
                const uint32_t part = pdiv.mod(~mm);
                const uint32_t nb = skm_nbytes(b % 3u, e - b + (uint32_t)k - 1u);
                if (nb > 65535u) raise_err(a.err, RGKC_ERR_SKM_TOO_LONG);
                atomicAdd(&L.hb[part], nb);
                atomicAdd(&L.hr[part], 1u);
                atomicAdd(&L.hk[part], e - b);
            });
        }
        __syncthreads();
    }
    // per-partition bytes/records of this tile -> LDS staging offsets and global run starts
    const size_t stride = (size_t)a.n_wg + 1;
    for (int c0 = 0; c0 < a.parts; c0 += T) {
        const int i = c0 + threadIdx.x;
        uint32_t nbp = 0, nrp = 0;
        if (i < a.parts) {
            if (STREAM) {
                nbp = L.hb[i];
                nrp = L.hr[i];
                uint64_t bo = 0, ro = 0;
                if (nrp) {
                    // one packed cursor (bytes << kRecBits | records) reserves the byte run and the
                    // record run together, so runs of different workgroups keep the same order in
                    // the store and in the side descriptors
                    const unsigned long long old =
                        atomicAdd(&a.pos_cursor[i], ((unsigned long long)nbp << kRecBits) | (unsigned long long)nrp);
                    bo = old >> kRecBits;
                    ro = old & kRecMask;
                    if (ro + nrp > kRecMask) raise_err(a.err, RGKC_ERR_PART_OVERFLOW);  // record field would carry
                    atomicAdd(&a.kmer_cursor[i], (unsigned long long)L.hk[i]);
                }
                L.gbyte[i] = a.part_off[i] + bo;
                L.grec[i] = a.rec_off[i] + ro;
            } else {
                const uint32_t bo = a.wg_byte_off[(size_t)i * stride + wg];
                const uint32_t ro = a.wg_rec_off[(size_t)i * stride + wg];
                nbp = a.wg_byte_off[(size_t)i * stride + wg + 1] - bo;
                nrp = a.wg_rec_off[(size_t)i * stride + wg + 1] - ro;
                L.gbyte[i] = a.part_off[i] + bo;
                L.grec[i] = a.rec_off[i] + ro;
            }
            // the run of this tile must fit the partition's region; ~0 marks a run that does not
            if (nbp && L.gbyte[i] + nbp > a.part_end[i]) { L.gbyte[i] = ~0ull; raise_err(a.err, RGKC_ERR_PART_OVERFLOW); }
            if (nrp && L.grec[i] + nrp > a.rec_end[i]) { L.grec[i] = ~0ull; raise_err(a.err, RGKC_ERR_PART_OVERFLOW); }
            L.pcur[i] = 0;
        }
        uint32_t tb, tr;
        const uint32_t eb = wg_exscan_u32(nbp, L.tmp, &tb);
        const uint32_t er = wg_exscan_u32(nrp, L.tmp, &tr);
        const uint32_t cb = c0 == 0 ? 0u : L.soff[c0], cr = c0 == 0 ? 0u : L.roff[c0];
        __syncthreads();
        if (i < a.parts) { L.soff[i] = cb + eb; L.roff[i] = cr + er; }
        if (threadIdx.x == 0) { L.soff[min(c0 + T, a.parts)] = cb + tb; L.roff[min(c0 + T, a.parts)] = cr + tr; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { L.totals[0] = L.soff[a.parts]; L.totals[1] = L.roff[a.parts]; }
    __syncthreads();
    const bool stage_b = L.totals[0] <= kStageB;
    const bool stage_r = L.totals[1] <= kSideN;
    if (!STREAM) stage_bases<T>(a, L.lw, T * 14, src, seg_lo, seg_hi);
    auto emit_one = [&](uint64_t R0, uint32_t len, uint32_t b, uint32_t e, uint32_t mm) __attribute__((always_inline)) {
        const uint32_t nmm = ~mm;
        const uint32_t q = pdiv.div(nmm);
        const uint32_t part = nmm - q * parts;
        const uint32_t sub = q % (uint32_t)kSubIds;
        const uint32_t kmers = e - b;
        const uint32_t nb = skm_nbytes(b % 3u, kmers + (uint32_t)k - 1u);
        const uint32_t sval = side_pack(sub, kmers > 0x3FFFFu ? 0x3FFFFu : kmers, b % 3u);
        const unsigned long long old = atomicAdd(&L.pcur[part], ((unsigned long long)nb << 32) | 1ull);
        const uint32_t slot = (uint32_t)old;
        const uint32_t pos = (uint32_t)(old >> 32);
        if (stage_r) {
            L.sside[L.roff[part] + slot] = sval;
        } else if (L.grec[part] != ~0ull) {
            a.side[L.grec[part] + slot] = sval;
        }
        if (stage_b) {
            write_record(src, R0, len, b, e, k, nb, L.sbytes + L.soff[part] + pos);
        } else if (L.gbyte[part] != ~0ull) {
            write_record(src, R0, len, b, e, k, nb, a.store + L.gbyte[part] + pos);
        }
    };
    if (s < seg_hi) {
        const uint32_t rd = a.seg_read[s];
        const uint64_t R0 = a.offs[rd];
        const uint32_t len = (uint32_t)(a.offs[rd + 1] - R0);
        // Superkmers wait in a per-lane queue of kEmitQ entries; the wave writes them out together
        // whenever some lane's queue is full, so the write path runs for many lanes at once instead
        // of for the few lanes that close a superkmer at a given base step.
        uint32_t qb[kEmitQ] = {}, qe[kEmitQ] = {}, qm[kEmitQ] = {};  // slot 0 is the newest entry
        uint32_t qn = 0;
        auto flush = [&]() __attribute__((always_inline)) {
            for (uint32_t t = 0; t < qn; t++) {  // newest first; the queue shifts down
                emit_one(R0, len, qb[0], qe[0], qm[0]);
#pragma unroll
                for (int u = 0; u + 1 < kEmitQ; u++) { qb[u] = qb[u + 1]; qe[u] = qe[u + 1]; qm[u] = qm[u + 1]; }
            }
            qn = 0;
        };
        auto push = [&](uint32_t b, uint32_t e, uint32_t mm) __attribute__((always_inline)) {
            if (qn == (uint32_t)kEmitQ) flush();
#pragma unroll
            for (int u = kEmitQ - 1; u > 0; u--) { qb[u] = qb[u - 1]; qe[u] = qe[u - 1]; qm[u] = qm[u - 1]; }
            qb[0] = b; qe[0] = e; qm[0] = mm;
            qn++;
        };
        scan_segment<RING, T>(src, L.ring, R0, len, a.seg_first[s], k, a.p, push, [&]() __attribute__((always_inline)) {
            if (__any(qn == (uint32_t)kEmitQ)) flush();
        });
        flush();
    }
    __syncthreads();
    // copy staged runs out, one wave per partition at a time
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int i = wave; i < a.parts; i += T / 64) {
        if (stage_b) {
            const uint32_t n = L.soff[i + 1] - L.soff[i];
            if (n && L.gbyte[i] != ~0ull) {
                uint8_t* dst = a.store + L.gbyte[i];
                const uint8_t* srcb = L.sbytes + L.soff[i];
                for (uint32_t x = lane; x < n; x += 64) dst[x] = srcb[x];
            }
        }
        if (stage_r) {
            const uint32_t n = L.roff[i + 1] - L.roff[i];
            if (n && L.grec[i] != ~0ull) {
                uint32_t* dst = a.side + L.grec[i];
                const uint32_t* srcs = L.sside + L.roff[i];
                for (uint32_t x = lane; x < n; x += 64) dst[x] = srcs[x];
            }
        }
    }
    __syncthreads();
}

template <int RING, int T>
__global__ __launch_bounds__(T) void k_p1_emit(P1Args a) {
    __shared__ EmitLds<RING, T> L;
    emit_tile<RING, T, false>(a, a.wg_map ? a.wg_map[blockIdx.x] : blockIdx.x, a.n_seg, L);
}

// Hardware places workgroup b on XCD b % 8: give each XCD one contiguous range of tiles, so the
// partition runs of neighbouring tiles (adjacent in the store) meet in one L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n) {
    const uint32_t x = b % 8u, q = n / 8u, r = n % 8u;
    return x * q + min(x, r) + b / 8u;
}

// Emit from the count pass's superkmer lists (exact mode): one workgroup per count-pass tile whose
// list fit. Each list entry is one record: its slot in the tile's partition run comes from an LDS
// cursor, and its side descriptor and bytes go straight to their place in HBM (the tile's runs
// hold its records only, so nothing is staged). A small LDS footprint keeps several tiles per CU.
// LDS of the list emit is sized so that three 256-thread workgroups share a CU (~51 KB each with
// partition tables for up to 512 partitions)
constexpr uint32_t kListStageW = 2560;  // staged 2-bit words per tile (the rest read from HBM)
constexpr uint32_t kReorderMax = 3840;  // list entries a tile can put in store order (more: list order)

template <int T, int PMAX>  // PMAX: partition-table size (parts <= PMAX)
__global__ __launch_bounds__(T) void k_p1_emit_list(P1Args a) {
    __shared__ uint32_t lw[kListStageW];
    __shared__ uint64_t seg_r0[T];
    __shared__ uint32_t seg_len[T], seg_s0[T];
    __shared__ uint64_t gbyte[PMAX], grec[PMAX];
    __shared__ unsigned long long pcur[PMAX];
    // store order of the tile's records: roff[p] = records of partitions < p, then each entry's
    // byte position in its run and the entry at each store-order place
    __shared__ uint32_t roff[PMAX];
    __shared__ uint32_t rpos[kReorderMax];
    __shared__ uint16_t order[kReorderMax];
    __shared__ uint32_t tmp[T / 64 + 1];
    const uint32_t wg = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t n_list = a.skm_n[wg];
    if (n_list > a.skm_cap) return;  // rescanned by the scanning emit
    const int k = a.k;
    const uint32_t parts = (uint32_t)a.parts;
    const FastDivU32 pdiv = FastDivU32::make(parts);
    const size_t stride = (size_t)a.n_wg + 1;
    for (uint32_t i = threadIdx.x; i < parts; i += T) {
        uint64_t bo, ro;
        uint32_t nbp, nrp;
        if (a.list_reserve) {
            // capacity-sized regions: the tile's counts, one packed cursor atomic per partition
            // (byte run and record run reserved together, as in the scanning emit)
            nbp = a.h_bytes[(size_t)i * stride + wg];
            nrp = a.h_recs[(size_t)i * stride + wg];
            bo = ro = 0;
            if (nrp) {
                const unsigned long long old =
                    atomicAdd(&a.pos_cursor[i], ((unsigned long long)nbp << kRecBits) | (unsigned long long)nrp);
                bo = old >> kRecBits;
                ro = old & kRecMask;
                if (ro + nrp > kRecMask) raise_err(a.err, RGKC_ERR_PART_OVERFLOW);  // record field would carry
                atomicAdd(&a.kmer_cursor[i], (unsigned long long)a.h_kmers[(size_t)i * stride + wg]);
            }
        } else {
            bo = a.wg_byte_off[(size_t)i * stride + wg];
            ro = a.wg_rec_off[(size_t)i * stride + wg];
            nbp = a.wg_byte_off[(size_t)i * stride + wg + 1] - (uint32_t)bo;
            nrp = a.wg_rec_off[(size_t)i * stride + wg + 1] - (uint32_t)ro;
        }
        uint64_t gb = a.part_off[i] + bo, gr = a.rec_off[i] + ro;
        // the run of this tile must fit the partition's region; ~0 marks a run that does not
        if (nbp && gb + nbp > a.part_end[i]) { gb = ~0ull; raise_err(a.err, RGKC_ERR_PART_OVERFLOW); }
        if (nrp && gr + nrp > a.rec_end[i]) { gr = ~0ull; raise_err(a.err, RGKC_ERR_PART_OVERFLOW); }
        gbyte[i] = gb;
        grec[i] = gr;
        pcur[i] = 0;
        roff[i] = nrp;
    }
    const bool reorder = n_list <= kReorderMax;
    if (reorder) {
        __syncthreads();
        uint32_t carry = 0;
        for (uint32_t c0 = 0; c0 < parts; c0 += T) {
            const uint32_t i = c0 + threadIdx.x;
            const uint32_t v = i < parts ? roff[i] : 0u;
            uint32_t tot;
            const uint32_t ex = wg_exscan_u32(v, tmp, &tot);
            if (i < parts) roff[i] = carry + ex;
            carry += tot;
        }
    }
    const uint64_t seg_lo = (uint64_t)wg * T;
    const uint64_t seg_hi = min(a.n_seg, seg_lo + T);
    const uint64_t s = seg_lo + threadIdx.x;
    if (s < seg_hi) {
        const uint32_t rd = a.seg_read[s];
        const uint64_t R0 = a.offs[rd];
        seg_r0[threadIdx.x] = R0;
        seg_len[threadIdx.x] = (uint32_t)(a.offs[rd + 1] - R0);
        seg_s0[threadIdx.x] = a.seg_first[s];
    }
    BaseSrc src;
    stage_bases<T>(a, lw, kListStageW, src, seg_lo, seg_hi);  // ends with a barrier
    const uint64_t* lst = a.skm_list + (size_t)wg * a.skm_cap;
    if (reorder) {
        // pass 1: slots; pass 2: lane i writes the tile's i-th record in store order, so the
        // lanes of one store cover neighbouring records instead of 64 unrelated places. Every
        // store-order place is filled when the count pass's histograms match its list; a place
        // left at the 0xFFFF mark means they do not, and is reported instead of read.
        for (uint32_t i = threadIdx.x; i < n_list; i += T) order[i] = 0xFFFFu;
        __syncthreads();
        for (uint32_t x = threadIdx.x; x < n_list; x += T) {
            const uint64_t v = lst[x];
            const uint32_t t = (uint32_t)(v >> 32) & 0xFFu;
            const uint32_t b = seg_s0[t] + ((uint32_t)(v >> 40) & 0x7Fu);
            if ((uint64_t)b + (v >> 47) + (uint64_t)k - 1u > seg_len[t]) {  // not a superkmer of its read
                raise_err(a.err, RGKC_ERR_BAD_STORE);
                rpos[x] = 0;
                continue;
            }
            const uint32_t part = pdiv.mod(~(uint32_t)v);
            const uint32_t nb = skm_nbytes(b % 3u, (uint32_t)(v >> 47) + (uint32_t)k - 1u);
            const unsigned long long old = atomicAdd(&pcur[part], ((unsigned long long)nb << 32) | 1ull);
            rpos[x] = (uint32_t)(old >> 32);
            const uint32_t oi = roff[part] + (uint32_t)old;  // < n_list: the count pass counted these records
            if (oi < n_list) order[oi] = (uint16_t)x;
            else raise_err(a.err, RGKC_ERR_BAD_STORE);
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n_list; i += T) {
            const uint32_t x = order[i];
            if (x >= n_list) {  // an unfilled place (0xFFFF): histograms and list disagree
                raise_err(a.err, RGKC_ERR_BAD_STORE);
                continue;
            }
            const uint64_t v = lst[x];
            const uint32_t t = (uint32_t)(v >> 32) & 0xFFu;
            const uint32_t b = seg_s0[t] + ((uint32_t)(v >> 40) & 0x7Fu);
            const uint32_t kmers = (uint32_t)(v >> 47);
            if ((uint64_t)b + kmers + (uint64_t)k - 1u > seg_len[t]) continue;  // reported by pass 1
            const uint32_t nmm = ~(uint32_t)v;
            const uint32_t q = pdiv.div(nmm);
            const uint32_t part = nmm - q * parts;
            const uint32_t sub = q % (uint32_t)kSubIds;
            const uint32_t nb = skm_nbytes(b % 3u, kmers + (uint32_t)k - 1u);
            if (grec[part] != ~0ull)
                a.side[grec[part] + (i - roff[part])] = side_pack(sub, kmers > 0x3FFFFu ? 0x3FFFFu : kmers, b % 3u);
            if (gbyte[part] != ~0ull)
                write_record(src, seg_r0[t], seg_len[t], b, b + kmers, k, nb, a.store + gbyte[part] + rpos[x]);
        }
        return;
    }
    if (threadIdx.x == 0 && a.n_listorder) atomicAdd(a.n_listorder, 1u);  // tile written in list order
    for (uint32_t x = threadIdx.x; x < n_list; x += T) {
        const uint64_t v = lst[x];
        const uint32_t t = (uint32_t)(v >> 32) & 0xFFu;
        const uint32_t b = seg_s0[t] + ((uint32_t)(v >> 40) & 0x7Fu);
        const uint32_t kmers = (uint32_t)(v >> 47);
        if ((uint64_t)b + kmers + (uint64_t)k - 1u > seg_len[t]) {  // not a superkmer of its read
            raise_err(a.err, RGKC_ERR_BAD_STORE);
            continue;
        }
        const uint32_t nmm = ~(uint32_t)v;
        const uint32_t q = pdiv.div(nmm);
        const uint32_t part = nmm - q * parts;
        const uint32_t sub = q % (uint32_t)kSubIds;
        const uint32_t nb = skm_nbytes(b % 3u, kmers + (uint32_t)k - 1u);
        const unsigned long long old = atomicAdd(&pcur[part], ((unsigned long long)nb << 32) | 1ull);
        if (grec[part] != ~0ull)
            a.side[grec[part] + (uint32_t)old] = side_pack(sub, kmers > 0x3FFFFu ? 0x3FFFFu : kmers, b % 3u);
        if (gbyte[part] != ~0ull)
            write_record(src, seg_r0[t], seg_len[t], b, b + kmers, k, nb, a.store + gbyte[part] + (uint32_t)(old >> 32));
    }
}

// Persistent: tiles of the batch whose segment count is read from device memory.
template <int RING, int T>
__global__ __launch_bounds__(T) void k_p1_emit_stream(P1Args a) {
    __shared__ EmitLds<RING, T> L;
    const uint64_t n_seg = *a.n_seg_dev;
    const uint64_t tiles = a.wg_map ? a.n_map : (n_seg + T - 1) / T;
    for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x)
        emit_tile<RING, T, true>(a, a.wg_map ? a.wg_map[t] : (uint32_t)t, n_seg, L);
}

}  // namespace

hipError_t launch_pack(const uint8_t* reads, uint64_t n_bytes, uint32_t* bases2, hipStream_t s) {
    const uint64_t n_words = (n_bytes >> 4) + 2;  // reads buffer is padded to (n_words * 16) zero bytes
    hipLaunchKernelGGL(k_pack, dim3((unsigned)((n_words + 255) / 256)), dim3(256), 0, s, (const uint4*)reads, n_words,
                       bases2);
    return hipGetLastError();
}

hipError_t launch_seg_count(const uint64_t* offs, uint64_t n_reads, int k, uint64_t* nseg, hipStream_t s) {
    if (!n_reads) return hipSuccess;
    hipLaunchKernelGGL(k_seg_count, dim3((unsigned)((n_reads + 255) / 256)), dim3(256), 0, s, offs, n_reads, k, nseg);
    return hipGetLastError();
}

hipError_t launch_seg_fill(const uint64_t* offs, const uint64_t* seg_off, uint64_t n_reads, int k, uint32_t* seg_read,
                           uint32_t* seg_first, hipStream_t s) {
    if (!n_reads) return hipSuccess;
    hipLaunchKernelGGL(k_seg_fill, dim3((unsigned)((n_reads + 255) / 256)), dim3(256), 0, s, offs, seg_off, n_reads, k,
                       seg_read, seg_first);
    return hipGetLastError();
}

// Tiles of 256 segments for windows w = k - p + 1 <= 32, of 128 for larger ones (the emit passes
// use the same tiles).
hipError_t launch_p1_count(const P1Args& a, hipStream_t s) {
    if (!a.n_seg) return hipSuccess;
    const int w = a.k - a.p + 1, T = w <= 32 ? 256 : 128;
    const size_t dyn = (((size_t)(w + 1) * T + 1) & ~(size_t)1) * 4 + (size_t)kEvQ * T * 8 + 3 * (size_t)a.parts * 4;
    if (w <= 32)
        hipLaunchKernelGGL((k_p1_count<256>), dim3(a.n_wg), dim3(256), dyn, s, a);
    else
        hipLaunchKernelGGL((k_p1_count<128>), dim3(a.n_wg), dim3(128), dyn, s, a);
    return hipGetLastError();
}

hipError_t launch_batch_prep(const uint8_t* reads, const uint64_t* offs, uint64_t* hdr, uint64_t cap_reads,
                             uint64_t cap_bases, int k, uint32_t* bases2, uint64_t* nseg, uint64_t* seg_off,
                             uint64_t* scan_tmp, uint32_t* seg_read, uint32_t* seg_first, hipStream_t s) {
    const uint64_t cap_words = (cap_bases >> 4) + 2;
    hipLaunchKernelGGL(k_pack_dev, dim3((unsigned)std::min<uint64_t>((cap_words + 255) / 256, 4096)), dim3(256), 0, s,
                       (const uint4*)reads, (const uint64_t*)hdr, cap_words, bases2);
    hipError_t e = hipGetLastError();
    if (e) return e;
    const unsigned g = (unsigned)std::min<uint64_t>((cap_reads + 256) / 256, 4096);
    hipLaunchKernelGGL(k_seg_count_dev, dim3(g), dim3(256), 0, s, offs, (const uint64_t*)hdr, cap_reads, k, nseg);
    if ((e = hipGetLastError())) return e;
    if ((e = scan_u64(nseg, seg_off, cap_reads + 1, scan_tmp, s))) return e;
    hipLaunchKernelGGL(k_seg_fill_dev, dim3(g), dim3(256), 0, s, offs, (const uint64_t*)seg_off, hdr, cap_reads, k,
                       seg_read, seg_first);
    return hipGetLastError();
}

// The list emit over all tiles, with partition tables sized for a.parts.
static hipError_t launch_list(const P1Args& a, bool big, hipStream_t s) {
    const bool small = a.parts <= 512;
    if (!big) {
        if (small) hipLaunchKernelGGL((k_p1_emit_list<256, 512>), dim3(a.n_wg), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_p1_emit_list<256, kMaxParts>), dim3(a.n_wg), dim3(256), 0, s, a);
    } else {
        if (small) hipLaunchKernelGGL((k_p1_emit_list<128, 512>), dim3(a.n_wg), dim3(128), 0, s, a);
        else hipLaunchKernelGGL((k_p1_emit_list<128, kMaxParts>), dim3(a.n_wg), dim3(128), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_p1_emit_sized(const P1Args& a, uint32_t n_rescan, hipStream_t s) {
    if (!a.n_seg) return hipSuccess;
    const bool big = (a.k - a.p + 1) > 32;
    hipError_t e = launch_list(a, big, s);
    if (e != hipSuccess || n_rescan == 0) return e;
    P1Args b = a;
    b.n_map = n_rescan;  // wg_map holds the tiles to rescan
    const unsigned grid = std::min(n_rescan, 1024u);
    if (!big) hipLaunchKernelGGL((k_p1_emit_stream<32, 256>), dim3(grid), dim3(256), 0, s, b);
    else hipLaunchKernelGGL((k_p1_emit_stream<64, 128>), dim3(grid), dim3(128), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_p1_emit_stream(const P1Args& a, unsigned grid, hipStream_t s) {
    if ((a.k - a.p + 1) <= 32)
        hipLaunchKernelGGL((k_p1_emit_stream<32, 256>), dim3(grid), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_p1_emit_stream<64, 128>), dim3(grid), dim3(128), 0, s, a);
    return hipGetLastError();
}

// Exact mode. With superkmer lists: the list kernel over all tiles, then the scanning kernel over
// the n_rescan tiles whose list overflowed (a.wg_map); without: the scanning kernel over all tiles.
hipError_t launch_p1_emit(const P1Args& a, uint32_t n_rescan, hipStream_t s) {
    if (!a.n_seg) return hipSuccess;
    const bool big = (a.k - a.p + 1) > 32;
    if (a.skm_list) {
        hipError_t e = launch_list(a, big, s);
        if (e != hipSuccess || n_rescan == 0) return e;
    }
    const unsigned grid = a.skm_list ? n_rescan : a.n_wg;
    if (!big)
        hipLaunchKernelGGL((k_p1_emit<32, 256>), dim3(grid), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_p1_emit<64, 128>), dim3(grid), dim3(128), 0, s, a);
    return hipGetLastError();
}

}  // namespace rgkc

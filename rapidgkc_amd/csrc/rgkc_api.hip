// C ABI of the gfx950 k-mer counting path: contexts, device buffers and stage orchestration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rgkc.h"
#include "rgkc_device.h"
#include "rgkc_internal.h"

using namespace rgkc;

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t need(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        if (p) { (void)hipFree(p); p = nullptr; cap = 0; }
        const size_t b = std::max<size_t>(bytes, 256);
        hipError_t e = hipMalloc(&p, b);
        if (e == hipSuccess) cap = b;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T> T* as() const { return (T*)p; }
};

// Working memory of rgkc_sort_table_noalloc: a fixed device area per GPU, held from the library's
// load until the process ends. Host code never uses the array's name as an address: it resolves
// each device's copy once (hipGetSymbolAddress, at context creation with that device current) and
// passes that address. Layout: scratch rows, segment descriptors, chunk maps, histograms + cursors.
constexpr size_t kNoAllocRowBytes = RGKC_NOALLOC_ROW_BYTES;
constexpr size_t kNoAllocSegBytes = 1u << 20;
constexpr size_t kNoAllocChunkBytes = 1u << 20;
constexpr size_t kNoAllocHistBytes = 14u << 20;
constexpr size_t kNoAllocBytes = kNoAllocRowBytes + kNoAllocSegBytes + kNoAllocChunkBytes + kNoAllocHistBytes;
static_assert(kNoAllocBytes == RGKC_NOALLOC_SCRATCH_BYTES, "header and library agree on the area");
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

// Address of `device`'s copy of the area (nullptr if it could not be resolved); `device` must be
// current on the first call for it.
uint8_t* device_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_done[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            ptr = nullptr;
        }
        g_scratch_ptr[device] = (uint8_t*)ptr;
        g_scratch_done[device] = true;
    }
    return g_scratch_ptr[device];
}

enum Ev { EV_START, EV_PACK, EV_SEG, EV_P1C, EV_P1E, EV_CB, EV_IDX, EV_PLAN, EV_EXP, EV_CNT, EV_N };

}  // namespace

struct rgkc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    hipEvent_t ev[EV_N] = {};
    bool ev_rec[EV_N] = {};
    // resident read batch
    DevBuf reads, offs, bases2;
    uint64_t n_reads = 0, n_bases = 0;
    bool reads_ok = false;
    // segments
    DevBuf nseg, seg_off, seg_read, seg_first, scan_tmp;
    uint64_t n_seg = 0;
    // phase 1
    DevBuf hist;  // 3 x parts x (n_wg+1) uint32
    DevBuf skm_list, skm_n;  // phase 1: superkmer lists of the count pass (n_wg x skm_cap entries)
    DevBuf totals;  // 3 x parts uint64
    DevBuf poffs;   // part_off, part_end, rec_off, rec_end, row_base, row_end: 6 x parts uint64
    DevBuf store, side, scratch;
    DevBuf err_word, stats_dev, cursors;
    int k = 0, p = 0, parts = 0;
    std::vector<uint64_t> h_part_off, h_rec_off, h_kmers;  // parts+1 prefix arrays
    bool store_ok = false;
    // buckets
    DevBuf b_byte, b_rec, b_part;
    DevBuf sub_start, g_hist, g_seg_bytes, sub_sinst, item_cnt, item_off, items, scan_tmp2, words, b_word;
    DevBuf key_cnt, key_off, exp_cnt, exp_off, sub_key_off, heavy_keys, exp_items;  // heavy sub-buckets
    DevBuf item_desc;  // count: first-sweep descriptor of each work item
    DevBuf sub_item0, sub_key_cs, ovf_n, ovf_off, ovf_keys;  // count: exact regions of overflowed slices
    uint32_t n_buckets = 0;
    uint64_t n_items = 0;
    // table
    DevBuf keys, counts;
    DevBuf tkeys, tcounts, sort_segs, sort_chunks, sort_hist;  // row sort: scratch rows and plans
    std::vector<uint64_t> h_row_base, h_rows;
    // per partition: the table holding its rows (0: keys/counts, 1: keys2/counts2, the predicted
    // path's recounted partitions) and its first row there
    std::vector<uint8_t> h_row_buf;
    std::vector<uint64_t> h_row_start;
    DevBuf keys2, counts2, part_distinct;
    bool table_ok = false;
    uint32_t inst_cap = 0;
    uint32_t count_wg_per_cu = 0;  // count kernel: resident workgroups per CU (0: four)
    uint64_t mem_limit = 0;  // count stage: device memory to plan with (0: what is free)
    bool compact_active = false;  // a count in compact mode holds its temporary table in tkeys/tcounts
    bool recount_active = false;  // the predicted path's recount writes keys2/counts2
    uint32_t list_cap = 0;  // phase-1 superkmer list entries per workgroup (0 default, ~0 off)
    int heavy_mode = 0;       // 0: expand heavy sub-buckets into keys when memory allows, 1: never
    // learned sizing: predicted distinct k-mers per partition (empty = exact sizing)
    std::vector<double> pred_distinct;
    float pred_margin = 0.1f;
    // stream mode (batched input into capacity-sized partition regions)
    bool stream_on = false, stream_graph = false, stream_ended = false;
    std::vector<uint64_t> h_used_bytes;  // bytes used in each stream region
    uint64_t cap_reads = 0, cap_bases = 0, n_batches = 0;
    DevBuf hdr, cursors3, b_ends;
    // two pinned staging sets (header, reads, offsets) used by alternate batches, so the host copies
    // the next batch while the device still works on the current one; stage_ev[i] marks the end of
    // the uploads from set i
    uint64_t* h_hdr[2] = {nullptr, nullptr};
    uint8_t* h_stage[2] = {nullptr, nullptr};
    uint64_t* h_offs_stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    bool stage_ev_rec[2] = {false, false};
    uint64_t last_batch_bases = 0;   // bases of the batch now in the device read buffer
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    P1Args sargs{};
    rgkc_stats st{};
    uint32_t h_listorder = 0;  // list-order tile count of the last phase-1 emit (read after a stream sync)
};

namespace {

int fail(rgkc_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    c->err = buf;
    return code;
}

#define HIPC(c, call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return fail((c), RGKC_E_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)
// Record a stage event; elapsed times are only read between events recorded in this run.
#define REC(c, e, s)                                   \
    do {                                               \
        HIPC(c, hipEventRecord((c)->ev[e], (s)));      \
        (c)->ev_rec[e] = true;                         \
    } while (0)
#define NEED(c, buf, bytes)                                                                     \
    do {                                                                                        \
        if ((buf).need(bytes) != hipSuccess)                                                    \
            return fail((c), RGKC_E_NOMEM, "device allocation of %zu bytes failed", (size_t)(bytes)); \
    } while (0)

// Count-stage allocations: the phase-1 superkmer lists and histograms are no longer needed once the
// store is written (they are only kept for reuse by the next batch), so an allocation that fails
// releases them and tries once more.
hipError_t need_count(rgkc_ctx* c, DevBuf& b, size_t bytes);
#define NEEDC(c, buf, bytes)                                                                    \
    do {                                                                                        \
        if (need_count((c), (buf), (bytes)) != hipSuccess)                                      \
            return fail((c), RGKC_E_NOMEM, "device allocation of %zu bytes failed", (size_t)(bytes)); \
    } while (0)

int p1_threads(int k, int p) { return (k - p + 1) <= 32 ? 256 : 128; }

// Phase-1 superkmer lists (unless switched off): room for 3x the superkmers a random sequence
// gives (2 / (w + 1) per k-mer) per workgroup; a workgroup with more is rescanned by the emit pass.
static int setup_lists(rgkc_ctx* c, P1Args& a, int T1, uint32_t n_wg) {
    if (c->list_cap == 0xFFFFFFFFu) return 0;
    hipStream_t s = c->stream;
    const uint64_t w = (uint64_t)(a.k - a.p + 1);
    uint64_t lcap = (uint64_t)T1 * kSegKmers * 3 / (w + 1);
    lcap = std::min<uint64_t>(std::max<uint64_t>((lcap + 63) / 64 * 64, 256), (uint64_t)T1 * kSegKmers);
    if (c->list_cap) lcap = std::min<uint64_t>(c->list_cap, (uint64_t)T1 * kSegKmers);
    const size_t lbytes = (size_t)n_wg * lcap * 8;
    if (lbytes > c->skm_list.cap) {
        // lists only when they fit in half the free device memory; otherwise the emit pass scans
        size_t fr = 0, tt = 0;
        if (hipMemGetInfo(&fr, &tt) != hipSuccess || lbytes > fr / 2) {
            (void)hipGetLastError();
            return 0;
        }
    }
    NEED(c, c->skm_list, lbytes);
    NEED(c, c->skm_n, ((size_t)2 * n_wg + 2) * 4);  // counts, rescan tiles, rescan count, list-order tiles
    HIPC(c, hipMemsetAsync(c->skm_n.as<uint32_t>() + 2 * (size_t)n_wg, 0, 8, s));
    a.skm_list = c->skm_list.as<uint64_t>();
    a.skm_n = c->skm_n.as<uint32_t>();
    a.skm_cap = (uint32_t)lcap;
    a.wg_rescan = a.skm_n + n_wg;
    a.n_rescan = a.skm_n + 2 * (size_t)n_wg;
    a.n_listorder = a.n_rescan + 1;
    return 0;
}

hipError_t need_count(rgkc_ctx* c, DevBuf& b, size_t bytes) {
    if (b.need(bytes) == hipSuccess) return hipSuccess;
    (void)hipGetLastError();
    // buffers no kernel of this count uses: the phase-1 lists and histograms, the temporary table
    // (outside compact mode) and the second table (before the recount)
    std::vector<DevBuf*> spare = {&c->skm_list, &c->hist};
    if (!c->compact_active) { spare.push_back(&c->tkeys); spare.push_back(&c->tcounts); }
    if (!c->recount_active) { spare.push_back(&c->keys2); spare.push_back(&c->counts2); }
    bool any = false;
    for (DevBuf* d : spare) any = any || (d != &b && d->p);
    if (!any) return hipErrorOutOfMemory;
    (void)hipStreamSynchronize(c->stream);  // kernels may still read them
    for (DevBuf* d : spare)
        if (d != &b) d->release();
    hipError_t e = b.need(bytes);
    if (e != hipSuccess) (void)hipGetLastError();
    return e;
}

int check_err_word(rgkc_ctx* c) {
    uint32_t e = 0;
    HIPC(c, hipMemcpyAsync(&e, c->err_word.p, 4, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    c->st.err_bits |= e;
    c->st.p1_list_order = c->h_listorder;
    if (e & (RGKC_ERR_PART_OVERFLOW | RGKC_ERR_ROW_OVERFLOW))
        return fail(c, RGKC_E_OVERFLOW, "region overflow (error bits 0x%x)", e);
    if (e & ~RGKC_ERR_STAGE_OVERFLOW) return fail(c, RGKC_E_KERNEL, "kernel error bits 0x%x", e);
    return 0;
}

// Phase 1 on the resident batch (mm_hist: optional device histogram for sample features).
int do_phase1(rgkc_ctx* c, int k, int p, int parts, uint32_t* mm_hist = nullptr) {
    if (!c->reads_ok) return fail(c, RGKC_E_STATE, "no read batch uploaded");
    if (k < 8 || k > 63 || p < 3 || p > 16 || p > k) return fail(c, RGKC_E_ARG, "unsupported k=%d p=%d", k, p);
    if (parts < 1 || parts > kMaxParts) return fail(c, RGKC_E_ARG, "parts must be in [1, %d]", kMaxParts);
    if (k - p + 1 > 64) return fail(c, RGKC_E_ARG, "window k-p+1 must be <= 64");
    hipStream_t s = c->stream;
    c->k = k; c->p = p; c->parts = parts;
    c->store_ok = false; c->table_ok = false; c->stream_ended = false;
    c->st = rgkc_stats{};
    c->st.n_reads = c->n_reads; c->st.n_bases = c->n_bases;
    NEED(c, c->err_word, 64);
    HIPC(c, hipMemsetAsync(c->err_word.p, 0, 64, s));
    for (auto& r : c->ev_rec) r = false;
    REC(c, EV_START, s);
    // (a) pack
    HIPC(c, launch_pack(c->reads.as<uint8_t>(), c->n_bases, c->bases2.as<uint32_t>(), s));
    REC(c, EV_PACK, s);
    // segments
    const uint64_t n = c->n_reads;
    NEED(c, c->nseg, (n + 1) * 8);
    NEED(c, c->seg_off, (n + 1) * 8);
    NEED(c, c->scan_tmp, scan_tmp_words(n + 1) * 8);
    HIPC(c, hipMemsetAsync(c->nseg.p, 0, (n + 1) * 8, s));
    HIPC(c, launch_seg_count(c->offs.as<uint64_t>(), n, k, c->nseg.as<uint64_t>(), s));
    HIPC(c, scan_u64(c->nseg.as<uint64_t>(), c->seg_off.as<uint64_t>(), n + 1, c->scan_tmp.as<uint64_t>(), s));
    uint64_t nseg_total = 0;
    HIPC(c, hipMemcpyAsync(&nseg_total, c->seg_off.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, s));
    HIPC(c, hipStreamSynchronize(s));
    if (nseg_total > 0xFFFFFFFFull) return fail(c, RGKC_E_ARG, "too many read segments in one batch");
    c->n_seg = nseg_total;
    NEED(c, c->seg_read, (nseg_total + 1) * 4);
    NEED(c, c->seg_first, (nseg_total + 1) * 4);
    HIPC(c, launch_seg_fill(c->offs.as<uint64_t>(), c->seg_off.as<uint64_t>(), n, k, c->seg_read.as<uint32_t>(),
                            c->seg_first.as<uint32_t>(), s));
    REC(c, EV_SEG, s);
    // (b) count pass
    const int T1 = p1_threads(k, p);
    const uint32_t n_wg = (uint32_t)((nseg_total + T1 - 1) / T1);
    const size_t cols = (size_t)n_wg + 1;
    const size_t hwords = (size_t)parts * cols;
    NEED(c, c->hist, 3 * hwords * 4);
    NEED(c, c->totals, 3 * (size_t)parts * 8);
    HIPC(c, hipMemsetAsync(c->hist.p, 0, 3 * hwords * 4, s));
    P1Args a{};
    a.bases2 = c->bases2.as<uint32_t>();
    a.offs = c->offs.as<uint64_t>();
    a.seg_read = c->seg_read.as<uint32_t>();
    a.seg_first = c->seg_first.as<uint32_t>();
    a.n_seg = nseg_total;
    a.n_bases = c->n_bases;
    a.k = k; a.p = p; a.parts = parts;
    a.h_bytes = c->hist.as<uint32_t>();
    a.h_recs = a.h_bytes + hwords;
    a.h_kmers = a.h_recs + hwords;
    a.n_wg = n_wg;
    a.err = c->err_word.as<uint32_t>();
    a.mm_hist = mm_hist;
    a.hist_shift = p > 8 ? 2 * (p - 8) : 0;
    if (int r = setup_lists(c, a, T1, n_wg)) return r;
    HIPC(c, launch_p1_count(a, s));
    a.mm_hist = nullptr;
    uint64_t* tot = c->totals.as<uint64_t>();
    HIPC(c, rowscan_u32(a.h_bytes, tot, parts, (uint32_t)cols, s));
    HIPC(c, rowscan_u32(a.h_recs, tot + parts, parts, (uint32_t)cols, s));
    HIPC(c, rowscan_u32(a.h_kmers, tot + 2 * parts, parts, (uint32_t)cols, s));
    REC(c, EV_P1C, s);
    std::vector<uint64_t> ht(3 * (size_t)parts);
    HIPC(c, hipMemcpyAsync(ht.data(), tot, ht.size() * 8, hipMemcpyDeviceToHost, s));
    uint32_t n_rescan = 0;
    if (a.skm_list) HIPC(c, hipMemcpyAsync(&n_rescan, a.n_rescan, 4, hipMemcpyDeviceToHost, s));
    HIPC(c, hipStreamSynchronize(s));
    if (int r = check_err_word(c)) return r;
    c->st.p1_rescans = n_rescan;
    c->h_part_off.assign(parts + 1, 0);
    c->h_rec_off.assign(parts + 1, 0);
    c->h_kmers.assign(parts + 1, 0);
    for (int i = 0; i < parts; i++) {
        if (ht[i] > 0xFFFFFFFFull) return fail(c, RGKC_E_ARG, "partition %d holds %llu bytes (> 4 GiB) in one batch", i,
                                                (unsigned long long)ht[i]);
        c->h_part_off[i + 1] = c->h_part_off[i] + ht[i];
        c->h_rec_off[i + 1] = c->h_rec_off[i] + ht[parts + i];
        c->h_kmers[i + 1] = c->h_kmers[i] + ht[2 * parts + i];
    }
    const uint64_t S = c->h_part_off[parts], R = c->h_rec_off[parts];
    c->st.store_bytes = S; c->st.n_records = R; c->st.n_kmers = c->h_kmers[parts];
    // partition regions: part_off, part_end, rec_off, rec_end, row_base, row_end
    std::vector<uint64_t> po(6 * (size_t)parts);
    for (int i = 0; i < parts; i++) {
        po[i] = c->h_part_off[i]; po[parts + i] = c->h_part_off[i + 1];
        po[2 * parts + i] = c->h_rec_off[i]; po[3 * parts + i] = c->h_rec_off[i + 1];
        po[4 * parts + i] = c->h_kmers[i]; po[5 * parts + i] = c->h_kmers[i + 1];
    }
    NEED(c, c->poffs, po.size() * 8);
    HIPC(c, hipMemcpyAsync(c->poffs.p, po.data(), po.size() * 8, hipMemcpyHostToDevice, s));
    NEED(c, c->store, S + 64);
    NEED(c, c->side, R * 4 + 64);
    const uint64_t* dpo = c->poffs.as<uint64_t>();
    a.part_off = dpo; a.part_end = dpo + parts; a.rec_off = dpo + 2 * parts; a.rec_end = dpo + 3 * parts;
    a.wg_byte_off = a.h_bytes;
    a.wg_rec_off = a.h_recs;
    a.store = c->store.as<uint8_t>();
    a.side = c->side.as<uint32_t>();
    a.wg_map = a.skm_list ? a.wg_rescan : nullptr;
    HIPC(c, launch_p1_emit(a, n_rescan, s));
    REC(c, EV_P1E, s);
    c->h_listorder = 0;
    if (a.skm_list) HIPC(c, hipMemcpyAsync(&c->h_listorder, a.n_listorder, 4, hipMemcpyDeviceToHost, s));
    // buckets: one per partition
    std::vector<uint32_t> bp(parts);
    for (int i = 0; i < parts; i++) bp[i] = i;
    NEED(c, c->b_byte, (parts + 1) * 8);
    NEED(c, c->b_rec, (parts + 1) * 8);
    NEED(c, c->b_part, parts * 4);
    HIPC(c, hipMemcpyAsync(c->b_byte.p, c->h_part_off.data(), (parts + 1) * 8, hipMemcpyHostToDevice, s));
    HIPC(c, hipMemcpyAsync(c->b_rec.p, c->h_rec_off.data(), (parts + 1) * 8, hipMemcpyHostToDevice, s));
    HIPC(c, hipMemcpyAsync(c->b_part.p, bp.data(), parts * 4, hipMemcpyHostToDevice, s));
    c->n_buckets = parts;
    c->store_ok = true;
    return 0;
}

// Count-stage working set of buckets [b0, b1) in bytes (2-bit words, record descriptors, key
// regions, per-bucket plan arrays; plus a temporary table of one row per k-mer instance in compact
// mode). An estimate used to cut the buckets into groups that fit the device memory.
uint64_t group_bytes(const rgkc_ctx* c, uint32_t b0, uint32_t b1, size_t W, bool temp_table) {
    const uint64_t kmi = c->h_kmers[b1] - c->h_kmers[b0], ri = c->h_rec_off[b1] - c->h_rec_off[b0];
    const uint64_t nb = b1 - b0;
    const uint64_t words = (kmi + ri * (uint64_t)(c->k + 14)) / 16 * 4 + nb * 128;
    const uint64_t keys = c->heavy_mode == 0 ? (kmi + kmi / 3) * W + nb * 8192 * W : 0;
    const uint64_t per_bucket = nb * ((uint64_t)kSubIds * 40 + (uint64_t)kGroupSegments * 2 * kGroups * 4 + 4096);
    const uint64_t items = kmi / 512 * 88;
    return words + ri * 8 + keys + per_bucket + items + (temp_table ? kmi * (W + 4) : 0);
}

// The count stage over buckets [b0, b1): grouping, plan, key expansion (and its overflow pass), item
// descriptors, count. Rows go to each partition's region as set in poffs (row base and end) by the
// caller, in `keys`/`counts`; the stage times add into c->st.
int count_buckets(rgkc_ctx* c, uint32_t b0, uint32_t b1, uint32_t cmin, uint32_t cmax, void* keys, uint32_t* counts,
                  bool first) {
    hipStream_t s = c->stream;
    const int parts = c->parts;
    const uint32_t nb = b1 - b0;
    const uint64_t R0 = c->h_rec_off[b0], R = c->h_rec_off[b1] - R0;
    const size_t W = c->k <= 32 ? 8 : 16;
    // sweeps of ~2x the table capacity in instances (distinct/instance ratios of read sets are
    // mostly below one half); the table holds 1536 (64-bit keys) or 768 (128-bit) distinct keys.
    // Long k-mers carry more sequencing errors (higher distinct ratios): their sweeps are 4/3 of
    // the table (a 200 Mbp k=55 set with 1% errors: 1024 instances beat 1536 by 13 ms).
    const uint32_t cap0 = c->inst_cap ? c->inst_cap : (c->k <= 32 ? 3328u : 1024u);
    NEEDC(c, c->scratch, R * 8 + 64);
    NEEDC(c, c->sub_start, (size_t)nb * (kGroups + 1) * 4);
    NEEDC(c, c->g_hist, (size_t)nb * kGroupSegments * 2 * kGroups * 4);
    NEEDC(c, c->g_seg_bytes, (size_t)nb * (kGroupSegments + 1) * 8);
    NEEDC(c, c->sub_sinst, (size_t)nb * kSubIds * 4);
    // 2-bit word regions: a record of L bases takes ceil(L / 16) words, at most (k-mers + records *
    // (k + 14)) / 16 per bucket (region record counts are upper bounds in the capacity-sized modes)
    std::vector<uint64_t> bw(nb + 1, 0);
    for (uint32_t i = 0; i < nb; i++) {
        const uint32_t b = b0 + i;
        const uint64_t kmi = c->h_kmers[b + 1] - c->h_kmers[b], ri = c->h_rec_off[b + 1] - c->h_rec_off[b];
        bw[i + 1] = bw[i] + (kmi + ri * (uint64_t)(c->k + 14)) / 16 + 8;
    }
    NEEDC(c, c->b_word, (nb + 1) * 8);
    NEEDC(c, c->words, (bw[nb] + 16) * 4);
    HIPC(c, hipMemcpyAsync(c->b_word.p, bw.data(), (nb + 1) * 8, hipMemcpyHostToDevice, s));
    NEEDC(c, c->item_cnt, ((size_t)nb + 1) * 8);
    NEEDC(c, c->item_off, ((size_t)nb + 1) * 8);
    NEEDC(c, c->scan_tmp2, scan_tmp_words(nb + 1) * 8);
    HIPC(c, hipMemsetAsync(c->item_cnt.p, 0, ((size_t)nb + 1) * 8, s));
    NEEDC(c, c->key_cnt, ((size_t)nb + 1) * 8);
    NEEDC(c, c->key_off, ((size_t)nb + 1) * 8);
    NEEDC(c, c->exp_cnt, ((size_t)nb + 1) * 8);
    NEEDC(c, c->exp_off, ((size_t)nb + 1) * 8);
    HIPC(c, hipMemsetAsync(c->key_cnt.p, 0, ((size_t)nb + 1) * 8, s));
    HIPC(c, hipMemsetAsync(c->exp_cnt.p, 0, ((size_t)nb + 1) * 8, s));
    const uint64_t* dpo = c->poffs.as<uint64_t>();
    CountArgs a{};
    a.store = c->store.as<uint8_t>();
    a.side = c->side.as<uint32_t>();
    a.b_byte = c->b_byte.as<uint64_t>() + b0;
    a.b_rec = c->b_rec.as<uint64_t>() + b0;
    a.b_part = c->b_part.as<uint32_t>() + b0;
    if (c->stream_ended) {
        a.b_byte_end = c->b_ends.as<uint64_t>() + b0;
        a.b_rec_end = c->b_ends.as<uint64_t>() + parts + b0;
    }
    a.n_buckets = nb;
    a.scratch = c->scratch.as<uint64_t>();
    a.scratch_base = R0;
    a.row_base = dpo + 4 * parts;
    a.row_end = dpo + 5 * parts;
    a.row_cursor = c->cursors.as<unsigned long long>();
    a.out_keys = keys;
    a.out_counts = counts;
    a.k = c->k;
    a.cmin = cmin; a.cmax = cmax;
    a.inst_cap = cap0;
    a.count_wg_per_cu = c->count_wg_per_cu;
    a.err = c->err_word.as<uint32_t>();
    a.stats = c->stats_dev.as<unsigned long long>();
    a.part_distinct = c->part_distinct.as<unsigned long long>();
    a.sub_start = c->sub_start.as<uint32_t>();
    a.g_hist = c->g_hist.as<uint32_t>();
    a.g_seg_bytes = c->g_seg_bytes.as<uint64_t>();
    HIPC(c, hipMemsetAsync(c->sub_sinst.p, 0, (size_t)nb * kSubIds * 4, s));
    a.sub_sinst = c->sub_sinst.as<uint32_t>();
    a.words = c->words.as<uint32_t>();
    a.b_word = c->b_word.as<uint64_t>();
    a.item_cnt = c->item_cnt.as<uint64_t>();
    a.heavy_on = c->heavy_mode == 0 ? 1u : 0u;
    a.key_cnt = c->key_cnt.as<uint64_t>();
    a.exp_cnt = c->exp_cnt.as<uint64_t>();
    REC(c, EV_IDX, s);
    HIPC(c, launch_group(a, s));
    HIPC(c, launch_plan(a, false, s));
    HIPC(c, scan_u64(a.item_cnt, c->item_off.as<uint64_t>(), nb + 1, c->scan_tmp2.as<uint64_t>(), s));
    uint64_t tot3[3] = {0, 0, 0};  // work items, key slots, expansion tasks
    HIPC(c, hipMemcpyAsync(&tot3[0], c->item_off.as<uint64_t>() + nb, 8, hipMemcpyDeviceToHost, s));
    if (a.heavy_on) {
        HIPC(c, scan_u64(a.key_cnt, c->key_off.as<uint64_t>(), nb + 1, c->scan_tmp2.as<uint64_t>(), s));
        HIPC(c, hipMemcpyAsync(&tot3[1], c->key_off.as<uint64_t>() + nb, 8, hipMemcpyDeviceToHost, s));
        HIPC(c, scan_u64(a.exp_cnt, c->exp_off.as<uint64_t>(), nb + 1, c->scan_tmp2.as<uint64_t>(), s));
        HIPC(c, hipMemcpyAsync(&tot3[2], c->exp_off.as<uint64_t>() + nb, 8, hipMemcpyDeviceToHost, s));
    }
    HIPC(c, hipStreamSynchronize(s));
    const uint64_t n_items = tot3[0];
    if (n_items > 0x7FFFFFFFull) return fail(c, RGKC_E_ARG, "too many count work items");
    NEEDC(c, c->items, n_items * 8 + 8);
    NEEDC(c, c->item_desc, n_items * 64 + 64);
    a.item_desc = c->item_desc.as<uint32_t>();
    a.item_cnt = c->item_off.as<uint64_t>();
    a.items = c->items.as<uint64_t>();
    a.n_items = n_items;
    c->n_items += n_items;
    uint64_t n_exp = 0;
    if (a.heavy_on && tot3[2] > 0 && tot3[2] <= 0x7FFFFFFFull) {
        // key regions (the groups were sized to hold them); if the device cannot hold them after
        // all, the items read records instead
        if (need_count(c, c->heavy_keys, (size_t)tot3[1] * W + 64) != hipSuccess) {
            (void)hipGetLastError();
            a.heavy_on = 0;
        }
    } else {
        a.heavy_on = 0;
    }
    if (a.heavy_on) {
        NEEDC(c, c->exp_items, tot3[2] * 8 + 8);
        NEEDC(c, c->sub_key_off, (size_t)nb * kSubIds * 8);
        NEEDC(c, c->sub_item0, (size_t)nb * kSubIds * 4);
        NEEDC(c, c->sub_key_cs, (size_t)nb * kSubIds * 4);
        a.sub_item0 = c->sub_item0.as<uint32_t>();
        a.sub_key_cs = c->sub_key_cs.as<uint32_t>();
        a.key_cnt = c->key_off.as<uint64_t>();
        a.exp_cnt = c->exp_off.as<uint64_t>();
        a.heavy_keys = c->heavy_keys.p;
        a.exp_items = c->exp_items.as<uint64_t>();
        a.sub_key_off = c->sub_key_off.as<uint64_t>();
        n_exp = tot3[2];
    }
    c->st.heavy_groups += (uint32_t)n_exp;
    HIPC(c, launch_plan(a, true, s));
    REC(c, EV_PLAN, s);
    HIPC(c, launch_expand(a, n_exp, s));
    if (a.heavy_on) {
        // slices that received more keys than their region holds: exact regions, second pass
        NEEDC(c, c->ovf_n, (n_items + 1) * 8);
        NEEDC(c, c->ovf_off, (n_items + 1) * 8);
        a.ovf_n = c->ovf_n.as<uint64_t>();
        NEEDC(c, c->scan_tmp2, scan_tmp_words(std::max<uint64_t>(n_items, nb) + 1) * 8);
        HIPC(c, launch_ovf_sizes(a, s));
        HIPC(c, scan_u64(a.ovf_n, c->ovf_off.as<uint64_t>(), n_items + 1, c->scan_tmp2.as<uint64_t>(), s));
        uint64_t n_ovf = 0;
        HIPC(c, hipMemcpyAsync(&n_ovf, c->ovf_off.as<uint64_t>() + n_items, 8, hipMemcpyDeviceToHost, s));
        HIPC(c, hipStreamSynchronize(s));
        c->st.ovf_keys += n_ovf;
        if (n_ovf) {
            NEEDC(c, c->ovf_keys, (size_t)n_ovf * W + 64);
            a.ovf_off = c->ovf_off.as<uint64_t>();
            a.ovf_keys = c->ovf_keys.p;
            HIPC(c, launch_expand_ovf(a, n_exp, s));
        } else {
            a.ovf_n = nullptr;
        }
    }
    REC(c, EV_EXP, s);
    HIPC(c, launch_count_items(a, s));
    REC(c, EV_CNT, s);
    HIPC(c, hipStreamSynchronize(s));
    auto el = [&](int e0, int e1) {
        float m = 0;
        if (c->ev_rec[e0] && c->ev_rec[e1] && hipEventElapsedTime(&m, c->ev[e0], c->ev[e1]) != hipSuccess) m = 0;
        (void)hipGetLastError();
        return m;
    };
    c->st.plan_ms += el(EV_IDX, EV_PLAN) + (first ? el(EV_CB, EV_IDX) : 0.0f);
    c->st.expand_ms += el(EV_PLAN, EV_EXP);
    c->st.count_ms += el(EV_EXP, EV_CNT);
    return 0;
}

// Error bits after the count kernels; a row region overflow is left to the caller when `rows_ok`
// (the predicted path recounts those partitions), every other bit fails the count.
int count_errors(rgkc_ctx* c, bool rows_ok, bool* row_overflow) {
    uint32_t e = 0;
    HIPC(c, hipMemcpyAsync(&e, c->err_word.p, 4, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    *row_overflow = (e & RGKC_ERR_ROW_OVERFLOW) != 0;
    if (rows_ok && e == RGKC_ERR_ROW_OVERFLOW) {
        HIPC(c, hipMemsetAsync(c->err_word.p, 0, 4, c->stream));
        return 0;
    }
    return check_err_word(c);
}

int do_count_once(rgkc_ctx* c, uint32_t cmin, uint32_t cmax, bool predicted);

// Predicted row regions that turned out too small: those partitions are counted again, alone and with
// exact (instance-bounded) regions, into the second table; the other partitions keep their rows.
int recount_overflowed(rgkc_ctx* c, std::vector<uint64_t>& rows, uint32_t cmin, uint32_t cmax) {
    hipStream_t s = c->stream;
    const int parts = c->parts;
    const size_t W = c->k <= 32 ? 8 : 16;
    std::vector<int> ov;
    for (int i = 0; i < parts; i++)
        if (rows[i] > c->h_row_base[i + 1] - c->h_row_base[i]) ov.push_back(i);
    if (ov.empty()) return 0;
    uint64_t tot = 0;
    for (int i : ov) tot += c->h_kmers[i + 1] - c->h_kmers[i];
    c->recount_active = true;
    struct Off { bool* f; ~Off() { *f = false; } } off_{&c->recount_active};
    NEEDC(c, c->keys2, tot * W + 64);
    NEEDC(c, c->counts2, tot * 4 + 64);
    c->st.fallbacks += (uint32_t)ov.size();
    uint64_t o = 0;
    for (int i : ov) {
        const uint64_t ni = c->h_kmers[i + 1] - c->h_kmers[i];
        const uint64_t rr[2] = {o, o + ni};
        HIPC(c, hipMemcpyAsync(c->poffs.as<uint64_t>() + 4 * parts + i, &rr[0], 8, hipMemcpyHostToDevice, s));
        HIPC(c, hipMemcpyAsync(c->poffs.as<uint64_t>() + 5 * parts + i, &rr[1], 8, hipMemcpyHostToDevice, s));
        HIPC(c, hipMemsetAsync(c->cursors.as<uint64_t>() + i, 0, 8, s));
        HIPC(c, hipMemsetAsync(c->part_distinct.as<uint64_t>() + i, 0, 8, s));  // counted again
        if (int r = count_buckets(c, (uint32_t)i, (uint32_t)i + 1, cmin, cmax, c->keys2.p, c->counts2.as<uint32_t>(), false))
            return r;
        bool row_ovf = false;
        if (int r = count_errors(c, false, &row_ovf)) return r;
        uint64_t n = 0;
        HIPC(c, hipMemcpyAsync(&n, c->cursors.as<uint64_t>() + i, 8, hipMemcpyDeviceToHost, s));
        HIPC(c, hipStreamSynchronize(s));
        rows[i] = n;
        c->h_row_buf[i] = 1;
        c->h_row_start[i] = o;
        o += ni;
    }
    return 0;
}

// Stage times of the phase 1 just run (its events are complete once the stream has synchronised).
void phase1_times(rgkc_ctx* c) {
    auto el = [&](int e0, int e1) {
        float v = 0;
        if (c->ev_rec[e0] && c->ev_rec[e1] && hipEventElapsedTime(&v, c->ev[e0], c->ev[e1]) != hipSuccess) v = 0;
        (void)hipGetLastError();
        return v;
    };
    c->st.pack_ms = el(EV_START, EV_PACK);
    c->st.segments_ms = el(EV_PACK, EV_SEG);
    c->st.p1_count_ms = el(EV_SEG, EV_P1C);
    c->st.p1_emit_ms = el(EV_P1C, EV_P1E);
}

int do_count(rgkc_ctx* c, uint32_t cmin, uint32_t cmax) {
    const bool pred = !c->pred_distinct.empty() && (int)c->pred_distinct.size() == c->parts;
    c->st.fallbacks = 0;
    REC(c, EV_CB, c->stream);
    return do_count_once(c, cmin, cmax, pred);
}

// The count stage: row regions per partition (exact: its k-mer instances; predicted: its
// predicted distinct k-mers with the margin), buckets counted in groups that fit the device memory.
// With exact regions too large for the memory, each group is counted into a temporary table and
// its rows are compacted into the final table (compact mode).
int do_count_once(rgkc_ctx* c, uint32_t cmin, uint32_t cmax, bool predicted) {
    if (!c->store_ok) return fail(c, RGKC_E_STATE, "no superkmer store");
    hipStream_t s = c->stream;
    const int parts = c->parts;
    const size_t W = c->k <= 32 ? 8 : 16;
    const uint32_t nbk = c->n_buckets;
    c->st.plan_ms = c->st.expand_ms = c->st.count_ms = 0;
    c->st.heavy_groups = 0;
    c->st.ovf_keys = 0;
    c->n_items = 0;
    std::vector<uint64_t> rb(parts + 1, 0);
    auto regions = [&](double margin) {
        for (int i = 0; i < parts; i++) {
            const uint64_t ni = c->h_kmers[i + 1] - c->h_kmers[i];
            uint64_t cap = ni;
            if (predicted) {
                const double d = std::max(0.0, c->pred_distinct[i]);
                cap = std::min<uint64_t>(ni, (uint64_t)(d * (1.0 + margin)) + 4096);
            }
            rb[i + 1] = rb[i] + cap;
        }
    };
    regions(c->pred_margin);
    NEEDC(c, c->cursors, parts * 8);
    NEEDC(c, c->stats_dev, 64);
    NEEDC(c, c->part_distinct, parts * 8);
    HIPC(c, hipMemsetAsync(c->cursors.p, 0, parts * 8, s));
    HIPC(c, hipMemsetAsync(c->part_distinct.p, 0, parts * 8, s));
    HIPC(c, hipMemsetAsync(c->stats_dev.p, 0, 64, s));
    // device memory this stage may use: what is free plus the count buffers it would reuse
    size_t fr = 0, tt = 0;
    if (hipMemGetInfo(&fr, &tt) != hipSuccess) { (void)hipGetLastError(); fr = 0; }
    uint64_t held = 0;
    // Every count-stage buffer counts as available, so the same batch gets the same plan each time
    // and the grow-only buffers are allocated once (allocations of tens of GB cost seconds). The
    // phase-1 buffers stay: the next batch would allocate them again.
    for (const DevBuf* d : {&c->scratch, &c->keys, &c->counts, &c->cursors, &c->stats_dev, &c->sub_start, &c->g_hist,
                            &c->g_seg_bytes, &c->sub_sinst, &c->item_cnt, &c->item_off, &c->items, &c->scan_tmp2,
                            &c->words, &c->b_word, &c->key_cnt, &c->key_off, &c->exp_cnt, &c->exp_off, &c->sub_key_off,
                            &c->heavy_keys, &c->exp_items, &c->item_desc, &c->sub_item0, &c->sub_key_cs, &c->ovf_n,
                            &c->ovf_off, &c->ovf_keys, &c->tkeys, &c->tcounts, &c->sort_segs, &c->sort_chunks,
                            &c->sort_hist, &c->keys2, &c->counts2})
        held += d->cap;
    const uint64_t reserve = 4ull << 30;  // other users of the device and the runtime
    uint64_t avail = (uint64_t)fr + held > reserve ? ((uint64_t)fr + held - reserve) / 100 * 92 : 0;
    if (c->mem_limit) avail = std::min<uint64_t>(avail, c->mem_limit);
    if (predicted) {
        // predicted regions that would take more than 45% of the memory: a smaller margin (the
        // partitions it leaves too small are counted again on their own)
        double m = c->pred_margin;
        while (rb[parts] * (W + 4) > avail / 100 * 45 && m > 0.0) {
            m = m > 0.02 ? m / 2 : 0.0;
            regions(m);
        }
    }
    const uint64_t table_bytes = rb[parts] * (W + 4);
    const bool compact = nbk > 1 && table_bytes + group_bytes(c, 0, nbk, W, false) > avail &&
                         table_bytes > avail / 100 * 45;
    // groups of consecutive buckets whose working set fits beside the table
    uint64_t final_cap = 0;
    uint64_t budget = 0;
    if (compact) {
        // the final table holds the rows of every partition; give it up to 40% of the memory (a
        // read set's distinct k-mers are well below its instances), the groups the rest
        final_cap = std::min<uint64_t>(rb[parts], avail / 100 * 35 / (W + 4));
        // a final table from an earlier batch that is nearly as large is kept as it is
        const uint64_t have = std::min<uint64_t>(c->keys.cap > 64 ? (c->keys.cap - 64) / W : 0,
                                                 c->counts.cap > 64 ? (c->counts.cap - 64) / 4 : 0);
        if (have >= final_cap / 4 * 3 && have <= rb[parts]) final_cap = have;
        budget = avail > final_cap * (W + 4) ? avail - final_cap * (W + 4) : 0;
    } else {
        budget = avail > table_bytes ? avail - table_bytes : 0;
    }
    std::vector<std::pair<uint32_t, uint32_t>> groups;
    for (uint32_t b = 0; b < nbk;) {
        uint32_t e = b + 1;
        while (e < nbk && group_bytes(c, b, e + 1, W, compact) <= budget) e++;
        groups.push_back({b, e});
        b = e;
    }
    // the largest group first: every grow-only buffer reaches its size on the first group
    std::stable_sort(groups.begin(), groups.end(), [&](const std::pair<uint32_t, uint32_t>& x,
                                                       const std::pair<uint32_t, uint32_t>& y) {
        return c->h_kmers[x.second] - c->h_kmers[x.first] > c->h_kmers[y.second] - c->h_kmers[y.first];
    });
    c->st.count_groups = (uint32_t)groups.size();
    c->h_row_buf.assign(parts, 0);
    c->h_row_start.assign(parts, 0);
    std::vector<uint64_t> rows(parts, 0);
    if (!compact) {
        c->st.table_rows_alloc = rb[parts];
        std::vector<uint64_t> rows2(2 * (size_t)parts);
        for (int i = 0; i < parts; i++) { rows2[i] = rb[i]; rows2[parts + i] = rb[i + 1]; }
        HIPC(c, hipMemcpyAsync(c->poffs.as<uint64_t>() + 4 * parts, rows2.data(), rows2.size() * 8,
                               hipMemcpyHostToDevice, s));
        NEEDC(c, c->keys, rb[parts] * W + 64);
        NEEDC(c, c->counts, rb[parts] * 4 + 64);
        c->h_row_base.assign(rb.begin(), rb.end());
        for (size_t gi = 0; gi < groups.size(); gi++)
            if (int r = count_buckets(c, groups[gi].first, groups[gi].second, cmin, cmax, c->keys.p,
                                      c->counts.as<uint32_t>(), gi == 0))
                return r;
        HIPC(c, hipMemcpyAsync(rows.data(), c->cursors.p, parts * 8, hipMemcpyDeviceToHost, s));
        HIPC(c, hipStreamSynchronize(s));
        for (int i = 0; i < parts; i++) c->h_row_start[i] = rb[i];
        bool row_ovf = false;
        if (int r = count_errors(c, predicted, &row_ovf)) return r;
        if (row_ovf) {
            // per-partition fallback: a predicted region that was too small is counted again
            if (int r = recount_overflowed(c, rows, cmin, cmax)) return r;
        }
    } else {
        // each group into a temporary table (instance-bounded regions), then its rows to the final
        // table, partition after partition
        c->compact_active = true;
        struct Off { bool* f; ~Off() { *f = false; } } off_{&c->compact_active};
        c->st.table_rows_alloc = final_cap;
        NEEDC(c, c->keys, final_cap * W + 64);
        NEEDC(c, c->counts, final_cap * 4 + 64);
        uint64_t fo = 0;
        for (size_t gi = 0; gi < groups.size(); gi++) {
            const uint32_t g0 = groups[gi].first, g1 = groups[gi].second;
            const uint64_t t0 = rb[g0], tn = rb[g1] - rb[g0];
            NEEDC(c, c->tkeys, tn * W + 64);
            NEEDC(c, c->tcounts, tn * 4 + 64);
            std::vector<uint64_t> rows2(2 * (size_t)parts, 0);
            for (int i = 0; i < parts; i++) {
                const bool in = (uint32_t)i >= g0 && (uint32_t)i < g1;
                rows2[i] = in ? rb[i] - t0 : 0;
                rows2[parts + i] = in ? rb[i + 1] - t0 : 0;
            }
            HIPC(c, hipMemcpyAsync(c->poffs.as<uint64_t>() + 4 * parts, rows2.data(), rows2.size() * 8,
                                   hipMemcpyHostToDevice, s));
            if (int r = count_buckets(c, g0, g1, cmin, cmax, c->tkeys.p, c->tcounts.as<uint32_t>(), gi == 0)) return r;
            bool row_ovf = false;
            if (int r = count_errors(c, false, &row_ovf)) return r;
            HIPC(c, hipMemcpyAsync(rows.data(), c->cursors.p, parts * 8, hipMemcpyDeviceToHost, s));
            HIPC(c, hipStreamSynchronize(s));
            for (uint32_t i = g0; i < g1; i++) {
                const uint64_t n = rows[i];
                if (fo + n > final_cap)
                    return fail(c, RGKC_E_NOMEM, "the table's %llu rows exceed the %llu rows the device holds",
                                (unsigned long long)(fo + n), (unsigned long long)final_cap);
                if (n) {
                    HIPC(c, hipMemcpyAsync((uint8_t*)c->keys.p + fo * W, (uint8_t*)c->tkeys.p + (rb[i] - t0) * W, n * W,
                                           hipMemcpyDeviceToDevice, s));
                    HIPC(c, hipMemcpyAsync(c->counts.as<uint32_t>() + fo, c->tcounts.as<uint32_t>() + (rb[i] - t0), n * 4,
                                           hipMemcpyDeviceToDevice, s));
                }
                c->h_row_start[i] = fo;
                fo += n;
            }
        }
        HIPC(c, hipStreamSynchronize(s));
        c->h_row_base.assign(parts + 1, 0);
        for (int i = 0; i < parts; i++) c->h_row_base[i + 1] = c->h_row_base[i] + rows[i];
    }
    uint64_t stats[3] = {0, 0, 0};
    std::vector<uint64_t> pdist(parts);
    HIPC(c, hipMemcpyAsync(stats, c->stats_dev.p, 24, hipMemcpyDeviceToHost, s));
    HIPC(c, hipMemcpyAsync(pdist.data(), c->part_distinct.p, parts * 8, hipMemcpyDeviceToHost, s));
    HIPC(c, hipStreamSynchronize(s));
    if (int r = check_err_word(c)) return r;
    c->h_rows = rows;
    uint64_t dsum = 0;
    for (uint64_t d : pdist) dsum += d;
    c->st.sweeps = stats[0]; c->st.sweep_retries = stats[1]; c->st.distinct = dsum;
    c->st.buckets = c->n_buckets;
    c->st.items = c->n_items;
    uint64_t tr = 0;
    for (auto r : rows) tr += r;
    c->st.rows = tr;
    float m = 0;
    if (c->ev_rec[EV_START] && c->ev_rec[EV_CNT] && hipEventElapsedTime(&m, c->ev[EV_START], c->ev[EV_CNT]) != hipSuccess) m = 0;
    (void)hipGetLastError();
    auto el = [&](int e0, int e1) {
        float v = 0;
        if (c->ev_rec[e0] && c->ev_rec[e1] && hipEventElapsedTime(&v, c->ev[e0], c->ev[e1]) != hipSuccess) v = 0;
        (void)hipGetLastError();
        return v;
    };
    c->st.pack_ms = el(EV_START, EV_PACK);
    c->st.segments_ms = el(EV_PACK, EV_SEG);
    c->st.p1_count_ms = el(EV_SEG, EV_P1C);
    c->st.p1_emit_ms = el(EV_P1C, EV_P1E);
    c->st.total_ms = m;
    c->table_ok = true;
    return 0;
}

// Device memory of the row sort: the table, a scratch table with one row per table row (indexed by
// the partitions' dense row prefix), and the plan areas (segment descriptors, chunk maps, bucket
// histograms and cursors). With fixed == false the plan areas are context buffers grown to each
// level's plan; with fixed == true they are areas of the given capacities and each level runs in
// as many rounds as its segments need to fit them.
struct SortMem {
    void* keys = nullptr;
    uint32_t* counts = nullptr;
    void* tkeys = nullptr;
    uint32_t* tcounts = nullptr;
    bool fixed = false;
    SortSeg* segs = nullptr;
    size_t segs_cap = 0;     // SortSeg entries
    uint32_t* chunks = nullptr;
    size_t chunks_cap = 0;   // uint32 entries (chunk map + first chunk of each big segment)
    uint32_t* hist = nullptr;
    size_t hist_cap = 0;     // uint32 entries (histogram + cursor, 2 x 4096 per big segment)
};

// MSD radix over the rows [base[i], base[i] + n[i]) of each partition i (see k_sort.hip): segments
// planned here, sorted in LDS or distributed by their next digit on the device, level by level.
int sort_core(rgkc_ctx* c, int W, int key_bits, const uint64_t* base, const uint64_t* n, int n_parts, SortMem& m) {
    hipStream_t s = c->stream;
    const uint32_t leaf_cap = (uint32_t)sort_leaf_cap(W);
    constexpr uint32_t nbins = 1u << kSortDigitBits;
    // digit width of a segment that does not fit a leaf: enough bits that its buckets average about
    // an eighth of a leaf (at least 4, at most kSortDigitBits and the key bits left). Children of
    // a few leaves' size get a few bits, not 12: their histograms stay small on both sides.
    auto digit_bits = [&](uint32_t len, int bits_left) {
        int b = 4;
        while (b < kSortDigitBits && ((uint64_t)leaf_cap << (b - 3)) < len) b++;
        return (uint8_t)std::min(b, bits_left);
    };
    std::vector<SortSeg> leaf, big;
    auto add = [&](std::vector<SortSeg>& lv, std::vector<SortSeg>& bg, SortSeg sg) {
        if (sg.len == 0 || (sg.len == 1 && sg.where == 0)) return;
        if (sg.len <= leaf_cap) lv.push_back(sg);
        else bg.push_back(sg);
    };
    uint64_t toff = 0;
    for (int i = 0; i < n_parts; i++) {
        if (n[i] > 0xFFFFFFFFull) return fail(c, RGKC_E_ARG, "partition %d has more than 2^32 rows", i);
        SortSeg sg{};
        sg.start = base[i];
        sg.toff = toff;
        sg.len = (uint32_t)n[i];
        sg.nbits = digit_bits(sg.len, key_bits);
        sg.shift = (uint16_t)(key_bits - sg.nbits);
        sg.where = 0;
        toff += n[i];
        add(leaf, big, sg);
    }
    SortArgs a{};
    a.keys = m.keys;
    a.counts = m.counts;
    a.tkeys = m.tkeys;
    a.tcounts = m.tcounts;
    std::vector<uint32_t> hist, chunk_seg, seg_chunk0;
    for (int level = 0; !leaf.empty() || !big.empty(); level++) {
        if (level > 64) return fail(c, RGKC_E_KERNEL, "row sort did not converge");
        std::vector<SortSeg> nleaf, nbig;
        // leaves of this level, in rounds of at most segs_cap segments
        for (size_t l0 = 0; l0 < leaf.size();) {
            size_t nl = leaf.size() - l0;
            if (m.fixed) nl = std::min(nl, m.segs_cap);
            else NEEDC(c, c->sort_segs, nl * sizeof(SortSeg) + 64);
            SortSeg* dseg = m.fixed ? m.segs : c->sort_segs.as<SortSeg>();
            HIPC(c, hipMemcpyAsync(dseg, leaf.data() + l0, nl * sizeof(SortSeg), hipMemcpyHostToDevice, s));
            a.leaf = dseg;
            HIPC(c, launch_sort_leaf(a, (uint32_t)nl, W, s));
            HIPC(c, hipStreamSynchronize(s));  // the segment area is reused by the next round
            l0 += nl;
        }
        // big segments: histogram, bucket starts on the host, scatter into the other buffer
        for (size_t b0 = 0; b0 < big.size();) {
            // the round: big segments whose descriptors, chunk maps and histograms fit
            size_t nbg = 0, nch_tot = 0;
            while (b0 + nbg < big.size()) {
                const size_t nch = (big[b0 + nbg].len + kSortChunk - 1) / kSortChunk;
                if (m.fixed && nbg > 0 &&
                    (nbg + 1 > m.segs_cap || nch_tot + nch + nbg + 1 > m.chunks_cap || 2 * (nbg + 1) * nbins > m.hist_cap))
                    break;
                nch_tot += nch;
                nbg++;
            }
            if (m.fixed && (nbg > m.segs_cap || nch_tot + nbg > m.chunks_cap || 2 * nbg * nbins > m.hist_cap))
                return fail(c, RGKC_E_ARG, "row sort plan exceeds the scratch areas");
            chunk_seg.clear();
            seg_chunk0.assign(nbg, 0);
            for (size_t i = 0; i < nbg; i++) {
                seg_chunk0[i] = (uint32_t)chunk_seg.size();
                const uint32_t nch = (big[b0 + i].len + kSortChunk - 1) / kSortChunk;
                for (uint32_t j = 0; j < nch; j++) chunk_seg.push_back((uint32_t)i);
            }
            if (!m.fixed) {
                NEEDC(c, c->sort_segs, nbg * sizeof(SortSeg) + 64);
                NEEDC(c, c->sort_chunks, (chunk_seg.size() + nbg) * 4 + 64);
                NEEDC(c, c->sort_hist, 2 * nbg * nbins * 4 + 64);
            }
            SortSeg* dseg = m.fixed ? m.segs : c->sort_segs.as<SortSeg>();
            uint32_t* dch = m.fixed ? m.chunks : c->sort_chunks.as<uint32_t>();
            // histogram rows as wide as the round's widest digit (narrower digits use a prefix)
            int nb_max = 0;
            for (size_t i = 0; i < nbg; i++) nb_max = std::max<int>(nb_max, big[b0 + i].nbits);
            const size_t bw = (size_t)1 << nb_max;
            a.hstride = (uint32_t)bw;
            a.hist = m.fixed ? m.hist : c->sort_hist.as<uint32_t>();
            a.cursor = a.hist + nbg * bw;
            HIPC(c, hipMemcpyAsync(dseg, big.data() + b0, nbg * sizeof(SortSeg), hipMemcpyHostToDevice, s));
            HIPC(c, hipMemcpyAsync(dch, chunk_seg.data(), chunk_seg.size() * 4, hipMemcpyHostToDevice, s));
            HIPC(c, hipMemcpyAsync(dch + chunk_seg.size(), seg_chunk0.data(), nbg * 4, hipMemcpyHostToDevice, s));
            a.big = dseg;
            a.chunk_seg = dch;
            a.seg_chunk0 = dch + chunk_seg.size();
            HIPC(c, hipMemsetAsync(a.hist, 0, nbg * bw * 4, s));
            HIPC(c, launch_sort_hist(a, (uint32_t)chunk_seg.size(), W, s));
            hist.resize(nbg * bw);
            HIPC(c, hipMemcpyAsync(hist.data(), a.hist, nbg * bw * 4, hipMemcpyDeviceToHost, s));
            HIPC(c, hipStreamSynchronize(s));
            // bucket starts, and the children: runs of consecutive buckets up to leaf_cap rows, and
            // single buckets above it (sorted by the next digit at the next level)
            for (size_t i = 0; i < nbg; i++) {
                const SortSeg& sg = big[b0 + i];
                uint32_t* h = hist.data() + i * bw;
                const uint32_t sbins = 1u << sg.nbits;
                uint32_t off = 0, run0 = 0, run_n = 0;
                auto close_run = [&]() {
                    if (!run_n) return;
                    SortSeg ch = sg;
                    ch.start = sg.start + run0;
                    ch.toff = sg.toff + run0;
                    ch.len = run_n;
                    ch.where = (uint8_t)(1 - sg.where);
                    nleaf.push_back(ch);
                    run_n = 0;
                };
                for (uint32_t b = 0; b < sbins; b++) {
                    const uint32_t nb = h[b];
                    h[b] = off;
                    if (nb == 0) continue;
                    if (nb > leaf_cap) {
                        close_run();
                        if (sg.shift == 0) return fail(c, RGKC_E_KERNEL, "row sort: repeated key in a table");
                        SortSeg ch = sg;
                        ch.start = sg.start + off;
                        ch.toff = sg.toff + off;
                        ch.len = nb;
                        ch.where = (uint8_t)(1 - sg.where);
                        ch.nbits = digit_bits(nb, sg.shift);
                        ch.shift = (uint16_t)(sg.shift - ch.nbits);
                        nbig.push_back(ch);
                    } else {
                        if (run_n && run_n + nb > leaf_cap) close_run();
                        if (!run_n) run0 = off;
                        run_n += nb;
                    }
                    off += nb;
                }
                close_run();
                if (off != sg.len) return fail(c, RGKC_E_KERNEL, "row sort: histogram of %u rows for %u", off, sg.len);
            }
            HIPC(c, hipMemcpyAsync(a.cursor, hist.data(), nbg * bw * 4, hipMemcpyHostToDevice, s));
            HIPC(c, launch_sort_scatter(a, (uint32_t)chunk_seg.size(), W, s));
            HIPC(c, hipStreamSynchronize(s));  // host vectors and plan areas are reused by the next round
            b0 += nbg;
        }
        leaf.swap(nleaf);
        big.swap(nbig);
    }
    HIPC(c, hipStreamSynchronize(s));
    return 0;
}

// Ascending key order inside every partition's rows of the context's table.
int do_sort_rows(rgkc_ctx* c) {
    const int W = c->k <= 32 ? 8 : 16;
    uint64_t total = 0;
    for (int i = 0; i < c->parts; i++) total += c->h_rows[i];
    NEEDC(c, c->tkeys, total * W + 64);
    NEEDC(c, c->tcounts, total * 4 + 64);
    // partitions in the main table, then those the predicted path recounted into the second one
    for (int tb = 0; tb < 2; tb++) {
        std::vector<uint64_t> n(c->parts, 0);
        bool any = false;
        for (int i = 0; i < c->parts; i++)
            if (c->h_row_buf[i] == tb) { n[i] = c->h_rows[i]; any = any || n[i]; }
        if (!any) continue;
        SortMem m;
        m.keys = tb ? c->keys2.p : c->keys.p;
        m.counts = (tb ? c->counts2 : c->counts).as<uint32_t>();
        m.tkeys = c->tkeys.p;
        m.tcounts = c->tcounts.as<uint32_t>();
        if (int r = sort_core(c, W, 2 * c->k, c->h_row_start.data(), n.data(), c->parts, m)) return r;
    }
    return check_err_word(c);
}

}  // namespace

extern "C" {

size_t rgkc_device_reset(int device) {
    if (hipSetDevice(device) != hipSuccess) return 0;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
    return fr;
}

int rgkc_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int rgkc_create(int device, rgkc_ctx** out) {
    if (!out) return RGKC_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return RGKC_E_NODEV;
    if (device < 0 || device >= n) return RGKC_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return RGKC_E_HIP;
    (void)device_scratch(device);  // this device's no-allocation area, resolved once with it current
    rgkc_ctx* c = new rgkc_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return RGKC_E_HIP; }
    for (int i = 0; i < EV_N; i++)
        if (hipEventCreate(&c->ev[i]) != hipSuccess) { delete c; return RGKC_E_HIP; }
    *out = c;
    return 0;
}

static int stream_free(rgkc_ctx* c);

int rgkc_destroy(rgkc_ctx* c) {
    if (!c) return 0;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    stream_free(c);
    for (int i = 0; i < EV_N; i++) if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

const char* rgkc_last_error(const rgkc_ctx* c) { return c ? c->err.c_str() : "null context"; }

int rgkc_set_reads(rgkc_ctx* c, const char* reads, const uint64_t* offs, int64_t n_reads) {
    if (!c || (!reads && n_reads > 0) || !offs || n_reads < 0) return RGKC_E_ARG;
    (void)hipGetLastError();
    HIPC(c, hipSetDevice(c->device));
    c->reads_ok = false;
    const uint64_t nb = offs[n_reads];
    if (offs[0] != 0) return fail(c, RGKC_E_ARG, "offs[0] must be 0");
    const uint64_t n_words = (nb >> 4) + 2;  // packed words written by the pack kernel
    NEED(c, c->reads, n_words * 16);
    NEED(c, c->offs, (n_reads + 1) * 8);
    NEED(c, c->bases2, (n_words + 2) * 4);
    hipStream_t s = c->stream;
    HIPC(c, hipMemsetAsync((uint8_t*)c->reads.p + nb, 0, n_words * 16 - nb, s));
    if (nb) HIPC(c, hipMemcpyAsync(c->reads.p, reads, nb, hipMemcpyHostToDevice, s));
    HIPC(c, hipMemcpyAsync(c->offs.p, offs, (n_reads + 1) * 8, hipMemcpyHostToDevice, s));
    HIPC(c, hipMemsetAsync(c->bases2.p, 0, (n_words + 2) * 4, s));
    HIPC(c, hipStreamSynchronize(s));
    c->n_reads = (uint64_t)n_reads;
    c->n_bases = nb;
    c->reads_ok = true;
    c->store_ok = false;
    c->table_ok = false;
    return 0;
}

int rgkc_superkmers(rgkc_ctx* c, int k, int p, int parts, uint64_t* part_offs, uint64_t* skm_cnt, uint64_t* kmer_cnt) {
    if (!c) return RGKC_E_ARG;
    (void)hipGetLastError();
    HIPC(c, hipSetDevice(c->device));
    if (int r = do_phase1(c, k, p, parts)) return r;
    HIPC(c, hipStreamSynchronize(c->stream));
    if (int r = check_err_word(c)) return r;
    phase1_times(c);
    for (int i = 0; i < parts; i++) {
        if (part_offs) part_offs[i] = c->h_part_off[i];
        if (skm_cnt) skm_cnt[i] = c->h_rec_off[i + 1] - c->h_rec_off[i];
        if (kmer_cnt) kmer_cnt[i] = c->h_kmers[i + 1] - c->h_kmers[i];
    }
    if (part_offs) part_offs[parts] = c->h_part_off[parts];
    return 0;
}

int rgkc_copy_store(rgkc_ctx* c, uint8_t* dst) {
    if (!c || !dst) return RGKC_E_ARG;
    if (!c->store_ok) return fail(c, RGKC_E_STATE, "no superkmer store");
    if (c->stream_ended) {  // regions have unused capacity: copy the used part of each
        uint64_t o = 0;
        for (int i = 0; i < c->parts; i++) {
            const uint64_t n = c->h_used_bytes[i];
            if (n) HIPC(c, hipMemcpyAsync(dst + o, c->store.as<uint8_t>() + c->h_part_off[i], n, hipMemcpyDeviceToHost,
                                          c->stream));
            o += n;
        }
        HIPC(c, hipStreamSynchronize(c->stream));
        return 0;
    }
    const uint64_t S = c->h_part_off[c->parts];
    if (S) HIPC(c, hipMemcpyAsync(dst, c->store.p, S, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return 0;
}

int rgkc_count(rgkc_ctx* c, uint32_t cmin, uint32_t cmax, uint64_t* row_counts, uint64_t* distinct) {
    if (!c) return RGKC_E_ARG;
    (void)hipGetLastError();
    HIPC(c, hipSetDevice(c->device));
    if (int r = do_count(c, cmin, cmax)) return r;
    if (row_counts) for (int i = 0; i < c->parts; i++) row_counts[i] = c->h_rows[i];
    if (distinct) *distinct = c->st.distinct;
    return 0;
}

// Records of a partition byte stream (cpu_funcs.cpp:516-525): the record at byte i, its start % 3,
// its bases and its byte count (0 if malformed).
static uint64_t host_record(const uint8_t* s, uint64_t i, uint64_t n, uint32_t& bm3, uint32_t& L) {
    const uint32_t ind0 = s[i] >> 6;
    uint64_t j = i + 1;
    while (j < n && s[j] >= 0xC0) j++;
    if (ind0 == 0 || j >= n) return 0;
    bm3 = 3u - ind0;
    L = ind0 + 3u * (uint32_t)(j - i - 1) + (uint32_t)(s[j] >> 6);
    return j - i + 1;
}

// Minimizer of the k-mer at base t0 of a record with minimizer length p: the minimum over its
// p-mers of min(forward, reverse complement), a p-mer that fails the filter counting as all ones
// (minimizer_filter.h:19-20, cpu_funcs.cpp:109-146).
static uint32_t host_kmer_mm(const uint8_t* r, uint32_t bm3, uint32_t t0, int k, int p) {
    const uint32_t pmask = (p >= 16) ? 0xFFFFFFFFu : ((1u << (2 * p)) - 1u);
    const uint32_t rsh = 2u * (uint32_t)(p - 1);
    uint32_t f = 0, rv = 0, mm = pmask;
    for (int t = 0; t < k; t++) {
        const uint32_t q = bm3 + t0 + (uint32_t)t;
        const uint32_t cb = ((uint32_t)r[q / 3u] >> (4u - 2u * (q % 3u))) & 3u;
        f = ((f << 2) | cb) & pmask;
        rv = (rv >> 2) | ((3u - cb) << rsh);
        if (t >= p - 1) {
            auto valid = [&](uint32_t x) { return ((x >> ((p - 3) * 2)) & 0x2Bu) != 0u && (x & 0x3Fu) != 0u; };
            const uint32_t vf = valid(f) ? f : pmask, vr = valid(rv) ? rv : pmask;
            mm = std::min(mm, std::min(vf, vr));
        }
    }
    return mm;
}

int rgkc_store_minimizer_length(const uint8_t* store, const uint64_t* part_offs, int n_parts, int k, int* p_out) {
    if (!part_offs || !p_out || n_parts < 1 || k < 8 || k > 63 || (!store && part_offs[n_parts] > 0)) return RGKC_E_ARG;
    // records of two or more k-mers: every k-mer of a superkmer has the superkmer's minimizer, so
    // the minimizer length is the one under which all of them agree
    struct Rec { const uint8_t* r; uint32_t bm3, L; };
    std::vector<Rec> recs;
    for (int pi = 0; pi < n_parts && recs.size() < 512; pi++) {
        const uint8_t* s = store + part_offs[pi];
        const uint64_t n = part_offs[pi + 1] - part_offs[pi];
        for (uint64_t i = 0; i < n && recs.size() < 512;) {
            uint32_t bm3 = 0, L = 0;
            const uint64_t nb = host_record(s, i, n, bm3, L);
            if (!nb) return RGKC_E_ARG;
            if (L >= (uint32_t)k + 1) recs.push_back({s + i, bm3, L});
            i += nb;
        }
    }
    if (recs.size() < 8) return RGKC_E_ARG;
    int found = 0, n_found = 0;
    for (int p = 3; p <= std::min(16, k); p++) {
        bool ok = true;
        for (const Rec& r : recs) {
            const uint32_t nk = r.L - (uint32_t)k + 1;
            const uint32_t m0 = host_kmer_mm(r.r, r.bm3, 0, k, p);
            for (uint32_t t = 1; t < nk && ok; t++) ok = host_kmer_mm(r.r, r.bm3, t, k, p) == m0;
            if (!ok) break;
        }
        if (ok) { found = p; n_found++; }
    }
    if (n_found != 1) return RGKC_E_ARG;
    *p_out = found;
    return 0;
}

int rgkc_count_host_store(rgkc_ctx* c, const uint8_t* store, const uint64_t* part_offs, int n_parts,
                          const int32_t* part_ids, int parts_total, int k, int p, uint32_t cmin, uint32_t cmax,
                          uint64_t* row_counts, uint64_t* distinct) {
    (void)part_ids;
    if (!c || !part_offs || n_parts < 1 || (!store && part_offs[n_parts] > 0)) return RGKC_E_ARG;
    if (p == 0) {
        const int rc = rgkc_store_minimizer_length(store, part_offs, n_parts, k, &p);
        if (rc) return fail(c, rc, "the minimizer length could not be recovered from the records");
    }
    if (k < 8 || k > 63 || p < 3 || p > 16 || p > k || parts_total < 0) return fail(c, RGKC_E_ARG, "bad k/p/parts");
    (void)hipGetLastError();
    HIPC(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    c->store_ok = false; c->table_ok = false; c->stream_ended = false;
    c->st = rgkc_stats{};
    const uint64_t S = part_offs[n_parts];
    NEED(c, c->err_word, 64);
    HIPC(c, hipMemsetAsync(c->err_word.p, 0, 64, s));
    NEED(c, c->store, S + 64);
    NEED(c, c->b_byte, (n_parts + 1) * 8);
    NEED(c, c->totals, 2 * (size_t)n_parts * 8);
    for (auto& r : c->ev_rec) r = false;
    REC(c, EV_START, s);
    if (S) HIPC(c, hipMemcpyAsync(c->store.p, store, S, hipMemcpyHostToDevice, s));
    HIPC(c, hipMemcpyAsync(c->b_byte.p, part_offs, (n_parts + 1) * 8, hipMemcpyHostToDevice, s));
    IndexArgs ia{};
    ia.store = c->store.as<uint8_t>();
    ia.part_byte = c->b_byte.as<uint64_t>();
    ia.n_parts = (uint32_t)n_parts;
    ia.k = k; ia.p = p; ia.parts_total = parts_total;
    ia.rec_count = c->totals.as<uint64_t>();
    ia.err = c->err_word.as<uint32_t>();
    HIPC(c, launch_index_count(ia, s));
    std::vector<uint64_t> rc(2 * (size_t)n_parts);
    HIPC(c, hipMemcpyAsync(rc.data(), c->totals.p, rc.size() * 8, hipMemcpyDeviceToHost, s));
    HIPC(c, hipStreamSynchronize(s));
    if (int r = check_err_word(c)) return r;
    c->k = k; c->p = p; c->parts = n_parts;
    c->h_part_off.assign(part_offs, part_offs + n_parts + 1);
    c->h_rec_off.assign(n_parts + 1, 0);
    c->h_kmers.assign(n_parts + 1, 0);
    for (int i = 0; i < n_parts; i++) {
        if (part_offs[i + 1] - part_offs[i] > 0xFFFFFFFFull) return fail(c, RGKC_E_ARG, "partition over 4 GiB");
        c->h_rec_off[i + 1] = c->h_rec_off[i] + rc[2 * i];
        c->h_kmers[i + 1] = c->h_kmers[i] + rc[2 * i + 1];
    }
    const uint64_t R = c->h_rec_off[n_parts];
    c->st.store_bytes = S; c->st.n_records = R; c->st.n_kmers = c->h_kmers[n_parts];
    std::vector<uint64_t> po(6 * (size_t)n_parts);
    for (int i = 0; i < n_parts; i++) {
        po[i] = part_offs[i]; po[n_parts + i] = part_offs[i + 1];
        po[2 * n_parts + i] = c->h_rec_off[i]; po[3 * n_parts + i] = c->h_rec_off[i + 1];
        po[4 * n_parts + i] = c->h_kmers[i]; po[5 * n_parts + i] = c->h_kmers[i + 1];
    }
    NEED(c, c->poffs, po.size() * 8);
    NEED(c, c->side, R * 4 + 64);
    NEED(c, c->b_rec, (n_parts + 1) * 8);
    NEED(c, c->b_part, n_parts * 4);
    std::vector<uint32_t> bp(n_parts);
    for (int i = 0; i < n_parts; i++) bp[i] = (uint32_t)i;
    HIPC(c, hipMemcpyAsync(c->poffs.p, po.data(), po.size() * 8, hipMemcpyHostToDevice, s));
    HIPC(c, hipMemcpyAsync(c->b_rec.p, c->h_rec_off.data(), (n_parts + 1) * 8, hipMemcpyHostToDevice, s));
    HIPC(c, hipMemcpyAsync(c->b_part.p, bp.data(), n_parts * 4, hipMemcpyHostToDevice, s));
    ia.side = c->side.as<uint32_t>();
    ia.rec_off = c->b_rec.as<uint64_t>();
    HIPC(c, launch_index_fill(ia, s));
    REC(c, EV_P1E, s);
    c->n_buckets = (uint32_t)n_parts;
    c->store_ok = true;
    if (int r = do_count(c, cmin, cmax)) return r;
    c->st.index_ms = 0;
    if (hipEventElapsedTime(&c->st.index_ms, c->ev[EV_START], c->ev[EV_P1E]) != hipSuccess) c->st.index_ms = 0;
    (void)hipGetLastError();
    c->st.pack_ms = c->st.segments_ms = c->st.p1_count_ms = c->st.p1_emit_ms = 0;
    if (row_counts) for (int i = 0; i < n_parts; i++) row_counts[i] = c->h_rows[i];
    if (distinct) *distinct = c->st.distinct;
    return 0;
}

int rgkc_copy_table(rgkc_ctx* c, void* keys, uint32_t* counts) {
    if (!c) return RGKC_E_ARG;
    if (!c->table_ok) return fail(c, RGKC_E_STATE, "no table");
    const size_t W = c->k <= 32 ? 8 : 16;
    uint64_t o = 0;
    for (int i = 0; i < c->parts; i++) {
        const uint64_t n = c->h_rows[i], b = c->h_row_start[i];
        const bool two = c->h_row_buf[i] != 0;
        if (n) {
            if (keys) HIPC(c, hipMemcpyAsync((uint8_t*)keys + o * W, (uint8_t*)(two ? c->keys2.p : c->keys.p) + b * W, n * W,
                                             hipMemcpyDeviceToHost, c->stream));
            if (counts) HIPC(c, hipMemcpyAsync(counts + o, (two ? c->counts2 : c->counts).as<uint32_t>() + b, n * 4,
                                               hipMemcpyDeviceToHost, c->stream));
        }
        o += n;
    }
    HIPC(c, hipStreamSynchronize(c->stream));
    return 0;
}

int rgkc_sort_rows(rgkc_ctx* c) {
    if (!c) return RGKC_E_ARG;
    if (!c->table_ok) return fail(c, RGKC_E_STATE, "no table");
    (void)hipGetLastError();
    HIPC(c, hipSetDevice(c->device));
    return do_sort_rows(c);
}

int rgkc_sort_table_noalloc(rgkc_ctx* c, int k, void* keys, uint32_t* counts, const uint64_t* row_base,
                            const uint64_t* row_counts, int n_parts) {
    if (!c) return RGKC_E_ARG;
    if (!row_base || !row_counts || n_parts < 1) return fail(c, RGKC_E_ARG, "row_base, row_counts and n_parts >= 1 required");
    if (k < 1 || k > 63) return fail(c, RGKC_E_ARG, "k must be in [1, 63]");
    (void)hipGetLastError();
    HIPC(c, hipSetDevice(c->device));
    uint8_t* const base = device_scratch(c->device);
    if (!base) return fail(c, RGKC_E_HIP, "the no-allocation area of device %d could not be resolved", c->device);
    const int W = k <= 32 ? 8 : 16;
    uint64_t total = 0;
    for (int i = 0; i < n_parts; i++) total += row_counts[i];
    const uint64_t max_rows = kNoAllocRowBytes / (uint64_t)(W + 4);
    if (total > max_rows)
        return fail(c, RGKC_E_ARG, "%llu rows exceed the no-allocation limit of %llu rows for %d-byte keys",
                    (unsigned long long)total, (unsigned long long)max_rows, W);
    if (total && (!keys || !counts)) return fail(c, RGKC_E_ARG, "keys and counts required");
    SortMem m;
    m.keys = keys;
    m.counts = counts;
    m.tkeys = base;
    m.tcounts = (uint32_t*)(base + max_rows * (uint64_t)W);
    m.fixed = true;
    m.segs = (SortSeg*)(base + kNoAllocRowBytes);
    m.segs_cap = kNoAllocSegBytes / sizeof(SortSeg);
    m.chunks = (uint32_t*)(base + kNoAllocRowBytes + kNoAllocSegBytes);
    m.chunks_cap = kNoAllocChunkBytes / 4;
    m.hist = (uint32_t*)(base + kNoAllocRowBytes + kNoAllocSegBytes + kNoAllocChunkBytes);
    m.hist_cap = kNoAllocHistBytes / 4;
    return sort_core(c, W, 2 * k, row_base, row_counts, n_parts, m);
}

int rgkc_run(rgkc_ctx* c, int k, int p, int parts, uint32_t cmin, uint32_t cmax) {
    if (!c) return RGKC_E_ARG;
    (void)hipGetLastError();
    HIPC(c, hipSetDevice(c->device));
    if (int r = do_phase1(c, k, p, parts)) return r;
    return do_count(c, cmin, cmax);
}

int rgkc_get_stats(const rgkc_ctx* c, rgkc_stats* out) {
    if (!c || !out) return RGKC_E_ARG;
    *out = c->st;
    return 0;
}

// After a capacity-sized emit: buckets = the used part of each partition region.
static int finish_regions(rgkc_ctx* c, const std::vector<uint64_t>& cur) {
    hipStream_t s = c->stream;
    const int parts = c->parts;
    c->h_used_bytes.assign(cur.begin(), cur.begin() + parts);
    c->h_kmers.assign(parts + 1, 0);
    std::vector<uint64_t> ends(2 * (size_t)parts);
    uint64_t S = 0, R = 0;
    for (int i = 0; i < parts; i++) {
        c->h_kmers[i + 1] = c->h_kmers[i] + cur[2 * parts + i];
        ends[i] = c->h_part_off[i] + cur[i];
        ends[parts + i] = c->h_rec_off[i] + cur[parts + i];
        S += cur[i]; R += cur[parts + i];
    }
    std::vector<uint32_t> bp(parts);
    for (int i = 0; i < parts; i++) bp[i] = (uint32_t)i;
    NEED(c, c->b_byte, (parts + 1) * 8);
    NEED(c, c->b_rec, (parts + 1) * 8);
    NEED(c, c->b_part, parts * 4);
    NEED(c, c->b_ends, 2 * (size_t)parts * 8);
    HIPC(c, hipMemcpyAsync(c->b_byte.p, c->h_part_off.data(), (parts + 1) * 8, hipMemcpyHostToDevice, s));
    HIPC(c, hipMemcpyAsync(c->b_rec.p, c->h_rec_off.data(), (parts + 1) * 8, hipMemcpyHostToDevice, s));
    HIPC(c, hipMemcpyAsync(c->b_part.p, bp.data(), parts * 4, hipMemcpyHostToDevice, s));
    HIPC(c, hipMemcpyAsync(c->b_ends.p, ends.data(), ends.size() * 8, hipMemcpyHostToDevice, s));
    HIPC(c, hipStreamSynchronize(s));
    c->n_buckets = (uint32_t)parts;
    c->store_ok = true;
    c->st.store_bytes = S; c->st.n_records = R; c->st.n_kmers = c->h_kmers[parts];
    c->stream_ended = true;
    return 0;
}

// Capacity-sized regions for phase-1 emission (stream mode and predicted single batches).
static int setup_regions(rgkc_ctx* c, const uint64_t* cap_bytes, const uint64_t* cap_recs) {
    const int parts = c->parts;
    hipStream_t s = c->stream;
    c->h_part_off.assign(parts + 1, 0);
    c->h_rec_off.assign(parts + 1, 0);
    for (int i = 0; i < parts; i++) {
        if (cap_recs[i] > kRecMask || cap_bytes[i] >= (1ull << (64 - kRecBits)))
            return fail(c, RGKC_E_ARG, "partition %d capacity exceeds 2^%d records or 2^%d bytes", i, kRecBits,
                        64 - kRecBits);
        c->h_part_off[i + 1] = c->h_part_off[i] + cap_bytes[i];
        c->h_rec_off[i + 1] = c->h_rec_off[i] + cap_recs[i];
    }
    std::vector<uint64_t> po(6 * (size_t)parts, 0);
    for (int i = 0; i < parts; i++) {
        po[i] = c->h_part_off[i]; po[parts + i] = c->h_part_off[i + 1];
        po[2 * parts + i] = c->h_rec_off[i]; po[3 * parts + i] = c->h_rec_off[i + 1];
    }
    NEED(c, c->poffs, po.size() * 8);
    HIPC(c, hipMemcpyAsync(c->poffs.p, po.data(), po.size() * 8, hipMemcpyHostToDevice, s));
    NEED(c, c->store, c->h_part_off[parts] + 64);
    NEED(c, c->side, c->h_rec_off[parts] * 4 + 64);
    NEED(c, c->cursors3, 3 * (size_t)parts * 8);
    HIPC(c, hipMemsetAsync(c->cursors3.p, 0, 3 * (size_t)parts * 8, s));
    return 0;
}

// The device keeps one packed cursor per partition (bytes << kRecBits | records) plus a k-mer
// cursor; this expands them in place to [bytes][records][k-mers].
static void unpack_cursors(std::vector<uint64_t>& cur, int parts) {
    for (int i = 0; i < parts; i++) {
        cur[parts + i] = cur[i] & kRecMask;
        cur[i] >>= kRecBits;
    }
}

static void region_args(rgkc_ctx* c, P1Args& a) {
    const int parts = c->parts;
    const uint64_t* dpo = c->poffs.as<uint64_t>();
    a.part_off = dpo; a.part_end = dpo + parts; a.rec_off = dpo + 2 * parts; a.rec_end = dpo + 3 * parts;
    unsigned long long* cur = c->cursors3.as<unsigned long long>();
    a.pos_cursor = cur; a.kmer_cursor = cur + 2 * parts;
    a.store = c->store.as<uint8_t>();
    a.side = c->side.as<uint32_t>();
    a.err = c->err_word.as<uint32_t>();
}

// ---- stream mode: many (uneven) batches appended into capacity-sized partition regions --------
static int stream_free(rgkc_ctx* c) {
    (void)hipStreamSynchronize(c->stream);  // uploads from the staging sets may be in flight
    if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
    for (int i = 0; i < 2; i++) {
        if (c->h_hdr[i]) { (void)hipHostFree(c->h_hdr[i]); c->h_hdr[i] = nullptr; }
        if (c->h_stage[i]) { (void)hipHostFree(c->h_stage[i]); c->h_stage[i] = nullptr; }
        if (c->h_offs_stage[i]) { (void)hipHostFree(c->h_offs_stage[i]); c->h_offs_stage[i] = nullptr; }
        if (c->stage_ev[i]) { (void)hipEventDestroy(c->stage_ev[i]); c->stage_ev[i] = nullptr; }
        c->stage_ev_rec[i] = false;
    }
    (void)hipGetLastError();
    c->stream_on = false;
    return 0;
}

static int stream_enqueue(rgkc_ctx* c) {
    hipStream_t s = c->stream;
    HIPC(c, launch_batch_prep(c->reads.as<uint8_t>(), c->offs.as<uint64_t>(), c->hdr.as<uint64_t>(), c->cap_reads,
                              c->cap_bases, c->k, c->bases2.as<uint32_t>(), c->nseg.as<uint64_t>(),
                              c->seg_off.as<uint64_t>(), c->scan_tmp.as<uint64_t>(), c->seg_read.as<uint32_t>(),
                              c->seg_first.as<uint32_t>(), s));
    HIPC(c, launch_p1_emit_stream(c->sargs, 512, s));
    return 0;
}

int rgkc_superkmers_sized(rgkc_ctx* c, int k, int p, int parts, const uint64_t* cap_bytes, const uint64_t* cap_recs,
                          uint64_t* part_bytes, uint64_t* skm_cnt, uint64_t* kmer_cnt) {
    if (!c || !cap_bytes || !cap_recs) return RGKC_E_ARG;
    (void)hipGetLastError();
    HIPC(c, hipSetDevice(c->device));
    if (!c->reads_ok) return fail(c, RGKC_E_STATE, "no read batch uploaded");
    if (k < 8 || k > 63 || p < 3 || p > 16 || p > k || k - p + 1 > 64) return fail(c, RGKC_E_ARG, "bad k/p");
    if (parts < 1 || parts > kMaxParts) return fail(c, RGKC_E_ARG, "parts must be in [1, %d]", kMaxParts);
    hipStream_t s = c->stream;
    c->k = k; c->p = p; c->parts = parts;
    c->store_ok = c->table_ok = false; c->stream_ended = false;
    c->st = rgkc_stats{};
    c->st.n_reads = c->n_reads; c->st.n_bases = c->n_bases;
    NEED(c, c->err_word, 64);
    HIPC(c, hipMemsetAsync(c->err_word.p, 0, 64, s));
    for (auto& r : c->ev_rec) r = false;
    REC(c, EV_START, s);
    HIPC(c, launch_pack(c->reads.as<uint8_t>(), c->n_bases, c->bases2.as<uint32_t>(), s));
    REC(c, EV_PACK, s);
    const uint64_t n = c->n_reads;
    const uint64_t max_seg = n + c->n_bases / kSegKmers + 2;
    NEED(c, c->nseg, (n + 1) * 8);
    NEED(c, c->seg_off, (n + 1) * 8);
    NEED(c, c->scan_tmp, scan_tmp_words(n + 1) * 8);
    NEED(c, c->seg_read, max_seg * 4);
    NEED(c, c->seg_first, max_seg * 4);
    HIPC(c, hipMemsetAsync(c->nseg.p, 0, (n + 1) * 8, s));
    HIPC(c, launch_seg_count(c->offs.as<uint64_t>(), n, k, c->nseg.as<uint64_t>(), s));
    HIPC(c, scan_u64(c->nseg.as<uint64_t>(), c->seg_off.as<uint64_t>(), n + 1, c->scan_tmp.as<uint64_t>(), s));
    HIPC(c, launch_seg_fill(c->offs.as<uint64_t>(), c->seg_off.as<uint64_t>(), n, k, c->seg_read.as<uint32_t>(),
                            c->seg_first.as<uint32_t>(), s));
    REC(c, EV_SEG, s);
    if (int r = setup_regions(c, cap_bytes, cap_recs)) return r;
    P1Args a{};
    a.bases2 = c->bases2.as<uint32_t>();
    a.offs = c->offs.as<uint64_t>();
    a.seg_read = c->seg_read.as<uint32_t>();
    a.seg_first = c->seg_first.as<uint32_t>();
    a.n_bases = c->n_bases;
    a.k = k; a.p = p; a.parts = parts;
    region_args(c, a);
    a.n_seg_dev = c->seg_off.as<uint64_t>() + n;
    if (c->list_cap != 0xFFFFFFFFu) {
        // one scan: the count pass keeps superkmer lists (its histograms are each tile's counts),
        // the list emit reserves each tile's runs against the capacities
        uint64_t nseg_total = 0;
        HIPC(c, hipMemcpyAsync(&nseg_total, c->seg_off.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, s));
        HIPC(c, hipStreamSynchronize(s));
        if (nseg_total > 0xFFFFFFFFull) return fail(c, RGKC_E_ARG, "too many read segments in one batch");
        const int T1 = p1_threads(k, p);
        const uint32_t n_wg = (uint32_t)((nseg_total + T1 - 1) / T1);
        const size_t hwords = (size_t)parts * ((size_t)n_wg + 1);
        NEED(c, c->hist, 3 * hwords * 4);
        HIPC(c, hipMemsetAsync(c->hist.p, 0, 3 * hwords * 4, s));
        a.n_seg = nseg_total;
        a.h_bytes = c->hist.as<uint32_t>();
        a.h_recs = a.h_bytes + hwords;
        a.h_kmers = a.h_recs + hwords;
        a.n_wg = n_wg;
        if (int r = setup_lists(c, a, T1, n_wg)) return r;
    }
    if (a.skm_list) {
        HIPC(c, launch_p1_count(a, s));
        REC(c, EV_P1C, s);
        uint32_t n_rescan = 0;
        HIPC(c, hipMemcpyAsync(&n_rescan, a.n_rescan, 4, hipMemcpyDeviceToHost, s));
        HIPC(c, hipStreamSynchronize(s));
        c->st.p1_rescans = n_rescan;
        a.list_reserve = 1;
        a.wg_map = a.wg_rescan;
        HIPC(c, launch_p1_emit_sized(a, n_rescan, s));
        HIPC(c, hipMemcpyAsync(&c->h_listorder, a.n_listorder, 4, hipMemcpyDeviceToHost, s));
    } else {
        REC(c, EV_P1C, s);
        HIPC(c, launch_p1_emit_stream(a, 1024, s));
    }
    REC(c, EV_P1E, s);
    std::vector<uint64_t> cur(3 * (size_t)parts);
    HIPC(c, hipMemcpyAsync(cur.data(), c->cursors3.p, cur.size() * 8, hipMemcpyDeviceToHost, s));
    HIPC(c, hipStreamSynchronize(s));
    unpack_cursors(cur, parts);
    if (int r = check_err_word(c)) return r;
    for (int i = 0; i < parts; i++)
        if (cur[i] > cap_bytes[i] || cur[parts + i] > cap_recs[i])
            return fail(c, RGKC_E_OVERFLOW, "partition %d exceeds its capacity", i);
    if (int r = finish_regions(c, cur)) return r;
    c->st.n_reads = n; c->st.n_bases = c->n_bases;
    for (int i = 0; i < parts; i++) {
        if (part_bytes) part_bytes[i] = cur[i];
        if (skm_cnt) skm_cnt[i] = cur[parts + i];
        if (kmer_cnt) kmer_cnt[i] = cur[2 * parts + i];
    }
    return 0;
}

int rgkc_stream_begin(rgkc_ctx* c, int k, int p, int parts, uint64_t max_reads, uint64_t max_bases,
                      const uint64_t* cap_bytes, const uint64_t* cap_recs, int use_graph) {
    if (!c || !cap_bytes || !cap_recs || max_reads < 1 || max_bases < 1) return RGKC_E_ARG;
    if (k < 8 || k > 63 || p < 3 || p > 16 || p > k || k - p + 1 > 64) return fail(c, RGKC_E_ARG, "bad k/p");
    if (parts < 1 || parts > kMaxParts) return fail(c, RGKC_E_ARG, "parts must be in [1, %d]", kMaxParts);
    (void)hipGetLastError();
    HIPC(c, hipSetDevice(c->device));
    stream_free(c);
    hipStream_t s = c->stream;
    c->k = k; c->p = p; c->parts = parts;
    c->cap_reads = max_reads; c->cap_bases = max_bases;
    c->store_ok = c->table_ok = c->reads_ok = false;
    c->st = rgkc_stats{};
    const uint64_t n_words = (max_bases >> 4) + 2;
    const uint64_t max_seg = max_reads + max_bases / kSegKmers + 2;
    NEED(c, c->reads, n_words * 16);
    NEED(c, c->offs, (max_reads + 1) * 8);
    NEED(c, c->bases2, (n_words + 2) * 4);
    NEED(c, c->nseg, (max_reads + 2) * 8);
    NEED(c, c->seg_off, (max_reads + 2) * 8);
    NEED(c, c->scan_tmp, scan_tmp_words(max_reads + 1) * 8);
    NEED(c, c->seg_read, max_seg * 4);
    NEED(c, c->seg_first, max_seg * 4);
    NEED(c, c->hdr, 64);
    NEED(c, c->err_word, 64);
    HIPC(c, hipMemsetAsync(c->reads.p, 0, n_words * 16, s));
    HIPC(c, hipMemsetAsync(c->bases2.p, 0, (n_words + 2) * 4, s));
    HIPC(c, hipMemsetAsync(c->err_word.p, 0, 64, s));
    for (int i = 0; i < 2; i++) {
        HIPC(c, hipHostMalloc((void**)&c->h_hdr[i], 64, hipHostMallocDefault));
        HIPC(c, hipHostMalloc((void**)&c->h_stage[i], n_words * 16, hipHostMallocDefault));
        HIPC(c, hipHostMalloc((void**)&c->h_offs_stage[i], (max_reads + 1) * 8, hipHostMallocDefault));
        HIPC(c, hipEventCreateWithFlags(&c->stage_ev[i], hipEventDisableTiming));
        c->stage_ev_rec[i] = false;
        memset(c->h_stage[i], 0, n_words * 16);
        memset(c->h_hdr[i], 0, 64);
    }
    c->last_batch_bases = 0;
    if (int r = setup_regions(c, cap_bytes, cap_recs)) return r;
    P1Args& a = c->sargs;
    a = P1Args{};
    a.bases2 = c->bases2.as<uint32_t>();
    a.offs = c->offs.as<uint64_t>();
    a.seg_read = c->seg_read.as<uint32_t>();
    a.seg_first = c->seg_first.as<uint32_t>();
    a.n_bases = max_bases;
    a.k = k; a.p = p; a.parts = parts;
    region_args(c, a);
    a.n_seg_dev = c->hdr.as<uint64_t>() + 2;
    HIPC(c, hipStreamSynchronize(s));
    c->stream_graph = use_graph != 0;
    if (c->stream_graph) {
        // capture the per-batch kernel sequence once; every batch replays it
        HIPC(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int r = stream_enqueue(c);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (r) return r;
        if (e != hipSuccess) return fail(c, RGKC_E_HIP, "graph capture: %s", hipGetErrorString(e));
        c->graph = g;
        HIPC(c, hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
    }
    c->n_batches = 0;
    c->stream_on = true;
    return 0;
}

int rgkc_stream_add(rgkc_ctx* c, const char* reads, const uint64_t* offs, int64_t n_reads) {
    if (!c || !offs || n_reads < 0) return RGKC_E_ARG;
    if (!c->stream_on) return fail(c, RGKC_E_STATE, "no stream begun");
    if ((uint64_t)n_reads > c->cap_reads || offs[n_reads] > c->cap_bases || offs[0] != 0)
        return fail(c, RGKC_E_ARG, "batch exceeds the stream capacity (%lld reads, %llu bases)", (long long)n_reads,
                    (unsigned long long)offs[n_reads]);
    hipStream_t s = c->stream;
    // this batch's staging set was last used two batches ago: wait only for its uploads
    const int sl = (int)(c->n_batches & 1);
    if (c->stage_ev_rec[sl]) HIPC(c, hipEventSynchronize(c->stage_ev[sl]));
    uint8_t* const stg = c->h_stage[sl];
    uint64_t* const hh = c->h_hdr[sl];
    const uint64_t nb = offs[n_reads];
    const uint64_t prev = c->last_batch_bases;  // the device read buffer still holds these bases
    // the upload covers the previous batch's bases too (zeros past this batch's end)
    const uint64_t copy = std::min<uint64_t>(std::max(nb, prev) + 16, ((c->cap_bases >> 4) + 2) * 16);
    memcpy(stg, reads, nb);
    if (copy > nb) memset(stg + nb, 0, copy - nb);
    memcpy(c->h_offs_stage[sl], offs, (n_reads + 1) * 8);
    hh[0] = (uint64_t)n_reads;
    hh[1] = nb;
    HIPC(c, hipMemcpyAsync(c->reads.p, stg, copy, hipMemcpyHostToDevice, s));
    HIPC(c, hipMemcpyAsync(c->offs.p, c->h_offs_stage[sl], (n_reads + 1) * 8, hipMemcpyHostToDevice, s));
    HIPC(c, hipMemcpyAsync(c->hdr.p, hh, 16, hipMemcpyHostToDevice, s));
    HIPC(c, hipEventRecord(c->stage_ev[sl], s));
    c->stage_ev_rec[sl] = true;
    c->last_batch_bases = nb;
    if (c->stream_graph) HIPC(c, hipGraphLaunch(c->gexec, s));
    else if (int r = stream_enqueue(c)) return r;
    c->n_batches++;
    c->st.n_reads += (uint64_t)n_reads;
    c->st.n_bases += nb;
    return 0;
}

int rgkc_stream_end(rgkc_ctx* c, uint64_t* part_bytes, uint64_t* skm_cnt, uint64_t* kmer_cnt) {
    if (!c) return RGKC_E_ARG;
    if (!c->stream_on) return fail(c, RGKC_E_STATE, "no stream begun");
    hipStream_t s = c->stream;
    const int parts = c->parts;
    std::vector<uint64_t> cur(3 * (size_t)parts);
    HIPC(c, hipMemcpyAsync(cur.data(), c->cursors3.p, cur.size() * 8, hipMemcpyDeviceToHost, s));
    HIPC(c, hipStreamSynchronize(s));
    unpack_cursors(cur, parts);
    const uint64_t nbat = c->n_batches, nr = c->st.n_reads, nbs = c->st.n_bases;
    stream_free(c);
    if (int r = check_err_word(c)) return r;  // RGKC_E_OVERFLOW when a capacity was too small
    for (int i = 0; i < parts; i++) {
        if (cur[i] > c->h_part_off[i + 1] - c->h_part_off[i] || cur[parts + i] > c->h_rec_off[i + 1] - c->h_rec_off[i])
            return fail(c, RGKC_E_OVERFLOW, "partition %d exceeds its capacity", i);
        if (part_bytes) part_bytes[i] = cur[i];
        if (skm_cnt) skm_cnt[i] = cur[parts + i];
        if (kmer_cnt) kmer_cnt[i] = cur[2 * parts + i];
    }
    if (int r = finish_regions(c, cur)) return r;
    c->st.n_reads = nr; c->st.n_bases = nbs; c->st.buckets = nbat;
    return 0;
}

int rgkc_set_sizing(rgkc_ctx* c, const double* distinct_pred, int parts, float margin) {
    if (!c) return RGKC_E_ARG;
    if (!distinct_pred || parts <= 0) { c->pred_distinct.clear(); return 0; }
    c->pred_distinct.assign(distinct_pred, distinct_pred + parts);
    c->pred_margin = margin;
    return 0;
}

int rgkc_set_count_workgroups(rgkc_ctx* c, uint32_t per_cu) {
    if (!c) return RGKC_E_ARG;
    if (per_cu > 4) return fail(c, RGKC_E_ARG, "at most four count workgroups fit a CU");
    c->count_wg_per_cu = per_cu;
    return 0;
}

int rgkc_set_memory_limit(rgkc_ctx* c, uint64_t bytes) {
    if (!c) return RGKC_E_ARG;
    c->mem_limit = bytes;
    return 0;
}

int rgkc_set_tuning(rgkc_ctx* c, uint32_t inst_per_sweep) {
    if (!c) return RGKC_E_ARG;
    c->inst_cap = inst_per_sweep;
    return 0;
}

int rgkc_set_heavy_expand(rgkc_ctx* c, int mode) {
    if (!c || mode < 0 || mode > 1) return RGKC_E_ARG;
    c->heavy_mode = mode;
    return 0;
}

int rgkc_set_p1_lists(rgkc_ctx* c, uint32_t list_cap) {
    if (!c) return RGKC_E_ARG;
    c->list_cap = list_cap;
    return 0;
}

// ---- learned sizing: sample features -----------------------------------------------------------
int rgkc_sample_features(rgkc_ctx* c, int k, int p, int parts, uint32_t* hist, uint64_t* skm_cnt, uint64_t* kmer_cnt,
                         uint64_t* store_bytes) {
    if (!c || !hist) return RGKC_E_ARG;
    (void)hipGetLastError();
    HIPC(c, hipSetDevice(c->device));
    const size_t nbins = (size_t)1 << (2 * (p > 8 ? 8 : p));
    DevBuf h;
    NEED(c, h, nbins * 4);
    HIPC(c, hipMemsetAsync(h.p, 0, nbins * 4, c->stream));
    if (int r = do_phase1(c, k, p, parts, h.as<uint32_t>())) return r;
    HIPC(c, hipMemcpyAsync(hist, h.p, nbins * 4, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < parts; i++) {
        if (skm_cnt) skm_cnt[i] = c->h_rec_off[i + 1] - c->h_rec_off[i];
        if (kmer_cnt) kmer_cnt[i] = c->h_kmers[i + 1] - c->h_kmers[i];
        if (store_bytes) store_bytes[i] = c->h_part_off[i + 1] - c->h_part_off[i];
    }
    return 0;
}

// ---- load predictor network -------------------------------------------------------------------
struct rgkc_mlp {
    rgkc_ctx* c = nullptr;
    MlpDims d;
    DevBuf w32, wb, g, m1, m2, X, T, M, H1, H2, Y, dY, dH1, dH2, loss;
    int step = 0;
};

int rgkc_mlp_create(rgkc_ctx* c, int n_in, int h1, int h2, int n_out, rgkc_mlp** out) {
    if (!c || !out || n_in < 1 || h1 < 1 || h2 < 1 || n_out < 1) return RGKC_E_ARG;
    (void)hipGetLastError();
    HIPC(c, hipSetDevice(c->device));
    rgkc_mlp* m = new rgkc_mlp();
    m->c = c;
    m->d.n_in = n_in; m->d.h1 = h1; m->d.h2 = h2; m->d.n_out = n_out;
    const size_t P = m->d.total();
    if (m->w32.need(P * 4) || m->wb.need(P * 2) || m->g.need(P * 4) || m->m1.need(P * 4) || m->m2.need(P * 4) ||
        m->loss.need(64)) { delete m; return fail(c, RGKC_E_NOMEM, "mlp parameter allocation failed"); }
    HIPC(c, hipMemsetAsync(m->w32.p, 0, P * 4, c->stream));
    HIPC(c, hipMemsetAsync(m->wb.p, 0, P * 2, c->stream));
    HIPC(c, hipMemsetAsync(m->m1.p, 0, P * 4, c->stream));
    HIPC(c, hipMemsetAsync(m->m2.p, 0, P * 4, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    *out = m;
    return 0;
}

int rgkc_mlp_destroy(rgkc_mlp* m) {
    if (m) { (void)hipStreamSynchronize(m->c->stream); delete m; }
    return 0;
}

int rgkc_mlp_num_params(const rgkc_mlp* m, uint64_t* n) {
    if (!m || !n) return RGKC_E_ARG;
    *n = m->d.total();
    return 0;
}

int rgkc_mlp_set_params(rgkc_mlp* m, const float* p) {
    if (!m || !p) return RGKC_E_ARG;
    rgkc_ctx* c = m->c;
    const size_t P = m->d.total();
    HIPC(c, hipMemcpyAsync(m->w32.p, p, P * 4, hipMemcpyHostToDevice, c->stream));
    HIPC(c, mlp_to_bf16(m->w32.as<float>(), m->wb.p, P, c->stream));
    HIPC(c, hipMemsetAsync(m->m1.p, 0, P * 4, c->stream));
    HIPC(c, hipMemsetAsync(m->m2.p, 0, P * 4, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    m->step = 0;
    return 0;
}

int rgkc_mlp_get_params(rgkc_mlp* m, float* p) {
    if (!m || !p) return RGKC_E_ARG;
    rgkc_ctx* c = m->c;
    HIPC(c, hipMemcpyAsync(p, m->w32.p, m->d.total() * 4, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return 0;
}

static int mlp_stage_fwd(rgkc_mlp* m, const float* x, int B) {
    rgkc_ctx* c = m->c;
    const MlpDims& d = m->d;
    NEED(c, m->X, (size_t)B * d.n_in * 4);
    NEED(c, m->H1, (size_t)B * d.h1 * 4);
    NEED(c, m->H2, (size_t)B * d.h2 * 4);
    NEED(c, m->Y, (size_t)B * d.n_out * 4);
    HIPC(c, hipMemcpyAsync(m->X.p, x, (size_t)B * d.n_in * 4, hipMemcpyHostToDevice, c->stream));
    HIPC(c, mlp_forward(d, m->wb.p, m->w32.as<float>(), m->X.as<float>(), B, m->H1.as<float>(), m->H2.as<float>(),
                        m->Y.as<float>(), c->stream));
    return 0;
}

int rgkc_mlp_forward(rgkc_mlp* m, const float* x, int B, float* y) {
    if (!m || !x || !y || B < 1) return RGKC_E_ARG;
    rgkc_ctx* c = m->c;
    (void)hipGetLastError();
    HIPC(c, hipSetDevice(c->device));
    if (int r = mlp_stage_fwd(m, x, B)) return r;
    HIPC(c, hipMemcpyAsync(y, m->Y.p, (size_t)B * m->d.n_out * 4, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return 0;
}

static int mlp_stage_bwd(rgkc_mlp* m, const float* t, const float* mask, int B, float* loss) {
    rgkc_ctx* c = m->c;
    const MlpDims& d = m->d;
    const size_t nY = (size_t)B * d.n_out;
    NEED(c, m->T, nY * 4);
    NEED(c, m->dY, nY * 4);
    NEED(c, m->dH1, (size_t)B * d.h1 * 4);
    NEED(c, m->dH2, (size_t)B * d.h2 * 4);
    float msum = (float)nY;
    if (mask) {
        NEED(c, m->M, nY * 4);
        double sm = 0;
        for (size_t i = 0; i < nY; i++) sm += mask[i];
        msum = (float)sm;
        HIPC(c, hipMemcpyAsync(m->M.p, mask, nY * 4, hipMemcpyHostToDevice, c->stream));
    }
    HIPC(c, hipMemcpyAsync(m->T.p, t, nY * 4, hipMemcpyHostToDevice, c->stream));
    HIPC(c, mlp_backward(d, m->wb.p, m->X.as<float>(), B, m->H1.as<float>(), m->H2.as<float>(), m->Y.as<float>(),
                         m->T.as<float>(), mask ? m->M.as<float>() : nullptr, msum, m->dY.as<float>(),
                         m->dH1.as<float>(), m->dH2.as<float>(), m->g.as<float>(), m->loss.as<float>(), c->stream));
    if (loss) HIPC(c, hipMemcpyAsync(loss, m->loss.p, 4, hipMemcpyDeviceToHost, c->stream));
    return 0;
}

int rgkc_mlp_grad(rgkc_mlp* m, const float* x, const float* t, const float* mask, int B, float* grad, float* loss) {
    if (!m || !x || !t || B < 1) return RGKC_E_ARG;
    rgkc_ctx* c = m->c;
    (void)hipGetLastError();
    HIPC(c, hipSetDevice(c->device));
    if (int r = mlp_stage_fwd(m, x, B)) return r;
    if (int r = mlp_stage_bwd(m, t, mask, B, loss)) return r;
    if (grad) HIPC(c, hipMemcpyAsync(grad, m->g.p, m->d.total() * 4, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return 0;
}

int rgkc_mlp_train_step(rgkc_mlp* m, const float* x, const float* t, const float* mask, int B, float lr, float wd,
                        float* loss) {
    if (!m || !x || !t || B < 1) return RGKC_E_ARG;
    rgkc_ctx* c = m->c;
    (void)hipGetLastError();
    HIPC(c, hipSetDevice(c->device));
    if (int r = mlp_stage_fwd(m, x, B)) return r;
    if (int r = mlp_stage_bwd(m, t, mask, B, loss)) return r;
    m->step++;
    HIPC(c, mlp_adam(m->d.total(), m->w32.as<float>(), m->wb.p, m->g.as<float>(), m->m1.as<float>(),
                     m->m2.as<float>(), lr, m->step, wd, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return 0;
}

}  // extern "C"

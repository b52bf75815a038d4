// Internal interfaces between the kernel translation units and the C-ABI host layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rgkc {

constexpr int kScanTile = 4096;      // elements per workgroup tile in the device scans
constexpr int kP1Threads = 256;      // phase-1 workgroup size (one read segment per thread)
constexpr int kSegKmers = 128;       // k-mers per read segment (long reads are cut into segments)
constexpr int kMaxParts = 1024;      // partitions supported by the LDS histograms of phase 1
constexpr int kSubIds = 4096;
constexpr uint32_t kGroupSegments = 16;  // count stage: record segments per bucket in the grouping kernels
constexpr int kGroups = 1024;         // record groups per bucket in the count stage (4 sub-buckets each)        // minimizer sub-bucket ids per partition (12 bits)
constexpr int kCountThreads = 1024;  // count workgroup size
// packed stream-mode cursor: records in the low kRecBits bits, bytes above
constexpr int kRecBits = 28;
constexpr unsigned long long kRecMask = (1ull << kRecBits) - 1;

// ---- utilities (k_util.hip) ------------------------------------------------------------------
// Exclusive scan of n uint64 values; `tmp` must hold scan_tmp_words(n) uint64 words.
size_t scan_tmp_words(size_t n);
hipError_t scan_u64(const uint64_t* in, uint64_t* out, size_t n, uint64_t* tmp, hipStream_t s);
// Per-row exclusive scan of a rows x cols uint32 matrix (row-major), in place; row totals to
// `totals` (uint64, rows entries).
hipError_t rowscan_u32(uint32_t* m, uint64_t* totals, int rows, uint32_t cols, hipStream_t s);
// out[i] = a[i] + b[i] (n entries)
hipError_t add_u64(const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n, hipStream_t s);

// ---- phase 1 (k_phase1.hip) ------------------------------------------------------------------
struct SegTable {
    uint32_t* seg_read = nullptr;   // read index of each segment
    uint32_t* seg_first = nullptr;  // first k-mer of each segment (read-relative)
    uint64_t n_seg = 0;
};

hipError_t launch_pack(const uint8_t* reads, uint64_t n_bytes, uint32_t* bases2, hipStream_t s);
hipError_t launch_seg_count(const uint64_t* offs, uint64_t n_reads, int k, uint64_t* nseg, hipStream_t s);
hipError_t launch_seg_fill(const uint64_t* offs, const uint64_t* seg_off, uint64_t n_reads, int k,
                           uint32_t* seg_read, uint32_t* seg_first, hipStream_t s);

struct P1Args {
    const uint32_t* bases2;
    const uint64_t* offs;
    const uint32_t* seg_read;
    const uint32_t* seg_first;
    uint64_t n_seg;
    uint64_t n_bases;  // bases in the batch (bases2 holds n_bases/16 + 2 words)
    int k, p, parts;
    // count mode: per-(partition, workgroup) histograms, partition-major [parts][n_wg]
    uint32_t* h_bytes;
    uint32_t* h_recs;
    uint32_t* h_kmers;
    uint32_t n_wg;
    // emit mode
    const uint64_t* part_off;   // byte region start of each partition in `store`
    const uint64_t* part_end;   // byte region end (capacity check)
    const uint64_t* rec_off;    // side-descriptor region start of each partition
    const uint64_t* rec_end;
    const uint32_t* wg_byte_off;  // exact mode: [parts][n_wg] exclusive prefixes (nullptr: atomics)
    const uint32_t* wg_rec_off;
    unsigned long long* pos_cursor;   // atomic mode: bytes << kRecBits | records (relative to part_off / rec_off)
    uint8_t* store;
    uint32_t* side;
    uint32_t* err;
    uint32_t* mm_hist;  // optional: k-mer-weighted histogram of minimizer prefixes (top 8 bases)
    int hist_shift;     // mm >> hist_shift is the bin
    // stream mode (persistent, batched input): segment count on the device, k-mer totals
    const uint64_t* n_seg_dev;
    unsigned long long* kmer_cursor;
    // superkmer lists (exact mode, optional): the count pass leaves each workgroup's superkmers in
    // skm_list[wg * skm_cap ...] as 8-byte entries (skm_entry) and their number in skm_n[wg]; the
    // emit pass writes records from the list instead of scanning again. skm_n[wg] > skm_cap marks
    // a workgroup whose list did not fit, or held a superkmer an entry cannot describe: the emit
    // pass rescans its segments.
    uint64_t* skm_list;
    uint32_t* skm_n;
    uint32_t skm_cap;
    uint32_t* wg_rescan;        // count pass: tiles whose list overflowed, n_rescan[0] of them
    uint32_t* n_rescan;
    uint32_t* n_listorder;      // list emit: tiles with too many entries for store order (written in list order)
    const uint32_t* wg_map;     // scanning emit: tile of each workgroup (nullptr: blockIdx.x)
    uint32_t n_map;             // persistent scanning emit: tiles in wg_map
    uint32_t list_reserve;      // list emit: reserve each tile's runs with pos_cursor from its raw
                                // histogram columns (capacity-sized regions) instead of exact offsets
};
// list entry: minimizer | segment within the workgroup << 32 | (b - segment's first k-mer) << 40 |
// k-mers << 47 (the owned superkmer [b, b + k-mers) starts inside its segment: b - s0 < kSegKmers)
constexpr int kSkmKmerBits = 17;
static_assert(kSegKmers <= 128, "list entries keep b - s0 in 7 bits");
hipError_t launch_p1_count(const P1Args& a, hipStream_t s);
hipError_t launch_p1_emit(const P1Args& a, uint32_t n_rescan, hipStream_t s);
hipError_t launch_p1_emit_stream(const P1Args& a, unsigned grid, hipStream_t s);
// Capacity-sized regions from superkmer lists: the list emit over all tiles (list_reserve), then
// the persistent scanning emit over the n_rescan tiles whose list overflowed (a.wg_map).
hipError_t launch_p1_emit_sized(const P1Args& a, uint32_t n_rescan, hipStream_t s);
// Batched input: pack + segment table from a device header ([0] reads, [1] bases; [2] set to the
// segment count), all launched over the batch capacity.
hipError_t launch_batch_prep(const uint8_t* reads, const uint64_t* offs, uint64_t* hdr, uint64_t cap_reads,
                             uint64_t cap_bases, int k, uint32_t* bases2, uint64_t* nseg, uint64_t* seg_off,
                             uint64_t* scan_tmp, uint32_t* seg_read, uint32_t* seg_first, hipStream_t s);

// ---- count (k_count.hip) ---------------------------------------------------------------------
struct CountArgs {
    const uint8_t* store;       // superkmer records
    const uint32_t* side;       // side descriptors, one per record, in store order
    const uint64_t* b_byte;     // bucket byte ranges [b_byte[i], b_byte[i+1])
    const uint64_t* b_rec;      // bucket record ranges [b_rec[i], b_rec[i+1])
    const uint64_t* b_byte_end; // optional explicit ends (regions with gaps); nullptr: next start
    const uint64_t* b_rec_end;
    const uint32_t* b_part;     // partition of each bucket
    uint32_t n_buckets;
    uint64_t* scratch;          // one uint64 per record: word offset << 32 | side word, group order
    uint64_t scratch_base;      // record index of scratch[0] (the first record of the buckets counted)
    const uint64_t* row_base;   // per partition: first row of its region
    const uint64_t* row_end;    // per partition: end of its region
    unsigned long long* row_cursor;  // per partition: rows written so far
    void* out_keys;             // uint64 (k <= 32) or 2 x uint64 (k > 32) per row
    uint32_t* out_counts;
    int k;
    uint32_t cmin, cmax;
    uint32_t inst_cap;          // k-mer instances planned per sweep (distinct-capacity proxy)
    uint32_t count_wg_per_cu;   // resident count workgroups per CU (0: four)
    uint32_t* err;
    unsigned long long* stats;  // [0] sweeps, [1] overflow retries, [2] distinct (unfiltered)
    unsigned long long* part_distinct;  // per partition: distinct keys (unfiltered)
    uint32_t* sub_start;        // [n_buckets][kGroups + 1] descriptor start of each record group
    uint32_t* sub_sinst;        // [n_buckets][kSubIds] k-mer instances of each sub-bucket
    uint32_t* g_hist;           // grouping: [n_buckets][segments][2][kGroups] records, words per group
    uint64_t* g_seg_bytes;      // grouping: [n_buckets][segments + 1] store bytes per segment, then prefix
    uint32_t* words;            // records re-encoded as word-aligned 2-bit bases, group order
    const uint64_t* b_word;     // [n_buckets + 1] word region of each bucket
    uint64_t* item_cnt;         // [n_buckets + 1] work items per bucket, then their prefix
    uint64_t* items;            // bucket << 40 | s_lo << 27 | s_hi << 14 | slice << 7 | nslices
    uint64_t n_items;
    // Heavy sub-buckets (more instances than one work item) are expanded into keys once: the keys
    // of each hash slice get a region of 1 + cs key slots (slot 0 holds the fill cursor), and the
    // work items of such a sub-bucket (s_hi field 0) read keys instead of records.
    uint32_t heavy_on;          // 0: heavy sub-buckets are sliced over their records
    uint64_t* key_cnt;          // [n_buckets + 1] key slots per bucket, then their prefix
    uint64_t* exp_cnt;          // [n_buckets + 1] expansion tasks (groups) per bucket, then their prefix
    uint64_t* sub_key_off;      // [n_buckets][kSubIds] first key slot of each heavy sub-bucket
    void* heavy_keys;           // key slots (8 or 16 bytes)
    uint64_t* exp_items;        // bucket << 32 | group, for each group holding a heavy sub-bucket
    uint32_t* item_desc;        // [n_items][16] first-sweep descriptor of each work item (k_item_prep)
    // Slices whose region overflowed (a very frequent key) get an exact region in a second key area:
    uint32_t* sub_item0;        // [n_buckets][kSubIds] first work item of each heavy sub-bucket
    uint32_t* sub_key_cs;       // [n_buckets][kSubIds] key slots of each slice region of the sub-bucket
    uint64_t* ovf_n;            // [n_items + 1] slots of the item's exact region (0: none), then their prefix in ovf_off
    uint64_t* ovf_off;
    void* ovf_keys;
};
hipError_t launch_group(const CountArgs& a, hipStream_t s);
hipError_t launch_plan(const CountArgs& a, bool fill, hipStream_t s);
hipError_t launch_expand(const CountArgs& a, uint64_t n_exp, hipStream_t s);
// ovf_n of every item (after launch_expand); then, with ovf_off and ovf_keys set, the exact regions
hipError_t launch_ovf_sizes(const CountArgs& a, hipStream_t s);
hipError_t launch_expand_ovf(const CountArgs& a, uint64_t n_exp, hipStream_t s);
hipError_t launch_count_items(const CountArgs& a, hipStream_t s);

// Side descriptors for a store that lacks them (host-provided partitions): one workgroup per
// partition finds record starts by the full/non-full byte rule and recomputes each record's
// minimizer from its first k-mer.
struct IndexArgs {
    const uint8_t* store;
    const uint64_t* part_byte;  // [n_parts+1]
    uint32_t n_parts;
    int k, p, parts_total;      // parts_total: the partition count the sub-bucket id is defined over
    uint32_t* side;             // out: one per record
    uint64_t* rec_count;        // out: records per partition (pass 1) / used as offsets (pass 2)
    const uint64_t* rec_off;    // pass 2: record offsets per partition
    uint32_t* err;
};
hipError_t launch_index_count(const IndexArgs& a, hipStream_t s);
hipError_t launch_index_fill(const IndexArgs& a, hipStream_t s);

// ---- row order (k_sort.hip) ------------------------------------------------------------------
constexpr int kSortDigitBits = 12;
constexpr uint32_t kSortChunk = 4096;  // rows per histogram / scatter workgroup
// A run of rows to sort: rows [start, start + len) of the table, or [toff, toff + len) of the
// scratch buffer when `where` is 1; the next digit is key bits [shift, shift + nbits).
struct SortSeg {
    uint64_t start, toff;
    uint32_t len;
    uint16_t shift;
    uint8_t nbits, where;
};
struct SortArgs {
    void* keys;           // the table (W bytes per row) and its counts
    uint32_t* counts;
    void* tkeys;          // scratch rows
    uint32_t* tcounts;
    const SortSeg* leaf;  // segments sorted in LDS
    const SortSeg* big;   // segments distributed by their next digit
    const uint32_t* chunk_seg;   // big segment of each chunk
    const uint32_t* seg_chunk0;  // first chunk of each big segment
    uint32_t* hist;       // [n_big][hstride] bucket sizes
    uint32_t* cursor;     // [n_big][hstride] next free row of each bucket (segment-relative)
    uint32_t hstride;     // bins per segment row: 1 << (widest digit of the round), <= 1 << kSortDigitBits
};
int sort_leaf_cap(int key_bytes);
hipError_t launch_sort_leaf(const SortArgs& a, uint32_t n_leaf, int key_bytes, hipStream_t s);
hipError_t launch_sort_hist(const SortArgs& a, uint32_t n_chunks, int key_bytes, hipStream_t s);
hipError_t launch_sort_scatter(const SortArgs& a, uint32_t n_chunks, int key_bytes, hipStream_t s);

// ---- load predictor network (k_mlp.hip) ---------------------------------------------------
// Flat parameter layout: W1[n_in x h1] b1[h1] W2[h1 x h2] b2[h2] W3[h2 x n_out] b3[n_out].
struct MlpDims {
    int n_in = 0, h1 = 0, h2 = 0, n_out = 0;
    size_t off_W1() const { return 0; }
    size_t off_b1() const { return (size_t)n_in * h1; }
    size_t off_W2() const { return off_b1() + h1; }
    size_t off_b2() const { return off_W2() + (size_t)h1 * h2; }
    size_t off_W3() const { return off_b2() + h2; }
    size_t off_b3() const { return off_W3() + (size_t)h2 * n_out; }
    size_t total() const { return off_b3() + n_out; }
};
hipError_t mlp_to_bf16(const float* w, void* wb, size_t n, hipStream_t s);
hipError_t mlp_forward(const MlpDims& d, const void* wb, const float* w32, const float* X, int B, float* H1, float* H2,
                       float* Y, hipStream_t s);
hipError_t mlp_backward(const MlpDims& d, const void* wb, const float* X, int B, const float* H1, const float* H2,
                        const float* Y, const float* T, const float* mask, float mask_sum, float* dY, float* dH1,
                        float* dH2, float* G, float* loss, hipStream_t s);
hipError_t mlp_adam(size_t n, float* w, void* wb, const float* g, float* m1, float* m2, float lr, int step, float wd,
                    hipStream_t s);

}  // namespace rgkc

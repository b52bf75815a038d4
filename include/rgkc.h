/* rgkc.h -- C ABI of the MI355X (gfx950) k-mer counting path.
 *
 * Plain pointers and sizes; no C++ or torch types. Every function returns 0 on success and a
 * negative RGKC_E_* code on failure (rgkc_last_error() describes it). Entry points map onto the
 * reference's GPU interface as noted per function (file:line in the reference tree).
 *
 * A context owns one device, one HIP stream and grow-only device buffers. Calls on one context
 * are not thread-safe; use one context per host thread (the reference creates streams per call,
 * gpu_skmgen.cu:532-533, gpu_kmercounting.cu:373; here they are created once per context).
 *
 * Keys: W = 8 bytes (uint64) for k <= 32, W = 16 bytes (two uint64, low word first) for
 * 33 <= k <= 63. Table rows are (key, uint32 count) with min_count <= count <= max_count.
 */
#ifndef RGKC_H
#define RGKC_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    RGKC_OK = 0,
    RGKC_E_HIP = -1,        /* a HIP runtime call failed */
    RGKC_E_ARG = -2,        /* invalid argument (k, p, parts, sizes) */
    RGKC_E_NOMEM = -3,      /* device allocation failed */
    RGKC_E_OVERFLOW = -4,   /* a sized region overflowed (predicted sizing missed): recount exactly */
    RGKC_E_KERNEL = -5,     /* a kernel reported an error bit */
    RGKC_E_STATE = -6,      /* call order (e.g. table requested before counting) */
    RGKC_E_NODEV = -7       /* no HIP device */
};

typedef struct rgkc_ctx rgkc_ctx;

/* Per-stage device times of the last run (milliseconds, HIP events on the context stream): plan_ms
 * spans grouping and planning, expand_ms the heavy-key expansion, count_ms the count kernel. */
typedef struct {
    float pack_ms, segments_ms, p1_count_ms, p1_emit_ms, index_ms, plan_ms, count_ms, total_ms;
    uint64_t n_reads, n_bases, n_kmers, n_records, store_bytes, distinct, rows;
    uint64_t sweeps, sweep_retries, buckets, items;
    uint64_t table_rows_alloc;  /* rows allocated for the table (sum of partition regions) */
    uint32_t fallbacks;         /* partitions whose predicted table region overflowed (counted again, exactly) */
    uint32_t err_bits;
    uint32_t p1_rescans;        /* phase-1 tiles whose superkmer list overflowed (rescanned by the emit pass) */
    uint32_t heavy_groups;      /* count: record groups whose heavy sub-buckets were expanded into keys */
    float expand_ms;            /* count: key expansion of heavy sub-buckets (k_expand); count_ms is k_count_items */
    uint32_t p1_list_order;     /* phase-1 tiles with too many superkmers to write in store order (list order) */
    uint64_t ovf_keys;          /* count: key slots of the exact second-pass regions (slices that overflowed) */
    uint32_t count_groups;      /* count: partition groups counted one after another to fit the device memory */
} rgkc_stats;

/* GPUReset (gpu_skmgen.h:52, gpu_skmgen.cu:487-497): select the device, return free bytes
 * (0 on failure). Unlike the reference it does not reset the device (other contexts may live). */
size_t rgkc_device_reset(int device);
int rgkc_device_count(void);

int rgkc_create(int device, rgkc_ctx** out);
int rgkc_destroy(rgkc_ctx* ctx);
const char* rgkc_last_error(const rgkc_ctx* ctx);

/* Upload a read batch (CSR: concatenated bases, offs[n_reads+1]) into device memory; the device
 * copy stays resident for the following calls. Same layout as PinnedCSR (gpu_skmgen.h:29-42). */
int rgkc_set_reads(rgkc_ctx* ctx, const char* reads, const uint64_t* offs, int64_t n_reads);

/* Phase 1 on the resident batch: GenSuperkmerGPU (gpu_skmgen.h:54-58, gpu_skmgen.cu:504-718).
 * Outputs per-partition store sizes; the store stays on the device. part_offs has parts+1
 * entries (byte offsets of each partition in the batch store), skm_cnt/kmer_cnt parts entries.
 * Any output pointer may be NULL. */
int rgkc_superkmers(rgkc_ctx* ctx, int k, int p, int parts, uint64_t* part_offs, uint64_t* skm_cnt,
                    uint64_t* kmer_cnt);
/* Copy the batch store (part_offs[parts] bytes, partition-major) to host memory. */
int rgkc_copy_store(rgkc_ctx* ctx, uint8_t* dst);

/* Phase 2 on the store produced by rgkc_superkmers (kmc_counting_GPU_streams,
 * gpu_kmercounting.h:14-17): count every partition. row_counts[parts] receives the rows of each
 * partition after the [min_count, max_count] filter; distinct (may be NULL) the unfiltered
 * distinct total (the reference's return value). */
int rgkc_count(rgkc_ctx* ctx, uint32_t min_count, uint32_t max_count, uint64_t* row_counts, uint64_t* distinct);

/* Phase 2 on host partitions (a batch store from the host, e.g. SKMStoreNoncon chunks): bytes
 * of n_parts partitions back to back, part_offs[n_parts+1]; part_ids[n_parts] their partition
 * ids out of parts_total (0 if unknown: the count then spreads each partition's minimizers over
 * its sub-buckets by a hash); p is the minimizer length used to build the store (0: recovered from
 * the records, rgkc_store_minimizer_length). */
int rgkc_count_host_store(rgkc_ctx* ctx, const uint8_t* store, const uint64_t* part_offs, int n_parts,
                          const int32_t* part_ids, int parts_total, int k, int p, uint32_t min_count,
                          uint32_t max_count, uint64_t* row_counts, uint64_t* distinct);

/* The minimizer length a store was built with, recovered from its records (no device needed; no
 * counterpart in the reference, whose counting call does not receive it, gpu_kmercounting.h:14-17):
 * the one length under which every k-mer of each sampled record has the same minimizer. Returns
 * RGKC_E_ARG if the records do not decide it. rgkc_count_host_store calls it for p == 0. */
int rgkc_store_minimizer_length(const uint8_t* store, const uint64_t* part_offs, int n_parts, int k, int* p);

/* Copy the table of the last count to host memory: partition i's rows go to
 * keys[row_offs[i] .. row_offs[i+1]) where row_offs is the prefix of row_counts. keys holds
 * W bytes per row. Within a partition, rows are grouped by minimizer sub-bucket and ascending by
 * key inside each group (rgkc_sort_rows orders them fully). */
int rgkc_copy_table(rgkc_ctx* ctx, void* keys, uint32_t* counts);

/* Order each partition's rows by ascending key on the device (the reference's table order,
 * cpu_funcs.cpp:614-640) before rgkc_copy_table. */
int rgkc_sort_rows(rgkc_ctx* ctx);

/* No-allocation variant of rgkc_sort_rows. It has no counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). It orders the rows of a caller-owned device table in place, ascending by
 * key within each partition: partition i holds rows [row_base[i], row_base[i] + row_counts[i]) of
 * keys (device, W bytes per row) and counts (device, uint32 per row); row_base and row_counts are
 * host arrays of n_parts entries. It allocates nothing: its working memory is a fixed device area
 * of RGKC_NOALLOC_SCRATCH_BYTES per GPU, held by the library until the process ends, so a call takes
 * at most RGKC_NOALLOC_MAX_ROWS(k) rows in all (RGKC_E_ARG beyond). Every call on one GPU shares that
 * area: calls that overlap on one GPU (other contexts, other host threads) race. */
#define RGKC_NOALLOC_ROW_BYTES (64u << 20)
#define RGKC_NOALLOC_SCRATCH_BYTES (80u << 20)
#define RGKC_NOALLOC_MAX_ROWS(k) (RGKC_NOALLOC_ROW_BYTES / ((k) <= 32 ? 12u : 20u))
int rgkc_sort_table_noalloc(rgkc_ctx* ctx, int k, void* keys, uint32_t* counts, const uint64_t* row_base,
                            const uint64_t* row_counts, int n_parts);

/* One full pass over the resident batch: phase 1 + phase 2, no host copies of results. */
int rgkc_run(rgkc_ctx* ctx, int k, int p, int parts, uint32_t min_count, uint32_t max_count);

int rgkc_get_stats(const rgkc_ctx* ctx, rgkc_stats* out);

/* Learned sizing (no counterpart in the reference; the reference sizes by instance counts,
 * gkc_main.cpp:324-352, gpu_hash_counter.cu:407-419): predicted distinct k-mers per partition.
 * The next count sizes each partition's table region to distinct*(1+margin) and its sweeps to
 * the predicted distinct/instance ratio; if a region overflows, the count is redone with exact
 * sizing and rgkc_stats.fallbacks is set. NULL clears the prediction. */
int rgkc_set_sizing(rgkc_ctx* ctx, const double* distinct_pred, int parts, float margin);

/* Phase 1 on the resident batch into capacity-sized partition regions (e.g. from the learned
 * predictor): one persistent emit pass, no sizing pass and no host round trip for sizes.
 * Returns RGKC_E_OVERFLOW if a capacity was too small (then size exactly with rgkc_superkmers). */
int rgkc_superkmers_sized(rgkc_ctx* ctx, int k, int p, int parts, const uint64_t* cap_bytes, const uint64_t* cap_recs,
                          uint64_t* part_bytes, uint64_t* skm_cnt, uint64_t* kmer_cnt);

/* Batched input (the reference's phase 1 runs per loader batch, gkc_main.cpp:77-89): many
 * batches of uneven reads are appended into device partition regions sized by cap_bytes /
 * cap_recs (e.g. from the learned predictor). Each batch is one H2D copy plus one replay of a
 * captured hipGraph (use_graph) of persistent kernels that read the batch size from device
 * memory. rgkc_stream_end returns RGKC_E_OVERFLOW if a region was too small (recount with exact
 * sizing), else the store is ready for rgkc_count. */
int rgkc_stream_begin(rgkc_ctx* ctx, int k, int p, int parts, uint64_t max_reads, uint64_t max_bases,
                      const uint64_t* cap_bytes, const uint64_t* cap_recs, int use_graph);
int rgkc_stream_add(rgkc_ctx* ctx, const char* reads, const uint64_t* offs, int64_t n_reads);
int rgkc_stream_end(rgkc_ctx* ctx, uint64_t* part_bytes, uint64_t* skm_cnt, uint64_t* kmer_cnt);

/* Device memory the count stage may plan with (0 = all it finds free; no counterpart in the
 * reference, whose VRAM grouping is gkc_main.cpp:324-352). Partitions are counted in groups whose
 * working set fits; when the table regions by k-mer instances do not fit beside them, each group is
 * counted into a temporary table and its rows are compacted into the final one. Results are the same
 * at every limit. */
int rgkc_set_memory_limit(rgkc_ctx* ctx, uint64_t bytes);

/* Tunables (0 = default): resident workgroups per CU of the persistent count kernel (1..4; its LDS
 * table allows four). Chosen by the autotuner (rapidgkc_amd/autotune.py). */
int rgkc_set_count_workgroups(rgkc_ctx* ctx, uint32_t per_cu);

/* Tunables (0 = default): instances planned per count sweep. */
int rgkc_set_tuning(rgkc_ctx* ctx, uint32_t inst_per_sweep);
/* Phase-1 superkmer lists (exact mode; no counterpart in the reference): the count pass keeps up to
 * list_cap superkmers per workgroup for the emit pass, which writes records from them instead of
 * scanning the reads again (a workgroup with more is rescanned). 0 = default (3x the superkmers a
 * random sequence gives), UINT32_MAX = no lists (the emit pass always scans). Lists that would
 * take more than half the free device memory are skipped by themselves. */
int rgkc_set_p1_lists(rgkc_ctx* ctx, uint32_t list_cap);
/* Heavy sub-buckets (more k-mer instances than one count work item; no counterpart in the
 * reference): mode 0 (default) extracts their keys once into per-hash-slice regions when half the
 * free device memory holds them, so each slice's sweep reads keys instead of every record of the
 * sub-bucket; mode 1 always reads records. */
int rgkc_set_heavy_expand(rgkc_ctx* ctx, int mode);

/* Features of a read sample for the learned sizing (no counterpart in the reference): runs phase 1
 * on the resident batch (the sample) and returns the k-mer-weighted histogram of minimizer prefixes
 * (the top min(p, 8) bases: 4^8 bins for p >= 8) plus the per-partition superkmer, k-mer and byte
 * counts of the sample. */
int rgkc_sample_features(rgkc_ctx* ctx, int k, int p, int parts, uint32_t* hist, uint64_t* skm_cnt, uint64_t* kmer_cnt,
                         uint64_t* store_bytes);

/* ---- learned load predictor network (no counterpart in the reference) -----------------------
 * A 3-layer ReLU MLP, BF16 weights with an FP32 master copy, FP32 accumulation, masked-MSE loss
 * and Adam, run by the library's own GEMM kernels. Parameters are one flat FP32 array:
 * W1[n_in x h1] b1[h1] W2[h1 x h2] b2[h2] W3[h2 x n_out] b3[n_out] (row-major, y = x W + b).
 * x: B x n_in, targets/mask/y: B x n_out (host arrays). Gradients are w.r.t. the BF16 weights. */
typedef struct rgkc_mlp rgkc_mlp;
int rgkc_mlp_create(rgkc_ctx* ctx, int n_in, int h1, int h2, int n_out, rgkc_mlp** out);
int rgkc_mlp_destroy(rgkc_mlp* m);
int rgkc_mlp_num_params(const rgkc_mlp* m, uint64_t* n);
int rgkc_mlp_set_params(rgkc_mlp* m, const float* params);
int rgkc_mlp_get_params(rgkc_mlp* m, float* params);
int rgkc_mlp_forward(rgkc_mlp* m, const float* x, int B, float* y);
int rgkc_mlp_grad(rgkc_mlp* m, const float* x, const float* target, const float* mask, int B, float* grad, float* loss);
int rgkc_mlp_train_step(rgkc_mlp* m, const float* x, const float* target, const float* mask, int B, float lr,
                        float weight_decay, float* loss);

#ifdef __cplusplus
}
#endif
#endif

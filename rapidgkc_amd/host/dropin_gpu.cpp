// Drop-in replacement for the reference's GPU translation units (gpu_skmgen.cu and
// gpu_kmercounting.cu): the same exported C++ symbols, implemented over the C ABI of the gfx950
// library (include/rgkc.h). Compiled against the reference's own headers, so a maintainer swaps
// the two .cu files for this file (plus librgkc.so) and keeps gkc_main.cpp, cpu_funcs.cpp and the
// rest unchanged.
//
// Symbols (reference declarations):
//   PinnedCSR::PinnedCSR / ~PinnedCSR          gpu_skmgen.h:29-42
//   size_t GPUReset(int)                        gpu_skmgen.h:52
//   void GenSuperkmerGPU(...)                   gpu_skmgen.h:54-58
//   size_t kmc_counting_GPU_streams(...)        gpu_kmercounting.h:14-17
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <unistd.h>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "gpu_kmercounting.h"
#include "gpu_skmgen.h"
#include "rgkc.h"

extern Logger* logger;

namespace {

// The counting call needs the minimizer length (to recover each record's minimizer, hence its
// sub-bucket) and, for the sub-bucket spread, the partition count; the reference's signature
// passes neither (gpu_kmercounting.h:14-17). Phase 1 on the GPU records both here; a count whose
// stores came from elsewhere (the CPU phase 1, .skm files) recovers the minimizer length from the
// records and spreads minimizers over sub-buckets by a hash (rgkc_count_host_store with 0, 0).
std::atomic<int> g_p1_p{0}, g_p1_parts{0};

[[noreturn]] void die(const char* what, rgkc_ctx* c, int rc) {
    std::cerr << "rgkc error calling \"" << what << "\", code is " << rc << ": " << (c ? rgkc_last_error(c) : "")
              << std::endl;  // the reference's CUDA_CHECK prints and exits (gpu_skmgen.h:4-9)
    std::cerr.flush();
    _exit(1);  // no static destructors while other worker threads still use the runtime
}

// One context per (host thread, device): created on first use, kept for the process.
rgkc_ctx* ctx_for(int device) {
    thread_local std::map<int, rgkc_ctx*> ctxs;
    auto it = ctxs.find(device);
    if (it != ctxs.end()) return it->second;
    rgkc_ctx* c = nullptr;
    const int rc = rgkc_create(device, &c);
    if (rc) die("rgkc_create", nullptr, rc);
    ctxs[device] = c;
    return c;
}

std::atomic<int> g_next_device{0};

}  // namespace

PinnedCSR::PinnedCSR(std::vector<ReadPtr>& reads, bool keep_original) {
    (void)keep_original;
    n_reads = (T_read_cnt)reads.size();
    size_capacity = 0;
    for (const ReadPtr& r : reads) size_capacity += r.len;
    if (hipHostMalloc((void**)&reads_offs, (n_reads + 1) * sizeof(T_CSR_cap), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&reads_CSR, size_capacity + 1, hipHostMallocDefault) != hipSuccess) {
        std::cerr << "hipHostMalloc failed for " << size_capacity << " bytes" << std::endl;
        exit(1);
    }
    char* p = reads_CSR;
    reads_offs[0] = 0;
    for (T_read_cnt i = 0; i < n_reads; i++) {
        memcpy(p, reads[i].read, reads[i].len);
        p += reads[i].len;
        reads_offs[i + 1] = reads_offs[i] + reads[i].len;
    }
}

PinnedCSR::~PinnedCSR() {
    (void)hipHostFree(reads_offs);
    (void)hipHostFree(reads_CSR);
}

size_t GPUReset(int did) { return rgkc_device_reset(did); }

void GenSuperkmerGPU(PinnedCSR& pinned_reads, const T_kvalue K_kmer, const T_kvalue P_minimizer, bool HPC,
                     CUDAParams& gpars, CountTask task, const int SKM_partitions,
                     std::vector<SKMStoreNoncon*> skm_partition_stores, int tid) {
    if (HPC) { std::cerr << "homopolymer compression is not supported by the gfx950 path" << std::endl; exit(1); }
    (void)task;
    const int n_dev = gpars.n_devices > 0 ? gpars.n_devices : 1;
    const int device = gpars.device_id.fetch_add(1) % n_dev;  // round-robin, gpu_skmgen.cu:515
    rgkc_ctx* c = ctx_for(device);
    g_p1_p.store(P_minimizer);
    g_p1_parts.store(SKM_partitions);
    WallClockTimer wct;
    int rc = rgkc_set_reads(c, pinned_reads.reads_CSR, pinned_reads.reads_offs, pinned_reads.n_reads);
    if (rc) die("rgkc_set_reads", c, rc);
    std::vector<uint64_t> po(SKM_partitions + 1), sc(SKM_partitions), kc(SKM_partitions);
    rc = rgkc_superkmers(c, K_kmer, P_minimizer, SKM_partitions, po.data(), sc.data(), kc.data());
    if (rc) die("rgkc_superkmers", c, rc);
    // The batch store is handed to the partition stores, which keep (memory mode) or write and
    // drop (disk mode) it, as save_batch_skms does for the reference (skmstore.hpp:222-236).
    u_char* buf = new u_char[po[SKM_partitions] + 1];
    rc = rgkc_copy_store(c, buf);
    if (rc) die("rgkc_copy_store", c, rc);
    std::vector<T_skm_partsize> skm_cnt(sc.begin(), sc.end()), kmer_cnt(kc.begin(), kc.end());
    std::vector<T_CSR_cap> offs(po.begin(), po.end());
    SKMStoreNoncon::save_batch_skms(skm_partition_stores, skm_cnt.data(), kmer_cnt.data(), offs.data(), buf);
    if (skm_partition_stores[0]->to_file) delete[] buf;
    logger->log("-- BATCH  GPU (T" + std::to_string(tid) + "): #reads: " + std::to_string(pinned_reads.n_reads) +
                " --  " + std::to_string(wct.stop()));
}

size_t kmc_counting_GPU_streams(T_kvalue k, std::vector<SKMStoreNoncon*> skms_stores, CUDAParams& gpars,
                                T_kmer_cnt kmer_min_freq, T_kmer_cnt kmer_max_freq,
                                std::vector<T_kmc> kmc_result_curthread[], int gpuid, int tid) {
    (void)gpars;
    rgkc_ctx* c = ctx_for(gpuid);
    const int n = (int)skms_stores.size();
    // Concatenate the partitions' bytes (chunks in memory mode, the .skm file in disk mode),
    // as Extract_Kmers does before its H2D copy (gpu_kmercounting.cu:247-307).
    std::vector<uint64_t> po(n + 1, 0);
    for (int i = 0; i < n; i++) po[i + 1] = po[i] + skms_stores[i]->tot_size_bytes;
    std::vector<uint8_t> buf(po[n] + 1);
    std::vector<int32_t> ids(n);
    for (int i = 0; i < n; i++) {
        SKMStoreNoncon* st = skms_stores[i];
        ids[i] = st->id;
        if (st->to_file) {
            st->load_from_file();
            memcpy(buf.data() + po[i], st->skms_from_file, st->tot_size_bytes);
        } else {
            size_t pos = po[i];
            SKM chunk;
            while (st->skms.try_dequeue(chunk)) {
                memcpy(buf.data() + pos, chunk.skm_chunk, chunk.chunk_bytes);
                pos += chunk.chunk_bytes;
            }
            if (pos != po[i + 1]) { std::cerr << "partition " << st->id << ": store size mismatch" << std::endl; exit(1); }
        }
    }
    std::vector<uint64_t> rows(n);
    uint64_t distinct = 0;
    int rc = rgkc_count_host_store(c, buf.data(), po.data(), n, ids.data(), g_p1_parts.load(), k, g_p1_p.load(),
                                   kmer_min_freq, kmer_max_freq, rows.data(), &distinct);
    if (rc) die("rgkc_count_host_store", c, rc);
    uint64_t tot = 0;
    for (auto r : rows) tot += r;
    const int W = k <= 32 ? 8 : 16;
    std::vector<uint8_t> keys(tot * W + 16);
    std::vector<uint32_t> cnts(tot + 1);
    // each partition's rows ascending by key, the order of the reference's table (cpu_funcs.cpp:614-640)
    rc = rgkc_sort_rows(c);
    if (rc) die("rgkc_sort_rows", c, rc);
    rc = rgkc_copy_table(c, keys.data(), cnts.data());
    if (rc) die("rgkc_copy_table", c, rc);
    // Fill the result vectors the reference leaves empty (gpu_kmercounting.cu:505-507).
    uint64_t o = 0;
    for (int i = 0; i < n; i++) {
        std::vector<T_kmc>& out = kmc_result_curthread[skms_stores[i]->id];
        out.reserve(out.size() + rows[i]);
        for (uint64_t r = 0; r < rows[i]; r++, o++) {
            T_kmer key = 0;
            memcpy(&key, keys.data() + o * W, sizeof(T_kmer) < (size_t)W ? sizeof(T_kmer) : (size_t)W);
            out.push_back(T_kmc{key, cnts[o]});
        }
    }
    std::string logs = "GPU " + std::to_string(gpuid) + "\t(T" + std::to_string(tid) + "):";
    for (int i = 0; i < n; i++) {
        logs += "\tPart " + std::to_string(skms_stores[i]->id) + " " + std::to_string(skms_stores[i]->tot_size_bytes) +
                "|" + std::to_string(skms_stores[i]->kmer_cnt);
        skms_stores[i]->clear_skm_data();
        delete skms_stores[i];  // the counting call owns its stores, gpu_kmercounting.cu:532-535
    }
    logger->log(logs + " " + std::to_string(distinct), Logger::LV_DEBUG);
    return (size_t)distinct;
}

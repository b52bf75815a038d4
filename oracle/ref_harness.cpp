// TEST INFRASTRUCTURE ONLY: a C ABI around the reference's own CPU translation unit
// (cpu_funcs.cpp, built from its source under the reference tree by oracle/Makefile into
// oracle/_ref/libref_cpu*.so). Used to pin the oracle restatement and as the timed CPU baseline.
// Nothing here reimplements the reference: every call goes into cpu_funcs.cpp.
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>
#include "cpu_funcs.h"
#include "utilities.hpp"

Logger* logger = nullptr;  // cpu_funcs.cpp logs through this global (defined by gkc_main.cpp upstream)

// Non-static functions of cpu_funcs.cpp that cpu_funcs.h does not declare.
T_read_len gen_skm_offs(u_char* read, T_read_len len, const T_kvalue K_kmer, const T_kvalue P_minimizer,
                        T_read_len* skm_offs, T_minimizer* minimizers);
void read_compression(u_char* read, T_read_len len);
void extract_kmers_cpu(SKMStoreNoncon& skms_store, T_kvalue k, T_kmer* kmers, unsigned int n_threads);
void sort_kmers(T_kmer* kmers, const T_skm_partsize n_kmers, const T_kvalue K_kmer, int i_base);

static Logger& quiet_logger() {
    static Logger lg(0, Logger::LV_ERROR, false, "./");
    logger = &lg;
    return lg;
}

extern "C" {

int ref_kmer_bytes() { return (int)sizeof(T_kmer); }

int ref_gen_skm_offs(const char* read, int len, int k, int p, int32_t* offs, uint32_t* mms) {
    quiet_logger();
    std::vector<u_char> buf(read, read + len);
    return gen_skm_offs(buf.data(), len, (T_kvalue)k, (T_kvalue)p, offs, mms);
}

void ref_read_compression(const char* read, int len, uint8_t* out) {
    std::vector<u_char> buf(read, read + len + 1);
    read_compression(buf.data(), len);
    memcpy(out, buf.data(), len / 3 + 1);
}

// GenSuperkmerCPU over one batch (in memory, like -im). Stores are returned as one
// partition-major byte buffer; caller frees with ref_free.
int ref_gen_superkmers(const char* reads, const uint64_t* offs, int64_t n_reads, int k, int p, int parts,
                       uint8_t** store_out, uint64_t* part_offs, uint64_t* skm_cnt, uint64_t* kmer_cnt) {
    quiet_logger();
    std::vector<char> copy(reads, reads + offs[n_reads]);  // GenSuperkmerCPU packs reads in place
    std::vector<ReadPtr> rv;
    for (int64_t i = 0; i < n_reads; i++) {
        int len = (int)(offs[i + 1] - offs[i]);
        if (len >= k) rv.push_back(ReadPtr{copy.data() + offs[i], len});
    }
    std::vector<SKMStoreNoncon*> stores;
    for (int i = 0; i < parts; i++) stores.push_back(new SKMStoreNoncon(i, false));
    GenSuperkmerCPU(rv, (T_kvalue)k, (T_kvalue)p, false, parts, stores, 0);
    uint64_t tot = 0;
    for (int i = 0; i < parts; i++) tot += stores[i]->tot_size_bytes;
    uint8_t* out = (uint8_t*)malloc(tot ? tot : 1);
    uint64_t pos = 0;
    for (int i = 0; i < parts; i++) {
        part_offs[i] = pos;
        skm_cnt[i] = stores[i]->skm_cnt;
        kmer_cnt[i] = stores[i]->kmer_cnt;
        SKM chunk;
        while (stores[i]->skms.try_dequeue(chunk)) { memcpy(out + pos, chunk.skm_chunk, chunk.chunk_bytes); pos += chunk.chunk_bytes; }
        stores[i]->clear_skm_data();
        delete stores[i];
    }
    part_offs[parts] = pos;
    *store_out = out;
    return pos == tot ? 0 : -1;
}

void ref_free(void* p) { free(p); }

// extract_kmers_cpu + sort_kmers on one partition's bytes; returns the number of keys written.
int64_t ref_extract_sorted(const uint8_t* skms, uint64_t nbytes, uint64_t skm_cnt, uint64_t kmer_cnt, int k, void* keys) {
    quiet_logger();
    SKMStoreNoncon st(0, false);
    u_char* chunk = new u_char[nbytes ? nbytes : 1];
    memcpy(chunk, skms, nbytes);
    st.add_skms(chunk, nbytes, skm_cnt, kmer_cnt);
    st.delete_list.enqueue(chunk);
    T_kmer* out = (T_kmer*)keys;
    extract_kmers_cpu(st, (T_kvalue)k, out, 1);
    sort_kmers(out, kmer_cnt, (T_kvalue)k, 0);
    st.clear_skm_data();
    return (int64_t)kmer_cnt;
}

// KmerCountingCPU on one partition; returns its distinct count (the reference's only result).
int64_t ref_count_partition(const uint8_t* skms, uint64_t nbytes, uint64_t skm_cnt, uint64_t kmer_cnt, int k) {
    quiet_logger();
    SKMStoreNoncon* st = new SKMStoreNoncon(0, false);
    u_char* chunk = new u_char[nbytes ? nbytes : 1];
    memcpy(chunk, skms, nbytes);
    st->add_skms(chunk, nbytes, skm_cnt, kmer_cnt);
    st->delete_list.enqueue(chunk);
    std::vector<T_kmc> res;
    return (int64_t)KmerCountingCPU((T_kvalue)k, st, 1, 0xFFFFFFFFu, res, 0, 1);  // deletes st
}

// The reference CPU pipeline as gkc_main.cpp runs it with -cpuonly -im: phase 1 on read batches
// of `batch` reads over `threads` workers, then KmerCountingCPU per partition over `threads`
// workers. Returns the total distinct count; phase times in seconds.
int64_t ref_cpu_pipeline(const char* reads, const uint64_t* offs, int64_t n_reads, int k, int p, int parts,
                         int threads, int batch, double* t_phase1, double* t_phase2) {
    quiet_logger();
    std::vector<char> copy(reads, reads + offs[n_reads]);
    std::vector<SKMStoreNoncon*> stores;
    for (int i = 0; i < parts; i++) stores.push_back(new SKMStoreNoncon(i, false));
    auto t0 = std::chrono::steady_clock::now();
    std::atomic<int64_t> next{0};
    auto w1 = [&](int tid) {
        for (;;) {
            int64_t b = next.fetch_add(batch);
            if (b >= n_reads) break;
            int64_t e = std::min<int64_t>(n_reads, b + batch);
            std::vector<ReadPtr> rv;
            for (int64_t i = b; i < e; i++) {
                int len = (int)(offs[i + 1] - offs[i]);
                if (len >= k) rv.push_back(ReadPtr{copy.data() + offs[i], len});
            }
            GenSuperkmerCPU(rv, (T_kvalue)k, (T_kvalue)p, false, parts, stores, tid);
        }
    };
    {
        std::vector<std::thread> th;
        for (int t = 0; t < threads; t++) th.emplace_back(w1, t);
        for (auto& t : th) t.join();
    }
    auto t1 = std::chrono::steady_clock::now();
    std::atomic<int> nextp{0};
    std::atomic<int64_t> distinct{0};
    auto w2 = [&](int tid) {
        for (;;) {
            int i = nextp.fetch_add(1);
            if (i >= parts) break;
            std::vector<T_kmc> res;
            distinct += (int64_t)KmerCountingCPU((T_kvalue)k, stores[i], 1, 0xFFFFFFFFu, res, tid, 1);
        }
    };
    {
        std::vector<std::thread> th;
        for (int t = 0; t < threads; t++) th.emplace_back(w2, t);
        for (auto& t : th) t.join();
    }
    auto t2 = std::chrono::steady_clock::now();
    *t_phase1 = std::chrono::duration<double>(t1 - t0).count();
    *t_phase2 = std::chrono::duration<double>(t2 - t1).count();
    return distinct.load();
}

}  // extern "C"

// rgkc -- command-line k-mer counter on the gfx950 path.
//
// Flags mirror the reference CLI (utilities.hpp:222-265: -t -tgpu -t2 -tgz -tsort -k -p -skm -hpc
// -tmp -im -fb -rb -grid -block -grid2 -block2 -ngpu -ns1 -ns2 -rps -cpuonly -gpuonly -log -lv
// -read <files...>), same defaults (utilities.hpp:187-218) and the same quirk that the loop stops
// at argc-1 (a trailing bare flag is ignored). Added: -ci/-cx count thresholds (inclusive,
// defaults 1/1000 as utilities.hpp:190 documents), -o table output, -gb bases per GPU batch, -ds 0
// to keep -im stores on the host.
// Launch-shape flags (-grid/-block/...) are accepted and have no effect: launch shapes are fixed
// per kernel on gfx950. -cpuonly and -hpc 1 are refused: this binary has no CPU counter and the
// reference's homopolymer mode is out of scope.
//
// Pipeline: load reads (FASTA 2-line / FASTQ 4-line records, optionally gzip, format chosen by the
// first file name as fileloader.hpp:430,731 do; reads shorter than k dropped) -> GPU batches ->
// phase 1 on the device -> partition stores: with -im on the device (the batched path appends
// every batch into device partition regions, and phase 2 counts them in place), or in the
// reference's byte format on the host, in memory (-im -ds 0, or when the device store does not fit)
// or appended to <tmp>/<id>.skm (skmstore.hpp:129-151), then counted per partition group sized to
// the device memory -> table rows filtered to [ci, cx] -> optional binary table file.
// Overlap: a loader thread reads and parses the next read batch while the device runs phase 1 on
// the current one (the reference's loader pipeline, fileloader.hpp:492-525), and in phase 2 the
// next partition group's stores are read on a second thread while the device counts the current
// group (gkc_main.cpp:332-336).
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <future>
#include <mutex>
#include <thread>
#include <unistd.h>
#include <cstdarg>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rgkc.h"

namespace {

struct Params {
    int K = 21, P = 7, parts = 31;
    unsigned cmin = 1, cmax = 1000;
    bool hpc = false, to_file = true, cpu_only = false, gpu_only = false, dev_store = true;
    std::string tmp = "/home/cyr/tmp/", log = "./log/", out;
    int lv = 5, threads = 4, n_devices = 1;
    long long rb = 2000, gpu_batch_bases = 1ll << 30, fb = 256;
    std::vector<std::string> reads;
};

void usage(const char* a0) {
    fprintf(stderr, "Usage: %s -t <threads_per_process> -k <K_kmer> -p <P_minimizer> -skm <SKM_partitions> -hpc <HPC> "
                    "-tmp <tmp_file_folder> [-im] [-ds 0|1] [-ci <min_count>] [-cx <max_count>] [-o <table_file>] "
                    "-read <read_files...>\n", a0);
}

Params parse(int argc, char** argv) {
    Params P;
    for (int i = 1; i < argc - 1; i++) {
        const char* a = argv[i];
        auto next = [&]() { return argv[++i]; };
        if (!strcmp(a, "-t")) P.threads = atoi(next());
        else if (!strcmp(a, "-tgpu") || !strcmp(a, "-t2") || !strcmp(a, "-tgz") || !strcmp(a, "-tsort") ||
                 !strcmp(a, "-grid") || !strcmp(a, "-block") || !strcmp(a, "-grid2") || !strcmp(a, "-block2") ||
                 !strcmp(a, "-ns1") || !strcmp(a, "-ns2") || !strcmp(a, "-rps")) (void)next();
        else if (!strcmp(a, "-k")) P.K = atoi(next());
        else if (!strcmp(a, "-p")) P.P = atoi(next());
        else if (!strcmp(a, "-skm")) P.parts = atoi(next());
        else if (!strcmp(a, "-hpc")) P.hpc = atoi(next()) != 0;
        else if (!strcmp(a, "-tmp")) P.tmp = next();
        else if (!strcmp(a, "-im")) P.to_file = false;
        else if (!strcmp(a, "-fb")) P.fb = atoll(next());
        else if (!strcmp(a, "-rb")) P.rb = atoll(next());
        else if (!strcmp(a, "-gb")) P.gpu_batch_bases = atoll(next());
        else if (!strcmp(a, "-ds")) P.dev_store = atoi(next()) != 0;
        else if (!strcmp(a, "-ngpu")) P.n_devices = atoi(next());
        else if (!strcmp(a, "-cpuonly")) P.cpu_only = true;
        else if (!strcmp(a, "-gpuonly")) P.gpu_only = true;
        else if (!strcmp(a, "-log")) P.log = next();
        else if (!strcmp(a, "-lv")) P.lv = atoi(next());
        else if (!strcmp(a, "-ci")) P.cmin = (unsigned)strtoul(next(), nullptr, 10);
        else if (!strcmp(a, "-cx")) P.cmax = (unsigned)strtoul(next(), nullptr, 10);
        else if (!strcmp(a, "-o")) P.out = next();
        else if (!strcmp(a, "-read")) {
            int j;
            for (j = i + 1; j < argc; j++, i++) {
                if (argv[j][0] == '-') { i = j - 1; break; }
                P.reads.push_back(argv[j]);
            }
        } else {
            fprintf(stderr, "Wrong argument: %s\n", a);
            usage(argv[0]);
            exit(1);
        }
    }
    if (P.reads.empty()) { usage(argv[0]); exit(1); }
    if (P.tmp.back() != '/') P.tmp.push_back('/');
    if (P.cpu_only) { fprintf(stderr, "-cpuonly: this binary counts on the GPU only\n"); exit(1); }
    if (P.hpc) { fprintf(stderr, "-hpc 1: homopolymer compression is not supported\n"); exit(1); }
    if (P.K > 63 || P.K < 8) { fprintf(stderr, "k must be in [8, 63]\n"); exit(1); }
    if (P.n_devices < 1) { fprintf(stderr, "-ngpu must be >= 1\n"); exit(1); }
    return P;
}

double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

void logf(const Params& P, int lv, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
void logf(const Params& P, int lv, const char* fmt, ...) {
    if (lv > P.lv) return;
    static const char* lab[6] = {"[FATAL]", "[ERROR]", "[NOTICE]", "[WARNING]", "[INFO]", "[DEBUG]"};
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    fprintf(stderr, "P0 %s\t%s\n", lab[lv], buf);
}

// Line reader over plain or gzip files.
struct LineReader {
    gzFile gz = nullptr;
    std::vector<char> buf;
    size_t pos = 0, len = 0;
    bool eof = false;
    explicit LineReader(const std::string& path) : buf(1 << 22) {
        gz = gzopen(path.c_str(), "rb");
        if (!gz) { fprintf(stderr, "cannot open %s\n", path.c_str()); exit(1); }
        gzbuffer(gz, 1 << 20);
    }
    ~LineReader() { if (gz) gzclose(gz); }
    bool line(std::string& out) {
        out.clear();
        for (;;) {
            if (pos == len) {
                if (eof) return !out.empty();
                const int n = gzread(gz, buf.data(), (unsigned)buf.size());
                if (n <= 0) { eof = true; len = pos = 0; continue; }
                len = (size_t)n; pos = 0;
            }
            const char* s = buf.data() + pos;
            const char* e = (const char*)memchr(s, '\n', len - pos);
            if (e) { out.append(s, e - s); pos += (e - s) + 1; return true; }
            out.append(s, len - pos);
            pos = len;
        }
    }
};

bool name_is_fasta(const std::string& f) {
    const bool gz = f.size() > 3 && (f.back() == 'z' || f.back() == 'Z');
    const char c = gz ? f[f.size() - 4] : f.back();
    return c == 'a' || c == 'A';
}

struct Store {  // partition stores in the reference's byte layout
    bool to_file;
    std::vector<std::vector<uint8_t>> mem;
    std::vector<FILE*> fp;
    std::vector<std::string> names;
    std::vector<uint64_t> bytes, skms, kmers;
    Store(const Params& P) : to_file(P.to_file), mem(P.parts), fp(P.parts, nullptr), names(P.parts),
                             bytes(P.parts, 0), skms(P.parts, 0), kmers(P.parts, 0) {
        if (to_file)
            for (int i = 0; i < P.parts; i++) {
                names[i] = P.tmp + std::to_string(i) + ".skm";
                fp[i] = fopen(names[i].c_str(), "wb");
                if (!fp[i]) { fprintf(stderr, "cannot create %s (use -tmp <dir> or -im)\n", names[i].c_str()); exit(1); }
            }
    }
    void add(int i, const uint8_t* d, uint64_t n, uint64_t s, uint64_t k) {
        if (to_file) { if (n && fwrite(d, 1, n, fp[i]) != n) { perror("fwrite"); exit(1); } }
        else mem[i].insert(mem[i].end(), d, d + n);
        bytes[i] += n; skms[i] += s; kmers[i] += k;
    }
    void close() { if (to_file) for (auto f : fp) fclose(f); }
    void load(int i, uint8_t* dst) {
        if (!to_file) { memcpy(dst, mem[i].data(), bytes[i]); std::vector<uint8_t>().swap(mem[i]); return; }
        FILE* f = fopen(names[i].c_str(), "rb");
        if (!f || fread(dst, 1, bytes[i], f) != bytes[i]) { fprintf(stderr, "cannot read %s\n", names[i].c_str()); exit(1); }
        fclose(f);
    }
};

// Read batches in CSR form (bases, offsets) handed from the loader thread to the device loop;
// line_bytes counts the input lines consumed up to the end of the batch (for size estimates).
struct Batch {
    std::vector<char> chars;
    std::vector<uint64_t> offs{0};
    uint64_t line_bytes = 0;
};

// Bounded queue of batches: the loader blocks while `depth` batches wait, the consumer until one
// arrives or the loader has finished.
class BatchQueue {
  public:
    explicit BatchQueue(size_t depth) : depth_(depth) {}
    void push(Batch&& b) {
        std::unique_lock<std::mutex> l(m_);
        cv_.wait(l, [&] { return q_.size() < depth_; });
        q_.push_back(std::move(b));
        cv_.notify_all();
    }
    void close() {
        std::lock_guard<std::mutex> l(m_);
        closed_ = true;
        cv_.notify_all();
    }
    bool pop(Batch& b) {
        std::unique_lock<std::mutex> l(m_);
        cv_.wait(l, [&] { return !q_.empty() || closed_; });
        if (q_.empty()) return false;
        b = std::move(q_.front());
        q_.pop_front();
        cv_.notify_all();
        return true;
    }

  private:
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<Batch> q_;
    size_t depth_;
    bool closed_ = false;
};

// Loader thread body: the files' sequence lines in batches of at most gpu_batch_bases bases (a
// longer read is a batch of its own), pushed to q; the queue is closed at the end.
void load_batches(const Params& P, BatchQueue& q) {
    const bool fasta = name_is_fasta(P.reads[0]);
    Batch b;
    std::string line;
    uint64_t consumed = 0;
    for (const std::string& f : P.reads) {
        LineReader lr(f);
        uint64_t li = 0;
        const uint64_t mask = fasta ? 1 : 3;
        while (lr.line(line)) {
            consumed += line.size() + 1;
            if ((li++ & mask) != 1) continue;
            if ((long long)line.size() < P.K) continue;
            if (b.offs.size() > 1 && (long long)(b.chars.size() + line.size()) > P.gpu_batch_bases) {
                q.push(std::move(b));
                b = Batch();
            }
            b.chars.insert(b.chars.end(), line.begin(), line.end());
            b.offs.push_back(b.chars.size());
            b.line_bytes = consumed;
        }
    }
    if (b.offs.size() > 1) q.push(std::move(b));
    q.close();
}

// Input size for the capacity estimate: plain files by size, gzip files by their trailer's
// uncompressed size (mod 2^32, so large or multi-member files are underestimated and fall back).
uint64_t input_bytes(const Params& P) {
    uint64_t t = 0;
    for (const std::string& f : P.reads) {
        FILE* fp = fopen(f.c_str(), "rb");
        if (!fp) continue;
        unsigned char m[2] = {0, 0};
        const bool gz = fread(m, 1, 2, fp) == 2 && m[0] == 0x1f && m[1] == 0x8b;
        fseek(fp, 0, SEEK_END);
        const long sz = ftell(fp);
        if (gz && sz >= 4) {
            unsigned char tr[4];
            fseek(fp, sz - 4, SEEK_SET);
            if (fread(tr, 1, 4, fp) == 4) t += (uint64_t)tr[0] | (uint64_t)tr[1] << 8 | (uint64_t)tr[2] << 16 | (uint64_t)tr[3] << 24;
        } else if (sz > 0) {
            t += (uint64_t)sz;
        }
        fclose(fp);
    }
    return t;
}

void check(rgkc_ctx* c, int rc, const char* what) {
    if (rc) { fprintf(stderr, "rgkc error calling \"%s\", code is %d: %s\n", what, rc, rgkc_last_error(c)); exit(1); }
}

}  // namespace

// The autotuner's cached choice for this (k, p, parts), the entry whose batch size is nearest this
// run's GPU batches (rapidgkc_amd/autotune.py writes autotune_cache.txt next to the library, one
// "<shape> inst_cap=.. count_wg_per_cu=.. p1_lists=.." line per shape).
void apply_tuning(rgkc_ctx* c, const Params& P) {
    char exe[4096];
    const ssize_t n = readlink("/proc/self/exe", exe, sizeof exe - 1);
    if (n <= 0) return;
    exe[n] = 0;
    std::string dir(exe);
    dir = dir.substr(0, dir.rfind('/'));
    FILE* f = fopen((dir + "/../autotune_cache.txt").c_str(), "r");
    if (!f) return;
    char prefix[128];
    snprintf(prefix, sizeof prefix, "gfx950:k%d:p%d:parts%d:n", P.K, P.P, P.parts);
    const int want = (int)std::log2((double)std::max<long long>(P.gpu_batch_bases, 1));
    char line[512];
    int best_d = 1 << 30, cap = 0, wg = 0, lists = 1;
    while (fgets(line, sizeof line, f)) {
        if (strncmp(line, prefix, strlen(prefix)) != 0) continue;
        int nb = 0, a = 0, b = 0, l = 1;
        if (sscanf(line + strlen(prefix), "%d", &nb) != 1) continue;
        const char* q = strstr(line, "inst_cap=");
        if (q) a = atoi(q + 9);
        q = strstr(line, "count_wg_per_cu=");
        if (q) b = atoi(q + 16);
        q = strstr(line, "p1_lists=");
        if (q) l = atoi(q + 9);
        if (std::abs(nb - want) < best_d) { best_d = std::abs(nb - want); cap = a; wg = b; lists = l; }
    }
    fclose(f);
    if (best_d == 1 << 30) return;
    rgkc_set_tuning(c, (uint32_t)cap);
    rgkc_set_count_workgroups(c, (uint32_t)wg);
    rgkc_set_p1_lists(c, lists ? 0u : 0xFFFFFFFFu);
    fprintf(stderr, "autotuned: inst_cap %d, count workgroups per CU %d, phase-1 lists %s\n", cap, wg, lists ? "on" : "off");
}

// Phase 1 with host stores: each batch's store is copied back and appended to the partition
// stores (memory or <tmp>/<id>.skm) while the loader thread reads the next batch.
void phase1_host(rgkc_ctx* c, const Params& P, Store& st, uint64_t& n_reads) {
    BatchQueue bq(2);
    std::thread loader(load_batches, std::cref(P), std::ref(bq));
    Batch b;
    std::vector<uint8_t> sbuf;
    while (bq.pop(b)) {
        const int64_t nr = (int64_t)b.offs.size() - 1;
        check(c, rgkc_set_reads(c, b.chars.data(), b.offs.data(), nr), "rgkc_set_reads");
        std::vector<uint64_t> po(P.parts + 1), sc(P.parts), kc(P.parts);
        check(c, rgkc_superkmers(c, P.K, P.P, P.parts, po.data(), sc.data(), kc.data()), "rgkc_superkmers");
        sbuf.resize(po[P.parts] + 1);
        check(c, rgkc_copy_store(c, sbuf.data()), "rgkc_copy_store");
        for (int i = 0; i < P.parts; i++) st.add(i, sbuf.data() + po[i], po[i + 1] - po[i], sc[i], kc[i]);
        n_reads += (uint64_t)nr;
        logf(P, 5, "-- BATCH  GPU: #reads: %lld --", (long long)nr);
    }
    loader.join();
}

// Phase 1 into a device-resident store: every batch is appended by the batched path (one pinned
// copy and one graph replay per batch, rgkc_stream_add) into partition regions sized from the
// first batch's exact phase 1, scaled to the input size, with a 30% margin. Returns false (the
// caller then uses host stores) when the regions would take more than 40% of the device memory,
// a batch exceeds the stream's batch capacity, or a region overflows.
bool phase1_device(rgkc_ctx* c, const Params& P, size_t vram, uint64_t& n_reads, std::vector<uint64_t>& bytes,
                   std::vector<uint64_t>& skms, std::vector<uint64_t>& kmers) {
    BatchQueue bq(2);
    std::thread loader(load_batches, std::cref(P), std::ref(bq));
    Batch b;
    bool good = bq.pop(b), begun = false;
    if (good) {
        int64_t nr = (int64_t)b.offs.size() - 1;
        check(c, rgkc_set_reads(c, b.chars.data(), b.offs.data(), nr), "rgkc_set_reads");
        std::vector<uint64_t> po(P.parts + 1), sc(P.parts), kc(P.parts);
        check(c, rgkc_superkmers(c, P.K, P.P, P.parts, po.data(), sc.data(), kc.data()), "rgkc_superkmers");
        const double scale = std::max(1.0, (double)input_bytes(P) / (double)std::max<uint64_t>(b.line_bytes, 1));
        std::vector<uint64_t> cap_b(P.parts), cap_r(P.parts);
        double need = 0;
        for (int i = 0; i < P.parts; i++) {
            cap_b[i] = (uint64_t)((double)(po[i + 1] - po[i]) * scale * 1.3) + 65536;
            cap_r[i] = (uint64_t)((double)sc[i] * scale * 1.3) + 1024;
            need += (double)cap_b[i] + 4.0 * (double)cap_r[i];
        }
        const uint64_t max_bases = (uint64_t)std::max<long long>(P.gpu_batch_bases, (long long)b.chars.size());
        good = need <= 0.4 * (double)(vram ? vram : (1ull << 34));
        if (good) good = rgkc_stream_begin(c, P.K, P.P, P.parts, max_bases / (uint64_t)P.K + 1, max_bases,
                                           cap_b.data(), cap_r.data(), 1) == 0;
        begun = good;
        while (good) {
            if (b.chars.size() > max_bases) { good = false; break; }
            if (rgkc_stream_add(c, b.chars.data(), b.offs.data(), nr) != 0) { good = false; break; }
            n_reads += (uint64_t)nr;
            logf(P, 5, "-- BATCH  GPU: #reads: %lld (device store) --", (long long)nr);
            if (!bq.pop(b)) break;
            nr = (int64_t)b.offs.size() - 1;
        }
    }
    while (bq.pop(b)) {}  // let the loader finish
    loader.join();
    if (begun) {
        const int rc = rgkc_stream_end(c, bytes.data(), skms.data(), kmers.data());
        if (good && rc == 0) return true;
        if (good && rc != RGKC_E_OVERFLOW) check(c, rc, "rgkc_stream_end");
    }
    if (n_reads || begun) logf(P, 3, "device store does not fit this input; using host stores");
    return false;
}

int main(int argc, char** argv) {
    const double t_all = now();
    Params P = parse(argc, argv);
    fprintf(stderr, "================ PROGRAM BEGINS ================\n");
    rgkc_ctx* c = nullptr;
    if (rgkc_create(0, &c)) { fprintf(stderr, "no HIP device\n"); return 1; }
    apply_tuning(c, P);
    const size_t vram = rgkc_device_reset(0);
    fprintf(stderr, "GPU 0 VRAM %zu\n", vram >> 20);
    logf(P, 2, "**** Phase 1: Loading and generate superkmers ****");
    const double t1 = now();
    Store st(P);
    uint64_t n_reads = 0;
    // -im keeps the store on the device (stream path) when it fits, else host stores
    std::vector<uint64_t> d_bytes(P.parts, 0), d_skms(P.parts, 0), d_kmers(P.parts, 0);
    const bool on_device = !P.to_file && P.dev_store && phase1_device(c, P, vram, n_reads, d_bytes, d_skms, d_kmers);
    if (!on_device) {
        n_reads = 0;
        phase1_host(c, P, st, n_reads);
    }
    st.close();
    if (on_device) {
        st.bytes = d_bytes;
        st.skms = d_skms;
        st.kmers = d_kmers;
    }
    uint64_t tot_skm = 0, tot_bytes = 0, tot_kmer = 0;
    for (int i = 0; i < P.parts; i++) { tot_skm += st.skms[i]; tot_bytes += st.bytes[i]; tot_kmer += st.kmers[i]; }
    const double p1 = now() - t1;
    logf(P, 2, "**** All reads loaded and SKMs generated (Phase 1 ends) ****");
    logf(P, 2, "     Phase 1 Time: %f sec", p1);
    logf(P, 4, "SKM TOT CNT = %llu BYTES = %llu", (unsigned long long)tot_skm, (unsigned long long)tot_bytes);
    logf(P, 4, "KMER TOT CNT = %llu", (unsigned long long)tot_kmer);
    fprintf(stderr, "Total reads loaded: %llu\n", (unsigned long long)n_reads);

    logf(P, 2, "**** Phase 2: Superkmer extraction and kmer counting ****");
    const double t2 = now();
    // Group partitions so one group's store, descriptors and table fit in a quarter of free VRAM.
    const int W = P.K <= 32 ? 8 : 16;
    const double budget = (double)(vram ? vram : (1ull << 34)) * 0.25;
    FILE* out = nullptr;
    std::vector<uint64_t> part_rows(P.parts, 0);
    std::vector<std::vector<uint8_t>> tkeys(P.parts);
    std::vector<std::vector<uint32_t>> tcnts(P.parts);
    uint64_t distinct_tot = 0, rows_tot = 0;
    std::vector<std::pair<int, int>> groups;  // partition ranges [i, j)
    if (on_device) {
        // the device store: one count call (the library streams partition groups itself)
        check(c, rgkc_count(c, P.cmin, P.cmax, part_rows.data(), &distinct_tot), "rgkc_count");
        for (auto r : part_rows) rows_tot += r;
        if (!P.out.empty()) {
            std::vector<uint8_t> keys(rows_tot * W + 16);
            std::vector<uint32_t> cnts(rows_tot + 1);
            check(c, rgkc_copy_table(c, keys.data(), cnts.data()), "rgkc_copy_table");
            uint64_t o = 0;
            for (int q = 0; q < P.parts; q++) {
                tkeys[q].assign(keys.begin() + o * W, keys.begin() + (o + part_rows[q]) * W);
                tcnts[q].assign(cnts.begin() + o, cnts.begin() + o + part_rows[q]);
                o += part_rows[q];
            }
        }
        logf(P, 5, "GPU 0: Part 0..%d %llu distinct (device store)", P.parts - 1, (unsigned long long)distinct_tot);
    }
    for (int i = 0; i < P.parts && !on_device;) {
        int j = i;
        double need = 0;
        while (j < P.parts) {
            const double n = st.bytes[j] * 1.0 + st.skms[j] * 12.0 + st.kmers[j] * (W + 4.0);
            if (j > i && need + n > budget) break;
            need += n;
            j++;
        }
        groups.emplace_back(i, j);
        i = j;
    }
    // a group's stores, read into one buffer (partition offsets po)
    struct Loaded {
        std::vector<uint64_t> po;
        std::vector<uint8_t> buf;
    };
    auto load_group = [&st](int i, int j) {
        Loaded L;
        L.po.assign(j - i + 1, 0);
        for (int q = i; q < j; q++) L.po[q - i + 1] = L.po[q - i] + st.bytes[q];
        L.buf.resize(L.po.back() + 1);
        for (int q = i; q < j; q++) st.load(q, L.buf.data() + L.po[q - i]);
        return L;
    };
    std::future<Loaded> next;
    if (!groups.empty()) next = std::async(std::launch::async, load_group, groups[0].first, groups[0].second);
    for (size_t gi = 0; gi < groups.size(); gi++) {
        const int i = groups[gi].first, j = groups[gi].second;
        Loaded L = next.get();
        if (gi + 1 < groups.size())  // the next group's stores load while this one counts
            next = std::async(std::launch::async, load_group, groups[gi + 1].first, groups[gi + 1].second);
        const std::vector<uint64_t>& po = L.po;
        std::vector<int32_t> ids;
        for (int q = i; q < j; q++) ids.push_back(q);
        std::vector<uint64_t> rows(j - i);
        uint64_t distinct = 0;
        check(c, rgkc_count_host_store(c, L.buf.data(), po.data(), j - i, ids.data(), P.parts, P.K, P.P, P.cmin, P.cmax,
                                       rows.data(), &distinct), "rgkc_count_host_store");
        uint64_t n = 0;
        for (auto r : rows) n += r;
        std::vector<uint8_t> keys(n * W + 16);
        std::vector<uint32_t> cnts(n + 1);
        check(c, rgkc_copy_table(c, keys.data(), cnts.data()), "rgkc_copy_table");
        uint64_t o = 0;
        for (int q = i; q < j; q++) {
            part_rows[q] = rows[q - i];
            if (!P.out.empty()) {
                tkeys[q].assign(keys.begin() + o * W, keys.begin() + (o + rows[q - i]) * W);
                tcnts[q].assign(cnts.begin() + o, cnts.begin() + o + rows[q - i]);
            }
            o += rows[q - i];
        }
        distinct_tot += distinct;
        rows_tot += n;
        logf(P, 5, "GPU 0: Part %d..%d %llu distinct", i, j - 1, (unsigned long long)distinct);
    }
    if (!P.out.empty()) {
        // Table file: "RGKCTAB1", u32 k, u32 key_bytes, u32 parts, u32 min, u32 max, u32 0,
        // u64 total_rows, u64 rows[parts], keys (key_bytes each, partition-major), u32 counts.
        out = fopen(P.out.c_str(), "wb");
        if (!out) { perror(P.out.c_str()); return 1; }
        uint32_t hdr[6] = {(uint32_t)P.K, (uint32_t)W, (uint32_t)P.parts, P.cmin, P.cmax, 0};
        fwrite("RGKCTAB1", 1, 8, out);
        fwrite(hdr, 4, 6, out);
        fwrite(&rows_tot, 8, 1, out);
        fwrite(part_rows.data(), 8, P.parts, out);
        for (int q = 0; q < P.parts; q++) fwrite(tkeys[q].data(), 1, tkeys[q].size(), out);
        for (int q = 0; q < P.parts; q++) fwrite(tcnts[q].data(), 4, tcnts[q].size(), out);
        fclose(out);
    }
    const double p2 = now() - t2;
    logf(P, 2, "Total number of distinct kmers: %llu", (unsigned long long)distinct_tot);
    logf(P, 2, "Rows in [%u, %u]: %llu", P.cmin, P.cmax, (unsigned long long)rows_tot);
    logf(P, 2, "**** Kmer counting finished (Phase 2 ends) ****");
    logf(P, 2, "     Phase 2 Time: %f sec (P1: %f s)", p2, p1);
    if (P.to_file) for (int q = 0; q < P.parts; q++) remove(st.names[q].c_str());
    rgkc_destroy(c);
    fprintf(stderr, "================ PROGRAM ENDS ================\n");
    printf("%f\n", now() - t_all);
    return 0;
}

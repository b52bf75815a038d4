// =====================================================================================
//  ORACLE -- TEST INFRASTRUCTURE ONLY.
//
//  CPU restatement of RapidGKC's reference CPU path for k-mer counting:
//    reads -> 3-bases-per-byte packing -> canonical-minimizer superkmers -> partition
//    store -> k-mer extraction -> canonical keys -> sort -> run-length (k-mer, count) table.
//
//  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library,
//  and only as the checker. The product (rapidgkc_amd) never links, loads or calls it.
//
//  Every function cites the reference file:line whose semantics it restates. The code is
//  written from those semantics, not copied. Pinning (what proves each function right)
//  is listed per function; tests/test_oracle.py holds the checks:
//    - closed forms / truth tables (filter, packing, partition, byte count, canonical),
//    - the reference's own sources built into oracle/_ref (when /root/reference exists),
//    - a brute-force canonical k-mer counter over the raw reads (independent of
//      minimizers and partitions), plus conservation invariants.
// =====================================================================================
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

typedef unsigned __int128 u128;

namespace {

// a1 -- base codes. cpu_funcs.cpp:20-53 (basemap) and :55-88 (basemap_compl):
// A/a -> 0, C/c -> 1, G/g -> 2, T/t -> 3, N/n -> 0; complement A->3 C->2 G->1 T->0 N->3.
// Bytes outside {ACGTN, acgtn} map to 255 in the reference (undefined downstream); this
// restatement treats them like N, the same rule the HIP path uses.
struct BaseTables {
    uint8_t fwd[256];
    BaseTables() {
        for (int i = 0; i < 256; i++) fwd[i] = 0;
        fwd['C'] = fwd['c'] = 1;
        fwd['G'] = fwd['g'] = 2;
        fwd['T'] = fwd['t'] = 3;
    }
};
const BaseTables g_bt;
inline uint32_t code(uint8_t c) { return g_bt.fwd[c]; }

// a4 -- minimizer filter, minimizer_filter.h:19-20 (the default "rapidgkc" filter):
// a p-mer is usable iff its first three bases are not [AC][AC]A and its last three are not AAA.
inline bool pmer_valid(uint32_t pmer, int p) {
    return (((pmer >> ((p - 3) * 2)) & 0x2Bu) != 0) && ((pmer & 0x3Fu) != 0);
}

// Filtered value of one p-mer strand: the p-mer itself if usable, else the all-ones mask
// (cpu_funcs.cpp:109-113, gpu_skmgen.cu:233-245).
inline uint32_t filtered(uint32_t pmer, int p, uint32_t mask) { return pmer_valid(pmer, p) ? pmer : mask; }

template <typename K> inline K kmer_mask(int k) {
    return (k * 2 >= (int)(sizeof(K) * 8)) ? (K)~(K)0 : (((K)1 << (2 * k)) - 1);
}

// a13 -- reverse complement of a k-mer: the k 2-bit digits of ~x in reverse order
// (cpu_funcs.cpp:427-436, gpu_kmercounting.cu:63-74).
template <typename K> inline K revcomp(K x, int k) {
    K r = 0;
    x = ~x;
    for (int i = 0; i < k; i++) { r = (r << 2) | (x & 3); x >>= 2; }
    return r;
}

// a14 -- ascending sort of keys. The reference sorts with an MSD radix (cpu_funcs.cpp:365-392)
// or RADULS (:394-422); a sorted multiset is unique, so any correct sort gives the same array.
void sort_u64(uint64_t* a, int64_t n) {
    if (n < 2) return;
    if (n < 65536) { std::sort(a, a + n); return; }
    std::vector<uint64_t> tmp(n);
    uint64_t* src = a; uint64_t* dst = tmp.data();
    for (int pass = 0; pass < 4; pass++) {
        const int sh = pass * 16;
        std::vector<int64_t> cnt(65537, 0);
        for (int64_t i = 0; i < n; i++) cnt[((src[i] >> sh) & 0xFFFF) + 1]++;
        for (int b = 0; b < 65536; b++) cnt[b + 1] += cnt[b];
        for (int64_t i = 0; i < n; i++) dst[cnt[(src[i] >> sh) & 0xFFFF]++] = src[i];
        std::swap(src, dst);
    }
    // 4 passes: the sorted data is back in `a`.
}
void sort_u128(u128* a, int64_t n) { std::sort(a, a + n); }

// a5 -- canonical minimizer of every k-mer of one read (closed form of cpu_funcs.cpp:124-191 and
// gpu_skmgen.cu:218-274): for k-mer i, the minimum over the w = k-p+1 p-mers starting at i..i+k-p
// of min(filtered(forward p-mer), filtered(reverse-complement p-mer)); the rc p-mer puts the
// complement of its last base in the most significant digit (cpu_funcs.cpp:133,139).
// The reference rolls the window and rescans when the minimum leaves it; the value per k-mer is
// the same as this direct window minimum.
int kmer_minimizers(const uint8_t* read, int len, int k, int p, uint32_t* mm) {
    if (len < k) return 0;
    const uint32_t mask = (p == 16) ? 0xFFFFFFFFu : ((1u << (2 * p)) - 1);
    const int npm = len - p + 1;
    std::vector<uint32_t> v(npm);
    uint32_t f = 0, r = 0;
    for (int j = 0; j < len; j++) {
        const uint32_t c = code(read[j]);
        f = ((f << 2) | c) & mask;
        r = (r >> 2) | ((3u - c) << (2 * p - 2));
        if (j >= p - 1) {
            const uint32_t a = filtered(f, p, mask), b = filtered(r, p, mask);
            v[j - p + 1] = a < b ? a : b;
        }
    }
    const int nk = len - k + 1, w = k - p + 1;
    for (int i = 0; i < nk; i++) {
        uint32_t m = mask;
        for (int j = 0; j < w; j++) m = v[i + j] < m ? v[i + j] : m;
        mm[i] = m;
    }
    return nk;
}

// a6 -- superkmer cuts (cpu_funcs.cpp:147-198; gpu_skmgen.cu:324-330): a new superkmer starts at
// k-mer i iff mm[i] != mm[i-1]. offs[t] is the first k-mer of superkmer t, offs[n] = len-k+1.
int skm_offsets(const uint32_t* mm, int nk, int32_t* offs, uint32_t* skm_mm) {
    int n = 0;
    offs[0] = 0;
    for (int i = 1; i < nk; i++) {
        if (mm[i] != mm[i - 1]) { skm_mm[n] = mm[i - 1]; n++; offs[n] = i; }
    }
    skm_mm[n] = mm[nk - 1];
    n++;
    offs[n] = nk;
    return n;
}

// a3 -- packing (cpu_funcs.cpp:200-211): byte i = 11|b[3i]|b[3i+1]|b[3i+2] (2 bits each, first
// base most significant); the tail byte holds (len%3)<<6 and its bases in the high bits, or is 0
// when len%3 == 0. Output has len/3+1 bytes.
int pack_read(const uint8_t* read, int len, uint8_t* out) {
    int i = 0;
    for (; i + 3 <= len; i += 3)
        out[i / 3] = (uint8_t)(0xC0 | (code(read[i]) << 4) | (code(read[i + 1]) << 2) | code(read[i + 2]));
    const int rem = len - i;
    uint8_t t = (uint8_t)(rem << 6);
    for (int j = 0; j < rem; j++) t |= (uint8_t)(code(read[i + j]) << (4 - 2 * j));
    out[i / 3] = t;
    return len / 3 + 1;
}

// a8 -- bytes of one superkmer record (cpu_funcs.cpp:216-218, gpu_skmgen.cu:279-281).
inline int skm_bytes(int beg, int end, int k) { return ((beg % 3) + end + (k - 1) - beg + 3) / 3; }

// a7 -- partition of a superkmer (cpu_funcs.cpp:213-215, gpu_skmgen.cu:276-278): (~mm) % parts,
// computed on the 32-bit minimizer.
inline uint32_t partition_of(uint32_t mm, uint32_t parts) { return (uint32_t)(~mm) % parts; }

// a9 -- one superkmer record (cpu_funcs.cpp:238-244, gpu_skmgen.cu:454-465): the packed bytes
// [beg/3, beg/3+nbytes) of the read, with the first byte's top two bits set to 3-beg%3 (valid
// bases sit in the low bits) and the last byte's top two bits set to (end-1+k)%3 (valid bases
// sit in the high bits; 0 marks an empty terminator). Other bits keep the neighbouring bases.
void encode_skm(const uint8_t* packed, int beg, int end, int k, uint8_t* out) {
    const int nb = skm_bytes(beg, end, k);
    memcpy(out, packed + beg / 3, nb);
    out[0] = (uint8_t)((out[0] & 0x3F) | ((3 - beg % 3) << 6));
    out[nb - 1] = (uint8_t)((out[nb - 1] & 0x3F) | (((end - 1 + k) % 3) << 6));
}

// a12 -- decode a partition byte stream into canonical k-mers (cpu_funcs.cpp:437-515). A record
// is: a first byte holding its low `ind` bases, full bytes (indicator 3) holding 3 bases each,
// then a last byte (indicator 0..2) holding its high `ind` bases. Records are back to back
// (the reference's boundary rule "full byte then non-full byte", cpu_funcs.cpp:516-525).
// Emits min(kmer, revcomp(kmer)) for each k-mer in record order; returns the number emitted,
// or -1 for a malformed stream.
template <typename K> int64_t extract_kmers(const uint8_t* s, uint64_t nbytes, int k, K* out) {
    const K km = kmer_mask<K>(k);
    const int top = 2 * (k - 1);
    uint64_t i = 0;
    int64_t n = 0;
    while (i < nbytes) {
        K f = 0, r = 0;
        int have = 0;
        // first byte
        int ind = s[i] >> 6;
        if (ind == 0) return -1;
        auto push = [&](uint32_t c) {
            f = ((f << 2) | c) & km;
            r = (r >> 2) | ((K)(3u - c) << top);
            have++;
            if (have >= k) out[n++] = f < r ? f : r;
        };
        for (int j = ind - 1; j >= 0; j--) push((s[i] >> (2 * j)) & 3);
        i++;
        for (;;) {
            if (i >= nbytes) return -1;
            ind = s[i] >> 6;
            for (int j = 0; j < ind; j++) push((s[i] >> (4 - 2 * j)) & 3);
            i++;
            if (ind != 3) break;
        }
        if (have < k) return -1;
    }
    return n;
}

// a15 -- run-length count of a sorted key array (cpu_funcs.cpp:614-640), counts saturated at
// 2^32-1 (gpu_kmercounting.cu:502), rows kept iff min <= count <= max (utilities.hpp:190; the
// reference computes the table but never applies the filter or emits it).
template <typename K>
int64_t rle(const K* a, int64_t n, uint32_t cmin, uint32_t cmax, K* ok, uint32_t* oc, int64_t* distinct) {
    int64_t rows = 0, d = 0;
    int64_t i = 0;
    while (i < n) {
        int64_t j = i + 1;
        while (j < n && a[j] == a[i]) j++;
        const uint64_t run = (uint64_t)(j - i);
        const uint32_t c = run > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)run;
        d++;
        if (c >= cmin && c <= cmax) { if (ok) ok[rows] = a[i]; if (oc) oc[rows] = c; rows++; }
        i = j;
    }
    if (distinct) *distinct = d;
    return rows;
}

template <typename K> void sort_keys(K* a, int64_t n);
template <> void sort_keys<uint64_t>(uint64_t* a, int64_t n) { sort_u64(a, n); }
template <> void sort_keys<u128>(u128* a, int64_t n) { sort_u128(a, n); }

// a17 -- one partition: extract, sort, run-length (cpu_funcs.cpp:586-651).
template <typename K>
int64_t count_partition(const uint8_t* skms, uint64_t nbytes, int k, uint32_t cmin, uint32_t cmax,
                        K* ok, uint32_t* oc, int64_t* distinct, int64_t* nkmers) {
    // a record of L bases carries L-k+1 k-mers in at least L/3 bytes, so 3*nbytes bounds the count
    std::vector<K> keys(nbytes * 3 + 1);
    const int64_t n = extract_kmers<K>(skms, nbytes, k, keys.data());
    if (n < 0) return -1;
    if (nkmers) *nkmers = n;
    sort_keys<K>(keys.data(), n);
    return rle<K>(keys.data(), n, cmin, cmax, ok, oc, distinct);
}

// Superkmers of one read into per-partition cursors (cpu_funcs.cpp:219-250, 252-298).
// Mode 0 sizes (bytes/skms/kmers per partition); mode 1 writes records at cursor[part].
int reads_to_store(const char* reads, const uint64_t* offs, int64_t n_reads, int k, int p, int parts,
                   int mode, uint64_t* part_bytes, uint64_t* part_skms, uint64_t* part_kmers,
                   uint64_t* cursor, uint8_t* store) {
    std::vector<uint32_t> mm;
    std::vector<int32_t> so;
    std::vector<uint32_t> smm;
    std::vector<uint8_t> packed;
    for (int64_t r = 0; r < n_reads; r++) {
        const uint8_t* rd = (const uint8_t*)reads + offs[r];
        const int len = (int)(offs[r + 1] - offs[r]);
        if (len < k) continue;  // the loader drops reads shorter than k (fileloader.hpp:304,430)
        mm.resize(len); so.resize(len + 2); smm.resize(len + 1); packed.resize(len / 3 + 1);
        const int nk = kmer_minimizers(rd, len, k, p, mm.data());
        const int ns = skm_offsets(mm.data(), nk, so.data(), smm.data());
        pack_read(rd, len, packed.data());
        for (int t = 0; t < ns; t++) {
            const uint32_t part = partition_of(smm[t], (uint32_t)parts);
            const int nb = skm_bytes(so[t], so[t + 1], k);
            if (nb > 65535) return -2;  // T_skm_len is 16 bits (types.h:19)
            if (mode == 0) {
                part_bytes[part] += nb; part_skms[part] += 1; part_kmers[part] += so[t + 1] - so[t];
            } else {
                encode_skm(packed.data(), so[t], so[t + 1], k, store + cursor[part]);
                cursor[part] += nb;
            }
        }
    }
    return 0;
}

template <typename K>
int64_t bruteforce(const char* reads, const uint64_t* offs, int64_t n_reads, int k, uint32_t cmin,
                   uint32_t cmax, K* ok, uint32_t* oc, int64_t* distinct) {
    int64_t total = 0;
    for (int64_t r = 0; r < n_reads; r++) {
        const int64_t len = (int64_t)(offs[r + 1] - offs[r]);
        if (len >= k) total += len - k + 1;
    }
    std::vector<K> keys(total > 0 ? total : 1);
    const K km = kmer_mask<K>(k);
    int64_t n = 0;
    for (int64_t r = 0; r < n_reads; r++) {
        const uint8_t* rd = (const uint8_t*)reads + offs[r];
        const int64_t len = (int64_t)(offs[r + 1] - offs[r]);
        if (len < k) continue;
        for (int64_t i = 0; i + k <= len; i++) {
            K f = 0;
            for (int j = 0; j < k; j++) f = (f << 2) | code(rd[i + j]);
            f &= km;
            const K rc = revcomp<K>(f, k);
            keys[n++] = f < rc ? f : rc;
        }
    }
    sort_keys<K>(keys.data(), n);
    return rle<K>(keys.data(), n, cmin, cmax, ok, oc, distinct);
}

}  // namespace

extern "C" {

// ---- a1..a9: per-read primitives --------------------------------------------------------
int or_pmer_valid(uint32_t pmer, int p) { return pmer_valid(pmer, p) ? 1 : 0; }
int or_pack_read(const char* read, int len, uint8_t* out) { return pack_read((const uint8_t*)read, len, out); }
int or_kmer_minimizers(const char* read, int len, int k, int p, uint32_t* mm) {
    return kmer_minimizers((const uint8_t*)read, len, k, p, mm);
}
int or_skm_offsets(const char* read, int len, int k, int p, int32_t* offs, uint32_t* skm_mm) {
    if (len < k) return 0;
    std::vector<uint32_t> mm(len);
    const int nk = kmer_minimizers((const uint8_t*)read, len, k, p, mm.data());
    return skm_offsets(mm.data(), nk, offs, skm_mm);
}
uint32_t or_partition(uint32_t mm, uint32_t parts) { return partition_of(mm, parts); }
int or_skm_bytes(int beg, int end, int k) { return skm_bytes(beg, end, k); }
void or_encode_skm(const uint8_t* packed, int beg, int end, int k, uint8_t* out) { encode_skm(packed, beg, end, k, out); }
uint64_t or_revcomp64(uint64_t x, int k) { return revcomp<uint64_t>(x, k); }
void or_revcomp128(const uint64_t* x, int k, uint64_t* out) {
    u128 v = ((u128)x[1] << 64) | x[0];
    v = revcomp<u128>(v, k);
    out[0] = (uint64_t)v; out[1] = (uint64_t)(v >> 64);
}

// ---- a10: partition store of a read batch (two-pass exact sizing) ---------------------------
// Sizes: part_bytes/part_skms/part_kmers[parts] are accumulated (caller zeroes them).
int or_store_sizes(const char* reads, const uint64_t* offs, int64_t n_reads, int k, int p, int parts,
                   uint64_t* part_bytes, uint64_t* part_skms, uint64_t* part_kmers) {
    return reads_to_store(reads, offs, n_reads, k, p, parts, 0, part_bytes, part_skms, part_kmers, nullptr, nullptr);
}
// Fill: part_offs[parts+1] from the sizes; records of each partition in read order.
int or_store_fill(const char* reads, const uint64_t* offs, int64_t n_reads, int k, int p, int parts,
                  const uint64_t* part_offs, uint8_t* store) {
    std::vector<uint64_t> cur(part_offs, part_offs + parts);
    return reads_to_store(reads, offs, n_reads, k, p, parts, 1, nullptr, nullptr, nullptr, cur.data(), store);
}

// ---- a12..a17: counting ----------------------------------------------------------------------
// Keys are W-byte little-endian integers: W = 8 for k <= 32, W = 16 (lo word first) for k > 32.
int64_t or_extract_kmers(const uint8_t* skms, uint64_t nbytes, int k, void* keys) {
    return k <= 32 ? extract_kmers<uint64_t>(skms, nbytes, k, (uint64_t*)keys)
                   : extract_kmers<u128>(skms, nbytes, k, (u128*)keys);
}
void or_sort_keys(void* keys, int64_t n, int k) {
    if (k <= 32) sort_u64((uint64_t*)keys, n); else sort_u128((u128*)keys, n);
}
int64_t or_rle(const void* sorted, int64_t n, int k, uint32_t cmin, uint32_t cmax, void* ok, uint32_t* oc,
               int64_t* distinct) {
    return k <= 32 ? rle<uint64_t>((const uint64_t*)sorted, n, cmin, cmax, (uint64_t*)ok, oc, distinct)
                   : rle<u128>((const u128*)sorted, n, cmin, cmax, (u128*)ok, oc, distinct);
}
int64_t or_count_partition(const uint8_t* skms, uint64_t nbytes, int k, uint32_t cmin, uint32_t cmax, void* ok,
                           uint32_t* oc, int64_t* distinct, int64_t* nkmers) {
    return k <= 32 ? count_partition<uint64_t>(skms, nbytes, k, cmin, cmax, (uint64_t*)ok, oc, distinct, nkmers)
                   : count_partition<u128>(skms, nbytes, k, cmin, cmax, (u128*)ok, oc, distinct, nkmers);
}
// Brute force: every canonical k-mer of every read of length >= k, sorted and run-length
// counted. Independent of minimizers and partitions; the union of the partition tables must
// equal it.
int64_t or_bruteforce_count(const char* reads, const uint64_t* offs, int64_t n_reads, int k, uint32_t cmin,
                            uint32_t cmax, void* ok, uint32_t* oc, int64_t* distinct) {
    return k <= 32 ? bruteforce<uint64_t>(reads, offs, n_reads, k, cmin, cmax, (uint64_t*)ok, oc, distinct)
                   : bruteforce<u128>(reads, offs, n_reads, k, cmin, cmax, (u128*)ok, oc, distinct);
}

}  // extern "C"

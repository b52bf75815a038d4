// Seeded synthetic genomes and reads for the k-mer counting workloads.
//
// PRNG: xoshiro256** seeded through splitmix64. Every read draws from its own stream seeded by
// (seed, read index), so output is identical for any thread count. Reads carry A/C/G/T only.
//
// Workloads (sizes are parameters; the Python side names the standard ones):
//   genome:  i.i.d. bases at a given GC fraction, then interspersed repeats (copies of earlier
//            segments with substitution divergence) and tandem repeats (a random unit of
//            unit_min..unit_max bp repeated back to back).
//   short reads: fixed length, uniform start, either strand with p = 0.5, uniform substitutions.
//   long reads: log-normal output lengths, substitution/insertion/deletion errors.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) { for (int i = 0; i < 4; i++) s[i] = splitmix64(seed); }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    inline uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    inline double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    inline uint64_t below(uint64_t n) { return n ? next() % n : 0; }
    inline double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

const char kBases[4] = {'A', 'C', 'G', 'T'};

inline char draw_base(Rng& r, double gc) {
    const double u = r.uniform();
    const double at = (1.0 - gc) * 0.5, cg = gc * 0.5;
    if (u < at) return 'A';
    if (u < at + cg) return 'C';
    if (u < at + 2 * cg) return 'G';
    return 'T';
}
inline char other_base(Rng& r, char b) {
    int c = b == 'A' ? 0 : b == 'C' ? 1 : b == 'G' ? 2 : 3;
    int o = (c + 1 + (int)r.below(3)) & 3;
    return kBases[o];
}
inline char comp(char b) { return b == 'A' ? 'T' : b == 'C' ? 'G' : b == 'G' ? 'C' : 'A'; }

uint64_t read_seed(uint64_t seed, int64_t i) {
    uint64_t x = seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(i + 1));
    splitmix64(x);
    return x;
}

template <class F> void parallel_for(int64_t n, int threads, F f) {
    if (threads <= 1 || n < 1024) { f(0, n); return; }
    std::vector<std::thread> th;
    const int64_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; t++) {
        const int64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([=] { f(b, e); });
    }
    for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

// Fills out[0..len) with a genome. Returns len.
int64_t sg_genome(uint64_t seed, int64_t len, double gc, double inter_frac, int rep_min, int rep_max,
                  double rep_div, double tandem_frac, int unit_min, int unit_max, char* out) {
    const int threads = 8;
    const int64_t nblk = (len + 65535) / 65536;  // one PRNG stream per 64 kbp block
    parallel_for(nblk, threads, [&](int64_t b, int64_t e) {
        for (int64_t blk = b; blk < e; blk++) {
            Rng r(read_seed(seed, -1 - blk));
            const int64_t s1 = std::min(len, (blk + 1) * 65536);
            for (int64_t i = blk * 65536; i < s1; i++) out[i] = draw_base(r, gc);
        }
    });
    Rng r(seed * 0x9E3779B97F4A7C15ull + 12345);
    // interspersed repeats: copies of earlier segments, substitution divergence rep_div
    if (inter_frac > 0 && len > 4 * (int64_t)rep_max) {
        int64_t covered = 0;
        const int64_t target = (int64_t)(inter_frac * (double)len);
        while (covered < target) {
            const int64_t L = rep_min + (int64_t)r.below((uint64_t)(rep_max - rep_min + 1));
            const int64_t dst = L + (int64_t)r.below((uint64_t)(len - 2 * L));
            const int64_t src = (int64_t)r.below((uint64_t)(dst - L + 1));
            for (int64_t i = 0; i < L; i++) {
                char b = out[src + i];
                if (r.uniform() < rep_div) b = other_base(r, b);
                out[dst + i] = b;
            }
            covered += L;
        }
    }
    // tandem repeats: a random unit repeated back to back over a region
    if (tandem_frac > 0 && len > 100000) {
        int64_t covered = 0;
        const int64_t target = (int64_t)(tandem_frac * (double)len);
        std::vector<char> unit;
        while (covered < target) {
            const int u = unit_min + (int)r.below((uint64_t)(unit_max - unit_min + 1));
            const int64_t copies = 5 + (int64_t)r.below(46);
            int64_t L = std::max<int64_t>(u * copies, 100);
            if (L > 20000) L = 20000;
            unit.resize(u);
            for (int i = 0; i < u; i++) unit[i] = draw_base(r, gc);
            const int64_t dst = (int64_t)r.below((uint64_t)(len - L));
            for (int64_t i = 0; i < L; i++) out[dst + i] = unit[i % u];
            covered += L;
        }
    }
    return len;
}

// Fixed-length reads. chars: n_reads*read_len bytes; offs: n_reads+1 (CSR).
int sg_short_reads(const char* genome, int64_t glen, uint64_t seed, int64_t n_reads, int read_len,
                   double sub_rate, char* chars, uint64_t* offs, int threads) {
    if (glen < read_len) return -1;
    for (int64_t i = 0; i <= n_reads; i++) offs[i] = (uint64_t)i * read_len;
    parallel_for(n_reads, threads, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; i++) {
            Rng r(read_seed(seed, i));
            const int64_t st = (int64_t)r.below((uint64_t)(glen - read_len + 1));
            const bool rc = (r.next() >> 63) != 0;
            char* o = chars + (uint64_t)i * read_len;
            for (int j = 0; j < read_len; j++) {
                char c = rc ? comp(genome[st + read_len - 1 - j]) : genome[st + j];
                if (sub_rate > 0 && r.uniform() < sub_rate) c = other_base(r, c);
                o[j] = c;
            }
        }
    });
    return 0;
}

// Log-normal lengths (natural-log mean mu, sd sigma), clamped to [min_len, max_len].
int64_t sg_lognormal_lengths(uint64_t seed, int64_t n, double mu, double sigma, int min_len, int max_len,
                             int32_t* lens) {
    int64_t tot = 0;
    for (int64_t i = 0; i < n; i++) {
        Rng r(read_seed(seed ^ 0xABCDEF, i));
        double l = std::exp(mu + sigma * r.normal());
        int64_t L = (int64_t)l;
        if (L < min_len) L = min_len;
        if (L > max_len) L = max_len;
        lens[i] = (int32_t)L;
        tot += L;
    }
    return tot;
}

// Reads of given output lengths with substitution / insertion / deletion errors at total rate
// err, split f_sub : f_ins : f_del. offs must hold the CSR offsets of lens (prefix sums).
int sg_long_reads(const char* genome, int64_t glen, uint64_t seed, int64_t n_reads, const int32_t* lens,
                  double err, double f_sub, double f_ins, double f_del, char* chars, const uint64_t* offs,
                  int threads) {
    const double fs = f_sub / (f_sub + f_ins + f_del), fi = f_ins / (f_sub + f_ins + f_del);
    parallel_for(n_reads, threads, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; i++) {
            Rng r(read_seed(seed, i));
            const int64_t L = lens[i];
            // template window: allow for deletions consuming extra template
            int64_t span = L + L / 8 + 64;
            if (span > glen) span = glen;
            const int64_t st = (int64_t)r.below((uint64_t)(glen - span + 1));
            const bool rc = (r.next() >> 63) != 0;
            char* o = chars + offs[i];
            int64_t t = 0, j = 0;
            auto tmpl = [&](int64_t x) -> char {
                x %= span;
                return rc ? comp(genome[st + span - 1 - x]) : genome[st + x];
            };
            while (j < L) {
                if (r.uniform() < err) {
                    const double u = r.uniform();
                    if (u < fs) { o[j++] = other_base(r, tmpl(t)); t++; }
                    else if (u < fs + fi) { o[j++] = kBases[r.below(4)]; }
                    else { t++; }
                } else {
                    o[j++] = tmpl(t); t++;
                }
            }
        }
    });
    return 0;
}

}  // extern "C"

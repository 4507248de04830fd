// datagen.cpp -- seeded synthetic inputs for the read-path workloads.
//
// Own generators only (splitmix64 seeding, xoshiro256** streams, inverse-CDF
// Zipf, Box-Muller normals), never std::*_distribution, so that the same seed
// gives the same bytes with any C++ standard library.
//  * vocabulary: n unique lowercase words, lengths uniform in [3,12];
//  * text: Zipf(s) ranks over the vocabulary (rank r -> word r-1), one space
//    between words, '\n' after every U[8,16] words, the tail padded with spaces
//    to the exact size; the generator tallies every word it emits;
//  * random bytes: 1 MiB chunks, chunk c drawn from its own stream, so any
//    byte range can be regenerated independently (and in parallel);
//  * matrix rows: row r drawn from its own stream of N(0,1) floats;
//  * file sizes: log-normal(median, sigma) clipped to [lo, hi].
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_set>
#include <vector>
#include <fcntl.h>
#include <unistd.h>

namespace {

inline uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed, uint64_t stream = 0) {
        uint64_t x = seed ^ (stream * 0xd1b54a32d192ed03ull);
        for (auto& v : s) v = splitmix64(x);
    }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }   // [0,1)
    double uniform_pos() { return (double)((next() >> 11) + 1) * (1.0 / 9007199254740992.0); }  // (0,1]
    uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
};

// Inverse-CDF Zipf sampler over ranks 1..n with weights r^-s.
struct Zipf {
    std::vector<double> cdf;
    Zipf(uint32_t n, double s) : cdf(n) {
        double acc = 0;
        for (uint32_t r = 1; r <= n; ++r) {
            acc += (s == 1.0) ? 1.0 / (double)r : std::pow((double)r, -s);
            cdf[r - 1] = acc;
        }
    }
    uint32_t sample(Xoshiro& g) const {
        double u = g.uniform() * cdf.back();
        return (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
    }
};

const Zipf& zipf_cached(uint32_t n, double s) {
    static std::mutex mu;
    static std::map<std::pair<uint32_t, double>, Zipf*> cache;
    std::lock_guard<std::mutex> g(mu);
    auto k = std::make_pair(n, s);
    auto it = cache.find(k);
    if (it != cache.end()) return *it->second;
    Zipf* z = new Zipf(n, s);
    cache[k] = z;
    return *z;
}

int write_all(int fd, const uint8_t* p, size_t n, uint64_t off) {
    size_t done = 0;
    while (done < n) {
        ssize_t w = pwrite(fd, p + done, n - done, (off_t)(off + done));
        if (w < 0) {
            if (errno == EINTR) continue;
            return -errno;
        }
        done += (size_t)w;
    }
    return 0;
}

template <class F> void parallel_for(uint64_t n, int threads, F fn) {
    std::atomic<uint64_t> next{0};
    auto work = [&]() {
        for (;;) {
            uint64_t i = next.fetch_add(1);
            if (i >= n) break;
            fn(i);
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < threads; ++t) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
}

const uint64_t kChunk = 1ull << 20;

} // namespace

extern "C" {

// n unique words, 32-byte NUL-padded records.
int dg_vocab(uint64_t seed, uint32_t n, char* out) {
    Xoshiro g(seed, 0x766f636162ull);
    std::unordered_set<std::string> seen;
    seen.reserve(n * 2);
    uint32_t i = 0;
    while (i < n) {
        uint32_t len = 3 + g.below(10);
        char w[33] = {0};
        for (uint32_t k = 0; k < len; ++k) w[k] = (char)('a' + g.below(26));
        if (!seen.insert(std::string(w, len)).second) continue;
        memset(out + (size_t)i * 32, 0, 32);
        memcpy(out + (size_t)i * 32, w, len);
        ++i;
    }
    return 0;
}

// nq distinct indices in [0, nvocab), uniformly (partial Fisher-Yates).
int dg_query(uint64_t seed, uint32_t nvocab, uint32_t nq, uint32_t* out) {
    if (nq > nvocab) return -EINVAL;
    Xoshiro g(seed, 0x7175657279ull);
    std::vector<uint32_t> p(nvocab);
    for (uint32_t i = 0; i < nvocab; ++i) p[i] = i;
    for (uint32_t i = 0; i < nq; ++i) {
        uint32_t j = i + g.below(nvocab - i);
        std::swap(p[i], p[j]);
        out[i] = p[i];
    }
    return 0;
}

// Text of exactly `size` bytes into `out`; tally[w] += occurrences of word w.
int dg_text(uint64_t seed, const char* vocab, uint32_t nvocab, double s, uint64_t size, uint8_t* out,
            uint32_t* tally) {
    const Zipf& z = zipf_cached(nvocab, s);
    Xoshiro g(seed, 0x74657874ull);
    std::vector<uint8_t> lens(nvocab);
    for (uint32_t i = 0; i < nvocab; ++i) lens[i] = (uint8_t)strnlen(vocab + (size_t)i * 32, 32);
    uint64_t pos = 0;
    uint32_t in_line = 0, line_len = 8 + g.below(9);
    for (;;) {
        uint32_t w = z.sample(g);
        uint32_t L = lens[w];
        if (pos + L + 1 > size) break;
        memcpy(out + pos, vocab + (size_t)w * 32, L);
        pos += L;
        if (++in_line == line_len) {
            out[pos++] = '\n';
            in_line = 0;
            line_len = 8 + g.below(9);
        } else {
            out[pos++] = ' ';
        }
        if (tally) tally[w]++;
    }
    memset(out + pos, ' ', size - pos);
    return 0;
}

// Many text files in parallel: file i uses seed (seed_base, i); tallies are
// per file (tally + i*nvocab) when `tally` is not NULL.
int dg_text_files(const char* names, uint32_t stride, uint32_t nfiles, const uint64_t* sizes, uint64_t seed_base,
                  const char* vocab, uint32_t nvocab, double s, uint32_t* tally, int threads) {
    std::atomic<int> err{0};
    zipf_cached(nvocab, s);
    parallel_for(nfiles, threads, [&](uint64_t i) {
        uint64_t x = seed_base + i;
        uint64_t fseed = splitmix64(x);
        std::vector<uint8_t> buf(sizes[i]);
        dg_text(fseed, vocab, nvocab, s, sizes[i], buf.data(), tally ? tally + i * (uint64_t)nvocab : nullptr);
        int fd = open(names + i * (uint64_t)stride, O_CREAT | O_TRUNC | O_WRONLY | O_CLOEXEC, 0644);
        if (fd < 0) { err = -errno; return; }
        int r = write_all(fd, buf.data(), buf.size(), 0);
        close(fd);
        if (r) err = r;
    });
    return err.load();
}

// Bytes [off, off+len) of the random stream `seed`.
int dg_random_fill(uint8_t* buf, uint64_t off, uint64_t len, uint64_t seed) {
    uint64_t end = off + len;
    for (uint64_t c = off / kChunk; c * kChunk < end; ++c) {
        Xoshiro g(seed, c + 1);
        uint64_t cs = c * kChunk;
        uint64_t lo = std::max(cs, off), hi = std::min(cs + kChunk, end);
        // skip the words before lo
        uint64_t wskip = (lo - cs) / 8;
        for (uint64_t k = 0; k < wskip; ++k) g.next();
        uint64_t p = cs + wskip * 8;
        while (p < hi) {
            uint64_t v = g.next();
            for (int b = 0; b < 8; ++b, ++p)
                if (p >= lo && p < hi) buf[p - off] = (uint8_t)(v >> (8 * b));
        }
    }
    return 0;
}

int dg_random_file(const char* path, uint64_t size, uint64_t seed, int threads) {
    int fd = open(path, O_CREAT | O_TRUNC | O_WRONLY | O_CLOEXEC, 0644);
    if (fd < 0) return -errno;
    std::atomic<int> err{0};
    const uint64_t block = 64 * kChunk;
    uint64_t nb = (size + block - 1) / block;
    parallel_for(nb, threads, [&](uint64_t b) {
        uint64_t off = b * block, len = std::min(block, size - off);
        std::vector<uint8_t> buf(len);
        dg_random_fill(buf.data(), off, len, seed);
        int r = write_all(fd, buf.data(), len, off);
        if (r) err = r;
    });
    close(fd);
    return err.load();
}

// n N(0,1) floats of stream `stream` (Box-Muller pairs).
int dg_normal_f32(float* out, uint64_t n, uint64_t seed, uint64_t stream) {
    Xoshiro g(seed, stream + 0x6e6f726dull);
    const double two_pi = 6.283185307179586476925286766559;
    for (uint64_t i = 0; i < n; i += 2) {
        double u1 = g.uniform_pos(), u2 = g.uniform();
        double r = std::sqrt(-2.0 * std::log(u1));
        out[i] = (float)(r * std::cos(two_pi * u2));
        if (i + 1 < n) out[i + 1] = (float)(r * std::sin(two_pi * u2));
    }
    return 0;
}

// rows x cols row-major fp32 matrix; row r = dg_normal_f32(stream = r).
int dg_matrix_file(const char* path, uint64_t rows, uint64_t cols, uint64_t seed, int threads) {
    int fd = open(path, O_CREAT | O_TRUNC | O_WRONLY | O_CLOEXEC, 0644);
    if (fd < 0) return -errno;
    std::atomic<int> err{0};
    const uint64_t rows_per = std::max<uint64_t>(1, (16ull << 20) / (cols * 4));
    uint64_t nb = (rows + rows_per - 1) / rows_per;
    parallel_for(nb, threads, [&](uint64_t b) {
        uint64_t r0 = b * rows_per, r1 = std::min(rows, r0 + rows_per);
        std::vector<float> buf((r1 - r0) * cols);
        for (uint64_t r = r0; r < r1; ++r) dg_normal_f32(buf.data() + (r - r0) * cols, cols, seed, r);
        int w = write_all(fd, (const uint8_t*)buf.data(), buf.size() * 4, r0 * cols * 4);
        if (w) err = w;
    });
    close(fd);
    return err.load();
}

// floor(exp(ln(median) + sigma * N(0,1))) clipped to [lo, hi].
int dg_lognormal_sizes(uint64_t seed, uint32_t n, double median, double sigma, uint64_t lo, uint64_t hi,
                       uint64_t* out) {
    std::vector<float> z(n + 1);
    dg_normal_f32(z.data(), n, seed, 0x73697a65ull);
    const double mu = std::log(median);
    for (uint32_t i = 0; i < n; ++i) {
        double v = std::floor(std::exp(mu + sigma * (double)z[i]));
        uint64_t s = v < (double)lo ? lo : (v > (double)hi ? hi : (uint64_t)v);
        out[i] = s;
    }
    return 0;
}

// count draws of Zipf(s) over items 0..nitems-1 (item 0 most popular).
int dg_zipf_sample(uint64_t seed, uint32_t nitems, double s, uint32_t count, uint32_t* out) {
    const Zipf& z = zipf_cached(nitems, s);
    Xoshiro g(seed, 0x7a697066ull);
    for (uint32_t i = 0; i < count; ++i) out[i] = z.sample(g);
    return 0;
}

// count exponential inter-arrival gaps (seconds) of a Poisson process at `rate`.
int dg_poisson_gaps(uint64_t seed, double rate, uint32_t count, double* out) {
    Xoshiro g(seed, 0x706f6973ull);
    for (uint32_t i = 0; i < count; ++i) out[i] = -std::log(g.uniform_pos()) / rate;
    return 0;
}

} // extern "C"

#include "generators.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <thread>

namespace gdb {
namespace infra {

uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

uint64_t rng(uint64_t seed, uint64_t stream, uint64_t counter) {
    const uint64_t state = mix64(seed * 0xD1B54A32D192ED03ull + mix64(stream));
    return mix64(state + counter * 0x9E3779B97F4A7C15ull);
}

int64_t uniformInt(uint64_t r, int64_t lo, int64_t hi) {
    const uint64_t span = static_cast<uint64_t>(hi - lo) + 1ull;
    if (span == 0) return lo + static_cast<int64_t>(r);  // full 64-bit range
    return lo + static_cast<int64_t>((static_cast<unsigned __int128>(r) * span) >> 64);
}

double uniform01(uint64_t r) { return static_cast<double>(r >> 11) * (1.0 / 9007199254740992.0); }

int defaultThreads() {
    unsigned hc = std::thread::hardware_concurrency();
    if (hc == 0) hc = 4;
    return static_cast<int>(std::min(hc, 16u));
}

namespace {

void parallelFor(uint64_t n, int threads, const std::function<void(uint64_t, uint64_t)>& body) {
    if (threads <= 0) threads = defaultThreads();
    if (n < 65536 || threads == 1) {
        body(0, n);
        return;
    }
    std::vector<std::thread> pool;
    const uint64_t chunk = (n + static_cast<uint64_t>(threads) - 1) / static_cast<uint64_t>(threads);
    for (int t = 0; t < threads; ++t) {
        const uint64_t b = static_cast<uint64_t>(t) * chunk;
        const uint64_t e = std::min(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back([=, &body]() { body(b, e); });
    }
    for (auto& th : pool) th.join();
}

// Balanced Feistel network over 2*h bits with 4 rounds.
uint64_t feistel(uint64_t seed, uint64_t stream, int h, uint64_t x) {
    const uint64_t mask = (1ull << h) - 1ull;
    uint64_t l = x >> h, r = x & mask;
    for (int round = 0; round < 4; ++round) {
        const uint64_t f = rng(seed, stream ^ 0xFE15ull, (static_cast<uint64_t>(round) << 40) ^ r) & mask;
        const uint64_t nl = r;
        r = l ^ f;
        l = nl;
    }
    return (l << h) | r;
}

}  // namespace

uint64_t permute(uint64_t seed, uint64_t stream, uint64_t n, uint64_t i) {
    if (n <= 1) return 0;
    int bits = 0;
    while ((1ull << bits) < n) ++bits;
    if (bits < 2) bits = 2;
    if (bits & 1) ++bits;
    const int h = bits / 2;
    uint64_t x = i;
    do {
        x = feistel(seed, stream, h, x);
    } while (x >= n);
    return x;
}

void genUniformI32(uint64_t seed, uint64_t stream, uint64_t n, int64_t lo, int64_t hi, int32_t* out, int threads) {
    parallelFor(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) out[i] = static_cast<int32_t>(uniformInt(rng(seed, stream, i), lo, hi));
    });
}

void genUniformI64(uint64_t seed, uint64_t stream, uint64_t n, int64_t lo, int64_t hi, int64_t* out, int threads) {
    parallelFor(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) out[i] = uniformInt(rng(seed, stream, i), lo, hi);
    });
}

void genCategoricalU8(uint64_t seed, uint64_t stream, uint64_t n, const uint8_t* values, const uint32_t* weights,
                      int nvalues, uint8_t* out, int threads) {
    if (nvalues <= 0) throw std::runtime_error("genCategoricalU8: no values");
    std::vector<uint64_t> cum(static_cast<size_t>(nvalues));
    uint64_t total = 0;
    for (int k = 0; k < nvalues; ++k) cum[static_cast<size_t>(k)] = (total += weights[k]);
    if (total == 0) throw std::runtime_error("genCategoricalU8: zero total weight");
    parallelFor(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            const uint64_t u = static_cast<uint64_t>(uniformInt(rng(seed, stream, i), 0, static_cast<int64_t>(total - 1)));
            int k = 0;
            while (u >= cum[static_cast<size_t>(k)]) ++k;
            out[i] = values[k];
        }
    });
}

void genPermutationI32(uint64_t seed, uint64_t stream, uint64_t n, int64_t base, int32_t* out, int threads) {
    parallelFor(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) out[i] = static_cast<int32_t>(base + static_cast<int64_t>(permute(seed, stream, n, i)));
    });
}

namespace {
int fanoutCount(uint64_t seed, uint64_t stream, int cmin, int cmax, uint64_t key) {
    return static_cast<int>(uniformInt(rng(seed, stream, key), cmin, cmax));
}
}  // namespace

uint64_t fanoutTotal(uint64_t seed, uint64_t stream, uint64_t nparent, int cmin, int cmax, int threads) {
    if (threads <= 0) threads = defaultThreads();
    std::vector<uint64_t> part(static_cast<size_t>(threads), 0);
    const uint64_t chunk = (nparent + static_cast<uint64_t>(threads) - 1) / static_cast<uint64_t>(threads);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        pool.emplace_back([&, t]() {
            const uint64_t b = static_cast<uint64_t>(t) * chunk, e = std::min(nparent, b + chunk);
            uint64_t s = 0;
            for (uint64_t k = b; k < e; ++k) s += static_cast<uint64_t>(fanoutCount(seed, stream, cmin, cmax, k + 1));
            part[static_cast<size_t>(t)] = s;
        });
    }
    for (auto& th : pool) th.join();
    uint64_t total = 0;
    for (auto v : part) total += v;
    return total;
}

void genFanoutFkeyI32(uint64_t seed, uint64_t stream, uint64_t nparent, int cmin, int cmax, uint64_t nchild,
                      int32_t* out, int threads) {
    // prefix[k] = number of children of keys 1..k
    std::vector<uint64_t> prefix(nparent + 1, 0);
    for (uint64_t k = 1; k <= nparent; ++k)
        prefix[k] = prefix[k - 1] + static_cast<uint64_t>(fanoutCount(seed, stream, cmin, cmax, k));
    if (prefix[nparent] != nchild) throw std::runtime_error("genFanoutFkeyI32: nchild does not match fanoutTotal");
    parallelFor(nchild, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            const uint64_t pos = permute(seed, stream ^ 0xC41Dull, nchild, i);
            // key k owns positions [prefix[k-1], prefix[k])
            const uint64_t k = static_cast<uint64_t>(std::upper_bound(prefix.begin(), prefix.end(), pos) - prefix.begin());
            out[i] = static_cast<int32_t>(k);
        }
    });
}

// Rejection-inversion sampling for a monotone discrete distribution
// P(k) ~ k^-s on 1..N (W. Hoermann, G. Derflinger, "Rejection-inversion to
// generate variates from monotone discrete distributions", ACM TOMACS 6(3),
// 1996). H is an antiderivative of h(x) = x^-s; helper functions keep the
// s -> 1 limit numerically stable.
namespace {
struct ZipfRI {
    double s;
    double n;
    double hX1, hN, sv;
    static double helper1(double x) { return std::fabs(x) > 1e-8 ? std::log1p(x) / x : 1.0 - x * (0.5 - x * (1.0 / 3.0 - 0.25 * x)); }
    static double helper2(double x) { return std::fabs(x) > 1e-8 ? std::expm1(x) / x : 1.0 + x * 0.5 * (1.0 + x * (1.0 / 3.0) * (1.0 + 0.25 * x)); }
    double h(double x) const { return std::exp(-s * std::log(x)); }
    double H(double x) const {
        const double lx = std::log(x);
        return helper2((1.0 - s) * lx) * lx;
    }
    double Hinv(double x) const {
        double t = x * (1.0 - s);
        if (t < -1.0) t = -1.0;
        return std::exp(helper1(t) * x);
    }
    ZipfRI(double s_, int64_t n_) : s(s_), n(static_cast<double>(n_)) {
        hX1 = H(1.5) - 1.0;
        hN = H(n + 0.5);
        sv = 2.0 - Hinv(H(2.5) - h(2.0));
    }
    int64_t sample(uint64_t seed, uint64_t stream, uint64_t row) const {
        for (uint64_t attempt = 0;; ++attempt) {
            const double r = uniform01(rng(seed, stream + (attempt << 40), row));
            const double u = hN + r * (hX1 - hN);
            const double x = Hinv(u);
            double kd = std::floor(x + 0.5);
            if (kd < 1.0) kd = 1.0;
            if (kd > n) kd = n;
            if (kd - x <= sv || u >= H(kd + 0.5) - h(kd)) return static_cast<int64_t>(kd);
        }
    }
};
}  // namespace

void genZipfI32(uint64_t seed, uint64_t stream, uint64_t n, double s, int64_t nkeys, int32_t* out, int threads) {
    if (nkeys < 1) throw std::runtime_error("genZipfI32: nkeys must be >= 1");
    if (s < 0.0) throw std::runtime_error("genZipfI32: exponent must be >= 0");
    ZipfRI z(s, nkeys);
    parallelFor(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) out[i] = static_cast<int32_t>(z.sample(seed, stream, i));
    });
}

void genVocabulary(uint64_t seed, int nwords, int minlen, int maxlen, char* out) {
    if (minlen < 1 || maxlen > 15 || minlen > maxlen) throw std::runtime_error("genVocabulary: bad lengths");
    std::memset(out, 0, static_cast<size_t>(nwords) * 16);
    for (int w = 0; w < nwords; ++w) {
        const int len = static_cast<int>(uniformInt(rng(seed, 0x70CAB, static_cast<uint64_t>(w)), minlen, maxlen));
        for (int c = 0; c < len; ++c)
            out[w * 16 + c] = static_cast<char>(
                'a' + uniformInt(rng(seed, 0x70CAC, static_cast<uint64_t>(w) * 16 + static_cast<uint64_t>(c)), 0, 25));
    }
}

void genText(uint64_t seed, uint64_t stream, uint64_t n, int width, const char* vocab, int nwords, char* out,
             int threads) {
    parallelFor(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            char* dst = out + i * static_cast<uint64_t>(width);
            std::memset(dst, 0, static_cast<size_t>(width));
            int pos = 0;
            for (int k = 0; k < 8; ++k) {
                const int w = static_cast<int>(uniformInt(rng(seed, stream, i * 8 + static_cast<uint64_t>(k)), 0, nwords - 1));
                const char* word = vocab + w * 16;
                const int len = static_cast<int>(strnlen(word, 16));
                const int need = len + (pos > 0 ? 1 : 0);
                if (pos + need > width) break;
                if (pos > 0) dst[pos++] = ' ';
                std::memcpy(dst + pos, word, static_cast<size_t>(len));
                pos += len;
            }
        }
    });
}

}  // namespace infra
}  // namespace gdb

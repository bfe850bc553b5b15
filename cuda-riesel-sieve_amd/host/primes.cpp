// primes.cpp -- CPU prime generator feeding the device sieve.
//
// Replaces the reference's generateGPUPrimes (kernel_v0.24.6.cu:1838-1987)
// and its small-prime set-up (:1271-1364).  The reference marks an odd-only
// window with 8 wheel masks and then every multiple of the first 4096*s odd
// primes, and takes the first arraySize unmarked positions; that is a
// partial sieve, exact only below (largest small prime)^2.  This generator
// keeps -s as "at least 4096*s sieving primes", adds sieving primes up to
// sqrt(hi) (capped at 2^25), proves survivors above the sieving bound
// squared with deterministic Miller-Rabin, and splits the window over
// threads.  Output keeps the reference's layout: uint32 offsets from an odd
// uint64 low (:1960-1966).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>

#include "host.h"

struct rs_primegen {
    std::vector<uint32_t> sp;  // odd sieving primes, ascending
    uint64_t bound2 = 0;       // survivors below this are prime without MR
    int threads = 1;
};

namespace {

const uint64_t kSegBits = 1u << 21;  // odd numbers per segment (256 KiB of bits)

uint64_t isqrt_ceil(uint64_t n) {
    uint64_t r = (uint64_t)std::sqrt((double)n);
    while ((rsh::u128)r * r < n) ++r;
    while (r > 0 && (rsh::u128)(r - 1) * (r - 1) >= n) --r;
    return r;
}

// Sieves the odd numbers base + 2i, i in [0, nbits), appending primes as
// offsets from `origin`.
void sieve_segment(const rs_primegen* g, uint64_t base, uint64_t nbits, uint64_t origin,
                   std::vector<uint64_t>& bits, std::vector<uint32_t>& out) {
    const size_t words = (nbits + 63) / 64;
    bits.assign(words, 0);
    const uint64_t top = base + 2 * (nbits - 1);
    for (uint32_t q : g->sp) {
        const uint64_t qq = (uint64_t)q * q;
        if (qq > top) break;
        // first i with q | base + 2i and base + 2i >= q^2
        uint64_t i;
        if (qq >= base) {
            i = (qq - base) / 2;
        } else {
            const uint64_t r = base % q;
            i = ((q - r) % q) * (uint64_t)((q + 1) / 2) % q;
        }
        for (; i < nbits; i += q) bits[i >> 6] |= 1ull << (i & 63);
    }
    for (size_t w = 0; w < words; ++w) {
        uint64_t free = ~bits[w];
        if (w == words - 1 && (nbits & 63)) free &= (1ull << (nbits & 63)) - 1;
        while (free) {
            const int b = __builtin_ctzll(free);
            free &= free - 1;
            const uint64_t v = base + 2 * (w * 64 + b);
            if (v < 3) continue;  // 1 is not prime (2 is never in an odd window)
            if (v >= g->bound2 && !rsh::is_prime_u64(v)) continue;
            out.push_back((uint32_t)(v - origin));
        }
    }
}

}  // namespace

extern "C" rs_status rs_primegen_create(int small_prime_scale, uint64_t hi, int threads,
                                        rs_primegen** out) {
    if (!out || small_prime_scale < 1 || small_prime_scale > 4096) return RS_ERR_ARG;
    *out = nullptr;
    rs_primegen* g = new rs_primegen();
    g->threads = rsh::hw_threads(threads);
    const uint64_t want = 4096ull * small_prime_scale;
    uint64_t lim = std::max<uint64_t>(isqrt_ceil(hi ? hi : 3), 64);
    lim = std::min<uint64_t>(lim, 1ull << 25);
    // grow the bound until it holds 4096*s odd primes (the -s meaning, :1274)
    while (true) {
        std::vector<uint32_t> ps = rsh::primes_upto((uint32_t)lim);
        if (ps.size() >= want + 1 || lim >= (1ull << 31)) {
            g->sp.assign(ps.begin() + 1, ps.end());  // drop 2
            break;
        }
        lim *= 2;
    }
    const uint64_t last = g->sp.back();
    g->bound2 = last * last;
    *out = g;
    return RS_OK;
}

extern "C" rs_status rs_primes_range(rs_primegen* g, uint64_t lo, uint64_t hi,
                                     uint32_t* offsets, int64_t cap, int64_t* count) {
    if (!g || !count || hi < lo || hi - lo > (1ull << 32) || (cap > 0 && !offsets))
        return RS_ERR_ARG;
    *count = 0;
    if (hi == lo) return RS_OK;
    std::vector<uint32_t> two;
    if (lo <= 2 && hi > 2) two.push_back((uint32_t)(2 - lo));
    const uint64_t base = (lo % 2) ? lo : lo + 1;  // first odd >= lo
    if (base >= hi) {
        *count = (int64_t)two.size();
        if ((int64_t)two.size() > cap) return RS_ERR_OVERFLOW;
        if (!two.empty()) offsets[0] = two[0];
        return RS_OK;
    }
    const uint64_t nodd = (hi - base + 1) / 2;
    const uint64_t nseg = (nodd + kSegBits - 1) / kSegBits;
    std::vector<std::vector<uint32_t>> parts(nseg);
    std::atomic<uint64_t> next(0);
    std::vector<std::thread> th;
    const int nt = (int)std::min<uint64_t>(g->threads, nseg);
    for (int t = 0; t < nt; ++t) {
        th.emplace_back([&]() {
            std::vector<uint64_t> bits;
            for (uint64_t s; (s = next.fetch_add(1)) < nseg;) {
                const uint64_t b0 = base + 2 * s * kSegBits;
                const uint64_t nb = std::min<uint64_t>(kSegBits, nodd - s * kSegBits);
                sieve_segment(g, b0, nb, lo, bits, parts[s]);
            }
        });
    }
    for (auto& x : th) x.join();
    int64_t total = (int64_t)two.size();
    for (auto& p : parts) total += (int64_t)p.size();
    *count = total;
    if (total > cap) return RS_ERR_OVERFLOW;
    int64_t c = 0;
    for (uint32_t v : two) offsets[c++] = v;
    for (auto& p : parts) {
        memcpy(offsets + c, p.data(), p.size() * sizeof(uint32_t));
        c += (int64_t)p.size();
    }
    return RS_OK;
}

extern "C" rs_status rs_primes_batch(rs_primegen* g, uint64_t low, uint32_t* offsets,
                                     int64_t count, uint64_t* next_low) {
    if (!g || !offsets || count < 1) return RS_ERR_ARG;
    if (low % 2 == 0) low -= 1;  // :1515-1519
    // window estimate: count primes need about count*ln(p) numbers
    double lnp = std::log((double)std::max<uint64_t>(low, 16));
    uint64_t width = (uint64_t)(count * lnp * 1.15) + 4096;
    while (true) {
        width = std::min<uint64_t>(width, 1ull << 32);
        std::vector<uint32_t> buf((size_t)(width / std::max(1.0, lnp - 1.2) * 1.1) + 1024);
        int64_t got = 0;
        rs_status st = rs_primes_range(g, low, low + width, buf.data(), (int64_t)buf.size(), &got);
        if (st == RS_ERR_OVERFLOW) {
            buf.resize((size_t)got);
            st = rs_primes_range(g, low, low + width, buf.data(), got, &got);
        }
        if (st != RS_OK) return st;
        if (got >= count) {
            memcpy(offsets, buf.data(), (size_t)count * sizeof(uint32_t));
            if (next_low) *next_low = low + offsets[count - 1] + 2;  // :1561
            return RS_OK;
        }
        if (width == (1ull << 32)) return RS_ERR_RANGE;
        width *= 2;
    }
}

extern "C" void rs_primegen_free(rs_primegen* g) { delete g; }

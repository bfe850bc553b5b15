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
//
// Sieve: odd numbers only, one bit each, in segments of kSegBits; a thread
// takes blocks of kBlockSegs consecutive segments and carries every sieving
// prime's next multiple from segment to segment (one remainder per prime
// per block, by a precomputed reciprocal, not a division per segment).  The multiples of 3, 5, 7, 11 and 13
// come from a precomputed pattern of period 15015 bits, copied word by word
// at the segment's phase.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>

#include "host.h"

namespace {

const uint32_t kWheel[] = {3, 5, 7, 11, 13};
const uint32_t kPeriod = 3 * 5 * 7 * 11 * 13;  // 15015 odd numbers

}  // namespace

struct rs_primegen {
    std::vector<uint32_t> sp;   // odd sieving primes above 13, ascending
    std::vector<uint64_t> pat;  // bit j: 2j + 1 has a factor in kWheel, for
                                // j < kPeriod + 128 (wrapped)
    std::vector<uint64_t> rq;   // floor((2^64 - 1) / q) per sieving prime
    size_t nsmall = 0;          // sieving primes below kSmallQ
    uint64_t bound2 = 0;        // survivors below this are prime without MR
    int threads = 1;
};

namespace {

const uint64_t kSegBits = 1u << 21;  // odd numbers per segment (256 KiB of bits)
const uint64_t kBlockSegs = 16;      // segments per thread block
const uint64_t kSliceBits = 1u << 18;  // 32 KiB slices for the small primes
const uint32_t kSmallQ = 1u << 15;

uint64_t isqrt_ceil(uint64_t n) {
    uint64_t r = (uint64_t)std::sqrt((double)n);
    while ((rsh::u128)r * r < n) ++r;
    while (r > 0 && (rsh::u128)(r - 1) * (r - 1) >= n) --r;
    return r;
}

// 64 pattern bits starting at bit `pos` (pos < kPeriod + 64).
inline uint64_t pattern_word(const std::vector<uint64_t>& pat, uint64_t pos) {
    const uint64_t w = pos >> 6, sh = pos & 63;
    return sh ? (pat[w] >> sh) | (pat[w + 1] << (64 - sh)) : pat[w];
}

// Sieves nseg segments of odd numbers from the odd `base` (nbits_total
// bits in all), appending primes as offsets from `origin`.
void sieve_block(const rs_primegen* g, uint64_t base, uint64_t nbits_total, uint64_t origin,
                 std::vector<uint64_t>& bits, std::vector<uint64_t>& next,
                 std::vector<uint32_t>& out) {
    const uint64_t top = base + 2 * (nbits_total - 1);
    // sieving primes that reach the block, and the bit index of the first
    // multiple of each at or above both base and q^2
    size_t np = 0;
    while (np < g->sp.size() && (uint64_t)g->sp[np] * g->sp[np] <= top) ++np;
    next.resize(np);
    for (size_t i = 0; i < np; ++i) {
        const uint64_t q = g->sp[i], qq = q * q;
        if (qq >= base) {
            next[i] = (qq - base) / 2;
        } else {
            // r = base mod q by the reciprocal (the estimated quotient is
            // short by at most one); base + d == 0 mod q for d = q - r (or
            // 0), and base + 2j for j = d / 2, or (d + q) / 2 with d odd
            uint64_t r = base - (uint64_t)(((rsh::u128)base * g->rq[i]) >> 64) * q;
            if (r >= q) r -= q;
            const uint64_t d = r ? q - r : 0;
            next[i] = (d & 1) ? (d + q) / 2 : d / 2;
        }
    }
    for (uint64_t done = 0; done < nbits_total; done += kSegBits) {
        const uint64_t nbits = std::min(kSegBits, nbits_total - done);
        const uint64_t b0 = base + 2 * done;
        const size_t words = (nbits + 63) / 64;
        bits.resize(words);
        // wheel primes from the pattern: bit i is 2j + 1 = b0 + 2i, j = (b0 - 1) / 2 + i
        uint64_t pos = ((b0 - 1) / 2) % kPeriod;
        for (size_t w = 0; w < words; ++w) {
            bits[w] = pattern_word(g->pat, pos);
            pos += 64;
            if (pos >= kPeriod) pos -= kPeriod;
        }
        if (b0 <= 13)  // the wheel primes themselves
            for (uint32_t q : kWheel)
                if (q >= b0 && (q - b0) / 2 < nbits) bits[(q - b0) / 2 >> 6] &= ~(1ull << ((q - b0) / 2 & 63));
        // primes below kSmallQ (many marks each) a cache-sized slice at a
        // time, the rest (at most a few marks each) over the whole segment
        const size_t ns = std::min(np, g->nsmall);
        for (uint64_t s0 = 0; s0 < nbits; s0 += kSliceBits) {
            const uint64_t s1 = std::min(nbits, s0 + kSliceBits);
            for (size_t i = 0; i < ns; ++i) {
                const uint64_t q = g->sp[i];
                uint64_t j = next[i];
                for (; j < s1; j += q) bits[j >> 6] |= 1ull << (j & 63);
                next[i] = j;
            }
        }
        for (size_t i = 0; i < ns; ++i) next[i] -= nbits;
        for (size_t i = ns; i < np; ++i) {
            const uint64_t q = g->sp[i];
            uint64_t j = next[i];
            for (; j < nbits; j += q) bits[j >> 6] |= 1ull << (j & 63);
            next[i] = j - nbits;
        }
        for (size_t w = 0; w < words; ++w) {
            uint64_t free = ~bits[w];
            if (w == words - 1 && (nbits & 63)) free &= (1ull << (nbits & 63)) - 1;
            while (free) {
                const int b = __builtin_ctzll(free);
                free &= free - 1;
                const uint64_t v = b0 + 2 * (w * 64 + b);
                if (v < 3) continue;  // 1 is not prime (2 is never in an odd window)
                if (v >= g->bound2 && !rsh::is_prime_u64(v)) continue;
                out.push_back((uint32_t)(v - origin));
            }
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
    // every prime <= lim sieves, so a survivor below (lim + 1)^2 is prime
    g->bound2 = (lim + 1) * (lim + 1);
    // the wheel primes are marked by the pattern
    g->sp.erase(g->sp.begin(), std::upper_bound(g->sp.begin(), g->sp.end(), 13u));
    g->nsmall = std::lower_bound(g->sp.begin(), g->sp.end(), kSmallQ) - g->sp.begin();
    g->rq.resize(g->sp.size());
    for (size_t i = 0; i < g->sp.size(); ++i) g->rq[i] = ~0ull / g->sp[i];
    g->pat.assign((kPeriod + 128 + 63) / 64 + 1, 0);
    for (uint32_t j = 0; j < kPeriod + 128; ++j) {
        const uint32_t v = 2 * (j % kPeriod) + 1;
        for (uint32_t q : kWheel)
            if (v % q == 0) g->pat[j >> 6] |= 1ull << (j & 63);
    }
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
    // blocks of kBlockSegs segments, fewer when that leaves threads idle
    const uint64_t bsegs =
        std::max<uint64_t>(1, std::min<uint64_t>(kBlockSegs, nseg / std::max(1, g->threads)));
    const uint64_t nblk = (nseg + bsegs - 1) / bsegs;
    std::vector<std::vector<uint32_t>> parts(nblk);
    std::atomic<uint64_t> next(0);
    std::vector<std::thread> th;
    const int nt = (int)std::min<uint64_t>(g->threads, nblk);
    auto work = [&]() {
        std::vector<uint64_t> bits, carry;
        for (uint64_t k; (k = next.fetch_add(1)) < nblk;) {
            const uint64_t b0 = base + 2 * k * bsegs * kSegBits;
            const uint64_t nb = std::min<uint64_t>(bsegs * kSegBits, nodd - k * bsegs * kSegBits);
            sieve_block(g, b0, nb, lo, bits, carry, parts[k]);
        }
    };
    for (int t = 1; t < nt; ++t) th.emplace_back(work);
    work();  // the calling thread takes blocks too
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

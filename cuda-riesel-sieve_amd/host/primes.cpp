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
// Sieve: a mod-30 wheel, one byte per 30 numbers (bit i: 30 j + kRes[i]), in
// segments of kSegBytes; a thread takes blocks of up to kBlockSegs
// consecutive segments and carries every sieving prime's next multiple
// (byte position and wheel index) from segment to segment: one remainder
// per prime per block, by a precomputed reciprocal.  The multiples of 7, 11
// and 13 come from a precomputed pattern of period 1001 bytes; primes below
// kSmallQ mark 32 KiB slices at a time, eight wheel steps per loop turn.
// Each block's primes go to a buffer of its own, and then, all threads
// copying at once, to their place in the output.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>

#include "host.h"

namespace {

// the residues coprime to 30, the gaps between them, and the index of each
const uint32_t kRes[8] = {1, 7, 11, 13, 17, 19, 23, 29};
const uint32_t kStep[8] = {6, 4, 2, 4, 2, 4, 6, 2};
const int kPatQ[3] = {7, 11, 13};
const uint32_t kPeriod = 7 * 11 * 13;  // pattern bytes

struct WheelTables {
    int8_t idx[30];         // index of a residue in kRes, or -1
    uint8_t up[30];         // distance from a residue to the next one prime to 30
    uint8_t mask[8][8];     // bit of q k, for q = kRes[qi] and k = kRes[wi] (mod 30)
    uint8_t add[8][8];      // extra bytes of the step from q k to q (k + kStep[wi])
    WheelTables() {
        for (int r = 0; r < 30; ++r) idx[r] = -1;
        for (int i = 0; i < 8; ++i) idx[kRes[i]] = (int8_t)i;
        for (int r = 0; r < 30; ++r) {
            int d = 0;
            while (idx[(r + d) % 30] < 0) ++d;
            up[r] = (uint8_t)d;
        }
        for (int qi = 0; qi < 8; ++qi)
            for (int wi = 0; wi < 8; ++wi) {
                const uint32_t rv = kRes[qi] * kRes[wi] % 30;
                mask[qi][wi] = (uint8_t)(1u << idx[rv]);
                add[qi][wi] = (uint8_t)((rv + kRes[qi] * kStep[wi]) / 30);
            }
    }
};
const WheelTables kW;

// A growable run of offsets whose new elements are left uninitialised (a
// vector zero-fills what it grows by, a second pass over every prime).
struct OffsBuf {
    std::unique_ptr<uint32_t[]> p;
    size_t n = 0, cap = 0;
    void reserve(size_t c) {
        if (c <= cap) return;
        c = std::max(c, cap + cap / 2);
        std::unique_ptr<uint32_t[]> q(new uint32_t[c]);
        if (n) memcpy(q.get(), p.get(), n * sizeof(uint32_t));
        p = std::move(q);
        cap = c;
    }
};

}  // namespace

struct rs_primegen {
    std::vector<uint32_t> sp;  // sieving primes above 13, ascending
    std::vector<uint64_t> rq;  // floor((2^64 - 1) / q) per sieving prime
    std::vector<uint32_t> q30;  // q / 30 per sieving prime
    std::vector<uint8_t> qi;    // the index of q mod 30 per sieving prime
    std::vector<uint8_t> pat;  // byte j: bit i set when 30 j + kRes[i] has a
                               // factor 7, 11 or 13, j < 2 kPeriod
    size_t nsmall = 0;         // sieving primes below kSmallQ
    uint64_t bound2 = 0;       // survivors below this are prime without MR
    int threads = 1;
    // per-block outputs kept from call to call (no page faults on the next
    // call's hundreds of MB); a call made while another holds them uses its own
    std::mutex parts_mu;
    std::vector<OffsBuf> parts;
};

namespace {

const uint64_t kSegBytes = 1u << 21;   // 2 MiB: 63M numbers per segment
const uint64_t kBlockSegs = 16;        // segments per thread block
const uint64_t kMinBlockBytes = 1u << 19;  // the smallest block: 15.7M numbers
const uint64_t kSliceBytes = 1u << 15;  // 32 KiB slices for the small primes
const uint32_t kSmallQ = 1u << 15;

uint64_t isqrt_ceil(uint64_t n) {
    uint64_t r = (uint64_t)std::sqrt((double)n);
    while ((rsh::u128)r * r < n) ++r;
    while (r > 0 && (rsh::u128)(r - 1) * (r - 1) >= n) --r;
    return r;
}

// Marks the multiples q k of one prime in seg[0, end) from byte `pos`,
// wheel index `wi` (k = kRes[wi] mod 30) on; leaves pos, wi at the first
// multiple at or past `end`.
inline void mark_prime(uint8_t* seg, uint64_t end, uint32_t q, uint64_t& pos, int& wi) {
    const uint32_t q30 = q / 30;
    const int qi = kW.idx[q % 30];
    const uint8_t* const mk = kW.mask[qi];
    const uint8_t* const ad = kW.add[qi];
    uint64_t p = pos;
    int w = wi;
    // eight wheel steps (one turn: q bytes) per loop while a whole turn fits
    if (p + q < end) {
        uint32_t inc[8];
        for (int i = 0; i < 8; ++i) inc[i] = q30 * kStep[(w + i) & 7] + ad[(w + i) & 7];
        uint8_t m[8];
        for (int i = 0; i < 8; ++i) m[i] = mk[(w + i) & 7];
        while (p + q < end) {
            seg[p] |= m[0]; p += inc[0];
            seg[p] |= m[1]; p += inc[1];
            seg[p] |= m[2]; p += inc[2];
            seg[p] |= m[3]; p += inc[3];
            seg[p] |= m[4]; p += inc[4];
            seg[p] |= m[5]; p += inc[5];
            seg[p] |= m[6]; p += inc[6];
            seg[p] |= m[7]; p += inc[7];
        }
    }
    while (p < end) {
        seg[p] |= mk[w];
        p += q30 * kStep[w] + ad[w];
        w = (w + 1) & 7;
    }
    pos = p;
    wi = w;
}

// Sieves the numbers [base, base + 30 nbytes_total) (base a multiple of 30)
// segment by segment, appending the primes of [lo, hi) as offsets from
// `origin`.
void sieve_block(const rs_primegen* g, uint64_t base, uint64_t nbytes_total, uint64_t lo,
                 uint64_t hi, uint64_t origin, std::vector<uint8_t>& seg,
                 std::vector<uint64_t>& next, std::vector<uint8_t>& wheel,
                 OffsBuf& out) {
    const uint64_t top = std::min<uint64_t>(base + 30 * nbytes_total, hi) - 1;
    // sieving primes that reach the block, and for each the first multiple
    // q k >= max(q^2, base) with k prime to 30, as a byte position and the
    // wheel index of k
    size_t np = 0;
    while (np < g->sp.size() && (uint64_t)g->sp[np] * g->sp[np] <= top) ++np;
    next.resize(np);
    wheel.resize(np);
    for (size_t i = 0; i < np; ++i) {
        const uint64_t q = g->sp[i];
        uint64_t k;
        if (q * q >= base) {
            k = q;
        } else {
            // k = ceil(base / q) by the reciprocal (the estimate is short by
            // at most one)
            uint64_t d = (uint64_t)(((rsh::u128)base * g->rq[i]) >> 64);
            if (base - d * q >= q) ++d;
            k = d + (base - d * q != 0);
        }
        const uint32_t kr0 = (uint32_t)(k % 30);
        k += kW.up[kr0];
        const uint32_t kr = (kr0 + kW.up[kr0]) % 30;
        next[i] = (q * k - base) / 30;
        wheel[i] = (uint8_t)kW.idx[kr];
    }
    const size_t ns = std::min(np, g->nsmall);
    // room for the block's primes up front (about 30 nbytes / ln(base)),
    // so the output does not move while it grows
    const double lnb = std::log((double)std::max<uint64_t>(base + 30 * nbytes_total, 16));
    out.reserve(out.n + (size_t)(30.0 * (double)nbytes_total / (lnb - 1.2) * 1.05) + 4096);
    for (uint64_t done = 0; done < nbytes_total; done += kSegBytes) {
        const uint64_t nb = std::min(kSegBytes, nbytes_total - done);
        const uint64_t b0 = base + 30 * done;
        seg.resize(nb + 8);
        // 7, 11, 13 from the pattern, at the segment's phase
        uint64_t ph = (b0 / 30) % kPeriod;
        for (uint64_t at = 0; at < nb;) {
            const uint64_t n = std::min<uint64_t>(nb - at, kPeriod);
            memcpy(seg.data() + at, g->pat.data() + ph, n);
            at += n;
            ph = (ph + n) % kPeriod;
        }
        memset(seg.data() + nb, 0xff, 8);  // past the end: never a prime
        if (b0 == 0) seg[0] = (uint8_t)((seg[0] & ~0x0e) | 1);  // 7, 11, 13 prime; 1 not
        for (uint64_t s0 = 0; s0 < nb; s0 += kSliceBytes) {
            const uint64_t s1 = std::min(nb, s0 + kSliceBytes);
            for (size_t i = 0; i < ns; ++i) {
                int w = wheel[i];
                mark_prime(seg.data(), s1, g->sp[i], next[i], w);
                wheel[i] = (uint8_t)w;
            }
        }
        for (size_t i = 0; i < ns; ++i) next[i] -= nb;
        // the rest: a few marks per segment each, most none
        uint8_t* const sg = seg.data();
        for (size_t i = ns; i < np; ++i) {
            uint64_t p = next[i];
            if (p < nb) {
                const uint32_t q30 = g->q30[i];
                const uint8_t* const mk = kW.mask[g->qi[i]];
                const uint8_t* const ad = kW.add[g->qi[i]];
                int w = wheel[i];
                do {
                    sg[p] |= mk[w];
                    p += q30 * kStep[w] + ad[w];
                    w = (w + 1) & 7;
                } while (p < nb);
                wheel[i] = (uint8_t)w;
            }
            next[i] = p - nb;
        }
        // survivors, in order: the words of 8 bytes (240 numbers) at both
        // ends of the window check the bounds, the test above bound2 only
        // where the segment reaches it; room for them from a popcount pass
        size_t nfree = 0;
        for (uint64_t w = 0; w < nb; w += 8) {
            uint64_t word;
            memcpy(&word, seg.data() + w, 8);
            nfree += (size_t)__builtin_popcountll(~word);
        }
        out.reserve(out.n + nfree);
        uint32_t* o = out.p.get() + out.n;
        const bool mr = b0 + 30 * nb > g->bound2;
        const uint64_t off = b0 - origin;
        for (uint64_t w = 0; w < nb; w += 8) {
            uint64_t word;
            memcpy(&word, seg.data() + w, 8);
            uint64_t free = ~word;
            const uint64_t vw = b0 + 30 * w;
            if (vw < lo || vw + 240 > hi || mr) {
                while (free) {
                    const int b = __builtin_ctzll(free);
                    free &= free - 1;
                    const uint64_t v = vw + 30 * (uint64_t)(b >> 3) + kRes[b & 7];
                    if (v < lo || v >= hi) continue;
                    if (v >= g->bound2 && !rsh::is_prime_u64(v)) continue;
                    *o++ = (uint32_t)(v - origin);
                }
            } else {
                const uint32_t ow = (uint32_t)(off + 30 * w);
                while (free) {
                    const int b = __builtin_ctzll(free);
                    free &= free - 1;
                    *o++ = ow + 30 * (uint32_t)(b >> 3) + kRes[b & 7];
                }
            }
        }
        out.n = (size_t)(o - out.p.get());
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
    // 2, 3 and 5 are outside the wheel, 7, 11 and 13 come from the pattern
    g->sp.erase(g->sp.begin(), std::upper_bound(g->sp.begin(), g->sp.end(), 13u));
    g->nsmall = std::lower_bound(g->sp.begin(), g->sp.end(), kSmallQ) - g->sp.begin();
    g->rq.resize(g->sp.size());
    g->q30.resize(g->sp.size());
    g->qi.resize(g->sp.size());
    for (size_t i = 0; i < g->sp.size(); ++i) {
        g->rq[i] = ~0ull / g->sp[i];
        g->q30[i] = g->sp[i] / 30;
        g->qi[i] = (uint8_t)kW.idx[g->sp[i] % 30];
    }
    g->pat.assign(2 * kPeriod, 0);
    for (uint32_t j = 0; j < 2 * kPeriod; ++j)
        for (int i = 0; i < 8; ++i) {
            const uint64_t v = 30ull * j + kRes[i];
            for (int q : kPatQ)
                if (v % q == 0) g->pat[j] |= (uint8_t)(1u << i);
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
    std::vector<uint32_t> small;  // 2, 3, 5 (outside the wheel)
    for (uint32_t q : {2u, 3u, 5u})
        if (q >= lo && q < hi) small.push_back((uint32_t)(q - lo));
    const uint64_t base = lo / 30 * 30;
    const uint64_t nbytes = (hi - base + 29) / 30;
    const uint64_t nseg = (nbytes + kSegBytes - 1) / kSegBytes;
    // equal blocks of at most about kBlockSegs segments, a whole number of
    // them per thread (multiples of 8 bytes, the last one takes the rest)
    const uint64_t nt0 = (uint64_t)std::max(1, g->threads);
    const uint64_t rounds = std::max<uint64_t>(1, (nseg + nt0 * kBlockSegs - 1) / (nt0 * kBlockSegs));
    // (each block computes every sieving prime's first multiple, about 3 ms
    // at 1e14: blocks of at least kMinBlockBytes, so that a narrow window
    // still keeps every thread busy, or one block)
    const uint64_t nblk =
        std::max<uint64_t>(1, std::min<uint64_t>(nt0 * rounds, nbytes / kMinBlockBytes));
    const uint64_t bbytes = (nbytes / nblk + 7) & ~7ull;
    std::unique_lock<std::mutex> own(g->parts_mu, std::try_to_lock);
    std::vector<OffsBuf> mine;
    std::vector<OffsBuf>& parts = own.owns_lock() ? g->parts : mine;
    if (parts.size() < nblk) parts.resize(nblk);
    for (uint64_t k = 0; k < nblk; ++k) parts[k].n = 0;
    const int nt = (int)std::min<uint64_t>(nt0, nblk);
    auto run = [&](const std::function<void(int)>& f) {
        std::vector<std::thread> th;
        for (int t = 1; t < nt; ++t) th.emplace_back(f, t);
        f(0);  // the calling thread takes a share too
        for (auto& x : th) x.join();
    };
    std::atomic<uint64_t> next(0);
    run([&](int) {
        std::vector<uint8_t> seg, wheel;
        std::vector<uint64_t> carry;
        for (uint64_t k; (k = next.fetch_add(1)) < nblk;) {
            const uint64_t first = k * bbytes;
            if (first >= nbytes) continue;
            const uint64_t nb = k + 1 == nblk ? nbytes - first : std::min(bbytes, nbytes - first);
            sieve_block(g, base + 30 * first, nb, lo, hi, lo, seg, carry, wheel, parts[k]);
        }
    });
    // each block's primes go to their place in the output, every thread
    // copying its share of the blocks
    std::vector<int64_t> at(nblk + 1);
    at[0] = (int64_t)small.size();
    for (uint64_t k = 0; k < nblk; ++k) at[k + 1] = at[k] + (int64_t)parts[k].n;
    *count = at[nblk];
    if (at[nblk] > cap) return RS_ERR_OVERFLOW;
    for (size_t i = 0; i < small.size(); ++i) offsets[i] = small[i];
    run([&](int t) {
        for (uint64_t k = (uint64_t)t; k < nblk; k += (uint64_t)nt)
            if (parts[k].n) memcpy(offsets + at[k], parts[k].p.get(), parts[k].n * sizeof(uint32_t));
    });
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

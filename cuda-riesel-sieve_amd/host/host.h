// host.h -- internal types of the host side of librsieve (ABCD input,
// planner, prime generator).  The C ABI over them is include/rsieve.h.
#pragma once
#include <cstdint>
#include <vector>

#include "rsieve.h"

struct rs_abcd {
    uint32_t base = 0;
    int c = -1;                   // k*b^n + c: -1 Riesel, +1 Sierpinski
    std::vector<int32_t> ks;      // one entry per header line
    std::vector<int32_t> kstart;  // CSR offsets into ns, size ks.size() + 1
    std::vector<int32_t> ns;      // strictly increasing per k
    int minN = 0, maxN = 0;
};

struct rs_plan {
    uint32_t base = 0;
    int c = -1;                   // form of the candidates, as in rs_abcd
    int Q = 0;
    int minN = 0, maxN = 0;
    int x0 = 0, x1 = 0;           // floor(minN/Q), floor(maxN/Q)
    int words = 0;                // 32-bit words per row bitmap
    std::vector<int32_t> rows_k;  // sorted by (k, r)
    std::vector<int32_t> rows_r;
    std::vector<uint32_t> bitmaps;  // rows * words, MSB-first
};

namespace rsh {

// 64-bit helpers shared by the host code.
typedef unsigned __int128 u128;

inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t p) {
    return (uint64_t)((u128)a * b % p);
}
uint64_t powmod(uint64_t b, uint64_t e, uint64_t p);
bool is_prime_u64(uint64_t n);  // deterministic Miller-Rabin
std::vector<uint32_t> primes_upto(uint32_t n);  // all primes <= n

// splitmix64 stream used for seeded draws.
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    uint64_t below(uint64_t n) { return next() % n; }
};

int hw_threads(int requested);

// Job-shape limits of the device kernels (the giant-step grid of a row
// slice is indexed in 16 bits, and a row's Y offset 8 r ys in 19 bits).
// giant_T is the giant steps per row for table size m and tile tw (1, 2 or
// 4): ceil(X/m) for the x-window X, rounded up to a multiple of tw, at
// least 2.  shape_ok is true iff a job of this plan with (m, tw) is within
// the limits: m a multiple of 64 in [64, 65536], T <= 65535, Q ys < 65536
// (ys = T, or T + 1 when T is even), min(rows, 512) T + 4096 < 65536.
int giant_T(const rs_plan* pl, int m, int tw);
bool shape_ok(const rs_plan* pl, int m, int tw);

}  // namespace rsh

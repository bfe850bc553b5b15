// util.cpp -- host arithmetic helpers, factor verification and output
// formatting.
#include <sched.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <thread>

#include "host.h"

namespace rsh {

uint64_t powmod(uint64_t b, uint64_t e, uint64_t p) {
    uint64_t r = 1 % p, x = b % p;
    while (e) {
        if (e & 1) r = mulmod(r, x, p);
        x = mulmod(x, x, p);
        e >>= 1;
    }
    return r;
}

// Deterministic Miller-Rabin for 64-bit n with the 7-base set of Jim
// Sinclair (2, 325, 9375, 28178, 450775, 9780504, 1795265022).
bool is_prime_u64(uint64_t n) {
    if (n < 2) return false;
    static const uint32_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (uint32_t q : small) {
        if (n == q) return true;
        if (n % q == 0) return false;
    }
    uint64_t d = n - 1;
    int s = 0;
    while ((d & 1) == 0) { d >>= 1; ++s; }
    static const uint64_t bases[] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
    for (uint64_t a : bases) {
        a %= n;
        if (a == 0) continue;
        uint64_t x = powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int i = 1; i < s && comp; ++i) {
            x = mulmod(x, x, n);
            if (x == n - 1) comp = false;
        }
        if (comp) return false;
    }
    return true;
}

std::vector<uint32_t> primes_upto(uint32_t n) {
    std::vector<uint32_t> out;
    if (n < 2) return out;
    std::vector<char> comp((size_t)n + 1, 0);
    for (uint64_t i = 2; i <= n; ++i) {
        if (comp[i]) continue;
        out.push_back((uint32_t)i);
        for (uint64_t j = i * i; j <= n; j += i) comp[j] = 1;
    }
    return out;
}

// requested > 0, else the CPUs this process may run on (its affinity set),
// at most OMP_NUM_THREADS when that is set (a host's share of its cores),
// at most 64.
int hw_threads(int requested) {
    if (requested > 0) return requested;
    int n = (int)std::thread::hardware_concurrency();
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof set, &set) == 0) n = CPU_COUNT(&set);
    if (const char* e = getenv("OMP_NUM_THREADS")) {
        const int v = atoi(e);
        if (v > 0) n = std::min(n, v);
    }
    return std::max(1, std::min(n, 64));
}

}  // namespace rsh

// kernel_v0.24.6.cu:1656-1671: b^n mod p by square-and-multiply in 128-bit,
// then k*b^n == 1 (mod p).
extern "C" int rs_verify_factor(uint64_t p, uint32_t k, uint32_t b, uint32_t n, int c) {
    if (p < 2 || (c != -1 && c != 1)) return 0;
    const uint64_t want = c == 1 ? p - 1 : 1 % p;  // k*b^n == -c (mod p)
    return rsh::mulmod(rsh::powmod(b, n, p), k % p, p) == want ? 1 : 0;
}

// writeFactorToFile's line, kernel_v0.24.6.cu:1833: "%llu | %d*%d^%d-1\n"
// (with "+1" for the Sierpinski form).
extern "C" int rs_format_factor(const rs_factor* f, uint32_t b, int c, char* buf, int cap) {
    if (!f || !buf || cap <= 0 || (c != -1 && c != 1)) return RS_ERR_ARG;
    int n = snprintf(buf, (size_t)cap, "%llu | %u*%u^%u%s\n", (unsigned long long)f->p,
                     f->k, b, f->n, c == 1 ? "+1" : "-1");
    if (n < 0) return RS_ERR_ARG;
    if (n >= cap) return RS_ERR_OVERFLOW;
    return n;
}

extern "C" const char* rs_strerror(rs_status s) {
    switch (s) {
        case RS_OK: return "ok";
        case RS_ERR_ARG: return "bad argument";
        case RS_ERR_HIP: return "HIP runtime error";
        case RS_ERR_NOMEM: return "out of memory";
        case RS_ERR_IO: return "file cannot be opened or written";
        case RS_ERR_FORMAT: return "malformed ABCD input";
        case RS_ERR_OVERFLOW: return "buffer too small";
        case RS_ERR_BUDGET: return "job does not fit the memory budget";
        case RS_ERR_STATE: return "call out of order";
        case RS_ERR_RANGE: return "too few primes in the window";
        default: return "unknown status";
    }
}

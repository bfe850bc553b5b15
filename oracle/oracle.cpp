// oracle.cpp -- CPU restatement of the reference Riesel BSGS sieve.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
// this file: it is used by tests/, by __graft_entry__.smoke() to compare its
// results, and by bench.py's cpu_baseline leg.  It restates the algorithm of
// kernel_v0.24.6.cu (rob147147/CUDA-Riesel-Sieve v0.24.6) in plain C++ with
// exact 128-bit arithmetic, applying these decisions on top of it:
//   * exact REDC instead of the PTX montmul whose limb-1 carry can be lost
//     (kernel_v0.24.6.cu:226);
//   * every hit is reported, not only the last one per prime (:615-621);
//   * the x-window [floor(minN/Q), floor(maxN/Q)] is covered in full
//     (the reference starts at tMin*m+1, :534/:578/:1155);
//   * Q is always even (the reference may fall back to Q=1, :1093/:1133,
//     while its Jacobi filter at :521 assumes even Q);
//   * rows are sorted lexicographically by (k, r) (:1220-1247);
//   * minN/maxN are taken over every n in the file (the reference updates
//     maxN only from delta lines, :1049-1054).
// Each function names the reference lines it follows.  What pins each part
// is listed in tests/test_oracle.py.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <unordered_map>
#include <fstream>
#include <climits>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;
typedef uint64_t u64;
typedef uint32_t u32;

extern "C" {

// ---------------------------------------------------------------- arithmetic

// (x*2^64 + y) mod z.  kernel_v0.24.6.cu:150-179 computes the same value by
// 64 shift-subtract steps (needs x < z < 2^63); here one 128-bit division.
u64 or_modul64(u64 x, u64 y, u64 z) {
    u128 v = ((u128)x << 64) | y;
    return (u64)(v % z);
}

// -p^-1 mod 2^64 for odd p.  kernel_v0.24.6.cu:92-111 (xbinGCDnew) gets it by
// a 64-step binary extended GCD; here by Newton iteration x <- x(2 - p x).
u64 or_mprime(u64 p) {
    u64 x = p;                      // p*p == 1 mod 8: 3 correct bits
    for (int i = 0; i < 5; ++i) x *= 2 - p * x;
    return (u64)0 - x;
}

// Exact Montgomery product a*b*2^-64 mod p (R = 2^64), p odd < 2^63,
// a < 2^64, b < p.  Replaces the PTX limb sequence of
// kernel_v0.24.6.cu:182-270 (which drops a carry at :226).
u64 or_montmul(u64 a, u64 b, u64 p, u64 mp) {
    u128 t = (u128)a * b;
    u64 m = (u64)t * mp;
    u128 u = (t >> 64) + (((u128)m * p + (u64)t) >> 64);
    // (u64)t + lo(m p) is 0 mod 2^64; the carry is folded into the shift above
    u64 r = (u64)u;
    if (r >= p) r -= p;
    return r;
}

// Jacobi symbol (a/n), n odd > 0: returns -1, 0 or +1.  The reference's
// legendre() (kernel_v0.24.6.cu:55-89, taken from sr2sieve) computes the same
// sign by reciprocity and returns 1 for +1 and 0 for -1; it has no gcd test,
// which only matters when p | a.  Restated with the gcd test.
int or_jacobi(u64 a, u64 n) {
    a %= n;
    int t = 1;
    while (a != 0) {
        while ((a & 1) == 0) {
            a >>= 1;
            u64 r = n & 7;
            if (r == 3 || r == 5) t = -t;
        }
        u64 tmp = a; a = n; n = tmp;
        if ((a & 3) == 3 && (n & 3) == 3) t = -t;
        a %= n;
    }
    return n == 1 ? t : 0;
}

// b^e mod p with plain 128-bit products (not Montgomery): the independent
// check the reference applies to each factor (kernel_v0.24.6.cu:1656-1671).
u64 or_powmod(u64 b, u64 e, u64 p) {
    u128 r = 1 % p, x = b % p;
    while (e) {
        if (e & 1) r = (r * x) % p;
        x = (x * x) % p;
        e >>= 1;
    }
    return (u64)r;
}

// 1 if p | k*b^n + c, c = -1 (Riesel, kernel_v0.24.6.cu:1656-1671) or +1
// (Sierpinski).
int or_verify(u64 p, u64 k, u64 b, u64 n, int c) {
    u128 v = (u128)or_powmod(b, n, p) * (k % p) % p;
    const u64 want = c == 1 ? (p - 1) % p : 1 % p;  // k*b^n == -c
    return v == want ? 1 : 0;
}

// ---------------------------------------------------------------- ABCD input

// Reads an ABCD file the way kernel_v0.24.6.cu:113-137 (countKs, also
// FileInput.cpp:12-36) and :938-1083 do: a line whose first space-separated
// token is "ABCD" starts a new k ("ABCD k*b^$a-1 [n0]"), the base comes from
// the first header only, and every other line is a positive delta added to
// the previous n.  Outputs (caller-sized by or_abcd_count): ks[nk],
// kstart[nk+1] (CSR offsets into ns), ns[].  Returns 0, or -1 if the file
// cannot be opened, -2 on a malformed line.
int or_abcd_count(const char* path, int* nk, int* nlines) {
    std::ifstream f(path);
    if (!f.is_open()) return -1;
    std::string line;
    int k = 0, l = 0;
    while (std::getline(f, line)) {
        ++l;
        std::string tok = line.substr(0, line.find(' '));
        if (tok == "ABCD") ++k;
    }
    *nk = k; *nlines = l;
    return 0;
}

int or_abcd_read(const char* path, u32* base, int* ks, int* kstart, int* ns,
                 int* minN, int* maxN) {
    std::ifstream f(path);
    if (!f.is_open()) return -1;
    std::string line;
    int nk = 0, nn = 0, total = 0;
    int b = 0;
    int lo = INT_MAX, hi = 0;
    bool have = false;
    while (std::getline(f, line)) {
        size_t sp = line.find(' ');
        std::string tok = line.substr(0, sp);
        if (tok == "ABCD") {
            std::string rest = line.substr(sp + 1);
            size_t st = rest.find('*');
            size_t car = rest.find('^');
            size_t lb = rest.find('[');
            size_t rb = rest.find(']');
            if (st == std::string::npos || car == std::string::npos ||
                lb == std::string::npos || rb == std::string::npos) return -2;
            int kv = atoi(rest.substr(0, st).c_str());
            if (b == 0) b = atoi(rest.substr(st + 1, car - st - 1).c_str());
            total = atoi(rest.substr(lb + 1, rb - lb - 1).c_str());
            kstart[nk] = nn;
            ks[nk++] = kv;
            ns[nn++] = total;
            have = true;
        } else {
            if (!have) return -2;
            int d = atoi(tok.c_str());
            if (d <= 0) return -2;
            total += d;
            ns[nn++] = total;
        }
        if (total < lo) lo = total;
        if (total > hi) hi = total;
    }
    kstart[nk] = nn;
    *base = (u32)b;
    *minN = lo;
    *maxN = hi;
    return 0;
}

// ---------------------------------------------------------------- planner

// Subsequence planner, kernel_v0.24.6.cu:1086-1255.  n = Q*x + r; Q is the
// even value in [2, 24) (or exactly Qin) with the least subs*R work, where
// R = maxN/Q - minN/Q + 1.  The reference keeps Q = 1 when no even Q beats
// K*(maxN-minN); this restatement then takes the least-work even Q instead.
// Rows are the distinct (k, r), sorted by (k, r); bitmap bit for
// x' = (n-r)/Q - minN/Q is word x'/32, bit 31-(x'%32) (MSB first).
// Outputs: *Q, *nrows, *words (32-bit words per row bitmap); rows_k/rows_r
// (cap rows); bitmaps (cap*words).  Call with cap=0 to size.
int or_plan(int nk, const int* ks, const int* kstart, const int* ns,
            int minN, int maxN, int Qin, int* Qout, int* nrows, int* words,
            int cap, int* rows_k, int* rows_r, u32* bitmaps) {
    long long minWork = (long long)nk * (maxN - minN);
    int minQ = 1;
    int qlo = 2, qhi = 24;
    if (Qin != 0) { qlo = Qin; qhi = Qin + 1; }
    long long bestEven = LLONG_MAX; int bestEvenQ = 0;
    for (int Q = qlo; Q < qhi; Q += 2) {
        std::vector<char> seen((size_t)nk * Q, 0);
        long long subs = 0;
        for (int i = 0; i < nk; ++i)
            for (int j = kstart[i]; j < kstart[i + 1]; ++j) {
                char& s = seen[(size_t)i * Q + ns[j] % Q];
                if (!s) { s = 1; ++subs; }
            }
        long long R = maxN / Q - minN / Q + 1;
        if (subs * R < minWork) { minWork = subs * R; minQ = Q; }
        if (subs * R < bestEven) { bestEven = subs * R; bestEvenQ = Q; }
    }
    if (minQ == 1) minQ = bestEvenQ;   // even-Q decision (see header)
    int Q = minQ;
    int x0 = minN / Q;
    int range = maxN / Q - x0 + 1;
    int w = range / 32 + 1;            // qRange of :1156 (one spare word)
    // distinct (k, r) sorted; k may repeat in the file, rows merge them
    std::vector<std::pair<int, int>> rows;
    for (int i = 0; i < nk; ++i)
        for (int j = kstart[i]; j < kstart[i + 1]; ++j)
            rows.push_back({ks[i], ns[j] % Q});
    std::sort(rows.begin(), rows.end());
    rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
    *Qout = Q; *nrows = (int)rows.size(); *words = w;
    if (cap == 0) return 0;
    if (cap < (int)rows.size()) return -3;
    for (size_t i = 0; i < rows.size(); ++i) {
        rows_k[i] = rows[i].first;
        rows_r[i] = rows[i].second;
    }
    memset(bitmaps, 0, sizeof(u32) * rows.size() * w);
    for (int i = 0; i < nk; ++i)
        for (int j = kstart[i]; j < kstart[i + 1]; ++j) {
            int n = ns[j], r = n % Q;
            auto it = std::lower_bound(rows.begin(), rows.end(),
                                       std::make_pair(ks[i], r));
            size_t row = it - rows.begin();
            int xp = (n - r) / Q - x0;
            bitmaps[row * w + xp / 32] |= 1u << (31 - (xp & 31));
        }
    return 0;
}

// tMin/tMax for m = 2^shift, kernel_v0.24.6.cu:1148-1158.
void or_tminmax(int minN, int maxN, int Q, int m, int* tmin, int* tmax) {
    int shift = 0;
    while ((1 << (shift + 1)) <= m) ++shift;
    *tmin = (minN / Q) >> shift;
    *tmax = (((maxN / Q) + 1) >> shift) + 1;
}

// ---------------------------------------------------------------- BSGS

struct OrHit { u64 p; u32 k; u32 n; };

struct OrJob {
    u32 base;
    int Q;
    int m;            // baby steps (any m >= 1; the reference uses 2^s)
    int minN, maxN;
    int nrows;
    const int* rows_k;   // sorted by (k, r)
    const int* rows_r;
    int c;               // -1: k*b^n-1 (Riesel), +1: k*b^n+1 (Sierpinski)
};

static bool hit_less(const OrHit& a, const OrHit& b) {
    if (a.p != b.p) return a.p < b.p;
    if (a.k != b.k) return a.k < b.k;
    return a.n < b.n;
}

// Per-prime BSGS, restating addKernel1 (kernel_v0.24.6.cu:273-684) with the
// decisions in the header.  Residues are in Montgomery form (R = 2^64):
//   oneMS = R mod p (:305), KB = b*R mod p (:343), g = b^Q (:350-354),
//   baby table {g^j : 0 <= j < m} (:358-411; here an exact multimap so
//   duplicate residues when ord(g) < m keep every j),
//   c1 = g^m (:428-432), beta = c1^tLo (:454-456),
//   per row (k, r) in sorted order: on a new k the Jacobi pair
//   J1 = (k/p), J2 = (kb/p) (:491-509) and fixedBeta = k*beta; a row is
//   searched iff (r even and J1 = +1) or (r odd and J2 = +1) (:516-521);
//   fixedBeta advances by b^(r - prevRem) (:525-528); giant steps
//   t in [tLo, tHi): sB *= c1, every j with g^j == sB gives
//   x = (t+1)m - j, n = Q x + r (:534-621).
// tLo = floor((x0-1)/m) (0 if x0 == 0) and tHi = ceil(x1/m) cover x in
// [x0, x1]; x = 0 (only when minN < Q) is tested directly.  Hits with n
// outside [minN, maxN] are dropped (:1606).  For c = +1 (k*b^n == -1) k is
// negated before the product with beta, the one-line change :505-506 notes,
// and the filter compares the Jacobi symbols with (-1/p) instead of +1
// (-k*b^r must be a square; the reference's unchanged filter would skip
// solvable rows for p = 3 mod 4).  Precondition: p odd, p < 2^63, p does
// not divide b.  Returns the number of hits (all written if <= cap).
int or_bsgs_prime(const OrJob* J, u64 p, OrHit* out, int cap,
                  long long* montmuls) {
    long long mm = 0;
    const u64 mp = or_mprime(p);
    const u64 one = or_modul64(1, 0, p);
    const u64 KB = or_modul64(J->base % p, 0, p);
    u64 g = one;
    for (int q = 0; q < J->Q; ++q) { g = or_montmul(KB, g, p, mp); ++mm; }
    std::unordered_multimap<u64, int> table;
    table.reserve((size_t)J->m * 2);
    u64 js = one;
    for (int j = 0; j < J->m; ++j) {
        table.emplace(js, j);
        js = or_montmul(js, g, p, mp); ++mm;
    }
    // c1 = g^m by square-and-multiply (the reference squares log2 m times)
    u64 c1 = one, base = g;
    for (int e = J->m; e; e >>= 1) {
        if (e & 1) { c1 = or_montmul(c1, base, p, mp); ++mm; }
        base = or_montmul(base, base, p, mp); ++mm;
    }
    const int Q = J->Q, m = J->m;
    const int x0 = J->minN / Q, x1 = J->maxN / Q;
    const int tLo = x0 >= 1 ? (x0 - 1) / m : 0;
    const int tHi = (x1 + m - 1) / m;
    u64 beta = one;
    for (int t = 0; t < tLo; ++t) { beta = or_montmul(beta, c1, p, mp); ++mm; }
    const int jb = or_jacobi(J->base, p);
    const int want = J->c == 1 ? ((p & 3) == 1 ? 1 : -1) : 1;  // (-1/p) for +1
    int nh = 0;
    int lastk = -1, prevRem = 0;
    bool j1 = false, j2 = false;
    u64 fixedBeta = 0, kM = 0;
    for (int row = 0; row < J->nrows; ++row) {
        const int k = J->rows_k[row], r = J->rows_r[row];
        if (k != lastk) {
            lastk = k;
            int jk = or_jacobi((u64)k, p);
            j1 = jk == want;
            j2 = jk * jb == want;
            kM = or_modul64((u64)k % p, 0, p);
            if (J->c == 1 && kM != 0) kM = p - kM;
            fixedBeta = or_montmul(kM, beta, p, mp); ++mm;
            prevRem = 0;
        }
        if (!((r % 2 == 0 && j1) || (r % 2 == 1 && j2))) continue;
        for (int rem = prevRem; rem < r; ++rem) {
            fixedBeta = or_montmul(fixedBeta, KB, p, mp); ++mm;
        }
        prevRem = r;
        if (x0 == 0) {
            // x = 0: k*b^r == 1 (fixedBeta is k*b^r*c1^0 here since tLo == 0)
            if (fixedBeta == one) {
                int n = r;
                if (n >= J->minN && n <= J->maxN) {
                    if (nh < cap) out[nh] = {p, (u32)k, (u32)n};
                    ++nh;
                }
            }
        }
        u64 sB = fixedBeta;
        for (int t = tLo; t < tHi; ++t) {
            sB = or_montmul(sB, c1, p, mp); ++mm;
            auto range = table.equal_range(sB);
            for (auto it = range.first; it != range.second; ++it) {
                long long x = (long long)(t + 1) * m - it->second;
                long long n = (long long)Q * x + r;
                if (n >= J->minN && n <= J->maxN) {
                    if (nh < cap) out[nh] = {p, (u32)k, (u32)n};
                    ++nh;
                }
            }
        }
    }
    if (nh <= cap) std::sort(out, out + nh, hit_less);
    if (montmuls) *montmuls = mm;
    return nh;
}

// Exhaustive check: every row, every x in [x0, x1], plain 128-bit products,
// no Jacobi filter and no table.  Defines the complete datless factor set
// {(p,k,n) : n in [minN,maxN], (k, n mod Q) a row, k*b^n == -c mod p}.
int or_brute_prime(const OrJob* J, u64 p, OrHit* out, int cap) {
    const int Q = J->Q;
    const u64 want = J->c == 1 ? (p - 1) % p : 1 % p;
    const int x0 = J->minN / Q, x1 = J->maxN / Q;
    const u64 bQ = or_powmod(J->base, Q, p);
    int nh = 0;
    for (int row = 0; row < J->nrows; ++row) {
        const u64 k = (u64)J->rows_k[row];
        const int r = J->rows_r[row];
        u64 v = (u64)((u128)(k % p) * or_powmod(J->base, (u64)Q * x0 + r, p) % p);
        for (int x = x0; x <= x1; ++x) {
            long long n = (long long)Q * x + r;
            if (v == want && n >= J->minN && n <= J->maxN) {
                if (nh < cap) out[nh] = {p, (u32)k, (u32)n};
                ++nh;
            }
            v = (u64)((u128)v * bQ % p);
        }
    }
    if (nh <= cap) std::sort(out, out + nh, hit_less);
    return nh;
}

// Sieve many primes p = low + offs[i] with the BSGS restatement, OpenMP over
// primes (threads <= 0: library default).  Hits are sorted by (p, k, n).
// Returns the hit count (all written if <= cap); *montmuls (optional) gets
// the total Montgomery products.
long long or_bsgs_many(const OrJob* J, u64 low, const u32* offs, long long count,
                       OrHit* out, long long cap, int threads,
                       long long* montmuls) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    std::vector<OrHit> all;
    long long mmtot = 0;
#pragma omp parallel
    {
        std::vector<OrHit> mine;
        OrHit buf[256];
        long long mmloc = 0;
#pragma omp for schedule(dynamic, 1024)
        for (long long i = 0; i < count; ++i) {
            long long mm = 0;
            u64 p = low + offs[i];
            int h = or_bsgs_prime(J, p, buf, 256, &mm);
            mmloc += mm;
            if (h <= 256) {
                mine.insert(mine.end(), buf, buf + h);
            } else {
                std::vector<OrHit> big(h);
                or_bsgs_prime(J, p, big.data(), h, nullptr);
                mine.insert(mine.end(), big.begin(), big.end());
            }
        }
#pragma omp critical
        {
            all.insert(all.end(), mine.begin(), mine.end());
            mmtot += mmloc;
        }
    }
    std::sort(all.begin(), all.end(), hit_less);
    long long n = (long long)all.size();
    for (long long i = 0; i < n && i < cap; ++i) out[i] = all[i];
    if (montmuls) *montmuls = mmtot;
    return n;
}

// ---------------------------------------------------------------- primes

// The reference prime-array generator, kernel_v0.24.6.cu:1271-1364 and
// :1838-1987: candidates low + 2i (low odd) in a window of testArraySize
// bits, marked by 8 wheel masks over the odd primes 3..107 (grouped as at
// :1329-1336, period = product of the group in 64-bit words, start word
// (low mod M) * 128^-1 mod M, :1844-1853) and then by the remaining small
// primes (index 27 on, i.e. from 109) of the first primeCount odd primes
// (:1890-1933).  The first arraySize unmarked positions give offsets 2i
// (:1949-1973).  A partial sieve: survivors are composite once
// low + offset >= (largest small prime)^2.  Returns the number of offsets
// written (< arraySize means the window ran out, the reference exits(1)).
static const int kGroups[8][4] = {{19, 13, 67, 11}, {103, 23, 17, 5},
                                  {107, 97, 3, 7},  {43, 41, 101, 0},
                                  {83, 73, 29, 0},  {61, 47, 59, 0},
                                  {31, 71, 79, 0},  {37, 89, 53, 0}};

// The first count odd primes (3, 5, 7, ...).
int or_small_primes(int count, u32* out) {
    long long lim = 64;
    while (true) {
        std::vector<char> comp(lim + 1, 0);
        int c = 0;
        for (long long i = 3; i <= lim && c < count; i += 2) {
            if (comp[i]) continue;
            out[c++] = (u32)i;
            for (long long j = i * i; j <= lim; j += 2 * i) comp[j] = 1;
        }
        if (c == count) return c;
        lim *= 2;
    }
}

// Inverse of 128 modulo an odd M (the reference's magicMultipliers, :1843).
long long or_inv128(long long M) {
    // extended Euclid on (128 mod M, M)
    long long r0 = 128 % M, r1 = M, s0 = 1, s1 = 0;
    while (r1) {
        long long q = r0 / r1;
        long long t = r0 - q * r1; r0 = r1; r1 = t;
        t = s0 - q * s1; s0 = s1; s1 = t;
    }
    long long inv = s0 % M;
    if (inv < 0) inv += M;
    return inv;
}

int or_ref_primes(u64 low, int smallPrimeScale, int arraySize,
                  int testArraySize, u32* out) {
    if (low % 2 == 0) low -= 1;                         // :1515-1519
    const int primeCount = 4096 * smallPrimeScale;      // :1274
    std::vector<u32> sp(primeCount);
    or_small_primes(primeCount, sp.data());
    const int words = testArraySize / 64 + 1;
    std::vector<u64> mark(words, 0);
    for (int g = 0; g < 8; ++g) {
        long long M = 1;
        for (int i = 0; i < 4 && kGroups[g][i]; ++i) M *= kGroups[g][i];
        // mask bit k (MSB-first in 64-bit words) set iff a group prime | k
        std::vector<u64> mask(M, 0);
        for (int i = 0; i < 4 && kGroups[g][i]; ++i) {
            long long q = kGroups[g][i];
            for (long long k = 0; k < 64 * M; k += q)
                mask[k >> 6] |= 0x8000000000000000ull >> (k & 63);
        }
        long long off = (long long)((u128)(low % M) * or_inv128(M) % M);
        for (int w = 0; w < words; ++w) {
            mark[w] |= mask[off];
            if (++off >= M) off -= M;
        }
    }
    for (int i = 27; i < primeCount; ++i) {
        u32 q = sp[i];
        u32 mod = (u32)(low % q);
        u32 diff = 0;
        if (mod != 0) {
            if (mod & 1) mod += q;
            mod >>= 1;
            diff = q - mod;
        }
        for (long long k = diff; k < testArraySize; k += q)
            mark[k >> 6] |= 0x8000000000000000ull >> (k & 63);
    }
    int c = 0;
    for (int i = 0; i < testArraySize / 64 && c < arraySize; ++i)
        for (int j = 0; j < 64 && c < arraySize; ++j)
            if (((mark[i] >> (63 - j)) & 1) == 0) out[c++] = 2u * (u32)(i * 64 + j);
    return c;
}

// testArraySize scale factor by -s, kernel_v0.24.6.cu:853-878.
int or_test_array_factor(int s) {
    if (s <= 4) return 12;
    if (s <= 8) return 13;
    if (s <= 16) return 14;
    if (s <= 32) return 16;
    if (s <= 64) return 18;
    if (s <= 128) return 20;
    if (s <= 256) return 22;
    return 24;
}

// All primes p in [lo, hi) (true primes, plain segmented Eratosthenes),
// written as offsets p - lo.  Returns the count (all written if <= cap).
long long or_true_primes(u64 lo, u64 hi, u32* out, long long cap) {
    if (hi <= lo) return 0;
    u64 r = 1;
    while ((u128)r * r < hi) ++r;
    std::vector<char> comp(r + 1, 0);
    std::vector<u64> ps;
    for (u64 i = 2; i <= r; ++i) {
        if (comp[i]) continue;
        ps.push_back(i);
        for (u64 j = i * i; j <= r; j += i) comp[j] = 1;
    }
    long long c = 0;
    const u64 SEG = 1 << 22;
    std::vector<char> seg(SEG);
    for (u64 a = lo; a < hi; a += SEG) {
        u64 b = std::min(hi, a + SEG);
        std::fill(seg.begin(), seg.begin() + (b - a), 0);
        for (u64 q : ps) {
            if (q * q >= b) break;
            u64 s = std::max(q * q, (a + q - 1) / q * q);
            for (u64 j = s; j < b; j += q) seg[j - a] = 1;
        }
        for (u64 v = a; v < b; ++v)
            if (v >= 2 && !seg[v - a]) {
                if (c < cap) out[c] = (u32)(v - lo);
                ++c;
            }
    }
    return c;
}

}  // extern "C"

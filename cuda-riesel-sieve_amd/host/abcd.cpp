// abcd.cpp -- ABCD sieve-file input/output and the synthetic input
// generator.
//
// Reading follows countKs (kernel_v0.24.6.cu:113-137, FileInput.cpp:12-36)
// and the parser in main (kernel_v0.24.6.cu:938-1083): the token before the
// first space decides the line type; "ABCD k*b^$a-1 [n0]" starts a k, the
// base is taken from the first header only (:1009-1016), and every other
// line is a positive delta added to the previous n (:1049-1054).  Two
// deliberate differences: malformed lines are an error instead of a stoi
// exception or silent garbage, and minN/maxN cover every n (the reference
// raises maxN only on delta lines).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>

#include "host.h"

namespace {

// Parses a non-negative decimal integer at s (leading blanks allowed);
// returns false if there is no digit or it does not fit in int32.
bool parse_int(const char* s, long long* out) {
    while (*s == ' ' || *s == '\t') ++s;
    if (!isdigit((unsigned char)*s)) return false;
    long long v = 0;
    while (isdigit((unsigned char)*s)) {
        v = v * 10 + (*s - '0');
        if (v > 0x7fffffffLL) return false;
        ++s;
    }
    *out = v;
    return true;
}

void finish(rs_abcd* a) {
    a->kstart.push_back((int32_t)a->ns.size());
    int lo = 0x7fffffff, hi = 0;
    for (int32_t n : a->ns) {
        lo = std::min(lo, (int)n);
        hi = std::max(hi, (int)n);
    }
    a->minN = a->ns.empty() ? 0 : lo;
    a->maxN = a->ns.empty() ? 0 : hi;
}

}  // namespace

extern "C" rs_status rs_abcd_read(const char* path, rs_abcd** out) {
    if (!path || !out) return RS_ERR_ARG;
    *out = nullptr;
    std::ifstream f(path);
    if (!f.is_open()) return RS_ERR_IO;
    rs_abcd* a = new rs_abcd();
    std::string line;
    long long total = 0;
    bool have = false;
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        size_t sp = line.find(' ');
        std::string tok = line.substr(0, sp);
        if (tok == "ABCD") {
            if (sp == std::string::npos) { delete a; return RS_ERR_FORMAT; }
            const std::string rest = line.substr(sp + 1);
            size_t st = rest.find('*'), car = rest.find('^');
            size_t lb = rest.find('['), rb = rest.find(']');
            if (st == std::string::npos || car == std::string::npos || car < st ||
                lb == std::string::npos || rb == std::string::npos || rb < lb) {
                delete a;
                return RS_ERR_FORMAT;
            }
            // "...^$a-1" (Riesel, the reference's form) or "...^$a+1"
            // (Sierpinski); one form per file
            const size_t mi = rest.find("$a-1", car), pl = rest.find("$a+1", car);
            int c = 0;
            if (mi != std::string::npos && mi < lb) c = -1;
            if (pl != std::string::npos && pl < lb) c = c ? 0 : 1;
            if (c == 0 || (have && c != a->c)) {
                delete a;
                return RS_ERR_FORMAT;
            }
            a->c = c;
            long long k, b, n0;
            if (!parse_int(rest.substr(0, st).c_str(), &k) || k < 1 ||
                !parse_int(rest.substr(st + 1, car - st - 1).c_str(), &b) || b < 2 ||
                !parse_int(rest.substr(lb + 1, rb - lb - 1).c_str(), &n0)) {
                delete a;
                return RS_ERR_FORMAT;
            }
            if (a->base == 0) a->base = (uint32_t)b;
            a->kstart.push_back((int32_t)a->ns.size());
            a->ks.push_back((int32_t)k);
            a->ns.push_back((int32_t)n0);
            total = n0;
            have = true;
        } else {
            long long d;
            if (!have || !parse_int(tok.c_str(), &d) || d <= 0 ||
                total + d > 0x7fffffffLL) {
                delete a;
                return RS_ERR_FORMAT;
            }
            total += d;
            a->ns.push_back((int32_t)total);
        }
    }
    finish(a);
    *out = a;
    return RS_OK;
}

extern "C" rs_status rs_abcd_from_lists(uint32_t base, int nk, const int32_t* ks,
                                        const int32_t* kstart, const int32_t* ns,
                                        rs_abcd** out) {
    if (!out || nk < 0 || base < 2 || (nk > 0 && (!ks || !kstart || !ns)))
        return RS_ERR_ARG;
    *out = nullptr;
    rs_abcd* a = new rs_abcd();
    a->base = base;
    for (int i = 0; i < nk; ++i) {
        if (ks[i] < 1 || kstart[i + 1] <= kstart[i]) { delete a; return RS_ERR_ARG; }
        a->kstart.push_back((int32_t)a->ns.size());
        a->ks.push_back(ks[i]);
        for (int j = kstart[i]; j < kstart[i + 1]; ++j) {
            if (ns[j] < 0 || (j > kstart[i] && ns[j] <= ns[j - 1])) {
                delete a;
                return RS_ERR_ARG;
            }
            a->ns.push_back(ns[j]);
        }
    }
    finish(a);
    *out = a;
    return RS_OK;
}

extern "C" rs_status rs_abcd_info(const rs_abcd* a, uint32_t* base, int* nk, int* nn,
                                  int* minN, int* maxN) {
    if (!a) return RS_ERR_ARG;
    if (base) *base = a->base;
    if (nk) *nk = (int)a->ks.size();
    if (nn) *nn = (int)a->ns.size();
    if (minN) *minN = a->minN;
    if (maxN) *maxN = a->maxN;
    return RS_OK;
}

extern "C" rs_status rs_abcd_form(const rs_abcd* a, int* c) {
    if (!a || !c) return RS_ERR_ARG;
    *c = a->c;
    return RS_OK;
}

extern "C" rs_status rs_abcd_set_form(rs_abcd* a, int c) {
    if (!a || (c != -1 && c != 1)) return RS_ERR_ARG;
    a->c = c;
    return RS_OK;
}

extern "C" rs_status rs_abcd_lists(const rs_abcd* a, int32_t* ks, int32_t* kstart,
                                   int32_t* ns) {
    if (!a) return RS_ERR_ARG;
    if (ks) std::copy(a->ks.begin(), a->ks.end(), ks);
    if (kstart) std::copy(a->kstart.begin(), a->kstart.end(), kstart);
    if (ns) std::copy(a->ns.begin(), a->ns.end(), ns);
    return RS_OK;
}

extern "C" rs_status rs_abcd_write(const rs_abcd* a, const char* path) {
    if (!a || !path) return RS_ERR_ARG;
    FILE* f = fopen(path, "w");
    if (!f) return RS_ERR_IO;
    for (size_t i = 0; i < a->ks.size(); ++i) {
        int s = a->kstart[i], e = a->kstart[i + 1];
        fprintf(f, "ABCD %d*%u^$a%s [%d]\n", a->ks[i], a->base, a->c == 1 ? "+1" : "-1",
                a->ns[s]);
        for (int j = s + 1; j < e; ++j) fprintf(f, "%d\n", a->ns[j] - a->ns[j - 1]);
    }
    return fclose(f) == 0 ? RS_OK : RS_ERR_IO;
}

extern "C" rs_status rs_abcd_remove(rs_abcd* a, const rs_factor* factors, int64_t count,
                                   int64_t* removed) {
    if (!a || count < 0 || (count > 0 && !factors)) return RS_ERR_ARG;
    // n to remove per k, sorted
    std::unordered_map<uint32_t, std::vector<int32_t>> drop;
    for (int64_t i = 0; i < count; ++i) drop[factors[i].k].push_back((int32_t)factors[i].n);
    for (auto& kv : drop) {
        std::sort(kv.second.begin(), kv.second.end());
        kv.second.erase(std::unique(kv.second.begin(), kv.second.end()), kv.second.end());
    }
    std::vector<int32_t> ks, kstart{0}, ns;
    ns.reserve(a->ns.size());
    int64_t gone = 0;
    for (size_t i = 0; i < a->ks.size(); ++i) {
        const int s0 = a->kstart[i], e0 = a->kstart[i + 1];
        auto it = drop.find((uint32_t)a->ks[i]);
        const size_t before = ns.size();
        for (int j = s0; j < e0; ++j) {
            const int32_t n = a->ns[j];
            if (it != drop.end() && std::binary_search(it->second.begin(), it->second.end(), n)) {
                ++gone;
                continue;
            }
            ns.push_back(n);
        }
        if (ns.size() == before) continue;  // list emptied: drop the k
        ks.push_back(a->ks[i]);
        kstart.push_back((int32_t)ns.size());
    }
    a->ks.swap(ks);
    a->kstart.swap(kstart);
    a->ns.swap(ns);
    a->minN = a->maxN = 0;
    if (!a->ns.empty()) {
        a->minN = *std::min_element(a->ns.begin(), a->ns.end());
        a->maxN = *std::max_element(a->ns.begin(), a->ns.end());
    }
    if (removed) *removed = gone;
    return RS_OK;
}

extern "C" void rs_abcd_free(rs_abcd* a) { delete a; }

// ------------------------------------------------------------ synthetic input

namespace {

uint64_t inv_mod(uint64_t a, uint64_t q) {  // q prime, a != 0 mod q
    return rsh::powmod(a % q, q - 2, q);
}

// Multiplicative order of b mod prime q (b != 0 mod q), from q-1's factors.
uint32_t order_mod(uint64_t b, uint32_t q) {
    uint32_t ord = q - 1, r = q - 1;
    for (uint32_t f = 2; (uint64_t)f * f <= r; ++f) {
        if (r % f) continue;
        while (r % f == 0) r /= f;
        while (ord % f == 0 && rsh::powmod(b, ord / f, q) == 1) ord /= f;
    }
    if (r > 1 && rsh::powmod(b, ord / r, q) == 1) ord /= r;
    return ord;
}

}  // namespace

// Presieve: n survives iff no odd prime q < bound (q ∤ b) divides k*b^n-1.
// For each q the subgroup generated by b mod q is walked once (b^i, i up to
// ord_q(b)), looking each element up in a small hash of the targets
// k^-1 mod q of the batch; a hit at i removes n == i (mod ord_q(b)) for that
// k.  Work is split over q across threads, which mark one shared removal
// byte per (k, n).
extern "C" rs_status rs_abcd_synth(uint32_t base, int nk, uint32_t kmin, uint32_t kmax,
                                   int odd_k, int c, int nmin, int nmax,
                                   uint32_t presieve_bound, uint64_t seed,
                                   rs_abcd** out) {
    if (!out || base < 2 || nk < 1 || kmin < 1 || kmax < kmin || nmin < 1 ||
        nmax < nmin || kmax > 0x7fffffffu || odd_k < 0 || odd_k > 2 || (c != -1 && c != 1))
        return RS_ERR_ARG;
    *out = nullptr;
    const uint64_t span = (uint64_t)(kmax - kmin) + 1;
    // k of the requested parity in [kmin, kmax]: odd_k 1 odd, 2 even, 0 any
    const uint32_t want_par = odd_k == 1 ? 1u : 0u;
    const uint64_t avail = odd_k ? (span + ((kmin & 1) == want_par ? 1 : 0)) / 2 : span;
    if ((uint64_t)nk > avail) return RS_ERR_ARG;
    std::vector<uint32_t> qs;
    for (uint32_t q : rsh::primes_upto(presieve_bound ? presieve_bound - 1 : 0))
        if (q != 2 && base % q != 0) qs.push_back(q);
    const int width = nmax - nmin + 1;
    rsh::SplitMix64 rng(seed);
    std::vector<uint32_t> chosen;
    std::vector<std::vector<int32_t>> lists;
    std::unordered_set<uint32_t> used;
    const int nthreads = rsh::hw_threads(0);
    int draws = 0;
    while ((int)chosen.size() < nk) {
        // draw a batch of candidate k and presieve them together
        std::vector<uint32_t> batch;
        while ((int)(chosen.size() + batch.size()) < nk) {
            if (++draws > 64 * nk + 1024) return RS_ERR_ARG;  // cannot fill
            uint32_t k = kmin + (uint32_t)rng.below(span);
            if (odd_k && (k & 1) != want_par) { if (k + 1 > kmax) continue; k += 1; }
            if (used.count(k)) continue;
            used.insert(k);
            batch.push_back(k);
        }
        const size_t B = batch.size();
        // one shared removal byte per (k, n); marks are idempotent, written
        // with relaxed atomic stores
        std::vector<uint8_t> removed(B * (size_t)width, 0);
        std::vector<std::thread> th;
        for (int t = 0; t < nthreads; ++t) {
            th.emplace_back([&, t]() {
                // Per q: x_k = dlog_b(k^-1) mod ord by baby-step/giant-step
                // over all k at once.  Baby entries t_k*b^j (j < m) go into
                // an open-addressing table with chains for equal values;
                // giant steps b^(i*m) then give x_k = i*m - j.  m is about
                // sqrt(ord/B), so the cost is ~2 sqrt(ord*B) per q instead
                // of ord (m = 1 is a plain walk of the subgroup).
                std::vector<uint32_t> hval, hidx, ej;
                std::vector<int32_t> next, ek;
                std::vector<uint32_t> xk(B);
                uint8_t* rm = removed.data();
                for (size_t qi = t; qi < qs.size(); qi += nthreads) {
                    const uint32_t q = qs[qi];
                    const uint64_t bq = base % q;
                    const uint32_t ord = order_mod(bq, q);
                    uint32_t m = (uint32_t)std::sqrt((double)ord / (double)B);
                    if (m < 1) m = 1;
                    if (m > ord) m = ord;
                    const size_t E = B * (size_t)m;
                    int hb = 1;
                    while (((size_t)1 << hb) < 2 * E) ++hb;
                    const uint32_t hmask = (1u << hb) - 1;
                    hval.assign(hmask + 1, 0);
                    hidx.assign(hmask + 1, 0xffffffffu);
                    next.resize(E);
                    ek.resize(E);
                    ej.resize(E);
                    std::fill(xk.begin(), xk.end(), 0xffffffffu);
                    size_t ne = 0, want = 0;
                    for (size_t i = 0; i < B; ++i) {
                        const uint64_t k = batch[i] % q;
                        if (k == 0) continue;  // k*b^n-1 == -1 mod q
                        ++want;
                        // q | k*b^n + c  <=>  b^n == -c * k^-1 (mod q)
                        uint64_t v = inv_mod(k, q);
                        if (c == 1) v = q - v;
                        for (uint32_t j = 0; j < m; ++j, v = v * bq % q) {
                            uint32_t h = ((uint32_t)v * 2654435761u) >> (32 - hb);
                            while (hidx[h] != 0xffffffffu && hval[h] != (uint32_t)v)
                                h = (h + 1) & hmask;
                            next[ne] = hidx[h] == 0xffffffffu ? -1 : (int32_t)hidx[h];
                            hval[h] = (uint32_t)v;
                            hidx[h] = (uint32_t)ne;
                            ek[ne] = (int32_t)i;
                            ej[ne] = j;
                            ++ne;
                        }
                    }
                    if (!want) continue;
                    const uint64_t gm = rsh::powmod(bq, m, q);
                    uint64_t g = gm;
                    size_t got = 0;
                    for (uint64_t i = 1; i * m < (uint64_t)ord + m && got < want; ++i, g = g * gm % q) {
                        uint32_t h = ((uint32_t)g * 2654435761u) >> (32 - hb);
                        while (hidx[h] != 0xffffffffu) {
                            if (hval[h] == (uint32_t)g) {
                                for (int32_t e = (int32_t)hidx[h]; e >= 0; e = next[e])
                                    if (xk[ek[e]] == 0xffffffffu) {
                                        xk[ek[e]] = (uint32_t)((i * m - ej[e]) % ord);
                                        ++got;
                                    }
                                break;
                            }
                            h = (h + 1) & hmask;
                        }
                    }
                    for (size_t i = 0; i < B; ++i) {
                        if (xk[i] == 0xffffffffu) continue;  // k^-1 not a power of b
                        // first n >= nmin with n == x (mod ord)
                        long long n = xk[i];
                        if (n < nmin) n += ((nmin - n + ord - 1) / ord) * (long long)ord;
                        uint8_t* row = rm + i * (size_t)width;
                        for (; n <= nmax; n += ord)
                            __atomic_store_n(&row[n - nmin], (uint8_t)1, __ATOMIC_RELAXED);
                    }
                }
            });
        }
        for (auto& x : th) x.join();
        for (size_t i = 0; i < B; ++i) {
            std::vector<int32_t> ns;
            const uint8_t* row = removed.data() + i * (size_t)width;
            for (int n = nmin; n <= nmax; ++n)
                if (!row[n - nmin]) ns.push_back(n);
            if (ns.empty()) continue;  // covered k: redraw
            chosen.push_back(batch[i]);
            lists.push_back(std::move(ns));
        }
    }
    // headers in ascending k, as srsieve writes them
    std::vector<int> order(nk);
    for (int i = 0; i < nk; ++i) order[i] = i;
    std::sort(order.begin(), order.end(),
              [&](int x, int y) { return chosen[x] < chosen[y]; });
    rs_abcd* a = new rs_abcd();
    a->base = base;
    a->c = c;
    for (int i : order) {
        a->kstart.push_back((int32_t)a->ns.size());
        a->ks.push_back((int32_t)chosen[i]);
        a->ns.insert(a->ns.end(), lists[i].begin(), lists[i].end());
    }
    finish(a);
    *out = a;
    return RS_OK;
}

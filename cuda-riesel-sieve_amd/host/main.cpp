// main.cpp -- rsieve, the command-line host: a drop-in for the reference
// program's main (kernel_v0.24.6.cu:688-1807) over librsieve.
//
// Flags as at :705-786: -i FILE (ABCD input, default sr_745.abcd), -m M
// (baby steps m = 2 << M, :931; without -m, or with M <= 1, the cost model
// picks m, where the reference sizes its tables from the free memory,
// :917-929; M = 2..4 runs m = 64, this build's smallest table), -p P0,
// -P P1, -Q Q (even), -O FILE (factor file, default factors.txt), -b/-t
// SCALE (primes per launch = 128*2^(b-1) * 128*2^(t-1), :817-818, :849-851),
// -s SCALE (at least 4096*s sieving primes, :1274), -v (no verification:
// datless output), -g (generate the primes on the GPU; the default is the
// CPU generator of the reference's design), -R FILE (write the input ABCD
// with every verified factor's candidate removed; the reference only
// appends factor lines), -h.  Unlike :705, -h and -v
// take no argument.  With no
// -p/-P the default test range of :837-839 is used.  The sieved range is
// clipped to [P0, P1] (the reference runs past P1 with its last batch).
// Output lines keep the reference's formats: factors "p | k*b^n-1"
// (:1833; "+1" for a Sierpinski file) and a progress line every 20
// launches (:1759).
#include <getopt.h>
#include <sched.h>

#include <algorithm>
#include <chrono>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <condition_variable>
#include <ctime>
#include <deque>
#include <future>
#include <mutex>
#include <thread>
#include <vector>

#include "rsieve.h"

static void usage() {
    printf(
        "\nrsieve -- a BSGS sieve for multiple sequences of the form k*b^n-1 (MI355X)\n\n"
        "-i FILE  : Read in abcd sieve file called FILE.\n"
        "-m m     : Use 2<<m baby steps (hash table elements); m <= 1 or no -m: cost model.\n"
        "-p P0    : Start sieving from P0. Must have a corresponding P1.\n"
        "-P P1    : Finish sieving at P1. Must have a corresponding P0.\n"
        "                If no -p and -P arguments then we default to 102297149770000 for testing.\n"
        "-Q Q     : Override subsequence value Q (even). Sieve k*b^n-1 as (k*b^d)*(b^Q)^m-1.\n"
        "-O FILE  : Write factors to a file called FILE.\n"
        "-b SCALE : Scale the primes per launch by 2^(SCALE-1) (blocks).\n"
        "-t SCALE : Scale the primes per launch by 2^(SCALE-1) (threads).\n"
        "-s SCALE : Use at least 4096*SCALE small primes for the prime generator. Default 256.\n"
        "-v       : Don't verify factors. This implies running datless.\n"
        "-g       : Generate the primes on the GPU instead of the CPU generator.\n"
        "-R FILE  : Write the ABCD file minus the candidates with a verified factor to FILE.\n"
        "-d DEV   : GPU device index (default 0).\n"
        "-h       : Prints this help\n");
}

// One launch's primes: p = low + offs[i], i < n (n = -1: the p-range
// [low, next) is generated on the device); next is the following batch's
// low; done marks the end of the range.
struct Batch {
    uint64_t low = 0, next = 0;
    long long n = 0;
    std::vector<uint32_t> offs;
    bool done = true;
};

// CPU-generated batches from a producer thread: windows of about 8 batches
// (at least 2^31 wide) are sieved at once by rs_primes_range (all host
// threads), one window ahead of the one being cut into batches of `size`
// primes as the reference cuts its prime arrays (:1561), at most 4 batches
// ahead of the sieve.
class BatchFeed {
   public:
    BatchFeed(rs_primegen* g, uint64_t low, uint64_t high, long long size)
        : g_(g), low_(low), high_(high), size_(size) {}
    ~BatchFeed() {
        {
            std::lock_guard<std::mutex> l(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        if (th_.joinable()) th_.join();
    }
    void start() { th_ = std::thread([this] { run(); }); }
    // The next batch (done = true after the last), or the generator's error.
    rs_status pop(Batch* b) {
        std::unique_lock<std::mutex> l(mu_);
        cv_.wait(l, [&] { return !q_.empty() || finished_; });
        if (!q_.empty()) {
            *b = std::move(q_.front());
            q_.pop_front();
            cv_.notify_all();
            return RS_OK;
        }
        b->done = true;
        return st_;
    }

   private:
    void push(Batch&& b) {
        std::unique_lock<std::mutex> l(mu_);
        cv_.wait(l, [&] { return q_.size() < 4 || stop_; });
        q_.push_back(std::move(b));
        cv_.notify_all();
    }
    void finish(rs_status st) {
        std::lock_guard<std::mutex> l(mu_);
        st_ = st;
        finished_ = true;
        cv_.notify_all();
    }
    // Primes of the window [lo, min(lo + width, P1 + 1)) as offsets from lo.
    struct Window {
        uint64_t lo = 0, hi = 0;
        std::vector<uint32_t> offs;
        rs_status st = RS_OK;
    };
    Window sieve_window(uint64_t wlo) {
        Window w;
        w.lo = wlo;
        const double lnp = std::log((double)std::max<uint64_t>(wlo, 16));
        // at least 2^31 wide, so that each generator thread sieves many
        // segments per block
        uint64_t width = (uint64_t)(8.0 * size_ * lnp);
        width = std::min<uint64_t>(std::max<uint64_t>(width, 1ull << 31), 1ull << 32);
        w.hi = std::min<uint64_t>(wlo + width, high_ + 1);
        w.offs.resize((size_t)((w.hi - wlo) / std::max(1.0, lnp - 1.5) * 1.2) + 4096);
        int64_t got = 0;
        w.st = rs_primes_range(g_, wlo, w.hi, w.offs.data(), (int64_t)w.offs.size(), &got);
        if (w.st == RS_ERR_OVERFLOW) {
            w.offs.resize((size_t)got);
            w.st = rs_primes_range(g_, wlo, w.hi, w.offs.data(), got, &got);
        }
        w.offs.resize(w.st == RS_OK ? (size_t)got : 0);
        return w;
    }
    void run() {
        uint64_t cur = low_;   // the next batch's low
        uint64_t wbase = low_; // the current window's low: its primes are wbase + win[i]
        std::vector<uint32_t> win;
        size_t head = 0, cnt = 0;
        // the next window is sieved while the current one is cut and sent
        std::future<Window> ahead;
        if (low_ <= high_)
            ahead = std::async(std::launch::async, [this] { return sieve_window(low_); });
        while (cur <= high_) {
            {
                std::lock_guard<std::mutex> l(mu_);
                if (stop_) break;
            }
            Batch b;
            b.low = cur;
            b.done = false;
            b.offs.resize((size_t)size_);
            size_t n = 0;
            while (n < (size_t)size_) {
                if (head == cnt) {  // take the next window
                    if (!ahead.valid()) break;
                    Window w = ahead.get();
                    if (w.st != RS_OK) return finish(w.st);
                    if (w.hi <= high_) {
                        const uint64_t nlo = w.hi;
                        ahead = std::async(std::launch::async, [this, nlo] { return sieve_window(nlo); });
                    }
                    wbase = w.lo;
                    win = std::move(w.offs);
                    head = 0;
                    cnt = win.size();
                    continue;
                }
                // the window's next primes, while their offsets from the
                // batch's low fit 32 bits (wbase + win[i] - cur < 2^32)
                size_t take = std::min((size_t)size_ - n, cnt - head);
                const uint64_t room = (1ull << 32) + cur - wbase;  // win[i] < room
                if (room <= 0xffffffffull)
                    take = std::min<size_t>(
                        take, std::lower_bound(win.begin() + head, win.begin() + head + take,
                                               (uint32_t)room) - (win.begin() + head));
                if (take == 0) break;
                const uint32_t adj = (uint32_t)(wbase - cur);  // mod 2^32
                const uint32_t* src = win.data() + head;
                uint32_t* dst = b.offs.data() + n;
                for (size_t i = 0; i < take; ++i) dst[i] = src[i] + adj;
                head += take;
                n += take;
                if (head < cnt && n < (size_t)size_) break;  // stopped by the 32-bit bound
            }
            if (n == 0) break;  // no prime left in [cur, P1]
            b.offs.resize(n);
            b.n = (long long)n;
            b.next = cur + b.offs.back() + 2;
            cur = b.next;
            push(std::move(b));
        }
        finish(RS_OK);
    }

    rs_primegen* g_;
    uint64_t low_, high_;
    long long size_;
    std::thread th_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<Batch> q_;
    bool stop_ = false, finished_ = false;
    rs_status st_ = RS_OK;
};

static int fail(const char* what, rs_status st) {
    fprintf(stderr, "%s failed: %s (%d)\n", what, rs_strerror(st), st);
    return 1;
}

int main(int argc, char* argv[]) {
    int blockScale = 1, threadScale = 1, Qin = 0, hashFlag = 0, verify = 1;
    int smallPrimeScale = 256, device = 0, gpuPrimes = 0;
    const char* abcdFile = "sr_745.abcd";
    const char* factorFile = "factors.txt";
    const char* removedFile = nullptr;
    unsigned long long low = 0, high = 0;
    int opt;
    while ((opt = getopt(argc, argv, "b:hi:m:p:P:Q:O:R:s:t:vd:g")) != -1) {
        switch (opt) {
            case 'b': blockScale = (int)strtol(optarg, nullptr, 0); break;
            case 'h': usage(); return 0;
            case 'i': abcdFile = optarg; break;
            case 'm': hashFlag = (int)strtol(optarg, nullptr, 0); break;
            case 'p': low = strtoull(optarg, nullptr, 0); break;
            case 'P': high = strtoull(optarg, nullptr, 0); break;
            case 'Q':
                Qin = (int)strtol(optarg, nullptr, 0);
                if (Qin < 1 || Qin % 2 == 1) {
                    printf("Bad input parameter : Q must be an even integer greater than 0.\n");
                    return 1;
                }
                break;
            case 'O': factorFile = optarg; break;
            case 's': smallPrimeScale = (int)strtol(optarg, nullptr, 0); break;
            case 't': threadScale = (int)strtol(optarg, nullptr, 0); break;
            case 'v': verify = 0; break;
            case 'd': device = (int)strtol(optarg, nullptr, 0); break;
            case 'g': gpuPrimes = 1; break;
            case 'R': removedFile = optarg; break;
            default: usage(); return 1;
        }
    }
    if (removedFile && !verify) {
        printf("Bad input parameters : -R needs verified factors (it cannot be used with -v).\n");
        return 1;
    }
    if (blockScale < 1 || blockScale > 31 || threadScale < 1 || threadScale > 31) {
        printf("Bad input parameter : 1 <= BlockScale, ThreadScale <= 31.\n");
        return 1;
    }
    // :917-931: M <= 1 sizes the table automatically, M >= 2 takes m = 2 << M
    if (hashFlag < 0 || hashFlag > 15) {
        printf("Bad input parameter : -m must be in [0, 15] (m = 2 << M; M <= 1: automatic).\n");
        return 1;
    }
    if (smallPrimeScale < 1 || smallPrimeScale > 4096) {
        printf("Bad input parameter : -s must be in [1, 4096].\n");
        return 1;
    }
    printf("BlockScale = %d\nThreadScale = %d\nLow = %llu\nHigh = %llu\nQ_In = %d\n", blockScale,
           threadScale, low, high, Qin);
    printf("HashTableSize = %d\nVerify = %d\nS = %d\nABCD = %s\nFactorFile = %s\n\n", hashFlag,
           verify, smallPrimeScale, abcdFile, factorFile);
    const long long bscale = 1ll << (blockScale - 1), tscale = 1ll << (threadScale - 1);
    if (low == 0 && high == 0) {  // :821-841
        low = 102297149770000ull;
        high = low + 2500000ull * (unsigned long long)(bscale * tscale);
    }
    if (high < low) {
        printf("Bad input parameters : P1 < P0. We want to sieve primes p in the range P0 <= p <= P1.\n");
        return 1;
    }
    const long long arraySize = 128 * bscale * 128 * tscale;
    if (arraySize > (1ll << 30)) {
        printf("Bad input parameter : primes per launch must stay below 2^30.\n");
        return 1;
    }

    rs_abcd* abcd = nullptr;
    rs_status st = rs_abcd_read(abcdFile, &abcd);
    if (st == RS_ERR_IO) {
        printf("Unable to open file %s\n", abcdFile);
        return 1;
    }
    if (st != RS_OK) return fail("reading the ABCD file", st);
    uint32_t base;
    int nk, nn, minN, maxN;
    rs_abcd_info(abcd, &base, &nk, &nn, &minN, &maxN);
    printf("The base is %u\n%d k-values, %d candidates\nMinN = %d and MaxN = %d\n", base, nk, nn,
           minN, maxN);
    int form = -1;
    rs_abcd_form(abcd, &form);
    if (form == 1) printf("Sierpinski form k*b^n+1\n");
    rs_plan* plan = nullptr;
    if ((st = rs_plan_create(abcd, Qin, &plan)) != RS_OK) return fail("planning", st);
    int Q, nrows, words;
    rs_plan_info(plan, nullptr, &Q, &nrows, &words, nullptr, nullptr);
    // The factor set does not depend on m, so the reference's small tables
    // (-m 2..4: m = 8..32) run with this build's smallest, m = 64.
    int m = hashFlag <= 1 ? 0 : (2 << hashFlag);
    if (m != 0 && m < 64) {
        printf("Note: -m %d asks for %d baby steps; this build's smallest table, 64, is used "
               "(the factors found are the same).\n", hashFlag, m);
        m = 64;
    }
    printf("Q=%d. Subsequences= %d. Range = %d\n", Q, nrows, maxN / Q - minN / Q + 1);
    if (low < 3) low = 3;
    if (low % 2 == 0) low -= 1;  // :1515-1519
    if (low <= base) {
        printf("Bad input parameters : P0 must be larger than the base %u.\n", base);
        return 1;
    }

    rs_sieve* sv = nullptr;
    if ((st = rs_sieve_create(device, plan, m, 0, arraySize, 1 << 20, &sv)) != RS_OK)
        return fail("device set-up", st);
    int mUsed, T, slots, ldsw, tile;
    rs_sieve_info(sv, &mUsed, &T, &slots, &ldsw, &tile);
    printf("Each prime uses %d baby steps, %d giant steps per row, a %d-slot LDS table "
           "(%d bytes of LDS per wavefront)\n\n", mUsed, T, slots, ldsw);
    // the generator's threads leave two of the process's CPUs to the batch
    // cutter and this submitting thread
    int cpus = (int)std::thread::hardware_concurrency();
    cpu_set_t cset;
    if (sched_getaffinity(0, sizeof cset, &cset) == 0) cpus = CPU_COUNT(&cset);
    if (const char* e = getenv("OMP_NUM_THREADS"))
        if (atoi(e) > 0) cpus = std::min(cpus, atoi(e));
    rs_primegen* gen = nullptr;
    if ((st = rs_primegen_create(smallPrimeScale, high + 1, std::max(1, cpus - 2), &gen)) !=
        RS_OK)
        return fail("prime generator set-up", st);

    const unsigned long long startLow = low;
    std::vector<rs_factor> hits(1 << 20);
    std::vector<rs_factor> verified;  // candidates to remove with -R
    // Batches as in the reference (:1537-1561): the first arraySize primes >=
    // low (low odd), offsets from low, the next low = last prime + 2; the
    // last batch is clipped to P1.  With the CPU generator they come from a
    // producer thread that sieves windows of about 8 batches at once with
    // all host threads (one window per call keeps every thread busy) and
    // cuts them into batches, so generation overlaps the sieve; with -g each
    // launch is a p-range sieved on the GPU (at least 2^30 wide: the device
    // generator needs wide ranges).
    BatchFeed feed(gen, low, high, arraySize);
    if (!gpuPrimes) feed.start();
    uint64_t gLow = low;
    auto next_batch = [&](Batch* b) -> rs_status {
        if (!gpuPrimes) return feed.pop(b);
        if (gLow > high) {
            b->done = true;
            return RS_OK;
        }
        const double lnp = std::log((double)std::max<uint64_t>(gLow, 16));
        uint64_t width = std::max<uint64_t>((uint64_t)(arraySize * lnp), 1ull << 30);
        width = std::min<uint64_t>(width, 1ull << 32);
        b->low = gLow;
        b->next = std::min<uint64_t>(gLow + width, high + 1);
        b->n = -1;
        b->done = false;
        gLow = b->next;
        return RS_OK;
    };
    // collects and writes the hits of everything submitted so far (the
    // batches are consecutive in p, so the sorted hits come out in the order
    // a collect per batch would give)
    long long factorCount = 0;
    auto drain = [&]() -> rs_status {
        int64_t got = 0, lost = 0;
        rs_status s2 = rs_sieve_collect(sv, hits.data(), (int64_t)hits.size(), &got, &lost);
        if (s2 == RS_ERR_OVERFLOW) {
            hits.resize((size_t)got);
            s2 = rs_sieve_collect(sv, hits.data(), (int64_t)hits.size(), &got, &lost);
        }
        if (s2 != RS_OK) return s2;
        if (lost > 0) fprintf(stderr, "Warning: %" PRId64 " hits dropped (append buffer full)\n", lost);
        FILE* ff = nullptr;
        for (int64_t i = 0; i < got; ++i) {
            const rs_factor& f = hits[i];
            if (verify) {
                if (!rs_plan_contains(plan, f.k, f.n)) continue;              // :1620-1647
                if (!rs_verify_factor(f.p, f.k, base, f.n, form)) {
                    fprintf(stderr, "Error: device reported %llu | %u*%u^%u%s, which fails "
                            "verification\n", (unsigned long long)f.p, f.k, base, f.n,
                            form == 1 ? "+1" : "-1");
                    exit(1);
                }
            }
            if (removedFile) verified.push_back(f);
            char line[128];
            rs_format_factor(&f, base, form, line, sizeof line);
            fputs(line, stdout);
            if (!ff) ff = fopen(factorFile, "a");
            if (!ff) {
                printf("Unable to open factor file %s\n", factorFile);
                exit(1);
            }
            fputs(line, ff);
            ++factorCount;
        }
        if (ff) fclose(ff);
        return RS_OK;
    };
    // hits are collected every kCollect launches (the device keeps them in
    // its append buffer meanwhile), so the host never waits for a launch it
    // has just queued
    const int kCollect = 8;
    int kernelCount = 0;
    auto t0 = std::chrono::steady_clock::now();
    auto tb = t0;
    Batch cur;
    if ((st = next_batch(&cur)) != RS_OK) return fail("prime generation", st);
    while (!cur.done) {
        ++kernelCount;
        unsigned long long progress = 0;
        if (cur.n < 0) {  // GPU-generated primes of [low, next)
            if ((st = rs_sieve_submit_range(sv, cur.low, cur.next)) != RS_OK)
                return fail("submit", st);
            progress = cur.next - cur.low;
        } else {
            if (cur.n > 0 &&
                (st = rs_sieve_submit(sv, cur.low, cur.offs.data(), cur.n)) != RS_OK)
                return fail("submit", st);
            progress = cur.n > 0 ? cur.offs[cur.n - 1] - cur.offs[0] : 0;
        }
        const uint64_t nLow = cur.next;
        Batch nxt;
        if ((st = next_batch(&nxt)) != RS_OK) return fail("prime generation", st);
        if (nxt.done || kernelCount % kCollect == 0)
            if ((st = drain()) != RS_OK) return fail("collect", st);
        auto te = std::chrono::steady_clock::now();
        const double dt = std::chrono::duration<double>(te - tb).count();
        tb = te;
        if (kernelCount % 20 == 0) {
            const double done = 100.0 * (double)(nLow - startLow) / (double)(high - startLow + 1);
            const double soFar = std::chrono::duration<double>(te - t0).count();
            const double left = done > 0 ? soFar * (100.0 / done - 1.0) : 0.0;
            time_t eta = time(nullptr) + (time_t)left;
            char etaBuf[64];
            strftime(etaBuf, sizeof etaBuf, "%a %b %d %H:%M:%S %Y", localtime(&eta));
            printf("p=%llu, %.0f p/sec,  %.2f%% done, %lld factors found, ETA %s \n",
                   (unsigned long long)nLow,
                   progress / dt, done > 100 ? 100.0 : done, factorCount, etaBuf);
            fflush(stdout);
        }
        cur = std::move(nxt);
    }
    if (kernelCount == 0 && (st = drain()) != RS_OK) return fail("collect", st);
    const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const unsigned long long swept = high + 1 - startLow;
    printf("\nTime taken: %gs\n", total);
    printf("Factors found: %lld\n", factorCount);
    printf("Time per kernel: %g\n", kernelCount ? total / kernelCount : 0.0);
    printf("Progress = %llu at %.0f p/sec\n\n", swept, swept / total);
    if (removedFile) {
        int64_t removed = 0;
        if ((st = rs_abcd_remove(abcd, verified.data(), (int64_t)verified.size(), &removed)) !=
                RS_OK ||
            (st = rs_abcd_write(abcd, removedFile)) != RS_OK)
            return fail("writing the reduced ABCD file", st);
        rs_abcd_info(abcd, nullptr, &nk, &nn, nullptr, nullptr);
        printf("Removed %lld candidates; %d k-values, %d candidates left in %s\n",
               (long long)removed, nk, nn, removedFile);
    }
    rs_primegen_free(gen);
    rs_sieve_destroy(sv);
    rs_plan_free(plan);
    rs_abcd_free(abcd);
    return 0;
}

// main.cpp -- rsieve, the command-line host: a drop-in for the reference
// program's main (kernel_v0.24.6.cu:688-1807) over librsieve.
//
// Flags as at :705-786: -i FILE (ABCD input, default sr_745.abcd), -m M
// (baby steps m = 2 << M, :931; without -m, or with M <= 1, the cost model
// picks m, where the reference sizes its tables from the free memory,
// :917-929; M = 2..4 runs m = 64, this build's smallest table), -p P0,
// -P P1, -Q Q (even), -O FILE (factor file, default factors.txt), -b/-t
// SCALE (primes per batch, the reference's launch, = 128*2^(b-1) *
// 128*2^(t-1), :817-818, :849-851; a device launch here holds up to 2^24
// primes of whole batches), -s SCALE (at least 4096*s sieving primes,
// :1274), -v (no verification:
// datless output), -g (generate the primes on the GPU; the default is the
// CPU generator of the reference's design), -R FILE (write the input ABCD
// with every verified factor's candidate removed; the reference only
// appends factor lines), -h.  Unlike :705, -h and -v
// take no argument.  With no
// -p/-P the default test range of :837-839 is used.  The sieved range is
// clipped to [P0, P1] (the reference runs past P1 with its last batch).
// Output lines keep the reference's formats: factors "p | k*b^n-1"
// (:1833; "+1" for a Sierpinski file) and a progress line every 20
// batches (:1759).
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
#include <memory>
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
        "-b SCALE : Scale the primes per batch by 2^(SCALE-1) (blocks).\n"
        "-t SCALE : Scale the primes per batch by 2^(SCALE-1) (threads).\n"
        "-s SCALE : Use at least 4096*SCALE small primes for the prime generator. Default 256.\n"
        "-v       : Don't verify factors. This implies running datless.\n"
        "-g       : Generate the primes on the GPU instead of the CPU generator.\n"
        "-R FILE  : Write the ABCD file minus the candidates with a verified factor to FILE.\n"
        "-d DEV   : GPU device index (default 0).\n"
        "-h       : Prints this help\n");
}

// Windows of true primes [lo, hi) (p = lo + offs[i], i < n) sieved ahead of
// the device by a producer thread with the generator's host threads.
// Windows are 2^30 wide: on 15 threads at 1e14 the generator sieves one in
// about 30 ms, the device about 40 (so after the first window the device
// never waits), and every thread gets a block of its own.  Three buffers
// are reused in turn (no allocation or page faults per window): one being
// submitted, one ready, one being sieved.
class WindowFeed {
   public:
    struct Window {
        uint64_t lo = 0, hi = 0;
        const uint32_t* offs = nullptr;
        size_t n = 0;
        int buf = -1;  // -1: no window (the range is done)
    };
    WindowFeed(rs_primegen* g, uint64_t low, uint64_t high) : g_(g), low_(low), high_(high) {
        for (int b = 0; b < kBufs; ++b) free_.push_back(b);
    }
    ~WindowFeed() {
        {
            std::lock_guard<std::mutex> l(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        if (th_.joinable()) th_.join();
    }
    void start() { th_ = std::thread([this] { run(); }); }
    // The next window (buf = -1 after the last), or the generator's error.
    rs_status pop(Window* w) {
        std::unique_lock<std::mutex> l(mu_);
        cv_.wait(l, [&] { return !ready_.empty() || finished_; });
        if (!ready_.empty()) {
            *w = ready_.front();
            ready_.pop_front();
            return RS_OK;
        }
        *w = Window{};
        return st_;
    }
    // The window's offsets have been copied out: its buffer is free again.
    void release(const Window& w) {
        if (w.buf < 0) return;
        std::lock_guard<std::mutex> l(mu_);
        free_.push_back(w.buf);
        cv_.notify_all();
    }

   private:
    static const int kBufs = 3;
    void run() {
        const uint64_t width = 1ull << 30;
        for (uint64_t lo = low_; lo <= high_;) {
            int b;
            {
                std::unique_lock<std::mutex> l(mu_);
                cv_.wait(l, [&] { return !free_.empty() || stop_; });
                if (stop_) return;
                b = free_.front();
                free_.pop_front();
            }
            const uint64_t hi = std::min<uint64_t>(lo + width, high_ + 1);
            // room for the window's primes: about (hi - lo) / ln lo, and never
            // more than pi(hi) < 1.25506 hi / ln hi (Rosser and Schoenfeld),
            // which bounds windows that start at small p (a short count is
            // retried below with the exact one)
            const double lnp = std::log((double)std::max<uint64_t>(lo, 16));
            const double lnh = std::log((double)std::max<uint64_t>(hi, 16));
            size_t need = (size_t)std::min((double)(hi - lo) / std::max(1.0, lnp - 1.5) * 1.2,
                                           1.25506 * (double)hi / lnh) + 4096;
            int64_t got = 0;
            rs_status st = RS_OK;
            for (int pass = 0; pass < 2; ++pass) {
                if (cap_[b] < need) {
                    buf_[b].reset(new uint32_t[need]);  // left uninitialised
                    cap_[b] = need;
                }
                st = rs_primes_range(g_, lo, hi, buf_[b].get(), (int64_t)cap_[b], &got);
                if (st != RS_ERR_OVERFLOW) break;
                need = (size_t)got;
            }
            std::lock_guard<std::mutex> l(mu_);
            if (st != RS_OK) {
                st_ = st;
                break;
            }
            ready_.push_back(Window{lo, hi, buf_[b].get(), (size_t)got, b});
            cv_.notify_all();
            lo = hi;
        }
        std::lock_guard<std::mutex> l(mu_);
        finished_ = true;
        cv_.notify_all();
    }

    rs_primegen* g_;
    uint64_t low_, high_;
    std::unique_ptr<uint32_t[]> buf_[kBufs];
    size_t cap_[kBufs] = {};
    std::thread th_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<int> free_;
    std::deque<Window> ready_;
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
        printf("Bad input parameter : primes per batch must stay below 2^30.\n");
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

    // Device launches hold whole reference batches, several at once when a
    // batch is small (at -b 5 a batch of 262144 primes is one 64-prime group
    // per resident wavefront, and a launch's fixed cost is then a third of
    // its time), up to kLaunchPrimes primes; the factors found do not depend
    // on how batches are grouped into launches.
    const long long kLaunchPrimes = std::max<long long>(arraySize, 1ll << 24);
    const int64_t kHitCap = 1 << 22;
    // Giant-step tile (the reference's thread takes one step at a time):
    // 4 steps per lane and item where that adds no steps over 2 (T, the
    // steps per row ceil(X/m) rounded up to a multiple of the tile, is the
    // same for both), with row-number lists for plans of at most 512 rows
    // (a smaller layout: five wavefronts per SIMD); else 2.  On sr_745.abcd
    // at m = 512: 3.06e10 p/s against 2.75e10 with tile 2.
    const int mPick = m ? m : rs_plan_auto_m(plan);
    const long long X = (long long)(maxN / Q) - (long long)(minN / Q) + 1;
    const long long steps = (X + mPick - 1) / mPick;
    auto round_to = [](long long t, int tw) { return std::max<long long>(2, (t + tw - 1) / tw * tw); };
    int tileSel = 2;
    if (round_to(steps, 4) == round_to(steps, 2))
        tileSel = nrows <= 512 ? (4 | RS_TILE_ROW_INDEX) : 4;
    rs_sieve* sv = nullptr;
    st = rs_sieve_create(device, plan, m, tileSel, kLaunchPrimes, kHitCap, &sv);
    if (st == RS_ERR_ARG && tileSel != 2)  // the wider tile's layout refused: tile 2
        st = rs_sieve_create(device, plan, m, 2, kLaunchPrimes, kHitCap, &sv);
    if (st != RS_OK) return fail("device set-up", st);
    int mUsed, T, slots, ldsw, tile;
    rs_sieve_info(sv, &mUsed, &T, &slots, &ldsw, &tile);
    printf("Each prime uses %d baby steps, %d giant steps per row (%d per lane and item%s), a "
           "%d-slot LDS table (%d bytes of LDS per wavefront)\n\n", mUsed, T, tile & 7,
           (tile & RS_TILE_ROW_INDEX) ? ", row-number lists" : "", slots, ldsw);
    // the generator's threads leave one of the process's CPUs to this
    // submitting thread
    int cpus = (int)std::thread::hardware_concurrency();
    cpu_set_t cset;
    if (sched_getaffinity(0, sizeof cset, &cset) == 0) cpus = CPU_COUNT(&cset);
    if (const char* e = getenv("OMP_NUM_THREADS"))
        if (atoi(e) > 0) cpus = std::min(cpus, atoi(e));
    rs_primegen* gen = nullptr;
    if ((st = rs_primegen_create(smallPrimeScale, high + 1, std::max(1, cpus - 1), &gen)) !=
        RS_OK)
        return fail("prime generator set-up", st);

    const unsigned long long startLow = low;
    std::vector<rs_factor> hits(1 << 20);
    std::vector<rs_factor> verified;  // candidates to remove with -R
    // collects and writes the hits of everything submitted so far (the
    // launches are consecutive in p, so the sorted hits come out in the
    // order a collect per launch would give); *raw = hits collected
    long long factorCount = 0;
    auto drain = [&](int64_t* raw) -> rs_status {
        int64_t got = 0, lost = 0;
        rs_status s2 = rs_sieve_collect(sv, hits.data(), (int64_t)hits.size(), &got, &lost);
        if (s2 == RS_ERR_OVERFLOW) {
            hits.resize((size_t)got);
            s2 = rs_sieve_collect(sv, hits.data(), (int64_t)hits.size(), &got, &lost);
        }
        if (s2 != RS_OK) return s2;
        *raw = got;
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
    // Hits wait in the device's append buffer between collects, and a
    // collect waits for the device, so collects are spaced by the hits they
    // bring: the first after one batch, then every drainEvery primes,
    // doubled (up to 2^28) while a collect brings fewer than kHitCap/16 hits
    // and halved (down to one batch) when it brings more than kHitCap/4.
    long long drainEvery = arraySize, sinceDrain = 0;
    // batches in the next launch: one after each collect, doubling after
    // every launch, so the copy of each launch overlaps the sieve of the
    // one before (a full-size launch right after a collect would leave the
    // device idle while the host copies its offsets)
    long long ramp = 1;
    auto paced_drain = [&](long long primes, bool last) -> rs_status {
        sinceDrain += primes;
        if (!last && sinceDrain < drainEvery) return RS_OK;
        int64_t raw = 0;
        rs_status s2 = drain(&raw);
        ramp = 1;
        if (raw < kHitCap / 16)
            drainEvery = std::min(drainEvery * 2, 1ll << 28);
        else if (raw > kHitCap / 4)
            drainEvery = std::max(drainEvery / 2, arraySize);
        sinceDrain = 0;
        return s2;
    };
    // the progress line of :1759, every 20 batches; the rate is over the
    // p-range since the previous line
    long long kernelCount = 0;  // batches (the reference's launches) submitted
    auto t0 = std::chrono::steady_clock::now();
    auto tLine = t0;
    uint64_t pLine = startLow;
    auto progress_line = [&](uint64_t nLow) {
        auto te = std::chrono::steady_clock::now();
        const double dt = std::chrono::duration<double>(te - tLine).count();
        const double done = 100.0 * (double)(nLow - startLow) / (double)(high - startLow + 1);
        const double soFar = std::chrono::duration<double>(te - t0).count();
        const double left = done > 0 ? soFar * (100.0 / done - 1.0) : 0.0;
        time_t eta = time(nullptr) + (time_t)left;
        char etaBuf[64];
        strftime(etaBuf, sizeof etaBuf, "%a %b %d %H:%M:%S %Y", localtime(&eta));
        printf("p=%llu, %.0f p/sec,  %.2f%% done, %lld factors found, ETA %s \n",
               (unsigned long long)nLow, dt > 0 ? (double)(nLow - pLine) / dt : 0.0,
               done > 100 ? 100.0 : done, factorCount, etaBuf);
        fflush(stdout);
        tLine = te;
        pLine = nLow;
    };
    // Batches as in the reference (:1537-1561): the first arraySize primes >=
    // low (low odd), the next batch from the last prime + 2, the last one
    // clipped to P1.  With the CPU generator the primes come in windows from
    // a producer thread, so generation overlaps the sieve, and each launch
    // is a slice of a window (offsets from the window's low: no copy beyond
    // the one into the pinned staging buffer) ending on a batch boundary or
    // at the window's end.  With -g each launch is a p-range sieved on the
    // GPU (at least 2^30 wide: the device generator needs wide ranges).
    if (!gpuPrimes) {
        WindowFeed feed(gen, low, high);
        feed.start();
        // the pinned staging buffers (64 MB each), allocated while the first
        // window of primes is sieved rather than at the first launches
        if ((st = rs_sieve_submit(sv, low, nullptr, 0)) != RS_OK) return fail("submit", st);
        long long inBatch = 0;  // primes of the open batch submitted so far
        for (;;) {
            WindowFeed::Window w;
            if ((st = feed.pop(&w)) != RS_OK) return fail("prime generation", st);
            if (w.buf < 0) break;
            for (size_t i = 0; i < w.n;) {
                const long long k = std::min(
                    ramp, std::max<long long>(1, std::min(kLaunchPrimes, drainEvery) / arraySize));
                ramp = 2 * k;
                const size_t take =
                    (size_t)std::min<long long>(k * arraySize - inBatch, (long long)(w.n - i));
                if ((st = rs_sieve_submit(sv, w.lo, w.offs + i, (int64_t)take)) != RS_OK)
                    return fail("submit", st);
                const long long ends = (inBatch + (long long)take) / arraySize;
                inBatch = (inBatch + (long long)take) % arraySize;
                i += take;
                if (ends > 0) {  // the last batch completed ends at prime i - inBatch - 1
                    const bool line = (kernelCount + ends) / 20 > kernelCount / 20;
                    kernelCount += ends;
                    if (line) progress_line(w.lo + w.offs[i - (size_t)inBatch - 1] + 2);
                }
                if ((st = paced_drain((long long)take, false)) != RS_OK)
                    return fail("collect", st);
            }
            feed.release(w);
        }
        if (inBatch > 0) ++kernelCount;  // the last batch, clipped to P1
    } else {
        for (uint64_t gLow = low; gLow <= high;) {
            const double lnp = std::log((double)std::max<uint64_t>(gLow, 16));
            uint64_t width = std::max<uint64_t>((uint64_t)(arraySize * lnp), 1ull << 30);
            width = std::min<uint64_t>(width, 1ull << 32);
            const uint64_t next = std::min<uint64_t>(gLow + width, high + 1);
            if ((st = rs_sieve_submit_range(sv, gLow, next)) != RS_OK) return fail("submit", st);
            if (++kernelCount % 20 == 0) progress_line(next);
            if ((st = paced_drain((long long)((double)(next - gLow) / lnp), false)) != RS_OK)
                return fail("collect", st);
            gLow = next;
        }
    }
    if ((st = paced_drain(0, true)) != RS_OK) return fail("collect", st);
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

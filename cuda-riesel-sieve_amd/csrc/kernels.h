// kernels.h -- job layout shared by the sieve kernels and their host side.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Record of one prime written by the set-up phase and read by the BSGS
// phase (64-bit words).  For p >= 2^50 residues are Montgomery forms
// (R = 2^64) unless marked plain.  For p < 2^50 ("fp": the BSGS phase
// multiplies in double precision) every residue but [3] is plain, held as
// a double (an exact integer in (-p, p)).
//   [0] p   [1] mp = -p^-1 mod 2^64   [2] fp: 1/p, else unused
//   [3] g^64 (Montgomery form, an integer in [0, p), in both cases)
//   [4] g^64 mod p (plain)
//   [5] its Shoup constant floor(g^64 2^64 / p) (fp: the double of the
//       second-entry bound (p >> 32) kmax + kmax + 1)
//   [6 .. 14)   g^c for c = 0..7          (g = b^Q)
//   [14 .. 22)  g^(8a) for a = 0..7       (plain)
//   [22]        p as a double
//   [23 .. 23+Q)        BH[r] = b^r * H * 2^32, H = g^(x0 + m - 1), negated
//                       for k*b^n+1
//   [23+Q .. 23+Q+2T)   G^t mod p (plain) and its Shoup constant (fp:
//                       unused), t < T, G = g^m, interleaved
//   [23+Q+2T .. +NW)    row masks: bit i of word w is set iff row 64w+i passes
//                       the Jacobi filter, (k/p) = +1 for even r and
//                       (k*b/p) = +1 for odd r (NW = ceil(nrows / 64))
//   then, for jobs whose baby table is shared by the NV wavefronts of a
//   workgroup (JobDev.nwv = NV > 0), at off_x:
//   [off_x .. +NV)      g^(64 w) for w < NV (Montgomery forms; fp: doubles)
//   [off_x + NV]        g^(64 NV), the baby-step stride, plain (fp: double)
//   [off_x + NV + 1]    its Shoup constant (fp: unused)
enum { REC_P = 0, REC_MP = 1, REC_PINV = 2, REC_G64 = 3, REC_GW = 4, REC_GWS = 5, REC_GP = 6,
       REC_G8P = 14, REC_PD = 22, REC_BH = 23 };

// Rows handled per compaction pass of the BSGS phase: bounds the LDS list
// of active rows whatever the job's row count.
enum { ROW_SLICE = 512 };

struct JobDev {
    uint32_t base;
    int Q;
    int m;        // baby steps, multiple of 64
    int T;        // giant steps per row: ceil((x1 - x0 + 1) / m), rounded up to tw
    int x0;       // floor(minN / Q)
    int minN, maxN;
    int nrows;
    int nk;       // distinct k
    int logNB;    // baby table: NB = 2^logNB buckets of 8 16-bit fingerprints
    uint32_t invT;  // ceil(2^32 / T) (T >= 2): idx / T for idx < 2^16
    int tw;         // giant-step tile: consecutive t of one row per lane, 1, 2 or 4
                    // (T is rounded up to a multiple of tw)
    uint32_t kmax;  // largest k: baby values below about p*kmax/2^32 get a
                    // second table entry (plus p) for the unreduced giant steps
    int plus;       // 1: k*b^n+1 (Sierpinski), 0: k*b^n-1
    int W;        // record words
    int off_gt;   // REC_BH + Q
    int off_j;    // off_gt + 2 T: the row masks
    int lds_words;  // 32-bit LDS words per wave (table, counts, Y, rows, record)
    int lds_words_idx;  // the same with the row list as uint16 row numbers
    int list_idx;   // 1: single-job launches list active rows by row number
                    // (RS_TILE_ROW_INDEX; at most ROW_SLICE rows)
    int tab_words;  // of which the table: 4 NB, a power of two
    int ys;         // Y row stride in entries: T, or T + 1 when T is even, so
                    // that rows of different r start on different LDS banks
    const uint4* kjac;        // per distinct k (ascending): odd part n, 2-adic
                              // valuation z, floor((2^64-1)/n) as lo, hi [nk]
    uint4 bjac;               // the same for the base
    const uint32_t* kjt;      // per distinct k: word offset of its quadratic-
                              // character table in jtab, or ~0u (no table)
    const uint32_t* jtab;     // bit x of a k's table: (x/n) = +1, x < n, n the
                              // odd part of k (and the prime-factor tables)
    const uint4* kfac;        // per distinct k without a table of its own:
                              // first factor entry, entry count, the largest
                              // prime factor of n, 1 (0: not factored)
    const uint4* fac;         // factor entries: prime q (or the cofactor),
                              // floor((2^64-1)/q) lo, hi, word offset of q's
                              // table in jtab (~0u: the cofactor, by the loop)
    const uint32_t* kw0;      // per distinct k: 64-row word holding its first row
    const uint64_t* kmask;    // per distinct k: its even-r rows in words w0 and
                              // w0 + 1, then its odd-r rows in the same words
    const uint2* rowent;      // per row: its giant-step entry (k, 8 r ys),
                              // in row order (ascending k, then r)
    const uint32_t* bitmap;   // optional n-list bitmaps (rows x words, MSB first)
    int words;                // 32-bit words per row bitmap
    int nwv;        // 0: one wavefront per prime with its own table; 4 or 8:
                    // the wavefronts of a workgroup share one prime's table
                    // (RS_TILE_GROUP4 / RS_TILE_GROUP8)
    int off_x;      // record words of the shared-table layout (nwv > 0)
    int lds_words_wg;  // 32-bit LDS words of one workgroup (nwv > 0)
    int sparse;        // 1: twice the buckets of the load rule (RS_TILE_SPARSE)
    int cubic;         // 1 (RS_TILE_CUBIC): for primes p = 1 (mod 3) below
                       // 2^50 the rows must also pass the cubic-residue test
                       // (Q a multiple of 3; per-wavefront layouts whose table
                       // holds the set-up's work area, cubic_area_bytes)
    uint32_t invnk;    // ceil(2^32 / nk) for nk > 1: j / nk for j < 2^16
    const uint32_t* kval;     // per distinct k (ascending): k
    const uint64_t* kmask3;   // per distinct k: its rows with r = 0, 1, 2 (mod 3)
                              // in words kw0 and kw0 + 1 (6 words)
};

// The set-up's cubic work area in a wavefront's table slot (bytes): 64 slot
// records of CUBIC_SLOT_BYTES (p, 1/p and the exponent (p-1)/3 of one prime,
// b^e and b^2e), then one class byte per (prime slot, k).
enum { CUBIC_SLOT_BYTES = 40, CUBIC_MAX_K = 1024 };
__host__ __device__ inline int cubic_area_bytes(int nk) { return 64 * CUBIC_SLOT_BYTES + 64 * nk; }

// A chunk of primes p = low + offs[i], i < count, of one job; group0 is
// its first 64-prime group in the launch's global numbering.  When
// count_dev is set (a single chunk whose primes were generated on the
// device), the count is read from there and `count` is only a bound.
struct ChunkDev {
    uint64_t low;
    const uint32_t* offs;
    long long count;
    long long group0;
    const uint32_t* count_dev;
    int job;
    int pad;
};

struct HitDev {
    uint64_t p;
    uint32_t k;
    uint32_t n;
    uint32_t job;
    uint32_t pad;
};

// Quadratic-character tables: k whose odd part n is at most JTAB_MAX_N get
// one, in order of k, while the job's tables stay within JTAB_MAX_WORDS.
// The odd parts of the other k are factored: the primes q <= FTAB_MAX_Q
// with an odd exponent get a table of (x/q) each (one per q, shared), the
// smallest first, while those tables stay within FTAB_MAX_WORDS; the rest
// (the product of the primes left) goes through the binary Jacobi loop.
enum : uint32_t {
    JTAB_MAX_N = 1u << 28,
    JTAB_MAX_WORDS = 1u << 26,
    FTAB_MAX_Q = 1u << 28,
    FTAB_MAX_WORDS = 1u << 27
};

static inline int rec_words(int Q, int T, int nrows, int nwv = 0) {
    return REC_BH + Q + 2 * T + (nrows + 63) / 64 + (nwv ? nwv + 2 : 0);
}

// LDS words of one wavefront: its table, placed apart from the rest so
// that it is aligned to its size,
//   4 NB                      NB buckets x 8 uint16 fingerprints (slots past
//                             a bucket's fill count hold stale values)
// and, from the base of the rest,
//   NB                        bucket fill counts, one uint32 per bucket
//   then 2 Q ys               Y entries (uint64), rows ys apart
//   then list_words           active rows of one slice: their entries (k,
//                             8 r ys), or with idx (single-job
//                             launches, entries in block-shared LDS) their
//                             uint16 row numbers
//   then 2 W                  the prime's record (uint64)
__host__ __device__ inline int list_words(int nrows, bool idx) {
    const int n = nrows < ROW_SLICE ? nrows : ROW_SLICE;  // one row slice
    return idx ? (n + 3) / 4 * 2 : 2 * n;                 // even: the record is 8-aligned
}
static inline int lds_words_per_wave(int logNB, int Q, int ys, int nrows, int W, bool idx) {
    const int NB = 1 << logNB;
    int w = 5 * NB + 2 * Q * ys + list_words(nrows, idx) + 2 * W;
    return (w + 3) & ~3;
}

// LDS words of one workgroup whose wavefronts share one prime's table
// (nwv > 0): the table first, at LDS address 0 (so aligned to its size),
//   4 NB                      NB buckets x 8 uint16 fingerprints (slots past
//                             a bucket's fill count hold stale values)
//   NB / 2                    bucket fill counts (uint16)
//   2 Q ys                    Y entries (uint64)
//   2 W                       the prime's record (uint64)
//   WG_MISC                   the group number and the chain flag
enum { WG_MISC = 8, GROUP_SLICE = 4096 };
static inline int lds_words_group(int logNB, int Q, int ys, int W) {
    const int NB = 1 << logNB;
    int w = 4 * NB + NB / 2 + 2 * Q * ys + 2 * W + WG_MISC;
    return (w + 3) & ~3;
}
// uint64 words of a shared-table workgroup's list of active rows: one slice
// of GROUP_SLICE rows, as uint16 row numbers from the slice's start.
__host__ __device__ inline int group_list_words(int nrows) {
    const int n = nrows < GROUP_SLICE ? nrows : GROUP_SLICE;
    return (n + 3) / 4;
}
// Global scratch of one shared-table workgroup, in uint64 words: the
// records of its 64 nwv primes, the low words of the current prime's m
// baby-step values (uint32, in order of j) for exact rescans, then its
// list of active rows.
static inline long long scratch_words_group(int nwv, int W, int m, int nrows) {
    return 64ll * nwv * W + (m + 1) / 2 + group_list_words(nrows);
}

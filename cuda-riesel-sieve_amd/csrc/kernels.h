// kernels.h -- job layout shared by the sieve kernels and their host side.
#pragma once

#include <stdint.h>

// Record of one prime written by the set-up kernel and read by the BSGS
// kernel (64-bit words).  All residues are Montgomery forms (R = 2^64).
//   [0] p   [1] mp = -p^-1 mod 2^64   [2] one = R mod p   [3] g^64
//   [4 .. 12)   g^c for c = 0..7          (g = b^Q)
//   [12 .. 20)  g^(8a) for a = 0..7
//   [20 .. 20+Q)        BH[r] = b^r * H * 2^32, H = g^(x0 + m - 1)
//   [20+Q .. 20+Q+T)    GT[t] = G^t, G = g^m
//   [20+Q+T .. +2*KW)   Jacobi bits: for each 64-k word w, word 2w has bit
//                       i set iff (k/p) = +1 and word 2w+1 iff (k*b/p) = +1
enum { REC_P = 0, REC_MP = 1, REC_ONE = 2, REC_G64 = 3, REC_GP = 4, REC_G8P = 12,
       REC_BH = 20 };

struct JobDev {
    uint32_t base;
    int Q;
    int m;        // baby steps, multiple of 64
    int T;        // giant steps per row: ceil((x1 - x0 + 1) / m)
    int x0;       // floor(minN / Q)
    int minN, maxN;
    int nrows;
    int nk;       // distinct k
    int logM;     // table slots M = 2^logM >= 2m
    int jbits;    // bit length of m: width of the j field of a table entry
    int W;        // record words
    int off_gt;   // REC_BH + Q
    int off_j;    // off_gt + T
    int lds_words;  // 32-bit LDS words per wave (table, Y, active list)
    const uint32_t* kvals;    // distinct k, ascending [nk]
    const uint32_t* rowinfo;  // per row: kidx | r << 16
    const uint32_t* rowk;     // per row: k
};

struct HitDev {
    uint64_t p;
    uint32_t k;
    uint32_t n;
};

static inline int rec_words(int Q, int T, int nk) {
    return REC_BH + Q + T + 2 * ((nk + 63) / 64);
}

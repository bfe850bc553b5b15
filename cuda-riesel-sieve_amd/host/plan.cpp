// plan.cpp -- subsequence planner: n = Q*x + r rows and the giant-step
// window.
//
// Follows kernel_v0.24.6.cu:1086-1255: Q is chosen among even values
// 2, 4, ..., 22 (or exactly the -Q value) by the least subs*R work, where
// subs = |{(k, n mod Q)}| and R = maxN/Q - minN/Q + 1 (:1107-1141); every
// (k, r) becomes a row with an MSB-first bitmap over x' = (n-r)/Q - minN/Q
// (:1160-1209), and rows are ordered by (k, r) (:1219-1255).  Differences:
// the reference keeps Q = 1 when no even Q beats K*(maxN-minN) (:1093), but
// its Jacobi filter (:521) needs n and r of equal parity, so the least-work
// even Q is taken instead; rows are sorted lexicographically even when the
// file does not list k in ascending order.
#include <algorithm>
#include <climits>
#include <cmath>
#include <utility>

#include "host.h"

extern "C" rs_status rs_plan_create(const rs_abcd* a, int Qin, rs_plan** out) {
    if (!a || !out) return RS_ERR_ARG;
    *out = nullptr;
    if (a->ks.empty()) return RS_ERR_ARG;
    if (Qin != 0 && (Qin < 2 || Qin % 2 != 0)) return RS_ERR_ARG;  // :756-759
    const int nk = (int)a->ks.size();
    const int minN = a->minN, maxN = a->maxN;
    long long minWork = (long long)nk * (maxN - minN);
    int bestQ = 0;
    long long bestEvenWork = LLONG_MAX;
    int bestEvenQ = 0;
    const int qlo = Qin ? Qin : 2, qhi = Qin ? Qin + 1 : 24;
    for (int Q = qlo; Q < qhi; Q += 2) {
        std::vector<char> seen((size_t)nk * Q, 0);
        long long subs = 0;
        for (int i = 0; i < nk; ++i)
            for (int j = a->kstart[i]; j < a->kstart[i + 1]; ++j) {
                char& s = seen[(size_t)i * Q + a->ns[j] % Q];
                if (!s) { s = 1; ++subs; }
            }
        const long long work = subs * (maxN / Q - minN / Q + 1);
        if (work < minWork) { minWork = work; bestQ = Q; }
        if (work < bestEvenWork) { bestEvenWork = work; bestEvenQ = Q; }
    }
    const int Q = bestQ ? bestQ : bestEvenQ;
    rs_plan* pl = new rs_plan();
    pl->base = a->base;
    pl->c = a->c;
    pl->Q = Q;
    pl->minN = minN;
    pl->maxN = maxN;
    pl->x0 = minN / Q;
    pl->x1 = maxN / Q;
    pl->words = (pl->x1 - pl->x0 + 1) / 32 + 1;
    std::vector<std::pair<int32_t, int32_t>> rows;
    rows.reserve(a->ns.size());
    for (int i = 0; i < nk; ++i)
        for (int j = a->kstart[i]; j < a->kstart[i + 1]; ++j)
            rows.push_back({a->ks[i], a->ns[j] % Q});
    std::sort(rows.begin(), rows.end());
    rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
    pl->rows_k.resize(rows.size());
    pl->rows_r.resize(rows.size());
    for (size_t i = 0; i < rows.size(); ++i) {
        pl->rows_k[i] = rows[i].first;
        pl->rows_r[i] = rows[i].second;
    }
    pl->bitmaps.assign(rows.size() * (size_t)pl->words, 0u);
    for (int i = 0; i < nk; ++i)
        for (int j = a->kstart[i]; j < a->kstart[i + 1]; ++j) {
            const int n = a->ns[j], r = n % Q;
            const size_t row =
                std::lower_bound(rows.begin(), rows.end(), std::make_pair(a->ks[i], r)) -
                rows.begin();
            const int xp = (n - r) / Q - pl->x0;
            pl->bitmaps[row * pl->words + xp / 32] |= 1u << (31 - (xp & 31));
        }
    *out = pl;
    return RS_OK;
}

extern "C" rs_status rs_plan_form(const rs_plan* pl, int* c) {
    if (!pl || !c) return RS_ERR_ARG;
    *c = pl->c;
    return RS_OK;
}

extern "C" rs_status rs_plan_info(const rs_plan* pl, uint32_t* base, int* Q, int* nrows,
                                  int* words, int* minN, int* maxN) {
    if (!pl) return RS_ERR_ARG;
    if (base) *base = pl->base;
    if (Q) *Q = pl->Q;
    if (nrows) *nrows = (int)pl->rows_k.size();
    if (words) *words = pl->words;
    if (minN) *minN = pl->minN;
    if (maxN) *maxN = pl->maxN;
    return RS_OK;
}

extern "C" rs_status rs_plan_rows(const rs_plan* pl, int32_t* rows_k, int32_t* rows_r,
                                  uint32_t* bitmaps) {
    if (!pl) return RS_ERR_ARG;
    if (rows_k) std::copy(pl->rows_k.begin(), pl->rows_k.end(), rows_k);
    if (rows_r) std::copy(pl->rows_r.begin(), pl->rows_r.end(), rows_r);
    if (bitmaps) std::copy(pl->bitmaps.begin(), pl->bitmaps.end(), bitmaps);
    return RS_OK;
}

// Bitmap membership, the lookup of kernel_v0.24.6.cu:1620-1647.
extern "C" int rs_plan_contains(const rs_plan* pl, uint32_t k, uint32_t n) {
    if (!pl || (int64_t)n < pl->minN || (int64_t)n > pl->maxN) return 0;
    const int r = (int)(n % (uint32_t)pl->Q);
    auto key = std::make_pair((int32_t)k, (int32_t)r);
    // rows are sorted by (k, r): binary search over the pair
    size_t lo = 0, hi = pl->rows_k.size();
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        auto v = std::make_pair(pl->rows_k[mid], pl->rows_r[mid]);
        if (v < key) lo = mid + 1; else hi = mid;
    }
    if (lo == pl->rows_k.size() || pl->rows_k[lo] != (int32_t)k || pl->rows_r[lo] != r)
        return 0;
    const int xp = (int)((n - r) / (uint32_t)pl->Q) - pl->x0;
    return (pl->bitmaps[lo * pl->words + xp / 32] >> (31 - (xp & 31))) & 1;
}

// kernel_v0.24.6.cu:1148-1158 (m a power of two).
extern "C" rs_status rs_plan_tminmax(const rs_plan* pl, int m, int* tmin, int* tmax) {
    if (!pl || m < 1 || (m & (m - 1)) != 0) return RS_ERR_ARG;
    int shift = 0;
    while ((1 << (shift + 1)) <= m) ++shift;
    if (tmin) *tmin = (pl->minN / pl->Q) >> shift;
    if (tmax) *tmax = (((pl->maxN / pl->Q) + 1) >> shift) + 1;
    return RS_OK;
}

int rsh::giant_T(const rs_plan* pl, int m, int tw) {
    if (tw < 1) tw = 1;
    const int X = pl->x1 - pl->x0 + 1;
    return std::max(2, ((X + m - 1) / m + tw - 1) / tw * tw);
}

bool rsh::shape_ok(const rs_plan* pl, int m, int tw) {
    if (!pl || m < 64 || m % 64 != 0 || m > (1 << 16) || pl->rows_k.empty()) return false;
    const long long T = giant_T(pl, m, tw);
    const long long ys = T + (T % 2 == 0 ? 1 : 0);
    const long long rows = std::min<long long>((long long)pl->rows_k.size(), 512);
    return T <= 65535 && pl->Q * ys < 65536 && rows * T + 4096 < 65536;
}

// Cost model for the table size m (multiple of 64): per prime the kernel
// does m baby steps and Q*T Y-table products (full 64x64 Montgomery, ~11
// 32-bit multiplies each), A*T giant steps (32x64 half products, ~5
// multiplies, A ~ rows/2 after the Jacobi filter) plus one table insert per
// baby step and one probe per giant step; T = ceil(X/m) with X the x-window.
// The table (2 slots per baby step, rounded up to a power of two, 4 bytes
// each) and the Y table live in LDS, so m is capped to keep a wave's share
// under 32 KiB.  Only shapes within the kernels' limits (shape_ok at the
// default tile 2) are considered; 0 when none is.
extern "C" int rs_plan_auto_m(const rs_plan* pl) {
    if (!pl) return 0;
    const double X = pl->x1 - pl->x0 + 1;
    const double A = 0.5 * pl->rows_k.size();
    int best = 0;
    double bestCost = 1e300;
    for (int m = 64; m <= 16384; m += 64) {
        if (!rsh::shape_ok(pl, m, 2)) continue;
        const int T = (int)std::ceil(X / m);
        int M = 1;
        while (M < 2 * m) M <<= 1;
        const double lds = 4.0 * M + 8.0 * pl->Q * T + 2.0 * pl->rows_k.size();
        if (lds > 32768) continue;
        const double cost = 14.0 * m + 11.0 * pl->Q * T + 9.0 * A * T;
        if (cost < bestCost) { bestCost = cost; best = m; }
    }
    return best;
}

extern "C" void rs_plan_free(rs_plan* pl) { delete pl; }

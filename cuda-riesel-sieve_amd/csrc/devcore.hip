// devcore.hip -- shared internals of the device sieve front ends (single
// jobs, batches, the server): job constants from a plan, the persistent
// kernel's launch resources, launching and draining hits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <map>
#include <vector>

#include "../host/host.h"
#include "devcore.h"
#include "kernels.h"

#define HIPCHK(x)                                 \
    do {                                          \
        if ((x) != hipSuccess) return RS_ERR_HIP; \
    } while (0)


extern "C" __global__ void rs_zero_u32(uint32_t*, int);
extern "C" __global__ void rs_jacobi_table(uint32_t n, uint32_t words, uint32_t* out);

namespace rsdev {

void free_job(DevJob& j) {
    (void)hipFree(j.d_kvals);
    (void)hipFree(j.d_kjt);
    (void)hipFree(j.d_jtab);
    j.d_kjt = j.d_jtab = nullptr;
    (void)hipFree(j.d_kfac);
    (void)hipFree(j.d_fac);
    j.d_kfac = j.d_fac = nullptr;
    (void)hipFree(j.d_kw0);
    (void)hipFree(j.d_kmask);
    (void)hipFree(j.d_kval);
    (void)hipFree(j.d_kmask3);
    j.d_kw0 = nullptr;
    j.d_kmask = nullptr;
    j.d_kval = nullptr;
    j.d_kmask3 = nullptr;
    (void)hipFree(j.d_rowent);
    (void)hipFree(j.d_bitmap);
    j.d_kvals = nullptr;
    j.d_rowent = nullptr;
    j.d_bitmap = nullptr;
    (void)hipFree(j.d_self);
    j.d_self = nullptr;
}

// Quadratic-character tables of a job: direct tables of the odd parts n of
// the distinct k (in order of k, n <= JTAB_MAX_N, within JTAB_MAX_WORDS),
// then for the other k the tables of the primes q <= FTAB_MAX_Q with an odd
// exponent in n (one per q, within FTAB_MAX_WORDS); what is left of n is its
// cofactor, for the binary loop.
struct JacobiLayout {
    std::vector<uint32_t> direct;  // per k: word offset of n's table, or ~0u
    std::vector<uint4> kfac;       // per k: first entry, count, largest prime factor, factored
    std::vector<uint4> fac;        // entries: q, floor((2^64-1)/q) lo, hi, table offset or ~0u
    std::vector<std::pair<uint32_t, uint32_t>> build;  // (n or q, word offset) to fill
    uint64_t words = 0;
};

static uint4 barrett_entry(uint32_t q, uint32_t off) {
    const uint64_t rn = ~0ull / q;
    return make_uint4(q, (uint32_t)rn, (uint32_t)(rn >> 32), off);
}

static JacobiLayout jacobi_layout(const std::vector<uint32_t>& kvals) {
    JacobiLayout L;
    L.direct.assign(kvals.size(), 0xffffffffu);
    L.kfac.assign(kvals.size(), make_uint4(0, 0, 0, 0));
    for (size_t i = 0; i < kvals.size(); ++i) {
        uint32_t n = kvals[i];
        while ((n & 1) == 0) n >>= 1;
        if (n < 3 || n > JTAB_MAX_N) continue;
        const uint64_t w = (n + 31) / 32;
        if (L.words + w > JTAB_MAX_WORDS) continue;
        L.direct[i] = (uint32_t)L.words;
        L.build.push_back({n, (uint32_t)L.words});
        L.words += w;
    }
    // factor the odd parts left: the odd primes of odd exponent (a square
    // factor's character is 1 for every p above it), the largest prime
    std::vector<std::vector<uint32_t>> odd(kvals.size());
    std::vector<uint32_t> small;  // odd primes up to sqrt(2^31)
    for (size_t i = 0; i < kvals.size(); ++i) {
        uint32_t x = kvals[i];
        while ((x & 1) == 0) x >>= 1;
        if (L.direct[i] != 0xffffffffu || x < 3) continue;
        if (small.empty()) {
            small = rsh::primes_upto(46341);
            small.erase(small.begin());  // drop 2
        }
        uint32_t top = 1;
        for (uint32_t q : small) {
            if ((uint64_t)q * q > x) break;
            int e = 0;
            while (x % q == 0) x /= q, ++e;
            if (e) top = q;
            if (e & 1) odd[i].push_back(q);
        }
        if (x > 1) odd[i].push_back(x), top = x;
        L.kfac[i] = make_uint4(0, 0, top, 1);
    }
    // tables for the smallest of those primes first (the most shared and
    // the cheapest), while they fit
    std::vector<uint32_t> qs;
    for (const auto& f : odd) qs.insert(qs.end(), f.begin(), f.end());
    std::sort(qs.begin(), qs.end());
    qs.erase(std::unique(qs.begin(), qs.end()), qs.end());
    std::map<uint32_t, uint32_t> qoff;
    uint64_t fwords = 0;
    for (uint32_t q : qs) {
        const uint64_t w = (q + 31) / 32;
        if (q > FTAB_MAX_Q || fwords + w > FTAB_MAX_WORDS) break;
        qoff.emplace(q, (uint32_t)L.words);
        L.build.push_back({q, (uint32_t)L.words});
        L.words += w;
        fwords += w;
    }
    for (size_t i = 0; i < kvals.size(); ++i) {
        if (!L.kfac[i].w) continue;
        L.kfac[i].x = (uint32_t)L.fac.size();
        uint32_t cof = 1;
        for (uint32_t q : odd[i]) {
            const auto it = qoff.find(q);
            if (it != qoff.end()) L.fac.push_back(barrett_entry(q, it->second));
            else cof *= q;
        }
        if (cof > 1) L.fac.push_back(barrett_entry(cof, 0xffffffffu));
        L.kfac[i].y = (uint32_t)L.fac.size() - L.kfac[i].x;
    }
    return L;
}

static std::vector<uint32_t> distinct_k(const rs_plan* pl) {
    std::vector<uint32_t> kvals;
    for (int32_t k : pl->rows_k)
        if (kvals.empty() || kvals.back() != (uint32_t)k) kvals.push_back((uint32_t)k);
    return kvals;
}

size_t job_device_bytes(const rs_plan* pl, bool dat) {
    const size_t rows = pl->rows_k.size();
    const JacobiLayout L = jacobi_layout(distinct_k(pl));
    // row tables, per k (at most one per row) Jacobi constants, masks and
    // table offsets, factor entries, the tables, and the n-list bitmaps
    // with dat
    return 8 * rows + 72 * rows + 16 * L.fac.size() + 4 * L.words +
           (dat ? 4 * rows * (size_t)pl->words : 0) + 256 + sizeof(JobDev);
}

// Jacobi constants of a 32-bit a > 0: odd part n, 2-adic valuation z and
// floor((2^64 - 1) / n) split in 32-bit halves.
uint4 jacobi_consts(uint32_t a) {
    uint32_t z = 0;
    while ((a & 1) == 0) a >>= 1, ++z;
    const uint64_t rn = ~0ull / a;
    return make_uint4(a, z, (uint32_t)rn, (uint32_t)(rn >> 32));
}

uint32_t magic_div(int d) {
    return (uint32_t)((0x100000000ull + d - 1) / d);  // d >= 2
}

// Fills the job constants from a plan, table size m (0: cost model) and
// giant-step tile tw (1, 2 or 4 items per lane; 0: 2); with dat the n-list
// bitmaps are uploaded too.
rs_status make_job(const rs_plan* pl, int m, int tw, bool dat, DevJob* dj) {
    if (!pl) return RS_ERR_ARG;
    if (m <= 0) m = rs_plan_auto_m(pl);
    const bool row_index = (tw & RS_TILE_ROW_INDEX) != 0;
    const int nwv = (tw & RS_TILE_GROUP8) ? 8 : (tw & RS_TILE_GROUP4) ? 4 : 0;
    if ((tw & RS_TILE_GROUP8) && (tw & RS_TILE_GROUP4)) return RS_ERR_ARG;
    // shared tables: m in whole strides of the workgroup's lanes, and at
    // most 16384 (bucket counts are 16-bit: at most 2m inserts per bucket)
    if (nwv && (row_index || m % (64 * nwv) != 0 || m > 16384)) return RS_ERR_ARG;
    const bool sparse = (tw & RS_TILE_SPARSE) != 0;
    const bool want_cubic = (tw & RS_TILE_CUBIC) != 0;
    tw &= ~(RS_TILE_ROW_INDEX | RS_TILE_GROUP4 | RS_TILE_GROUP8 | RS_TILE_SPARSE |
            RS_TILE_CUBIC);
    if (tw < 0) return RS_ERR_ARG;
    if (tw == 0) tw = 2;
    if (tw > 4) return RS_ERR_ARG;
    if (tw == 3) tw = 2;
    if (pl->Q < 2 || pl->Q % 2 != 0 || pl->minN < 1 || pl->rows_k.empty()) return RS_ERR_ARG;
    // index limits of the giant-step grid (rsh::shape_ok)
    if (!rsh::shape_ok(pl, m, tw)) return RS_ERR_ARG;
    const int nrows = (int)pl->rows_k.size();
    std::vector<uint32_t> kvals, rowinfo(nrows), rowk(nrows);
    for (int32_t k : pl->rows_k)
        if (kvals.empty() || kvals.back() != (uint32_t)k) kvals.push_back((uint32_t)k);
    if (kvals.size() > 65535) return RS_ERR_ARG;
    for (int i = 0; i < nrows; ++i) {
        const uint32_t kidx =
            (uint32_t)(std::lower_bound(kvals.begin(), kvals.end(), (uint32_t)pl->rows_k[i]) -
                       kvals.begin());
        rowinfo[i] = kidx | ((uint32_t)pl->rows_r[i] << 16);
        rowk[i] = (uint32_t)pl->rows_k[i];
    }
    JobDev& J = dj->J;
    J.base = pl->base;
    J.Q = pl->Q;
    J.m = m;
    J.x0 = pl->x0;
    // T in whole giant-step tiles, at least 2 (so idx / T is one high
    // product): the extra steps search x past x1, whose n exceed maxN and
    // are dropped
    J.T = rsh::giant_T(pl, m, tw);
    J.minN = pl->minN;
    J.maxN = pl->maxN;
    J.nrows = nrows;
    J.nk = (int)kvals.size();
    // buckets: average load between 1.5 and 3 of the 8 slots, counting the
    // second entries of baby values below p*kmax/2^32 (a kmax/2^32 share)
    const int entries = m + (int)(((uint64_t)m * kvals.back() >> 32) + 1);
    int logNB = 0;
    while ((1 << logNB) * 3 < entries) ++logNB;
    J.logNB = std::max(logNB, 2) + (sparse ? 1 : 0);
    J.sparse = sparse ? 1 : 0;
    J.ys = J.T + (J.T % 2 == 0 ? 1 : 0);
    J.invT = magic_div(J.T);
    J.tw = tw;
    J.kmax = kvals.back();
    J.plus = pl->c == 1;
    J.nwv = nwv;
    J.W = rec_words(J.Q, J.T, nrows, nwv);
    J.off_gt = REC_BH + J.Q;
    J.off_j = J.off_gt + 2 * J.T;
    J.off_x = J.off_j + (nrows + 63) / 64;
    J.lds_words = lds_words_per_wave(J.logNB, J.Q, J.ys, nrows, J.W, false);
    J.lds_words_idx = lds_words_per_wave(J.logNB, J.Q, J.ys, nrows, J.W, true);
    J.lds_words_wg = nwv ? lds_words_group(J.logNB, J.Q, J.ys, J.W) : 0;
    J.list_idx = row_index && nrows <= ROW_SLICE ? 1 : 0;
    J.tab_words = 4 << J.logNB;
    // per k: the 64-row word of its first row and its even-r / odd-r rows
    // in that word and the next (a k has at most Q < 64 rows)
    // (and its rows by r mod 3, for the cubic-residue test)
    std::vector<uint32_t> kw0(kvals.size(), 0xffffffffu);
    std::vector<uint64_t> kmask(4 * kvals.size(), 0), kmask3(6 * kvals.size(), 0);
    for (int i = 0; i < nrows; ++i) {
        const uint32_t kidx = rowinfo[i] & 0xffffu, r = rowinfo[i] >> 16;
        if (kw0[kidx] == 0xffffffffu) kw0[kidx] = (uint32_t)(i >> 6);
        const uint32_t dw = (uint32_t)(i >> 6) - kw0[kidx];
        if (dw > 1) return RS_ERR_ARG;  // rows of one k are contiguous and < 64
        kmask[4 * kidx + 2 * (r & 1) + dw] |= 1ull << (i & 63);
        kmask3[6 * kidx + 2 * (r % 3) + dw] |= 1ull << (i & 63);
    }
    // the cubic-residue test: n = r (mod 3) when 3 | Q; its work area must
    // fit the wavefront's table slot during the set-up
    J.cubic = want_cubic && J.Q % 3 == 0 && nwv == 0 && J.nk <= CUBIC_MAX_K &&
                      cubic_area_bytes(J.nk) <= 4 * J.tab_words
                  ? 1
                  : 0;
    J.invnk = J.nk > 1 ? magic_div(J.nk) : 0;
    std::vector<uint4> kjac(kvals.size());
    for (size_t i = 0; i < kvals.size(); ++i) kjac[i] = jacobi_consts(kvals[i]);
    J.bjac = jacobi_consts(pl->base);
    HIPCHK(hipMalloc(&dj->d_kvals, kjac.size() * sizeof(uint4)));
    // giant-step entries of the rows: k and the byte offset of r's Y run
    std::vector<uint2> rowent(nrows);
    for (int i = 0; i < nrows; ++i)
        rowent[i] = make_uint2(rowk[i], (rowinfo[i] >> 16) * (uint32_t)J.ys * 8u);
    HIPCHK(hipMalloc(&dj->d_rowent, rowent.size() * sizeof(uint2)));
    HIPCHK(hipMemcpy(dj->d_kvals, kjac.data(), kjac.size() * sizeof(uint4),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dj->d_rowent, rowent.data(), rowent.size() * sizeof(uint2),
                     hipMemcpyHostToDevice));
    J.kjac = dj->d_kvals;
    {
        const JacobiLayout TL = jacobi_layout(kvals);
        HIPCHK(hipMalloc(&dj->d_kjt, TL.direct.size() * 4));
        HIPCHK(hipMemcpy(dj->d_kjt, TL.direct.data(), TL.direct.size() * 4,
                         hipMemcpyHostToDevice));
        HIPCHK(hipMalloc(&dj->d_kfac, TL.kfac.size() * sizeof(uint4)));
        HIPCHK(hipMemcpy(dj->d_kfac, TL.kfac.data(), TL.kfac.size() * sizeof(uint4),
                         hipMemcpyHostToDevice));
        HIPCHK(hipMalloc(&dj->d_fac, std::max<size_t>(TL.fac.size(), 1) * sizeof(uint4)));
        if (!TL.fac.empty())
            HIPCHK(hipMemcpy(dj->d_fac, TL.fac.data(), TL.fac.size() * sizeof(uint4),
                             hipMemcpyHostToDevice));
        if (hipMalloc(&dj->d_jtab, std::max<uint64_t>(TL.words, 1) * 4) != hipSuccess)
            return RS_ERR_NOMEM;
        hipStream_t bs = nullptr;
        HIPCHK(hipStreamCreateWithFlags(&bs, hipStreamNonBlocking));
        for (const auto& t : TL.build) {
            const uint32_t w = (t.first + 31) / 32;
            hipLaunchKernelGGL(rs_jacobi_table, dim3((w + 255) / 256), dim3(256), 0, bs,
                               t.first, w, dj->d_jtab + t.second);
        }
        const hipError_t e1 = hipGetLastError(), e2 = hipStreamSynchronize(bs);
        (void)hipStreamDestroy(bs);
        if (e1 != hipSuccess || e2 != hipSuccess) return RS_ERR_HIP;
        J.kjt = dj->d_kjt;
        J.jtab = dj->d_jtab;
        J.kfac = dj->d_kfac;
        J.fac = dj->d_fac;
    }
    HIPCHK(hipMalloc(&dj->d_kw0, kw0.size() * 4));
    HIPCHK(hipMalloc(&dj->d_kmask, kmask.size() * 8));
    HIPCHK(hipMalloc(&dj->d_kval, kvals.size() * 4));
    HIPCHK(hipMalloc(&dj->d_kmask3, kmask3.size() * 8));
    HIPCHK(hipMemcpy(dj->d_kw0, kw0.data(), kw0.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dj->d_kmask, kmask.data(), kmask.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dj->d_kval, kvals.data(), kvals.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dj->d_kmask3, kmask3.data(), kmask3.size() * 8, hipMemcpyHostToDevice));
    J.kw0 = dj->d_kw0;
    J.kmask = dj->d_kmask;
    J.kval = dj->d_kval;
    J.kmask3 = dj->d_kmask3;
    J.rowent = dj->d_rowent;
    J.words = pl->words;
    J.bitmap = nullptr;
    if (dat) {
        const size_t bytes = pl->bitmaps.size() * 4;
        HIPCHK(hipMalloc(&dj->d_bitmap, std::max<size_t>(bytes, 4)));
        HIPCHK(hipMemcpy(dj->d_bitmap, pl->bitmaps.data(), bytes, hipMemcpyHostToDevice));
        J.bitmap = dj->d_bitmap;
    }
    // the constants themselves in device memory: single-job launches read
    // them from there like batched ones (not as a kernel argument, whose
    // words would stay in scalar registers for the whole kernel)
    HIPCHK(hipMalloc(&dj->d_self, sizeof(JobDev)));
    HIPCHK(hipMemcpy(dj->d_self, &J, sizeof(JobDev), hipMemcpyHostToDevice));
    return RS_OK;
}

void free_launcher(Launcher& L) {
    (void)hipFree(L.scratch);
    (void)hipFree(L.work);
    (void)hipFree(L.hits);
    (void)hipFree(L.hit_count);
    L.scratch = nullptr;
    L.work = nullptr;
    L.hits = nullptr;
    L.hit_count = nullptr;
}

// Sizes the grid for lds_words per wave and record size W (both maxima over
// the jobs of the launch) and allocates scratch, counter and hits.
static const void* kernel_of(bool w5, bool row_index) {
    return row_index ? reinterpret_cast<const void*>(rs_sieve_kernel5_rix)
           : w5      ? reinterpret_cast<const void*>(rs_sieve_kernel5)
                     : reinterpret_cast<const void*>(rs_sieve_kernel);
}

rs_status launch_geometry(int device, int lds_words, Launcher* L, int shared_words,
                          bool row_index) {
    L->device = device;
    int max_lds = 0, cus = 0;
    HIPCHK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    const size_t per_wave = (size_t)lds_words * 4, shared = (size_t)shared_words * 4;
    if (lds_words < 1 || shared_words < 0 || per_wave + shared > (size_t)max_lds)
        return RS_ERR_ARG;
    // waves of a block are independent (one table each); 1 to 4 per block,
    // whichever keeps the most waves resident per CU (ties: more per block,
    // so the block-shared part is loaded fewer times).  rs_sieve_kernel5 only
    // when it keeps a whole wavefront per SIMD more resident (4 per CU):
    // its tighter register budget costs more than 1-3 extra wavefronts gain
    // (measured on sr_745.abcd: 18 against 16 resident was 7% slower)
    int w_of[2] = {0, 0}, wpb_of[2] = {1, 1}, cu_of[2] = {1, 1};
    for (bool w5 : {false, true}) {
        if (row_index && !w5) continue;
        // the kernel addresses its tables from LDS address 0: it must have
        // no static LDS before the dynamic region
        hipFuncAttributes fa{};
        HIPCHK(hipFuncGetAttributes(&fa, kernel_of(w5, row_index)));
        if (fa.sharedSizeBytes != 0) return RS_ERR_HIP;
        // the attribute is per function: always the device limit, so
        // launches of different LDS sizes never restrict one another
        HIPCHK(hipFuncSetAttribute(kernel_of(w5, row_index),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
        for (int wpb = 1; wpb <= kMaxWavesPerBlock; ++wpb) {
            const size_t bytes = per_wave * wpb + shared;
            if (bytes > (size_t)max_lds) break;
            int per_cu = 0;
            HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, kernel_of(w5, row_index), 64 * wpb, bytes));
            if (per_cu * wpb >= w_of[w5]) {
                w_of[w5] = per_cu * wpb;
                wpb_of[w5] = wpb;
                cu_of[w5] = per_cu;
            }
        }
    }
    const bool best_w5 = row_index || w_of[1] >= w_of[0] + 4;
    const int best_wpb = wpb_of[best_w5], best_cu = cu_of[best_w5];
    L->wpb = best_wpb;
    L->w5 = best_w5;
    L->row_index = row_index;
    L->lds_words = lds_words;
    L->shared_words = shared_words;
    L->lds_bytes = per_wave * L->wpb + shared;
    L->grid = std::max(1, cus * std::max(1, best_cu));
    return RS_OK;
}

static const void* group_kernel(int nwv) {
    return nwv == 8 ? reinterpret_cast<const void*>(rs_sieve_group8_kernel)
                    : reinterpret_cast<const void*>(rs_sieve_group4_kernel);
}

rs_status launch_geometry_group(int device, const JobDev& J, Launcher* L) {
    if (J.nwv != 4 && J.nwv != 8) return RS_ERR_ARG;
    L->device = device;
    int max_lds = 0, cus = 0;
    HIPCHK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    const size_t bytes = (size_t)J.lds_words_wg * 4;
    if (bytes > (size_t)max_lds) return RS_ERR_ARG;
    const void* k = group_kernel(J.nwv);
    hipFuncAttributes fa{};
    HIPCHK(hipFuncGetAttributes(&fa, k));
    if (fa.sharedSizeBytes != 0) return RS_ERR_HIP;  // the table sits at LDS address 0
    HIPCHK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
    int per_cu = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 64 * J.nwv, bytes));
    if (per_cu < 1) return RS_ERR_ARG;
    L->nwv = J.nwv;
    L->wpb = J.nwv;
    L->w5 = false;
    L->row_index = false;
    L->lds_words = J.lds_words_wg;
    L->tab_words = J.tab_words;
    L->shared_words = 0;
    L->lds_bytes = bytes;
    L->grid = cus * per_cu;
    L->scratch_words = (int)scratch_words_group(J.nwv, J.W, J.m, J.nrows);
    return RS_OK;
}

static rs_status alloc_buffers(Launcher* L, size_t scratch_words_total, int64_t hit_capacity);

rs_status init_launcher_group(Launcher* L, int device, const JobDev& J, int64_t hit_capacity) {
    rs_status st = launch_geometry_group(device, J, L);
    if (st != RS_OK) return st;
    return alloc_buffers(L, (size_t)L->grid * L->scratch_words, hit_capacity);
}

static rs_status alloc_buffers(Launcher* L, size_t scratch_words_total, int64_t hit_capacity) {
    if (hipMalloc(&L->scratch, scratch_words_total * 8) != hipSuccess) return RS_ERR_NOMEM;
    HIPCHK(hipMalloc(&L->work, 8));
    L->hit_cap = (uint32_t)hit_capacity;
    if (hipMalloc(&L->hits, (size_t)hit_capacity * sizeof(HitDev)) != hipSuccess)
        return RS_ERR_NOMEM;
    HIPCHK(hipMalloc(&L->hit_count, 4));
    HIPCHK(hipMemset(L->hit_count, 0, 4));
    return RS_OK;
}

rs_status init_launcher(Launcher* L, int device, int lds_words, int tab_words, int W,
                        int64_t hit_capacity, int shared_words, bool row_index) {
    rs_status st = launch_geometry(device, lds_words, L, shared_words, row_index);
    if (st != RS_OK) return st;
    L->tab_words = tab_words;
    L->scratch_words = 64 * W;
    return alloc_buffers(L, (size_t)L->grid * L->wpb * L->scratch_words, hit_capacity);
}

rs_status init_multi(MultiLauncher* M, int device, int w_cap, int64_t hit_capacity) {
    M->w_cap = w_cap;
    // the smallest layout gives the most resident waves (of either kernel):
    // size scratch for it
    M->max_waves = 0;
    for (bool row_index : {false, true}) {
        Launcher g;
        rs_status st = launch_geometry(device, 64, &g, 0, row_index);
        if (st != RS_OK) return st;
        M->max_waves = std::max(M->max_waves, (size_t)g.grid * g.wpb);
    }
    M->base.device = device;
    return alloc_buffers(&M->base, M->max_waves * 64 * (size_t)w_cap, hit_capacity);
}

rs_status plan_classes(MultiLauncher* M, const std::vector<JobDev>& jobs,
                       std::vector<ChunkDev>* chunks, std::vector<ClassRange>* out) {
    out->clear();
    // occupancy key of each job's own layout
    std::vector<std::pair<long long, int>> key(jobs.size());
    for (size_t j = 0; j < jobs.size(); ++j) {
        if (jobs[j].nwv > 0) {  // shared-table jobs run as single launches
            key[j] = {-1, 0};
            continue;
        }
        auto it = M->geo.find(jobs[j].lds_words);
        if (it == M->geo.end()) {
            Launcher g;
            rs_status st = launch_geometry(M->base.device, jobs[j].lds_words, &g);
            if (st != RS_OK) return st;
            it = M->geo.emplace(jobs[j].lds_words, g).first;
        }
        if (jobs[j].W > M->w_cap) return RS_ERR_ARG;
        // occupancy, then table size: a class then takes its members' table
        // and largest rest, whose sum is at most its largest member's layout
        key[j] = {((long long)it->second.grid * kMaxWavesPerBlock + it->second.wpb) << 24 |
                      jobs[j].tab_words,
                  0};
    }
    // classes in order of decreasing occupancy; a class takes the largest
    // table, the largest rest of the layout and the largest record of its
    // members (their sum may give a grid of fewer waves than the members')
    std::vector<long long> keys;
    for (auto& k : key)
        if (k.first >= 0) keys.push_back(k.first);
    std::sort(keys.begin(), keys.end(), std::greater<long long>());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    std::vector<ChunkDev> sorted;
    sorted.reserve(chunks->size());
    for (long long k : keys) {
        ClassRange cr{};
        int tabw = 0, rest = 0, W = 0;
        for (size_t j = 0; j < jobs.size(); ++j)
            if (key[j].first == k) {
                tabw = std::max(tabw, jobs[j].tab_words);
                rest = std::max(rest, jobs[j].lds_words - jobs[j].tab_words);
                W = std::max(W, jobs[j].W);
            }
        const int lds = tabw + rest;
        cr.chunk0 = (int)sorted.size();
        long long g = 0;
        for (const ChunkDev& c : *chunks) {
            if (key[c.job].first != k || c.count == 0) continue;
            ChunkDev d = c;
            d.group0 = g;
            g += (c.count + 63) / 64;
            sorted.push_back(d);
        }
        cr.nchunks = (int)sorted.size() - cr.chunk0;
        cr.ngroups = g;
        if (cr.nchunks == 0) continue;
        auto it = M->geo.find(lds);
        if (it == M->geo.end()) {
            Launcher g;
            rs_status st = launch_geometry(M->base.device, lds, &g);
            if (st != RS_OK) return st;
            it = M->geo.emplace(lds, g).first;
        }
        cr.L = it->second;
        cr.L.tab_words = tabw;
        cr.L.scratch_words = 64 * W;
        cr.L.scratch = M->base.scratch;
        cr.L.work = M->base.work;
        cr.L.hits = M->base.hits;
        cr.L.hit_count = M->base.hit_count;
        cr.L.hit_cap = M->base.hit_cap;
        if ((size_t)cr.L.grid * cr.L.wpb * cr.L.scratch_words > M->max_waves * 64 * (size_t)M->w_cap)
            return RS_ERR_ARG;
        out->push_back(cr);
    }
    *chunks = std::move(sorted);
    return RS_OK;
}

rs_status plan_launches(MultiLauncher* M, const std::vector<JobDev>& jobs,
                        std::vector<ChunkDev>* chunks, std::vector<SingleLaunch>* singles,
                        std::vector<ClassRange>* cls) {
    singles->clear();
    cls->clear();
    std::vector<ChunkDev> rest;
    size_t gwords = 0;  // scratch the shared-table launches need
    for (const ChunkDev& c : *chunks) {
        if (c.count == 0) continue;
        const JobDev& J = jobs[c.job];
        if (J.nwv > 0) {  // shared tables: always a launch of their own
            SingleLaunch s;
            rs_status st = launch_geometry_group(M->base.device, J, &s.L);
            if (st != RS_OK) return st;
            gwords = std::max(gwords, (size_t)s.L.grid * s.L.scratch_words);
            s.L.work = M->base.work;
            s.L.hits = M->base.hits;
            s.L.hit_count = M->base.hit_count;
            s.L.hit_cap = M->base.hit_cap;
            s.c0 = c;
            s.c0.group0 = 0;
            singles->push_back(s);
            continue;
        }
        const int shared = J.nrows <= ROW_SLICE ? (2 * J.nrows + 3) & ~3 : 0;
        SingleLaunch s;
        const bool rix = shared && J.list_idx;
        rs_status st = launch_geometry(M->base.device, rix ? J.lds_words_idx : J.lds_words,
                                       &s.L, shared, rix);
        if (st != RS_OK) return st;
        const long long groups = (c.count + 63) / 64;
        if (groups < (long long)kSingleGroupsPerWave * s.L.grid * s.L.wpb ||
            J.W > M->w_cap) {
            rest.push_back(c);
            continue;
        }
        s.L.tab_words = J.tab_words;
        s.L.scratch_words = 64 * J.W;
        s.L.scratch = M->base.scratch;
        s.L.work = M->base.work;
        s.L.hits = M->base.hits;
        s.L.hit_count = M->base.hit_count;
        s.L.hit_cap = M->base.hit_cap;
        if ((size_t)s.L.grid * s.L.wpb * s.L.scratch_words > M->max_waves * 64 * (size_t)M->w_cap) {
            rest.push_back(c);
            continue;
        }
        s.c0 = c;
        s.c0.group0 = 0;
        singles->push_back(s);
    }
    if (gwords > M->gscratch_words) {
        (void)hipFree(M->gscratch);
        M->gscratch = nullptr;
        M->gscratch_words = 0;
        if (hipMalloc(&M->gscratch, gwords * 8) != hipSuccess) return RS_ERR_NOMEM;
        M->gscratch_words = gwords;
    }
    for (SingleLaunch& s : *singles)
        if (s.L.nwv > 0) s.L.scratch = M->gscratch;
    *chunks = std::move(rest);
    return chunks->empty() ? RS_OK : plan_classes(M, jobs, chunks, cls);
}

rs_status launch_singles(const std::vector<SingleLaunch>& singles, hipStream_t stream,
                         const JobDev* jobs) {
    for (const SingleLaunch& s : singles) {
        rs_status st = launch(s.L, stream, jobs, s.c0, nullptr, 1, groups_of(s.L, s.c0.count));
        if (st != RS_OK) return st;
    }
    return RS_OK;
}

rs_status launch_classes(const std::vector<ClassRange>& cls, hipStream_t stream,
                         const JobDev* jobs, const ChunkDev* d_chunks) {
    for (const ClassRange& c : cls) {
        ChunkDev c0{};
        rs_status st = launch(c.L, stream, jobs, c0, d_chunks + c.chunk0, c.nchunks,
                              c.ngroups);
        if (st != RS_OK) return st;
    }
    return RS_OK;
}

void free_multi(MultiLauncher& M) {
    free_launcher(M.base);
    (void)hipFree(M.gscratch);
    M.gscratch = nullptr;
    M.gscratch_words = 0;
    M.geo.clear();
}

rs_status launch(const Launcher& L, hipStream_t stream, const JobDev* jobs, const ChunkDev& c0,
                 const ChunkDev* chunks, int nchunks, long long ngroups) {
    hipLaunchKernelGGL(rs_zero_u32, dim3(1), dim3(64), 0, stream, (uint32_t*)L.work, 2);
    HIPCHK(hipGetLastError());
    // a block takes groups of 64 primes per wavefront (shared tables: of
    // 64 nwv primes per block)
    const int grid = (int)std::min<long long>(
        L.grid, L.nwv > 0 ? ngroups : (ngroups + L.wpb - 1) / L.wpb);
    hipLaunchKernelGGL(L.nwv == 8    ? rs_sieve_group8_kernel
                       : L.nwv == 4  ? rs_sieve_group4_kernel
                       : L.row_index ? rs_sieve_kernel5_rix
                       : L.w5        ? rs_sieve_kernel5
                                     : rs_sieve_kernel,
                       dim3(std::max(1, grid)),
                       dim3(64 * L.wpb), L.lds_bytes, stream, c0, jobs, chunks, nchunks, ngroups,
                       L.lds_words, L.tab_words, L.shared_words,
                       L.scratch_words, L.scratch, L.work, L.hits, L.hit_count, L.hit_cap);
    HIPCHK(hipGetLastError());
    return RS_OK;
}

// Copies the hits out (appending to `out`) and zeroes the device counter.
rs_status drain_hits(const Launcher& L, hipStream_t stream, std::vector<HitDev>* out,
                     int64_t* lost) {
    HIPCHK(hipStreamSynchronize(stream));
    uint32_t n = 0;
    HIPCHK(hipMemcpy(&n, L.hit_count, 4, hipMemcpyDeviceToHost));
    const uint32_t got = std::min(n, L.hit_cap);
    if (n > L.hit_cap) *lost += (int64_t)(n - L.hit_cap);
    if (got) {
        const size_t old = out->size();
        out->resize(old + got);
        HIPCHK(hipMemcpy(out->data() + old, L.hits, (size_t)got * sizeof(HitDev),
                         hipMemcpyDeviceToHost));
    }
    hipLaunchKernelGGL(rs_zero_u32, dim3(1), dim3(64), 0, stream, L.hit_count, 1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream));
    return RS_OK;
}

bool hit_less(const HitDev& a, const HitDev& b) {
    if (a.job != b.job) return a.job < b.job;
    if (a.p != b.p) return a.p < b.p;
    if (a.k != b.k) return a.k < b.k;
    return a.n < b.n;
}

}  // namespace

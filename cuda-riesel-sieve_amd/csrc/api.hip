// api.hip -- C ABI of the device sieve (include/rsieve.h, "device sieve").
//
// Replaces the reference's device set-up and launch sequence
// (kernel_v0.24.6.cu:1367-1512 allocations and constant copies,
// :1537-1611 per-batch copy, memsets, addKernel1 launch and result copy):
//  * job constants and the row table are copied once at create;
//  * a batch runs as two kernels on one compute stream: set-up (one thread
//    per prime) and BSGS (one wavefront per prime, table in LDS);
//  * host prime batches go through two pinned buffers on a copy stream, so
//    the host may refill one while the other is copied or sieved (the
//    reference reuses one pinned buffer, racing :1573 against :1592);
//  * hits go to a device append buffer (every hit, not one per prime) and
//    are returned sorted at collect.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../host/host.h"
#include "kernels.h"

extern "C" __global__ void rs_prep_kernel(JobDev J, uint64_t low, const uint32_t* offs,
                                          long long count, uint64_t* rec);
extern "C" __global__ void rs_bsgs_kernel(JobDev J, const uint64_t* rec, long long count,
                                          HitDev* hits, uint32_t* hit_count, uint32_t hit_cap);
extern "C" __global__ void rs_test_montmul(const uint64_t*, const uint64_t*, const uint64_t*,
                                           uint64_t*, long long);
extern "C" __global__ void rs_test_halfmul(const uint32_t*, const uint64_t*, const uint64_t*,
                                           uint64_t*, long long);
extern "C" __global__ void rs_test_jacobi(const uint32_t*, const uint64_t*, int32_t*, long long);
extern "C" __global__ void rs_zero_u32(uint32_t*, int);
extern "C" __global__ void rs_mad_probe(uint64_t*, int);

#define HIPCHK(x)                                 \
    do {                                          \
        if ((x) != hipSuccess) return RS_ERR_HIP; \
    } while (0)

namespace {

const int kMaxWavesPerBlock = 4;

struct Timing {
    hipEvent_t a, b, c;  // prep start, bsgs start, bsgs end
    long long primes;
};

}  // namespace

struct rs_sieve {
    int device = 0;
    JobDev J{};
    uint32_t* d_kvals = nullptr;
    uint32_t* d_rowinfo = nullptr;
    uint32_t* d_rowk = nullptr;
    uint64_t* d_rec = nullptr;
    int64_t max_batch = 0;
    HitDev* d_hits = nullptr;
    uint32_t* d_hit_count = nullptr;
    uint32_t hit_cap = 0;
    hipStream_t stream = nullptr, copy_stream = nullptr;
    uint32_t* h_pin[2] = {nullptr, nullptr};
    uint32_t* d_offs[2] = {nullptr, nullptr};
    hipEvent_t copied[2] = {nullptr, nullptr};
    hipEvent_t consumed[2] = {nullptr, nullptr};
    bool consumed_valid[2] = {false, false};
    int buf = 0;
    int bsgs_grid = 0;
    int waves_per_block = 1;
    size_t lds_bytes = 0;
    std::vector<Timing> timings;
    double acc_bsgs_ms = 0, acc_prep_ms = 0;
    long long acc_primes = 0;
    std::vector<rs_factor> held;
    int64_t lost = 0;
};

static void free_all(rs_sieve* s) {
    if (!s) return;
    hipSetDevice(s->device);
    if (s->stream) hipStreamSynchronize(s->stream);
    if (s->copy_stream) hipStreamSynchronize(s->copy_stream);
    (void)hipFree(s->d_kvals);
    (void)hipFree(s->d_rowinfo);
    (void)hipFree(s->d_rowk);
    (void)hipFree(s->d_rec);
    (void)hipFree(s->d_hits);
    (void)hipFree(s->d_hit_count);
    for (int b = 0; b < 2; ++b) {
        if (s->h_pin[b]) hipHostFree(s->h_pin[b]);
        (void)hipFree(s->d_offs[b]);
        if (s->copied[b]) hipEventDestroy(s->copied[b]);
        if (s->consumed[b]) hipEventDestroy(s->consumed[b]);
    }
    for (auto& t : s->timings) {
        hipEventDestroy(t.a);
        hipEventDestroy(t.b);
        hipEventDestroy(t.c);
    }
    if (s->stream) hipStreamDestroy(s->stream);
    if (s->copy_stream) hipStreamDestroy(s->copy_stream);
    delete s;
}

extern "C" int rs_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// Fills the job constants from a plan and table size m.
static rs_status make_job(const rs_plan* pl, int m, JobDev* J, std::vector<uint32_t>* kvals,
                          std::vector<uint32_t>* rowinfo, std::vector<uint32_t>* rowk) {
    if (m <= 0) m = rs_plan_auto_m(pl);
    if (m < 64 || m % 64 != 0 || m > (1 << 16)) return RS_ERR_ARG;
    if (pl->Q < 2 || pl->Q % 2 != 0 || pl->minN < 1 || pl->rows_k.empty()) return RS_ERR_ARG;
    const int nrows = (int)pl->rows_k.size();
    kvals->clear();
    for (int32_t k : pl->rows_k)
        if (kvals->empty() || kvals->back() != (uint32_t)k) kvals->push_back((uint32_t)k);
    if (kvals->size() > 65535) return RS_ERR_ARG;
    rowinfo->resize(nrows);
    rowk->resize(nrows);
    for (int i = 0; i < nrows; ++i) {
        const uint32_t kidx =
            (uint32_t)(std::lower_bound(kvals->begin(), kvals->end(), (uint32_t)pl->rows_k[i]) -
                       kvals->begin());
        (*rowinfo)[i] = kidx | ((uint32_t)pl->rows_r[i] << 16);
        (*rowk)[i] = (uint32_t)pl->rows_k[i];
    }
    J->base = pl->base;
    J->Q = pl->Q;
    J->m = m;
    J->x0 = pl->x0;
    J->T = (pl->x1 - pl->x0 + 1 + m - 1) / m;
    J->minN = pl->minN;
    J->maxN = pl->maxN;
    J->nrows = nrows;
    J->nk = (int)kvals->size();
    // buckets: average load between 1.5 and 3 of the 8 slots
    int logNB = 0;
    while ((1 << logNB) * 3 < m) ++logNB;
    J->logNB = std::max(logNB, 2);
    J->invT = J->T == 1 ? 0u : (uint32_t)((0x100000000ull + J->T - 1) / J->T);
    if (J->Q * J->T >= 65536 || J->T > 65535) return RS_ERR_ARG;
    J->W = rec_words(J->Q, J->T, J->nk);
    J->off_gt = REC_BH + J->Q;
    J->off_j = J->off_gt + J->T;
    const int words = lds_words_per_wave(J->logNB, J->Q, J->T, nrows, J->W);
    J->lds_words = words;
    return RS_OK;
}

extern "C" rs_status rs_sieve_create(int device, const rs_plan* pl, int m, int64_t max_batch,
                                     int64_t hit_capacity, rs_sieve** out) {
    if (!pl || !out || max_batch < 1 || hit_capacity < 1 || hit_capacity > 0x7fffffff)
        return RS_ERR_ARG;
    *out = nullptr;
    rs_sieve* s = new rs_sieve();
    s->device = device;
    std::vector<uint32_t> kvals, rowinfo, rowk;
    rs_status st = make_job(pl, m, &s->J, &kvals, &rowinfo, &rowk);
    if (st != RS_OK) { delete s; return st; }
    if (hipSetDevice(device) != hipSuccess) { delete s; return RS_ERR_HIP; }
    // waves of a block are independent (one table each); use up to 4 per
    // block while their tables fit the per-block LDS limit
    int max_lds = 0;
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) !=
        hipSuccess) {
        delete s;
        return RS_ERR_HIP;
    }
    const size_t per_wave = (size_t)s->J.lds_words * 4;
    if (per_wave > (size_t)max_lds) { delete s; return RS_ERR_ARG; }
    s->waves_per_block = (int)std::min<size_t>(kMaxWavesPerBlock, (size_t)max_lds / per_wave);
    s->lds_bytes = per_wave * s->waves_per_block;
    if (s->lds_bytes > 65536 &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(rs_bsgs_kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)s->lds_bytes) != hipSuccess) {
        delete s;
        return RS_ERR_HIP;
    }
#define CK(x)                        \
    do {                             \
        if ((x) != hipSuccess) {     \
            free_all(s);             \
            return RS_ERR_HIP;       \
        }                            \
    } while (0)
    CK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    CK(hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking));
    CK(hipMalloc(&s->d_kvals, kvals.size() * 4));
    CK(hipMalloc(&s->d_rowinfo, rowinfo.size() * 4));
    CK(hipMalloc(&s->d_rowk, rowk.size() * 4));
    CK(hipMemcpy(s->d_kvals, kvals.data(), kvals.size() * 4, hipMemcpyHostToDevice));
    CK(hipMemcpy(s->d_rowinfo, rowinfo.data(), rowinfo.size() * 4, hipMemcpyHostToDevice));
    CK(hipMemcpy(s->d_rowk, rowk.data(), rowk.size() * 4, hipMemcpyHostToDevice));
    s->J.kvals = s->d_kvals;
    s->J.rowinfo = s->d_rowinfo;
    s->J.rowk = s->d_rowk;
    s->max_batch = max_batch;
    if (hipMalloc(&s->d_rec, (size_t)max_batch * s->J.W * 8) != hipSuccess) {
        free_all(s);
        return RS_ERR_NOMEM;
    }
    s->hit_cap = (uint32_t)hit_capacity;
    CK(hipMalloc(&s->d_hits, (size_t)hit_capacity * sizeof(HitDev)));
    CK(hipMalloc(&s->d_hit_count, 4));
    CK(hipMemset(s->d_hit_count, 0, 4));
    for (int b = 0; b < 2; ++b) {
        CK(hipEventCreateWithFlags(&s->copied[b], hipEventDisableTiming));
        CK(hipEventCreateWithFlags(&s->consumed[b], hipEventDisableTiming));
    }
    int cus = 0, per_cu = 0;
    CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    CK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(rs_bsgs_kernel), 64 * s->waves_per_block,
        s->lds_bytes));
    s->bsgs_grid = std::max(1, cus * std::max(1, per_cu));
#undef CK
    *out = s;
    return RS_OK;
}

extern "C" rs_status rs_sieve_info(const rs_sieve* s, int* m, int* T, int* table_slots,
                                   int* lds_bytes_per_wave) {
    if (!s) return RS_ERR_ARG;
    if (m) *m = s->J.m;
    if (T) *T = s->J.T;
    if (table_slots) *table_slots = 8 << s->J.logNB;
    if (lds_bytes_per_wave) *lds_bytes_per_wave = s->J.lds_words * 4;
    return RS_OK;
}

// Launches set-up + BSGS over count primes whose offsets are in device
// memory, on the compute stream, recording timing events.
static rs_status launch(rs_sieve* s, uint64_t low, const uint32_t* d_offs, int64_t count) {
    Timing t{};
    HIPCHK(hipEventCreate(&t.a));
    HIPCHK(hipEventCreate(&t.b));
    HIPCHK(hipEventCreate(&t.c));
    t.primes = count;
    HIPCHK(hipEventRecord(t.a, s->stream));
    const int pb = 256;
    const unsigned pg = (unsigned)((count + pb - 1) / pb);
    hipLaunchKernelGGL(rs_prep_kernel, dim3(pg), dim3(pb), 0, s->stream, s->J, low, d_offs,
                       (long long)count, s->d_rec);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(t.b, s->stream));
    // one wavefront per prime, at most the resident grid (the kernel strides)
    const int wpb = s->waves_per_block;
    const int grid = (int)std::min<long long>(s->bsgs_grid, (count + wpb - 1) / wpb);
    hipLaunchKernelGGL(rs_bsgs_kernel, dim3(std::max(1, grid)), dim3(64 * wpb),
                       s->lds_bytes, s->stream, s->J, s->d_rec, (long long)count, s->d_hits,
                       s->d_hit_count, s->hit_cap);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(t.c, s->stream));
    s->timings.push_back(t);
    return RS_OK;
}

static rs_status check_range(const rs_sieve* s, uint64_t low, uint64_t max_off) {
    const uint64_t hi = low + max_off;
    if (hi < low || hi >= (1ull << 63) || low < 3) return RS_ERR_ARG;
    (void)s;
    return RS_OK;
}

extern "C" rs_status rs_sieve_submit_device(rs_sieve* s, uint64_t low,
                                            const uint32_t* offsets_dev, int64_t count) {
    if (!s || count < 0 || (count > 0 && !offsets_dev)) return RS_ERR_ARG;
    if (count == 0) return RS_OK;
    // offsets are not read back on this path: bound p by the widest offset
    if (low < 3 || low >= (1ull << 63) - (1ull << 32)) return RS_ERR_ARG;
    HIPCHK(hipSetDevice(s->device));
    for (int64_t done = 0; done < count; done += s->max_batch) {
        const int64_t n = std::min<int64_t>(s->max_batch, count - done);
        rs_status st = launch(s, low, offsets_dev + done, n);
        if (st != RS_OK) return st;
    }
    return RS_OK;
}

extern "C" rs_status rs_sieve_submit(rs_sieve* s, uint64_t low, const uint32_t* offsets,
                                     int64_t count) {
    if (!s || count < 0 || (count > 0 && !offsets)) return RS_ERR_ARG;
    if (count == 0) return RS_OK;
    uint32_t mx = 0;
    for (int64_t i = 0; i < count; ++i) mx = std::max(mx, offsets[i]);
    rs_status st = check_range(s, low, mx);
    if (st != RS_OK) return st;
    HIPCHK(hipSetDevice(s->device));
    for (int64_t done = 0; done < count; done += s->max_batch) {
        const int64_t n = std::min<int64_t>(s->max_batch, count - done);
        const int b = s->buf;
        if (!s->h_pin[b]) {
            HIPCHK(hipHostMalloc(&s->h_pin[b], (size_t)s->max_batch * 4, hipHostMallocDefault));
            HIPCHK(hipMalloc(&s->d_offs[b], (size_t)s->max_batch * 4));
        }
        // the buffer's previous batch must be fully sieved before reuse
        if (s->consumed_valid[b]) HIPCHK(hipEventSynchronize(s->consumed[b]));
        memcpy(s->h_pin[b], offsets + done, (size_t)n * 4);
        HIPCHK(hipMemcpyAsync(s->d_offs[b], s->h_pin[b], (size_t)n * 4, hipMemcpyHostToDevice,
                              s->copy_stream));
        HIPCHK(hipEventRecord(s->copied[b], s->copy_stream));
        HIPCHK(hipStreamWaitEvent(s->stream, s->copied[b], 0));
        st = launch(s, low, s->d_offs[b], n);
        if (st != RS_OK) return st;
        HIPCHK(hipEventRecord(s->consumed[b], s->stream));
        s->consumed_valid[b] = true;
        s->buf ^= 1;
    }
    return RS_OK;
}

static bool factor_less(const rs_factor& a, const rs_factor& b) {
    if (a.p != b.p) return a.p < b.p;
    if (a.k != b.k) return a.k < b.k;
    return a.n < b.n;
}

static rs_status drain(rs_sieve* s) {
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    uint32_t n = 0;
    HIPCHK(hipMemcpy(&n, s->d_hit_count, 4, hipMemcpyDeviceToHost));
    const uint32_t got = std::min(n, s->hit_cap);
    if (n > s->hit_cap) s->lost += (int64_t)(n - s->hit_cap);
    if (got) {
        const size_t old = s->held.size();
        s->held.resize(old + got);
        HIPCHK(hipMemcpy(s->held.data() + old, s->d_hits, (size_t)got * sizeof(HitDev),
                         hipMemcpyDeviceToHost));
    }
    hipLaunchKernelGGL(rs_zero_u32, dim3(1), dim3(64), 0, s->stream, s->d_hit_count, 1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    for (auto& t : s->timings) {
        float x = 0, y = 0;
        HIPCHK(hipEventElapsedTime(&x, t.a, t.b));
        HIPCHK(hipEventElapsedTime(&y, t.b, t.c));
        s->acc_prep_ms += x;
        s->acc_bsgs_ms += y;
        s->acc_primes += t.primes;
        hipEventDestroy(t.a);
        hipEventDestroy(t.b);
        hipEventDestroy(t.c);
    }
    s->timings.clear();
    return RS_OK;
}

extern "C" rs_status rs_sieve_sync(rs_sieve* s) {
    if (!s) return RS_ERR_ARG;
    return drain(s);
}

extern "C" rs_status rs_sieve_collect(rs_sieve* s, rs_factor* out, int64_t cap, int64_t* count,
                                      int64_t* lost) {
    if (!s || !count || cap < 0 || (cap > 0 && !out)) return RS_ERR_ARG;
    rs_status st = drain(s);
    if (st != RS_OK) return st;
    std::sort(s->held.begin(), s->held.end(), factor_less);
    *count = (int64_t)s->held.size();
    if (lost) *lost = s->lost;
    if ((int64_t)s->held.size() > cap) return RS_ERR_OVERFLOW;
    std::copy(s->held.begin(), s->held.end(), out);
    s->held.clear();
    s->lost = 0;
    return RS_OK;
}

extern "C" rs_status rs_sieve_timing(rs_sieve* s, double* bsgs_ms, double* prep_ms,
                                     int64_t* primes) {
    if (!s) return RS_ERR_ARG;
    rs_status st = drain(s);
    if (st != RS_OK) return st;
    if (bsgs_ms) *bsgs_ms = s->acc_bsgs_ms;
    if (prep_ms) *prep_ms = s->acc_prep_ms;
    if (primes) *primes = s->acc_primes;
    s->acc_bsgs_ms = s->acc_prep_ms = 0;
    s->acc_primes = 0;
    return RS_OK;
}

extern "C" void rs_sieve_destroy(rs_sieve* s) { free_all(s); }

// ------------------------------------------------------------ test helpers

template <typename F>
static rs_status run_elementwise(int device, int64_t count, F launch_fn) {
    if (count < 0) return RS_ERR_ARG;
    if (count == 0) return RS_OK;
    HIPCHK(hipSetDevice(device));
    return launch_fn();
}

extern "C" rs_status rs_dev_montmul(int device, const uint64_t* a, const uint64_t* b,
                                    const uint64_t* p, uint64_t* out, int64_t count) {
    if (!a || !b || !p || !out) return RS_ERR_ARG;
    for (int64_t i = 0; i < count; ++i)
        if (p[i] % 2 == 0 || p[i] < 3 || p[i] >= (1ull << 63) || b[i] >= p[i]) return RS_ERR_ARG;
    return run_elementwise(device, count, [&]() -> rs_status {
        uint64_t *da, *db, *dp, *dout;
        const size_t sz = (size_t)count * 8;
        HIPCHK(hipMalloc(&da, sz));
        HIPCHK(hipMalloc(&db, sz));
        HIPCHK(hipMalloc(&dp, sz));
        HIPCHK(hipMalloc(&dout, sz));
        HIPCHK(hipMemcpy(da, a, sz, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(db, b, sz, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dp, p, sz, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(rs_test_montmul, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                           0, da, db, dp, dout, (long long)count);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(out, dout, sz, hipMemcpyDeviceToHost));
        (void)hipFree(da);
        (void)hipFree(db);
        (void)hipFree(dp);
        (void)hipFree(dout);
        return RS_OK;
    });
}

extern "C" rs_status rs_dev_halfmul(int device, const uint32_t* k, const uint64_t* y,
                                    const uint64_t* p, uint64_t* out, int64_t count) {
    if (!k || !y || !p || !out) return RS_ERR_ARG;
    for (int64_t i = 0; i < count; ++i)
        if (p[i] % 2 == 0 || p[i] < 3 || p[i] >= (1ull << 63) || y[i] >= p[i]) return RS_ERR_ARG;
    return run_elementwise(device, count, [&]() -> rs_status {
        uint32_t* dk;
        uint64_t *dy, *dp, *dout;
        const size_t sz = (size_t)count * 8;
        HIPCHK(hipMalloc(&dk, (size_t)count * 4));
        HIPCHK(hipMalloc(&dy, sz));
        HIPCHK(hipMalloc(&dp, sz));
        HIPCHK(hipMalloc(&dout, sz));
        HIPCHK(hipMemcpy(dk, k, (size_t)count * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dy, y, sz, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dp, p, sz, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(rs_test_halfmul, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                           0, dk, dy, dp, dout, (long long)count);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(out, dout, sz, hipMemcpyDeviceToHost));
        (void)hipFree(dk);
        (void)hipFree(dy);
        (void)hipFree(dp);
        (void)hipFree(dout);
        return RS_OK;
    });
}

extern "C" rs_status rs_dev_jacobi(int device, const uint32_t* a, const uint64_t* p,
                                   int32_t* out, int64_t count) {
    if (!a || !p || !out) return RS_ERR_ARG;
    for (int64_t i = 0; i < count; ++i)
        if (p[i] % 2 == 0 || p[i] < 3) return RS_ERR_ARG;
    return run_elementwise(device, count, [&]() -> rs_status {
        uint32_t* da;
        uint64_t* dp;
        int32_t* dout;
        HIPCHK(hipMalloc(&da, (size_t)count * 4));
        HIPCHK(hipMalloc(&dp, (size_t)count * 8));
        HIPCHK(hipMalloc(&dout, (size_t)count * 4));
        HIPCHK(hipMemcpy(da, a, (size_t)count * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dp, p, (size_t)count * 8, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(rs_test_jacobi, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, 0,
                           da, dp, dout, (long long)count);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(out, dout, (size_t)count * 4, hipMemcpyDeviceToHost));
        (void)hipFree(da);
        (void)hipFree(dp);
        (void)hipFree(dout);
        return RS_OK;
    });
}

// ------------------------------------------------------------ device buffers

extern "C" rs_status rs_dev_alloc(int device, int64_t bytes, void** ptr) {
    if (!ptr || bytes <= 0) return RS_ERR_ARG;
    *ptr = nullptr;
    HIPCHK(hipSetDevice(device));
    if (hipMalloc(ptr, (size_t)bytes) != hipSuccess) return RS_ERR_NOMEM;
    return RS_OK;
}

extern "C" rs_status rs_dev_upload(int device, void* dst, const void* src, int64_t bytes) {
    if (!dst || !src || bytes < 0) return RS_ERR_ARG;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
    return RS_OK;
}

extern "C" rs_status rs_dev_free(int device, void* ptr) {
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipFree(ptr));
    return RS_OK;
}

extern "C" rs_status rs_dev_mad_rate(int device, int blocks, int iters, double* ops_per_s) {
    if (!ops_per_s || blocks < 1 || iters < 1) return RS_ERR_ARG;
    HIPCHK(hipSetDevice(device));
    uint64_t* out = nullptr;
    HIPCHK(hipMalloc(&out, (size_t)blocks * 256 * 8));
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    hipLaunchKernelGGL(rs_mad_probe, dim3(blocks), dim3(256), 0, 0, out, iters);  // warm-up
    HIPCHK(hipEventRecord(a, 0));
    hipLaunchKernelGGL(rs_mad_probe, dim3(blocks), dim3(256), 0, 0, out, iters);
    HIPCHK(hipEventRecord(b, 0));
    HIPCHK(hipEventSynchronize(b));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(out);
    *ops_per_s = (double)blocks * 256.0 * 8.0 * iters / (ms * 1e-3);
    return RS_OK;
}

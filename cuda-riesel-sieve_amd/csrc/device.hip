// device.hip -- C ABI of the device sieve: single jobs (rs_sieve_*) and
// batched launches of many jobs (rs_batch_*), both around the persistent
// kernel rs_sieve_kernel.
//
// Replaces the reference's device set-up and launch sequence
// (kernel_v0.24.6.cu:1367-1512 allocations and constant copies,
// :1537-1611 per-batch copy, memsets, addKernel1 launch and result copy):
//  * job constants and the row table are copied once at create;
//  * a chunk of primes runs as one persistent launch whose wavefronts take
//    64-prime groups from a device work counter (no per-launch batch limit,
//    no hash-table memsets: the tables live in LDS);
//  * host prime batches go through two pinned buffers on a copy stream, so
//    the host may refill one while the other is copied or sieved (the
//    reference reuses one pinned buffer, racing :1573 against :1592);
//  * hits go to a device append buffer (every hit, not one per prime) and
//    are returned sorted at collect;
//  * batched launches take job and chunk tables from device memory and are
//    captured once in a hipGraph, then replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../host/host.h"
#include "devcore.h"
#include "gpuprimes.h"
#include "kernels.h"

extern "C" __global__ void rs_accum_count(const uint32_t*, unsigned long long*);

#define HIPCHK(x)                                 \
    do {                                          \
        if ((x) != hipSuccess) return RS_ERR_HIP; \
    } while (0)

using namespace rsdev;


extern "C" int rs_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// ------------------------------------------------------------ single job

struct rs_sieve {
    int device = 0;
    DevJob job;
    Launcher L;
    int64_t max_batch = 0;
    hipStream_t stream = nullptr, copy_stream = nullptr;
    uint32_t* h_pin[2] = {nullptr, nullptr};
    uint32_t* d_offs[2] = {nullptr, nullptr};
    hipEvent_t copied[2] = {nullptr, nullptr};
    hipEvent_t consumed[2] = {nullptr, nullptr};
    bool consumed_valid[2] = {false, false};
    int buf = 0;
    std::vector<Timing> timings;
    std::vector<Timing> spare;  // timing events read and free for reuse (creating
                                // and destroying a pair per launch costs host time)
    double acc_ms = 0;
    long long acc_primes = 0;
    std::vector<HitDev> held;
    int64_t lost = 0;
    // device-generated chunks (rs_sieve_submit_range)
    rs_gpu_primegen* gpg = nullptr;
    uint32_t* d_range = nullptr;
    int64_t range_cap = 0;
    unsigned long long* d_range_primes = nullptr;  // primes of those chunks so far
};

static void free_sieve(rs_sieve* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->copy_stream) (void)hipStreamSynchronize(s->copy_stream);
    free_job(s->job);
    free_launcher(s->L);
    if (s->gpg) rs_gpu_primegen_free(s->gpg);
    (void)hipFree(s->d_range);
    (void)hipFree(s->d_range_primes);
    for (int b = 0; b < 2; ++b) {
        if (s->h_pin[b]) (void)hipHostFree(s->h_pin[b]);
        (void)hipFree(s->d_offs[b]);
        if (s->copied[b]) (void)hipEventDestroy(s->copied[b]);
        if (s->consumed[b]) (void)hipEventDestroy(s->consumed[b]);
    }
    for (auto* v : {&s->timings, &s->spare})
        for (auto& t : *v) {
            (void)hipEventDestroy(t.a);
            (void)hipEventDestroy(t.b);
        }
    if (s->stream) (void)hipStreamDestroy(s->stream);
    if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
    delete s;
}

extern "C" rs_status rs_sieve_create(int device, const rs_plan* pl, int m, int tile,
                                     int64_t max_batch, int64_t hit_capacity, rs_sieve** out) {
    if (!pl || !out || max_batch < 1 || hit_capacity < 1 || hit_capacity > 0x7fffffff)
        return RS_ERR_ARG;
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return RS_ERR_HIP;
    rs_sieve* s = new rs_sieve();
    s->device = device;
    s->max_batch = max_batch;
    rs_status st = make_job(pl, m, tile, false, &s->job);
    // the row table goes to block-shared LDS when it fits one row slice
    const int nrows = s->job.J.nrows;
    const int shared = nrows <= ROW_SLICE ? (2 * nrows + 3) & ~3 : 0;
    if (st == RS_OK && s->job.J.nwv > 0)
        st = init_launcher_group(&s->L, device, s->job.J, hit_capacity);
    else if (st == RS_OK)
        st = init_launcher(&s->L, device,
                           shared && s->job.J.list_idx ? s->job.J.lds_words_idx
                                                       : s->job.J.lds_words,
                           s->job.J.tab_words, s->job.J.W, hit_capacity, shared,
                           shared && s->job.J.list_idx);
    if (st == RS_OK &&
        (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
         hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking) != hipSuccess))
        st = RS_ERR_HIP;
    for (int b = 0; b < 2 && st == RS_OK; ++b)
        if (hipEventCreateWithFlags(&s->copied[b], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s->consumed[b], hipEventDisableTiming) != hipSuccess)
            st = RS_ERR_HIP;
    if (st != RS_OK) {
        free_sieve(s);
        return st;
    }
    *out = s;
    return RS_OK;
}

extern "C" rs_status rs_sieve_info(const rs_sieve* s, int* m, int* T, int* table_slots,
                                   int* lds_bytes_per_wave, int* tile) {
    if (!s) return RS_ERR_ARG;
    if (tile)
        *tile = s->job.J.tw | (s->job.J.list_idx ? RS_TILE_ROW_INDEX : 0) |
                (s->job.J.nwv == 4 ? RS_TILE_GROUP4 : s->job.J.nwv == 8 ? RS_TILE_GROUP8 : 0) |
                (s->job.J.sparse ? RS_TILE_SPARSE : 0) | (s->job.J.cubic ? RS_TILE_CUBIC : 0);
    if (m) *m = s->job.J.m;
    if (T) *T = s->job.J.T;
    if (table_slots) *table_slots = 8 << s->job.J.logNB;
    if (lds_bytes_per_wave) *lds_bytes_per_wave = s->L.lds_words * 4;
    return RS_OK;
}

extern "C" const char* rs_sieve_kernel_name(const rs_sieve* s) {
    if (!s) return "";
    if (s->L.nwv == 8) return "rs_sieve_group8_kernel";
    if (s->L.nwv == 4) return "rs_sieve_group4_kernel";
    return s->L.row_index ? "rs_sieve_kernel5_rix" : s->L.w5 ? "rs_sieve_kernel5" : "rs_sieve_kernel";
}

// One persistent launch of chunk c0 (ngroups groups of 64 primes) between
// two timing events; primes: the count credited to the timing (0 when the
// device counts them).
static rs_status timed_launch(rs_sieve* s, const ChunkDev& c0, long long ngroups,
                              int64_t primes) {
    Timing t{};
    if (!s->spare.empty()) {
        t = s->spare.back();
        s->spare.pop_back();
    } else {
        if (hipEventCreate(&t.a) != hipSuccess) return RS_ERR_HIP;
        if (hipEventCreate(&t.b) != hipSuccess) {
            (void)hipEventDestroy(t.a);
            return RS_ERR_HIP;
        }
    }
    t.primes = primes;
    rs_status st = hipEventRecord(t.a, s->stream) == hipSuccess ? RS_OK : RS_ERR_HIP;
    if (st == RS_OK) st = launch(s->L, s->stream, s->job.d_self, c0, nullptr, 1, ngroups);
    if (st == RS_OK && hipEventRecord(t.b, s->stream) != hipSuccess) st = RS_ERR_HIP;
    if (st != RS_OK) {  // the events are kept only on success
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
        return st;
    }
    s->timings.push_back(t);
    return RS_OK;
}

// One persistent launch over count primes (offsets in device memory).
static rs_status sieve_launch(rs_sieve* s, uint64_t low, const uint32_t* d_offs, int64_t count) {
    ChunkDev c0{};
    c0.low = low;
    c0.offs = d_offs;
    c0.count = count;
    c0.group0 = 0;
    c0.job = 0;
    return timed_launch(s, c0, groups_of(s->L, count), count);
}

extern "C" rs_status rs_sieve_submit_device(rs_sieve* s, uint64_t low,
                                            const uint32_t* offsets_dev, int64_t count) {
    if (!s || count < 0 || (count > 0 && !offsets_dev)) return RS_ERR_ARG;
    if (count == 0) return RS_OK;
    // offsets are not read back on this path: bound p by the widest offset
    if (low < 3 || low >= (1ull << 63) - (1ull << 32)) return RS_ERR_ARG;
    HIPCHK(hipSetDevice(s->device));
    return sieve_launch(s, low, offsets_dev, count);
}

// ------------------------------------------------------------ no allocation
//
// rs_sieve_submit_device without allocating: the record scratch and the
// work counter live in the file-scope array `scratch` (one copy per device,
// held until the process ends), hits go to the caller's buffer.  Host code
// never passes `scratch` itself to a kernel (its name is a host-side stand-in
// on the host): the device address of each device's copy is taken once with
// hipGetSymbolAddress and kept in g_scratch_dev.
namespace {

const size_t kNoAllocWords = (size_t)16 << 20;  // 128 MiB of uint64
const int kNoAllocCounterWords = 32;             // work counter, padded
const int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) uint64_t scratch[kNoAllocWords];

static std::mutex g_scratch_mu;
static uint64_t* g_scratch_dev[kMaxDevices];

extern "C" __global__ void rs_zero_u32(uint32_t*, int);

static rs_status noalloc_scratch(int device, uint64_t** ptr) {
    if (device < 0 || device >= kMaxDevices) return RS_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[device] = (uint64_t*)p;
    }
    *ptr = g_scratch_dev[device];
    return RS_OK;
}

extern "C" rs_status rs_sieve_submit_device_noalloc(rs_sieve* s, uint64_t low,
                                                    const uint32_t* offsets_dev, int64_t count,
                                                    rs_job_factor* hits_dev, int64_t hit_cap,
                                                    uint32_t* nhits_dev) {
    if (!s || count < 0 || (count > 0 && !offsets_dev) || !hits_dev || !nhits_dev ||
        hit_cap < 1 || hit_cap > 0x7fffffff)
        return RS_ERR_ARG;
    if (low < 3 || low >= (1ull << 63) - (1ull << 32)) return RS_ERR_ARG;
    HIPCHK(hipSetDevice(s->device));
    uint64_t* base = nullptr;
    rs_status st = noalloc_scratch(s->device, &base);
    if (st != RS_OK) return st;
    // as many resident workgroups as the array's records hold
    Launcher L = s->L;
    const size_t per_block = (size_t)L.scratch_words * (L.nwv > 0 ? 1 : L.wpb);
    const size_t blocks = (kNoAllocWords - kNoAllocCounterWords) / per_block;
    if (blocks < 1) return RS_ERR_ARG;  // one workgroup's records exceed the array
    L.grid = (int)std::min<size_t>((size_t)L.grid, blocks);
    L.work = (unsigned long long*)base;
    L.scratch = base + kNoAllocCounterWords;
    L.hits = (HitDev*)hits_dev;
    L.hit_count = nhits_dev;
    L.hit_cap = (uint32_t)hit_cap;
    // the hit count starts at zero (a kernel, not a memset)
    hipLaunchKernelGGL(rs_zero_u32, dim3(1), dim3(64), 0, s->stream, nhits_dev, 1);
    HIPCHK(hipGetLastError());
    if (count == 0) return RS_OK;
    ChunkDev c0{};
    c0.low = low;
    c0.offs = offsets_dev;
    c0.count = count;
    return launch(L, s->stream, s->job.d_self, c0, nullptr, 1, groups_of(L, count));
}

extern "C" rs_status rs_sieve_submit(rs_sieve* s, uint64_t low, const uint32_t* offsets,
                                     int64_t count) {
    if (!s || count < 0 || (count > 0 && !offsets)) return RS_ERR_ARG;
    if (count == 0) {  // only the staging buffers, made ready before the first primes
        HIPCHK(hipSetDevice(s->device));
        for (int b = 0; b < 2; ++b)
            if (!s->h_pin[b]) {
                HIPCHK(hipHostMalloc(&s->h_pin[b], (size_t)s->max_batch * 4,
                                     hipHostMallocDefault));
                HIPCHK(hipMalloc(&s->d_offs[b], (size_t)s->max_batch * 4));
            }
        return RS_OK;
    }
    if (low < 3) return RS_ERR_ARG;
    // the largest offset is read only when low + 2^32 - 1 could reach 2^63
    if (low >= (1ull << 63) - 0xffffffffull) {
        uint32_t mx = 0;
        for (int64_t i = 0; i < count; ++i) mx = std::max(mx, offsets[i]);
        if (low + mx < low || low + mx >= (1ull << 63)) return RS_ERR_ARG;
    }
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
        rs_status st = sieve_launch(s, low, s->d_offs[b], n);
        if (st != RS_OK) return st;
        HIPCHK(hipEventRecord(s->consumed[b], s->stream));
        s->consumed_valid[b] = true;
        s->buf ^= 1;
    }
    return RS_OK;
}

extern "C" rs_status rs_sieve_submit_range(rs_sieve* s, uint64_t lo, uint64_t hi) {
    if (!s || hi < lo || lo < 3 || hi > (1ull << 63)) return RS_ERR_ARG;
    HIPCHK(hipSetDevice(s->device));
    for (uint64_t a = lo; a < hi;) {
        const uint64_t b = std::min<uint64_t>(hi, a + (1ull << 32));
        if (!s->gpg || rs_gpu_primegen_hi(s->gpg) < b) {
            // sieving primes for up to 4x this bound, so later chunks reuse them
            if (s->gpg) {
                HIPCHK(hipStreamSynchronize(s->stream));
                rs_gpu_primegen_free(s->gpg);
                s->gpg = nullptr;
            }
            const uint64_t cap_hi = b > (1ull << 61) ? (1ull << 63) : 4 * b;
            rs_status st = rs_gpu_primegen_create(s->device, cap_hi, &s->gpg);
            if (st != RS_OK) return st;
        }
        const int64_t need = gpu_primes_capacity(b - a);
        if (need > s->range_cap) {
            HIPCHK(hipStreamSynchronize(s->stream));
            (void)hipFree(s->d_range);
            s->d_range = nullptr;
            if (hipMalloc(&s->d_range, (size_t)need * 4) != hipSuccess) return RS_ERR_NOMEM;
            s->range_cap = need;
        }
        if (!s->d_range_primes) {
            HIPCHK(hipMalloc(&s->d_range_primes, 8));
            HIPCHK(hipMemset(s->d_range_primes, 0, 8));
        }
        // generation and sieve are ordered on the sieve stream: the next
        // chunk's generation starts after this chunk's sieve has read d_range
        const uint32_t* cnt_dev = nullptr;
        rs_status st = gpu_primes_enqueue(s->gpg, a, b, s->d_range, s->range_cap, s->stream,
                                          &cnt_dev);
        if (st != RS_OK) return st;
        ChunkDev c0{};
        c0.low = a;
        c0.offs = s->d_range;
        c0.count = s->range_cap;
        c0.count_dev = cnt_dev;
        // the primes are counted on the device, see d_range_primes
        st = timed_launch(s, c0, groups_of(s->L, s->range_cap), 0);
        if (st != RS_OK) return st;
        hipLaunchKernelGGL(rs_accum_count, dim3(1), dim3(64), 0, s->stream, cnt_dev,
                           s->d_range_primes);
        HIPCHK(hipGetLastError());
        a = b;
    }
    return RS_OK;
}

static rs_status sieve_drain(rs_sieve* s) {
    HIPCHK(hipSetDevice(s->device));
    rs_status st = drain_hits(s->L, s->stream, &s->held, &s->lost);
    if (st != RS_OK) return st;
    if (s->d_range_primes) {
        unsigned long long n = 0;
        HIPCHK(hipMemcpy(&n, s->d_range_primes, 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemset(s->d_range_primes, 0, 8));
        s->acc_primes += (long long)n;
    }
    for (auto& t : s->timings) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, t.a, t.b));
        s->acc_ms += ms;
        s->acc_primes += t.primes;
        s->spare.push_back(t);
    }
    s->timings.clear();
    return RS_OK;
}

extern "C" rs_status rs_sieve_sync(rs_sieve* s) {
    if (!s) return RS_ERR_ARG;
    return sieve_drain(s);
}

extern "C" rs_status rs_sieve_collect(rs_sieve* s, rs_factor* out, int64_t cap, int64_t* count,
                                      int64_t* lost) {
    if (!s || !count || cap < 0 || (cap > 0 && !out)) return RS_ERR_ARG;
    rs_status st = sieve_drain(s);
    if (st != RS_OK) return st;
    std::sort(s->held.begin(), s->held.end(), hit_less);
    *count = (int64_t)s->held.size();
    if (lost) *lost = s->lost;
    if ((int64_t)s->held.size() > cap) return RS_ERR_OVERFLOW;
    for (size_t i = 0; i < s->held.size(); ++i) {
        out[i].p = s->held[i].p;
        out[i].k = s->held[i].k;
        out[i].n = s->held[i].n;
    }
    s->held.clear();
    s->lost = 0;
    return RS_OK;
}

extern "C" rs_status rs_sieve_timing(rs_sieve* s, double* kernel_ms, int64_t* primes) {
    if (!s) return RS_ERR_ARG;
    rs_status st = sieve_drain(s);
    if (st != RS_OK) return st;
    if (kernel_ms) *kernel_ms = s->acc_ms;
    if (primes) *primes = s->acc_primes;
    s->acc_ms = 0;
    s->acc_primes = 0;
    return RS_OK;
}

extern "C" void rs_sieve_destroy(rs_sieve* s) { free_sieve(s); }

// ------------------------------------------------------------ row masks

extern "C" __global__ void rs_test_row_masks(const JobDev*, const uint64_t*, uint64_t*,
                                             uint64_t*, long long);

extern "C" rs_status rs_dev_row_masks(int device, const rs_plan* pl, const uint64_t* p,
                                      int64_t count, uint64_t* masks) {
    if (!pl || count < 0 || (count > 0 && (!p || !masks))) return RS_ERR_ARG;
    for (int64_t i = 0; i < count; ++i)
        if (p[i] % 2 == 0 || p[i] < 3 || p[i] >= (1ull << 63)) return RS_ERR_ARG;
    if (count == 0) return RS_OK;
    HIPCHK(hipSetDevice(device));
    // the masks do not depend on m: any table size within the shape limits
    int m = rs_plan_auto_m(pl);
    for (int mm = 64; m < 64 && mm <= (1 << 16); mm += 64)
        if (rsh::shape_ok(pl, mm, 2)) m = mm;
    if (m < 64) return RS_ERR_ARG;
    DevJob dj;
    rs_status st = make_job(pl, m, 0, false, &dj);
    // the cubic test wherever it applies (the test kernel brings its own
    // work area, whatever the table size)
    if (st == RS_OK && dj.J.Q % 3 == 0 && dj.J.nk <= CUBIC_MAX_K && !dj.J.cubic) {
        dj.J.cubic = 1;
        if (hipMemcpy(dj.d_self, &dj.J, sizeof(JobDev), hipMemcpyHostToDevice) != hipSuccess)
            st = RS_ERR_HIP;
    }
    const int nw = (dj.J.nrows + 63) / 64;
    uint64_t *dp = nullptr, *drec = nullptr, *dout = nullptr;
    if (st == RS_OK &&
        (hipMalloc(&dp, (size_t)count * 8) != hipSuccess ||
         hipMalloc(&drec, (size_t)count * dj.J.W * 8) != hipSuccess ||
         hipMalloc(&dout, (size_t)count * nw * 8) != hipSuccess))
        st = RS_ERR_NOMEM;
    if (st == RS_OK && hipMemcpy(dp, p, (size_t)count * 8, hipMemcpyHostToDevice) != hipSuccess)
        st = RS_ERR_HIP;
    if (st == RS_OK) {
        hipLaunchKernelGGL(rs_test_row_masks, dim3((unsigned)((count + 63) / 64)), dim3(64),
                           dj.J.cubic ? cubic_area_bytes(dj.J.nk) : 0, 0, dj.d_self, dp, drec,
                           dout, (long long)count);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpy(masks, dout, (size_t)count * nw * 8, hipMemcpyDeviceToHost) != hipSuccess)
            st = RS_ERR_HIP;
    }
    (void)hipFree(dp);
    (void)hipFree(drec);
    (void)hipFree(dout);
    free_job(dj);
    return st;
}

// ------------------------------------------------------------ autotuner

namespace {

struct Tuner {
    int device;
    const rs_plan* pl;
    uint64_t low;
    const uint32_t* offs;
    int64_t count;
    int max_lds;
    bool warm = false;
    int n = 0, cap;
    int *cand_m, *cand_tile;
    double* ns_out;
    double best = 1e300;
    int best_m = 0, best_tile = 0;
    struct Cand {
        int m, tile;
        double ns;  // mean over its timings
        int reps;
    };
    std::vector<Cand> all;

    // Kernel ns per prime of one launch of (m, tile) on the sample, or the
    // status: RS_ERR_ARG when the layout does not fit one wavefront's LDS
    // limit (nor will a larger m), kSkip when the shape is outside the
    // kernel's index limits (m too small for the x-window: a larger m may
    // fit).
    static const rs_status kSkip = 1;
    rs_status measure(int m, int tile, double* ns) {
        if (!rsh::shape_ok(pl, m, std::max(1, tile & 7))) return kSkip;
        rs_sieve* s = nullptr;
        rs_status st = rs_sieve_create(device, pl, m, tile, 1, 1 << 16, &s);
        if (st != RS_OK) return st;
        if (s->job.J.nwv == 0 && s->job.J.lds_words * 4 > max_lds) {
            rs_sieve_destroy(s);
            return RS_ERR_ARG;
        }
        double ms_time = 0;
        int64_t primes = 0;
        if (!warm) {  // first launch of the process pays code-object loading
            st = rs_sieve_submit_device(s, low, offs, count);
            if (st == RS_OK) st = rs_sieve_timing(s, nullptr, nullptr);
            warm = true;
        }
        if (st == RS_OK) st = rs_sieve_submit_device(s, low, offs, count);
        if (st == RS_OK) st = rs_sieve_timing(s, &ms_time, &primes);
        rs_sieve_destroy(s);
        if (st != RS_OK) return st;
        *ns = ms_time * 1e6 / (double)count;
        return RS_OK;
    }

    // Times (m, tile) once and records it as a candidate.
    rs_status time(int m, int tile, double* ns) {
        rs_status st = measure(m, tile, ns);
        if (st != RS_OK) return st;
        all.push_back({m, tile, *ns, 1});
        ++n;
        if (*ns < best) {
            best = *ns;
            best_m = m;
            best_tile = tile;
        }
        return RS_OK;
    }

    // Times the `top` fastest candidates `extra` more times each and ranks
    // them by their mean: one timing on the sample varies by a few percent
    // from launch to launch, as much as the leading candidates differ.
    rs_status refine(int top, int extra) {
        std::vector<size_t> idx(all.size());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
        std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return all[a].ns < all[b].ns; });
        for (size_t r = 0; r < idx.size() && (int)r < top; ++r) {
            Cand& c = all[idx[r]];
            for (int e = 0; e < extra; ++e) {
                double ns = 0;
                rs_status st = measure(c.m, c.tile, &ns);
                if (st != RS_OK) return st;
                c.ns = (c.ns * c.reps + ns) / (c.reps + 1);
                ++c.reps;
            }
        }
        best = 1e300;
        for (const Cand& c : all)
            if (c.ns < best) best = c.ns, best_m = c.m, best_tile = c.tile;
        return RS_OK;
    }

    // The candidates (mean ns per prime) into the caller's arrays.
    void report() const {
        for (int i = 0; i < (int)all.size() && i < cap; ++i) {
            if (cand_m) cand_m[i] = all[i].m;
            if (cand_tile) cand_tile[i] = all[i].tile;
            if (ns_out) ns_out[i] = all[i].ns;
        }
    }
};

int round64(int m) { return std::max(64, (m + 32) / 64 * 64); }

// Shared-table layouts are worth timing when the balance of baby and giant
// steps, m ~ sqrt(A X / 2) (A = rows/2 Jacobi survivors, X the x-window,
// a giant step costing about half a baby step), lies well above the best
// table one wavefront could hold.
bool groups_worth(const rs_plan* pl, int best_m) {
    const double A = 0.5 * (double)pl->rows_k.size(), X = pl->x1 - pl->x0 + 1;
    const double mopt = std::sqrt(A * X * 0.5);
    return mopt >= 2048 && mopt > 1.5 * best_m;
}

// Times shared-table layouts (4 or 8 wavefronts per prime): effort 0 the
// grid of m in {2048 .. 16384} x both group sizes x tiles 2 and 4 (up to the
// first m whose layouts do not fit LDS); effort 1 m upwards with 8
// wavefronts and tile 4 while the time improves, then 4 wavefronts and tile
// 2 at the best m.
rs_status tune_groups(Tuner& tu, int effort) {
    static const int ms[] = {2048, 3072, 4096, 6144, 8192, 12288, 16384};
    double ns = 0;
    auto soft = [](rs_status st) {
        return st == RS_OK || st == RS_ERR_ARG || st == Tuner::kSkip || st == RS_ERR_NOMEM;
    };
    if (effort == 0) {
        for (int m : ms) {
            bool any_fits = false;
            for (int g : {RS_TILE_GROUP8, RS_TILE_GROUP4})
                for (int tile : {4, 2}) {
                    rs_status st = tu.time(m, tile | g, &ns);
                    if (st == RS_OK || st == Tuner::kSkip) any_fits = true;
                    else if (!soft(st)) return st;
                }
            if (!any_fits) break;
        }
        return RS_OK;
    }
    double prev = 1e300;
    int bm = 0;
    for (int m : ms) {
        rs_status st = tu.time(m, 4 | RS_TILE_GROUP8, &ns);
        if (st == Tuner::kSkip) continue;
        if (st != RS_OK) {
            if (!soft(st)) return st;
            break;  // larger tables do not fit either
        }
        if (ns >= prev) break;
        prev = ns;
        bm = m;
    }
    if (bm) {
        for (int t : {4 | RS_TILE_GROUP4, 2 | RS_TILE_GROUP8}) {
            rs_status st = tu.time(bm, t, &ns);
            if (!soft(st)) return st;
        }
    }
    return RS_OK;
}

}  // namespace

extern "C" rs_status rs_autotune(int device, const rs_plan* pl, uint64_t low,
                                 const uint32_t* offsets_dev, int64_t count, int effort,
                                 int* best_m, int* best_tile, int* cand_m, int* cand_tile,
                                 double* ns_per_prime, int cap, int* ncand) {
    if (!pl || !offsets_dev || count < 1 || !best_m || !best_tile || effort < 0 || effort > 1)
        return RS_ERR_ARG;
    HIPCHK(hipSetDevice(device));
    Tuner tu{device, pl, low, offsets_dev, count, 0};
    HIPCHK(hipDeviceGetAttribute(&tu.max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    tu.cap = cap;
    tu.cand_m = cand_m;
    tu.cand_tile = cand_tile;
    tu.ns_out = ns_per_prime;
    const int m0 = rs_plan_auto_m(pl);
    *best_m = 0;
    *best_tile = 0;
    double ns = 0;
    if (effort == 0) {
        // full grid: m0 * num / 8 for num = 4..16 (rounded to 64) x tiles
        std::vector<int> ms;
        for (int num = 4; num <= 16; ++num) {
            const int m = round64(m0 * num / 8);
            if (std::find(ms.begin(), ms.end(), m) == ms.end()) ms.push_back(m);
        }
        const int rix = pl->rows_k.size() <= (size_t)ROW_SLICE ? RS_TILE_ROW_INDEX : 0;
        for (int m : ms) {
            // an m none of whose layouts fits LDS ends the grid (larger ones
            // do not fit either); shapes outside the index limits are skipped
            bool any_fits = false;
            const int tiles[6] = {1, 2, 4, 1 | rix, 2 | rix, 4 | rix};
            for (int i = 0; i < (rix ? 6 : 3); ++i) {
                rs_status st = tu.time(m, tiles[i], &ns);
                if (st == RS_OK || st == Tuner::kSkip) any_fits = true;
                else if (st != RS_ERR_ARG) return st;
            }
            if (!any_fits) break;
        }
    } else if (m0 >= 64) {
        // coordinate descent: the tile at m0 (or the largest m that fits),
        // then m in steps of m0/4 up and down while the time improves
        int m = m0;
        rs_status st = RS_ERR_ARG;
        while (m >= 64) {
            st = tu.time(m, 2, &ns);
            if (st != RS_ERR_ARG) break;
            m = m > 64 ? round64(m * 3 / 4) : 0;
        }
        if (st != RS_OK && st != RS_ERR_ARG && st != Tuner::kSkip) return st;
        if (st == RS_OK) {
            for (int tile : {1, 4})
                if ((st = tu.time(m, tile, &ns)) != RS_OK && st != RS_ERR_ARG && st != Tuner::kSkip)
                    return st;
            const int tile = tu.best_tile;
            const int step = std::max(64, round64(m / 4));
            for (int dir : {+1, -1}) {
                double prev = tu.best;
                for (int mm = m + dir * step; mm >= 64; mm += dir * step) {
                    st = tu.time(mm, tile, &ns);
                    if (st == RS_ERR_ARG || st == Tuner::kSkip) break;
                    if (st != RS_OK) return st;
                    if (ns >= prev) break;
                    prev = ns;
                }
                if (tu.best_m != m) break;  // improved going up: skip going down
            }
            // the row-index layout at the best pair
            if (pl->rows_k.size() <= (size_t)ROW_SLICE) {
                st = tu.time(tu.best_m, tu.best_tile | RS_TILE_ROW_INDEX, &ns);
                if (st != RS_OK && st != RS_ERR_ARG && st != Tuner::kSkip) return st;
            }
        }
    }
    if (groups_worth(pl, tu.best_m)) {
        rs_status st = tune_groups(tu, effort);
        if (st != RS_OK) return st;
    }
    // the sparse table at the best pair (fewer bucket overflows for more LDS)
    if (tu.best_m && !(tu.best_tile & RS_TILE_SPARSE)) {
        double ns = 0;
        rs_status st = tu.time(tu.best_m, tu.best_tile | RS_TILE_SPARSE, &ns);
        if (st != RS_OK && st != RS_ERR_ARG && st != Tuner::kSkip && st != RS_ERR_NOMEM)
            return st;
    }
    // the cubic-residue row test at the best per-wavefront pair, where it
    // applies (Q a multiple of 3, few enough distinct k)
    if (tu.best_m && pl->Q % 3 == 0 && !(tu.best_tile & (RS_TILE_GROUP4 | RS_TILE_GROUP8))) {
        size_t nk = 0;
        for (size_t i = 0; i < pl->rows_k.size(); ++i)
            nk += i == 0 || pl->rows_k[i] != pl->rows_k[i - 1];
        if (nk <= CUBIC_MAX_K) {
            double ns = 0;
            rs_status st = tu.time(tu.best_m, tu.best_tile | RS_TILE_CUBIC, &ns);
            if (st != RS_OK && st != RS_ERR_ARG && st != Tuner::kSkip && st != RS_ERR_NOMEM)
                return st;
        }
    }
    rs_status st = tu.refine(4, 2);
    if (st != RS_OK && st != RS_ERR_NOMEM) return st;
    tu.report();
    if (ncand) *ncand = tu.n;
    *best_m = tu.best_m;
    *best_tile = tu.best_tile;
    return *best_m ? RS_OK : RS_ERR_ARG;
}

// ------------------------------------------------------------ batched jobs

struct rs_batch {
    int device = 0;
    std::vector<DevJob> jobs;
    JobDev* d_jobs = nullptr;
    MultiLauncher M;
    std::vector<ClassRange> classes;
    std::vector<SingleLaunch> singles;  // large chunks, own launch
    hipStream_t stream = nullptr;
    std::vector<ChunkDev> chunks;      // as added; ordered by class at capture
    ChunkDev* d_chunks = nullptr;
    size_t d_chunks_cap = 0;
    bool dirty = true;                 // chunk table changed since capture
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    bool timed = false;
    double acc_ms = 0;
    long long acc_primes = 0, pending_primes = 0;
    std::vector<HitDev> held;
    int64_t lost = 0;
};

static void free_batch(rs_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto& j : b->jobs) free_job(j);
    (void)hipFree(b->d_jobs);
    (void)hipFree(b->d_chunks);
    free_multi(b->M);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->ev_a) (void)hipEventDestroy(b->ev_a);
    if (b->ev_b) (void)hipEventDestroy(b->ev_b);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

extern "C" rs_status rs_batch_create(int device, int njobs, const rs_plan* const* plans,
                                     const int* ms, const int* tiles, int64_t hit_capacity,
                                     rs_batch** out) {
    if (!out || njobs < 1 || !plans || hit_capacity < 1 || hit_capacity > 0x7fffffff)
        return RS_ERR_ARG;
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return RS_ERR_HIP;
    rs_batch* b = new rs_batch();
    b->device = device;
    b->jobs.resize(njobs);
    rs_status st = RS_OK;
    int W = 0;
    for (int i = 0; i < njobs && st == RS_OK; ++i) {
        st = make_job(plans[i], ms ? ms[i] : 0, tiles ? tiles[i] : 0, false, &b->jobs[i]);
        if (st == RS_OK) W = std::max(W, b->jobs[i].J.W);
    }
    if (st == RS_OK) st = init_multi(&b->M, device, W, hit_capacity);
    if (st == RS_OK) {
        std::vector<JobDev> host(njobs);
        for (int i = 0; i < njobs; ++i) host[i] = b->jobs[i].J;
        if (hipMalloc(&b->d_jobs, sizeof(JobDev) * njobs) != hipSuccess ||
            hipMemcpy(b->d_jobs, host.data(), sizeof(JobDev) * njobs, hipMemcpyHostToDevice) !=
                hipSuccess ||
            hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreate(&b->ev_a) != hipSuccess || hipEventCreate(&b->ev_b) != hipSuccess)
            st = RS_ERR_HIP;
    }
    if (st != RS_OK) {
        free_batch(b);
        return st;
    }
    *out = b;
    return RS_OK;
}

extern "C" rs_status rs_batch_job_info(const rs_batch* b, int job, int* m, int* T,
                                       int* lds_bytes, int* tile) {
    if (!b || job < 0 || job >= (int)b->jobs.size()) return RS_ERR_ARG;
    if (tile) *tile = b->jobs[job].J.tw | (b->jobs[job].J.list_idx ? RS_TILE_ROW_INDEX : 0);
    if (m) *m = b->jobs[job].J.m;
    if (T) *T = b->jobs[job].J.T;
    if (lds_bytes) *lds_bytes = b->jobs[job].J.lds_words * 4;
    return RS_OK;
}

extern "C" rs_status rs_batch_add_chunk(rs_batch* b, int job, uint64_t low,
                                        const uint32_t* offsets_dev, int64_t count) {
    if (!b || job < 0 || job >= (int)b->jobs.size() || count < 0 || (count > 0 && !offsets_dev))
        return RS_ERR_ARG;
    if (low < 3 || low >= (1ull << 63) - (1ull << 32)) return RS_ERR_ARG;
    if (count == 0) return RS_OK;
    ChunkDev c{};
    c.low = low;
    c.offs = offsets_dev;
    c.count = count;
    c.job = job;
    b->chunks.push_back(c);
    b->dirty = true;
    return RS_OK;
}

extern "C" rs_status rs_batch_clear(rs_batch* b) {
    if (!b) return RS_ERR_ARG;
    b->chunks.clear();
    b->dirty = true;
    return RS_OK;
}

// Gives large chunks single-job launches and orders the rest by occupancy
// class, uploads the chunk table and captures one zero + kernel pair per
// launch into a graph.
static rs_status batch_capture(rs_batch* b) {
    std::vector<JobDev> host(b->jobs.size());
    for (size_t i = 0; i < b->jobs.size(); ++i) host[i] = b->jobs[i].J;
    std::vector<ChunkDev> ordered = b->chunks;
    rs_status st = plan_launches(&b->M, host, &ordered, &b->singles, &b->classes);
    if (st != RS_OK) return st;
    if (ordered.size() > b->d_chunks_cap) {
        (void)hipFree(b->d_chunks);
        b->d_chunks = nullptr;
        b->d_chunks_cap = std::max<size_t>(64, ordered.size());
        if (hipMalloc(&b->d_chunks, sizeof(ChunkDev) * b->d_chunks_cap) != hipSuccess)
            return RS_ERR_NOMEM;
    }
    if (!ordered.empty())
        HIPCHK(hipMemcpy(b->d_chunks, ordered.data(), sizeof(ChunkDev) * ordered.size(),
                         hipMemcpyHostToDevice));
    if (b->exec) {
        HIPCHK(hipGraphExecDestroy(b->exec));
        b->exec = nullptr;
    }
    if (b->graph) {
        HIPCHK(hipGraphDestroy(b->graph));
        b->graph = nullptr;
    }
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    st = launch_singles(b->singles, b->stream, b->d_jobs);
    if (st == RS_OK) st = launch_classes(b->classes, b->stream, b->d_jobs, b->d_chunks);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &graph);
    if (st != RS_OK) return st;
    HIPCHK(e);
    b->graph = graph;
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->dirty = false;
    return RS_OK;
}

extern "C" rs_status rs_batch_run(rs_batch* b) {
    if (!b) return RS_ERR_ARG;
    if (b->chunks.empty()) return RS_ERR_STATE;
    HIPCHK(hipSetDevice(b->device));
    if (b->timed) return RS_ERR_STATE;  // collect the previous run first
    if (b->dirty) {
        rs_status st = batch_capture(b);
        if (st != RS_OK) return st;
    }
    HIPCHK(hipEventRecord(b->ev_a, b->stream));
    HIPCHK(hipGraphLaunch(b->exec, b->stream));
    HIPCHK(hipEventRecord(b->ev_b, b->stream));
    b->timed = true;
    b->pending_primes = 0;
    for (auto& c : b->chunks) b->pending_primes += c.count;
    return RS_OK;
}

static rs_status batch_drain(rs_batch* b) {
    HIPCHK(hipSetDevice(b->device));
    rs_status st = drain_hits(b->M.base, b->stream, &b->held, &b->lost);
    if (st != RS_OK) return st;
    if (b->timed) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, b->ev_a, b->ev_b));
        b->acc_ms += ms;
        b->acc_primes += b->pending_primes;
        b->timed = false;
    }
    return RS_OK;
}

extern "C" rs_status rs_batch_collect(rs_batch* b, rs_job_factor* out, int64_t cap,
                                      int64_t* count, int64_t* lost) {
    if (!b || !count || cap < 0 || (cap > 0 && !out)) return RS_ERR_ARG;
    rs_status st = batch_drain(b);
    if (st != RS_OK) return st;
    std::sort(b->held.begin(), b->held.end(), hit_less);
    *count = (int64_t)b->held.size();
    if (lost) *lost = b->lost;
    if ((int64_t)b->held.size() > cap) return RS_ERR_OVERFLOW;
    for (size_t i = 0; i < b->held.size(); ++i) {
        out[i].p = b->held[i].p;
        out[i].k = b->held[i].k;
        out[i].n = b->held[i].n;
        out[i].job = b->held[i].job;
        out[i].pad = 0;
    }
    b->held.clear();
    b->lost = 0;
    return RS_OK;
}

extern "C" rs_status rs_batch_timing(rs_batch* b, double* kernel_ms, int64_t* primes) {
    if (!b) return RS_ERR_ARG;
    rs_status st = batch_drain(b);
    if (st != RS_OK) return st;
    if (kernel_ms) *kernel_ms = b->acc_ms;
    if (primes) *primes = b->acc_primes;
    b->acc_ms = 0;
    b->acc_primes = 0;
    return RS_OK;
}

extern "C" void rs_batch_destroy(rs_batch* b) { free_batch(b); }

// server.hip -- continuous-batching sieve server on one GPU (rs_server_*).
// No reference counterpart: the reference runs one job per process.
//
// Jobs (a plan, a table size and a p-range) arrive at any time and wait in
// a FIFO queue.  Each step:
//  1. admits queued jobs while a slot is free and the job's device
//     footprint (row tables, n-list bitmaps, its prime buffer) fits the HBM
//     budget; a job larger than the whole budget is rejected;
//  2. generates on the GPU the primes of every distinct p-window of the
//     active jobs (jobs at the same progress share one window and one
//     generation);
//  3. sieves every active job's window in one persistent launch (job and
//     chunk tables in device memory);
//  4. moves each job's hits to its result list and retires jobs that
//     reached the end of their range, freeing slot and budget.
// Job constants, row tables and bitmaps stay resident on the device from
// admission to retirement; only the p-window advances between steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <deque>
#include <map>
#include <vector>

#include "../host/host.h"
#include "devcore.h"
#include "gpuprimes.h"
#include "kernels.h"

#define HIPCHK(x)                                 \
    do {                                          \
        if ((x) != hipSuccess) return RS_ERR_HIP; \
    } while (0)

using namespace rsdev;

namespace {

enum { ST_QUEUED = 0, ST_RUNNING = 1, ST_DONE = 2, ST_REJECTED = -1 };

struct SrvJob {
    int64_t id = 0;
    rs_plan plan;
    int m = 0, tile = 0;
    bool dat = false;
    uint64_t p_lo = 0, p_hi = 0, p_next = 0;
    double ns_hint = 0, ns = 0;  // kernel ns per prime: caller's hint, estimate used
    int state = ST_QUEUED;
    DevJob dj;
    size_t footprint = 0;
    std::vector<rs_factor> hits;
};

struct Window {
    uint64_t lo, hi;
    uint32_t* buf;
    const uint32_t* cnt_dev;
    uint32_t count;
};

}  // namespace

struct rs_server {
    int device = 0;
    int64_t budget = 0, used = 0;
    int max_active = 0;
    uint64_t width = 0;
    double quantum_ms = 0;
    int lds_cap_words = 0, w_cap = 0;
    MultiLauncher M;
    hipStream_t stream = nullptr;
    rs_gpu_primegen* gpg = nullptr;
    std::vector<uint32_t*> bufs;        // one prime buffer per slot
    int64_t buf_cap = 0;
    uint32_t* h_counts = nullptr;       // pinned, one per slot
    JobDev* d_jobs = nullptr;
    ChunkDev* d_chunks = nullptr;
    std::map<int64_t, SrvJob> jobs;
    std::deque<int64_t> queue;
    std::vector<int64_t> active;
    int64_t next_id = 1;
    double kernel_ms = 0;
    long long primes = 0;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
};

static void free_server(rs_server* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto& kv : s->jobs) free_job(kv.second.dj);
    for (auto* b : s->bufs) (void)hipFree(b);
    if (s->h_counts) (void)hipHostFree(s->h_counts);
    (void)hipFree(s->d_jobs);
    (void)hipFree(s->d_chunks);
    free_multi(s->M);
    if (s->gpg) rs_gpu_primegen_free(s->gpg);
    if (s->ev_a) (void)hipEventDestroy(s->ev_a);
    if (s->ev_b) (void)hipEventDestroy(s->ev_b);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

extern "C" rs_status rs_server_create(int device, int64_t hbm_budget, int max_active,
                                      uint64_t chunk_width, int lds_cap_bytes,
                                      int64_t hit_capacity, double quantum_ms,
                                      rs_server** out) {
    if (!out || hbm_budget < 1 || max_active < 1 || max_active > 1024 || chunk_width < 2 ||
        chunk_width > (1ull << 32) || lds_cap_bytes < 4096 || hit_capacity < 1 ||
        hit_capacity > 0x7fffffff || !(quantum_ms >= 0) || quantum_ms > 1e6)
        return RS_ERR_ARG;
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return RS_ERR_HIP;
    rs_server* s = new rs_server();
    s->device = device;
    s->budget = hbm_budget;
    s->max_active = max_active;
    s->width = chunk_width;
    s->quantum_ms = quantum_ms;
    s->lds_cap_words = lds_cap_bytes / 4;
    s->w_cap = 256;  // record words: 22 + Q + T + ceil(rows/64)
    s->buf_cap = gpu_primes_capacity(chunk_width);
    Launcher g;
    rs_status st = launch_geometry(device, s->lds_cap_words, &g);  // cap fits the device
    if (st == RS_OK) st = init_multi(&s->M, device, s->w_cap, hit_capacity);
    if (st == RS_OK &&
        (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
         hipEventCreate(&s->ev_a) != hipSuccess || hipEventCreate(&s->ev_b) != hipSuccess ||
         hipHostMalloc(&s->h_counts, 4 * max_active, hipHostMallocDefault) != hipSuccess ||
         hipMalloc(&s->d_jobs, sizeof(JobDev) * max_active) != hipSuccess ||
         hipMalloc(&s->d_chunks, sizeof(ChunkDev) * max_active) != hipSuccess))
        st = RS_ERR_HIP;
    if (st != RS_OK) {
        free_server(s);
        return st;
    }
    *out = s;
    return RS_OK;
}

// Device bytes a job holds while admitted: tables, bitmaps, its prime buffer.
static size_t footprint(const rs_server* s, const rs_plan& pl, bool dat) {
    return job_device_bytes(&pl, dat) + (size_t)s->buf_cap * 4;
}

extern "C" rs_status rs_server_job_bytes(const rs_server* s, const rs_plan* pl, int dat,
                                        int64_t* bytes) {
    if (!s || !pl || !bytes) return RS_ERR_ARG;
    *bytes = (int64_t)footprint(s, *pl, dat != 0);
    return RS_OK;
}

extern "C" rs_status rs_server_submit(rs_server* s, const rs_plan* pl, int m, int tile, int dat,
                                      uint64_t p_lo, uint64_t p_hi, double ns_per_prime,
                                      int64_t* job_id) {
    if (!s || !pl || !job_id || p_lo < 3 || p_hi <= p_lo || p_hi > (1ull << 63) - (1ull << 32) ||
        !(ns_per_prime >= 0))
        return RS_ERR_ARG;
    SrvJob j;
    j.id = s->next_id++;
    j.plan = *pl;
    j.m = m;
    j.tile = tile;
    j.dat = dat != 0;
    j.p_lo = j.p_next = p_lo;
    j.p_hi = p_hi;
    j.ns_hint = ns_per_prime;
    j.footprint = footprint(s, j.plan, j.dat);
    *job_id = j.id;
    s->queue.push_back(j.id);
    s->jobs.emplace(j.id, std::move(j));
    return RS_OK;
}

// Kernel ns per prime of a job on MI355X, fitted to rs_sieve_kernel timings
// of 64 jobs of mixed shape (base 3..1030, 1..200 k, n-widths 1e4..5e5;
// median error 8%): a fixed part, baby steps, the Q*T table of giant-step
// starts and A*T giant steps with A = rows/2 Jacobi survivors.
static double est_ns_per_prime(const rs_plan& pl, const JobDev& J) {
    const double A = 0.5 * (double)pl.rows_k.size();
    return 0.282 + 0.001445 * J.m + 0.000274 * (double)pl.Q * J.T + 0.002097 * A * J.T;
}

// The job's next window: chunk_width, or about quantum_ms of kernel time.
static uint64_t window_width(const rs_server* s, const SrvJob& j) {
    if (s->quantum_ms <= 0) return s->width;
    // primes per unit of p near p_next ~ 1/ln p
    const double lnp = std::log((double)j.p_next + 16.0);
    const double w = s->quantum_ms * 1e6 / j.ns * lnp;
    uint64_t width = w >= (double)s->width ? s->width : (uint64_t)w;
    width = std::max<uint64_t>(width & ~0xffffull, std::min<uint64_t>(1ull << 20, s->width));
    return width;
}

// Builds the job's device state with the largest m (from the requested or
// cost-model value down) whose LDS layout fits the server's per-wave cap.
static rs_status admit(rs_server* s, SrvJob& j) {
    int m = j.m > 0 ? j.m : rs_plan_auto_m(&j.plan);
    // shared-table jobs (one workgroup per prime) need m in multiples of 64
    // nwv and a workgroup layout that fits the device; others a wavefront
    // layout within the server's cap
    const int nwv = (j.tile & RS_TILE_GROUP8) ? 8 : (j.tile & RS_TILE_GROUP4) ? 4 : 0;
    const int mstep = 64 * std::max(1, nwv);
    while (true) {
        DevJob dj;
        rs_status st = make_job(&j.plan, m, j.tile, j.dat, &dj);
        bool fits = false;
        if (st == RS_OK && nwv > 0) {
            Launcher g;
            fits = launch_geometry_group(s->device, dj.J, &g) == RS_OK;
        } else if (st == RS_OK) {
            fits = dj.J.lds_words <= s->lds_cap_words && dj.J.W <= s->w_cap;
        }
        if (fits) {
            j.dj = dj;
            j.m = m;
            j.ns = j.ns_hint > 0 ? j.ns_hint : est_ns_per_prime(j.plan, dj.J);
            return RS_OK;
        }
        free_job(dj);
        if (st != RS_OK && st != RS_ERR_ARG) return st;
        if (m <= mstep) return RS_ERR_BUDGET;
        m -= mstep;
    }
}

extern "C" rs_status rs_server_step(rs_server* s, int* active, int* queued,
                                    int64_t* primes_step) {
    if (!s) return RS_ERR_ARG;
    HIPCHK(hipSetDevice(s->device));
    if (primes_step) *primes_step = 0;
    // 1. admission, FIFO
    bool changed = false;
    while ((int)s->active.size() < s->max_active && !s->queue.empty()) {
        SrvJob& j = s->jobs.at(s->queue.front());
        if ((int64_t)j.footprint > s->budget) {
            j.state = ST_REJECTED;
            s->queue.pop_front();
            continue;
        }
        if (s->used + (int64_t)j.footprint > s->budget) break;
        rs_status st = admit(s, j);
        if (st == RS_ERR_BUDGET) {
            j.state = ST_REJECTED;
            s->queue.pop_front();
            continue;
        }
        if (st != RS_OK) return st;
        j.state = ST_RUNNING;
        s->used += (int64_t)j.footprint;
        s->active.push_back(j.id);
        s->queue.pop_front();
        changed = true;
    }
    if (s->active.empty()) {
        if (active) *active = 0;
        if (queued) *queued = (int)s->queue.size();
        return RS_OK;
    }
    if (changed) {
        std::vector<JobDev> host;
        for (int64_t id : s->active) host.push_back(s->jobs.at(id).dj.J);
        HIPCHK(hipMemcpyAsync(s->d_jobs, host.data(), sizeof(JobDev) * host.size(),
                              hipMemcpyHostToDevice, s->stream));
    }
    // 2. one prime generation per distinct window
    uint64_t need_hi = 0;
    for (int64_t id : s->active) need_hi = std::max(need_hi, s->jobs.at(id).p_hi);
    if (!s->gpg || rs_gpu_primegen_hi(s->gpg) < need_hi) {
        HIPCHK(hipStreamSynchronize(s->stream));
        if (s->gpg) rs_gpu_primegen_free(s->gpg);
        s->gpg = nullptr;
        rs_status st = rs_gpu_primegen_create(
            s->device, need_hi > (1ull << 61) ? (1ull << 63) : 4 * need_hi, &s->gpg);
        if (st != RS_OK) return st;
    }
    std::vector<Window> wins;
    std::vector<int> win_of(s->active.size());
    for (size_t a = 0; a < s->active.size(); ++a) {
        const SrvJob& j = s->jobs.at(s->active[a]);
        const uint64_t lo = j.p_next, hi = std::min(j.p_next + window_width(s, j), j.p_hi);
        int w = -1;
        for (size_t k = 0; k < wins.size(); ++k)
            if (wins[k].lo == lo && wins[k].hi == hi) w = (int)k;
        if (w < 0) {
            w = (int)wins.size();
            while ((int)s->bufs.size() <= w) {
                uint32_t* b = nullptr;
                if (hipMalloc(&b, (size_t)s->buf_cap * 4) != hipSuccess) return RS_ERR_NOMEM;
                s->bufs.push_back(b);
            }
            wins.push_back({lo, hi, s->bufs[w], nullptr, 0});
        }
        win_of[a] = w;
    }
    for (size_t k = 0; k < wins.size(); ++k) {
        rs_status st = gpu_primes_enqueue(s->gpg, wins[k].lo, wins[k].hi, wins[k].buf, s->buf_cap,
                                          s->stream, &wins[k].cnt_dev);
        if (st != RS_OK) return st;
        // the generator reuses its count slot: copy each count out in order
        HIPCHK(hipMemcpyAsync(s->h_counts + k, wins[k].cnt_dev, 4, hipMemcpyDeviceToHost,
                              s->stream));
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    for (size_t k = 0; k < wins.size(); ++k) {
        wins[k].count = s->h_counts[k];
        if ((int64_t)wins[k].count > s->buf_cap) return RS_ERR_OVERFLOW;
    }
    // 3. every active job's window: large windows in single-job launches,
    //    the rest in one persistent launch per occupancy class
    std::vector<ChunkDev> chunks;
    std::vector<JobDev> jhost;
    long long step_primes = 0;
    for (size_t a = 0; a < s->active.size(); ++a) {
        jhost.push_back(s->jobs.at(s->active[a]).dj.J);
        const Window& w = wins[win_of[a]];
        if (w.count == 0) continue;
        ChunkDev c{};
        c.low = w.lo;
        c.offs = w.buf;
        c.count = w.count;
        c.job = (int)a;
        step_primes += c.count;
        chunks.push_back(c);
    }
    const bool launched = !chunks.empty();
    if (launched) {
        std::vector<ClassRange> cls;
        std::vector<SingleLaunch> singles;
        rs_status st = plan_launches(&s->M, jhost, &chunks, &singles, &cls);
        if (st != RS_OK) return st;
        if (!chunks.empty())
            HIPCHK(hipMemcpyAsync(s->d_chunks, chunks.data(), sizeof(ChunkDev) * chunks.size(),
                                  hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipEventRecord(s->ev_a, s->stream));
        st = launch_singles(singles, s->stream, s->d_jobs);
        if (st == RS_OK) st = launch_classes(cls, s->stream, s->d_jobs, s->d_chunks);
        if (st != RS_OK) return st;
        HIPCHK(hipEventRecord(s->ev_b, s->stream));
    }
    // 4. hits to their jobs, progress, retirement
    std::vector<HitDev> hits;
    int64_t lost = 0;
    rs_status st = drain_hits(s->M.base, s->stream, &hits, &lost);
    if (st != RS_OK) return st;
    if (lost) return RS_ERR_OVERFLOW;
    if (launched) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, s->ev_a, s->ev_b));
        s->kernel_ms += ms;
    }
    s->primes += step_primes;
    if (primes_step) *primes_step = step_primes;
    for (const HitDev& h : hits) {
        SrvJob& j = s->jobs.at(s->active[h.job]);
        j.hits.push_back(rs_factor{h.p, h.k, h.n});
    }
    std::vector<int64_t> still;
    for (size_t a = 0; a < s->active.size(); ++a) {
        SrvJob& j = s->jobs.at(s->active[a]);
        j.p_next = wins[win_of[a]].hi;
        if (j.p_next >= j.p_hi) {
            j.state = ST_DONE;
            free_job(j.dj);
            s->used -= (int64_t)j.footprint;
            std::sort(j.hits.begin(), j.hits.end(), [](const rs_factor& x, const rs_factor& y) {
                if (x.p != y.p) return x.p < y.p;
                if (x.k != y.k) return x.k < y.k;
                return x.n < y.n;
            });
        } else {
            still.push_back(j.id);
        }
    }
    if (still.size() != s->active.size()) {
        s->active = still;
        std::vector<JobDev> host;
        for (int64_t id : s->active) host.push_back(s->jobs.at(id).dj.J);
        if (!host.empty())
            HIPCHK(hipMemcpyAsync(s->d_jobs, host.data(), sizeof(JobDev) * host.size(),
                                  hipMemcpyHostToDevice, s->stream));
    }
    if (active) *active = (int)s->active.size();
    if (queued) *queued = (int)s->queue.size();
    return RS_OK;
}

extern "C" rs_status rs_server_poll(rs_server* s, int64_t job_id, int* state, uint64_t* p_done,
                                    int64_t* nhits) {
    if (!s) return RS_ERR_ARG;
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return RS_ERR_ARG;
    const SrvJob& j = it->second;
    if (state) *state = j.state;
    if (p_done) *p_done = j.p_next - j.p_lo;
    if (nhits) *nhits = (int64_t)j.hits.size();
    return RS_OK;
}

extern "C" rs_status rs_server_results(rs_server* s, int64_t job_id, rs_factor* out,
                                       int64_t cap, int64_t* count) {
    if (!s || !count || cap < 0 || (cap > 0 && !out)) return RS_ERR_ARG;
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return RS_ERR_ARG;
    const SrvJob& j = it->second;
    *count = (int64_t)j.hits.size();
    if (*count > cap) return RS_ERR_OVERFLOW;
    std::copy(j.hits.begin(), j.hits.end(), out);
    return RS_OK;
}

extern "C" rs_status rs_server_release(rs_server* s, int64_t job_id) {
    if (!s) return RS_ERR_ARG;
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return RS_ERR_ARG;
    if (it->second.state == ST_RUNNING || it->second.state == ST_QUEUED) return RS_ERR_STATE;
    s->jobs.erase(it);
    return RS_OK;
}

extern "C" rs_status rs_server_stats(rs_server* s, int64_t* used_bytes, int64_t* budget,
                                     double* kernel_ms, int64_t* primes) {
    if (!s) return RS_ERR_ARG;
    if (used_bytes) *used_bytes = s->used;
    if (budget) *budget = s->budget;
    if (kernel_ms) *kernel_ms = s->kernel_ms;
    if (primes) *primes = s->primes;
    return RS_OK;
}

extern "C" void rs_server_destroy(rs_server* s) { free_server(s); }

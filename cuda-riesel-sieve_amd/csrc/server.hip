// server.hip -- continuous-batching sieve server on one GPU (rs_server_*).
// No reference counterpart: the reference runs one job per process.
//
// Jobs (a plan, a table size and a p-range) arrive at any time and wait in
// a FIFO queue.  Each call of rs_server_step
//  1. admits queued jobs while a slot is free and the job's device
//     footprint (row tables, n-list bitmaps, its prime buffers) fits the
//     HBM budget; a job larger than the whole budget is rejected;
//  2. queues the next step behind the one in flight, without waiting:
//     generation on the GPU of the primes of every distinct p-window of the
//     active jobs (jobs at the same progress share one window and one
//     generation; the counts stay on the device, where the launches read
//     them), then the launches (large windows in single-job launches, the
//     rest one persistent launch per occupancy class), with job and chunk
//     tables uploaded from pinned staging;
//  3. drains the previous step: waits for it, moves each job's hits to its
//     result list and retires jobs that reached the end of their range,
//     freeing slot and budget.
// So the device always has the next step queued while the host drains,
// plans and admits.  Two step slots alternate (tables, counts, hit
// buffers); job constants, row tables and bitmaps stay resident on the
// device from admission to retirement.
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
    uint64_t p_lo = 0, p_hi = 0;
    uint64_t p_next = 0;  // start of the next window to launch
    uint64_t p_done = 0;  // end of the windows whose hits are in
    int inflight = 0;     // launched steps not yet drained
    double ns_hint = 0, ns = 0;  // kernel ns per prime: caller's hint, estimate used
    int state = ST_QUEUED;
    DevJob dj;
    size_t footprint = 0;
    std::vector<rs_factor> hits;
};

// One of the two alternating steps: its prime windows, job and chunk
// tables (pinned staging and device copies) and hit buffer.
struct Slot {
    bool busy = false;              // launched, not drained
    std::vector<int64_t> ids;       // job of each job index of the step
    std::vector<uint64_t> win_hi;   // each job's window end
    std::vector<int> win_of;        // each job's window
    int nwin = 0;
    bool launched = false;          // the step launched a sieve kernel
    uint32_t* d_counts = nullptr;   // primes per window (device, read by the launches)
    uint32_t* h_counts = nullptr;   // the same, copied back (pinned)
    JobDev* h_jobs = nullptr;       // pinned staging
    ChunkDev* h_chunks = nullptr;
    JobDev* d_jobs = nullptr;
    ChunkDev* d_chunks = nullptr;
    HitDev* d_hits = nullptr;
    uint32_t* d_hit_count = nullptr;
    uint32_t* h_hit_count = nullptr;  // pinned
    hipEvent_t ev_a = nullptr, ev_b = nullptr, ev_done = nullptr;
};

}  // namespace

struct rs_server {
    int device = 0;
    int64_t budget = 0, used = 0;
    int max_active = 0;
    uint64_t width = 0;
    double quantum_ms = 0;
    int lds_cap_words = 0, w_cap = 0;
    uint32_t hit_cap = 0;
    MultiLauncher M;
    hipStream_t stream = nullptr;   // generation and sieve, in order
    rs_gpu_primegen* gpg = nullptr;
    std::vector<uint32_t*> bufs;    // prime windows (one per distinct window of a step)
    int64_t buf_cap = 0;
    Slot slots[2];
    std::deque<int> inflight;       // busy slots, oldest first
    std::map<int64_t, SrvJob> jobs;
    std::deque<int64_t> queue;
    std::vector<int64_t> active;    // admitted, not retired
    int64_t next_id = 1;
    double kernel_ms = 0;
    long long primes = 0;
};

static void free_slot(Slot& sl) {
    if (sl.h_counts) (void)hipHostFree(sl.h_counts);
    if (sl.h_jobs) (void)hipHostFree(sl.h_jobs);
    if (sl.h_chunks) (void)hipHostFree(sl.h_chunks);
    if (sl.h_hit_count) (void)hipHostFree(sl.h_hit_count);
    (void)hipFree(sl.d_counts);
    (void)hipFree(sl.d_jobs);
    (void)hipFree(sl.d_chunks);
    (void)hipFree(sl.d_hits);
    (void)hipFree(sl.d_hit_count);
    for (hipEvent_t* e : {&sl.ev_a, &sl.ev_b, &sl.ev_done})
        if (*e) (void)hipEventDestroy(*e);
}

static void free_server(rs_server* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto& kv : s->jobs) free_job(kv.second.dj);
    for (Slot& sl : s->slots) free_slot(sl);
    for (auto* b : s->bufs) (void)hipFree(b);
    free_multi(s->M);
    if (s->gpg) rs_gpu_primegen_free(s->gpg);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

static rs_status init_slot(Slot& sl, int max_active, int64_t hit_capacity) {
    const size_t n = (size_t)max_active;
    if (hipHostMalloc(&sl.h_counts, 4 * n, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&sl.h_jobs, sizeof(JobDev) * n, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&sl.h_chunks, sizeof(ChunkDev) * n, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&sl.h_hit_count, 4, hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&sl.d_counts, 4 * n) != hipSuccess ||
        hipMalloc(&sl.d_jobs, sizeof(JobDev) * n) != hipSuccess ||
        hipMalloc(&sl.d_chunks, sizeof(ChunkDev) * n) != hipSuccess ||
        hipMalloc(&sl.d_hits, sizeof(HitDev) * (size_t)hit_capacity) != hipSuccess ||
        hipMalloc(&sl.d_hit_count, 4) != hipSuccess)
        return RS_ERR_NOMEM;
    if (hipEventCreate(&sl.ev_a) != hipSuccess || hipEventCreate(&sl.ev_b) != hipSuccess ||
        hipEventCreateWithFlags(&sl.ev_done, hipEventDisableTiming) != hipSuccess)
        return RS_ERR_HIP;
    return RS_OK;
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
    s->w_cap = 256;  // record words: 23 + Q + 2T + ceil(rows/64)
    s->hit_cap = (uint32_t)hit_capacity;
    s->buf_cap = gpu_primes_capacity(chunk_width);
    Launcher g;
    rs_status st = launch_geometry(device, s->lds_cap_words, &g);  // cap fits the device
    // the launches write to the slots' hit buffers, not the launcher's
    if (st == RS_OK) st = init_multi(&s->M, device, s->w_cap, 1);
    for (int i = 0; i < 2 && st == RS_OK; ++i) st = init_slot(s->slots[i], max_active, hit_capacity);
    if (st == RS_OK && hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        st = RS_ERR_HIP;
    if (st != RS_OK) {
        free_server(s);
        return st;
    }
    *out = s;
    return RS_OK;
}

// Device bytes a job holds while admitted: tables, bitmaps, one prime
// window buffer.
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
    j.p_lo = j.p_next = j.p_done = p_lo;
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
// cost-model value down) whose LDS layout fits: shared-table jobs (one
// workgroup per prime) a workgroup layout within the device, m in
// multiples of 64 nwv; others a wavefront layout within the server's cap.
static rs_status admit(rs_server* s, SrvJob& j) {
    int m = j.m > 0 ? j.m : rs_plan_auto_m(&j.plan);
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

// Waits for a launched step, moves its hits to their jobs and retires the
// jobs whose last window it was.
static rs_status drain(rs_server* s, Slot& sl, long long* primes) {
    HIPCHK(hipEventSynchronize(sl.ev_done));
    const uint32_t n = *sl.h_hit_count;
    if (n > s->hit_cap) return RS_ERR_OVERFLOW;  // hits were dropped: raise hit_capacity
    std::vector<HitDev> hits(n);
    if (n) HIPCHK(hipMemcpy(hits.data(), sl.d_hits, sizeof(HitDev) * n, hipMemcpyDeviceToHost));
    if (sl.launched) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, sl.ev_a, sl.ev_b));
        s->kernel_ms += ms;
    }
    // primes sieved, counted per job (jobs sharing a window each sieve it)
    long long step_primes = 0;
    for (int w = 0; w < sl.nwin; ++w)
        if ((int64_t)sl.h_counts[w] > s->buf_cap) return RS_ERR_OVERFLOW;
    for (size_t i = 0; i < sl.ids.size(); ++i) step_primes += sl.h_counts[sl.win_of[i]];
    s->primes += step_primes;
    if (primes) *primes += step_primes;
    for (const HitDev& h : hits) s->jobs.at(sl.ids[h.job]).hits.push_back(rs_factor{h.p, h.k, h.n});
    for (size_t i = 0; i < sl.ids.size(); ++i) {
        SrvJob& j = s->jobs.at(sl.ids[i]);
        j.p_done = std::max(j.p_done, sl.win_hi[i]);
        if (--j.inflight == 0 && j.p_done >= j.p_hi) {
            j.state = ST_DONE;
            free_job(j.dj);
            s->used -= (int64_t)j.footprint;
            std::sort(j.hits.begin(), j.hits.end(), [](const rs_factor& x, const rs_factor& y) {
                if (x.p != y.p) return x.p < y.p;
                if (x.k != y.k) return x.k < y.k;
                return x.n < y.n;
            });
            s->active.erase(std::find(s->active.begin(), s->active.end(), j.id));
        }
    }
    sl.busy = false;
    sl.launched = false;
    return RS_OK;
}

// Queues the next step of the jobs of `run` into slot sl, all behind the
// step in flight on the one stream: the primes of each distinct window are
// generated on the GPU (their counts stay on the device, where the launches
// read them: the host does not wait), then the sieve launches, whose groups
// are numbered up to the window capacity (groups past a window's count are
// empty and skipped).
static rs_status launch_step(rs_server* s, Slot& sl, const std::vector<int64_t>& run) {
    uint64_t need_hi = 0;
    for (int64_t id : run) need_hi = std::max(need_hi, s->jobs.at(id).p_hi);
    if (!s->gpg || rs_gpu_primegen_hi(s->gpg) < need_hi) {
        // a new generator: nothing may still use the old one
        HIPCHK(hipStreamSynchronize(s->stream));
        if (s->gpg) rs_gpu_primegen_free(s->gpg);
        s->gpg = nullptr;
        rs_status st = rs_gpu_primegen_create(
            s->device, need_hi > (1ull << 61) ? (1ull << 63) : 4 * need_hi, &s->gpg);
        if (st != RS_OK) return st;
    }
    // distinct windows (jobs at equal progress share one)
    struct Win {
        uint64_t lo, hi;
    };
    std::vector<Win> wins;
    sl.win_of.assign(run.size(), 0);
    for (size_t a = 0; a < run.size(); ++a) {
        const SrvJob& j = s->jobs.at(run[a]);
        const uint64_t lo = j.p_next, hi = std::min(j.p_next + window_width(s, j), j.p_hi);
        int w = -1;
        for (size_t k = 0; k < wins.size(); ++k)
            if (wins[k].lo == lo && wins[k].hi == hi) w = (int)k;
        if (w < 0) {
            w = (int)wins.size();
            wins.push_back({lo, hi});
        }
        sl.win_of[a] = w;
    }
    sl.nwin = (int)wins.size();
    while (s->bufs.size() < wins.size()) {
        uint32_t* b = nullptr;
        if (hipMalloc(&b, (size_t)s->buf_cap * 4) != hipSuccess) return RS_ERR_NOMEM;
        s->bufs.push_back(b);
    }
    // the window buffers are rewritten behind the step in flight, which
    // reads them: the stream orders the two
    for (size_t k = 0; k < wins.size(); ++k) {
        const uint32_t* cnt_dev = nullptr;
        rs_status st = gpu_primes_enqueue(s->gpg, wins[k].lo, wins[k].hi, s->bufs[k], s->buf_cap,
                                          s->stream, &cnt_dev);
        if (st != RS_OK) return st;
        // the generator reuses its count slot: each count to the step's own
        HIPCHK(hipMemcpyAsync(sl.d_counts + k, cnt_dev, 4, hipMemcpyDeviceToDevice, s->stream));
    }
    HIPCHK(hipMemcpyAsync(sl.h_counts, sl.d_counts, 4 * wins.size(), hipMemcpyDeviceToHost,
                          s->stream));
    std::vector<ChunkDev> chunks;
    std::vector<JobDev> jhost;
    sl.ids = run;
    sl.win_hi.assign(run.size(), 0);
    size_t gneed = 0;  // scratch of shared-table launches
    for (size_t a = 0; a < run.size(); ++a) {
        SrvJob& j = s->jobs.at(run[a]);
        jhost.push_back(j.dj.J);
        if (j.dj.J.nwv > 0) {
            Launcher g;
            rs_status st = launch_geometry_group(s->device, j.dj.J, &g);
            if (st != RS_OK) return st;
            gneed = std::max(gneed, (size_t)g.grid * g.scratch_words);
        }
        const int w = sl.win_of[a];
        j.p_next = wins[w].hi;
        sl.win_hi[a] = wins[w].hi;
        ++j.inflight;
        ChunkDev c{};
        c.low = wins[w].lo;
        c.offs = s->bufs[w];
        c.count = s->buf_cap;           // a bound for the group numbering
        c.count_dev = sl.d_counts + w;  // the count the kernel reads
        c.job = (int)a;
        chunks.push_back(c);
    }
    // the shared-table scratch is reallocated while no launch can use it
    if (gneed > s->M.gscratch_words) HIPCHK(hipStreamSynchronize(s->stream));
    hipLaunchKernelGGL(rs_zero_u32, dim3(1), dim3(64), 0, s->stream, sl.d_hit_count, 1);
    HIPCHK(hipGetLastError());
    std::vector<ClassRange> cls;
    std::vector<SingleLaunch> singles;
    rs_status st = plan_launches(&s->M, jhost, &chunks, &singles, &cls);
    if (st != RS_OK) return st;
    // tables from pinned staging (this slot's previous step is drained)
    std::copy(jhost.begin(), jhost.end(), sl.h_jobs);
    std::copy(chunks.begin(), chunks.end(), sl.h_chunks);
    HIPCHK(hipMemcpyAsync(sl.d_jobs, sl.h_jobs, sizeof(JobDev) * jhost.size(),
                          hipMemcpyHostToDevice, s->stream));
    if (!chunks.empty())
        HIPCHK(hipMemcpyAsync(sl.d_chunks, sl.h_chunks, sizeof(ChunkDev) * chunks.size(),
                              hipMemcpyHostToDevice, s->stream));
    for (auto& x : singles) {
        x.L.hits = sl.d_hits;
        x.L.hit_count = sl.d_hit_count;
        x.L.hit_cap = s->hit_cap;
    }
    for (auto& x : cls) {
        x.L.hits = sl.d_hits;
        x.L.hit_count = sl.d_hit_count;
        x.L.hit_cap = s->hit_cap;
    }
    HIPCHK(hipEventRecord(sl.ev_a, s->stream));
    st = launch_singles(singles, s->stream, sl.d_jobs);
    if (st == RS_OK) st = launch_classes(cls, s->stream, sl.d_jobs, sl.d_chunks);
    if (st != RS_OK) return st;
    HIPCHK(hipEventRecord(sl.ev_b, s->stream));
    HIPCHK(hipMemcpyAsync(sl.h_hit_count, sl.d_hit_count, 4, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipEventRecord(sl.ev_done, s->stream));
    sl.launched = true;
    sl.busy = true;
    return RS_OK;
}

extern "C" rs_status rs_server_step(rs_server* s, int* active, int* queued,
                                    int64_t* primes_step) {
    if (!s) return RS_ERR_ARG;
    HIPCHK(hipSetDevice(s->device));
    long long drained = 0;
    // 1. admission, FIFO.  A slot is held by a job with range left to
    // launch: a job whose last window is in flight frees its slot for the
    // next step (its budget only when its hits are drained)
    auto runnable = [&]() {
        int n = 0;
        for (int64_t id : s->active) n += s->jobs.at(id).p_next < s->jobs.at(id).p_hi;
        return n;
    };
    int slots_used = runnable();
    while (slots_used < s->max_active && !s->queue.empty()) {
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
        ++slots_used;
    }
    // 2. the next step, queued behind the one in flight
    std::vector<int64_t> run;
    for (int64_t id : s->active)
        if (s->jobs.at(id).p_next < s->jobs.at(id).p_hi) run.push_back(id);
    bool launched = false;
    if (!run.empty()) {
        int idle = !s->slots[0].busy ? 0 : !s->slots[1].busy ? 1 : -1;
        if (idle < 0) {  // both in flight: drain the older first
            const int old = s->inflight.front();
            s->inflight.pop_front();
            rs_status st = drain(s, s->slots[old], &drained);
            if (st != RS_OK) return st;
            idle = old;
        }
        rs_status st = launch_step(s, s->slots[idle], run);
        if (st != RS_OK) return st;
        s->inflight.push_back(idle);
        launched = true;
    }
    // 3. drain the previous step (or, with nothing launched, the last one)
    if (!s->inflight.empty() && (s->inflight.size() > 1 || !launched)) {
        const int old = s->inflight.front();
        s->inflight.pop_front();
        rs_status st = drain(s, s->slots[old], &drained);
        if (st != RS_OK) return st;
    }
    if (primes_step) *primes_step = drained;
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
    if (p_done) *p_done = j.p_done - j.p_lo;
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

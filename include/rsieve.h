/* rsieve.h -- C ABI of the MI355X (gfx950) Riesel BSGS sieve, librsieve.so.
 *
 * Sieves candidates k*b^n-1: for every prime p of a chunk it finds every
 * (k, n) with n in [minN, maxN], (k, n mod Q) a planned row and
 * k*b^n == 1 (mod p).  Plain integers and pointers only; every call returns
 * an rs_status (0 = RS_OK, negative = error) and never throws.  Handles are
 * opaque; one caller thread per handle.
 *
 * Reference interface each group replaces (rob147147/CUDA-Riesel-Sieve
 * v0.24.6, file "CUDA Riesel Sieve/kernel_v0.24.6.cu" unless noted):
 *   ABCD input ........ countKs :113-137 (FileInput.cpp:12-36, FileInput.h:1)
 *                       and the parser in main :938-1083
 *   planner ........... Q choice :1086-1146, tMin/tMax :1148-1158,
 *                       row matrix + bitmaps :1160-1209, row sort :1219-1255
 *   prime generator ... small primes :1271-1303, wheel masks :1306-1364,
 *                       generateGPUPrimes :1838-1987
 *   device sieve ...... device set-up :1367-1512 (cudaMalloc/cudaMemcpy/
 *                       cudaMemcpyToSymbol of base, tMin, tMax, Q, m, rows),
 *                       per-batch launch :1537-1598 (lowGPU, KernelP copy,
 *                       memsets, addKernel1<<<>>> at :1578),
 *                       result copy + scan :1599-1611
 *   verification ...... bitmap lookup + __int128 modpow :1612-1719,
 *                       writeFactorToFile :1830-1835
 */
#ifndef RSIEVE_H
#define RSIEVE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int rs_status;
#define RS_OK 0
#define RS_ERR_ARG (-1)      /* bad argument or unsupported value */
#define RS_ERR_HIP (-2)      /* a HIP runtime call failed */
#define RS_ERR_NOMEM (-3)    /* host or device allocation failed */
#define RS_ERR_IO (-4)       /* file cannot be opened or written */
#define RS_ERR_FORMAT (-5)   /* malformed ABCD input */
#define RS_ERR_OVERFLOW (-6) /* a caller buffer is too small (count says how big) */
#define RS_ERR_BUDGET (-7)   /* serving: job does not fit the memory budget */
#define RS_ERR_STATE (-8)    /* call out of order (e.g. collect with nothing submitted) */
#define RS_ERR_RANGE (-9)    /* prime window has too few primes (reference exit(1), :1975-1979) */

/* Text for a status code (static string). */
const char* rs_strerror(rs_status s);

/* One factor: p | k*b^n-1. */
typedef struct rs_factor {
    uint64_t p;
    uint32_t k;
    uint32_t n;
} rs_factor;

/* One hit of a batched launch: p | k*b^n-1 for job `job`. */
typedef struct rs_job_factor {
    uint64_t p;
    uint32_t k;
    uint32_t n;
    uint32_t job;
    uint32_t pad;
} rs_job_factor;

/* ------------------------------------------------------------ ABCD input */
/* Replaces countKs (:113-137) + the parser (:938-1083): a line whose first
 * space-separated token is "ABCD" starts a k ("ABCD k*b^$a-1 [n0]"), the
 * base comes from the first header, other lines are positive n deltas. */
typedef struct rs_abcd rs_abcd;

rs_status rs_abcd_read(const char* path, rs_abcd** out);
/* Builds a file image from explicit lists: ks[nk], kstart[nk+1] CSR offsets
 * into ns (strictly increasing per k). */
rs_status rs_abcd_from_lists(uint32_t base, int nk, const int32_t* ks,
                             const int32_t* kstart, const int32_t* ns,
                             rs_abcd** out);
rs_status rs_abcd_info(const rs_abcd* a, uint32_t* base, int* nk, int* nn,
                       int* minN, int* maxN);
/* The form k*b^n + c of the file: c = -1 (Riesel, the reference's only form,
 * :1009-1016) or +1 (Sierpinski, "ABCD k*b^$a+1 [n0]" headers; a file mixing
 * both is rejected by rs_abcd_read).  set_form changes it for images built
 * from lists. */
rs_status rs_abcd_form(const rs_abcd* a, int* c);
rs_status rs_abcd_set_form(rs_abcd* a, int c);
/* Copies the lists out (ks[nk], kstart[nk+1], ns[nn]). */
rs_status rs_abcd_lists(const rs_abcd* a, int32_t* ks, int32_t* kstart,
                        int32_t* ns);
/* Writes the ABCD text form (srsieve layout, one header per k). */
rs_status rs_abcd_write(const rs_abcd* a, const char* path);
/* Factor removal ("dat" mode; no reference counterpart: the reference only
 * appends factor lines, :1606-1611, :1832).  Removes every candidate (k, n)
 * named by factors[0..count) from the lists (pairs not in the lists are
 * ignored), drops k whose list becomes empty, recomputes minN/maxN and sets
 * *removed (may be NULL) to the number of candidates removed.  Removing
 * every candidate leaves an image with no k, which rs_abcd_write writes as
 * an empty file and rs_plan_create rejects. */
rs_status rs_abcd_remove(rs_abcd* a, const rs_factor* factors, int64_t count,
                         int64_t* removed);
void rs_abcd_free(rs_abcd* a);

/* Synthetic input (no reference counterpart: the reference only reads ABCD
 * made by srsieve, sr_745.abcd:1).  Draws nk distinct k uniformly from
 * [kmin, kmax] (odd_k: 0 any parity, 1 odd only, 2 even only, as Riesel
 * k*b^n-1 with odd b needs even k) with a seeded splitmix64 stream, keeps
 * n in [nmin, nmax] such that no odd prime q < presieve_bound with q ∤ b
 * divides k*b^n+c (c = -1 or +1), and redraws a k whose list comes out
 * empty. */
rs_status rs_abcd_synth(uint32_t base, int nk, uint32_t kmin, uint32_t kmax,
                        int odd_k, int c, int nmin, int nmax,
                        uint32_t presieve_bound, uint64_t seed, rs_abcd** out);

/* ------------------------------------------------------------ planner */
/* Replaces :1086-1255.  Q is the even value in [2, 24) (or Qin, even) with
 * the least rows*R work, R = maxN/Q - minN/Q + 1; rows are the distinct
 * (k, n mod Q) sorted by (k, r), each with an MSB-first bitmap over
 * x' = (n-r)/Q - floor(minN/Q).  Unlike :1093, Q never falls back to 1. */
typedef struct rs_plan rs_plan;

rs_status rs_plan_create(const rs_abcd* a, int Qin, rs_plan** out);
/* The plan's form c (-1 or +1), from its ABCD image. */
rs_status rs_plan_form(const rs_plan* pl, int* c);
rs_status rs_plan_info(const rs_plan* pl, uint32_t* base, int* Q, int* nrows,
                       int* words, int* minN, int* maxN);
/* rows_k[nrows], rows_r[nrows], bitmaps[nrows*words] (any may be NULL). */
rs_status rs_plan_rows(const rs_plan* pl, int32_t* rows_k, int32_t* rows_r,
                       uint32_t* bitmaps);
/* 1 if n is in k's list (the bitmap test of :1620-1647), else 0. */
int rs_plan_contains(const rs_plan* pl, uint32_t k, uint32_t n);
/* tMin/tMax of :1148-1158 for m = 2^s (reported for the -m table). */
rs_status rs_plan_tminmax(const rs_plan* pl, int m, int* tmin, int* tmax);
/* Table size m picked by the cost model for this plan (a multiple of 64
 * within the job-shape limits below), or 0 when no m is within them. */
int rs_plan_auto_m(const rs_plan* pl);
void rs_plan_free(rs_plan* pl);

/* ------------------------------------------------------------ primes */
/* True primes (the reference's generator is a partial sieve, :1274 and
 * :1890, whose survivors are composite past (largest small prime)^2).
 * small_prime_scale keeps the meaning of -s: at least 4096*s sieving primes
 * are used; more are added up to sqrt(hi), and survivors above the sieving
 * bound squared are proved by deterministic Miller-Rabin.  threads <= 0
 * uses the hardware count. */
typedef struct rs_primegen rs_primegen;

rs_status rs_primegen_create(int small_prime_scale, uint64_t hi, int threads,
                             rs_primegen** out);
/* Every prime in [lo, hi) as offsets p - lo (hi - lo <= 2^32).  If cap is
 * too small returns RS_ERR_OVERFLOW with *count = the number needed.
 * Threads may call it on one generator at once: one call at a time reuses
 * the generator's per-block buffers, the others allocate their own. */
rs_status rs_primes_range(rs_primegen* g, uint64_t lo, uint64_t hi,
                          uint32_t* offsets, int64_t cap, int64_t* count);
/* generateGPUPrimes (:1838-1987) replacement: the first count primes >= low
 * (low made odd as at :1515-1519) as offsets from low, plus the next low
 * (last prime + 2, :1561).  RS_ERR_RANGE if 2^32 offsets do not hold count. */
rs_status rs_primes_batch(rs_primegen* g, uint64_t low, uint32_t* offsets,
                          int64_t count, uint64_t* next_low);
void rs_primegen_free(rs_primegen* g);

/* Prime generation on the GPU (the step before the sieve): segmented
 * sieve in LDS with the odd primes up to sqrt(hi_max) when that is at most
 * 2^25 (no Miller-Rabin at all below 2^50), else up to 2^22 with
 * deterministic Miller-Rabin above that bound squared.  Offsets p - lo go to
 * DEVICE memory.  No reference counterpart (the reference generates on the
 * CPU, :1838-1987). */
typedef struct rs_gpu_primegen rs_gpu_primegen;

rs_status rs_gpu_primegen_create(int device, uint64_t hi_max,
                                 rs_gpu_primegen** out);
/* Every prime in [lo, hi) (hi - lo <= 2^32, 3 <= lo, hi <= hi_max) into
 * offsets_dev; RS_ERR_OVERFLOW with *count = the number found if it exceeds
 * cap (nothing past cap is written).  Synchronous. */
rs_status rs_gpu_primes_range(rs_gpu_primegen* g, uint64_t lo, uint64_t hi,
                              uint32_t* offsets_dev, int64_t cap,
                              int64_t* count);
void rs_gpu_primegen_free(rs_gpu_primegen* g);

/* ------------------------------------------------------------ verification */
/* 1 if p | k*b^n + c, c = -1 (Riesel; 128-bit square-and-multiply as at
 * :1656-1671) or +1 (Sierpinski). */
int rs_verify_factor(uint64_t p, uint32_t k, uint32_t b, uint32_t n, int c);
/* Formats "p | k*b^n-1\n" (writeFactorToFile, :1833; "+1" for c = +1) into
 * buf; returns the length or a negative status. */
int rs_format_factor(const rs_factor* f, uint32_t b, int c, char* buf, int cap);

/* ------------------------------------------------------------ device sieve */
/* One sieve job on one GPU: base, Q, table size m (multiple of 64, or 0 for
 * the cost model's pick), rows and n window from a plan.  tile (1..4, 0 for
 * 2) is the number of giant steps of one row a lane takes per item,
 * optionally | RS_TILE_ROW_INDEX: for a job of at most 512 rows, each
 * wavefront lists its active rows as row numbers into the workgroup's
 * shared row table instead of copies of the rows, a smaller LDS layout
 * (more resident wavefronts) for one more LDS read per giant-step round;
 * which is faster depends on the job (the autotuner times both).
 * Replaces the set-up at :1367-1512.  max_batch bounds the primes copied
 * from host memory per launch (the reference's arraySize, :851);
 * hit_capacity bounds the device append buffer (the reference keeps one
 * slot per prime, :615-621). */
typedef struct rs_sieve rs_sieve;
#define RS_TILE_ROW_INDEX 16
/* Shared-table layouts (tile flags, with tile 1, 2 or 4; not with
 * RS_TILE_ROW_INDEX): the 4 or 8 wavefronts of a workgroup work on one prime
 * together with one baby table in the workgroup's LDS, for tables larger
 * than one wavefront's share of LDS allows (jobs with many rows over wide n
 * windows, whose best m runs into the thousands).  m must be a multiple of
 * 64 * 4 or 64 * 8 and at most 16384; rs_sieve_info then reports the
 * workgroup's LDS bytes. */
#define RS_TILE_GROUP4 32
#define RS_TILE_GROUP8 64
/* Sparse table (tile flag, any layout): twice the buckets, so the average
 * bucket load halves and fewer primes see a bucket overflow (whose giant
 * steps then also follow overflow chains), for more LDS. */
#define RS_TILE_SPARSE 128
/* Cubic-residue row test (tile flag, per-wavefront layouts).  When Q is a
 * multiple of 3 the set-up also drops, for primes p = 1 (mod 3) below
 * 2^50, the rows (k, r) whose k b^r is not a cube mod p (k b^n = 1 needs
 * it, and n = r mod 3): on those primes two thirds of the rows that pass
 * the quadratic test, for one exponentiation per distinct k and prime.  It
 * pays only where the dropped rows shorten the giant steps by whole rounds
 * (measured within a few percent either way), so it is off unless asked
 * for; rs_autotune tries it at its best (m, tile).  The factor set is the
 * same either way.  Ignored where it does not apply (Q not a multiple of
 * 3, more than 1024 distinct k, shared-table layouts, or a table too small
 * for the set-up's work area: 2560 + 64 * (distinct k) bytes);
 * rs_sieve_info reports the flag only where it applies. */
#define RS_TILE_CUBIC 256
/* Job-shape limits (rs_sieve_create, rs_batch_create, rs_autotune and the
 * server return RS_ERR_ARG outside them; rs_plan_auto_m only picks shapes
 * inside them): with X = floor(maxN/Q) - floor(minN/Q) + 1 and T the giant
 * steps per row, ceil(X/m) rounded up to a multiple of the tile and at
 * least 2,
 *   m a multiple of 64 in [64, 65536],
 *   T <= 65535 and Q * ys < 65536, ys = T (T odd) or T + 1 (T even),
 *   min(rows, 512) * T + 4096 < 65536 (a row slice's giant steps are
 *     numbered in 16 bits),
 *   at most 65535 distinct k,
 * and the layout of one wavefront (table, Y table, row list, record; with
 * RS_TILE_GROUP4/8 that of one workgroup; see rs_sieve_info) must fit the
 * device's LDS.  A small m raises T: a wide n window with many rows needs a
 * larger m. */

rs_status rs_sieve_create(int device, const rs_plan* pl, int m, int tile,
                          int64_t max_batch, int64_t hit_capacity,
                          rs_sieve** out);
/* lds_bytes_per_wave: the LDS bytes of one wavefront's layout, or of one
 * workgroup's for the shared-table layouts (RS_TILE_GROUP4/8). */
rs_status rs_sieve_info(const rs_sieve* s, int* m, int* T, int* table_slots,
                        int* lds_bytes_per_wave, int* tile);
/* Name of the kernel build the job's launches run (static string; "" for
 * NULL): rs_sieve_kernel, rs_sieve_kernel5 (layouts small enough for five
 * wavefronts per SIMD), rs_sieve_kernel5_rix (RS_TILE_ROW_INDEX) or
 * rs_sieve_group4_kernel / rs_sieve_group8_kernel (shared tables). */
const char* rs_sieve_kernel_name(const rs_sieve* s);
/* Primes p = low + offsets[i] from HOST memory (p odd, 3 <= p < 2^63, p ∤ b).
 * Copied through pinned double buffers on a copy stream, then sieved on the
 * compute stream; returns once the copy of this call's data is issued.
 * count = 0 only allocates the pinned buffers (max_batch primes each) if
 * they are not there yet, so a caller can have them ready before its first
 * primes (offsets may then be NULL).  Replaces :1543-1578. */
rs_status rs_sieve_submit(rs_sieve* s, uint64_t low, const uint32_t* offsets,
                          int64_t count);
/* Same, with offsets already in device memory (HBM-resident input): one
 * persistent launch, asynchronous. */
rs_status rs_sieve_submit_device(rs_sieve* s, uint64_t low,
                                 const uint32_t* offsets_dev, int64_t count);
/* rs_sieve_submit_device that allocates nothing (no reference counterpart:
 * for callers that cannot allocate device memory, such as code under HIP
 * graph capture or tight serving loops).  The record scratch and the work
 * counter live in a static 128 MiB device array of the library (one per
 * device, held until the process ends; the wavefronts launched are limited
 * to what it holds, RS_ERR_ARG if not even one workgroup's records fit);
 * hits go to the caller's device buffer hits_dev[hit_cap] (job 0, in no
 * particular order), their number to the device word *nhits_dev (cleared
 * by the call; a value above hit_cap means hits were dropped).  Enqueued on
 * the sieve's stream: rs_sieve_sync waits for it.  Calls that overlap on one
 * device race on the array. */
rs_status rs_sieve_submit_device_noalloc(rs_sieve* s, uint64_t low,
                                         const uint32_t* offsets_dev, int64_t count,
                                         rs_job_factor* hits_dev, int64_t hit_cap,
                                         uint32_t* nhits_dev);
/* Generates every prime of [lo, hi) on the device and sieves them: no host
 * prime generation and no PCIe copy (3 <= lo, hi <= 2^63; ranges wider than
 * 2^32 are split).  Asynchronous.  Replaces generateGPUPrimes (:1838-1987)
 * + :1543-1578 in one call. */
rs_status rs_sieve_submit_range(rs_sieve* s, uint64_t lo, uint64_t hi);
/* Waits for everything submitted and returns all hits found since the last
 * collect, sorted by (p, k, n).  If cap is too small: RS_ERR_OVERFLOW with
 * *count = hits held (kept for the next call).  *lost (optional) > 0 means
 * the device append buffer overflowed and hits were dropped (raise
 * hit_capacity).  Replaces :1596-1611. */
rs_status rs_sieve_collect(rs_sieve* s, rs_factor* out, int64_t cap,
                           int64_t* count, int64_t* lost);
/* Waits for all submitted work. */
rs_status rs_sieve_sync(rs_sieve* s);
/* Device time of the sieve kernel since the last call, in ms (HIP events on
 * the compute stream), and the number of primes it processed. */
rs_status rs_sieve_timing(rs_sieve* s, double* kernel_ms, int64_t* primes);
void rs_sieve_destroy(rs_sieve* s);

/* Autotuner (no reference counterpart: the reference takes m from -m).
 * Times candidate (m, tile) pairs for a plan on `count` primes
 * p = low + offsets_dev[i] (device memory) and returns the fastest.
 * effort 0 times the full grid: multiples of 64 from 1/2 to 2 times the
 * cost model's m that fit one wavefront's LDS limit, each with tiles 1, 2
 * and 4, each with and without RS_TILE_ROW_INDEX (plans of at most 512
 * rows).  effort 1 is a coordinate descent (tile at the model's m, then m
 * in steps of m/4 while the time improves, then the row-index layout at
 * the best pair), typically 6-9 timings.  Both then time the shared-table
 * layouts where the cost model's balance of baby and giant steps lies far
 * above what one wavefront's table holds, and, at the best pair, the
 * sparse table (RS_TILE_SPARSE) and, where it applies, the cubic-residue
 * test (RS_TILE_CUBIC).  Finally the four fastest are timed twice more
 * and ranked by the mean of their three timings.  cand_m,
 * cand_tile, ns_per_prime (optional, capacity cap) receive the measured
 * candidates (each pair once, with its mean time), *ncand their number. */
rs_status rs_autotune(int device, const rs_plan* pl, uint64_t low,
                      const uint32_t* offsets_dev, int64_t count, int effort, int* best_m,
                      int* best_tile, int* cand_m, int* cand_tile,
                      double* ns_per_prime, int cap, int* ncand);

/* ------------------------------------------------------------ batched jobs */
/* Many sieve jobs (different bases, k counts, n windows and table sizes)
 * in one persistent launch on one GPU.  No reference counterpart: the
 * reference runs one job per process.  ms[i] = 0 takes the cost model's m
 * for plans[i] and tiles[i] = 0 tile 2 (tiles may be NULL); the launch
 * shares one LDS layout sized for the largest job. */
typedef struct rs_batch rs_batch;


rs_status rs_batch_create(int device, int njobs, const rs_plan* const* plans,
                          const int* ms, const int* tiles, int64_t hit_capacity,
                          rs_batch** out);
rs_status rs_batch_job_info(const rs_batch* b, int job, int* m, int* T,
                            int* lds_bytes, int* tile);
/* Queues primes p = low + offsets_dev[i] (device memory) for job `job`. */
rs_status rs_batch_add_chunk(rs_batch* b, int job, uint64_t low,
                             const uint32_t* offsets_dev, int64_t count);
/* Drops every queued chunk. */
rs_status rs_batch_clear(rs_batch* b);
/* Sieves every queued chunk of every job in one persistent launch, captured
 * in a hipGraph when the chunk list changed and replayed otherwise.
 * Asynchronous; call rs_batch_collect before the next run. */
rs_status rs_batch_run(rs_batch* b);
/* Waits for the run and returns its hits sorted by (job, p, k, n); same
 * overflow contract as rs_sieve_collect. */
rs_status rs_batch_collect(rs_batch* b, rs_job_factor* out, int64_t cap,
                           int64_t* count, int64_t* lost);
/* Kernel time (HIP events) and primes of the runs since the last call. */
rs_status rs_batch_timing(rs_batch* b, double* kernel_ms, int64_t* primes);
void rs_batch_destroy(rs_batch* b);

/* ------------------------------------------------------------ serving */
/* Continuous-batching server on one GPU (no reference counterpart).  Jobs
 * arrive at any time and queue FIFO; each rs_server_step admits queued jobs
 * while a slot is free (max_active counts jobs with range left to launch)
 * and their device footprint (row tables, n-list bitmaps with dat, one
 * prime buffer of chunk_width) fits hbm_budget, queues the next step behind
 * the one in flight (GPU generation of the primes of every distinct
 * p-window of the running jobs, jobs at equal progress sharing one, then
 * one persistent launch per occupancy class of their LDS layouts), then
 * waits for the step before it, moves its hits to their jobs and retires
 * finished jobs (freeing their budget).  So the device always holds the next
 * step while the host admits, plans and drains.  Job state stays resident on
 * the device between steps.  lds_cap_bytes bounds one wavefront's LDS (a
 * job's m is lowered to fit).  A job larger than the whole budget is
 * rejected (state -1).
 *
 * Windows: with quantum_ms = 0 every job advances chunk_width per step.
 * With quantum_ms > 0 each job's window is sized to about quantum_ms of
 * kernel time from its cost estimate (ns per prime: the submit hint, else
 * a cost model of m, Q*T and rows*T), at most chunk_width and at least
 * 2^20, so cheap and costly jobs get equal device time per step (processor
 * sharing) and short jobs finish early. */
typedef struct rs_server rs_server;

rs_status rs_server_create(int device, int64_t hbm_budget, int max_active,
                           uint64_t chunk_width, int lds_cap_bytes,
                           int64_t hit_capacity, double quantum_ms,
                           rs_server** out);
/* Queues a job: the plan is copied; m/tile as for rs_sieve_create; dat != 0
 * keeps only n in the job's lists; primes p in [p_lo, p_hi).  ns_per_prime
 * is the job's measured kernel cost (rs_autotune), 0 if unknown. */
/* Device bytes a job of this plan holds while admitted (row tables, Jacobi
 * constants and character tables, n-list bitmaps with dat, one prime
 * window buffer): what admission charges against hbm_budget. */
rs_status rs_server_job_bytes(const rs_server* s, const rs_plan* pl, int dat,
                              int64_t* bytes);
rs_status rs_server_submit(rs_server* s, const rs_plan* pl, int m, int tile,
                           int dat, uint64_t p_lo, uint64_t p_hi,
                           double ns_per_prime, int64_t* job_id);
/* One scheduling step: returns once the previous step is drained (the new
 * one stays queued on the device).  Reports the admitted, unfinished jobs
 * and the queued ones after the call, and the primes sieved (per job) in the
 * step it drained. */
rs_status rs_server_step(rs_server* s, int* active, int* queued,
                         int64_t* primes_step);
/* state: 0 queued, 1 running, 2 done, -1 rejected; p_done: p-range sieved. */
rs_status rs_server_poll(rs_server* s, int64_t job_id, int* state,
                         uint64_t* p_done, int64_t* nhits);
/* The job's hits so far (sorted once the job is done). */
rs_status rs_server_results(rs_server* s, int64_t job_id, rs_factor* out,
                            int64_t cap, int64_t* count);
/* Forgets a finished or rejected job. */
rs_status rs_server_release(rs_server* s, int64_t job_id);
rs_status rs_server_stats(rs_server* s, int64_t* used_bytes, int64_t* budget,
                          double* kernel_ms, int64_t* primes);
void rs_server_destroy(rs_server* s);

/* Device-side helpers for tests (exact arithmetic on the GPU): out[i] =
 * montmul(a[i], b[i]) mod p[i] (R = 2^64) and out[i] = k[i]*y[i]*2^-32
 * mod p[i]; count elements, host buffers. */
rs_status rs_dev_montmul(int device, const uint64_t* a, const uint64_t* b,
                         const uint64_t* p, uint64_t* out, int64_t count);
rs_status rs_dev_halfmul(int device, const uint32_t* k, const uint64_t* y,
                         const uint64_t* p, uint64_t* out, int64_t count);
/* out[i] = Jacobi (a[i] / p[i]) in {-1, 0, 1} computed on the device. */
rs_status rs_dev_jacobi(int device, const uint32_t* a, const uint64_t* p,
                        int32_t* out, int64_t count);
/* The sieve kernel's own row filter (the Legendre test of :491-521, as the
 * set-up phase computes it: per-k Jacobi constants, quadratic-character
 * tables, (2/p)^z, reciprocity and, for k*b^n+1, (-1/p); and, when Q is a
 * multiple of 3 and the plan has at most 1024 distinct k, the cubic-residue
 * test for p = 1 (mod 3) below 2^50) for the plan's rows and odd primes
 * 3 <= p[i] < 2^63: bit j of masks[i * NW + w], NW = ceil(rows / 64), is
 * set iff row 64 w + j, (k, r), can hold a factor of p[i], i.e. c k b^r is
 * a quadratic residue mod p[i] (c = 1 for k*b^n-1, -1 for k*b^n+1) and,
 * where the cubic test applies, a cubic residue too.  Host buffers; for
 * tests. */
rs_status rs_dev_row_masks(int device, const rs_plan* pl, const uint64_t* p,
                           int64_t count, uint64_t* masks);

/* Device buffers for HBM-resident inputs (plain hipMalloc/hipMemcpy; no
 * reference counterpart beyond the cudaMalloc/cudaMemcpy of :1387-1512). */
rs_status rs_dev_alloc(int device, int64_t bytes, void** ptr);
rs_status rs_dev_upload(int device, void* dst, const void* src, int64_t bytes);
rs_status rs_dev_download(int device, void* dst, const void* src, int64_t bytes);
rs_status rs_dev_free(int device, void* ptr);

/* Measured rate of the 32x32+64 -> 64 multiply-add (v_mad_u64_u32) on the
 * device, in lane-operations per second, from a dependent-chain kernel of
 * `iters` mads per lane over `blocks` x 256 lanes. */
rs_status rs_dev_mad_rate(int device, int blocks, int iters, double* ops_per_s);

/* HBM bandwidth of a streaming copy of `bytes` (read + write counted), in
 * bytes per second; also the known-size reference for traffic counters. */
rs_status rs_dev_copy_rate(int device, int64_t bytes, double* bytes_per_s);

/* Number of visible HIP devices (0 on a machine without a GPU). */
int rs_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* RSIEVE_H */

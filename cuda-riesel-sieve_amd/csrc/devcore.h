// devcore.h -- internal interface shared by the device sieve front ends.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <vector>

#include "../host/host.h"
#include "kernels.h"

#define RS_SIEVE_KERNEL_ARGS                                                                  \
    ChunkDev c0, const JobDev *jobs, const ChunkDev *chunks, int nchunks, long long ngroups, \
        int lds_words, int tab_words, int shared_words, int scratch_words, uint64_t *scratch, \
        unsigned long long *work, HitDev *hits, uint32_t *hit_count, uint32_t hit_cap
// 4 and 5 wavefronts per SIMD with full row entries, and 5 with row-number
// lists (RS_TILE_ROW_INDEX, single-job launches)
extern "C" __global__ void rs_sieve_kernel(RS_SIEVE_KERNEL_ARGS);
extern "C" __global__ void rs_sieve_kernel5(RS_SIEVE_KERNEL_ARGS);
extern "C" __global__ void rs_sieve_kernel5_rix(RS_SIEVE_KERNEL_ARGS);
// one shared table per workgroup of 4 or 8 wavefronts (single-chunk launches)
extern "C" __global__ void rs_sieve_group4_kernel(RS_SIEVE_KERNEL_ARGS);
extern "C" __global__ void rs_sieve_group8_kernel(RS_SIEVE_KERNEL_ARGS);
extern "C" __global__ void rs_zero_u32(uint32_t*, int);
extern "C" __global__ void rs_jacobi_table(uint32_t n, uint32_t words, uint32_t* out);

namespace rsdev {

const int kMaxWavesPerBlock = 4;

// Job constants and device row tables built from a plan.
struct DevJob {
    JobDev J{};
    uint4* d_kvals = nullptr;    // Jacobi constants of each distinct k
    uint32_t* d_kjt = nullptr;   // quadratic-character table offsets of each k
    uint32_t* d_jtab = nullptr;  // the tables
    uint4* d_kfac = nullptr;     // per k: its prime-factor entries
    uint4* d_fac = nullptr;
    uint32_t* d_kw0 = nullptr;   // row-mask words of each k
    uint64_t* d_kmask = nullptr;
    uint32_t* d_kval = nullptr;   // each distinct k
    uint64_t* d_kmask3 = nullptr; // its rows by r mod 3
    uint2* d_rowent = nullptr;
    uint32_t* d_bitmap = nullptr;
    JobDev* d_self = nullptr;    // J itself in device memory
};

// Resources of the persistent kernel on one device: grid shape, per-wave
// record scratch, work counter and hit buffer.
struct Launcher {
    int device = 0;
    int grid = 0;
    int wpb = 1;
    int lds_words = 0;       // per wavefront
    int tab_words = 0;       // of which the table slot (a power of two, at least
                             // the table of every job launched)
    int shared_words = 0;    // block-shared words after the wavefront regions
    bool w5 = false;         // rs_sieve_kernel5 (row_index: rs_sieve_kernel5_rix)
    bool row_index = false;
    int nwv = 0;             // > 0: rs_sieve_group<nwv>_kernel, one table per block
    size_t lds_bytes = 0;
    int scratch_words = 0;  // per wavefront: 64 records (nwv > 0: per block,
                            // scratch_words_group)
    uint64_t* scratch = nullptr;
    unsigned long long* work = nullptr;
    HitDev* hits = nullptr;
    uint32_t* hit_count = nullptr;
    uint32_t hit_cap = 0;
};

// Launch resources shared by jobs of different LDS layouts: one persistent
// launch per occupancy class (jobs whose layouts give the same grid shape),
// each sized for its largest member, all sharing one scratch, work counter
// and hit buffer (the launches run in order on one stream).
struct MultiLauncher {
    Launcher base;                // owns the buffers
    int w_cap = 0;                // largest record (words) the scratch holds
    size_t max_waves = 0;         // resident waves of the smallest layout
    std::map<int, Launcher> geo;  // per-wave LDS words -> grid shape
    uint64_t* gscratch = nullptr; // scratch of shared-table launches (grown on demand)
    size_t gscratch_words = 0;
};

struct ClassRange {
    Launcher L;                   // class grid shape + shared buffers
    int chunk0 = 0, nchunks = 0;  // slice of the class-ordered chunk table
    long long ngroups = 0;
};

struct Timing {
    hipEvent_t a, b;
    long long primes;
};

uint32_t magic_div(int d);
uint4 jacobi_consts(uint32_t a);
void free_job(DevJob& j);
// Job constants from a plan, table size m (0: cost model), giant-step tile
// (1, 2 or 4; 0: 2); with dat, the plan's n-list bitmaps go to the device
// and hits outside the lists are dropped by the kernel.
rs_status make_job(const rs_plan* pl, int m, int tw, bool dat, DevJob* dj);
// Device bytes make_job allocates for a plan (row tables, bitmaps if dat).
size_t job_device_bytes(const rs_plan* pl, bool dat);
void free_launcher(Launcher& L);
// shared_words: block-shared LDS (the row table of a single-job launch).
// row_index: the layout lists rows by number (needs rs_sieve_kernel5).
rs_status init_launcher(Launcher* L, int device, int lds_words, int tab_words, int W,
                        int64_t hit_capacity, int shared_words = 0, bool row_index = false);
// Grid shape for a per-wave LDS layout (no allocation): of the two kernels
// the one keeping more wavefronts resident (rs_sieve_kernel on a tie), or
// rs_sieve_kernel5 when row_index.
rs_status launch_geometry(int device, int lds_words, Launcher* L, int shared_words = 0,
                          bool row_index = false);
rs_status init_multi(MultiLauncher* M, int device, int w_cap, int64_t hit_capacity);
// Grid shape of a shared-table job (J.nwv > 0): one workgroup of nwv
// wavefronts per prime, lds_words_group words of LDS per workgroup.
rs_status launch_geometry_group(int device, const JobDev& J, Launcher* L);
// The same, allocating its scratch, work counter and hit buffer.
rs_status init_launcher_group(Launcher* L, int device, const JobDev& J, int64_t hit_capacity);
// Launch groups of count primes: 64 per group, 64 nwv for shared tables.
inline long long groups_of(const Launcher& L, long long count) {
    const long long g = 64ll * (L.nwv > 0 ? L.nwv : 1);
    return (count + g - 1) / g;
}
// Orders `chunks` (each with .job indexing `jobs`) by occupancy class,
// numbering groups from 0 within each class, and returns one range per
// non-empty class.
rs_status plan_classes(MultiLauncher* M, const std::vector<JobDev>& jobs,
                       std::vector<ChunkDev>* chunks, std::vector<ClassRange>* out);
// A chunk with at least kSingleGroupsPerWave 64-prime groups per resident
// wave of its job's own grid runs as a single-job launch: its launch tail is
// negligible, and the job's row table can sit in block-shared LDS, which a
// launch mixing jobs cannot do (2-8% faster on 2^30-wide chunks).
const int kSingleGroupsPerWave = 32;
struct SingleLaunch {
    Launcher L;   // the job's grid shape + the shared buffers of M
    ChunkDev c0;  // the chunk (its .job indexes the launch's job table)
};
// Splits `chunks` into single-job launches (large chunks) and occupancy
// classes of the rest (`chunks` becomes the class-ordered rest).
rs_status plan_launches(MultiLauncher* M, const std::vector<JobDev>& jobs,
                        std::vector<ChunkDev>* chunks, std::vector<SingleLaunch>* singles,
                        std::vector<ClassRange>* cls);
rs_status launch_singles(const std::vector<SingleLaunch>& singles, hipStream_t stream,
                         const JobDev* jobs);
// Enqueues one persistent launch per class; d_chunks holds the ordered table.
rs_status launch_classes(const std::vector<ClassRange>& cls, hipStream_t stream,
                         const JobDev* jobs, const ChunkDev* d_chunks);
void free_multi(MultiLauncher& M);
// jobs: device job constants, indexed by the chunks' .job (c0 when chunks
// is null); a launcher with shared_words is for a single job.
rs_status launch(const Launcher& L, hipStream_t stream, const JobDev* jobs, const ChunkDev& c0,
                 const ChunkDev* chunks, int nchunks, long long ngroups);
rs_status drain_hits(const Launcher& L, hipStream_t stream, std::vector<HitDev>* out,
                     int64_t* lost);
bool hit_less(const HitDev& a, const HitDev& b);

}  // namespace rsdev

// gpuprimes.hip -- prime offsets generated on the GPU: the step before the
// sieve, replacing the CPU generator of the reference (generateGPUPrimes,
// kernel_v0.24.6.cu:1838-1987) when the chunk should never leave HBM.
//
// Every prime p in [lo, hi) (hi - lo <= 2^32, 3 <= lo, hi <= 2^63) is
// written as the uint32 offset p - lo, ascending.  The odd numbers of the
// range are split into segments (kSegSmall or kSegWide odd numbers); one
// workgroup sieves a segment
// in an LDS bitmap with the odd primes q <= S (S = sqrt(hi) up to 2^25,
// else 2^22),
// primes below SMALL_Q marked by the whole block and larger ones by one
// thread each (each prime's first multiple in the segment from a remainder
// in double precision, about a third of the instructions of a 64-bit
// integer remainder); survivors at or above S^2 are proved by deterministic
// Miller-Rabin (bases 2, 325, 9375, 28178, 450775, 9780504, 1795265022,
// exact for 64-bit n).  A scan of the segment counts then places each
// segment's offsets.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../host/host.h"
#include "gpuprimes.h"
#include "mont.h"

// Segment shapes: 2^18 odd numbers (a 32 KiB LDS bitmap) per workgroup of
// 256 threads, or, for ranges of at least 512 such wide segments (2^30
// numbers), 2^20 (128 KiB) per workgroup of 1024.  Every sieving prime is
// visited once per segment, so the wide shape makes a quarter of the visits
// (2.4 ms against 4.8 per 2^30 numbers at 1e14, with the remainders below);
// narrow ranges keep the small shape, whose segments spread over more CUs.
constexpr u32 kSegSmall = 1u << 18, kSegWide = 1u << 20;
#define SMALL_Q 4096u  // primes below this are marked block-wide
constexpr int kMaxSmall = 1024;  // room for them (563 odd primes below SMALL_Q)

// v mod q for any 64-bit v and an odd q < 2^26, in double precision: with
// v = a 2^32 + b, every quotient below is under 2^32 (x = (a mod q)
// (2^32 mod q) + b < q^2 + 2^32), so a reciprocal within 2^-48 of 1/q makes
// each estimate off by less than one, and the remainders, in (-q, 2q), are
// exact in 32-bit arithmetic.  The reciprocal is the hardware estimate with
// two Newton steps (each squares the relative error).
__device__ __forceinline__ u32 mod_small(u64 v, u32 q) {
    const double qd = (double)q;
    double rq = __builtin_amdgcn_rcp(qd);
    rq = __builtin_fma(__builtin_fma(-qd, rq, 1.0), rq, rq);
    rq = __builtin_fma(__builtin_fma(-qd, rq, 1.0), rq, rq);
    auto fix = [&](u32 t, u32 qe) -> u32 {  // t - qe q, taken into [0, q)
        int r = (int)(t - qe * q);
        if (r < 0) r += (int)q;
        else if (r >= (int)q) r -= (int)q;
        return (u32)r;
    };
    auto red32 = [&](u32 t) { return fix(t, (u32)((double)t * rq)); };
    const u32 c = red32(0u - q);  // 2^32 mod q = (2^32 - q) mod q
    const u64 x = (u64)red32((u32)(v >> 32)) * c + (v & 0xffffffffull);
    return fix((u32)x, (u32)((double)x * rq));
}

__device__ __forceinline__ bool mr_prime(u64 n) {
    // n odd, n > 2^20 here (only survivors above S^2 are tested)
    const u64 mp = mont_mprime(n);
    const u64 one = (0ull - n) % n;
    const u64 mone = n - one;                      // Montgomery form of n - 1
    u64 r2 = one << 1;
    if (r2 >= n) r2 -= n;
    for (int s = 0; s < 6; ++s) r2 = montmul(r2, r2, n, mp);
    u64 d = n - 1;
    int s = 0;
    while ((d & 1) == 0) {
        d >>= 1;
        ++s;
    }
    const u64 bases[7] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
    for (int k = 0; k < 7; ++k) {
        const u64 a = bases[k] % n;
        if (a == 0) continue;
        u64 x = one, b = montmul(a, r2, n, mp);
        for (u64 e = d; e; e >>= 1) {
            if (e & 1) x = montmul(x, b, n, mp);
            b = montmul(b, b, n, mp);
        }
        if (x == one || x == mone) continue;
        bool comp = true;
        for (int i = 1; i < s && comp; ++i) {
            x = montmul(x, x, n, mp);
            if (x == mone) comp = false;
        }
        if (comp) return false;
    }
    return true;
}

// Sieves segment blockIdx.x of the odd numbers base + 2i, i < nodd; writes
// the bitmap (bit set = composite or excluded) and the prime count.
template <u32 SEG_BITS, int NT>
__device__ __forceinline__ void prime_sieve(uint64_t base, long long nodd,
                                            const uint32_t* __restrict__ sp, int np, int nsmall,
                                            uint64_t s2, uint32_t* __restrict__ bitmap,
                                            uint32_t* __restrict__ counts) {
    constexpr u32 SEG_WORDS = SEG_BITS / 32;
    __shared__ uint32_t bm[SEG_WORDS];
    __shared__ uint32_t red[NT];
    __shared__ uint32_t i0s[kMaxSmall];  // first index of each small prime
    const int tid = threadIdx.x;
    const long long i_lo = (long long)blockIdx.x * SEG_BITS;
    const u32 nbits = (u32)min((long long)SEG_BITS, nodd - i_lo);
    const u64 v_lo = base + 2ull * (u64)i_lo;
    const u64 v_hi = v_lo + 2ull * (nbits - 1);
    for (u32 w = tid; w < SEG_WORDS; w += NT) bm[w] = 0;
    __syncthreads();
    // first index i with q | v_lo + 2i and v_lo + 2i >= q^2 (q odd)
    auto first = [&](u32 q) -> u64 {
        const u64 qq = (u64)q * q;
        if (qq >= v_lo) return (qq - v_lo) >> 1;
        const u32 r = mod_small(v_lo, q);
        const u32 d = r == 0 ? 0u : q - r;
        return (d & 1) ? (u64)((d + q) >> 1) : (u64)(d >> 1);
    };
    // small primes: each one's first index once per block (a thread per
    // prime; ~0u where q^2 is past the segment, and so for every later
    // prime), then the whole block marks each one
    for (int k = tid; k < nsmall; k += NT) {
        const u32 q = sp[k];
        i0s[k] = (u64)q * q > v_hi ? ~0u : (u32)min(first(q), (u64)nbits);
    }
    __syncthreads();
    for (int k = 0; k < nsmall; ++k) {
        const u32 i0 = i0s[k];
        if (i0 == ~0u) break;
        const u32 q = sp[k];
        for (u64 i = i0 + (u64)tid * q; i < nbits; i += (u64)NT * q)
            atomicOr(&bm[i >> 5], 1u << (i & 31));
    }
    // large primes: one thread each (at most SEG_BITS / SMALL_Q marks), four
    // of a thread's primes loaded together so their loads overlap
    bool done = false;
    for (int k0 = nsmall + tid; k0 < np && !done; k0 += 4 * NT) {
        u32 qv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) qv[j] = k0 + j * NT < np ? sp[k0 + j * NT] : 0u;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const u32 q = qv[j];
            if (done || q == 0 || (u64)q * q > v_hi) {
                done = true;
                continue;
            }
            for (u64 i = first(q); i < nbits; i += q) atomicOr(&bm[i >> 5], 1u << (i & 31));
        }
    }
    __syncthreads();
    // exclusions: 1, the tail past nbits, survivors >= S^2 that fail MR
    u32 cnt = 0;
    for (u32 w = tid; w < SEG_WORDS; w += NT) {
        u32 x = bm[w];
        const u32 b0 = w * 32;
        if (b0 >= nbits) x = ~0u;
        else if (nbits - b0 < 32) x |= ~0u << (nbits - b0);
        if (b0 == 0 && v_lo == 1) x |= 1u;
        u32 free_bits = ~x;
        if (free_bits && v_lo + 2ull * (b0 + 31) >= s2) {
            while (free_bits) {
                const int b = __builtin_ctz(free_bits);
                free_bits &= free_bits - 1;
                const u64 v = v_lo + 2ull * (b0 + b);
                if (v >= s2 && !mr_prime(v)) x |= 1u << b;
            }
        }
        bitmap[(i_lo >> 5) + w] = x;
        cnt += __popc(~x);
    }
    red[tid] = cnt;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    if (tid == 0) counts[blockIdx.x] = red[0];
}

extern "C" __global__ void __launch_bounds__(256)
rs_prime_sieve_kernel(uint64_t base, long long nodd, const uint32_t* __restrict__ sp, int np,
                      int nsmall, uint64_t s2, uint32_t* __restrict__ bitmap,
                      uint32_t* __restrict__ counts) {
    prime_sieve<kSegSmall, 256>(base, nodd, sp, np, nsmall, s2, bitmap, counts);
}

extern "C" __global__ void __launch_bounds__(1024)
rs_prime_sieve_wide_kernel(uint64_t base, long long nodd, const uint32_t* __restrict__ sp, int np,
                           int nsmall, uint64_t s2, uint32_t* __restrict__ bitmap,
                           uint32_t* __restrict__ counts) {
    prime_sieve<kSegWide, 1024>(base, nodd, sp, np, nsmall, s2, bitmap, counts);
}

// Exclusive scan of the segment counts in place (one block); total in
// counts[nseg].
extern "C" __global__ void __launch_bounds__(1024)
rs_prime_scan_kernel(uint32_t* __restrict__ counts, int nseg) {
    __shared__ uint32_t part[1024];
    const int tid = threadIdx.x;
    const int per = (nseg + 1023) / 1024;
    const int a = tid * per, b = min(nseg, a + per);
    u32 s = 0;
    for (int i = a; i < b; ++i) s += counts[i];
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const u32 v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    u32 run = part[tid] - s;  // exclusive prefix of this thread's slice
    for (int i = a; i < b; ++i) {
        const u32 c = counts[i];
        counts[i] = run;
        run += c;
    }
    if (tid == 1023) counts[nseg] = part[1023];
}

// Writes the offsets of segment blockIdx.x starting at counts[seg].
template <u32 SEG_BITS, int NT>
__device__ __forceinline__ void prime_emit(uint64_t base, uint64_t lo,
                                           const uint32_t* __restrict__ bitmap,
                                           const uint32_t* __restrict__ counts,
                                           uint32_t* __restrict__ out, long long cap) {
    constexpr u32 SEG_WORDS = SEG_BITS / 32;
    __shared__ uint32_t part[NT];
    const int tid = threadIdx.x;
    const long long w0 = (long long)blockIdx.x * SEG_WORDS;
    // each thread owns SEG_WORDS / NT consecutive words
    constexpr int per = SEG_WORDS / NT;
    u32 c = 0;
    for (int k = 0; k < per; ++k) c += __popc(~bitmap[w0 + tid * per + k]);
    part[tid] = c;
    __syncthreads();
    for (int off = 1; off < NT; off <<= 1) {
        const u32 v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    long long at = (long long)counts[blockIdx.x] + part[tid] - c;
    const u64 d0 = base - lo;
    for (int k = 0; k < per; ++k) {
        const long long w = w0 + tid * per + k;
        u32 f = ~bitmap[w];
        while (f) {
            const int b = __builtin_ctz(f);
            f &= f - 1;
            if (at < cap) out[at] = (uint32_t)(d0 + 2ull * (u64)(w * 32 + b));
            ++at;
        }
    }
}

extern "C" __global__ void __launch_bounds__(256)
rs_prime_emit_kernel(uint64_t base, uint64_t lo, const uint32_t* __restrict__ bitmap,
                     const uint32_t* __restrict__ counts, uint32_t* __restrict__ out,
                     long long cap) {
    prime_emit<kSegSmall, 256>(base, lo, bitmap, counts, out, cap);
}

extern "C" __global__ void __launch_bounds__(256)
rs_prime_emit_wide_kernel(uint64_t base, uint64_t lo, const uint32_t* __restrict__ bitmap,
                          const uint32_t* __restrict__ counts, uint32_t* __restrict__ out,
                          long long cap) {
    prime_emit<kSegWide, 256>(base, lo, bitmap, counts, out, cap);
}

#define HIPCHK(x)                                 \
    do {                                          \
        if ((x) != hipSuccess) return RS_ERR_HIP; \
    } while (0)

extern "C" __global__ void rs_zero_u32(uint32_t*, int);

struct rs_gpu_primegen {
    int device = 0;  // buffers below are reused by every call
    uint64_t hi_max = 0;
    uint64_t s2 = 0;           // survivors >= s2 need Miller-Rabin
    uint32_t* d_sp = nullptr;  // odd sieving primes
    int np = 0, nsmall = 0;
    uint32_t* d_bitmap = nullptr;
    size_t bitmap_words = 0;
    uint32_t* d_counts = nullptr;
    int counts_cap = 0;
    hipStream_t stream = nullptr;
};

extern "C" rs_status rs_gpu_primegen_create(int device, uint64_t hi_max, rs_gpu_primegen** out) {
    if (!out || hi_max < 3 || hi_max > (1ull << 63)) return RS_ERR_ARG;
    *out = nullptr;
    HIPCHK(hipSetDevice(device));
    rs_gpu_primegen* g = new rs_gpu_primegen();
    g->device = device;
    g->hi_max = hi_max;
    // sieving primes up to sqrt(hi_max) when that is at most 2^25 (then no
    // survivor needs Miller-Rabin: below 2^50 a window of 2^30 costs more
    // in Miller-Rabin than in sieving by the larger primes); above, MR is
    // needed anyway and the primes up to 2^22 keep the sieve light
    uint64_t S = (uint64_t)std::sqrt((double)hi_max) + 2;
    if (S > (1ull << 25)) S = 1ull << 22;
    std::vector<uint32_t> ps = rsh::primes_upto((uint32_t)S);
    std::vector<uint32_t> odd(ps.begin() + 1, ps.end());
    g->np = (int)odd.size();
    g->nsmall = (int)(std::lower_bound(odd.begin(), odd.end(), SMALL_Q) - odd.begin());
    if (g->nsmall > kMaxSmall) {  // the sieve kernel keeps one LDS word per small prime
        rs_gpu_primegen_free(g);
        return RS_ERR_ARG;
    }
    const uint64_t last = odd.empty() ? 3 : odd.back();
    g->s2 = last * last;
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&g->d_sp, std::max<size_t>(4, odd.size() * 4)) != hipSuccess ||
        (odd.size() && hipMemcpy(g->d_sp, odd.data(), odd.size() * 4, hipMemcpyHostToDevice) !=
                           hipSuccess)) {
        rs_gpu_primegen_free(g);
        return RS_ERR_HIP;
    }
    *out = g;
    return RS_OK;
}

rs_status gpu_primes_enqueue(rs_gpu_primegen* g, uint64_t lo, uint64_t hi, uint32_t* out,
                             int64_t cap, hipStream_t stream, const uint32_t** count_dev) {
    if (!g || hi < lo || hi - lo > (1ull << 32) || lo < 3 || hi > g->hi_max ||
        (cap > 0 && !out) || !count_dev)
        return RS_ERR_ARG;
    const uint64_t base = lo | 1;
    const long long nodd = base >= hi ? 0 : (long long)((hi - base + 1) / 2);
    const bool wide = nodd >= 512ll * kSegWide;
    const long long seg_bits = wide ? kSegWide : kSegSmall;
    const int nseg = (int)std::max<long long>(1, (nodd + seg_bits - 1) / seg_bits);
    const size_t words = (size_t)nseg * (size_t)(seg_bits / 32);
    // buffers are reused: earlier work on any stream must not still read them
    if (words > g->bitmap_words || nseg + 1 > g->counts_cap) HIPCHK(hipDeviceSynchronize());
    if (words > g->bitmap_words) {
        (void)hipFree(g->d_bitmap);
        g->d_bitmap = nullptr;
        if (hipMalloc(&g->d_bitmap, words * 4) != hipSuccess) return RS_ERR_NOMEM;
        g->bitmap_words = words;
    }
    if (nseg + 1 > g->counts_cap) {
        (void)hipFree(g->d_counts);
        g->d_counts = nullptr;
        if (hipMalloc(&g->d_counts, (size_t)(nseg + 1) * 4) != hipSuccess) return RS_ERR_NOMEM;
        g->counts_cap = nseg + 1;
    }
    if (nodd == 0) {
        hipLaunchKernelGGL(rs_zero_u32, dim3(1), dim3(64), 0, stream, g->d_counts + nseg, 1);
        HIPCHK(hipGetLastError());
    } else {
        if (wide)
            hipLaunchKernelGGL(rs_prime_sieve_wide_kernel, dim3(nseg), dim3(1024), 0, stream, base,
                               nodd, g->d_sp, g->np, g->nsmall, g->s2, g->d_bitmap, g->d_counts);
        else
            hipLaunchKernelGGL(rs_prime_sieve_kernel, dim3(nseg), dim3(256), 0, stream, base, nodd,
                               g->d_sp, g->np, g->nsmall, g->s2, g->d_bitmap, g->d_counts);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(rs_prime_scan_kernel, dim3(1), dim3(1024), 0, stream, g->d_counts,
                           nseg);
        HIPCHK(hipGetLastError());
        if (wide)
            hipLaunchKernelGGL(rs_prime_emit_wide_kernel, dim3(nseg), dim3(256), 0, stream, base,
                               lo, g->d_bitmap, g->d_counts, out, (long long)cap);
        else
            hipLaunchKernelGGL(rs_prime_emit_kernel, dim3(nseg), dim3(256), 0, stream, base, lo,
                               g->d_bitmap, g->d_counts, out, (long long)cap);
        HIPCHK(hipGetLastError());
    }
    *count_dev = g->d_counts + nseg;
    return RS_OK;
}

uint64_t rs_gpu_primegen_hi(const rs_gpu_primegen* g) { return g->hi_max; }

int64_t gpu_primes_capacity(uint64_t width) {
    // pi(x + y) - pi(x) <= 2 y / ln y (Montgomery-Vaughan), plus slack
    if (width < 64) return 64;
    return (int64_t)(2.0 * (double)width / std::log((double)width)) + 64;
}

// Sieves [lo, hi) into offsets_dev (device memory, capacity cap) and returns
// the count (RS_ERR_OVERFLOW if it exceeds cap; nothing past cap is written).
extern "C" rs_status rs_gpu_primes_range(rs_gpu_primegen* g, uint64_t lo, uint64_t hi,
                                         uint32_t* offsets_dev, int64_t cap, int64_t* count) {
    if (!g || !count) return RS_ERR_ARG;
    *count = 0;
    HIPCHK(hipSetDevice(g->device));
    const uint32_t* cnt_dev = nullptr;
    rs_status st = gpu_primes_enqueue(g, lo, hi, offsets_dev, cap, g->stream, &cnt_dev);
    if (st != RS_OK) return st;
    uint32_t total = 0;
    HIPCHK(hipMemcpyAsync(&total, cnt_dev, 4, hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    *count = total;
    return (int64_t)total > cap ? RS_ERR_OVERFLOW : RS_OK;
}

extern "C" void rs_gpu_primegen_free(rs_gpu_primegen* g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    (void)hipFree(g->d_sp);
    (void)hipFree(g->d_bitmap);
    (void)hipFree(g->d_counts);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

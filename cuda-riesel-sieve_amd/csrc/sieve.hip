// sieve.hip -- BSGS discrete-log sieve kernel for gfx950.
//
// Solves k*b^n == 1 (mod p) for every row (k, r), n = Q*x + r, of a job and
// every prime p of a chunk: the computation of the reference's addKernel1
// (kernel_v0.24.6.cu:273-684), laid out for CDNA4 as one persistent kernel.
// Each wavefront repeatedly takes a group of 64 consecutive primes of one
// chunk from a device work counter and runs two phases:
//
//  set-up, one lane per prime: Montgomery constants, g = b^Q, the baby-step
//    start powers, the BH[r] and G^t factors of the giant steps and the
//    Jacobi bits of every k (the reference computes these per thread at
//    :288-356, :425-456, :491-509), written as a record to the wave's global
//    scratch;
//  BSGS, the whole wavefront per prime: the 64 lanes fill the baby table
//    {g^j : 0 <= j < m} in LDS (8-slot buckets of 16-bit fingerprints; the
//    reference keeps a chained table per thread in global memory,
//    :358-411), form the Y table b^r*H*G^t, compact the rows that pass the
//    Jacobi filter (:516-521) and spread the (row, t) giant steps over the
//    lanes; a giant step is one 32x64 Montgomery product k*Y and one bucket
//    read (:534-608).  A fingerprint match is resolved by an exact rescan of
//    the baby steps (the reference rescans on a 16-bit key match too, :578)
//    and every solution is appended to the hit buffer (the reference keeps
//    the last hit per prime only, :615-621).
//
// Giant step t of row (k, r) holds s = k*b^r*g^(x0 + m - 1 + t*m); s = g^j
// means x = x0 + t*m + (m - 1 - j), so t in [0, T) covers x in
// [x0, x0 + T*m - 1] which contains [floor(minN/Q), floor(maxN/Q)].
//
// Chunks and jobs come either by value (c0, j0: one job, one chunk) or from
// device tables (batched launches of many jobs).
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "mont.h"

// Orders LDS traffic between the lanes of one wavefront.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Makes the wavefront's global stores visible to its other lanes.
__device__ __forceinline__ void wave_sync_global() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ u64 readlane64(u64 x, int l) {
    const u32 lo = __builtin_amdgcn_readlane(lo32(x), l);
    const u32 hi = __builtin_amdgcn_readlane(hi32(x), l);
    return ((u64)hi << 32) | lo;
}

// idx / T for idx < 2^16: umulhi with invT = ceil(2^32 / T) when T >= 2
// (invT is 0 for T == 1, where the quotient is idx itself).
__device__ __forceinline__ int div_T(u32 idx, u32 invT) {
    return invT ? (int)__umulhi(idx, invT) : (int)idx;
}

// 16-bit table fingerprint of a residue: 15 bits above the bucket index
// with the top bit set, so it is never 0, the empty-slot value.
__device__ __forceinline__ u32 fingerprint(u64 v, int logNB) {
    return (lo32(v >> logNB) & 0x7fffu) | 0x8000u;
}

// Nonzero iff some 16-bit half of x is zero (exact as a boolean: the
// lowest zero half takes no borrow from below).
__device__ __forceinline__ u32 has_zero_half(u32 x) {
    return (x - 0x10001u) & ~x & 0x80008000u;
}

// Emits every baby step j with g^j == target (all 64 lanes take part:
// lane l checks j = l + 64 v) as a hit of giant step tt of row `row`
// (k, r).  With n-list bitmaps only n in the row's list is emitted.
__device__ __forceinline__ void rescan(const JobDev& J, int job, u64 target, u32 k, u32 ry,
                                       int tt, u64 vstart, u64 g64, u64 p, u64 mp,
                                       HitDev* __restrict__ hits, u32* __restrict__ hit_count,
                                       u32 hit_cap) {
    const int lane = threadIdx.x & 63;
    const u32 r = ry & 0xffu, row = ry >> 8;
    u64 w = vstart;
    for (int v = 0; v < (J.m >> 6); ++v) {
        if (w == target) {
            const int j = lane + 64 * v;
            const long long xx = (long long)J.x0 + (long long)tt * J.m + (J.m - 1 - j);
            const long long n = (long long)J.Q * xx + (int)r;
            bool keep = n >= J.minN && n <= J.maxN;
            if (keep && J.bitmap) {
                const long long xp = xx - J.x0;  // x' of the bitmap, < 32 words
                keep = (J.bitmap[(long long)row * J.words + (xp >> 5)] >> (31 - (xp & 31))) & 1;
            }
            if (keep) {
                const u32 at = atomicAdd(hit_count, 1u);
                if (at < hit_cap) {
                    HitDev h;
                    h.p = p;
                    h.k = k;
                    h.n = (u32)n;
                    h.job = (u32)job;
                    h.pad = 0;
                    hits[at] = h;
                }
            }
        }
        w = montmul(w, g64, p, mp);
    }
}

// Giant steps of one prime: the A*T grid of (active row a, step t), t
// fastest, is cut into runs of TW consecutive items per lane and 64*TW per
// round.  Each item is one 32x64 Montgomery product s = k*Y[r][t] and one
// bucket read; all TW of a lane are issued before any is tested.  A full
// bucket (slot 7 taken) continues into the next bucket, and a fingerprint
// match is resolved by rescan(); both are rare and handled after the fast
// path.  The loop is wave-uniform because rescan() needs every lane.
template <int TW>
__device__ __forceinline__ void giant_steps(const JobDev& J, int job, int A,
                                            const uint4* __restrict__ tab,
                                            const u64* __restrict__ Y,
                                            const uint2* __restrict__ list, u64 vstart, u64 g64,
                                            u64 p, u64 mp, HitDev* __restrict__ hits,
                                            u32* __restrict__ hit_count, u32 hit_cap) {
    const int lane = threadIdx.x & 63;
    const u32 NB = 1u << J.logNB;
    const int T = J.T;
    const int total = A * T;
    const u32 mp0 = lo32(mp);
    for (int base = 0; base < total; base += 64 * TW) {
        const int i0 = base + TW * lane;
        int a = div_T((u32)i0, J.invT), t = i0 - a * T;
        u64 s[TW];
        uint4 e[TW];
        u32 fp2[TW], bk[TW], ku[TW], ru[TW];
        int tu[TW];
        bool ok[TW];
#pragma unroll
        for (int u = 0; u < TW; ++u) {
            ok[u] = i0 + u < total;
            const int ac = min(a, A - 1);
            const uint2 row = list[ac];
            ku[u] = row.x;
            ru[u] = row.y;
            tu[u] = t;
            s[u] = halfmul(row.x, Y[(row.y & 0xffu) * T + min(t, T - 1)], p, mp0);
            bk[u] = lo32(s[u]) & (NB - 1);
            fp2[u] = fingerprint(s[u], J.logNB) * 0x10001u;
            if (++t == T) { t = 0; ++a; }
        }
#pragma unroll
        for (int u = 0; u < TW; ++u) e[u] = tab[bk[u]];
        u32 mbits = 0, fbits = 0;
#pragma unroll
        for (int u = 0; u < TW; ++u) {
            const u32 z = has_zero_half(e[u].x ^ fp2[u]) | has_zero_half(e[u].y ^ fp2[u]) |
                          has_zero_half(e[u].z ^ fp2[u]) | has_zero_half(e[u].w ^ fp2[u]);
            mbits |= (u32)(ok[u] && z != 0) << u;
            fbits |= (u32)(ok[u] && e[u].w >= 0x10000u) << u;  // slot 7 taken
        }
        // full buckets: follow the chain to the first bucket with slot 7 free
        if (__ballot(fbits != 0)) {
#pragma unroll
            for (int u = 0; u < TW; ++u) {
                if ((fbits >> u) & 1) {
                    u32 b = bk[u];
                    while (true) {
                        b = (b + 1) & (NB - 1);
                        const uint4 f = tab[b];
                        const u32 z = has_zero_half(f.x ^ fp2[u]) | has_zero_half(f.y ^ fp2[u]) |
                                      has_zero_half(f.z ^ fp2[u]) | has_zero_half(f.w ^ fp2[u]);
                        if (z) mbits |= 1u << u;
                        if (f.w < 0x10000u) break;
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < TW; ++u) {
            u64 mb = __ballot((mbits >> u) & 1);
            while (mb) {
                const int src = __builtin_ctzll(mb);
                mb &= mb - 1;
                rescan(J, job, readlane64(s[u], src), __builtin_amdgcn_readlane(ku[u], src),
                       __builtin_amdgcn_readlane(ru[u], src),
                       __builtin_amdgcn_readlane(tu[u], src), vstart, g64, p, mp, hits,
                       hit_count, hit_cap);
            }
        }
    }
}

// Set-up of one prime (one lane): writes the record described in kernels.h.
__device__ __forceinline__ void prep_record(const JobDev& J, u64 p, u64* __restrict__ R) {
    const u64 mp = mont_mprime(p);
    const u64 one = (0ull - p) % p;                  // 2^64 mod p
    u64 r2 = one << 1;                               // Mont(2)
    if (r2 >= p) r2 -= p;
#pragma unroll
    for (int s = 0; s < 6; ++s) r2 = montmul(r2, r2, p, mp);  // Mont(2^64) = R^2
    const u64 B = montmul(J.base, r2, p, mp);        // Mont(b)
    const u64 g = mont_pow(B, (u32)J.Q, one, p, mp);
    // baby-step start powers: lane l of the BSGS wave starts at
    // g^(l & 7) * g^(8 (l >> 3))
    u64 x = one;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        R[REC_GP + c] = x;
        x = montmul(x, g, p, mp);
    }
    const u64 g8 = x;
    x = one;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        R[REC_G8P + a] = x;
        x = montmul(x, g8, p, mp);
    }
    const u64 g64 = x;
    const u64 G = mont_pow(g64, (u32)(J.m >> 6), one, p, mp);
    const u64 H = mont_pow(g, (u32)(J.x0 + J.m - 1), one, p, mp);
    const u64 c32 = montmul(1ull << 32, r2, p, mp);  // Mont(2^32)
    u64 bh = montmul(H, c32, p, mp);
    for (int r = 0; r < J.Q; ++r) {
        R[REC_BH + r] = bh;
        bh = montmul(bh, B, p, mp);
    }
    u64 gt = one;
    for (int t = 0; t < J.T; ++t) {
        R[J.off_gt + t] = gt;
        gt = montmul(gt, G, p, mp);
    }
    // Jacobi bits; a prime dividing b gets none, so no row is searched
    const int jb = B == 0 ? 0 : jacobi32(J.base, p);
    for (int w = 0; w < J.nk; w += 64) {
        u64 jk = 0, jkb = 0;
        const int e = min(64, J.nk - w);
        for (int i2 = 0; i2 < e; ++i2) {
            const int jkv = jb == 0 ? 0 : jacobi32(J.kvals[w + i2], p);
            jk |= (u64)(jkv == 1) << i2;
            jkb |= (u64)(jkv * jb == 1) << i2;
        }
        R[J.off_j + 2 * (w >> 6)] = jk;
        R[J.off_j + 2 * (w >> 6) + 1] = jkb;
    }
    R[REC_P] = p;
    R[REC_MP] = mp;
    R[REC_ONE] = one;
    R[REC_G64] = g64;
}

// BSGS of one prime by the whole wavefront.  `pre` holds record words
// 0..63 (lane l has word l), the rest are read from Rg; LDS at wbase.
__device__ __forceinline__ void bsgs_prime(const JobDev& J, int job, u64 pre,
                                           const u64* __restrict__ Rg, u32* __restrict__ wbase,
                                           HitDev* __restrict__ hits,
                                           u32* __restrict__ hit_count, u32 hit_cap) {
    const int lane = threadIdx.x & 63;
    const u32 NB = 1u << J.logNB;
    const int Q = J.Q, T = J.T, m = J.m, W = J.W;
    uint4* const tab = (uint4*)wbase;
    uint16_t* const tab16 = (uint16_t*)wbase;
    u32* const cnt = wbase + 4 * NB;
    u64* const Y = (u64*)(cnt + NB / 2);
    uint2* const list = (uint2*)(Y + Q * T);
    u64* const R = (u64*)(list + min(J.nrows, ROW_SLICE));
    const int L = m >> 6;

    if (lane < W) R[lane] = pre;
    for (int w = 64 + lane; w < W; w += 64) R[w] = Rg[w];
    const u64 p = readlane64(pre, REC_P), mp = readlane64(pre, REC_MP);
    const u64 g64 = readlane64(pre, REC_G64);

    // clear table and counts
    for (u32 b = lane; b < NB; b += 64) tab[b] = make_uint4(0, 0, 0, 0);
    for (u32 c = lane; c < NB / 2; c += 64) cnt[c] = 0;
    wave_sync();

    // baby steps: lane l inserts g^j, j = l + 64 s, as a fingerprint in the
    // first free slot of its bucket (or of the next non-full bucket)
    const u64 vstart = montmul(R[REC_GP + (lane & 7)], R[REC_G8P + (lane >> 3)], p, mp);
    u64 v = vstart;
    for (int s = 0; s < L; ++s) {
        u32 b = lo32(v) & (NB - 1);
        const u32 fp = fingerprint(v, J.logNB);
        while (true) {
            const u32 sh = (b & 1) * 16;
            const u32 c = (atomicAdd(&cnt[b >> 1], 1u << sh) >> sh) & 0xffffu;
            if (c < 8) {
                tab16[b * 8 + c] = (uint16_t)fp;
                break;
            }
            b = (b + 1) & (NB - 1);
        }
        if (s + 1 < L) v = montmul(v, g64, p, mp);
    }

    // Y[r*T + t] = b^r H G^t 2^32 (Montgomery form)
    for (int idx = lane; idx < Q * T; idx += 64) {
        const int r = div_T((u32)idx, J.invT), t = idx - r * T;
        Y[idx] = montmul(R[REC_BH + r], R[J.off_gt + t], p, mp);
    }

    // rows in slices of at most ROW_SLICE: those passing the Jacobi filter
    // are compacted in order as (k, row << 8 | r), then their giant steps run
    for (int r0 = 0; r0 < J.nrows; r0 += ROW_SLICE) {
        const int r1 = min(J.nrows, r0 + ROW_SLICE);
        int A = 0;
        for (int c = r0; c < r1; c += 64) {
            const int row = c + lane;
            bool pass = false;
            u32 k = 0, r = 0;
            if (row < r1) {
                const u32 info = J.rowinfo[row];
                const u32 kidx = info & 0xffffu;
                r = info >> 16;
                k = J.rowk[row];
                const u64 w = R[J.off_j + 2 * (kidx >> 6) + (r & 1)];
                pass = (w >> (kidx & 63)) & 1;
            }
            const u64 mask = __ballot(pass);
            if (pass)
                list[A + __popcll(mask & ((1ull << lane) - 1))] = make_uint2(k, ((u32)row << 8) | r);
            A += __popcll(mask);
        }
        wave_sync();
        if (J.tw >= 4)
            giant_steps<4>(J, job, A, tab, Y, list, vstart, g64, p, mp, hits, hit_count, hit_cap);
        else if (J.tw == 2)
            giant_steps<2>(J, job, A, tab, Y, list, vstart, g64, p, mp, hits, hit_count, hit_cap);
        else
            giant_steps<1>(J, job, A, tab, Y, list, vstart, g64, p, mp, hits, hit_count, hit_cap);
        wave_sync();  // the next slice rewrites the list
    }
    wave_sync();  // the next prime reuses the table
}

extern "C" __global__ void __launch_bounds__(256)
rs_sieve_kernel(JobDev j0, ChunkDev c0, const JobDev* __restrict__ jobs,
                const ChunkDev* __restrict__ chunks, int nchunks, long long ngroups,
                int lds_words, int scratch_words, uint64_t* __restrict__ scratch,
                unsigned long long* __restrict__ work, HitDev* __restrict__ hits,
                uint32_t* __restrict__ hit_count, uint32_t hit_cap) {
    extern __shared__ uint32_t lds[];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wpb = blockDim.x >> 6;
    u32* const wbase = lds + wv * lds_words;
    u64* const S = scratch + ((long long)blockIdx.x * wpb + wv) * scratch_words;

    while (true) {
        unsigned long long grp = 0;
        if (lane == 0) grp = atomicAdd(work, 1ull);
        grp = ((unsigned long long)__builtin_amdgcn_readfirstlane((u32)(grp >> 32)) << 32) |
              __builtin_amdgcn_readfirstlane((u32)grp);
        if ((long long)grp >= ngroups) break;
        // the chunk holding this group (chunks sorted by group0)
        int ci = 0;
        if (chunks) {
            int hi = nchunks - 1;
            while (ci < hi) {
                const int mid = (ci + hi + 1) >> 1;
                if (chunks[mid].group0 <= (long long)grp) ci = mid;
                else hi = mid - 1;
            }
        }
        const ChunkDev C = chunks ? chunks[ci] : c0;
        const JobDev J = jobs ? jobs[C.job] : j0;
        const long long q0 = ((long long)grp - C.group0) * 64;
        const long long cnt = C.count_dev ? (long long)*C.count_dev : C.count;
        const int n = (int)min(64ll, cnt - q0);
        if (n <= 0) break;  // device-counted chunk: groups past the end are empty

        // set-up, one lane per prime
        if (lane < n) prep_record(J, C.low + C.offs[q0 + lane], S + lane * J.W);
        wave_sync_global();

        // BSGS, the wavefront per prime; record words 0..63 of the next prime
        // are prefetched into one register per lane
        u64 pre = lane < J.W ? S[lane] : 0;
        for (int i = 0; i < n; ++i) {
            const u64 cur = pre;
            if (i + 1 < n) pre = lane < J.W ? S[(i + 1) * J.W + lane] : 0;
            bsgs_prime(J, C.job, cur, S + i * J.W, wbase, hits, hit_count, hit_cap);
        }
        wave_sync_global();  // records are rewritten by the next group
    }
}

// ------------------------------------------------------------ test kernels

extern "C" __global__ void rs_test_montmul(const uint64_t* a, const uint64_t* b,
                                           const uint64_t* p, uint64_t* out, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = montmul(a[i], b[i], p[i], mont_mprime(p[i]));
}

extern "C" __global__ void rs_test_halfmul(const uint32_t* k, const uint64_t* y,
                                           const uint64_t* p, uint64_t* out, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = halfmul(k[i], y[i], p[i], lo32(mont_mprime(p[i])));
}

extern "C" __global__ void rs_test_jacobi(const uint32_t* a, const uint64_t* p, int32_t* out,
                                          long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = jacobi32(a[i], p[i]);
}

// Throughput probe for v_mad_u64_u32: 8 independent chains per lane, each
// step acc = lo32(acc) * x + c with a loop-invariant 64-bit addend c, so the
// loop body is mads only (8*iters per lane).  The result is stored so the
// chains are not removed.
extern "C" __global__ void __launch_bounds__(256) rs_mad_probe(uint64_t* out, int iters) {
    u64 acc[8];
    const u32 a = threadIdx.x * 2654435761u + blockIdx.x;
    const u64 c = ((u64)a << 17) ^ 0x9E3779B97F4A7C15ull;
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = (u64)(a + k) * 0x9E3779B9u;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[k] = mad32(lo32(acc[k]), a ^ (u32)k, c);
    }
    u64 s = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) s ^= acc[k];
    out[(long long)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// Streaming copy of n uint4 (16 B) elements, grid-stride: the HBM traffic
// calibration reference (reads and writes exactly 16 n bytes).
extern "C" __global__ void __launch_bounds__(256) rs_copy_probe(const uint4* __restrict__ src,
                                                                uint4* __restrict__ dst,
                                                                long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// acc += *c (prime counts of device-generated chunks, for timing reports).
extern "C" __global__ void rs_accum_count(const uint32_t* c, unsigned long long* acc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *acc += *c;
}

extern "C" __global__ void rs_zero_u32(uint32_t* x, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0;
}

// sieve.hip -- BSGS discrete-log sieve kernels for gfx950.
//
// Solves k*b^n == 1 (mod p) for every row (k, r), n = Q*x + r, of a job and
// every prime p of a batch, the computation of the reference's addKernel1
// (kernel_v0.24.6.cu:273-684), laid out for CDNA4:
//
//  rs_prep_kernel (one thread per prime): Montgomery constants, g = b^Q,
//    the baby-step start powers, BH[r] and G^t factors of the giant steps and
//    the Jacobi bits of every k (the reference computes these per thread at
//    :288-356, :425-456, :491-509).  Writes one record per prime.
//
//  rs_bsgs_kernel (one wavefront per prime): the 64 lanes build the baby
//    table {g^j : 0 <= j < m} in LDS (open addressing on the low residue
//    bits; the reference keeps a chained table per thread in global memory,
//    :358-411), form the Y table b^r*H*G^t, compact the rows that pass the
//    Jacobi filter (:516-521), and spread the (row, t) giant steps over the
//    lanes; each giant step is one 32x64 Montgomery product k*Y and one
//    probe (:534-608).  Every fingerprint match is checked exactly against
//    g^j before it is appended to the hit buffer (the reference keeps the
//    last hit per prime only, :615-621).
//
// Giant step t of row (k, r) holds s = k*b^r*g^(x0 + m - 1 + t*m); s = g^j
// means x = x0 + t*m + (m - 1 - j), so t in [0, T) covers x in
// [x0, x0 + T*m - 1] which contains [floor(minN/Q), floor(maxN/Q)].
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "mont.h"

#define EMPTY_SLOT 0xffffffffu

// Orders LDS traffic between the lanes of one wavefront.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

extern "C" __global__ void __launch_bounds__(256)
rs_prep_kernel(JobDev J, uint64_t low, const uint32_t* __restrict__ offs,
               long long count, uint64_t* __restrict__ rec) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const u64 p = low + offs[i];
    u64* R = rec + i * (long long)J.W;
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

extern "C" __global__ void __launch_bounds__(256)
rs_bsgs_kernel(JobDev J, const uint64_t* __restrict__ rec, long long count,
               HitDev* __restrict__ hits, uint32_t* __restrict__ hit_count,
               uint32_t hit_cap) {
    extern __shared__ uint32_t lds[];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wpb = blockDim.x >> 6;
    const u32 M = 1u << J.logM;
    const u32 jmask = (1u << J.jbits) - 1;
    const int Q = J.Q, T = J.T, m = J.m;
    u32* tab = lds + wv * J.lds_words;
    u64* Y = (u64*)(tab + M);
    u32* list = (u32*)(Y + Q * T);
    const int L = m >> 6;
    const int dA = 64 / T, dT = 64 - dA * T;

    for (long long i = (long long)blockIdx.x * wpb + wv; i < count;
         i += (long long)gridDim.x * wpb) {
        const u64* R = rec + i * (long long)J.W;
        const u64 p = R[REC_P], mp = R[REC_MP], one = R[REC_ONE], g64 = R[REC_G64];
        const u32 mp0 = lo32(mp);

        // clear the table
        for (u32 s = lane * 4; s < M; s += 256)
            *(uint4*)(tab + s) = make_uint4(EMPTY_SLOT, EMPTY_SLOT, EMPTY_SLOT, EMPTY_SLOT);
        wave_sync();

        // baby steps: lane l inserts g^j for j = l + 64 s
        u64 v = montmul(R[REC_GP + (lane & 7)], R[REC_G8P + (lane >> 3)], p, mp);
        for (int s = 0; s < L; ++s) {
            const u32 j = (u32)(lane + 64 * s);
            u32 slot = lo32(v) & (M - 1);
            const u32 entry = (lo32(v >> J.logM) << J.jbits) | j;
            while (atomicCAS(&tab[slot], EMPTY_SLOT, entry) != EMPTY_SLOT)
                slot = (slot + 1) & (M - 1);
            if (s + 1 < L) v = montmul(v, g64, p, mp);
        }

        // Y[r*T + t] = b^r H G^t 2^32 (Montgomery form)
        for (int idx = lane; idx < Q * T; idx += 64) {
            const int r = idx / T, t = idx - r * T;
            Y[idx] = montmul(R[REC_BH + r], R[J.off_gt + t], p, mp);
        }

        // rows passing the Jacobi filter, compacted in order
        int A = 0;
        for (int c = 0; c < J.nrows; c += 64) {
            const int row = c + lane;
            bool pass = false;
            if (row < J.nrows) {
                const u32 info = J.rowinfo[row];
                const u32 kidx = info & 0xffffu, r = info >> 16;
                const u64 w = R[J.off_j + 2 * (kidx >> 6) + (r & 1)];
                pass = (w >> (kidx & 63)) & 1;
            }
            const u64 mask = __ballot(pass);
            if (pass) list[A + __popcll(mask & ((1ull << lane) - 1))] = (u32)row;
            A += __popcll(mask);
        }
        wave_sync();

        // giant steps: item (a, t) of the A*T grid, lanes stride by 64
        const int total = A * T;
        int a = lane / T, t = lane - (lane / T) * T;
        for (int it = lane; it < total; it += 64) {
            const u32 row = list[a];
            const u32 k = J.rowk[row];
            const int r = (int)(J.rowinfo[row] >> 16);
            const u64 s = halfmul(k, Y[r * T + t], p, mp0);
            u32 slot = lo32(s) & (M - 1);
            const u32 fp = lo32(s >> J.logM) << J.jbits;
            while (true) {
                const u32 e = tab[slot];
                if (e == EMPTY_SLOT) break;
                if ((e & ~jmask) == fp) {
                    // exact check: g^j == s
                    const u32 j = e & jmask;
                    u64 gj = montmul(R[REC_GP + (j & 7)], R[REC_G8P + ((j >> 3) & 7)], p, mp);
                    if (j >> 6) gj = montmul(gj, mont_pow(g64, j >> 6, one, p, mp), p, mp);
                    if (gj == s) {
                        const long long xx = (long long)J.x0 + (long long)t * m + (m - 1 - (int)j);
                        const long long n = (long long)Q * xx + r;
                        if (n >= J.minN && n <= J.maxN) {
                            const u32 at = atomicAdd(hit_count, 1u);
                            if (at < hit_cap) {
                                HitDev h;
                                h.p = p;
                                h.k = k;
                                h.n = (u32)n;
                                hits[at] = h;
                            }
                        }
                    }
                }
                slot = (slot + 1) & (M - 1);
            }
            t += dT;
            a += dA;
            if (t >= T) { t -= T; a += 1; }
        }
        wave_sync();  // the next prime reuses the table
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

// Throughput probe for v_mad_u64_u32: 8 independent dependent chains per
// lane, iters steps each (8*iters mads per lane).  The result is stored so
// the chains are not removed.
extern "C" __global__ void __launch_bounds__(256) rs_mad_probe(uint64_t* out, int iters) {
    u64 acc[8];
    u32 a = threadIdx.x * 2654435761u + blockIdx.x;
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c] = (u64)(a + c) * 0x9E3779B9u;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[c] = mad32(lo32(acc[c]), a ^ (u32)c, hi32(acc[c]));
    }
    u64 s = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) s ^= acc[c];
    out[(long long)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

extern "C" __global__ void rs_zero_u32(uint32_t* x, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0;
}

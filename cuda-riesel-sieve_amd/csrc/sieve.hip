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
//    lanes; a giant step is the low word of one Shoup product k*Y (three
//    32-bit multiplies against the precomputed quotient of Y) and one bucket
//    read (:534-608).  Below p = 2^50 the baby chain, the Y table and the
//    set-up multiply in double precision (fp_mulmod), above it with exact
//    64-bit Montgomery and Shoup products.  A fingerprint match is resolved by an exact rescan of
//    the baby steps (the reference rescans on a 16-bit key match too, :578)
//    and every solution is appended to the hit buffer (the reference keeps
//    the last hit per prime only, :615-621).
//
// Giant step t of row (k, r) holds s = k*b^r*g^(x0 + m - 1 + t*m); s = g^j
// means x = x0 + t*m + (m - 1 - j), so t in [0, T) covers x in
// [x0, x0 + T*m - 1] which contains [floor(minN/Q), floor(maxN/Q)].
//
// Job constants are always read from a device table (one entry for a
// single-job launch); chunks come by value (c0: one chunk) or from a device
// table (batched launches of many jobs).
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

// Loads from the job's tables, which no kernel writes, through the
// constant address space for wave-uniform addresses (scalar loads); and
// per-lane loads through the global address space (instead of flat loads,
// which also count as LDS operations in the wait counters).  Vector types
// go through plain integer loads.
template <typename T>
__device__ __forceinline__ T const_ld(const T* p) {
    return *(const __attribute__((address_space(4))) T*)p;
}
template <typename T>
__device__ __forceinline__ T global_ld(const T* p) {
    return *(const __attribute__((address_space(1))) T*)p;
}
__device__ __forceinline__ uint2 global_ld(const uint2* p) {
    const u64 v = global_ld((const u64*)p);
    return make_uint2((u32)v, (u32)(v >> 32));
}
__device__ __forceinline__ uint4 const_ld(const uint4* p) {
    const u64 a = const_ld((const u64*)p), b = const_ld((const u64*)p + 1);
    return make_uint4((u32)a, (u32)(a >> 32), (u32)b, (u32)(b >> 32));
}

// idx / T for idx < 2^16 and 2 <= T < 2^16: umulhi with invT = ceil(2^32 / T)
// (invT = (2^32 + e) / T with e < T, so idx * invT / 2^32 exceeds idx / T by
// less than 1/T).
__device__ __forceinline__ int div_T(u32 idx, u32 invT) { return (int)__umulhi(idx, invT); }

// Table hash of a residue: the bucket is taken from bits 4 and up of its
// low word and the 16-bit fingerprint is its high half.  Only the fill
// counts are cleared for each prime, not the slots: a slot past its
// bucket's count holds a fingerprint of another prime's baby step (or
// whatever the LDS held at launch), which equals a probe's fingerprint with
// chance 2^-16 like any unrelated value.  The probe compares all eight
// slots at once; only when one matches does the bucket's count decide
// whether a filled slot matched (bucket_match_live), and only then does the
// exact rescan run.  (Clearing the table would cost more LDS write cycles
// than that rare test.)  The bucket's byte offset in the table
// is lo & (nbm << 4), which a table aligned to its size adds by OR.  (Above
// 4096 buckets the bucket bits reach into the fingerprint, whose distinct
// bits within a bucket drop below 16: only the rescan rate grows.)
__device__ __forceinline__ u32 bucket_of(u32 lo, u32 logNB) {
    return __builtin_amdgcn_ubfe(lo, 4, logNB);
}
__device__ __forceinline__ unsigned short fingerprint(u32 lo) { return (unsigned short)(lo >> 16); }

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const u32x4 lds_u32x4;

// The bucket at LDS byte address a.
__device__ __forceinline__ uint4 lds_bucket(u32 a) {
    const u32x4 v = *(lds_u32x4*)(size_t)a;
    return make_uint4(v.x, v.y, v.z, v.w);
}

// The LDS byte address of a pointer into LDS, and a uint2 at one.
template <typename T>
__device__ __forceinline__ u32 lds_addr(const T* p) {
    return (u32)(size_t)(const __attribute__((address_space(3))) T*)p;
}
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint2 lds_ld2(u32 a) {
    const u32x2 v = *(const __attribute__((address_space(3))) u32x2*)(size_t)a;
    return make_uint2(v.x, v.y);
}

// The 8 16-bit slots of e minus the fingerprint of the low word lo (its
// high half; packed 16-bit subtractions whose second operand takes the high
// half of lo for both halves, so the fingerprint needs no instruction of its
// own), reduced to their packed minimum: a zero half means a slot matched.
__device__ __forceinline__ u16x2 bucket_min(uint4 e, u32 lo) {
    const u16x2 l = __builtin_bit_cast(u16x2, lo);
    const u16x2 f = __builtin_shufflevector(l, l, 1, 1);
    const u16x2 a = __builtin_bit_cast(u16x2, e.x) - f, b = __builtin_bit_cast(u16x2, e.y) - f;
    const u16x2 c = __builtin_bit_cast(u16x2, e.z) - f, d = __builtin_bit_cast(u16x2, e.w) - f;
    return __builtin_elementwise_min(__builtin_elementwise_min(a, b),
                                     __builtin_elementwise_min(c, d));
}
// Whether a packed minimum has a zero half: folded with its halves
// swapped, its low half is zero.
__device__ __forceinline__ bool has_zero_half(u16x2 mn) {
    const u16x2 g = __builtin_elementwise_min(mn, __builtin_shufflevector(mn, mn, 1, 0));
    return g.x == 0;
}
// True iff one of the 8 16-bit slots of e equals the fingerprint of lo.
__device__ __forceinline__ bool bucket_match(uint4 e, u32 lo) {
    return has_zero_half(bucket_min(e, lo));
}

// Whether one of the first c slots of bucket e (the slots its inserts have
// filled) holds the fingerprint of lo: the exact test behind a
// bucket_match, run only when one matched (a match in a slot past the fill
// count is a stale or empty slot, which no baby step of this prime wrote).
__device__ __forceinline__ bool bucket_match_live(uint4 e, u32 lo, u32 c) {
    const u32 f = lo >> 16;
    const u32 w[4] = {e.x, e.y, e.z, e.w};
    bool m = false;
#pragma unroll
    for (int s = 0; s < 8; ++s) m |= (u32)s < c && ((w[s >> 1] >> (16 * (s & 1))) & 0xffffu) == f;
    return m;
}

// Active-row entries of the giant steps (built on the host, JobDev.rowent):
// .x = k, .y = 8*r*ys (byte offset of the row's Y run, ys the Y row stride).

// The job row of the entry (k, yoff): rows are ordered by k, then r (so by
// Y offset), and no two are equal, so a binary search of the job's row
// table finds it (only a hit in a job with n-list bitmaps needs it).
__device__ __forceinline__ int find_row(const uint2* __restrict__ rowent, int nrows, u32 k,
                                        u32 yoff) {
    int lo = 0, hi = nrows - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const uint2 e = global_ld(rowent + mid);
        if (e.x < k || (e.x == k && e.y < yoff)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Emits every baby step j with g^j == target (all 64 lanes take part:
// lane l checks j = l + 64 v) as a hit of giant step tt of the row whose
// entry is (k, yent).  With n-list bitmaps only n in the row's list is
// emitted.  p, mp and g^64 come from the
// prime's record R in LDS (a hit is rare: they need not stay in registers
// through the giant steps); g^64 there is a Montgomery form, so the chain
// keeps vstart's form: plain for p < 2^50, else Montgomery.
__device__ __forceinline__ void rescan(const JobDev& J, const JobDev* __restrict__ jobs, int job,
                                       const u64* R, u64 yv, u32 k, u32 yent, int tt,
                                       u64 vstart, HitDev* __restrict__ hits,
                                       u32* __restrict__ hit_count, u32 hit_cap) {
    const int lane = threadIdx.x & 63;
    const u32 r = (yent >> 3) / (u32)J.ys;
    const u64 p = R[REC_P], mp = R[REC_MP], g64 = R[REC_G64];
    const u64 target = shoup32_full(k, yv, p);  // the exact product of the item
    u64 w = vstart;
    for (int v = 0; v < (J.m >> 6); ++v) {
        if (w == target) {
            const int j = lane + 64 * v;
            // the job constants only a hit needs are read here, not kept
            // in registers through the giant steps
            const JobDev* const jp = jobs + job;
            const int x0 = const_ld(&jp->x0);
            const long long xx = (long long)x0 + (long long)tt * J.m + (J.m - 1 - j);
            const long long n = (long long)J.Q * xx + (int)r;
            bool keep = n >= const_ld(&jp->minN) && n <= const_ld(&jp->maxN);
            const uint32_t* const bitmap = const_ld(&jp->bitmap);
            if (keep && bitmap) {
                const long long xp = xx - x0;  // x' of the bitmap, < 32 words
                const int row = find_row(const_ld(&jp->rowent), J.nrows, k, yent);
                keep = (global_ld(bitmap + (long long)row * const_ld(&jp->words) + (xp >> 5)) >>
                        (31 - (xp & 31))) & 1;
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

// A copy of the struct at a wave-uniform address, held in scalar registers
// (read once by every lane, then taken from the first lane): job and chunk
// constants then steer loops and offsets without vector instructions.
template <typename S>
__device__ __forceinline__ S uniform_load(const S* p) {
    static_assert(sizeof(S) % 4 == 0, "struct of 32-bit words");
    S out;
    const u32* src = (const u32*)p;
    u32* dst = (u32*)&out;
#pragma unroll
    for (int i = 0; i < (int)(sizeof(S) / 4); ++i)
        dst[i] = (u32)__builtin_amdgcn_readfirstlane(src[i]);
    return out;
}

// Lane mask of a per-lane condition (straight to a scalar register pair).
__device__ __forceinline__ u64 lanes(bool c) { return __builtin_amdgcn_ballot_w64(c); }

// Whether this lane's bit of a wave-uniform mask is set (the mask becomes
// the execution mask of the guarded code, no vector instruction).
__device__ __forceinline__ bool in_mask(u64 m) { return __builtin_amdgcn_inverse_ballot_w64(m); }

// Giant steps of one prime: the A*T grid of (active row a, step t) is cut
// into runs of TW consecutive t of one row (T is a multiple of TW), one run
// per lane and 64 runs per round.  A lane reads its row entry once and its
// TW Y entries at fixed offsets; each item is the low word of one Shoup
// product s = k*Y[r][t] in [0, 2p) (shoup32_lo: three multiplies) and one
// bucket read, all TW issued before any is tested.  When s exceeds p the
// true residue is below p*kmax/2^32, and the baby table holds such
// residues a second time, plus p (bsgs_prime), so the unreduced value
// finds them.  With CHAIN (some bucket of this prime
// overflowed) an item whose bucket is full also follows the chain to the
// next buckets (fullness is read from the fill counts `cnt`).  A
// fingerprint match is resolved by an exact rescan() of
// the reduced product.  The loop is wave-uniform because rescan() needs
// every lane.
template <int TW, bool CHAIN, bool EXACT, bool PF>
__device__ __forceinline__ void giant_steps(const JobDev& J, const JobDev* __restrict__ jobs,
                                            int job, int A, int r0,
                                            u32 tab_addr,
                                            const u32* __restrict__ cnt,
                                            const char* __restrict__ Yb,
                                            const uint2* __restrict__ list,
                                            const uint2* __restrict__ rows_lds, bool idx,
                                            const u64* R, u32 negp0, u64 vstart,
                                            HitDev* __restrict__ hits, u32* __restrict__ hit_count,
                                            u32 hit_cap) {
    const int lane = threadIdx.x & 63;
    const u32 nbm = (1u << J.logNB) - 1;
    const uint16_t* const list16 = (const uint16_t*)list;
    u32 bmask = nbm << 4;
    asm volatile("" : "+v"(bmask));  // in a vector register: one and-or per address
    const int T = J.T;
    const int total = A * T;
    // each lane's run (row a, first step t0) is carried from round to
    // round: a round of 64*TW items moves it by rstep rows and tstep steps,
    // tstep = 0 when EXACT (t0 < T always, so only a is clamped for lanes
    // past the end)
    int a = div_T((u32)(TW * lane), J.invT);
    int t0 = TW * lane - a * T;
    const int rstep = div_T(64u * TW, J.invT), tstep = 64 * TW - rstep * T;
    auto advance = [&](int& aa, int& tt) {
        aa += rstep;
        if (!EXACT) {
            tt += tstep;
            if (tt >= T) tt -= T, aa += 1;
        }
    };
    // PF: the lane's list element is kept as its LDS byte address la (the
    // lanes past the end read the last element, llast), and the next
    // round's entry is read during this round, unconditionally (after the
    // round's products, into the registers they free), so its wait is a
    // round later; with idx the next round's row number is read with the
    // round's bucket reads and its entry after the match tests.  A rescan
    // reads its entry again.
    const u32 es = idx ? 2u : 8u;  // bytes per list element
    const u32 llast = lds_addr(list) + es * (A - 1);  // A >= 1 when the loop runs
    const u32 lstep = es * rstep;
    u32 la = lds_addr(list) + es * a;
    auto lds_u16 = [](u32 x) -> u32 {
        return *(const __attribute__((address_space(3))) uint16_t*)(size_t)x;
    };
    auto entry_at = [&](u32 x) { return idx ? rows_lds[lds_u16(x)] : lds_ld2(x); };
    uint2 rown = make_uint2(0, 0);
    if (PF && A > 0) rown = entry_at(min(la, llast));
    // EXACT: each lane keeps its step t0, so its Y run starts at a fixed
    // offset into every row's Y entries
    const char* const ybase = Yb + 8 * t0;
    for (int base = 0; base < total; base += 64 * TW) {
        // the row's entry: in the list, or (idx) behind its row number
        uint2 row;
        const int tcur = t0;
        const u32 lcur = la;
        if (PF) {
            row = rown;
            la += lstep;
            if (!EXACT) {
                t0 += tstep;
                if (t0 >= T) t0 -= T, la += es;
            }
        } else {
            const int ac = min(a, A - 1);
            row = idx ? rows_lds[list16[ac]] : list[ac];
        }
        // (EXACT: tested on the advanced address, so lcur need not be kept)
        const bool ok = PF ? (EXACT ? la <= llast + lstep : lcur <= llast) : a < A;
        const u64 okm = lanes(ok);
        const u64* const yp = (const u64*)((EXACT ? ybase : Yb + 8 * tcur) + row.y);
        uint4 e[TW];
        u32 lw[TW], ba[TW];
#pragma unroll
        for (int u = 0; u < TW; ++u) {
            const u32 lo = shoup32_lo(row.x, lo32(yp[u]), hi32(yp[u]), negp0);
            ba[u] = (lo & bmask) | tab_addr;  // the bucket's LDS byte address
            lw[u] = lo;
        }
        u32 rnum = 0;
        if (PF) {
            if (idx) rnum = lds_u16(min(la, llast));
            else rown = lds_ld2(min(la, llast));
        }
#pragma unroll
        for (int u = 0; u < TW; ++u) e[u] = lds_bucket(ba[u]);
        u64 mm[TW], full[TW];
        u64 anyf = 0, anym = 0;
        if (CHAIN) {
#pragma unroll
            for (int u = 0; u < TW; ++u) {
                mm[u] = lanes(bucket_match(e[u], lw[u])) & okm;
                anym |= mm[u];
                full[u] = lanes(cnt[(ba[u] - tab_addr) >> 4] >= 8) & okm;
                anyf |= full[u];
            }
        } else {
            // one test for the TW items together (a match anywhere is rare):
            // the packed minimum over all their slots has a zero half; the
            // items are told apart below, only when one matched
            u16x2 acc = bucket_min(e[0], lw[0]);
#pragma unroll
            for (int u = 1; u < TW; ++u)
                acc = __builtin_elementwise_min(acc, bucket_min(e[u], lw[u]));
            anym = lanes(has_zero_half(acc)) & okm;
#pragma unroll
            for (int u = 0; u < TW; ++u) mm[u] = 0;
        }
        if (PF && idx) rown = rows_lds[rnum];
        // rare: keep only the matches in filled slots
        if (anym) {
#pragma unroll
            for (int u = 0; u < TW; ++u) {
                if (!CHAIN) mm[u] = lanes(bucket_match(e[u], lw[u])) & okm;
                u32 c = 0;
                if (in_mask(mm[u])) c = cnt[(ba[u] - tab_addr) >> 4];
                mm[u] &= lanes(bucket_match_live(e[u], lw[u], c));
            }
        }
        // full buckets: follow the chain to the first bucket that is not full
        if (CHAIN && anyf) {
#pragma unroll
            for (int u = 0; u < TW; ++u) {
                bool hit = false;
                if (in_mask(full[u])) {
                    u32 b = (ba[u] - tab_addr) >> 4;
                    while (true) {
                        b = (b + 1) & nbm;
                        const uint4 f = lds_bucket(tab_addr + 16 * b);
                        const u32 cb = cnt[b];
                        if (bucket_match(f, lw[u]) && bucket_match_live(f, lw[u], cb)) hit = true;
                        if (cb < 8) break;
                    }
                }
                mm[u] |= lanes(hit);
            }
        }
#pragma unroll
        for (int u = 0; u < TW; ++u) {
            u64 mb = mm[u];
            while (mb) {
                const int src = __builtin_ctzll(mb);
                mb &= mb - 1;
                uint2 re;
                if (PF) {
                    const u32 at = EXACT ? __builtin_amdgcn_readlane(la, src) - lstep
                                         : __builtin_amdgcn_readlane(lcur, src);
                    re = entry_at(at);
                } else {
                    re = make_uint2(__builtin_amdgcn_readlane(row.x, src),
                                    __builtin_amdgcn_readlane(row.y, src));
                }
                const u32 k = __builtin_amdgcn_readfirstlane(re.x);
                const u32 yent = __builtin_amdgcn_readfirstlane(re.y);
                const int tt = __builtin_amdgcn_readlane(tcur, src) + u;
                const u64 yv = *(const u64*)(Yb + yent + 8 * tt);
                rescan(J, jobs, job, R, yv, k, yent, tt, vstart, hits, hit_count, hit_cap);
            }
        }
        if (!PF) advance(a, t0);
    }
}

// Rows [r0, r1) that pass the Jacobi filter of this prime (row masks at
// R + off_j, r0 a multiple of 64), compacted in order into `list` as
// giant-step entries from the job's global table or (FROM_LDS) from the
// block-shared table rows_lds, which then holds the whole job (one slice,
// r0 = 0); with idx (FROM_LDS only) as uint16 row numbers into rows_lds.
// Returns their number.
template <bool FROM_LDS>
__device__ __forceinline__ int compact_rows(const JobDev& J, const uint2* rows_lds, bool idx,
                                            const u64* __restrict__ R, uint2* __restrict__ list,
                                            int r0, int r1, int T) {
    const int lane = threadIdx.x & 63;
    int A = 0;
    // the slice's mask words (at most 8), lane w holding word w
    const int nw = (r1 - r0 + 63) >> 6;
    const u64 wm = lane < nw ? R[J.off_j + (r0 >> 6) + lane] : 0;
    // passing lanes (rows below r1 only) write their entry at their rank
    auto put = [&](int w, auto entry) {
        const u32 mlo = (u32)__builtin_amdgcn_readlane(lo32(wm), w);
        const u32 mhi = (u32)__builtin_amdgcn_readlane(hi32(wm), w);
        const u64 mask = ((u64)mhi << 32) | mlo;
        if (in_mask(mask)) {
            const u32 below = __builtin_amdgcn_mbcnt_hi(hi32(mask),
                                                        __builtin_amdgcn_mbcnt_lo(lo32(mask), 0));
            entry(A + below);
        }
        A += __popcll(mask);
    };
    if (FROM_LDS) {
        // finished entries in block-shared LDS
        uint16_t* const list16 = (uint16_t*)list;
        if (idx) {
            for (int c = r0, w = 0; c < r1; c += 64, ++w)
                put(w, [&](int at) { list16[at] = (uint16_t)(c + lane); });
        } else {
            for (int c = r0, w = 0; c < r1; c += 64, ++w)
                put(w, [&](int at) { list[at] = rows_lds[c + lane]; });
        }
    } else {
        // from the job's table in global memory: every load of the slice
        // issued before the first is used (ROW_SLICE = 8 x 64 rows)
        uint2 ent[ROW_SLICE / 64];
#pragma unroll
        for (int w = 0; w < ROW_SLICE / 64; ++w) {
            const int row = r0 + 64 * w + lane;
            ent[w] = row < r1 ? global_ld(J.rowent + row) : make_uint2(0, 0);
        }
#pragma unroll
        for (int w = 0; w < ROW_SLICE / 64; ++w)
            if (w < nw) put(w, [&](int at) { list[at] = ent[w]; });
    }
    return A;
}

// Powers of the record for p >= 2^50, as Montgomery forms (exact 64-bit
// products).  Returns whether p divides b.
__device__ __forceinline__ bool prep_powers_int(const JobDev& J, u64 p, u64 mp,
                                                u64* __restrict__ R) {
    const u64 one = (0ull - p) % p;                  // 2^64 mod p
    u64 r2 = one << 1;                               // Mont(2)
    if (r2 >= p) r2 -= p;
#pragma unroll
    for (int s = 0; s < 6; ++s) r2 = montmul(r2, r2, p, mp);  // Mont(2^64) = R^2
    const u64 B = montmul(J.base, r2, p, mp);        // Mont(b)
    const u64 g = mont_pow(B, (u32)J.Q, one, p, mp);
    // baby-step start powers: lane l of the BSGS wave starts at
    // g^(l & 7) * g^(8 (l >> 3))
    u64 w, ws;
    shoup_pair(g, p, mp, &w, &ws);
    u64 x = one;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        R[REC_GP + c] = x;
        x = shoup_mul(x, w, ws, p);
    }
    const u64 g8 = x;
    shoup_pair(g8, p, mp, &w, &ws);
    x = one;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        R[REC_G8P + a] = x;
        x = shoup_mul(x, w, ws, p);
    }
    const u64 g64 = x;
    shoup_pair(g64, p, mp, &w, &ws);
    R[REC_GW] = w;
    R[REC_GWS] = ws;
    R[REC_PD] = (u64)__double_as_longlong((double)p);
    R[REC_PINV] = 0;
    R[REC_G64] = g64;
    const u64 G = mont_pow(g64, (u32)(J.m >> 6), one, p, mp);
    const u64 H = mont_pow(g, (u32)(J.x0 + J.m - 1), one, p, mp);
    const u64 c32 = montmul(1ull << 32, r2, p, mp);  // Mont(2^32)
    // k*b^n+1: solve (-k) b^r g^x == 1, i.e. negate the giant-step starts
    u64 bh = montmul(H, c32, p, mp);
    if (J.plus && bh != 0) bh = p - bh;
    shoup_pair(B, p, mp, &w, &ws);
    for (int r = 0; r < J.Q; ++r) {
        R[REC_BH + r] = bh;
        bh = shoup_mul(bh, w, ws, p);
    }
    shoup_pair(G, p, mp, &w, &ws);
    u64 gt = one;  // Montgomery form of G^t
    for (int t = 0; t < J.T; ++t) {
        u64 gp, gs;
        shoup_pair(gt, p, mp, &gp, &gs);
        R[J.off_gt + 2 * t] = gp;
        R[J.off_gt + 2 * t + 1] = gs;
        gt = shoup_mul(gt, w, ws, p);
    }
    return B == 0;
}

// x^e mod p < 2^50 in double precision (fp_mulmod), for |x| < p and a
// wave-uniform e; the result lies in (-p, p).
__device__ __forceinline__ double fp_pow(double x, u32 e, double pd, double pinv) {
    double r = 1.0;
    for (int i = 31 - __builtin_clz(e | 1); i >= 0; --i) {
        r = fp_mulmod(r, r, pd, pinv);
        if ((e >> i) & 1) r = fp_mulmod(r, x, pd, pinv);
    }
    return r;
}

// Powers of the record for p < 2^50, as plain residues held as doubles in
// (-p, p) (the BSGS phase multiplies them with fp_mulmod), but for the
// rescan's g^64.  Returns whether p divides b.
__device__ __forceinline__ bool prep_powers_fp(const JobDev& J, u64 p, u64* __restrict__ R) {
    const double pd = (double)p, pinv = 1.0 / pd;
    auto put = [&](int w, double v) { R[w] = (u64)__double_as_longlong(v); };
    // b mod p (b < 2^32; a reduction only for p <= b) and 2^32 mod p
    const u32 b = J.base;
    const double bd = p > b ? (double)b : fp_mulmod((double)b, 1.0, pd, pinv);
    const double c32 = fp_mulmod(4294967296.0, 1.0, pd, pinv);
    const double g = fp_pow(bd, (u32)J.Q, pd, pinv);
    // baby-step start powers, as in prep_powers_int
    double x = 1.0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        put(REC_GP + c, x);
        x = fp_mulmod(x, g, pd, pinv);
    }
    const double g8 = x;
    x = 1.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        put(REC_G8P + a, x);
        x = fp_mulmod(x, g8, pd, pinv);
    }
    const double g64 = x;
    put(REC_GW, g64);
    // the rescan's multiplier: the Montgomery form g^64 2^64 mod p, an
    // integer, keeps its chain of plain residues plain
    const double g64m = fp_mulmod(g64, fp_mulmod(c32, c32, pd, pinv), pd, pinv);
    R[REC_G64] = fp_bits(g64m < 0.0 ? g64m + pd : g64m) & ((1ull << 52) - 1);
    put(REC_GWS, (double)((p >> 32) * J.kmax + J.kmax + 1));  // second-entry bound
    put(REC_PD, pd);
    put(REC_PINV, pinv);
    const double G = fp_pow(g64, (u32)(J.m >> 6), pd, pinv);
    const double H = fp_pow(g, (u32)(J.x0 + J.m - 1), pd, pinv);
    // k*b^n+1: negate the giant-step starts (-0.0 is harmless: fp_mulmod
    // products of it are zeros)
    double bh = fp_mulmod(H, c32, pd, pinv);
    if (J.plus) bh = -bh;
    for (int r = 0; r < J.Q; ++r) {
        put(REC_BH + r, bh);
        bh = fp_mulmod(bh, bd, pd, pinv);
    }
    double gt = 1.0;
    for (int t = 0; t < J.T; ++t) {
        put(J.off_gt + 2 * t, gt);
        gt = fp_mulmod(gt, G, pd, pinv);
    }
    return bd == 0.0;
}

// Cubic-residue classes of the job's k for the primes of one wavefront's
// set-up (lane l holds prime p; valid lanes only).  For p = 1 (mod 3), the
// cubic character x^e, e = (p-1)/3, is 1, w or w^2 for a cube root of
// unity w; k b^n = 1 (mod p) needs k^e b^(n e) = 1, and n = r (mod 3) when
// 3 | Q.  With u = b^e: if u = 1 a row passes iff k^e = 1 (every r or
// none); else, taking w = u, k^e = u^c passes exactly the r = -c (mod 3).
// Each such prime below 2^50 gets a slot: its lane computes u and u^2 (in
// double precision, fp_mulmod) into the slot record at LDS byte address
// `area`; then the (slot, k) pairs are spread over the 64 lanes, each
// computing k^e and writing one class byte, the set of residues r mod 3 its
// rows may have (bit rho: r = rho), after the slot records.  pmax bounds the
// wavefront's primes (it fixes the exponent's bit length).  Returns the
// lane's slot, or -1 (no cubic test for its prime).
__device__ __forceinline__ int cubic_classes(const JobDev& J, u64 p, bool valid, u64 pmax,
                                             u32 area) {
    const int lane = threadIdx.x & 63;
    const bool c3 = valid && p < kFpMaxP && (hi32(p) % 3u + lo32(p) % 3u) % 3u == 1;
    const u64 cm = lanes(c3);
    if (cm == 0) return -1;
    const int nc = __popcll(cm);
    const int slot =
        (int)__builtin_amdgcn_mbcnt_hi(hi32(cm), __builtin_amdgcn_mbcnt_lo(lo32(cm), 0));
    const u64 emax = (min(pmax, kFpMaxP) - 1) / 3;
    const int top = 63 - __builtin_clzll(emax | 1);
    // x^e mod p, canonical, for x in [0, p) (left to right over the bits)
    auto powe = [&](double x, u64 e, double pd, double pinv) {
        double r = 1.0;
        for (int i = top; i >= 0; --i) {
            r = fp_mulmod(r, r, pd, pinv);
            if ((e >> i) & 1) r = fp_mulmod(r, x, pd, pinv);
        }
        return r < 0.0 ? r + pd : r;
    };
    // x mod p for a 32-bit x (p <= x only for p below 2^32)
    auto reduce = [](u32 x, u64 pp) { return pp > x ? (double)x : (double)(x % (u32)pp); };
    typedef __attribute__((address_space(3))) double lds_f64;
    typedef __attribute__((address_space(3))) u64 lds_u64;
    if (c3) {
        const u32 a = area + CUBIC_SLOT_BYTES * slot;
        const double pd = (double)p, pinv = 1.0 / pd;
        const u64 e = (p - 1) * 0xAAAAAAAAAAAAAAABull;  // (p - 1) / 3, exact
        const double u = powe(reduce(J.base, p), e, pd, pinv);
        double u2 = fp_mulmod(u, u, pd, pinv);
        u2 = u2 < 0.0 ? u2 + pd : u2;
        *(lds_f64*)(size_t)a = pd;
        *(lds_f64*)(size_t)(a + 8) = pinv;
        *(lds_u64*)(size_t)(a + 16) = e;
        *(lds_f64*)(size_t)(a + 24) = u;
        *(lds_f64*)(size_t)(a + 32) = u2;
    }
    wave_sync();
    const int nk = J.nk, total = nc * nk;  // total < 2^16 (nk <= CUBIC_MAX_K)
    for (int j = lane; j < total; j += 64) {
        const int s = nk > 1 ? div_T((u32)j, J.invnk) : j;
        const u32 a = area + CUBIC_SLOT_BYTES * s;
        const double pd = *(lds_f64*)(size_t)a, pinv = *(lds_f64*)(size_t)(a + 8);
        const u64 e = *(lds_u64*)(size_t)(a + 16);
        const double u = *(lds_f64*)(size_t)(a + 24), u2 = *(lds_f64*)(size_t)(a + 32);
        const double y = powe(reduce(global_ld(J.kval + (j - s * nk)), (u64)pd), e, pd, pinv);
        u32 m3 = 0;
        if (y == 1.0) m3 = u == 1.0 ? 7u : 1u;  // k a cube: r = 0 (mod 3), or any r
        else if (u != 1.0 && y == u) m3 = 4u;   // c = 1: r = 2
        else if (u != 1.0 && y == u2) m3 = 2u;  // c = 2: r = 1
        *(__attribute__((address_space(3))) unsigned char*)(size_t)(area + 64 * CUBIC_SLOT_BYTES +
                                                                    j) = (unsigned char)m3;
    }
    wave_sync();
    return c3 ? slot : -1;
}

// Set-up of one prime (one lane): writes the record described in kernels.h.
// cslot >= 0: the prime's slot in the cubic work area at LDS byte address
// carea (cubic_classes), whose class bytes narrow each k's rows.
__device__ __forceinline__ void prep_record(const JobDev& J, u64 p, u64* __restrict__ R,
                                            int cslot = -1, u32 carea = 0) {
    const u64 mp = mont_mprime(p);
    const bool bzero = p < kFpMaxP ? prep_powers_fp(J, p, R) : prep_powers_int(J, p, mp, R);
    // Jacobi bits; a prime dividing b gets none, so no row is searched
    const int jb = bzero ? 0
                         : jacobi_pre(J.bjac.x, J.bjac.y,
                                       ((u64)J.bjac.w << 32) | J.bjac.z, p);
    // row masks: k (in order) adds its even-r rows if (k/p) = w and its
    // odd-r rows if (kb/p) = w to words w0, w0 + 1, where w = +1, or (-1/p)
    // for k*b^n+1 (-k b^r must be a square); words are written once, in
    // order, when no later k can reach them
    const int want = J.plus && (lo32(p) & 3) == 3 ? -1 : 1;
    const int nw = (J.nrows + 63) >> 6;
    u64* const RM = R + J.off_j;
    int cur = 0;
    u64 lo = 0, hi = 0;  // words cur, cur + 1
    for (int i = 0; i < J.nk; ++i) {
        const int w0 = (int)const_ld(J.kw0 + i);
        if (w0 != cur) {
            RM[cur] = lo;
            if (w0 == cur + 1) {
                lo = hi;
            } else {
                RM[cur + 1] = hi;
                lo = 0;
            }
            hi = 0;
            cur = w0;
        }
        const uint4 kj = const_ld(J.kjac + i);
        const u32 tabw = const_ld(J.kjt + i);
        int jkv;
        if (jb == 0) {
            jkv = 0;
        } else if (tabw != 0xffffffffu && p > kj.x) {
            // (k/p) = (2/p)^z (n/p), (n/p) = (p mod n / n) (-1)^((n-1)/2 (p-1)/2);
            // p > n is prime, so p mod n is a unit and the table bit decides
            const u32 p8 = lo32(p) & 7;
            int t = 1;
            if ((kj.y & 1) && (p8 == 3 || p8 == 5)) t = -t;
            if ((kj.x & 3) == 3 && (p8 & 3) == 3) t = -t;
            const u64 rn = ((u64)kj.w << 32) | kj.z;
            const u64 q = umulhi64(p, rn);
            u64 r = p - mad32(lo32(q), kj.x, (u64)(hi32(q) * kj.x) << 32);
            if (r >= kj.x) r -= kj.x;
            if (r >= kj.x) r -= kj.x;
            const u32 x = (u32)r;
            jkv = (global_ld(J.jtab + tabw + (x >> 5)) >> (x & 31)) & 1 ? t : -t;
        } else if (const uint4 kf = const_ld(J.kfac + i); kf.w && p > kf.z) {
            // (k/p) = (2/p)^z prod (q/p) over the primes q of odd exponent in
            // n, (q/p) = (p mod q / q) (-1)^((q-1)/2 (p-1)/2) from q's table;
            // the primes without one come as their product, by the loop
            const u32 p8 = lo32(p) & 7;
            int t = (kj.y & 1) && (p8 == 3 || p8 == 5) ? -1 : 1;
            for (u32 f = 0; f < kf.y; ++f) {
                const uint4 e = const_ld(J.fac + kf.x + f);
                if (e.w == 0xffffffffu) {
                    t *= jacobi_pre(e.x, 0, ((u64)e.z << 32) | e.y, p);
                    continue;
                }
                if ((e.x & 3) == 3 && (p8 & 3) == 3) t = -t;
                const u64 q = umulhi64(p, ((u64)e.z << 32) | e.y);
                u64 r = p - mad32(lo32(q), e.x, (u64)(hi32(q) * e.x) << 32);
                if (r >= e.x) r -= e.x;
                if (r >= e.x) r -= e.x;
                const u32 x = (u32)r;
                if (!((global_ld(J.jtab + e.w + (x >> 5)) >> (x & 31)) & 1)) t = -t;
            }
            jkv = t;
        } else {
            jkv = jacobi_pre(kj.x, kj.y, ((u64)kj.w << 32) | kj.z, p);
        }
        const u64* const km = J.kmask + 4 * i;
        u64 qlo = 0, qhi = 0;
        if (jkv == want) qlo |= const_ld(km), qhi |= const_ld(km + 1);
        if (jkv * jb == want) qlo |= const_ld(km + 2), qhi |= const_ld(km + 3);
        if (cslot >= 0) {  // the cubic test: the rows of the residues r mod 3 it passes
            const u32 m3 = *(__attribute__((address_space(3))) const unsigned char*)(size_t)(
                carea + 64 * CUBIC_SLOT_BYTES + cslot * J.nk + i);
            const u64* const k3 = J.kmask3 + 6 * i;
            u64 clo = 0, chi = 0;
            if (m3 & 1) clo |= const_ld(k3), chi |= const_ld(k3 + 1);
            if (m3 & 2) clo |= const_ld(k3 + 2), chi |= const_ld(k3 + 3);
            if (m3 & 4) clo |= const_ld(k3 + 4), chi |= const_ld(k3 + 5);
            qlo &= clo;
            qhi &= chi;
        }
        lo |= qlo;
        hi |= qhi;
    }
    RM[cur] = lo;
    if (cur + 1 < nw) RM[cur + 1] = hi;
    R[REC_P] = p;
    R[REC_MP] = mp;
}

// BSGS of one prime by the whole wavefront.  `pre` holds record words
// 0..63 (lane l has word l), the rest are read from Rg.  The table is at
// tab, LDS byte address tab_addr, a multiple of its size; the rest of the
// wavefront's LDS (counts, Y, row list, record) at wbase.  With idx (only
// with rows_in_lds) the row list holds row numbers into rows_lds.
template <int BB, bool PF>
__device__ __forceinline__ void bsgs_prime(const JobDev& J, const JobDev* __restrict__ jobs,
                                           int job, u64 pre,
                                           const u64* __restrict__ Rg, uint4* __restrict__ tab,
                                           u32 tab_addr, u32* __restrict__ wbase,
                                           const uint2* __restrict__ rows_lds,
                                           bool rows_in_lds, bool idx,
                                           HitDev* __restrict__ hits,
                                           u32* __restrict__ hit_count, u32 hit_cap) {
    const int lane = threadIdx.x & 63;
    const u32 NB = 1u << J.logNB;
    const u32 nbm = NB - 1;
    const int Q = J.Q, T = J.T, m = J.m, W = J.W;
    u32* const cnt = wbase;
    u64* const Y = (u64*)(cnt + NB);
    uint2* const list = (uint2*)(Y + Q * J.ys);
    u64* const R = (u64*)((u32*)list + list_words(J.nrows, idx));
    const int L = m >> 6;

    if (lane < W) R[lane] = pre;
    for (int w = 64 + lane; w < W; w += 64) R[w] = Rg[w];
    const u64 p = readlane64(pre, REC_P), mp = readlane64(pre, REC_MP);
    const u64 gw = readlane64(pre, REC_GW), gws = readlane64(pre, REC_GWS);

    // clear the fill counts (the slots are not cleared: see bucket_of)
    for (u32 c = 4 * lane; c < NB; c += 256) *(uint4*)(cnt + c) = make_uint4(0, 0, 0, 0);
    wave_sync();

    // p < 2^50: products in double precision (fp_mulmod: about a third of
    // the integer Shoup step's instructions); values are made canonical and
    // their low words read from the bits of value + 2^52.  The record holds
    // the constants as doubles.
    const bool fpm = p < kFpMaxP;
    auto rec_d = [&](int w) { return __longlong_as_double((long long)readlane64(pre, w)); };
    auto lds_d = [](u64 v) { return __longlong_as_double((long long)v); };
    const double pd = fpm ? rec_d(REC_PD) : 0.0, pinv = fpm ? rec_d(REC_PINV) : 0.0;

    // baby steps: lane l inserts g^j, j = l + 64 s, as a fingerprint in the
    // first free slot of its bucket (or of the next non-full bucket); it
    // starts from g^(l & 7) g^(8 (l >> 3))
    u64 vstart;
    double vd = 0.0;
    if (fpm) {
        vd = fp_mulmod(lds_d(R[REC_GP + (lane & 7)]), lds_d(R[REC_G8P + (lane >> 3)]), pd, pinv);
        vd = vd < 0.0 ? vd + pd : vd;
        vstart = fp_bits(vd) & ((1ull << 52) - 1);
    } else {
        vstart = montmul(R[REC_GP + (lane & 7)], R[REC_G8P + (lane >> 3)], p, mp);
    }
    u64 v = vstart;
    bool ovf = false;  // some insert went past a full bucket
    // the first free slot at or after bucket b, given the slot count c
    // that b's fill counter returned: rare, walk to the first bucket with a
    // free slot; returns the slot's LDS byte address
    auto place = [&](u32 b, u32 ta, u32 c) {
        if (c >= 8) {
            ovf = true;
            do {
                b = (b + 1) & nbm;
                c = atomicAdd(&cnt[b], 1u);
            } while (c >= 8);
            ta = tab_addr + 16 * b;
        }
        return ta + 2 * c;
    };
    auto store_fp = [](u32 at, u32 lo) {
        *(__attribute__((address_space(3))) unsigned short*)(size_t)at = fingerprint(lo);
    };
    auto insert = [&](u32 lo) {
        const u32 b = bucket_of(lo, (u32)J.logNB);
        store_fp(place(b, tab_addr + 16 * b, atomicAdd(&cnt[b], 1u)), lo);
    };
    // giant steps leave products in [0, 2p): an unreduced one is s + p with
    // s < p*kmax/2^32 <= bound, so such baby values go in twice (the second
    // entry's low word is lo(value) + lo(p))
    const u64 bound = (p >> 32) * J.kmax + J.kmax + 1;
    const u32 plo = lo32(p);
    // steps in blocks of BB: the block's values, then its fill-counter
    // atomics (all in flight together), then its fingerprints; step(lo, dup)
    // yields the next value's low word and whether it is below the bound,
    // and advances the chain (one product past the last step, cheaper than
    // selecting it away)
    auto baby = [&](auto step) {
        int s0 = 0;
        for (; s0 + BB <= L; s0 += BB) {
            u32 lo[BB], ta[BB], c[BB], bk[BB];
            bool dup[BB];
#pragma unroll
            for (int u = 0; u < BB; ++u) step(lo[u], dup[u]);
#pragma unroll
            for (int u = 0; u < BB; ++u) {
                bk[u] = bucket_of(lo[u], (u32)J.logNB);
                // the bucket's LDS byte address, kept as one value (the empty
                // asm stops the slot offset being folded back into the bucket
                // index, one more add)
                ta[u] = tab_addr + 16 * bk[u];
                asm volatile("" : "+v"(ta[u]));
                c[u] = atomicAdd(&cnt[bk[u]], 1u);
            }
#pragma unroll
            for (int u = 0; u < BB; ++u) store_fp(place(bk[u], ta[u], c[u]), lo[u]);
#pragma unroll
            for (int u = 0; u < BB; ++u)
                if (dup[u]) insert(lo[u] + plo);
        }
        for (; s0 < L; ++s0) {
            u32 lo;
            bool dup;
            step(lo, dup);
            insert(lo);
            if (dup) insert(lo + plo);
        }
    };
    if (fpm) {
        const double wd = rec_d(REC_GW), bd = rec_d(REC_GWS);  // g^64 and the bound
        baby([&](u32& lo, bool& dup) {
            const double c = vd < 0.0 ? vd + pd : vd;
            lo = lo32(fp_bits(c));
            dup = c < bd;
            vd = fp_mulmod(vd, wd, pd, pinv);
        });
    } else {
        baby([&](u32& lo, bool& dup) {
            lo = lo32(v);
            dup = v < bound;
            v = shoup_mul(v, gw, gws, p);
        });
    }
    const bool chain = __ballot(ovf) != 0;

    // Y[r*ys + t]: the Shoup entry of b^r H G^t, made from its form times
    // 2^32, which is BH[r] times the plain G^t by Shoup's method
    const u32 mp0 = lo32(mp);
    if (fpm) {  // BH and G^t are doubles in the record
        auto yent = [&](int idx) {
            const int r = div_T((u32)idx, J.invT), t = idx - r * T;
            double zd = fp_mulmod(__longlong_as_double((long long)R[REC_BH + r]),
                                  __longlong_as_double((long long)R[J.off_gt + 2 * t]), pd, pinv);
            zd = zd < 0.0 ? zd + pd : zd;
            const u64 z = fp_bits(zd) & ((1ull << 52) - 1);
            Y[r * J.ys + t] = shoup32_entry(z, p, mp0);
        };
        if (Q * T <= 192) {  // at most three rounds, unrolled
#pragma unroll
            for (int j = 0; j < 3; ++j)
                if (64 * j < Q * T && 64 * j + lane < Q * T) yent(64 * j + lane);
        } else {
            for (int idx = lane; idx < Q * T; idx += 64) yent(idx);
        }
    } else {
        for (int idx = lane; idx < Q * T; idx += 64) {
            const int r = div_T((u32)idx, J.invT), t = idx - r * T;
            const u64 z =
                shoup_mul(R[REC_BH + r], R[J.off_gt + 2 * t], R[J.off_gt + 2 * t + 1], p);
            Y[r * J.ys + t] = shoup32_entry(z, p, mp0);
        }
    }

    // rows in slices of at most ROW_SLICE: those passing the Jacobi filter
    // are compacted in order as (k, row << 8 | r), then their giant steps run
    for (int r0 = 0; r0 < J.nrows; r0 += ROW_SLICE) {
        const int r1 = min(J.nrows, r0 + ROW_SLICE);
        const int A = rows_in_lds ? compact_rows<true>(J, rows_lds, idx, R, list, r0, r1, T)
                                  : compact_rows<false>(J, rows_lds, false, R, list, r0, r1, T);
        wave_sync();
        const char* const Yb = (const char*)Y;
#define RS_GIANT(TWV, CH, EX)                                                                \
    giant_steps<TWV, CH, EX, PF>(J, jobs, job, A, r0, tab_addr, cnt, Yb, list, rows_lds, idx,       \
                             R, 0u - lo32(p), vstart, hits, hit_count, hit_cap)
#define RS_GIANT_TW(CH, EX)      \
    if (J.tw >= 4)               \
        RS_GIANT(4, CH, EX);     \
    else if (J.tw == 2)          \
        RS_GIANT(2, CH, EX);     \
    else                         \
        RS_GIANT(1, CH, EX);
        // rounds of 64*TW items cover whole rows when T/TW divides 64
        // (runs then keep their step t0); chain mode is rare and general
        if (chain) {
            RS_GIANT_TW(true, false)
        } else if ((64 * J.tw) % T == 0) {
            RS_GIANT_TW(false, true)
        } else {
            RS_GIANT_TW(false, false)
        }
#undef RS_GIANT_TW
#undef RS_GIANT
        wave_sync();  // the next slice rewrites the list
    }
    wave_sync();  // the next prime reuses the table
}

// The persistent kernel.  IDX: single-job launches whose wavefronts list
// active rows by row number (RS_TILE_ROW_INDEX), a smaller LDS layout.  BB:
// baby steps per block of fill-counter atomics in flight.
template <bool IDX, int BB, bool PF>
__device__ __forceinline__ void sieve_body(
    ChunkDev c0, const JobDev* __restrict__ jobs, const ChunkDev* __restrict__ chunks,
    int nchunks, long long ngroups, int lds_words, int tab_words, int shared_words,
    int scratch_words, uint64_t* __restrict__ scratch, unsigned long long* __restrict__ work,
    HitDev* __restrict__ hits, uint32_t* __restrict__ hit_count, uint32_t hit_cap) {
    extern __shared__ uint32_t lds[];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wpb = blockDim.x >> 6;
    // block LDS: the wavefronts' tables (tab_words each, a power of two at
    // least the table of every job of the launch, so each is aligned to its
    // size), then the rest of each wavefront's words (lds_words - tab_words),
    // then the block-shared part: single-job launches whose rows fit one
    // slice keep the job's active-row entries (JobDev.rowent) there
    // (shared_words >= 2 * nrows), read by every prime
    uint2* const rows_lds = (uint2*)(lds + wpb * lds_words);
    // (shared_words > 0 only for single-chunk launches, whose job is c0.job)
    const bool rows_in_lds = shared_words > 0;
    if (rows_in_lds) {
        const JobDev J0 = uniform_load(jobs + c0.job);
        for (int i = threadIdx.x; i < J0.nrows; i += blockDim.x)
            rows_lds[i] = global_ld(J0.rowent + i);
        __syncthreads();
    }
    // the kernel has no static LDS (checked by the host), so the dynamic
    // region starts at LDS address 0 and each table is aligned to its slot
    uint4* const tab = (uint4*)(lds + wv * tab_words);
    const u32 tab_addr = (u32)(wv * tab_words) * 4u;
    u32* const wbase = lds + wpb * tab_words + wv * (lds_words - tab_words);
    u64* const S = scratch + ((long long)blockIdx.x * wpb + wv) * scratch_words;

    int ci = 0;                       // chunk of the previous group
    long long c_lo = 0, c_hi = -1;    // its group range [c_lo, c_hi)
    while (true) {
        unsigned long long grp = 0;
        if (lane == 0) grp = atomicAdd(work, 1ull);
        grp = ((unsigned long long)(u32)__builtin_amdgcn_readfirstlane((u32)(grp >> 32)) << 32) |
              (u32)__builtin_amdgcn_readfirstlane((u32)grp);
        if ((long long)grp >= ngroups) break;
        // the chunk holding this group (chunks sorted by group0); a wave's
        // next group is almost always in the same chunk, so the last one
        // found and its group range are kept
        if (chunks && ((long long)grp < c_lo || (long long)grp >= c_hi)) {
            int hi = nchunks - 1;
            ci = 0;
            while (ci < hi) {
                const int mid = (ci + hi + 1) >> 1;
                if (chunks[mid].group0 <= (long long)grp) ci = mid;
                else hi = mid - 1;
            }
            ci = __builtin_amdgcn_readfirstlane(ci);
            c_lo = chunks[ci].group0;
            c_hi = ci + 1 < nchunks ? chunks[ci + 1].group0 : ngroups;
            c_lo = ((long long)__builtin_amdgcn_readfirstlane((int)(c_lo >> 32)) << 32) |
                   (u32)__builtin_amdgcn_readfirstlane((u32)c_lo);
            c_hi = ((long long)__builtin_amdgcn_readfirstlane((int)(c_hi >> 32)) << 32) |
                   (u32)__builtin_amdgcn_readfirstlane((u32)c_hi);
        }
        const ChunkDev C = chunks ? uniform_load(chunks + ci) : c0;
        const JobDev J = uniform_load(jobs + C.job);
        const long long q0 = ((long long)grp - C.group0) * 64;
        const long long cnt =
            C.count_dev ? (long long)(u32)__builtin_amdgcn_readfirstlane(*C.count_dev) : C.count;
        const int n = (int)min(64ll, cnt - q0);
        // a device-counted chunk numbers its groups up to a bound and the
        // groups past its count are empty: past the last chunk the loop ends,
        // in a table of chunks the next group may belong to the next chunk
        if (n <= 0) {
            if (chunks) continue;
            break;
        }

        // set-up, one lane per prime (the cubic classes first, with the
        // wavefront's table slot as their work area; offsets ascend, so the
        // last prime is the largest)
        const u64 p = lane < n ? C.low + global_ld(C.offs + q0 + lane) : 0;
        const int cslot = J.cubic ? cubic_classes(J, p, lane < n, readlane64(p, n - 1), tab_addr)
                                  : -1;
        if (lane < n) prep_record(J, p, S + lane * J.W, cslot, tab_addr);
        wave_sync_global();

        // BSGS, the wavefront per prime; record words 0..63 of the next prime
        // are prefetched into one register per lane
        u64 pre = lane < J.W ? S[lane] : 0;
        for (int i = 0; i < n; ++i) {
            const u64 cur = pre;
            if (i + 1 < n) pre = lane < J.W ? S[(i + 1) * J.W + lane] : 0;
            bsgs_prime<BB, PF>(J, jobs, C.job, cur, S + i * J.W, tab, tab_addr, wbase, rows_lds, rows_in_lds,
                       IDX, hits, hit_count, hit_cap);
        }
        wave_sync_global();  // records are rewritten by the next group
    }
}

// ------------------------------------------------------------ shared tables
//
// Jobs whose best table is larger than one wavefront's share of LDS allows
// (many rows over a wide n window: the best m grows as sqrt(rows * X)) run
// with one table per workgroup: the NV wavefronts of a workgroup work on one
// prime together.  Each wavefront sets up 64 primes of the workgroup's group
// of 64 NV; then, prime after prime, the 64 NV lanes fill the shared table
// (lane t inserts g^(t + 64 NV s)), form the Y table, compact each row slice
// (wavefront w takes the slice's mask words w, w + NV, ...) and spread the
// giant steps over all lanes, with workgroup barriers between the phases.
// The baby steps also write the low words of their values, in order of j,
// to the workgroup's global scratch, so that a fingerprint match is settled
// by one wavefront reading m/64 words per lane instead of recomputing the
// m baby steps (with m in the thousands that rescan would cost about as
// much as the giant steps themselves).

// The shared-table record words at off_x (kernels.h): g^(64 w) for w < NV,
// then the stride g^(64 NV) with its Shoup constant.
template <int NV>
__device__ __forceinline__ void prep_group_powers(const JobDev& J, u64* __restrict__ R) {
    const u64 p = R[REC_P], mp = R[REC_MP];
    if (p < kFpMaxP) {
        const double pd = __longlong_as_double((long long)R[REC_PD]);
        const double pinv = __longlong_as_double((long long)R[REC_PINV]);
        const double g64 = __longlong_as_double((long long)R[REC_GW]);
        double x = 1.0;
        for (int w = 0; w < NV; ++w) {
            R[J.off_x + w] = (u64)__double_as_longlong(x);
            x = fp_mulmod(x, g64, pd, pinv);
        }
        R[J.off_x + NV] = (u64)__double_as_longlong(x);
        R[J.off_x + NV + 1] = 0;
    } else {
        const u64 g64 = R[REC_G64];  // Montgomery form
        u64 x = R[REC_GP];           // g^0: the Montgomery form of 1
        for (int w = 0; w < NV; ++w) {
            R[J.off_x + w] = x;
            x = montmul(x, g64, p, mp);
        }
        u64 w, ws;
        shoup_pair(x, p, mp, &w, &ws);
        R[J.off_x + NV] = w;
        R[J.off_x + NV + 1] = ws;
    }
}

// Emits every baby step j whose value equals the exact product of a giant
// step item (entry (k, yent) of row `row`, step tt, Y entry yv): the lanes
// of one wavefront compare the low words `low` (lane l reads j = l + 64 v)
// and confirm a low-word match exactly, by g^j from the record
// (square-and-multiply, in the form of the baby values).
__device__ __forceinline__ void rescan_group(const JobDev& J, const JobDev* __restrict__ jobs,
                                             int job, const u64* R, u64 yv, u32 k, u32 yent,
                                             int row, int tt, const u32* __restrict__ low,
                                             HitDev* __restrict__ hits,
                                             u32* __restrict__ hit_count, u32 hit_cap) {
    const int lane = threadIdx.x & 63;
    const u64 p = R[REC_P];
    const u64 target = shoup32_full(k, yv, p);
    const u32 tlo = lo32(target);
    const int L = J.m >> 6;
    // sixteen loads in flight per lane, then the (rare) matches
    for (int v = 0; v < L; v += 16) {
        u32 w[16];
#pragma unroll
        for (int u = 0; u < 16; ++u)
            w[u] = v + u < L ? global_ld(low + lane + 64 * (v + u)) : ~tlo;
        u32 found = 0;
#pragma unroll
        for (int u = 0; u < 16; ++u) found |= (u32)(w[u] == tlo) << u;
        while (found) {
            const int u = __builtin_ctz(found);
            found &= found - 1;
            const int j = lane + 64 * (v + u);
            u64 gj;
            if (p < kFpMaxP) {
                const double pd = __longlong_as_double((long long)R[REC_PD]);
                const double pinv = __longlong_as_double((long long)R[REC_PINV]);
                double d = fp_pow(__longlong_as_double((long long)R[REC_GP + 1]), (u32)j, pd, pinv);
                d = d < 0.0 ? d + pd : d;
                gj = fp_bits(d) & ((1ull << 52) - 1);
            } else {
                gj = mont_pow(R[REC_GP + 1], (u32)j, R[REC_GP], p, R[REC_MP]);
            }
            if (gj != target) continue;
            const u32 r = (yent >> 3) / (u32)J.ys;
            const JobDev* const jp = jobs + job;
            const int x0 = const_ld(&jp->x0);
            const long long xx = (long long)x0 + (long long)tt * J.m + (J.m - 1 - j);
            const long long n = (long long)J.Q * xx + (int)r;
            bool keep = n >= const_ld(&jp->minN) && n <= const_ld(&jp->maxN);
            const uint32_t* const bitmap = const_ld(&jp->bitmap);
            if (keep && bitmap) {
                const long long xp = xx - x0;
                keep = (global_ld(bitmap + (long long)row * const_ld(&jp->words) + (xp >> 5)) >>
                        (31 - (xp & 31))) & 1;
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
    }
}

// Giant steps of one prime over the lanes of the whole workgroup: runs of TW
// consecutive t of one row as in giant_steps, but each lane keeps its run's
// t for the whole prime: a round covers RR = floor(64 NV TW / T) whole rows
// (the lanes past RR T / TW rows' worth of runs idle, less than T items of
// 64 NV TW) and every lane moves RR rows on.  The table is at LDS address 0.
// The active rows of the slice starting at row r0 are uint16 numbers
// (row - r0) in list16, in the workgroup's global scratch; each lane reads
// its row number two rounds ahead and its row's entry from the job's row
// table (a few KB, cache-resident) one round ahead.
template <int NV, int TW, bool CHAIN>
__device__ __forceinline__ void giant_steps_group(const JobDev& J, const JobDev* __restrict__ jobs,
                                                  int job, int A, int r0,
                                                  const unsigned short* __restrict__ cnt,
                                                  const char* __restrict__ Yb,
                                                  const uint16_t* list16,
                                                  const u64* R, u32 negp0,
                                                  const u32* __restrict__ low,
                                                  HitDev* __restrict__ hits,
                                                  u32* __restrict__ hit_count, u32 hit_cap) {
    const int lid = threadIdx.x;  // lane of the workgroup
    const u32 nbm = (1u << J.logNB) - 1;
    u32 bmask = nbm << 4;
    asm volatile("" : "+v"(bmask));
    const int T = J.T;
    const int RR = div_T(64u * NV * TW, J.invT);  // rows per round (T <= 64 NV TW)
    int a = div_T((u32)(TW * lid), J.invT);       // >= RR: an idle lane
    const int tcur = TW * lid - a * T;
    const bool on = a < RR;
    if (!on) a = A;
    const uint2* const rowent = J.rowent + r0;
    // (the loads past the last round re-read the last row: unconditional
    // loads keep the loop free of branches, so the wait for each load is the
    // round after its issue, not the next merge point)
    int rnum = global_ld(list16 + min(a, A - 1));
    int rnum1 = global_ld(list16 + min(a + RR, A - 1));
    uint2 rown = global_ld(rowent + rnum);
    for (int base = 0; base < A; base += RR) {
        const bool ok = a < A;
        const u64 okm = lanes(ok);
        const uint2 row = rown;
        const int rcur = rnum;
        // the row number two rounds ahead, the row entry one round ahead
        const int rnum2 = global_ld(list16 + min(a + 2 * RR, A - 1));
        rown = global_ld(rowent + rnum1);
        rnum = rnum1;
        rnum1 = rnum2;
        if (on) a += RR;
        const u64* const yp = (const u64*)(Yb + row.y + 8 * tcur);
        uint4 e[TW];
        u32 lw[TW], ba[TW];
#pragma unroll
        for (int u = 0; u < TW; ++u) {
            const u32 lo = shoup32_lo(row.x, lo32(yp[u]), hi32(yp[u]), negp0);
            ba[u] = lo & bmask;  // the bucket's LDS byte address
            lw[u] = lo;
        }
#pragma unroll
        for (int u = 0; u < TW; ++u) e[u] = lds_bucket(ba[u]);
        // one fingerprint test for the TW items together (a match anywhere is
        // rare): the packed minimum over all their slots has a zero half; the
        // items are told apart below, only when one matched
        u64 mm[TW], full[TW];
        u64 anyf = 0;
        u16x2 acc = bucket_min(e[0], lw[0]);
#pragma unroll
        for (int u = 1; u < TW; ++u) acc = __builtin_elementwise_min(acc, bucket_min(e[u], lw[u]));
        const u64 anym = lanes(has_zero_half(acc)) & okm;
#pragma unroll
        for (int u = 0; u < TW; ++u) {
            mm[u] = 0;
            if (CHAIN) {
                full[u] = lanes(cnt[ba[u] >> 4] >= 8) & okm;
                anyf |= full[u];
            }
        }
        // rare: keep only the matches in filled slots
        if (anym) {
#pragma unroll
            for (int u = 0; u < TW; ++u) {
                mm[u] = lanes(bucket_match(e[u], lw[u])) & okm;
                u32 c = 0;
                if (in_mask(mm[u])) c = cnt[ba[u] >> 4];
                mm[u] &= lanes(bucket_match_live(e[u], lw[u], c));
            }
        }
        if (CHAIN && anyf) {
#pragma unroll
            for (int u = 0; u < TW; ++u) {
                bool hit = false;
                if (in_mask(full[u])) {
                    u32 b = ba[u] >> 4;
                    while (true) {
                        b = (b + 1) & nbm;
                        const uint4 f = lds_bucket(16 * b);
                        const u32 cb = cnt[b];
                        if (bucket_match(f, lw[u]) && bucket_match_live(f, lw[u], cb)) hit = true;
                        if (cb < 8) break;
                    }
                }
                mm[u] |= lanes(hit);
            }
        }
#pragma unroll
        for (int u = 0; u < TW; ++u) {
            u64 mb = mm[u];
            while (mb) {
                const int src = __builtin_ctzll(mb);
                mb &= mb - 1;
                const u32 k = __builtin_amdgcn_readlane(row.x, src);
                const u32 yent = __builtin_amdgcn_readlane(row.y, src);
                const int tt = __builtin_amdgcn_readlane(tcur, src) + u;
                const int rw = r0 + __builtin_amdgcn_readlane(rcur, src);
                const u64 yv = *(const u64*)(Yb + yent + 8 * tt);
                rescan_group(J, jobs, job, R, yv, k, yent, rw, tt, low, hits, hit_count,
                             hit_cap);
            }
        }
    }
}

// BSGS of one prime by the NV wavefronts of the workgroup.  Rg: the
// prime's record in global scratch (word tid already in pre); the table at
// LDS address 0, the rest of the layout (lds_words_group) from wbase; low:
// the workgroup's m low words and list16 its row list, both in its global
// scratch.
template <int NV, int BB>
__device__ __forceinline__ void bsgs_prime_group(const JobDev& J, const JobDev* __restrict__ jobs,
                                                 int job, const u64* __restrict__ Rg,
                                                 uint4* __restrict__ tab, u32* __restrict__ wbase,
                                                 u32* __restrict__ low, uint16_t* list16,
                                                 u64 pre,
                                                 HitDev* __restrict__ hits,
                                                 u32* __restrict__ hit_count, u32 hit_cap) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const u32 NB = 1u << J.logNB;
    const u32 nbm = NB - 1;
    const int Q = J.Q, T = J.T, m = J.m, W = J.W;
    // bucket fill counts as 16-bit halves of 32-bit words (at most 2m <=
    // 32768 inserts reach one bucket, so a half never carries into the other)
    u32* const cnt = wbase;
    const unsigned short* const cnt16 = (const unsigned short*)cnt;
    auto cnt_add = [&](u32 b) -> u32 {
        const u32 sh = (b & 1) << 4;
        return (atomicAdd(&cnt[b >> 1], 1u << sh) >> sh) & 0xffffu;
    };
    u64* const Y = (u64*)(cnt + NB / 2);
    u64* const R = Y + Q * J.ys;
    u32* const misc = (u32*)(R + W);  // [0]: some insert overflowed its bucket

    // the record: word tid came in `pre`, loaded one prime ahead
    if (tid < W) R[tid] = pre;
    for (int w = tid + 64 * NV; w < W; w += 64 * NV) R[w] = Rg[w];
    for (u32 c = tid; c < NB / 2; c += 64 * NV) cnt[c] = 0;  // the slots are not cleared
    if (tid == 0) misc[0] = 0;
    __syncthreads();
    const u64 p = R[REC_P], mp = R[REC_MP];
    const bool fpm = p < kFpMaxP;
    auto lds_d = [](u64 v) { return __longlong_as_double((long long)v); };
    const double pd = fpm ? lds_d(R[REC_PD]) : 0.0, pinv = fpm ? lds_d(R[REC_PINV]) : 0.0;

    // baby steps: lane t inserts g^j, j = t + 64 NV s, starting from
    // g^(t & 7) g^(8 ((t >> 3) & 7)) g^(64 w)
    const int L = m / (64 * NV);
    double vd = 0.0;
    u64 v = 0;
    if (fpm) {
        vd = fp_mulmod(fp_mulmod(lds_d(R[REC_GP + (lane & 7)]), lds_d(R[REC_G8P + (lane >> 3)]),
                                 pd, pinv),
                       lds_d(R[J.off_x + wv]), pd, pinv);
    } else {
        v = montmul(montmul(R[REC_GP + (lane & 7)], R[REC_G8P + (lane >> 3)], p, mp),
                    R[J.off_x + wv], p, mp);
    }
    bool ovf = false;
    auto place = [&](u32 b, u32 ta, u32 c) {
        if (c >= 8) {
            ovf = true;
            do {
                b = (b + 1) & nbm;
                c = cnt_add(b);
            } while (c >= 8);
            ta = 16 * b;
        }
        return ta + 2 * c;
    };
    auto store_fp = [](u32 at, u32 lo) {
        *(__attribute__((address_space(3))) unsigned short*)(size_t)at = fingerprint(lo);
    };
    auto insert = [&](u32 lo) {
        const u32 b = bucket_of(lo, (u32)J.logNB);
        store_fp(place(b, 16 * b, cnt_add(b)), lo);
    };
    const u64 bound = (p >> 32) * J.kmax + J.kmax + 1;
    const u32 plo = lo32(p);
    auto baby = [&](auto step) {
        int s0 = 0;
        for (; s0 + BB <= L; s0 += BB) {
            u32 lo[BB], ta[BB], c[BB], bk[BB];
            bool dup[BB];
#pragma unroll
            for (int u = 0; u < BB; ++u) {
                step(lo[u], dup[u]);
                low[tid + 64 * NV * (s0 + u)] = lo[u];
            }
#pragma unroll
            for (int u = 0; u < BB; ++u) {
                bk[u] = bucket_of(lo[u], (u32)J.logNB);
                ta[u] = 16 * bk[u];
                asm volatile("" : "+v"(ta[u]));
                c[u] = cnt_add(bk[u]);
            }
#pragma unroll
            for (int u = 0; u < BB; ++u) store_fp(place(bk[u], ta[u], c[u]), lo[u]);
#pragma unroll
            for (int u = 0; u < BB; ++u)
                if (dup[u]) insert(lo[u] + plo);
        }
        for (; s0 < L; ++s0) {
            u32 lo;
            bool dup;
            step(lo, dup);
            low[tid + 64 * NV * s0] = lo;
            insert(lo);
            if (dup) insert(lo + plo);
        }
    };
    if (fpm) {
        const double wd = lds_d(R[J.off_x + NV]), bd = lds_d(R[REC_GWS]);
        baby([&](u32& lo, bool& dup) {
            const double c = vd < 0.0 ? vd + pd : vd;
            lo = lo32(fp_bits(c));
            dup = c < bd;
            vd = fp_mulmod(vd, wd, pd, pinv);
        });
    } else {
        const u64 gw = R[J.off_x + NV], gws = R[J.off_x + NV + 1];
        baby([&](u32& lo, bool& dup) {
            lo = lo32(v);
            dup = v < bound;
            v = shoup_mul(v, gw, gws, p);
        });
    }
    if (__ballot(ovf) != 0 && lane == 0) atomicOr(&misc[0], 1u);

    // Y table over all lanes of the workgroup (as in bsgs_prime)
    const u32 mp0 = lo32(mp);
    for (int idx = tid; idx < Q * T; idx += 64 * NV) {
        const int r = div_T((u32)idx, J.invT), t = idx - r * T;
        u64 z;
        if (fpm) {
            double zd = fp_mulmod(lds_d(R[REC_BH + r]), lds_d(R[J.off_gt + 2 * t]), pd, pinv);
            zd = zd < 0.0 ? zd + pd : zd;
            z = fp_bits(zd) & ((1ull << 52) - 1);
        } else {
            z = shoup_mul(R[REC_BH + r], R[J.off_gt + 2 * t], R[J.off_gt + 2 * t + 1], p);
        }
        Y[r * J.ys + t] = shoup32_entry(z, p, mp0);
    }
    __syncthreads();  // table, low words and Y complete
    const bool chain = misc[0] != 0;

    for (int r0 = 0; r0 < J.nrows; r0 += GROUP_SLICE) {
        const int r1 = min(J.nrows, r0 + GROUP_SLICE);
        const int nw = (r1 - r0 + 63) >> 6;
        // mask words of the slice; wavefront w compacts words w, w + NV, ...
        // after the rows of the words before them, as row numbers from r0
        int A = 0;
        for (int w = 0; w < nw; ++w) {
            const u64 mask = R[J.off_j + (r0 >> 6) + w];
            if ((w % NV) == wv) {
                if (in_mask(mask)) {
                    const u32 below = __builtin_amdgcn_mbcnt_hi(
                        hi32(mask), __builtin_amdgcn_mbcnt_lo(lo32(mask), 0));
                    *(__attribute__((address_space(1))) uint16_t*)(list16 + A + below) =
                        (uint16_t)(64 * w + lane);
                }
            }
            A += __popcll(mask);
        }
        __syncthreads();
        if (A == 0) continue;  // no barrier needed: the list was not read
        const char* const Yb = (const char*)Y;
#define RS_GIANT_G(TWV, CH)                                                                 \
    giant_steps_group<NV, TWV, CH>(J, jobs, job, A, r0, cnt16, Yb, list16, R, 0u - lo32(p), \
                                   low, hits, hit_count, hit_cap)
#define RS_GIANT_G_TW(CH)  \
    if (J.tw >= 4)         \
        RS_GIANT_G(4, CH); \
    else if (J.tw == 2)    \
        RS_GIANT_G(2, CH); \
    else                   \
        RS_GIANT_G(1, CH);
        if (chain) {
            RS_GIANT_G_TW(true)
        } else {
            RS_GIANT_G_TW(false)
        }
#undef RS_GIANT_G_TW
#undef RS_GIANT_G
        __syncthreads();  // the next slice rewrites the list, the next prime the table
    }
}

// The persistent shared-table kernel body: single-chunk launches (c0).
// Each workgroup takes groups of 64 NV consecutive primes from the work
// counter; its scratch holds their records and the low words.
template <int NV, int BB>
__device__ __forceinline__ void sieve_body_group(
    ChunkDev c0, const JobDev* __restrict__ jobs, long long ngroups, int lds_words,
    int scratch_words, uint64_t* __restrict__ scratch, unsigned long long* __restrict__ work,
    HitDev* __restrict__ hits, uint32_t* __restrict__ hit_count, uint32_t hit_cap) {
    extern __shared__ uint32_t lds[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const JobDev J = uniform_load(jobs + c0.job);
    uint4* const tab = (uint4*)lds;
    u32* const wbase = lds + 4 * (1 << J.logNB);
    u32* const gmisc = lds + lds_words - 4;  // the group number (WG_MISC words at the end)
    u64* const S = scratch + (long long)blockIdx.x * scratch_words;
    u32* const low = (u32*)(S + 64ll * NV * J.W);
    uint16_t* const list16 = (uint16_t*)(S + 64ll * NV * J.W + (J.m + 1) / 2);
    while (true) {
        if (tid == 0) {
            const unsigned long long g = atomicAdd(work, 1ull);
            gmisc[0] = (u32)g;
            gmisc[1] = (u32)(g >> 32);
        }
        __syncthreads();
        const long long grp =
            ((long long)(u32)__builtin_amdgcn_readfirstlane(gmisc[1]) << 32) |
            (u32)__builtin_amdgcn_readfirstlane(gmisc[0]);
        __syncthreads();  // every lane has read it before it is rewritten
        if (grp >= ngroups) break;
        const long long q0 = grp * 64 * NV;
        const long long cnt =
            c0.count_dev ? (long long)(u32)__builtin_amdgcn_readfirstlane(*c0.count_dev)
                         : c0.count;
        const int n = (int)min((long long)(64 * NV), cnt - q0);
        if (n <= 0) break;  // device-counted chunk: groups past the end are empty
        // set-up, one lane per prime: wavefront w takes primes 64 w ..
        const int i = 64 * wv + lane;
        if (i < n) {
            u64* const R = S + (long long)i * J.W;
            prep_record(J, c0.low + global_ld(c0.offs + q0 + i), R);
            prep_group_powers<NV>(J, R);
        }
        __syncthreads();
        u64 pre = tid < J.W ? S[tid] : 0;
        for (int k = 0; k < n; ++k) {
            const u64 cur = pre;
            if (k + 1 < n) pre = tid < J.W ? S[(long long)(k + 1) * J.W + tid] : 0;
            bsgs_prime_group<NV, BB>(J, jobs, c0.job, S + (long long)k * J.W, tab, wbase, low,
                                     list16, cur, hits, hit_count, hit_cap);
        }
    }
}

#ifndef RS_BB4
#define RS_BB4 4
#endif
#ifndef RS_BB5
#define RS_BB5 2
#endif
#ifndef RS_RIX_PF
#define RS_RIX_PF true
#endif
// baby-step blocks of the 128- and the 96-VGPR builds
constexpr int BB4 = RS_BB4, BB5 = RS_BB5;

// At least 4 wavefronts per SIMD (128 VGPRs): full row entries, each giant-
// step round's entry read a round ahead (the registers for it fit this
// build's budget, not the 96 of the 5-wavefront builds); the LDS layouts of
// typical jobs allow 4, and fewer leave the VALU idle on LDS waits.  The
// host picks this or rs_sieve_kernel5, whichever keeps more wavefronts
// resident for the launch's layout.
extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
rs_sieve_kernel(ChunkDev c0, const JobDev* __restrict__ jobs,
                const ChunkDev* __restrict__ chunks, int nchunks, long long ngroups,
                int lds_words, int tab_words, int shared_words, int scratch_words,
                uint64_t* __restrict__ scratch, unsigned long long* __restrict__ work,
                HitDev* __restrict__ hits, uint32_t* __restrict__ hit_count, uint32_t hit_cap) {
    sieve_body<false, BB4, true>(c0, jobs, chunks, nchunks, ngroups, lds_words, tab_words,
                                 shared_words, scratch_words, scratch, work, hits, hit_count,
                                 hit_cap);
}


// At least 5 wavefronts per SIMD (96 VGPRs), for layouts small enough to
// keep 5 resident: full row entries, and row-number lists (single-job
// launches with RS_TILE_ROW_INDEX only).
extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
rs_sieve_kernel5(ChunkDev c0, const JobDev* __restrict__ jobs,
                 const ChunkDev* __restrict__ chunks, int nchunks, long long ngroups,
                 int lds_words, int tab_words, int shared_words, int scratch_words,
                 uint64_t* __restrict__ scratch, unsigned long long* __restrict__ work,
                 HitDev* __restrict__ hits, uint32_t* __restrict__ hit_count, uint32_t hit_cap) {
    sieve_body<false, BB5, false>(c0, jobs, chunks, nchunks, ngroups, lds_words, tab_words,
                                  shared_words, scratch_words, scratch, work, hits, hit_count,
                                  hit_cap);
}

extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
rs_sieve_kernel5_rix(ChunkDev c0, const JobDev* __restrict__ jobs,
                     const ChunkDev* __restrict__ chunks, int nchunks, long long ngroups,
                     int lds_words, int tab_words, int shared_words, int scratch_words,
                     uint64_t* __restrict__ scratch, unsigned long long* __restrict__ work,
                     HitDev* __restrict__ hits, uint32_t* __restrict__ hit_count,
                     uint32_t hit_cap) {
    sieve_body<true, BB5, RS_RIX_PF>(c0, jobs, chunks, nchunks, ngroups, lds_words, tab_words,
                                 shared_words, scratch_words, scratch, work, hits, hit_count,
                                 hit_cap);
}

// One workgroup of 4 or 8 wavefronts per prime, one shared table
// (RS_TILE_GROUP4 / RS_TILE_GROUP8), at least 4 wavefronts per SIMD.
#define RS_GROUP_KERNEL(NAME, NV)                                                            \
    extern "C" __global__ void __launch_bounds__(64 * NV)                                   \
        __attribute__((amdgpu_waves_per_eu(4)))                                             \
        NAME(ChunkDev c0, const JobDev* __restrict__ jobs, const ChunkDev* __restrict__ chunks, \
             int nchunks, long long ngroups, int lds_words, int tab_words, int shared_words,   \
             int scratch_words, uint64_t* __restrict__ scratch,                                \
             unsigned long long* __restrict__ work, HitDev* __restrict__ hits,                 \
             uint32_t* __restrict__ hit_count, uint32_t hit_cap) {                             \
        (void)chunks;                                                                       \
        (void)nchunks;                                                                      \
        (void)tab_words;                                                                    \
        (void)shared_words;                                                                 \
        sieve_body_group<NV, BB4>(c0, jobs, ngroups, lds_words, scratch_words, scratch, work, \
                                  hits, hit_count, hit_cap);                                \
    }
RS_GROUP_KERNEL(rs_sieve_group4_kernel, 4)
RS_GROUP_KERNEL(rs_sieve_group8_kernel, 8)
#undef RS_GROUP_KERNEL

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

// The set-up phase's row masks for primes p[i] (one lane per prime, the
// sieve's own prep_record on the job's constants): out[i * NW + w] = mask
// word w of p[i]'s record, NW = ceil(rows / 64); rec holds W words of
// scratch per prime.
extern "C" __global__ void rs_test_row_masks(const JobDev* __restrict__ jobs,
                                             const uint64_t* __restrict__ p,
                                             uint64_t* __restrict__ rec,
                                             uint64_t* __restrict__ out, long long n) {
    // one wavefront per block; with J.cubic its cubic work area is the
    // block's dynamic LDS (cubic_area_bytes(nk) bytes, at address 0)
    extern __shared__ uint32_t lds[];
    const JobDev J = uniform_load(jobs);  // before any lane leaves
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const u64 pi = i < n ? p[i] : 0;
    // primes in any order: the exponent bound of the largest below 2^50
    const int cslot = J.cubic ? cubic_classes(J, pi, i < n, kFpMaxP - 1, lds_addr(lds)) : -1;
    if (i >= n) return;
    u64* const R = rec + i * J.W;
    prep_record(J, pi, R, cslot, lds_addr(lds));
    const int nw = (J.nrows + 63) >> 6;
    for (int w = 0; w < nw; ++w) out[i * nw + w] = R[J.off_j + w];
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

// Quadratic-character table of an odd n > 1: bit x of the table is set iff
// the Jacobi symbol (x/n) = +1, for x < n; one thread per 32-bit word.
extern "C" __global__ void __launch_bounds__(256) rs_jacobi_table(uint32_t n, uint32_t words,
                                                                  uint32_t* __restrict__ out) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= words) return;
    uint32_t bits = 0;
    for (uint32_t i = 0; i < 32; ++i) {
        const uint32_t x = w * 32 + i;
        if (x < n && jacobi_loop(x, n, 1) == 1) bits |= 1u << i;
    }
    out[w] = bits;
}

extern "C" __global__ void rs_zero_u32(uint32_t* x, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0;
}

// mont.h -- exact 64-bit modular arithmetic for gfx950.
//
// Residues are Montgomery forms x*R mod p with R = 2^64, for odd
// 3 <= p < 2^63.  Products are built from 32x32+64 -> 64 multiply-adds,
// which the AMDGPU backend selects as v_mad_u64_u32, so every carry is
// exact (the reference's PTX version, kernel_v0.24.6.cu:205-241, loses the
// limb-1 carry at :226).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef uint64_t u64;
typedef uint32_t u32;

// a*b + c, exact in 64 bits for 32-bit a, b, c (one v_mad_u64_u32).
__device__ __forceinline__ u64 mad32(u32 a, u32 b, u64 c) {
    return (u64)a * (u64)b + c;
}

__device__ __forceinline__ u32 lo32(u64 x) { return (u32)x; }
__device__ __forceinline__ u32 hi32(u64 x) { return (u32)(x >> 32); }

// -p^-1 mod 2^64 (p odd): Newton steps x <- x(2 - p x), four in 32 bits
// (3 -> 48 correct bits) and one in 64 bits.
__device__ __forceinline__ u64 mont_mprime(u64 p) {
    u32 p0 = lo32(p);
    u32 x = p0;
    x *= 2u - p0 * x;
    x *= 2u - p0 * x;
    x *= 2u - p0 * x;
    x *= 2u - p0 * x;
    u64 y = x;
    y *= 2ull - p * y;
    return 0ull - y;
}

// r - p if r >= p, else r (r < 2p): the borrow of the 64-bit subtraction
// selects, so the compare is the subtraction itself.
__device__ __forceinline__ u64 csub(u64 r, u64 p) {
    u64 d;
    const bool borrow = __builtin_sub_overflow(r, p, &d);
    return borrow ? r : d;
}

// Montgomery product a*b*2^-64 mod p.  Needs b < p and a < 2^64 (a may be
// an unreduced integer such as k or b): then a*b + m*p < 2^128 and the
// result before the final subtraction is < 2p.
//
// T = a*b as four partial products (t3:t2.lo:t0.lo), m = lo64(T)*mp mod 2^64
// (three products), then hi64(T + m*p) column by column (four products);
// the low 64 bits of T + m*p are zero by construction.
__device__ __forceinline__ u64 montmul(u64 a, u64 b, u64 p, u64 mp) {
    const u32 a0 = lo32(a), a1 = hi32(a), b0 = lo32(b), b1 = hi32(b);
    const u64 t0 = mad32(a0, b0, 0);
    const u64 t1 = mad32(a0, b1, hi32(t0));
    const u64 t2 = mad32(a1, b0, lo32(t1));
    const u64 t3 = mad32(a1, b1, hi32(t1)) + hi32(t2);
    const u32 l0 = lo32(t0), l1 = lo32(t2);
    const u32 mp0 = lo32(mp), mp1 = hi32(mp);
    const u64 q0 = mad32(l0, mp0, 0);
    const u32 m0 = lo32(q0);
    const u32 m1 = hi32(q0) + l0 * mp1 + l1 * mp0;
    const u32 p0 = lo32(p), p1 = hi32(p);
    const u64 s0 = mad32(m0, p0, l0);               // low word is 0
    const u64 s1 = mad32(m0, p1, hi32(s0));
    const u64 s2 = mad32(m1, p0, lo32(s1)) + l1;    // low word is 0
    const u64 s3 = mad32(m1, p1, hi32(s1)) + hi32(s2);
    return csub(t3 + s3, p);
}

// k*y*2^-32 mod p before the final subtraction, in [0, 2p), for 32-bit k
// and y < p: one 32-bit Montgomery step.  With a0 = k*lo(y), m = lo(a0)*mp
// and s0 = m*lo(p) + lo(a0) (whose low word is 0), the quotient
// (k*y + m*p)/2^32 is k*hi(y) + m*hi(p) + hi(a0) + hi(s0).  It is < 2p <
// 2^64, so no partial sum overflows and the last two terms enter as the
// 64-bit addend of the two multiply-adds.
__device__ __forceinline__ u64 halfmul_lazy(u32 k, u64 y, u64 p, u32 mp0) {
    const u64 a0 = mad32(k, lo32(y), 0);
    const u32 m = lo32(a0) * mp0;
    const u64 s0 = mad32(m, lo32(p), lo32(a0));     // low word is 0
    // k*hi(y) + m*hi(p) + hi(a0) + hi(s0), every partial sum below 2p
    return mad32(k, hi32(y), mad32(m, hi32(p), (u64)hi32(a0) + hi32(s0)));
}

// The low 32 bits of halfmul_lazy(k, y, p, mp0): bits 32..63 of
// k*y + m*p, which are the high word of (k*lo(y) + m*lo(p)) mod 2^64 plus
// the low words of k*hi(y) and m*hi(p).
__device__ __forceinline__ u32 halfmul_lazy_lo(u32 k, u64 y, u64 p, u32 mp0) {
    const u64 a0 = mad32(k, lo32(y), 0);
    const u32 m = lo32(a0) * mp0;
    const u64 x = mad32(m, lo32(p), a0);  // mod 2^64: only its high word is used
    return hi32(x) + k * hi32(y) + m * hi32(p);
}

// k*y*2^-32 mod p for 32-bit k and y < p.
__device__ __forceinline__ u64 halfmul(u32 k, u64 y, u64 p, u32 mp0) {
    return csub(halfmul_lazy(k, y, p, mp0), p);
}

// Shoup entry of a residue given as z = y 2^32 mod p: yq:lo(y) with the
// plain y = z 2^-32 mod p and yq = floor(y 2^32 / p).  One 32-bit
// Montgomery step gives both: its multiplier m = lo(z)*mp0 satisfies
// y 2^32 = z + m p with 0 <= z < p, so m is yq and y < p exactly.
__device__ __forceinline__ u64 shoup32_entry(u64 z, u64 p, u32 mp0) {
    const u32 yq = lo32(z) * mp0;
    const u64 s0 = mad32(yq, lo32(p), lo32(z));  // low word is 0
    const u32 ylo = yq * hi32(p) + hi32(z) + hi32(s0);
    return ((u64)yq << 32) | ylo;
}

// Low word of Shoup's product k*y - floor(k*yq/2^32)*p for a 32-bit k and
// the entry yq:lo(y): the product lies in [0, 2p) and exceeds p only when
// the reduced product is below p*k/2^32.  negp0 = -lo(p) mod 2^32.
// (Three instructions: k*ylo as a 64-bit product, the high product, and
// one multiply-add whose low word is the result.  Written in C, the
// compiler turns the last step into a multiply and a subtract because only
// the low word is used; the one-instruction asm keeps the multiply-add.)
// negp0 must be wave-uniform: it is an SGPR operand of the asm.
__device__ __forceinline__ u32 shoup32_lo(u32 k, u32 ylo, u32 yq, u32 negp0) {
    const u64 ky = mad32(k, ylo, 0);
    const u32 q = __umulhi(k, yq);
    u64 r, carry;
    asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(r), "=s"(carry) : "v"(q), "s"(negp0), "v"(ky));
    return lo32(r);
}

// k*y mod p from the entry yq:lo(y).  y is recovered from it: since
// y 2^32 = yq p + z with z < p, y - floor(yq p / 2^32) lies in
// [0, p/2^32 + 1), below 2^32, so the low word of y fixes it.
__device__ __forceinline__ u64 shoup32_full(u32 k, u64 ent, u64 p) {
    const u32 ylo = lo32(ent), yq = hi32(ent);
    const u64 a = mad32(yq, hi32(p), hi32(mad32(yq, lo32(p), 0)));
    const u64 y = a + (u32)(ylo - lo32(a));
    const u32 q = __umulhi(k, yq);
    return csub((u64)k * y - (u64)q * p, p);
}

// Products mod p < 2^50 in double precision, on residues kept as exact
// integers in (-p, p).  With P = a*b (|a|, |b| < p), h = fl(a*b) and
// l = P - h exactly (the FMA error term).  t = fl(h/p + 1.5*2^52) rounds
// h/p to an integer q (the constant fixes the rounding position, |h/p| <
// 2^50); the rounding of 1/p and of the product move h/p from P/p by less
// than 0.375, so |P - q p| < 0.875 p.  Then fma(-q, p, h) = h - q p, an
// integer below 2^53 in magnitude, is exact, and adding l is exact too.
constexpr double kFpRound = 6755399441055744.0;  // 1.5 * 2^52
constexpr double kFpLow = 4503599627370496.0;    // 2^52
constexpr uint64_t kFpMaxP = 1ull << 50;

__device__ __forceinline__ double fp_mulmod(double a, double b, double p, double pinv) {
#pragma clang fp contract(off)
    const double h = a * b;
    const double l = __fma_rn(a, b, -h);
    const double q = __fma_rn(h, pinv, kFpRound) - kFpRound;
    return __fma_rn(-q, p, h) + l;
}

// The low 64 bits of the residue r + 2^52 for an integer 0 <= r < 2^52:
// bits 0..51 are r itself.
__device__ __forceinline__ u64 fp_bits(double r) { return (u64)__double_as_longlong(r + kFpLow); }

// High 64 bits of x*y.
__device__ __forceinline__ u64 umulhi64(u64 x, u64 y) {
    const u64 t0 = mad32(lo32(x), lo32(y), 0);
    const u64 t1 = mad32(lo32(x), hi32(y), hi32(t0));
    const u64 t2 = mad32(hi32(x), lo32(y), lo32(t1));
    return mad32(hi32(x), hi32(y), hi32(t1)) + hi32(t2);
}

// Multiplication by a fixed residue (Shoup): x*w mod p for x < 2^64 and a
// plain w < p with ws = floor(w 2^64 / p).  q = hi64(x ws) is the quotient
// or one less, so x*w - q*p (computed mod 2^64) lies in [0, 2p).  With x a
// Montgomery form the result is the Montgomery form of the product.  For
// w~ = w R mod p (w's Montgomery form), w = redc(w~) and ws = w~ * mp mod
// 2^64 exactly, because w 2^64 = ws p + w~ and mp = -p^-1 mod 2^64.
__device__ __forceinline__ u64 shoup_mul(u64 x, u64 w, u64 ws, u64 p) {
    const u64 q = umulhi64(x, ws);
    return csub(x * w - q * p, p);
}

// a * 2^-64 mod p (Montgomery reduction of a 64-bit a): with m = a*mp mod
// 2^64, a + m*p == 0 mod 2^64, so (a + m*p) / 2^64 = hi64(m*p) + (a != 0),
// which is at most p.
__device__ __forceinline__ u64 redc(u64 a, u64 p, u64 mp) {
    const u64 r = umulhi64(a * mp, p) + (a != 0 ? 1 : 0);
    return r == p ? 0 : r;
}

// The Shoup pair (w, ws) of a Montgomery form wm.
__device__ __forceinline__ void shoup_pair(u64 wm, u64 p, u64 mp, u64* w, u64* ws) {
    *w = redc(wm, p, mp);
    *ws = wm * mp;
}

// x^e in Montgomery form (x and one = R mod p in Montgomery form).
__device__ __forceinline__ u64 mont_pow(u64 x, u32 e, u64 one, u64 p, u64 mp) {
    u64 r = one;
    while (e) {
        if (e & 1) r = montmul(r, x, p, mp);
        e >>= 1;
        if (e) x = montmul(x, x, p, mp);
    }
    return r;
}

// Binary Jacobi loop: t * (a/n) for odd n > 1 and a < 2^32.
__device__ __forceinline__ int jacobi_loop(u32 a, u32 n, int t) {
    while (a != 0) {
        const u32 s = __builtin_ctz(a);
        a >>= s;
        const u32 n8 = n & 7;
        if ((s & 1) && (n8 == 3 || n8 == 5)) t = -t;
        if (a < n) {
            const u32 tmp = a;
            a = n;
            n = tmp;
            if ((a & 3) == 3 && (n & 3) == 3) t = -t;
        }
        a -= n;
    }
    return n == 1 ? t : 0;
}

// Jacobi symbol (a/p) for a = 2^z * n (n odd, both 32-bit) and odd p > 1,
// with rn = floor((2^64 - 1) / n) precomputed: (2/p)^z, then reciprocity
// (n/p) = (p/n) (-1)^((n-1)/2 (p-1)/2) and p mod n by a Barrett quotient
// (at most 2 below the true one).
__device__ __forceinline__ int jacobi_pre(u32 n, u32 z, u64 rn, u64 p) {
    int t = 1;
    const u32 p8 = lo32(p) & 7;
    if ((z & 1) && (p8 == 3 || p8 == 5)) t = -t;
    if (n == 1) return t;
    if ((n & 3) == 3 && (p8 & 3) == 3) t = -t;
    const u64 q = umulhi64(p, rn);
    u64 r = p - mad32(lo32(q), n, (u64)(hi32(q) * n) << 32);
    if (r >= n) r -= n;
    if (r >= n) r -= n;
    return jacobi_loop((u32)r, n, t);
}

// Jacobi symbol (a/p) for 32-bit a and odd p > 1, in {-1, 0, 1}.
// Binary algorithm after one reduction p mod a.
__device__ __forceinline__ int jacobi32(u32 a, u64 p) {
    int t = 1;
    u32 n;
    if (p <= 0xffffffffull && a >= (u32)p) a %= (u32)p;
    if (a == 0) return p == 1 ? 1 : 0;
    const u32 z = __builtin_ctz(a);
    a >>= z;
    const u32 p8 = lo32(p) & 7;
    if ((z & 1) && (p8 == 3 || p8 == 5)) t = -t;
    if (a == 1) return t;
    // reciprocity: (a/p) = (p/a) (-1)^((a-1)/2 (p-1)/2), a odd > 1
    if ((a & 3) == 3 && (p8 & 3) == 3) t = -t;
    n = a;
    a = (u32)(p % n);
    while (a != 0) {
        const u32 s = __builtin_ctz(a);
        a >>= s;
        const u32 n8 = n & 7;
        if ((s & 1) && (n8 == 3 || n8 == 5)) t = -t;
        if (a < n) {
            const u32 tmp = a;
            a = n;
            n = tmp;
            if ((a & 3) == 3 && (n & 3) == 3) t = -t;
        }
        a -= n;
    }
    return n == 1 ? t : 0;
}

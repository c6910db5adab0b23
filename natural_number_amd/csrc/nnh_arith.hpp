// nnh_arith.hpp — modular arithmetic for the three NTT primes on CDNA4.
//
// The primes are the reference's p_i = k_i * 2^56 + 1 with k = 27, 58, 87
// (ntt_info.h:69-73).  Device code keeps residues in the ordinary domain and
// multiplies by constants held in Montgomery form (R = 2^64).  Because
// p^-1 mod 2^64 = 1 - k*2^56, the Montgomery quotient m differs from T's low
// word only in its top byte, and the reduction needs three 32x32 multiplies
// and no carry flags.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace nnh {

typedef uint64_t u64;
typedef uint32_t u32;
typedef int64_t i64;

constexpr u32 PRIME_K[3] = {27u, 58u, 87u};
__host__ __device__ constexpr u64 prime_p(u32 k) { return ((u64)k << 56) | 1u; }

__host__ __device__ __forceinline__ u32 lo32(u64 x) { return (u32)x; }
__host__ __device__ __forceinline__ u32 hi32(u64 x) { return (u32)(x >> 32); }
__host__ __device__ __forceinline__ u64 mk64(u32 lo, u32 hi) { return ((u64)hi << 32) | lo; }

// Residue words are read once and written once per pass: the transform
// passes move them with nontemporal loads and stores, so that the streamed
// data does not evict the twiddle tables every work-group re-reads from
// the caches (same-box A/B on the 2^30-bit product: 3.53 -> 3.38 ms).
// Twiddles and limbs keep ordinary loads.
// (Plain accesses where the header is compiled by a host compiler without
// the builtins, as the CPU check of this arithmetic does.)
// Every buffer these accesses reach is global memory (hipMalloc), and the
// accesses say so with an address-space cast: a pointer the kernels read
// from a job table is generic to the compiler, which would otherwise use
// flat instructions, and those count against the LDS wait counter as well
// as the vector-memory one, so each LDS transpose of a tile would also wait
// for the tile's outstanding loads and stores.
#if defined(__clang__)
typedef __attribute__((address_space(1))) u64 gu64;
__device__ __forceinline__ u64 ld_stream(const u64* p) { return __builtin_nontemporal_load((const gu64*)p); }
__device__ __forceinline__ void st_stream(u64* p, u64 v) { __builtin_nontemporal_store(v, (gu64*)p); }
template <class T>
__device__ __forceinline__ T ld_g(const T* p) { return *(const __attribute__((address_space(1))) T*)p; }
template <class T>
__device__ __forceinline__ void st_g(T* p, T v) { *(__attribute__((address_space(1))) T*)p = v; }
#else
__device__ __forceinline__ u64 ld_stream(const u64* p) { return *p; }
__device__ __forceinline__ void st_stream(u64* p, u64 v) { *p = v; }
template <class T>
__device__ __forceinline__ T ld_g(const T* p) { return *p; }
template <class T>
__device__ __forceinline__ void st_g(T* p, T v) { *p = v; }
#endif

// Full 64x64 -> 128-bit product from four 32x32+64 multiply-adds.
__host__ __device__ __forceinline__ void mul_wide(u64 a, u64 b, u64& lo, u64& hi) {
    u32 a0 = lo32(a), a1 = hi32(a), b0 = lo32(b), b1 = hi32(b);
    u64 p00 = (u64)a0 * b0;
    u64 p10 = (u64)a1 * b0 + hi32(p00);
    u64 p01 = (u64)a0 * b1 + lo32(p10);
    hi = (u64)a1 * b1 + hi32(p10) + hi32(p01);
    lo = mk64(lo32(p00), lo32(p01));
}

// One NTT prime p = k*2^56 + 1 (held in scalar registers inside kernels).
struct Prime {
    u64 p;
    u64 p2;   // 2p
    u64 np;   // -p mod 2^64
    u64 n2p;  // -2p mod 2^64
    u32 k;
    u32 k24;  // k << 24
};
__host__ __device__ constexpr Prime make_prime(u32 k) {
    return Prime{prime_p(k), prime_p(k) << 1, 0 - prime_p(k), 0 - (prime_p(k) << 1), k, k << 24};
}
// 4p < 2^64 holds for p0 and p1 (k = 27, 58) but not p2 (k = 87): only the
// first two primes use the lazy [0, 4p) butterflies below.
__host__ __device__ constexpr bool lazy_prime(int i) { return i < 2; }
__host__ __device__ constexpr Prime prime_of(int i) { return make_prime(PRIME_K[i]); }

// Montgomery reduction for p = k*2^56+1, R = 2^64: returns T * 2^-64 mod p
// in (0, 2p) for T = hi:lo < p * 2^64.
//   m = T_lo * p^-1 mod 2^64 = T_lo - (T_lo*k mod 2^8) * 2^56, so T - m p is
//   a multiple of 2^64 whose low word cancels exactly: (T - m p) / 2^64 =
//   T_hi - N with N = floor(m p / 2^64) in [0, p).  With m = mh*2^32 + l:
//   m p = m + m k 2^56, N = mh*k*2^24 + floor((l*k + (mh >> 24)) / 2^8).
// The result T_hi - N + p lies in (0, 2p) because T_hi < p.
__host__ __device__ __forceinline__ u64 redc(const Prime& P, u64 lo, u64 hi) {
    u32 l = lo32(lo);
    u32 mh = hi32(lo) - l * P.k24;                 // high word of m
    u64 s = (u64)l * P.k + (mh >> 24);             // < 2^40
    u64 N = (u64)mh * P.k24 + (u32)(s >> 8);
    return hi + P.p - N;
}

#if defined(__HIP_DEVICE_COMPILE__)
// Device form of mont(): the same reduction with the 64x64 product's
// column sums carried through SGPR carry flags instead of zero-extended
// 64-bit addends (no register moves), and floor((l k + (mh >> 24)) / 2^8)
// taken from the one multiply G = l * (k << 24) = (byte << 24) + 2^32 *
// floor(l k / 2^8) plus the carry of (byte << 24) + (mh & 0xff000000).
// Result (T_hi + p) - mh*k*2^24 - q in (0, 2p), as redc().  Every SGPR carry
// is read at least two instructions after the VALU instruction that wrote
// it, counting only instructions inside these blocks (gfx950 requires two
// wait states between a VALU SGPR write and a VALU read of it), so the
// order in which the compiler places the blocks cannot break it.
// mont_dev_add adds a 64-bit addend c in place of p: the result is
// T_hi + c - mh*k*2^24 - q = c - p + (T 2^-64 mod p, in (0, 2p)).
// Requires T = a*b < p 2^64, y1*w1 + c < 2^64 and T_hi + c < 2^64 (for
// c = p: b < 2^63).
template <bool VADD>
__device__ __forceinline__ u64 mont_dev_add(const Prime& P, u64 y, u64 w, u64 c) {
    u64 A, B, C, D, cM, d1;
    if constexpr (VADD) {
        asm("v_mad_u64_u32 %[A], %[d1], %[y0], %[w0], 0\n\t"
            "v_mad_u64_u32 %[B], %[d1], %[y1], %[w0], 0\n\t"
            "v_mad_u64_u32 %[C], %[cM], %[y0], %[w1], %[B]\n\t"
            "v_mad_u64_u32 %[D], %[d1], %[y1], %[w1], %[pp]"
            : [A] "=&v"(A), [B] "=&v"(B), [C] "=&v"(C), [D] "=&v"(D), [cM] "=&s"(cM), [d1] "=&s"(d1)
            : [y0] "v"(lo32(y)), [y1] "v"(hi32(y)), [w0] "v"(lo32(w)), [w1] "v"(hi32(w)), [pp] "v"(c));
    } else {
        asm("v_mad_u64_u32 %[A], %[d1], %[y0], %[w0], 0\n\t"
            "v_mad_u64_u32 %[B], %[d1], %[y1], %[w0], 0\n\t"
            "v_mad_u64_u32 %[C], %[cM], %[y0], %[w1], %[B]\n\t"
            "v_mad_u64_u32 %[D], %[d1], %[y1], %[w1], %[pp]"
            : [A] "=&v"(A), [B] "=&v"(B), [C] "=&v"(C), [D] "=&v"(D), [cM] "=&s"(cM), [d1] "=&s"(d1)
            : [y0] "v"(lo32(y)), [y1] "v"(hi32(y)), [w0] "v"(lo32(w)), [w1] "v"(hi32(w)), [pp] "s"(c));
    }
    // A = y0 w0, C = y0 w1 + y1 w0 (carry cM = 2^64), D = y1 w1 + c
    u32 L1;
    u64 c2, G, dG;
    asm("v_add_co_u32_e64 %[L1], %[c2], %[h0], %[c0]\n\t"
        "v_mad_u64_u32 %[G], %[dG], %[l0], %[k24], 0"
        : [L1] "=&v"(L1), [c2] "=&s"(c2), [G] "=&v"(G), [dG] "=&s"(dG)
        : [h0] "v"(hi32(A)), [c0] "v"(lo32(C)), [l0] "v"(lo32(A)), [k24] "s"(P.k24));
    // L1 = T bits 32-63; E = T_hi + c; mh = L1 - byte << 24; cq = carry of the byte sum
    u32 E0, E1, mh, shh, dq;
    u64 c3, cq, d2;
    asm("s_nop 0\n\t"
        "v_addc_co_u32_e64 %[E0], %[c3], %[d0], %[c1], %[c2]\n\t"
        "v_sub_u32 %[mh], %[L1], %[glo]\n\t"
        "v_and_b32 %[shh], 0xff000000, %[mh]\n\t"
        "v_addc_co_u32_e64 %[E1], %[d2], %[dd1], 0, %[c3]\n\t"
        "v_add_co_u32_e64 %[dq], %[cq], %[glo], %[shh]\n\t"
        "v_addc_co_u32_e64 %[E1], %[d2], %[E1], 0, %[cM]"
        : [E0] "=&v"(E0), [E1] "=&v"(E1), [mh] "=&v"(mh), [shh] "=&v"(shh), [dq] "=&v"(dq), [c3] "=&s"(c3),
          [cq] "=&s"(cq), [d2] "=&s"(d2)
        : [d0] "v"(lo32(D)), [c1] "v"(hi32(C)), [c2] "s"(c2), [L1] "v"(L1), [glo] "v"(lo32(G)), [dd1] "v"(hi32(D)),
          [cM] "s"(cM));
    // F = E + mh (2^32 - k 2^24) = E - mh k 2^24 + mh 2^32
    u64 F, dF;
    asm("v_mad_u64_u32 %[F], %[dF], %[mh], %[nk], %[E]"
        : [F] "=&v"(F), [dF] "=&s"(dF)
        : [mh] "v"(mh), [nk] "s"(0u - P.k24), [E] "v"(mk64(E0, E1)));
    // t = F - (q0 + cq) - mh 2^32
    u32 T0, T1;
    u64 b1, b2;
    asm("v_subb_co_u32_e64 %[T0], %[b1], %[f0], %[q0], %[cq]\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32_e64 %[T1], %[b2], %[f1], %[mh], %[b1]"
        : [T0] "=&v"(T0), [T1] "=&v"(T1), [b1] "=&s"(b1), [b2] "=&s"(b2)
        : [f0] "v"(lo32(F)), [f1] "v"(hi32(F)), [q0] "v"(hi32(G)), [mh] "v"(mh), [cq] "s"(cq));
    return mk64(T0, T1);
}
__device__ __forceinline__ u64 mont_dev(const Prime& P, u64 y, u64 w) { return mont_dev_add<false>(P, y, w, P.p); }

// (a*b + c*d) * 2^-64 mod p in (0, 2p) for p0 and p1 and canonical inputs
// (below p, so every high word is below 2^30): the four middle-column
// products then sum below 2^64 with no carry, the low column a0 b0 + c0 d0
// carries once into the high word, and the high column starts at p, which
// leaves E = T_hi + p for the reduction of mont_dev.  Carries are read at
// least two instructions after they are written, counting only
// instructions inside these blocks.
__device__ __forceinline__ u64 mont2_dev(const Prime& P, u64 a, u64 b, u64 c, u64 d) {
    u64 L, M, H, cL, sx;
    asm("v_mad_u64_u32 %[L], %[sx], %[a0], %[b0], 0\n\t"
        "v_mad_u64_u32 %[M], %[sx], %[a1], %[b0], 0\n\t"
        "v_mad_u64_u32 %[L], %[cL], %[c0], %[d0], %[L]\n\t"
        "v_mad_u64_u32 %[H], %[sx], %[a1], %[b1], %[pp]\n\t"
        "v_mad_u64_u32 %[M], %[sx], %[a0], %[b1], %[M]\n\t"
        "v_mad_u64_u32 %[H], %[sx], %[c1], %[d1], %[H]\n\t"
        "v_mad_u64_u32 %[M], %[sx], %[c1], %[d0], %[M]\n\t"
        "v_mad_u64_u32 %[M], %[sx], %[c0], %[d1], %[M]"
        : [L] "=&v"(L), [M] "=&v"(M), [H] "=&v"(H), [cL] "=&s"(cL), [sx] "=&s"(sx)
        : [a0] "v"(lo32(a)), [a1] "v"(hi32(a)), [b0] "v"(lo32(b)), [b1] "v"(hi32(b)), [c0] "v"(lo32(c)),
          [c1] "v"(hi32(c)), [d0] "v"(lo32(d)), [d1] "v"(hi32(d)), [pp] "s"(P.p));
    // T = L + M 2^32 + (H - p) 2^64, L < 2^65 (carry cL), M < 2^64
    u32 W1, H0, E0, E1p, E1, mh, shh, dq;
    u64 G, c1, ch, ce, cq, s2;
    asm("v_addc_co_u32_e64 %[H0], %[ch], %[h0], 0, %[cL]\n\t"
        "v_add_co_u32_e64 %[W1], %[c1], %[l1], %[m0]\n\t"
        "v_mad_u64_u32 %[G], %[s2], %[l0], %[k24], 0\n\t"
        "v_addc_co_u32_e64 %[E1p], %[s2], %[h1], 0, %[ch]\n\t"
        "v_addc_co_u32_e64 %[E0], %[ce], %[H0], %[m1], %[c1]"
        : [W1] "=&v"(W1), [H0] "=&v"(H0), [E0] "=&v"(E0), [E1p] "=&v"(E1p), [G] "=&v"(G), [c1] "=&s"(c1),
          [ch] "=&s"(ch), [ce] "=&s"(ce), [s2] "=&s"(s2)
        : [l0] "v"(lo32(L)), [l1] "v"(hi32(L)), [m0] "v"(lo32(M)), [m1] "v"(hi32(M)), [h0] "v"(lo32(H)),
          [h1] "v"(hi32(H)), [cL] "s"(cL), [k24] "s"(P.k24));
    asm("v_sub_u32 %[mh], %[W1], %[g0]\n\t"
        "v_and_b32 %[shh], 0xff000000, %[mh]\n\t"
        "v_add_co_u32_e64 %[dq], %[cq], %[g0], %[shh]\n\t"
        "v_addc_co_u32_e64 %[E1], %[s2], %[E1p], 0, %[ce]"
        : [mh] "=&v"(mh), [shh] "=&v"(shh), [dq] "=&v"(dq), [E1] "=&v"(E1), [cq] "=&s"(cq), [s2] "=&s"(s2)
        : [W1] "v"(W1), [g0] "v"(lo32(G)), [E1p] "v"(E1p), [ce] "s"(ce));
    // W1 = T bits 32-63, E = T_hi + p; then as mont_dev
    u64 F, dF;
    asm("v_mad_u64_u32 %[F], %[dF], %[mh], %[nk], %[E]"
        : [F] "=&v"(F), [dF] "=&s"(dF)
        : [mh] "v"(mh), [nk] "s"(0u - P.k24), [E] "v"(mk64(E0, E1)));
    u32 T0, T1;
    u64 b1, b2;
    asm("v_subb_co_u32_e64 %[T0], %[b1], %[f0], %[q0], %[cq]\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32_e64 %[T1], %[b2], %[f1], %[mh], %[b1]"
        : [T0] "=&v"(T0), [T1] "=&v"(T1), [b1] "=&s"(b1), [b2] "=&s"(b2)
        : [f0] "v"(lo32(F)), [f1] "v"(hi32(F)), [q0] "v"(hi32(G)), [mh] "v"(mh), [cq] "s"(cq));
    return mk64(T0, T1);
}
#endif

__host__ __device__ __forceinline__ u64 mont(const Prime& P, u64 a, u64 b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return mont_dev(P, a, b);
#else
    u64 lo, hi;
    mul_wide(a, b, lo, hi);
    return redc(P, lo, hi);
#endif
}

// Sum of two products, reduced: (a*b + c*d) * 2^-64 mod p in [0, 2p);
// requires a*b + c*d < p * 2^64.
__host__ __device__ __forceinline__ u64 mont2(const Prime& P, u64 a, u64 b, u64 c, u64 d) {
    u64 l1, h1, l2, h2;
    mul_wide(a, b, l1, h1);
    mul_wide(c, d, l2, h2);
    u64 lo = l1 + l2;
    u64 hi = h1 + h2 + (lo < l1 ? 1u : 0u);
    return redc(P, lo, hi);
}

// x in [0, 2p) -> [0, p): x - p (mod 2^64) is negative as a signed value
// exactly when x < p (p < 2^63), and its sign is the sign of its high word.
// On the device the subtraction is one 64-bit add of -p (v_lshl_add_u64
// with the scalar pair) and the test a 32-bit compare.
// (The empty asm hides the constant from the optimizer, which would
// otherwise turn the add back into a two-instruction subtraction and a
// 64-bit compare.)
__host__ __device__ __forceinline__ u64 opaque_s(u64 v) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+s"(v));
#endif
    return v;
}
__host__ __device__ __forceinline__ u64 csub(const Prime& P, u64 x) {
    u64 y = x + opaque_s(P.np);
    return (int32_t)hi32(y) < 0 ? x : y;
}

// x in [0, 4p) -> [0, 2p) (requires 2p < 2^63)
__host__ __device__ __forceinline__ u64 csub2(const Prime& P, u64 x) {
    u64 y = x + opaque_s(P.n2p);
    return (int32_t)hi32(y) < 0 ? x : y;
}

// Lazy butterflies after Harvey ("Faster arithmetic for number-theoretic
// transforms", J. Symbolic Comput. 60, 2014) for primes with 4p < 2^64:
// Montgomery products of inputs below 4p stay below 2p, so one conditional
// subtraction of 2p per butterfly replaces two of p.
// Forward: (x, y) in [0, 4p) -> (x + w*y, x - w*y) in [0, 4p).
// On the device a = csub2(x) and p are the Montgomery product's addend, so
// the reduction yields x' = a + t directly (t = w*y in (0, 2p)) and
// y' = 2(a + p) - x' = a + 2p - t.  The addend bound holds: y1*w1 <
// 4p^2/2^64 < 2^61.8 and a + p < 3p, so y1*w1 + a + p < 2^64; T_hi + a + p
// < 4p < 2^64.  a + p is formed as x + (x >= 2p ? -p : p): one 64-bit
// compare, two selects of constant halves and one 64-bit add, one VALU
// instruction fewer than the subtraction, sign test, selects and add of
// csub2(x) + p (the 2^30-bit product: 3.35-3.37 -> 3.33-3.34 ms, same box).
__host__ __device__ __forceinline__ void ct_bfly_lazy(const Prime& P, u64& x, u64& y, u64 wm) {
#if defined(__HIP_DEVICE_COMPILE__)
    const u64 ap = x + (x >= P.p2 ? P.np : P.p);  // csub2(x) + p
    const u64 X = mont_dev_add<true>(P, y, wm, ap);  // a + t
    y = (ap << 1) - X;
    x = X;
#else
    u64 t = mont(P, y, wm);  // y*w < 4p*p < p*2^64: t in [0, 2p)
    u64 a = csub2(P, x);
    x = a + t;
    y = a + P.p2 - t;
#endif
}
// Inverse: (x, y) in [0, 2p) -> (x + y, (x - y) * w) in [0, 2p).
__host__ __device__ __forceinline__ void gs_bfly_lazy(const Prime& P, u64& x, u64& y, u64 wm) {
    u64 a = csub2(P, x + y);
    u64 d = (x + P.p2) - y;  // (0, 4p)
    y = mont(P, d, wm);
    x = a;
}

// Cooley-Tukey butterfly (x, y) -> (x + w*y, x - w*y), inputs and outputs in
// [0, 2p); wm = w * 2^64 mod p.
__host__ __device__ __forceinline__ void ct_bfly(const Prime& P, u64& x, u64& y, u64 wm) {
    u64 t = csub(P, mont(P, y, wm));
    u64 a = csub(P, x);
    x = a + t;
    y = (a + P.p) - t;
}
// Gentleman-Sande butterfly (x, y) -> (x + y, (x - y) * w), canonical in/out.
__host__ __device__ __forceinline__ void gs_bfly(const Prime& P, u64& x, u64& y, u64 wm) {
    u64 a = csub(P, x + y);
    u64 d = (x + P.p) - y;
    y = csub(P, mont(P, d, wm));
    x = a;
}

// The same butterflies for twiddle 1 (block 0 of a level): no product;
// w*y is y brought into the range the Montgomery product would give.
__host__ __device__ __forceinline__ void ct_bfly_lazy_one(const Prime& P, u64& x, u64& y) {
    const u64 a = csub2(P, x), t = csub2(P, y);
    x = a + t;
    y = (a + P.p2) - t;
}
__host__ __device__ __forceinline__ void gs_bfly_lazy_one(const Prime& P, u64& x, u64& y) {
    const u64 a = csub2(P, x + y);
    y = csub2(P, (x + P.p2) - y);
    x = a;
}
__host__ __device__ __forceinline__ void ct_bfly_one(const Prime& P, u64& x, u64& y) {
    const u64 a = csub(P, x), t = csub(P, y);
    x = a + t;
    y = (a + P.p) - t;
}
__host__ __device__ __forceinline__ void gs_bfly_one(const Prime& P, u64& x, u64& y) {
    const u64 a = csub(P, x + y);
    y = csub(P, (x + P.p) - y);
    x = a;
}

// Base multiply of one 4-word group (reference asmNttMul0/1/2,
// asmLinux.asm:910-1367): (x0+x1X)(y0+y1X) mod (X^2 - w) and
// (x2+x3X)(y2+y3X) mod (X^2 + w).  Inputs in [0, 2p), or [0, 4p) when LAZY;
// outputs carry a factor 2^-64 (absorbed by the load scaling) and are
// canonical, or in (0, 2p) when LAZY on the device (the lazy inverse levels
// accept that).
template <bool LAZY = false>
__host__ __device__ __forceinline__ void base_mul4(const Prime& P, u64& x0, u64& x1, u64& x2, u64& x3, u64 y0,
                                                   u64 y1, u64 y2, u64 y3, u64 wm) {
    if constexpr (LAZY) {
        x0 = csub2(P, x0); x1 = csub2(P, x1); x2 = csub2(P, x2); x3 = csub2(P, x3);
        y0 = csub2(P, y0); y1 = csub2(P, y1); y2 = csub2(P, y2); y3 = csub2(P, y3);
    }
    x0 = csub(P, x0); x1 = csub(P, x1); x2 = csub(P, x2); x3 = csub(P, x3);
    y0 = csub(P, y0); y1 = csub(P, y1); y2 = csub(P, y2); y3 = csub(P, y3);
    const u64 t1 = csub(P, mont(P, y1, wm));        // w*y1
    const u64 t3 = P.p - csub(P, mont(P, y3, wm));  // -w*y3, in (0, p]
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (LAZY) {
        const u64 z0 = mont2_dev(P, x0, y0, x1, t1);
        const u64 z1 = mont2_dev(P, x0, y1, x1, y0);
        const u64 z2 = mont2_dev(P, x2, y2, x3, t3);
        const u64 z3 = mont2_dev(P, x2, y3, x3, y2);
        x0 = z0; x1 = z1; x2 = z2; x3 = z3;
        return;
    }
#endif
    const u64 z0 = mont2(P, x0, y0, x1, t1);
    const u64 z1 = mont2(P, x0, y1, x1, y0);
    const u64 z2 = mont2(P, x2, y2, x3, t3);
    const u64 z3 = mont2(P, x2, y3, x3, y2);
    x0 = csub(P, z0); x1 = csub(P, z1); x2 = csub(P, z2); x3 = csub(P, z3);
}

// Garner reconstruction of X < p0*p1*p2 < 2^186 from canonical residues, in
// the order of the reference's asmCRT (asmLinux.asm:1369-1450): first p0 and
// p2, then p1.  Constants: c02m = (p0^-1 mod p2)*2^64 mod p2,
// c1m = ((p0 p2)^-1 mod p1)*2^64 mod p1, km = (-p0 (p0 p2)^-1 mod p1)*2^64
// mod p1.  Output X as three words x0 + x1 2^64 + x2 2^128.
//   u = (r2 - r0) / p0 mod p2,  x02 = r0 + p0 u,
//   v = (r1 - x02) / (p0 p2) mod p1 = (r1 - r0) c1 + (u mod p1) k  (mod p1),
//   X = x02 + p0 p2 v,
// the last with p0 p2 = k0 k2 2^112 + (k0 + k2) 2^56 + 1 known at compile
// time, and v as one double Montgomery product (one reduction for both
// terms) instead of two products and a subtraction.
__host__ __device__ __forceinline__ void garner(u64 r0, u64 r1, u64 r2, u64 c02m, u64 c1m, u64 km, u64& x0,
                                                u64& x1, u64& x2) {
    constexpr Prime Q0 = prime_of(0), Q1 = prime_of(1), Q2 = prime_of(2);
    constexpr u64 P0 = Q0.p, P1 = Q1.p, P2 = Q2.p;
    static_assert(PRIME_K[0] + PRIME_K[2] < 256, "p0 p2 middle term fits one byte");
    constexpr u64 M02LO = ((u64)(PRIME_K[0] + PRIME_K[2]) << 56) + 1;  // p0 p2 mod 2^64
    constexpr u64 M02HI = (u64)(PRIME_K[0] * PRIME_K[2]) << 48;         // p0 p2 >> 64
    u64 d = r2 + P2 - r0;                                  // (0, 2 p2)
    u64 u = csub(Q2, mont(Q2, d, c02m));                   // (r2 - r0)/p0 mod p2
    u64 lo, hi;
    mul_wide(u, P0, lo, hi);
    lo += r0;
    hi += lo < r0 ? 1u : 0u;                               // x02 = r0 + p0 u
    const u64 e = csub(Q1, r1 + P1 - r0);                  // r1 - r0 mod p1
    const u64 uu = csub(Q1, u);                            // u mod p1 (u < p2 < 2 p1)
#if defined(__HIP_DEVICE_COMPILE__)
    const u64 v = csub(Q1, mont2_dev(Q1, e, c1m, uu, km));
#else
    const u64 v = csub(Q1, mont2(Q1, e, c1m, uu, km));
#endif
    u64 l1, h1, l2, h2;
    mul_wide(v, M02LO, l1, h1);
    mul_wide(v, M02HI, l2, h2);
    x0 = lo + l1;
    u64 mid = hi + h1 + (x0 < lo ? 1u : 0u);
    x1 = mid + l2;
    x2 = h2 + (x1 < mid ? 1u : 0u);
}

}  // namespace nnh

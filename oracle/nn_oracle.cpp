// nn_oracle.cpp — CPU oracle (test infrastructure only; see nn_oracle.hpp).
//
// Restates the reference CPU multiply path of Natural_Number in portable
// C++20 with unsigned __int128 in place of the x86-64 mulx/adc sequences.
// Citations are reference file:line.
#include "nn_oracle.hpp"

#include <algorithm>
#include <bit>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

namespace nno {

static inline u64 hi64(u128 x) { return (u64)(x >> 64); }
static inline bool neg(u64 x) { return (std::int64_t)x < 0; }

// ============================================================================
// Limbs: basic_natural.h:9-97
// ============================================================================
static u64* xalloc(u64 words) {
    u64* p = (u64*)std::malloc(words * sizeof(u64));
    if (!p) { std::fprintf(stderr, "oracle: out of memory\n"); std::abort(); }
    return p;
}
Limbs::Limbs() : data(xalloc(1)), size(1), capacity(1) {}
Limbs::Limbs(const Limbs& o) {
    size = o.size;
    capacity = std::bit_ceil(size);
    data = xalloc(capacity);
    std::memcpy(data, o.data, size * sizeof(u64));
}
Limbs::Limbs(Limbs&& o) noexcept : data(o.data), size(o.size), capacity(o.capacity) {
    o.data = nullptr;
    o.size = o.capacity = 0;
}
Limbs::~Limbs() { std::free(data); }
Limbs& Limbs::operator=(const Limbs& o) {
    if (this == &o) return *this;
    resize(o.size);
    std::memcpy(data, o.data, size * sizeof(u64));
    return *this;
}
Limbs& Limbs::operator=(Limbs&& o) noexcept {
    if (this != &o) {
        std::free(data);
        data = o.data; size = o.size; capacity = o.capacity;
        o.data = nullptr; o.size = o.capacity = 0;
    }
    return *this;
}
void Limbs::resize(u64 s) {
    reserve(s);
    size = s;
}
void Limbs::reserve(u64 s) {
    if (capacity < s) {
        capacity = std::bit_ceil(s);
        data = (u64*)std::realloc(data, capacity * sizeof(u64));
        if (!data) { std::fprintf(stderr, "oracle: out of memory\n"); std::abort(); }
    }
}

// ============================================================================
// mpn layer: asmLinux.asm:7-344 and base.h:1-91
// ============================================================================
// asmAdd0 (asmLinux.asm:7-28): d = a + b over n words, returns carry.
static int add_n(u64 n, u64* d, const u64* a, const u64* b) {
    u64 c = 0;
    for (u64 i = 0; i < n; i++) {
        u128 s = (u128)a[i] + b[i] + c;
        d[i] = (u64)s;
        c = hi64(s);
    }
    return (int)c;
}
// asmAdd1 (asmLinux.asm:30-62): d = s + cf over n words, returns carry.
static int add_1(u64 n, const u64* s, u64* d, int cf) {
    for (u64 i = 0; i < n; i++) {
        u64 x = s[i] + (u64)cf;
        cf = cf && x == 0;
        d[i] = x;
    }
    return cf;
}
// asmSub0 (asmLinux.asm:64-85): d = a - b over n words, returns borrow.
static int sub_n(u64 n, u64* d, const u64* a, const u64* b) {
    u64 c = 0;
    for (u64 i = 0; i < n; i++) {
        u64 x = a[i], y = b[i];
        u64 r = x - y - c;
        c = (x < y) || (x == y && c);
        d[i] = r;
    }
    return (int)c;
}
// asmSub1 (asmLinux.asm:87-119): d = s - cf over n words, returns borrow.
static int sub_1(u64 n, const u64* s, u64* d, int cf) {
    for (u64 i = 0; i < n; i++) {
        u64 x = s[i];
        d[i] = x - (u64)cf;
        cf = cf && x == 0;
    }
    return cf;
}
// base.h:7-10 (add), sizes s1 >= s2.
static int add(u64* d, const u64* a, u64 s1, const u64* b, u64 s2) {
    int cf = add_n(s2, d, a, b);
    return add_1(s1 - s2, a + s2, d + s2, cf);
}
// base.h:13-16 (sub), a >= b.
static int sub(u64* d, const u64* a, u64 s1, const u64* b, u64 s2) {
    int cf = sub_n(s2, d, a, b);
    return sub_1(s1 - s2, a + s2, d + s2, cf);
}
// asmShl (asmLinux.asm:121-139): d[0..n] = s << c, 1 <= c <= 63 (n+1 words).
static void shl_words(unsigned c, u64 n, const u64* s, u64* d) {
    u64 carry = 0;
    for (u64 i = n; i-- > 0;) {
        u64 x = s[i];
        d[i + 1] = (carry << c) | (x >> (64 - c));
        carry = x;
    }
    d[0] = carry << c;
}
// asmShr (asmLinux.asm:141-164): d[0..n) = s >> c, 1 <= c <= 63.
static void shr_words(unsigned c, u64 n, const u64* s, u64* d) {
    for (u64 i = 0; i + 1 < n; i++) d[i] = (s[i] >> c) | (s[i + 1] << (64 - c));
    d[n - 1] = s[n - 1] >> c;
}
// asmMul0 (asmLinux.asm:166-186): d[0..n] = s * m (n+1 words).
static void mul_1(u64 n, u64 m, const u64* s, u64* d) {
    u64 c = 0;
    for (u64 i = 0; i < n; i++) {
        u128 t = (u128)s[i] * m + c;
        d[i] = (u64)t;
        c = hi64(t);
    }
    d[n] = c;
}
// asmMul1 (asmLinux.asm:188-278): column-wise schoolbook, n1 <= n2,
// d[0..n1+n2) = a*b with a 3-word column accumulator.
static void mul_school(u64 n1, u64 n2, const u64* a, const u64* b, u64* d) {
    u64 c0 = 0, c1 = 0, c2 = 0;
    for (u64 k = 0; k + 1 < n1 + n2; k++) {
        u64 lo = k >= n2 ? k - n2 + 1 : 0;
        u64 hi = std::min(k, n1 - 1);
        for (u64 i = lo; i <= hi; i++) {
            u128 t = (u128)a[i] * b[k - i];
            u128 s = (u128)c0 + (u64)t;
            c0 = (u64)s;
            s = (u128)c1 + hi64(t) + hi64(s);
            c1 = (u64)s;
            c2 += hi64(s);
        }
        d[k] = c0;
        c0 = c1; c1 = c2; c2 = 0;
    }
    d[n1 + n2 - 1] = c0;
}
// asmSqr (asmLinux.asm:280-344): d[0..2n) = a^2, doubling the cross terms.
static void sqr_school(u64 n, u64* d, const u64* a) {
    u64 c0 = 0, c1 = 0;
    for (u64 k = 0; k + 1 < 2 * n; k++) {
        u64 lo = k >= n ? k - n + 1 : 0;
        u64 hi = std::min(k, n - 1);
        u64 s0 = 0, s1 = 0, s2 = 0;
        u64 i = lo, j = hi;
        while (i < j) {
            u128 t = (u128)a[i] * a[j];
            u128 s = (u128)s0 + (u64)t; s0 = (u64)s;
            s = (u128)s1 + hi64(t) + hi64(s); s1 = (u64)s; s2 += hi64(s);
            i++; j--;
        }
        // double
        s2 = (s2 << 1) | (s1 >> 63); s1 = (s1 << 1) | (s0 >> 63); s0 <<= 1;
        if (i == j) {
            u128 t = (u128)a[i] * a[i];
            u128 s = (u128)s0 + (u64)t; s0 = (u64)s;
            s = (u128)s1 + hi64(t) + hi64(s); s1 = (u64)s; s2 += hi64(s);
        }
        u128 s = (u128)s0 + c0; s0 = (u64)s;
        s = (u128)s1 + c1 + hi64(s); s1 = (u64)s; s2 += hi64(s);
        d[k] = s0;
        c0 = s1; c1 = s2;
    }
    d[2 * n - 1] = c0;
}

static const u64 TOOM22_MUL = 32;  // base.h:18
static const u64 TOOM22_SQR = 48;  // base.h:37

// toom22mul (base.h:21-35): Karatsuba, both operands s words.
static void karatsuba(u64* d, const u64* a, const u64* b, u64 s) {
    if (s < TOOM22_MUL) return mul_school(s, s, a, b, d);
    u64 s0 = (s + 1) >> 1, s1 = s - s0;
    karatsuba(d, a, b, s0);
    karatsuba(d + 2 * s0, a + s0, b + s0, s1);
    std::vector<u64> a1(s0 + 1), b1(s0 + 1), d1(2 * (s0 + 1));
    a1[s0] = add(a1.data(), a, s0, a + s0, s1);
    b1[s0] = add(b1.data(), b, s0, b + s0, s1);
    karatsuba(d1.data(), a1.data(), b1.data(), s0 + 1);
    sub(d1.data(), d1.data(), 2 * (s0 + 1), d, 2 * s0);
    sub(d1.data(), d1.data(), 2 * (s0 + 1), d + 2 * s0, 2 * s1);
    add(d + s0, d + s0, s + s1, d1.data(), 2 * (s0 + 1));
}
// toom22sqr (base.h:40-53)
static void karatsuba_sqr(u64* d, const u64* a, u64 s) {
    if (s < TOOM22_SQR) return sqr_school(s, d, a);
    u64 s0 = (s + 1) >> 1, s1 = s - s0;
    karatsuba_sqr(d, a, s0);
    karatsuba_sqr(d + 2 * s0, a + s0, s1);
    std::vector<u64> a1(s0 + 1), d1(2 * (s0 + 1));
    a1[s0] = add(a1.data(), a, s0, a + s0, s1);
    karatsuba_sqr(d1.data(), a1.data(), s0 + 1);
    sub(d1.data(), d1.data(), 2 * (s0 + 1), d, 2 * s0);
    sub(d1.data(), d1.data(), 2 * (s0 + 1), d + 2 * s0, 2 * s1);
    add(d + s0, d + s0, s + s1, d1.data(), 2 * (s0 + 1));
}
// mul_base (base.h:56-81): n1 >= n2, unbalanced operands chunked by n2.
static void mul_base(u64* d, const u64* a, u64 n1, const u64* b, u64 n2) {
    if (n2 <= 1) {
        mul_1(n1, *b, a, d);
    } else if (n2 < TOOM22_MUL) {
        mul_school(n2, n1, b, a, d);
    } else {
        karatsuba(d, a, b, n2);
        a += n2; d += n2; n1 -= n2;
        std::memset(d + n2, 0, sizeof(u64) * n1);
        std::vector<u64> t(2 * n2);
        while (n1 >= n2) {
            karatsuba(t.data(), a, b, n2);
            add(d, d, 2 * n2, t.data(), 2 * n2);
            a += n2; d += n2; n1 -= n2;
        }
        if (n1 > 0) {
            mul_base(t.data(), b, n2, a, n1);
            add(d, d, n1 + n2, t.data(), n1 + n2);
        }
    }
}
// sqr_base (base.h:84-91)
static void sqr_base(u64* d, const u64* a, u64 n) {
    if (n <= 1) mul_1(n, *a, a, d);
    else if (n < TOOM22_MUL) sqr_school(n, d, a);
    else karatsuba_sqr(d, a, n);
}

// ============================================================================
// modular helpers: asmLinux.asm:720-774
// ============================================================================
// asmDiv (asmLinux.asm:720-727): floor((hi:lo) / m), hi < m.
static inline u64 div_2w(u64 lo, u64 hi, u64 m) { return (u64)((((u128)hi << 64) | lo) / m); }
// asmMod (asmLinux.asm:739-751): a mod p via q = hi(a * r), r = ceil(2^64/p).
static inline u64 mod_r64(u64 a, u64 r, u64 p) {
    u64 q = hi64((u128)a * r);
    u64 t = a - q * p;
    return neg(t) ? t + p : t;
}
// asmMulMod (asmLinux.asm:753-774): a*b mod p, Barrett with r = ceil(2^124/p).
static inline u64 mulmod_r124(u64 a, u64 b, u64 p, u64 r) {
    u128 t = (u128)a * b;
    u64 lo = (u64)t, hi = hi64(t);
    u128 m = (u128)hi * r + hi64((u128)lo * r);
    u64 q = (u64)(m >> 60);
    u64 x = lo - q * p;
    return neg(x) ? x + p : x;
}
static inline u64 shoup(u64 w, u64 p) { return div_2w(0, w, p) + 1; }
// exact helpers for constant derivation
static inline u64 mulmod_exact(u64 a, u64 b, u64 p) { return (u64)((u128)a * b % p); }
static u64 powmod(u64 b, u64 e, u64 p) {
    u64 r = 1;
    for (; e; e >>= 1, b = mulmod_exact(b, b, p))
        if (e & 1) r = mulmod_exact(r, b, p);
    return r;
}

// ============================================================================
// NTT engine constants: ntt_info.h:69-145
// ============================================================================
const u64 NttInfo::P[3] = {27ull << 56 | 1, 58ull << 56 | 1, 87ull << 56 | 1};
// ceil(2^124 / p) (ntt_info.h:75-79)
const u64 NttInfo::R124[3] = {
    (u64)(((u128)1 << 124) / (27ull << 56 | 1)) + 1,
    (u64)(((u128)1 << 124) / (58ull << 56 | 1)) + 1,
    (u64)(((u128)1 << 124) / (87ull << 56 | 1)) + 1,
};
// Primitive 2^56-th roots of unity: the last entry of each root_base row
// (ntt_info.h:89, 98, 107).  Every other entry is a repeated square of it.
static const u64 ROOT_GEN[3] = {11, 2, 310};
u64 NttInfo::RB[3][56];
u64 NttInfo::IRB[3][56];
Limbs NttInfo::root[3];
Limbs NttInfo::iroot[3];
int NttInfo::root_scale = 0;  // 0 = not yet initialised; becomes 1 (ntt_info.h:145)

static u64 R64(int i) { return (u64)(((u128)1 << 64) / NttInfo::P[i]) + 1; }  // ntt_info.h:148

static void ensure_constants() {
    if (NttInfo::root_scale) return;
    for (int i = 0; i < 3; i++) {
        u64 p = NttInfo::P[i];
        NttInfo::RB[i][55] = ROOT_GEN[i];
        for (int j = 54; j >= 0; j--) NttInfo::RB[i][j] = mulmod_exact(NttInfo::RB[i][j + 1], NttInfo::RB[i][j + 1], p);
        for (int j = 0; j < 56; j++) NttInfo::IRB[i][j] = powmod(NttInfo::RB[i][j], p - 2, p);
        // root[i] = iroot[i] = { 1, ceil(2^64/p) } (ntt_info.h:141-143)
        NttInfo::root[i].resize(2);
        NttInfo::root[i][0] = 1; NttInfo::root[i][1] = R64(i);
        NttInfo::iroot[i].resize(2);
        NttInfo::iroot[i][0] = 1; NttInfo::iroot[i][1] = R64(i);
    }
    NttInfo::root_scale = 1;
}

// ntt_info.h:22-41: grow the (w, shoup(w)) pair tables to 2^(scale-2) pairs.
// Pair t holds w_t = prod over set bits q of t of RB[q+1].
void NttInfo::init(int s) {
    ensure_constants();
    s -= 1;
    if (s <= root_scale) return;
    for (int i = 0; i < 3; i++) {
        for (int dir = 0; dir < 2; dir++) {
            Limbs& t = dir ? iroot[i] : root[i];
            const u64* base = dir ? IRB[i] : RB[i];
            t.resize(1ull << s);
            for (int j = root_scale; j < s; j++)
                for (u64 k = 0, n = 1ull << j; k < n; k += 2) {
                    t[n + k] = mulmod_r124(t[k], base[j], P[i], R124[i]);
                    t[n + k + 1] = shoup(t[n + k], P[i]);
                }
        }
    }
    root_scale = s;
}

// ntt_info.h:147-173 with asmLoad (asmLinux.asm:877-908) / asmMod.
void NttInfo::load(const u64* src, u64 size, int s) {
    scale = s;
    u64 N = 1ull << (s - 1);
    for (int i = 0; i < 3; i++) v[i].resize(N + N);
    for (int i = 0; i < 3; i++) {
        u64 p = P[i], r = R64(i);
        u64* x = v[i].data;
        if (size <= N) {
            for (u64 j = 0; j < size; j++) x[j] = x[j + N] = mod_r64(src[j], r, p);
            std::memset(x + size, 0, sizeof(u64) * (N - size));
            std::memset(x + N + size, 0, sizeof(u64) * (N - size));
        } else {
            u64 j = 0;
            for (; j < size - N; j++) {
                u64 n0 = mod_r64(src[j], r, p), n1 = mod_r64(src[j + N], r, p);
                x[j] = n0 + n1;
                x[j + N] = n0 - n1 + p;
            }
            for (; j < N; j++) x[j] = x[j + N] = mod_r64(src[j], r, p);
        }
    }
    init(scale);
}

// One Cooley-Tukey level of half-size h over blocks [beg, end) — asmNtt
// (asmLinux.asm:346-458, unit first block) and asmNtt2 (:460-534).
// tw points at the (w, shoup) pair of the first block.  Values stay in [0,2p).
static void ct_level(u64 h, u64 p, u64* beg, const u64* tw, u64* end, bool unit_first) {
    for (u64* blk = beg; blk < end; blk += 2 * h, tw += 2) {
        if (unit_first && blk == beg) {
            for (u64 k = 0; k < h; k++) {
                u64 x = blk[k], y = blk[k + h];
                if (x >= p) x -= p;
                if (y >= p) y -= p;
                blk[k] = x + y;
                blk[k + h] = x + p - y;
            }
            continue;
        }
        u64 w = tw[0], ws = tw[1];
        for (u64 k = 0; k < h; k++) {
            u64 x = blk[k], y = blk[k + h];
            u64 q = hi64((u128)ws * y);
            u64 t = w * y - q * p;
            if (neg(t)) t += p;
            if (x >= p) x -= p;
            blk[k] = x + t;
            blk[k + h] = x + p - t;
        }
    }
}
// One Gentleman-Sande level — asmINtt (asmLinux.asm:536-643) and asmINtt2
// (:645-718).  Inputs and outputs canonical [0,p).
static void gs_level(u64 h, u64 p, u64* beg, const u64* tw, u64* end, bool unit_first) {
    for (u64* blk = beg; blk < end; blk += 2 * h, tw += 2) {
        if (unit_first && blk == beg) {
            for (u64 k = 0; k < h; k++) {
                u64 x = blk[k], y = blk[k + h];
                u64 a = x + y, b = x - y;
                if (!neg(a - p)) a -= p;
                if (neg(b)) b += p;
                blk[k] = a;
                blk[k + h] = b;
            }
            continue;
        }
        u64 w = tw[0], ws = tw[1];
        for (u64 k = 0; k < h; k++) {
            u64 x = blk[k], y = blk[k + h];
            u64 a = x + y, d = x - y + p;
            if (!neg(a - p)) a -= p;
            u64 q = hi64((u128)ws * d);
            u64 t = w * d - q * p;
            if (neg(t)) t += p;
            blk[k] = a;
            blk[k + h] = t;
        }
    }
}
// asmINttShr (asmLinux.asm:776-875): last GS level (w = 1) fused with the
// multiplication by 2^-c, c = scale-1, using p = 1 (mod 2^56):
//   v * 2^-c = ((v-1) >>arith c) + 1 + ((-v) & (2^c - 1)) * (p >> c).
static void gs_last_scaled(int s, u64 p, u64* v) {
    unsigned c = (unsigned)(s - 1);
    u64 N = 1ull << c, mask = N - 1, ps = p >> c;
    auto fix = [&](u64 x) {
        u64 t = x - 1;
        return (u64)((std::int64_t)t >> c) + 1 + (mask & ~t) * ps;
    };
    for (u64 k = 0; k < N; k++) {
        u64 x = v[k], y = v[k + N];
        u64 a = x + y, b = x - y;
        if (!neg(a - p)) a -= p;
        if (neg(b)) b += p;
        v[k] = fix(a);
        v[k + N] = fix(b);
    }
}

static const int L3_SCALE = 20;  // ntt_info.h:179

// ntt_info.h:180-198
void NttInfo::forward() {
    for (int i = 0; i < 3; i++) {
        u64* x = v[i].data;
        const u64* r = root[i].data;
        u64 n = v[i].size;
        if (scale <= L3_SCALE) {
            for (u64 h = 1ull << (scale - 2); h > 1; h >>= 1) ct_level(h, P[i], x, r, x + n, true);
            continue;
        }
        for (u64 h = 1ull << (scale - 2); h >= 1ull << L3_SCALE; h >>= 1) ct_level(h, P[i], x, r, x + n, true);
        for (int j = L3_SCALE - 1; j > 0; j--) ct_level(1ull << j, P[i], x, r, x + (1ull << L3_SCALE), true);
        for (u64 k = 1; k < n >> L3_SCALE; k++)
            for (int j = L3_SCALE - 1; j > 0; j--)
                ct_level(1ull << j, P[i], x + (k << L3_SCALE), r + (k << (L3_SCALE - j)), x + ((k + 1) << L3_SCALE), false);
    }
}

// ntt_info.h:200-221
void NttInfo::inverse() {
    for (int i = 0; i < 3; i++) {
        u64* x = v[i].data;
        const u64* r = iroot[i].data;
        u64 n = v[i].size;
        if (scale <= L3_SCALE) {
            for (u64 h = 2; h < 1ull << (scale - 1); h <<= 1) gs_level(h, P[i], x, r, x + n, true);
            gs_last_scaled(scale, P[i], x);
            continue;
        }
        for (int j = 1; j < L3_SCALE; j++) gs_level(1ull << j, P[i], x, r, x + (1ull << L3_SCALE), true);
        for (u64 k = 1; k < n >> L3_SCALE; k++)
            for (int j = 1; j < L3_SCALE; j++)
                gs_level(1ull << j, P[i], x + (k << L3_SCALE), r + (k << (L3_SCALE - j)), x + ((k + 1) << L3_SCALE), false);
        for (u64 h = 1ull << L3_SCALE; h < 1ull << (scale - 1); h <<= 1) gs_level(h, P[i], x, r, x + n, true);
        gs_last_scaled(scale, P[i], x);
    }
}

// Base multiply of asmNttMul0/1/2 (asmLinux.asm:910-1367): per 4-word group g
// with twiddle w_g,  (x0+x1X)(y0+y1X) mod (X^2 - w_g)  and
// (x2+x3X)(y2+y3X) mod (X^2 + w_g).  128-bit sums reduced by Barrett with
// floor(2^(124+i)/p_i) (asmLinux.asm:926, 1075, 1236).
static const int BAR_SHIFT[3] = {4, 3, 2};
static const int BAR_FIX[3] = {1, 2, 1};
static u64 bar_const(int i) { return (u64)(((u128)1 << (124 + i)) / NttInfo::P[i]); }
static inline u64 barrett(int i, u64 br, u128 T) {
    u64 p = NttInfo::P[i];
    u64 t = (u64)(T >> (64 - BAR_SHIFT[i]));
    u64 q = hi64((u128)t * br);
    u64 r = (u64)T - q * p;
    for (int n = 0; n < BAR_FIX[i]; n++)
        if (!neg(r - p)) r -= p;
    return r;
}
static void base_mul(int i, u64 L, u64* d, const u64* a, const u64* b, const u64* tw) {
    u64 p = NttInfo::P[i], br = bar_const(i);
    auto red = [p](u64 x) { return x >= p ? x - p : x; };
    for (u64 g = 0; g < L / 4; g++, a += 4, b += 4, d += 4, tw += 2) {
        u64 x0 = red(a[0]), x1 = red(a[1]), x2 = red(a[2]), x3 = red(a[3]);
        u64 y0 = red(b[0]), y1 = red(b[1]), y2 = red(b[2]), y3 = red(b[3]);
        u64 w = tw[0], ws = tw[1];
        u64 r1 = barrett(i, br, (u128)x0 * y1 + (u128)x1 * y0);
        u64 q = hi64((u128)y1 * ws);
        u64 t1 = y1 * w - q * p;
        if (neg(t1)) t1 += p;
        u64 r0 = barrett(i, br, (u128)x0 * y0 + (u128)t1 * x1);
        q = hi64((u128)y3 * ws);
        u64 t3 = q * p - y3 * w;
        if (neg(t3)) t3 += p;
        u64 r2 = barrett(i, br, (u128)x3 * t3 + (u128)x2 * y2);
        u64 r3 = barrett(i, br, (u128)x2 * y3 + (u128)x3 * y2);
        d[0] = r0; d[1] = r1; d[2] = r2; d[3] = r3;
    }
}

// ntt_info.h:223-231
void NttInfo::mul(const NttInfo& a, const NttInfo& b) {
    scale = a.scale;
    u64 L = 1ull << scale;
    for (int i = 0; i < 3; i++) {
        v[i].resize(a.v[i].size);
        base_mul(i, L, v[i].data, a.v[i].data, b.v[i].data, root[i].data);
    }
}

// Garner constants of asmCRT (asmLinux.asm:1394-1425), derived.
struct CrtConst {
    u64 c02, c02s, c01q, c1, c1s, m02lo, m02hi;
    CrtConst() {
        u64 m0 = NttInfo::P[0], m1 = NttInfo::P[1], m2 = NttInfo::P[2];
        c02 = powmod(m0 % m2, m2 - 2, m2);                 // m0^-1 mod m2
        c02s = shoup(c02, m2);
        c01q = div_2w(0, m0, m1) + 1;                      // floor(m0 * 2^64 / m1) + 1
        c1 = powmod(mulmod_exact(m0 % m1, m2 % m1, m1), m1 - 2, m1);  // (m0 m2)^-1 mod m1
        c1s = shoup(c1, m1);
        u128 m02 = (u128)m0 * m2;
        m02lo = (u64)m02; m02hi = hi64(m02);
    }
};
static const CrtConst& crtc() { static CrtConst c; return c; }

// asmCRT (asmLinux.asm:1369-1450): Garner reconstruction (p0,p2 then p1) of
// X_k < 2^186 and the serial 128-bit carry.  Output word k overwrites v[0][k];
// the final 2-word carry lands in v[1][0..1] and its low word is returned.
u64 NttInfo::crt(u64 n) {
    const CrtConst& C = crtc();
    u64 m0 = P[0], m1 = P[1], m2 = P[2];
    u64* o = v[0].data;
    const u64* r1p = v[1].data;
    const u64* r2p = v[2].data;
    u64 clo = 0, chi = 0;
    for (u64 k = 0; k < n; k++) {
        u64 r0 = o[k], r1 = r1p[k], r2 = r2p[k];
        u64 d = r2 - r0 + m2;
        u64 q = hi64((u128)C.c02s * d);
        u64 u = d * C.c02 - q * m2;
        if (neg(u)) u += m2;
        u128 x = (u128)u * m0;
        u64 q2 = hi64((u128)C.c01q * u);
        u64 e = (q2 + 2) * m1 + (r1 - (u64)x - r0);
        q = hi64((u128)C.c1s * e);
        u64 vv = e * C.c1 - q * m1;
        if (neg(vv)) vv += m1;
        x += r0;
        u128 plo = (u128)vv * C.m02lo, phi = (u128)vv * C.m02hi;
        // X = x + plo + (phi << 64), 3 words (w0, w1, w2); then add carry.
        u128 s = (u128)(u64)x + (u64)plo;
        u64 w0 = (u64)s;
        u128 mid = (u128)hi64(x) + hi64(plo) + hi64(s);   // fits (x < 2^124)
        u128 s2 = (u128)(u64)phi + (u64)mid;
        u64 w1 = (u64)s2;
        u64 w2 = hi64(phi) + hi64(mid) + hi64(s2);
        s = (u128)w0 + clo;
        o[k] = (u64)s;
        s2 = (u128)w1 + chi + hi64(s);
        clo = (u64)s2;
        chi = w2 + hi64(s2);
    }
    v[1].data[0] = clo;
    v[1].data[1] = chi;
    return clo;
}

// ntt_info.h:237-248: exact mode writes the final carry as the top word;
// cyclic mode (size-1 > L) folds the carry and the end-around carry, giving
// the product mod (2^(64L) - 1).
void NttInfo::save(Limbs& out, u64 size) {
    u64 L = 1ull << scale;
    if (size - 1 > L) {
        crt(L);
        int cf = add(v[0].data, v[0].data, v[0].size, v[1].data, 2);
        add_1(v[0].size, v[0].data, v[0].data, cf);
    } else {
        v[0].resize(size);
        v[0][size - 1] = crt(size - 1);
    }
    out = std::move(v[0]);
}

// ============================================================================
// natural: natural.h
// ============================================================================
static const u64 NTT_MUL_THRESHOLD = 384;  // ntt_info.h:6
static const u64 NTT_DIV_THRESHOLD = 256;  // ntt_info.h:7
static const int NTT_SCALE_THRESHOLD = 10;  // ntt_info.h:8

Nat::Nat(u64 x) { data[0] = x; }
Nat::Nat(const u64* w, u64 n) {
    resize(n ? n : 1);
    if (n) std::memcpy(data, w, n * sizeof(u64));
    else data[0] = 0;
}
// natural.h:82-93
Nat& Nat::normalize() {
    if (!size) { size = 1; data[0] = 0; return *this; }
    while (--size)
        if (data[size]) break;
    size++;
    return *this;
}

// natural.h:145-179
bool operator<(const Nat& a, const Nat& b) {
    if (a.size != b.size) return a.size < b.size;
    for (u64 i = a.size; i-- > 0;)
        if (a[i] != b[i]) return a[i] < b[i];
    return false;
}
bool operator==(const Nat& a, const Nat& b) {
    if (a.size != b.size) return false;
    for (u64 i = a.size; i-- > 0;)
        if (a[i] != b[i]) return false;
    return true;
}
static bool is_zero(const Nat& a) { return a.size == 1 && a[0] == 0; }

// natural.h:244-267 (_shl)
Nat shl(const Nat& a, u64 bits) {
    Nat d;
    if (is_zero(a)) return d;
    u64 as = a.size, n = bits >> 6;
    unsigned b = bits & 63;
    if (b == 0) {
        d.resize(as + n);
        std::memmove(d.data + n, a.data, as * sizeof(u64));
        std::memset(d.data, 0, n * sizeof(u64));
        return d;
    }
    d.resize(as + n + 1);
    shl_words(b, as, a.data, d.data + n);
    std::memset(d.data, 0, n * sizeof(u64));
    d.normalize();
    return d;
}
// natural.h:281-295 (_shr)
Nat shr(const Nat& a, u64 bits) {
    Nat d;
    u64 as = a.size, n = bits >> 6;
    if (as <= n) return d;
    unsigned b = bits & 63;
    d.resize(as - n);
    if (b == 0) {
        std::memmove(d.data, a.data + n, (as - n) * sizeof(u64));
        return d;
    }
    shr_words(b, as - n, a.data + n, d.data);
    d.normalize();
    return d;
}
// natural.h:303-311 (_add)
Nat add(const Nat& x, const Nat& y) {
    const Nat *a = &x, *b = &y;
    if (a->size < b->size) std::swap(a, b);
    Nat d;
    d.resize(a->size + 1);
    d[a->size] = add(d.data, a->data, a->size, b->data, b->size);
    d.normalize();
    return d;
}
// natural.h:324-342 (_sub); returns |x - y| like the reference.
Nat sub(const Nat& x, const Nat& y) {
    const Nat *a = &x, *b = &y;
    if (a->size < b->size) std::swap(a, b);
    u64 as = a->size, bs = b->size;
    Nat d;
    if (as == bs) {
        while (--bs)
            if ((*a)[bs] != (*b)[bs]) break;
        if ((*a)[bs] < (*b)[bs]) std::swap(a, b);
        d.resize(++bs);
        sub_n(bs, d.data, a->data, b->data);
    } else {
        d.resize(as);
        sub(d.data, a->data, as, b->data, bs);
    }
    d.normalize();
    return d;
}

// natural.h:356-369 (_mul_base)
void mul_basecase(Nat& d, const Nat& x, const Nat& y) {
    const Nat *a = &x, *b = &y;
    if (a->size < b->size) std::swap(a, b);
    Nat r;
    r.resize(a->size + b->size);
    mul_base(r.data, a->data, a->size, b->data, b->size);
    r.normalize();
    d = std::move(r);
}
// natural.h:371-382 (_sqr_base)
static void sqr_basecase(Nat& d, const Nat& a) {
    Nat r;
    r.resize(a.size << 1);
    sqr_base(r.data, a.data, a.size);
    r.normalize();
    d = std::move(r);
}
// natural.h:384-386
int ntt_scale(u64 s1, u64 s2) { return 64 - std::countl_zero(s1 + s2 - 2); }

// natural.h:452-467 (mul_save_NTT_info, transform of a kept in ni_a)
static void mul_save(Nat& d, const Nat& a, const Nat& b, NttInfo& ni_a) {
    if (ni_a.scale < NTT_SCALE_THRESHOLD) {
        mul_basecase(d, a, b);
        ni_a.scale = 0;
        return;
    }
    u64 as = a.size, bs = b.size;
    ni_a.load(a.data, as, ni_a.scale);
    NttInfo ni_b;
    ni_b.load(b.data, bs, ni_a.scale);
    ni_a.forward();
    ni_b.forward();
    ni_b.mul(ni_a, ni_b);
    ni_b.inverse();
    ni_b.save(d, as + bs);
    d.normalize();
}
// natural.h:469-486 (both transforms kept)
static void mul_save2(Nat& d, const Nat& a, const Nat& b, NttInfo& ni_a, NttInfo& ni_b) {
    if (ni_a.scale < NTT_SCALE_THRESHOLD) {
        mul_basecase(d, a, b);
        ni_a.scale = 0;
        ni_b.scale = 0;
        return;
    }
    u64 as = a.size, bs = b.size;
    ni_a.load(a.data, as, ni_a.scale);
    ni_b.load(b.data, bs, ni_a.scale);
    ni_a.forward();
    ni_b.forward();
    NttInfo ni;
    ni.mul(ni_a, ni_b);
    ni.inverse();
    ni.save(d, as + bs);
    d.normalize();
}
// natural.h:488-502 (sqr_save_NTT_info)
static void sqr_save(Nat& d, const Nat& a, NttInfo& ni_a) {
    if (ni_a.scale < NTT_SCALE_THRESHOLD) {
        sqr_basecase(d, a);
        ni_a.scale = 0;
        return;
    }
    u64 as = a.size;
    ni_a.load(a.data, as, ni_a.scale);
    NttInfo ni_b;
    ni_a.forward();
    ni_b.mul(ni_a, ni_a);
    ni_b.inverse();
    ni_b.save(d, as * 2);
    d.normalize();
}
// natural.h:504-515 (mul_load_NTT_info, reuses ni_a)
static void mul_load(Nat& d, const Nat& a, const Nat& b, const NttInfo& ni_a) {
    if (ni_a.scale == 0) {
        mul_basecase(d, a, b);
        return;
    }
    u64 as = a.size, bs = b.size;
    NttInfo ni_b;
    ni_b.load(b.data, bs, ni_a.scale);
    ni_b.forward();
    ni_b.mul(ni_a, ni_b);
    ni_b.inverse();
    ni_b.save(d, as + bs);
    d.normalize();
}
// natural.h:517-527 (both transforms reused)
static void mul_load2(Nat& d, const Nat& a, const Nat& b, const NttInfo& ni_a, const NttInfo& ni_b) {
    if (ni_a.scale == 0) {
        mul_basecase(d, a, b);
        return;
    }
    u64 as = a.size, bs = b.size;
    NttInfo ni;
    ni.mul(ni_a, ni_b);
    ni.inverse();
    ni.save(d, as + bs);
    d.normalize();
}

// natural.h:391-431 (_mul_NTT), including the one-level unbalanced split.
void mul_ntt(Nat& d, const Nat& x, const Nat& y) {
    const Nat *a = &x, *b = &y;
    if (a->size < b->size) std::swap(a, b);
    u64 a0_size = (a->size + 1) >> 1;
    int s = ntt_scale(a->size, b->size);
    int s_opt = ntt_scale(a0_size, b->size);
    if (s == s_opt) {
        NttInfo ni_a, ni_b;
        ni_a.load(a->data, a->size, s);
        ni_b.load(b->data, b->size, s);
        ni_a.forward();
        ni_b.forward();
        ni_a.mul(ni_a, ni_b);
        ni_a.inverse();
        Nat r;
        ni_a.save(r, a->size + b->size);
        r.normalize();
        d = std::move(r);
        return;
    }
    Nat a0, a1;
    a0.resize(a0_size);
    a1.resize(a->size - a0_size);
    std::memcpy(a0.data, a->data, sizeof(u64) * a0_size);
    std::memcpy(a1.data, a->data + a0_size, sizeof(u64) * a1.size);
    Nat bb = *b;
    NttInfo ni_b;
    ni_b.scale = s_opt;
    mul_save(a0, bb, a0, ni_b);
    mul_load(a1, bb, a1, ni_b);
    d = add(shl(a1, a0_size << 6), a0);
}
// natural.h:433-442 (_sqr_NTT)
void sqr_ntt(Nat& d, const Nat& a) {
    int s = ntt_scale(a.size, a.size);
    NttInfo ni;
    ni.load(a.data, a.size, s);
    ni.forward();
    ni.mul(ni, ni);
    ni.inverse();
    Nat r;
    ni.save(r, a.size * 2);
    r.normalize();
    d = std::move(r);
}
// natural.h:444-450 (_mul)
Nat mul(const Nat& a, const Nat& b) {
    Nat d;
    if (std::min(a.size, b.size) < NTT_MUL_THRESHOLD) mul_basecase(d, a, b);
    else mul_ntt(d, a, b);
    return d;
}
// natural.h:529-535 (_sqr)
Nat sqr(const Nat& a) {
    Nat d;
    if (a.size < NTT_MUL_THRESHOLD) sqr_basecase(d, a);
    else sqr_ntt(d, a);
    return d;
}
// natural.h:554-564
Nat pow(const Nat& b, u64 e) {
    if (e == 0) return Nat(1);
    if (e == 1 || b < Nat(2)) return b;
    Nat p = pow(b, e >> 1);
    p = sqr(p);
    if (e & 1) p = mul(p, b);
    return p;
}

// ---- division: natural.h:566-769 ------------------------------------------
// natural.h:566-573 (div_base): single-word divisor.
static u64 div_base(const Nat& a, u64 b, Nat& q) {
    Nat r;
    r.resize(a.size);
    u64 rem = 0;
    for (u64 i = a.size; i-- > 0;) {
        u128 num = ((u128)rem << 64) | a[i];
        r[i] = (u64)(num / b);
        rem = (u64)(num % b);
    }
    r.normalize();
    q = std::move(r);
    return rem;
}

// natural.h:576-617 (reciprocal without a cached divisor transform)
static Nat reciprocal_plain(const Nat& divs, u64 divd_size) {
    u64 expected_acc = divd_size + 1 - divs.size;
    Nat ans = add(Nat((1ull << 63) | (divs[divs.size - 2] >> 1)), Nat(1));
    if (ans[0] == 0) ans[0] = 1ull << 63;
    else ans[0] = div_2w(0, 1ull << 63, ans[0]);
    for (;;) {
        u64 cur_acc = ans.size;
        if (cur_acc >= expected_acc + 1) return shr(ans, (cur_acc - expected_acc) << 6);
        Nat tmp;
        if (divs.size > 2 * cur_acc + 1) tmp = shr(divs, (divs.size - 2 * cur_acc - 1) << 6);
        else tmp = shl(divs, (2 * cur_acc + 1 - divs.size) << 6);
        NttInfo ni_ans;
        ni_ans.scale = ntt_scale(tmp.size, 1);
        mul_save(tmp, ans, tmp, ni_ans);
        tmp.resize(tmp.size - cur_acc);
        for (u64 i = 0; i < tmp.size; i++) tmp[i] = ~tmp[i + cur_acc];
        tmp.normalize();
        int lz = std::countl_zero(tmp[tmp.size - 1]);
        u64 nxt_acc = std::min(2 * cur_acc, (2 * cur_acc - tmp.size) * 2 + (u64)((lz + 31) >> 5));
        if (nxt_acc == NTT_DIV_THRESHOLD) nxt_acc = NTT_DIV_THRESHOLD - 1;
        tmp = shr(tmp, (2 * cur_acc - nxt_acc) << 6);
        if (is_zero(tmp)) tmp = Nat(1);
        mul_load(tmp, ans, tmp, ni_ans);
        ans = shl(ans, (nxt_acc - cur_acc) << 6);
        add(ans.data, ans.data, ans.size, tmp.data + cur_acc, tmp.size - cur_acc);
    }
}

struct RefDefect {};  // input on which the reference's algorithm is undefined

// natural.h:620-687 (reciprocal keeping the divisor transform)
static Nat reciprocal_ntt(const Nat& divisor, u64 dividend_size, NttInfo& ni_divisor) {
    u64 expected_acc = dividend_size + 1 - divisor.size;
    Nat ans = add(Nat((1ull << 63) | (divisor[divisor.size - 2] >> 1)), Nat(1));
    ans[0] = ans[0] ? div_2w(0, 1ull << 63, ans[0]) : 1ull << 63;
    for (u64 i = 1; i < divisor.size; i *= 2) {
        u64 prev_acc = ans.size, seg_size = 2 * prev_acc + 1;
        if (2 * i >= divisor.size) seg_size = divisor.size;
        const u64* seg = divisor.data + divisor.size - seg_size;
        Nat tmp;
        NttInfo ni_ans;
        if (2 * i <= NTT_DIV_THRESHOLD) {
            tmp.resize(seg_size + prev_acc);
            mul_base(tmp.data, seg, seg_size, ans.data, prev_acc);
            tmp.normalize();
        } else {
            int s = ntt_scale(i, i);
            ni_ans.load(ans.data, ans.size, s);
            ni_divisor.load(seg, seg_size, s);
            ni_ans.forward();
            ni_divisor.forward();
            NttInfo ni_tmp;
            ni_tmp.mul(ni_ans, ni_divisor);
            ni_tmp.inverse();
            ni_tmp.save(tmp, seg_size + prev_acc);
            tmp.normalize();
        }
        u64 n_shl = seg_size - 1 - prev_acc;
        while (tmp.size > n_shl + 1 && ~tmp[tmp.size - 1] == 0) tmp.size -= 1;
        int lz = std::countl_zero(~tmp[tmp.size - 1]);
        u64 cur_acc = std::min(2 * prev_acc, (2 * prev_acc - tmp.size + n_shl) * 2 + (u64)((lz + 31) >> 5));
        if (cur_acc == NTT_DIV_THRESHOLD) cur_acc -= 1;
        // The reference continues with inc_acc = cur_acc - prev_acc wrapped
        // to ~2^64 and memsets that many words (natural.h:675-677): it faults
        // for divisors whose normalised leading words are all ones (the
        // one-word start estimate overflows to two words).  Report it.
        if (cur_acc < prev_acc) throw RefDefect{};
        n_shl += 2 * prev_acc - cur_acc;
        if (tmp.size <= n_shl) {
            tmp = Nat(1);
        } else {
            tmp.resize(tmp.size - n_shl);
            for (u64 k = 0; k < tmp.size; k++) tmp[k] = ~tmp[k + n_shl];
            tmp.normalize();
        }
        mul_load(tmp, ans, tmp, ni_ans);
        tmp = shr(tmp, prev_acc << 6);
        u64 inc_acc = cur_acc - prev_acc;
        // the reference writes up to cur_acc words before resizing (natural.h:675-680)
        tmp.reserve(std::max(cur_acc, inc_acc + prev_acc));
        if (inc_acc >= tmp.size) {
            std::memset(tmp.data + tmp.size, 0, sizeof(u64) * (inc_acc - tmp.size));
            std::memcpy(tmp.data + inc_acc, ans.data, sizeof(u64) * prev_acc);
        } else {
            add(tmp.data + inc_acc, ans.data, prev_acc, tmp.data + inc_acc, tmp.size - inc_acc);
        }
        tmp.resize(cur_acc);
        ans = std::move(tmp);
        if (cur_acc >= expected_acc + 1) return shr(ans, (cur_acc - expected_acc) << 6);
    }
    return shr(ans, 64);
}

// natural.h:689-747
static void div_iterative(const Nat& divd, const Nat& divs, Nat& q, Nat& r) {
    int n_shl = (std::countl_zero(divs[divs.size - 1]) + 1) & 63;
    Nat b = shl(divs, n_shl);
    r = shl(divd, n_shl);
    NttInfo ni_b;
    Nat rec = reciprocal_ntt(b, r.size, ni_b);
    NttInfo ni_rec;
    ni_rec.scale = ntt_scale(rec.size, rec.size);
    int sb = ntt_scale(b.size, 1);
    u64 Lb = 1ull << sb;
    if (ni_b.scale != sb && sb >= NTT_SCALE_THRESHOLD - 1) {
        ni_b.load(b.data, b.size, sb);
        ni_b.forward();
    }
    q.resize(r.size - b.size + 1);
    std::memset(q.data, 0, sizeof(u64) * q.size);
    Nat tmp;
    bool first = true;
    for (;;) {
        u64 idx = 0;
        if (r.size > rec.size + b.size - 1) idx = r.size - rec.size - b.size + 1;
        Nat top = shr(r, (idx + b.size - 1) << 6);
        if (first) mul_save(tmp, rec, top, ni_rec);
        else mul_load(tmp, rec, top, ni_rec);
        first = false;
        tmp = shr(tmp, rec.size << 6);
        add(q.data + idx, q.data + idx, q.size - idx, tmp.data, tmp.size);
        mul_load(tmp, b, tmp, ni_b);
        sub(r.data + idx, r.data + idx, r.size - idx, tmp.data, tmp.size);
        r.normalize();
        if (ni_b.scale) {
            if (r.size - idx > Lb) {
                int cf = add(r.data + idx, r.data + idx, Lb, r.data + idx + Lb, r.size - idx - Lb);
                add_1(Lb, r.data + idx, r.data + idx, cf);
                r.resize(idx + Lb);
            }
            if (r.size == idx + Lb && r[r.size - 1] >> 63) r.resize(idx);
            r.normalize();
        }
        if (idx == 0) break;
    }
    while (r >= b) {
        q = add(q, Nat(1));
        r = sub(r, b);
    }
    r = shr(r, n_shl);
}

// natural.h:749-769 (_div)
void divmod(const Nat& a, const Nat& b, Nat& q, Nat& r) {
    if (is_zero(b)) {
        q = Nat(is_zero(a) ? 1 : 0);
        r = Nat(0);
        return;
    }
    if (a < b) {
        Nat aa = a;
        q = Nat(0);
        r = std::move(aa);
        return;
    }
    if (b.size == 1) {
        Nat qq;
        u64 rem = div_base(a, b[0], qq);
        q = std::move(qq);
        r = Nat(rem);
        return;
    }
    Nat qq, rr;
    div_iterative(a, b, qq, rr);
    q = std::move(qq);
    r = std::move(rr);
}

// natural.h:841-857
void calc_e(Nat& num, Nat& den, u64 begin, u64 end) {
    if (end - begin == 2) {
        num = Nat(end);
        den = Nat(begin + 1);
        if (begin) den = mul(den, Nat(begin));
        return;
    }
    Nat num1, den1;
    calc_e(num, den, begin, (begin + end) >> 1);
    calc_e(num1, den1, (begin + end) >> 1, end);
    NttInfo ni_den1;
    ni_den1.scale = ntt_scale(std::max(den.size, num.size), den1.size);
    mul_save(num, den1, num, ni_den1);
    mul_load(den, den1, den, ni_den1);
    num = add(num, num1);
}

// natural.h:859-872
void sqrt_10005(Nat& num, Nat& den, u64 acc) {
    num = Nat(4001);
    den = Nat(40);
    Nat cache1, cache2;
    while (num.size + den.size - 2 < acc) {
        NttInfo ni_num, ni_den;
        ni_num.scale = ni_den.scale = ntt_scale(num.size, num.size);
        sqr_save(cache1, den, ni_den);
        sqr_save(cache2, num, ni_num);
        mul_load2(den, num, den, ni_num, ni_den);
        den = shl(den, 1);
        num = add(mul(cache1, Nat(10005)), cache2);
    }
}

// natural.h:874-902
void calc_pi(Nat& num, Nat& den, Nat& numc, u64 begin, u64 end) {
    if (end - begin == 1) {
        den = Nat(begin);
        den = mul(mul(mul(den, den), den), Nat(0x26dd041d878000ull));
        numc = Nat(2 * begin - 1);
        numc = mul(numc, Nat(6 * begin - 1));
        numc = mul(numc, Nat(6 * begin - 5));
        num = Nat(begin);
        num = mul(add(mul(Nat(0x207e2da6ull), num), Nat(0xcf6371)), numc);
        return;
    }
    Nat num1, den1, numc1;
    calc_pi(num, den, numc, begin, (begin + end) >> 1);
    calc_pi(num1, den1, numc1, (begin + end) >> 1, end);
    if (end - begin == 2) {
        num = sub(mul(numc, num1), mul(num, den1));  // natural subtraction: |a - b|
        den = mul(den, den1);
        numc = mul(numc, numc1);
        return;
    }
    NttInfo ni_den1, ni_numc;
    ni_den1.scale = ntt_scale(std::max(den.size, num.size), den1.size);
    mul_save(num, den1, num, ni_den1);
    mul_load(den, den1, den, ni_den1);
    ni_numc.scale = ntt_scale(std::max(numc1.size, num1.size), numc.size);
    mul_save(num1, numc, num1, ni_numc);
    mul_load(numc, numc, numc1, ni_numc);
    num = add(num, num1);
}

// ---- decimal output: natural.h:904-1039 ------------------------------------
static const u64 DEC_LEN = 18, DEC_BASE0 = 1000000000000000000ull;  // natural.h:904
static Nat dec_base[64], dec_base_r[64];                             // natural.h:905
static NttInfo ni_dec_base[64], ni_dec_base_r[64];                    // natural.h:906
static bool dec_ready = false;

// natural.h:909-923
static void calc_dec_base(int i) {
    if (!dec_ready) { dec_base[0] = Nat(DEC_BASE0); dec_ready = true; }
    if (is_zero(dec_base[i + 1])) {
        ni_dec_base[i].scale = ntt_scale(dec_base[i].size, dec_base[i].size);
        sqr_save(dec_base[i + 1], dec_base[i], ni_dec_base[i]);
        int n_shl = (std::countl_zero(dec_base[i][dec_base[i].size - 1]) + 1) & 63;
        dec_base_r[i] = reciprocal_plain(shl(dec_base[i], n_shl), dec_base[i].size * 2 + 1);
        ni_dec_base_r[i].scale = ntt_scale(dec_base_r[i].size, dec_base[i].size + 2);
        if (ni_dec_base_r[i].scale < NTT_SCALE_THRESHOLD) {
            ni_dec_base_r[i].scale = 0;
        } else {
            ni_dec_base_r[i].load(dec_base_r[i].data, dec_base_r[i].size, ni_dec_base_r[i].scale);
            ni_dec_base_r[i].forward();
        }
    }
}
// natural.h:953-967
static void dec_split(const Nat& x, int i, Nat& hi, Nat& lo) {
    if (i == 0) {
        lo = Nat(div_base(x, DEC_BASE0, hi));
        return;
    }
    int n_shl = (std::countl_zero(dec_base[i][dec_base[i].size - 1]) + 1) & 63;
    mul_load(hi, dec_base_r[i], shr(x, (dec_base[i].size - 1) << 6), ni_dec_base_r[i]);
    hi = shr(hi, ((dec_base[i].size + 2) << 6) - n_shl);
    mul_load(lo, dec_base[i], hi, ni_dec_base[i]);
    lo = sub(x, lo);
    while (lo >= dec_base[i]) {
        hi = add(hi, Nat(1));
        lo = sub(lo, dec_base[i]);
    }
}
// natural.h:969-992
static void print_dec(const Nat& x, std::string& out, int i, bool first) {
    if (i == -1) {
        char buf[32];
        if (first) std::snprintf(buf, sizeof buf, "%llu", (unsigned long long)x[0]);
        else std::snprintf(buf, sizeof buf, "%018llu", (unsigned long long)x[0]);
        out += buf;
        return;
    }
    Nat hi, lo;
    dec_split(x, i, hi, lo);
    if (first && is_zero(hi)) {
        print_dec(lo, out, i - 1, true);
        return;
    }
    print_dec(hi, out, i - 1, first);
    print_dec(lo, out, i - 1, false);
}
// natural.h:1013-1018
static int dec_init(const Nat& n) {
    if (!dec_ready) { dec_base[0] = Nat(DEC_BASE0); dec_ready = true; }
    int i = 0;
    for (; n >= dec_base[i]; i++) calc_dec_base(i);
    return i - 1;
}
// natural.h:1020-1039 (decimal branch of operator<<)
std::string to_dec(const Nat& n) {
    std::string out;
    print_dec(n, out, dec_init(n), true);
    return out;
}

}  // namespace nno

// ============================================================================
// C-ABI for the test suite (ctypes).  Sizes are in 64-bit words.
// ============================================================================
using namespace nno;

static u64 emit(const Nat& r, u64* dst, u64 cap) {
    u64 n = std::min(r.size, cap);
    std::memcpy(dst, r.data, n * sizeof(u64));
    if (cap > n) std::memset(dst + n, 0, (cap - n) * sizeof(u64));
    return r.size;
}

extern "C" {

// Product a*b through the reference dispatch (natural::_mul).  dst holds
// an+bn words (zero-filled above the normalised size); returns that size.
u64 nno_mul(u64* dst, const u64* a, u64 an, const u64* b, u64 bn) {
    return emit(mul(Nat(a, an), Nat(b, bn)), dst, an + bn);
}
u64 nno_sqr(u64* dst, const u64* a, u64 an) { return emit(sqr(Nat(a, an)), dst, 2 * an); }
// Individual algorithm arms.
u64 nno_mul_base(u64* dst, const u64* a, u64 an, const u64* b, u64 bn) {
    Nat d;
    mul_basecase(d, Nat(a, an), Nat(b, bn));
    return emit(d, dst, an + bn);
}
u64 nno_mul_ntt(u64* dst, const u64* a, u64 an, const u64* b, u64 bn) {
    Nat d;
    mul_ntt(d, Nat(a, an), Nat(b, bn));
    return emit(d, dst, an + bn);
}
u64 nno_sqr_ntt(u64* dst, const u64* a, u64 an) {
    Nat d;
    sqr_ntt(d, Nat(a, an));
    return emit(d, dst, 2 * an);
}
// One direct transform product at a given scale (load, NTT, mul, INTT,
// save(size=an+bn)).  With an+bn-1 > 2^scale this is the cyclic product
// mod 2^(64*2^scale)-1 (ntt_info.h:238-242): dst then holds 2^scale words.
u64 nno_ntt_product(u64* dst, u64 cap, const u64* a, u64 an, const u64* b, u64 bn, int scale) {
    NttInfo x, y;
    x.load(a, an, scale);
    y.load(b, bn, scale);
    x.forward();
    y.forward();
    x.mul(x, y);
    x.inverse();
    Limbs out;
    x.save(out, an + bn);
    u64 n = std::min(out.size, cap);
    std::memcpy(dst, out.data, n * sizeof(u64));
    return out.size;
}
// Stage outputs for parity checks; out holds 3 * 2^scale words (prime-major).
void nno_stage_load(u64* out, const u64* a, u64 an, int scale) {
    NttInfo x;
    x.load(a, an, scale);
    u64 L = 1ull << scale;
    for (int i = 0; i < 3; i++) std::memcpy(out + i * L, x.v[i].data, L * sizeof(u64));
}
void nno_stage_forward(u64* out, const u64* a, u64 an, int scale) {
    NttInfo x;
    x.load(a, an, scale);
    x.forward();
    u64 L = 1ull << scale;
    for (int i = 0; i < 3; i++) std::memcpy(out + i * L, x.v[i].data, L * sizeof(u64));
}
void nno_stage_pointwise(u64* out, const u64* a, u64 an, const u64* b, u64 bn, int scale) {
    NttInfo x, y;
    x.load(a, an, scale);
    y.load(b, bn, scale);
    x.forward();
    y.forward();
    x.mul(x, y);
    u64 L = 1ull << scale;
    for (int i = 0; i < 3; i++) std::memcpy(out + i * L, x.v[i].data, L * sizeof(u64));
}
void nno_stage_inverse(u64* out, const u64* a, u64 an, const u64* b, u64 bn, int scale) {
    NttInfo x, y;
    x.load(a, an, scale);
    y.load(b, bn, scale);
    x.forward();
    y.forward();
    x.mul(x, y);
    x.inverse();
    u64 L = 1ull << scale;
    for (int i = 0; i < 3; i++) std::memcpy(out + i * L, x.v[i].data, L * sizeof(u64));
}
// Twiddle table of prime i (pairs w, shoup) grown to 'scale'; out holds
// 2^(scale-1) words.  inverse selects iroot.
void nno_twiddles(u64* out, int i, int scale, int inverse) {
    NttInfo::init(scale);
    const Limbs& t = inverse ? NttInfo::iroot[i] : NttInfo::root[i];
    std::memcpy(out, t.data, (1ull << (scale - 1)) * sizeof(u64));
}
void nno_root_bases(u64* out_rb, u64* out_irb) {
    ensure_constants();
    std::memcpy(out_rb, NttInfo::RB, sizeof NttInfo::RB);
    std::memcpy(out_irb, NttInfo::IRB, sizeof NttInfo::IRB);
}
// Derived constants: [P0..2, R124 0..2, BAR 0..2, c02, c02s, c01q, c1, c1s, m02lo, m02hi]
void nno_constants(u64* out) {
    const CrtConst& c = crtc();
    for (int i = 0; i < 3; i++) {
        out[i] = NttInfo::P[i];
        out[3 + i] = NttInfo::R124[i];
        out[6 + i] = bar_const(i);
    }
    out[9] = c.c02; out[10] = c.c02s; out[11] = c.c01q; out[12] = c.c1; out[13] = c.c1s;
    out[14] = c.m02lo; out[15] = c.m02hi;
}
// Quotient and remainder (natural::_div).  q holds an words, r holds bn words.
// Returns 0, or 2 when the reference's Newton division is undefined for this
// divisor (see reciprocal_ntt).
int nno_divmod(u64* q, u64* r, const u64* a, u64 an, const u64* b, u64 bn, u64* qn, u64* rn) {
    Nat qq, rr;
    try {
        divmod(Nat(a, an), Nat(b, bn), qq, rr);
    } catch (const RefDefect&) {
        *qn = *rn = 0;
        return 2;
    }
    *qn = emit(qq, q, an);
    *rn = emit(rr, r, bn);
    return 0;
}
// b^e (natural.h:554-564).  Returns the size; dst holds cap words.
u64 nno_pow(u64* dst, u64 cap, u64 base, u64 e) { return emit(pow(Nat(base), e), dst, cap); }
// Decimal string of a (natural.h:1020-1039).  Returns the digit count;
// writes at most cap-1 digits plus a NUL.
u64 nno_to_dec(char* out, u64 cap, const u64* a, u64 an) {
    std::string s = to_dec(Nat(a, an));
    if (cap) {
        u64 n = std::min<u64>(s.size(), cap - 1);
        std::memcpy(out, s.data(), n);
        out[n] = 0;
    }
    return s.size();
}
// The e recipe of test.cpp:12-19: calc_e(num, den, 0, 2^k); num *= 10^digits;
// num /= den; decimal output.  Returns the digit count.
u64 nno_e_digits(char* out, u64 cap, int k, u64 digits) {
    Nat num, den;
    calc_e(num, den, 0, 1ull << k);
    num = mul(num, pow(Nat(10), digits));
    Nat q, r;
    divmod(num, den, q, r);
    return nno_to_dec(out, cap, q.data, q.size);
}
// The reference's pi recipe (test.cpp:22-32): calc_pi over [1, 2^k + 1),
// sqrt(10005) to acc words, num = 13591409 den - num, den *= 426880 numc
// 10^digits, num *= denc, den /= num, decimal digits of den.
u64 nno_pi_digits(char* out, u64 cap, int k, u64 digits, u64 acc) {
    Nat num, den, numc, denc;
    calc_pi(num, den, numc, 1, (1ull << k) | 1);
    sqrt_10005(numc, denc, acc);
    num = sub(mul(Nat(0xcf6371), den), num);
    den = mul(den, mul(mul(Nat(0x68380), numc), pow(Nat(10), digits)));
    num = mul(num, denc);
    Nat q, r;
    divmod(den, num, q, r);
    return nno_to_dec(out, cap, q.data, q.size);
}
// calc_pi over [1, 2^k + 1): num, den, numc (natural.h:874-902).
void nno_calc_pi(u64* num, u64* den, u64* numc, u64 cap, int k, u64* sizes) {
    Nat a, b, c;
    calc_pi(a, b, c, 1, (1ull << k) | 1);
    sizes[0] = emit(a, num, cap);
    sizes[1] = emit(b, den, cap);
    sizes[2] = emit(c, numc, cap);
}
// sqrt_10005(num, den, acc) (natural.h:859-872).
void nno_sqrt_10005(u64* num, u64* den, u64 cap, u64 acc, u64* sizes) {
    Nat a, b;
    sqrt_10005(a, b, acc);
    sizes[0] = emit(a, num, cap);
    sizes[1] = emit(b, den, cap);
}
// calc_e numerator/denominator for [0, 2^k); sizes returned through pointers.
void nno_calc_e(u64* num, u64 num_cap, u64* den, u64 den_cap, int k, u64* num_n, u64* den_n) {
    Nat a, b;
    calc_e(a, b, 0, 1ull << k);
    *num_n = emit(a, num, num_cap);
    *den_n = emit(b, den, den_cap);
}

}  // extern "C"

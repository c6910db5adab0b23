// Host-compiled check of the device arithmetic in nnh_arith.hpp (the same
// __host__ __device__ functions the kernels use) against exact 128-bit math.
// Built and run by tests/test_arith.py.  Exit status 0 = all checks pass.
#include "../../natural_number_amd/csrc/nnh_arith.hpp"

#include <cstdio>
#include <random>

using namespace nnh;
typedef unsigned __int128 u128;

static u64 powmod(u64 b, u64 e, u64 p) {
    u64 r = 1;
    for (b %= p; e; e >>= 1, b = (u64)((u128)b * b % p))
        if (e & 1) r = (u64)((u128)r * b % p);
    return r;
}
static u64 to_mont(u64 x, u64 p) { return (u64)(((u128)(x % p) << 64) % p); }

static int check_prime(std::mt19937_64& rng, int iters, int pi) {
    const Prime P = prime_of(pi);
    const u64 p = P.p;
    const u32 K = P.k;
    const u64 rinv = powmod((u64)(((u128)1 << 64) % p), p - 2, p);  // 2^-64 mod p
    int bad = 0;
    auto expect = [&](u128 t) { return (u64)(t % p * rinv % p); };
    for (int it = 0; it < iters; it++) {
        u64 a, b;
        switch (it % 4) {
            case 0: a = rng() % (2 * p); b = rng() % p; break;             // butterfly operands
            case 1: a = rng(); b = rng() % p; break;                       // limb * constant (load)
            case 2: a = 2 * p - 1 - (rng() & 0xff); b = p - 1 - (rng() & 0xff); break;  // bound
            default: a = rng() & 0xffff; b = rng() % p; break;
        }
        u64 r = mont(P, a, b);
        if (r >= 2 * p || r % p != expect((u128)a * b)) bad++;
        u64 x0 = rng() % p, y0 = rng() % p, x1 = rng() % p, y1 = rng() % p;
        u64 r2 = mont2(P, x0, y0, x1, y1);
        if (r2 >= 2 * p || r2 % p != expect((u128)x0 * y0 % p + (u128)x1 * y1 % p)) bad++;
        u64 x = rng() % (2 * p);
        if (csub(P, x) != x % p) bad++;
        // butterflies
        u64 w = rng() % p, wm = to_mont(w, p);
        u64 bx = rng() % (2 * p), by = rng() % (2 * p), cx = bx, cy = by;
        ct_bfly(P, cx, cy, wm);
        u64 t = (u64)((u128)(by % p) * w % p);
        if (cx >= 2 * p || cy >= 2 * p || cx % p != (bx % p + t) % p || cy % p != (bx % p + p - t) % p) bad++;
        u64 gx = rng() % p, gy = rng() % p, ex = gx, ey = gy;
        gs_bfly(P, ex, ey, wm);
        if (ex != (gx + gy) % p || ey != (u64)((u128)((gx + p - gy) % p) * w % p)) bad++;
        if (lazy_prime(pi)) {
            // forward lazy butterfly: inputs and outputs in [0, 4p), boundary inputs included
            u64 lx = it % 5 == 0 ? 4 * p - 1 : rng() % (4 * p), ly = it % 7 == 0 ? 4 * p - 1 : rng() % (4 * p);
            u64 lcx = lx, lcy = ly;
            ct_bfly_lazy(P, lcx, lcy, wm);
            u64 lt = (u64)((u128)(ly % p) * w % p);
            if (lcx >= 4 * p || lcy >= 4 * p || lcx % p != (lx % p + lt) % p || lcy % p != (lx % p + p - lt) % p)
                bad++;
            // inverse lazy butterfly: inputs and outputs in [0, 2p)
            u64 ix = it % 5 == 1 ? 2 * p - 1 : rng() % (2 * p), iy = it % 7 == 1 ? 2 * p - 1 : rng() % (2 * p);
            u64 lex = ix, ley = iy;
            gs_bfly_lazy(P, lex, ley, wm);
            if (lex >= 2 * p || ley >= 2 * p || lex % p != (ix + iy) % p ||
                ley % p != (u64)((u128)((ix % p + p - iy % p) % p) * w % p))
                bad++;
            // base multiply: lazy inputs in [0, 4p) give the canonical result of reduced inputs
            u64 in[8], red[8];
            for (int q = 0; q < 8; q++) {
                in[q] = (it + q) % 9 == 0 ? 4 * p - 1 : rng() % (4 * p);
                red[q] = in[q] % p;
            }
            u64 l0 = in[0], l1 = in[1], l2 = in[2], l3 = in[3];
            u64 s0 = red[0], s1 = red[1], s2 = red[2], s3 = red[3];
            base_mul4<true>(P, l0, l1, l2, l3, in[4], in[5], in[6], in[7], wm);
            base_mul4<false>(P, s0, s1, s2, s3, red[4], red[5], red[6], red[7], wm);
            if (l0 != s0 || l1 != s1 || l2 != s2 || l3 != s3) bad++;
            // twiddle-1 butterflies (block 0) and the 2p conditional subtraction
            u64 ox = lx, oy = ly;
            ct_bfly_lazy_one(P, ox, oy);
            if (ox >= 4 * p || oy >= 4 * p || ox % p != (lx % p + ly % p) % p || oy % p != (lx % p + p - ly % p) % p)
                bad++;
            u64 qx = ix, qy = iy;
            gs_bfly_lazy_one(P, qx, qy);
            if (qx >= 2 * p || qy >= 2 * p || qx % p != (ix % p + iy % p) % p || qy % p != (ix % p + p - iy % p) % p)
                bad++;
            if (csub2(P, lx) >= 2 * p || csub2(P, lx) % p != lx % p) bad++;
        }
        {
            u64 ox = bx, oy = by;
            ct_bfly_one(P, ox, oy);
            if (ox >= 2 * p || oy >= 2 * p || ox % p != (bx % p + by % p) % p || oy % p != (bx % p + p - by % p) % p)
                bad++;
            u64 qx = gx, qy = gy;
            gs_bfly_one(P, qx, qy);
            if (qx != (gx + gy) % p || qy != (gx + p - gy) % p) bad++;
        }
    }
    // reduction edge cases: zero words, all-ones words, top byte patterns of the quotient
    const u64 lows[] = {0, 1, 0xffffffffull, 1ull << 32, 0xff00000000000000ull, ~0ull, 0x00ffffffffffffffull,
                        0x8000000080000000ull};
    for (u64 lo : lows)
        for (u64 hi : {(u64)0, (u64)1, p - 1, p / 2}) {
            u64 r = redc(P, lo, hi);
            if (r == 0 || r >= 2 * p || r % p != expect((u128)hi << 64 | lo)) bad++;
        }
    std::printf("prime k=%u: %d failures in %d iterations\n", K, bad, iters);
    return bad;
}

static int check_garner(std::mt19937_64& rng, int iters) {
    const u64 m0 = prime_of(0).p, m1 = prime_of(1).p, m2 = prime_of(2).p;
    u64 c02m = to_mont(powmod(m0, m2 - 2, m2), m2);
    u64 c1 = powmod((u64)((u128)(m0 % m1) * (m2 % m1) % m1), m1 - 2, m1);
    u64 c1m = to_mont(c1, m1);
    u64 km = to_mont((m1 - (u64)((u128)(m0 % m1) * c1 % m1)) % m1, m1);
    int bad = 0;
    for (int it = 0; it < iters; it++) {
        // random X < m0 m1 m2 built from random residues, or extreme residues
        u64 r0 = rng() % m0, r1 = rng() % m1, r2 = rng() % m2;
        if (it % 8 == 1) { r0 = m0 - 1; r1 = m1 - 1; r2 = m2 - 1; }
        if (it % 8 == 2) { r0 = 0; r1 = 0; r2 = 0; }
        if (it % 8 == 3) { r0 = m0 - 1; r1 = 0; r2 = m2 - 1; }
        u64 x0, x1, x2;
        garner(r0, r1, r2, c02m, c1m, km, x0, x1, x2);
        // check X mod each prime via the 3-word value
        auto modp = [&](u64 p) {
            u128 t = ((u128)(x2 % p) << 64 | x1) % p;
            return (u64)(((t << 64) | x0) % p);
        };
        if (modp(m0) != r0 || modp(m1) != r1 || modp(m2) != r2) bad++;
        // X < m0 m1 m2 < 2^186: x2 < 2^58
        if (x2 >> 58) bad++;
    }
    std::printf("garner: %d failures in %d iterations\n", bad, iters);
    return bad;
}

int main() {
    std::mt19937_64 rng(12345);
    int bad = check_prime(rng, 1000000, 0) + check_prime(rng, 1000000, 1) + check_prime(rng, 1000000, 2) +
              check_garner(rng, 1000000);
    return bad ? 1 : 0;
}

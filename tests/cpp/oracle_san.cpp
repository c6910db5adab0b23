// oracle_san.cpp — TEST INFRASTRUCTURE.  The oracle (oracle/nn_oracle.cpp)
// built for the CPU under the address and undefined-behaviour sanitizers by
// tests/test_oracle_sanitizers.py, exercising every path the GPU tests lean
// on: products at and around every threshold (schoolbook, the transform,
// the one-level split, the cyclic save), squares, shifts, additions,
// subtractions, the Newton division, pow, decimal output and calc_e.  Each
// result is checked by an identity that a dropped term or a wrong index
// breaks (transform product = schoolbook product, (q b + r) = a with r < b,
// a + b - b = a, decimal round trip of a small power).
// Prints "oracle_san ok" or the first failed check; exit 1 on failure.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "nn_oracle.hpp"

using namespace nno;

static uint64_t sm_next(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static Nat rnd(uint64_t seed, uint64_t n) {
    Nat x(0);
    x.resize(n);
    for (uint64_t i = 0; i < n; i++) x[i] = sm_next(seed);
    if (!x[n - 1]) x[n - 1] = 1;
    return x;
}
static int bad = 0;
static void check(bool ok, const std::string& what) {
    if (!ok) {
        std::printf("FAIL %s\n", what.c_str());
        bad++;
    }
}

int main() {
    // products: the transform against schoolbook across the thresholds and the split
    const uint64_t shapes[][2] = {{1, 1}, {383, 383}, {384, 384}, {385, 200}, {500, 384}, {1000, 999},
                                  {3000, 400}, {4096, 4096}, {5000, 700}, {20000, 385}, {9000, 8999}};
    for (auto& sh : shapes) {
        Nat a = rnd(sh[0] * 7 + 1, sh[0]), b = rnd(sh[1] * 11 + 2, sh[1]);
        Nat d1, d2;
        mul_basecase(d1, a, b);
        d1.normalize();
        Nat p = mul(a, b);
        if (std::min(sh[0], sh[1]) >= 2) {
            mul_ntt(d2, a, b);
            d2.normalize();
            check(d1 == d2, "transform product " + std::to_string(sh[0]) + " x " + std::to_string(sh[1]));
        }
        check(p == d1, "dispatched product " + std::to_string(sh[0]) + " x " + std::to_string(sh[1]));
        // additions and subtractions
        check(sub(add(a, b), b) == a, "a + b - b");
        check(shr(shl(a, 131), 131) == a, "shifts");
    }
    for (uint64_t n : {1ull, 383ull, 384ull, 2500ull}) {
        Nat a = rnd(n + 99, n), s = sqr(a);
        check(s == mul(a, a), "square " + std::to_string(n));
    }
    // division: q b + r = a, r < b
    const uint64_t divs[][2] = {{100, 1}, {3000, 1}, {700, 300}, {4000, 1500}, {12000, 5000}, {9000, 4500}};
    for (auto& sh : divs) {
        Nat a = rnd(sh[0] + 5, sh[0]), b = rnd(sh[1] + 6, sh[1]), q, r;
        divmod(a, b, q, r);
        q.normalize();
        r.normalize();
        check(r < b && add(mul(q, b), r) == a, "division " + std::to_string(sh[0]) + " / " + std::to_string(sh[1]));
    }
    // pow and decimal output: 2^65536 has 19729 digits starting 2003529930 (test.cpp:9)
    Nat p = pow(Nat(2), 65536);
    std::string d = to_dec(p);
    check(d.size() == 19729 && d.compare(0, 10, "2003529930") == 0, "2^65536 in decimal");
    check(to_dec(Nat(1234567890123456789ull)) == "1234567890123456789", "one-word decimal");
    // calc_e over [0, 2^10): den = (2^10 - 1)! (natural.h:841-857), num / den ~ e
    Nat num, den;
    calc_e(num, den, 0, 1 << 10);
    Nat q, r;
    divmod(shl(num, 64 * 4), den, q, r);
    q.normalize();
    // e * 2^256 = 0x2b7e151628aed2a6abf7158809cf4f3c762e7160f38b4da56a784d9045190cfef...
    check(q.size == 5 && q[4] == 2 && q[3] == 0xb7e151628aed2a6aull, "calc_e");
    if (bad) {
        std::printf("oracle_san: %d failures\n", bad);
        return 1;
    }
    std::printf("oracle_san ok\n");
    return 0;
}

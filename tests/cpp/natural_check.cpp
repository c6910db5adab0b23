// The C++ class of include/nnh_natural.hpp driven the way the reference's
// test.cpp drives natural (test.cpp:8-40): 2^65536 in decimal, the e recipe,
// a factorial, plus seeded operand pairs.  It prints one "key value" line per
// result; tests/test_gpu_natural_cpp.py checks every line against Python's
// integers and the golden digests.  Test infrastructure, linked to libnnh.so.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>

#include "nnh_natural.hpp"

using nnh::natural;

static uint64_t g_state;
static uint64_t splitmix64() {
    uint64_t z = (g_state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static natural random_natural(uint64_t words) {
    std::vector<uint64_t> v(words);
    for (auto& w : v) w = splitmix64();
    if (words) v.back() |= 1ull << 63;
    return natural(std::move(v));
}

static std::string hex(const natural& x) {
    std::ostringstream os;
    os << std::hex << x;
    return os.str();
}

int main(int argc, char** argv) {
    try {
        // test.cpp:9
        std::cout << "pow2_65536 " << pow(natural(2), 65536) << "\n";

        // Seeded pairs: (a words, b words).  Sizes fall on both sides of the
        // 384-word schoolbook / transform threshold of the product.
        const uint64_t shapes[][2] = {{1, 1}, {3, 2}, {17, 5}, {64, 64}, {300, 40}, {1000, 999}, {4096, 1500},
                                      {20000, 7000}, {70000, 30000}};
        g_state = 12345;
        int i = 0;
        for (const auto& sh : shapes) {
            const natural a = random_natural(sh[0]), b = random_natural(sh[1]);
            natural q, r;
            natural::divmod(a, b, q, r);
            std::cout << "case" << i << "_a " << hex(a) << "\n"
                      << "case" << i << "_b " << hex(b) << "\n"
                      << "case" << i << "_mul " << hex(a * b) << "\n"
                      << "case" << i << "_sqr " << hex(sqr(a)) << "\n"
                      << "case" << i << "_div " << hex(a / b) << "\n"
                      << "case" << i << "_mod " << hex(a % b) << "\n"
                      << "case" << i << "_add " << hex(a + b) << "\n"
                      << "case" << i << "_sub " << hex(a < b ? b - a : a - b) << "\n"
                      << "case" << i << "_cmp " << cmp(a, b) << " " << cmp(b, a) << " " << cmp(a, a) << "\n";
            // q b + r == a with r < b, entirely through the class
            std::cout << "case" << i << "_identity " << ((q * b + r == a) && (r < b) ? 1 : 0) << "\n";
            if (sh[0] <= 1000) std::cout << "case" << i << "_dec " << a * b << "\n";
            i++;
        }
        natural x = random_natural(9);
        x *= x;
        x /= natural(0x1234567);
        x %= random_natural(3);
        std::cout << "compound_ops " << hex(x) << "\n";
        std::ostringstream up;
        up << std::hex << std::uppercase << std::showbase << natural(0xabcdef) * natural(1ull << 40);
        std::cout << "hex_upper " << up.str() << "\n";

        // test.cpp:12-19 at 2^14 terms and 10^4 digits
        {
            natural num, den;
            nnh::calc_e(num, den, 14);
            num *= pow(natural(10), 10000);
            num /= den;
            std::cout << "e_k14_D10000 " << num << "\n";
        }
        // test.cpp:35-40 at n = 10^6: the limbs go to a file
        if (argc > 1) {
            const natural f = natural::factorial(1000000);
            std::ofstream out(argv[1], std::ios::binary);
            out.write(reinterpret_cast<const char*>(f.data()), f.size() * 8);
            std::cout << "factorial_1000000_words " << f.size() << "\n";
        }

        // The reference's defined edge cases: a - b with a < b is |a - b|
        // (natural.h:324-340); x / 0 = 0, 0 / 0 = 1, x % 0 = 0 (natural.h:750-756).
        std::cout << "sub_smaller " << natural(1) - natural(2) << "\n";
        std::cout << "div_by_zero " << natural(5) / natural(0) << " " << natural(0) / natural(0) << " "
                  << natural(5) % natural(0) << "\n";
    } catch (const std::exception& e) {
        std::fprintf(stderr, "natural_check: %s\n", e.what());
        return 1;
    }
    return 0;
}

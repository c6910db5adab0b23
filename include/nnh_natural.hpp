/*
 * nnh_natural.hpp — a C++ natural-number class over the C ABI of libnnh.so,
 * with the reference natural's interface on the multiply path:
 *   operator* / operator*=      natural.h:537-552 (-> nnh_mul)
 *   sqr, pow                    natural.h:553-564 (-> nnh_sqr, nnh_pow)
 *   operator/ % /= %=           natural.h:749-793 (-> nnh_divmod)
 *   operator<< (dec / hex)      natural.h:1020-1039 (decimal -> nnh_to_dec)
 *   factorial                   natural.h:795-839 (-> nnh_factorial)
 *   calc_e, calc_pi             natural.h:841-857, 874-902 (-> nnh_calc_e, nnh_calc_pi)
 * Products, powers, division and decimal output run on the GPU; the O(n)
 * additions, subtraction and comparisons here run on the host, as the
 * reference's own base.h/asm loops do.  Limbs are little-endian uint64_t
 * words, normalised (top word nonzero unless the value is 0), as in
 * basic_natural (basic_natural.h:11-13, natural.h:82-93).
 *
 * Errors: the reference never reports one on this path (it exits on
 * allocation failure, basic_natural.h:18; division by zero and a - b with
 * a < b have defined results, kept here); a non-zero status of the library
 * throws std::runtime_error carrying nnh_last_error().  There is no CPU
 * fallback: without a usable GPU every device operation throws.
 */
#ifndef NNH_NATURAL_HPP
#define NNH_NATURAL_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <ios>
#include <ostream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "nnh.h"

namespace nnh {

class natural {
   public:
    natural(uint64_t w = 0) : d_{w} {}
    explicit natural(std::vector<uint64_t> limbs) : d_(std::move(limbs)) { trim(); }

    const uint64_t* data() const { return d_.data(); }
    uint64_t size() const { return d_.size(); }
    const std::vector<uint64_t>& limbs() const { return d_; }
    bool is_zero() const { return d_.size() == 1 && d_[0] == 0; }
    uint64_t bits() const {  // bit length, at least 1
        const uint64_t top = d_.back();
        return top ? 64 * (d_.size() - 1) + 64 - __builtin_clzll(top) : 1;
    }

    // ---- device operations ---------------------------------------------
    friend natural operator*(const natural& a, const natural& b) {
        ready();
        std::vector<uint64_t> r(a.size() + b.size());
        check(nnh_mul(r.data(), a.data(), a.size(), b.data(), b.size()));
        return natural(std::move(r));
    }
    natural& operator*=(const natural& b) { return *this = *this * b; }
    friend natural sqr(const natural& a) {
        ready();
        std::vector<uint64_t> r(2 * a.size());
        check(nnh_sqr(r.data(), a.data(), a.size()));
        return natural(std::move(r));
    }
    friend natural pow(const natural& b, uint64_t e) {
        ready();
        if (e == 0) return natural(1);
        if (e > UINT64_MAX / b.bits()) throw std::length_error("natural: pow result too large");
        const uint64_t cap = (b.bits() * e) / 64 + 2;
        std::vector<uint64_t> r(cap);
        uint64_t n = 0;
        check(nnh_pow(r.data(), cap, &n, b.data(), b.size(), e));
        r.resize(n);
        return natural(std::move(r));
    }
    // q = a / b, r = a % b (the reference's Newton division).  As in
    // natural::_div (natural.h:750-756), b = 0 gives q = (a == 0), r = 0.
    static void divmod(const natural& a, const natural& b, natural& q, natural& r) {
        ready();
        const uint64_t qc = a.size() + 1, rc = b.size() + 1;
        std::vector<uint64_t> qv(qc), rv(rc);
        uint64_t qn = 0, rn = 0;
        check(nnh_divmod(qv.data(), qc, &qn, rv.data(), rc, &rn, a.data(), a.size(), b.data(), b.size()));
        qv.resize(qn);
        rv.resize(rn);
        q = natural(std::move(qv));
        r = natural(std::move(rv));
    }
    friend natural operator/(const natural& a, const natural& b) {
        natural q, r;
        divmod(a, b, q, r);
        return q;
    }
    friend natural operator%(const natural& a, const natural& b) {
        natural q, r;
        divmod(a, b, q, r);
        return r;
    }
    natural& operator/=(const natural& b) { return *this = *this / b; }
    natural& operator%=(const natural& b) { return *this = *this % b; }
    std::string to_string() const {
        ready();
        const uint64_t cap = 20 * size() + 64;  // 64 log10(2) < 20 digits per word
        std::string s(cap, '\0');
        uint64_t len = 0;
        check(nnh_to_dec(s.data(), cap, &len, data(), size()));
        s.resize(len);
        return s;
    }
    // Hexadecimal digits (host side): the top word unpadded, every lower
    // word as 16 digits, as natural.h:1023-1032 prints them.
    std::string to_hex_string(bool upper = false) const {
        std::string s;
        char w[17];
        for (uint64_t i = size(); i-- > 0;) {
            std::snprintf(w, sizeof w, i + 1 == size() ? (upper ? "%llX" : "%llx") : (upper ? "%016llX" : "%016llx"),
                          (unsigned long long)d_[i]);
            s += w;
        }
        return s;
    }
    // std::hex selects hexadecimal (std::showbase adds 0x / 0X), otherwise
    // decimal, as operator<< of natural.h:1020-1039.
    friend std::ostream& operator<<(std::ostream& os, const natural& x) {
        const auto f = os.flags();
        if (f & std::ios::hex) {
            const bool upper = f & std::ios::uppercase;
            if (f & std::ios::showbase) os << (upper ? "0X" : "0x");
            return os << x.to_hex_string(upper);
        }
        return os << x.to_string();
    }
    static natural factorial(uint64_t n) {
        ready();
        const double bits = (double)n * std::max(1.0, std::log2((double)std::max<uint64_t>(n, 2))) + 64;
        const uint64_t cap = (uint64_t)(bits / 64) + 2;
        std::vector<uint64_t> r(cap);
        uint64_t m = 0;
        check(nnh_factorial(r.data(), cap, &m, n));
        r.resize(m);
        return natural(std::move(r));
    }

    // ---- host operations (O(n) word loops) ------------------------------
    friend natural operator+(const natural& a, const natural& b) {
        const natural& x = a.size() >= b.size() ? a : b;
        const natural& y = a.size() >= b.size() ? b : a;
        std::vector<uint64_t> r(x.size() + 1);
        unsigned carry = 0;
        for (uint64_t i = 0; i < x.size(); i++) {
            const uint64_t yi = i < y.size() ? y.d_[i] : 0;
            const uint64_t s = x.d_[i] + yi;
            const uint64_t t = s + carry;
            carry = (s < yi) | (t < s);
            r[i] = t;
        }
        r[x.size()] = carry;
        return natural(std::move(r));
    }
    natural& operator+=(const natural& b) { return *this = *this + b; }
    // |a - b|: the reference's natural has no negative values and its
    // subtraction takes the larger operand first (natural::_sub,
    // natural.h:324-340).
    friend natural operator-(const natural& a, const natural& b) {
        const natural& x = cmp(a, b) >= 0 ? a : b;
        const natural& y = cmp(a, b) >= 0 ? b : a;
        std::vector<uint64_t> r(x.size());
        unsigned borrow = 0;
        for (uint64_t i = 0; i < x.size(); i++) {
            const uint64_t yi = i < y.size() ? y.d_[i] : 0;
            const uint64_t d = x.d_[i] - yi;
            r[i] = d - borrow;
            borrow = (x.d_[i] < yi) | (d < borrow);
        }
        return natural(std::move(r));
    }
    natural& operator-=(const natural& b) { return *this = *this - b; }
    friend int cmp(const natural& a, const natural& b) {
        if (a.size() != b.size()) return a.size() < b.size() ? -1 : 1;
        for (uint64_t i = a.size(); i-- > 0;)
            if (a.d_[i] != b.d_[i]) return a.d_[i] < b.d_[i] ? -1 : 1;
        return 0;
    }
    friend bool operator==(const natural& a, const natural& b) { return cmp(a, b) == 0; }
    friend bool operator!=(const natural& a, const natural& b) { return cmp(a, b) != 0; }
    friend bool operator<(const natural& a, const natural& b) { return cmp(a, b) < 0; }

    // Library plumbing shared with calc_e / calc_pi below.
    static void check(int rc) {
        if (rc != NNH_OK) throw std::runtime_error(std::string("libnnh: ") + nnh_last_error());
    }
    static void ready() {
        static const int rc = nnh_init(0);
        check(rc);
    }

   private:
    std::vector<uint64_t> d_;
    void trim() {
        while (d_.size() > 1 && d_.back() == 0) d_.pop_back();
        if (d_.empty()) d_.push_back(0);
    }
};

// calc_e(num, den, 0, 2^k) (natural.h:841-857): e ~ num / den.
inline void calc_e(natural& num, natural& den, int k) {
    natural::ready();
    const uint64_t cap = k <= 24 ? (uint64_t(24) << k) / 64 + 64 : (uint64_t(k) << k) / 64 + 64;
    std::vector<uint64_t> nv(cap), dv(cap);
    uint64_t nn = 0, dn = 0;
    natural::check(nnh_calc_e(nv.data(), cap, &nn, dv.data(), cap, &dn, k));
    nv.resize(nn);
    dv.resize(dn);
    num = natural(std::move(nv));
    den = natural(std::move(dv));
}

// calc_pi(num, den, numc, 1, 2^k + 1) (natural.h:874-902).
inline void calc_pi(natural& num, natural& den, natural& numc, int k) {
    natural::ready();
    const uint64_t cap = (uint64_t(340) << k) / 64 + 64;
    std::vector<uint64_t> v[3] = {std::vector<uint64_t>(cap), std::vector<uint64_t>(cap), std::vector<uint64_t>(cap)};
    uint64_t sizes[3] = {0, 0, 0};
    natural::check(nnh_calc_pi(v[0].data(), v[1].data(), v[2].data(), cap, k, sizes));
    natural* out[3] = {&num, &den, &numc};
    for (int i = 0; i < 3; i++) {
        v[i].resize(sizes[i]);
        *out[i] = natural(std::move(v[i]));
    }
}

}  // namespace nnh

#endif  // NNH_NATURAL_HPP

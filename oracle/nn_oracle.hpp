// nn_oracle.hpp — CPU oracle for the three-prime NTT multiply.
//
// TEST INFRASTRUCTURE ONLY.  This is a portable C++20 restatement of the
// reference library's CPU multiply path (Natural_Number: ntt_info.h, base.h,
// basic_natural.h, natural.h and the x86-64 routines in asmLinux.asm).  It is
// the correctness checker for the HIP product path and the timed CPU baseline.
// Only the test suite, the smoke check and the benchmark's CPU-baseline leg
// may load it; the shipped library never links or calls it.
//
// Every function cites the reference file:line it restates.  The restatement
// keeps the reference's arithmetic (Shoup twiddles, Barrett constants, lazy
// [0,2p) ranges, the Garner sequence of asmCRT) so that intermediate transform
// words are comparable, not just final products.
#pragma once
#include <cstdint>
#include <cstddef>
#include <string>

namespace nno {

using u64 = std::uint64_t;
using u128 = unsigned __int128;

// ---- limb container: basic_natural.h:9-97 --------------------------------
// malloc-owned little-endian limbs, power-of-two capacity, shrink keeps data.
struct Limbs {
    u64* data;
    u64 size;
    u64 capacity;
    Limbs();
    Limbs(const Limbs&);
    Limbs(Limbs&&) noexcept;
    ~Limbs();
    Limbs& operator=(const Limbs&);
    Limbs& operator=(Limbs&&) noexcept;
    void resize(u64 s);
    void reserve(u64 s);  // grow capacity only (used where the reference writes past size)
    u64& operator[](u64 i) { return data[i]; }
    const u64& operator[](u64 i) const { return data[i]; }
};

// ---- NTT engine: ntt_info.h:10-67 ------------------------------------------
struct NttInfo {
    static const u64 P[3];          // ntt_info.h:69-73
    static const u64 R124[3];       // ntt_info.h:75-79, ceil(2^124/p)
    static u64 RB[3][56];           // ntt_info.h:81-109 (derived, see root_bases())
    static u64 IRB[3][56];          // ntt_info.h:111-139
    static Limbs root[3], iroot[3]; // ntt_info.h:141-143, (w, shoup) pairs
    static int root_scale;          // ntt_info.h:145
    Limbs v[3];
    int scale = 0;

    static void init(int scale);                        // ntt_info.h:22-41
    void load(const u64* src, u64 size, int scale);     // ntt_info.h:147-173
    void forward();                                     // ntt_info.h:179-198
    void inverse();                                     // ntt_info.h:200-221
    void mul(const NttInfo& a, const NttInfo& b);       // ntt_info.h:223-231
    u64 crt(u64 n);                                     // ntt_info.h:233-235
    void save(Limbs& out, u64 size);                    // ntt_info.h:237-248
};

// ---- natural: natural.h:16-80 ----------------------------------------------
struct Nat : Limbs {
    Nat(u64 x = 0);
    Nat(const u64* w, u64 n);
    Nat& normalize();  // natural.h:82-93 (std)
};

bool operator<(const Nat& a, const Nat& b);
bool operator==(const Nat& a, const Nat& b);
inline bool operator!=(const Nat& a, const Nat& b) { return !(a == b); }
inline bool operator>=(const Nat& a, const Nat& b) { return !(a < b); }

Nat shl(const Nat& a, u64 bits);
Nat shr(const Nat& a, u64 bits);
Nat add(const Nat& a, const Nat& b);
Nat sub(const Nat& a, const Nat& b);
Nat mul(const Nat& a, const Nat& b);  // natural::_mul, natural.h:444-450
Nat sqr(const Nat& a);                // natural::_sqr, natural.h:529-535
Nat pow(const Nat& b, u64 e);         // natural.h:554-564
void divmod(const Nat& a, const Nat& b, Nat& q, Nat& r);  // natural.h:749-769
void calc_e(Nat& num, Nat& den, u64 begin, u64 end);      // natural.h:841-857
std::string to_dec(const Nat& n);                          // natural.h:1020-1039

// algorithm arms, exposed for tests
void mul_ntt(Nat& d, const Nat& a, const Nat& b);   // natural.h:391-431
void sqr_ntt(Nat& d, const Nat& a);                 // natural.h:433-442
void mul_basecase(Nat& d, const Nat& a, const Nat& b);  // natural.h:356-369
int ntt_scale(u64 s1, u64 s2);                      // natural.h:384-386

}  // namespace nno

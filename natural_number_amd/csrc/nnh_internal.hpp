// nnh_internal.hpp — descriptors shared by the kernels and the host driver.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "nnh_jobs.hpp"

namespace nnh {

// Root bases in Montgomery form (x * 2^64 mod p), by value in kernel args.
struct RootBases {
    u64 rb_m[3][56];
    u64 irb_m[3][56];
    u64 one_m[3];
};

// Garner constants (Montgomery form where they multiply residues).
struct CrtConsts {
    u64 c02m;   // m0^-1 mod m2, times 2^64 mod m2
    u64 c1m;    // (m0 m2)^-1 mod m1, times 2^64 mod m1
    u64 km;     // -m0 (m0 m2)^-1 mod m1, times 2^64 mod m1
};

struct TopPassDesc {
    u64* data[3];
    const u64* tw[3];
    const u64* limbs;   // non-null: first forward pass loads from limbs
    u64 an;
    u64 alpha_m[3];
    u64 L;
    u32 g_lo, nlev;
    bool inverse;
    // Optional second operand of a forward pass, transformed by the same
    // launch (grid z = 2): its arrays, and its limbs when limbs is set.
    bool two;
    u64* data2[3];
    const u64* limbs2;
    u64 an2;
};

struct MidDesc {
    u64* fa[3];
    const u64* fb[3];
    const u64* twf[3];
    const u64* twi[3];
    const u64* la;      // non-null: load limbs (L == 2^cb)
    const u64* lb;
    u64 an, bn;
    u64 alpha_m[3];
    u64 L;
    int cb;
    bool sqr;
    bool fwd_only;   // store A's finished transform (no product)
    int a_done;      // A already fully transformed
    int b_done;
};

struct CrtDesc {
    const u64* r[3];
    int columns;
    u64 n_res;
    u64* out;
    u64 n_out;
    const u64* addend;
    u64 n_add;
    u8* tile_fn;
    u8* tile_cin;
    CrtConsts C;
};

struct TopBatchParams {
    const u64* tw[3];
    u64 alpha_m[3];
    u64 L;
    u32 g_lo, nlev;
};

struct MidBatchParams {
    const u64* twf[3];
    const u64* twi[3];
    u64 alpha_m[3];
    u64 L;
};

// (The p0/p1 kernels go to st, the p2 kernels to st2, st when null.)
void launch_top_pass_batch(const BJob* jobs, u32 njobs, int operand, const TopBatchParams& P, bool inverse,
                           bool from_limbs, hipStream_t st, hipStream_t st2 = nullptr);
// fwd_only: the jobs are forward-only entries that store shared B transforms.
void launch_mid_batch(const BJob* jobs, u32 njobs, const MidBatchParams& P, int cb, bool from_limbs, bool fwd_only,
                      hipStream_t st, hipStream_t st2 = nullptr);
void launch_school_batch(const BJob* jobs, const u32* sj, u32 nsj, u64 ncols, hipStream_t st);
void launch_crt_batch(const BJob* jobs, u32 njobs, u64 ntiles, u8* tile_fn, u8* tile_cin, const CrtConsts& C,
                      hipStream_t st);

void launch_twiddles(u64* tw, u64 n, const RootBases& rb, int pr, int inv, hipStream_t st);
// The p0/p1 kernels go to st, the p2 kernels to st2 (st when null).
void launch_top_pass(const TopPassDesc& d, hipStream_t st, hipStream_t st2 = nullptr);
void launch_mid(const MidDesc& d, hipStream_t st, hipStream_t st2 = nullptr);
void launch_crt(const CrtDesc& d, hipStream_t st);
void launch_school_cols(const u64* a, u64 an, const u64* b, u64 bn, u64* c0, u64* c1, u64* c2, hipStream_t st);
void launch_cyclic_fold(u64* w, u64 L, hipStream_t st);
// One Karatsuba level over schoolbook columns (an >= bn >= 2): writes the
// three column planes of a*b (an+bn-1 columns each) into cols; needs
// kara_ws_words(an, bn) words of workspace.
void launch_kara(const u64* a, u64 an, const u64* b, u64 bn, u64* ws, u64* cols, hipStream_t st);
u64 kara_ws_words(u64 an, u64 bn);

}  // namespace nnh

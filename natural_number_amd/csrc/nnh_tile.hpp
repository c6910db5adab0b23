// nnh_tile.hpp — the tile engine shared by the transform kernels.
//
// Transform layout (per prime, L = 2^s words, reference ntt_info.h:147-248):
// natural-order input, Cooley-Tukey (DIF) levels h = L/2 .. 2 whose block b
// uses twiddle w_b = prod_{bit q of b} rb[q+1] (ntt_info.h:22-41), leaving
// the incomplete block-bit-reversed transform: group g of 4 words holds
// a mod (X^2 - w_g) and a mod (X^2 + w_g).  The base multiply, the
// Gentleman-Sande inverse and a fused Garner CRT with a carry-lookahead scan
// recover the exact product.
//
// All butterflies run on a "tile" of 2^CB elements held as 16 registers in
// each of 2^(CB-4) threads.  A tile layout is a 4-bit register window
// [lo, lo+4) of the local element index y:
//     y(t, j) = (t mod 2^lo) | (j << lo) | ((t >> lo) << (lo + 4)),
// so every level whose index bit lies in the window is a register-to-register
// butterfly; moving the window is one LDS transpose (padded by one word per
// 16 so that 16-lane quarter-waves hit distinct banks).
//
// For element y at level (local bit l) the block index is
//     b = tile_id * 2^(CB-l-1) + (y >> (l+1)),
// the same formula for the fused chunk kernel (global index = tile_id*2^CB+y)
// and for the strided top passes (tile = columns x rows of the big array).
#pragma once
#include <type_traits>

#include "nnh_arith.hpp"
#include "nnh_internal.hpp"

namespace nnh {

template <int B, int E>
struct CFor {
    template <class F>
    __device__ __forceinline__ static void run(F&& f) {
        if constexpr (B < E) {
            f(std::integral_constant<int, B>{});
            CFor<B + 1, E>::run(f);
        }
    }
};

__host__ __device__ constexpr u32 lay(u32 t, u32 j, u32 lo) {
    return (t & ((1u << lo) - 1u)) | (j << lo) | ((t >> lo) << (lo + 4));
}
__host__ __device__ constexpr u32 pad(u32 y) { return y + (y >> 4); }
// lay(t, j, lo) = T(t, lo) | (j << lo) with disjoint bits, and
// pad(T | J) = pad(T) + pad(J): a per-thread base plus a constant per j.
__device__ __forceinline__ u32 lay_t(u32 t, u32 lo) { return (t & ((1u << lo) - 1u)) | ((t >> lo) << (lo + 4)); }
template <int J, int LO>
__host__ __device__ constexpr u32 padj() { return pad((u32)J << LO); }

template <int CB, int LF, int LT>
__device__ __forceinline__ void xchg(u64 (&v)[16], u64* lds, u32 t) {
    const u32 bf = pad(lay_t(t, LF)), bt = pad(lay_t(t, LT));
    __syncthreads();
    CFor<0, 16>::run([&](auto jc) { lds[bf + padj<decltype(jc)::value, LF>()] = v[decltype(jc)::value]; });
    __syncthreads();
    CFor<0, 16>::run([&](auto jc) { v[decltype(jc)::value] = lds[bt + padj<decltype(jc)::value, LT>()]; });
}
// The same transpose through half the LDS: low words, then high words.
// Kernels whose occupancy the 64-bit buffer limits (the top passes) use it.
template <int CB, int LF, int LT>
__device__ __forceinline__ void xchg(u64 (&v)[16], u32* lds, u32 t) {
    const u32 bf = pad(lay_t(t, LF)), bt = pad(lay_t(t, LT));
    u32 lo[16];
    __syncthreads();
    CFor<0, 16>::run([&](auto jc) { lds[bf + padj<decltype(jc)::value, LF>()] = lo32(v[decltype(jc)::value]); });
    __syncthreads();
    CFor<0, 16>::run([&](auto jc) { lo[decltype(jc)::value] = lds[bt + padj<decltype(jc)::value, LT>()]; });
    __syncthreads();
    CFor<0, 16>::run([&](auto jc) { lds[bf + padj<decltype(jc)::value, LF>()] = hi32(v[decltype(jc)::value]); });
    __syncthreads();
    CFor<0, 16>::run([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        v[j] = mk64(lo[j], lds[bt + padj<j, LT>()]);
    });
}

// One level at local bit L with the register window at LO (L-LO in 0..3).
// Block index of element y: b = tile_id 2^(CB-L-1) + (y >> (L+1)); for the
// thread's elements this is a per-thread base plus j >> (L+1-LO).
// ONE (the pass holds the transform's top level, so tile_id is 0 in every
// work-group): butterflies in block 0 (twiddle 1) skip the product.  In the
// top window (LO = CB-4, every thread's t >> LO = 0) their block is 0 when
// j >> (L+1-LO) = 0, known at compile time.
template <bool INV, int CB, int L, int LO, bool LAZY, bool ONE = false>
__device__ __forceinline__ void level_ct(const Prime& P, u64 (&v)[16], u32 t, u64 tile_id, const u64* __restrict__ tw) {
    static_assert(L >= LO && L <= LO + 3, "level outside the register window");
    constexpr int D = 1 << (L - LO);
    const u64* twb = tw + ((tile_id << (CB - 1 - L)) + ((u64)(t >> LO) << (LO + 3 - L)));
    CFor<0, 16>::run([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        if constexpr ((j & D) == 0) {
            if constexpr (ONE && LO == CB - 4 && (j >> (L + 1 - LO)) == 0) {
                if constexpr (INV && LAZY) gs_bfly_lazy_one(P, v[j], v[j + D]);
                else if constexpr (INV) gs_bfly_one(P, v[j], v[j + D]);
                else if constexpr (LAZY) ct_bfly_lazy_one(P, v[j], v[j + D]);
                else ct_bfly_one(P, v[j], v[j + D]);
            } else {
                const u64 w = twb[j >> (L + 1 - LO)];
                if constexpr (INV && LAZY) gs_bfly_lazy(P, v[j], v[j + D], w);
                else if constexpr (INV) gs_bfly(P, v[j], v[j + D], w);
                else if constexpr (LAZY) ct_bfly_lazy(P, v[j], v[j + D], w);
                else ct_bfly(P, v[j], v[j + D], w);
            }
        }
    });
}

// Window sequence of the forward levels L_HI..L_LO starting at window LO0,
// never below LO_MIN: a level below the window moves it to max(l-3, LO_MIN).
__host__ __device__ constexpr int fwd_window(int l, int l_hi, int lo0, int lo_min) {
    int lo = lo0;
    for (int x = l_hi; x >= l; x--)
        if (x < lo) lo = x - 3 < lo_min ? lo_min : x - 3;
    return lo;
}
// Inverse levels ascending: a level above the window moves it to min(l, CB-4).
__host__ __device__ constexpr int inv_window(int l, int l_lo, int lo0, int cb) {
    int lo = lo0;
    for (int x = l_lo; x <= l; x++)
        if (x > lo + 3) lo = x > cb - 4 ? cb - 4 : x;
    return lo;
}

template <int CB, int L, int L_LO, int L_HI, int LO0, int LO_MIN, bool LAZY, class LT, bool ONE = false>
__device__ __forceinline__ void fwd_rec(const Prime& P, u64 (&v)[16], LT* lds, u32 t, u64 tile_id, const u64* __restrict__ tw) {
    if constexpr (L >= L_LO) {
        constexpr int prev = L == L_HI ? LO0 : fwd_window(L + 1, L_HI, LO0, LO_MIN);
        constexpr int cur = fwd_window(L, L_HI, LO0, LO_MIN);
        if constexpr (cur != prev) xchg<CB, prev, cur>(v, lds, t);
        level_ct<false, CB, L, cur, LAZY, ONE>(P, v, t, tile_id, tw);
        fwd_rec<CB, L - 1, L_LO, L_HI, LO0, LO_MIN, LAZY, LT, ONE>(P, v, lds, t, tile_id, tw);
    }
}
template <int CB, int L, int L_LO, int L_HI, int LO0, bool LAZY, class LT, bool ONE = false>
__device__ __forceinline__ void inv_rec(const Prime& P, u64 (&v)[16], LT* lds, u32 t, u64 tile_id, const u64* __restrict__ tw) {
    if constexpr (L <= L_HI) {
        constexpr int prev = L == L_LO ? LO0 : inv_window(L - 1, L_LO, LO0, CB);
        constexpr int cur = inv_window(L, L_LO, LO0, CB);
        if constexpr (cur != prev) xchg<CB, prev, cur>(v, lds, t);
        level_ct<true, CB, L, cur, LAZY, ONE>(P, v, t, tile_id, tw);
        inv_rec<CB, L + 1, L_LO, L_HI, LO0, LAZY, LT, ONE>(P, v, lds, t, tile_id, tw);
    }
}

// Element i of a three-entry (per prime) array without dynamic indexing,
// which would force the argument struct into scratch memory.
template <class T>
__device__ __forceinline__ T sel3(const T (&a)[3], u32 i) {
    return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]);
}

// ---------------------------------------------------------------------------
// Operand load: residue of limb x scaled by alpha (Montgomery form alpha*R),
// zero beyond the operand.  Reference ntt_info.h:147-173 reduces by Barrett;
// here the scale alpha^2 = 2^64 * 2^-(s-1) absorbs both the base-multiply
// Montgomery factor and the inverse transform's 2^(s-1).
// ---------------------------------------------------------------------------
// Limbs of a thread's first N elements (element j at index xs(j)) as
// residues: all loads first, then the reductions, so that the loads issue
// back to back.  The index is clamped into the operand and the limb masked
// to zero beyond it (the reduction of 0 is p, which is 0 mod p).  an >= 1.
template <int N, class XS>
__device__ __forceinline__ void load_limbs(const Prime& P, u64 (&v)[16], const u64* __restrict__ a, u64 an, u64 alpha_m,
                                           XS&& xs) {
    u64 w[N];
    CFor<0, N>::run([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const u64 x = xs(j);
        w[j] = ld_g(&a[x < an ? x : an - 1]);
    });
    CFor<0, N>::run([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        v[j] = mont(P, xs(j) < an ? w[j] : 0, alpha_m);
    });
}

struct TopArgs {
    u64* data[3];          // per-prime arrays of L words (in/out)
    const u64* tw[3];      // per-prime twiddle tables (forward or inverse)
    const u64* limbs;      // operand limbs when loading (first forward pass)
    u64 an;
    u64 alpha_m[3];
    u32 g_lo;              // global bit of the pass's lowest level
    u32 nlev;              // levels in this pass (1..8)
};

// Top pass over a 4096-element tile of Q = 2^(12-NLEV) consecutive columns
// by R = 2^NLEV rows of stride d = 2^g_lo (local index y = col + Q row).
// TOP: the pass holds the transform's top level (g_lo + NLEV = s), so every
// tile lies in block 0 of its top level.  The limb-loading pass always does.
template <bool INV, bool FROM_LIMBS, int NLEV, bool LAZY, bool TOP, class LT>
__device__ __forceinline__ void top_pass_body(const TopArgs& A, u32 pr, u64 tile, LT* lds) {
    static_assert(INV || TOP == FROM_LIMBS, "the first forward pass loads the limbs");
    const Prime P = make_prime(PRIME_K[0] * (pr == 0) + PRIME_K[1] * (pr == 1) + PRIME_K[2] * (pr == 2));
    constexpr int CB = 12, Q = 12 - NLEV;
    constexpr u32 QM = (1u << Q) - 1u;
    constexpr int LO_IO = CB - 4;                           // coalesced window for loads/stores
    constexpr int LO_INV0 = Q < CB - 4 ? Q : CB - 4;        // inverse starts at the bottom row bit
    constexpr int LO_END = INV ? inv_window(CB - 1, Q, LO_INV0, CB) : fwd_window(Q, CB - 1, LO_IO, Q);
    constexpr int LO_ST = LO_END < Q ? LO_IO : LO_END;      // store window (columns in lane bits)
    // The inverse loads straight into its first window: lanes still cover
    // 2^min(Q, 8) consecutive columns of a row, as in LO_IO.
    constexpr int LO_LD = INV ? LO_INV0 : LO_IO;
    const u32 t = threadIdx.x;
    const u64 d = 1ull << A.g_lo;
    const u64 tiles_per_blk = d >> Q;
    const u64 blk = tile / tiles_per_blk;
    const u64 cbase = (tile % tiles_per_blk) << Q;
    const u64 xbase = (blk << (A.g_lo + NLEV)) + cbase;
    u64* __restrict__ data = sel3(A.data, pr);
    const u64* __restrict__ tw = sel3(A.tw, pr);
    const u64 alpha = sel3(A.alpha_m, pr);
    u64 v[16];
    // The limb-loading pass is the first one: its top level (local bit CB-1)
    // pairs x with x + L/2 in block 0, whose twiddle is 1.  When the operand
    // fits in the lower half the partner is zero and the level is a copy.
    // (Uniform per work-group.)
    const bool top_copy = FROM_LIMBS && !INV && blk == 0 && (d << (NLEV - 1)) >= A.an;
    // Element y of the tile sits at xbase + x0(y), x0(y) = (y & QM) + (y >> Q) d,
    // and x0 adds over disjoint bit sets: for y = yt | (j << lo) that is a
    // per-thread base plus an offset that is uniform per register j (one
    // 64-bit add per access instead of a shift and adds per element).
    auto x0 = [&](u32 y) -> u64 { return (u64)(y & QM) + ((u64)(y >> Q) << A.g_lo); };
    {
        const u32 yt = lay_t(t, LO_LD);
        const u64 xt = xbase + x0(yt);
        if constexpr (FROM_LIMBS) {
            // (the limb loads keep the per-element index: measured faster
            // there than the base-plus-offset form)
            auto xs = [&](int j) {
                const u32 y = yt | ((u32)j << LO_LD);
                return xbase + (y & QM) + (u64)(y >> Q) * d;
            };
            if (A.an == 0) CFor<0, 16>::run([&](auto jc) { v[decltype(jc)::value] = 0; });
            else if (top_copy) load_limbs<8>(P, v, A.limbs, A.an, alpha, xs);
            else load_limbs<16>(P, v, A.limbs, A.an, alpha, xs);
        } else {
            const u64* const dt = data + xt;
            CFor<0, 16>::run([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                v[j] = ld_stream(&dt[x0((u32)j << LO_LD)]);
            });
        }
    }
    if constexpr (!INV) {
        static_assert(LO_IO == CB - 4, "the top level's partner is register j + 8");
        if (top_copy) {
            CFor<0, 8>::run([&](auto jc) { v[decltype(jc)::value + 8] = v[decltype(jc)::value]; });
            fwd_rec<CB, CB - 2, Q, CB - 1, LO_IO, Q, LAZY, LT, TOP>(P, v, lds, t, blk, tw);
        } else {
            fwd_rec<CB, CB - 1, Q, CB - 1, LO_IO, Q, LAZY, LT, TOP>(P, v, lds, t, blk, tw);
        }
    } else {
        inv_rec<CB, Q, Q, CB - 1, LO_INV0, LAZY, LT, TOP>(P, v, lds, t, blk, tw);
    }
    if constexpr (LO_ST != LO_END) xchg<CB, LO_END, LO_ST>(v, lds, t);
    {
        u64* const dt = data + xbase + x0(lay_t(t, LO_ST));
        CFor<0, 16>::run([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            st_stream(&dt[x0((u32)j << LO_ST)], v[j]);
        });
    }
}

// ---------------------------------------------------------------------------
// Fused chunk kernel: for one 2^CB chunk, the remaining forward levels
// (local bits CB-1..1) of A and B, the base multiply, and the inverse levels
// (1..CB-1).  With FROM_LIMBS (L == 2^CB) it also loads the operands.
// FWD_ONLY stores A's finished transform instead (for transform reuse);
// a_done / b_done mark operands whose transform is already complete.
// ---------------------------------------------------------------------------
struct MidArgs {
    u64* fa[3];            // A transform (in), product or transform (out)
    const u64* fb[3];      // B transform (in); == fa for squaring
    const u64* twf[3];
    const u64* twi[3];
    const u64* la;         // limbs when FROM_LIMBS
    const u64* lb;
    u64 an, bn;
    u64 alpha_m[3];
    int a_done, b_done;
};

// SEL: pick the per-prime pointers with selects (an argument struct built
// in registers, batched kernels) instead of indexing the kernel-argument
// struct directly (single-product kernel, where that keeps fewer registers).
// slot: the entry of A's per-prime arrays to use when SEL (pr if omitted; a
// caller that stores one prime's pointers in every entry passes 0, which
// keeps the selection out of a run-time-indexed stack array).
template <int CB, bool FROM_LIMBS, bool SQR, bool FWD_ONLY, bool SEL, bool LAZY>
__device__ __forceinline__ void mid_body(const MidArgs& A, u32 pr, u64 chunk, u64* lds, bool sqr_rt = false,
                                         u32 slot = ~0u) {
    const Prime P = make_prime(PRIME_K[0] * (pr == 0) + PRIME_K[1] * (pr == 1) + PRIME_K[2] * (pr == 2));
    constexpr int LO_IO = CB - 4;
    constexpr int LO_F = fwd_window(1, CB - 1, LO_IO, 0);    // window after the forward levels
    const u32 t = threadIdx.x;
    const u64 base = chunk << CB;
    const u32 yt = lay_t(t, LO_IO);
    // SEL: selected once; otherwise indexed at each use, which lets the
    // compiler rematerialise the kernel-argument loads
    const u32 sl = slot == ~0u ? pr : slot;
    u64* const fa_s = SEL ? sel3(A.fa, sl) : nullptr;
    const u64* const fb_s = SEL ? sel3(A.fb, sl) : nullptr;
    const u64* const twf_s = SEL ? sel3(A.twf, sl) : nullptr;
    const u64* const twi_s = SEL ? sel3(A.twi, sl) : nullptr;
    const u64 alpha_s = SEL ? sel3(A.alpha_m, sl) : 0;
    auto fa_ = [&]() -> u64* { if constexpr (SEL) return fa_s; else return A.fa[pr]; };
    auto fb_ = [&]() -> const u64* { if constexpr (SEL) return fb_s; else return A.fb[pr]; };
    auto twf_ = [&]() -> const u64* { if constexpr (SEL) return twf_s; else return A.twf[pr]; };
    auto twi_ = [&]() -> const u64* { if constexpr (SEL) return twi_s; else return A.twi[pr]; };
    auto alpha_ = [&]() -> u64 { if constexpr (SEL) return alpha_s; else return A.alpha_m[pr]; };
    u64 a[16], b[16];
    auto xs = [&](int j) { return base + (yt | ((u32)j << LO_IO)); };
    if constexpr (FROM_LIMBS) {
        if (A.an == 0) CFor<0, 16>::run([&](auto jc) { a[decltype(jc)::value] = 0; });
        else load_limbs<16>(P, a, A.la, A.an, alpha_(), xs);
    } else {
        CFor<0, 16>::run([&](auto jc) { a[decltype(jc)::value] = ld_stream(&fa_()[xs(decltype(jc)::value)]); });
    }
    if constexpr (FWD_ONLY) {
        fwd_rec<CB, CB - 1, 1, CB - 1, LO_IO, 0, LAZY>(P, a, lds, t, chunk, twf_());
        if constexpr (LO_F != LO_IO) xchg<CB, LO_F, LO_IO>(a, lds, t);
        CFor<0, 16>::run([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            st_stream(&fa_()[base + (yt | ((u32)j << LO_IO))], a[j]);
        });
        return;
    } else {
        if (!A.a_done) {
            fwd_rec<CB, CB - 1, 1, CB - 1, LO_IO, 0, LAZY>(P, a, lds, t, chunk, twf_());
            if constexpr (LO_F != 0) xchg<CB, LO_F, 0>(a, lds, t);
        } else {
            xchg<CB, LO_IO, 0>(a, lds, t);
        }
        if (!SQR && !sqr_rt) {
            // B comes from limbs only in the kernels that load them
            // (transforms of one chunk); the others always read fb.
            if (FROM_LIMBS && A.lb && A.bn) load_limbs<16>(P, b, A.lb, A.bn, alpha_(), xs);
            else if (FROM_LIMBS && A.lb) CFor<0, 16>::run([&](auto jc) { b[decltype(jc)::value] = 0; });
            else CFor<0, 16>::run([&](auto jc) { b[decltype(jc)::value] = ld_stream(&fb_()[xs(decltype(jc)::value)]); });
            if (!A.b_done) {
                fwd_rec<CB, CB - 1, 1, CB - 1, LO_IO, 0, LAZY>(P, b, lds, t, chunk, twf_());
                if constexpr (LO_F != 0) xchg<CB, LO_F, 0>(b, lds, t);
            } else {
                xchg<CB, LO_IO, 0>(b, lds, t);
            }
        } else {
            CFor<0, 16>::run([&](auto jc) { b[decltype(jc)::value] = a[decltype(jc)::value]; });
        }
        // base multiply: window 0 holds y = 16t + j, groups g = 4t + j/4
        const u64* twg = twf_() + (base >> 2) + 4 * (u64)t;
        CFor<0, 4>::run([&](auto gc) {
            constexpr int g = decltype(gc)::value, j = 4 * g;
            base_mul4<LAZY>(P, a[j], a[j + 1], a[j + 2], a[j + 3], b[j], b[j + 1], b[j + 2], b[j + 3], twg[g]);
        });
        inv_rec<CB, 1, 1, CB - 1, 0, LAZY>(P, a, lds, t, chunk, twi_());
        constexpr int LO_I = inv_window(CB - 1, 1, 0, CB);
        if constexpr (LO_I != LO_IO) xchg<CB, LO_I, LO_IO>(a, lds, t);
        CFor<0, 16>::run([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            st_stream(&fa_()[base + (yt | ((u32)j << LO_IO))], a[j]);
        });
    }
}

// Work items of a persistent launch: at most `cap` work-groups.
inline unsigned persist_grid(u64 items, unsigned cap) { return (unsigned)(items < cap ? (items ? items : 1) : cap); }

}  // namespace nnh

// nnh_carry.hpp — carry machinery shared by the CRT, the limb arithmetic
// and the small-product kernels: Garner reconstruction with a wave-level
// carry-lookahead over 4096-word tiles, the single work-group scan of tile
// carry functions and the carry fix-up.
#pragma once
#include "nnh_arith.hpp"
#include "nnh_internal.hpp"

namespace nnh {

// CRT + carry.  Garner (p0, p2, then p1 as in reference asmCRT,
// asmLinux.asm:1369-1450) gives X_k < 2^186 as three words; the product
// words are  sum_k X_k B^k (+ an optional addend).  With
//   D_k = x0_k + x1_{k-1} + x2_{k-2} + add_k = dlo_k + 2^64 dhi_k,
//   E_k = dlo_k + dhi_{k-1} <= 2^64 + 2,
// every position generates (E_k >= 2^64) or propagates (E_k == 2^64-1) a
// single carry bit; a wave resolves 64 of them with one 64-bit add of its
// ballots (carry-ins = (A + G + cin) ^ P, A = G|P).  Tiles resolve with
// carry-in 0 and publish their carry function; a scan and a fix-up pass
// apply the tile carry-ins.
// ---------------------------------------------------------------------------
struct CrtArgs {
    const u64* r[3];      // residues, or the three words of column sums
    int columns;          // 1: r[] already holds X_k as three words
    u64 n_res;            // positions carrying residues
    u64* out;
    u64 n_out;            // output words
    const u64* addend;    // optional, n_add words added at position 0
    u64 n_add;
    u8* tile_fn;          // per tile: bit0 = carry-out with cin 0, bit1 = with cin 1
    CrtConsts C;
};

constexpr int CRT_TILE_ROUNDS = 16;  // 16 x 256 positions per tile

// The three words at position k as loaded (residues, or column words);
// zero past n_res.  GMEM: the residues, addend and output are global memory
// (the CRT kernels; accesses say so, see ld_stream); otherwise they may be
// LDS (work-group products of the natural layer) and take generic accesses.
struct CrtIn {
    u64 r0, r1, r2;
};
template <bool GMEM>
__device__ __forceinline__ u64 crt_ld(const u64* p) {
    if constexpr (GMEM) return ld_stream(p);
    else return *p;
}
template <bool GMEM>
__device__ __forceinline__ CrtIn crt_load(const CrtArgs& A, u64 k) {
    if (k >= A.n_res) return CrtIn{0, 0, 0};
    return CrtIn{crt_ld<GMEM>(&A.r[0][k]), crt_ld<GMEM>(&A.r[1][k]), crt_ld<GMEM>(&A.r[2][k])};
}
__device__ __forceinline__ void crt_finish(const CrtArgs& A, u64 k, const CrtIn& in, u64& x0, u64& x1, u64& x2) {
    if (k >= A.n_res) x0 = x1 = x2 = 0;
    else if (A.columns) { x0 = in.r0; x1 = in.r1; x2 = in.r2; }
    else {
        // p0 and p1 leave the inverse transform in [0, 2p) (lazy butterflies)
        constexpr Prime Q0 = prime_of(0), Q1 = prime_of(1);
        garner(csub(Q0, in.r0), csub(Q1, in.r1), in.r2, A.C.c02m, A.C.c1m, A.C.km, x0, x1, x2);
    }
}
template <bool GMEM>
__device__ __forceinline__ void crt_x(const CrtArgs& A, u64 k, u64& x0, u64& x1, u64& x2) {
    crt_finish(A, k, crt_load<GMEM>(A, k), x0, x1, x2);
}

// One round of a tile: positions k = tile_base + 256 r + t.  FULL: every
// position of the tile carries residues and an output word and there is no
// addend, so the round needs no range checks (the common case of a product's
// tiles; the last tile and column or addend jobs take the general form).
template <bool FULL, bool GMEM>
__device__ __forceinline__ void crt_round(const CrtArgs& A, u64 k, const CrtIn& cur, u64 (*xs)[256 + 3],
                                          u32 (*wco)[4], u32& cin0, u32& cin1) {
    const u32 t = threadIdx.x, lane = t & 63, wv = t >> 6;
    u64 x0, x1, x2;
    if constexpr (FULL) {
        // p0 and p1 leave the inverse transform in [0, 2p) (lazy butterflies)
        constexpr Prime Q0 = prime_of(0), Q1 = prime_of(1);
        garner(csub(Q0, cur.r0), csub(Q1, cur.r1), cur.r2, A.C.c02m, A.C.c1m, A.C.km, x0, x1, x2);
    } else {
        crt_finish(A, k, cur, x0, x1, x2);
    }
    __syncthreads();
    xs[0][3 + t] = x0; xs[1][3 + t] = x1; xs[2][3 + t] = x2;
    __syncthreads();
    // D_k and D_{k-1}
    const u64 ad = (!FULL && A.addend && k < A.n_add) ? (GMEM ? ld_g(&A.addend[k]) : A.addend[k]) : 0;
    const u64 a1 = xs[1][2 + t], a2 = xs[2][1 + t];
    // low word of D_k; its high part reaches position k+1, which
    // recomputes it from the same LDS words (cp below)
    const u64 dlo = x0 + a1 + a2 + ad;
    const u64 ad1 =
        (!FULL && A.addend && k >= 1 && k - 1 < A.n_add) ? (GMEM ? ld_g(&A.addend[k - 1]) : A.addend[k - 1]) : 0;
    const u64 p0 = xs[0][2 + t], p1 = xs[1][1 + t], p2 = xs[2][t];
    const u64 q = p0 + p1;
    u32 cp = q < p0 ? 1u : 0u;
    const u64 q2 = q + p2;
    cp += q2 < q ? 1u : 0u;
    const u64 q3 = q2 + ad1;
    cp += q3 < q2 ? 1u : 0u;
    const u64 e = dlo + cp;
    const bool g = e < dlo;
    const bool pp = e == ~0ull;
    const u64 G = __ballot(g), P = __ballot(pp), Am = G | P;
    // wave carry functions
    const u64 s0 = Am + G;
    const u32 co0 = s0 < Am ? 1u : 0u;
    const u64 s1 = s0 + 1;
    const u32 co1 = co0 | (s1 == 0 ? 1u : 0u);
    if (lane == 0) { wco[0][wv] = co0; wco[1][wv] = co1; }
    __syncthreads();
    // compose the four wave functions in order (a loop of constant length
    // over wave-uniform values): the carry into this wave is the state
    // before wave wv, the carry out of the round the state after wave 3
    const u32 wvu = __builtin_amdgcn_readfirstlane(wv);
    u32 c0 = cin0, c1 = cin1, n0 = cin0, n1 = cin1;
    for (u32 w = 0; w < 4; w++) {
        if (w == wvu) { c0 = n0; c1 = n1; }
        const u32 f0 = wco[0][w], f1 = wco[1][w];
        n0 = n0 ? f1 : f0;
        n1 = n1 ? f1 : f0;
    }
    const u64 sw = c0 ? s1 : s0;
    const u32 cin_lane = (u32)(((sw ^ P) >> lane) & 1u);
    if (FULL || k < A.n_out) {
        if constexpr (GMEM) st_g(&A.out[k], e + cin_lane);
        else A.out[k] = e + cin_lane;
    }
    cin0 = n0; cin1 = n1;
    __syncthreads();
    if (t >= 253) {
        xs[0][t - 253] = x0; xs[1][t - 253] = x1; xs[2][t - 253] = x2;
    }
}

template <bool GMEM, int ROUNDS = CRT_TILE_ROUNDS>
__device__ __forceinline__ void crt_tile_body(const CrtArgs& A, u64 tile, u8* fn_slot, u64 (*xs)[256 + 3],
                                              u32 (*wco)[4]) {
    const u32 t = threadIdx.x;
    const u64 tile_base = tile * (256 * ROUNDS);
    __syncthreads();  // LDS reuse across persistent iterations
    if (t < 3) {
        u64 x0 = 0, x1 = 0, x2 = 0;
        if (tile_base >= 3 - t) crt_x<GMEM>(A, tile_base - 3 + t, x0, x1, x2);
        xs[0][t] = x0; xs[1][t] = x1; xs[2][t] = x2;
    }
    u32 cin0 = 0, cin1 = 1;  // running carry assuming tile carry-in 0 / 1
    // each round's words are loaded one round ahead, so that their HBM
    // latency overlaps the previous round's reconstruction and carries
    const u64 tile_end = tile_base + 256 * (u64)ROUNDS;
    if (!A.columns && !A.addend && tile_end <= A.n_res && tile_end <= A.n_out) {  // uniform
        CrtIn cur{crt_ld<GMEM>(&A.r[0][tile_base + t]), crt_ld<GMEM>(&A.r[1][tile_base + t]),
                  crt_ld<GMEM>(&A.r[2][tile_base + t])};
#pragma unroll 2
        for (int r = 0; r < ROUNDS; r++) {
            const u64 k = tile_base + (u64)r * 256 + t;
            CrtIn next{};
            if (r + 1 < ROUNDS)
                next = CrtIn{crt_ld<GMEM>(&A.r[0][k + 256]), crt_ld<GMEM>(&A.r[1][k + 256]),
                             crt_ld<GMEM>(&A.r[2][k + 256])};
            crt_round<true, GMEM>(A, k, cur, xs, wco, cin0, cin1);
            cur = next;
        }
    } else {
        CrtIn next = crt_load<GMEM>(A, tile_base + t);
        for (int r = 0; r < ROUNDS; r++) {
            const u64 k = tile_base + (u64)r * 256 + t;
            if (k >= A.n_out && tile_base + (u64)r * 256 >= A.n_out) break;  // uniform
            const CrtIn cur = next;
            if (r + 1 < ROUNDS) next = crt_load<GMEM>(A, k + 256);
            crt_round<false, GMEM>(A, k, cur, xs, wco, cin0, cin1);
        }
    }
    if (t == 0) *fn_slot = (u8)(cin0 | (cin1 << 1));
}


// Single work-group scan of tile carry functions -> carry-in per tile.
// Each tile function is kill (00), generate (11) or propagate (10).  cin0 is
// the carry into the first tile; the carry out of the last is stored to
// *cout when cout is non-null.  1024 threads: each composes a run of
// consecutive tiles, a wave resolves its 64 runs with one 64-bit add of
// its generate/propagate ballots (as the CRT does for words), and the 16
// wave functions are composed in order.
__device__ __forceinline__ void carry_scan_body(const u8* tile_fn, u64 n_tiles, u8* tile_cin, u32* chunk_fn,
                                                u32 cin0 = 0, u32* cout = nullptr) {
    const u32 t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const u64 per = (n_tiles + 1023) / 1024;
    const u64 b0 = (u64)t * per, b1 = b0 + per < n_tiles ? b0 + per : n_tiles;
    u32 f0 = 0, f1 = 1;  // composed function of this thread's tiles
    for (u64 i = b0; i < b1; i++) {
        u32 f = tile_fn[i];
        u32 o0 = f & 1u, o1 = (f >> 1) & 1u;
        f0 = f0 ? o1 : o0;
        f1 = f1 ? o1 : o0;
    }
    const u64 G = __ballot(f0 != 0), P = __ballot(f0 == 0 && f1 != 0), Am = G | P;
    const u64 s0 = Am + G, s1 = s0 + 1;
    const u32 co0 = s0 < Am ? 1u : 0u, co1 = co0 | (s1 == 0 ? 1u : 0u);
    if (lane == 0) chunk_fn[wv] = co0 | (co1 << 1);
    __syncthreads();
    // the 16 wave functions composed over wave-uniform values (constant trip
    // count): cw = the carry into this wave, cc = out of the last wave
    const u32 wvu = __builtin_amdgcn_readfirstlane(wv);
    u32 cw = cin0, cc = cin0;
    for (u32 w = 0; w < 16; w++) {
        if (w == wvu) cw = cc;
        const u32 f = chunk_fn[w];
        cc = cc ? (f >> 1) & 1u : f & 1u;
    }
    if (cout && t == 0) st_g(cout, cc);
    u32 c = (u32)((((cw ? s1 : s0) ^ P) >> lane) & 1u);
    for (u64 i = b0; i < b1; i++) {
        tile_cin[i] = (u8)c;
        u32 f = tile_fn[i];
        c = c ? (f >> 1) & 1u : f & 1u;
    }
}

// Add the tile carry-in (0/1) at the tile's first word and propagate it
// through the tile's all-ones words.
template <int ROUNDS = CRT_TILE_ROUNDS>
__device__ __forceinline__ void carry_fix_body(u64* out, u64 n_out, u64 tile, u32& first_nz) {
    const u32 t = threadIdx.x;
    const u64 tile_base = tile * (256 * ROUNDS);
    for (int r = 0; r < ROUNDS; r++) {
        u64 base = tile_base + (u64)r * 256;
        if (base >= n_out) return;
        u64 k = base + t;
        u64 w = k < n_out ? ld_g(&out[k]) : 0;
        if (t == 0) first_nz = 256;
        __syncthreads();
        if (k < n_out && w != ~0ull) atomicMin(&first_nz, t);
        __syncthreads();
        u32 f = first_nz;
        if (k < n_out) {
            if (t < f) st_g(&out[k], (u64)0);
            else if (t == f) st_g(&out[k], w + 1);
        }
        if (f < 256) return;
        __syncthreads();
    }
}

}  // namespace nnh

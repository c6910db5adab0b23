// nnh_kernels.hip — CDNA4 kernels for the three-prime NTT product.
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
#include "nnh_arith.hpp"
#include "nnh_internal.hpp"

namespace nnh {

__device__ __forceinline__ u32 lay(u32 t, u32 j, u32 lo) {
    u32 m = (1u << lo) - 1u;
    return (t & m) | (j << lo) | ((t >> lo) << (lo + 4));
}
__device__ __forceinline__ u32 pad(u32 y) { return y + (y >> 4); }

template <int CB>
__device__ __forceinline__ void xchg(u64 (&v)[16], u64* lds, u32 t, u32 lo_from, u32 lo_to) {
    __syncthreads();
#pragma unroll
    for (u32 j = 0; j < 16; j++) lds[pad(lay(t, j, lo_from))] = v[j];
    __syncthreads();
#pragma unroll
    for (u32 j = 0; j < 16; j++) v[j] = lds[pad(lay(t, j, lo_to))];
}

// One level (local bit l, register distance D = 2^(l-lo)).
template <u32 K, bool INV, int D, int CB>
__device__ __forceinline__ void level_d(u64 (&v)[16], u32 t, u32 lo, u32 l, u64 tile_id,
                                        const u64* __restrict__ tw) {
    u64 bbase = tile_id << (CB - 1 - l);
#pragma unroll
    for (u32 j = 0; j < 16; j++) {
        if (j & D) continue;
        u32 y = lay(t, j, lo);
        u64 w = tw[bbase + (y >> (l + 1))];
        if (INV) gs_bfly<K>(v[j], v[j + D], w);
        else ct_bfly<K>(v[j], v[j + D], w);
    }
}
template <u32 K, bool INV, int CB>
__device__ __forceinline__ void level(u64 (&v)[16], u32 t, u32 lo, u32 l, u64 tile_id,
                                      const u64* __restrict__ tw) {
    switch (l - lo) {
        case 0: level_d<K, INV, 1, CB>(v, t, lo, l, tile_id, tw); break;
        case 1: level_d<K, INV, 2, CB>(v, t, lo, l, tile_id, tw); break;
        case 2: level_d<K, INV, 4, CB>(v, t, lo, l, tile_id, tw); break;
        default: level_d<K, INV, 8, CB>(v, t, lo, l, tile_id, tw); break;
    }
}

// Forward levels for local bits l_hi..l_lo (descending); the window starts at
// *lo and never drops below lo_min.  Returns with *lo = final window.
template <u32 K, int CB>
__device__ __forceinline__ void fwd_levels(u64 (&v)[16], u64* lds, u32 t, u32& lo, int l_hi, int l_lo,
                                           int lo_min, u64 tile_id, const u64* __restrict__ tw) {
    for (int l = l_hi; l >= l_lo; l--) {
        if (l < (int)lo) {
            int nl = l - 3 < lo_min ? lo_min : l - 3;
            xchg<CB>(v, lds, t, lo, (u32)nl);
            lo = (u32)nl;
        }
        level<K, false, CB>(v, t, lo, (u32)l, tile_id, tw);
    }
}
// Inverse levels for local bits l_lo..l_hi (ascending).
template <u32 K, int CB>
__device__ __forceinline__ void inv_levels(u64 (&v)[16], u64* lds, u32 t, u32& lo, int l_lo, int l_hi,
                                           u64 tile_id, const u64* __restrict__ tw) {
    for (int l = l_lo; l <= l_hi; l++) {
        if (l > (int)lo + 3) {
            int nl = l > CB - 4 ? CB - 4 : l;
            xchg<CB>(v, lds, t, lo, (u32)nl);
            lo = (u32)nl;
        }
        level<K, true, CB>(v, t, lo, (u32)l, tile_id, tw);
    }
}

// ---------------------------------------------------------------------------
// Operand load: residue of limb x scaled by alpha (Montgomery form alpha*R),
// zero beyond the operand.  Reference ntt_info.h:147-173 reduces by Barrett;
// here the scale alpha^2 = 2^64 * 2^-(s-1) absorbs both the base-multiply
// Montgomery factor and the inverse transform's 2^(s-1).
// ---------------------------------------------------------------------------
template <u32 K>
__device__ __forceinline__ u64 load_limb(const u64* __restrict__ a, u64 an, u64 x, u64 alpha_m) {
    return x < an ? mont<K>(a[x], alpha_m) : 0;
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

// Top pass over a 4096-element tile of Q = 2^(12-nlev) consecutive columns
// by R = 2^nlev rows of stride d = 2^g_lo.  blockIdx.x = tile, blockIdx.y = prime.
template <u32 K, bool INV, bool FROM_LIMBS>
__device__ __forceinline__ void top_pass_body(const TopArgs& A, u32 pr, u64* lds) {
    constexpr int CB = 12;
    const u32 t = threadIdx.x;
    const u32 q = 12 - A.nlev;
    const u64 d = 1ull << A.g_lo;
    const u64 tiles_per_blk = d >> q;
    const u64 tile = blockIdx.x;
    const u64 blk = tile / tiles_per_blk;
    const u64 cbase = (tile % tiles_per_blk) << q;
    const u64 xbase = (blk << (A.g_lo + A.nlev)) + cbase;
    const u32 qm = (1u << q) - 1u;
    u64* __restrict__ data = A.data[pr];
    u64 v[16];
    u32 lo = CB - 4;
#pragma unroll
    for (u32 j = 0; j < 16; j++) {
        u32 y = lay(t, j, lo);
        u64 x = xbase + (y & qm) + (u64)(y >> q) * d;
        if (FROM_LIMBS) v[j] = load_limb<K>(A.limbs, A.an, x, A.alpha_m[pr]);
        else v[j] = data[x];
    }
    if (!INV) {
        fwd_levels<K, CB>(v, lds, t, lo, CB - 1, (int)q, (int)q, blk, A.tw[pr]);
    } else {
        // inverse: start from a window at the bottom of the row bits
        u32 nl = q < CB - 4 ? q : CB - 4;
        if (nl != lo) {
            xchg<CB>(v, lds, t, lo, nl);
            lo = nl;
        }
        inv_levels<K, CB>(v, lds, t, lo, (int)q, CB - 1, blk, A.tw[pr]);
    }
    if (lo < q) {
        xchg<CB>(v, lds, t, lo, CB - 4);
        lo = CB - 4;
    }
#pragma unroll
    for (u32 j = 0; j < 16; j++) {
        u32 y = lay(t, j, lo);
        u64 x = xbase + (y & qm) + (u64)(y >> q) * d;
        data[x] = v[j];
    }
}

template <bool INV, bool FROM_LIMBS>
__global__ __launch_bounds__(256) void k_top_pass(TopArgs A) {
    __shared__ u64 lds[4096 + 256];
    const u32 pr = blockIdx.y;
    if (pr == 0) top_pass_body<PRIME_K[0], INV, FROM_LIMBS>(A, 0, lds);
    else if (pr == 1) top_pass_body<PRIME_K[1], INV, FROM_LIMBS>(A, 1, lds);
    else top_pass_body<PRIME_K[2], INV, FROM_LIMBS>(A, 2, lds);
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

template <u32 K, int CB, bool FROM_LIMBS, bool SQR, bool FWD_ONLY>
__device__ __forceinline__ void mid_body(const MidArgs& A, u32 pr, u64* lds) {
    const u32 t = threadIdx.x;
    const u64 chunk = blockIdx.x;
    const u64 base = chunk << CB;
    u64 a[16], b[16];
    u32 lo = CB - 4;
#pragma unroll
    for (u32 j = 0; j < 16; j++) {
        u64 x = base + lay(t, j, lo);
        if (FROM_LIMBS) a[j] = load_limb<K>(A.la, A.an, x, A.alpha_m[pr]);
        else a[j] = A.fa[pr][x];
    }
    if (!A.a_done) fwd_levels<K, CB>(a, lds, t, lo, CB - 1, 1, 0, chunk, A.twf[pr]);
    if (FWD_ONLY) {
        if (lo != CB - 4) { xchg<CB>(a, lds, t, lo, CB - 4); lo = CB - 4; }
#pragma unroll
        for (u32 j = 0; j < 16; j++) A.fa[pr][base + lay(t, j, lo)] = a[j];
        return;
    }
    if (lo != 0) { xchg<CB>(a, lds, t, lo, 0); lo = 0; }
    if (!SQR) {
        u32 lob = CB - 4;
#pragma unroll
        for (u32 j = 0; j < 16; j++) {
            u64 x = base + lay(t, j, lob);
            if (A.lb) b[j] = load_limb<K>(A.lb, A.bn, x, A.alpha_m[pr]);
            else b[j] = A.fb[pr][x];
        }
        if (!A.b_done) fwd_levels<K, CB>(b, lds, t, lob, CB - 1, 1, 0, chunk, A.twf[pr]);
        if (lob != 0) xchg<CB>(b, lds, t, lob, 0);
    } else {
#pragma unroll
        for (u32 j = 0; j < 16; j++) b[j] = a[j];
    }
    // base multiply: window lo = 0 holds y = 16t + j, groups g = 4t + j/4
    const u64* twf = A.twf[pr];
#pragma unroll
    for (u32 j = 0; j < 16; j += 4) {
        u64 g = (base >> 2) + (u64)(4 * t + j / 4);
        base_mul4<K>(a[j], a[j + 1], a[j + 2], a[j + 3], b[j], b[j + 1], b[j + 2], b[j + 3], twf[g]);
    }
    inv_levels<K, CB>(a, lds, t, lo, 1, CB - 1, chunk, A.twi[pr]);
    if (lo != CB - 4) { xchg<CB>(a, lds, t, lo, CB - 4); lo = CB - 4; }
#pragma unroll
    for (u32 j = 0; j < 16; j++) A.fa[pr][base + lay(t, j, lo)] = a[j];
}

template <int CB, bool FROM_LIMBS, bool SQR, bool FWD_ONLY>
__global__ __launch_bounds__(1 << (CB - 4)) void k_mid(MidArgs A) {
    __shared__ u64 lds[(1 << CB) + (1 << (CB - 4))];
    const u32 pr = blockIdx.y;
    if (pr == 0) mid_body<PRIME_K[0], CB, FROM_LIMBS, SQR, FWD_ONLY>(A, 0, lds);
    else if (pr == 1) mid_body<PRIME_K[1], CB, FROM_LIMBS, SQR, FWD_ONLY>(A, 1, lds);
    else mid_body<PRIME_K[2], CB, FROM_LIMBS, SQR, FWD_ONLY>(A, 2, lds);
}

// ---------------------------------------------------------------------------
// Twiddle tables: entry t = w_t * 2^64 mod p (Montgomery form), where
// w_t = prod over set bits q of t of rb[q+1] (reference ntt_info.h:22-41).
// ---------------------------------------------------------------------------
template <u32 K>
__global__ void k_twiddles(u64* tw, u64 n, RootBases rb, int inv, int pr) {
    u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const u64* base = inv ? rb.irb_m[pr] : rb.rb_m[pr];
    u64 acc = rb.one_m[pr];
    for (int q = 0; (t >> q) != 0; q++)
        if ((t >> q) & 1) acc = csub<K>(mont<K>(acc, base[q + 1]));
    tw[t] = acc;
}

// ---------------------------------------------------------------------------
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

__device__ __forceinline__ void crt_x(const CrtArgs& A, u64 k, u64& x0, u64& x1, u64& x2) {
    if (k >= A.n_res) x0 = x1 = x2 = 0;
    else if (A.columns) { x0 = A.r[0][k]; x1 = A.r[1][k]; x2 = A.r[2][k]; }
    else garner(A.r[0][k], A.r[1][k], A.r[2][k], A.C.c02m, A.C.m0m, A.C.c1m, A.C.m02lo, A.C.m02hi, x0, x1, x2);
}

__global__ __launch_bounds__(256) void k_crt_tile(CrtArgs A) {
    __shared__ u64 xs[3][256 + 3];
    __shared__ u32 wco[2][4];
    const u32 t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const u64 tile_base = (u64)blockIdx.x * (256 * CRT_TILE_ROUNDS);
    if (t < 3) {
        u64 x0 = 0, x1 = 0, x2 = 0;
        if (tile_base >= 3 - t) crt_x(A, tile_base - 3 + t, x0, x1, x2);
        xs[0][t] = x0; xs[1][t] = x1; xs[2][t] = x2;
    }
    u32 cin0 = 0, cin1 = 1;  // running carry assuming tile carry-in 0 / 1
    for (int r = 0; r < CRT_TILE_ROUNDS; r++) {
        const u64 k = tile_base + (u64)r * 256 + t;
        if (k >= A.n_out && tile_base + (u64)r * 256 >= A.n_out) break;  // uniform
        u64 x0, x1, x2;
        crt_x(A, k, x0, x1, x2);
        __syncthreads();
        xs[0][3 + t] = x0; xs[1][3 + t] = x1; xs[2][3 + t] = x2;
        __syncthreads();
        // D_k and D_{k-1}
        u64 ad = (A.addend && k < A.n_add) ? A.addend[k] : 0;
        u64 a1 = xs[1][2 + t], a2 = xs[2][1 + t];
        u64 s = x0 + a1;
        u32 c = s < x0 ? 1u : 0u;
        u64 s2 = s + a2;
        c += s2 < s ? 1u : 0u;
        u64 dlo = s2 + ad;
        c += dlo < s2 ? 1u : 0u;
        u64 ad1 = (A.addend && k >= 1 && k - 1 < A.n_add) ? A.addend[k - 1] : 0;
        u64 p0 = xs[0][2 + t], p1 = xs[1][1 + t], p2 = xs[2][t];
        u64 q = p0 + p1;
        u32 cp = q < p0 ? 1u : 0u;
        u64 q2 = q + p2;
        cp += q2 < q ? 1u : 0u;
        u64 q3 = q2 + ad1;
        cp += q3 < q2 ? 1u : 0u;
        u64 e = dlo + cp;
        bool g = e < dlo;
        bool pp = e == ~0ull;
        u64 G = __ballot(g), P = __ballot(pp), Am = G | P;
        // wave carry functions
        u64 s0 = Am + G;
        u32 co0 = s0 < Am ? 1u : 0u;
        u64 s1 = s0 + 1;
        u32 co1 = co0 | (s1 == 0 ? 1u : 0u);
        if (lane == 0) { wco[0][wv] = co0; wco[1][wv] = co1; }
        __syncthreads();
        u32 c0 = cin0, c1 = cin1;
        for (u32 w = 0; w < wv; w++) {
            c0 = c0 ? wco[1][w] : wco[0][w];
            c1 = c1 ? wco[1][w] : wco[0][w];
        }
        u64 sw = c0 ? s1 : s0;
        u32 cin_lane = (u32)(((sw ^ P) >> lane) & 1u);
        if (k < A.n_out) A.out[k] = e + cin_lane;
        // round carry-out for the next round
        u32 n0 = cin0, n1 = cin1;
        for (u32 w = 0; w < 4; w++) {
            n0 = n0 ? wco[1][w] : wco[0][w];
            n1 = n1 ? wco[1][w] : wco[0][w];
        }
        cin0 = n0; cin1 = n1;
        __syncthreads();
        if (t >= 253) {
            xs[0][t - 253] = x0; xs[1][t - 253] = x1; xs[2][t - 253] = x2;
        }
    }
    if (t == 0) A.tile_fn[blockIdx.x] = (u8)(cin0 | (cin1 << 1));
}

// Single work-group scan of tile carry functions -> carry-in per tile.
// Each tile function is kill (00), generate (11) or propagate (10).
__global__ __launch_bounds__(1024) void k_carry_scan(u8* tile_fn, u64 n_tiles, u8* tile_cin) {
    __shared__ u32 chunk_fn[1024];
    const u32 t = threadIdx.x;
    const u64 per = (n_tiles + 1023) / 1024;
    const u64 b0 = (u64)t * per, b1 = b0 + per < n_tiles ? b0 + per : n_tiles;
    u32 f0 = 0, f1 = 1;  // composed function of this thread's tiles
    for (u64 i = b0; i < b1; i++) {
        u32 f = tile_fn[i];
        u32 o0 = f & 1u, o1 = (f >> 1) & 1u;
        f0 = f0 ? o1 : o0;
        f1 = f1 ? o1 : o0;
    }
    chunk_fn[t] = f0 | (f1 << 1);
    __syncthreads();
    // serial exclusive scan over 1024 thread functions (cheap: 1024 steps)
    if (t == 0) {
        u32 c = 0;
        for (u32 i = 0; i < 1024; i++) {
            u32 f = chunk_fn[i];
            chunk_fn[i] = c;
            c = c ? (f >> 1) & 1u : f & 1u;
        }
    }
    __syncthreads();
    u32 c = chunk_fn[t];
    for (u64 i = b0; i < b1; i++) {
        tile_cin[i] = (u8)c;
        u32 f = tile_fn[i];
        c = c ? (f >> 1) & 1u : f & 1u;
    }
}

// Add the tile carry-in (0/1) at the tile's first word and propagate it
// through the tile's all-ones words.
__global__ __launch_bounds__(256) void k_carry_fix(u64* out, u64 n_out, const u8* tile_cin) {
    __shared__ u32 first_nz;
    if (!tile_cin[blockIdx.x]) return;
    const u32 t = threadIdx.x;
    const u64 tile_base = (u64)blockIdx.x * (256 * CRT_TILE_ROUNDS);
    for (int r = 0; r < CRT_TILE_ROUNDS; r++) {
        u64 base = tile_base + (u64)r * 256;
        if (base >= n_out) return;
        u64 k = base + t;
        u64 w = k < n_out ? out[k] : 0;
        if (t == 0) first_nz = 256;
        __syncthreads();
        if (k < n_out && w != ~0ull) atomicMin(&first_nz, t);
        __syncthreads();
        u32 f = first_nz;
        if (k < n_out) {
            if (t < f) out[k] = 0;
            else if (t == f) out[k] = w + 1;
        }
        if (f < 256) return;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Schoolbook columns (reference asmMul1, asmLinux.asm:188-278, column-wise
// with a 3-word accumulator): X_k = sum_i a_i b_{k-i} as three words, one
// thread per column; k_crt_tile turns the columns into product words.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_school_cols(const u64* __restrict__ a, u64 an,
                                                     const u64* __restrict__ b, u64 bn,
                                                     u64* c0, u64* c1, u64* c2) {
    u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= an + bn - 1) return;
    u64 i0 = k >= bn ? k - bn + 1 : 0;
    u64 i1 = k < an - 1 ? k : an - 1;
    u64 s0 = 0, s1 = 0, s2 = 0;
    for (u64 i = i0; i <= i1; i++) {
        u64 lo, hi;
        mul_wide(a[i], b[k - i], lo, hi);
        s0 += lo;
        u64 c = s0 < lo ? 1u : 0u;
        u64 h = hi + c;  // hi <= 2^64-2, no overflow
        s1 += h;
        s2 += s1 < h ? 1u : 0u;
    }
    c0[k] = s0; c1[k] = s1; c2[k] = s2;
}

// Cyclic fold (reference NTT_info::save, ntt_info.h:238-242): w[0..L) +=
// (w[L], w[L+1]) with carry, then the carry out is added once more at word 0
// (its own carry out is dropped, as in the reference).  One work-group.
__device__ void wg_add_at(u64* w, u64 n, u64 start, u32& carry, u32* sh) {
    // add 1 at position `start` and propagate through all-ones words
    const u32 t = threadIdx.x;
    for (u64 base = start; base < n && carry; base += 256) {
        u64 k = base + t;
        u64 v = k < n ? w[k] : 0;
        if (t == 0) *sh = 256;
        __syncthreads();
        if (k < n && v != ~0ull) atomicMin(sh, t);
        __syncthreads();
        u32 f = *sh;
        if (k < n) {
            if (t < f) w[k] = 0;
            else if (t == f) w[k] = v + 1;
        }
        if (f < 256 && base + f < n) carry = 0;
        __syncthreads();
    }
}
__global__ __launch_bounds__(256) void k_cyclic_fold(u64* w, u64 L) {
    __shared__ u32 sh;
    __shared__ u32 cf_sh;
    if (threadIdx.x == 0) {
        u64 clo = w[L], chi = w[L + 1];
        u64 x0 = w[0] + clo;
        u32 c = x0 < clo ? 1u : 0u;
        u64 x1 = w[1] + chi;
        u32 c2 = x1 < chi ? 1u : 0u;
        u64 x1c = x1 + c;
        c2 += x1c < x1 ? 1u : 0u;
        w[0] = x0;
        w[1] = x1c;
        cf_sh = c2;
    }
    __syncthreads();
    u32 carry = cf_sh;
    wg_add_at(w, L, 2, carry, &sh);   // carry out of the 2-word add
    __syncthreads();
    wg_add_at(w, L, 0, carry, &sh);   // end-around carry (final carry dropped)
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
void launch_twiddles(u64* tw, u64 n, const RootBases& rb, int pr, int inv, hipStream_t st) {
    dim3 grid((unsigned)((n + 255) / 256)), block(256);
    if (pr == 0) hipLaunchKernelGGL(k_twiddles<PRIME_K[0]>, grid, block, 0, st, tw, n, rb, inv, pr);
    else if (pr == 1) hipLaunchKernelGGL(k_twiddles<PRIME_K[1]>, grid, block, 0, st, tw, n, rb, inv, pr);
    else hipLaunchKernelGGL(k_twiddles<PRIME_K[2]>, grid, block, 0, st, tw, n, rb, inv, pr);
}

void launch_top_pass(const TopPassDesc& D, hipStream_t st) {
    TopArgs A;
    for (int i = 0; i < 3; i++) {
        A.data[i] = D.data[i];
        A.tw[i] = D.tw[i];
        A.alpha_m[i] = D.alpha_m[i];
    }
    A.limbs = D.limbs;
    A.an = D.an;
    A.g_lo = D.g_lo;
    A.nlev = D.nlev;
    dim3 grid((unsigned)(D.L >> 12), 3), block(256);
    if (D.inverse) hipLaunchKernelGGL((k_top_pass<true, false>), grid, block, 0, st, A);
    else if (D.limbs) hipLaunchKernelGGL((k_top_pass<false, true>), grid, block, 0, st, A);
    else hipLaunchKernelGGL((k_top_pass<false, false>), grid, block, 0, st, A);
}

template <int CB>
static void mid_dispatch(const MidArgs& A, u64 L, bool from_limbs, bool sqr, bool fwd_only, hipStream_t st) {
    dim3 grid((unsigned)(L >> CB), 3), block(1 << (CB - 4));
    if (fwd_only) {
        if (from_limbs) hipLaunchKernelGGL((k_mid<CB, true, false, true>), grid, block, 0, st, A);
        else hipLaunchKernelGGL((k_mid<CB, false, false, true>), grid, block, 0, st, A);
    } else if (from_limbs) {
        if (sqr) hipLaunchKernelGGL((k_mid<CB, true, true, false>), grid, block, 0, st, A);
        else hipLaunchKernelGGL((k_mid<CB, true, false, false>), grid, block, 0, st, A);
    } else {
        if (sqr) hipLaunchKernelGGL((k_mid<CB, false, true, false>), grid, block, 0, st, A);
        else hipLaunchKernelGGL((k_mid<CB, false, false, false>), grid, block, 0, st, A);
    }
}

void launch_mid(const MidDesc& D, hipStream_t st) {
    MidArgs A;
    for (int i = 0; i < 3; i++) {
        A.fa[i] = D.fa[i];
        A.fb[i] = D.fb[i];
        A.twf[i] = D.twf[i];
        A.twi[i] = D.twi[i];
        A.alpha_m[i] = D.alpha_m[i];
    }
    A.la = D.la; A.lb = D.lb; A.an = D.an; A.bn = D.bn;
    A.a_done = D.a_done; A.b_done = D.b_done;
    bool fl = D.la != nullptr;
    switch (D.cb) {
        case 8: mid_dispatch<8>(A, D.L, fl, D.sqr, D.fwd_only, st); break;
        case 9: mid_dispatch<9>(A, D.L, fl, D.sqr, D.fwd_only, st); break;
        case 10: mid_dispatch<10>(A, D.L, fl, D.sqr, D.fwd_only, st); break;
        case 11: mid_dispatch<11>(A, D.L, fl, D.sqr, D.fwd_only, st); break;
        default: mid_dispatch<12>(A, D.L, fl, D.sqr, D.fwd_only, st); break;
    }
}

void launch_crt(const CrtDesc& D, hipStream_t st) {
    CrtArgs A;
    for (int i = 0; i < 3; i++) A.r[i] = D.r[i];
    A.columns = D.columns;
    A.n_res = D.n_res;
    A.out = D.out;
    A.n_out = D.n_out;
    A.addend = D.addend;
    A.n_add = D.n_add;
    A.tile_fn = D.tile_fn;
    A.C = D.C;
    u64 tiles = (D.n_out + 256 * CRT_TILE_ROUNDS - 1) / (256 * CRT_TILE_ROUNDS);
    hipLaunchKernelGGL(k_crt_tile, dim3((unsigned)tiles), dim3(256), 0, st, A);
    if (tiles > 1) {
        hipLaunchKernelGGL(k_carry_scan, dim3(1), dim3(1024), 0, st, D.tile_fn, tiles, D.tile_cin);
        hipLaunchKernelGGL(k_carry_fix, dim3((unsigned)tiles), dim3(256), 0, st, D.out, D.n_out,
                           (const u8*)D.tile_cin);
    }
}

void launch_school_cols(const u64* a, u64 an, const u64* b, u64 bn, u64* c0, u64* c1, u64* c2,
                        hipStream_t st) {
    u64 n = an + bn - 1;
    hipLaunchKernelGGL(k_school_cols, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, an, b, bn, c0, c1,
                       c2);
}

void launch_cyclic_fold(u64* w, u64 L, hipStream_t st) {
    hipLaunchKernelGGL(k_cyclic_fold, dim3(1), dim3(256), 0, st, w, L);
}

u64 crt_tiles(u64 n_out) { return (n_out + 256 * CRT_TILE_ROUNDS - 1) / (256 * CRT_TILE_ROUNDS); }

}  // namespace nnh

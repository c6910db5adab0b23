// nnh_crt.hip — Garner CRT with the carry-lookahead scan, schoolbook
// columns, the cyclic fold and the twiddle tables.
#include "nnh_carry.hpp"
#include "nnh_tile.hpp"

#include <algorithm>

namespace nnh {

// ---------------------------------------------------------------------------
// Twiddle tables: entry t = w_t * 2^64 mod p (Montgomery form), where
// w_t = prod over set bits q of t of rb[q+1] (reference ntt_info.h:22-41).
// ---------------------------------------------------------------------------
__global__ void k_twiddles(u64* tw, u64 n, RootBases rb, int inv, int pr) {
    const Prime P = prime_of(pr);
    u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const u64* base = inv ? rb.irb_m[pr] : rb.rb_m[pr];
    u64 acc = rb.one_m[pr];
    for (int q = 0; (t >> q) != 0; q++)
        if ((t >> q) & 1) acc = csub(P, mont(P, acc, base[q + 1]));
    tw[t] = acc;
}

template <int ROUNDS>
__global__ __launch_bounds__(256) void k_crt_tile(CrtArgs A) {
    __shared__ u64 xs[3][256 + 3];
    __shared__ u32 wco[2][4];
    crt_tile_body<true, ROUNDS>(A, blockIdx.x, A.tile_fn + blockIdx.x, xs, wco);
}

__global__ __launch_bounds__(1024) void k_carry_scan(u8* tile_fn, u64 n_tiles, u8* tile_cin) {
    __shared__ u32 chunk_fn[1024];
    carry_scan_body(tile_fn, n_tiles, tile_cin, chunk_fn);
}

template <int ROUNDS>
__global__ __launch_bounds__(256) void k_carry_fix(u64* out, u64 n_out, const u8* tile_cin) {
    __shared__ u32 first_nz;
    if (!tile_cin[blockIdx.x]) return;
    carry_fix_body<ROUNDS>(out, n_out, blockIdx.x, first_nz);
}

__device__ __forceinline__ u32 find_job_by_tile(const BJob* jobs, u32 njobs, u64 tile) {
    u32 lo = 0, hi = njobs - 1;
    while (lo < hi) {
        u32 mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile0 <= tile) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_crt_batch(const BJob* __restrict__ jobs, u32 njobs, u64 ntiles,
                                                   u8* tile_fn, CrtConsts C) {
    __shared__ u64 xs[3][256 + 3];
    __shared__ u32 wco[2][4];
    for (u64 it = blockIdx.x; it < ntiles; it += gridDim.x) {
        const BJob& J = jobs[find_job_by_tile(jobs, njobs, it)];
        CrtArgs A;
        for (int i = 0; i < 3; i++) A.r[i] = J.r[i];
        A.columns = J.school;
        A.n_res = J.n_res;
        A.out = J.out;
        A.n_out = J.n_out;
        A.addend = J.add;
        A.n_add = J.n_add;
        A.tile_fn = nullptr;
        A.C = C;
        crt_tile_body<true>(A, it - J.tile0, tile_fn + it, xs, wco);
    }
}

__global__ __launch_bounds__(1024) void k_carry_scan_batch(const BJob* __restrict__ jobs, u32 njobs, u64 ntiles,
                                                           u8* tile_fn, u8* tile_cin) {
    __shared__ u32 chunk_fn[1024];
    for (u32 j = blockIdx.x; j < njobs; j += gridDim.x) {
        const u64 t0 = jobs[j].tile0;
        const u64 t1 = j + 1 < njobs ? jobs[j + 1].tile0 : ntiles;
        __syncthreads();
        carry_scan_body(tile_fn + t0, t1 - t0, tile_cin + t0, chunk_fn);
    }
}

__global__ __launch_bounds__(256) void k_carry_fix_batch(const BJob* __restrict__ jobs, u32 njobs, u64 ntiles,
                                                         const u8* tile_cin) {
    __shared__ u32 first_nz;
    for (u64 it = blockIdx.x; it < ntiles; it += gridDim.x) {
        if (!tile_cin[it]) continue;
        const BJob& J = jobs[find_job_by_tile(jobs, njobs, it)];
        __syncthreads();
        carry_fix_body(J.out, J.n_out, it - J.tile0, first_nz);
    }
}

// Schoolbook columns for the batch's small jobs: one thread per column over
// the concatenated column ranges (job.col0 = first global column).
__global__ __launch_bounds__(256) void k_school_batch(const BJob* __restrict__ jobs, const u32* __restrict__ sj,
                                                      u32 nsj, u64 ncols) {
    for (u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x; k < ncols; k += (u64)gridDim.x * blockDim.x) {
        u32 lo = 0, hi = nsj - 1;
        while (lo < hi) {
            u32 mid = (lo + hi + 1) >> 1;
            if (jobs[sj[mid]].col0 <= k) lo = mid;
            else hi = mid - 1;
        }
        const BJob& J = jobs[sj[lo]];
        const u64 c = k - J.col0;
        const u64 an = J.an, bn = J.bn;
        const u64* a = J.a;
        const u64* b = J.b;
        u64 i0 = c >= bn ? c - bn + 1 : 0;
        u64 i1 = c < an - 1 ? c : an - 1;
        u64 s0 = 0, s1 = 0, s2 = 0;
        for (u64 i = i0; i <= i1; i++) {
            u64 lo64, hi64;
            mul_wide(ld_g(&a[i]), ld_g(&b[c - i]), lo64, hi64);
            s0 += lo64;
            u64 h = hi64 + (s0 < lo64 ? 1u : 0u);
            s1 += h;
            s2 += s1 < h ? 1u : 0u;
        }
        const u64 n = J.n_res;
        st_g(&J.fa[c], s0);
        st_g(&J.fa[n + c], s1);
        st_g(&J.fa[2 * n + c], s2);
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

// ---------------------------------------------------------------------------
// One Karatsuba level in the polynomial domain (the reference's toom22mul,
// base.h:21-35, without its carry passes): with h = ceil(bn/2),
//   a = a0 + a1 x^h, b = b0 + b1 x^h (x = 2^64),
//   ab = z0 + (zm - z0 - z2) x^h + z2 x^2h,  zm = (a0 + a1)(b0 + b1),
// where every term is a vector of column sums.  a0 + a1 is formed
// coefficient-wise (65-bit coefficients, no carry chain), so zm's columns
// stay below 2^192 and the combined columns C_k = sum a_i b_{k-i} (three
// words, non-negative) go through the same carry kernel as schoolbook.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_kara_prep(const u64* __restrict__ a, u64 an, const u64* __restrict__ b,
                                                   u64 bn, u64 h, u64* s_lo, u64* s_hi, u64 ns, u64* t_lo,
                                                   u64* t_hi, u64 nt) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ns) {
        u64 x = i < h ? a[i] : 0, y = i < an - h ? a[h + i] : 0;
        u64 z = x + y;
        s_lo[i] = z;
        s_hi[i] = z < x ? 1u : 0u;
    }
    if (i < nt) {
        u64 x = i < h ? b[i] : 0, y = i < bn - h ? b[h + i] : 0;
        u64 z = x + y;
        t_lo[i] = z;
        t_hi[i] = z < x ? 1u : 0u;
    }
}

// Column sums of two polynomials with 65-bit coefficients (lo word + 0/1).
__global__ __launch_bounds__(256) void k_cols_wide(const u64* __restrict__ x_lo, const u64* __restrict__ x_hi,
                                                   u64 nx, const u64* __restrict__ y_lo,
                                                   const u64* __restrict__ y_hi, u64 ny, u64* c0, u64* c1,
                                                   u64* c2) {
    u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nx + ny - 1) return;
    u64 i0 = k >= ny ? k - ny + 1 : 0;
    u64 i1 = k < nx - 1 ? k : nx - 1;
    u64 s0 = 0, s1 = 0, s2 = 0;
    for (u64 i = i0; i <= i1; i++) {
        u64 xl = x_lo[i], xh = x_hi[i], yl = y_lo[k - i], yh = y_hi[k - i];
        u64 lo, hi;
        mul_wide(xl, yl, lo, hi);
        // + (xh*yl + yh*xl) 2^64 + xh*yh 2^128
        u64 top = xh & yh;
        u64 m1 = xh ? yl : 0, m2 = yh ? xl : 0;
        u64 h1 = hi + m1;
        top += h1 < m1 ? 1u : 0u;
        u64 h2 = h1 + m2;
        top += h2 < m2 ? 1u : 0u;
        s0 += lo;
        u64 c = s0 < lo ? 1u : 0u;
        u64 hh = h2 + c;
        top += hh < c ? 1u : 0u;
        s1 += hh;
        s2 += top + (s1 < hh ? 1u : 0u);
    }
    c0[k] = s0; c1[k] = s1; c2[k] = s2;
}

__device__ __forceinline__ void add3(u64& a0, u64& a1, u64& a2, u64 b0, u64 b1, u64 b2) {
    u64 r0 = a0 + b0;
    u64 c = r0 < a0 ? 1u : 0u;
    u64 t = a1 + b1;
    u64 c1 = t < a1 ? 1u : 0u;
    u64 r1 = t + c;
    c1 += r1 < t ? 1u : 0u;
    a0 = r0; a1 = r1; a2 = a2 + b2 + c1;
}
__device__ __forceinline__ void sub3(u64& a0, u64& a1, u64& a2, u64 b0, u64 b1, u64 b2) {
    u64 r0 = a0 - b0;
    u64 br = a0 < b0 ? 1u : 0u;
    u64 t = a1 - b1;
    u64 br1 = a1 < b1 ? 1u : 0u;
    u64 r1 = t - br;
    br1 += t < br ? 1u : 0u;
    a0 = r0; a1 = r1; a2 = a2 - b2 - br1;
}

// C_k = z0_k + zm_{k-h} - z0_{k-h} - z2_{k-h} + z2_{k-2h}, arithmetic mod 2^192
// (the true column is in [0, 2^192)).  z arrays are three word planes.
__global__ __launch_bounds__(256) void k_kara_combine(const u64* z0, u64 n0, const u64* zm, u64 nm, const u64* z2,
                                                      u64 n2, u64 h, u64* c, u64 nc) {
    u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc) return;
    u64 r0 = 0, r1 = 0, r2 = 0;
    if (k < n0) add3(r0, r1, r2, z0[k], z0[n0 + k], z0[2 * n0 + k]);
    if (k >= h && k - h < nm) add3(r0, r1, r2, zm[k - h], zm[nm + k - h], zm[2 * nm + k - h]);
    if (k >= h && k - h < n0) sub3(r0, r1, r2, z0[k - h], z0[n0 + k - h], z0[2 * n0 + k - h]);
    if (k >= h && k - h < n2) sub3(r0, r1, r2, z2[k - h], z2[n2 + k - h], z2[2 * n2 + k - h]);
    if (k >= 2 * h && k - 2 * h < n2) add3(r0, r1, r2, z2[k - 2 * h], z2[n2 + k - 2 * h], z2[2 * n2 + k - 2 * h]);
    c[k] = r0; c[nc + k] = r1; c[2 * nc + k] = r2;
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

void launch_twiddles(u64* tw, u64 n, const RootBases& rb, int pr, int inv, hipStream_t st) {
    hipLaunchKernelGGL(k_twiddles, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, tw, n, rb, inv, pr);
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
    const u64 tiles = crt_tiles_single(D.n_out);
    const bool sh = crt_tile_words_single(D.n_out) == CRT_SHORT_TILE_WORDS;
    static_assert(CRT_SHORT_TILE_WORDS == 2 * 256, "short CRT tiles are two rounds");
    if (sh) hipLaunchKernelGGL(k_crt_tile<2>, dim3((unsigned)tiles), dim3(256), 0, st, A);
    else hipLaunchKernelGGL(k_crt_tile<CRT_TILE_ROUNDS>, dim3((unsigned)tiles), dim3(256), 0, st, A);
    if (tiles > 1) {
        hipLaunchKernelGGL(k_carry_scan, dim3(1), dim3(1024), 0, st, D.tile_fn, tiles, D.tile_cin);
        if (sh)
            hipLaunchKernelGGL(k_carry_fix<2>, dim3((unsigned)tiles), dim3(256), 0, st, D.out, D.n_out,
                               (const u8*)D.tile_cin);
        else
            hipLaunchKernelGGL(k_carry_fix<CRT_TILE_ROUNDS>, dim3((unsigned)tiles), dim3(256), 0, st, D.out,
                               D.n_out, (const u8*)D.tile_cin);
    }
}

void launch_school_cols(const u64* a, u64 an, const u64* b, u64 bn, u64* c0, u64* c1, u64* c2,
                        hipStream_t st) {
    u64 n = an + bn - 1;
    hipLaunchKernelGGL(k_school_cols, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, an, b, bn, c0, c1,
                       c2);
}

void launch_kara(const u64* a, u64 an, const u64* b, u64 bn, u64* ws, u64* cols, hipStream_t st) {
    // an >= bn >= 2; h = ceil(bn/2); workspace layout documented in nnh_internal.hpp
    u64 h = (bn + 1) / 2;
    u64 ns = std::max(h, an - h), nt = h;
    u64 *s_lo = ws, *s_hi = s_lo + ns, *t_lo = s_hi + ns, *t_hi = t_lo + nt;
    u64 n0 = 2 * h - 1, n2 = (an - h) + (bn - h) - 1, nm = ns + nt - 1;
    u64 *z0 = t_hi + nt, *z2 = z0 + 3 * n0, *zm = z2 + 3 * n2;
    u64 m = std::max(ns, nt);
    hipLaunchKernelGGL(k_kara_prep, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, a, an, b, bn, h, s_lo, s_hi,
                       ns, t_lo, t_hi, nt);
    launch_school_cols(a, h, b, h, z0, z0 + n0, z0 + 2 * n0, st);
    launch_school_cols(a + h, an - h, b + h, bn - h, z2, z2 + n2, z2 + 2 * n2, st);
    hipLaunchKernelGGL(k_cols_wide, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, st, s_lo, s_hi, ns, t_lo, t_hi,
                       nt, zm, zm + nm, zm + 2 * nm);
    u64 nc = an + bn - 1;
    hipLaunchKernelGGL(k_kara_combine, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, z0, n0, zm, nm, z2, n2, h,
                       cols, nc);
}

u64 kara_ws_words(u64 an, u64 bn) {
    u64 h = (bn + 1) / 2;
    u64 ns = std::max(h, an - h), nt = h;
    u64 n0 = 2 * h - 1, n2 = (an - h) + (bn - h) - 1, nm = ns + nt - 1;
    return 2 * ns + 2 * nt + 3 * (n0 + n2 + nm);
}

void launch_cyclic_fold(u64* w, u64 L, hipStream_t st) {
    hipLaunchKernelGGL(k_cyclic_fold, dim3(1), dim3(256), 0, st, w, L);
}

void launch_school_batch(const BJob* jobs, const u32* sj, u32 nsj, u64 ncols, hipStream_t st) {
    u64 blocks = (ncols + 255) / 256;
    hipLaunchKernelGGL(k_school_batch, dim3(persist_grid(blocks, 8192)), dim3(256), 0, st, jobs, sj, nsj, ncols);
}

void launch_crt_batch(const BJob* jobs, u32 njobs, u64 ntiles, u8* tile_fn, u8* tile_cin, const CrtConsts& C,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_crt_batch, dim3(persist_grid(ntiles, 8192)), dim3(256), 0, st, jobs, njobs, ntiles, tile_fn,
                       C);
    hipLaunchKernelGGL(k_carry_scan_batch, dim3(persist_grid(njobs, 1024)), dim3(1024), 0, st, jobs, njobs, ntiles,
                       tile_fn, tile_cin);
    hipLaunchKernelGGL(k_carry_fix_batch, dim3(persist_grid(ntiles, 8192)), dim3(256), 0, st, jobs, njobs, ntiles,
                       (const u8*)tile_cin);
}

static_assert(256 * CRT_TILE_ROUNDS == CRT_TILE_WORDS, "CRT tile size");

}  // namespace nnh

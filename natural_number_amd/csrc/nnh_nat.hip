// nnh_nat.hip — limb arithmetic on device-resident naturals: batched
// modular add/sub with tile carry functions, shifts, complement, searches,
// and the kernels of the binary-splitting and radix-conversion workloads
// (reference natural.h:841-857 and 904-1039).
#include <algorithm>

#include "nnh_carry.hpp"
#include "nnh_nat.hpp"
#include "nnh_tile.hpp"

namespace nnh {

// ---------------------------------------------------------------------------
// Batched modular addition.  Tiles of 4096 words resolve their carries with
// wave ballots (the same lookahead as the CRT: generate = overflow,
// propagate = all-ones word) assuming carry-in 0 and publish their carry
// function; one scan per job applies the job's carry-in and one fix-up pass
// adds the tile carry-ins.
// ---------------------------------------------------------------------------
__device__ __forceinline__ u32 find_add_job(const AddJob* jobs, u32 njobs, u64 tile) {
    u32 lo = 0, hi = njobs - 1;
    while (lo < hi) {
        u32 mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile0 <= tile) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// jobs == nullptr: the single job `one` passed by value.
__global__ __launch_bounds__(256) void k_add_tiles(const AddJob* __restrict__ jobs, AddJob one, u32 njobs, u64 ntiles,
                                                   u8* tile_fn) {
    __shared__ u32 wco[2][4];
    const u32 t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (u64 it = blockIdx.x; it < ntiles; it += gridDim.x) {
        const AddJob J = jobs ? jobs[find_add_job(jobs, njobs, it)] : one;
        const u64 tile_base = (it - J.tile0) * LIMB_TILE;
        u32 cin0 = 0, cin1 = 1;
        for (int r = 0; r < CRT_TILE_ROUNDS; r++) {
            const u64 base = tile_base + (u64)r * 256;
            if (base >= J.n) break;  // uniform
            const u64 k = base + t;
            u64 x = k < J.na ? ld_g(&J.a[k]) : 0;
            u64 y = k < J.nb ? ld_g(&J.b[k]) : 0;
            if (J.sub) y = ~y;
            u64 e = x + y;
            // positions past n pass the carry through, so the tile's carry
            // out is the carry out of word n-1
            bool g = k < J.n && e < x;
            bool pp = k >= J.n || e == ~0ull;
            u64 G = __ballot(g), P = __ballot(pp), Am = G | P;
            u64 s0 = Am + G;
            u32 co0 = s0 < Am ? 1u : 0u;
            u64 s1 = s0 + 1;
            u32 co1 = co0 | (s1 == 0 ? 1u : 0u);
            __syncthreads();  // previous round's readers of wco are done
            if (lane == 0) { wco[0][wv] = co0; wco[1][wv] = co1; }
            __syncthreads();
            // wave functions composed in a loop of constant length over
            // wave-uniform values; c0 = the carry into this wave
            const u32 wvu = __builtin_amdgcn_readfirstlane(wv);
            u32 c0 = cin0, n0 = cin0, n1 = cin1;
            for (u32 w = 0; w < 4; w++) {
                if (w == wvu) c0 = n0;
                n0 = n0 ? wco[1][w] : wco[0][w];
                n1 = n1 ? wco[1][w] : wco[0][w];
            }
            u64 sw = c0 ? s1 : s0;
            u32 cl = (u32)(((sw ^ P) >> lane) & 1u);
            if (k < J.n) st_g(&J.out[k], e + cl);
            cin0 = n0;
            cin1 = n1;
        }
        if (t == 0) tile_fn[it] = (u8)(cin0 | (cin1 << 1));
    }
}

__global__ __launch_bounds__(1024) void k_add_scan(const AddJob* __restrict__ jobs, AddJob one, u32 njobs,
                                                   u64 ntiles, const u8* tile_fn, u8* tile_cin) {
    __shared__ u32 chunk_fn[1024];
    for (u32 j = blockIdx.x; j < njobs; j += gridDim.x) {
        const AddJob J = jobs ? jobs[j] : one;
        const u64 t0 = J.tile0, t1 = j + 1 < njobs ? jobs[j + 1].tile0 : ntiles;
        const u32 cin = J.sub ? 1u : (J.cin_ptr ? ld_g(J.cin_ptr) : J.cin);
        __syncthreads();
        carry_scan_body(tile_fn + t0, t1 - t0, tile_cin + t0, chunk_fn, cin, J.cout);
    }
}

__global__ __launch_bounds__(256) void k_add_fix(const AddJob* __restrict__ jobs, AddJob one, u32 njobs, u64 ntiles,
                                                 const u8* tile_cin) {
    __shared__ u32 first_nz;
    for (u64 it = blockIdx.x; it < ntiles; it += gridDim.x) {
        if (!tile_cin[it]) continue;  // uniform
        const AddJob J = jobs ? jobs[find_add_job(jobs, njobs, it)] : one;
        __syncthreads();
        carry_fix_body(J.out, J.n, it - J.tile0, first_nz);
    }
}

u64 add_tiles_needed(const AddJob* jobs, u64 njobs) {
    u64 t = 0;
    for (u64 j = 0; j < njobs; j++) t += (jobs[j].n + LIMB_TILE - 1) / LIMB_TILE;
    return t;
}

static void add_launch(const AddJob* d_jobs, const AddJob& one, u32 njobs, u64 ntiles, u8* d_tiles,
                       hipStream_t st) {
    hipLaunchKernelGGL(k_add_tiles, dim3(persist_grid(ntiles, 8192)), dim3(256), 0, st, d_jobs, one, njobs, ntiles,
                       d_tiles);
    hipLaunchKernelGGL(k_add_scan, dim3(persist_grid(njobs, 1024)), dim3(1024), 0, st, d_jobs, one, njobs, ntiles,
                       (const u8*)d_tiles, d_tiles + ntiles);
    hipLaunchKernelGGL(k_add_fix, dim3(persist_grid(ntiles, 8192)), dim3(256), 0, st, d_jobs, one, njobs, ntiles,
                       (const u8*)(d_tiles + ntiles));
}
void launch_add_batch(AddJob* d_jobs, u32 njobs, u64 ntiles, u8* d_tiles, hipStream_t st) {
    add_launch(d_jobs, AddJob{}, njobs, ntiles, d_tiles, st);
}
void launch_add_one(AddJob job, u8* d_tiles, hipStream_t st) {
    job.tile0 = 0;
    add_launch(nullptr, job, 1, (job.n + LIMB_TILE - 1) / LIMB_TILE, d_tiles, st);
}

// ---------------------------------------------------------------------------
// Word-parallel helpers
// ---------------------------------------------------------------------------
__global__ void k_not(u64* out, const u64* a, u64 n) {
    for (u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (u64)gridDim.x * blockDim.x) out[k] = ~a[k];
}
// out[k] = splitmix64 output for state seed + (k + 1) * gamma (random-looking words)
__global__ void k_fill_hash(u64* out, u64 n, u64 seed) {
    for (u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (u64)gridDim.x * blockDim.x) {
        u64 z = seed + (k + 1) * 0x9e3779b97f4a7c15ull;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        out[k] = z ^ (z >> 31);
    }
}
__global__ void k_fill(u64* out, u64 n, u64 v) {
    for (u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (u64)gridDim.x * blockDim.x) out[k] = v;
}
__global__ void k_shl_bits(u64* out, const u64* a, u64 n, u32 b) {
    for (u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x; k <= n; k += (u64)gridDim.x * blockDim.x) {
        u64 lo = k < n ? a[k] << b : 0;
        u64 hi = k > 0 ? a[k - 1] >> (64 - b) : 0;
        out[k] = lo | hi;
    }
}
__global__ void k_shr_bits(u64* out, const u64* a, u64 n, u32 b) {
    for (u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (u64)gridDim.x * blockDim.x) {
        u64 lo = a[k] >> b;
        u64 hi = k + 1 < n ? a[k + 1] << (64 - b) : 0;
        out[k] = lo | hi;
    }
}
__global__ void k_set_word(u64* p, u64 v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}
// Highest qualifying index, scanned top-down in chunks: a block whose chunk
// lies below an index already found exits at once, so normalising a number
// whose top word is nonzero reads one chunk.
constexpr u64 FIND_CHUNK = 2048;

__global__ __launch_bounds__(256) void k_find(const u64* a, const u64* b, u64 n, int mode, u64* res3) {
    __shared__ u32 skip;
    for (u64 c = blockIdx.x; c * FIND_CHUNK < n; c += gridDim.x) {
        const u64 hi = n - c * FIND_CHUNK;
        const u64 lo = hi > FIND_CHUNK ? hi - FIND_CHUNK : 0;
        if (threadIdx.x == 0)
            skip = __hip_atomic_load((unsigned long long*)res3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > hi;
        __syncthreads();
        if (skip) return;  // every later chunk of this block is lower still
        unsigned long long best = 0;
        for (u64 k = lo + threadIdx.x; k < hi; k += blockDim.x) {
            u64 x = a[k];
            bool hit = mode == 0 ? x != 0 : (mode == 1 ? x != ~0ull : x != b[k]);
            if (hit) best = k + 1;
        }
        for (int off = 32; off > 0; off >>= 1) {
            unsigned long long o = __shfl_xor(best, off);
            best = o > best ? o : best;
        }
        if ((threadIdx.x & 63) == 0 && best) atomicMax((unsigned long long*)res3, best);
        __syncthreads();
    }
}
// Publish {index, a[index-1], b[index-1]} to host-visible memory with a
// sequence number, and clear the device result for the next search.
__global__ void k_fetch(u64* res3, const u64* a, const u64* b, u64* hout, u64 seq) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    u64 i = res3[0];
    hout[0] = i;
    hout[1] = i ? a[i - 1] : 0;
    hout[2] = (i && b) ? b[i - 1] : 0;
    res3[0] = 0;
    __threadfence_system();
    hout[3] = seq;
}
// Publish up to three words of p to host-visible memory with a sequence number.
__global__ void k_peek(const u64* p, u32 n, u64* hout, u64 seq) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (u32 i = 0; i < n; i++) hout[i] = p[i];
    __threadfence_system();
    hout[3] = seq;
}

static unsigned grid_for(u64 n) { return persist_grid((n + 255) / 256, 4096); }

void launch_not(u64* out, const u64* a, u64 n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_not, dim3(grid_for(n)), dim3(256), 0, st, out, a, n);
}
void launch_fill(u64* out, u64 n, u64 v, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(256), 0, st, out, n, v);
}
void launch_fill_hash(u64* out, u64 n, u64 seed, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_hash, dim3((unsigned)std::min<u64>((n + 255) / 256, 8192)), dim3(256), 0, st, out, n, seed);
}
void launch_shl_bits(u64* out, const u64* a, u64 n, u32 b, hipStream_t st) {
    hipLaunchKernelGGL(k_shl_bits, dim3(grid_for(n + 1)), dim3(256), 0, st, out, a, n, b);
}
void launch_shr_bits(u64* out, const u64* a, u64 n, u32 b, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_shr_bits, dim3(grid_for(n)), dim3(256), 0, st, out, a, n, b);
}
void launch_set_word(u64* p, u64 v, hipStream_t st) { hipLaunchKernelGGL(k_set_word, dim3(1), dim3(64), 0, st, p, v); }
void launch_find(const u64* a, const u64* b, u64 n, int mode, u64* res3, u64* hout, u64 seq, hipStream_t st) {
    if (n)
        hipLaunchKernelGGL(k_find, dim3(persist_grid((n + FIND_CHUNK - 1) / FIND_CHUNK, 1024)), dim3(256), 0, st, a, b,
                           n, mode, res3);
    hipLaunchKernelGGL(k_fetch, dim3(1), dim3(64), 0, st, res3, a, b, hout, seq);
}
void launch_peek(const u64* p, u32 n, u64* hout, u64 seq, hipStream_t st) {
    hipLaunchKernelGGL(k_peek, dim3(1), dim3(64), 0, st, p, n, hout, seq);
}

// ---------------------------------------------------------------------------
// Work-group multiword helpers (256 threads; operands in LDS or global).
// ---------------------------------------------------------------------------
// x[0..n) = x + (sub ? ~y : y) + cin, y zero-extended beyond ny; returns the
// carry out of word n-1 (uniform).
__device__ u32 wg_addsub(u64* x, u64 n, const u64* y, u64 ny, u32 sub, u32 cin, u32 (*wco)[4]) {
    const u32 t = threadIdx.x, lane = t & 63, wv = t >> 6;
    u32 c = cin;
    for (u64 base = 0; base < n; base += 256) {
        const u64 k = base + t;
        u64 xv = k < n ? x[k] : 0;
        u64 yv = k < ny ? y[k] : 0;
        if (sub) yv = ~yv;
        u64 e = xv + yv;
        bool g = k < n && e < xv;
        bool pp = k >= n || e == ~0ull;  // past the end: carry passes through
        u64 G = __ballot(g), P = __ballot(pp), Am = G | P;
        u64 s0 = Am + G;
        u32 co0 = s0 < Am ? 1u : 0u;
        u64 s1 = s0 + 1;
        u32 co1 = co0 | (s1 == 0 ? 1u : 0u);
        __syncthreads();
        if (lane == 0) { wco[0][wv] = co0; wco[1][wv] = co1; }
        __syncthreads();
        // wave functions composed over wave-uniform values (constant trip
        // count): cw = the carry into this wave, c = out of the round
        const u32 wvu = __builtin_amdgcn_readfirstlane(wv);
        u32 cw = c;
        for (u32 w = 0; w < 4; w++) {
            if (w == wvu) cw = c;
            c = c ? wco[1][w] : wco[0][w];
        }
        u64 sw = cw ? s1 : s0;
        u32 cl = (u32)(((sw ^ P) >> lane) & 1u);
        if (k < n) x[k] = e + cl;
    }
    __syncthreads();
    return c;
}

// Sign of x - y (x: nx words, y: ny words, both zero-extended); uniform.
__device__ int wg_cmp(const u64* x, u64 nx, const u64* y, u64 ny, u32* sh) {
    const u32 t = threadIdx.x;
    u64 top = nx > ny ? nx : ny;
    while (top > 0) {
        const u64 lo = top > 256 ? top - 256 : 0;
        __syncthreads();
        if (t == 0) *sh = 0;
        __syncthreads();
        const u64 k = lo + t;
        if (k < top) {
            u64 a = k < nx ? x[k] : 0, b = k < ny ? y[k] : 0;
            if (a != b) atomicMax(sh, (u32)(k - lo + 1));
        }
        __syncthreads();
        const u32 h = *sh;
        if (h) {
            const u64 kk = lo + h - 1;
            u64 a = kk < nx ? x[kk] : 0, b = kk < ny ? y[kk] : 0;
            return a > b ? 1 : -1;
        }
        top = lo;
    }
    return 0;
}

// out[0..n_out) = a * b (+ add) by column sums (3 words each) and the CRT
// carry lookahead in columns mode; n_out <= 4096 and the value must fit.
__device__ void wg_mul(u64* out, u64 n_out, const u64* a, u64 na, const u64* b, u64 nb, const u64* add, u64 n_add,
                       u64* c0, u64* c1, u64* c2, u64 (*xs)[256 + 3], u32 (*wco)[4], u8* fn) {
    const u64 nc = na + nb - 1;
    for (u64 col = threadIdx.x; col < nc; col += 256) {
        const u64 i0 = col >= nb - 1 ? col - (nb - 1) : 0;
        const u64 i1 = col < na - 1 ? col : na - 1;
        u64 lo = 0, hi = 0, top = 0;
        for (u64 i = i0; i <= i1; i++) {
            u64 pl, ph;
            mul_wide(a[i], b[col - i], pl, ph);
            lo += pl;
            u64 c = lo < pl ? 1 : 0;
            hi += ph;
            top += hi < ph ? 1 : 0;
            hi += c;
            top += hi < c ? 1 : 0;
        }
        c0[col] = lo;
        c1[col] = hi;
        c2[col] = top;
    }
    __syncthreads();
    CrtArgs A;
    A.r[0] = c0;
    A.r[1] = c1;
    A.r[2] = c2;
    A.columns = 1;
    A.n_res = nc;
    A.out = out;
    A.n_out = n_out;
    A.addend = add;
    A.n_add = n_add;
    A.tile_fn = nullptr;
    A.C = CrtConsts{};
    crt_tile_body<false>(A, 0, fn, xs, wco);  // columns, addend and output may be LDS
    __syncthreads();
}

// ---------------------------------------------------------------------------
// e by binary splitting: num/den of a range [b, e) satisfy
//   leaf [b, b+2): num = b+2, den = (b+1) b (den = 1 for b = 0),
//   combine: num = num_l den_r + num_r, den = den_l den_r,
// which is associative, so a block of pairs folds right to left with
// one-word multipliers.
// ---------------------------------------------------------------------------
constexpr int E_LEAF_WORDS = 14;

__global__ __launch_bounds__(256) void k_e_leaves(u64* num, u64* den, u64 nblocks, u64 stride) {
    const u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nblocks) return;
    const u64 b0 = 32 * j;
    u64 N[E_LEAF_WORDS], D[E_LEAF_WORDS];
#pragma unroll
    for (int k = 0; k < E_LEAF_WORDS; k++) N[k] = D[k] = 0;
    N[0] = b0 + 32;
    D[0] = (b0 + 31) * (b0 + 30);
#pragma unroll
    for (int p = 14; p >= 0; p--) {
        const u64 beg = b0 + 2 * (u64)p;
        const u64 np = beg + 2;
        const u64 dp = beg ? (beg + 1) * beg : 1;
        u64 c1 = 0, c2 = 0;
#pragma unroll
        for (int k = 0; k < E_LEAF_WORDS; k++) {
            u64 lo, hi;
            mul_wide(np, D[k], lo, hi);
            u64 s = lo + N[k];
            hi += s < lo ? 1 : 0;
            u64 s2 = s + c1;
            hi += s2 < s ? 1 : 0;
            N[k] = s2;
            c1 = hi;
            mul_wide(dp, D[k], lo, hi);
            u64 d2 = lo + c2;
            hi += d2 < lo ? 1 : 0;
            D[k] = d2;
            c2 = hi;
        }
    }
    u64* on = num + j * stride;
    u64* od = den + j * stride;
#pragma unroll
    for (int k = 0; k < E_LEAF_WORDS; k++)
        if ((u64)k < stride) {
            on[k] = N[k];
            od[k] = D[k];
        }
    for (u64 k = E_LEAF_WORDS; k < stride; k++) on[k] = od[k] = 0;
}

constexpr int E_WG_MAX = 256;

// WG: largest operand (words) the instance holds in LDS; the 512-word
// instance (46 KB of LDS) takes the levels the 256-word one (23 KB, twice
// the work-groups per CU) cannot.
template <int WG>
__global__ __launch_bounds__(256) void k_e_combine(const u64* __restrict__ num, const u64* __restrict__ den, u64 S,
                                                   u64* num2, u64* den2, u64 S2, u64 ncomb) {
    __shared__ u64 nl[WG], dl[WG], nr[WG], dr[WG];
    __shared__ u64 c0[2 * WG], c1[2 * WG], c2[2 * WG];
    __shared__ u64 xs[3][256 + 3];
    __shared__ u32 wco[2][4];
    __shared__ u8 fn;
    const u32 t = threadIdx.x;
    for (u64 c = blockIdx.x; c < ncomb; c += gridDim.x) {
        const u64 L = 2 * c * S, R = (2 * c + 1) * S;
        __syncthreads();
        for (u64 k = t; k < S; k += 256) {
            nl[k] = num[L + k];
            dl[k] = den[L + k];
            nr[k] = num[R + k];
            dr[k] = den[R + k];
        }
        __syncthreads();
        wg_mul(num2 + c * S2, S2, nl, S, dr, S, nr, S, c0, c1, c2, xs, wco, &fn);
        wg_mul(den2 + c * S2, S2, dl, S, dr, S, nullptr, 0, c0, c1, c2, xs, wco, &fn);
    }
}

void launch_e_leaves(u64* num, u64* den, u64 nblocks, u64 stride, hipStream_t st) {
    hipLaunchKernelGGL(k_e_leaves, dim3((unsigned)((nblocks + 255) / 256)), dim3(256), 0, st, num, den, nblocks,
                       stride);
}
void launch_e_combine(const u64* num, const u64* den, u64 S, u64* num2, u64* den2, u64 S2, u64 ncomb,
                      hipStream_t st) {
    const dim3 grid(persist_grid(ncomb, 16384));
    if (S <= E_WG_MAX)
        hipLaunchKernelGGL(k_e_combine<E_WG_MAX>, grid, dim3(256), 0, st, num, den, S, num2, den2, S2, ncomb);
    else
        hipLaunchKernelGGL(k_e_combine<2 * E_WG_MAX>, grid, dim3(256), 0, st, num, den, S, num2, den2, S2, ncomb);
}

// ---------------------------------------------------------------------------
// pi by binary splitting (reference calc_pi, natural.h:874-902).
// Term b: den = b^3 C, numc = (2b-1)(6b-1)(6b-5), num = (A b + B) numc with
// C = 640320^3 / 24, A = 545140134, B = 13591409.  Pairs combine with
// num = |numc_l num_r - num_l den_r|, larger ranges with
// num = num_l den_r + numc_l num_r; den and numc multiply.
// ---------------------------------------------------------------------------
typedef unsigned __int128 u128;

// 128 x 128 -> 256-bit product, little-endian words.
__device__ __forceinline__ void mul_2x2(u64 a0, u64 a1, u64 b0, u64 b1, u64 (&r)[4]) {
    u64 l00, h00, l01, h01, l10, h10, l11, h11;
    mul_wide(a0, b0, l00, h00);
    mul_wide(a0, b1, l01, h01);
    mul_wide(a1, b0, l10, h10);
    mul_wide(a1, b1, l11, h11);
    r[0] = l00;
    u128 m = (u128)h00 + l01 + l10;
    r[1] = (u64)m;
    u128 n = (m >> 64) + (u128)h01 + h10 + l11;
    r[2] = (u64)n;
    r[3] = (u64)(n >> 64) + h11;
}

// b < 2^21: den < 2^117, numc < 2^70, num < 2^121.
__device__ __forceinline__ void pi_term(u64 b, u128& num, u128& den, u128& numc) {
    u64 lo, hi;
    mul_wide(b * b * b, 0x26dd041d878000ull, lo, hi);
    den = (u128)hi << 64 | lo;
    mul_wide((2 * b - 1) * (6 * b - 1), 6 * b - 5, lo, hi);
    numc = (u128)hi << 64 | lo;
    const u64 f = 0x207e2da6ull * b + 0xcf6371ull;  // < 2^51
    mul_wide(f, lo, lo, hi);                         // f * numc_lo
    hi += f * (u64)(numc >> 64);                     // + f * numc_hi * 2^64 (numc_hi < 2^6)
    num = (u128)hi << 64 | lo;
}

__global__ __launch_bounds__(256) void k_pi_pairs(u64 first, u64 npairs, u64* num, u64* den, u64* numc, u64 S) {
    const u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npairs) return;
    u128 nl, dl, cl, nr, dr, cr;
    pi_term(first + 2 * j, nl, dl, cl);
    pi_term(first + 2 * j + 1, nr, dr, cr);
    u64 A[4], B[4], D[4], C[4];
    mul_2x2((u64)cl, (u64)(cl >> 64), (u64)nr, (u64)(nr >> 64), A);  // numc_l num_r
    mul_2x2((u64)nl, (u64)(nl >> 64), (u64)dr, (u64)(dr >> 64), B);  // num_l den_r
    mul_2x2((u64)dl, (u64)(dl >> 64), (u64)dr, (u64)(dr >> 64), D);
    mul_2x2((u64)cl, (u64)(cl >> 64), (u64)cr, (u64)(cr >> 64), C);
    // |A - B|
    bool a_ge = true;
    for (int k = 3; k >= 0; k--)
        if (A[k] != B[k]) {
            a_ge = A[k] > B[k];
            break;
        }
    u64 X[4], br = 0;
    for (int k = 0; k < 4; k++) {
        u64 x = a_ge ? A[k] : B[k], y = a_ge ? B[k] : A[k];
        u64 d = x - y - br;
        br = (x < y || (x == y && br)) ? 1 : 0;
        X[k] = d;
    }
    u64* on = num + j * S;
    u64* od = den + j * S;
    u64* oc = numc + j * S;
    for (u64 k = 0; k < S; k++) {
        on[k] = k < 4 ? X[k] : 0;
        od[k] = k < 4 ? D[k] : 0;
        oc[k] = k < 4 ? C[k] : 0;
    }
}

__global__ __launch_bounds__(256) void k_pi_combine(const u64* __restrict__ num, const u64* __restrict__ den,
                                                    const u64* __restrict__ numc, u64 S, u64* num2, u64* den2,
                                                    u64* numc2, u64 S2, u64 ncomb) {
    __shared__ u64 nl[E_WG_MAX], dl[E_WG_MAX], cl[E_WG_MAX], nr[E_WG_MAX], dr[E_WG_MAX], cr[E_WG_MAX];
    __shared__ u64 X[2 * E_WG_MAX];
    __shared__ u64 c0[2 * E_WG_MAX], c1[2 * E_WG_MAX], c2[2 * E_WG_MAX];
    __shared__ u64 xs[3][256 + 3];
    __shared__ u32 wco[2][4];
    __shared__ u8 fn;
    const u32 t = threadIdx.x;
    for (u64 c = blockIdx.x; c < ncomb; c += gridDim.x) {
        const u64 L = 2 * c * S, R = (2 * c + 1) * S;
        __syncthreads();
        for (u64 k = t; k < S; k += 256) {
            nl[k] = num[L + k];
            dl[k] = den[L + k];
            cl[k] = numc[L + k];
            nr[k] = num[R + k];
            dr[k] = den[R + k];
            cr[k] = numc[R + k];
        }
        __syncthreads();
        const u64 nx = 2 * S < S2 ? 2 * S : S2;
        wg_mul(X, nx, cl, S, nr, S, nullptr, 0, c0, c1, c2, xs, wco, &fn);
        wg_mul(num2 + c * S2, S2, nl, S, dr, S, X, nx, c0, c1, c2, xs, wco, &fn);
        wg_mul(den2 + c * S2, S2, dl, S, dr, S, nullptr, 0, c0, c1, c2, xs, wco, &fn);
        wg_mul(numc2 + c * S2, S2, cl, S, cr, S, nullptr, 0, c0, c1, c2, xs, wco, &fn);
    }
}

void launch_pi_pairs(u64 first, u64 npairs, u64* num, u64* den, u64* numc, u64 S, hipStream_t st) {
    hipLaunchKernelGGL(k_pi_pairs, dim3((unsigned)((npairs + 255) / 256)), dim3(256), 0, st, first, npairs, num, den,
                       numc, S);
}
void launch_pi_combine(const u64* num, const u64* den, const u64* numc, u64 S, u64* num2, u64* den2, u64* numc2,
                       u64 S2, u64 ncomb, hipStream_t st) {
    hipLaunchKernelGGL(k_pi_combine, dim3(persist_grid(ncomb, 16384)), dim3(256), 0, st, num, den, numc, S, num2,
                       den2, numc2, S2, ncomb);
}

// ---------------------------------------------------------------------------
// Product trees of one-word factors (the prime products of the reference
// factorial, natural.h:795-839).  Leaves: thread j multiplies factors
// [16j, 16j+16) (each < 2^32) into TREE_LEAF_WORDS words; levels with
// operands of at most 256 words run one work-group per product.
// ---------------------------------------------------------------------------
constexpr int TREE_LEAF = 16;
constexpr int TREE_LEAF_WORDS = 9;

__global__ __launch_bounds__(256) void k_tree_leaves(const u64* __restrict__ f, u64 nleaves, u64* out, u64 stride) {
    const u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nleaves) return;
    u64 acc[TREE_LEAF_WORDS];
#pragma unroll
    for (int k = 0; k < TREE_LEAF_WORDS; k++) acc[k] = k == 0 ? 1 : 0;
#pragma unroll
    for (int g = 0; g < TREE_LEAF; g++) {
        const u64 m = f[j * TREE_LEAF + g];
        u64 c = 0;
#pragma unroll
        for (int k = 0; k < TREE_LEAF_WORDS; k++) {
            u64 lo, hi;
            mul_wide(acc[k], m, lo, hi);
            u64 s = lo + c;
            hi += s < lo ? 1 : 0;
            acc[k] = s;
            c = hi;
        }
    }
    u64* o = out + j * stride;
#pragma unroll
    for (int k = 0; k < TREE_LEAF_WORDS; k++)
        if ((u64)k < stride) o[k] = acc[k];
    for (u64 k = TREE_LEAF_WORDS; k < stride; k++) o[k] = 0;
}

__global__ __launch_bounds__(256) void k_tree_combine(const u64* __restrict__ in, u64 S, u64* out, u64 S2,
                                                      u64 ncomb) {
    __shared__ u64 l[E_WG_MAX], r[E_WG_MAX];
    __shared__ u64 c0[2 * E_WG_MAX], c1[2 * E_WG_MAX], c2[2 * E_WG_MAX];
    __shared__ u64 xs[3][256 + 3];
    __shared__ u32 wco[2][4];
    __shared__ u8 fn;
    const u32 t = threadIdx.x;
    for (u64 c = blockIdx.x; c < ncomb; c += gridDim.x) {
        __syncthreads();
        for (u64 k = t; k < S; k += 256) {
            l[k] = in[2 * c * S + k];
            r[k] = in[(2 * c + 1) * S + k];
        }
        __syncthreads();
        wg_mul(out + c * S2, S2, l, S, r, S, nullptr, 0, c0, c1, c2, xs, wco, &fn);
    }
}

void launch_tree_leaves(const u64* f, u64 nleaves, u64* out, u64 stride, hipStream_t st) {
    hipLaunchKernelGGL(k_tree_leaves, dim3((unsigned)((nleaves + 255) / 256)), dim3(256), 0, st, f, nleaves, out,
                       stride);
}
void launch_tree_combine(const u64* in, u64 S, u64* out, u64 S2, u64 ncomb, hipStream_t st) {
    hipLaunchKernelGGL(k_tree_combine, dim3(persist_grid(ncomb, 16384)), dim3(256), 0, st, in, S, out, S2, ncomb);
}

// ---------------------------------------------------------------------------
// Radix conversion.
// Division of a two-word numerator by 10^18 with a precomputed reciprocal:
// Moller and Granlund, "Improved division by invariant integers" (IEEE
// Trans. Computers 60(2), 2011), Algorithm 4, on the normalised divisor
// d = 10^18 * 2^4 with v = floor((2^128 - 1) / d) - 2^64.
// ---------------------------------------------------------------------------
constexpr u64 DEC_DN = 0xde0b6b3a76400000ull;   // 10^18 << 4
constexpr u64 DEC_V = 0x2725dd1d243aba0eull;

// (u1 * 2^64 + u0) / d for normalised d (top bit set), u1 < d, and
// v = floor((2^128 - 1) / d) - 2^64: returns the quotient, r <- remainder.
__device__ __forceinline__ u64 div_preinv(u64 u1, u64 u0, u64 d, u64 v, u64& r) {
    u64 q0, q1;
    mul_wide(v, u1, q0, q1);
    u64 s = q0 + u0;
    q1 += (s < q0 ? 1 : 0) + u1 + 1;
    q0 = s;
    u64 rr = u0 - q1 * d;
    if (rr > q0) {
        q1--;
        rr += d;
    }
    if (rr >= d) {
        q1++;
        rr -= d;
    }
    r = rr;
    return q1;
}

// (r * 2^64 + u) / 10^18 for r < 10^18: returns the quotient word, r <- remainder.
__device__ __forceinline__ u64 div_1e18(u64& r, u64 u) {
    u64 rr;
    u64 q = div_preinv((r << 4) | (u >> 60), u << 4, DEC_DN, DEC_V, rr);
    r = rr >> 4;
    return q;
}

// q[0..n) = a / d0, *rem = a mod d0 for a one-word divisor (reference
// div_base, natural.h:566-574); d = d0 << l normalised, the dividend taken
// as a << l (its words u_k below, plus a top remainder of the bits shifted
// out).  The remainder chain runs top-down, one word after another; across
// chunks of DIVW_CHUNK words it is an affine map r -> (r B^len + rho) mod d
// of the remainder entering the chunk (B = 2^64, rho the chunk's own value
// mod d), and affine maps compose.  So: every chunk's rho in parallel
// (k_divw_local), the remainder entering every chunk by a scan of the
// composed maps in one work-group (k_divw_scan), then every chunk divides
// its words in parallel from that remainder (k_divw_final).
constexpr u64 DIVW_CHUNK = 64;
constexpr u64 DIVW_SERIAL_MAX = 2048;  // up to this many words one thread does it all

__device__ __forceinline__ u64 shifted_word(const u64* a, u64 k, u32 l) {
    u64 u = a[k] << l;
    if (l && k > 0) u |= a[k - 1] >> (64 - l);
    return u;
}
__device__ __forceinline__ u64 mulmod_pre(u64 x, u64 y, u64 d, u64 v) {  // x, y < d
    u64 lo, hi, r;
    mul_wide(x, y, lo, hi);
    (void)div_preinv(hi, lo, d, v, r);
    return r;
}
__device__ __forceinline__ u64 addmod_d(u64 x, u64 y, u64 d) {  // x, y < d, d >= 2^63
    const u64 s = x + y;
    return (s < x || s >= d) ? s - d : s;
}

__global__ void k_div_word(u64* q, const u64* a, u64 n, u64 d, u32 l, u64 v, u64* rem) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    u64 r = l ? a[n - 1] >> (64 - l) : 0;
    for (u64 k = n; k-- > 0;) q[k] = div_preinv(r, shifted_word(a, k, l), d, v, r);
    *rem = r >> l;
}

__global__ void k_divw_local(const u64* a, u64 n, u64 d, u32 l, u64 v, u64 nch, u64* rho) {
    for (u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x; j < nch; j += (u64)gridDim.x * blockDim.x) {
        const u64 lo = j * DIVW_CHUNK, hi = lo + DIVW_CHUNK < n ? lo + DIVW_CHUNK : n;
        u64 r = 0;
        for (u64 k = hi; k-- > lo;) (void)div_preinv(r, shifted_word(a, k, l), d, v, r);
        rho[j] = r;
    }
}

// R[j] = the remainder entering chunk j.  beta = B^DIVW_CHUNK mod d,
// beta_top = B^(length of the top chunk) mod d.
__global__ __launch_bounds__(1024) void k_divw_scan(const u64* a, u64 n, u32 l, const u64* rho, u64 nch, u64 beta,
                                                    u64 beta_top, u64 d, u64 v, u64* R) {
    __shared__ u64 sm[2][1024], sc[2][1024];
    const u32 t = threadIdx.x;
    const u64 seg = (nch + 1023) / 1024;
    const u64 lo = t * seg < nch ? t * seg : nch, hi = lo + seg < nch ? lo + seg : nch;
    // this thread's chunks applied from the top down: r -> r m + c
    u64 m = 1, c = 0;
    for (u64 j = hi; j-- > lo;) {
        const u64 bj = j == nch - 1 ? beta_top : beta;
        m = mulmod_pre(m, bj, d, v);
        c = addmod_d(mulmod_pre(c, bj, d, v), rho[j], d);
    }
    sm[0][t] = m;
    sc[0][t] = c;
    __syncthreads();
    // suffix composition: after the steps, thread t holds every segment at or
    // above t applied from the top one down (its own last)
    int cur = 0;
    for (u32 off = 1; off < 1024; off <<= 1) {
        u64 mm = sm[cur][t], cc = sc[cur][t];
        if (t + off < 1024) {
            const u64 my = sm[cur][t + off], cy = sc[cur][t + off];  // applied first
            cc = addmod_d(mulmod_pre(cy, mm, d, v), cc, d);
            mm = mulmod_pre(my, mm, d, v);
        }
        sm[cur ^ 1][t] = mm;
        sc[cur ^ 1][t] = cc;
        __syncthreads();
        cur ^= 1;
    }
    const u64 r_init = l ? a[n - 1] >> (64 - l) : 0;
    u64 r = t == 1023 ? r_init : addmod_d(mulmod_pre(r_init, sm[cur][t + 1], d, v), sc[cur][t + 1], d);
    for (u64 j = hi; j-- > lo;) {
        R[j] = r;
        const u64 bj = j == nch - 1 ? beta_top : beta;
        r = addmod_d(mulmod_pre(r, bj, d, v), rho[j], d);
    }
}

__global__ void k_divw_final(u64* q, const u64* a, u64 n, u64 d, u32 l, u64 v, u64 nch, const u64* R, u64* rem) {
    for (u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x; j < nch; j += (u64)gridDim.x * blockDim.x) {
        const u64 lo = j * DIVW_CHUNK, hi = lo + DIVW_CHUNK < n ? lo + DIVW_CHUNK : n;
        u64 r = R[j];
        for (u64 k = hi; k-- > lo;) q[k] = div_preinv(r, shifted_word(a, k, l), d, v, r);
        if (j == 0) *rem = r >> l;
    }
}

u64 div_word_scratch_words(u64 n) { return n <= DIVW_SERIAL_MAX ? 1 : 2 * ((n + DIVW_CHUNK - 1) / DIVW_CHUNK); }

void launch_div_word(u64* q, const u64* a, u64 n, u64 d, u32 l, u64 v, u64* rem, u64* scratch, hipStream_t st) {
    if (n <= DIVW_SERIAL_MAX) {
        hipLaunchKernelGGL(k_div_word, dim3(1), dim3(64), 0, st, q, a, n, d, l, v, rem);
        return;
    }
    const u64 nch = (n + DIVW_CHUNK - 1) / DIVW_CHUNK, top_len = n - (nch - 1) * DIVW_CHUNK;
    // B mod d = 2^64 - d (d >= 2^63), and its powers
    typedef unsigned __int128 u128;
    auto powmod = [&](u64 e) {
        u64 b = 0 - d, r = 1;
        for (; e; e >>= 1, b = (u64)((u128)b * b % d))
            if (e & 1) r = (u64)((u128)r * b % d);
        return r;
    };
    u64* rho = scratch;
    u64* R = scratch + nch;
    const unsigned blocks = (unsigned)std::min<u64>((nch + 255) / 256, 4096);
    hipLaunchKernelGGL(k_divw_local, dim3(blocks), dim3(256), 0, st, a, n, d, l, v, nch, rho);
    hipLaunchKernelGGL(k_divw_scan, dim3(1), dim3(1024), 0, st, a, n, l, (const u64*)rho, nch, powmod(DIVW_CHUNK),
                       powmod(top_len), d, v, R);
    hipLaunchKernelGGL(k_divw_final, dim3(blocks), dim3(256), 0, st, q, a, n, d, l, v, nch, (const u64*)R, rem);
}

__device__ __forceinline__ void put9(char* o, u32 v) {
#pragma unroll
    for (int d = 8; d >= 0; d--) {
        o[d] = (char)('0' + v % 10u);
        v /= 10u;
    }
}
__device__ __forceinline__ void put18(char* o, u64 v) {
    u64 hi = v / 1000000000ull;
    put9(o, (u32)hi);
    put9(o + 9, (u32)(v - hi * 1000000000ull));
}

constexpr int DEC_LEAF_WORDS = 16;

__global__ __launch_bounds__(256) void k_dec_leaves(const u64* __restrict__ x, u64 stride, u64 count, int nchunks,
                                                    char* out) {
    const u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    u64 v[DEC_LEAF_WORDS];
#pragma unroll
    for (int k = 0; k < DEC_LEAF_WORDS; k++) v[k] = (u64)k < stride ? x[j * stride + k] : 0;
    char* o = out + j * 18 * (u64)nchunks;
#pragma unroll
    for (int c = 0; c < DEC_LEAF_WORDS; c++) {
        if (c < nchunks) {
            u64 r = 0;
#pragma unroll
            for (int k = DEC_LEAF_WORDS - 1; k >= 0; k--) v[k] = div_1e18(r, v[k]);
            put18(o + 18 * (nchunks - 1 - c), r);
        }
    }
}

constexpr int DEC_WG_MAX = 256;  // largest base (words) split by one work-group

__global__ __launch_bounds__(256) void k_dec_split_wg(const u64* __restrict__ x, u64 sx, u64 count,
                                                      const u64* __restrict__ base, u64 w, const u64* __restrict__ r,
                                                      u64 rn, u32 n_shl, u64* out, u64* err) {
    __shared__ u64 X[2 * DEC_WG_MAX + 16];
    __shared__ u64 Rr[DEC_WG_MAX + 16];
    __shared__ u64 Bs[DEC_WG_MAX];
    __shared__ u64 Pq[2 * DEC_WG_MAX + 32];
    __shared__ u64 H[DEC_WG_MAX];
    __shared__ u64 c0[2 * DEC_WG_MAX + 32], c1[2 * DEC_WG_MAX + 32], c2[2 * DEC_WG_MAX + 32];
    __shared__ u64 xs[3][256 + 3];
    __shared__ u32 wco[2][4];
    __shared__ u32 sh;
    __shared__ u8 fn;
    const u32 t = threadIdx.x;
    for (u64 k = t; k < rn; k += 256) Rr[k] = r[k];
    for (u64 k = t; k < w; k += 256) Bs[k] = base[k];
    const u64 xn = sx - (w - 1);               // words of x >> (w-1) words
    const u64 sbits = (w + 2) * 64 - n_shl;    // hi = (xt * r) >> sbits
    const u64 ws = sbits >> 6;
    const u32 bs = (u32)(sbits & 63);
    const u64 pn = xn + rn;
    for (u64 j = blockIdx.x; j < count; j += gridDim.x) {
        __syncthreads();
        for (u64 k = t; k < 2 * w; k += 256) X[k] = k < sx ? x[j * sx + k] : 0;
        __syncthreads();
        wg_mul(Pq, pn, X + (w - 1), xn, Rr, rn, nullptr, 0, c0, c1, c2, xs, wco, &fn);
        for (u64 k = t; k < w; k += 256) {
            u64 i = k + ws;
            u64 lo = i < pn ? Pq[i] >> bs : 0;
            u64 hi = (bs && i + 1 < pn) ? Pq[i + 1] << (64 - bs) : 0;
            H[k] = lo | hi;
        }
        __syncthreads();
        wg_mul(Pq, 2 * w, H, w, Bs, w, nullptr, 0, c0, c1, c2, xs, wco, &fn);
        u32 nb = wg_addsub(X, 2 * w, Pq, 2 * w, 1, 1, wco);  // lo = x - hi*base
        int bad = nb ? 0 : 1;
        // at most 8 corrections, the bound the batched split levels apply
        int it = 0;
        for (; it < 8 && wg_cmp(X, 2 * w, Bs, w, &sh) >= 0; it++) {
            wg_addsub(X, 2 * w, Bs, w, 1, 1, wco);
            wg_addsub(H, w, nullptr, 0, 0, 1, wco);
        }
        if (it == 8 && wg_cmp(X, 2 * w, Bs, w, &sh) >= 0) bad = 1;
        if (bad && t == 0) *err = 1;
        for (u64 k = t; k < w; k += 256) {
            out[(2 * j) * w + k] = H[k];
            out[(2 * j + 1) * w + k] = X[k];
        }
    }
}

__global__ void k_shr_pieces(u64* out, u64 so, const u64* in, u64 si, u64 n_in, u64 shift, u64 n_out, u64 count) {
    const u64 ws = shift >> 6;
    const u32 bs = (u32)(shift & 63);
    const u64 total = n_out * count;
    for (u64 g = (u64)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (u64)gridDim.x * blockDim.x) {
        const u64 j = g / n_out, k = g - j * n_out;
        const u64* p = in + j * si;
        const u64 i = k + ws;
        u64 lo = i < n_in ? p[i] >> bs : 0;
        u64 hi = (bs && i + 1 < n_in) ? p[i + 1] << (64 - bs) : 0;
        out[j * so + k] = lo | hi;
    }
}

__global__ void k_nonzero_pieces(const u64* x, u64 sx, u64 lo, u64 len, u64 count, u64* flag) {
    const u64 total = len * count;
    bool nz = false;
    for (u64 g = (u64)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (u64)gridDim.x * blockDim.x) {
        const u64 j = g / len, k = g - j * len;
        nz |= x[j * sx + lo + k] != 0;
    }
    if (__ballot(nz) && (threadIdx.x & 63) == 0) *flag = 1;
}

__global__ __launch_bounds__(256) void k_cmp_pieces(const u64* x, u64 sx, u64 nx, const u64* y, u64 ny, u64 count,
                                                    u8* flags) {
    __shared__ u32 sh;
    for (u64 j = blockIdx.x; j < count; j += gridDim.x) {
        int c = wg_cmp(x + j * sx, nx, y, ny, &sh);
        if (threadIdx.x == 0) flags[j] = c >= 0 ? 1 : 0;
    }
}

void launch_dec_leaves(const u64* x, u64 stride, u64 count, int lvl, char* out, hipStream_t st) {
    hipLaunchKernelGGL(k_dec_leaves, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, x, stride, count,
                       1 << (lvl + 1), out);
}
void launch_dec_split_wg(const u64* x, u64 sx, u64 count, const u64* base, u64 w, const u64* r, u64 rn, u32 n_shl,
                         u64* out, u64* err, hipStream_t st) {
    hipLaunchKernelGGL(k_dec_split_wg, dim3(persist_grid(count, 16384)), dim3(256), 0, st, x, sx, count, base, w, r,
                       rn, n_shl, out, err);
}
void launch_shr_pieces(u64* out, u64 so, const u64* in, u64 si, u64 n_in, u64 shift, u64 n_out, u64 count,
                       hipStream_t st) {
    u64 total = n_out * count;
    if (total)
        hipLaunchKernelGGL(k_shr_pieces, dim3(grid_for(total)), dim3(256), 0, st, out, so, in, si, n_in, shift, n_out,
                           count);
}
void launch_nonzero_pieces(const u64* x, u64 sx, u64 lo, u64 hi, u64 count, u64* flag, hipStream_t st) {
    if (hi > lo && count)
        hipLaunchKernelGGL(k_nonzero_pieces, dim3(grid_for((hi - lo) * count)), dim3(256), 0, st, x, sx, lo, hi - lo,
                           count, flag);
}
void launch_cmp_pieces(const u64* x, u64 sx, u64 nx, const u64* y, u64 ny, u64 count, u8* flags, hipStream_t st) {
    hipLaunchKernelGGL(k_cmp_pieces, dim3(persist_grid(count, 16384)), dim3(256), 0, st, x, sx, nx, y, ny, count,
                       flags);
}

}  // namespace nnh

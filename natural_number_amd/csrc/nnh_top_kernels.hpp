// nnh_top_kernels.hpp — strided "top" passes of the transform (levels
// h >= 4096): forward (fused with the limb load on the first pass) and
// inverse, for one product or a bucket of a batch.  LAZY kernels serve the
// primes p0 and p1 (grid dimension y = 2, lazy butterflies); the others
// serve p2.  Each translation unit instantiates one LAZY value.
#pragma once
#include "nnh_tile.hpp"

namespace nnh {

// The limb-loading pass asks for 6 waves per SIMD (80 VGPRs; it needs 81
// unconstrained, which would allow 5).  A forward pass over both operands of
// a product runs as one launch: blockIdx.z picks the operand (A, or B), so
// the two transforms share the GPU and the launch count halves.
// (The second operand's fields replace the first's one by one: selecting
// a whole argument struct at run time would put a copy in scratch memory.)
struct TopOperand {
    u64* data[3];
    const u64* limbs;
    u64 an;
};
template <bool INV, bool FROM_LIMBS, int NLEV, bool LAZY, bool TOP>
__global__ __launch_bounds__(256, FROM_LIMBS ? 6 : 1) void k_top_pass(TopArgs A, TopOperand B, u32 pr0) {
    __shared__ u32 lds[4096 + 256];  // 32-bit transposes: half the LDS, more work-groups per CU
    TopArgs X = A;
    if (blockIdx.z) {
        X.data[0] = B.data[0];
        X.data[1] = B.data[1];
        X.data[2] = B.data[2];
        X.limbs = B.limbs;
        X.an = B.an;
    }
    top_pass_body<INV, FROM_LIMBS, NLEV, LAZY, TOP>(X, pr0 + blockIdx.y, blockIdx.x, lds);
}

// Batched: one work-group per (job, tile) item of a bucket of equal
// transform length (blockIdx.x), prime pr0 + blockIdx.y.  One item per
// work-group keeps the register budget of the single-product kernel (a
// grid-stride loop over items kept its loop state live and cost ~60 VGPRs,
// one wave per SIMD); the per-prime pointers are resolved before the body,
// so nothing is indexed by the prime at run time.
template <bool INV, bool FROM_LIMBS, int NLEV, bool LAZY, bool TOP>
__global__ __launch_bounds__(256) void k_top_pass_batch(const BJob* __restrict__ jobs, u32 njobs, int operand,
                                                        TopBatchParams P, u32 pr0) {
    __shared__ u32 lds[4096 + 256];  // 32-bit transposes: half the LDS, more work-groups per CU
    const u64 tpj = P.L >> 12;
    const u32 pr = pr0 + blockIdx.y;
    const u64 tile = blockIdx.x % tpj;
    const BJob& J = jobs[blockIdx.x / tpj];
    if (operand && (J.sqr || J.share == 2)) return;  // uniform per work-group
    TopArgs A;
    u64* const data = (operand ? J.fb : J.fa) + (u64)pr * P.L;
    const u64* const tw = sel3(P.tw, pr);
    const u64 alpha = sel3(P.alpha_m, pr);
    for (int i = 0; i < 3; i++) {
        A.data[i] = data;
        A.tw[i] = tw;
        A.alpha_m[i] = alpha;
    }
    A.limbs = operand ? J.b : J.a;
    A.an = operand ? J.bn : J.an;
    A.g_lo = P.g_lo;
    A.nlev = P.nlev;
    top_pass_body<INV, FROM_LIMBS, NLEV, LAZY, TOP>(A, pr, tile, lds);
}

// Host dispatch on the pass depth (1..8 levels), known per launch.
template <bool INV, bool FL, bool LAZY, bool TOP>
static void top_single(const TopArgs& A, const TopOperand& B, u32 nlev, dim3 grid, u32 pr0, hipStream_t st) {
    switch (nlev) {
#define NNH_TP(n) \
    case n: hipLaunchKernelGGL((k_top_pass<INV, FL, n, LAZY, TOP>), grid, dim3(256), 0, st, A, B, pr0); break;
        NNH_TP(1) NNH_TP(2) NNH_TP(3) NNH_TP(4) NNH_TP(5) NNH_TP(6) NNH_TP(7)
        default: hipLaunchKernelGGL((k_top_pass<INV, FL, 8, LAZY, TOP>), grid, dim3(256), 0, st, A, B, pr0); break;
#undef NNH_TP
    }
}
template <bool INV, bool FL, bool LAZY, bool TOP>
static void top_batch(const BJob* jobs, u32 njobs, int operand, const TopBatchParams& P, dim3 grid, u32 pr0,
                      hipStream_t st) {
    switch (P.nlev) {
#define NNH_TP(n)                                                                                          \
    case n:                                                                                                \
        hipLaunchKernelGGL((k_top_pass_batch<INV, FL, n, LAZY, TOP>), grid, dim3(256), 0, st, jobs, njobs, \
                           operand, P, pr0);                                                               \
        break;
        NNH_TP(1) NNH_TP(2) NNH_TP(3) NNH_TP(4) NNH_TP(5) NNH_TP(6) NNH_TP(7)
        default: hipLaunchKernelGGL((k_top_pass_batch<INV, FL, 8, LAZY, TOP>), grid, dim3(256), 0, st, jobs,
                                    njobs, operand, P, pr0);
            break;
#undef NNH_TP
    }
}

template <bool LAZY>
static void top_pass_primes(const TopPassDesc& D, hipStream_t st) {
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
    TopOperand B{};
    if (D.two) {
        for (int i = 0; i < 3; i++) B.data[i] = D.data2[i];
        B.limbs = D.limbs2;
        B.an = D.an2;
    }
    const u32 pr0 = LAZY ? 0 : 2, npr = LAZY ? 2 : 1;
    dim3 grid((unsigned)(D.L >> 12), npr, D.two ? 2 : 1);
    const bool top = (1ull << (D.g_lo + D.nlev)) == D.L;  // the pass holds the top level
    if (D.inverse && top) top_single<true, false, LAZY, true>(A, B, D.nlev, grid, pr0, st);
    else if (D.inverse) top_single<true, false, LAZY, false>(A, B, D.nlev, grid, pr0, st);
    else if (D.limbs) top_single<false, true, LAZY, true>(A, B, D.nlev, grid, pr0, st);
    else top_single<false, false, LAZY, false>(A, B, D.nlev, grid, pr0, st);
}

template <bool LAZY>
static void top_pass_batch_primes(const BJob* jobs, u32 njobs, int operand, const TopBatchParams& P, bool inverse,
                                  bool from_limbs, hipStream_t st) {
    const u32 pr0 = LAZY ? 0 : 2, npr = LAZY ? 2 : 1;
    dim3 grid((unsigned)((u64)njobs * (P.L >> 12)), npr);
    const bool top = (1ull << (P.g_lo + P.nlev)) == P.L;  // the pass holds the top level
    if (inverse && top) top_batch<true, false, LAZY, true>(jobs, njobs, operand, P, grid, pr0, st);
    else if (inverse) top_batch<true, false, LAZY, false>(jobs, njobs, operand, P, grid, pr0, st);
    else if (from_limbs) top_batch<false, true, LAZY, true>(jobs, njobs, operand, P, grid, pr0, st);
    else top_batch<false, false, LAZY, false>(jobs, njobs, operand, P, grid, pr0, st);
}

}  // namespace nnh

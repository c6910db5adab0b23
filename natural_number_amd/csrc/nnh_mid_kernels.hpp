// nnh_mid_kernels.hpp — the fused chunk kernel (remaining forward levels of
// both operands, base multiply, inverse levels h < 4096), single and
// batched.  LAZY kernels serve p0 and p1 (lazy butterflies), the others p2;
// each translation unit instantiates one LAZY value.
#pragma once
#include "nnh_tile.hpp"

namespace nnh {

template <int CB, bool FROM_LIMBS, bool SQR, bool FWD_ONLY, bool LAZY>
__global__ __launch_bounds__(1 << (CB - 4)) void k_mid(MidArgs A, u32 pr0) {
    __shared__ u64 lds[(1 << CB) + (1 << (CB - 4))];
    const u32 pr = pr0 + blockIdx.y;
    mid_body<CB, FROM_LIMBS, SQR, FWD_ONLY, false, LAZY>(A, pr, blockIdx.x, lds);
}

// Batched: one work-group per (job, chunk) item (blockIdx.x), prime
// pr0 + blockIdx.y, as in the batched top pass (one item per work-group
// keeps the single-product register budget; per-prime pointers are resolved
// before the body).  FWD_ONLY: the jobs are the forward-only entries of
// shared B transforms (fa = the owner's fb, a = its B limbs); the finished
// transform is stored.  Otherwise a job with share != 0 reads B's finished
// transform from fb.
template <int CB, bool FROM_LIMBS, bool FWD_ONLY, bool LAZY>
__global__ __launch_bounds__(1 << (CB - 4)) void k_mid_batch(const BJob* __restrict__ jobs, u32 njobs,
                                                            MidBatchParams P, u32 pr0) {
    __shared__ u64 lds[(1 << CB) + (1 << (CB - 4))];
    const u64 cpj = P.L >> CB;
    const u32 pr = pr0 + blockIdx.y;
    const u64 chunk = blockIdx.x % cpj;
    const BJob& J = jobs[blockIdx.x / cpj];
    MidArgs A;
    u64* const fa = J.fa + (u64)pr * P.L;
    const u64* const fb = J.fb + (u64)pr * P.L;
    const u64* const twf = sel3(P.twf, pr);
    const u64* const twi = sel3(P.twi, pr);
    const u64 alpha = sel3(P.alpha_m, pr);
    for (int i = 0; i < 3; i++) {
        A.fa[i] = fa;
        A.fb[i] = fb;
        A.twf[i] = twf;
        A.twi[i] = twi;
        A.alpha_m[i] = alpha;
    }
    A.la = J.a;
    A.lb = FROM_LIMBS && !J.share ? J.b : nullptr;
    A.an = J.an;
    A.bn = J.bn;
    A.a_done = 0;
    A.b_done = J.share != 0;
    bool sq = J.sqr != 0;
    mid_body<CB, FROM_LIMBS, false, FWD_ONLY, true, LAZY>(A, pr, chunk, lds, sq, 0);
}

template <int CB, bool LAZY>
static void mid_dispatch(const MidArgs& A, u64 L, bool from_limbs, bool sqr, bool fwd_only, hipStream_t st) {
    const u32 pr0 = LAZY ? 0 : 2, npr = LAZY ? 2 : 1;
    dim3 grid((unsigned)(L >> CB), npr), block(1 << (CB - 4));
    if (fwd_only) {
        if (from_limbs) hipLaunchKernelGGL((k_mid<CB, true, false, true, LAZY>), grid, block, 0, st, A, pr0);
        else hipLaunchKernelGGL((k_mid<CB, false, false, true, LAZY>), grid, block, 0, st, A, pr0);
    } else if (from_limbs) {
        if (sqr) hipLaunchKernelGGL((k_mid<CB, true, true, false, LAZY>), grid, block, 0, st, A, pr0);
        else hipLaunchKernelGGL((k_mid<CB, true, false, false, LAZY>), grid, block, 0, st, A, pr0);
    } else {
        if (sqr) hipLaunchKernelGGL((k_mid<CB, false, true, false, LAZY>), grid, block, 0, st, A, pr0);
        else hipLaunchKernelGGL((k_mid<CB, false, false, false, LAZY>), grid, block, 0, st, A, pr0);
    }
}

template <bool LAZY>
static void mid_primes(const MidDesc& D, hipStream_t st) {
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
        case 8: mid_dispatch<8, LAZY>(A, D.L, fl, D.sqr, D.fwd_only, st); break;
        case 9: mid_dispatch<9, LAZY>(A, D.L, fl, D.sqr, D.fwd_only, st); break;
        case 10: mid_dispatch<10, LAZY>(A, D.L, fl, D.sqr, D.fwd_only, st); break;
        case 11: mid_dispatch<11, LAZY>(A, D.L, fl, D.sqr, D.fwd_only, st); break;
        default: mid_dispatch<12, LAZY>(A, D.L, fl, D.sqr, D.fwd_only, st); break;
    }
}

template <int CB, bool LAZY>
static void mid_batch_dispatch(const BJob* jobs, u32 njobs, const MidBatchParams& P, bool from_limbs, bool fwd_only,
                               hipStream_t st) {
    const u32 pr0 = LAZY ? 0 : 2, npr = LAZY ? 2 : 1;
    dim3 grid((unsigned)((u64)njobs * (P.L >> CB)), npr), block(1 << (CB - 4));
#define NNH_MB(fl, fo) hipLaunchKernelGGL((k_mid_batch<CB, fl, fo, LAZY>), grid, block, 0, st, jobs, njobs, P, pr0)
    if (fwd_only) {
        if (from_limbs) NNH_MB(true, true);
        else NNH_MB(false, true);
    } else {
        if (from_limbs) NNH_MB(true, false);
        else NNH_MB(false, false);
    }
#undef NNH_MB
}

template <bool LAZY>
static void mid_batch_primes(const BJob* jobs, u32 njobs, const MidBatchParams& P, int cb, bool from_limbs,
                             bool fwd_only, hipStream_t st) {
    switch (cb) {
        case 8: mid_batch_dispatch<8, LAZY>(jobs, njobs, P, from_limbs, fwd_only, st); break;
        case 9: mid_batch_dispatch<9, LAZY>(jobs, njobs, P, from_limbs, fwd_only, st); break;
        case 10: mid_batch_dispatch<10, LAZY>(jobs, njobs, P, from_limbs, fwd_only, st); break;
        case 11: mid_batch_dispatch<11, LAZY>(jobs, njobs, P, from_limbs, fwd_only, st); break;
        default: mid_batch_dispatch<12, LAZY>(jobs, njobs, P, from_limbs, fwd_only, st); break;
    }
}

}  // namespace nnh

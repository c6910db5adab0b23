// nnh_top.hip — strided "top" passes of the transform (levels h >= 4096):
// forward (fused with the limb load on the first pass) and inverse, for one
// product or a bucket of a batch.
#include "nnh_tile.hpp"

namespace nnh {

template <bool INV, bool FROM_LIMBS>
__global__ __launch_bounds__(256) void k_top_pass(TopArgs A) {
    __shared__ u64 lds[4096 + 256];
    const u32 pr = blockIdx.y;
    top_pass_nlev<INV, FROM_LIMBS>(A, pr, blockIdx.x, lds);
}


// Batched: persistent grid-stride loop over (job, tile, prime) items of a
// bucket of equal transform length.
template <bool INV, bool FROM_LIMBS>
__global__ __launch_bounds__(256) void k_top_pass_batch(const BJob* __restrict__ jobs, u32 njobs, int operand,
                                                        TopBatchParams P) {
    __shared__ u64 lds[4096 + 256];
    const u64 tpj = P.L >> 12;
    const u64 items = (u64)njobs * tpj * 3;
    for (u64 it = blockIdx.x; it < items; it += gridDim.x) {
        const u32 pr = (u32)(it % 3);
        const u64 r = it / 3;
        const u64 tile = r % tpj;
        const BJob& J = jobs[r / tpj];
        if (operand && J.sqr) continue;  // uniform per work-group
        TopArgs A;
        u64* base = operand ? J.fb : J.fa;
        for (int i = 0; i < 3; i++) {
            A.data[i] = base + (u64)i * P.L;
            A.tw[i] = P.tw[i];
            A.alpha_m[i] = P.alpha_m[i];
        }
        A.limbs = operand ? J.b : J.a;
        A.an = operand ? J.bn : J.an;
        A.g_lo = P.g_lo;
        A.nlev = P.nlev;
        top_pass_nlev<INV, FROM_LIMBS>(A, pr, tile, lds);
        __syncthreads();
    }
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

void launch_top_pass_batch(const BJob* jobs, u32 njobs, int operand, const TopBatchParams& P, bool inverse,
                           bool from_limbs, hipStream_t st) {
    u64 items = (u64)njobs * (P.L >> 12) * 3;
    dim3 grid(persist_grid(items, 4096)), block(256);
    if (inverse) hipLaunchKernelGGL((k_top_pass_batch<true, false>), grid, block, 0, st, jobs, njobs, operand, P);
    else if (from_limbs) hipLaunchKernelGGL((k_top_pass_batch<false, true>), grid, block, 0, st, jobs, njobs, operand, P);
    else hipLaunchKernelGGL((k_top_pass_batch<false, false>), grid, block, 0, st, jobs, njobs, operand, P);
}

}  // namespace nnh

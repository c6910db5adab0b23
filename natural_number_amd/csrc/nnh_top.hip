// nnh_top.hip — top passes for p2 (strict butterflies) and the dispatch of
// one top pass over all three primes.
#include "nnh_top_kernels.hpp"

namespace nnh {

void launch_top_pass_part(const TopPassDesc& d, bool lazy, hipStream_t st);
void launch_top_pass_batch_part(const BJob* jobs, u32 njobs, int operand, const TopBatchParams& P, bool inverse,
                                bool from_limbs, bool lazy, hipStream_t st);

void launch_top_pass(const TopPassDesc& D, hipStream_t st, hipStream_t st2) {
    launch_top_pass_part(D, true, st);
    top_pass_primes<false>(D, st2 ? st2 : st);
}

void launch_top_pass_batch(const BJob* jobs, u32 njobs, int operand, const TopBatchParams& P, bool inverse,
                           bool from_limbs, hipStream_t st, hipStream_t st2) {
    launch_top_pass_batch_part(jobs, njobs, operand, P, inverse, from_limbs, true, st);
    top_pass_batch_primes<false>(jobs, njobs, operand, P, inverse, from_limbs, st2 ? st2 : st);
}

}  // namespace nnh

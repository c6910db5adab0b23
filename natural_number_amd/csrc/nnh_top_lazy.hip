// nnh_top_lazy.hip — top passes for p0 and p1 (lazy butterflies).
#include "nnh_top_kernels.hpp"

namespace nnh {

void launch_top_pass_part(const TopPassDesc& d, bool lazy, hipStream_t st) {
    if (lazy) top_pass_primes<true>(d, st);
}

void launch_top_pass_batch_part(const BJob* jobs, u32 njobs, int operand, const TopBatchParams& P, bool inverse,
                                bool from_limbs, bool lazy, hipStream_t st) {
    if (lazy) top_pass_batch_primes<true>(jobs, njobs, operand, P, inverse, from_limbs, st);
}

}  // namespace nnh

// nnh_mid_lazy.hip — fused chunk kernels for p0 and p1 (lazy butterflies).
#include "nnh_mid_kernels.hpp"

namespace nnh {

void launch_mid_part(const MidDesc& d, bool lazy, hipStream_t st) {
    if (lazy) mid_primes<true>(d, st);
}

void launch_mid_batch_part(const BJob* jobs, u32 njobs, const MidBatchParams& P, int cb, bool from_limbs,
                           bool fwd_only, bool lazy, hipStream_t st) {
    if (lazy) mid_batch_primes<true>(jobs, njobs, P, cb, from_limbs, fwd_only, st);
}

}  // namespace nnh

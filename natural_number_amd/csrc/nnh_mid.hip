// nnh_mid.hip — fused chunk kernels for p2 (strict butterflies) and the
// dispatch of one chunk pass over all three primes.
#include "nnh_mid_kernels.hpp"

namespace nnh {

void launch_mid_part(const MidDesc& d, bool lazy, hipStream_t st);
void launch_mid_batch_part(const BJob* jobs, u32 njobs, const MidBatchParams& P, int cb, bool from_limbs,
                           bool fwd_only, bool lazy, hipStream_t st);

void launch_mid(const MidDesc& D, hipStream_t st, hipStream_t st2) {
    launch_mid_part(D, true, st);
    mid_primes<false>(D, st2 ? st2 : st);
}

void launch_mid_batch(const BJob* jobs, u32 njobs, const MidBatchParams& P, int cb, bool from_limbs, bool fwd_only,
                      hipStream_t st, hipStream_t st2) {
    launch_mid_batch_part(jobs, njobs, P, cb, from_limbs, fwd_only, true, st);
    mid_batch_primes<false>(jobs, njobs, P, cb, from_limbs, fwd_only, st2 ? st2 : st);
}

}  // namespace nnh

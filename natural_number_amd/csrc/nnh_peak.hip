// nnh_peak.hip — issue ceiling of the transform butterflies: the same
// ct_bfly_lazy / ct_bfly functions the transform kernels use, on
// register-resident data with no memory or LDS traffic.  The rate it
// measures bounds every transform kernel from above.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/nnh.h"
#include "nnh_arith.hpp"
#include "nnh_host.hpp"

namespace nnh {

constexpr int PEAK_THREADS = 256;
constexpr int PEAK_BLOCKS = 256 * 8;  // 8 work-groups per CU on 256 CUs

// Each thread holds 16 residues and 8 twiddles and runs 16 butterflies per
// iteration (two levels of 8 independent pairs).
template <bool LAZY>
__global__ __launch_bounds__(PEAK_THREADS) void k_bfly_peak(u64* data, const u64* tw, u32 pr, u32 iters) {
    const Prime P = make_prime(PRIME_K[0] * (pr == 0) + PRIME_K[1] * (pr == 1) + PRIME_K[2] * (pr == 2));
    u64 r[16], w[8];
    const u64 base = ((u64)blockIdx.x * PEAK_THREADS + threadIdx.x) * 16;
    for (int i = 0; i < 16; i++) r[i] = data[base + i];
    for (int i = 0; i < 8; i++) w[i] = tw[(threadIdx.x + 37 * i) & 1023];
    for (u32 it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < 8; i++) {
            if constexpr (LAZY) ct_bfly_lazy(P, r[i], r[i + 8], w[i]);
            else ct_bfly(P, r[i], r[i + 8], w[i]);
        }
#pragma unroll
        for (int i = 0; i < 8; i++) {
            if constexpr (LAZY) ct_bfly_lazy(P, r[2 * i], r[2 * i + 1], w[(i + 3) & 7]);
            else ct_bfly(P, r[2 * i], r[2 * i + 1], w[(i + 3) & 7]);
        }
    }
    for (int i = 0; i < 16; i++) data[base + i] = r[i];
}

__global__ void k_peak_fill(u64* d, u64 n, u64 p) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
        u64 z = (i + 1) * 0x9e3779b97f4a7c15ull;
        z = (z ^ (z >> 31)) * 0xbf58476d1ce4e5b9ull;
        d[i] = (z ^ (z >> 29)) % p;
    }
}

int butterfly_peak(int pr, u32 iters, double* rate) {
    const u64 n = (u64)PEAK_BLOCKS * PEAK_THREADS * 16;
    const u64 p = prime_of(pr).p;
    hipStream_t st = host_stream();
    u64 *d = nullptr, *tw = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = NNH_OK;
    auto chk = [&](hipError_t e) {
        if (e != hipSuccess && rc == NNH_OK)
            rc = set_error(NNH_ERR_HIP, std::string("butterfly peak: ") + hipGetErrorString(e));
        return e == hipSuccess;
    };
    if (chk(hipMalloc(&d, n * 8)) && chk(hipMalloc(&tw, 1024 * 8)) && chk(hipEventCreate(&e0)) &&
        chk(hipEventCreate(&e1))) {
        hipLaunchKernelGGL(k_peak_fill, dim3(1024), dim3(256), 0, st, d, n, p);
        hipLaunchKernelGGL(k_peak_fill, dim3(4), dim3(256), 0, st, tw, (u64)1024, p);
        chk(hipGetLastError());
        float best = 1e30f;
        for (int rep = 0; rep < 4 && rc == NNH_OK; rep++) {
            chk(hipEventRecord(e0, st));
            if (lazy_prime(pr))
                hipLaunchKernelGGL(k_bfly_peak<true>, dim3(PEAK_BLOCKS), dim3(PEAK_THREADS), 0, st, d, tw, (u32)pr, iters);
            else
                hipLaunchKernelGGL(k_bfly_peak<false>, dim3(PEAK_BLOCKS), dim3(PEAK_THREADS), 0, st, d, tw, (u32)pr, iters);
            chk(hipGetLastError());
            chk(hipEventRecord(e1, st));
            chk(hipEventSynchronize(e1));
            float ms = 0;
            chk(hipEventElapsedTime(&ms, e0, e1));
            if (ms > 0 && ms < best) best = ms;
        }
        // n / 16 threads, 16 butterflies per thread and iteration
        if (rc == NNH_OK) *rate = (double)n * iters / (best * 1e-3);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (d) (void)hipFree(d);
    if (tw) (void)hipFree(tw);
    return rc;
}

}  // namespace nnh

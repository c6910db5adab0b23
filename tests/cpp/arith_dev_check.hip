// Device check of the modular arithmetic in nnh_arith.hpp as the kernels
// compile it (the asm Montgomery products, the fused-addend lazy forward
// butterfly, the twiddle-1 butterflies, the asm base multiply, Garner's
// reconstruction): each primitive runs on the GPU over random and boundary
// inputs, and the host checks every output's range and its value mod p
// against exact 128-bit arithmetic.  Built by the package build next to this file; run by
// tests/test_gpu_arith.py.  Exit status 0 = every check passes.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../natural_number_amd/csrc/nnh_arith.hpp"

using namespace nnh;
typedef unsigned __int128 u128;

constexpr int NIN = 9, NOUT = 20;

// in: x0..x3, y0..y3, w (Montgomery twiddle) per case
__global__ void k_check(const u64* __restrict__ in, u64* __restrict__ out, u32 pr, u32 n) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Prime P = make_prime(PRIME_K[0] * (pr == 0) + PRIME_K[1] * (pr == 1) + PRIME_K[2] * (pr == 2));
    const u64* c = in + (u64)NIN * i;
    u64* o = out + (u64)NOUT * i;
    const u64 w = c[8];
    o[0] = mont(P, c[0], w);
    if (lazy_prime(pr)) {
        u64 x = c[0], y = c[1];  // [0, 4p)
        ct_bfly_lazy(P, x, y, w);
        o[1] = x; o[2] = y;
        x = csub2(P, c[2]); y = csub2(P, c[3]);  // [0, 2p)
        gs_bfly_lazy(P, x, y, w);
        o[3] = x; o[4] = y;
        x = c[0]; y = c[1];
        ct_bfly_lazy_one(P, x, y);
        o[5] = x; o[6] = y;
        x = csub2(P, c[2]); y = csub2(P, c[3]);
        gs_bfly_lazy_one(P, x, y);
        o[7] = x; o[8] = y;
        u64 x0 = c[0], x1 = c[1], x2 = c[2], x3 = c[3];
        base_mul4<true>(P, x0, x1, x2, x3, c[4], c[5], c[6], c[7], w);
        o[9] = x0; o[10] = x1; o[11] = x2; o[12] = x3;
        o[13] = csub2(P, c[0]);
    } else {
        u64 x = c[0] % (2 * P.p), y = c[1] % (2 * P.p);  // [0, 2p)
        ct_bfly(P, x, y, w);
        o[1] = x; o[2] = y;
        x = c[2] % P.p; y = c[3] % P.p;
        gs_bfly(P, x, y, w);
        o[3] = x; o[4] = y;
        x = c[0] % (2 * P.p); y = c[1] % (2 * P.p);
        ct_bfly_one(P, x, y);
        o[5] = x; o[6] = y;
        x = c[2] % P.p; y = c[3] % P.p;
        gs_bfly_one(P, x, y);
        o[7] = x; o[8] = y;
        u64 x0 = c[0] % (2 * P.p), x1 = c[1] % (2 * P.p), x2 = c[2] % (2 * P.p), x3 = c[3] % (2 * P.p);
        base_mul4<false>(P, x0, x1, x2, x3, c[4] % (2 * P.p), c[5] % (2 * P.p), c[6] % (2 * P.p),
                         c[7] % (2 * P.p), w);
        o[9] = x0; o[10] = x1; o[11] = x2; o[12] = x3;
        o[13] = csub(P, c[0] % (2 * P.p));
    }
    // mont2 on canonical operands, the last one in (0, p] as in the base multiply
    const u64 a = c[0] % P.p, b = c[4] % P.p, cc = c[1] % P.p, d = c[5] % P.p + 1;
    o[14] = mont2(P, a, b, cc, d);
    o[15] = o[14];
    o[16] = o[17] = o[18] = o[19] = 0;
#if defined(__HIP_DEVICE_COMPILE__)
    if (lazy_prime(pr)) {
        o[15] = mont2_dev(P, a, b, cc, d);
        // the lazy forward butterfly's product: y in [0, 4p), addend a + p with a in [0, 2p)
        const u64 ap = csub2(P, c[0]) + P.p;
        o[16] = mont_dev_add<true>(P, c[1], w, ap);
        o[17] = c[1];
        o[18] = ap;
    }
#endif
}

static u64 powmod(u64 b, u64 e, u64 p) {
    u64 r = 1;
    for (b %= p; e; e >>= 1, b = (u64)((u128)b * b % p))
        if (e & 1) r = (u64)((u128)r * b % p);
    return r;
}

// Garner reconstruction as the CRT kernels compile it: canonical residues
// (r0, r1, r2) per case in, X as three words out.
__global__ void k_garner(const u64* __restrict__ in, u64* __restrict__ out, u64 c02m, u64 c1m, u64 km, u32 n) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    garner(in[3 * (u64)i], in[3 * (u64)i + 1], in[3 * (u64)i + 2], c02m, c1m, km, out[3 * (u64)i],
           out[3 * (u64)i + 1], out[3 * (u64)i + 2]);
}

static u64 to_mont(u64 x, u64 p) { return (u64)(((u128)x << 64) % p); }

// X < p0 p1 p2 with X mod p_i = r_i for random and boundary residues.
static int check_garner() {
    const u64 m0 = prime_of(0).p, m1 = prime_of(1).p, m2 = prime_of(2).p;
    const u64 c02m = to_mont(powmod(m0, m2 - 2, m2), m2);
    const u64 c1 = powmod((u64)((u128)(m0 % m1) * (m2 % m1) % m1), m1 - 2, m1);
    const u64 c1m = to_mont(c1, m1);
    const u64 km = to_mont((m1 - (u64)((u128)(m0 % m1) * c1 % m1)) % m1, m1);
    const u32 n = 1u << 18;
    std::vector<u64> in(3 * (size_t)n), out(3 * (size_t)n);
    std::mt19937_64 rng(4321);
    const u64 ms[3] = {m0, m1, m2};
    for (u32 i = 0; i < n; i++)
        for (int j = 0; j < 3; j++) {
            const u64 p = ms[j];
            const u64 edge[] = {0, 1, 2, p - 1, p - 2, p / 2, (p - 1) & ~0xffffffffull, 0xffffffffull};
            in[3 * (size_t)i + j] = i < n / 4 ? edge[rng() % 8] : rng() % p;
        }
    u64 *din = nullptr, *dout = nullptr;
    if (hipMalloc(&din, in.size() * 8) != hipSuccess || hipMalloc(&dout, out.size() * 8) != hipSuccess ||
        hipMemcpy(din, in.data(), in.size() * 8, hipMemcpyHostToDevice) != hipSuccess) {
        std::printf("garner: device set-up failed\n");
        return 1;
    }
    hipLaunchKernelGGL(k_garner, dim3(n / 256), dim3(256), 0, 0, din, dout, c02m, c1m, km, n);
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(out.data(), dout, out.size() * 8, hipMemcpyDeviceToHost) != hipSuccess) {
        std::printf("garner: kernel failed\n");
        return 1;
    }
    (void)hipFree(din);
    (void)hipFree(dout);
    // M = p0 p1 p2 as three words, for the range check
    const u128 m01 = (u128)m0 * m1;
    const u128 lo = (u128)(u64)m01 * m2, hi = (u128)(u64)(m01 >> 64) * m2 + (lo >> 64);
    const u64 M[3] = {(u64)lo, (u64)hi, (u64)(hi >> 64)};
    long bad = 0;
    for (u32 i = 0; i < n; i++) {
        const u64* x = &out[3 * (size_t)i];
        bool ok = true;
        for (int j = 0; j < 3; j++) {
            const u64 p = ms[j];
            const u128 t = ((u128)(x[2] % p) << 64 | x[1]) % p;
            ok = ok && (u64)(((t << 64) | x[0]) % p) == in[3 * (size_t)i + j];
        }
        const bool below = x[2] != M[2] ? x[2] < M[2] : (x[1] != M[1] ? x[1] < M[1] : x[0] < M[0]);
        if (!ok || !below) bad++;
    }
    std::printf("garner: %u cases; %ld failing\n", n, bad);
    return bad ? 1 : 0;
}

int main() {
    int bad_total = 0;
    for (u32 pr = 0; pr < 3; pr++) {
        const Prime P = prime_of(pr);
        const u64 p = P.p;
        const bool lazy = lazy_prime(pr);
        const u64 rinv = powmod((u64)(((u128)1 << 64) % p), p - 2, p);  // 2^-64 mod p
        const u64 top = lazy ? 4 * p : 2 * p;                            // forward input bound
        const u32 n = 1u << 18;
        std::vector<u64> in((size_t)NIN * n), out((size_t)NOUT * n);
        std::mt19937_64 rng(1234 + pr);
        const u64 edge[] = {0, 1, p - 1, p, p + 1, 2 * p - 1, 2 * p - 2, top - 1, top - 2, top / 2, p - 2,
                            (p - 1) & ~0xffffffffull};
        const int ne = sizeof(edge) / sizeof(edge[0]);
        for (u32 i = 0; i < n; i++) {
            for (int j = 0; j < 8; j++) {
                u64 v = i < n / 4 ? edge[rng() % ne] : (i < n / 2 ? top - 1 - (rng() & 0xfff) : rng() % top);
                in[(size_t)NIN * i + j] = v % top;
            }
            u64 wv = i % 5 == 0 ? p - 1 : (i % 7 == 0 ? 0 : rng() % p);
            in[(size_t)NIN * i + 8] = (u64)(((u128)wv << 64) % p);  // Montgomery form
        }
        u64 *din = nullptr, *dout = nullptr;
        if (hipMalloc(&din, in.size() * 8) != hipSuccess || hipMalloc(&dout, out.size() * 8) != hipSuccess) {
            std::printf("hipMalloc failed\n");
            return 2;
        }
        if (hipMemcpy(din, in.data(), in.size() * 8, hipMemcpyHostToDevice) != hipSuccess) {
            std::printf("copy failed\n");
            return 2;
        }
        hipLaunchKernelGGL(k_check, dim3(n / 256), dim3(256), 0, 0, din, dout, pr, n);
        if (hipDeviceSynchronize() != hipSuccess) {
            std::printf("kernel failed\n");
            return 2;
        }
        if (hipMemcpy(out.data(), dout, out.size() * 8, hipMemcpyDeviceToHost) != hipSuccess) {
            std::printf("copy failed\n");
            return 2;
        }
        (void)hipFree(din);
        (void)hipFree(dout);
        long bad[NOUT] = {0};
        auto m = [&](u128 t) { return (u64)(t % p); };
        for (u32 i = 0; i < n; i++) {
            const u64* c = &in[(size_t)NIN * i];
            const u64* o = &out[(size_t)NOUT * i];
            const u64 wm = c[8], wp = m((u128)wm * rinv);  // plain twiddle
            // Montgomery product of a limb-sized value
            if (o[0] >= 2 * p || o[0] % p != m((u128)m((u128)c[0] * wm) * rinv)) bad[0]++;
            const u64 xa = lazy ? c[0] : c[0] % (2 * p), ya = lazy ? c[1] : c[1] % (2 * p);
            const u64 t = m((u128)(ya % p) * wp);
            const u64 fb = lazy ? 4 * p : 2 * p;  // forward output bound
            if (o[1] >= fb || o[2] >= fb || o[1] % p != (xa % p + t) % p || o[2] % p != (xa % p + p - t) % p) bad[1]++;
            const u64 gx = (lazy ? (c[2] >= 2 * p ? c[2] - 2 * p : c[2]) : c[2] % p) % p;
            const u64 gy = (lazy ? (c[3] >= 2 * p ? c[3] - 2 * p : c[3]) : c[3] % p) % p;
            const u64 ib = lazy ? 2 * p : p;  // inverse output bound
            if (o[3] >= ib || o[4] >= ib || o[3] % p != (gx + gy) % p || o[4] % p != m((u128)((gx + p - gy) % p) * wp))
                bad[3]++;
            if (o[5] >= fb || o[6] >= fb || o[5] % p != (xa % p + ya % p) % p || o[6] % p != (xa % p + p - ya % p) % p)
                bad[5]++;
            if (o[7] >= ib || o[8] >= ib || o[7] % p != (gx + gy) % p || o[8] % p != (gx + p - gy) % p) bad[7]++;
            // base multiply: (x0 + x1 X)(y0 + y1 X) mod X^2 - w, (x2 + x3 X)(y2 + y3 X) mod X^2 + w, times 2^-64
            u64 xs[4], ys[4];
            for (int j = 0; j < 4; j++) {
                xs[j] = (lazy ? c[j] : c[j] % (2 * p)) % p;
                ys[j] = (lazy ? c[4 + j] : c[4 + j] % (2 * p)) % p;
            }
            const u64 z0 = m((u128)xs[0] * ys[0] + (u128)m((u128)xs[1] * ys[1]) * wp);
            const u64 z1 = m((u128)xs[0] * ys[1] + (u128)xs[1] * ys[0]);
            const u64 z2 = m((u128)xs[2] * ys[2] + (u128)m((u128)xs[3] * ys[3]) * (p - wp));
            const u64 z3 = m((u128)xs[2] * ys[3] + (u128)xs[3] * ys[2]);
            const u64 zs[4] = {z0, z1, z2, z3};
            const u64 zb = lazy ? 2 * p : p;
            for (int j = 0; j < 4; j++)
                if (o[9 + j] >= zb || o[9 + j] % p != m((u128)zs[j] * rinv)) bad[9]++;
            const u64 cs = lazy ? c[0] : c[0] % (2 * p);
            if (o[13] >= (lazy ? 2 * p : p) || o[13] % p != cs % p) bad[13]++;
            const u64 a = c[0] % p, b = c[4] % p, cc = c[1] % p, d = c[5] % p + 1;
            const u64 e2 = m((u128)m((u128)a * b + (u128)cc * d) * rinv);
            if (o[14] >= 2 * p || o[14] % p != e2 || o[15] != o[14]) bad[14]++;
            // fused-addend product: c - p + (y w 2^-64 mod p, in (0, 2p))
            if (lazy) {
                const u64 add = o[18], y = o[17];
                const u64 r = o[16] - (add - p);
                if (r >= 2 * p || r % p != m((u128)m((u128)y * wm) * rinv)) bad[16]++;
            }
        }
        int bp = 0;
        for (int k = 0; k < NOUT; k++) bp += bad[k] != 0;
        std::printf("p%u (%s): %u cases; failing checks:", pr, lazy ? "lazy" : "strict", n);
        for (int k = 0; k < NOUT; k++)
            if (bad[k]) std::printf(" [%d]=%ld", k, bad[k]);
        std::printf(bp ? "\n" : " none\n");
        bad_total += bp;
    }
    bad_total += check_garner();
    return bad_total ? 1 : 0;
}

// nnh_mlp.hip — the learned cost model: an FP32 MLP  x -> relu(W1 x + b1)
// -> relu(W2 h1 + b2) -> W3 h2 + b3  whose outputs predict log2(time) of
// each candidate multiplication algorithm.  Hand-written forward, backward
// (per-sample deltas, then a deterministic per-weight reduction over the
// batch), Adam update and arg-min selection kernels.  No counterpart in the
// reference, whose algorithm choice is a fixed threshold.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/nnh.h"

namespace mlpk {

constexpr int H = 64;          // hidden width
constexpr int MAX_IN = 8;
constexpr int MAX_OUT = 16;

struct Dims {
    int ni, no;
    // offsets into the flat parameter vector
    int w1, b1, w2, b2, w3, b3, n;
};

__host__ __device__ inline Dims make_dims(int ni, int no) {
    Dims d;
    d.ni = ni;
    d.no = no;
    d.w1 = 0;
    d.b1 = d.w1 + H * ni;
    d.w2 = d.b1 + H;
    d.b2 = d.w2 + H * H;
    d.w3 = d.b2 + H;
    d.b3 = d.w3 + no * H;
    d.n = d.b3 + no;
    return d;
}

// One thread per sample.  Stores h1, h2 (post-ReLU) for the backward pass.
__global__ __launch_bounds__(64) void k_forward(const float* __restrict__ P, Dims d, const float* __restrict__ x,
                                                int n, float* __restrict__ y, float* __restrict__ h1s,
                                                float* __restrict__ h2s) {
    __shared__ float sp[H * MAX_IN + H + H * H + H + MAX_OUT * H + MAX_OUT];
    // per-thread activations live in LDS rows (indexed by loop counters that
    // a register array could only serve through scratch memory)
    __shared__ float sh1[64][H + 1], sh2[64][H + 1];
    for (int i = threadIdx.x; i < d.n; i += blockDim.x) sp[i] = P[i];
    __syncthreads();
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float xi[MAX_IN];
#pragma unroll
    for (int k = 0; k < MAX_IN; k++) xi[k] = k < d.ni ? x[(size_t)s * d.ni + k] : 0.f;
    float* h1 = sh1[threadIdx.x];
    float* h2 = sh2[threadIdx.x];
    for (int j = 0; j < H; j++) {
        float a = sp[d.b1 + j];
#pragma unroll
        for (int k = 0; k < MAX_IN; k++)
            if (k < d.ni) a = fmaf(sp[d.w1 + j * d.ni + k], xi[k], a);
        h1[j] = a > 0.f ? a : 0.f;
    }
    for (int j = 0; j < H; j++) {
        float a = sp[d.b2 + j];
        for (int k = 0; k < H; k++) a = fmaf(sp[d.w2 + j * H + k], h1[k], a);
        h2[j] = a > 0.f ? a : 0.f;
    }
    for (int o = 0; o < d.no; o++) {
        float a = sp[d.b3 + o];
        for (int k = 0; k < H; k++) a = fmaf(sp[d.w3 + o * H + k], h2[k], a);
        y[(size_t)s * d.no + o] = a;
    }
    if (h1s) {
        for (int j = 0; j < H; j++) h1s[(size_t)s * H + j] = h1[j];
        for (int j = 0; j < H; j++) h2s[(size_t)s * H + j] = h2[j];
    }
}

// Per-sample deltas of the masked mean-squared error
//   L = sum_s sum_o m_so (y_so - t_so)^2 / sum m.
// dy = 2 m (y - t) / M ; dh2 = W3^T dy * [h2 > 0] ; dh1 = W2^T dh2 * [h1 > 0].
__global__ __launch_bounds__(64) void k_deltas(const float* __restrict__ P, Dims d, int n,
                                               const float* __restrict__ y, const float* __restrict__ t,
                                               const float* __restrict__ m, float inv_m,
                                               const float* __restrict__ h1s, const float* __restrict__ h2s,
                                               float* __restrict__ dys, float* __restrict__ dh1s,
                                               float* __restrict__ dh2s, float* __restrict__ sq) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float dy[MAX_OUT];
    float l = 0.f;
    for (int o = 0; o < d.no; o++) {
        float e = y[(size_t)s * d.no + o] - t[(size_t)s * d.no + o];
        float mm = m[(size_t)s * d.no + o];
        dy[o] = 2.f * mm * e * inv_m;
        l += mm * e * e;
        dys[(size_t)s * d.no + o] = dy[o];
    }
    sq[s] = l;
    float dh2[H];
    for (int j = 0; j < H; j++) {
        float a = 0.f;
        for (int o = 0; o < d.no; o++) a = fmaf(P[d.w3 + o * H + j], dy[o], a);
        dh2[j] = h2s[(size_t)s * H + j] > 0.f ? a : 0.f;
        dh2s[(size_t)s * H + j] = dh2[j];
    }
    for (int j = 0; j < H; j++) {
        float a = 0.f;
        for (int k = 0; k < H; k++) a = fmaf(P[d.w2 + k * H + j], dh2[k], a);
        dh1s[(size_t)s * H + j] = h1s[(size_t)s * H + j] > 0.f ? a : 0.f;
    }
}

// One thread per parameter: sum over samples in a fixed order.
__global__ __launch_bounds__(256) void k_weight_grads(Dims d, int n, const float* __restrict__ x,
                                                      const float* __restrict__ h1s, const float* __restrict__ h2s,
                                                      const float* __restrict__ dys, const float* __restrict__ dh1s,
                                                      const float* __restrict__ dh2s, float* __restrict__ g) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.n) return;
    float acc = 0.f;
    if (p < d.b1) {                 // W1[j][k]
        int j = p / d.ni, k = p % d.ni;
        for (int s = 0; s < n; s++) acc = fmaf(dh1s[(size_t)s * H + j], x[(size_t)s * d.ni + k], acc);
    } else if (p < d.w2) {          // b1[j]
        int j = p - d.b1;
        for (int s = 0; s < n; s++) acc += dh1s[(size_t)s * H + j];
    } else if (p < d.b2) {          // W2[j][k]
        int q = p - d.w2, j = q / H, k = q % H;
        for (int s = 0; s < n; s++) acc = fmaf(dh2s[(size_t)s * H + j], h1s[(size_t)s * H + k], acc);
    } else if (p < d.w3) {          // b2[j]
        int j = p - d.b2;
        for (int s = 0; s < n; s++) acc += dh2s[(size_t)s * H + j];
    } else if (p < d.b3) {          // W3[o][k]
        int q = p - d.w3, o = q / H, k = q % H;
        for (int s = 0; s < n; s++) acc = fmaf(dys[(size_t)s * d.no + o], h2s[(size_t)s * H + k], acc);
    } else {                        // b3[o]
        int o = p - d.b3;
        for (int s = 0; s < n; s++) acc += dys[(size_t)s * d.no + o];
    }
    g[p] = acc;
}

__global__ __launch_bounds__(256) void k_adam(float* P, const float* g, float* m1, float* m2, int n, float lr,
                                              float b1, float b2, float c1, float c2, float eps) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float gg = g[p];
    float a = b1 * m1[p] + (1.f - b1) * gg;
    float b = b2 * m2[p] + (1.f - b2) * gg * gg;
    m1[p] = a;
    m2[p] = b;
    P[p] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// Arg-min of the predicted cost over the feasible outputs (mask > 0).
__global__ __launch_bounds__(64) void k_choose(const float* __restrict__ y, const float* __restrict__ feas, int n,
                                               int no, int* __restrict__ choice) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int best = -1;
    float bv = 0.f;
    for (int o = 0; o < no; o++) {
        if (feas[(size_t)s * no + o] <= 0.f) continue;
        float v = y[(size_t)s * no + o];
        if (best < 0 || v < bv) {
            best = o;
            bv = v;
        }
    }
    choice[s] = best;
}

}  // namespace mlpk

using namespace mlpk;

struct nnh_mlp {
    Dims d;
    float *P = nullptr, *G = nullptr, *M1 = nullptr, *M2 = nullptr;
    int step = 0;
    float inv_m = 0.f;  // 1 / (number of unmasked targets) of the uploaded batch
    // per-batch scratch (grown)
    int cap = 0;
    float *x = nullptr, *t = nullptr, *m = nullptr, *y = nullptr, *h1 = nullptr, *h2 = nullptr, *dy = nullptr,
          *dh1 = nullptr, *dh2 = nullptr, *sq = nullptr, *feas = nullptr;
    int* ch = nullptr;
};

#define MCHK(x)                                             \
    do {                                                    \
        if ((x) != hipSuccess) return NNH_ERR_HIP;          \
    } while (0)

static int mlp_grow(nnh_mlp* M, int n) {
    if (n <= M->cap) return NNH_OK;
    float** bufs[] = {&M->x, &M->t, &M->m, &M->y, &M->h1, &M->h2, &M->dy, &M->dh1, &M->dh2, &M->sq, &M->feas};
    for (float** b : bufs)
        if (*b) (void)hipFree(*b);
    if (M->ch) (void)hipFree(M->ch);
    const Dims& d = M->d;
    size_t sz[] = {(size_t)n * d.ni, (size_t)n * d.no, (size_t)n * d.no, (size_t)n * d.no, (size_t)n * H,
                   (size_t)n * H, (size_t)n * d.no, (size_t)n * H, (size_t)n * H, (size_t)n, (size_t)n * d.no};
    for (int i = 0; i < 11; i++) MCHK(hipMalloc((void**)bufs[i], sz[i] * sizeof(float)));
    MCHK(hipMalloc((void**)&M->ch, (size_t)n * sizeof(int)));
    M->cap = n;
    return NNH_OK;
}

extern "C" {

int nnh_mlp_create(int n_in, int n_out, uint64_t seed, nnh_mlp** out) {
    if (!out || n_in < 1 || n_in > MAX_IN || n_out < 1 || n_out > MAX_OUT) return NNH_ERR_ARG;
    nnh_mlp* M = new nnh_mlp();
    M->d = make_dims(n_in, n_out);
    int n = M->d.n;
    MCHK(hipMalloc((void**)&M->P, n * sizeof(float)));
    MCHK(hipMalloc((void**)&M->G, n * sizeof(float)));
    MCHK(hipMalloc((void**)&M->M1, n * sizeof(float)));
    MCHK(hipMalloc((void**)&M->M2, n * sizeof(float)));
    // He-style initialisation on the host (deterministic from seed)
    std::vector<float> p(n, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int cnt, int fan_in) {
        std::normal_distribution<float> nd(0.f, std::sqrt(2.f / fan_in));
        for (int i = 0; i < cnt; i++) p[off + i] = nd(rng);
    };
    fill(M->d.w1, H * n_in, n_in);
    fill(M->d.w2, H * H, H);
    fill(M->d.w3, n_out * H, H);
    MCHK(hipMemcpy(M->P, p.data(), n * sizeof(float), hipMemcpyHostToDevice));
    MCHK(hipMemset(M->M1, 0, n * sizeof(float)));
    MCHK(hipMemset(M->M2, 0, n * sizeof(float)));
    *out = M;
    return NNH_OK;
}

int nnh_mlp_destroy(nnh_mlp* M) {
    if (!M) return NNH_OK;
    (void)hipDeviceSynchronize();
    float* bufs[] = {M->P, M->G, M->M1, M->M2, M->x, M->t, M->m, M->y, M->h1, M->h2, M->dy, M->dh1, M->dh2, M->sq,
                     M->feas};
    for (float* b : bufs)
        if (b) (void)hipFree(b);
    if (M->ch) (void)hipFree(M->ch);
    delete M;
    return NNH_OK;
}

int nnh_mlp_num_params(nnh_mlp* M) { return M ? M->d.n : 0; }

int nnh_mlp_set_params(nnh_mlp* M, const float* p, uint64_t n) {
    if (!M || !p || n != (uint64_t)M->d.n) return NNH_ERR_ARG;
    MCHK(hipMemcpy(M->P, p, n * sizeof(float), hipMemcpyHostToDevice));
    return NNH_OK;
}

int nnh_mlp_get_params(nnh_mlp* M, float* p, uint64_t n) {
    if (!M || !p || n != (uint64_t)M->d.n) return NNH_ERR_ARG;
    MCHK(hipMemcpy(p, M->P, n * sizeof(float), hipMemcpyDeviceToHost));
    return NNH_OK;
}

int nnh_mlp_forward(nnh_mlp* M, const float* x, uint64_t n, float* y) {
    if (!M || !x || !y || !n) return NNH_ERR_ARG;
    int rc = mlp_grow(M, (int)n);
    if (rc) return rc;
    MCHK(hipMemcpy(M->x, x, n * M->d.ni * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_forward, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, 0, M->P, M->d, M->x, (int)n, M->y,
                       (float*)nullptr, (float*)nullptr);
    MCHK(hipGetLastError());
    MCHK(hipMemcpy(y, M->y, n * M->d.no * sizeof(float), hipMemcpyDeviceToHost));
    return NNH_OK;
}

static int mlp_grad_dev(nnh_mlp* M, int n, float* loss) {
    const Dims& d = M->d;
    const float inv_m = M->inv_m;
    hipLaunchKernelGGL(k_forward, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, 0, M->P, d, M->x, n, M->y, M->h1,
                       M->h2);
    hipLaunchKernelGGL(k_deltas, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, 0, M->P, d, n, M->y, M->t, M->m, inv_m,
                       M->h1, M->h2, M->dy, M->dh1, M->dh2, M->sq);
    hipLaunchKernelGGL(k_weight_grads, dim3((unsigned)((d.n + 255) / 256)), dim3(256), 0, 0, d, n, M->x, M->h1, M->h2,
                       M->dy, M->dh1, M->dh2, M->G);
    MCHK(hipGetLastError());
    if (loss) {
        std::vector<float> sq(n);
        MCHK(hipMemcpy(sq.data(), M->sq, n * sizeof(float), hipMemcpyDeviceToHost));
        double l = 0;
        for (float v : sq) l += v;
        *loss = (float)(l * inv_m);
    }
    return NNH_OK;
}

static int mlp_upload(nnh_mlp* M, const float* x, const float* t, const float* m, uint64_t n) {
    int rc = mlp_grow(M, (int)n);
    if (rc) return rc;
    MCHK(hipMemcpy(M->x, x, n * M->d.ni * sizeof(float), hipMemcpyHostToDevice));
    MCHK(hipMemcpy(M->t, t, n * M->d.no * sizeof(float), hipMemcpyHostToDevice));
    MCHK(hipMemcpy(M->m, m, n * M->d.no * sizeof(float), hipMemcpyHostToDevice));
    double msum = 0;
    for (uint64_t i = 0; i < n * M->d.no; i++) msum += m[i];
    M->inv_m = msum > 0 ? (float)(1.0 / msum) : 0.f;
    return NNH_OK;
}

int nnh_mlp_gradients(nnh_mlp* M, const float* x, const float* t, const float* m, uint64_t n, float* grad,
                      float* loss) {
    if (!M || !x || !t || !m || !grad || !n) return NNH_ERR_ARG;
    int rc = mlp_upload(M, x, t, m, n);
    if (rc || (rc = mlp_grad_dev(M, (int)n, loss))) return rc;
    MCHK(hipMemcpy(grad, M->G, M->d.n * sizeof(float), hipMemcpyDeviceToHost));
    return NNH_OK;
}

int nnh_mlp_train(nnh_mlp* M, const float* x, const float* t, const float* m, uint64_t n, int steps, float lr,
                  float* final_loss) {
    if (!M || !x || !t || !m || !n || steps < 0) return NNH_ERR_ARG;
    int rc = mlp_upload(M, x, t, m, n);
    if (rc) return rc;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-7f;
    for (int k = 0; k < steps; k++) {
        if ((rc = mlp_grad_dev(M, (int)n, k + 1 == steps ? final_loss : nullptr))) return rc;
        M->step++;
        float c1 = 1.f - std::pow(b1, (float)M->step), c2 = 1.f - std::pow(b2, (float)M->step);
        hipLaunchKernelGGL(k_adam, dim3((unsigned)((M->d.n + 255) / 256)), dim3(256), 0, 0, M->P, M->G, M->M1, M->M2,
                           M->d.n, lr, b1, b2, c1, c2, eps);
    }
    MCHK(hipGetLastError());
    MCHK(hipDeviceSynchronize());
    return NNH_OK;
}

int nnh_mlp_choose(nnh_mlp* M, const float* x, const float* feasible, uint64_t n, int* choice) {
    if (!M || !x || !feasible || !choice || !n) return NNH_ERR_ARG;
    int rc = mlp_grow(M, (int)n);
    if (rc) return rc;
    const Dims& d = M->d;
    MCHK(hipMemcpy(M->x, x, n * d.ni * sizeof(float), hipMemcpyHostToDevice));
    MCHK(hipMemcpy(M->feas, feasible, n * d.no * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_forward, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, 0, M->P, d, M->x, (int)n, M->y,
                       (float*)nullptr, (float*)nullptr);
    hipLaunchKernelGGL(k_choose, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, 0, M->y, M->feas, (int)n, d.no,
                       M->ch);
    MCHK(hipGetLastError());
    MCHK(hipMemcpy(choice, M->ch, n * sizeof(int), hipMemcpyDeviceToHost));
    return NNH_OK;
}

}  // extern "C"

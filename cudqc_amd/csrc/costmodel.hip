// HIP kernels of the subtree-cost model (forward, full-data gradient, Adam training)
// and their C-ABI (include/dqc.h). The model orders and splits the task batches of the
// learned schedule; it changes only the order in which tasks run, never the output.
//
// Training is deterministic: each workgroup processes chunks of 64 samples through LDS
// activation tiles, every thread accumulates the gradient of a fixed set of parameters
// in registers, per-workgroup partial gradients are reduced in a fixed order, and Adam
// applies the step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/dqc.h"
#include "costmodel.hpp"

namespace {

using namespace dqc_cost;

constexpr int GT = 256;                              // threads per training workgroup
constexpr int CH = 64;                               // samples per chunk
constexpr int OWN = (NPARAM + GT - 1) / GT;          // parameters owned per thread (19)
constexpr int MAX_BLOCKS = 512;

static_assert(NPARAM == DQC_MLP_PARAMS, "parameter count");
static_assert(IN == DQC_MLP_IN && H == DQC_MLP_HIDDEN, "model shape");

__global__ __launch_bounds__(256) void forward_kernel(const float* __restrict__ params, const float* __restrict__ x,
                                                      long long n, float* __restrict__ y) {
    __shared__ float p[NPARAM];
    for (int i = threadIdx.x; i < NPARAM; i += blockDim.x) p[i] = params[i];
    __syncthreads();
    for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (long long)gridDim.x * blockDim.x)
        y[s] = forward(p, x + IN * s);
}

// Gradient of sum_s scale * (y_s - t_s)^2 over samples idx[0..m) (idx null: 0..m-1).
// partial[block][p] gets the block's share, lpart[block] its sum of squared errors.
__global__ __launch_bounds__(GT) void grad_kernel(const float* __restrict__ p, const float* __restrict__ x,
                                                  const float* __restrict__ t, const int32_t* __restrict__ idx,
                                                  long long m, float scale, float* __restrict__ partial,
                                                  double* __restrict__ lpart) {
    __shared__ float xs[CH][IN];
    __shared__ float ts[CH];
    __shared__ float dy[CH];
    __shared__ float A[CH][H];   // first hidden layer (after ReLU)
    __shared__ float B[CH][H];   // second hidden layer, then its pre-activation gradient
    __shared__ float C[CH][H];   // pre-activation gradient of the first hidden layer
    __shared__ double lred[GT];
    const int tid = threadIdx.x;
    float acc[OWN];
#pragma unroll
    for (int k = 0; k < OWN; ++k) acc[k] = 0.f;
    double loss = 0.0;
    const long long nchunk = (m + CH - 1) / CH;
    for (long long c = blockIdx.x; c < nchunk; c += gridDim.x) {
        const long long base = c * CH;
        const int cnt = (int)std::min<long long>(CH, m - base);
        for (int e = tid; e < CH * IN; e += GT) {
            const int s = e / IN, f = e % IN;
            float v = 0.f;
            if (s < cnt) {
                const long long g = idx ? (long long)idx[base + s] : base + s;
                v = x[g * IN + f];
            }
            xs[s][f] = v;
        }
        if (tid < CH) {
            float v = 0.f;
            if (tid < cnt) v = t[idx ? (long long)idx[base + tid] : base + tid];
            ts[tid] = v;
        }
        __syncthreads();
        for (int e = tid; e < CH * H; e += GT) {  // layer 1
            const int s = e / H, j = e % H;
            float z = p[OFF_B1 + j];
#pragma unroll
            for (int f = 0; f < IN; ++f) z += p[OFF_W1 + j * IN + f] * xs[s][f];
            A[s][j] = z > 0.f ? z : 0.f;
        }
        __syncthreads();
        for (int e = tid; e < CH * H; e += GT) {  // layer 2
            const int s = e / H, j = e % H;
            float z = p[OFF_B2 + j];
            for (int i = 0; i < H; ++i) z += p[OFF_W2 + j * H + i] * A[s][i];
            B[s][j] = z > 0.f ? z : 0.f;
        }
        __syncthreads();
        if (tid < CH) {  // output, loss and its gradient
            float y = p[OFF_B3];
            for (int j = 0; j < H; ++j) y += p[OFF_W3 + j] * B[tid][j];
            const float r = tid < cnt ? y - ts[tid] : 0.f;
            dy[tid] = 2.f * scale * r;
            loss += (double)r * (double)r;
        }
        __syncthreads();
        // dW3, db3 (B still holds the second hidden layer)
#pragma unroll
        for (int k = 0; k < OWN; ++k) {
            const int q = tid + k * GT;
            if (q >= OFF_W3 && q < NPARAM) {
                float g = 0.f;
                if (q == OFF_B3) {
                    for (int s = 0; s < CH; ++s) g += dy[s];
                } else {
                    const int j = q - OFF_W3;
                    for (int s = 0; s < CH; ++s) g += dy[s] * B[s][j];
                }
                acc[k] += g;
            }
        }
        __syncthreads();
        for (int e = tid; e < CH * H; e += GT) {
            const int s = e / H, j = e % H;
            B[s][j] = B[s][j] > 0.f ? dy[s] * p[OFF_W3 + j] : 0.f;
        }
        __syncthreads();
        // dW2, db2
#pragma unroll
        for (int k = 0; k < OWN; ++k) {
            const int q = tid + k * GT;
            if (q >= OFF_W2 && q < OFF_W3) {
                float g = 0.f;
                if (q >= OFF_B2) {
                    const int j = q - OFF_B2;
                    for (int s = 0; s < CH; ++s) g += B[s][j];
                } else {
                    const int j = (q - OFF_W2) / H, i = (q - OFF_W2) % H;
                    for (int s = 0; s < CH; ++s) g += B[s][j] * A[s][i];
                }
                acc[k] += g;
            }
        }
        for (int e = tid; e < CH * H; e += GT) {
            const int s = e / H, i = e % H;
            float g = 0.f;
            if (A[s][i] > 0.f)
                for (int j = 0; j < H; ++j) g += B[s][j] * p[OFF_W2 + j * H + i];
            C[s][i] = g;
        }
        __syncthreads();
        // dW1, db1
#pragma unroll
        for (int k = 0; k < OWN; ++k) {
            const int q = tid + k * GT;
            if (q < OFF_W2) {
                float g = 0.f;
                if (q >= OFF_B1) {
                    const int j = q - OFF_B1;
                    for (int s = 0; s < CH; ++s) g += C[s][j];
                } else {
                    const int j = q / IN, f = q % IN;
                    for (int s = 0; s < CH; ++s) g += C[s][j] * xs[s][f];
                }
                acc[k] += g;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < OWN; ++k) {
        const int q = tid + k * GT;
        if (q < NPARAM) partial[(long long)blockIdx.x * NPARAM + q] = acc[k];
    }
    lred[tid] = loss;
    __syncthreads();
    if (tid == 0) {
        double sum = 0.0;
        for (int i = 0; i < GT; ++i) sum += lred[i];
        lpart[blockIdx.x] = sum;
    }
}

// grad[q] = sum over blocks of partial[b][q], in block order.
__global__ void reduce_kernel(const float* __restrict__ partial, int nblocks, float* __restrict__ grad) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= NPARAM) return;
    float g = 0.f;
    for (int b = 0; b < nblocks; ++b) g += partial[(long long)b * NPARAM + q];
    grad[q] = g;
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ mo,
                            float* __restrict__ ve, float lr, float c1, float c2) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= NPARAM) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float m = b1 * mo[q] + (1.f - b1) * g[q];
    const float v = b2 * ve[q] + (1.f - b2) * g[q] * g[q];
    mo[q] = m;
    ve[q] = v;
    p[q] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// Device buffers of one call; freed on scope exit.
struct Buf {
    std::vector<void*> ptrs;
    ~Buf() {
        for (void* q : ptrs) hipFree(q);
    }
    template <class T>
    T* alloc(size_t count) {
        void* q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return nullptr;
        ptrs.push_back(q);
        return (T*)q;
    }
};

int grad_blocks(long long m) {
    return (int)std::max<long long>(1, std::min<long long>(MAX_BLOCKS, (m + CH - 1) / CH));
}

// Full-data gradient on device buffers; returns the sum of squared errors.
bool device_gradient(const float* dp, const float* dx, const float* dt, long long n, float* dgrad, float* dpart,
                     double* dl, double* sse) {
    const int nb = grad_blocks(n);
    grad_kernel<<<nb, GT>>>(dp, dx, dt, nullptr, n, 1.f / (float)n, dpart, dl);
    reduce_kernel<<<(NPARAM + 255) / 256, 256>>>(dpart, nb, dgrad);
    std::vector<double> l(nb);
    if (hipMemcpy(l.data(), dl, sizeof(double) * nb, hipMemcpyDeviceToHost) != hipSuccess) return false;
    double s = 0.0;
    for (double v : l) s += v;
    *sse = s;
    return hipGetLastError() == hipSuccess;
}

}  // namespace

extern "C" {

dqc_status dqc_task_features(const dqc_task_log* r, int64_t n, float* x) {
    if (n < 0 || (n > 0 && (!r || !x))) return DQC_E_ARG;
    for (int64_t i = 0; i < n; ++i)
        features(r[i].nm, r[i].tot, r[i].depth, r[i].dmin_out, r[i].dmax_out, r[i].dmin_in, r[i].dmax_in, x + IN * i);
    return DQC_OK;
}

dqc_status dqc_mlp_forward(const float* params, const float* x, int64_t n, float* y) {
    if (!params || n < 0 || (n > 0 && (!x || !y))) return DQC_E_ARG;
    if (n == 0) return DQC_OK;
    Buf b;
    float *dp = b.alloc<float>(NPARAM), *dx = b.alloc<float>((size_t)n * IN), *dy = b.alloc<float>((size_t)n);
    if (!dp || !dx || !dy) return DQC_E_HIP;
    if (hipMemcpy(dp, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dx, x, sizeof(float) * IN * (size_t)n, hipMemcpyHostToDevice) != hipSuccess)
        return DQC_E_HIP;
    const int blocks = (int)std::min<long long>(4096, (n + 255) / 256);
    forward_kernel<<<blocks, 256>>>(dp, dx, n, dy);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(y, dy, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess)
        return DQC_E_HIP;
    return DQC_OK;
}

dqc_status dqc_mlp_gradient(const float* params, const float* x, const float* t, int64_t n, float* grad,
                            double* loss) {
    if (!params || !grad || n <= 0 || !x || !t) return DQC_E_ARG;
    Buf b;
    const int nb = grad_blocks(n);
    float *dp = b.alloc<float>(NPARAM), *dx = b.alloc<float>((size_t)n * IN), *dt = b.alloc<float>((size_t)n);
    float *dg = b.alloc<float>(NPARAM), *dpart = b.alloc<float>((size_t)nb * NPARAM);
    double* dl = b.alloc<double>(nb);
    if (!dp || !dx || !dt || !dg || !dpart || !dl) return DQC_E_HIP;
    if (hipMemcpy(dp, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dx, x, sizeof(float) * IN * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dt, t, sizeof(float) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess)
        return DQC_E_HIP;
    double sse = 0.0;
    if (!device_gradient(dp, dx, dt, n, dg, dpart, dl, &sse)) return DQC_E_HIP;
    if (hipMemcpy(grad, dg, sizeof(float) * NPARAM, hipMemcpyDeviceToHost) != hipSuccess) return DQC_E_HIP;
    if (loss) *loss = sse / (double)n;
    return DQC_OK;
}

dqc_status dqc_mlp_train(float* params, const float* x, const float* t, int64_t n, int32_t epochs, int32_t batch,
                         float lr, uint64_t seed, double* loss) {
    if (!params || !x || !t || n <= 0 || epochs < 0 || batch < 1 || !(lr > 0.f) || n > INT32_MAX) return DQC_E_ARG;
    Buf b;
    const int nb = grad_blocks(std::min<long long>(n, batch));
    const int nbf = grad_blocks(n);
    float *dp = b.alloc<float>(NPARAM), *dx = b.alloc<float>((size_t)n * IN), *dt = b.alloc<float>((size_t)n);
    float *dg = b.alloc<float>(NPARAM), *dm = b.alloc<float>(NPARAM), *dv = b.alloc<float>(NPARAM);
    float* dpart = b.alloc<float>((size_t)std::max(nb, nbf) * NPARAM);
    double* dl = b.alloc<double>(std::max(nb, nbf));
    int32_t* didx = b.alloc<int32_t>((size_t)n);
    if (!dp || !dx || !dt || !dg || !dm || !dv || !dpart || !dl || !didx) return DQC_E_HIP;
    if (hipMemcpy(dp, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dx, x, sizeof(float) * IN * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dt, t, sizeof(float) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(dm, 0, sizeof(float) * NPARAM) != hipSuccess || hipMemset(dv, 0, sizeof(float) * NPARAM) != hipSuccess)
        return DQC_E_HIP;
    std::vector<int32_t> order((size_t)n);
    for (int64_t i = 0; i < n; ++i) order[i] = (int32_t)i;
    std::mt19937_64 rng(seed);
    long long step = 0;
    for (int e = 0; e < epochs; ++e) {
        for (int64_t i = n - 1; i > 0; --i) {  // Fisher-Yates with a bounded draw
            const uint64_t r = rng() % (uint64_t)(i + 1);
            std::swap(order[i], order[r]);
        }
        if (hipMemcpy(didx, order.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess)
            return DQC_E_HIP;
        for (long long s0 = 0; s0 < n; s0 += batch) {
            const long long m = std::min<long long>(batch, n - s0);
            const int blocks = grad_blocks(m);
            grad_kernel<<<blocks, GT>>>(dp, dx, dt, didx + s0, m, 1.f / (float)m, dpart, dl);
            reduce_kernel<<<(NPARAM + 255) / 256, 256>>>(dpart, blocks, dg);
            ++step;
            const float c1 = 1.f - powf(0.9f, (float)step), c2 = 1.f - powf(0.999f, (float)step);
            adam_kernel<<<(NPARAM + 255) / 256, 256>>>(dp, dg, dm, dv, lr, c1, c2);
        }
        if (hipGetLastError() != hipSuccess) return DQC_E_HIP;
    }
    double sse = 0.0;
    if (!device_gradient(dp, dx, dt, n, dg, dpart, dl, &sse)) return DQC_E_HIP;
    if (hipMemcpy(params, dp, sizeof(float) * NPARAM, hipMemcpyDeviceToHost) != hipSuccess) return DQC_E_HIP;
    if (loss) *loss = sse / (double)n;
    return DQC_OK;
}

}  // extern "C"

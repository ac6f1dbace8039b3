// Cost model of the schedule tuner: a one-hidden-layer perceptron (4 inputs,
// 32 tanh units, one linear output) trained on the GPU by full-batch Adam on
// the mean squared error, all in FP64.  It has no counterpart in the
// reference (whose schedule is StarPU's); the tuner (mfdag/tune.py) feeds it
// the factorization times it measures.
//
// Parameters, 193 doubles: W1[j][i] at j*4 + i (j < 32, i < 4), then b1[32],
// w2[32], b2.  y(x) = b2 + sum_j w2[j] tanh(b1[j] + sum_i W1[j][i] x[i]).
// Loss = mean over samples of (y - t)^2.
//
// Every sum runs in a fixed order (samples of a workgroup in index order,
// then the workgroups' partials in index order), so results do not depend
// on scheduling.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "mfdag.h"

namespace {

constexpr int kIn = MFD_MLP_INPUTS;
constexpr int kHid = MFD_MLP_HIDDEN;
constexpr int kPar = MFD_MLP_PARAMS;
constexpr int kChunk = 128;  // samples per gradient workgroup
constexpr int kMaxSamples = 1 << 20;

static_assert(kPar == kHid * (kIn + 2) + 1, "parameter count");

// y for one sample; h receives the hidden activations when given.
__device__ __forceinline__ double mlp_eval(const double* p, const double* x, double* h)
{
    const double* W1 = p;
    const double* b1 = p + kHid * kIn;
    const double* w2 = b1 + kHid;
    double y = w2[kHid];
    for (int j = 0; j < kHid; ++j) {
        double a = b1[j];
        for (int i = 0; i < kIn; ++i) a = fma(W1[j * kIn + i], x[i], a);
        const double t = tanh(a);
        if (h) h[j] = t;
        y = fma(w2[j], t, y);
    }
    return y;
}

__global__ __launch_bounds__(256) void k_mlp_predict(const double* params, const double* x, int n, double* y)
{
    __shared__ double p[kPar];
    for (int k = threadIdx.x; k < kPar; k += blockDim.x) p[k] = params[k];
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    double xs[kIn];
    for (int i = 0; i < kIn; ++i) xs[i] = x[(int64_t)s * kIn + i];
    y[s] = mlp_eval(p, xs, nullptr);
}

// Workgroup b sums, over its samples s in [b*kChunk, (b+1)*kChunk), the
// gradient of (y_s - t_s)^2 * scale and the loss term into
// part[b*(kPar+1) + k] (k = kPar: the loss).  Phase 1: one thread per
// sample stores its error e = 2 scale (y - t), the hidden values and the
// back-propagated g_j = e w2[j] (1 - h_j^2) in LDS; phase 2: one thread per
// parameter adds the samples' terms in sample order.
__global__ __launch_bounds__(256) void k_mlp_grad(const double* params, const double* x, const double* t, int n,
                                                  double scale, double* part)
{
    __shared__ double p[kPar];
    __shared__ double H[kChunk][kHid + 1];
    __shared__ double G[kChunk][kHid + 1];
    __shared__ double X[kChunk][kIn];
    __shared__ double E[kChunk], Q[kChunk];
    for (int k = threadIdx.x; k < kPar; k += blockDim.x) p[k] = params[k];
    __syncthreads();
    const int s0 = blockIdx.x * kChunk, cnt = min(kChunk, n - s0);
    const int tid = threadIdx.x;
    if (tid < kChunk) {
        if (tid < cnt) {
            const int s = s0 + tid;
            double xs[kIn], h[kHid];
            for (int i = 0; i < kIn; ++i) xs[i] = X[tid][i] = x[(int64_t)s * kIn + i];
            const double y = mlp_eval(p, xs, h);
            const double r = y - t[s];
            const double e = 2.0 * scale * r;
            E[tid] = e;
            Q[tid] = scale * r * r;
            const double* w2 = p + kHid * kIn + kHid;
            for (int j = 0; j < kHid; ++j) {
                H[tid][j] = h[j];
                G[tid][j] = e * w2[j] * (1.0 - h[j] * h[j]);
            }
        }
    }
    __syncthreads();
    double* out = part + (int64_t)blockIdx.x * (kPar + 1);
    for (int k = tid; k <= kPar; k += blockDim.x) {
        double acc = 0.0;
        if (k < kHid * kIn) {  // W1[j][i]: g_j x_i
            const int j = k / kIn, i = k % kIn;
            for (int q = 0; q < cnt; ++q) acc = fma(G[q][j], X[q][i], acc);
        } else if (k < kHid * kIn + kHid) {  // b1[j]: g_j
            const int j = k - kHid * kIn;
            for (int q = 0; q < cnt; ++q) acc += G[q][j];
        } else if (k < kHid * kIn + 2 * kHid) {  // w2[j]: e h_j
            const int j = k - kHid * kIn - kHid;
            for (int q = 0; q < cnt; ++q) acc = fma(E[q], H[q][j], acc);
        } else if (k == kPar - 1) {  // b2: e
            for (int q = 0; q < cnt; ++q) acc += E[q];
        } else {  // the loss
            for (int q = 0; q < cnt; ++q) acc += Q[q];
        }
        out[k] = acc;
    }
}

// grad[k] = sum over the nblk partials in order (k = kPar: the loss).  With
// lr > 0, one Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected
// with step number it) on params with moments m, v.
__global__ __launch_bounds__(256) void k_mlp_reduce(const double* part, int nblk, double* grad, double* params,
                                                    double* m, double* v, double lr, int it)
{
    for (int k = threadIdx.x; k <= kPar; k += blockDim.x) {
        double g = 0.0;
        for (int b = 0; b < nblk; ++b) g += part[(int64_t)b * (kPar + 1) + k];
        grad[k] = g;
        if (lr > 0.0 && k < kPar) {
            const double mk = 0.9 * m[k] + 0.1 * g, vk = 0.999 * v[k] + 0.001 * g * g;
            m[k] = mk;
            v[k] = vk;
            const double mh = mk / (1.0 - pow(0.9, (double)it)), vh = vk / (1.0 - pow(0.999, (double)it));
            params[k] -= lr * mh / (sqrt(vh) + 1e-8);
        }
    }
}

bool device_ready()
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return false;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess || strncmp(prop.gcnArchName, "gfx950", 6) != 0) return false;
    return hipSetDevice(0) == hipSuccess;
}

// Device buffers of one call, freed on every return path.
struct Buffers {
    std::vector<void*> ptrs;
    hipStream_t st = nullptr;
    ~Buffers()
    {
        if (st) (void)hipStreamSynchronize(st);
        for (void* p : ptrs) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
    }
    template <typename T> bool get(T** p, size_t count)
    {
        void* q = nullptr;
        if (hipMalloc(&q, count * sizeof(T)) != hipSuccess) return false;
        ptrs.push_back(q);
        *p = (T*)q;
        return true;
    }
};

bool args_ok(int n, const double* x, const double* params)
{
    return n >= 1 && n <= kMaxSamples && x && params;
}

// Gradient (lr = 0, one pass) or Adam training (epochs passes) on the GPU.
int run(int n, const double* x, const double* t, double* params, int epochs, double lr, double* grad_out,
        double* loss)
{
    if (!device_ready()) return MFD_ERR_NODEVICE;
    Buffers B;
    if (hipStreamCreateWithFlags(&B.st, hipStreamNonBlocking) != hipSuccess) return MFD_ERR_HIP;
    const int nblk = (n + kChunk - 1) / kChunk;
    double *dx, *dt, *dp, *dpart, *dg, *dm, *dv;
    if (!B.get(&dx, (size_t)n * kIn) || !B.get(&dt, n) || !B.get(&dp, kPar) || !B.get(&dpart, (size_t)nblk * (kPar + 1)) ||
        !B.get(&dg, kPar + 1) || !B.get(&dm, kPar) || !B.get(&dv, kPar))
        return MFD_ERR_NOMEM;
    if (hipMemcpyAsync(dx, x, sizeof(double) * n * kIn, hipMemcpyHostToDevice, B.st) != hipSuccess ||
        hipMemcpyAsync(dt, t, sizeof(double) * n, hipMemcpyHostToDevice, B.st) != hipSuccess ||
        hipMemcpyAsync(dp, params, sizeof(double) * kPar, hipMemcpyHostToDevice, B.st) != hipSuccess ||
        hipMemsetAsync(dm, 0, sizeof(double) * kPar, B.st) != hipSuccess ||
        hipMemsetAsync(dv, 0, sizeof(double) * kPar, B.st) != hipSuccess)
        return MFD_ERR_HIP;
    const double scale = 1.0 / n;
    const int passes = grad_out ? 1 : epochs;
    for (int it = 1; it <= passes; ++it) {
        k_mlp_grad<<<nblk, 256, 0, B.st>>>(dp, dx, dt, n, scale, dpart);
        k_mlp_reduce<<<1, 256, 0, B.st>>>(dpart, nblk, dg, dp, dm, dv, grad_out ? 0.0 : lr, it);
    }
    if (hipGetLastError() != hipSuccess) return MFD_ERR_HIP;
    std::vector<double> g(kPar + 1);
    if (hipMemcpyAsync(g.data(), dg, sizeof(double) * (kPar + 1), hipMemcpyDeviceToHost, B.st) != hipSuccess)
        return MFD_ERR_HIP;
    if (!grad_out && hipMemcpyAsync(params, dp, sizeof(double) * kPar, hipMemcpyDeviceToHost, B.st) != hipSuccess)
        return MFD_ERR_HIP;
    if (hipStreamSynchronize(B.st) != hipSuccess) return MFD_ERR_HIP;
    if (grad_out) std::memcpy(grad_out, g.data(), sizeof(double) * kPar);
    if (loss) *loss = g[kPar];
    return MFD_OK;
}

}  // namespace

extern "C" int mfd_mlp_predict(int n, const double* x, const double* params, double* y)
{
    if (!args_ok(n, x, params) || !y) return MFD_ERR_ARG;
    if (!device_ready()) return MFD_ERR_NODEVICE;
    Buffers B;
    if (hipStreamCreateWithFlags(&B.st, hipStreamNonBlocking) != hipSuccess) return MFD_ERR_HIP;
    double *dx, *dp, *dy;
    if (!B.get(&dx, (size_t)n * kIn) || !B.get(&dp, kPar) || !B.get(&dy, n)) return MFD_ERR_NOMEM;
    if (hipMemcpyAsync(dx, x, sizeof(double) * n * kIn, hipMemcpyHostToDevice, B.st) != hipSuccess ||
        hipMemcpyAsync(dp, params, sizeof(double) * kPar, hipMemcpyHostToDevice, B.st) != hipSuccess)
        return MFD_ERR_HIP;
    k_mlp_predict<<<(n + 255) / 256, 256, 0, B.st>>>(dp, dx, n, dy);
    if (hipGetLastError() != hipSuccess) return MFD_ERR_HIP;
    if (hipMemcpyAsync(y, dy, sizeof(double) * n, hipMemcpyDeviceToHost, B.st) != hipSuccess ||
        hipStreamSynchronize(B.st) != hipSuccess)
        return MFD_ERR_HIP;
    return MFD_OK;
}

extern "C" int mfd_mlp_gradient(int n, const double* x, const double* t, const double* params, double* grad,
                                double* loss)
{
    if (!args_ok(n, x, params) || !t || !grad) return MFD_ERR_ARG;
    return run(n, x, t, const_cast<double*>(params), 1, 0.0, grad, loss);
}

extern "C" int mfd_mlp_train(int n, const double* x, const double* t, double* params, int epochs, double lr,
                             double* loss)
{
    if (!args_ok(n, x, params) || !t || epochs < 1 || !(lr > 0.0)) return MFD_ERR_ARG;
    return run(n, x, t, params, epochs, lr, nullptr, loss);
}

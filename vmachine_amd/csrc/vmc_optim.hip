// vmc_optim.hip -- the device side of an optimisation loop of one system that
// never returns to the host between iterations: from the iteration's sums
// (per-walker accumulator totals and the RBM gradient-window rows) the
// reference's energy gradient g = 2(<G E> - <E><G>) (ref
// src/Optimization/optimization.cpp:10-16, <E> over all samples, <G> over the
// gradient window, ref src/sampler.cpp:216-217) and the optimiser step of
// every entry of the parameter matrix (ADAM, ref src/Optimization/adam.cpp:23-38;
// SGD / gradient descent, ref sgd.cpp:23-29, gradientdescent.cpp:24-30), the
// updated parameters written where the walk reads them, then the advance of
// the random counter and gradient window.  The walk kernel reads the
// iteration's values from OptState, so one captured iteration replays.
#include <hip/hip_runtime.h>

#include <mutex>

#include "vmc_device.h"

namespace vmc {

// one thread per entry of the rows x mp parameter matrix (padding entries are
// updated too, as the reference's optimisers do)
__global__ __launch_bounds__(256) void opt_step_kernel(const OptLayout L, const double* __restrict__ acc,
                                                       const double* __restrict__ rbm, double* theta,
                                                       double* m, double* v, const OptState* st,
                                                       double* ag, double* rbm_a, double* rbm_b,
                                                       double* rbm_W, double* hist, int it0)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = L.rows * L.mp;
    double* h = hist + (size_t)(st->iteration - it0) * kOptOut;
    if (i == 0) {
        for (int k = 0; k < kNumWalkerAcc; k++) h[k] = acc[k];
        // every entry of the row is written: alpha / gamma are 0 without their element
        if (L.row_gauss < 0) h[kOptAlpha] = 0.0;
        if (L.row_pade < 0) h[kOptGamma] = 0.0;
    }
    if (i >= n) return;
    const int r = i / L.mp, c = i - r * L.mp;
    const int half = L.nd + L.H + L.nd * L.H;   // RBM rows: [a | b | W] sums, then E x the same
    double G = 0, GE = 0;
    if (r == L.row_gauss && c == 0) {
        G = acc[kAccGAlpha];
        GE = acc[kAccGAlphaE];
    } else if (r == L.row_pade && c == 0) {
        G = acc[kAccGGamma];
        GE = acc[kAccGGammaE];
    } else if (r == L.row_rbmg && c < L.nd) {
        G = rbm[c];
        GE = rbm[half + c];
    } else if (r == L.row_rbmp && c < L.H + L.nd * L.H) {
        G = rbm[L.nd + c];
        GE = rbm[half + L.nd + c];
    }
    const double E = acc[kAccE] / L.M;
    const double g = 2.0 * (GE / L.Sw - E * (G / L.Sw));
    const int step = st->step + 1;
    double upd;
    if (L.kind == 0) {
        const double b1 = 0.89, b2 = 0.999, eps = 1e-8;   // ref adam.h
        m[i] = b1 * m[i] + (1 - b1) * g;
        v[i] = b2 * v[i] + (1 - b2) * g * g;
        const double mh = m[i] / (1 - pow(b1, double(step)));
        const double vh = v[i] / (1 - pow(b2, double(step)));
        upd = L.lr * mh / (sqrt(vh) + eps);
    } else {
        const double mono = 1.0 / pow(double(step), L.monotonic);
        v[i] = L.gamma * v[i] + L.lr * g * mono;
        upd = v[i];
    }
    const double t = theta[i] - upd;   // ref system.cpp:114
    theta[i] = t;
    // where the walk reads it
    if (r == L.row_gauss && c == 0) {
        ag[0] = t;
        h[kOptAlpha] = t;
    } else if (r == L.row_pade && c == 0) {
        ag[1] = t;
        h[kOptGamma] = t;
    } else if (r == L.row_rbmg && c < L.nd) {
        rbm_a[c] = t;
    } else if (r == L.row_rbmp && c < L.H) {
        rbm_b[c] = t;
    } else if (r == L.row_rbmp && c < L.H + L.nd * L.H) {
        const int q = c - L.H, k = q % L.nd, hh = q / L.nd;   // W column-major (ND x H)
        rbm_W[k * L.H + hh] = t;
    }
}

__global__ void opt_advance_kernel(OptState* st, int steps_total)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st->counter0 += (uint64_t)steps_total;
        st->iteration += 1;
        st->fresh = 0;
        st->step += 1;
    }
}

hipError_t launch_opt_tail(const OptLayout& L, const double* acc, const double* rbm, double* theta,
                           double* m, double* v, OptState* st, double* ag, double* rbm_a,
                           double* rbm_b, double* rbm_W, double* hist, int it0, int steps_total,
                           hipStream_t stream)
{
    const int n = L.rows * L.mp;
    // one block at least: thread 0 writes the iteration's history row
    hipLaunchKernelGGL(opt_step_kernel, dim3(n > 0 ? (n + 255) / 256 : 1), dim3(256), 0, stream, L, acc, rbm,
                       theta, m, v, st, ag, rbm_a, rbm_b, rbm_W, hist, it0);
    hipLaunchKernelGGL(opt_advance_kernel, dim3(1), dim3(64), 0, stream, st, steps_total);
    return hipGetLastError();
}

// ---- device state of the allocation-free loop ----
// One copy per GPU, held until the process ends.  Host code never passes the
// array's name to a kernel: it uses the device address hipGetSymbolAddress
// returns for the current device.
static __device__ __align__(256) double scratch[kStaticDoubles];

__global__ void fill_kernel(double* p, int n, double value)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

hipError_t launch_fill(double* p, int n, double value, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, p, n, value);
    return hipGetLastError();
}

// The device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
hipError_t static_scratch(double** out)
{
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static double* addr[kMaxDevices] = {};
    int dev = 0;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess) return err;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess) return err;
        if (!p) return hipErrorInvalidSymbol;
        addr[dev] = static_cast<double*>(p);
    }
    *out = addr[dev];
    return hipSuccess;
}

}  // namespace vmc

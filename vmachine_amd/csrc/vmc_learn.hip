// vmc_learn.hip -- training kernels of the learned Metropolis proposal.
//
// The proposal network (18 -> 64 -> 64 -> 4, tanh) maps features of one
// electron (its coordinates and the relative coordinates of its 8 nearest
// electrons, units of 1/sqrt(omega)) to a Gaussian move: mean mu and scale
// sigma.  It is trained to maximise the expected squared jump distance
//   J = A |dx|^2 omega,  A = min(1, exp(L)),
//   L = log|psi(x')/psi(x)|^2 + log q(x|x') - log q(x'|x),
// with the reparameterised move dx = mu(x) + sigma(x) xi.  One training sample
// per walker: electron i and normals xi from Philox, positions from the
// walker.  The pipeline (all on the engine's stream):
//   lp_draws_kernel        i, xi per sample
//   lp_propose_kernel      forward passes at x and x' (one thread per sample,
//                          FP32, weights in LDS), the proposed position
//   (walk eval kernel)     |psi(x')/psi(x)|^2 and the drift at x' of electron i
//   lp_backward_kernel     L, A, J and the backward passes: per-sample layer
//                          deltas of both network passes
//   lp_wgrad_kernel        dJ/dtheta = sum over samples of delta x activation,
//                          one partial row per block of samples
//   (reduce_rows_kernel)   the rows, in a fixed order
//   lp_adam_kernel         gradient-ascent ADAM step on the FP32 master
//                          weights and their bfloat16 copy for the walk
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vmc_device.h"

namespace vmc {

namespace {

constexpr int kAct = kLpIn + 2 * kLpHid + kLpOut;    // f, h1, h2, o of one pass
constexpr int kDelta = 2 * kLpHid + kLpOut;          // d1, d2, d3 of one pass

__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
}

// features of electron i placed at (xi, yi); returns the neighbour count
__device__ int features(const double* x, int N, double s, int i, double xi, double yi, float* f)
{
    int idx[8];
    double dd[8];
    int n = 0;
    for (int j = 0; j < N; j++) {
        if (j == i) continue;
        const double e0 = x[2 * j] - xi, e1 = x[2 * j + 1] - yi;
        const double d2 = e0 * e0 + e1 * e1;
        // insertion into the sorted list of the 8 nearest (ties keep the lower index first)
        if (n < 8 || d2 < dd[n - 1]) {
            int p = n < 8 ? n : 7;
            while (p > 0 && d2 < dd[p - 1]) {
                dd[p] = dd[p - 1];
                idx[p] = idx[p - 1];
                p--;
            }
            dd[p] = d2;
            idx[p] = j;
            if (n < 8) n++;
        }
    }
    f[0] = float(s * xi);
    f[1] = float(s * yi);
    for (int m = 0; m < 8; m++) {
        f[2 + 2 * m] = m < n ? float(s * (x[2 * idx[m]] - xi)) : 0.0f;
        f[3 + 2 * m] = m < n ? float(s * (x[2 * idx[m] + 1] - yi)) : 0.0f;
    }
    return n;
}

__device__ __forceinline__ float tanhf_(float a)
{
    const float e = __expf(2.0f * a);
    return 1.0f - 2.0f / (e + 1.0f);
}

// forward of one pass; act = [f | h1 | h2 | o]
__device__ void forward(const float* th, float* act)
{
    const float* f = act;
    float* h1 = act + kLpIn;
    float* h2 = h1 + kLpHid;
    float* o = h2 + kLpHid;
    for (int u = 0; u < kLpHid; u++) {
        float a = th[kLpB1 + u];
        for (int m = 0; m < kLpIn; m++) a += th[kLpW1 + u * kLpIn + m] * f[m];
        h1[u] = tanhf_(a);
    }
    for (int u = 0; u < kLpHid; u++) {
        float a = th[kLpB2 + u];
        for (int c = 0; c < kLpHid; c++) a += th[kLpW2 + u * kLpHid + c] * h1[c];
        h2[u] = tanhf_(a);
    }
    for (int r = 0; r < kLpOut; r++) {
        float a = th[kLpB3 + r];
        for (int u = 0; u < kLpHid; u++) a += th[kLpW3 + r * kLpHid + u] * h2[u];
        o[r] = a;
    }
}

// backward of one pass: g_o = dX/do; writes scale * (d1, d2, d3) into delta
// and, if g_f is non-null, dX/df (unscaled)
__device__ void backward(const float* th, const float* act, const float* g_o, float scale,
                         float* delta, float* g_f)
{
    const float* h1 = act + kLpIn;
    const float* h2 = h1 + kLpHid;
    float* d1 = delta;
    float* d2 = delta + kLpHid;
    float* d3 = d2 + kLpHid;
    for (int r = 0; r < kLpOut; r++) d3[r] = scale * g_o[r];
    for (int u = 0; u < kLpHid; u++) {
        float g = 0;
        for (int r = 0; r < kLpOut; r++) g += th[kLpW3 + r * kLpHid + u] * g_o[r];
        d2[u] = g * (1.0f - h2[u] * h2[u]);   // unscaled for now
    }
    for (int c = 0; c < kLpHid; c++) {
        float g = 0;
        for (int u = 0; u < kLpHid; u++) g += th[kLpW2 + u * kLpHid + c] * d2[u];
        d1[c] = g * (1.0f - h1[c] * h1[c]);
    }
    if (g_f) {
        for (int m = 0; m < kLpIn; m++) {
            float g = 0;
            for (int c = 0; c < kLpHid; c++) g += th[kLpW1 + c * kLpIn + m] * d1[c];
            g_f[m] = g;
        }
    }
    for (int u = 0; u < kLpHid; u++) {
        d1[u] *= scale;
        d2[u] *= scale;
    }
}

__device__ __forceinline__ void head(const LpTrainParams& T, const float* o, float* mu, float* sg,
                                     float* tm, float* ts)
{
    const float ell = float(1.0 / sqrt(T.omega));
    for (int d = 0; d < 2; d++) {
        tm[d] = tanhf_(o[d] / float(T.mu_max));
        ts[d] = tanhf_(o[2 + d] / float(T.ls_max));
        mu[d] = ell * float(T.mu_max) * tm[d];
        sg[d] = float(T.sigma0) * ell * __expf(float(T.ls_max) * ts[d]);
    }
}

__device__ void load_theta(const float* theta, float* th)
{
    for (int i = threadIdx.x; i < kLpParams; i += blockDim.x) th[i] = theta[i];
    __syncthreads();
}

}  // namespace

// electron and normals of each training sample: Philox counter (iteration,
// sample), stream 2 (the walk uses stream 0, initial positions stream 1)
__global__ void lp_draws_kernel(LpTrainParams T, uint64_t iteration, int* electrons,
                                double* normals)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= T.B) return;
    uint32_t c[4] = {(uint32_t)iteration, (uint32_t)(iteration >> 32), (uint32_t)b, 2u};
    philox(c, (uint32_t)T.seed, (uint32_t)(T.seed >> 32));
    electrons[b] = (int)__umulhi(c[0], (uint32_t)T.N);
    const double U1 = (double(c[1] >> 9) + 0.5) * (1.0 / 8388608.0);
    const double U2 = (double(c[2] >> 9) + 0.5) * (1.0 / 8388608.0);
    const double r = sqrt(-2.0 * log(U1));
    normals[2 * b] = r * cos(6.283185307179586476925286766559 * U2);
    normals[2 * b + 1] = r * sin(6.283185307179586476925286766559 * U2);
}

__global__ __launch_bounds__(64) void lp_propose_kernel(LpTrainParams T)
{
    __shared__ float th[kLpParams];
    load_theta(T.theta, th);
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= T.B) return;
    const int nd = 2 * T.N, i = T.electrons[b];
    const double* x = T.pos + (size_t)b * nd;
    const double s = sqrt(T.omega);
    float* actA = T.act + (size_t)b * 2 * kAct;
    float* actB = actA + kAct;
    features(x, T.N, s, i, x[2 * i], x[2 * i + 1], actA);
    forward(th, actA);
    float mu[2], sg[2], tm[2], ts[2];
    head(T, actA + kAct - kLpOut, mu, sg, tm, ts);
    const double xn = x[2 * i] + double(mu[0] + sg[0] * float(T.normals[2 * b]));
    const double yn = x[2 * i + 1] + double(mu[1] + sg[1] * float(T.normals[2 * b + 1]));
    T.xnew[2 * b] = xn;
    T.xnew[2 * b + 1] = yn;
    T.nnb[b] = features(x, T.N, s, i, xn, yn, actB);
    forward(th, actB);
}

__global__ __launch_bounds__(64) void lp_backward_kernel(LpTrainParams T)
{
    __shared__ float th[kLpParams];
    load_theta(T.theta, th);
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= T.B) return;
    const int nd = 2 * T.N, i = T.electrons[b];
    const double* x = T.pos + (size_t)b * nd;
    const float omega = float(T.omega), s = sqrtf(omega), ell = 1.0f / s;
    const float* actA = T.act + (size_t)b * 2 * kAct;
    const float* actB = actA + kAct;
    float* dA = T.delta + (size_t)b * 2 * kDelta;
    float* dB = dA + kDelta;
    float mu[2], sg[2], tm[2], ts[2], mub[2], sgb[2], tmb[2], tsb[2];
    head(T, actA + kAct - kLpOut, mu, sg, tm, ts);
    head(T, actB + kAct - kLpOut, mub, sgb, tmb, tsb);
    const float xi[2] = {float(T.normals[2 * b]), float(T.normals[2 * b + 1])};
    float dx[2], r[2];
    dx[0] = float(T.xnew[2 * b] - x[2 * i]);
    dx[1] = float(T.xnew[2 * b + 1] - x[2 * i + 1]);
    float lqf = 0, lqr = 0;
    for (int d = 0; d < 2; d++) {
        r[d] = -dx[d] - mub[d];
        lqf -= 0.5f * xi[d] * xi[d] + __logf(sg[d]);
        lqr -= r[d] * r[d] / (2.0f * sgb[d] * sgb[d]) + __logf(sgb[d]);
    }
    // |psi(x')/psi(x)|^2 and the drift of electron i at x' from the walk's eval kernel
    const double* o1 = T.eval + (size_t)b * (9 + nd);
    const double ratio = o1[6] * o1[7] * o1[8];
    const float logR = float(log(ratio));
    const float F[2] = {float(o1[9 + 2 * i]), float(o1[9 + 2 * i + 1])};
    const float L = logR + lqr - lqf;
    const float A = L < 0.0f ? __expf(L) : 1.0f;
    const float d2 = (dx[0] * dx[0] + dx[1] * dx[1]) * omega;
    const float J = A * d2;
    const float gL = (L < 0.0f ? A : 0.0f) * d2;
    // d log q_r / d o_b, then the backward of pass B (scaled by dJ/dL)
    float gob[kLpOut], gfb[kLpIn];
    for (int d = 0; d < 2; d++) {
        const float is2 = 1.0f / (sgb[d] * sgb[d]);
        gob[d] = r[d] * is2 * ell * (1.0f - tmb[d] * tmb[d]);
        gob[2 + d] = (r[d] * r[d] * is2 - 1.0f) * (1.0f - tsb[d] * tsb[d]);
    }
    backward(th, actB, gob, gL, dB, gfb);
    // dJ/d dx, then the backward of pass A
    const int nnb = T.nnb[b];
    float goa[kLpOut];
    for (int d = 0; d < 2; d++) {
        float viaf = gfb[d];
        for (int m = 0; m < nnb; m++) viaf -= gfb[2 + 2 * m + d];
        const float dLdx = F[d] + r[d] / (sgb[d] * sgb[d]) + s * viaf;
        const float gdx = A * 2.0f * dx[d] * omega + gL * dLdx;
        const float dJdsig = gdx * xi[d] + gL / sg[d];
        goa[d] = gdx * ell * (1.0f - tm[d] * tm[d]);
        goa[2 + d] = dJdsig * sg[d] * (1.0f - ts[d] * ts[d]);
    }
    backward(th, actA, goa, 1.0f, dA, nullptr);
    double* st = T.stats + (size_t)b * 5;
    st[0] = J;
    st[1] = A;
    st[2] = logR;
    st[3] = T.xnew[2 * b];
    st[4] = T.xnew[2 * b + 1];
}

// one row of dJ/dtheta per block of kChunk samples: sum over the samples and
// both passes of delta x activation (W1: d1 f, W2: d2 h1, W3: d3 h2; biases: deltas)
constexpr int kChunk = 16;

__global__ __launch_bounds__(256) void lp_wgrad_kernel(LpTrainParams T, double* rows)
{
    __shared__ float act[kChunk * 2 * kAct];
    __shared__ float del[kChunk * 2 * kDelta];
    const int b0 = blockIdx.x * kChunk;
    const int nb = (T.B - b0) < kChunk ? (T.B - b0) : kChunk;
    for (int i = threadIdx.x; i < nb * 2 * kAct; i += blockDim.x)
        act[i] = T.act[(size_t)b0 * 2 * kAct + i];
    for (int i = threadIdx.x; i < nb * 2 * kDelta; i += blockDim.x)
        del[i] = T.delta[(size_t)b0 * 2 * kDelta + i];
    __syncthreads();
    double* row = rows + (size_t)blockIdx.x * kLpParams;
    for (int p = threadIdx.x; p < kLpParams; p += blockDim.x) {
        // which delta and which activation this parameter multiplies
        int di, ai;   // offsets within one pass's delta / activation vectors (ai < 0: bias)
        if (p < kLpB1) {
            di = p / kLpIn;                      // d1[u]
            ai = p % kLpIn;                      // f[m]
        } else if (p < kLpW2) {
            di = p - kLpB1;
            ai = -1;
        } else if (p < kLpB2) {
            const int q = p - kLpW2;
            di = kLpHid + q / kLpHid;            // d2[u]
            ai = kLpIn + q % kLpHid;             // h1[c]
        } else if (p < kLpW3) {
            di = kLpHid + (p - kLpB2);
            ai = -1;
        } else if (p < kLpB3) {
            const int q = p - kLpW3;
            di = 2 * kLpHid + q / kLpHid;        // d3[r]
            ai = kLpIn + kLpHid + q % kLpHid;    // h2[u]
        } else {
            di = 2 * kLpHid + (p - kLpB3);
            ai = -1;
        }
        float sum = 0.0f;
        for (int b = 0; b < nb; b++)
            for (int pass = 0; pass < 2; pass++) {
                const float dv = del[(b * 2 + pass) * kDelta + di];
                sum += ai < 0 ? dv : dv * act[(b * 2 + pass) * kAct + ai];
            }
        row[p] = sum;
    }
}

// gradient ascent: theta += lr mhat / (sqrt(vhat) + eps), grad = sum / B
__global__ void lp_adam_kernel(float* theta, uint16_t* bf16, float* m, float* v, const double* grad,
                               double inv_b, double lr, int step)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kLpParams) return;
    const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
    const double g = grad[p] * inv_b;
    const double mm = b1 * m[p] + (1 - b1) * g;
    const double vv = b2 * v[p] + (1 - b2) * g * g;
    m[p] = float(mm);
    v[p] = float(vv);
    const double mh = mm / (1 - pow(b1, step)), vh = vv / (1 - pow(b2, step));
    const float t = float(double(theta[p]) + lr * mh / (sqrt(vh) + eps));
    theta[p] = t;
    // round to nearest even bfloat16
    const uint32_t u = __float_as_uint(t);
    bf16[p] = (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

hipError_t launch_lp_draws(const LpTrainParams& T, uint64_t iteration, int* electrons,
                           double* normals, hipStream_t stream)
{
    hipLaunchKernelGGL(lp_draws_kernel, dim3((T.B + 255) / 256), dim3(256), 0, stream, T, iteration,
                       electrons, normals);
    return hipGetLastError();
}

hipError_t launch_lp_propose(const LpTrainParams& T, hipStream_t stream)
{
    hipLaunchKernelGGL(lp_propose_kernel, dim3((T.B + 63) / 64), dim3(64), 0, stream, T);
    return hipGetLastError();
}

hipError_t launch_lp_backward(const LpTrainParams& T, hipStream_t stream)
{
    hipLaunchKernelGGL(lp_backward_kernel, dim3((T.B + 63) / 64), dim3(64), 0, stream, T);
    return hipGetLastError();
}

int lp_wgrad_rows(int B) { return (B + kChunk - 1) / kChunk; }

hipError_t launch_lp_wgrad(const LpTrainParams& T, double* rows, hipStream_t stream)
{
    hipLaunchKernelGGL(lp_wgrad_kernel, dim3(lp_wgrad_rows(T.B)), dim3(256), 0, stream, T, rows);
    return hipGetLastError();
}

hipError_t launch_lp_adam(float* theta, uint16_t* bf16, float* m, float* v, const double* grad,
                          double inv_b, double lr, int step, hipStream_t stream)
{
    hipLaunchKernelGGL(lp_adam_kernel, dim3((kLpParams + 255) / 256), dim3(256), 0, stream, theta,
                       bf16, m, v, grad, inv_b, lr, step);
    return hipGetLastError();
}

int lp_act_floats() { return 2 * kAct; }
int lp_delta_floats() { return 2 * kDelta; }

}  // namespace vmc

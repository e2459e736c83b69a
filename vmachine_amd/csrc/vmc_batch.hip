// vmc_batch.hip -- several quantum dots (different N and omega) walked by one
// persistent launch, and the device-side pieces of one optimisation iteration
// that a hipGraph replays: per-system reduction, the reference's ADAM update
// of the Gaussian and Pade-Jastrow parameters (ref src/Optimization/adam.cpp:12-38,
// src/Optimization/optimization.cpp:10-16) and the advance of the random
// counter and gradient window.
//
// The walk kernel's device code (walk_block) is shared with vmc_walk.hip: one
// workgroup (one wave) takes tiles from a work queue; a tile is `waves` blocks
// of one system, walked one after the other with that system's (np, K)
// instantiation.  Tiles of the costliest systems come first.
#define VMC_WALK_PRECISION 99   // the walker templates only
#include "vmc_walk.hip"
#undef VMC_WALK_PRECISION

namespace vmc {

// the (np, K) shapes the batch kernel carries (Slater x Gaussian x Pade only)
#define VMC_BATCH_SHAPES(X)                                                                  \
    X(0, 1, 1) X(1, 1, 2) X(2, 3, 4) X(3, 3, 8) X(4, 6, 8) X(5, 6, 16) X(6, 10, 16)           \
    X(7, 10, 32) X(8, 15, 16) X(9, 15, 32) X(10, 21, 32) X(11, 28, 32)

// the system table is read through the constant address space: with the
// system index uniform across the wave every field loads with a scalar load
// when the walk needs it, as the single-system kernel reads its kernel
// arguments, instead of living in registers (or in scratch) for the whole walk
#define VMC_CONST_AS __attribute__((address_space(4)))

// FP32: two waves per SIMD, as the single-system N=56 kernel runs (the union
// of all shapes would otherwise take more than 256 registers and one wave)
template <typename R>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 2 : 1)))
void batch_walk_kernel(const BatchTable* table, int* queue, int t_begin, int t_end, int acc_add,
                       const BatchState* __restrict__ st, const double* __restrict__ theta)
{
    extern __shared__ double lds[];
    __shared__ int tile_sh;
    const VMC_CONST_AS BatchTable& T = *(const VMC_CONST_AS BatchTable*)table;
    for (;;) {
        if (threadIdx.x == 0) tile_sh = atomicAdd(queue, 1);
        __syncthreads();
        const int tile = __builtin_amdgcn_readfirstlane(tile_sh);
        __syncthreads();
        if (tile >= T.total_tiles) break;
        int s = 0;
        while (s + 1 < T.nsys && tile >= T.sys[s + 1].tile0) s++;
        const BatchSystem& B = *(const BatchSystem*)&T.sys[s];
        WalkDyn D;
        D.alpha = theta[2 * s];
        D.gamma = theta[2 * s + 1];
        D.counter0 = st->counter0;
        D.t_begin = t_begin;
        D.t_end = t_end;
        D.acc_add = acc_add;
        const int S = B.P.steps / st->batches;                 // ref sampler.cpp:104
        const int kb = st->iteration % st->batches;
        D.window_lo = S * kb;
        D.window_hi = S * (kb + 1);
        D.init_qf = st->fresh;
        const int G = 64 / B.K;
        const int nblocks = (B.P.W + G - 1) / G;
        const int b0 = (tile - B.tile0) * B.waves;
        const int b1 = b0 + B.waves < nblocks ? b0 + B.waves : nblocks;
        for (int b = b0; b < b1; b++) {
            switch (B.shape) {
#define X(id, np, k)                                                                          \
    case id:                                                                                  \
        walk_block<R, np, k, false, false>(B.P, D, b, lds, lds + (64 / k) * np * np, nullptr); \
        break;
                VMC_BATCH_SHAPES(X)
#undef X
            }
        }
    }
}

// per-system column sums of the per-walker accumulators, walkers in order;
// out[iteration - it0][s][0..kNumWalkerAcc)
__global__ __launch_bounds__(256) void batch_reduce_kernel(const BatchTable* table, const double* acc,
                                                           double* out, const BatchState* st,
                                                           int it0)
{
    const BatchTable& T = *table;
    __shared__ double part[256];
    const int s = blockIdx.x;
    const BatchSystem& B = T.sys[s];
    const double* a0 = acc + (size_t)B.walker0 * kNumWalkerAcc;
    double* o = out + ((size_t)(st->iteration - it0) * T.nsys + s) * kBatchOut;
    for (int a = 0; a < kNumWalkerAcc; a++) {
        double v = 0;
        for (int i = threadIdx.x; i < B.P.W; i += 256) v += a0[(size_t)i * kNumWalkerAcc + a];
        part[threadIdx.x] = v;
        __syncthreads();
        for (int m = 128; m > 0; m >>= 1) {
            if ((int)threadIdx.x < m) part[threadIdx.x] += part[threadIdx.x + m];
            __syncthreads();
        }
        if (threadIdx.x == 0) o[a] = part[0];
        __syncthreads();
    }
}

// the reference's optimiser step on (alpha, gamma) of every system, from the
// iteration's sums: g = 2(<G E> - <E><G>) with <E> = sum/M and <G> = sum/(W S),
// ADAM beta1 = 0.89, beta2 = 0.999, eps = 1e-8, moments starting at 1
// (ref src/Optimization/adam.cpp:12-38); theta -= update (ref src/system.cpp:114)
__global__ void batch_adam_kernel(const BatchTable* table, BatchOpt* opt, double* theta,
                                  double* out, const BatchState* st, int it0)
{
    const BatchTable& T = *table;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= T.nsys) return;
    const BatchSystem& B = T.sys[s];
    BatchOpt& O = opt[s];
    double* o = out + ((size_t)(st->iteration - it0) * T.nsys + s) * kBatchOut;
    const double M = double(B.P.W) * B.P.steps;
    const double Sw = double(B.P.W) * (B.P.steps / st->batches);
    const double E = o[kAccE] / M;
    const double ga = o[kAccGAlpha] / Sw, gae = o[kAccGAlphaE] / Sw;
    const double gg = o[kAccGGamma] / Sw, gge = o[kAccGGammaE] / Sw;
    const double grad[2] = {2.0 * (gae - E * ga), 2.0 * (gge - E * gg)};
    const bool on[2] = {B.P.has_gauss != 0, B.P.has_pade != 0};
    if (st->learning_rate > 0) {
        O.step += 1;
        for (int p = 0; p < 2; p++) {
            if (!on[p]) continue;
            const double b1 = 0.89, b2 = 0.999, eps = 1e-8;
            O.m[p] = b1 * O.m[p] + (1 - b1) * grad[p];
            O.v[p] = b2 * O.v[p] + (1 - b2) * grad[p] * grad[p];
            const double mh = O.m[p] / (1 - pow(b1, O.step)), vh = O.v[p] / (1 - pow(b2, O.step));
            theta[2 * s + p] -= st->learning_rate * mh / (sqrt(vh) + eps);
        }
    }
    o[kBatchAlpha] = theta[2 * s];
    o[kBatchGamma] = theta[2 * s + 1];
}

// the next iteration: random counter and gradient window
__global__ void batch_advance_kernel(BatchState* st, int steps_total)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st->counter0 += (uint64_t)steps_total;
        st->iteration += 1;
        st->fresh = 0;
    }
}

// the work-queue counter, reset by a kernel so that the iteration can be
// captured in a graph without a memset node
__global__ void batch_reset_kernel(int* queue)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *queue = 0;
}

hipError_t launch_batch_walk(int precision, const BatchTable* T, int* queue, int grid, size_t lds,
                             int t_begin, int t_end, int acc_add, const BatchState* st,
                             const double* theta, hipStream_t stream)
{
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, stream, queue);
    if (precision == 32)
        hipLaunchKernelGGL(batch_walk_kernel<float>, dim3(grid), dim3(64), lds, stream, T, queue,
                           t_begin, t_end, acc_add, st, theta);
    else
        hipLaunchKernelGGL(batch_walk_kernel<double>, dim3(grid), dim3(64), lds, stream, T, queue,
                           t_begin, t_end, acc_add, st, theta);
    return hipGetLastError();
}

hipError_t launch_batch_tail(const BatchTable* T, int nsys, BatchOpt* opt, double* theta,
                             const double* acc, double* out, BatchState* st, int it0,
                             int steps_total, hipStream_t stream)
{
    hipLaunchKernelGGL(batch_reduce_kernel, dim3(nsys), dim3(256), 0, stream, T, acc, out, st, it0);
    hipLaunchKernelGGL(batch_adam_kernel, dim3(1), dim3(64), 0, stream, T, opt, theta, out, st, it0);
    hipLaunchKernelGGL(batch_advance_kernel, dim3(1), dim3(64), 0, stream, st, steps_total);
    return hipGetLastError();
}

// the batch shape id of (np, K), or -1
int batch_shape_id(int np, int K)
{
#define X(id, a, b) \
    if (np == a && K == b) return id;
    VMC_BATCH_SHAPES(X)
#undef X
    return -1;
}

// LDS of one batch workgroup: the largest Slater staging over the shapes used
size_t batch_lds_bytes(int np, int K)
{
    return sizeof(double) * ((size_t)(64 / K) * np * np + 5 * (size_t)(64 / K) * np);
}

// occupancy of the batch kernel: resident one-wave workgroups per CU with
// `lds` bytes of dynamic LDS each
int batch_resident_per_cu(int precision, size_t lds)
{
    int n = 0;
    if (precision == 32)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_walk_kernel<float>, 64, lds);
    else
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_walk_kernel<double>, 64, lds);
    return n;
}

}  // namespace vmc

// vmc_batch.hip -- several quantum dots (different N and omega) walked side
// by side, one kernel per system on streams of their own, and the
// device-side pieces of one optimisation iteration that a hipGraph replays:
// per-system reduction, the reference's ADAM update of the Gaussian and
// Pade-Jastrow parameters (ref src/Optimization/adam.cpp:12-38,
// src/Optimization/optimization.cpp:10-16) and the advance of the random
// counter and gradient window.
//
// The walk's device code (walk_block) is shared with vmc_walk.hip: each
// workgroup (one wave) walks one block of 64/K walkers of its system.
#define VMC_WALK_PRECISION 99   // the walker templates only
#include "vmc_walk.hip"
#undef VMC_WALK_PRECISION

namespace vmc {

// the (np, K) shapes the batch kernel carries (Slater x Gaussian x Pade only)
#define VMC_BATCH_SHAPES(X)                                                                  \
    X(0, 1, 1) X(1, 1, 2) X(2, 3, 4) X(3, 3, 8) X(4, 6, 8) X(5, 6, 16) X(6, 10, 16)           \
    X(7, 10, 32) X(8, 15, 16) X(9, 15, 32) X(10, 21, 32) X(11, 28, 32)

// one kernel per system, all of a batch launched side by side: the system's
// parameters arrive as a kernel argument and each workgroup walks one block
// of walkers, exactly as the single-system kernel (whose registers it keeps),
// while the values that change between iterations (alpha, gamma, random
// counter, gradient window) come from device memory, so that a captured
// iteration can be replayed
// SG: Slater x Gaussian x Pade-Jastrow, Coulomb, exact Langevin, asserted to
// the compiler (the single-system kernel's specialisation of the same name)
template <typename R, int NP, int K, bool SG>
__global__ __launch_bounds__(64) void batch_walk_kernel(const WalkParams P, int user, int t_begin,
                                                        int t_end, int acc_add,
                                                        const BatchState* __restrict__ st,
                                                        const double* __restrict__ theta)
{
    constexpr int G = 64 / K;
    if constexpr (SG) {
        __builtin_assume(P.sampling == kImportanceExact);
        __builtin_assume(P.has_slater == 1 && P.has_gauss == 1 && P.has_pade == 1 && P.coulomb == 1);
    }
    __shared__ double lds[walk_static_lds_doubles<NP, K>()];
    WalkDyn D;
    D.alpha = theta[2 * user];
    D.gamma = theta[2 * user + 1];
    D.counter0 = st->counter0;
    D.t_begin = t_begin;
    D.t_end = t_end;
    D.acc_add = acc_add;
    const int S = P.steps / st->batches;                 // ref sampler.cpp:104
    const int kb = st->iteration % st->batches;
    D.window_lo = S * kb;
    D.window_hi = S * (kb + 1);
    D.init_qf = st->fresh;
    walk_block<R, NP, K, false, false>(P, D, blockIdx.x, lds, lds + walk_t_doubles<NP, K>(), nullptr);
}

// ---- one persistent launch per iteration ----
// A tile is one wave's block of 64/K walkers of one system, walked through
// every refresh segment of the iteration (a rebuild at the start of each).
// Tiles are numbered in table order, costliest system first, and the waves of
// one grid that fills the GPU take them from a device counter: the big
// systems' tiles start first and the small systems' tiles fill the tail, with
// no launch boundary between systems or refresh segments.  Each shape's walk
// is a function of its own, called rather than inlined, so the kernel's
// register count is the largest shape's, not that of all shapes' code merged
// into one body.

template <typename T>
__device__ __forceinline__ T* uniform_ptr(T* p)
{
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (T*)(((uint64_t)hi << 32) | lo);
}

// largest static LDS of the batch shapes (N=56: 2 walkers x 28 x 28 + rows)
constexpr int batch_lds_doubles()
{
    int m = 0;
#define X(id, np, k) m = m > walk_static_lds_doubles<np, k>() ? m : walk_static_lds_doubles<np, k>();
    VMC_BATCH_SHAPES(X)
#undef X
    return m;
}

// the walk of tile `tile` of system s (table order) through all segments; the
// arguments are wave-uniform, which readfirstlane tells the compiler
template <int NP, int K>
__device__ __attribute__((noinline)) void batch_tile(const BatchTable* table, int s, int tile,
                                                     const BatchState* state,
                                                     const double* theta_in, int total, int seg,
                                                     double* lds)
{
    // the table is read through the constant address space at a uniform
    // address, so the walk's parameter loads are scalar loads, as from a kernel
    // argument
    typedef const __attribute__((address_space(4))) BatchTable ConstTable;
    s = __builtin_amdgcn_readfirstlane(s);
    tile = __builtin_amdgcn_readfirstlane(tile);
    constexpr int G = 64 / K;
    for (int t0 = 0; t0 < total; t0 += seg) {
        // the pointers are made opaque once per segment: nothing derived from
        // them is hoisted out of the segment loop and held live across the
        // walk (which spilled registers)
        uint64_t tv = (uint64_t)uniform_ptr(table);
        uint64_t sv = (uint64_t)uniform_ptr(state);
        uint64_t hv = (uint64_t)uniform_ptr(theta_in);
        asm volatile("" : "+s"(tv), "+s"(sv), "+s"(hv));
        const ConstTable* T = (const ConstTable*)tv;
        const BatchState* __restrict__ st = (const BatchState*)sv;
        const double* __restrict__ theta = (const double*)hv;
        const WalkParams& P = (const WalkParams&)T->sys[s].P;
        __builtin_assume(P.sampling == kImportanceExact);
        __builtin_assume(P.has_slater == 1 && P.has_gauss == 1 && P.has_pade == 1 && P.coulomb == 1);
        const int user = T->sys[s].user;
        WalkDyn D;
        D.alpha = theta[2 * user];
        D.gamma = theta[2 * user + 1];
        D.counter0 = st->counter0;
        const int S = P.steps / st->batches;                 // ref sampler.cpp:104
        const int kb = st->iteration % st->batches;
        D.window_lo = S * kb;
        D.window_hi = S * (kb + 1);
        D.init_qf = st->fresh;
        D.t_begin = t0;
        D.t_end = t0 + seg < total ? t0 + seg : total;
        D.acc_add = t0 > 0 ? 1 : 0;
        walk_block<float, NP, K, false, false>(P, D, tile, lds, lds + walk_t_doubles<NP, K>(), nullptr);
    }
}

// FP32, Slater x Gaussian x Pade-Jastrow, Coulomb, exact Langevin; every wave
// takes tiles until the counter passes the last one, then leaves
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2)))
void batch_persistent_kernel(const BatchTable* __restrict__ T, int* queue,
                             const BatchState* st, const double* theta, int total, int seg)
{
    __shared__ double lds[batch_lds_doubles()];
    const int nsys = T->nsys;
    for (;;) {
        int t = 0;
        if (threadIdx.x == 0) t = atomicAdd(queue, 1);
        t = __builtin_amdgcn_readfirstlane(__shfl(t, 0));
        int s = 0, t0 = 0;
        for (; s < nsys; s++) {
            if (t < t0 + T->sys[s].blocks) break;
            t0 += T->sys[s].blocks;
        }
        if (s >= nsys) break;
        switch (T->sys[s].shape) {
#define X(id, np, k)                                                                \
    case id:                                                                        \
        batch_tile<np, k>(T, s, t - t0, st, theta, total, seg, lds);                \
        break;
            VMC_BATCH_SHAPES(X)
#undef X
        default:
            break;
        }
    }
}

__global__ void batch_queue_reset_kernel(int* queue)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *queue = 0;
}

// the number of 256-walker chunks of each system before system s (table
// order) and of system s itself
__device__ __forceinline__ int batch_chunk0(const BatchTable& T, int s)
{
    int c0 = 0;
    for (int i = 0; i < s; i++) c0 += (T.sys[i].P.W + 255) / 256;
    return c0;
}

// per-system column sums of the per-walker accumulators in the two stages of
// the single-system reduction (reduce_acc_partial_kernel, reduce_acc_kernel),
// so that a system's sums come out as a single-system engine's would: stage
// one gives each 256-walker chunk of every system a workgroup (coalesced rows
// into LDS, fixed tree), part[chunk][kNumWalkerAcc]
__global__ __launch_bounds__(256) void batch_reduce_partial_kernel(const BatchTable* table,
                                                                   const double* acc, double* part)
{
    const BatchTable& T = *table;
    __shared__ double tile[256 * kNumWalkerAcc];
    int s = 0, c0 = 0;
    for (; s < T.nsys; s++) {
        const int nc = (T.sys[s].P.W + 255) / 256;
        if ((int)blockIdx.x < c0 + nc) break;
        c0 += nc;
    }
    if (s >= T.nsys) return;   // the launch covers exactly the chunks; uniform per workgroup
    const BatchSystem& B = T.sys[s];
    const int w0 = ((int)blockIdx.x - c0) * 256;
    const int n = (B.P.W - w0 < 256 ? B.P.W - w0 : 256) * kNumWalkerAcc;
    const double* src = acc + ((size_t)B.walker0 + w0) * kNumWalkerAcc;
    for (int i = threadIdx.x; i < 256 * kNumWalkerAcc; i += 256) tile[i] = i < n ? src[i] : 0.0;
    __syncthreads();
    for (int m = 128; m > 0; m >>= 1) {
        if ((int)threadIdx.x < m) {
#pragma unroll
            for (int a = 0; a < kNumWalkerAcc; a++)
                tile[threadIdx.x * kNumWalkerAcc + a] += tile[(threadIdx.x + m) * kNumWalkerAcc + a];
        }
        __syncthreads();
    }
    if ((int)threadIdx.x < kNumWalkerAcc)
        part[(size_t)blockIdx.x * kNumWalkerAcc + threadIdx.x] = tile[threadIdx.x];
}

// stage two: workgroup s sums system s's chunk partials in order with the same
// tree; out[iteration - it0][user][0..kNumWalkerAcc)
__global__ __launch_bounds__(256) void batch_reduce_kernel(const BatchTable* table,
                                                           const double* part, double* out,
                                                           const BatchState* st, int it0)
{
    const BatchTable& T = *table;
    __shared__ double tile[256 * kNumWalkerAcc];
    const int s = blockIdx.x;
    const BatchSystem& B = T.sys[s];
    const int c0 = batch_chunk0(T, s), nb = (B.P.W + 255) / 256;
    const double* p0 = part + (size_t)c0 * kNumWalkerAcc;
    for (int a = 0; a < kNumWalkerAcc; a++) {
        double v = 0;
        for (int i = threadIdx.x; i < nb; i += 256) v += p0[(size_t)i * kNumWalkerAcc + a];
        tile[threadIdx.x * kNumWalkerAcc + a] = v;
    }
    __syncthreads();
    for (int m = 128; m > 0; m >>= 1) {
        if ((int)threadIdx.x < m) {
#pragma unroll
            for (int a = 0; a < kNumWalkerAcc; a++)
                tile[threadIdx.x * kNumWalkerAcc + a] += tile[(threadIdx.x + m) * kNumWalkerAcc + a];
        }
        __syncthreads();
    }
    double* o = out + ((size_t)(st->iteration - it0) * T.nsys + B.user) * kBatchOut;
    if ((int)threadIdx.x < kNumWalkerAcc) o[threadIdx.x] = tile[threadIdx.x];
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
    const int u = B.user;
    BatchOpt& O = opt[u];
    double* o = out + ((size_t)(st->iteration - it0) * T.nsys + u) * kBatchOut;
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
            theta[2 * u + p] -= st->learning_rate * mh / (sqrt(vh) + eps);
        }
    }
    o[kBatchAlpha] = theta[2 * u];
    o[kBatchGamma] = theta[2 * u + 1];
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


hipError_t launch_batch_walk(int precision, int shape, const WalkParams& P, int user, int t_begin,
                             int t_end, int acc_add, const BatchState* st, const double* theta,
                             hipStream_t stream)
{
    const bool sg = P.sampling == kImportanceExact && P.has_slater == 1 && P.has_gauss == 1 &&
                    P.has_pade == 1 && P.coulomb == 1;
    switch (shape) {
#define X(id, np, k)                                                                               \
    case id: {                                                                                     \
        const int blocks = (P.W + 64 / k - 1) / (64 / k);                                          \
        if (precision == 32 && sg)                                                                 \
            hipLaunchKernelGGL((batch_walk_kernel<float, np, k, true>), dim3(blocks), dim3(64), 0, \
                               stream, P, user, t_begin, t_end, acc_add, st, theta);              \
        else if (precision == 32)                                                                  \
            hipLaunchKernelGGL((batch_walk_kernel<float, np, k, false>), dim3(blocks), dim3(64), 0,\
                               stream, P, user, t_begin, t_end, acc_add, st, theta);              \
        else                                                                                       \
            hipLaunchKernelGGL((batch_walk_kernel<double, np, k, false>), dim3(blocks), dim3(64),  \
                               0, stream, P, user, t_begin, t_end, acc_add, st, theta);           \
        break;                                                                                     \
    }
        VMC_BATCH_SHAPES(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// one iteration's walk as one persistent launch (the counter reset first, by a
// kernel, so that the pair can be captured in a graph); grid: resident waves
hipError_t launch_batch_persistent(const BatchTable* T, int* queue, const BatchState* st,
                                   const double* theta, int total, int seg, int grid,
                                   hipStream_t stream)
{
    hipLaunchKernelGGL(batch_queue_reset_kernel, dim3(1), dim3(64), 0, stream, queue);
    hipLaunchKernelGGL(batch_persistent_kernel, dim3(grid), dim3(64), 0, stream, T, queue, st,
                       theta, total, seg);
    return hipGetLastError();
}

// workgroups of the persistent kernel resident on the whole GPU at once
hipError_t batch_persistent_grid(int* grid)
{
    int dev = 0, per_cu = 0;
    hipDeviceProp_t prop;
    hipError_t err;
    if ((err = hipGetDevice(&dev)) != hipSuccess ||
        (err = hipGetDeviceProperties(&prop, dev)) != hipSuccess ||
        (err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_persistent_kernel, 64,
                                                             0)) != hipSuccess)
        return err;
    *grid = (per_cu < 1 ? 1 : per_cu) * prop.multiProcessorCount;
    return hipSuccess;
}

// part: [chunks][kNumWalkerAcc], chunks = the systems' 256-walker chunks in all
hipError_t launch_batch_tail(const BatchTable* T, int nsys, int chunks, BatchOpt* opt,
                             double* theta, const double* acc, double* part, double* out,
                             BatchState* st, int it0, int steps_total, hipStream_t stream)
{
    hipLaunchKernelGGL(batch_reduce_partial_kernel, dim3(chunks), dim3(256), 0, stream, T, acc,
                       part);
    hipLaunchKernelGGL(batch_reduce_kernel, dim3(nsys), dim3(256), 0, stream, T, part, out, st, it0);
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


}  // namespace vmc

// vmc_engine.cpp -- host side of the C ABI in include/vmc.h.
//
// Owns the device buffers of one engine (walker positions, stored drifts,
// per-walker accumulators) and one HIP stream, validates configurations, maps
// the reference's parameter matrix onto kernel arguments and splits a sampling
// iteration into launches of the walk kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "vmc.h"
#include "vmc_device.h"

static_assert(vmc::kStaticMaxParams == VMC_STATIC_MAX_PARAMS &&
                  vmc::kStaticMaxIterations == VMC_STATIC_MAX_ITERATIONS,
              "the header's limits of vmc_optimize_static are the scratch array's");

namespace vmc {
hipError_t launch_walk(int precision, const KernelShape& sh, const WalkParams& P,
                       hipStream_t stream);
hipError_t launch_eval(int precision, const KernelShape& sh, const WalkParams& P,
                       const double* pos, const int* moves, const double* newv, const EvalOut& O,
                       hipStream_t stream);
// out: [1 + kRowChunks][width] (the sums, then partial rows)
hipError_t launch_reduce_rows(const double* rows, int nrows, int width, double* out,
                              hipStream_t stream);
hipError_t launch_opt_tail(const OptLayout& L, const double* acc, const double* rbm, double* theta,
                           double* m, double* v, OptState* st, double* ag, double* rbm_a,
                           double* rbm_b, double* rbm_W, double* hist, int it0, int steps_total,
                           hipStream_t stream);
hipError_t launch_lp_draws(const LpTrainParams& T, uint64_t iteration, int* electrons,
                           double* normals, hipStream_t stream);
hipError_t launch_lp_propose(const LpTrainParams& T, hipStream_t stream);
hipError_t launch_lp_backward(const LpTrainParams& T, hipStream_t stream);
hipError_t launch_lp_wgrad(const LpTrainParams& T, double* rows, hipStream_t stream);
hipError_t launch_lp_adam(float* theta, uint16_t* bf16, float* m, float* v, const double* grad,
                          double inv_b, double lr, int step, hipStream_t stream);
int lp_wgrad_rows(int B);
int lp_act_floats();
int lp_delta_floats();
hipError_t launch_batch_walk(int precision, int shape, const WalkParams& P, int user, int t_begin,
                             int t_end, int acc_add, const BatchState* st, const double* theta,
                             hipStream_t stream);
hipError_t launch_batch_tail(const BatchTable* T, int nsys, int chunks, BatchOpt* opt,
                             double* theta, const double* acc, double* part, double* out,
                             BatchState* st, int it0, int steps_total, hipStream_t stream);
int batch_shape_id(int np, int K);
hipError_t launch_batch_persistent(const BatchTable* T, int* queue, const BatchState* st,
                                   const double* theta, int total, int seg, int grid,
                                   hipStream_t stream);
hipError_t batch_persistent_grid(int* grid);
// out: [1 + ceil(W/256)][kNumWalkerAcc] (the sums, then block partials)
hipError_t launch_reduce(const double* acc, int W, double* out, hipStream_t stream);
hipError_t launch_init_positions(double* pos, int W, int nd, double omega, uint64_t seed,
                                 hipStream_t stream);
hipError_t launch_block_means(const float* rec, int walkers, int steps, int block, double* means,
                              hipStream_t stream);
hipError_t launch_fill(double* p, int n, double value, hipStream_t stream);
hipError_t static_scratch(double** out);
hipError_t launch_draws(int precision, uint64_t seed, uint32_t walker, uint64_t counter0, int nd,
                        int n, int* k, double* z, double* um, double* ua, hipStream_t stream);
}  // namespace vmc

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(VMC_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// orbitals per spin of the closed shells N = 2, 6, ..., 110 (the last three,
// N = 72, 90, 110, one walker per wave)
const int kNP[] = {1, 3, 6, 10, 15, 21, 28, 36, 45, 55};
constexpr int kRowChunks = 64;   // partial rows of launch_reduce_rows
const int kLanes[] = {1, 4, 8, 16, 16, 32, 32, 64, 64, 64};   // lanes per walker for each kNP entry
constexpr int kMaxParticles = 110;
constexpr int kLearnedMaxOrbitals = 28;   // the learned proposal: N <= 56
constexpr int kBatchMaxOrbitals = 28;     // batches: N <= 56

// RBM instantiations: (np, K) pairs compiled with the RBM elements
const int kRbmShapes[][2] = {{1, 1}, {1, 2}, {1, 4}, {1, 8}, {3, 4}, {3, 8}, {6, 8}, {10, 16}};
constexpr int kRbmMaxDof = 32, kRbmMaxHidden = 32;

// closed-shell 2-D states per spin: 2 * C(c+2, 2) == N (ref src/Basis/basis.cpp:106-123)
int closed_shell_orbitals(int N)
{
    for (int c = 0; c < 16; c++) {
        int b = (c + 1) * (c + 2) / 2;
        if (2 * b == N) return b;
        if (2 * b > N) return -1;
    }
    return -1;
}

bool has(const vmc_config* c, int kind)
{
    for (int e = 0; e < c->num_elements; e++)
        if (c->kinds[e] == kind) return true;
    return false;
}

int element_row(const vmc_config* c, int kind)
{
    for (int e = 0; e < c->num_elements; e++)
        if (c->kinds[e] == kind) return e;
    return -1;
}

}  // namespace

struct vmc_engine {
    vmc_config cfg;
    int np = 0;             // particles per ownership slot (Slater orbitals per spin)
    vmc::KernelShape shape{};
    int rbm_mf = 0;                // RBM-product contractions: 0 lanes, 1 FP32 MFMA, 2 BF16 MFMA
    int nd = 0;
    int maxpar = 0;
    int row_gauss = -1, row_pade = -1, row_rbmg = -1, row_rbmp = -1;
    double* d_rbm_a = nullptr;     // [nd]
    double* d_rbm_b = nullptr;     // [H]
    double* d_rbm_W = nullptr;     // [nd][H]
    double* d_rbm_acc = nullptr;   // [blocks][row width]
    double* d_rbm_out = nullptr;   // [row width]
    int rbm_blocks = 0, rbm_width = 0;
    // learned proposal
    bool lp_set = false;
    vmc_lp_hyper lp_hp{0.2, 1.0, 2.0};
    float* d_lp_theta = nullptr;   // FP32 master weights
    uint16_t* d_lp_bf16 = nullptr; // bfloat16 copy used by the walk
    float* d_lp_m = nullptr;       // ADAM moments
    float* d_lp_v = nullptr;
    int lp_step = 0;
    uint64_t lp_iter = 0;
    // training buffers for one sample per walker
    int lp_B = 0;
    int* d_lp_el = nullptr;
    int* d_lp_nnb = nullptr;
    double *d_lp_nrm = nullptr, *d_lp_xnew = nullptr, *d_lp_o0 = nullptr, *d_lp_o1 = nullptr;
    double *d_lp_stats = nullptr, *d_lp_rows = nullptr, *d_lp_grad = nullptr;
    float *d_lp_act = nullptr, *d_lp_delta = nullptr;

    void free_lp_train()
    {
        for (void* b : {(void*)d_lp_el, (void*)d_lp_nnb, (void*)d_lp_nrm, (void*)d_lp_xnew,
                        (void*)d_lp_o0, (void*)d_lp_o1, (void*)d_lp_stats, (void*)d_lp_rows,
                        (void*)d_lp_grad, (void*)d_lp_act, (void*)d_lp_delta})
            if (b) (void)hipFree(b);
        d_lp_el = d_lp_nnb = nullptr;
        d_lp_nrm = d_lp_xnew = d_lp_o0 = d_lp_o1 = d_lp_stats = d_lp_rows = d_lp_grad = nullptr;
        d_lp_act = d_lp_delta = nullptr;
        lp_B = 0;
    }
    double alpha = 1.0, gamma = 1.0;
    bool params_set = false;
    bool walkers_set = false;
    bool init_qf = true;
    uint64_t counter = 0;
    uint64_t seed = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double* d_pos = nullptr;
    double* d_qf = nullptr;
    double* d_acc = nullptr;
    double* d_out = nullptr;
    float* d_rec = nullptr;
    size_t rec_cap = 0;
    double* d_means = nullptr;   // block means of the recorded series
    size_t means_cap = 0;
    // electron densities: one allocation, radial [bins] | spatial [bins^2] | pair [bins^2]
    unsigned long long* d_dens = nullptr;
    int dens_flags = 0, dens_bins = 0;
    double dens_max_radius = 0;

    ~vmc_engine()
    {
        if (d_pos) (void)hipFree(d_pos);
        if (d_qf) (void)hipFree(d_qf);
        if (d_acc) (void)hipFree(d_acc);
        if (d_out) (void)hipFree(d_out);
        if (d_rec) (void)hipFree(d_rec);
        if (d_means) (void)hipFree(d_means);
        if (d_dens) (void)hipFree(d_dens);
        for (double* b : {d_rbm_a, d_rbm_b, d_rbm_W, d_rbm_acc, d_rbm_out})
            if (b) (void)hipFree(b);
        for (void* b : {(void*)d_lp_theta, (void*)d_lp_bf16, (void*)d_lp_m, (void*)d_lp_v})
            if (b) (void)hipFree(b);
        free_lp_train();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" {

const char* vmc_last_error(void) { return g_err.c_str(); }

int vmc_num_parameters(const vmc_config* c, int e)
{
    if (!c || e < 0 || e >= c->num_elements) return 0;
    const int ND = c->num_particles * c->num_dimensions;
    switch (c->kinds[e]) {
    case VMC_GAUSSIAN: return 1;
    case VMC_SLATER: return 0;
    case VMC_PADE: return 1;
    case VMC_RBM_GAUSSIAN: return ND;
    case VMC_RBM_PRODUCT: return c->num_hidden * ND + c->num_hidden;
    }
    return 0;
}

int vmc_max_parameters(const vmc_config* c)
{
    int m = 0;
    for (int e = 0; c && e < c->num_elements; e++) {
        const int p = vmc_num_parameters(c, e);
        if (p > m) m = p;
    }
    return m;
}

int vmc_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int vmc_engine_create(const vmc_config* cfg, vmc_engine** out)
{
    if (!cfg || !out) return fail(VMC_E_INVALID, "null argument");
    *out = nullptr;
    const vmc_config& c = *cfg;
    if (c.num_dimensions != 2)
        return fail(VMC_E_INVALID, "the engine supports numDimensions: 2");
    if (c.num_particles < 1 || c.num_particles > kMaxParticles)
        return fail(VMC_E_INVALID, "numParticles must be in 1..110");
    if (c.num_elements < 1 || c.num_elements > 6)
        return fail(VMC_E_INVALID, "between 1 and 6 wave-function elements");
    for (int e = 0; e < c.num_elements; e++) {
        const int k = c.kinds[e];
        if (k != VMC_GAUSSIAN && k != VMC_SLATER && k != VMC_PADE && k != VMC_RBM_GAUSSIAN &&
            k != VMC_RBM_PRODUCT)
            return fail(VMC_E_INVALID, "unknown wave-function element");
        for (int f = 0; f < e; f++)
            if (c.kinds[f] == k) return fail(VMC_E_INVALID, "an element appears twice");
    }
    if (c.sampling < VMC_BRUTE_FORCE || c.sampling > VMC_LEARNED)
        return fail(VMC_E_INVALID, "unknown sampling method");
    if (c.precision != 32 && c.precision != 64)
        return fail(VMC_E_INVALID, "precision must be 32 or 64");
    if (c.num_walkers < 1) return fail(VMC_E_INVALID, "num_walkers must be positive");
    if (!(c.omega > 0)) return fail(VMC_E_INVALID, "omega must be positive");
    if (!(c.step_length > 0)) return fail(VMC_E_INVALID, "stepLength must be positive");
    if (c.refresh_interval < 0) return fail(VMC_E_INVALID, "refresh_interval must be >= 0");
    if (!(c.drift_limit >= 0)) return fail(VMC_E_INVALID, "drift_limit must be >= 0");

    const bool rbm = has(&c, VMC_RBM_GAUSSIAN) || has(&c, VMC_RBM_PRODUCT);
    if (rbm && c.sampling == VMC_LEARNED)
        return fail(VMC_E_INVALID, "the learned proposal is not available with RBM elements");
    if (rbm) {
        if (2 * c.num_particles > kRbmMaxDof)
            return fail(VMC_E_INVALID, "RBM elements: numParticles * numDimensions must be <= 32");
        if (has(&c, VMC_RBM_PRODUCT) && (c.num_hidden < 1 || c.num_hidden > kRbmMaxHidden))
            return fail(VMC_E_INVALID, "RBM elements: numHiddenNodes must be in 1..32");
        if (!(c.sigma > 0)) return fail(VMC_E_INVALID, "RBM elements: sigma must be positive");
    }
    int np;
    if (has(&c, VMC_SLATER)) {
        np = closed_shell_orbitals(c.num_particles);
        if (np < 0)
            return fail(VMC_E_INVALID, "slaterDeterminant needs a closed-shell numParticles "
                                       "(2, 6, 12, 20, 30, 42, 56, 72, 90, 110)");
    } else {
        const int half = (c.num_particles + 1) / 2;
        np = 0;
        for (int v : kNP)
            if (v >= half) { np = v; break; }
    }

    if (c.sampling == VMC_LEARNED && np > kLearnedMaxOrbitals)
        return fail(VMC_E_INVALID, "the learned proposal is available for numParticles <= 56");
    int K = 0;
    for (size_t i = 0; i < sizeof(kNP) / sizeof(kNP[0]); i++)
        if (kNP[i] == np) K = kLanes[i];
    if (rbm) {
        // four hidden units per lane: K >= H/4
        const int need = has(&c, VMC_RBM_PRODUCT) ? (c.num_hidden + 3) / 4 : 1;
        while (K < need) K *= 2;
        bool found = false;
        for (const auto& sh : kRbmShapes) found |= (sh[0] == np && sh[1] == K);
        if (!found) return fail(VMC_E_INVALID, "no RBM kernel for this particle/hidden-unit count");
    }
    if (c.rbm_matrix < VMC_RBM_AUTO || c.rbm_matrix > VMC_RBM_MFMA_BF16)
        return fail(VMC_E_INVALID, "rbm_matrix must be one of the VMC_RBM_* values");
    const bool mf_ok = has(&c, VMC_RBM_PRODUCT) && c.precision == 32 && K >= 4;
    if (c.rbm_matrix >= VMC_RBM_MFMA_F32 && !mf_ok)
        return fail(VMC_E_INVALID, "matrix-core RBM contractions need an FP32 walk with an "
                                   "RBM-product element and >= 4 lanes per walker");
    int rbm_mf = 0;
    if (mf_ok && c.rbm_matrix != VMC_RBM_LANES) rbm_mf = c.rbm_matrix == VMC_RBM_MFMA_BF16 ? 2 : 1;

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(VMC_E_DEVICE, "no HIP device visible");

    vmc_engine* e = new vmc_engine();
    e->cfg = c;
    e->np = np;
    e->rbm_mf = rbm_mf;
    e->shape = vmc::KernelShape{np, K, rbm};
    e->nd = 2 * c.num_particles;
    e->maxpar = vmc_max_parameters(&c);
    e->row_gauss = element_row(&c, VMC_GAUSSIAN);
    e->row_pade = element_row(&c, VMC_PADE);
    e->row_rbmg = element_row(&c, VMC_RBM_GAUSSIAN);
    e->row_rbmp = element_row(&c, VMC_RBM_PRODUCT);
    auto cleanup = [&](int code) {
        delete e;
        return code;
    };
    if (hipGetDevice(&e->device) != hipSuccess) return cleanup(fail(VMC_E_DEVICE, "hipGetDevice"));
    const size_t W = (size_t)c.num_walkers;
    hipError_t err;
    if ((err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking)) != hipSuccess ||
        (err = hipEventCreate(&e->ev0)) != hipSuccess ||
        (err = hipEventCreate(&e->ev1)) != hipSuccess ||
        (err = hipMalloc(&e->d_pos, W * e->nd * sizeof(double))) != hipSuccess ||
        (err = hipMalloc(&e->d_qf, W * e->nd * sizeof(double))) != hipSuccess ||
        (err = hipMalloc(&e->d_acc, W * vmc::kNumWalkerAcc * sizeof(double))) != hipSuccess ||
        (err = hipMalloc(&e->d_out, (size_t)(1 + (c.num_walkers + 255) / 256) *
                                        vmc::kNumWalkerAcc * sizeof(double))) != hipSuccess)
        return cleanup(fail(VMC_E_DEVICE, std::string("allocation: ") + hipGetErrorString(err)));
    if (rbm) {
        const int H = has(&c, VMC_RBM_PRODUCT) ? c.num_hidden : 0, nd = e->nd;
        e->rbm_blocks = (c.num_walkers + 64 / K - 1) / (64 / K);
        e->rbm_width = vmc::rbm_row_width(nd, H);
        if ((err = hipMalloc(&e->d_rbm_a, nd * sizeof(double))) != hipSuccess ||
            (err = hipMalloc(&e->d_rbm_b, (H + 1) * sizeof(double))) != hipSuccess ||
            (err = hipMalloc(&e->d_rbm_W, ((size_t)nd * H + 1) * sizeof(double))) != hipSuccess ||
            (err = hipMalloc(&e->d_rbm_acc, (size_t)e->rbm_blocks * e->rbm_width * sizeof(double))) !=
                hipSuccess ||
            (err = hipMalloc(&e->d_rbm_out, (size_t)(1 + kRowChunks) * e->rbm_width * sizeof(double))) !=
                hipSuccess)
            return cleanup(fail(VMC_E_DEVICE, std::string("allocation: ") + hipGetErrorString(err)));
    }
    *out = e;
    g_err.clear();
    return VMC_OK;
}

void vmc_engine_destroy(vmc_engine* e) { delete e; }

int vmc_set_parameters(vmc_engine* e, const double* params, int rows, int cols)
{
    if (!e || !params) return fail(VMC_E_INVALID, "null argument");
    if (rows != e->cfg.num_elements || cols != e->maxpar)
        return fail(VMC_E_INVALID, "parameter matrix must be num_elements x max_parameters");
    if (e->row_gauss >= 0) e->alpha = params[e->row_gauss * cols + 0];
    if (e->row_pade >= 0) e->gamma = params[e->row_pade * cols + 0];
    const int nd = e->nd, H = e->cfg.num_hidden;
    if (e->row_rbmg >= 0)   // a_k at column k (ref rbmgaussian.cpp:19-22)
        HIP_TRY(hipMemcpy(e->d_rbm_a, params + (size_t)e->row_rbmg * cols, nd * sizeof(double),
                          hipMemcpyHostToDevice));
    if (e->row_rbmp >= 0) {
        // b_h at column h, W[k][h] at column H + k + nd*h (ref rbmproduct.cpp:18-27)
        const double* row = params + (size_t)e->row_rbmp * cols;
        std::vector<double> W((size_t)nd * H);
        for (int k = 0; k < nd; k++)
            for (int h = 0; h < H; h++) W[(size_t)k * H + h] = row[H + k + nd * h];
        HIP_TRY(hipMemcpy(e->d_rbm_b, row, H * sizeof(double), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(e->d_rbm_W, W.data(), W.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    e->params_set = true;
    return VMC_OK;
}

int vmc_init_walkers(vmc_engine* e, uint64_t seed)
{
    if (!e) return fail(VMC_E_INVALID, "null engine");
    HIP_TRY(vmc::launch_init_positions(e->d_pos, e->cfg.num_walkers, e->nd, e->cfg.omega, seed,
                                       e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->seed = seed;
    e->counter = 0;
    e->init_qf = true;
    e->walkers_set = true;
    return VMC_OK;
}

static int check_walkers(vmc_engine* e, const void* p, int walkers, int dof)
{
    if (!e || !p) return fail(VMC_E_INVALID, "null argument");
    if (walkers != e->cfg.num_walkers || dof != e->nd)
        return fail(VMC_E_INVALID, "expected num_walkers x (num_particles*2) values");
    return VMC_OK;
}

int vmc_set_positions(vmc_engine* e, const double* pos, int walkers, int dof)
{
    if (int rc = check_walkers(e, pos, walkers, dof)) return rc;
    HIP_TRY(hipMemcpyAsync(e->d_pos, pos, sizeof(double) * walkers * dof, hipMemcpyHostToDevice,
                           e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->walkers_set = true;
    e->init_qf = true;
    return VMC_OK;
}

int vmc_get_positions(vmc_engine* e, double* pos, int walkers, int dof)
{
    if (int rc = check_walkers(e, pos, walkers, dof)) return rc;
    HIP_TRY(hipMemcpyAsync(pos, e->d_pos, sizeof(double) * walkers * dof, hipMemcpyDeviceToHost,
                           e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return VMC_OK;
}

int vmc_set_quantum_force(vmc_engine* e, const double* qf, int walkers, int dof)
{
    if (int rc = check_walkers(e, qf, walkers, dof)) return rc;
    HIP_TRY(hipMemcpyAsync(e->d_qf, qf, sizeof(double) * walkers * dof, hipMemcpyHostToDevice,
                           e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->init_qf = false;
    return VMC_OK;
}

int vmc_get_quantum_force(vmc_engine* e, double* qf, int walkers, int dof)
{
    if (int rc = check_walkers(e, qf, walkers, dof)) return rc;
    HIP_TRY(hipMemcpyAsync(qf, e->d_qf, sizeof(double) * walkers * dof, hipMemcpyDeviceToHost,
                           e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return VMC_OK;
}

int vmc_set_random_state(vmc_engine* e, uint64_t seed, uint64_t counter)
{
    if (!e) return fail(VMC_E_INVALID, "null engine");
    e->seed = seed;
    e->counter = counter;
    return VMC_OK;
}

int vmc_get_random_state(vmc_engine* e, uint64_t* seed, uint64_t* counter)
{
    if (!e || !seed || !counter) return fail(VMC_E_INVALID, "null argument");
    *seed = e->seed;
    *counter = e->counter;
    return VMC_OK;
}

static vmc::WalkParams base_params(const vmc_engine* e)
{
    const vmc_config& c = e->cfg;
    vmc::WalkParams P{};
    P.N = c.num_particles;
    P.n_up = c.num_particles / 2;
    P.omega = c.omega;
    P.sqrt_omega = std::sqrt(c.omega);
    P.inv_sqrt_omega = 1.0 / std::sqrt(c.omega);
    P.sqrt_dt = std::sqrt(c.step_length);
    P.has_gauss = e->row_gauss >= 0;
    P.has_slater = has(&c, VMC_SLATER);
    P.has_pade = e->row_pade >= 0;
    P.coulomb = c.interaction ? 1 : 0;
    P.alpha = e->alpha;
    P.gamma = e->gamma;
    P.a_same = 1.0 / 3.0;   // 2-D spin-parallel Pade a (ref padejastrow.cpp:143-146)
    P.a_anti = 1.0;
    P.sampling = c.sampling;
    P.dt = c.step_length;
    P.q2_upper = (c.quirks & VMC_Q2_PADE_RATIO_UPPER) ? 1 : 0;
    P.q3_half = (c.quirks & VMC_Q3_PADE_GAMMA_HALF) ? 1 : 0;
    P.drift_a = c.drift_limit;
    P.has_rbmg = e->row_rbmg >= 0;
    P.has_rbmp = e->row_rbmp >= 0;
    P.ND = e->nd;
    P.H = P.has_rbmp ? c.num_hidden : 0;
    P.sorting = c.sorting ? 1 : 0;
    P.sig2_g = c.sigma * c.sigma;                              // ref rbmgaussian.cpp:16-17
    P.isig2_p = 1.0 / ((10.0 * c.sigma) * (10.0 * c.sigma));  // ref rbmproduct.cpp:15-17
    P.q13_psi_ratio = (c.quirks & VMC_Q13_RBMG_PSI_RATIO) ? 1 : 0;
    P.q6_w_unscaled = (c.quirks & VMC_Q6_RBM_W_UNSCALED) ? 1 : 0;
    P.rbm_mf = e->rbm_mf;
    P.rbm_a = e->d_rbm_a;
    P.rbm_b = e->d_rbm_b;
    P.rbm_W = e->d_rbm_W;
    P.rbm_acc = e->d_rbm_acc;
    P.lp_w = e->d_lp_bf16;
    P.lp_sigma0 = e->lp_hp.sigma0;
    P.lp_mu_max = e->lp_hp.mu_max;
    P.lp_ls_max = e->lp_hp.ls_max;
    if (c.sampling == VMC_LEARNED) P.q2_upper = 0;   // the exact |psi'/psi|^2
    P.seed = e->seed;
    P.W = c.num_walkers;
    P.rec_parts = 1;
    if (e->dens_flags) {
        const size_t b = (size_t)e->dens_bins;
        P.dens_flags = e->dens_flags;
        P.dens_bins = e->dens_bins;
        P.dens_step = e->dens_max_radius / e->dens_bins;   // ref sampler.cpp:41
        P.dens_max_radius = e->dens_max_radius;
        P.dens_radial = e->d_dens;
        P.dens_spatial = e->d_dens + b;
        P.dens_pair = e->d_dens + b + b * b;
    }
    return P;
}

static int run_impl(vmc_engine* e, int steps, int equil, int iteration, int batches,
                    vmc_sums* sums, double* grad, double* grad_e, int rec_walkers, float* energies,
                    int rec_parts = 1, int rec_block = 1, double* means = nullptr)
{
    if (!e) return fail(VMC_E_INVALID, "null engine");
    if (!e->params_set) return fail(VMC_E_STATE, "vmc_set_parameters has not been called");
    if (!e->walkers_set) return fail(VMC_E_STATE, "walkers have no positions yet");
    if (steps < 1 || equil < 0 || batches < 1 || iteration < 0)
        return fail(VMC_E_INVALID, "steps >= 1, equilibration >= 0, batches >= 1 required");
    if (batches > steps)   // an empty gradient window would divide by zero samples
        return fail(VMC_E_INVALID, "batches must not exceed steps (each gradient window needs a step)");
    if (e->cfg.sampling == VMC_LEARNED && !e->lp_set)
        return fail(VMC_E_STATE, "the learned proposal has no weights (vmc_lp_set)");
    if (rec_walkers < 0 || rec_walkers > e->cfg.num_walkers)
        return fail(VMC_E_INVALID, "record_walkers out of range");
    const int W = e->cfg.num_walkers;
    if (rec_walkers > 0) {
        const size_t need = (size_t)rec_parts * rec_walkers * steps;
        if (need > e->rec_cap) {
            if (e->d_rec) (void)hipFree(e->d_rec);
            e->d_rec = nullptr;
            e->rec_cap = 0;
            HIP_TRY(hipMalloc(&e->d_rec, need * sizeof(float)));
            e->rec_cap = need;
        }
        const size_t nm = rec_block > 1 ? (size_t)rec_walkers * (steps / rec_block) : 0;
        if (nm > e->means_cap) {
            if (e->d_means) (void)hipFree(e->d_means);
            e->d_means = nullptr;
            e->means_cap = 0;
            HIP_TRY(hipMalloc(&e->d_means, nm * sizeof(double)));
            e->means_cap = nm;
        }
    }
    vmc::WalkParams P = base_params(e);
    P.counter0 = e->counter;
    P.steps = steps;
    P.equil = equil;
    const int S = steps / batches;                 // ref sampler.cpp:104
    const int kb = iteration % batches;
    P.window_lo = S * kb;
    P.window_hi = S * (kb + 1);
    P.pos = e->d_pos;
    P.qf = (e->cfg.sampling == VMC_IMPORTANCE) ? e->d_qf : nullptr;
    P.acc = e->d_acc;
    P.rec = rec_walkers > 0 ? e->d_rec : nullptr;
    P.rec_parts = rec_parts;
    P.rec_walkers = rec_walkers;
    P.init_qf = e->init_qf ? 1 : 0;
    const int total = equil + steps;
    const int seg = (e->cfg.refresh_interval > 0) ? e->cfg.refresh_interval : total;
    HIP_TRY(hipEventRecord(e->ev0, e->stream));
    for (int t0 = 0; t0 < total; t0 += seg) {
        P.t_begin = t0;
        P.t_end = (t0 + seg < total) ? t0 + seg : total;
        P.acc_add = t0 > 0 ? 1 : 0;
        HIP_TRY(vmc::launch_walk(e->cfg.precision, e->shape, P, e->stream));
    }
    HIP_TRY(hipEventRecord(e->ev1, e->stream));
    HIP_TRY(vmc::launch_reduce(e->d_acc, W, e->d_out, e->stream));
    double out[vmc::kNumWalkerAcc];
    HIP_TRY(hipMemcpyAsync(out, e->d_out, sizeof(out), hipMemcpyDeviceToHost, e->stream));
    std::vector<double> rbm_out(e->rbm_width);
    if (e->shape.rbm) {
        HIP_TRY(vmc::launch_reduce_rows(e->d_rbm_acc, e->rbm_blocks, e->rbm_width, e->d_rbm_out,
                                        e->stream));
        HIP_TRY(hipMemcpyAsync(rbm_out.data(), e->d_rbm_out, sizeof(double) * e->rbm_width,
                               hipMemcpyDeviceToHost, e->stream));
    }
    if (rec_walkers > 0 && rec_block > 1) {
        HIP_TRY(vmc::launch_block_means(e->d_rec, rec_walkers, steps, rec_block, e->d_means,
                                        e->stream));
        HIP_TRY(hipMemcpyAsync(means, e->d_means,
                               sizeof(double) * (size_t)rec_walkers * (steps / rec_block),
                               hipMemcpyDeviceToHost, e->stream));
    } else if (rec_walkers > 0 && energies)
        HIP_TRY(hipMemcpyAsync(energies, e->d_rec,
                               sizeof(float) * (size_t)rec_parts * rec_walkers * steps,
                               hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    e->counter += (uint64_t)total;
    e->init_qf = false;
    if (sums) {
        sums->energy = out[vmc::kAccE];
        sums->energy2 = out[vmc::kAccE2];
        sums->kinetic = out[vmc::kAccT];
        sums->external = out[vmc::kAccVext];
        sums->interaction = out[vmc::kAccVint];
        sums->kinetic2 = out[vmc::kAccT2];
        sums->external2 = out[vmc::kAccVext2];
        sums->interaction2 = out[vmc::kAccVint2];
        sums->accepted = out[vmc::kAccAccept];
        sums->samples = (double)W * steps;
        sums->gradient_samples = (double)W * S;
        sums->kernel_ms = ms;
        sums->nonfinite = out[vmc::kAccBad];
    }
    const int E = e->cfg.num_elements, mp = e->maxpar;
    for (double* g : {grad, grad_e}) {
        if (!g) continue;
        for (int i = 0; i < E * mp; i++) g[i] = 0.0;
    }
    if (mp > 0) {
        if (e->row_gauss >= 0) {
            if (grad) grad[e->row_gauss * mp] = out[vmc::kAccGAlpha];
            if (grad_e) grad_e[e->row_gauss * mp] = out[vmc::kAccGAlphaE];
        }
        if (e->row_pade >= 0) {
            if (grad) grad[e->row_pade * mp] = out[vmc::kAccGGamma];
            if (grad_e) grad_e[e->row_pade * mp] = out[vmc::kAccGGammaE];
        }
        // RBM rows: the block-row layout is [a (nd) | b (H) | W (nd*H)] for the sums
        // of dlnpsi/dtheta, then the same for E_L * dlnpsi/dtheta
        const int nd = e->nd, H = e->shape.rbm ? base_params(e).H : 0;
        const int half = nd + H + nd * H;
        for (int hf = 0; hf < 2; hf++) {
            double* g = hf ? grad_e : grad;
            if (!g) continue;
            if (e->row_rbmg >= 0)
                for (int k = 0; k < nd; k++) g[e->row_rbmg * mp + k] = rbm_out[hf * half + k];
            if (e->row_rbmp >= 0)
                for (int c2 = 0; c2 < H + nd * H; c2++)
                    g[e->row_rbmp * mp + c2] = rbm_out[hf * half + nd + c2];
        }
    }
    return VMC_OK;
}

// One implementation of the device optimisation loop for both entry points:
// vmc_optimize allocates the loop's device state (optimiser moments, state,
// history staging) per call; vmc_optimize_static (use_static) takes it from
// the per-device array of vmc_optim.hip and allocates nothing.
static int optimize_impl(vmc_engine* e, double* params, int iterations, int steps,
                         int equilibration, const vmc_optimizer_config* opt, int use_graph,
                         double* history, double* kernel_ms, bool use_static)
{
    if (!e || !params || !opt) return fail(VMC_E_INVALID, "null argument");
    if (!e->walkers_set) return fail(VMC_E_STATE, "walkers have no positions yet");
    if (e->cfg.sampling == VMC_LEARNED && !e->lp_set)
        return fail(VMC_E_STATE, "the learned proposal has no weights (vmc_lp_set)");
    if (iterations < 1 || steps < 1 || equilibration < 0 || opt->batches < 1 ||
        opt->batches > steps)
        return fail(VMC_E_INVALID, "iterations, steps >= 1; equilibration >= 0; 1 <= batches <= steps");
    if (opt->kind < VMC_OPT_ADAM || opt->kind > VMC_OPT_GD || !(opt->learning_rate >= 0))
        return fail(VMC_E_INVALID, "optimizer kind VMC_OPT_ADAM/SGD/GD, learning_rate >= 0");
    const int E = e->cfg.num_elements, mp = e->maxpar, n = E * mp;
    if (use_static && (n > vmc::kStaticMaxParams || iterations > vmc::kStaticMaxIterations))
        return fail(VMC_E_INVALID,
                    "vmc_optimize_static: at most " + std::to_string(vmc::kStaticMaxParams) +
                        " parameter-matrix entries and " +
                        std::to_string(vmc::kStaticMaxIterations) + " iterations per call (got " +
                        std::to_string(n) + " and " + std::to_string(iterations) + ")");
    if (int rc = vmc_set_parameters(e, params, E, mp)) return rc;
    vmc::OptLayout L{};
    L.rows = E;
    L.mp = mp;
    L.nd = e->nd;
    L.H = e->row_rbmp >= 0 ? e->cfg.num_hidden : 0;
    L.row_gauss = e->row_gauss;
    L.row_pade = e->row_pade;
    L.row_rbmg = e->row_rbmg;
    L.row_rbmp = e->row_rbmp;
    L.kind = opt->kind == VMC_OPT_ADAM ? 0 : 1;
    L.lr = opt->learning_rate;
    L.gamma = opt->gamma;
    L.monotonic = opt->monotonic;
    const int W = e->cfg.num_walkers, total = equilibration + steps;
    L.M = double(W) * steps;
    L.Sw = double(W) * (steps / opt->batches);   // ref sampler.cpp:104, 216-217

    double *d_theta = nullptr, *d_m = nullptr, *d_v = nullptr, *d_ag = nullptr, *d_hist = nullptr;
    vmc::OptState* d_st = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    auto done = [&](int code) {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (!use_static)
            for (void* p : {(void*)d_theta, (void*)d_m, (void*)d_v, (void*)d_ag, (void*)d_hist,
                            (void*)d_st})
                if (p) (void)hipFree(p);
        return code;
    };
    const double ag[2] = {e->alpha, e->gamma};
    const vmc::OptState st0{e->counter, 0, opt->batches, e->init_qf ? 1 : 0, 0};
    hipError_t err;
    if (use_static) {
        double* s = nullptr;
        if ((err = hipSetDevice(e->device)) != hipSuccess || (err = vmc::static_scratch(&s)) != hipSuccess)
            return fail(VMC_E_DEVICE, std::string("vmc_optimize_static scratch: ") + hipGetErrorString(err));
        d_theta = s + vmc::kStaticTheta;
        d_m = s + vmc::kStaticM;
        d_v = s + vmc::kStaticV;
        d_ag = s + vmc::kStaticAG;
        d_st = reinterpret_cast<vmc::OptState*>(s + vmc::kStaticState);
        d_hist = s + vmc::kStaticHist;
        // the array keeps the previous call's contents: every value the loop
        // reads is written here first (the history rows are written by the
        // iterations before they are copied out)
        if ((n > 0 && (err = hipMemcpyAsync(d_theta, params, sizeof(double) * n, hipMemcpyHostToDevice,
                                            e->stream)) != hipSuccess) ||
            (err = vmc::launch_fill(d_m, n, 1.0, e->stream)) != hipSuccess ||   // ref adam.cpp:17-19
            (err = vmc::launch_fill(d_v, n, 1.0, e->stream)) != hipSuccess ||
            (err = hipMemcpyAsync(d_ag, ag, sizeof(ag), hipMemcpyHostToDevice, e->stream)) != hipSuccess ||
            (err = hipMemcpyAsync(d_st, &st0, sizeof(st0), hipMemcpyHostToDevice, e->stream)) != hipSuccess ||
            (err = hipStreamSynchronize(e->stream)) != hipSuccess)
            return done(fail(VMC_E_DEVICE, std::string("optimize set-up: ") + hipGetErrorString(err)));
    } else {
        const std::vector<double> ones(std::max(n, 1), 1.0);   // ref adam.cpp:17-19, sgd.cpp:20
        if ((err = hipMalloc(&d_theta, sizeof(double) * std::max(n, 1))) != hipSuccess ||
            (err = hipMalloc(&d_m, sizeof(double) * std::max(n, 1))) != hipSuccess ||
            (err = hipMalloc(&d_v, sizeof(double) * std::max(n, 1))) != hipSuccess ||
            (err = hipMalloc(&d_ag, sizeof(ag))) != hipSuccess ||
            (err = hipMalloc(&d_st, sizeof(st0))) != hipSuccess ||
            (err = hipMalloc(&d_hist, sizeof(double) * (size_t)iterations * vmc::kOptOut)) != hipSuccess ||
            (n > 0 && (err = hipMemcpy(d_theta, params, sizeof(double) * n, hipMemcpyHostToDevice)) != hipSuccess) ||
            (err = hipMemcpy(d_m, ones.data(), sizeof(double) * ones.size(), hipMemcpyHostToDevice)) != hipSuccess ||
            (err = hipMemcpy(d_v, ones.data(), sizeof(double) * ones.size(), hipMemcpyHostToDevice)) != hipSuccess ||
            (err = hipMemcpy(d_ag, ag, sizeof(ag), hipMemcpyHostToDevice)) != hipSuccess ||
            (err = hipMemcpy(d_st, &st0, sizeof(st0), hipMemcpyHostToDevice)) != hipSuccess)
            return done(fail(VMC_E_DEVICE, std::string("optimize set-up: ") + hipGetErrorString(err)));
    }

    vmc::WalkParams P = base_params(e);
    P.steps = steps;
    P.equil = equilibration;
    P.pos = e->d_pos;
    P.qf = (e->cfg.sampling == VMC_IMPORTANCE) ? e->d_qf : nullptr;
    P.acc = e->d_acc;
    P.rec = nullptr;
    P.dyn_ag = d_ag;
    P.dyn_st = d_st;
    const int seg = (e->cfg.refresh_interval > 0) ? e->cfg.refresh_interval : total;
    auto enqueue = [&]() -> hipError_t {
        hipError_t r;
        for (int t0 = 0; t0 < total; t0 += seg) {
            P.t_begin = t0;
            P.t_end = (t0 + seg < total) ? t0 + seg : total;
            P.acc_add = t0 > 0 ? 1 : 0;
            if ((r = vmc::launch_walk(e->cfg.precision, e->shape, P, e->stream)) != hipSuccess) return r;
        }
        if ((r = vmc::launch_reduce(e->d_acc, W, e->d_out, e->stream)) != hipSuccess) return r;
        if (e->shape.rbm &&
            (r = vmc::launch_reduce_rows(e->d_rbm_acc, e->rbm_blocks, e->rbm_width, e->d_rbm_out,
                                         e->stream)) != hipSuccess)
            return r;
        return vmc::launch_opt_tail(L, e->d_out, e->d_rbm_out, d_theta, d_m, d_v, d_st, d_ag,
                                    e->d_rbm_a, e->d_rbm_b, e->d_rbm_W, d_hist, 0, total, e->stream);
    };
    if (use_graph) {
        if ((err = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
            return done(fail(VMC_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(err)));
        const hipError_t e1 = enqueue();
        const hipError_t e2 = hipStreamEndCapture(e->stream, &graph);
        if (e1 != hipSuccess || e2 != hipSuccess)
            return done(fail(VMC_E_DEVICE, std::string("graph capture: ") +
                                               hipGetErrorString(e1 != hipSuccess ? e1 : e2)));
        if ((err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0)) != hipSuccess)
            return done(fail(VMC_E_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(err)));
    }
    if ((err = hipEventRecord(e->ev0, e->stream)) != hipSuccess)
        return done(fail(VMC_E_DEVICE, std::string("optimize: ") + hipGetErrorString(err)));
    for (int it = 0; it < iterations; it++) {
        err = use_graph ? hipGraphLaunch(exec, e->stream) : enqueue();
        if (err != hipSuccess) return done(fail(VMC_E_DEVICE, std::string("optimize: ") + hipGetErrorString(err)));
    }
    double agf[2];
    float ms = 0;
    if ((err = hipEventRecord(e->ev1, e->stream)) != hipSuccess ||
        (n > 0 && (err = hipMemcpyAsync(params, d_theta, sizeof(double) * n, hipMemcpyDeviceToHost,
                                        e->stream)) != hipSuccess) ||
        (err = hipMemcpyAsync(agf, d_ag, sizeof(agf), hipMemcpyDeviceToHost, e->stream)) != hipSuccess ||
        (history && (err = hipMemcpyAsync(history, d_hist,
                                          sizeof(double) * (size_t)iterations * vmc::kOptOut,
                                          hipMemcpyDeviceToHost, e->stream)) != hipSuccess) ||
        (err = hipStreamSynchronize(e->stream)) != hipSuccess ||
        (err = hipEventElapsedTime(&ms, e->ev0, e->ev1)) != hipSuccess)
        return done(fail(VMC_E_DEVICE, std::string("optimize: ") + hipGetErrorString(err)));
    e->alpha = agf[0];
    e->gamma = agf[1];
    e->counter += (uint64_t)iterations * total;
    e->init_qf = false;
    if (kernel_ms) *kernel_ms = ms;
    return done(VMC_OK);
}

int vmc_optimize(vmc_engine* e, double* params, int iterations, int steps, int equilibration,
                 const vmc_optimizer_config* opt, int use_graph, double* history, double* kernel_ms)
{
    return optimize_impl(e, params, iterations, steps, equilibration, opt, use_graph, history,
                         kernel_ms, false);
}

int vmc_optimize_static(vmc_engine* e, double* params, int iterations, int steps,
                        int equilibration, const vmc_optimizer_config* opt, double* history,
                        double* kernel_ms)
{
    return optimize_impl(e, params, iterations, steps, equilibration, opt, 0, history, kernel_ms,
                         true);
}

int vmc_run(vmc_engine* e, int steps, int equilibration, int iteration, int batches,
            vmc_sums* sums, double* grad, double* grad_e)
{
    return run_impl(e, steps, equilibration, iteration, batches, sums, grad, grad_e, 0, nullptr);
}

int vmc_run_recorded(vmc_engine* e, int steps, int equilibration, int iteration, int batches,
                     vmc_sums* sums, double* grad, double* grad_e, int record_walkers,
                     float* energies)
{
    if (record_walkers > 0 && !energies) return fail(VMC_E_INVALID, "null energies buffer");
    return run_impl(e, steps, equilibration, iteration, batches, sums, grad, grad_e,
                    record_walkers, energies);
}

int vmc_run_blocked(vmc_engine* e, int steps, int equilibration, int iteration, int batches,
                    vmc_sums* sums, double* grad, double* grad_e, int record_walkers, int block,
                    double* means)
{
    if (record_walkers < 1 || !means) return fail(VMC_E_INVALID, "record_walkers >= 1 and a means buffer");
    if (block < 2 || steps % block != 0)
        return fail(VMC_E_INVALID, "block >= 2 dividing steps");
    if ((size_t)record_walkers * (size_t)(steps / block) >= (size_t)1 << 31)
        return fail(VMC_E_INVALID, "record_walkers * steps / block must stay below 2^31");
    return run_impl(e, steps, equilibration, iteration, batches, sums, grad, grad_e,
                    record_walkers, nullptr, 1, block, means);
}

int vmc_run_instant(vmc_engine* e, int steps, int equilibration, int iteration, int batches,
                    vmc_sums* sums, double* grad, double* grad_e, int record_walkers, float* out)
{
    if (record_walkers < 1 || !out) return fail(VMC_E_INVALID, "record_walkers >= 1 and an output buffer");
    return run_impl(e, steps, equilibration, iteration, batches, sums, grad, grad_e,
                    record_walkers, out, 4);
}

int vmc_set_density(vmc_engine* e, int flags, int bins, double max_radius)
{
    if (!e) return fail(VMC_E_INVALID, "null engine");
    if (flags & ~(VMC_DENSITY_RADIAL | VMC_DENSITY_SPATIAL | VMC_DENSITY_TWO_BODY))
        return fail(VMC_E_INVALID, "unknown density flags");
    if (flags && (bins < 1 || bins > VMC_DENSITY_MAX_BINS || !(max_radius > 0)))
        return fail(VMC_E_INVALID, "numberOfBins in 1..4096 and maxRadius > 0");
    if (e->d_dens) (void)hipFree(e->d_dens);
    e->d_dens = nullptr;
    e->dens_flags = 0;
    if (!flags) return VMC_OK;
    const size_t n = (size_t)bins + 2 * (size_t)bins * bins;
    HIP_TRY(hipMalloc(&e->d_dens, n * sizeof(unsigned long long)));
    HIP_TRY(hipMemsetAsync(e->d_dens, 0, n * sizeof(unsigned long long), e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->dens_flags = flags;
    e->dens_bins = bins;
    e->dens_max_radius = max_radius;
    return VMC_OK;
}

int vmc_get_density(vmc_engine* e, int64_t* radial, int64_t* spatial, int64_t* two_body)
{
    if (!e) return fail(VMC_E_INVALID, "null engine");
    if (!e->dens_flags) return fail(VMC_E_STATE, "densities are off (vmc_set_density)");
    const size_t b = (size_t)e->dens_bins;
    int64_t* dst[3] = {radial, spatial, two_body};
    const size_t off[3] = {0, b, b + b * b}, len[3] = {b, b * b, b * b};
    for (int i = 0; i < 3; i++)
        if (dst[i])
            HIP_TRY(hipMemcpy(dst[i], e->d_dens + off[i], len[i] * sizeof(int64_t),
                              hipMemcpyDeviceToHost));
    return VMC_OK;
}

static int evaluate_impl(vmc_engine* e, const double* positions, int walkers, const int* moves,
                         const double* new_values, double* out_before, double* out_after,
                         double* ratios, double* grads_before, double* grads_after)
{
    if (!e || !positions || !moves || !new_values || !out_before || !out_after)
        return fail(VMC_E_INVALID, "null argument");
    if (!e->params_set) return fail(VMC_E_STATE, "vmc_set_parameters has not been called");
    if (walkers < 1) return fail(VMC_E_INVALID, "walkers must be positive");
    for (int w = 0; w < walkers; w++)
        if (moves[w] < 0 || moves[w] >= e->nd) return fail(VMC_E_INVALID, "move index out of range");
    const int width = 9 + e->nd;
    const int E = e->cfg.num_elements, gw = E * e->maxpar;
    const bool want_g = grads_before && grads_after && gw > 0;
    std::vector<double*> bufs;
    auto release = [&]() {
        for (double* b : bufs)
            if (b) (void)hipFree(b);
    };
    auto dalloc = [&](size_t n, double** out) {
        *out = nullptr;
        hipError_t err = hipMalloc(out, sizeof(double) * (n > 0 ? n : 1));
        bufs.push_back(*out);
        return err;
    };
    double *d_p, *d_v, *d_o0, *d_o1, *d_r = nullptr, *d_g0 = nullptr, *d_g1 = nullptr, *d_mi;
    hipError_t err;
    vmc::WalkParams P = base_params(e);
    P.W = walkers;
    vmc::EvalOut O{};
    O.num_elements = E;
    O.maxpar = e->maxpar;
    for (int i = 0; i < E; i++) O.rows[i] = e->cfg.kinds[i];
    if ((err = dalloc((size_t)walkers * e->nd, &d_p)) != hipSuccess ||
        (err = dalloc(walkers, &d_v)) != hipSuccess ||
        (err = dalloc(walkers, &d_mi)) != hipSuccess ||
        (err = dalloc((size_t)walkers * width, &d_o0)) != hipSuccess ||
        (err = dalloc((size_t)walkers * width, &d_o1)) != hipSuccess ||
        (ratios && (err = dalloc((size_t)walkers * E, &d_r)) != hipSuccess) ||
        (want_g && (err = dalloc((size_t)walkers * gw, &d_g0)) != hipSuccess) ||
        (want_g && (err = dalloc((size_t)walkers * gw, &d_g1)) != hipSuccess)) {
        release();
        return fail(VMC_E_DEVICE, std::string("evaluate: ") + hipGetErrorString(err));
    }
    int* d_m = reinterpret_cast<int*>(d_mi);
    O.o0 = d_o0;
    O.o1 = d_o1;
    O.ratios = d_r;
    O.g0 = d_g0;
    O.g1 = d_g1;
    if ((err = hipMemcpyAsync(d_p, positions, sizeof(double) * walkers * e->nd,
                              hipMemcpyHostToDevice, e->stream)) != hipSuccess ||
        (err = hipMemcpyAsync(d_v, new_values, sizeof(double) * walkers, hipMemcpyHostToDevice,
                              e->stream)) != hipSuccess ||
        (err = hipMemcpyAsync(d_m, moves, sizeof(int) * walkers, hipMemcpyHostToDevice,
                              e->stream)) != hipSuccess ||
        (want_g && (err = hipMemsetAsync(d_g0, 0, sizeof(double) * walkers * gw, e->stream)) !=
                       hipSuccess) ||
        (want_g && (err = hipMemsetAsync(d_g1, 0, sizeof(double) * walkers * gw, e->stream)) !=
                       hipSuccess) ||
        (err = vmc::launch_eval(e->cfg.precision, e->shape, P, d_p, d_m, d_v, O, e->stream)) !=
            hipSuccess ||
        (err = hipMemcpyAsync(out_before, d_o0, sizeof(double) * walkers * width,
                              hipMemcpyDeviceToHost, e->stream)) != hipSuccess ||
        (err = hipMemcpyAsync(out_after, d_o1, sizeof(double) * walkers * width,
                              hipMemcpyDeviceToHost, e->stream)) != hipSuccess ||
        (ratios && (err = hipMemcpyAsync(ratios, d_r, sizeof(double) * walkers * E,
                                         hipMemcpyDeviceToHost, e->stream)) != hipSuccess) ||
        (want_g && (err = hipMemcpyAsync(grads_before, d_g0, sizeof(double) * walkers * gw,
                                         hipMemcpyDeviceToHost, e->stream)) != hipSuccess) ||
        (want_g && (err = hipMemcpyAsync(grads_after, d_g1, sizeof(double) * walkers * gw,
                                         hipMemcpyDeviceToHost, e->stream)) != hipSuccess) ||
        (err = hipStreamSynchronize(e->stream)) != hipSuccess) {
        release();
        return fail(VMC_E_DEVICE, std::string("evaluate: ") + hipGetErrorString(err));
    }
    release();
    return VMC_OK;
}

int vmc_evaluate(vmc_engine* e, const double* positions, int walkers, const int* moves,
                 const double* new_values, double* out_before, double* out_after)
{
    return evaluate_impl(e, positions, walkers, moves, new_values, out_before, out_after, nullptr,
                         nullptr, nullptr);
}

int vmc_evaluate_elements(vmc_engine* e, const double* positions, int walkers, const int* moves,
                          const double* new_values, double* out_before, double* out_after,
                          double* ratios, double* grads_before, double* grads_after)
{
    if (!ratios || !grads_before || !grads_after) return fail(VMC_E_INVALID, "null argument");
    return evaluate_impl(e, positions, walkers, moves, new_values, out_before, out_after, ratios,
                         grads_before, grads_after);
}

static uint16_t to_bf16(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);   // round to nearest even
}

int vmc_lp_set(vmc_engine* e, const float* theta, int n, const vmc_lp_hyper* hp)
{
    if (!e || !theta) return fail(VMC_E_INVALID, "null argument");
    if (n != vmc::kLpParams) return fail(VMC_E_INVALID, "the network has VMC_LP_NPARAM parameters");
    if (hp) {
        if (!(hp->sigma0 > 0) || !(hp->mu_max > 0) || !(hp->ls_max > 0))
            return fail(VMC_E_INVALID, "vmc_lp_hyper fields must be positive");
        e->lp_hp = *hp;
    }
    const size_t n4 = sizeof(float) * vmc::kLpParams;
    if (!e->d_lp_theta) {
        HIP_TRY(hipMalloc(&e->d_lp_theta, n4));
        HIP_TRY(hipMalloc(&e->d_lp_bf16, sizeof(uint16_t) * vmc::kLpParams));
        HIP_TRY(hipMalloc(&e->d_lp_m, n4));
        HIP_TRY(hipMalloc(&e->d_lp_v, n4));
    }
    std::vector<uint16_t> b(vmc::kLpParams);
    for (int i = 0; i < vmc::kLpParams; i++) b[i] = to_bf16(theta[i]);
    HIP_TRY(hipMemcpy(e->d_lp_theta, theta, n4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->d_lp_bf16, b.data(), sizeof(uint16_t) * vmc::kLpParams,
                      hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(e->d_lp_m, 0, n4));
    HIP_TRY(hipMemset(e->d_lp_v, 0, n4));
    e->lp_step = 0;
    e->lp_set = true;
    return VMC_OK;
}

int vmc_lp_get(vmc_engine* e, float* theta, int n)
{
    if (!e || !theta) return fail(VMC_E_INVALID, "null argument");
    if (n != vmc::kLpParams) return fail(VMC_E_INVALID, "the network has VMC_LP_NPARAM parameters");
    if (!e->lp_set) return fail(VMC_E_STATE, "vmc_lp_set has not been called");
    HIP_TRY(hipMemcpy(theta, e->d_lp_theta, sizeof(float) * n, hipMemcpyDeviceToHost));
    return VMC_OK;
}

// training buffers for B samples
static int lp_buffers(vmc_engine* e, int B)
{
    if (e->lp_B == B) return VMC_OK;
    e->free_lp_train();
    const size_t w9 = (size_t)B * (9 + e->nd);
    HIP_TRY(hipMalloc(&e->d_lp_el, sizeof(int) * B));
    HIP_TRY(hipMalloc(&e->d_lp_nnb, sizeof(int) * B));
    HIP_TRY(hipMalloc(&e->d_lp_nrm, sizeof(double) * 2 * B));
    HIP_TRY(hipMalloc(&e->d_lp_xnew, sizeof(double) * 2 * B));
    HIP_TRY(hipMalloc(&e->d_lp_o0, sizeof(double) * w9));
    HIP_TRY(hipMalloc(&e->d_lp_o1, sizeof(double) * w9));
    HIP_TRY(hipMalloc(&e->d_lp_stats, sizeof(double) * 5 * B));
    HIP_TRY(hipMalloc(&e->d_lp_rows, sizeof(double) * (size_t)vmc::lp_wgrad_rows(B) * vmc::kLpParams));
    HIP_TRY(hipMalloc(&e->d_lp_grad, sizeof(double) * (1 + kRowChunks) * vmc::kLpParams));
    HIP_TRY(hipMalloc(&e->d_lp_act, sizeof(float) * (size_t)B * vmc::lp_act_floats()));
    HIP_TRY(hipMalloc(&e->d_lp_delta, sizeof(float) * (size_t)B * vmc::lp_delta_floats()));
    e->lp_B = B;
    return VMC_OK;
}

// the training pipeline on B configurations at pos (device); electrons and
// normals already in the training buffers; leaves the summed gradient in d_lp_grad
static int lp_pipeline(vmc_engine* e, const double* pos, int B)
{
    vmc::LpTrainParams T{};
    T.B = B;
    T.N = e->cfg.num_particles;
    T.omega = e->cfg.omega;
    T.seed = e->seed;
    T.sigma0 = e->lp_hp.sigma0;
    T.mu_max = e->lp_hp.mu_max;
    T.ls_max = e->lp_hp.ls_max;
    T.theta = e->d_lp_theta;
    T.pos = pos;
    T.electrons = e->d_lp_el;
    T.normals = e->d_lp_nrm;
    T.xnew = e->d_lp_xnew;
    T.nnb = e->d_lp_nnb;
    T.act = e->d_lp_act;
    T.delta = e->d_lp_delta;
    T.eval = e->d_lp_o1;
    T.stats = e->d_lp_stats;
    HIP_TRY(vmc::launch_lp_propose(T, e->stream));
    // |psi(x')/psi(x)|^2 and the drift at x': the walk's own evaluation of the
    // whole-electron move (FP64 or FP32 as the engine), exact Pade ratio
    vmc::WalkParams P = base_params(e);
    P.W = B;
    P.q2_upper = 0;
    vmc::EvalOut O{};
    O.newv2 = e->d_lp_xnew;
    O.o0 = e->d_lp_o0;
    O.o1 = e->d_lp_o1;
    O.num_elements = e->cfg.num_elements;
    O.maxpar = e->maxpar;
    for (int i = 0; i < e->cfg.num_elements; i++) O.rows[i] = e->cfg.kinds[i];
    HIP_TRY(vmc::launch_eval(e->cfg.precision, e->shape, P, pos, e->d_lp_el, e->d_lp_xnew, O,
                             e->stream));
    HIP_TRY(vmc::launch_lp_backward(T, e->stream));
    HIP_TRY(vmc::launch_lp_wgrad(T, e->d_lp_rows, e->stream));
    HIP_TRY(vmc::launch_reduce_rows(e->d_lp_rows, vmc::lp_wgrad_rows(B), vmc::kLpParams,
                                    e->d_lp_grad, e->stream));
    return VMC_OK;
}

static bool lp_usable(const vmc_engine* e) { return !e->shape.rbm; }

int vmc_lp_gradient(vmc_engine* e, const double* positions, int walkers, const int* electrons,
                    const double* normals, double* per_sample, double* grad)
{
    if (!e || !positions || !electrons || !normals || !per_sample || !grad)
        return fail(VMC_E_INVALID, "null argument");
    if (!e->lp_set) return fail(VMC_E_STATE, "vmc_lp_set has not been called");
    if (!e->params_set) return fail(VMC_E_STATE, "vmc_set_parameters has not been called");
    if (!lp_usable(e)) return fail(VMC_E_INVALID, "the learned proposal is not available with RBM elements");
    if (walkers < 1) return fail(VMC_E_INVALID, "walkers must be positive");
    for (int w = 0; w < walkers; w++)
        if (electrons[w] < 0 || electrons[w] >= e->cfg.num_particles)
            return fail(VMC_E_INVALID, "electron index out of range");
    int rc = lp_buffers(e, walkers);
    if (rc != VMC_OK) return rc;
    double* d_pos = nullptr;
    HIP_TRY(hipMalloc(&d_pos, sizeof(double) * walkers * e->nd));
    hipError_t err;
    if ((err = hipMemcpyAsync(d_pos, positions, sizeof(double) * walkers * e->nd,
                              hipMemcpyHostToDevice, e->stream)) != hipSuccess ||
        (err = hipMemcpyAsync(e->d_lp_el, electrons, sizeof(int) * walkers, hipMemcpyHostToDevice,
                              e->stream)) != hipSuccess ||
        (err = hipMemcpyAsync(e->d_lp_nrm, normals, sizeof(double) * 2 * walkers,
                              hipMemcpyHostToDevice, e->stream)) != hipSuccess) {
        (void)hipFree(d_pos);
        return fail(VMC_E_DEVICE, std::string("lp_gradient: ") + hipGetErrorString(err));
    }
    rc = lp_pipeline(e, d_pos, walkers);
    if (rc == VMC_OK) {
        if ((err = hipMemcpyAsync(per_sample, e->d_lp_stats, sizeof(double) * 5 * walkers,
                                  hipMemcpyDeviceToHost, e->stream)) != hipSuccess ||
            (err = hipMemcpyAsync(grad, e->d_lp_grad, sizeof(double) * vmc::kLpParams,
                                  hipMemcpyDeviceToHost, e->stream)) != hipSuccess ||
            (err = hipStreamSynchronize(e->stream)) != hipSuccess)
            rc = fail(VMC_E_DEVICE, std::string("lp_gradient: ") + hipGetErrorString(err));
    }
    (void)hipFree(d_pos);
    return rc;
}

int vmc_lp_train(vmc_engine* e, int iterations, int walk_steps, double learning_rate,
                 double* history)
{
    if (!e) return fail(VMC_E_INVALID, "null engine");
    if (!e->lp_set) return fail(VMC_E_STATE, "vmc_lp_set has not been called");
    if (!e->params_set) return fail(VMC_E_STATE, "vmc_set_parameters has not been called");
    if (!e->walkers_set) return fail(VMC_E_STATE, "walkers have no positions yet");
    if (!lp_usable(e)) return fail(VMC_E_INVALID, "the learned proposal is not available with RBM elements");
    if (iterations < 0 || walk_steps < 0 || !(learning_rate > 0))
        return fail(VMC_E_INVALID, "iterations >= 0, walk_steps >= 0, learning_rate > 0 required");
    const int B = e->cfg.num_walkers;
    int rc = lp_buffers(e, B);
    if (rc != VMC_OK) return rc;
    std::vector<double> st(history ? (size_t)5 * B : 0);
    for (int it = 0; it < iterations; it++) {
        vmc::LpTrainParams T{};
        T.B = B;
        T.N = e->cfg.num_particles;
        T.seed = e->seed;
        HIP_TRY(vmc::launch_lp_draws(T, e->lp_iter++, e->d_lp_el, e->d_lp_nrm, e->stream));
        rc = lp_pipeline(e, e->d_pos, B);
        if (rc != VMC_OK) return rc;
        e->lp_step += 1;
        HIP_TRY(vmc::launch_lp_adam(e->d_lp_theta, e->d_lp_bf16, e->d_lp_m, e->d_lp_v, e->d_lp_grad,
                                    1.0 / B, learning_rate, e->lp_step, e->stream));
        if (history) {
            HIP_TRY(hipMemcpyAsync(st.data(), e->d_lp_stats, sizeof(double) * 5 * B,
                                   hipMemcpyDeviceToHost, e->stream));
            HIP_TRY(hipStreamSynchronize(e->stream));
            double sJ = 0, sA = 0, sD = 0;
            for (int b = 0; b < B; b++) {
                sJ += st[5 * b];
                sA += st[5 * b + 1];
                sD += st[5 * b + 1] > 0 ? st[5 * b] / st[5 * b + 1] : 0.0;
            }
            history[3 * it] = sJ / B;
            history[3 * it + 1] = sA / B;
            history[3 * it + 2] = sD / B;
        }
        if (walk_steps > 0) {
            rc = run_impl(e, walk_steps, 0, 0, 1, nullptr, nullptr, nullptr, 0, nullptr);
            if (rc != VMC_OK) return rc;
        }
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    return VMC_OK;
}

int vmc_draws(int precision, uint64_t seed, uint32_t walker, uint64_t counter0, int n, int dof,
              int* coords, double* normals, double* u_move, double* u_accept)
{
    if (!coords || !normals || !u_move || !u_accept || n < 1 || dof < 1)
        return fail(VMC_E_INVALID, "bad argument");
    if (precision != 32 && precision != 64) return fail(VMC_E_INVALID, "precision must be 32 or 64");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(VMC_E_DEVICE, "no HIP device visible");
    int* d_k = nullptr;
    double* d_f = nullptr;
    hipError_t err;
    if ((err = hipMalloc(&d_k, sizeof(int) * n)) != hipSuccess ||
        (err = hipMalloc(&d_f, sizeof(double) * 3 * n)) != hipSuccess ||
        (err = vmc::launch_draws(precision, seed, walker, counter0, dof, n, d_k, d_f, d_f + n,
                                 d_f + 2 * n, nullptr)) != hipSuccess ||
        (err = hipMemcpy(coords, d_k, sizeof(int) * n, hipMemcpyDeviceToHost)) != hipSuccess ||
        (err = hipMemcpy(normals, d_f, sizeof(double) * n, hipMemcpyDeviceToHost)) != hipSuccess ||
        (err = hipMemcpy(u_move, d_f + n, sizeof(double) * n, hipMemcpyDeviceToHost)) != hipSuccess ||
        (err = hipMemcpy(u_accept, d_f + 2 * n, sizeof(double) * n, hipMemcpyDeviceToHost)) !=
            hipSuccess) {
        if (d_k) (void)hipFree(d_k);
        if (d_f) (void)hipFree(d_f);
        return fail(VMC_E_DEVICE, std::string("draws: ") + hipGetErrorString(err));
    }
    (void)hipFree(d_k);
    (void)hipFree(d_f);
    return VMC_OK;
}


}  // extern "C"

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
struct vmc_batch {
    vmc_batch_config cfg{};
    std::vector<vmc_batch_system> sys;     // caller order
    std::vector<vmc::BatchSystem> hs;      // table order: costliest first
    std::vector<int> dev_of;               // caller index -> table index
    std::vector<double> cost;              // caller order: time per walker-step (relative)
    std::vector<double> theta;             // caller order: alpha, gamma
    int total_walkers = 0;
    uint64_t seed = 0, counter = 0;
    int iteration = 0;
    bool walkers_set = false, fresh = true;
    hipStream_t stream = nullptr;
    hipStream_t side[vmc::kBatchMaxSystems] = {};   // one more stream per further system
    hipEvent_t ev0 = nullptr, ev1 = nullptr, fork = nullptr;
    hipEvent_t join[vmc::kBatchMaxSystems] = {};
    vmc::BatchTable* d_table = nullptr;    // [2]: the run's table, the autotuner's
    vmc::BatchState* d_st = nullptr;
    vmc::BatchOpt* d_opt = nullptr;        // caller order
    double *d_theta = nullptr, *d_pos = nullptr, *d_qf = nullptr, *d_acc = nullptr;
    double* d_part = nullptr;              // [chunks][kNumWalkerAcc]: reduction partials
    int chunks = 0;                        // 256-walker chunks of all systems
    int launch = VMC_BATCH_SIDE_BY_SIDE;   // how an iteration's walk is launched
    int* d_queue = nullptr;                // the persistent launch's tile counter
    int persistent_grid = 0;               // its resident workgroups
    std::vector<size_t> pos_off;           // caller index -> first double of its positions
    size_t pos_total = 0;                  // doubles of all systems' positions

    ~vmc_batch()
    {
        for (void* p : {(void*)d_table, (void*)d_st, (void*)d_opt, (void*)d_theta, (void*)d_pos,
                        (void*)d_qf, (void*)d_acc, (void*)d_part, (void*)d_queue})
            if (p) (void)hipFree(p);
        for (hipEvent_t e : {ev0, ev1, fork})
            if (e) (void)hipEventDestroy(e);
        for (int i = 0; i < vmc::kBatchMaxSystems; i++) {
            if (join[i]) (void)hipEventDestroy(join[i]);
            if (side[i]) (void)hipStreamDestroy(side[i]);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// the launch order for the systems' current shapes: the costliest system
// (walkers x time per walker-step, measured or modelled) first, so that the
// longest kernel starts first and the shorter ones fill in beside it
void batch_plan(vmc_batch* b)
{
    auto work = [&](const vmc::BatchSystem& B) { return B.P.W * b->cost[B.user]; };
    std::stable_sort(b->hs.begin(), b->hs.end(),
                     [&](const vmc::BatchSystem& x, const vmc::BatchSystem& y) { return work(x) > work(y); });
    for (int d = 0; d < (int)b->hs.size(); d++) {
        vmc::BatchSystem& B = b->hs[d];
        b->dev_of[B.user] = d;
        B.blocks = (B.P.W + 64 / B.K - 1) / (64 / B.K);
    }
}

// the persistent launch carries FP32 walks of Slater x Gaussian x Pade-Jastrow
// with Coulomb under exact Langevin (the mixed-size batch's configuration)
bool batch_persistent_ok(const vmc_batch* b)
{
    if (b->cfg.precision != 32 || b->cfg.sampling != VMC_IMPORTANCE_EXACT || !b->cfg.interaction)
        return false;
    const vmc::WalkParams& P = b->hs[0].P;
    return P.has_slater == 1 && P.has_gauss == 1 && P.has_pade == 1;
}

vmc::BatchTable batch_table(const vmc_batch* b)
{
    vmc::BatchTable T{};
    T.nsys = (int)b->hs.size();
    for (int d = 0; d < T.nsys; d++) {
        T.sys[d] = b->hs[d];
    }
    return T;
}

// one iteration's walk: per refresh segment the systems' kernels side by
// side (the main stream forks to one stream per further system and joins
// them again)
int batch_enqueue_walk(vmc_batch* b, int total)
{
    const int seg = b->cfg.refresh_interval > 0 ? b->cfg.refresh_interval : total;
    if (b->launch == VMC_BATCH_PERSISTENT) {
        HIP_TRY(vmc::launch_batch_persistent(b->d_table, b->d_queue, b->d_st, b->d_theta, total, seg,
                                             b->persistent_grid, b->stream));
        return VMC_OK;
    }
    const int n = (int)b->hs.size();
    for (int t0 = 0; t0 < total; t0 += seg) {
        const int t1 = t0 + seg < total ? t0 + seg : total;
        if (n > 1) HIP_TRY(hipEventRecord(b->fork, b->stream));
        for (int d = 0; d < n; d++) {
            const vmc::BatchSystem& B = b->hs[d];
            hipStream_t st = d == 0 ? b->stream : b->side[d - 1];
            if (d > 0) HIP_TRY(hipStreamWaitEvent(st, b->fork, 0));
            HIP_TRY(vmc::launch_batch_walk(b->cfg.precision, B.shape, B.P, B.user, t0, t1,
                                           t0 > 0 ? 1 : 0, b->d_st, b->d_theta, st));
            if (d > 0) HIP_TRY(hipEventRecord(b->join[d - 1], st));
        }
        for (int d = 1; d < n; d++) HIP_TRY(hipStreamWaitEvent(b->stream, b->join[d - 1], 0));
    }
    return VMC_OK;
}

}  // namespace

extern "C" {

int vmc_batch_create(const vmc_batch_config* cfg, vmc_batch** out)
{
    if (!cfg || !out || !cfg->systems) return fail(VMC_E_INVALID, "null argument");
    *out = nullptr;
    const vmc_batch_config& c = *cfg;
    if (c.num_systems < 1 || c.num_systems > vmc::kBatchMaxSystems)
        return fail(VMC_E_INVALID, "1.." + std::to_string(vmc::kBatchMaxSystems) + " systems");
    if (c.num_elements < 1 || c.num_elements > 3) return fail(VMC_E_INVALID, "1..3 elements");
    bool has_slater = false;
    for (int e = 0; e < c.num_elements; e++) {
        const int k = c.kinds[e];
        if (k != VMC_GAUSSIAN && k != VMC_SLATER && k != VMC_PADE)
            return fail(VMC_E_INVALID, "batches take Slater, Gaussian and Pade-Jastrow elements");
        for (int f = 0; f < e; f++)
            if (c.kinds[f] == k) return fail(VMC_E_INVALID, "an element appears twice");
        has_slater |= k == VMC_SLATER;
    }
    if (c.sampling < VMC_BRUTE_FORCE || c.sampling > VMC_IMPORTANCE_EXACT)
        return fail(VMC_E_INVALID, "batches take brute force or Langevin sampling");
    if (c.precision != 32 && c.precision != 64) return fail(VMC_E_INVALID, "precision 32 or 64");
    if (c.refresh_interval < 0 || !(c.drift_limit >= 0))
        return fail(VMC_E_INVALID, "refresh_interval >= 0 and drift_limit >= 0");
    vmc_batch* b = new vmc_batch();
    auto cleanup = [&](int code) {
        delete b;
        return code;
    };
    b->cfg = c;
    b->sys.assign(c.systems, c.systems + c.num_systems);
    b->cfg.systems = nullptr;
    std::vector<vmc::BatchSystem> hs(c.num_systems);
    size_t pos = 0, acc = 0;
    b->pos_off.resize(c.num_systems);
    for (int s = 0; s < c.num_systems; s++) {
        const vmc_batch_system& y = b->sys[s];
        const std::string tag = "system " + std::to_string(s) + ": ";
        if (y.num_walkers < 1 || y.num_walkers > (1 << 24) || !(y.omega > 0) ||
            !(y.step_length > 0))
            return cleanup(fail(VMC_E_INVALID, tag + "num_walkers in 1..2^24, omega and step_length positive"));
        int np = -1;
        if (has_slater) {
            np = closed_shell_orbitals(y.num_particles);
            if (np < 0 || np > kBatchMaxOrbitals)
                return cleanup(fail(VMC_E_INVALID, tag + "slaterDeterminant needs a closed shell "
                                                         "of at most 56 particles in a batch"));
        } else if (y.num_particles >= 1 && y.num_particles <= 56) {
            const int half = (y.num_particles + 1) / 2;
            for (int v : kNP)
                if (v >= half) { np = v; break; }
        }
        if (np < 0) return cleanup(fail(VMC_E_INVALID, tag + "numParticles must be in 1..56"));
        int K = 0;
        for (size_t i = 0; i < sizeof(kNP) / sizeof(kNP[0]); i++)
            if (kNP[i] == np) K = kLanes[i];
        vmc::BatchSystem& B = hs[s];
        B = vmc::BatchSystem{};
        vmc::WalkParams& P = B.P;
        P.N = y.num_particles;
        P.n_up = y.num_particles / 2;
        P.omega = y.omega;
        P.sqrt_omega = std::sqrt(y.omega);
        P.inv_sqrt_omega = 1.0 / std::sqrt(y.omega);
        P.sqrt_dt = std::sqrt(y.step_length);
        for (int e = 0; e < c.num_elements; e++) {
            P.has_gauss |= c.kinds[e] == VMC_GAUSSIAN;
            P.has_slater |= c.kinds[e] == VMC_SLATER;
            P.has_pade |= c.kinds[e] == VMC_PADE;
        }
        P.coulomb = c.interaction ? 1 : 0;
        P.alpha = y.alpha;   // the walk reads the device copy of (alpha, gamma)
        P.gamma = y.gamma;
        P.a_same = 1.0 / 3.0;   // 2-D spin-parallel Pade a (ref padejastrow.cpp:143-146)
        P.a_anti = 1.0;
        P.ND = 2 * y.num_particles;
        P.sampling = c.sampling;
        P.dt = y.step_length;
        P.q2_upper = (c.quirks & VMC_Q2_PADE_RATIO_UPPER) ? 1 : 0;
        P.q3_half = (c.quirks & VMC_Q3_PADE_GAMMA_HALF) ? 1 : 0;
        P.drift_a = c.drift_limit;
        P.W = y.num_walkers;
        B.np = np;
        B.K = K;
        B.shape = vmc::batch_shape_id(np, K);
        B.user = s;
        B.walker0 = (int)acc;
        b->pos_off[s] = pos;
        pos += (size_t)y.num_walkers * 2 * y.num_particles;
        acc += (size_t)y.num_walkers;
    }
    b->total_walkers = (int)acc;
    b->pos_total = pos;
    for (const auto& y : b->sys) b->chunks += (y.num_walkers + 255) / 256;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return cleanup(fail(VMC_E_DEVICE, "no HIP device visible"));
    hipError_t err;
    int dev = 0;
    if ((err = hipGetDevice(&dev)) != hipSuccess ||
        (err = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess ||
        (err = hipEventCreate(&b->ev0)) != hipSuccess || (err = hipEventCreate(&b->ev1)) != hipSuccess ||
        (err = hipMalloc(&b->d_pos, sizeof(double) * pos)) != hipSuccess ||
        (err = hipMalloc(&b->d_qf, sizeof(double) * pos)) != hipSuccess ||
        (err = hipMalloc(&b->d_acc, sizeof(double) * acc * vmc::kNumWalkerAcc)) != hipSuccess ||
        (err = hipMalloc(&b->d_part, sizeof(double) * b->chunks * vmc::kNumWalkerAcc)) != hipSuccess ||
        (err = hipMalloc(&b->d_queue, sizeof(int))) != hipSuccess ||
        (err = hipMalloc(&b->d_theta, sizeof(double) * 2 * c.num_systems)) != hipSuccess ||
        (err = hipMalloc(&b->d_table, 2 * sizeof(vmc::BatchTable))) != hipSuccess ||
        (err = hipMalloc(&b->d_st, sizeof(vmc::BatchState))) != hipSuccess ||
        (err = hipMalloc(&b->d_opt, sizeof(vmc::BatchOpt) * c.num_systems)) != hipSuccess ||
        (err = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming)) != hipSuccess)
        return cleanup(fail(VMC_E_DEVICE, std::string("batch set-up: ") + hipGetErrorString(err)));
    for (int i = 0; i + 1 < c.num_systems; i++)
        if ((err = hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking)) != hipSuccess ||
            (err = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming)) != hipSuccess)
            return cleanup(fail(VMC_E_DEVICE, std::string("batch set-up: ") + hipGetErrorString(err)));
    for (int s = 0; s < c.num_systems; s++) {
        vmc::WalkParams& P = hs[s].P;
        P.pos = b->d_pos + b->pos_off[s];
        P.qf = (c.sampling == VMC_IMPORTANCE) ? b->d_qf + b->pos_off[s] : nullptr;
        P.acc = b->d_acc + (size_t)hs[s].walker0 * vmc::kNumWalkerAcc;
        b->theta.push_back(hs[s].P.alpha);
        b->theta.push_back(hs[s].P.gamma);
    }
    // the cost model until the autotuner measures: time per walker-step ~ N^2
    b->hs = hs;
    b->dev_of.assign(c.num_systems, 0);
    b->cost.resize(c.num_systems);
    for (int s = 0; s < c.num_systems; s++)
        b->cost[s] = double(b->sys[s].num_particles) * b->sys[s].num_particles;
    batch_plan(b);
    // ADAM moments start at one (ref src/Optimization/adam.cpp:17-19)
    std::vector<vmc::BatchOpt> opt(c.num_systems, vmc::BatchOpt{{1.0, 1.0}, {1.0, 1.0}, 0});
    if ((err = hipMemcpy(b->d_opt, opt.data(), sizeof(vmc::BatchOpt) * opt.size(),
                         hipMemcpyHostToDevice)) != hipSuccess ||
        (err = hipMemcpy(b->d_theta, b->theta.data(), sizeof(double) * b->theta.size(),
                         hipMemcpyHostToDevice)) != hipSuccess)
        return cleanup(fail(VMC_E_DEVICE, std::string("batch set-up: ") + hipGetErrorString(err)));
    *out = b;
    g_err.clear();
    return VMC_OK;
}

void vmc_batch_destroy(vmc_batch* b) { delete b; }

int vmc_batch_init_walkers(vmc_batch* b, uint64_t seed)
{
    if (!b) return fail(VMC_E_INVALID, "null batch");
    for (const auto& B : b->hs)
        HIP_TRY(vmc::launch_init_positions(b->d_pos + b->pos_off[B.user], B.P.W, B.P.ND, B.P.omega,
                                           seed + (uint64_t)B.user, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->seed = seed;
    b->counter = 0;
    b->iteration = 0;
    b->walkers_set = true;
    b->fresh = true;
    return VMC_OK;
}

static int batch_check_system(vmc_batch* b, int system, const double* p, int walkers, int dof)
{
    if (!b || !p) return fail(VMC_E_INVALID, "null argument");
    if (system < 0 || system >= (int)b->sys.size()) return fail(VMC_E_INVALID, "no such system");
    const auto& y = b->sys[system];
    if (walkers != y.num_walkers || dof != 2 * y.num_particles)
        return fail(VMC_E_INVALID, "expected the system's num_walkers x (num_particles*2) values");
    return VMC_OK;
}

int vmc_batch_set_positions(vmc_batch* b, int system, const double* positions, int walkers, int dof)
{
    if (int rc = batch_check_system(b, system, positions, walkers, dof)) return rc;
    HIP_TRY(hipMemcpy(b->d_pos + b->pos_off[system], positions, sizeof(double) * walkers * dof,
                      hipMemcpyHostToDevice));
    b->walkers_set = true;
    b->fresh = true;
    return VMC_OK;
}

int vmc_batch_get_positions(vmc_batch* b, int system, double* positions, int walkers, int dof)
{
    if (int rc = batch_check_system(b, system, positions, walkers, dof)) return rc;
    HIP_TRY(hipMemcpy(positions, b->d_pos + b->pos_off[system], sizeof(double) * walkers * dof,
                      hipMemcpyDeviceToHost));
    return VMC_OK;
}

int vmc_batch_get_parameters(vmc_batch* b, double* alpha_gamma)
{
    if (!b || !alpha_gamma) return fail(VMC_E_INVALID, "null argument");
    for (size_t i = 0; i < b->theta.size(); i++) alpha_gamma[i] = b->theta[i];
    return VMC_OK;
}

int vmc_batch_set_shape(vmc_batch* b, int system, int lanes_per_walker)
{
    if (!b) return fail(VMC_E_INVALID, "null batch");
    if (system < 0 || system >= (int)b->sys.size()) return fail(VMC_E_INVALID, "no such system");
    vmc::BatchSystem& B = b->hs[b->dev_of[system]];
    const int id = vmc::batch_shape_id(B.np, lanes_per_walker);
    if (id < 0) return fail(VMC_E_INVALID, "no batch kernel shape for these lanes per walker");
    B.K = lanes_per_walker;
    B.shape = id;
    batch_plan(b);
    return VMC_OK;
}

int vmc_batch_set_launch(vmc_batch* b, int launch)
{
    if (!b) return fail(VMC_E_INVALID, "null batch");
    if (launch != VMC_BATCH_SIDE_BY_SIDE && launch != VMC_BATCH_PERSISTENT)
        return fail(VMC_E_INVALID, "launch: VMC_BATCH_SIDE_BY_SIDE or VMC_BATCH_PERSISTENT");
    if (launch == VMC_BATCH_PERSISTENT) {
        if (!batch_persistent_ok(b))
            return fail(VMC_E_INVALID, "the persistent launch walks FP32 Slater x Gaussian x "
                                       "Pade-Jastrow with Coulomb under exact Langevin only");
        if (!b->persistent_grid) HIP_TRY(vmc::batch_persistent_grid(&b->persistent_grid));
    }
    b->launch = launch;
    return VMC_OK;
}

int vmc_batch_get_launch(vmc_batch* b, int* launch)
{
    if (!b || !launch) return fail(VMC_E_INVALID, "null argument");
    *launch = b->launch;
    return VMC_OK;
}

int vmc_batch_run(vmc_batch* b, int iterations, int steps, int equilibration, int batches,
                  double learning_rate, int use_graph, double* out, double* kernel_ms)
{
    if (!b) return fail(VMC_E_INVALID, "null batch");
    if (!b->walkers_set) return fail(VMC_E_STATE, "walkers have no positions yet");
    if (iterations < 1 || steps < 1 || equilibration < 0 || batches < 1 || !(learning_rate >= 0))
        return fail(VMC_E_INVALID,
                    "iterations, steps, batches >= 1; equilibration, learning_rate >= 0");
    const int nsys = (int)b->hs.size();
    for (auto& B : b->hs) {
        B.P.seed = b->seed + (uint64_t)B.user;   // system s retraces an engine seeded seed + s
        B.P.steps = steps;
        B.P.equil = equilibration;
    }
    const vmc::BatchTable T = batch_table(b);
    const int it0 = b->iteration;
    const int total = steps + equilibration;
    const vmc::BatchState st{b->counter, it0, batches, b->fresh ? 1 : 0, learning_rate};
    double* d_out = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    auto done = [&](int code) {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (d_out) (void)hipFree(d_out);
        return code;
    };
    hipError_t err;
    if ((err = hipMalloc(&d_out, sizeof(double) * (size_t)iterations * nsys * vmc::kBatchOut)) != hipSuccess ||
        (err = hipMemcpyAsync(b->d_table, &T, sizeof(T), hipMemcpyHostToDevice, b->stream)) != hipSuccess ||
        (err = hipMemcpyAsync(b->d_st, &st, sizeof(st), hipMemcpyHostToDevice, b->stream)) != hipSuccess ||
        (err = hipStreamSynchronize(b->stream)) != hipSuccess)
        return done(fail(VMC_E_DEVICE, std::string("batch run: ") + hipGetErrorString(err)));
    auto enqueue = [&]() -> int {
        if (int rc = batch_enqueue_walk(b, total)) return rc;
        HIP_TRY(vmc::launch_batch_tail(b->d_table, nsys, b->chunks, b->d_opt, b->d_theta,
                                       b->d_acc, b->d_part, d_out, b->d_st, it0, total,
                                       b->stream));
        return VMC_OK;
    };
    if (use_graph) {
        if ((err = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
            return done(fail(VMC_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(err)));
        const int rc = enqueue();
        const hipError_t e2 = hipStreamEndCapture(b->stream, &graph);
        if (rc != VMC_OK) return done(rc);
        if (e2 != hipSuccess) return done(fail(VMC_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(e2)));
        if ((err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0)) != hipSuccess)
            return done(fail(VMC_E_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(err)));
    }
    HIP_TRY(hipEventRecord(b->ev0, b->stream));
    for (int it = 0; it < iterations; it++) {
        if (use_graph) {
            if ((err = hipGraphLaunch(exec, b->stream)) != hipSuccess)
                return done(fail(VMC_E_DEVICE, std::string("graph launch: ") + hipGetErrorString(err)));
        } else if (int rc = enqueue()) {
            return done(rc);
        }
    }
    std::vector<double> h((size_t)iterations * nsys * vmc::kBatchOut);
    float ms = 0;
    if ((err = hipEventRecord(b->ev1, b->stream)) != hipSuccess ||
        (err = hipMemcpyAsync(h.data(), d_out, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                              b->stream)) != hipSuccess ||
        (err = hipMemcpyAsync(b->theta.data(), b->d_theta, sizeof(double) * b->theta.size(),
                              hipMemcpyDeviceToHost, b->stream)) != hipSuccess ||
        (err = hipStreamSynchronize(b->stream)) != hipSuccess ||
        (err = hipEventElapsedTime(&ms, b->ev0, b->ev1)) != hipSuccess)
        return done(fail(VMC_E_DEVICE, std::string("batch run: ") + hipGetErrorString(err)));
    b->counter += (uint64_t)iterations * total;
    b->iteration += iterations;
    b->fresh = false;
    if (out) std::copy(h.begin(), h.end(), out);   // [iteration][caller's system][16]
    if (kernel_ms) *kernel_ms = ms;
    return done(VMC_OK);
}

int vmc_batch_autotune(vmc_batch* b, int steps, int* lanes_out, double* ns_per_step_out)
{
    if (!b) return fail(VMC_E_INVALID, "null batch");
    if (!b->walkers_set) return fail(VMC_E_STATE, "walkers have no positions yet");
    if (steps < 1) return fail(VMC_E_INVALID, "steps >= 1");
    const int nsys = (int)b->hs.size();
    // each system alone on the whole GPU, walked from its saved positions
    // and drifts, which are restored after every candidate
    size_t maxpos = 0;
    for (const auto& y : b->sys) maxpos = std::max(maxpos, (size_t)y.num_walkers * 2 * y.num_particles);
    double* d_save = nullptr;
    HIP_TRY(hipMalloc(&d_save, sizeof(double) * 2 * maxpos));
    auto finish = [&](int code) {
        (void)hipFree(d_save);
        return code;
    };
    hipError_t err;
    const vmc::BatchState st{b->counter, 0, 1, b->fresh ? 1 : 0, 0.0};
    if ((err = hipMemcpy(b->d_st, &st, sizeof(st), hipMemcpyHostToDevice)) != hipSuccess)
        return finish(fail(VMC_E_DEVICE, std::string("autotune: ") + hipGetErrorString(err)));
    for (int d = 0; d < nsys; d++) {   // table order is kept until the re-plan below
        const vmc::BatchSystem B = b->hs[d];
        const size_t n = (size_t)B.P.W * B.P.ND;
        const size_t off = b->pos_off[B.user];
        if ((err = hipMemcpy(d_save, b->d_pos + off, sizeof(double) * n, hipMemcpyDeviceToDevice)) != hipSuccess ||
            (err = hipMemcpy(d_save + n, b->d_qf + off, sizeof(double) * n, hipMemcpyDeviceToDevice)) != hipSuccess)
            return finish(fail(VMC_E_DEVICE, std::string("autotune: ") + hipGetErrorString(err)));
        double best = 1e300;
        int bestK = B.K;
        for (int K : {B.K, 2 * B.K}) {
            const int id = vmc::batch_shape_id(B.np, K);
            if (id < 0) continue;
            vmc::WalkParams P = B.P;
            P.seed = b->seed + (uint64_t)B.user;
            P.steps = steps;
            P.equil = 0;
            float ms = 0;
            for (int rep = 0; rep < 2; rep++)   // the second launch is timed
                if ((err = hipEventRecord(b->ev0, b->stream)) != hipSuccess ||
                    (err = vmc::launch_batch_walk(b->cfg.precision, id, P, B.user, 0, steps, 0,
                                                  b->d_st, b->d_theta, b->stream)) != hipSuccess ||
                    (err = hipEventRecord(b->ev1, b->stream)) != hipSuccess ||
                    (err = hipStreamSynchronize(b->stream)) != hipSuccess ||
                    (err = hipEventElapsedTime(&ms, b->ev0, b->ev1)) != hipSuccess ||
                    (err = hipMemcpy(b->d_pos + off, d_save, sizeof(double) * n, hipMemcpyDeviceToDevice)) != hipSuccess ||
                    (err = hipMemcpy(b->d_qf + off, d_save + n, sizeof(double) * n, hipMemcpyDeviceToDevice)) != hipSuccess)
                    return finish(fail(VMC_E_DEVICE, std::string("autotune: ") + hipGetErrorString(err)));
            const double ns = 1e6 * ms / (double(P.W) * steps);
            if (ns < best) {
                best = ns;
                bestK = K;
            }
        }
        b->hs[d].K = bestK;
        b->hs[d].shape = vmc::batch_shape_id(B.np, bestK);
        b->cost[B.user] = best;
        if (lanes_out) lanes_out[B.user] = bestK;
        if (ns_per_step_out) ns_per_step_out[B.user] = best;
    }
    batch_plan(b);
    if (!batch_persistent_ok(b)) return finish(VMC_OK);
    // the launch: the whole batch walked side by side and as one persistent
    // launch (the faster of the second and third run of each), from the same
    // positions and drifts, over `steps` steps plus up to one refresh
    // interval, so that the trial crosses a refresh boundary as an iteration
    // does; the faster form is kept
    (void)hipFree(d_save);
    d_save = nullptr;
    if ((!b->persistent_grid && (err = vmc::batch_persistent_grid(&b->persistent_grid)) != hipSuccess) ||
        (err = hipMalloc(&d_save, sizeof(double) * 2 * b->pos_total)) != hipSuccess)
        return finish(fail(VMC_E_DEVICE, std::string("autotune: ") + hipGetErrorString(err)));
    const int refresh = b->cfg.refresh_interval;
    const int trial = steps + (refresh > 0 ? std::min(refresh, 4 * steps) : 0);
    for (auto& B : b->hs) {
        B.P.seed = b->seed + (uint64_t)B.user;
        B.P.steps = trial;
        B.P.equil = 0;
    }
    const vmc::BatchTable T = batch_table(b);
    const size_t np = b->pos_total;
    if ((err = hipMemcpy(b->d_table, &T, sizeof(T), hipMemcpyHostToDevice)) != hipSuccess ||
        (err = hipMemcpy(d_save, b->d_pos, sizeof(double) * np, hipMemcpyDeviceToDevice)) != hipSuccess ||
        (err = hipMemcpy(d_save + np, b->d_qf, sizeof(double) * np, hipMemcpyDeviceToDevice)) != hipSuccess)
        return finish(fail(VMC_E_DEVICE, std::string("autotune: ") + hipGetErrorString(err)));
    double best_ms = 1e300;
    int best_launch = VMC_BATCH_SIDE_BY_SIDE;
    for (int launch : {VMC_BATCH_SIDE_BY_SIDE, VMC_BATCH_PERSISTENT}) {
        b->launch = launch;
        float ms = 0, best_rep = 1e30f;
        for (int rep = 0; rep < 3; rep++) {   // the faster of the second and third run
            if ((err = hipEventRecord(b->ev0, b->stream)) != hipSuccess)
                return finish(fail(VMC_E_DEVICE, std::string("autotune: ") + hipGetErrorString(err)));
            if (int rc = batch_enqueue_walk(b, trial)) return finish(rc);
            if ((err = hipEventRecord(b->ev1, b->stream)) != hipSuccess ||
                (err = hipStreamSynchronize(b->stream)) != hipSuccess ||
                (err = hipEventElapsedTime(&ms, b->ev0, b->ev1)) != hipSuccess ||
                (err = hipMemcpy(b->d_pos, d_save, sizeof(double) * np, hipMemcpyDeviceToDevice)) != hipSuccess ||
                (err = hipMemcpy(b->d_qf, d_save + np, sizeof(double) * np, hipMemcpyDeviceToDevice)) != hipSuccess)
                return finish(fail(VMC_E_DEVICE, std::string("autotune: ") + hipGetErrorString(err)));
            if (rep > 0 && ms < best_rep) best_rep = ms;
        }
        ms = best_rep;
        if (ms < best_ms) {
            best_ms = ms;
            best_launch = launch;
        }
    }
    b->launch = best_launch;
    return finish(VMC_OK);
}

}  // extern "C"

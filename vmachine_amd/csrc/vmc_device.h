// vmc_device.h -- launch parameters shared by the host engine and the HIP kernels.
#pragma once
#include <stdint.h>

namespace vmc {

// sampling methods (same values as include/vmc.h)
enum : int { kBruteForce = 0, kImportanceCompat = 1, kImportanceExact = 2, kLearned = 3 };

// learned proposal network: 18 inputs, two tanh layers of 64, 4 outputs;
// parameters W1[u][m], b1[u], W2[u][c], b2[u], W3[r][u], b3[r]
enum : int {
    kLpIn = 18, kLpHid = 64, kLpOut = 4,
    kLpW1 = 0,
    kLpB1 = kLpW1 + kLpHid * kLpIn,
    kLpW2 = kLpB1 + kLpHid,
    kLpB2 = kLpW2 + kLpHid * kLpHid,
    kLpW3 = kLpB2 + kLpHid,
    kLpB3 = kLpW3 + kLpOut * kLpHid,
    kLpParams = kLpB3 + kLpOut
};

// per-walker accumulator slots written by the walk kernel
enum : int {
    kAccE = 0, kAccE2, kAccT, kAccVext, kAccVint, kAccT2, kAccVext2, kAccVint2, kAccAccept,
    kAccGAlpha, kAccGGamma, kAccGAlphaE, kAccGGammaE, kAccBad,
    kNumWalkerAcc
};

// width of one block's row of RBM gradient-window sums: the RBM-Gaussian row
// (a_k at k), then the RBM-product row (b_h at h, W[k][h] at H + k + ND*h, the
// reference's parameter layout, ref rbmproduct.cpp:18-27), for the sums of
// dlnpsi/dtheta and then of E_L * dlnpsi/dtheta
constexpr int rbm_row_width(int nd, int h) { return 2 * (nd + h + nd * h); }

struct WalkParams {
    // system
    int N;              // particles
    int n_up;           // particles 0..n_up-1 are spin up (reference: i < N/2)
    double omega;
    double sqrt_omega;      // sqrt(omega), 1/sqrt(omega), sqrt(dt): set by the host
    double inv_sqrt_omega;
    double sqrt_dt;
    // elements present
    int has_gauss, has_slater, has_pade, coulomb;
    double alpha;       // Gaussian alpha
    double gamma;       // Pade gamma
    double a_same, a_anti;
    // RBM elements (ref WaveFunctions/rbmgaussian.cpp, rbmproduct.cpp)
    int has_rbmg, has_rbmp;
    int ND;             // degrees of freedom, 2N
    int H;              // hidden units
    int sorting;        // RBM-product sees the sorted coordinate vector
    double sig2_g;      // RBM-Gaussian sigma^2
    double isig2_p;     // 1 / sigma_p^2, sigma_p = 10 sigma (ref rbmproduct.cpp:15)
    int q13_psi_ratio;  // reference RBM-Gaussian ratio exp(d/(2 sigma^2))
    int q6_w_unscaled;  // reference dlnpsi/dW without 1/sigma_p^2
    int rbm_mf;         // RBM-product contractions: 0 lanes, 1 FP32 matrix cores, 2 BF16 matrix cores
    const double* rbm_a;    // [ND]
    const double* rbm_b;    // [H]
    const double* rbm_W;    // [ND][H], W[k][h] at k*H + h
    double* rbm_acc;        // [blocks][rbm_row_width(ND, H)] gradient-window sums per block
    // optimisation loop on the device (null otherwise): alpha, gamma, the
    // random counter, gradient window and drift flag of the current iteration
    const double* dyn_ag;   // [2] alpha, gamma
    const struct OptState* dyn_st;
    // learned proposal (sampling == kLearned)
    const uint16_t* lp_w;   // [kLpParams] network weights, bfloat16
    double lp_sigma0;       // untrained move scale, units of 1/sqrt(omega)
    double lp_mu_max;       // |move mean| < lp_mu_max / sqrt(omega)
    double lp_ls_max;       // |log(scale / sigma0)| < lp_ls_max
    // Metropolis
    int sampling;
    double dt;          // importance dt or brute-force step length
    int q2_upper;       // reference Pade ratio over pairs j > i_p only
    int q3_half;        // reference 1/2 in dlnpsi/dgamma
    double drift_a;     // drift limit a (exact Langevin only; 0 = none)
    // random numbers
    uint64_t seed;
    uint64_t counter0;  // step counter of the first step of this launch
    // schedule: this launch runs steps t_begin <= t < t_end of an iteration of
    // equil burn-in + steps sampled steps; every launch starts by rebuilding the
    // walker state (Slater inverse in FP64) from the positions
    int steps;          // sampled steps of the iteration
    int equil;          // burn-in steps before sampling
    int t_begin, t_end;
    int acc_add;        // 0: overwrite the per-walker accumulators, 1: add to them
    int window_lo;      // gradient window: step index t with window_lo < t < window_hi
    int window_hi;
    // walkers
    int W;
    double* pos;        // [W][N*2] FP64, read at launch start, written at the end
    double* qf;         // [W][N*2] stale quantum force (importance compat), may be null
    double* acc;        // [W][kNumWalkerAcc] per-walker accumulators (overwritten)
    float* rec;         // [rec_parts][rec_walkers][steps] instant values, may be null:
                        // E_L, then (rec_parts == 4) kinetic, external, interaction
    int rec_walkers;
    int rec_parts;      // 1 or 4
    // electron densities counted after every sampled step (ref
    // src/sampler.cpp:639-692): bit 0 radial one-body, bit 1 spatial
    // one-body, bit 2 radial two-body; 0 = off
    int dens_flags;
    int dens_bins;
    double dens_step;       // maxRadius / bins
    double dens_max_radius;
    unsigned long long* dens_radial;    // [bins]
    unsigned long long* dens_spatial;   // [bins][bins], row = y cell, column = x cell
    unsigned long long* dens_pair;      // [bins][bins], row = bin of the lower-index particle
    int init_qf;        // 1: set the stored drift to the drift at the loaded positions
};

// the values of one walk launch that change between launches and iterations
// while the rest of WalkParams stays fixed (the batch kernel reads them from
// device state; the single-system kernel copies them from its WalkParams)
struct WalkDyn {
    double alpha, gamma;  // Gaussian alpha, Pade gamma
    uint64_t counter0;
    int t_begin, t_end, acc_add;
    int window_lo, window_hi;
    int init_qf;
};

// training of the learned proposal: one sample per walker (vmc_learn.hip)
struct LpTrainParams {
    int B, N;               // samples, electrons
    double omega;
    uint64_t seed;
    double sigma0, mu_max, ls_max;
    const float* theta;     // [kLpParams] FP32 master weights
    const double* pos;      // [B][2N] configurations
    const int* electrons;   // [B] moved electron
    const double* normals;  // [B][2]
    double* xnew;           // [B][2] proposed position of the electron
    int* nnb;               // [B] neighbours in the features at x'
    float* act;             // [B][2][f | h1 | h2 | o] activations of both passes
    float* delta;           // [B][2][d1 | d2 | d3] layer deltas of both passes
    const double* eval;     // [B][9 + 2N] the walk's evaluation after the move
    double* stats;          // [B][5]: J, A, log|psi'/psi|^2, x', y'
};

// per-walker outputs of the parity kernel
struct EvalOut {
    const double* newv2;   // [W][2]: force a move of particle moves[w] to these coordinates (or null)
    double* o0;        // [W][9 + nd]
    double* o1;
    double* ratios;    // [W][num_elements] per-element ratio of the forced move, or null
    double* g0;        // [W][num_elements * maxpar] dlnpsi/dtheta before, or null (zeroed by the host)
    double* g1;        // after
    int rows[6];       // element kind of each parameter row
    int num_elements, maxpar;
};

// ---- optimisation loop on the device (vmc_optim.hip) ----
struct OptState {
    uint64_t counter0;  // random counter of the iteration's first step
    int iteration;      // iteration (gradient window = iteration % batches)
    int batches;        // gradient batches
    int fresh;          // 1: set the stored drift from the positions (cleared after an iteration)
    int step;           // optimiser steps taken
};

// where the sums of each parameter of the row-major elements x maxParams
// matrix are, and the optimiser's constants
struct OptLayout {
    int rows, mp, nd, H;
    int row_gauss, row_pade, row_rbmg, row_rbmp;
    int kind;           // 0 ADAM, 1 SGD / gradient descent (momentum form)
    double lr, gamma, monotonic;
    double M, Sw;       // samples and gradient-window samples per iteration
};
enum : int { kOptAlpha = kNumWalkerAcc, kOptGamma, kOptOut };

// allocation-free optimisation loop (vmc_optimize_static): its device state
// lives in one file-scope array per GPU, laid out as
// theta [max params] | m [max params] | v [max params] | alpha, gamma (4) |
// OptState (4) | history [max iterations][kOptOut]
constexpr int kStaticMaxParams = 8192;       // VMC_STATIC_MAX_PARAMS
constexpr int kStaticMaxIterations = 1024;   // VMC_STATIC_MAX_ITERATIONS
constexpr int kStaticTheta = 0;
constexpr int kStaticM = kStaticTheta + kStaticMaxParams;
constexpr int kStaticV = kStaticM + kStaticMaxParams;
constexpr int kStaticAG = kStaticV + kStaticMaxParams;
constexpr int kStaticState = kStaticAG + 4;
constexpr int kStaticHist = kStaticState + 4;
constexpr int kStaticDoubles = kStaticHist + kStaticMaxIterations * kOptOut;
static_assert(sizeof(OptState) <= 4 * sizeof(double), "OptState fits its slot");

// ---- batch of systems walked side by side (vmc_batch.hip) ----
enum : int { kBatchAlpha = kNumWalkerAcc, kBatchGamma, kBatchOut };

struct BatchSystem {
    WalkParams P;       // the system's parameters and buffers; per-iteration fields set in-kernel
    int shape;          // instantiation id (np, K)
    int np, K;
    int blocks;         // one-wave workgroups of the system's kernel
    int walker0;        // first row of the system in the shared walker buffers
    int user;           // the system's index in the caller's list
};

// the static description of every system of a batch (device memory, written
// by the host before a run, read-only for the kernels)
constexpr int kBatchMaxSystems = 8;
struct BatchTable {
    BatchSystem sys[kBatchMaxSystems];   // launch order: costliest system first
    int nsys;
};

struct BatchState {
    uint64_t counter0;  // random counter of the iteration's first step
    int iteration;      // optimisation iteration (gradient window = iteration % batches)
    int batches;        // gradient batches
    int fresh;          // 1: set the stored drift from the positions (cleared after an iteration)
    double learning_rate;
};

struct BatchOpt {
    double m[2], v[2];  // ADAM moments of alpha and gamma
    int step;
};

// which kernel instantiation runs a configuration: np particle slots per spin
// (Slater orbitals per spin), K lanes per walker, RBM elements or not
struct KernelShape {
    int np, K;
    bool rbm;
};

}  // namespace vmc

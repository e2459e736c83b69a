// vmc_device.h -- launch parameters shared by the host engine and the HIP kernels.
#pragma once
#include <stdint.h>

namespace vmc {

// sampling methods (same values as include/vmc.h)
enum : int { kBruteForce = 0, kImportanceCompat = 1, kImportanceExact = 2 };

// per-walker accumulator slots written by the walk kernel
enum : int {
    kAccE = 0, kAccE2, kAccT, kAccVext, kAccVint, kAccT2, kAccVext2, kAccVint2, kAccAccept,
    kAccGAlpha, kAccGGamma, kAccGAlphaE, kAccGGammaE, kAccBad,
    kNumWalkerAcc
};

struct WalkParams {
    // system
    int N;              // particles
    int n_up;           // particles 0..n_up-1 are spin up (reference: i < N/2)
    double omega;
    // elements present
    int has_gauss, has_slater, has_pade, coulomb;
    double alpha;       // Gaussian alpha
    double gamma;       // Pade gamma
    double a_same, a_anti;
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
    float* rec;         // [rec_walkers][steps] local energies, may be null
    int rec_walkers;
    int init_qf;        // 1: set the stored drift to the drift at the loaded positions
};

}  // namespace vmc

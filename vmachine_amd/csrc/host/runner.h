// runner.h -- the optimisation loop of `vmachine input.in` on the engine.
#pragma once
#include <iosfwd>
#include <string>
#include <vector>

#include "config.h"

namespace vmc_host {

struct RunOptions {
    int walkers = 0;            // 0: chosen from numSteps (64 sweeps per walker per iteration)
    int precision = 64;         // 64 or 32
    int refresh = 100;          // FP32: steps between FP64 inverse rebuilds
    bool exact = false;         // exact Metropolis-Hastings Langevin instead of the
                                // reference's Green's-function ratio
    bool reference_quirks = true;
    double drift_limit = 0.0;   // with exact: Umrigar drift limit a
    // Unsampled Metropolis sweeps (ND steps each) every walker takes before the
    // first iteration.  The reference's one chain runs numSteps consecutive
    // steps per iteration, so its start-up transient is a small part of the
    // first iteration; numSteps/W steps per walker are not, and without this
    // burn-in the first iterations' parameter gradients come from the
    // transient of the initial positions.
    int burn_in_sweeps = 100;
    bool have_seed = false;
    unsigned long long seed = 0;
};

// Host-side optimisers (ref src/Optimization/optimization.cpp:10-16,
// adam.cpp:12-38, sgd.cpp:15-29, gradientdescent.cpp:15-30).
struct Optimizer {
    std::string kind;           // adam, sgd, gd
    double eta = 0.1, gamma = 0, monotonic = 0;
    double beta1 = 0.89, beta2 = 0.999, epsilon = 1e-8;
    int step = 0;
    std::vector<double> m, v;
    void init(const std::string& k, int size, double lr, double g, double mono);
    int batches() const { return kind == "adam" ? 10 : 1; }
    // g = 2(<G E> - <E><G>); returns the update the caller subtracts
    std::vector<double> update(double avgE, const std::vector<double>& avgG,
                               const std::vector<double>& avgGE);
};

// Initial parameter matrix (rows = elements, row-major) for the configured
// initialWeights (ref src/InitialWeights/automatize.cpp:55-105, constant.cpp,
// xavier.cpp, randomuniform.cpp, fromfile.cpp).
std::vector<double> initial_weights(const Config& c, int rows, int cols, unsigned long long seed);
// the reference's label of the initial weights (the method automatize resolves to)
std::string initial_weights_label(const Config& c);

// Runs the simulation and writes the reference's terminal output to `out`.
// Returns the process exit status (0, or 143 for configuration errors as the
// reference's MPI_Abort(..., 143)).
int run_simulation(const Config& c, const RunOptions& o, std::ostream& out, std::ostream& err);

// Automated blocking (ref include/block/blocker.cpp): mean and standard error
// of a 2^d-long correlated series.  Returns false if the length is not a
// power of two.
bool blocking(std::vector<double> x, double& mean, double& stderr_);

}  // namespace vmc_host

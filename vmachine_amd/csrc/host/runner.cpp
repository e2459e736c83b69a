// runner.cpp -- `vmachine input.in` on the MI355X engine: the reference's
// System::runSimulation loop (ref src/system.cpp:81-139) with the sampling
// iteration (runMetropolisCycles + Sampler::computeTotals) replaced by vmc_run.
#include "runner.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <random>
#include <sstream>

#include "vmc.h"

namespace vmc_host {

// ---------------------------------------------------------------------------
void Optimizer::init(const std::string& k, int size, double lr, double g, double mono)
{
    kind = k;
    eta = lr;
    gamma = g;
    monotonic = mono;
    step = 0;
    m.assign(size, 1.0);   // ref adam.cpp:17-19 and sgd.cpp:20 start from ones
    v.assign(size, 1.0);
}

std::vector<double> Optimizer::update(double avgE, const std::vector<double>& avgG,
                                      const std::vector<double>& avgGE)
{
    step += 1;
    std::vector<double> out(avgG.size());
    for (size_t i = 0; i < avgG.size(); i++) {
        const double g = 2.0 * (avgGE[i] - avgE * avgG[i]);
        if (kind == "adam") {
            m[i] = beta1 * m[i] + (1 - beta1) * g;
            v[i] = beta2 * v[i] + (1 - beta2) * g * g;
            const double mh = m[i] / (1 - std::pow(beta1, step));
            const double vh = v[i] / (1 - std::pow(beta2, step));
            out[i] = eta * mh / (std::sqrt(vh) + epsilon);
        } else {
            const double mono = 1.0 / std::pow(double(step), monotonic);
            v[i] = gamma * v[i] + eta * g * mono;
            out[i] = v[i];
        }
    }
    return out;
}

// ---------------------------------------------------------------------------
namespace {

// regularised lower incomplete gamma P(a, x) (series / Lentz continued fraction)
double gamma_p(double a, double x)
{
    if (x <= 0) return 0.0;
    const double lg = std::lgamma(a);
    if (x < a + 1.0) {
        double sum = 1.0 / a, del = sum, ap = a;
        for (int i = 0; i < 100000 && std::fabs(del) > std::fabs(sum) * 1e-17; i++) {
            ap += 1.0;
            del *= x / ap;
            sum += del;
        }
        return sum * std::exp(-x + a * std::log(x) - lg);
    }
    double b = x + 1.0 - a, c = 1e300, d = 1.0 / b, h = d;
    for (int i = 1; i < 100000; i++) {
        const double an = -i * (i - a);
        b += 2.0;
        d = an * d + b;
        if (std::fabs(d) < 1e-300) d = 1e-300;
        c = b + an / c;
        if (std::fabs(c) < 1e-300) c = 1e-300;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (std::fabs(del - 1.0) < 1e-17) break;
    }
    return 1.0 - std::exp(-x + a * std::log(x) - lg) * h;
}

double chi2_q95(int dof)
{
    double lo = 0, hi = 10.0 * dof + 100.0;
    for (int i = 0; i < 200; i++) {
        const double mid = 0.5 * (lo + hi);
        (gamma_p(0.5 * dof, 0.5 * mid) < 0.95 ? lo : hi) = mid;
    }
    return 0.5 * (lo + hi);
}

int element_kind(const std::string& key)
{
    if (key == "gaussian") return VMC_GAUSSIAN;
    if (key == "slaterDeterminant") return VMC_SLATER;
    if (key == "padeJastrow") return VMC_PADE;
    if (key == "RBMGaussian") return VMC_RBM_GAUSSIAN;
    if (key == "RBMProduct") return VMC_RBM_PRODUCT;
    return -1;
}

std::string element_name(int kind)
{
    switch (kind) {
    case VMC_GAUSSIAN: return "Gaussian";
    case VMC_SLATER: return "Slater determinant";
    case VMC_PADE: return "Padé-Jastrow";
    case VMC_RBM_GAUSSIAN: return "RBM-Gaussian";
    case VMC_RBM_PRODUCT: return "RBM-product";
    }
    return "?";
}

struct Averages {
    double E, T, V, I, E2, T2, V2, I2, varE, varT, varV, varI, acc;
};

}  // namespace

// an integer matrix as the reference writes it, Eigen's default stream format:
// every coefficient right-aligned to the widest one, one space between
// columns, one row per line (ref tests/Eigen/src/Core/IO.h:129-199)
static void write_counts(const std::string& path, const std::vector<int64_t>& v, int rows,
                         int cols)
{
    size_t w = 1;
    for (int64_t x : v) w = std::max(w, std::to_string(x).size());
    std::ofstream f(path);
    for (int i = 0; i < rows; i++) {
        for (int j = 0; j < cols; j++) {
            if (j) f << ' ';
            f << std::setw((int)w) << v[(size_t)i * cols + j];
        }
        f << '\n';
    }
}

bool blocking(std::vector<double> x, double& mean, double& stderr_)
{
    const size_t n = x.size();
    if (n == 0 || (n & (n - 1)) != 0) return false;
    int d = 0;
    while ((size_t(1) << d) < n) d++;
    double s0 = 0;
    for (double v : x) s0 += v;
    const double mu = s0 / double(n);
    std::vector<double> gam(d), s(d), X(n);
    for (size_t i = 0; i < n; i++) X[i] = x[i] - mu;
    size_t m = n;
    for (int k = 0; k < d; k++) {           // ref blocker.cpp:42-54, 73-80
        double g1 = 0, v0 = 0;
        for (size_t i = 0; i + 1 < m; i++) g1 += X[i] * X[i + 1];
        for (size_t i = 0; i < m; i++) v0 += X[i] * X[i];
        gam[k] = g1 / double(m);
        s[k] = v0 / double(m);
        for (size_t i = 0; i < m / 2; i++) {
            x[i] = 0.5 * (x[2 * i] + x[2 * i + 1]);
            X[i] = x[i] - mu;
        }
        m /= 2;
    }
    std::vector<double> cs(d);
    double run = 0;
    for (int j = 0; j < d; j++) {           // cumsum(flip(M)), ref blocker.cpp:55-60
        const int i = d - 1 - j;
        run += (gam[i] / s[i]) * (gam[i] / s[i]) * std::ldexp(1.0, d - i);
        cs[j] = run;
    }
    int k;
    for (k = d - 1; k >= 0; k--)
        if (cs[k] < chi2_q95(k + 1)) break;
    int K = d - (k + 1);
    if (K > d - 1) K = d - 1;
    mean = mu;
    stderr_ = std::sqrt(s[K] / std::ldexp(1.0, d - K));
    return true;
}

// the preset initialWeights: automatize resolves to (ref automatize.cpp:83-101)
static std::string resolved_weights_method(const Config& c)
{
    if (c.initialWeights != "automatize") return c.initialWeights;
    const std::string label = trial_wave_function_label(c.elements);
    if (label == "VMC" || label == "BVMC" || label == "PRBM") return "constant";
    if (label == "RBM" || label == "RBMSJ" || label == "RBMPJ") return "xavier";
    return "randomuniform";
}

std::string initial_weights_label(const Config& c)
{
    const std::string m = resolved_weights_method(c);   // labels of ref src/InitialWeights/*.h
    if (m == "constant") return "constant";
    if (m == "xavier") return "Xavier";
    if (m == "fromfile") return "from-file";
    return "random uniform";
}

std::vector<double> initial_weights(const Config& c, int rows, int cols, unsigned long long seed)
{
    std::vector<double> P((size_t)rows * cols, 0.0);
    std::mt19937_64 gen(seed ^ 0x5eed5eedULL);
    auto uniform = [&](double lo, double hi) {
        return std::uniform_real_distribution<double>(lo, hi)(gen);
    };
    std::string method = c.initialWeights;
    double factor = c.initialWeightsValue;
    if (method == "automatize") {            // ref automatize.cpp:83-101
        const std::string label = trial_wave_function_label(c.elements);
        if (label == "VMC" || label == "BVMC") { method = "constant"; factor = 1.0; }
        else if (label == "RBM" || label == "RBMSJ" || label == "RBMPJ") method = "xavier";
        else if (label == "PRBM") { method = "constant"; factor = 0.0; }
        else if (label == "SSJ") { method = "randomuniform"; factor = 0.2; }
        else { method = "randomuniform"; factor = 0.01; }
    }
    if (method == "constant") {              // ref constant.cpp:11-26
        std::fill(P.begin(), P.end(), factor);
    } else if (method == "randomuniform") {  // ref randomuniform.cpp (factor x U(0,1))
        if (!c.initialWeightsHasValue && c.initialWeights == "randomuniform") factor = 1.0;
        for (double& p : P) p = factor * uniform(0.0, 1.0);
    } else if (method == "xavier") {         // ref xavier.cpp:10-23
        const double a = 1.0 / std::sqrt(double(c.numParticles * c.numDimensions));
        for (double& p : P) p = uniform(-a, a);
        if (rows > 3)
            for (int j = 0; j < cols; j++) P[3 * cols + j] = uniform(0.0, 0.2);
    } else if (method == "fromfile") {       // ref fromfile.cpp:30-74
        std::ifstream in(c.weightsFile);
        if (!in.is_open()) throw ConfigError("File '" + c.weightsFile + "' was not found");
        for (int i = 0; i < rows; i++)
            for (int j = 0; j < cols; j++)
                if (!(in >> P[(size_t)i * cols + j]))
                    throw ConfigError("weights file '" + c.weightsFile + "' is too short");
    }
    return P;
}

int run_simulation(const Config& c, const RunOptions& o, std::ostream& out, std::ostream& err)
{
    const auto start = std::chrono::system_clock::now();
    // ---- map the configuration onto the engine ----
    if (c.hamiltonian != "harmonicOscillator") {
        err << "The engine runs the harmonic-oscillator Hamiltonian (hamiltonian: "
               "harmonicOscillator); '" << c.hamiltonian << "' is outside it" << std::endl;
        return 143;
    }
    if (c.numDimensions != 2) {
        err << "The engine runs numDimensions: 2" << std::endl;
        return 143;
    }
    if (c.initialState != "randomNormal") {
        err << "The engine draws initialState: randomNormal" << std::endl;
        return 143;
    }
    const int dens_flags = (c.computeOneBodyDensity ? VMC_DENSITY_RADIAL : 0) |
                           (c.computeOneBodyDensity2 ? VMC_DENSITY_SPATIAL : 0) |
                           (c.computeTwoBodyDensity ? VMC_DENSITY_TWO_BODY : 0);
    if (dens_flags && (c.numberOfBins < 1 || c.numberOfBins > VMC_DENSITY_MAX_BINS ||
                       !(c.maxRadius > 0))) {
        err << "electron densities need numberOfBins in 1.." << VMC_DENSITY_MAX_BINS
            << " and maxRadius > 0" << std::endl;
        return 143;
    }
    vmc_config ec = {};
    ec.num_particles = c.numParticles;
    ec.num_dimensions = c.numDimensions;
    ec.omega = c.omega;
    ec.interaction = c.interaction == "coulomb" ? 1 : 0;
    ec.num_elements = (int)c.elements.size();
    if (ec.num_elements < 1 || ec.num_elements > 6) {
        err << "between 1 and 6 wave-function elements are needed" << std::endl;
        return 143;
    }
    for (int e = 0; e < ec.num_elements; e++) {
        ec.kinds[e] = element_kind(c.elements[e]);
        if (ec.kinds[e] < 0) {
            err << "wave-function element '" << c.elements[e] << "' is outside the engine"
                << std::endl;
            return 143;
        }
        if (ec.kinds[e] == VMC_SLATER && c.basis != "hermite" && c.numParticles > 2) {
            err << "slaterDeterminant needs basis: hermite" << std::endl;
            return 143;
        }
    }
    ec.num_hidden = c.numHiddenUnits;
    ec.sigma = c.sigma;
    ec.sorting = c.sorting ? 1 : 0;
    ec.sampling = c.sampling == "bruteForce" ? VMC_BRUTE_FORCE
                                              : (o.exact ? VMC_IMPORTANCE_EXACT : VMC_IMPORTANCE);
    ec.step_length = c.stepLength;
    ec.quirks = o.reference_quirks ? VMC_QUIRKS_REFERENCE : 0u;
    ec.precision = o.precision;
    ec.refresh_interval = o.precision == 32 ? o.refresh : 0;
    ec.drift_limit = o.exact ? o.drift_limit : 0.0;
    const int total0 = c.totalStepsWOEqui;
    int W = o.walkers;
    // default: 64 sweeps per walker per iteration.  Walkers lag behind each
    // parameter update by their equilibration time; with only a few sweeps
    // per iteration that lag feeds back into the optimiser's momentum and the
    // parameters oscillate, while 64 sweeps let the ensemble follow the
    // updates as the reference's single long chain does.
    if (W <= 0) W = std::max(1, std::min(32768, total0 / (64 * c.numParticles * c.numDimensions)));
    ec.num_walkers = W;

    unsigned long long seed = o.have_seed ? o.seed : (c.seeded ? (unsigned long long)c.seed : 0ULL);
    if (!o.have_seed && !c.seeded) seed = std::random_device{}();

    vmc_engine* eng = nullptr;
    const int created = vmc_engine_create(&ec, &eng);
    if (created != VMC_OK) {
        err << vmc_last_error() << std::endl;
        return created == VMC_E_INVALID ? 143 : 1;   // configuration error, else device error
    }
    const int rows = ec.num_elements, cols = vmc_max_parameters(&ec);
    std::vector<double> P;
    try {
        P = initial_weights(c, rows, cols, seed);
    } catch (const ConfigError& ex) {
        err << ex.what() << std::endl;
        vmc_engine_destroy(eng);
        return 143;
    }
    auto fail = [&](const char* what) {
        err << what << ": " << vmc_last_error() << std::endl;
        vmc_engine_destroy(eng);
        return 1;
    };
    double none = 0;   // a Slater-only product has no parameters
    auto params = [&]() { return P.empty() ? &none : P.data(); };
    if (vmc_set_parameters(eng, params(), rows, cols) != VMC_OK) return fail("parameters");
    if (vmc_init_walkers(eng, seed) != VMC_OK) return fail("walkers");
    const int dof = c.numParticles * c.numDimensions;
    if (o.burn_in_sweeps > 0) {
        vmc_sums s;
        std::vector<double> g((size_t)rows * std::max(cols, 1)), ge(g.size());
        if (vmc_run(eng, 1, o.burn_in_sweeps * dof, 0, 1, &s, g.data(), ge.data()) != VMC_OK)
            return fail("burn-in");
    }

    Optimizer opt;
    opt.init(c.optimization, rows * cols, c.learningRate, c.optGamma, c.optMonotonic);
    const int B = opt.batches();
    if (std::max(1, total0 / W) < B) {
        // every gradient window needs at least one step per walker
        err << "numSteps / walkers = " << std::max(1, total0 / W) << " steps per walker is fewer "
            << "than the optimiser's " << B << " gradient batches: use fewer walkers" << std::endl;
        vmc_engine_destroy(eng);
        return 143;
    }

    // ---- header: the reference's logo and information blocks with this run's
    // values (ref src/printinfo.cpp:14-140, src/system.cpp:83-91) ----
    out << "__      ____  __          _____ _    _ _____ _   _ ______ " << std::endl;
    out << "\\ \\    / /  \\/  |   /\\   / ____| |  | |_   _| \\ | |  ____|" << std::endl;
    out << " \\ \\  / /| \\  / |  /  \\ | |    | |__| | | | |  \\| | |__   " << std::endl;
    out << "  \\ \\/ / | |\\/| | / /\\ \\| |    |  __  | | | | . ` |  __|  " << std::endl;
    out << "   \\  /  | |  | |/ ____ \\ |____| |  | |_| |_| |\\  | |____ " << std::endl;
    out << "    \\/   |_|  |_/_/    \\_\\_____|_|  |_|_____|_| \\_|______|" << std::endl;
    out << "==========================================================" << std::endl << std::endl;
    std::time_t t0 = std::chrono::system_clock::to_time_t(start);
    out << "Started computation at " << std::ctime(&t0);
    out << "Running on 1 MI355X GPU (HIP), " << W << " walkers, FP" << o.precision << " walk"
        << std::endl;
    out << "Simulation is run from the directory: " << c.path << std::endl;
    int nparams = 0;
    for (int e = 0; e < rows; e++) nparams += vmc_num_parameters(&ec, e);
    auto block = [&](const char* title) {
        out << std::endl << std::endl << title << std::endl
            << "==============================================" << std::endl;
    };
    out << std::fixed << std::boolalpha << std::setprecision(6);
    block("             SYSTEM INFORMATION");
    out << "Number of particles:      " << c.numParticles << std::endl;
    out << "Number of dimensions:     " << c.numDimensions << std::endl;
    out << "Interaction style:        " << (ec.interaction ? "Coulomb" : "no interaction") << std::endl;
    out << "Hamiltonian:              harmonic oscillator" << std::endl;
    out << "Initial state:            random normal" << std::endl;
    out << "Oscillator frequency:     " << c.omega << std::endl;
    block("           WAVE FUNCTION INFORMATION");
    for (int e = 0; e < rows; e++)
        out << "Element " << e << ":                " << element_name(ec.kinds[e]) << std::endl;
    out << "Basis:                    " << (c.basis == "hermite" ? "Hermite" : "no basis") << std::endl;
    out << "Initial parameters:       " << initial_weights_label(c) << std::endl;
    out << "Number of parameters:     " << nparams << std::endl;
    out << "Number of hidden nodes:   " << c.numHiddenUnits << std::endl;
    out << "Print parameters to file: " << c.printParametersToFile << std::endl;
    block("            SIMULATION INFORMATION");
    out << "Max number of iterations: " << c.numIterations << std::endl;
    out << "Number of MC cycles:      " << c.totalStepsWOEqui << std::endl;
    out << "Number of burn-in cycles: " << c.equilibrationSteps << std::endl;
    out << "Learning rate:            " << c.learningRate << std::endl;
    out << "Step length:              " << c.stepLength << std::endl;
    out << "Equilibration fraction:   " << c.equilibrationFraction << std::endl;
    out << "Sampling:                 "
        << (ec.sampling == VMC_BRUTE_FORCE
                ? "brute force Metropolis"
                : (o.exact ? "Metropolis-Hastings (Langevin, exact ratio)" : "Metropolis-Hastings"))
        << std::endl;
    out << "Optimization:             "
        << (c.optimization == "adam" ? "ADAM"
                                     : (c.optimization == "sgd" ? "stochastic gradient descent"
                                                                : "gradient descent"))
        << std::endl;
    out << "Random number generator:  Philox4x32-10 (seed " << seed << ")" << std::endl;
    out << "Walkers and burn-in:      " << W << " walkers, " << o.burn_in_sweeps << " sweeps ("
        << o.burn_in_sweeps * dof << " steps per walker)" << std::endl;
    block("               PARTICLE DENSITY");
    out << "Compute radial one-body density:  " << c.computeOneBodyDensity << std::endl;
    out << "Compute spatial one-body density: " << c.computeOneBodyDensity2 << std::endl;
    out << "Compute radial two-body density:  " << c.computeTwoBodyDensity << std::endl;
    out << "Max radius of electron density:   " << c.maxRadius << std::endl;
    out << "Number of bins in each direction: " << c.numberOfBins << std::endl;
    block("                  RESAMPLING");
    out << "Do resampling:            " << c.doResampling << std::endl;
    out << "Print energies to file:   " << c.printEnergyToFile << std::endl;
    block("                  CONVERGENCE");
    out << "Check convergence:        " << c.checkConvergence << std::endl;
    out << "Tolerance:                " << c.tolerance << std::endl;
    out << "Number of energies:       " << c.numberOfEnergies << std::endl;
    block("                 ADAPTIVE STEPS");
    out << "Apply adaptive steps:       " << c.applyAdaptiveSteps << std::endl;
    out << "Range of adaptive steps:    " << c.rangeOfAdaptiveSteps << std::endl;
    out << "Additional steps:           " << c.additionalSteps << std::endl;
    out << "Additional steps last iter: " << c.additionalStepsLastIter << std::endl;
    out << std::endl << std::endl;
    out << std::noboolalpha;
    out << "Step  Energy  Energy_STD  Kinetic  Kinetic_STD  External  External_STD  "
           "Interaction  Interaction_STD  Acceptence  CPU_time" << std::endl;

    // ---- iteration loop (ref src/system.cpp:92-138) ----
    int numberOfNormalIterations = c.numIterations - c.rangeOfAdaptiveSteps - 1;
    const int stepsInit = total0;
    double globalTime = 0;
    double cumT2 = 0, cumV2 = 0, cumI2 = 0;   // never reset by the reference (sampler.cpp:78-80)
    std::vector<double> energies(c.numberOfEnergies, 0.0);
    bool checkConvergence = c.checkConvergence;
    Averages av{};
    std::vector<float> rec;
    int printed = 0;   // the sampler's own iteration counter (ref sampler.cpp:234)
    for (int iter = 0; iter < c.numIterations; iter++) {
        int total = stepsInit;
        const bool last = iter == numberOfNormalIterations + c.rangeOfAdaptiveSteps;
        if (c.applyAdaptiveSteps) {           // ref src/system.cpp:214-224
            int ratio = 1;
            if (last) ratio = 1 << c.additionalStepsLastIter;
            else if (iter >= numberOfNormalIterations) ratio = 1 << c.additionalSteps;
            total = stepsInit * ratio;
        }
        const int S = std::max(1, total / W);
        // burn-in is fixed once from the unscaled steps (ref src/userinput.cpp:74):
        // adaptive iterations lengthen only the sampled part (ref src/system.cpp:94-97)
        const int equil = int(std::max(1, stepsInit / W) * c.equilibrationFraction);
        vmc_sums s;
        std::vector<double> g((size_t)rows * std::max(cols, 1)), ge(g.size());
        int rc;
        // the last iteration counts the densities (ref src/system.cpp:155-161)
        if (last && dens_flags &&
            vmc_set_density(eng, dens_flags, c.numberOfBins, c.maxRadius) != VMC_OK)
            return fail("densities");
        int rw = 0;
        if (last && c.doResampling) {
            const long long cap = 1LL << 24;
            rw = (int)std::max(1LL, std::min<long long>(W, cap / S));
            rec.assign((size_t)4 * rw * S, 0.f);
            rc = vmc_run_instant(eng, S, equil, printed, B, &s, g.data(), ge.data(), rw, rec.data());
        } else {
            rc = vmc_run(eng, S, equil, printed, B, &s, g.data(), ge.data());
        }
        if (rc != VMC_OK) return fail("run");
        const double time = s.kernel_ms / 1000.0;
        if (iter < numberOfNormalIterations) globalTime += time;
        // Sampler::computeAverages (ref src/sampler.cpp:194-226)
        const double M = s.samples;
        cumT2 += s.kinetic2;
        cumV2 += s.external2;
        cumI2 += s.interaction2;
        av.E = s.energy / M;
        av.T = s.kinetic / M;
        av.V = s.external / M;
        av.I = s.interaction / M;
        av.E2 = s.energy2 / M;
        av.T2 = cumT2 / M;
        av.V2 = cumV2 / M;
        av.I2 = cumI2 / M;
        av.varE = (av.E2 - av.E * av.E) / M;
        av.varT = (av.T2 - av.T * av.T) / M;
        av.varV = (av.V2 - av.V * av.V) / M;
        av.varI = (av.I2 - av.I * av.I) / M;
        av.acc = s.accepted / M;
        if (std::isnan(av.T)) {
            err << "Energy exploded, please decrease the learning rate" << std::endl;
            vmc_engine_destroy(eng);
            return 143;
        }
        std::vector<double> avgG(g.size()), avgGE(g.size());
        for (size_t i = 0; i < g.size(); i++) {
            avgG[i] = g[i] / s.gradient_samples;
            avgGE[i] = ge[i] / s.gradient_samples;
        }
        if (cols > 0) {
            const std::vector<double> upd = opt.update(av.E, avgG, avgGE);
            for (size_t i = 0; i < P.size(); i++) P[i] -= upd[i];
        }
        if (c.printParametersToFile && iter % std::max(1, c.checkpointFreq) == 0) {
            std::ofstream f(c.path + "weights_" + std::to_string(iter) + ".dat");
            f << std::setprecision(10);
            for (int i = 0; i < rows; i++) {
                for (int j = 0; j < cols; j++) f << (j ? " " : "") << P[(size_t)i * cols + j];
                f << "\n";
            }
        }
        if (c.printEnergyToFile) {
            auto app = [&](const char* name, double v) {
                std::ofstream f(c.path + name, iter ? std::ios::app : std::ios::trunc);
                f << v << "\n";
            };
            app("energy.dat", av.E);
            app("kinetic.dat", av.T);
            app("external.dat", av.V);
            app("interaction.dat", av.I);
        }
        if (last && dens_flags) {
            // ref Sampler::print*DensityToFile (src/sampler.cpp:522-581)
            const int nb = c.numberOfBins;
            std::vector<int64_t> radial(nb), spatial((size_t)nb * nb), pair((size_t)nb * nb);
            if (vmc_get_density(eng, radial.data(), spatial.data(), pair.data()) != VMC_OK)
                return fail("densities");
            if (c.computeOneBodyDensity) write_counts(c.path + "onebody.dat", radial, nb, 1);
            if (c.computeOneBodyDensity2) write_counts(c.path + "onebody2.dat", spatial, nb, nb);
            if (c.computeTwoBodyDensity) write_counts(c.path + "twobody.dat", pair, nb, nb);
        }
        if (last) {
            // with resampling: blocking over the instant series written to
            // <n>_0.dat, <n>_0_kin.dat, <n>_0_ext.dat, <n>_0_int.dat, one value per
            // line, the recorded walkers' chains one after another as the
            // reference appends its ranks' files; the three part files are
            // removed afterwards, the energy file is kept (ref
            // src/sampler.cpp:318-430, 466-482, 624-632)
            double mean[4] = {av.E, av.T, av.V, av.I};
            double se[4] = {std::sqrt(av.varE), std::sqrt(av.varT), std::sqrt(av.varV),
                            std::sqrt(av.varI)};
            if (c.doResampling && rw > 0) {
                std::mt19937 pick((uint32_t)(seed ^ 0x1a57a17u));
                const std::string stem =
                    c.path + std::to_string(pick() % 1000000u) + "_0";
                const char* ext[4] = {".dat", "_kin.dat", "_ext.dat", "_int.dat"};
                const size_t plane = (size_t)rw * S;
                for (int q = 0; q < 4; q++) {
                    const std::string name = stem + ext[q];
                    {
                        std::ofstream f(name);
                        for (size_t i = 0; i < plane; i++) f << double(rec[q * plane + i]) << "\n";
                    }
                    std::vector<double> x;
                    {
                        std::ifstream f(name);
                        std::string line;
                        while (std::getline(f, line)) x.push_back(std::strtod(line.c_str(), nullptr));
                    }
                    size_t n = 1;
                    while (n * 2 <= x.size()) n *= 2;
                    x.resize(n);
                    double m, e;
                    if (blocking(x, m, e)) {
                        mean[q] = m;
                        se[q] = e;
                    }
                    if (q > 0) std::remove(name.c_str());
                }
            }
            out << std::endl << std::fixed << std::setprecision(10);
            out << "               Final results" << std::endl;
            out << "==============================================" << std::endl;
            out << " Total energy            : " << mean[0] << " (with STD = " << se[0] << ")" << std::endl;
            out << " Kinetic energy          : " << mean[1] << " (with STD = " << se[1] << ")" << std::endl;
            out << " External energy         : " << mean[2] << " (with STD = " << se[2] << ")" << std::endl;
            out << " Interaction energy      : " << mean[3] << " (with STD = " << se[3] << ")" << std::endl;
            out << " Acceptence Ratio        : " << av.acc << std::endl << std::endl;
            const auto end = std::chrono::system_clock::now();
            const std::chrono::duration<double> el = end - start;
            std::time_t te = std::chrono::system_clock::to_time_t(end);
            out << "Finished computation at " << std::ctime(&te) << "Elapsed time: " << el.count()
                << "s" << std::endl;
            out << std::endl << "Average CPU time before convergence: "
                << globalTime / numberOfNormalIterations << std::endl;
            break;
        }
        printed += 1;
        out << std::fixed << std::setprecision(8) << printed << "  " << av.E << "  "
            << std::sqrt(av.varE) << "  " << av.T << "  " << std::sqrt(av.varT) << "  " << av.V
            << "  " << std::sqrt(av.varV) << "  " << av.I << "  " << std::sqrt(av.varI) << "  "
            << av.acc << "  " << time << std::endl;
        if (checkConvergence) {               // ref src/system.cpp:197-207
            std::rotate(energies.begin(), energies.begin() + 1, energies.end());
            energies.back() = av.E;
            if (std::fabs(energies.front() - energies.back()) < c.tolerance) {
                out << "The system has converged! Let's run one more cycle to collect data"
                    << std::endl;
                numberOfNormalIterations = iter + 1;
                checkConvergence = false;
            }
        }
        if (vmc_set_parameters(eng, params(), rows, cols) != VMC_OK) return fail("parameters");
    }
    vmc_engine_destroy(eng);
    return 0;
}

}  // namespace vmc_host

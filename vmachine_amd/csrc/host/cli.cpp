// cli.cpp -- the C ABI of the host layer (vmc_host.h) and the `vmachine`
// command line: `vmachine [options] input.in`, the drop-in for the
// reference's `mpiexec -n P vmachine input.in` (ref src/main.cpp:3-16).
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>

#include "config.h"
#include "runner.h"
#include "vmc.h"
#include "vmc_host.h"

using vmc_host::Config;
using vmc_host::ConfigError;

struct vmc_host_optimizer {
    vmc_host::Optimizer opt;
};

namespace {

void usage(std::ostream& os)
{
    os << "usage: vmachine [options] input.in\n"
          "  --walkers W       independent chains on the GPU (default: 64 sweeps per walker\n"
          "                    per iteration, numSteps/(64*numParticles*numDimensions),\n"
          "                    clamped to [1, 32768])\n"
          "  --precision P     64 (FP64 walk, default) or 32 (FP32 walk, FP64 rebuilds)\n"
          "  --refresh K       FP32 walk: steps between FP64 Slater-inverse rebuilds (100)\n"
          "  --exact           Langevin proposal with the exact Metropolis-Hastings ratio\n"
          "  --drift-limit A   with --exact: Umrigar drift limit (0 = off)\n"
          "  --no-quirks       Pade-Jastrow ratio over all pairs and d/dgamma without 1/2\n"
          "  --burn-in-sweeps K  unsampled sweeps per walker before the first iteration (100)\n"
          "  --seed S          Philox key (default: the rng line's seed, else random)\n"
          "  --dry-run         print the resolved configuration as JSON and exit\n";
}

void copy_out(const std::string& s, char* buf, int cap)
{
    if (!buf || cap <= 0) return;
    const size_t n = std::min<size_t>(s.size(), (size_t)cap - 1);
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
}

// engine shape of the parameter matrix for a parsed configuration
bool matrix_shape(const Config& c, int& rows, int& cols)
{
    vmc_config ec = {};
    ec.num_particles = c.numParticles;
    ec.num_dimensions = c.numDimensions;
    ec.num_hidden = c.numHiddenUnits;
    ec.num_elements = (int)c.elements.size();
    if (ec.num_elements > 6) return false;
    for (int e = 0; e < ec.num_elements; e++) {
        const std::string& k = c.elements[e];
        ec.kinds[e] = k == "gaussian" ? VMC_GAUSSIAN
                      : k == "slaterDeterminant" ? VMC_SLATER
                      : k == "padeJastrow" ? VMC_PADE
                      : k == "RBMGaussian" ? VMC_RBM_GAUSSIAN
                      : k == "RBMProduct" ? VMC_RBM_PRODUCT : -1;
        if (ec.kinds[e] < 0) return false;
    }
    rows = ec.num_elements;
    cols = vmc_max_parameters(&ec);
    return true;
}

}  // namespace

extern "C" {

int vmc_host_blocking(const double* x, int64_t n, double* mean, double* std_error)
{
    if (!x || n < 2 || !mean || !std_error) return -1;
    std::vector<double> v(x, x + n);
    return vmc_host::blocking(std::move(v), *mean, *std_error) ? 0 : -1;
}

vmc_host_optimizer* vmc_host_optimizer_create(const char* kind, int size, double learning_rate,
                                              double gamma, double monotonic)
{
    if (!kind || size < 0) return nullptr;
    const std::string k = kind;
    if (k != "adam" && k != "sgd" && k != "gd") return nullptr;
    auto* o = new vmc_host_optimizer;
    o->opt.init(k, size, learning_rate, gamma, monotonic);
    return o;
}

int vmc_host_optimizer_batches(const vmc_host_optimizer* o) { return o ? o->opt.batches() : 0; }

void vmc_host_optimizer_update(vmc_host_optimizer* o, double avg_energy, const double* avg_grad,
                               const double* avg_grad_e, double* update)
{
    if (!o || !avg_grad || !avg_grad_e || !update) return;
    const size_t n = o->opt.m.size();
    std::vector<double> g(avg_grad, avg_grad + n), ge(avg_grad_e, avg_grad_e + n);
    const std::vector<double> u = o->opt.update(avg_energy, g, ge);
    std::copy(u.begin(), u.end(), update);
}

void vmc_host_optimizer_destroy(vmc_host_optimizer* o) { delete o; }

int vmc_host_parse(const char* path, char* buf, int cap)
{
    try {
        const Config c = vmc_host::parse_config(path ? path : "");
        copy_out(vmc_host::to_json(c), buf, cap);
        return 0;
    } catch (const ConfigError& ex) {
        copy_out(ex.what(), buf, cap);
        return 143;
    }
}

int vmc_host_initial_weights(const char* path, uint64_t seed, double* out, int rows, int cols)
{
    try {
        const Config c = vmc_host::parse_config(path ? path : "");
        int r = 0, k = 0;
        if (!matrix_shape(c, r, k) || r != rows || k != cols || !out) return 143;
        const std::vector<double> P = vmc_host::initial_weights(c, rows, cols, seed);
        std::copy(P.begin(), P.end(), out);
        return 0;
    } catch (const ConfigError&) {
        return 143;
    }
}

int vmc_host_main(int argc, char** argv)
{
    vmc_host::RunOptions o;
    std::string file;
    bool dry = false;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        const bool has_value = i + 1 < argc;
        try {
            if (a == "--walkers" && has_value) o.walkers = std::stoi(argv[++i]);
            else if (a == "--precision" && has_value) o.precision = std::stoi(argv[++i]);
            else if (a == "--refresh" && has_value) o.refresh = std::stoi(argv[++i]);
            else if (a == "--exact") o.exact = true;
            else if (a == "--drift-limit" && has_value) o.drift_limit = std::stod(argv[++i]);
            else if (a == "--no-quirks") o.reference_quirks = false;
            else if (a == "--burn-in-sweeps" && has_value) o.burn_in_sweeps = std::stoi(argv[++i]);
            else if (a == "--seed" && has_value) {
                o.seed = std::stoull(argv[++i]);
                o.have_seed = true;
            } else if (a == "--dry-run") dry = true;
            else if (a == "-h" || a == "--help") {
                usage(std::cout);
                return 0;
            } else if (!a.empty() && a[0] == '-') {
                std::cerr << "unknown option or missing value: " << a << std::endl;
                usage(std::cerr);
                return 2;
            } else if (file.empty()) file = a;
            else {
                std::cerr << "one configuration file, please" << std::endl;
                return 2;
            }
        } catch (const std::exception&) {
            std::cerr << "bad value for " << a << std::endl;
            return 2;
        }
    }
    if (o.precision != 32 && o.precision != 64) {
        std::cerr << "--precision is 32 or 64" << std::endl;
        return 2;
    }
    Config c;
    try {
        if (!file.empty()) c = vmc_host::parse_config(file);
    } catch (const ConfigError& ex) {
        std::cerr << ex.what() << std::endl;
        return 143;
    }
    if (dry) {
        std::cout << vmc_host::to_json(c) << std::endl;
        return 0;
    }
    return vmc_host::run_simulation(c, o, std::cout, std::cerr);
}

}  // extern "C"

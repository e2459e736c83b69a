/* vmc_host.h -- C ABI of the host layer that drives the engine the way the
 * reference's System/Sampler/Optimization classes drive their plugins.
 *
 * libvmc_host.so holds the input.in parser, the optimisers, the initial-weight
 * presets, automated blocking and the iteration loop behind the `vmachine`
 * executable; it calls the GPU only through vmc.h.  Everything here except
 * vmc_host_main runs without a GPU.
 */
#ifndef VMC_HOST_H
#define VMC_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Automated blocking of a correlated series of length n = 2^d: mean and
 * standard error.  Replaces Blocker::Blocker (ref include/block/blocker.cpp:15-118).
 * Returns 0, or -1 if n is not a power of two (n >= 2). */
int vmc_host_blocking(const double* x, int64_t n, double* mean, double* std_error);

/* Optimiser state for `size` parameters (the row-major elements x max-parameters
 * matrix).  kind: "adam", "sgd" or "gd"; gamma and monotonic are the sgd/gd
 * arguments of the `optimization:` key.  Replaces Optimization::initialize
 * (ref src/Optimization/adam.cpp:12-20, sgd.cpp:15-21, gradientdescent.cpp:15-21).
 * Returns NULL for an unknown kind. */
typedef struct vmc_host_optimizer vmc_host_optimizer;
vmc_host_optimizer* vmc_host_optimizer_create(const char* kind, int size, double learning_rate,
                                              double gamma, double monotonic);
/* Gradient batches per iteration the sampler splits its steps into
 * (ref src/Optimization/adam.h:15: 10 for adam, 1 otherwise). */
int vmc_host_optimizer_batches(const vmc_host_optimizer* o);
/* One update from <E>, <dlnpsi/dtheta> and <E dlnpsi/dtheta>; writes the
 * amount the caller subtracts from the parameters.  Replaces
 * Optimization::getEnergyGradient + updateParameters
 * (ref src/Optimization/optimization.cpp:10-16, adam.cpp:23-38,
 * sgd.cpp:23-29, gradientdescent.cpp:24-30). */
void vmc_host_optimizer_update(vmc_host_optimizer* o, double avg_energy, const double* avg_grad,
                               const double* avg_grad_e, double* update);
void vmc_host_optimizer_destroy(vmc_host_optimizer* o);

/* Parses an input.in file and writes the resolved configuration as one JSON
 * object into buf (NUL-terminated, truncated to cap bytes).  Returns 0, or 143
 * with the reference's error message in buf.  Replaces System::parser
 * (ref src/parser.cpp:25-333). */
int vmc_host_parse(const char* path, char* buf, int cap);

/* Initial parameter matrix (row-major, rows = elements, cols = max
 * parameters) for the input.in at `path` and its initialWeights key.
 * Replaces InitialWeights::setupInitialWeights (ref
 * src/InitialWeights/automatize.cpp:83-101, constant.cpp:11-26,
 * xavier.cpp:10-23, randomuniform.cpp, fromfile.cpp:30-74).
 * Returns 0, or 143 on a configuration error. */
int vmc_host_initial_weights(const char* path, uint64_t seed, double* out, int rows, int cols);

/* The `vmachine` command line (argv[0] is the program name).  Returns the
 * process exit status.  Replaces main (ref src/main.cpp:3-16). */
int vmc_host_main(int argc, char** argv);

#ifdef __cplusplus
}
#endif

#endif /* VMC_HOST_H */

/*
 * vmc.h -- C ABI of the MI355X VMC sampling engine (libvmc_hip.so).
 *
 * The reference (evenmn/VMaChine) has no foreign-function interface: its hot
 * path is a set of C++ plugin classes called once per Metropolis step from
 * System::runMetropolisCycles.  A per-step, per-element virtual call cannot
 * cross a device boundary, so this ABI sits one level up, at the granularity
 * of "advance W walkers through one sampling iteration and return the
 * Sampler's sums".  Each entry point names the reference interface it
 * replaces.  Plain C types only; every call returns 0 on success or a negative
 * VMC_E* code, with a message from vmc_last_error().
 */
#ifndef VMC_H
#define VMC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---- */
#define VMC_OK 0
#define VMC_E_INVALID (-1)   /* bad argument or unsupported configuration */
#define VMC_E_DEVICE (-2)    /* HIP runtime error / no device */
#define VMC_E_STATE (-3)     /* call out of order (e.g. run before parameters) */

/* ---- wave-function elements (ref src/parser.cpp:263-286 keys) ---- */
#define VMC_GAUSSIAN 1       /* waveFunctionElement: gaussian          (ref WaveFunctions/gaussian.cpp) */
#define VMC_SLATER 2         /* waveFunctionElement: slaterDeterminant (ref WaveFunctions/slaterdeterminant.cpp) */
#define VMC_PADE 3           /* waveFunctionElement: padeJastrow       (ref WaveFunctions/padejastrow.cpp) */
#define VMC_RBM_GAUSSIAN 4   /* waveFunctionElement: RBMGaussian       (ref WaveFunctions/rbmgaussian.cpp) */
#define VMC_RBM_PRODUCT 5    /* waveFunctionElement: RBMProduct        (ref WaveFunctions/rbmproduct.cpp) */

/* ---- sampling (ref src/parser.cpp:210-222) ---- */
#define VMC_BRUTE_FORCE 0         /* sampling: bruteForce (ref Metropolis/bruteforce.cpp:26-52) */
#define VMC_IMPORTANCE 1          /* sampling: importanceSampling, the reference's Green's-function ratio
                                     (ref Metropolis/importancesampling.cpp:43-47, 68-81) */
#define VMC_IMPORTANCE_EXACT 2    /* Langevin proposal with the exact Metropolis-Hastings ratio */
#define VMC_LEARNED 3             /* learned whole-electron proposal (vmc_lp_*), exact MH ratio */

/* ---- reference behaviours that change results (set bit = reproduce) ---- */
#define VMC_Q2_PADE_RATIO_UPPER (1u << 2)  /* ref padejastrow.cpp:214 */
#define VMC_Q3_PADE_GAMMA_HALF (1u << 3)   /* ref padejastrow.cpp:114 */
#define VMC_Q6_RBM_W_UNSCALED (1u << 6)   /* ref rbmproduct.cpp:77-83: dlnpsi/dW = x n^T, no 1/sigma_p^2 */
#define VMC_Q13_RBMG_PSI_RATIO (1u << 13) /* ref rbmgaussian.cpp:44-45: ratio psi'/psi, not |psi'/psi|^2 */
#define VMC_QUIRKS_REFERENCE                                                   \
    (VMC_Q2_PADE_RATIO_UPPER | VMC_Q3_PADE_GAMMA_HALF | VMC_Q6_RBM_W_UNSCALED | \
     VMC_Q13_RBMG_PSI_RATIO)

typedef struct vmc_config {
    int num_particles;      /* numParticles: 1..110 (with a Slater determinant the closed
                               shells 2, 6, 12, 20, 30, 42, 56, 72, 90, 110; the learned
                               proposal and batches up to 56) */
    int num_dimensions;     /* numDimensions (the engine supports 2) */
    double omega;           /* omega */
    int interaction;        /* interactionStyle: 0 noInteraction, 1 coulomb */
    int num_elements;       /* waveFunctionElement lines, in order: row e of the parameter matrix */
    int kinds[6];
    int num_hidden;         /* numHiddenNodes */
    double sigma;           /* RBM width (1/sqrt(omega) once omega is set) */
    int sorting;            /* sorting */
    int sampling;           /* VMC_BRUTE_FORCE / VMC_IMPORTANCE / VMC_IMPORTANCE_EXACT */
    double step_length;     /* stepLength */
    unsigned quirks;        /* VMC_Q* bits */
    int precision;          /* 64: FP64 walk; 32: FP32 walk with FP64 inverse rebuilds */
    int refresh_interval;   /* steps between FP64 Slater-inverse rebuilds (0: launch start only) */
    int num_walkers;        /* independent chains on the device */
    double drift_limit;     /* VMC_IMPORTANCE_EXACT only: a > 0 replaces the drift F by
                               2F / (1 + sqrt(1 + a F^2 dt / 2)) in the proposal and in its
                               density (Umrigar, Nightingale, Runge, J. Chem. Phys. 99, 2865
                               (1993)), so walkers next to a node of the Slater determinant
                               are not frozen by its divergent drift; 0 = textbook Langevin */
    int rbm_matrix;         /* RBM-product hidden-layer contractions (pre-activation of every
                               proposal, d ln psi/dx after every accepted move):
                               VMC_RBM_AUTO, VMC_RBM_LANES, VMC_RBM_MFMA_F32, VMC_RBM_MFMA_BF16 */
} vmc_config;

/* rbm_matrix: AUTO = FP32 matrix cores where they apply (FP32 walk, RBM-product,
 * >= 4 lanes per walker), lanes otherwise; LANES = FMA chains on the walker's
 * lanes; MFMA_F32 = 16x16x4 FP32 matrix-core tiles; MFMA_BF16 = 16x16x32 BF16
 * tiles (inputs rounded to BF16, FP32 accumulation).  The MFMA values are
 * rejected where they do not apply. */
#define VMC_RBM_AUTO 0
#define VMC_RBM_LANES 1
#define VMC_RBM_MFMA_F32 2
#define VMC_RBM_MFMA_BF16 3

/* Sums over all walkers of one sampling iteration (ref src/sampler.cpp:52-90,
 * the quantities Sampler::computeTotals reduces, ref src/sampler.cpp:111-187). */
typedef struct vmc_sums {
    double energy, energy2;
    double kinetic, external, interaction;
    double kinetic2, external2, interaction2;
    double accepted;          /* accepted moves among the sampled steps */
    double samples;           /* walkers x sampled steps */
    double gradient_samples;  /* walkers x gradient-window length (divisor of <G>, <GE>) */
    double kernel_ms;         /* walk-kernel time, HIP events on the engine stream */
    double nonfinite;         /* sampled non-finite local energies (all walkers) */
} vmc_sums;

typedef struct vmc_engine vmc_engine;

/* Last error message of the calling thread ("" if none). */
const char* vmc_last_error(void);

/* Parameter-matrix shape: ref WaveFunction::getNumberOfParameters and
 * System::setMaxParameters (ref src/elementoperation.cpp:173-186). Host only. */
int vmc_num_parameters(const vmc_config* cfg, int element);
int vmc_max_parameters(const vmc_config* cfg);

/* Number of HIP devices visible (0 without a GPU); never fails. */
int vmc_device_count(void);

/* Creates the engine on the current HIP device.  Replaces the construction
 * and System::initializeSystem wiring of the plugins (ref src/system.cpp:50-73). */
int vmc_engine_create(const vmc_config* cfg, vmc_engine** out);
void vmc_engine_destroy(vmc_engine* e);

/* Parameter matrix, row-major [rows = num_elements][cols = max parameters].
 * Replaces System::updateAllParameters (ref src/elementoperation.cpp:68-73). */
int vmc_set_parameters(vmc_engine* e, const double* params, int rows, int cols);

/* Draws every walker's positions x ~ N(0, (1/omega)^2) from Philox(seed) and
 * resets the step counter.  Replaces RandomNormal::setupInitialState
 * (ref src/InitialStates/randomnormal.cpp:10-25) and
 * ImportanceSampling::initialize (ref src/Metropolis/importancesampling.cpp:11-32). */
int vmc_init_walkers(vmc_engine* e, uint64_t seed);

/* Walker positions, row-major [walkers][num_particles*num_dimensions], FP64. */
int vmc_set_positions(vmc_engine* e, const double* positions, int walkers, int dof);
int vmc_get_positions(vmc_engine* e, double* positions, int walkers, int dof);
/* The importance sampler's stored drift vector per walker (same layout). */
int vmc_set_quantum_force(vmc_engine* e, const double* qf, int walkers, int dof);
int vmc_get_quantum_force(vmc_engine* e, double* qf, int walkers, int dof);
/* Random-number state: the Philox key (seed) and the step counter of the next
 * step (all walkers share both; walker w draws from counter words (step, w)).
 * Replaces RandomNumberGenerator::setSeed (ref src/RNG/rng.h:10). */
int vmc_set_random_state(vmc_engine* e, uint64_t seed, uint64_t counter);
int vmc_get_random_state(vmc_engine* e, uint64_t* seed, uint64_t* counter);

/* One sampling iteration: every walker runs `equilibration` burn-in steps and
 * `steps` sampled steps; the gradient sums cover the steps t with
 * S*(iteration % batches) < t < S*(iteration % batches + 1), S = steps/batches,
 * t counted from the first burn-in step.  grad and grad_e (row-major,
 * num_elements x max parameters) receive the window sums of dlnpsi/dtheta and
 * E_L*dlnpsi/dtheta.  Replaces System::runMetropolisCycles +
 * Sampler::sample/computeTotals (ref src/system.cpp:147-164,
 * src/sampler.cpp:52-90, 111-187). */
int vmc_run(vmc_engine* e, int steps, int equilibration, int iteration, int batches,
            vmc_sums* sums, double* grad, double* grad_e);

/* As vmc_run, and also returns the sampled local energies of the first
 * `record_walkers` walkers, row-major [record_walkers][steps] (FP32).
 * Replaces Sampler::printInstantValuesToFile (ref src/sampler.cpp:624-632). */
int vmc_run_recorded(vmc_engine* e, int steps, int equilibration, int iteration, int batches,
                     vmc_sums* sums, double* grad, double* grad_e, int record_walkers,
                     float* energies);

/* As vmc_run_recorded, returning instead the means of `block` consecutive
 * sampled local energies of each of the first `record_walkers` walkers,
 * row-major [record_walkers][steps / block] (FP64; block >= 2 divides steps).
 * The reduction runs on the device, so long series need no per-step copy to
 * the host.  The block means are the input of the reference's blocking
 * resampler, which reads every instant energy from the file
 * Sampler::printInstantValuesToFile writes (ref src/sampler.cpp:624-632,
 * include/block/blocker.cpp:15-33). */
int vmc_run_blocked(vmc_engine* e, int steps, int equilibration, int iteration, int batches,
                    vmc_sums* sums, double* grad, double* grad_e, int record_walkers, int block,
                    double* means);

/* As vmc_run_recorded with all four instant values of the first
 * `record_walkers` walkers: out [4][record_walkers][steps] (FP32) = local
 * energy, kinetic, external, interaction, the four series
 * Sampler::printInstantValuesToFile writes (ref src/sampler.cpp:624-632). */
int vmc_run_instant(vmc_engine* e, int steps, int equilibration, int iteration, int batches,
                    vmc_sums* sums, double* grad, double* grad_e, int record_walkers, float* out);

/* Electron densities, counted by the walk after every sampled step of every
 * walker while enabled (ref Sampler::computeOneBodyDensity,
 * computeOneBodyDensity2, computeTwoBodyDensity, src/sampler.cpp:639-692),
 * with dr = max_radius / bins and r_i = |x_i|:
 *   VMC_DENSITY_RADIAL: radial [bins], r < max_radius at int(r/dr);
 *   VMC_DENSITY_SPATIAL: spatial [bins][bins] (row = y), cell
 *     (int(y/2dr) + bins/2, int(x/2dr) + bins/2) with C truncation, both
 *     |int(.)| < bins/2;
 *   VMC_DENSITY_TWO_BODY: two_body [bins][bins], every pair i < j at
 *     (int(r_i/dr) + 1, int(r_j/dr) + 1), both < bins.
 * vmc_set_density clears the counts (flags 0 switches them off);
 * vmc_get_density copies them out (any pointer may be null). */
#define VMC_DENSITY_RADIAL 1
#define VMC_DENSITY_SPATIAL 2
#define VMC_DENSITY_TWO_BODY 4
#define VMC_DENSITY_MAX_BINS 4096
int vmc_set_density(vmc_engine* e, int flags, int bins, double max_radius);
int vmc_get_density(vmc_engine* e, int64_t* radial, int64_t* spatial, int64_t* two_body);

/* Parity entry point.  For each of `walkers` configurations (positions
 * row-major [walkers][dof]) the engine rebuilds its state, evaluates the local
 * energy and its parts, dlnpsi/dtheta and the drift, then forces the move of
 * coordinate moves[w] to new_values[w] through its incremental update path and
 * evaluates again.  out_before / out_after: [walkers][VMC_EVAL_WIDTH] FP64.
 * Row layout: see VMC_EVAL_*.  Replaces the per-element calls
 * evaluateRatio / computeGradient / computeLaplacian / computeParameterGradient
 * and System::getKineticEnergy (ref src/WaveFunctions/wavefunction.h:29-32,
 * src/elementoperation.cpp:80-123). */
#define VMC_EVAL_E 0
#define VMC_EVAL_T 1
#define VMC_EVAL_VEXT 2
#define VMC_EVAL_VINT 3
#define VMC_EVAL_GALPHA 4
#define VMC_EVAL_GGAMMA 5
#define VMC_EVAL_RATIO_SLATER 6   /* |det'/det|^2 of the forced move (after only) */
#define VMC_EVAL_RATIO_PADE 7     /* Pade ratio of the forced move (after only) */
#define VMC_EVAL_RATIO_GAUSS 8    /* Gaussian ratio of the forced move (after only) */
#define VMC_EVAL_DRIFT 9          /* then dof drift components 2 dlnpsi/dx_k */
#define VMC_EVAL_WIDTH(dof) (9 + (dof))
int vmc_evaluate(vmc_engine* e, const double* positions, int walkers, const int* moves,
                 const double* new_values, double* out_before, double* out_after);

/* As vmc_evaluate, and per element (any kind, including the RBM elements):
 * ratios [walkers][num_elements] = the element's evaluateRatio of the forced
 * move, grads_before / grads_after [walkers][num_elements x max parameters] =
 * the parameter-gradient matrix (row e = element e's computeParameterGradient)
 * before and after it.  Replaces WaveFunction::evaluateRatio and
 * computeParameterGradient of each element (ref src/WaveFunctions/wavefunction.h:29-32,
 * src/elementoperation.cpp:80-87, 111-123). */
int vmc_evaluate_elements(vmc_engine* e, const double* positions, int walkers, const int* moves,
                          const double* new_values, double* out_before, double* out_after,
                          double* ratios, double* grads_before, double* grads_after);

/* ---- Learned proposal (no counterpart in the reference) ----
 * With sampling = VMC_LEARNED a step moves one whole electron i (drawn
 * uniformly) by dx = mu + sigma * xi, xi ~ N(0, 1), where (mu, sigma) come from
 * a per-electron MLP: 18 inputs (the electron's coordinates and the relative
 * coordinates of its 8 nearest electrons, units of 1/sqrt(omega)), two tanh
 * layers of 64, 4 outputs (mean, log-scale).  Acceptance is Metropolis-
 * Hastings with the proposal ratio q(x|x')/q(x'|x) and the exact |psi'/psi|^2
 * (the Pade ratio over all pairs whatever the quirk bits), so |psi|^2 stays
 * the target.  The walk evaluates the network with bfloat16 weights and
 * arithmetic in the walk precision.  Not available with RBM elements. */
#define VMC_LP_NPARAM 5636   /* W1[64][18], b1[64], W2[64][64], b2[64], W3[4][64], b3[4] */
typedef struct vmc_lp_hyper {
    double sigma0;   /* move scale of the untrained (zero-head) network, units of 1/sqrt(omega) */
    double mu_max;   /* |mean| < mu_max / sqrt(omega) */
    double ls_max;   /* |log(scale / sigma0)| < ls_max */
} vmc_lp_hyper;

/* Network weights (FP32 master copy; the walk uses them rounded to bfloat16)
 * and output bounds; resets the training state. */
int vmc_lp_set(vmc_engine* e, const float* theta, int n, const vmc_lp_hyper* hp);
int vmc_lp_get(vmc_engine* e, float* theta, int n);

/* Parity entry point of the training kernels: for each configuration
 * (positions [walkers][dof]), electron and normals [walkers][2], the proposal,
 * J = A |dx|^2 omega with A = min(1, |psi'/psi|^2 q(x|x')/q(x'|x)), and dJ/dtheta.
 * per_sample [walkers][5] = J, A, log|psi'/psi|^2, x', y'; grad [VMC_LP_NPARAM]
 * = the sum over the configurations. */
int vmc_lp_gradient(vmc_engine* e, const double* positions, int walkers, const int* electrons,
                    const double* normals, double* per_sample, double* grad);

/* Online training on the engine's own walkers: per iteration one sample per
 * walker (electron and normals from Philox), gradient ascent on the mean J
 * (ADAM, learning_rate), then walk_steps steps of the engine's sampling.
 * history [iterations][3] (may be null) = mean J, mean A, mean |dx|^2 omega. */
int vmc_lp_train(vmc_engine* e, int iterations, int walk_steps, double learning_rate,
                 double* history);

/* ---- Optimisation on the device ----
 * `iterations` optimisation iterations of the engine's system without a return
 * to the host in between: per iteration equilibration + steps steps of every
 * walker, the sums, the reference's energy gradient g = 2(<G E> - <E><G>) and
 * optimiser step on every entry of the parameter matrix, and the updated
 * parameters written where the next iteration's walk reads them.  With
 * use_graph one iteration is captured as a hipGraph and replayed.  Replaces
 * the iteration loop of System::runSimulation (ref src/system.cpp:81-139:
 * runMetropolisCycles, Sampler::computeTotals, Optimization::updateParameters;
 * ref src/Optimization/optimization.cpp:10-16, adam.cpp:12-38, sgd.cpp:15-29,
 * gradientdescent.cpp:15-30).  The optimiser starts as
 * Optimization::initialize leaves it (moments of one, step 0) and the
 * gradient window of iteration i is i % batches.
 * params [num_elements][max_parameters]: in, the starting matrix; out, the
 * final one, which the engine keeps (as after vmc_set_parameters).
 * history [iterations][VMC_OPT_OUT] (may be null): the iteration's 14 sums
 * over walkers and sampled steps (E, E^2, T, Vext, Vint, T^2, Vext^2, Vint^2,
 * accepted, G_alpha, G_gamma, G_alpha E, G_gamma E over the gradient window,
 * non-finite E_L), then alpha and gamma after its step.  kernel_ms (may be
 * null): device time of all iterations. */
typedef struct vmc_optimizer_config {
    int kind;               /* VMC_OPT_ADAM, VMC_OPT_SGD, VMC_OPT_GD */
    double learning_rate;   /* learningRate */
    double gamma;           /* SGD / GD momentum (the optimization key's gamma) */
    double monotonic;       /* SGD / GD step exponent */
    int batches;            /* gradient batches: 10 for ADAM in the reference, 1 otherwise */
} vmc_optimizer_config;
#define VMC_OPT_ADAM 0
#define VMC_OPT_SGD 1
#define VMC_OPT_GD 2
#define VMC_OPT_OUT 16
int vmc_optimize(vmc_engine* e, double* params, int iterations, int steps, int equilibration,
                 const vmc_optimizer_config* opt, int use_graph, double* history,
                 double* kernel_ms);

/* As vmc_optimize with use_graph = 0, allocating no device memory: no
 * counterpart in the reference.  For callers that cannot allocate once the
 * engine exists (tight optimisation loops, harnesses that forbid hipMalloc).
 * The loop's device state (parameter matrix, optimiser moments, iteration
 * state, history staging) lives in one file-scope array per GPU, set up on the
 * first call on that GPU and held until the process ends, so a call accepts at
 * most VMC_STATIC_MAX_PARAMS parameter-matrix entries (num_elements x max
 * parameters) and VMC_STATIC_MAX_ITERATIONS iterations (VMC_E_INVALID
 * beyond).  Every call on a GPU reuses that GPU's array: calls that overlap on
 * one GPU (two threads, two engines) race and must be serialised by the
 * caller.  The engine's own buffers come from vmc_engine_create. */
#define VMC_STATIC_MAX_PARAMS 8192
#define VMC_STATIC_MAX_ITERATIONS 1024
int vmc_optimize_static(vmc_engine* e, double* params, int iterations, int steps,
                        int equilibration, const vmc_optimizer_config* opt, double* history,
                        double* kernel_ms);

/* ---- Batches: several quantum dots walked side by side ----
 * Systems of different particle numbers and frequencies (Slater / Gaussian /
 * Pade-Jastrow elements, the same for all) are walked by one kernel each, the
 * kernels launched together on streams of their own (costliest system first)
 * so that the GPU's workgroup dispatcher fills the machine from all of them.
 * vmc_batch_run runs whole optimisation iterations on the device: the walk,
 * per-system sums, the reference's ADAM step on (alpha, gamma) of every system
 * (ref src/Optimization/adam.cpp:12-38) and the advance of the random counter
 * and gradient window; with use_graph the iteration is captured once as a
 * hipGraph and replayed.  No counterpart in the reference (it runs one system
 * per process). */
#define VMC_BATCH_MAX_SYSTEMS 8
typedef struct vmc_batch vmc_batch;
typedef struct vmc_batch_system {
    int num_particles;
    double omega;
    int num_walkers;
    double step_length;   /* stepLength of this system */
    double alpha, gamma;  /* initial Gaussian and Pade parameters */
} vmc_batch_system;
typedef struct vmc_batch_config {
    int num_systems;
    const vmc_batch_system* systems;
    int num_elements;     /* element kinds shared by the systems (VMC_SLATER, VMC_GAUSSIAN, VMC_PADE) */
    int kinds[6];
    int interaction;
    int sampling;         /* VMC_BRUTE_FORCE, VMC_IMPORTANCE or VMC_IMPORTANCE_EXACT */
    unsigned quirks;
    int precision;        /* 32 or 64 */
    int refresh_interval;
    double drift_limit;
} vmc_batch_config;
#define VMC_BATCH_OUT 16  /* per system and iteration: the 14 VMC_EVAL-order sums, then alpha, gamma */

int vmc_batch_create(const vmc_batch_config* cfg, vmc_batch** out);
void vmc_batch_destroy(vmc_batch* b);
/* System s draws from seed + s (its walkers retrace a single-system engine
 * created with that seed). */
int vmc_batch_init_walkers(vmc_batch* b, uint64_t seed);
int vmc_batch_set_positions(vmc_batch* b, int system, const double* positions, int walkers, int dof);
int vmc_batch_get_positions(vmc_batch* b, int system, double* positions, int walkers, int dof);
/* alpha_gamma [num_systems][2]: the Gaussian alpha and Pade gamma of every
 * system after the last vmc_batch_run (the device optimiser's values). */
int vmc_batch_get_parameters(vmc_batch* b, double* alpha_gamma);
/* Kernel shape of one system: lanes per walker (a power of two >= the
 * orbitals per spin, at most twice the smallest such). */
int vmc_batch_set_shape(vmc_batch* b, int system, int lanes_per_walker);
/* Times each system's candidate shapes on its own walkers (positions and
 * random state are restored) and keeps the fastest; lanes_out [num_systems]
 * receives the choices, ns_per_step_out [num_systems] (may be null) their time
 * per walker-step, which then orders the launches.  Where the configuration
 * allows the persistent launch, it then times the whole batch side by side and
 * as one persistent launch over steps + min(refresh_interval, 4 steps) steps
 * (a trial that crosses a refresh boundary) and keeps the faster
 * (vmc_batch_get_launch). */
int vmc_batch_autotune(vmc_batch* b, int steps, int* lanes_out, double* ns_per_step_out);
/* How vmc_batch_run launches an iteration's walk.  VMC_BATCH_SIDE_BY_SIDE:
 * per refresh segment one kernel per system, on streams of their own.
 * VMC_BATCH_PERSISTENT: one launch of one persistent kernel per iteration,
 * whose waves (as many as fit on the GPU at once) take tiles of 64/K walkers
 * of one system from a device counter, costliest system's tiles first, and
 * walk each through every refresh segment (FP32 Slater x Gaussian x
 * Pade-Jastrow with Coulomb under exact Langevin; VMC_E_INVALID otherwise).
 * Both give the same chains. */
#define VMC_BATCH_SIDE_BY_SIDE 0
#define VMC_BATCH_PERSISTENT 1
int vmc_batch_set_launch(vmc_batch* b, int launch);
int vmc_batch_get_launch(vmc_batch* b, int* launch);
/* `iterations` optimisation iterations of equilibration + steps steps each.
 * out [iterations][num_systems][VMC_BATCH_OUT]; kernel_ms (may be null) = the
 * device time of all iterations.  learning_rate 0 keeps the parameters. */
int vmc_batch_run(vmc_batch* b, int iterations, int steps, int equilibration, int batches,
                  double learning_rate, int use_graph, double* out, double* kernel_ms);

/* Parity entry point for the random-number protocol: the draws of steps
 * counter0 .. counter0+n-1 of one walker exactly as the walk computes them in
 * the given precision (32 or 64): coordinate index in [0, dof), standard
 * normal, move uniform and accept uniform.  Replaces the calls
 * RandomNumberGenerator::nextInt/nextGaussian/nextDouble of one Metropolis step
 * (ref src/Metropolis/importancesampling.cpp:51-75, src/RNG/rng.h:11-13). */
int vmc_draws(int precision, uint64_t seed, uint32_t walker, uint64_t counter0, int n, int dof,
              int* coords, double* normals, double* u_move, double* u_accept);

#ifdef __cplusplus
}
#endif
#endif

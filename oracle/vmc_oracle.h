/*
 * vmc_oracle.h -- FP64 CPU restatement of VMaChine's sampling hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (vmachine_amd/, include/,
 * the vmachine CLI) links or calls this code.  Only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load it, and only
 * as the checker.
 *
 * Every function restates a reference function; the reference file:line is
 * cited next to it in vmc_oracle.c.  Pins (what checks each function against
 * something other than itself) are listed in tests/test_oracle_*.py.
 */
#ifndef VMC_ORACLE_H
#define VMC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VMCO_MAXE 6

/* element kinds, in the reference's vocabulary */
enum {
    VMCO_GAUSSIAN = 1,      /* ref src/WaveFunctions/gaussian.cpp          */
    VMCO_SLATER = 2,        /* ref src/WaveFunctions/slaterdeterminant.cpp */
    VMCO_PADE = 3,          /* ref src/WaveFunctions/padejastrow.cpp       */
    VMCO_RBM_GAUSSIAN = 4,  /* ref src/WaveFunctions/rbmgaussian.cpp       */
    VMCO_RBM_PRODUCT = 5    /* ref src/WaveFunctions/rbmproduct.cpp        */
};

/* sampling methods */
enum {
    VMCO_BRUTE_FORCE = 0,       /* ref src/Metropolis/bruteforce.cpp:26-52            */
    VMCO_IMPORTANCE_COMPAT = 1, /* ref src/Metropolis/importancesampling.cpp:49-81    */
    VMCO_IMPORTANCE_EXACT = 2   /* Langevin with the exact MH Green's-function ratio   */
};

/* reference quirks that can be switched off (set bit = reproduce the reference) */
enum {
    VMCO_Q2_PADE_RATIO_UPPER = 1u << 2,  /* padejastrow.cpp:214: only pairs j > i_p in the ratio */
    VMCO_Q3_PADE_GAMMA_HALF = 1u << 3,   /* padejastrow.cpp:114: extra 1/2 in dlnpsi/dgamma     */
    VMCO_Q6_RBM_W_NO_SIGMA = 1u << 6,    /* rbmproduct.cpp:77-83: dlnpsi/dW lacks 1/sigma_p^2    */
    VMCO_Q13_RBMG_SQRT_RATIO = 1u << 13, /* rbmgaussian.cpp:44-45: ratio uses 2 sigma^2            */
    VMCO_QUIRKS_ALL = (1u << 2) | (1u << 3) | (1u << 6) | (1u << 13)
};

typedef struct vmco_config {
    int N;               /* particles */
    int D;               /* dimensions (2 or 3) */
    double omega;        /* oscillator frequency */
    int interaction;     /* 0 none, 1 Coulomb */
    int num_elements;
    int kinds[VMCO_MAXE];
    int num_hidden;      /* RBM hidden units */
    double sigma;        /* System::getWidth() (1/sqrt(omega) once omega is set) */
    int sorting;         /* RBM input sorting (reference default true) */
    int sampling;        /* VMCO_BRUTE_FORCE / VMCO_IMPORTANCE_* */
    double step_length;  /* dt (importance) or brute-force step */
    unsigned quirks;     /* VMCO_Q* bits */
    int basis_scaled;    /* 0: reference Hermite H_n(x); 1: H_n(sqrt(omega) x) */
    double drift_limit;  /* exact Langevin: F -> 2F/(1 + sqrt(1 + a F^2 dt/2)); 0 = none */
} vmco_config;

typedef struct vmco_walker vmco_walker;

/* number of parameters of each element (ref getNumberOfParameters) and the
 * maximum over elements (ref System::setMaxParameters, elementoperation.cpp:173-186) */
int vmco_num_parameters(const vmco_config* c, int element);
int vmco_max_parameters(const vmco_config* c);

vmco_walker* vmco_walker_new(const vmco_config* c);
void vmco_walker_free(vmco_walker* w);
/* params: num_elements x maxpar, row-major */
void vmco_set_parameters(vmco_walker* w, const double* params);
/* initializeArrays + setArrays for every element (randomnormal.cpp:21-24) */
void vmco_init_positions(vmco_walker* w, const double* positions);
void vmco_get_positions(const vmco_walker* w, double* out);

/* System::updateAllArrays (elementoperation.cpp:38-47): move coordinate k to newval */
void vmco_propose(vmco_walker* w, int k, double newval);
void vmco_reject(vmco_walker* w);  /* System::resetAllArrays */
void vmco_accept(vmco_walker* w);

double vmco_element_ratio(vmco_walker* w, int e);
double vmco_element_gradient(vmco_walker* w, int e, int k);
double vmco_element_laplacian(vmco_walker* w, int e);
void vmco_element_parameter_gradient(vmco_walker* w, int e, double* out /* maxpar */);
double vmco_probability_ratio(vmco_walker* w);
double vmco_kinetic_energy(vmco_walker* w);
double vmco_external_energy(vmco_walker* w);
double vmco_interaction_energy(vmco_walker* w);

/* quantum-force vector of the importance sampler (ND) */
void vmco_get_quantum_force(const vmco_walker* w, double* out);
void vmco_set_quantum_force(vmco_walker* w, const double* qf);
void vmco_reset_quantum_force(vmco_walker* w);

/* one Metropolis step with explicit random numbers:
 *   k = coordinate, z = standard normal (importance), u_move = U(0,1) (brute force),
 *   u_acc = U(0,1) acceptance draw.  Returns 1 if accepted. */
int vmco_step(vmco_walker* w, int k, double z, double u_move, double u_acc);

/* ---- Sampler accumulators (sampler.cpp:52-90) ---- */
enum {
    VMCO_ACC_E = 0, VMCO_ACC_E2, VMCO_ACC_T, VMCO_ACC_VEXT, VMCO_ACC_VINT,
    VMCO_ACC_T2, VMCO_ACC_VEXT2, VMCO_ACC_VINT2, VMCO_ACC_ACCEPT, VMCO_ACC_NSAMPLES,
    VMCO_NACC
};

/* One walker's iteration: equil + steps Metropolis steps with the engine's
 * Philox protocol for (seed, walker) starting at step counter `counter`.
 * Samples after equil; gradient window = steps with index i in
 * (S*(iter%B), S*(iter%B+1)), S = steps/B, exactly as sampler.cpp:82-86 does on
 * one rank.  acc[VMCO_NACC] and grad/gradE[num_elements*maxpar] are ADDED to.
 * If energies != NULL, the sampled local energies are written there.
 * Returns the next step counter. */
uint64_t vmco_run(vmco_walker* w, uint64_t seed, uint32_t walker, uint64_t counter,
                  int steps, int equil, int iter, int batches,
                  double* acc, double* grad, double* gradE, double* energies);

/* many walkers, multithreaded over walkers (pthreads); positions[W*ND] updated
 * in place, counters[W] likewise.  qf[W*ND] holds the importance sampler's
 * quantum-force vectors (may be NULL: then freshly initialised). */
void vmco_run_many(const vmco_config* c, const double* params, int W, double* positions,
                   double* qf, uint64_t seed, uint64_t* counters, int steps, int equil,
                   int iter, int batches, int threads, double* acc, double* grad,
                   double* gradE);

/* ---- Engine random-number protocol ---- */
void vmco_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* the four draws of step `counter` of `walker`: coordinate index, normal, move uniform, accept uniform */
void vmco_step_draws(uint64_t seed, uint32_t walker, uint64_t counter, int nd,
                     int* k, double* z, double* u_move, double* u_acc);
/* initial positions x ~ N(0, (1/omega)^2) per coordinate (randomnormal.cpp:17) */
void vmco_initial_positions(uint64_t seed, uint32_t walker, int nd, double omega, double* out);

/* ---- Optimisers (Optimization/{optimization,adam,sgd,gradientdescent}.cpp) ---- */
typedef struct vmco_optimizer {
    int kind;            /* 0 ADAM, 1 SGD, 2 GD */
    int size;            /* rows*cols */
    int step;
    double eta, gamma, monotonic_exp;
    double beta1, beta2, epsilon;
    double m[4096], v[4096];
} vmco_optimizer;
void vmco_optimizer_init(vmco_optimizer* o, int kind, int size, double eta, double gamma,
                         double monotonic_exp);
/* g = 2(<G E> - <E><G>); writes the update that the caller subtracts from P */
void vmco_optimizer_update(vmco_optimizer* o, double avgE, const double* avgG,
                           const double* avgGE, double* update);

/* ---- Blocking (include/block/blocker.cpp) ---- */
/* x is overwritten (as the reference does).  Returns 0 on success, -1 if n is
 * not a power of two. */
int vmco_blocking(double* x, uint64_t n, double* mean, double* stderr_, uint64_t* nK);
double vmco_chi2_quantile95(int dof);

/* ---- Hermite basis (Basis/hermite.cpp, Basis/basis.cpp) ---- */
/* ---- learned Metropolis proposal (no reference counterpart) ----
 * A per-electron MLP maps features of the moved electron (its own coordinates
 * and the relative coordinates of its 8 nearest electrons, in units of
 * 1/sqrt(omega)) through two tanh layers of width 64 to a Gaussian move mean
 * and a diagonal scale.  The proposal moves one whole electron; acceptance is
 * Metropolis-Hastings with the proposal ratio.  The network is trained to
 * maximise the expected squared jump distance E[A |dx|^2].  Pins: finite
 * differences of the gradient, the zero-network closed form (a symmetric
 * Gaussian random walk, q ratio 1). */
#define VMCO_LP_IN 18
#define VMCO_LP_HID 64
#define VMCO_LP_OUT 4
#define VMCO_LP_NPARAM (VMCO_LP_HID * VMCO_LP_IN + VMCO_LP_HID + VMCO_LP_HID * VMCO_LP_HID + \
                        VMCO_LP_HID + VMCO_LP_OUT * VMCO_LP_HID + VMCO_LP_OUT)
/* parameter offsets: W1[u][m], b1[u], W2[u][c], b2[u], W3[r][u], b3[r] */
#define VMCO_LP_W1 0
#define VMCO_LP_B1 (VMCO_LP_W1 + VMCO_LP_HID * VMCO_LP_IN)
#define VMCO_LP_W2 (VMCO_LP_B1 + VMCO_LP_HID)
#define VMCO_LP_B2 (VMCO_LP_W2 + VMCO_LP_HID * VMCO_LP_HID)
#define VMCO_LP_W3 (VMCO_LP_B2 + VMCO_LP_HID)
#define VMCO_LP_B3 (VMCO_LP_W3 + VMCO_LP_OUT * VMCO_LP_HID)

typedef struct vmco_lp_hyper {
    double sigma0;   /* scale of the untrained proposal, in units of 1/sqrt(omega) */
    double mu_max;   /* |mean| < mu_max / sqrt(omega) */
    double ls_max;   /* |log(scale / sigma0)| < ls_max */
} vmco_lp_hyper;

/* features of electron i placed at (xi, yi), all other electrons at x; returns
 * the number of neighbours used (min(8, N-1)) */
int vmco_lp_features(const vmco_config* c, const double* x, int i, double xi, double yi,
                     double* f /* 18 */);
/* network forward: h1, h2 (64 each, may be NULL), o (4) */
void vmco_lp_forward(const double* theta, const double* f, double* h1, double* h2, double* o);
/* move mean and scale of the head outputs o */
void vmco_lp_head(const vmco_lp_hyper* hp, double omega, const double* o, double* mu, double* sig);
/* one learned-proposal step of the walker: electron i, normals xi0, xi1,
 * acceptance uniform ua.  The walker's configuration must use exact ratios
 * (quirks 0).  Returns 1 if accepted. */
int vmco_lp_step(vmco_walker* w, const double* theta, const vmco_lp_hyper* hp, int i, double xi0,
                 double xi1, double ua);
/* one training sample at the walker's current configuration: J = A |dx|^2 omega
 * and, added into grad (VMCO_LP_NPARAM), dJ/dtheta.  Also returns log|psi'/psi|^2,
 * the acceptance A, the proposed position and the log weight L (A = min(1, e^L)).
 * The walker is left unchanged. */
double vmco_lp_sample_grad(vmco_walker* w, const double* theta, const vmco_lp_hyper* hp, int i,
                           double xi0, double xi1, double* grad,
                           double* out /* 6: J, A, logR, x', y', L */);

/* ---- Electron densities (src/sampler.cpp:639-692) ----
 * Add the counts of one configuration x [N*2] to the histograms (any may be
 * NULL): radial [bins], spatial [bins][bins] (row = y cell), two_body
 * [bins][bins]; dr = max_radius / bins. */
void vmco_density_add(const double* x, int N, int bins, double max_radius, int64_t* radial,
                      int64_t* spatial, int64_t* two_body);

int vmco_num_orbitals(int N, int D);  /* n = N/2 for closed shells, -1 if open */
void vmco_orbital_list(int n, int D, int* states /* n*D */);
/* value / derivative / second derivative of orbital j along dimension d at r */
double vmco_basis_element(const vmco_walker* w, int j, const double* r);
double vmco_basis_element_der(const vmco_walker* w, int j, int d, const double* r);
double vmco_basis_element_secder(const vmco_walker* w, int j, int d, const double* r);

#ifdef __cplusplus
}
#endif
#endif

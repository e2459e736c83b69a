/*
 * vmc_oracle.c -- FP64 CPU restatement of VMaChine's sampling hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine, never the thing
 * measured or shipped.  See vmc_oracle.h.
 *
 * The restatement follows the reference's element state machines (initialise,
 * update on a proposed move, reset on rejection) and its arithmetic; every
 * function cites the reference file:line it follows.  Where the reference has
 * a quirk that changes results, the quirk is reproduced when its VMCO_Q* bit is
 * set and the mathematically exact form is used otherwise.
 *
 * Pins (checked in tests/test_oracle_*.py):
 *   elements, kinetic/external/interaction energy  -> golden vectors written by
 *       the reference itself (oracle/ref/harness.cpp), finite differences,
 *       closed forms, closed-shell zero variance;
 *   Metropolis step + sampler + optimiser        -> reference trajectories
 *       driven by the same scripted random numbers;
 *   Philox4x32-10                                -> Random123 known-answer vectors;
 *   blocking                                     -> the reference's own blocking
 *       fixture (examples/resampling/523861_0.dat, log line 182);
 *   chi-square quantiles                         -> scipy.stats.chi2.
 */
#include "vmc_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
 * as 1, 2, 3", SC'11).  Published round function and constants.              */
/* ------------------------------------------------------------------------- */
void vmco_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 23-bit uniform in (0,1): (m + 1/2) 2^-23 with m < 2^23 needs 24 significant
 * bits, so it is exact in float and double alike */
static double u23(uint32_t u) { return ((double)(u >> 9) + 0.5) * (1.0 / 8388608.0); }

static const double TWO_PI = 6.283185307179586476925286766559;

void vmco_step_draws(uint64_t seed, uint32_t walker, uint64_t counter, int nd,
                     int* k, double* z, double* u_move, double* u_acc)
{
    uint32_t ctr[4] = {(uint32_t)counter, (uint32_t)(counter >> 32), walker, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    vmco_philox4x32_10(ctr, key, o);
    *k = (int)(((uint64_t)o[0] * (uint64_t)nd) >> 32);
    double U1 = u23(o[1]), U2 = u23(o[2]);
    *z = sqrt(-2.0 * log(U1)) * cos(TWO_PI * U2);
    *u_move = U1;
    *u_acc = u23(o[3]);
}

void vmco_initial_positions(uint64_t seed, uint32_t walker, int nd, double omega, double* out)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int blk = 0; 2 * blk < nd; blk++) {
        uint32_t ctr[4] = {(uint32_t)blk, 0u, walker, 1u};
        uint32_t o[4];
        vmco_philox4x32_10(ctr, key, o);
        double r = sqrt(-2.0 * log(u23(o[0])));
        double th = TWO_PI * u23(o[1]);
        /* std::normal_distribution(mean, 1/omega) takes a standard deviation
         * (ref InitialStates/randomnormal.cpp:17, RNG/mersennetwister.cpp:67) */
        out[2 * blk] = r * cos(th) / omega;
        if (2 * blk + 1 < nd) out[2 * blk + 1] = r * sin(th) / omega;
    }
}

/* ------------------------------------------------------------------------- */
/* Hermite basis                                                              */
/* ------------------------------------------------------------------------- */

/* closed-shell orbital count (ref Basis/basis.cpp:106-123): smallest shell
 * count c+1 with 2*C(c+D, D) >= N; returns the number of states n. */
int vmco_num_orbitals(int N, int D)
{
    for (int c = 0; c < 64; c++) {
        long long b = 1;
        for (int i = 1; i <= D; i++) b = b * (c + i) / i; /* C(c+D, D) */
        if (2 * b == N) return (int)b;
        if (2 * b > N) return -1;
    }
    return -1;
}

/* list of (n_x, n_y[, n_z]) states, shell by shell (ref Basis/basis.cpp:125-160) */
void vmco_orbital_list(int n, int D, int* states)
{
    int counter = 0;
    for (int i = 0; counter < n; i++) {
        if (D == 2) {
            for (int j = 0; j <= i && counter < n; j++) {
                states[counter * 2 + 0] = i - j;
                states[counter * 2 + 1] = j;
                counter++;
            }
        } else {
            for (int j = 0; j <= i && counter < n; j++)
                for (int k = 0; k <= i - j && counter < n; k++) {
                    states[counter * 3 + 0] = i - j - k;
                    states[counter * 3 + 1] = j;
                    states[counter * 3 + 2] = k;
                    counter++;
                }
        }
    }
}

/* physicists' Hermite polynomial H_n(t) by the three-term recurrence
 * H_{n+1} = 2t H_n - 2n H_{n-1}; derivatives H_n' = 2n H_{n-1},
 * H_n'' = 4n(n-1) H_{n-2}. */
static double hermite_H(int n, double t)
{
    if (n == 0) return 1.0;
    double hm = 1.0, h = 2.0 * t;
    for (int m = 1; m < n; m++) {
        double hn = 2.0 * t * h - 2.0 * m * hm;
        hm = h;
        h = hn;
    }
    return h;
}

/* normalisation 1/sqrt(2^n n!) (ref Basis/hermite.cpp:13-17) */
static double hermite_c(int n)
{
    double f = 1.0;
    for (int i = 1; i <= n; i++) f *= 2.0 * i;
    return 1.0 / sqrt(f);
}

/* ------------------------------------------------------------------------- */
/* walker state                                                               */
/* ------------------------------------------------------------------------- */
struct vmco_walker {
    vmco_config c;
    int N, D, ND, n, H, maxpar;
    int* states;       /* n x D */
    double* P;         /* num_elements x maxpar */
    double scale;      /* sqrt(omega) for the scaled basis, 1 otherwise */
    double* hc;        /* hermite_c(0..) */

    /* element parameters */
    double alpha, gamma;
    double* beta;      /* N x N Pade a_ij (ref padejastrow.cpp:140-170) */
    double* a_rbm;     /* ND */
    double* b_rbm;     /* H */
    double* W;         /* ND x H, W[k + ND*h] (ref rbmproduct.cpp:21-30) */
    double* W2;
    double sig2_g, sig2_p, sig4_p;

    /* state */
    double* x;         /* ND */
    double* dist;      /* N x N symmetric */
    double r_gauss;
    double* A;         /* N x n, row q = particle q */
    double* Ainv;      /* column q of the inverse at Ainv[q*n .. q*n+n) */
    double* der;       /* ND x n */
    double* sec;       /* ND x n */
    double* detDer;    /* ND */
    double* detSec;    /* ND */
    double R;
    double* h;         /* N x N symmetric, r/(1+gamma r) */
    double* fp;        /* N x N symmetric, a/(1+gamma r)^2 */
    double r_pade;
    double* Xa;        /* ND */
    double r_rbmg;
    double* xs;        /* ND (sorted if sorting) */
    double* v;         /* H */
    double* nn;        /* H */
    double* pp;        /* H */
    double r_rbmp;
    double* qf;        /* ND importance-sampling quantum force */

    /* saved slices for resetArrays */
    int mk, mp;
    double sv_x;
    double* sv_dist;   /* N */
    double* sv_A;      /* n */
    double* sv_der;    /* D*n */
    double* sv_sec;    /* D*n */
    double* sv_Ainv;   /* n*n */
    double* sv_detDer; /* n*D */
    double* sv_detSec; /* n*D */
    double sv_R;
    double* sv_h;      /* N */
    double* sv_fp;     /* N */
    double* sv_Xa;     /* ND */
    double* sv_xs;     /* ND */
    double* sv_v;      /* H */
    double* sv_nn;
    double* sv_pp;
    double sv_ratio[VMCO_MAXE];
};

static int has_kind(const vmco_config* c, int kind)
{
    for (int e = 0; e < c->num_elements; e++)
        if (c->kinds[e] == kind) return 1;
    return 0;
}

int vmco_num_parameters(const vmco_config* c, int e)
{
    int ND = c->N * c->D;
    switch (c->kinds[e]) {
    case VMCO_GAUSSIAN: return 1;      /* gaussian.h */
    case VMCO_SLATER: return 0;        /* slaterdeterminant.h */
    case VMCO_PADE: return 1;          /* padejastrow.h */
    case VMCO_RBM_GAUSSIAN: return ND; /* rbmgaussian.cpp:22 */
    case VMCO_RBM_PRODUCT: return c->num_hidden * ND + c->num_hidden; /* rbmproduct.cpp:12 */
    }
    return 0;
}

int vmco_max_parameters(const vmco_config* c)
{
    int m = 0;
    for (int e = 0; e < c->num_elements; e++) {
        int p = vmco_num_parameters(c, e);
        if (p > m) m = p;
    }
    return m;
}

static double* dalloc(size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }

vmco_walker* vmco_walker_new(const vmco_config* c)
{
    vmco_walker* w = (vmco_walker*)calloc(1, sizeof(vmco_walker));
    w->c = *c;
    w->N = c->N;
    w->D = c->D;
    w->ND = c->N * c->D;
    w->H = c->num_hidden;
    w->maxpar = vmco_max_parameters(c);
    int N = w->N, D = w->D, ND = w->ND, H = w->H;
    w->n = N / 2;
    int n = w->n;
    w->states = (int*)calloc((size_t)(n ? n : 1) * D, sizeof(int));
    if (has_kind(c, VMCO_SLATER)) vmco_orbital_list(n, D, w->states);
    w->P = dalloc((size_t)c->num_elements * (w->maxpar ? w->maxpar : 1));
    w->scale = c->basis_scaled ? sqrt(c->omega) : 1.0;
    w->hc = dalloc(32);
    for (int i = 0; i < 32; i++) w->hc[i] = hermite_c(i);

    /* Pade spin-dependent a (ref padejastrow.cpp:140-170) */
    w->beta = dalloc((size_t)N * N);
    double sym = 1.0 / 3, anti = 1.0;
    if (D == 3) { sym = 1.0 / 4; anti = 1.0 / 2; }
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            int same = (i < N / 2 && j < N / 2) || (i >= N / 2 && j >= N / 2);
            w->beta[i * N + j] = same ? sym : anti;
        }
    w->a_rbm = dalloc(ND);
    w->b_rbm = dalloc(H);
    w->W = dalloc((size_t)ND * H);
    w->W2 = dalloc((size_t)ND * H);
    w->sig2_g = c->sigma * c->sigma;                       /* rbmgaussian.cpp:16-17 */
    w->sig2_p = (10.0 * c->sigma) * (10.0 * c->sigma);     /* rbmproduct.cpp:15-17 (Q5) */
    w->sig4_p = w->sig2_p * w->sig2_p;

    w->x = dalloc(ND);
    w->dist = dalloc((size_t)N * N);
    w->A = dalloc((size_t)N * n);
    w->Ainv = dalloc((size_t)N * n);
    w->der = dalloc((size_t)ND * n);
    w->sec = dalloc((size_t)ND * n);
    w->detDer = dalloc(ND);
    w->detSec = dalloc(ND);
    w->h = dalloc((size_t)N * N);
    w->fp = dalloc((size_t)N * N);
    w->Xa = dalloc(ND);
    w->xs = dalloc(ND);
    w->v = dalloc(H);
    w->nn = dalloc(H);
    w->pp = dalloc(H);
    w->qf = dalloc(ND);

    w->sv_dist = dalloc(N);
    w->sv_A = dalloc(n);
    w->sv_der = dalloc((size_t)D * n);
    w->sv_sec = dalloc((size_t)D * n);
    w->sv_Ainv = dalloc((size_t)n * n);
    w->sv_detDer = dalloc((size_t)n * D);
    w->sv_detSec = dalloc((size_t)n * D);
    w->sv_h = dalloc(N);
    w->sv_fp = dalloc(N);
    w->sv_Xa = dalloc(ND);
    w->sv_xs = dalloc(ND);
    w->sv_v = dalloc(H);
    w->sv_nn = dalloc(H);
    w->sv_pp = dalloc(H);
    return w;
}

void vmco_walker_free(vmco_walker* w)
{
    if (!w) return;
    double* ptrs[] = {w->P, w->hc, w->beta, w->a_rbm, w->b_rbm, w->W, w->W2, w->x, w->dist,
                      w->A, w->Ainv, w->der, w->sec, w->detDer, w->detSec, w->h, w->fp,
                      w->Xa, w->xs, w->v, w->nn, w->pp, w->qf, w->sv_dist, w->sv_A,
                      w->sv_der, w->sv_sec, w->sv_Ainv, w->sv_detDer, w->sv_detSec,
                      w->sv_h, w->sv_fp, w->sv_Xa, w->sv_xs, w->sv_v, w->sv_nn, w->sv_pp};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    free(w->states);
    free(w);
}

/* WaveFunction::updateParameters of every element (elementoperation.cpp:68-73) */
void vmco_set_parameters(vmco_walker* w, const double* params)
{
    const vmco_config* c = &w->c;
    int ND = w->ND, H = w->H, mp = w->maxpar;
    memcpy(w->P, params, sizeof(double) * (size_t)c->num_elements * mp);
    for (int e = 0; e < c->num_elements; e++) {
        const double* row = params + (size_t)e * mp;
        switch (c->kinds[e]) {
        case VMCO_GAUSSIAN: w->alpha = row[0]; break;               /* gaussian.cpp:52-56 */
        case VMCO_PADE: w->gamma = row[0]; break;                   /* padejastrow.cpp:65-68 */
        case VMCO_RBM_GAUSSIAN:                                     /* rbmgaussian.cpp:19-22 */
            for (int k = 0; k < ND; k++) w->a_rbm[k] = row[k];
            break;
        case VMCO_RBM_PRODUCT:                                      /* rbmproduct.cpp:18-27 */
            for (int h = 0; h < H; h++) w->b_rbm[h] = row[h];
            for (int i = 0; i < ND * H; i++) {
                w->W[i] = row[H + i];
                w->W2[i] = row[H + i] * row[H + i];
            }
            break;
        }
    }
}

void vmco_get_positions(const vmco_walker* w, double* out) { memcpy(out, w->x, sizeof(double) * w->ND); }

/* ---- basis functions (ref Basis/hermite.cpp:33-71) ---- */
static double h1d(const vmco_walker* w, int nq, double x)
{
    return w->hc[nq] * hermite_H(nq, w->scale * x);
}
static double dh1d(const vmco_walker* w, int nq, double x)
{
    if (nq == 0) return 0.0;
    return w->scale * w->hc[nq] * 2.0 * nq * hermite_H(nq - 1, w->scale * x);
}
static double ddh1d(const vmco_walker* w, int nq, double x)
{
    if (nq < 2) return 0.0;
    return w->scale * w->scale * w->hc[nq] * 4.0 * nq * (nq - 1) * hermite_H(nq - 2, w->scale * x);
}

double vmco_basis_element(const vmco_walker* w, int j, const double* r)
{
    double prod = 1.0;
    for (int d = 0; d < w->D; d++) prod *= h1d(w, w->states[j * w->D + d], r[d]);
    return prod;
}

double vmco_basis_element_der(const vmco_walker* w, int j, int d, const double* r)
{
    double prod = dh1d(w, w->states[j * w->D + d], r[d]);
    for (int e = 0; e < w->D; e++)
        if (e != d) prod *= h1d(w, w->states[j * w->D + e], r[e]);
    return prod;
}

double vmco_basis_element_secder(const vmco_walker* w, int j, int d, const double* r)
{
    double prod = ddh1d(w, w->states[j * w->D + d], r[d]);
    for (int e = 0; e < w->D; e++)
        if (e != d) prod *= h1d(w, w->states[j * w->D + e], r[e]);
    return prod;
}

/* ---- distances (ref InitialStates/initialstate.cpp:20-49, Metropolis/metropolis.cpp:9-29) ---- */
static double pair_dist(const vmco_walker* w, int i, int j)
{
    double s = 0;
    for (int d = 0; d < w->D; d++) {
        double df = w->x[i * w->D + d] - w->x[j * w->D + d];
        s += df * df;
    }
    return sqrt(s);
}

/* ---- Slater (ref WaveFunctions/slaterdeterminant.cpp) ---- */
static void slater_row(vmco_walker* w, int q)
{
    const double* r = w->x + q * w->D;
    for (int j = 0; j < w->n; j++) w->A[q * w->n + j] = vmco_basis_element(w, j, r);   /* :99-104 */
    for (int d = 0; d < w->D; d++)
        for (int j = 0; j < w->n; j++) {
            w->der[(q * w->D + d) * w->n + j] = vmco_basis_element_der(w, j, d, r);      /* :106-115 */
            w->sec[(q * w->D + d) * w->n + j] = vmco_basis_element_secder(w, j, d, r);   /* :117-126 */
        }
}

/* dense inverse by LU with partial pivoting; in: a (n x n row-major), out: inv (row-major) */
static void lu_inverse(int n, const double* a, double* inv)
{
    double* lu = dalloc((size_t)n * n);
    int* piv = (int*)malloc(sizeof(int) * (size_t)(n ? n : 1));
    memcpy(lu, a, sizeof(double) * (size_t)n * n);
    for (int i = 0; i < n; i++) piv[i] = i;
    for (int k = 0; k < n; k++) {
        int p = k;
        double best = fabs(lu[k * n + k]);
        for (int i = k + 1; i < n; i++)
            if (fabs(lu[i * n + k]) > best) { best = fabs(lu[i * n + k]); p = i; }
        if (p != k) {
            for (int j = 0; j < n; j++) {
                double t = lu[k * n + j]; lu[k * n + j] = lu[p * n + j]; lu[p * n + j] = t;
            }
            int t = piv[k]; piv[k] = piv[p]; piv[p] = t;
        }
        for (int i = k + 1; i < n; i++) {
            lu[i * n + k] /= lu[k * n + k];
            for (int j = k + 1; j < n; j++) lu[i * n + j] -= lu[i * n + k] * lu[k * n + j];
        }
    }
    /* solve for each column of the identity */
    double* col = dalloc(n);
    for (int c = 0; c < n; c++) {
        for (int i = 0; i < n; i++) col[i] = (piv[i] == c) ? 1.0 : 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < i; j++) col[i] -= lu[i * n + j] * col[j];
        for (int i = n - 1; i >= 0; i--) {
            for (int j = i + 1; j < n; j++) col[i] -= lu[i * n + j] * col[j];
            col[i] /= lu[i * n + i];
        }
        for (int i = 0; i < n; i++) inv[i * n + c] = col[i];
    }
    free(col);
    free(piv);
    free(lu);
}

/* updateSlaterDeterminantDerivatives (ref slaterdeterminant.cpp:133-141) */
static void slater_derivatives(vmco_walker* w, int start, int end)
{
    int n = w->n, D = w->D;
    for (int i = start * D; i < end * D; i++) {
        int q = i / D;
        double s1 = 0, s2 = 0;
        for (int j = 0; j < n; j++) {
            s1 += w->der[i * n + j] * w->Ainv[q * n + j];
            s2 += w->sec[i * n + j] * w->Ainv[q * n + j];
        }
        w->detDer[i] = s1;
        w->detSec[i] = s2;
    }
}

/* initializeArrays (ref slaterdeterminant.cpp:45-62, 90-97) */
static void slater_init(vmco_walker* w)
{
    int n = w->n;
    for (int q = 0; q < w->N; q++) slater_row(w, q);
    double* blk = dalloc((size_t)n * n);
    double* inv = dalloc((size_t)n * n);
    for (int b = 0; b < 2; b++) {
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) blk[i * n + j] = w->A[(b * n + i) * n + j];
        lu_inverse(n, blk, inv);
        /* inverse[k][i] becomes column (b*n + i), row k */
        for (int i = 0; i < n; i++)
            for (int k = 0; k < n; k++) w->Ainv[(b * n + i) * n + k] = inv[k * n + i];
    }
    free(blk);
    free(inv);
    w->R = 1;
    slater_derivatives(w, 0, w->N);
}

/* updateArrays + updateSlaterMatrixInverse (ref slaterdeterminant.cpp:17-35, 128-153) */
static void slater_update(vmco_walker* w, int p)
{
    int n = w->n;
    slater_row(w, p);
    int start = (p >= n) ? n : 0, end = start + n;
    const double* row = w->A + p * n;
    double* colp = w->Ainv + p * n;
    double R = 0;
    for (int j = 0; j < n; j++) R += row[j] * colp[j];
    w->R = R;
    for (int q = start; q < end; q++) {
        if (q == p) continue;
        double* colq = w->Ainv + q * n;
        double S = 0;
        for (int j = 0; j < n; j++) S += row[j] * colq[j];
        for (int j = 0; j < n; j++) colq[j] -= S * colp[j] / R;
    }
    for (int j = 0; j < n; j++) colp[j] /= R;
    slater_derivatives(w, start, end);
}

/* ---- Pade-Jastrow (ref WaveFunctions/padejastrow.cpp) ---- */
static void pade_pair(vmco_walker* w, int i, int j)
{
    double r = w->dist[i * w->N + j];
    double f = 1.0 / (1.0 + w->gamma * r);              /* :192-209 */
    double hh = r * f;
    double ff = w->beta[i * w->N + j] * f * f;
    w->h[i * w->N + j] = w->h[j * w->N + i] = hh;
    w->fp[i * w->N + j] = w->fp[j * w->N + i] = ff;
}

static void pade_init(vmco_walker* w)   /* :17-27, 131-138 */
{
    for (int i = 0; i < w->N; i++)
        for (int j = i + 1; j < w->N; j++) pade_pair(w, i, j);
}

static double pade_ratio(vmco_walker* w, int p, const double* h_old)
{
    /* calculateProbabilityRatio (:211-218): the reference sums only the
     * upper-triangular entries of row i_p (pairs j > i_p), Q2 */
    double s = 0;
    int lo = (w->c.quirks & VMCO_Q2_PADE_RATIO_UPPER) ? p + 1 : 0;
    for (int j = lo; j < w->N; j++) {
        if (j == p) continue;
        s += w->beta[p * w->N + j] * (w->h[p * w->N + j] - h_old[j]);
    }
    return exp(2.0 * s);
}

static double pade_gradient(vmco_walker* w, int k)    /* :75-91 */
{
    int N = w->N, D = w->D;
    int kp = k / D, kd = k % D;
    double g = 0;
    for (int jp = 0; jp < N; jp++) {
        if (jp == kp) continue;
        double r = w->dist[kp * N + jp];
        g += w->fp[kp * N + jp] * (w->x[k] - w->x[jp * D + kd]) / r;
    }
    return g;
}

static double pade_laplacian(vmco_walker* w)           /* :93-109 */
{
    int N = w->N, D = w->D;
    double s = 0;
    for (int k = 0; k < w->ND; k++) {
        int kp = k / D, kd = k % D;
        for (int jp = kp + 1; jp < N; jp++) {
            double r = w->dist[kp * N + jp];
            double pd = (w->x[k] - w->x[jp * D + kd]) / r;
            s += w->fp[kp * N + jp] * (1.0 - (1.0 + 2.0 * w->gamma * w->h[kp * N + jp]) * pd * pd) / r;
        }
    }
    return 2.0 * s;
}

static double pade_gamma_gradient(vmco_walker* w)      /* :111-116 */
{
    double s = 0;
    for (int i = 0; i < w->N; i++)
        for (int j = i + 1; j < w->N; j++) s += w->beta[i * w->N + j] * w->h[i * w->N + j] * w->h[i * w->N + j];
    return (w->c.quirks & VMCO_Q3_PADE_GAMMA_HALF) ? -s / 2.0 : -s;
}

/* ---- RBM elements (ref WaveFunctions/rbmgaussian.cpp, rbmproduct.cpp) ---- */
static int cmp_double(const void* a, const void* b)
{
    double x = *(const double*)a, y = *(const double*)b;
    return (x > y) - (x < y);
}

static void rbmp_vectors(vmco_walker* w)   /* rbmproduct.cpp:104-110 */
{
    int ND = w->ND, H = w->H;
    memcpy(w->xs, w->x, sizeof(double) * ND);
    /* WaveFunction::argsort sorts the full coordinate vector; the returned index map is the
     * identity (ref wavefunction.cpp:40-54), so only the values are reordered */
    if (w->c.sorting) qsort(w->xs, (size_t)ND, sizeof(double), cmp_double);
    for (int h = 0; h < H; h++) {
        double s = 0;
        for (int k = 0; k < ND; k++) s += w->W[k + ND * h] * w->xs[k];
        w->v[h] = w->b_rbm[h] + s / w->sig2_p;
        double e = exp(w->v[h]);
        w->pp[h] = 1.0 / (e + 1.0);
        w->nn[h] = e * w->pp[h];
    }
}

/* ------------------------------------------------------------------------- */
void vmco_init_positions(vmco_walker* w, const double* positions)
{
    const vmco_config* c = &w->c;
    int N = w->N;
    memcpy(w->x, positions, sizeof(double) * w->ND);
    for (int i = 0; i < N; i++)
        for (int j = i + 1; j < N; j++) w->dist[i * N + j] = w->dist[j * N + i] = pair_dist(w, i, j);
    for (int e = 0; e < c->num_elements; e++) {
        w->sv_ratio[e] = 1.0;
        switch (c->kinds[e]) {
        case VMCO_GAUSSIAN: w->r_gauss = 1; break;
        case VMCO_SLATER: slater_init(w); break;
        case VMCO_PADE: pade_init(w); w->r_pade = 1; break;
        case VMCO_RBM_GAUSSIAN:
            for (int k = 0; k < w->ND; k++) w->Xa[k] = w->x[k] - w->a_rbm[k];
            w->r_rbmg = 1;
            break;
        case VMCO_RBM_PRODUCT: rbmp_vectors(w); w->r_rbmp = 1; break;
        }
    }
    vmco_reset_quantum_force(w);
}

void vmco_propose(vmco_walker* w, int k, double newval)
{
    const vmco_config* c = &w->c;
    int N = w->N, D = w->D, n = w->n, ND = w->ND, H = w->H;
    int p = k / D;
    w->mk = k;
    w->mp = p;
    /* save what the move touches (the reference copies whole objects in setArrays) */
    w->sv_x = w->x[k];
    memcpy(w->sv_dist, w->dist + p * N, sizeof(double) * N);
    int hasS = has_kind(c, VMCO_SLATER), hasP = has_kind(c, VMCO_PADE);
    int hasRG = has_kind(c, VMCO_RBM_GAUSSIAN), hasRP = has_kind(c, VMCO_RBM_PRODUCT);
    int start = (p >= n) ? n : 0;
    if (hasS) {
        memcpy(w->sv_A, w->A + p * n, sizeof(double) * n);
        memcpy(w->sv_der, w->der + p * D * n, sizeof(double) * D * n);
        memcpy(w->sv_sec, w->sec + p * D * n, sizeof(double) * D * n);
        memcpy(w->sv_Ainv, w->Ainv + start * n, sizeof(double) * n * n);
        memcpy(w->sv_detDer, w->detDer + start * D, sizeof(double) * n * D);
        memcpy(w->sv_detSec, w->detSec + start * D, sizeof(double) * n * D);
        w->sv_R = w->R;
    }
    if (hasP) {
        memcpy(w->sv_h, w->h + p * N, sizeof(double) * N);
        memcpy(w->sv_fp, w->fp + p * N, sizeof(double) * N);
    }
    if (hasRG) memcpy(w->sv_Xa, w->Xa, sizeof(double) * ND);
    if (hasRP) {
        memcpy(w->sv_xs, w->xs, sizeof(double) * ND);
        memcpy(w->sv_v, w->v, sizeof(double) * H);
        memcpy(w->sv_nn, w->nn, sizeof(double) * H);
        memcpy(w->sv_pp, w->pp, sizeof(double) * H);
    }
    w->sv_ratio[0] = w->r_gauss;
    w->sv_ratio[1] = w->r_pade;
    w->sv_ratio[2] = w->r_rbmg;
    w->sv_ratio[3] = w->r_rbmp;

    double xold = w->x[k];
    w->x[k] = newval;
    /* Metropolis::calculateDistanceMatrixCross (metropolis.cpp:21-29) */
    for (int j = 0; j < N; j++)
        if (j != p) w->dist[p * N + j] = w->dist[j * N + p] = pair_dist(w, p, j);

    for (int e = 0; e < c->num_elements; e++) {
        switch (c->kinds[e]) {
        case VMCO_GAUSSIAN:  /* gaussian.cpp:23-37 */
            w->r_gauss = exp(c->omega * w->alpha * (xold * xold - newval * newval));
            break;
        case VMCO_SLATER: slater_update(w, p); break;
        case VMCO_PADE:
            for (int j = 0; j < N; j++)
                if (j != p) pade_pair(w, p, j);
            w->r_pade = pade_ratio(w, p, w->sv_h);
            break;
        case VMCO_RBM_GAUSSIAN: { /* rbmgaussian.cpp:34-46 */
            double old2 = w->Xa[k] * w->Xa[k];
            for (int i = 0; i < ND; i++) w->Xa[i] = w->x[i] - w->a_rbm[i];
            double diff = old2 - w->Xa[k] * w->Xa[k];
            double den = (c->quirks & VMCO_Q13_RBMG_SQRT_RATIO) ? 2.0 * w->sig2_g : w->sig2_g;
            w->r_rbmg = exp(diff / den);
            break;
        }
        case VMCO_RBM_PRODUCT: { /* rbmproduct.cpp:43-52, 111-116 */
            rbmp_vectors(w);
            double po = 1, pn = 1;
            for (int h = 0; h < H; h++) { po *= w->sv_pp[h]; pn *= w->pp[h]; }
            double q = po / pn;
            w->r_rbmp = q * q;
            break;
        }
        }
    }
}

void vmco_reject(vmco_walker* w)
{
    const vmco_config* c = &w->c;
    int N = w->N, D = w->D, n = w->n, ND = w->ND, H = w->H;
    int p = w->mp, k = w->mk;
    w->x[k] = w->sv_x;
    for (int j = 0; j < N; j++) {
        w->dist[p * N + j] = w->dist[j * N + p] = w->sv_dist[j];
    }
    int start = (p >= n) ? n : 0;
    if (has_kind(c, VMCO_SLATER)) {
        memcpy(w->A + p * n, w->sv_A, sizeof(double) * n);
        memcpy(w->der + p * D * n, w->sv_der, sizeof(double) * D * n);
        memcpy(w->sec + p * D * n, w->sv_sec, sizeof(double) * D * n);
        memcpy(w->Ainv + start * n, w->sv_Ainv, sizeof(double) * n * n);
        memcpy(w->detDer + start * D, w->sv_detDer, sizeof(double) * n * D);
        memcpy(w->detSec + start * D, w->sv_detSec, sizeof(double) * n * D);
        w->R = w->sv_R;
    }
    if (has_kind(c, VMCO_PADE)) {
        for (int j = 0; j < N; j++) {
            w->h[p * N + j] = w->h[j * N + p] = w->sv_h[j];
            w->fp[p * N + j] = w->fp[j * N + p] = w->sv_fp[j];
        }
    }
    if (has_kind(c, VMCO_RBM_GAUSSIAN)) memcpy(w->Xa, w->sv_Xa, sizeof(double) * ND);
    if (has_kind(c, VMCO_RBM_PRODUCT)) {
        memcpy(w->xs, w->sv_xs, sizeof(double) * ND);
        memcpy(w->v, w->sv_v, sizeof(double) * H);
        memcpy(w->nn, w->sv_nn, sizeof(double) * H);
        memcpy(w->pp, w->sv_pp, sizeof(double) * H);
    }
    w->r_gauss = w->sv_ratio[0];
    w->r_pade = w->sv_ratio[1];
    w->r_rbmg = w->sv_ratio[2];
    w->r_rbmp = w->sv_ratio[3];
}

void vmco_accept(vmco_walker* w) { (void)w; }

/* evaluateRatio of element e */
double vmco_element_ratio(vmco_walker* w, int e)
{
    switch (w->c.kinds[e]) {
    case VMCO_GAUSSIAN: return w->r_gauss;
    case VMCO_SLATER: return w->R * w->R;            /* slaterdeterminant.cpp:155-158 */
    case VMCO_PADE: return w->r_pade;
    case VMCO_RBM_GAUSSIAN: return w->r_rbmg;
    case VMCO_RBM_PRODUCT: return w->r_rbmp;
    }
    return 1.0;
}

double vmco_element_gradient(vmco_walker* w, int e, int k)
{
    switch (w->c.kinds[e]) {
    case VMCO_GAUSSIAN: return -w->c.omega * w->alpha * w->x[k];      /* gaussian.cpp:62-65 */
    case VMCO_SLATER: return w->detDer[k];                              /* slaterdeterminant.cpp:160-163 */
    case VMCO_PADE: return pade_gradient(w, k);
    case VMCO_RBM_GAUSSIAN: return -w->Xa[k] / w->sig2_g;               /* rbmgaussian.cpp:66-74 */
    case VMCO_RBM_PRODUCT: {                                            /* rbmproduct.cpp:62-70 */
        double s = 0;
        for (int h = 0; h < w->H; h++) s += w->W[k + w->ND * h] * w->nn[h];
        return s / w->sig2_p;
    }
    }
    return 0.0;
}

double vmco_element_laplacian(vmco_walker* w, int e)
{
    switch (w->c.kinds[e]) {
    case VMCO_GAUSSIAN: return -w->c.omega * w->alpha * w->ND;          /* gaussian.cpp:67-71 */
    case VMCO_SLATER: {                                                 /* slaterdeterminant.cpp:165-169 */
        double s = 0, g = 0;
        for (int k = 0; k < w->ND; k++) { s += w->detSec[k]; g += w->detDer[k] * w->detDer[k]; }
        return s - g;
    }
    case VMCO_PADE: return pade_laplacian(w);
    case VMCO_RBM_GAUSSIAN: return -w->ND / w->sig2_g;                  /* rbmgaussian.cpp:76-79 */
    case VMCO_RBM_PRODUCT: {                                            /* rbmproduct.cpp:72-75 */
        double s = 0;
        for (int h = 0; h < w->H; h++) {
            double a = w->pp[h] * w->nn[h];
            for (int k = 0; k < w->ND; k++) s += w->W2[k + w->ND * h] * a;
        }
        return s / w->sig4_p;
    }
    }
    return 0.0;
}

void vmco_element_parameter_gradient(vmco_walker* w, int e, double* out)
{
    int mp = w->maxpar, ND = w->ND, H = w->H;
    for (int i = 0; i < mp; i++) out[i] = 0;
    switch (w->c.kinds[e]) {
    case VMCO_GAUSSIAN: {                                               /* gaussian.cpp:73-78 */
        double s = 0;
        for (int k = 0; k < ND; k++) s += w->x[k] * w->x[k];
        if (mp > 0) out[0] = -0.5 * w->c.omega * s;
        break;
    }
    case VMCO_PADE: if (mp > 0) out[0] = pade_gamma_gradient(w); break;
    case VMCO_RBM_GAUSSIAN:                                             /* rbmgaussian.cpp:81-86 */
        for (int k = 0; k < ND; k++) out[k] = w->Xa[k] / w->sig2_g;
        break;
    case VMCO_RBM_PRODUCT: {                                            /* rbmproduct.cpp:77-83 */
        double scale = (w->c.quirks & VMCO_Q6_RBM_W_NO_SIGMA) ? 1.0 : 1.0 / w->sig2_p;
        for (int h = 0; h < H; h++) {
            out[h] = w->nn[h];
            for (int k = 0; k < ND; k++) out[H + k + ND * h] = w->xs[k] * w->nn[h] * scale;
        }
        break;
    }
    }
}

double vmco_probability_ratio(vmco_walker* w)   /* elementoperation.cpp:80-87 */
{
    double r = 1;
    for (int e = 0; e < w->c.num_elements; e++) r *= vmco_element_ratio(w, e);
    return r;
}

double vmco_kinetic_energy(vmco_walker* w)      /* elementoperation.cpp:95-109 */
{
    double t = 0;
    for (int e = 0; e < w->c.num_elements; e++) t += vmco_element_laplacian(w, e);
    for (int k = 0; k < w->ND; k++) {
        double g = 0;
        for (int e = 0; e < w->c.num_elements; e++) g += vmco_element_gradient(w, e, k);
        t += g * g;
    }
    return -0.5 * t;
}

double vmco_external_energy(vmco_walker* w)     /* harmonicoscillator.cpp:16-20 */
{
    double s = 0;
    for (int k = 0; k < w->ND; k++) s += w->x[k] * w->x[k];
    return 0.5 * w->c.omega * w->c.omega * s;
}

double vmco_interaction_energy(vmco_walker* w)  /* coulomb.cpp:17-48 */
{
    if (!w->c.interaction) return 0.0;
    double s = 0;
    for (int i = 0; i < w->N; i++)
        for (int j = i + 1; j < w->N; j++) s += 1.0 / w->dist[i * w->N + j];
    return s;
}

/* ---- Metropolis (ref Metropolis/importancesampling.cpp, bruteforce.cpp) ---- */
static double quantum_force(vmco_walker* w, int k)   /* importancesampling.cpp:34-41 */
{
    double g = 0;
    for (int e = 0; e < w->c.num_elements; e++) g += vmco_element_gradient(w, e, k);
    return 2.0 * g;
}

void vmco_reset_quantum_force(vmco_walker* w)          /* importancesampling.cpp:28-31 */
{
    for (int k = 0; k < w->ND; k++) w->qf[k] = quantum_force(w, k);
}
void vmco_get_quantum_force(const vmco_walker* w, double* out) { memcpy(out, w->qf, sizeof(double) * w->ND); }
void vmco_set_quantum_force(vmco_walker* w, const double* qf) { memcpy(w->qf, qf, sizeof(double) * w->ND); }

/* drift limit of Umrigar, Nightingale and Runge (J. Chem. Phys. 99, 2865 (1993)):
 * velocity v = F/2 -> v (sqrt(1 + 2 a v^2 dt) - 1) / (a v^2 dt), written stably */
static double limited_drift(double F, double a, double dt)
{
    if (!(a > 0)) return F;
    return 2.0 * F / (1.0 + sqrt(1.0 + 0.5 * a * F * F * dt));
}

int vmco_step(vmco_walker* w, int k, double z, double u_move, double u_acc)
{
    const vmco_config* c = &w->c;
    double dt = c->step_length;
    if (c->sampling == VMCO_BRUTE_FORCE) {                  /* bruteforce.cpp:26-52 */
        double newval = w->x[k] + 10.0 * (u_move - 0.5) * dt;
        vmco_propose(w, k, newval);
        double p = vmco_probability_ratio(w);
        if (p < u_acc) { vmco_reject(w); return 0; }
        return 1;
    }
    if (c->sampling == VMCO_IMPORTANCE_COMPAT) {            /* importancesampling.cpp:49-81 */
        double stale = w->qf[k];
        double fresh = quantum_force(w, k);
        double dx = 0.5 * dt * fresh + z * sqrt(dt);
        vmco_propose(w, k, w->x[k] + dx);
        double p = vmco_probability_ratio(w);
        /* GreenRatio (:43-47): exp(1/2 (F_stale - F_fresh) dx) + 1, Q1 */
        double wgt = (exp(0.5 * (stale - fresh) * dx) + 1.0) * p;
        if (wgt < u_acc) { vmco_reject(w); return 0; }
        w->qf[k] = fresh;
        return 1;
    }
    /* exact Metropolis-Hastings Langevin: single-coordinate proposal
     * x_k' = x_k + dt/2 F_k(x) + sqrt(dt) z, acceptance
     * |psi'/psi|^2 exp(-dx (F+F')/2 + dt/8 (F^2 - F'^2)) */
    double F = limited_drift(quantum_force(w, k), c->drift_limit, dt);
    double dx = 0.5 * dt * F + z * sqrt(dt);
    vmco_propose(w, k, w->x[k] + dx);
    double p = vmco_probability_ratio(w);
    double Fn = limited_drift(quantum_force(w, k), c->drift_limit, dt);
    double lg = -0.5 * dx * (F + Fn) + dt / 8.0 * (F * F - Fn * Fn);
    double a = p * exp(lg);
    if (!(a >= u_acc)) { vmco_reject(w); return 0; }   /* an undefined weight is rejected */
    return 1;
}

/* ---- Sampler (ref sampler.cpp:52-90) ---- */
uint64_t vmco_run(vmco_walker* w, uint64_t seed, uint32_t walker, uint64_t counter,
                  int steps, int equil, int iter, int batches,
                  double* acc, double* grad, double* gradE, double* energies)
{
    int E = w->c.num_elements, mp = w->maxpar;
    int S = steps / batches;                                  /* sampler.cpp:104 */
    int kb = iter % batches;
    double* g = dalloc((size_t)mp);
    for (int i = 0; i < equil + steps; i++) {
        int k; double z, um, ua;
        vmco_step_draws(seed, walker, counter++, w->ND, &k, &z, &um, &ua);
        int accepted = vmco_step(w, k, z, um, ua);
        if (i < equil) continue;
        double T = vmco_kinetic_energy(w);
        double V = vmco_external_energy(w);
        double Vi = vmco_interaction_energy(w);
        double El = T + V + Vi;
        if (energies) energies[i - equil] = El;
        acc[VMCO_ACC_E] += El;
        acc[VMCO_ACC_E2] += El * El;
        acc[VMCO_ACC_T] += T;
        acc[VMCO_ACC_VEXT] += V;
        acc[VMCO_ACC_VINT] += Vi;
        acc[VMCO_ACC_T2] += T * T;
        acc[VMCO_ACC_VEXT2] += V * V;
        acc[VMCO_ACC_VINT2] += Vi * Vi;
        acc[VMCO_ACC_ACCEPT] += accepted;
        acc[VMCO_ACC_NSAMPLES] += 1;
        if (i > S * kb && i < S * (kb + 1)) {                 /* sampler.cpp:82-86 */
            for (int e = 0; e < E; e++) {
                vmco_element_parameter_gradient(w, e, g);
                for (int j = 0; j < mp; j++) {
                    grad[e * mp + j] += g[j];
                    gradE[e * mp + j] += g[j] * El;
                }
            }
        }
    }
    free(g);
    return counter;
}

typedef struct {
    const vmco_config* c;
    const double* params;
    int W, threads, tid;
    double* positions;
    double* qf;
    uint64_t seed;
    uint64_t* counters;
    int steps, equil, iter, batches;
    double* acc_w;   /* W x NACC */
    double* g_w;     /* W x E*mp */
    double* ge_w;
} run_job;

static void* run_worker(void* arg)
{
    run_job* j = (run_job*)arg;
    int ND = j->c->N * j->c->D;
    int gsz = j->c->num_elements * vmco_max_parameters(j->c);
    vmco_walker* w = vmco_walker_new(j->c);
    vmco_set_parameters(w, j->params);
    for (int i = j->tid; i < j->W; i += j->threads) {
        vmco_init_positions(w, j->positions + (size_t)i * ND);
        if (j->qf) vmco_set_quantum_force(w, j->qf + (size_t)i * ND);
        j->counters[i] = vmco_run(w, j->seed, (uint32_t)i, j->counters[i], j->steps, j->equil,
                                  j->iter, j->batches, j->acc_w + (size_t)i * VMCO_NACC,
                                  j->g_w + (size_t)i * gsz, j->ge_w + (size_t)i * gsz, NULL);
        vmco_get_positions(w, j->positions + (size_t)i * ND);
        if (j->qf) vmco_get_quantum_force(w, j->qf + (size_t)i * ND);
    }
    vmco_walker_free(w);
    return NULL;
}

void vmco_run_many(const vmco_config* c, const double* params, int W, double* positions,
                   double* qf, uint64_t seed, uint64_t* counters, int steps, int equil,
                   int iter, int batches, int threads, double* acc, double* grad,
                   double* gradE)
{
    int gsz = c->num_elements * vmco_max_parameters(c);
    if (threads < 1) threads = 1;
    double* acc_w = dalloc((size_t)W * VMCO_NACC);
    double* g_w = dalloc((size_t)W * gsz);
    double* ge_w = dalloc((size_t)W * gsz);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
    run_job* jobs = (run_job*)malloc(sizeof(run_job) * threads);
    for (int t = 0; t < threads; t++) {
        run_job jb = {c, params, W, threads, t, positions, qf, seed, counters,
                      steps, equil, iter, batches, acc_w, g_w, ge_w};
        jobs[t] = jb;
        pthread_create(&th[t], NULL, run_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    /* deterministic reduction in walker order */
    for (int i = 0; i < W; i++) {
        for (int a = 0; a < VMCO_NACC; a++) acc[a] += acc_w[(size_t)i * VMCO_NACC + a];
        for (int a = 0; a < gsz; a++) {
            grad[a] += g_w[(size_t)i * gsz + a];
            gradE[a] += ge_w[(size_t)i * gsz + a];
        }
    }
    free(jobs); free(th); free(acc_w); free(g_w); free(ge_w);
}

/* ---- Optimisers ---- */
void vmco_optimizer_init(vmco_optimizer* o, int kind, int size, double eta, double gamma,
                         double monotonic_exp)
{
    memset(o, 0, sizeof(*o));
    o->kind = kind;
    o->size = size;
    o->eta = eta;
    o->gamma = gamma;
    o->monotonic_exp = monotonic_exp;
    o->beta1 = 0.89;      /* adam.h: beta1, beta2, epsilon */
    o->beta2 = 0.999;
    o->epsilon = 1e-8;
    for (int i = 0; i < size; i++) { o->m[i] = 1.0; o->v[i] = 1.0; }  /* adam.cpp:17-19, sgd.cpp:20 */
}

void vmco_optimizer_update(vmco_optimizer* o, double avgE, const double* avgG,
                           const double* avgGE, double* update)
{
    o->step += 1;
    for (int i = 0; i < o->size; i++) {
        double g = 2.0 * (avgGE[i] - avgE * avgG[i]);          /* optimization.cpp:10-16 */
        if (o->kind == 0) {                                      /* adam.cpp:22-38 */
            o->m[i] = o->beta1 * o->m[i] + (1 - o->beta1) * g;
            o->v[i] = o->beta2 * o->v[i] + (1 - o->beta2) * g * g;
            double mh = o->m[i] / (1 - pow(o->beta1, o->step));
            double vh = o->v[i] / (1 - pow(o->beta2, o->step));
            update[i] = o->eta * mh / (sqrt(vh) + o->epsilon);
        } else {                                                 /* sgd.cpp:23-29, gradientdescent.cpp:24-30 */
            double mono = 1.0 / pow((double)o->step, o->monotonic_exp);
            o->v[i] = o->gamma * o->v[i] + o->eta * g * mono;
            update[i] = o->v[i];
        }
    }
}

/* ---- Blocking (ref include/block/blocker.cpp; Jonsson, Phys. Rev. E 98, 043304 (2018)) ---- */

/* regularised lower incomplete gamma P(a, x) by its series / continued fraction */
static double gamma_p(double a, double x)
{
    if (x <= 0) return 0.0;
    double lg = lgamma(a);
    if (x < a + 1.0) {
        double sum = 1.0 / a, del = sum, ap = a;
        for (int i = 0; i < 100000; i++) {
            ap += 1.0;
            del *= x / ap;
            sum += del;
            if (fabs(del) < fabs(sum) * 1e-17) break;
        }
        return sum * exp(-x + a * log(x) - lg);
    }
    /* Lentz continued fraction for Q(a, x) */
    double b = x + 1.0 - a, cc = 1.0 / 1e-300, d = 1.0 / b, hh = d;
    for (int i = 1; i < 100000; i++) {
        double an = -i * (i - a);
        b += 2.0;
        d = an * d + b;
        if (fabs(d) < 1e-300) d = 1e-300;
        cc = b + an / cc;
        if (fabs(cc) < 1e-300) cc = 1e-300;
        d = 1.0 / d;
        double del = d * cc;
        hh *= del;
        if (fabs(del - 1.0) < 1e-17) break;
    }
    return 1.0 - exp(-x + a * log(x) - lg) * hh;
}

/* 95 % quantile of the chi-square distribution with dof degrees of freedom:
 * the x with P(dof/2, x/2) = 0.95, by bisection */
double vmco_chi2_quantile95(int dof)
{
    double lo = 0.0, hi = 10.0 * dof + 100.0;
    for (int i = 0; i < 200; i++) {
        double mid = 0.5 * (lo + hi);
        if (gamma_p(0.5 * dof, 0.5 * mid) < 0.95) lo = mid; else hi = mid;
    }
    return 0.5 * (lo + hi);
}

int vmco_blocking(double* x, uint64_t n, double* mean, double* stderr_, uint64_t* nK)
{
    if (n == 0 || (n & (n - 1)) != 0) {                  /* blocker.cpp:15-33 */
        *mean = *stderr_ = -1.0;
        *nK = 0;
        return -1;
    }
    double s0 = 0;
    for (uint64_t i = 0; i < n; i++) s0 += x[i];
    double mu = s0 / (double)n;
    int d = 0;
    while (((uint64_t)1 << d) < n) d++;
    double* gam = dalloc(d + 1);
    double* s = dalloc(d + 1);
    double* X = dalloc(n);
    for (uint64_t i = 0; i < n; i++) X[i] = x[i] - mu;
    uint64_t m = n;
    for (int k = 0; k < d; k++) {                        /* blocker.cpp:42-54 */
        double g1 = 0, v0 = 0;
        for (uint64_t i = 0; i + 1 < m; i++) g1 += X[i] * X[i + 1];
        for (uint64_t i = 0; i < m; i++) v0 += X[i] * X[i];
        gam[k] = g1 / (double)m;
        s[k] = v0 / (double)m;
        for (uint64_t i = 0; i < m / 2; i++) {           /* blocker.cpp:73-80 */
            x[i] = 0.5 * (x[2 * i] + x[2 * i + 1]);
            X[i] = x[i] - mu;
        }
        m /= 2;
    }
    /* test statistic M_j and the first k with M_k < chi2 quantile (blocker.cpp:55-69) */
    double* M = dalloc(d + 1);
    for (int j = 0; j < d; j++) M[j] = (gam[j] / s[j]) * (gam[j] / s[j]) * ldexp(1.0, d - j);
    double* cs = dalloc(d + 1);
    double run = 0;
    for (int j = 0; j < d; j++) { run += M[d - 1 - j]; cs[j] = run; }   /* cumsum(flip(M)) */
    int k;
    for (k = d - 1; k >= 0; k--)
        if (cs[k] < vmco_chi2_quantile95(k + 1)) break;
    int K = d - (k + 1);
    if (K > d - 1) K = d - 1;   /* the reference reads one past the end here */
    uint64_t nk = (uint64_t)1 << (d - K);
    *mean = mu;
    *stderr_ = sqrt(s[K] / (double)nk);
    *nK = nk;
    free(gam); free(s); free(X); free(M); free(cs);
    return 0;
}

/* ------------------------------------------------------------------------- *
 * Learned Metropolis proposal (no reference counterpart; see vmc_oracle.h)
 * ------------------------------------------------------------------------- */
int vmco_lp_features(const vmco_config* c, const double* x, int i, double xi, double yi, double* f)
{
    const double s = sqrt(c->omega);
    int N = c->N;
    for (int m = 0; m < VMCO_LP_IN; m++) f[m] = 0.0;
    f[0] = s * xi;
    f[1] = s * yi;
    /* the 8 nearest other electrons by squared distance, ties to the lower index */
    int taken[64] = {0};
    int nn = N - 1 < 8 ? N - 1 : 8;
    for (int m = 0; m < nn; m++) {
        int best = -1;
        double bd = 0;
        for (int j = 0; j < N; j++) {
            if (j == i || taken[j]) continue;
            double dx = x[2 * j] - xi, dy = x[2 * j + 1] - yi;
            double d2 = dx * dx + dy * dy;
            if (best < 0 || d2 < bd) { best = j; bd = d2; }
        }
        taken[best] = 1;
        f[2 + 2 * m] = s * (x[2 * best] - xi);
        f[3 + 2 * m] = s * (x[2 * best + 1] - yi);
    }
    return nn;
}

void vmco_lp_forward(const double* th, const double* f, double* h1, double* h2, double* o)
{
    double a1[VMCO_LP_HID], a2[VMCO_LP_HID];
    for (int u = 0; u < VMCO_LP_HID; u++) {
        double a = th[VMCO_LP_B1 + u];
        for (int m = 0; m < VMCO_LP_IN; m++) a += th[VMCO_LP_W1 + u * VMCO_LP_IN + m] * f[m];
        a1[u] = tanh(a);
    }
    for (int u = 0; u < VMCO_LP_HID; u++) {
        double a = th[VMCO_LP_B2 + u];
        for (int c = 0; c < VMCO_LP_HID; c++) a += th[VMCO_LP_W2 + u * VMCO_LP_HID + c] * a1[c];
        a2[u] = tanh(a);
    }
    for (int r = 0; r < VMCO_LP_OUT; r++) {
        double a = th[VMCO_LP_B3 + r];
        for (int u = 0; u < VMCO_LP_HID; u++) a += th[VMCO_LP_W3 + r * VMCO_LP_HID + u] * a2[u];
        o[r] = a;
    }
    if (h1) memcpy(h1, a1, sizeof(a1));
    if (h2) memcpy(h2, a2, sizeof(a2));
}

/* backward of the network: g_o = dX/do; adds scale * dX/dtheta into grad and
 * returns dX/df (unscaled) in g_f if non-NULL */
static void lp_backward(const double* th, const double* f, const double* h1, const double* h2,
                        const double* g_o, double scale, double* grad, double* g_f)
{
    double ga2[VMCO_LP_HID], ga1[VMCO_LP_HID];
    for (int u = 0; u < VMCO_LP_HID; u++) {
        double g = 0;
        for (int r = 0; r < VMCO_LP_OUT; r++) g += th[VMCO_LP_W3 + r * VMCO_LP_HID + u] * g_o[r];
        ga2[u] = g * (1.0 - h2[u] * h2[u]);
    }
    for (int r = 0; r < VMCO_LP_OUT; r++) {
        for (int u = 0; u < VMCO_LP_HID; u++) grad[VMCO_LP_W3 + r * VMCO_LP_HID + u] += scale * g_o[r] * h2[u];
        grad[VMCO_LP_B3 + r] += scale * g_o[r];
    }
    for (int c = 0; c < VMCO_LP_HID; c++) {
        double g = 0;
        for (int u = 0; u < VMCO_LP_HID; u++) g += th[VMCO_LP_W2 + u * VMCO_LP_HID + c] * ga2[u];
        ga1[c] = g * (1.0 - h1[c] * h1[c]);
    }
    for (int u = 0; u < VMCO_LP_HID; u++) {
        for (int c = 0; c < VMCO_LP_HID; c++) grad[VMCO_LP_W2 + u * VMCO_LP_HID + c] += scale * ga2[u] * h1[c];
        grad[VMCO_LP_B2 + u] += scale * ga2[u];
    }
    for (int c = 0; c < VMCO_LP_HID; c++) {
        for (int m = 0; m < VMCO_LP_IN; m++) grad[VMCO_LP_W1 + c * VMCO_LP_IN + m] += scale * ga1[c] * f[m];
        grad[VMCO_LP_B1 + c] += scale * ga1[c];
    }
    if (g_f) {
        for (int m = 0; m < VMCO_LP_IN; m++) {
            double g = 0;
            for (int c = 0; c < VMCO_LP_HID; c++) g += th[VMCO_LP_W1 + c * VMCO_LP_IN + m] * ga1[c];
            g_f[m] = g;
        }
    }
}

void vmco_lp_head(const vmco_lp_hyper* hp, double omega, const double* o, double* mu, double* sig)
{
    const double ell = 1.0 / sqrt(omega);
    for (int d = 0; d < 2; d++) {
        mu[d] = ell * hp->mu_max * tanh(o[d] / hp->mu_max);
        sig[d] = hp->sigma0 * ell * exp(hp->ls_max * tanh(o[2 + d] / hp->ls_max));
    }
}

/* log|psi(x')/psi(x)|^2 of moving electron i to (xn, yn) as two coordinate
 * moves, and the drift 2 grad ln|psi| of electron i at x'; the walker is left
 * at x' */
static double lp_move(vmco_walker* w, int i, double xn, double yn, double* F)
{
    vmco_propose(w, 2 * i, xn);
    double r1 = vmco_probability_ratio(w);
    vmco_accept(w);
    vmco_propose(w, 2 * i + 1, yn);
    double r2 = vmco_probability_ratio(w);
    vmco_accept(w);
    if (F) {
        F[0] = quantum_force(w, 2 * i);
        F[1] = quantum_force(w, 2 * i + 1);
    }
    return log(r1) + log(r2);
}

typedef struct {
    double fa[VMCO_LP_IN], h1a[VMCO_LP_HID], h2a[VMCO_LP_HID], oa[VMCO_LP_OUT];
    double fb[VMCO_LP_IN], h1b[VMCO_LP_HID], h2b[VMCO_LP_HID], ob[VMCO_LP_OUT];
    double mu[2], sig[2], mub[2], sigb[2], dx[2], r[2];
    double xn, yn, lqf, lqr;
    int nnb;
} lp_pass;

/* both network passes of a proposal of electron i with normals xi */
static void lp_propose(const vmco_walker* w, const double* th, const vmco_lp_hyper* hp, int i,
                       const double* xi, lp_pass* q)
{
    const double* x = w->x;
    vmco_lp_features(&w->c, x, i, x[2 * i], x[2 * i + 1], q->fa);
    vmco_lp_forward(th, q->fa, q->h1a, q->h2a, q->oa);
    vmco_lp_head(hp, w->c.omega, q->oa, q->mu, q->sig);
    q->lqf = 0;
    for (int d = 0; d < 2; d++) {
        q->dx[d] = q->mu[d] + q->sig[d] * xi[d];
        q->lqf -= 0.5 * xi[d] * xi[d] + log(q->sig[d]);
    }
    q->xn = x[2 * i] + q->dx[0];
    q->yn = x[2 * i + 1] + q->dx[1];
    q->nnb = vmco_lp_features(&w->c, x, i, q->xn, q->yn, q->fb);
    vmco_lp_forward(th, q->fb, q->h1b, q->h2b, q->ob);
    vmco_lp_head(hp, w->c.omega, q->ob, q->mub, q->sigb);
    q->lqr = 0;
    for (int d = 0; d < 2; d++) {
        q->r[d] = -q->dx[d] - q->mub[d];
        q->lqr -= q->r[d] * q->r[d] / (2.0 * q->sigb[d] * q->sigb[d]) + log(q->sigb[d]);
    }
}

int vmco_lp_step(vmco_walker* w, const double* th, const vmco_lp_hyper* hp, int i, double xi0,
                 double xi1, double ua)
{
    const double xi[2] = {xi0, xi1};
    lp_pass q;
    lp_propose(w, th, hp, i, xi, &q);
    double* old = dalloc(w->ND);
    memcpy(old, w->x, sizeof(double) * w->ND);
    double logR = lp_move(w, i, q.xn, q.yn, NULL);
    int acc = (logR + q.lqr - q.lqf >= log(ua)) ? 1 : 0;
    if (!acc) vmco_init_positions(w, old);
    free(old);
    return acc;
}

double vmco_lp_sample_grad(vmco_walker* w, const double* th, const vmco_lp_hyper* hp, int i,
                           double xi0, double xi1, double* grad, double* out)
{
    const double xi[2] = {xi0, xi1};
    const double omega = w->c.omega, s = sqrt(omega), ell = 1.0 / s;
    lp_pass q;
    lp_propose(w, th, hp, i, xi, &q);
    double* old = dalloc(w->ND);
    memcpy(old, w->x, sizeof(double) * w->ND);
    double F[2];
    double logR = lp_move(w, i, q.xn, q.yn, F);
    vmco_init_positions(w, old);
    free(old);

    double L = logR + q.lqr - q.lqf;
    double A = L < 0 ? exp(L) : 1.0;
    double d2 = (q.dx[0] * q.dx[0] + q.dx[1] * q.dx[1]) * omega;
    double J = A * d2;
    double gL = (L < 0 ? A : 0.0) * d2;                 /* dJ/dL */
    /* d log q_r / d o_b */
    double gob[VMCO_LP_OUT], gfb[VMCO_LP_IN];
    for (int d = 0; d < 2; d++) {
        double t = tanh(q.ob[d] / hp->mu_max), ts = tanh(q.ob[2 + d] / hp->ls_max);
        double is2 = 1.0 / (q.sigb[d] * q.sigb[d]);
        gob[d] = q.r[d] * is2 * ell * (1.0 - t * t);
        gob[2 + d] = (q.r[d] * q.r[d] * is2 - 1.0) * (1.0 - ts * ts);
    }
    lp_backward(th, q.fb, q.h1b, q.h2b, gob, gL, grad, gfb);
    /* dJ/d dx: explicit |dx|^2, and through L (psi at x', q_r explicit and via f_b) */
    double gdx[2];
    for (int d = 0; d < 2; d++) {
        double viaf = gfb[d];
        for (int m = 0; m < q.nnb; m++) viaf -= gfb[2 + 2 * m + d];
        double dLdx = F[d] + q.r[d] / (q.sigb[d] * q.sigb[d]) + s * viaf;
        gdx[d] = A * 2.0 * q.dx[d] * omega + gL * dLdx;
    }
    double goa[VMCO_LP_OUT];
    for (int d = 0; d < 2; d++) {
        double t = tanh(q.oa[d] / hp->mu_max), ts = tanh(q.oa[2 + d] / hp->ls_max);
        double dJdmu = gdx[d];
        double dJdsig = gdx[d] * xi[d] + gL / q.sig[d];
        goa[d] = dJdmu * ell * (1.0 - t * t);
        goa[2 + d] = dJdsig * q.sig[d] * (1.0 - ts * ts);
    }
    lp_backward(th, q.fa, q.h1a, q.h2a, goa, 1.0, grad, NULL);
    if (out) {
        out[0] = J;
        out[1] = A;
        out[2] = logR;
        out[3] = q.xn;
        out[4] = q.yn;
        out[5] = L;
    }
    return J;
}

/* ---- electron densities ----
 * r_i = sqrt(sum_d x_id^2) (ref Metropolis::calculateRadialVectorElement,
 * src/Metropolis/metropolis.cpp:31-39), dr = maxRadius / numberOfBins (ref
 * src/sampler.cpp:41). */
void vmco_density_add(const double* x, int N, int bins, double max_radius, int64_t* radial,
                      int64_t* spatial, int64_t* two_body)
{
    const double dr = max_radius / bins;
    const int half = bins / 2;   /* ref sampler.cpp:35 */
    for (int i = 0; i < N; i++) {
        const double ri = sqrt(x[2 * i] * x[2 * i] + x[2 * i + 1] * x[2 * i + 1]);
        /* radial one-body (ref sampler.cpp:639-652) */
        if (radial && ri < max_radius) {
            const int bin = (int)(ri / dr);
            if (bin < bins) radial[bin]++;
        }
        /* spatial one-body (ref sampler.cpp:658-673): indices by truncation of
         * x_d / (2 dr), row = y index + half, column = x index + half */
        if (spatial) {
            const double fx = x[2 * i] / (2 * dr), fy = x[2 * i + 1] / (2 * dr);
            if (fabs(fx) < 2.0e9 && fabs(fy) < 2.0e9) {
                const int ix = (int)fx, iy = (int)fy;
                if (abs(ix) < half && abs(iy) < half)
                    spatial[(int64_t)(iy + half) * bins + (ix + half)]++;
            }
        }
        /* radial two-body (ref sampler.cpp:679-692): bins shifted by one */
        if (two_body) {
            const double qi = ri / dr;
            if (qi >= 2.0e9) continue;
            const int bin_i = (int)qi + 1;
            for (int j = i + 1; j < N; j++) {
                const double rj = sqrt(x[2 * j] * x[2 * j] + x[2 * j + 1] * x[2 * j + 1]);
                const double qj = rj / dr;
                if (qj >= 2.0e9) continue;
                const int bin_j = (int)qj + 1;
                if (bin_i < bins && bin_j < bins) two_body[(int64_t)bin_i * bins + bin_j]++;
            }
        }
    }
}

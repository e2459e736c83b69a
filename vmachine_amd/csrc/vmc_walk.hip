// vmc_walk.hip -- the persistent Metropolis walk kernel for gfx950 (MI355X).
//
// One walker is a group of K lanes of a 64-lane wavefront (K = the smallest
// power of two >= NP, NP = particles per spin block), so a wave carries 64/K
// walkers.  Lane l of a walker owns
//   * particle l (slot 0) and particle NP+l (slot 1): positions, Slater
//     gradient and second-derivative sums, Pade gradient, stale drift;
//   * column l of the spin-up and of the spin-down Slater inverse, held in
//     registers for the whole launch (so a Sherman-Morrison update, a ratio
//     and the derivative refresh of particle l touch no memory at all).
// Pair terms of a moved particle are lane-parallel over the other particles;
// sums over a walker are butterfly reductions inside the K-lane group.  The
// only HBM traffic is the walker's positions at launch start and end and one
// row of per-walker accumulators.
//
// The walk reproduces, per step, the reference's single-coordinate Metropolis
// step and its per-step sampling:
//   ref src/system.cpp:147-164            System::runMetropolisCycles
//   ref src/Metropolis/importancesampling.cpp:34-81, bruteforce.cpp:26-52
//   ref src/elementoperation.cpp:80-123   ratio, kinetic energy, parameter gradients
//   ref src/WaveFunctions/gaussian.cpp, slaterdeterminant.cpp, padejastrow.cpp
//   ref src/Hamiltonians/harmonicoscillator.cpp:16-20, src/Interaction/coulomb.cpp:17-48
//   ref src/sampler.cpp:52-90              per-step accumulation and gradient window
// The Slater inverse is built in FP64 by Gauss-Jordan elimination with partial
// pivoting at the start of every launch (the host splits an iteration into
// launches of `refresh` steps); in between it is kept by Sherman-Morrison
// updates in the walk precision R.  The orbitals are
// the normalised Hermite polynomials h_n(sqrt(omega) x): for closed shells they
// span the same polynomial space as the reference's H_n(x) (ref
// src/Basis/hermite.cpp), so every ratio and log-derivative is identical, and
// they stay well conditioned at small omega.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "vmc_device.h"

namespace vmc {

// ------------------------------------------------------------------------
// compile-time tables
// ------------------------------------------------------------------------
constexpr double csqrt_iter(double x, double cur, double prev)
{
    return cur == prev ? cur : csqrt_iter(x, 0.5 * (cur + x / cur), cur);
}
constexpr double csqrt(double x) { return x <= 0 ? 0.0 : csqrt_iter(x, x > 1 ? x : 1.0, 0.0); }

// degrees 0..17, as the reference's table (ref src/Basis/hermite.cpp:342-386)
constexpr int kHermiteN = 18;
struct HermiteConst {
    double c1[kHermiteN];  // sqrt(2/(m+1))          h_{m+1} = c1 t h_m - c2 h_{m-1}
    double c2[kHermiteN];  // sqrt(m/(m+1))
    double d1[kHermiteN];  // sqrt(2m):      h_m'  = d1 h_{m-1}
    double d2[kHermiteN];  // 2 sqrt(m(m-1)): h_m'' = d2 h_{m-2}
};
constexpr HermiteConst make_hermite_const()
{
    HermiteConst h{};
    for (int m = 0; m < kHermiteN; m++) {
        h.c1[m] = csqrt(2.0 / (m + 1));
        h.c2[m] = csqrt(double(m) / (m + 1));
        h.d1[m] = csqrt(2.0 * m);
        h.d2[m] = 2.0 * csqrt(double(m) * (m - 1 > 0 ? m - 1 : 0));
    }
    return h;
}
constexpr HermiteConst kHC = make_hermite_const();

// closed-shell 2-D orbital list: shell i holds (n_x, n_y) = (i-j, j), j = 0..i
// (ref src/Basis/basis.cpp:125-143)
constexpr int orb_shell(int k)
{
    int i = 0;
    while ((i + 1) * (i + 2) / 2 <= k) i++;
    return i;
}
constexpr int orb_ny(int k) { return k - orb_shell(k) * (orb_shell(k) + 1) / 2; }
constexpr int orb_nx(int k) { return orb_shell(k) - orb_ny(k); }
constexpr int max_degree(int np) { return np <= 1 ? 0 : orb_shell(np - 1); }

// ------------------------------------------------------------------------
// precision-dependent math
// ------------------------------------------------------------------------
template <typename R> struct M;
template <> struct M<float> {
    static __device__ __forceinline__ float sqrt(float x) { return __builtin_sqrtf(x); }
    // the hardware square root and base-2 logarithm (v_sqrt_f32, v_log_f32, about
    // 1 ulp) for the per-step distances, drift limits and acceptance tests; the
    // correctly rounded forms above stay in the draws and the rebuilds
    static __device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ __forceinline__ float flog(float x)
    {
        return __builtin_amdgcn_logf(x) * 0.693147180559945309417f;
    }
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    // distance r and its inverse from r^2 with one reciprocal square root
    // (v_rsq_f32) instead of a square root and a reciprocal; a masked pair
    // (val false: the particle itself, padding lanes) gets r = 1/r = 0
    static __device__ __forceinline__ void dist(float r2, bool val, float& r, float& ir)
    {
        ir = val ? __builtin_amdgcn_rsqf(r2) : 0.f;
        r = r2 * ir;
    }
    static __device__ __forceinline__ float exp(float x) { return __expf(x); }
    static __device__ __forceinline__ float log(float x) { return __logf(x); }
    // v_cos_f32 takes its argument in revolutions: cos(2 pi u)
    static __device__ __forceinline__ float cos2pi(float u) { return __builtin_amdgcn_cosf(u); }
};
template <> struct M<double> {
    static __device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
    static __device__ __forceinline__ double fsqrt(double x) { return ::sqrt(x); }
    static __device__ __forceinline__ double flog(double x) { return ::log(x); }
    static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
    static __device__ __forceinline__ void dist(double r2, bool val, double& r, double& ir)
    {
        r = ::sqrt(r2);
        ir = val ? 1.0 / r : 0.0;
    }
    static __device__ __forceinline__ double exp(double x) { return ::exp(x); }
    static __device__ __forceinline__ double log(double x) { return ::log(x); }
    static __device__ __forceinline__ double cos2pi(double u) { return ::cos(6.283185307179586476925286766559 * u); }
};

// normalised Hermite values h_0..h_S at t
template <typename R, int S>
__device__ __forceinline__ void hermite(R t, R (&h)[S + 1])
{
    h[0] = R(1);
#pragma unroll
    for (int m = 0; m < S; m++)
        h[m + 1] = R(kHC.c1[m]) * t * h[m] - (m ? R(kHC.c2[m]) * h[m - 1] : R(0));
}

// ------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) and the engine's draw protocol
// ------------------------------------------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
}

// 23-bit uniform in (0,1): (m + 1/2) 2^-23 with m < 2^23 needs 24 significant
// bits, so it is exact in float and double alike
template <typename R>
__device__ __forceinline__ R u23(uint32_t u)
{
    return (R(u >> 9) + R(0.5)) * R(1.0 / 8388608.0);
}

// step draws: coordinate, standard normal (Box-Muller), move uniform, accept uniform,
// and the Box-Muller partner normal z2 (used by particle moves)
template <typename R>
__device__ __forceinline__ void step_draws(uint64_t seed, uint32_t walker, uint64_t counter,
                                           int nd, int& k, R& z, R& um, R& ua, R& z2,
                                           uint32_t stream = 0u)
{
    uint32_t c[4] = {(uint32_t)counter, (uint32_t)(counter >> 32), walker, stream};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    k = (int)__umulhi(c[0], (uint32_t)nd);
    const R U1 = u23<R>(c[1]), U2 = u23<R>(c[2]);
    const R r = M<R>::sqrt(R(-2) * M<R>::log(U1));
    z = r * M<R>::cos2pi(U2);
    z2 = r * M<R>::cos2pi(U2 - R(0.25));   // sin(2 pi U2)
    um = U1;
    ua = u23<R>(c[3]);
}

// ------------------------------------------------------------------------
// LDS exchange between the lanes of one wave.  Every kernel that walks runs
// one-wave workgroups (launch bounds 64), and one wave's LDS instructions
// are performed in the order it issues them, so a store by one lane is seen
// by a later load of another lane with no wait in between.  The fences at
// wavefront scope emit no instruction; with the wave barrier they keep the
// compiler from moving the loads above the stores.  A workgroup barrier here
// would first wait for every outstanding LDS store to complete.  The
// Gauss-Jordan rebuild keeps the barrier: its pivot-row exchange measured
// 0.5 % slower without the wait, the network's exchanges 4 % faster.
// ------------------------------------------------------------------------
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------------------
// K-lane group collectives
// ------------------------------------------------------------------------
// data-parallel-primitive lane permutations (one VALU op, no LDS crossbar):
// quad_perm(1,0,3,2), quad_perm(2,3,0,1), row_half_mirror, row_mirror
enum : int { kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140 };

template <int CTRL>
__device__ __forceinline__ float dpp(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL,
                                                              0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), CTRL, 0xF, 0xF, true);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

template <int CTRL>
__device__ __forceinline__ int dpp(int v)
{
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ uint64_t dpp(uint64_t v)
{
    const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)v, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(v >> 32), CTRL, 0xF, 0xF, true);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// sum over the K lanes of a group, every lane receives it.  Each step adds the
// value of a partner lane, and partners pair up symmetrically, so all lanes
// form the same floating-point sums.  Steps 1, 2, 4, 8 use DPP (after the
// quad steps, the half mirror pairs the two quads of 8 lanes and the mirror
// the two halves of 16); 16 and 32 cross DPP rows and use a lane shuffle.
template <int K, typename T>
__device__ __forceinline__ T gsum(T v)
{
    if constexpr (K >= 2) v += dpp<kDppXor1>(v);
    if constexpr (K >= 4) v += dpp<kDppXor2>(v);
    if constexpr (K >= 8) v += dpp<kDppHalfMirror>(v);
    if constexpr (K >= 16) v += dpp<kDppMirror>(v);
    if constexpr (K >= 32) v += __shfl_xor(v, 16, 64);
    if constexpr (K >= 64) v += __shfl_xor(v, 32, 64);
    return v;
}

// product over the K lanes of a group (same pairing as gsum)
template <int K, typename T>
__device__ __forceinline__ T gprod(T v)
{
    if constexpr (K >= 2) v *= dpp<kDppXor1>(v);
    if constexpr (K >= 4) v *= dpp<kDppXor2>(v);
    if constexpr (K >= 8) v *= dpp<kDppHalfMirror>(v);
    if constexpr (K >= 16) v *= dpp<kDppMirror>(v);
    if constexpr (K >= 32) v *= __shfl_xor(v, 16, 64);
    if constexpr (K >= 64) v *= __shfl_xor(v, 32, 64);
    return v;
}

// the values of lanes l and l ^ 16 (and of l and l ^ 32) by a VALU lane swap
// between neighbouring 16-lane rows (v_permlane16_swap_b32; v_permlane32_swap_b32
// between the wave's halves): swapping v with itself leaves, in every lane,
// its own value in one result and its partner's in the other, in an order
// that depends on the row, so only symmetric operations of the two may follow.
// Lower latency than a lane shuffle through the LDS crossbar, for dependent
// chains; all lanes must be active.
__device__ __forceinline__ void xchg16(uint32_t v, uint32_t& a, uint32_t& b)
{
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    a = r[0];
    b = r[1];
}
__device__ __forceinline__ void xchg32(uint32_t v, uint32_t& a, uint32_t& b)
{
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    a = r[0];
    b = r[1];
}

template <int K>
__device__ __forceinline__ uint32_t gmax32(uint32_t v)
{
    auto take = [&](uint32_t o) { v = o > v ? o : v; };
    if constexpr (K >= 2) take((uint32_t)dpp<kDppXor1>((int)v));
    if constexpr (K >= 4) take((uint32_t)dpp<kDppXor2>((int)v));
    if constexpr (K >= 8) take((uint32_t)dpp<kDppHalfMirror>((int)v));
    if constexpr (K >= 16) take((uint32_t)dpp<kDppMirror>((int)v));
    uint32_t a, b;
    if constexpr (K >= 32) {
        xchg16(v, a, b);
        v = a > b ? a : b;
    }
    if constexpr (K >= 64) {
        xchg32(v, a, b);
        v = a > b ? a : b;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T vsel(bool c, T a, T b) { return c ? a : b; }

// compile-time loop: f(std::integral_constant<int, i>) for i = B..E-1, so that
// register arrays are only ever indexed by constants
template <int B, int E, typename F>
__device__ __forceinline__ void sfor(F&& f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>());
        sfor<B + 1, E>(f);
    }
}

// ------------------------------------------------------------------------
// 16 x 16 matrix-core tiles (CDNA3/4 lane layout): element e of lane t holds
// contraction index mf_index(q, e), q = t / 16, of row / column t % 16 of the
// A / B operand; result element r of lane t is row 4q + r, column t % 16.
// MODE 1: FP32 inputs, eight V_MFMA_F32_16X16X4_F32 (index 4e + q);
// MODE 2: BF16 inputs (round to nearest even), one V_MFMA_F32_16X16X32_BF16
// (index 8q + e).  Both accumulate in FP32.
// ------------------------------------------------------------------------
typedef float mf_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 mf_b8 __attribute__((ext_vector_type(8)));

// LDS row strides of the RBM state, padded so that the lanes of a tile operand
// (one row each) fall into different banks: the walk's copy of W, [32][4K + 1]
// (unit h = l + K j of row k at k (4K + 1) + 4 l + j); the tile results, one row
// of 36 per walker (16-byte aligned rows for the four-element stores); the
// FP64 window sums of xs_k n_h, row k of H + 1
template <int K>
__host__ __device__ constexpr int rbm_wst() { return 4 * K + 1; }
constexpr int kMfX = 36;
__host__ __device__ constexpr int rbm_hst(int H) { return H + 1; }

template <int MODE>
__device__ __forceinline__ int mf_index(int q, int e) { return MODE == 1 ? 4 * e + q : 8 * q + e; }

template <int MODE>
__device__ __forceinline__ mf_f4 mf_tile(const float (&a)[8], const float (&b)[8])
{
    mf_f4 c = {0.f, 0.f, 0.f, 0.f};
    if constexpr (MODE == 1) {
#pragma unroll
        for (int e = 0; e < 8; e++) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[e], b[e], c, 0, 0, 0);
    } else {
        mf_b8 av, bv;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            av[e] = (__bf16)a[e];
            bv[e] = (__bf16)b[e];
        }
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, c, 0, 0, 0);
    }
    return c;
}

// ------------------------------------------------------------------------
// orbital rows of one particle: h_n(t) of both coordinates and their first
// derivatives h_n' = d1_n h_{n-1} (dx[0] = dy[0] = 0), t = sqrt(omega) x
// ------------------------------------------------------------------------
template <typename R, int S>
struct OrbRows {
    R hx[S + 1], hy[S + 1], dx[S + 1], dy[S + 1];
};

template <typename R, int S>
__device__ __forceinline__ void orb_rows(R t0, R t1, OrbRows<R, S>& o)
{
    hermite<R, S>(t0, o.hx);
    hermite<R, S>(t1, o.hy);
#pragma unroll
    for (int i = 0; i <= S; i++) {
        o.dx[i] = i >= 1 ? R(kHC.d1[i]) * o.hx[i >= 1 ? i - 1 : 0] : R(0);
        o.dy[i] = i >= 1 ? R(kHC.d1[i]) * o.hy[i >= 1 ? i - 1 : 0] : R(0);
    }
}

// ------------------------------------------------------------------------
// walker state
// ------------------------------------------------------------------------
template <typename R, int NP, int K, bool RB>
struct Walker {
    static constexpr int S = max_degree(NP);
    static constexpr int G = 64 / K;
    static constexpr int HJ = 4;          // RBM hidden units per lane: h = l + K j, j < HJ

    // owned particles: slot 0 = particle l, slot 1 = particle NP + l
    R x[2][2];
    // column l of the two Slater inverses: a0 holds the block of spin `act`
    // (0 up, 1 down), a1 the other one.  A step brings the moved block into a0
    // first (a register swap, needed on about half the steps), so the ratio,
    // the Sherman-Morrison update and the derivative refresh address a0 only.
    R a0[NP], a1[NP];
    int act;
    R gS[2][2];            // d ln det / dx of the owned particles
    R sS[2];               // sum_d d^2 det / dx_d^2 / det of the owned particles
    R gJ[2][2];            // d ln J / dx of the owned particles (Pade)
    R qf[2][2];            // stale quantum force (importance compat)
    double sumx2, vint, lj, jg;   // walker sums: sum x^2, Coulomb, Pade Laplacian, sum a h^2
    double el, tk, vext, ga, gg;  // local energy parts and parameter gradients of the current state
    bool valid[2];
    bool upspin[2];
    int l, gbase, g;
    double th_alpha, th_gamma;   // Gaussian alpha and Pade gamma of the walk
    R pa_same, pa_anti;          // Pade a of a spin-parallel / antiparallel pair

    // RBM state (RB only).  Lane l owns hidden units h = l + K j; the sorted
    // coordinate vector of the walker and its n = sigmoid(v) live in LDS.
    R rv[HJ], rn[HJ], rp[HJ];   // pre-activation v, n, p = 1 - n of the current state
    R pv[HJ], pn[HJ], pp[HJ];   // the same for the pending proposal
    R* lbc;                     // b_h and sum_k W[k][h]^2 of the lane's units in LDS, [2][K][4]
    __device__ __forceinline__ R rbv(int j) const { return lbc[4 * l + j]; }
    __device__ __forceinline__ R rcv(int j) const { return lbc[4 * K + 4 * l + j]; }
    R ra[2][2];                 // RBM-Gaussian a of the owned coordinates
    R gR[2][2];                 // RBM-product d ln psi / dx of the owned coordinates
    int rk[2][2];               // rank of each owned coordinate in the sorted vector
    double lr;                  // RBM-product Laplacian of ln psi (walker total)
    int mk, r_old, r_new;       // the pending proposal: coordinate, its old and new rank
    R* lW;                      // W[k][l' + K j] at lW[k * 4K + 4 l' + j]
    R* xs_cur;                  // this walker's sorted coordinates
    R* xs_prop;                 // the proposal's sorted coordinates (scratch)
    R* ln;                      // this walker's n_h, h = 0..H-1
    R* mf_x;                    // [G][kMfX] matrix-core results of the wave's walkers
    R* mf_xs;                   // every walker's sorted vectors, [G][2][ND]
    R* mf_n;                    // every walker's n, [G][H]

    __device__ __forceinline__ bool hvalid(const WalkParams& P, int j) const { return l + K * j < P.H; }
    __device__ __forceinline__ R isp2(const WalkParams& P) const { return R(P.isig2_p); }

    __device__ __forceinline__ R sw(const WalkParams& P) const { return R(P.sqrt_omega); }

    // Slater log-derivatives of a particle with orbital rows `o` from its
    // column `col` of the block inverse
    __device__ __forceinline__ void slater_derivs_rows(const WalkParams& P, const OrbRows<R, S>& o,
                                                       const R (&col)[NP], R& gx, R& gy, R& sec) const
    {
        const R s = sw(P);
        // second-derivative rows h_n'' = d2_n h_{n-2}
        R ddx[S + 1], ddy[S + 1];
#pragma unroll
        for (int i = 0; i <= S; i++) {
            ddx[i] = i >= 2 ? R(kHC.d2[i]) * o.hx[i >= 2 ? i - 2 : 0] : R(0);
            ddy[i] = i >= 2 ? R(kHC.d2[i]) * o.hy[i >= 2 ? i - 2 : 0] : R(0);
        }
        R u0[S + 1], u1[S + 1], u2[S + 1];
#pragma unroll
        for (int i = 0; i <= S; i++) u0[i] = u1[i] = u2[i] = R(0);
#pragma unroll
        for (int k = 0; k < NP; k++) {
            const int nx = orb_nx(k), ny = orb_ny(k);
            const R B = col[k];
            u0[nx] += B * o.hy[ny];
            if (ny >= 1) u1[nx] += B * o.dy[ny];
            if (ny >= 2) u2[nx] += B * ddy[ny];
        }
        R ax = R(0), ay = R(0), ss = R(0);
#pragma unroll
        for (int i = 0; i <= S; i++) {
            if (i >= 1) ax += o.dx[i] * u0[i];
            if (i >= 2) ss += ddx[i] * u0[i];
            ay += o.hx[i] * u1[i];
            ss += o.hx[i] * u2[i];
        }
        gx = s * ax;
        gy = s * ay;
        sec = s * s * ss;
    }

    // the same from the particle's position (px, py)
    __device__ __forceinline__ void slater_derivs(const WalkParams& P, R px, R py,
                                                  const R (&col)[NP], R& gx, R& gy, R& sec) const
    {
        const R s = sw(P);
        OrbRows<R, S> o;
        orb_rows<R, S>(s * px, s * py, o);
        slater_derivs_rows(P, o, col, gx, gy, sec);
    }

    // FP64 Gauss-Jordan inverses of NBK spin blocks (b0, b0 + 1, ...) with
    // partial pivoting, the pivot chains of the blocks interleaved; leaves
    // column l of the inverse of block b0 in out0 (and of block 1 in out1 when
    // NBK = 2).  Lane l holds row l of each block (orbital values of particle
    // b NP + l).  lds_row: [2 buffers][2 blocks][G][NP] pivot rows, then
    // [2][G][NP] pivot lanes as int.
    template <int NBK>
    __device__ __forceinline__ void invert_blocks(const WalkParams& P, int b0, double* lds_T,
                                                  double* lds_row, R (&out0)[NP], R (&out1)[NP])
    {
        const double s = P.sqrt_omega;
        double a[NBK][NP];
#pragma unroll
        for (int b = 0; b < NBK; b++) {
            const double t0 = s * double(vsel(b0 + b != 0, x[1][0], x[0][0]));
            const double t1 = s * double(vsel(b0 + b != 0, x[1][1], x[0][1]));
            double hx[S + 1], hy[S + 1];
            hermite<double, S>(t0, hx);
            hermite<double, S>(t1, hy);
#pragma unroll
            for (int k = 0; k < NP; k++) a[b][k] = hx[orb_nx(k)] * hy[orb_ny(k)];
        }
        int* piv = reinterpret_cast<int*>(lds_row + 4 * G * NP);   // [2][G][NP]
        bool used[NBK];
        int mycol[NBK];
#pragma unroll
        for (int b = 0; b < NBK; b++) {
            used[b] = l >= NP;
            mycol[b] = 0;
        }
        sfor<0, NP>([&](auto ci) {
            constexpr int c = decltype(ci)::value;
            // the pivot: the group maximum of one 32-bit key per lane, the high
            // word of |a| (sign, exponent and 20 mantissa bits, which order as
            // the value) with the low 5 bits (6 for 64 lanes) replaced by the
            // lane, found with DPP steps; magnitudes within 2^-15 (2^-14) of
            // each other count as equal
            constexpr uint32_t kLaneMask = K > 32 ? 63u : 31u;
            int idx[NBK];
#pragma unroll
            for (int b = 0; b < NBK; b++) {
                const uint32_t hiw = (uint32_t)(__builtin_bit_cast(uint64_t, fabs(a[b][c])) >> 32);
                const uint32_t key = used[b] ? 0u : ((hiw & ~kLaneMask) | (uint32_t)l);
                idx[b] = (int)(gmax32<K>(key) & kLaneMask);
            }
            if constexpr (NBK == 2) {
                // both blocks' pivot rows leave in one set of (masked) LDS
                // writes: each pivot lane writes its block's row; a lane that
                // pivots both blocks writes the second row separately
                double* row0 = lds_row + (((c & 1) * 2 + b0) * G + g) * NP;
                double* row1 = lds_row + (((c & 1) * 2 + b0 + 1) * G + g) * NP;
                const bool w0 = l == idx[0], w1 = l == idx[1];
                if (w0 || w1) {
                    double* dst = w0 ? row0 : row1;
#pragma unroll
                    for (int j = 0; j < NP; j++) dst[j] = w0 ? a[0][j] : a[1][j];
                    piv[((b0 + (w0 ? 0 : 1)) * G + g) * NP + c] = w0 ? idx[0] : idx[1];
                }
                if (w0 && w1) {
#pragma unroll
                    for (int j = 0; j < NP; j++) row1[j] = a[1][j];
                    piv[((b0 + 1) * G + g) * NP + c] = idx[1];
                }
            } else {
#pragma unroll
                for (int b = 0; b < NBK; b++) {
                    double* row = lds_row + (((c & 1) * 2 + b0 + b) * G + g) * NP;
                    if (l == idx[b]) {
#pragma unroll
                        for (int j = 0; j < NP; j++) row[j] = a[b][j];
                        piv[((b0 + b) * G + g) * NP + c] = idx[b];
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int b = 0; b < NBK; b++) {
                const double* row = lds_row + (((c & 1) * 2 + b0 + b) * G + g) * NP;
                // reciprocal of the pivot: hardware estimate and one Newton step
                const double pv = row[c];
                double ip = __builtin_amdgcn_rcp(pv);
                ip = fma(ip, fma(-pv, ip, 1.0), ip);
                // one update for every lane: the pivot lane's row equals `row`,
                // so a - (1 - ip) row scales it by ip; the others eliminate
                // column c with f = a[c] ip
                const bool me = l == idx[b];
                const double f = me ? 1.0 - ip : a[b][c] * ip;
                const double cv = me ? ip : -f;
                if (me) {
                    used[b] = true;
                    mycol[b] = c;
                }
#pragma unroll
                for (int j = 0; j < NP; j++) a[b][j] = (j == c) ? cv : a[b][j] - f * row[j];
            }
        });
        // lane piv[c] now holds row c of the inverse with its columns permuted:
        // a[j] = inv[c][piv[j]].  Transpose through LDS so lane q gets inv[.][q]
        // (rows of NP + 1: the lanes' reads of one column fall into different banks)
        constexpr int TS = NP + 1;
        double* T = lds_T + g * NP * TS;
        auto transpose = [&](int b, R(&out)[NP]) {
            if (l < NP) {
#pragma unroll
                for (int j = 0; j < NP; j++)
                    T[piv[((b0 + b) * G + g) * NP + j] * TS + mycol[b]] = a[b][j];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < NP; k++) out[k] = l < NP ? R(T[l * TS + k]) : R(0);
            __syncthreads();
        };
        transpose(0, out0);
        if constexpr (NBK == 2) transpose(1, out1);
    }

    __device__ __forceinline__ void rebuild_slater(const WalkParams& P, double* lds_T, double* lds_row)
    {
        if constexpr (NP > 28) {
            // one block at a time: two interleaved FP64 blocks of NP > 28 rows
            // would not fit the registers
            invert_blocks<1>(P, 0, lds_T, lds_row, a0, a0);
            __syncthreads();   // the first transpose's reads are done
            invert_blocks<1>(P, 1, lds_T, lds_row, a1, a1);
        } else {
            invert_blocks<2>(P, 0, lds_T, lds_row, a0, a1);
        }
        act = 0;
        R gx, gy, sec;
        slater_derivs(P, x[0][0], x[0][1], a0, gx, gy, sec);
        gS[0][0] = gx; gS[0][1] = gy; sS[0] = sec;
        slater_derivs(P, x[1][0], x[1][1], a1, gx, gy, sec);
        gS[1][0] = gx; gS[1][1] = gy; sS[1] = sec;
        if (l >= NP) {
            gS[0][0] = gS[0][1] = gS[1][0] = gS[1][1] = R(0);
            sS[0] = sS[1] = R(0);
        }
    }

    // Pade gradients and the walker sums, from scratch (O(N^2) per walker)
    __device__ __forceinline__ void rebuild_pairs(const WalkParams& P)
    {
        const R gam = R(th_gamma);
        const R a_same = R(P.a_same), a_anti = R(P.a_anti);
        double vi = 0, ljs = 0, jgs = 0;
#pragma unroll
        for (int s = 0; s < 2; s++) gJ[s][0] = gJ[s][1] = R(0);
        sfor<0, 2 * NP>([&](auto qi) {
            constexpr int qq = decltype(qi)::value;
            constexpr int own = qq % NP, sl = qq / NP;
            const R y0 = __shfl(x[sl][0], gbase + own, 64);
            const R y1 = __shfl(x[sl][1], gbase + own, 64);
            if (qq >= P.N) return;
            const bool qq_up = qq < P.n_up;
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const int q = s * NP + l;
                if (!valid[s] || q == qq) continue;
                const R e0 = x[s][0] - y0, e1 = x[s][1] - y1;
                R r, ir;
                M<R>::dist(e0 * e0 + e1 * e1, true, r, ir);
                const R a = (upspin[s] == qq_up) ? a_same : a_anti;
                const R f = M<R>::rcp(R(1) + gam * r);
                const R u1 = a * f * f;
                if (P.has_pade) {
                    gJ[s][0] += u1 * e0 * ir;
                    gJ[s][1] += u1 * e1 * ir;
                }
                if (qq > q) {
                    vi += double(ir);
                    ljs += double(R(2) * u1 * (ir - R(2) * gam * f));
                    const R h = r * f;
                    jgs += double(a * h * h);
                }
            }
        });
        vint = P.coulomb ? gsum<K>(vi) : 0.0;
        lj = P.has_pade ? gsum<K>(ljs) : 0.0;
        jg = P.has_pade ? gsum<K>(jgs) : 0.0;
        if (!P.has_pade) {
#pragma unroll
            for (int s = 0; s < 2; s++) gJ[s][0] = gJ[s][1] = R(0);
        }
        double sx = 0;
#pragma unroll
        for (int s = 0; s < 2; s++)
            if (valid[s]) sx += double(x[s][0]) * double(x[s][0]) + double(x[s][1]) * double(x[s][1]);
        sumx2 = gsum<K>(sx);
    }

    // local energy of the current state (ref src/elementoperation.cpp:95-109, src/sampler.cpp:66-70)
    __device__ __forceinline__ void local_energy(const WalkParams& P)
    {
        const R oa = R(P.omega * th_alpha);
        // the lane terms and their group sum in the walk's precision (FP32
        // values of O(1) each: the sum over 28 lanes keeps 1e-7 relative);
        // the walker totals and the accumulation stay FP64
        R t = R(0);
#pragma unroll
        for (int s = 0; s < 2; s++) {
            if (!valid[s]) continue;
            // lap ln det + |grad ln psi|^2 with the |grad ln det|^2 parts cancelled
            // analytically: sS - gS^2 + (gS + o)^2 = sS + 2 gS.o + o^2, o = the
            // other elements' gradient.  Near a node gS ~ 1/distance, and the
            // explicit difference would lose every digit in FP32.
            R o0 = R(0), o1 = R(0);
            if (P.has_pade) {
                o0 += gJ[s][0];
                o1 += gJ[s][1];
            }
            if (P.has_gauss) {
                o0 -= oa * x[s][0];
                o1 -= oa * x[s][1];
            }
            if constexpr (RB) {
                if (P.has_rbmp) {
                    o0 += gR[s][0];
                    o1 += gR[s][1];
                }
                if (P.has_rbmg) {
                    const R is2 = R(1.0 / P.sig2_g);
                    o0 -= (x[s][0] - ra[s][0]) * is2;
                    o1 -= (x[s][1] - ra[s][1]) * is2;
                }
            }
            R v = o0 * o0 + o1 * o1;
            if (P.has_slater) v += sS[s] + R(2) * (gS[s][0] * o0 + gS[s][1] * o1);
            t += v;
        }
        double tt = double(gsum<K>(t));
        if (P.has_pade) tt += lj;
        if (P.has_gauss) tt -= P.omega * th_alpha * 2.0 * P.N;
        if constexpr (RB) {
            if (P.has_rbmp) tt += lr;
            if (P.has_rbmg) tt -= double(P.ND) / P.sig2_g;
        }
        tk = -0.5 * tt;
        vext = 0.5 * P.omega * P.omega * sumx2;
        el = tk + vext + vint;
        ga = P.has_gauss ? -0.5 * P.omega * sumx2 : 0.0;
        gg = P.has_pade ? (P.q3_half ? -0.5 : -1.0) * jg : 0.0;
    }

    // ---- RBM elements (ref src/WaveFunctions/rbmgaussian.cpp, rbmproduct.cpp) ----

    // RBM-product contractions on the matrix cores (FP32 walks, K >= 4, ND and
    // H <= 32).  The wave's G <= 16 walkers are the columns of 16 x 16 tiles;
    // hidden units (pre-activation) or coordinates (gradient) are the rows.
    __device__ __forceinline__ bool mf_on(const WalkParams& P) const
    {
        if constexpr (sizeof(R) == 4 && K >= 4) return P.rbm_mf != 0;
        return false;
    }

    // the pre-activation sums sum_k W[k][h] x'_k of the proposal's sorted
    // vector (xs_prop, all walkers staged) for the lane's hidden units
    template <int MODE>
    __device__ __forceinline__ void rbm_mf_pre(const WalkParams& P, R (&acc)[HJ]) const
    {
        if constexpr (sizeof(R) == 4 && K >= 4) {
            const int t = threadIdx.x & 63, c = t & 15, q = t >> 4;
            const int wv = c < G ? c : G - 1;
            const R* xp = mf_xs + wv * 2 * P.ND + P.ND;
            float b[8];
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int k = mf_index<MODE>(q, e);
                const float v = xp[k < P.ND ? k : P.ND - 1];
                b[e] = (c < G && k < P.ND) ? v : 0.f;
            }
#pragma unroll
            for (int th = 0; th < 2; th++)
                if (16 * th < P.H) {
                    const int h = 16 * th + c;   // < 4K: H <= 4K
                    const int ho = 4 * (h % K) + h / K;
                    float a[8];
#pragma unroll
                    for (int e = 0; e < 8; e++) a[e] = lW[mf_index<MODE>(q, e) * rbm_wst<K>() + ho];
                    const mf_f4 d = mf_tile<MODE>(a, b);
                    if (c < G) *reinterpret_cast<mf_f4*>(mf_x + c * kMfX + 16 * th + 4 * q) = d;
                }
            wave_sync();
#pragma unroll
            for (int j = 0; j < HJ; j++) {
                const int h = l + K * j;
                acc[j] = hvalid(P, j) ? mf_x[g * kMfX + (h < 32 ? h : 31)] : R(0);
            }
            wave_sync();
        }
    }

    // the gradient sums sum_h W[k][h] n_h of the staged n (all walkers) for
    // the lane's owned coordinates
    template <int MODE>
    __device__ __forceinline__ void rbm_mf_grad(const WalkParams& P)
    {
        if constexpr (sizeof(R) == 4 && K >= 4) {
            const int t = threadIdx.x & 63, c = t & 15, q = t >> 4;
            const int wv = c < G ? c : G - 1;
            const R* np_ = mf_n + wv * P.H;
            float b[8];
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int h = mf_index<MODE>(q, e);
                const float v = np_[h < P.H ? h : P.H - 1];
                b[e] = (c < G && h < P.H) ? v : 0.f;
            }
#pragma unroll
            for (int tk = 0; tk < 2; tk++)
                if (16 * tk < P.ND) {
                    const int k = 16 * tk + c;   // rows up to 32 exist (zero from ND on)
                    float a[8];
#pragma unroll
                    for (int e = 0; e < 8; e++) {
                        const int h = mf_index<MODE>(q, e);
                        const int hc = h < 4 * K ? h : 4 * K - 1;
                        const float v = lW[k * rbm_wst<K>() + 4 * (hc % K) + hc / K];
                        a[e] = h < 4 * K ? v : 0.f;
                    }
                    const mf_f4 d = mf_tile<MODE>(a, b);
                    if (c < G) *reinterpret_cast<mf_f4*>(mf_x + c * kMfX + 16 * tk + 4 * q) = d;
                }
            wave_sync();
#pragma unroll
            for (int s = 0; s < 2; s++)
#pragma unroll
                for (int d = 0; d < 2; d++)
                    {
                    const int k = 2 * (s * NP + l) + d;
                    gR[s][d] = valid[s] ? isp2(P) * mf_x[g * kMfX + (k < 32 ? k : 31)] : R(0);
                }
            wave_sync();
        }
    }

    // parameters of the lane's hidden units and owned coordinates; W into LDS
    // (block-wide, before any other RBM call)
    __device__ __forceinline__ void rbm_setup(const WalkParams& P)
    {
        const int ND = P.ND, H = P.H, K4 = 4 * K, WS = rbm_wst<K>();
        if (P.has_rbmp) {
            const int rows = ND > 32 ? ND : 32;   // zero rows pad the matrix-core tiles
            for (int i = threadIdx.x; i < rows * WS; i += 64) {
                const int k = i / WS, r = i - k * WS;
                const int h = (r >> 2) + K * (r & 3);
                lW[i] = (r < K4 && h < H && k < ND) ? R(P.rbm_W[k * H + h]) : R(0);
            }
#pragma unroll
            for (int j = 0; j < HJ; j++) {
                const int h = l + K * j;
                double c = 0;
                if (h < H)
                    for (int k = 0; k < ND; k++) c += P.rbm_W[k * H + h] * P.rbm_W[k * H + h];
                if (g == 0) {   // the same for every walker of the wave
                    lbc[4 * l + j] = h < H ? R(P.rbm_b[h]) : R(0);
                    lbc[4 * K + 4 * l + j] = R(c);
                }
            }
        }
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int d = 0; d < 2; d++)
                ra[s][d] = (P.has_rbmg && valid[s]) ? R(P.rbm_a[2 * (s * NP + l) + d]) : R(0);
    }

    // v, n, p of the lane's units from the current sorted vector
    __device__ __forceinline__ void rbm_activate(const WalkParams& P)
    {
        const int K4 = rbm_wst<K>();
        R acc[HJ];
#pragma unroll
        for (int j = 0; j < HJ; j++) acc[j] = R(0);
        for (int i = 0; i < P.ND; i++) {
            const R xi = xs_cur[i];
#pragma unroll
            for (int j = 0; j < HJ; j++) acc[j] += lW[i * K4 + 4 * l + j] * xi;
        }
#pragma unroll
        for (int j = 0; j < HJ; j++) {
            const bool ok = hvalid(P, j);
            rv[j] = rbv(j) + isp2(P) * acc[j];
            const R e = M<R>::exp(rv[j]);
            rp[j] = ok ? M<R>::rcp(R(1) + e) : R(0);
            rn[j] = ok ? e * rp[j] : R(0);
        }
    }

    // RBM-product gradient of the owned coordinates, d ln psi/dx_k =
    // sum_h W[k][h] n_h / sigma_p^2 with the unsorted row k (ref rbmproduct.cpp:62-70),
    // and the Laplacian sum_h (sum_k W[k][h]^2) n_h p_h / sigma_p^4 (ref :72-75).
    // Reads the walker's n from LDS: call after a wave_sync that follows the staging.
    __device__ __forceinline__ void rbm_refresh_grad(const WalkParams& P)
    {
        const int K4 = rbm_wst<K>();
        if (mf_on(P)) {
            if (P.rbm_mf == 2)
                rbm_mf_grad<2>(P);
            else
                rbm_mf_grad<1>(P);
        } else {
#pragma unroll
            for (int s = 0; s < 2; s++)
#pragma unroll
                for (int d = 0; d < 2; d++) {
                    R sum = R(0);
                    if (valid[s]) {
                        const int k = 2 * (s * NP + l) + d;
                        for (int h = 0; h < P.H; h++)
                            sum += lW[k * K4 + 4 * (h & (K - 1)) + h / K] * ln[h];
                    }
                    gR[s][d] = isp2(P) * sum;
                }
        }
        double t = 0;
#pragma unroll
        for (int j = 0; j < HJ; j++) t += double(rcv(j)) * double(rn[j]) * double(rp[j]);
        lr = P.isig2_p * P.isig2_p * gsum<K>(t);
    }

    __device__ __forceinline__ void rbm_stage_n(const WalkParams& P)
    {
#pragma unroll
        for (int j = 0; j < HJ; j++)
            if (hvalid(P, j)) ln[l + K * j] = rn[j];
    }

    // ranks of the owned coordinates, the sorted vector, v, n, p, gradients (block-wide)
    __device__ __forceinline__ void rbm_init(const WalkParams& P)
    {
        if (!P.has_rbmp) return;
        // stage the unsorted coordinates in the scratch row
#pragma unroll
        for (int s = 0; s < 2; s++)
            if (valid[s]) {
                xs_prop[2 * (s * NP + l)] = x[s][0];
                xs_prop[2 * (s * NP + l) + 1] = x[s][1];
            }
        wave_sync();
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int d = 0; d < 2; d++) {
                const int idx = 2 * (s * NP + l) + d;
                int r = idx;
                if (P.sorting && valid[s]) {
                    // rank with ties broken by index (ref wavefunction.cpp:40-54 sorts values)
                    r = 0;
                    const R xv = x[s][d];
                    for (int i = 0; i < P.ND; i++) {
                        const R xi = xs_prop[i];
                        r += (xi < xv || (xi == xv && i < idx)) ? 1 : 0;
                    }
                }
                rk[s][d] = r;
                if (valid[s]) xs_cur[r] = x[s][d];
            }
        wave_sync();
        rbm_activate(P);
        rbm_stage_n(P);
        wave_sync();
        rbm_refresh_grad(P);
    }

    // after a step (block-wide): commit an accepted proposal's RBM state,
    // refresh the gradients and the local energy
    __device__ __forceinline__ void rbm_after_step(const WalkParams& P, bool accepted)
    {
        if (P.has_rbmp) {
            if (accepted) {
                for (int i = l; i < P.ND; i += K) xs_cur[i] = xs_prop[i];
#pragma unroll
                for (int s = 0; s < 2; s++)
#pragma unroll
                    for (int d = 0; d < 2; d++) {
                        const int idx = 2 * (s * NP + l) + d;
                        int r = rk[s][d];
                        r -= (r > r_old) ? 1 : 0;
                        r += (r >= r_new) ? 1 : 0;
                        rk[s][d] = (idx == mk) ? r_new : r;
                    }
#pragma unroll
                for (int j = 0; j < HJ; j++) {
                    rv[j] = pv[j];
                    rn[j] = pn[j];
                    rp[j] = pp[j];
                }
                rbm_stage_n(P);
            }
            wave_sync();
            // the matrix cores take the whole wave: walkers that kept their
            // state get their previous values again (the same n, the same sums)
            if (mf_on(P) || accepted) rbm_refresh_grad(P);
        }
        if (accepted) local_energy(P);
    }
};

// ------------------------------------------------------------------------
// one Metropolis step of a walker (ref importancesampling.cpp:49-81,
// bruteforce.cpp:26-52, elementoperation.cpp:38-87)
// ------------------------------------------------------------------------
template <typename R>
struct StepInfo {
    R rs2, rp, rg;   // Slater |det'/det|^2, Pade and Gaussian ratios of the proposal
    R rrg, rrp;      // RBM-Gaussian and RBM-product ratios
};

// a whole-particle move (learned proposal): particle p to (x0, x1), accepted if
// log|psi'/psi|^2 + logq >= log u (logq = log q(x|x') - log q(x'|x))
template <typename R>
struct ParticleMove {
    int p;
    R x0, x1;
    R logq;
};

template <typename R, int NP, int K, bool RB>
__device__ __forceinline__ bool walker_step(Walker<R, NP, K, RB>& w, const WalkParams& P, int k,
                                            R z, R um, R ua, R lua, bool forced, R xforced,
                                            StepInfo<R>& info,
                                            const ParticleMove<R>* pm = nullptr)
{
    using WK = Walker<R, NP, K, RB>;
    if (pm) k = 2 * pm->p;
    const R dt = R(P.dt), sqdt = R(P.sqrt_dt);
    const R oa = R(P.omega * w.th_alpha);
    const R gam = R(w.th_gamma);
    const int p = k >> 1, d = k & 1;
    const int slot = p >= NP ? 1 : 0;
    const int lp = p - slot * NP;
    const int osrc = w.gbase + lp;
    const bool pup = p < P.n_up;

    // current position and drift of the moved particle
    // (value selects throughout: a conditional between two register-array
    // elements must not become a select of addresses, which would force the
    // array into scratch memory)
    const R xp0 = __shfl(vsel(slot, w.x[1][0], w.x[0][0]), osrc, 64);
    const R xp1 = __shfl(vsel(slot, w.x[1][1], w.x[0][1]), osrc, 64);
    R mg = R(0);
    if (P.has_slater) mg += vsel(slot, vsel(d, w.gS[1][1], w.gS[1][0]), vsel(d, w.gS[0][1], w.gS[0][0]));
    if (P.has_pade) mg += vsel(slot, vsel(d, w.gJ[1][1], w.gJ[1][0]), vsel(d, w.gJ[0][1], w.gJ[0][0]));
    if constexpr (RB) {
        if (P.has_rbmp) mg += vsel(slot, vsel(d, w.gR[1][1], w.gR[1][0]), vsel(d, w.gR[0][1], w.gR[0][0]));
        if (P.has_rbmg) {
            const R xo = vsel(slot, vsel(d, w.x[1][1], w.x[1][0]), vsel(d, w.x[0][1], w.x[0][0]));
            const R ao = vsel(slot, vsel(d, w.ra[1][1], w.ra[1][0]), vsel(d, w.ra[0][1], w.ra[0][0]));
            mg -= (xo - ao) * R(1.0 / P.sig2_g);
        }
    }
    const R xpd = d ? xp1 : xp0;
    R Gd = __shfl(mg, osrc, 64);
    if (P.has_gauss) Gd -= oa * xpd;
    const R F = R(2) * Gd;
    const R stale =
        __shfl(vsel(slot, vsel(d, w.qf[1][1], w.qf[1][0]), vsel(d, w.qf[0][1], w.qf[0][0])), osrc, 64);

    // drift of the proposal: the reference's F, or in exact Langevin with a drift
    // limit a > 0, 2F / (1 + sqrt(1 + a F^2 dt / 2))
    const bool limit = P.sampling == kImportanceExact && P.drift_a > 0;
    const R Fp =
        limit ? R(2) * F * M<R>::rcp(R(1) + M<R>::fsqrt(R(1) + R(0.5 * P.drift_a) * F * F * dt)) : F;
    R dx;
    if (forced) dx = xforced - xpd;
    else if (P.sampling == kBruteForce) dx = R(10) * (um - R(0.5)) * dt;
    else dx = R(0.5) * dt * Fp + sqdt * z;
    const R xnd = forced ? xforced : xpd + dx;
    const R xn0 = pm ? pm->x0 : (d ? xp0 : xnd);
    const R xn1 = pm ? pm->x1 : (d ? xnd : xp1);

    // pair terms of the moved particle against the owned ones
    R jexp = R(0), dvi = R(0), dlj = R(0), djg = R(0), gp0 = R(0), gp1 = R(0);
    R dg[2][2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const int q = s * NP + w.l;
        const bool val = w.valid[s] && q != p;
        const R e0 = w.x[s][0] - xp0, e1 = w.x[s][1] - xp1;
        const R n0 = w.x[s][0] - xn0, n1 = w.x[s][1] - xn1;
        R ro, rn, iro, irn;
        M<R>::dist(e0 * e0 + e1 * e1, val, ro, iro);
        M<R>::dist(n0 * n0 + n1 * n1, val, rn, irn);
        const R a = (w.upspin[s] == pup) ? w.pa_same : w.pa_anti;
        const R fo = M<R>::rcp(R(1) + gam * ro), fn = M<R>::rcp(R(1) + gam * rn);
        const R ho = ro * fo, hn = rn * fn;
        const R uo = a * fo * fo, un = a * fn * fn;
        const bool in_ratio = val && (!P.q2_upper || q > p);
        jexp += in_ratio ? a * (hn - ho) : R(0);
        dvi += irn - iro;
        dlj += val ? R(2) * (un * (irn - R(2) * gam * fn) - uo * (iro - R(2) * gam * fo)) : R(0);
        djg += val ? a * (hn * hn - ho * ho) : R(0);
        gp0 -= un * n0 * irn;
        gp1 -= un * n1 * irn;
        dg[s][0] = un * n0 * irn - uo * e0 * iro;
        dg[s][1] = un * n1 * irn - uo * e1 * iro;
    }
    if (P.has_pade) {
        jexp = gsum<K>(jexp);
        dlj = gsum<K>(dlj);
        djg = gsum<K>(djg);
        gp0 = gsum<K>(gp0);
        gp1 = gsum<K>(gp1);
    }
    if (P.coulomb) dvi = gsum<K>(dvi);

    // Slater ratio R = phi(r_p') . Ainv[:,p] (ref slaterdeterminant.cpp:128-132)
    R Rs = R(1), Sl = R(0), gsn = R(0), iR = R(1);
    OrbRows<R, WK::S> nrow;   // orbital rows of the proposed position
    if (P.has_slater) {
        // bring the moved spin block into a0 (wave-uniform branch: skipped when
        // no walker of the wave needs it)
        const bool need = slot != w.act;
        if (__any(need)) {
            if constexpr (std::is_same<R, float>::value) {
                // one register swap per entry on the lanes that need it
                // (exec-masked) instead of two selects per entry
                if (need) {
#pragma unroll
                    for (int kk = 0; kk < NP; kk++)
                        asm("v_swap_b32 %0, %1" : "+v"(w.a0[kk]), "+v"(w.a1[kk]));
                }
            } else {
#pragma unroll
                for (int kk = 0; kk < NP; kk++) {
                    const R t0 = w.a0[kk], t1 = w.a1[kk];
                    w.a0[kk] = need ? t1 : t0;
                    w.a1[kk] = need ? t0 : t1;
                }
            }
            w.act = slot;
        }
        const R s = w.sw(P);
        orb_rows<R, WK::S>(s * xn0, s * xn1, nrow);
        // phi(r').col = sum_nx h_nx(x') u[nx], u[nx] = sum_ny h_ny(y') col(nx, ny); the
        // d/dx' row reuses u, and v[nx] = sum_ny h_ny'(y') col(nx, ny) gives the
        // d/dy' row (both for exact Langevin only)
        const bool exact = P.sampling == kImportanceExact;
        R u[WK::S + 1], v[WK::S + 1];
#pragma unroll
        for (int i = 0; i <= WK::S; i++) u[i] = v[i] = R(0);
#pragma unroll
        for (int kk = 0; kk < NP; kk++) {
            const int nx = orb_nx(kk), ny = orb_ny(kk);
            u[nx] += nrow.hy[ny] * w.a0[kk];
            if (exact && ny >= 1) v[nx] += nrow.dy[ny] * w.a0[kk];
        }
        R Tx = R(0), Ty = R(0);
#pragma unroll
        for (int i = 0; i <= WK::S; i++) {
            Sl += nrow.hx[i] * u[i];
            if (exact) {
                if (i >= 1) Tx += nrow.dx[i] * u[i];
                Ty += nrow.hx[i] * v[i];
            }
        }
        Rs = __shfl(Sl, osrc, 64);
        iR = M<R>::rcp(Rs);
        if (exact) gsn = s * __shfl(d ? Ty : Tx, osrc, 64) * iR;
    }

    // RBM ratios of the proposal (ref rbmgaussian.cpp:34-46, rbmproduct.cpp:43-52, 104-116)
    R rrg = R(1), rrp = R(1), gRn = R(0), ak = R(0);
    if constexpr (RB) {
        if (P.has_rbmg) {
            ak = __shfl(vsel(slot, vsel(d, w.ra[1][1], w.ra[1][0]), vsel(d, w.ra[0][1], w.ra[0][0])),
                        osrc, 64);
            const R xo = xpd - ak, xq = xnd - ak;
            // the reference's ratio is psi'/psi, exp(d / (2 sigma^2)); |psi'/psi|^2 is exp(d / sigma^2)
            const R den = R(P.q13_psi_ratio ? 0.5 / P.sig2_g : 1.0 / P.sig2_g);
            rrg = M<R>::exp((xo * xo - xq * xq) * den);
        }
        if (P.has_rbmp) {
            // the proposal's sorted vector: remove the moved value at its rank, insert
            // the new one at r_new = #(other coordinates < x') (identity when unsorted)
            int ro = k, rq = k;
            if (P.sorting) {
                ro = __shfl(vsel(slot, vsel(d, w.rk[1][1], w.rk[1][0]), vsel(d, w.rk[0][1], w.rk[0][0])),
                            osrc, 64);
                int cnt = 0;
#pragma unroll
                for (int s = 0; s < 2; s++)
#pragma unroll
                    for (int dd = 0; dd < 2; dd++) {
                        const int idx = 2 * (s * NP + w.l) + dd;
                        cnt += (w.valid[s] && idx != k && w.x[s][dd] < xnd) ? 1 : 0;
                    }
                rq = gsum<K>(cnt);
            }
            w.mk = k;
            w.r_old = ro;
            w.r_new = rq;
            constexpr int K4 = rbm_wst<K>();   // row stride of the walk's W
            R acc[WK::HJ];
#pragma unroll
            for (int j = 0; j < WK::HJ; j++) acc[j] = R(0);
            // entry i of the proposal's sorted vector
            auto sorted_at = [&](int i) {
                const int ii = i - (i > rq ? 1 : 0);
                int src = ii + (ii >= ro ? 1 : 0);
                src = src < P.ND ? src : P.ND - 1;
                const R xs = w.xs_cur[src];
                return (i == rq) ? xnd : xs;
            };
            if (w.mf_on(P)) {
                for (int i = w.l; i < P.ND; i += K) w.xs_prop[i] = sorted_at(i);
                wave_sync();
                if (P.rbm_mf == 2)
                    w.template rbm_mf_pre<2>(P, acc);
                else
                    w.template rbm_mf_pre<1>(P, acc);
            } else {
                for (int i = 0; i < P.ND; i++) {
                    const R xi = sorted_at(i);
                    if ((i & (K - 1)) == w.l) w.xs_prop[i] = xi;
#pragma unroll
                    for (int j = 0; j < WK::HJ; j++) acc[j] += w.lW[i * K4 + 4 * w.l + j] * xi;
                }
            }
            R q = R(1), gt = R(0);
#pragma unroll
            for (int j = 0; j < WK::HJ; j++) {
                const bool ok = w.hvalid(P, j);
                w.pv[j] = w.rbv(j) + w.isp2(P) * acc[j];
                const R e = M<R>::exp(w.pv[j]);
                w.pp[j] = ok ? M<R>::rcp(R(1) + e) : R(0);
                w.pn[j] = ok ? e * w.pp[j] : R(0);
                // p_old / p_new = p_old (1 + e^{v'}) (ref rbmproduct.cpp:112-116)
                q *= ok ? w.rp[j] * (R(1) + e) : R(1);
                if (P.sampling == kImportanceExact) gt += w.lW[k * K4 + 4 * w.l + j] * w.pn[j];
            }
            q = gprod<K>(q);
            rrp = q * q;
            if (P.sampling == kImportanceExact) gRn = w.isp2(P) * gsum<K>(gt);
        }
    }

    // acceptance (ref importancesampling.cpp:43-47, 68-81; bruteforce.cpp:42-52)
    info.rs2 = Rs * Rs;
    info.rp = P.has_pade ? M<R>::exp(R(2) * jexp) : R(1);
    info.rg = P.has_gauss ? M<R>::exp(pm ? oa * (xp0 * xp0 + xp1 * xp1 - xn0 * xn0 - xn1 * xn1)
                                         : oa * (xpd * xpd - xnd * xnd))
                          : R(1);
    info.rrg = rrg;
    info.rrp = rrp;
    const R ratio = info.rs2 * info.rp * info.rg * rrg * rrp;
    // log of the ratio as a sum of logs: whole-electron moves change the Pade
    // exponent by tens, and the product of the factors would leave the FP32 range
    auto log_ratio = [&]() {
        R lr = R(0);
        if (P.has_slater) lr += R(2) * M<R>::flog(Rs < R(0) ? -Rs : Rs);
        if (P.has_pade) lr += R(2) * jexp;
        if (P.has_gauss)
            lr += pm ? oa * (xp0 * xp0 + xp1 * xp1 - xn0 * xn0 - xn1 * xn1)
                     : oa * (xpd * xpd - xnd * xnd);
        if constexpr (RB) lr += M<R>::flog(rrg) + M<R>::flog(rrp);
        return lr;
    };
    bool accept;
    if (forced) {
        accept = true;
    } else if (pm) {
        accept = log_ratio() + pm->logq >= lua;
    } else if (P.sampling == kBruteForce) {
        accept = !(ratio < ua);
    } else if (P.sampling == kImportanceCompat) {
        const R xg = R(0.5) * (stale - F) * dx;
        if constexpr (std::is_same<R, double>::value) {
            // the reference's weight and test, `if (w < u) reject` (a NaN weight is accepted)
            const R wgt = (M<R>::exp(xg) + R(1)) * ratio;
            accept = !(wgt < ua);
        } else {
            // FP32: the same test in the log domain, log w = softplus(xg) + log ratio,
            // which neither overflows nor forms inf x 0 next to a node
            const R sp = (xg > R(0) ? xg : R(0)) + M<R>::flog(R(1) + M<R>::exp(-(xg > R(0) ? xg : -xg)));
            accept = sp + log_ratio() >= lua;
        }
    } else {
        R Gn = R(0);
        if (P.has_slater) Gn += gsn;
        if (P.has_pade) Gn += d ? gp1 : gp0;
        if (P.has_gauss) Gn -= oa * xnd;
        if constexpr (RB) {
            if (P.has_rbmp) Gn += gRn;
            if (P.has_rbmg) Gn -= (xnd - ak) * R(1.0 / P.sig2_g);
        }
        const R Fr = R(2) * Gn;
        const R Fn =
            limit ? R(2) * Fr * M<R>::rcp(R(1) + M<R>::fsqrt(R(1) + R(0.5 * P.drift_a) * Fr * Fr * dt))
                  : Fr;
        const R lg = R(-0.5) * dx * (Fp + Fn) + dt * R(0.125) * (Fp * Fp - Fn * Fn);
        // an undefined weight (0 x inf next to a node) is rejected
        if constexpr (std::is_same<R, double>::value)
            accept = ratio * M<R>::exp(lg) >= ua;
        else
            accept = log_ratio() + lg >= lua;
    }
    if (!accept) return false;

    const bool owner = (w.l == lp);
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const bool me = owner && s == slot;
        if (me && pm) {
            w.x[s][0] = xn0;
            w.x[s][1] = xn1;
        } else if (me) {
            if (d) w.x[s][1] = xnd; else w.x[s][0] = xnd;
            if (P.sampling == kImportanceCompat) {
                if (d) w.qf[s][1] = F; else w.qf[s][0] = F;
            }
        }
        if (P.has_pade) {
            w.gJ[s][0] = me ? gp0 : w.gJ[s][0] + (w.valid[s] ? dg[s][0] : R(0));
            w.gJ[s][1] = me ? gp1 : w.gJ[s][1] + (w.valid[s] ? dg[s][1] : R(0));
        }
    }
    if (pm)
        w.sumx2 += double(xn0) * double(xn0) + double(xn1) * double(xn1) - double(xp0) * double(xp0) -
                   double(xp1) * double(xp1);
    else
        w.sumx2 += double(xnd) * double(xnd) - double(xpd) * double(xpd);
    // only group-reduced increments may enter the walker sums
    if (P.coulomb) w.vint += double(dvi);
    if (P.has_pade) {
        w.lj += double(dlj);
        w.jg += double(djg);
    }

    if (P.has_slater) {
        // Sherman-Morrison update of the moved block (ref slaterdeterminant.cpp:143-153):
        // column j -= (S_j/R) column p, column p /= R, written as one FMA per entry
        const R cf = owner ? (R(1) - iR) : Sl * iR;
#pragma unroll
        for (int kk = 0; kk < NP; kk++) {
            const R cp = __shfl(w.a0[kk], osrc, 64);
            w.a0[kk] -= cf * cp;
        }
        R gx, gy, sec;
        w.slater_derivs(P, vsel(slot, w.x[1][0], w.x[0][0]), vsel(slot, w.x[1][1], w.x[0][1]),
                        w.a0, gx, gy, sec);
        if (w.l < NP) {
            if (slot) { w.gS[1][0] = gx; w.gS[1][1] = gy; w.sS[1] = sec; }
            else { w.gS[0][0] = gx; w.gS[0][1] = gy; w.sS[0] = sec; }
        }
    }
    // with RBM elements the state is completed (and E_L computed) block-wide
    // in rbm_after_step
    if constexpr (!RB) w.local_energy(P);
    return true;
}

template <typename R, int NP, int K, bool RB>
__device__ __forceinline__ void walker_load(Walker<R, NP, K, RB>& w, const WalkParams& P,
                                            const double* pos, const double* qf)
{
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const int q = s * NP + w.l;
        w.valid[s] = (w.l < NP) && (q < P.N);
        w.upspin[s] = q < P.n_up;
        w.x[s][0] = w.valid[s] ? R(pos[2 * q]) : R(0);
        w.x[s][1] = w.valid[s] ? R(pos[2 * q + 1]) : R(0);
        if (qf && w.valid[s]) {
            w.qf[s][0] = R(qf[2 * q]);
            w.qf[s][1] = R(qf[2 * q + 1]);
        } else {
            w.qf[s][0] = w.qf[s][1] = R(0);
        }
        w.gS[s][0] = w.gS[s][1] = w.sS[s] = R(0);
    }
#pragma unroll
    for (int k = 0; k < NP; k++) w.a0[k] = w.a1[k] = R(0);
    w.act = 0;
}

// full gradient d ln psi / dx of owned particle s, dimension dd
template <typename R, int NP, int K, bool RB>
__device__ __forceinline__ R walker_gradient(const Walker<R, NP, K, RB>& w, const WalkParams& P,
                                             int s, int dd)
{
    R g = R(0);
    if (P.has_slater) g += w.gS[s][dd];
    if (P.has_pade) g += w.gJ[s][dd];
    if (P.has_gauss) g -= R(P.omega * w.th_alpha) * w.x[s][dd];
    if constexpr (RB) {
        if (P.has_rbmp) g += w.gR[s][dd];
        if (P.has_rbmg) g -= (w.x[s][dd] - w.ra[s][dd]) * R(1.0 / P.sig2_g);
    }
    return g;
}

// ------------------------------------------------------------------------
// learned proposal: a per-electron MLP evaluated by the walker's K lanes.
// Features of electron i: its coordinates and the relative coordinates of its
// 8 nearest electrons, in units of 1/sqrt(omega); two tanh layers of 64 (lane
// l computes units l + K j); 4 outputs: move mean and log-scale.  Weights are
// bfloat16 in LDS, arithmetic in the walk precision R.
// ------------------------------------------------------------------------
__device__ __forceinline__ float bf16_to_float(uint16_t b)
{
    return __uint_as_float((uint32_t)b << 16);
}

template <typename R>
__device__ __forceinline__ R lp_tanh(R a)
{
    if constexpr (std::is_same<R, double>::value) {
        return ::tanh(a);
    } else {
        // e^(2a) as one base-2 exponential of a scaled by 2 log2(e)
        const R e = __builtin_amdgcn_exp2f(a * R(2.0 * 1.4426950408889634));
        return R(1) - R(2) * M<R>::rcp(e + R(1));
    }
}

// the network's weights in LDS: the parameter vector with the rows of W1
// padded from 18 to 24 bf16 (48 B, zeros after the 18th: a lane's four-word
// operand of the first layer is then one aligned load, no masked words) and the
// rows of W2 from 64 to 72 bf16 (144 B, so that the sixteen rows a
// matrix-core operand, or a wave's lanes, reads at one column do not share
// their banks); the blocks keep their order
constexpr int kLpW1S = 24;
constexpr int kLpW2S = 72;
constexpr int kLdsW1 = 0;
constexpr int kLdsB1 = kLdsW1 + kLpHid * kLpW1S;
constexpr int kLdsW2 = kLdsB1 + kLpHid;
constexpr int kLdsB2 = kLdsW2 + kLpHid * kLpW2S;
constexpr int kLdsW3 = kLdsB2 + kLpHid;
constexpr int kLdsB3 = kLdsW3 + kLpOut * kLpHid;
constexpr int kLdsParams = kLdsB3 + kLpOut;
// LDS position of parameter i
__device__ __forceinline__ int lp_lds_of(int i)
{
    if (i < kLpB1) return kLdsW1 + (i / kLpIn) * kLpW1S + i % kLpIn;
    if (i < kLpW2) return i - kLpB1 + kLdsB1;
    if (i < kLpB2) return kLdsW2 + ((i - kLpW2) / kLpHid) * kLpW2S + (i - kLpW2) % kLpHid;
    return i - kLpB2 + kLdsB2;
}
// the weights into LDS (lanes of one wave), W1's row padding zeroed
__device__ __forceinline__ void lp_load_weights(uint16_t* W, const uint16_t* w, int lane)
{
    for (int i = lane; i < kLpParams; i += 64) W[lp_lds_of(i)] = w[i];
    constexpr int pad = kLpW1S - kLpIn;
    for (int i = lane; i < kLpHid * pad; i += 64) W[kLdsW1 + (i / pad) * kLpW1S + kLpIn + i % pad] = 0;
}

template <typename R, int K>
struct LpLds {
    uint16_t* W;   // [kLdsParams], see kLpW2S
    R* h;          // [G][kLpHid] first-layer activations of each walker
    float* xa;     // [G][kLpHid] matrix-core pre-activations (FP32 walks, K >= 16)
    uint32_t* fs;  // [G][16] features as bf16 pairs, zero from 18 on (FP32 walks, K >= 16)
    __device__ __forceinline__ R w(int i) const { return R(bf16_to_float(W[i])); }
};

// FP32 walks with K >= 16 (G <= 4 walkers per wave) run the network's hidden
// layers on BF16 matrix-core tiles
constexpr bool lp_mf(size_t rsize, int K) { return rsize == 4 && K >= 16; }

// bytes of the network's LDS (weights, activations, matrix-core staging)
constexpr size_t lp_lds_bytes_c(int K, size_t rsize)
{
    const size_t base = rsize * (size_t)(64 / K) * kLpHid + sizeof(uint16_t) * (size_t)kLdsParams;
    return lp_mf(rsize, K) ? ((base + 15) & ~size_t(15)) + sizeof(float) * (64 / K) * kLpHid +
                                 sizeof(uint32_t) * (64 / K) * 16
                           : base;
}
inline size_t lp_lds_bytes(int K, size_t rsize) { return lp_lds_bytes_c(K, rsize); }

template <typename R, int K>
__device__ __forceinline__ void lp_carve(LpLds<R, K>& L, double* base)
{
    constexpr int G = 64 / K;
    L.h = reinterpret_cast<R*>(base);
    L.W = reinterpret_cast<uint16_t*>(L.h + G * kLpHid);
    if constexpr (lp_mf(sizeof(R), K)) {
        const size_t off = ((sizeof(R) * G * kLpHid + sizeof(uint16_t) * kLdsParams) + 15) & ~size_t(15);
        L.xa = reinterpret_cast<float*>(reinterpret_cast<char*>(base) + off);
        L.fs = reinterpret_cast<uint32_t*>(L.xa + G * kLpHid);
    } else {
        L.xa = nullptr;
        L.fs = nullptr;
    }
}

template <int NP, int K>
constexpr int walk_static_lds_doubles();
// the rebuild's transpose area: per walker NP rows of NP + 1 doubles
template <int NP, int K>
constexpr int walk_t_doubles() { return (64 / K) * NP * (NP + 1); }

// the learned proposal's LDS fits in the Slater rebuild's static LDS, which is
// free once the rebuild is done (then a launch needs no dynamic LDS: 14.8 KB
// instead of 26 KB per wave at N=56, 10 waves per CU instead of 6)
template <typename R, int NP, int K>
constexpr bool lp_fits_static()
{
    return walk_static_lds_doubles<NP, K>() * sizeof(double) >= lp_lds_bytes_c(K, sizeof(R));
}

// minimum over the K lanes of a group of 32-bit keys (same pairing as gsum)
template <int K>
__device__ __forceinline__ uint32_t gmin32(uint32_t v)
{
    auto take = [&](uint32_t o) { v = o < v ? o : v; };
    if constexpr (K >= 2) take((uint32_t)dpp<kDppXor1>((int)v));
    if constexpr (K >= 4) take((uint32_t)dpp<kDppXor2>((int)v));
    if constexpr (K >= 8) take((uint32_t)dpp<kDppHalfMirror>((int)v));
    if constexpr (K >= 16) take((uint32_t)dpp<kDppMirror>((int)v));
    uint32_t a, b;
    if constexpr (K >= 32) {
        xchg16(v, a, b);
        v = a < b ? a : b;
    }
    if constexpr (K >= 64) {
        xchg32(v, a, b);
        v = a < b ? a : b;
    }
    return v;
}

template <typename R, int NP, int K, bool RB>
__device__ __forceinline__ void lp_features(const Walker<R, NP, K, RB>& w, const WalkParams& P, int i,
                                            R xi, R yi, R (&f)[kLpIn])
{
    const R s = R(P.sqrt_omega);
    f[0] = s * xi;
    f[1] = s * yi;
    if constexpr (std::is_same<R, float>::value) {
        // FP32: each candidate as one 32-bit key, the squared distance's bits
        // with the low 6 replaced by the electron's index (the bits of a
        // non-negative float order as the value): the nearest remaining
        // electron is the group minimum of the keys, taken out of its owner
        // lane after each round.  Distances within 2^-17 of each other order
        // by index; the selection is a fixed function of the configuration,
        // the same in both passes of the proposal.
        uint32_t key[2];
#pragma unroll
        for (int sl = 0; sl < 2; sl++) {
            const int q = sl * NP + w.l;
            const float e0 = w.x[sl][0] - xi, e1 = w.x[sl][1] - yi;
            const float d2 = e0 * e0 + e1 * e1;
            key[sl] = (w.valid[sl] && q != i) ? ((__float_as_uint(d2) & ~63u) | (uint32_t)q) : ~0u;
        }
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const uint32_t best = gmin32<K>(key[0] < key[1] ? key[0] : key[1]);
            const bool found = best != ~0u;
            const int j = found ? (int)(best & 63u) : 0;
            const int js = j >= NP ? 1 : 0, jl = j - js * NP;
            const R xj0 = __shfl(vsel(js != 0, w.x[1][0], w.x[0][0]), w.gbase + jl, 64);
            const R xj1 = __shfl(vsel(js != 0, w.x[1][1], w.x[0][1]), w.gbase + jl, 64);
            f[2 + 2 * m] = found ? s * (xj0 - xi) : R(0);
            f[3 + 2 * m] = found ? s * (xj1 - yi) : R(0);
            key[0] = key[0] == best ? ~0u : key[0];
            key[1] = key[1] == best ? ~0u : key[1];
        }
        return;
    }
    bool taken0 = false, taken1 = false;
    constexpr int kNone = 1 << 30;
    sfor<0, 8>([&](auto mi) {
        constexpr int m = decltype(mi)::value;
        // nearest remaining electron: squared distance, ties to the lower index
        R bv = R(0);
        int bi = kNone;
#pragma unroll
        for (int sl = 0; sl < 2; sl++) {
            const int q = sl * NP + w.l;
            const bool tk = sl ? taken1 : taken0;
            const R e0 = w.x[sl][0] - xi, e1 = w.x[sl][1] - yi;
            const R d2 = e0 * e0 + e1 * e1;
            const bool ok = w.valid[sl] && q != i && !tk;
            if (ok && (bi == kNone || d2 < bv || (d2 == bv && q < bi))) {
                bv = d2;
                bi = q;
            }
        }
#pragma unroll
        for (int msk = K / 2; msk >= 1; msk >>= 1) {
            const R ov = __shfl_xor(bv, msk, 64);
            const int oi = __shfl_xor(bi, msk, 64);
            if (oi != kNone && (bi == kNone || ov < bv || (ov == bv && oi < bi))) {
                bv = ov;
                bi = oi;
            }
        }
        const bool found = bi != kNone;
        const int j = found ? bi : 0;
        const int js = j / NP, jl = j - js * NP;
        const R xj0 = __shfl(vsel(js != 0, w.x[1][0], w.x[0][0]), w.gbase + jl, 64);
        const R xj1 = __shfl(vsel(js != 0, w.x[1][1], w.x[0][1]), w.gbase + jl, 64);
        f[2 + 2 * m] = found ? s * (xj0 - xi) : R(0);
        f[3 + 2 * m] = found ? s * (xj1 - yi) : R(0);
        if (found && w.l == jl) {
            if (js) taken1 = true;
            else taken0 = true;
        }
    });
}

// network outputs o[4] of features f (all K lanes of the walker take part).
// Weight rows are read from LDS 8 bfloat16 at a time (W1 rows 2 at a time).
__device__ __forceinline__ float bf16_lo(uint32_t v) { return __uint_as_float(v << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t v) { return __uint_as_float(v & 0xFFFF0000u); }

// FP32 walks: activations rounded to bfloat16 (nearest even) and the hidden
// layers as bf16 dot products with FP32 accumulation (v_dot2c_f32_bf16), two
// products per instruction with no weight conversion.  The proposal stays an
// exact Metropolis-Hastings proposal: both passes evaluate the same function.
typedef __bf16 lp_bf16x2 __attribute__((ext_vector_type(2)));
typedef float lp_float2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ lp_bf16x2 lp_pack(float a, float b)
{
    return __builtin_convertvector((lp_float2){a, b}, lp_bf16x2);
}
__device__ __forceinline__ lp_bf16x2 lp_word(uint32_t v) { return __builtin_bit_cast(lp_bf16x2, v); }
__device__ __forceinline__ float lp_dot2(uint32_t w, uint32_t h, float c)
{
    return __builtin_amdgcn_fdot2_f32_bf16(lp_word(w), lp_word(h), c, false);
}

template <int K>
__device__ __forceinline__ void lp_mlp_bf16(const LpLds<float, K>& L, int g, int l,
                                            const float (&f)[kLpIn], float (&o)[kLpOut])
{
    constexpr int UPL = kLpHid / K;
    constexpr int UNR = UPL <= 4 ? UPL : 1;
    uint16_t* hb = reinterpret_cast<uint16_t*>(L.h) + g * kLpHid;   // bf16 activations
    const uint32_t* W32 = reinterpret_cast<const uint32_t*>(L.W);
    uint32_t fp[kLpIn / 2];
#pragma unroll
    for (int m = 0; m < kLpIn / 2; m++) fp[m] = __builtin_bit_cast(uint32_t, lp_pack(f[2 * m], f[2 * m + 1]));
#pragma unroll UNR
    for (int j = 0; j < UPL; j++) {
        const int u = l + K * j;
        float a = L.w(kLdsB1 + u);
        const uint32_t* row = W32 + (kLdsW1 + u * kLpW1S) / 2;
#pragma unroll
        for (int m = 0; m < kLpIn / 2; m++) a = lp_dot2(row[m], fp[m], a);
        hb[u] = __builtin_bit_cast(uint16_t, (__bf16)lp_tanh(a));
    }
    wave_sync();
    float t[kLpOut] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll UNR
    for (int j = 0; j < UPL; j++) {
        const int u = l + K * j;
        float a2 = L.w(kLdsB2 + u);
#pragma unroll
        for (int c = 0; c < kLpHid; c += 8) {
            const uint4 h = *reinterpret_cast<const uint4*>(hb + c);
            const uint4 v = *reinterpret_cast<const uint4*>(L.W + kLdsW2 + u * kLpW2S + c);
            a2 = lp_dot2(v.x, h.x, a2);
            a2 = lp_dot2(v.y, h.y, a2);
            a2 = lp_dot2(v.z, h.z, a2);
            a2 = lp_dot2(v.w, h.w, a2);
        }
        const float h2 = lp_tanh(a2);
#pragma unroll
        for (int r = 0; r < kLpOut; r++) t[r] += L.w(kLdsW3 + r * kLpHid + u) * h2;
    }
    wave_sync();
#pragma unroll
    for (int r = 0; r < kLpOut; r++) o[r] = gsum<K>(t[r]) + L.w(kLdsB3 + r);
}

// FP32 walks, K >= 16: the three layers as BF16 matrix-core tiles (FP32
// accumulation), the G <= 4 walkers of the wave the columns of 16 x 16 tiles
// (V_MFMA_F32_16X16X32_BF16: result element r of lane t is unit 4q + r of
// the tile, q = t / 16, column t % 16), 4 tiles of 16 units per hidden layer;
// layer 1 contracts the 18 features (zero-padded to 32), layers 2 and 3 the
// 64 units in two steps of 32, layer 3 on one tile whose rows 0..3 are the
// outputs.  The hidden layers' biases and tanh stay on the walker's lanes
// (unit l + K j), reading the tiles' pre-activations through LDS.  Features
// and both layers' activations are rounded to bfloat16, and both passes of a
// proposal evaluate this same function, so the Metropolis-Hastings ratio
// stays exact.
typedef __bf16 lp_bf16x8 __attribute__((ext_vector_type(8)));
typedef float lp_f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ lp_f4 lp_tile(uint4 a, uint4 b, lp_f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(lp_bf16x8, a),
                                                   __builtin_bit_cast(lp_bf16x8, b), c, 0, 0, 0);
}

// STAGED: the features are already in L.fs (lp_features_staged), f unused
template <int K, bool STAGED = false>
__device__ __forceinline__ void lp_mlp_mfma(const LpLds<float, K>& L, int g, int l,
                                            const float (&f)[kLpIn], float (&o)[kLpOut])
{
    constexpr int G = 64 / K, UPL = kLpHid / K;
    static_assert(G <= 16, "one tile column per walker");
    const int t = threadIdx.x & 63, c = t & 15, q = t >> 4;
    const int wv = c < G ? c : G - 1;
    uint16_t* hb = reinterpret_cast<uint16_t*>(L.h);   // [G][64] bf16 activations
    const uint32_t* W32 = reinterpret_cast<const uint32_t*>(L.W);
    // the walker's features, bf16 pairs, zero-padded to 32
    if (!STAGED && l == 0) {
        uint32_t p[16];
#pragma unroll
        for (int m = 0; m < 16; m++)
            p[m] = m < kLpIn / 2 ? __builtin_bit_cast(uint32_t, lp_pack(f[2 * m], f[2 * m + 1])) : 0u;
#pragma unroll
        for (int m = 0; m < 16; m += 4)
            *reinterpret_cast<uint4*>(L.fs + g * 16 + m) = make_uint4(p[m], p[m + 1], p[m + 2], p[m + 3]);
    }
    wave_sync();
    // column c of a tile's result depends on column c of B only: the columns
    // past the wave's walkers repeat walker G - 1 and are not read
    const uint4 b = *reinterpret_cast<const uint4*>(L.fs + wv * 16 + 4 * q);   // features 8q .. 8q+7
    // layer 1: the four words from index 8q of W1's 48-byte row; past the
    // 18th feature they are the row's zero padding (q = 2) or the next row's
    // first words (q = 3, the last row: the first-layer biases), finite
    // values the features' zero padding multiplies
#pragma unroll
    for (int tu = 0; tu < 4; tu++) {
        const int u = 16 * tu + c;
        const uint4 a = *reinterpret_cast<const uint4*>(W32 + (kLdsW1 + u * kLpW1S) / 2 + 4 * q);
        const lp_f4 d = lp_tile(a, b, (lp_f4){0.f, 0.f, 0.f, 0.f});
        if (c < G) *reinterpret_cast<lp_f4*>(L.xa + c * kLpHid + 16 * tu + 4 * q) = d;
    }
    wave_sync();
#pragma unroll
    for (int j = 0; j < UPL; j++) {
        const int u = l + K * j;
        const float a = L.xa[g * kLpHid + u] + L.w(kLdsB1 + u);
        hb[g * kLpHid + u] = __builtin_bit_cast(uint16_t, (__bf16)lp_tanh(a));
    }
    wave_sync();
    // layer 2: B = the walker's first-layer activations, two steps of 32 units
    const uint4 b0 = *reinterpret_cast<const uint4*>(hb + wv * kLpHid + 8 * q);
    const uint4 b1 = *reinterpret_cast<const uint4*>(hb + wv * kLpHid + 32 + 8 * q);
#pragma unroll
    for (int tu = 0; tu < 4; tu++) {
        const uint16_t* row = L.W + kLdsW2 + (16 * tu + c) * kLpW2S;
        lp_f4 d = lp_tile(*reinterpret_cast<const uint4*>(row + 8 * q), b0, (lp_f4){0.f, 0.f, 0.f, 0.f});
        d = lp_tile(*reinterpret_cast<const uint4*>(row + 32 + 8 * q), b1, d);
        if (c < G) *reinterpret_cast<lp_f4*>(L.xa + c * kLpHid + 16 * tu + 4 * q) = d;
    }
    wave_sync();
    // second-layer activations, rounded to bf16, over the first ones (read)
#pragma unroll
    for (int j = 0; j < UPL; j++) {
        const int u = l + K * j;
        const float a = L.xa[g * kLpHid + u] + L.w(kLdsB2 + u);
        hb[g * kLpHid + u] = __builtin_bit_cast(uint16_t, (__bf16)lp_tanh(a));
    }
    wave_sync();
    // layer 3 on one tile: rows 0..3 are W3's (rows 4..15 repeat them and are
    // not read), B the second-layer activations in two steps of 32; lane
    // c < G of the first row of lanes then holds walker c's four outputs
    const uint4 c0 = *reinterpret_cast<const uint4*>(hb + wv * kLpHid + 8 * q);
    const uint4 c1 = *reinterpret_cast<const uint4*>(hb + wv * kLpHid + 32 + 8 * q);
    const uint16_t* row3 = L.W + kLdsW3 + (c & (kLpOut - 1)) * kLpHid;
    lp_f4 d = lp_tile(*reinterpret_cast<const uint4*>(row3 + 8 * q), c0, (lp_f4){0.f, 0.f, 0.f, 0.f});
    d = lp_tile(*reinterpret_cast<const uint4*>(row3 + 32 + 8 * q), c1, d);
    static_assert(kLpOut == 4, "the output rows are lanes c & 3");
#pragma unroll
    for (int r = 0; r < kLpOut; r++) o[r] = __shfl(d[r], g, 64) + L.w(kLdsB3 + r);
}

// FP32 walks on matrix-core tiles (lp_mf): the features of electron i at
// (xi, yi) written straight into the walker's bf16 staging row L.fs (word 0 the
// electron's coordinates, word 1 + m the m-th nearest electron's relative
// coordinates, words 9..15 the tile's zero padding, cleared once per launch).
// The same selection and the same FP32 values as lp_features, but the lane that
// owns the selected electron writes its word, so no round moves coordinates
// between lanes.
template <int NP, int K, bool RB>
__device__ __forceinline__ void lp_features_staged(const Walker<float, NP, K, RB>& w,
                                                   const WalkParams& P, const LpLds<float, K>& L,
                                                   int i, float xi, float yi)
{
    const float s = float(P.sqrt_omega);
    uint32_t* fs = L.fs + w.g * 16;
    if (w.l == 0) fs[0] = __builtin_bit_cast(uint32_t, lp_pack(s * xi, s * yi));
    uint32_t key[2], word[2];
#pragma unroll
    for (int sl = 0; sl < 2; sl++) {
        const int q = sl * NP + w.l;
        const float e0 = w.x[sl][0] - xi, e1 = w.x[sl][1] - yi;
        const float d2 = e0 * e0 + e1 * e1;
        key[sl] = (w.valid[sl] && q != i) ? ((__float_as_uint(d2) & ~63u) | (uint32_t)q) : ~0u;
        word[sl] = __builtin_bit_cast(uint32_t, lp_pack(s * e0, s * e1));
    }
    // with NP >= 10 orbitals per spin (N >= 13) every electron has at least
    // eight others, so every round finds one
    constexpr bool always = NP >= 10;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const uint32_t best = gmin32<K>(key[0] < key[1] ? key[0] : key[1]);
        const bool found = always || best != ~0u;
        const bool own0 = found && key[0] == best, own1 = found && key[1] == best;
        if (own0 || own1) fs[1 + m] = own0 ? word[0] : word[1];
        else if (!always && !found && w.l == 0) fs[1 + m] = 0u;
        key[0] = own0 ? ~0u : key[0];
        key[1] = own1 ? ~0u : key[1];
    }
}

template <typename R, int K>
__device__ __forceinline__ void lp_mlp(const LpLds<R, K>& L, int g, int l, const R (&f)[kLpIn],
                                       R (&o)[kLpOut])
{
    if constexpr (std::is_same<R, float>::value) {
        if constexpr (lp_mf(sizeof(R), K))
            lp_mlp_mfma<K>(L, g, l, f, o);
        else
            lp_mlp_bf16<K>(L, g, l, f, o);
        return;
    }
    constexpr int UPL = kLpHid / K;
    constexpr int UNR = UPL <= 4 ? UPL : 1;   // unroll only the few-units-per-lane shapes
    R* hb = L.h + g * kLpHid;
    const uint32_t* W32 = reinterpret_cast<const uint32_t*>(L.W);
#pragma unroll UNR
    for (int j = 0; j < UPL; j++) {
        const int u = l + K * j;
        R a = L.w(kLdsB1 + u);
        const uint32_t* row = W32 + (kLdsW1 + u * kLpW1S) / 2;   // 18 weights = 9 words, 4-byte aligned
#pragma unroll
        for (int m = 0; m < kLpIn / 2; m++) {
            const uint32_t v = row[m];
            a += R(bf16_lo(v)) * f[2 * m];
            a += R(bf16_hi(v)) * f[2 * m + 1];
        }
        hb[u] = lp_tanh(a);
    }
    wave_sync();
    R t[kLpOut] = {R(0), R(0), R(0), R(0)};
    if constexpr (UPL <= 4) {
        // c-blocks of 8: the activations once, then each unit's 8 weights in one read
        R a2[UPL];
#pragma unroll
        for (int j = 0; j < UPL; j++) a2[j] = L.w(kLdsB2 + l + K * j);
        for (int c = 0; c < kLpHid; c += 8) {
            R h[8];
#pragma unroll
            for (int q = 0; q < 8; q++) h[q] = hb[c + q];
#pragma unroll
            for (int j = 0; j < UPL; j++) {
                const int u = l + K * j;
                const uint4 v = *reinterpret_cast<const uint4*>(L.W + kLdsW2 + u * kLpW2S + c);
                a2[j] += R(bf16_lo(v.x)) * h[0] + R(bf16_hi(v.x)) * h[1] + R(bf16_lo(v.y)) * h[2] +
                         R(bf16_hi(v.y)) * h[3] + R(bf16_lo(v.z)) * h[4] + R(bf16_hi(v.z)) * h[5] +
                         R(bf16_lo(v.w)) * h[6] + R(bf16_hi(v.w)) * h[7];
            }
        }
#pragma unroll
        for (int j = 0; j < UPL; j++) {
            const R h2 = lp_tanh(a2[j]);
#pragma unroll
            for (int r = 0; r < kLpOut; r++) t[r] += L.w(kLdsW3 + r * kLpHid + l + K * j) * h2;
        }
    } else {
        // many units per lane: one unit at a time, the head folded in
        for (int j = 0; j < UPL; j++) {
            const int u = l + K * j;
            R a2 = L.w(kLdsB2 + u);
            for (int c = 0; c < kLpHid; c += 8) {
                const uint4 v = *reinterpret_cast<const uint4*>(L.W + kLdsW2 + u * kLpW2S + c);
                a2 += R(bf16_lo(v.x)) * hb[c] + R(bf16_hi(v.x)) * hb[c + 1] +
                      R(bf16_lo(v.y)) * hb[c + 2] + R(bf16_hi(v.y)) * hb[c + 3] +
                      R(bf16_lo(v.z)) * hb[c + 4] + R(bf16_hi(v.z)) * hb[c + 5] +
                      R(bf16_lo(v.w)) * hb[c + 6] + R(bf16_hi(v.w)) * hb[c + 7];
            }
            const R h2 = lp_tanh(a2);
#pragma unroll
            for (int r = 0; r < kLpOut; r++) t[r] += L.w(kLdsW3 + r * kLpHid + u) * h2;
        }
    }
    wave_sync();
#pragma unroll
    for (int r = 0; r < kLpOut; r++) o[r] = gsum<K>(t[r]) + L.w(kLdsB3 + r);
}

// move mean and scale: mu = l_mu tanh(o/l_mu) / sqrt(omega), scale =
// sigma0 exp(ls tanh(o/ls)) / sqrt(omega)
// (lt = ls tanh(o/ls), the log of the scale relative to sigma0/sqrt(omega))
template <typename R>
__device__ __forceinline__ void lp_head(const WalkParams& P, const R (&o)[kLpOut], R (&mu)[2],
                                        R (&sg)[2], R (&lt)[2])
{
    const R ell = R(P.inv_sqrt_omega);
    const R mm = R(P.lp_mu_max), ls = R(P.lp_ls_max);
    // FP32: the quotients by the two bounds as products with their
    // reciprocals (an IEEE division is about ten instructions; both passes of
    // a proposal evaluate this same function, so the Metropolis-Hastings
    // ratio stays exact)
    const bool f32 = std::is_same<R, float>::value;
    const R imm = R(1.0 / P.lp_mu_max), ils = R(1.0 / P.lp_ls_max);
#pragma unroll
    for (int d = 0; d < 2; d++) {
        mu[d] = ell * mm * lp_tanh(f32 ? o[d] * imm : o[d] / mm);
        lt[d] = ls * lp_tanh(f32 ? o[2 + d] * ils : o[2 + d] / ls);
        sg[d] = R(P.lp_sigma0) * ell * M<R>::exp(lt[d]);
    }
}

// the learned proposal of electron i with normals z0, z1: the move and
// log q(x|x') - log q(x'|x)
template <typename R, int NP, int K, bool RB>
__device__ __forceinline__ ParticleMove<R> lp_propose(const Walker<R, NP, K, RB>& w,
                                                      const WalkParams& P, const LpLds<R, K>& L,
                                                      int i, R z0, R z1)
{
    const int slot = i >= NP ? 1 : 0, li = i - slot * NP;
    const R xi = __shfl(vsel(slot != 0, w.x[1][0], w.x[0][0]), w.gbase + li, 64);
    const R yi = __shfl(vsel(slot != 0, w.x[1][1], w.x[0][1]), w.gbase + li, 64);
    R f[kLpIn], o[kLpOut], mu[2], sg[2], lt[2];
    if constexpr (lp_mf(sizeof(R), K)) {
        // features staged in LDS by the lanes that own them
        lp_features_staged(w, P, L, i, xi, yi);
        lp_mlp_mfma<K, true>(L, w.g, w.l, f, o);
        lp_head(P, o, mu, sg, lt);
        const R xn = xi + mu[0] + sg[0] * z0, yn = yi + mu[1] + sg[1] * z1;
        const R lqf = -R(0.5) * (z0 * z0 + z1 * z1) - (lt[0] + lt[1]);
        lp_features_staged(w, P, L, i, xn, yn);
        lp_mlp_mfma<K, true>(L, w.g, w.l, f, o);
        lp_head(P, o, mu, sg, lt);
        const R r0 = (xi - xn - mu[0]) * M<R>::rcp(sg[0]), r1 = (yi - yn - mu[1]) * M<R>::rcp(sg[1]);
        const R lqr = -R(0.5) * (r0 * r0 + r1 * r1) - (lt[0] + lt[1]);
        return ParticleMove<R>{i, xn, yn, lqr - lqf};
    }
    lp_features(w, P, i, xi, yi, f);
    lp_mlp(L, w.g, w.l, f, o);
    lp_head(P, o, mu, sg, lt);
    const R xn = xi + mu[0] + sg[0] * z0, yn = yi + mu[1] + sg[1] * z1;
    if constexpr (std::is_same<R, float>::value) {
        // log sigma = log(sigma0/sqrt(omega)) + lt: the constant cancels in the
        // ratio, which then needs no logarithm
        const R lqf = -R(0.5) * (z0 * z0 + z1 * z1) - (lt[0] + lt[1]);
        lp_features(w, P, i, xn, yn, f);
        lp_mlp(L, w.g, w.l, f, o);
        lp_head(P, o, mu, sg, lt);
        const R r0 = (xi - xn - mu[0]) * M<R>::rcp(sg[0]), r1 = (yi - yn - mu[1]) * M<R>::rcp(sg[1]);
        const R lqr = -R(0.5) * (r0 * r0 + r1 * r1) - (lt[0] + lt[1]);
        return ParticleMove<R>{i, xn, yn, lqr - lqf};
    }
    const R lqf = -(R(0.5) * z0 * z0 + M<R>::log(sg[0])) - (R(0.5) * z1 * z1 + M<R>::log(sg[1]));
    lp_features(w, P, i, xn, yn, f);
    lp_mlp(L, w.g, w.l, f, o);
    lp_head(P, o, mu, sg, lt);
    const R r0 = xi - xn - mu[0], r1 = yi - yn - mu[1];
    const R lqr = -(r0 * r0 / (R(2) * sg[0] * sg[0]) + M<R>::log(sg[0])) -
                  (r1 * r1 / (R(2) * sg[1] * sg[1]) + M<R>::log(sg[1]));
    return ParticleMove<R>{i, xn, yn, lqr - lqf};
}

// ------------------------------------------------------------------------
// RBM block-level LDS: gradient-window sums of the wave, then W and the
// walkers' sorted coordinate vectors and n.  Sized at launch (dynamic LDS).
// ------------------------------------------------------------------------
template <typename R, int K>
struct RbmLds {
    double* accW;    // [2][ND][H + 1] sum over the wave's walkers of xs_k n_h, and of E xs_k n_h
    double* accAB;   // [2][ND + H] the a and b parts of the same sums
    double* E;       // [G] staged local energies
    R* W;            // [max(ND, 32)][4K], rows from ND on zero
    R* X;            // [G][kMfX] matrix-core results, one row per walker
    R* xs;           // [G][2][ND] sorted coordinates: current, proposal
    R* n;            // [G][H]
    R* bc;           // [2][K][4]: b_h and sum_k W[k][h]^2, unit h = l + K j at [l][j]
    // W_elsewhere: W in other LDS (the rebuild's), not in the dynamic area
    __device__ __forceinline__ void carve(double* base, int ND, int H, R* W_elsewhere = nullptr)
    {
        constexpr int G = 64 / K;
        accW = base;
        accAB = accW + 2 * ND * rbm_hst(H);
        E = accAB + 2 * (ND + H);
        R* p = reinterpret_cast<R*>(E + G);   // 16-byte aligned: ND and G are even
        if (W_elsewhere) {
            W = W_elsewhere;
        } else {
            W = p;
            p += (ND > 32 ? ND : 32) * rbm_wst<K>();
        }
        X = p;
        xs = X + G * kMfX;
        n = xs + G * 2 * ND;
        bc = n + G * H;
    }
};

inline size_t rbm_lds_bytes(int K, int ND, int H, size_t rsize, bool w_elsewhere = false)
{
    const int G = 64 / K;
    const size_t wsz = w_elsewhere ? 0 : (size_t)(ND > 32 ? ND : 32) * (4 * K + 1);
    return sizeof(double) * (2 * ND * rbm_hst(H) + 2 * (ND + H) + G) +
           rsize * (wsz + (size_t)G * kMfX + (size_t)G * 2 * ND + (size_t)G * H + (size_t)8 * K);
}

// LDS the walk of one block needs besides the dynamic RBM / learned-proposal
// area: the Gauss-Jordan transpose and pivot rows of the Slater rebuild
template <int NP, int K>
constexpr int walk_static_lds_doubles() { return walk_t_doubles<NP, K>() + 5 * (64 / K) * NP; }

// the walk's copy of the RBM-product W (32 rows: N D <= 32) fits in the Slater
// rebuild's static LDS, free once the launch's rebuild is done (N=12 FP32:
// 17.5 KB instead of 21.6 KB per wave, 9 waves per CU instead of 7)
template <typename R, int NP, int K>
constexpr bool rbm_w_fits_static()
{
    return walk_static_lds_doubles<NP, K>() * sizeof(double) >= 32 * rbm_wst<K>() * sizeof(R);
}

template <typename R, int NP, int K, bool RB>
__device__ __forceinline__ void rbm_bind(Walker<R, NP, K, RB>& w, RbmLds<R, K>& L, const WalkParams& P)
{
    w.lW = L.W;
    w.mf_x = L.X;
    w.mf_xs = L.xs;
    w.mf_n = L.n;
    w.xs_cur = L.xs + (size_t)w.g * 2 * P.ND;
    w.xs_prop = w.xs_cur + P.ND;
    w.ln = L.n + (size_t)w.g * P.H;
    w.lbc = L.bc;
}

// the gradient-window sums of dlnpsi/dW = xs n^T over the wave's walkers on
// FP32 matrix-core tiles (V_MFMA_F32_16X16X4_F32): coordinates k are the tile
// rows, hidden units h the columns, the walkers the contraction index (4 per
// instruction); FP32 products and 4-walker sums, each step's tile added to
// the FP64 LDS sums (every (k, h) belongs to one lane).  Reads the staged
// current sorted vectors, n and E_L of the live walkers.
// BT: row ND of the tiles (zero padding otherwise, ND % 16 != 0) is a row of
// ones, so the same tiles also give the window sums of dlnpsi/db_h = n_h and of
// E_L n_h over the wave's walkers, added to the b part of L.accAB
template <int K, bool BT = false>
__device__ __forceinline__ void rbm_mf_wgrad(const WalkParams& P, const RbmLds<float, K>& L, int glive)
{
    constexpr int G = 64 / K, OPS = G / 4;
    static_assert(G >= 4 && G <= 16, "4..16 walkers per wave");
    const int t = threadIdx.x & 63, c = t & 15, q = t >> 4;
    const int ND = P.ND, H = P.H;
    // one tile at a time, its operands read from LDS where they are used (few
    // registers live: the walk kernel is at its register cap)
#pragma unroll
    for (int tk = 0; tk < 2; tk++)
#pragma unroll
        for (int th = 0; th < 2; th++) {
            if (16 * tk >= ND || 16 * th >= H) continue;   // wave-uniform
            const int kr = 16 * tk + c, hc = 16 * th + c;
            mf_f4 d1 = {0.f, 0.f, 0.f, 0.f}, d2 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int op = 0; op < OPS; op++) {
                const int gw = 4 * op + q;   // walker: the contraction index of this lane
                const bool live = gw < glive;
                const float x = L.xs[gw * 2 * ND + (kr < ND ? kr : ND - 1)];
                const float n = L.n[gw * H + (hc < H ? hc : H - 1)];
                const float xa = (live && kr < ND) ? x : ((BT && live && kr == ND) ? 1.f : 0.f);
                const float nb = (live && hc < H) ? n : 0.f;
                d1 = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, nb, d1, 0, 0, 0);
                d2 = __builtin_amdgcn_mfma_f32_16x16x4f32(float(L.E[gw]) * xa, nb, d2, 0, 0, 0);
            }
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int k = 16 * tk + 4 * q + i;
                if (k < ND && hc < H) {
                    L.accW[k * rbm_hst(H) + hc] += double(d1[i]);
                    L.accW[ND * rbm_hst(H) + k * rbm_hst(H) + hc] += double(d2[i]);
                }
                if (BT && k == ND && hc < H) {
                    L.accAB[ND + hc] += double(d1[i]);
                    L.accAB[2 * ND + H + hc] += double(d2[i]);
                }
            }
        }
}

// ------------------------------------------------------------------------
// the walk kernel: equil + steps Metropolis steps per walker
// ------------------------------------------------------------------------
// the per-launch values of a single-system launch
__device__ __forceinline__ WalkDyn walk_dyn(const WalkParams& P)
{
    return WalkDyn{P.alpha,   P.gamma,     P.counter0,  P.t_begin, P.t_end,
                   P.acc_add, P.window_lo, P.window_hi, P.init_qf};
}

// electron densities of the walker's current state (ref src/sampler.cpp:639-692;
// r_i = |x_i|, ref Metropolis::calculateRadialVectorElement, metropolis.cpp:31-39):
//   radial one-body: r < maxRadius, bin int(r/dr) < bins;
//   spatial one-body: cell (int(y/2dr) + bins/2, int(x/2dr) + bins/2) with C
//     truncation toward zero, both |int(.)| < bins/2 (the centre cell spans 4dr);
//   radial two-body: for i < j, (int(r_i/dr) + 1, int(r_j/dr) + 1), both < bins.
// Lanes count their own particles; the pair loop takes the other particles'
// bins from their owner lanes.  Positions enter in FP64 without contraction, as
// the reference computes them.
template <typename R, int NP, int K, bool RB>
__device__ __forceinline__ void density_sample(const Walker<R, NP, K, RB>& w, const WalkParams& P,
                                               bool live)
{
#pragma clang fp contract(off)
    const int nb = P.dens_bins, half = nb / 2;
    const double dr = P.dens_step;
    int b2[2] = {nb, nb};
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const double x = double(w.x[s][0]), y = double(w.x[s][1]);
        const double r = ::sqrt(x * x + y * y);
        const double q = r / dr;
        const bool own = live && w.valid[s];
        b2[s] = (q < double(nb - 1)) ? int(q) + 1 : nb;
        if (own && (P.dens_flags & 1) && r < P.dens_max_radius && q < double(nb))
            atomicAdd(P.dens_radial + int(q), 1ull);
        if (own && (P.dens_flags & 2)) {
            const double fx = x / (2.0 * dr), fy = y / (2.0 * dr);
            // |int(f)| < half exactly when |f| < half (truncation toward zero)
            if (fabs(fx) < double(half) && fabs(fy) < double(half))
                atomicAdd(P.dens_spatial + (size_t)(int(fy) + half) * nb + (int(fx) + half), 1ull);
        }
    }
    if (P.dens_flags & 4) {
        for (int j = 0; j < P.N; j++) {
            const int sj = j >= NP ? 1 : 0;
            const int bj = __shfl(sj ? b2[1] : b2[0], w.gbase + (j - sj * NP), 64);
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const int q = s * NP + w.l;
                if (live && w.valid[s] && q < j && b2[s] < nb && bj < nb)
                    atomicAdd(P.dens_pair + (size_t)b2[s] * nb + bj, 1ull);
            }
        }
    }
}

// the walk of the G = 64/K walkers of block `block` (one wave); lds_T and
// lds_row as walk_static_lds_doubles, rbm_dyn the dynamic area.  DN: the
// instantiation counts electron densities when P.dens_flags asks for them
// (kept out of the plain instantiations the walk runs without densities)
// BT (FP32 matrix-core RBM walks with ND % 16 != 0): the window sums of
// dlnpsi/db come from the W-gradient tiles (rbm_mf_wgrad) instead of per-lane
// FP64 registers
template <typename R, int NP, int K, bool RB, bool LP, bool DN = false, bool BT = false>
__device__ __forceinline__ void walk_block(const WalkParams& P, const WalkDyn& D, int block,
                                           double* lds_T, double* lds_row, double* rbm_dyn)
{
    using WK = Walker<R, NP, K, RB>;
    constexpr int G = WK::G;

    WK w;
    w.th_alpha = D.alpha;
    w.th_gamma = D.gamma;
    w.pa_same = R(P.a_same);
    w.pa_anti = R(P.a_anti);
    const int lane = threadIdx.x;
    w.l = lane & (K - 1);
    w.gbase = lane & ~(K - 1);
    w.g = lane / K;
    const int walker_raw = block * G + w.g;
    const bool live = walker_raw < P.W;
    const int walker = live ? walker_raw : P.W - 1;
    const int nd = 2 * P.N;
    const int glive = (P.W - block * G) < G ? (P.W - block * G) : G;

    RbmLds<R, K> L;
    LpLds<R, K> LL;
    walker_load(w, P, P.pos + (size_t)walker * nd, P.qf ? P.qf + (size_t)walker * nd : nullptr);
    // the network's weights and activations: in the rebuild's LDS once the
    // rebuild is done where they fit (large NP), in the dynamic area otherwise
    constexpr bool lp_static = lp_fits_static<R, NP, K>();
    if constexpr (LP && !lp_static) {
        lp_carve(LL, rbm_dyn);
        lp_load_weights(LL.W, P.lp_w, lane);
        if constexpr (lp_mf(sizeof(R), K))   // the feature rows' zero padding
            for (int i = lane; i < G * 16; i += 64) LL.fs[i] = 0u;
        wave_sync();
    }
    // the RBM state; W in the rebuild's LDS after the rebuild where it fits
    constexpr bool w_static = rbm_w_fits_static<R, NP, K>();
    auto rbm_start = [&](R* W_elsewhere) {
        L.carve(rbm_dyn, P.ND, P.H, W_elsewhere);
        rbm_bind(w, L, P);
        for (int i = lane; i < 2 * P.ND * rbm_hst(P.H); i += 64) L.accW[i] = 0.0;
        for (int i = lane; i < 2 * (P.ND + P.H); i += 64) L.accAB[i] = 0.0;
        w.rbm_setup(P);
        wave_sync();
        w.rbm_init(P);
    };
    if constexpr (RB && !w_static) rbm_start(nullptr);
    if (P.has_slater) w.rebuild_slater(P, lds_T, lds_row);
    if constexpr (RB && w_static) {
        wave_sync();   // the rebuild's reads of lds_T are done
        rbm_start(reinterpret_cast<R*>(lds_T));
    }
    w.rebuild_pairs(P);
    w.local_energy(P);
    if constexpr (LP && lp_static) {
        wave_sync();   // the rebuild's reads of lds_T are done
        lp_carve(LL, lds_T);
        lp_load_weights(LL.W, P.lp_w, lane);
        if constexpr (lp_mf(sizeof(R), K))   // the feature rows' zero padding
            for (int i = lane; i < G * 16; i += 64) LL.fs[i] = 0u;
        wave_sync();
    }
    if (D.init_qf && D.t_begin == 0) {
#pragma unroll
        for (int s = 0; s < 2; s++) {
            w.qf[s][0] = R(2) * walker_gradient(w, P, s, 0);
            w.qf[s][1] = R(2) * walker_gradient(w, P, s, 1);
        }
    }

    double aE = 0, aE2 = 0, aT = 0, aV = 0, aI = 0, aT2 = 0, aV2 = 0, aI2 = 0, aAcc = 0;
    double aGa = 0, aGg = 0, aGaE = 0, aGgE = 0;
    double rGa[2][2] = {{0, 0}, {0, 0}}, rGaE[2][2] = {{0, 0}, {0, 0}};
    double rGb[WK::HJ], rGbE[WK::HJ];
#pragma unroll
    for (int j = 0; j < WK::HJ; j++) rGb[j] = rGbE[j] = 0.0;
    int bad = 0;

    // random draws for K steps at a time, one step per lane of the group
    int rk = 0;
    R rz = R(0), rum = R(0), rua = R(0), rz2 = R(0), rlua = R(0);
    // the FP32 walks and the learned moves test acceptance in the log domain
    constexpr bool log_accept = sizeof(R) == 4 || LP;

    for (int t = D.t_begin; t < D.t_end; t++) {
        const int tk = t & (K - 1);
        if (tk == 0 || t == D.t_begin) {
            step_draws<R>(P.seed, (uint32_t)walker, D.counter0 + (uint64_t)(t - tk + w.l), nd, rk,
                          rz, rum, rua, rz2);
            if constexpr (log_accept) rlua = M<R>::flog(rua);
        }
        const int src = w.gbase + tk;
        const int k = __shfl(rk, src, 64);
        const R z = __shfl(rz, src, 64);
        const R um = __shfl(rum, src, 64);
        const R ua = __shfl(rua, src, 64);
        const R lua = log_accept ? __shfl(rlua, src, 64) : R(0);

        StepInfo<R> info;
        bool accept;
        if constexpr (LP) {
            // electron k/2, normals (z, z2) of the same Box-Muller pair
            const R z2 = __shfl(rz2, src, 64);
            const ParticleMove<R> mv = lp_propose(w, P, LL, k >> 1, z, z2);
            accept = walker_step(w, P, k, z, um, ua, lua, false, R(0), info, &mv);
        } else {
            accept = walker_step(w, P, k, z, um, ua, lua, false, R(0), info);
        }
        if constexpr (RB) w.rbm_after_step(P, accept);

        if (t >= P.equil) {
            const double E = w.el;
            if (!(E - E == 0.0)) bad++;
            aE += E;
            aE2 += E * E;
            aT += w.tk;
            aV += w.vext;
            aI += w.vint;
            aT2 += w.tk * w.tk;
            aV2 += w.vext * w.vext;
            aI2 += w.vint * w.vint;
            aAcc += accept ? 1.0 : 0.0;
            if (t > D.window_lo && t < D.window_hi) {      // ref sampler.cpp:82-86
                aGa += w.ga;
                aGg += w.gg;
                aGaE += w.ga * E;
                aGgE += w.gg * E;
                if constexpr (RB) {
                    // RBM-Gaussian: dlnpsi/da_k = (x_k - a_k)/sigma^2 (ref rbmgaussian.cpp:81-86)
                    if (P.has_rbmg) {
#pragma unroll
                        for (int s = 0; s < 2; s++)
#pragma unroll
                            for (int d = 0; d < 2; d++) {
                                const double xa =
                                    w.valid[s] ? double(w.x[s][d] - w.ra[s][d]) / P.sig2_g : 0.0;
                                rGa[s][d] += xa;
                                rGaE[s][d] += E * xa;
                            }
                    }
                    if (P.has_rbmp) {
                        // dlnpsi/db_h = n_h (ref rbmproduct.cpp:77-83)
                        if constexpr (!BT) {
#pragma unroll
                            for (int j = 0; j < WK::HJ; j++) {
                                rGb[j] += double(w.rn[j]);
                                rGbE[j] += E * double(w.rn[j]);
                            }
                        }
                        // dlnpsi/dW[k][h] = xs_k n_h, summed over the wave's walkers
                        if (w.l == 0) L.E[w.g] = E;
                        wave_sync();
                        bool wg_done = false;
                        if constexpr (sizeof(R) == 4 && K >= 4) {
                            if (w.mf_on(P)) {
                                rbm_mf_wgrad<K, BT>(P, L, glive);
                                wg_done = true;
                            }
                        }
                        if (!wg_done) {
                        // lane (kb, hb) of an 8 x 8 grid owns a block of (k, h) pairs
                        const int kbs = (P.ND + 7) >> 3, hbs = (P.H + 7) >> 3;
                        const int kb = lane >> 3, hb = lane & 7;
                        for (int kk = 0; kk < kbs; kk++) {
                            const int kc = kb * kbs + kk;
                            if (kc >= P.ND) break;
                            double s1[4] = {0, 0, 0, 0}, s2[4] = {0, 0, 0, 0};
                            for (int gg = 0; gg < glive; gg++) {
                                const double xk = double(L.xs[(size_t)gg * 2 * P.ND + kc]);
                                const double ex = L.E[gg] * xk;
#pragma unroll
                                for (int hh = 0; hh < 4; hh++) {
                                    const int h = hb * hbs + hh;
                                    if (hh < hbs && h < P.H) {
                                        const double nv = double(L.n[(size_t)gg * P.H + h]);
                                        s1[hh] += xk * nv;
                                        s2[hh] += ex * nv;
                                    }
                                }
                            }
#pragma unroll
                            for (int hh = 0; hh < 4; hh++) {
                                const int h = hb * hbs + hh;
                                if (hh < hbs && h < P.H) {
                                    L.accW[kc * rbm_hst(P.H) + h] += s1[hh];
                                    L.accW[P.ND * rbm_hst(P.H) + kc * rbm_hst(P.H) + h] += s2[hh];
                                }
                            }
                        }
                        }
                        wave_sync();
                    }
                }
            }
            if (P.rec && live && walker < P.rec_walkers && w.l == 0) {
                const size_t at = (size_t)walker * P.steps + (t - P.equil);
                P.rec[at] = float(E);
                if (P.rec_parts == 4) {
                    const size_t plane = (size_t)P.rec_walkers * P.steps;
                    P.rec[plane + at] = float(w.tk);
                    P.rec[2 * plane + at] = float(w.vext);
                    P.rec[3 * plane + at] = float(w.vint);
                }
            }
            if constexpr (DN) {
                if (P.dens_flags) density_sample(w, P, live);
            }
        }
    }

    if constexpr (RB) {
        // the wave's a and b sums, walker by walker in a fixed order, then the
        // block row of the gradient-window sums
        const int nab = P.ND + P.H;
        for (int gg = 0; gg < G; gg++) {
            if (w.g == gg && live) {
                if (P.has_rbmg) {
#pragma unroll
                    for (int s = 0; s < 2; s++)
#pragma unroll
                        for (int d = 0; d < 2; d++)
                            if (w.valid[s]) {
                                const int kc = 2 * (s * NP + w.l) + d;
                                L.accAB[kc] += rGa[s][d];
                                L.accAB[nab + kc] += rGaE[s][d];
                            }
                }
                if (P.has_rbmp && !BT) {
#pragma unroll
                    for (int j = 0; j < WK::HJ; j++)
                        if (w.hvalid(P, j)) {
                            L.accAB[P.ND + w.l + K * j] += rGb[j];
                            L.accAB[nab + P.ND + w.l + K * j] += rGbE[j];
                        }
                }
            }
            wave_sync();
        }
        const int half = P.ND + P.H + P.ND * P.H;
        const double wscale = P.q6_w_unscaled ? 1.0 : P.isig2_p;
        double* row = P.rbm_acc + (size_t)block * rbm_row_width(P.ND, P.H);
        for (int i = lane; i < 2 * half; i += 64) {
            const int hf = i >= half ? 1 : 0;
            const int r = i - hf * half;
            double v;
            if (r < nab) {
                v = L.accAB[hf * nab + r];
            } else {
                const int m = r - nab, kc = m % P.ND, h = m / P.ND;
                v = L.accW[hf * P.ND * rbm_hst(P.H) + kc * rbm_hst(P.H) + h] * wscale;
            }
            row[i] = D.acc_add ? row[i] + v : v;
        }
    }

    if (!live) return;
#pragma unroll
    for (int s = 0; s < 2; s++) {
        if (!w.valid[s]) continue;
        const int q = s * NP + w.l;
        double* dst = P.pos + (size_t)walker * nd + 2 * q;
        dst[0] = double(w.x[s][0]);
        dst[1] = double(w.x[s][1]);
        if (P.qf) {
            double* qd = P.qf + (size_t)walker * nd + 2 * q;
            qd[0] = double(w.qf[s][0]);
            qd[1] = double(w.qf[s][1]);
        }
    }
    if (w.l == 0) {
        double* a = P.acc + (size_t)walker * kNumWalkerAcc;
        const double v[kNumWalkerAcc] = {aE, aE2, aT, aV, aI, aT2, aV2, aI2, aAcc,
                                         aGa, aGg, aGaE, aGgE, double(bad)};
#pragma unroll
        for (int i = 0; i < kNumWalkerAcc; i++) a[i] = D.acc_add ? a[i] + v[i] : v[i];
    }
}

// FP32 RBM instantiations: capped at 256 registers for 2 waves/SIMD (they
// need ~330; the spilled ones are cold enough that the second wave wins,
// 1.2x at N=12, H=24)
// MF: the RBM-product contraction mode P.rbm_mf (0 lanes, 1 FP32 tiles, 2
// BF16 tiles); an instantiation per mode lets the compiler drop the others' code
// SG >= 0: the walk of Slater x Gaussian x Pade-Jastrow with Coulomb
// interaction and sampler SG (exact Langevin, the reference's importance
// sampling, or learned moves with LP), the configurations of the quantum-dot
// studies, asserted to the compiler so that the other elements' and
// samplers' code drops out of the instantiation; SG = -1: any configuration
template <typename R, int NP, int K, bool RB, bool LP, bool DN, int MF = 0, int SG = -1>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu((RB && sizeof(R) == 4) ? 2 : 1)))
void walk_kernel(const WalkParams P)
{
    __builtin_assume(P.rbm_mf == MF);
    if constexpr (SG >= 0 && !RB) {
        __builtin_assume(P.sampling == SG);
        __builtin_assume(P.has_slater == 1 && P.has_gauss == 1 && P.has_pade == 1);
        __builtin_assume(P.coulomb == 1 && P.has_rbmg == 0 && P.has_rbmp == 0);
    }
    if constexpr (SG >= 0 && RB) {   // Slater x RBM-Gaussian x RBM-product x Pade-Jastrow
        __builtin_assume(P.sampling == SG);
        __builtin_assume(P.has_slater == 1 && P.has_rbmg == 1 && P.has_rbmp == 1 && P.has_pade == 1);
        __builtin_assume(P.coulomb == 1 && P.has_gauss == 0);
    }
    __shared__ __align__(16) double lds[walk_static_lds_doubles<NP, K>()];
    extern __shared__ __align__(16) double rbm_dyn[];
    WalkDyn D = walk_dyn(P);
    if (P.dyn_st) {   // an optimisation loop on the device: this iteration's values
        const OptState& st = *P.dyn_st;
        D.alpha = P.dyn_ag[0];
        D.gamma = P.dyn_ag[1];
        D.counter0 = st.counter0;
        const int S = P.steps / st.batches;   // ref sampler.cpp:104
        const int kb = st.iteration % st.batches;
        D.window_lo = S * kb;
        D.window_hi = S * (kb + 1);
        D.init_qf = st.fresh;
    }
    // the RBM x Slater x Pade-Jastrow instantiations on tiles take the b sums
    // from the tiles (their launcher checks ND % 16 != 0)
    constexpr bool BT = RB && MF != 0 && SG >= 0 && sizeof(R) == 4;
    walk_block<R, NP, K, RB, LP, DN, BT>(P, D, blockIdx.x, lds, lds + walk_t_doubles<NP, K>(), rbm_dyn);
}

// ------------------------------------------------------------------------
// parity kernel: evaluate, force one move through the incremental path,
// evaluate again (see vmc_evaluate and vmc_evaluate_elements in include/vmc.h)
// ------------------------------------------------------------------------

template <typename R, int NP, int K, bool RB>
__device__ __forceinline__ void eval_gradients(const Walker<R, NP, K, RB>& w, const WalkParams& P,
                                               const EvalOut& O, double* g)
{
    const int mp = O.maxpar;
    for (int e = 0; e < O.num_elements; e++) {
        double* row = g + (size_t)e * mp;
        const int kind = O.rows[e];
        if (kind == 1 && w.l == 0 && mp > 0) row[0] = w.ga;          // Gaussian alpha
        if (kind == 3 && w.l == 0 && mp > 0) row[0] = w.gg;          // Pade gamma
        if constexpr (RB) {
            if (kind == 4) {                                          // RBM-Gaussian a_k
#pragma unroll
                for (int s = 0; s < 2; s++)
#pragma unroll
                    for (int d = 0; d < 2; d++)
                        if (w.valid[s])
                            row[2 * (s * NP + w.l) + d] = double(w.x[s][d] - w.ra[s][d]) / P.sig2_g;
            }
            if (kind == 5) {                                          // RBM-product b_h, W[k][h]
                const double ws = P.q6_w_unscaled ? 1.0 : P.isig2_p;
#pragma unroll
                for (int j = 0; j < Walker<R, NP, K, RB>::HJ; j++) {
                    const int h = w.l + K * j;
                    if (h >= P.H) continue;
                    row[h] = double(w.rn[j]);
                    for (int kc = 0; kc < P.ND; kc++)
                        row[P.H + kc + P.ND * h] = double(w.xs_cur[kc]) * double(w.rn[j]) * ws;
                }
            }
        }
    }
}

template <typename R, int NP, int K, bool RB>
__global__ __launch_bounds__(64) void eval_kernel(const WalkParams P, const double* pos,
                                                  const int* moves, const double* newv,
                                                  const EvalOut O)
{
    using WK = Walker<R, NP, K, RB>;
    constexpr int G = WK::G;
    __shared__ __align__(16) double lds_T[walk_t_doubles<NP, K>()];
    __shared__ double lds_row[5 * G * NP];
    extern __shared__ __align__(16) double rbm_dyn[];
    WK w;
    w.th_alpha = P.alpha;
    w.th_gamma = P.gamma;
    w.pa_same = R(P.a_same);
    w.pa_anti = R(P.a_anti);
    const int lane = threadIdx.x;
    w.l = lane & (K - 1);
    w.gbase = lane & ~(K - 1);
    w.g = lane / K;
    const int walker_raw = blockIdx.x * G + w.g;
    const bool live = walker_raw < P.W;
    const int walker = live ? walker_raw : P.W - 1;
    const int nd = 2 * P.N;
    const int width = 9 + nd;
    RbmLds<R, K> L;
    walker_load(w, P, pos + (size_t)walker * nd, nullptr);
    if constexpr (RB) {
        L.carve(rbm_dyn, P.ND, P.H);
        rbm_bind(w, L, P);
        w.rbm_setup(P);
        wave_sync();
        w.rbm_init(P);
    }
    if (P.has_slater) w.rebuild_slater(P, lds_T, lds_row);
    w.rebuild_pairs(P);
    w.local_energy(P);
    for (int pass = 0; pass < 2; pass++) {
        StepInfo<R> info{R(0), R(0), R(0), R(0), R(0)};
        if (pass == 1) {
            if (O.newv2) {
                const ParticleMove<R> mv{moves[walker], R(O.newv2[2 * walker]),
                                         R(O.newv2[2 * walker + 1]), R(0)};
                walker_step(w, P, 0, R(0), R(0), R(0), R(0), true, R(0), info, &mv);
            } else {
                walker_step(w, P, moves[walker], R(0), R(0), R(0), R(0), true, R(newv[walker]), info);
            }
            if constexpr (RB) w.rbm_after_step(P, true);
        }
        double* o = (pass ? O.o1 : O.o0) + (size_t)walker * width;
        if (live && w.l == 0) {
            o[0] = w.el;
            o[1] = w.tk;
            o[2] = w.vext;
            o[3] = w.vint;
            o[4] = w.ga;
            o[5] = w.gg;
            o[6] = double(info.rs2);
            o[7] = double(info.rp);
            o[8] = double(info.rg);
            if (pass == 1 && O.ratios) {
                for (int e = 0; e < O.num_elements; e++) {
                    const int kind = O.rows[e];
                    const double r = kind == 1 ? double(info.rg)
                                     : kind == 2 ? double(info.rs2)
                                     : kind == 3 ? double(info.rp)
                                     : kind == 4 ? double(info.rrg)
                                                 : double(info.rrp);
                    O.ratios[(size_t)walker * O.num_elements + e] = r;
                }
            }
        }
        if (live) {
#pragma unroll
            for (int s = 0; s < 2; s++) {
                if (!w.valid[s]) continue;
                const int q = s * NP + w.l;
                o[9 + 2 * q] = 2.0 * double(walker_gradient(w, P, s, 0));
                o[9 + 2 * q + 1] = 2.0 * double(walker_gradient(w, P, s, 1));
            }
            double* g = pass ? O.g1 : O.g0;
            if (g) eval_gradients(w, P, O, g + (size_t)walker * O.num_elements * O.maxpar);
        }
    }
}

#ifndef VMC_WALK_PRECISION
#define VMC_WALK_PRECISION 0
#endif

#if VMC_WALK_PRECISION == 32 || VMC_WALK_PRECISION == 64
// ------------------------------------------------------------------------
// launch tables of one precision (this translation unit is compiled once per
// precision so the two halves build in parallel)
// ------------------------------------------------------------------------
using RT = std::conditional_t<VMC_WALK_PRECISION == 32, float, double>;

template <int NP, int K, bool RB, bool LP, bool DN, int SG = -1>
static hipError_t launch_walk_t(const WalkParams& P, hipStream_t stream)
{
    constexpr int G = 64 / K;
    const int blocks = (P.W + G - 1) / G;
    const size_t lds = RB ? rbm_lds_bytes(K, P.ND, P.H, sizeof(RT), rbm_w_fits_static<RT, NP, K>())
                          : ((LP && !lp_fits_static<RT, NP, K>()) ? lp_lds_bytes(K, sizeof(RT)) : 0);
    if constexpr (RB && sizeof(RT) == 4 && K >= 4) {
        // the RBM x Slater x Pade-Jastrow walk under the reference's importance
        // sampling with the tiles has instantiations of its own
        if constexpr (!DN) {
            const bool set = P.sampling == kImportanceCompat && P.has_slater == 1 && P.has_rbmg == 1 &&
                             P.has_rbmp == 1 && P.has_pade == 1 && P.coulomb == 1 && P.has_gauss == 0 &&
                             P.ND % 16 != 0;
            if (set && P.rbm_mf == 1) {
                hipLaunchKernelGGL((walk_kernel<RT, NP, K, RB, LP, DN, 1, kImportanceCompat>),
                                   dim3(blocks), dim3(64), lds, stream, P);
                return hipGetLastError();
            }
            if (set && P.rbm_mf == 2) {
                hipLaunchKernelGGL((walk_kernel<RT, NP, K, RB, LP, DN, 2, kImportanceCompat>),
                                   dim3(blocks), dim3(64), lds, stream, P);
                return hipGetLastError();
            }
        }
        if (P.rbm_mf == 1) {
            hipLaunchKernelGGL((walk_kernel<RT, NP, K, RB, LP, DN, 1>), dim3(blocks), dim3(64), lds,
                               stream, P);
            return hipGetLastError();
        }
        if (P.rbm_mf == 2) {
            hipLaunchKernelGGL((walk_kernel<RT, NP, K, RB, LP, DN, 2>), dim3(blocks), dim3(64), lds,
                               stream, P);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((walk_kernel<RT, NP, K, RB, LP, DN, 0, SG>), dim3(blocks), dim3(64), lds,
                       stream, P);
    return hipGetLastError();
}

template <int NP, int K, bool RB>
static hipError_t launch_eval_t(const WalkParams& P, const double* pos, const int* moves,
                                const double* newv, const EvalOut& O, hipStream_t stream)
{
    constexpr int G = 64 / K;
    const int blocks = (P.W + G - 1) / G;
    const size_t lds = RB ? rbm_lds_bytes(K, P.ND, P.H, sizeof(RT)) : 0;
    hipLaunchKernelGGL((eval_kernel<RT, NP, K, RB>), dim3(blocks), dim3(64), lds, stream, P, pos,
                       moves, newv, O);
    return hipGetLastError();
}

// (np, K): Slater orbitals per spin (or particle slots per spin without a
// determinant) and lanes per walker; rbm selects the RBM instantiations.  The
// closed shells past N = 56 (N = 72, 90, 110: one walker per wave, 64 lanes)
// are a translation unit of their own per precision (VMC_WALK_LARGE), without
// RBM elements and the learned proposal.
#ifdef VMC_WALK_LARGE
#if VMC_WALK_PRECISION == 32
#define VMC_SUFFIX(name) name##_large_f32
#else
#define VMC_SUFFIX(name) name##_large_f64
#endif
#define VMC_SHAPES(X) X(36, 64, false) X(45, 64, false) X(55, 64, false)
#define VMC_LP_SHAPES(X)
#else
#if VMC_WALK_PRECISION == 32
#define VMC_SUFFIX(name) name##_f32
#else
#define VMC_SUFFIX(name) name##_f64
#endif
#define VMC_SHAPES(X)                                                        \
    X(1, 1, false) X(3, 4, false) X(6, 8, false) X(10, 16, false)            \
    X(15, 16, false) X(21, 32, false) X(28, 32, false)                       \
    X(1, 1, true) X(1, 2, true) X(1, 4, true) X(1, 8, true) X(3, 4, true)    \
    X(3, 8, true) X(6, 8, true) X(10, 16, true)

// shapes with the learned proposal (no RBM elements)
#define VMC_LP_SHAPES(X) X(1, 1) X(3, 4) X(6, 8) X(10, 16) X(15, 16) X(21, 32) X(28, 32)
#endif

// one shape's launch: density counting, then the specialised samplers of the
// plain shapes, then the generic instantiation
template <int NP, int K, bool RB>
static hipError_t launch_walk_shape(const WalkParams& P, bool sg, hipStream_t stream)
{
    if (P.dens_flags) return launch_walk_t<NP, K, RB, false, true>(P, stream);
    if constexpr (!RB) {
        if (sg && P.sampling == kImportanceExact)
            return launch_walk_t<NP, K, RB, false, false, kImportanceExact>(P, stream);
        if (sg && P.sampling == kImportanceCompat)
            return launch_walk_t<NP, K, RB, false, false, kImportanceCompat>(P, stream);
    }
    return launch_walk_t<NP, K, RB, false, false>(P, stream);
}

hipError_t VMC_SUFFIX(launch_walk)(int np, int K, bool rbm, bool lp, const WalkParams& P,
                                   hipStream_t stream)
{
    if (lp) {
        if (rbm) return hipErrorInvalidValue;
    const bool sgl = P.has_slater == 1 && P.has_gauss == 1 && P.has_pade == 1 && P.coulomb == 1;
#define Y(a, b)                                                                       \
    if (np == a && K == b) {                                                          \
        if (P.dens_flags) return launch_walk_t<a, b, false, true, true>(P, stream);   \
        if (sgl) return launch_walk_t<a, b, false, true, false, kLearned>(P, stream); \
        return launch_walk_t<a, b, false, true, false>(P, stream);                    \
    }
        VMC_LP_SHAPES(Y)
#undef Y
        return hipErrorInvalidValue;
    }
    // every shape has an instantiation without the density code (its registers
    // and scheduling stay out of the walks that count no densities), and the
    // plain shapes ones for Slater x Gaussian x Pade-Jastrow with Coulomb under
    // exact Langevin and under the reference's importance sampling
    const bool sg = P.has_slater == 1 && P.has_gauss == 1 && P.has_pade == 1 && P.coulomb == 1 &&
                    P.has_rbmg == 0 && P.has_rbmp == 0;
#define X(a, b, c) \
    if (np == a && K == b && rbm == c) return launch_walk_shape<a, b, c>(P, sg, stream);
    VMC_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t VMC_SUFFIX(launch_eval)(int np, int K, bool rbm, const WalkParams& P,
                                   const double* pos, const int* moves, const double* newv,
                                   const EvalOut& O, hipStream_t stream)
{
#define X(a, b, c) \
    if (np == a && K == b && rbm == c) return launch_eval_t<a, b, c>(P, pos, moves, newv, O, stream);
    VMC_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

#ifndef VMC_WALK_LARGE
// the draws of n consecutive steps of one walker, in the walk precision
template <typename R>
__global__ void draws_kernel(uint64_t seed, uint32_t walker, uint64_t counter0, int nd, int n,
                             int* k, double* z, double* um, double* ua)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int kk;
    R zz, m, a, z2;
    step_draws<R>(seed, walker, counter0 + (uint64_t)i, nd, kk, zz, m, a, z2);
    k[i] = kk;
    z[i] = double(zz);
    um[i] = double(m);
    ua[i] = double(a);
}

hipError_t VMC_SUFFIX(launch_draws)(uint64_t seed, uint32_t walker, uint64_t counter0, int nd,
                                    int n, int* k, double* z, double* um, double* ua,
                                    hipStream_t stream)
{
    hipLaunchKernelGGL(draws_kernel<RT>, dim3((n + 255) / 256), dim3(256), 0, stream, seed, walker,
                       counter0, nd, n, k, z, um, ua);
    return hipGetLastError();
}
#endif
#endif

#if VMC_WALK_PRECISION == 0
// ------------------------------------------------------------------------
// deterministic reduction of the per-walker accumulators: one block, fixed order
// ------------------------------------------------------------------------
// column sums of the per-walker accumulators in two deterministic stages:
// block b sums walkers [256 b, 256 b + 256) with a fixed tree (coalesced loads
// of the block's rows into LDS), then one block sums the block partials in
// order with the same tree
__global__ __launch_bounds__(256) void reduce_acc_partial_kernel(const double* acc, int W,
                                                                 double* part)
{
    __shared__ double tile[256 * kNumWalkerAcc];
    const int w0 = blockIdx.x * 256;
    const int n = (W - w0 < 256 ? W - w0 : 256) * kNumWalkerAcc;
    const double* src = acc + (size_t)w0 * kNumWalkerAcc;
    for (int i = threadIdx.x; i < 256 * kNumWalkerAcc; i += 256) tile[i] = i < n ? src[i] : 0.0;
    __syncthreads();
    for (int m = 128; m > 0; m >>= 1) {
        if ((int)threadIdx.x < m) {
#pragma unroll
            for (int a = 0; a < kNumWalkerAcc; a++)
                tile[threadIdx.x * kNumWalkerAcc + a] += tile[(threadIdx.x + m) * kNumWalkerAcc + a];
        }
        __syncthreads();
    }
    if ((int)threadIdx.x < kNumWalkerAcc)
        part[(size_t)blockIdx.x * kNumWalkerAcc + threadIdx.x] = tile[threadIdx.x];
}

__global__ __launch_bounds__(256) void reduce_acc_kernel(const double* part, int nb, double* out)
{
    __shared__ double tile[256 * kNumWalkerAcc];
    for (int a = 0; a < kNumWalkerAcc; a++) {
        double s = 0;
        for (int i = threadIdx.x; i < nb; i += 256) s += part[(size_t)i * kNumWalkerAcc + a];
        tile[threadIdx.x * kNumWalkerAcc + a] = s;
    }
    __syncthreads();
    for (int m = 128; m > 0; m >>= 1) {
        if ((int)threadIdx.x < m) {
#pragma unroll
            for (int a = 0; a < kNumWalkerAcc; a++)
                tile[threadIdx.x * kNumWalkerAcc + a] += tile[(threadIdx.x + m) * kNumWalkerAcc + a];
        }
        __syncthreads();
    }
    if ((int)threadIdx.x < kNumWalkerAcc) out[threadIdx.x] = tile[threadIdx.x];
}

// initial positions x ~ N(0, (1/omega)^2) (ref InitialStates/randomnormal.cpp:17)
__global__ void init_positions_kernel(double* pos, int W, int nd, double omega, uint64_t seed)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int blocks = (nd + 1) / 2;
    if (i >= W * blocks) return;
    const int w = i / blocks, blk = i % blocks;
    uint32_t c[4] = {(uint32_t)blk, 0u, (uint32_t)w, 1u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double r = ::sqrt(-2.0 * ::log(u23<double>(c[0])));
    const double th = 6.283185307179586476925286766559 * u23<double>(c[1]);
    pos[(size_t)w * nd + 2 * blk] = r * ::cos(th) / omega;
    if (2 * blk + 1 < nd) pos[(size_t)w * nd + 2 * blk + 1] = r * ::sin(th) / omega;
}

hipError_t launch_walk_f32(int np, int K, bool rbm, bool lp, const WalkParams& P,
                           hipStream_t stream);
hipError_t launch_walk_f64(int np, int K, bool rbm, bool lp, const WalkParams& P,
                           hipStream_t stream);
hipError_t launch_eval_f32(int np, int K, bool rbm, const WalkParams& P, const double* pos,
                           const int* moves, const double* newv, const EvalOut& O,
                           hipStream_t stream);
hipError_t launch_eval_f64(int np, int K, bool rbm, const WalkParams& P, const double* pos,
                           const int* moves, const double* newv, const EvalOut& O,
                           hipStream_t stream);

hipError_t launch_walk_large_f32(int np, int K, bool rbm, bool lp, const WalkParams& P,
                                 hipStream_t stream);
hipError_t launch_walk_large_f64(int np, int K, bool rbm, bool lp, const WalkParams& P,
                                 hipStream_t stream);
hipError_t launch_eval_large_f32(int np, int K, bool rbm, const WalkParams& P, const double* pos,
                                 const int* moves, const double* newv, const EvalOut& O,
                                 hipStream_t stream);
hipError_t launch_eval_large_f64(int np, int K, bool rbm, const WalkParams& P, const double* pos,
                                 const int* moves, const double* newv, const EvalOut& O,
                                 hipStream_t stream);

hipError_t launch_draws_f32(uint64_t seed, uint32_t walker, uint64_t counter0, int nd, int n,
                            int* k, double* z, double* um, double* ua, hipStream_t stream);
hipError_t launch_draws_f64(uint64_t seed, uint32_t walker, uint64_t counter0, int nd, int n,
                            int* k, double* z, double* um, double* ua, hipStream_t stream);

hipError_t launch_walk(int precision, const KernelShape& sh, const WalkParams& P,
                       hipStream_t stream)
{
    const bool lp = P.sampling == kLearned;
    if (sh.np > 28)   // the closed shells past N = 56
        return precision == 32 ? launch_walk_large_f32(sh.np, sh.K, sh.rbm, lp, P, stream)
                               : launch_walk_large_f64(sh.np, sh.K, sh.rbm, lp, P, stream);
    return precision == 32 ? launch_walk_f32(sh.np, sh.K, sh.rbm, lp, P, stream)
                           : launch_walk_f64(sh.np, sh.K, sh.rbm, lp, P, stream);
}

hipError_t launch_draws(int precision, uint64_t seed, uint32_t walker, uint64_t counter0, int nd,
                        int n, int* k, double* z, double* um, double* ua, hipStream_t stream)
{
    return precision == 32 ? launch_draws_f32(seed, walker, counter0, nd, n, k, z, um, ua, stream)
                           : launch_draws_f64(seed, walker, counter0, nd, n, k, z, um, ua, stream);
}

hipError_t launch_eval(int precision, const KernelShape& sh, const WalkParams& P,
                       const double* pos, const int* moves, const double* newv, const EvalOut& O,
                       hipStream_t stream)
{
    if (sh.np > 28)
        return precision == 32
                   ? launch_eval_large_f32(sh.np, sh.K, sh.rbm, P, pos, moves, newv, O, stream)
                   : launch_eval_large_f64(sh.np, sh.K, sh.rbm, P, pos, moves, newv, O, stream);
    return precision == 32 ? launch_eval_f32(sh.np, sh.K, sh.rbm, P, pos, moves, newv, O, stream)
                           : launch_eval_f64(sh.np, sh.K, sh.rbm, P, pos, moves, newv, O, stream);
}

// column sums of a row-major [rows][width] matrix, rows added in order
// column sums of a [nrows][width] array in two deterministic stages: chunk r
// of the rows into partial row r, then the partial rows in order
constexpr int kRowChunks = 64;

__global__ __launch_bounds__(256) void reduce_rows_partial_kernel(const double* rows, int nrows,
                                                                  int width, int chunk,
                                                                  double* part)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= width) return;
    const int r0 = blockIdx.y * chunk, r1 = r0 + chunk < nrows ? r0 + chunk : nrows;
    double s = 0;
    for (int r = r0; r < r1; r++) s += rows[(size_t)r * width + c];
    part[(size_t)blockIdx.y * width + c] = s;
}

__global__ __launch_bounds__(256) void reduce_rows_kernel(const double* part, int nparts, int width,
                                                          double* out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= width) return;
    double s = 0;
    for (int r = 0; r < nparts; r++) s += part[(size_t)r * width + c];
    out[c] = s;
}

// out: [1 + kRowChunks][width]: the sums, then the partial rows
hipError_t launch_reduce_rows(const double* rows, int nrows, int width, double* out,
                              hipStream_t stream)
{
    const int nparts = nrows / 32 < 1 ? 1 : (nrows / 32 < kRowChunks ? nrows / 32 : kRowChunks);
    const int chunk = (nrows + nparts - 1) / nparts;
    double* part = out + width;
    const int cb = (width + 255) / 256;
    hipLaunchKernelGGL(reduce_rows_partial_kernel, dim3(cb, nparts), dim3(256), 0, stream, rows,
                       nrows, width, chunk, part);
    hipLaunchKernelGGL(reduce_rows_kernel, dim3(cb), dim3(256), 0, stream, part, nparts, width, out);
    return hipGetLastError();
}

hipError_t launch_reduce(const double* acc, int W, double* out, hipStream_t stream)
{
    const int nb = (W + 255) / 256;
    double* part = out + kNumWalkerAcc;
    hipLaunchKernelGGL(reduce_acc_partial_kernel, dim3(nb), dim3(256), 0, stream, acc, W, part);
    hipLaunchKernelGGL(reduce_acc_kernel, dim3(1), dim3(256), 0, stream, part, nb, out);
    return hipGetLastError();
}

hipError_t launch_init_positions(double* pos, int W, int nd, double omega, uint64_t seed,
                                 hipStream_t stream)
{
    const int n = W * ((nd + 1) / 2);
    hipLaunchKernelGGL(init_positions_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, pos, W,
                       nd, omega, seed);
    return hipGetLastError();
}

// means of `block` consecutive recorded values: one wave per (block, walker),
// lane j sums values j, j + 64, ... in FP64, then a fixed butterfly over the
// wave; rec [walkers][steps], means [walkers][steps / block]
__global__ __launch_bounds__(64) void block_means_kernel(const float* rec, int steps, int block,
                                                         double* means)
{
    const int nb = steps / block;
    const int walker = blockIdx.x / nb, b = blockIdx.x - walker * nb;
    const float* src = rec + (size_t)walker * steps + (size_t)b * block;
    double s = 0.0;
    for (int j = threadIdx.x; j < block; j += 64) s += double(src[j]);
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) s += __shfl_xor(s, m);
    if (threadIdx.x == 0) means[(size_t)walker * nb + b] = s / double(block);
}

hipError_t launch_block_means(const float* rec, int walkers, int steps, int block, double* means,
                              hipStream_t stream)
{
    // one workgroup per mean; the caller keeps walkers * (steps / block) < 2^31
    hipLaunchKernelGGL(block_means_kernel, dim3(walkers * (steps / block)), dim3(64), 0, stream,
                       rec, steps, block, means);
    return hipGetLastError();
}
#endif

}  // namespace vmc

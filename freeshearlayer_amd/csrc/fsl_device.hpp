// fsl_device.hpp -- device building blocks of the per-step update (gfx950).
//
// Each function restates one piece of the reference's arithmetic in the
// reference's association order, so that with T = double and FP contraction
// off the results are bit-identical to src/ compiled with real = double.
// FP32 kernels may instantiate the templates with T = float (Divisor and
// root are then plain division and sqrt).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsl {

// Workgroup barrier that orders LDS only.  __syncthreads() is a workgroup
// fence over global memory as well, which makes every wave wait for all of
// its outstanding global loads (s_waitcnt vmcnt(0)) at each barrier, so
// loads issued ahead for the next plane would complete there.  The stage
// kernels exchange data between threads through LDS only (their global
// inputs and outputs are distinct arrays), so an LDS-scoped fence suffices:
// s_waitcnt lgkmcnt(0) + s_barrier.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// src/def.h:30-49
constexpr double kK = 1.4;
constexpr double kK_1 = 0.4;
constexpr double kNK_1 = -0.4;
constexpr double kH05K = 0.7;
constexpr double kK_K_1 = 3.5;
constexpr double kK_1_2 = 0.2;
constexpr double kI2K_1 = 1.25;
constexpr double kRair = 287.10;
constexpr double kC1 = 0.333333333333333;
constexpr double kC2 = 0.166666666666667;
constexpr double kSmall = 1e-30;
constexpr double kTwo3 = 0.666666666666667;

// Cell-array geometry: one ghost layer, [i][j][k] with k fastest.
struct Geo {
    int L, H, D;
    long long sj;   // D + 2
    long long si;   // (H + 2)(D + 2)
    long long G;    // (L + 2)(H + 2)(D + 2)
};

__host__ __device__ inline long long cidx(const Geo& g, int i, int j, int k)
{
    return (long long)i * g.si + (long long)j * g.sj + k;
}

// Constants of the case (src/initialize.c:67-94)
struct Phys {
    double rdel[3];    // 1/dx, 1/dy, 1/dz
    double r2del[3];   // 1/(2 dx) built as 2 * 1/(4 dx)
    double r4del[3];   // 1/(4 dx)
    double mu_L, lambda_L, cp_Pr_T, CsDD, DD;
    int dynamic;       // mu_T from mu_SGS (1) or constant Smagorinsky (0)
    int z_quirk;       // src/reconstruction.c:328,340
};

// Inflow / outflow boundary data (src/boundCondInGhostCells.c:17-50,
// src/bounCondOnInterfaces.c:17-71)
struct Bnd {
    const double* u_in;   // HIG values, cell row j = 1..HIG
    double rho_in, p_in, p_out;
    int BL_HIG;
    unsigned k_min, k_max;
    double Ud, Vd, Wd;
};

// Five field pointers of one state (rho, rho u, rho v, rho w, E)
struct State5 {
    double* f[5];
};
struct CState5 {
    const double* f[5];
};

// ---------------------------------------------------------------------
// Limiter, src/reconstruction.c:17-24 (Chakravarthy-Osher minmod, b = 4)
// x y <= 0 gives 0; otherwise the reference returns min(x, |4y|) for x > 0
// and, for x < 0, x when -x < |4y|, else -|4y|: sign(x) min(|x|, |4y|), with
// a tie choosing the same value.  Written as min and copysign (no compare /
// select chains); on finite operands the result is the reference's bit for
// bit.  (With a NaN operand the reference may return NaN where min returns
// the other operand; a NaN state fails the time-step guard.)
template <class T>
__device__ __forceinline__ T minmod(T x, T y)
{
    const T m = __builtin_fmin(__builtin_fabs(x), __builtin_fabs(T(4.0) * y));
    return (x * y <= T(0)) ? T(0) : __builtin_copysign(m, x);
}

// Quotients over one denominator.  For T = float: plain division.  For
// T = double: the compiler's correctly rounded n / d is v_div_scale of both
// operands, v_rcp of d, two Newton steps on the reciprocal, m = n r, one
// residual step (v_div_fmas) and v_div_fixup.  The reciprocal and its Newton
// steps depend on d alone: Divisor<double> does them once and div() does the
// rest per numerator, with the same operations on the same values.  The two
// steps left out change nothing in the normal range: v_div_scale rescales
// only subnormal or near-overflow operands or quotients (|d| outside
// [2^-1022, 2^1022), 0 < |n| < 2^-969, quotients outside the normal range),
// and for finite non-zero operands v_div_fixup only returns the signed zero
// of a zero numerator, which copysign(q, m) gives (m carries the quotient's
// sign).  Densities, momenta, pressures and their differences in the flow lie
// far inside that range, where div() is n / d bit for bit.  The fast build of
// the FP64 stage kernel (FSL_FP64_FAST) takes n r alone, within an ulp or so.
template <class T>
struct Divisor {
    T d;
    __device__ __forceinline__ explicit Divisor(T den) : d(den) {}
    __device__ __forceinline__ T div(T n) const { return n / d; }
};

template <>
struct Divisor<double> {
    double d, r;
    __device__ __forceinline__ explicit Divisor(double den) : d(den)
    {
        r = __builtin_amdgcn_rcp(den);
        double e = __builtin_fma(-den, r, 1.0);
        r = __builtin_fma(r, e, r);
        e = __builtin_fma(-den, r, 1.0);
        r = __builtin_fma(r, e, r);
    }
    __device__ __forceinline__ double div(double n) const
    {
#ifdef FSL_FP64_FAST
        return n * r;
#else
        const double m = n * r;
        const double rem = __builtin_fma(-d, m, n);
        return __builtin_copysign(__builtin_fma(rem, r, m), m);
#endif
    }
};

template <class T>
__device__ __forceinline__ T divide(T n, T d) { return Divisor<T>(d).div(n); }

// Square root.  For T = double: the compiler's correctly rounded sqrt is a
// pre-scaling of arguments below 2^-767 by 2^256, v_rsq, Goldschmidt steps on
// g = x y and h = y / 2 with two residual corrections, the matching rescale
// of the root and v_cmp_class passing +-0 and +inf through.  root() keeps
// every step but the scaling, which is an identity from 2^-767 up, so for
// those arguments and for zero, infinity, negatives and NaN it is sqrt(x) bit
// for bit.  (Squared sound speeds of the flow are of order 10^5.)
template <class T>
__device__ __forceinline__ T root(T x) { return sqrt(x); }

template <>
__device__ __forceinline__ double root(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    double d = __builtin_fma(-g, g, x);
    h = __builtin_fma(h, r, h);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return __builtin_amdgcn_class(x, 0x260) ? x : g;   // -0, +0, +inf
}

// Per-cell quantities of the characteristic transform,
// src/reconstruction.c:64-88, 104, 142-145
template <class T>
struct Prim {
    T q[5];
    T R, vel[3], P, C;
    T aa, ll, ee, gg, hh, nn, S11;
    T vhh[3], vee[3];
};

// R = 1 / q[0] and vel = q[1..3] R given: the values vel_recip computes for
// the same cell (the same operations), which a caller may carry over
template <class T>
__device__ __forceinline__ void cell_prims_rv(const T q[5], T R, const T vel[3], Prim<T>& c)
{
#pragma unroll
    for (int r = 0; r < 5; r++) c.q[r] = q[r];
    c.R = R;
    c.vel[0] = vel[0];
    c.vel[1] = vel[1];
    c.vel[2] = vel[2];
    c.P = (q[4] - T(0.5) * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) * c.R) * T(kK_1);
    c.C = root(T(kK) * c.P * c.R);
    T ff = c.vel[0] * c.vel[0] + c.vel[1] * c.vel[1] + c.vel[2] * c.vel[2];
    c.aa = T(kK_1_2) * ff;
    c.ll = divide(T(1.), c.C + c.C);
    c.ee = divide(c.ll, c.C);
    c.gg = -ff * c.ee;
    c.hh = -(c.ee + c.ee);
    c.S11 = c.aa - c.C * c.C;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        c.vhh[r] = c.vel[r] * c.hh;
        c.vee[r] = c.vel[r] * c.ee;
    }
    c.nn = T(-0.5) * c.gg;
}

template <class T>
__device__ __forceinline__ void cell_prims(const T q[5], Prim<T>& c)
{
    const T R = divide(T(1.), q[0]);
    const T vel[3] = {q[1] * R, q[2] * R, q[3] * R};
    cell_prims_rv(q, R, vel, c);
}

// Reconstruction along one direction (src/reconstruction.c:59-170 for x;
// y and z rotate the momenta, lines 176-256 and 261-341).  N, T1, T2 are the
// normal and tangential momentum indices (1..3): x (1,2,3), y (2,3,1),
// z (3,1,2).  qp / qm: conserved state of the +1 / -1 neighbour.
// plus: state on the cell's + face, minus: state on its - face.
template <int N, int T1, int T2, class T>
__device__ __forceinline__ void recon_dir(const Prim<T>& c, const T qp[5], const T qm[5], bool quirk,
                                          T plus[5], T minus[5])
{
    const int perm[5] = {0, N, T1, T2, 4};
    T m[5], w[5], M[5], W[5], Mh[5], Wh[5], kv[5];
#pragma unroll
    for (int r = 0; r < 5; r++) {
        m[r] = qp[perm[r]] - c.q[perm[r]];
        w[r] = c.q[perm[r]] - qm[perm[r]];
    }
    const T bb = T(kNK_1) * c.vel[N - 1];
    const T cc = T(kNK_1) * c.vel[T1 - 1];
    const T dd = T(kNK_1) * c.vel[T2 - 1];
    const T mm = c.vel[N - 1] * c.ll;
    T kk = c.C * c.vel[N - 1];
    const T S41 = c.aa - kk;
    const T S51 = c.aa + kk;
    const T S12 = bb;
    const T S42 = bb + c.C;
    const T S52 = bb - c.C;
    T k2, k3;
    kk = T(kK_1) * m[4]; k2 = cc * m[2]; k3 = dd * m[3];
    M[0] = c.S11 * m[0] + S12 * m[1] + k2 + k3 + kk;
    M[1] = -c.vel[T1 - 1] * m[0] + m[2];
    M[2] = -c.vel[T2 - 1] * m[0] + m[3];
    M[3] = S41 * m[0] + S42 * m[1] + k2 + k3 + kk;
    M[4] = S51 * m[0] + S52 * m[1] + k2 + k3 + kk;
    kk = T(kK_1) * w[4]; k2 = cc * w[2]; k3 = dd * w[3];
    W[0] = c.S11 * w[0] + S12 * w[1] + k2 + k3 + kk;
    W[1] = -c.vel[T1 - 1] * w[0] + w[2];
    W[2] = -c.vel[T2 - 1] * w[0] + w[3];
    W[3] = S41 * w[0] + S42 * w[1] + k2 + k3 + kk;
    W[4] = S51 * w[0] + S52 * w[1] + k2 + k3 + kk;
#pragma unroll
    for (int r = 0; r < 5; r++) {
        Mh[r] = minmod(M[r], W[r]);
        Wh[r] = minmod(W[r], M[r]);
    }
    const T S24 = c.vee[N - 1] + c.ll;
    const T S25 = c.vee[N - 1] - c.ll;
    const T S5 = c.nn + T(kI2K_1);
    const T S54 = S5 + mm;
    const T S55 = S5 - mm;
    const T ee_t2 = quirk ? c.vee[N - 1] : c.vee[T2 - 1];
#pragma unroll
    for (int r = 0; r < 5; r++) kv[r] = Mh[r] * T(kC1) + Wh[r] * T(kC2);
    kk = kv[3] + kv[4];
    plus[0] = c.q[0] + c.hh * kv[0] + c.ee * kk;
    plus[N] = c.q[N] + c.vhh[N - 1] * kv[0] + S24 * kv[3] + S25 * kv[4];
    plus[T1] = c.q[T1] + c.vhh[T1 - 1] * kv[0] + kv[1] + c.vee[T1 - 1] * kk;
    plus[T2] = c.q[T2] + c.vhh[T2 - 1] * kv[0] + kv[2] + ee_t2 * kk;
    plus[4] = c.q[4] + c.gg * kv[0] + c.vel[T1 - 1] * kv[1] + c.vel[T2 - 1] * kv[2] + S54 * kv[3] +
              S55 * kv[4];
#pragma unroll
    for (int r = 0; r < 5; r++) kv[r] = Mh[r] * T(kC2) + Wh[r] * T(kC1);
    kk = kv[3] + kv[4];
    minus[0] = c.q[0] - c.hh * kv[0] - c.ee * kk;
    minus[N] = c.q[N] - c.vhh[N - 1] * kv[0] - S24 * kv[3] - S25 * kv[4];
    minus[T1] = c.q[T1] - c.vhh[T1 - 1] * kv[0] - kv[1] - c.vee[T1 - 1] * kk;
    minus[T2] = c.q[T2] - c.vhh[T2 - 1] * kv[0] - kv[2] - ee_t2 * kk;
    minus[4] = c.q[4] - c.gg * kv[0] - c.vel[T1 - 1] * kv[1] - c.vel[T2 - 1] * kv[2] - S54 * kv[3] -
               S55 * kv[4];
}

// Linearised characteristic Riemann solver, src/fluxes.c:61-113 (x),
// 230-282 (y), 401-453 (z).  N: normal velocity index 0..2.
template <int N, class T>
__device__ __forceinline__ void riemann(const T l[5], const T r[5], T& Ro, T vel[3], T& Po)
{
    T lv[3], rv[3];
    const T lR = l[0], rR = r[0];
    const Divisor<T> byL(lR), byR(rR);   // four quotients over each density
    lv[0] = byL.div(l[1]); lv[1] = byL.div(l[2]); lv[2] = byL.div(l[3]);
    const T lP = (l[4] - T(0.5) * lR * (lv[0] * lv[0] + lv[1] * lv[1] + lv[2] * lv[2])) * T(kK_1);
    const T lC = root(byL.div(T(kK) * lP));
    const T ljo = T(kNK_1) * lP;
    T RU = divide(lP, lR * lC);
    const T ljp = lv[N] + RU;
    const T ljm = lv[N] - RU;
    rv[0] = byR.div(r[1]); rv[1] = byR.div(r[2]); rv[2] = byR.div(r[3]);
    const T rP = (r[4] - T(0.5) * rR * (rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2])) * T(kK_1);
    const T rC = root(byR.div(T(kK) * rP));
    const T rjo = T(kNK_1) * rP;
    RU = divide(rP, rR * rC);
    const T rjp = rv[N] + RU;
    const T rjm = rv[N] - RU;
    T Un = T(0.5) * (lv[N] + rv[N]);
    const T Cm = T(0.5) * (lC + rC);
    const T C_p = Un + Cm, C_m = Un - Cm, C_o = Un;
    // each coefficient from the side its characteristic comes from: the
    // operands are chosen first and divided once (the same operations on the
    // same values as the reference's branches, without evaluating both)
    const bool sp = C_p > T(0.), sm = C_m > T(0.), so = C_o > T(0.);
    const T al_p = divide(T(kK_1_2) * (sp ? ljm - ljp : rjm - rjp), sp ? ljo : rjo);
    const T J_p = sp ? ljp : rjp;
    const T al_m = divide(T(kK_1_2) * (sm ? ljp - ljm : rjp - rjm), sm ? ljo : rjo);
    const T J_m = sm ? ljm : rjm;
    const T a = T(kH05K) * (so ? ljp - ljm : rjp - rjm);
    const T al_o = -a * a;
    const T J_o = so ? ljo : rjo;
    const T P = divide(J_p - J_m, al_p - al_m);
    Un = J_p - al_p * P;
    Ro = divide(J_o - P, al_o);
    const bool left = Un > T(0);
#pragma unroll
    for (int t = 0; t < 3; t++) vel[t] = left ? lv[t] : rv[t];
    vel[N] = Un;
    Po = P;
}

// Cell velocity by reciprocal density, src/fluxes.c:124-127 (and the
// reconstruction's U = u2 * (1/u1), src/reconstruction.c:71-74).
template <class T>
__device__ __forceinline__ void vel_recip(T r, T m0, T m1, T m2, T v[3], T* rr = nullptr)
{
    const T RU = divide(T(1.), r);
    v[0] = m0 * RU; v[1] = m1 * RU; v[2] = m2 * RU;
    if (rr) *rr = RU;
}

// Cell velocity by division and temperature, src/fluxes.c:117-122
template <class T>
__device__ __forceinline__ void vel_div_T(T r, T m0, T m1, T m2, T e, T v[3], T& Tc)
{
    const Divisor<T> byR(r);
    v[0] = byR.div(m0); v[1] = byR.div(m1); v[2] = byR.div(m2);
    const T P = (e - T(0.5) * r * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2])) * T(kK_1);
    Tc = divide(P, T(kRair) * r);
}

// Viscous + SGS part and flux assembly of one face, src/fluxes.c:171-220
// (x), 340-391 (y), 511-568 (z).  g[c][e] = d vel_c / d x_e at the face.
// vf / vb: division velocities of the forward / backward cell, Tf / Tb their
// temperatures, Rf / Rb their densities.  mu_dyn = mu_T of the dynamic model.
template <int DIR, class T>
__device__ __forceinline__ void face_flux(const Phys& ph, T R, const T vel[3], T P, const T vf[3],
                                          const T vb[3], T Tf, T Tb, T Rf, T Rb, const T g[3][3],
                                          T mu_dyn, T F[5])
{
    const T uu = T(0.5) * (vf[0] + vb[0]);
    const T vv = T(0.5) * (vf[1] + vb[1]);
    const T ww = T(0.5) * (vf[2] + vb[2]);
    T mu_T;
    if (ph.dynamic) {
        mu_T = mu_dyn;
    } else {
        const T rr = T(0.5) * (Rf + Rb);
        const T S12 = T(0.5) * (g[0][1] + g[1][0]);
        const T S13 = T(0.5) * (g[0][2] + g[2][0]);
        const T S23 = T(0.5) * (g[1][2] + g[2][1]);
        const T S = sqrt(T(2) * (g[0][0] * g[0][0] + g[1][1] * g[1][1] + g[2][2] * g[2][2] +
                                 T(2) * (S12 * S12 + S13 * S13 + S23 * S23)));
        mu_T = rr * T(ph.CsDD) * S;
    }
    const T mu_E = T(ph.mu_L) + mu_T;
    T s0, s1, s2;
    if (DIR == 0) {
        s0 = T(kTwo3) * mu_E * (g[0][0] + g[0][0] - g[1][1] - g[2][2]);
        s1 = mu_E * (g[0][1] + g[1][0]);
        s2 = mu_E * (g[0][2] + g[2][0]);
    } else if (DIR == 1) {
        s0 = mu_E * (g[1][0] + g[0][1]);
        s1 = T(kTwo3) * mu_E * (g[1][1] + g[1][1] - g[0][0] - g[2][2]);
        s2 = mu_E * (g[1][2] + g[2][1]);
    } else {
        s0 = mu_E * (g[2][0] + g[0][2]);
        s1 = mu_E * (g[2][1] + g[1][2]);
        s2 = T(kTwo3) * mu_E * (g[2][2] + g[2][2] - g[0][0] - g[1][1]);
    }
    const T lambda_E = T(ph.lambda_L) + mu_T * T(ph.cp_Pr_T);
    const T q = -lambda_E * (Tf - Tb) * T(ph.rdel[DIR]);
    const T RU = R * vel[DIR];
    F[0] = RU;
    F[1] = (DIR == 0) ? RU * vel[0] + P - s0 : RU * vel[0] - s0;
    F[2] = (DIR == 1) ? RU * vel[1] + P - s1 : RU * vel[1] - s1;
    F[3] = (DIR == 2) ? RU * vel[2] + P - s2 : RU * vel[2] - s2;
    F[4] = vel[DIR] * (T(kK_K_1) * P + T(0.5) * R * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2])) -
           uu * s0 - vv * s1 - ww * s2 + q;
}

// Inflow state of ghost cells (src/boundCondInGhostCells.c:17-31):
// rho_in, u_in[row], v = w = 0, E = p/(K-1) + rho u^2 / 2
__device__ __forceinline__ void inflow_cell(const Bnd& b, int row, double q[5])
{
    const double R = b.rho_in, P = b.p_in, u = b.u_in[row - 1];
    q[0] = R;
    q[1] = R * u;
    q[2] = 0.;
    q[3] = 0.;
    q[4] = P / kK_1 + 0.5 * R * (u * u);
}

// Left state of inflow face 0 at face row j (0-based), z index k
// (src/bounCondOnInterfaces.c:17-58): the disturbance block covers rows
// [H/2, H/2 + BL_HIG) and k in [k_min, k_max) (or its complement when
// k_max <= k_min).
__device__ __forceinline__ void inflow_face(const Bnd& b, int H, int j, int k, double q[5])
{
    const double R = b.rho_in, P = b.p_in;
    double U = b.u_in[j], V = 0., W = 0.;
    if (j >= H / 2 && j < H / 2 + b.BL_HIG) {
        const unsigned ku = (unsigned)k;
        bool in_block;
        if (b.k_max > b.k_min) in_block = (ku >= b.k_min && ku < b.k_max);
        else in_block = !(ku >= b.k_max && ku < b.k_min);
        if (in_block) { U = U + b.Ud; V = b.Vd; W = b.Wd; }
    }
    q[0] = R;
    q[1] = R * U;
    q[2] = R * V;
    q[3] = R * W;
    q[4] = P / kK_1 + 0.5 * R * (U * U + V * V + W * W);
}

// Outflow: momenta kept, E rebuilt with p_out (src/boundCondInGhostCells.c:46-50,
// src/bounCondOnInterfaces.c:63-71)
__device__ __forceinline__ void outflow_state(double p_out, double q[5])
{
    const double R = q[0], u = q[1], v = q[2], w = q[3];
    q[4] = p_out / kK_1 + 0.5 * (u * u + v * v + w * w) / R;
}

// Per-cell quantity of the CFL reduction: mode 0 the reference's signed
// streamwise Courant number (rho u / rho) dt / dx (src/helpers.c:23-52),
// mode 1 the acoustic rate sum_d (|u_d| + c) / delta_d.
__device__ __forceinline__ double cfl_cell(int mode, const double q[5], const Phys& ph, double dtX)
{
    const double R = q[0];
    if (mode == 0) {
        const double u = q[1] / R;
        return u * dtX;
    }
    const double u = q[1] / R, vv = q[2] / R, w = q[3] / R;
    const double p = (q[4] - 0.5 * R * (u * u + vv * vv + w * w)) * kK_1;
    const double cs = sqrt(kK * p / R);
    return (fabs(u) + cs) * ph.rdel[0] + (fabs(vv) + cs) * ph.rdel[1] + (fabs(w) + cs) * ph.rdel[2];
}

// Running maximum of the CFL quantity.  The reference seeds the maximum with
// 0 and its comparison skips NaN cells (src/helpers.c:39-41); here a NaN
// cell makes the maximum NaN, so a partly non-finite state cannot give a
// finite time step.  Finite maxima are the reference's, bit for bit.
__device__ __forceinline__ double nan_max(double m, double v) { return (v > m || v != v) ? v : m; }

// Every lane contributes max(+0, value) and the wave maximum is folded into
// a 64-bit atomic max on the bit pattern, which orders non-negative doubles
// and puts a NaN (exponent all ones, non-zero mantissa) above +inf.
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = nan_max(v, __shfl_xor(v, o, 64));
    return v;
}

// wave maximum of m into *out (every lane of the wave calls this)
__device__ __forceinline__ void cfl_fold(double m, unsigned long long* out)
{
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

}  // namespace fsl

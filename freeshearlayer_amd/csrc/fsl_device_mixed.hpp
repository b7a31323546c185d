// fsl_device_mixed.hpp -- building blocks of the mixed-precision path.
//
// Same algorithm as fsl_device.hpp (src/reconstruction.c, src/fluxes.c), with
// the characteristic transform, limiter, Riemann solver and viscous/SGS flux
// in FP32 and these kept in FP64:
//   * the conserved state and every difference taken from it (the
//     reconstruction's forward/backward differences are formed in FP64 and
//     rounded once),
//   * the face energy (cell value in FP64 plus an FP32 correction); face
//     density and momenta are FP32 (FaceF),
//   * pressure recovered from total energy, p = (K-1)(E - rho|u|^2/2): E is
//     FP64 and only the kinetic part is FP32, so the cancellation at low Mach
//     costs no FP32 digits,
//   * the Runge-Kutta accumulation (fsl_fused_mixed.hip).
// FP32 reciprocals and square roots use the hardware approximations
// (v_rcp_f32, v_sqrt_f32, about 1 ulp) instead of correctly rounded
// sequences; contraction into FMA is allowed in these functions.
#pragma once

#include "fsl_device.hpp"

namespace fsl {
namespace mx {

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Limiter of src/reconstruction.c:17-24 as the median of {x, 0, 4y}: for x and
// y of one sign it is the one of x and 4y nearer zero, for opposite signs or a
// zero operand it is 0 -- the reference's cases, in one v_med3_f32 (no
// compare/select chain).
__device__ __forceinline__ float minmod(float x, float y) { return __builtin_amdgcn_fmed3f(x, 0.f, 4.f * y); }

// pressure from FP64 total energy and an FP32 rho |u|^2, halved and
// subtracted in one FP64 fma
__device__ __forceinline__ float pressure2(double E, float rho_uu)
{
    return (float)(__builtin_fma((double)rho_uu, -0.5, E) * kK_1);
}

// face state: FP32 density and momenta, FP64 total energy (pressure recovery)
struct FaceF {
    float f[4];
    double E;
};

__device__ __forceinline__ FaceF to_face(const double q[5])
{
    FaceF o;
#pragma unroll
    for (int r = 0; r < 4; r++) o.f[r] = (float)q[r];
    o.E = q[4];
    return o;
}

// outflow: momenta kept, E rebuilt with p_out (src/bounCondOnInterfaces.c:63-71)
__device__ __forceinline__ void outflow_face(double p_out, FaceF& q)
{
    const double R = q.f[0], u = q.f[1], v = q.f[2], w = q.f[3];
    q.E = p_out / kK_1 + 0.5 * (u * u + v * v + w * w) / R;
}

struct PrimF {
    float R, vel[3], P, C, aa, ll, ee, gg, hh, nn, S11;
    float vhh[3], vee[3], vc[3];
};

// Limiter weights (1 + kappa)/4 and (1 - kappa)/4 at kappa = 1/3: the FP32
// values satisfy C1 = 2 C2 exactly, so a face increment C1 a + C2 b is
// C2 (2a + b), and C2 is carried by the right-eigenvector coefficients.
constexpr float kC2f = (float)kC2;
static_assert((float)kC1 == 2.f * (float)kC2, "limiter weights");

// Per-cell quantities of the characteristic transform from the state and its
// 1/rho and velocity (computed once per cell by the caller); C and 1/C come
// from one reciprocal square root of K p / rho.  The right-eigenvector
// coefficients (ll, ee, gg, hh, nn, vhh, vee, vc) include the factor C2.
__device__ __forceinline__ void cell_prims(const double q[5], float Ri, const float v[3], PrimF& c)
{
#pragma clang fp contract(fast)
    const float r = (float)q[0];
    c.R = Ri;
    c.vel[0] = v[0];
    c.vel[1] = v[1];
    c.vel[2] = v[2];
    const float ff = c.vel[0] * c.vel[0] + c.vel[1] * c.vel[1] + c.vel[2] * c.vel[2];
    c.P = pressure2(q[4], r * ff);
    const float C2 = 1.4f * c.P * c.R;
    const float rC = __builtin_amdgcn_rsqf(C2);
    c.C = C2 * rC;
    c.aa = 0.2f * ff;
    c.ll = (0.5f * kC2f) * rC;
    c.ee = c.ll * rC;
    c.gg = -ff * c.ee;
    c.hh = -(c.ee + c.ee);
    c.S11 = c.aa - C2;
#pragma unroll
    for (int r2 = 0; r2 < 3; r2++) {
        c.vhh[r2] = c.vel[r2] * c.hh;
        c.vee[r2] = c.vel[r2] * c.ee;
        c.vc[r2] = c.vel[r2] * kC2f;
    }
    c.nn = -0.5f * c.gg;
}

// Characteristic reconstruction in direction N from the forward and backward
// differences dp = q(+1) - q, dm = q - q(-1) of the five fields (field order,
// FP64 differences rounded once).
template <int N, int T1, int T2>
__device__ __forceinline__ void recon_dir_d(const PrimF& c, const double q[5], const float qf[4], const float dp[5],
                                            const float dm[5], bool quirk, FaceF& plus, FaceF& minus)
{
#pragma clang fp contract(fast)
    const int perm[5] = {0, N, T1, T2, 4};
    float m[5], w[5], M[5], W[5], Mh[5], Wh[5], kv[5];
#pragma unroll
    for (int r = 0; r < 5; r++) {
        m[r] = dp[perm[r]];
        w[r] = dm[perm[r]];
    }
    const float bb = -0.4f * c.vel[N - 1];
    const float cc = -0.4f * c.vel[T1 - 1];
    const float dd = -0.4f * c.vel[T2 - 1];
    const float mm = c.vel[N - 1] * c.ll;
    float kk = c.C * c.vel[N - 1];
    const float S41 = c.aa - kk, S51 = c.aa + kk;
    const float S12 = bb, S42 = bb + c.C, S52 = bb - c.C;
    // rows 1, 4, 5 of the left eigenvectors share the terms of columns 3-5
    float t = dd * m[3] + (cc * m[2] + 0.4f * m[4]);
    M[0] = c.S11 * m[0] + (S12 * m[1] + t);
    M[1] = -c.vel[T1 - 1] * m[0] + m[2];
    M[2] = -c.vel[T2 - 1] * m[0] + m[3];
    M[3] = S41 * m[0] + (S42 * m[1] + t);
    M[4] = S51 * m[0] + (S52 * m[1] + t);
    t = dd * w[3] + (cc * w[2] + 0.4f * w[4]);
    W[0] = c.S11 * w[0] + (S12 * w[1] + t);
    W[1] = -c.vel[T1 - 1] * w[0] + w[2];
    W[2] = -c.vel[T2 - 1] * w[0] + w[3];
    W[3] = S41 * w[0] + (S42 * w[1] + t);
    W[4] = S51 * w[0] + (S52 * w[1] + t);
#pragma unroll
    for (int r = 0; r < 5; r++) {
        Mh[r] = minmod(M[r], W[r]);
        Wh[r] = minmod(W[r], M[r]);
    }
    const float S24 = c.vee[N - 1] + c.ll, S25 = c.vee[N - 1] - c.ll;
    const float S5 = c.nn + 1.25f * kC2f;
    const float S54 = S5 + mm, S55 = S5 - mm;
    const float ee_t2 = quirk ? c.vee[N - 1] : c.vee[T2 - 1];
    float d[5];
    // kv = (C1 Mh + C2 Wh) / C2 for the + face, (C2 Mh + C1 Wh) / C2 for the - face
#pragma unroll
    for (int r = 0; r < 5; r++) kv[r] = 2.f * Mh[r] + Wh[r];
    kk = kv[3] + kv[4];
    d[0] = c.hh * kv[0] + c.ee * kk;
    d[N] = c.vhh[N - 1] * kv[0] + S24 * kv[3] + S25 * kv[4];
    d[T1] = c.vhh[T1 - 1] * kv[0] + kC2f * kv[1] + c.vee[T1 - 1] * kk;
    d[T2] = c.vhh[T2 - 1] * kv[0] + kC2f * kv[2] + ee_t2 * kk;
    d[4] = c.gg * kv[0] + c.vc[T1 - 1] * kv[1] + c.vc[T2 - 1] * kv[2] + S54 * kv[3] + S55 * kv[4];
#pragma unroll
    for (int r = 0; r < 4; r++) plus.f[r] = qf[r] + d[r];
    plus.E = q[4] + (double)d[4];
#pragma unroll
    for (int r = 0; r < 5; r++) kv[r] = Mh[r] + 2.f * Wh[r];
    kk = kv[3] + kv[4];
    d[0] = c.hh * kv[0] + c.ee * kk;
    d[N] = c.vhh[N - 1] * kv[0] + S24 * kv[3] + S25 * kv[4];
    d[T1] = c.vhh[T1 - 1] * kv[0] + kC2f * kv[1] + c.vee[T1 - 1] * kk;
    d[T2] = c.vhh[T2 - 1] * kv[0] + kC2f * kv[2] + ee_t2 * kk;
    d[4] = c.gg * kv[0] + c.vc[T1 - 1] * kv[1] + c.vc[T2 - 1] * kv[2] + S54 * kv[3] + S55 * kv[4];
#pragma unroll
    for (int r = 0; r < 4; r++) minus.f[r] = qf[r] - d[r];
    minus.E = q[4] - (double)d[4];
}

template <int N, int T1, int T2>
__device__ __forceinline__ void recon_dir(const PrimF& c, const double q[5], const float qf[4], const double qp[5],
                                          const double qm[5], bool quirk, FaceF& plus, FaceF& minus)
{
    float dp[5], dm[5];
#pragma unroll
    for (int r = 0; r < 5; r++) {
        dp[r] = (float)(qp[r] - q[r]);
        dm[r] = (float)(q[r] - qm[r]);
    }
    recon_dir_d<N, T1, T2>(c, q, qf, dp, dm, quirk, plus, minus);
}

// One side of a face for the Riemann solver: velocity, pressure (from the
// FP64 energy), sound speed C, t = 1/C, A = 1/(rho C) and p/(rho C).
struct SideF {
    float v[3], P, C, t, A, PA;
};

__device__ __forceinline__ SideF face_side(const FaceF& q)
{
#pragma clang fp contract(fast)
    SideF s;
    const float R = q.f[0], Ri = rcp(R);
    s.v[0] = q.f[1] * Ri; s.v[1] = q.f[2] * Ri; s.v[2] = q.f[3] * Ri;
    // rho |u|^2 as momentum . velocity
    s.P = pressure2(q.E, q.f[1] * s.v[0] + q.f[2] * s.v[1] + q.f[3] * s.v[2]);
    const float C2 = 1.4f * s.P * Ri;
    s.t = __builtin_amdgcn_rsqf(C2);
    s.C = C2 * s.t;
    s.A = Ri * s.t;
    s.PA = s.P * s.A;
    return s;
}

// Linearised Riemann solver of src/fluxes.c:129-170 (as fsl_device.hpp's
// riemann), written with C^2 = K p / rho so that its terms need one
// reciprocal and one reciprocal square root per side:
//   RU = p / (rho C)                    (the Riemann invariants J+- = u_n +- RU)
//   al+- = (K-1)/2 (J-+ - J+-) / J_o = +-RU / p = +-1 / (rho C),  J_o = -(K-1) p
//   al_o = -(K/2 (J+ - J-))^2 = -C^2, so (J_o - P) / al_o = (P - J_o) t^2
// Same upwind choices (on u_n + C_m, u_n - C_m, u_n and the new u_n).
template <int N>
__device__ __forceinline__ void riemann(const FaceF& l, const FaceF& r, float& Ro, float vel[3], float& Po)
{
#pragma clang fp contract(fast)
    const SideF L = face_side(l), Rs = face_side(r);
    float Un = 0.5f * (L.v[N] + Rs.v[N]);
    const float Cm = 0.5f * (L.C + Rs.C);
    const bool sp = Un + Cm > 0.f, sm = Un - Cm > 0.f, so = Un > 0.f;
    const float al_p = sp ? L.A : Rs.A;
    const float J_p = sp ? L.v[N] + L.PA : Rs.v[N] + Rs.PA;
    const float am = sm ? L.A : Rs.A;                     // al_m = -am
    const float J_m = sm ? L.v[N] - L.PA : Rs.v[N] - Rs.PA;
    const float to = so ? L.t : Rs.t;
    const float J_o = -0.4f * (so ? L.P : Rs.P);
    const float P = (J_p - J_m) * rcp(al_p + am);
    Un = J_p - al_p * P;
    Ro = (P - J_o) * to * to;
    const bool left = Un > 0.f;
#pragma unroll
    for (int t = 0; t < 3; t++) vel[t] = left ? L.v[t] : Rs.v[t];
    vel[N] = Un;
    Po = P;
}

// cell temperature from the pressure and 1/rho of cell_prims
__device__ __forceinline__ float cell_temp(const PrimF& c) { return c.P * c.R * (float)(1.0 / kRair); }

template <int DIR>
__device__ __forceinline__ void face_flux(const Phys& ph, float R, const float vel[3], float P, const float vf[3],
                                          const float vb[3], float Tf, float Tb, float Rf, float Rb,
                                          const float g[3][3], float mu_dyn, float F[5])
{
#pragma clang fp contract(fast)
    // twice the face velocity (the stress work takes the factor 1/2 once)
    const float uu = vf[0] + vb[0];
    const float vv = vf[1] + vb[1];
    const float ww = vf[2] + vb[2];
    float mu_T;
    if (ph.dynamic) {
        mu_T = mu_dyn;
    } else {
        const float rr = 0.5f * (Rf + Rb);
        const float S12 = 0.5f * (g[0][1] + g[1][0]);
        const float S13 = 0.5f * (g[0][2] + g[2][0]);
        const float S23 = 0.5f * (g[1][2] + g[2][1]);
        const float S = fsqrt(2.f * (g[0][0] * g[0][0] + g[1][1] * g[1][1] + g[2][2] * g[2][2] +
                                     2.f * (S12 * S12 + S13 * S13 + S23 * S23)));
        mu_T = rr * (float)ph.CsDD * S;
    }
    const float mu_E = (float)ph.mu_L + mu_T;
    const float t3 = (float)kTwo3;
    float s0, s1, s2;
    if (DIR == 0) {
        s0 = t3 * mu_E * (g[0][0] + g[0][0] - g[1][1] - g[2][2]);
        s1 = mu_E * (g[0][1] + g[1][0]);
        s2 = mu_E * (g[0][2] + g[2][0]);
    } else if (DIR == 1) {
        s0 = mu_E * (g[1][0] + g[0][1]);
        s1 = t3 * mu_E * (g[1][1] + g[1][1] - g[0][0] - g[2][2]);
        s2 = mu_E * (g[1][2] + g[2][1]);
    } else {
        s0 = mu_E * (g[2][0] + g[0][2]);
        s1 = mu_E * (g[2][1] + g[1][2]);
        s2 = t3 * mu_E * (g[2][2] + g[2][2] - g[0][0] - g[1][1]);
    }
    // heat flux -(lambda_L + mu_T cp/Pr_T) (Tf - Tb) / delta, the constants scaled by -1/delta
    const float lq = mu_T * (float)(-ph.cp_Pr_T * ph.rdel[DIR]) + (float)(-ph.lambda_L * ph.rdel[DIR]);
    const float q = lq * (Tf - Tb);
    const float RU = R * vel[DIR];
    F[0] = RU;
    F[1] = RU * vel[0] - s0;
    F[2] = RU * vel[1] - s1;
    F[3] = RU * vel[2] - s2;
    F[1 + DIR] += P;
    F[4] = vel[DIR] * (3.5f * P + 0.5f * R * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2])) -
           0.5f * (uu * s0 + vv * s1 + ww * s2) + q;
}

// CFL quantity of a cell (fsl_device.hpp's cfl_cell, src/helpers.c:23-52) in
// FP64: 1/rho and sqrt(K p / rho) from v_rcp_f64 and v_rsq_f64 (about 23
// correct bits) refined by one Newton step each, about 1e-14 relative (the
// quantity only sets dt = CFL / max); zero and negative K p / rho give 0 and
// NaN as sqrt does.
__device__ __forceinline__ double cfl_cell(int mode, const double q[5], const Phys& ph, double dtX)
{
    const double R = q[0];
    double Ri = __builtin_amdgcn_rcp(R);
    Ri = __builtin_fma(Ri, __builtin_fma(-R, Ri, 1.0), Ri);
    if (mode == 0) return q[1] * Ri * dtX;
    const double u = q[1] * Ri, vv = q[2] * Ri, w = q[3] * Ri;
    const double p = (q[4] - 0.5 * R * (u * u + vv * vv + w * w)) * kK_1;
    const double t = kK * p * Ri;
    double y = __builtin_amdgcn_rsq(t);
    y = y * __builtin_fma(-0.5 * t * y, y, 1.5);
    const double cs = t > 0.0 ? t * y : (t == 0.0 ? 0.0 : __builtin_nan(""));
    return (fabs(u) + cs) * ph.rdel[0] + (fabs(vv) + cs) * ph.rdel[1] + (fabs(w) + cs) * ph.rdel[2];
}

}  // namespace mx
}  // namespace fsl

// fsl_sgs.hip -- dynamic Smagorinsky model (Germano-Lilly, pointwise Cd
// clipped to [0, 0.15]), src/turbulence.c:17-507.
//
// The reference filters 21 fields per call.  The velocity products and the
// A tensor are symmetric and their filtered copies are bit-identical
// (u v == v u in IEEE arithmetic), so 15 filters give the same result.  The
// model reads the step-start state only (src/fluxes.c:53 passes U1..U4,
// whose ghost layer is filled at stage 1 and unchanged until the end of
// stage 3), so the fused path evaluates it once per step.
//
// Filter semantics reproduced exactly (src/turbulence.c:47-95): three
// separable passes over index boxes that leave zeros outside them:
//   y pass  t1 on i in [0,L], j in [1,H], k in [0,D]
//   z pass  t2 on i in [0,L], j in [0,H], k in [1,D]
//   x pass  out on i in [1,L], j in [0,H], k in [0,D]; elsewhere out = t1
// with each sum accumulated as ((0 + a h) + b h) + c h.
//
// The FP64 model (the reference's operations, bit-identical) is one x-marching
// kernel, fsl_sgs_march.hip.  The FP32 model of mixed-precision contexts is
// the two tiled kernels here.  Each workgroup owns a 2 x 8 x 32 tile of
// interior cells and stages a one-cell-halo box in LDS:
//   k_sgs_a  U -> u, v, w on the box; per field (u, v, w and the six
//            products) the y, z and x passes on the box; Leonard stress,
//            |S| and A = |S| S at the tile's cells; writes u^, L, A, |S|
//            (and u^ on the i = 0 plane, which keeps its y-pass value).
//   k_sgs_b  A^ by the same three passes, S^ from u^, B = |S^| S^,
//            M = Delta^2 (4 B - A^), Cd and mu_SGS.
// Every intermediate value is computed by the same operations on the same
// inputs as the pass-by-pass form, so mu_SGS is bit-identical to it.
#include <algorithm>

#include "fsl_internal.hpp"

namespace fsl {

namespace {

constexpr int kBlock = 256;

inline unsigned nblk(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

struct Ptr9 {
    double* p[9];
};
struct CPtr9 {
    const double* p[9];
};
struct H3 {
    double h[3];
};

__device__ __forceinline__ void decode(const Geo& g, long long t, int& i, int& j, int& k)
{
    k = (int)(t % (g.D + 2));
    j = (int)((t / (g.D + 2)) % (g.H + 2));
    i = (int)(t / g.si);
}

// ---- pass-by-pass filter (fsl_filter: any cell array, any kernel h) ----
__global__ void __launch_bounds__(kBlock) k_filt_y(CPtr9 in, Ptr9 t1, Geo g, H3 w)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.G) return;
    const int a = blockIdx.y;
    int i, j, k;
    decode(g, t, i, j, k);
    double r = 0.;
    if (i <= g.L && j >= 1 && j <= g.H && k <= g.D) {
        const double* s = in.p[a];
        double tmp = 0.;
        tmp += s[t - g.sj] * w.h[0];
        tmp += s[t] * w.h[1];
        tmp += s[t + g.sj] * w.h[2];
        r = tmp;
    }
    t1.p[a][t] = r;
}

__global__ void __launch_bounds__(kBlock) k_filt_z(CPtr9 t1, Ptr9 t2, Geo g, H3 w)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.G) return;
    const int a = blockIdx.y;
    int i, j, k;
    decode(g, t, i, j, k);
    double r = 0.;
    if (i <= g.L && j <= g.H && k >= 1 && k <= g.D) {
        const double* s = t1.p[a];
        double tmp = 0.;
        tmp += s[t - 1] * w.h[0];
        tmp += s[t] * w.h[1];
        tmp += s[t + 1] * w.h[2];
        r = tmp;
    }
    t2.p[a][t] = r;
}

__global__ void __launch_bounds__(kBlock) k_filt_x(CPtr9 t2, CPtr9 t1, Ptr9 out, Geo g, H3 w)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.G) return;
    const int a = blockIdx.y;
    int i, j, k;
    decode(g, t, i, j, k);
    double r;
    if (i >= 1 && i <= g.L && j <= g.H && k <= g.D) {
        const double* s = t2.p[a];
        double tmp = 0.;
        tmp += s[t - g.si] * w.h[0];
        tmp += s[t] * w.h[1];
        tmp += s[t + g.si] * w.h[2];
        r = tmp;
    } else {
        r = t1.p[a][t];
    }
    out.p[a][t] = r;
}

// ---- tiled model ----
constexpr int TI = 2, TJ = 8, TK = 32;                 // interior cells per workgroup
constexpr int BI = TI + 2, BJ = TJ + 2, BK = TK + 2;   // box with a one-cell halo
constexpr int NB = BI * BJ * BK;
constexpr int NT1 = BI * TJ * BK;                      // y-pass values on the box
constexpr int NT2 = BI * TJ * TK;                      // z-pass values on the box
constexpr int NC = TI * TJ * TK;                       // tile cells
constexpr int CPT = NC / kBlock;                       // cells per thread
static_assert(NC % kBlock == 0, "tile must divide the workgroup");

struct Tile {
    int i0, j0, k0;
};

__device__ __forceinline__ Tile tile_of(const Geo& g)
{
    const int ntk = (g.D + TK - 1) / TK, ntj = (g.H + TJ - 1) / TJ;
    int b = blockIdx.x;
    const int tk = b % ntk;
    b /= ntk;
    const int tj = b % ntj;
    const int ti = b / ntj;
    return Tile{1 + ti * TI, 1 + tj * TJ, 1 + tk * TK};
}

__device__ __forceinline__ int bidx(int bi, int bj, int bk) { return (bi * BJ + bj) * BK + bk; }

// field a of {u, v, w, uu, uv, uw, vv, vw, ww} at box cell b
template <class T>
__device__ __forceinline__ T field9(const T (*s)[NB], int a, int b)
{
    switch (a) {
    case 0: return s[0][b];
    case 1: return s[1][b];
    case 2: return s[2][b];
    case 3: return s[0][b] * s[0][b];
    case 4: return s[0][b] * s[1][b];
    case 5: return s[0][b] * s[2][b];
    case 6: return s[1][b] * s[1][b];
    case 7: return s[1][b] * s[2][b];
    default: return s[2][b] * s[2][b];
    }
}

// The three passes of the box filter for one field whose box values come
// from src(b).  out[q] receives the filtered value at this thread's tile
// cells; t1i0 (optional) the y-pass value on the i = 0 plane.
template <class R, class Src>
__device__ __forceinline__ void box_filter(const Geo& g, const Tile& T, R h, Src src, R* s_t1, R* s_t2,
                                           R out[CPT], R* t1i0_out, int tid, R t1i0_add = 0)
{
    for (int e = tid; e < NT1; e += kBlock) {
        const int bk = e % BK, tj = (e / BK) % TJ, bi = e / (BK * TJ);
        const int gi = T.i0 - 1 + bi, gj = T.j0 + tj, gk = T.k0 - 1 + bk;
        R r = 0;
        if (gi <= g.L && gj <= g.H && gk <= g.D) {
            R tmp = 0;
            tmp += src(bidx(bi, tj, bk)) * h;
            tmp += src(bidx(bi, tj + 1, bk)) * h;
            tmp += src(bidx(bi, tj + 2, bk)) * h;
            r = tmp;
        }
        s_t1[e] = r;
    }
    __syncthreads();
    for (int e = tid; e < NT2; e += kBlock) {
        const int tk = e % TK, tj = (e / TK) % TJ, bi = e / (TK * TJ);
        const int gi = T.i0 - 1 + bi, gj = T.j0 + tj, gk = T.k0 + tk;
        R r = 0;
        if (gi <= g.L && gj <= g.H && gk <= g.D) {
            const int b1 = (bi * TJ + tj) * BK + tk;
            R tmp = 0;
            tmp += s_t1[b1] * h;
            tmp += s_t1[b1 + 1] * h;
            tmp += s_t1[b1 + 2] * h;
            r = tmp;
        }
        s_t2[e] = r;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int e = tid + q * kBlock;
        const int tk = e % TK, tj = (e / TK) % TJ, ti = e / (TK * TJ);
        const int b2 = (ti * TJ + tj) * TK + tk;
        R tmp = 0;
        tmp += s_t2[b2] * h;
        tmp += s_t2[b2 + TJ * TK] * h;
        tmp += s_t2[b2 + 2 * TJ * TK] * h;
        out[q] = tmp;
    }
    if (t1i0_out && T.i0 == 1) {
        // u^ on the i = 0 plane keeps the y-pass value (the x pass starts at i = 1)
        for (int e = tid; e < TJ * TK; e += kBlock) {
            const int tk = e % TK, tj = e / TK;
            const int gj = T.j0 + tj, gk = T.k0 + tk;
            if (gj <= g.H && gk <= g.D) {
                if (sizeof(R) == sizeof(double)) t1i0_out[(long long)gj * g.sj + gk] = s_t1[(0 * TJ + tj) * BK + tk + 1];
                else t1i0_out[(long long)gj * g.sj + gk] = s_t1[(0 * TJ + tj) * BK + tk + 1] + t1i0_add;
            }
        }
    }
    __syncthreads();
}

template <class R>
__device__ __forceinline__ void strain_box(const R (*s)[NB], int b, const Phys& ph, R S[9])
{
    R gr[3][3];
    constexpr int sI = BJ * BK, sJ = BK;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        gr[r][0] = (s[r][b + sI] - s[r][b - sI]) * (R)ph.r2del[0];
        gr[r][1] = (s[r][b + sJ] - s[r][b - sJ]) * (R)ph.r2del[1];
        gr[r][2] = (s[r][b + 1] - s[r][b - 1]) * (R)ph.r2del[2];
    }
    S[0] = gr[0][0];
    S[1] = (R)0.5 * (gr[0][1] + gr[1][0]);
    S[2] = (R)0.5 * (gr[0][2] + gr[2][0]);
    S[3] = S[1];
    S[4] = gr[1][1];
    S[5] = (R)0.5 * (gr[1][2] + gr[2][1]);
    S[6] = S[2];
    S[7] = S[5];
    S[8] = gr[2][2];
}

// FP64: correctly rounded sqrt, 1/x and x/y (the reference's operations);
// FP32: the hardware approximations (v_sqrt_f32, v_rcp_f32, about 1 ulp)
__device__ __forceinline__ double sqrt_r(double x) { return sqrt(x); }
__device__ __forceinline__ float sqrt_r(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double rcp_r(double x) { return 1. / x; }
__device__ __forceinline__ float rcp_r(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double div_r(double a, double b) { return a / b; }
__device__ __forceinline__ float div_r(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

template <class R>
__device__ __forceinline__ R mag9(const R S[9])
{
    return sqrt_r((R)2 * (S[0] * S[0] + S[1] * S[1] + S[2] * S[2] + S[3] * S[3] + S[4] * S[4] + S[5] * S[5] +
                          S[6] * S[6] + S[7] * S[7] + S[8] * S[8]));
}

template <class R>
struct SgsA {
    CState5 U;
    R* uf[3];
    R* Lij[6];
    R* Aij[6];
    R* mag;
};

template <class R>
__global__ void __launch_bounds__(kBlock) k_sgs_a(SgsA<R> a, Geo g, Phys ph)
{
    __shared__ R s_u[3][NB];
    __shared__ R s_t1[NT1];
    __shared__ R s_t2[NT2];
    const Tile T = tile_of(g);
    const int tid = threadIdx.x;
    // u, v, w = (rho u, rho v, rho w) * (1 / rho) on the box (src/turbulence.c:197-209)
    for (int e = tid; e < NB; e += kBlock) {
        const int bk = e % BK, bj = (e / BK) % BJ, bi = e / (BK * BJ);
        const int gi = T.i0 - 1 + bi, gj = T.j0 - 1 + bj, gk = T.k0 - 1 + bk;
        R u = 0, v = 0, w = 0;
        if (gi <= g.L + 1 && gj <= g.H + 1 && gk <= g.D + 1) {
            const long long c = cidx(g, gi, gj, gk);
            const R rr = rcp_r((R)a.U.f[0][c]);
            u = (R)a.U.f[1][c] * rr;
            v = (R)a.U.f[2][c] * rr;
            w = (R)a.U.f[3][c] * rr;
        }
        s_u[0][e] = u;
        s_u[1][e] = v;
        s_u[2][e] = w;
    }
    __syncthreads();
    const R h = (R)(1. / 3.);
    // FP32: the box's velocities are shifted by the tile-centre velocity a,
    // u = u' + a, so the products lose no digits to |u| >> |u'|.  With W the
    // filter's weight sum at a cell (the three passes applied to a field of
    // ones: about 1 where every window is complete, 2/3 on the planes k = D
    // and i = L, whose z and x windows the truncated passes cut),
    //   u^ = f(u') + a W
    //   L_ij = f(u'_i u'_j) - f(u'_i) f(u'_j) + (1 - W)(a_i f(u'_j) + a_j f(u'_i)) + a_i a_j (W - W^2),
    // the same values as f(u_i u_j) - u^_i u^_j.  The strain is a difference
    // and does not see the shift.
    R sh[3] = {0, 0, 0}, w1 = 0, w2f = 0, w2t = 0;
    if (sizeof(R) == sizeof(float)) {
        const int bc = bidx(1 + TI / 2, 1 + TJ / 2, 1 + TK / 2);
#pragma unroll
        for (int q = 0; q < 3; q++) sh[q] = s_u[q][bc];
        __syncthreads();
        for (int e = tid; e < NB; e += kBlock) {
#pragma unroll
            for (int q = 0; q < 3; q++) s_u[q][e] -= sh[q];
        }
        __syncthreads();
        w1 = (((R)0 + h) + h) + h;                       // y pass: always three valid rows
        w2f = (((R)0 + w1 * h) + w1 * h) + w1 * h;       // z pass, k < D
        w2t = ((R)0 + w1 * h) + w1 * h;                  // z pass, k = D (plane D + 1 of the y pass is 0)
    }
    R fil[9][CPT];
#pragma unroll
    for (int f = 0; f < 9; f++) {
        box_filter<R>(g, T, h, [&](int b) { return field9(s_u, f, b); }, s_t1, s_t2, fil[f],
                      f < 3 ? a.uf[f] : nullptr, tid, f < 3 ? sh[f] * w1 : (R)0);
    }
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int e = tid + q * kBlock;
        const int tk = e % TK, tj = (e / TK) % TJ, ti = e / (TK * TJ);
        const int gi = T.i0 + ti, gj = T.j0 + tj, gk = T.k0 + tk;
        if (gi > g.L || gj > g.H || gk > g.D) continue;
        const long long c = cidx(g, gi, gj, gk);
        const R uf = fil[0][q], vf = fil[1][q], wf = fil[2][q];
        if (sizeof(R) == sizeof(double)) {
            a.uf[0][c] = uf;
            a.uf[1][c] = vf;
            a.uf[2][c] = wf;
            a.Lij[0][c] = fil[3][q] - uf * uf;
            a.Lij[1][c] = fil[4][q] - uf * vf;
            a.Lij[2][c] = fil[5][q] - uf * wf;
            a.Lij[3][c] = fil[6][q] - vf * vf;
            a.Lij[4][c] = fil[7][q] - vf * wf;
            a.Lij[5][c] = fil[8][q] - wf * wf;
        } else {
            const R w2 = gk < g.D ? w2f : w2t;                                  // t2 on planes i-1, i, i+1
            const R W = (((R)0 + w2 * h) + w2 * h) + (gi < g.L ? w2 : (R)0) * h;   // plane L + 1 of t2 is 0
            const R om = (R)1 - W, ww = W - W * W;
            const R fu[3] = {uf, vf, wf};
#pragma unroll
            for (int r = 0; r < 3; r++) a.uf[r][c] = fu[r] + sh[r] * W;
            const int pr[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
#pragma unroll
            for (int m = 0; m < 6; m++) {
                const int x = pr[m][0], y = pr[m][1];
                a.Lij[m][c] = fil[3 + m][q] - fu[x] * fu[y] + om * (sh[x] * fu[y] + sh[y] * fu[x]) + sh[x] * sh[y] * ww;
            }
        }
        R S[9];
        strain_box(s_u, bidx(ti + 1, tj + 1, tk + 1), ph, S);
        const R Sm = mag9(S);
        a.mag[c] = Sm;
        const int uniq[6] = {0, 1, 2, 4, 5, 8};   // 11, 12, 13, 22, 23, 33
#pragma unroll
        for (int m = 0; m < 6; m++) a.Aij[m][c] = Sm * S[uniq[m]];
    }
}

template <class R>
struct SgsB {
    const double* rho;
    const R* uf[3];
    const R* Lij[6];
    const R* Aij[6];
    const R* mag;
    double* mu;
    float* cd;   // optional: the clipped coefficient (training targets)
    double* lm;  // averaged model: L:M and M:M per cell (mu then holds rho Delta^2 |S|), or null
    double* mm;
};

template <class R>
__global__ void __launch_bounds__(kBlock) k_sgs_b(SgsB<R> a, Geo g, Phys ph)
{
    __shared__ R s_uf[3][NB];
    __shared__ R s_a[NB];
    __shared__ R s_t1[NT1];
    __shared__ R s_t2[NT2];
    const Tile T = tile_of(g);
    const int tid = threadIdx.x;
    for (int e = tid; e < NB; e += kBlock) {
        const int bk = e % BK, bj = (e / BK) % BJ, bi = e / (BK * BJ);
        const int gi = T.i0 - 1 + bi, gj = T.j0 - 1 + bj, gk = T.k0 - 1 + bk;
        R u = 0, v = 0, w = 0;
        if (gi <= g.L + 1 && gj <= g.H + 1 && gk <= g.D + 1) {
            const long long c = cidx(g, gi, gj, gk);
            u = a.uf[0][c];
            v = a.uf[1][c];
            w = a.uf[2][c];
        }
        s_uf[0][e] = u;
        s_uf[1][e] = v;
        s_uf[2][e] = w;
    }
    const R h = (R)(1. / 3.);
    R Af[6][CPT];
#pragma unroll
    for (int m = 0; m < 6; m++) {
        __syncthreads();   // previous users of s_a are done
        for (int e = tid; e < NB; e += kBlock) {
            const int bk = e % BK, bj = (e / BK) % BJ, bi = e / (BK * BJ);
            const int gi = T.i0 - 1 + bi, gj = T.j0 - 1 + bj, gk = T.k0 - 1 + bk;
            s_a[e] = (gi <= g.L + 1 && gj <= g.H + 1 && gk <= g.D + 1) ? a.Aij[m][cidx(g, gi, gj, gk)] : (R)0;
        }
        __syncthreads();
        box_filter<R>(g, T, h, [&](int b) { return s_a[b]; }, s_t1, s_t2, Af[m], nullptr, tid);
    }
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int e = tid + q * kBlock;
        const int tk = e % TK, tj = (e / TK) % TJ, ti = e / (TK * TJ);
        const int gi = T.i0 + ti, gj = T.j0 + tj, gk = T.k0 + tk;
        if (gi > g.L || gj > g.H || gk > g.D) continue;
        const long long c = cidx(g, gi, gj, gk);
        R Sf[9];
        strain_box(s_uf, bidx(ti + 1, tj + 1, tk + 1), ph, Sf);
        const R Sfm = mag9(Sf);
        // 3x3 from the six distinct components: 11 12 13 / 21=12 22 23 / 31=13 32=23 33
        const int map9[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
        const int uniq[6] = {0, 1, 2, 4, 5, 8};
        R B6[6], M[9], Lm[9];
#pragma unroll
        for (int m = 0; m < 6; m++) B6[m] = Sfm * Sf[uniq[m]];
#pragma unroll
        for (int r = 0; r < 9; r++) {
            M[r] = (R)ph.DD * ((R)4 * B6[map9[r]] - Af[map9[r]][q]);
            Lm[r] = a.Lij[map9[r]][c];
        }
        const R tr = -(R)(1.0 / 3.0) * (Lm[0] + Lm[4] + Lm[8]);
        Lm[0] += tr;
        Lm[4] += tr;
        Lm[8] += tr;
        const R LLMM = (Lm[0] * M[0] + Lm[1] * M[1] + Lm[2] * M[2] + Lm[3] * M[3] + Lm[4] * M[4] +
                             Lm[5] * M[5] + Lm[6] * M[6] + Lm[7] * M[7] + Lm[8] * M[8]);
        const R MMMM = (M[0] * M[0] + M[1] * M[1] + M[2] * M[2] + M[3] * M[3] + M[4] * M[4] +
                             M[5] * M[5] + M[6] * M[6] + M[7] * M[7] + M[8] * M[8]);
        if (a.lm) {   // averaged model: the terms, and the factor of the averaged Cd
            a.lm[c] = (double)LLMM;
            a.mm[c] = (double)MMMM;
            a.mu[c] = a.rho[c] * ph.DD * (double)a.mag[c];
            continue;
        }
        R Cd = -(R)0.5 * div_r(LLMM, MMMM + (R)kSmall);
        if (Cd > (R)0.15) Cd = (R)0.15;
        else if (Cd < (R)0.0) Cd = (R)0.0;
        if (sizeof(R) == sizeof(double)) a.mu[c] = a.rho[c] * (double)Cd * ph.DD * (double)a.mag[c];
        else a.mu[c] = (double)((R)a.rho[c] * Cd * (R)ph.DD * a.mag[c]);
        if (a.cd) a.cd[c] = (float)Cd;
    }
}

void run_filters(hipStream_t s, int nf, const double* const* in, double* const* out, double* const* t1,
                 double* const* t2, const Geo& g, const double h[3], void (*prof)(void*, int, int),
                 void* pa)
{
    const H3 w{{h[0], h[1], h[2]}};
    CPtr9 cin{}, ct1{}, ct2{};
    Ptr9 pt1{}, pt2{}, pout{};
    for (int a = 0; a < nf; a++) {
        cin.p[a] = in[a];
        pt1.p[a] = t1[a];
        ct1.p[a] = t1[a];
        pt2.p[a] = t2[a];
        ct2.p[a] = t2[a];
        pout.p[a] = out[a];
    }
    const dim3 grid(nblk(g.G), nf);
    if (prof) prof(pa, K_FILT_Y, 0);
    hipLaunchKernelGGL(k_filt_y, grid, dim3(kBlock), 0, s, cin, pt1, g, w);
    if (prof) prof(pa, K_FILT_Y, 1);
    if (prof) prof(pa, K_FILT_Z, 0);
    hipLaunchKernelGGL(k_filt_z, grid, dim3(kBlock), 0, s, ct1, pt2, g, w);
    if (prof) prof(pa, K_FILT_Z, 1);
    if (prof) prof(pa, K_FILT_X, 0);
    hipLaunchKernelGGL(k_filt_x, grid, dim3(kBlock), 0, s, ct2, ct1, pout, g, w);
    if (prof) prof(pa, K_FILT_X, 1);
}

}  // namespace

void launch_filter(hipStream_t s, const double* in, double* out, double* t1, double* t2, const Geo& g,
                   const double h[3], void (*prof)(void*, int, int), void* pa)
{
    run_filters(s, 1, &in, &out, &t1, &t2, g, h, prof, pa);
}

namespace {

// Averaged coefficient: z-line sums of L:M and M:M per (i, j) -- one
// wavefront per line, a lane-strided sum and a fixed butterfly, so the sums
// are reproducible -- into sums[(i - 1) H + j - 1] and sums[L H + ...].
__global__ void __launch_bounds__(256) k_avg_lines(const double* lm, const double* mm, Geo g, double* sums)
{
    const long long line = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (line >= (long long)g.L * g.H) return;
    const int i = (int)(line / g.H) + 1, j = (int)(line % g.H) + 1;
    const long long c0 = cidx(g, i, j, 0);
    double a = 0., b = 0.;
    for (int k = 1 + lane; k <= g.D; k += 64) {
        a += lm[c0 + k];
        b += mm[c0 + k];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    if (lane == 0) {
        sums[line] = a;
        sums[(long long)g.L * g.H + line] = b;
    }
}

// x-z plane sums per j from the line sums, over i in order, into
// sums[2 L H + j - 1] and sums[2 L H + H + j - 1]
__global__ void k_avg_planes(Geo g, double* sums)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= g.H) return;
    const long long LH = (long long)g.L * g.H;
    double a = 0., b = 0.;
    for (int i = 0; i < g.L; i++) {
        a += sums[(long long)i * g.H + j];
        b += sums[LH + (long long)i * g.H + j];
    }
    sums[2 * LH + j] = a;
    sums[2 * LH + g.H + j] = b;
}

// Cd from the sums of the cell's line or plane, clipped to [0, 0.15]; mu
// (rho Delta^2 |S| on entry) times Cd
__global__ void __launch_bounds__(256) k_avg_apply(double* mu, float* cd, const double* sums, Geo g, int mode)
{
    const long long n = (long long)g.L * g.H * g.D;
    const long long LH = (long long)g.L * g.H;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(t % g.D) + 1;
        const long long line = t / g.D;
        const int j = (int)(line % g.H) + 1, i = (int)(line / g.H) + 1;
        const double a = mode == FSL_SGS_AVG_Z ? sums[line] : sums[2 * LH + j - 1];
        const double b = mode == FSL_SGS_AVG_Z ? sums[LH + line] : sums[2 * LH + g.H + j - 1];
        double Cd = -0.5 * (a / (b + kSmall));
        if (Cd > 0.15) Cd = 0.15;
        else if (Cd < 0.0) Cd = 0.0;
        const long long c = cidx(g, i, j, k);
        mu[c] *= Cd;
        if (cd) cd[c] = (float)Cd;
    }
}

template <class R>
void sgs_pair(hipStream_t s, CState5 U, double* mu, R* const uf[3], R* const Lij[6], R* const Aij[6], R* mag,
              const Geo& g, const Phys& ph, void (*prof)(void*, int, int), void* pa, float* cd_out,
              const SgsAvg* avg)
{
    const int ntiles = ((g.L + TI - 1) / TI) * ((g.H + TJ - 1) / TJ) * ((g.D + TK - 1) / TK);
    SgsA<R> A{};
    A.U = U;
    for (int r = 0; r < 3; r++) A.uf[r] = uf[r];
    for (int m = 0; m < 6; m++) {
        A.Lij[m] = Lij[m];
        A.Aij[m] = Aij[m];
    }
    A.mag = mag;
    if (prof) prof(pa, K_SGS_PRIM, 0);
    hipLaunchKernelGGL(k_sgs_a<R>, dim3(ntiles), dim3(kBlock), 0, s, A, g, ph);
    if (prof) prof(pa, K_SGS_PRIM, 1);
    SgsB<R> Bk{};
    Bk.rho = U.f[0];
    for (int r = 0; r < 3; r++) Bk.uf[r] = uf[r];
    for (int m = 0; m < 6; m++) {
        Bk.Lij[m] = Lij[m];
        Bk.Aij[m] = Aij[m];
    }
    Bk.mag = mag;
    Bk.mu = mu;
    Bk.cd = cd_out;
    const bool av = avg && avg->mode != FSL_SGS_AVG_NONE;
    Bk.lm = av ? avg->lm : nullptr;
    Bk.mm = av ? avg->mm : nullptr;
    if (prof) prof(pa, K_SGS_FINAL, 0);
    hipLaunchKernelGGL(k_sgs_b<R>, dim3(ntiles), dim3(kBlock), 0, s, Bk, g, ph);
    if (av) launch_sgs_average(s, *avg, mu, cd_out, g);
    if (prof) prof(pa, K_SGS_FINAL, 1);
}

}  // namespace

size_t sgs_avg_bytes(const Geo& g)
{
    return (2 * (size_t)g.G + 2 * (size_t)g.L * g.H + 2 * (size_t)g.H) * sizeof(double);
}

void launch_sgs_average(hipStream_t s, const SgsAvg& av, double* mu, float* cd_out, const Geo& g)
{
    const long long LH = (long long)g.L * g.H;
    hipLaunchKernelGGL(k_avg_lines, dim3((unsigned)((LH + 3) / 4)), dim3(256), 0, s, av.lm, av.mm, g, av.sums);
    if (av.mode == FSL_SGS_AVG_XZ)
        hipLaunchKernelGGL(k_avg_planes, dim3((unsigned)((g.H + 63) / 64)), dim3(64), 0, s, g, av.sums);
    const long long n = LH * g.D;
    const unsigned nb = (unsigned)std::min<long long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_avg_apply, dim3(nb), dim3(256), 0, s, mu, cd_out, av.sums, g, av.mode);
}

void launch_sgs(hipStream_t s, CState5 U, double* mu, const Geo& g, const Phys& ph, void (*prof)(void*, int, int),
                void* pa, float* cd_out, const SgsAvg* avg)
{
    const bool av = avg && avg->mode != FSL_SGS_AVG_NONE;
    if (prof) prof(pa, K_SGS_FINAL, 0);
    launch_sgs_march(s, U, mu, g, ph, cd_out, av ? avg->lm : nullptr, av ? avg->mm : nullptr);
    if (av) launch_sgs_average(s, *avg, mu, cd_out, g);
    if (prof) prof(pa, K_SGS_FINAL, 1);
}

size_t sgs_f32_scratch_bytes(const Geo& g) { return (size_t)(3 + 6 + 6 + 1) * (size_t)g.G * sizeof(float); }

void launch_sgs_f32(hipStream_t s, CState5 U, double* mu, void* scratch, const Geo& g, const Phys& ph,
                    void (*prof)(void*, int, int), void* pa, float* cd_out, const SgsAvg* avg)
{
    // 16 float cell arrays: u^ (3), Leonard stress (6), A (6), |S|
    float* p = (float*)scratch;
    float *uf[3], *Lij[6], *Aij[6];
    for (int r = 0; r < 3; r++) { uf[r] = p; p += g.G; }
    for (int m = 0; m < 6; m++) { Lij[m] = p; p += g.G; }
    for (int m = 0; m < 6; m++) { Aij[m] = p; p += g.G; }
    sgs_pair<float>(s, U, mu, uf, Lij, Aij, p, g, ph, prof, pa, cd_out, avg);
}

}  // namespace fsl

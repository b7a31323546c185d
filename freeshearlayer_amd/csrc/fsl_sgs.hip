// fsl_sgs.hip -- dynamic Smagorinsky model (Germano-Lilly, pointwise Cd
// clipped to [0, 0.15]), src/turbulence.c:17-507.
//
// The reference filters 21 fields per call.  The velocity products and the
// A tensor are symmetric and their filtered copies are bit-identical
// (u v == v u in IEEE arithmetic), so 15 filters give the same result.  The
// model reads the step-start state only (src/fluxes.c:53 passes U1..U4,
// whose ghost layer is filled at stage 1 and unchanged until the end of
// stage 3), so the context evaluates it once per step.
//
// Filter semantics reproduced exactly (src/turbulence.c:47-95): three
// separable passes over index boxes that leave zeros outside them:
//   y pass  t1 on i in [0,L], j in [1,H], k in [0,D]
//   z pass  t2 on i in [0,L], j in [0,H], k in [1,D]
//   x pass  out on i in [1,L], j in [0,H], k in [0,D]; elsewhere out = t1
// with each sum accumulated as ((0 + a h) + b h) + c h.
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

__device__ __forceinline__ void decode(const Geo& g, long long t, int& i, int& j, int& k)
{
    k = (int)(t % (g.D + 2));
    j = (int)((t / (g.D + 2)) % (g.H + 2));
    i = (int)(t / g.si);
}

// src/turbulence.c:196-230: primitives and the six distinct products on the
// whole ghosted box
__global__ void __launch_bounds__(kBlock) k_sgs_prims(CState5 U, Ptr9 out, long long G)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= G) return;
    const double rr = 1. / U.f[0][t];
    const double u = U.f[1][t] * rr, v = U.f[2][t] * rr, w = U.f[3][t] * rr;
    out.p[0][t] = u;
    out.p[1][t] = v;
    out.p[2][t] = w;
    out.p[3][t] = u * u;
    out.p[4][t] = u * v;
    out.p[5][t] = u * w;
    out.p[6][t] = v * v;
    out.p[7][t] = v * w;
    out.p[8][t] = w * w;
}

struct H3 {
    double h[3];
};

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

// src/turbulence.c:250-382: Leonard stress, strain rates of the resolved and
// the filtered velocity, A = |S| S and B = |S^| S^ (interior cells)
struct StrainArgs {
    const double* prim[3];   // u, v, w
    const double* filt[9];   // u^, v^, w^, (uu)^, (uv)^, (uw)^, (vv)^, (vw)^, (ww)^
    double* Lij[6];
    double* Bij[6];
    double* Aij[6];
    double* mag;
};

__device__ __forceinline__ void strain9(const double* const vel[3], long long c, const Geo& g,
                                        const Phys& ph, double S[9])
{
    double gr[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        gr[r][0] = (vel[r][c + g.si] - vel[r][c - g.si]) * ph.r2del[0];
        gr[r][1] = (vel[r][c + g.sj] - vel[r][c - g.sj]) * ph.r2del[1];
        gr[r][2] = (vel[r][c + 1] - vel[r][c - 1]) * ph.r2del[2];
    }
    S[0] = gr[0][0];
    S[1] = 0.5 * (gr[0][1] + gr[1][0]);
    S[2] = 0.5 * (gr[0][2] + gr[2][0]);
    S[3] = S[1];
    S[4] = gr[1][1];
    S[5] = 0.5 * (gr[1][2] + gr[2][1]);
    S[6] = S[2];
    S[7] = S[5];
    S[8] = gr[2][2];
}

__device__ __forceinline__ double mag9(const double S[9])
{
    return sqrt(2 * (S[0] * S[0] + S[1] * S[1] + S[2] * S[2] + S[3] * S[3] + S[4] * S[4] + S[5] * S[5] +
                     S[6] * S[6] + S[7] * S[7] + S[8] * S[8]));
}

__global__ void __launch_bounds__(kBlock) k_sgs_strain(StrainArgs a, Geo g, Phys ph)
{
    const long long n = (long long)g.L * g.H * g.D;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int k = (int)(t % g.D) + 1;
    const int j = (int)((t / g.D) % g.H) + 1;
    const int i = (int)(t / ((long long)g.D * g.H)) + 1;
    const long long c = cidx(g, i, j, k);
    const double uf = a.filt[0][c], vf = a.filt[1][c], wf = a.filt[2][c];
    a.Lij[0][c] = a.filt[3][c] - uf * uf;
    a.Lij[1][c] = a.filt[4][c] - uf * vf;
    a.Lij[2][c] = a.filt[5][c] - uf * wf;
    a.Lij[3][c] = a.filt[6][c] - vf * vf;
    a.Lij[4][c] = a.filt[7][c] - vf * wf;
    a.Lij[5][c] = a.filt[8][c] - wf * wf;
    double S[9], Sf[9];
    const double* vel[3] = {a.prim[0], a.prim[1], a.prim[2]};
    const double* velf[3] = {a.filt[0], a.filt[1], a.filt[2]};
    strain9(vel, c, g, ph, S);
    const double Sm = mag9(S);
    a.mag[c] = Sm;
    strain9(velf, c, g, ph, Sf);
    const double Sfm = mag9(Sf);
    const int uniq[6] = {0, 1, 2, 4, 5, 8};   // 11, 12, 13, 22, 23, 33
#pragma unroll
    for (int q = 0; q < 6; q++) {
        a.Bij[q][c] = Sfm * Sf[uniq[q]];
        a.Aij[q][c] = Sm * S[uniq[q]];
    }
}

// src/turbulence.c:398-445: M = Delta^2 (4 B - A^), deviatoric L, Cd, mu_SGS
struct FinalArgs {
    const double* rho;
    const double* Lij[6];
    const double* Bij[6];
    const double* Af[6];
    const double* mag;
    double* mu;
};

__global__ void __launch_bounds__(kBlock) k_sgs_final(FinalArgs a, Geo g, Phys ph)
{
    const long long n = (long long)g.L * g.H * g.D;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int k = (int)(t % g.D) + 1;
    const int j = (int)((t / g.D) % g.H) + 1;
    const int i = (int)(t / ((long long)g.D * g.H)) + 1;
    const long long c = cidx(g, i, j, k);
    // 3x3 from the six distinct components: 11 12 13 / 21=12 22 23 / 31=13 32=23 33
    const int map9[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    double M[9], Lm[9];
#pragma unroll
    for (int q = 0; q < 9; q++) {
        M[q] = ph.DD * (4 * a.Bij[map9[q]][c] - a.Af[map9[q]][c]);
        Lm[q] = a.Lij[map9[q]][c];
    }
    const double tr = -(1.0 / 3.0) * (Lm[0] + Lm[4] + Lm[8]);
    Lm[0] += tr;
    Lm[4] += tr;
    Lm[8] += tr;
    const double LLMM = (Lm[0] * M[0] + Lm[1] * M[1] + Lm[2] * M[2] + Lm[3] * M[3] + Lm[4] * M[4] +
                         Lm[5] * M[5] + Lm[6] * M[6] + Lm[7] * M[7] + Lm[8] * M[8]);
    const double MMMM = (M[0] * M[0] + M[1] * M[1] + M[2] * M[2] + M[3] * M[3] + M[4] * M[4] +
                         M[5] * M[5] + M[6] * M[6] + M[7] * M[7] + M[8] * M[8]);
    double Cd = -0.5 * (LLMM / (MMMM + kSmall));
    if (Cd > 0.15) Cd = 0.15;
    else if (Cd < 0.0) Cd = 0.0;
    a.mu[c] = a.rho[c] * Cd * ph.DD * a.mag[c];
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

size_t sgs_scratch_arrays() { return 9 * 4 + 6 * 4 + 1; }

void launch_filter(hipStream_t s, const double* in, double* out, double* t1, double* t2, const Geo& g,
                   const double h[3], void (*prof)(void*, int, int), void* pa)
{
    run_filters(s, 1, &in, &out, &t1, &t2, g, h, prof, pa);
}

void launch_sgs(hipStream_t s, CState5 U, double* mu, const SgsBufs& B, const Geo& g, const Phys& ph,
                void (*prof)(void*, int, int), void* pa)
{
    // src/turbulence.c:113: box kernel, 1./3. stored as real
    const double box[3] = {1. / 3., 1. / 3., 1. / 3.};
    Ptr9 prim{};
    for (int a = 0; a < 9; a++) prim.p[a] = B.prim[a];
    if (prof) prof(pa, K_SGS_PRIM, 0);
    hipLaunchKernelGGL(k_sgs_prims, dim3(nblk(g.G)), dim3(kBlock), 0, s, U, prim, g.G);
    if (prof) prof(pa, K_SGS_PRIM, 1);

    const double* in9[9];
    for (int a = 0; a < 9; a++) in9[a] = B.prim[a];
    run_filters(s, 9, in9, B.filt, B.t1, B.t2, g, box, prof, pa);

    StrainArgs sa{};
    for (int r = 0; r < 3; r++) sa.prim[r] = B.prim[r];
    for (int a = 0; a < 9; a++) sa.filt[a] = B.filt[a];
    for (int q = 0; q < 6; q++) {
        sa.Lij[q] = B.Lij[q];
        sa.Bij[q] = B.Bij[q];
        sa.Aij[q] = B.Aij[q];
    }
    sa.mag = B.mag;
    const long long n = (long long)g.L * g.H * g.D;
    if (prof) prof(pa, K_SGS_STRAIN, 0);
    hipLaunchKernelGGL(k_sgs_strain, dim3(nblk(n)), dim3(kBlock), 0, s, sa, g, ph);
    if (prof) prof(pa, K_SGS_STRAIN, 1);

    const double* a6[6];
    for (int q = 0; q < 6; q++) a6[q] = B.Aij[q];
    run_filters(s, 6, a6, B.Af, B.t1, B.t2, g, box, prof, pa);

    FinalArgs fa{};
    fa.rho = U.f[0];
    for (int q = 0; q < 6; q++) {
        fa.Lij[q] = B.Lij[q];
        fa.Bij[q] = B.Bij[q];
        fa.Af[q] = B.Af[q];
    }
    fa.mag = B.mag;
    fa.mu = mu;
    if (prof) prof(pa, K_SGS_FINAL, 0);
    hipLaunchKernelGGL(k_sgs_final, dim3(nblk(n)), dim3(kBlock), 0, s, fa, g, ph);
    if (prof) prof(pa, K_SGS_FINAL, 1);
}

}  // namespace fsl

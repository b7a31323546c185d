// fsl_sgs_march.hip -- the FP64 dynamic Smagorinsky model (src/turbulence.c:
// 17-507) in one kernel that marches along x, with no intermediate arrays in
// HBM.
//
// A workgroup owns a 16 x 16 tile of (j, k) columns and walks a segment of x
// planes.  Per loaded plane P it advances a pipeline whose stages lag by one
// plane each:
//   plane P     u = (rho u)(1/rho) on the tile and a two-cell ring; the y and z
//               passes of the box filter for u, v, w (on the tile and a
//               one-cell ring) and for the six velocity products (tile only)
//   plane P-1   the x passes close: u^ (tile and ring) and the filtered
//               products (tile), so the Leonard stress L = (u_i u_j)^ - u^_i u^_j;
//               the strain S, |S| and A = |S| S on the tile and ring, and the
//               y and z passes of the six components of A
//   plane P-2   the x pass of A closes: A^; S^ from u^ of planes P-3..P-1,
//               B = |S^| S^, M = Delta^2 (4 B - A^), Cd = -(L:M) / 2 (M:M + small)
//               clipped to [0, 0.15], mu_SGS = rho Cd Delta^2 |S|
// The x pass of a field is kept as running sums: its three terms arrive in
// three consecutive planes.  Values needed by other columns (u, the y-pass
// results, u^, A) go through LDS; per-column and per-point ones (the x sums
// of u^, the products and A, L, |S|) stay in registers.  The six products
// and the six components of A each take one y pass and one z pass (six
// workgroup barriers per plane).
//
// Every value is formed by the operations, in the order, of the reference
// with real = double (each filter sum ((0 + a h) + b h) + c h; FP contraction
// is off in this build), and the truncated filter boxes are the reference's
// (src/turbulence.c:47-95): y pass on i in [0, L], j in [1, H], k in [0, D];
// z pass on i in [0, L], j in [0, H], k in [1, D]; x pass on i in [1, L],
// j in [0, H], k in [0, D], elsewhere the y-pass value (the i = 0 plane) or
// 0.  A and the Leonard stress exist on interior cells only; mu_SGS is
// written on interior cells only.  So mu_SGS is bit-identical to the
// reference's.
#include <hip/hip_runtime.h>

#include "fsl_internal.hpp"

namespace fsl {

namespace {

constexpr int kThreads = 256;
constexpr int MJ = 16, MK = 16;                 // tile columns (j, k)
constexpr int RJ2 = MJ + 4, RK2 = MK + 4;       // tile + two-cell ring
constexpr int RJ1 = MJ + 2, RK1 = MK + 2;       // tile + one-cell ring
constexpr int N2 = RJ2 * RK2;                   // 400
constexpr int N1 = RJ1 * RK1;                   // 324
constexpr int NY1 = RJ1 * RK2;                  // y-pass of u^: rows of the one-cell ring, columns of the two-cell ring
constexpr int NYT = MJ * RK1;                   // y-pass on tile rows, one-cell-ring columns
static_assert(MJ * MK == kThreads, "one tile column per thread");
static_assert(3 * NY1 <= 6 * NYT && NYT <= NY1, "y-pass scratch");
static_assert(N1 <= 2 * kThreads, "at most two ring points per thread");

struct MarchArgs {
    CState5 U;
    double* mu;
    float* cd;
    double* lm;   // averaged model: L:M and M:M per cell (mu then holds rho Delta^2 |S|), or null
    double* mm;
    Geo g;
    Phys ph;
    int segL, ntj, ntk, nblocks, xcd_remap;
};

__device__ __forceinline__ int mod3(int p) { return ((p % 3) + 3) % 3; }

// the six distinct components, in the order of the Leonard-stress and A arrays
// (11, 12, 13, 22, 23, 33), as products of velocity components
__device__ __forceinline__ int pa(int m) { return m < 3 ? 0 : (m < 5 ? 1 : 2); }
__device__ __forceinline__ int pb(int m) { return m < 3 ? m : (m < 5 ? m - 2 : 2); }

__global__ void __launch_bounds__(kThreads, 2) k_sgs_march(MarchArgs a)
{
    __shared__ double s_u[3][3][N2];      // u, v, w of planes P-2, P-1, P (slot P mod 3)
    __shared__ double s_t1[6 * NYT];      // y-pass results: u^ (3 x NY1) or the products / A (6 x NYT)
    __shared__ double s_hu[2][3][N1];     // completed u^ of planes P-1, P-2 (slot by plane parity)
    __shared__ double s_A[6][N1];         // the six components of A on the tile and ring
    __shared__ double s_mag[MJ * MK];     // |S| of the tile's cells

    const Geo& g = a.g;
    const int L = g.L, H = g.H, D = g.D;
    int bid = blockIdx.x;
    if (a.xcd_remap) {
        const int per = gridDim.x >> 3;
        bid = (bid & 7) * per + (bid >> 3);
    }
    if (bid >= a.nblocks) return;
    const int tk = bid % a.ntk;
    const int rem = bid / a.ntk;
    const int tj = rem % a.ntj;
    const int seg = rem / a.ntj;
    const int j0 = 1 + tj * MJ, k0 = 1 + tk * MK;
    const int i0 = 1 + seg * a.segL;
    const int i1 = min(i0 + a.segL, L + 1);   // planes [i0, i1) get mu_SGS
    const int tid = threadIdx.x;
    const int cj = tid / MK, ck = tid % MK;   // this thread's tile column
    const int gjc = j0 + cj, gkc = k0 + ck;
    const bool own = gjc <= H && gkc <= D;
    const double h = 1. / 3.;
    const double r2x = a.ph.r2del[0], r2y = a.ph.r2del[1], r2z = a.ph.r2del[2];

    // per-column pipeline state (registers)
    double pacc2[6], pacc1[6];   // product x sums: 2 terms (plane P-1), 1 term (plane P)
    double aacc2[6], aacc1[6];   // A x sums: 2 terms (plane P-2), 1 term (plane P-1)
    double Lr[6];                // Leonard stress of plane P-2 (completed last iteration)
    double mag_prev = 0.;        // |S| of plane P-2
    double hu3[3];               // u^ of plane P-3 at this column
    // running x sums of u^ at this thread's ring points e = tid, tid + 256:
    // hacc2 of plane P-1 (two terms), hacc1 of plane P (one term)
    double hacc2[2][3], hacc1[2][3];
#pragma unroll
    for (int m = 0; m < 6; m++) { pacc2[m] = pacc1[m] = aacc2[m] = aacc1[m] = Lr[m] = 0.; }
#pragma unroll
    for (int it = 0; it < 2; it++)
#pragma unroll
        for (int f = 0; f < 3; f++) hacc2[it][f] = hacc1[it][f] = 0.;
    hu3[0] = hu3[1] = hu3[2] = 0.;
    for (int e = tid; e < 3 * 3 * N2; e += kThreads) (&s_u[0][0][0])[e] = 0.;
    for (int e = tid; e < 2 * 3 * N1; e += kThreads) (&s_hu[0][0][0])[e] = 0.;

    // The state of the next plane is loaded one iteration ahead (registers:
    // this thread's cells e = tid and tid + 256 of the two-cell ring).  The
    // workgroup barriers order LDS only (lds_barrier), so these loads stay in
    // flight across them.
    constexpr int NPF = (N2 + kThreads - 1) / kThreads;
    double pf[NPF][4];
    bool pfv[NPF];
    long long pfc[NPF];   // in-plane cell offsets of this thread's ring cells (-1: outside the ghosted box)
#pragma unroll
    for (int it = 0; it < NPF; it++) {
        const int e = tid + it * kThreads;
        const int gj = j0 - 2 + e / RK2, gk = k0 - 2 + e % RK2;
        pfc[it] = (e < N2 && gj >= 0 && gj <= H + 1 && gk >= 0 && gk <= D + 1) ? (long long)gj * g.sj + gk : -1;
    }
    auto prefetch = [&](int P) {
#pragma unroll
        for (int it = 0; it < NPF; it++) {
            pfv[it] = pfc[it] >= 0 && P >= 0 && P <= L + 1;
            const long long c = (long long)(pfv[it] ? P : 0) * g.si + (pfc[it] >= 0 ? pfc[it] : 0);
#pragma unroll
            for (int f = 0; f < 4; f++) pf[it][f] = pfv[it] ? a.U.f[f][c] : 1.;
        }
    };
    prefetch(i0 - 2);
    // rho of this column two planes ahead of its use (mu_SGS of plane P-2)
    const long long ocol = (long long)min(gjc, H + 1) * g.sj + min(gkc, D + 1);
    double rho1 = 0., rho2 = 0.;   // planes P-1, P-2
    lds_barrier();   // the zeroed LDS arrays

    for (int P = i0 - 2; P <= i1 + 1; P++) {
        // No barrier here: the u slot rewritten below (plane P - 3) was last
        // read by the previous plane's strain, before its last two barriers;
        // every other array written before this plane's second barrier was
        // last read before the previous plane's last barrier.
        const int s0 = mod3(P), s1 = mod3(P - 1), s2 = mod3(P - 2);
        // ---- plane P: u on the tile and two-cell ring (0 outside the ghosted box)
#pragma unroll
        for (int it = 0; it < NPF; it++) {
            const int e = tid + it * kThreads;
            if (e >= N2) break;
            double u = 0., v = 0., w = 0.;
            if (pfv[it]) {
                const double rr = 1. / pf[it][0];   // src/turbulence.c:197-209
                u = pf[it][1] * rr;
                v = pf[it][2] * rr;
                w = pf[it][3] * rr;
            }
            s_u[s0][0][e] = u;
            s_u[s0][1][e] = v;
            s_u[s0][2][e] = w;
        }
        prefetch(P + 1);
        const double rho0 = (P >= 0 && P <= L + 1) ? a.U.f[0][(long long)P * g.si + ocol] : 0.;
        lds_barrier();
        const bool pv = P >= 0 && P <= L;   // plane of the y and z passes' boxes
        // ---- plane P: u^ -- y pass (rows of the one-cell ring, columns of the two-cell ring)
        for (int e = tid; e < NY1; e += kThreads) {
            const int rj = e / RK2, rk = e % RK2;
            const int gj = j0 - 1 + rj, gk = k0 - 2 + rk;
            const bool ok = pv && gj >= 1 && gj <= H && gk >= 0 && gk <= D;
            const int b = (rj + 1) * RK2 + rk;   // (gj, gk) in the two-cell ring
#pragma unroll
            for (int f = 0; f < 3; f++) {
                double r = 0.;
                if (ok) {
                    double t = 0.;
                    t += s_u[s0][f][b - RK2] * h;
                    t += s_u[s0][f][b] * h;
                    t += s_u[s0][f][b + RK2] * h;
                    r = t;
                }
                s_t1[f * NY1 + e] = r;
            }
        }
        lds_barrier();
        // z pass on the tile and one-cell ring, then the x sums; the sums of
        // plane P-1 close into u^ (P-1).  u^ of the i = 0 plane is its y-pass
        // value (the x pass starts at i = 1): kept in the running sum of plane
        // 0 at P = 0 and taken as it is at P = 1.
        {
            const int hs = (P - 1) & 1;                 // slot of u^ (P-1)
#pragma unroll
            for (int it = 0; it < 2; it++) {
                const int e = tid + it * kThreads;
                if (e >= N1) break;
                const int rj = e / RK1, rk = e % RK1;
                const int gj = j0 - 1 + rj, gk = k0 - 1 + rk;
                const bool ok = pv && gj >= 0 && gj <= H && gk >= 1 && gk <= D;
                const int b = rj * RK2 + rk + 1;        // (gj, gk) in the y-pass array
#pragma unroll
                for (int f = 0; f < 3; f++) {
                    const double* t1f = s_t1 + f * NY1;
                    double t2 = 0.;
                    if (ok) {
                        double t = 0.;
                        t += t1f[b - 1] * h;
                        t += t1f[b] * h;
                        t += t1f[b + 1] * h;
                        t2 = t;
                    }
                    const double th = t2 * h;
                    const double acc2 = hacc2[it][f], acc1 = hacc1[it][f];
                    double done = acc2 + th;
                    if (P - 1 == 0) done = acc2;              // u^ (0) = y-pass value
                    else if (P - 1 > L || P - 1 < 0) done = 0.;
                    s_hu[hs][f][e] = done;
                    if (P == 0) {
                        const bool oky = gj >= 1 && gj <= H && gk >= 0 && gk <= D;
                        hacc2[it][f] = oky ? t1f[b] : 0.;     // u^ (0), closed at P = 1
                    } else {
                        hacc2[it][f] = acc1 + th;
                    }
                    double t1 = 0.;
                    t1 += th;
                    hacc1[it][f] = t1;                        // 0 + t2 h: first term of plane P+1
                }
            }
        }
        // ---- plane P: the six products -- y pass on the tile rows, then z pass
        // and x sums at this column
        double fil[6];   // filtered products of plane P-1 at this column
        lds_barrier();   // s_t1 free; u^ (P-1) complete
        for (int e = tid; e < NYT; e += kThreads) {
            const int rj = e / RK1, rk = e % RK1;
            const int gj = j0 + rj, gk = k0 - 1 + rk;
            const bool ok = pv && gj >= 1 && gj <= H && gk >= 0 && gk <= D;
            const int b = (rj + 2) * RK2 + rk + 1;   // (gj, gk) in the two-cell ring
#pragma unroll
            for (int m = 0; m < 6; m++) {
                double r = 0.;
                if (ok) {
                    const double* ua = s_u[s0][pa(m)];
                    const double* ub = s_u[s0][pb(m)];
                    double t = 0.;
                    t += ua[b - RK2] * ub[b - RK2] * h;
                    t += ua[b] * ub[b] * h;
                    t += ua[b + RK2] * ub[b + RK2] * h;
                    r = t;
                }
                s_t1[m * NYT + e] = r;
            }
        }
        lds_barrier();
        {
            const bool ok = pv && gjc <= H && gkc >= 1 && gkc <= D;
            const int b = cj * RK1 + ck + 1;
#pragma unroll
            for (int m = 0; m < 6; m++) {
                const double* t1m = s_t1 + m * NYT;
                double t2 = 0.;
                if (ok) {
                    double t = 0.;
                    t += t1m[b - 1] * h;
                    t += t1m[b] * h;
                    t += t1m[b + 1] * h;
                    t2 = t;
                }
                const double th = t2 * h;
                fil[m] = pacc2[m] + th;
                pacc2[m] = pacc1[m] + th;
                double t1 = 0.;
                t1 += th;
                pacc1[m] = t1;
            }
        }
        // Leonard stress of plane P-1 (interior cells): (u_i u_j)^ - u^_i u^_j
        double Lnew[6];
        const int eo = (cj + 1) * RK1 + ck + 1;   // this column in the one-cell ring
        const int hs1 = (P - 1) & 1, hs2 = P & 1; // u^ slots of planes P-1 and P-2
        {
            const double uf[3] = {s_hu[hs1][0][eo], s_hu[hs1][1][eo], s_hu[hs1][2][eo]};
#pragma unroll
            for (int m = 0; m < 6; m++) Lnew[m] = fil[m] - uf[pa(m)] * uf[pb(m)];
        }
        // ---- plane P-1: strain, |S| and A on the tile and ring; y and z passes
        // of A, x sums at this column.  (s_A was last read before this plane's
        // first barrier; s_t1, read by the products' z pass above, is
        // rewritten after the next barrier.)
        const int PA = P - 1;
        const bool pva = PA >= 0 && PA <= L;
        double Af[6];         // A^ of plane P-2 at this column
#pragma unroll
        for (int it = 0; it < 2; it++) {
            const int e = tid + it * kThreads;
            if (e >= N1) break;
            const int rj = e / RK1, rk = e % RK1;
            const int gj = j0 - 1 + rj, gk = k0 - 1 + rk;
            double A6[6] = {0., 0., 0., 0., 0., 0.};
            double Sm = 0.;
            if (PA >= 1 && PA <= L && gj >= 1 && gj <= H && gk >= 1 && gk <= D) {
                const int b = (rj + 1) * RK2 + rk + 1;   // (gj, gk) in the two-cell ring
                double gr[3][3];
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    gr[r][0] = (s_u[s0][r][b] - s_u[s2][r][b]) * r2x;
                    gr[r][1] = (s_u[s1][r][b + RK2] - s_u[s1][r][b - RK2]) * r2y;
                    gr[r][2] = (s_u[s1][r][b + 1] - s_u[s1][r][b - 1]) * r2z;
                }
                double S[9];
                S[0] = gr[0][0];
                S[1] = 0.5 * (gr[0][1] + gr[1][0]);
                S[2] = 0.5 * (gr[0][2] + gr[2][0]);
                S[3] = S[1];
                S[4] = gr[1][1];
                S[5] = 0.5 * (gr[1][2] + gr[2][1]);
                S[6] = S[2];
                S[7] = S[5];
                S[8] = gr[2][2];
                Sm = sqrt(2. * (S[0] * S[0] + S[1] * S[1] + S[2] * S[2] + S[3] * S[3] + S[4] * S[4] +
                                S[5] * S[5] + S[6] * S[6] + S[7] * S[7] + S[8] * S[8]));
                const int uniq[6] = {0, 1, 2, 4, 5, 8};
#pragma unroll
                for (int m = 0; m < 6; m++) A6[m] = Sm * S[uniq[m]];
            }
            if (rj >= 1 && rj <= MJ && rk >= 1 && rk <= MK) s_mag[(rj - 1) * MK + rk - 1] = Sm;
#pragma unroll
            for (int m = 0; m < 6; m++) s_A[m][e] = A6[m];
        }
        lds_barrier();
        for (int e = tid; e < NYT; e += kThreads) {
            const int rj = e / RK1, rk = e % RK1;
            const int gj = j0 + rj, gk = k0 - 1 + rk;
            const bool ok = pva && gj >= 1 && gj <= H && gk >= 0 && gk <= D;
            const int b = (rj + 1) * RK1 + rk;   // (gj, gk) in the one-cell ring
#pragma unroll
            for (int m = 0; m < 6; m++) {
                double r = 0.;
                if (ok) {
                    double t = 0.;
                    t += s_A[m][b - RK1] * h;
                    t += s_A[m][b] * h;
                    t += s_A[m][b + RK1] * h;
                    r = t;
                }
                s_t1[m * NYT + e] = r;
            }
        }
        lds_barrier();
        {
            const bool ok = pva && gjc <= H && gkc >= 1 && gkc <= D;
            const int b = cj * RK1 + ck + 1;
#pragma unroll
            for (int m = 0; m < 6; m++) {
                const double* t1m = s_t1 + m * NYT;
                double t2 = 0.;
                if (ok) {
                    double t = 0.;
                    t += t1m[b - 1] * h;
                    t += t1m[b] * h;
                    t += t1m[b + 1] * h;
                    t2 = t;
                }
                const double th = t2 * h;
                Af[m] = aacc2[m] + th;
                aacc2[m] = aacc1[m] + th;
                double t1 = 0.;
                t1 += th;
                aacc1[m] = t1;
            }
        }
        const double mag_new = s_mag[tid];   // |S| of plane P-1 at this column (0 off the interior)
        // ---- plane P-2: S^, M, Cd and mu_SGS
        const int PM = P - 2;
        const double hu1[3] = {s_hu[hs1][0][eo], s_hu[hs1][1][eo], s_hu[hs1][2][eo]};   // u^ (P-1)
        if (PM >= i0 && PM < i1 && own) {
            double gr[3][3];
#pragma unroll
            for (int r = 0; r < 3; r++) {
                gr[r][0] = (hu1[r] - hu3[r]) * r2x;
                gr[r][1] = (s_hu[hs2][r][eo + RK1] - s_hu[hs2][r][eo - RK1]) * r2y;
                gr[r][2] = (s_hu[hs2][r][eo + 1] - s_hu[hs2][r][eo - 1]) * r2z;
            }
            double Sf[9];
            Sf[0] = gr[0][0];
            Sf[1] = 0.5 * (gr[0][1] + gr[1][0]);
            Sf[2] = 0.5 * (gr[0][2] + gr[2][0]);
            Sf[3] = Sf[1];
            Sf[4] = gr[1][1];
            Sf[5] = 0.5 * (gr[1][2] + gr[2][1]);
            Sf[6] = Sf[2];
            Sf[7] = Sf[5];
            Sf[8] = gr[2][2];
            const double Sfm = sqrt(2. * (Sf[0] * Sf[0] + Sf[1] * Sf[1] + Sf[2] * Sf[2] + Sf[3] * Sf[3] + Sf[4] * Sf[4] +
                                          Sf[5] * Sf[5] + Sf[6] * Sf[6] + Sf[7] * Sf[7] + Sf[8] * Sf[8]));
            // 3 x 3 from the six distinct components: 11 12 13 / 21=12 22 23 / 31=13 32=23 33
            const int map9[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
            const int uniq[6] = {0, 1, 2, 4, 5, 8};
            double B6[6], M[9], Lm[9];
#pragma unroll
            for (int m = 0; m < 6; m++) B6[m] = Sfm * Sf[uniq[m]];
#pragma unroll
            for (int r = 0; r < 9; r++) {
                M[r] = a.ph.DD * (4. * B6[map9[r]] - Af[map9[r]]);
                Lm[r] = Lr[map9[r]];
            }
            const double tr = -(1.0 / 3.0) * (Lm[0] + Lm[4] + Lm[8]);
            Lm[0] += tr;
            Lm[4] += tr;
            Lm[8] += tr;
            const double LLMM = (Lm[0] * M[0] + Lm[1] * M[1] + Lm[2] * M[2] + Lm[3] * M[3] + Lm[4] * M[4] +
                                 Lm[5] * M[5] + Lm[6] * M[6] + Lm[7] * M[7] + Lm[8] * M[8]);
            const double MMMM = (M[0] * M[0] + M[1] * M[1] + M[2] * M[2] + M[3] * M[3] + M[4] * M[4] +
                                 M[5] * M[5] + M[6] * M[6] + M[7] * M[7] + M[8] * M[8]);
            const long long c = cidx(g, PM, gjc, gkc);
            if (a.lm) {   // averaged model: the terms, and the factor of the averaged Cd
                a.lm[c] = LLMM;
                a.mm[c] = MMMM;
                a.mu[c] = rho2 * a.ph.DD * mag_prev;
            } else {
                double Cd = -0.5 * (LLMM / (MMMM + kSmall));
                if (Cd > 0.15) Cd = 0.15;
                else if (Cd < 0.0) Cd = 0.0;
                a.mu[c] = rho2 * Cd * a.ph.DD * mag_prev;
                if (a.cd) a.cd[c] = (float)Cd;
            }
        }
        // rotate the per-column pipeline
#pragma unroll
        for (int r = 0; r < 3; r++) hu3[r] = s_hu[hs2][r][eo];   // u^ (P-2) becomes P-3 of the next plane
#pragma unroll
        for (int m = 0; m < 6; m++) Lr[m] = Lnew[m];
        mag_prev = mag_new;
        rho2 = rho1;
        rho1 = rho0;
    }
}

}  // namespace

void launch_sgs_march(hipStream_t s, CState5 U, double* mu, const Geo& g, const Phys& ph, float* cd_out,
                      double* lm, double* mm)
{
    MarchArgs a{};
    a.U = U;
    a.mu = mu;
    a.cd = cd_out;
    a.lm = lm;
    a.mm = mm;
    a.g = g;
    a.ph = ph;
    a.ntj = (g.H + MJ - 1) / MJ;
    a.ntk = (g.D + MK - 1) / MK;
    const int cols = a.ntj * a.ntk;
    int segL = 0;
    const int nseg = fused_segments(g.L, cols, 2, &segL);
    a.segL = segL;
    a.nblocks = nseg * cols;
    a.xcd_remap = 1;
    const dim3 grid((a.nblocks + 7) / 8 * 8);
    hipLaunchKernelGGL(k_sgs_march, grid, dim3(kThreads), 0, s, a);
}

}  // namespace fsl

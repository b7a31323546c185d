// fsl_kernels.hip -- kernels of the unfused path: one kernel per reference
// entry point, over device-resident SoA cell and face arrays.
//
//   k_ghost_fill   src/boundCondInGhostCells.c:10-96
//   k_recon        src/reconstruction.c:26-345
//   k_iface_bc     src/bounCondOnInterfaces.c:11-110
//   k_flux<DIR>    src/fluxes.c:57-223 (x), 226-394 (y), 397-571 (z)
//   k_rk<KIND>     src/evolution.c:29-245
//   k_cfl<MODE>    src/helpers.c:23-52 (and the acoustic rule)
//
// All kernels are HBM streaming stencils: one thread per cell or face, k (the
// contiguous index) mapped to consecutive lanes so every wavefront reads
// coalesced 64-element runs.
#include "fsl_internal.hpp"

#include <stdio.h>

#include <mutex>

#include "fsl.h"

namespace fsl {

namespace {

constexpr int kBlock = 256;

__host__ __device__ inline long long xf(const Geo& g, int f, int j, int k)
{
    return ((long long)f * g.H + j) * g.D + k;
}
__host__ __device__ inline long long yf(const Geo& g, int i, int f, int k)
{
    return ((long long)i * (g.H + 1) + f) * g.D + k;
}
__host__ __device__ inline long long zf(const Geo& g, int i, int j, int f)
{
    return ((long long)i * g.H + j) * (g.D + 1) + f;
}

inline unsigned nblk(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// ---------------------------------------------------------------------
// Ghost layer.  The reference fills it in four ordered sweeps (x inflow,
// x outflow, z periodic over i in [0, L+1], y slip over i in [0, L+1] and
// k in [0, D+1]), later sweeps copying what earlier ones wrote.  Composing
// the sweeps gives each ghost cell a closed form in terms of interior cells,
// so one gather kernel fills every ghost cell with no ordering between
// threads: map j (slip: 0 -> 1, H+1 -> H, flip rho v), then k (periodic:
// 0 -> D, D+1 -> 1), then i (0: inflow of row j, L+1: outflow of cell L).
__global__ void __launch_bounds__(kBlock) k_ghost_fill(State5 U, Geo g, Bnd b)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int L = g.L, H = g.H, D = g.D;
    // each section's size is below 2^32 (a mesh holds at most 4e9 cells), so
    // the divisions within a section are 32-bit
    const unsigned W = (unsigned)(D + 2);
    const long long px = (long long)(H + 2) * W, py = (long long)L * W, pz = (long long)L * H;
    int i, j, k;
    if (t < 2 * px) {
        const bool side = t >= px;
        const unsigned r = (unsigned)(side ? t - px : t), q = r / W;
        i = side ? L + 1 : 0;
        j = (int)q;
        k = (int)(r - q * W);
    } else if ((t -= 2 * px) < 2 * py) {
        const bool side = t >= py;
        const unsigned r = (unsigned)(side ? t - py : t), q = r / W;
        j = side ? H + 1 : 0;
        i = 1 + (int)q;
        k = (int)(r - q * W);
    } else if ((t -= 2 * py) < pz) {
        // z periodic ghosts of one interior row (i, j): k = 0 <- D and k = D+1 <- 1.
        // One thread per row touches the row's first and last cache lines once;
        // all ten loads are issued before the stores (one memory round trip).
        const unsigned r = (unsigned)t, q = r / (unsigned)H;
        i = 1 + (int)q;
        j = 1 + (int)(r - q * (unsigned)H);
        const long long row = cidx(g, i, j, 0);
        double lo[5], hi[5];
#pragma unroll
        for (int f = 0; f < 5; f++) {
            lo[f] = U.f[f][row + D];
            hi[f] = U.f[f][row + 1];
        }
#pragma unroll
        for (int f = 0; f < 5; f++) {
            U.f[f][row] = lo[f];
            U.f[f][row + D + 1] = hi[f];
        }
        return;
    } else {
        return;
    }
    int jm = j;
    bool flip = false;
    if (j == 0) { jm = 1; flip = true; }
    else if (j == H + 1) { jm = H; flip = true; }
    int km = k;
    if (k == 0) km = D;
    else if (k == D + 1) km = 1;
    double q[5];
    if (i == 0) {
        inflow_cell(b, jm, q);
    } else {
        const long long s = cidx(g, (i == L + 1) ? L : i, jm, km);
#pragma unroll
        for (int f = 0; f < 5; f++) q[f] = U.f[f][s];
        if (i == L + 1) outflow_state(b.p_out, q);
    }
    if (flip) q[2] = -q[2];
    const long long d = cidx(g, i, j, k);
#pragma unroll
    for (int f = 0; f < 5; f++) U.f[f][d] = q[f];
}

// ---------------------------------------------------------------------
// Reconstruction: every interior cell writes its + and - face states in all
// three directions (src/reconstruction.c:154-170, 240-256, 325-341).
__global__ void __launch_bounds__(kBlock) k_recon(CState5 U, Faces F, Geo g, int quirk)
{
    const long long n = (long long)g.L * g.H * g.D;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int k = (int)(t % g.D) + 1;
    const int j = (int)((t / g.D) % g.H) + 1;
    const int i = (int)(t / ((long long)g.D * g.H)) + 1;
    const long long c = cidx(g, i, j, k);
    double q[5], qp[5], qm[5], plus[5], minus[5];
#pragma unroll
    for (int f = 0; f < 5; f++) q[f] = U.f[f][c];
    Prim<double> P;
    cell_prims(q, P);
#pragma unroll
    for (int f = 0; f < 5; f++) { qp[f] = U.f[f][c + g.si]; qm[f] = U.f[f][c - g.si]; }
    recon_dir<1, 2, 3>(P, qp, qm, false, plus, minus);
#pragma unroll
    for (int f = 0; f < 5; f++) {
        F.a[f][xf(g, i, j - 1, k - 1)] = plus[f];
        F.a[5 + f][xf(g, i - 1, j - 1, k - 1)] = minus[f];
    }
#pragma unroll
    for (int f = 0; f < 5; f++) { qp[f] = U.f[f][c + g.sj]; qm[f] = U.f[f][c - g.sj]; }
    recon_dir<2, 3, 1>(P, qp, qm, false, plus, minus);
#pragma unroll
    for (int f = 0; f < 5; f++) {
        F.a[10 + f][yf(g, i - 1, j, k - 1)] = plus[f];
        F.a[15 + f][yf(g, i - 1, j - 1, k - 1)] = minus[f];
    }
#pragma unroll
    for (int f = 0; f < 5; f++) { qp[f] = U.f[f][c + 1]; qm[f] = U.f[f][c - 1]; }
    recon_dir<3, 1, 2>(P, qp, qm, quirk != 0, plus, minus);
#pragma unroll
    for (int f = 0; f < 5; f++) {
        F.a[20 + f][zf(g, i - 1, j - 1, k)] = plus[f];
        F.a[25 + f][zf(g, i - 1, j - 1, k - 1)] = minus[f];
    }
}

// ---------------------------------------------------------------------
// Boundary faces (src/bounCondOnInterfaces.c:11-110).  The four parts touch
// disjoint face entries and read only reconstruction output, so they run as
// one kernel.
__global__ void __launch_bounds__(kBlock) k_iface_bc(Faces F, Geo g, Bnd b)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int L = g.L, H = g.H, D = g.D;
    const long long nx = (long long)H * D, ny = (long long)L * D, nz = (long long)L * H;
    if (t < nx) {   // inflow: left state of face 0
        const int j = (int)(t / D), k = (int)(t % D);
        double q[5];
        inflow_face(b, H, j, k, q);
        const long long c = xf(g, 0, j, k);
#pragma unroll
        for (int f = 0; f < 5; f++) F.a[f][c] = q[f];
        return;
    }
    t -= nx;
    if (t < nx) {   // outflow: right state of face L from its left state
        const int j = (int)(t / D), k = (int)(t % D);
        const long long c = xf(g, L, j, k);
        double q[5];
#pragma unroll
        for (int f = 0; f < 5; f++) q[f] = F.a[f][c];
        outflow_state(b.p_out, q);
#pragma unroll
        for (int f = 0; f < 5; f++) F.a[5 + f][c] = q[f];
        return;
    }
    t -= nx;
    if (t < ny) {   // y slip: mirror with rho v negated
        const int i = (int)(t / D), k = (int)(t % D);
        const long long lo = yf(g, i, 0, k), hi = yf(g, i, H, k);
#pragma unroll
        for (int f = 0; f < 5; f++) {
            const double a = F.a[15 + f][lo], c = F.a[10 + f][hi];
            F.a[10 + f][lo] = (f == 2) ? -a : a;
            F.a[15 + f][hi] = (f == 2) ? -c : c;
        }
        return;
    }
    t -= ny;
    if (t < nz) {   // z periodic
        const int i = (int)(t / H), j = (int)(t % H);
        const long long f0 = zf(g, i, j, 0), fD = zf(g, i, j, D);
#pragma unroll
        for (int f = 0; f < 5; f++) {
            F.a[20 + f][f0] = F.a[20 + f][fD];
            F.a[25 + f][fD] = F.a[25 + f][f0];
        }
    }
}

// ---------------------------------------------------------------------
// Fluxes of one direction: Riemann solver on the face states, face velocity
// gradients from the two adjacent cells and their eight tangential
// neighbours, mu_T, viscous stress and heat flux.  The flux overwrites the
// left-state array of the face (src/def.h:6-22).
template <int DIR>
__global__ void __launch_bounds__(kBlock) k_flux(CState5 U, Faces F, const double* __restrict__ mu,
                                                 Geo g, Phys ph)
{
    const int n0 = g.L + (DIR == 0), n1 = g.H + (DIR == 1), n2 = g.D + (DIR == 2);
    const long long n = (long long)n0 * n1 * n2;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int a2 = (int)(t % n2);
    const int a1 = (int)((t / n2) % n1);
    const int a0 = (int)(t / ((long long)n2 * n1));
    // backward cell: face index along DIR, +1 along the others
    const int ci = a0 + (DIR != 0), cj = a1 + (DIR != 1), ck = a2 + (DIR != 2);
    const long long st[3] = {g.si, g.sj, 1};
    const long long cb = cidx(g, ci, cj, ck), cf = cb + st[DIR];
    const int base = DIR * 10;
    double l[5], r[5];
#pragma unroll
    for (int f = 0; f < 5; f++) { l[f] = F.a[base + f][t]; r[f] = F.a[base + 5 + f][t]; }
    double R, vel[3], P;
    riemann<DIR>(l, r, R, vel, P);
    double vf[3], vb[3], Tf, Tb;
    const double Rf = U.f[0][cf], Rb = U.f[0][cb];
    vel_div_T(Rf, U.f[1][cf], U.f[2][cf], U.f[3][cf], U.f[4][cf], vf, Tf);
    vel_div_T(Rb, U.f[1][cb], U.f[2][cb], U.f[3][cb], U.f[4][cb], vb, Tb);
    double gr[3][3];
#pragma unroll
    for (int c = 0; c < 3; c++) gr[c][DIR] = (vf[c] - vb[c]) * ph.rdel[DIR];
#pragma unroll
    for (int e = 0; e < 3; e++) {
        if (e == DIR) continue;
        const long long s = st[e];
        double fp[3], bp[3], fm[3], bm[3];
        vel_recip(U.f[0][cf + s], U.f[1][cf + s], U.f[2][cf + s], U.f[3][cf + s], fp);
        vel_recip(U.f[0][cb + s], U.f[1][cb + s], U.f[2][cb + s], U.f[3][cb + s], bp);
        vel_recip(U.f[0][cf - s], U.f[1][cf - s], U.f[2][cf - s], U.f[3][cf - s], fm);
        vel_recip(U.f[0][cb - s], U.f[1][cb - s], U.f[2][cb - s], U.f[3][cb - s], bm);
#pragma unroll
        for (int c = 0; c < 3; c++) gr[c][e] = (fp[c] + bp[c] - fm[c] - bm[c]) * ph.r4del[e];
    }
    double mu_dyn = 0.0;
    if (ph.dynamic) {
        // samples (a0+1, a1+1, a2+1) and (a0, a1+1, a2+1) of mu_SGS
        // (src/fluxes.c:191, 360, 531: y and z use the x neighbour)
        const long long ma = cidx(g, a0 + 1, a1 + 1, a2 + 1);
        mu_dyn = 0.5 * (mu[ma] + mu[ma - g.si]);
    }
    double Fl[5];
    face_flux<DIR>(ph, R, vel, P, vf, vb, Tf, Tb, Rf, Rb, gr, mu_dyn, Fl);
#pragma unroll
    for (int f = 0; f < 5; f++) F.a[base + f][t] = Fl[f];
}

// ---------------------------------------------------------------------
// Runge-Kutta stage updates (src/evolution.c:29-161 three-stage,
// 164-245 two-stage).  L = dtX (Fx[i-1] - Fx[i]) + dtY (..) + dtZ (..).
template <int KIND>
__global__ void __launch_bounds__(kBlock) k_rk(State5 U, State5 Up, Faces F, Geo g, Dt3 dt)
{
    const long long n = (long long)g.L * g.H * g.D;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int k = (int)(t % g.D) + 1;
    const int j = (int)((t / g.D) % g.H) + 1;
    const int i = (int)(t / ((long long)g.D * g.H)) + 1;
    const long long c = cidx(g, i, j, k);
    const double one_third = 1. / 3., two_thirds = 2. / 3.;
#pragma unroll
    for (int f = 0; f < 5; f++) {
        const double lx = dt.x * (F.a[f][xf(g, i - 1, j - 1, k - 1)] - F.a[f][xf(g, i, j - 1, k - 1)]);
        const double ly = dt.y * (F.a[10 + f][yf(g, i - 1, j - 1, k - 1)] - F.a[10 + f][yf(g, i - 1, j, k - 1)]);
        const double lz = dt.z * (F.a[20 + f][zf(g, i - 1, j - 1, k - 1)] - F.a[20 + f][zf(g, i - 1, j - 1, k)]);
        if (KIND == 31 || KIND == 21)
            Up.f[f][c] = U.f[f][c] + lx + ly + lz;
        else if (KIND == 32)
            Up.f[f][c] = 0.75 * U.f[f][c] + 0.25 * (Up.f[f][c] + lx + ly + lz);
        else if (KIND == 33)
            U.f[f][c] = one_third * U.f[f][c] + two_thirds * (Up.f[f][c] + lx + ly + lz);
        else
            U.f[f][c] = 0.5 * (U.f[f][c] + Up.f[f][c] + lx + ly + lz);
    }
}

// ---------------------------------------------------------------------
// CFL reduction (cfl_cell, wave_max: fsl_device.hpp)
template <int MODE>
__global__ void __launch_bounds__(kBlock) k_cfl(CState5 U, Geo g, Phys ph, double dtX,
                                                unsigned long long* out)
{
    __shared__ double red[kBlock / 64];
    const long long n = (long long)g.L * g.H * g.D;
    double m = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(t % g.D) + 1;
        const int j = (int)((t / g.D) % g.H) + 1;
        const int i = (int)(t / ((long long)g.D * g.H)) + 1;
        const long long c = cidx(g, i, j, k);
        const double R = U.f[0][c];
        double v;
        if (MODE == 0) {
            const double u = U.f[1][c] / R;
            v = u * dtX;
        } else {
            const double u = U.f[1][c] / R, vv = U.f[2][c] / R, w = U.f[3][c] / R;
            const double p = (U.f[4][c] - 0.5 * R * (u * u + vv * vv + w * w)) * kK_1;
            const double cs = sqrt(kK * p / R);
            v = (fabs(u) + cs) * ph.rdel[0] + (fabs(vv) + cs) * ph.rdel[1] + (fabs(w) + cs) * ph.rdel[2];
        }
        m = nan_max(m, v);
    }
    m = wave_max(m);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = red[0];
        for (int w = 1; w < kBlock / 64; w++) b = nan_max(b, red[w]);
        atomicMax(out, (unsigned long long)__double_as_longlong(b));
    }
}

// ---------------------------------------------------------------------
// Conservation sums: per-block partials in a fixed order, then one block
// adds the partials in index order, so the result does not depend on
// scheduling.
__global__ void __launch_bounds__(kBlock) k_sums_partial(CState5 U, Geo g, double* partial)
{
    __shared__ double red[5][kBlock];
    const long long n = (long long)g.L * g.H * g.D;
    double s[5] = {0, 0, 0, 0, 0};
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(t % g.D) + 1;
        const int j = (int)((t / g.D) % g.H) + 1;
        const int i = (int)(t / ((long long)g.D * g.H)) + 1;
        const long long c = cidx(g, i, j, k);
#pragma unroll
        for (int f = 0; f < 5; f++) s[f] += U.f[f][c];
    }
#pragma unroll
    for (int f = 0; f < 5; f++) red[f][threadIdx.x] = s[f];
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
#pragma unroll
            for (int f = 0; f < 5; f++) red[f][threadIdx.x] += red[f][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) partial[(long long)blockIdx.x * 5 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_sums_final(const double* partial, int nblocks, double* out)
{
    if (threadIdx.x < 5) {
        double s = 0.0;
        for (int b = 0; b < nblocks; b++) s += partial[(long long)b * 5 + threadIdx.x];
        out[threadIdx.x] = s;
    }
}

__global__ void __launch_bounds__(kBlock) k_copy(const double* __restrict__ a, double* __restrict__ b,
                                                 long long n)
{
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        b[t] = a[t];
}

}  // namespace

// ---------------------------------------------------------------------
void launch_ghost_fill(hipStream_t s, State5 U, const Geo& g, const Bnd& b, bool z)
{
    // the z section (last) is left out for readers that take the periodic images
    const long long n = 2LL * (g.H + 2) * (g.D + 2) + 2LL * g.L * (g.D + 2) + (z ? (long long)g.L * g.H : 0LL);
    hipLaunchKernelGGL(k_ghost_fill, dim3(nblk(n)), dim3(kBlock), 0, s, U, g, b);
}

void launch_recon(hipStream_t s, CState5 U, const Faces& F, const Geo& g, int quirk)
{
    const long long n = (long long)g.L * g.H * g.D;
    hipLaunchKernelGGL(k_recon, dim3(nblk(n)), dim3(kBlock), 0, s, U, F, g, quirk);
}

void launch_iface_bc(hipStream_t s, const Faces& F, const Geo& g, const Bnd& b)
{
    const long long n = 2LL * g.H * g.D + (long long)g.L * g.D + (long long)g.L * g.H;
    hipLaunchKernelGGL(k_iface_bc, dim3(nblk(n)), dim3(kBlock), 0, s, F, g, b);
}

void launch_flux(hipStream_t s, int dir, CState5 U, const Faces& F, const double* mu, const Geo& g,
                 const Phys& ph)
{
    const long long n = (long long)(g.L + (dir == 0)) * (g.H + (dir == 1)) * (g.D + (dir == 2));
    if (dir == 0) hipLaunchKernelGGL(k_flux<0>, dim3(nblk(n)), dim3(kBlock), 0, s, U, F, mu, g, ph);
    else if (dir == 1) hipLaunchKernelGGL(k_flux<1>, dim3(nblk(n)), dim3(kBlock), 0, s, U, F, mu, g, ph);
    else hipLaunchKernelGGL(k_flux<2>, dim3(nblk(n)), dim3(kBlock), 0, s, U, F, mu, g, ph);
}

void launch_rk(hipStream_t s, int kind, State5 U, State5 Up, const Faces& F, const Geo& g, Dt3 dt)
{
    const long long n = (long long)g.L * g.H * g.D;
    const dim3 gr(nblk(n)), bl(kBlock);
    switch (kind) {
    case 31: hipLaunchKernelGGL(k_rk<31>, gr, bl, 0, s, U, Up, F, g, dt); break;
    case 32: hipLaunchKernelGGL(k_rk<32>, gr, bl, 0, s, U, Up, F, g, dt); break;
    case 33: hipLaunchKernelGGL(k_rk<33>, gr, bl, 0, s, U, Up, F, g, dt); break;
    case 21: hipLaunchKernelGGL(k_rk<21>, gr, bl, 0, s, U, Up, F, g, dt); break;
    default: hipLaunchKernelGGL(k_rk<22>, gr, bl, 0, s, U, Up, F, g, dt); break;
    }
}

void launch_cfl(hipStream_t s, int mode, CState5 U, const Geo& g, const Phys& ph, double dtX,
                unsigned long long* out)
{
    const long long n = (long long)g.L * g.H * g.D;
    unsigned blocks = nblk(n);
    if (blocks > 2048) blocks = 2048;
    if (mode == 0) hipLaunchKernelGGL(k_cfl<0>, dim3(blocks), dim3(kBlock), 0, s, U, g, ph, dtX, out);
    else hipLaunchKernelGGL(k_cfl<1>, dim3(blocks), dim3(kBlock), 0, s, U, g, ph, dtX, out);
}

int sums_nblocks(const Geo& g)
{
    const long long n = (long long)g.L * g.H * g.D;
    long long b = (n + kBlock - 1) / kBlock;
    return (int)(b > 1024 ? 1024 : b);
}

void launch_sums(hipStream_t s, CState5 U, const Geo& g, double* partial, int nblocks, double* out)
{
    hipLaunchKernelGGL(k_sums_partial, dim3(nblocks), dim3(kBlock), 0, s, U, g, partial);
    hipLaunchKernelGGL(k_sums_final, dim3(1), dim3(64), 0, s, (const double*)partial, nblocks, out);
}

void launch_copy(hipStream_t s, const double* a, double* b, size_t n)
{
    unsigned blocks = nblk((long long)n);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(kBlock), 0, s, a, b, (long long)n);
}

// ---------------------------------------------------------------------
// Allocation-free conservation sums.  The per-block partials of
// k_sums_partial live in a file-scope device array instead of a context
// allocation: at most kSumsMaxBlocks blocks x 5 fields (40 KiB), held by
// every device the library runs on until the process ends.  Each device
// has its own copy; its address is looked up once per device
// (hipGetSymbolAddress, with that device current) and kept here -- the
// host-side name `scratch` is only a stand-in and never reaches a kernel.
namespace {

constexpr int kSumsMaxBlocks = 1024;
constexpr int kScratchDoubles = kSumsMaxBlocks * 5;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

double* g_scratch[kMaxDevices];
std::mutex g_scratch_mu;

int scratch_of_current_device(double** out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return FSL_ERR_HIP;
    if (dev < 0 || dev >= kMaxDevices) return FSL_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "fsl_conservation_sums_noalloc: hipGetSymbolAddress failed on device %d: %s\n", dev,
                    hipGetErrorString(e));
            return FSL_ERR_HIP;
        }
        g_scratch[dev] = (double*)p;
    }
    *out = g_scratch[dev];
    return FSL_OK;
}

}  // namespace
}  // namespace fsl

extern "C" int fsl_conservation_sums_noalloc(const double* U, unsigned LEN, unsigned HIG, unsigned DEP, double* sums,
                                             void* stream)
{
    using namespace fsl;
    if (!U || !sums || LEN < 1 || HIG < 1 || DEP < 1) {
        fprintf(stderr, "fsl_conservation_sums_noalloc: null buffer or empty mesh (%u x %u x %u)\n", LEN, HIG,
                DEP);
        return FSL_ERR_ARG;
    }
    Geo g;
    g.L = (int)LEN;
    g.H = (int)HIG;
    g.D = (int)DEP;
    g.sj = g.D + 2;
    g.si = (long long)(g.H + 2) * (g.D + 2);
    g.G = (long long)(g.L + 2) * g.si;
    const int nb = sums_nblocks(g);
    if (nb > kSumsMaxBlocks) {   // sums_nblocks caps the partial count at the scratch's size
        fprintf(stderr, "fsl_conservation_sums_noalloc: %d partial blocks exceed the %d of the scratch array\n", nb,
                kSumsMaxBlocks);
        return FSL_ERR_UNSUPPORTED;
    }
    double* part = nullptr;
    const int rc = scratch_of_current_device(&part);
    if (rc != FSL_OK) return rc;
    CState5 s;
    for (int f = 0; f < 5; f++) s.f[f] = U + f * g.G;
    // every partial the final kernel reads is written first by this call, so
    // nothing in the scratch needs clearing
    launch_sums((hipStream_t)stream, s, g, part, nb, sums);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        fprintf(stderr, "fsl_conservation_sums_noalloc: launch failed: %s\n", hipGetErrorString(e));
        return FSL_ERR_HIP;
    }
    return FSL_OK;
}

// fsl_fused_mixed.hip -- mixed-precision fused stage kernel.
//
// The decomposition of fsl_fused.hip (16 x 16 workgroup tile with a one-cell
// halo ring marching along x; x faces in registers, y/z faces and viscous
// data through LDS) with the FP32 arithmetic of fsl_device_mixed.hpp.  The
// state, face energies, pressure recovery and the RK accumulation U + dU stay
// in FP64; the increment dU = -dt div F is summed in FP32 from the FP32
// fluxes.  Fluxes, 1/rho-velocities and viscous cell data travel through LDS
// in FP32, which halves the kernel's register and LDS footprint.
#include <stdlib.h>

#include "fsl.h"
#include "fsl_buffer.hpp"
#include "fsl_device_mixed.hpp"
#include "fsl_internal.hpp"

namespace fsl {


namespace {

struct MixedArgs {
    CState5 in;
    CState5 base;
    State5 out;
    const double* mu;
    Geo g;
    Phys ph;
    Bnd b;
    double dtX, dtY, dtZ;
    int segL, nseg, ntj, ntk, nblocks, xcd_remap;
    unsigned fb;                   // bytes of one field (G doubles) in the PK form
    unsigned long long* cfl_out;   // stage 3: CFL maximum of the new state (FP64; or null)
    unsigned long long* cfl_zero;  // stage 2: CFL cell to clear for the following stage 3 (or null)
    int cfl_mode;
};

// v of the lane after / before this one in its row of 16 lanes (DPP
// row_shl:1 / row_shr:1 on both halves); the row's last / first lane, whose
// source lies outside the row, keeps `edge` (bound_ctrl off: the DPP move
// leaves such a lane's destination, preloaded with `edge`, unchanged).
// Every lane of the wave must execute these.
template <int CTRL>
__device__ __forceinline__ double row_dpp(double v, double edge)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned long long e = __builtin_bit_cast(unsigned long long, edge);
    const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)e, (int)(unsigned)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(unsigned)(e >> 32), (int)(unsigned)(b >> 32), CTRL, 0xf, 0xf,
                                               false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double row_from_next(double v, double edge) { return row_dpp<0x101>(v, edge); }
__device__ __forceinline__ double row_from_prev(double v, double edge) { return row_dpp<0x111>(v, edge); }

template <int STAGE, int BJ, int BK, int MINW, bool PK>
__global__ void __launch_bounds__(BJ* BK, MINW) k_fused_mixed(MixedArgs a)
{
    constexpr int TJ = BJ - 2, TK = BK - 2, NT = BJ * BK, NW = (BJ + 2) * BK;
    __shared__ float s_vr[2][3][NT];    // 1/rho velocities of planes i, i + 1 (slot p & 1)
    __shared__ float s_d[3][3][NT];     // their central differences along x, y, z at plane i
    __shared__ float s_visc[5][NT];     // velocity (3), T, rho of plane i
    __shared__ float s_ym[4][NT];       // - face states along y: density, momenta
    __shared__ double s_ymE[NT];        //   and energy
    __shared__ float s_zm[4][NT];       // - face states along z
    __shared__ double s_zmE[NT];
    __shared__ float s_fy[5][NT];       // y flux through the + face
    __shared__ float s_fz[5][NT];       // z flux through the + face
    __shared__ float s_mu[NT];          // mu_SGS of the faces to the cell's x neighbour: (mu(i) + mu(i - 1)) / 2
    __shared__ double s_w[2][5][NW];    // state of planes i, i+1 (slot p & 1) with the rows of outer
                                        // neighbours above and below the tile: y neighbours

    const Geo& g = a.g;
    const int L = g.L, H = g.H, D = g.D;
    int bid = blockIdx.x;
    if (a.xcd_remap) {
        const int per = gridDim.x >> 3;
        bid = (bid & 7) * per + (bid >> 3);
    }
    // stage 2 clears the CFL cell that the following stage 3 reduces into
    // (one vector store; the stream orders it before stage 3)
    if (STAGE == 2 && a.cfl_zero && blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) *a.cfl_zero = 0ull;
    if (bid >= a.nblocks) return;
    const int tk = bid % a.ntk;
    const int rem = bid / a.ntk;
    const int tj = rem % a.ntj;
    const int seg = rem / a.ntj;
    const int k0 = 1 + tk * TK, j0 = 1 + tj * TJ;
    const int i0 = 1 + seg * a.segL;
    const int i1 = min(i0 + a.segL, L + 1);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int c = ty * BK + tx;
    const int j = j0 - 1 + ty, k = k0 - 1 + tx;
    const int jc = min(j, H + 1), kc = min(k, D + 1);
    const bool interior = tx >= 1 && tx <= TK && ty >= 1 && ty <= TJ && j <= H && k <= D;
    const bool yface = tx >= 1 && tx <= TK && ty <= TJ && k <= D && j <= H;
    const bool zface = ty >= 1 && ty <= TJ && tx <= TK && j <= H && k <= D;
    const int jm = max(jc - 1, 0), jp = min(jc + 1, H + 1);
    const long long si = g.si;
    const int pb = (int)(si * 8);                         // bytes of one plane of one field
    const unsigned sjb = (unsigned)g.sj * 8u;            // byte stride of j within a plane
    const unsigned fb = a.fb;
    // State reads take the z-periodic image of a column, k -> (k - 1) mod D + 1
    // (ghost columns k = 0 -> D, k = D + 1 -> 1: the values the ghost fill
    // copies there, so the inputs of stages 2 and 3 need no z section of the
    // ghost fill; lanes beyond the domain, k > D + 1, hold the image of their
    // column, so that the lane of k = D + 1 finds its z + 1 neighbour, the
    // image of k = D + 2, in the next lane).  mu_SGS is read at its own ghost
    // columns (colmu).
    const int kim = ((k - 1) % D + D) % D + 1;
    const unsigned colb = (unsigned)jc * sjb + (unsigned)kim * 8u, colmu = (unsigned)jc * sjb + (unsigned)kc * 8u;
    // Neighbours of a lane's cell in its plane.  z: the cells of the lanes
    // beside it in its row of BK = 16 lanes (one DPP row), moved across by
    // row shifts of the registers; the row's end lanes load the outer
    // neighbour.  y: the cells of the rows above and below through s_w, a
    // copy of the plane in 16-double rows (two rows of one half-wave fill all
    // banks once) with a row of outer neighbours above and below the tile
    // that the tile's edge rows load.  In y the rows are clamped (j - 1 >= 0,
    // j + 1 <= H + 1), in z the columns periodic; every neighbour a used face
    // state needs is the one the reference's ghost layer holds there.
    static_assert(BK == 16, "z neighbours move along one 16-lane DPP row");
    const int cw = (ty + 1) * BK + tx;
    const bool yedge = ty == 0 || ty == BJ - 1, zedge = tx == 0 || tx == BK - 1;
    const int cwy = ty == 0 ? cw - BK : cw + BK;
    // in-tile neighbours of the central differences (clamped at the tile's
    // edges, where the differences are not used)
    const int oyp = ty < BJ - 1 ? c + BK : c, oym = ty > 0 ? c - BK : c;
    const int ozp = tx < BK - 1 ? c + 1 : c, ozm = tx > 0 ? c - 1 : c;
    const int kie = (((tx == 0 ? k - 1 : k + 1) - 1) % D + D) % D + 1;
    const unsigned oye = (unsigned)(ty == 0 ? jm : jp) * sjb + (unsigned)kim * 8u;
    const unsigned oze = (unsigned)jc * sjb + (unsigned)kie * 8u;
    const float r4[3] = {(float)a.ph.r4del[0], (float)a.ph.r4del[1], (float)a.ph.r4del[2]};
    const float rd[3] = {(float)a.ph.rdel[0], (float)a.ph.rdel[1], (float)a.ph.rdel[2]};
    const float dtXf = (float)a.dtX, dtYf = (float)a.dtY, dtZf = (float)a.dtZ;

    double cfl_m = 0.0;   // stage 3: CFL maximum over the cells this thread updates
    // step state of the cell the next plane step updates (stages 2 and 3),
    // loaded one plane step ahead
    double bq[5];
    // stage input of the current and the next plane (that of the previous
    // plane, which the update needs, is read back from the plane copy)
    double w1[5], w2[5];
    double pye[5], pze[5];                   // edge lanes: outer y/z neighbours of the next plane
#pragma unroll
    for (int f = 0; f < 5; f++) pze[f] = 0.;
    float dxp[5];                            // x difference q(i) - q(i - 1) of the current plane
    float ri1, v1[3];                        // 1/rho and velocity of the current plane's cell
    float vm1[3] = {0.f, 0.f, 0.f};          // velocity of the previous plane's cell
    float vm2[3] = {0.f, 0.f, 0.f};          //   and of the one before
    float Tprev = 0.f, Rprev = 0.f;          // temperature and density of the previous plane's cell
    double mu_prev = 0.;                     // its mu_SGS
    float Dy_prev[3] = {0.f, 0.f, 0.f}, Dz_prev[3] = {0.f, 0.f, 0.f};   // y / z differences, plane i - 1
    mx::FaceF xp_prev;
    float Fx_prev[5], ly[5], lz[5];
#pragma unroll
    for (int f = 0; f < 4; f++) xp_prev.f[f] = 0.f;
    xp_prev.E = 0.;
#pragma unroll
    for (int f = 0; f < 5; f++) { Fx_prev[f] = 0.f; ly[f] = 0.f; lz[f] = 0.f; }
    {
        const int ia = max(i0 - 2, 0), ib = i0 - 1;
        load5<PK>(a.in, (long long)ia * si, pb, fb, colb, w1);
        load5<PK>(a.in, (long long)ib * si, pb, fb, colb, w2);
#pragma unroll
        for (int f = 0; f < 5; f++) dxp[f] = (float)(w2[f] - w1[f]);
        ri1 = mx::rcp((float)w2[0]);
#pragma unroll
        for (int e = 0; e < 3; e++) {
            v1[e] = (float)w2[1 + e] * ri1;
            s_vr[ib & 1][e][c] = v1[e];
        }
    }
    lds_barrier();   // the first plane's velocities, for its y / z differences

    for (int i = i0 - 1; i <= i1; i++) {
#pragma unroll
        for (int f = 0; f < 5; f++) w1[f] = w2[f];
        // Loads in the order of first use (the wait counter retires in issue
        // order): next plane's cell and this cell's mu_SGS (end of the
        // reconstruction), the next plane's outer y/z neighbours (before the
        // second barrier); the step state of the cell updated in this plane
        // step was loaded in the previous one.
        load5<PK>(a.in, (long long)min(i + 1, L + 1) * si, pb, fb, colb, w2);
        // own cell's mu_SGS; the faces take theirs from the LDS ring
        const double mu_c = a.ph.dynamic ? bload(plane_rsrc(a.mu, (long long)i * si, pb), colmu) : 0.;
        // y neighbours of this plane's cell (written to s_w in the previous
        // iteration); z neighbours from the lanes beside it, and at a row's
        // ends the outer neighbour loaded in the previous iteration
        double nyp[5], nym[5], nzp[5], nzm[5], w0[5];
        {
            const int sl = i & 1;
#pragma unroll
            for (int f = 0; f < 5; f++) {
                nyp[f] = s_w[sl][f][cw + BK];
                nym[f] = s_w[sl][f][cw - BK];
                // all lanes; the row's end lanes keep their outer neighbour
                nzp[f] = row_from_next(w1[f], pze[f]);
                nzm[f] = row_from_prev(w1[f], pze[f]);
            }
            const long long pn = (long long)min(i + 1, L + 1) * si;
            // stage input of the cell the update below writes (plane i - 1)
#pragma unroll
            for (int f = 0; f < 5; f++) w0[f] = s_w[(i + 1) & 1][f][cw];
            if (yedge) load5<PK>(a.in, pn, pb, fb, oye, pye);
            if (zedge) load5<PK>(a.in, pn, pb, fb, oze, pze);
        }
        // y and z central differences of the velocity at this plane (the
        // tangential gradients of the faces are sums of two of them)
        float Dx[3], Dy[3], Dz[3];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            Dy[q] = s_vr[i & 1][q][oyp] - s_vr[i & 1][q][oym];
            Dz[q] = s_vr[i & 1][q][ozp] - s_vr[i & 1][q][ozm];
            s_d[1][q][c] = Dy[q];
            s_d[2][q][c] = Dz[q];
        }
        // No barrier here: the LDS written before the next barrier (s_vr of
        // plane i + 1, read for the previous plane's differences before its
        // first barrier; s_d, s_visc/s_mu slot of i - 2 and the y/z minus
        // states, read by the previous plane's face loops) was last read
        // before the previous plane's last barrier; s_fy/s_fz, read after
        // that barrier, are rewritten only after the next one.
        float mu_sum = 0.f;
        mx::PrimF P;
        mx::cell_prims(w1, ri1, v1, P);
        const float Tcur = mx::cell_temp(P), Rcur = (float)w1[0];
        s_visc[0][c] = v1[0];
        s_visc[1][c] = v1[1];
        s_visc[2][c] = v1[2];
        s_visc[3][c] = Tcur;
        s_visc[4][c] = Rcur;
        // the y/z reconstruction needs nothing loaded in this iteration; the
        // uses of w2 and mu_c follow it, so that it covers their latency
        const float qf[4] = {(float)w1[0], (float)w1[1], (float)w1[2], (float)w1[3]};
        const bool inseg = (i >= i0 && i < i1);
        mx::FaceF yp, zp;
        if (inseg) {
            mx::FaceF mi;
            mx::recon_dir<2, 3, 1>(P, w1, qf, nyp, nym, false, yp, mi);
#pragma unroll
            for (int f = 0; f < 4; f++) s_ym[f][c] = mi.f[f];
            s_ymE[c] = mi.E;
            mx::recon_dir<3, 1, 2>(P, w1, qf, nzp, nzm, a.ph.z_quirk != 0, zp, mi);
#pragma unroll
            for (int f = 0; f < 4; f++) s_zm[f][c] = mi.f[f];
            s_zmE[c] = mi.E;
        }
        mx::FaceF xp, xm;
        {
            // x: the backward difference is the previous plane's forward one
            float dp[5];
#pragma unroll
            for (int f = 0; f < 5; f++) dp[f] = (float)(w2[f] - w1[f]);
            mx::recon_dir_d<1, 2, 3>(P, w1, qf, dp, dxp, false, xp, xm);
#pragma unroll
            for (int f = 0; f < 5; f++) dxp[f] = dp[f];
        }
        {
            // 1/rho and velocity of plane i + 1 (LDS, and this cell's next
            // iteration); the x difference of this plane
            const float ri = mx::rcp((float)w2[0]);
#pragma unroll
            for (int e = 0; e < 3; e++) {
                const float vn = (float)w2[1 + e] * ri;
                Dx[e] = vn - vm1[e];
                s_d[0][e][c] = Dx[e];
                vm2[e] = vm1[e];
                vm1[e] = v1[e];
                v1[e] = vn;
                s_vr[(i + 1) & 1][e][c] = vn;
            }
            ri1 = ri;
            mu_sum = (float)(0.5 * (mu_c + mu_prev));   // src/fluxes.c:191
            s_mu[c] = mu_sum;
        }
        lds_barrier();

        if (i >= i0 && interior) {
            mx::FaceF l, r;
            if (i == 1) {
                double q[5];
                inflow_face(a.b, H, j - 1, k - 1, q);
                l = mx::to_face(q);
            } else {
                l = xp_prev;
            }
            if (i == L + 1) {
                r = xp_prev;
                mx::outflow_face(a.b.p_out, r);
            } else {
                r = xm;
            }
            float R, vel[3], Pf;
            mx::riemann<0>(l, r, R, vel, Pf);
            // this cell (plane i) and its x neighbour (plane i - 1)
            const float vf[3] = {vm1[0], vm1[1], vm1[2]};
            const float vbk[3] = {vm2[0], vm2[1], vm2[2]};
            float gr[3][3];
#pragma unroll
            for (int q = 0; q < 3; q++) {
                gr[q][0] = (vf[q] - vbk[q]) * rd[0];
                gr[q][1] = (Dy[q] + Dy_prev[q]) * r4[1];
                gr[q][2] = (Dz[q] + Dz_prev[q]) * r4[2];
            }
            float mu_dyn = 0.f;
            if (a.ph.dynamic) mu_dyn = mu_sum;
            float Fx[5];
            mx::face_flux<0>(a.ph, R, vel, Pf, vf, vbk, Tcur, Tprev, Rcur, Rprev, gr, mu_dyn, Fx);
            if (i - 1 >= i0) {
                const long long po = (long long)(i - 1) * si;
                double ov[5];
#pragma unroll
                for (int f = 0; f < 5; f++) {
                    // the increment -dt div F in FP32 (its own rounding is below that of
                    // the FP32 fluxes), accumulated onto the FP64 state
                    const float inc = dtXf * (Fx_prev[f] - Fx[f]) + ly[f] + lz[f];
                    const double s = w0[f] + (double)inc;
                    double o;
                    if (STAGE == 1) o = s;
                    else if (STAGE == 2) o = 0.75 * bq[f] + 0.25 * s;
                    else o = (1. / 3.) * bq[f] + (2. / 3.) * s;
                    bstore(field_rsrc<PK>(a.out, f, po, pb, fb), colb, PK ? f * fb : 0u, o);
                    ov[f] = o;
                }
                if (STAGE == 3 && a.cfl_out) {
                    const double v = mx::cfl_cell(a.cfl_mode, ov, a.ph, a.dtX);
                    cfl_m = nan_max(cfl_m, v);
                }
            }
#pragma unroll
            for (int f = 0; f < 5; f++) Fx_prev[f] = Fx[f];
        }
        // step state of the cell the next plane step updates (i): interior lanes
        // only; loaded here, after this plane step's update, so that its
        // latency is covered by a whole plane step
        if (STAGE > 1 && interior && i >= i0 && i < i1) load5<PK>(a.base, (long long)i * si, pb, fb, colb, bq);
        xp_prev = xp;
        Tprev = Tcur;
        Rprev = Rcur;
        mu_prev = mu_c;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            Dy_prev[q] = Dy[q];
            Dz_prev[q] = Dz[q];
        }

        if (inseg) {
            const float vbk[3] = {vm1[0], vm1[1], vm1[2]};
            const float Tb = Tcur, Rb = Rcur;
            if (yface) {
                mx::FaceF up;   // - state of the face above: the neighbour's
#pragma unroll
                for (int f = 0; f < 4; f++) up.f[f] = s_ym[f][c + BK];
                up.E = s_ymE[c + BK];
                mx::FaceF l, r;
                if (j == 0) {
                    l = up;
                    l.f[2] = -l.f[2];
                } else {
                    l = yp;
                }
                if (j == H) {
                    r = yp;
                    r.f[2] = -r.f[2];
                } else {
                    r = up;
                }
                float R, vel[3], Pf;
                mx::riemann<1>(l, r, R, vel, Pf);
                const float vf[3] = {s_visc[0][c + BK], s_visc[1][c + BK], s_visc[2][c + BK]};
                float gr[3][3];
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    gr[q][1] = (vf[q] - vbk[q]) * rd[1];
                    gr[q][0] = (s_d[0][q][c + BK] + Dx[q]) * r4[0];
                    gr[q][2] = (s_d[2][q][c + BK] + Dz[q]) * r4[2];
                }
                float mu_dyn = 0.f;   // cell (i, j + 1) and its x neighbour (src/fluxes.c:360)
                if (a.ph.dynamic) mu_dyn = s_mu[c + BK];
                float Fy[5];
                mx::face_flux<1>(a.ph, R, vel, Pf, vf, vbk, s_visc[3][c + BK], Tb, s_visc[4][c + BK], Rb,
                                 gr, mu_dyn, Fy);
#pragma unroll
                for (int f = 0; f < 5; f++) s_fy[f][c] = Fy[f];
            }
            if (zface) {
                const mx::FaceF l = zp;
                mx::FaceF r;
#pragma unroll
                for (int f = 0; f < 4; f++) r.f[f] = s_zm[f][c + 1];
                r.E = s_zmE[c + 1];
                float R, vel[3], Pf;
                mx::riemann<2>(l, r, R, vel, Pf);
                const float vf[3] = {s_visc[0][c + 1], s_visc[1][c + 1], s_visc[2][c + 1]};
                float gr[3][3];
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    gr[q][2] = (vf[q] - vbk[q]) * rd[2];
                    gr[q][0] = (s_d[0][q][c + 1] + Dx[q]) * r4[0];
                    gr[q][1] = (s_d[1][q][c + 1] + Dy[q]) * r4[1];
                }
                float mu_dyn = 0.f;   // cell (i, k + 1) and its x neighbour (src/fluxes.c:531)
                if (a.ph.dynamic) mu_dyn = s_mu[c + 1];
                float Fz[5];
                mx::face_flux<2>(a.ph, R, vel, Pf, vf, vbk, s_visc[3][c + 1], Tb, s_visc[4][c + 1], Rb, gr,
                                 mu_dyn, Fz);
#pragma unroll
                for (int f = 0; f < 5; f++) s_fz[f][c] = Fz[f];
            }
        }
        // plane i + 1 into the plane copy (the slot of plane i - 1: its y
        // neighbours were read at the start of the previous plane, before its
        // barriers, and this lane's own cell in the update above), with the
        // outer y neighbours that the edge rows loaded
#pragma unroll
        for (int f = 0; f < 5; f++) s_w[(i + 1) & 1][f][cw] = w2[f];
        if (yedge) {
#pragma unroll
            for (int f = 0; f < 5; f++) s_w[(i + 1) & 1][f][cwy] = pye[f];
        }
        lds_barrier();
        if (inseg && interior) {
#pragma unroll
            for (int f = 0; f < 5; f++) {
                ly[f] = dtYf * (s_fy[f][c - BK] - s_fy[f][c]);
                lz[f] = dtZf * (s_fz[f][c - 1] - s_fz[f][c]);
            }
        }
    }
    if (STAGE == 3 && a.cfl_out) cfl_fold(cfl_m, a.cfl_out);
}

template <int BJ, int BK, int MINW, bool PK>
void launch_mixed(hipStream_t s, int stage, MixedArgs a, int per_cu)
{
    const int TJ = BJ - 2, TK = BK - 2;
    a.ntj = (a.g.H + TJ - 1) / TJ;
    a.ntk = (a.g.D + TK - 1) / TK;
    const int cols = a.ntj * a.ntk;
    int segL = 0;
    const int nseg = fused_segments(a.g.L, cols, per_cu, &segL);
    a.segL = segL;
    a.nseg = nseg;
    a.nblocks = nseg * cols;
    a.xcd_remap = 1;
    const dim3 grid((a.nblocks + 7) / 8 * 8), block(BK, BJ);
    if (stage == 1) hipLaunchKernelGGL((k_fused_mixed<1, BJ, BK, MINW, PK>), grid, block, 0, s, a);
    else if (stage == 2) hipLaunchKernelGGL((k_fused_mixed<2, BJ, BK, MINW, PK>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_fused_mixed<3, BJ, BK, MINW, PK>), grid, block, 0, s, a);
}

}  // namespace

void launch_fused_mixed(const CtxView& v, int stage, const double* in, double* out)
{
    MixedArgs a{};
    for (int f = 0; f < 5; f++) {
        a.in.f[f] = in + f * v.g.G;
        a.base.f[f] = v.U + f * v.g.G;
        a.out.f[f] = out + f * v.g.G;
    }
    a.mu = v.mu;
    a.g = v.g;
    a.ph = v.ph;
    a.b = v.bnd;
    a.dtX = v.dtX;
    a.dtY = v.dtY;
    a.dtZ = v.dtZ;
    a.cfl_out = stage == 3 ? v.red : nullptr;
    a.cfl_zero = stage == 2 ? v.red : nullptr;
    a.cfl_mode = v.cfl_mode;
    // 16 x 16 workgroups at two waves per SIMD (two workgroups per CU, whose
    // barriers overlap); 32 x 16 and 16 x 32 tiles (one workgroup per CU)
    // and three waves per SIMD (register spills) measured slower
    // one buffer resource per state plane where the five fields fit 32-bit
    // offsets, (4 G + one plane) * 8 < 2^32 bytes: up to about 1.3e8 cells;
    // FSL_MIXED_FIELD_RSRC=1 selects the per-field form at any size
    const unsigned long long span = (4ull * (unsigned long long)v.g.G + (unsigned long long)v.g.si) * 8ull;
    const char* e = getenv("FSL_MIXED_FIELD_RSRC");
    if (span <= 0xffffffffull && !(e && atoi(e) == 1)) {
        a.fb = (unsigned)(v.g.G * 8);
        launch_mixed<16, 16, 2, true>(v.stream, stage, a, 2);
    } else {
        launch_mixed<16, 16, 2, false>(v.stream, stage, a, 2);
    }
}

}  // namespace fsl

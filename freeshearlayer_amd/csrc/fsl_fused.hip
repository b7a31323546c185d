// fsl_fused.hip -- fused stage kernel: reconstruction, interface boundary
// states, Riemann solver, viscous/SGS fluxes and the TVD-RK update of one
// stage in a single pass over the stage state.
//
// Reference functions fused here (nikola-m/freeShearLayer):
//   src/reconstruction.c:26-345, src/bounCondOnInterfaces.c:11-110,
//   src/fluxes.c:57-571, src/evolution.c:29-161 (three-stage scheme).
// The arithmetic is the device building blocks of fsl_device.hpp in the
// reference's operation order, so the result is bit-identical to the unfused
// path (and to src/ with real = double).
//
// Decomposition.  A workgroup owns a 14 x 14 tile of (j, k) columns plus a
// one-cell halo ring (16 x 16 threads, k on the fast lane index so every row
// access is a coalesced 128-byte run) and marches along x over a segment of
// cells.  Per plane i each thread computes, once, its cell's primitives and
// its x, y and z face states; the x faces never leave registers (the march
// carries the previous plane's + state and flux), the y and z face states,
// the fluxes and the per-cell viscous data are exchanged through LDS, and a
// three-plane LDS ring of 1/rho-velocities feeds the tangential gradients.
// HBM traffic per cell and stage is the stage state (plus halo re-reads that
// hit L2), the step state for stages 2 and 3, mu_SGS and the new state.
//
// Addressing.  Every access goes through a buffer resource of one x plane
// (fsl_buffer.hpp): the plane is uniform over the workgroup, so its base is
// scalar, and a lane adds a 32-bit in-plane byte offset.  The offsets of a
// lane's own cell and of its y and z neighbours do not change along the
// march and are computed once, so the 30 state loads of a plane step carry
// no per-lane address arithmetic.
#include <stdlib.h>

#include "fsl.h"
#include "fsl_buffer.hpp"
#include "fsl_internal.hpp"

namespace fsl {

CtxView ctx_view(fsl_ctx* c);
void ctx_prof(fsl_ctx* c, int id, int end);
int ctx_check(fsl_ctx* c, const char* what);
int ctx_sgs(fsl_ctx* c);
int learned_sgs_step(fsl_ctx* c);
int ctx_cfl_reset(fsl_ctx* c);
void ctx_cfl_ready(fsl_ctx* c);
void ctx_cfl_zeroed(fsl_ctx* c, int z);
int ctx_ghost(fsl_ctx* c, const double* state, bool z = true);
void ctx_set_cur(fsl_ctx* c, int is_p);
int ctx_err(fsl_ctx* c, int code, const char* msg);

#ifdef FSL_FP64_FAST
#define k_fused k_fused_fast   // the fast build's kernels under a name of their own (profiles)
#endif

namespace {

struct FusedArgs {
    CState5 in;      // stage input, ghost layer filled
    CState5 base;    // step state (stages 2 and 3)
    State5 out;      // new state (interior written)
    const double* mu;
    Geo g;
    Phys ph;
    Bnd b;
    double dtX, dtY, dtZ;
    int segL, nseg, ntj, ntk, nblocks, xcd_remap;
    unsigned fb;                   // bytes of one field (G doubles) in the one-resource-per-plane form
    unsigned long long* cfl_out;   // stage 3: CFL maximum of the new state (or null)
    unsigned long long* cfl_zero;  // stage 2: CFL cell to clear for the following stage 3 (or null)
    int cfl_mode;
};

__device__ __forceinline__ int ring(int p) { return ((p % 3) + 3) % 3; }

// BJ x BK threads (BK on the fast lane index), MINW: minimum waves per SIMD
// the register allocation must allow; PK: one buffer resource per state plane.
template <int STAGE, int BJ, int BK, int MINW, bool PK>
__global__ void __launch_bounds__(BJ* BK, MINW) k_fused(FusedArgs a)
{
    constexpr int TJ = BJ - 2, TK = BK - 2, NT = BJ * BK;
    __shared__ double s_vr[3][3][NT];   // 1/rho velocities, planes i-1, i, i+1
    __shared__ double s_visc[2][5][NT]; // rho u / rho (3), T, rho of planes i-1, i
    __shared__ double s_ym[5][NT];      // - face states along y
    __shared__ double s_zm[5][NT];      // - face states along z
    __shared__ double s_fy[5][NT];      // own + y state, then the y flux through it
    __shared__ double s_fz[5][NT];      // own + z state, then the z flux through it

    const Geo& g = a.g;
    const int L = g.L, H = g.H, D = g.D;

    int bid = blockIdx.x;
    if (a.xcd_remap) {
        // consecutive logical tiles on one XCD (workgroups are dealt to the
        // eight XCDs round-robin), so neighbours share that XCD's L2
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
    // y-recon neighbours (clamped: ghost rows never use their y states)
    const int jm = max(jc - 1, 0), jp = min(jc + 1, H + 1);
    // z-recon neighbours; the ghost columns reconstruct their periodic partner
    // (k = 0 -> cell D's + state, k = D+1 -> cell 1's - state)
    int zkm = kc - 1, zkp = kc + 1;
    if (kc == 0) { zkm = D - 1; zkp = D + 1; }
    else if (kc == D + 1) { zkm = 0; zkp = 2; }
    const long long si = g.si;
    const int pb = (int)(si * 8);                 // bytes of one plane of one field
    const unsigned fb = a.fb;
    const unsigned sjb = (unsigned)g.sj * 8u;
    // in-plane byte offsets of this lane's cell and of its neighbours
    const unsigned oc = (unsigned)jc * sjb + (unsigned)kc * 8u;
    const unsigned oyp = (unsigned)jp * sjb + (unsigned)kc * 8u, oym = (unsigned)jm * sjb + (unsigned)kc * 8u;
    const unsigned ozp = (unsigned)jc * sjb + (unsigned)zkp * 8u, ozm = (unsigned)jc * sjb + (unsigned)zkm * 8u;
    // mu_SGS of the y and z face neighbours (cells (j + 1) and (k + 1), read
    // by face lanes only, whose j <= HIG and k <= DEP)
    const unsigned omy = oc + sjb, omz = oc + 8u;

    double cfl_m = 0.0;           // stage 3: CFL maximum over the cells this thread updates
    double bq[5];                 // step state of the cell the next plane step updates (stages 2, 3)
    double w0[5], w1[5], w2[5];   // stage state at planes i-1, i, i+1 (own column)
    double xp_prev[5];            // + x state of plane i-1
    double Fx_prev[5];
    double ly[5], lz[5];
    double mu_x = 0., mu_y = 0., mu_z = 0.;   // mu_SGS of plane i-1: own cell, (j+1), (k+1)
#pragma unroll
    for (int f = 0; f < 5; f++) { xp_prev[f] = 0.; Fx_prev[f] = 0.; ly[f] = 0.; lz[f] = 0.; }

    // 1/rho and velocity of the next plane's cell (src/reconstruction.c:71-74
    // and src/fluxes.c:124-127 compute them alike), carried to the plane
    // where that cell is reconstructed
    double rr2, vr2[3];
    {
        const int ia = max(i0 - 2, 0), ib = i0 - 1;
        load5<PK>(a.in, (long long)ia * si, pb, fb, oc, w1);
        load5<PK>(a.in, (long long)ib * si, pb, fb, oc, w2);
        vel_recip(w2[0], w2[1], w2[2], w2[3], vr2, &rr2);
        const int r = ring(ib);
#pragma unroll
        for (int e = 0; e < 3; e++) s_vr[r][e][c] = vr2[e];
    }

    for (int i = i0 - 1; i <= i1; i++) {
#pragma unroll
        for (int f = 0; f < 5; f++) { w0[f] = w1[f]; w1[f] = w2[f]; }
        const long long pl = (long long)i * si;
        // loads in the order of first use: this plane's y neighbours (y
        // reconstruction), the next plane's cell (1/rho velocity), mu_SGS of
        // this plane (the step state of the cell this plane step updates was
        // loaded in the previous one)
        const bool inseg = (i >= i0 && i < i1);
        double qyp[5], qym[5];
        if (inseg) {
            load5<PK>(a.in, pl, pb, fb, oyp, qyp);
            load5<PK>(a.in, pl, pb, fb, oym, qym);
        }
        load5<PK>(a.in, (long long)min(i + 1, L + 1) * si, pb, fb, oc, w2);
        double m_x = 0., m_y = 0., m_z = 0.;
        if (a.ph.dynamic) {
            const __amdgpu_buffer_rsrc_t mr = plane_rsrc(a.mu, pl, pb);
            m_x = bload(mr, oc);
            // also on the plane before the segment: the segment's first y/z faces need it
            if (yface) m_y = bload(mr, omy);
            if (zface) m_z = bload(mr, omz);
        }
        lds_barrier();   // the previous plane's LDS readers are done
        const int vc = i & 1, vb = vc ^ 1;   // s_visc slots of planes i and i-1
        {
            double vd[3], Tc;
            vel_div_T(w1[0], w1[1], w1[2], w1[3], w1[4], vd, Tc);
            s_visc[vc][0][c] = vd[0];
            s_visc[vc][1][c] = vd[1];
            s_visc[vc][2][c] = vd[2];
            s_visc[vc][3][c] = Tc;
            s_visc[vc][4][c] = w1[0];
        }
        Prim<double> P;
        cell_prims_rv(w1, rr2, vr2, P);
        if (inseg) {
            double qzp[5], qzm[5], pl5[5], mi5[5];
            load5<PK>(a.in, pl, pb, fb, ozp, qzp);
            load5<PK>(a.in, pl, pb, fb, ozm, qzm);
            recon_dir<2, 3, 1>(P, qyp, qym, false, pl5, mi5);
#pragma unroll
            for (int f = 0; f < 5; f++) { s_fy[f][c] = pl5[f]; s_ym[f][c] = mi5[f]; }
            recon_dir<3, 1, 2>(P, qzp, qzm, a.ph.z_quirk != 0, pl5, mi5);
#pragma unroll
            for (int f = 0; f < 5; f++) { s_fz[f][c] = pl5[f]; s_zm[f][c] = mi5[f]; }
        }
        {
            vel_recip(w2[0], w2[1], w2[2], w2[3], vr2, &rr2);
            const int r = ring(i + 1);
#pragma unroll
            for (int e = 0; e < 3; e++) s_vr[r][e][c] = vr2[e];
        }
        double xp[5], xm[5];
        recon_dir<1, 2, 3>(P, w2, w0, false, xp, xm);
        lds_barrier();

        // ---- x face i - 1/2 and the update of cell i - 1 ----
        if (i >= i0 && interior) {
            double l[5], r[5];
            if (i == 1) {
                inflow_face(a.b, H, j - 1, k - 1, l);
            } else {
#pragma unroll
                for (int f = 0; f < 5; f++) l[f] = xp_prev[f];
            }
            if (i == L + 1) {
#pragma unroll
                for (int f = 0; f < 5; f++) r[f] = xp_prev[f];
                outflow_state(a.b.p_out, r);
            } else {
#pragma unroll
                for (int f = 0; f < 5; f++) r[f] = xm[f];
            }
            double R, vel[3], Pf;
            riemann<0>(l, r, R, vel, Pf);
            const int rc = ring(i), rb = ring(i - 1);
            const double vf[3] = {s_visc[vc][0][c], s_visc[vc][1][c], s_visc[vc][2][c]};
            const double vbk[3] = {s_visc[vb][0][c], s_visc[vb][1][c], s_visc[vb][2][c]};
            double gr[3][3];
#pragma unroll
            for (int q = 0; q < 3; q++) {
                gr[q][0] = (vf[q] - vbk[q]) * a.ph.rdel[0];
                gr[q][1] = (s_vr[rc][q][c + BK] + s_vr[rb][q][c + BK] - s_vr[rc][q][c - BK] -
                            s_vr[rb][q][c - BK]) * a.ph.r4del[1];
                gr[q][2] = (s_vr[rc][q][c + 1] + s_vr[rb][q][c + 1] - s_vr[rc][q][c - 1] -
                            s_vr[rb][q][c - 1]) * a.ph.r4del[2];
            }
            const double mu_dyn = a.ph.dynamic ? 0.5 * (m_x + mu_x) : 0.;
            double Fx[5];
            face_flux<0>(a.ph, R, vel, Pf, vf, vbk, s_visc[vc][3][c], s_visc[vb][3][c], s_visc[vc][4][c],
                         s_visc[vb][4][c], gr, mu_dyn, Fx);
            if (i - 1 >= i0) {
                const long long po = (long long)(i - 1) * si;
                double ov[5];
#pragma unroll
                for (int f = 0; f < 5; f++) {
                    const double lx = a.dtX * (Fx_prev[f] - Fx[f]);
                    const double s = w0[f] + lx + ly[f] + lz[f];
                    double o;
                    if (STAGE == 1) o = s;
                    else if (STAGE == 2) o = 0.75 * bq[f] + 0.25 * s;
                    else o = (1. / 3.) * bq[f] + (2. / 3.) * s;
                    bstore(field_rsrc<PK>(a.out, f, po, pb, fb), oc, PK ? f * fb : 0u, o);
                    ov[f] = o;
                }
                if (STAGE == 3 && a.cfl_out) {
                    const double v = cfl_cell(a.cfl_mode, ov, a.ph, a.dtX);
                    cfl_m = nan_max(cfl_m, v);
                }
            }
#pragma unroll
            for (int f = 0; f < 5; f++) Fx_prev[f] = Fx[f];
        }
#pragma unroll
        for (int f = 0; f < 5; f++) xp_prev[f] = xp[f];
        // step state of the cell the next plane step updates (i), interior
        // lanes: loaded one plane step ahead, after this step's update
        if (STAGE > 1 && interior && i >= i0 && i < i1) load5<PK>(a.base, (long long)i * si, pb, fb, oc, bq);

        // ---- y and z faces of plane i ----
        if (inseg) {
            const int rp = ring(i + 1), rc = ring(i), rm = ring(i - 1);
            const double vbk[3] = {s_visc[vc][0][c], s_visc[vc][1][c], s_visc[vc][2][c]};
            const double Tb = s_visc[vc][3][c], Rb = s_visc[vc][4][c];
            if (yface) {
                double l[5], r[5];
                if (j == 0) {
#pragma unroll
                    for (int f = 0; f < 5; f++) l[f] = s_ym[f][c + BK];
                    l[2] = -l[2];
                } else {
#pragma unroll
                    for (int f = 0; f < 5; f++) l[f] = s_fy[f][c];
                }
                if (j == H) {
#pragma unroll
                    for (int f = 0; f < 5; f++) r[f] = s_fy[f][c];
                    r[2] = -r[2];
                } else {
#pragma unroll
                    for (int f = 0; f < 5; f++) r[f] = s_ym[f][c + BK];
                }
                double R, vel[3], Pf;
                riemann<1>(l, r, R, vel, Pf);
                const double vf[3] = {s_visc[vc][0][c + BK], s_visc[vc][1][c + BK], s_visc[vc][2][c + BK]};
                double gr[3][3];
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    gr[q][1] = (vf[q] - vbk[q]) * a.ph.rdel[1];
                    gr[q][0] = (s_vr[rp][q][c + BK] + s_vr[rp][q][c] - s_vr[rm][q][c + BK] -
                                s_vr[rm][q][c]) * a.ph.r4del[0];
                    gr[q][2] = (s_vr[rc][q][c + BK + 1] + s_vr[rc][q][c + 1] - s_vr[rc][q][c + BK - 1] -
                                s_vr[rc][q][c - 1]) * a.ph.r4del[2];
                }
                // cell (i, j + 1) and its x neighbour (src/fluxes.c:360)
                const double mu_dyn = a.ph.dynamic ? 0.5 * (m_y + mu_y) : 0.;
                double Fy[5];
                face_flux<1>(a.ph, R, vel, Pf, vf, vbk, s_visc[vc][3][c + BK], Tb, s_visc[vc][4][c + BK], Rb,
                             gr, mu_dyn, Fy);
#pragma unroll
                for (int f = 0; f < 5; f++) s_fy[f][c] = Fy[f];
            }
            if (zface) {
                double l[5], r[5];
#pragma unroll
                for (int f = 0; f < 5; f++) { l[f] = s_fz[f][c]; r[f] = s_zm[f][c + 1]; }
                double R, vel[3], Pf;
                riemann<2>(l, r, R, vel, Pf);
                const double vf[3] = {s_visc[vc][0][c + 1], s_visc[vc][1][c + 1], s_visc[vc][2][c + 1]};
                double gr[3][3];
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    gr[q][2] = (vf[q] - vbk[q]) * a.ph.rdel[2];
                    gr[q][0] = (s_vr[rp][q][c + 1] + s_vr[rp][q][c] - s_vr[rm][q][c + 1] -
                                s_vr[rm][q][c]) * a.ph.r4del[0];
                    gr[q][1] = (s_vr[rc][q][c + 1 + BK] + s_vr[rc][q][c + BK] - s_vr[rc][q][c + 1 - BK] -
                                s_vr[rc][q][c - BK]) * a.ph.r4del[1];
                }
                // cell (i, k + 1) and its x neighbour (src/fluxes.c:531)
                const double mu_dyn = a.ph.dynamic ? 0.5 * (m_z + mu_z) : 0.;
                double Fz[5];
                face_flux<2>(a.ph, R, vel, Pf, vf, vbk, s_visc[vc][3][c + 1], Tb, s_visc[vc][4][c + 1], Rb,
                             gr, mu_dyn, Fz);
#pragma unroll
                for (int f = 0; f < 5; f++) s_fz[f][c] = Fz[f];
            }
        }
        mu_x = m_x;
        mu_y = m_y;
        mu_z = m_z;
        lds_barrier();
        if (inseg && interior) {
#pragma unroll
            for (int f = 0; f < 5; f++) {
                ly[f] = a.dtY * (s_fy[f][c - BK] - s_fy[f][c]);
                lz[f] = a.dtZ * (s_fz[f][c - 1] - s_fz[f][c]);
            }
        }
    }
    if (STAGE == 3 && a.cfl_out) cfl_fold(cfl_m, a.cfl_out);
}

#ifndef FSL_FP64_FAST   // (shared with the fast build, which links to these)
// Segments along x.  Workgroups of one launch run in rounds of `resident`
// (per_cu per CU); a round takes about (segment length + 2) plane steps
// (the march starts two planes early).  Pick the count that minimises
// rounds x (length + 2), i.e. long segments without a mostly idle last round.
int device_cus()
{
    static int n = 0;
    if (!n) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1)
            n = 256;
    }
    return n;
}
#endif

}  // namespace

#ifndef FSL_FP64_FAST
int fused_segments(int L, int cols, int per_cu, int* seg_len)
{
    const long long resident = (long long)per_cu * device_cus();
    int best = 1, best_len = L;
    long long best_cost = -1;
    for (int nseg = 1; nseg <= L; nseg++) {
        const int len = (L + nseg - 1) / nseg;
        if (len < 4 && L >= 4) break;
        if ((L + len - 1) / len != nseg) continue;   // same length as a smaller count
        const long long rounds = ((long long)nseg * cols + resident - 1) / resident;
        const long long cost = rounds * (len + 2);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = nseg;
            best_len = len;
        }
    }
    *seg_len = best_len;
    return best;
}
#endif

namespace {

struct Launch {
    dim3 grid, block;
    int segL, nseg, ntj, ntk, nblocks, remap;
};

Launch plan(const Geo& g, int BJ, int BK, int per_cu)
{
    const int TJ = BJ - 2, TK = BK - 2;
    Launch p;
    p.ntj = (g.H + TJ - 1) / TJ;
    p.ntk = (g.D + TK - 1) / TK;
    const int cols = p.ntj * p.ntk;
    int segL = 0;
    const int nseg = fused_segments(g.L, cols, per_cu, &segL);
    p.segL = segL;
    p.nseg = nseg;
    p.nblocks = nseg * cols;
    p.remap = 1;
    const int padded = (p.nblocks + 7) / 8 * 8;
    p.grid = dim3(padded);
    p.block = dim3(BK, BJ);
    return p;
}

template <int BJ, int BK, int MINW, bool PK>
void launch_variant(hipStream_t s, int stage, FusedArgs a, int per_cu)
{
    const Launch p = plan(a.g, BJ, BK, per_cu);
    a.segL = p.segL;
    a.nseg = p.nseg;
    a.ntj = p.ntj;
    a.ntk = p.ntk;
    a.nblocks = p.nblocks;
    a.xcd_remap = p.remap;
    if (stage == 1) hipLaunchKernelGGL((k_fused<1, BJ, BK, MINW, PK>), p.grid, p.block, 0, s, a);
    else if (stage == 2) hipLaunchKernelGGL((k_fused<2, BJ, BK, MINW, PK>), p.grid, p.block, 0, s, a);
    else hipLaunchKernelGGL((k_fused<3, BJ, BK, MINW, PK>), p.grid, p.block, 0, s, a);
}

// Occupancy variants of the 16 x 16 tile: 0 = 2 waves/SIMD (default),
// 1 = 1 wave/SIMD (no register limit below 512); read at each launch
int fused_variant()
{
    const char* e = getenv("FSL_FUSED_VARIANT");
    const int v = e ? atoi(e) : 0;
    return (v < 0 || v > 1) ? 0 : v;
}

}  // namespace

// The FP64 launch of one stage.  Built twice: as the parity kernel
// (launch_fused_fp64) and, with FSL_FP64_FAST defined and FMA contraction
// enabled, as the fast build of the same source (launch_fused_fp64_fast).
#ifdef FSL_FP64_FAST
void launch_fused_fp64_fast(const CtxView& v, int stage, const double* in, double* out)
#else
void launch_fused_fp64(const CtxView& v, int stage, const double* in, double* out)
#endif
{
    FusedArgs a{};
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
    // stage 3 also reduces the CFL quantity of the new state for the next
    // step's time step (fsl_check_co_num uses it while U and dt are unchanged)
    a.cfl_out = stage == 3 ? v.red : nullptr;
    a.cfl_zero = stage == 2 ? v.red : nullptr;
    a.cfl_mode = v.cfl_mode;
    // one buffer resource per state plane where the five fields fit 32-bit
    // offsets, (4 G + one plane) * 8 < 2^32 bytes; otherwise one per field
    const unsigned long long span = (4ull * (unsigned long long)v.g.G + (unsigned long long)v.g.si) * 8ull;
    const char* e = getenv("FSL_FP64_FIELD_RSRC");
    const bool pk = span <= 0xffffffffull && !(e && atoi(e) == 1);
    a.fb = (unsigned)(pk ? v.g.G * 8 : 0);
    if (pk) {
        switch (fused_variant()) {
        case 1: launch_variant<16, 16, 1, true>(v.stream, stage, a, 1); break;
        default: launch_variant<16, 16, 2, true>(v.stream, stage, a, 2); break;
        }
    } else {
        launch_variant<16, 16, 2, false>(v.stream, stage, a, 2);
    }
}

#ifndef FSL_FP64_FAST
int fused_stage(fsl_ctx* ctx, int stage)
{
    CtxView v = ctx_view(ctx);
    const double* in = stage == 1 ? v.U : (stage == 2 ? v.Up : v.Uq);
    double* out = stage == 1 ? v.Up : (stage == 2 ? v.Uq : v.U);
    if (!v.Uq) return ctx_err(ctx, FSL_ERR_ARG, "fused path needs the third stage buffer (create with fused = 1)");
    // the mixed stage kernel reads the z-periodic images of the ghost columns
    // of Up and Uq (the stage-1 input U keeps its full ghost layer: the learned
    // closure's features read it)
    int rc = ctx_ghost(ctx, in, !(v.precision == FSL_PREC_MIXED && stage > 1));
    if (rc) return rc;
    if (stage == 1 && v.sgs_model == FSL_SGS_DYNAMIC) {
        // the model reads the step state only: once per step (src/fluxes.c:53)
        rc = ctx_sgs(ctx);
        if (rc) return rc;
    } else if (stage == 1 && v.sgs_model == FSL_SGS_LEARNED) {
        rc = learned_sgs_step(ctx);
        if (rc) return rc;
    }
    v = ctx_view(ctx);   // the learned closure may have swapped the mu_SGS array
    if (stage == 3 && (rc = ctx_cfl_reset(ctx))) return rc;
    ctx_prof(ctx, K_FUSED_XYZ, 0);
    if (v.precision == FSL_PREC_MIXED) launch_fused_mixed(v, stage, in, out);
    else if (v.precision == FSL_PREC_FP64_FAST) launch_fused_fp64_fast(v, stage, in, out);
    else launch_fused_fp64(v, stage, in, out);
    ctx_prof(ctx, K_FUSED_XYZ, 1);
    rc = ctx_check(ctx, "fused stage");
    if (rc) return rc;
    if (stage == 3) ctx_cfl_ready(ctx);
    if (stage == 2) ctx_cfl_zeroed(ctx, 1);
    // stage inputs: U, then Up, then Uq; the state after stage 3 is U
    ctx_set_cur(ctx, stage == 3 ? 0 : 1);
    return FSL_OK;
}
#endif

}  // namespace fsl

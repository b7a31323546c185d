/*
 * fsl_oracle.c -- CPU restatement of the freeShearLayer per-step update (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never part of it.
 * See fsl_oracle.h for the file:line map to the reference and for the layout.
 *
 * Every arithmetic expression keeps the reference's association order so the
 * result is bit-identical to src/ compiled with real=double and no FP
 * contraction (gcc -O2 on x86-64 emits no FMA).  Build this file with
 * -ffp-contract=off.
 */
#include "fsl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* src/def.h:30-41: gamma combinations and gas constants */
#define K_      1.4
#define K_1     0.4
#define NK_1    (-0.4)
#define H05K    0.7
#define K_K_1   3.5
#define K_1_2   0.2
#define I2K_1   1.25
#define CV      717.75
#define R_AIR   287.10
/* src/def.h:45-49: 15-digit truncations of 1/3, 1/6, 2/3 */
#define C1      0.333333333333333
#define C2      0.166666666666667
#define SMALL   1e-30
#define TWO3    0.666666666666667

int or_fault_drop_heat_flux = 0;

size_t or_ncell(const or_params *p)  { return (size_t)(p->L + 2) * (p->H + 2) * (p->D + 2); }
size_t or_nxface(const or_params *p) { return (size_t)(p->L + 1) * p->H * p->D; }
size_t or_nyface(const or_params *p) { return (size_t)p->L * (p->H + 1) * p->D; }
size_t or_nzface(const or_params *p) { return (size_t)p->L * p->H * (p->D + 1); }

#define GI(i, j, k) ((((size_t)(i)) * (H + 2) + (size_t)(j)) * (D + 2) + (size_t)(k))
#define XI(f, j, k) ((((size_t)(f)) * H + (size_t)(j)) * D + (size_t)(k))
#define YI(i, f, k) ((((size_t)(i)) * (H + 1) + (size_t)(f)) * D + (size_t)(k))
#define ZI(i, j, f) ((((size_t)(i)) * H + (size_t)(j)) * (D + 1) + (size_t)(f))

void or_params_default_ref(or_params *p, int L, int H, int D, const double *u_in_storage)
{
    memset(p, 0, sizeof(*p));
    p->L = L; p->H = H; p->D = D;
    /* layer2.ini values */
    p->dx = 1.6666667; p->dy = 1.0; p->dz = 1.5; p->dt = 0.0007;
    p->mu_L = 0.253358; p->Pr_L = 0.72; p->Pr_T = 0.8; p->Cs = 0.01;
    p->BL_HIG = 5; p->Ua = 35.; p->Va = 35.; p->Wa = 35.;
    p->Ns_min = 10; p->Nst = 50; p->nStages = 3; p->maxCoNum = 0.1;
    p->dynamic = 1; p->z_quirk = 1;
    p->rho_in = 1.0; p->p_in = 100000.; p->p_out = 100000.;
    p->u_in = u_in_storage;
    p->X = 0.317391239817;
    p->counter = 0;
}

void or_params_derive(or_params *p)
{
    /* src/initialize.c:67-94 */
    p->dtX = p->dt / p->dx;
    p->dtY = p->dt / p->dy;
    p->dtZ = p->dt / p->dz;
    p->rdx = 1. / p->dx;
    p->rdy = 1. / p->dy;
    p->rdz = 1. / p->dz;
    p->r4dx = 1. / (4 * p->dx);
    p->r4dy = 1. / (4 * p->dy);
    p->r4dz = 1. / (4 * p->dz);
    p->r2dx = p->r4dx + p->r4dx;
    p->r2dy = p->r4dy + p->r4dy;
    p->r2dz = p->r4dz + p->r4dz;
    p->lambda_L = p->mu_L * CV * K_ / p->Pr_L;
    p->cp_Pr_T = CV * K_ / p->Pr_T;
    p->DD = pow(p->dx * p->dy * p->dz, TWO3);
    p->CsDD = p->Cs * p->DD;
}

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

int or_state_alloc(const or_params *p, or_state *s)
{
    int f;
    memset(s, 0, sizeof(*s));
    for (f = 0; f < 5; f++) {
        s->U[f] = zalloc(or_ncell(p));
        s->Up[f] = zalloc(or_ncell(p));
        s->xL[f] = zalloc(or_nxface(p)); s->xR[f] = zalloc(or_nxface(p));
        s->yL[f] = zalloc(or_nyface(p)); s->yR[f] = zalloc(or_nyface(p));
        s->zL[f] = zalloc(or_nzface(p)); s->zR[f] = zalloc(or_nzface(p));
        if (!s->U[f] || !s->Up[f] || !s->xL[f] || !s->xR[f] || !s->yL[f] || !s->yR[f] ||
            !s->zL[f] || !s->zR[f])
            return -1;
    }
    s->mu_sgs = zalloc(or_ncell(p));
    return s->mu_sgs ? 0 : -1;
}

void or_state_free(or_state *s)
{
    int f;
    for (f = 0; f < 5; f++) {
        free(s->U[f]); free(s->Up[f]);
        free(s->xL[f]); free(s->xR[f]); free(s->yL[f]); free(s->yR[f]);
        free(s->zL[f]); free(s->zR[f]);
    }
    free(s->mu_sgs);
    memset(s, 0, sizeof(*s));
}

void or_ic_reference(const or_params *p, or_state *s)
{
    /* src/initialize.c:159-187: two streams, 76.4 below j = H/2, 200 above */
    const int L = p->L, H = p->H, D = p->D;
    int i, j, k;
    for (i = 1; i <= L; i++)
        for (j = 1; j <= H; j++)
            for (k = 1; k <= D; k++) {
                double R = 1.0, P = 100000.;
                double U = (j <= H / 2) ? 76.4 : 200;
                size_t c = GI(i, j, k);
                s->U[0][c] = R;
                s->U[1][c] = R * U;
                s->U[2][c] = 0.;
                s->U[3][c] = 0.;
                s->U[4][c] = P / K_1 + 0.5 * R * (U * U);
            }
    s->cur_is_p = 0;
}

double or_random(double x)
{
    /* src/helpers.c:13-17 */
    x = 11. * x + 3.14159265358979323846;
    return x - (double)((int)x);
}

double or_minmod(double x, double y)
{
    /* src/reconstruction.c:17-24, Chakravarthy-Osher minmod with b = 4 */
    double ay;
    if (x * y <= 0) return 0;
    y *= 4.0;
    ay = (y > 0) ? y : -y;
    if (x > 0) return (x < ay) ? x : ay;
    return (-x < ay) ? x : -ay;
}

void or_block_draw(or_params *p)
{
    /* src/layer2.c:152-162 */
    if (p->counter > 0) { p->counter--; return; }
    p->counter = (int)(p->Ns_min + p->Nst * (p->X = or_random(p->X)));
    p->k_min = (unsigned)((unsigned)p->D * (p->X = or_random(p->X)));
    p->k_max = (unsigned)(p->k_min + (unsigned)p->D * (p->X = or_random(p->X)));
    if (p->k_max > (unsigned)p->D) p->k_max -= (unsigned)p->D;
    p->X = or_random(p->X); p->Ud = p->Ua * (1 - p->X - p->X);
    p->X = or_random(p->X); p->Vd = p->Va * (1 - p->X - p->X);
    p->X = or_random(p->X); p->Wd = p->Wa * (1 - p->X - p->X);
}

double or_check_co_num(or_params *p, const or_state *s)
{
    /* src/helpers.c:23-52: signed streamwise Courant number only */
    const int L = p->L, H = p->H, D = p->D;
    int i, j, k;
    double maxCo = 0.0, co;
    for (k = 1; k <= D; k++)
        for (j = 1; j <= H; j++)
            for (i = 1; i <= L; i++) {
                size_t c = GI(i, j, k);
                double R = s->U[0][c];
                double U = s->U[1][c] / R;
                co = U * p->dtX;
                maxCo = (co > maxCo) ? co : maxCo;
            }
    p->dt *= (p->maxCoNum / maxCo);
    p->dtX = p->dt / p->dx;
    p->dtY = p->dt / p->dy;
    p->dtZ = p->dt / p->dz;
    return maxCo;
}

void or_ghost_fill(const or_params *p, double *const U[5])
{
    /* src/boundCondInGhostCells.c:10-96; the four sweeps run in this order and
       the later sweeps read what the earlier ones wrote (edges, corners). */
    const int L = p->L, H = p->H, D = p->D;
    int i, j, k, f;
    /* x inflow, i = 0 */
    for (j = 1; j <= H; j++)
        for (k = 1; k <= D; k++) {
            double R = p->rho_in, P = p->p_in, u = p->u_in[j - 1];
            size_t c = GI(0, j, k);
            U[0][c] = R;
            U[1][c] = R * u;
            U[2][c] = 0.;
            U[3][c] = 0.;
            U[4][c] = P / K_1 + 0.5 * R * (u * u);
        }
    /* x outflow, i = L+1: momenta copied, E rebuilt with the outflow pressure */
    for (j = 1; j <= H; j++)
        for (k = 1; k <= D; k++) {
            size_t c = GI(L + 1, j, k), s = GI(L, j, k);
            double R, u, v, w;
            U[0][c] = R = U[0][s];
            U[1][c] = u = U[1][s];
            U[2][c] = v = U[2][s];
            U[3][c] = w = U[3][s];
            U[4][c] = p->p_out / K_1 + 0.5 * (u * u + v * v + w * w) / R;
        }
    /* z periodic */
    for (i = 0; i <= L + 1; i++)
        for (j = 1; j <= H; j++)
            for (f = 0; f < 5; f++) {
                U[f][GI(i, j, 0)] = U[f][GI(i, j, D)];
                U[f][GI(i, j, D + 1)] = U[f][GI(i, j, 1)];
            }
    /* y slip: rho v mirrored */
    for (i = 0; i <= L + 1; i++)
        for (k = 0; k <= D + 1; k++)
            for (f = 0; f < 5; f++) {
                U[f][GI(i, 0, k)] = (f == 2) ? -U[f][GI(i, 1, k)] : U[f][GI(i, 1, k)];
                U[f][GI(i, H + 1, k)] = (f == 2) ? -U[f][GI(i, H, k)] : U[f][GI(i, H, k)];
            }
}

/* Characteristic reconstruction of one cell along one direction.
 * n, t1, t2 index the normal and the two tangential momenta (1..3 -> rho u,
 * rho v, rho w); x: (1,2,3), y: (2,3,1), z: (3,1,2) (src/reconstruction.c
 * rotates U->V->W the same way, lines 176, 261).  q is the cell's conserved
 * state, qp / qm the neighbours at +1 / -1 along the direction.  Outputs:
 * plus[5] = state on the + face (left state there), minus[5] = state on the -
 * face (right state there). */
typedef struct cellprim {
    double u1, u2, u3, u4, u5;
    double R, vel[3], P, C;
    double ff, aa, ll, ee, gg, hh, nn;
    double vhh[3], vee[3];
    double S11;
} cellprim;

static void cell_prims(const double q[5], cellprim *c)
{
    /* src/reconstruction.c:64-88, 104, 142-145 */
    c->u1 = q[0]; c->u2 = q[1]; c->u3 = q[2]; c->u4 = q[3]; c->u5 = q[4];
    c->R = 1. / c->u1;
    c->vel[0] = c->u2 * c->R;
    c->vel[1] = c->u3 * c->R;
    c->vel[2] = c->u4 * c->R;
    c->P = (c->u5 - 0.5 * (c->u2 * c->u2 + c->u3 * c->u3 + c->u4 * c->u4) * c->R) * K_1;
    c->C = sqrt(K_ * c->P * c->R);
    c->ff = c->vel[0] * c->vel[0] + c->vel[1] * c->vel[1] + c->vel[2] * c->vel[2];
    c->aa = K_1_2 * c->ff;
    c->ll = 1. / (c->C + c->C);
    c->ee = c->ll / c->C;
    c->gg = -c->ff * c->ee;
    c->hh = -(c->ee + c->ee);
    c->S11 = c->aa - c->C * c->C;
    c->vhh[0] = c->vel[0] * c->hh; c->vhh[1] = c->vel[1] * c->hh; c->vhh[2] = c->vel[2] * c->hh;
    c->vee[0] = c->vel[0] * c->ee; c->vee[1] = c->vel[1] * c->ee; c->vee[2] = c->vel[2] * c->ee;
    c->nn = -0.5 * c->gg;
}

static void recon_dir(const cellprim *c, const double qp[5], const double qm[5],
                      int n, int t1, int t2, int quirk_t2_uses_n,
                      double plus[5], double minus[5])
{
    const double q[5] = { c->u1, c->u2, c->u3, c->u4, c->u5 };
    double m[5], w[5], M[5], W[5], Mh[5], Wh[5], kv[5];
    double bb = NK_1 * c->vel[n - 1], cc = NK_1 * c->vel[t1 - 1], dd = NK_1 * c->vel[t2 - 1];
    double mm = c->vel[n - 1] * c->ll;
    double kk, k2, k3, S41, S51, S12, S42, S52, S24, S25, S54, S55, ee_t2;
    int r;
    /* conserved differences, permuted to (rho, normal, t1, t2, E) */
    const int perm[5] = { 0, n, t1, t2, 4 };
    for (r = 0; r < 5; r++) {
        m[r] = qp[perm[r]] - q[perm[r]];
        w[r] = q[perm[r]] - qm[perm[r]];
    }
    S41 = c->aa - (kk = c->C * c->vel[n - 1]);
    S51 = c->aa + kk;
    S12 = S42 = S52 = bb;
    S42 += c->C;
    S52 -= c->C;
    /* dW = S dU, forward (src/reconstruction.c:113-120) */
    kk = K_1 * m[4]; k2 = cc * m[2]; k3 = dd * m[3];
    M[0] = c->S11 * m[0] + S12 * m[1] + k2 + k3 + kk;
    M[1] = -c->vel[t1 - 1] * m[0] + m[2];
    M[2] = -c->vel[t2 - 1] * m[0] + m[3];
    M[3] = S41 * m[0] + S42 * m[1] + k2 + k3 + kk;
    M[4] = S51 * m[0] + S52 * m[1] + k2 + k3 + kk;
    /* backward */
    kk = K_1 * w[4]; k2 = cc * w[2]; k3 = dd * w[3];
    W[0] = c->S11 * w[0] + S12 * w[1] + k2 + k3 + kk;
    W[1] = -c->vel[t1 - 1] * w[0] + w[2];
    W[2] = -c->vel[t2 - 1] * w[0] + w[3];
    W[3] = S41 * w[0] + S42 * w[1] + k2 + k3 + kk;
    W[4] = S51 * w[0] + S52 * w[1] + k2 + k3 + kk;
    for (r = 0; r < 5; r++) {
        Mh[r] = or_minmod(M[r], W[r]);
        Wh[r] = or_minmod(W[r], M[r]);
    }
    S24 = c->vee[n - 1] + c->ll;
    S25 = c->vee[n - 1] - c->ll;
    S54 = S55 = c->nn + I2K_1;
    S54 += mm; S55 -= mm;
    /* z quirk: the v row uses W*ee (src/reconstruction.c:328,340) */
    ee_t2 = quirk_t2_uses_n ? c->vee[n - 1] : c->vee[t2 - 1];
    /* + face */
    for (r = 0; r < 5; r++) kv[r] = Mh[r] * C1 + Wh[r] * C2;
    kk = kv[3] + kv[4];
    plus[0]  = c->u1 + c->hh * kv[0] + c->ee * kk;
    plus[n]  = q[n]  + c->vhh[n - 1] * kv[0] + S24 * kv[3] + S25 * kv[4];
    plus[t1] = q[t1] + c->vhh[t1 - 1] * kv[0] + kv[1] + c->vee[t1 - 1] * kk;
    plus[t2] = q[t2] + c->vhh[t2 - 1] * kv[0] + kv[2] + ee_t2 * kk;
    plus[4]  = c->u5 + c->gg * kv[0] + c->vel[t1 - 1] * kv[1] + c->vel[t2 - 1] * kv[2]
             + S54 * kv[3] + S55 * kv[4];
    /* - face */
    for (r = 0; r < 5; r++) kv[r] = Mh[r] * C2 + Wh[r] * C1;
    kk = kv[3] + kv[4];
    minus[0]  = c->u1 - c->hh * kv[0] - c->ee * kk;
    minus[n]  = q[n]  - c->vhh[n - 1] * kv[0] - S24 * kv[3] - S25 * kv[4];
    minus[t1] = q[t1] - c->vhh[t1 - 1] * kv[0] - kv[1] - c->vee[t1 - 1] * kk;
    minus[t2] = q[t2] - c->vhh[t2 - 1] * kv[0] - kv[2] - ee_t2 * kk;
    minus[4]  = c->u5 - c->gg * kv[0] - c->vel[t1 - 1] * kv[1] - c->vel[t2 - 1] * kv[2]
              - S54 * kv[3] - S55 * kv[4];
}

void or_reconstruct(const or_params *p, or_state *s)
{
    /* src/reconstruction.c:55-344 */
    const int L = p->L, H = p->H, D = p->D;
    double *const *U = s->cur_is_p ? s->Up : s->U;
    int i, j, k, f;
    for (i = 1; i <= L; i++)
        for (j = 1; j <= H; j++)
            for (k = 1; k <= D; k++) {
                double q[5], qp[5], qm[5], plus[5], minus[5];
                cellprim c;
                for (f = 0; f < 5; f++) q[f] = U[f][GI(i, j, k)];
                cell_prims(q, &c);
                /* x */
                for (f = 0; f < 5; f++) { qp[f] = U[f][GI(i + 1, j, k)]; qm[f] = U[f][GI(i - 1, j, k)]; }
                recon_dir(&c, qp, qm, 1, 2, 3, 0, plus, minus);
                for (f = 0; f < 5; f++) {
                    s->xL[f][XI(i, j - 1, k - 1)] = plus[f];
                    s->xR[f][XI(i - 1, j - 1, k - 1)] = minus[f];
                }
                /* y */
                for (f = 0; f < 5; f++) { qp[f] = U[f][GI(i, j + 1, k)]; qm[f] = U[f][GI(i, j - 1, k)]; }
                recon_dir(&c, qp, qm, 2, 3, 1, 0, plus, minus);
                for (f = 0; f < 5; f++) {
                    s->yL[f][YI(i - 1, j, k - 1)] = plus[f];
                    s->yR[f][YI(i - 1, j - 1, k - 1)] = minus[f];
                }
                /* z */
                for (f = 0; f < 5; f++) { qp[f] = U[f][GI(i, j, k + 1)]; qm[f] = U[f][GI(i, j, k - 1)]; }
                recon_dir(&c, qp, qm, 3, 1, 2, p->z_quirk, plus, minus);
                for (f = 0; f < 5; f++) {
                    s->zL[f][ZI(i - 1, j - 1, k)] = plus[f];
                    s->zR[f][ZI(i - 1, j - 1, k - 1)] = minus[f];
                }
            }
}

void or_iface_bc(const or_params *p, or_state *s)
{
    /* src/bounCondOnInterfaces.c:11-110 */
    const int L = p->L, H = p->H, D = p->D;
    int i, j, k, f;
    /* x inflow, left state of face 0; rows [H/2, H/2+BL_HIG) carry the block */
    for (j = 0; j < H; j++)
        for (k = 0; k < D; k++) {
            double R = p->rho_in, P = p->p_in, U = p->u_in[j], V = 0., W = 0.;
            size_t c = XI(0, j, k);
            if (j >= H / 2 && j < H / 2 + p->BL_HIG) {
                int in_block;
                if (p->k_max > p->k_min)
                    in_block = ((unsigned)k >= p->k_min && (unsigned)k < p->k_max);
                else
                    in_block = !((unsigned)k >= p->k_max && (unsigned)k < p->k_min);
                if (in_block) { U = U + p->Ud; V = p->Vd; W = p->Wd; }
            }
            s->xL[0][c] = R;
            s->xL[1][c] = R * U;
            s->xL[2][c] = R * V;
            s->xL[3][c] = R * W;
            s->xL[4][c] = P / K_1 + 0.5 * R * (U * U + V * V + W * W);
        }
    /* x outflow, right state of face L from its left state */
    for (j = 0; j < H; j++)
        for (k = 0; k < D; k++) {
            size_t c = XI(L, j, k);
            double R, u, v, w;
            s->xR[0][c] = R = s->xL[0][c];
            s->xR[1][c] = u = s->xL[1][c];
            s->xR[2][c] = v = s->xL[2][c];
            s->xR[3][c] = w = s->xL[3][c];
            s->xR[4][c] = p->p_out / K_1 + 0.5 * (u * u + v * v + w * w) / R;
        }
    /* y slip */
    for (i = 0; i < L; i++)
        for (k = 0; k < D; k++)
            for (f = 0; f < 5; f++) {
                double lo = s->yR[f][YI(i, 0, k)], hi = s->yL[f][YI(i, H, k)];
                s->yL[f][YI(i, 0, k)] = (f == 2) ? -lo : lo;
                s->yR[f][YI(i, H, k)] = (f == 2) ? -hi : hi;
            }
    /* z periodic */
    for (i = 0; i < L; i++)
        for (j = 0; j < H; j++)
            for (f = 0; f < 5; f++) {
                s->zL[f][ZI(i, j, 0)] = s->zL[f][ZI(i, j, D)];
                s->zR[f][ZI(i, j, D)] = s->zR[f][ZI(i, j, 0)];
            }
}

/* Linearised characteristic Riemann solver (src/fluxes.c:61-113).  n is the
 * normal velocity index (0..2).  Returns the interface rho, velocity (normal
 * from the solver, tangential upwinded) and p. */
static void riemann(const double l[5], const double r[5], int n, double *Ro, double vel[3], double *Po)
{
    double lv[3], rv[3], lP, rP, lC, rC, ljo, ljp, ljm, rjo, rjp, rjm, RU;
    double Un, Cm, C_p, C_m, C_o, al_p, al_m, al_o, J_p, J_m, J_o, P, R, a;
    double lR = l[0], rR = r[0];
    int t;
    lv[0] = l[1] / lR; lv[1] = l[2] / lR; lv[2] = l[3] / lR;
    lP = (l[4] - 0.5 * lR * (lv[0] * lv[0] + lv[1] * lv[1] + lv[2] * lv[2])) * K_1;
    lC = sqrt(K_ * lP / lR);
    ljo = NK_1 * lP;
    ljp = lv[n] + (RU = lP / (lR * lC));
    ljm = lv[n] - RU;
    rv[0] = r[1] / rR; rv[1] = r[2] / rR; rv[2] = r[3] / rR;
    rP = (r[4] - 0.5 * rR * (rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2])) * K_1;
    rC = sqrt(K_ * rP / rR);
    rjo = NK_1 * rP;
    rjp = rv[n] + (RU = rP / (rR * rC));
    rjm = rv[n] - RU;
    Un = 0.5 * (lv[n] + rv[n]);
    Cm = 0.5 * (lC + rC);
    C_p = Un + Cm; C_m = Un - Cm; C_o = Un;
    if (C_p > 0.) { al_p = K_1_2 * (ljm - ljp) / ljo; J_p = ljp; }
    else          { al_p = K_1_2 * (rjm - rjp) / rjo; J_p = rjp; }
    if (C_m > 0.) { al_m = K_1_2 * (ljp - ljm) / ljo; J_m = ljm; }
    else          { al_m = K_1_2 * (rjp - rjm) / rjo; J_m = rjm; }
    if (C_o > 0.) { a = H05K * (ljp - ljm); al_o = -a * a; J_o = ljo; }
    else          { a = H05K * (rjp - rjm); al_o = -a * a; J_o = rjo; }
    P = (J_p - J_m) / (al_p - al_m);
    Un = J_p - al_p * P;
    R = (J_o - P) / al_o;
    for (t = 0; t < 3; t++) vel[t] = (Un > 0) ? lv[t] : rv[t];
    vel[n] = Un;
    *Ro = R; *Po = P;
}

/* velocity of a cell by reciprocal density (src/fluxes.c:124-127 style) */
static void vel_recip(double *const U[5], size_t c, double v[3])
{
    double RU = 1. / U[0][c];
    v[0] = U[1][c] * RU; v[1] = U[2][c] * RU; v[2] = U[3][c] * RU;
}

/* velocity, T of a cell by division (src/fluxes.c:117-122) */
static void vel_div_T(double *const U[5], size_t c, double v[3], double *T, double *R)
{
    double r = U[0][c], P;
    v[0] = U[1][c] / r; v[1] = U[2][c] / r; v[2] = U[3][c] / r;
    P = (U[4][c] - 0.5 * r * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2])) * K_1;
    *T = P / (R_AIR * r);
    *R = r;
}

/* Flux through one face of direction d (0 x, 1 y, 2 z) between ghosted cells
 * cb (backward) and cf (forward).  lft/rgt: face states.  Strides s[3] of the
 * cell array move one cell along x, y, z.  mua, mub: the two mu_SGS samples
 * averaged in dynamic mode.  src/fluxes.c:57-223 (x), 226-394 (y), 397-571 (z). */
static void face_flux(const or_params *p, double *const U[5], int d,
                      const double lft[5], const double rgt[5],
                      size_t cb, size_t cf, const size_t st[3],
                      double mua, double mub, double F[5])
{
    double R, vel[3], P, vf[3], vb[3], Tf, Tb, Rf, Rb, g[3][3];
    double uu, vv, ww, mu_T, mu_E, lambda_E, s0, s1, s2, q, RU;
    const double rdel[3] = { p->rdx, p->rdy, p->rdz };
    const double r4del[3] = { p->r4dx, p->r4dy, p->r4dz };
    int e, c;
    riemann(lft, rgt, d, &R, vel, &P);
    vel_div_T(U, cf, vf, &Tf, &Rf);
    vel_div_T(U, cb, vb, &Tb, &Rb);
    /* g[c][e] = d(vel_c)/d(x_e) */
    for (e = 0; e < 3; e++) {
        if (e == d) {
            for (c = 0; c < 3; c++) g[c][e] = (vf[c] - vb[c]) * rdel[e];
        } else {
            double fp[3], bp[3], fm[3], bm[3];
            vel_recip(U, cf + st[e], fp);
            vel_recip(U, cb + st[e], bp);
            vel_recip(U, cf - st[e], fm);
            vel_recip(U, cb - st[e], bm);
            for (c = 0; c < 3; c++) g[c][e] = (fp[c] + bp[c] - fm[c] - bm[c]) * r4del[e];
        }
    }
    uu = 0.5 * (vf[0] + vb[0]);
    vv = 0.5 * (vf[1] + vb[1]);
    ww = 0.5 * (vf[2] + vb[2]);
    if (p->dynamic) {
        mu_T = 0.5 * (mua + mub);
    } else {
        /* src/fluxes.c:195-203 */
        double rr = 0.5 * (Rf + Rb);
        double S12 = 0.5 * (g[0][1] + g[1][0]);
        double S13 = 0.5 * (g[0][2] + g[2][0]);
        double S23 = 0.5 * (g[1][2] + g[2][1]);
        double S = sqrt(2 * (g[0][0] * g[0][0] + g[1][1] * g[1][1] + g[2][2] * g[2][2]
                             + 2 * (S12 * S12 + S13 * S13 + S23 * S23)));
        mu_T = rr * p->CsDD * S;
    }
    mu_E = p->mu_L + mu_T;
    if (d == 0) {
        s0 = TWO3 * mu_E * (g[0][0] + g[0][0] - g[1][1] - g[2][2]);   /* sigma_xx */
        s1 = mu_E * (g[0][1] + g[1][0]);                               /* sigma_xy */
        s2 = mu_E * (g[0][2] + g[2][0]);                               /* sigma_xz */
    } else if (d == 1) {
        s0 = mu_E * (g[1][0] + g[0][1]);                               /* sigma_yx */
        s1 = TWO3 * mu_E * (g[1][1] + g[1][1] - g[0][0] - g[2][2]);   /* sigma_yy */
        s2 = mu_E * (g[1][2] + g[2][1]);                               /* sigma_yz */
    } else {
        s0 = mu_E * (g[2][0] + g[0][2]);                               /* sigma_zx */
        s1 = mu_E * (g[2][1] + g[1][2]);                               /* sigma_zy */
        s2 = TWO3 * mu_E * (g[2][2] + g[2][2] - g[0][0] - g[1][1]);   /* sigma_zz */
    }
    lambda_E = p->lambda_L + mu_T * p->cp_Pr_T;
    q = -lambda_E * (Tf - Tb) * rdel[d];
    if (or_fault_drop_heat_flux) q = 0.0;
    /* the normal momentum row carries + p (src/fluxes.c:215-220, 386-391, 563-568) */
    F[0] = RU = R * vel[d];
    F[1] = (d == 0) ? RU * vel[0] + P - s0 : RU * vel[0] - s0;
    F[2] = (d == 1) ? RU * vel[1] + P - s1 : RU * vel[1] - s1;
    F[3] = (d == 2) ? RU * vel[2] + P - s2 : RU * vel[2] - s2;
    F[4] =vel[d] * (K_K_1 * P + 0.5 * R * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]))
         - uu * s0 - vv * s1 - ww * s2 + q;
}

void or_fluxes(const or_params *p, or_state *s)
{
    const int L = p->L, H = p->H, D = p->D;
    double *const *U = s->cur_is_p ? s->Up : s->U;
    const size_t st[3] = { (size_t)(H + 2) * (D + 2), (size_t)(D + 2), 1 };
    int i, j, k, f;
    double lft[5], rgt[5], F[5];
    if (p->dynamic) or_dynamic_smagorinsky(p, s->U, s->mu_sgs);   /* src/fluxes.c:52-54 */
    /* x faces */
    for (i = 0; i <= L; i++)
        for (j = 0; j < H; j++)
            for (k = 0; k < D; k++) {
                size_t c = XI(i, j, k);
                for (f = 0; f < 5; f++) { lft[f] = s->xL[f][c]; rgt[f] = s->xR[f][c]; }
                face_flux(p, U, 0, lft, rgt, GI(i, j + 1, k + 1), GI(i + 1, j + 1, k + 1), st,
                          s->mu_sgs[GI(i + 1, j + 1, k + 1)], s->mu_sgs[GI(i, j + 1, k + 1)], F);
                for (f = 0; f < 5; f++) s->xL[f][c] = F[f];
            }
    /* y faces; mu_SGS samples (i+1, j+1, k+1) and (i, j+1, k+1) as shipped */
    for (i = 0; i < L; i++)
        for (j = 0; j <= H; j++)
            for (k = 0; k < D; k++) {
                size_t c = YI(i, j, k);
                for (f = 0; f < 5; f++) { lft[f] = s->yL[f][c]; rgt[f] = s->yR[f][c]; }
                face_flux(p, U, 1, lft, rgt, GI(i + 1, j, k + 1), GI(i + 1, j + 1, k + 1), st,
                          s->mu_sgs[GI(i + 1, j + 1, k + 1)], s->mu_sgs[GI(i, j + 1, k + 1)], F);
                for (f = 0; f < 5; f++) s->yL[f][c] = F[f];
            }
    /* z faces; same mu_SGS sampling */
    for (i = 0; i < L; i++)
        for (j = 0; j < H; j++)
            for (k = 0; k <= D; k++) {
                size_t c = ZI(i, j, k);
                for (f = 0; f < 5; f++) { lft[f] = s->zL[f][c]; rgt[f] = s->zR[f][c]; }
                face_flux(p, U, 2, lft, rgt, GI(i + 1, j + 1, k), GI(i + 1, j + 1, k + 1), st,
                          s->mu_sgs[GI(i + 1, j + 1, k + 1)], s->mu_sgs[GI(i, j + 1, k + 1)], F);
                for (f = 0; f < 5; f++) s->zL[f][c] = F[f];
            }
}

static void evo_three(const or_params *p, or_state *s, int stage)
{
    /* src/evolution.c:29-161 */
    const int L = p->L, H = p->H, D = p->D;
    double one_third = 1. / 3., two_thirds = 2. / 3.;
    int i, j, k, f;
    for (i = 1; i <= L; i++)
        for (j = 1; j <= H; j++)
            for (k = 1; k <= D; k++) {
                size_t c = GI(i, j, k);
                for (f = 0; f < 5; f++) {
                    double lx = p->dtX * (s->xL[f][XI(i - 1, j - 1, k - 1)] - s->xL[f][XI(i, j - 1, k - 1)]);
                    double ly = p->dtY * (s->yL[f][YI(i - 1, j - 1, k - 1)] - s->yL[f][YI(i - 1, j, k - 1)]);
                    double lz = p->dtZ * (s->zL[f][ZI(i - 1, j - 1, k - 1)] - s->zL[f][ZI(i - 1, j - 1, k)]);
                    if (stage == 1)
                        s->Up[f][c] = s->U[f][c] + lx + ly + lz;
                    else if (stage == 2)
                        s->Up[f][c] = 0.75 * s->U[f][c] + 0.25 * (s->Up[f][c] + lx + ly + lz);
                    else
                        s->U[f][c] = one_third * s->U[f][c] + two_thirds * (s->Up[f][c] + lx + ly + lz);
                }
            }
    s->cur_is_p = (stage == 1 || stage == 2) ? 1 : 0;
}

static void evo_two(const or_params *p, or_state *s, int stage)
{
    /* src/evolution.c:164-245 */
    const int L = p->L, H = p->H, D = p->D;
    int i, j, k, f;
    for (i = 1; i <= L; i++)
        for (j = 1; j <= H; j++)
            for (k = 1; k <= D; k++) {
                size_t c = GI(i, j, k);
                for (f = 0; f < 5; f++) {
                    double lx = p->dtX * (s->xL[f][XI(i - 1, j - 1, k - 1)] - s->xL[f][XI(i, j - 1, k - 1)]);
                    double ly = p->dtY * (s->yL[f][YI(i - 1, j - 1, k - 1)] - s->yL[f][YI(i - 1, j, k - 1)]);
                    double lz = p->dtZ * (s->zL[f][ZI(i - 1, j - 1, k - 1)] - s->zL[f][ZI(i - 1, j - 1, k)]);
                    if (stage == 1)
                        s->Up[f][c] = s->U[f][c] + lx + ly + lz;
                    else
                        s->U[f][c] = 0.5 * (s->U[f][c] + s->Up[f][c] + lx + ly + lz);
                }
            }
    s->cur_is_p = (stage == 1) ? 1 : 0;
}

void or_evolution(const or_params *p, or_state *s, int numStages, int stage)
{
    /* src/evolution.c:15-27: case 2 falls through into the 3-stage update */
    if (numStages == 2) evo_two(p, s, stage);
    evo_three(p, s, stage);
}

/* filter_ (src/turbulence.c:17-100).  Three separable passes y, z, x over
 * index ranges that leave the scratch arrays zero outside them; the scratch
 * arrays of the reference are zero-initialised here. */
void or_filter(const or_params *p, const double *in, double *out, const double h[3])
{
    const int L = p->L, H = p->H, D = p->D;
    size_t n = or_ncell(p);
    double *t1 = out;                 /* U__ of the reference */
    double *t2 = zalloc(n);           /* U_  of the reference */
    int i, j, k;
    for (i = 0; i <= L; i++)
        for (j = 1; j <= H; j++)
            for (k = 0; k <= D; k++) {
                double tmp = 0.;
                tmp += in[GI(i, j - 1, k)] * h[0];
                tmp += in[GI(i, j, k)] * h[1];
                tmp += in[GI(i, j + 1, k)] * h[2];
                t1[GI(i, j, k)] = tmp;
            }
    for (i = 0; i <= L; i++)
        for (j = 0; j <= H; j++)
            for (k = 1; k <= D; k++) {
                double tmp = 0.;
                tmp += t1[GI(i, j, k - 1)] * h[0];
                tmp += t1[GI(i, j, k)] * h[1];
                tmp += t1[GI(i, j, k + 1)] * h[2];
                t2[GI(i, j, k)] = tmp;
            }
    for (i = 1; i <= L; i++)
        for (j = 0; j <= H; j++)
            for (k = 0; k <= D; k++) {
                double tmp = 0.;
                tmp += t2[GI(i - 1, j, k)] * h[0];
                tmp += t2[GI(i, j, k)] * h[1];
                tmp += t2[GI(i + 1, j, k)] * h[2];
                out[GI(i, j, k)] = tmp;
            }
    free(t2);
}

/* src/turbulence.c:103-507.  With lm (and mm, fac) given, also the Germano
 * terms L:M, M:M and rho DD |S| of every interior cell (the pointwise model's
 * intermediates, from which the averaged coefficients are formed). */
static void dynamic_core(const or_params *p, double *const U[5], double *mu_sgs, double *lm, double *mm,
                         double *fac)
{
    /* src/turbulence.c:103-507, Germano-Lilly, pointwise Cd clipped to [0,0.15] */
    const int L = p->L, H = p->H, D = p->D;
    const size_t n = or_ncell(p);
    const double h[3] = { 1. / 3., 1. / 3., 1. / 3. };
    double *u = zalloc(n), *v = zalloc(n), *w = zalloc(n);
    double *prod[9], *fil[9], *uf = zalloc(n), *vf = zalloc(n), *wf = zalloc(n);
    double *A[9], *Af[9], *B[9], *mag = zalloc(n);
    size_t c;
    int a, i, j, k;
    for (a = 0; a < 9; a++) {
        prod[a] = zalloc(n); fil[a] = zalloc(n);
        A[a] = zalloc(n); Af[a] = zalloc(n); B[a] = zalloc(n);
    }
    for (c = 0; c < n; c++) {
        double rr = 1. / U[0][c];
        u[c] = U[1][c] * rr;
        v[c] = U[2][c] * rr;
        w[c] = U[3][c] * rr;
    }
    for (c = 0; c < n; c++) {
        double q[3] = { u[c], v[c], w[c] };
        for (a = 0; a < 9; a++) prod[a][c] = q[a / 3] * q[a % 3];
    }
    or_filter(p, u, uf, h);
    or_filter(p, v, vf, h);
    or_filter(p, w, wf, h);
    for (a = 0; a < 9; a++) or_filter(p, prod[a], fil[a], h);
    for (i = 1; i <= L; i++)
        for (j = 1; j <= H; j++)
            for (k = 1; k <= D; k++) {
                const size_t sx = (size_t)(H + 2) * (D + 2), sy = (size_t)(D + 2);
                double g[3][3], gf[3][3], S[9], Sf[9], Smag, Sfmag;
                double *vel[3] = { u, v, w }, *velf[3] = { uf, vf, wf };
                int r;
                c = GI(i, j, k);
                {
                    double q[3] = { uf[c], vf[c], wf[c] };
                    for (a = 0; a < 9; a++) prod[a][c] = fil[a][c] - q[a / 3] * q[a % 3];   /* L_ij */
                }
                for (r = 0; r < 3; r++) {
                    g[r][0] = (vel[r][c + sx] - vel[r][c - sx]) * p->r2dx;
                    g[r][1] = (vel[r][c + sy] - vel[r][c - sy]) * p->r2dy;
                    g[r][2] = (vel[r][c + 1] - vel[r][c - 1]) * p->r2dz;
                    gf[r][0] = (velf[r][c + sx] - velf[r][c - sx]) * p->r2dx;
                    gf[r][1] = (velf[r][c + sy] - velf[r][c - sy]) * p->r2dy;
                    gf[r][2] = (velf[r][c + 1] - velf[r][c - 1]) * p->r2dz;
                }
                S[0] = g[0][0]; S[1] = 0.5 * (g[0][1] + g[1][0]); S[2] = 0.5 * (g[0][2] + g[2][0]);
                S[3] = S[1]; S[4] = g[1][1]; S[5] = 0.5 * (g[1][2] + g[2][1]);
                S[6] = S[2]; S[7] = S[5]; S[8] = g[2][2];
                Smag = sqrt(2 * (S[0] * S[0] + S[1] * S[1] + S[2] * S[2]
                               + S[3] * S[3] + S[4] * S[4] + S[5] * S[5]
                               + S[6] * S[6] + S[7] * S[7] + S[8] * S[8]));
                mag[c] = Smag;
                Sf[0] = gf[0][0]; Sf[1] = 0.5 * (gf[0][1] + gf[1][0]); Sf[2] = 0.5 * (gf[0][2] + gf[2][0]);
                Sf[3] = Sf[1]; Sf[4] = gf[1][1]; Sf[5] = 0.5 * (gf[1][2] + gf[2][1]);
                Sf[6] = Sf[2]; Sf[7] = Sf[5]; Sf[8] = gf[2][2];
                Sfmag = sqrt(2 * (Sf[0] * Sf[0] + Sf[1] * Sf[1] + Sf[2] * Sf[2]
                                + Sf[3] * Sf[3] + Sf[4] * Sf[4] + Sf[5] * Sf[5]
                                + Sf[6] * Sf[6] + Sf[7] * Sf[7] + Sf[8] * Sf[8]));
                for (a = 0; a < 9; a++) {
                    B[a][c] = Sfmag * Sf[a];
                    A[a][c] = Smag * S[a];
                }
            }
    for (a = 0; a < 9; a++) or_filter(p, A[a], Af[a], h);
    for (i = 1; i <= L; i++)
        for (j = 1; j <= H; j++)
            for (k = 1; k <= D; k++) {
                double M[9], Lij[9], tr, LLMM, MMMM, Cd;
                c = GI(i, j, k);
                for (a = 0; a < 9; a++) {
                    M[a] = p->DD * (4 * B[a][c] - Af[a][c]);
                    Lij[a] = prod[a][c];
                }
                tr = -(1.0 / 3.0) * (Lij[0] + Lij[4] + Lij[8]);
                Lij[0] += tr; Lij[4] += tr; Lij[8] += tr;
                LLMM = (Lij[0] * M[0] + Lij[1] * M[1] + Lij[2] * M[2]
                      + Lij[3] * M[3] + Lij[4] * M[4] + Lij[5] * M[5]
                      + Lij[6] * M[6] + Lij[7] * M[7] + Lij[8] * M[8]);
                MMMM = (M[0] * M[0] + M[1] * M[1] + M[2] * M[2]
                      + M[3] * M[3] + M[4] * M[4] + M[5] * M[5]
                      + M[6] * M[6] + M[7] * M[7] + M[8] * M[8]);
                Cd = -0.5 * (LLMM / (MMMM + SMALL));
                if (Cd > 0.15) Cd = 0.15;
                else if (Cd < 0.0) Cd = 0.0;
                if (mu_sgs) mu_sgs[c] = U[0][c] * Cd * p->DD * mag[c];
                if (lm) {
                    lm[c] = LLMM;
                    mm[c] = MMMM;
                    fac[c] = U[0][c] * p->DD * mag[c];
                }
            }
    free(u); free(v); free(w); free(uf); free(vf); free(wf); free(mag);
    for (a = 0; a < 9; a++) { free(prod[a]); free(fil[a]); free(A[a]); free(Af[a]); free(B[a]); }
}

void or_dynamic_smagorinsky(const or_params *p, double *const U[5], double *mu_sgs)
{
    dynamic_core(p, U, mu_sgs, NULL, NULL, NULL);
}

void or_dynamic_terms(const or_params *p, double *const U[5], double *lm, double *mm, double *fac)
{
    dynamic_core(p, U, NULL, lm, mm, fac);
}

void or_step(or_params *p, or_state *s)
{
    /* src/layer2.c:141-185 */
    int stage;
    or_check_co_num(p, s);
    or_block_draw(p);
    for (stage = 1; stage <= p->nStages; stage++) {
        double *const *Ucur = s->cur_is_p ? s->Up : s->U;
        or_ghost_fill(p, Ucur);
        or_reconstruct(p, s);
        or_iface_bc(p, s);
        or_fluxes(p, s);
        or_evolution(p, s, p->nStages, stage);
    }
}

/* ---- handle surface for ctypes ---- */
struct or_sim {
    or_params p;
    or_state s;
    double *u_in;
};

or_sim *or_sim_create(int L, int H, int D, int dynamic)
{
    or_sim *m = (or_sim *)calloc(1, sizeof(or_sim));
    int j;
    if (!m) return NULL;
    m->u_in = zalloc((size_t)H);
    for (j = 0; j < H; j++) m->u_in[j] = (j + 1 <= H / 2) ? 76.4 : 200;
    or_params_default_ref(&m->p, L, H, D, m->u_in);
    m->p.dynamic = dynamic;
    or_params_derive(&m->p);
    if (or_state_alloc(&m->p, &m->s) != 0) { or_sim_destroy(m); return NULL; }
    return m;
}

void or_sim_destroy(or_sim *m)
{
    if (!m) return;
    or_state_free(&m->s);
    free(m->u_in);
    free(m);
}

/* src/output.c:88-136 (with real = double): primitives, central
   differences of u = rho u / rho with the 1/(2 delta) factors, vorticity,
   strain magnitude, Q and the constant-coefficient mu_T / mu_L ratio, in the
   reference's association order. */
void or_output_fields(const or_params *p, double *const U[5], double *out)
{
    const int L = p->L, H = p->H, D = p->D;
    const long sj = D + 2, si = (long)(H + 2) * (D + 2);
    int i, j, k;
    double *o = out;
    for (k = 1; k <= D; k++)
        for (j = 1; j <= H; j++)
            for (i = 1; i <= L; i++, o += 9) {
                const long c = i * si + j * sj + k;
                const double *r = U[0], *mu = U[1], *mv = U[2], *mw = U[3];
                double R, Uc, Vc, Wc, P, T, du_dx, du_dy, du_dz, dv_dx, dv_dy, dv_dz, dw_dx, dw_dy, dw_dz;
                double ox, oy, oz, Omega, S12, S13, S23, Strain, Q, muT;
                R = r[c];
                Uc = mu[c] / R;
                Vc = mv[c] / R;
                Wc = mw[c] / R;
                P = (U[4][c] - 0.5 * R * (Uc * Uc + Vc * Vc + Wc * Wc)) * 0.4;
                T = P / (287.10 * R);
                du_dx = (mu[c + si] / r[c + si] - mu[c - si] / r[c - si]) * p->r2dx;
                du_dy = (mu[c + sj] / r[c + sj] - mu[c - sj] / r[c - sj]) * p->r2dy;
                du_dz = (mu[c + 1] / r[c + 1] - mu[c - 1] / r[c - 1]) * p->r2dz;
                dv_dx = (mv[c + si] / r[c + si] - mv[c - si] / r[c - si]) * p->r2dx;
                dv_dy = (mv[c + sj] / r[c + sj] - mv[c - sj] / r[c - sj]) * p->r2dy;
                dv_dz = (mv[c + 1] / r[c + 1] - mv[c - 1] / r[c - 1]) * p->r2dz;
                dw_dx = (mw[c + si] / r[c + si] - mw[c - si] / r[c - si]) * p->r2dx;
                dw_dy = (mw[c + sj] / r[c + sj] - mw[c - sj] / r[c - sj]) * p->r2dy;
                dw_dz = (mw[c + 1] / r[c + 1] - mw[c - 1] / r[c - 1]) * p->r2dz;
                ox = dw_dy - dv_dz;
                oy = du_dz - dw_dx;
                oz = dv_dx - du_dy;
                Omega = sqrt(ox * ox + oy * oy + oz * oz);
                S12 = 0.5 * (du_dy + dv_dx);
                S13 = 0.5 * (du_dz + dw_dx);
                S23 = 0.5 * (dv_dz + dw_dy);
                Strain = sqrt(2 * (du_dx * du_dx + dv_dy * dv_dy + dw_dz * dw_dz +
                                   2 * (S12 * S12 + S13 * S13 + S23 * S23)));
                Q = 0.5 * (Omega * Omega - Strain * Strain);
                muT = R * p->CsDD * Strain / p->mu_L;
                o[0] = R; o[1] = Uc; o[2] = Vc; o[3] = Wc; o[4] = P; o[5] = T;
                o[6] = Omega; o[7] = Q; o[8] = muT;
            }
}

/* src/probes.c:17-40: stations i = LEN/2, 3*LEN/4, LEN (unsigned
   arithmetic), mid-z plane k = DEP/2 + 1 */
void or_probe_values(const or_params *p, double *const U[5], double *out)
{
    const unsigned L = (unsigned)p->L, H = (unsigned)p->H, D = (unsigned)p->D;
    const long sj = D + 2, si = (long)(H + 2) * (D + 2);
    const unsigned st[3] = {L / 2, 3 * L / 4, L};
    const unsigned k = D / 2 + 1;
    unsigned s, f, j;
    for (s = 0; s < 3; s++)
        for (f = 0; f < 2; f++)
            for (j = 1; j <= H; j++) *out++ = U[f][st[s] * si + j * sj + k];
}

#define OR_KEYS(X_) \
    X_(dx) X_(dy) X_(dz) X_(dt) X_(mu_L) X_(Pr_L) X_(Pr_T) X_(Cs) X_(Ua) X_(Va) X_(Wa) \
    X_(maxCoNum) X_(rho_in) X_(p_in) X_(p_out) X_(rdx) X_(rdy) X_(rdz) X_(r2dx) X_(r2dy) \
    X_(r2dz) X_(r4dx) X_(r4dy) X_(r4dz) X_(lambda_L) X_(cp_Pr_T) X_(DD) X_(CsDD) X_(dtX) \
    X_(dtY) X_(dtZ) X_(Ud) X_(Vd) X_(Wd) X_(X)
#define OR_IKEYS(X_) \
    X_(BL_HIG) X_(Ns_min) X_(Nst) X_(nStages) X_(dynamic) X_(z_quirk) X_(counter)

int or_sim_set(or_sim *m, const char *key, double v)
{
#define SETD(n) if (!strcmp(key, #n)) { m->p.n = v; return 0; }
#define SETI(n) if (!strcmp(key, #n)) { m->p.n = (int)v; return 0; }
    OR_KEYS(SETD)
    OR_IKEYS(SETI)
    if (!strcmp(key, "k_min")) { m->p.k_min = (unsigned)v; return 0; }
    if (!strcmp(key, "k_max")) { m->p.k_max = (unsigned)v; return 0; }
    return -1;
#undef SETD
#undef SETI
}

double or_sim_get(or_sim *m, const char *key)
{
#define GETD(n) if (!strcmp(key, #n)) return m->p.n;
#define GETI(n) if (!strcmp(key, #n)) return (double)m->p.n;
    OR_KEYS(GETD)
    OR_IKEYS(GETI)
    if (!strcmp(key, "k_min")) return (double)m->p.k_min;
    if (!strcmp(key, "k_max")) return (double)m->p.k_max;
    if (!strcmp(key, "L")) return m->p.L;
    if (!strcmp(key, "H")) return m->p.H;
    if (!strcmp(key, "D")) return m->p.D;
    return 0.0 / 0.0;
#undef GETD
#undef GETI
}

void or_sim_derive(or_sim *m) { or_params_derive(&m->p); }
double *or_sim_u_in(or_sim *m) { return m->u_in; }

double *or_sim_cell(or_sim *m, int id)
{
    if (id >= 0 && id < 5) return m->s.U[id];
    if (id >= 5 && id < 10) return m->s.Up[id - 5];
    if (id == 10) return m->s.mu_sgs;
    return NULL;
}

double *or_sim_face(or_sim *m, int id)
{
    int f = id % 5;
    switch (id / 5) {
    case 0: return m->s.xL[f]; case 1: return m->s.xR[f];
    case 2: return m->s.yL[f]; case 3: return m->s.yR[f];
    case 4: return m->s.zL[f]; case 5: return m->s.zR[f];
    default: return NULL;
    }
}

void or_sim_ghost(or_sim *m) { or_ghost_fill(&m->p, m->s.cur_is_p ? m->s.Up : m->s.U); }
void or_sim_recon(or_sim *m) { or_reconstruct(&m->p, &m->s); }
void or_sim_iface(or_sim *m) { or_iface_bc(&m->p, &m->s); }
void or_sim_fluxes(or_sim *m) { or_fluxes(&m->p, &m->s); }
void or_sim_evolution(or_sim *m, int n, int st) { or_evolution(&m->p, &m->s, n, st); }
double or_sim_check_co_num(or_sim *m) { return or_check_co_num(&m->p, &m->s); }
void or_sim_block_draw(or_sim *m) { or_block_draw(&m->p); }
void or_sim_dynamic(or_sim *m) { or_dynamic_smagorinsky(&m->p, m->s.U, m->s.mu_sgs); }
void or_sim_dynamic_terms(or_sim *m, double *lm, double *mm, double *fac)
{
    or_dynamic_terms(&m->p, m->s.U, lm, mm, fac);
}
void or_sim_filter(or_sim *m, const double *in, double *out)
{
    const double h[3] = { 1. / 3., 1. / 3., 1. / 3. };
    memset(out, 0, or_ncell(&m->p) * sizeof(double));
    or_filter(&m->p, in, out, h);
}
void or_sim_step(or_sim *m) { or_step(&m->p, &m->s); }
void or_sim_ic_reference(or_sim *m) { or_ic_reference(&m->p, &m->s); }
int or_sim_cur_is_p(or_sim *m) { return m->s.cur_is_p; }
void or_sim_set_cur(or_sim *m, int is_p) { m->s.cur_is_p = is_p; }

void or_sim_output_fields(or_sim *m, double *out) { or_output_fields(&m->p, m->s.U, out); }
void or_sim_probe_values(or_sim *m, double *out) { or_probe_values(&m->p, m->s.U, out); }

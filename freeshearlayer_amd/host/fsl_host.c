/*
 * fsl_host.c -- host side of the per-step update (ANSI C): case input,
 * inflow LCG and disturbance block, initial conditions, backup.dat I/O.
 *
 *   fsl_case_read_ini    src/initialize.c:18-60
 *   fsl_case_read_bin    src-par/initialize.c:26-85
 *   fsl_case_ini_to_bin  src-par/ini2bin/ini2bin.c:11-81
 *   fsl_random           src/helpers.c:13-17
 *   fsl_block_advance    src/layer2.c:152-162
 *   fsl_ic_reference     src/initialize.c:159-187
 *   fsl_backup_write     src/finalize.c:14-34
 *   fsl_backup_read      src/initialize.c:133-158
 */
#include "fsl.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define FSL_PI 3.14159265358979323846

int fsl_case_read_ini(const char *path, fsl_case *c)
{
    /* 22 positional lines; each line is read with fgets(str, 79, ...) and
       converted with atoi / atof, which take the leading number and ignore
       the rest of the line (the item names and comments). */
    char str[80];
    FILE *f;
    int i;
    if (!path || !c) return FSL_ERR_ARG;
    f = fopen(path, "r");
    if (!f) return FSL_ERR_IO;
    memset(c, 0, sizeof *c);
    for (i = 0; i < 22; i++) {
        if (fgets(str, 79, f) == NULL) {
            fclose(f);
            return FSL_ERR_IO;
        }
        switch (i) {
        case 0: c->LEN = (unsigned)atoi(str); break;
        case 1: c->HIG = (unsigned)atoi(str); break;
        case 2: c->DEP = (unsigned)atoi(str); break;
        case 3: c->deltaX = atof(str); break;
        case 4: c->deltaY = atof(str); break;
        case 5: c->deltaZ = atof(str); break;
        case 6: c->deltaT = atof(str); break;
        case 7: c->Answer = (unsigned)atoi(str); break;
        case 8: c->numstep = (unsigned)atoi(str); break;
        case 9: c->mu_L = atof(str); break;
        case 10: c->Pr_L = atof(str); break;
        case 11: c->Pr_T = atof(str); break;
        case 12: c->Cs = atof(str); break;
        case 13: c->BL_HIG = (unsigned)atoi(str); break;
        case 14: c->Ua = atof(str); break;
        case 15: c->Va = atof(str); break;
        case 16: c->Wa = atof(str); break;
        case 17: c->Ns_min = atoi(str); break;
        case 18: c->Nst = atoi(str); break;
        case 19: c->f_step = (unsigned)atoi(str); break;
        case 20: c->nStages = atoi(str); break;
        case 21: c->maxCoNum = atof(str); break;
        default: break;
        }
    }
    fclose(f);
    return FSL_OK;
}

int fsl_case_read_bin(const char *path, fsl_case *c)
{
    /* 22 native float32 values in .ini order; integer items are converted
       from float as the MPI variant's assignments do. */
    float v[22];
    FILE *f;
    if (!path || !c) return FSL_ERR_ARG;
    f = fopen(path, "rb");
    if (!f) return FSL_ERR_IO;
    if (fread(v, sizeof(float), 22, f) != 22) {
        fclose(f);
        return FSL_ERR_IO;
    }
    fclose(f);
    memset(c, 0, sizeof *c);
    c->LEN = (unsigned)v[0];
    c->HIG = (unsigned)v[1];
    c->DEP = (unsigned)v[2];
    c->deltaX = v[3];
    c->deltaY = v[4];
    c->deltaZ = v[5];
    c->deltaT = v[6];
    c->Answer = (unsigned)v[7];
    c->numstep = (unsigned)v[8];
    c->mu_L = v[9];
    c->Pr_L = v[10];
    c->Pr_T = v[11];
    c->Cs = v[12];
    c->BL_HIG = (unsigned)v[13];
    c->Ua = v[14];
    c->Va = v[15];
    c->Wa = v[16];
    c->Ns_min = (int)v[17];
    c->Nst = (int)v[18];
    c->f_step = (unsigned)v[19];
    c->nStages = (int)v[20];
    c->maxCoNum = v[21];
    return FSL_OK;
}

int fsl_case_ini_to_bin(const char *ini_path, const char *bin_path)
{
    char str[80];
    FILE *fi, *fo;
    int i;
    fi = fopen(ini_path, "r");
    if (!fi) return FSL_ERR_IO;
    fo = fopen(bin_path, "wb");
    if (!fo) {
        fclose(fi);
        return FSL_ERR_IO;
    }
    for (i = 0; i < 22; i++) {
        float buf;
        if (fgets(str, 79, fi) == NULL) {
            fclose(fi);
            fclose(fo);
            return FSL_ERR_IO;
        }
        buf = (float)atof(str);
        if (fwrite(&buf, sizeof(float), 1, fo) != 1) {
            fclose(fi);
            fclose(fo);
            return FSL_ERR_IO;
        }
    }
    fclose(fi);
    fclose(fo);
    return FSL_OK;
}

int fsl_case_validate(const fsl_case *c, char *msg, size_t n)
{
    const char *e = NULL;
    if (!c) e = "null case";
    else if (c->LEN < 1 || c->HIG < 2 || c->DEP < 1) e = "LEN >= 1, HIG >= 2, DEP >= 1 required";
    else if (!(c->deltaX > 0) || !(c->deltaY > 0) || !(c->deltaZ > 0)) e = "cell sizes must be positive";
    else if (c->BL_HIG > c->HIG - c->HIG / 2)
        e = "BL_HIG > HIG - HIG/2 would write past the inflow face (src/bounCondOnInterfaces.c:31)";
    else if (c->nStages != 2 && c->nStages != 3) e = "nStages must be 2 or 3";
    else if ((double)(c->LEN + 2) * (c->HIG + 2) * (c->DEP + 2) > 4.0e9) e = "mesh too large";
    if (msg && n) {
        if (e) {
            strncpy(msg, e, n - 1);
            msg[n - 1] = 0;
        } else {
            msg[0] = 0;
        }
    }
    return e ? FSL_ERR_ARG : FSL_OK;
}

double fsl_random(double x)
{
    x = 11. * x + FSL_PI;
    return x - (double)((int)x);
}

double fsl_minmod(double x, double y)
{
    double ay;
    if (x * y <= 0) return 0;
    y *= 4.0;
    ay = (y > 0) ? y : -y;
    if (x > 0) return (x < ay) ? x : ay;
    return (-x < ay) ? x : -ay;
}

void fsl_block_init(fsl_block *b, double seed)
{
    memset(b, 0, sizeof *b);
    b->X = seed;
    b->counter = 0;
}

int fsl_block_advance(fsl_block *b, const fsl_case *c)
{
    /* a block lives counter + 1 steps; the draw consumes six LCG values */
    if (b->counter > 0) {
        b->counter--;
        return 0;
    }
    b->counter = (int)(c->Ns_min + c->Nst * (b->X = fsl_random(b->X)));
    b->k_min = (unsigned)(c->DEP * (b->X = fsl_random(b->X)));
    b->k_max = (unsigned)(b->k_min + c->DEP * (b->X = fsl_random(b->X)));
    if (b->k_max > c->DEP) b->k_max -= c->DEP;
    b->X = fsl_random(b->X);
    b->Ud = c->Ua * (1 - b->X - b->X);
    b->X = fsl_random(b->X);
    b->Vd = c->Va * (1 - b->X - b->X);
    b->X = fsl_random(b->X);
    b->Wd = c->Wa * (1 - b->X - b->X);
    return 1;
}

static size_t ncell(const fsl_case *c) { return (size_t)(c->LEN + 2) * (c->HIG + 2) * (c->DEP + 2); }

static size_t gidx(const fsl_case *c, unsigned i, unsigned j, unsigned k)
{
    return ((size_t)i * (c->HIG + 2) + j) * (c->DEP + 2) + k;
}

void fsl_ic_reference(const fsl_case *c, double *U)
{
    const size_t G = ncell(c);
    unsigned i, j, k;
    memset(U, 0, 5 * G * sizeof(double));
    for (i = 1; i <= c->LEN; i++)
        for (j = 1; j <= c->HIG; j++)
            for (k = 1; k <= c->DEP; k++) {
                const double R = 1.0, P = 100000.;
                const double u = (j <= c->HIG / 2) ? 76.4 : 200;
                const size_t n = gidx(c, i, j, k);
                U[n] = R;
                U[G + n] = R * u;
                U[2 * G + n] = 0.;
                U[3 * G + n] = 0.;
                U[4 * G + n] = P / 0.4 + 0.5 * R * (u * u);
            }
}

void fsl_shear_defaults(fsl_shear_case *s)
{
    s->mach_c = 0.3;
    s->ratio = 0.5;
    s->delta_w0 = 0.0;   /* 0: four cell heights */
    s->reynolds = 500.0;
    s->pert = 0.02;
    s->rho0 = 1.0;
    s->p0 = 100000.;
    s->seed = 20261203ULL;
}

static unsigned long long splitmix64(unsigned long long x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

/* uniform in [-1, 1) from the top 53 bits */
static double unit_pm1(unsigned long long r)
{
    return (double)(r >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

double fsl_ic_shear_layer(const fsl_case *c, const fsl_shear_case *s, double *U, double *u_in)
{
    const size_t G = ncell(c);
    const double c0 = sqrt(1.4 * s->p0 / s->rho0);
    const double dU = 2.0 * s->mach_c * c0;
    const double U1 = dU / (1.0 - s->ratio);
    const double U2 = s->ratio * U1;
    const double Um = 0.5 * (U1 + U2);
    const double dw = (s->delta_w0 > 0) ? s->delta_w0 : 4.0 * c->deltaY;
    unsigned i, j, k;
    memset(U, 0, 5 * G * sizeof(double));
    for (j = 1; j <= c->HIG; j++) {
        const double y = ((double)j - (double)(c->HIG / 2) - 0.5) * c->deltaY;
        const double ub = Um + 0.5 * dU * tanh(2.0 * y / dw);
        if (u_in) u_in[j - 1] = ub;
        for (i = 1; i <= c->LEN; i++)
            for (k = 1; k <= c->DEP; k++) {
                const size_t n = gidx(c, i, j, k);
                double u = ub, v = 0.0, w = 0.0;
                if (fabs(y) < dw && s->pert > 0) {
                    const unsigned long long key = s->seed ^ (3ULL * n);
                    u += s->pert * dU * unit_pm1(splitmix64(key));
                    v += s->pert * dU * unit_pm1(splitmix64(key + 1));
                    w += s->pert * dU * unit_pm1(splitmix64(key + 2));
                }
                U[n] = s->rho0;
                U[G + n] = s->rho0 * u;
                U[2 * G + n] = s->rho0 * v;
                U[3 * G + n] = s->rho0 * w;
                U[4 * G + n] = s->p0 / 0.4 + 0.5 * s->rho0 * (u * u + v * v + w * w);
            }
    }
    return s->rho0 * dU * dw / s->reynolds;
}

/* element size of the backup file: 8 (real = double) or 4 (real = float, as
   the reference ships) */
static int put_real(FILE *f, double v, int rb)
{
    if (rb == 4) {
        float x = (float)v;
        return fwrite(&x, 4, 1, f) == 1;
    }
    return fwrite(&v, 8, 1, f) == 1;
}

static int get_real(FILE *f, double *v, int rb)
{
    if (rb == 4) {
        float x;
        if (fread(&x, 4, 1, f) != 1) return 0;
        *v = x;
        return 1;
    }
    return fread(v, 8, 1, f) == 1;
}

int fsl_backup_write(const char *path, const fsl_case *c, const double *U, double deltaT, unsigned step,
                     int real_bytes)
{
    const size_t G = ncell(c);
    unsigned i, j, k, f;
    FILE *fp;
    if (real_bytes != 4 && real_bytes != 8) return FSL_ERR_ARG;
    fp = fopen(path, "wb");
    if (!fp) return FSL_ERR_IO;
    for (i = 1; i <= c->LEN; i++)
        for (j = 1; j <= c->HIG; j++)
            for (f = 0; f < 5; f++)
                for (k = 1; k <= c->DEP; k++)
                    if (!put_real(fp, U[f * G + gidx(c, i, j, k)], real_bytes)) {
                        fclose(fp);
                        return FSL_ERR_IO;
                    }
    if (!put_real(fp, deltaT, real_bytes) || fwrite(&step, sizeof step, 1, fp) != 1) {
        fclose(fp);
        return FSL_ERR_IO;
    }
    fclose(fp);
    return FSL_OK;
}

int fsl_backup_read(const char *path, const fsl_case *c, double *U, double *deltaT, unsigned *step,
                    int ref_layout, int real_bytes)
{
    const size_t G = ncell(c);
    unsigned i, j, k, f;
    FILE *fp;
    if (real_bytes != 4 && real_bytes != 8) return FSL_ERR_ARG;
    fp = fopen(path, "rb");
    if (!fp) return FSL_ERR_IO;
    for (i = 1; i <= c->LEN; i++)
        for (j = 1; j <= c->HIG; j++)
            for (f = 0; f < 5; f++)
                for (k = 1; k <= c->DEP; k++)
                    if (!get_real(fp, &U[f * G + gidx(c, i, j, k)], real_bytes)) {
                        fclose(fp);
                        return FSL_ERR_IO;
                    }
    if (ref_layout) {
        /* src/initialize.c:153-154 reads step where Finalize wrote deltaT */
        if (fread(step, sizeof *step, 1, fp) != 1) {
            fclose(fp);
            return FSL_ERR_IO;
        }
    } else {
        if (!get_real(fp, deltaT, real_bytes) || fread(step, sizeof *step, 1, fp) != 1) {
            fclose(fp);
            return FSL_ERR_IO;
        }
    }
    fclose(fp);
    return FSL_OK;
}

/* fsl_stats.c -- host side of the learned closure (parameter init and file
   I/O) and the shear-layer statistics computed from the device plane
   profiles (ANSI C). */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "fsl.h"

static unsigned long long mix(unsigned long long x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

/* uniform in [-1, 1) */
static double upm1(unsigned long long r) { return (double)(r >> 11) * (2.0 / 9007199254740992.0) - 1.0; }

void fsl_mlp_init_params(unsigned long long seed, float *params, float *norm)
{
    const int nin = FSL_MLP_NIN, nh = FSL_MLP_NHID;
    const int ow1 = 0, ob1 = nin * nh, ow2 = ob1 + nh, ob2 = ow2 + nh * nh, ow3 = ob2 + nh, ob3 = ow3 + nh;
    const double a1 = sqrt(6.0 / (nin + nh)), a2 = sqrt(6.0 / (nh + nh)), a3 = sqrt(6.0 / (nh + 1));
    unsigned long long s = mix(seed);
    int p;
    memset(params, 0, FSL_MLP_NPARAM * sizeof(float));
    for (p = 0; p < nin * nh; p++) params[ow1 + p] = (float)(a1 * upm1(s = mix(s)));
    for (p = 0; p < nh * nh; p++) params[ow2 + p] = (float)(a2 * upm1(s = mix(s)));
    for (p = 0; p < nh; p++) params[ow3 + p] = (float)(a3 * upm1(s = mix(s)));
    params[ob3] = 0.1f;
    for (p = 0; p < 8; p++) {
        norm[p] = 0.0f;
        norm[8 + p] = 1.0f;
    }
}

static const char kMagic[8] = {'F', 'S', 'L', 'M', 'L', 'P', '0', '1'};
static const char kMagic2[8] = {'F', 'S', 'L', 'M', 'L', 'P', '0', '2'};

int fsl_mlp_file_write_act(const char *path, const float *params, const float *norm, int activation)
{
    const int hdr[4] = {FSL_MLP_NIN, FSL_MLP_NHID, FSL_MLP_NPARAM, activation};
    const int v2 = activation != FSL_MLP_TANH;
    FILE *f;
    int ok;
    if (activation != FSL_MLP_TANH && activation != FSL_MLP_RELU) return FSL_ERR_ARG;
    f = fopen(path, "wb");
    if (!f) return FSL_ERR_IO;
    ok = fwrite(v2 ? kMagic2 : kMagic, 1, 8, f) == 8 && fwrite(hdr, sizeof(int), v2 ? 4 : 3, f) == (size_t)(v2 ? 4 : 3) &&
         fwrite(norm, sizeof(float), 16, f) == 16 &&
         fwrite(params, sizeof(float), FSL_MLP_NPARAM, f) == FSL_MLP_NPARAM;
    if (fclose(f) != 0) ok = 0;
    return ok ? FSL_OK : FSL_ERR_IO;
}

int fsl_mlp_file_write(const char *path, const float *params, const float *norm)
{
    return fsl_mlp_file_write_act(path, params, norm, FSL_MLP_TANH);
}

int fsl_mlp_file_read_act(const char *path, float *params, float *norm, int *activation)
{
    char m[8];
    int hdr[4] = {0, 0, 0, FSL_MLP_TANH}, i, v2;
    FILE *f = fopen(path, "rb");
    if (!f) return FSL_ERR_IO;
    if (fread(m, 1, 8, f) != 8 || (memcmp(m, kMagic, 8) != 0 && memcmp(m, kMagic2, 8) != 0)) {
        fclose(f);
        return FSL_ERR_IO;
    }
    v2 = memcmp(m, kMagic2, 8) == 0;
    if (fread(hdr, sizeof(int), v2 ? 4 : 3, f) != (size_t)(v2 ? 4 : 3) || hdr[0] != FSL_MLP_NIN ||
        hdr[1] != FSL_MLP_NHID || hdr[2] != FSL_MLP_NPARAM ||
        (hdr[3] != FSL_MLP_TANH && hdr[3] != FSL_MLP_RELU) || fread(norm, sizeof(float), 16, f) != 16 ||
        fread(params, sizeof(float), FSL_MLP_NPARAM, f) != FSL_MLP_NPARAM) {
        fclose(f);
        return FSL_ERR_IO;
    }
    fclose(f);
    for (i = 0; i < 16; i++)
        if (!isfinite(norm[i])) return FSL_ERR_NONFINITE;
    for (i = 0; i < FSL_MLP_NPARAM; i++)
        if (!isfinite(params[i])) return FSL_ERR_NONFINITE;
    if (activation) *activation = hdr[3];
    return FSL_OK;
}

int fsl_mlp_file_read(const char *path, float *params, float *norm)
{
    int act = FSL_MLP_TANH;
    const int rc = fsl_mlp_file_read_act(path, params, norm, &act);
    if (rc != FSL_OK) return rc;
    return act == FSL_MLP_TANH ? FSL_OK : FSL_ERR_UNSUPPORTED;
}

int fsl_shear_statistics(const fsl_case *c, const double *prof, double U1, double U2, double rho0,
                         double *dtheta_x, double *dw_x, fsl_shear_stats *out)
{
    const unsigned L = c->LEN, H = c->HIG, D = c->DEP;
    const size_t nrow = (size_t)L * H;
    const double dU = U1 - U2, dy = c->deltaY, vol = c->deltaX * c->deltaY * c->deltaZ;
    const double *sr = prof, *su = prof + nrow, *sv = prof + 2 * nrow, *sw = prof + 3 * nrow;
    const double *suu = prof + 4 * nrow, *svv = prof + 5 * nrow, *sww = prof + 6 * nrow;
    double th_sum = 0, dw_sum = 0, tke = 0;
    unsigned i, j, nhalf = 0;
    if (!(dU > 0) || !(rho0 > 0) || L < 1 || H < 3) return FSL_ERR_ARG;
    for (i = 0; i < L; i++) {
        double th = 0, gmax = 0, dw;
        for (j = 0; j < H; j++) {
            const size_t r = (size_t)i * H + j;
            const double ut = su[r] / sr[r], rb = sr[r] / D;
            th += rb * (U1 - ut) * (ut - U2) * dy;
            tke += 0.5 * ((suu[r] - su[r] * su[r] / sr[r]) + (svv[r] - sv[r] * sv[r] / sr[r]) +
                          (sww[r] - sw[r] * sw[r] / sr[r])) * vol;
            if (j >= 1 && j + 1 < H) {
                const double g = fabs(su[r + 1] / sr[r + 1] - su[r - 1] / sr[r - 1]) / (2 * dy);
                if (g > gmax) gmax = g;
            }
        }
        th /= rho0 * dU * dU;
        dw = gmax > 0 ? dU / gmax : 0;
        if (dtheta_x) dtheta_x[i] = th;
        if (dw_x) dw_x[i] = dw;
        if (i + 1 > L / 2) {
            th_sum += th;
            dw_sum += dw;
            nhalf++;
        }
        if (i + 1 == L) out->delta_theta_out = th;
    }
    out->delta_theta = th_sum / nhalf;
    out->delta_w = dw_sum / nhalf;
    out->tke = tke;
    return FSL_OK;
}

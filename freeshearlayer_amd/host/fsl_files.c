/*
 * fsl_files.c -- the reference's output files from the device state
 * (ANSI C host side; the fields and probe columns come from
 * csrc/fsl_output.hip).
 *
 *   <step>.plt    Tecplot POINT frame of src/output.c:16-143
 *   probes.dat    appended record per step of src/probes.c:13-43
 *   probe{1,2,3}.dat  Favre-averaged profiles of favre-average-src/layer2_p.c
 *
 * The text is formatted with the reference's own printf conversions from
 * the same FP64 values, so for the same state the files are byte-identical.
 */
#include "fsl.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NCOL 9

static int plt_header(FILE *f, const fsl_case *c)
{
    /* src/output.c:71-86 */
    fprintf(f, "title     = \" 3-D compressible case \"\n");
    fprintf(f, "variables = \" x \"\n");
    fprintf(f, "\"y\"\n");
    fprintf(f, "\"z\"\n");
    fprintf(f, "\"rho\"\n");
    fprintf(f, "\"u\"\n");
    fprintf(f, "\"v\"\n");
    fprintf(f, "\"w\"\n");
    fprintf(f, "\"p\"\n");
    fprintf(f, "\"T\"\n");
    fprintf(f, "\"Vort. mag.\"\n");
    fprintf(f, "\"Q-criteria.\"\n");
    fprintf(f, "\"muSgs-mu-ratio\"\n");
    fprintf(f, "zone t=\" \"\n");
    return fprintf(f, "i=%d, j=%d, k=%d, f=point\n", c->LEN, c->HIG, c->DEP) < 0 ? FSL_ERR_IO : FSL_OK;
}

static int plt_rows(FILE *f, const fsl_case *c, int k0, int nk, const double *fields)
{
    /* src/output.c:88-136: cell centres from the (1-based, unsigned) indices,
       then the nine computed columns */
    unsigned i, j, k;
    const double *o = fields;
    for (k = (unsigned)k0; k < (unsigned)(k0 + nk); k++)
        for (j = 1; j <= c->HIG; j++)
            for (i = 1; i <= c->LEN; i++, o += NCOL) {
                const double xc = (i - 1) * c->deltaX + 0.5 * c->deltaX;
                const double yc = (j - 1) * c->deltaY + 0.5 * c->deltaY;
                const double zc = (k - 1) * c->deltaZ + 0.5 * c->deltaZ;
                if (fprintf(f, "%g %g %g %g %g %g %g %g %g %g %g %g\n", xc, yc, zc, o[0], o[1], o[2], o[3], o[4],
                            o[5], o[6], o[7], o[8]) < 0)
                    return FSL_ERR_IO;
            }
    return FSL_OK;
}

int fsl_plt_write_fields(const char *path, const fsl_case *c, const double *fields)
{
    FILE *f = fopen(path, "w");
    int rc;
    if (!f) return FSL_ERR_IO;
    rc = plt_header(f, c);
    if (rc == FSL_OK) rc = plt_rows(f, c, 1, (int)c->DEP, fields);
    if (fclose(f) != 0 && rc == FSL_OK) rc = FSL_ERR_IO;
    return rc;
}

int fsl_write_plt(fsl_ctx *ctx, const char *path)
{
    fsl_case c;
    char name[64];
    FILE *f;
    double *buf;
    int rc = FSL_OK, k0, nk;
    size_t plane;
    fsl_get_case(ctx, &c);
    if (!path) {   /* src/output.c:66: "<step>.plt" in the working directory */
        snprintf(name, sizeof name, "%u.plt", fsl_get_step(ctx));
        path = name;
    }
    plane = (size_t)c.LEN * c.HIG;
    /* stream the frame in slabs of k planes of about 32 MB of fields */
    nk = (int)(((size_t)4 << 20) / (plane * NCOL) ? ((size_t)4 << 20) / (plane * NCOL) : 1);
    if (nk > (int)c.DEP) nk = (int)c.DEP;
    buf = (double *)malloc((size_t)nk * plane * NCOL * sizeof(double));
    if (!buf) return FSL_ERR_NOMEM;
    f = fopen(path, "w");
    if (!f) {
        free(buf);
        return FSL_ERR_IO;
    }
    rc = plt_header(f, &c);
    for (k0 = 1; rc == FSL_OK && k0 <= (int)c.DEP; k0 += nk) {
        const int n = (k0 + nk - 1 <= (int)c.DEP) ? nk : (int)c.DEP - k0 + 1;
        rc = fsl_output_fields(ctx, k0, n, buf);
        if (rc == FSL_OK) rc = plt_rows(f, &c, k0, n, buf);
    }
    if (fclose(f) != 0 && rc == FSL_OK) rc = FSL_ERR_IO;
    free(buf);
    return rc;
}

int fsl_probes_append(const char *path, unsigned step, unsigned HIG, const double *vals, int real_bytes)
{
    /* src/probes.c:17-42: the step (unsigned), then rho and rho u along j at
       each of the three stations, as `real` */
    FILE *f;
    unsigned q;
    float *tmp = 0;
    if (real_bytes != 4 && real_bytes != 8) return FSL_ERR_ARG;
    f = fopen(path ? path : "probes.dat", "ab");
    if (!f) return FSL_ERR_IO;
    if (fwrite(&step, sizeof step, 1, f) != 1) {
        fclose(f);
        return FSL_ERR_IO;
    }
    if (real_bytes == 4) {
        tmp = (float *)malloc(HIG * sizeof(float));
        if (!tmp) {
            fclose(f);
            return FSL_ERR_NOMEM;
        }
    }
    for (q = 0; q < 6; q++) {
        size_t w;
        if (real_bytes == 8) {
            w = fwrite(vals + (size_t)q * HIG, sizeof(double) * HIG, 1, f);
        } else {
            unsigned j;
            for (j = 0; j < HIG; j++) tmp[j] = (float)vals[(size_t)q * HIG + j];
            w = fwrite(tmp, sizeof(float) * HIG, 1, f);
        }
        if (w != 1 && HIG > 0) {
            free(tmp);
            fclose(f);
            return FSL_ERR_IO;
        }
    }
    free(tmp);
    return fclose(f) == 0 ? FSL_OK : FSL_ERR_IO;
}

int fsl_write_probes(fsl_ctx *ctx, const char *path, int real_bytes)
{
    fsl_case c;
    double *vals;
    int rc;
    fsl_get_case(ctx, &c);
    vals = (double *)malloc(6 * (size_t)c.HIG * sizeof(double));
    if (!vals) return FSL_ERR_NOMEM;
    rc = fsl_probe_values(ctx, vals);
    if (rc == FSL_OK) rc = fsl_probes_append(path, fsl_get_step(ctx), c.HIG, vals, real_bytes);
    free(vals);
    return rc;
}

int fsl_favre_average(const char *path, unsigned HIG, unsigned step_start, unsigned step_end, int real_bytes,
                      double *ru, unsigned *nrec)
{
    /* favre-average-src/layer2_p.c:50-113: sum rho and rho u over the
       records with step_start <= step <= step_end (in float, as the
       post-processor does for the float layout; in double for the 8-byte
       layout), stop at the first record past step_end, then u~ = sum(rho u) /
       sum(rho) per station and row.  The post-processor keeps reading at the
       end of the file (its loop has no end-of-file test); here the end of
       the file ends the sums. */
    FILE *f;
    unsigned step, q, j, used = 0;
    size_t rec = (size_t)6 * HIG * (size_t)real_bytes;
    unsigned char *buf;
    float *sf = 0;
    double *sd = 0;
    if (real_bytes != 4 && real_bytes != 8) return FSL_ERR_ARG;
    f = fopen(path ? path : "probes.dat", "rb");
    if (!f) return FSL_ERR_IO;
    buf = (unsigned char *)malloc(rec ? rec : 1);
    sf = (float *)calloc(6 * (size_t)HIG + 1, sizeof(float));
    sd = (double *)calloc(6 * (size_t)HIG + 1, sizeof(double));
    if (!buf || !sf || !sd) {
        free(buf); free(sf); free(sd);
        fclose(f);
        return FSL_ERR_NOMEM;
    }
    while (fread(&step, sizeof step, 1, f) == 1) {
        if (step > step_end) break;
        if (fread(buf, 1, rec, f) != rec) break;
        if (step < step_start) continue;
        for (q = 0; q < 6; q++)
            for (j = 0; j < HIG; j++) {
                const size_t e = (size_t)q * HIG + j;
                if (real_bytes == 4) {
                    float v;
                    memcpy(&v, buf + e * 4, 4);
                    sf[e] += v;
                } else {
                    double v;
                    memcpy(&v, buf + e * 8, 8);
                    sd[e] += v;
                }
            }
        used++;
    }
    fclose(f);
    /* record order per station: rho (q = 2s), rho u (q = 2s + 1) */
    for (q = 0; q < 3; q++)
        for (j = 0; j < HIG; j++) {
            const size_t r = (size_t)(2 * q) * HIG + j, m = (size_t)(2 * q + 1) * HIG + j;
            if (real_bytes == 4) {
                float v = sf[m];
                v /= sf[r];
                ru[(size_t)q * HIG + j] = v;
            } else {
                ru[(size_t)q * HIG + j] = sd[m] / sd[r];
            }
        }
    if (nrec) *nrec = used;
    free(buf); free(sf); free(sd);
    return FSL_OK;
}

int fsl_favre_write_profiles(const char *dir, unsigned HIG, const double *ru)
{
    /* favre-average-src/layer2_p.c:115-147: probe1.dat .. probe3.dat */
    char path[4096];
    unsigned q, j;
    for (q = 0; q < 3; q++) {
        FILE *f;
        snprintf(path, sizeof path, "%s%sprobe%u.dat", dir ? dir : "", (dir && *dir) ? "/" : "", q + 1);
        f = fopen(path, "w");
        if (!f) return FSL_ERR_IO;
        fprintf(f, "%d\n1\n", HIG + 2);
        fprintf(f, "%f %3.2f \n", -0.00001, 0.);
        for (j = 0; j < HIG; j++) fprintf(f, "%d %3.2f \n", j, ru[(size_t)q * HIG + j]);
        fprintf(f, "%f %3.2f \n", HIG - 1 + 0.00001, 210.);
        if (fclose(f) != 0) return FSL_ERR_IO;
    }
    return FSL_OK;
}

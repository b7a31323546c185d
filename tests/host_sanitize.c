/* Driver of the ANSI-C host layer for the sanitizer test (tests/test_host_sanitizers.py):
   case files (.ini, .bin, validation), the LCG and block draw, both initial
   conditions, the backup writer and reader (FP64 and FP32 layouts), the
   network parameter files and the shear-layer statistics, on four meshes.
   argv[1]: directory for the files it writes. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "fsl.h"

static int check(int rc, const char* what) { if (rc) { printf("FAIL %s rc=%d\n", what, rc); exit(1); } return rc; }

static void run(const char* dir, const char* name, int L, int H, int D)
{
    char ini[512];
    snprintf(ini, sizeof ini, "%s/%s", dir, name);
    fsl_case c; memset(&c, 0, sizeof c);
    FILE* f = fopen(ini, "w");
    fprintf(f, "%d LEN\n%d HIG\n%d DEP\n1.6666667 dx\n1. dy\n1.5 dz\n0.0007 dt\n0 Answer\n3000 numstep\n0.253358 mu_L\n0.72 Pr_L\n0.8 Pr_T\n0.01 Cs\n%d BL_HIG\n35. Ua\n35. Va\n35. Wa\n10 Nst_min\n50 Nst\n100000 f_step\n3 nStages\n0.1 maxCoNum\n",
            L, H, D, H - H / 2 < 5 ? H - H / 2 : 5);
    fclose(f);
    check(fsl_case_read_ini(ini, &c), "read_ini");
    char msg[256];
    fsl_case_validate(&c, msg, sizeof msg);
    char bin[600]; snprintf(bin, sizeof bin, "%s.bin", ini);
    check(fsl_case_ini_to_bin(ini, bin), "ini_to_bin");
    fsl_case cb; memset(&cb, 0, sizeof cb);
    check(fsl_case_read_bin(bin, &cb), "read_bin");
    fsl_block b; fsl_block_init(&b, 0.317391239817);
    for (int s = 0; s < 500; s++) fsl_block_advance(&b, &c);
    double x = 0.3; for (int s = 0; s < 1000; s++) x = fsl_random(x);
    (void)fsl_minmod(1.0, -2.0);
    const size_t G = (size_t)(L + 2) * (H + 2) * (D + 2);
    double* U = calloc(5 * G, sizeof(double));
    double* V = calloc(5 * G, sizeof(double));
    double* u_in = calloc(H, sizeof(double));
    fsl_ic_reference(&c, U);
    fsl_shear_case sh; fsl_shear_defaults(&sh);
    double mu = fsl_ic_shear_layer(&c, &sh, U, u_in);
    char bk[600]; snprintf(bk, sizeof bk, "%s.backup", ini);
    check(fsl_backup_write(bk, &c, U, 1e-3, 7, 8), "backup_write");
    double dt; unsigned step;
    check(fsl_backup_read(bk, &c, V, &dt, &step, 0, 8), "backup_read");
    check(fsl_backup_write(bk, &c, U, 1e-3, 7, 4), "backup_write4");
    check(fsl_backup_read(bk, &c, V, &dt, &step, 1, 4), "backup_read4");
    float* p = calloc(FSL_MLP_NPARAM, sizeof(float)); float nrm[16];
    fsl_mlp_init_params(1, p, nrm);
    char mf[600]; snprintf(mf, sizeof mf, "%s.mlp", ini);
    check(fsl_mlp_file_write(mf, p, nrm), "mlp_write");
    check(fsl_mlp_file_read(mf, p, nrm), "mlp_read");
    /* plane profiles: 7 sums per (i, j) row of the interior */
    double* prof = calloc((size_t)7 * L * H, sizeof(double));
    for (size_t r = 0; r < (size_t)L * H; r++) {
        prof[r] = D; prof[L * H + r] = 300. * D; prof[2 * L * H + r] = 1.; prof[3 * L * H + r] = 0.;
        prof[4 * L * H + r] = 9e4 * D; prof[5 * L * H + r] = 1. * D; prof[6 * L * H + r] = 1. * D;
    }
    double* dth = calloc(L + 2, sizeof(double)); double* dw = calloc(L + 2, sizeof(double));
    fsl_shear_stats st;
    fsl_shear_statistics(&c, prof, 449., 224.5, 1.0, dth, dw, &st);
    printf("mesh %dx%dx%d ok (mu %.3g, dtheta %.3g)\n", L, H, D, mu, st.delta_theta);
    free(U); free(V); free(u_in); free(p); free(prof); free(dth); free(dw);
}

int main(int argc, char** argv)
{
    const char* dir = argc > 1 ? argv[1] : ".";
    run(dir, "a.ini", 16, 12, 8);
    run(dir, "b.ini", 5, 10, 3);
    run(dir, "c.ini", 40, 10, 1);
    run(dir, "d.ini", 9, 44, 40);
    return 0;
}

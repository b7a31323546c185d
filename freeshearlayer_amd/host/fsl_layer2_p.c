/*
 * fsl_layer2_p.c -- the Favre-averaging post-processor of the reference
 * (favre-average-src/layer2_p.c:27-152) over a probes.dat written by
 * fsl_layer2 or by the reference: reads stepStart, stepEnd and HIG from
 * stdin as the reference does, writes probe1.dat .. probe3.dat.
 *
 *   fsl_layer2_p [--real8] [probes.dat] < "stepStart stepEnd HIG"
 *
 * The default record layout is the float one the reference's post-processor
 * reads (src/ as shipped, real = float); --real8 reads 8-byte reals.
 */
#include "fsl.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int main(int argc, char **argv)
{
    unsigned HIG = 0, start = 0, end = 0, nrec = 0;
    int real_bytes = 4, a, rc;
    const char *path = "probes.dat";
    double *ru;
    for (a = 1; a < argc; a++) {
        if (!strcmp(argv[a], "--real8")) real_bytes = 8;
        else path = argv[a];
    }
    printf("stepStart = ");
    if (scanf("%u", &start) != 1) return 2;
    printf("stepEnd = ");
    if (scanf("%u", &end) != 1) return 2;
    printf("HIG = ");
    if (scanf("%u", &HIG) != 1 || HIG == 0) return 2;
    ru = (double *)malloc(3 * (size_t)HIG * sizeof(double));
    if (!ru) return 1;
    rc = fsl_favre_average(path, HIG, start, end, real_bytes, ru, &nrec);
    if (rc != FSL_OK) {
        printf("Cannot open file \"%s\"\n", path);
        free(ru);
        return 1;
    }
    printf("\n%u records averaged\n", nrec);
    rc = fsl_favre_write_profiles("", HIG, ru);
    free(ru);
    return rc == FSL_OK ? 0 : 1;
}

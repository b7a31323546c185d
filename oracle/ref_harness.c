/*
 * ref_harness.c -- drives the compiled reference (nikola-m/freeShearLayer src/)
 * so the CPU restatement in fsl_oracle.c can be checked against it.
 *
 * TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles the reference's own
 * sources where they lie (/root/reference/src) with
 *   -DTYPE_H -Dreal=double      real = double (the shipped typedef is float)
 *   -Dmain=fslref_layer2_main   layer2.c supplies the globals, not the entry
 *   -Wl,--wrap=malloc           Array3D rows zero-initialised (helpers.c:63-73
 *                               leaves them uninitialised; the dynamic model
 *                               reads never-written cells)
 * and, for the constant-Cs variant, a forced include (ref_const_cs.h) that
 * sets DynamicSmagorinskySGS to 0 after the reference's def.h.  The output goes
 * only to oracle/_ref/.  Nothing of the reference is copied into this tree.
 *
 * The harness replays main()'s per-step order (src/layer2.c:141-185),
 * exposes the reference's global arrays to Python through ctypes, and runs
 * the reference's own main (fslref_main) for the output-file fixtures.
 */
#include <fcntl.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define API __attribute__((visibility("default")))

/* globals defined by src/layer2.c:34-101 (with real = double) */
extern double ***U1_, ***U2_, ***U3_, ***U4_, ***U5_;
extern double ***U1, ***U2, ***U3, ***U4, ***U5;
extern double ***U1p, ***U2p, ***U3p, ***U4p, ***U5p;
extern double ***xU1, ***xU2, ***xU3, ***xU4, ***xU5;
extern double ***U1x, ***U2x, ***U3x, ***U4x, ***U5x;
extern double ***yU1, ***yU2, ***yU3, ***yU4, ***yU5;
extern double ***U1y, ***U2y, ***U3y, ***U4y, ***U5y;
extern double ***zU1, ***zU2, ***zU3, ***zU4, ***zU5;
extern double ***U1z, ***U2z, ***U3z, ***U4z, ***U5z;
extern double ***mu_SGS;
extern double deltaX, deltaY, deltaZ, deltaT, deltaT_X, deltaT_Y, deltaT_Z;
extern double _deltaX, _deltaY, _deltaZ, _2deltaX, _2deltaY, _2deltaZ, _4deltaX, _4deltaY, _4deltaZ;
extern double mu_L, lambda_L, Pr_L, mu_T, Pr_T, cp_Pr_T, mu_E, lambda_E, Cs, CsDD, DD;
extern double maxCoNum;
extern int nStages;
extern unsigned step, numstep, LEN, HIG, DEP, LENN, HIGG, DEPP;
extern unsigned BL_HIG, k_min, k_max;
extern double Ud, Vd, Wd, Ua, Va, Wa;
extern int Ns_min, Nst;
extern double X;

/* reference entry points (declared in the src/ headers; several are empty) */
void Initialize(void);
void BounCondInGhostCells(void);
void Reconstruction(void);
void BounCondOnInterfaces(void);
void Fluxes(void);
void Evolution(int numStages, int Stage);
void checkCoNum(void);
double Random(double x);
double minmod(double x, double y);
double ***filter_(double ***U, double h[3]);
void DynamicSmagorinsky(double ***rho, double ***ru, double ***rv, double ***rw, double ***mu);
double ***Array3D(int columns, int rows, int floors);
void free3D(double ***arr, int columns, int rows);

/* Array3D rows are malloc'd; route them to calloc (see header comment). */
void *__real_malloc(size_t n);
void *__wrap_malloc(size_t n) { return calloc(1, n ? n : 1); }

static int g_counter = 0;
static int g_quiet = 1;
static int g_saved_fd = -1;

static void mute(void)
{
    int fd;
    if (!g_quiet) return;
    fflush(stdout);
    g_saved_fd = dup(1);
    fd = open("/dev/null", O_WRONLY);
    if (fd >= 0) { dup2(fd, 1); close(fd); }
}

static void unmute(void)
{
    if (!g_quiet || g_saved_fd < 0) return;
    fflush(stdout);
    dup2(g_saved_fd, 1);
    close(g_saved_fd);
    g_saved_fd = -1;
}

API void fslref_set_quiet(int q) { g_quiet = q; }

/* Calls the reference's Initialize() with dir as working directory (it reads
   layer2.ini there, and backup.dat when Answer = 1). */
API int fslref_initialize(const char *dir)
{
    char cwd[4096];
    if (!getcwd(cwd, sizeof cwd)) return -1;
    if (chdir(dir) != 0) return -2;
    mute();
    Initialize();
    unmute();
    if (chdir(cwd) != 0) return -3;
    g_counter = 0;
    return 0;
}

static double ***cell_array(int id)
{
    switch (id) {
    case 0: return U1;  case 1: return U2;  case 2: return U3;  case 3: return U4;  case 4: return U5;
    case 5: return U1p; case 6: return U2p; case 7: return U3p; case 8: return U4p; case 9: return U5p;
    case 10: return mu_SGS;
    case 11: return U1_; case 12: return U2_; case 13: return U3_; case 14: return U4_; case 15: return U5_;
    default: return NULL;
    }
}

/* face arrays: 0..4 xU*, 5..9 U*x, 10..14 yU*, 15..19 U*y, 20..24 zU*, 25..29 U*z */
static double ***face_array(int id, unsigned *n0, unsigned *n1, unsigned *n2)
{
    static double ****tab[30];
    int dir = id / 10;
    if (id < 0 || id >= 30) return NULL;
    tab[0] = &xU1; tab[1] = &xU2; tab[2] = &xU3; tab[3] = &xU4; tab[4] = &xU5;
    tab[5] = &U1x; tab[6] = &U2x; tab[7] = &U3x; tab[8] = &U4x; tab[9] = &U5x;
    tab[10] = &yU1; tab[11] = &yU2; tab[12] = &yU3; tab[13] = &yU4; tab[14] = &yU5;
    tab[15] = &U1y; tab[16] = &U2y; tab[17] = &U3y; tab[18] = &U4y; tab[19] = &U5y;
    tab[20] = &zU1; tab[21] = &zU2; tab[22] = &zU3; tab[23] = &zU4; tab[24] = &zU5;
    tab[25] = &U1z; tab[26] = &U2z; tab[27] = &U3z; tab[28] = &U4z; tab[29] = &U5z;
    *n0 = LEN + (dir == 0); *n1 = HIG + (dir == 1); *n2 = DEP + (dir == 2);
    return *tab[id];
}

API int fslref_get_cell(int id, double *out)
{
    double ***a = cell_array(id);
    unsigned i, j;
    if (!a) return -1;
    for (i = 0; i < LEN + 2; i++)
        for (j = 0; j < HIG + 2; j++)
            memcpy(out + ((size_t)i * (HIG + 2) + j) * (DEP + 2), a[i][j], (DEP + 2) * sizeof(double));
    return 0;
}

API int fslref_set_cell(int id, const double *in)
{
    double ***a = cell_array(id);
    unsigned i, j;
    if (!a) return -1;
    for (i = 0; i < LEN + 2; i++)
        for (j = 0; j < HIG + 2; j++)
            memcpy(a[i][j], in + ((size_t)i * (HIG + 2) + j) * (DEP + 2), (DEP + 2) * sizeof(double));
    return 0;
}

API int fslref_get_face(int id, double *out)
{
    unsigned n0, n1, n2, i, j;
    double ***a = face_array(id, &n0, &n1, &n2);
    if (!a) return -1;
    for (i = 0; i < n0; i++)
        for (j = 0; j < n1; j++)
            memcpy(out + ((size_t)i * n1 + j) * n2, a[i][j], n2 * sizeof(double));
    return 0;
}

API int fslref_set_face(int id, const double *in)
{
    unsigned n0, n1, n2, i, j;
    double ***a = face_array(id, &n0, &n1, &n2);
    if (!a) return -1;
    for (i = 0; i < n0; i++)
        for (j = 0; j < n1; j++)
            memcpy(a[i][j], in + ((size_t)i * n1 + j) * n2, n2 * sizeof(double));
    return 0;
}

/* U_ points at U (0) or at Up (1) */
API int fslref_cur_is_p(void) { return U1_ == U1p; }

API void fslref_set_cur(int is_p)
{
    if (is_p) { U1_ = U1p; U2_ = U2p; U3_ = U3p; U4_ = U4p; U5_ = U5p; }
    else      { U1_ = U1;  U2_ = U2;  U3_ = U3;  U4_ = U4;  U5_ = U5; }
}

/* scalars: 0 deltaT 1 deltaT_X 2 deltaT_Y 3 deltaT_Z 4 k_min 5 k_max 6 Ud 7 Vd
   8 Wd 9 X 10 DD 11 CsDD 12 lambda_L 13 cp_Pr_T 14 counter 15 step 16 nStages
   17 maxCoNum 18 _2deltaX 19 _4deltaX 20 _deltaX 21 mu_L */
API void fslref_get_scalars(double *o)
{
    o[0] = deltaT; o[1] = deltaT_X; o[2] = deltaT_Y; o[3] = deltaT_Z;
    o[4] = k_min; o[5] = k_max; o[6] = Ud; o[7] = Vd; o[8] = Wd; o[9] = X;
    o[10] = DD; o[11] = CsDD; o[12] = lambda_L; o[13] = cp_Pr_T;
    o[14] = g_counter; o[15] = step; o[16] = nStages; o[17] = maxCoNum;
    o[18] = _2deltaX; o[19] = _4deltaX; o[20] = _deltaX; o[21] = mu_L;
}

API void fslref_set_time(double dt, double x, int counter)
{
    deltaT = dt;
    deltaT_X = deltaT / deltaX; deltaT_Y = deltaT / deltaY; deltaT_Z = deltaT / deltaZ;
    X = x;
    g_counter = counter;
}

API void fslref_set_block(unsigned kmin, unsigned kmax, double ud, double vd, double wd)
{
    k_min = kmin; k_max = kmax; Ud = ud; Vd = vd; Wd = wd;
}

API void fslref_set_nstages(int n) { nStages = n; }

API void fslref_ghost(void)    { BounCondInGhostCells(); }
API void fslref_recon(void)    { Reconstruction(); }
API void fslref_iface(void)    { BounCondOnInterfaces(); }
API void fslref_fluxes(void)   { Fluxes(); }
API void fslref_evolution(int numStages, int stage) { Evolution(numStages, stage); }
API void fslref_check_co_num(void) { mute(); checkCoNum(); unmute(); }
API double fslref_random(double x) { return Random(x); }
API double fslref_minmod(double x, double y) { return minmod(x, y); }

API void fslref_dynamic(void) { DynamicSmagorinsky(U1, U2, U3, U4, mu_SGS); }

/* filter_ with the box kernel on a caller-given cell array */
API int fslref_filter(const double *in, double *out)
{
    double h[3] = { 1. / 3., 1. / 3., 1. / 3. };
    double ***a = Array3D(LEN + 2, HIG + 2, DEP + 2), ***r;
    unsigned i, j;
    for (i = 0; i < LEN + 2; i++)
        for (j = 0; j < HIG + 2; j++)
            memcpy(a[i][j], in + ((size_t)i * (HIG + 2) + j) * (DEP + 2), (DEP + 2) * sizeof(double));
    r = filter_(a, h);
    for (i = 0; i < LEN + 2; i++)
        for (j = 0; j < HIG + 2; j++)
            memcpy(out + ((size_t)i * (HIG + 2) + j) * (DEP + 2), r[i][j], (DEP + 2) * sizeof(double));
    free3D(a, LEN + 2, HIG + 2);
    free3D(r, LEN + 2, HIG + 2);
    return 0;
}

/* disturbance-block draw of src/layer2.c:152-162 (main's own code) */
API void fslref_block_draw(void)
{
    if (g_counter > 0) { g_counter--; return; }
    g_counter = (int)(Ns_min + Nst * (X = Random(X)));
    k_min = (unsigned)(DEP * (X = Random(X)));
    k_max = (unsigned)(k_min + DEP * (X = Random(X)));
    if (k_max > DEP) k_max -= DEP;
    X = Random(X); Ud = Ua * (1 - X - X);
    X = Random(X); Vd = Va * (1 - X - X);
    X = Random(X); Wd = Wa * (1 - X - X);
}

/* The reference's own main (src/layer2.c:117-195, renamed by the build):
   with dir as working directory (layer2.ini there; probes.dat, <step>.plt
   and backup.dat written there) and the LCG seed read from seed_path as its
   stdin.  For a fresh process: main's globals are not reset afterwards. */
int fslref_layer2_main(void);

API int fslref_main(const char *dir, const char *seed_path)
{
    char cwd[4096];
    int rc;
    if (!getcwd(cwd, sizeof cwd)) return -1;
    if (!freopen(seed_path, "r", stdin)) return -2;
    if (chdir(dir) != 0) return -3;
    mute();
    rc = fslref_layer2_main();
    unmute();
    if (chdir(cwd) != 0) return -4;
    return rc;
}

/* one time step in main()'s order, src/layer2.c:141-185 (no probes/output) */
API void fslref_step(void)
{
    int stage;
    fslref_check_co_num();
    ++step;
    fslref_block_draw();
    for (stage = 1; stage <= nStages; stage++) {
        BounCondInGhostCells();
        Reconstruction();
        BounCondOnInterfaces();
        mute();
        Fluxes();
        unmute();
        Evolution(nStages, stage);
    }
}

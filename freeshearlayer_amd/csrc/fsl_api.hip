// fsl_api.hip -- device context and the C-ABI of include/fsl.h.
//
// The context owns device-resident SoA state (U, Up and, for the fused path,
// a third stage buffer), the face arrays of the unfused path (allocated on
// first use), mu_SGS and the dynamic-model scratch, one HIP stream, and the
// per-step scalars that the reference keeps in globals (src/global.h).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <new>
#include <vector>

#include "fsl_ctx.hpp"

using namespace fsl;

namespace fsl {
// fsl_fused.hip
int fused_stage(fsl_ctx* ctx, int stage);
}

namespace fsl {

Bnd bnd_now(fsl_ctx* c)
{
    Bnd b = c->bnd;
    b.k_min = c->blk.k_min;
    b.k_max = c->blk.k_max;
    b.Ud = c->blk.Ud;
    b.Vd = c->blk.Vd;
    b.Wd = c->blk.Wd;
    return b;
}

const char* kNames[K_NKERNELS] = {
    "ghost_fill", "recon", "iface_bc", "flux_x", "flux_y", "flux_z", "rk_update", "cfl_reduce",
    "sgs_filter_a", "filter_y", "filter_z", "filter_x", "sgs_strain", "sgs_filter_b", "fused_x",
    "fused_y", "fused_z", "sums", "copy", "fused_stage", "mlp_mu", "mlp_train", "plane_profiles"};

int set_err(fsl_ctx* c, int code, const char* fmt, ...)
{
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof c->err, fmt, ap);
        va_end(ap);
    }
    return code;
}

hipEvent_t pool_get(fsl_ctx* c)
{
    if (!c->pool.empty()) {
        hipEvent_t e = c->pool.back();
        c->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void prof_drain(fsl_ctx* c)
{
    if (c->pending.empty()) return;
    (void)hipEventSynchronize(c->pending.back().b);
    for (auto& r : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            c->prof_ms[r.id] += ms;
            c->prof_cnt[r.id] += 1;
        }
        c->pool.push_back(r.a);
        c->pool.push_back(r.b);
    }
    c->pending.clear();
}

void prof_mark(void* arg, int id, int end)
{
    fsl_ctx* c = (fsl_ctx*)arg;
    if (!c->prof) return;
    hipEvent_t e = pool_get(c);
    if (!e) return;
    (void)hipEventRecord(e, c->stream);
    if (!end) {
        c->prof_open_id = id;
        c->prof_open_ev = e;
    } else {
        c->pending.push_back({c->prof_open_id, c->prof_open_ev, e});
        if (c->pending.size() >= 4096) prof_drain(c);
    }
}

State5 st5(double* base, long long G)
{
    State5 s;
    for (int f = 0; f < 5; f++) s.f[f] = base + f * G;
    return s;
}

CState5 cst5(const double* base, long long G)
{
    CState5 s;
    for (int f = 0; f < 5; f++) s.f[f] = base + f * G;
    return s;
}

double* cur_state(fsl_ctx* c) { return c->cur_is_p ? c->Up : c->U; }

int ensure_faces(fsl_ctx* c)
{
    if (c->faces) return FSL_OK;
    const size_t n = 10 * (c->nface[0] + c->nface[1] + c->nface[2]);
    HIPCHK(c, hipMalloc(&c->faces, n * sizeof(double)));
    HIPCHK(c, hipMemsetAsync(c->faces, 0, n * sizeof(double), c->stream));
    c->bytes += n * sizeof(double);
    return FSL_OK;
}

Faces faces_of(fsl_ctx* c)
{
    Faces F;
    size_t off = 0;
    for (int d = 0; d < 3; d++)
        for (int a = 0; a < 10; a++) {
            F.a[d * 10 + a] = c->faces + off;
            off += c->nface[d];
        }
    return F;
}

const SgsAvg* sgs_avg_of(fsl_ctx* c)
{
    static thread_local SgsAvg a;
    if (c->sgs_avg == FSL_SGS_AVG_NONE || !c->avg_buf) return nullptr;
    a.mode = c->sgs_avg;
    a.lm = c->avg_buf;
    a.mm = c->avg_buf + c->g.G;
    a.sums = c->avg_buf + 2 * c->g.G;
    return &a;
}

// The dynamic model of every context is the FP64 marching kernel, mixed
// contexts included: at 384x256x256 it takes 1.36 ms where the FP32
// two-kernel form (filters and Germano contraction in FP32, 16 float scratch
// arrays) takes 2.34 ms, and it is the parity arithmetic.  FSL_SGS_FP32=1
// selects the FP32 form in mixed contexts (read at each call).
bool sgs_fp32_form(const fsl_ctx* c)
{
    if (c->cfg.precision != FSL_PREC_MIXED) return false;
    const char* e = getenv("FSL_SGS_FP32");
    return e && e[0] == '1' && e[1] == 0;
}

int ensure_sgs(fsl_ctx* c)
{
    // the FP64 model keeps its intermediates on chip; the FP32 form has 16
    // float scratch arrays, allocated on its first use
    if (c->sgs || !sgs_fp32_form(c)) return FSL_OK;
    const size_t n = sgs_f32_scratch_bytes(c->g);
    HIPCHK(c, hipMalloc(&c->sgs, n));
    // the A tensor's ghost layer must read as 0 (src/turbulence.c:369-379
    // writes the interior only, into fresh zeroed arrays)
    HIPCHK(c, hipMemsetAsync(c->sgs, 0, n, c->stream));
    c->bytes += n;
    return FSL_OK;
}

int ensure_filt(fsl_ctx* c)
{
    if (c->filt) return FSL_OK;
    const size_t n = 4 * (size_t)c->g.G;
    HIPCHK(c, hipMalloc(&c->filt, n * sizeof(double)));
    c->bytes += n * sizeof(double);
    return FSL_OK;
}

int check_launch(fsl_ctx* c, const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_err(c, FSL_ERR_HIP, "%s launch failed: %s", what, hipGetErrorString(e));
    return FSL_OK;
}

void derive(fsl_ctx* c)
{
    // src/initialize.c:67-94
    const fsl_case& cs = c->cs;
    Phys& ph = c->ph;
    const double d[3] = {cs.deltaX, cs.deltaY, cs.deltaZ};
    for (int e = 0; e < 3; e++) {
        ph.rdel[e] = 1. / d[e];
        ph.r4del[e] = 1. / (4 * d[e]);
        ph.r2del[e] = ph.r4del[e] + ph.r4del[e];
    }
    ph.mu_L = cs.mu_L;
    ph.lambda_L = cs.mu_L * 717.75 * 1.4 / cs.Pr_L;
    ph.cp_Pr_T = 717.75 * 1.4 / cs.Pr_T;
    ph.DD = pow(cs.deltaX * cs.deltaY * cs.deltaZ, 0.666666666666667);
    ph.CsDD = cs.Cs * ph.DD;
    ph.dynamic = (c->cfg.sgs_model != FSL_SGS_CONSTANT);
    ph.z_quirk = c->cfg.z_quirk;
}

}  // namespace fsl

// ----------------------------------------------------------------------
extern "C" {

void fsl_config_defaults(fsl_config* cfg)
{
    memset(cfg, 0, sizeof *cfg);
    cfg->sgs_model = FSL_SGS_DYNAMIC;
    cfg->precision = FSL_PREC_FP64;
    cfg->cfl_mode = FSL_CFL_CONVECTIVE;
    cfg->cfl = 0.5;
    cfg->z_quirk = 1;
    cfg->fused = 0;
    cfg->rho_in = 1.0;
    cfg->p_in = 100000.;
    cfg->p_out = 100000.;
    cfg->device = 0;
}

int fsl_create(const fsl_case* cs, const fsl_config* cfg, fsl_ctx** out)
{
    if (!cs || !out) return FSL_ERR_ARG;
    *out = nullptr;
    char msg[256];
    if (fsl_case_validate(cs, msg, sizeof msg) != FSL_OK) {
        fprintf(stderr, "fsl_create: %s\n", msg);
        return FSL_ERR_ARG;
    }
    fsl_ctx* c = new (std::nothrow) fsl_ctx();
    if (!c) return FSL_ERR_NOMEM;
    c->cs = *cs;
    if (cfg) c->cfg = *cfg;
    else fsl_config_defaults(&c->cfg);
    if ((c->cfg.precision != FSL_PREC_FP64 && c->cfg.precision != FSL_PREC_MIXED &&
         c->cfg.precision != FSL_PREC_FP64_FAST) ||
        c->cfg.sgs_model < FSL_SGS_CONSTANT || c->cfg.sgs_model > FSL_SGS_LEARNED) {
        delete c;
        return FSL_ERR_ARG;
    }
    if (c->cfg.precision == FSL_PREC_FP64_FAST && !c->cfg.fused) {
        // the fast build is one of the fused stage kernel; the unfused entry
        // points are the parity kernels only
        fprintf(stderr, "fsl_create: precision FSL_PREC_FP64_FAST needs the fused path (fused = 1)\n");
        delete c;
        return FSL_ERR_UNSUPPORTED;
    }
    if (c->cfg.fused && cs->nStages != 3) {
        // the fused stage kernels implement the three-stage TVD RK only; the
        // reference's nStages = 2 (two-stage update falling through into the
        // three-stage one, src/evolution.c:15-27) runs on the unfused path
        fprintf(stderr, "fsl_create: the fused path needs nStages = 3 (case has %d)\n", cs->nStages);
        delete c;
        return FSL_ERR_UNSUPPORTED;
    }
    const int L = cs->LEN, H = cs->HIG, D = cs->DEP;
    if (L < 1 || H < 1 || D < 1) {
        fprintf(stderr, "fsl_create: mesh %d x %d x %d has an empty direction\n", L, H, D);
        delete c;
        return FSL_ERR_ARG;
    }
    c->g.L = L; c->g.H = H; c->g.D = D;
    c->g.sj = D + 2;
    c->g.si = (long long)(H + 2) * (D + 2);
    c->g.G = (long long)(L + 2) * c->g.si;
    if (c->cfg.sgs_model == FSL_SGS_LEARNED && c->g.si * 8 > 0x7fffffffLL) {
        // the network kernel addresses one (j, k) plane of a field with 32-bit byte offsets
        fprintf(stderr, "fsl_create: a %d x %d plane exceeds the learned closure's 2^31-byte planes\n", H, D);
        delete c;
        return FSL_ERR_UNSUPPORTED;
    }
    if (c->cfg.precision == FSL_PREC_MIXED && c->g.si * 8 > 0x7fffffffLL) {
        // the mixed stage kernel addresses one (j, k) plane of a field with 32-bit byte offsets
        fprintf(stderr, "fsl_create: a %d x %d plane is too large for the mixed-precision path\n", H, D);
        delete c;
        return FSL_ERR_UNSUPPORTED;
    }
    c->nface[0] = (size_t)(L + 1) * H * D;
    c->nface[1] = (size_t)L * (H + 1) * D;
    c->nface[2] = (size_t)L * H * (D + 1);
    derive(c);
    c->dt = cs->deltaT;
    c->dtX = c->dt / cs->deltaX;
    c->dtY = c->dt / cs->deltaY;
    c->dtZ = c->dt / cs->deltaZ;
    fsl_block_init(&c->blk, 0.0);

    int rc = FSL_OK;
    do {
        if (hipSetDevice(c->cfg.device) != hipSuccess) { rc = set_err(c, FSL_ERR_HIP, "hipSetDevice"); break; }
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            rc = set_err(c, FSL_ERR_HIP, "hipStreamCreate");
            break;
        }
        const size_t sb = 5 * (size_t)c->g.G * sizeof(double);
        if (hipMalloc(&c->U, sb) != hipSuccess || hipMalloc(&c->Up, sb) != hipSuccess) {
            rc = set_err(c, FSL_ERR_NOMEM, "state allocation failed");
            break;
        }
        c->bytes += 2 * sb;
        if (c->cfg.fused) {
            if (hipMalloc(&c->Uq, sb) != hipSuccess) { rc = set_err(c, FSL_ERR_NOMEM, "Uq"); break; }
            c->bytes += sb;
            if (hipMemsetAsync(c->Uq, 0, sb, c->stream) != hipSuccess) { rc = set_err(c, FSL_ERR_HIP, "clear Uq"); break; }
        }
        if (hipMalloc(&c->mu, c->g.G * sizeof(double)) != hipSuccess ||
            hipMalloc(&c->u_in, (size_t)H * sizeof(double)) != hipSuccess ||
            hipMalloc(&c->red, 64) != hipSuccess ||
            hipHostMalloc((void**)&c->red_host, 64, hipHostMallocDefault) != hipSuccess) {
            rc = set_err(c, FSL_ERR_NOMEM, "allocation failed");
            break;
        }
        c->bytes += c->g.G * sizeof(double) + H * sizeof(double) + 64;
        const int nb = sums_nblocks(c->g);
        if (hipMalloc(&c->sums, ((size_t)nb * 5 + 8) * sizeof(double)) != hipSuccess) {
            rc = set_err(c, FSL_ERR_NOMEM, "sums");
            break;
        }
        // ghost layers and mu_SGS read as 0 until written (the dynamic model
        // and the network's features read U's ghosts, the fluxes mu's)
        if (hipMemsetAsync(c->U, 0, sb, c->stream) != hipSuccess ||
            hipMemsetAsync(c->Up, 0, sb, c->stream) != hipSuccess ||
            hipMemsetAsync(c->mu, 0, c->g.G * sizeof(double), c->stream) != hipSuccess) {
            rc = set_err(c, FSL_ERR_HIP, "clearing the state buffers failed");
            break;
        }
        // reference inflow: 76.4 m/s for rows j <= HIG/2, 200 above
        std::vector<double> u(H);
        for (int j = 1; j <= H; j++) u[j - 1] = (j <= H / 2) ? 76.4 : 200;
        c->dU = 200 - 76.4;
        c->dw0 = 2 * cs->deltaY;   // one-cell jump over a two-cell central difference
        if (hipMemcpyAsync(c->u_in, u.data(), H * sizeof(double), hipMemcpyHostToDevice, c->stream) !=
            hipSuccess) {
            rc = set_err(c, FSL_ERR_HIP, "inflow upload");
            break;
        }
        if (c->cfg.sgs_model != FSL_SGS_CONSTANT && (rc = ensure_sgs(c)) != FSL_OK) break;
        if (!c->cfg.fused && (rc = ensure_faces(c)) != FSL_OK) break;
        if (hipStreamSynchronize(c->stream) != hipSuccess) { rc = set_err(c, FSL_ERR_HIP, "sync"); break; }
    } while (0);
    if (rc != FSL_OK) {
        fprintf(stderr, "fsl_create: %s\n", c->err);
        fsl_destroy(c);
        return rc;
    }
    c->bnd.u_in = c->u_in;
    c->bnd.rho_in = c->cfg.rho_in;
    c->bnd.p_in = c->cfg.p_in;
    c->bnd.p_out = c->cfg.p_out;
    c->bnd.BL_HIG = (int)cs->BL_HIG;
    *out = c;
    return FSL_OK;
}

void fsl_destroy(fsl_ctx* c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    prof_drain(c);
    for (auto e : c->pool) (void)hipEventDestroy(e);
    (void)hipFree(c->U);
    (void)hipFree(c->Up);
    (void)hipFree(c->Uq);
    (void)hipFree(c->faces);
    (void)hipFree(c->mu);
    (void)hipFree(c->sgs);
    (void)hipFree(c->filt);
    (void)hipFree(c->u_in);
    (void)hipFree(c->sums);
    (void)hipFree(c->red);
    (void)hipFree(c->probe_buf);
    (void)hipFree(c->avg_buf);
    if (c->red_host) (void)hipHostFree(c->red_host);
    learned_free(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* fsl_last_error(const fsl_ctx* c) { return c ? c->err : "null context"; }
size_t fsl_ncell(const fsl_ctx* c) { return (size_t)c->g.G; }
size_t fsl_nface(const fsl_ctx* c, int dir) { return (dir >= 0 && dir < 3) ? c->nface[dir] : 0; }
size_t fsl_device_bytes(const fsl_ctx* c) { return c->bytes; }

int fsl_get_case(const fsl_ctx* c, fsl_case* out)
{
    *out = c->cs;
    return FSL_OK;
}

int fsl_get_config(const fsl_ctx* c, fsl_config* out)
{
    *out = c->cfg;
    return FSL_OK;
}

int fsl_set_inflow(fsl_ctx* c, const double* u_in)
{
    double lo = u_in[0], hi = u_in[0];
    for (int j = 1; j < c->g.H; j++) {
        lo = fmin(lo, u_in[j]);
        hi = fmax(hi, u_in[j]);
    }
    c->dU = hi - lo;
    double gmax = 0;
    for (int j = 1; j + 1 < c->g.H; j++) gmax = fmax(gmax, fabs(u_in[j + 1] - u_in[j - 1]) / (2 * c->cs.deltaY));
    c->dw0 = gmax > 0 ? c->dU / gmax : c->g.H * c->cs.deltaY;
    c->lrn.dw_valid = 0;
    HIPCHK(c, hipMemcpyAsync(c->u_in, u_in, c->g.H * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSL_OK;
}

int fsl_upload_state(fsl_ctx* c, int which, const double* U)
{
    c->cfl_ready = 0;
    double* d = which ? c->Up : c->U;
    HIPCHK(c, hipMemcpyAsync(d, U, 5 * c->g.G * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSL_OK;
}

int fsl_download_state(fsl_ctx* c, int which, double* U)
{
    const double* d = which ? c->Up : c->U;
    HIPCHK(c, hipMemcpyAsync(U, d, 5 * c->g.G * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSL_OK;
}

static int face_ptr(fsl_ctx* c, int id, double** p, size_t* n)
{
    if (id < 0 || id >= 30) return set_err(c, FSL_ERR_ARG, "face id %d", id);
    int rc = ensure_faces(c);
    if (rc) return rc;
    Faces F = faces_of(c);
    *p = F.a[id];
    *n = c->nface[id / 10];
    return FSL_OK;
}

int fsl_download_face(fsl_ctx* c, int id, double* out)
{
    double* p;
    size_t n;
    int rc = face_ptr(c, id, &p, &n);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(out, p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSL_OK;
}

int fsl_upload_face(fsl_ctx* c, int id, const double* in)
{
    double* p;
    size_t n;
    int rc = face_ptr(c, id, &p, &n);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(p, in, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSL_OK;
}

int fsl_download_mu_sgs(fsl_ctx* c, double* out)
{
    HIPCHK(c, hipMemcpyAsync(out, c->mu, c->g.G * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSL_OK;
}

int fsl_upload_mu_sgs(fsl_ctx* c, const double* in)
{
    HIPCHK(c, hipMemcpyAsync(c->mu, in, c->g.G * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSL_OK;
}

int fsl_set_dt(fsl_ctx* c, double deltaT)
{
    c->cfl_ready = 0;
    c->dt = deltaT;
    c->dtX = c->dt / c->cs.deltaX;
    c->dtY = c->dt / c->cs.deltaY;
    c->dtZ = c->dt / c->cs.deltaZ;
    return FSL_OK;
}

double fsl_get_dt(const fsl_ctx* c) { return c->dt; }

int fsl_set_block(fsl_ctx* c, unsigned k_min, unsigned k_max, double Ud, double Vd, double Wd)
{
    c->blk.k_min = k_min;
    c->blk.k_max = k_max;
    c->blk.Ud = Ud;
    c->blk.Vd = Vd;
    c->blk.Wd = Wd;
    return FSL_OK;
}

int fsl_get_cur(const fsl_ctx* c) { return c->cur_is_p; }

int fsl_set_cur(fsl_ctx* c, int is_p)
{
    c->cur_is_p = is_p ? 1 : 0;
    return FSL_OK;
}

int fsl_set_lcg(fsl_ctx* c, double X, int counter)
{
    c->blk.X = X;
    c->blk.counter = counter;
    return FSL_OK;
}

int fsl_get_lcg(const fsl_ctx* c, double* X, int* counter, unsigned* k_min, unsigned* k_max, double* Ud,
                double* Vd, double* Wd)
{
    if (X) *X = c->blk.X;
    if (counter) *counter = c->blk.counter;
    if (k_min) *k_min = c->blk.k_min;
    if (k_max) *k_max = c->blk.k_max;
    if (Ud) *Ud = c->blk.Ud;
    if (Vd) *Vd = c->blk.Vd;
    if (Wd) *Wd = c->blk.Wd;
    return FSL_OK;
}

int fsl_get_block_state(const fsl_ctx* c, fsl_block* b)
{
    *b = c->blk;
    return FSL_OK;
}

int fsl_set_block_state(fsl_ctx* c, const fsl_block* b)
{
    c->blk = *b;
    return FSL_OK;
}

unsigned fsl_get_step(const fsl_ctx* c) { return c->step; }

int fsl_set_step(fsl_ctx* c, unsigned step)
{
    c->step = step;
    return FSL_OK;
}

int fsl_synchronize(fsl_ctx* c)
{
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSL_OK;
}

// --- operations ------------------------------------------------------

int fsl_ghost_fill(fsl_ctx* c)
{
    prof_mark(c, K_GHOST, 0);
    launch_ghost_fill(c->stream, st5(cur_state(c), c->g.G), c->g, bnd_now(c));
    prof_mark(c, K_GHOST, 1);
    return check_launch(c, "ghost_fill");
}

int fsl_reconstruct(fsl_ctx* c)
{
    int rc = ensure_faces(c);
    if (rc) return rc;
    prof_mark(c, K_RECON, 0);
    launch_recon(c->stream, cst5(cur_state(c), c->g.G), faces_of(c), c->g, c->cfg.z_quirk);
    prof_mark(c, K_RECON, 1);
    return check_launch(c, "recon");
}

int fsl_iface_bc(fsl_ctx* c)
{
    int rc = ensure_faces(c);
    if (rc) return rc;
    prof_mark(c, K_IFACE, 0);
    launch_iface_bc(c->stream, faces_of(c), c->g, bnd_now(c));
    prof_mark(c, K_IFACE, 1);
    return check_launch(c, "iface_bc");
}

int fsl_dynamic_smagorinsky(fsl_ctx* c)
{
    int rc = ensure_sgs(c);
    if (rc) return rc;
    if (sgs_fp32_form(c))
        launch_sgs_f32(c->stream, cst5(c->U, c->g.G), c->mu, c->sgs, c->g, c->ph, prof_mark, c, nullptr,
                       sgs_avg_of(c));
    else
        launch_sgs(c->stream, cst5(c->U, c->g.G), c->mu, c->g, c->ph, prof_mark, c, nullptr, sgs_avg_of(c));
    return check_launch(c, "dynamic_smagorinsky");
}

int fsl_set_sgs_average(fsl_ctx* c, int mode)
{
    if (mode != FSL_SGS_AVG_NONE && mode != FSL_SGS_AVG_Z && mode != FSL_SGS_AVG_XZ)
        return set_err(c, FSL_ERR_ARG, "averaging mode %d", mode);
    if (mode != FSL_SGS_AVG_NONE && !c->avg_buf) {
        const size_t n = sgs_avg_bytes(c->g);
        HIPCHK(c, hipMalloc(&c->avg_buf, n));
        c->bytes += n;
    }
    c->sgs_avg = mode;
    return FSL_OK;
}

int fsl_get_sgs_average(const fsl_ctx* c) { return c->sgs_avg; }

int fsl_fluxes(fsl_ctx* c)
{
    int rc = ensure_faces(c);
    if (rc) return rc;
    if (c->cfg.sgs_model == FSL_SGS_DYNAMIC) {
        // src/fluxes.c:52-54: the model runs on the step state U1..U4
        rc = fsl_dynamic_smagorinsky(c);
        if (rc) return rc;
    } else if (c->cfg.sgs_model == FSL_SGS_LEARNED && !c->cur_is_p) {
        // once per step, at the stage that reads the step state
        rc = learned_sgs_step(c);
        if (rc) return rc;
    }
    const Faces F = faces_of(c);
    const CState5 U = cst5(cur_state(c), c->g.G);
    for (int d = 0; d < 3; d++) {
        prof_mark(c, K_FLUX_X + d, 0);
        launch_flux(c->stream, d, U, F, c->mu, c->g, c->ph);
        prof_mark(c, K_FLUX_X + d, 1);
    }
    return check_launch(c, "fluxes");
}

int fsl_rk_update(fsl_ctx* c, int scheme, int stage)
{
    c->cfl_ready = 0;   // writes U or Up
    int rc = ensure_faces(c);
    if (rc) return rc;
    if (stage < 1 || stage > 3 || (scheme != 2 && scheme != 3))
        return set_err(c, FSL_ERR_ARG, "scheme %d stage %d", scheme, stage);
    const Faces F = faces_of(c);
    const Dt3 dt{c->dtX, c->dtY, c->dtZ};
    State5 U = st5(c->U, c->g.G), Up = st5(c->Up, c->g.G);
    prof_mark(c, K_RK, 0);
    if (scheme == 2) {
        launch_rk(c->stream, stage == 1 ? 21 : 22, U, Up, F, c->g, dt);
        c->cur_is_p = (stage == 1) ? 1 : 0;
    } else {
        launch_rk(c->stream, 30 + stage, U, Up, F, c->g, dt);
        c->cur_is_p = (stage == 1 || stage == 2) ? 1 : 0;
    }
    prof_mark(c, K_RK, 1);
    return check_launch(c, "evolution");
}

int fsl_evolution(fsl_ctx* c, int numStages, int stage)
{
    int rc;
    // src/evolution.c:15-27: no break after case 2, so the two-stage update
    // is followed by the three-stage update of the same stage
    if (numStages == 2 && (rc = fsl_rk_update(c, 2, stage)) != FSL_OK) return rc;
    return fsl_rk_update(c, 3, stage);
}

int fsl_check_co_num(fsl_ctx* c, double* maxCo_out)
{
    const int mode = c->cfg.cfl_mode == FSL_CFL_ACOUSTIC ? 1 : 0;
    if (!c->cfl_ready) {
        // otherwise the last fused stage 3 reduced the same quantity while writing U
        HIPCHK(c, hipMemsetAsync(c->red, 0, sizeof(unsigned long long), c->stream));
        c->cfl_zeroed = 0;   // the cell now holds this reduction
        prof_mark(c, K_CFL, 0);
        launch_cfl(c->stream, mode, cst5(c->U, c->g.G), c->g, c->ph, c->dtX, c->red);
        prof_mark(c, K_CFL, 1);
        int rc = check_launch(c, "cfl");
        if (rc) return rc;
    }
    c->cfl_ready = 0;
    // the maximum comes back through pinned memory (a pageable destination
    // is staged by the runtime and costs tens of microseconds per step)
    HIPCHK(c, hipMemcpyAsync(c->red_host, c->red, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const unsigned long long bits = *c->red_host;
    double m;
    memcpy(&m, &bits, sizeof m);
    if (maxCo_out) *maxCo_out = m;
    // the reduction carries NaN cells through, so a partly non-finite state
    // gives NaN here; the context keeps its last valid time step
    const double dt = (mode == 0) ? c->dt * (c->cs.maxCoNum / m)   // src/helpers.c:47-50
                                  : c->cfg.cfl / m;
    if (!(m > 0.0) || !isfinite(m) || !isfinite(dt) || !(dt > 0.0))
        return set_err(c, FSL_ERR_NONFINITE, "CFL reduction gave %g (dt %g)", m, dt);
    c->dt = dt;
    c->dtX = c->dt / c->cs.deltaX;
    c->dtY = c->dt / c->cs.deltaY;
    c->dtZ = c->dt / c->cs.deltaZ;
    return FSL_OK;
}

int fsl_filter(fsl_ctx* c, const double* in, double* out, const double h[3])
{
    int rc = ensure_filt(c);
    if (rc) return rc;
    const long long G = c->g.G;
    double *d_in = c->filt, *d_out = c->filt + G, *t1 = c->filt + 2 * G, *t2 = c->filt + 3 * G;
    const size_t nb = G * sizeof(double);
    HIPCHK(c, hipMemcpyAsync(d_in, in, nb, hipMemcpyHostToDevice, c->stream));
    launch_filter(c->stream, d_in, d_out, t1, t2, c->g, h, prof_mark, c);
    rc = check_launch(c, "filter");
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(out, d_out, nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSL_OK;
}

int fsl_conservation_sums(fsl_ctx* c, double out[5])
{
    const int nb = sums_nblocks(c->g);
    prof_mark(c, K_SUM, 0);
    launch_sums(c->stream, cst5(c->U, c->g.G), c->g, c->sums, nb, c->sums + (size_t)nb * 5);
    prof_mark(c, K_SUM, 1);
    int rc = check_launch(c, "sums");
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(out, c->sums + (size_t)nb * 5, 5 * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSL_OK;
}

int fsl_stage(fsl_ctx* c, int stage)
{
    if (stage < 1 || stage > 3) return set_err(c, FSL_ERR_ARG, "stage %d", stage);
    if (!c->cfg.fused) {
        int rc;
        if ((rc = fsl_ghost_fill(c)) || (rc = fsl_reconstruct(c)) || (rc = fsl_iface_bc(c)) ||
            (rc = fsl_fluxes(c)) || (rc = fsl_evolution(c, c->cs.nStages, stage)))
            return rc;
        return FSL_OK;
    }
    return fused_stage(c, stage);
}

int fsl_profile_enable(fsl_ctx* c, int on)
{
    c->prof = on ? 1 : 0;
    return FSL_OK;
}

int fsl_profile_reset(fsl_ctx* c)
{
    prof_drain(c);
    memset(c->prof_ms, 0, sizeof c->prof_ms);
    memset(c->prof_cnt, 0, sizeof c->prof_cnt);
    return FSL_OK;
}

int fsl_profile_read(fsl_ctx* c, int nmax, double* ms, long long* count)
{
    prof_drain(c);
    for (int i = 0; i < nmax && i < K_NKERNELS; i++) {
        ms[i] = c->prof_ms[i];
        count[i] = c->prof_cnt[i];
    }
    return K_NKERNELS;
}

const char* fsl_kernel_name(int id) { return (id >= 0 && id < K_NKERNELS) ? kNames[id] : "?"; }
int fsl_kernel_count(void) { return K_NKERNELS; }

int fsl_copy_kernel(fsl_ctx* c, size_t n, int reps)
{
    double *a = nullptr, *b = nullptr;
    HIPCHK(c, hipMalloc(&a, n * sizeof(double)));
    if (hipMalloc(&b, n * sizeof(double)) != hipSuccess) {
        (void)hipFree(a);
        return set_err(c, FSL_ERR_NOMEM, "copy buffer");
    }
    if (hipMemsetAsync(a, 0, n * sizeof(double), c->stream) != hipSuccess) {
        (void)hipFree(a);
        (void)hipFree(b);
        return set_err(c, FSL_ERR_HIP, "clear copy buffer");
    }
    for (int r = 0; r < reps; r++) {
        prof_mark(c, K_COPY, 0);
        launch_copy(c->stream, a, b, n);
        prof_mark(c, K_COPY, 1);
    }
    int rc = check_launch(c, "copy");
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

}  // extern "C"

// --- accessors for the fused path (fsl_fused.hip) ---------------------
namespace fsl {

CtxView ctx_view(fsl_ctx* c)
{
    CtxView v;
    v.stream = c->stream;
    v.g = c->g;
    v.ph = c->ph;
    v.bnd = bnd_now(c);
    v.U = c->U;
    v.Up = c->Up;
    v.Uq = c->Uq;
    v.mu = c->mu;
    v.dtX = c->dtX;
    v.dtY = c->dtY;
    v.dtZ = c->dtZ;
    v.cur_is_p = c->cur_is_p;
    v.sgs_model = c->cfg.sgs_model;
    v.precision = c->cfg.precision;
    v.red = c->red;
    v.cfl_mode = c->cfg.cfl_mode == FSL_CFL_ACOUSTIC ? 1 : 0;
    return v;
}

void ctx_prof(fsl_ctx* c, int id, int end) { prof_mark(c, id, end); }
int ctx_ghost(fsl_ctx* c, const double* state, bool z)
{
    prof_mark(c, K_GHOST, 0);
    launch_ghost_fill(c->stream, st5(const_cast<double*>(state), c->g.G), c->g, bnd_now(c), z);
    prof_mark(c, K_GHOST, 1);
    return check_launch(c, "ghost_fill");
}
int ctx_check(fsl_ctx* c, const char* what) { return check_launch(c, what); }
int ctx_cfl_reset(fsl_ctx* c)
{
    c->cfl_ready = 0;
    // a fused stage 2 on the same stream has already cleared the cell
    if (!c->cfl_zeroed) HIPCHK(c, hipMemsetAsync(c->red, 0, sizeof(unsigned long long), c->stream));
    c->cfl_zeroed = 0;
    return FSL_OK;
}
void ctx_cfl_zeroed(fsl_ctx* c, int z) { c->cfl_zeroed = z; }
void ctx_cfl_ready(fsl_ctx* c) { c->cfl_ready = 1; }
int ctx_sgs(fsl_ctx* c) { return fsl_dynamic_smagorinsky(c); }
void ctx_set_cur(fsl_ctx* c, int is_p) { c->cur_is_p = is_p; }
int ctx_err(fsl_ctx* c, int code, const char* msg) { return set_err(c, code, "%s", msg); }
}  // namespace fsl

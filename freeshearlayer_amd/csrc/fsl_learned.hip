// fsl_learned.hip -- the learned SGS closure behind the C-ABI: parameter
// upload and BF16 packing, the device sample store that the dynamic model
// fills (features of the step state, targets Cd / 0.15), the training loop
// (MFMA gradient kernel + Adam), and the per-step selection between the
// learned and the dynamic mu_SGS.
//
// Selection (FSL_SGS_LEARNED).  Every `interval` steps (steps 1, 1 +
// interval, ...) the step is judged: the dynamic model of src/turbulence.c
// and the network both evaluate mu_SGS on the step state, and the relative
// difference of their SGS dissipation sum(mu |S|^2) over the interior is the
// error.  Above `tol` the dynamic mu_SGS is used and the closure stays in
// fallback (dynamic model every step) until a later judge passes; otherwise
// the network's mu_SGS is used.  Between judges the network alone runs.
// The judge also refreshes delta_w(x), the thickness in the y / delta_w
// feature, from the Favre-averaged profile of each x plane, and sets the
// level of the network's Cd for the following steps from the dynamic
// model's dissipation: one global factor (adapt 1) or one per y row
// (adapt 2).
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "fsl_ctx.hpp"

using namespace fsl;

namespace {

int ensure_params(fsl_ctx* c)
{
    LearnedState& L = c->lrn;
    if (L.master) return FSL_OK;
    HIPCHK(c, hipMalloc(&L.master, mlp_nparam() * sizeof(float)));
    HIPCHK(c, hipMalloc(&L.pack, (size_t)mlp_npack() * 8));
    HIPCHK(c, hipMalloc(&L.norm, 16 * sizeof(float)));
    c->bytes += mlp_nparam() * sizeof(float) + (size_t)mlp_npack() * 8 + 64;
    return FSL_OK;
}

int ensure_train(fsl_ctx* c)
{
    LearnedState& L = c->lrn;
    if (L.grad) return FSL_OK;
    const size_t nb = mlp_nparam() * sizeof(float);
    HIPCHK(c, hipMalloc(&L.grad, nb));
    HIPCHK(c, hipMalloc(&L.adam_m, nb));
    HIPCHK(c, hipMalloc(&L.adam_v, nb));
    HIPCHK(c, hipMalloc(&L.loss, sizeof(float)));
    HIPCHK(c, hipMemsetAsync(L.grad, 0, nb, c->stream));
    HIPCHK(c, hipMemsetAsync(L.adam_m, 0, nb, c->stream));
    HIPCHK(c, hipMemsetAsync(L.adam_v, 0, nb, c->stream));
    HIPCHK(c, hipMemsetAsync(L.loss, 0, sizeof(float), c->stream));
    L.adam_t = 0;
    c->bytes += 3 * nb + sizeof(float);
    return FSL_OK;
}

int ensure_diag(fsl_ctx* c)
{
    LearnedState& L = c->lrn;
    if (L.zs) return FSL_OK;
    const size_t nz = 7 * (size_t)c->g.L * c->g.H;
    HIPCHK(c, hipMalloc(&L.zs, nz * sizeof(double)));
    HIPCHK(c, hipMalloc(&L.dw, (size_t)(c->g.L + 2) * sizeof(float)));
    c->bytes += nz * sizeof(double) + (c->g.L + 2) * sizeof(float);
    return FSL_OK;
}

int ensure_judge(fsl_ctx* c)
{
    LearnedState& L = c->lrn;
    if (L.mu_alt) return FSL_OK;
    HIPCHK(c, hipMalloc(&L.mu_alt, c->g.G * sizeof(double)));
    // the ghost layer of mu_SGS reads as 0 (src/turbulence.c:436-445 writes the interior only)
    HIPCHK(c, hipMemsetAsync(L.mu_alt, 0, c->g.G * sizeof(double), c->stream));
    const size_t nr = (size_t)dissipation_splits(c->g) * c->g.H;
    HIPCHK(c, hipMalloc(&L.part, 5 * nr * sizeof(double)));
    HIPCHK(c, hipHostMalloc((void**)&L.hpart, 5 * nr * sizeof(double), hipHostMallocDefault));
    c->bytes += c->g.G * sizeof(double) + 5 * nr * sizeof(double);
    return FSL_OK;
}

// Level factors of the network's scale (adapt 2 and 3): band_f holds one
// factor per (x band, row), the bands those of the judge's dissipation sums
// (band s: planes s L / NB + 1 .. (s + 1) L / NB, integer division); the
// kernel reads them per (plane, row), linear in x between band centres and
// constant beyond the first and last centre (one band: the row's factor on
// every plane).  Ghost planes and rows take their neighbour's factor.
int upload_levels(fsl_ctx* c)
{
    LearnedState& L = c->lrn;
    const int Lx = c->g.L, H = c->g.H, NB = L.nband;
    const size_t n = (size_t)(Lx + 2) * (H + 2);
    if (!L.fac) {
        HIPCHK(c, hipMalloc(&L.fac, n * sizeof(float)));
        c->bytes += n * sizeof(float);
    }
    L.hfac.assign(n, 1.f);
    std::vector<double> cs(NB);
    for (int b = 0; b < NB; b++)
        cs[b] = 0.5 * ((double)((long long)b * Lx / NB) + 1. + (double)((long long)(b + 1) * Lx / NB));
    for (int i = 1; i <= Lx; i++) {
        int b0 = 0;
        double w = 0.;
        if (NB > 1) {
            if (i <= cs[0]) {
                b0 = 0;
            } else if (i >= cs[NB - 1]) {
                b0 = NB - 1;
            } else {
                while (b0 + 1 < NB - 1 && i > cs[b0 + 1]) b0++;
                w = (i - cs[b0]) / (cs[b0 + 1] - cs[b0]);
            }
        }
        for (int j = 1; j <= H; j++) {
            const double f0 = L.band_f[(size_t)b0 * H + j - 1];
            const double f = w > 0. ? (1. - w) * f0 + w * L.band_f[(size_t)(b0 + 1) * H + j - 1] : f0;
            L.hfac[(size_t)i * (H + 2) + j] = (float)f;
        }
        L.hfac[(size_t)i * (H + 2)] = L.hfac[(size_t)i * (H + 2) + 1];
        L.hfac[(size_t)i * (H + 2) + H + 1] = L.hfac[(size_t)i * (H + 2) + H];
    }
    for (int j = 0; j < H + 2; j++) {
        L.hfac[j] = L.hfac[(size_t)(H + 2) + j];
        L.hfac[(size_t)(Lx + 1) * (H + 2) + j] = L.hfac[(size_t)Lx * (H + 2) + j];
    }
    HIPCHK(c, hipMemcpyAsync(L.fac, L.hfac.data(), n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSL_OK;
}

// all level factors 1, with the band count of the adapt mode
int reset_levels(fsl_ctx* c)
{
    LearnedState& L = c->lrn;
    if (L.adapt < 2) return FSL_OK;
    L.nband = L.adapt == 3 ? dissipation_splits(c->g) : 1;
    L.band_f.assign((size_t)L.nband * c->g.H, 1.f);
    return upload_levels(c);
}

int ensure_cd(fsl_ctx* c)
{
    LearnedState& L = c->lrn;
    if (L.cd) return FSL_OK;
    HIPCHK(c, hipMalloc(&L.cd, c->g.G * sizeof(float)));
    HIPCHK(c, hipMemsetAsync(L.cd, 0, c->g.G * sizeof(float), c->stream));
    c->bytes += c->g.G * sizeof(float);
    return FSL_OK;
}

// delta_w(x) from the step state U
int update_thickness(fsl_ctx* c)
{
    int rc = ensure_diag(c);
    if (rc) return rc;
    prof_mark(c, K_DIAG, 0);
    launch_diag_zsum(c->stream, cst5(c->U, c->g.G), c->g, c->lrn.zs);
    launch_diag_dw(c->stream, c->lrn.zs, c->g, c->dU, c->cs.deltaY, (float)c->dw0, c->lrn.dw);
    prof_mark(c, K_DIAG, 1);
    c->lrn.dw_valid = 1;
    return check_launch(c, "plane profiles");
}

int need_ready(fsl_ctx* c)
{
    if (!c->lrn.ready)
        return set_err(c, FSL_ERR_ARG, "learned closure has no parameters (fsl_mlp_set_params)");
    if (!(c->dU > 0)) return set_err(c, FSL_ERR_ARG, "learned closure needs two distinct inflow streams");
    return FSL_OK;
}

int learned_mu(fsl_ctx* c, double* mu)
{
    LearnedState& L = c->lrn;
    if (!L.dw_valid) {
        int rc = update_thickness(c);
        if (rc) return rc;
    }
    prof_mark(c, K_MLP_MU, 0);
    launch_mlp_mu(c->stream, cst5(c->U, c->g.G), c->g, c->ph, c->dU, L.dw, c->cs.deltaY, L.pack, L.master,
                  L.norm, mu, nullptr, L.scale, L.adapt >= 2 ? L.fac : nullptr, L.nband > 1, L.act);
    prof_mark(c, K_MLP_MU, 1);
    return check_launch(c, "learned mu_SGS");
}

int dynamic_mu(fsl_ctx* c, float* cd)
{
    int rc = ensure_sgs(c);
    if (rc) return rc;
    if (sgs_fp32_form(c))
        launch_sgs_f32(c->stream, cst5(c->U, c->g.G), c->mu, c->sgs, c->g, c->ph, prof_mark, c, cd, sgs_avg_of(c));
    else
        launch_sgs(c->stream, cst5(c->U, c->g.G), c->mu, c->g, c->ph, prof_mark, c, cd, sgs_avg_of(c));
    return check_launch(c, "dynamic_smagorinsky");
}

// The judge's measures of mu_b (the network's) against mu_a (the dynamic
// model's), with e = mu |S|^2 the SGS dissipation density and D its sum over
// the interior:
//   err  = |D_b - D_a| / D_a                              (global level)
//   perr = || s P_b - P_a ||_2 / || P_a ||_2, P(j) the x-z sum of e in row j
//   lerr = || s e_b - e_a ||_2 / || e_a ||_2 over the cells
// with s = D_a / D_b, the level the judge's rescale sets: perr measures the
// distribution of the dissipation across the layer, lerr its cell-by-cell
// distribution.
struct JudgeErr {
    double err, perr, lerr, da, db;
};

int judge_error(fsl_ctx* c, const double* mu_a, const double* mu_b, JudgeErr* out)
{
    LearnedState& L = c->lrn;
    const int ns = dissipation_splits(c->g), H = c->g.H;
    const size_t nr = (size_t)ns * H;
    launch_sgs_dissipation(c->stream, cst5(c->U, c->g.G), c->g, c->ph, mu_a, mu_b, L.part);
    int rc = check_launch(c, "dissipation");
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(L.hpart, L.part, 5 * nr * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // rows[q][j]: the NS partials of each row added in order
    std::vector<double> rows(5 * (size_t)H, 0.);
    for (int q = 0; q < 5; q++)
        for (int p = 0; p < ns; p++)
            for (int j = 0; j < H; j++) rows[(size_t)q * H + j] += L.hpart[((size_t)q * ns + p) * H + j];
    double tot[5] = {0., 0., 0., 0., 0.};
    for (int q = 0; q < 5; q++)
        for (int j = 0; j < H; j++) tot[q] += rows[(size_t)q * H + j];
    const double sa = tot[0], sb = tot[1];
    out->da = sa;
    out->db = sb;
    L.jrows.assign(rows.begin(), rows.begin() + 2 * (size_t)H);
    if (sa > 0) out->err = fabs(sb - sa) / sa;
    else out->err = sb > 0 ? INFINITY : 0.;
    if (sa > 0 && sb > 0) {
        const double sc = sa / sb;
        double num = 0., den = 0.;
        for (int j = 0; j < H; j++) {
            const double d = sc * rows[(size_t)H + j] - rows[j];
            num += d * d;
            den += rows[j] * rows[j];
        }
        out->perr = sqrt(num / den);
        const double l2 = sc * sc * tot[3] - 2. * sc * tot[4] + tot[2];
        out->lerr = sqrt(fmax(l2, 0.) / tot[2]);
    } else {
        out->perr = out->lerr = (sa > 0 || sb > 0) ? INFINITY : 0.;
    }
    return FSL_OK;
}

unsigned long long mix64(unsigned long long x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

unsigned long long gcd64(unsigned long long a, unsigned long long b)
{
    while (b) {
        const unsigned long long t = a % b;
        a = b;
        b = t;
    }
    return a;
}

}  // namespace

namespace fsl {

void learned_free(fsl_ctx* c)
{
    LearnedState& L = c->lrn;
    (void)hipFree(L.master);
    (void)hipFree(L.pack);
    (void)hipFree(L.norm);
    (void)hipFree(L.grad);
    (void)hipFree(L.adam_m);
    (void)hipFree(L.adam_v);
    (void)hipFree(L.loss);
    (void)hipFree(L.X);
    (void)hipFree(L.t);
    (void)hipFree(L.cd);
    (void)hipFree(L.zs);
    (void)hipFree(L.dw);
    (void)hipFree(L.mu_alt);
    (void)hipFree(L.part);
    (void)hipFree(L.fac);
    if (L.hpart) (void)hipHostFree(L.hpart);
    L = LearnedState();
}

int learned_sgs_step(fsl_ctx* c)
{
    int rc = need_ready(c);
    if (rc) return rc;
    LearnedState& L = c->lrn;
    const long long s = (long long)c->step;
    const bool judge = L.interval > 0 && s >= 1 && (s - 1) % L.interval == 0;
    if (judge) {
        if ((rc = ensure_judge(c)) || (rc = update_thickness(c))) return rc;
        if ((rc = dynamic_mu(c, nullptr))) return rc;           // into c->mu
        if ((rc = learned_mu(c, L.mu_alt))) return rc;
        JudgeErr je;
        if ((rc = judge_error(c, c->mu, L.mu_alt, &je))) return rc;
        const double err = je.err, perr = je.perr, dd = je.da, dl = je.db;
        L.judged++;
        L.last_error = err;
        L.last_profile_error = perr;
        L.last_local_error = je.lerr;
        L.hist_err.push_back((float)err);
        L.hist_perr.push_back((float)perr);
        L.hist_lerr.push_back((float)je.lerr);
        L.hist_scale.push_back(L.scale);
        // global level of the network's coefficient from the dynamic model
        // (the network supplies the spatial distribution): used from the next step
        const float next = (L.adapt && dl > 0 && dd > 0) ? std::min(10.f, std::max(0.1f, (float)(L.scale * dd / dl)))
                                                         : L.scale;
        // adapt 2 / 3: each (band, row) factor moves by its own ratio of the
        // dynamic to the network's dissipation over the global one, so its
        // level follows its own ratio; eps (1 % of the mean dissipation of a
        // band's row) keeps rows of negligible dissipation, outside the layer,
        // near their factor
        if (L.adapt >= 2 && dl > 0 && dd > 0 && L.fac) {
            const int H = c->g.H, NB = L.nband, ns = dissipation_splits(c->g);
            const double g = dd / dl, eps = 0.01 * dd / ((double)NB * H);
            for (int b = 0; b < NB; b++)
                for (int j = 0; j < H; j++) {
                    const double pa = NB == 1 ? L.jrows[j] : L.hpart[(size_t)b * H + j];
                    const double pb = NB == 1 ? L.jrows[H + j] : L.hpart[((size_t)ns + b) * H + j];
                    float& f = L.band_f[(size_t)b * H + j];
                    f = (float)std::min(10., std::max(0.1, f * ((pa + eps) / (pb + eps) / g)));
                }
            if ((rc = upload_levels(c))) return rc;
        }
        if (err > L.tol || perr > L.profile_tol || je.lerr > L.local_tol) {
            L.fallbacks++;
            L.in_fallback = 1;
            L.dynamic_steps++;
        } else {
            std::swap(c->mu, L.mu_alt);   // both arrays keep a zero ghost layer
            L.in_fallback = 0;
            L.learned_steps++;
        }
        L.scale = next;
    } else if (L.in_fallback) {
        if ((rc = dynamic_mu(c, nullptr))) return rc;
        L.dynamic_steps++;
    } else {
        if ((rc = learned_mu(c, c->mu))) return rc;
        L.learned_steps++;
    }
    return FSL_OK;
}

}  // namespace fsl

extern "C" {

int fsl_mlp_set_params(fsl_ctx* c, const float* params, const float* norm)
{
    int rc = ensure_params(c);
    if (rc) return rc;
    for (int p = 0; p < mlp_nparam(); p++)
        if (!isfinite(params[p])) return set_err(c, FSL_ERR_NONFINITE, "parameter %d is not finite", p);
    for (int q = 0; q < 16; q++)
        if (!isfinite(norm[q])) return set_err(c, FSL_ERR_NONFINITE, "normalisation %d is not finite", q);
    LearnedState& L = c->lrn;
    HIPCHK(c, hipMemcpyAsync(L.master, params, mlp_nparam() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(L.norm, norm, 16 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    launch_mlp_pack(c->stream, L.master, L.pack, L.act);
    rc = check_launch(c, "mlp pack");
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    L.ready = 1;
    L.adam_t = 0;
    L.scale = 1.f;   // the judge recalibrates the new network from scratch
    if ((rc = reset_levels(c))) return rc;
    if (L.grad) {   // a new parameter set restarts the optimizer
        const size_t nb = mlp_nparam() * sizeof(float);
        HIPCHK(c, hipMemsetAsync(L.adam_m, 0, nb, c->stream));
        HIPCHK(c, hipMemsetAsync(L.adam_v, 0, nb, c->stream));
        HIPCHK(c, hipMemsetAsync(L.grad, 0, nb, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return FSL_OK;
}

int fsl_mlp_set_activation(fsl_ctx* c, int activation)
{
    if (activation != FSL_MLP_TANH && activation != FSL_MLP_RELU)
        return set_err(c, FSL_ERR_ARG, "activation %d", activation);
    LearnedState& L = c->lrn;
    L.act = activation;
    if (L.ready) {   // the inference operands depend on the activation
        launch_mlp_pack(c->stream, L.master, L.pack, L.act);
        int rc = check_launch(c, "mlp pack");
        if (rc) return rc;
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    L.scale = 1.f;   // a different network: the judge recalibrates it
    return reset_levels(c);
}

int fsl_mlp_get_activation(const fsl_ctx* c) { return c->lrn.act; }

int fsl_mlp_get_params(fsl_ctx* c, float* params, float* norm)
{
    if (!c->lrn.ready) return set_err(c, FSL_ERR_ARG, "no parameters");
    if (params)
        HIPCHK(c, hipMemcpyAsync(params, c->lrn.master, mlp_nparam() * sizeof(float), hipMemcpyDeviceToHost,
                                 c->stream));
    if (norm) HIPCHK(c, hipMemcpyAsync(norm, c->lrn.norm, 16 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSL_OK;
}

int fsl_mlp_forward(fsl_ctx* c, const float* X, long long n, float* y)
{
    if (n < 0) return set_err(c, FSL_ERR_ARG, "n %lld", n);
    if (!c->lrn.ready) return set_err(c, FSL_ERR_ARG, "no parameters");
    if (n == 0) return FSL_OK;
    float *dX = nullptr, *dy = nullptr;
    HIPCHK(c, hipMalloc(&dX, n * 8 * sizeof(float)));
    if (hipMalloc(&dy, n * sizeof(float)) != hipSuccess) {
        (void)hipFree(dX);
        return set_err(c, FSL_ERR_NOMEM, "forward buffers");
    }
    int rc = FSL_OK;
    if (hipMemcpyAsync(dX, X, n * 8 * sizeof(float), hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = set_err(c, FSL_ERR_HIP, "upload");
    if (!rc) {
        launch_mlp_forward(c->stream, dX, n, c->lrn.pack, c->lrn.master, c->lrn.norm, dy, c->lrn.act);
        rc = check_launch(c, "mlp forward");
    }
    if (!rc && (hipMemcpyAsync(y, dy, n * sizeof(float), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess))
        rc = set_err(c, FSL_ERR_HIP, "download");
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(dX);
    (void)hipFree(dy);
    return rc;
}

int fsl_mlp_update_thickness(fsl_ctx* c) { return update_thickness(c); }

int fsl_mlp_get_thickness(fsl_ctx* c, float* dw)
{
    if (!c->lrn.dw_valid) return set_err(c, FSL_ERR_ARG, "thickness not computed");
    HIPCHK(c, hipMemcpyAsync(dw, c->lrn.dw, (c->g.L + 2) * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSL_OK;
}

int fsl_mlp_set_thickness(fsl_ctx* c, const float* dw)
{
    int rc = ensure_diag(c);
    if (rc) return rc;
    for (int i = 0; i < c->g.L + 2; i++)
        if (!(dw[i] > 0)) return set_err(c, FSL_ERR_ARG, "thickness %d must be positive", i);
    HIPCHK(c, hipMemcpyAsync(c->lrn.dw, dw, (c->g.L + 2) * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->lrn.dw_valid = 1;
    return FSL_OK;
}

int fsl_plane_profiles(fsl_ctx* c, double* out)
{
    int rc = ensure_diag(c);
    if (rc) return rc;
    prof_mark(c, K_DIAG, 0);
    launch_diag_zsum(c->stream, cst5(c->U, c->g.G), c->g, c->lrn.zs);
    prof_mark(c, K_DIAG, 1);
    if ((rc = check_launch(c, "plane profiles"))) return rc;
    HIPCHK(c, hipMemcpyAsync(out, c->lrn.zs, 7 * (size_t)c->g.L * c->g.H * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSL_OK;
}

int fsl_mlp_features(fsl_ctx* c, long long stride, long long offset, long long n, float* X)
{
    const long long N = (long long)c->g.L * c->g.H * c->g.D;
    if (stride < 1 || offset < 0 || n < 0 || (n > 0 && offset + (n - 1) * stride >= N))
        return set_err(c, FSL_ERR_ARG, "sample range outside the interior");
    if (n == 0) return FSL_OK;
    int rc;
    if (!c->lrn.dw_valid && (rc = update_thickness(c))) return rc;
    float* dX = nullptr;
    HIPCHK(c, hipMalloc(&dX, n * 8 * sizeof(float)));
    launch_mlp_features(c->stream, cst5(c->U, c->g.G), c->g, c->ph, c->dU, c->lrn.dw, c->cs.deltaY, nullptr,
                        stride, offset, dX, nullptr, n);
    rc = check_launch(c, "features");
    if (!rc && (hipMemcpyAsync(X, dX, n * 8 * sizeof(float), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess))
        rc = set_err(c, FSL_ERR_HIP, "download");
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(dX);
    return rc;
}

int fsl_learned_smagorinsky(fsl_ctx* c)
{
    int rc = need_ready(c);
    if (rc) return rc;
    rc = learned_mu(c, c->mu);
    return rc;
}

int fsl_mlp_reserve(fsl_ctx* c, long long cap)
{
    LearnedState& L = c->lrn;
    if (cap < 0 || cap > (1LL << 31)) return set_err(c, FSL_ERR_ARG, "capacity %lld", cap);
    if (cap <= L.cap) return FSL_OK;
    float *X = nullptr, *t = nullptr;
    HIPCHK(c, hipMalloc(&X, cap * 8 * sizeof(float)));
    if (hipMalloc(&t, cap * sizeof(float)) != hipSuccess) {
        (void)hipFree(X);
        return set_err(c, FSL_ERR_NOMEM, "sample store of %lld", cap);
    }
    if (L.n > 0) {
        HIPCHK(c, hipMemcpyAsync(X, L.X, L.n * 8 * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(t, L.t, L.n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(L.X);
    (void)hipFree(L.t);
    c->bytes += (size_t)(cap - L.cap) * 9 * sizeof(float);
    L.X = X;
    L.t = t;
    L.cap = cap;
    return FSL_OK;
}

long long fsl_mlp_sample_count(const fsl_ctx* c) { return c->lrn.n; }

int fsl_mlp_clear_samples(fsl_ctx* c)
{
    c->lrn.n = 0;
    return FSL_OK;
}

int fsl_mlp_collect(fsl_ctx* c, long long stride, long long offset, long long* n_total)
{
    LearnedState& L = c->lrn;
    const long long N = (long long)c->g.L * c->g.H * c->g.D;
    if (stride < 1 || offset < 0 || offset >= N) return set_err(c, FSL_ERR_ARG, "stride %lld offset %lld", stride, offset);
    if (!(c->dU > 0)) return set_err(c, FSL_ERR_ARG, "features need two distinct inflow streams");
    long long nn = (N - offset + stride - 1) / stride;
    if (nn > L.cap - L.n) nn = L.cap - L.n;
    int rc;
    if ((rc = ensure_cd(c))) return rc;
    // the model reads the ghost layer of the step state
    if ((rc = ctx_ghost(c, c->U))) return rc;
    if ((rc = update_thickness(c))) return rc;
    if ((rc = dynamic_mu(c, L.cd))) return rc;
    if (nn > 0) {
        launch_mlp_features(c->stream, cst5(c->U, c->g.G), c->g, c->ph, c->dU, L.dw, c->cs.deltaY, L.cd, stride,
                            offset, L.X + L.n * 8, L.t + L.n, nn);
        if ((rc = check_launch(c, "features"))) return rc;
        L.n += nn;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (n_total) *n_total = L.n;
    return FSL_OK;
}

int fsl_mlp_samples(fsl_ctx* c, long long first, long long n, float* X, float* t)
{
    LearnedState& L = c->lrn;
    if (first < 0 || n < 0 || first + n > L.n) return set_err(c, FSL_ERR_ARG, "sample range");
    if (n == 0) return FSL_OK;
    if (X) HIPCHK(c, hipMemcpyAsync(X, L.X + first * 8, n * 8 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (t) HIPCHK(c, hipMemcpyAsync(t, L.t + first, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSL_OK;
}

int fsl_mlp_load_samples(fsl_ctx* c, const float* X, const float* t, long long n)
{
    int rc = fsl_mlp_reserve(c, n);
    if (rc) return rc;
    LearnedState& L = c->lrn;
    if (n > 0) {
        HIPCHK(c, hipMemcpyAsync(L.X, X, n * 8 * sizeof(float), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(L.t, t, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    L.n = n;
    return FSL_OK;
}

// loss weight (|S| Delta / dU)^power of each sample, normalised to mean 1 over the store
int fsl_mlp_set_loss_weight(fsl_ctx* c, float power)
{
    LearnedState& L = c->lrn;
    if (!(power >= 0.f) || power > 8.f) return set_err(c, FSL_ERR_ARG, "weight power %g", (double)power);
    if (power == 0.f) {
        L.wpow = 0.f;
        L.winv = 1.f;
        return FSL_OK;
    }
    if (L.n < 1) return set_err(c, FSL_ERR_ARG, "need samples");
    const long long chunk = 1 << 20;
    std::vector<float> buf((size_t)chunk * 8);
    double sw = 0;
    for (long long f = 0; f < L.n; f += chunk) {
        const long long m = std::min(chunk, L.n - f);
        int rc = fsl_mlp_samples(c, f, m, buf.data(), nullptr);
        if (rc) return rc;
        for (long long r = 0; r < m; r++) sw += pow(std::max((double)buf[r * 8 + 6], 0.), (double)power);
    }
    if (!(sw > 0)) return set_err(c, FSL_ERR_ARG, "all weights vanish");
    L.wpow = power;
    L.winv = (float)(L.n / sw);
    return FSL_OK;
}

// feature means and reciprocal standard deviations of the stored samples
int fsl_mlp_fit_norm(fsl_ctx* c, float* norm)
{
    LearnedState& L = c->lrn;
    if (L.n < 2) return set_err(c, FSL_ERR_ARG, "need samples");
    const long long chunk = 1 << 20;
    std::vector<float> buf((size_t)chunk * 8);
    double s1[8] = {0}, s2[8] = {0};
    for (long long f = 0; f < L.n; f += chunk) {
        const long long m = std::min(chunk, L.n - f);
        int rc = fsl_mlp_samples(c, f, m, buf.data(), nullptr);
        if (rc) return rc;
        for (long long r = 0; r < m; r++)
            for (int q = 0; q < 8; q++) {
                const double v = buf[r * 8 + q];
                s1[q] += v;
                s2[q] += v * v;
            }
    }
    for (int q = 0; q < 8; q++) {
        const double mean = s1[q] / L.n;
        const double var = std::max(s2[q] / L.n - mean * mean, 0.);
        norm[q] = (float)mean;
        norm[8 + q] = (float)(var > 1e-20 ? 1. / sqrt(var) : 1.);
    }
    return FSL_OK;
}

int fsl_mlp_train(fsl_ctx* c, int epochs, long long batch, float lr, unsigned long long seed, float* epoch_loss)
{
    LearnedState& L = c->lrn;
    int rc = FSL_OK;
    if (!L.ready) return set_err(c, FSL_ERR_ARG, "no parameters");
    if (epochs < 0 || batch < 256 || batch % 256 || batch > L.n || !(lr > 0))
        return set_err(c, FSL_ERR_ARG, "epochs %d batch %lld (multiple of 256, <= %lld samples) lr %g", epochs,
                       batch, L.n, (double)lr);
    if ((rc = ensure_train(c))) return rc;
    const unsigned long long n = (unsigned long long)L.n;
    const long long steps = L.n / batch;
    for (int e = 0; e < epochs; e++) {
        // sample order of the epoch: idx -> (a idx + b) mod n with gcd(a, n) = 1
        unsigned long long a = (mix64(seed * 2 + 2 * e + 1) % n) | 1ULL, r = 1;
        while (gcd64(a, n) != 1) a = (mix64(seed + 977 * (++r) + e) % n) | 1ULL;
        const unsigned long long b = mix64(seed ^ (0xA5A5ULL + e)) % n;
        prof_mark(c, K_MLP_TRAIN, 0);
        for (long long s = 0; s < steps; s++) {
            launch_mlp_train_step(c->stream, L.X, L.t, L.n, batch, s, a, b, L.pack, L.master, L.norm, L.grad,
                                  L.loss, L.wpow, L.winv, L.act);
            launch_mlp_adam(c->stream, L.master, L.grad, L.adam_m, L.adam_v, lr, ++L.adam_t);
            launch_mlp_pack(c->stream, L.master, L.pack, L.act);
        }
        prof_mark(c, K_MLP_TRAIN, 1);
        if ((rc = check_launch(c, "training"))) return rc;
        float loss = 0.f;
        HIPCHK(c, hipMemcpyAsync(&loss, L.loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemsetAsync(L.loss, 0, sizeof(float), c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (epoch_loss) epoch_loss[e] = loss / (float)(steps * batch);
        if (!isfinite(loss)) return set_err(c, FSL_ERR_NONFINITE, "training diverged in epoch %d", e);
    }
    return FSL_OK;
}

int fsl_mlp_gradient(fsl_ctx* c, long long batch, float* grad, float* loss)
{
    LearnedState& L = c->lrn;
    if (!L.ready) return set_err(c, FSL_ERR_ARG, "no parameters");
    if (batch < 256 || batch % 256 || batch > L.n) return set_err(c, FSL_ERR_ARG, "batch %lld", batch);
    int rc = ensure_train(c);
    if (rc) return rc;
    HIPCHK(c, hipMemsetAsync(L.grad, 0, mlp_nparam() * sizeof(float), c->stream));
    HIPCHK(c, hipMemsetAsync(L.loss, 0, sizeof(float), c->stream));
    // samples 0 .. batch-1 in order (permutation a = 1, b = 0)
    launch_mlp_train_step(c->stream, L.X, L.t, L.n, batch, 0, 1, 0, L.pack, L.master, L.norm, L.grad, L.loss,
                          L.wpow, L.winv, L.act);
    if ((rc = check_launch(c, "gradient"))) return rc;
    HIPCHK(c, hipMemcpyAsync(grad, L.grad, mlp_nparam() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (loss) HIPCHK(c, hipMemcpyAsync(loss, L.loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemsetAsync(L.grad, 0, mlp_nparam() * sizeof(float), c->stream));
    HIPCHK(c, hipMemsetAsync(L.loss, 0, sizeof(float), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSL_OK;
}

int fsl_learned_configure(fsl_ctx* c, int interval, double tol)
{
    if (interval < 0 || !(tol >= 0)) return set_err(c, FSL_ERR_ARG, "interval %d tol %g", interval, tol);
    c->lrn.interval = interval;
    c->lrn.tol = tol;
    c->lrn.in_fallback = 0;
    return FSL_OK;
}

int fsl_learned_set_tolerances(fsl_ctx* c, double profile_tol, double local_tol)
{
    if (!(profile_tol >= 0) || !(local_tol >= 0))
        return set_err(c, FSL_ERR_ARG, "tolerances %g %g", profile_tol, local_tol);
    c->lrn.profile_tol = profile_tol;
    c->lrn.local_tol = local_tol;
    return FSL_OK;
}

int fsl_learned_set_scale(fsl_ctx* c, double scale, int adapt)
{
    if (!(scale >= 0.1 && scale <= 10.)) return set_err(c, FSL_ERR_ARG, "scale %g outside [0.1, 10]", scale);
    if (adapt < 0 || adapt > 3) return set_err(c, FSL_ERR_ARG, "adapt %d (0 .. 3)", adapt);
    c->lrn.scale = (float)scale;
    c->lrn.adapt = adapt;
    return reset_levels(c);
}

long long fsl_learned_level_factors(const fsl_ctx* c, long long nmax, float* out)
{
    const LearnedState& L = c->lrn;
    const int Lx = c->g.L, H = c->g.H;
    const bool on = L.adapt >= 2 && L.fac;
    long long q = 0;
    for (int i = 1; i <= Lx; i++)
        for (int j = 1; j <= H; j++, q++)
            if (q < nmax) out[q] = on ? L.hfac[(size_t)i * (H + 2) + j] : 1.f;
    return (long long)Lx * H;
}

int fsl_learned_history(const fsl_ctx* c, int nmax, double* err, double* scale)
{
    return fsl_learned_history_errors(c, nmax, err, nullptr, nullptr, scale);
}

int fsl_learned_history_errors(const fsl_ctx* c, int nmax, double* err, double* profile_err, double* local_err,
                               double* scale)
{
    const LearnedState& L = c->lrn;
    const int n = (int)L.hist_err.size();
    for (int q = 0; q < n && q < nmax; q++) {
        if (err) err[q] = L.hist_err[q];
        if (profile_err) profile_err[q] = L.hist_perr[q];
        if (local_err) local_err[q] = L.hist_lerr[q];
        if (scale) scale[q] = L.hist_scale[q];
    }
    return n;
}

int fsl_learned_get_status(const fsl_ctx* c, fsl_learned_status* s)
{
    const LearnedState& L = c->lrn;
    s->scale = L.scale;
    s->adapt = L.adapt;
    s->interval = L.interval;
    s->tol = L.tol;
    s->in_fallback = L.in_fallback;
    s->last_error = L.last_error;
    s->profile_tol = L.profile_tol;
    s->local_tol = L.local_tol;
    s->last_profile_error = L.last_profile_error;
    s->last_local_error = L.last_local_error;
    s->judged = L.judged;
    s->fallbacks = L.fallbacks;
    s->learned_steps = L.learned_steps;
    s->dynamic_steps = L.dynamic_steps;
    return FSL_OK;
}

int fsl_learned_judge_errors(fsl_ctx* c, double* err, double* profile_err, double* local_err, double* d_dynamic,
                             double* d_learned)
{
    int rc = need_ready(c);
    if (rc) return rc;
    if ((rc = ensure_judge(c)) || (rc = update_thickness(c))) return rc;
    if ((rc = dynamic_mu(c, nullptr)) || (rc = learned_mu(c, c->lrn.mu_alt))) return rc;
    JudgeErr je;
    if ((rc = judge_error(c, c->mu, c->lrn.mu_alt, &je))) return rc;
    if (err) *err = je.err;
    if (profile_err) *profile_err = je.perr;
    if (local_err) *local_err = je.lerr;
    if (d_dynamic) *d_dynamic = je.da;
    if (d_learned) *d_learned = je.db;
    return FSL_OK;
}

int fsl_learned_judge(fsl_ctx* c, double* err, double* d_dynamic, double* d_learned)
{
    return fsl_learned_judge_errors(c, err, nullptr, nullptr, d_dynamic, d_learned);
}

int fsl_mfma_probe(fsl_ctx* c, const float* A, const float* B, float* D)
{
    float* d = nullptr;
    HIPCHK(c, hipMalloc(&d, 3 * 256 * sizeof(float)));
    int rc = FSL_OK;
    if (hipMemcpyAsync(d, A, 256 * sizeof(float), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(d + 256, B, 256 * sizeof(float), hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = set_err(c, FSL_ERR_HIP, "upload");
    if (!rc) {
        launch_mfma_probe(c->stream, d, d + 256, d + 512);
        rc = check_launch(c, "mfma probe");
    }
    if (!rc && (hipMemcpyAsync(D, d + 512, 256 * sizeof(float), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess))
        rc = set_err(c, FSL_ERR_HIP, "download");
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    return rc;
}

}  // extern "C"

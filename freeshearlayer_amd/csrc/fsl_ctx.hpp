// fsl_ctx.hpp -- the device context behind the C-ABI of include/fsl.h
// (private to the library: fsl_api.hip and fsl_learned.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "fsl.h"
#include "fsl_internal.hpp"

using fsl::Geo;
using fsl::Phys;
using fsl::Bnd;
using fsl::K_NKERNELS;

struct ProfRec {
    int id;
    hipEvent_t a, b;
};

// learned closure (FSL_SGS_LEARNED): parameters, sample store, judge
struct LearnedState {
    float* master = nullptr;   // FP32 parameters
    void* pack = nullptr;      // BF16 MFMA operands
    float* norm = nullptr;     // feature means, reciprocal std (16)
    float* grad = nullptr;     // training: gradient, Adam moments, loss
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    float* loss = nullptr;
    int adam_t = 0;
    float wpow = 0.f, winv = 1.f;   // loss weight (feature 6)^wpow, normalised to mean 1
    int ready = 0;             // parameters set
    int act = 0;               // hidden activation, FSL_MLP_TANH or FSL_MLP_RELU
    float* X = nullptr;        // sample store: features (8 per sample), targets
    float* t = nullptr;
    long long n = 0, cap = 0;
    float* cd = nullptr;       // dynamic coefficient (G floats), sample targets
    double* zs = nullptr;      // plane profiles, 7 L H
    float* dw = nullptr;       // delta_w(x), L + 2
    int dw_valid = 0;
    double* mu_alt = nullptr;  // second mu_SGS array for the judge
    double* part = nullptr;    // dissipation partials per x band and row (5 sums)
    double* hpart = nullptr;   // pinned host copy of the partials
    int interval = 50;         // judge every interval steps (0: never)
    double tol = 0.10;         // relative dissipation error that triggers the fallback
    int in_fallback = 0;
    int adapt = 2;             // judge rescales the network's Cd to the dynamic model's dissipation:
                               // 1 one global level, 2 a level per y row, 3 per x band and y row
    float scale = 1.f;         // current output scale of the network's Cd
    float* fac = nullptr;      // adapt >= 2: factors of the scale per (plane i, row j), (L + 2)(H + 2) (device)
    std::vector<float> hfac;   //   and their host copy
    int nband = 1;             // x bands of the level factors (1 for adapt 2)
    std::vector<float> band_f; // factor per (band, row j = 1 .. H)
    std::vector<double> jrows; // last judge's per-row dissipation sums: dynamic (H), network (H)
    double last_error = -1.;
    double profile_tol = 1e30;       // bounds on the judge's profile and local errors (judge_error)
    double local_tol = 1e30;
    double last_profile_error = -1.;
    double last_local_error = -1.;
    std::vector<float> hist_err, hist_perr, hist_lerr, hist_scale;   // per judge: errors, scale in use
    long long judged = 0, fallbacks = 0, learned_steps = 0, dynamic_steps = 0;
};

struct fsl_ctx {
    fsl_case cs;
    fsl_config cfg;
    Geo g;
    Phys ph;
    Bnd bnd;
    size_t nface[3];
    hipStream_t stream;

    double* U;          // 5 G
    double* Up;         // 5 G
    double* Uq;         // 5 G, fused path only
    double* faces;      // 10 (NX + NY + NZ), unfused path only
    double* mu;         // G
    double* sgs;        // scratch for the FP32 dynamic model (mixed contexts)
    double* filt;       // scratch of fsl_filter (in, out and two pass buffers)
    double* u_in;       // HIG
    double* sums;       // partials + 5 results
    unsigned long long* red;
    unsigned long long* red_host;   // pinned copy of the CFL maximum (hipHostMalloc)
    double* probe_buf;  // 6 HIG probe values (allocated on first use)
    int sgs_avg;        // FSL_SGS_AVG_*: averaging of the dynamic coefficient
    double* avg_buf;    // its terms and sums (sgs_avg_bytes; allocated on first use)
    size_t bytes;

    int cur_is_p;       // stage input: U (0) or Up (1)
    double dt, dtX, dtY, dtZ;
    fsl_block blk;
    unsigned step;
    int cfl_ready;      // c->red holds the CFL maximum of U (from fused stage 3) for dtX
    int cfl_zeroed;     // a fused stage 2 has cleared c->red for the next stage 3

    int prof;
    int prof_open_id;
    hipEvent_t prof_open_ev;
    std::vector<ProfRec> pending;
    std::vector<hipEvent_t> pool;
    double prof_ms[K_NKERNELS];
    long long prof_cnt[K_NKERNELS];

    double dU;          // velocity difference of the inflow streams
    double dw0;         // vorticity thickness of the inflow profile
    LearnedState lrn;

    char err[512];
};


namespace fsl {
int set_err(fsl_ctx* c, int code, const char* fmt, ...);
void prof_mark(void* arg, int id, int end);
State5 st5(double* base, long long G);
CState5 cst5(const double* base, long long G);
int ensure_sgs(fsl_ctx* c);
// the FP32 two-kernel form of the dynamic model runs (mixed contexts with FSL_SGS_FP32=1)
bool sgs_fp32_form(const fsl_ctx* c);
// the averaging of the dynamic model in force (null: pointwise)
const SgsAvg* sgs_avg_of(fsl_ctx* c);
int check_launch(fsl_ctx* c, const char* what);
Bnd bnd_now(fsl_ctx* c);
int ctx_ghost(fsl_ctx* c, const double* state, bool z = true);
// learned closure (fsl_learned.hip)
int learned_sgs_step(fsl_ctx* c);   // mu_SGS of the step from U (call once per step), judge, fallback
void learned_free(fsl_ctx* c);
}  // namespace fsl

#define HIPCHK(ctx, call)                                                                        \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return set_err((ctx), FSL_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// fsl_internal.hpp -- launch-level interface between the kernel translation
// units and the context layer (fsl_api.hip).  Not part of the C-ABI.
#pragma once

#include "fsl.h"
#include "fsl_device.hpp"

namespace fsl {

// The 30 face arrays of the unfused path, in the id order of the C-ABI:
// 0..4 xU*, 5..9 U*x, 10..14 yU*, 15..19 U*y, 20..24 zU*, 25..29 U*z.
struct Faces {
    double* a[30];
};

// Kernel ids for the event profiler (names in fsl_api.hip)
enum KernelId {
    K_GHOST = 0,
    K_RECON,
    K_IFACE,
    K_FLUX_X,
    K_FLUX_Y,
    K_FLUX_Z,
    K_RK,
    K_CFL,
    K_SGS_PRIM,
    K_FILT_Y,
    K_FILT_Z,
    K_FILT_X,
    K_SGS_STRAIN,
    K_SGS_FINAL,
    K_FUSED_X,
    K_FUSED_Y,
    K_FUSED_Z,
    K_SUM,
    K_COPY,
    K_FUSED_XYZ,
    K_MLP_MU,
    K_MLP_TRAIN,
    K_DIAG,
    K_NKERNELS
};

// Time-step ratios deltaT_X, deltaT_Y, deltaT_Z
struct Dt3 {
    double x, y, z;
};

// --- unfused path (fsl_kernels.hip) ---
// ghost layer of U (src/boundCondInGhostCells.c:10-96); z = false leaves the
// z-periodic columns of the interior rows unwritten (for readers that take the
// periodic images k = D and k = 1 instead)
void launch_ghost_fill(hipStream_t s, State5 U, const Geo& g, const Bnd& b, bool z = true);
void launch_recon(hipStream_t s, CState5 U, const Faces& F, const Geo& g, int quirk);
void launch_iface_bc(hipStream_t s, const Faces& F, const Geo& g, const Bnd& b);
void launch_flux(hipStream_t s, int dir, CState5 U, const Faces& F, const double* mu, const Geo& g,
                 const Phys& ph);
// kind: 31, 32, 33 = three-stage update of stage 1..3; 21, 22 = two-stage
void launch_rk(hipStream_t s, int kind, State5 U, State5 Up, const Faces& F, const Geo& g, Dt3 dt);
// max-reduction of the Courant measure into *out (bits of a non-negative double);
// mode 0: (rho u / rho) dtX, mode 1: (|u|+c)/dx + (|v|+c)/dy + (|w|+c)/dz
void launch_cfl(hipStream_t s, int mode, CState5 U, const Geo& g, const Phys& ph, double dtX,
                unsigned long long* out);
// deterministic FP64 sums of the five fields over the interior
void launch_sums(hipStream_t s, CState5 U, const Geo& g, double* partial, int nblocks, double* out);
int sums_nblocks(const Geo& g);
void launch_copy(hipStream_t s, const double* a, double* b, size_t n);

// --- dynamic Smagorinsky (fsl_sgs.hip) ---
// Averaged coefficient (opt-in, not the reference's pointwise model): the
// model kernels write the Germano terms L:M and M:M of every interior cell to
// lm, mm and rho Delta^2 |S| to mu; the averaging kernels sum the terms over
// z lines (mode FSL_SGS_AVG_Z, sums per (i, j)) or x-z planes
// (FSL_SGS_AVG_XZ, per j) in a fixed order, form Cd = -1/2 sum(L:M) /
// (sum(M:M) + small) clipped to [0, 0.15], and scale mu by it.  sums holds
// 2 LEN HIG + 2 HIG doubles.  A null SgsAvg (or mode 0): pointwise Cd.
struct SgsAvg {
    int mode;
    double* lm;
    double* mm;
    double* sums;
};
size_t sgs_avg_bytes(const Geo& g);
void launch_sgs_average(hipStream_t s, const SgsAvg& av, double* mu, float* cd_out, const Geo& g);
// FP64: the reference's operations in one x-marching kernel without
// intermediate arrays (fsl_sgs_march.hip)
void launch_sgs(hipStream_t s, CState5 U, double* mu, const Geo& g, const Phys& ph,
                void (*prof)(void*, int, int), void* prof_arg, float* cd_out, const SgsAvg* avg = nullptr);
// bytes of scratch the FP32 model needs (16 float cell arrays)
size_t sgs_f32_scratch_bytes(const Geo& g);
void launch_sgs_march(hipStream_t s, CState5 U, double* mu, const Geo& g, const Phys& ph, float* cd_out,
                      double* lm = nullptr, double* mm = nullptr);
// the same model in FP32 (filters, Germano contraction) with two tiled
// kernels and 16 float scratch arrays; mu_SGS stays an FP64 array
void launch_sgs_f32(hipStream_t s, CState5 U, double* mu, void* scratch, const Geo& g, const Phys& ph,
                    void (*prof)(void*, int, int), void* prof_arg, float* cd_out, const SgsAvg* avg = nullptr);

// x segments of the fused stage kernels (fsl_fused.hip): count and length
int fused_segments(int L, int cols, int per_cu, int* seg_len);

// --- learned closure (fsl_mlp.hip) ---
int mlp_nparam();
int mlp_npack();   // packed BF16 operand entries (8 bytes each)
// act: FSL_MLP_TANH or FSL_MLP_RELU, the network's hidden activation
void launch_mlp_pack(hipStream_t s, const float* master, void* pack, int act);
void launch_mlp_forward(hipStream_t s, const float* X, long long n, const void* pack, const float* master,
                        const float* norm, float* y, int act);
// fac: factors of scale per (plane i, row j), index i (H + 2) + j, or null
// (all 1); fac_planes 0: every plane holds the same factors (plane 0's used)
void launch_mlp_mu(hipStream_t s, CState5 U, const Geo& g, const Phys& ph, double dU, const float* dw,
                   double dy, const void* pack, const float* master, const float* norm, double* mu,
                   float* cd_out, float scale, const float* fac, int fac_planes, int act);
void launch_mlp_features(hipStream_t s, CState5 U, const Geo& g, const Phys& ph, double dU, const float* dw,
                         double dy, const float* cd, long long stride, long long offset, float* X, float* t,
                         long long nout);
void launch_mlp_train_step(hipStream_t s, const float* X, const float* t, long long n, long long batch,
                           long long step, unsigned long long pa, unsigned long long pc, const void* pack,
                           const float* master, const float* norm, float* grad, float* loss, float wpow,
                           float winv, int act);
void launch_mlp_adam(hipStream_t s, float* master, float* grad, float* m, float* v, float lr, int t);
void launch_mfma_probe(hipStream_t s, const float* A, const float* B, float* D);
// per-row (y) sums over x band s of the SGS dissipation densities e_a, e_b
// of two mu_SGS fields and of e_a^2, e_b^2, e_a e_b:
// rows[(q NS + s) HIG + j - 1], q = 0..4, s = 0 .. NS - 1 with
// NS = dissipation_splits(g) = min(kDissBands, LEN) x bands
constexpr int kDissBands = 16;
int dissipation_splits(const Geo& g);
void launch_sgs_dissipation(hipStream_t s, CState5 U, const Geo& g, const Phys& ph, const double* mu_a,
                            const double* mu_b, double* rows);
void launch_diag_zsum(hipStream_t s, CState5 U, const Geo& g, double* out);   // 7 L H doubles
void launch_diag_dw(hipStream_t s, const double* zs, const Geo& g, double dU, double dy, float fallback,
                    float* dw);
// one box filter of a cell array (y, z, x passes)
void launch_filter(hipStream_t s, const double* in, double* out, double* t1, double* t2, const Geo& g,
                   const double h[3], void (*prof)(void*, int, int), void* prof_arg);

// Snapshot of a context's launch parameters for the fused-path translation
// units (fsl_api.hip builds it).
struct CtxView {
    hipStream_t stream;
    Geo g;
    Phys ph;
    Bnd bnd;
    double *U, *Up, *Uq, *mu;
    double dtX, dtY, dtZ;
    int cur_is_p;
    int sgs_model;
    int precision;
    unsigned long long* red;   // CFL reduction cell
    int cfl_mode;
};
void launch_fused_mixed(const CtxView& v, int stage, const double* in, double* out);
void launch_fused_fp64(const CtxView& v, int stage, const double* in, double* out);
void launch_fused_fp64_fast(const CtxView& v, int stage, const double* in, double* out);

}  // namespace fsl

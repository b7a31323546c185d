// fsl_output.hip -- the per-cell output fields and the probe columns of the
// step state, computed on the device for the host writers (host/fsl_files.c).
//
//   k_output_fields  the nine computed columns of the reference's Tecplot
//                    frame (src/output.c:88-136): rho, u, v, w, p, T, vorticity
//                    magnitude, Q and mu_T / mu_L, in FP64 and in the
//                    reference's association order (FP contraction is off in
//                    this build), so the formatted text equals the reference's
//                    for the same state.  Rows in the frame's point order: i
//                    fastest, then j, then k.
//   k_probes         rho and rho u along j at the three x stations of
//                    src/probes.c:20-40 (i = LEN/2, 3 LEN/4, LEN) on the
//                    mid-z plane k = DEP/2 + 1.
#include <hip/hip_runtime.h>

#include "fsl_ctx.hpp"

using namespace fsl;

namespace {

constexpr int kOutBlock = 256;
constexpr int kOutCols = 9;

__global__ void __launch_bounds__(kOutBlock) k_output_fields(CState5 U, Geo g, Phys ph, int k0, int nk,
                                                              double* out)
{
    const long long n = (long long)nk * g.H * g.L;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % g.L) + 1;
        const int j = (int)((t / g.L) % g.H) + 1;
        const int k = (int)(t / ((long long)g.L * g.H)) + k0;
        const long long c = cidx(g, i, j, k);
        const long long xp = c + g.si, xm = c - g.si, yp = c + g.sj, ym = c - g.sj, zp = c + 1, zm = c - 1;
        const double* r = U.f[0];
        const double* mu = U.f[1];
        const double* mv = U.f[2];
        const double* mw = U.f[3];
        const double R = r[c];
        const double Uc = mu[c] / R;
        const double Vc = mv[c] / R;
        const double Wc = mw[c] / R;
        const double P = (U.f[4][c] - 0.5 * R * (Uc * Uc + Vc * Vc + Wc * Wc)) * kK_1;
        const double T = P / (kRair * R);
        // src/output.c:105-117: central differences of u = (rho u) / rho
        const double du_dx = (mu[xp] / r[xp] - mu[xm] / r[xm]) * ph.r2del[0];
        const double du_dy = (mu[yp] / r[yp] - mu[ym] / r[ym]) * ph.r2del[1];
        const double du_dz = (mu[zp] / r[zp] - mu[zm] / r[zm]) * ph.r2del[2];
        const double dv_dx = (mv[xp] / r[xp] - mv[xm] / r[xm]) * ph.r2del[0];
        const double dv_dy = (mv[yp] / r[yp] - mv[ym] / r[ym]) * ph.r2del[1];
        const double dv_dz = (mv[zp] / r[zp] - mv[zm] / r[zm]) * ph.r2del[2];
        const double dw_dx = (mw[xp] / r[xp] - mw[xm] / r[xm]) * ph.r2del[0];
        const double dw_dy = (mw[yp] / r[yp] - mw[ym] / r[ym]) * ph.r2del[1];
        const double dw_dz = (mw[zp] / r[zp] - mw[zm] / r[zm]) * ph.r2del[2];
        const double ox = dw_dy - dv_dz;
        const double oy = du_dz - dw_dx;
        const double oz = dv_dx - du_dy;
        const double Omega = sqrt(ox * ox + oy * oy + oz * oz);
        const double S12 = 0.5 * (du_dy + dv_dx);
        const double S13 = 0.5 * (du_dz + dw_dx);
        const double S23 = 0.5 * (dv_dz + dw_dy);
        const double Strain = sqrt(2 * (du_dx * du_dx + dv_dy * dv_dy + dw_dz * dw_dz +
                                        2 * (S12 * S12 + S13 * S13 + S23 * S23)));
        const double Q = 0.5 * (Omega * Omega - Strain * Strain);
        // src/output.c:134: the constant-coefficient ratio in every SGS mode
        const double muT = R * ph.CsDD * Strain / ph.mu_L;
        double* o = out + t * kOutCols;
        o[0] = R; o[1] = Uc; o[2] = Vc; o[3] = Wc; o[4] = P; o[5] = T; o[6] = Omega; o[7] = Q; o[8] = muT;
    }
}

__global__ void k_probes(CState5 U, Geo g, double* out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 6 * g.H) return;
    const int q = t / g.H, j = t % g.H + 1;
    const int st = q >> 1;
    const int i = st == 0 ? g.L / 2 : st == 1 ? 3 * g.L / 4 : g.L;
    out[t] = U.f[q & 1][cidx(g, i, j, g.D / 2 + 1)];
}

}  // namespace

extern "C" {

int fsl_output_fields(fsl_ctx* c, int k0, int nk, double* out)
{
    if (k0 < 1 || nk < 1 || k0 + nk - 1 > c->g.D)
        return set_err(c, FSL_ERR_ARG, "output planes %d..%d outside 1..%d", k0, k0 + nk - 1, c->g.D);
    const size_t n = (size_t)nk * c->g.H * c->g.L;
    double* d = nullptr;
    HIPCHK(c, hipMalloc(&d, n * kOutCols * sizeof(double)));
    const long long want = ((long long)n + kOutBlock - 1) / kOutBlock;
    const int grid = (int)(want < 65536 ? want : 65536);
    k_output_fields<<<grid, kOutBlock, 0, c->stream>>>(cst5(c->U, c->g.G), c->g, c->ph, k0, nk, d);
    int rc = check_launch(c, "output fields");
    if (rc == FSL_OK) {
        const hipError_t e = hipMemcpyAsync(out, d, n * kOutCols * sizeof(double), hipMemcpyDeviceToHost,
                                            c->stream);
        if (e != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
            rc = set_err(c, FSL_ERR_HIP, "output fields download failed");
    }
    (void)hipFree(d);
    return rc;
}

int fsl_probe_values(fsl_ctx* c, double* out)
{
    const int n = 6 * c->g.H;
    if (!c->probe_buf) {   // kept for the run: the probes are read every step
        HIPCHK(c, hipMalloc(&c->probe_buf, n * sizeof(double)));
        c->bytes += n * sizeof(double);
    }
    k_probes<<<(n + 255) / 256, 256, 0, c->stream>>>(cst5(c->U, c->g.G), c->g, c->probe_buf);
    int rc = check_launch(c, "probes");
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(out, c->probe_buf, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSL_OK;
}

}  // extern "C"

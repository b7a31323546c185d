// fsl_mlp.hip -- learned SGS closure: a per-cell MLP (8 -> 64 -> 64 -> 1,
// tanh or ReLU hidden layers) that predicts the dynamic-Smagorinsky coefficient Cd = Cs^2 from
// normalised invariants of the resolved strain and rotation, a grid-scale
// roughness ratio, |S| Delta and the distance from the layer centre,
// replacing the test filter and the
// Germano-Lilly least squares of src/turbulence.c:103-507.  mu_SGS is then
// rho Cd Delta^2 |S| as in src/turbulence.c:442.
//
// All dense contractions run on the matrix cores: v_mfma_f32_16x16x16_bf16
// (BF16 operands, FP32 accumulation).  One wavefront pushes 64 cells through
// the network as 4 x 4 tiles of 16 x 16; activations move between layers
// through LDS in BF16.  Training uses the same tiles for the forward pass,
// the transposed products of the backward pass and the weight gradients,
// which are reduced with FP32 atomics and applied by an Adam step.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include <algorithm>

#include "fsl_internal.hpp"

namespace fsl {

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int NIN = 8, NH = 64;
constexpr int KP1 = 16;   // layer-1 K padded to the MFMA depth

// parameter offsets in the FP32 master vector
constexpr int OW1 = 0, OB1 = OW1 + NIN * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH, NPARAM = OB3 + 1;
static_assert(NPARAM == 4801, "parameter count");

// packed BF16 B operands, lane order (16x16x16 MFMA: lane l holds
// B[k = 4 (l / 16) + e][n = l % 16], e = 0..3)
constexpr int PB1 = 0;                  // W1: 4 n-tiles
constexpr int PB2 = PB1 + 4 * 64;       // W2: 4 k-steps x 4 n-tiles
constexpr int PB2T = PB2 + 16 * 64;     // W2 transposed (backward): 4 k-steps x 4 n-tiles
constexpr int PB3 = PB2T + 16 * 64;     // w3 in column 0: 4 k-steps
// BF16 operands of the inference kernels (K = 2 log2 e, the tanh argument
// scale; the hidden layers carry r = 1 / (2^u + 1), tanh = 1 - 2 r):
constexpr int PH1 = PB3 + 4 * 64;       // K W1
constexpr int PH2 = PH1 + 4 * 64;       // -2 K W2, 16x16x32 operands (below)
constexpr int PH3 = PH2 + 16 * 64;      // -2 w3 in column 0, 16x16x32 operands
// Layers 2 and 3 use v_mfma_f32_16x16x32_bf16: the operand of K slab s (input
// units 32 s .. 32 s + 31), output tile t and lane l is the pair of 16x16x16
// operands of K tiles 2 s and 2 s + 1 (units 32 s + 4 (l / 16) + e and
// 32 s + 16 + 4 (l / 16) + e), 16 bytes: the A and B elements an MFMA pairs
// share lane group and element, so this unit order is the product's K order,
// and the previous layer's two output tiles concatenate into the B operand.
constexpr int PHF = PH3 + 4 * 64;       // FP32: K (b2 + sum_k W2[k][n]) (64), b3 + sum_k w3[k] (1)
constexpr int NPACK = PHF + 64;

typedef __bf16 bfv2 __attribute__((ext_vector_type(2)));
typedef float f2 __attribute__((ext_vector_type(2)));
typedef unsigned int u2 __attribute__((ext_vector_type(2)));

// FP32 -> BF16, round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ unsigned short f2bf(float f) { return __builtin_bit_cast(unsigned short, (__bf16)f); }

// Operands of the inference products: BF16 (as in training), FP32
// accumulation.
typedef __bf16 iv4 __attribute__((ext_vector_type(4)));
typedef __bf16 iv2 __attribute__((ext_vector_type(2)));
typedef short s4v __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even, two values per v_cvt_pk_bf16_f32
__device__ __forceinline__ unsigned cvt2i(float a, float b)
{
    return __builtin_bit_cast(unsigned, __builtin_convertvector((f2){a, b}, iv2));
}

__device__ __forceinline__ iv4 cvt4i(float a, float b, float c, float d)
{
    return __builtin_bit_cast(iv4, (u2){cvt2i(a, b), cvt2i(c, d)});
}

__device__ __forceinline__ unsigned short f2i(float f) { return __builtin_bit_cast(unsigned short, (__bf16)f); }

__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ f4 mfma(bf4 a, bf4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
__device__ __forceinline__ f4 mfma_i(iv4 a, iv4 b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(bf4, a), __builtin_bit_cast(bf4, b), c, 0, 0, 0);
}
typedef __bf16 iv8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ f4 mfma_i32(iv8 a, iv8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
__device__ __forceinline__ iv8 cat8(iv4 a, iv4 b) { return __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7); }
// 16x16x32 operand (K slab s, tile t) of lane l in a packed region (two 8-byte units)
__device__ __forceinline__ iv8 pack8(const bf4* region, int s, int t, int ntiles, int lane)
{
    return *reinterpret_cast<const iv8*>(region + 2 * ((s * ntiles + t) * 64 + lane));
}
__device__ __forceinline__ iv4 as_i4(bf4 v) { return __builtin_bit_cast(iv4, v); }

// tanh(z + b) = 1 - 2 / (2^(K (z + b)) + 1), K = 2 log2(e), with kb = K b:
// one FMA for the scaled argument, v_exp_f32, v_rcp_f32 and one FMA.  No
// clamp: 2^u saturates to +inf or 0 and the reciprocal maps those to +-1.
// BF16 activations keep 8 bits, far coarser than this evaluation's error.
constexpr float kTanhK = 2.8853900817779268f;
__device__ __forceinline__ float act(float z, float kb)
{
    const float e = __builtin_amdgcn_exp2f(__builtin_fmaf(z, kTanhK, kb));
    return __builtin_fmaf(-2.f, __builtin_amdgcn_rcpf(e + 1.f), 1.f);
}

// Inference activation r = 1 / (2^u + 1) of the scaled argument
// u = K (z + b) (tanh(z + b) = 1 - 2 r): v_exp_f32, an add, v_rcp_f32.  The
// inference kernels multiply BF16 operands: the features, K W1 onto K b1;
// r1 and -2 K W2 onto K (b2 + sum_k W2[k][.]); r2 and -2 w3 onto
// b3 + sum_k w3[k] -- the network's products with h = 1 - 2 r.  BF16 keeps
// r to 2^-9 relative: h to 2e-3 near +1 (as the BF16 h of training) and to
// 4e-3 near -1.
__device__ __forceinline__ float act_r(float u)
{
    return __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(u) + 1.f);
}

// ReLU networks (ACT = FSL_MLP_RELU): h = max(z + b, 0).  The inference form
// multiplies BF16 operands W1, W2, w3 onto accumulators that start at b1,
// b2, b3, and takes the maximum after the conversion to BF16 (max and the
// rounding commute), so its hidden values are those of the training form
// up to the accumulation order.
constexpr int kTanh = 0, kRelu = 1;
// max(x, 0) of four BF16 values as a signed 16-bit maximum of their bit
// patterns (a negative BF16 value, -0 included, is a negative integer): one
// v_pk_max_i16 per two units
__device__ __forceinline__ iv4 relu4(iv4 v)
{
    return __builtin_bit_cast(iv4, __builtin_elementwise_max(__builtin_bit_cast(s4v, v), (s4v){0, 0, 0, 0}));
}

// BF16 bits of element e of v (through the packed 32-bit words)
__device__ __forceinline__ unsigned short i4_bits(iv4 v, int e)
{
    const u2 w = __builtin_bit_cast(u2, v);
    return (unsigned short)((e & 1) ? (w[e >> 1] >> 16) : (w[e >> 1] & 0xffffu));
}

// activation of the training form (BF16 tiles, FP32 accumulator z, bias b)
template <int ACT>
__device__ __forceinline__ float act_train(float z, float b)
{
    if (ACT == kRelu) return fmaxf(z + b, 0.f);
    return act(z, kTanhK * b);
}

// derivative of the activation from its value h
template <int ACT>
__device__ __forceinline__ float act_deriv(float h)
{
    if (ACT == kRelu) return h > 0.f ? 1.f : 0.f;
    return 1.f - h * h;
}

// bias scale of the inference form's first-layer accumulator
template <int ACT>
__device__ __forceinline__ float infer_b1(float b)
{
    return ACT == kRelu ? b : kTanhK * b;
}

// four hidden values of the inference form from an output tile's accumulators
template <int ACT>
__device__ __forceinline__ iv4 act4(const f4 z)
{
    if (ACT == kRelu) return relu4(cvt4i(z[0], z[1], z[2], z[3]));
    return cvt4i(act_r(z[0]), act_r(z[1]), act_r(z[2]), act_r(z[3]));
}

// Each wavefront of the inference kernel stages its own 64 cells' features in
// its own LDS slice and reads them back across its lanes: ordering the
// wave's LDS accesses suffices (s_waitcnt lgkmcnt(0)); no workgroup barrier,
// so the block's waves run independently.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

struct MlpW {
    const bf4* pack;      // NPACK / 64 operand tiles, bf4 per lane
    const float* master;  // biases (and w3 for the backward pass)
    const float* norm;    // mean[8], 1/std[8]
};

// LDS workspace of one wavefront
struct WaveLds {
    unsigned short X[64 * KP1];
    unsigned short H1[64 * NH];
    unsigned short H2[64 * NH];
    float y[64];
};

__device__ __forceinline__ bf4 lds_row4(const unsigned short* p)
{
    const unsigned short* q = p;
    bf4 v;
    v.x = (short)q[0]; v.y = (short)q[1]; v.z = (short)q[2]; v.w = (short)q[3];
    return v;
}

// Forward pass of 64 samples held in ws.X (already normalised, BF16);
// leaves H1, H2 in LDS (BF16 h; for INFER BF16 r of tanh networks, BF16 h
// of ReLU networks) and the outputs in ws.y.  Training
// (INFER false) uses the BF16 tiles; the inference form matches
// k_mlp_mu_tr's arithmetic.  Caller syncs.
template <bool INFER, int ACT>
__device__ __forceinline__ void forward64(WaveLds& ws, const MlpW& w, int lane)
{
    const int r16 = lane & 15, q4 = lane >> 4;
    const float* cf = reinterpret_cast<const float*>(w.pack + PHF);
    // layer 1: X (64 x 16) W1 (16 x 64)
    for (int tm = 0; tm < 4; tm++) {
        const bf4 a = lds_row4(&ws.X[(tm * 16 + r16) * KP1 + 4 * q4]);
        for (int tn = 0; tn < 4; tn++) {
            const float b = w.master[OB1 + tn * 16 + r16], kb = infer_b1<ACT>(b);
            f4 acc;
            if (INFER) acc = mfma_i(as_i4(a), as_i4(w.pack[PH1 + tn * 64 + lane]), f4{kb, kb, kb, kb});
            else acc = mfma(a, w.pack[PB1 + tn * 64 + lane], f4{0.f, 0.f, 0.f, 0.f});
            const iv4 hi = act4<ACT>(acc);
#pragma unroll
            for (int e = 0; e < 4; e++)
                ws.H1[(tm * 16 + 4 * q4 + e) * NH + tn * 16 + r16] =
                    INFER ? i4_bits(hi, e) : f2bf(act_train<ACT>(acc[e], b));
        }
    }
    __syncthreads();
    // layer 2: H1 (64 x 64) W2 (64 x 64)
    for (int tm = 0; tm < 4; tm++) {
        f4 acc[4];
        for (int tn = 0; tn < 4; tn++) {
            const float c2 = cf[tn * 16 + r16];
            acc[tn] = INFER ? f4{c2, c2, c2, c2} : f4{0.f, 0.f, 0.f, 0.f};
        }
        if (INFER) {
            for (int sl = 0; sl < 2; sl++) {
                const iv8 a = cat8(as_i4(lds_row4(&ws.H1[(tm * 16 + r16) * NH + 32 * sl + 4 * q4])),
                                  as_i4(lds_row4(&ws.H1[(tm * 16 + r16) * NH + 32 * sl + 16 + 4 * q4])));
                for (int tn = 0; tn < 4; tn++) acc[tn] = mfma_i32(a, pack8(w.pack + PH2, sl, tn, 4, lane), acc[tn]);
            }
        } else {
            for (int ks = 0; ks < 4; ks++) {
                const bf4 a = lds_row4(&ws.H1[(tm * 16 + r16) * NH + ks * 16 + 4 * q4]);
                for (int tn = 0; tn < 4; tn++) acc[tn] = mfma(a, w.pack[PB2 + (ks * 4 + tn) * 64 + lane], acc[tn]);
            }
        }
        for (int tn = 0; tn < 4; tn++) {
            const float b = w.master[OB2 + tn * 16 + r16];
            const iv4 hi = act4<ACT>(acc[tn]);
#pragma unroll
            for (int e = 0; e < 4; e++)
                ws.H2[(tm * 16 + 4 * q4 + e) * NH + tn * 16 + r16] =
                    INFER ? i4_bits(hi, e) : f2bf(act_train<ACT>(acc[tn][e], b));
        }
    }
    __syncthreads();
    // layer 3: H2 (64 x 64) w3 (64 x 16, column 0)
    for (int tm = 0; tm < 4; tm++) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        if (INFER) {
            for (int sl = 0; sl < 2; sl++) {
                const iv8 a = cat8(as_i4(lds_row4(&ws.H2[(tm * 16 + r16) * NH + 32 * sl + 4 * q4])),
                                  as_i4(lds_row4(&ws.H2[(tm * 16 + r16) * NH + 32 * sl + 16 + 4 * q4])));
                acc = mfma_i32(a, pack8(w.pack + PH3, sl, 0, 1, lane), acc);
            }
        } else {
            for (int ks = 0; ks < 4; ks++) {
                const bf4 a = lds_row4(&ws.H2[(tm * 16 + r16) * NH + ks * 16 + 4 * q4]);
                acc = mfma(a, w.pack[PB3 + ks * 64 + lane], acc);
            }
        }
        if (r16 == 0) {
            const float b3 = INFER ? cf[64] : w.master[OB3];
#pragma unroll
            for (int e = 0; e < 4; e++) ws.y[tm * 16 + 4 * q4 + e] = acc[e] + b3;
        }
    }
    __syncthreads();
}

// ---- features ----
struct FeatArgs {
    CState5 U;
    Geo g;
    Phys ph;
    double dU;        // velocity difference of the two streams
    const float* dw;  // vorticity thickness delta_w(x) per x index (L + 2)
    double dy;
};

// Velocity u = (rho u) / rho of one stencil point in FP32: v_rcp_f32 of the
// density and one product per component, each rounded (this function is not
// under FP contraction), so every kernel that forms the features -- from a
// plane ring in LDS or from global memory -- sees the same velocity values.
__device__ __forceinline__ void point_velocity(double r, double m0, double m1, double m2, float u[3])
{
    const float rr = __builtin_amdgcn_rcpf((float)r);
    u[0] = (float)m0 * rr;
    u[1] = (float)m1 * rr;
    u[2] = (float)m2 * rr;
}

// Velocity at a cell and its six neighbours: up[e] / um[e] the neighbour at
// +1 / -1 along axis e (x, y, z).
struct Stencil {
    float uc[3], up[3][3], um[3][3];
};

// the stencil of cell c from global memory (strides st of x, y, z)
__device__ __forceinline__ void gather_stencil(const CState5& U, long long c, const long long st[3], Stencil& s)
{
    point_velocity(U.f[0][c], U.f[1][c], U.f[2][c], U.f[3][c], s.uc);
#pragma unroll
    for (int e = 0; e < 3; e++) {
        const long long p = c + st[e], m = c - st[e];
        point_velocity(U.f[0][p], U.f[1][p], U.f[2][p], U.f[3][p], s.up[e]);
        point_velocity(U.f[0][m], U.f[1][m], U.f[2][m], U.f[3][m], s.um[e]);
    }
}

// Central-difference velocity gradient gr[r][e] = du_r/dx_e and the
// second-difference Laplacian of u from a stencil (FP32, FMA contraction).
__device__ __forceinline__ void stencil_gradient(const Stencil& s, const Phys& ph, float gr[3][3], float lap[3])
{
#pragma clang fp contract(fast)
#pragma unroll
    for (int r = 0; r < 3; r++) lap[r] = 0.f;
#pragma unroll
    for (int e = 0; e < 3; e++) {
        const float h = (float)ph.r2del[e];
        const float iv2 = (float)(ph.rdel[e] * ph.rdel[e]);
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const float up = s.up[e][r], um = s.um[e][r];
            gr[r][e] = (up - um) * h;
            lap[r] += (up - 2.f * s.uc[r] + um) * iv2;
        }
    }
}

// S:S of the symmetric part of gr (the sum that |S| = sqrt(2 S:S) uses)
__device__ __forceinline__ float strain_sq(const float gr[3][3])
{
#pragma clang fp contract(fast)
    float s2 = 0.f;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int e = 0; e < 3; e++) {
            const float Sre = 0.5f * (gr[r][e] + gr[e][r]);
            s2 += Sre * Sre;
        }
    return s2;
}

// Features of interior cell (i, j, k) from the 7-point stencil of u =
// (rho u) / rho: invariants of the central-difference strain S and rotation
// W, the grid-scale roughness Delta |lap u| / |S| (second differences), |S|
// Delta / dU and y / delta_w(x).  Returns |S| = sqrt(2 S:S) (the model's
// strain magnitude, src/turbulence.c:299-301).
__device__ __forceinline__ float cell_features(const Stencil& st, const FeatArgs& a, int i, int j, float f[NIN])
{
    // FP32 with FMA contraction and the hardware square root (1 ulp): the
    // features reach the network rounded to BF16
#pragma clang fp contract(fast)
    const Geo& g = a.g;
    float gr[3][3], lap[3];
    stencil_gradient(st, a.ph, gr, lap);
    // Invariants from the six entries of the symmetric S and the vorticity
    // om (W_ij = -eps_ijk om_k / 2, so W^2 = (om om^T - |om|^2 I) / 4):
    //   |W|^2 = |om|^2 / 2,  tr(S W^2) = (om.S.om - |om|^2 tr S) / 4,
    //   tr(S^2 W^2) = (om.S^2.om - |om|^2 S:S) / 4.
    const float s00 = gr[0][0], s11 = gr[1][1], s22 = gr[2][2];
    const float s01 = 0.5f * (gr[0][1] + gr[1][0]), s02 = 0.5f * (gr[0][2] + gr[2][0]);
    const float s12 = 0.5f * (gr[1][2] + gr[2][1]);
    const float om[3] = {gr[2][1] - gr[1][2], gr[0][2] - gr[2][0], gr[1][0] - gr[0][1]};
    const float q00 = s00 * s00 + s01 * s01 + s02 * s02;   // S^2
    const float q11 = s01 * s01 + s11 * s11 + s12 * s12;
    const float q22 = s02 * s02 + s12 * s12 + s22 * s22;
    const float q01 = s00 * s01 + s01 * s11 + s02 * s12;
    const float q02 = s00 * s02 + s01 * s12 + s02 * s22;
    const float q12 = s01 * s02 + s11 * s12 + s12 * s22;
    const float s2 = q00 + q11 + q22;                         // S:S = tr S^2
    const float s3 = q00 * s00 + q11 * s11 + q22 * s22 + 2.f * (q01 * s01 + q02 * s02 + q12 * s12);
    const float oo = om[0] * om[0] + om[1] * om[1] + om[2] * om[2];
    const float w2 = 0.5f * oo;
    const float ws = s00 * om[0] * om[0] + s11 * om[1] * om[1] + s22 * om[2] * om[2] +
                     2.f * (s01 * om[0] * om[1] + s02 * om[0] * om[2] + s12 * om[1] * om[2]);
    const float wqw = q00 * om[0] * om[0] + q11 * om[1] * om[1] + q22 * om[2] * om[2] +
                      2.f * (q01 * om[0] * om[1] + q02 * om[0] * om[2] + q12 * om[1] * om[2]);
    const float sww = 0.25f * (ws - oo * (s00 + s11 + s22));
    const float ssww = 0.25f * (wqw - oo * s2);
    const float tiny = 1e-30f;
    const float smag = __builtin_amdgcn_sqrtf(2.f * s2);
    const float sn = __builtin_amdgcn_sqrtf(s2);
    const float yv = ((float)j - (float)(g.H / 2) - 0.5f) * (float)a.dy;
    // ratios through v_rcp_f32 (1 ulp): the network sees them rounded to BF16
    f[0] = (w2 - s2) * __builtin_amdgcn_rcpf(w2 + s2 + tiny);
    f[1] = s3 * __builtin_amdgcn_rcpf(s2 * sn + tiny);
    f[2] = sww * __builtin_amdgcn_rcpf(sn * w2 + tiny);
    f[3] = ssww * __builtin_amdgcn_rcpf(s2 * w2 + tiny);
    const float delta = __builtin_amdgcn_sqrtf((float)a.ph.DD);   // (dx dy dz)^(1/3)
    const float lapn = __builtin_amdgcn_sqrtf(lap[0] * lap[0] + lap[1] * lap[1] + lap[2] * lap[2]);
    f[4] = lapn * delta * __builtin_amdgcn_rcpf(smag + tiny);
    f[5] = ws * __builtin_amdgcn_rcpf(oo * smag + tiny);
    f[6] = smag * delta * (float)(1. / a.dU);
    f[7] = yv * __builtin_amdgcn_rcpf(a.dw[i]);
    return smag;
}

// interior cell t (k fastest) -> (i, j, k), 1-based
__device__ __forceinline__ void decode_interior(const Geo& g, long long t, int& i, int& j, int& k)
{
    k = (int)(t % g.D) + 1;
    j = (int)((t / g.D) % g.H) + 1;
    i = (int)(t / ((long long)g.D * g.H)) + 1;
}

// features (raw) and training target Cd / 0.15 of every `stride`-th interior cell
__global__ void __launch_bounds__(256) k_mlp_features(FeatArgs a, const float* cd, long long stride,
                                                      long long offset, float* X, float* t, long long nout)
{
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nout) return;
    const long long cell = offset + s * stride;
    int i, j, k;
    decode_interior(a.g, cell, i, j, k);
    float f[NIN];
    const long long st[3] = {a.g.si, a.g.sj, 1};
    Stencil sv;
    gather_stencil(a.U, cidx(a.g, i, j, k), st, sv);
    cell_features(sv, a, i, j, f);
#pragma unroll
    for (int q = 0; q < NIN; q++) X[s * NIN + q] = f[q];
    if (t) t[s] = cd[cidx(a.g, i, j, k)] / 0.15f;
}

// Learned mu_SGS, register-resident network on a plane march.  A workgroup
// (4 waves, one thread per cell) owns a 16 x 16 tile of (j, k) columns and
// marches along x over a segment of planes.  The velocity of every point of
// the tile and its one-cell ring (18 x 18) passes through a four-plane LDS
// ring, so each point's four fields are read from global memory once per
// plane (the ring's 68 points also by the neighbouring tiles, mostly from
// L2) and its velocity is formed once; a cell's 7-point stencil comes from
// the ring.  Each plane's loads are issued one iteration ahead.  One
// workgroup barrier per plane: plane i + 2 goes into the ring slot of plane
// i - 2, last read in iteration i - 1, before the barrier of iteration i.
//
// The network runs transposed: Z^T = W^T X^T with units as MFMA rows (M) and
// cells as columns (N).  The output tile D[unit][cell] then sits in the lanes
// exactly as the next layer's B operand B[unit][cell] (lane = cell + 16
// (unit / 4), element = unit % 4), so activations go from the accumulators
// through the activation (act4) and BF16 rounding straight into the next
// product.  The weight tiles are the A operands (same lane layout as the
// packed B tiles) and stay in registers for the whole march; only the 8
// features of each cell pass through LDS (one 8-byte read per lane and
// 16-cell tile).
constexpr int MT = 16;                   // tile edge (j and k)
constexpr int MW = MT + 2;               // tile and ring
constexpr int MFP = MW * MW;             // 324 ring points
constexpr int MHALO = MFP - MT * MT;     // 68 of them outside the tile
constexpr int MRING = 4;                 // planes in the LDS ring
static_assert(MT * MT == 256 && MHALO <= 256 && (MRING & (MRING - 1)) == 0, "mu kernel tiling");
// LDS layout of one field of one ring plane: columns kk = 1 .. MT of every
// ring row jj as rows of MT floats (a wave reads four consecutive rows at
// one offset: each fills its own quarter of the 64 banks), and the side
// columns kk = 0 and MW - 1 in blocks of 64 floats at the one bank of the
// row's quarter that a z-neighbour read leaves free (kk = MW - 1: the
// quarter's first bank, read with columns 2 .. MT; kk = 0: its last,
// read with columns 1 .. MT - 1), so no read of the stencil has a bank
// conflict.
constexpr int MFL = MW * MT + (MW + 3) / 4 * 64;   // 608 floats
__device__ __forceinline__ int ring_pos(int jj, int kk)
{
    if (kk >= 1 && kk <= MT) return jj * MT + kk - 1;
    const int side = MW * MT + (jj >> 2) * 64 + 16 * (jj & 3);
    return kk == 0 ? side + 15 : side;
}

constexpr int kRsrcWord3 = 0x00020000;   // gfx9 buffer resource, untyped access

__device__ __forceinline__ double bload(__amdgpu_buffer_rsrc_t r, unsigned boff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)boff, 0, 0));
}

struct MuArgs {
    FeatArgs f;
    MlpW w;
    double* mu;
    float* cd_out;
    float scale;
    const float* fac;    // factor of scale per (plane i, row j), index i fsi + j (or null)
    int fsi;             // plane stride of fac: H + 2, or 0 where the factors do not depend on x
    int segL, ntj, ntk, nblocks;
};

template <int TP, int ACT>   // 16-cell tiles in flight per layer; hidden activation
__global__ void __launch_bounds__(256, 3) k_mlp_mu_plane(MuArgs ma)
{
    __shared__ float s_u[MRING][3][MFL];   // velocity of the tile and ring, plane p in slot p % MRING
    __shared__ double s_rho[MRING][MT * MT];   // density of the tile's cells (mu_SGS = rho Cd Delta^2 |S|)
    __shared__ unsigned short sX[4][64 * NIN];
    const FeatArgs& a = ma.f;
    const MlpW& w = ma.w;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r16 = lane & 15, q4 = lane >> 4;
    iv4 A1[4];
    iv8 A2[2][4], A3[2];
    float kb1[4][4], kb2[4][4];
    const float* cf = reinterpret_cast<const float*>(w.pack + PHF);
#pragma unroll
    for (int t = 0; t < 4; t++) {
        A1[t] = as_i4(w.pack[PH1 + t * 64 + lane]);
        A2[0][t] = pack8(w.pack + PH2, 0, t, 4, lane);
        A2[1][t] = pack8(w.pack + PH2, 1, t, 4, lane);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            kb1[t][e] = infer_b1<ACT>(w.master[OB1 + t * 16 + 4 * q4 + e]);
            kb2[t][e] = cf[t * 16 + 4 * q4 + e];
        }
    }
#pragma unroll
    for (int sl = 0; sl < 2; sl++) A3[sl] = pack8(w.pack + PH3, sl, 0, 1, lane);
    const float b3 = cf[64];
    float nm[NIN], ns[NIN];
#pragma unroll
    for (int q = 0; q < NIN; q++) {
        nm[q] = w.norm[q];
        ns[q] = w.norm[NIN + q];
    }
    unsigned short* X = sX[wv];

    const Geo& g = a.g;
    const int L = g.L, H = g.H, D = g.D;
    // XCD-aware order: workgroups are dispatched round-robin over the 8 XCDs
    // (workgroup b on XCD b % 8), so XCD x takes a contiguous range of tiles,
    // whose shared ring points stay in its L2.  The grid is a multiple of 8.
    int bid = blockIdx.x;
    {
        const int per = gridDim.x >> 3;
        bid = (bid & 7) * per + (bid >> 3);
    }
    if (bid >= ma.nblocks) return;   // the whole workgroup, before any barrier
    const int tk = bid % ma.ntk;
    const int rem = bid / ma.ntk;
    const int tj = rem % ma.ntj;
    const int seg = rem / ma.ntj;
    const int j0 = 1 + tj * MT, k0 = 1 + tk * MT;
    const int i0 = 1 + seg * ma.segL, i1 = min(i0 + ma.segL, L + 1);
    const int t = threadIdx.x, tr = t >> 4, tc = t & 15;
    const int j = j0 + tr, k = k0 + tc;
    const bool valid = j <= H && k <= D;
    const int po = ring_pos(tr + 1, tc + 1);   // own point in the ring
    const int pyp = ring_pos(tr + 2, tc + 1), pym = ring_pos(tr, tc + 1);   // y neighbours
    const int pzp = ring_pos(tr + 1, tc + 2), pzm = ring_pos(tr + 1, tc);   // z neighbours
    // ring point outside the tile of threads t < MHALO: rows 0 and MW - 1,
    // then columns 0 and MW - 1 of rows 1 .. MT
    int jj, kk;
    if (t < MW) {
        jj = 0;
        kk = t;
    } else if (t < 2 * MW) {
        jj = MW - 1;
        kk = t - MW;
    } else if (t < 2 * MW + MT) {
        jj = t - 2 * MW + 1;
        kk = 0;
    } else {
        jj = t - (2 * MW + MT) + 1;
        kk = MW - 1;
    }
    const bool hal = t < MHALO;
    const int ph = hal ? ring_pos(jj, kk) : po;
    // in-plane byte offsets (rows and columns beyond the ghost layer clamped:
    // only invalid lanes read those points)
    const unsigned sjb = (unsigned)g.sj * 8u;
    const unsigned oo = (unsigned)min(j, H + 1) * sjb + (unsigned)min(k, D + 1) * 8u;
    const unsigned oh = (unsigned)min(j0 - 1 + jj, H + 1) * sjb + (unsigned)min(k0 - 1 + kk, D + 1) * 8u;
    const int pb = (int)(g.si * 8);
    auto ld4 = [&](int ip, unsigned off, double q[4]) {
        const long long plane = (long long)min(ip, L + 1) * g.si;
#pragma unroll
        for (int f = 0; f < 4; f++)
            q[f] = bload(__builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(a.U.f[f] + plane), (short)0, pb,
                                                           kRsrcWord3),
                         off);
    };
    auto put = [&](int ip, const double qo[4], const double qh[4]) {
        const int sl = ip & (MRING - 1);
        float u[3];
        point_velocity(qo[0], qo[1], qo[2], qo[3], u);
#pragma unroll
        for (int e = 0; e < 3; e++) s_u[sl][e][po] = u[e];
        s_rho[sl][t] = qo[0];
        if (hal) {
            point_velocity(qh[0], qh[1], qh[2], qh[3], u);
#pragma unroll
            for (int e = 0; e < 3; e++) s_u[sl][e][ph] = u[e];
        }
    };
    double qo[4], qh[4] = {1., 0., 0., 0.};
    {
        double a0[4], a1[4], a2[4], h0[4] = {1., 0., 0., 0.}, h1[4] = {1., 0., 0., 0.}, iv2[4] = {1., 0., 0., 0.};
        ld4(i0 - 1, oo, a0);
        ld4(i0, oo, a1);
        ld4(i0 + 1, oo, a2);
        if (hal) {
            ld4(i0 - 1, oh, h0);
            ld4(i0, oh, h1);
            ld4(i0 + 1, oh, iv2);
        }
        put(i0 - 1, a0, h0);
        put(i0, a1, h1);
        put(i0 + 1, a2, iv2);
    }
    ld4(i0 + 2, oo, qo);
    if (hal) ld4(i0 + 2, oh, qh);
    const float scale0 = ma.fac && valid ? ma.scale * ma.fac[j] : ma.scale;
    for (int i = i0; i < i1; i++) {
        // level of the network's Cd in this cell (used after the network)
        const float scale = ma.fac && ma.fsi && valid ? ma.scale * ma.fac[(long long)i * ma.fsi + j] : scale0;
        lds_barrier();
        Stencil st;
        {
            const int sm = (i - 1) & (MRING - 1), sc = i & (MRING - 1), sp = (i + 1) & (MRING - 1);
#pragma unroll
            for (int r = 0; r < 3; r++) {
                st.uc[r] = s_u[sc][r][po];
                st.up[0][r] = s_u[sp][r][po];
                st.um[0][r] = s_u[sm][r][po];
                st.up[1][r] = s_u[sc][r][pyp];
                st.um[1][r] = s_u[sc][r][pym];
                st.up[2][r] = s_u[sc][r][pzp];
                st.um[2][r] = s_u[sc][r][pzm];
            }
        }
        float f[NIN], smag = 0.f;
        if (valid) {
            smag = cell_features(st, a, i, j, f);
        } else {
#pragma unroll
            for (int q = 0; q < NIN; q++) f[q] = nm[q];
        }
        const double rho = s_rho[i & (MRING - 1)][t];
        // plane i + 2 (loaded during the previous iteration) into the ring,
        // plane i + 3 in flight
        put(i + 2, qo, qh);
        ld4(i + 3, oo, qo);
        if (hal) ld4(i + 3, oh, qh);
        float fn[NIN];
#pragma unroll
        for (int q = 0; q < NIN; q++) fn[q] = (f[q] - nm[q]) * ns[q];
        *(iv4*)&X[lane * NIN] = cvt4i(fn[0], fn[1], fn[2], fn[3]);
        *(iv4*)&X[lane * NIN + 4] = cvt4i(fn[4], fn[5], fn[6], fn[7]);
        wave_lds_sync();
        float ymine = 0.f;
#pragma unroll
        for (int s0 = 0; s0 < 4; s0 += TP) {
            // B operand of layer 1: features 4 q4 .. 4 q4 + 3 of cell st * 16 + r16 (K padded to 16)
            iv4 xb[TP];
#pragma unroll
            for (int u = 0; u < TP; u++) {
                xb[u] = iv4{0, 0, 0, 0};
                if (q4 < 2) xb[u] = *(const iv4*)&X[((s0 + u) * 16 + r16) * NIN + 4 * q4];
            }
            iv4 h1[TP][4], iv2[TP][4];
#pragma unroll
            for (int tt = 0; tt < 4; tt++)
#pragma unroll
                for (int u = 0; u < TP; u++) {
                    const f4 z = mfma_i(A1[tt], xb[u], f4{kb1[tt][0], kb1[tt][1], kb1[tt][2], kb1[tt][3]});
                    h1[u][tt] = act4<ACT>(z);
                }
#pragma unroll
            for (int tt = 0; tt < 4; tt++)
#pragma unroll
                for (int u = 0; u < TP; u++) {
                    f4 z = {kb2[tt][0], kb2[tt][1], kb2[tt][2], kb2[tt][3]};
                    z = mfma_i32(A2[0][tt], cat8(h1[u][0], h1[u][1]), z);
                    z = mfma_i32(A2[1][tt], cat8(h1[u][2], h1[u][3]), z);
                    iv2[u][tt] = act4<ACT>(z);
                }
#pragma unroll
            for (int u = 0; u < TP; u++) {
                f4 z = {0.f, 0.f, 0.f, 0.f};
                z = mfma_i32(A3[0], cat8(iv2[u][0], iv2[u][1]), z);
                z = mfma_i32(A3[1], cat8(iv2[u][2], iv2[u][3]), z);
                // row 0 of the output tile: y of cell st * 16 + r16 sits in lane r16, element 0
                const float yt = __shfl(z[0], r16, 64);
                if (q4 == s0 + u) ymine = yt + b3;
            }
        }
        if (valid) {
            const float Cd = fminf(fmaxf(0.15f * scale * ymine, 0.f), 0.15f);
            const long long c = cidx(g, i, j, k);
            ma.mu[c] = rho * (double)Cd * a.ph.DD * (double)smag;
            if (ma.cd_out) ma.cd_out[c] = Cd;
        }
        wave_lds_sync();   // X is rewritten by the next iteration
    }
}

// forward pass on caller-given raw features (normalised as in k_mlp_mu_plane)
template <int ACT>
__global__ void __launch_bounds__(256) k_mlp_forward(const float* X, long long n, MlpW w, float* y)
{
    __shared__ WaveLds lds[4];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    WaveLds& ws = lds[wv];
    const long long s = (long long)blockIdx.x * 256 + threadIdx.x;
#pragma unroll
    for (int q = 0; q < NIN; q++) ws.X[lane * KP1 + q] = s < n ? f2i((X[s * NIN + q] - w.norm[q]) * w.norm[NIN + q]) : 0;
#pragma unroll
    for (int q = NIN; q < KP1; q++) ws.X[lane * KP1 + q] = 0;
    __syncthreads();
    forward64<true, ACT>(ws, w, lane);
    if (s < n) y[s] = ws.y[lane];
}

// ---- training ----
struct TrainArgs {
    const float* X;      // raw features
    const float* t;      // targets
    long long n;         // samples
    long long batch;     // samples per optimizer step (multiple of 256)
    long long step;      // step index inside the epoch
    unsigned long long perm_a, perm_c;   // sample permutation: (a * idx + c) mod n
    float* grad;         // NPARAM accumulators
    float* loss;         // sum of weighted squared errors
    float wpow, winv;    // sample weight (|S| Delta / dU)^wpow * winv (feature 6); wpow 0: 1
};

template <int ACT>
__global__ void __launch_bounds__(256) k_mlp_grad(TrainArgs T, MlpW w)
{
    __shared__ WaveLds lds[4];
    __shared__ unsigned short sdz[4][64 * NH];
    __shared__ float sdy[4][64];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r16 = lane & 15, q4 = lane >> 4;
    WaveLds& ws = lds[wv];
    unsigned short* dz = sdz[wv];
    const long long idx = T.step * T.batch + (long long)blockIdx.x * 256 + threadIdx.x;
    const long long s = (long long)((T.perm_a * (unsigned long long)(idx % T.n) + T.perm_c) % (unsigned long long)T.n);
#pragma unroll
    for (int q = 0; q < NIN; q++) ws.X[lane * KP1 + q] = f2bf((T.X[s * NIN + q] - w.norm[q]) * w.norm[NIN + q]);
#pragma unroll
    for (int q = NIN; q < KP1; q++) ws.X[lane * KP1 + q] = 0;
    __syncthreads();
    forward64<false, ACT>(ws, w, lane);
    const float wgt = T.wpow != 0.f ? powf(fmaxf(T.X[s * NIN + 6], 0.f), T.wpow) * T.winv : 1.f;
    const float err = ws.y[lane] - T.t[s];
    const float dy = wgt * err * (2.f / (float)T.batch);
    sdy[wv][lane] = dy;
    atomicAdd(T.loss, wgt * err * err);
    __syncthreads();
    // layer 3: grad w3[k] = sum_s H2[s][k] dy[s], grad b3; dZ2 = dy w3 act'(H2)
    {
        float gw = 0.f, gb = 0.f;
        for (int q = 0; q < 64; q++) {
            gw += bf2f(ws.H2[q * NH + lane]) * sdy[wv][q];
            gb += sdy[wv][q];
        }
        atomicAdd(&T.grad[OW3 + lane], gw);
        if (lane == 0) atomicAdd(&T.grad[OB3], gb);
        for (int kk = 0; kk < NH; kk++) {
            const float h = bf2f(ws.H2[lane * NH + kk]);
            dz[lane * NH + kk] = f2bf(dy * w.master[OW3 + kk] * act_deriv<ACT>(h));
        }
    }
    __syncthreads();
    // grad W2 = H1^T dZ2 (M = j, N = k, K = samples); grad b2 = column sums of dZ2
    {
        float gb = 0.f;
        for (int q = 0; q < 64; q++) gb += bf2f(dz[q * NH + lane]);
        atomicAdd(&T.grad[OB2 + lane], gb);
        for (int tj = 0; tj < 4; tj++)
            for (int tk = 0; tk < 4; tk++) {
                f4 acc = {0.f, 0.f, 0.f, 0.f};
                for (int ks = 0; ks < 4; ks++) {
                    bf4 a, b;
                    const int s0 = ks * 16 + 4 * q4;
                    a.x = (short)ws.H1[(s0 + 0) * NH + tj * 16 + r16];
                    a.y = (short)ws.H1[(s0 + 1) * NH + tj * 16 + r16];
                    a.z = (short)ws.H1[(s0 + 2) * NH + tj * 16 + r16];
                    a.w = (short)ws.H1[(s0 + 3) * NH + tj * 16 + r16];
                    b.x = (short)dz[(s0 + 0) * NH + tk * 16 + r16];
                    b.y = (short)dz[(s0 + 1) * NH + tk * 16 + r16];
                    b.z = (short)dz[(s0 + 2) * NH + tk * 16 + r16];
                    b.w = (short)dz[(s0 + 3) * NH + tk * 16 + r16];
                    acc = mfma(a, b, acc);
                }
#pragma unroll
                for (int e = 0; e < 4; e++)
                    atomicAdd(&T.grad[OW2 + (tj * 16 + 4 * q4 + e) * NH + tk * 16 + r16], acc[e]);
            }
    }
    // dZ1 = (dZ2 W2^T) act'(H1) -> H2 buffer
    {
        f4 acc[4][4];
        for (int tm = 0; tm < 4; tm++)
            for (int tn = 0; tn < 4; tn++) acc[tm][tn] = f4{0.f, 0.f, 0.f, 0.f};
        for (int tm = 0; tm < 4; tm++)
            for (int ks = 0; ks < 4; ks++) {
                const bf4 a = lds_row4(&dz[(tm * 16 + r16) * NH + ks * 16 + 4 * q4]);
                for (int tn = 0; tn < 4; tn++) acc[tm][tn] = mfma(a, w.pack[PB2T / 64 * 64 + (ks * 4 + tn) * 64 + lane], acc[tm][tn]);
            }
        __syncthreads();
        for (int tm = 0; tm < 4; tm++)
            for (int tn = 0; tn < 4; tn++)
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int sr = tm * 16 + 4 * q4 + e, jc = tn * 16 + r16;
                    const float h = bf2f(ws.H1[sr * NH + jc]);
                    ws.H2[sr * NH + jc] = f2bf(acc[tm][tn][e] * act_deriv<ACT>(h));
                }
    }
    __syncthreads();
    // grad W1 = X^T dZ1 (M = features padded to 16, N = j, K = samples); grad b1
    {
        float gb = 0.f;
        for (int q = 0; q < 64; q++) gb += bf2f(ws.H2[q * NH + lane]);
        atomicAdd(&T.grad[OB1 + lane], gb);
        for (int tn = 0; tn < 4; tn++) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int ks = 0; ks < 4; ks++) {
                bf4 a, b;
                const int s0 = ks * 16 + 4 * q4;
                a.x = (short)ws.X[(s0 + 0) * KP1 + r16];
                a.y = (short)ws.X[(s0 + 1) * KP1 + r16];
                a.z = (short)ws.X[(s0 + 2) * KP1 + r16];
                a.w = (short)ws.X[(s0 + 3) * KP1 + r16];
                b.x = (short)ws.H2[(s0 + 0) * NH + tn * 16 + r16];
                b.y = (short)ws.H2[(s0 + 1) * NH + tn * 16 + r16];
                b.z = (short)ws.H2[(s0 + 2) * NH + tn * 16 + r16];
                b.w = (short)ws.H2[(s0 + 3) * NH + tn * 16 + r16];
                acc = mfma(a, b, acc);
            }
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int fr = 4 * q4 + e;
                if (fr < NIN) atomicAdd(&T.grad[OW1 + fr * NH + tn * 16 + r16], acc[e]);
            }
        }
    }
}

struct AdamArgs {
    float* master;
    float* grad;
    float* m;
    float* v;
    float lr, b1, b2, eps;
    float c1, c2;   // 1 / (1 - b1^t), 1 / (1 - b2^t)
};

__global__ void k_mlp_adam(AdamArgs a)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    const float g = a.grad[p];
    const float m = a.b1 * a.m[p] + (1.f - a.b1) * g;
    const float v = a.b2 * a.v[p] + (1.f - a.b2) * g * g;
    a.m[p] = m;
    a.v[p] = v;
    a.master[p] -= a.lr * (m * a.c1) / (sqrtf(v * a.c2) + a.eps);
    a.grad[p] = 0.f;
}

// BF16 B operands in lane order from the FP32 master weights, and the BF16
// operands and FP32 constants of the inference form of activation act
__global__ void k_mlp_pack(const float* master, bf4* pack, int act)
{
    const bool relu = act == kRelu;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;   // one (tile, lane) per thread
    if (t >= NPACK) return;
    if (t >= PHF) {
        // FP32 constants of the inference form, two per 8-byte slot:
        // K (b2[n] + sum_k W2[k][n]) for n < 64, b3 + sum_k w3[k] at 64
        // (tanh); b2[n], b3 (ReLU)
        float c[2] = {0.f, 0.f};
        for (int e = 0; e < 2; e++) {
            const int n = 2 * (t - PHF) + e;
            if (n < NH) {
                float sum = master[OB2 + n];
                if (!relu)
                    for (int k2 = 0; k2 < NH; k2++) sum += master[OW2 + k2 * NH + n];
                c[e] = relu ? sum : kTanhK * sum;
            } else if (n == NH) {
                float sum = master[OB3];
                if (!relu)
                    for (int k2 = 0; k2 < NH; k2++) sum += master[OW3 + k2];
                c[e] = sum;
            }
        }
        pack[t] = __builtin_bit_cast(bf4, (f2){c[0], c[1]});
        return;
    }
    const int tile = t / 64, lane = t % 64;
    const int r16 = lane & 15, q4 = lane >> 4;
    float v[4];
    for (int e = 0; e < 4; e++) {
        const int kk = 4 * q4 + e;
        float x = 0.f;
        if (tile < PB2 / 64) {                      // W1, n-tile = tile
            const int n = tile * 16 + r16;
            x = kk < NIN ? master[OW1 + kk * NH + n] : 0.f;
        } else if (tile < PB2T / 64) {              // W2 [k][n]
            const int q = tile - PB2 / 64, ks = q / 4, tn = q % 4;
            x = master[OW2 + (ks * 16 + kk) * NH + tn * 16 + r16];
        } else if (tile < PB3 / 64) {               // W2^T: B[k][n] = W2[n][k]
            const int q = tile - PB2T / 64, ks = q / 4, tn = q % 4;
            x = master[OW2 + (tn * 16 + r16) * NH + ks * 16 + kk];
        } else if (tile < PH1 / 64) {               // w3 in column 0
            const int ks = tile - PB3 / 64;
            x = r16 == 0 ? master[OW3 + ks * 16 + kk] : 0.f;
        } else if (tile < PH2 / 64) {               // K W1 (W1 for ReLU)
            const int n = (tile - PH1 / 64) * 16 + r16;
            x = kk < NIN ? (relu ? 1.f : kTanhK) * master[OW1 + kk * NH + n] : 0.f;
        } else if (tile < PH3 / 64) {               // -2 K W2 [k][n] (W2 for ReLU), 16x16x32 order
            const int u = t - PH2, pos = u >> 1, half = u & 1;
            const int sl = pos / 256, tn = (pos / 64) % 4, ln = pos % 64;
            const int ks = 2 * sl + half, kq = 4 * (ln >> 4) + e, nn = tn * 16 + (ln & 15);
            x = (relu ? 1.f : -2.f * kTanhK) * master[OW2 + (ks * 16 + kq) * NH + nn];
        } else {                                    // -2 w3 in column 0 (w3 for ReLU), 16x16x32 order
            const int u = t - PH3, pos = u >> 1, half = u & 1;
            const int sl = pos / 64, ln = pos % 64;
            const int ks = 2 * sl + half, kq = 4 * (ln >> 4) + e;
            x = (ln & 15) == 0 ? (relu ? 1.f : -2.f) * master[OW3 + ks * 16 + kq] : 0.f;
        }
        v[e] = x;
    }
    if (t >= PH1) {
        pack[t] = __builtin_bit_cast(bf4, cvt4i(v[0], v[1], v[2], v[3]));
        return;
    }
    bf4 o;
    o.x = (short)f2bf(v[0]);
    o.y = (short)f2bf(v[1]);
    o.z = (short)f2bf(v[2]);
    o.w = (short)f2bf(v[3]);
    pack[t] = o;
}

// one 16x16x16 product with the assumed operand layouts (layout check)
__global__ void k_mfma_probe(const float* A, const float* B, float* D)
{
    const int lane = threadIdx.x, r16 = lane & 15, q4 = lane >> 4;
    bf4 a, b;
    a.x = (short)f2bf(A[r16 * 16 + 4 * q4 + 0]);
    a.y = (short)f2bf(A[r16 * 16 + 4 * q4 + 1]);
    a.z = (short)f2bf(A[r16 * 16 + 4 * q4 + 2]);
    a.w = (short)f2bf(A[r16 * 16 + 4 * q4 + 3]);
    b.x = (short)f2bf(B[(4 * q4 + 0) * 16 + r16]);
    b.y = (short)f2bf(B[(4 * q4 + 1) * 16 + r16]);
    b.z = (short)f2bf(B[(4 * q4 + 2) * 16 + r16]);
    b.w = (short)f2bf(B[(4 * q4 + 3) * 16 + r16]);
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(a, b, acc);
#pragma unroll
    for (int e = 0; e < 4; e++) D[(4 * q4 + e) * 16 + r16] = acc[e];
}

struct SumArgs {
    CState5 U;
    Geo g;
    Phys ph;
    const double* mu_a;
    const double* mu_b;
};

// SGS dissipation densities e = mu |S|^2 (|S| = sqrt(2 S:S) of the
// central-difference strain, FP32; the sums FP64) of two mu_SGS fields a, b,
// summed over z and over the planes of one x band for every y row j: sums
// e_a, e_b, e_a^2, e_b^2, e_a e_b.  Block (j, s) sums the planes
// i in [s L / NS, (s + 1) L / NS) + 1 of row j into rows[(q NS + s) H + j - 1],
// q = 0..4; the host adds the NS bands of each row in order (deterministic).
__global__ void __launch_bounds__(256) k_sgs_dissipation_rows(SumArgs a, double* rows)
{
    __shared__ double red[5][256];
    const int j = blockIdx.x + 1, s = blockIdx.y, ns = gridDim.y;
    const int i0 = (int)((long long)s * a.g.L / ns), i1 = (int)((long long)(s + 1) * a.g.L / ns);
    const long long n = (long long)(i1 - i0) * a.g.D;   // cells of this block
    double acc[5] = {0., 0., 0., 0., 0.};
    for (long long t = threadIdx.x; t < n; t += blockDim.x) {
        const int k = (int)(t % a.g.D) + 1;
        const int i = i0 + 1 + (int)(t / a.g.D);
        const long long st[3] = {a.g.si, a.g.sj, 1};
        Stencil sv;
        gather_stencil(a.U, cidx(a.g, i, j, k), st, sv);
        float gr[3][3], lap[3];
        stencil_gradient(sv, a.ph, gr, lap);
        const float sm = __builtin_amdgcn_sqrtf(2.f * strain_sq(gr));
        const long long c = cidx(a.g, i, j, k);
        const double ea = a.mu_a[c] * (double)sm * (double)sm, eb = a.mu_b[c] * (double)sm * (double)sm;
        acc[0] += ea;
        acc[1] += eb;
        acc[2] += ea * ea;
        acc[3] += eb * eb;
        acc[4] += ea * eb;
    }
#pragma unroll
    for (int q = 0; q < 5; q++) red[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
#pragma unroll
            for (int q = 0; q < 5; q++) red[q][threadIdx.x] += red[q][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) rows[((long long)threadIdx.x * ns + s) * a.g.H + j - 1] = red[threadIdx.x][0];
}

// z sums of rho, rho u, rho v, rho w, (rho u)^2 / rho, (rho v)^2 / rho,
// (rho w)^2 / rho for every interior (i, j): one wavefront per (i, j) row,
// fixed lane-strided order and a fixed butterfly, so the sums are
// reproducible.  Layout out[q][(i - 1) H + (j - 1)], q = 0..6.
__global__ void __launch_bounds__(256) k_diag_zsum(CState5 U, Geo g, double* out)
{
    const long long row = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const long long nrow = (long long)g.L * g.H;
    if (row >= nrow) return;
    const int i = (int)(row / g.H) + 1, j = (int)(row % g.H) + 1;
    const long long base = cidx(g, i, j, 0);
    double s[7] = {0., 0., 0., 0., 0., 0., 0.};
    for (int k = 1 + lane; k <= g.D; k += 64) {
        const double r = U.f[0][base + k], mu = U.f[1][base + k], mv = U.f[2][base + k], mw = U.f[3][base + k];
        const double ir = 1. / r;
        s[0] += r;
        s[1] += mu;
        s[2] += mv;
        s[3] += mw;
        s[4] += mu * mu * ir;
        s[5] += mv * mv * ir;
        s[6] += mw * mw * ir;
    }
#pragma unroll
    for (int q = 0; q < 7; q++) {
        double v = s[q];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) out[q * nrow + row] = v;
    }
}

// delta_w(x) = dU / max_j |d u~ / dy| from the Favre-averaged profile
// u~ = sum(rho u) / sum(rho) of each x plane (one block per i); planes whose
// profile is flat keep the fallback value.
__global__ void __launch_bounds__(256) k_diag_dw(const double* zs, Geo g, double dU, double dy, float fallback,
                                                 float* dw)
{
    __shared__ double red[256];
    const int i = blockIdx.x + 1;
    const long long nrow = (long long)g.L * g.H;
    const double* sr = zs + (long long)(i - 1) * g.H;
    const double* su = zs + nrow + (long long)(i - 1) * g.H;
    double m = 0.;
    for (int j = 1 + threadIdx.x; j < g.H - 1; j += 256) {
        const double up = su[j + 1] / sr[j + 1], um = su[j - 1] / sr[j - 1];
        m = fmax(m, fabs(up - um) / (2. * dy));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double d = red[0] > 0. ? dU / red[0] : (double)fallback;
        dw[i] = (float)d;
        if (i == 1) dw[0] = (float)d;
        if (i == g.L) dw[g.L + 1] = (float)d;
    }
}

}  // namespace

// ---- launch layer ----
int mlp_nparam() { return NPARAM; }
int mlp_npack() { return NPACK; }

void launch_mlp_pack(hipStream_t s, const float* master, void* pack, int act)
{
    hipLaunchKernelGGL(k_mlp_pack, dim3((NPACK + 255) / 256), dim3(256), 0, s, master, (bf4*)pack, act);
}

static MlpW mlpw(const void* pack, const float* master, const float* norm)
{
    MlpW w;
    w.pack = (const bf4*)pack;
    w.master = master;
    w.norm = norm;
    return w;
}

void launch_mlp_forward(hipStream_t s, const float* X, long long n, const void* pack, const float* master,
                        const float* norm, float* y, int act)
{
    const dim3 grid((unsigned)((n + 255) / 256));
    if (act == kRelu) hipLaunchKernelGGL(k_mlp_forward<kRelu>, grid, dim3(256), 0, s, X, n, mlpw(pack, master, norm), y);
    else hipLaunchKernelGGL(k_mlp_forward<kTanh>, grid, dim3(256), 0, s, X, n, mlpw(pack, master, norm), y);
}

void launch_mlp_mu(hipStream_t s, CState5 U, const Geo& g, const Phys& ph, double dU, const float* dw,
                   double dy, const void* pack, const float* master, const float* norm, double* mu, float* cd_out,
                   float scale, const float* fac, int fac_planes, int act)
{
    MuArgs ma;
    ma.f.U = U;
    ma.f.g = g;
    ma.f.ph = ph;
    ma.f.dU = dU;
    ma.f.dw = dw;
    ma.f.dy = dy;
    ma.w = mlpw(pack, master, norm);
    ma.mu = mu;
    ma.cd_out = cd_out;
    ma.scale = scale;
    ma.fac = fac;
    ma.fsi = fac_planes ? g.H + 2 : 0;
    ma.ntj = (g.H + MT - 1) / MT;
    ma.ntk = (g.D + MT - 1) / MT;
    const int cols = ma.ntj * ma.ntk;
    int segL = 0;
    // three workgroups per CU at this kernel's occupancy
    const int nseg = fused_segments(g.L, cols, 3, &segL);
    ma.segL = segL;
    ma.nblocks = nseg * cols;
    const dim3 grid((unsigned)((ma.nblocks + 7) / 8 * 8));   // a multiple of the 8 XCDs
    const char* e = getenv("FSL_MLP_TILES");   // read at each launch
    const int tp = e && atoi(e) == 2 ? 2 : 1;
    if (act == kRelu) {
        if (tp == 2) hipLaunchKernelGGL((k_mlp_mu_plane<2, kRelu>), grid, dim3(256), 0, s, ma);
        else hipLaunchKernelGGL((k_mlp_mu_plane<1, kRelu>), grid, dim3(256), 0, s, ma);
    } else {
        if (tp == 2) hipLaunchKernelGGL((k_mlp_mu_plane<2, kTanh>), grid, dim3(256), 0, s, ma);
        else hipLaunchKernelGGL((k_mlp_mu_plane<1, kTanh>), grid, dim3(256), 0, s, ma);
    }
}

void launch_mlp_features(hipStream_t s, CState5 U, const Geo& g, const Phys& ph, double dU, const float* dw,
                         double dy, const float* cd, long long stride, long long offset, float* X, float* t,
                         long long nout)
{
    FeatArgs a;
    a.U = U;
    a.g = g;
    a.ph = ph;
    a.dU = dU;
    a.dw = dw;
    a.dy = dy;
    hipLaunchKernelGGL(k_mlp_features, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, s, a, cd, stride,
                       offset, X, t, nout);
}

void launch_mlp_train_step(hipStream_t s, const float* X, const float* t, long long n, long long batch,
                           long long step, unsigned long long pa, unsigned long long pc, const void* pack,
                           const float* master, const float* norm, float* grad, float* loss, float wpow,
                           float winv, int act)
{
    TrainArgs T;
    T.wpow = wpow;
    T.winv = winv;
    T.X = X;
    T.t = t;
    T.n = n;
    T.batch = batch;
    T.step = step;
    T.perm_a = pa;
    T.perm_c = pc;
    T.grad = grad;
    T.loss = loss;
    const dim3 grid((unsigned)(batch / 256));
    if (act == kRelu) hipLaunchKernelGGL(k_mlp_grad<kRelu>, grid, dim3(256), 0, s, T, mlpw(pack, master, norm));
    else hipLaunchKernelGGL(k_mlp_grad<kTanh>, grid, dim3(256), 0, s, T, mlpw(pack, master, norm));
}

void launch_mlp_adam(hipStream_t s, float* master, float* grad, float* m, float* v, float lr, int t)
{
    AdamArgs a;
    a.master = master;
    a.grad = grad;
    a.m = m;
    a.v = v;
    a.lr = lr;
    a.b1 = 0.9f;
    a.b2 = 0.999f;
    a.eps = 1e-8f;
    a.c1 = 1.f / (1.f - powf(a.b1, (float)t));
    a.c2 = 1.f / (1.f - powf(a.b2, (float)t));
    hipLaunchKernelGGL(k_mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, s, a);
}

void launch_mfma_probe(hipStream_t s, const float* A, const float* B, float* D)
{
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, A, B, D);
}

int dissipation_splits(const Geo& g)
{
    return g.L < kDissBands ? g.L : kDissBands;
}

void launch_sgs_dissipation(hipStream_t s, CState5 U, const Geo& g, const Phys& ph, const double* mu_a,
                            const double* mu_b, double* rows)
{
    SumArgs a;
    a.U = U;
    a.g = g;
    a.ph = ph;
    a.mu_a = mu_a;
    a.mu_b = mu_b;
    hipLaunchKernelGGL(k_sgs_dissipation_rows, dim3(g.H, dissipation_splits(g)), dim3(256), 0, s, a, rows);
}

void launch_diag_zsum(hipStream_t s, CState5 U, const Geo& g, double* out)
{
    const long long nrow = (long long)g.L * g.H;
    hipLaunchKernelGGL(k_diag_zsum, dim3((unsigned)((nrow + 3) / 4)), dim3(256), 0, s, U, g, out);
}

void launch_diag_dw(hipStream_t s, const double* zs, const Geo& g, double dU, double dy, float fallback, float* dw)
{
    hipLaunchKernelGGL(k_diag_dw, dim3((unsigned)g.L), dim3(256), 0, s, zs, g, dU, dy, fallback, dw);
}

}  // namespace fsl

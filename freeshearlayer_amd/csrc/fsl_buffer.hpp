// fsl_buffer.hpp -- plane-relative buffer access of the fused stage kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "fsl_device.hpp"

namespace fsl {

// Plane-relative buffer access.  The x index of a plane is the same for the
// whole workgroup, so a state buffer's plane is a buffer resource in scalar
// registers (field_rsrc below) and a lane supplies only its 32-bit in-plane
// byte offset: no per-lane 64-bit address arithmetic, and the hardware range
// check confines every access to the plane's fields.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
constexpr int kRsrcWord3 = 0x00020000;   // gfx9 buffer resource, untyped access

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const double* field, long long plane, int bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(field + plane), (short)0, bytes, kRsrcWord3);
}

__device__ __forceinline__ double bload(__amdgpu_buffer_rsrc_t r, unsigned boff, unsigned soff = 0)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)boff, (int)soff, 0));
}

__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, unsigned boff, unsigned soff, double v)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)boff, (int)soff, 0);
}

// Field f of a state buffer's plane.  PK: one resource per plane (base =
// field 0 + plane, range = the four further fields + one plane) and the
// field's byte offset f * fb in the scalar offset, which needs 4 fb + one
// plane below 2^32 bytes; otherwise one resource per field and plane.  PK
// holds 4 scalar registers per buffer instead of 20.
template <bool PK, class S5>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t field_rsrc(const S5& s, int f, long long plane, int pbytes,
                                                           unsigned fb)
{
    return PK ? plane_rsrc(s.f[0], plane, (int)(4u * fb + (unsigned)pbytes)) : plane_rsrc(s.f[f], plane, pbytes);
}

template <bool PK>
__device__ __forceinline__ void load5(const CState5& s, long long plane, int pbytes, unsigned fb, unsigned boff,
                                      double q[5])
{
#pragma unroll
    for (int f = 0; f < 5; f++) q[f] = bload(field_rsrc<PK>(s, f, plane, pbytes, fb), boff, PK ? f * fb : 0u);
}

}  // namespace fsl

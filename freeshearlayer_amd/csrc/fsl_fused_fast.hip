// fsl_fused_fast.hip -- the fast build of the FP64 fused stage kernel
// (contexts created with precision FSL_PREC_FP64_FAST).
//
// The source of fsl_fused.hip compiled with FMA contraction
// (-ffp-contract=fast: a product feeding a sum becomes one fused
// multiply-add, rounded once) and with every quotient of the Riemann solver,
// the reconstruction's primitives and the cell velocities taken as n times
// the Newton-refined reciprocal of its denominator (FSL_FP64_FAST in
// fsl_device.hpp) instead of the correctly rounded division.  The
// operations are the reference's otherwise; the results differ from the
// parity kernel by a few units in the last place per operation, not bit for
// bit (about 1e-15 relative after one step on the shear-layer cases).
#define FSL_FP64_FAST 1
#include "fsl_fused.hip"

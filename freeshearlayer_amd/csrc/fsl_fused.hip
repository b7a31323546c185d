// fsl_fused.hip -- fused stage kernels (placeholder until implemented).
#include "fsl.h"
#include "fsl_internal.hpp"

namespace fsl {
int ctx_err(fsl_ctx* c, int code, const char* msg);
int fused_stage(fsl_ctx* ctx, int stage)
{
    (void)stage;
    return ctx_err(ctx, FSL_ERR_UNSUPPORTED, "fused path not built");
}
}  // namespace fsl

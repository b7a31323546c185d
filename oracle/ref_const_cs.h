/*
 * ref_const_cs.h -- forced include (gcc -include) for the constant-Cs build of
 * the reference.  It pulls in the reference's own def.h first, whose include
 * guard then keeps the sources from re-reading it, and switches the compiled-in
 * SGS option (src/def.h:47) to the constant-coefficient Smagorinsky model.
 * TEST INFRASTRUCTURE ONLY.
 */
#include "def.h"
#undef DynamicSmagorinskySGS
#define DynamicSmagorinskySGS 0

/* Learned conserved-to-primitive initial guess (placeholder until the network kernels land). */
#include <hip/hip_runtime.h>
#include "../../include/maragpu.h"

extern "C" void maragpu_mlp_free(void *) {}
extern "C" int maragpu_mlp_c2p_launch(maragpu_ctx *, int, int) { return MARAGPU_E_NOMODEL; }

// sam2d_adjoint.cpp -- FWI gradient entry point (filled in with the adjoint kernels).
#include "sam2d_internal.h"

using namespace sam2d;

void sam2d_adjoint_release(sam2d_ctx *ctx) { (void)ctx; }

extern "C" int sam2d_gradient(sam2d_ctx *ctx, const sam2d_shot *shots, int nshots,
                              const int *windows, const float *adj, int ckpt_k, float *seis,
                              float *g_vp, float *g_rho)
{
    (void)ctx; (void)shots; (void)nshots; (void)windows; (void)adj; (void)ckpt_k;
    (void)seis; (void)g_vp; (void)g_rho;
    return fail(SAM2D_ERR_PARAM, "gradient not built yet");
}

// vc_boundary.hip -- boundary sites, quads and cube-complex Euler (placeholder)
#include "vc_internal.h"

int vc_boundary_run(vc_ctx* ctx, vc_vol* vol, float* sites, int64_t site_cap,
                    int64_t* n_sites, int32_t* quads, int64_t quad_cap,
                    int64_t* n_quads, vc_euler* eu)
{
    return vc_fail(ctx, VC_ERR_INVALID, "boundary extraction not built yet");
}

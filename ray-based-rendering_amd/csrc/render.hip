// Wavefront path tracer (filled in by the integrator work; traversal-only build for now).
#include <hip/hip_runtime.h>

#include "context.h"

using namespace rbr;

void rbr::freeShading(rbr_scene* s) {
    freeBuffer(s->dMaterials);
    freeBuffer(s->dPrimMaterial);
    freeBuffer(s->dLights);
}

void rbr::destroyRenderState(rbr_ctx*) {}

extern "C" {

int rbr_scene_set_shading(rbr_scene* s, const rbr_material*, uint32_t, const int32_t*, uint32_t, const rbr_light*,
                          uint32_t, const rbr_medium*, const float[3]) {
    return s ? fail(s->ctx, RBR_ERR_UNSUPPORTED, "shading not built yet") : RBR_ERR_INVALID;
}

int rbr_render(rbr_ctx* ctx, const rbr_job*, int) {
    return ctx ? fail(ctx, RBR_ERR_UNSUPPORTED, "render not built yet") : RBR_ERR_INVALID;
}

int rbr_render_moments(rbr_ctx* ctx, const rbr_job*, double*, double*) {
    return ctx ? fail(ctx, RBR_ERR_UNSUPPORTED, "render not built yet") : RBR_ERR_INVALID;
}

int rbr_last_render_timing(rbr_ctx* ctx, double*, double*) {
    return ctx ? fail(ctx, RBR_ERR_UNSUPPORTED, "render not built yet") : RBR_ERR_INVALID;
}

}  // extern "C"

// Headless frame driver: the reference's PathTracer (src/pathTracer.cpp)
// without window, swapchain or UI, calling the HIP layer's C ABI.
#include <string.h>

#include <algorithm>
#include <new>
#include <string>

#include "xr_host.h"

struct xrh_app {
    xrh_settings settings;
    const xrh_scene* scene = nullptr;
    const xrh_camera_path* camera = nullptr;
    xr_ctx* ctx = nullptr;
    uint32_t width = 0, height = 0;
    // PathTracer state (pathTracer.h:36-56)
    uint32_t spp = 0;
    uint32_t total_samples = 0;
    bool rebuild_rays = true;
    uint32_t current_frame_total = 0;
    bool has_prev = false;
    float view_prev[16];
    float last_view[16];
    uint32_t path_frame = 0;
    xr_ubo last_ubo;
    std::string err;
};

extern "C" {

int xrh_app_create(const xrh_settings* st, const xrh_scene* scene, const xrh_camera_path* camera, uint32_t w,
                   uint32_t h, xrh_app** out) {
    if (!st || !scene || !camera || !out) return XR_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (st->rendering_mode != XR_MODE_PLAIN && st->rendering_mode != XR_MODE_RESTIR) return XR_ERR_INVALID_ARGUMENT;
    xrh_app* a = new (std::nothrow) xrh_app();
    if (!a) return XR_ERR_OUT_OF_MEMORY;
    a->settings = *st;
    a->scene = scene;
    a->camera = camera;
    a->width = w;
    a->height = h;
    memset(&a->last_ubo, 0, sizeof(a->last_ubo));
    xr_scene_desc d;
    int rc = xrh_scene_desc(scene, &d);
    if (rc == XR_OK) rc = xr_create(&d, w, h, &a->ctx);
    if (rc != XR_OK) {
        // keep the reason reachable through xr_global_last_error()
        delete a;
        return rc;
    }
    // CreateAccelerationStructures (pathTracer.cpp:39 -> ApplicationRT.cpp:55-66)
    // the scene's textures (descriptor binding 8), when it has any
    const xr_texture* tex = nullptr;
    uint32_t ntex = 0;
    rc = xrh_scene_textures(scene, &tex, &ntex);
    if (rc == XR_OK && ntex) rc = xr_set_textures(a->ctx, tex, ntex);
    if (rc == XR_OK) rc = xr_build_accel(a->ctx, XR_BUILDER_LBVH);
    if (rc != XR_OK) {
        a->err = xr_last_error(a->ctx);
        xr_destroy(a->ctx);
        delete a;
        return rc;
    }
    *out = a;
    return XR_OK;
}

int xrh_app_draw_frame(xrh_app* a, const xr_restir_params* params) {
    if (!a) return XR_ERR_INVALID_ARGUMENT;
    // PathTracer::DrawFrame (pathTracer.cpp:71-79)
    if (a->rebuild_rays || !a->settings.enable_ray_accumulation) {
        a->total_samples = 0;
        a->rebuild_rays = false;
    }
    // glm::clamp(maxSpp - totalSamples, 0u, spp) in uint32 arithmetic (the
    // difference wraps if maxSpp is below totalSamples, as in the reference)
    a->spp = std::min(a->settings.max_spp - a->total_samples, a->settings.spp);
    a->total_samples += a->spp;
    a->current_frame_total = (a->current_frame_total + 1) % 10000;
    // PathTracer::Render: camera update (pathTracer.cpp:88); a moved camera
    // restarts accumulation on the next frame, as rebuildRays does.
    float view[16];
    xrh_camera_view(a->camera, a->path_frame, view);
    a->path_frame++;
    if (a->has_prev && memcmp(view, a->last_view, sizeof(view)) != 0) a->rebuild_rays = true;
    memcpy(a->last_view, view, sizeof(view));
    // GetUniformBufferObject (pathTracer.cpp:169-201).  The reference's
    // m_viewPrev is uninitialised on the first frame (currentFrameTotal is
    // already 1 there); the driver uses the current view instead.
    const float* prev = a->has_prev ? a->view_prev : view;
    xr_ubo ubo;
    xrh_make_ubo(view, prev, a->settings.fov, a->width, a->height, a->spp, a->total_samples,
                 a->settings.number_of_bounces, a->settings.enable_sky_lighting, a->settings.enable_reuse,
                 a->current_frame_total, &ubo);
    memcpy(a->view_prev, view, sizeof(view));
    a->has_prev = true;
    a->last_ubo = ubo;
    const int rc = xr_render(a->ctx, &ubo, (int)a->settings.rendering_mode, params);
    if (rc != XR_OK) a->err = xr_last_error(a->ctx);
    return rc;
}

int xrh_app_last_ubo(const xrh_app* a, xr_ubo* out) {
    if (!a || !out) return XR_ERR_INVALID_ARGUMENT;
    *out = a->last_ubo;
    return XR_OK;
}

xr_ctx* xrh_app_ctx(xrh_app* a) { return a ? a->ctx : nullptr; }

const char* xrh_app_last_error(const xrh_app* a) { return a ? a->err.c_str() : ""; }

void xrh_app_destroy(xrh_app* a) {
    if (!a) return;
    xr_destroy(a->ctx);
    delete a;
}

}  // extern "C"

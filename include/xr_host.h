/*
 * xr_host.h - C ABI of the C++ host side: synthetic scenes, camera paths,
 * UBO construction and a headless frame driver.
 *
 * The frame driver plays the role of the reference's PathTracer
 * (src/pathTracer.cpp:57-201) without a window, swapchain or ImGui: it keeps
 * the same per-frame bookkeeping (spp / totalSamples / currentFrameTotal,
 * :71-79), builds the UniformBufferObject the same way
 * (GetUniformBufferObject, :169-201) from a seeded camera path instead of
 * the interactive Camera, and calls xr_render() where the reference records
 * ApplicationRT::Render / ApplicationReSTIRGI::Render (:93-108).
 */
#ifndef XR_HOST_H
#define XR_HOST_H

#include <stddef.h>
#include <stdint.h>
#include "xr_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- synthetic scenes (AScene layout: concatenated buffers + per-instance offsets) ---- */
typedef struct xrh_scene xrh_scene;

/* Cornell-style box: 12 axis-aligned quads in the unit cube (floor, back,
 * left, right as two quads each; the ceiling as four quads around a 0.2 x 0.2
 * opening), albedo U[0.2,0.8]^3 per quad, a 1,280-triangle icosphere of radius
 * 0.25 at the centre, and a 0.2 x 0.2 ceiling light of radiance 15 (baseColor
 * 15/4: the shaders emit baseColor * 4). */
int xrh_scene_cornell(uint64_t seed, xrh_scene** out);

/* Terrain: grid x grid vertex heightfield over [0,1]^2 with 4-octave fBm
 * heights (amplitude 0.1), n_spheres icospheres (1,280 triangles each, one
 * instance each) resting on it with radii U[0.02,0.1] and albedo U[0.1,0.9],
 * and n_lights downward-facing quad lights (side U[0.03,0.08], height
 * U[0.6,0.9]) of radiance U[5,50] above it (one
 * instance). */
int xrh_scene_terrain(uint64_t seed, uint32_t grid, uint32_t n_spheres, uint32_t n_lights, xrh_scene** out);

/* glTF 2.0 files (.gltf JSON with external or base64 data-URI buffers; GLB
 * is not read), one instance per file, loaded as AglTFModel::loadglTFFile
 * and AScene do (src/Asset/Model.cpp:69-260, src/Asset/Scene.cpp:16-50):
 * scene 0's nodes depth first, children before the node's own mesh; node
 * matrix T*R*S or `matrix`, composed with its parents in glm's float order;
 * positions world*(p,1); normals transpose(inverse(world))*(normalize(n),1),
 * not renormalised (NaN when a primitive has no NORMAL, as in the
 * reference); TEXCOORD_0 or 0; material index of the primitive (-1 when
 * absent) plus the file's material offset; baseColorFactor, else zeros;
 * baseColorTexture index, else -1; material model DiffuseLight for a
 * material named "light", else Lambertian.
 * Base-colour textures: the file's `textures` in order, each the PNG image
 * its `source` names (uri relative to the file, data URI, or bufferView),
 * decoded to RGBA8 (xrh_scene_textures); other image formats (JPEG) fail.
 * XRH_GLTF_UNTEXTURED loads no images and sets every texture index to -1,
 * so the base colour factor is shaded where the reference samples a texture.
 * Returns XR_ERR_INVALID_ARGUMENT with a message in err (when err_len > 0)
 * for unreadable or malformed files and for inputs the reference would read
 * out of bounds (no indices, index types other than u8/u16/u32, non-float
 * attributes, accessors past their buffer). */
#define XRH_GLTF_UNTEXTURED 0x1u
int xrh_scene_gltf(const char* const* paths, uint32_t count, uint32_t flags, xrh_scene** out, char* err,
                   size_t err_len);

/* Pointers into the scene's arrays (valid until xrh_scene_destroy). */
int xrh_scene_desc(const xrh_scene* s, xr_scene_desc* out);
/* The scene's textures (glTF scenes loaded without XRH_GLTF_UNTEXTURED: the
 * PNG images of every file's textures, in file order, as AScene collects them,
 * Scene.cpp:56-63; material texture indices are file-local, as in the
 * reference).  *textures is valid until xrh_scene_destroy; count may be 0. */
int xrh_scene_textures(const xrh_scene* s, const xr_texture** textures, uint32_t* count);
/* PNG file bytes -> RGBA8 (every colour type and depth except interlaced
 * images; grey replicated, alpha 255 unless the image carries one, 16-bit
 * samples keep their high byte).  rgba8 NULL: only the size is returned. */
int xrh_png_decode(const uint8_t* data, size_t bytes, uint32_t* width, uint32_t* height, uint8_t* rgba8,
                   size_t rgba_bytes);
uint32_t xrh_scene_triangle_count(const xrh_scene* s);
void xrh_scene_destroy(xrh_scene* s);

/* ---- camera ---- */
typedef struct xrh_camera_path xrh_camera_path;

/* Static camera at eye looking at target (config-1 style). */
int xrh_camera_static(const float eye[3], const float target[3], xrh_camera_path** out);
/* Closed Catmull-Rom path through n_ctrl seeded control points over the
 * terrain (eye at heights 0.32-0.45: above every sphere, below every light;
 * looking at a second seeded path
 * near the ground), sampled at `frames` evenly spaced parameters. */
int xrh_camera_orbit(uint64_t seed, uint32_t n_ctrl, uint32_t frames, xrh_camera_path** out);
/* Explicit per-frame eye and target positions (3 floats each per frame):
 * stands in for the reference's interactive camera input. */
int xrh_camera_keyframes(const float* eyes, const float* targets, uint32_t frames, xrh_camera_path** out);
/* View matrix (column-major) of path frame k (k taken modulo the frame count). */
int xrh_camera_view(const xrh_camera_path* p, uint32_t k, float view[16]);
void xrh_camera_destroy(xrh_camera_path* p);

/* glm::perspective(radians(fovy_deg), aspect, znear, zfar) (right-handed,
 * depth 0..1) with [1][1] negated, as pathTracer.cpp:173-175 builds it. */
void xrh_perspective(float fovy_deg, float aspect, float znear, float zfar, float m[16]);
/* General 4x4 inverse (glm::inverse), computed in FP64 and rounded to FP32.
 * Returns 0 when the matrix is singular. */
int xrh_inverse(const float m[16], float inv[16]);

/* ---- frame driver ---- */

/* UserSettings (src/Editor/UserSettings.h:5-30), the fields the path reads. */
typedef struct xrh_settings {
    uint32_t rendering_mode;          /* XR_MODE_PLAIN or XR_MODE_RESTIR */
    uint32_t spp;                     /* per frame (main.cpp:46 uses 12) */
    uint32_t number_of_bounces;       /* main.cpp:47 uses 10 */
    uint32_t max_spp;                 /* main.cpp:48 uses 1000 */
    float fov;                        /* vertical degrees, main.cpp:45 uses 45 */
    uint32_t enable_ray_accumulation; /* main.cpp:43 false */
    uint32_t enable_sky_lighting;     /* main.cpp:50 false */
    uint32_t enable_reuse;            /* value-initialised false in main.cpp:41 */
} xrh_settings;

typedef struct xrh_app xrh_app;

/* Creates the HIP context for the scene at width x height and builds the
 * acceleration structure.  The scene and the camera path must outlive the
 * app. */
int xrh_app_create(const xrh_settings* settings, const xrh_scene* scene, const xrh_camera_path* camera,
                   uint32_t width, uint32_t height, xrh_app** out);
/* One PathTracer::DrawFrame: bookkeeping, camera step, UBO, xr_render.
 * Asynchronous; params may be NULL (reference defaults). */
int xrh_app_draw_frame(xrh_app* app, const xr_restir_params* params);
/* The UBO the last frame was rendered with. */
int xrh_app_last_ubo(const xrh_app* app, xr_ubo* out);
xr_ctx* xrh_app_ctx(xrh_app* app);
const char* xrh_app_last_error(const xrh_app* app);
void xrh_app_destroy(xrh_app* app);

/* ---- image output ---- */

/* PNG (8-bit RGB, stored deflate blocks) of a B8G8R8A8_UNORM image as
 * xr_read_image returns it: the file-side counterpart of the reference's copy
 * of its output image to the swapchain (ApplicationReSTIRGI.cpp:247-267).
 * Alpha is dropped.  xrh_png_encode writes the file's bytes to dst (dst NULL:
 * only *written, the size, is set); xrh_write_png writes them to path. */
int xrh_png_encode(const uint32_t* bgra8, uint32_t width, uint32_t height, uint8_t* dst, size_t dst_bytes,
                   size_t* written);
int xrh_write_png(const char* path, const uint32_t* bgra8, uint32_t width, uint32_t height);

/* Fills ubo for a given view and frame counter exactly as the driver does
 * (prev matrices from the arguments). */
void xrh_make_ubo(const float view[16], const float view_prev[16], float fov_deg, uint32_t width, uint32_t height,
                  uint32_t spp, uint32_t total_samples, uint32_t bounces, uint32_t sky, uint32_t reuse,
                  uint32_t frame, xr_ubo* ubo);

#ifdef __cplusplus
}
#endif
#endif /* XR_HOST_H */

/*
 * xr_api.h - C ABI of the MI355X (gfx950) path-tracing / ReSTIR GI layer.
 *
 * This library is a drop-in for the ray-tracing half of the reference's
 * frame: PathTracer::Render (src/pathTracer.cpp:82-116) calls
 *   case 1: ApplicationRT::Render        (src/Vulkan/RayTracing/ApplicationRT.cpp:117-183)
 *   case 2: ApplicationReSTIRGI::Render  (src/Vulkan/RayTracing/ApplicationReSTIRGI.cpp:150-275)
 * and both are replaced by xr_render().  The resources those classes create
 * (acceleration structures, ApplicationRT.cpp:55-66; ReSTIR buffers and
 * G-buffer images, ApplicationReSTIRGI.cpp:304-374) are owned by an xr_ctx.
 * The operator contract is the reference's 19-binding descriptor table
 * (src/Vulkan/RayTracing/ReSTIRGIPipeline.cpp:40-74); every binding maps to a
 * buffer id below.
 *
 * Conventions: plain C, opaque context, int status (XR_OK = 0), no call
 * throws or aborts, one context is not thread safe, each context owns one
 * HIP stream, xr_render is asynchronous and xr_sync waits for it.  Host
 * pointers unless a function says "device".
 */
#ifndef XR_API_H
#define XR_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference throws std::runtime_error, XVKCommon.h:10-16) ---- */
#define XR_OK 0
#define XR_ERR_INVALID_ARGUMENT 1 /* null pointer, bad size, bad enum */
#define XR_ERR_DEVICE 2           /* HIP runtime error (no device, launch failure) */
#define XR_ERR_OUT_OF_MEMORY 3
#define XR_ERR_STATE 4            /* call out of order, e.g. render before xr_build_accel */
#define XR_ERR_UNSUPPORTED 5      /* feature outside this layer (textures, non-identity instances) */

typedef struct xr_ctx xr_ctx;

/* AVertex (src/Asset/Vertex.h:11-16), 36 bytes, read as 9 dwords by Vertex.hlsl:12-33. */
typedef struct xr_vertex {
    float pos[3];
    float normal[3];
    float tex_coord[2];
    int32_t material_index;
} xr_vertex;

/* AMaterial (src/Asset/Material.h:46-57), 32 bytes. */
#define XR_MATERIAL_LAMBERTIAN 0u
#define XR_MATERIAL_METALLIC 1u
#define XR_MATERIAL_DIELECTRIC 2u
#define XR_MATERIAL_ISOTROPIC 3u
#define XR_MATERIAL_DIFFUSE_LIGHT 4u
typedef struct xr_material {
    float base_color[4];
    int32_t base_color_texture_index; /* -1, or an index into the textures given to xr_set_textures */
    float fuzziness;
    float refraction_index;
    uint32_t material_model;
} xr_material;

/* UniformBufferObject (src/Asset/UniformBuffer.h:16-32), 412 bytes, matrices
 * column-major as glm stores them; built by PathTracer::GetUniformBufferObject
 * (src/pathTracer.cpp:169-201). */
typedef struct xr_ubo {
    float view[16];
    float view_prev[16];
    float projection[16];
    float projection_prev[16];
    float view_inverse[16];
    float projection_inverse[16];
    uint32_t spp;
    uint32_t total_number_of_samples;
    uint32_t number_of_bounces;
    uint32_t random_seed;
    uint32_t has_sky_box;
    uint32_t reuse;
    uint32_t current_frame;
} xr_ubo;

/* ReSTIRGISample / ReSTIRGIReservoir (src/Vulkan/RayTracing/ReSTIRGICommon.h:8-32):
 * 96 and 112 bytes.  The reference-layout form of xr_read_buffer/xr_write_buffer. */
typedef struct xr_restir_sample {
    float x_view[3];
    uint32_t material_id;
    float n_view[3];
    uint32_t rand_seed;
    float x_sample[3];
    float p_q;
    float n_sample[3];
    uint32_t pad1;
    float l_out[3];
    uint32_t pad2;
    float f[3];
    uint32_t pad3;
} xr_restir_sample;

typedef struct xr_restir_reservoir {
    float w_sum;
    float w;
    uint32_t m;
    float pad;
    xr_restir_sample s;
} xr_restir_reservoir;

/* The scene as AScene concatenates it (src/Asset/Scene.cpp:16-87): one
 * vertex, index and material buffer, plus (indexOffset, vertexOffset) per
 * instance (Scene.cpp:37).  Instance i owns the triangles of index range
 * [offsets[2i], next instance's offsets[2i] or index_count), which must be
 * increasing.  Instances have identity transforms, as the reference's TLAS
 * (XVKTLAccelerationStructure.cpp:68-93).  All arrays are copied to the device
 * by xr_create. */
typedef struct xr_scene_desc {
    const xr_vertex* vertices;
    uint32_t vertex_count;
    const uint32_t* indices;
    uint32_t index_count;
    const xr_material* materials;
    uint32_t material_count;
    const uint32_t* offsets; /* 2 x uint32 per instance */
    uint32_t instance_count;
} xr_scene_desc;

/* ---- render modes (UserSettings::renderingMode, src/pathTracer.cpp:93-108) ---- */
#define XR_MODE_PLAIN 1  /* ApplicationRT::Render: rgen.hlsl */
#define XR_MODE_RESTIR 2 /* ApplicationReSTIRGI::Render: initial -> temporal -> spatial */

/* visibility test of the spatial pass */
#define XR_VIS_REFERENCE 0 /* CheckVisibility as the reference binds it: always visible (Spatial.rgen:44-62) */
#define XR_VIS_TRACE 1     /* trace the any-hit shadow ray */

/* flags: which passes of mode 2 run (all when none is set) */
#define XR_PASS_INITIAL 0x1u
#define XR_PASS_TEMPORAL 0x2u
#define XR_PASS_SPATIAL 0x4u

typedef struct xr_restir_params {
    uint32_t temporal_m_cap;   /* M_MAX of ReSTIRGITemporalReuse.rgen.hlsl:3 (reference 30) */
    uint32_t spatial_m_cap;    /* M_MAX of ReSTIRGISpatialReuse.rgen.hlsl:6 (reference 300) */
    uint32_t spatial_samples;  /* NUM_SPATIAL_SAMPLES, Spatial.rgen:4 (reference 6), at most 64; with 0
                                  and both reuse passes run in one call, the spatial pass (the centre's
                                  reservoir alone) runs in the temporal pass's kernel and its time counts
                                  as temporal */
    float spatial_radius;      /* SPATIAL_RADIUS in pixels, Spatial.rgen:5 (reference 5.0) */
    uint32_t visibility;       /* XR_VIS_* (reference XR_VIS_REFERENCE) */
    uint32_t plain_max_depth;  /* mode 1 bounce cap, 0 = unbounded as rgen.hlsl:59 */
    uint32_t passes;           /* XR_PASS_* mask for mode 2, 0 = all three */
    uint32_t flags;            /* XR_FLAG_* */
    float mlp_spread;          /* cache: stop at x_s when its spread > mlp_spread * primary footprint (0.01) */
    float mlp_lr;              /* cache: Adam learning rate (1e-3) */
    uint32_t mlp_train_pixels; /* cache: training samples per frame (65,536) of the paths reaching a diffuse x_s */
    float mlp_loss_scale;      /* cache: loss scale of the BF16 backward pass (128) */
    uint32_t reserved[4];      /* must be zero */
} xr_restir_params;

/* flags: also write the linear (pre-tone-map) colour of every output pixel to
 * XR_BUF_RADIANCE (0 for the black / white early-out pixels of the spatial
 * pass), for error measurements against the converged image. */
#define XR_FLAG_RADIANCE 0x1u
/* flags: MLP radiance cache (no reference counterpart).  In mode 2 the
 * initial-sample pass ends most paths at their first diffuse vertex x_s and
 * takes L_out from a 96-64x5-3 MLP evaluated there.  Of the frame's paths
 * that reach a diffuse x_s, exactly mlp_train_pixels (or all, if fewer),
 * spread evenly over them in pixel order and rotating with the frame, keep
 * tracing and their path-suffix radiance trains the network online (one Adam
 * step per frame, BF16 MFMA, FP32 accumulation).  Other paths whose x_s is
 * too close to their primary footprint fall back to plain tracing. */
#define XR_FLAG_MLP_CACHE 0x2u
/* flags: with XR_FLAG_MLP_CACHE, the cache's hidden-layer weights are FP8
 * E4M3 with one power-of-two scale per layer (dequantised exactly to BF16 for
 * the MFMA; activations stay BF16, accumulation FP32).  Training keeps the
 * FP32 master weights and backpropagates through the quantised values.
 * Without it the weights are BF16. */
#define XR_FLAG_MLP_FP8 0x4u

/* Fills p with the reference's compile-time constants (30 / 300 / 6 / 5.0,
 * always-visible, unbounded plain depth, all passes). */
void xr_default_params(xr_restir_params* p);

/* ---- lifecycle ---- */

/* Copies the scene to device 0 and allocates the frame buffers for a
 * width x height target (InitReSTIRGITracer, ApplicationReSTIRGI.cpp:304-374).
 * Vertex coordinates must be finite and within +-2^40 (XR_ERR_UNSUPPORTED
 * otherwise: the traversal's slab arithmetic stays finite in that range).
 * On failure *out is NULL and xr_global_last_error() says why. */
int xr_create(const xr_scene_desc* scene, uint32_t width, uint32_t height, xr_ctx** out);

/* Builds the acceleration structure (CreateAccelerationStructures,
 * ApplicationRT.cpp:55-66).  builder: XR_BUILDER_LBVH (Morton-order radix
 * tree on the GPU, fastest build), XR_BUILDER_PLOC (locally-ordered
 * clustering on the GPU), XR_BUILDER_SAH (binned surface-area heuristic, top
 * down on the GPU: usually the best trees) or XR_BUILDER_SAH_HOST (the same
 * binned SAH on host threads; the same tree, for comparison).  Every builder
 * gives the same ray/triangle hits. */
#define XR_BUILDER_LBVH 0
#define XR_BUILDER_PLOC 1
#define XR_BUILDER_SAH 2
#define XR_BUILDER_SAH_HOST 3
int xr_build_accel(xr_ctx* ctx, int builder);

/* Base-colour textures (descriptor binding 8, TextureSamplers[],
 * BSDF.hlsl:4-5; images uploaded as R8G8B8A8_UNORM, TextureImage.cpp:25, and
 * sampled with a LINEAR, CLAMP_TO_EDGE sampler at level 0, XVKSampler.h:9-26,
 * BSDF.hlsl:22).  Copies `count` RGBA8 images (rows top to bottom, texel
 * (x, y) at rgba8[4 * (y * width + x)]) to the device, replacing earlier
 * ones.  A frame whose scene names a texture index >= count fails with
 * XR_ERR_STATE.  Sampling: x = u * width - 0.5, i0 = floor(x), a = x - i0
 * (x first clamped to [-1, width]), likewise y / j0 / b; the four texels
 * clamped to the edge; colour = ((w00 t00 + w10 t10) + w01 t01) + w11 t11
 * with w00 = (1-a)(1-b), w10 = a(1-b), w01 = (1-a)b, w11 = ab and texels as
 * byte / 255, in FP32 without contraction (the oracle's definition; the
 * reference's hardware sampler rounds its weights to a fixed-point grid of
 * its own). */
typedef struct xr_texture {
    uint32_t width, height;
    const uint8_t* rgba8;
} xr_texture;
int xr_set_textures(xr_ctx* ctx, const xr_texture* textures, uint32_t count);

/* Reallocates the frame buffers (swapchain recreation,
 * ApplicationReSTIRGI.cpp:115-148); reservoir history is cleared. */
int xr_resize(xr_ctx* ctx, uint32_t width, uint32_t height);

/* Records one frame on the context's stream.  mode: XR_MODE_*; params may be
 * NULL (reference defaults). */
int xr_render(xr_ctx* ctx, const xr_ubo* ubo, int mode, const xr_restir_params* params);

/* Waits for every call queued on the context. */
int xr_sync(xr_ctx* ctx);

/* Destroys the context and frees its device memory.  NULL is allowed. */
void xr_destroy(xr_ctx* ctx);

/* ---- read back / inject (parity dumps, fixtures) ---- */

/* buffer ids = reference bindings */
#define XR_BUF_ACCUMULATION 1 /* binding 1, RGBA32F, 16 B/px */
#define XR_BUF_OUTPUT 2       /* binding 2, B8G8R8A8_UNORM, 4 B/px */
#define XR_BUF_SAMPLES 9      /* binding 9, ReSTIRGISample, N entries */
#define XR_BUF_TEMPORAL 10    /* binding 10, ReSTIRGIReservoir, 2N entries (two ping-pong slots) */
#define XR_BUF_SPATIAL 11     /* binding 11, ReSTIRGIReservoir, 2N entries */
#define XR_BUF_GBUF_POS 12    /* binding 12 (= 16), RGBA32F */
#define XR_BUF_GBUF_NRM 13    /* binding 13 (= 17), RGBA32F */
#define XR_BUF_GBUF_MATUV 14  /* binding 14, RGBA32F */
#define XR_BUF_MOTION 15      /* binding 15, RG16F, 4 B/px */
#define XR_BUF_RADIANCE 16    /* no binding: linear colour behind the output, RGBA32F (XR_FLAG_RADIANCE) */
#define XR_BUF_CONVERGED 17   /* no binding: FP64 RGB sums of xr_converge, 24 B/px */

#define XR_LAYOUT_REFERENCE 0 /* AoS exactly as the reference binds it (96 / 112-byte structs) */
#define XR_LAYOUT_NATIVE 1    /* as stored on the device: six float4 planes per entry */

/* bytes must equal the buffer's size in the chosen layout.  Synchronous. */
int xr_read_buffer(xr_ctx* ctx, int which, void* dst, size_t bytes, int layout);
int xr_write_buffer(xr_ctx* ctx, int which, const void* src, size_t bytes, int layout);
/* Shorthand for xr_read_buffer(ctx, XR_BUF_OUTPUT, ...). */
int xr_read_image(xr_ctx* ctx, uint32_t* dst_bgra8, size_t bytes);
/* Size in bytes of a buffer in a layout (0 for an unknown id). */
size_t xr_buffer_bytes(const xr_ctx* ctx, int which, int layout);

/* ---- ray batches (TraceRay with the reference's flags) ---- */

typedef struct xr_ray {
    float origin[3];
    float tmin;
    float direction[3];
    float tmax;
} xr_ray;

typedef struct xr_hit {
    float t; /* -1 on miss (Ray.rmiss.hlsl:6-7) */
    float u; /* hit-attribute barycentrics: weight of v1 */
    float v; /* weight of v2 */
    uint32_t pad0;
    uint32_t instance;  /* InstanceIndex(), 0xFFFFFFFF on miss */
    uint32_t primitive; /* PrimitiveIndex() */
    uint32_t triangle;  /* canonical id = first index of the triangle / 3 */
    uint32_t pad1;
} xr_hit;

/* any_hit = 0: closest hit (RAY_FLAG_NONE, InitialSample.rgen:90-98).
 * any_hit = 1: first hit found (ACCEPT_FIRST_HIT_AND_END_SEARCH, Spatial.rgen:58); t/u/v/ids of
 * some hit in [tmin, tmax].  Host arrays, synchronous. */
int xr_trace_rays(xr_ctx* ctx, const xr_ray* rays, uint32_t n, xr_hit* hits, int any_hit);
/* Same on device arrays, queued on the context's stream (no copies). */
int xr_trace_rays_device(xr_ctx* ctx, const xr_ray* d_rays, uint32_t n, xr_hit* d_hits, int any_hit);

/* ---- measurement ---- */

/* Milliseconds of the last xr_render's kernels (HIP events on the context's
 * stream): mode 2 -> {initial, temporal, spatial, cache (inference + training
 * + Adam; two adjacent events without XR_FLAG_MLP_CACHE)}, mode 1 -> {plain}.  Returns the
 * number written.  Waits for the frame. */
int xr_pass_times(xr_ctx* ctx, float* ms, uint32_t n);

/* Per-pass milliseconds summed over every frame rendered since the last
 * xr_reset_pass_times (same pass order), and the number of frames.  Each
 * frame's events are kept until read, so a run of frames is timed without a
 * host sync per frame.  Waits for the frames. */
int xr_pass_time_totals(xr_ctx* ctx, double* ms_sum, uint32_t n, uint32_t* frames);
int xr_reset_pass_times(xr_ctx* ctx);

/* Device counters since xr_create / the last reset: [0] extend rays
 * (camera + bounce), [1] shadow rays, [2] pixels whose primary ray found a
 * non-emissive surface (primitiveHitValid of the initial sample pass),
 * [3] radiance-cache queries, [4] radiance-cache training samples.  Writes
 * min(n, 5) values.  Waits for queued work. */
int xr_counters(xr_ctx* ctx, uint64_t* out, uint32_t n);
int xr_reset_counters(xr_ctx* ctx);

typedef struct xr_accel_info {
    uint32_t triangle_count;
    uint32_t node_count;
    uint32_t max_depth;
    uint32_t builder;
    float build_ms;
    uint64_t node_bytes;
    uint64_t triangle_bytes;
} xr_accel_info;
int xr_get_accel_info(const xr_ctx* ctx, xr_accel_info* info);

/* Copies the acceleration structure to host memory for validation: node_count
 * 64-byte nodes of up to four children {float4 q; uint4 p; uint4 r; int4
 * link}.  q.xyz is the node's grid origin; the grid steps (powers of two) are
 * the floats q.w (x), r.z (y) and r.w (z, both as float bits); child k's box
 * is origin + g * step per axis with the planes g in byte k of p.x (lo.x),
 * p.y (hi.x), p.z (lo.y), p.w (hi.y), r.x (lo.z), r.y (hi.z).  link[k] >= 0
 * names a node, link[k] < 0 a leaf ~(first * 16 + count - 1).  The children
 * fill slots 0 .. count-1; an unused slot holds the empty box lo = 255,
 * hi = 0 on every axis and link -1.
 * Node 0 is the root and max_depth counts the nodes on the longest
 * root-to-leaf path.  Triangles are triangle_count 48-byte BVH-ordered
 * records {v0.xyz, canonical id; v1.xyz, instance; v2.xyz, primitive}. */
int xr_read_accel(xr_ctx* ctx, void* nodes, size_t node_bytes, void* triangles, size_t triangle_bytes);

/* Resources of the frame kernels (no reference counterpart; measurement
 * support): index 0..xr_kernel_count()-1 names k_initial_sample,
 * k_tail_bounce, k_temporal, k_spatial, k_plain, k_converge.  vgprs and
 * scratch_bytes (per lane) and lds_bytes (per workgroup) as the runtime's
 * function attributes report them; blocks_per_cu resident workgroups of
 * `block` lanes per CU, waves_per_simd = blocks_per_cu * block / 256.
 * Needs a device: XR_ERR_DEVICE without one. */
typedef struct xr_kernel_info {
    char name[32];
    uint32_t vgprs;
    uint32_t scratch_bytes;
    uint32_t lds_bytes;
    uint32_t block;
    uint32_t blocks_per_cu;
    uint32_t waves_per_simd;
} xr_kernel_info;
uint32_t xr_kernel_count(void);
int xr_kernel_info_get(uint32_t index, xr_kernel_info* out);

/* HIP stream the context records on (a hipStream_t). */
void* xr_stream(xr_ctx* ctx);

/* ---- converged image and error (no reference counterpart) ---- */

/* Adds spp samples per pixel of the quantity the ReSTIR output estimates (the
 * initial sample's one-sample RIS estimate L_out * f * cos / p_q at the
 * primary hit, 0 where the primary ray misses or sees an emitter) into FP64
 * per-pixel sums.  Sample s uses frame number frame_base + s for its RNG seed,
 * so successive calls with disjoint ranges accumulate one progressive image.
 * ubo supplies camera and bounce count.  Asynchronous. */
int xr_converge(xr_ctx* ctx, const xr_ubo* ubo, uint32_t spp, uint32_t frame_base);
/* Clears the sums. */
int xr_converge_reset(xr_ctx* ctx);
/* Samples per pixel accumulated so far. */
uint64_t xr_converged_spp(const xr_ctx* ctx);
/* Root-mean-square error over pixels and RGB channels of XR_BUF_RADIANCE (the
 * last frame rendered with XR_FLAG_RADIANCE) against the converged mean,
 * reduced in FP64 on the device.  Synchronous. */
int xr_rmse(xr_ctx* ctx, double* rmse);

/* ---- MLP radiance cache (no reference counterpart) ---- */

#define XR_MLP_INPUTS 96
#define XR_MLP_PARAMS 23043

/* (Re)initialises the network (He-uniform weights from seed, zero biases and
 * Adam state).  Runs implicitly on the first use with seed 1. */
int xr_mlp_reset(xr_ctx* ctx, uint64_t seed);
/* Forward pass on rows x 96 inputs (host arrays) -> rows x 3 outputs, on the
 * same MFMA path the cache uses.  Synchronous. */
int xr_mlp_eval(xr_ctx* ctx, const float* inputs, uint32_t rows, float* outputs);
/* Gradient of the relative-L2 loss mean((p - y)^2 / (p^2 + 0.01)) over rows x 3
 * on the given data (BF16 backward, FP32 accumulation), without a step. */
int xr_mlp_gradient(xr_ctx* ctx, const float* inputs, const float* targets, uint32_t rows, float* grad_out);
/* One Adam step (lr; loss scale as in the frame) on the given data. */
int xr_mlp_train_rows(xr_ctx* ctx, const float* inputs, const float* targets, uint32_t rows, float lr);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  The gradient
 * of xr_mlp_gradient for rows x 96 inputs and rows x 3 targets already in
 * device memory, written to d_grad_out (XR_MLP_PARAMS floats, device memory).
 * Calls no hipMalloc / hipFree: the 32 partial gradient copies live in a
 * static device array of the library (2,949,504 bytes per device, held until
 * the process ends), cleared by a kernel at the start of every call.  The
 * network must exist (xr_mlp_reset) and rows <= XR_MLP_NOALLOC_MAX_ROWS
 * (3 * rows exact in FP32, the loss normalisation), else an error.
 * Asynchronous on the context's stream: make the inputs ready before the
 * call and xr_sync before reading d_grad_out.  Calls that overlap on one
 * device race on the shared array. */
#define XR_MLP_NOALLOC_MAX_ROWS 4194304u
int xr_mlp_gradient_noalloc(xr_ctx* ctx, const float* d_inputs, const float* d_targets, uint32_t rows,
                            float* d_grad_out);
#define XR_MLP_WEIGHTS_BF16 0
#define XR_MLP_WEIGHTS_FP8 1
#define XR_MLP_FP8_BYTES 22528 /* E4M3 W^T of the hidden layers: [64][96] then 4 x [64][64] */
/* Weight format of the network's GEMM operands for xr_mlp_eval,
 * xr_mlp_gradient and xr_mlp_train_rows (a frame with XR_FLAG_MLP_CACHE
 * selects FP8 when XR_FLAG_MLP_FP8 is set, BF16 otherwise). */
int xr_mlp_set_weight_format(xr_ctx* ctx, int format);
/* In the FP8 format: the E4M3 codes (XR_MLP_FP8_BYTES; layer l's W^T is
 * row j = output unit, column k = input) and the 5 per-layer scales.
 * XR_ERR_STATE in the BF16 format. */
int xr_mlp_fp8_weights(xr_ctx* ctx, uint8_t* codes, float* scales);
/* FP32 master parameters (XR_MLP_PARAMS floats; per layer W[in][out]
 * row-major, then the bias).  set != 0: the network takes params first (and
 * its GEMM copies are rebuilt); then the parameters are copied back into
 * params.  Synchronous. */
int xr_mlp_params(xr_ctx* ctx, float* params, int set);

/* ---- batched views (no reference counterpart) ----
 * Several cameras on the scene of one context, each with its own frame
 * buffers and reservoir history, rendered in ReSTIR GI mode with one kernel
 * launch per pass for all views (one workgroup per tile, or persistent
 * workgroups looping over the views' tiles); the frame is captured once as a
 * HIP graph and replayed.  Each view's results equal those of rendering it alone.
 * Errors are reported through xr_last_error(ctx).  Destroy a batch before
 * its context. */
typedef struct xr_batch xr_batch;
int xr_batch_create(xr_ctx* ctx, uint32_t view_count, const uint32_t* widths, const uint32_t* heights,
                    xr_batch** out);
/* One ReSTIR GI frame of every view; ubos[view_count].  params as for
 * xr_render (every pass).  With XR_FLAG_MLP_CACHE each view chooses its own
 * training samples exactly as xr_render would for it alone (up to
 * mlp_train_pixels, rotated by its UBO's frame number), then one training
 * step of the context's network covers the samples of all views (the mean
 * loss over them), then every view's query and training pixels take the
 * updated network's radiance.  With the graph enabled the initial-sample
 * pass and the temporal + spatial passes replay as two graphs and the
 * training step runs between them.  Asynchronous. */
int xr_batch_render(xr_batch* batch, const xr_ubo* ubos, const xr_restir_params* params);
/* Milliseconds of the last xr_batch_render (events around it).  Waits. */
int xr_batch_frame_time(xr_batch* batch, float* ms);
/* enable != 0 (default): replay a captured HIP graph; 0: launch eagerly.
 * The graph is captured again on the next xr_batch_render whenever the
 * parameters (cache settings included) or the kernel configuration change,
 * or xr_set_textures / xr_build_accel replaced the scene's device memory
 * since the capture. */
int xr_batch_set_graph(xr_batch* batch, int enable);
/* Kernel configurations: tile shape, LDS traversal-stack depth and launch
 * form.  A grid configuration launches one workgroup per tile; a persistent
 * one launches as many workgroups as the GPU holds at once, each looping
 * over tiles taken from a per-XCD queue (workgroup g works in XCD g % 8's
 * band of tile rows).  Results are identical in every configuration. */
uint32_t xr_batch_config_count(void);
int xr_batch_config_info(uint32_t config, uint32_t* tile_w, uint32_t* tile_h, uint32_t* lds_depth);
/* *persistent = 1: persistent launches, 0: one workgroup per tile. */
int xr_batch_config_persistent(uint32_t config, int* persistent);
int xr_batch_set_config(xr_batch* batch, uint32_t config);
uint32_t xr_batch_config(const xr_batch* batch);
/* Autotuner: renders frames (1..64) of the given views with every
 * configuration, keeps the fastest (ms_per_config, if not NULL, receives
 * xr_batch_config_count() averages), then clears the views' buffers and the
 * counters. */
int xr_batch_autotune(xr_batch* batch, const xr_ubo* ubos, const xr_restir_params* params, uint32_t frames,
                      float* ms_per_config);
/* Clears every view's buffers (reservoir history included). */
int xr_batch_reset(xr_batch* batch);
/* Counters summed over the views, as xr_counters. */
int xr_batch_counters(xr_batch* batch, uint64_t* out, uint32_t n);
int xr_batch_reset_counters(xr_batch* batch);
/* Per-view buffers: XR_BUF_OUTPUT, XR_BUF_MOTION, XR_BUF_GBUF_POS,
 * XR_BUF_SAMPLES, XR_BUF_TEMPORAL, XR_BUF_SPATIAL (0 bytes: not kept). */
size_t xr_batch_buffer_bytes(const xr_batch* batch, uint32_t view, int which, int layout);
int xr_batch_read_buffer(xr_batch* batch, uint32_t view, int which, void* dst, size_t bytes, int layout);
void xr_batch_destroy(xr_batch* batch);

/* ---- errors ---- */
const char* xr_last_error(const xr_ctx* ctx);
/* Error of the last failed call that had no context (xr_create). Thread local. */
const char* xr_global_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* XR_API_H */

// C ABI of the HIP layer (include/xr_api.h).
//
// Owns what the reference's ApplicationRT / ApplicationReSTIRGI own: the
// scene buffers (AScene, src/Asset/Scene.cpp:16-87), the acceleration
// structure (ApplicationRT.cpp:55-66) and the per-resolution frame resources
// (InitReSTIRGITracer, ApplicationReSTIRGI.cpp:304-374), and records the
// frame that ApplicationReSTIRGI::Render records (:150-275): initial sample,
// temporal reuse, spatial reuse, in order on one stream, each pass ordered
// after the previous one by the stream (the reference's memory barriers,
// :180-215).
//
// Deliberate differences from the reference frame, none observable in the
// image or in the reservoirs:
//  * the end-of-frame copy samples -> old samples (:272-274) is not made; the
//    reference binds the old buffer as the sample buffer and never reads the
//    copy before the next initial pass overwrites it (ReSTIRGIPipeline.cpp:141-143);
//  * the "prev" G-buffer bindings 16/17 alias bindings 12/13 in the reference
//    (ReSTIRGIPipeline.cpp:171-199, ApplicationReSTIRGI.cpp:367-372), so they
//    are one buffer here and the self-copy of InitialSample.rgen:193-194 is
//    dropped;
//  * buffers start zeroed (Vulkan leaves them undefined).
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <new>
#include <string>
#include <vector>

#include "../../include/xr_api.h"
#include "xr_kernels.h"
#include "xr_mlp.h"

using namespace xr;

static_assert(sizeof(xr_vertex) == sizeof(AVertex), "xr_vertex layout");
static_assert(sizeof(xr_material) == sizeof(AMaterial), "xr_material layout");
static_assert(sizeof(xr_ubo) == sizeof(UniformBufferObject), "xr_ubo layout");
static_assert(sizeof(xr_restir_sample) == sizeof(ReSTIRGISample), "xr_restir_sample layout");
static_assert(sizeof(xr_restir_reservoir) == sizeof(ReSTIRGIReservoir), "xr_restir_reservoir layout");
static_assert(sizeof(xr_ray) == 32 && sizeof(xr_hit) == 32, "ray / hit records are two float4");

// Device counters: extend rays, shadow rays, valid primary hits, cache
// queries, cache training samples.
constexpr uint32_t kCounters = 5;

struct xr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    float* noalloc_scratch = nullptr;  // the device's static scratch array (xr_mlp_gradient_noalloc)
    // Bumped whenever device memory that a SceneDev points at is freed or
    // replaced (textures, BVH): captured batch graphs hold a SceneDev by value.
    uint64_t scene_gen = 0;
    std::string err;

    // scene
    float* d_vertices = nullptr;
    uint32_t* d_indices = nullptr;
    float4* d_materials = nullptr;
    uint2* d_offsets = nullptr;
    uint32_t* d_inst_first = nullptr;
    uint32_t vertex_count = 0, index_count = 0, material_count = 0, instance_count = 0, tri_count = 0;
    // base-colour textures (xr_set_textures)
    uint4* d_tex_info = nullptr;  // (first texel, width, height, 0) per texture
    uint32_t* d_texels = nullptr;  // RGBA8, R in the low byte
    uint32_t texture_count = 0;
    int32_t max_texture_index = -1;  // largest index a material names

    // acceleration structure
    WideNode* d_nodes = nullptr;
    BvhTri* d_tris = nullptr;
    bool built = false;
    xr_accel_info accel{};

    // frame resources
    uint32_t width = 0, height = 0;
    float4* d_samples = nullptr;
    float4* d_temporal = nullptr;
    float4* d_spatial = nullptr;
    float4* d_gpos = nullptr;
    float4* d_gnrm = nullptr;
    float4* d_gmatuv = nullptr;
    float4* d_accum = nullptr;
    uint32_t* d_motion = nullptr;
    uint32_t* d_output = nullptr;
    float4* d_radiance = nullptr;
    uint32_t* d_cache_list = nullptr;  // radiance-cache work lists (2 + 2n words)
    uint32_t* d_cache_work = nullptr;  // training choice: n pending slots + block scan (nblocks + 1)
    float4* d_nbr = nullptr;           // spatial-pass neighbour records (4 float4 per pixel)
    // The records hold the samples and temporal reservoirs of slot nbr_ping as
    // the temporal pass left them; a buffer write or resize invalidates them.
    bool nbr_ok = false;
    uint32_t nbr_ping = 0;
    float4* d_tail = nullptr;          // initial-pass paths past the head bounces (2 queues of n x 4 float4)
    uint32_t* d_tail_count = nullptr;
    unsigned long long* d_counters = nullptr;

    // converged image: FP64 RGB sums per pixel, samples accumulated, error scratch
    double* d_conv = nullptr;
    uint64_t conv_spp = 0;
    double* d_err = nullptr;

    // Per-pass timing: a ring of event sets, one per recorded frame.  A set is
    // harvested into the running totals when it is about to be reused or when
    // the totals are read, so timing a run of frames needs no host sync per
    // frame.
    // ReSTIR sets record e0 initial e1 cache e2 temporal e3 spatial e4.
    struct EventSet {
        hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
        int n = 0;           // events recorded (2: plain, 5: ReSTIR)
        bool pending = false;
    };
    static constexpr int kRing = 8;
    EventSet ring[kRing];
    int ring_cursor = 0;
    int last_set = -1;
    double pass_sum_ms[4] = {0, 0, 0, 0};  // initial (or plain), temporal, spatial, cache
    uint32_t pass_frames = 0;

    // scene box (position normalisation of the cache inputs)
    float scene_lo[3] = {0, 0, 0}, scene_hi[3] = {1, 1, 1};

    // MLP radiance cache
    void* d_mlp = nullptr;
    MlpState mlp;
    float* d_mlp_io = nullptr;  // staging for xr_mlp_eval / train_rows
    size_t mlp_io_cap = 0;

    // staging for ray batches
    float4* d_rays = nullptr;
    float4* d_hits = nullptr;
    size_t ray_cap = 0;
};

namespace {

thread_local std::string g_err;

int set_err(xr_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    else g_err = buf;
    return code;
}

int hip_code(hipError_t e) { return e == hipErrorOutOfMemory ? XR_ERR_OUT_OF_MEMORY : XR_ERR_DEVICE; }

#define XR_HIP(ctx, expr)                                                                         \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return set_err((ctx), hip_code(e_), "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

bool fault_inject(const char* what) {
    const char* v = getenv("XR_FAULT_INJECT");
    return v && strstr(v, what) != nullptr;
}

// Device allocation with optional injected failure (XR_FAULT_INJECT=alloc).
template <typename T>
hipError_t dev_alloc(T** p, size_t bytes) {
    *p = nullptr;
    if (fault_inject("alloc")) return hipErrorOutOfMemory;
    if (bytes == 0) bytes = 16;
    return hipMalloc((void**)p, bytes);
}

template <typename T>
void dev_free(T** p) {
    if (*p) hipFree((void*)*p);
    *p = nullptr;
}

void free_frame(xr_ctx* c) {
    dev_free(&c->d_samples);
    dev_free(&c->d_temporal);
    dev_free(&c->d_spatial);
    dev_free(&c->d_gpos);
    dev_free(&c->d_gnrm);
    dev_free(&c->d_gmatuv);
    dev_free(&c->d_accum);
    dev_free(&c->d_motion);
    dev_free(&c->d_output);
    dev_free(&c->d_radiance);
    dev_free(&c->d_cache_list);
    dev_free(&c->d_cache_work);
    dev_free(&c->d_nbr);
    c->nbr_ok = false;
    dev_free(&c->d_tail);
    dev_free(&c->d_tail_count);
    dev_free(&c->d_conv);
    c->conv_spp = 0;
    c->width = c->height = 0;
}

int alloc_frame(xr_ctx* c, uint32_t w, uint32_t h) {
    free_frame(c);
    const size_t n = (size_t)w * h;
    const size_t f4 = sizeof(float4);
    XR_HIP(c, dev_alloc(&c->d_samples, kPlanes * n * f4));
    XR_HIP(c, dev_alloc(&c->d_temporal, 2 * kPlanes * n * f4));
    XR_HIP(c, dev_alloc(&c->d_spatial, 2 * kPlanes * n * f4));
    XR_HIP(c, dev_alloc(&c->d_gpos, n * f4));
    XR_HIP(c, dev_alloc(&c->d_gnrm, n * f4));
    XR_HIP(c, dev_alloc(&c->d_gmatuv, n * f4));
    XR_HIP(c, dev_alloc(&c->d_accum, n * f4));
    XR_HIP(c, dev_alloc(&c->d_motion, n * 4));
    XR_HIP(c, dev_alloc(&c->d_output, n * 4));
    XR_HIP(c, dev_alloc(&c->d_radiance, n * f4));
    XR_HIP(c, dev_alloc(&c->d_cache_list, (2 + 2 * n) * sizeof(uint32_t)));
    XR_HIP(c, dev_alloc(&c->d_tail, 2 * 4 * n * f4));  // two path queues, one per bounce parity
    XR_HIP(c, dev_alloc(&c->d_tail_count, kTailCounters * sizeof(uint32_t)));
    XR_HIP(c, dev_alloc(&c->d_nbr, 4 * n * f4));
    XR_HIP(c, dev_alloc(&c->d_cache_work, (n + (n + kCacheSelBlock - 1) / kCacheSelBlock + 1) * sizeof(uint32_t)));
    XR_HIP(c, hipMemsetAsync(c->d_cache_list, 0, (2 + 2 * n) * sizeof(uint32_t), c->stream));
    XR_HIP(c, hipMemsetAsync(c->d_radiance, 0, n * f4, c->stream));
    XR_HIP(c, hipMemsetAsync(c->d_samples, 0, kPlanes * n * f4, c->stream));
    XR_HIP(c, hipMemsetAsync(c->d_temporal, 0, 2 * kPlanes * n * f4, c->stream));
    XR_HIP(c, hipMemsetAsync(c->d_spatial, 0, 2 * kPlanes * n * f4, c->stream));
    XR_HIP(c, hipMemsetAsync(c->d_gpos, 0, n * f4, c->stream));
    XR_HIP(c, hipMemsetAsync(c->d_gnrm, 0, n * f4, c->stream));
    XR_HIP(c, hipMemsetAsync(c->d_gmatuv, 0, n * f4, c->stream));
    XR_HIP(c, hipMemsetAsync(c->d_accum, 0, n * f4, c->stream));
    XR_HIP(c, hipMemsetAsync(c->d_motion, 0, n * 4, c->stream));
    XR_HIP(c, hipMemsetAsync(c->d_output, 0, n * 4, c->stream));
    XR_HIP(c, hipStreamSynchronize(c->stream));
    c->width = w;
    c->height = h;
    return XR_OK;
}

SceneDev scene_dev(const xr_ctx* c) {
    SceneDev s;
    s.vertices = c->d_vertices;
    s.indices = c->d_indices;
    s.materials = c->d_materials;
    s.offsets = c->d_offsets;
    s.tex_info = c->d_tex_info;
    s.texels = c->d_texels;
    s.nodes = c->d_nodes;
    s.tris = c->d_tris;
    s.material_count = c->material_count;
    return s;
}

FrameDev frame_dev(const xr_ctx* c) {
    FrameDev f;
    f.width = c->width;
    f.height = c->height;
    f.n = c->width * c->height;
    f.samples = c->d_samples;
    f.temporal = c->d_temporal;
    f.spatial = c->d_spatial;
    f.gpos = c->d_gpos;
    f.gnrm = c->d_gnrm;
    f.gmatuv = c->d_gmatuv;
    f.motion = c->d_motion;
    f.output = c->d_output;
    f.accum = c->d_accum;
    f.nbr = c->d_nbr;
    f.radiance = nullptr;
    f.counters = c->d_counters;
    f.cache_list = c->d_cache_list;
    f.cache_slot = c->d_cache_work;
    f.cache_blocks = c->d_cache_work ? c->d_cache_work + f.n : nullptr;
    f.tail = c->d_tail;
    f.tail_count = c->d_tail_count;
    f.pending = c->d_tail ? c->d_tail + 4 * (size_t)f.n : nullptr;
    f.pending_count = c->d_tail_count ? c->d_tail_count + kPendingCount : nullptr;
    return f;
}

// Milliseconds per pass of one recorded set, in reporting order
// (initial, temporal, spatial, cache); *count = number of passes.
int set_times(xr_ctx* c, const xr_ctx::EventSet& set, float* ms, int* count) {
    XR_HIP(c, hipEventSynchronize(set.ev[set.n - 1]));
    if (set.n == 2) {
        XR_HIP(c, hipEventElapsedTime(&ms[0], set.ev[0], set.ev[1]));
        *count = 1;
        return XR_OK;
    }
    static const int slot[4] = {0, 3, 1, 2};  // interval i -> reporting slot
    for (int i = 0; i < 4; ++i) XR_HIP(c, hipEventElapsedTime(&ms[slot[i]], set.ev[i], set.ev[i + 1]));
    *count = 4;
    return XR_OK;
}

// Folds a recorded event set into the per-pass totals (waits for it).
int harvest(xr_ctx* c, xr_ctx::EventSet& set) {
    if (!set.pending) return XR_OK;
    float ms[4] = {0, 0, 0, 0};
    int k = 0;
    const int rc = set_times(c, set, ms, &k);
    if (rc) return rc;
    for (int i = 0; i < k; ++i) c->pass_sum_ms[i] += ms[i];
    c->pass_frames++;
    set.pending = false;
    return XR_OK;
}

// He-uniform initial weights (U(-a, a), a = sqrt(6 / fan_in)) from a
// splitmix64 stream, zero biases, zero Adam state; then the BF16 copies.
int mlp_init(xr_ctx* c, uint64_t seed) {
    std::vector<float> w(kMlpParams, 0.0f);
    uint64_t s = seed;
    auto next = [&]() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    for (int l = 0; l < kMlpLayers; ++l) {
        const float a = sqrtf(6.0f / (float)mlp_in(l));
        for (int i = 0; i < mlp_in(l) * mlp_out(l); ++i) {
            const double u = (double)(next() >> 11) * (1.0 / 9007199254740992.0);
            w[mlp_woff(l) + i] = (float)((2.0 * u - 1.0) * a);
        }
    }
    XR_HIP(c, hipMemsetAsync(c->d_mlp, 0, mlp_bytes(), c->stream));
    XR_HIP(c, hipMemcpyAsync(c->mlp.w, w.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice, c->stream));
    c->mlp.steps = 0;
    XR_HIP(c, launch_mlp_adam(c->mlp, 0.0f, 1.0f, false, c->stream));
    XR_HIP(c, hipStreamSynchronize(c->stream));
    return XR_OK;
}

int ensure_mlp(xr_ctx* c) {
    if (c->d_mlp) return XR_OK;
    XR_HIP(c, dev_alloc(&c->d_mlp, mlp_bytes()));
    mlp_bind(&c->mlp, c->d_mlp);
    for (int a = 0; a < 3; ++a) {
        const float ext = c->scene_hi[a] - c->scene_lo[a];
        c->mlp.lo[a] = c->scene_lo[a];
        c->mlp.inv_ext[a] = ext > 0.0f ? 1.0f / ext : 1.0f;
    }
    return mlp_init(c, 1);
}

int ensure_mlp_io(xr_ctx* c, size_t floats) {
    if (floats <= c->mlp_io_cap) return XR_OK;
    dev_free(&c->d_mlp_io);
    c->mlp_io_cap = 0;
    XR_HIP(c, dev_alloc(&c->d_mlp_io, floats * sizeof(float)));
    c->mlp_io_cap = floats;
    return XR_OK;
}

int check_textures(xr_ctx* c) {
    if (c->max_texture_index >= (int32_t)c->texture_count)
        return set_err(c, XR_ERR_STATE, "a material names texture %d; xr_set_textures has given %u",
                       c->max_texture_index, c->texture_count);
    return XR_OK;
}

int check_params(xr_ctx* c, const xr_restir_params* p) {
    if (int rc = check_textures(c)) return rc;
    for (int i = 0; i < 4; ++i)
        if (p->reserved[i] != 0) return set_err(c, XR_ERR_INVALID_ARGUMENT, "params.reserved[%d] must be zero", i);
    if (p->flags & ~(XR_FLAG_RADIANCE | XR_FLAG_MLP_CACHE | XR_FLAG_MLP_FP8))
        return set_err(c, XR_ERR_INVALID_ARGUMENT, "params.flags has unknown bits");
    if (p->flags & XR_FLAG_MLP_CACHE) {
        if (!(p->mlp_spread >= 0.0f) || !(p->mlp_lr >= 0.0f && p->mlp_lr < 1.0f) || p->mlp_train_pixels == 0 ||
            !(p->mlp_loss_scale > 0.0f))
            return set_err(c, XR_ERR_INVALID_ARGUMENT, "params.mlp_* out of range");
    }
    if (p->visibility > XR_VIS_TRACE) return set_err(c, XR_ERR_INVALID_ARGUMENT, "params.visibility %u", p->visibility);
    if (p->spatial_samples > 64) return set_err(c, XR_ERR_INVALID_ARGUMENT, "params.spatial_samples > 64");
    if (!(p->spatial_radius >= 0.0f && p->spatial_radius <= 4096.0f))
        return set_err(c, XR_ERR_INVALID_ARGUMENT, "params.spatial_radius out of [0, 4096]");
    if (p->passes & ~7u) return set_err(c, XR_ERR_INVALID_ARGUMENT, "params.passes has unknown bits");
    return XR_OK;
}

}  // namespace

extern "C" {

void xr_default_params(xr_restir_params* p) {
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->temporal_m_cap = 30;
    p->spatial_m_cap = 300;
    p->spatial_samples = 6;
    p->spatial_radius = 5.0f;
    p->visibility = XR_VIS_REFERENCE;
    p->plain_max_depth = 0;
    p->passes = 0;
    p->mlp_spread = 0.01f;
    p->mlp_lr = 1e-3f;
    p->mlp_train_pixels = 65536;
    p->mlp_loss_scale = 128.0f;
}

const char* xr_global_last_error(void) { return g_err.c_str(); }

const char* xr_last_error(const xr_ctx* ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

void* xr_stream(xr_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

uint32_t xr_kernel_count(void) { return (uint32_t)kKernelInfoCount; }

int xr_kernel_info_get(uint32_t index, xr_kernel_info* out) {
    if (!out) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "out is NULL");
    if (index >= (uint32_t)kKernelInfoCount)
        return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "kernel index %u out of range (%d kernels)", index,
                       kKernelInfoCount);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_err(nullptr, XR_ERR_DEVICE, "no HIP device");
    const char* name = nullptr;
    hipFuncAttributes a;
    int blocks = 0, block = 0;
    const hipError_t e = kernel_resources(index, &name, &a, &blocks, &block);
    if (e != hipSuccess) return set_err(nullptr, XR_ERR_DEVICE, "kernel attributes: %s", hipGetErrorString(e));
    memset(out, 0, sizeof(*out));
    snprintf(out->name, sizeof(out->name), "%s", name);
    out->vgprs = (uint32_t)a.numRegs;
    out->scratch_bytes = (uint32_t)a.localSizeBytes;
    out->lds_bytes = (uint32_t)a.sharedSizeBytes;
    out->block = (uint32_t)block;
    out->blocks_per_cu = (uint32_t)blocks;
    out->waves_per_simd = (uint32_t)(blocks * block / 256);
    return XR_OK;
}

void xr_destroy(xr_ctx* c) {
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    free_frame(c);
    dev_free(&c->d_vertices);
    dev_free(&c->d_indices);
    dev_free(&c->d_materials);
    dev_free(&c->d_offsets);
    dev_free(&c->d_tex_info);
    dev_free(&c->d_texels);
    dev_free(&c->d_inst_first);
    dev_free(&c->d_nodes);
    dev_free(&c->d_tris);
    dev_free(&c->d_counters);
    dev_free(&c->d_rays);
    dev_free(&c->d_hits);
    dev_free(&c->d_err);
    dev_free(&c->d_mlp);
    dev_free(&c->d_mlp_io);
    for (auto& set : c->ring)
        for (auto& e : set.ev)
            if (e) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int xr_create(const xr_scene_desc* s, uint32_t width, uint32_t height, xr_ctx** out) {
    if (!out) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (!s) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "scene is NULL");
    if (width == 0 || height == 0 || (uint64_t)width * height > (1ull << 28))
        return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "bad resolution %ux%u", width, height);
    if (!s->vertices || !s->indices || !s->materials || !s->offsets || s->vertex_count == 0 ||
        s->index_count == 0 || s->material_count == 0 || s->instance_count == 0)
        return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "scene arrays must be non-empty");
    if (s->index_count % 3 != 0) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "index_count %% 3 != 0");
    if (s->index_count / 3 >= (1u << 27))
        return set_err(nullptr, XR_ERR_UNSUPPORTED, "more than 2^27 triangles");
    int32_t max_tex = -1;
    for (uint32_t m = 0; m < s->material_count; ++m) {
        const int32_t t = s->materials[m].base_color_texture_index;
        if (t < -1) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "material %u texture index %d", m, t);
        max_tex = t > max_tex ? t : max_tex;
    }
    for (uint32_t v = 0; v < s->vertex_count; ++v) {
        const int32_t mi = s->vertices[v].material_index;
        if (mi != -1 && (mi < 0 || (uint32_t)mi >= s->material_count))
            return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "vertex %u material %d out of range", v, mi);
    }
    // instance ranges and index bounds
    std::vector<uint32_t> first(s->instance_count + 1);
    for (uint32_t i = 0; i < s->instance_count; ++i) {
        const uint32_t io = s->offsets[2 * i], vo = s->offsets[2 * i + 1];
        const uint32_t end = (i + 1 < s->instance_count) ? s->offsets[2 * (i + 1)] : s->index_count;
        if (io % 3 != 0 || end < io || end > s->index_count || (end - io) % 3 != 0)
            return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "instance %u index range [%u,%u) invalid", i, io, end);
        if (i == 0 && io != 0) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "instance 0 must start at index 0");
        for (uint32_t k = io; k < end; ++k) {
            if ((uint64_t)vo + s->indices[k] >= s->vertex_count)
                return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "index %u of instance %u out of range", k, i);
        }
        first[i] = io / 3;
    }
    first[s->instance_count] = s->index_count / 3;
    // Coordinates within +-2^40 keep every wide node's grid step (at most
    // 2^34) times a reciprocal direction (at most 1e20) finite.
    float box_lo[3] = {3e38f, 3e38f, 3e38f}, box_hi[3] = {-3e38f, -3e38f, -3e38f};
    for (uint32_t v = 0; v < s->vertex_count; ++v)
        for (int a = 0; a < 3; ++a) {
            if (!(fabsf(s->vertices[v].pos[a]) <= 0x1p40f))
                return set_err(nullptr, XR_ERR_UNSUPPORTED, "vertex %u coordinate %g is not finite or beyond +-2^40", v,
                               (double)s->vertices[v].pos[a]);
            box_lo[a] = fminf(box_lo[a], s->vertices[v].pos[a]);
            box_hi[a] = fmaxf(box_hi[a], s->vertices[v].pos[a]);
        }
    for (uint32_t i = 0; i < s->instance_count; ++i)
        if (first[i + 1] == first[i])
            return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "instance %u has no triangles", i);

    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0)
        return set_err(nullptr, XR_ERR_DEVICE, "no HIP device: %s", e != hipSuccess ? hipGetErrorString(e) : "0 devices");

    xr_ctx* c = new (std::nothrow) xr_ctx();
    if (!c) return set_err(nullptr, XR_ERR_OUT_OF_MEMORY, "host allocation failed");
    auto fail = [&](int code) {
        g_err = c->err;
        xr_destroy(c);
        return code;
    };
    int rc;
#define XR_CREATE_HIP(expr)                                                                                  \
    do {                                                                                                     \
        hipError_t e2 = (expr);                                                                              \
        if (e2 != hipSuccess) return fail(set_err(c, hip_code(e2), "%s: %s", #expr, hipGetErrorString(e2))); \
    } while (0)
    XR_CREATE_HIP(hipSetDevice(0));
    XR_CREATE_HIP(mlp_noalloc_scratch(&c->noalloc_scratch));
    XR_CREATE_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto& set : c->ring)
        for (auto& ev : set.ev) XR_CREATE_HIP(hipEventCreate(&ev));
    c->vertex_count = s->vertex_count;
    c->index_count = s->index_count;
    c->material_count = s->material_count;
    c->max_texture_index = max_tex;
    c->instance_count = s->instance_count;
    c->tri_count = s->index_count / 3;
    for (int a = 0; a < 3; ++a) {
        c->scene_lo[a] = box_lo[a];
        c->scene_hi[a] = box_hi[a];
    }
    XR_CREATE_HIP(dev_alloc(&c->d_vertices, sizeof(xr_vertex) * (size_t)s->vertex_count));
    XR_CREATE_HIP(dev_alloc(&c->d_indices, sizeof(uint32_t) * (size_t)s->index_count));
    XR_CREATE_HIP(dev_alloc(&c->d_materials, sizeof(xr_material) * (size_t)s->material_count));
    XR_CREATE_HIP(dev_alloc(&c->d_offsets, sizeof(uint32_t) * 2 * (size_t)s->instance_count));
    XR_CREATE_HIP(dev_alloc(&c->d_inst_first, sizeof(uint32_t) * first.size()));
    XR_CREATE_HIP(dev_alloc(&c->d_counters, sizeof(unsigned long long) * kCounters));
    XR_CREATE_HIP(hipMemcpyAsync(c->d_vertices, s->vertices, sizeof(xr_vertex) * (size_t)s->vertex_count,
                                 hipMemcpyHostToDevice, c->stream));
    XR_CREATE_HIP(hipMemcpyAsync(c->d_indices, s->indices, sizeof(uint32_t) * (size_t)s->index_count,
                                 hipMemcpyHostToDevice, c->stream));
    XR_CREATE_HIP(hipMemcpyAsync(c->d_materials, s->materials, sizeof(xr_material) * (size_t)s->material_count,
                                 hipMemcpyHostToDevice, c->stream));
    XR_CREATE_HIP(hipMemcpyAsync(c->d_offsets, s->offsets, sizeof(uint32_t) * 2 * (size_t)s->instance_count,
                                 hipMemcpyHostToDevice, c->stream));
    XR_CREATE_HIP(hipMemcpyAsync(c->d_inst_first, first.data(), sizeof(uint32_t) * first.size(),
                                 hipMemcpyHostToDevice, c->stream));
    XR_CREATE_HIP(hipMemsetAsync(c->d_counters, 0, sizeof(unsigned long long) * kCounters, c->stream));
    XR_CREATE_HIP(hipStreamSynchronize(c->stream));
#undef XR_CREATE_HIP
    rc = alloc_frame(c, width, height);
    if (rc != XR_OK) return fail(rc);
    *out = c;
    return XR_OK;
}

int xr_set_textures(xr_ctx* c, const xr_texture* t, uint32_t count) {
    if (!c) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (count && !t) return set_err(c, XR_ERR_INVALID_ARGUMENT, "textures is NULL");
    std::vector<uint4> info(count);
    size_t total = 0;
    for (uint32_t i = 0; i < count; ++i) {
        if (!t[i].rgba8 || t[i].width == 0 || t[i].height == 0 || t[i].width > 16384 || t[i].height > 16384)
            return set_err(c, XR_ERR_INVALID_ARGUMENT, "texture %u: bad size or NULL data", i);
        info[i] = make_uint4((uint32_t)total, t[i].width, t[i].height, 0u);
        total += (size_t)t[i].width * t[i].height;
        if (total >= (1ull << 32)) return set_err(c, XR_ERR_UNSUPPORTED, "more than 2^32 texels");
    }
    XR_HIP(c, hipStreamSynchronize(c->stream));  // frames in flight may still read the old textures
    ++c->scene_gen;
    dev_free(&c->d_tex_info);
    dev_free(&c->d_texels);
    c->texture_count = 0;
    if (count) {
        XR_HIP(c, dev_alloc(&c->d_tex_info, sizeof(uint4) * count));
        XR_HIP(c, dev_alloc(&c->d_texels, sizeof(uint32_t) * total));
        XR_HIP(c, hipMemcpy(c->d_tex_info, info.data(), sizeof(uint4) * count, hipMemcpyHostToDevice));
        for (uint32_t i = 0; i < count; ++i)
            XR_HIP(c, hipMemcpy(c->d_texels + info[i].x, t[i].rgba8, 4ull * t[i].width * t[i].height,
                                hipMemcpyHostToDevice));
    }
    c->texture_count = count;
    return XR_OK;
}

int xr_build_accel(xr_ctx* c, int builder) {
    if (!c) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (builder != XR_BUILDER_LBVH && builder != XR_BUILDER_PLOC && builder != XR_BUILDER_SAH &&
        builder != XR_BUILDER_SAH_HOST)
        return set_err(c, XR_ERR_INVALID_ARGUMENT, "unknown builder %d", builder);
    dev_free(&c->d_nodes);
    dev_free(&c->d_tris);
    c->built = false;
    ++c->scene_gen;
    BvhBuildResult r;
    char msg[256] = {0};
    hipError_t e = (builder == XR_BUILDER_PLOC       ? build_ploc
                    : builder == XR_BUILDER_SAH      ? build_sah_gpu
                    : builder == XR_BUILDER_SAH_HOST ? build_sah
                                                     : build_lbvh)(
        c->d_vertices, c->d_indices, c->d_offsets, c->d_inst_first, c->instance_count, c->tri_count, c->stream, &r,
        msg, sizeof(msg));
    if (e != hipSuccess) return set_err(c, XR_ERR_DEVICE, "BVH build failed: %s", msg);
    c->d_tris = r.tris;
    if (r.max_depth > (uint32_t)kMaxTraversalDepth) {
        hipFree(r.nodes);
        dev_free(&c->d_tris);
        return set_err(c, XR_ERR_UNSUPPORTED, "BVH depth %u exceeds the traversal stack (%d)", r.max_depth,
                       kMaxTraversalDepth);
    }
    // Four children per node where the tree's stack bound allows it; the
    // binary tree's own bound is its depth, checked above, so width 2 fits.
    WideBuildResult w;
    for (int width = 4; width >= 2; --width) {
        e = build_wide(r.nodes, r.node_count, width, c->stream, &w, msg, sizeof(msg));
        if (e != hipSuccess) break;
        if (w.stack_bound <= (uint32_t)kMaxTraversalDepth) break;
        hipFree(w.nodes);
        w.nodes = nullptr;
    }
    hipFree(r.nodes);
    if (e != hipSuccess) {
        dev_free(&c->d_tris);
        return set_err(c, XR_ERR_DEVICE, "wide BVH build failed: %s", msg);
    }
    if (!w.nodes) {
        dev_free(&c->d_tris);
        return set_err(c, XR_ERR_UNSUPPORTED, "BVH stack bound exceeds the traversal stack (%d) at every width",
                       kMaxTraversalDepth);
    }
    c->d_nodes = w.nodes;
    c->accel.triangle_count = r.tri_count;
    c->accel.node_count = w.node_count;
    c->accel.max_depth = w.depth;
    c->accel.builder = (uint32_t)builder;
    c->accel.build_ms = r.build_ms + w.build_ms;
    c->accel.node_bytes = (uint64_t)w.node_count * sizeof(WideNode);
    c->accel.triangle_bytes = (uint64_t)r.tri_count * sizeof(BvhTri);
    c->built = true;
    return XR_OK;
}

int xr_get_accel_info(const xr_ctx* c, xr_accel_info* info) {
    if (!c || !info) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!c->built) return XR_ERR_STATE;
    *info = c->accel;
    return XR_OK;
}

int xr_read_accel(xr_ctx* c, void* nodes, size_t node_bytes, void* tris, size_t tri_bytes) {
    if (!c || !nodes || !tris) return set_err(c, XR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!c->built) return set_err(c, XR_ERR_STATE, "xr_build_accel has not run");
    if (node_bytes != c->accel.node_bytes || tri_bytes != c->accel.triangle_bytes)
        return set_err(c, XR_ERR_INVALID_ARGUMENT, "need %llu node bytes and %llu triangle bytes",
                       (unsigned long long)c->accel.node_bytes, (unsigned long long)c->accel.triangle_bytes);
    XR_HIP(c, hipMemcpyAsync(nodes, c->d_nodes, node_bytes, hipMemcpyDeviceToHost, c->stream));
    XR_HIP(c, hipMemcpyAsync(tris, c->d_tris, tri_bytes, hipMemcpyDeviceToHost, c->stream));
    XR_HIP(c, hipStreamSynchronize(c->stream));
    return XR_OK;
}

int xr_resize(xr_ctx* c, uint32_t w, uint32_t h) {
    if (!c) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (w == 0 || h == 0 || (uint64_t)w * h > (1ull << 28))
        return set_err(c, XR_ERR_INVALID_ARGUMENT, "bad resolution %ux%u", w, h);
    XR_HIP(c, hipStreamSynchronize(c->stream));
    return alloc_frame(c, w, h);
}

int xr_render(xr_ctx* c, const xr_ubo* ubo, int mode, const xr_restir_params* params) {
    if (!c) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!ubo) return set_err(c, XR_ERR_INVALID_ARGUMENT, "ubo is NULL");
    if (!c->built) return set_err(c, XR_ERR_STATE, "xr_build_accel has not run");
    xr_restir_params p;
    if (params) p = *params;
    else xr_default_params(&p);
    int rc = check_params(c, &p);
    if (rc) return rc;
    PassParams pp;
    pp.temporal_m_cap = p.temporal_m_cap;
    pp.spatial_m_cap = p.spatial_m_cap;
    pp.spatial_samples = p.spatial_samples;
    pp.spatial_radius = p.spatial_radius;
    pp.visibility = p.visibility;
    pp.plain_max_depth = p.plain_max_depth;
    pp.history = 0;
    UniformBufferObject u;
    memcpy(&u, ubo, sizeof(u));
    const SceneDev sc = scene_dev(c);
    FrameDev fr = frame_dev(c);
    if (p.flags & XR_FLAG_RADIANCE) fr.radiance = c->d_radiance;
    if (fault_inject("launch")) return set_err(c, XR_ERR_DEVICE, "injected launch failure");
    if (mode != XR_MODE_PLAIN && mode != XR_MODE_RESTIR)
        return set_err(c, XR_ERR_INVALID_ARGUMENT, "unknown mode %d", mode);
    xr_ctx::EventSet& set = c->ring[c->ring_cursor];
    rc = harvest(c, set);
    if (rc) return rc;
    hipEvent_t* ev = set.ev;
    if (mode == XR_MODE_PLAIN) {
        XR_HIP(c, hipEventRecord(ev[0], c->stream));
        XR_HIP(c, launch_plain(sc, fr, u, p.plain_max_depth, c->stream));
        XR_HIP(c, hipEventRecord(ev[1], c->stream));
        set.n = 2;
    } else {
        const uint32_t passes = p.passes ? p.passes : 7u;
        const bool cache_on = (p.flags & XR_FLAG_MLP_CACHE) != 0;
        CacheCfg cache = {0u, 0u, 0u, 0.0f};
        if (cache_on) {
            rc = ensure_mlp(c);
            if (rc) return rc;
            const bool fp8 = (p.flags & XR_FLAG_MLP_FP8) != 0;
            if (c->mlp.fp8 != fp8) {  // rebuild the GEMM copies in the requested format
                c->mlp.fp8 = fp8;
                XR_HIP(c, launch_mlp_adam(c->mlp, 0.0f, 1.0f, false, c->stream));
            }
            cache.enabled = 1u;
            cache.train_target = p.mlp_train_pixels;
            cache.frame = u.currentFrame;
            cache.spread_c = p.mlp_spread;
        }
        XR_HIP(c, hipEventRecord(ev[0], c->stream));
        if (cache_on && (passes & XR_PASS_INITIAL)) XR_HIP(c, launch_clear_u32(fr.cache_list, 2, c->stream));
        if (passes & XR_PASS_INITIAL) {
            XR_HIP(c, launch_initial_sample(sc, fr, u, cache, c->stream));
            c->nbr_ok = false;  // new samples
        }
        XR_HIP(c, hipEventRecord(ev[1], c->stream));
        if (cache_on && (passes & XR_PASS_INITIAL)) {
            // one training step on this frame's path suffixes, then the
            // updated network's radiance for query and training pixels alike
            XR_HIP(c, launch_mlp_train_frame(c->mlp, fr, c->d_materials, p.mlp_loss_scale, c->stream));
            c->mlp.steps++;
            XR_HIP(c, launch_mlp_adam(c->mlp, p.mlp_lr, 1.0f / p.mlp_loss_scale, true, c->stream));
            XR_HIP(c, launch_mlp_infer(c->mlp, fr, c->d_materials, c->stream));
        }
        XR_HIP(c, hipEventRecord(ev[2], c->stream));
        // without spatial neighbours the spatial pass runs in the temporal
        // pass's kernel (its time counts as temporal)
        const bool spatial0 = (passes & XR_PASS_TEMPORAL) && (passes & XR_PASS_SPATIAL) && pp.spatial_samples == 0u;
        if (spatial0) {
            XR_HIP(c, launch_temporal_spatial0(fr, u.currentFrame, u.reuse, pp, c->d_gpos, c->stream));
            c->nbr_ok = false;  // no neighbour records written
        } else if (passes & XR_PASS_TEMPORAL) {
            XR_HIP(c, launch_temporal(fr, u.currentFrame, pp, c->d_gpos, c->stream));
            c->nbr_ok = true;
            c->nbr_ping = u.currentFrame % 2u;
        }
        XR_HIP(c, hipEventRecord(ev[3], c->stream));
        if ((passes & XR_PASS_SPATIAL) && !spatial0) {
            if (!c->nbr_ok || c->nbr_ping != u.currentFrame % 2u) {
                XR_HIP(c, launch_nbr_build(fr, u.currentFrame, c->stream));
                c->nbr_ok = true;
                c->nbr_ping = u.currentFrame % 2u;
            }
            XR_HIP(c, launch_spatial(sc, fr, u.currentFrame, u.reuse, pp, c->stream));
        }
        XR_HIP(c, hipEventRecord(ev[4], c->stream));
        set.n = 5;
    }
    set.pending = true;
    c->last_set = c->ring_cursor;
    c->ring_cursor = (c->ring_cursor + 1) % xr_ctx::kRing;
    return XR_OK;
}

int xr_sync(xr_ctx* c) {
    if (!c) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "ctx is NULL");
    XR_HIP(c, hipStreamSynchronize(c->stream));
    return XR_OK;
}

int xr_pass_times(xr_ctx* c, float* ms, uint32_t n) {
    if (!c || !ms) return set_err(c, XR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (c->last_set < 0) return 0;
    float t[4] = {0, 0, 0, 0};
    int k = 0;
    const int rc = set_times(c, c->ring[c->last_set], t, &k);
    if (rc) return -rc;
    int w = 0;
    for (; w < k && (uint32_t)w < n; ++w) ms[w] = t[w];
    return w;
}

int xr_pass_time_totals(xr_ctx* c, double* ms_sum, uint32_t n, uint32_t* frames) {
    if (!c || !ms_sum) return set_err(c, XR_ERR_INVALID_ARGUMENT, "NULL argument");
    for (auto& set : c->ring) {
        int rc = harvest(c, set);
        if (rc) return rc;
    }
    for (uint32_t i = 0; i < n && i < 4; ++i) ms_sum[i] = c->pass_sum_ms[i];
    if (frames) *frames = c->pass_frames;
    return XR_OK;
}

int xr_reset_pass_times(xr_ctx* c) {
    if (!c) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "ctx is NULL");
    for (auto& set : c->ring) {
        int rc = harvest(c, set);
        if (rc) return rc;
    }
    c->pass_sum_ms[0] = c->pass_sum_ms[1] = c->pass_sum_ms[2] = c->pass_sum_ms[3] = 0.0;
    c->pass_frames = 0;
    return XR_OK;
}

int xr_counters(xr_ctx* c, uint64_t* out, uint32_t n) {
    if (!c || !out) return set_err(c, XR_ERR_INVALID_ARGUMENT, "NULL argument");
    unsigned long long h[kCounters];
    XR_HIP(c, hipMemcpyAsync(h, c->d_counters, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    XR_HIP(c, hipStreamSynchronize(c->stream));
    for (uint32_t i = 0; i < n && i < kCounters; ++i) out[i] = h[i];
    return XR_OK;
}

int xr_reset_counters(xr_ctx* c) {
    if (!c) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "ctx is NULL");
    XR_HIP(c, hipMemsetAsync(c->d_counters, 0, sizeof(unsigned long long) * kCounters, c->stream));
    return XR_OK;
}

size_t xr_buffer_bytes(const xr_ctx* c, int which, int layout) {
    if (!c) return 0;
    const size_t n = (size_t)c->width * c->height;
    switch (which) {
        case XR_BUF_ACCUMULATION:
        case XR_BUF_GBUF_POS:
        case XR_BUF_GBUF_NRM:
        case XR_BUF_GBUF_MATUV:
        case XR_BUF_RADIANCE: return n * 16;
        case XR_BUF_CONVERGED: return n * 24;
        case XR_BUF_OUTPUT:
        case XR_BUF_MOTION: return n * 4;
        case XR_BUF_SAMPLES: return layout == XR_LAYOUT_REFERENCE ? n * sizeof(ReSTIRGISample) : n * kPlanes * 16;
        case XR_BUF_TEMPORAL:
        case XR_BUF_SPATIAL:
            return layout == XR_LAYOUT_REFERENCE ? 2 * n * sizeof(ReSTIRGIReservoir) : 2 * n * kPlanes * 16;
        default: return 0;
    }
}

static void* buffer_ptr(xr_ctx* c, int which) {
    switch (which) {
        case XR_BUF_ACCUMULATION: return c->d_accum;
        case XR_BUF_OUTPUT: return c->d_output;
        case XR_BUF_SAMPLES: return c->d_samples;
        case XR_BUF_TEMPORAL: return c->d_temporal;
        case XR_BUF_SPATIAL: return c->d_spatial;
        case XR_BUF_GBUF_POS: return c->d_gpos;
        case XR_BUF_GBUF_NRM: return c->d_gnrm;
        case XR_BUF_GBUF_MATUV: return c->d_gmatuv;
        case XR_BUF_MOTION: return c->d_motion;
        case XR_BUF_RADIANCE: return c->d_radiance;
        case XR_BUF_CONVERGED: return c->d_conv;
        default: return nullptr;
    }
}

static int ensure_conv(xr_ctx* c) {
    if (c->d_conv) return XR_OK;
    const size_t bytes = (size_t)c->width * c->height * 3 * sizeof(double);
    XR_HIP(c, dev_alloc(&c->d_conv, bytes));
    XR_HIP(c, hipMemsetAsync(c->d_conv, 0, bytes, c->stream));
    c->conv_spp = 0;
    return XR_OK;
}

int xr_converge(xr_ctx* c, const xr_ubo* ubo, uint32_t spp, uint32_t frame_base) {
    if (!c) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!ubo) return set_err(c, XR_ERR_INVALID_ARGUMENT, "ubo is NULL");
    if (!c->built) return set_err(c, XR_ERR_STATE, "xr_build_accel has not run");
    int rc = check_textures(c);
    if (rc) return rc;
    rc = ensure_conv(c);
    if (rc) return rc;
    UniformBufferObject u;
    memcpy(&u, ubo, sizeof(u));
    // Launch in slices of at most 64 samples so no single kernel runs long.
    for (uint32_t done = 0; done < spp;) {
        const uint32_t k = spp - done < 64u ? spp - done : 64u;
        XR_HIP(c, launch_converge(scene_dev(c), frame_dev(c), u, frame_base + done, k, c->d_conv, c->stream));
        done += k;
    }
    c->conv_spp += spp;
    return XR_OK;
}

int xr_converge_reset(xr_ctx* c) {
    if (!c) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (c->d_conv)
        XR_HIP(c, hipMemsetAsync(c->d_conv, 0, (size_t)c->width * c->height * 3 * sizeof(double), c->stream));
    c->conv_spp = 0;
    return XR_OK;
}

uint64_t xr_converged_spp(const xr_ctx* c) { return c ? c->conv_spp : 0; }

int xr_rmse(xr_ctx* c, double* rmse) {
    if (!c || !rmse) return set_err(c, XR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!c->d_conv || c->conv_spp == 0) return set_err(c, XR_ERR_STATE, "no converged samples (xr_converge)");
    if (!c->d_err) XR_HIP(c, dev_alloc(&c->d_err, sizeof(double)));
    const uint32_t n = c->width * c->height;
    XR_HIP(c, hipMemsetAsync(c->d_err, 0, sizeof(double), c->stream));
    XR_HIP(c, launch_sq_error(c->d_radiance, c->d_conv, 1.0 / (double)c->conv_spp, n, c->d_err, c->stream));
    double sum = 0.0;
    XR_HIP(c, hipMemcpyAsync(&sum, c->d_err, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    XR_HIP(c, hipStreamSynchronize(c->stream));
    *rmse = sqrt(sum / (3.0 * (double)n));
    return XR_OK;
}

// plane p of entry j (planes stored as [slot][plane][pixel])
static inline const float* plane_at(const std::vector<float4>& v, size_t n, size_t j, int p) {
    const size_t slot = j / n, i = j % n;
    return &v[(slot * kPlanes + p) * n + i].x;
}

}  // extern "C"

// Copies buffer `which` of an image of n pixels at device address src to dst
// (bytes already checked), converting the sample / reservoir planes to the
// reference AoS layout when asked.
static int read_device_buffer(xr_ctx* c, const void* src, size_t n, int which, void* dst, size_t bytes,
                              int layout) {
    const bool planar = which == XR_BUF_SAMPLES || which == XR_BUF_TEMPORAL || which == XR_BUF_SPATIAL;
    if (!planar || layout == XR_LAYOUT_NATIVE) {
        XR_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
        XR_HIP(c, hipStreamSynchronize(c->stream));
        return XR_OK;
    }
    const size_t entries = which == XR_BUF_SAMPLES ? n : 2 * n;
    std::vector<float4> tmp;
    try {
        tmp.resize(entries * kPlanes);
    } catch (const std::bad_alloc&) {
        return set_err(c, XR_ERR_OUT_OF_MEMORY, "host staging allocation failed");
    }
    XR_HIP(c, hipMemcpyAsync(tmp.data(), src, tmp.size() * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    XR_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t j = 0; j < entries; ++j) {
        const float* q[kPlanes];
        for (int p = 0; p < kPlanes; ++p) q[p] = plane_at(tmp, n, j, p);
        ReSTIRGISample s;
        memset(&s, 0, sizeof(s));
        memcpy(s.x_view, q[0], 12);
        memcpy(&s.materialID, q[0] + 3, 4);
        memcpy(s.n_view, q[1], 12);
        memcpy(&s.randSeed, q[1] + 3, 4);
        memcpy(s.x_sample, q[2], 12);
        memcpy(&s.p_q, q[2] + 3, 4);
        memcpy(s.n_sample, q[3], 12);
        memcpy(s.L_out, q[4], 12);
        memcpy(s.f, q[5], 12);
        if (which == XR_BUF_SAMPLES) {
            // spare words: 0 on the reference path; with the radiance cache,
            // pad1 = material at x_s and pad2 = 1 (query) / 2 (training)
            memcpy(&s.pad1, q[3] + 3, 4);
            memcpy(&s.pad2, q[4] + 3, 4);
            memcpy(&s.pad3, q[5] + 3, 4);
            memcpy((char*)dst + j * sizeof(ReSTIRGISample), &s, sizeof(s));
        } else {
            ReSTIRGIReservoir r;
            memset(&r, 0, sizeof(r));
            memcpy(&r.w_sum, q[3] + 3, 4);
            memcpy(&r.W, q[4] + 3, 4);
            memcpy(&r.m, q[5] + 3, 4);
            r.s = s;
            memcpy((char*)dst + j * sizeof(ReSTIRGIReservoir), &r, sizeof(r));
        }
    }
    return XR_OK;
}

extern "C" {

int xr_read_buffer(xr_ctx* c, int which, void* dst, size_t bytes, int layout) {
    if (!c || !dst) return set_err(c, XR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (layout != XR_LAYOUT_REFERENCE && layout != XR_LAYOUT_NATIVE)
        return set_err(c, XR_ERR_INVALID_ARGUMENT, "unknown layout %d", layout);
    const size_t want = xr_buffer_bytes(c, which, layout);
    if (want == 0) return set_err(c, XR_ERR_INVALID_ARGUMENT, "unknown buffer %d", which);
    if (bytes != want) return set_err(c, XR_ERR_INVALID_ARGUMENT, "buffer %d needs %zu bytes, got %zu", which, want, bytes);
    const void* src = buffer_ptr(c, which);
    if (!src) return set_err(c, XR_ERR_STATE, "buffer %d has not been created yet", which);
    return read_device_buffer(c, src, (size_t)c->width * c->height, which, dst, bytes, layout);
}

int xr_write_buffer(xr_ctx* c, int which, const void* src, size_t bytes, int layout) {
    if (!c || !src) return set_err(c, XR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (layout != XR_LAYOUT_REFERENCE && layout != XR_LAYOUT_NATIVE)
        return set_err(c, XR_ERR_INVALID_ARGUMENT, "unknown layout %d", layout);
    const size_t want = xr_buffer_bytes(c, which, layout);
    if (want == 0) return set_err(c, XR_ERR_INVALID_ARGUMENT, "unknown buffer %d", which);
    if (bytes != want) return set_err(c, XR_ERR_INVALID_ARGUMENT, "buffer %d needs %zu bytes, got %zu", which, want, bytes);
    void* dst = buffer_ptr(c, which);
    if (!dst) return set_err(c, XR_ERR_STATE, "buffer %d has not been created yet", which);
    const bool planar = which == XR_BUF_SAMPLES || which == XR_BUF_TEMPORAL || which == XR_BUF_SPATIAL;
    if (which == XR_BUF_SAMPLES || which == XR_BUF_TEMPORAL) c->nbr_ok = false;
    if (!planar || layout == XR_LAYOUT_NATIVE) {
        XR_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
        XR_HIP(c, hipStreamSynchronize(c->stream));
        return XR_OK;
    }
    const size_t n = (size_t)c->width * c->height;
    const size_t entries = which == XR_BUF_SAMPLES ? n : 2 * n;
    std::vector<float4> tmp;
    try {
        tmp.resize(entries * kPlanes);
    } catch (const std::bad_alloc&) {
        return set_err(c, XR_ERR_OUT_OF_MEMORY, "host staging allocation failed");
    }
    for (size_t j = 0; j < entries; ++j) {
        ReSTIRGISample s;
        ReSTIRGIReservoir r;
        memset(&r, 0, sizeof(r));
        if (which == XR_BUF_SAMPLES) {
            memcpy(&s, (const char*)src + j * sizeof(s), sizeof(s));
        } else {
            memcpy(&r, (const char*)src + j * sizeof(r), sizeof(r));
            s = r.s;
        }
        const size_t slot = j / n, i = j % n;
        float4* q[kPlanes];
        for (int p = 0; p < kPlanes; ++p) q[p] = &tmp[(slot * kPlanes + p) * n + i];
        memcpy(&q[0]->x, s.x_view, 12);
        memcpy(&q[0]->w, &s.materialID, 4);
        memcpy(&q[1]->x, s.n_view, 12);
        memcpy(&q[1]->w, &s.randSeed, 4);
        memcpy(&q[2]->x, s.x_sample, 12);
        q[2]->w = s.p_q;
        memcpy(&q[3]->x, s.n_sample, 12);
        memcpy(&q[4]->x, s.L_out, 12);
        memcpy(&q[5]->x, s.f, 12);
        if (which == XR_BUF_SAMPLES) {
            memcpy(&q[3]->w, &s.pad1, 4);
            memcpy(&q[4]->w, &s.pad2, 4);
            memcpy(&q[5]->w, &s.pad3, 4);
        } else {
            q[3]->w = r.w_sum;
            q[4]->w = r.W;
            memcpy(&q[5]->w, &r.m, 4);
        }
    }
    XR_HIP(c, hipMemcpyAsync(dst, tmp.data(), tmp.size() * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    XR_HIP(c, hipStreamSynchronize(c->stream));
    return XR_OK;
}

int xr_read_image(xr_ctx* c, uint32_t* dst, size_t bytes) {
    return xr_read_buffer(c, XR_BUF_OUTPUT, dst, bytes, XR_LAYOUT_NATIVE);
}

int xr_trace_rays_device(xr_ctx* c, const xr_ray* d_rays, uint32_t n, xr_hit* d_hits, int any_hit) {
    if (!c) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (n && (!d_rays || !d_hits)) return set_err(c, XR_ERR_INVALID_ARGUMENT, "NULL ray or hit array");
    if (!c->built) return set_err(c, XR_ERR_STATE, "xr_build_accel has not run");
    XR_HIP(c, launch_trace(scene_dev(c), (const float4*)d_rays, n, (float4*)d_hits, any_hit != 0, c->stream));
    return XR_OK;
}

int xr_trace_rays(xr_ctx* c, const xr_ray* rays, uint32_t n, xr_hit* hits, int any_hit) {
    if (!c) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (n == 0) return XR_OK;
    if (!rays || !hits) return set_err(c, XR_ERR_INVALID_ARGUMENT, "NULL ray or hit array");
    if (!c->built) return set_err(c, XR_ERR_STATE, "xr_build_accel has not run");
    if (n > c->ray_cap) {
        dev_free(&c->d_rays);
        dev_free(&c->d_hits);
        c->ray_cap = 0;
        XR_HIP(c, dev_alloc(&c->d_rays, sizeof(xr_ray) * (size_t)n));
        XR_HIP(c, dev_alloc(&c->d_hits, sizeof(xr_hit) * (size_t)n));
        c->ray_cap = n;
    }
    XR_HIP(c, hipMemcpyAsync(c->d_rays, rays, sizeof(xr_ray) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    XR_HIP(c, launch_trace(scene_dev(c), c->d_rays, n, c->d_hits, any_hit != 0, c->stream));
    XR_HIP(c, hipMemcpyAsync(hits, c->d_hits, sizeof(xr_hit) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    XR_HIP(c, hipStreamSynchronize(c->stream));
    return XR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// MLP radiance cache entry points
// ---------------------------------------------------------------------------
extern "C" {

int xr_mlp_reset(xr_ctx* c, uint64_t seed) {
    if (!c) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!c->d_mlp) {
        const int rc = ensure_mlp(c);
        if (rc) return rc;
    }
    return mlp_init(c, seed);
}

int xr_mlp_eval(xr_ctx* c, const float* inputs, uint32_t rows, float* outputs) {
    if (!c || (rows && (!inputs || !outputs))) return set_err(c, XR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (rows == 0) return XR_OK;
    int rc = ensure_mlp(c);
    if (rc) return rc;
    rc = ensure_mlp_io(c, (size_t)rows * (kMlpIn + 3));
    if (rc) return rc;
    float* din = c->d_mlp_io;
    float* dout = c->d_mlp_io + (size_t)rows * kMlpIn;
    XR_HIP(c, hipMemcpyAsync(din, inputs, sizeof(float) * rows * kMlpIn, hipMemcpyHostToDevice, c->stream));
    XR_HIP(c, launch_mlp_eval(c->mlp, din, rows, dout, c->stream));
    XR_HIP(c, hipMemcpyAsync(outputs, dout, sizeof(float) * rows * 3, hipMemcpyDeviceToHost, c->stream));
    XR_HIP(c, hipStreamSynchronize(c->stream));
    return XR_OK;
}

static int stage_rows(xr_ctx* c, const float* inputs, const float* targets, uint32_t rows) {
    int rc = ensure_mlp_io(c, (size_t)rows * (kMlpIn + 3));
    if (rc) return rc;
    XR_HIP(c, hipMemcpyAsync(c->d_mlp_io, inputs, sizeof(float) * rows * kMlpIn, hipMemcpyHostToDevice, c->stream));
    XR_HIP(c, hipMemcpyAsync(c->d_mlp_io + (size_t)rows * kMlpIn, targets, sizeof(float) * rows * 3,
                             hipMemcpyHostToDevice, c->stream));
    return XR_OK;
}

int xr_mlp_gradient(xr_ctx* c, const float* inputs, const float* targets, uint32_t rows, float* grad_out) {
    if (!c || !inputs || !targets || !grad_out || rows == 0)
        return set_err(c, XR_ERR_INVALID_ARGUMENT, "NULL argument or no rows");
    int rc = ensure_mlp(c);
    if (rc) return rc;
    rc = stage_rows(c, inputs, targets, rows);
    if (rc) return rc;
    const size_t gbytes = sizeof(float) * (size_t)kMlpParams * 32;
    XR_HIP(c, hipMemsetAsync(c->mlp.grad, 0, gbytes, c->stream));
    XR_HIP(c, launch_mlp_train_rows(c->mlp, c->d_mlp_io, c->d_mlp_io + (size_t)rows * kMlpIn, rows, 1.0f, c->stream));
    std::vector<float> parts((size_t)kMlpParams * 32);
    XR_HIP(c, hipMemcpyAsync(parts.data(), c->mlp.grad, gbytes, hipMemcpyDeviceToHost, c->stream));
    XR_HIP(c, hipMemsetAsync(c->mlp.grad, 0, gbytes, c->stream));
    XR_HIP(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < kMlpParams; ++i) {
        double g = 0.0;
        for (int p = 0; p < 32; ++p) g += parts[(size_t)p * kMlpParams + i];
        grad_out[i] = (float)g;
    }
    return XR_OK;
}

int xr_mlp_gradient_noalloc(xr_ctx* c, const float* d_inputs, const float* d_targets, uint32_t rows,
                            float* d_grad_out) {
    if (!c) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!d_inputs || !d_targets || !d_grad_out || rows == 0)
        return set_err(c, XR_ERR_INVALID_ARGUMENT, "NULL device pointer or no rows");
    if (rows > XR_MLP_NOALLOC_MAX_ROWS)
        return set_err(c, XR_ERR_INVALID_ARGUMENT, "xr_mlp_gradient_noalloc: %u rows exceed the maximum of %u", rows,
                       (unsigned)XR_MLP_NOALLOC_MAX_ROWS);
    if (!c->d_mlp) return set_err(c, XR_ERR_STATE, "xr_mlp_gradient_noalloc: no network yet (call xr_mlp_reset)");
    if (!c->noalloc_scratch) return set_err(c, XR_ERR_STATE, "xr_mlp_gradient_noalloc: scratch address not resolved");
    XR_HIP(c, launch_mlp_gradient_noalloc(c->mlp, c->noalloc_scratch, d_inputs, d_targets, rows, d_grad_out,
                                          c->stream));
    return XR_OK;
}

int xr_mlp_train_rows(xr_ctx* c, const float* inputs, const float* targets, uint32_t rows, float lr) {
    if (!c || !inputs || !targets || rows == 0) return set_err(c, XR_ERR_INVALID_ARGUMENT, "NULL argument or no rows");
    if (!(lr >= 0.0f && lr < 1.0f)) return set_err(c, XR_ERR_INVALID_ARGUMENT, "lr out of range");
    int rc = ensure_mlp(c);
    if (rc) return rc;
    rc = stage_rows(c, inputs, targets, rows);
    if (rc) return rc;
    const float scale = 128.0f;
    XR_HIP(c, launch_mlp_train_rows(c->mlp, c->d_mlp_io, c->d_mlp_io + (size_t)rows * kMlpIn, rows, scale,
                                    c->stream));
    c->mlp.steps++;
    XR_HIP(c, launch_mlp_adam(c->mlp, lr, 1.0f / scale, true, c->stream));
    XR_HIP(c, hipStreamSynchronize(c->stream));
    return XR_OK;
}

int xr_mlp_params(xr_ctx* c, float* params, int set) {
    if (!c || !params) return set_err(c, XR_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = ensure_mlp(c);
    if (rc) return rc;
    if (set) {
        XR_HIP(c, hipMemcpyAsync(c->mlp.w, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, c->stream));
        XR_HIP(c, launch_mlp_adam(c->mlp, 0.0f, 1.0f, false, c->stream));
    }
    XR_HIP(c, hipMemcpyAsync(params, c->mlp.w, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, c->stream));
    XR_HIP(c, hipStreamSynchronize(c->stream));
    return XR_OK;
}

int xr_mlp_set_weight_format(xr_ctx* c, int format) {
    if (!c) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (format != XR_MLP_WEIGHTS_BF16 && format != XR_MLP_WEIGHTS_FP8)
        return set_err(c, XR_ERR_INVALID_ARGUMENT, "unknown weight format %d", format);
    int rc = ensure_mlp(c);
    if (rc) return rc;
    c->mlp.fp8 = format == XR_MLP_WEIGHTS_FP8;
    XR_HIP(c, launch_mlp_adam(c->mlp, 0.0f, 1.0f, false, c->stream));
    XR_HIP(c, hipStreamSynchronize(c->stream));
    return XR_OK;
}

int xr_mlp_fp8_weights(xr_ctx* c, uint8_t* codes, float* scales) {
    if (!c || !codes || !scales) return set_err(c, XR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!c->d_mlp || !c->mlp.fp8) return set_err(c, XR_ERR_STATE, "the network is not in the FP8 weight format");
    XR_HIP(c, hipMemcpyAsync(codes, c->mlp.wt_f8, XR_MLP_FP8_BYTES, hipMemcpyDeviceToHost, c->stream));
    XR_HIP(c, hipMemcpyAsync(scales, c->mlp.f8_scale, 5 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    XR_HIP(c, hipStreamSynchronize(c->stream));
    return XR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched views
// ---------------------------------------------------------------------------
struct xr_batch {
    xr_ctx* ctx = nullptr;  // scene, BVH, stream, errors
    uint32_t nviews = 0;
    struct View {
        uint32_t w = 0, h = 0;
        float4 *samples = nullptr, *temporal = nullptr, *spatial = nullptr, *gpos = nullptr, *gnrm = nullptr,
               *gmatuv = nullptr, *nbr = nullptr;
        uint32_t *motion = nullptr, *output = nullptr;
        // radiance cache (allocated by the first frame that uses it)
        uint32_t *cache_list = nullptr, *cache_work = nullptr;
        float4* pending = nullptr;
    };
    std::vector<View> views;
    std::vector<FrameDev> frames;  // host copy of d_frames
    FrameDev* d_frames = nullptr;
    UniformBufferObject* d_ubos = nullptr;
    float4* d_tail = nullptr;
    uint32_t* d_tail_count = nullptr;
    unsigned long long* d_counters = nullptr;
    uint32_t* d_pending_counts = nullptr;  // per view (radiance cache)
    bool cache_ready = false;
    uint32_t cfg = 0;
    bool use_graph = true;
    // Captured graphs: [0] the whole frame; with the radiance cache [1] the
    // initial-sample pass and [2] the temporal and spatial passes, the
    // training step running eagerly between them (its Adam bias correction
    // changes every step).
    struct Graph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        PassParams pp{};
        CacheCfg cache{};
        uint32_t cfg = 0;
        uint64_t scene_gen = ~0ull;  // the context's scene generation at the capture
    };
    Graph graphs[3];
    hipEvent_t e0 = nullptr, e1 = nullptr;
};

namespace {

void batch_free(xr_batch* b) {
    for (auto& v : b->views) {
        dev_free(&v.samples);
        dev_free(&v.temporal);
        dev_free(&v.spatial);
        dev_free(&v.gpos);
        dev_free(&v.gnrm);
        dev_free(&v.gmatuv);
        dev_free(&v.motion);
        dev_free(&v.output);
        dev_free(&v.nbr);
        dev_free(&v.cache_list);
        dev_free(&v.cache_work);
        dev_free(&v.pending);
    }
    dev_free(&b->d_pending_counts);
    dev_free(&b->d_frames);
    dev_free(&b->d_ubos);
    dev_free(&b->d_tail);
    dev_free(&b->d_tail_count);
    dev_free(&b->d_counters);
    for (auto& g : b->graphs) {
        if (g.exec) hipGraphExecDestroy(g.exec);
        if (g.graph) hipGraphDestroy(g.graph);
        g = xr_batch::Graph();
    }
    if (b->e0) hipEventDestroy(b->e0);
    if (b->e1) hipEventDestroy(b->e1);
}

// Zeroes every view's buffers (history included).
int batch_clear(xr_batch* b) {
    xr_ctx* c = b->ctx;
    const size_t f4 = sizeof(float4);
    for (auto& v : b->views) {
        const size_t n = (size_t)v.w * v.h;
        XR_HIP(c, hipMemsetAsync(v.samples, 0, kPlanes * n * f4, c->stream));
        XR_HIP(c, hipMemsetAsync(v.temporal, 0, 2 * kPlanes * n * f4, c->stream));
        XR_HIP(c, hipMemsetAsync(v.spatial, 0, 2 * kPlanes * n * f4, c->stream));
        XR_HIP(c, hipMemsetAsync(v.gpos, 0, n * f4, c->stream));
        XR_HIP(c, hipMemsetAsync(v.gnrm, 0, n * f4, c->stream));
        XR_HIP(c, hipMemsetAsync(v.gmatuv, 0, n * f4, c->stream));
        XR_HIP(c, hipMemsetAsync(v.motion, 0, n * 4, c->stream));
        XR_HIP(c, hipMemsetAsync(v.output, 0, n * 4, c->stream));
    }
    XR_HIP(c, hipStreamSynchronize(c->stream));
    return XR_OK;
}

BatchDev batch_dev(const xr_batch* b) {
    BatchDev bd;
    bd.views = b->d_frames;
    bd.ubos = b->d_ubos;
    bd.nviews = b->nviews;
    bd.tail = b->d_tail;
    bd.tail_count = b->d_tail_count;
    bd.tile_queue = b->d_tail_count + 1;
    bd.counters = b->d_counters;
    return bd;
}

PassParams pass_params(const xr_restir_params& p) {
    PassParams pp;
    pp.temporal_m_cap = p.temporal_m_cap;
    pp.spatial_m_cap = p.spatial_m_cap;
    pp.spatial_samples = p.spatial_samples;
    pp.spatial_radius = p.spatial_radius;
    pp.visibility = p.visibility;
    pp.plain_max_depth = p.plain_max_depth;
    pp.history = 0;
    return pp;
}

bool same_pp(const PassParams& a, const PassParams& b) { return memcmp(&a, &b, sizeof(a)) == 0; }

// Per-view radiance-cache buffers (work lists, training choice, pending
// records), allocated by the first frame that uses the cache.
int ensure_batch_cache(xr_batch* b) {
    if (b->cache_ready) return XR_OK;
    xr_ctx* c = b->ctx;
    XR_HIP(c, hipStreamSynchronize(c->stream));
    XR_HIP(c, dev_alloc(&b->d_pending_counts, sizeof(uint32_t) * b->nviews));
    for (uint32_t v = 0; v < b->nviews; ++v) {
        xr_batch::View& vw = b->views[v];
        const size_t n = (size_t)vw.w * vw.h;
        XR_HIP(c, dev_alloc(&vw.cache_list, (2 + 2 * n) * sizeof(uint32_t)));
        XR_HIP(c, dev_alloc(&vw.cache_work, (n + (n + kCacheSelBlock - 1) / kCacheSelBlock + 1) * sizeof(uint32_t)));
        XR_HIP(c, dev_alloc(&vw.pending, 4 * n * sizeof(float4)));
        FrameDev& f = b->frames[v];
        f.cache_list = vw.cache_list;
        f.cache_slot = vw.cache_work;
        f.cache_blocks = vw.cache_work + n;
        f.pending = vw.pending;
        f.pending_count = b->d_pending_counts + v;
    }
    XR_HIP(c, hipMemcpyAsync(b->d_frames, b->frames.data(), sizeof(FrameDev) * b->nviews, hipMemcpyHostToDevice,
                             c->stream));
    XR_HIP(c, hipStreamSynchronize(c->stream));
    for (auto& g : b->graphs) {  // their launches took the view table by value
        if (g.exec) hipGraphExecDestroy(g.exec);
        if (g.graph) hipGraphDestroy(g.graph);
        g = xr_batch::Graph();
    }
    b->cache_ready = true;
    return XR_OK;
}

// Replays graph slot g, capturing it again first when the parameters, the
// cache settings, the kernel configuration or the scene's device memory
// (textures, BVH) changed since its capture.
template <class Launch>
int batch_graph(xr_batch* b, xr_batch::Graph& g, const PassParams& pp, const CacheCfg& cache, Launch launch) {
    xr_ctx* c = b->ctx;
    const bool same = g.exec && same_pp(pp, g.pp) && g.cfg == b->cfg && g.scene_gen == c->scene_gen &&
                      g.cache.enabled == cache.enabled && g.cache.train_target == cache.train_target &&
                      g.cache.spread_c == cache.spread_c;
    if (!same) {
        if (g.exec) hipGraphExecDestroy(g.exec);
        if (g.graph) hipGraphDestroy(g.graph);
        g = xr_batch::Graph();
        XR_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const hipError_t le = launch();
        hipGraph_t graph = nullptr;
        const hipError_t ce = hipStreamEndCapture(c->stream, &graph);
        XR_HIP(c, le);
        XR_HIP(c, ce);
        g.graph = graph;
        XR_HIP(c, hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
        g.pp = pp;
        g.cache = cache;
        g.cfg = b->cfg;
        g.scene_gen = c->scene_gen;
    }
    XR_HIP(c, hipGraphLaunch(g.exec, c->stream));
    return XR_OK;
}

// One frame of every view: eager launches or captured graphs.  With the
// radiance cache the initial-sample pass (each view choosing its own training
// samples, as a single-view frame would) is followed by one training step on
// the samples of all views (the mean loss over their union), then the
// network's radiance for every view's query and training pixels, then the
// temporal and spatial passes.
int batch_frame(xr_batch* b, const PassParams& pp, const CacheCfg& cache, float loss_scale, float lr) {
    xr_ctx* c = b->ctx;
    const SceneDev sc = scene_dev(c);
    const BatchDev bd = batch_dev(b);
    const FrameDev* views = b->frames.data();
    hipStream_t st = c->stream;
    if (!cache.enabled) {
        if (!b->use_graph) {
            XR_HIP(c, launch_batch_frame(sc, bd, views, b->cfg, pp, st));
            return XR_OK;
        }
        return batch_graph(b, b->graphs[0], pp, cache, [&] { return launch_batch_frame(sc, bd, views, b->cfg, pp, st); });
    }
    if (!b->use_graph) {
        XR_HIP(c, launch_batch_initial(sc, bd, views, b->cfg, cache, st));
    } else {
        const int rc = batch_graph(b, b->graphs[1], pp, cache,
                                   [&] { return launch_batch_initial(sc, bd, views, b->cfg, cache, st); });
        if (rc) return rc;
    }
    uint32_t* total = b->d_tail_count + kBatchTrainTotal;
    XR_HIP(c, launch_batch_train_total(bd, total, st));
    for (uint32_t v = 0; v < b->nviews; ++v)
        XR_HIP(c, launch_mlp_train_frame(c->mlp, views[v], c->d_materials, loss_scale, st, total));
    c->mlp.steps++;
    XR_HIP(c, launch_mlp_adam(c->mlp, lr, 1.0f / loss_scale, true, st));
    for (uint32_t v = 0; v < b->nviews; ++v) XR_HIP(c, launch_mlp_infer(c->mlp, views[v], c->d_materials, st));
    if (!b->use_graph) {
        XR_HIP(c, launch_batch_reuse(sc, bd, views, b->cfg, pp, st));
        return XR_OK;
    }
    return batch_graph(b, b->graphs[2], pp, cache, [&] { return launch_batch_reuse(sc, bd, views, b->cfg, pp, st); });
}

}  // namespace

extern "C" {

int xr_batch_create(xr_ctx* c, uint32_t view_count, const uint32_t* widths, const uint32_t* heights,
                    xr_batch** out) {
    if (!out) return set_err(c, XR_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (!c || !widths || !heights) return set_err(c, XR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (view_count == 0 || view_count > 64) return set_err(c, XR_ERR_INVALID_ARGUMENT, "view_count must be 1..64");
    if (!c->built) return set_err(c, XR_ERR_STATE, "build the acceleration structure first");
    size_t total = 0;
    for (uint32_t v = 0; v < view_count; ++v) {
        if (widths[v] == 0 || heights[v] == 0 || widths[v] > 16384 || heights[v] > 16384)
            return set_err(c, XR_ERR_INVALID_ARGUMENT, "view %u size %ux%u out of range", v, widths[v], heights[v]);
        total += (size_t)widths[v] * heights[v];
    }
    if (total > 0xFFFFFFFFull / 4) return set_err(c, XR_ERR_INVALID_ARGUMENT, "batch has too many pixels");
    xr_batch* b = new (std::nothrow) xr_batch;
    if (!b) return set_err(c, XR_ERR_OUT_OF_MEMORY, "host allocation failed");
    b->ctx = c;
    b->nviews = view_count;
    b->views.resize(view_count);
    b->frames.resize(view_count);
    const size_t f4 = sizeof(float4);
    int rc = XR_OK;
#define XR_BATCH_HIP(expr)                                                                                \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) {                                                                           \
            rc = set_err(c, hip_code(e_), "%s: %s", #expr, hipGetErrorString(e_));                       \
            goto fail;                                                                                    \
        }                                                                                                 \
    } while (0)
    for (uint32_t v = 0; v < view_count; ++v) {
        xr_batch::View& vw = b->views[v];
        vw.w = widths[v];
        vw.h = heights[v];
        const size_t n = (size_t)vw.w * vw.h;
        XR_BATCH_HIP(dev_alloc(&vw.samples, kPlanes * n * f4));
        XR_BATCH_HIP(dev_alloc(&vw.temporal, 2 * kPlanes * n * f4));
        XR_BATCH_HIP(dev_alloc(&vw.spatial, 2 * kPlanes * n * f4));
        XR_BATCH_HIP(dev_alloc(&vw.gpos, n * f4));
        XR_BATCH_HIP(dev_alloc(&vw.gnrm, n * f4));
        XR_BATCH_HIP(dev_alloc(&vw.gmatuv, n * f4));
        XR_BATCH_HIP(dev_alloc(&vw.motion, n * 4));
        XR_BATCH_HIP(dev_alloc(&vw.output, n * 4));
        XR_BATCH_HIP(dev_alloc(&vw.nbr, 4 * n * f4));
    }
    XR_BATCH_HIP(dev_alloc(&b->d_frames, sizeof(FrameDev) * view_count));
    XR_BATCH_HIP(dev_alloc(&b->d_ubos, sizeof(UniformBufferObject) * view_count));
    XR_BATCH_HIP(dev_alloc(&b->d_tail, 4 * total * f4));
    XR_BATCH_HIP(dev_alloc(&b->d_tail_count, kBatchWords * sizeof(uint32_t)));
    XR_BATCH_HIP(dev_alloc(&b->d_counters, sizeof(unsigned long long) * kCounters));
    XR_BATCH_HIP(hipMemsetAsync(b->d_counters, 0, sizeof(unsigned long long) * kCounters, c->stream));
    XR_BATCH_HIP(hipEventCreate(&b->e0));
    XR_BATCH_HIP(hipEventCreate(&b->e1));
    for (uint32_t v = 0; v < view_count; ++v) {
        const xr_batch::View& vw = b->views[v];
        FrameDev f{};
        f.width = vw.w;
        f.height = vw.h;
        f.n = vw.w * vw.h;
        f.samples = vw.samples;
        f.temporal = vw.temporal;
        f.spatial = vw.spatial;
        f.gpos = vw.gpos;
        f.gnrm = vw.gnrm;
        f.gmatuv = vw.gmatuv;
        f.motion = vw.motion;
        f.output = vw.output;
        f.accum = nullptr;
        f.nbr = vw.nbr;
        f.radiance = nullptr;
        f.counters = b->d_counters;
        f.cache_list = nullptr;
        f.cache_slot = nullptr;
        f.cache_blocks = nullptr;
        f.pending = nullptr;
        f.pending_count = nullptr;
        f.tail = b->d_tail;
        f.tail_count = b->d_tail_count;
        b->frames[v] = f;
    }
    XR_BATCH_HIP(hipMemcpyAsync(b->d_frames, b->frames.data(), sizeof(FrameDev) * view_count, hipMemcpyHostToDevice,
                                c->stream));
#undef XR_BATCH_HIP
    rc = batch_clear(b);
    if (rc) goto fail;
    *out = b;
    return XR_OK;
fail:
    batch_free(b);
    delete b;
    return rc;
}

int xr_batch_render(xr_batch* b, const xr_ubo* ubos, const xr_restir_params* params) {
    if (!b || !ubos) return set_err(b ? b->ctx : nullptr, XR_ERR_INVALID_ARGUMENT, "NULL argument");
    xr_ctx* c = b->ctx;
    xr_restir_params p;
    if (params) p = *params;
    else xr_default_params(&p);
    int rc = check_params(c, &p);
    if (rc) return rc;
    if (p.passes && p.passes != 7u) return set_err(c, XR_ERR_UNSUPPORTED, "batched views run every pass");
    if (!c->built) return set_err(c, XR_ERR_STATE, "acceleration structure not built");
    CacheCfg cache = {0u, 0u, 0u, 0.0f};
    if (p.flags & XR_FLAG_MLP_CACHE) {
        rc = ensure_mlp(c);
        if (rc) return rc;
        const bool fp8 = (p.flags & XR_FLAG_MLP_FP8) != 0;
        if (c->mlp.fp8 != fp8) {  // rebuild the GEMM copies in the requested format
            c->mlp.fp8 = fp8;
            XR_HIP(c, launch_mlp_adam(c->mlp, 0.0f, 1.0f, false, c->stream));
        }
        rc = ensure_batch_cache(b);
        if (rc) return rc;
        cache.enabled = 1u;
        cache.train_target = p.mlp_train_pixels;
        cache.spread_c = p.mlp_spread;  // the frame number comes from each view's UBO
    }
    // pageable source: the copy is staged before the call returns
    XR_HIP(c, hipMemcpyAsync(b->d_ubos, ubos, sizeof(UniformBufferObject) * b->nviews, hipMemcpyHostToDevice,
                             c->stream));
    XR_HIP(c, hipEventRecord(b->e0, c->stream));
    rc = batch_frame(b, pass_params(p), cache, p.mlp_loss_scale, p.mlp_lr);
    if (rc) return rc;
    XR_HIP(c, hipEventRecord(b->e1, c->stream));
    return XR_OK;
}

int xr_batch_frame_time(xr_batch* b, float* ms) {
    if (!b || !ms) return set_err(b ? b->ctx : nullptr, XR_ERR_INVALID_ARGUMENT, "NULL argument");
    XR_HIP(b->ctx, hipEventSynchronize(b->e1));
    XR_HIP(b->ctx, hipEventElapsedTime(ms, b->e0, b->e1));
    return XR_OK;
}

int xr_batch_set_graph(xr_batch* b, int enable) {
    if (!b) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "batch is NULL");
    b->use_graph = enable != 0;
    return XR_OK;
}

uint32_t xr_batch_config_count(void) { return batch_config_count(); }

int xr_batch_config_info(uint32_t config, uint32_t* tile_w, uint32_t* tile_h, uint32_t* lds_depth) {
    if (config >= batch_config_count() || !tile_w || !tile_h || !lds_depth)
        return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "config %u / NULL argument", config);
    bool persistent = false;
    batch_config_info(config, tile_w, tile_h, lds_depth, &persistent);
    return XR_OK;
}

int xr_batch_config_persistent(uint32_t config, int* persistent) {
    if (config >= batch_config_count() || !persistent)
        return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "config %u / NULL argument", config);
    uint32_t tw, th, lds;
    bool p = false;
    batch_config_info(config, &tw, &th, &lds, &p);
    *persistent = p ? 1 : 0;
    return XR_OK;
}

int xr_batch_set_config(xr_batch* b, uint32_t config) {
    if (!b) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (config >= batch_config_count()) return set_err(b->ctx, XR_ERR_INVALID_ARGUMENT, "config %u", config);
    b->cfg = config;
    return XR_OK;
}

uint32_t xr_batch_config(const xr_batch* b) { return b ? b->cfg : 0u; }

int xr_batch_autotune(xr_batch* b, const xr_ubo* ubos, const xr_restir_params* params, uint32_t frames,
                      float* ms_per_config) {
    if (!b || !ubos) return set_err(b ? b->ctx : nullptr, XR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (frames == 0 || frames > 64) return set_err(b->ctx, XR_ERR_INVALID_ARGUMENT, "frames must be 1..64");
    xr_ctx* c = b->ctx;
    const uint32_t ncfg = batch_config_count();
    const bool graph = b->use_graph;
    b->use_graph = false;
    float best = 0.0f;
    uint32_t best_cfg = 0;
    std::vector<xr_ubo> u(ubos, ubos + b->nviews);
    int rc = XR_OK;
    for (uint32_t k = 0; k < ncfg && rc == XR_OK; ++k) {
        b->cfg = k;
        float sum = 0.0f;
        for (uint32_t f = 0; f <= frames && rc == XR_OK; ++f) {  // frame 0 warms up
            for (uint32_t v = 0; v < b->nviews; ++v) u[v].current_frame = ubos[v].current_frame + f;
            rc = xr_batch_render(b, u.data(), params);
            float ms = 0.0f;
            if (rc == XR_OK) rc = xr_batch_frame_time(b, &ms);
            if (f > 0) sum += ms;
        }
        if (rc) break;
        const float avg = sum / (float)frames;
        if (ms_per_config) ms_per_config[k] = avg;
        if (k == 0 || avg < best) {
            best = avg;
            best_cfg = k;
        }
    }
    b->use_graph = graph;
    b->cfg = best_cfg;
    if (rc) return rc;
    rc = batch_clear(b);  // tuning frames leave no history behind
    if (rc) return rc;
    XR_HIP(c, hipMemsetAsync(b->d_counters, 0, sizeof(unsigned long long) * kCounters, c->stream));
    return XR_OK;
}

int xr_batch_reset(xr_batch* b) {
    if (!b) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "batch is NULL");
    return batch_clear(b);
}

int xr_batch_counters(xr_batch* b, uint64_t* out, uint32_t n) {
    if (!b || !out) return set_err(b ? b->ctx : nullptr, XR_ERR_INVALID_ARGUMENT, "NULL argument");
    unsigned long long h[kCounters];
    XR_HIP(b->ctx, hipMemcpyAsync(h, b->d_counters, sizeof(h), hipMemcpyDeviceToHost, b->ctx->stream));
    XR_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
    for (uint32_t i = 0; i < n && i < kCounters; ++i) out[i] = h[i];
    return XR_OK;
}

int xr_batch_reset_counters(xr_batch* b) {
    if (!b) return set_err(nullptr, XR_ERR_INVALID_ARGUMENT, "batch is NULL");
    XR_HIP(b->ctx, hipMemsetAsync(b->d_counters, 0, sizeof(unsigned long long) * kCounters, b->ctx->stream));
    return XR_OK;
}

size_t xr_batch_buffer_bytes(const xr_batch* b, uint32_t view, int which, int layout) {
    if (!b || view >= b->nviews) return 0;
    const size_t n = (size_t)b->views[view].w * b->views[view].h;
    switch (which) {
        case XR_BUF_OUTPUT:
        case XR_BUF_MOTION: return n * 4;
        case XR_BUF_GBUF_POS: return n * 16;
        case XR_BUF_SAMPLES: return layout == XR_LAYOUT_REFERENCE ? n * sizeof(ReSTIRGISample) : n * kPlanes * 16;
        case XR_BUF_TEMPORAL:
        case XR_BUF_SPATIAL:
            return layout == XR_LAYOUT_REFERENCE ? 2 * n * sizeof(ReSTIRGIReservoir) : 2 * n * kPlanes * 16;
        default: return 0;
    }
}

int xr_batch_read_buffer(xr_batch* b, uint32_t view, int which, void* dst, size_t bytes, int layout) {
    if (!b || !dst) return set_err(b ? b->ctx : nullptr, XR_ERR_INVALID_ARGUMENT, "NULL argument");
    xr_ctx* c = b->ctx;
    if (view >= b->nviews) return set_err(c, XR_ERR_INVALID_ARGUMENT, "view %u of %u", view, b->nviews);
    if (layout != XR_LAYOUT_REFERENCE && layout != XR_LAYOUT_NATIVE)
        return set_err(c, XR_ERR_INVALID_ARGUMENT, "unknown layout %d", layout);
    const size_t want = xr_batch_buffer_bytes(b, view, which, layout);
    if (want == 0) return set_err(c, XR_ERR_INVALID_ARGUMENT, "buffer %d is not kept per view", which);
    if (bytes != want) return set_err(c, XR_ERR_INVALID_ARGUMENT, "buffer %d needs %zu bytes, got %zu", which, want, bytes);
    const xr_batch::View& vw = b->views[view];
    const void* src = which == XR_BUF_OUTPUT     ? (const void*)vw.output
                      : which == XR_BUF_MOTION   ? (const void*)vw.motion
                      : which == XR_BUF_GBUF_POS ? (const void*)vw.gpos
                      : which == XR_BUF_SAMPLES  ? (const void*)vw.samples
                      : which == XR_BUF_TEMPORAL ? (const void*)vw.temporal
                                                 : (const void*)vw.spatial;
    return read_device_buffer(c, src, (size_t)vw.w * vw.h, which, dst, bytes, layout);
}

void xr_batch_destroy(xr_batch* b) {
    if (!b) return;
    if (b->ctx && b->ctx->stream) hipStreamSynchronize(b->ctx->stream);
    batch_free(b);
    delete b;
}

}  // extern "C"

// gfx950 kernels of the ReSTIR GI frame and of the plain path tracer.
//
//   k_initial_sample  ReSTIRGIInitialSample.rgen.hlsl:35-214
//   k_temporal        ReSTIRGITemporalReuse.rgen.hlsl:42-151
//   k_spatial         ReSTIRGISpatialReuse.rgen.hlsl:64-183
//   k_plain, k_plain_head, k_plain_bounce   rgen.hlsl:31-151 (+ rhit.hlsl, sampling.hlsl)
//   k_trace           TraceRay on a caller-supplied ray batch
//
// Image kernels run 256-lane workgroups on 32x8 pixel tiles; each wave64
// covers a 32x2 strip.  Tiles are dealt to the eight XCDs in contiguous
// bands (the hardware hands workgroup i to XCD i % 8), so each XCD's L2 holds
// one band of the image, its G-buffer/reservoir neighbourhoods and the BVH
// nodes its rays touch.
#include <algorithm>
#include <mutex>

#include "xr_internal.h"
#include "xr_kernels.h"

namespace xr {

// Workgroups of 256 lanes cover 32 x 8 pixel tiles: a row of a tile reads
// 128 contiguous bytes of each float4 plane (a whole cache line; 16 x 16 tiles
// read half lines), which took the spatial pass from 1.91 to 1.85 ms and the
// initial pass from 1.47 to 1.46 ms on the 1080p frame (64 x 4 was slower).
constexpr int kTileW = 32;
constexpr int kTileH = 8;
constexpr int kBlock = kTileW * kTileH;
constexpr int kXcds = 8;

struct TileCoord {
    uint32_t x, y;
    bool in_image;
};

// Logical tile of this workgroup.  Workgroups are dispatched round-robin over
// the 8 XCDs (blockIdx.x % 8), each with its own L2: XCD k gets the band of
// tile rows [k*rows, (k+1)*rows) and walks it column by column, so the tiles
// resident on one XCD at a time form a compact patch (the spatial pass
// gathers neighbours up to 30 px away from it).
// Tiles of kTw x kTh pixels (kTw * kTh = kBlock); b is the workgroup's index
// within the image's grid (a multiple of 8 workgroups, see image_grid).
template <int kTw, int kTh>
__device__ __forceinline__ TileCoord tile_coord(uint32_t b, uint32_t width, uint32_t height) {
    static_assert(kTw * kTh == kBlock, "one thread per pixel of the tile");
    const uint32_t tiles_x = (width + kTw - 1) / kTw;
    const uint32_t tiles_y = (height + kTh - 1) / kTh;
    const uint32_t rows = (tiles_y + kXcds - 1) / kXcds;
    const uint32_t j = b / kXcds;
    const uint32_t ty = (b % kXcds) * rows + j % rows;
    const uint32_t tx = j / rows;
    TileCoord tc;
    const uint32_t lx = threadIdx.x % kTw, ly = threadIdx.x / kTw;
    tc.x = tx * kTw + lx;
    tc.y = ty * kTh + ly;
    tc.in_image = tx < tiles_x && ty < tiles_y && tc.x < width && tc.y < height;
    return tc;
}
__device__ __forceinline__ TileCoord tile_pixel(uint32_t width, uint32_t height) {
    return tile_coord<kTileW, kTileH>(blockIdx.x, width, height);
}

uint32_t image_grid(uint32_t width, uint32_t height, uint32_t tw, uint32_t th) {
    const uint32_t tiles_x = (width + tw - 1) / tw;
    const uint32_t tiles_y = (height + th - 1) / th;
    return kXcds * ((tiles_y + kXcds - 1) / kXcds) * tiles_x;
}
uint32_t image_grid(uint32_t width, uint32_t height) { return image_grid(width, height, kTileW, kTileH); }

// Appends `value` of every lane whose pred is set to list (one global atomic
// per wave, slots in lane order).  Called by every lane of the wave.
__device__ __forceinline__ void wave_append(uint32_t* count, uint32_t* list, bool pred, uint32_t value) {
    const unsigned long long mask = __ballot(pred);
    if (mask == 0ull) return;
    const int lane = (int)__lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0u;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
    base = (uint32_t)__shfl((int)base, leader);
    if (pred) list[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = value;
}

// Slot in a device list for every thread of the workgroup that wants one: one
// LDS atomic per thread, one global atomic per workgroup.  Called by every
// thread of the workgroup.
__device__ __forceinline__ uint32_t block_slot(bool want, uint32_t* count, uint32_t* lds_count,
                                               uint32_t* lds_base) {
    if (threadIdx.x == 0) *lds_count = 0;
    __syncthreads();
    const uint32_t local = want ? atomicAdd(lds_count, 1u) : 0u;
    __syncthreads();
    if (threadIdx.x == 0) *lds_base = *lds_count ? atomicAdd(count, *lds_count) : 0u;
    __syncthreads();
    return *lds_base + local;
}

// One LDS atomic per lane, one global atomic per workgroup.
__device__ __forceinline__ void count_rays(unsigned long long* counter, uint32_t n, uint32_t* lds_count) {
    if (threadIdx.x == 0) *lds_count = 0;
    __syncthreads();
    if (n) atomicAdd(lds_count, n);
    __syncthreads();
    if (threadIdx.x == 0 && *lds_count) atomicAdd(counter, (unsigned long long)*lds_count);
}

__device__ __forceinline__ v3 xyz(float4 a) { return mk3(a.x, a.y, a.z); }
__device__ __forceinline__ float4 f4(v3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

// Camera ray for pixel (x, y) with a jitter drawn from the stream
// (InitialSample.rgen:43-52, rgen.hlsl:44-54).
__device__ __forceinline__ void camera_ray(const UniformBufferObject& ubo, uint32_t x, uint32_t y, uint32_t w,
                                           uint32_t h, uint32_t* rng, v3* org, v3* dir) {
    const float jx = rand01(rng);
    const float jy = rand01(rng);
    const float px = (float)x + jx, py = (float)y + jy;
    const float uvx = (px / (float)w) * 2.0f - 1.0f;
    const float uvy = (py / (float)h) * 2.0f - 1.0f;
    float o4[4], t4[4], d4[4];
    mat_vec(ubo.viewInverse, 0.0f, 0.0f, 0.0f, 1.0f, o4);
    mat_vec(ubo.projectionInverse, uvx, uvy, 1.0f, 1.0f, t4);
    const v3 tn = normalize(mk3(t4[0], t4[1], t4[2]));
    mat_vec(ubo.viewInverse, tn.x, tn.y, tn.z, 0.0f, d4);
    *org = mk3(o4[0], o4[1], o4[2]);
    *dir = normalize(mk3(d4[0], d4[1], d4[2]));
}

// ---------------------------------------------------------------------------
// Initial sample pass
// ---------------------------------------------------------------------------
// What the bounce loop of InitialSample.rgen:72-181 leaves for one pixel.
struct PathOut {
    v3 x_view, n_view, x_s, n_s, L, f0;
    float p_q, uv_u, uv_v;
    uint32_t mat_idx, rand_seed;
    uint32_t cache;   // 0: plain path, 1: ends in a cache query at x_s, 2: training path, 3 / 4: pending
    uint32_t xs_mat;  // material at x_s (cache modes)
    bool valid;       // primitiveHitValid
};

// Camera ray + bounce loop of the initial sample pass (InitialSample.rgen:41-181),
// split so that a kernel can run a range of bounces and hand the path on.
// With the radiance cache enabled (no reference counterpart) a path that
// reaches a diffuse first secondary vertex x_s is a candidate: exactly
// min(T, candidates) of the frame's candidates, spread evenly over them in
// pixel order, keep tracing to full depth as training samples; every other
// candidate stops at x_s when that vertex's spread t1^2 / (p_q |cos theta_s|)
// exceeds spread_c times the primary footprint t0^2 / (4 pi |cos theta_v|)
// (Mueller et al. 2021, "Real-time Neural Radiance Caching"), leaving L_out
// to the cache, and traces on as a plain path otherwise.  The head kernel
// runs a candidate through bounce 1 and parks it as pending (kind 3, or 4
// when it would stop into the cache); k_cache_select decides once every
// pixel's candidacy is known.
struct PathState {
    v3 origin, dir, thr;
    uint32_t rng;
    float a0;  // primary footprint (cache mode)
};

// Seed, camera ray and the empty PathOut of pixel (x, y) in frame `frame`.
__device__ __forceinline__ void path_begin(const UniformBufferObject& ubo, uint32_t x, uint32_t y, uint32_t w,
                                           uint32_t h, uint32_t frame, PathState* ps, PathOut* o) {
    ps->rng = init_seed(x, y, w, frame);
    camera_ray(ubo, x, y, w, h, &ps->rng, &ps->origin, &ps->dir);
    ps->thr = mk3(1.0f, 1.0f, 1.0f);
    ps->a0 = 0.0f;
    const v3 zero = mk3(0.f, 0.f, 0.f);
    o->x_view = zero; o->n_view = zero; o->x_s = zero; o->n_s = zero; o->L = zero; o->f0 = zero;
    o->p_q = 0.0f; o->uv_u = 0.0f; o->uv_v = 0.0f;
    o->mat_idx = 0xFFFFFFFFu; o->rand_seed = 0;
    o->cache = 0u; o->xs_mat = 0u;
    o->valid = false;
}

constexpr uint32_t kCachePending = 3u;       // candidate that traces on unless chosen for training
constexpr uint32_t kCachePendingQuery = 4u;  // candidate that stops into the cache unless chosen
constexpr uint32_t kRecordAlive = 0x100u;    // pending record: the path is alive after the head bounces

// Bounces b0 .. b1-1 of the loop.  Returns true when the path is still alive
// at bounce b1 (< bounces), false when it ended (miss, emitter, cache query,
// Russian roulette or the bounce limit).  *rays counts the rays traced.
template <int kLds = kLdsDepth>
__device__ __forceinline__ bool path_bounces(const SceneDev& sc, uint32_t x, uint32_t y, uint32_t w,
                                             uint32_t bounces, uint32_t b0, uint32_t b1, bool sky, PathState* ps,
                                             int* col, PathOut* o, const CacheCfg& cache, uint32_t* rays) {
    for (uint32_t bounce = b0;; ++bounce) {
        if (bounce == bounces) return false;
        if (bounce == b1) return true;
        const RayPre rp = make_ray_pre(ps->origin, ps->dir);
        Hit hit;
        ++*rays;
        // A hit whose material index is -1 reads as a miss (InitialSample.rgen:102-103).
        Payload pl;
        bool found = trace_ray<kBlock, false, kLds>(sc.nodes, sc.tris, rp, 0.0001f, 10000.0f, col, &hit);
        if (found) {
            closest_hit(sc, hit, &pl);
            found = pl.material != -1;
        }
        if (!found) {
            if (sky) o->L = add(o->L, ps->thr);
            if (bounce == 1) {
                o->x_s = add(ps->origin, muls(ps->dir, 1000.0f));
                o->n_s = neg(ps->dir);
            }
            return false;
        }
        const float4 m0 = sc.materials[2 * pl.material];
        const float4 m1 = sc.materials[2 * pl.material + 1];
        // Emitters shade with the untextured baseColor.rgb * 4 (InitialSample.rgen:117-120).
        if (__float_as_uint(m1.w) == kDiffuseLight) {
            o->L = add(o->L, muls(mul(ps->thr, mk3(m0.x, m0.y, m0.z)), 4.0f));
            if (bounce == 1) { o->x_s = pl.position; o->n_s = pl.normal; }
            if (bounce == 0) { o->x_view = pl.position; o->n_view = pl.normal; }
            return false;
        }
        if (bounce == 1) { o->x_s = pl.position; o->n_s = pl.normal; }
        if (cache.enabled) {
            if (bounce == 0) {
                ps->a0 = pl.t * pl.t / (4.0f * XR_PI * fabsf(dot(pl.normal, ps->dir)));
            } else if (bounce == 1) {
                o->xs_mat = (uint32_t)pl.material;
                const float a1 = pl.t * pl.t / (o->p_q * fabsf(dot(pl.normal, ps->dir)));
                o->cache = a1 > cache.spread_c * ps->a0 ? kCachePendingQuery : kCachePending;
            }
        }
        const uint32_t seed_here = ps->rng;
        v3 wi;
        float pdf, cos_t;
        const v3 f = sample_bsdf(sc, pl.normal, m0, m1, pl.u, pl.v, &wi, &pdf, &cos_t, &ps->rng);
        if (bounce == 0) {
            o->valid = true;
            o->x_view = pl.position;
            o->n_view = pl.normal;
            o->rand_seed = seed_here;
            o->mat_idx = (uint32_t)pl.material;
            o->uv_u = pl.u;
            o->uv_v = pl.v;
            o->p_q = pdf;
            o->f0 = f;
        } else {
            ps->thr = mul(ps->thr, divs(muls(f, cos_t), pdf));
        }
        if (bounce > 3) {
            const float p = maxf(ps->thr.x, maxf(ps->thr.y, ps->thr.z));
            if (rand01(&ps->rng) > p) return false;
            ps->thr = muls(ps->thr, 1.0f / p);
        }
        const v3 hp = add(ps->origin, muls(ps->dir, pl.t));
        const v3 g = pl.geo_normal;
        const v3 od = dot(wi, g) > 0.0f ? g : neg(g);
        ps->origin = add(hp, muls(od, 0.01f));
        ps->dir = wi;
    }
}

// What the loop leaves for a path that never found a diffuse primary hit.
__device__ __forceinline__ void path_end(PathOut* o) {
    if (!o->valid) {
        const v3 zero = mk3(0.f, 0.f, 0.f);
        o->L = zero;
        o->n_s = zero;
    }
}

// The whole loop for one pixel; returns the number of rays traced.
__device__ __forceinline__ uint32_t trace_path(const SceneDev& sc, const UniformBufferObject& ubo, uint32_t x,
                                               uint32_t y, uint32_t w, uint32_t h, uint32_t frame,
                                               uint32_t bounces, int* col, PathOut* o, const CacheCfg& cache) {
    PathState ps;
    path_begin(ubo, x, y, w, h, frame, &ps, o);
    uint32_t rays = 0;
    path_bounces(sc, x, y, w, bounces, 0u, bounces, ubo.hasSkyBox != 0u, &ps, col, o, cache, &rays);
    path_end(o);
    return rays;
}

// Initial-sample pass in two kernels.  k_initial_sample runs bounces 0 and 1
// of every pixel (all of them have a camera ray, most a first bounce) and
// writes everything the pass produces; a path still alive after bounce 1 is
// handed on as a PathRecord (compacted: one global atomic per workgroup) and
// k_tail_bounce runs its remaining bounces, one launch per bounce, and writes
// its L_out.  Beyond bounce 1 few paths survive on an outdoor scene (most
// bounce rays leave it), so compacting them keeps whole waves from tracing
// for one or two live lanes; the batched views keep one kernel for all the
// remaining bounces (k_batch_tail).
constexpr uint32_t kHeadBounces = 2;

// Path handed from the head to the tail kernel: 4 float4.
//   [0] origin, pixel   [1] direction, rng   [2] throughput, cache kind   [3] L_out so far, view
__device__ __forceinline__ void store_path(float4* rec, uint32_t pix, uint32_t view, const PathState& ps,
                                           const PathOut& o) {
    rec[0] = f4(ps.origin, __uint_as_float(pix));
    rec[1] = f4(ps.dir, __uint_as_float(ps.rng));
    rec[2] = f4(ps.thr, __uint_as_float(o.cache));
    rec[3] = f4(o.L, __uint_as_float(view));
}

// Head of the initial-sample pass for pixel (x, y): bounces 0 and 1 and every
// write of the pass except, for a path still alive, L_out (written by the
// tail).  Returns whether the path is alive.  The loop body is that of
// path_bounces for bounces 0 and 1, in one loop with one traversal; what
// bounce 0 settles (visible point, normal, seed, material, f, G-buffer,
// motion) is written before the bounce-1 ray is traced, so it is not live
// across that traversal (the head kernel's registers spilled to scratch:
// 288 bytes per lane, 620 MB of writes per 1080p frame against 282 MB
// compulsory).
template <int kLds>
__device__ __forceinline__ bool initial_head_pixel(const SceneDev& sc, const FrameDev& fr,
                                                   const UniformBufferObject& ubo, uint32_t x, uint32_t y,
                                                   const CacheCfg& cache, int* col, PathState* ps, PathOut* po,
                                                   uint32_t* rays) {
    PathOut& o = *po;
    const uint32_t pix = y * fr.width + x;
    const uint32_t n = fr.n;
    float4* S = fr.samples;
    path_begin(ubo, x, y, fr.width, fr.height, ubo.currentFrame, ps, &o);
    const uint32_t nb = ubo.numberOfBounces;
    const bool sky = ubo.hasSkyBox != 0u;
    float p_q = 0.0f;
    bool alive = false;
#pragma unroll
    for (uint32_t bounce = 0; bounce < kHeadBounces; ++bounce) {
        if (bounce >= nb) break;
        alive = false;
        const RayPre rp = make_ray_pre(ps->origin, ps->dir);
        Hit hit;
        ++*rays;
        // A hit whose material index is -1 reads as a miss (InitialSample.rgen:102-103).
        Payload pl;
        bool found = trace_ray<kBlock, false, kLds>(sc.nodes, sc.tris, rp, 0.0001f, 10000.0f, col, &hit);
        if (found) {
            closest_hit(sc, hit, &pl);
            found = pl.material != -1;
        }
        bool go_on = false;
        if (!found) {
            if (sky) o.L = add(o.L, ps->thr);
            if (bounce == 1) {
                o.x_s = add(ps->origin, muls(ps->dir, 1000.0f));
                o.n_s = neg(ps->dir);
            }
        } else {
            const float4 m0 = sc.materials[2 * pl.material];
            const float4 m1 = sc.materials[2 * pl.material + 1];
            if (__float_as_uint(m1.w) == kDiffuseLight) {
                // untextured baseColor.rgb * 4 (InitialSample.rgen:117-120)
                o.L = add(o.L, muls(mul(ps->thr, mk3(m0.x, m0.y, m0.z)), 4.0f));
                if (bounce == 1) { o.x_s = pl.position; o.n_s = pl.normal; }
                if (bounce == 0) { o.x_view = pl.position; o.n_view = pl.normal; }
            } else {
                if (bounce == 1) { o.x_s = pl.position; o.n_s = pl.normal; }
                if (cache.enabled) {
                    if (bounce == 0) {
                        ps->a0 = pl.t * pl.t / (4.0f * XR_PI * fabsf(dot(pl.normal, ps->dir)));
                    } else {
                        o.xs_mat = (uint32_t)pl.material;
                        const float a1 = pl.t * pl.t / (p_q * fabsf(dot(pl.normal, ps->dir)));
                        o.cache = a1 > cache.spread_c * ps->a0 ? kCachePendingQuery : kCachePending;
                    }
                }
                const uint32_t seed_here = ps->rng;
                v3 wi;
                float pdf, cos_t;
                const v3 f = sample_bsdf(sc, pl.normal, m0, m1, pl.u, pl.v, &wi, &pdf, &cos_t, &ps->rng);
                if (bounce == 0) {
                    o.valid = true;
                    p_q = pdf;
                    // what bounce 0 settles, written now
                    S[0 * n + pix] = f4(pl.position, __uint_as_float((uint32_t)pl.material));
                    S[1 * n + pix] = f4(pl.normal, __uint_as_float(seed_here));
                    S[5 * n + pix] = f4(f, 0.0f);
                    fr.gpos[pix] = f4(pl.position, 1.0f);
                    fr.gnrm[pix] = f4(pl.normal, 0.0f);
                    fr.gmatuv[pix] = make_float4((float)pl.material, pl.u, pl.v, 0.0f);
                    const float cu = ((float)x + 0.5f) / (float)fr.width;
                    const float cv = ((float)y + 0.5f) / (float)fr.height;
                    float vp[4], cp[4];
                    mat_vec(ubo.viewPrev, pl.position.x, pl.position.y, pl.position.z, 1.0f, vp);
                    mat_vec(ubo.projectionPrev, vp[0], vp[1], vp[2], vp[3], cp);
                    const float pu = (cp[0] / cp[3]) * 0.5f + 0.5f;
                    const float pv = (cp[1] / cp[3]) * 0.5f + 0.5f;
                    const uint32_t mx = f32_to_f16_bits(cu - pu), my = f32_to_f16_bits(cv - pv);
                    fr.motion[pix] = mx | (my << 16);
                } else {
                    ps->thr = mul(ps->thr, divs(muls(f, cos_t), pdf));
                }
                const v3 hp = add(ps->origin, muls(ps->dir, pl.t));
                const v3 g = pl.geo_normal;
                const v3 od = dot(wi, g) > 0.0f ? g : neg(g);
                ps->origin = add(hp, muls(od, 0.01f));
                ps->dir = wi;
                go_on = true;
            }
        }
        if (!go_on) break;
        alive = bounce + 1u < nb;  // a path past bounce 1 goes on when the loop has bounces left
    }
    if (!alive) path_end(&o);
    // Spare words (zero without the cache): plane 3 .w = material at x_s,
    // plane 4 .w = 1 cache query / 2 training sample.
    if (!o.valid) {
        S[0 * n + pix] = f4(o.x_view, __uint_as_float(0xFFFFFFFFu));
        S[1 * n + pix] = f4(o.n_view, 0.0f);
        S[5 * n + pix] = make_float4(0.f, 0.f, 0.f, 0.f);
        fr.gpos[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
        fr.gnrm[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
        fr.gmatuv[pix] = make_float4(-1.f, 0.f, 0.f, 0.f);
        fr.motion[pix] = 0u;
    }
    S[2 * n + pix] = f4(o.x_s, p_q);
    S[3 * n + pix] = f4(o.n_s, __uint_as_float(o.cache ? o.xs_mat : 0u));
    if (!alive && o.cache < kCachePending) S[4 * n + pix] = f4(o.L, __uint_as_float(o.cache));
    return alive;
}

// Hands the workgroup's live paths to the tail list: one global atomic per
// workgroup.  Called by every thread of the workgroup.
__device__ __forceinline__ void hand_on(bool alive, float4* tail, uint32_t* tail_count, uint32_t pix, uint32_t view,
                                        const PathState& ps, const PathOut& o, uint32_t* lds_count,
                                        uint32_t* lds_base) {
    const uint32_t q = block_slot(alive, tail_count, lds_count, lds_base);
    if (alive) store_path(tail + 4 * (size_t)q, pix, view, ps, o);
}

// The head kernel keeps 20 stack levels in LDS (20 KiB per workgroup) and
// asks for 7 waves per SIMD (72 VGPRs, 224 bytes of scratch per lane): the
// initial pass of the 1080p terrain frame took 1.003 ms against 1.035 at 24
// levels and 6 waves (18 levels at 8 waves: 1.005, 22 at 7: 1.003).
constexpr int kHeadLds = 20;
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_initial_sample(
    SceneDev sc, FrameDev fr, UniformBufferObject ubo, CacheCfg cache) {
    __shared__ int stack[kHeadLds * kBlock];
    __shared__ uint32_t lds_count;
    __shared__ uint32_t lds_base;
    const TileCoord tc = tile_pixel(fr.width, fr.height);
    uint32_t rays = 0, valid = 0, kind = 0;
    bool alive = false;
    PathState ps;
    PathOut o;
    if (tc.in_image) {
        alive = initial_head_pixel<kHeadLds>(sc, fr, ubo, tc.x, tc.y, cache, stack + threadIdx.x, &ps, &o, &rays);
        kind = o.cache;
        valid = o.valid ? 1u : 0u;
    }
    const uint32_t pix = tc.y * fr.width + tc.x;
    const bool pending = kind >= kCachePending;
    hand_on(alive && !pending, fr.tail, fr.tail_count, pix, 0u, ps, o, &lds_count, &lds_base);
    if (cache.enabled) {  // uniform over the launch
        // candidates park their record (alive or not) in the second queue
        const uint32_t q = block_slot(pending, fr.pending_count, &lds_count, &lds_base);
        if (pending) {
            o.cache = kind | (alive ? kRecordAlive : 0u);
            store_path(fr.pending + 4 * (size_t)q, pix, 0u, ps, o);
        }
        if (tc.in_image) fr.cache_slot[pix] = pending ? q + 1u : 0u;
    }
    count_rays(&fr.counters[0], rays, &lds_count);
    count_rays(&fr.counters[2], valid, &lds_count);
}

// ---------------------------------------------------------------------------
// Training choice of the radiance cache: candidates ranked in pixel order by
// an exclusive scan of per-pixel flags over 4096-pixel blocks; the chosen ones
// and the plain candidates join the tail queue, the others end as queries.
// ---------------------------------------------------------------------------
// Candidates among the pixels of this workgroup's block (pixel
// block*kCacheSelBlock + r*256 + thread in round r: rounds read coalesced rows).
__global__ __launch_bounds__(256) void k_cache_count(FrameDev fr) {
    __shared__ uint32_t part[4];
    const uint32_t base = blockIdx.x * kCacheSelBlock;
    uint32_t c = 0;
    for (uint32_t r = 0; r < kCacheSelBlock / 256; ++r) {
        const uint32_t p = base + r * 256 + threadIdx.x;
        if (p < fr.n && fr.cache_slot[p] != 0u) ++c;
    }
    for (int off = 32; off > 0; off >>= 1) c += (uint32_t)__shfl_down((int)c, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) fr.cache_blocks[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// Exclusive scan of the nblocks block counts in place (one workgroup of
// 1024 threads, 1024 counts per step); the total goes to entry nblocks.
__global__ __launch_bounds__(1024) void k_cache_scan(FrameDev fr, uint32_t nblocks) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0u;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (uint32_t base = 0; base < nblocks; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nblocks ? fr.cache_blocks[i] : 0u;
        uint32_t x = v;  // inclusive scan within the wave
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) wsum[wv] = x;
        __syncthreads();
        uint32_t before = carry;
        for (uint32_t k = 0; k < wv; ++k) before += wsum[k];
        if (i < nblocks) fr.cache_blocks[i] = before + x - v;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (int k = 0; k < 16; ++k) t += wsum[k];
            carry += t;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) fr.cache_blocks[nblocks] = carry;
}

// Decides every pending candidate: rank r in pixel order, r' = (r + o) mod C
// with o = (frame * 7919) mod C, training when C <= T or
// floor((r' + 1) T / C) > floor(r' T / C) (exactly min(T, C) of them).
// Training and plain candidates still alive join the tail queue (tail_count[0]);
// queries and finished paths write their L_out and kind.  Appends the
// frame's query / training pixel lists and counts them.  A thread of the
// 1024-thread workgroup holds its 4 pixels (block*4096 + r*1024 + thread) in
// registers: their slots and record words are loaded before any ranking, one
// wave scans the (round, wave) candidate counts for the ranks, list slots
// within the workgroup come from one LDS atomic per list and wave-round, and
// the workgroup takes one global slot range per list (a wave-level global
// append hammered two counters 65 k times per frame: 0.58 ms; 1024-pixel
// workgroups made four times as many same-address range atomics and took
// 79 against 49 us).
// frame_ubo: when not null, the frame number comes from that device UBO (a
// batched view's, which a replayed graph cannot take as an argument).
constexpr uint32_t kSelThreads = 1024;
__global__ __launch_bounds__(kSelThreads) void k_cache_select(FrameDev fr, CacheCfg cache,
                                                              const UniformBufferObject* frame_ubo) {
    constexpr uint32_t kRounds = kCacheSelBlock / kSelThreads, kWaves = kSelThreads / 64;
    static_assert(kRounds * kWaves <= 64, "one wave scans the (round, wave) counts");
    __shared__ uint32_t wpre[kRounds * kWaves];  // candidates of (round, wave), then their exclusive prefix
    __shared__ uint32_t cnt[3];             // go on, query list, training list
    __shared__ uint32_t base[3];
    const uint32_t nblocks = gridDim.x;
    const uint32_t C = fr.cache_blocks[nblocks];
    const uint64_t T = cache.train_target;
    const uint32_t frame = frame_ubo ? frame_ubo->currentFrame : cache.frame;
    const uint32_t o = C ? (frame * 7919u) % C : 0u;
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    if (threadIdx.x < 3) cnt[threadIdx.x] = 0u;
    const float4* pend = fr.pending;
    const uint32_t p0 = blockIdx.x * kCacheSelBlock + threadIdx.x;
    uint32_t slot[kRounds], word[kRounds], lower[kRounds];
#pragma unroll
    for (uint32_t r = 0; r < kRounds; ++r) {
        const uint32_t p = p0 + r * kSelThreads;
        slot[r] = p < fr.n ? fr.cache_slot[p] : 0u;
    }
#pragma unroll
    for (uint32_t r = 0; r < kRounds; ++r) {
        word[r] = slot[r] ? __float_as_uint(pend[4 * (size_t)(slot[r] - 1u) + 2].w) : 0u;
        const unsigned long long m = __ballot(slot[r] != 0u);
        lower[r] = (uint32_t)__popcll(m & below);
        if (lane == 0) wpre[r * kWaves + wv] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (wv == 0) {
        const uint32_t v = lane < kRounds * kWaves ? wpre[lane] : 0u;
        uint32_t x = v;  // inclusive scan within the wave (every lane shuffles)
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)x, off);
            if (lane >= off) x += y;
        }
        if (lane < kRounds * kWaves) wpre[lane] = x - v;
    }
    __syncthreads();
    const uint32_t running = fr.cache_blocks[blockIdx.x];
    // per pixel: bit 0 go on, bits 1-2 kind (1 query, 2 training), bits 3-15
    // slot in the workgroup's tail range, bits 16-28 slot in its list range;
    // all ones: not a candidate
    uint32_t dec[kRounds];
#pragma unroll
    for (uint32_t r = 0; r < kRounds; ++r) {
        const bool cand = slot[r] != 0u;
        uint32_t kind = 0u;
        bool go_on = false;
        if (cand) {
            uint32_t rr = running + wpre[r * kWaves + wv] + lower[r] + o;
            if (rr >= C) rr -= C;
            const bool train = (uint64_t)C <= T || ((uint64_t)rr + 1u) * T / C > (uint64_t)rr * T / C;
            kind = train ? 2u : ((word[r] & 0xFFu) == kCachePendingQuery ? 1u : 0u);
            go_on = kind != 1u && (word[r] & kRecordAlive) != 0u;
        }
        const unsigned long long mg = __ballot(cand && go_on);
        const unsigned long long mq = __ballot(cand && kind == 1u);
        const unsigned long long mt = __ballot(cand && kind == 2u);
        uint32_t bg = 0u, bq = 0u, bt = 0u;
        if (lane == 0) {
            if (mg) bg = atomicAdd(&cnt[0], (uint32_t)__popcll(mg));
            if (mq) bq = atomicAdd(&cnt[1], (uint32_t)__popcll(mq));
            if (mt) bt = atomicAdd(&cnt[2], (uint32_t)__popcll(mt));
        }
        bg = (uint32_t)__shfl((int)bg, 0);
        bq = (uint32_t)__shfl((int)bq, 0);
        bt = (uint32_t)__shfl((int)bt, 0);
        const uint32_t ig = bg + (uint32_t)__popcll(mg & below);
        const uint32_t il = kind == 2u ? bt + (uint32_t)__popcll(mt & below) : bq + (uint32_t)__popcll(mq & below);
        dec[r] = cand ? ((go_on ? 1u : 0u) | (kind << 1) | (ig << 3) | (il << 16)) : 0xFFFFFFFFu;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        uint32_t* ctr = threadIdx.x == 0 ? fr.tail_count : fr.cache_list + (threadIdx.x - 1u);
        base[threadIdx.x] = cnt[threadIdx.x] ? atomicAdd(ctr, cnt[threadIdx.x]) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < kRounds; ++r) {
        const uint32_t d = dec[r];
        if (d == 0xFFFFFFFFu) continue;
        const uint32_t p = p0 + r * kSelThreads;
        const uint32_t kind = (d >> 1) & 3u;
        const float4* src = pend + 4 * (size_t)(slot[r] - 1u);
        if (d & 1u) {
            float4* dst = fr.tail + 4 * (size_t)(base[0] + ((d >> 3) & 0x1FFFu));
            dst[0] = src[0];
            dst[1] = src[1];
            const float4 r2 = src[2];
            dst[2] = make_float4(r2.x, r2.y, r2.z, __uint_as_float(kind));
            dst[3] = src[3];
        } else {
            const float4 r3 = src[3];
            fr.samples[4 * (size_t)fr.n + p] = make_float4(r3.x, r3.y, r3.z, __uint_as_float(kind));
        }
        // the material at x_s is kept for cache kinds only (spare word of plane 3)
        if (kind == 0u) reinterpret_cast<uint32_t*>(fr.samples + 3 * (size_t)fr.n + p)[3] = 0u;
        else fr.cache_list[2 + (kind == 2u ? fr.n : 0u) + base[kind] + (d >> 16)] = p;
    }
    __shared__ uint32_t lds_count;
    count_rays(&fr.counters[3], threadIdx.x == 0 ? cnt[1] : 0u, &lds_count);
    count_rays(&fr.counters[4], threadIdx.x == 0 ? cnt[2] : 0u, &lds_count);
}

// Remaining bounces of one handed-on path; writes its L_out.
template <int kLds>
__device__ __forceinline__ void initial_tail_path(const SceneDev& sc, const FrameDev& fr,
                                                  const UniformBufferObject& ubo, const float4* rec,
                                                  const CacheCfg& cache, int* col, uint32_t* rays) {
    const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3];
    const uint32_t pix = __float_as_uint(r0.w);
    PathState ps;
    ps.origin = xyz(r0);
    ps.dir = xyz(r1);
    ps.rng = __float_as_uint(r1.w);
    ps.thr = xyz(r2);
    ps.a0 = 0.0f;
    PathOut o;
    o.L = xyz(r3);
    o.cache = __float_as_uint(r2.w);
    o.valid = true;  // only paths with a diffuse primary hit get past bounce 0
    path_bounces<kLds>(sc, pix % fr.width, pix / fr.width, fr.width, ubo.numberOfBounces, kHeadBounces,
                       ubo.numberOfBounces, ubo.hasSkyBox != 0u, &ps, col, &o, cache, rays);
    fr.samples[4 * (size_t)fr.n + pix] = f4(o.L, __uint_as_float(o.cache));
}

// Remaining bounces of the handed-on paths, from `bounce` (>= kHeadBounces)
// on: queue (bounce - kHeadBounces) % 2 in, the other queue out.  Where many
// paths are alive (over a quarter of the pixels, as in a closed interior) the
// launch runs one bounce and hands the survivors on: the next launch takes
// the next bounce, so the waves of a launch stay at the same bounce instead
// of each waiting for its longest path at every bounce (the reference's sb2
// scene: initial pass 5.57 -> 4.40 ms).  Fewer paths run to their end in this
// launch (the later launches find an empty queue): on an outdoor scene the
// extra launches and path records cost more than they save.  A path that
// ends writes its L_out.  The grid covers the queue's capacity; workgroups
// past the count return.
__global__ __launch_bounds__(kBlock) XR_TRAVERSAL_OCCUPANCY void k_tail_bounce(SceneDev sc, FrameDev fr,
                                                                                UniformBufferObject ubo,
                                                                                CacheCfg cache, uint32_t bounce) {
    __shared__ int stack[kLdsDepth * kBlock];
    __shared__ uint32_t lds_count;
    __shared__ uint32_t lds_base;
    const uint32_t j = bounce - kHeadBounces, side = j & 1u;
    const float4* in = fr.tail + (size_t)side * 4 * fr.n;
    float4* out = fr.tail + (size_t)(side ^ 1u) * 4 * fr.n;
    const uint32_t count = fr.tail_count[j % 3u];
    if (blockIdx.x == 0u && threadIdx.x == 0u) fr.tail_count[(j + 2u) % 3u] = 0u;  // the next launch's output
    if (blockIdx.x * kBlock >= count) return;  // uniform over the workgroup
    const bool one = (uint64_t)count * 4u > (uint64_t)fr.n;  // uniform over the launch
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    uint32_t rays = 0;
    bool alive = false;
    PathState ps;
    PathOut o;
    uint32_t pix = 0;
    if (q < count) {
        const float4* rec = in + 4 * (size_t)q;
        const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3];
        pix = __float_as_uint(r0.w);
        ps.origin = xyz(r0);
        ps.dir = xyz(r1);
        ps.rng = __float_as_uint(r1.w);
        ps.thr = xyz(r2);
        ps.a0 = 0.0f;
        o.L = xyz(r3);
        o.cache = __float_as_uint(r2.w);
        o.valid = true;  // only paths with a diffuse primary hit get past bounce 0
        alive = path_bounces<kLdsDepth>(sc, pix % fr.width, pix / fr.width, fr.width, ubo.numberOfBounces, bounce,
                                        one ? bounce + 1u : ubo.numberOfBounces, ubo.hasSkyBox != 0u, &ps,
                                        stack + threadIdx.x, &o, cache, &rays);
        if (!alive) fr.samples[4 * (size_t)fr.n + pix] = f4(o.L, __uint_as_float(o.cache));
    }
    if (one) hand_on(alive, out, fr.tail_count + (j + 1u) % 3u, pix, 0u, ps, o, &lds_count, &lds_base);
    count_rays(&fr.counters[0], rays, &lds_count);
}

__global__ void k_clear_u32(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

hipError_t launch_clear_u32(uint32_t* p, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Converged image (FP64 progressive accumulation; no reference counterpart)
// ---------------------------------------------------------------------------
// One sample of the quantity the ReSTIR output estimates: the initial
// sample's own RIS estimate with M = 1, W = 1/p_q, i.e.
// L_out * (1/p_q) * f * max(0, n_v . normalize(x_s - x_v)) for pixels with a
// valid primary hit and 0 elsewhere (the black / white early-out pixels of
// the spatial pass).  Sample s of pixel p uses frame number frame_base + s.
// Each pixel adds its spp samples into three FP64 sums.
__global__ __launch_bounds__(kBlock) XR_TRAVERSAL_OCCUPANCY void k_converge(SceneDev sc, FrameDev fr, UniformBufferObject ubo,
                                                     uint32_t frame_base, uint32_t spp, double* __restrict__ acc) {
    __shared__ int stack[kLdsDepth * kBlock];
    __shared__ uint32_t lds_count;
    const TileCoord tc = tile_pixel(fr.width, fr.height);
    uint32_t rays = 0;
    if (tc.in_image) {
        const uint32_t pix = tc.y * fr.width + tc.x;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (uint32_t s = 0; s < spp; ++s) {
            PathOut o;
            const CacheCfg plain = {0u, 1u, 0u, 0.0f};
            rays += trace_path(sc, ubo, tc.x, tc.y, fr.width, fr.height, frame_base + s, ubo.numberOfBounces,
                               stack + threadIdx.x, &o, plain);
            if (o.valid && o.p_q > 1e-6f) {
                const v3 wi = normalize(sub(o.x_s, o.x_view));
                const float c = maxf(0.0f, dot(o.n_view, wi));
                const float W = 1.0f / o.p_q;
                sx += (double)(o.L.x * W * o.f0.x * c);
                sy += (double)(o.L.y * W * o.f0.y * c);
                sz += (double)(o.L.z * W * o.f0.z * c);
            }
        }
        acc[3 * (size_t)pix + 0] += sx;
        acc[3 * (size_t)pix + 1] += sy;
        acc[3 * (size_t)pix + 2] += sz;
    }
    count_rays(&fr.counters[0], rays, &lds_count);
}

// Sum over pixels and channels of (img - acc/spp)^2 in FP64; one double
// atomic per workgroup.
__global__ __launch_bounds__(256) void k_sq_error(const float4* __restrict__ img, const double* __restrict__ acc,
                                                  double inv_spp, uint32_t n, double* __restrict__ out) {
    __shared__ double part[256];
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    double e = 0.0;
    if (i < n) {
        const float4 v = img[i];
        const double dx = (double)v.x - acc[3 * (size_t)i] * inv_spp;
        const double dy = (double)v.y - acc[3 * (size_t)i + 1] * inv_spp;
        const double dz = (double)v.z - acc[3 * (size_t)i + 2] * inv_spp;
        e = dx * dx + dy * dy + dz * dz;
    }
    part[threadIdx.x] = e;
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, part[0]);
}

// ---------------------------------------------------------------------------
// Reservoir helpers (ReSTIRGICommon.hlsl:32-78)
// ---------------------------------------------------------------------------
struct Res {
    float4 p[kPlanes];  // planes as in xr_layout.h (w of 3,4,5 = w_sum, W, M)
};

__device__ __forceinline__ void res_init(Res* r) {
    r->p[0] = make_float4(0.f, 0.f, 0.f, __uint_as_float(0xFFFFFFFFu));
    r->p[1] = make_float4(0.f, 0.f, 0.f, 0.f);
    r->p[2] = make_float4(0.f, 0.f, 0.f, 0.f);
    r->p[3] = make_float4(0.f, 0.f, 0.f, 0.f);
    r->p[4] = make_float4(0.f, 0.f, 0.f, 0.f);
    r->p[5] = make_float4(0.f, 0.f, 0.f, 0.f);
}
__device__ __forceinline__ float r_wsum(const Res& r) { return r.p[3].w; }
__device__ __forceinline__ float r_W(const Res& r) { return r.p[4].w; }
__device__ __forceinline__ uint32_t r_m(const Res& r) { return __float_as_uint(r.p[5].w); }
__device__ __forceinline__ void set_wsum(Res* r, float v) { r->p[3].w = v; }
__device__ __forceinline__ void set_W(Res* r, float v) { r->p[4].w = v; }
__device__ __forceinline__ void set_m(Res* r, uint32_t v) { r->p[5].w = __uint_as_float(v); }
// r.s = src (sample planes 0-2 whole, xyz of 3-5)
__device__ __forceinline__ void take_sample(Res* r, const float4* s) {
    r->p[0] = s[0];
    r->p[1] = s[1];
    r->p[2] = s[2];
    r->p[3] = make_float4(s[3].x, s[3].y, s[3].z, r->p[3].w);
    r->p[4] = make_float4(s[4].x, s[4].y, s[4].z, r->p[4].w);
    r->p[5] = make_float4(s[5].x, s[5].y, s[5].z, r->p[5].w);
}
// UpdateReservoir: w_sum += w; m++; one draw; take the sample if xi < w/max(1e-10, w_sum).
__device__ __forceinline__ void res_update(Res* r, const float4* s, float w, uint32_t* seed) {
    const float ws = r_wsum(*r) + w;
    set_wsum(r, ws);
    set_m(r, r_m(*r) + 1u);
    if (rand01(seed) < w / maxf(1e-10f, ws)) take_sample(r, s);
}
// MergeReservoir: w = p_hat * src.W * src.m; update; m = M0 + src.m.
__device__ __forceinline__ void res_merge(Res* r, const Res& src, float p_hat, uint32_t* seed) {
    const uint32_t m0 = r_m(*r);
    const float w = p_hat * r_W(src) * (float)r_m(src);
    res_update(r, src.p, w, seed);
    set_m(r, m0 + r_m(src));
}
__device__ __forceinline__ void load_planes(const float4* base, uint32_t n, uint32_t i, float4* out) {
#pragma unroll
    for (int k = 0; k < kPlanes; ++k) out[k] = base[(size_t)k * n + i];
}
__device__ __forceinline__ void store_planes(float4* base, uint32_t n, uint32_t i, const float4* in) {
#pragma unroll
    for (int k = 0; k < kPlanes; ++k) base[(size_t)k * n + i] = in[k];
}

// EvaluateTargetFunction (Temporal.rgen:33-40): |L| if the sample lies in
// front of the current surface, else 0.
__device__ __forceinline__ float target_fn(const float4* s, v3 x_curr, v3 n_curr) {
    const v3 wi = normalize(sub(xyz(s[2]), x_curr));
    const float c = dot(n_curr, wi);
    if (c <= 0.0f) return 0.0f;
    return length(xyz(s[4]));
}

// Final shading of a resampled reservoir: L * W * f * max(0, n.wi).
__device__ __forceinline__ v3 shade_reservoir(const Res& r, v3 x_curr, v3 n_curr) {
    const v3 wi = normalize(sub(xyz(r.p[2]), x_curr));
    const float c = maxf(0.0f, dot(n_curr, wi));
    const float W = r_W(r);
    const v3 Lo = xyz(r.p[4]), f = xyz(r.p[5]);
    return mk3(Lo.x * W * f.x * c, Lo.y * W * f.y * c, Lo.z * W * f.z * c);
}

// The spatial pass's neighbour record of pixel pix: its sample's planes 0 and
// 1 (x_view + material, n_view) and, of its temporal reservoir r, |L_out|
// (the p_hat of Spatial.rgen:131, computed as there), W, x_sample, n_sample, M.
__device__ __forceinline__ void store_nbr(float4* __restrict__ nbr, uint32_t pix, float4 s0, float4 s1, const Res& r) {
    float4* q = nbr + 4 * (size_t)pix;
    q[0] = s0;
    q[1] = make_float4(s1.x, s1.y, s1.z, length(xyz(r.p[4])));
    q[2] = make_float4(r.p[2].x, r.p[2].y, r.p[2].z, r_W(r));
    q[3] = make_float4(r.p[3].x, r.p[3].y, r.p[3].z, r.p[5].w);
}

// Neighbour records from the samples and the temporal reservoirs of slot
// `ping` (the spatial pass run without the temporal pass before it).
__global__ __launch_bounds__(256) void k_nbr_build(FrameDev fr, uint32_t ping) {
    const uint32_t pix = blockIdx.x * 256u + threadIdx.x;
    if (pix >= fr.n) return;
    const uint32_t n = fr.n;
    const float4* rd = fr.temporal + (size_t)ping * kPlanes * n;
    Res r;
    load_planes(rd, n, pix, r.p);
    store_nbr(fr.nbr, pix, fr.samples[pix], fr.samples[n + pix], r);
}

// ---------------------------------------------------------------------------
// Temporal reuse pass (HBM bound; no rays)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void spatial_after_temporal(const FrameDev& fr, uint32_t frame, uint32_t reuse,
                                                       const PassParams& pp, uint32_t x, uint32_t y,
                                                       const float4* s, const Res& r, v3 t_color);

// kSpatial0: the spatial pass without neighbours (spatial_samples = 0) runs
// here on the values just computed, instead of a second kernel reading them
// back (its result depends on this pixel's temporal reservoir alone); no
// neighbour records are written then.
template <bool kSpatial0 = false>
__device__ __forceinline__ void temporal_pixel(const FrameDev& fr, uint32_t frame, const PassParams& pp,
                                               const float4* __restrict__ gpos_hist, uint32_t pix,
                                               uint32_t reuse = 0u) {
    const uint32_t n = fr.n, w = fr.width, h = fr.height;
    const uint32_t x = pix % w, y = pix / w;
    uint32_t rng = init_seed(x, y, w, frame);
    const uint32_t ping = frame % 2u;
    float4* cur = fr.temporal + (size_t)ping * kPlanes * n;
    const float4* prev = fr.temporal + (size_t)(1u - ping) * kPlanes * n;
    float4* spat_cur = fr.spatial + (size_t)ping * kPlanes * n;

    float4 s[kPlanes];
    load_planes(fr.samples, n, pix, s);
    if (length(xyz(s[3])) == 0.0f) {
        Res r0;
        res_init(&r0);
        store_planes(cur, n, pix, r0.p);
        store_planes(spat_cur, n, pix, r0.p);
        if (fr.radiance) fr.radiance[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (kSpatial0) spatial_after_temporal(fr, frame, reuse, pp, x, y, s, r0, mk3(0.f, 0.f, 0.f));
        else store_nbr(fr.nbr, pix, s[0], s[1], r0);
        return;
    }
    bool found = false;
    Res rp;
    const uint32_t mbits = fr.motion[pix];
    const float mx = f16_bits_to_f32((uint16_t)(mbits & 0xFFFFu));
    const float my = f16_bits_to_f32((uint16_t)(mbits >> 16));
    const float u = ((float)x + 0.5f) / (float)w;
    const float v = ((float)y + 0.5f) / (float)h;
    const float pu = u - mx, pv = v - my;
    const int px = (int)(pu * (float)w);
    const int py = (int)(pv * (float)h);
    if (px >= 0 && px < (int)w && py >= 0 && py < (int)h) {
        const uint32_t pidx = (uint32_t)py * w + (uint32_t)px;
        if (gpos_hist[pidx].w != 0.0f) {
            load_planes(prev, n, pidx, rp.p);
            found = true;
        } else {
            Res r0;
            res_init(&r0);
            store_planes(spat_cur, n, pix, r0.p);
        }
    }
    const v3 x_curr = xyz(s[0]);
    const v3 n_curr = normalize(xyz(s[1]));
    Res r;
    res_init(&r);
    const float tp_new = target_fn(s, x_curr, n_curr);
    const float w_new = s[2].w > 1e-6f ? tp_new / s[2].w : 0.0f;
    res_update(&r, s, w_new, &rng);
    if (found) {
        const float tp_prev = target_fn(rp.p, x_curr, n_curr);
        res_merge(&r, rp, tp_prev, &rng);
    } else {
        set_m(&r, 1u);
    }
    const uint32_t cap = pp.temporal_m_cap;
    if (r_m(r) > cap) {
        set_wsum(&r, r_wsum(r) * ((float)cap / (float)r_m(r)));
        set_m(&r, cap);
    }
    const float tp_final = target_fn(r.p, x_curr, n_curr);
    if (tp_final > 1e-6f && r_m(r) > 0u) set_W(&r, r_wsum(r) / ((float)r_m(r) * tp_final));
    else set_W(&r, 0.0f);
    store_planes(cur, n, pix, r.p);
    if (!kSpatial0) store_nbr(fr.nbr, pix, s[0], s[1], r);
    v3 color = mk3(0.f, 0.f, 0.f);
    if (r_W(r) > 0.0f) color = shade_reservoir(r, x_curr, xyz(s[1]));
    if (kSpatial0) {
        spatial_after_temporal(fr, frame, reuse, pp, x, y, s, r, color);
        return;
    }
    fr.output[pix] = resolve_bgra8(color, 1.0f);
    if (fr.radiance) fr.radiance[pix] = f4(color, 0.0f);
}

__global__ __launch_bounds__(256) void k_temporal(FrameDev fr, uint32_t frame, PassParams pp,
                                                  const float4* __restrict__ gpos_hist) {
    const uint32_t pix = blockIdx.x * 256u + threadIdx.x;
    if (pix < fr.n) temporal_pixel(fr, frame, pp, gpos_hist, pix);
}

__global__ __launch_bounds__(256) void k_temporal_spatial0(FrameDev fr, uint32_t frame, PassParams pp,
                                                           const float4* __restrict__ gpos_hist, uint32_t reuse) {
    const uint32_t pix = blockIdx.x * 256u + threadIdx.x;
    if (pix < fr.n) temporal_pixel<true>(fr, frame, pp, gpos_hist, pix, reuse);
}

// ---------------------------------------------------------------------------
// Spatial reuse pass
// ---------------------------------------------------------------------------
// IsSimilar (ReSTIRGICommon.hlsl:99-109).
__device__ __forceinline__ bool similar(float4 cx, float4 cn, float4 hx, float4 hn) {
    const v3 d = sub(xyz(cx), xyz(hx));
    if (dot(d, d) > 0.1f) return false;
    if (dot(xyz(cn), xyz(hn)) < 0.8f) return false;
    if (__float_as_uint(cx.w) != __float_as_uint(hx.w)) return false;
    return true;
}

// The reservoir r of Spatial.rgen is kept as its w_sum and M plus the index of
// the pixel whose temporal reservoir holds the selected sample (MergeReservoir
// copies that sample; the copy is made once, after the loop) and that
// sample's visible point and normal (the origin of the next visibility ray).
// Per neighbour only the planes that p_hat and the merge weight read are
// loaded (x_s, n_s, L_out + W, M).
//
// One pixel's pass is a resumable task: spatial_step runs the neighbour loop
// until a neighbour needs a visibility ray (the ray and the merge it decides
// are left pending) or the loop ends; the caller traces the ray and applies
// the merge.  The draws and float operations are those of the loop in
// Spatial.rgen:64-183, in the same order, whoever drives the task.
struct SpatialTask {
    uint32_t pix, x, y, rng, i, sel, m;
    float ws;
    float4 s0, s1;  // centre x_view / material, n_view
    v3 sel_x, sel_n;
    v3 ro, rdir;    // pending visibility ray
    float tmax;
    float pph, pW;  // pending merge: p_hat, W, M, source pixel
    uint32_t pM, pnidx;
};

// The centre's early outs (Spatial.rgen:66-74): no visible surface (black)
// or no sample (white), written without a reservoir; true when taken.
__device__ __forceinline__ bool spatial_early_out(const FrameDev& fr, uint32_t pix, float4 s1, float4 s3) {
    if (length(xyz(s1)) == 0.0f) {
        fr.output[pix] = 0xFF000000u;
        if (fr.radiance) fr.radiance[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
        return true;
    }
    if (length(xyz(s3)) == 0.0f) {
        fr.output[pix] = 0xFFFFFFFFu;
        if (fr.radiance) fr.radiance[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
        return true;
    }
    return false;
}

// Starts the task on an empty reservoir and merges the centre's temporal
// reservoir (MergeReservoir(r, centerRes, |L|)) from its |L_out| (t.s1.w),
// W and M; true when the draw takes the centre's sample.
__device__ __forceinline__ bool spatial_centre_merge(SpatialTask& t, uint32_t width, uint32_t frame, float Wc,
                                                     uint32_t mc) {
    t.rng = init_seed(t.x, t.y, width, frame);
    t.sel = 0xFFFFFFFFu;  // none yet: InitReservoir's sample
    t.sel_x = mk3(0.f, 0.f, 0.f);
    t.sel_n = mk3(0.f, 0.f, 0.f);
    t.ws = 0.0f;
    t.i = 0;
    const float wc = t.s1.w * Wc * (float)mc;
    t.ws = t.ws + wc;
    const bool take = rand01(&t.rng) < wc / maxf(1e-10f, t.ws);
    t.m = 0u + mc;
    return take;
}

// Loads the centre and merges its reservoir; false when the pixel takes an
// early out (written here) and needs nothing more.
__device__ __forceinline__ bool spatial_begin(SpatialTask& t, const FrameDev& fr, const float4* __restrict__ rd,
                                              uint32_t frame, uint32_t x, uint32_t y) {
    const uint32_t n = fr.n;
    const uint32_t pix = y * fr.width + x;
    t.pix = pix;
    t.x = x;
    t.y = y;
    const float4* c = fr.nbr + 4 * (size_t)pix;
    t.s0 = c[0];
    t.s1 = c[1];
    if (spatial_early_out(fr, pix, t.s1, fr.samples[3 * n + pix])) return false;
    t.rng = init_seed(x, y, fr.width, frame);
    t.sel = 0xFFFFFFFFu;  // none yet: InitReservoir's sample
    t.sel_x = mk3(0.f, 0.f, 0.f);
    t.sel_n = mk3(0.f, 0.f, 0.f);
    t.ws = 0.0f;
    t.i = 0;
    // MergeReservoir(r, centerRes, |L|), as spatial_centre_merge
    const uint32_t mc = __float_as_uint(c[3].w);
    const float wc = t.s1.w * c[2].w * (float)mc;
    t.ws = t.ws + wc;
    if (rand01(&t.rng) < wc / maxf(1e-10f, t.ws)) {
        t.sel = pix;
        t.sel_x = xyz(rd[pix]);
        t.sel_n = xyz(rd[n + pix]);
    }
    t.m = 0u + mc;
    return true;
}

// MergeReservoir(r, r_neighbor, ph): w = ph * W * M, update (one draw),
// m = M0 + M; then m -= M if ph == 0.
__device__ __forceinline__ void spatial_merge(SpatialTask& t, const float4* __restrict__ rd, uint32_t n, float ph,
                                              float W, uint32_t M, uint32_t nidx) {
    const uint32_t m0 = t.m;
    const float wn = ph * W * (float)M;
    t.ws = t.ws + wn;
    if (rand01(&t.rng) < wn / maxf(1e-10f, t.ws)) {
        t.sel = nidx;
        t.sel_x = xyz(rd[nidx]);
        t.sel_n = xyz(rd[n + nidx]);
    }
    t.m = m0 + M;
    if (ph == 0.0f) t.m = t.m - M;
}

// Runs neighbours until one needs a visibility ray (returns true, ray and
// merge pending) or the loop ends (returns false).
__device__ __forceinline__ bool spatial_step(SpatialTask& t, const FrameDev& fr, const float4* __restrict__ rd,
                                             const PassParams& pp) {
    const uint32_t n = fr.n, w = fr.width, h = fr.height;
    while (t.i < pp.spatial_samples) {
        ++t.i;
        const float ra = rand01(&t.rng) * 2.0f * XR_PI;
        const float rr = sqrtf(rand01(&t.rng)) * pp.spatial_radius;
        float sa, ca;
        sincos_xr(ra, &sa, &ca);
        int nx = (int)t.x + (int)floorf(ca * rr);
        int ny = (int)t.y + (int)floorf(sa * rr);
        nx = nx < 0 ? 0 : (nx > (int)w - 1 ? (int)w - 1 : nx);
        ny = ny < 0 ? 0 : (ny > (int)h - 1 ? (int)h - 1 : ny);
        const uint32_t nidx = (uint32_t)ny * w + (uint32_t)nx;
        const float4* nb = fr.nbr + 4 * (size_t)nidx;
        const float4 q0 = nb[0];
        const float4 q1 = nb[1];
        if (length(xyz(q1)) == 0.0f || !similar(t.s0, t.s1, q0, q1)) continue;
        const float4 rn2 = nb[2], rn3 = nb[3];
        const uint32_t mn = __float_as_uint(rn3.w);
        const v3 xs = xyz(rn2);
        const v3 ns = xyz(rn3);
        const v3 vo = sub(xyz(q0), xs);
        const float d2o = dot(vo, vo);
        const v3 vc = sub(xyz(t.s0), xs);
        const float d2c = dot(vc, vc);
        float jac = 1.0f;
        if (d2o > 1e-6f && d2c > 1e-6f) {
            const v3 dir_o = muls(vo, 1.0f / sqrtf(d2o));
            const v3 dir_c = muls(vc, 1.0f / sqrtf(d2c));
            const float co = maxf(0.0f, dot(ns, dir_o));
            const float cc = maxf(0.0f, dot(ns, dir_c));
            if (co > 1e-6f) jac = (cc * d2o) / (co * d2c);
        }
        jac = clampf(jac, 0.1f, 10.0f);
        const float ph = q1.w * jac;
        // CheckVisibility (Spatial.rgen:44-62) from the currently selected
        // sample's visible point, any-hit.  Its only effect is to zero ph,
        // so it is traced only when ph is not already zero (same result,
        // fewer rays: samples that found no light carry L_out = 0).
        if (pp.visibility && ph != 0.0f) {
            const v3 dv = sub(xs, t.sel_x);
            const float dist = length(dv);
            t.ro = add(t.sel_x, muls(t.sel_n, 0.01f));
            t.rdir = normalize(dv);
            t.tmax = dist - 0.01f;
            t.pph = ph;
            t.pW = rn2.w;
            t.pM = mn;
            t.pnidx = nidx;
            return true;
        }
        spatial_merge(t, rd, n, ph, rn2.w, mn, nidx);
    }
    return false;
}

// After the loop: take_sel(r) copies the selected sample in (if any), then
// M cap, W, shading, writes.
template <class TakeSel>
__device__ __forceinline__ void spatial_write(const SpatialTask& t, const FrameDev& fr, TakeSel take_sel,
                                              float4* __restrict__ wr, uint32_t reuse, const PassParams& pp) {
    const uint32_t n = fr.n;
    Res r;
    res_init(&r);
    take_sel(r);
    set_wsum(&r, t.ws);
    set_m(&r, t.m);
    if (r_m(r) > pp.spatial_m_cap) set_m(&r, pp.spatial_m_cap);
    const float phf = length(xyz(r.p[4]));
    if (phf > 1e-6f && r_m(r) > 0u) set_W(&r, r_wsum(r) / ((float)r_m(r) * phf));
    else set_W(&r, 0.0f);
    v3 color = mk3(0.f, 0.f, 0.f);
    if (r_W(r) > 0.0f) color = shade_reservoir(r, xyz(t.s0), xyz(t.s1));
    store_planes(wr, n, t.pix, r.p);
    if (reuse) {
        fr.output[t.pix] = resolve_bgra8(color, 1.0f);
        if (fr.radiance) fr.radiance[t.pix] = f4(color, 0.0f);
    }
}

__device__ __forceinline__ void spatial_finish(const SpatialTask& t, const FrameDev& fr, const float4* __restrict__ rd,
                                               float4* __restrict__ wr, uint32_t reuse, const PassParams& pp) {
    spatial_write(
        t, fr,
        [&](Res& r) {
            if (t.sel != 0xFFFFFFFFu) {
                float4 sp[kPlanes];
                load_planes(rd, fr.n, t.sel, sp);
                take_sample(&r, sp);
            }
        },
        wr, reuse, pp);
}

// The spatial pass of one pixel without neighbours, in the temporal pass's
// kernel: s = the pixel's sample planes, r = the temporal reservoir just
// written for it (what spatial_begin and spatial_finish would read back from
// the neighbour record and the temporal reservoirs), t_color = the temporal
// pass's shading, which stays the output unless reuse.
__device__ __forceinline__ void spatial_after_temporal(const FrameDev& fr, uint32_t frame, uint32_t reuse,
                                                       const PassParams& pp, uint32_t x, uint32_t y,
                                                       const float4* s, const Res& r, v3 t_color) {
    const uint32_t n = fr.n, pix = y * fr.width + x;
    SpatialTask t;
    t.pix = pix;
    t.x = x;
    t.y = y;
    t.s0 = s[0];
    t.s1 = make_float4(s[1].x, s[1].y, s[1].z, length(xyz(r.p[4])));
    if (spatial_early_out(fr, pix, t.s1, s[3])) return;
    const bool take = spatial_centre_merge(t, fr.width, frame, r_W(r), r_m(r));
    float4* wr = fr.spatial + (size_t)(1u - frame % 2u) * kPlanes * n;
    spatial_write(
        t, fr,
        [&](Res& rs) {
            if (take) take_sample(&rs, r.p);
        },
        wr, reuse, pp);
    if (!reuse) {
        fr.output[pix] = resolve_bgra8(t_color, 1.0f);
        if (fr.radiance) fr.radiance[pix] = f4(t_color, 0.0f);
    }
}

template <int kLds>
__device__ __forceinline__ uint32_t spatial_pixel(const SceneDev& sc, const FrameDev& fr, uint32_t frame,
                                                  uint32_t reuse, const PassParams& pp, uint32_t x, uint32_t y,
                                                  int* col) {
    const uint32_t n = fr.n, ping = frame % 2u;
    const float4* rd = fr.temporal + (size_t)ping * kPlanes * n;
    float4* wr = fr.spatial + (size_t)(1u - ping) * kPlanes * n;
    SpatialTask t;
    if (!spatial_begin(t, fr, rd, frame, x, y)) return 0u;
    uint32_t rays = 0;
    while (spatial_step(t, fr, rd, pp)) {
        const RayPre rp = make_ray_pre(t.ro, t.rdir);
        Hit hh;
        ++rays;
        const bool blocked = trace_ray<kBlock, true, kLds>(sc.nodes, sc.tris, rp, 0.0f, t.tmax, col, &hh);
        spatial_merge(t, rd, n, blocked ? 0.0f : t.pph, t.pW, t.pM, t.pnidx);
    }
    spatial_finish(t, fr, rd, wr, reuse, pp);
    return rays;
}

// The spatial pass keeps 20 stack levels in LDS (20 KiB per workgroup) and
// runs at 7 waves per SIMD (72 VGPRs): its any-hit rays rarely go deeper, and
// the extra wave took the 1080p pass from 1.99 to 1.91 ms (16 levels at 7
// or 8 waves, 12 at 8 and 24 at 5 or 6 measured slower).
constexpr int kSpatialLds = 20;
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_spatial(SceneDev sc, FrameDev fr,
                                                                                         uint32_t frame, uint32_t reuse,
                                                                                         PassParams pp) {
    __shared__ int stack[kSpatialLds * kBlock];
    __shared__ uint32_t lds_count;
    const TileCoord tc = tile_pixel(fr.width, fr.height);
    uint32_t rays = 0;
    if (tc.in_image)
        rays = spatial_pixel<kSpatialLds>(sc, fr, frame, reuse, pp, tc.x, tc.y, stack + threadIdx.x);
    count_rays(&fr.counters[1], rays, &lds_count);
}

// ---------------------------------------------------------------------------
// Batched views (no reference counterpart): several cameras on one scene, each
// with its own frame buffers; one launch per pass covers every view.  The
// tiled kernels are instantiated for several tile shapes and LDS stack depths
// (the autotuner times them on the actual views and keeps the fastest); all
// give identical results.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t view_blocks(uint32_t w, uint32_t h, uint32_t tw, uint32_t th) {
    const uint32_t tiles_x = (w + tw - 1) / tw;
    const uint32_t tiles_y = (h + th - 1) / th;
    return kXcds * ((tiles_y + kXcds - 1) / kXcds) * tiles_x;
}

// View of workgroup b of a tiled batched grid; *local = b within the view's grid.
__device__ __forceinline__ uint32_t batch_view(const BatchDev& bd, uint32_t b, uint32_t tw, uint32_t th,
                                               uint32_t* local) {
    uint32_t v = 0;
    for (; v + 1 < bd.nviews; ++v) {
        const uint32_t g = view_blocks(bd.views[v].width, bd.views[v].height, tw, th);
        if (b < g) break;
        b -= g;
    }
    *local = b;
    return v;
}

// Next tile of a batched tiled kernel.  Grid form: the workgroup's own tile,
// once.  Persistent form (a grid of as many workgroups as the GPU holds at
// once, each looping): workgroup g serves XCD g % 8 (the hardware deals
// workgroup g to XCD g % 8, checked with the XCC_ID register on a 1,792- and
// a 32,768-workgroup grid) and takes that XCD's tiles t * 8 + g % 8 from the
// XCD's queue counter in order, so each XCD still walks its own band of tile
// rows while its workgroups balance the load among themselves; false when
// the XCD has no tile left (tiles is a multiple of 8).  Called by every
// thread of the workgroup.
template <bool kPersistent>
__device__ __forceinline__ bool batch_tile(uint32_t* queue, uint32_t tiles, uint32_t it, uint32_t* lds_tile,
                                           uint32_t* b) {
    if (!kPersistent) {
        *b = blockIdx.x;
        return it == 0u;
    }
    const uint32_t x = blockIdx.x % kXcds;
    __syncthreads();  // every thread has read the previous tile
    if (threadIdx.x == 0) *lds_tile = atomicAdd(queue + x, 1u) * kXcds + x;
    __syncthreads();
    *b = *lds_tile;
    return *b < tiles;
}

// With the radiance cache (cache.enabled; its frame number comes from each
// view's UBO) the candidates park their records in their view's pending
// queue, as k_initial_sample does, for the view's training choice.
template <int kTw, int kTh, int kLds, bool kPersistent>
__global__ __launch_bounds__(kBlock) XR_TRAVERSAL_OCCUPANCY void k_batch_initial(SceneDev sc, BatchDev bd,
                                                                                  uint32_t tiles, CacheCfg cache) {
    __shared__ int stack[kLds * kBlock];
    __shared__ uint32_t lds_count;
    __shared__ uint32_t lds_base;
    __shared__ uint32_t lds_tile;
    uint32_t rays = 0, valid = 0, b = 0;
    for (uint32_t it = 0; batch_tile<kPersistent>(bd.tile_queue, tiles, it, &lds_tile, &b); ++it) {
        uint32_t local;
        const uint32_t v = batch_view(bd, b, kTw, kTh, &local);
        const FrameDev fr = bd.views[v];
        const UniformBufferObject& ubo = bd.ubos[v];
        const TileCoord tc = tile_coord<kTw, kTh>(local, fr.width, fr.height);
        bool alive = false;
        uint32_t kind = 0;
        PathState ps;
        PathOut o;
        if (tc.in_image) {
            alive = initial_head_pixel<kLds>(sc, fr, ubo, tc.x, tc.y, cache, stack + threadIdx.x, &ps, &o, &rays);
            kind = o.cache;
            valid += o.valid ? 1u : 0u;
        }
        const uint32_t pix = tc.y * fr.width + tc.x;
        const bool pending = kind >= kCachePending;
        hand_on(alive && !pending, bd.tail, bd.tail_count, pix, v, ps, o, &lds_count, &lds_base);
        if (cache.enabled) {  // uniform over the launch
            const uint32_t q = block_slot(pending, fr.pending_count, &lds_count, &lds_base);
            if (pending) {
                o.cache = kind | (alive ? kRecordAlive : 0u);
                store_path(fr.pending + 4 * (size_t)q, pix, v, ps, o);
            }
            if (tc.in_image) fr.cache_slot[pix] = pending ? q + 1u : 0u;
        }
    }
    count_rays(&bd.counters[0], rays, &lds_count);
    count_rays(&bd.counters[2], valid, &lds_count);
}

template <int kLds>
__global__ __launch_bounds__(kBlock) XR_TRAVERSAL_OCCUPANCY void k_batch_tail(SceneDev sc, BatchDev bd) {
    __shared__ int stack[kLds * kBlock];
    __shared__ uint32_t lds_count;
    const uint32_t count = *bd.tail_count;
    const CacheCfg nocache = {0u, 1u, 0u, 0.0f};
    uint32_t rays = 0;
    for (uint32_t q = blockIdx.x * kBlock + threadIdx.x; q < count; q += gridDim.x * kBlock) {
        const float4* rec = bd.tail + 4 * (size_t)q;
        const uint32_t v = __float_as_uint(rec[3].w);
        initial_tail_path<kLds>(sc, bd.views[v], bd.ubos[v], rec, nocache, stack + threadIdx.x, &rays);
    }
    count_rays(&bd.counters[0], rays, &lds_count);
}

__global__ __launch_bounds__(256) void k_batch_temporal(BatchDev bd, PassParams pp) {
    uint32_t p = blockIdx.x * 256u + threadIdx.x;
    uint32_t v = 0;
    for (; v < bd.nviews; ++v) {
        if (p < bd.views[v].n) break;
        p -= bd.views[v].n;
    }
    if (v == bd.nviews) return;
    const FrameDev fr = bd.views[v];
    temporal_pixel(fr, bd.ubos[v].currentFrame, pp, fr.gpos, p);
}

// With at most 20 LDS stack levels (20 KiB) 7 workgroups fit a CU: ask for 7
// waves, as the single-view spatial kernel does.
template <int kTw, int kTh, int kLds, bool kPersistent>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kLds <= 20 ? 7 : 6))) void k_batch_spatial(
    SceneDev sc, BatchDev bd, PassParams pp, uint32_t tiles) {
    __shared__ int stack[kLds * kBlock];
    __shared__ uint32_t lds_count;
    __shared__ uint32_t lds_tile;
    uint32_t rays = 0, b = 0;
    for (uint32_t it = 0; batch_tile<kPersistent>(bd.tile_queue + kXcds, tiles, it, &lds_tile, &b); ++it) {
        uint32_t local;
        const uint32_t v = batch_view(bd, b, kTw, kTh, &local);
        const FrameDev fr = bd.views[v];
        const UniformBufferObject& ubo = bd.ubos[v];
        const TileCoord tc = tile_coord<kTw, kTh>(local, fr.width, fr.height);
        if (tc.in_image)
            rays += spatial_pixel<kLds>(sc, fr, ubo.currentFrame, ubo.reuse, pp, tc.x, tc.y, stack + threadIdx.x);
    }
    count_rays(&bd.counters[1], rays, &lds_count);
}

// Tile shapes x LDS stack depths x launch forms the autotuner chooses from.
struct BatchKernels {
    uint32_t tw, th, lds;
    bool persistent;
    void (*initial)(SceneDev, BatchDev, uint32_t, CacheCfg);
    void (*tail)(SceneDev, BatchDev);
    void (*spatial)(SceneDev, BatchDev, PassParams, uint32_t);
};
#define XR_BATCH_CFG(TW, TH, L, P) \
    {TW, TH, L, P, k_batch_initial<TW, TH, L, P>, k_batch_tail<L>, k_batch_spatial<TW, TH, L, P>}
static const BatchKernels kBatchKernels[] = {
    XR_BATCH_CFG(32, 8, 20, false),  XR_BATCH_CFG(32, 8, 16, false),  XR_BATCH_CFG(32, 8, 24, false),
    XR_BATCH_CFG(16, 16, 20, false), XR_BATCH_CFG(16, 16, 16, false), XR_BATCH_CFG(16, 16, 24, false),
    XR_BATCH_CFG(8, 32, 20, false),  XR_BATCH_CFG(8, 32, 16, false),  XR_BATCH_CFG(8, 32, 24, false),
    XR_BATCH_CFG(32, 8, 20, true),   XR_BATCH_CFG(32, 8, 16, true),   XR_BATCH_CFG(32, 8, 24, true),
    XR_BATCH_CFG(16, 16, 20, true),  XR_BATCH_CFG(16, 16, 16, true),  XR_BATCH_CFG(16, 16, 24, true),
    XR_BATCH_CFG(8, 32, 20, true),   XR_BATCH_CFG(8, 32, 16, true),   XR_BATCH_CFG(8, 32, 24, true),
};
#undef XR_BATCH_CFG
constexpr uint32_t kBatchConfigs = sizeof(kBatchKernels) / sizeof(kBatchKernels[0]);

// Workgroups of kernel k resident on the whole GPU at once (a multiple of 8),
// found once per configuration and kernel.
uint32_t resident_grid(uint32_t cfg, int which) {
    static std::mutex mu;
    static uint32_t cached[kBatchConfigs][2] = {};
    std::lock_guard<std::mutex> lock(mu);
    if (!cached[cfg][which]) {
        int dev = 0, cus = 0, per_cu = 0;
        const void* fn = which == 0 ? (const void*)kBatchKernels[cfg].initial : (const void*)kBatchKernels[cfg].spatial;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, 0) != hipSuccess || per_cu < 1)
            per_cu = 1, cus = cus > 0 ? cus : 256;
        const uint32_t g = (uint32_t)(cus * per_cu) / kXcds * kXcds;
        cached[cfg][which] = g >= kXcds ? g : kXcds;
    }
    return cached[cfg][which];
}

uint32_t batch_config_count() { return kBatchConfigs; }

void batch_config_info(uint32_t cfg, uint32_t* tw, uint32_t* th, uint32_t* lds, bool* persistent) {
    *tw = kBatchKernels[cfg].tw;
    *th = kBatchKernels[cfg].th;
    *lds = kBatchKernels[cfg].lds;
    *persistent = kBatchKernels[cfg].persistent;
}

// Per-frame counts of the radiance cache in every view: query and training
// list lengths and pending records.
__global__ void k_batch_clear_cache(BatchDev bd) {
    const uint32_t v = threadIdx.x;
    if (v >= bd.nviews) return;
    bd.views[v].cache_list[0] = 0u;
    bd.views[v].cache_list[1] = 0u;
    *bd.views[v].pending_count = 0u;
}

// Training samples of all views together (the mean loss of the batch's
// training step divides by it).
__global__ void k_batch_train_total(BatchDev bd, uint32_t* total) {
    uint32_t t = 0;
    for (uint32_t v = threadIdx.x; v < bd.nviews; v += 64u) t += bd.views[v].cache_list[1];
    for (int off = 32; off > 0; off >>= 1) t += (uint32_t)__shfl_xor((int)t, off);
    if (threadIdx.x == 0) *total = t;
}

hipError_t launch_batch_train_total(const BatchDev& bd, uint32_t* total, hipStream_t st) {
    hipLaunchKernelGGL(k_batch_train_total, dim3(1), dim3(64), 0, st, bd, total);
    return hipGetLastError();
}

// The initial-sample pass of every view (head, the views' training choices
// in cache mode, tail).  views: host copy of the view table, bd: its device
// twin.  One launch clears the tail count and the persistent kernels' tile
// queues.
hipError_t launch_batch_initial(const SceneDev& sc, const BatchDev& bd, const FrameDev* views, uint32_t cfg,
                                const CacheCfg& cache, hipStream_t st) {
    const BatchKernels& k = kBatchKernels[cfg];
    uint32_t tiles = 0;
    for (uint32_t v = 0; v < bd.nviews; ++v) tiles += image_grid(views[v].width, views[v].height, k.tw, k.th);
    const uint32_t gi = k.persistent ? std::min(tiles, resident_grid(cfg, 0)) : tiles;
    hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(64), 0, st, bd.tail_count, kBatchCountWords);
    if (cache.enabled) hipLaunchKernelGGL(k_batch_clear_cache, dim3(1), dim3(64), 0, st, bd);
    hipLaunchKernelGGL(k.initial, dim3(gi), dim3(kBlock), 0, st, sc, bd, tiles, cache);
    if (cache.enabled)
        for (uint32_t v = 0; v < bd.nviews; ++v) {
            const hipError_t e = launch_cache_choice(views[v], cache, bd.ubos + v, st);
            if (e != hipSuccess) return e;
        }
    hipLaunchKernelGGL(k.tail, dim3(2048), dim3(kBlock), 0, st, sc, bd);
    return hipGetLastError();
}

// Temporal and spatial passes of every view.
hipError_t launch_batch_reuse(const SceneDev& sc, const BatchDev& bd, const FrameDev* views, uint32_t cfg,
                              const PassParams& pp, hipStream_t st) {
    const BatchKernels& k = kBatchKernels[cfg];
    uint32_t tiles = 0, pixels = 0;
    for (uint32_t v = 0; v < bd.nviews; ++v) {
        tiles += image_grid(views[v].width, views[v].height, k.tw, k.th);
        pixels += views[v].n;
    }
    const uint32_t gs = k.persistent ? std::min(tiles, resident_grid(cfg, 1)) : tiles;
    hipLaunchKernelGGL(k_batch_temporal, dim3((pixels + 255) / 256), dim3(256), 0, st, bd, pp);
    hipLaunchKernelGGL(k.spatial, dim3(gs), dim3(kBlock), 0, st, sc, bd, pp, tiles);
    return hipGetLastError();
}

hipError_t launch_batch_frame(const SceneDev& sc, const BatchDev& bd, const FrameDev* views, uint32_t cfg,
                              const PassParams& pp, hipStream_t st) {
    const CacheCfg nocache = {0u, 1u, 0u, 0.0f};
    const hipError_t e = launch_batch_initial(sc, bd, views, cfg, nocache, st);
    if (e != hipSuccess) return e;
    return launch_batch_reuse(sc, bd, views, cfg, pp, st);
}

// ---------------------------------------------------------------------------
// Plain path tracer (rgen.hlsl:31-151)
// ---------------------------------------------------------------------------
// The plain closest-hit program (rhit.hlsl:29-59) returns the same position,
// normal, UV and material as the ReSTIR one; it does not compute the
// geometric normal, which the plain loop does not use.
// State of one plain path between bounces.
struct PlainPath {
    v3 origin, dir, thr, color;
    uint32_t rng;
};

// Bounces b0, b0+1, ... of a plain path (the loop of rgen.hlsl:59-136);
// returns true when the path is still alive at bounce b1.
__device__ __forceinline__ bool plain_bounces(const SceneDev& sc, const UniformBufferObject& ubo, uint32_t b0,
                                              uint32_t b1, uint32_t max_depth, PlainPath* pp, int* col,
                                              uint32_t* rays) {
    for (uint32_t bounce = b0;; ++bounce) {
        if (max_depth != 0u && bounce == max_depth) return false;
        if (bounce == b1) return true;
        const RayPre rp = make_ray_pre(pp->origin, pp->dir);
        Hit hit;
        ++*rays;
        Payload pl;
        bool found = trace_ray<kBlock, false>(sc.nodes, sc.tris, rp, 0.0001f, 10000.0f, col, &hit);
        if (found) {
            closest_hit(sc, hit, &pl);
            found = pl.material != -1;
        }
        if (!found) {
            const float t = 0.5f * (pp->dir.y + 1.0f);
            // lerp(a, b, t) = a + t*(b - a)
            const v3 sky = mk3(1.0f + t * (0.5f - 1.0f), 1.0f + t * (0.7f - 1.0f), 1.0f + t * (1.0f - 1.0f));
            if (ubo.hasSkyBox) pp->color = add(pp->color, mul(pp->thr, sky));
            return false;
        }
        const float4 m0 = sc.materials[2 * pl.material];
        const float4 m1 = sc.materials[2 * pl.material + 1];
        // untextured baseColor.rgb * 4 (rgen.hlsl:96-99)
        if (__float_as_uint(m1.w) == kDiffuseLight) {
            pp->color = add(pp->color, muls(mul(pp->thr, mk3(m0.x, m0.y, m0.z)), 4.0f));
            return false;
        }
        v3 wi;
        float pdf, cos_t;
        const v3 f = sample_bsdf(sc, pl.normal, m0, m1, pl.u, pl.v, &wi, &pdf, &cos_t, &pp->rng);
        if (pdf <= 1e-6f) return false;
        pp->thr = mul(pp->thr, divs(muls(f, cos_t), pdf));
        if (bounce > 3) {
            const float p = maxf(pp->thr.x, maxf(pp->thr.y, pp->thr.z));
            if (rand01(&pp->rng) > p) return false;
            pp->thr = muls(pp->thr, 1.0f / p);
        }
        const v3 hp = add(pp->origin, muls(pp->dir, pl.t));
        const v3 od = dot(wi, pl.normal) > 0.0f ? pl.normal : neg(pl.normal);
        pp->origin = add(hp, muls(od, 0.001f));
        pp->dir = wi;
    }
}

// Accumulation and output of one pixel (rgen.hlsl:139-150).
__device__ __forceinline__ void plain_finish(const FrameDev& fr, const UniformBufferObject& ubo, uint32_t pix,
                                             v3 pixel) {
    const bool accumulate = ubo.spp != ubo.totalNumberOfSamples;
    v3 acc = mk3(0.f, 0.f, 0.f);
    if (accumulate) acc = xyz(fr.accum[pix]);
    acc = add(acc, pixel);
    const v3 fin = divs(acc, (float)ubo.totalNumberOfSamples);
    fr.accum[pix] = f4(acc, 0.0f);
    fr.output[pix] = resolve_bgra8(fin, 0.0f);
}

__device__ __forceinline__ void plain_camera(const UniformBufferObject& ubo, uint32_t x, uint32_t y,
                                             const FrameDev& fr, PlainPath* pp) {
    camera_ray(ubo, x, y, fr.width, fr.height, &pp->rng, &pp->origin, &pp->dir);
    pp->color = mk3(0.f, 0.f, 0.f);
    pp->thr = mk3(1.f, 1.f, 1.f);
}

// Every sample of every pixel in one kernel (any spp).
__global__ __launch_bounds__(kBlock) XR_TRAVERSAL_OCCUPANCY void k_plain(SceneDev sc, FrameDev fr, UniformBufferObject ubo,
                                                  uint32_t max_depth) {
    __shared__ int stack[kLdsDepth * kBlock];
    __shared__ uint32_t lds_count;
    const TileCoord tc = tile_pixel(fr.width, fr.height);
    uint32_t rays = 0;
    if (tc.in_image) {
        const uint32_t x = tc.x, y = tc.y;
        PlainPath pp;
        pp.rng = init_seed(x, y, fr.width, ubo.currentFrame);
        v3 pixel = mk3(0.f, 0.f, 0.f);
        for (uint32_t i = 0; i < ubo.spp; ++i) {
            plain_camera(ubo, x, y, fr, &pp);
            plain_bounces(sc, ubo, 0u, 0xFFFFFFFFu, max_depth, &pp, stack + threadIdx.x, &rays);
            pixel = add(pixel, pp.color);
        }
        plain_finish(fr, ubo, y * fr.width + x, pixel);
    }
    count_rays(&fr.counters[0], rays, &lds_count);
}

// One sample per pixel as a head kernel (bounces 0-1) and one launch per
// further bounce, the paths handed on through the two path queues as in the
// initial-sample pass (a sample's RNG continues from the previous sample's,
// so only spp = 1 splits).  Record: [0] origin, pixel [1] direction, rng
// [2] throughput [3] colour so far.
__device__ __forceinline__ void plain_store(float4* rec, uint32_t pix, const PlainPath& pp) {
    rec[0] = f4(pp.origin, __uint_as_float(pix));
    rec[1] = f4(pp.dir, __uint_as_float(pp.rng));
    rec[2] = f4(pp.thr, 0.0f);
    rec[3] = f4(pp.color, 0.0f);
}

__device__ __forceinline__ void plain_hand_on(bool alive, float4* queue, uint32_t* count, uint32_t pix,
                                              const PlainPath& pp, uint32_t* lds_count, uint32_t* lds_base) {
    const uint32_t q = block_slot(alive, count, lds_count, lds_base);
    if (alive) plain_store(queue + 4 * (size_t)q, pix, pp);
}

__global__ __launch_bounds__(kBlock) XR_TRAVERSAL_OCCUPANCY void k_plain_head(SceneDev sc, FrameDev fr,
                                                                               UniformBufferObject ubo,
                                                                               uint32_t max_depth) {
    __shared__ int stack[kLdsDepth * kBlock];
    __shared__ uint32_t lds_count;
    __shared__ uint32_t lds_base;
    const TileCoord tc = tile_pixel(fr.width, fr.height);
    uint32_t rays = 0;
    bool alive = false;
    PlainPath pp;
    const uint32_t pix = tc.y * fr.width + tc.x;
    if (tc.in_image) {
        pp.rng = init_seed(tc.x, tc.y, fr.width, ubo.currentFrame);
        plain_camera(ubo, tc.x, tc.y, fr, &pp);
        alive = plain_bounces(sc, ubo, 0u, kHeadBounces, max_depth, &pp, stack + threadIdx.x, &rays);
        if (!alive) plain_finish(fr, ubo, pix, add(mk3(0.f, 0.f, 0.f), pp.color));
    }
    plain_hand_on(alive, fr.tail, fr.tail_count, pix, pp, &lds_count, &lds_base);
    count_rays(&fr.counters[0], rays, &lds_count);
}

// Bounce `bounce` of the handed-on plain paths, or (few paths left, or the
// last launch) all their remaining bounces; as k_tail_bounce.
__global__ __launch_bounds__(kBlock) XR_TRAVERSAL_OCCUPANCY void k_plain_bounce(SceneDev sc, FrameDev fr,
                                                                                 UniformBufferObject ubo,
                                                                                 uint32_t max_depth, uint32_t bounce,
                                                                                 uint32_t last) {
    __shared__ int stack[kLdsDepth * kBlock];
    __shared__ uint32_t lds_count;
    __shared__ uint32_t lds_base;
    const uint32_t j = bounce - kHeadBounces, side = j & 1u;
    const float4* in = fr.tail + (size_t)side * 4 * fr.n;
    float4* out = fr.tail + (size_t)(side ^ 1u) * 4 * fr.n;
    const uint32_t count = fr.tail_count[j % 3u];
    if (blockIdx.x == 0u && threadIdx.x == 0u) fr.tail_count[(j + 2u) % 3u] = 0u;  // the next launch's output
    if (blockIdx.x * kBlock >= count) return;  // uniform over the workgroup
    const bool one = !last && (uint64_t)count * 4u > (uint64_t)fr.n;  // uniform over the launch
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    uint32_t rays = 0;
    bool alive = false;
    PlainPath pp;
    uint32_t pix = 0;
    if (q < count) {
        const float4* rec = in + 4 * (size_t)q;
        const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3];
        pix = __float_as_uint(r0.w);
        pp.origin = xyz(r0);
        pp.dir = xyz(r1);
        pp.rng = __float_as_uint(r1.w);
        pp.thr = xyz(r2);
        pp.color = xyz(r3);
        alive = plain_bounces(sc, ubo, bounce, one ? bounce + 1u : 0xFFFFFFFFu, max_depth, &pp,
                              stack + threadIdx.x, &rays);
        if (!alive) plain_finish(fr, ubo, pix, add(mk3(0.f, 0.f, 0.f), pp.color));
    }
    if (one) plain_hand_on(alive, out, fr.tail_count + (j + 1u) % 3u, pix, pp, &lds_count, &lds_base);
    count_rays(&fr.counters[0], rays, &lds_count);
}

// ---------------------------------------------------------------------------
// Ray batch tracing (parity and microbenchmarks)
// ---------------------------------------------------------------------------
template <bool kAnyHit>
__global__ __launch_bounds__(kBlock) void k_trace(SceneDev sc, const float4* __restrict__ rays, uint32_t nrays,
                                                  float4* __restrict__ hits) {
    __shared__ int stack[kLdsDepth * kBlock];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nrays) return;
    const float4 a = rays[2 * i], b = rays[2 * i + 1];
    const RayPre rp = make_ray_pre(mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z));
    Hit hit;
    const bool found = trace_ray<kBlock, kAnyHit>(sc.nodes, sc.tris, rp, a.w, b.w, stack + threadIdx.x, &hit);
    float4 o0, o1;
    if (found) {
        const BvhTri tr = sc.tris[hit.idx];
        o0 = make_float4(hit.t, hit.u, hit.v, 0.0f);
        o1 = make_float4(tr.v1.w, tr.v2.w, __uint_as_float(hit.id), 0.0f);
    } else {
        o0 = make_float4(-1.0f, 0.0f, 0.0f, 0.0f);
        o1 = make_float4(__uint_as_float(0xFFFFFFFFu), __uint_as_float(0xFFFFFFFFu), __uint_as_float(0xFFFFFFFFu),
                         0.0f);
    }
    hits[2 * i] = o0;
    hits[2 * i + 1] = o1;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
hipError_t launch_cache_choice(const FrameDev& fr, const CacheCfg& cache, const UniformBufferObject* frame_ubo,
                               hipStream_t st) {
    const uint32_t nblocks = (fr.n + kCacheSelBlock - 1) / kCacheSelBlock;
    hipLaunchKernelGGL(k_cache_count, dim3(nblocks), dim3(256), 0, st, fr);
    hipLaunchKernelGGL(k_cache_scan, dim3(1), dim3(1024), 0, st, fr, nblocks);
    hipLaunchKernelGGL(k_cache_select, dim3(nblocks), dim3(kSelThreads), 0, st, fr, cache, frame_ubo);
    return hipGetLastError();
}

hipError_t launch_initial_sample(const SceneDev& sc, const FrameDev& fr, const UniformBufferObject& ubo,
                                 const CacheCfg& cache, hipStream_t st) {
    hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(64), 0, st, fr.tail_count, kTailCounters);
    hipLaunchKernelGGL(k_initial_sample, dim3(image_grid(fr.width, fr.height)), dim3(kBlock), 0, st, sc, fr, ubo,
                       cache);
    if (cache.enabled) launch_cache_choice(fr, cache, nullptr, st);
    const uint32_t nb = ubo.numberOfBounces;
    for (uint32_t b = kHeadBounces; b < nb; ++b)
        hipLaunchKernelGGL(k_tail_bounce, dim3((fr.n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, sc, fr, ubo,
                           cache, b);
    return hipGetLastError();
}
hipError_t launch_temporal(const FrameDev& fr, uint32_t frame, const PassParams& pp, const float4* gpos_hist,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_temporal, dim3((fr.n + 255) / 256), dim3(256), 0, st, fr, frame, pp, gpos_hist);
    return hipGetLastError();
}
hipError_t launch_temporal_spatial0(const FrameDev& fr, uint32_t frame, uint32_t reuse, const PassParams& pp,
                                    const float4* gpos_hist, hipStream_t st) {
    hipLaunchKernelGGL(k_temporal_spatial0, dim3((fr.n + 255) / 256), dim3(256), 0, st, fr, frame, pp, gpos_hist,
                       reuse);
    return hipGetLastError();
}
hipError_t launch_nbr_build(const FrameDev& fr, uint32_t frame, hipStream_t st) {
    hipLaunchKernelGGL(k_nbr_build, dim3((fr.n + 255) / 256), dim3(256), 0, st, fr, frame % 2u);
    return hipGetLastError();
}
hipError_t launch_spatial(const SceneDev& sc, const FrameDev& fr, uint32_t frame, uint32_t reuse,
                          const PassParams& pp, hipStream_t st) {
    hipLaunchKernelGGL(k_spatial, dim3(image_grid(fr.width, fr.height)), dim3(kBlock), 0, st, sc, fr, frame, reuse,
                       pp);
    return hipGetLastError();
}
hipError_t launch_plain(const SceneDev& sc, const FrameDev& fr, const UniformBufferObject& ubo, uint32_t max_depth,
                        hipStream_t st) {
    // One kernel for several samples per pixel, and where at most one bounce
    // would follow the head (the extra launches cost more than they save:
    // 0.136 against 0.159 ms for the 640x360 Cornell frame at depth 3).
    if (ubo.spp != 1u || (max_depth != 0u && max_depth <= kHeadBounces + 1u)) {
        hipLaunchKernelGGL(k_plain, dim3(image_grid(fr.width, fr.height)), dim3(kBlock), 0, st, sc, fr, ubo,
                           max_depth);
        return hipGetLastError();
    }
    // Launches past the head: one per bounce up to the depth cap, at most
    // kPlainLaunches (unbounded paths: the last launch finishes every path).
    constexpr uint32_t kPlainLaunches = 12;
    hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(64), 0, st, fr.tail_count, kTailCounters);
    hipLaunchKernelGGL(k_plain_head, dim3(image_grid(fr.width, fr.height)), dim3(kBlock), 0, st, sc, fr, ubo,
                       max_depth);
    const uint32_t end = max_depth != 0u ? max_depth : kHeadBounces + kPlainLaunches;
    for (uint32_t b = kHeadBounces; b < end; ++b)
        hipLaunchKernelGGL(k_plain_bounce, dim3((fr.n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, sc, fr, ubo,
                           max_depth, b, b + 1u == end ? 1u : 0u);
    return hipGetLastError();
}
hipError_t launch_converge(const SceneDev& sc, const FrameDev& fr, const UniformBufferObject& ubo,
                           uint32_t frame_base, uint32_t spp, double* acc, hipStream_t st) {
    hipLaunchKernelGGL(k_converge, dim3(image_grid(fr.width, fr.height)), dim3(kBlock), 0, st, sc, fr, ubo,
                       frame_base, spp, acc);
    return hipGetLastError();
}
hipError_t launch_sq_error(const float4* img, const double* acc, double inv_spp, uint32_t n, double* out,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_sq_error, dim3((n + 255) / 256), dim3(256), 0, st, img, acc, inv_spp, n, out);
    return hipGetLastError();
}
hipError_t launch_trace(const SceneDev& sc, const float4* rays, uint32_t n, float4* hits, bool any_hit,
                        hipStream_t st) {
    if (n == 0) return hipSuccess;
    const dim3 grid((n + kBlock - 1) / kBlock);
    if (any_hit) hipLaunchKernelGGL(k_trace<true>, grid, dim3(kBlock), 0, st, sc, rays, n, hits);
    else hipLaunchKernelGGL(k_trace<false>, grid, dim3(kBlock), 0, st, sc, rays, n, hits);
    return hipGetLastError();
}


// Resources of the frame kernels as the runtime reports them (registers,
// scratch and LDS per workgroup, resident workgroups per CU), for the
// measurement output.
hipError_t kernel_resources(uint32_t index, const char** name, hipFuncAttributes* attr, int* blocks_per_cu,
                            int* block) {
    struct Entry {
        const char* name;
        const void* fn;
        int block;
    };
    static const Entry table[kKernelInfoCount] = {
        {"k_initial_sample", reinterpret_cast<const void*>(&k_initial_sample), kBlock},
        {"k_tail_bounce", reinterpret_cast<const void*>(&k_tail_bounce), kBlock},
        {"k_temporal", reinterpret_cast<const void*>(&k_temporal), 256},
        {"k_spatial", reinterpret_cast<const void*>(&k_spatial), kBlock},
        {"k_plain", reinterpret_cast<const void*>(&k_plain), kBlock},
        {"k_converge", reinterpret_cast<const void*>(&k_converge), kBlock},
    };
    if (index >= (uint32_t)kKernelInfoCount) return hipErrorInvalidValue;
    *name = table[index].name;
    *block = table[index].block;
    hipError_t e = hipFuncGetAttributes(attr, table[index].fn);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, table[index].fn, table[index].block, 0);
}

}  // namespace xr

// Device-side views of the scene and of the per-frame buffers, plus the
// traversal routine shared by every kernel that traces rays.
#pragma once
#include <hip/hip_runtime.h>
#include "xr_layout.h"
#include "xr_math.h"
#include "xr_bvh.h"

namespace xr {

// Number of float4 planes of a sample / reservoir (see xr_layout.h).
constexpr int kPlanes = 6;

struct SceneDev {
    const float* vertices;      // AVertex stream, 9 dwords per vertex (binding 4)
    const uint32_t* indices;    // binding 5
    const float4* materials;    // AMaterial as 2 float4 (binding 6)
    const uint2* offsets;       // (indexOffset, vertexOffset) per instance (binding 7)
    const WideNode* nodes;      // four-child tree (xr_bvh.h)
    const BvhTri* tris;
    uint32_t material_count;
    const uint4* tex_info;    // (first texel, width, height, 0) per texture (binding 8)
    const uint32_t* texels;   // RGBA8, R in the low byte
};

struct FrameDev {
    uint32_t width, height, n;  // n = width*height
    float4* samples;            // kPlanes planes of n            (binding 9)
    float4* temporal;           // 2 slots x kPlanes planes of n  (binding 10)
    float4* spatial;            // 2 slots x kPlanes planes of n  (binding 11)
    float4* gpos;               // binding 12 (16 aliases it, see ReSTIRGIPipeline.cpp:171-199)
    float4* gnrm;               // binding 13 (17 aliases it)
    float4* gmatuv;             // binding 14
    uint32_t* motion;           // binding 15, RG16F
    uint32_t* output;           // binding 2, B8G8R8A8_UNORM
    float4* accum;              // binding 1, RGBA32F
    // Spatial-reuse neighbour records (no reference binding), one 64-byte
    // record per pixel, written by the temporal pass: [0] the pixel's sample
    // x_view + material, [1] its n_view + |L_out| of its temporal reservoir,
    // [2] that reservoir's x_sample + W, [3] its n_sample + M.  A neighbour
    // gather reads one record instead of six 16-byte planes.
    float4* nbr;
    float4* radiance;           // linear colour behind the output (no reference binding), or null
    unsigned long long* counters;  // [0] extend rays, [1] shadow rays, [2] valid hits, [3] / [4] cache
    // Radiance-cache work lists of the frame: [0] query count, [1] training
    // count, [2, 2+n) query pixels, [2+n, 2+2n) training pixels.
    uint32_t* cache_list;
    // Training choice of the frame (cache mode): per pixel 1 + the index of its
    // pending path record (0: not a candidate), and per 4096-pixel block the
    // candidates before it (exclusive scan; entry nblocks = the total).
    uint32_t* cache_slot;
    uint32_t* cache_blocks;
    // The candidates' pending path records (cache mode: 4 float4 each, n
    // capacity) and their count.  The single-view pass keeps them in its
    // second tail queue (tail + 4n, count tail_count[kPendingCount]); each
    // batched view has its own.
    float4* pending;
    uint32_t* pending_count;
    // Paths alive after the head bounces of the initial-sample pass (4 float4
    // each, n capacity) and their count; the single-view pass keeps a second
    // queue of n paths at tail + 4n.  Tail
    // launch j (bounce kHeadBounces + j) reads queue j % 2 with count
    // tail_count[j % 3], appends to the other queue with count
    // tail_count[(j + 1) % 3] and clears tail_count[(j + 2) % 3], the next
    // launch's output count (the previous launch's input, read by no one
    // during launch j), so one clear before the head serves every launch.
    float4* tail;
    uint32_t* tail_count;
};
constexpr uint32_t kPendingCount = 3;  // word of tail_count counting pending cache records
constexpr uint32_t kTailCounters = 4;  // words of tail_count

// Parameters of one ReSTIR GI / plain frame as the kernels see them.
struct PassParams {
    uint32_t temporal_m_cap;
    uint32_t spatial_m_cap;
    uint32_t spatial_samples;
    float spatial_radius;
    uint32_t visibility;       // 0: CheckVisibility returns true (reference), 1: trace
    uint32_t plain_max_depth;  // 0: unbounded (rgen.hlsl:59)
    uint32_t history;          // 0: validity read from the current G-buffer (reference aliasing), 1: previous frame's
};

// Column-major 4x4 times (v0, v1, v2, v3), rows summed left to right
// (HLSL mul(M, v) on a glm matrix).
XR_HD void mat_vec(const float* M, float v0, float v1, float v2, float v3, float* out4) {
    for (int r = 0; r < 4; ++r) out4[r] = M[r] * v0 + M[4 + r] * v1 + M[8 + r] * v2 + M[12 + r] * v3;
}

// --------------------------------------------------------------------------
// Traversal.  The stack lives in LDS, one column per lane
// (stack[level * kBlock + lane] -> consecutive lanes hit consecutive banks);
// levels beyond kLdsDepth spill to a private array.  A ray never holds more
// entries than the tree's stack bound (the largest sum of "children - 1" over
// the nodes of a root-to-leaf path), and xr_build_accel rejects trees whose
// bound exceeds kLdsDepth + kSpillDepth, so the stack cannot overflow.
// --------------------------------------------------------------------------
constexpr int kLdsDepth = 24;
constexpr int kSpillDepth = 40;
constexpr int kMaxTraversalDepth = kLdsDepth + kSpillDepth;
constexpr uint32_t kMaxNodeVisits = 1u << 24;

// Occupancy of the traversal kernels: the LDS stack (kLdsDepth x kBlock ints,
// 24 KiB per workgroup) admits 6 workgroups per CU, i.e. 6 waves per SIMD;
// left alone the register allocator uses 112-116 VGPRs (4 waves).  Asking for
// 6 waves caps them at 80 and spills the rest to scratch; traversal is
// latency-bound and the extra waves win (5.00 -> 4.63 ms per 1080p frame).
#define XR_TRAVERSAL_OCCUPANCY __attribute__((amdgpu_waves_per_eu(6)))

// Top-of-stack read.  Written as "sp < kLds ? lds[..] : spill[..]" the
// compiler merges the two loads into one flat load through a selected
// pointer, which waits on both the vector-memory and the LDS counters in
// every iteration; the volatile private read (levels past kLds only, rare)
// cannot be merged, so the common case stays a plain LDS read.
template <int kBlock, int kLds>
__device__ __forceinline__ int stack_top(const int* __restrict__ lds_col, const int* spill, int sp) {
    if (sp < kLds) return lds_col[sp * kBlock];
    return static_cast<const volatile int*>(spill)[sp - kLds];
}

// Triangles of leaf `code` against the ray (closest-hit rule above, or the
// first hit for kAnyHit); true when an any-hit ray is done.  The kept hit's
// barycentrics are held as V, W and det (u = V / det, v = W / det).
template <bool kAnyHit>
__device__ __forceinline__ bool leaf_hits(const BvhTri* __restrict__ tris, const RayPre& r, float tmin, int code,
                                          float* best_t, uint32_t* best_id, uint32_t* best_idx, float* best_V,
                                          float* best_W, float* best_det, Hit* hit) {
    const int c = ~code;
    const int first = c >> 4;
    const int cnt = (c & 15) + 1;
    for (int k = 0; k < cnt; ++k) {
        const BvhTri tr = tris[first + k];
        float t, V, W, det;
        if (intersect_tri(r, mk3(tr.v0.x, tr.v0.y, tr.v0.z), mk3(tr.v1.x, tr.v1.y, tr.v1.z),
                            mk3(tr.v2.x, tr.v2.y, tr.v2.z), tmin, *best_t, &t, &V, &W, &det)) {
            const uint32_t id = as_uint(tr.v0.w);
            if (kAnyHit) {
                hit->t = t; hit->u = V / det; hit->v = W / det; hit->id = id; hit->idx = first + k;
                return true;
            }
            // intersect_tri guarantees t <= best_t, so this is
            // "t < best_t, or t == best_t and a smaller id".
            if (t < *best_t || id < *best_id) {
                *best_t = t; *best_id = id; *best_idx = first + k; *best_V = V; *best_W = W; *best_det = det;
            }
        }
    }
    return false;
}

// Compare-exchange on (entry distance, link) pairs, nearer first.
__device__ __forceinline__ void order2(float& ka, int& va, float& kb, int& vb) {
    const bool s = kb < ka;
    const float k = s ? kb : ka;
    const int v = s ? vb : va;
    kb = s ? ka : kb;
    vb = s ? va : vb;
    ka = k;
    va = v;
}

// Four-element sorting network (5 compare-exchanges).
__device__ __forceinline__ void order4(float& k0, int& v0, float& k1, int& v1, float& k2, int& v2, float& k3,
                                       int& v3) {
    order2(k0, v0, k1, v1);
    order2(k2, v2, k3, v3);
    order2(k0, v0, k2, v2);
    order2(k1, v1, k3, v3);
    order2(k1, v1, k2, v2);
}

// The four slab tests of a wide node (unused slots hold empty boxes).
// On axis a the distance to grid plane g is o_a * (1/d) + g * (step / d),
// one fma from two terms formed once per node, o_a * (1/d) itself one fma
// from the ray's origin * (1/d) (step is a power of two, so step / d is
// exact).  The roundings move a plane by a few ulp of the scene's
// coordinates, far less than the padding the builders add to every box
// (at least extent * 2^-16), so no child box that holds a hit is culled; the
// far distance keeps the 1 + 2*gamma(3) scaling of the plain slab test.
__device__ __forceinline__ void wide_slabs(const WideNode* __restrict__ nd, const RayPre& r, float tmin, float tb,
                                           int4* lk, bool h[4], float e[4]) {
    const float4 q = nd->q;
    const uint4 p = nd->p;
    const uint4 pz = nd->r;
    *lk = nd->link;
    const float ax = fmaf(q.x, r.idir.x, -r.oi.x), bx = q.w * r.idir.x;
    const float ay = fmaf(q.y, r.idir.y, -r.oi.y), by = __uint_as_float(pz.z) * r.idir.y;
    const float az = fmaf(q.z, r.idir.z, -r.oi.z), bz = __uint_as_float(pz.w) * r.idir.z;
    // A plane's distance fma(g, b, a) is monotonic in g, increasing where the
    // reciprocal direction is positive (it is never zero or NaN, see
    // make_ray_pre): the near plane of every child is its low plane on such
    // an axis and its high plane otherwise, which is what min/max of the two
    // distances would pick, with 6 selects per node instead of 6 min/max per
    // child.
    const bool px = r.idir.x >= 0.0f, py = r.idir.y >= 0.0f, pzp = r.idir.z >= 0.0f;
    const uint32_t nx = px ? p.x : p.y, fx = px ? p.y : p.x;
    const uint32_t ny = py ? p.z : p.w, fy = py ? p.w : p.z;
    const uint32_t nz = pzp ? pz.x : pz.y, fz = pzp ? pz.y : pz.x;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int sh = 8 * k;
        const float t0x = fmaf((float)((nx >> sh) & 0xFFu), bx, ax), t1x = fmaf((float)((fx >> sh) & 0xFFu), bx, ax);
        const float t0y = fmaf((float)((ny >> sh) & 0xFFu), by, ay), t1y = fmaf((float)((fy >> sh) & 0xFFu), by, ay);
        const float t0z = fmaf((float)((nz >> sh) & 0xFFu), bz, az), t1z = fmaf((float)((fz >> sh) & 0xFFu), bz, az);
        const float tn = fmaxf(fmaxf(t0x, t0y), fmaxf(t0z, tmin));
        float tf = fminf(fminf(t1x, t1y), t1z);
        tf = fminf(tf * 1.0000003576f, tb);
        h[k] = tn <= tf;
        e[k] = tn;
    }
    // The links are needed only on some paths of the callers; without a use
    // here the compiler sinks their load behind the slab tests, and every
    // node visit waits for memory twice.
    asm volatile("" ::"v"(lk->x), "v"(lk->y), "v"(lk->z), "v"(lk->w));
}

// Traversal for any-hit rays (trace_ray below picks the loop per kind of
// ray).  kLds: stack levels kept in LDS (the rest, up to kMaxTraversalDepth,
// in a private array); the batched kernels pick it by measurement.
//
// Loop structure: "while-while" with postponed leaves and speculative
// traversal (Aila and Laine, "Understanding the Efficiency of Ray Traversal on
// GPUs", HPG 2009).  A lane that reaches a leaf parks it and keeps descending
// inner nodes until half of the wave's lanes hold a leaf (or have run out of
// nodes); then the wave intersects triangles together (waiting for every
// lane: 1.40 ms for the 1080p spatial pass, three quarters 1.40, a quarter
// 1.375, half 1.36).  Without it a wave
// runs the triangle loop in almost every iteration, for the few lanes that
// sit on a leaf, while the others wait.  Leaf codes travel on the stack like
// node indices: a node's hit children are taken in slot order, the first next
// and the others pushed (sorting them by entry distance, as the closest-hit
// loop does, costs more than it saves on visibility rays: 1.57 against
// 1.40 ms for the 1080p spatial pass).  The
// closest-hit rule is order independent and an any-hit ray reports only
// whether something blocks it, so the results do not depend on the visiting
// order.  The tree's stack bound (the largest sum of "children - 1" along a
// root-to-leaf path, checked when the tree is built) keeps sp within the
// stack.
template <int kBlock, bool kAnyHit, int kLds>
__device__ __forceinline__ bool trace_ray_ww(const WideNode* __restrict__ nodes, const BvhTri* __restrict__ tris,
                                             const RayPre& r, float tmin, float tmax, int* __restrict__ lds_col,
                                             Hit* hit) {
    // The bottom stack entry is a sentinel: popping it ends the ray.
    constexpr int kDone = 0x7FFFFFFF;
    int spill[kMaxTraversalDepth + 1 - kLds];
    lds_col[0] = kDone;
    int sp = 1;
    auto push = [&](int v) {
        if (sp < kLds) lds_col[sp * kBlock] = v;
        else spill[sp - kLds] = v;
        ++sp;
    };
    int node = 0;  // >= 0: inner node, < 0: leaf code, kDone: finished
    int leaf = 0;  // parked leaf code, or 0
    float best_t = tmax;
    uint32_t best_id = 0xFFFFFFFFu;
    uint32_t best_idx = 0xFFFFFFFFu;
    float best_V = 0.0f, best_W = 0.0f, best_det = 1.0f;
    // A well-formed tree visits each node at most once per ray; the cap only
    // guarantees termination should a malformed tree ever reach the kernel.
    uint32_t guard = 0;
    // an any-hit ray returns at its first hit, so its far bound never changes
    const float tb_any = tmax * 1.0000003576f;
    while (node != kDone) {
        while (node >= 0 && node != kDone) {
            if (++guard >= kMaxNodeVisits) {
                node = kDone;
                leaf = 0;
                break;
            }
            int4 lk;
            bool h[4];
            float e[4];
            wide_slabs(nodes + node, r, tmin, kAnyHit ? tb_any : best_t * 1.0000003576f, &lk, h, e);
            int next = 0;
            bool have = false;
            if (h[3]) { next = lk.w; have = true; }
            if (h[2]) { if (have) push(next); next = lk.z; have = true; }
            if (h[1]) { if (have) push(next); next = lk.y; have = true; }
            if (h[0]) { if (have) push(next); next = lk.x; have = true; }
            if (!have) {
                --sp;
                next = stack_top<kBlock, kLds>(lds_col, spill, sp);
            }
            node = next;
            if (node < 0 && leaf == 0) {  // park the first leaf, keep descending
                leaf = node;
                --sp;
                node = stack_top<kBlock, kLds>(lds_col, spill, sp);
            }
            // half of the active lanes hold a leaf
            if (__popcll(__ballot(leaf != 0)) * 2 >= __popcll(__ballot(1))) break;
        }
        while (leaf < 0) {
            if (leaf_hits<kAnyHit>(tris, r, tmin, leaf, &best_t, &best_id, &best_idx, &best_V, &best_W, &best_det, hit))
                return true;
            leaf = 0;
            if (node < 0) {  // the current node is a leaf as well: take it
                leaf = node;
                --sp;
                node = stack_top<kBlock, kLds>(lds_col, spill, sp);
            }
        }
    }
    if (kAnyHit) return false;
    hit->t = best_idx == 0xFFFFFFFFu ? -1.0f : best_t;
    hit->u = best_V / best_det;
    hit->v = best_W / best_det;
    hit->id = best_id;
    hit->idx = best_idx;
    return best_idx != 0xFFFFFFFFu;
}
// Single loop: a lane intersects the hit leaf children of a node as soon as
// it tests the node (in slot order, in one triangle loop), then descends to
// the nearest hit inner child and pushes the other hit inner children,
// farthest deepest; inner children that a leaf hit has put beyond the new
// closest t are dropped.
template <int kBlock, bool kAnyHit, int kLds>
__device__ __forceinline__ bool trace_ray_inline(const WideNode* __restrict__ nodes, const BvhTri* __restrict__ tris,
                                                 const RayPre& r, float tmin, float tmax, int* __restrict__ lds_col,
                                                 Hit* hit) {
    // The bottom stack entry is a sentinel: popping it ends the ray.
    constexpr int kDone = 0x7FFFFFFF;
    constexpr float kInf = __builtin_huge_valf();
    int spill[kMaxTraversalDepth + 1 - kLds];
    lds_col[0] = kDone;
    int sp = 1;
    auto push = [&](int v) {
        if (sp < kLds) lds_col[sp * kBlock] = v;
        else spill[sp - kLds] = v;
        ++sp;
    };
    int node = 0;
    float best_t = tmax;
    uint32_t best_id = 0xFFFFFFFFu;
    uint32_t best_idx = 0xFFFFFFFFu;
    float best_V = 0.0f, best_W = 0.0f, best_det = 1.0f;
    // A well-formed tree visits each node at most once per ray; the cap only
    // guarantees termination should a malformed tree ever reach the kernel.
    uint32_t guard = 0;
    while (node != kDone && ++guard < kMaxNodeVisits) {
        int4 lk;
        bool h[4];
        float e[4];
        wide_slabs(nodes + node, r, tmin, best_t * 1.0000003576f, &lk, h, e);
        int lm = (h[0] && lk.x < 0 ? 1 : 0) | (h[1] && lk.y < 0 ? 2 : 0) | (h[2] && lk.z < 0 ? 4 : 0) |
                 (h[3] && lk.w < 0 ? 8 : 0);
        while (lm) {
            const int k = __builtin_ctz(lm);
            lm &= lm - 1;
            const int code = k == 0 ? lk.x : (k == 1 ? lk.y : (k == 2 ? lk.z : lk.w));
            if (leaf_hits<kAnyHit>(tris, r, tmin, code, &best_t, &best_id, &best_idx, &best_V, &best_W, &best_det, hit))
                return true;
        }
        // The slab test against the updated closest t, as the node's own test
        // would make it: no box holding a hit at t <= best_t is dropped.
        const float tc = best_t * 1.0000003576f;
        float k0 = (h[0] && lk.x >= 0 && e[0] <= tc) ? e[0] : kInf;
        float k1 = (h[1] && lk.y >= 0 && e[1] <= tc) ? e[1] : kInf;
        float k2 = (h[2] && lk.z >= 0 && e[2] <= tc) ? e[2] : kInf;
        float k3 = (h[3] && lk.w >= 0 && e[3] <= tc) ? e[3] : kInf;
        int c0 = lk.x, c1 = lk.y, c2 = lk.z, c3 = lk.w;
        order4(k0, c0, k1, c1, k2, c2, k3, c3);
        int next;
        if (k0 == kInf) {
            --sp;
            next = stack_top<kBlock, kLds>(lds_col, spill, sp);
        } else {
            next = c0;
            if (k3 != kInf) push(c3);
            if (k2 != kInf) push(c2);
            if (k1 != kInf) push(c1);
        }
        node = next;
    }
    if (kAnyHit) return false;
    hit->t = best_idx == 0xFFFFFFFFu ? -1.0f : best_t;
    hit->u = best_V / best_det;
    hit->v = best_W / best_det;
    hit->id = best_id;
    hit->idx = best_idx;
    return best_idx != 0xFFFFFFFFu;
}

// Measured on the 1080p terrain frame: the parked-leaf loop makes the
// spatial pass's any-hit visibility rays faster (2.27 -> 2.06 ms) and the
// initial pass's closest-hit rays slower (1.55 -> 1.71 ms, nodes that a
// leaf's hit would have culled get visited first), so each kind of ray takes
// the loop that suits it.
template <int kBlock, bool kAnyHit, int kLds = kLdsDepth>
__device__ __forceinline__ bool trace_ray(const WideNode* __restrict__ nodes, const BvhTri* __restrict__ tris,
                                          const RayPre& r, float tmin, float tmax, int* __restrict__ lds_col,
                                          Hit* hit) {
    if constexpr (kAnyHit) return trace_ray_ww<kBlock, true, kLds>(nodes, tris, r, tmin, tmax, lds_col, hit);
    else return trace_ray_inline<kBlock, false, kLds>(nodes, tris, r, tmin, tmax, lds_col, hit);
}

// Closest-hit attributes (Ray.rchit.hlsl:26-61): world position from
// barycentric interpolation, twice-normalized interpolated normal, geometric
// normal normalize(cross(v2-v0, v1-v0)), UV, material of v0.  Instances carry
// identity transforms (XVKTLAccelerationStructure.cpp:68-93), so the
// object-to-world products are exact copies and are omitted.
struct Payload {
    v3 position;
    v3 normal;
    v3 geo_normal;
    float u, v;
    int32_t material;
    float t;
};

__device__ __forceinline__ void load_vertex(const float* __restrict__ vb, uint32_t i, v3* p, v3* n, float* tu,
                                            float* tv, int32_t* mat) {
    const float* q = vb + (size_t)i * 9u;
    *p = mk3(q[0], q[1], q[2]);
    *n = mk3(q[3], q[4], q[5]);
    *tu = q[6];
    *tv = q[7];
    *mat = __float_as_int(q[8]);
}

__device__ __forceinline__ void closest_hit(const SceneDev& sc, const Hit& h, Payload* pl) {
    const BvhTri tr = sc.tris[h.idx];
    const uint32_t inst = as_uint(tr.v1.w);
    const uint32_t prim = as_uint(tr.v2.w);
    const uint2 off = sc.offsets[inst];
    const uint32_t base = off.x + prim * 3u;
    const uint32_t i0 = sc.indices[base + 0], i1 = sc.indices[base + 1], i2 = sc.indices[base + 2];
    v3 p0, p1, p2, n0, n1, n2;
    float u0, v0, u1, v1, u2, v2;
    int32_t m0, m1, m2;
    load_vertex(sc.vertices, off.y + i0, &p0, &n0, &u0, &v0, &m0);
    load_vertex(sc.vertices, off.y + i1, &p1, &n1, &u1, &v1, &m1);
    load_vertex(sc.vertices, off.y + i2, &p2, &n2, &u2, &v2, &m2);
    const float bx = 1.0f - h.u - h.v, by = h.u, bz = h.v;
    v3 nrm = add(add(muls(n0, bx), muls(n1, by)), muls(n2, bz));
    nrm = normalize(normalize(nrm));
    pl->position = add(add(muls(p0, bx), muls(p1, by)), muls(p2, bz));
    pl->normal = nrm;
    pl->u = u0 * bx + u1 * by + u2 * bz;
    pl->v = v0 * bx + v1 * by + v2 * bz;
    pl->geo_normal = normalize(cross(sub(p2, p0), sub(p1, p0)));
    pl->material = m0;
    pl->t = h.t;
}

// BranchlessONB (Sampling.hlsl:4-12; Duff et al. 2017).
XR_HD void onb(v3 n, v3* b1, v3* b2) {
    const float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    *b1 = mk3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    *b2 = mk3(b, sign + n.y * n.y * a, -n.y);
}

// TextureSamplers[t].SampleLevel(LinearSampler, uv, 0).rgb (BSDF.hlsl:22):
// bilinear, clamp to edge, R8G8B8A8_UNORM texels as byte / 255.  The weights
// and the sum order are this layer's definition (the oracle's is the same);
// the reference's sampler hardware rounds its weights to a fixed-point grid.
__device__ __forceinline__ float unorm8(uint32_t p, int shift) { return (float)((p >> shift) & 0xFFu) / 255.0f; }
__device__ __forceinline__ v3 tex_bilinear(const SceneDev& sc, int t, float u, float v) {
    const uint4 ti = sc.tex_info[t];
    const int w = (int)ti.y, h = (int)ti.z;
    // x clamped to [-1, w] first (NaN -> -1): beyond it both taps are the edge texel
    const float x = fminf(fmaxf(u * (float)w - 0.5f, -1.0f), (float)w);
    const float y = fminf(fmaxf(v * (float)h - 0.5f, -1.0f), (float)h);
    const float fx = floorf(x), fy = floorf(y);
    const float a = x - fx, b = y - fy;
    const int i0 = (int)fx, j0 = (int)fy;
    const int x0 = min(max(i0, 0), w - 1), x1 = min(max(i0 + 1, 0), w - 1);
    const int y0 = min(max(j0, 0), h - 1), y1 = min(max(j0 + 1, 0), h - 1);
    const uint32_t* tx = sc.texels + ti.x;
    const uint32_t p00 = tx[y0 * w + x0], p10 = tx[y0 * w + x1], p01 = tx[y1 * w + x0], p11 = tx[y1 * w + x1];
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    float c[3];
    for (int k = 0; k < 3; ++k) {
        const int sh = 8 * k;
        c[k] = ((w00 * unorm8(p00, sh) + w10 * unorm8(p10, sh)) + w01 * unorm8(p01, sh)) + w11 * unorm8(p11, sh);
    }
    return mk3(c[0], c[1], c[2]);
}

// Base colour of a material at texture coordinates (u, v): its texture when
// it names one, else baseColor.rgb (BSDF.hlsl:22-23); used for the Lambertian
// albedo only, since the bounce loops shade emitters untextured.
__device__ __forceinline__ v3 material_color(const SceneDev& sc, const float4 m0, const float4 m1, float u,
                                             float v) {
    const int t = __float_as_int(m1.x);
    return t >= 0 ? tex_bilinear(sc, t, u, v) : mk3(m0.x, m0.y, m0.z);
}

// SampleBSDF (BSDF.hlsl:34-57) with SampleLambertian (:7-25); the albedo is
// read only for a sample above the surface, as the reference reads it.
__device__ __forceinline__ v3 sample_bsdf(const SceneDev& sc, v3 n, const float4 mat0, const float4 mat1, float tu,
                                          float tv, v3* wi_world, float* pdf, float* cos_theta, uint32_t* seed) {
    v3 T, B;
    onb(n, &T, &B);
    v3 f = mk3(0.0f, 0.0f, 0.0f);
    v3 wl = mk3(0.0f, 0.0f, 0.0f);
    *pdf = 0.0f;
    *cos_theta = 0.0f;
    const uint32_t model = __float_as_uint(mat1.w);
    if (model == kLambertian) {
        const float xi0 = rand01(seed);
        const float xi1 = rand01(seed);
        const float phi = XR_TWO_PI * xi0;
        const float ct = sqrtf(1.0f - xi1);
        const float st = sqrtf(xi1);
        float s, c;
        sincos_xr(phi, &s, &c);
        wl = mk3(st * c, st * s, ct);
        *cos_theta = wl.z;
        *pdf = wl.z * XR_INV_PI;
        if (wl.z <= 0.0f) {
            *pdf = 0.0f;
        } else {
            const v3 al = material_color(sc, mat0, mat1, tu, tv);
            f = mk3(al.x * XR_INV_PI, al.y * XR_INV_PI, al.z * XR_INV_PI);
        }
    }
    *wi_world = mk3(wl.x * T.x + wl.y * B.x + wl.z * n.x, wl.x * T.y + wl.y * B.y + wl.z * n.y,
                    wl.x * T.z + wl.y * B.z + wl.z * n.z);
    return f;
}

}  // namespace xr

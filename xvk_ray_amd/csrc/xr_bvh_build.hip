// GPU BVH builders for gfx950: LBVH (Karras, "Maximizing Parallelism in the
// Construction of BVHs, Octrees, and k-d Trees", HPG 2012) and PLOC (Meister
// & Bittner, "Parallel Locally-Ordered Clustering for Bounding Volume
// Hierarchy Construction", IEEE TVCG 2018).
//
// Replaces vkCmdBuildAccelerationStructuresKHR of the reference
// (XVKBLAccelerationStructure.cpp:44-58, XVKTLAccelerationStructure.cpp:49-66),
// whose BVH arithmetic lives in the Vulkan driver.  All instances carry
// identity transforms in the reference, so the two-level structure collapses
// into one BVH over every triangle of every instance.
//
// Passes (one kernel each, all on the caller's stream):
//   1. gather   canonical triangles from the reference buffers, centroid and vertex bounds
//   2. morton   63-bit Morton code of each centroid
//   3. sort     (code, id) pairs, rocPRIM/hipCUB radix sort
//   4. karras   internal-node topology from the sorted codes
//   5. refit    bottom-up boxes (padded, see box_pad) and SAH costs, one launch per tree level
//   6. emit     64-byte two-box nodes and the BVH-ordered triangle array
// PLOC shares 1-3 and replaces 4-5 by bottom-up clustering of the sorted
// triangles (each cluster merges with its nearest neighbour by merged-box
// area within 16 positions when the choice is mutual), computing boxes, SAH
// costs and leaf collapses as nodes are created; a depth-first numbering of
// the leaves then makes every collapsed subtree a contiguous triangle range,
// and the linked nodes are stored in depth-first preorder (a node's left
// child follows it), with no storage for nodes inside collapsed subtrees.
#include <hipcub/hipcub.hpp>
#include <math.h>
#include <stdio.h>

#include <algorithm>
#include <atomic>
#include <new>
#include <chrono>
#include <thread>
#include <vector>

#include "xr_kernels.h"

namespace xr {

namespace {

__device__ __forceinline__ int ord_enc(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float ord_dec(int o) { return __int_as_float(o >= 0 ? o : o ^ 0x7FFFFFFF); }

struct Box {
    float lo[3], hi[3];
};

// bounds[0..2] / [3..5]: centroid min / max; bounds[6..8] / [9..11]: vertex min / max
// (order-preserving int encodings of the floats).  Grid-stride: each
// workgroup reduces its triangles' bounds in registers and shuffles and adds
// one atomic per value (at most kGatherBlocks workgroups contend per address).
constexpr uint32_t kGatherBlocks = 2048;
__global__ __launch_bounds__(256) void k_gather(const float* __restrict__ vtx, const uint32_t* __restrict__ idx,
                                                const uint2* __restrict__ offsets,
                                                const uint32_t* __restrict__ first_tri, uint32_t ninst,
                                                uint32_t ntri, BvhTri* __restrict__ ctris,
                                                float4* __restrict__ cent, int* __restrict__ bounds) {
    int v[12];
    for (int a = 0; a < 3; ++a) {
        v[a] = v[6 + a] = 0x7FFFFFFF;
        v[3 + a] = v[9 + a] = (int)0x80000000;
    }
    for (uint32_t g = blockIdx.x * 256u + threadIdx.x; g < ntri; g += gridDim.x * 256u) {
        // instance = last i with first_tri[i] <= g
        uint32_t lo = 0, hi = ninst;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (first_tri[mid] <= g) lo = mid; else hi = mid;
        }
        const uint32_t inst = lo;
        const uint32_t prim = g - first_tri[inst];
        const uint2 off = offsets[inst];
        float p[3][3];
        for (int k = 0; k < 3; ++k) {
            const uint32_t vi = off.y + idx[off.x + prim * 3u + k];
            const float* q = vtx + (size_t)vi * 9u;
            p[k][0] = q[0]; p[k][1] = q[1]; p[k][2] = q[2];
        }
        BvhTri t;
        t.v0 = make_float4(p[0][0], p[0][1], p[0][2], __uint_as_float(g));
        t.v1 = make_float4(p[1][0], p[1][1], p[1][2], __uint_as_float(inst));
        t.v2 = make_float4(p[2][0], p[2][1], p[2][2], __uint_as_float(prim));
        ctris[g] = t;
        float c[3];
        for (int a = 0; a < 3; ++a) {
            c[a] = (p[0][a] + p[1][a] + p[2][a]) * (1.0f / 3.0f);
            v[a] = min(v[a], ord_enc(c[a]));
            v[3 + a] = max(v[3 + a], ord_enc(c[a]));
            v[6 + a] = min(v[6 + a], ord_enc(fminf(fminf(p[0][a], p[1][a]), p[2][a])));
            v[9 + a] = max(v[9 + a], ord_enc(fmaxf(fmaxf(p[0][a], p[1][a]), p[2][a])));
        }
        cent[g] = make_float4(c[0], c[1], c[2], 0.f);
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int a = 0; a < 3; ++a) {
            v[a] = min(v[a], __shfl_xor(v[a], off));
            v[3 + a] = max(v[3 + a], __shfl_xor(v[3 + a], off));
            v[6 + a] = min(v[6 + a], __shfl_xor(v[6 + a], off));
            v[9 + a] = max(v[9 + a], __shfl_xor(v[9 + a], off));
        }
    __shared__ int part[4][12];
    if ((threadIdx.x & 63u) == 0u)
        for (int k = 0; k < 12; ++k) part[threadIdx.x >> 6][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 12u) {
        const int k = (int)threadIdx.x;
        const bool is_min = k < 3 || (k >= 6 && k < 9);
        int r = part[0][k];
        for (int w = 1; w < 4; ++w) r = is_min ? min(r, part[w][k]) : max(r, part[w][k]);
        if (is_min) atomicMin(&bounds[k], r);
        else atomicMax(&bounds[k], r);
    }
}

void launch_gather(const float* vtx, const uint32_t* idx, const uint2* offsets, const uint32_t* first_tri,
                   uint32_t ninst, uint32_t ntri, BvhTri* ctris, float4* cent, int* bounds, hipStream_t st) {
    const uint32_t blocks = std::max(1u, std::min((ntri + 255u) / 256u, kGatherBlocks));
    hipLaunchKernelGGL(k_gather, dim3(blocks), dim3(256), 0, st, vtx, idx, offsets, first_tri, ninst, ntri, ctris,
                       cent, bounds);
}

// Box padding.  The watertight FP32 test is consistent across shared edges
// but not exact: it may report a hit up to a few ulps of |vertex - origin|
// outside the triangle, i.e. outside its exact box.  Every box is therefore
// widened by pad = max(2^-16 * scene extent, 2^-20 * max |coordinate|), well
// above that error for ray origins within the scene, so no hit the triangle
// test would report is culled by a box.
__device__ __forceinline__ float box_pad(const int* __restrict__ bounds) {
    float ext = 0.0f, mag = 0.0f;
    for (int a = 0; a < 3; ++a) {
        const float lo = ord_dec(bounds[6 + a]), hi = ord_dec(bounds[9 + a]);
        ext = fmaxf(ext, hi - lo);
        mag = fmaxf(mag, fmaxf(fabsf(lo), fabsf(hi)));
    }
    return fmaxf(ext * (1.0f / 65536.0f), mag * (1.0f / 1048576.0f));
}

__device__ __forceinline__ unsigned long long spread21(uint32_t v) {
    unsigned long long x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x001F00000000FFFFull;
    x = (x | (x << 16)) & 0x001F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__global__ void k_morton(const float4* __restrict__ cent, uint32_t ntri, const int* __restrict__ bounds,
                         unsigned long long* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ntri) return;
    const float4 c = cent[g];
    float lo[3], ext[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = ord_dec(bounds[a]);
        const float hi = ord_dec(bounds[3 + a]);
        ext[a] = hi > lo[a] ? hi - lo[a] : 1.0f;
    }
    const float cc[3] = {c.x, c.y, c.z};
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
        float u = (cc[a] - lo[a]) / ext[a];
        u = fminf(fmaxf(u, 0.0f), 1.0f);
        q[a] = (uint32_t)fminf(u * 2097152.0f, 2097151.0f);
    }
    keys[g] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
    vals[g] = g;
}

__device__ __forceinline__ int delta(const unsigned long long* __restrict__ keys, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const unsigned long long a = keys[i], b = keys[j];
    if (a == b) return 64 + __clz((uint32_t)(i ^ j));
    return __clzll(a ^ b);
}

// Children encoding during the build: value < n-1 internal, else leaf (value - (n-1)).
// range[i] = the contiguous sorted-triangle range [lo, hi] below internal node i.
__global__ void k_karras(const unsigned long long* __restrict__ keys, int n, int2* __restrict__ child,
                         int* __restrict__ parent, int2* __restrict__ range) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0;
    for (int div = 2;; div *= 2) {
        const int t = (l + div - 1) / div;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    const int gamma = i + s * d + (d < 0 ? -1 : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const int left = (lo == gamma) ? (n - 1) + gamma : gamma;
    const int right = (hi == gamma + 1) ? (n - 1) + gamma + 1 : gamma + 1;
    child[i] = make_int2(left, right);
    range[i] = make_int2(lo, hi);
    parent[left] = i;
    parent[right] = i;
}

// Surface-area-heuristic constants of the leaf collapse: relative cost of
// visiting a node (two slab tests, one 64-byte fetch) and of one triangle
// test (one 48-byte fetch, the watertight test, run by few lanes of a wave at
// a time).  Measured: a triangle cost of 3 against 1 took the 1080p terrain
// frame from 3.42 to 3.35 ms (LBVH) and plain tracing of the sb2 scene from
// 10.8 to 9.6 ms (LBVH) and 7.6 to 7.2 ms (PLOC); 1.5 and 2 fall in between,
// 0.7 is slower, and a 4-triangle leaf cap changes nothing measurable.
constexpr float kSahNode = 1.0f;
constexpr float kSahTri = 3.0f;
constexpr int kMaxLeafTris = 8;

__device__ __forceinline__ float box_area(const Box& b) {
    const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// Bottom-up boxes and SAH costs, one depth level of the tree per launch
// (deepest first), so every child is complete before its parent reads it and
// no lane waits on another.  Node ids: internal 0..n-2, leaf j -> n-1+j.
// cost(leaf of k tris) = kSahTri*k*A; cost(split) = kSahNode*A + cost(left) +
// cost(right); a non-root node whose subtree holds at most kMaxLeafTris
// triangles becomes one leaf when that is cheaper (collapse[p]).
__global__ void k_leaf_boxes(const BvhTri* __restrict__ ctris, const uint32_t* __restrict__ sorted_ids, int n,
                             const int* __restrict__ bounds, Box* __restrict__ boxes, float* __restrict__ cost) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const BvhTri t = ctris[sorted_ids[j]];
    const float pad = box_pad(bounds);
    Box b;
    b.lo[0] = fminf(fminf(t.v0.x, t.v1.x), t.v2.x) - pad;
    b.lo[1] = fminf(fminf(t.v0.y, t.v1.y), t.v2.y) - pad;
    b.lo[2] = fminf(fminf(t.v0.z, t.v1.z), t.v2.z) - pad;
    b.hi[0] = fmaxf(fmaxf(t.v0.x, t.v1.x), t.v2.x) + pad;
    b.hi[1] = fmaxf(fmaxf(t.v0.y, t.v1.y), t.v2.y) + pad;
    b.hi[2] = fmaxf(fmaxf(t.v0.z, t.v1.z), t.v2.z) + pad;
    const int node = (n - 1) + j;
    boxes[node] = b;
    cost[node] = kSahTri * box_area(b);
}

// Depth (edges from the root) of every internal node, the sort key of the
// level order; a walk longer than kMaxLbvhDepth marks the build as failed.
constexpr int kMaxLbvhDepth = 255;
__global__ void k_node_depth(const int* __restrict__ parent, int ninner, uint32_t* __restrict__ depth,
                             uint32_t* __restrict__ ids, uint32_t* __restrict__ too_deep) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ninner) return;
    int p = i, d = 0;
    while (p != 0 && d <= kMaxLbvhDepth) {
        p = parent[p];
        ++d;
    }
    if (d > kMaxLbvhDepth) *too_deep = 1u;
    depth[i] = (uint32_t)min(d, kMaxLbvhDepth);
    ids[i] = (uint32_t)i;
}

// level_start[d] = first position of depth d in the sorted keys (levels are
// never empty between 0 and the deepest).
__global__ void k_level_starts(const uint32_t* __restrict__ sorted_depth, int ninner,
                               uint32_t* __restrict__ level_start) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ninner) return;
    const uint32_t d = sorted_depth[i];
    if (i == 0 || sorted_depth[i - 1] != d) level_start[d] = (uint32_t)i;
    if (i == ninner - 1) level_start[kMaxLbvhDepth + 1] = d;  // the deepest level
}

// One level: every node's children are leaves or nodes one level deeper.
// height[p] counts edges down to the deepest leaf of the emitted tree (a
// collapsed node is a leaf); the root's height is the tree's depth.
__global__ void k_refit_level(const uint32_t* __restrict__ level_ids, int count, int n,
                              const int2* __restrict__ child, const int2* __restrict__ range, Box* __restrict__ boxes,
                              float* __restrict__ cost, int* __restrict__ collapse, int* __restrict__ height) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int p = (int)level_ids[k];
    const int leaf_base = n - 1;
    const int2 c = child[p];
    const Box bx = boxes[c.x], by = boxes[c.y];
    Box m;
    for (int a = 0; a < 3; ++a) {
        m.lo[a] = fminf(bx.lo[a], by.lo[a]);
        m.hi[a] = fmaxf(bx.hi[a], by.hi[a]);
    }
    boxes[p] = m;
    const float area = box_area(m);
    const float split = kSahNode * area + cost[c.x] + cost[c.y];
    const int2 r = range[p];
    const int cnt = r.y - r.x + 1;
    const float as_leaf = kSahTri * (float)cnt * area;
    const bool leaf = p != 0 && cnt <= kMaxLeafTris && as_leaf <= split;
    collapse[p] = leaf ? 1 : 0;
    cost[p] = leaf ? as_leaf : split;
    const int hx = c.x >= leaf_base ? 0 : height[c.x];
    const int hy = c.y >= leaf_base ? 0 : height[c.y];
    height[p] = leaf ? 0 : 1 + max(hx, hy);
}

// Internal nodes below a collapsed node are emitted too but never linked.
__global__ void k_emit(int n, const int2* __restrict__ child, const Box* __restrict__ boxes,
                       BvhNode* __restrict__ nodes, const BvhTri* __restrict__ ctris,
                       const uint32_t* __restrict__ sorted_ids, BvhTri* __restrict__ tris,
                       const int2* __restrict__ range, const int* __restrict__ collapse) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) tris[i] = ctris[sorted_ids[i]];
    const int ninner = n > 1 ? n - 1 : 1;
    if (i >= ninner) return;
    int c0, c1;
    if (n == 1) {
        c0 = c1 = 0;  // leaf node id 0 -> both links name leaf 0
    } else {
        c0 = child[i].x;
        c1 = child[i].y;
    }
    const int leaf_base = n > 1 ? n - 1 : 0;
    const Box b0 = boxes[n > 1 ? c0 : 0];
    const Box b1 = boxes[n > 1 ? c1 : 0];
    BvhNode nd;
    nd.a = make_float4(b0.lo[0], b0.hi[0], b0.lo[1], b0.hi[1]);
    nd.b = make_float4(b1.lo[0], b1.hi[0], b1.lo[1], b1.hi[1]);
    nd.c = make_float4(b0.lo[2], b0.hi[2], b1.lo[2], b1.hi[2]);
    auto enc = [&](int c) -> int {
        if (n == 1) return ~0;
        if (c >= leaf_base) return ~((c - leaf_base) * 16);
        if (collapse[c]) {
            const int2 r = range[c];
            return ~(r.x * 16 + (r.y - r.x));
        }
        return c;
    };
    nd.d = make_int4(enc(c0), enc(c1), 0, 0);
    nodes[i] = nd;
}

// ---------------------------------------------------------------------------
// PLOC
// ---------------------------------------------------------------------------
// Search radius 16 (8: slower trees; 32 and 64: slower on the terrain, no
// better on sb2).
constexpr int kPlocRadius = 16;

__device__ __forceinline__ Box box_union(const Box& a, const Box& b) {
    Box m;
    for (int k = 0; k < 3; ++k) {
        m.lo[k] = fminf(a.lo[k], b.lo[k]);
        m.hi[k] = fmaxf(a.hi[k], b.hi[k]);
    }
    return m;
}

// Node ids: internal 0..n-2 in creation order (the root is n-2), leaf of
// sorted position j -> (n-1) + j.
__global__ void k_ploc_leaves(const BvhTri* __restrict__ ctris, const uint32_t* __restrict__ sorted_ids, int n,
                              const int* __restrict__ bounds, Box* __restrict__ boxes, float* __restrict__ cost,
                              int* __restrict__ count, int* __restrict__ linked, int* __restrict__ clusters) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const BvhTri t = ctris[sorted_ids[j]];
    const float pad = box_pad(bounds);
    Box b;
    b.lo[0] = fminf(fminf(t.v0.x, t.v1.x), t.v2.x) - pad;
    b.lo[1] = fminf(fminf(t.v0.y, t.v1.y), t.v2.y) - pad;
    b.lo[2] = fminf(fminf(t.v0.z, t.v1.z), t.v2.z) - pad;
    b.hi[0] = fmaxf(fmaxf(t.v0.x, t.v1.x), t.v2.x) + pad;
    b.hi[1] = fmaxf(fmaxf(t.v0.y, t.v1.y), t.v2.y) + pad;
    b.hi[2] = fmaxf(fmaxf(t.v0.z, t.v1.z), t.v2.z) + pad;
    const int id = (n - 1) + j;
    boxes[id] = b;
    cost[id] = kSahTri * box_area(b);
    count[id] = 1;
    linked[id] = 0;
    clusters[j] = id;
}

// Nearest neighbour of each cluster among the 2 * kPlocRadius around it, by
// the area of the merged box; among equal areas the position's buddy
// (positions 2k and 2k + 1 are buddies) comes first, then the lower position.
// That orders candidates exactly as the pairs (area, not buddies, lower,
// higher) are ordered, a strict order that does not depend on which side
// looks, so the globally best pair is always mutual and every iteration
// merges.  The buddy rule matters for runs of identical boxes (coincident
// triangles): with ties going to the lowest position alone, only the run's
// first pair was mutual, one pair merged per iteration and the run became a
// chain as deep as it was long; with it the run pairs up completely.
__global__ void k_ploc_nn(const int* __restrict__ clusters, int C, const Box* __restrict__ boxes,
                          int* __restrict__ nn) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const Box bi = boxes[clusters[i]];
    float best_d = __builtin_huge_valf();
    int best = -1;
    const int lo = i - kPlocRadius < 0 ? 0 : i - kPlocRadius;
    const int hi = i + kPlocRadius > C - 1 ? C - 1 : i + kPlocRadius;
    for (int j = lo; j <= hi; ++j) {
        if (j == i) continue;
        const float d = box_area(box_union(bi, boxes[clusters[j]]));
        if (d < best_d || (d == best_d && (j ^ 1) == i)) {
            best_d = d;
            best = j;
        }
    }
    nn[i] = best;
}

// flags[i]: high word 1 when cluster i creates a node (mutual pair, i the
// lower), low word 1 when position i survives (it is not the higher of a
// mutual pair).
__global__ void k_ploc_flags(const int* __restrict__ nn, int C, unsigned long long* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const int j = nn[i];
    const bool mutual = j >= 0 && nn[j] == i;
    const unsigned long long make = (mutual && i < j) ? 1ull : 0ull;
    const unsigned long long keep = (mutual && i > j) ? 0ull : 1ull;
    flags[i] = (make << 32) | keep;
}

__global__ void k_ploc_merge(int n, const int* __restrict__ clusters, const int* __restrict__ nn, int C,
                             const unsigned long long* __restrict__ flags, const unsigned long long* __restrict__ offs,
                             int node_base, Box* __restrict__ boxes, float* __restrict__ cost,
                             int* __restrict__ count, int* __restrict__ linked, int2* __restrict__ child,
                             int* __restrict__ parent, int* __restrict__ collapse, int* __restrict__ next) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const unsigned long long f = flags[i];
    if (!(f & 1ull)) return;
    const unsigned long long o = offs[i];
    const int pos = (int)(o & 0xFFFFFFFFull);
    if (!(f >> 32)) {
        next[pos] = clusters[i];
        return;
    }
    const int a = clusters[i], b = clusters[nn[i]];
    const int id = node_base + (int)(o >> 32);
    const Box m = box_union(boxes[a], boxes[b]);
    boxes[id] = m;
    const float area = box_area(m);
    const int cnt = count[a] + count[b];
    const float split = kSahNode * area + cost[a] + cost[b];
    const float as_leaf = kSahTri * (float)cnt * area;
    const bool leaf = cnt <= kMaxLeafTris && as_leaf <= split;
    collapse[id] = leaf ? 1 : 0;
    cost[id] = leaf ? as_leaf : split;
    count[id] = cnt;
    // linked[x]: nodes stored for the subtree of x when x itself is stored;
    // leaves and collapsed nodes are encoded in their parent and store none
    auto stored = [&](int x) { return (x >= n - 1 || collapse[x]) ? 0 : linked[x]; };
    linked[id] = 1 + stored(a) + stored(b);
    child[id] = make_int2(a, b);
    parent[a] = id;
    parent[b] = id;
    next[pos] = id;
}

// Depth-first position of each leaf (left subtree first) and its depth.
__global__ void k_ploc_leaf_order(int n, const int* __restrict__ parent, const int2* __restrict__ child,
                                  const int* __restrict__ count, int* __restrict__ leaf_off,
                                  uint32_t* __restrict__ max_depth) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int root = n - 2;
    int node = (n - 1) + j, off = 0;
    uint32_t depth = 0;
    while (node != root && depth < (uint32_t)n) {
        const int p = parent[node];
        const int2 c = child[p];
        if (c.y == node) off += count[c.x];
        node = p;
        ++depth;
    }
    leaf_off[j] = off;
    atomicMax(max_depth, depth);
}

// First depth-first triangle position below each internal node (its leftmost leaf).
__global__ void k_ploc_range(int n, const int2* __restrict__ child, const int* __restrict__ leaf_off,
                             int* __restrict__ first) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int node = i;
    for (int steps = 0; node < n - 1 && steps < n; ++steps) node = child[node].x;
    first[i] = leaf_off[node - (n - 1)];
}

// Depth-first preorder position of each stored internal node (-1 for nodes
// inside a collapsed subtree): the left child of a stored node follows it,
// the right child follows the left child's stored subtree.  The root is 0.
__global__ void k_ploc_preorder(int n, const int* __restrict__ parent, const int2* __restrict__ child,
                                const int* __restrict__ collapse, const int* __restrict__ linked,
                                int* __restrict__ pre) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int root = n - 2;
    if (i != root && collapse[i]) {
        pre[i] = -1;
        return;
    }
    int node = i, pos = 0;
    for (int steps = 0; node != root && steps < n; ++steps) {
        const int p = parent[node];
        if (p != root && collapse[p]) {
            pre[i] = -1;
            return;
        }
        const int2 c = child[p];
        if (c.x == node) pos += 1;
        else pos += 1 + ((c.x >= n - 1 || collapse[c.x]) ? 0 : linked[c.x]);
        node = p;
    }
    pre[i] = pos;
}

// Stored nodes at their preorder positions; triangles at their depth-first
// positions.
__global__ void k_ploc_emit(int n, const int2* __restrict__ child, const Box* __restrict__ boxes,
                            const int* __restrict__ count, const int* __restrict__ collapse,
                            const int* __restrict__ leaf_off, const int* __restrict__ first,
                            const int* __restrict__ pre, const BvhTri* __restrict__ ctris,
                            const uint32_t* __restrict__ sorted_ids, BvhNode* __restrict__ nodes,
                            BvhTri* __restrict__ tris) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) tris[leaf_off[i]] = ctris[sorted_ids[i]];
    if (i >= n - 1 || pre[i] < 0) return;
    const int2 c = child[i];
    const Box b0 = boxes[c.x], b1 = boxes[c.y];
    auto enc = [&](int cid) -> int {
        if (cid >= n - 1) return ~(leaf_off[cid - (n - 1)] * 16);
        if (collapse[cid]) return ~(first[cid] * 16 + (count[cid] - 1));
        return pre[cid];
    };
    BvhNode nd;
    nd.a = make_float4(b0.lo[0], b0.hi[0], b0.lo[1], b0.hi[1]);
    nd.b = make_float4(b1.lo[0], b1.hi[0], b1.lo[1], b1.hi[1]);
    nd.c = make_float4(b0.lo[2], b0.hi[2], b1.lo[2], b1.hi[2]);
    nd.d = make_int4(enc(c.x), enc(c.y), 0, 0);
    nodes[pre[i]] = nd;
}

}  // namespace

#define XR_TRY(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            snprintf(err, err_len, "%s failed: %s", #expr, hipGetErrorString(e_));           \
            goto fail;                                                                       \
        }                                                                                    \
    } while (0)

hipError_t build_lbvh(const float* vertices, const uint32_t* indices, const uint2* offsets,
                      const uint32_t* inst_first_tri, uint32_t instance_count, uint32_t tri_count,
                      hipStream_t st, BvhBuildResult* out, char* err, size_t err_len) {
    hipError_t rc = hipSuccess;
    const int n = (int)tri_count;
    const uint32_t ninner = n > 1 ? (uint32_t)(n - 1) : 1u;
    BvhTri* ctris = nullptr;
    float4* cent = nullptr;
    int* bounds = nullptr;
    unsigned long long *keys = nullptr, *keys2 = nullptr;
    uint32_t *vals = nullptr, *vals2 = nullptr;
    int2* child = nullptr;
    int* parent = nullptr;
    Box* boxes = nullptr;
    int2* range = nullptr;
    float* cost = nullptr;
    int* collapse = nullptr;
    int* height = nullptr;
    uint32_t* lvl = nullptr;  // level starts [0, kMaxLbvhDepth], deepest level, too-deep flag
    std::vector<uint32_t> hl(kMaxLbvhDepth + 3, 0u);
    void* temp = nullptr;
    size_t temp_bytes = 0, level_sort_bytes = 0;
    BvhNode* nodes = nullptr;
    BvhTri* tris = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const int init_bounds[12] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF, (int)0x80000000, (int)0x80000000,
                                 (int)0x80000000, 0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF, (int)0x80000000,
                                 (int)0x80000000, (int)0x80000000};
    const uint32_t blocks = (tri_count + 255) / 256;
    uint32_t host_depth = 0;
    float ms = 0.f;
    if (tri_count == 0) {
        snprintf(err, err_len, "scene has no triangles");
        return hipErrorInvalidValue;
    }
    XR_TRY(hipEventCreate(&e0));
    XR_TRY(hipEventCreate(&e1));
    XR_TRY(hipEventRecord(e0, st));
    XR_TRY(hipMalloc(&ctris, sizeof(BvhTri) * tri_count));
    XR_TRY(hipMalloc(&cent, sizeof(float4) * tri_count));
    XR_TRY(hipMalloc(&bounds, sizeof(int) * 12));
    XR_TRY(hipMalloc(&keys, sizeof(unsigned long long) * tri_count));
    XR_TRY(hipMalloc(&keys2, sizeof(unsigned long long) * tri_count));
    XR_TRY(hipMalloc(&vals, sizeof(uint32_t) * tri_count));
    XR_TRY(hipMalloc(&vals2, sizeof(uint32_t) * tri_count));
    XR_TRY(hipMalloc(&child, sizeof(int2) * ninner));
    XR_TRY(hipMalloc(&parent, sizeof(int) * (ninner + tri_count)));
    XR_TRY(hipMalloc(&boxes, sizeof(Box) * (ninner + tri_count)));
    XR_TRY(hipMalloc(&range, sizeof(int2) * ninner));
    XR_TRY(hipMalloc(&cost, sizeof(float) * (ninner + tri_count)));
    XR_TRY(hipMalloc(&collapse, sizeof(int) * ninner));
    XR_TRY(hipMalloc(&height, sizeof(int) * ninner));
    XR_TRY(hipMemsetAsync(collapse, 0, sizeof(int) * ninner, st));
    XR_TRY(hipMalloc(&lvl, sizeof(uint32_t) * (kMaxLbvhDepth + 3)));
    XR_TRY(hipMalloc(&nodes, sizeof(BvhNode) * ninner));
    XR_TRY(hipMalloc(&tris, sizeof(BvhTri) * tri_count));
    XR_TRY(hipMemcpyAsync(bounds, init_bounds, sizeof(init_bounds), hipMemcpyHostToDevice, st));
    XR_TRY(hipMemsetAsync(lvl, 0, sizeof(uint32_t) * (kMaxLbvhDepth + 3), st));
    XR_TRY(hipMemsetAsync(parent, 0, sizeof(int) * (ninner + tri_count), st));

    launch_gather(vertices, indices, offsets, inst_first_tri, instance_count, tri_count, ctris, cent, bounds, st);
    XR_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_morton, dim3(blocks), dim3(256), 0, st, cent, tri_count, bounds, keys, vals);
    XR_TRY(hipGetLastError());
    XR_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys, keys2, vals, vals2, n, 0, 63, st));
    XR_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, level_sort_bytes, (uint32_t*)keys, (uint32_t*)keys2, vals,
                                              vals, (int)ninner, 0, 8, st));
    temp_bytes = std::max(temp_bytes, level_sort_bytes);
    XR_TRY(hipMalloc(&temp, temp_bytes));
    XR_TRY(hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys, keys2, vals, vals2, n, 0, 63, st));
    if (n > 1) {
        hipLaunchKernelGGL(k_karras, dim3((n - 1 + 255) / 256), dim3(256), 0, st, keys2, n, child, parent, range);
        XR_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_leaf_boxes, dim3(blocks), dim3(256), 0, st, ctris, vals2, n, bounds, boxes, cost);
    XR_TRY(hipGetLastError());
    if (n > 1) {
        // the Morton keys are no longer needed: depth keys and node ids in
        // their place, sorted into keys2 (keys first, ids after them)
        uint32_t* dkey = (uint32_t*)keys;
        uint32_t* dkey2 = (uint32_t*)keys2;
        uint32_t* nid = vals;
        uint32_t* nid2 = (uint32_t*)keys2 + tri_count;
        const uint32_t ib = (ninner + 255) / 256;
        hipLaunchKernelGGL(k_node_depth, dim3(ib), dim3(256), 0, st, parent, (int)ninner, dkey, nid,
                           lvl + kMaxLbvhDepth + 2);
        XR_TRY(hipGetLastError());
        XR_TRY(hipcub::DeviceRadixSort::SortPairs(temp, level_sort_bytes, dkey, dkey2, nid, nid2, (int)ninner, 0, 8,
                                                  st));
        hipLaunchKernelGGL(k_level_starts, dim3(ib), dim3(256), 0, st, dkey2, (int)ninner, lvl);
        XR_TRY(hipGetLastError());
        XR_TRY(hipMemcpyAsync(hl.data(), lvl, sizeof(uint32_t) * hl.size(), hipMemcpyDeviceToHost, st));
        XR_TRY(hipStreamSynchronize(st));
        if (hl[kMaxLbvhDepth + 2]) {
            snprintf(err, err_len, "LBVH deeper than %d levels", kMaxLbvhDepth);
            goto fail;
        }
        const int deepest = (int)hl[kMaxLbvhDepth + 1];
        for (int d = deepest; d >= 0; --d) {
            const uint32_t lb = hl[d], le = d == deepest ? ninner : hl[d + 1];
            hipLaunchKernelGGL(k_refit_level, dim3((le - lb + 255) / 256), dim3(256), 0, st, nid2 + lb,
                               (int)(le - lb), n, child, range, boxes, cost, collapse, height);
        }
        XR_TRY(hipGetLastError());
        XR_TRY(hipMemcpyAsync(&host_depth, height, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    }
    hipLaunchKernelGGL(k_emit, dim3((tri_count + 255) / 256), dim3(256), 0, st, n, child, boxes, nodes, ctris,
                       vals2, tris, range, collapse);
    XR_TRY(hipGetLastError());
    XR_TRY(hipEventRecord(e1, st));
    XR_TRY(hipStreamSynchronize(st));
    XR_TRY(hipEventElapsedTime(&ms, e0, e1));

    out->nodes = nodes;
    out->tris = tris;
    out->node_count = ninner;
    out->tri_count = tri_count;
    out->max_depth = host_depth;
    out->build_ms = ms;
    nodes = nullptr;
    tris = nullptr;
    rc = hipSuccess;
    goto done;
fail:
    rc = hipErrorUnknown;
done:
    hipFree(ctris);
    hipFree(cent);
    hipFree(bounds);
    hipFree(keys);
    hipFree(keys2);
    hipFree(vals);
    hipFree(vals2);
    hipFree(child);
    hipFree(parent);
    hipFree(boxes);
    hipFree(range);
    hipFree(cost);
    hipFree(collapse);
    hipFree(height);
    hipFree(lvl);
    hipFree(temp);
    hipFree(nodes);
    hipFree(tris);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    return rc;
}

hipError_t build_ploc(const float* vertices, const uint32_t* indices, const uint2* offsets,
                      const uint32_t* inst_first_tri, uint32_t instance_count, uint32_t tri_count,
                      hipStream_t st, BvhBuildResult* out, char* err, size_t err_len) {
    if (tri_count < 2)  // one triangle: the single-node tree of the LBVH path
        return build_lbvh(vertices, indices, offsets, inst_first_tri, instance_count, tri_count, st, out, err,
                          err_len);
    hipError_t rc = hipSuccess;
    const int n = (int)tri_count;
    const size_t nall = 2 * (size_t)n - 1;
    BvhTri* ctris = nullptr;
    float4* cent = nullptr;
    int* bounds = nullptr;
    unsigned long long *keys = nullptr, *keys2 = nullptr;
    uint32_t *vals = nullptr, *vals2 = nullptr;
    Box* boxes = nullptr;
    float* cost = nullptr;
    int* count = nullptr;
    int* linked = nullptr;
    int* pre = nullptr;
    int2* child = nullptr;
    int* parent = nullptr;
    int* collapse = nullptr;
    int *clusters = nullptr, *next = nullptr, *nn = nullptr;
    int stored = 0;
    unsigned long long *flags = nullptr, *offs = nullptr;
    int *leaf_off = nullptr, *first = nullptr;
    uint32_t* dmax = nullptr;
    void* temp = nullptr;
    size_t temp_bytes = 0, scan_bytes = 0;
    BvhNode* nodes = nullptr;
    BvhTri* tris = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const int init_bounds[12] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF, (int)0x80000000, (int)0x80000000,
                                 (int)0x80000000, 0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF, (int)0x80000000,
                                 (int)0x80000000, (int)0x80000000};
    const uint32_t blocks = (tri_count + 255) / 256;
    uint32_t host_depth = 0;
    float ms = 0.f;
    int C = n, node_base = 0, iterations = 0;
    XR_TRY(hipEventCreate(&e0));
    XR_TRY(hipEventCreate(&e1));
    XR_TRY(hipEventRecord(e0, st));
    XR_TRY(hipMalloc(&ctris, sizeof(BvhTri) * tri_count));
    XR_TRY(hipMalloc(&cent, sizeof(float4) * tri_count));
    XR_TRY(hipMalloc(&bounds, sizeof(int) * 12));
    XR_TRY(hipMalloc(&keys, sizeof(unsigned long long) * tri_count));
    XR_TRY(hipMalloc(&keys2, sizeof(unsigned long long) * tri_count));
    XR_TRY(hipMalloc(&vals, sizeof(uint32_t) * tri_count));
    XR_TRY(hipMalloc(&vals2, sizeof(uint32_t) * tri_count));
    XR_TRY(hipMalloc(&boxes, sizeof(Box) * nall));
    XR_TRY(hipMalloc(&cost, sizeof(float) * nall));
    XR_TRY(hipMalloc(&count, sizeof(int) * nall));
    XR_TRY(hipMalloc(&linked, sizeof(int) * nall));
    XR_TRY(hipMalloc(&pre, sizeof(int) * (n - 1)));
    XR_TRY(hipMalloc(&child, sizeof(int2) * (n - 1)));
    XR_TRY(hipMalloc(&parent, sizeof(int) * nall));
    XR_TRY(hipMalloc(&collapse, sizeof(int) * (n - 1)));
    XR_TRY(hipMalloc(&clusters, sizeof(int) * n));
    XR_TRY(hipMalloc(&next, sizeof(int) * n));
    XR_TRY(hipMalloc(&nn, sizeof(int) * n));
    XR_TRY(hipMalloc(&flags, sizeof(unsigned long long) * n));
    XR_TRY(hipMalloc(&offs, sizeof(unsigned long long) * n));
    XR_TRY(hipMalloc(&leaf_off, sizeof(int) * n));
    XR_TRY(hipMalloc(&first, sizeof(int) * (n - 1)));
    XR_TRY(hipMalloc(&dmax, sizeof(uint32_t)));
    XR_TRY(hipMalloc(&nodes, sizeof(BvhNode) * (n - 1)));
    XR_TRY(hipMalloc(&tris, sizeof(BvhTri) * tri_count));
    XR_TRY(hipMemcpyAsync(bounds, init_bounds, sizeof(init_bounds), hipMemcpyHostToDevice, st));
    XR_TRY(hipMemsetAsync(dmax, 0, sizeof(uint32_t), st));
    XR_TRY(hipMemsetAsync(parent, 0, sizeof(int) * nall, st));

    launch_gather(vertices, indices, offsets, inst_first_tri, instance_count, tri_count, ctris, cent, bounds, st);
    XR_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_morton, dim3(blocks), dim3(256), 0, st, cent, tri_count, bounds, keys, vals);
    XR_TRY(hipGetLastError());
    XR_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys, keys2, vals, vals2, n, 0, 63, st));
    XR_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, flags, offs, n, st));
    if (scan_bytes > temp_bytes) temp_bytes = scan_bytes;
    XR_TRY(hipMalloc(&temp, temp_bytes));
    {
        size_t tb = temp_bytes;
        XR_TRY(hipcub::DeviceRadixSort::SortPairs(temp, tb, keys, keys2, vals, vals2, n, 0, 63, st));
    }
    hipLaunchKernelGGL(k_ploc_leaves, dim3(blocks), dim3(256), 0, st, ctris, vals2, n, bounds, boxes, cost, count,
                       linked, clusters);
    XR_TRY(hipGetLastError());
    // Clustering: every iteration merges at least the globally closest pair,
    // so at most n - 1 iterations run (in practice a few dozen).
    while (C > 1) {
        if (++iterations > n) {
            snprintf(err, err_len, "PLOC made no progress at %d clusters", C);
            goto fail;
        }
        const uint32_t cb = (uint32_t)(C + 255) / 256;
        hipLaunchKernelGGL(k_ploc_nn, dim3(cb), dim3(256), 0, st, clusters, C, boxes, nn);
        XR_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_ploc_flags, dim3(cb), dim3(256), 0, st, nn, C, flags);
        XR_TRY(hipGetLastError());
        {
            size_t tb = temp_bytes;
            XR_TRY(hipcub::DeviceScan::ExclusiveSum(temp, tb, flags, offs, C, st));
        }
        unsigned long long last[2];
        XR_TRY(hipMemcpyAsync(&last[0], flags + (C - 1), sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        XR_TRY(hipMemcpyAsync(&last[1], offs + (C - 1), sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        hipLaunchKernelGGL(k_ploc_merge, dim3(cb), dim3(256), 0, st, n, clusters, nn, C, flags, offs, node_base,
                           boxes, cost, count, linked, child, parent, collapse, next);
        XR_TRY(hipGetLastError());
        XR_TRY(hipStreamSynchronize(st));
        const unsigned long long tot = last[0] + last[1];
        const int made = (int)(tot >> 32), kept = (int)(tot & 0xFFFFFFFFull);
        if (made <= 0 || node_base + made > n - 1) {
            snprintf(err, err_len, "PLOC merge count %d invalid at %d clusters", made, C);
            goto fail;
        }
        node_base += made;
        C = kept;
        int* t = clusters;
        clusters = next;
        next = t;
    }
    if (node_base != n - 1) {
        snprintf(err, err_len, "PLOC built %d internal nodes for %d triangles", node_base, n);
        goto fail;
    }
    hipLaunchKernelGGL(k_ploc_leaf_order, dim3(blocks), dim3(256), 0, st, n, parent, child, count, leaf_off, dmax);
    XR_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_ploc_range, dim3(blocks), dim3(256), 0, st, n, child, leaf_off, first);
    XR_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_ploc_preorder, dim3(blocks), dim3(256), 0, st, n, parent, child, collapse, linked, pre);
    XR_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_ploc_emit, dim3(blocks), dim3(256), 0, st, n, child, boxes, count, collapse, leaf_off, first,
                       pre, ctris, vals2, nodes, tris);
    XR_TRY(hipGetLastError());
    XR_TRY(hipEventRecord(e1, st));
    XR_TRY(hipMemcpyAsync(&host_depth, dmax, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    XR_TRY(hipMemcpyAsync(&stored, linked + (n - 2), sizeof(int), hipMemcpyDeviceToHost, st));
    XR_TRY(hipStreamSynchronize(st));
    XR_TRY(hipEventElapsedTime(&ms, e0, e1));

    out->nodes = nodes;
    out->tris = tris;
    if (stored < 1 || stored > n - 1) {
        snprintf(err, err_len, "PLOC stored-node count %d out of range", stored);
        goto fail;
    }
    out->node_count = (uint32_t)stored;
    out->tri_count = tri_count;
    out->max_depth = host_depth;
    out->build_ms = ms;
    nodes = nullptr;
    tris = nullptr;
    rc = hipSuccess;
    goto done;
fail:
    rc = hipErrorUnknown;
done:
    hipFree(ctris);
    hipFree(cent);
    hipFree(bounds);
    hipFree(keys);
    hipFree(keys2);
    hipFree(vals);
    hipFree(vals2);
    hipFree(boxes);
    hipFree(cost);
    hipFree(count);
    hipFree(linked);
    hipFree(pre);
    hipFree(child);
    hipFree(parent);
    hipFree(collapse);
    hipFree(clusters);
    hipFree(next);
    hipFree(nn);
    hipFree(flags);
    hipFree(offs);
    hipFree(leaf_off);
    hipFree(first);
    hipFree(dmax);
    hipFree(temp);
    hipFree(nodes);
    hipFree(tris);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    return rc;
}

// ---------------------------------------------------------------------------
// Binned SAH, top down on the host (the driver builders behind the
// reference's PREFER_FAST_TRACE flag, ApplicationRT.cpp:186-224, are of this
// family).  The canonical triangles come from k_gather on the device; boxes
// are padded as in k_refit; each node is split at the best of 64 centroid bins
// per axis by SAH with the leaf-collapse constants above (a range of at most
// kMaxLeafTris triangles becomes a leaf when that is cheaper); ranges whose
// centroids fall in one bin are split at the centroid median of their widest
// axis.  Nodes are stored in depth-first preorder (a node's left child follows
// it) within each subtree; on scenes above kSahParallel triangles the top
// levels are split on the calling thread and the subtrees below them built by
// a thread pool and appended (each spliced subtree is contiguous, preorder).
// ---------------------------------------------------------------------------
namespace {

// 64 bins per axis (32: 0.5 % slower frames, 16: 3 % slower; build time
// grows a few percent per doubling).
constexpr int kSahBins = 64;
constexpr uint32_t kSahParallel = 1u << 16;

float ord_dec_host(int o) {
    const int i = o >= 0 ? o : o ^ 0x7FFFFFFF;
    float f;
    memcpy(&f, &i, sizeof(f));
    return f;
}

struct HBox {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const HBox& b) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], b.lo[a]);
            hi[a] = fmaxf(hi[a], b.hi[a]);
        }
    }
    float area() const {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct SahPrim {
    HBox box;   // padded triangle box
    float c[3]; // centroid
};

struct SahTree {
    std::vector<BvhNode> nodes;
    uint32_t depth = 0;  // edges on the longest root-to-leaf path
};

struct SahJob {
    uint32_t b, e;  // triangle range [b, e) of ids
    HBox box;
    int parent;     // node index in the tree being built, -1 for its root
    int side;
    uint32_t depth; // edges from the tree's root to this job's node
};

void set_link(BvhNode& n, int side, int v) {
    if (side == 0) n.d.x = v; else n.d.y = v;
}

void set_child_box(BvhNode& n, int side, const HBox& b) {
    if (side == 0) {
        n.a = make_float4(b.lo[0], b.hi[0], b.lo[1], b.hi[1]);
        n.c.x = b.lo[2];
        n.c.y = b.hi[2];
    } else {
        n.b = make_float4(b.lo[0], b.hi[0], b.lo[1], b.hi[1]);
        n.c.z = b.lo[2];
        n.c.w = b.hi[2];
    }
}

// Splits ids[b, e) in place; returns the split position m (b < m < e) and
// the two sides' boxes.
uint32_t sah_split(const std::vector<SahPrim>& prims, std::vector<uint32_t>& ids, uint32_t b, uint32_t e,
                   HBox* left, HBox* right) {
    HBox cb;
    for (uint32_t i = b; i < e; ++i) {
        const SahPrim& p = prims[ids[i]];
        for (int a = 0; a < 3; ++a) {
            cb.lo[a] = fminf(cb.lo[a], p.c[a]);
            cb.hi[a] = fmaxf(cb.hi[a], p.c[a]);
        }
    }
    float best = INFINITY;
    int best_axis = -1, best_bin = -1;
    for (int a = 0; a < 3; ++a) {
        const float ext = cb.hi[a] - cb.lo[a];
        if (!(ext > 0.0f)) continue;
        const float scale = (float)kSahBins / ext;
        HBox bb[kSahBins];
        uint32_t cnt[kSahBins] = {};
        for (uint32_t i = b; i < e; ++i) {
            const SahPrim& p = prims[ids[i]];
            const int k = std::min(kSahBins - 1, (int)((p.c[a] - cb.lo[a]) * scale));
            bb[k].grow(p.box);
            ++cnt[k];
        }
        float right_area[kSahBins];
        uint32_t right_cnt[kSahBins];
        HBox acc;
        uint32_t n = 0;
        for (int k = kSahBins - 1; k > 0; --k) {
            acc.grow(bb[k]);
            n += cnt[k];
            right_area[k] = acc.area();
            right_cnt[k] = n;
        }
        acc = HBox();
        n = 0;
        for (int k = 0; k < kSahBins - 1; ++k) {
            acc.grow(bb[k]);
            n += cnt[k];
            if (n == 0 || right_cnt[k + 1] == 0) continue;
            const float c = (float)n * acc.area() + (float)right_cnt[k + 1] * right_area[k + 1];
            if (c < best) {
                best = c;
                best_axis = a;
                best_bin = k;
            }
        }
    }
    uint32_t m;
    if (best_axis >= 0) {
        const int a = best_axis;
        const float scale = (float)kSahBins / (cb.hi[a] - cb.lo[a]);
        const float lo = cb.lo[a];
        auto it = std::partition(ids.begin() + b, ids.begin() + e, [&](uint32_t id) {
            return std::min(kSahBins - 1, (int)((prims[id].c[a] - lo) * scale)) <= best_bin;
        });
        m = (uint32_t)(it - ids.begin());
    } else {  // every centroid in one place: median of the widest vertex-box axis
        HBox vb;
        for (uint32_t i = b; i < e; ++i) vb.grow(prims[ids[i]].box);
        int a = 0;
        for (int k = 1; k < 3; ++k)
            if (vb.hi[k] - vb.lo[k] > vb.hi[a] - vb.lo[a]) a = k;
        m = b + (e - b) / 2;
        std::nth_element(ids.begin() + b, ids.begin() + m, ids.begin() + e,
                         [&](uint32_t x, uint32_t y) { return prims[x].c[a] < prims[y].c[a]; });
    }
    *left = HBox();
    *right = HBox();
    for (uint32_t i = b; i < m; ++i) left->grow(prims[ids[i]].box);
    for (uint32_t i = m; i < e; ++i) right->grow(prims[ids[i]].box);
    return m;
}

// Builds the subtree of ids[b, e) (an internal node: the caller has decided
// not to make it a leaf) into `tree`, nodes numbered from 0.  With chunk > 0,
// subtrees of at most `chunk` triangles are built by a pool of threads and
// spliced in afterwards.
void sah_subtree(const std::vector<SahPrim>& prims, std::vector<uint32_t>& ids, uint32_t b, uint32_t e,
                 const HBox& box, SahTree* tree, uint32_t chunk) {
    std::vector<SahJob> stack;
    stack.push_back(SahJob{b, e, box, -1, 0, 0});
    struct Deferred {
        uint32_t b, e;
        HBox box;
        int parent, side;
        uint32_t depth;
        SahTree sub;
    };
    std::vector<Deferred> deferred;
    while (!stack.empty()) {
        SahJob j = stack.back();
        stack.pop_back();
        const uint32_t cnt = j.e - j.b;
        const float area = j.box.area();
        if (j.parent >= 0 && cnt <= (uint32_t)kMaxLeafTris) {
            // leaf if cheaper than the best split (always for one triangle)
            bool leaf = cnt == 1;
            if (!leaf) {
                std::vector<uint32_t> save(ids.begin() + j.b, ids.begin() + j.e);
                HBox l, r;
                const uint32_t m = sah_split(prims, ids, j.b, j.e, &l, &r);
                const float split = kSahNode * area + kSahTri * ((float)(m - j.b) * l.area() + (float)(j.e - m) * r.area());
                leaf = kSahTri * (float)cnt * area <= split;
                std::copy(save.begin(), save.end(), ids.begin() + j.b);
            }
            if (leaf) {
                set_link(tree->nodes[j.parent], j.side, ~(int)(j.b * 16u + (cnt - 1u)));
                tree->depth = std::max(tree->depth, j.depth);
                continue;
            }
        }
        if (chunk && j.parent >= 0 && cnt <= chunk) {
            // built on its own thread below, spliced in afterwards
            deferred.push_back(Deferred{j.b, j.e, j.box, j.parent, j.side, j.depth, SahTree()});
            continue;
        }
        const int k = (int)tree->nodes.size();
        BvhNode nd;
        memset(&nd, 0, sizeof(nd));
        tree->nodes.push_back(nd);
        if (j.parent >= 0) set_link(tree->nodes[j.parent], j.side, k);
        HBox l, r;
        const uint32_t m = sah_split(prims, ids, j.b, j.e, &l, &r);
        set_child_box(tree->nodes[k], 0, l);
        set_child_box(tree->nodes[k], 1, r);
        stack.push_back(SahJob{m, j.e, r, k, 1, j.depth + 1});
        stack.push_back(SahJob{j.b, m, l, k, 0, j.depth + 1});
    }
    if (deferred.empty()) return;
    std::atomic<size_t> next{0};
    std::atomic<bool> failed{false};
    const unsigned hw = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    auto work = [&] {
        try {
            for (size_t k = next++; k < deferred.size() && !failed; k = next++) {
                Deferred& d = deferred[k];
                sah_subtree(prims, ids, d.b, d.e, d.box, &d.sub, 0);
            }
        } catch (...) {  // out of host memory: reported by the caller
            failed = true;
        }
    };
    try {
        for (unsigned t = 0; t < std::min<size_t>(hw, deferred.size()); ++t) pool.emplace_back(work);
    } catch (...) {  // no more threads: the calling thread does the rest
    }
    work();
    for (auto& t : pool) t.join();
    if (failed) throw std::bad_alloc();
    for (auto& d : deferred) {
        const int base = (int)tree->nodes.size();
        set_link(tree->nodes[d.parent], d.side, base);
        for (BvhNode n : d.sub.nodes) {
            if (n.d.x >= 0) n.d.x += base;
            if (n.d.y >= 0) n.d.y += base;
            tree->nodes.push_back(n);
        }
        tree->depth = std::max(tree->depth, d.depth + d.sub.depth);
    }
}

}  // namespace

hipError_t build_sah(const float* vertices, const uint32_t* indices, const uint2* offsets,
                     const uint32_t* inst_first_tri, uint32_t instance_count, uint32_t tri_count, hipStream_t st,
                     BvhBuildResult* out, char* err, size_t err_len) {
    if (tri_count < 2)
        return build_lbvh(vertices, indices, offsets, inst_first_tri, instance_count, tri_count, st, out, err,
                          err_len);
    hipError_t rc = hipSuccess;
    const uint32_t n = tri_count;
    BvhTri* ctris = nullptr;
    float4* cent = nullptr;
    int* bounds = nullptr;
    BvhNode* nodes = nullptr;
    BvhTri* tris = nullptr;
    const int init_bounds[12] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF, (int)0x80000000, (int)0x80000000,
                                 (int)0x80000000, 0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF, (int)0x80000000,
                                 (int)0x80000000, (int)0x80000000};
    int hb[12];
    std::vector<BvhTri> h;
    std::vector<SahPrim> prims;
    std::vector<uint32_t> ids;
    std::vector<BvhTri> sorted;
    SahTree tree;
    HBox all;
    float pad = 0.0f, ext = 0.0f, mag = 0.0f;
    const auto t0 = std::chrono::steady_clock::now();
    try {
        h.resize(n);
        prims.resize(n);
        ids.resize(n);
        sorted.resize(n);
    } catch (...) {
        snprintf(err, err_len, "host memory for the SAH build");
        return hipErrorOutOfMemory;
    }
    XR_TRY(hipMalloc(&ctris, sizeof(BvhTri) * n));
    XR_TRY(hipMalloc(&cent, sizeof(float4) * n));
    XR_TRY(hipMalloc(&bounds, sizeof(int) * 12));
    XR_TRY(hipMemcpyAsync(bounds, init_bounds, sizeof(init_bounds), hipMemcpyHostToDevice, st));
    launch_gather(vertices, indices, offsets, inst_first_tri, instance_count, n, ctris, cent, bounds, st);
    XR_TRY(hipGetLastError());
    XR_TRY(hipMemcpyAsync(h.data(), ctris, sizeof(BvhTri) * n, hipMemcpyDeviceToHost, st));
    XR_TRY(hipMemcpyAsync(hb, bounds, sizeof(hb), hipMemcpyDeviceToHost, st));
    XR_TRY(hipStreamSynchronize(st));
    // box_pad, on the host
    for (int a = 0; a < 3; ++a) {
        const float lo = ord_dec_host(hb[6 + a]), hi = ord_dec_host(hb[9 + a]);
        ext = fmaxf(ext, hi - lo);
        mag = fmaxf(mag, fmaxf(fabsf(lo), fabsf(hi)));
    }
    pad = fmaxf(ext * (1.0f / 65536.0f), mag * (1.0f / 1048576.0f));
    for (uint32_t i = 0; i < n; ++i) {
        const BvhTri& t = h[i];
        SahPrim& p = prims[i];
        const float v[3][3] = {{t.v0.x, t.v0.y, t.v0.z}, {t.v1.x, t.v1.y, t.v1.z}, {t.v2.x, t.v2.y, t.v2.z}};
        for (int a = 0; a < 3; ++a) {
            p.box.lo[a] = fminf(fminf(v[0][a], v[1][a]), v[2][a]) - pad;
            p.box.hi[a] = fmaxf(fmaxf(v[0][a], v[1][a]), v[2][a]) + pad;
            p.c[a] = (v[0][a] + v[1][a] + v[2][a]) * (1.0f / 3.0f);
        }
        ids[i] = i;
        all.grow(p.box);
    }
    try {
        // subtrees handed to threads: about 256 of them on large scenes
        const uint32_t chunk = n > kSahParallel ? std::max(n / 256u, 4096u) : 0u;
        sah_subtree(prims, ids, 0, n, all, &tree, chunk);
    } catch (...) {
        snprintf(err, err_len, "SAH build failed (host memory or threads)");
        goto fail;
    }
    for (uint32_t i = 0; i < n; ++i) sorted[i] = h[ids[i]];
    XR_TRY(hipMalloc(&nodes, sizeof(BvhNode) * tree.nodes.size()));
    XR_TRY(hipMalloc(&tris, sizeof(BvhTri) * n));
    XR_TRY(hipMemcpyAsync(nodes, tree.nodes.data(), sizeof(BvhNode) * tree.nodes.size(), hipMemcpyHostToDevice, st));
    XR_TRY(hipMemcpyAsync(tris, sorted.data(), sizeof(BvhTri) * n, hipMemcpyHostToDevice, st));
    XR_TRY(hipStreamSynchronize(st));
    out->nodes = nodes;
    out->tris = tris;
    out->node_count = (uint32_t)tree.nodes.size();
    out->tri_count = n;
    out->max_depth = tree.depth;
    out->build_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    nodes = nullptr;
    tris = nullptr;
    rc = hipSuccess;
    goto done;
fail:
    rc = hipErrorUnknown;
done:
    hipFree(ctris);
    hipFree(cent);
    hipFree(bounds);
    hipFree(nodes);
    hipFree(tris);
    return rc;
}

// ---------------------------------------------------------------------------
// Binned SAH on the GPU.  Makes the host builder's decisions (64 centroid bins
// per axis, the first lowest-cost bin in axis then bin order, the same float
// arithmetic for boxes, areas and costs, leaves of at most kMaxLeafTris
// triangles where they are cheaper than their best split), so it builds the
// same tree as build_sah: every split is a function of the set of triangles
// in the node, which does not depend on the order either builder keeps them
// in.  (Ranges whose centroids all coincide are halved in index order; the
// host takes a median by its own order there, so such ranges may be cut
// differently.)
//
// Levels of large ranges (more than kGsSmall triangles) are split with every
// workgroup on the device, one kernel per step: centroid bounds, binning
// (LDS bins per 2048-triangle chunk, merged with one global atomic per bin
// and value), the SAH sweep (one wave per range, a lane per bin), a stable
// partition (per-chunk counts, a scan per range, a scatter).  The host reads
// back the next level's ranges (one synchronisation per level) to lay out its
// chunks.  Ranges of at most kGsSmall triangles are then built one per wave,
// depth first with a job stack, bins and triangle ids in LDS.
// ---------------------------------------------------------------------------
namespace {

constexpr uint32_t kGsChunk = 2048;
constexpr uint32_t kGsSmall = 512;
constexpr int kGsAxisBins = 3 * kSahBins;

struct GsSeg {
    uint32_t b, e;    // triangle range [b, e) of the id array
    int parent;       // node index, -1 for the root
    uint32_t sd;      // side | depth << 1 (depth: edges from the root)
    float lo[3], hi[3];  // the range's box, as its parent stored it
};

// Per large range of a level: centroid bounds and the 192 bins
// (order-preserving int encodings of the box floats), then the split.
struct GsBins {
    int cb[6];
    int box[kGsAxisBins][6];
    uint32_t cnt[kGsAxisBins];
};
struct GsSplit {
    int axis;          // -1: halve in index order
    int bin;
    float lo, scale;   // bin of centroid c: min(63, (int)((c - lo) * scale))
    uint32_t m;        // first triangle of the right range
    uint32_t node;     // the range's node
};
struct GsChunk {
    uint32_t seg, b, e;
    uint32_t first_chunk;  // first chunk of the same range
};

__device__ __forceinline__ float gs_area(const float lo[3], const float hi[3]) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

__device__ __forceinline__ int gs_bin(float c, float lo, float scale) {
    const int k = (int)((c - lo) * scale);
    return k < kSahBins - 1 ? k : kSahBins - 1;
}

// Padded triangle boxes (the host's SahPrim.box).
__global__ void k_gs_prims(const BvhTri* __restrict__ ctris, uint32_t n, float pad, float4* __restrict__ plo,
                           float4* __restrict__ phi, uint32_t* __restrict__ ids) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const BvhTri t = ctris[i];
    plo[i] = make_float4(fminf(fminf(t.v0.x, t.v1.x), t.v2.x) - pad, fminf(fminf(t.v0.y, t.v1.y), t.v2.y) - pad,
                         fminf(fminf(t.v0.z, t.v1.z), t.v2.z) - pad, 0.0f);
    phi[i] = make_float4(fmaxf(fmaxf(t.v0.x, t.v1.x), t.v2.x) + pad, fmaxf(fmaxf(t.v0.y, t.v1.y), t.v2.y) + pad,
                         fmaxf(fmaxf(t.v0.z, t.v1.z), t.v2.z) + pad, 0.0f);
    ids[i] = i;
}

__global__ void k_gs_init_bins(GsBins* __restrict__ bins, uint32_t nseg) {
    const uint32_t s = blockIdx.x;
    if (s >= nseg) return;
    GsBins& B = bins[s];
    for (uint32_t i = threadIdx.x; i < 6; i += blockDim.x) B.cb[i] = i < 3 ? 0x7FFFFFFF : (int)0x80000000;
    for (uint32_t i = threadIdx.x; i < kGsAxisBins; i += blockDim.x) {
        for (int k = 0; k < 3; ++k) {
            B.box[i][k] = 0x7FFFFFFF;
            B.box[i][3 + k] = (int)0x80000000;
        }
        B.cnt[i] = 0u;
    }
}

// Centroid bounds of each chunk's range.
__global__ __launch_bounds__(256) void k_gs_cbounds(const GsChunk* __restrict__ chunks, const uint32_t* __restrict__ ids,
                                                    const float4* __restrict__ cent, GsBins* __restrict__ bins) {
    const GsChunk ch = chunks[blockIdx.x];
    int v[6] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF, (int)0x80000000, (int)0x80000000, (int)0x80000000};
    for (uint32_t i = ch.b + threadIdx.x; i < ch.e; i += 256u) {
        const float4 c = cent[ids[i]];
        const int o[3] = {ord_enc(c.x), ord_enc(c.y), ord_enc(c.z)};
        for (int a = 0; a < 3; ++a) {
            v[a] = min(v[a], o[a]);
            v[3 + a] = max(v[3 + a], o[a]);
        }
    }
    for (int a = 0; a < 3; ++a)
        for (int off = 32; off > 0; off >>= 1) {
            v[a] = min(v[a], __shfl_xor(v[a], off));
            v[3 + a] = max(v[3 + a], __shfl_xor(v[3 + a], off));
        }
    if ((threadIdx.x & 63u) == 0u) {
        GsBins& B = bins[ch.seg];
        for (int a = 0; a < 3; ++a) {
            atomicMin(&B.cb[a], v[a]);
            atomicMax(&B.cb[3 + a], v[3 + a]);
        }
    }
}

__device__ __forceinline__ void gs_axis_frame(const int cb[6], int a, float* lo, float* scale, bool* valid) {
    const float l = ord_dec(cb[a]), h = ord_dec(cb[3 + a]);
    const float ext = h - l;
    *valid = ext > 0.0f;
    *lo = l;
    *scale = *valid ? (float)kSahBins / ext : 0.0f;
}

// Bins of each chunk in LDS, merged into its range's bins.
__global__ __launch_bounds__(256) void k_gs_bin(const GsChunk* __restrict__ chunks, const uint32_t* __restrict__ ids,
                                                const float4* __restrict__ cent, const float4* __restrict__ plo,
                                                const float4* __restrict__ phi, GsBins* __restrict__ bins) {
    __shared__ int sbox[kGsAxisBins][6];
    __shared__ uint32_t scnt[kGsAxisBins];
    const GsChunk ch = chunks[blockIdx.x];
    GsBins& B = bins[ch.seg];
    for (uint32_t i = threadIdx.x; i < kGsAxisBins; i += 256u) {
        for (int k = 0; k < 3; ++k) {
            sbox[i][k] = 0x7FFFFFFF;
            sbox[i][3 + k] = (int)0x80000000;
        }
        scnt[i] = 0u;
    }
    float lo[3], scale[3];
    bool valid[3];
    for (int a = 0; a < 3; ++a) gs_axis_frame(B.cb, a, &lo[a], &scale[a], &valid[a]);
    __syncthreads();
    for (uint32_t i = ch.b + threadIdx.x; i < ch.e; i += 256u) {
        const uint32_t id = ids[i];
        const float4 c = cent[id], l = plo[id], h = phi[id];
        const float cc[3] = {c.x, c.y, c.z};
        const int ol[3] = {ord_enc(l.x), ord_enc(l.y), ord_enc(l.z)};
        const int oh[3] = {ord_enc(h.x), ord_enc(h.y), ord_enc(h.z)};
        for (int a = 0; a < 3; ++a) {
            if (!valid[a]) continue;
            const int k = a * kSahBins + gs_bin(cc[a], lo[a], scale[a]);
            for (int q = 0; q < 3; ++q) {
                atomicMin(&sbox[k][q], ol[q]);
                atomicMax(&sbox[k][3 + q], oh[q]);
            }
            atomicAdd(&scnt[k], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kGsAxisBins; i += 256u) {
        if (scnt[i] == 0u) continue;
        for (int q = 0; q < 3; ++q) {
            atomicMin(&B.box[i][q], sbox[i][q]);
            atomicMax(&B.box[i][3 + q], sbox[i][3 + q]);
        }
        atomicAdd(&B.cnt[i], scnt[i]);
    }
}

// The SAH sweep of one axis with a lane per bin: the left side of split k is
// bins 0..k, the right side bins k+1..63 (unions are exact, so these are the
// host's accumulated boxes).  Returns per lane its split's cost (infinite
// where a side is empty or k = 63) and the two sides' boxes and counts.
struct GsSide {
    float lo[3], hi[3];
    uint32_t n;
};
__device__ __forceinline__ void gs_sweep(const int (*box)[6], const uint32_t* cnt, int lane, float* cost, GsSide* L,
                                         GsSide* R) {
    int bl[6], br[6];
    for (int q = 0; q < 6; ++q) bl[q] = br[q] = box[lane][q];
    uint32_t nl = cnt[lane], nr = cnt[lane];
    // inclusive prefix (left) and suffix (right) over the lanes
    for (int off = 1; off < 64; off <<= 1) {
        int ul[6], ur[6];
        for (int q = 0; q < 6; ++q) {
            ul[q] = __shfl_up(bl[q], off);
            ur[q] = __shfl_down(br[q], off);
        }
        const uint32_t cl = (uint32_t)__shfl_up((int)nl, off), cr = (uint32_t)__shfl_down((int)nr, off);
        if (lane >= off) {
            for (int q = 0; q < 3; ++q) {
                bl[q] = min(bl[q], ul[q]);
                bl[3 + q] = max(bl[3 + q], ul[3 + q]);
            }
            nl += cl;
        }
        if (lane + off < 64) {
            for (int q = 0; q < 3; ++q) {
                br[q] = min(br[q], ur[q]);
                br[3 + q] = max(br[3 + q], ur[3 + q]);
            }
            nr += cr;
        }
    }
    // right side of split k: the suffix from bin k + 1
    int rb[6];
    for (int q = 0; q < 6; ++q) rb[q] = __shfl_down(br[q], 1);
    // every lane takes part in the shuffle (lane 63 is the source of lane 62)
    const uint32_t rn_next = (uint32_t)__shfl_down((int)nr, 1);
    const uint32_t rn = lane + 1 < 64 ? rn_next : 0u;
    for (int q = 0; q < 3; ++q) {
        L->lo[q] = ord_dec(bl[q]);
        L->hi[q] = ord_dec(bl[3 + q]);
        R->lo[q] = ord_dec(rb[q]);
        R->hi[q] = ord_dec(rb[3 + q]);
    }
    L->n = nl;
    R->n = lane + 1 < 64 ? rn : 0u;
    if (lane == 63 || L->n == 0u || R->n == 0u) {
        *cost = INFINITY;
    } else {
        *cost = (float)L->n * gs_area(L->lo, L->hi) + (float)R->n * gs_area(R->lo, R->hi);
    }
}

// The first lowest (cost, axis, bin) of the wave: returns the winning lane's
// packed key (axis * 64 + bin), or -1 when no split is finite.
__device__ __forceinline__ int gs_argmin(float cost, int key) {
    float c = cost < INFINITY ? cost : INFINITY;  // NaN never wins
    int k = cost < INFINITY ? key : 0x7FFFFFFF;
    for (int off = 32; off > 0; off >>= 1) {
        const float oc = __shfl_xor(c, off);
        const int ok = __shfl_xor(k, off);
        if (oc < c || (oc == c && ok < k)) {
            c = oc;
            k = ok;
        }
    }
    return k == 0x7FFFFFFF ? -1 : k;
}

__device__ __forceinline__ void gs_set_child(BvhNode* nd, int side, const float lo[3], const float hi[3]) {
    if (side == 0) {
        nd->a = make_float4(lo[0], hi[0], lo[1], hi[1]);
        nd->c.x = lo[2];
        nd->c.y = hi[2];
    } else {
        nd->b = make_float4(lo[0], hi[0], lo[1], hi[1]);
        nd->c.z = lo[2];
        nd->c.w = hi[2];
    }
}
__device__ __forceinline__ void gs_set_link(BvhNode* nodes, int parent, int side, int v) {
    if (side == 0) nodes[parent].d.x = v;
    else nodes[parent].d.y = v;
}

// One wave per large range: the split, its node, its two children (appended
// to the next level's large ranges or to the small ones).  Ranges whose
// centroids coincide are halved in index order (their child boxes come from
// k_gs_median_boxes).
__global__ __launch_bounds__(64) void k_gs_split(const GsSeg* __restrict__ segs, uint32_t nseg,
                                                 const GsBins* __restrict__ bins, GsSplit* __restrict__ split,
                                                 BvhNode* __restrict__ nodes, uint32_t* __restrict__ counters,
                                                 GsSeg* __restrict__ next, GsSeg* __restrict__ small) {
    const uint32_t s = blockIdx.x;
    if (s >= nseg) return;
    const int lane = (int)threadIdx.x;
    const GsSeg g = segs[s];
    const GsBins& B = bins[s];
    float best_cost = INFINITY;
    int best_key = -1;
    GsSide L, R;
    GsSide bestL, bestR;
    for (int a = 0; a < 3; ++a) {
        float lo, scale;
        bool valid;
        gs_axis_frame(B.cb, a, &lo, &scale, &valid);
        if (!valid) continue;
        float c;
        gs_sweep(&B.box[a * kSahBins], &B.cnt[a * kSahBins], lane, &c, &L, &R);
        const int key = gs_argmin(c, a * kSahBins + lane);
        if (key < 0) continue;
        const int wl = key % kSahBins;
        const float wc = __shfl(c, wl);
        if (wc < best_cost) {  // axes in order, strictly lower: the host's rule
            best_cost = wc;
            best_key = key;
            for (int q = 0; q < 3; ++q) {
                bestL.lo[q] = __shfl(L.lo[q], wl);
                bestL.hi[q] = __shfl(L.hi[q], wl);
                bestR.lo[q] = __shfl(R.lo[q], wl);
                bestR.hi[q] = __shfl(R.hi[q], wl);
            }
            bestL.n = (uint32_t)__shfl((int)L.n, wl);
            bestR.n = (uint32_t)__shfl((int)R.n, wl);
        }
    }
    if (lane != 0) return;
    GsSplit sp;
    const uint32_t cnt = g.e - g.b;
    const uint32_t k = atomicAdd(&counters[0], 1u);
    sp.node = k;
    if (best_key >= 0) {
        sp.axis = best_key / kSahBins;
        sp.bin = best_key % kSahBins;
        bool valid;
        gs_axis_frame(B.cb, sp.axis, &sp.lo, &sp.scale, &valid);
        sp.m = g.b + bestL.n;
    } else {
        sp.axis = -1;
        sp.bin = 0;
        sp.lo = 0.0f;
        sp.scale = 0.0f;
        sp.m = g.b + cnt / 2u;
    }
    split[s] = sp;
    BvhNode nd;
    nd.a = nd.b = nd.c = make_float4(0.f, 0.f, 0.f, 0.f);
    nd.d = make_int4(0, 0, 0, 0);
    if (best_key >= 0) {
        gs_set_child(&nd, 0, bestL.lo, bestL.hi);
        gs_set_child(&nd, 1, bestR.lo, bestR.hi);
    }
    nodes[k] = nd;
    if (g.parent >= 0) gs_set_link(nodes, g.parent, (int)(g.sd & 1u), (int)k);
    const uint32_t depth = (g.sd >> 1) + 1u;
    for (int side = 0; side < 2; ++side) {
        GsSeg c;
        c.b = side == 0 ? g.b : sp.m;
        c.e = side == 0 ? sp.m : g.e;
        c.parent = (int)k;
        c.sd = (uint32_t)side | (depth << 1);
        const GsSide& S = side == 0 ? bestL : bestR;
        for (int q = 0; q < 3; ++q) {
            c.lo[q] = best_key >= 0 ? S.lo[q] : 0.0f;
            c.hi[q] = best_key >= 0 ? S.hi[q] : 0.0f;
        }
        if (c.e - c.b > kGsSmall) next[atomicAdd(&counters[1], 1u)] = c;
        else small[atomicAdd(&counters[2], 1u)] = c;
    }
}

// Goes-left test of triangle id under range split sp.
__device__ __forceinline__ bool gs_left(const GsSplit& sp, const float4* __restrict__ cent, uint32_t id, uint32_t pos) {
    if (sp.axis < 0) return pos < sp.m;
    const float4 c = cent[id];
    const float cc = sp.axis == 0 ? c.x : (sp.axis == 1 ? c.y : c.z);
    return gs_bin(cc, sp.lo, sp.scale) <= sp.bin;
}

// Per chunk: how many of its triangles go left.
__global__ __launch_bounds__(256) void k_gs_count(const GsChunk* __restrict__ chunks, const uint32_t* __restrict__ ids,
                                                  const float4* __restrict__ cent, const GsSplit* __restrict__ split,
                                                  uint32_t* __restrict__ chunk_left) {
    __shared__ uint32_t part[4];
    const GsChunk ch = chunks[blockIdx.x];
    const GsSplit sp = split[ch.seg];
    uint32_t n = 0;
    for (uint32_t i = ch.b + threadIdx.x; i < ch.e; i += 256u) n += gs_left(sp, cent, ids[i], i) ? 1u : 0u;
    for (int off = 32; off > 0; off >>= 1) n += (uint32_t)__shfl_xor((int)n, off);
    if ((threadIdx.x & 63u) == 0u) part[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) chunk_left[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// Exclusive scan of chunk_left over each range's chunks (one workgroup per
// range; its chunks are consecutive): chunk_off[c] = left triangles of the
// range's chunks before c.
__global__ __launch_bounds__(256) void k_gs_chunk_scan(const GsChunk* __restrict__ chunks,
                                                       const uint32_t* __restrict__ seg_first_chunk,
                                                       const uint32_t* __restrict__ chunk_left,
                                                       uint32_t* __restrict__ chunk_off) {
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t carry;
    const uint32_t c0 = seg_first_chunk[blockIdx.x], c1 = seg_first_chunk[blockIdx.x + 1];
    const int lane = (int)(threadIdx.x & 63u), wv = (int)(threadIdx.x >> 6);
    if (threadIdx.x == 0) carry = 0u;
    __syncthreads();
    for (uint32_t base = c0; base < c1; base += 256u) {
        const uint32_t c = base + threadIdx.x;
        const uint32_t v = c < c1 ? chunk_left[c] : 0u;
        uint32_t x = v;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) wsum[wv] = x;
        __syncthreads();
        uint32_t before = carry;
        for (int k = 0; k < wv; ++k) before += wsum[k];
        if (c < c1) chunk_off[c] = before + x - v;
        __syncthreads();
        if (threadIdx.x == 0) carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
}

// Stable partition of every chunk into ids_out; each chunk's left triangles go
// to b + chunk_off + rank, its right ones to m + (chunk start - b - chunk_off) + rank.
__global__ __launch_bounds__(256) void k_gs_scatter(const GsChunk* __restrict__ chunks, const GsSeg* __restrict__ segs,
                                                    const uint32_t* __restrict__ ids, const float4* __restrict__ cent,
                                                    const GsSplit* __restrict__ split,
                                                    const uint32_t* __restrict__ chunk_off,
                                                    uint32_t* __restrict__ ids_out) {
    __shared__ uint32_t wl[4], wr[4];
    const GsChunk ch = chunks[blockIdx.x];
    const GsSeg g = segs[ch.seg];
    const GsSplit sp = split[ch.seg];
    const int lane = (int)(threadIdx.x & 63u), wv = (int)(threadIdx.x >> 6);
    uint32_t lbase = g.b + chunk_off[blockIdx.x];
    uint32_t rbase = sp.m + (ch.b - g.b - chunk_off[blockIdx.x]);
    for (uint32_t base = ch.b; base < ch.e; base += 256u) {
        const uint32_t i = base + threadIdx.x;
        const bool live = i < ch.e;
        const uint32_t id = live ? ids[i] : 0u;
        const bool left = live && gs_left(sp, cent, id, i);
        const unsigned long long ml = __ballot(left), mr = __ballot(live && !left);
        if (lane == 0) {
            wl[wv] = (uint32_t)__popcll(ml);
            wr[wv] = (uint32_t)__popcll(mr);
        }
        __syncthreads();
        uint32_t pl = lbase, pr = rbase;
        for (int k = 0; k < wv; ++k) {
            pl += wl[k];
            pr += wr[k];
        }
        const unsigned long long below = (1ull << lane) - 1ull;
        if (left) ids_out[pl + (uint32_t)__popcll(ml & below)] = id;
        else if (live) ids_out[pr + (uint32_t)__popcll(mr & below)] = id;
        lbase += wl[0] + wl[1] + wl[2] + wl[3];
        rbase += wr[0] + wr[1] + wr[2] + wr[3];
        __syncthreads();
    }
}

__global__ void k_gs_copy_back(const GsChunk* __restrict__ chunks, const uint32_t* __restrict__ src,
                               uint32_t* __restrict__ dst) {
    const GsChunk ch = chunks[blockIdx.x];
    for (uint32_t i = ch.b + threadIdx.x; i < ch.e; i += blockDim.x) dst[i] = src[i];
}

// Child boxes of ranges halved in index order: unions of their triangles' boxes.
__global__ __launch_bounds__(64) void k_gs_median_boxes(const GsSeg* __restrict__ segs, uint32_t nseg,
                                                        const GsSplit* __restrict__ split,
                                                        const uint32_t* __restrict__ ids,
                                                        const float4* __restrict__ plo, const float4* __restrict__ phi,
                                                        BvhNode* __restrict__ nodes, GsSeg* __restrict__ next,
                                                        GsSeg* __restrict__ small, const uint32_t* __restrict__ counters) {
    const uint32_t s = blockIdx.x;
    if (s >= nseg || split[s].axis >= 0) return;
    const GsSeg g = segs[s];
    const GsSplit sp = split[s];
    for (int side = 0; side < 2; ++side) {
        const uint32_t b = side == 0 ? g.b : sp.m, e = side == 0 ? sp.m : g.e;
        int v[6] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF, (int)0x80000000, (int)0x80000000, (int)0x80000000};
        for (uint32_t i = b + threadIdx.x; i < e; i += 64u) {
            const float4 l = plo[ids[i]], h = phi[ids[i]];
            v[0] = min(v[0], ord_enc(l.x)); v[1] = min(v[1], ord_enc(l.y)); v[2] = min(v[2], ord_enc(l.z));
            v[3] = max(v[3], ord_enc(h.x)); v[4] = max(v[4], ord_enc(h.y)); v[5] = max(v[5], ord_enc(h.z));
        }
        for (int off = 32; off > 0; off >>= 1)
            for (int q = 0; q < 3; ++q) {
                v[q] = min(v[q], __shfl_xor(v[q], off));
                v[3 + q] = max(v[3 + q], __shfl_xor(v[3 + q], off));
            }
        if (threadIdx.x != 0) continue;
        float lo[3], hi[3];
        for (int q = 0; q < 3; ++q) {
            lo[q] = ord_dec(v[q]);
            hi[q] = ord_dec(v[3 + q]);
        }
        gs_set_child(&nodes[sp.node], side, lo, hi);
        // the child range carries its box to its own node (either list)
        GsSeg* lists[2] = {next, small};
        const uint32_t ns[2] = {counters[1], counters[2]};
        for (int l = 0; l < 2; ++l)
            for (uint32_t j = 0; j < ns[l]; ++j)
                if (lists[l][j].parent == (int)sp.node && (lists[l][j].sd & 1u) == (uint32_t)side)
                    for (int q = 0; q < 3; ++q) {
                        lists[l][j].lo[q] = lo[q];
                        lists[l][j].hi[q] = hi[q];
                    }
    }
}

// ---- ranges of at most kGsSmall triangles: one wave each ----
struct GsJob {
    uint32_t b, e;  // local range of the wave's id copy
    int parent;
    uint32_t sd;
    float lo[3], hi[3];
};
// Depth first, the larger child pushed first and the smaller taken first:
// at most log2(kGsSmall) + 1 jobs wait at any time.
constexpr int kGsJobs = 16;

struct GsWave {
    int box[kGsAxisBins][6];
    uint32_t cnt[kGsAxisBins];
    uint32_t ids[kGsSmall];
    uint32_t tmp[kGsSmall];
    GsJob jobs[kGsJobs];
};

__device__ __forceinline__ void gs_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The host's sah_split on W.ids[b, b + cnt): returns the left count m
// (0 < m < cnt) and the sides' boxes; with `partition` the range is
// reordered left side first (stably).
__device__ uint32_t gs_wave_split(GsWave& W, uint32_t b, uint32_t cnt, const float4* __restrict__ cent,
                                  const float4* __restrict__ plo, const float4* __restrict__ phi, GsSide* Lo,
                                  GsSide* Ro, bool partition) {
    const int lane = (int)threadIdx.x & 63;
    int cb[6] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF, (int)0x80000000, (int)0x80000000, (int)0x80000000};
    for (uint32_t i = lane; i < cnt; i += 64u) {
        const float4 c = cent[W.ids[b + i]];
        cb[0] = min(cb[0], ord_enc(c.x)); cb[1] = min(cb[1], ord_enc(c.y)); cb[2] = min(cb[2], ord_enc(c.z));
        cb[3] = max(cb[3], ord_enc(c.x)); cb[4] = max(cb[4], ord_enc(c.y)); cb[5] = max(cb[5], ord_enc(c.z));
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int q = 0; q < 3; ++q) {
            cb[q] = min(cb[q], __shfl_xor(cb[q], off));
            cb[3 + q] = max(cb[3 + q], __shfl_xor(cb[3 + q], off));
        }
    for (int i = lane; i < kGsAxisBins; i += 64) {
        for (int q = 0; q < 3; ++q) {
            W.box[i][q] = 0x7FFFFFFF;
            W.box[i][3 + q] = (int)0x80000000;
        }
        W.cnt[i] = 0u;
    }
    float alo[3], ascale[3];
    bool avalid[3];
    for (int a = 0; a < 3; ++a) gs_axis_frame(cb, a, &alo[a], &ascale[a], &avalid[a]);
    gs_wave_sync();
    for (uint32_t i = lane; i < cnt; i += 64u) {
        const uint32_t id = W.ids[b + i];
        const float4 c = cent[id], l = plo[id], h = phi[id];
        const float cc[3] = {c.x, c.y, c.z};
        const int ol[3] = {ord_enc(l.x), ord_enc(l.y), ord_enc(l.z)};
        const int oh[3] = {ord_enc(h.x), ord_enc(h.y), ord_enc(h.z)};
        for (int a = 0; a < 3; ++a) {
            if (!avalid[a]) continue;
            const int k = a * kSahBins + gs_bin(cc[a], alo[a], ascale[a]);
            for (int q = 0; q < 3; ++q) {
                atomicMin(&W.box[k][q], ol[q]);
                atomicMax(&W.box[k][3 + q], oh[q]);
            }
            atomicAdd(&W.cnt[k], 1u);
        }
    }
    gs_wave_sync();
    float best_cost = INFINITY;
    int best_key = -1;
    GsSide L, R;
    for (int a = 0; a < 3; ++a) {
        if (!avalid[a]) continue;
        float c;
        gs_sweep(&W.box[a * kSahBins], &W.cnt[a * kSahBins], lane, &c, &L, &R);
        const int key = gs_argmin(c, a * kSahBins + lane);
        if (key < 0) continue;
        const int wl = key % kSahBins;
        const float wc = __shfl(c, wl);
        if (wc < best_cost) {
            best_cost = wc;
            best_key = key;
            for (int q = 0; q < 3; ++q) {
                Lo->lo[q] = __shfl(L.lo[q], wl);
                Lo->hi[q] = __shfl(L.hi[q], wl);
                Ro->lo[q] = __shfl(R.lo[q], wl);
                Ro->hi[q] = __shfl(R.hi[q], wl);
            }
            Lo->n = (uint32_t)__shfl((int)L.n, wl);
            Ro->n = (uint32_t)__shfl((int)R.n, wl);
        }
    }
    if (best_key < 0) {  // every centroid in one place: halve in index order
        const uint32_t m = cnt / 2u;
        for (int side = 0; side < 2; ++side) {
            const uint32_t sb = side == 0 ? 0u : m, se = side == 0 ? m : cnt;
            int v[6] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF, (int)0x80000000, (int)0x80000000, (int)0x80000000};
            for (uint32_t i = sb + lane; i < se; i += 64u) {
                const float4 l = plo[W.ids[b + i]], h = phi[W.ids[b + i]];
                v[0] = min(v[0], ord_enc(l.x)); v[1] = min(v[1], ord_enc(l.y)); v[2] = min(v[2], ord_enc(l.z));
                v[3] = max(v[3], ord_enc(h.x)); v[4] = max(v[4], ord_enc(h.y)); v[5] = max(v[5], ord_enc(h.z));
            }
            for (int off = 32; off > 0; off >>= 1)
                for (int q = 0; q < 3; ++q) {
                    v[q] = min(v[q], __shfl_xor(v[q], off));
                    v[3 + q] = max(v[3 + q], __shfl_xor(v[3 + q], off));
                }
            GsSide* S = side == 0 ? Lo : Ro;
            for (int q = 0; q < 3; ++q) {
                S->lo[q] = ord_dec(v[q]);
                S->hi[q] = ord_dec(v[3 + q]);
            }
            S->n = se - sb;
        }
        return m;
    }
    const uint32_t m = Lo->n;
    if (partition) {
        const int a = best_key / kSahBins, bin = best_key % kSahBins;
        uint32_t pl = 0u, pr = m;
        for (uint32_t base = 0; base < cnt; base += 64u) {
            const uint32_t i = base + lane;
            const bool live = i < cnt;
            const uint32_t id = live ? W.ids[b + i] : 0u;
            bool left = false;
            if (live) {
                const float4 c = cent[id];
                const float cc = a == 0 ? c.x : (a == 1 ? c.y : c.z);
                left = gs_bin(cc, alo[a], ascale[a]) <= bin;
            }
            const unsigned long long ml = __ballot(left), mr = __ballot(live && !left);
            const unsigned long long below = (1ull << lane) - 1ull;
            if (left) W.tmp[pl + (uint32_t)__popcll(ml & below)] = id;
            else if (live) W.tmp[pr + (uint32_t)__popcll(mr & below)] = id;
            pl += (uint32_t)__popcll(ml);
            pr += (uint32_t)__popcll(mr);
        }
        gs_wave_sync();
        for (uint32_t i = lane; i < cnt; i += 64u) W.ids[b + i] = W.tmp[i];
        gs_wave_sync();
    }
    return m;
}

// Builds each small range's subtree with one wave (4 waves per workgroup):
// the host's job loop (sah_subtree) with its leaf rule.
__global__ __launch_bounds__(256) void k_gs_small(const GsSeg* __restrict__ segs, uint32_t nseg,
                                                  const float4* __restrict__ cent, const float4* __restrict__ plo,
                                                  const float4* __restrict__ phi, uint32_t* __restrict__ ids,
                                                  BvhNode* __restrict__ nodes, uint32_t* __restrict__ counters) {
    __shared__ GsWave waves[4];
    const int wv = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63u);
    const uint32_t s = blockIdx.x * 4u + (uint32_t)wv;
    if (s >= nseg) return;
    GsWave& W = waves[wv];
    const GsSeg g = segs[s];
    const uint32_t n = g.e - g.b;
    for (uint32_t i = lane; i < n; i += 64u) W.ids[i] = ids[g.b + i];
    if (lane == 0) {
        GsJob& j = W.jobs[0];
        j.b = 0u;
        j.e = n;
        j.parent = g.parent;
        j.sd = g.sd;
        for (int q = 0; q < 3; ++q) {
            j.lo[q] = g.lo[q];
            j.hi[q] = g.hi[q];
        }
    }
    int top = 1;
    uint32_t max_depth = 0;
    gs_wave_sync();
    while (top > 0) {
        --top;
        const GsJob j = W.jobs[top];
        gs_wave_sync();
        const uint32_t cnt = j.e - j.b;
        const uint32_t depth = j.sd >> 1;
        GsSide L, R;
        if (cnt == 1u) {  // always a leaf (the root range holds at least two triangles)
            if (lane == 0) gs_set_link(nodes, j.parent, (int)(j.sd & 1u), ~(int)((g.b + j.b) * 16u));
            max_depth = max(max_depth, depth);
            continue;
        }
        // the split (the range reordered left side first: a leaf's triangle
        // set does not depend on the order)
        const uint32_t m = gs_wave_split(W, j.b, cnt, cent, plo, phi, &L, &R, true);
        if (j.parent >= 0 && cnt <= (uint32_t)kMaxLeafTris) {
            // a leaf if that is cheaper than the best split
            const float area = gs_area(j.lo, j.hi);
            const float split =
                kSahNode * area + kSahTri * ((float)m * gs_area(L.lo, L.hi) + (float)(cnt - m) * gs_area(R.lo, R.hi));
            if (kSahTri * (float)cnt * area <= split) {
                if (lane == 0)
                    gs_set_link(nodes, j.parent, (int)(j.sd & 1u), ~(int)((g.b + j.b) * 16u + (cnt - 1u)));
                max_depth = max(max_depth, depth);
                continue;
            }
        }
        uint32_t k = 0;
        if (lane == 0) k = atomicAdd(&counters[0], 1u);
        k = (uint32_t)__shfl((int)k, 0);
        if (lane == 0) {
            BvhNode nd;
            nd.a = nd.b = nd.c = make_float4(0.f, 0.f, 0.f, 0.f);
            nd.d = make_int4(0, 0, 0, 0);
            gs_set_child(&nd, 0, L.lo, L.hi);
            gs_set_child(&nd, 1, R.lo, R.hi);
            nodes[k] = nd;
            if (j.parent >= 0) gs_set_link(nodes, j.parent, (int)(j.sd & 1u), (int)k);
            GsJob cl, cr;
            cl.b = j.b;
            cl.e = j.b + m;
            cr.b = j.b + m;
            cr.e = j.e;
            cl.parent = cr.parent = (int)k;
            cl.sd = 0u | ((depth + 1u) << 1);
            cr.sd = 1u | ((depth + 1u) << 1);
            for (int q = 0; q < 3; ++q) {
                cl.lo[q] = L.lo[q];
                cl.hi[q] = L.hi[q];
                cr.lo[q] = R.lo[q];
                cr.hi[q] = R.hi[q];
            }
            const bool left_small = m <= cnt - m;
            W.jobs[top] = left_small ? cr : cl;      // the larger waits
            W.jobs[top + 1] = left_small ? cl : cr;  // the smaller goes next
        }
        top += 2;
        gs_wave_sync();
    }
    for (uint32_t i = lane; i < n; i += 64u) ids[g.b + i] = W.ids[i];
    for (int off = 32; off > 0; off >>= 1) max_depth = max(max_depth, (uint32_t)__shfl_xor((int)max_depth, off));
    if (lane == 0) atomicMax(&counters[3], max_depth);
}

__global__ void k_gs_gather_tris(const BvhTri* __restrict__ ctris, const uint32_t* __restrict__ ids, uint32_t n,
                                 BvhTri* __restrict__ out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) out[i] = ctris[ids[i]];
}

}  // namespace

hipError_t build_sah_gpu(const float* vertices, const uint32_t* indices, const uint2* offsets,
                         const uint32_t* inst_first_tri, uint32_t instance_count, uint32_t tri_count, hipStream_t st,
                         BvhBuildResult* out, char* err, size_t err_len) {
    if (tri_count < 2)
        return build_lbvh(vertices, indices, offsets, inst_first_tri, instance_count, tri_count, st, out, err,
                          err_len);
    hipError_t rc = hipSuccess;
    const uint32_t n = tri_count;
    const uint32_t big_cap = n / kGsSmall + 2u;
    const uint32_t chunk_cap = n / kGsChunk + big_cap + 2u;
    BvhTri* ctris = nullptr;
    float4 *cent = nullptr, *plo = nullptr, *phi = nullptr;
    int* bounds = nullptr;
    uint32_t *ids = nullptr, *ids_tmp = nullptr, *counters = nullptr, *chunk_left = nullptr, *chunk_off = nullptr,
             *seg_first = nullptr;
    GsSeg *segs = nullptr, *segs_next = nullptr, *small = nullptr;
    GsBins* bins = nullptr;
    GsSplit* split = nullptr;
    GsChunk* chunks = nullptr;
    BvhNode* nodes = nullptr;
    BvhTri* tris = nullptr;
    const int init_bounds[12] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF, (int)0x80000000, (int)0x80000000,
                                 (int)0x80000000, 0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF, (int)0x80000000,
                                 (int)0x80000000, (int)0x80000000};
    int hb[12];
    uint32_t hc[4] = {0u, 0u, 0u, 0u};
    float pad = 0.0f, ext = 0.0f, mag = 0.0f;
    std::vector<GsSeg> level;
    std::vector<GsChunk> hchunks;
    std::vector<uint32_t> hfirst;
    GsSeg root;
    const auto t0 = std::chrono::steady_clock::now();
    XR_TRY(hipMalloc(&ctris, sizeof(BvhTri) * n));
    XR_TRY(hipMalloc(&cent, sizeof(float4) * n));
    XR_TRY(hipMalloc(&plo, sizeof(float4) * n));
    XR_TRY(hipMalloc(&phi, sizeof(float4) * n));
    XR_TRY(hipMalloc(&bounds, sizeof(int) * 12));
    XR_TRY(hipMalloc(&ids, sizeof(uint32_t) * n));
    XR_TRY(hipMalloc(&ids_tmp, sizeof(uint32_t) * n));
    XR_TRY(hipMalloc(&counters, sizeof(uint32_t) * 4));
    XR_TRY(hipMalloc(&segs, sizeof(GsSeg) * big_cap));
    XR_TRY(hipMalloc(&segs_next, sizeof(GsSeg) * big_cap));
    XR_TRY(hipMalloc(&small, sizeof(GsSeg) * n));
    XR_TRY(hipMalloc(&bins, sizeof(GsBins) * big_cap));
    XR_TRY(hipMalloc(&split, sizeof(GsSplit) * big_cap));
    XR_TRY(hipMalloc(&chunks, sizeof(GsChunk) * chunk_cap));
    XR_TRY(hipMalloc(&chunk_left, sizeof(uint32_t) * chunk_cap));
    XR_TRY(hipMalloc(&chunk_off, sizeof(uint32_t) * chunk_cap));
    XR_TRY(hipMalloc(&seg_first, sizeof(uint32_t) * (big_cap + 1)));
    XR_TRY(hipMalloc(&nodes, sizeof(BvhNode) * n));
    XR_TRY(hipMemcpyAsync(bounds, init_bounds, sizeof(init_bounds), hipMemcpyHostToDevice, st));
    XR_TRY(hipMemsetAsync(counters, 0, sizeof(uint32_t) * 4, st));
    launch_gather(vertices, indices, offsets, inst_first_tri, instance_count, n, ctris, cent, bounds, st);
    XR_TRY(hipGetLastError());
    XR_TRY(hipMemcpyAsync(hb, bounds, sizeof(hb), hipMemcpyDeviceToHost, st));
    XR_TRY(hipStreamSynchronize(st));
    // box_pad, as the host builder computes it
    for (int a = 0; a < 3; ++a) {
        const float lo = ord_dec_host(hb[6 + a]), hi = ord_dec_host(hb[9 + a]);
        ext = fmaxf(ext, hi - lo);
        mag = fmaxf(mag, fmaxf(fabsf(lo), fabsf(hi)));
    }
    pad = fmaxf(ext * (1.0f / 65536.0f), mag * (1.0f / 1048576.0f));
    hipLaunchKernelGGL(k_gs_prims, dim3((n + 255) / 256), dim3(256), 0, st, ctris, n, pad, plo, phi, ids);
    XR_TRY(hipGetLastError());
    root.b = 0u;
    root.e = n;
    root.parent = -1;
    root.sd = 0u;
    for (int a = 0; a < 3; ++a) {  // the root's box is never stored
        root.lo[a] = 0.0f;
        root.hi[a] = 0.0f;
    }
    if (n > kGsSmall) {
        level.push_back(root);
    } else {
        XR_TRY(hipMemcpyAsync(small, &root, sizeof(GsSeg), hipMemcpyHostToDevice, st));
        hc[2] = 1u;
        XR_TRY(hipMemcpyAsync(counters + 2, &hc[2], sizeof(uint32_t), hipMemcpyHostToDevice, st));
    }
    while (!level.empty()) {
        const uint32_t ns = (uint32_t)level.size();
        hchunks.clear();
        hfirst.assign(ns + 1, 0u);
        for (uint32_t s = 0; s < ns; ++s) {
            hfirst[s] = (uint32_t)hchunks.size();
            for (uint32_t b = level[s].b; b < level[s].e; b += kGsChunk)
                hchunks.push_back(GsChunk{s, b, std::min(b + kGsChunk, level[s].e), hfirst[s]});
        }
        hfirst[ns] = (uint32_t)hchunks.size();
        const uint32_t nc = (uint32_t)hchunks.size();
        XR_TRY(hipMemcpyAsync(segs, level.data(), sizeof(GsSeg) * ns, hipMemcpyHostToDevice, st));
        XR_TRY(hipMemcpyAsync(chunks, hchunks.data(), sizeof(GsChunk) * nc, hipMemcpyHostToDevice, st));
        XR_TRY(hipMemcpyAsync(seg_first, hfirst.data(), sizeof(uint32_t) * (ns + 1), hipMemcpyHostToDevice, st));
        XR_TRY(hipMemsetAsync(counters + 1, 0, sizeof(uint32_t), st));
        hipLaunchKernelGGL(k_gs_init_bins, dim3(ns), dim3(64), 0, st, bins, ns);
        hipLaunchKernelGGL(k_gs_cbounds, dim3(nc), dim3(256), 0, st, chunks, ids, cent, bins);
        hipLaunchKernelGGL(k_gs_bin, dim3(nc), dim3(256), 0, st, chunks, ids, cent, plo, phi, bins);
        hipLaunchKernelGGL(k_gs_split, dim3(ns), dim3(64), 0, st, segs, ns, bins, split, nodes, counters, segs_next,
                           small);
        hipLaunchKernelGGL(k_gs_count, dim3(nc), dim3(256), 0, st, chunks, ids, cent, split, chunk_left);
        hipLaunchKernelGGL(k_gs_chunk_scan, dim3(ns), dim3(256), 0, st, chunks, seg_first, chunk_left, chunk_off);
        hipLaunchKernelGGL(k_gs_scatter, dim3(nc), dim3(256), 0, st, chunks, segs, ids, cent, split, chunk_off,
                           ids_tmp);
        hipLaunchKernelGGL(k_gs_copy_back, dim3(nc), dim3(256), 0, st, chunks, ids_tmp, ids);
        hipLaunchKernelGGL(k_gs_median_boxes, dim3(ns), dim3(64), 0, st, segs, ns, split, ids, plo, phi, nodes,
                           segs_next, small, counters);
        XR_TRY(hipGetLastError());
        XR_TRY(hipMemcpyAsync(hc, counters, sizeof(hc), hipMemcpyDeviceToHost, st));
        XR_TRY(hipStreamSynchronize(st));
        level.resize(hc[1]);
        if (hc[1]) {
            XR_TRY(hipMemcpyAsync(level.data(), segs_next, sizeof(GsSeg) * hc[1], hipMemcpyDeviceToHost, st));
            XR_TRY(hipStreamSynchronize(st));
        }
    }
    XR_TRY(hipMemcpyAsync(hc, counters, sizeof(hc), hipMemcpyDeviceToHost, st));
    XR_TRY(hipStreamSynchronize(st));
    if (hc[2]) hipLaunchKernelGGL(k_gs_small, dim3((hc[2] + 3u) / 4u), dim3(256), 0, st, small, hc[2], cent, plo,
                                  phi, ids, nodes, counters);
    XR_TRY(hipGetLastError());
    XR_TRY(hipMalloc(&tris, sizeof(BvhTri) * n));
    hipLaunchKernelGGL(k_gs_gather_tris, dim3((n + 255) / 256), dim3(256), 0, st, ctris, ids, n, tris);
    XR_TRY(hipGetLastError());
    XR_TRY(hipMemcpyAsync(hc, counters, sizeof(hc), hipMemcpyDeviceToHost, st));
    XR_TRY(hipStreamSynchronize(st));
    out->nodes = nodes;
    out->tris = tris;
    out->node_count = hc[0];
    out->tri_count = n;
    out->max_depth = hc[3];
    out->build_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    nodes = nullptr;
    tris = nullptr;
    rc = hipSuccess;
    goto done;
fail:
    rc = hipErrorUnknown;
done:
    hipFree(ctris);
    hipFree(cent);
    hipFree(plo);
    hipFree(phi);
    hipFree(bounds);
    hipFree(ids);
    hipFree(ids_tmp);
    hipFree(counters);
    hipFree(segs);
    hipFree(segs_next);
    hipFree(small);
    hipFree(bins);
    hipFree(split);
    hipFree(chunks);
    hipFree(chunk_left);
    hipFree(chunk_off);
    hipFree(seg_first);
    hipFree(nodes);
    hipFree(tris);
    return rc;
}

// ---------------------------------------------------------------------------
// Wide nodes.  The binary tree is collapsed top-down, one level of wide nodes
// per pass: each item (a binary node and the stack entries its ancestors can
// leave) opens its children greedily, always the inner child with the largest
// box surface, until it holds `width` children; a first kernel counts the
// inner children, a scan gives every node's first child slot on the next
// level, and a second kernel writes the node (child boxes rounded outwards
// onto the node's 8-bit grid) and the next level's items.
// Children of one node are consecutive, levels are stored one after another.
// ---------------------------------------------------------------------------
namespace {

struct WideSlots {
    float lo[4][3], hi[4][3];
    int link[4];
    int count;
};

__device__ __forceinline__ void binary_child(const BvhNode& n, int side, float lo[3], float hi[3], int* link) {
    if (side == 0) {
        lo[0] = n.a.x; hi[0] = n.a.y; lo[1] = n.a.z; hi[1] = n.a.w; lo[2] = n.c.x; hi[2] = n.c.y;
        *link = n.d.x;
    } else {
        lo[0] = n.b.x; hi[0] = n.b.y; lo[1] = n.b.z; hi[1] = n.b.w; lo[2] = n.c.z; hi[2] = n.c.w;
        *link = n.d.y;
    }
}

__device__ void open_wide(const BvhNode* __restrict__ bn, int b, int width, WideSlots* s) {
    const BvhNode n = bn[b];
    binary_child(n, 0, s->lo[0], s->hi[0], &s->link[0]);
    binary_child(n, 1, s->lo[1], s->hi[1], &s->link[1]);
    int cnt = 2;
    while (cnt < width) {
        int best = -1;
        float area = -1.0f;
        for (int k = 0; k < cnt; ++k) {
            if (s->link[k] < 0) continue;
            const float dx = s->hi[k][0] - s->lo[k][0], dy = s->hi[k][1] - s->lo[k][1],
                        dz = s->hi[k][2] - s->lo[k][2];
            const float a = dx * dy + dy * dz + dz * dx;
            if (a > area) {
                area = a;
                best = k;
            }
        }
        if (best < 0) break;
        const BvhNode m = bn[s->link[best]];
        binary_child(m, 0, s->lo[best], s->hi[best], &s->link[best]);
        binary_child(m, 1, s->lo[cnt], s->hi[cnt], &s->link[cnt]);
        ++cnt;
    }
    s->count = cnt;
}

// items[i] = (binary node, stack entries above it)
__global__ void k_wide_count(const BvhNode* __restrict__ bn, const int2* __restrict__ items, int count, int width,
                             int* __restrict__ inner) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    WideSlots s;
    open_wide(bn, items[i].x, width, &s);
    int m = 0;
    for (int k = 0; k < s.count; ++k) m += s.link[k] >= 0 ? 1 : 0;
    inner[i] = m;
}

// Grid of one axis of a wide node: origin lo and the smallest power-of-two
// step with lo + 255 * step >= hi in float arithmetic; returns the step's
// biased exponent (1..254, so the step is a normal float).
__device__ uint32_t grid_exponent(float lo, float hi) {
    uint32_t eb = 1;
    const float ext = hi - lo;
    if (ext > 0.0f) {
        const uint32_t bits = __float_as_uint(ext * (1.0f / 255.0f));
        eb = bits >> 23;
        if (bits & 0x7FFFFFu) ++eb;  // round up to a power of two
        if (eb < 1) eb = 1;
    }
    while (eb < 254 && lo + 255.0f * __uint_as_float(eb << 23) < hi) ++eb;
    return eb;
}

// Grid planes of a child box, rounded outwards and checked in the arithmetic
// that decodes them (origin + g * step; g * step is exact).
__device__ uint32_t grid_below(float v, float origin, float step) {
    float g = fminf(fmaxf(floorf((v - origin) / step), 0.0f), 255.0f);
    while (g > 0.0f && origin + g * step > v) g -= 1.0f;
    return (uint32_t)g;
}
__device__ uint32_t grid_above(float v, float origin, float step) {
    float g = fminf(fmaxf(ceilf((v - origin) / step), 0.0f), 255.0f);
    while (g < 255.0f && origin + g * step < v) g += 1.0f;
    return (uint32_t)g;
}

__global__ void k_wide_emit(const BvhNode* __restrict__ bn, const int2* __restrict__ items, int count, int width,
                            uint32_t self_base, uint32_t next_base, const int* __restrict__ inner,
                            const int* __restrict__ incl, WideNode* __restrict__ out, int2* __restrict__ next,
                            uint32_t* __restrict__ bound) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    WideSlots s;
    open_wide(bn, items[i].x, width, &s);
    const int acc = items[i].y + s.count - 1;
    int slot = incl[i] - inner[i];
    float nlo[3], nhi[3];
    for (int a = 0; a < 3; ++a) {
        nlo[a] = s.lo[0][a];
        nhi[a] = s.hi[0][a];
        for (int k = 1; k < s.count; ++k) {
            nlo[a] = fminf(nlo[a], s.lo[k][a]);
            nhi[a] = fmaxf(nhi[a], s.hi[k][a]);
        }
    }
    uint32_t eb[3];
    float step[3];
    for (int a = 0; a < 3; ++a) {
        eb[a] = grid_exponent(nlo[a], nhi[a]);
        step[a] = __uint_as_float(eb[a] << 23);
    }
    uint32_t planes[6] = {0, 0, 0, 0, 0, 0};  // lox, hix, loy, hiy, loz, hiz; byte k = child k
    int link[4];
    for (int k = 0; k < 4; ++k) {
        if (k < s.count) {
            for (int a = 0; a < 3; ++a) {
                planes[2 * a] |= grid_below(s.lo[k][a], nlo[a], step[a]) << (8 * k);
                planes[2 * a + 1] |= grid_above(s.hi[k][a], nlo[a], step[a]) << (8 * k);
            }
            if (s.link[k] >= 0) {
                next[slot] = make_int2(s.link[k], acc);
                link[k] = (int)(next_base + (uint32_t)slot);
                ++slot;
            } else {
                link[k] = s.link[k];
            }
        } else {  // empty box (lo 255 > hi 0 on every axis), leaf of triangle 0
            for (int a = 0; a < 3; ++a) planes[2 * a] |= 255u << (8 * k);
            link[k] = -1;
        }
    }
    WideNode w;
    w.q = make_float4(nlo[0], nlo[1], nlo[2], step[0]);
    w.p = make_uint4(planes[0], planes[1], planes[2], planes[3]);
    w.r = make_uint4(planes[4], planes[5], __float_as_uint(step[1]), __float_as_uint(step[2]));
    w.link = make_int4(link[0], link[1], link[2], link[3]);
    out[self_base + (uint32_t)i] = w;
    atomicMax(bound, (uint32_t)acc);
}

}  // namespace

hipError_t build_wide(const BvhNode* binary, uint32_t binary_count, int width, hipStream_t st, WideBuildResult* out,
                      char* err, size_t err_len) {
    hipError_t rc = hipSuccess;
    const uint32_t cap = binary_count > 0 ? binary_count : 1u;
    WideNode* wide = nullptr;
    WideNode* exact = nullptr;
    int2* items[2] = {nullptr, nullptr};
    int* inner = nullptr;
    int* incl = nullptr;
    uint32_t* bound = nullptr;
    void* temp = nullptr;
    size_t temp_bytes = 0;
    const int2 root = make_int2(0, 0);
    uint32_t self_base = 0, total = 1, count = 1, depth = 0, host_bound = 0;
    int cur = 0;
    const auto t0 = std::chrono::steady_clock::now();
    if (width < 2 || width > 4 || binary_count == 0) {
        snprintf(err, err_len, "bad wide-tree request (width %d, %u nodes)", width, binary_count);
        return hipErrorInvalidValue;
    }
    XR_TRY(hipMalloc(&wide, sizeof(WideNode) * cap));
    XR_TRY(hipMalloc(&items[0], sizeof(int2) * cap));
    XR_TRY(hipMalloc(&items[1], sizeof(int2) * cap));
    XR_TRY(hipMalloc(&inner, sizeof(int) * cap));
    XR_TRY(hipMalloc(&incl, sizeof(int) * cap));
    XR_TRY(hipMalloc(&bound, sizeof(uint32_t)));
    XR_TRY(hipMemsetAsync(bound, 0, sizeof(uint32_t), st));
    XR_TRY(hipMemcpyAsync(items[0], &root, sizeof(int2), hipMemcpyHostToDevice, st));
    XR_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, temp_bytes, inner, incl, (int)cap, st));
    XR_TRY(hipMalloc(&temp, temp_bytes));
    while (count > 0) {
        const uint32_t blocks = (count + 255) / 256;
        int last = 0;
        hipLaunchKernelGGL(k_wide_count, dim3(blocks), dim3(256), 0, st, binary, items[cur], (int)count, width, inner);
        XR_TRY(hipGetLastError());
        XR_TRY(hipcub::DeviceScan::InclusiveSum(temp, temp_bytes, inner, incl, (int)count, st));
        hipLaunchKernelGGL(k_wide_emit, dim3(blocks), dim3(256), 0, st, binary, items[cur], (int)count, width,
                           self_base, total, inner, incl, wide, items[cur ^ 1], bound);
        XR_TRY(hipGetLastError());
        XR_TRY(hipMemcpyAsync(&last, incl + (count - 1), sizeof(int), hipMemcpyDeviceToHost, st));
        XR_TRY(hipStreamSynchronize(st));
        ++depth;
        // every wide node consumes at least one binary node of its own
        if (last < 0 || (uint64_t)total + (uint32_t)last > cap || depth > cap) {
            snprintf(err, err_len, "malformed binary tree (%u wide nodes at level %u)", total, depth);
            goto fail;
        }
        self_base = total;
        total += (uint32_t)last;
        count = (uint32_t)last;
        cur ^= 1;
    }
    XR_TRY(hipMalloc(&exact, sizeof(WideNode) * total));
    XR_TRY(hipMemcpyAsync(exact, wide, sizeof(WideNode) * total, hipMemcpyDeviceToDevice, st));
    XR_TRY(hipMemcpyAsync(&host_bound, bound, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    XR_TRY(hipStreamSynchronize(st));
    out->nodes = exact;
    out->node_count = total;
    out->depth = depth;
    out->stack_bound = host_bound;
    out->build_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    exact = nullptr;
    rc = hipSuccess;
    goto done;
fail:
    rc = hipErrorUnknown;
done:
    hipFree(wide);
    hipFree(exact);
    hipFree(items[0]);
    hipFree(items[1]);
    hipFree(inner);
    hipFree(incl);
    hipFree(bound);
    hipFree(temp);
    return rc;
}

}  // namespace xr

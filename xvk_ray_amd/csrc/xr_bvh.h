// Acceleration structure layout and wave64 traversal for gfx950.
//
// Replaces the driver-side TraceRay of the reference (KHR ray tracing; BLAS
// built in ApplicationRT.cpp:186-224, TLAS with identity instances and
// CULL_DISABLE in XVKTLAccelerationStructure.cpp:68-93).  CDNA has no RT
// cores, so traversal is plain VALU code: a four-child BVH with quantised
// child boxes (64 bytes per visited node, collapsed from the binary tree the
// builders make), a per-lane traversal stack kept in LDS, and the watertight
// ray/triangle test of Woop, Benthin and Wald (JCGT 2013).
//
// Closest-hit rule (order independent, so any BVH and the brute-force check
// agree): the hit with the smallest t in [tmin, tmax]; equal t goes to the
// smaller canonical triangle id.  The canonical id of triangle p of instance i
// is (first triangle of instance i) + p, i.e. the position of the triangle in
// the concatenated index buffer divided by 3.
#pragma once
#include "xr_math.h"

namespace xr {

// 64-byte node: the boxes of both children plus their links.
//   a = (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y)
//   b = (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y)
//   c = (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)
//   d = (link0, link1, -, -)
// link >= 0: index of an internal node.  link < 0: leaf, ~link = first*16 +
// (count-1) over the BVH-ordered triangle array.  Every node has two real
// children; a one-triangle scene gets a root whose two links name the same
// leaf (the tie rule keeps the second visit from changing the hit).
struct alignas(16) BvhNode {
    float4 a, b, c;
    int4 d;
};
static_assert(sizeof(BvhNode) == 64, "BVH node must be one 64-byte line");

// The tree the kernels traverse: the binary tree above with up to four
// children per node (each node opens its largest-area inner children until it
// has four), in 64 bytes per node: the child boxes are quantised to 8 bits
// per plane on a per-node grid (Ylitie, Karras and Laine, "Efficient
// Incoherent Ray Traversal on GPUs Through Compressed Wide BVHs", HPG 2017).
//   q.xyz = grid origin (floats, the min corner of the union of the children)
//   q.w   = grid step on x, a power of two (float)
//   p     = (lox, hix, loy, hiy): byte k holds child k's plane on that grid,
//   r     = (loz, hiz, step on y, step on z) (steps as float bits)
//   link  = (link0 .. link3): >= 0 inner node, < 0 leaf ~(first*16 + count-1)
// Slots past the child count hold the empty box lo = 255, hi = 0 on every
// axis and link -1 (the leaf of BVH triangle 0).  Their near plane lies 255
// steps beyond their far plane, which the roundings of the slab test (far
// distance scaled by 1 + 2*gamma(3)) cannot undo for a ray whose origin is
// within 2^29 grid steps of the node: 64 scene extents, a step being at
// least 2^-23 of the extent since every box is padded.  A ray from farther
// away that entered one would only test triangle 0, a real triangle, which
// cannot change what it finds.  The traversal therefore needs no child
// count.
// A child's box is [origin + lo * step, origin + hi * step] per axis, rounded
// outwards by the builder so it holds the child's padded box.  Traversal reads
// four 16-byte words per node; its cost follows the bytes fetched per visit
// (the 128-byte uncompressed form halved the visits of the binary tree and
// gained 3 %), and the quantised node halves those bytes.
struct alignas(64) WideNode {
    float4 q;
    uint4 p;
    uint4 r;
    int4 link;
};
static_assert(sizeof(WideNode) == 64, "wide node must be 64 bytes");

// BVH-ordered triangle: three float4.  v0.w = canonical triangle id (bits),
// v1.w = instance index (bits), v2.w = primitive index within the instance.
struct alignas(16) BvhTri {
    float4 v0, v1, v2;
};

struct Hit {
    float t;      // -1 on miss
    float u, v;   // hit attributes: weights of v1 and v2
    uint32_t id;  // canonical triangle id, 0xFFFFFFFF on miss
    uint32_t idx; // position in the BVH-ordered triangle array
};

// Per-ray constants of the watertight test (permutation and shear).
struct RayPre {
    v3 org;
    v3 idir;
    v3 oi;  // org * idir: a slab plane's distance is fma(plane, idir, -oi)
    int kx, ky, kz;
    float sx, sy, sz;
};

XR_HD float comp(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

XR_HD RayPre make_ray_pre(v3 org, v3 dir) {
    RayPre p;
    p.org = org;
    // Reciprocal direction for the slab test; exact zeros become +-1e-20 so
    // the slab products never form 0*inf.
    float dx = fabsf(dir.x) > 1e-20f ? dir.x : copysignf(1e-20f, dir.x);
    float dy = fabsf(dir.y) > 1e-20f ? dir.y : copysignf(1e-20f, dir.y);
    float dz = fabsf(dir.z) > 1e-20f ? dir.z : copysignf(1e-20f, dir.z);
    p.idir = mk3(1.0f / dx, 1.0f / dy, 1.0f / dz);
    p.oi = mk3(org.x * p.idir.x, org.y * p.idir.y, org.z * p.idir.z);
    float ax = fabsf(dir.x), ay = fabsf(dir.y), az = fabsf(dir.z);
    int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (comp(dir, kz) < 0.0f) { int t = kx; kx = ky; ky = t; }
    p.kx = kx; p.ky = ky; p.kz = kz;
    float dkz = comp(dir, kz);
    p.sx = comp(dir, kx) / dkz;
    p.sy = comp(dir, ky) / dkz;
    p.sz = 1.0f / dkz;
    return p;
}

// Watertight ray/triangle test.  Returns true and t when the ray hits the
// triangle with tmin <= t <= tmax, plus the scaled barycentrics V, W and det:
// the hit attributes are u = V / det, v = W / det, divided out by the caller
// for the one hit it keeps.  Same operation sequence on host and device (no
// contraction), so both produce identical bits.
XR_HD bool intersect_tri(const RayPre& r, v3 p0, v3 p1, v3 p2, float tmin, float tmax, float* t_out,
                           float* V_out, float* W_out, float* det_out) {
    v3 A = sub(p0, r.org), B = sub(p1, r.org), C = sub(p2, r.org);
    float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
    float Ax = fmaf(-r.sx, Akz, comp(A, r.kx));
    float Ay = fmaf(-r.sy, Akz, comp(A, r.ky));
    float Bx = fmaf(-r.sx, Bkz, comp(B, r.kx));
    float By = fmaf(-r.sy, Bkz, comp(B, r.ky));
    float Cx = fmaf(-r.sx, Ckz, comp(C, r.kx));
    float Cy = fmaf(-r.sy, Ckz, comp(C, r.ky));
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        // Edge-on case: redo the edge functions in FP64 (exact products).
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = r.sz * Akz, Bz = r.sz * Bkz, Cz = r.sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    float t = T / det;
    if (!(t >= tmin && t <= tmax)) return false;
    *t_out = t;
    *V_out = V;
    *W_out = W;
    *det_out = det;
    return true;
}

// Robust slab test against one box (Ize, "Robust BVH Ray Traversal", JCGT
// 2013): (lo - o) * (1/d) per axis, far distance scaled by 1 + 2*gamma(3), so
// no box that contains a hit is culled.  Returns the entry distance or +inf.
XR_HD float slab(const RayPre& r, float lox, float hix, float loy, float hiy, float loz, float hiz,
                 float tmin, float tmax) {
    float t0x = (lox - r.org.x) * r.idir.x, t1x = (hix - r.org.x) * r.idir.x;
    float t0y = (loy - r.org.y) * r.idir.y, t1y = (hiy - r.org.y) * r.idir.y;
    float t0z = (loz - r.org.z) * r.idir.z, t1z = (hiz - r.org.z) * r.idir.z;
    float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
    float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
    tf = fminf(tf * 1.0000003576f, tmax);
    return tn <= tf ? tn : __builtin_huge_valf();
}

// The same test returning the hit flag, with the entry distance in *tn_out.
XR_HD bool slab_hit(const RayPre& r, float lox, float hix, float loy, float hiy, float loz, float hiz, float tmin,
                    float tmax, float* tn_out) {
    float t0x = (lox - r.org.x) * r.idir.x, t1x = (hix - r.org.x) * r.idir.x;
    float t0y = (loy - r.org.y) * r.idir.y, t1y = (hiy - r.org.y) * r.idir.y;
    float t0z = (loz - r.org.z) * r.idir.z, t1z = (hiz - r.org.z) * r.idir.z;
    float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
    float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
    tf = fminf(tf * 1.0000003576f, tmax);
    *tn_out = tn;
    return tn <= tf;
}

}  // namespace xr

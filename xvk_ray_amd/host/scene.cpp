// Synthetic scenes in the reference's AScene layout (src/Asset/Scene.cpp:16-87):
// one concatenated vertex buffer (AVertex, 36 B), one index buffer, one
// material buffer, and (indexOffset, vertexOffset) per instance.  Emitters
// follow the reference convention (DiffuseLight material, radiance =
// baseColor * 4, BSDF.hlsl:27-32), so a light of radiance R has baseColor R/4.
//
// All randomness comes from one splitmix64 stream per scene, seeded by the
// caller, so a seed names a scene exactly.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <map>
#include <new>
#include <vector>

#include "scene_internal.h"

namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform in [lo, hi)
    float uni(float lo, float hi) { return lo + (hi - lo) * (float)((next() >> 40) * (1.0 / 16777216.0)); }
};

xr_material lambert(float r, float g, float b) {
    xr_material m;
    memset(&m, 0, sizeof(m));
    m.base_color[0] = r;
    m.base_color[1] = g;
    m.base_color[2] = b;
    m.base_color[3] = 1.0f;
    m.base_color_texture_index = -1;
    m.material_model = XR_MATERIAL_LAMBERTIAN;
    return m;
}

xr_material light(float radiance) {
    xr_material m = lambert(radiance / 4.0f, radiance / 4.0f, radiance / 4.0f);
    m.material_model = XR_MATERIAL_DIFFUSE_LIGHT;
    return m;
}

struct Mesh {
    std::vector<xr_vertex> v;
    std::vector<uint32_t> i;
};

void add_instance(xrh_scene* s, const Mesh& m) {
    s->offsets.push_back((uint32_t)s->indices.size());
    s->offsets.push_back((uint32_t)s->vertices.size());
    s->vertices.insert(s->vertices.end(), m.v.begin(), m.v.end());
    s->indices.insert(s->indices.end(), m.i.begin(), m.i.end());
}

xr_vertex vert(double x, double y, double z, double nx, double ny, double nz, float u, float v, int mat) {
    xr_vertex o;
    o.pos[0] = (float)x; o.pos[1] = (float)y; o.pos[2] = (float)z;
    o.normal[0] = (float)nx; o.normal[1] = (float)ny; o.normal[2] = (float)nz;
    o.tex_coord[0] = u; o.tex_coord[1] = v;
    o.material_index = mat;
    return o;
}

// Axis-aligned quad with corner c, edges e1, e2 and normal n (two triangles).
void quad(Mesh& m, const double c[3], const double e1[3], const double e2[3], const double n[3], int mat) {
    const uint32_t b = (uint32_t)m.v.size();
    m.v.push_back(vert(c[0], c[1], c[2], n[0], n[1], n[2], 0, 0, mat));
    m.v.push_back(vert(c[0] + e1[0], c[1] + e1[1], c[2] + e1[2], n[0], n[1], n[2], 1, 0, mat));
    m.v.push_back(vert(c[0] + e1[0] + e2[0], c[1] + e1[1] + e2[1], c[2] + e1[2] + e2[2], n[0], n[1], n[2], 1, 1, mat));
    m.v.push_back(vert(c[0] + e2[0], c[1] + e2[1], c[2] + e2[2], n[0], n[1], n[2], 0, 1, mat));
    const uint32_t t[6] = {b, b + 1, b + 2, b, b + 2, b + 3};
    m.i.insert(m.i.end(), t, t + 6);
}

// Icosphere: icosahedron subdivided `levels` times (20 * 4^levels triangles),
// vertices on the sphere, smooth normals.
Mesh icosphere(double cx, double cy, double cz, double r, int levels, int mat) {
    const double t = (1.0 + sqrt(5.0)) / 2.0;
    std::vector<double> p = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                             0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
    std::vector<uint32_t> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6,
                               7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10,
                               8, 6, 7, 9, 8, 1};
    auto norm = [&](uint32_t k) {
        double l = sqrt(p[3 * k] * p[3 * k] + p[3 * k + 1] * p[3 * k + 1] + p[3 * k + 2] * p[3 * k + 2]);
        p[3 * k] /= l; p[3 * k + 1] /= l; p[3 * k + 2] /= l;
    };
    for (uint32_t k = 0; k < 12; ++k) norm(k);
    for (int l = 0; l < levels; ++l) {
        std::map<uint64_t, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            const uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            const uint32_t k = (uint32_t)(p.size() / 3);
            for (int a2 = 0; a2 < 3; ++a2) p.push_back(0.5 * (p[3 * a + a2] + p[3 * b + a2]));
            norm(k);
            mid[key] = k;
            return k;
        };
        std::vector<uint32_t> nf;
        for (size_t q = 0; q < f.size(); q += 3) {
            const uint32_t a = f[q], b = f[q + 1], c = f[q + 2];
            const uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            const uint32_t tri[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nf.insert(nf.end(), tri, tri + 12);
        }
        f.swap(nf);
    }
    Mesh m;
    const size_t nv = p.size() / 3;
    for (size_t k = 0; k < nv; ++k) {
        const double nx = p[3 * k], ny = p[3 * k + 1], nz = p[3 * k + 2];
        const float u = (float)(0.5 + atan2(nz, nx) / (2 * M_PI)), v = (float)(0.5 - asin(ny) / M_PI);
        m.v.push_back(vert(cx + r * nx, cy + r * ny, cz + r * nz, nx, ny, nz, u, v, mat));
    }
    m.i = f;
    return m;
}

// Value-noise fBm: a seeded lattice of values in [-1,1], smoothstep
// interpolation, 4 octaves from frequency 4 with gain 1/2, scaled to
// [-amplitude, amplitude].
struct Fbm {
    uint64_t seed;
    double amplitude;
    double lattice(int64_t i, int64_t j, int oct) const {
        uint64_t h = seed ^ ((uint64_t)(i * 73856093) ^ (uint64_t)(j * 19349663) ^ ((uint64_t)oct * 83492791ull));
        Rng r{h};
        r.next();
        return (double)(r.next() >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
    double noise(double x, double y, int oct) const {
        const double fx = floor(x), fy = floor(y);
        const int64_t i = (int64_t)fx, j = (int64_t)fy;
        double tx = x - fx, ty = y - fy;
        tx = tx * tx * (3 - 2 * tx);
        ty = ty * ty * (3 - 2 * ty);
        const double a = lattice(i, j, oct), b = lattice(i + 1, j, oct), c = lattice(i, j + 1, oct),
                     d = lattice(i + 1, j + 1, oct);
        return (a * (1 - tx) + b * tx) * (1 - ty) + (c * (1 - tx) + d * tx) * ty;
    }
    double height(double x, double z) const {
        double sum = 0, amp = 1, freq = 4, norm = 0;
        for (int o = 0; o < 4; ++o) {
            sum += amp * noise(x * freq, z * freq, o);
            norm += amp;
            amp *= 0.5;
            freq *= 2;
        }
        return amplitude * sum / norm;
    }
};

}  // namespace

extern "C" {

int xrh_scene_cornell(uint64_t seed, xrh_scene** out) {
    if (!out) return XR_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    xrh_scene* s = new (std::nothrow) xrh_scene();
    if (!s) return XR_ERR_OUT_OF_MEMORY;
    Rng rng{seed};
    Mesh box;
    int mat = 0;
    auto next_mat = [&]() {
        s->materials.push_back(lambert(rng.uni(0.2f, 0.8f), rng.uni(0.2f, 0.8f), rng.uni(0.2f, 0.8f)));
        return mat++;
    };
    const double up[3] = {0, 1, 0}, down[3] = {0, -1, 0}, fwd[3] = {0, 0, -1}, px[3] = {1, 0, 0}, nx[3] = {-1, 0, 0};
    // floor (y = 0) and back wall (z = 1), two quads each
    for (int k = 0; k < 2; ++k) {
        const double c[3] = {0.5 * k, 0, 0}, e1[3] = {0.5, 0, 0}, e2[3] = {0, 0, 1};
        quad(box, c, e1, e2, up, next_mat());
    }
    for (int k = 0; k < 2; ++k) {
        const double c[3] = {0.5 * k, 0, 1}, e1[3] = {0.5, 0, 0}, e2[3] = {0, 1, 0};
        quad(box, c, e1, e2, fwd, next_mat());
    }
    // left (x = 0) and right (x = 1) walls, two quads each
    for (int k = 0; k < 2; ++k) {
        const double c[3] = {0, 0, 0.5 * k}, e1[3] = {0, 1, 0}, e2[3] = {0, 0, 0.5};
        quad(box, c, e1, e2, px, next_mat());
    }
    for (int k = 0; k < 2; ++k) {
        const double c[3] = {1, 0, 0.5 * k}, e1[3] = {0, 1, 0}, e2[3] = {0, 0, 0.5};
        quad(box, c, e1, e2, nx, next_mat());
    }
    // ceiling (y = 1): four quads framing the 0.2 x 0.2 opening at the centre
    {
        const double a = 0.4, b = 0.6;
        const double c0[3] = {0, 1, 0}, e10[3] = {1, 0, 0}, e20[3] = {0, 0, a};
        quad(box, c0, e10, e20, down, next_mat());
        const double c1[3] = {0, 1, b}, e11[3] = {1, 0, 0}, e21[3] = {0, 0, 1 - b};
        quad(box, c1, e11, e21, down, next_mat());
        const double c2[3] = {0, 1, a}, e12[3] = {a, 0, 0}, e22[3] = {0, 0, b - a};
        quad(box, c2, e12, e22, down, next_mat());
        const double c3[3] = {b, 1, a}, e13[3] = {1 - b, 0, 0}, e23[3] = {0, 0, b - a};
        quad(box, c3, e13, e23, down, next_mat());
    }
    add_instance(s, box);
    // light in the opening
    {
        Mesh lm;
        s->materials.push_back(light(15.0f));
        const int lmat = mat++;
        const double c[3] = {0.4, 1.0, 0.4}, e1[3] = {0.2, 0, 0}, e2[3] = {0, 0, 0.2};
        quad(lm, c, e1, e2, down, lmat);
        add_instance(s, lm);
    }
    // icosphere, 3 subdivisions = 1,280 triangles
    {
        s->materials.push_back(lambert(rng.uni(0.2f, 0.8f), rng.uni(0.2f, 0.8f), rng.uni(0.2f, 0.8f)));
        const int smat = mat++;
        add_instance(s, icosphere(0.5, 0.5, 0.5, 0.25, 3, smat));
    }
    *out = s;
    return XR_OK;
}

int xrh_scene_terrain(uint64_t seed, uint32_t grid, uint32_t n_spheres, uint32_t n_lights, xrh_scene** out) {
    if (!out || grid < 2 || grid > 8192) return XR_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    xrh_scene* s = new (std::nothrow) xrh_scene();
    if (!s) return XR_ERR_OUT_OF_MEMORY;
    try {
        Rng rng{seed};
        Fbm fbm{rng.next(), 0.1};
        // heightfield over [0,1]^2, material 0
        s->materials.push_back(lambert(0.5f, 0.5f, 0.5f));
        {
            Mesh hf;
            hf.v.reserve((size_t)grid * grid);
            const double h = 1.0 / (grid - 1);
            for (uint32_t j = 0; j < grid; ++j) {
                for (uint32_t i = 0; i < grid; ++i) {
                    const double x = i * h, z = j * h;
                    const double y = fbm.height(x, z);
                    const double dx = fbm.height(x + 0.5 * h, z) - fbm.height(x - 0.5 * h, z);
                    const double dz = fbm.height(x, z + 0.5 * h) - fbm.height(x, z - 0.5 * h);
                    double nx = -dx / h, ny = 1.0, nz = -dz / h;
                    const double l = sqrt(nx * nx + ny * ny + nz * nz);
                    nx /= l; ny /= l; nz /= l;
                    hf.v.push_back(vert(x, y, z, nx, ny, nz, (float)x, (float)z, 0));
                }
            }
            hf.i.reserve((size_t)(grid - 1) * (grid - 1) * 6);
            for (uint32_t j = 0; j + 1 < grid; ++j) {
                for (uint32_t i = 0; i + 1 < grid; ++i) {
                    const uint32_t a = j * grid + i, b = a + 1, c = a + grid, d = c + 1;
                    const uint32_t t[6] = {a, c, b, b, c, d};
                    hf.i.insert(hf.i.end(), t, t + 6);
                }
            }
            add_instance(s, hf);
        }
        // spheres, one instance each
        const Mesh unit = icosphere(0, 0, 0, 1, 3, 0);
        for (uint32_t k = 0; k < n_spheres; ++k) {
            const double r = rng.uni(0.02f, 0.1f);
            const double x = rng.uni(0.05f, 0.95f), z = rng.uni(0.05f, 0.95f);
            const double y = fbm.height(x, z) + r;
            s->materials.push_back(lambert(rng.uni(0.1f, 0.9f), rng.uni(0.1f, 0.9f), rng.uni(0.1f, 0.9f)));
            const int mat = (int)s->materials.size() - 1;
            Mesh m = unit;
            for (auto& v : m.v) {
                v.pos[0] = (float)(x + r * v.normal[0]);
                v.pos[1] = (float)(y + r * v.normal[1]);
                v.pos[2] = (float)(z + r * v.normal[2]);
                v.material_index = mat;
            }
            add_instance(s, m);
        }
        // lights: one instance holding every quad, facing down
        if (n_lights) {
            Mesh lm;
            const double down[3] = {0, -1, 0};
            for (uint32_t k = 0; k < n_lights; ++k) {
                const double size = rng.uni(0.03f, 0.08f);
                // above the camera band of xrh_camera_orbit (eye heights 0.32-0.45)
                const double x = rng.uni(0.0f, 1.0f), z = rng.uni(0.0f, 1.0f), y = rng.uni(0.6f, 0.9f);
                s->materials.push_back(light(rng.uni(5.0f, 50.0f)));
                const int mat = (int)s->materials.size() - 1;
                const double c[3] = {x - 0.5 * size, y, z - 0.5 * size}, e1[3] = {0, 0, size}, e2[3] = {size, 0, 0};
                quad(lm, c, e1, e2, down, mat);
            }
            add_instance(s, lm);
        }
    } catch (const std::bad_alloc&) {
        delete s;
        return XR_ERR_OUT_OF_MEMORY;
    }
    *out = s;
    return XR_OK;
}

int xrh_scene_desc(const xrh_scene* s, xr_scene_desc* d) {
    if (!s || !d) return XR_ERR_INVALID_ARGUMENT;
    d->vertices = s->vertices.data();
    d->vertex_count = (uint32_t)s->vertices.size();
    d->indices = s->indices.data();
    d->index_count = (uint32_t)s->indices.size();
    d->materials = s->materials.data();
    d->material_count = (uint32_t)s->materials.size();
    d->offsets = s->offsets.data();
    d->instance_count = (uint32_t)(s->offsets.size() / 2);
    return XR_OK;
}

int xrh_scene_textures(const xrh_scene* s, const xr_texture** textures, uint32_t* count) {
    if (!s || !textures || !count) return XR_ERR_INVALID_ARGUMENT;
    *textures = s->texture_desc.empty() ? nullptr : s->texture_desc.data();
    *count = (uint32_t)s->texture_desc.size();
    return XR_OK;
}

uint32_t xrh_scene_triangle_count(const xrh_scene* s) { return s ? (uint32_t)(s->indices.size() / 3) : 0; }

void xrh_scene_destroy(xrh_scene* s) { delete s; }

}  // extern "C"

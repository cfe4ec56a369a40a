// Camera, projection and UBO math of the host side.
//
// Restates from the mathematics what the reference takes from glm 0.9.9:
// perspectiveRH_ZO (GLM_FORCE_DEPTH_ZERO_TO_ONE, src/Asset/Vertex.h:2-3),
// lookAtRH for the view of a seeded camera path (the reference's interactive
// Camera::GetViewMatrix, src/Editor/Camera.cpp:21-27, is replaced), and the
// general 4x4 inverse of GetUniformBufferObject (src/pathTracer.cpp:188-189).
#include <math.h>
#include <string.h>

#include <new>
#include <vector>

#include "xr_host.h"

struct xrh_camera_path {
    std::vector<double> eye;     // 3 per frame
    std::vector<double> target;  // 3 per frame
};

namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni(double lo, double hi) { return lo + (hi - lo) * ((next() >> 11) * (1.0 / 9007199254740992.0)); }
};

void look_at(const double* eye, const double* center, const double* up, float* m) {
    double f[3] = {center[0] - eye[0], center[1] - eye[1], center[2] - eye[2]};
    double l = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& v : f) v /= l;
    double s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    l = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (double& v : s) v /= l;
    const double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    double r[16] = {0};
    r[0] = s[0]; r[4] = s[1]; r[8] = s[2];
    r[1] = u[0]; r[5] = u[1]; r[9] = u[2];
    r[2] = -f[0]; r[6] = -f[1]; r[10] = -f[2];
    r[12] = -(s[0] * eye[0] + s[1] * eye[1] + s[2] * eye[2]);
    r[13] = -(u[0] * eye[0] + u[1] * eye[1] + u[2] * eye[2]);
    r[14] = f[0] * eye[0] + f[1] * eye[1] + f[2] * eye[2];
    r[15] = 1.0;
    for (int i = 0; i < 16; ++i) m[i] = (float)r[i];
}

double catmull_rom(double p0, double p1, double p2, double p3, double t) {
    const double t2 = t * t, t3 = t2 * t;
    return 0.5 * ((2 * p1) + (-p0 + p2) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t2 + (-p0 + 3 * p1 - 3 * p2 + p3) * t3);
}

}  // namespace

extern "C" {

void xrh_perspective(float fovy_deg, float aspect, float znear, float zfar, float m[16]) {
    const double fovy = (double)fovy_deg * M_PI / 180.0;
    const double th = tan(fovy / 2.0);
    double r[16] = {0};
    r[0] = 1.0 / ((double)aspect * th);
    r[5] = 1.0 / th;
    r[10] = (double)zfar / ((double)znear - (double)zfar);
    r[11] = -1.0;
    r[14] = -((double)zfar * (double)znear) / ((double)zfar - (double)znear);
    r[5] = -r[5];  // currentProj[1][1] *= -1 (pathTracer.cpp:175)
    for (int i = 0; i < 16; ++i) m[i] = (float)r[i];
}

int xrh_inverse(const float in[16], float out[16]) {
    double m[16], inv[16];
    for (int i = 0; i < 16; ++i) m[i] = in[i];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    const double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    if (det == 0.0) return 0;
    for (int i = 0; i < 16; ++i) out[i] = (float)(inv[i] / det);
    return 1;
}

int xrh_camera_static(const float eye[3], const float target[3], xrh_camera_path** out) {
    if (!eye || !target || !out) return XR_ERR_INVALID_ARGUMENT;
    xrh_camera_path* p = new (std::nothrow) xrh_camera_path();
    if (!p) return XR_ERR_OUT_OF_MEMORY;
    p->eye = {eye[0], eye[1], eye[2]};
    p->target = {target[0], target[1], target[2]};
    *out = p;
    return XR_OK;
}

int xrh_camera_keyframes(const float* eyes, const float* targets, uint32_t frames, xrh_camera_path** out) {
    if (!eyes || !targets || !out || frames == 0) return XR_ERR_INVALID_ARGUMENT;
    xrh_camera_path* p = new (std::nothrow) xrh_camera_path();
    if (!p) return XR_ERR_OUT_OF_MEMORY;
    p->eye.assign(eyes, eyes + 3 * (size_t)frames);
    p->target.assign(targets, targets + 3 * (size_t)frames);
    *out = p;
    return XR_OK;
}

int xrh_camera_orbit(uint64_t seed, uint32_t n_ctrl, uint32_t frames, xrh_camera_path** out) {
    if (!out || n_ctrl < 4 || frames == 0) return XR_ERR_INVALID_ARGUMENT;
    xrh_camera_path* p = new (std::nothrow) xrh_camera_path();
    if (!p) return XR_ERR_OUT_OF_MEMORY;
    Rng rng{seed};
    std::vector<double> ce(3 * n_ctrl), ct(3 * n_ctrl);
    for (uint32_t k = 0; k < n_ctrl; ++k) {
        const double th = 2.0 * M_PI * (k + rng.uni(-0.2, 0.2)) / n_ctrl;
        const double rho = rng.uni(0.25, 0.45);
        ce[3 * k] = 0.5 + rho * cos(th);
        // above every sphere top of xrh_scene_terrain (<= 0.1 + 2 * 0.1) and
        // below its lights (0.6-0.9), so the eye never sits inside geometry
        ce[3 * k + 1] = rng.uni(0.32, 0.45);
        ce[3 * k + 2] = 0.5 + rho * sin(th);
        const double tr = rng.uni(0.0, 0.2);
        ct[3 * k] = 0.5 + tr * cos(th + M_PI);
        ct[3 * k + 1] = rng.uni(-0.02, 0.05);
        ct[3 * k + 2] = 0.5 + tr * sin(th + M_PI);
    }
    p->eye.resize(3 * (size_t)frames);
    p->target.resize(3 * (size_t)frames);
    for (uint32_t f = 0; f < frames; ++f) {
        const double u = (double)f * n_ctrl / frames;
        const uint32_t i1 = (uint32_t)u % n_ctrl;
        const double t = u - floor(u);
        const uint32_t i0 = (i1 + n_ctrl - 1) % n_ctrl, i2 = (i1 + 1) % n_ctrl, i3 = (i1 + 2) % n_ctrl;
        for (int a = 0; a < 3; ++a) {
            p->eye[3 * f + a] = catmull_rom(ce[3 * i0 + a], ce[3 * i1 + a], ce[3 * i2 + a], ce[3 * i3 + a], t);
            p->target[3 * f + a] = catmull_rom(ct[3 * i0 + a], ct[3 * i1 + a], ct[3 * i2 + a], ct[3 * i3 + a], t);
        }
    }
    *out = p;
    return XR_OK;
}

int xrh_camera_view(const xrh_camera_path* p, uint32_t k, float view[16]) {
    if (!p || !view || p->eye.empty()) return XR_ERR_INVALID_ARGUMENT;
    const size_t n = p->eye.size() / 3;
    const size_t f = k % n;
    const double up[3] = {0, 1, 0};
    look_at(&p->eye[3 * f], &p->target[3 * f], up, view);
    return XR_OK;
}

void xrh_camera_destroy(xrh_camera_path* p) { delete p; }

void xrh_make_ubo(const float view[16], const float view_prev[16], float fov_deg, uint32_t width, uint32_t height,
                  uint32_t spp, uint32_t total_samples, uint32_t bounces, uint32_t sky, uint32_t reuse,
                  uint32_t frame, xr_ubo* ubo) {
    memset(ubo, 0, sizeof(*ubo));
    float proj[16];
    xrh_perspective(fov_deg, (float)width / (float)height, 0.001f, 1000.0f, proj);
    memcpy(ubo->view, view, 64);
    memcpy(ubo->view_prev, view_prev, 64);
    memcpy(ubo->projection, proj, 64);
    memcpy(ubo->projection_prev, proj, 64);
    xrh_inverse(view, ubo->view_inverse);
    xrh_inverse(proj, ubo->projection_inverse);
    ubo->spp = spp;
    ubo->total_number_of_samples = total_samples;
    ubo->number_of_bounces = bounces;
    ubo->random_seed = 1;  // written but never read by the shaders (pathTracer.cpp:196)
    ubo->has_sky_box = sky;
    ubo->reuse = reuse;
    ubo->current_frame = frame;
}

}  // extern "C"

// CPU oracle of the MLP radiance cache (test infrastructure only: the tests,
// the smoke check and the CPU-baseline leg of the benchmark call it for
// reference results; the product never links it).
//
// The cache has no counterpart in the reference.  What is restated here is
// the published method the HIP kernels implement:
//   * network: 96 inputs -> 5 x 64 ReLU -> 3 linear outputs; per layer l the
//     parameter vector holds W_l[in][out] row-major, then b_l[out];
//   * input encoding: sin / cos(2^k pi u), k = 0..11, of each position
//     coordinate u normalised by the scene box (Mildenhall et al. 2020,
//     "NeRF", eq. 4), then view direction, normal, albedo, roughness;
//   * loss: relative L2, mean over samples and channels of
//     (p - y)^2 / (sg(p)^2 + 0.01) with the denominator held constant
//     (Mueller et al. 2021, "Real-time Neural Radiance Caching", section 5);
//   * Adam (Kingma & Ba 2015, algorithm 1) with bias correction;
//   * BF16 round to nearest even, defined by choosing the nearer of the two
//     BF16 neighbours of x (ties to the even significand), not by bit tricks;
//   * FP8 E4M3 (OCP E4M3FN: bias 7, no infinities, largest 448) round to
//     nearest even with saturation, defined as the nearest entry of the
//     table of all finite E4M3 values (ties to the even code), and the
//     per-layer scale: the smallest power of two 2^k with max|W| / 2^k <= 448.
//
// Precision modes of the forward pass:
//   mode 0  exact network in FP64 (the finite-difference reference);
//   mode 1  the rounding points of the HIP kernels: inputs, hidden weights and
//           hidden activations rounded to BF16 (FP64 dot products, bias added
//           in FP32), output layer in FP32 fmaf from the bias;
//   mode 2  as mode 1 with the hidden weights quantised to FP8 E4M3 times the
//           layer scale (exact in BF16).
// The backward pass runs in FP64 on the activations of the forward pass; in
// mode 1 the backpropagated pre-activation gradients are rounded to BF16 where
// the kernels round them.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" void orc_sincos(const float* x, uint32_t n, float* s, float* c);

namespace {

constexpr int kIn = 96, kHid = 64, kOut = 3, kLayers = 6, kBands = 12;
int lin(int l) { return l == 0 ? kIn : kHid; }
int lout(int l) { return l == kLayers - 1 ? kOut : kHid; }
int woff(int l) {
    int o = 0;
    for (int i = 0; i < l; ++i) o += lin(i) * lout(i) + lout(i);
    return o;
}
int boff(int l) { return woff(l) + lin(l) * lout(l); }
constexpr int kParams = 23043;

// Nearest BF16 value of a finite float (ties to even significand).
float bf16(float x) {
    if (!std::isfinite(x) || x == 0.0f) return x;
    uint32_t u;
    std::memcpy(&u, &x, 4);
    const uint32_t lo_bits = u & 0xFFFF0000u;  // truncation toward zero
    float lo, hi;
    std::memcpy(&lo, &lo_bits, 4);
    const uint32_t hi_bits = lo_bits + 0x10000u;  // next BF16 away from zero
    std::memcpy(&hi, &hi_bits, 4);
    const double dl = std::fabs((double)x - (double)lo), dh = std::fabs((double)hi - (double)x);
    if (dl < dh) return lo;
    if (dh < dl) return hi;
    return ((lo_bits >> 16) & 1u) ? hi : lo;  // tie: even significand
}

// The 127 finite non-negative E4M3 values (codes 0x00-0x7E) in increasing order.
struct E4m3Table {
    double value[127];
    uint8_t code[127];
    int n = 0;
    E4m3Table() {
        for (int c = 0; c < 0x7F; ++c) {  // 0x7F is NaN
            const int E = c >> 3, m = c & 7;
            value[n] = E == 0 ? m * std::ldexp(1.0, -9) : (1.0 + m / 8.0) * std::ldexp(1.0, E - 7);
            code[n] = (uint8_t)c;
            ++n;
        }
    }
};
const E4m3Table& e4m3_table() {
    static const E4m3Table t;
    return t;
}

uint8_t e4m3_code(float x) {
    const E4m3Table& t = e4m3_table();
    const double a = std::fabs((double)x);
    const uint8_t sign = std::signbit(x) ? 0x80 : 0x00;
    if (a >= t.value[t.n - 1]) return sign | t.code[t.n - 1];  // saturate at 448
    int lo = 0;
    while (lo + 1 < t.n && t.value[lo + 1] <= a) ++lo;
    if (t.value[lo] == a) return sign | t.code[lo];
    const double dl = a - t.value[lo], dh = t.value[lo + 1] - a;
    const int pick = dl < dh ? lo : dh < dl ? lo + 1 : ((t.code[lo] & 1) ? lo + 1 : lo);
    return sign | t.code[pick];
}

double e4m3_value(uint8_t c) {
    const double v = e4m3_table().value[c & 0x7F];
    return (c & 0x80) ? -v : v;
}

float layer_scale(const float* w, int l) {
    float mx = 0.0f;
    for (int i = 0; i < lin(l) * lout(l); ++i) mx = std::max(mx, std::fabs(w[woff(l) + i]));
    if (!(mx > 0.0f)) return 1.0f;
    double s = 1.0;  // smallest power of two with mx / s <= 448
    while ((double)mx / s > 448.0) s *= 2.0;
    while ((double)mx / (s / 2.0) <= 448.0) s /= 2.0;
    return (float)s;
}

// Parameters with the hidden-layer weights replaced by their FP8 values.
std::vector<float> fp8_params(const float* w) {
    std::vector<float> q(w, w + kParams);
    for (int l = 0; l < 5; ++l) {
        const float s = layer_scale(w, l);
        for (int i = 0; i < lin(l) * kHid; ++i) {
            const int k = woff(l) + i;
            q[k] = (float)(e4m3_value(e4m3_code((float)((double)w[k] / s))) * s);
        }
    }
    return q;
}

struct Trace {
    std::vector<double> act[6];  // act[0]: inputs, act[l+1]: output of hidden layer l
    double pred[3];
};

void forward_row(const float* w, const float* x, int mode, Trace& t) {
    t.act[0].resize(kIn);
    for (int k = 0; k < kIn; ++k) t.act[0][k] = mode ? (double)bf16(x[k]) : (double)x[k];
    for (int l = 0; l < 5; ++l) {
        const std::vector<double>& in = t.act[l];
        std::vector<double>& out = t.act[l + 1];
        out.assign(kHid, 0.0);
        for (int j = 0; j < kHid; ++j) {
            double z = 0.0;
            for (int k = 0; k < lin(l); ++k) {
                const float wk = w[woff(l) + k * kHid + j];
                z += in[k] * (mode ? (double)bf16(wk) : (double)wk);
            }
            if (mode) {
                const float zb = (float)z + w[boff(l) + j];
                out[j] = (double)bf16(zb > 0.0f ? zb : 0.0f);
            } else {
                z += (double)w[boff(l) + j];
                out[j] = z > 0.0 ? z : 0.0;
            }
        }
    }
    const std::vector<double>& h = t.act[5];
    for (int c = 0; c < kOut; ++c) {
        if (mode) {
            float s = w[boff(5) + c];
            for (int k = 0; k < kHid; ++k) s = std::fmaf((float)h[k], w[woff(5) + 3 * k + c], s);
            t.pred[c] = s;
        } else {
            double s = w[boff(5) + c];
            for (int k = 0; k < kHid; ++k) s += h[k] * (double)w[woff(5) + 3 * k + c];
            t.pred[c] = s;
        }
    }
}

}  // namespace

extern "C" {

int orc_mlp_param_count(void) { return kParams; }
int orc_mlp_layer_offset(int l, int bias) { return bias ? boff(l) : woff(l); }

void orc_bf16_round(const float* x, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = bf16(x[i]);
}

// Network inputs of one vertex (layout as documented in the header above):
// [0,72) sin, cos of 2^k pi u for each axis (axis-major, sin before cos),
// [72,75) w, [75,78) n, [78,81) albedo, [81] roughness 0, [82,96) zero.
void orc_mlp_encode(const float* p, const float* w, const float* n, const float* a, const float* lo,
                    const float* inv_ext, float* out) {
    for (int ax = 0; ax < 3; ++ax) {
        const float u = (p[ax] - lo[ax]) * inv_ext[ax];
        float arg[kBands], s[kBands], c[kBands];
        float scale = 3.14159265359f;
        for (int k = 0; k < kBands; ++k) {
            arg[k] = scale * u;
            scale = scale * 2.0f;
        }
        orc_sincos(arg, kBands, s, c);
        for (int k = 0; k < kBands; ++k) {
            out[ax * 2 * kBands + 2 * k] = s[k];
            out[ax * 2 * kBands + 2 * k + 1] = c[k];
        }
    }
    for (int i = 0; i < 3; ++i) {
        out[72 + i] = w[i];
        out[75 + i] = n[i];
        out[78 + i] = a[i];
    }
    for (int i = 81; i < kIn; ++i) out[i] = 0.0f;
}

void orc_e4m3_round(const float* x, uint32_t n, uint8_t* codes, float* values) {
    for (uint32_t i = 0; i < n; ++i) {
        codes[i] = e4m3_code(x[i]);
        values[i] = (float)e4m3_value(codes[i]);
    }
}

// FP8 form of the hidden layers: codes of W^T ([64][in] per layer, layers
// back to back, 22,528 bytes) and the 5 scales.
void orc_mlp_quantize(const float* w, uint8_t* codes, float* scales) {
    size_t o = 0;
    for (int l = 0; l < 5; ++l) {
        const float s = layer_scale(w, l);
        scales[l] = s;
        for (int j = 0; j < kHid; ++j)
            for (int k = 0; k < lin(l); ++k) codes[o++] = e4m3_code((float)((double)w[woff(l) + k * kHid + j] / s));
    }
}

// rows x 96 inputs -> rows x 3 outputs.
void orc_mlp_forward(const float* w_in, const float* x, uint32_t rows, int mode, double* out) {
    std::vector<float> wq;
    const float* w = w_in;
    if (mode == 2) {
        wq = fp8_params(w_in);
        w = wq.data();
        mode = 1;
    }
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < (int64_t)rows; ++r) {
        Trace t;
        forward_row(w, x + (size_t)r * kIn, mode, t);
        for (int c = 0; c < 3; ++c) out[(size_t)r * 3 + c] = t.pred[c];
    }
}

// Loss and its gradient over the parameters (FP64).  grad may be NULL.
// Returns the loss mean((p - y)^2 / (p^2 + 0.01)) over rows x 3.
double orc_mlp_loss_grad(const float* w_in, const float* x, const float* y, uint32_t rows, int mode, double* grad) {
    std::vector<float> wq;
    const float* w = w_in;
    if (mode == 2) {  // straight-through: gradients w.r.t. the master weights through the quantised ones
        wq = fp8_params(w_in);
        w = wq.data();
        mode = 1;
    }
    const double inv_count = 1.0 / (3.0 * (double)rows);
    double loss = 0.0;
    std::vector<double> g_total(grad ? kParams : 0, 0.0);
#pragma omp parallel
    {
        std::vector<double> g_local(grad ? kParams : 0, 0.0);
        double loss_local = 0.0;
        Trace t;
        std::vector<double> dz(kHid), dh(kHid);
#pragma omp for schedule(static)
        for (int64_t r = 0; r < (int64_t)rows; ++r) {
            forward_row(w, x + (size_t)r * kIn, mode, t);
            double g[3];
            for (int c = 0; c < 3; ++c) {
                const double p = t.pred[c], d = p - (double)y[(size_t)r * 3 + c];
                const double den = p * p + 0.01;
                loss_local += d * d / den * inv_count;
                g[c] = 2.0 * d / den * inv_count;
            }
            if (!grad) continue;
            // output layer
            const std::vector<double>& h4 = t.act[5];
            for (int k = 0; k < kHid; ++k)
                for (int c = 0; c < 3; ++c) g_local[woff(5) + 3 * k + c] += h4[k] * g[c];
            for (int c = 0; c < 3; ++c) g_local[boff(5) + c] += g[c];
            for (int k = 0; k < kHid; ++k) {
                double s = 0.0;
                for (int c = 0; c < 3; ++c) s += g[c] * (double)w[woff(5) + 3 * k + c];
                const double v = h4[k] > 0.0 ? s : 0.0;
                dz[k] = mode ? (double)bf16((float)v) : v;
            }
            // hidden layers 4..0
            for (int l = 4; l >= 0; --l) {
                const std::vector<double>& in = t.act[l];
                const int K = lin(l);
                for (int k = 0; k < K; ++k) {
                    if (in[k] == 0.0) continue;
                    for (int j = 0; j < kHid; ++j) g_local[woff(l) + k * kHid + j] += in[k] * dz[j];
                }
                for (int j = 0; j < kHid; ++j) g_local[boff(l) + j] += dz[j];
                if (l == 0) break;
                for (int k = 0; k < kHid; ++k) {
                    double s = 0.0;
                    for (int j = 0; j < kHid; ++j) {
                        const float wk = w[woff(l) + k * kHid + j];
                        s += dz[j] * (mode ? (double)bf16(wk) : (double)wk);
                    }
                    const double v = in[k] > 0.0 ? s : 0.0;
                    dh[k] = mode ? (double)bf16((float)v) : v;
                }
                dz.swap(dh);
            }
        }
#pragma omp critical
        {
            loss += loss_local;
            for (size_t i = 0; i < g_local.size(); ++i) g_total[i] += g_local[i];
        }
    }
    if (grad)
        for (int i = 0; i < kParams; ++i) grad[i] = g_total[i];
    return loss;
}

// One Adam step (FP64) at step number t >= 1 (bias correction 1 - beta^t).
void orc_adam_step(double* w, double* m, double* v, const double* g, uint32_t n, double lr, double beta1,
                   double beta2, double eps, uint32_t t) {
    const double bc1 = 1.0 - std::pow(beta1, (double)t), bc2 = 1.0 - std::pow(beta2, (double)t);
    for (uint32_t i = 0; i < n; ++i) {
        m[i] = beta1 * m[i] + (1.0 - beta1) * g[i];
        v[i] = beta2 * v[i] + (1.0 - beta2) * g[i] * g[i];
        w[i] -= lr * (m[i] / bc1) / (std::sqrt(v[i] / bc2) + eps);
    }
}

}  // extern "C"

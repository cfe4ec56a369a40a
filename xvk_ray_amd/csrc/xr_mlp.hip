// MLP radiance cache kernels for gfx950 (no counterpart in the reference).
//
// Every GEMM of the 96-64-64-64-64-64-3 network runs on the CDNA4 MFMA
// v_mfma_f32_32x32x16_bf16 (BF16 operands, FP32 accumulation): one wave64
// owns a tile of 64 rows (samples) and computes each 64x64 layer as 2x2
// output tiles of 32x32, K in steps of 16.  Fragment layouts (verified on the
// device): A lane l holds row l%32, k = 8*(l/32)..+7; B lane l holds
// k = 8*(l/32)..+7 of column l%32; D element j of lane l is row
// 8*(j/4) + 4*(l/32) + j%4, column l%32.  Activations live in LDS as BF16
// rows padded to 16-byte-aligned strides (104 / 72 elements) so the 16-byte
// fragment loads of consecutive lanes hit distinct banks.  The 64->3 output
// layer is 192 FMAs per row on the VALU.
//
//   k_mlp_infer   cache queries of the initial-sample pass -> L_out (persistent,
//                 weights in LDS, 8 waves per workgroup)
//   k_mlp_train   forward + backward on the frame's training samples (4 waves per
//                 64-sample tile), FP32 gradients accumulated in registers over
//                 the workgroup's tiles, then added with atomics into 32
//                 partial copies
//   k_mlp_adam    sums the partials, removes the loss scale, Adam step on the
//                 FP32 master weights, repacks the BF16 copies
//   k_mlp_eval    forward on caller-given inputs (testing)
#include "xr_internal.h"
#include "xr_kernels.h"
#include "xr_mlp.h"

#include <cstddef>
#include <mutex>

namespace xr {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kStrideIn = kMlpIn + 8;       // 104 (input rows)
constexpr int kStrideHid = kMlpHidden + 8;  // 72  (hidden rows)
// BF16 W^T copies (forward): layer 0 [64][96], layers 1..4 [64][64].
constexpr int kWtSize = 64 * 96 + 4 * 64 * 64;
XR_HD int wt_off(int l) { return l == 0 ? 0 : 64 * 96 + (l - 1) * 64 * 64; }
// BF16 W copies (backward input gradients): layers 1..4 [64][64] row-major.
constexpr int kWSize = 4 * 64 * 64;
constexpr int kGradCopies = 32;

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// FP32 -> BF16 bits on the device: one v_cvt_pk_bf16_f32, round to nearest
// even like bf16_bits (the two agree on every finite value; the network's
// activations, gradients and weights are finite).
__device__ __forceinline__ uint16_t bf16_dev(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

__device__ __forceinline__ bf16x8 frag16(const uint16_t* p) { return __builtin_bit_cast(bf16x8, *(const u32x4*)p); }

__device__ __forceinline__ bf16x8 frag_gather(const uint16_t* base, int stride) {
    bf16x8 r;
#pragma unroll
    for (int i = 0; i < 8; ++i) r[i] = __builtin_bit_cast(__bf16, base[i * stride]);
    return r;
}

// out = relu(in . W + b) for one hidden layer on a 64-row tile.
//   in  [64][sin] bf16 (LDS), K columns used
//   wt  W^T [64][wstride] bf16 (LDS or global)
//   out [64][sout] bf16 (LDS)
template <int K>
__device__ __forceinline__ void fwd_layer(const uint16_t* in, int sin, const uint16_t* wt, int wstride,
                                          const float* bias, uint16_t* out, int sout, int lane) {
    const int lr = lane & 31, lk = 8 * (lane >> 5);
#pragma unroll
    for (int tr = 0; tr < 2; ++tr) {
#pragma unroll
        for (int tc = 0; tc < 2; ++tc) {
            f32x16 acc = {};
#pragma unroll
            for (int ks = 0; ks < K / 16; ++ks) {
                const bf16x8 a = frag16(in + (tr * 32 + lr) * sin + ks * 16 + lk);
                const bf16x8 b = frag16(wt + (tc * 32 + lr) * wstride + ks * 16 + lk);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
            }
            const int col = tc * 32 + lr;
            const float bj = bias[col];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int row = tr * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
                out[row * sout + col] = bf16_dev(fmaxf(acc[j] + bj, 0.0f));
            }
        }
    }
}

// The same layer computed in place in a [64][stride] buffer: a row block's
// two 32-column output tiles are both accumulated before either is stored,
// and no later output tile reads those rows, so one buffer per wave serves
// every layer.
template <int K>
__device__ __forceinline__ void fwd_layer_inplace(uint16_t* buf, int stride, const uint16_t* wt, int wstride,
                                                  const float* bias, int lane) {
    const int lr = lane & 31, lk = 8 * (lane >> 5);
#pragma unroll
    for (int tr = 0; tr < 2; ++tr) {
        f32x16 acc[2];
#pragma unroll
        for (int tc = 0; tc < 2; ++tc) {
            acc[tc] = f32x16{};
#pragma unroll
            for (int ks = 0; ks < K / 16; ++ks) {
                const bf16x8 a = frag16(buf + (tr * 32 + lr) * stride + ks * 16 + lk);
                const bf16x8 b = frag16(wt + (tc * 32 + lr) * wstride + ks * 16 + lk);
                acc[tc] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[tc], 0, 0, 0);
            }
        }
        wave_sync();
#pragma unroll
        for (int tc = 0; tc < 2; ++tc) {
            const int col = tc * 32 + lr;
            const float bj = bias[col];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int row = tr * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
                buf[row * stride + col] = bf16_dev(fmaxf(acc[tc][j] + bj, 0.0f));
            }
        }
    }
}

// Output layer (64 -> 3) for the row owned by this lane.
__device__ __forceinline__ void out_layer(const uint16_t* h, const float* w5, const float* b5, float* pred) {
    float s0 = b5[0], s1 = b5[1], s2 = b5[2];
    for (int k = 0; k < kMlpHidden; ++k) {
        const float x = bf16_float(h[k]);
        s0 = fmaf(x, w5[3 * k + 0], s0);
        s1 = fmaf(x, w5[3 * k + 1], s1);
        s2 = fmaf(x, w5[3 * k + 2], s2);
    }
    pred[0] = s0;
    pred[1] = s1;
    pred[2] = s2;
}

struct MlpDev {
    const float* w;         // FP32 master parameters (kMlpParams)
    const uint16_t* wt_bf;  // BF16 W^T per hidden layer (forward)
    const uint16_t* w_bf;   // BF16 W per layer 1..4 (backward)
    float* grad;            // kGradCopies x kMlpParams partial gradients
    float* m;               // Adam first moment
    float* v;               // Adam second moment
    const uint8_t* wt_f8;   // FP8 E4M3 W^T per hidden layer, or null (BF16 weight format)
    const float* f8_scale;  // its per-layer scales
    float lo[3], inv_ext[3];
};

// Features of the path vertex recorded by the initial-sample pass for pixel
// pix: x_s, outgoing direction towards x_v, n_s, albedo of x_s's material.
__device__ __forceinline__ void pixel_features(const MlpDev& m, const float4* __restrict__ samples, uint32_t n,
                                               uint32_t pix, const float4* __restrict__ materials, float* f) {
    const float4 s0 = samples[pix], s2 = samples[2 * n + pix], s3 = samples[3 * n + pix];
    const v3 xs = mk3(s2.x, s2.y, s2.z);
    const v3 w = normalize(sub(mk3(s0.x, s0.y, s0.z), xs));
    const uint32_t mat = __float_as_uint(s3.w);
    const float4 mc = materials[2 * mat];
    mlp_encode(xs, w, mk3(s3.x, s3.y, s3.z), mk3(mc.x, mc.y, mc.z), mk3(m.lo[0], m.lo[1], m.lo[2]),
               mk3(m.inv_ext[0], m.inv_ext[1], m.inv_ext[2]), f);
}

__device__ __forceinline__ void store_row(uint16_t* row, const float* f, int nf) {
    for (int i = 0; i < nf; i += 2) {
        const uint32_t pk = (uint32_t)bf16_dev(f[i]) | ((uint32_t)bf16_dev(f[i + 1]) << 16);
        *(uint32_t*)(row + i) = pk;
    }
}

// ---------------------------------------------------------------------------
// Inference (cache queries) and evaluation
// ---------------------------------------------------------------------------
// LDS: padded W^T of the five hidden layers, their biases and the output
// layer, plus one activation buffer per wave (layers computed in place), for
// kInferWaves waves: two per SIMD with one workgroup per CU.
// Every array starts on a 16-byte boundary (the fragment loads are 16 bytes).
constexpr int kInferWaves = 8;
struct InferLds {
    alignas(16) uint16_t wt[64 * kStrideIn + 4 * 64 * kStrideHid];
    alignas(16) float bias[5 * 64];
    alignas(16) float w5[64 * 3 + 3];
    alignas(16) uint16_t act[kInferWaves][64 * kStrideIn];
};
static_assert(offsetof(InferLds, act) % 16 == 0 && sizeof(InferLds) <= 160 * 1024, "InferLds layout");

// FP8 weight format: the E4M3 codes are read (1 byte per weight) and
// dequantised into the BF16 LDS copy; E4M3 times a power of two is exact in
// BF16, so the MFMA sees exactly the quantised weights.
// Weights into LDS 8 at a time (16 bytes of BF16, or 8 FP8 codes): a row of
// W^T is 96 (layer 0) or 64 elements, so no 8-element chunk straddles a row.
// Element-wise 2-byte loads made this the launch's floor (about 40 us, paid
// even when the frame has few queries).
__device__ __forceinline__ void load_wt_lds(const MlpDev& m, bool f8, uint16_t* wt, float* bias, float* w5) {
    constexpr int kChunks0 = 64 * kMlpIn / 8;
    constexpr int kChunks = kChunks0 + 4 * 64 * 64 / 8;
    for (int c = threadIdx.x; c < kChunks; c += blockDim.x) {
        int l, src, dst;
        if (c < kChunks0) {
            const int r = c / (kMlpIn / 8), col = (c % (kMlpIn / 8)) * 8;
            l = 0;
            src = r * kMlpIn + col;
            dst = r * kStrideIn + col;
        } else {
            const int d = c - kChunks0, e = d % 512, r = e / 8, col = (e % 8) * 8;
            l = 1 + d / 512;
            src = 64 * kMlpIn + (l - 1) * 4096 + r * 64 + col;
            dst = 64 * kStrideIn + (l - 1) * 64 * kStrideHid + r * kStrideHid + col;
        }
        u32x4 v;
        if (f8) {
            const uint2 q = *(const uint2*)(m.wt_f8 + src);
            const float sc = m.f8_scale[l];
            uint32_t h[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t word = j < 2 ? q.x : q.y;
                const int sh = 16 * (j & 1);
                h[j] = (uint32_t)bf16_bits(e4m3_float((uint8_t)(word >> sh)) * sc) |
                       ((uint32_t)bf16_bits(e4m3_float((uint8_t)(word >> (sh + 8))) * sc) << 16);
            }
            v = u32x4{h[0], h[1], h[2], h[3]};
        } else {
            v = *(const u32x4*)(m.wt_bf + src);
        }
        *(u32x4*)(wt + dst) = v;
    }
    for (int i = threadIdx.x; i < 5 * 64; i += blockDim.x) bias[i] = m.w[mlp_boff(i / 64) + i % 64];
    for (int i = threadIdx.x; i < 64 * 3 + 3; i += blockDim.x) w5[i] = m.w[mlp_woff(5) + i];
}
__device__ __forceinline__ void load_weights_lds(const MlpDev& m, InferLds& s) {
    load_wt_lds(m, m.wt_f8 != nullptr, s.wt, s.bias, s.w5);
}

// Runs the five hidden layers and the output layer on the tile in act[w]
// (inputs already stored there with stride kStrideIn).
__device__ __forceinline__ void forward_tile(InferLds& s, int w, int lane, float* pred) {
    uint16_t* A = s.act[w];
    const uint16_t* wt1 = s.wt + 64 * kStrideIn;
    wave_sync();
    fwd_layer_inplace<kMlpIn>(A, kStrideIn, s.wt, kStrideIn, s.bias, lane);
    wave_sync();
    fwd_layer_inplace<64>(A, kStrideIn, wt1 + 0 * 64 * kStrideHid, kStrideHid, s.bias + 64, lane);
    wave_sync();
    fwd_layer_inplace<64>(A, kStrideIn, wt1 + 1 * 64 * kStrideHid, kStrideHid, s.bias + 128, lane);
    wave_sync();
    fwd_layer_inplace<64>(A, kStrideIn, wt1 + 2 * 64 * kStrideHid, kStrideHid, s.bias + 192, lane);
    wave_sync();
    fwd_layer_inplace<64>(A, kStrideIn, wt1 + 3 * 64 * kStrideHid, kStrideHid, s.bias + 256, lane);
    wave_sync();
    out_layer(A + lane * kStrideIn, s.w5, s.w5 + 192, pred);
    wave_sync();
}

__global__ __launch_bounds__(64 * kInferWaves) void k_mlp_infer(MlpDev m, FrameDev fr, const float4* __restrict__ materials) {
    __shared__ InferLds s;
    load_weights_lds(m, s);
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t n = fr.n;
    // the initial pass listed this frame's query pixels and training pixels;
    // both take the network's prediction (training pixels after the step has
    // read their traced radiance, as the training paths of neural radiance
    // caching only feed the network)
    const uint32_t nq = fr.cache_list[0], nt = fr.cache_list[1];
    const uint32_t count = nq + nt;
    const uint32_t ntiles = (count + 63) / 64;
    for (uint32_t tile = blockIdx.x * kInferWaves + w; tile < ntiles; tile += gridDim.x * kInferWaves) {
        const uint32_t idx = tile * 64 + lane;
        const uint32_t kind = idx < nq ? 1u : 2u;
        const uint32_t pix = idx < nq ? fr.cache_list[2 + idx] : (idx < count ? fr.cache_list[2 + n + (idx - nq)] : 0u);
        bool query = false;
        float f[kMlpIn];
        if (idx < count) query = __float_as_uint(fr.samples[4 * n + pix].w) == kind;
        if (query) pixel_features(m, fr.samples, n, pix, materials, f);
        else
            for (int i = 0; i < kMlpIn; ++i) f[i] = 0.0f;
        store_row(s.act[w] + lane * kStrideIn, f, kMlpIn);
        float pred[3];
        forward_tile(s, w, lane, pred);
        if (query) {
            const float4 old = fr.samples[4 * n + pix];
            fr.samples[4 * n + pix] =
                make_float4(fmaxf(pred[0], 0.0f), fmaxf(pred[1], 0.0f), fmaxf(pred[2], 0.0f), old.w);
        }
    }
}

__global__ __launch_bounds__(64 * kInferWaves) void k_mlp_eval(MlpDev m, const float* __restrict__ feats, uint32_t rows,
                                                  float* __restrict__ out) {
    __shared__ InferLds s;
    load_weights_lds(m, s);
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t ntiles = (rows + 63) / 64;
    for (uint32_t tile = blockIdx.x * kInferWaves + w; tile < ntiles; tile += gridDim.x * kInferWaves) {
        const uint32_t r = tile * 64 + lane;
        float f[kMlpIn];
        for (int i = 0; i < kMlpIn; ++i) f[i] = r < rows ? feats[(size_t)r * kMlpIn + i] : 0.0f;
        store_row(s.act[w] + lane * kStrideIn, f, kMlpIn);
        float pred[3];
        forward_tile(s, w, lane, pred);
        if (r < rows)
            for (int c = 0; c < 3; ++c) out[(size_t)r * 3 + c] = pred[c];
    }
}

// ---------------------------------------------------------------------------
// Training: forward + backward of 64-sample tiles, four waves per workgroup
// ---------------------------------------------------------------------------
// LDS of a training workgroup (one per CU: the kernel's registers allow one
// wave per SIMD): the tile's activations and the network.  The MFMA operands
// of the weights come from padded LDS copies loaded once per workgroup (W^T
// for the forward products, W of layers 1..4 for dH = dz . W^T) instead of
// from the BF16 copies in global memory once per tile and layer.
struct TrainLds {
    alignas(16) uint16_t x[64 * kStrideIn];
    alignas(16) uint16_t h[5][64 * kStrideHid];  // outputs of hidden layers 0..4
    alignas(16) uint16_t dz[64 * kStrideHid];
    alignas(16) float g[64 * 4];
    alignas(16) uint16_t wt[64 * kStrideIn + 4 * 64 * kStrideHid];
    alignas(16) uint16_t wb[4 * 64 * kStrideHid];
    alignas(16) float bias[5 * 64];
    alignas(16) float w5[64 * 3 + 3];
};
static_assert(sizeof(TrainLds) <= 160 * 1024, "TrainLds layout");

// W of hidden layers 1..4 (row-major [64][64]) into rows of kStrideHid.
__device__ __forceinline__ void load_w_lds(const MlpDev& m, uint16_t* wb) {
    for (int c = threadIdx.x; c < 4 * 512; c += blockDim.x) {
        const int l = c / 512, e = c % 512, r = e / 8, col = (e % 8) * 8;
        *(u32x4*)(wb + l * 64 * kStrideHid + r * kStrideHid + col) =
            *(const u32x4*)(m.w_bf + l * 4096 + r * 64 + col);
    }
}

// One 32x32 output tile (row block tr, column block tc) of a hidden layer:
// the arithmetic of fwd_layer for that tile.
template <int K>
__device__ __forceinline__ void fwd_subtile(const uint16_t* in, int sin, const uint16_t* wt, int wstride,
                                            const float* bias, uint16_t* out, int sout, int lane, int tr, int tc) {
    const int lr = lane & 31, lk = 8 * (lane >> 5);
    f32x16 acc = {};
#pragma unroll
    for (int ks = 0; ks < K / 16; ++ks) {
        const bf16x8 a = frag16(in + (tr * 32 + lr) * sin + ks * 16 + lk);
        const bf16x8 b = frag16(wt + (tc * 32 + lr) * wstride + ks * 16 + lk);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    const int col = tc * 32 + lr;
    const float bj = bias[col];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int row = tr * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
        out[row * sout + col] = bf16_dev(fmaxf(acc[j] + bj, 0.0f));
    }
}

// Four waves per 64-sample tile: each wave encodes a quarter of the input
// columns, wave 0 owns the rows for the output layer and the loss (lane =
// sample), and every layer's four 32x32 output tiles, the weight-gradient
// tiles and the row loops are shared out between the waves.  With one wave per tile the tile's VALU work ran serially on one
// SIMD (68 us per frame); the frame has only ~131 tiles.
constexpr int kTrainWaves = 4;

// Loss: relative L2, sum over samples and channels of (p - y)^2 / (sg(p)^2 + 0.01),
// times inv_count.  g = loss_scale * dLoss/dp.
// kFromFrame: the samples are the frame's training pixels listed by the
// initial pass (*count of them in list); otherwise nslots rows of
// caller-given features and targets.  inv_count = 1 / (3 * sample count),
// where in frame mode the count is *total when given (the training samples
// of several batched views, trained as one step over their union).
template <bool kFromFrame>
__global__ __launch_bounds__(64 * kTrainWaves) void k_mlp_train(MlpDev m, const float4* __restrict__ samples,
                                                                uint32_t n, const uint32_t* __restrict__ count,
                                                                const uint32_t* __restrict__ list, uint32_t nslots,
                                                                const float4* __restrict__ materials,
                                                                const float* __restrict__ feats,
                                                                const float* __restrict__ targets, float loss_scale,
                                                                float inv_count, const uint32_t* __restrict__ total) {
    __shared__ TrainLds s;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lr = lane & 31, lk = 8 * (lane >> 5);
    const int tr = wv >> 1, tc = wv & 1;
    float* grad = m.grad + (size_t)(blockIdx.x % kGradCopies) * kMlpParams;
    if (kFromFrame) {
        nslots = *count;
        const uint32_t all = total ? *total : nslots;
        inv_count = 1.0f / (3.0f * (float)(all ? all : 1u));
    }
    const uint32_t ntiles = (nslots + 63) / 64;
    if ((uint32_t)blockIdx.x >= ntiles) return;  // uniform over the workgroup
    load_wt_lds(m, false, s.wt, s.bias, s.w5);
    load_w_lds(m, s.wb);
    __syncthreads();
    // This wave's gradient tiles, accumulated over every tile the workgroup
    // trains and added into the partial copy once at the end: dw[l - 1] is
    // output tile (wv >> 1, wv & 1) of dW_l for l = 1..4, dw[4] that tile of
    // dW_0 and dw[5] (waves 0 and 1) tile (2, wv) of dW_0; db[l] is this
    // lane's column sum of dz_l over rows 16 wv .. 16 wv + 15; d5[c] is
    // dW5[lane][c] and d5[3] db5[lane] (wave 0).
    f32x16 dw[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) dw[i] = f32x16{};
    float db[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    float d5[4] = {0.f, 0.f, 0.f, 0.f};
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        // inputs: wave w < 3 writes the 24 frequency features of axis w, wave 3
        // columns 72..95 (direction, normal, albedo, padding)
        const uint32_t slot = tile * 64 + lane;
        const int c0 = 24 * wv;
        float y[3] = {0.f, 0.f, 0.f};
        float f[24];
        bool valid = false;
        if (kFromFrame) {
            const uint32_t pix = slot < nslots ? list[slot] : n;
            if (pix < n) {
                const float4 s4 = samples[4 * n + pix];
                if (__float_as_uint(s4.w) == 2u) {
                    valid = true;
                    y[0] = s4.x; y[1] = s4.y; y[2] = s4.z;
                    const float4 s2 = samples[2 * n + pix];
                    if (wv < 3) {
                        const float pa = wv == 0 ? s2.x : (wv == 1 ? s2.y : s2.z);
                        mlp_encode_axis((pa - m.lo[wv]) * m.inv_ext[wv], f);
                    } else {
                        const float4 s0 = samples[pix], s3 = samples[3 * n + pix];
                        const v3 w = normalize(sub(mk3(s0.x, s0.y, s0.z), mk3(s2.x, s2.y, s2.z)));
                        const float4 mc = materials[2 * __float_as_uint(s3.w)];
                        mlp_encode_tail(w, mk3(s3.x, s3.y, s3.z), mk3(mc.x, mc.y, mc.z), f);
                    }
                }
            }
        } else if (slot < nslots) {
            valid = true;
            for (int i = 0; i < 24; ++i) f[i] = feats[(size_t)slot * kMlpIn + c0 + i];
            for (int c = 0; c < 3; ++c) y[c] = targets[(size_t)slot * 3 + c];
        }
        if (!valid)
            for (int i = 0; i < 24; ++i) f[i] = 0.0f;
        store_row(s.x + lane * kStrideIn + c0, f, 24);
        __syncthreads();
        // forward, keeping every hidden activation: wave wv computes tile (tr, tc)
        fwd_subtile<kMlpIn>(s.x, kStrideIn, s.wt, kStrideIn, s.bias, s.h[0], kStrideHid, lane, tr, tc);
        for (int l = 1; l < 5; ++l) {
            __syncthreads();
            fwd_subtile<64>(s.h[l - 1], kStrideHid, s.wt + 64 * kStrideIn + (l - 1) * 64 * kStrideHid, kStrideHid,
                            s.bias + 64 * l, s.h[l], kStrideHid, lane, tr, tc);
        }
        __syncthreads();
        const float* w5 = s.w5;
        if (wv == 0) {
            float pred[3];
            out_layer(s.h[4] + lane * kStrideHid, w5, w5 + 192, pred);
            for (int c = 0; c < 3; ++c) {
                const float d = pred[c] - y[c];
                const float gc = valid ? loss_scale * inv_count * 2.0f * d / (pred[c] * pred[c] + 0.01f) : 0.0f;
                s.g[lane * 4 + c] = gc;
            }
        }
        __syncthreads();
        // output layer: dW5[k][c] = sum_r h4[r][k] g[r][c]; db5[c] = sum_r g[r][c]  (wave 0, lane = k)
        if (wv == 0) {
            float a0 = 0.f, a1 = 0.f, a2 = 0.f;
            for (int r = 0; r < 64; ++r) {
                const float x = bf16_float(s.h[4][r * kStrideHid + lane]);
                a0 = fmaf(x, s.g[r * 4 + 0], a0);
                a1 = fmaf(x, s.g[r * 4 + 1], a1);
                a2 = fmaf(x, s.g[r * 4 + 2], a2);
            }
            d5[0] += a0;
            d5[1] += a1;
            d5[2] += a2;
            if (lane < 3) {
                float b = 0.f;
                for (int r = 0; r < 64; ++r) b += s.g[r * 4 + lane];
                d5[3] += b;
            }
        }
        // dZ4 = (g . W5^T) masked by h4 > 0: lane = row, wave wv takes columns 16 wv .. 16 wv + 15
        {
            const float g0 = s.g[lane * 4 + 0], g1 = s.g[lane * 4 + 1], g2 = s.g[lane * 4 + 2];
            for (int k = 16 * wv; k < 16 * wv + 16; ++k) {
                const float dh = g0 * w5[3 * k] + g1 * w5[3 * k + 1] + g2 * w5[3 * k + 2];
                const bool on = bf16_float(s.h[4][lane * kStrideHid + k]) > 0.0f;
                s.dz[lane * kStrideHid + k] = bf16_dev(on ? dh : 0.0f);
            }
        }
        __syncthreads();
        // hidden layers 4..0: dW_l = in^T . dz, db_l = colsum(dz), dz_prev = (dz . W_l^T) * (in > 0)
#pragma unroll
        for (int l = 4; l >= 0; --l) {
            const uint16_t* in = l == 0 ? s.x : s.h[l - 1];
            const int sin = l == 0 ? kStrideIn : kStrideHid;
            // dW_l tile (tm, tn) += in[:, 32 tm ..]^T . dz[:, 32 tn ..]
            auto dw_tile = [&](f32x16& acc, int tm, int tn) {
#pragma unroll
                for (int ks = 0; ks < 4; ++ks) {
                    const bf16x8 a = frag_gather(in + (ks * 16 + lk) * sin + tm * 32 + lr, sin);
                    const bf16x8 b = frag_gather(s.dz + (ks * 16 + lk) * kStrideHid + tn * 32 + lr, kStrideHid);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
                }
            };
            if (l > 0) {
                dw_tile(dw[l - 1], tr, tc);
            } else {
                dw_tile(dw[4], tr, tc);
                if (wv < 2) dw_tile(dw[5], 2, wv);
            }
            {  // column = lane, rows 16 wv .. 16 wv + 15
                float b = 0.f;
                for (int r = 16 * wv; r < 16 * wv + 16; ++r) b += bf16_float(s.dz[r * kStrideHid + lane]);
                db[l] += b;
            }
            if (l == 0) break;
            // dH_in = dz . W_l^T  (W_l row-major [64][64] from the BF16 copy): wave wv computes tile (tr, tc)
            const uint16_t* wl = s.wb + (l - 1) * 64 * kStrideHid;
            f32x16 acc = {};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                const bf16x8 a = frag16(s.dz + (tr * 32 + lr) * kStrideHid + ks * 16 + lk);
                const bf16x8 b = frag16(wl + (tc * 32 + lr) * kStrideHid + ks * 16 + lk);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
            }
            __syncthreads();  // every wave is done reading dz
            const int col = tc * 32 + lr;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int row = tr * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
                const bool on = bf16_float(in[row * sin + col]) > 0.0f;
                s.dz[row * kStrideHid + col] = bf16_dev(on ? acc[j] : 0.0f);
            }
            __syncthreads();
        }
        __syncthreads();
    }
    auto flush = [&](const f32x16& acc, int l, int tm, int tn) {
        const int col = tn * 32 + lr;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int row = tm * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
            atomicAdd(&grad[mlp_woff(l) + row * 64 + col], acc[j]);
        }
    };
#pragma unroll
    for (int l = 1; l < 5; ++l) flush(dw[l - 1], l, tr, tc);
    flush(dw[4], 0, tr, tc);
    if (wv < 2) flush(dw[5], 0, 2, wv);
#pragma unroll
    for (int l = 0; l < 5; ++l) atomicAdd(&grad[mlp_boff(l) + lane], db[l]);
    if (wv == 0) {
        atomicAdd(&grad[mlp_woff(5) + 3 * lane + 0], d5[0]);
        atomicAdd(&grad[mlp_woff(5) + 3 * lane + 1], d5[1]);
        atomicAdd(&grad[mlp_woff(5) + 3 * lane + 2], d5[2]);
        if (lane < 3) atomicAdd(&grad[mlp_boff(5) + lane], d5[3]);
    }
}

// Adam on the FP32 master weights (one thread per parameter), then the BF16
// copies.  The partial gradients are summed, unscaled and cleared.
__global__ __launch_bounds__(256) void k_mlp_adam(MlpDev m, float* __restrict__ w, uint16_t* __restrict__ wt_bf,
                                                  uint16_t* __restrict__ w_bf, float lr, float beta1, float beta2,
                                                  float eps, float inv_scale, float bc1, float bc2, int do_step) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kMlpParams) return;
    float wi = w[i];
    if (do_step) {
        float g = 0.0f;
        for (int p = 0; p < kGradCopies; ++p) {
            g += m.grad[(size_t)p * kMlpParams + i];
            m.grad[(size_t)p * kMlpParams + i] = 0.0f;
        }
        g *= inv_scale;
        if (isfinite(g)) {
            const float mi = beta1 * m.m[i] + (1.0f - beta1) * g;
            const float vi = beta2 * m.v[i] + (1.0f - beta2) * g * g;
            m.m[i] = mi;
            m.v[i] = vi;
            wi -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
            w[i] = wi;
        }
    }
    // repack: which layer / weight or bias
    for (int l = 0; l < 5; ++l) {
        const int wo = mlp_woff(l), in = mlp_in(l);
        if (i >= wo && i < wo + in * 64) {
            const int k = (i - wo) / 64, j = (i - wo) % 64;
            wt_bf[wt_off(l) + j * in + k] = bf16_bits(wi);
            if (l >= 1) w_bf[(l - 1) * 4096 + k * 64 + j] = bf16_bits(wi);
        }
    }
}

// FP8 weight format: one workgroup per hidden layer finds max|W|, picks the
// power-of-two scale, writes the E4M3 codes of W^T and overwrites the BF16
// forward / backward copies with the dequantised values (straight-through:
// training keeps the FP32 master weights and backpropagates through the
// quantised ones).
__global__ __launch_bounds__(256) void k_mlp_quant(const float* __restrict__ w, uint8_t* __restrict__ wt_f8,
                                                   float* __restrict__ scales, uint16_t* __restrict__ wt_bf,
                                                   uint16_t* __restrict__ w_bf) {
    __shared__ float red[256];
    const int l = blockIdx.x;
    const int in = mlp_in(l), cnt = in * 64, wo = mlp_woff(l);
    float mx = 0.0f;
    for (int i = threadIdx.x; i < cnt; i += 256) mx = fmaxf(mx, fabsf(w[wo + i]));
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
    }
    const float scale = e4m3_layer_scale(red[0]);
    if (threadIdx.x == 0) scales[l] = scale;
    for (int i = threadIdx.x; i < cnt; i += 256) {
        const int k = i / 64, j = i % 64;
        const uint8_t q = e4m3_bits(w[wo + i] / scale);
        const uint16_t b = bf16_bits(e4m3_float(q) * scale);
        wt_f8[wt_off(l) + j * in + k] = q;
        wt_bf[wt_off(l) + j * in + k] = b;
        if (l >= 1) w_bf[(l - 1) * 4096 + k * 64 + j] = b;
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static MlpDev mlp_dev(const MlpState& st) {
    MlpDev m;
    m.w = st.w;
    m.wt_bf = st.wt_bf;
    m.w_bf = st.w_bf;
    m.wt_f8 = st.fp8 ? st.wt_f8 : nullptr;
    m.f8_scale = st.f8_scale;
    m.grad = st.grad;
    m.m = st.m;
    m.v = st.v;
    for (int a = 0; a < 3; ++a) {
        m.lo[a] = st.lo[a];
        m.inv_ext[a] = st.inv_ext[a];
    }
    return m;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
static const size_t kParamBytes = align256(sizeof(float) * kMlpParams);

size_t mlp_bytes() {
    return 3 * kParamBytes + align256(sizeof(float) * (size_t)kMlpParams * kGradCopies) +
           align256(sizeof(uint16_t) * kWtSize) + align256(sizeof(uint16_t) * kWSize) + align256(kWtSize) +
           align256(sizeof(float) * 5);
}

size_t mlp_grad_offset_bytes() { return 3 * kParamBytes; }

// Regions (each 256-byte aligned, so every 16-byte BF16 fragment load is
// aligned): w | m | v | grad[32] | wt_bf | w_bf | wt_f8 | f8_scale.
hipError_t mlp_bind(MlpState* st, void* base) {
    char* p = (char*)base;
    st->w = (float*)p;
    p += kParamBytes;
    st->m = (float*)p;
    p += kParamBytes;
    st->v = (float*)p;
    p += kParamBytes;
    st->grad = (float*)p;
    p += align256(sizeof(float) * (size_t)kMlpParams * kGradCopies);
    st->wt_bf = (uint16_t*)p;
    p += align256(sizeof(uint16_t) * kWtSize);
    st->w_bf = (uint16_t*)p;
    p += align256(sizeof(uint16_t) * kWSize);
    st->wt_f8 = (uint8_t*)p;
    p += align256(kWtSize);
    st->f8_scale = (float*)p;
    return hipSuccess;
}

hipError_t launch_mlp_adam(const MlpState& st, float lr, float inv_scale, bool step, hipStream_t s) {
    MlpDev m = mlp_dev(st);
    const float b1 = 0.9f, b2 = 0.99f;
    const float bc1 = 1.0f - powf(b1, (float)st.steps);
    const float bc2 = 1.0f - powf(b2, (float)st.steps);
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, m, st.w, st.wt_bf, st.w_bf, lr,
                       b1, b2, 1e-8f, inv_scale, step ? bc1 : 1.0f, step ? bc2 : 1.0f, step ? 1 : 0);
    if (st.fp8) hipLaunchKernelGGL(k_mlp_quant, dim3(5), dim3(256), 0, s, st.w, st.wt_f8, st.f8_scale, st.wt_bf, st.w_bf);
    return hipGetLastError();
}

hipError_t launch_mlp_infer(const MlpState& st, const FrameDev& fr, const float4* materials, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_infer, dim3(256), dim3(64 * kInferWaves), 0, s, mlp_dev(st), fr, materials);
    return hipGetLastError();
}

hipError_t launch_mlp_eval(const MlpState& st, const float* feats, uint32_t rows, float* out, hipStream_t s) {
    const uint32_t per = 64u * kInferWaves;
    const uint32_t blocks = (rows + per - 1) / per < 256 ? (rows + per - 1) / per : 256;
    hipLaunchKernelGGL(k_mlp_eval, dim3(blocks ? blocks : 1), dim3(per), 0, s, mlp_dev(st), feats, rows, out);
    return hipGetLastError();
}

hipError_t launch_mlp_train_frame(const MlpState& st, const FrameDev& fr, const float4* materials, float loss_scale,
                                  hipStream_t s, const uint32_t* total) {
    // The sample count is on the device (cache_list[1]); 256 workgroups
    // stride over its 64-sample tiles.
    hipLaunchKernelGGL(k_mlp_train<true>, dim3(256), dim3(64 * kTrainWaves), 0, s, mlp_dev(st), fr.samples, fr.n,
                       (const uint32_t*)(fr.cache_list + 1), (const uint32_t*)(fr.cache_list + 2 + fr.n), 0u,
                       materials, (const float*)nullptr, (const float*)nullptr, loss_scale, 0.0f, total);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Gradient without allocation: the 32 partial gradient copies live in a
// file-scope device array instead of the network's state, so a caller that
// cannot allocate (graph capture, serving loops) gets the gradient of
// caller-given device rows.  One array per device, held until the process
// ends; every call on a device shares it (overlapping calls race).
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchFloats = (size_t)kGradCopies * kMlpParams;  // 2,949,504 bytes
static __device__ __align__(256) float scratch[kNoAllocScratchFloats];

namespace {
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

hipError_t mlp_noalloc_scratch(float** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = (float*)p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

__global__ __launch_bounds__(256) void k_clear_f32(float* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = 0.0f;
}

// grad_out[i] = sum over the copies in copy order, in FP64, rounded once.
__global__ __launch_bounds__(256) void k_mlp_grad_sum(const float* __restrict__ parts, float* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kMlpParams) return;
    double g = 0.0;
    for (int p = 0; p < kGradCopies; ++p) g += (double)parts[(size_t)p * kMlpParams + i];
    out[i] = (float)g;
}

hipError_t launch_mlp_gradient_noalloc(const MlpState& st, float* scratch_ptr, const float* feats,
                                       const float* targets, uint32_t rows, float* grad_out, hipStream_t s) {
    MlpState g = st;
    g.grad = scratch_ptr;
    hipLaunchKernelGGL(k_clear_f32, dim3(1024), dim3(256), 0, s, scratch_ptr, kNoAllocScratchFloats);
    hipError_t e = launch_mlp_train_rows(g, feats, targets, rows, 1.0f, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlp_grad_sum, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, scratch_ptr, grad_out);
    return hipGetLastError();
}

hipError_t launch_mlp_train_rows(const MlpState& st, const float* feats, const float* targets, uint32_t rows,
                                 float loss_scale, hipStream_t s) {
    const uint32_t tiles = (rows + 63) / 64;
    const uint32_t blocks = tiles < 512 ? tiles : 512;
    hipLaunchKernelGGL(k_mlp_train<false>, dim3(blocks ? blocks : 1), dim3(64 * kTrainWaves), 0, s, mlp_dev(st),
                       (const float4*)nullptr, 0u, (const uint32_t*)nullptr, (const uint32_t*)nullptr, rows,
                       (const float4*)nullptr, feats,
                       targets, loss_scale, 1.0f / (3.0f * (float)(rows ? rows : 1)), (const uint32_t*)nullptr);
    return hipGetLastError();
}

}  // namespace xr

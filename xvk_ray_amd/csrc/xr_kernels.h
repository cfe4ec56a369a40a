// Host-side launchers of the frame kernels and of the BVH builder.
#pragma once
#include <hip/hip_runtime.h>
#include "xr_internal.h"

namespace xr {

uint32_t image_grid(uint32_t width, uint32_t height);

struct CacheCfg;
hipError_t launch_initial_sample(const SceneDev& sc, const FrameDev& fr, const UniformBufferObject& ubo,
                                 const CacheCfg& cache, hipStream_t st);
hipError_t launch_temporal(const FrameDev& fr, uint32_t frame, const PassParams& pp, const float4* gpos_hist,
                           hipStream_t st);
// The temporal pass followed by the spatial pass without neighbours
// (pp.spatial_samples = 0) in one kernel: the same writes as the two passes,
// except the neighbour records, which nothing reads then.
hipError_t launch_temporal_spatial0(const FrameDev& fr, uint32_t frame, uint32_t reuse, const PassParams& pp,
                                    const float4* gpos_hist, hipStream_t st);
// Views rendered together: device arrays of their frame buffers and UBOs, a
// shared handed-on path list (records name their view) and shared counters.
struct BatchDev {
    const FrameDev* views;
    const UniformBufferObject* ubos;
    uint32_t nviews;
    float4* tail;
    uint32_t* tail_count;   // 1 + 16 words: the tail count, then the tile queues
    uint32_t* tile_queue;   // tail_count + 1: per-XCD tile counters of the persistent initial / spatial kernels
    unsigned long long* counters;
};
constexpr uint32_t kBatchCountWords = 17;
// The batch's per-frame words after the tail count and tile queues: the
// training-sample total of the views (cache mode).
constexpr uint32_t kBatchTrainTotal = kBatchCountWords;
constexpr uint32_t kBatchWords = kBatchCountWords + 1;
uint32_t batch_config_count();
void batch_config_info(uint32_t cfg, uint32_t* tw, uint32_t* th, uint32_t* lds, bool* persistent);
hipError_t launch_batch_frame(const SceneDev& sc, const BatchDev& bd, const FrameDev* views, uint32_t cfg,
                              const PassParams& pp, hipStream_t st);
// The two halves of a batched frame, for the radiance cache's training step
// between them: the initial-sample pass of every view (with each view's
// training choice when cache.enabled), then the temporal and spatial passes.
hipError_t launch_batch_initial(const SceneDev& sc, const BatchDev& bd, const FrameDev* views, uint32_t cfg,
                                const CacheCfg& cache, hipStream_t st);
hipError_t launch_batch_reuse(const SceneDev& sc, const BatchDev& bd, const FrameDev* views, uint32_t cfg,
                              const PassParams& pp, hipStream_t st);
// *total = the views' training samples of this frame (cache mode).
hipError_t launch_batch_train_total(const BatchDev& bd, uint32_t* total, hipStream_t st);
// Neighbour records of the spatial pass from the samples and the temporal
// reservoirs of frame's slot (when the temporal pass did not write them).
hipError_t launch_nbr_build(const FrameDev& fr, uint32_t frame, hipStream_t st);
hipError_t launch_spatial(const SceneDev& sc, const FrameDev& fr, uint32_t frame, uint32_t reuse,
                          const PassParams& pp, hipStream_t st);
hipError_t launch_plain(const SceneDev& sc, const FrameDev& fr, const UniformBufferObject& ubo, uint32_t max_depth,
                        hipStream_t st);
hipError_t launch_converge(const SceneDev& sc, const FrameDev& fr, const UniformBufferObject& ubo,
                           uint32_t frame_base, uint32_t spp, double* acc, hipStream_t st);
hipError_t launch_sq_error(const float4* img, const double* acc, double inv_spp, uint32_t n, double* out,
                           hipStream_t st);
hipError_t launch_trace(const SceneDev& sc, const float4* rays, uint32_t n, float4* hits, bool any_hit,
                        hipStream_t st);
// Resources of frame kernel `index` (0 .. kKernelInfoCount-1): name, the
// runtime's attributes, resident workgroups per CU and the workgroup size.
constexpr int kKernelInfoCount = 6;
hipError_t kernel_resources(uint32_t index, const char** name, hipFuncAttributes* attr, int* blocks_per_cu,
                            int* block);

// MLP radiance cache state: one device allocation of mlp_bytes() bytes carved
// by mlp_bind into 256-byte-aligned regions.
struct MlpState {
    float* w = nullptr;         // FP32 master parameters (23,043)
    float* m = nullptr;         // Adam first moment
    float* v = nullptr;         // Adam second moment
    float* grad = nullptr;      // 32 partial gradient copies
    uint16_t* wt_bf = nullptr;  // BF16 W^T of the hidden layers (forward)
    uint16_t* w_bf = nullptr;   // BF16 W of hidden layers 1..4 (backward)
    uint8_t* wt_f8 = nullptr;   // FP8 E4M3 W^T of the hidden layers (FP8 weight format)
    float* f8_scale = nullptr;  // per hidden layer power-of-two scale of wt_f8
    bool fp8 = false;           // weight format: the BF16 copies hold dequantised FP8 values
    float lo[3] = {0, 0, 0}, inv_ext[3] = {1, 1, 1};  // position normalisation
    uint32_t steps = 0;         // Adam steps taken
};
size_t mlp_bytes();
size_t mlp_grad_offset_bytes();
hipError_t mlp_bind(MlpState* st, void* base);
hipError_t launch_mlp_adam(const MlpState& st, float lr, float inv_scale, bool step, hipStream_t s);
hipError_t launch_clear_u32(uint32_t* p, uint32_t n, hipStream_t s);
hipError_t launch_mlp_infer(const MlpState& st, const FrameDev& fr, const float4* materials, hipStream_t s);
hipError_t launch_mlp_eval(const MlpState& st, const float* feats, uint32_t rows, float* out, hipStream_t s);
// Gradient of the frame's training samples, added into the partial copies;
// total (device, may be null): the loss divides by *total samples instead of
// the frame's own count (a batch's views trained as one step).
hipError_t launch_mlp_train_frame(const MlpState& st, const FrameDev& fr, const float4* materials, float loss_scale,
                                  hipStream_t s, const uint32_t* total = nullptr);
hipError_t launch_mlp_train_rows(const MlpState& st, const float* feats, const float* targets, uint32_t rows,
                                 float loss_scale, hipStream_t s);
// Gradient of caller-given device rows with the partial copies in the
// library's static scratch array (no allocation).  mlp_noalloc_scratch
// resolves that array's address on the current device (once per device).
constexpr int kMaxScratchDevices = 64;
hipError_t mlp_noalloc_scratch(float** out);
hipError_t launch_mlp_gradient_noalloc(const MlpState& st, float* scratch_ptr, const float* feats,
                                       const float* targets, uint32_t rows, float* grad_out, hipStream_t s);

// Cache termination of the initial-sample pass (MLP mode).
struct CacheCfg {
    uint32_t enabled;
    uint32_t train_target;  // training samples per frame (of the candidates with a diffuse x_s)
    uint32_t frame;         // rotates the training choice
    float spread_c;         // terminate at x_s when its spread exceeds spread_c * primary footprint
};
// Pixels per workgroup of the training-choice kernels.
constexpr uint32_t kCacheSelBlock = 4096;
// The frame's training choice among its pending candidates (after the head
// kernel); frame_ubo, when not null, is the device UBO holding the frame
// number (else cache.frame).
hipError_t launch_cache_choice(const FrameDev& fr, const CacheCfg& cache, const UniformBufferObject* frame_ubo,
                               hipStream_t st);

// Output of the GPU builder.  All arrays are device allocations owned by the
// caller after a successful build (free with hipFree).
struct BvhBuildResult {
    BvhNode* nodes = nullptr;
    BvhTri* tris = nullptr;
    uint32_t node_count = 0;
    uint32_t tri_count = 0;
    uint32_t max_depth = 0;   // edges on the longest root-to-leaf path
    float build_ms = 0.0f;
};

// LBVH (Karras 2012) over all triangles of the scene.
//   vertices/indices/offsets: device copies of the reference buffers
//   inst_first_tri: device array, first canonical triangle of each instance
//   (instance_count + 1 entries; the last is the total triangle count)
hipError_t build_lbvh(const float* vertices, const uint32_t* indices, const uint2* offsets,
                      const uint32_t* inst_first_tri, uint32_t instance_count, uint32_t tri_count,
                      hipStream_t st, BvhBuildResult* out, char* err, size_t err_len);
hipError_t build_ploc(const float* vertices, const uint32_t* indices, const uint2* offsets,
                      const uint32_t* inst_first_tri, uint32_t instance_count, uint32_t tri_count,
                      hipStream_t st, BvhBuildResult* out, char* err, size_t err_len);
// Binned SAH on the host (canonical triangles gathered on the device).
hipError_t build_sah(const float* vertices, const uint32_t* indices, const uint2* offsets,
                     const uint32_t* inst_first_tri, uint32_t instance_count, uint32_t tri_count,
                     hipStream_t st, BvhBuildResult* out, char* err, size_t err_len);
// The same binned SAH built on the GPU (the same tree: every split is a
// function of the node's triangle set).
hipError_t build_sah_gpu(const float* vertices, const uint32_t* indices, const uint2* offsets,
                         const uint32_t* inst_first_tri, uint32_t instance_count, uint32_t tri_count,
                         hipStream_t st, BvhBuildResult* out, char* err, size_t err_len);

// The traversed tree: the binary tree of a build collapsed top-down into
// nodes of up to `width` (2..4) children.  `nodes` is a device allocation
// owned by the caller after a successful call.
struct WideBuildResult {
    WideNode* nodes = nullptr;
    uint32_t node_count = 0;
    uint32_t depth = 0;        // wide nodes on the longest root-to-leaf path
    uint32_t stack_bound = 0;  // largest sum of (children - 1) over such a path
    float build_ms = 0.0f;
};
hipError_t build_wide(const BvhNode* binary, uint32_t binary_count, int width, hipStream_t st, WideBuildResult* out,
                      char* err, size_t err_len);

}  // namespace xr

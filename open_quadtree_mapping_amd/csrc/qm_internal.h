// qm_internal.h -- helpers shared by the host translation units of libqm.
#pragma once

#include <hip/hip_runtime.h>
#include <string>

#define QM_HIP(call)                                                                                       \
  do {                                                                                                     \
    hipError_t e_ = (call);                                                                                \
    if (e_ != hipSuccess) return ::qm::fail(QM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace qm {
struct FrameDev;
// geometry fields of a stage-level descriptor (sizes, pyramid levels, chain
// splits, weight edges); every other field zero
void init_geometry(FrameDev& F, int W, int H, int pitch);
// set the thread-local message returned by qm_last_error and pass the code through
int fail(int code, const std::string& msg);
// device allocation that honours the test cap and the byte accounting
int dev_alloc(void** p, size_t bytes);
void dev_free(void* p);
// remap a SW x SH source into a W x H destination through per-pixel maps
void launch_remap_u8(const uint8_t* src, int SW, int SH, int spitch, int W, int H, const float* mapx,
                     const float* mapy, uint8_t* dst, hipStream_t s);
// semi-dense keyframe seeds (k_semidense.hip); seeds are W x H x 2 float4
void sd_gradient(const uint8_t* img, int W, int H, int pitch, float2* grad, hipStream_t s);
void sd_init(const float2* grad, int W, int H, float4* seeds, hipStream_t s);
void sd_step(float4* seeds[2], const uint8_t* key, const float2* key_grad, const uint8_t* inc,
             const float2* inc_grad, int W, int H, int pitch, const float cam[4], const float T_key_world[12],
             const float T_inc_world[12], bool switch_keyframe, float* prefill, float2* prior,
             unsigned long long* u64, float4* info, hipStream_t s);
// world-frame point cloud of a depth map (k_output.hip); xyzi holds W*H
// float4 (x, y, z, intensity); *count (host) receives the number of points
int point_cloud(const float* depth, const uint8_t* ref, int W, int H, const float cam[4], const float T[12],
                float* xyzi, int* count, hipStream_t s);
}  // namespace qm

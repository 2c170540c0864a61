// k_output.hip -- back-projection of a fused depth map into a world-frame
// point cloud with intensities, replacing the CPU loop of
// Publisher::publishPointCloud (src/publisher.cpp:129-164).
//
// Points keep the reference's raster order: pixel (x, y) with depth d not
// below 0.1 (a NaN depth is kept, as the reference's "d < 0.1" test lets it
// through) becomes T_world_ref * (normalize(K^-1 (x, y, 1)) * d) and the
// reference image's grey level. Ordered compaction: per-block counts, one
// exclusive scan of the block counts, then each block writes its points at its
// offset plus the wave-ballot prefix of each lane.
#include <string>

#include "../../include/qm.h"
#include "qm_device.h"
#include "qm_internal.h"

namespace qm {

__device__ __forceinline__ bool keep_point(float d) { return !((double)d < 0.1); }

__global__ __launch_bounds__(256) void k_pc_count(const float* __restrict__ depth, int n, int* __restrict__ counts) {
  __shared__ int wsum[4];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const bool keep = i < n && keep_point(depth[i]);
  const unsigned long long m = __ballot(keep);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) counts[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// exclusive scan of nb block counts by one block of 1024 threads (chunks of 1024)
__global__ __launch_bounds__(1024) void k_pc_scan(int* __restrict__ counts, int nb, int* __restrict__ total) {
  __shared__ int s[1024];
  int carry = 0;
  for (int base = 0; base < nb; base += 1024) {
    const int i = base + threadIdx.x;
    const int v = i < nb ? counts[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) counts[i] = carry + s[threadIdx.x] - v;
    const int chunk = s[1023];
    __syncthreads();
    carry += chunk;
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(256) void k_pc_write(const float* __restrict__ depth, const uint8_t* __restrict__ ref,
                                                  int W, int H, float fx, float fy, float cx, float cy, Se3 T,
                                                  const int* __restrict__ offsets, float4* __restrict__ out) {
  __shared__ int wsum[4];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int n = W * H;
  const float d = i < n ? depth[i] : -1.0f;
  const bool keep = i < n && keep_point(d);
  const unsigned long long m = __ballot(keep);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) wsum[w] = __popcll(m);
  __syncthreads();
  if (!keep) return;
  int off = offsets[blockIdx.x];
  for (int k = 0; k < w; ++k) off += wsum[k];
  off += __popcll(m & ((1ull << lane) - 1ull));
  const int y = i / W, x = i - y * W;
  const V3 f = ray_dir(x, y, fx, fy, cx, cy);
  V3 p;
  p.x = f.x * d;
  p.y = f.y * d;
  p.z = f.z * d;
  const V3 q = se3_apply(T.m, p);
  out[off] = make_float4(q.x, q.y, q.z, (float)ref[i]);
}

int point_cloud(const float* depth, const uint8_t* ref, int W, int H, const float cam[4], const float T[12],
                float* xyzi, int* count, hipStream_t s) {
  const int n = W * H;
  const int nb = (n + 255) / 256;
  void* scratch;
  int rc = dev_alloc(&scratch, sizeof(int) * (nb + 1));
  if (rc) return rc;
  int* counts = (int*)scratch;
  Se3 t;
  for (int k = 0; k < 12; ++k) t.m[k] = T[k];
  hipLaunchKernelGGL(k_pc_count, dim3(nb), dim3(256), 0, s, depth, n, counts);
  hipLaunchKernelGGL(k_pc_scan, dim3(1), dim3(1024), 0, s, counts, nb, counts + nb);
  hipLaunchKernelGGL(k_pc_write, dim3(nb), dim3(256), 0, s, depth, ref, W, H, cam[0], cam[1], cam[2], cam[3], t,
                     counts, (float4*)xyzi);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(count, counts + nb, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dev_free(scratch);
  if (e != hipSuccess) return fail(QM_ERR_HIP, std::string("point cloud: ") + hipGetErrorString(e));
  return QM_OK;
}

}  // namespace qm

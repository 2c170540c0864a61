// qm_noalloc.hip -- qm_fuse_noalloc: the fusion stage of qm_fuse without any
// device allocation, for callers that cannot allocate (tight serving loops,
// harnesses that forbid hipMalloc). Its scratch -- the stage descriptor and
// the 64-bit z-buffer -- lives in one static device array per GPU.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "../../include/qm.h"
#include "qm_device.h"
#include "qm_internal.h"
#include "qm_launch.h"

using namespace qm;

namespace {
constexpr size_t kDescBytes = 4096;                  // the FrameDev descriptor slot
constexpr size_t kMaxPixels = (size_t)1 << 21;       // 2097152 pixels (1920x1080 fits)
constexpr size_t kScratchBytes = kDescBytes + 8 * kMaxPixels;  // + z-buffer: 16 MiB + 4 KiB
static_assert(sizeof(FrameDev) <= kDescBytes, "descriptor slot too small");
}  // namespace

// One copy per GPU, held until the process ends. Host code never uses this
// name as an address: it takes the device address from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices];  // device address of `scratch` on each device, once looked up

// The address of the current device's copy of `scratch`, looked up on the
// first call on that device (with the device current) and kept.
int scratch_address(void** out) {
  int dev = 0;
  QM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(QM_ERR_INVALID, "qm_fuse_noalloc: device index beyond 63");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    QM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(QM_ERR_HIP, "qm_fuse_noalloc: no device address for the static scratch");
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return QM_OK;
}
}  // namespace

extern "C" int qm_fuse_noalloc(float* seeds, const float* depth_in, int W, int H, const float cam[4],
                               const float T_last_to_cur[12], float* depth_out, float* seeds_out, int32_t* table,
                               void* stream) {
  if (W < 16 || H < 16 || W > 16384 || H > 16384) return fail(QM_ERR_INVALID, "image size out of range [16, 16384]");
  if ((size_t)W * (size_t)H > kMaxPixels)
    return fail(QM_ERR_INVALID,
                "qm_fuse_noalloc: W * H above 2097152 pixels, the capacity of its static z-buffer (use qm_fuse)");
  if (!seeds || !depth_in || !cam || !T_last_to_cur || !seeds_out) return fail(QM_ERR_INVALID, "NULL pointer");
  void* base = nullptr;
  int rc = scratch_address(&base);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  FrameDev F;
  init_geometry(F, W, H, W);
  F.fx = cam[0];
  F.fy = cam[1];
  F.cx = cam[2];
  F.cy = cam[3];
  std::memcpy(F.T_fuse, T_last_to_cur, sizeof(float) * 12);
  F.seeds = (float4*)seeds;
  F.seeds_new = (float4*)seeds_out;
  F.dense = const_cast<float*>(depth_in);
  F.depth_out = depth_out;
  F.table_out = table;
  F.zbuf = (unsigned long long*)((char*)base + kDescBytes);
  // descriptor into the scratch's first slot; the z-buffer is filled by a
  // kernel before the warp (it keeps the previous call's contents until then)
  QM_HIP(hipMemcpyAsync(base, &F, sizeof(F), hipMemcpyHostToDevice, s));
  launch_fuse((const FrameDev*)base, &F, 1, true, s);
  QM_HIP(hipGetLastError());
  QM_HIP(hipStreamSynchronize(s));
  return QM_OK;
}

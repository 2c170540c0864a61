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
// set the thread-local message returned by qm_last_error and pass the code through
int fail(int code, const std::string& msg);
// device allocation that honours the test cap and the byte accounting
int dev_alloc(void** p, size_t bytes);
void dev_free(void* p);
void launch_remap_u8(const uint8_t* src, int W, int H, int spitch, const float* mapx, const float* mapy, uint8_t* dst,
                     hipStream_t s);
}  // namespace qm

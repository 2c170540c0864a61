// qm_ctx.hip -- the Depthmap-level state machine of one camera stream
// (reference: quadmap::Depthmap, src/depthmap.cpp:19-68, over
// SeedMatrix::add_frames and friends, src/seed_matrix.cu:81-198,448-612).
//
// Per frame: upload the uint8 image (optionally remapped on the GPU), update
// the semi-dense keyframe seeds (the sparse map's pre-fill), run the frame
// gates on the host, and when a depth map is due run the whole keyframe
// pipeline as one hipGraph replay of a one-stream qm_batch. Reference frames
// live in a ring of device slots (no per-frame allocation); the fusion seeds
// live in the batch and persist across keyframes.
#include <hip/hip_runtime.h>

#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "../../include/qm.h"
#include "qm_internal.h"
#include "qm_se3.h"

using namespace qm;

namespace {
struct RefEntry {
  int slot;
  float T[12];  // T_ref_world
};
}  // namespace

struct qm_ctx {
  int W = 0, H = 0;
  float cam[4] = {0, 0, 0, 0};
  qm_params p;
  int ratio = 1;
  int frame_index = 0;
  bool initialized = false;
  int nslots = 0;
  uint8_t* slots = nullptr;  // nslots x H x W
  std::vector<int> free_slots;
  std::deque<RefEntry> refs;  // front = newest
  int income_slot = -1;
  float income_T[12];
  float fuse_pose[12];  // world->camera pose of the last fused keyframe
  qm_batch* batch = nullptr;
  float* d_depth = nullptr;
  float* d_sparse = nullptr;
  float* d_mapx = nullptr;
  float* d_mapy = nullptr;
  uint8_t* d_raw = nullptr;
  size_t raw_bytes = 0;
  uint8_t* h_stage = nullptr;
  float* h_depth = nullptr;
  float* h_debug = nullptr;
  uint8_t* h_ref = nullptr;
  float T_world_ref[12];
  bool have_result = false;
  hipStream_t stream = nullptr;
  // semi-dense keyframe seeds (the sparse map's pre-fill)
  bool semidense = true;
  uint8_t* sd_key = nullptr;        // keyframe image
  float2* sd_key_grad = nullptr;
  float2* sd_inc_grad = nullptr;
  float4* sd_seeds[2] = {nullptr, nullptr};
  unsigned long long* sd_u64 = nullptr;
  float4* sd_info = nullptr;
  float* sd_prefill = nullptr;
  float2* sd_prior = nullptr;  // (inverse depth, variance) prior of the income frame
  bool prior_cost = false, prior_quadtree = false;
  float sd_key_T[12];
  uint8_t* d_ref = nullptr;  // reference image of the last result

  uint8_t* slot_ptr(int s) const { return slots + (size_t)s * W * H; }
};

namespace {

void ctx_free(qm_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->batch) qm_batch_destroy(c->batch);
  dev_free(c->slots);
  dev_free(c->d_depth);
  dev_free(c->d_sparse);
  dev_free(c->d_mapx);
  dev_free(c->d_mapy);
  dev_free(c->d_raw);
  dev_free(c->sd_key);
  dev_free(c->sd_key_grad);
  dev_free(c->sd_inc_grad);
  dev_free(c->sd_seeds[0]);
  dev_free(c->sd_seeds[1]);
  dev_free(c->sd_u64);
  dev_free(c->sd_info);
  dev_free(c->sd_prefill);
  dev_free(c->sd_prior);
  dev_free(c->d_ref);
  if (c->h_stage) (void)hipHostFree(c->h_stage);
  if (c->h_depth) (void)hipHostFree(c->h_depth);
  if (c->h_debug) (void)hipHostFree(c->h_debug);
  if (c->h_ref) (void)hipHostFree(c->h_ref);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int take_slot(qm_ctx* c) {
  const int s = c->free_slots.back();
  c->free_slots.pop_back();
  return s;
}

// keyframe descriptor: key = income image, slots {0,0,1,2,3} -> the four newest references
void make_desc(qm_ctx* c, qm_frame_desc* d) {
  std::memset(d, 0, sizeof(*d));
  d->W = c->W;
  d->H = c->H;
  d->pitch = c->W;
  std::memcpy(d->cam, c->cam, sizeof(d->cam));
  d->key = c->slot_ptr(c->income_slot);
  float inv_income[12];
  se3_inv(c->income_T, inv_income);
  for (int r = 0; r < 4; ++r) {
    const RefEntry& e = c->refs[r];
    d->refs[r] = c->slot_ptr(e.slot);
    se3_mul(e.T, inv_income, d->T_ref_key[r]);  // src/seed_matrix.cu:462
  }
  float inv_fuse[12];
  se3_inv(c->fuse_pose, inv_fuse);
  se3_mul(c->income_T, inv_fuse, d->T_last_to_cur);  // src/seed_matrix.cu:574
  d->prefill = c->semidense ? c->sd_prefill : nullptr;
  d->prior = (c->semidense && c->prior_cost) ? (const float*)c->sd_prior : nullptr;
  d->prior_depth = (c->semidense && c->prior_quadtree) ? c->sd_prefill : nullptr;
  d->depth_out = c->d_depth;
  d->sparse_out = c->d_sparse;
}

}  // namespace

extern "C" {

int qm_create(int W, int H, float fx, float fy, float cx, float cy, const qm_params* p, qm_ctx** out) {
  if (!out) return fail(QM_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (W < 16 || H < 16 || W > 16384 || H > 16384) return fail(QM_ERR_INVALID, "image size out of range [16, 16384]");
  if (!(fx != 0.0f) || !(fy != 0.0f)) return fail(QM_ERR_INVALID, "focal length is zero");
  qm_params prm;
  if (p)
    prm = *p;
  else
    qm_default_params(&prm);
  if (prm.min_refs < 4) return fail(QM_ERR_INVALID, "min_refs must be >= 4 (slots read four references)");
  if (prm.max_refs < prm.min_refs) return fail(QM_ERR_INVALID, "max_refs < min_refs");
  qm_ctx* c = new qm_ctx();
  c->W = W;
  c->H = H;
  c->cam[0] = fx;
  c->cam[1] = fy;
  c->cam[2] = cx;
  c->cam[3] = cy;
  c->p = prm;
  se3_identity(c->income_T);
  se3_identity(c->fuse_pose);
  se3_identity(c->T_world_ref);
  const size_t px = (size_t)W * H;
  c->nslots = prm.max_refs + 1;
  int rc;
  void* v;
  if ((rc = dev_alloc(&v, px * c->nslots))) { ctx_free(c); return rc; }
  c->slots = (uint8_t*)v;
  if ((rc = dev_alloc(&v, px * 4))) { ctx_free(c); return rc; }
  c->d_depth = (float*)v;
  if ((rc = dev_alloc(&v, px * 4))) { ctx_free(c); return rc; }
  c->d_sparse = (float*)v;
  if ((rc = dev_alloc(&v, px))) { ctx_free(c); return rc; }
  c->sd_key = (uint8_t*)v;
  if ((rc = dev_alloc(&v, px * 8))) { ctx_free(c); return rc; }
  c->sd_key_grad = (float2*)v;
  if ((rc = dev_alloc(&v, px * 8))) { ctx_free(c); return rc; }
  c->sd_inc_grad = (float2*)v;
  for (int k = 0; k < 2; ++k) {
    if ((rc = dev_alloc(&v, px * 32))) { ctx_free(c); return rc; }
    c->sd_seeds[k] = (float4*)v;
  }
  if ((rc = dev_alloc(&v, px * 8))) { ctx_free(c); return rc; }
  c->sd_u64 = (unsigned long long*)v;
  if ((rc = dev_alloc(&v, px * 16))) { ctx_free(c); return rc; }
  c->sd_info = (float4*)v;
  if ((rc = dev_alloc(&v, px * 4))) { ctx_free(c); return rc; }
  c->sd_prefill = (float*)v;
  if ((rc = dev_alloc(&v, px * 8))) { ctx_free(c); return rc; }
  c->sd_prior = (float2*)v;
  if ((rc = dev_alloc(&v, px))) { ctx_free(c); return rc; }
  c->d_ref = (uint8_t*)v;
  se3_identity(c->sd_key_T);
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_stage, px, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_depth, px * 4, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_debug, px * 4, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_ref, px, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMemset(c->slots, 0, px * c->nslots);
  if (e != hipSuccess) {
    ctx_free(c);
    return fail(QM_ERR_HIP, std::string("qm_create: ") + hipGetErrorString(e));
  }
  for (int s = c->nslots - 1; s >= 0; --s) c->free_slots.push_back(s);
  // the keyframe batch is planned once; its inputs are re-pointed per keyframe
  qm_frame_desc d;
  std::memset(&d, 0, sizeof(d));
  d.W = W;
  d.H = H;
  d.pitch = W;
  std::memcpy(d.cam, c->cam, sizeof(d.cam));
  d.key = c->slot_ptr(0);
  for (int r = 0; r < 4; ++r) {
    d.refs[r] = c->slot_ptr(0);
    se3_identity(d.T_ref_key[r]);
  }
  se3_identity(d.T_last_to_cur);
  d.depth_out = c->d_depth;
  d.sparse_out = c->d_sparse;
  if ((rc = qm_batch_create(&d, 1, &c->p, &c->batch))) {
    ctx_free(c);
    return rc;
  }
  *out = c;
  return QM_OK;
}

int qm_destroy(qm_ctx* c) {
  ctx_free(c);
  return QM_OK;
}

int qm_set_semi2dense_ratio(qm_ctx* c, int ratio) {
  if (!c) return fail(QM_ERR_INVALID, "NULL context");
  if (ratio < 1) return fail(QM_ERR_INVALID, "ratio must be >= 1");
  c->ratio = ratio;
  return QM_OK;
}

int qm_set_semidense(qm_ctx* c, int on) {
  if (!c) return fail(QM_ERR_INVALID, "NULL context");
  if (c->initialized && (on != 0) != c->semidense)
    return fail(QM_ERR_STATE, "the semi-dense setting is fixed once frames have been added");
  c->semidense = on != 0;
  return QM_OK;
}

int qm_set_prior_alternates(qm_ctx* c, int prior_cost, int prior_quadtree) {
  if (!c) return fail(QM_ERR_INVALID, "NULL context");
  if (c->initialized) return fail(QM_ERR_STATE, "the prior alternates are fixed once frames have been added");
  c->prior_cost = prior_cost != 0;
  c->prior_quadtree = prior_quadtree != 0;
  return QM_OK;
}

int qm_set_remap(qm_ctx* c, const float* map_x, const float* map_y) {
  if (!c) return fail(QM_ERR_INVALID, "NULL context");
  if ((map_x == nullptr) != (map_y == nullptr)) return fail(QM_ERR_INVALID, "give both maps or neither");
  const size_t px = (size_t)c->W * c->H;
  if (!map_x) {
    dev_free(c->d_mapx);
    dev_free(c->d_mapy);
    c->d_mapx = c->d_mapy = nullptr;
    return QM_OK;
  }
  int rc;
  void* v;
  if (!c->d_mapx) {
    if ((rc = dev_alloc(&v, px * 4))) return rc;
    c->d_mapx = (float*)v;
    if ((rc = dev_alloc(&v, px * 4))) return rc;
    c->d_mapy = (float*)v;
  }
  QM_HIP(hipMemcpy(c->d_mapx, map_x, px * 4, hipMemcpyHostToDevice));
  QM_HIP(hipMemcpy(c->d_mapy, map_y, px * 4, hipMemcpyHostToDevice));
  return QM_OK;
}

int qm_add_frame(qm_ctx* c, const uint8_t* img, int img_w, int img_h, int pitch, const float T_curr_world[12],
                 int* has_result) {
  if (!c || !img || !T_curr_world || !has_result) return fail(QM_ERR_INVALID, "NULL pointer");
  if (img_w <= 0 || img_h <= 0 || pitch < img_w) return fail(QM_ERR_INVALID, "bad image size or pitch");
  if (!c->d_mapx && (img_w != c->W || img_h != c->H))
    return fail(QM_ERR_INVALID, "without a remap the image must be the depth map size");
  for (int i = 0; i < 12; ++i)
    if (!(T_curr_world[i] == T_curr_world[i])) return fail(QM_ERR_INVALID, "pose has NaN entries");
  *has_result = 0;
  const int W = c->W, H = c->H;
  const size_t px = (size_t)W * H;
  c->frame_index++;
  if (c->income_slot < 0) c->income_slot = take_slot(c);
  uint8_t* slot = c->slot_ptr(c->income_slot);
  // SeedMatrix::input_raw / add_income_image: (remap) + upload
  if (c->d_mapx) {
    const size_t raw = (size_t)img_w * img_h;
    if (raw > c->raw_bytes) {
      QM_HIP(hipStreamSynchronize(c->stream));
      dev_free(c->d_raw);
      c->d_raw = nullptr;
      c->raw_bytes = 0;
      if (c->h_stage) (void)hipHostFree(c->h_stage);
      c->h_stage = nullptr;
      void* v;
      int rc = dev_alloc(&v, raw);
      if (rc) return rc;
      c->d_raw = (uint8_t*)v;
      QM_HIP(hipHostMalloc((void**)&c->h_stage, raw > px ? raw : px, hipHostMallocDefault));
      c->raw_bytes = raw;
    }
    for (int y = 0; y < img_h; ++y) std::memcpy(c->h_stage + (size_t)y * img_w, img + (size_t)y * pitch, img_w);
    QM_HIP(hipMemcpyAsync(c->d_raw, c->h_stage, raw, hipMemcpyHostToDevice, c->stream));
    launch_remap_u8(c->d_raw, img_w, img_h, img_w, W, H, c->d_mapx, c->d_mapy, slot, c->stream);
    QM_HIP(hipGetLastError());
  } else {
    for (int y = 0; y < H; ++y) std::memcpy(c->h_stage + (size_t)y * W, img + (size_t)y * pitch, W);
    QM_HIP(hipMemcpyAsync(slot, c->h_stage, px, hipMemcpyHostToDevice, c->stream));
  }
  std::memcpy(c->income_T, T_curr_world, sizeof(float) * 12);
  // semi-dense keyframe seeds, every frame (src/seed_matrix.cu:113-139)
  if (c->semidense) {
    sd_gradient(slot, W, H, W, c->sd_inc_grad, c->stream);
    if (!c->initialized) {
      // set_income_as_keyframe + initial_keyframe (src/seed_matrix.cu:96-104,211-245)
      QM_HIP(hipMemcpyAsync(c->sd_key, slot, px, hipMemcpyDeviceToDevice, c->stream));
      QM_HIP(hipMemcpyAsync(c->sd_key_grad, c->sd_inc_grad, px * 8, hipMemcpyDeviceToDevice, c->stream));
      std::memcpy(c->sd_key_T, c->income_T, sizeof(c->sd_key_T));
      sd_init(c->sd_key_grad, W, H, c->sd_seeds[0], c->stream);
    } else {
      // update_keyframe, then need_switchkeyframe / create_new_keyframe (src/seed_matrix.cu:130-139)
      float cosv, base;
      se3_axis_cos_baseline(c->sd_key_T, c->income_T, &cosv, &base);
      const bool sw = ((double)cosv < 0.86) || ((double)base > 0.5);
      sd_step(c->sd_seeds, c->sd_key, c->sd_key_grad, slot, c->sd_inc_grad, W, H, W, c->cam, c->sd_key_T, c->income_T,
              sw, c->sd_prefill, c->prior_cost ? c->sd_prior : nullptr, c->sd_u64, c->sd_info, c->stream);
      if (sw) {
        QM_HIP(hipMemcpyAsync(c->sd_key, slot, px, hipMemcpyDeviceToDevice, c->stream));
        QM_HIP(hipMemcpyAsync(c->sd_key_grad, c->sd_inc_grad, px * 8, hipMemcpyDeviceToDevice, c->stream));
        std::memcpy(c->sd_key_T, c->income_T, sizeof(c->sd_key_T));
      }
    }
    QM_HIP(hipGetLastError());
  }
  if (!c->initialized) {
    // src/seed_matrix.cu:122-128: the first frame only initialises and reports
    // a result; the reference's map is uninitialised memory there, here zeros
    c->initialized = true;
    std::memset(c->h_depth, 0, px * 4);
    std::memset(c->h_debug, 0, px * 4);
    QM_HIP(hipMemsetAsync(c->d_depth, 0, px * 4, c->stream));
    QM_HIP(hipMemcpyAsync(c->d_ref, slot, px, hipMemcpyDeviceToDevice, c->stream));
    QM_HIP(hipMemcpyAsync(c->h_ref, slot, px, hipMemcpyDeviceToHost, c->stream));
    QM_HIP(hipStreamSynchronize(c->stream));
    se3_inv(T_curr_world, c->T_world_ref);
    c->have_result = true;
    *has_result = 1;
    return QM_OK;
  }
  if (c->frame_index % c->ratio != 0) {  // src/seed_matrix.cu:144
    QM_HIP(hipStreamSynchronize(c->stream));
    return QM_OK;
  }
  const bool has_output = (int)c->refs.size() >= c->p.min_refs;  // size() > 10
  if (has_output) {
    qm_frame_desc d;
    make_desc(c, &d);
    int rc = qm_batch_set_frame(c->batch, 0, &d);
    if (rc) return rc;
    if ((rc = qm_batch_run(c->batch, 1, c->stream))) return rc;
  }
  // need_add_reference / add_reference (src/seed_matrix.cu:171-198)
  bool add = c->refs.empty();
  if (!add) {
    float cosv, base;
    se3_axis_cos_baseline(c->refs.front().T, c->income_T, &cosv, &base);
    add = (cosv < c->p.ref_cos) || (base > c->p.ref_baseline);
  }
  const int used_slot = c->income_slot;
  if (add) {
    RefEntry e;
    e.slot = c->income_slot;
    std::memcpy(e.T, c->income_T, sizeof(e.T));
    c->refs.push_front(e);
    c->income_slot = -1;
    if ((int)c->refs.size() > c->p.max_refs) {
      c->free_slots.push_back(c->refs.back().slot);
      c->refs.pop_back();
    }
  }
  if (has_output) {
    std::memcpy(c->fuse_pose, c->income_T, sizeof(float) * 12);  // src/seed_matrix.cu:592
    QM_HIP(hipMemcpyAsync(c->h_depth, c->d_depth, px * 4, hipMemcpyDeviceToHost, c->stream));
    QM_HIP(hipMemcpyAsync(c->h_debug, c->d_sparse, px * 4, hipMemcpyDeviceToHost, c->stream));
    QM_HIP(hipMemcpyAsync(c->d_ref, c->slot_ptr(used_slot), px, hipMemcpyDeviceToDevice, c->stream));
    QM_HIP(hipMemcpyAsync(c->h_ref, c->slot_ptr(used_slot), px, hipMemcpyDeviceToHost, c->stream));
    QM_HIP(hipStreamSynchronize(c->stream));
    se3_inv(T_curr_world, c->T_world_ref);
    c->have_result = true;
    *has_result = 1;
  } else {
    QM_HIP(hipStreamSynchronize(c->stream));
  }
  return QM_OK;
}

int qm_get_depth(qm_ctx* c, float* out) {
  if (!c || !out) return fail(QM_ERR_INVALID, "NULL pointer");
  if (!c->have_result) return fail(QM_ERR_STATE, "no result yet");
  std::memcpy(out, c->h_depth, (size_t)c->W * c->H * 4);
  return QM_OK;
}

int qm_get_debug(qm_ctx* c, float* out) {
  if (!c || !out) return fail(QM_ERR_INVALID, "NULL pointer");
  if (!c->have_result) return fail(QM_ERR_STATE, "no result yet");
  std::memcpy(out, c->h_debug, (size_t)c->W * c->H * 4);
  return QM_OK;
}

int qm_get_reference(qm_ctx* c, uint8_t* out) {
  if (!c || !out) return fail(QM_ERR_INVALID, "NULL pointer");
  if (!c->have_result) return fail(QM_ERR_STATE, "no result yet");
  std::memcpy(out, c->h_ref, (size_t)c->W * c->H);
  return QM_OK;
}

int qm_get_T_world_ref(qm_ctx* c, float out[12]) {
  if (!c || !out) return fail(QM_ERR_INVALID, "NULL pointer");
  std::memcpy(out, c->T_world_ref, sizeof(float) * 12);
  return QM_OK;
}

int qm_get_point_cloud(qm_ctx* c, float* xyzi, int* count) {
  if (!c || !xyzi || !count) return fail(QM_ERR_INVALID, "NULL pointer");
  if (!c->have_result) return fail(QM_ERR_STATE, "no result yet");
  const size_t px = (size_t)c->W * c->H;
  void* v;
  int rc = dev_alloc(&v, px * 16);
  if (rc) return rc;
  rc = point_cloud(c->d_depth, c->d_ref, c->W, c->H, c->cam, c->T_world_ref, (float*)v, count, c->stream);
  if (rc == QM_OK && *count > 0) {
    const hipError_t e = hipMemcpy(xyzi, v, (size_t)*count * 16, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = fail(QM_ERR_HIP, std::string("point cloud download: ") + hipGetErrorString(e));
  }
  dev_free(v);
  return rc;
}

int qm_num_references(qm_ctx* c) {
  if (!c) return fail(QM_ERR_INVALID, "NULL context");
  return (int)c->refs.size();
}

}  // extern "C"

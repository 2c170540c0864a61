// quadmap/depthmap.h -- drop-in quadmap::Depthmap over libqm.so.
//
// Same class, constructor and methods as the reference's
// include/quadmap/depthmap.h:31-89 (note the constructor order fx, cx, fy, cy),
// implemented inline over the C ABI of include/qm.h; the pipeline runs on the
// GPU. Failures throw quadmap::QmException (the reference throws CudaException).
// Link with -lqm (open_quadtree_mapping_amd/libqm.so).
#pragma once

#include <cstddef>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../qm.h"
#include "cv_compat.h"
#include "se3.h"

namespace quadmap {

class QmException : public std::runtime_error {
 public:
  QmException(const std::string& what, int code) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void qm_throw_if(int rc, const char* where) {
  if (rc != QM_OK) throw QmException(std::string(where) + ": " + qm_last_error(), rc);
}

class Depthmap {
 public:
  Depthmap(size_t width, size_t height, float fx, float cx, float fy, float cy, cv::Mat remap_1, cv::Mat remap_2,
           int semi2dense_ratio)
      : width_(width), height_(height), fx_(fx), fy_(fy), cx_(cx), cy_(cy) {
    qm_throw_if(qm_create((int)width, (int)height, fx, fy, cx, cy, nullptr, &ctx_), "qm_create");
    qm_throw_if(qm_set_semi2dense_ratio(ctx_, semi2dense_ratio), "qm_set_semi2dense_ratio");
    if (!remap_1.empty() && !remap_2.empty()) {
      if (remap_1.type() != CV_32FC1 || remap_2.type() != CV_32FC1 || remap_1.rows != (int)height ||
          remap_1.cols != (int)width || remap_2.rows != (int)height || remap_2.cols != (int)width)
        throw QmException("remap maps must be CV_32FC1 height x width coordinate maps", QM_ERR_INVALID);
      std::vector<float> mx(width * height), my(width * height);
      for (int y = 0; y < (int)height; ++y) {
        std::memcpy(&mx[(size_t)y * width], remap_1.ptr<float>(y), width * 4);
        std::memcpy(&my[(size_t)y * width], remap_2.ptr<float>(y), width * 4);
      }
      qm_throw_if(qm_set_remap(ctx_, mx.data(), my.data()), "qm_set_remap");
    }
  }
  ~Depthmap() { qm_destroy(ctx_); }
  Depthmap(const Depthmap&) = delete;
  Depthmap& operator=(const Depthmap&) = delete;

  bool add_frames(const cv::Mat& img_curr, const SE3<float>& T_curr_world) {
    if (img_curr.type() != CV_8UC1) throw QmException("add_frames expects a CV_8UC1 image", QM_ERR_INVALID);
    int has = 0;
    qm_throw_if(qm_add_frame(ctx_, img_curr.data, img_curr.cols, img_curr.rows, (int)img_curr.step,
                             T_curr_world.data.data, &has),
                "qm_add_frame");
    if (has) {
      std::lock_guard<std::mutex> lk(ref_img_mutex_);
      depth_out_.create((int)height_, (int)width_, CV_32FC1);
      debug_out_.create((int)height_, (int)width_, CV_32FC1);
      reference_out_.create((int)height_, (int)width_, CV_8UC1);
      std::vector<float> buf(width_ * height_);
      qm_throw_if(qm_get_depth(ctx_, buf.data()), "qm_get_depth");
      for (int y = 0; y < (int)height_; ++y) std::memcpy(depth_out_.ptr<float>(y), &buf[(size_t)y * width_], width_ * 4);
      qm_throw_if(qm_get_debug(ctx_, buf.data()), "qm_get_debug");
      for (int y = 0; y < (int)height_; ++y) std::memcpy(debug_out_.ptr<float>(y), &buf[(size_t)y * width_], width_ * 4);
      std::vector<unsigned char> ref(width_ * height_);
      qm_throw_if(qm_get_reference(ctx_, ref.data()), "qm_get_reference");
      for (int y = 0; y < (int)height_; ++y)
        std::memcpy(reference_out_.ptr<unsigned char>(y), &ref[(size_t)y * width_], width_);
      qm_throw_if(qm_get_T_world_ref(ctx_, T_world_ref_.data.data), "qm_get_T_world_ref");
    }
    return has != 0;
  }

  const cv::Mat_<float> getDepthmap() const { return cv::Mat_<float>(depth_out_); }
  const cv::Mat_<float> getDebugmap() const { return cv::Mat_<float>(debug_out_); }
  const cv::Mat getReferenceImage() const { return reference_out_; }
  float getFx() const { return fx_; }
  float getFy() const { return fy_; }
  float getCx() const { return cx_; }
  float getCy() const { return cy_; }
  std::mutex& getRefImgMutex() { return ref_img_mutex_; }
  size_t getWidth() const { return width_; }
  size_t getHeight() const { return height_; }
  SE3<float> getT_world_ref() const { return T_world_ref_; }

 private:
  qm_ctx* ctx_ = nullptr;
  size_t width_, height_;
  float fx_, fy_, cx_, cy_;
  std::mutex ref_img_mutex_;
  SE3<float> T_world_ref_;
  cv::Mat depth_out_, reference_out_, debug_out_;
};

}  // namespace quadmap

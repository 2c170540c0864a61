// quadmap/cv_compat.h -- the cv::Mat surface quadmap::Depthmap uses.
//
// With OpenCV installed this just includes it. Without it (as here), it
// provides a minimal stand-in: a reference-counted 2-D single-channel buffer
// with rows, cols, type(), data, step, ptr<T>(row), at<T>(row, col), clone(),
// create() and the Mat_<float> view, so caller code written against the
// reference header compiles unchanged for the subset Depthmap needs.
#pragma once

#if defined(QUADMAP_USE_OPENCV) || __has_include(<opencv2/core.hpp>)
#include <opencv2/core.hpp>
#define QUADMAP_HAVE_OPENCV 1
#else
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#ifndef CV_8UC1
#define CV_8UC1 0
#endif
#ifndef CV_32FC1
#define CV_32FC1 5
#endif

namespace cv {

class Mat {
 public:
  int rows = 0, cols = 0;
  unsigned char* data = nullptr;
  size_t step = 0;

  Mat() = default;
  Mat(int r, int c, int t) { create(r, c, t); }
  Mat(int r, int c, int t, void* ext, size_t st = 0) : rows(r), cols(c), data((unsigned char*)ext), type_(t) {
    step = st ? st : (size_t)c * elem_size(t);
  }
  void create(int r, int c, int t) {
    rows = r;
    cols = c;
    type_ = t;
    step = (size_t)c * elem_size(t);
    buf_ = std::make_shared<std::vector<unsigned char>>((size_t)r * step);
    data = buf_->data();
  }
  int type() const { return type_; }
  bool empty() const { return data == nullptr || rows == 0 || cols == 0; }
  bool isContinuous() const { return step == (size_t)cols * elem_size(type_); }
  size_t elemSize() const { return elem_size(type_); }
  Mat clone() const {
    Mat m(rows, cols, type_);
    for (int y = 0; y < rows; ++y) std::memcpy(m.data + y * m.step, data + y * step, (size_t)cols * elemSize());
    return m;
  }
  template <typename T> T* ptr(int r) { return reinterpret_cast<T*>(data + (size_t)r * step); }
  template <typename T> const T* ptr(int r) const { return reinterpret_cast<const T*>(data + (size_t)r * step); }
  template <typename T> T& at(int r, int c) { return ptr<T>(r)[c]; }
  template <typename T> const T& at(int r, int c) const { return ptr<T>(r)[c]; }
  static size_t elem_size(int t) { return t == CV_32FC1 ? 4 : 1; }

 protected:
  int type_ = CV_8UC1;
  std::shared_ptr<std::vector<unsigned char>> buf_;
};

template <typename T>
class Mat_ : public Mat {
 public:
  Mat_() = default;
  Mat_(const Mat& m) : Mat(m) {}
  T& operator()(int r, int c) { return at<T>(r, c); }
  const T& operator()(int r, int c) const { return at<T>(r, c); }
};

}  // namespace cv
#endif

// Drives quadmap::Depthmap (include/quadmap/depthmap.h) exactly as the
// reference's ROS node does (construct, add_frames per image + pose, read the
// getters), from a raw input file, and writes the results for the test to
// compare. Input: int32 W, H, n, ratio; float fx, fy, cx, cy; n x (W*H uint8
// image, 12 float T_curr_world). Output: per frame int32 has_result, then when
// set W*H float depth, W*H float debug, 12 float T_world_ref.
#include <cstdio>
#include <cstdint>
#include <vector>

#include "quadmap/depthmap.h"

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
    return 2;
  }
  FILE* in = std::fopen(argv[1], "rb");
  FILE* out = std::fopen(argv[2], "wb");
  if (!in || !out) return 2;
  int32_t hdr[4];
  float cam[4];
  if (std::fread(hdr, 4, 4, in) != 4 || std::fread(cam, 4, 4, in) != 4) return 2;
  const int W = hdr[0], H = hdr[1], n = hdr[2], ratio = hdr[3];
  try {
    quadmap::Depthmap dm(W, H, cam[0], cam[2], cam[1], cam[3], cv::Mat(), cv::Mat(), ratio);
    cv::Mat img(H, W, CV_8UC1);
    for (int k = 0; k < n; ++k) {
      if (std::fread(img.data, 1, (size_t)W * H, in) != (size_t)W * H) return 2;
      float T[12];
      if (std::fread(T, 4, 12, in) != 12) return 2;
      quadmap::SE3<float> pose;
      for (int i = 0; i < 12; ++i) pose.data[i] = T[i];
      const int32_t has = dm.add_frames(img, pose) ? 1 : 0;
      std::fwrite(&has, 4, 1, out);
      if (has) {
        const cv::Mat_<float> d = dm.getDepthmap();
        const cv::Mat_<float> g = dm.getDebugmap();
        for (int y = 0; y < H; ++y) std::fwrite(d.ptr<float>(y), 4, W, out);
        for (int y = 0; y < H; ++y) std::fwrite(g.ptr<float>(y), 4, W, out);
        const quadmap::SE3<float> Twr = dm.getT_world_ref();
        std::fwrite(Twr.data.data, 4, 12, out);
      }
    }
  } catch (const quadmap::QmException& e) {
    std::fprintf(stderr, "QmException %d: %s\n", e.code(), e.what());
    return 1;
  }
  std::fclose(in);
  std::fclose(out);
  return 0;
}

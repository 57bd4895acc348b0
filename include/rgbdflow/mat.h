// mat.h -- minimal stand-ins for the OpenCV types at the reference's API
// boundary (cv::Mat_<T>, cv::Vec3b, cv::Vec3f, cv::Size).  OpenCV is not part
// of this build; these carry the same element layouts (row-major, contiguous,
// channels interleaved) so a caller that has OpenCV can wrap its buffers
// without conversion.
#pragma once
#include <cstdint>
#include <vector>

namespace rgbdflow {

template <class T, int N>
struct Vec {
  T val[N];
  T& operator[](int i) { return val[i]; }
  const T& operator[](int i) const { return val[i]; }
};
typedef Vec<uint8_t, 3> Vec3b;
typedef Vec<float, 3> Vec3f;

struct Size {
  int width = 0, height = 0;
  Size() = default;
  Size(int w, int h) : width(w), height(h) {}
};

template <class T>
struct Mat_ {
  int rows = 0, cols = 0;
  std::vector<T> buf;
  Mat_() = default;
  Mat_(int r, int c) { create(r, c); }
  void create(int r, int c) {
    rows = r;
    cols = c;
    buf.assign((size_t)r * c, T{});
  }
  bool empty() const { return buf.empty(); }
  Size size() const { return Size(cols, rows); }
  T& operator()(int i, int j) { return buf[(size_t)i * cols + j]; }
  const T& operator()(int i, int j) const { return buf[(size_t)i * cols + j]; }
  T* data() { return buf.data(); }
  const T* data() const { return buf.data(); }
};

}  // namespace rgbdflow

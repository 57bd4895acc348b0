// Image.h -- the container part of Ce Liu's Image<T> / DImage
// (include/celiu_flow/Image.h:458 `typedef Image<double> DImage`) as used at
// the scene-flow API boundary: row-major pixels, channels interleaved,
// zero-initialised on allocate, deep copies.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

template <class T>
class Image {
 public:
  Image() = default;
  Image(int width, int height, int nchannels = 1) { allocate(width, height, nchannels); }

  // Image.h:506-520: allocate zero-filled storage
  void allocate(int width, int height, int nchannels = 1) {
    imWidth = width;
    imHeight = height;
    nChannels = nchannels;
    pData.assign((size_t)width * height * nchannels, T(0));
  }
  template <class T1>
  void allocate(const Image<T1>& other) {
    allocate(other.width(), other.height(), other.nchannels());
  }
  void copyData(const Image<T>& other) { *this = other; }
  void clear() { allocate(0, 0, 1); }

  int width() const { return imWidth; }
  int height() const { return imHeight; }
  int nchannels() const { return nChannels; }
  int npixels() const { return imWidth * imHeight; }
  int nelements() const { return imWidth * imHeight * nChannels; }
  bool matchDimension(int w, int h, int c) const { return w == imWidth && h == imHeight && c == nChannels; }
  template <class T1>
  bool matchDimension(const Image<T1>& o) const {
    return matchDimension(o.width(), o.height(), o.nchannels());
  }

  T* data() { return pData.data(); }
  const T* data() const { return pData.data(); }
  T& operator()(uint32_t r, uint32_t c) { return pData[(size_t)r * imWidth + c]; }
  T operator()(uint32_t r, uint32_t c) const { return pData[(size_t)r * imWidth + c]; }
  T& operator()(uint32_t r, uint32_t c, uint32_t k) { return pData[((size_t)r * imWidth + c) * nChannels + k]; }
  T operator()(uint32_t r, uint32_t c, uint32_t k) const { return pData[((size_t)r * imWidth + c) * nChannels + k]; }
  T& operator[](size_t i) { return pData[i]; }
  const T& operator[](size_t i) const { return pData[i]; }

  // Image.h:1778-1782
  void Multiplywith(double value) {
    for (auto& x : pData) x *= value;
  }

 private:
  int imWidth = 0, imHeight = 0, nChannels = 1;
  std::vector<T> pData;
};

typedef Image<double> DImage;
typedef Image<float> FImage;

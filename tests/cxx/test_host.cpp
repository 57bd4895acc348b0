// Drives the drop-in C++ API's host functions on inputs written by
// tests/test_host_cxx.py: the reference-order hole fills
// (recursiveMedianFilter, recursiveMedianFilterSmallInvalidRegionsUsingColor)
// and the demo's ingestion helpers (readDepthMapValuesImg, downsizeFrame).
//
// usage: test_host <op> <input.bin> <output.bin>
//   fill        in: int32 W, H; f64 depth[H][W]            out: f64 depth[H][W]
//   colourfill  in: int32 W, H, maxRegion; f64 depth[H][W]; f64 bgr[H][W][3]
//                                                          out: f64 depth[H][W]
//   readdepth   in: int32 W, H; uint16 mm[H][W]             out: f32 metres[H][W]
//   downsize    in: int32 W, H; uint8 bgr[H][W][3]; f32 depth[H][W]
//               out: int32 w, h; uint8 bgr[h][w][3]; f32 depth[h][w]
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "rgbdflow/OpticalFlow.h"
#include "rgbdflow/runFlowTwoFrames.h"

using rgbdflow::Mat_;
using rgbdflow::Vec3b;

template <class T>
void rd(std::ifstream& in, T* p, size_t n) {
  if (!in.read(reinterpret_cast<char*>(p), sizeof(T) * n)) throw std::runtime_error("short input");
}
template <class T>
void wr(std::ofstream& out, const T* p, size_t n) {
  if (!out.write(reinterpret_cast<const char*>(p), sizeof(T) * n)) throw std::runtime_error("write failed");
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s fill|colourfill|readdepth|downsize input.bin output.bin\n", argv[0]);
    return 2;
  }
  try {
    const std::string op = argv[1];
    std::ifstream in(argv[2], std::ios::binary);
    std::ofstream out(argv[3], std::ios::binary);
    int32_t W, H;
    rd(in, &W, 1);
    rd(in, &H, 1);
    const size_t N = (size_t)W * H;
    if (op == "fill" || op == "colourfill") {
      int32_t maxRegion = 0;
      if (op == "colourfill") rd(in, &maxRegion, 1);
      DImage depth(W, H, 1);
      rd(in, depth.data(), N);
      if (op == "fill") {
        recursiveMedianFilter(depth);
      } else {
        DImage bgr(W, H, 3);
        rd(in, bgr.data(), 3 * N);
        recursiveMedianFilterSmallInvalidRegionsUsingColor(depth, bgr, (uint32_t)maxRegion);
      }
      wr(out, depth.data(), N);
    } else if (op == "readdepth") {
      Mat_<uint16_t> img(H, W);
      rd(in, img.data(), N);
      Mat_<float> depth;
      readDepthMapValuesImg(img, depth);
      wr(out, depth.data(), N);
    } else if (op == "downsize") {
      Mat_<Vec3b> img(H, W);
      Mat_<float> depth(H, W);
      rd(in, reinterpret_cast<uint8_t*>(img.data()), 3 * N);
      rd(in, depth.data(), N);
      downsizeFrame(img, depth);
      const int32_t w = img.cols, h = img.rows;
      wr(out, &w, 1);
      wr(out, &h, 1);
      wr(out, reinterpret_cast<const uint8_t*>(img.data()), 3 * (size_t)w * h);
      wr(out, depth.data(), (size_t)w * h);
    } else {
      throw std::runtime_error("unknown op " + op);
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "test_host: %s\n", e.what());
    return 1;
  }
  return 0;
}

// Drives the drop-in C++ API the way the reference's demo does (main.cpp ->
// runFlowTwoFramesSceneFlowExperiments -> writeSceneFlow), on inputs written
// by tests/test_gpu_dropin.py.
//
// usage: test_dropin <input.bin> <output.flo3> <maxIters> <relax_variant>
// input.bin: int32 W, H; float32 f, cx, cy; uint8 img1[H][W][3] (BGR);
// float32 depth1[H][W] (m); uint8 img2; float32 depth2; int32 n_levels;
// per level int32 w, h and float32 hpb[h][w][2].
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "rgbdflow/runFlowTwoFrames.h"

using rgbdflow::Mat_;
using rgbdflow::Vec3b;
using rgbdflow::Vec3f;

template <class T>
void rd(std::ifstream& in, T* p, size_t n) {
  if (!in.read(reinterpret_cast<char*>(p), sizeof(T) * n)) throw std::runtime_error("short input");
}

int main(int argc, char** argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s input.bin output.flo3 maxIters relax_variant\n", argv[0]);
    return 2;
  }
  try {
    std::ifstream in(argv[1], std::ios::binary);
    int32_t W, H;
    float f, cx, cy;
    rd(in, &W, 1);
    rd(in, &H, 1);
    rd(in, &f, 1);
    rd(in, &cx, 1);
    rd(in, &cy, 1);
    Mat_<Vec3b> img1(H, W), img2(H, W);
    Mat_<float> d1(H, W), d2(H, W);
    rd(in, img1.data(), (size_t)W * H);
    rd(in, d1.data(), (size_t)W * H);
    rd(in, img2.data(), (size_t)W * H);
    rd(in, d2.data(), (size_t)W * H);
    int32_t L;
    rd(in, &L, 1);
    std::map<int, std::vector<float>> maps;
    for (int k = 0; k < L; k++) {
      int32_t w, h;
      rd(in, &w, 1);
      rd(in, &h, 1);
      std::vector<float> m((size_t)w * h * 2);
      rd(in, m.data(), m.size());
      maps[w] = m;
    }
    OpticalFlow::setHeuristicPbFunction([&](const DImage& img, const DImage&) {
      auto it = maps.find(img.width());
      if (it == maps.end()) throw std::runtime_error("no boundary map for this level");
      return it->second;
    });
    OpticalFlow::schedule().relax_variant = std::atoi(argv[4]);
    rgbd::CameraParams cam{};
    cam.xRes = W;
    cam.yRes = H;
    cam.centerX = cx;
    cam.centerY = cy;
    cam.focalLength = f;
    const Mat_<Vec3f> flow = runFlowTwoFramesSceneFlowExperiments(
        cam, img1, d1, img2, d2, .02, 1, OpticalFlow::sceneFlowRegularizationType::HPB, std::atoi(argv[3]),
        std::nullopt, std::nullopt);
    writeSceneFlow(flow, argv[2]);
    const Mat_<Vec3f> back = readSceneFlow(argv[2]);
    for (size_t l = 0; l < flow.buf.size(); l++)
      for (int k = 0; k < 3; k++)
        if (back.buf[l][k] != flow.buf[l][k]) throw std::runtime_error(".flo3 round trip differs");
    // without an installed boundary function the maps are made on the device
    OpticalFlow::setHeuristicPbFunction(nullptr);
    const Mat_<Vec3f> dev = runFlowTwoFramesSceneFlowExperiments(
        cam, img1, d1, img2, d2, .02, 1, OpticalFlow::sceneFlowRegularizationType::HPB, std::atoi(argv[3]),
        std::nullopt, std::nullopt);
    writeSceneFlow(dev, std::string(argv[2]) + ".device.flo3");
    std::printf("ok %dx%d\n", W, H);
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "test_dropin: %s\n", e.what());
    return 1;
  }
}

// runFlowTwoFrames.h -- the two-frame scene-flow entry of
// include/rgbd_flow/runFlowTwoFrames.h:49-51, plus the conversions and the
// .flo3 format it is used with (src/ceLiuOpticalFlow.cpp:19-91,
// src/sceneFlowIO.cpp:18-61, src/main.cpp:32-48, src/rgbdFrameUtils.cpp:11-42).
#pragma once
#include <optional>
#include <string>

#include "OpticalFlow.h"

// runFlowTwoFrames.cpp:14-24
rgbd::CameraParams getResizedCamParams(const rgbd::CameraParams& initialParams, const rgbdflow::Size& sz);

// runFlowTwoFrames.cpp:66-100 (depth stdevs of both frames on the GPU)
void setUpSceneFlowParams(const rgbd::CameraParams& camParams, const rgbdflow::Mat_<rgbdflow::Vec3b>& prevImg,
                          const rgbdflow::Mat_<float>& prevDepth, const rgbdflow::Mat_<rgbdflow::Vec3b>& curImg,
                          const rgbdflow::Mat_<float>& curDepth, OpticalFlow::sceneFlowParams& params,
                          const double smoothnessWeight, const double depthVsColorDataWeight = 1,
                          const OpticalFlow::sceneFlowRegularizationType regularizationType =
                              OpticalFlow::sceneFlowRegularizationType::HPB,
                          const uint32_t maxIters = 10000);

// runFlowTwoFrames.cpp:109-121.  prevImg/curImg: 8-bit BGR; depths in metres
// (<= 0 invalid); result: per pixel (u, v in pixels, w in metres).
rgbdflow::Mat_<rgbdflow::Vec3f> runFlowTwoFramesSceneFlowExperiments(
    const rgbd::CameraParams& fullSizeCamParams, const rgbdflow::Mat_<rgbdflow::Vec3b>& prevImg,
    const rgbdflow::Mat_<float>& prevDepth, const rgbdflow::Mat_<rgbdflow::Vec3b>& curImg,
    const rgbdflow::Mat_<float>& curDepth, const double smoothnessWeight, const double depthVsColorDataWeight,
    const OpticalFlow::sceneFlowRegularizationType regularizationType, const uint32_t maxIters,
    const std::optional<rgbdflow::Mat_<rgbdflow::Vec3f>>& estFlow,
    const std::optional<std::string>& outdirPlusFilebase);

// ceLiuOpticalFlow.cpp:19-48, 64-91
DImage cv2liuBGR(const rgbdflow::Mat_<rgbdflow::Vec3b>& img);
DImage cv2liuDepth(const rgbdflow::Mat_<float>& dm, const bool fillIn);
DImage cv2liuFlow3d(const rgbdflow::Mat_<rgbdflow::Vec3f>& flow);
rgbdflow::Mat_<rgbdflow::Vec3f> liu2cvFlow3d(const DImage& vx, const DImage& vy, const DImage& vz);

// ceLiuOpticalFlow.cpp:190-215
rgbdflow::Mat_<rgbdflow::Vec3f> ceLiuSceneFlow(const rgbdflow::Mat_<rgbdflow::Vec3b>& img1,
                                               const rgbdflow::Mat_<float>& depth1,
                                               const rgbdflow::Mat_<rgbdflow::Vec3b>& img2,
                                               const rgbdflow::Mat_<float>& depth2,
                                               const std::optional<rgbdflow::Mat_<rgbdflow::Vec3f>>& estFlow,
                                               const OpticalFlow::sceneFlowParams& params,
                                               const std::optional<std::string>& outdirPlusFilebase);

// sceneFlowIO.cpp:18-61: 8-byte magic, int32 width, int32 height, then W*H*3 float32.
rgbdflow::Mat_<rgbdflow::Vec3f> readSceneFlow(const std::string& filepath);
void writeSceneFlow(const rgbdflow::Mat_<rgbdflow::Vec3f>& flow, const std::string& filepath);

// main.cpp:32-48: 16-bit depth values in mm -> metres
void readDepthMapValuesImg(const rgbdflow::Mat_<uint16_t>& img, rgbdflow::Mat_<float>& depth);
// rgbdFrameUtils.cpp:11-42 without SMOOTH_DEPTH: depth subsampled at even pixels; colour
// halved with a 2x2 box filter (cv::resize's INTER_LINEAR at exactly half size).
void downsizeFrame(rgbdflow::Mat_<rgbdflow::Vec3b>& img, rgbdflow::Mat_<float>& depth);

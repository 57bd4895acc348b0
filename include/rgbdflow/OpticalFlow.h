// OpticalFlow.h -- the drop-in C++ API of the scene-flow path, with the names,
// fields and argument meanings of include/celiu_flow/OpticalFlow.h (scene-flow
// part).  Implemented in librgbdflow_cxx.so, which reaches the GPU only through
// the C ABI of include/rgbdflow/rgbdflow.h.  Errors throw std::runtime_error,
// as the reference's ASSERT_ALWAYS does (include/rgbd_util/assert.h:57).
#pragma once
#include <functional>
#include <optional>
#include <string>
#include <vector>

#include "CameraParams.h"
#include "Image.h"
#include "mat.h"
#include "rgbdflow.h"

// OpticalFlow.h:17-22 (defined at src/sceneFlowSingleScaleJacobiCUDA.cu:660): one
// pyramid level.  Im1/Im2/warpIm2 are 5-channel feature images; u, v, w are
// in/out; warpIm2/warpDm2 are read (non-const in name only, as in the reference).
void sceneFlowSingleScale(const rgbd::CameraParams& camParams, const DImage& Im1, const DImage& dm1,
                          const DImage& Im2, const DImage& dm2, DImage& warpIm2, DImage& warpDm2, DImage& u,
                          DImage& v, DImage& w, const double colorDataTermWeight,
                          const DImage& depthDataTermWeights, const double alpha,
                          const DImage& depthSmoothnessTermWeights, const double gamma,
                          const DImage& depthMagTermWeights, const DImage& regularizationWeights,
                          int nOuterFPIterations, int nInnerFPIterations, int nSORIterations,
                          const bool allowEarlyTermination);

class OpticalFlow {
 public:
  static bool IsDisplay;

  // OpticalFlow.h:93-104
  enum class sceneFlowVariant { NONROBUST, LINEARIZED, FULL };
  enum class sceneFlowRegularizationType {
    ISOTROPIC,
    LETOUZEY,  // depth-dependent gaussian
    HPB        // "heuristic probability of boundary"
  };

  // OpticalFlow.h:105-129, same fields and defaults
  struct sceneFlowParams {
    sceneFlowParams()
        : variant(sceneFlowVariant::FULL), regularizationType(sceneFlowRegularizationType::HPB),
          camParams{}, colorDataWeight(1), depthVsColorDataWeight(1), alpha(0), gamma(0), pyramidRatio(.75),
          minWidth(20), nOuterFPIterations(7), nInnerFPIterations(1), nSORIterations(30) {}

    sceneFlowVariant variant;
    sceneFlowRegularizationType regularizationType;
    rgbd::CameraParams camParams;
    double colorDataWeight;
    double depthVsColorDataWeight;
    double alpha;  // smoothness weight
    double gamma;  // zero-flow prior weight
    rgbdflow::Mat_<float> depthMap;                  // for regularization; float
    rgbdflow::Mat_<float> depthSigma1, depthSigma2;  // depth stdev at each pixel
    double pyramidRatio;
    int minWidth;
    int nOuterFPIterations, nInnerFPIterations, nSORIterations;
  };

  // The boundary-probability map of one pyramid level ([H][W][2], right and
  // down neighbours) from that level's colour image (BGR in [0,1]) and raw
  // depth: the role of computeHeuristicPb (src/heuristicProbBoundary.cpp:21-41).
  // Without an installed function the device's restatement
  // (rgbdflow_heuristic_pb) is used; installing one (e.g. the reference's own)
  // replaces it.
  typedef std::function<std::vector<float>(const DImage& img, const DImage& depth)> HeuristicPbFunction;
  static void setHeuristicPbFunction(HeuristicPbFunction f);

  // Schedule of the GPU solve (relaxation variant, tiles); does not change results.
  static rgbdflow_params& schedule();

  // OpticalFlow.h:133 (src/sceneFlowMultiscale.cpp:30-201)
  static void sceneFlowMultiscale(const DImage& Im1, const DImage& dm1, const DImage& Im2, const DImage& dm2,
                                  const sceneFlowParams& params, DImage& u, DImage& v, DImage& w,
                                  const std::optional<std::string>& outdirPlusFilebase);

  // OpticalFlow.h:145 (src/sceneFlowWithInit.cpp:7-81)
  static void sceneFlowWithInit(const DImage& Im1, const DImage& dm1, const DImage& Im2, const DImage& dm2,
                                const sceneFlowParams& params, const DImage& estFlow, DImage& u, DImage& v,
                                DImage& w, const std::optional<std::string>& outdirPlusFilebase);

  // OpticalFlow.h:154-170 (src/OpticalFlow.cpp:71-141): host helpers that
  // build sceneFlowSingleScale's weight images.  hpb: [H][W][2].
  static void calculateDepthDataTermWeights(const sceneFlowParams& params, DImage& depthDataTermWeights);
  static void calculateDepthSmoothnessTermWeights(const sceneFlowParams& params, DImage& depthSmoothnessTermWeights);
  static void calculateDepthMagnitudeTermWeights(const DImage& depth, const sceneFlowParams& params,
                                                 DImage& depthMagnitudeTermWeights);
  static void calculateRegularizationWeights(const DImage& img, const DImage& depth, const sceneFlowParams& params,
                                             const std::vector<float>& hpb, DImage& regularizationWeights);

  // OpticalFlow.cpp:148-220 im2feature(RGB_WITH_GRADIENT), on the GPU.
  static void im2feature(DImage& imfeature, const DImage& im);
};

// src/recursiveMedianFiltering.cpp:117-120 and :127-222 (host; order-dependent greedy fills)
void recursiveMedianFilter(DImage& depth, const int32_t windowHalfwidth = 1);
void recursiveMedianFilterSmallInvalidRegionsUsingColor(DImage& depth, const DImage& col,
                                                        const uint32_t maxRegionSizeToFill);

// staticDepthNoiseModeling.h:40 (src/staticDepthNoiseModeling.cpp:225-258), on the GPU.
// stdevs: [rows][cols].
void computeDepthMapStdevsPointwise(const rgbdflow::Mat_<float>& depth, std::vector<float>& stdevs,
                                    const uint32_t increaseNearDepthBoundariesMaxDist, const bool multithread);

// warpBy3DFlow (src/opticalFlowUtils.cpp:76-99) for 5-channel feature images, on the GPU.
void warpBy3DFlow(const DImage& Im1, const DImage& dm1, const DImage& Im2, const DImage& dm2, DImage& warpIm2,
                  DImage& warpDm2, const DImage& vx, const DImage& vy, const DImage& vz);

// dropin.cpp -- the reference's scene-flow C++ entry points on top of the C ABI.
//
// Mirrors src/sceneFlowMultiscale.cpp, src/sceneFlowWithInit.cpp,
// src/runFlowTwoFrames.cpp, src/ceLiuOpticalFlow.cpp and src/sceneFlowIO.cpp:
// host-side conversions, the depth hole fills, and calls into librgbdflow for
// every image-sized computation.  One librgbdflow context per host thread.
#include <cmath>
#include <cstring>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>

#include "../../include/rgbdflow/OpticalFlow.h"
#include "../../include/rgbdflow/runFlowTwoFrames.h"

using rgbdflow::Mat_;
using rgbdflow::Vec3b;
using rgbdflow::Vec3f;

bool OpticalFlow::IsDisplay = false;

namespace {

struct CtxHolder {
  rgbdflow_ctx* ctx = nullptr;
  ~CtxHolder() {
    if (ctx) rgbdflow_destroy(ctx);
  }
};

rgbdflow_ctx* context() {
  thread_local CtxHolder h;
  if (!h.ctx) {
    const int rc = rgbdflow_create(&h.ctx, 0, 0);
    if (rc != RGBDFLOW_OK) throw std::runtime_error("rgbdflow_create failed (" + std::to_string(rc) + "): no HIP device?");
  }
  return h.ctx;
}

void check(int rc, const char* what) {
  if (rc != RGBDFLOW_OK)
    throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + rgbdflow_last_error(context()));
}

OpticalFlow::HeuristicPbFunction& hpb_function() {
  static OpticalFlow::HeuristicPbFunction f;
  return f;
}

rgbdflow_params& schedule_params() {
  static rgbdflow_params p = [] {
    rgbdflow_params q;
    rgbdflow_default_params(&q);
    q.relax_variant = RGBDFLOW_RELAX_TILED;
    return q;
  }();
  return p;
}

double sqr(double x) { return x * x; }

// primesensorUtils.cpp:53-56
double stereoError(double z) { return sqr(z) / (1 / .00285 + z); }

rgbdflow_params to_c(const OpticalFlow::sceneFlowParams& p) {
  rgbdflow_params c = schedule_params();
  c.variant = (int)p.variant;
  c.regularization_type = (int)p.regularizationType;
  c.cam.xRes = p.camParams.xRes;
  c.cam.yRes = p.camParams.yRes;
  c.cam.centerX = p.camParams.centerX;
  c.cam.centerY = p.camParams.centerY;
  c.cam.focalLength = p.camParams.focalLength;
  c.cam.stereoBaseline = p.camParams.stereoBaseline;
  c.cam.fovX = p.camParams.fovX;
  c.cam.fovY = p.camParams.fovY;
  c.color_data_weight = p.colorDataWeight;
  c.depth_vs_color_data_weight = p.depthVsColorDataWeight;
  c.alpha = p.alpha;
  c.gamma = p.gamma;
  c.pyramid_ratio = p.pyramidRatio;
  c.min_width = p.minWidth;
  c.n_outer = p.nOuterFPIterations;
  c.n_inner = p.nInnerFPIterations;
  c.n_sor = p.nSORIterations;
  return c;
}

// computeHeuristicPb of one image: the installed function, else the device's
// (rgbdflow_heuristic_pb).
std::vector<float> heuristic_pb(const DImage& img, const DImage& dep) {
  if (hpb_function()) return hpb_function()(img, dep);
  std::vector<float> m((size_t)img.width() * img.height() * 2);
  check(rgbdflow_heuristic_pb(context(), img.data(), dep.data(), img.width(), img.height(), m.data()),
        "rgbdflow_heuristic_pb");
  return m;
}

// The boundary maps of every level: computeHeuristicPb of each level of the
// colour pyramid and of the raw-depth pyramid, as sceneFlowMultiscale.cpp:78-83,
// 116 feeds it.
std::vector<float> boundary_maps(const DImage& Im1, const DImage& dm1, const OpticalFlow::sceneFlowParams& params) {
  const int W = Im1.width(), H = Im1.height();
  int ws[16], hs[16];
  const int L = rgbdflow_pyramid_levels(W, H, params.pyramidRatio, params.minWidth, ws, hs);
  if (L < 1) throw std::runtime_error("bad pyramid shape");
  size_t tot = 0;
  for (int k = 0; k < L; k++) tot += (size_t)ws[k] * hs[k];
  // DepthMapPyramid's level sizes are (int)(W r^k) (depthMapPyramid.cpp:15-30);
  // computeHeuristicPb pairs them with the colour levels, which must match
  const double r = (params.pyramidRatio > 0.98 || params.pyramidRatio < 0.4) ? 0.75 : params.pyramidRatio;
  for (int k = 0; k < L; k++)
    if ((int)(W * std::pow(r, k)) != ws[k] || (int)(H * std::pow(r, k)) != hs[k])
      throw std::runtime_error("HPB: the depth-map and colour pyramids' level sizes differ at this ratio");
  std::vector<double> colour(tot * 3);
  check(rgbdflow_gaussian_pyramid(context(), Im1.data(), W, H, 3, params.pyramidRatio, params.minWidth,
                                  colour.data()),
        "rgbdflow_gaussian_pyramid");
  DImage filled = dm1;
  recursiveMedianFilterSmallInvalidRegionsUsingColor(filled, Im1, 60);
  std::vector<double> raw(tot);
  check(rgbdflow_depth_map_pyramid(context(), filled.data(), W, H, params.pyramidRatio, params.minWidth, raw.data()),
        "rgbdflow_depth_map_pyramid");
  std::vector<float> out;
  out.reserve(tot * 2);
  size_t off = 0;
  for (int k = 0; k < L; k++) {
    DImage img(ws[k], hs[k], 3), dep(ws[k], hs[k], 1);
    std::memcpy(img.data(), colour.data() + off * 3, sizeof(double) * ws[k] * hs[k] * 3);
    std::memcpy(dep.data(), raw.data() + off, sizeof(double) * ws[k] * hs[k]);
    const std::vector<float> m = heuristic_pb(img, dep);
    if (m.size() != (size_t)ws[k] * hs[k] * 2) throw std::runtime_error("HPB function returned a map of the wrong size");
    out.insert(out.end(), m.begin(), m.end());
    off += (size_t)ws[k] * hs[k];
  }
  return out;
}

std::vector<float> sigma_of(const OpticalFlow::sceneFlowParams& params, int W, int H) {
  const Mat_<float>& s = params.depthSigma1;
  if (s.rows != H || s.cols != W) throw std::runtime_error("params.depthSigma1 must be set (setUpSceneFlowParams)");
  for (float x : s.buf)
    if (!(x > 0)) throw std::runtime_error("params.depthSigma1 must be positive");
  return s.buf;
}

}  // namespace

void OpticalFlow::setHeuristicPbFunction(HeuristicPbFunction f) { hpb_function() = std::move(f); }
rgbdflow_params& OpticalFlow::schedule() { return schedule_params(); }

// src/sceneFlowSingleScaleJacobiCUDA.cu:660-969
void sceneFlowSingleScale(const rgbd::CameraParams& camParams, const DImage& Im1, const DImage& dm1, const DImage& Im2,
                          const DImage& dm2, DImage& warpIm2, DImage& warpDm2, DImage& u, DImage& v, DImage& w,
                          const double colorDataTermWeight, const DImage& depthDataTermWeights, const double alpha,
                          const DImage& depthSmoothnessTermWeights, const double gamma,
                          const DImage& depthMagTermWeights, const DImage& regularizationWeights,
                          int nOuterFPIterations, int nInnerFPIterations, int nSORIterations,
                          const bool allowEarlyTermination) {
  const int W = Im1.width(), H = Im1.height();
  if (Im1.nchannels() != 5 || Im2.nchannels() != 5 || warpIm2.nchannels() != 5)
    throw std::runtime_error("sceneFlowSingleScale expects 5-channel feature images");
  if (u.width() != W || v.width() != W || w.width() != W || regularizationWeights.nchannels() != 2)
    throw std::runtime_error("sceneFlowSingleScale: argument sizes differ");
  rgbdflow_params opts = schedule_params();
  rgbdflow_stats st;
  check(rgbdflow_solve_level(context(), W, H, camParams.focalLength, Im1.data(), dm1.data(), Im2.data(), dm2.data(),
                             warpIm2.data(), warpDm2.data(), u.data(), v.data(), w.data(), colorDataTermWeight,
                             depthDataTermWeights.data(), alpha, depthSmoothnessTermWeights.data(), gamma,
                             depthMagTermWeights.data(), regularizationWeights.data(), nOuterFPIterations,
                             nInnerFPIterations, nSORIterations, allowEarlyTermination ? 1 : 0, &opts, &st),
        "rgbdflow_solve_level");
}

// src/sceneFlowMultiscale.cpp:30-201
void OpticalFlow::sceneFlowMultiscale(const DImage& Im1, const DImage& dm1, const DImage& Im2, const DImage& dm2,
                                      const sceneFlowParams& params, DImage& u, DImage& v, DImage& w,
                                      const std::optional<std::string>&) {
  if (params.variant != sceneFlowVariant::FULL) throw std::runtime_error("bad scene flow variant");
  const int W = Im1.width(), H = Im1.height();
  if (Im1.nchannels() != 3 || Im2.nchannels() != 3) throw std::runtime_error("sceneFlowMultiscale expects BGR images");
  DImage depth1 = dm1, depth2 = dm2;
  recursiveMedianFilter(depth1);
  recursiveMedianFilter(depth2);
  std::vector<float> hpb;
  if (params.regularizationType == sceneFlowRegularizationType::HPB) hpb = boundary_maps(Im1, dm1, params);
  const std::vector<float> sigma = sigma_of(params, W, H);
  const rgbdflow_params p = to_c(params);
  u.allocate(W, H);
  v.allocate(W, H);
  w.allocate(W, H);
  rgbdflow_stats st;
  check(rgbdflow_multiscale(context(), W, H, Im1.data(), depth1.data(), Im2.data(), depth2.data(), sigma.data(),
                            hpb.empty() ? nullptr : hpb.data(), &p, u.data(), v.data(), w.data(), &st),
        "rgbdflow_multiscale");
}

// src/sceneFlowWithInit.cpp:7-81
void OpticalFlow::sceneFlowWithInit(const DImage& Im1, const DImage& dm1, const DImage& Im2, const DImage& dm2,
                                    const sceneFlowParams& params, const DImage& estFlow, DImage& u, DImage& v,
                                    DImage& w, const std::optional<std::string>&) {
  if (estFlow.nchannels() != 3) throw std::runtime_error("estFlow must have 3 channels");
  const int W = Im1.width(), H = Im1.height();
  DImage depth1 = dm1, depth2 = dm2;
  recursiveMedianFilter(depth1);
  recursiveMedianFilter(depth2);
  std::vector<float> hpb0;
  if (params.regularizationType == sceneFlowRegularizationType::HPB) hpb0 = heuristic_pb(Im1, dm1);
  const std::vector<float> sigma = sigma_of(params, W, H);
  const rgbdflow_params p = to_c(params);
  u.allocate(W, H);
  v.allocate(W, H);
  w.allocate(W, H);
  rgbdflow_stats st;
  check(rgbdflow_with_init(context(), W, H, Im1.data(), depth1.data(), Im2.data(), depth2.data(), sigma.data(),
                           hpb0.empty() ? nullptr : hpb0.data(), estFlow.data(), &p, u.data(), v.data(), w.data(),
                           &st),
        "rgbdflow_with_init");
}

// src/OpticalFlow.cpp:71-81
void OpticalFlow::calculateDepthDataTermWeights(const sceneFlowParams& params, DImage& out) {
  out.allocate(params.camParams.xRes, params.camParams.yRes, 1);
  double* d = out.data();
  for (uint32_t i = 0; i < params.camParams.yRes; i++)
    for (uint32_t j = 0; j < params.camParams.xRes; j++) {
      if (!(params.depthSigma1(i, j) > 0)) throw std::runtime_error("depthSigma1 must be positive");
      *d++ = stereoError(1) / params.depthSigma1(i, j) * params.depthVsColorDataWeight;
    }
}

// src/OpticalFlow.cpp:90-97
void OpticalFlow::calculateDepthSmoothnessTermWeights(const sceneFlowParams& params, DImage& out) {
  out.allocate(params.camParams.xRes, params.camParams.yRes, 1);
  for (int i = 0; i < out.nelements(); i++) out[i] = sqr(params.camParams.focalLength);
}

// src/OpticalFlow.cpp:104-111.  `depth` of a pyramid level evaluated under the
// full-size camera (as sceneFlowMultiscale.cpp:144 calls it): level pixel l
// reads depth[(l / xRes) * width + l % xRes]; only the level's first W*H
// entries are computed (the reference computes xRes*yRes, reading past the
// level image; the solver uses the first W*H, .cu:686).
void OpticalFlow::calculateDepthMagnitudeTermWeights(const DImage& depth, const sceneFlowParams& params, DImage& out) {
  const int W = depth.width(), H = depth.height();
  out.allocate(W, H);
  const size_t N = (size_t)W * H, Wf = params.camParams.xRes;
  for (size_t l = 0; l < N; l++) {
    const size_t idx = (l / Wf) * W + l % Wf;
    const double d = idx < N ? depth[idx] : 0.0;
    out[l] = d > 0 ? sqr(params.camParams.focalLength / d) : sqr(params.camParams.focalLength / 2);
  }
}

// src/OpticalFlow.cpp:113-141
void OpticalFlow::calculateRegularizationWeights(const DImage& img, const DImage& depth, const sceneFlowParams& params,
                                                 const std::vector<float>& hpb, DImage& out) {
  const int W = img.width(), H = img.height();
  out.allocate(W, H, 2);
  if (params.regularizationType == sceneFlowRegularizationType::HPB && hpb.size() != (size_t)W * H * 2)
    throw std::runtime_error("calculateRegularizationWeights: hpb must be [H][W][2]");
  auto weight = [&](int l, int l2, float h) -> double {
    switch (params.regularizationType) {
      case sceneFlowRegularizationType::ISOTROPIC: return 1;
      case sceneFlowRegularizationType::LETOUZEY: return std::exp(-sqr((depth[l] - depth[l2]) / .02));
      case sceneFlowRegularizationType::HPB: return 1 - h * .99;
    }
    throw std::runtime_error("bad regularization type");
  };
  for (int i = 0, l = 0; i < H; i++)
    for (int j = 0; j < W; j++, l++) {
      if (j < W - 1) out(i, j, 0) = weight(l, l + 1, hpb.empty() ? 0.f : hpb[(size_t)l * 2 + 0]);
      if (i < H - 1) out(i, j, 1) = weight(l, l + W, hpb.empty() ? 0.f : hpb[(size_t)l * 2 + 1]);
    }
}

void OpticalFlow::im2feature(DImage& imfeature, const DImage& im) {
  if (im.nchannels() != 3) throw std::runtime_error("im2feature: RGB_WITH_GRADIENT needs 3 channels");
  imfeature.allocate(im.width(), im.height(), 5);
  check(rgbdflow_im2feature(context(), im.data(), im.width(), im.height(), imfeature.data()), "rgbdflow_im2feature");
}

void computeDepthMapStdevsPointwise(const Mat_<float>& depth, std::vector<float>& stdevs,
                                    const uint32_t increaseNearDepthBoundariesMaxDist, const bool) {
  if (depth.rows <= 0 || depth.cols <= 0) throw std::runtime_error("computeDepthMapStdevsPointwise: empty depth");
  stdevs.resize((size_t)depth.rows * depth.cols);
  check(rgbdflow_depth_stdevs_pointwise(context(), depth.data(), depth.cols, depth.rows,
                                        (int)increaseNearDepthBoundariesMaxDist, stdevs.data()),
        "rgbdflow_depth_stdevs_pointwise");
}

void warpBy3DFlow(const DImage& Im1, const DImage& dm1, const DImage& Im2, const DImage& dm2, DImage& warpIm2,
                  DImage& warpDm2, const DImage& vx, const DImage& vy, const DImage& vz) {
  warpIm2.allocate(Im2);
  warpDm2.allocate(dm2);
  check(rgbdflow_warp_by_3d_flow(context(), Im1.data(), dm1.data(), Im2.data(), dm2.data(), vx.data(), vy.data(),
                                 vz.data(), Im1.width(), Im1.height(), Im1.nchannels(), warpIm2.data(),
                                 warpDm2.data()),
        "rgbdflow_warp_by_3d_flow");
}

// ---------------------------------------------------------------- two frames
rgbd::CameraParams getResizedCamParams(const rgbd::CameraParams& initialParams, const rgbdflow::Size& sz) {
  rgbd::CameraParams camParams = initialParams;
  const double sizeFactor = (double)sz.width / initialParams.xRes;
  camParams.xRes = std::rint(camParams.xRes * sizeFactor);
  camParams.yRes = std::rint(camParams.yRes * sizeFactor);
  camParams.centerX = std::rint(camParams.centerX * sizeFactor);
  camParams.centerY = std::rint(camParams.centerY * sizeFactor);
  camParams.focalLength *= sizeFactor;
  return camParams;
}

void setUpSceneFlowParams(const rgbd::CameraParams& camParams, const Mat_<Vec3b>&, const Mat_<float>& prevDepth,
                          const Mat_<Vec3b>&, const Mat_<float>& curDepth, OpticalFlow::sceneFlowParams& params,
                          const double smoothnessWeight, const double depthVsColorDataWeight,
                          const OpticalFlow::sceneFlowRegularizationType regularizationType, const uint32_t maxIters) {
  params.regularizationType = regularizationType;
  params.camParams = camParams;
  params.minWidth = 60;
  params.colorDataWeight = 1;
  params.depthVsColorDataWeight = depthVsColorDataWeight;
  params.alpha = smoothnessWeight;
  params.gamma = 1e-5;
  params.nOuterFPIterations = 50;
  params.nInnerFPIterations = 1;
  params.nSORIterations = maxIters;
  params.depthMap = prevDepth;
  std::vector<float> s1, s2;
  computeDepthMapStdevsPointwise(prevDepth, s1, 4, true);
  computeDepthMapStdevsPointwise(curDepth, s2, 4, true);
  params.depthSigma1.create(prevDepth.rows, prevDepth.cols);
  params.depthSigma2.create(prevDepth.rows, prevDepth.cols);
  params.depthSigma1.buf = s1;
  params.depthSigma2.buf = s2;
}

DImage cv2liuBGR(const Mat_<Vec3b>& img) {
  DImage im(img.cols, img.rows, 3);
  double* d = im.data();
  for (int i = 0; i < img.rows; i++)
    for (int j = 0; j < img.cols; j++)
      for (int k = 0; k < 3; k++) *d++ = img(i, j)[k] / 255.0;
  return im;
}

DImage cv2liuDepth(const Mat_<float>& dm, const bool fillIn) {
  DImage dim(dm.cols, dm.rows, 1);
  for (size_t l = 0; l < dm.buf.size(); l++) dim[l] = dm.buf[l];
  if (fillIn) recursiveMedianFilter(dim);
  return dim;
}

DImage cv2liuFlow3d(const Mat_<Vec3f>& flow) {
  DImage f(flow.cols, flow.rows, 3);
  for (size_t l = 0; l < flow.buf.size(); l++)
    for (int k = 0; k < 3; k++) f[l * 3 + k] = flow.buf[l][k];
  return f;
}

Mat_<Vec3f> liu2cvFlow3d(const DImage& vx, const DImage& vy, const DImage& vz) {
  Mat_<Vec3f> m(vx.height(), vx.width());
  for (size_t l = 0; l < m.buf.size(); l++) {
    m.buf[l][0] = vx[l];
    m.buf[l][1] = vy[l];
    m.buf[l][2] = vz[l];
  }
  return m;
}

Mat_<Vec3f> ceLiuSceneFlow(const Mat_<Vec3b>& img1, const Mat_<float>& depth1, const Mat_<Vec3b>& img2,
                           const Mat_<float>& depth2, const std::optional<Mat_<Vec3f>>& estFlow,
                           const OpticalFlow::sceneFlowParams& params, const std::optional<std::string>& outdir) {
  DImage Im1 = cv2liuBGR(img1), Im2 = cv2liuBGR(img2);
  DImage dm1 = cv2liuDepth(depth1, false), dm2 = cv2liuDepth(depth2, false);
  DImage vx, vy, vz;
  if (estFlow) OpticalFlow::sceneFlowWithInit(Im1, dm1, Im2, dm2, params, cv2liuFlow3d(*estFlow), vx, vy, vz, outdir);
  else OpticalFlow::sceneFlowMultiscale(Im1, dm1, Im2, dm2, params, vx, vy, vz, outdir);
  return liu2cvFlow3d(vx, vy, vz);
}

Mat_<Vec3f> runFlowTwoFramesSceneFlowExperiments(const rgbd::CameraParams& fullSizeCamParams,
                                                 const Mat_<Vec3b>& prevImg, const Mat_<float>& prevDepth,
                                                 const Mat_<Vec3b>& curImg, const Mat_<float>& curDepth,
                                                 const double smoothnessWeight, const double depthVsColorDataWeight,
                                                 const OpticalFlow::sceneFlowRegularizationType regularizationType,
                                                 const uint32_t maxIters, const std::optional<Mat_<Vec3f>>& estFlow,
                                                 const std::optional<std::string>& outdir) {
  const rgbd::CameraParams camParams = getResizedCamParams(fullSizeCamParams, prevImg.size());
  OpticalFlow::sceneFlowParams params;
  setUpSceneFlowParams(camParams, prevImg, prevDepth, curImg, curDepth, params, smoothnessWeight,
                       depthVsColorDataWeight, regularizationType, maxIters);
  return ceLiuSceneFlow(prevImg, prevDepth, curImg, curDepth, estFlow, params, outdir);
}

// ---------------------------------------------------------------- .flo3
namespace {
const uint8_t kMagic[8] = {237, 159, 43, 26, 202, 94, 218, 68};
}

Mat_<Vec3f> readSceneFlow(const std::string& filepath) {
  std::ifstream in(filepath, std::ios::binary);
  if (!in) throw std::runtime_error("can't open " + filepath);
  uint8_t magic[8];
  int32_t w, h;
  if (!in.read((char*)magic, 8) || !in.read((char*)&w, 4) || !in.read((char*)&h, 4))
    throw std::runtime_error("can't read metadata for flow image");
  if (std::memcmp(magic, kMagic, 8) != 0) throw std::invalid_argument("magic number doesn't match .flo3 format");
  if (w < 0 || h < 0) throw std::runtime_error("bad .flo3 size");
  Mat_<Vec3f> img(h, w);
  for (auto& p : img.buf)
    if (!in.read((char*)p.val, 12)) throw std::runtime_error("can't read pixel values for flow image");
  return img;
}

void writeSceneFlow(const Mat_<Vec3f>& flow, const std::string& filepath) {
  std::ofstream out(filepath, std::ios::binary);
  if (!out) throw std::runtime_error("can't open " + filepath);
  const int32_t w = flow.cols, h = flow.rows;
  if (!out.write((const char*)kMagic, 8) || !out.write((const char*)&w, 4) || !out.write((const char*)&h, 4))
    throw std::runtime_error("can't write metadata for flow image");
  for (const auto& p : flow.buf)
    if (!out.write((const char*)p.val, 12)) throw std::runtime_error("can't write pixel values for flow image");
}

void readDepthMapValuesImg(const Mat_<uint16_t>& img, Mat_<float>& depth) {
  depth.create(img.rows, img.cols);
  for (size_t l = 0; l < img.buf.size(); l++) depth.buf[l] = img.buf[l] * .001;
}

void downsizeFrame(Mat_<Vec3b>& img, Mat_<float>& depth) {
  Mat_<Vec3b> small(img.rows / 2, img.cols / 2);
  for (int i = 0; i < small.rows; i++)
    for (int j = 0; j < small.cols; j++)
      for (int k = 0; k < 3; k++) {
        const int s = img(2 * i, 2 * j)[k] + img(2 * i, 2 * j + 1)[k] + img(2 * i + 1, 2 * j)[k] +
                      img(2 * i + 1, 2 * j + 1)[k];
        small(i, j)[k] = (uint8_t)((s + 2) >> 2);
      }
  Mat_<float> dsmall(depth.rows / 2, depth.cols / 2);
  for (int i = 0; i < dsmall.rows; i++)
    for (int j = 0; j < dsmall.cols; j++) dsmall(i, j) = depth(i * 2, j * 2);
  img = small;
  depth = dsmall;
}

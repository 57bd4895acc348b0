// CameraParams.h -- rgbd::CameraParams (include/rgbd_util/CameraParams.h:17-28),
// field for field.
#pragma once

namespace rgbd {

struct CameraParams {
  unsigned int xRes;     // pixels
  unsigned int yRes;     // pixels
  float centerX;         // pixels
  float centerY;         // pixels
  float focalLength;     // pixels
  float stereoBaseline;  // metres (depth camera: IR projector to IR camera)
  float fovX, fovY;      // radians
};

}  // namespace rgbd

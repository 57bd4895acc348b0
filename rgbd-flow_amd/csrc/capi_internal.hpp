// capi_internal.hpp -- entry points of capi.cpp for the serving runtime
// (server.cpp): batches whose frames are already resident on the device.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "../../include/rgbdflow/rgbdflow.h"

namespace rgbdflow {

// A frame resident on the device.
struct FrameDev {
  const double* im;   // [3][N] BGR planes in [0, 1]
  const double* dm;   // [N] depth, metres
  const double* pim;  // colour pyramid levels >= 1 ([3][N_k] per level, finest first), or null
  const double* pd;   // depth pyramid levels >= 1, or null
  const float* hpb;   // boundary maps of every level ([N_k][2], finest first), or null
  const float* sigma; // [N] depth stdevs as frame 1 (from the depth before any hole fill), or null:
                      // computed from dm
  const double* hpb_dm; // [N] depth the boundary maps made on the device start from, or null: dm
};

// One pair whose frames are on the device.  Frame 1's pyramids, when given,
// are used as they are (they are the pyramids frame 1 had as a frame 2).
// init: a flow estimate ([3][N] planes) -- the pair is then solved as
// sceneFlowWithInit (finest level only), else null.
struct PairDev {
  int W, H, min_width;
  rgbdflow_camera cam;
  FrameDev f1, f2;
  const double* init;
};

int batch_load_device(rgbdflow_ctx* c, const rgbdflow_params* p, int n, const PairDev* pairs);
// After a run: pair i's frame 2 with its pyramids and its flow ([3][N] planes).
int pair_outputs(rgbdflow_ctx* c, int i, FrameDev* f2, const double** flow, size_t* pyr_elems);
hipStream_t ctx_stream(rgbdflow_ctx* c);
size_t pyramid_rest_elems(int W, int H, double ratio, int min_width);
size_t pyramid_all_elems(int W, int H, double ratio, int min_width);

}  // namespace rgbdflow

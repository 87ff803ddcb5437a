// Handle types shared by the host sources of libamcl3d_gpu.
#pragma once

#include "a3d_internal.h"

struct a3d_grid {
  a3d::GridDev g{};
  float* prob = nullptr;
  float* dist = nullptr;
  double sensor_dev = 0.0;
  hipStream_t stream = nullptr;
};

// A captured chain of _dev calls, replayed on the stream it was captured on.
struct a3d_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t stream = nullptr;
};

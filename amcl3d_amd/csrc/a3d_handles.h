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

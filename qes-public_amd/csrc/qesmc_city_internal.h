// Pieces of the flat-city input builder shared by the host builder
// (qesmc_city.cpp) and the device builder (qesmc_city_gpu.hip): the vertical
// grid, one rectangle's polygon and cell bounds, the cell-centre
// crossing-number test and the one-dimensional sensor profile.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "qesmc.h"

namespace qesmc {

// One rectangle as PolyBuilding::setCellFlags sees it (PolyBuilding.cpp:130-236):
// a closed ring of 5 vertices, the (i, j) search box and the k range [k0, k1).
struct Footprint {
  float xv[5], yv[5];
  int i0, i1, j0, j1, k0, k1;
};

// Domain::defineVerticalGrid (Domain.cpp:146-160): z_face (nz) and z (nz-1).
void vertical_grid(int nz, const float *dz, std::vector<float> &z_face, std::vector<float> &z);

// RectangularBuilding's polygon (RectangularBuilding.h:86-103) and the
// bounds PolyBuilding::setCellFlags loops over.
Footprint building_footprint(int nz, float dx, const std::vector<float> &z_face, const std::vector<float> &z,
                             const qesmc_building &b);

// Single-sensor profile on flat ground (WindProfilerSensorType.cpp:290-333,
// 571-621): up[k], vp[k] for k in [0, nz).  Returns QESMC_ERR_ARG for an
// unknown profile kind.
int inflow_profile(int nz, const std::vector<float> &z, const qesmc_inflow &in, std::vector<float> &up,
                   std::vector<float> &vp);

// Cell-centre crossing-number test of PolyBuilding::setCellFlags (the
// vertex loop restarts after a ring closes).  Same float operations on the
// host and the device; translation units are built without FMA contraction.
__host__ __device__ inline bool footprint_contains(const Footprint &F, float xc, float yc)
{
  int crossings = 0;
  unsigned v = 0, start = 0;
  while (v < 4) {
    const bool straddles = (F.yv[v] <= yc && F.yv[v + 1] > yc) || (F.yv[v] > yc && F.yv[v + 1] <= yc);
    if (straddles) {
      const float t = (yc - F.yv[v]) / (F.yv[v + 1] - F.yv[v]);
      if (xc < (F.xv[v] + t * (F.xv[v + 1] - F.xv[v]))) ++crossings;
    }
    ++v;
    if (F.xv[v] == F.xv[start] && F.yv[v] == F.yv[start]) {
      ++v;
      start = v;
    }
  }
  return (crossings & 1) != 0;
}

__host__ __device__ inline bool is_solid_flag(int fl) { return fl == 0 || fl == 2; }

}  // namespace qesmc

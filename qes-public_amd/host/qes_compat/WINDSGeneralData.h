/*
 * Minimal re-declaration of WINDSGeneralData for the library's own tests:
 * only the arrays Solver::solve reads and writes (src/winds/WINDSGeneralData.h
 * :231-274), allocated as WINDSGeneralData::allocateMemory does
 * (WINDSGeneralData.cpp:1064-1120).  Not the reference's header.
 */
#pragma once

#include <vector>

#include "qes/Domain.h"

class WINDSGeneralData
{
public:
  explicit WINDSGeneralData(const qes::Domain &d)
    : domain(d),
      e(d.numCellCentered(), 1.0f), f(d.numCellCentered(), 1.0f), g(d.numCellCentered(), 1.0f),
      h(d.numCellCentered(), 1.0f), m(d.numCellCentered(), 1.0f), n(d.numCellCentered(), 1.0f),
      icellflag(d.numCellCentered(), 1),
      u0(d.numFaceCentered(), 0.0f), v0(d.numFaceCentered(), 0.0f), w0(d.numFaceCentered(), 0.0f),
      u(d.numFaceCentered(), 0.0f), v(d.numFaceCentered(), 0.0f), w(d.numFaceCentered(), 0.0f)
  {}

  qes::Domain domain;
  std::vector<float> e, f, g, h, m, n;  /**< solver coefficients */
  std::vector<int> icellflag;           /**< 0 building, 1 fluid, 2 terrain */
  std::vector<float> u0, v0, w0;        /**< initial wind (faces) */
  std::vector<float> u, v, w;           /**< final wind (faces) */
  float z0{};                           /**< roughness used by wallLogBC (WINDSGeneralData.h:208) */
};

/*
 * Minimal SolverFactory (src/winds/SolverFactory.h:48-57, SolverFactory.cpp:
 * 35-60) showing the registration of the HIP solver as solveType 5.  Only the
 * HIP solver exists in this re-declaration; the reference's factory keeps its
 * CPU and CUDA branches and gains the HIP one (see the integration notes).
 */
#pragma once

#include "Solver_HIP.h"
#include "qes/Domain.h"
#include "util/QESout.h"

enum solverTypes : int { CPU_Type = 1,
                         DYNAMIC_P = 2,
                         Global_M = 3,
                         Shared_M = 4,
                         HIP_MI355X = 5 };

class SolverFactory
{
public:
  static Solver *create(const int &solveType, const qes::Domain &domain, const float &tolerance)
  {
    Solver *solver = nullptr;
    if (solveType == HIP_MI355X) {
      solver = new Solver_HIP(domain, tolerance);
    } else {
      QESout::error("Invalid solver type");
    }
    return solver;
  }
};

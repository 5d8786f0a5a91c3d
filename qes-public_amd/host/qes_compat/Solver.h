/*
 * Minimal re-declaration of the QES-Winds Solver base class
 * (src/winds/Solver.h:61-120, Solver.cpp:77-91) with the same members,
 * constants and resetLambda semantics, for the library's own tests of
 * Solver_HIP.  Not the reference's header.
 */
#pragma once

#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "WINDSGeneralData.h"
#include "qes/Domain.h"

class Solver
{
protected:
  qes::Domain domain;

  const int alpha1; /**< Gaussian precision moduli */
  const int alpha2;
  const float eta;
  const float A;
  const float B;

  float tol;                 /**< Error tolerance */
  const float omega = 1.78f; /**< Over-relaxation factor */

  std::vector<float> R;
  std::vector<float> lambda, lambda_old;

  Solver(qes::Domain domain_in, const float &tolerance)
    : domain(std::move(domain_in)),
      alpha1(1),
      alpha2(1),
      eta(std::pow((alpha1 / alpha2), 2.0)),
      A(std::pow((domain.dx() / domain.dy()), 2.0)),
      B(eta * std::pow((domain.dx() / domain.dz()), 2.0))
  {
    tol = tolerance;
    lambda.resize(domain.numCellCentered(), 0.0);
    lambda_old.resize(domain.numCellCentered(), 0.0);
    R.resize(domain.numCellCentered(), 0.0);
  }

public:
  virtual ~Solver() = default;

  void resetLambda()
  {
    std::fill(lambda.begin(), lambda.end(), 0.0);
    std::fill(lambda_old.begin(), lambda_old.end(), 0.0);
    std::fill(R.begin(), R.end(), 0.0);
  }

  /** Access for tests (the reference keeps lambda protected). */
  const std::vector<float> &getLambda() const { return lambda; }

  virtual void solve(WINDSGeneralData *, const int &) = 0;
};

// Drives Solver_HIP the way qesWinds does (qesWinds/qesWindsMain.cpp:140-170):
// a solver from SolverFactory by solveType, then per time step
// resetLambda (or not) + solve(WGD, itermax), and checks u, v, w and lambda
// bit for bit against the CPU oracle (test infrastructure, linked here only
// as the reference).  Prints PASS and returns 0 on success.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "SolverFactory.h"
#include "Solver_HIP.h"
#include "WINDSGeneralData.h"
#include "qesmc.h"

extern "C" int oracle_solve_rb_f32(int nx, int ny, int nz, float dx, float dy, const float *dz_array,
                                   const float *e, const float *f, const float *g, const float *h,
                                   const float *m, const float *n, const int *icellflag, const float *u0,
                                   const float *v0, const float *w0, float *R, float *lambda,
                                   float *lambda_old, float omega, float tol, int itermax, int alpha1,
                                   int alpha2, float *u, float *v, float *w, float *max_error_out);

extern "C" void oracle_divergence_f64(int nx, int ny, int nz, double dx, double dy, const float *dz_array,
                                      const float *e, const float *f, const float *g, const float *h,
                                      const float *m, const float *n, const float *u0, const float *v0,
                                      const float *w0, int alpha1, double *R);
extern "C" int oracle_sor_rb_f64(int nx, int ny, int nz, const float *e, const float *f, const float *g,
                                 const float *h, const float *m, const float *n, const double *R, double *lambda,
                                 double *lambda_old, double omega, double tol, int itermax, double *max_error_out);
extern "C" void oracle_velocity_f64(int nx, int ny, int nz, double dx, double dy, const float *dz_array,
                                    const float *f, const float *h, const float *n, const int *icellflag,
                                    const float *u0, const float *v0, const float *w0, const double *lambda,
                                    int alpha1, int alpha2, double *u, double *v, double *w);

extern "C" long oracle_wall_log_bc(int nx, int ny, int nz, float dx, float dz, float z0, const int *icellflag,
                                   float *u, float *v, float *w);

static int failures = 0;

// Device and host math libraries may differ by a few ulp in atan2f / cosf /
// sinf, which wallLogBC evaluates per wall face.
static void expect_close(const std::vector<float> &a, const std::vector<float> &b, const char *what)
{
  size_t bad = 0;
  for (size_t q = 0; q < a.size(); ++q) {
    const double d = std::fabs((double)a[q] - (double)b[q]);
    if (!(d <= 2e-6 * (1.0 + std::fabs((double)b[q])))) ++bad;
  }
  if (bad) {
    std::printf("MISMATCH %s: %zu of %zu entries beyond 2e-6 relative\n", what, bad, a.size());
    ++failures;
  }
}

static void expect_equal(const std::vector<float> &a, const std::vector<float> &b, const char *what)
{
  size_t bad = 0;
  for (size_t q = 0; q < a.size(); ++q)
    if (!(a[q] == b[q]) && !(std::isnan(a[q]) && std::isnan(b[q]))) ++bad;
  if (bad) {
    std::printf("MISMATCH %s: %zu of %zu entries differ\n", what, bad, a.size());
    ++failures;
  }
}

static void set_inflow(WINDSGeneralData &wgd, const qes::Domain &dom, float wind_dir,
                       const std::vector<qesmc_building> &bl)
{
  qesmc_inflow inflow{1, 3.8f, 10.0f, 0.5f, wind_dir};
  const int rc = qesmc_city_build(dom.nx(), dom.ny(), dom.nz(), dom.dx(), dom.dz_array.data(), (int)bl.size(),
                                  bl.data(), &inflow, wgd.e.data(), wgd.f.data(), wgd.g.data(), wgd.h.data(),
                                  wgd.m.data(), wgd.n.data(), wgd.icellflag.data(), wgd.u0.data(),
                                  wgd.v0.data(), wgd.w0.data());
  if (rc != QESMC_OK) {
    std::printf("city build failed: %s\n", qesmc_status_string(rc));
    std::exit(2);
  }
}

int main()
{
  // 90 x 70 x 24 cells, dx = dy = 2 m, stretched dz (1 m at ratio 1.03)
  qes::Domain dom(90, 70, 24, 2.0f, 2.0f, 1.0f);
  std::vector<float> dzc(24);
  for (int k = 0; k < 24; ++k) dzc[k] = (float)std::pow(1.03, k);
  dom.setVerticalStretching(dzc);
  std::vector<qesmc_building> bl(12);
  const int nb = qesmc_city_random_buildings(99, dom.nx(), dom.ny(), dom.dx(), 12, 6.0f, 24.0f, 4.0f, 20.0f,
                                             0, 0.0f, 2, bl.data());
  bl.resize(nb);

  WINDSGeneralData wgd(dom);
  std::unique_ptr<Solver> solver(SolverFactory::create(HIP_MI355X, dom, 1e-9f));
  auto *hip = dynamic_cast<Solver_HIP *>(solver.get());

  const long nc = dom.numCellCentered(), nf = dom.numFaceCentered();
  std::vector<float> R(nc), lam(nc, 0.f), lam_old(nc), u(nf), v(nf), w(nf);
  float me = 0.f;
  const float dirs[3] = {270.0f, 247.5f, 300.0f};
  const bool reset[3] = {true, false, true};
  const int iters[3] = {300, 120, 0};
  for (int step = 0; step < 3; ++step) {
    set_inflow(wgd, dom, dirs[step], bl);
    if (reset[step]) {
      solver->resetLambda();
      std::fill(lam.begin(), lam.end(), 0.f);
    }
    solver->solve(&wgd, iters[step]);
    const int it = oracle_solve_rb_f32(dom.nx(), dom.ny(), dom.nz(), dom.dx(), dom.dy(), dom.dz_array.data(),
                                       wgd.e.data(), wgd.f.data(), wgd.g.data(), wgd.h.data(), wgd.m.data(),
                                       wgd.n.data(), wgd.icellflag.data(), wgd.u0.data(), wgd.v0.data(),
                                       wgd.w0.data(), R.data(), lam.data(), lam_old.data(), 1.78f, 1e-9f,
                                       iters[step], 1, 1, u.data(), v.data(), w.data(), &me);
    if (hip->lastResult().iterations != it) {
      std::printf("MISMATCH iterations step %d: %d vs %d\n", step, hip->lastResult().iterations, it);
      ++failures;
    }
    expect_equal(wgd.u, u, "u");
    expect_equal(wgd.v, v, "v");
    expect_equal(wgd.w, w, "w");
    expect_equal(solver->getLambda(), lam, "lambda");
    std::printf("step %d: %d iterations, %.3f ms on device, %.1f GLUPS\n", step, hip->lastResult().iterations,
                hip->lastResult().ms_total, hip->lastResult().glups);
  }
  // the logLawFlag == 1 branch (qesWindsMain.cpp:157-165): solve, then
  // wallLogBC on the solved field, u0 left equal to u
  set_inflow(wgd, dom, 285.0f, bl);
  wgd.z0 = 0.1f;
  const std::vector<float> u0 = wgd.u0, v0 = wgd.v0, w0 = wgd.w0;
  solver->resetLambda();
  std::fill(lam.begin(), lam.end(), 0.f);
  hip->solveWithLogLaw(&wgd, 200);
  oracle_solve_rb_f32(dom.nx(), dom.ny(), dom.nz(), dom.dx(), dom.dy(), dom.dz_array.data(), wgd.e.data(),
                      wgd.f.data(), wgd.g.data(), wgd.h.data(), wgd.m.data(), wgd.n.data(), wgd.icellflag.data(),
                      u0.data(), v0.data(), w0.data(), R.data(), lam.data(), lam_old.data(), 1.78f, 1e-9f, 200, 1,
                      1, u.data(), v.data(), w.data(), &me);
  float dzmin = dom.dz_array[0];
  for (float d : dom.dz_array) dzmin = d < dzmin ? d : dzmin;
  const long nwall = oracle_wall_log_bc(dom.nx(), dom.ny(), dom.nz(), dom.dx(), dzmin, wgd.z0,
                                        wgd.icellflag.data(), u.data(), v.data(), w.data());
  expect_close(wgd.u, u, "u (log law)");
  expect_close(wgd.v, v, "v (log law)");
  expect_close(wgd.w, w, "w (log law)");
  expect_equal(wgd.u0, wgd.u, "u0 == u after the log-law branch");
  expect_equal(wgd.w0, wgd.w, "w0 == w after the log-law branch");
  expect_equal(solver->getLambda(), lam, "lambda (log law)");
  std::printf("log-law step: %ld wall faces, %d iterations\n", nwall, hip->lastResult().iterations);

  // field-accuracy mode: estimated omega, production stop rule at 1e-4,
  // FP64 residual every 100 iterations; the field within 1e-4 m/s of the
  // FP64 oracle's converged one, in far fewer iterations than the cap
  set_inflow(wgd, dom, 233.0f, bl);
  {
    std::vector<double> R64(nc), l64(nc, 0.0), l64o(nc), u64(nf), v64(nf), w64(nf);
    double me64 = 0.0;
    oracle_divergence_f64(dom.nx(), dom.ny(), dom.nz(), dom.dx(), dom.dy(), dom.dz_array.data(), wgd.e.data(),
                          wgd.f.data(), wgd.g.data(), wgd.h.data(), wgd.m.data(), wgd.n.data(), wgd.u0.data(),
                          wgd.v0.data(), wgd.w0.data(), 1, R64.data());
    oracle_sor_rb_f64(dom.nx(), dom.ny(), dom.nz(), wgd.e.data(), wgd.f.data(), wgd.g.data(), wgd.h.data(),
                      wgd.m.data(), wgd.n.data(), R64.data(), l64.data(), l64o.data(), 1.78, 1e-10, 200000, &me64);
    oracle_velocity_f64(dom.nx(), dom.ny(), dom.nz(), dom.dx(), dom.dy(), dom.dz_array.data(), wgd.f.data(),
                        wgd.h.data(), wgd.n.data(), wgd.icellflag.data(), wgd.u0.data(), wgd.v0.data(),
                        wgd.w0.data(), l64.data(), 1, 1, u64.data(), v64.data(), w64.data());
    solver->resetLambda();
    hip->setFieldAccuracyMode(true);
    solver->solve(&wgd, 20000);
    double err = 0.0;
    for (long i = 0; i < nf; ++i) {
      err = std::fmax(err, std::fabs((double)wgd.u[i] - u64[i]));
      err = std::fmax(err, std::fabs((double)wgd.v[i] - v64[i]));
      err = std::fmax(err, std::fabs((double)wgd.w[i] - w64[i]));
    }
    const int its = hip->lastResult().iterations;
    std::printf("field-accuracy mode: omega %.4f, %d iterations, L_inf %.3e m/s from FP64\n", hip->omegaUsed(),
                its, err);
    if (!(err <= 1e-4) || its >= 5000 || !(hip->omegaUsed() > 1.78f && hip->omegaUsed() < 2.f)) {
      std::printf("MISMATCH field-accuracy mode\n");
      ++failures;
    }
    hip->setFieldAccuracyMode(false);
  }
  std::printf(failures ? "FAIL\n" : "PASS\n");
  return failures ? 1 : 0;
}

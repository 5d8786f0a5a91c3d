/*
 * qesmc_oracle.c -- CPU restatement of QES-Winds' mass-consistency step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the reference results are compared
 * against: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product (libqesmc, HIP) never links, loads or calls anything here.
 *
 * Every function restates one piece of the reference (UtahEFD/QES-Public,
 * CMake project 2.3.0) with the same loop nests, bounds, evaluation order and
 * float/double promotions, so that with -ffp-contract=off on x86-64 (SSE,
 * FLT_EVAL_METHOD == 0, no FMA) the fp32 paths reproduce the reference's
 * arithmetic operation for operation.  Citations are file:line in the
 * reference tree.
 *
 * Index conventions (src/qes/Domain.h:143-175):
 *   nx = ncx+1, ny = ncy+1, nz = ncz+2 (staggered sizes, Domain.cpp:12-14)
 *   cell(i,j,k) = i + j*(nx-1) + k*(nx-1)*(ny-1)        (cell-centred arrays)
 *   face(i,j,k) = i + j*nx + k*nx*ny                     (face arrays)
 *
 * Pinning status of each function is recorded in the header comment of the
 * function and in tests/test_oracle_*.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CELL(i, j, k) ((long)(i) + (long)(j) * (nx - 1) + (long)(k) * (long)(nx - 1) * (ny - 1))
#define FACE(i, j, k) ((long)(i) + (long)(j) * nx + (long)(k) * (long)nx * ny)

/* ------------------------------------------------------------------------ */
/* a6: divergence source R                                                   */
/* ------------------------------------------------------------------------ */

/* Solver_CPU_RB.cpp:56-80 (identical in Solver_CPU.cpp:56-77).
 * Quirk A1 kept: the vertical term is m*dz_k*0.5*(dz_k+dz_{k+1})*w0, i.e. the
 * plain difference of w0, and its first product runs in double because of the
 * double literal 0.5.  Only layers k in [1, nz-3] are written. */
void oracle_divergence_f32(int nx, int ny, int nz, float dx, float dy, const float *dz_array,
                           const float *e, const float *f, const float *g, const float *h,
                           const float *m, const float *n, const float *u0, const float *v0,
                           const float *w0, int alpha1, float *R)
{
  const double pre = -2.0f * pow(alpha1, 2.0);
#pragma omp parallel for schedule(static)
  for (int k = 1; k < nz - 2; ++k) {
    for (int j = 0; j < ny - 1; ++j) {
      for (int i = 0; i < nx - 1; ++i) {
        long c = CELL(i, j, k);
        long fc = FACE(i, j, k);
        R[c] = pre
               * (((e[c] * u0[fc + 1] - f[c] * u0[fc]) * dx)
                  + ((g[c] * v0[fc + nx] - h[c] * v0[fc]) * dy)
                  + ((m[c] * dz_array[k] * 0.5 * (dz_array[k] + dz_array[k + 1]) * w0[fc + (long)nx * ny]
                      - n[c] * dz_array[k] * 0.5f * (dz_array[k] + dz_array[k - 1]) * w0[fc])));
      }
    }
  }
}

/* FP64 restatement of the same operator: every operation of
 * Solver_CPU_RB.cpp:56-80 carried out in double on the (exactly promoted)
 * fp32 inputs.  This defines the discrete problem the FP64 converged oracle
 * solves. */
void oracle_divergence_f64(int nx, int ny, int nz, double dx, double dy, const float *dz_array,
                           const float *e, const float *f, const float *g, const float *h,
                           const float *m, const float *n, const float *u0, const float *v0,
                           const float *w0, int alpha1, double *R)
{
  const double pre = -2.0 * pow(alpha1, 2.0);
#pragma omp parallel for schedule(static)
  for (int k = 1; k < nz - 2; ++k) {
    for (int j = 0; j < ny - 1; ++j) {
      for (int i = 0; i < nx - 1; ++i) {
        long c = CELL(i, j, k);
        long fc = FACE(i, j, k);
        double dzk = dz_array[k], dzp = dz_array[k + 1], dzm = dz_array[k - 1];
        R[c] = pre
               * ((((double)e[c] * u0[fc + 1] - (double)f[c] * u0[fc]) * dx)
                  + (((double)g[c] * v0[fc + nx] - (double)h[c] * v0[fc]) * dy)
                  + (((double)m[c] * dzk * 0.5 * (dzk + dzp) * w0[fc + (long)nx * ny]
                      - (double)n[c] * dzk * 0.5 * (dzk + dzm) * w0[fc])));
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* a7-a11: SOR loops                                                          */
/* ------------------------------------------------------------------------ */

/* One red-black colour pass, Solver_CPU_RB.cpp:109-129 (colour 0 = red) and
 * :132-151 (colour 1 = black).  Expression kept operand for operand. */
static void rb_pass_f32(int nx, int ny, int nz, int colour, const float *e, const float *f,
                        const float *g, const float *h, const float *m, const float *n,
                        const float *R, float *lambda, float omega)
{
  const long sy = nx - 1, sz = (long)(nx - 1) * (ny - 1);
#pragma omp for schedule(static)
  for (int k = 1; k < nz - 2; ++k) {
    for (int j = 1; j < ny - 2; ++j) {
      for (int i = 1; i < nx - 2; ++i) {
        if (((i + j + k) % 2) == colour) {
          long c = CELL(i, j, k);
          lambda[c] = (omega / (e[c] + f[c] + g[c] + h[c] + m[c] + n[c]))
                        * (e[c] * lambda[c + 1] + f[c] * lambda[c - 1] + g[c] * lambda[c + sy]
                           + h[c] * lambda[c - sy] + m[c] * lambda[c + sz] + n[c] * lambda[c - sz]
                           - R[c])
                      + (1.0f - omega) * lambda[c];
        }
      }
    }
  }
}

/* Solver_CPU_RB::solve SOR loop, Solver_CPU_RB.cpp:88-184:
 *   while (iter < itermax && max_error > tol) {
 *     lambda_old = lambda (:100-104); red (:109-129); black (:132-151);
 *     mirror lambda(k=0) = lambda(k=1) (:154-162);
 *     max_error = max over k in [1,nz-2], all i,j of fabs(lambda-lambda_old) (:164-180)
 *   }
 * lambda_old is caller-provided scratch of the same size.  Returns the
 * iteration count (printed, not returned, by the reference, :192) and stores
 * the last max_error.  OpenMP work-sharing as in the reference; the result is
 * independent of the thread count because updates within a colour are
 * independent. */
int oracle_sor_rb_f32(int nx, int ny, int nz, const float *e, const float *f, const float *g,
                      const float *h, const float *m, const float *n, const float *R,
                      float *lambda, float *lambda_old, float omega, float tol, int itermax,
                      float *max_error_out)
{
  int iter = 0;
  float max_error = 1.0f;
  const long ncell = (long)(nx - 1) * (ny - 1) * (nz - 1);
  while (iter < itermax && max_error > tol) {
    float me = 0.0f;
#pragma omp parallel
    {
#pragma omp for schedule(static)
      for (long q = 0; q < ncell; ++q) lambda_old[q] = lambda[q];
      rb_pass_f32(nx, ny, nz, 0, e, f, g, h, m, n, R, lambda, omega);
      rb_pass_f32(nx, ny, nz, 1, e, f, g, h, m, n, R, lambda, omega);
#pragma omp for schedule(static)
      for (int j = 0; j < ny - 1; ++j)
        for (int i = 0; i < nx - 1; ++i) lambda[CELL(i, j, 0)] = lambda[CELL(i, j, 1)];
#pragma omp for schedule(static) reduction(max : me)
      for (int k = 1; k < nz - 1; ++k)
        for (int j = 0; j < ny - 1; ++j)
          for (int i = 0; i < nx - 1; ++i) {
            long c = CELL(i, j, k);
            float err = fabsf(lambda[c] - lambda_old[c]);
            if (err > me) me = err;
          }
    }
    max_error = me;
    iter += 1;
  }
  if (max_error_out) *max_error_out = max_error;
  return iter;
}

/* The red-black loop of oracle_sor_rb_f32 with a factor per colour pass
 * (omegas[2t] for the red pass of iteration t, omegas[2t+1] for the black
 * pass) and a fixed count of iterations: the over-relaxation schedules the
 * library offers beside the reference's fixed factor (cyclic Chebyshev
 * semi-iteration).  No counterpart in the reference; each pass is the
 * reference's pass, Solver_CPU_RB.cpp:109-162, at that pass's factor. */
void oracle_sor_rb_sched_f32(int nx, int ny, int nz, const float *e, const float *f, const float *g,
                             const float *h, const float *m, const float *n, const float *R,
                             float *lambda, const float *omegas, int iters)
{
  for (int t = 0; t < iters; ++t) {
#pragma omp parallel
    {
      rb_pass_f32(nx, ny, nz, 0, e, f, g, h, m, n, R, lambda, omegas[2 * t]);
      rb_pass_f32(nx, ny, nz, 1, e, f, g, h, m, n, R, lambda, omegas[2 * t + 1]);
#pragma omp for schedule(static)
      for (int j = 0; j < ny - 1; ++j)
        for (int i = 0; i < nx - 1; ++i) lambda[CELL(i, j, 0)] = lambda[CELL(i, j, 1)];
    }
  }
}

/* Solver_CPU::solve SOR loop, Solver_CPU.cpp:88-140: one in-place
 * lexicographic sweep (k, then j, then i) per iteration, then the mirror and
 * the same max|dlambda| stop test.  Single-threaded as in the reference. */
int oracle_sor_serial_f32(int nx, int ny, int nz, const float *e, const float *f, const float *g,
                          const float *h, const float *m, const float *n, const float *R,
                          float *lambda, float *lambda_old, float omega, float tol, int itermax,
                          float *max_error_out)
{
  int iter = 0;
  float max_error = 1.0f;
  const long sy = nx - 1, sz = (long)(nx - 1) * (ny - 1);
  const long ncell = (long)(nx - 1) * (ny - 1) * (nz - 1);
  while (iter < itermax && max_error > tol) {
    memcpy(lambda_old, lambda, sizeof(float) * ncell);
    for (int k = 1; k < nz - 2; ++k)
      for (int j = 1; j < ny - 2; ++j)
        for (int i = 1; i < nx - 2; ++i) {
          long c = CELL(i, j, k);
          lambda[c] = (omega / (e[c] + f[c] + g[c] + h[c] + m[c] + n[c]))
                        * (e[c] * lambda[c + 1] + f[c] * lambda[c - 1] + g[c] * lambda[c + sy]
                           + h[c] * lambda[c - sy] + m[c] * lambda[c + sz] + n[c] * lambda[c - sz]
                           - R[c])
                      + (1.0f - omega) * lambda[c];
        }
    for (int j = 0; j < ny - 1; ++j)
      for (int i = 0; i < nx - 1; ++i) lambda[CELL(i, j, 0)] = lambda[CELL(i, j, 1)];
    max_error = 0.0f;
    for (int k = 1; k < nz - 1; ++k)
      for (int j = 0; j < ny - 1; ++j)
        for (int i = 0; i < nx - 1; ++i) {
          long c = CELL(i, j, k);
          float err = fabsf(lambda[c] - lambda_old[c]);
          if (err > max_error) max_error = err;
        }
    iter += 1;
  }
  if (max_error_out) *max_error_out = max_error;
  return iter;
}

/* FP64 red-black SOR: the loop of Solver_CPU_RB.cpp:95-184 with lambda, R,
 * the relaxation arithmetic and the stop test in double.  Used as the
 * converged oracle (tol 1e-9 reachable in double).  Coefficients stay the
 * caller's fp32 values, promoted exactly. */
int oracle_sor_rb_f64(int nx, int ny, int nz, const float *e, const float *f, const float *g,
                      const float *h, const float *m, const float *n, const double *R,
                      double *lambda, double *lambda_old, double omega, double tol, int itermax,
                      double *max_error_out)
{
  int iter = 0;
  double max_error = 1.0;
  const long sy = nx - 1, sz = (long)(nx - 1) * (ny - 1);
  const long ncell = (long)(nx - 1) * (ny - 1) * (nz - 1);
  while (iter < itermax && max_error > tol) {
    double me = 0.0;
#pragma omp parallel
    {
#pragma omp for schedule(static)
      for (long q = 0; q < ncell; ++q) lambda_old[q] = lambda[q];
      for (int colour = 0; colour < 2; ++colour) {
#pragma omp for schedule(static)
        for (int k = 1; k < nz - 2; ++k)
          for (int j = 1; j < ny - 2; ++j)
            for (int i = 1; i < nx - 2; ++i) {
              if (((i + j + k) % 2) != colour) continue;
              long c = CELL(i, j, k);
              double ec = e[c], fc = f[c], gc = g[c], hc = h[c], mc = m[c], nc = n[c];
              lambda[c] = (omega / (ec + fc + gc + hc + mc + nc))
                            * (ec * lambda[c + 1] + fc * lambda[c - 1] + gc * lambda[c + sy]
                               + hc * lambda[c - sy] + mc * lambda[c + sz] + nc * lambda[c - sz]
                               - R[c])
                          + (1.0 - omega) * lambda[c];
            }
      }
#pragma omp for schedule(static)
      for (int j = 0; j < ny - 1; ++j)
        for (int i = 0; i < nx - 1; ++i) lambda[CELL(i, j, 0)] = lambda[CELL(i, j, 1)];
#pragma omp for schedule(static) reduction(max : me)
      for (int k = 1; k < nz - 1; ++k)
        for (int j = 0; j < ny - 1; ++j)
          for (int i = 0; i < nx - 1; ++i) {
            long c = CELL(i, j, k);
            double err = fabs(lambda[c] - lambda_old[c]);
            if (err > me) me = err;
          }
    }
    max_error = me;
    iter += 1;
  }
  if (max_error_out) *max_error_out = max_error;
  return iter;
}

/* ------------------------------------------------------------------------ */
/* a12-a13: velocity correction and solid zeroing                            */
/* ------------------------------------------------------------------------ */

/* Solver_CPU_RB.cpp:194-284: u = u0 (all faces), then for k in [1,nz-3],
 * j in [1,ny-2], i in [1,nx-2]:
 *   u = u0 + (1/(2 a1^2)) * f * dx * (lambda_c - lambda_{i-1})    (:214-227)
 *   v = v0 + (1/(2 a1^2)) * h * dy * (lambda_c - lambda_{j-1})    (:230-243)
 *   w = w0 + (1/(2 a2^2)) * n * dz_k * (lambda_c - lambda_{k-1})  (:246-259)
 * then every face of a cell with icellflag 0 or 2 (k in [1,nz-2]) is zeroed
 * (:262-284). */
void oracle_velocity_f32(int nx, int ny, int nz, float dx, float dy, const float *dz_array,
                         const float *f, const float *h, const float *n, const int *icellflag,
                         const float *u0, const float *v0, const float *w0, const float *lambda,
                         int alpha1, int alpha2, float *u, float *v, float *w)
{
  const long nface = (long)nx * ny * nz;
  const long sy = nx - 1, sz = (long)(nx - 1) * (ny - 1);
  const float p1 = 1.0f / (2.0f * (float)pow(alpha1, 2.0));
  const float p2 = 1.0f / (2.0f * (float)pow(alpha2, 2.0));
#pragma omp parallel
  {
#pragma omp for schedule(static)
    for (long q = 0; q < nface; ++q) {
      u[q] = u0[q];
      v[q] = v0[q];
      w[q] = w0[q];
    }
#pragma omp for schedule(static)
    for (int k = 1; k < nz - 2; ++k)
      for (int j = 1; j < ny - 1; ++j)
        for (int i = 1; i < nx - 1; ++i) {
          long c = CELL(i, j, k), fc = FACE(i, j, k);
          u[fc] = u0[fc] + p1 * f[c] * dx * (lambda[c] - lambda[c - 1]);
          v[fc] = v0[fc] + p1 * h[c] * dy * (lambda[c] - lambda[c - sy]);
          w[fc] = w0[fc] + p2 * n[c] * dz_array[k] * (lambda[c] - lambda[c - sz]);
        }
#pragma omp for schedule(static)
    for (int k = 1; k < nz - 1; ++k)
      for (int j = 0; j < ny - 1; ++j)
        for (int i = 0; i < nx - 1; ++i) {
          long c = CELL(i, j, k), fc = FACE(i, j, k);
          if (icellflag[c] == 0 || icellflag[c] == 2) {
            u[fc] = 0;
            u[fc + 1] = 0;
            v[fc] = 0;
            v[fc + nx] = 0;
            w[fc] = 0;
            w[fc + (long)nx * ny] = 0;
          }
        }
  }
}

/* The same correction and zeroing in double, from an FP64 lambda
 * (Solver_CPU_RB.cpp:194-284 with every operation in double). */
void oracle_velocity_f64(int nx, int ny, int nz, double dx, double dy, const float *dz_array,
                         const float *f, const float *h, const float *n, const int *icellflag,
                         const float *u0, const float *v0, const float *w0, const double *lambda,
                         int alpha1, int alpha2, double *u, double *v, double *w)
{
  const long nface = (long)nx * ny * nz;
  const long sy = nx - 1, sz = (long)(nx - 1) * (ny - 1);
  const double p1 = 1.0 / (2.0 * pow(alpha1, 2.0));
  const double p2 = 1.0 / (2.0 * pow(alpha2, 2.0));
#pragma omp parallel
  {
#pragma omp for schedule(static)
    for (long q = 0; q < nface; ++q) {
      u[q] = u0[q];
      v[q] = v0[q];
      w[q] = w0[q];
    }
#pragma omp for schedule(static)
    for (int k = 1; k < nz - 2; ++k)
      for (int j = 1; j < ny - 1; ++j)
        for (int i = 1; i < nx - 1; ++i) {
          long c = CELL(i, j, k), fc = FACE(i, j, k);
          u[fc] = u0[fc] + p1 * f[c] * dx * (lambda[c] - lambda[c - 1]);
          v[fc] = v0[fc] + p1 * h[c] * dy * (lambda[c] - lambda[c - sy]);
          w[fc] = w0[fc] + p2 * n[c] * (double)dz_array[k] * (lambda[c] - lambda[c - sz]);
        }
#pragma omp for schedule(static)
    for (int k = 1; k < nz - 1; ++k)
      for (int j = 0; j < ny - 1; ++j)
        for (int i = 0; i < nx - 1; ++i) {
          long c = CELL(i, j, k), fc = FACE(i, j, k);
          if (icellflag[c] == 0 || icellflag[c] == 2) {
            u[fc] = 0;
            u[fc + 1] = 0;
            v[fc] = 0;
            v[fc + nx] = 0;
            w[fc] = 0;
            w[fc + (long)nx * ny] = 0;
          }
        }
  }
}

/* Whole Solver_CPU_RB::solve (Solver_CPU_RB.cpp:41-291) for convenience:
 * R, SOR loop, correction.  R, lambda_old are caller scratch; lambda is
 * in/out (warm start when non-zero, as when resetLambda is not called). */
int oracle_solve_rb_f32(int nx, int ny, int nz, float dx, float dy, const float *dz_array,
                        const float *e, const float *f, const float *g, const float *h,
                        const float *m, const float *n, const int *icellflag, const float *u0,
                        const float *v0, const float *w0, float *R, float *lambda,
                        float *lambda_old, float omega, float tol, int itermax, int alpha1,
                        int alpha2, float *u, float *v, float *w, float *max_error_out)
{
  oracle_divergence_f32(nx, ny, nz, dx, dy, dz_array, e, f, g, h, m, n, u0, v0, w0, alpha1, R);
  int it = oracle_sor_rb_f32(nx, ny, nz, e, f, g, h, m, n, R, lambda, lambda_old, omega, tol,
                             itermax, max_error_out);
  oracle_velocity_f32(nx, ny, nz, dx, dy, dz_array, f, h, n, icellflag, u0, v0, w0, lambda, alpha1,
                      alpha2, u, v, w);
  return it;
}

/* Divergence of a face field measured with the reference operator, i.e.
 * R(u)/(-2 a1^2) of Solver_CPU_RB.cpp:56-80 evaluated in double on the
 * corrected field, over fluid cells (icellflag 1) of the updated interior
 * k in [1,nz-3], j in [1,ny-3], i in [1,nx-3].  Returns the L-infinity norm.
 * With physical != 0 the vertical term is divided by dz_k (the physically
 * consistent divergence, quirk A1 removed). */
double oracle_div_linf(int nx, int ny, int nz, double dx, double dy, const float *dz_array,
                       const float *e, const float *f, const float *g, const float *h,
                       const float *m, const float *n, const int *icellflag, const float *u,
                       const float *v, const float *w, int physical)
{
  double mx = 0.0;
#pragma omp parallel for schedule(static) reduction(max : mx)
  for (int k = 1; k < nz - 2; ++k)
    for (int j = 1; j < ny - 2; ++j)
      for (int i = 1; i < nx - 2; ++i) {
        long c = CELL(i, j, k), fc = FACE(i, j, k);
        if (icellflag[c] != 1) continue;
        double d;
        if (physical) {
          d = ((double)u[fc + 1] - u[fc]) / dx + ((double)v[fc + nx] - v[fc]) / dy
              + ((double)w[fc + (long)nx * ny] - w[fc]) / dz_array[k];
        } else {
          double dzk = dz_array[k];
          d = (((double)e[c] * u[fc + 1] - (double)f[c] * u[fc]) * dx)
              + (((double)g[c] * v[fc + nx] - (double)h[c] * v[fc]) * dy)
              + ((double)m[c] * dzk * 0.5 * (dzk + dz_array[k + 1]) * w[fc + (long)nx * ny]
                 - (double)n[c] * dzk * 0.5 * (dzk + dz_array[k - 1]) * w[fc]);
        }
        d = fabs(d);
        if (d > mx) mx = d;
      }
  return mx;
}

/* ------------------------------------------------------------------------ */
/* Inputs of the path (a1, a3, a4, a5): grid, stair-step flags, wall         */
/* coefficients, single-sensor inflow profile.                               */
/* ------------------------------------------------------------------------ */

/* Domain::defineVerticalGrid, Domain.cpp:146-160.  dz_array has nz-1
 * entries; z_face nz entries; z nz-1 entries. */
void oracle_vertical_grid(int nz, const float *dz_array, float *z_face, float *z)
{
  z_face[0] = -dz_array[0];
  z_face[1] = 0.0f;
  for (int k = 2; k < nz; ++k) z_face[k] = z_face[k - 1] + dz_array[k - 1];
  z[0] = -0.5f * dz_array[0];
  for (int k = 1; k < nz - 1; ++k) z[k] = 0.5f * (z_face[k] + z_face[k + 1]);
}

/* WINDSGeneralData::allocateMemory (WINDSGeneralData.cpp:1077-1095: e..n = 1,
 * icellflag = 1) and the ground ghost layer icellflag(k=0) = 2
 * (WINDSGeneralData.cpp:439-444). */
void oracle_init_cells(int nx, int ny, int nz, float *e, float *f, float *g, float *h, float *m,
                       float *n, int *icellflag)
{
  const long ncell = (long)(nx - 1) * (ny - 1) * (nz - 1);
  for (long q = 0; q < ncell; ++q) {
    e[q] = f[q] = g[q] = h[q] = m[q] = n[q] = 1.0f;
    icellflag[q] = 1;
  }
  for (int j = 0; j < ny - 1; ++j)
    for (int i = 0; i < nx - 1; ++i) icellflag[CELL(i, j, 0)] = 2;
}

/* One rectangular building: RectangularBuilding::parseValues vertex list
 * (RectangularBuilding.h:86-103, rotation in degrees) followed by
 * PolyBuilding::setCellFlags with meshTypeFlag 0 (PolyBuilding.cpp:130-236):
 * bounding-box index ranges, k_start/k_end searches over z_face / z, and a
 * PNPOLY crossing test at cell centres.  Cells inside get icellflag 0. */
void oracle_rect_building(int nx, int ny, int nz, float dx, float dy_unused, const float *z_face,
                          const float *z, float x_start, float y_start, float L, float W,
                          float H, float base_height, float rotation_deg, int *icellflag)
{
  (void)dy_unused;
  /* getDomainSize() returns (dx, dx, dz): dy is dx (Domain.h:104). */
  const float dy = dx;
  float height_eff = H + base_height;
  double rot = rotation_deg * M_PI / 180.0;
  float rotf = (float)rot;
  float xv[5], yv[5];
  xv[0] = xv[4] = x_start;
  yv[0] = yv[4] = y_start;
  /* sin/cos of a float resolve to the float overloads in the reference (C++) */
  xv[1] = x_start - W * sinf(rotf);
  yv[1] = y_start + W * cosf(rotf);
  xv[2] = xv[1] + L * cosf(rotf);
  yv[2] = yv[1] + L * sinf(rotf);
  xv[3] = x_start + L * cosf(rotf);
  yv[3] = y_start + L * sinf(rotf);

  float x_min = xv[0], x_max = xv[0], y_min = yv[0], y_max = yv[0];
  for (int id = 1; id < 5; ++id) {
    if (xv[id] > x_max) x_max = xv[id];
    if (xv[id] < x_min) x_min = xv[id];
    if (yv[id] > y_max) y_max = yv[id];
    if (yv[id] < y_min) y_min = yv[id];
  }
  int i_start = (int)(x_min / dx);
  int i_end = (int)((x_max / dx) + 1);
  int j_start = (int)(y_min / dy);
  int j_end = (int)((y_max / dy) + 1);

  const int zsize = nz - 1;
  int k_start = 1, k_end = 1;
  for (int k = 1; k < zsize; ++k) {
    k_start = k;
    if (base_height <= z_face[k + 1]) break;
  }
  for (int k = 0; k < zsize; ++k) {
    k_end = k + 1;
    if (height_eff < z[k + 1]) break;
  }

  for (int j = j_start; j <= j_end; ++j) {
    float y_cent = (j + 0.5) * dy;
    for (int i = i_start; i <= i_end; ++i) {
      float x_cent = (i + 0.5) * dx;
      unsigned vert_id = 0;
      unsigned start_poly = 0;
      int num_crossing = 0;
      while (vert_id < 5 - 1) {
        if ((yv[vert_id] <= y_cent && yv[vert_id + 1] > y_cent)
            || (yv[vert_id] > y_cent && yv[vert_id + 1] <= y_cent)) {
          float ray_intersect = (y_cent - yv[vert_id]) / (yv[vert_id + 1] - yv[vert_id]);
          if (x_cent < (xv[vert_id] + ray_intersect * (xv[vert_id + 1] - xv[vert_id]))) {
            num_crossing += 1;
          }
        }
        vert_id += 1;
        if (xv[vert_id] == xv[start_poly] && yv[vert_id] == yv[start_poly]) {
          vert_id += 1;
          start_poly = vert_id;
        }
      }
      if ((num_crossing % 2) != 0) {
        if (i < 0 || j < 0 || i >= nx - 1 || j >= ny - 1) continue; /* outside: never in configs */
        for (int k = k_start; k < k_end; ++k) {
          long c = CELL(i, j, k);
          if (icellflag[c] != 7) icellflag[c] = 0;
        }
      }
    }
  }
}

static inline int solid(int flag) { return flag == 0 || flag == 2; }

/* Wall::defineWalls (Wall.cpp:44-349) for stair-step geometry (icellflag in
 * {0,1,2}; the cut-cell flags 7/8 and their branches are out of scope):
 * clamp (:48-91), wall rules (:128-163, repeated at :288-348), and the
 * all-zero reset of solid cells (:251-286, a no-op for stair-step cells whose
 * coefficients are never zeroed).  The east and north lookups use cellAdd and
 * so wrap into the next row / layer at i = nx-2 / j = ny-2 (quirk A11). */
void oracle_define_walls(int nx, int ny, int nz, const int *icellflag, float *e, float *f,
                         float *g, float *h, float *m, float *n)
{
  const long sy = nx - 1, sz = (long)(nx - 1) * (ny - 1);
  for (int k = 1; k < nz - 1; k++)
    for (int i = 0; i < nx - 1; i++)
      for (int j = 0; j < ny - 1; j++) {
        long c = CELL(i, j, k);
        float *arr[6] = {e, f, g, h, m, n};
        for (int a = 0; a < 6; ++a) {
          if (arr[a][c] < 0.05) arr[a][c] = 0.0f;
          if (arr[a][c] > 1.0) arr[a][c] = 1.0f;
        }
      }
  for (int k = 1; k < nz - 2; k++)
    for (int i = 0; i < nx - 1; i++)
      for (int j = 0; j < ny - 1; j++) {
        long c = CELL(i, j, k);
        if (icellflag[c] != 0 && icellflag[c] != 2) {
          if (solid(icellflag[c - sz])) n[c] = 0.0f;
          if (solid(icellflag[c + sz])) m[c] = 0.0f;
          if (solid(icellflag[c - 1]) && i > 0) f[c] = 0.0f;
          if (solid(icellflag[c + 1])) e[c] = 0.0f;
          if (solid(icellflag[c - sy]) && j > 0) h[c] = 0.0f;
          if (solid(icellflag[c + sy])) g[c] = 0.0f;
        }
      }
}

/* Wall::solverCoefficients (Wall.cpp:827-846) for k in [1,nz-3]; dy is dx
 * (quirk A2, getDomainSize at Domain.h:104). */
void oracle_solver_coefficients(int nx, int ny, int nz, float dx, const float *dz_array, float *e,
                                float *f, float *g, float *h, float *m, float *n)
{
  const float dy = dx;
  for (int k = 1; k < nz - 2; k++)
    for (int j = 0; j < ny - 1; j++)
      for (int i = 0; i < nx - 1; i++) {
        long c = CELL(i, j, k);
        e[c] = e[c] / (dx * dx);
        f[c] = f[c] / (dx * dx);
        g[c] = g[c] / (dy * dy);
        h[c] = h[c] / (dy * dy);
        m[c] = m[c] / (dz_array[k] * 0.5f * (dz_array[k] + dz_array[k + 1]));
        n[c] = n[c] / (dz_array[k] * 0.5f * (dz_array[k] + dz_array[k - 1]));
      }
}

/* Single-sensor vertical profile on flat terrain (terrain_face_id = 1,
 * z_terrain = z_face[1] = 0), neutral stratification (1/L = 0), from
 * WindProfilerSensorType::sensorsProfiles (WindProfilerSensorType.cpp:220,
 * :290-295 uniform, :299-325 log, :328-333 power; the power-law exponent is
 * the siteZ0 field, quirk A7).  kind: 0 uniform, 1 log, 2 power.  Writes
 * u_prof/v_prof for k in [1, nz-2] (the layers the interpolation reads). */
void oracle_profile(int nz, const float *z, int kind, float U_ref, float z_ref, float z0_or_p,
                    float wind_dir, float *u_prof, float *v_prof)
{
  const float vk = 0.4f;
  float theta = (float)((270.0 - wind_dir) * M_PI / 180.0);
  float one_overL = 0.0f;
  float z_terrain = 0.0f;
  for (int k = 0; k < nz; ++k) u_prof[k] = v_prof[k] = 0.0f;
  if (kind == 0) {
    for (int k = 1; k < nz; ++k) {
      u_prof[k] = cosf(theta) * U_ref;
      v_prof[k] = sinf(theta) * U_ref;
    }
  } else if (kind == 1) {
    float psi = (float)(4.7 * z_ref * one_overL);
    float u_star = U_ref * vk / (logf((z_ref + z0_or_p) / z0_or_p) + psi);
    for (int k = 1; k < nz - 1; ++k) {
      psi = (float)(4.7 * (z[k] - z_terrain) * one_overL);
      u_prof[k] = (cosf(theta) * u_star / vk) * (logf(((z[k] - z_terrain) + z0_or_p) / z0_or_p) + psi);
      v_prof[k] = (sinf(theta) * u_star / vk) * (logf(((z[k] - z_terrain) + z0_or_p) / z0_or_p) + psi);
      if (fabsf(u_star) >= 2 * U_ref) {
        u_prof[k] = (cosf(theta) * U_ref);
        v_prof[k] = (sinf(theta) * U_ref);
      }
    }
  } else {
    for (int k = 1; k < nz - 1; ++k) {
      u_prof[k] = cosf(theta) * U_ref * powf(((z[k] - z_terrain) / z_ref), z0_or_p);
      v_prof[k] = sinf(theta) * U_ref * powf(((z[k] - z_terrain) / z_ref), z0_or_p);
    }
  }
}

/* WindProfilerSensorType::singleSensorInterpolation on flat terrain
 * (WindProfilerSensorType.cpp:571-621): u0 = u_prof[k], v0 = v_prof[k],
 * w0 = 0 at every face of layers k in [1, nz-2]; other faces stay 0
 * (applyWindProfile resizes to 0, WINDSGeneralData.cpp:1204-1209). */
void oracle_extrude_profile(int nx, int ny, int nz, const float *u_prof, const float *v_prof,
                            float *u0, float *v0, float *w0)
{
  const long nface = (long)nx * ny * nz;
  for (long q = 0; q < nface; ++q) u0[q] = v0[q] = w0[q] = 0.0f;
  for (int k = 1; k < nz - 1; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        long fc = FACE(i, j, k);
        u0[fc] = u_prof[k];
        v0[fc] = v_prof[k];
        w0[fc] = 0.0f;
      }
}

/* Wall::setVelocityZero (Wall.cpp:805-824): k in [0,nz-2], j in [1,ny-2],
 * i in [1,nx-2]; all six faces of a solid cell set to 0. */
void oracle_set_velocity_zero(int nx, int ny, int nz, const int *icellflag, float *u0, float *v0,
                              float *w0)
{
  for (int k = 0; k < nz - 1; k++)
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) {
        long c = CELL(i, j, k), fc = FACE(i, j, k);
        if (icellflag[c] == 0 || icellflag[c] == 2) {
          u0[fc] = 0.0f;
          u0[fc + 1] = 0.0f;
          v0[fc] = 0.0f;
          v0[fc + nx] = 0.0f;
          w0[fc] = 0.0f;
          w0[fc + (long)nx * ny] = 0.0f;
        }
      }
}

/* ------------------------------------------------------------------------ */
/* f4: turbulence model, velocity gradient tensor at cell centres            */
/* ------------------------------------------------------------------------ */

/* TURBGeneralData::derivativeVelocity, TURBGeneralData.cpp:479-552 (the wall
 * log-law corrections that follow at :553-556 are not restated).  Cells:
 * the fluid cells (icellflag not 0 or 2) with i in [1, nx-3], j in [1, ny-3],
 * k in [1, nz-3] (the icellfluid list, TURBGeneralData.cpp:264-283); every
 * other cell of the nine outputs is set to 0.  uniform_z selects the
 * uniform-grid branch (:521-533, dz = domain.dz()), else the stretched one
 * (:534-551) with z the cell-centre heights of Domain::defineVerticalGrid.
 * Float throughout; 4.0f and 0.5f as in the source.  G holds the nine arrays
 * in the order Gxx, Gxy, Gxz, Gyx, Gyy, Gyz, Gzx, Gzy, Gzz.
 * Pinned by: linear velocity fields give their exact gradients on uniform
 * and stretched grids (the stretched dudz/dvdz weights are exact for linear
 * profiles), with distinct coefficients per component so that a swapped
 * component or operand fails (tests/test_oracle.py). */
void oracle_velocity_gradient(int nx, int ny, int nz, float dx, float dy, float dz, const float *dz_array,
                              int uniform_z, const int *icellflag, const float *u, const float *v,
                              const float *w, float *G)
{
  const long ncell = (long)(nx - 1) * (ny - 1) * (nz - 1);
  float *Gxx = G, *Gxy = G + ncell, *Gxz = G + 2 * ncell, *Gyx = G + 3 * ncell, *Gyy = G + 4 * ncell;
  float *Gyz = G + 5 * ncell, *Gzx = G + 6 * ncell, *Gzy = G + 7 * ncell, *Gzz = G + 8 * ncell;
  float *z_face = (float *)malloc(sizeof(float) * nz), *z = (float *)malloc(sizeof(float) * (nz - 1));
  oracle_vertical_grid(nz, dz_array, z_face, z);
  memset(G, 0, sizeof(float) * 9 * ncell);
#pragma omp parallel for schedule(static)
  for (int k = 1; k < nz - 2; ++k)
    for (int j = 1; j < ny - 2; ++j)
      for (int i = 1; i < nx - 2; ++i) {
        const long c = CELL(i, j, k);
        if (icellflag[c] == 0 || icellflag[c] == 2) continue;
        Gxx[c] = (u[FACE(i + 1, j, k)] - u[FACE(i, j, k)]) / (dx);
        Gyx[c] = ((v[FACE(i + 1, j, k)] + v[FACE(i + 1, j + 1, k)]) - (v[FACE(i - 1, j, k)] + v[FACE(i - 1, j + 1, k)]))
                 / (4.0f * dx);
        Gzx[c] = ((w[FACE(i + 1, j, k)] + w[FACE(i + 1, j, k + 1)]) - (w[FACE(i - 1, j, k)] + w[FACE(i - 1, j, k + 1)]))
                 / (4.0f * dx);
        Gxy[c] = ((u[FACE(i, j + 1, k)] + u[FACE(i + 1, j + 1, k)]) - (u[FACE(i, j - 1, k)] + u[FACE(i + 1, j - 1, k)]))
                 / (4.0f * dy);
        Gyy[c] = (v[FACE(i, j + 1, k)] - v[FACE(i, j, k)]) / (dy);
        Gzy[c] = ((w[FACE(i, j + 1, k)] + w[FACE(i, j + 1, k + 1)]) - (w[FACE(i, j - 1, k)] + w[FACE(i, j - 1, k + 1)]))
                 / (4.0f * dy);
        if (uniform_z) {
          Gxz[c] = ((u[FACE(i, j, k + 1)] + u[FACE(i + 1, j, k + 1)]) - (u[FACE(i, j, k - 1)] + u[FACE(i + 1, j, k - 1)]))
                   / (4.0f * dz);
          Gyz[c] = ((v[FACE(i, j, k + 1)] + v[FACE(i, j + 1, k + 1)]) - (v[FACE(i, j, k - 1)] + v[FACE(i, j + 1, k - 1)]))
                   / (4.0f * dz);
          Gzz[c] = (w[FACE(i, j, k + 1)] - w[FACE(i, j, k)]) / (dz);
        } else {
          Gxz[c] = (0.5f * (z[k] - z[k - 1]) / (z[k + 1] - z[k])
                      * ((u[FACE(i, j, k + 1)] + u[FACE(i + 1, j, k + 1)]) - (u[FACE(i, j, k)] + u[FACE(i + 1, j, k)]))
                    + 0.5f * (z[k + 1] - z[k]) / (z[k] - z[k - 1])
                        * ((u[FACE(i, j, k)] + u[FACE(i + 1, j, k)]) - (u[FACE(i, j, k - 1)] + u[FACE(i + 1, j, k - 1)])))
                   / (z[k + 1] - z[k - 1]);
          Gyz[c] = (0.5f * (z[k] - z[k - 1]) / (z[k + 1] - z[k])
                      * ((v[FACE(i, j, k + 1)] + v[FACE(i, j + 1, k + 1)]) - (v[FACE(i, j, k)] + v[FACE(i, j + 1, k)]))
                    + 0.5f * (z[k + 1] - z[k]) / (z[k] - z[k - 1])
                        * ((v[FACE(i, j, k)] + v[FACE(i, j + 1, k)]) - (v[FACE(i, j, k - 1)] + v[FACE(i, j + 1, k - 1)])))
                   / (z[k + 1] - z[k - 1]);
          Gzz[c] = (w[FACE(i, j, k + 1)] - w[FACE(i, j, k)]) / (dz_array[k]);
        }
      }
  free(z_face);
  free(z);
}

/* ------------------------------------------------------------------------ */
/* f4: Wall::wallLogBC on stair-step walls                                  */
/* ------------------------------------------------------------------------ */

/* One wall-parallel log-law correction (the body of every loop of
 * Wall.cpp:677-797): 20 fixed-point steps of the friction velocity from the
 * second cell's velocity (a, b: its two wall-parallel components at face
 * `nb`), writing the normal component (*pn = 0) and the parallel ones of face
 * `f`.  Math as the reference evaluates it: Wall.h includes <math.h> with
 * `using namespace std` (Wall.h:38,46), so log/atan2/cos/sin of float
 * arguments are the float overloads, pow(float, 2.0) is double, and the
 * magnitude is a double sqrt stored to float.  vk = 0.4f
 * (WINDSGeneralData.h:161). */
static void log_law_face(float *A, float *B, float *N, long f, long nb, float dist1, float dist2, float z0)
{
  const float vk = 0.4f;
  float ustar_wall = 0.1f;
  for (int iter = 0; iter < 20; iter++) {
    float wind_dir = atan2f(B[nb], A[nb]);
    float vel_mag2 = (float)sqrt(pow((double)A[nb], 2.0) + pow((double)B[nb], 2.0));
    float vel_mag1 = vel_mag2 - (ustar_wall / vk) * logf(dist2 / dist1);
    N[f] = 0;
    A[f] = vel_mag1 * cosf(wind_dir);
    B[f] = vel_mag1 * sinf(wind_dir);
    float new_ustar = vk * vel_mag1 / logf(dist1 / z0);
    ustar_wall = new_ustar;
  }
}

/* Wall::wallLogBC (Wall.cpp:352-797) for stair-step geometry, applied the way
 * qesWindsMain.cpp:157-165 applies it to a solved field: u0 = u; wallLogBC;
 * u = u0 (so u, v, w are updated in place here).  The cut-cell loop over
 * wall_indices (Wall.cpp:381-655) visits only flags 7/8, which stair-step
 * geometry does not have; returns -1 (nothing done) when a 7/8 flag exists.
 *
 * The six index lists are those of Wall::defineWalls (Wall.cpp:288-340): for
 * k in [1,nz-3], i in [0,nx-2], j in [0,ny-2] (i outer, j inner), every cell
 * with icellflag not in {0,2} (and not 8) whose neighbour below / above / back
 * (i > 0) / front / right (j > 0) / left is solid (flag 0 or 2) contributes
 * its face index face(i,j,k); neighbours are read with cellAdd, so front and
 * left wrap at i = nx-2 / j = ny-2 (quirk A11), and every fluid cell at k = 1
 * has the ground ghost layer below it.  The passes run in the reference's
 * order -- below, above (dist 0.5 dz / 1.5 dz), back, front (0.5 dx / 1.5 dx),
 * right, left (0.5 dy / 1.5 dy, with dy = dx: getDomainSize, Domain.h:104)
 * -- each reading its second-cell face at +nx*ny, -nx*ny, +1, -1, +nx, -nx,
 * so a later pass sees the faces an earlier one rewrote.  dz is the Domain's
 * dz, the smallest dz_array entry for a stretched grid (Domain.cpp:89,131).
 * z0 is WGD->z0 (WINDSGeneralData.cpp:396-406: the sensor's site_z0).
 * Returns the number of wall faces processed. */
long oracle_wall_log_bc(int nx, int ny, int nz, float dx, float dz, float z0, const int *icellflag, float *u,
                        float *v, float *w)
{
  const long sy = nx - 1, sz = (long)(nx - 1) * (ny - 1);
  const long ncell = sz * (nz - 1);
  for (long c = 0; c < ncell; ++c)
    if (icellflag[c] == 7 || icellflag[c] == 8) return -1;
  long cap = ncell, cnt[6] = {0, 0, 0, 0, 0, 0};
  long *list[6];
  for (int p = 0; p < 6; ++p) list[p] = (long *)malloc(sizeof(long) * (size_t)(cap > 0 ? cap : 1));
  for (int k = 1; k < nz - 2; k++)
    for (int i = 0; i < nx - 1; i++)
      for (int j = 0; j < ny - 1; j++) {
        long c = CELL(i, j, k);
        long fc = FACE(i, j, k);
        int fl = icellflag[c];
        if (fl != 0 && fl != 2 && fl != 8) {
          if (solid(icellflag[c - sz])) list[0][cnt[0]++] = fc;
          if (solid(icellflag[c + sz])) list[1][cnt[1]++] = fc;
          if (solid(icellflag[c - 1]) && i > 0) list[2][cnt[2]++] = fc;
          if (solid(icellflag[c + 1])) list[3][cnt[3]++] = fc;
          if (solid(icellflag[c - sy]) && j > 0) list[4][cnt[4]++] = fc;
          if (solid(icellflag[c + sy])) list[5][cnt[5]++] = fc;
        }
      }
  const long fz = (long)nx * ny;
  const float dy = dx;
  float d1 = 0.5f * dz, d2 = 1.5f * dz;
  for (long q = 0; q < cnt[0]; ++q) log_law_face(u, v, w, list[0][q], list[0][q] + fz, d1, d2, z0);
  for (long q = 0; q < cnt[1]; ++q) log_law_face(u, v, w, list[1][q], list[1][q] - fz, d1, d2, z0);
  d1 = 0.5f * dx;
  d2 = 1.5f * dx;
  for (long q = 0; q < cnt[2]; ++q) log_law_face(v, w, u, list[2][q], list[2][q] + 1, d1, d2, z0);
  for (long q = 0; q < cnt[3]; ++q) log_law_face(v, w, u, list[3][q], list[3][q] - 1, d1, d2, z0);
  d1 = 0.5f * dy;
  d2 = 1.5f * dy;
  for (long q = 0; q < cnt[4]; ++q) log_law_face(u, w, v, list[4][q], list[4][q] + nx, d1, d2, z0);
  for (long q = 0; q < cnt[5]; ++q) log_law_face(u, w, v, list[5][q], list[5][q] - nx, d1, d2, z0);
  long total = 0;
  for (int p = 0; p < 6; ++p) {
    total += cnt[p];
    free(list[p]);
  }
  return total;
}

/* ------------------------------------------------------------------------ */
/* f4: wall corrections of the turbulence velocity gradient (stair-step)    */
/* ------------------------------------------------------------------------ */

/* The wall step of TURBGeneralData::derivativeVelocity (TURBGeneralData.cpp:
 * 553-556): for each wall object in order -- buildings (TURBWallBuilding,
 * cell flag 0) then terrain (TURBWallTerrain, cell flag 2) -- the velocity
 * wall treatment its flag selects (TURBWallBuilding.cpp:37-70,
 * TURBWallTerrain.cpp:37-70; turbParams buildingWallFlag / terrainWallFlag,
 * default 2, TURBParams.h:74):
 *   mode 1  set_loglaw_stairstep (TURBWallBuilding.cpp:87-169,
 *           TURBWallTerrain.cpp:88-...): the wall-normal derivatives of the
 *           wall-parallel components from the log law, 0.5 (q_f + q_f') /
 *           (0.5 d log(0.5 d / z0)), with d = dz_array[k], dx or dy = dx
 *           (getDomainSize, Domain.h:104); the building object uses WGD->z0,
 *           the terrain object the mean of the four surrounding entries of
 *           z0_domain_u / z0_domain_v; its sixth pass ("left of the cell")
 *           walks wall_below_indices, as the reference does;
 *   mode 2  comp_velocity_deriv_finitediff_stairstep (TURBWall.cpp:206-300):
 *           one-sided differences into the fluid, over z[k+1]-z[k],
 *           z[k]-z[k-1], dx, dy;
 *   mode 0  none.
 * The index lists are TURBWall::get_stairstep_wall_id's (TURBWall.cpp:
 * 39-121): i in [1,nx-2], j in [1,ny-2], k in [1,nz-3], a cell whose flag is
 * not 0 or 2, whose neighbour (flat index -(nx-1)(ny-1), +(nx-1)(ny-1), -1,
 * +1, -(nx-1), +(nx-1): so +1 at i = nx-2 and +(nx-1) at j = ny-2 reach the
 * next row / plane) has the object's flag, with face(i,j,k).  Passes run in
 * the order below, above, back, front, right, left, so a later one
 * overwrites what an earlier one wrote.  Arithmetic as written there: float
 * sums of two faces, 0.5 (double) times them, the quotient in double, stored
 * to float.  G holds Gxx..Gzz (9 ncell blocks) and is updated in place. */
static void turb_wall_object(int nx, int ny, int nz, float dx, const float *dz_array, const float *z, int cf,
                             int mode, float z0b, const float *z0u, const float *z0v, const int *icellflag,
                             const float *u, const float *v, const float *w, float *G)
{
  if (mode != 1 && mode != 2) return;
  const long sy = nx - 1, sz = (long)(nx - 1) * (ny - 1), ncell = sz * (nz - 1);
  float *Gxy = G + ncell, *Gxz = G + 2 * ncell, *Gyx = G + 3 * ncell, *Gyz = G + 5 * ncell;
  float *Gzx = G + 6 * ncell, *Gzy = G + 7 * ncell;
  long cnt[6] = {0, 0, 0, 0, 0, 0};
  long *lc[6], *lf[6];
  for (int p = 0; p < 6; ++p) {
    lc[p] = (long *)malloc(sizeof(long) * (size_t)(ncell > 0 ? ncell : 1));
    lf[p] = (long *)malloc(sizeof(long) * (size_t)(ncell > 0 ? ncell : 1));
  }
  for (int i = 1; i < nx - 1; i++)
    for (int j = 1; j < ny - 1; j++)
      for (int k = 1; k < nz - 2; k++) {
        const long c = CELL(i, j, k), f = FACE(i, j, k);
        if (icellflag[c] == 0 || icellflag[c] == 2) continue;
        const long nb[6] = {c - sz, c + sz, c - 1, c + 1, c - sy, c + sy};
        for (int p = 0; p < 6; ++p)
          if (icellflag[nb[p]] == cf) {
            lc[p][cnt[p]] = c;
            lf[p][cnt[p]] = f;
            cnt[p]++;
          }
      }
  const long fy = nx, fz = (long)nx * ny;
  const float dy = dx;
  for (int p = 0; p < 6; ++p) {
    /* mode 1's sixth pass walks the list of walls below (TURBWallBuilding.cpp:153,
       TURBWallTerrain.cpp: the same) */
    const int lp = (mode == 1 && p == 5) ? 0 : p;
    for (long q = 0; q < cnt[lp]; ++q) {
      const long c = lc[lp][q], f = lf[lp][q];
      const int k = (int)(c / sz);
      const int j = (int)((c - (long)k * sz) / sy);
      const int i = (int)(c - (long)j * sy - (long)k * sz);
      float z0 = z0b;
      if (mode == 1 && cf == 2)
        z0 = (z0u[i + j * nx] + z0u[i + 1 + j * nx] + z0v[i + j * nx] + z0v[i + (j + 1) * nx]) / 4.0;
      if (mode == 1) {
        if (p <= 1) {
          const double den = 0.5 * dz_array[k] * log(0.5 * dz_array[k] / z0);
          Gxz[c] = 0.5 * (u[f] + u[f + 1]) / den;
          Gyz[c] = 0.5 * (v[f] + v[f + fy]) / den;
        } else if (p <= 3) {
          const double den = 0.5 * dx * log(0.5 * dx / z0);
          Gyx[c] = 0.5 * (v[f] + v[f + fy]) / den;
          Gzx[c] = 0.5 * (w[f] + w[f + fz]) / den;
        } else {
          const double den = 0.5 * dy * log(0.5 * dy / z0);
          Gxy[c] = 0.5 * (u[f] + u[f + 1]) / den;
          Gzy[c] = 0.5 * (w[f] + w[f + fz]) / den;
        }
      } else if (p == 0) { /* wall below: forward difference up */
        Gxz[c] = (0.5 * (u[f + fz] + u[f + 1 + fz]) - 0.5 * (u[f] + u[f + 1])) / (z[k + 1] - z[k]);
        Gyz[c] = (0.5 * (v[f + fz] + v[f + fy + fz]) - 0.5 * (v[f] + v[f + fy])) / (z[k + 1] - z[k]);
      } else if (p == 1) { /* wall above: backward difference */
        Gxz[c] = (0.5 * (u[f] + u[f + 1]) - 0.5 * (u[f - fz] + u[f + 1 - fz])) / (z[k] - z[k - 1]);
        Gyz[c] = (0.5 * (v[f] + v[f + fy]) - 0.5 * (v[f - fz] + v[f + fy - fz])) / (z[k] - z[k - 1]);
      } else if (p == 2) { /* wall in back */
        Gyx[c] = (0.5 * (v[f + 1] + v[f + 1 + fy]) - 0.5 * (v[f] + v[f + fy])) / dx;
        Gzx[c] = (0.5 * (w[f + 1] + w[f + 1 + fz]) - 0.5 * (w[f] + w[f + fz])) / dx;
      } else if (p == 3) { /* wall in front */
        Gyx[c] = (0.5 * (v[f] + v[f + fy]) - 0.5 * (v[f - 1] + v[f - 1 + fy])) / dx;
        Gzx[c] = (0.5 * (w[f] + w[f + fz]) - 0.5 * (w[f - 1] + w[f - 1 + fz])) / dx;
      } else if (p == 4) { /* wall on the right (j-1) */
        Gxy[c] = (0.5 * (u[f + fy] + u[f + 1 + fy]) - 0.5 * (u[f] + u[f + 1])) / dy;
        Gzy[c] = (0.5 * (w[f + fy] + w[f + fy + fz]) - 0.5 * (w[f] + w[f + fz])) / dy;
      } else { /* wall on the left (j+1) */
        Gxy[c] = (0.5 * (u[f] + u[f + 1]) - 0.5 * (u[f - fy] + u[f + 1 - fy])) / dy;
        Gzy[c] = (0.5 * (w[f] + w[f + fz]) - 0.5 * (w[f - fy] + w[f - fy + fz])) / dy;
      }
    }
  }
  for (int p = 0; p < 6; ++p) {
    free(lc[p]);
    free(lf[p]);
  }
}

/* Returns -1 (nothing done) when a cut cell (flag 7/8) exists or a mode is
 * not 0, 1, 2; z0_domain_u / z0_domain_v (nx*ny each) are read only when the
 * terrain mode is 1. */
int oracle_turb_wall_stairstep(int nx, int ny, int nz, float dx, const float *dz_array, float z0,
                               const float *z0_domain_u, const float *z0_domain_v, int building_mode,
                               int terrain_mode, const int *icellflag, const float *u, const float *v, const float *w,
                               float *G)
{
  const long ncell = (long)(nx - 1) * (ny - 1) * (nz - 1);
  for (long c = 0; c < ncell; ++c)
    if (icellflag[c] == 7 || icellflag[c] == 8) return -1;
  if (building_mode < 0 || building_mode > 2 || terrain_mode < 0 || terrain_mode > 2) return -1;
  float *z_face = (float *)malloc(sizeof(float) * nz), *z = (float *)malloc(sizeof(float) * (nz - 1));
  oracle_vertical_grid(nz, dz_array, z_face, z);
  turb_wall_object(nx, ny, nz, dx, dz_array, z, 0, building_mode, z0, z0_domain_u, z0_domain_v, icellflag, u, v,
                   w, G);
  turb_wall_object(nx, ny, nz, dx, dz_array, z, 2, terrain_mode, z0, z0_domain_u, z0_domain_v, icellflag, u, v, w,
                   G);
  free(z_face);
  free(z);
  return 0;
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void oracle_set_num_threads(int t)
{
#ifdef _OPENMP
  omp_set_num_threads(t);
#else
  (void)t;
#endif
}

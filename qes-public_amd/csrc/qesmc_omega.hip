// Over-relaxation factor for time to tolerance: an estimate of the Jacobi
// spectral radius of the resident problem and the SOR factor it implies.
//
// Red-black SOR (a consistently ordered iteration) converges fastest at
// omega_opt = 2 / (1 + sqrt(1 - rho_J^2)), rho_J the spectral radius of the
// Jacobi iteration matrix D^-1 C (Young's theorem).  The reference runs at a
// fixed omega = 1.78 (Solver.h:83), which on large domains is far from it:
// with rho_J close to 1 the slowest, domain-scale modes then decay thousands
// of times slower per iteration.  No counterpart in the reference.
//
// The operator.  The cells the sweep relaxes that are fluid (flag not 0 or 2)
// form the active set; everything else holds lambda fixed (the Dirichlet ring
// and top) or is a building cell, which no fluid cell reads (its coefficient
// toward a solid is zero, Wall.cpp:288-340).  On the active set C couples a
// cell to its six neighbours with e..n and, at k = 1, to its own mirror
// image (lambda(i,j,0) = lambda(i,j,1), Solver_CPU_RB.cpp:154-162): a
// diagonal term n.  Weighted by W = dz_k, W C is symmetric (dz_k m_k =
// 2 / (dz_k + dz_k+1) = dz_k+1 n_k+1; the horizontal coefficients are equal
// pairwise among fluid cells), so M = (D W)^-1/2 (W C) (D W)^-1/2 is a
// symmetric matrix similar to D^-1 C, and rho_J = lambda_max(M) (the largest
// eigenvalue; C is non-negative).
//
// The estimate.  Lanczos on M in FP64 from a smooth start (the product of
// the lowest Dirichlet modes in x and y and a vertical half cosine, which is
// close to the dominant eigenvector, so few steps resolve it), dot products
// reduced per work-group and summed on the host in a fixed order, then the
// largest eigenvalue of the tridiagonal Lanczos matrix by bisection on its
// Sturm sequence.  The Ritz value never exceeds lambda_max.
#include <cmath>
#include <vector>

#include "qesmc_internal.h"

namespace {

using qesmc::Dims;

#define OCHK(expr)                                         \
  do {                                                     \
    hipError_t _e = (expr);                                \
    if (_e != hipSuccess) return qesmc::map_hip_error(_e); \
  } while (0)

constexpr int kDotBlocks = 1024;

__device__ inline bool active_cell(const Dims &d, int i, int j, int k, int flag)
{
  return i >= 1 && i <= d.NX - 2 && j >= 1 && j <= d.NY - 2 && k >= 1 && k <= d.NZ - 2 && flag != 0 && flag != 2;
}

// s[c] = sqrt(D_c W_c) on active cells, 0 elsewhere; v[c] = the smooth start.
__global__ __launch_bounds__(256) void k_omega_init(Dims d, const float *__restrict__ e, const float *__restrict__ f,
                                                    const float *__restrict__ g, const float *__restrict__ h,
                                                    const float *__restrict__ m, const float *__restrict__ n,
                                                    const int *__restrict__ flag, const float *__restrict__ dz,
                                                    double *__restrict__ s, double *__restrict__ v)
{
  const long long total = (long long)d.NX * d.NY * d.NZ;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const double pi = 3.14159265358979323846;
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < total; c += stride) {
    const int k = (int)(c / d.sz);
    const long long p = c - (long long)k * d.sz;
    const int j = (int)(p / d.NX);
    const int i = (int)(p - (long long)j * d.NX);
    if (!active_cell(d, i, j, k, flag[c])) {
      s[c] = 0.0;
      v[c] = 0.0;
      continue;
    }
    const double D = (double)e[c] + f[c] + g[c] + h[c] + m[c] + n[c];
    s[c] = sqrt(D * (double)dz[k]);
    v[c] = sin(pi * i / (d.NX - 1)) * sin(pi * j / (d.NY - 1)) * cos(0.5 * pi * (k - 0.5) / (d.NZ - 1.5));
  }
}

// w = M v on active cells (0 elsewhere); partial[b] / partial[kDotBlocks + b]
// = this block's sums of w.v and w.w.
__global__ __launch_bounds__(256) void k_omega_apply(Dims d, const float *__restrict__ e, const float *__restrict__ f,
                                                     const float *__restrict__ g, const float *__restrict__ h,
                                                     const float *__restrict__ m, const float *__restrict__ n,
                                                     const float *__restrict__ dz, const double *__restrict__ s,
                                                     const double *__restrict__ v, double *__restrict__ w,
                                                     double *__restrict__ partial)
{
  const long long total = (long long)d.NX * d.NY * d.NZ;
  const long long stride = (long long)gridDim.x * blockDim.x;
  double acc = 0.0, acc2 = 0.0;
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < total; c += stride) {
    const double sc = s[c];
    if (sc == 0.0) {
      w[c] = 0.0;
      continue;
    }
    const int k = (int)(c / d.sz);
    // y = v / s on the neighbours (0 off the active set, where s = 0 and v = 0)
    auto y = [&](long long q) { return s[q] != 0.0 ? v[q] / s[q] : 0.0; };
    double cy = (double)e[c] * y(c + 1) + (double)f[c] * y(c - 1) + (double)g[c] * y(c + d.sy) +
                (double)h[c] * y(c - d.sy) + (double)m[c] * y(c + d.sz);
    cy += (k == 1) ? (double)n[c] * y(c) : (double)n[c] * y(c - d.sz);  // the k = 0 mirror
    const double wc = (double)dz[k] * cy / sc;
    w[c] = wc;
    acc += wc * v[c];
    acc2 += wc * wc;
  }
  for (int off = 32; off > 0; off >>= 1) {
    acc += __shfl_xor(acc, off, 64);
    acc2 += __shfl_xor(acc2, off, 64);
  }
  __shared__ double red[2][4];
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = acc;
    red[1][threadIdx.x >> 6] = acc2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    partial[kDotBlocks + blockIdx.x] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  }
}

// The step's Lanczos coefficients from the partial sums of k_omega_apply,
// summed in a fixed order by one work-group: hist[2 it] = alpha = w.v,
// hist[2 it + 1] = beta' = sqrt(|w|^2 - alpha^2 - beta^2) (v, u orthonormal,
// u.w = beta, the previous beta'), or 0 when that is not positive (the
// Krylov space is exhausted; the host stops there).
__global__ __launch_bounds__(256) void k_omega_scalars(const double *__restrict__ partial, int it,
                                                       double *__restrict__ hist)
{
  __shared__ double red[2][256];
  double s0 = 0.0, s1 = 0.0;
  for (int q = 0; q < kDotBlocks / 256; ++q) {
    s0 += partial[threadIdx.x * (kDotBlocks / 256) + q];
    s1 += partial[kDotBlocks + threadIdx.x * (kDotBlocks / 256) + q];
  }
  red[0][threadIdx.x] = s0;
  red[1][threadIdx.x] = s1;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      red[0][threadIdx.x] += red[0][threadIdx.x + h];
      red[1][threadIdx.x] += red[1][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double alpha = red[0][0];
    const double beta = it > 0 ? hist[2 * it - 1] : 0.0;
    const double b2 = red[1][0] - alpha * alpha - beta * beta;
    hist[2 * it] = alpha;
    hist[2 * it + 1] = b2 > 0.0 ? sqrt(b2) : 0.0;
  }
}

// The next Lanczos vector: v' = (w - alpha v - beta u) / beta', u' = v, with
// the coefficients of step `it` from hist (nothing when beta' is 0).
__global__ __launch_bounds__(256) void k_omega_next(long long total, int it, const double *__restrict__ hist,
                                                    double *__restrict__ u, double *__restrict__ v,
                                                    const double *__restrict__ w)
{
  const double alpha = hist[2 * it], beta = it > 0 ? hist[2 * it - 1] : 0.0, nb = hist[2 * it + 1];
  if (!(nb > 0.0)) return;
  const double inv = 1.0 / nb;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < total; c += stride) {
    const double vc = v[c];
    v[c] = (w[c] - alpha * vc - beta * u[c]) * inv;
    u[c] = vc;
  }
}

// v *= inv.
__global__ __launch_bounds__(256) void k_omega_scale(long long total, double inv, double *__restrict__ v)
{
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < total; c += stride) v[c] *= inv;
}

// sum of squares partials of v.
__global__ __launch_bounds__(256) void k_omega_norm(long long total, const double *__restrict__ v,
                                                    double *__restrict__ partial)
{
  const long long stride = (long long)gridDim.x * blockDim.x;
  double acc = 0.0;
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < total; c += stride) acc += v[c] * v[c];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

struct Buf {
  void *p = nullptr;
  ~Buf()
  {
    if (p) (void)hipFree(p);
  }
};

// Number of eigenvalues of the symmetric tridiagonal (a, b) below x (Sturm).
int sturm_below(const std::vector<double> &a, const std::vector<double> &b, double x)
{
  int count = 0;
  double q = 1.0;
  for (size_t i = 0; i < a.size(); ++i) {
    const double bb = i ? b[i - 1] * b[i - 1] : 0.0;
    q = (a[i] - x) - (i ? bb / q : 0.0);
    if (q == 0.0) q = -1e-300;
    if (q < 0.0) ++count;
  }
  return count;
}

// Largest eigenvalue of the symmetric tridiagonal (a, b) by bisection.
double tridiag_max_eig(const std::vector<double> &a, const std::vector<double> &b)
{
  double lo = 1e300, hi = -1e300;
  for (size_t i = 0; i < a.size(); ++i) {  // Gershgorin bounds
    const double r = (i ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < a.size() ? std::fabs(b[i]) : 0.0);
    lo = std::fmin(lo, a[i] - r);
    hi = std::fmax(hi, a[i] + r);
  }
  const int n = (int)a.size();
  for (int it = 0; it < 200 && hi - lo > 1e-15 * std::fmax(1.0, std::fabs(hi)); ++it) {
    const double mid = 0.5 * (lo + hi);
    if (sturm_below(a, b, mid) < n) lo = mid;  // some eigenvalue is above mid
    else hi = mid;
  }
  return 0.5 * (lo + hi);
}

}  // namespace

extern "C" {

int qesmc_estimate_omega(qesmc_ctx *x, int steps, double *rho, float *omega)
{
  if (!x || steps < 1 || steps > 100000 || (!rho && !omega)) return QESMC_ERR_ARG;
  if (!x->have_geom) return QESMC_ERR_STATE;
  OCHK(hipSetDevice(x->device));
  const Dims &d = x->d;
  const long long total = x->ncell;
  hipStream_t st = x->stream;
  Buf bs, bv, bu, bw, bp;
  OCHK(hipMalloc(&bs.p, sizeof(double) * (size_t)total));
  OCHK(hipMalloc(&bv.p, sizeof(double) * (size_t)total));
  OCHK(hipMalloc(&bu.p, sizeof(double) * (size_t)total));
  OCHK(hipMalloc(&bw.p, sizeof(double) * (size_t)total));
  OCHK(hipMalloc(&bp.p, 2 * sizeof(double) * kDotBlocks));
  double *s = (double *)bs.p, *v = (double *)bv.p, *u = (double *)bu.p, *w = (double *)bw.p, *part = (double *)bp.p;
  std::vector<double> hp(2 * kDotBlocks);
  // the partial sums of one or two block reductions, summed on the host in a
  // fixed order, so the estimate is deterministic
  auto sums = [&](int count, double *out) -> hipError_t {
    hipError_t e = hipMemcpyAsync(hp.data(), part, sizeof(double) * kDotBlocks * count, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    for (int q = 0; q < count; ++q) {
      double acc = 0.0;
      for (int b = 0; b < kDotBlocks; ++b) acc += hp[(size_t)q * kDotBlocks + b];
      out[q] = acc;
    }
    return e;
  };
  const dim3 grid(kDotBlocks), block(256);
  float *const *cf = x->d_coef;
  hipLaunchKernelGGL(k_omega_init, grid, block, 0, st, d, cf[0], cf[1], cf[2], cf[3], cf[4], cf[5], x->d_flag,
                     x->d_dz, s, v);
  OCHK(hipGetLastError());
  OCHK(hipMemsetAsync(u, 0, sizeof(double) * (size_t)total, st));
  hipLaunchKernelGGL(k_omega_norm, grid, block, 0, st, total, v, part);
  OCHK(hipGetLastError());
  double nrm2 = 0.0;
  OCHK(sums(1, &nrm2));
  if (!(nrm2 > 0.0)) return QESMC_ERR_UNSUPPORTED;  // no active cell
  hipLaunchKernelGGL(k_omega_scale, grid, block, 0, st, total, 1.0 / std::sqrt(nrm2), v);
  OCHK(hipGetLastError());
  // Lanczos with one reduction per step: w = M v gives alpha = w.v and
  // |w|^2 together, the next beta from |w - alpha v - beta u|^2 =
  // |w|^2 - alpha^2 - beta^2.  The coefficients stay on the device; the host
  // reads them every 10 steps to test the largest Ritz value for
  // convergence.
  Buf bh;
  OCHK(hipMalloc(&bh.p, sizeof(double) * 2 * (size_t)steps));
  double *hist = (double *)bh.p;
  std::vector<double> hh(2 * (size_t)steps);
  double last = -1.0, r = 0.0;
  for (int it = 0; it < steps; ++it) {
    hipLaunchKernelGGL(k_omega_apply, grid, block, 0, st, d, cf[0], cf[1], cf[2], cf[3], cf[4], cf[5], x->d_dz, s, v,
                       w, part);
    hipLaunchKernelGGL(k_omega_scalars, dim3(1), block, 0, st, (const double *)part, it, hist);
    hipLaunchKernelGGL(k_omega_next, grid, block, 0, st, total, it, (const double *)hist, u, v, (const double *)w);
    OCHK(hipGetLastError());
    if ((it + 1) % 10 != 0 && it + 1 != steps) continue;
    OCHK(hipMemcpyAsync(hh.data(), hist, sizeof(double) * 2 * (size_t)(it + 1), hipMemcpyDeviceToHost, st));
    OCHK(hipStreamSynchronize(st));
    // the tridiagonal Lanczos matrix up to the first exhausted step
    std::vector<double> a, b;
    bool done = false;
    for (int q = 0; q <= it; ++q) {
      a.push_back(hh[2 * (size_t)q]);
      if (!(hh[2 * (size_t)q + 1] > 0.0)) {
        done = true;
        break;
      }
      if (q < it) b.push_back(hh[2 * (size_t)q + 1]);
    }
    r = tridiag_max_eig(a, b);
    if (done || std::fabs(r - last) <= 1e-12 || it + 1 == steps) break;
    last = r;
  }
  if (rho) *rho = r;
  if (omega) {
    // Young's optimum, moved back from 2 by a fifth of its distance: past the
    // optimum the slowest modes oscillate at omega - 1 and fp32 rounding noise
    // is amplified, so a small margin costs little and keeps clear of both
    // (measured on 128x128x64 cities: 1.95 fastest, 1.97 slower; on the
    // 1000x1000x150 city 1.99 reached 1e-4 m/s, 1.995 did not)
    const double rr = r < 1.0 ? r : 1.0;
    const double opt = 2.0 / (1.0 + std::sqrt(1.0 - rr * rr));
    *omega = (float)(2.0 - 1.2 * (2.0 - opt));
  }
  return QESMC_OK;
}

}  // extern "C"

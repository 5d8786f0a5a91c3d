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
// reduced per work-group and the partials summed in a fixed order, then the
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

constexpr int kDotBlocks = 2048;  // a multiple of kXcds
constexpr int kXcds = 8;

__device__ inline bool active_cell(const Dims &d, int i, int j, int k, int flag)
{
  return i >= 1 && i <= d.NX - 2 && j >= 1 && j <= d.NY - 2 && k >= 1 && k <= d.NZ - 2 && flag != 0 && flag != 2;
}

// The iteration runs in the coordinates y = S^-1 v, S = (D W)^1/2, so that the
// operator needs no square root or division per neighbour: with
// q = W C y, M v = S^-1 q, v.(M v) = q.y and |M v|^2 = q.z for z = q / (D W)
// = S^-1 (M v), and the Lanczos recurrence on v is the same recurrence on y
// (divided through by S).

// inv[c] = 1 / (D_c W_c) on active cells, 0 elsewhere; y[c] = the smooth
// start (the lowest Jacobi mode's shape, which in these coordinates is the
// dominant eigenvector's direction).
__global__ __launch_bounds__(256) void k_omega_init(Dims d, const float *__restrict__ e, const float *__restrict__ f,
                                                    const float *__restrict__ g, const float *__restrict__ h,
                                                    const float *__restrict__ m, const float *__restrict__ n,
                                                    const int *__restrict__ flag, const float *__restrict__ dz,
                                                    double *__restrict__ inv, double *__restrict__ y)
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
      inv[c] = 0.0;
      y[c] = 0.0;
      continue;
    }
    const double D = (double)e[c] + f[c] + g[c] + h[c] + m[c] + n[c];
    inv[c] = 1.0 / (D * (double)dz[k]);
    y[c] = sin(pi * i / (d.NX - 1)) * sin(pi * j / (d.NY - 1)) * cos(0.5 * pi * (k - 0.5) / (d.NZ - 1.5));
  }
}

// z = S^-1 M S y on active cells; partial[b] / partial[kDotBlocks + b] = this
// block's sums of v.(M v) and |M v|^2.  Work-groups take the interior rows
// (j, k) in turn and their threads the row's interior cells; every other
// cell of z is left as it is (0: the buffers start at 0 off the active set
// and the recurrence keeps them there).  With MASKED the coefficients come
// from the wall mask (split layout: cell (i,j,k) at half-index i/2 of row
// (j,k) of colour (i+j+k)&1) and the per-layer wall values, which give the
// resident coefficients bit for bit on a mask-encoded geometry and read 1
// byte per cell instead of 24; otherwise from the six coefficient arrays.
template <bool MASKED>
__global__ __launch_bounds__(256) void k_omega_apply(Dims d, const float *__restrict__ e, const float *__restrict__ f,
                                                     const float *__restrict__ g, const float *__restrict__ h,
                                                     const float *__restrict__ m, const float *__restrict__ n,
                                                     const unsigned char *__restrict__ mask0,
                                                     const unsigned char *__restrict__ mask1,
                                                     const unsigned *__restrict__ canon,
                                                     const float *__restrict__ dz, const double *__restrict__ inv,
                                                     const double *__restrict__ y, double *__restrict__ z,
                                                     double *__restrict__ partial)
{
  const int ny = d.NY - 2, rows = ny > 0 && d.NZ > 2 ? ny * (d.NZ - 2) : 0;
  // the rows split into one contiguous range per XCD (work-groups go to the
  // XCDs in turn), so the neighbouring rows a work-group reads are in its
  // own XCD's L2
  const int xcd = (int)blockIdx.x % kXcds, per = (int)gridDim.x / kXcds;
  const int span = (rows + kXcds - 1) / kXcds, r0 = xcd * span, r1 = min(rows, r0 + span);
  double acc = 0.0, acc2 = 0.0;
  for (int r = r0 + (int)blockIdx.x / kXcds; r < r1; r += per) {
    const int k = 1 + r / ny, j = 1 + r % ny;
    const long long row = (long long)k * d.NY + j;
    const double dzk = (double)dz[k];
    double canon_k[6];
    if (MASKED) {
#pragma unroll
      for (int a = 0; a < 6; ++a) canon_k[a] = (double)__uint_as_float(canon[a * d.NZ + k]);
    }
    for (int i = 1 + (int)threadIdx.x; i <= d.NX - 2; i += 256) {
      const long long c = row * d.NX + i;
      const double ic = inv[c];
      if (ic == 0.0) continue;
      double cf[6];
      if (MASKED) {
        const long long x = row * d.HP + (i >> 1);
        const unsigned mk = ((i + j + k) & 1) ? mask1[x] : mask0[x];
#pragma unroll
        for (int a = 0; a < 6; ++a) cf[a] = ((mk >> a) & 1u) ? canon_k[a] : 0.0;
      } else {
        cf[0] = e[c];
        cf[1] = f[c];
        cf[2] = g[c];
        cf[3] = h[c];
        cf[4] = m[c];
        cf[5] = n[c];
      }
      // y is 0 off the active set, so those neighbours drop out by themselves
      const double yc = y[c];
      double cy =
          cf[0] * y[c + 1] + cf[1] * y[c - 1] + cf[2] * y[c + d.sy] + cf[3] * y[c - d.sy] + cf[4] * y[c + d.sz];
      cy += cf[5] * ((k == 1) ? yc : y[c - d.sz]);  // the k = 0 mirror
      const double q = dzk * cy;
      const double zc = q * ic;
      z[c] = zc;
      acc += q * yc;
      acc2 += q * zc;
    }
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

// The next Lanczos vector, written over z: z <- (z - alpha y - beta yp) /
// beta' (the host then rotates yp <- y, y <- z), with the coefficients of
// step `it` from hist (nothing when beta' is 0).  Two cells per thread; the
// thread after the last pair takes an odd count's last cell.
__global__ __launch_bounds__(256) void k_omega_next(long long total, int it, const double *__restrict__ hist,
                                                    const double *__restrict__ yp, const double *__restrict__ y,
                                                    double *__restrict__ z)
{
  const double alpha = hist[2 * it], beta = it > 0 ? hist[2 * it - 1] : 0.0, nb = hist[2 * it + 1];
  if (!(nb > 0.0)) return;
  const double inb = 1.0 / nb;
  const long long t = blockIdx.x * 256LL + threadIdx.x, pairs = total / 2;
  if (t < pairs) {
    const double2 a = reinterpret_cast<const double2 *>(z)[t], b = reinterpret_cast<const double2 *>(y)[t],
                  c = reinterpret_cast<const double2 *>(yp)[t];
    reinterpret_cast<double2 *>(z)[t] =
        make_double2((a.x - alpha * b.x - beta * c.x) * inb, (a.y - alpha * b.y - beta * c.y) * inb);
  } else if (t == pairs && (total & 1)) {
    z[total - 1] = (z[total - 1] - alpha * y[total - 1] - beta * yp[total - 1]) * inb;
  }
}

// v *= inv.
__global__ __launch_bounds__(256) void k_omega_scale(long long total, double inv, double *__restrict__ v)
{
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < total; c += stride) v[c] *= inv;
}

// partials of |v|^2 = sum y^2 D W.
__global__ __launch_bounds__(256) void k_omega_norm(long long total, const double *__restrict__ inv,
                                                    const double *__restrict__ y, double *__restrict__ partial)
{
  const long long stride = (long long)gridDim.x * blockDim.x;
  double acc = 0.0;
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < total; c += stride)
    if (inv[c] != 0.0) acc += y[c] * y[c] / inv[c];
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
  // inv = 1 / (D W); y, the previous vector yp and z rotate through three
  // buffers
  double *inv = (double *)bs.p, *y = (double *)bv.p, *yp = (double *)bu.p, *z = (double *)bw.p,
         *part = (double *)bp.p;
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
  const dim3 grid(kDotBlocks), block(256), next_grid((unsigned)((total / 2 + 256) / 256));
  float *const *cf = x->d_coef;
  // the wall-mask encoding built when the geometry was committed, when the
  // coefficients have one
  const bool masked = x->compressible;
  hipLaunchKernelGGL(k_omega_init, grid, block, 0, st, d, cf[0], cf[1], cf[2], cf[3], cf[4], cf[5], x->d_flag,
                     x->d_dz, inv, y);
  OCHK(hipGetLastError());
  OCHK(hipMemsetAsync(yp, 0, sizeof(double) * (size_t)total, st));
  OCHK(hipMemsetAsync(z, 0, sizeof(double) * (size_t)total, st));
  hipLaunchKernelGGL(k_omega_norm, grid, block, 0, st, total, (const double *)inv, (const double *)y, part);
  OCHK(hipGetLastError());
  double nrm2 = 0.0;
  OCHK(sums(1, &nrm2));
  if (!(nrm2 > 0.0)) return QESMC_ERR_UNSUPPORTED;  // no active cell
  hipLaunchKernelGGL(k_omega_scale, grid, block, 0, st, total, 1.0 / std::sqrt(nrm2), y);
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
    if (masked)
      hipLaunchKernelGGL(k_omega_apply<true>, grid, block, 0, st, d, cf[0], cf[1], cf[2], cf[3], cf[4], cf[5],
                         x->d_mask[0], x->d_mask[1], x->d_canon, x->d_dz, (const double *)inv, (const double *)y,
                         z, part);
    else
      hipLaunchKernelGGL(k_omega_apply<false>, grid, block, 0, st, d, cf[0], cf[1], cf[2], cf[3], cf[4], cf[5],
                         nullptr, nullptr, nullptr, x->d_dz, (const double *)inv, (const double *)y, z, part);
    hipLaunchKernelGGL(k_omega_scalars, dim3(1), block, 0, st, (const double *)part, it, hist);
    hipLaunchKernelGGL(k_omega_next, next_grid, block, 0, st, total, it, (const double *)hist,
                       (const double *)yp, (const double *)y, z);
    OCHK(hipGetLastError());
    double *const old = yp;
    yp = y;
    y = z;
    z = old;
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

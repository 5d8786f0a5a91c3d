// One step of mixed-precision iterative refinement of the resident lambda
// (qesmc_refine).
//
// The solver keeps lambda in fp32, as the reference does.  Near the fixed
// point every relaxation rounds to a few ulp of |lambda| (about 1e-5 at
// |lambda| ~ 100), and SOR amplifies that noise the more, the closer omega is
// to 2: on large grids, where the best omega is within 1e-2 of 2, the noise
// alone holds the velocity field about 1e-4 m/s away from the exact solution
// of the discrete problem.  Refinement removes that floor.  The residual of
// the current lambda is formed in FP64 (exact inputs, so its only error is
// the final rounding to fp32 of a small number), the same red-black SOR
// solves for the correction delta from zero (|delta| is of the size of the
// remaining error, so its own rounding noise is negligible), and lambda +
// delta is rounded once.
//
// In the SOR's own form, cell c relaxes to (sum_nb c_nb lambda_nb - r_c) / D
// (src/winds/Solver_CPU_RB.cpp:117-124), so lambda = lambda1 + delta with
// D delta = sum_nb c_nb delta_nb - r'_c, r'_c = r_c + D lambda1_c -
// sum_nb c_nb lambda1_nb: the correction solve is the ordinary solve with
// R replaced by r' and lambda started at zero.
#include <climits>

#include "qesmc_internal.h"

hipError_t qesmc_launch_fill_u32(qesmc_ctx *x, unsigned *p, long long count, unsigned v);

namespace {

using qesmc::Dims;

#define RCHK(expr)                                         \
  do {                                                     \
    hipError_t _e = (expr);                                \
    if (_e != hipSuccess) return qesmc::map_hip_error(_e); \
  } while (0)

// r' of every relaxed cell of colour `colour` from its R (Rin), in FP64,
// into Rout (which may be Rin); own/oth are the two colours of lambda1 on
// the split layout (cell (i,j,k) of colour c at x = (k*NY + j)*HP + i/2 of
// array c).  Coefficients as the
// sweep uses them: the wall mask and the per-layer wall values, or the full
// split arrays.  The k = 1 cells read plane 0 as stored (the k = 0 mirror).
template <bool MASKED>
__global__ __launch_bounds__(256) void k_refine_residual(Dims d, int colour, const float *__restrict__ own,
                                                         const float *__restrict__ oth, const float *Rin,
                                                         float *Rout, const unsigned char *__restrict__ mask,
                                                         const unsigned *__restrict__ canon,
                                                         const float *__restrict__ ce, const float *__restrict__ cf,
                                                         const float *__restrict__ cg, const float *__restrict__ ch,
                                                         const float *__restrict__ cm, const float *__restrict__ cn)
{
  const long long n = d.splane * d.NZ;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long x = blockIdx.x * 256LL + threadIdx.x; x < n; x += stride) {
    const int k = (int)(x / d.splane);
    const long long rem = x - (long long)k * d.splane;
    const int j = (int)(rem / d.HP);
    const int q = (int)(rem - (long long)j * d.HP);
    const int s = (j + k + colour) & 1;
    const int i = 2 * q + s;
    if (k < 1 || k > d.NZ - 2 || j < 1 || j > d.NY - 2 || i < 1 || i > d.NX - 2) continue;
    double c[6];
    if (MASKED) {
      const unsigned m = mask[x];
#pragma unroll
      for (int a = 0; a < 6; ++a) c[a] = (m >> a) & 1u ? (double)__uint_as_float(canon[a * d.NZ + k]) : 0.0;
    } else {
      c[0] = ce[x];
      c[1] = cf[x];
      c[2] = cg[x];
      c[3] = ch[x];
      c[4] = cm[x];
      c[5] = cn[x];
    }
    // x-neighbours: i+1 and i-1 sit at half-index q+s and q-1+s of the other colour
    const double xp = oth[x + s], xm = oth[x - 1 + s];
    const double sum = c[0] * xp + c[1] * xm + c[2] * (double)oth[x + d.HP] + c[3] * (double)oth[x - d.HP] +
                       c[4] * (double)oth[x + d.splane] + c[5] * (double)oth[x - d.splane];
    const double D = ((c[0] + c[1]) + (c[2] + c[3])) + (c[4] + c[5]);
    Rout[x] = (float)((double)Rin[x] + D * (double)own[x] - sum);
  }
}

// dst += src over one colour array (every position: the ring and the
// planes that are never relaxed hold zero in the correction).  With cur the
// buffer set *cur selects (dst1 / src1 for set 1) stands for the one the
// two-colour march wrote last; without it set 0.
__global__ __launch_bounds__(256) void k_refine_add(long long n, float *dst0, float *dst1, const float *src0,
                                                    const float *src1, const unsigned *__restrict__ cur)
{
  const bool one = cur != nullptr && *cur != 0u;
  float *__restrict__ dst = one ? dst1 : dst0;
  const float *__restrict__ src = one ? src1 : src0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long x = blockIdx.x * 256LL + threadIdx.x; x < n; x += stride) dst[x] = src[x] + dst[x];
}

// The k = 0 mirror of a base lambda between refreshes: ghost (i0,j,0) of
// colour c (half-index q of plane 0 of array c) takes lambda(i0,j,1), the
// same half-index of plane 1 of the other colour, for the columns the
// sweep mirrors (1 <= i0 <= NX-2, 1 <= j <= NY-2).  The sweep writes each
// ghost one pass before its reader, with the other colour as it stood then;
// after a refresh the correction's ghosts restart from zero, so the base's
// must hold the mirror of the base itself.
__global__ __launch_bounds__(256) void k_refine_mirror(Dims d, int colour, float *__restrict__ own,
                                                       const float *__restrict__ oth)
{
  const long long n = d.splane;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long x = blockIdx.x * 256LL + threadIdx.x; x < n; x += stride) {
    const int j = (int)(x / d.HP);
    const int q = (int)(x - (long long)j * d.HP);
    const int i0 = 2 * q + ((j + colour) & 1);
    if (j < 1 || j > d.NY - 2 || i0 < 1 || i0 > d.NX - 2) continue;
    own[x] = oth[x + d.splane];
  }
}

unsigned grid_for(long long n)
{
  const long long b = (n + 255) / 256;
  return (unsigned)(b < 8192 ? (b < 1 ? 1 : b) : 8192);
}

hipError_t residual(qesmc_ctx *x, bool masked, float *const *base, float *const *Rin)
{
  const unsigned g = grid_for(x->nsplit);
  for (int c = 0; c < 2; ++c) {
    const float *const *cc = x->d_cc[c];
    if (masked)
      hipLaunchKernelGGL(k_refine_residual<true>, dim3(g), dim3(256), 0, x->stream, x->d, c, base[c], base[c ^ 1],
                         Rin[c], x->d_Rc[c], x->d_mask[c], x->d_canon, nullptr, nullptr, nullptr, nullptr, nullptr,
                         nullptr);
    else
      hipLaunchKernelGGL(k_refine_residual<false>, dim3(g), dim3(256), 0, x->stream, x->d, c, base[c], base[c ^ 1],
                         Rin[c], x->d_Rc[c], nullptr, nullptr, cc[0], cc[1], cc[2], cc[3], cc[4], cc[5]);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// the correction = 0: d_lc, and with `fused` both of the march's buffer sets
hipError_t zero_correction(qesmc_ctx *x, bool fused)
{
  for (int set = 0; set < (fused ? 2 : 1); ++set)
    for (int c = 0; c < 2; ++c) {
      float *a = set ? x->d_lc2[c] : x->d_lc[c];
      const hipError_t e = qesmc_launch_fill_u32(x, reinterpret_cast<unsigned *>(a), x->nsplit, 0u);
      if (e != hipSuccess) return e;
    }
  return hipSuccess;
}

// base += correction (to_base) or correction += base, over both colour
// arrays; the correction is the march's latest buffer set with `fused`
hipError_t add(qesmc_ctx *x, bool to_base, bool fused)
{
  const unsigned g = grid_for(x->nsplit);
  const unsigned *cur = fused ? x->d_cur : nullptr;
  for (int c = 0; c < 2; ++c) {
    float *d0 = x->d_lc[c], *d1 = fused ? x->d_lc2[c] : x->d_lc[c];
    if (to_base)
      hipLaunchKernelGGL(k_refine_add, dim3(g), dim3(256), 0, x->stream, x->nsplit, x->d_base[c], x->d_base[c],
                         (const float *)d0, (const float *)d1, cur);
    else
      hipLaunchKernelGGL(k_refine_add, dim3(g), dim3(256), 0, x->stream, x->nsplit, d0, d1,
                         (const float *)x->d_base[c], (const float *)x->d_base[c], cur);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace

namespace qesmc {

// After R: base = lambda (d_lc; the march's second set holds the same), R0
// = R, R = residual of base, delta = 0.
hipError_t refine_setup(qesmc_ctx *x, bool masked, bool fused)
{
  const size_t bytes = sizeof(float) * (size_t)x->nsplit;
  for (int c = 0; c < 2; ++c) {
    hipError_t e = hipMemcpyAsync(x->d_R0[c], x->d_Rc[c], bytes, hipMemcpyDeviceToDevice, x->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(x->d_base[c], x->d_lc[c], bytes, hipMemcpyDeviceToDevice, x->stream);
    if (e != hipSuccess) return e;
  }
  hipError_t e = residual(x, masked, x->d_base, x->d_R0);
  return e == hipSuccess ? zero_correction(x, fused) : e;
}

// Between iterations: base += delta, the base's k = 0 mirror, R = residual
// of base, delta = 0.
hipError_t refine_refresh(qesmc_ctx *x, bool masked, bool fused)
{
  hipError_t e = add(x, true, fused);
  if (e != hipSuccess) return e;
  const unsigned g = grid_for(x->d.splane);
  for (int c = 0; c < 2; ++c) {
    hipLaunchKernelGGL(k_refine_mirror, dim3(g), dim3(256), 0, x->stream, x->d, c, x->d_base[c],
                       (const float *)x->d_base[c ^ 1]);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  e = residual(x, masked, x->d_base, x->d_R0);
  return e == hipSuccess ? zero_correction(x, fused) : e;
}

// lambda = base + delta, in the set that holds delta (the one the unpacking
// reads).
hipError_t refine_fold_final(qesmc_ctx *x, bool fused) { return add(x, false, fused); }

}  // namespace qesmc

extern "C" {

int qesmc_refine(qesmc_ctx *x, const qesmc_params *pp, qesmc_result *res)
{
  if (!x || !pp) return QESMC_ERR_ARG;
  qesmc_params p = *pp;
  if (p.path == QESMC_PATH_PLAIN) return QESMC_ERR_UNSUPPORTED;
  if (p.refine_every < 0) return QESMC_ERR_ARG;
  // the resident lambda is the base; no refresh inside unless asked for
  if (p.refine_every == 0) p.refine_every = INT_MAX;
  return qesmc::run_solve(x, &p, res);
}

}  // extern "C"

// Ensembles on one geometry: lambda as a linear combination of resident
// solutions, and the L_inf distance between two resident velocity fields.
//
// The discrete mass-consistency problem is linear: R is linear in (u0, v0, w0)
// (Solver_CPU_RB.cpp:56-80), the SOR fixed point is linear in R, and a
// single-sensor inflow from direction theta is cos(theta) times the inflow
// along x plus sin(theta) times the inflow along y (WindProfilerSensorType.cpp
// :571-621, the profile is scaled by cos/sin of the direction).  A combination
// of converged basis solutions is therefore an initial lambda for any other
// direction that is already close to its solution; SOR still runs to the same
// stop rule from there.
#include "qesmc_internal.h"

namespace {

#define ECHK(expr)                                         \
  do {                                                     \
    hipError_t _e = (expr);                                \
    if (_e != hipSuccess) return qesmc::map_hip_error(_e); \
  } while (0)

constexpr int kMaxTerms = 8;

struct Terms {
  const float *src[kMaxTerms];
  float coef[kMaxTerms];
  int n;
};

// dst = sum_t coef[t] * src[t], accumulated in float in term order.
__global__ __launch_bounds__(256) void k_combine(Terms T, long long count, float *__restrict__ dst)
{
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < count; c += stride) {
    float acc = T.coef[0] * T.src[0][c];
    for (int t = 1; t < T.n; ++t) acc = acc + T.coef[t] * T.src[t][c];
    dst[c] = acc;
  }
}

// max |a - b| over u, v, w in double (a NaN difference counts as +inf); one
// 64-bit atomic per work-group on the bit pattern (monotonic for
// non-negative doubles).
__global__ __launch_bounds__(256) void k_vel_diff(const float *__restrict__ a0, const float *__restrict__ a1,
                                                  const float *__restrict__ a2, const float *__restrict__ b0,
                                                  const float *__restrict__ b1, const float *__restrict__ b2,
                                                  long long count, unsigned long long *__restrict__ out)
{
  double m = 0.0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < count; c += stride) {
    const double d0 = fabs((double)a0[c] - (double)b0[c]);
    const double d1 = fabs((double)a1[c] - (double)b1[c]);
    const double d2 = fabs((double)a2[c] - (double)b2[c]);
    const double d = fmax(fmax(d0, d1), d2);
    m = (d != d || d0 != d0 || d1 != d1 || d2 != d2) ? __builtin_huge_val() : fmax(m, d);
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  __shared__ double wmax[4];
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double b = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
    atomicMax(out, (unsigned long long)__double_as_longlong(b));
  }
}

// out[q] = u[idx[q]], out[n + q] = v[idx[q]], out[2n + q] = w[idx[q]].
__global__ __launch_bounds__(256) void k_probe(const float *__restrict__ u, const float *__restrict__ v,
                                               const float *__restrict__ w, const long long *__restrict__ idx,
                                               long long n, float *__restrict__ out)
{
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += stride) {
    const long long f = idx[q];
    out[q] = u[f];
    out[n + q] = v[f];
    out[2 * n + q] = w[f];
  }
}

bool same_grid(const qesmc_ctx *a, const qesmc_ctx *b)
{
  return a->d.nx == b->d.nx && a->d.ny == b->d.ny && a->d.nz == b->d.nz;
}

}  // namespace

extern "C" {

int qesmc_lambda_combine(qesmc_ctx *x, int n, qesmc_ctx *const *srcs, const float *coefs)
{
  if (!x || n < 1 || n > kMaxTerms || !srcs || !coefs) return QESMC_ERR_ARG;
  if (!x->d_lam) return QESMC_ERR_STATE;
  Terms T{};
  T.n = n;
  for (int t = 0; t < n; ++t) {
    if (!srcs[t] || srcs[t]->device != x->device || !same_grid(srcs[t], x)) return QESMC_ERR_ARG;
    if (!srcs[t]->d_lam) return QESMC_ERR_STATE;
    T.src[t] = srcs[t]->d_lam;
    T.coef[t] = coefs[t];
  }
  ECHK(hipSetDevice(x->device));
  // the sources' own work (a run still writing lambda) completes first
  for (int t = 0; t < n; ++t)
    if (srcs[t] != x) ECHK(hipStreamSynchronize(srcs[t]->stream));
  long long blocks = (x->ncell + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_combine, dim3((unsigned)blocks), dim3(256), 0, x->stream, T, x->ncell, x->d_lam);
  ECHK(hipGetLastError());
  ECHK(hipStreamSynchronize(x->stream));
  return QESMC_OK;
}

int qesmc_velocity_linf_diff(qesmc_ctx *a, qesmc_ctx *b, double *out)
{
  if (!a || !b || !out) return QESMC_ERR_ARG;
  if (a->device != b->device || !same_grid(a, b)) return QESMC_ERR_ARG;
  if (!a->d_u || !b->d_u) return QESMC_ERR_STATE;
  ECHK(hipSetDevice(a->device));
  if (b != a) ECHK(hipStreamSynchronize(b->stream));
  ECHK(hipMemsetAsync(a->d_norm, 0, sizeof(double), a->stream));
  long long blocks = (a->nface + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_vel_diff, dim3((unsigned)blocks), dim3(256), 0, a->stream, a->d_u, a->d_v, a->d_w, b->d_u,
                     b->d_v, b->d_w, a->nface, (unsigned long long *)a->d_norm);
  ECHK(hipGetLastError());
  ECHK(hipMemcpyAsync(out, a->d_norm, sizeof(double), hipMemcpyDeviceToHost, a->stream));
  ECHK(hipStreamSynchronize(a->stream));
  return QESMC_OK;
}

int qesmc_set_probes(qesmc_ctx *x, long long n, const long long *faces)
{
  if (!x || n < 0 || (n > 0 && !faces)) return QESMC_ERR_ARG;
  if (!x->d_u) return QESMC_ERR_STATE;
  for (long long q = 0; q < n; ++q)
    if (faces[q] < 0 || faces[q] >= x->nface) return QESMC_ERR_ARG;
  ECHK(hipSetDevice(x->device));
  if (n > x->probe_cap) {
    ECHK(hipStreamSynchronize(x->stream));
    if (x->d_probe) (void)hipFree(x->d_probe);
    if (x->d_probe_val) (void)hipFree(x->d_probe_val);
    x->d_probe = nullptr;
    x->d_probe_val = nullptr;
    x->probe_cap = x->n_probe = 0;
    ECHK(hipMalloc((void **)&x->d_probe, sizeof(long long) * (size_t)n));
    ECHK(hipMalloc((void **)&x->d_probe_val, 3 * sizeof(float) * (size_t)n));
    x->probe_cap = n;
  }
  if (n > 0)
    ECHK(hipMemcpyAsync(x->d_probe, faces, sizeof(long long) * (size_t)n, hipMemcpyHostToDevice, x->stream));
  ECHK(hipStreamSynchronize(x->stream));
  x->n_probe = n;
  x->probe_nface = x->nface;
  return QESMC_OK;
}

int qesmc_get_probes(qesmc_ctx *x, float *u, float *v, float *w)
{
  if (!x) return QESMC_ERR_ARG;
  if (x->n_probe > 0 && (!u || !v || !w)) return QESMC_ERR_ARG;
  if (!x->d_u || x->probe_nface != x->nface) return QESMC_ERR_STATE;
  if (x->n_probe == 0) return QESMC_OK;
  ECHK(hipSetDevice(x->device));
  const long long n = x->n_probe;
  long long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_probe, dim3((unsigned)blocks), dim3(256), 0, x->stream, x->d_u, x->d_v, x->d_w, x->d_probe, n,
                     x->d_probe_val);
  ECHK(hipGetLastError());
  float *dst[3] = {u, v, w};
  for (int a = 0; a < 3; ++a)
    ECHK(hipMemcpyAsync(dst[a], x->d_probe_val + a * n, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, x->stream));
  ECHK(hipStreamSynchronize(x->stream));
  return QESMC_OK;
}

}  // extern "C"

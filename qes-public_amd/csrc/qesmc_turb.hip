// Velocity gradient tensor of the corrected wind field at cell centres: the
// derivative step of QES-Turb, TURBGeneralData::derivativeVelocity
// (src/winds/TURBGeneralData.cpp:479-552; its device counterpart is
// TURBGeneralData.cu:37-306), computed on the resident u, v, w right after
// the solve, without a host round trip.  The wall log-law corrections that
// follow in the reference (:553-556) are not part of this step.
//
// One thread per cell column (i, j), marching kb planes: a gather over the
// 3x3x3 face neighbourhood, whose planes k and k+1 are the previous step's
// k+1 and k+2.  Work-groups are mapped XCD-aware: XCD x (work-group id % 8)
// takes the x-th eighth of each plane's cells, so the rows j-1 and j+1 a
// work-group reads are in the same L2 as the work-groups that own them (with
// consecutive work-groups dealt round-robin over the XCDs, every face row was
// fetched by three XCDs).  Every cell is written, so that the nine outputs
// are fully defined: fluid cells with
// i in [1, nx-3], j in [1, ny-3], k in [1, nz-3] (the icellfluid list,
// TURBGeneralData.cpp:264-283) get the reference's float expressions,
// evaluated in the same order without contraction; every other cell gets 0.
// The stretched-grid weights depend on k only and are evaluated on the host
// with the reference's expressions (same float operations).
#include <vector>

#include "qesmc_city_internal.h"
#include "qesmc_internal.h"

namespace {

using qesmc::Dims;

#define TCHK(expr)                                         \
  do {                                                     \
    hipError_t _e = (expr);                                \
    if (_e != hipSuccess) return qesmc::map_hip_error(_e); \
  } while (0)

// zc[k] = {0.5f (z_k - z_{k-1}) / (z_{k+1} - z_k), 0.5f (z_{k+1} - z_k) / (z_k - z_{k-1}),
//          z_{k+1} - z_{k-1}, dz_array[k]}
__global__ __launch_bounds__(256) void k_vel_grad(Dims d, const int *__restrict__ flag, const float *__restrict__ u,
                                                  const float *__restrict__ v, const float *__restrict__ w,
                                                  const float4 *__restrict__ zc, float dx, float dy, float dz,
                                                  int uniform_z, long long ncell, int kb, float *__restrict__ G)
{
  const long long plane = d.sz;
  const unsigned per = gridDim.x / 8;  // the grid is a multiple of 8 work-groups
  const long long p = (long long)((blockIdx.x % 8) * per + blockIdx.x / 8) * 256 + threadIdx.x;
  if (p >= plane) return;
  const int j = (int)(p / d.NX);
  const int i = (int)(p - (long long)j * d.NX);
  const bool column = i >= 1 && i <= d.nx - 3 && j >= 1 && j <= d.ny - 3;
  const int k0 = blockIdx.y * kb;
  const int k1 = k0 + kb < d.NZ ? k0 + kb : d.NZ;
#pragma unroll 1
  for (int k = k0; k < k1; ++k) {
    const long long c = p + k * plane;
    float gxx = 0.f, gxy = 0.f, gxz = 0.f, gyx = 0.f, gyy = 0.f, gyz = 0.f, gzx = 0.f, gzy = 0.f, gzz = 0.f;
    if (column && k >= 1 && k <= d.nz - 3 && !qesmc::is_solid_flag(flag[c])) {
      const long long nx = d.nx, fz = d.fz;
      const long long f = i + j * nx + k * fz;
      gxx = (u[f + 1] - u[f]) / (dx);
      gyx = ((v[f + 1] + v[f + 1 + nx]) - (v[f - 1] + v[f - 1 + nx])) / (4.0f * dx);
      gzx = ((w[f + 1] + w[f + 1 + fz]) - (w[f - 1] + w[f - 1 + fz])) / (4.0f * dx);
      gxy = ((u[f + nx] + u[f + 1 + nx]) - (u[f - nx] + u[f + 1 - nx])) / (4.0f * dy);
      gyy = (v[f + nx] - v[f]) / (dy);
      gzy = ((w[f + nx] + w[f + nx + fz]) - (w[f - nx] + w[f - nx + fz])) / (4.0f * dy);
      if (uniform_z) {
        gxz = ((u[f + fz] + u[f + 1 + fz]) - (u[f - fz] + u[f + 1 - fz])) / (4.0f * dz);
        gyz = ((v[f + fz] + v[f + nx + fz]) - (v[f - fz] + v[f + nx - fz])) / (4.0f * dz);
        gzz = (w[f + fz] - w[f]) / (dz);
      } else {
        const float4 q = zc[k];
        gxz = (q.x * ((u[f + fz] + u[f + 1 + fz]) - (u[f] + u[f + 1])) +
               q.y * ((u[f] + u[f + 1]) - (u[f - fz] + u[f + 1 - fz]))) /
              q.z;
        gyz = (q.x * ((v[f + fz] + v[f + nx + fz]) - (v[f] + v[f + nx])) +
               q.y * ((v[f] + v[f + nx]) - (v[f - fz] + v[f + nx - fz]))) /
              q.z;
        gzz = (w[f + fz] - w[f]) / (q.w);
      }
    }
    // nine output streams that this kernel never reads back: non-temporal
    // stores keep them from displacing the face rows still to be read
    __builtin_nontemporal_store(gxx, G + c);
    __builtin_nontemporal_store(gxy, G + c + ncell);
    __builtin_nontemporal_store(gxz, G + c + 2 * ncell);
    __builtin_nontemporal_store(gyx, G + c + 3 * ncell);
    __builtin_nontemporal_store(gyy, G + c + 4 * ncell);
    __builtin_nontemporal_store(gyz, G + c + 5 * ncell);
    __builtin_nontemporal_store(gzx, G + c + 6 * ncell);
    __builtin_nontemporal_store(gzy, G + c + 7 * ncell);
    __builtin_nontemporal_store(gzz, G + c + 8 * ncell);
  }
}

struct Tmp {
  void *p = nullptr;
  ~Tmp()
  {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

extern "C" {

int qesmc_velocity_gradient(qesmc_ctx *x, int uniform_z, float dz, float *G, double *ms)
{
  if (!x) return QESMC_ERR_ARG;
  if (uniform_z && !(dz > 0.f)) return QESMC_ERR_ARG;
  if (!x->have_geom || !x->have_velocity) return QESMC_ERR_STATE;
  if (x->d.nz > 65535) return QESMC_ERR_UNSUPPORTED;
  TCHK(hipSetDevice(x->device));
  const qesmc::Dims &d = x->d;
  if (!x->d_grad) TCHK(hipMalloc((void **)&x->d_grad, sizeof(float) * 9 * (size_t)x->ncell));
  // stretched-grid weights per layer, as the reference evaluates them
  std::vector<float> z_face, z;
  qesmc::vertical_grid(d.nz, x->dz_host.data(), z_face, z);
  std::vector<float> zc(4 * (size_t)d.NZ, 0.f);
  for (int k = 1; k + 1 < d.NZ; ++k) {
    zc[4 * k + 0] = 0.5f * (z[k] - z[k - 1]) / (z[k + 1] - z[k]);
    zc[4 * k + 1] = 0.5f * (z[k + 1] - z[k]) / (z[k] - z[k - 1]);
    zc[4 * k + 2] = z[k + 1] - z[k - 1];
    zc[4 * k + 3] = x->dz_host[k];
  }
  Tmp tmp;
  TCHK(hipMalloc(&tmp.p, sizeof(float) * zc.size()));
  TCHK(hipMemcpyAsync(tmp.p, zc.data(), sizeof(float) * zc.size(), hipMemcpyHostToDevice, x->stream));
  TCHK(hipEventRecord(x->ev[0], x->stream));
  const int kb = 8;
  const long long wgs = (d.sz + 255) / 256;
  const dim3 grid((unsigned)((wgs + 7) / 8 * 8), (unsigned)((d.NZ + kb - 1) / kb));
  hipLaunchKernelGGL(k_vel_grad, grid, dim3(256), 0, x->stream, d, x->d_flag, x->d_u, x->d_v, x->d_w,
                     (const float4 *)tmp.p, x->dx, x->dy, dz, uniform_z ? 1 : 0, x->ncell, kb, x->d_grad);
  TCHK(hipGetLastError());
  TCHK(hipEventRecord(x->ev[1], x->stream));
  if (G) TCHK(hipMemcpyAsync(G, x->d_grad, sizeof(float) * 9 * (size_t)x->ncell, hipMemcpyDeviceToHost, x->stream));
  TCHK(hipStreamSynchronize(x->stream));
  if (ms) {
    float t = 0.f;
    TCHK(hipEventElapsedTime(&t, x->ev[0], x->ev[1]));
    *ms = t;
  }
  return QESMC_OK;
}

}  // extern "C"

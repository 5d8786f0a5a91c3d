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

// The wall step of derivativeVelocity (TURBGeneralData.cpp:553-556) on
// stair-step walls: for the building object (flag 0) and then the terrain
// object (flag 2), the velocity wall treatment of its mode -- 1: the log law
// of set_loglaw_stairstep (TURBWallBuilding.cpp:87-169, TURBWallTerrain.cpp:
// 88-...), 2: the one-sided differences of
// comp_velocity_deriv_finitediff_stairstep (TURBWall.cpp:206-300), 0: none.
// The reference walks six index lists per object (TURBWall.cpp:76-118:
// i in [1,nx-2], j in [1,ny-2], k in [1,nz-3], a cell whose flag is not 0 or
// 2 and whose flat neighbour -sz, +sz, -1, +1, -sy, +sy has the object's
// flag) in the order below, above, back, front, right, left, each pass
// overwriting two components; here one thread per cell evaluates the same
// passes in the same order and writes what the last applicable pass of each
// component wrote.  Mode 1's sixth pass walks the cells with a wall below
// (the reference's loop over wall_below_indices there).  Arithmetic as the
// reference's: float sums of two faces, times 0.5 in double, quotient in
// double, stored to float; dy is dx (getDomainSize, Domain.h:104); z0 is
// WGD->z0 for buildings and, for terrain, the mean of the four surrounding
// z0_domain_u / z0_domain_v entries (float sum, divided by 4.0).
struct WallPass {
  float val[6];  // Gxy, Gxz, Gyx, Gyz, Gzx, Gzy
  bool set[6];
};

__device__ __forceinline__ void turb_wall_object(const Dims &d, long long c, long long f, int i, int j, int k,
                                                 const int *__restrict__ flag, int cf, int mode, float z0,
                                                 float dx, const float *__restrict__ dzc, const float *__restrict__ zc,
                                                 const float *__restrict__ u, const float *__restrict__ v,
                                                 const float *__restrict__ w, WallPass &o)
{
  const long long sy = d.NX, sz = (long long)d.NX * d.NY;
  const long long fy = d.nx, fz = (long long)d.nx * d.ny;
  const bool below = flag[c - sz] == cf, above = flag[c + sz] == cf, back = flag[c - 1] == cf,
             front = flag[c + 1] == cf, right = flag[c - sy] == cf, left = flag[c + sy] == cf;
  const float dy = dx;
  enum { XY = 0, XZ = 1, YX = 2, YZ = 3, ZX = 4, ZY = 5 };
  auto put = [&](int a, double x) {
    o.val[a] = (float)x;
    o.set[a] = true;
  };
  if (mode == 1) {
    const double dzk = dzc[k];
    if (below || above) {  // passes 1 and 2 write the same values
      const double den = 0.5 * dzk * log(0.5 * dzk / (double)z0);
      put(XZ, 0.5 * (double)(u[f] + u[f + 1]) / den);
      put(YZ, 0.5 * (double)(v[f] + v[f + fy]) / den);
    }
    if (back || front) {
      const double den = 0.5 * (double)dx * log(0.5 * (double)dx / (double)z0);
      put(YX, 0.5 * (double)(v[f] + v[f + fy]) / den);
      put(ZX, 0.5 * (double)(w[f] + w[f + fz]) / den);
    }
    if (right || below) {  // pass 6 walks the cells with a wall below
      const double den = 0.5 * (double)dy * log(0.5 * (double)dy / (double)z0);
      put(XY, 0.5 * (double)(u[f] + u[f + 1]) / den);
      put(ZY, 0.5 * (double)(w[f] + w[f + fz]) / den);
    }
  } else if (mode == 2) {
    if (above) {  // the later pass wins
      const double h = (double)(zc[k] - zc[k - 1]);
      put(XZ, (0.5 * (double)(u[f] + u[f + 1]) - 0.5 * (double)(u[f - fz] + u[f + 1 - fz])) / h);
      put(YZ, (0.5 * (double)(v[f] + v[f + fy]) - 0.5 * (double)(v[f - fz] + v[f + fy - fz])) / h);
    } else if (below) {
      const double h = (double)(zc[k + 1] - zc[k]);
      put(XZ, (0.5 * (double)(u[f + fz] + u[f + 1 + fz]) - 0.5 * (double)(u[f] + u[f + 1])) / h);
      put(YZ, (0.5 * (double)(v[f + fz] + v[f + fy + fz]) - 0.5 * (double)(v[f] + v[f + fy])) / h);
    }
    if (front) {
      put(YX, (0.5 * (double)(v[f] + v[f + fy]) - 0.5 * (double)(v[f - 1] + v[f - 1 + fy])) / (double)dx);
      put(ZX, (0.5 * (double)(w[f] + w[f + fz]) - 0.5 * (double)(w[f - 1] + w[f - 1 + fz])) / (double)dx);
    } else if (back) {
      put(YX, (0.5 * (double)(v[f + 1] + v[f + 1 + fy]) - 0.5 * (double)(v[f] + v[f + fy])) / (double)dx);
      put(ZX, (0.5 * (double)(w[f + 1] + w[f + 1 + fz]) - 0.5 * (double)(w[f] + w[f + fz])) / (double)dx);
    }
    if (left) {
      put(XY, (0.5 * (double)(u[f] + u[f + 1]) - 0.5 * (double)(u[f - fy] + u[f + 1 - fy])) / (double)dy);
      put(ZY, (0.5 * (double)(w[f] + w[f + fz]) - 0.5 * (double)(w[f - fy] + w[f - fy + fz])) / (double)dy);
    } else if (right) {
      put(XY, (0.5 * (double)(u[f + fy] + u[f + 1 + fy]) - 0.5 * (double)(u[f] + u[f + 1])) / (double)dy);
      put(ZY, (0.5 * (double)(w[f + fy] + w[f + fy + fz]) - 0.5 * (double)(w[f] + w[f + fz])) / (double)dy);
    }
  }
  (void)i;
  (void)j;
}

__global__ __launch_bounds__(256) void k_turb_walls(Dims d, const int *__restrict__ flag,
                                                    const float *__restrict__ u, const float *__restrict__ v,
                                                    const float *__restrict__ w, const float *__restrict__ dzc,
                                                    const float *__restrict__ zc, float dx, int building_mode,
                                                    int terrain_mode, float z0, const float *__restrict__ z0u,
                                                    const float *__restrict__ z0v, long long ncell,
                                                    float *__restrict__ G)
{
  // cells i in [1, NX-1], j in [1, NY-1], k in [1, NZ-2]
  const long long ni = d.NX - 1, nj = d.NY - 1, nk = d.NZ - 2;
  const long long n = ni * nj * nk;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < n; e += stride) {
    const int i = 1 + (int)(e % ni);
    const long long r = e / ni;
    const int j = 1 + (int)(r % nj);
    const int k = 1 + (int)(r / nj);
    const long long c = i + (long long)j * d.NX + (long long)k * d.NX * d.NY;
    const int fl = flag[c];
    if (fl == 0 || fl == 2) continue;
    const long long f = i + (long long)j * d.nx + (long long)k * d.nx * d.ny;
    WallPass o;
#pragma unroll
    for (int a = 0; a < 6; ++a) o.set[a] = false;
    turb_wall_object(d, c, f, i, j, k, flag, 0, building_mode, z0, dx, dzc, zc, u, v, w, o);
    if (terrain_mode != 0) {
      float z0t = z0;
      if (terrain_mode == 1)
        z0t = (z0u[i + j * d.nx] + z0u[i + 1 + j * d.nx] + z0v[i + j * d.nx] + z0v[i + (j + 1) * d.nx]) / 4.0;
      turb_wall_object(d, c, f, i, j, k, flag, 2, terrain_mode, z0t, dx, dzc, zc, u, v, w, o);
    }
    // G blocks: Gxx, Gxy, Gxz, Gyx, Gyy, Gyz, Gzx, Gzy, Gzz
    const int blockof[6] = {1, 2, 3, 5, 6, 7};
#pragma unroll
    for (int a = 0; a < 6; ++a)
      if (o.set[a]) G[blockof[a] * ncell + c] = o.val[a];
  }
}

}  // namespace

extern "C" {

int qesmc_velocity_gradient_walls(qesmc_ctx *x, int building_mode, int terrain_mode, float z0,
                                  const float *z0_domain_u, const float *z0_domain_v, float *G, double *ms)
{
  if (!x) return QESMC_ERR_ARG;
  if (building_mode < 0 || building_mode > 2 || terrain_mode < 0 || terrain_mode > 2) return QESMC_ERR_ARG;
  if (terrain_mode == 1 && (!z0_domain_u || !z0_domain_v)) return QESMC_ERR_ARG;
  if (!x->have_geom || !x->have_velocity || !x->d_grad) return QESMC_ERR_STATE;
  TCHK(hipSetDevice(x->device));
  {
    bool cut = false;
    TCHK(qesmc::find_cut_cells(x, &cut));
    if (cut) return QESMC_ERR_UNSUPPORTED;
  }
  const qesmc::Dims &d = x->d;
  std::vector<float> z_face, z;
  qesmc::vertical_grid(d.nz, x->dz_host.data(), z_face, z);
  const size_t nzz = (size_t)d.NZ, nxy = (size_t)d.nx * d.ny;
  Tmp tz, tdz, tu, tv;
  TCHK(hipMalloc(&tz.p, sizeof(float) * nzz));
  TCHK(hipMalloc(&tdz.p, sizeof(float) * nzz));
  TCHK(hipMemcpyAsync(tz.p, z.data(), sizeof(float) * nzz, hipMemcpyHostToDevice, x->stream));
  TCHK(hipMemcpyAsync(tdz.p, x->dz_host.data(), sizeof(float) * nzz, hipMemcpyHostToDevice, x->stream));
  if (terrain_mode == 1) {
    TCHK(hipMalloc(&tu.p, sizeof(float) * nxy));
    TCHK(hipMalloc(&tv.p, sizeof(float) * nxy));
    TCHK(hipMemcpyAsync(tu.p, z0_domain_u, sizeof(float) * nxy, hipMemcpyHostToDevice, x->stream));
    TCHK(hipMemcpyAsync(tv.p, z0_domain_v, sizeof(float) * nxy, hipMemcpyHostToDevice, x->stream));
  }
  TCHK(hipEventRecord(x->ev[0], x->stream));
  const long long n = (long long)(d.NX - 1) * (d.NY - 1) * (d.NZ - 2);
  if (n > 0 && (building_mode != 0 || terrain_mode != 0)) {
    long long blocks = (n + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(k_turb_walls, dim3((unsigned)blocks), dim3(256), 0, x->stream, d, x->d_flag, x->d_u, x->d_v,
                       x->d_w, (const float *)tdz.p, (const float *)tz.p, x->dx, building_mode, terrain_mode, z0,
                       (const float *)tu.p, (const float *)tv.p, x->ncell, x->d_grad);
    TCHK(hipGetLastError());
  }
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

// Device builder of the solver inputs for flat stair-step cities: the same
// arrays as qesmc_city_build, written straight into a context's device
// buffers, so a large domain (or thousands of training cities) never builds
// and uploads 24 B/cell of coefficients from the host.
//
//   ground ghost flags           WINDSGeneralData.cpp:439-444       k_city_flags
//   rectangular building flags   PolyBuilding.cpp:130-236           k_city_stamp (one work-group per building)
//   wall coefficients            Wall.cpp:44-349, 827-846           k_city_coefs (one thread per cell, gather)
//   inflow extrusion + zeroing   WindProfilerSensorType.cpp:571-621, Wall.cpp:805-824
//                                                                   k_city_inflow (one thread per face, gather)
//
// Per-building polygons, layer denominators and the one-dimensional inflow
// profile (O(buildings + nz) values) are evaluated on the host by the same
// functions the host builder uses; every per-cell and per-face operation runs
// on the device with the host builder's float operations, so the arrays are
// bitwise equal to qesmc_city_build's.
#include <cmath>
#include <vector>

#include "qesmc_city_internal.h"
#include "qesmc_internal.h"

namespace {

using qesmc::Dims;
using qesmc::Footprint;
using qesmc::is_solid_flag;

#define CCHK(expr)                                         \
  do {                                                     \
    hipError_t _e = (expr);                                \
    if (_e != hipSuccess) return qesmc::map_hip_error(_e); \
  } while (0)

// allocateMemory: every cell fluid, the ground ghost layer k = 0 terrain.
__global__ __launch_bounds__(256) void k_city_flags(Dims d, int *__restrict__ flag)
{
  const long long plane = d.sz;
  const long long p = blockIdx.x * 256LL + threadIdx.x;
  if (p >= plane) return;
  const int k = blockIdx.y;
  flag[p + k * plane] = (k == 0) ? 2 : 1;
}

// One work-group per building: every (i, j) of its clamped search box whose
// centre lies inside the polygon becomes solid over [k0, k1).  Overlapping
// buildings write the same value, so the order between work-groups does not
// matter.
__global__ __launch_bounds__(256) void k_city_stamp(Dims d, const Footprint *__restrict__ fps, float dx,
                                                    int *__restrict__ flag)
{
  const Footprint F = fps[blockIdx.x];
  const int ni = F.i1 - F.i0 + 1, nj = F.j1 - F.j0 + 1;
  const long long tot = (long long)ni * nj;
  for (long long q = threadIdx.x; q < tot; q += blockDim.x) {
    const int j = F.j0 + (int)(q / ni);
    const int i = F.i0 + (int)(q % ni);
    const float yc = (float)((j + 0.5) * (double)dx);
    const float xc = (float)((i + 0.5) * (double)dx);
    if (!qesmc::footprint_contains(F, xc, yc)) continue;
    for (int k = F.k0; k < F.k1; ++k) {
      int *fl = flag + i + j * d.sy + k * d.sz;
      if (*fl != 7) *fl = 0;
    }
  }
}

// Wall::defineWalls + Wall::solverCoefficients as a gather: a fluid cell in
// k in [1, nz-3] has the coefficient toward each solid neighbour zeroed
// (east / north by flat index +1 / +(nx-1), which wraps at the last column /
// row; west / south only for i > 0 / j > 0), then layers [1, nz-3] are
// divided by dx^2 and the per-layer vertical denominators.  Other layers
// keep 1.  lay[2k] / lay[2k+1] hold the m / n denominators of layer k.
__global__ __launch_bounds__(256) void k_city_coefs(Dims d, const int *__restrict__ flag,
                                                    const float *__restrict__ lay, float dxx, float *__restrict__ e,
                                                    float *__restrict__ f, float *__restrict__ g,
                                                    float *__restrict__ h, float *__restrict__ m,
                                                    float *__restrict__ n)
{
  const long long plane = d.sz;
  const long long p = blockIdx.x * 256LL + threadIdx.x;
  if (p >= plane) return;
  const int k = blockIdx.y;
  const int j = (int)(p / d.NX);
  const int i = (int)(p - (long long)j * d.NX);
  const long long c = p + k * plane;
  float ce = 1.0f, cf = 1.0f, cg = 1.0f, ch = 1.0f, cm = 1.0f, cn = 1.0f;
  if (k >= 1 && k <= d.nz - 3) {
    if (!is_solid_flag(flag[c])) {
      if (is_solid_flag(flag[c - d.sz])) cn = 0.0f;
      if (is_solid_flag(flag[c + d.sz])) cm = 0.0f;
      if (i > 0 && is_solid_flag(flag[c - 1])) cf = 0.0f;
      if (is_solid_flag(flag[c + 1])) ce = 0.0f;
      if (j > 0 && is_solid_flag(flag[c - d.sy])) ch = 0.0f;
      if (is_solid_flag(flag[c + d.sy])) cg = 0.0f;
    }
    ce = ce / dxx;
    cf = cf / dxx;
    cg = cg / dxx;
    ch = ch / dxx;
    cm = cm / lay[2 * k];
    cn = cn / lay[2 * k + 1];
  }
  e[c] = ce;
  f[c] = cf;
  g[c] = cg;
  h[c] = ch;
  m[c] = cm;
  n[c] = cn;
}

// Profile extrusion (faces k in [1, nz-2] take the sensor profile, w0 = 0)
// and Wall::setVelocityZero as a gather: a solid cell (i, j, k) with
// i in [1, nx-2], j in [1, ny-2] zeroes its west/east u, south/north v and
// bottom/top w faces.
__global__ __launch_bounds__(256) void k_city_inflow(Dims d, const int *__restrict__ flag,
                                                     const float *__restrict__ up, const float *__restrict__ vp,
                                                     float *__restrict__ u0, float *__restrict__ v0,
                                                     float *__restrict__ w0)
{
  const long long plane = d.fz;
  const long long p = blockIdx.x * 256LL + threadIdx.x;
  if (p >= plane) return;
  const int k = blockIdx.y;
  const int j = (int)(p / d.nx);
  const int i = (int)(p - (long long)j * d.nx);
  auto solid = [&](int ci, int cj) {
    return ci >= 1 && ci <= d.nx - 2 && cj >= 1 && cj <= d.ny - 2 && k <= d.nz - 2 &&
           is_solid_flag(flag[ci + cj * d.sy + k * d.sz]);
  };
  float uu = 0.0f, vv = 0.0f;
  if (k >= 1 && k <= d.nz - 2) {
    uu = up[k];
    vv = vp[k];
  }
  const bool here = solid(i, j);
  if (here || solid(i - 1, j)) uu = 0.0f;
  if (here || solid(i, j - 1)) vv = 0.0f;
  const long long fc = p + k * plane;
  u0[fc] = uu;
  v0[fc] = vv;
  w0[fc] = 0.0f;
}

bool finite_building(const qesmc_building &b)
{
  const float v[7] = {b.x_start, b.y_start, b.length, b.width, b.height, b.base_height, b.rotation_deg};
  for (float a : v)
    if (!std::isfinite(a) || std::fabs(a) > 1e8f) return false;
  return true;
}

// Profile on the host (nz values), extrusion and zeroing on the device.
int launch_inflow(qesmc_ctx *x, const qesmc_inflow &in, float *d_tmp)
{
  const qesmc::Dims &d = x->d;
  std::vector<float> z_face, z, up, vp;
  qesmc::vertical_grid(d.nz, x->dz_host.data(), z_face, z);
  if (qesmc::inflow_profile(d.nz, z, in, up, vp) != QESMC_OK) return QESMC_ERR_ARG;
  up.insert(up.end(), vp.begin(), vp.end());
  CCHK(hipMemcpyAsync(d_tmp, up.data(), sizeof(float) * up.size(), hipMemcpyHostToDevice, x->stream));
  const dim3 grid((unsigned)((d.fz + 255) / 256), (unsigned)d.nz);
  hipLaunchKernelGGL(k_city_inflow, grid, dim3(256), 0, x->stream, d, x->d_flag, d_tmp, d_tmp + d.nz, x->d_u0,
                     x->d_v0, x->d_w0);
  CCHK(hipGetLastError());
  CCHK(hipStreamSynchronize(x->stream));  // the staged host vectors die with this frame
  x->have_inflow = true;
  return QESMC_OK;
}

// A device buffer for the O(buildings + nz) host-evaluated values of one call.
struct Tmp {
  void *p = nullptr;
  ~Tmp()
  {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

extern "C" {

int qesmc_set_city(qesmc_ctx *x, int nx, int ny, int nz, float dx, const float *dz_array, int nb,
                   const qesmc_building *bs, const qesmc_inflow *inflow)
{
  if (!x || nx < 2 || ny < 2 || nz < 3 || !(dx > 0.f) || !std::isfinite(dx) || !dz_array || !inflow)
    return QESMC_ERR_ARG;
  if (nb < 0 || (nb > 0 && !bs)) return QESMC_ERR_ARG;
  if (inflow->kind < 0 || inflow->kind > 2) return QESMC_ERR_ARG;
  for (int k = 0; k < nz - 1; ++k)
    if (!(dz_array[k] > 0.f) || !std::isfinite(dz_array[k])) return QESMC_ERR_ARG;
  for (int b = 0; b < nb; ++b)
    if (!finite_building(bs[b])) return QESMC_ERR_ARG;
  const long long ncell = (long long)(nx - 1) * (ny - 1) * (nz - 1);
  if (ncell > (1LL << 31) - 1 || nz > 65535) return QESMC_ERR_UNSUPPORTED;  // grid.y = nz
  CCHK(hipSetDevice(x->device));
  int rc = qesmc::geometry_reserve(x, nx, ny, nz);
  if (rc != QESMC_OK) return rc;
  x->have_geom = x->have_inflow = false;
  x->dx = dx;
  x->dy = dx;  // getDomainSize() returns (dx, dx, dz), Domain.h:104
  x->dz_host.assign(dz_array, dz_array + (nz - 1));
  const qesmc::Dims &d = x->d;

  std::vector<float> z_face, z;
  qesmc::vertical_grid(nz, dz_array, z_face, z);
  // footprints clamped to the domain (cells outside it are skipped anyway)
  std::vector<Footprint> fps;
  fps.reserve(nb);
  for (int b = 0; b < nb; ++b) {
    Footprint F = qesmc::building_footprint(nz, dx, z_face, z, bs[b]);
    F.i0 = F.i0 < 0 ? 0 : F.i0;
    F.j0 = F.j0 < 0 ? 0 : F.j0;
    F.i1 = F.i1 > d.NX - 1 ? d.NX - 1 : F.i1;
    F.j1 = F.j1 > d.NY - 1 ? d.NY - 1 : F.j1;
    if (F.i1 < F.i0 || F.j1 < F.j0 || F.k1 <= F.k0) continue;
    fps.push_back(F);
  }
  // vertical denominators of Wall::solverCoefficients, layers [1, nz-3]
  std::vector<float> lay(2 * (size_t)(nz - 1), 1.0f);
  for (int k = 1; k < nz - 2; ++k) {
    lay[2 * k] = dz_array[k] * 0.5f * (dz_array[k] + dz_array[k + 1]);
    lay[2 * k + 1] = dz_array[k] * 0.5f * (dz_array[k] + dz_array[k - 1]);
  }
  const float dxx = dx * dx;

  Tmp tmp;
  const size_t fp_bytes = (sizeof(Footprint) * fps.size() + 255) / 256 * 256;
  const size_t lay_bytes = (sizeof(float) * lay.size() + 255) / 256 * 256;
  const size_t prof_bytes = sizeof(float) * 2 * (size_t)nz;
  CCHK(hipMalloc(&tmp.p, fp_bytes + lay_bytes + prof_bytes));
  char *base = (char *)tmp.p;
  Footprint *d_fp = (Footprint *)base;
  float *d_lay = (float *)(base + fp_bytes);
  float *d_prof = (float *)(base + fp_bytes + lay_bytes);
  if (!fps.empty())
    CCHK(hipMemcpyAsync(d_fp, fps.data(), sizeof(Footprint) * fps.size(), hipMemcpyHostToDevice, x->stream));
  CCHK(hipMemcpyAsync(d_lay, lay.data(), sizeof(float) * lay.size(), hipMemcpyHostToDevice, x->stream));
  CCHK(hipMemcpyAsync(x->d_dz, dz_array, sizeof(float) * (nz - 1), hipMemcpyHostToDevice, x->stream));

  const dim3 cgrid((unsigned)((d.sz + 255) / 256), (unsigned)d.NZ);
  hipLaunchKernelGGL(k_city_flags, cgrid, dim3(256), 0, x->stream, d, x->d_flag);
  CCHK(hipGetLastError());
  if (!fps.empty()) {
    hipLaunchKernelGGL(k_city_stamp, dim3((unsigned)fps.size()), dim3(256), 0, x->stream, d, d_fp, dx, x->d_flag);
    CCHK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_city_coefs, cgrid, dim3(256), 0, x->stream, d, x->d_flag, d_lay, dxx, x->d_coef[0],
                     x->d_coef[1], x->d_coef[2], x->d_coef[3], x->d_coef[4], x->d_coef[5]);
  CCHK(hipGetLastError());
  rc = launch_inflow(x, *inflow, d_prof);  // synchronises the stream
  if (rc != QESMC_OK) return rc;
  rc = qesmc::geometry_commit(x);
  if (rc != QESMC_OK) x->have_inflow = false;
  return rc;
}

int qesmc_set_city_inflow(qesmc_ctx *x, const qesmc_inflow *inflow)
{
  if (!x || !inflow) return QESMC_ERR_ARG;
  if (inflow->kind < 0 || inflow->kind > 2) return QESMC_ERR_ARG;
  if (!x->have_geom || !x->d_flag) return QESMC_ERR_STATE;
  CCHK(hipSetDevice(x->device));
  Tmp tmp;
  CCHK(hipMalloc(&tmp.p, sizeof(float) * 2 * (size_t)x->d.nz));
  return launch_inflow(x, *inflow, (float *)tmp.p);
}

int qesmc_get_inputs(qesmc_ctx *x, float *e, float *f, float *g, float *h, float *m, float *n, int *icellflag,
                     float *u0, float *v0, float *w0)
{
  if (!x) return QESMC_ERR_ARG;
  if (!x->have_geom) return QESMC_ERR_STATE;
  if ((u0 || v0 || w0) && !x->have_inflow) return QESMC_ERR_STATE;
  CCHK(hipSetDevice(x->device));
  const size_t cb = sizeof(float) * (size_t)x->ncell, fb = sizeof(float) * (size_t)x->nface;
  float *co[6] = {e, f, g, h, m, n};
  for (int a = 0; a < 6; ++a)
    if (co[a]) CCHK(hipMemcpyAsync(co[a], x->d_coef[a], cb, hipMemcpyDeviceToHost, x->stream));
  if (icellflag)
    CCHK(hipMemcpyAsync(icellflag, x->d_flag, sizeof(int) * (size_t)x->ncell, hipMemcpyDeviceToHost, x->stream));
  if (u0) CCHK(hipMemcpyAsync(u0, x->d_u0, fb, hipMemcpyDeviceToHost, x->stream));
  if (v0) CCHK(hipMemcpyAsync(v0, x->d_v0, fb, hipMemcpyDeviceToHost, x->stream));
  if (w0) CCHK(hipMemcpyAsync(w0, x->d_w0, fb, hipMemcpyDeviceToHost, x->stream));
  CCHK(hipStreamSynchronize(x->stream));
  return QESMC_OK;
}

}  // extern "C"

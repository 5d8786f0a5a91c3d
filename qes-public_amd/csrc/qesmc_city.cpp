// Input builder for flat stair-step cities: the arrays WINDSGeneralData's
// constructor and per-step wind set-up hand to Solver::solve.
//
//   ground ghost flags           WINDSGeneralData.cpp:439-444
//   rectangular building flags   RectangularBuilding.h:86-103, PolyBuilding.cpp:130-236
//   wall coefficients            Wall.cpp:44-349 (stair-step branches), 827-846
//   single-sensor inflow         WindProfilerSensorType.cpp:220, 290-333, 571-621
//   solid zeroing of u0          Wall.cpp:805-824
//
// Host code (it runs once per geometry); float/double promotions follow the
// reference so the arrays equal the reference's for the same inputs.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "qesmc_city_internal.h"

namespace qesmc {

void vertical_grid(int nz, const float *dz, std::vector<float> &z_face, std::vector<float> &z)
{
  z_face.assign(nz, 0.f);
  z.assign(nz - 1, 0.f);
  z_face[0] = -dz[0];
  z_face[1] = 0.0f;
  for (int k = 2; k < nz; ++k) z_face[k] = z_face[k - 1] + dz[k - 1];
  z[0] = -0.5f * dz[0];
  for (int k = 1; k < nz - 1; ++k) z[k] = 0.5f * (z_face[k] + z_face[k + 1]);
}

Footprint building_footprint(int nz, float dx, const std::vector<float> &z_face, const std::vector<float> &z,
                             const qesmc_building &b)
{
  Footprint F;
  const float dy = dx;  // getDomainSize() returns (dx, dx, dz), Domain.h:104
  const float height_eff = b.height + b.base_height;
  const float rot = (float)(b.rotation_deg * M_PI / 180.0);
  F.xv[0] = F.xv[4] = b.x_start;
  F.yv[0] = F.yv[4] = b.y_start;
  F.xv[1] = b.x_start - b.width * std::sin(rot);
  F.yv[1] = b.y_start + b.width * std::cos(rot);
  F.xv[2] = F.xv[1] + b.length * std::cos(rot);
  F.yv[2] = F.yv[1] + b.length * std::sin(rot);
  F.xv[3] = b.x_start + b.length * std::cos(rot);
  F.yv[3] = b.y_start + b.length * std::sin(rot);
  float x_min = F.xv[0], x_max = F.xv[0], y_min = F.yv[0], y_max = F.yv[0];
  for (int v = 1; v < 5; ++v) {
    x_min = F.xv[v] < x_min ? F.xv[v] : x_min;
    x_max = F.xv[v] > x_max ? F.xv[v] : x_max;
    y_min = F.yv[v] < y_min ? F.yv[v] : y_min;
    y_max = F.yv[v] > y_max ? F.yv[v] : y_max;
  }
  F.i0 = (int)(x_min / dx);
  F.i1 = (int)((x_max / dx) + 1);
  F.j0 = (int)(y_min / dy);
  F.j1 = (int)((y_max / dy) + 1);
  const int zn = nz - 1;
  F.k0 = F.k1 = 1;
  for (int k = 1; k < zn; ++k) {
    F.k0 = k;
    if (b.base_height <= z_face[k + 1]) break;
  }
  for (int k = 0; k < zn; ++k) {
    F.k1 = k + 1;
    if (k + 1 >= zn || height_eff < z[k + 1]) break;
  }
  return F;
}

int inflow_profile(int nz, const std::vector<float> &z, const qesmc_inflow &in, std::vector<float> &up,
                   std::vector<float> &vp)
{
  up.assign(nz, 0.f);
  vp.assign(nz, 0.f);
  const float vk = 0.4f;
  const float theta = (float)((270.0 - in.wind_dir_deg) * M_PI / 180.0);
  const float ct = std::cos(theta), st = std::sin(theta);
  const float U = in.U_ref, zr = in.z_ref, zp = in.z0_or_p;
  if (in.kind == 0) {
    for (int k = 1; k < nz; ++k) {
      up[k] = ct * U;
      vp[k] = st * U;
    }
  } else if (in.kind == 1) {
    const float psi = 0.0f;  // neutral stability
    const float u_star = U * vk / (std::log((zr + zp) / zp) + psi);
    for (int k = 1; k < nz - 1; ++k) {
      const float lg = std::log(((z[k] - 0.0f) + zp) / zp) + psi;
      up[k] = (ct * u_star / vk) * lg;
      vp[k] = (st * u_star / vk) * lg;
      if (std::fabs(u_star) >= 2 * U) {
        up[k] = ct * U;
        vp[k] = st * U;
      }
    }
  } else if (in.kind == 2) {
    for (int k = 1; k < nz - 1; ++k) {
      const float pw = std::pow(((z[k] - 0.0f) / zr), zp);
      up[k] = ct * U * pw;
      vp[k] = st * U * pw;
    }
  } else {
    return QESMC_ERR_ARG;
  }
  return QESMC_OK;
}

}  // namespace qesmc

namespace {

using qesmc::is_solid_flag;

struct Grid {
  int nx, ny, nz;
  long long sy, sz, fy, fz;
  long long cell(int i, int j, int k) const { return i + j * sy + k * sz; }
  long long face(int i, int j, int k) const { return i + j * fy + k * fz; }
};

// PolyBuilding::setCellFlags for one rectangle (meshTypeFlag 0).
void stamp_building(const Grid &G, float dx, const std::vector<float> &z_face,
                    const std::vector<float> &z, const qesmc_building &b, int *flag)
{
  const qesmc::Footprint F = qesmc::building_footprint(G.nz, dx, z_face, z, b);
  for (int j = F.j0; j <= F.j1; ++j) {
    const float yc = (j + 0.5) * dx;
    for (int i = F.i0; i <= F.i1; ++i) {
      const float xc = (i + 0.5) * dx;
      if (!qesmc::footprint_contains(F, xc, yc)) continue;
      if (i < 0 || j < 0 || i >= G.nx - 1 || j >= G.ny - 1) continue;
      for (int k = F.k0; k < F.k1; ++k) {
        int &fl = flag[G.cell(i, j, k)];
        if (fl != 7) fl = 0;
      }
    }
  }
}

// splitmix64: small, seedable, identical on every platform
struct Rng {
  uint64_t s;
  uint64_t next()
  {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal()
  {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

}  // namespace

extern "C" {

int qesmc_city_build(int nx, int ny, int nz, float dx, const float *dz_array, int nb,
                     const qesmc_building *bs, const qesmc_inflow *inflow, float *e, float *f,
                     float *g, float *h, float *m, float *n, int *flag, float *u0, float *v0,
                     float *w0)
{
  if (nx < 2 || ny < 2 || nz < 3 || !(dx > 0.f) || !dz_array || !inflow) return QESMC_ERR_ARG;
  if (!e || !f || !g || !h || !m || !n || !flag || !u0 || !v0 || !w0) return QESMC_ERR_ARG;
  if (nb > 0 && !bs) return QESMC_ERR_ARG;
  Grid G{nx, ny, nz, nx - 1, (long long)(nx - 1) * (ny - 1), nx, (long long)nx * ny};
  const long long ncell = (long long)(nx - 1) * (ny - 1) * (nz - 1);
  const long long nface = (long long)nx * ny * nz;
  std::vector<float> z_face, z;
  qesmc::vertical_grid(nz, dz_array, z_face, z);

  // allocateMemory: coefficients 1, cells fluid; ground ghost layer terrain
  float *coef[6] = {e, f, g, h, m, n};
  for (long long c = 0; c < ncell; ++c) {
    for (auto *a : coef) a[c] = 1.0f;
    flag[c] = 1;
  }
  for (int j = 0; j < ny - 1; ++j)
    for (int i = 0; i < nx - 1; ++i) flag[G.cell(i, j, 0)] = 2;

  for (int b = 0; b < nb; ++b) stamp_building(G, dx, z_face, z, bs[b], flag);

  // Wall::defineWalls: clamp, then zero the coefficient toward each solid
  // neighbour of a fluid cell (k in [1, nz-3]).  East / north lookups use
  // the flat index +1 / +(nx-1), wrapping at the last column / row.
  for (int k = 1; k < nz - 1; ++k)
    for (int j = 0; j < ny - 1; ++j)
      for (int i = 0; i < nx - 1; ++i) {
        const long long c = G.cell(i, j, k);
        for (auto *a : coef) {
          if (a[c] < 0.05) a[c] = 0.0f;
          if (a[c] > 1.0) a[c] = 1.0f;
        }
      }
  for (int k = 1; k < nz - 2; ++k)
    for (int j = 0; j < ny - 1; ++j)
      for (int i = 0; i < nx - 1; ++i) {
        const long long c = G.cell(i, j, k);
        if (is_solid_flag(flag[c])) continue;
        if (is_solid_flag(flag[c - G.sz])) n[c] = 0.0f;
        if (is_solid_flag(flag[c + G.sz])) m[c] = 0.0f;
        if (i > 0 && is_solid_flag(flag[c - 1])) f[c] = 0.0f;
        if (is_solid_flag(flag[c + 1])) e[c] = 0.0f;
        if (j > 0 && is_solid_flag(flag[c - G.sy])) h[c] = 0.0f;
        if (is_solid_flag(flag[c + G.sy])) g[c] = 0.0f;
      }

  // Wall::solverCoefficients (dy taken as dx, Domain.h:104)
  const float dxx = dx * dx;
  for (int k = 1; k < nz - 2; ++k) {
    const float mden = dz_array[k] * 0.5f * (dz_array[k] + dz_array[k + 1]);
    const float nden = dz_array[k] * 0.5f * (dz_array[k] + dz_array[k - 1]);
    for (int j = 0; j < ny - 1; ++j)
      for (int i = 0; i < nx - 1; ++i) {
        const long long c = G.cell(i, j, k);
        e[c] = e[c] / dxx;
        f[c] = f[c] / dxx;
        g[c] = g[c] / dxx;
        h[c] = h[c] / dxx;
        m[c] = m[c] / mden;
        n[c] = n[c] / nden;
      }
  }

  // sensor profile on flat ground (z_terrain = 0, neutral: psi = 0)
  std::vector<float> up, vp;
  if (qesmc::inflow_profile(nz, z, *inflow, up, vp) != QESMC_OK) return QESMC_ERR_ARG;
  std::memset(u0, 0, sizeof(float) * nface);
  std::memset(v0, 0, sizeof(float) * nface);
  std::memset(w0, 0, sizeof(float) * nface);
  for (int k = 1; k < nz - 1; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long fc = G.face(i, j, k);
        u0[fc] = up[k];
        v0[fc] = vp[k];
      }

  // Wall::setVelocityZero
  for (int k = 0; k < nz - 1; ++k)
    for (int j = 1; j < ny - 1; ++j)
      for (int i = 1; i < nx - 1; ++i) {
        if (!is_solid_flag(flag[G.cell(i, j, k)])) continue;
        const long long fc = G.face(i, j, k);
        u0[fc] = u0[fc + 1] = 0.0f;
        v0[fc] = v0[fc + G.fy] = 0.0f;
        w0[fc] = w0[fc + G.fz] = 0.0f;
      }
  return QESMC_OK;
}

int qesmc_city_random_buildings(unsigned long long seed, int nx, int ny, float dx, int nb,
                                float lmin, float lmax, float hmin, float hmax, int lognormal,
                                float hcap, int margin, qesmc_building *out)
{
  if (nx < 2 || ny < 2 || !(dx > 0.f) || nb < 0 || (nb > 0 && !out) || margin < 0) return QESMC_ERR_ARG;
  if (!(lmax >= lmin) || !(lmin > 0.f)) return QESMC_ERR_ARG;
  const int NX = nx - 1, NY = ny - 1;
  std::vector<unsigned char> occ((size_t)NX * NY, 0);
  Rng rng{seed};
  int placed = 0;
  const long long max_tries = 200LL * (nb > 0 ? nb : 1);
  for (long long tries = 0; placed < nb && tries < max_tries; ++tries) {
    const float L = (float)(lmin + (lmax - lmin) * rng.uniform());
    const float W = (float)(lmin + (lmax - lmin) * rng.uniform());
    float H;
    if (lognormal) {
      H = (float)(hmin * std::exp(hmax * rng.normal()));
      if (H > hcap) H = hcap;
      if (H < dx) H = dx;
    } else {
      H = (float)(hmin + (hmax - hmin) * rng.uniform());
    }
    const double xlo = margin * (double)dx, xhi = (NX - margin) * (double)dx - L;
    const double ylo = margin * (double)dx, yhi = (NY - margin) * (double)dx - W;
    if (xhi <= xlo || yhi <= ylo) continue;
    // corners on whole cells keep the stair-step footprint exactly L x W
    const float xs = (float)std::floor((xlo + (xhi - xlo) * rng.uniform()) / dx) * dx;
    const float ys = (float)std::floor((ylo + (yhi - ylo) * rng.uniform()) / dx) * dx;
    // footprint cells plus a one-cell gap on every side must be free
    const int i0 = (int)std::floor(xs / dx) - 1, i1 = (int)std::ceil((xs + L) / dx) + 1;
    const int j0 = (int)std::floor(ys / dx) - 1, j1 = (int)std::ceil((ys + W) / dx) + 1;
    if (i0 < 0 || j0 < 0 || i1 > NX || j1 > NY) continue;
    bool free_ = true;
    for (int j = j0; j < j1 && free_; ++j)
      for (int i = i0; i < i1; ++i)
        if (occ[(size_t)j * NX + i]) {
          free_ = false;
          break;
        }
    if (!free_) continue;
    for (int j = j0 + 1; j < j1 - 1; ++j)
      for (int i = i0 + 1; i < i1 - 1; ++i) occ[(size_t)j * NX + i] = 1;
    out[placed++] = qesmc_building{xs, ys, L, W, H, 0.0f, 0.0f};
  }
  return placed;
}

}  // extern "C"

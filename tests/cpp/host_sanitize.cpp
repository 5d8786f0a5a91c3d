// Host code of libqesmc (the NetCDF workspace reader/writer and the flat-city
// builder) and the CPU oracle, driven under AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/test_host_sanitizers.py builds this file
// with g++ -fsanitize=address,undefined and the sources it names).
//
//   host_sanitize <dir> [extra.nc ...]
//
// 1. builds a small city with qesmc_city_build and with the oracle's
//    builders and requires them to agree bitwise; runs the oracle's solves;
// 2. writes a workspace file and reads it back;
// 3. feeds the reader malformed files: every truncation of the header, a
//    hand-made header whose variable names a dimension id out of range, huge
//    name / attribute / variable counts, offsets past the end of the file,
//    and seeded random byte flips of the header.  Each must return a status
//    code (no crash, no sanitizer report);
// 4. reads the extra files given on the command line (made by the Python
//    test, e.g. an icellflag variable holding NaN), which must be refused.
// Prints PASS on success.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "qesmc.h"

extern "C" {
void oracle_vertical_grid(int nz, const float *dz_array, float *z_face, float *z);
void oracle_init_cells(int nx, int ny, int nz, float *e, float *f, float *g, float *h, float *m, float *n,
                       int *icellflag);
void oracle_rect_building(int nx, int ny, int nz, float dx, float dy_unused, const float *z_face, const float *z,
                          float x_start, float y_start, float L, float W, float H, float base_height,
                          float rotation_deg, int *icellflag);
void oracle_define_walls(int nx, int ny, int nz, const int *icellflag, float *e, float *f, float *g, float *h,
                         float *m, float *n);
void oracle_solver_coefficients(int nx, int ny, int nz, float dx, const float *dz_array, float *e, float *f,
                                float *g, float *h, float *m, float *n);
void oracle_profile(int nz, const float *z, int kind, float U_ref, float z_ref, float z0_or_p, float wind_dir,
                    float *u_prof, float *v_prof);
void oracle_extrude_profile(int nx, int ny, int nz, const float *u_prof, const float *v_prof, float *u0, float *v0,
                            float *w0);
void oracle_set_velocity_zero(int nx, int ny, int nz, const int *icellflag, float *u0, float *v0, float *w0);
int oracle_solve_rb_f32(int nx, int ny, int nz, float dx, float dy, const float *dz_array, const float *e,
                        const float *f, const float *g, const float *h, const float *m, const float *n,
                        const int *icellflag, const float *u0, const float *v0, const float *w0, float *R,
                        float *lambda, float *lambda_old, float omega, float tol, int itermax, int alpha1, int alpha2,
                        float *u, float *v, float *w, float *max_error_out);
void oracle_divergence_f64(int nx, int ny, int nz, double dx, double dy, const float *dz_array, const float *e,
                           const float *f, const float *g, const float *h, const float *m, const float *n,
                           const float *u0, const float *v0, const float *w0, int alpha1, double *R);
int oracle_sor_rb_f64(int nx, int ny, int nz, const float *e, const float *f, const float *g, const float *h,
                      const float *m, const float *n, const double *R, double *lambda, double *lambda_old,
                      double omega, double tol, int itermax, double *max_error_out);
double oracle_div_linf(int nx, int ny, int nz, double dx, double dy, const float *dz_array, const float *e,
                       const float *f, const float *g, const float *h, const float *m, const float *n,
                       const int *icellflag, const float *u, const float *v, const float *w, int physical);
long oracle_wall_log_bc(int nx, int ny, int nz, float dx, float dz, float z0, const int *icellflag, float *u,
                        float *v, float *w);
void oracle_velocity_gradient(int nx, int ny, int nz, float dx, float dy, float dz, const float *dz_array,
                              int uniform_z, const int *icellflag, const float *u, const float *v, const float *w,
                              float *G);
int oracle_turb_wall_stairstep(int nx, int ny, int nz, float dx, const float *dz_array, float z0,
                               const float *z0_domain_u, const float *z0_domain_v, int building_mode,
                               int terrain_mode, const int *icellflag, const float *u, const float *v, const float *w,
                               float *G);
}

static int failures = 0;
#define EXPECT(c)                                                         \
  do {                                                                    \
    if (!(c)) {                                                           \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);            \
      ++failures;                                                         \
    }                                                                     \
  } while (0)

static std::vector<unsigned char> slurp(const std::string &p)
{
  std::vector<unsigned char> b;
  FILE *f = std::fopen(p.c_str(), "rb");
  if (!f) return b;
  unsigned char buf[65536];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) b.insert(b.end(), buf, buf + n);
  std::fclose(f);
  return b;
}

static void spit(const std::string &p, const unsigned char *d, size_t n)
{
  FILE *f = std::fopen(p.c_str(), "wb");
  if (n) std::fwrite(d, 1, n, f);
  std::fclose(f);
}

// every call on a malformed file must come back with a status, never crash
static int read_all(const std::string &path, int nx, int ny, int nz, std::vector<float> *bufs, std::vector<int> &flag)
{
  int fx = 0, fy = 0, fz = 0;
  int rc = qesmc_workspace_dims(path.c_str(), &fx, &fy, &fz);
  if (rc != QESMC_OK) return rc;
  float dx = 0, dy = 0;
  return qesmc_workspace_read(path.c_str(), nx, ny, nz, &dx, &dy, bufs[0].data(), bufs[1].data(), bufs[2].data(),
                              bufs[3].data(), bufs[4].data(), bufs[5].data(), bufs[6].data(), flag.data(),
                              bufs[7].data(), bufs[8].data(), bufs[9].data(), bufs[10].data(), bufs[11].data(),
                              bufs[12].data());
}

struct Be {
  std::vector<unsigned char> b;
  void u32(uint32_t v)
  {
    for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
  }
  void name(const char *s)
  {
    const uint32_t n = (uint32_t)std::strlen(s);
    u32(n);
    b.insert(b.end(), s, s + n);
    while (b.size() % 4) b.push_back(0);
  }
};

int main(int argc, char **argv)
{
  if (argc < 2) return 2;
  const std::string dir = argv[1];
  const int ncx = 23, ncy = 17, ncz = 9, nx = ncx + 1, ny = ncy + 1, nz = ncz + 2;
  const float dx = 2.0f;
  const long ncell = (long)(nx - 1) * (ny - 1) * (nz - 1), nface = (long)nx * ny * nz;
  std::vector<float> dz(nz - 1);
  for (int k = 1; k < nz - 1; ++k) dz[k] = (float)std::pow(1.05, k - 1);
  dz[0] = dz[1];

  // ---- 1. host builder vs oracle builder, oracle solves -------------------
  qesmc_building blds[4];
  const int nb = qesmc_city_random_buildings(3, nx, ny, dx, 4, 4.0f, 12.0f, 3.0f, 9.0f, 0, 0.0f, 1, blds);
  EXPECT(nb > 0);
  blds[0].rotation_deg = 30.0f;  // one rotated footprint
  const qesmc_inflow inflow = {1, 4.0f, 10.0f, 0.5f, 233.0f};
  std::vector<float> co[6], cu[3];
  std::vector<int> flag(ncell), oflag(ncell);
  for (auto &a : co) a.assign(ncell, 0.0f);
  for (auto &a : cu) a.assign(nface, 0.0f);
  EXPECT(qesmc_city_build(nx, ny, nz, dx, dz.data(), nb, blds, &inflow, co[0].data(), co[1].data(), co[2].data(),
                          co[3].data(), co[4].data(), co[5].data(), flag.data(), cu[0].data(), cu[1].data(),
                          cu[2].data()) == QESMC_OK);
  std::vector<float> oc[6], ou[3], zf(nz), zc(nz - 1), up(nz), vp(nz);
  for (auto &a : oc) a.assign(ncell, 0.0f);
  for (auto &a : ou) a.assign(nface, 0.0f);
  oracle_vertical_grid(nz, dz.data(), zf.data(), zc.data());
  oracle_init_cells(nx, ny, nz, oc[0].data(), oc[1].data(), oc[2].data(), oc[3].data(), oc[4].data(), oc[5].data(),
                    oflag.data());
  for (int b = 0; b < nb; ++b)
    oracle_rect_building(nx, ny, nz, dx, dx, zf.data(), zc.data(), blds[b].x_start, blds[b].y_start, blds[b].length,
                         blds[b].width, blds[b].height, blds[b].base_height, blds[b].rotation_deg, oflag.data());
  oracle_define_walls(nx, ny, nz, oflag.data(), oc[0].data(), oc[1].data(), oc[2].data(), oc[3].data(), oc[4].data(),
                      oc[5].data());
  oracle_solver_coefficients(nx, ny, nz, dx, dz.data(), oc[0].data(), oc[1].data(), oc[2].data(), oc[3].data(),
                             oc[4].data(), oc[5].data());
  oracle_profile(nz, zc.data(), inflow.kind, inflow.U_ref, inflow.z_ref, inflow.z0_or_p, inflow.wind_dir_deg,
                 up.data(), vp.data());
  oracle_extrude_profile(nx, ny, nz, up.data(), vp.data(), ou[0].data(), ou[1].data(), ou[2].data());
  oracle_set_velocity_zero(nx, ny, nz, oflag.data(), ou[0].data(), ou[1].data(), ou[2].data());
  EXPECT(flag == oflag);
  for (int a = 0; a < 6; ++a) EXPECT(std::memcmp(co[a].data(), oc[a].data(), 4 * ncell) == 0);
  for (int a = 0; a < 3; ++a) EXPECT(std::memcmp(cu[a].data(), ou[a].data(), 4 * nface) == 0);

  std::vector<float> R(ncell), lam(ncell, 0.0f), old(ncell), u(nface), v(nface), w(nface), G(9 * ncell);
  float me = 0;
  const int it = oracle_solve_rb_f32(nx, ny, nz, dx, dx, dz.data(), co[0].data(), co[1].data(), co[2].data(),
                                     co[3].data(), co[4].data(), co[5].data(), flag.data(), cu[0].data(),
                                     cu[1].data(), cu[2].data(), R.data(), lam.data(), old.data(), 1.78f, 1e-9f, 200,
                                     1, 1, u.data(), v.data(), w.data(), &me);
  EXPECT(it == 200 && std::isfinite(me));
  std::vector<double> R64(ncell), l64(ncell, 0.0), o64(ncell);
  double me64 = 0;
  oracle_divergence_f64(nx, ny, nz, dx, dx, dz.data(), co[0].data(), co[1].data(), co[2].data(), co[3].data(),
                        co[4].data(), co[5].data(), cu[0].data(), cu[1].data(), cu[2].data(), 1, R64.data());
  oracle_sor_rb_f64(nx, ny, nz, co[0].data(), co[1].data(), co[2].data(), co[3].data(), co[4].data(), co[5].data(),
                    R64.data(), l64.data(), o64.data(), 1.78, 1e-9, 5000, &me64);
  EXPECT(me64 <= 1e-9);
  EXPECT(std::isfinite(oracle_div_linf(nx, ny, nz, dx, dx, dz.data(), co[0].data(), co[1].data(), co[2].data(),
                                       co[3].data(), co[4].data(), co[5].data(), flag.data(), u.data(), v.data(),
                                       w.data(), 1)));
  oracle_velocity_gradient(nx, ny, nz, dx, dx, 1.0f, dz.data(), 0, flag.data(), u.data(), v.data(), w.data(),
                           G.data());
  {
    // wallLogBC restatement (reads neighbour faces at +-1, +-nx, +-nx*ny of
    // every wall face, including the wrapped ones at the domain edges)
    std::vector<float> lu = u, lv = v, lw = w;
    EXPECT(oracle_wall_log_bc(nx, ny, nz, dx, 1.0f, 0.1f, flag.data(), lu.data(), lv.data(), lw.data()) > 0);
  }
  {
    // the gradient's wall step in every mode pair (faces at +-1, +-nx, +-nx*ny
    // of the wall cells, z0 arrays at the four surrounding entries)
    std::vector<float> z0u((size_t)nx * ny, 0.1f), z0v((size_t)nx * ny, 0.12f);
    for (int bm = 0; bm < 3; ++bm)
      for (int tm = 0; tm < 3; ++tm) {
        std::vector<float> Gw = G;
        EXPECT(oracle_turb_wall_stairstep(nx, ny, nz, dx, dz.data(), 0.1f, z0u.data(), z0v.data(), bm, tm,
                                          flag.data(), u.data(), v.data(), w.data(), Gw.data()) == 0);
      }
  }

  // ---- 2. a valid workspace file round trip --------------------------------
  const std::string good = dir + "/good.nc";
  qesmc_grid grid = {nx, ny, nz, dx, dx, dz.data()};
  qesmc_coeffs cs = {co[0].data(), co[1].data(), co[2].data(), co[3].data(), co[4].data(), co[5].data(), flag.data()};
  qesmc_fields fs = {cu[0].data(), cu[1].data(), cu[2].data(), u.data(), v.data(), w.data()};
  EXPECT(qesmc_workspace_write(good.c_str(), &grid, &cs, &fs) == QESMC_OK);
  std::vector<float> bufs[13];
  bufs[0].assign(nz - 1, 0.0f);
  for (int a = 1; a <= 6; ++a) bufs[a].assign(ncell, 0.0f);
  for (int a = 7; a <= 12; ++a) bufs[a].assign(nface, 0.0f);
  std::vector<int> rflag(ncell);
  EXPECT(read_all(good, nx, ny, nz, bufs, rflag) == QESMC_OK);
  EXPECT(rflag == flag);
  EXPECT(std::memcmp(bufs[10].data(), u.data(), 4 * nface) == 0);

  // ---- 3. malformed files ---------------------------------------------------
  const std::vector<unsigned char> data = slurp(good);
  const std::string bad = dir + "/bad.nc";
  int refused = 0, accepted = 0;
  auto attempt = [&](const std::vector<unsigned char> &bytes) {
    spit(bad, bytes.data(), bytes.size());
    const int rc = read_all(bad, nx, ny, nz, bufs, rflag);
    (rc == QESMC_OK ? accepted : refused) += 1;
    return rc;
  };
  // the header ends where the first variable's data begins (a few KB)
  const size_t hdr = 4096 < data.size() ? 4096 : data.size();
  for (size_t len = 0; len < hdr; len += 3) {
    std::vector<unsigned char> t(data.begin(), data.begin() + len);
    EXPECT(attempt(t) != QESMC_OK);
  }
  {  // a variable naming dimension id 7 of a one-dimension header
    Be h;
    h.b = {'C', 'D', 'F', 1};
    h.u32(0);
    h.u32(0x0A);
    h.u32(1);
    h.name("x_face");
    h.u32(5);
    h.u32(0);
    h.u32(0);
    h.u32(0x0B);
    h.u32(1);
    h.name("u");
    h.u32(1);
    h.u32(7);
    h.u32(0);
    h.u32(0);
    h.u32(5);
    h.u32(20);
    h.u32(200);
    h.b.resize(400, 0);
    EXPECT(attempt(h.b) != QESMC_OK);
    int a = 0, b = 0, c = 0;
    spit(bad, h.b.data(), h.b.size());
    EXPECT(qesmc_workspace_dims(bad.c_str(), &a, &b, &c) != QESMC_OK);
  }
  {  // huge name length, dimension count, attribute count, variable count
    const uint32_t huge[] = {0xffffffffu, 0x7fffffffu, 0x10000000u};
    for (uint32_t hv : huge)
      for (int which = 0; which < 4; ++which) {
        Be h;
        h.b = {'C', 'D', 'F', 2};
        h.u32(0);
        h.u32(0x0A);
        if (which == 1) {
          h.u32(hv);
        } else {
          h.u32(3);
          for (const char *nm : {"x_face", "y_face", "z_face"}) {
            if (which == 0 && nm[0] == 'y') {
              h.u32(hv);  // name length
            } else {
              h.name(nm);
            }
            h.u32(4);
          }
        }
        h.u32(0x0C);
        h.u32(which == 2 ? hv : 0);
        h.u32(0x0B);
        h.u32(which == 3 ? hv : 0);
        h.b.resize(1024, 0);
        EXPECT(attempt(h.b) != QESMC_OK);
      }
  }
  {  // every variable's begin offset pointed past the end of the file
    std::vector<unsigned char> t = data;
    t.resize(hdr);
    EXPECT(attempt(t) != QESMC_OK);
  }
  // seeded random byte flips of the header region
  std::mt19937_64 rng(20261203);
  for (int trial = 0; trial < 1500; ++trial) {
    std::vector<unsigned char> t = data;
    const int flips = 1 + (int)(rng() % 4);
    for (int q = 0; q < flips; ++q) {
      const size_t at = rng() % hdr;
      t[at] = (unsigned char)(rng() & 0xff);
    }
    attempt(t);
  }

  // ---- 4. files made elsewhere that must be refused ------------------------
  for (int a = 2; a < argc; ++a) EXPECT(read_all(argv[a], nx, ny, nz, bufs, rflag) != QESMC_OK);

  std::printf("malformed files: %d refused, %d read\n", refused, accepted);
  if (failures) {
    std::printf("FAILED %d checks\n", failures);
    return 1;
  }
  std::printf("PASS\n");
  return 0;
}

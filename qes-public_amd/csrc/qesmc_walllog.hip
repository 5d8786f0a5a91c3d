// Log-law wall correction of the solved wind field on stair-step walls:
// Wall::wallLogBC (src/winds/Wall.cpp:352-797) as QES-Winds applies it after
// a solve when logLawFlag is set (qesWinds/qesWindsMain.cpp:157-165:
// u0 = u; wallLogBC; u = u0), on the resident u, v, w.
//
// The reference keeps six lists of wall-adjacent fluid cells, filled by
// Wall::defineWalls (Wall.cpp:288-340), and walks them in six passes (wall
// below, above, back, front, right, left).  For every listed cell the face
// (i,j,k) gets its wall-normal component set to 0 and its two parallel
// components set to the log-law speed of a 20-step fixed point for the
// friction velocity, from the velocity of the face one cell further from
// the wall.  Within a pass no face is both written and read (the neighbour
// face belongs to a cell that cannot be in the same list), so a pass is a
// parallel map; passes depend on each other (a later pass reads faces an
// earlier one rewrote), so each is one launch, in the reference's order.
// List membership is decided on the fly from the flags, one thread per cell
// of the listed range k in [1, nz-3]: the lists' order never matters.
//
// Arithmetic follows the reference's types: Wall.h includes <math.h> with
// `using namespace std`, so log/atan2/cos/sin of float arguments are float
// functions, pow(float, 2.0) is double (x*x, exact for a float x) and the
// speed is a double sqrt (rounded to nearest, see sqrt_rn) stored to float.
// log(dist2/dist1) and
// log(dist1/z0) depend on the pass only and are evaluated on the host with
// the host's logf; atan2f, cosf and sinf run on the device and agree with
// the host's to a few ulp.
#include <cmath>

#include "qesmc_internal.h"

namespace {

using qesmc::Dims;

#define WCHK(expr)                                         \
  do {                                                     \
    hipError_t _e = (expr);                                \
    if (_e != hipSuccess) return qesmc::map_hip_error(_e); \
  } while (0)

constexpr unsigned ST_CUTCELL = 8u;

__device__ inline bool solid_flag(int f) { return f == 0 || f == 2; }

// sqrt(x) rounded to nearest, as the host's sqrt: the device's double sqrt
// can be one ulp off, which the log-law recurrence amplifies (|d m/d V| reaches
// 1e6 when it diverges), so its result is checked with Tuckerman's rounding
// test -- s is the rounded root iff s*pred(s) < x <= s*succ(s), the products'
// signs against x computed exactly by fma -- and moved by one ulp if needed.
__device__ inline double sqrt_rn(double x)
{
  double s = sqrt(x);
  if (!(x > 0.0) || !isfinite(x)) return s;
  const double sm = nextafter(s, 0.0), sp = nextafter(s, __builtin_huge_val());
  if (fma(s, sm, -x) >= 0.0) return sm;
  if (fma(s, sp, -x) < 0.0) return sp;
  return s;
}

// Flags 7/8 (cut cells) take the reference's first loop (Wall.cpp:381-655),
// which is not built: raise a status bit if any cell has one.
__global__ __launch_bounds__(256) void k_find_cut_cells(const int *__restrict__ flag, long long ncell,
                                                        unsigned *__restrict__ status)
{
  const long long stride = (long long)gridDim.x * blockDim.x;
  bool cut = false;
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < ncell; c += stride) {
    const int f = flag[c];
    cut |= (f == 7 || f == 8);
  }
  if (__any(cut) && (threadIdx.x & 63) == 0) atomicOr(status, ST_CUTCELL);
}

struct Pass {
  long long cell_off;  // neighbour cell that must be solid (cellAdd offset)
  long long face_off;  // face read: the second cell from the wall
  int guard;           // 0 none, 1 i > 0, 2 j > 0
  int a, b, n;         // parallel components (atan2(b, a)) and the normal one: 0 u, 1 v, 2 w
  float log_d2_d1;     // logf(dist2 / dist1)
  float log_d1_z0;     // logf(dist1 / z0)
};

// One pass over the listed range; every thread one cell.
__global__ __launch_bounds__(256) void k_wall_log(Dims d, Pass P, const int *__restrict__ flag, float *u, float *v,
                                                  float *w)
{
  const long long plane = d.sz;
  const long long total = plane * (long long)(d.nz - 3);
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = blockIdx.x * 256LL + threadIdx.x; q < total; q += stride) {
    const int k = 1 + (int)(q / plane);
    const long long p = q - (long long)(k - 1) * plane;
    const int j = (int)(p / d.NX);
    const int i = (int)(p - (long long)j * d.NX);
    const long long c = p + (long long)k * plane;
    const int fl = flag[c];
    if (fl == 0 || fl == 2 || fl == 8) continue;
    if ((P.guard == 1 && i == 0) || (P.guard == 2 && j == 0)) continue;
    if (!solid_flag(flag[c + P.cell_off])) continue;
    const long long f = i + (long long)j * d.nx + (long long)k * d.fz;
    const long long nb = f + P.face_off;
    float *A = P.a == 0 ? u : P.a == 1 ? v : w;
    float *B = P.b == 0 ? u : P.b == 1 ? v : w;
    float *N = P.n == 0 ? u : P.n == 1 ? v : w;
    const float an = A[nb], bn = B[nb];
    const float vk = 0.4f;
    // loop-invariant in the reference's loop (its neighbour face is never
    // written by the same pass): direction and speed of the second cell
    const float wind_dir = atan2f(bn, an);
    const float vel_mag2 = (float)sqrt_rn((double)an * (double)an + (double)bn * (double)bn);
    float ustar_wall = 0.1f, vel_mag1 = 0.f;
    for (int iter = 0; iter < 20; iter++) {
      vel_mag1 = vel_mag2 - (ustar_wall / vk) * P.log_d2_d1;
      ustar_wall = vk * vel_mag1 / P.log_d1_z0;
    }
    N[f] = 0.f;
    A[f] = vel_mag1 * cosf(wind_dir);
    B[f] = vel_mag1 * sinf(wind_dir);
  }
}

}  // namespace

namespace qesmc {

// Whether the resident flags hold a cut cell (7 or 8), on the context's stream.
hipError_t find_cut_cells(qesmc_ctx *x, bool *cut)
{
  hipStream_t s = x->stream;
  hipError_t e = hipMemsetAsync(x->d_status, 0, sizeof(unsigned), s);
  if (e != hipSuccess) return e;
  long long blocks = (x->ncell + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_find_cut_cells, dim3((unsigned)blocks), dim3(256), 0, s, x->d_flag, x->ncell, x->d_status);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  unsigned *hst = x->h_poll + 2 * qesmc::kErrSlots;
  if ((e = hipMemcpyAsync(hst, x->d_status, sizeof(unsigned), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  *cut = (*hst & ST_CUTCELL) != 0;
  return hipSuccess;
}

}  // namespace qesmc

extern "C" {

int qesmc_wall_log_bc(qesmc_ctx *x, float z0, double *ms)
{
  if (!x) return QESMC_ERR_ARG;
  if (!x->have_geom || !x->have_velocity) return QESMC_ERR_STATE;
  WCHK(hipSetDevice(x->device));
  const Dims &d = x->d;
  hipStream_t s = x->stream;
  // cut cells first: their branch of wallLogBC is not built
  {
    bool cut = false;
    WCHK(qesmc::find_cut_cells(x, &cut));
    if (cut) return QESMC_ERR_UNSUPPORTED;
  }
  if (d.nz < 4) return QESMC_OK;  // the listed range k in [1, nz-3] is empty
  // Domain dz: the smallest dz_array entry (Domain.cpp:89,131); dy is dx
  // (getDomainSize, Domain.h:104)
  float dz = x->dz_host[0];
  for (float v : x->dz_host) dz = v < dz ? v : dz;
  const float dx = x->dx, dy = x->dx;
  auto pass = [&](long long cell_off, long long face_off, int guard, int a, int b, int n, float h) {
    const float dist1 = 0.5f * h, dist2 = 1.5f * h;
    return Pass{cell_off, face_off, guard, a, b, n, logf(dist2 / dist1), logf(dist1 / z0)};
  };
  const Pass passes[6] = {
      pass(-d.sz, d.fz, 0, 0, 1, 2, dz),   // wall below (Wall.cpp:681-697)
      pass(d.sz, -d.fz, 0, 0, 1, 2, dz),   // wall above (:700-716)
      pass(-1, 1, 1, 1, 2, 0, dx),         // wall in back, i > 0 (:722-738)
      pass(1, -1, 0, 1, 2, 0, dx),         // wall in front (:742-758)
      pass(-d.sy, d.nx, 2, 0, 2, 1, dy),   // wall on right, j > 0 (:765-781)
      pass(d.sy, -d.nx, 0, 0, 2, 1, dy)};  // wall on left (:784-800)
  const long long total = d.sz * (long long)(d.nz - 3);
  long long blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  WCHK(hipEventRecord(x->ev[0], s));
  for (const Pass &P : passes) {
    hipLaunchKernelGGL(k_wall_log, dim3((unsigned)blocks), dim3(256), 0, s, d, P, x->d_flag, x->d_u, x->d_v, x->d_w);
    WCHK(hipGetLastError());
  }
  WCHK(hipEventRecord(x->ev[1], s));
  WCHK(hipStreamSynchronize(s));
  if (ms) {
    float t = 0.f;
    WCHK(hipEventElapsedTime(&t, x->ev[0], x->ev[1]));
    *ms = t;
  }
  return QESMC_OK;
}

}  // extern "C"

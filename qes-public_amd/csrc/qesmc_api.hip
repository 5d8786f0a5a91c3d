// C ABI of libqesmc: context, device buffers, the solve schedule.
//
// Schedule of one solve (qesmc_run), all on the context's stream:
//   [encode]  wall-mask stencil encoding (once per geometry, or per run on request)
//   [pack]    lambda -> colour-split layout                      (split path)
//   R         divergence of u0                                    Solver_CPU_RB.cpp:56-80
//   loop t    red pass, black pass (gated on err[t] > tol)        Solver_CPU_RB.cpp:95-184
//   [unpack]  lambda -> x-fastest layout                          (split path)
//   mirror, correction, solid zeroing                             Solver_CPU_RB.cpp:154-162,194-284
// The host never waits inside the loop except to poll convergence one chunk
// behind the device, so no iteration pays a device->host round trip
// (the reference copies max_error back every iteration,
// Solver_GPU_GlobalMemory.cu:327).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>

#include "qesmc_internal.h"

hipError_t qesmc_launch_divergence(qesmc_ctx *x, bool split, int alpha1);
hipError_t qesmc_launch_stencil_encode(qesmc_ctx *x);
hipError_t qesmc_launch_split_coefs(qesmc_ctx *x);
hipError_t qesmc_launch_pack(qesmc_ctx *x, bool to_split);
hipError_t qesmc_launch_unpack_selected(qesmc_ctx *x);
hipError_t qesmc_launch_iteration(qesmc_ctx *x, int path, bool masked, int t, int itermax, float omega_red,
                                  float omega_black, float tol);
hipError_t qesmc_launch_finish(qesmc_ctx *x, bool mirror, int alpha1, int alpha2);
hipError_t qesmc_launch_div_linf(qesmc_ctx *x, int physical);
hipError_t qesmc_launch_set_u32(qesmc_ctx *x, unsigned *p, unsigned v);
hipError_t qesmc_launch_fill_u32(qesmc_ctx *x, unsigned *p, long long count, unsigned v);
hipError_t qesmc_launch_rows_summary(qesmc_ctx *x, int launched, float tol);
hipError_t qesmc_fused_prepare(qesmc_ctx *x);

namespace {

int map_hip(hipError_t e)
{
  switch (e) {
  case hipSuccess: return QESMC_OK;
  case hipErrorOutOfMemory: return QESMC_ERR_OOM;
  case hipErrorNoDevice:
  case hipErrorInvalidDevice: return QESMC_ERR_NO_DEVICE;
  default: return QESMC_ERR_HIP;
  }
}

#define HIPCHK(expr)                      \
  do {                                    \
    hipError_t _e = (expr);               \
    if (_e != hipSuccess) return map_hip(_e); \
  } while (0)

template <typename T>
hipError_t dalloc(T **p, long long count)
{
  *p = nullptr;
  if (count <= 0) return hipSuccess;
  return hipMalloc((void **)p, sizeof(T) * (size_t)count);
}

template <typename T>
void dfree(T *&p)
{
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

void free_geometry(qesmc_ctx *x)
{
  for (auto &p : x->d_coef) dfree(p);
  dfree(x->d_flag);
  dfree(x->d_dz);
  dfree(x->d_u0);
  dfree(x->d_v0);
  dfree(x->d_w0);
  dfree(x->d_u);
  dfree(x->d_v);
  dfree(x->d_w);
  dfree(x->d_R);
  dfree(x->d_lam);
  dfree(x->d_grad);
  x->have_velocity = false;
  for (int c = 0; c < 2; ++c) {
    dfree(x->d_lc[c]);
    dfree(x->d_Rc[c]);
    dfree(x->d_mask[c]);
    dfree(x->d_lc2[c]);
    dfree(x->d_R0[c]);
    dfree(x->d_base[c]);
    for (auto &p : x->d_cc[c]) dfree(p);
  }
  dfree(x->d_canon);
  dfree(x->d_canon8);
  x->have_geom = x->have_inflow = false;
  x->geom_checked = false;
  x->d = qesmc::Dims{};
}

int alloc_geometry(qesmc_ctx *x, int nx, int ny, int nz)
{
  qesmc::Dims d{};
  d.nx = nx;
  d.ny = ny;
  d.nz = nz;
  d.NX = nx - 1;
  d.NY = ny - 1;
  d.NZ = nz - 1;
  const int half = (d.NX + 1) / 2;
  // half-rows padded to a multiple of 4 half-cells: the split kernels' vector
  // accesses (up to 4 half-cells per lane) then never straddle a row and stay
  // 16-byte aligned; wider padding only adds traffic (a 400-cell row padded
  // to 64-byte alignment stores 208 half-cells for 200)
  d.HP = (half + 3) / 4 * 4;
  d.sy = d.NX;
  d.sz = (long long)d.NX * d.NY;
  d.fy = nx;
  d.fz = (long long)nx * ny;
  d.splane = (long long)d.NY * d.HP;
  x->ncell = (long long)d.NX * d.NY * d.NZ;
  x->nface = (long long)nx * ny * nz;
  x->nsplit = d.splane * d.NZ;
  x->d = d;
  for (auto &p : x->d_coef) HIPCHK(dalloc(&p, x->ncell));
  HIPCHK(dalloc(&x->d_flag, x->ncell));
  HIPCHK(dalloc(&x->d_dz, nz - 1));
  HIPCHK(dalloc(&x->d_u0, x->nface));
  HIPCHK(dalloc(&x->d_v0, x->nface));
  HIPCHK(dalloc(&x->d_w0, x->nface));
  HIPCHK(dalloc(&x->d_u, x->nface));
  HIPCHK(dalloc(&x->d_v, x->nface));
  HIPCHK(dalloc(&x->d_w, x->nface));
  HIPCHK(dalloc(&x->d_R, x->ncell));
  HIPCHK(dalloc(&x->d_lam, x->ncell));
  for (int c = 0; c < 2; ++c) {
    HIPCHK(dalloc(&x->d_lc[c], x->nsplit));
    HIPCHK(dalloc(&x->d_Rc[c], x->nsplit));
    HIPCHK(dalloc(&x->d_mask[c], x->nsplit));
  }
  HIPCHK(dalloc(&x->d_canon, 6LL * d.NZ));
  HIPCHK(hipMemsetAsync(x->d_lam, 0, sizeof(float) * x->ncell, x->stream));
  return QESMC_OK;
}

// Build the wall-mask encoding from the resident natural coefficients and
// decide whether the masked stencil reproduces them exactly.
int encode_stencil(qesmc_ctx *x)
{
  HIPCHK(hipMemsetAsync(x->d_status, 0, sizeof(unsigned), x->stream));
  HIPCHK(hipMemsetAsync(x->d_canon, 0, sizeof(unsigned) * 6 * x->d.NZ, x->stream));
  HIPCHK(qesmc_launch_stencil_encode(x));
  unsigned *hst = x->h_poll + 2 * qesmc::kErrSlots;
  HIPCHK(hipMemcpyAsync(hst, x->d_status, sizeof(unsigned), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  const unsigned st = *hst;
  if (st & qesmc::ST_ZERO_DIAG) return QESMC_ERR_ZERO_DIAG;
  x->compressible = !(st & qesmc::ST_INCOMPRESSIBLE);
  if (!x->compressible) {
    for (int c = 0; c < 2; ++c)
      for (auto &p : x->d_cc[c])
        if (!p) HIPCHK(dalloc(&p, x->nsplit));
    HIPCHK(qesmc_launch_split_coefs(x));
  }
  x->geom_checked = true;
  return QESMC_OK;
}

}  // namespace

namespace qesmc {

int geometry_reserve(qesmc_ctx *x, int nx, int ny, int nz)
{
  if (x->d.nx != nx || x->d.ny != ny || x->d.nz != nz || !x->d_lam) {
    HIPCHK(hipStreamSynchronize(x->stream));
    free_geometry(x);
    int rc = alloc_geometry(x, nx, ny, nz);
    if (rc != QESMC_OK) {
      free_geometry(x);
      return rc;
    }
  }
  return QESMC_OK;
}

int geometry_commit(qesmc_ctx *x)
{
  x->have_geom = false;
  x->have_velocity = false;
  int rc = encode_stencil(x);
  if (rc != QESMC_OK) return rc;
  x->have_geom = true;
  return QESMC_OK;
}

}  // namespace qesmc

extern "C" {

const char *qesmc_status_string(int s)
{
  switch (s) {
  case QESMC_OK: return "ok";
  case QESMC_ERR_ARG: return "invalid argument";
  case QESMC_ERR_NO_DEVICE: return "no HIP device";
  case QESMC_ERR_OOM: return "device out of memory";
  case QESMC_ERR_HIP: return "HIP runtime error";
  case QESMC_ERR_ZERO_DIAG: return "zero stencil diagonal (e+f+g+h+m+n == 0 in an updated cell)";
  case QESMC_ERR_NONFINITE: return "lambda became non-finite";
  case QESMC_ERR_STATE: return "inputs not set";
  case QESMC_ERR_UNSUPPORTED: return "unsupported option";
  case QESMC_ERR_IO: return "file missing, unwritable or not a NetCDF classic file";
  case QESMC_ERR_LIMIT: return "beyond the fixed capacity of an entry point that allocates nothing";
  default: return "unknown status";
  }
}

void qesmc_default_params(qesmc_params *p)
{
  if (!p) return;
  p->omega = 1.78f;
  p->tol = 1e-9f;
  p->itermax = 500;
  p->alpha1 = 1;
  p->alpha2 = 1;
  p->path = QESMC_PATH_AUTO;
  p->check_every = 64;
  p->flags = 0;
  p->tile_k = 0;
  p->tile_v = 0;
  p->wg_per_cu = 0;
  p->slab_planes = 0;
  p->stop_extend = 0.f;
  p->refine_every = 0;
  p->chebyshev_rho = 0.0;
  p->chebyshev_start = 0;
}

int qesmc_create(qesmc_ctx **out, int device)
{
  if (!out) return QESMC_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return QESMC_ERR_NO_DEVICE;
  if (device < 0 || device >= n) return QESMC_ERR_NO_DEVICE;
  HIPCHK(hipSetDevice(device));
  qesmc_ctx *x = new (std::nothrow) qesmc_ctx();
  if (!x) return QESMC_ERR_OOM;
  x->device = device;
  int rc = QESMC_OK;
  auto fail = [&](hipError_t e) {
    rc = map_hip(e);
    return e != hipSuccess;
  };
  if (fail(hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking))) goto bad;
  for (auto &ev : x->ev)
    if (fail(hipEventCreate(&ev))) goto bad;
  {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
      x->n_cu = ncu;
  }
  // the no-allocation entry point's per-device scratch address (errors are
  // reported by qesmc_run_static itself, which repeats the lookup)
  (void)qesmc::static_scratch_init(device);
  if (fail(hipHostMalloc((void **)&x->h_poll, (2 * qesmc::kErrSlots + 8) * sizeof(unsigned), hipHostMallocDefault))) goto bad;
  if (fail(dalloc(&x->d_status, qesmc::kStatusWords))) goto bad;
  if (fail(dalloc(&x->d_norm, 1))) goto bad;
  *out = x;
  return QESMC_OK;
bad:
  qesmc_destroy(x);
  return rc == QESMC_OK ? QESMC_ERR_HIP : rc;
}

int qesmc_destroy(qesmc_ctx *x)
{
  if (!x) return QESMC_OK;
  (void)hipSetDevice(x->device);
  if (x->stream) (void)hipStreamSynchronize(x->stream);
  free_geometry(x);
  dfree(x->d_err);
  dfree(x->d_status);
  dfree(x->d_norm);
  dfree(x->d_cur);
  dfree(x->d_probe);
  dfree(x->d_probe_val);
  if (x->h_poll) (void)hipHostFree(x->h_poll);
  for (auto &ev : x->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (x->stream) (void)hipStreamDestroy(x->stream);
  delete x;
  return QESMC_OK;
}

int qesmc_set_geometry(qesmc_ctx *x, const qesmc_grid *grid, const qesmc_coeffs *cf)
{
  if (!x || !grid || !cf || !grid->dz_array) return QESMC_ERR_ARG;
  if (!cf->e || !cf->f || !cf->g || !cf->h || !cf->m || !cf->n || !cf->icellflag) return QESMC_ERR_ARG;
  if (grid->nx < 2 || grid->ny < 2 || grid->nz < 3) return QESMC_ERR_ARG;
  if (!(grid->dx > 0.f) || !(grid->dy > 0.f)) return QESMC_ERR_ARG;
  const long long ncell = (long long)(grid->nx - 1) * (grid->ny - 1) * (grid->nz - 1);
  if (ncell > (1LL << 31) - 1) return QESMC_ERR_UNSUPPORTED;  // per-plane int math
  HIPCHK(hipSetDevice(x->device));
  {
    int rc = qesmc::geometry_reserve(x, grid->nx, grid->ny, grid->nz);
    if (rc != QESMC_OK) return rc;
  }
  x->dx = grid->dx;
  x->dy = grid->dy;
  x->dz_host.assign(grid->dz_array, grid->dz_array + (grid->nz - 1));
  const size_t cb = sizeof(float) * (size_t)x->ncell;
  const float *src[6] = {cf->e, cf->f, cf->g, cf->h, cf->m, cf->n};
  for (int a = 0; a < 6; ++a) HIPCHK(hipMemcpyAsync(x->d_coef[a], src[a], cb, hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipMemcpyAsync(x->d_flag, cf->icellflag, sizeof(int) * (size_t)x->ncell, hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipMemcpyAsync(x->d_dz, grid->dz_array, sizeof(float) * (grid->nz - 1), hipMemcpyHostToDevice, x->stream));
  return qesmc::geometry_commit(x);
}

int qesmc_set_inflow(qesmc_ctx *x, const float *u0, const float *v0, const float *w0)
{
  if (!x || !u0 || !v0 || !w0) return QESMC_ERR_ARG;
  if (!x->d_u0) return QESMC_ERR_STATE;
  HIPCHK(hipSetDevice(x->device));
  const size_t fb = sizeof(float) * (size_t)x->nface;
  HIPCHK(hipMemcpyAsync(x->d_u0, u0, fb, hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipMemcpyAsync(x->d_v0, v0, fb, hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipMemcpyAsync(x->d_w0, w0, fb, hipMemcpyHostToDevice, x->stream));
  x->have_inflow = true;
  return QESMC_OK;
}

int qesmc_set_lambda(qesmc_ctx *x, const float *lambda)
{
  if (!x) return QESMC_ERR_ARG;
  if (!x->d_lam) return QESMC_ERR_STATE;
  HIPCHK(hipSetDevice(x->device));
  if (lambda)
    HIPCHK(hipMemcpyAsync(x->d_lam, lambda, sizeof(float) * (size_t)x->ncell, hipMemcpyHostToDevice, x->stream));
  else
    HIPCHK(hipMemsetAsync(x->d_lam, 0, sizeof(float) * (size_t)x->ncell, x->stream));
  return QESMC_OK;
}

}  // extern "C"

namespace qesmc {

float row_max(const unsigned *row)
{
  float mx = 0.f;
  for (int q = 0; q < kErrSlots; ++q) {
    float v;
    std::memcpy(&v, row + q, 4);
    mx = v > mx ? v : mx;
  }
  return mx;
}

// Validation, optional re-encoding, tiling, convergence rows, lambda packing
// and R, on the context's current stream (ev[0] .. ev[1]).
int run_prepare(qesmc_ctx *x, const qesmc_params *pp, RunState &st)
{
  if (!x || !pp) return QESMC_ERR_ARG;
  if (!x->have_geom || !x->have_inflow) return QESMC_ERR_STATE;
  qesmc_params &p = st.p;
  p = *pp;
  if (p.itermax < 0) return QESMC_ERR_ARG;
  if (p.check_every < 1) p.check_every = 64;
  if (p.path != QESMC_PATH_AUTO && p.path != QESMC_PATH_PLAIN && p.path != QESMC_PATH_SPLIT &&
      p.path != QESMC_PATH_FUSED)
    return QESMC_ERR_ARG;
  if (p.tile_k < 0 || p.wg_per_cu < 0 || p.slab_planes < -1) return QESMC_ERR_ARG;
  if (!(p.stop_extend >= 0.f && p.stop_extend <= 1000.f)) return QESMC_ERR_ARG;  // also refuses NaN
  if (p.tile_v != 0 && p.tile_v != 1 && p.tile_v != 2 && p.tile_v != 4) return QESMC_ERR_ARG;
  if (p.refine_every < 0) return QESMC_ERR_ARG;
  if (!(p.chebyshev_rho >= 0.0 && p.chebyshev_rho < 1.0) || p.chebyshev_start < 0) return QESMC_ERR_ARG;
  st.path = (p.path == QESMC_PATH_PLAIN || p.path == QESMC_PATH_FUSED) ? p.path : QESMC_PATH_SPLIT;
  // the refinement runs on the split kernel and on the march
  if (p.refine_every > 0 && st.path == QESMC_PATH_PLAIN) return QESMC_ERR_UNSUPPORTED;
  HIPCHK(hipSetDevice(x->device));
  if (p.flags & QESMC_RUN_REENCODE) {
    int rc = encode_stencil(x);
    if (rc != QESMC_OK) return rc;
  }
  st.masked = x->compressible;
  // QESMC_PATH_AUTO: the two-colour march where the colour arrays (lambda and
  // R of both colours, the masks) exceed MI355X's 256 MB Infinity Cache, so
  // that each iteration streams them from HBM, and where it applies (wall
  // mask, 32-bit offsets, its buffers available without allocating in a
  // no-allocation run); the split kernel otherwise.
  // Measured: 568 against 626 us per iteration on 1000 x 1000 x 150; 64
  // against 58 us on 400 x 400 x 120, whose arrays stay in the cache.
  if (p.path == QESMC_PATH_AUTO && st.masked &&
      x->nsplit < (1LL << 31) && (double)x->nsplit * 2.0 * (4.0 + 4.0 + 1.0) > 256.0 * (1 << 20) &&
      (!x->no_alloc_run || (x->d_lc2[0] && x->d_lc2[1] && x->d_cur && x->d_canon8)))
    st.path = QESMC_PATH_FUSED;
  if (st.path == QESMC_PATH_FUSED && !st.masked) return QESMC_ERR_UNSUPPORTED;
  // the fused march addresses its arrays with 32-bit element offsets
  if (st.path == QESMC_PATH_FUSED && x->nsplit >= (1LL << 31)) return QESMC_ERR_UNSUPPORTED;
  x->kplanes = p.tile_k > 0 ? p.tile_k : x->kplanes_auto;
  x->tile_v = p.tile_v > 0 ? p.tile_v : x->tile_v_auto;
  if (x->kplanes == 0 || x->tile_v == 0) {
    // Untuned default: the most work per wave tile (planes x half-cells per
    // lane, at most 4 x 2) that still leaves >= 6 wave tiles per CU; small
    // domains are latency-bound and need the parallelism more than long
    // k-marches.  Measured on MI355X: 4 x 2 matches 8 x 2 on 400x400x120 and
    // 1000x1000x150 and is 25 % faster on 256x256x96.
    const long long flat = (long long)(x->d.NY - 2) * x->d.HP;
    auto tiles = [&](int kb, int vq) {
      return ((flat + 64LL * vq - 1) / (64LL * vq)) * ((x->d.NZ - 2 + kb - 1) / kb);
    };
    static const int order[][2] = {{4, 2}, {2, 2}, {4, 1}, {2, 1}, {1, 2}, {1, 1}};
    int bk = 0, bv = 0;
    long long most = -1;
    for (const auto &o : order) {
      if ((x->kplanes && o[0] != x->kplanes) || (x->tile_v && o[1] != x->tile_v)) continue;
      const long long n = tiles(o[0], o[1]);
      if (n >= 6LL * x->n_cu) {
        bk = o[0];
        bv = o[1];
        break;
      }
      if (n > most) {
        most = n;
        bk = o[0];
        bv = o[1];
      }
    }
    if (!x->kplanes) x->kplanes = bk ? bk : 8;
    if (!x->tile_v) x->tile_v = bv ? bv : 1;
  }
  // untuned: 4 resident work-groups per CU for two or four half-cells per
  // lane, 5 for one (fewer, longer-lived work-groups beat full occupancy on
  // every measured size but the HBM-bound 1000x1000x150, where they lose 1 %)
  x->wg_per_cu = p.wg_per_cu > 0 ? p.wg_per_cu : x->wg_per_cu_auto > 0 ? x->wg_per_cu_auto : x->tile_v >= 2 ? 4 : 5;
  // k-slab schedule of the split path (results never depend on it)
  x->slab_planes = p.slab_planes > 0 ? p.slab_planes : p.slab_planes < 0 ? 0 : x->slab_planes_auto;
  if (st.path == QESMC_PATH_PLAIN) {
    // the classic layout's own default: 8 planes per work-group (its
    // measured best; the split path's choice above does not apply to it)
    x->kplanes = p.tile_k > 0 ? p.tile_k : 8;
  }
  if (st.path == QESMC_PATH_FUSED) {
    // one wave per column tile of tile_k planes (4 or 8; tile_v and
    // wg_per_cu unused)
    x->kplanes = p.tile_k > 0 ? p.tile_k : 8;
    x->tile_v = 1;
    x->wg_per_cu = 0;
    for (int c = 0; c < 2; ++c)
      if (!x->d_lc2[c]) HIPCHK(dalloc(&x->d_lc2[c], x->nsplit));
    if (!x->d_cur) HIPCHK(dalloc(&x->d_cur, 1));
    if (!x->d_canon8) HIPCHK(dalloc(&x->d_canon8, 8LL * x->d.NZ));
    HIPCHK(qesmc_fused_prepare(x));
  }
  if (p.refine_every > 0) {
    // base lambda and the run's R, besides the correction and the residual
    for (int c = 0; c < 2; ++c) {
      if (!x->d_base[c]) HIPCHK(dalloc(&x->d_base[c], x->nsplit));
      if (!x->d_R0[c]) HIPCHK(dalloc(&x->d_R0[c], x->nsplit));
    }
  }

  const long long rows = (long long)p.itermax + 1;
  if (x->err_cap < rows) {
    dfree(x->d_err);
    x->err_cap = 0;
    const long long cap = rows > 1024 ? rows : 1024;
    HIPCHK(dalloc(&x->d_err, cap * kErrSlots));
    x->err_cap = (int)cap;
  }
  hipStream_t s = x->stream;
  HIPCHK(hipEventRecord(x->ev[0], s));
  HIPCHK(qesmc_launch_fill_u32(x, x->d_status, 1, 0u));
  // the stop rule: no end recorded yet, and the extension factor (0 for the
  // reference's rule) that every kernel's gate reads
  {
    unsigned ext;
    std::memcpy(&ext, &p.stop_extend, 4);
    HIPCHK(qesmc_launch_fill_u32(x, x->d_status + kStopAt, 1, 0xffffffffu));
    HIPCHK(qesmc_launch_fill_u32(x, x->d_status + kStopExtend, 1, ext));
  }
  HIPCHK(qesmc_launch_fill_u32(x, x->d_err, rows * kErrSlots, 0u));
  HIPCHK(qesmc_launch_set_u32(x, x->d_err, 0x3f800000u));  // row 0: 1.0f
  if (st.path != QESMC_PATH_PLAIN) HIPCHK(qesmc_launch_pack(x, true));
  if (st.path == QESMC_PATH_FUSED) {
    // both buffer sets start equal: cells that are never relaxed keep their
    // values in both, and iteration 0 reads set 0
    for (int c = 0; c < 2; ++c)
      HIPCHK(hipMemcpyAsync(x->d_lc2[c], x->d_lc[c], sizeof(float) * (size_t)x->nsplit, hipMemcpyDeviceToDevice, s));
    HIPCHK(qesmc_launch_set_u32(x, x->d_cur, 0u));
  }
  HIPCHK(qesmc_launch_divergence(x, st.path != QESMC_PATH_PLAIN, p.alpha1));
  if (p.refine_every > 0) HIPCHK(refine_setup(x, st.masked, st.path == QESMC_PATH_FUSED));
  HIPCHK(hipEventRecord(x->ev[1], s));
  st.first_runs = p.itermax > 0 && (1.0f > p.tol);
  return QESMC_OK;
}

// Unpack, mirror, correction, zeroing (ev[2] .. ev[3]), then the iteration
// count from the convergence rows: iteration t ran iff max(row t) > tol.
int run_finish(qesmc_ctx *x, const RunState &st, int launched, qesmc_result *res)
{
  const qesmc_params &p = st.p;
  hipStream_t s = x->stream;
  HIPCHK(hipEventRecord(x->ev[2], s));
  if (st.path == QESMC_PATH_SPLIT) HIPCHK(qesmc_launch_pack(x, false));
  if (st.path == QESMC_PATH_FUSED) HIPCHK(qesmc_launch_unpack_selected(x));
  HIPCHK(qesmc_launch_finish(x, st.first_runs, p.alpha1, p.alpha2));
  HIPCHK(hipEventRecord(x->ev[3], s));
  // iteration t ran iff max(row t) > tol: the count and the last residual are
  // reduced on the device, so only four words come back
  HIPCHK(qesmc_launch_rows_summary(x, launched, p.tol));
  unsigned *sw = x->h_poll + 2 * kErrSlots;
  HIPCHK(hipMemcpyAsync(sw, x->d_status, 4 * sizeof(unsigned), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  x->have_velocity = true;
  const int count = (int)sw[2];
  float last;
  std::memcpy(&last, sw + 3, 4);
  const int rc = (sw[0] & ST_NONFINITE) ? QESMC_ERR_NONFINITE : QESMC_OK;
  if (res) {
    float ms_total = 0.f, ms_sor = 0.f;
    (void)hipEventElapsedTime(&ms_total, x->ev[0], x->ev[3]);
    (void)hipEventElapsedTime(&ms_sor, x->ev[1], x->ev[2]);
    res->iterations = count;
    res->path = st.path;
    res->compressed = (st.path != QESMC_PATH_PLAIN && st.masked) ? 1 : 0;
    res->status = rc;
    res->final_max_dlambda = last;
    res->ms_total = ms_total;
    res->ms_sor = ms_sor;
    const double lups = (double)(x->d.nx - 3) * (x->d.ny - 3) * (x->d.nz - 3) * count;
    res->glups = ms_sor > 0.f ? lups / (ms_sor * 1e6) : 0.0;
    res->tile_k = x->kplanes;
    res->tile_v = x->tile_v;
    res->wg_per_cu = x->wg_per_cu;
    res->reserved = 0;
  }
  return rc;
}

}  // namespace qesmc

extern "C" {

int qesmc_run(qesmc_ctx *x, const qesmc_params *pp, qesmc_result *res) { return qesmc::run_solve(x, pp, res); }

}  // extern "C"

namespace qesmc {

// The iteration loop of a prepared run, with convergence polling one chunk
// behind the device; *launched counts the iterations queued.
int run_loop(qesmc_ctx *x, const RunState &st, int *launched_out)
{
  const qesmc_params &p = st.p;
  hipStream_t s = x->stream;
  int rc = QESMC_OK;
  int launched = 0;
  if (st.first_runs) {
    // Poll convergence one chunk behind the device: after chunk n is queued,
    // wait for chunk n-1's row (already computed or about to be).  The gate
    // inside the kernels makes iterations after convergence no-ops, so the
    // result never depends on when the host notices.
    int chunk = 0;
    bool stop = false;
    // the cyclic Chebyshev schedule, advanced past the passes of earlier calls;
    // once a factor repeats the previous one the recurrence has reached its
    // fixed point (Young's optimum) and every later factor is that value, so
    // the advance stops there
    const double rho2 = p.chebyshev_rho * p.chebyshev_rho;
    long long pass = 0;
    double w_prev = 1.0;
    bool fixed = false;
    auto next_omega = [&]() -> float {
      if (!fixed) {
        const double w =
            pass == 0 ? 1.0 : pass == 1 ? 1.0 / (1.0 - 0.5 * rho2) : 1.0 / (1.0 - 0.25 * rho2 * w_prev);
        fixed = pass >= 2 && w == w_prev;
        ++pass;
        w_prev = w;
      }
      return (float)w_prev;
    };
    if (p.chebyshev_rho > 0.0)
      for (long long h = 0; h < 2LL * p.chebyshev_start && !fixed; ++h) (void)next_omega();
    for (int t = 0; t < p.itermax && !stop; ++t) {
      float w_red = p.omega, w_black = p.omega;
      if (p.chebyshev_rho > 0.0) {
        w_red = next_omega();
        w_black = next_omega();
      }
      hipError_t e = qesmc_launch_iteration(x, st.path, st.masked, t, p.itermax, w_red, w_black, p.tol);
      if (e != hipSuccess) {
        rc = map_hip(e);
        break;
      }
      ++launched;
      if (p.refine_every > 0 && (t + 1) % p.refine_every == 0 && t + 1 < p.itermax &&
          (e = refine_refresh(x, st.masked, st.path == QESMC_PATH_FUSED)) != hipSuccess) {
        rc = map_hip(e);
        break;
      }
      if ((t + 1) % p.check_every == 0 && t + 1 < p.itermax) {
        const int slot = chunk & 1;
        unsigned *dst = x->h_poll + slot * qesmc::kErrSlots;
        if (chunk >= 2 && (e = hipEventSynchronize(x->ev[4 + slot])) != hipSuccess) { rc = map_hip(e); break; }
        e = hipMemcpyAsync(dst, x->d_err + (long long)(t + 1) * qesmc::kErrSlots,
                           sizeof(unsigned) * qesmc::kErrSlots, hipMemcpyDeviceToHost, s);
        // the production rule's recorded end (all ones until fixed)
        unsigned *end_h = x->h_poll + 2 * qesmc::kErrSlots + 4 + slot;
        if (e == hipSuccess)
          e = hipMemcpyAsync(end_h, x->d_status + qesmc::kStopAt, sizeof(unsigned), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipEventRecord(x->ev[4 + slot], s);
        if (e != hipSuccess) { rc = map_hip(e); break; }
        if (chunk >= 1) {
          const int prev = slot ^ 1;
          if ((e = hipEventSynchronize(x->ev[4 + prev])) != hipSuccess) { rc = map_hip(e); break; }
          if (p.stop_extend > 0.f) {
            // every iteration from the recorded end on is a no-op
            if (x->h_poll[2 * qesmc::kErrSlots + 4 + prev] <= (unsigned)(t + 1)) stop = true;
          } else if (!(qesmc::row_max(x->h_poll + prev * qesmc::kErrSlots) > p.tol)) {
            stop = true;
          }
        }
        ++chunk;
      }
    }
    if (rc != QESMC_OK) (void)hipStreamSynchronize(s);
  }
  *launched_out = launched;
  return rc;
}

// One resident solve: preparation, the iteration loop and the finish
// (qesmc_run).
int run_solve(qesmc_ctx *x, const qesmc_params *pp, qesmc_result *res)
{
  qesmc::RunState st;
  int rc = qesmc::run_prepare(x, pp, st);
  if (rc != QESMC_OK) return rc;
  int launched = 0;
  rc = run_loop(x, st, &launched);
  if (rc != QESMC_OK) return rc;
  if (st.p.refine_every > 0) {
    hipError_t e = refine_fold_final(x, st.path == QESMC_PATH_FUSED);
    if (e != hipSuccess) return map_hip(e);
  }
  return qesmc::run_finish(x, st, launched, res);
}

}  // namespace qesmc

extern "C" {

int qesmc_get_velocity(qesmc_ctx *x, float *u, float *v, float *w)
{
  if (!x || !u || !v || !w) return QESMC_ERR_ARG;
  if (!x->d_u) return QESMC_ERR_STATE;
  HIPCHK(hipSetDevice(x->device));
  const size_t fb = sizeof(float) * (size_t)x->nface;
  HIPCHK(hipMemcpyAsync(u, x->d_u, fb, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipMemcpyAsync(v, x->d_v, fb, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipMemcpyAsync(w, x->d_w, fb, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return QESMC_OK;
}

int qesmc_get_lambda(qesmc_ctx *x, float *lambda)
{
  if (!x || !lambda) return QESMC_ERR_ARG;
  if (!x->d_lam) return QESMC_ERR_STATE;
  HIPCHK(hipSetDevice(x->device));
  HIPCHK(hipMemcpyAsync(lambda, x->d_lam, sizeof(float) * (size_t)x->ncell, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return QESMC_OK;
}

int qesmc_divergence_linf(qesmc_ctx *x, int physical, double *out)
{
  if (!x || !out) return QESMC_ERR_ARG;
  if (!x->have_geom || !x->d_u) return QESMC_ERR_STATE;
  HIPCHK(hipSetDevice(x->device));
  HIPCHK(qesmc_launch_div_linf(x, physical));
  HIPCHK(hipMemcpyAsync(out, x->d_norm, sizeof(double), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return QESMC_OK;
}

int qesmc_solve(qesmc_ctx *x, const qesmc_grid *grid, const qesmc_coeffs *cf, qesmc_fields *fl,
                float *lambda_inout, const qesmc_params *p, qesmc_result *res)
{
  if (!x || !grid || !cf || !fl || !p) return QESMC_ERR_ARG;
  if (!fl->u0 || !fl->v0 || !fl->w0 || !fl->u || !fl->v || !fl->w) return QESMC_ERR_ARG;
  int rc = qesmc_set_geometry(x, grid, cf);
  if (rc != QESMC_OK) return rc;
  if ((rc = qesmc_set_inflow(x, fl->u0, fl->v0, fl->w0)) != QESMC_OK) return rc;
  if ((rc = qesmc_set_lambda(x, lambda_inout)) != QESMC_OK) return rc;
  const int run_rc = qesmc_run(x, p, res);
  if (run_rc != QESMC_OK && run_rc != QESMC_ERR_NONFINITE) return run_rc;
  if ((rc = qesmc_get_velocity(x, fl->u, fl->v, fl->w)) != QESMC_OK) return rc;
  if (lambda_inout && (rc = qesmc_get_lambda(x, lambda_inout)) != QESMC_OK) return rc;
  return run_rc;
}

}  // extern "C"

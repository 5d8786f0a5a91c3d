// Batched solves: many independent problems (uneven sizes, one context each)
// relaxed by one kernel launch per colour pass, and the tile autotuner.
//
// Every instance keeps its own colour-split arrays, stencil encoding and
// convergence rows; a device table describes them.  Wave tiles are numbered
// over all instances (instance-major, then plane chunk, then position); a
// wave finds its instance by binary search of the table's tile offsets,
// checks that instance's gate (its iteration counter and its previous max
// |dlambda|), relaxes the tile exactly as the single-problem kernel does and
// publishes its max into the instance's striped slots.  Converged instances
// cost one gate read per tile, so a batch behaves as if each problem were
// solved alone.  Iterations can be replayed as a captured HIP graph: the
// iteration number then comes from a device counter advanced by the graph.
#include <algorithm>
#include <cstring>
#include <vector>

#include "qesmc_sor_tile.h"

hipError_t qesmc_launch_set_u32(qesmc_ctx *x, unsigned *p, unsigned v);

namespace qesmc {

namespace {

constexpr int kXcdsB = 8;

struct BatchItem {
  Dims d;
  float *lc[2];
  const float *Rc[2];
  const unsigned char *mask[2];
  const unsigned *canon;
  const float *cc[2][6];
  unsigned *err;
  unsigned *status;
  int masked;
  int kb;
  int tiles_per_plane;
  int pad;
  long long tile_begin;  // first global wave tile of this instance
};

__device__ inline float wave_max_b(float v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// t = t_base[0] + t_off when t_base is given (graph replay), else t_off.
//
// Tiles are assigned as in the single-problem kernel (XCD x owns the x-th
// eighth of the global tile range; the waves of an XCD stride through it
// together, so concurrently active waves work on neighbouring tiles).  A
// wave's tile indices only increase, so its instance is found by advancing
// from the previous one; the instance's gate is read and its max|dlambda| is
// published once per instance the wave visits, not per tile.
template <bool ALL_MASKED, int VQ>
__global__ __launch_bounds__(256) void k_sor_batch(const BatchItem *__restrict__ items, int n, long long total,
                                                   int colour, const int *t_base, int t_off, int itermax,
                                                   float omega, float one_m_omega, float tol)
{
  const int lane = threadIdx.x & 63;
  // the wave index is uniform across the wave: say so, so that everything
  // derived from the tile index (the instance and its fields) lives in
  // scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int t = (t_base ? t_base[0] : 0) + t_off;
  if (t >= itermax) return;
  const int xcd = blockIdx.x % kXcdsB;
  const int local = blockIdx.x / kXcdsB;
  const int nblk = (gridDim.x - xcd + kXcdsB - 1) / kXcdsB;
  const long long per = (total + kXcdsB - 1) / kXcdsB;
  const long long tbeg = xcd * per;
  const long long tend = tbeg + per < total ? tbeg + per : total;
  const long long step = (long long)nblk * 4;
  const int slot = (blockIdx.x * 4 + wave) % kErrSlots;
  // the current instance, kept in (wave-uniform) registers
  int cur = -1;
  long long cbeg = 0, cend = 0;
  bool open = false, masked = true;
  Dims d{};
  int kb = 1, tpp = 1;
  QESMC_GLOBAL float *own = nullptr;
  const QESMC_GLOBAL float *oth = nullptr, *Rc = nullptr;
  const QESMC_GLOBAL unsigned char *mk = nullptr;
  const QESMC_CONSTANT unsigned *canon = nullptr;
  unsigned *status = nullptr;
  float dl = 0.f;
  for (long long wt = tbeg + (long long)local * 4 + wave; wt < tend; wt += step) {
    if (wt >= cend) {
      // publish the finished instance, then advance to the one holding wt
      if (cur >= 0) {
        const float wm = wave_max_b(dl);
        if (lane == 0 && wm > 0.f)
          atomicMax(items[cur].err + (long long)(t + 1) * kErrSlots + slot, __float_as_uint(wm));
      }
      dl = 0.f;
      int nx;
      if (cur < 0) {
        int lo = 0, hi = n - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (items[mid].tile_begin <= wt) lo = mid;
          else hi = mid - 1;
        }
        nx = lo;
      } else {
        nx = cur + 1;
        while (nx + 1 < n && items[nx + 1].tile_begin <= wt) ++nx;
      }
      cur = nx;
      const BatchItem &it = items[cur];
      cbeg = it.tile_begin;
      cend = cur + 1 < n ? items[cur + 1].tile_begin : total;
      open = gate_pass(wave_max_b(__uint_as_float(it.err[(long long)t * kErrSlots + lane])), t, tol, it.status);
      d = it.d;
      kb = it.kb;
      tpp = it.tiles_per_plane;
      own = as_global(it.lc[colour]);
      oth = as_global(it.lc[colour ^ 1]);
      Rc = as_global(it.Rc[colour]);
      mk = as_global(it.mask[colour]);
      canon = as_constant(it.canon);
      status = it.status;
      masked = ALL_MASKED || it.masked != 0;
    }
    if (!open) continue;
    const long long lt = wt - cbeg;
    const int kc = (int)(lt / tpp);
    const int pt = (int)(lt - (long long)kc * tpp);
    const int p = pt * (64 * VQ) + lane * VQ;
    const int jr = p / d.HP;
    const int j = jr + 1;
    const int q = p - jr * d.HP;
    if (j > d.NY - 2) continue;
    const int k0 = 1 + kc * kb;
    const int k1 = k0 + kb < d.NZ - 1 ? k0 + kb : d.NZ - 1;
    // generic pointers for the call; inlined, the compiler still sees that
    // they point to global memory
    float *o_ = (float *)own;
    const float *oth_ = (const float *)oth, *Rc_ = (const float *)Rc;
    const unsigned char *mk_ = (const unsigned char *)mk;
    if (masked) {
      dl = sor_tile<true, VQ>(d, colour, t, j, q, k0, k1, o_, oth_, Rc_, mk_, canon, nullptr, nullptr, nullptr,
                              nullptr, nullptr, nullptr, omega, one_m_omega, status, dl);
    } else if (!ALL_MASKED) {
      const float *const *cc = items[cur].cc[colour];
      dl = sor_tile<false, VQ>(d, colour, t, j, q, k0, k1, o_, oth_, Rc_, mk_, canon,
                              (const float *)as_global(cc[0]), (const float *)as_global(cc[1]),
                              (const float *)as_global(cc[2]), (const float *)as_global(cc[3]),
                              (const float *)as_global(cc[4]), (const float *)as_global(cc[5]), omega, one_m_omega,
                              status, dl);
    }
  }
  // last instance of each wave: combine the work-group's four waves first, so
  // that each instance gets one atomic per work-group (device-scope atomics on
  // a few slots serialise at the memory side)
  __shared__ int wcur[4];
  __shared__ float wmx[4];
  const float wm = cur >= 0 ? wave_max_b(dl) : 0.f;
  if (lane == 0) {
    wcur[wave] = cur;
    wmx[wave] = wm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int a = 0; a < 4; ++a) {
      if (wcur[a] < 0) continue;
      bool seen = false;
      for (int b = 0; b < a; ++b) seen = seen || wcur[b] == wcur[a];
      if (seen) continue;
      float m = wmx[a];
      for (int b = a + 1; b < 4; ++b)
        if (wcur[b] == wcur[a]) m = fmaxf(m, wmx[b]);
      if (m > 0.f)
        atomicMax(items[wcur[a]].err + (long long)(t + 1) * kErrSlots + blockIdx.x % kErrSlots, __float_as_uint(m));
    }
  }
}

__global__ void k_advance(int *t_base, int by) { t_base[0] += by; }

// number of instances whose row t is still above tol (t = t_base[0] + t_off)
__global__ __launch_bounds__(64) void k_count_active(const BatchItem *items, int n, const int *t_base, int t_off,
                                                     int itermax, float tol, int *out)
{
  const int t = (t_base ? t_base[0] : 0) + t_off;
  int c = 0;
  if (t < itermax)
    for (int i = threadIdx.x; i < n; i += 64) {
      float mx = 0.f;
      for (int q = 0; q < kErrSlots; ++q) mx = fmaxf(mx, __uint_as_float(items[i].err[(long long)t * kErrSlots + q]));
      // the gate of iteration t (gate_pass, without recording the end)
      const unsigned *st = items[i].status;
      const float e = __uint_as_float(st[kStopExtend]);
      bool open;
      if (!(e > 0.f)) open = mx > tol;
      else if ((unsigned)t >= *(volatile const unsigned *)(st + kStopAt)) open = false;
      else open = mx > tol || (unsigned)t + (unsigned)ceilf(e * (float)t) > (unsigned)t;
      c += open ? 1 : 0;
    }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
  if (threadIdx.x == 0) out[0] = c;
}

int map_err_b(hipError_t e)
{
  if (e == hipSuccess) return QESMC_OK;
  if (e == hipErrorOutOfMemory) return QESMC_ERR_OOM;
  return QESMC_ERR_HIP;
}

#define BCHK(expr)                                 \
  do {                                             \
    hipError_t _e = (expr);                        \
    if (_e != hipSuccess) return map_err_b(_e);    \
  } while (0)

}  // namespace

}  // namespace qesmc

using namespace qesmc;

extern "C" {

int qesmc_run_batch(qesmc_ctx **ctxs, int n, const qesmc_params *params, qesmc_result *results, int use_graph,
                    double *ms_loop)
{
  if (!ctxs || n < 1 || !params) return QESMC_ERR_ARG;
  for (int i = 0; i < n; ++i) {
    if (!ctxs[i]) return QESMC_ERR_ARG;
    if (ctxs[i]->device != ctxs[0]->device) return QESMC_ERR_UNSUPPORTED;
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == ctxs[i]) return QESMC_ERR_ARG;
  }
  if (params->path == QESMC_PATH_PLAIN || params->path == QESMC_PATH_FUSED) return QESMC_ERR_UNSUPPORTED;
  // the batched loop has no refinement refresh
  if (params->refine_every != 0 || params->chebyshev_rho != 0.0) return QESMC_ERR_UNSUPPORTED;
  // One tiling for the whole batch (half-cells per lane 1 or 2, planes per
  // wave tile) unless the caller sets it: the most work per wave tile, at
  // most 4 planes x 2 half-cells, that leaves >= 6 wave tiles per CU over all
  // members, the single-problem rule applied to the batch.  On a batch of 64
  // uneven cities that is 4 x 2, 13-28 % faster than taking each member's
  // tiling autotuned for its single solves.
  int kb = params->tile_k > 0 ? params->tile_k : 0;
  int vq = params->tile_v;
  if (vq != 0 && vq != 1 && vq != 2) return QESMC_ERR_UNSUPPORTED;
  if (kb == 0 || vq == 0) {
    static const int order[][2] = {{4, 2}, {2, 2}, {4, 1}, {2, 1}, {1, 2}, {1, 1}};
    int bk = 0, bv = 0;
    long long most = -1;
    for (const auto &o : order) {
      if ((kb && o[0] != kb) || (vq && o[1] != vq)) continue;
      long long tiles = 0;
      for (int i = 0; i < n; ++i) {
        const Dims &d = ctxs[i]->d;
        if (d.NY < 3 || d.NZ < 3) continue;
        const long long flat = (long long)(d.NY - 2) * d.HP;
        tiles += ((flat + 64LL * o[1] - 1) / (64LL * o[1])) * ((d.NZ - 2 + o[0] - 1) / o[0]);
      }
      if (tiles >= 6LL * ctxs[0]->n_cu) {
        bk = o[0];
        bv = o[1];
        break;
      }
      if (tiles > most) {
        most = tiles;
        bk = o[0];
        bv = o[1];
      }
    }
    if (!kb) kb = bk ? bk : 4;
    if (!vq) vq = bv ? bv : 2;
  }
  qesmc_params pv = *params;
  pv.tile_v = vq;
  pv.tile_k = kb;
  params = &pv;
  qesmc_ctx *lead = ctxs[0];
  BCHK(hipSetDevice(lead->device));
  hipStream_t bs = lead->stream;
  std::vector<hipStream_t> own_streams(n);
  // every instance's own work completes before any context is moved to the
  // batch stream, so a failure here leaves every context untouched
  for (int i = 0; i < n; ++i) BCHK(hipStreamSynchronize(ctxs[i]->stream));
  for (int i = 0; i < n; ++i) {
    own_streams[i] = ctxs[i]->stream;
    ctxs[i]->stream = bs;  // every phase of every instance on the batch stream
  }
  auto restore = [&]() {
    for (int i = 0; i < n; ++i) ctxs[i]->stream = own_streams[i];
  };
  std::vector<RunState> st(n);
  int rc = QESMC_OK;
  for (int i = 0; i < n && rc == QESMC_OK; ++i) rc = run_prepare(ctxs[i], params, st[i]);
  if (rc != QESMC_OK) {
    (void)hipStreamSynchronize(bs);
    restore();
    return rc;
  }
  const qesmc_params &p = st[0].p;
  // device table
  std::vector<BatchItem> items(n);
  long long total = 0;
  for (int i = 0; i < n; ++i) {
    qesmc_ctx *x = ctxs[i];
    BatchItem &it = items[i];
    std::memset(&it, 0, sizeof(it));
    it.d = x->d;
    for (int c = 0; c < 2; ++c) {
      it.lc[c] = x->d_lc[c];
      it.Rc[c] = x->d_Rc[c];
      it.mask[c] = x->d_mask[c];
      for (int a = 0; a < 6; ++a) it.cc[c][a] = x->d_cc[c][a];
    }
    it.canon = x->d_canon;
    it.err = x->d_err;
    it.status = x->d_status;
    it.masked = st[i].masked ? 1 : 0;
    it.kb = x->kplanes < 1 ? 1 : x->kplanes;
    const long long flat = (long long)(x->d.NY - 2) * x->d.HP;
    it.tiles_per_plane = (int)((flat + 64LL * vq - 1) / (64LL * vq));
    const int nkc = x->d.NZ - 2 >= 1 && x->d.NY - 2 >= 1 ? (x->d.NZ - 2 + it.kb - 1) / it.kb : 0;
    it.tile_begin = total;
    total += (long long)it.tiles_per_plane * nkc;
  }
  BatchItem *d_items = nullptr;
  int *d_t = nullptr, *d_active = nullptr, *h_active = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&]() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (d_items) (void)hipFree(d_items);
    if (d_t) (void)hipFree(d_t);
    if (d_active) (void)hipFree(d_active);
    if (h_active) (void)hipHostFree(h_active);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
  auto fail = [&](hipError_t e) {
    (void)hipStreamSynchronize(bs);
    cleanup();
    restore();
    return map_err_b(e);
  };
  hipError_t e;
  if ((e = hipMalloc((void **)&d_items, sizeof(BatchItem) * n)) != hipSuccess) return fail(e);
  if ((e = hipMalloc((void **)&d_t, sizeof(int))) != hipSuccess) return fail(e);
  if ((e = hipMalloc((void **)&d_active, sizeof(int))) != hipSuccess) return fail(e);
  if ((e = hipHostMalloc((void **)&h_active, sizeof(int), hipHostMallocDefault)) != hipSuccess) return fail(e);
  if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) return fail(e);
  if ((e = hipMemcpyAsync(d_items, items.data(), sizeof(BatchItem) * n, hipMemcpyHostToDevice, bs)) != hipSuccess)
    return fail(e);
  if ((e = qesmc_launch_set_u32(lead, (unsigned *)d_t, 0u)) != hipSuccess) return fail(e);

  bool all_masked = true;
  for (int i = 0; i < n; ++i) all_masked = all_masked && st[i].masked;
  // persistent grid: never more work-groups than can be resident at once.
  // Unless the caller sets it: 5 per CU when the batch has >= 55 wave tiles
  // per CU (measured 3 % faster than 4 at 61 tiles per CU and 5-8 % at
  // 74-76: one 400 x 400 x 120 city, two 280 x 280 x 120 or four
  // 200 x 200 x 120, and on the 64-city batch), else the single-problem
  // kernel's choice, 4 for two half-cells per lane and 5 for one (4 measured
  // faster at 25, 32, 39 and 49 tiles per CU, by 7 %, 0.6 %, 8 % and 4 %).
  // Members' autotuned values tune the single-problem kernel and are not
  // used here.
  int wg = p.wg_per_cu > 0 ? p.wg_per_cu : total >= 55LL * lead->n_cu ? 5 : vq == 2 ? 4 : 5;
  {
    int fit = 0;
    const hipError_t oe =
        all_masked ? (vq == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, k_sor_batch<true, 2>, 256, 0)
                              : hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, k_sor_batch<true, 1>, 256, 0))
                   : (vq == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, k_sor_batch<false, 2>, 256, 0)
                              : hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, k_sor_batch<false, 1>, 256, 0));
    if (oe == hipSuccess && fit > 0 && fit < wg) wg = fit;
  }
  long long blocks = (total + 3) / 4;
  const long long cap = (long long)lead->n_cu * (wg < 1 ? 1 : wg);
  if (blocks > cap) blocks = cap;
  if (blocks < kXcdsB) blocks = kXcdsB;
  const float omo = 1.0f - p.omega;
  const bool any_runs = st[0].first_runs && total > 0;
  const int C = p.check_every;
  int launched = 0;
  if ((e = hipEventRecord(e0, bs)) != hipSuccess) return fail(e);
  if (any_runs) {
    auto launch_iter = [&](const int *tb, int toff) {
      for (int colour = 0; colour < 2; ++colour) {
#define QESMC_BATCH_LAUNCH(M, V)                                                                           \
  hipLaunchKernelGGL((k_sor_batch<M, V>), dim3((unsigned)blocks), dim3(256), 0, bs, d_items, n, total, colour, \
                     tb, toff, p.itermax, p.omega, omo, p.tol)
        if (all_masked) {
          if (vq == 2) QESMC_BATCH_LAUNCH(true, 2);
          else QESMC_BATCH_LAUNCH(true, 1);
        } else {
          if (vq == 2) QESMC_BATCH_LAUNCH(false, 2);
          else QESMC_BATCH_LAUNCH(false, 1);
        }
#undef QESMC_BATCH_LAUNCH
      }
    };
    if (use_graph) {
      // one captured chunk of C iterations, advancing the device counter
      if ((e = hipStreamBeginCapture(bs, hipStreamCaptureModeThreadLocal)) != hipSuccess) return fail(e);
      for (int j = 0; j < C; ++j) launch_iter(d_t, j);
      hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, bs, d_t, C);
      if ((e = hipStreamEndCapture(bs, &graph)) != hipSuccess) return fail(e);
      if ((e = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0)) != hipSuccess) return fail(e);
    }
    bool stop = false;
    while (launched < p.itermax && !stop) {
      const int chunk = std::min(C, p.itermax - launched);
      if (use_graph) {
        if ((e = hipGraphLaunch(gexec, bs)) != hipSuccess) return fail(e);
      } else {
        for (int j = 0; j < chunk; ++j) launch_iter(nullptr, launched + j);
        if ((e = hipGetLastError()) != hipSuccess) return fail(e);
      }
      launched += use_graph ? C : chunk;
      if (launched >= p.itermax) break;
      // instances still open at the next iteration
      hipLaunchKernelGGL(k_count_active, dim3(1), dim3(64), 0, bs, d_items, n, nullptr, launched, p.itermax, p.tol,
                         d_active);
      if ((e = hipMemcpyAsync(h_active, d_active, sizeof(int), hipMemcpyDeviceToHost, bs)) != hipSuccess)
        return fail(e);
      if ((e = hipStreamSynchronize(bs)) != hipSuccess) return fail(e);
      if (*h_active == 0) stop = true;
    }
    if (launched > p.itermax) launched = p.itermax;
  }
  if ((e = hipEventRecord(e1, bs)) != hipSuccess) return fail(e);
  for (int i = 0; i < n; ++i) {
    const int r = run_finish(ctxs[i], st[i], any_runs ? launched : 0, results ? &results[i] : nullptr);
    if (results) results[i].wg_per_cu = wg;  // the batch's launch, not the context's own default
    if (r != QESMC_OK && rc == QESMC_OK) rc = r;
  }
  if (ms_loop) {
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    *ms_loop = t;
  }
  cleanup();
  restore();
  return rc;
}

// Time `iterations` SOR iterations for each candidate tiling on the resident
// problem and keep the fastest as the context's default (tile fields 0).
// lambda is saved and restored, so the problem's state is unchanged.
int qesmc_autotune(qesmc_ctx *x, const qesmc_params *base, int iterations, qesmc_params *best)
{
  if (!x || !base || iterations < 1) return QESMC_ERR_ARG;
  if (!x->have_geom || !x->have_inflow) return QESMC_ERR_STATE;
  BCHK(hipSetDevice(x->device));
  float *saved = nullptr;
  BCHK(hipMalloc((void **)&saved, sizeof(float) * (size_t)x->ncell));
  {
    const hipError_t e =
        hipMemcpyAsync(saved, x->d_lam, sizeof(float) * (size_t)x->ncell, hipMemcpyDeviceToDevice, x->stream);
    if (e != hipSuccess) {
      (void)hipFree(saved);
      return map_hip_error(e);
    }
  }
  static const int cand[][3] = {{4, 2, 4}, {8, 2, 4}, {10, 2, 6}, {6, 2, 4}, {7, 2, 5},
                                {3, 2, 4}, {10, 2, 8}, {6, 2, 8}, {2, 2, 4}, {4, 1, 5},
                                {6, 1, 6}, {10, 1, 6}, {2, 1, 5}, {1, 2, 4}, {1, 1, 6}};
  double best_ms = 1e300;
  int bi = 0, rc = QESMC_OK;
  for (int c = 0; c < (int)(sizeof(cand) / sizeof(cand[0])); ++c) {
    qesmc_params p = *base;
    p.path = QESMC_PATH_SPLIT;
    p.itermax = iterations;
    p.tol = -1.0f;  // never converges: every candidate runs the same iterations
    p.tile_k = cand[c][0];
    p.tile_v = cand[c][1];
    p.wg_per_cu = cand[c][2];
    p.flags &= ~QESMC_RUN_REENCODE;
    qesmc_result r;
    double ms = 0.0;
    for (int rep = 0; rep < 2 && rc == QESMC_OK; ++rep) {  // first repetition warms caches
      const hipError_t e =
          hipMemcpyAsync(x->d_lam, saved, sizeof(float) * (size_t)x->ncell, hipMemcpyDeviceToDevice, x->stream);
      rc = e != hipSuccess ? map_hip_error(e) : qesmc_run(x, &p, &r);
      ms = r.ms_sor;
    }
    if (rc != QESMC_OK) break;
    if (ms < best_ms) {
      best_ms = ms;
      bi = c;
    }
  }
  (void)hipMemcpyAsync(x->d_lam, saved, sizeof(float) * (size_t)x->ncell, hipMemcpyDeviceToDevice, x->stream);
  (void)hipStreamSynchronize(x->stream);
  (void)hipFree(saved);
  if (rc != QESMC_OK) return rc;
  x->kplanes_auto = cand[bi][0];
  x->tile_v_auto = cand[bi][1];
  x->wg_per_cu_auto = cand[bi][2];
  if (best) {
    *best = *base;
    best->tile_k = cand[bi][0];
    best->tile_v = cand[bi][1];
    best->wg_per_cu = cand[bi][2];
  }
  return QESMC_OK;
}

}  // extern "C"

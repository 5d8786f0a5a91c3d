// Internal declarations shared by the HIP translation units of libqesmc.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "qesmc.h"

namespace qesmc {

// Pointers read from memory (tables, argument structs) are generic; casting
// them to the global address space lets loads and stores use global rather
// than flat instructions, as for pointers passed directly as kernel arguments.
#define QESMC_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ inline QESMC_GLOBAL T *as_global(T *p)
{
  return (QESMC_GLOBAL T *)p;
}

// Read-only data that a kernel never writes, addressed through a pointer read
// from memory: in the constant address space, wave-uniform loads from it go
// through the scalar cache (s_load), as they do for a const __restrict__
// kernel argument; through a global pointer they would be vector loads.
#define QESMC_CONSTANT __attribute__((address_space(4)))
template <typename T>
__device__ inline const QESMC_CONSTANT T *as_constant(const T *p)
{
  return (const QESMC_CONSTANT T *)p;
}

// Status bits raised by kernels into ctx->d_status.
enum : unsigned {
  ST_NONFINITE = 1u,      // a relaxed lambda was NaN or Inf
  ST_INCOMPRESSIBLE = 2u, // a coefficient is neither 0 nor its layer's wall value
  ST_ZERO_DIAG = 4u,      // an updated cell has e+f+g+h+m+n == 0
};

// Words of ctx->d_status: [0] status bits, [1..3] convergence-row summary
// of the last run, [4] the production stop rule's end iteration (all ones
// until fixed), [5] its extension factor (float bits, 0 = reference rule).
constexpr int kStatusWords = 8;
constexpr int kStopAt = 4;
constexpr int kStopExtend = 5;

// Iteration gate, decided identically by every wave of a launch.  g is the
// max of convergence row t (the previous iteration's max|dlambda|, 1.0 for
// t = 0).  Reference rule (extension 0): iteration t runs iff g > tol
// (Solver_CPU_RB.cpp:95).  Production rule (extension e > 0): the first
// iteration t* whose row is not above tol fixes the end at
// t* + ceil(e * t*), recorded in status[kStopAt] with atomicMin (every wave
// that sees t* proposes the same value, and later launches read it);
// iterations before the end run whatever their row says.  Within the launch
// of t* the recorded end is either unset or that value, both above t*, so
// the decision cannot differ between waves.
__device__ inline bool gate_pass(float g, int t, float tol, unsigned *status)
{
  const float e = __uint_as_float(status[kStopExtend]);
  if (!(e > 0.f)) return g > tol;
  const unsigned tu = (unsigned)t;
  const unsigned recorded = *(volatile unsigned *)(status + kStopAt);
  if (tu >= recorded) return false;
  if (g > tol) return true;
  const unsigned end = tu + (unsigned)ceilf(e * (float)t);
  if (end <= tu) return false;
  // record only while unset: every later iteration proposes a later end, and
  // thousands of waves contending for one address serialise at the memory side
  if (recorded > end && (threadIdx.x & 63) == 0) atomicMin(status + kStopAt, end);
  return true;
}

// max|dlambda| of one iteration is accumulated into kErrSlots striped
// words (block b -> slot b % kErrSlots): device-scope atomics from every
// XCD on one address serialise at the memory side, so they are spread.
constexpr int kErrSlots = 64;

// Grid constants passed by value to every kernel.
//   nx, ny, nz : staggered sizes (qes::Domain)
//   NX, NY, NZ : cell counts per axis = nx-1, ny-1, nz-1
//   HP         : padded length of one colour's half-row in the split layout
struct Dims {
  int nx, ny, nz;
  int NX, NY, NZ;
  int HP;
  long long sy, sz;    // cell strides (NX, NX*NY)
  long long fy, fz;    // face strides (nx, nx*ny)
  long long splane;    // split-layout plane stride NY*HP
};

}  // namespace qesmc

struct qesmc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[6] = {};  // 0-3 timing marks, 4-5 convergence polling

  qesmc::Dims d{};
  float dx = 0.f, dy = 0.f;
  long long ncell = 0, nface = 0, nsplit = 0;
  std::vector<float> dz_host;
  bool have_geom = false, have_inflow = false;
  bool have_velocity = false;  // d_u, d_v, d_w hold a corrected field of the resident geometry
  bool geom_checked = false;   // stencil encoding built for the current geometry
  bool no_alloc_run = false;   // inside qesmc_run_static: the run may not allocate
  bool compressible = false;

  // natural (x-fastest) device arrays
  float *d_coef[6] = {};
  int *d_flag = nullptr;
  float *d_dz = nullptr;
  float *d_u0 = nullptr, *d_v0 = nullptr, *d_w0 = nullptr;
  float *d_u = nullptr, *d_v = nullptr, *d_w = nullptr;
  float *d_R = nullptr;
  float *d_lam = nullptr;
  float *d_grad = nullptr;     // velocity gradient tensor, 9 cell arrays (allocated on first use)
  long long *d_probe = nullptr; // velocity probes: face indices (qesmc_set_probes)
  float *d_probe_val = nullptr; // gathered u, v, w of the probes (3 x n_probe)
  long long n_probe = 0, probe_cap = 0, probe_nface = 0;

  // colour-split arrays (index = colour)
  float *d_lc[2] = {};
  float *d_Rc[2] = {};
  unsigned char *d_mask[2] = {};
  float *d_cc[2][6] = {};     // full coefficients, split (only if not compressible)
  unsigned *d_canon = nullptr; // 6*NZ wall values per layer (float bits)
  unsigned *d_canon8 = nullptr; // fused path: the same, 8 words per layer (6 used), layer-major
  float *d_lc2[2] = {};       // second lambda buffer set of the fused path (allocated on first use)
  unsigned *d_cur = nullptr;   // fused path: buffer set holding the latest iterate (0: d_lc, 1: d_lc2)
  float *d_R0[2] = {};        // refine_every: R of the run (d_Rc then holds the residual of the base)
  float *d_base[2] = {};      // refine_every: the base lambda (the iterates hold the correction)

  // convergence slots, kErrSlots words per iteration: row 0 holds the 1.0f
  // sentinel, row t+1 the striped max|dlambda| of iteration t
  unsigned *d_err = nullptr;
  int err_cap = 0;             // rows allocated
  unsigned *d_status = nullptr; // kStatusWords words, see kStopAt
  double *d_norm = nullptr;
  unsigned *h_poll = nullptr;  // pinned host words: 2 polling rows + status
  int kplanes = 8;             // k planes marched by one SOR wave tile (this run)
  int kplanes_auto = 0;        // default when qesmc_params.tile_k == 0 (0: chosen per run)
  int tile_v = 1;              // half-cells per lane of the split kernel (this run)
  int tile_v_auto = 0;         // 0: chosen per run from the domain size
  int wg_per_cu = 8;           // resident work-groups per CU of the split kernel
  int wg_per_cu_auto = 0;      // 0: chosen per run from tile_v
  int slab_planes = 0;         // split path: k-slab of the interleaved red/black schedule (this run; 0: whole passes)
  int slab_planes_auto = 0;    // default when qesmc_params.slab_planes == 0
  int n_cu = 256;              // compute units of the device
};

namespace qesmc {

// Phases of one solve shared by qesmc_run and qesmc_run_batch.
struct RunState {
  qesmc_params p{};
  int path = QESMC_PATH_SPLIT;
  bool masked = false;
  bool first_runs = false;
};
int run_prepare(qesmc_ctx *x, const qesmc_params *pp, RunState &st);
int run_finish(qesmc_ctx *x, const RunState &st, int launched, qesmc_result *res);
int run_solve(qesmc_ctx *x, const qesmc_params *pp, qesmc_result *res);
int run_loop(qesmc_ctx *x, const RunState &st, int *launched);
// refine_every: set up base = lambda, R0 = R, R = residual of base, delta
// = 0 (refine_setup); fold delta into base and re-form the residual
// (refine_refresh); lambda = base + delta at the end (refine_fold_final).
// With `fused` the correction lives in whichever buffer set the march wrote
// last (d_cur, read on the device); otherwise in d_lc.
hipError_t refine_setup(qesmc_ctx *x, bool masked, bool fused);
hipError_t refine_refresh(qesmc_ctx *x, bool masked, bool fused);
hipError_t refine_fold_final(qesmc_ctx *x, bool fused);
// Whether the resident flags hold a cut cell (7 or 8).
hipError_t find_cut_cells(qesmc_ctx *x, bool *cut);
int static_scratch_init(int device);
float row_max(const unsigned *row);

// Geometry set-up shared by qesmc_set_geometry and the device city builder:
// reserve (re)allocates the device arrays when the grid size changes;
// commit encodes the stencil of the resident natural coefficients.
int geometry_reserve(qesmc_ctx *x, int nx, int ny, int nz);
int geometry_commit(qesmc_ctx *x);

inline int map_hip_error(hipError_t e)
{
  switch (e) {
  case hipSuccess: return QESMC_OK;
  case hipErrorOutOfMemory: return QESMC_ERR_OOM;
  case hipErrorNoDevice:
  case hipErrorInvalidDevice: return QESMC_ERR_NO_DEVICE;
  default: return QESMC_ERR_HIP;
  }
}

}  // namespace qesmc

// Red-black SOR colour pass on the colour-split layout: vectorised,
// persistent and XCD-aware.
//
// Layout (both colours): cell (i,j,k) of colour c = (i+j+k)&1 lives at
// half-index q = i>>1 of row (j,k) of colour array c, x = (k*NY + j)*HP + q.
// A colour-c cell at q has i = 2q + s, s = (j+k+c)&1; its neighbours in the
// other colour's array sit at q-1+s, q+s (x-1, x+1) and at q in rows j+-1
// and planes k+-1.
//
// Work decomposition.  The updated rows j in [1, NY-2] of one plane are seen
// as a flat run of (NY-2)*HP half-cells; a wave tile is 64*VQ consecutive
// half-cells (VQ per lane, one vector load per array) marched over KB planes.
// Tiles are numbered plane-chunk-major; XCD x (work-group id % 8 on MI355X)
// takes the x-th eighth of the tile range, so each L2 serves a contiguous
// slab and neighbour rows/planes are shared inside one XCD.  The grid is a
// fixed number of work-groups per CU that stride through their XCD's range.
//
// Arithmetic is the reference's operand for operand
// (src/winds/Solver_CPU_RB.cpp:109-151, compiled without contraction).
#include "qesmc_sor_tile.h"

namespace qesmc {

namespace {

constexpr int kXcds = 8;  // MI355X: 8 XCDs, work-groups dispatched round-robin

__device__ inline float wave_max(float v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

}  // namespace

template <bool MASKED, int VQ>
__global__ __launch_bounds__(256) void k_sor_split(Dims d, int colour, int kb, int k_lo, int k_hi,
                                                   int tiles_per_plane, long long ntiles,
                                                   float *__restrict__ own,
                                                   const float *__restrict__ oth,
                                                   const float *__restrict__ Rc,
                                                   const unsigned char *__restrict__ mask,
                                                   const unsigned *__restrict__ canon,
                                                   const float *__restrict__ ce,
                                                   const float *__restrict__ cf,
                                                   const float *__restrict__ cg,
                                                   const float *__restrict__ ch,
                                                   const float *__restrict__ cm,
                                                   const float *__restrict__ cn, float omega,
                                                   float one_m_omega, float tol, unsigned *err, int t,
                                                   int itermax, unsigned *status, int reverse)
{
  const int lane = threadIdx.x & 63;
  // uniform across the wave: tile indices then live in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (t >= itermax) return;
  {
    // gate: iteration t runs iff max(err row t) > tol (row 0 = 1.0f)
    const float g = wave_max(__uint_as_float(err[(long long)t * kErrSlots + lane]));
    if (!gate_pass(g, t, tol, status)) return;  // identical in every wave of the grid
  }
  const int xcd = blockIdx.x % kXcds;
  const int local = blockIdx.x / kXcds;
  const int nblk = (gridDim.x - xcd + kXcds - 1) / kXcds;
  const long long per = (ntiles + kXcds - 1) / kXcds;
  const long long tbeg = xcd * per;
  const long long tend = tbeg + per < ntiles ? tbeg + per : ntiles;

  float dl = 0.f;
  for (long long wt = tbeg + (long long)local * 4 + wave; wt < tend; wt += (long long)nblk * 4) {
    // with `reverse`, the black pass walks each XCD's tile range backwards:
    // each pass then starts on the planes the previous pass touched last,
    // still in that XCD's L2 and in the Infinity Cache (reversing the whole
    // range instead would hand each XCD another XCD's slab)
    const long long wr = (colour && reverse) ? tbeg + tend - 1 - wt : wt;
    const int kc = (int)(wr / tiles_per_plane);
    const int pt = (int)(wr - (long long)kc * tiles_per_plane);
    const int p = pt * (64 * VQ) + lane * VQ;
    const int jr = p / d.HP;
    const int j = jr + 1;
    const int q = p - jr * d.HP;
    if (j > d.NY - 2) continue;
    const int k0 = k_lo + kc * kb;
    const int k1 = k0 + kb < k_hi ? k0 + kb : k_hi;
    dl = sor_tile<MASKED, VQ>(d, colour, t, j, q, k0, k1, own, oth, Rc, mask, canon, ce, cf, cg, ch, cm, cn, omega,
                              one_m_omega, status, dl);
  }
  // block max -> one striped slot
  __shared__ float wmax[4];
  const float wm = wave_max(dl);
  if (lane == 0) wmax[wave] = wm;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float bm = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
    if (bm > 0.f) atomicMax(err + (long long)(t + 1) * kErrSlots + (blockIdx.x % kErrSlots), __float_as_uint(bm));
  }
}

}  // namespace qesmc

using namespace qesmc;

namespace {

// Work-groups of one k_sor_split variant that fit on a CU at once (queried
// once per variant): the persistent grid never asks for more, or the extra
// work-groups would start only after others finish.
int resident_per_cu(bool masked, int vq)
{
  static int cache[2][5] = {};
  int &c = cache[masked ? 1 : 0][vq];
  if (c == 0) {
    int fit = 0;
    hipError_t e;
    if (masked)
      e = vq == 4 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, k_sor_split<true, 4>, 256, 0)
        : vq == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, k_sor_split<true, 2>, 256, 0)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, k_sor_split<true, 1>, 256, 0);
    else
      e = vq == 4 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, k_sor_split<false, 4>, 256, 0)
        : vq == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, k_sor_split<false, 2>, 256, 0)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, k_sor_split<false, 1>, 256, 0);
    c = (e == hipSuccess && fit > 0) ? fit : 1 << 20;
  }
  return c;
}

}  // namespace

namespace {

// One colour pass over planes [k_lo, k_hi) of the updated range.
hipError_t launch_split_pass(qesmc_ctx *x, bool masked, int colour, int k_lo, int k_hi, int t, int itermax,
                             float omega, float tol, int reverse)
{
  const Dims &d = x->d;
  if (k_hi <= k_lo) return hipSuccess;
  const int vq = x->tile_v;
  const int kb = x->kplanes < 1 ? 1 : x->kplanes;
  const long long flat = (long long)(d.NY - 2) * d.HP;
  const int tiles_per_plane = (int)((flat + 64LL * vq - 1) / (64LL * vq));
  const int nkc = (k_hi - k_lo + kb - 1) / kb;
  const long long ntiles = (long long)tiles_per_plane * nkc;
  long long blocks = (ntiles + 3) / 4;
  int wg = x->wg_per_cu < 1 ? 1 : x->wg_per_cu;
  const int fit = resident_per_cu(masked, vq == 4 ? 4 : vq == 2 ? 2 : 1);
  if (fit < wg) wg = fit;
  const long long cap = (long long)x->n_cu * wg;
  if (blocks > cap) blocks = cap;
  if (blocks < kXcds) blocks = kXcds;  // every XCD's tile range needs a work-group
  const float omo = 1.0f - omega;
  float *own = x->d_lc[colour], *oth = x->d_lc[colour ^ 1];
  const float *const *cc = x->d_cc[colour];
  const unsigned char *mk = x->d_mask[colour];
#define QESMC_SPLIT_LAUNCH(M, V)                                                                          \
  hipLaunchKernelGGL((k_sor_split<M, V>), dim3((unsigned)blocks), dim3(256), 0, x->stream, d, colour, kb, \
                     k_lo, k_hi, tiles_per_plane, ntiles, own, oth, x->d_Rc[colour], mk, x->d_canon, cc[0], \
                     cc[1], cc[2], cc[3], cc[4], cc[5], omega, omo, tol, x->d_err, t, itermax, x->d_status,   \
                     reverse)
  if (masked) {
    if (vq == 4) QESMC_SPLIT_LAUNCH(true, 4);
    else if (vq == 2) QESMC_SPLIT_LAUNCH(true, 2);
    else QESMC_SPLIT_LAUNCH(true, 1);
  } else {
    if (vq == 4) QESMC_SPLIT_LAUNCH(false, 4);
    else if (vq == 2) QESMC_SPLIT_LAUNCH(false, 2);
    else QESMC_SPLIT_LAUNCH(false, 1);
  }
#undef QESMC_SPLIT_LAUNCH
  return hipGetLastError();
}

}  // namespace

// Launch one red-black iteration (red pass, black pass) on the split layout.
//
// With x->slab_planes = S > 0 the two passes are interleaved over k-slabs of
// S planes: red over slab s, then black over the same planes shifted down by
// one (black plane k needs red planes k-1..k+1 of this iteration, so it can
// run once red has done plane k+1), then red over slab s+1, and so on; the
// last black launch finishes the range.  Every red update still reads the
// previous iteration's black values and every black update this iteration's
// red values, exactly as with two whole-domain passes (a red launch over
// planes [a, b) reads black planes a-1..b, which no black launch of this
// iteration has touched yet; a black launch over [a', b-1) reads red planes
// up to b-1, all relaxed), so results are unchanged.  What changes is where
// the bytes come from: the black launch re-reads the lambda planes that the
// red launch just read and wrote, from the Infinity Cache instead of HBM,
// when a slab's footprint fits in it.
hipError_t qesmc_launch_split_iteration(qesmc_ctx *x, bool masked, int t, int itermax, float omega_red,
                                        float omega_black, float tol)
{
  const Dims &d = x->d;
  if (d.NZ - 2 < 1 || d.NY - 2 < 1) return hipSuccess;
  // Reverse the black pass only when the pass's arrays (lambda and R of both
  // colours and the masks) exceed MI355X's 256 MB Infinity Cache: measured
  // 2.7 % faster per iteration on 1000 x 1000 x 150, 0.8 % slower on
  // 400 x 400 x 120, whose arrays stay cache-resident either way.
  const int reverse = (double)x->nsplit * 2.0 * (4.0 + 4.0 + 1.0) > 256.0 * (1 << 20) ? 1 : 0;
  const int lo = 1, hi = d.NZ - 1;  // updated planes [1, NZ-2]
  const int S = x->slab_planes;
  if (S <= 0 || S >= hi - lo) {
    hipError_t e = launch_split_pass(x, masked, 0, lo, hi, t, itermax, omega_red, tol, reverse);
    if (e != hipSuccess) return e;
    return launch_split_pass(x, masked, 1, lo, hi, t, itermax, omega_black, tol, reverse);
  }
  int black_lo = lo;
  for (int a = lo; a < hi; a += S) {
    const int b = a + S < hi ? a + S : hi;
    hipError_t e = launch_split_pass(x, masked, 0, a, b, t, itermax, omega_red, tol, reverse);
    if (e != hipSuccess) return e;
    const int black_hi = b == hi ? hi : b - 1;
    e = launch_split_pass(x, masked, 1, black_lo, black_hi, t, itermax, omega_black, tol, reverse);
    if (e != hipSuccess) return e;
    black_lo = black_hi;
  }
  return hipSuccess;
}

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
#include "qesmc_internal.h"

namespace qesmc {

namespace {

constexpr int kXcds = 8;  // MI355X: 8 XCDs, work-groups dispatched round-robin

template <int VQ>
struct Vec;
template <>
struct Vec<1> {
  __device__ static inline void load(const float *p, float *v) { v[0] = p[0]; }
  __device__ static inline void store(float *p, const float *v) { p[0] = v[0]; }
  __device__ static inline void loadb(const unsigned char *p, unsigned *v) { v[0] = p[0]; }
};
template <>
struct Vec<2> {
  __device__ static inline void load(const float *p, float *v)
  {
    const float2 t = *reinterpret_cast<const float2 *>(p);
    v[0] = t.x;
    v[1] = t.y;
  }
  __device__ static inline void store(float *p, const float *v)
  {
    *reinterpret_cast<float2 *>(p) = make_float2(v[0], v[1]);
  }
  __device__ static inline void loadb(const unsigned char *p, unsigned *v)
  {
    const unsigned short t = *reinterpret_cast<const unsigned short *>(p);
    v[0] = t & 0xffu;
    v[1] = t >> 8;
  }
};
template <>
struct Vec<4> {
  __device__ static inline void load(const float *p, float *v)
  {
    const float4 t = *reinterpret_cast<const float4 *>(p);
    v[0] = t.x;
    v[1] = t.y;
    v[2] = t.z;
    v[3] = t.w;
  }
  __device__ static inline void store(float *p, const float *v)
  {
    *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
  }
  __device__ static inline void loadb(const unsigned char *p, unsigned *v)
  {
    const unsigned t = *reinterpret_cast<const unsigned *>(p);
    v[0] = t & 0xffu;
    v[1] = (t >> 8) & 0xffu;
    v[2] = (t >> 16) & 0xffu;
    v[3] = t >> 24;
  }
};

__device__ inline float wave_max(float v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

}  // namespace

template <bool MASKED, int VQ>
__global__ __launch_bounds__(256) void k_sor_split(Dims d, int colour, int kb, int tiles_per_plane,
                                                   long long ntiles, float *__restrict__ own,
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
                                                   int itermax, unsigned *status)
{
  const int lane = threadIdx.x & 63;
  // uniform across the wave: tile indices then live in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (t >= itermax) return;
  {
    // gate: iteration t runs iff max(err row t) > tol (row 0 = 1.0f)
    const float g = wave_max(__uint_as_float(err[(long long)t * kErrSlots + lane]));
    if (!(g > tol)) return;  // identical in every wave of the grid
  }
  const int xcd = blockIdx.x % kXcds;
  const int local = blockIdx.x / kXcds;
  const int nblk = (gridDim.x - xcd + kXcds - 1) / kXcds;
  const long long per = (ntiles + kXcds - 1) / kXcds;
  const long long tbeg = xcd * per;
  const long long tend = tbeg + per < ntiles ? tbeg + per : ntiles;

  float dl = 0.f;
  for (long long wt = tbeg + (long long)local * 4 + wave; wt < tend; wt += (long long)nblk * 4) {
    const int kc = (int)(wt / tiles_per_plane);
    const int pt = (int)(wt - (long long)kc * tiles_per_plane);
    const int p = pt * (64 * VQ) + lane * VQ;
    const int jr = p / d.HP;
    const int j = jr + 1;
    const int q = p - jr * d.HP;
    if (j > d.NY - 2) continue;
    const int k0 = 1 + kc * kb;
    const int k1 = k0 + kb < d.NZ - 1 ? k0 + kb : d.NZ - 1;
#pragma unroll 1
    for (int k = k0; k < k1; ++k) {
      const int s = (j + k + colour) & 1;
      const long long x = (long long)(k * d.NY + j) * d.HP + q;
      float o[VQ], r[VQ], c[VQ], yp[VQ], ym[VQ], zp[VQ], zm[VQ];
      Vec<VQ>::load(own + x, o);
      Vec<VQ>::load(Rc + x, r);
      Vec<VQ>::load(oth + x, c);
      Vec<VQ>::load(oth + x + d.HP, yp);
      Vec<VQ>::load(oth + x - d.HP, ym);
      Vec<VQ>::load(oth + x + d.splane, zp);
      Vec<VQ>::load(oth + x - d.splane, zm);
      const float left = oth[x - 1], right = oth[x + VQ];
      unsigned mk[VQ];
      if (MASKED) Vec<VQ>::loadb(mask + x, mk);
      float nvv[VQ];
      bool upd[VQ];
#pragma unroll
      for (int v = 0; v < VQ; ++v) {
        const int i = 2 * (q + v) + s;
        upd[v] = (i >= 1 && i <= d.NX - 2);
        const float xm = s ? c[v] : (v > 0 ? c[v - 1] : left);
        const float xp = s ? (v + 1 < VQ ? c[v + 1] : right) : c[v];
        float ec, fc, gc, hc, mc, nc;
        if (MASKED) {
          const unsigned m = mk[v];
          ec = (m & 1u) ? __uint_as_float(canon[0 * d.NZ + k]) : 0.f;
          fc = (m & 2u) ? __uint_as_float(canon[1 * d.NZ + k]) : 0.f;
          gc = (m & 4u) ? __uint_as_float(canon[2 * d.NZ + k]) : 0.f;
          hc = (m & 8u) ? __uint_as_float(canon[3 * d.NZ + k]) : 0.f;
          mc = (m & 16u) ? __uint_as_float(canon[4 * d.NZ + k]) : 0.f;
          nc = (m & 32u) ? __uint_as_float(canon[5 * d.NZ + k]) : 0.f;
        } else {
          ec = upd[v] ? ce[x + v] : 0.f;
          fc = upd[v] ? cf[x + v] : 0.f;
          gc = upd[v] ? cg[x + v] : 0.f;
          hc = upd[v] ? ch[x + v] : 0.f;
          mc = upd[v] ? cm[x + v] : 0.f;
          nc = upd[v] ? cn[x + v] : 0.f;
        }
        const float nv = (omega / (ec + fc + gc + hc + mc + nc))
                             * (ec * xp + fc * xm + gc * yp[v] + hc * ym[v] + mc * zp[v] + nc * zm[v] - r[v])
                         + one_m_omega * o[v];
        nvv[v] = nv;
        if (upd[v]) {
          if (!isfinite(nv)) atomicOr(status, ST_NONFINITE);
          const float dd = fabsf(nv - o[v]);
          dl = (dd > dl) ? dd : dl;  // NaN never wins, as in the reference's max test
          o[v] = nv;
        }
      }
      // positions that are not updated are written back unchanged: every
      // half-cell of this row segment belongs to this lane in this pass
      Vec<VQ>::store(own + x, o);
      // k = 0 mirror (Solver_CPU_RB.cpp:154-162), kept in the other colour:
      // own plane 0 at this half-index is ghost (i0,j,0), whose only reader is
      // cell (i0,j,1) of the other colour, read in the *next* pass.  Its value
      // must be lambda(i0,j,1) at the end of the previous iteration, which
      // is oth plane 1 right now (that colour is not relaxed until the next
      // pass).  In the first red pass the ghosts keep their initial values,
      // as the reference mirrors only after an iteration.
      if (k == 1 && !(colour == 0 && t == 0)) {
#pragma unroll
        for (int v = 0; v < VQ; ++v) {
          const int i0 = 2 * (q + v) + (s ^ 1);
          if (i0 >= 1 && i0 <= d.NX - 2) own[x - d.splane + v] = c[v];
        }
      }
      (void)nvv;
    }
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

// Launch one red-black iteration (red pass, black pass) on the split layout.
hipError_t qesmc_launch_split_iteration(qesmc_ctx *x, bool masked, int t, int itermax, float omega,
                                        float tol)
{
  const Dims &d = x->d;
  if (d.NZ - 2 < 1 || d.NY - 2 < 1) return hipSuccess;
  const int vq = x->tile_v;
  const int kb = x->kplanes < 1 ? 1 : x->kplanes;
  const long long flat = (long long)(d.NY - 2) * d.HP;
  const int tiles_per_plane = (int)((flat + 64LL * vq - 1) / (64LL * vq));
  const int nkc = (d.NZ - 2 + kb - 1) / kb;
  const long long ntiles = (long long)tiles_per_plane * nkc;
  long long blocks = (ntiles + 3) / 4;
  int wg = x->wg_per_cu < 1 ? 1 : x->wg_per_cu;
  const int fit = resident_per_cu(masked, vq == 4 ? 4 : vq == 2 ? 2 : 1);
  if (fit < wg) wg = fit;
  const long long cap = (long long)x->n_cu * wg;
  if (blocks > cap) blocks = cap;
  if (blocks < kXcds) blocks = kXcds;  // every XCD's tile range needs a work-group
  const float omo = 1.0f - omega;
  for (int colour = 0; colour < 2; ++colour) {
    float *own = x->d_lc[colour], *oth = x->d_lc[colour ^ 1];
    const float *const *cc = x->d_cc[colour];
    const unsigned char *mk = x->d_mask[colour];
#define QESMC_SPLIT_LAUNCH(M, V)                                                                          \
  hipLaunchKernelGGL((k_sor_split<M, V>), dim3((unsigned)blocks), dim3(256), 0, x->stream, d, colour, kb, \
                     tiles_per_plane, ntiles, own, oth, x->d_Rc[colour], mk, x->d_canon, cc[0], cc[1],    \
                     cc[2], cc[3], cc[4], cc[5], omega, omo, tol, x->d_err, t, itermax, x->d_status)
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
  }
  return hipGetLastError();
}
